// Batched local CC: many graphs of uneven size solved by one replay of a
// captured HIP graph.
//
// Each graph keeps its own CSR and label array in HBM (GraphDesc).  A
// block-to-graph schedule (WorkItem) gives every workgroup one graph and a
// vertex range of at most `item_vertices` vertices, so a batch of a 4M-vertex
// graph and a thousand 1K-vertex graphs fills the chip with no per-graph
// launches.  The stages are those of the single-graph solve
// (components.hip): parent = min(v, first neighbour), exact sampled hooking
// of neighbours 0 and 1 with L2-confirmed roots (in one pass by default,
// else a pass per neighbour with pointer jumping between), the most
// frequent root of each graph (one workgroup per graph), the frontier sweep
// of every vertex outside its graph's giant tree, pointer jumping.  Labels
// are per graph: the minimum vertex id of the component inside that graph.
//
// The tables (header, graph descriptors, schedule) live in device memory
// and are refreshed by a host-to-device copy node at the head of the
// captured graph, from pinned host mirrors, so one instantiated graph
// serves every batch up to its capacity.  The graph holds no memset nodes.

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "ccx_internal.h"
#include "../../../include/cc_hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kSamples = 1024;

struct GraphDesc {
  const uint64_t *row_ptr;
  const uint32_t *col;
  uint32_t *labels;
  uint32_t n;
  uint32_t giant;  // written on the device by kb_sample_giant
};

struct WorkItem {
  uint32_t graph;
  uint32_t v_begin;
  uint32_t v_end;
  uint32_t pad;
};

struct Header {
  uint32_t n_items;
  uint32_t n_graphs;
  uint32_t pad[2];
};

__device__ __forceinline__ uint32_t load_l2(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t root_of(uint32_t x, const uint32_t *p) {
  uint32_t px = p[x];
  while (px != x) {
    x = px;
    px = p[x];
  }
  return x;
}

__device__ __forceinline__ uint32_t root_via(uint32_t x, const uint32_t *p, uint32_t g, bool &hit) {
  bool h = (x == g);
  uint32_t px = p[x];
  while (px != x) {
    x = px;
    h |= (x == g);
    px = p[x];
  }
  hit = h;
  return x;
}

// unite() continued from two known (possibly stale) roots: the larger root
// is confirmed with an L2 read before the atomic, and a changed root is
// re-walked with plain reads (the single-graph solve's default mode; every
// walk still ends at or below its start, so max(ra, rb) keeps dropping).
__device__ __forceinline__ void unite_roots(uint32_t ra, uint32_t rb, uint32_t *p) {
  while (ra != rb) {
    const uint32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
    const uint32_t cur = load_l2(&p[hi]);
    if (cur != hi) {
      ra = root_of(cur, p);
      rb = root_of(lo, p);
      continue;
    }
    const uint32_t old = atomicMin(&p[hi], lo);
    if (old == hi || old == lo) return;
    ra = root_of(old, p);
    rb = root_of(lo, p);
  }
}

__device__ __forceinline__ void unite(uint32_t a, uint32_t b, uint32_t *p) {
  unite_roots(root_of(a, p), root_of(b, p), p);
}

// Work items are taken grid-stride: a launch has a few workgroups per CU
// whatever the capacity of the engine, so a small batch does not pay for
// thousands of workgroups that find no item and exit.
#define FOR_EACH_ITEM \
  for (uint32_t item_ = blockIdx.x; item_ < hdr->n_items; item_ += gridDim.x)
#define ITEM_VARS                        \
  const WorkItem it = items[item_];      \
  const GraphDesc gd = graphs[it.graph];

__global__ __launch_bounds__(kBlock) void kb_init_min(const Header *__restrict__ hdr,
                                                      const GraphDesc *__restrict__ graphs,
                                                      const WorkItem *__restrict__ items) {
  FOR_EACH_ITEM {
    ITEM_VARS
    for (uint32_t v = it.v_begin + threadIdx.x; v < it.v_end; v += kBlock) {
      uint32_t x = v;
      const uint64_t b = gd.row_ptr[v];
      if (b < gd.row_ptr[v + 1]) {
        const uint32_t w = gd.col[b];
        if (w < x) x = w;
      }
      gd.labels[v] = x;
    }
  }
}

__global__ __launch_bounds__(kBlock) void kb_hook_neighbour(const Header *__restrict__ hdr,
                                                            const GraphDesc *__restrict__ graphs,
                                                            const WorkItem *__restrict__ items,
                                                            uint32_t r) {
  FOR_EACH_ITEM {
    ITEM_VARS
    for (uint32_t v = it.v_begin + threadIdx.x; v < it.v_end; v += kBlock) {
      const uint64_t b = gd.row_ptr[v] + r;
      if (b >= gd.row_ptr[v + 1]) continue;
      const uint32_t w = gd.col[b];
      if (w < gd.n && !(r == 0 && w < v)) unite(v, w, gd.labels);
    }
  }
}

// Union by Rem's algorithm (the single-graph solve's rem_unite without
// splicing): both paths are climbed together, the side with the larger
// parent first, until the parents meet or the larger side is a root, which
// is then linked under the other side's parent by a CAS (after one L2 read
// confirms it is still a root).  Links only ever point a root at a smaller
// id of its component, so roots stay tree minima; each iteration returns or
// moves to a smaller id, so it ends.
__device__ __forceinline__ void rem_unite(uint32_t rx, uint32_t ry, uint32_t *p) {
  uint32_t px = p[rx], py = p[ry];
  while (px != py) {
    if (px < py) {
      uint32_t t = rx;
      rx = ry;
      ry = t;
      t = px;
      px = py;
      py = t;
    }
    if (rx == px) {
      const uint32_t cur = load_l2(&p[rx]);
      if (cur != rx) {
        px = cur;
        continue;
      }
      const uint32_t old = atomicCAS(&p[rx], rx, py);
      if (old == rx) return;
      px = old;
      continue;
    }
    rx = px;
    px = p[rx];
  }
}

// Both sampled neighbours in one pass (neighbour 0 only when it is above v;
// init linked it otherwise).  Rem: each by rem_unite.  Otherwise v's root
// and the roots of neighbours 0 and 1 are walked in lockstep, then v's tree
// is hooked with each.  One pass and one pointer-jumping launch fewer than
// a round per neighbour.
template <bool Rem>
__global__ __launch_bounds__(kBlock) void kb_hook_pair(const Header *__restrict__ hdr,
                                                       const GraphDesc *__restrict__ graphs,
                                                       const WorkItem *__restrict__ items) {
  FOR_EACH_ITEM {
    ITEM_VARS
    uint32_t *p = gd.labels;
    for (uint32_t v = it.v_begin + threadIdx.x; v < it.v_end; v += kBlock) {
      const uint64_t b = gd.row_ptr[v], e = gd.row_ptr[v + 1];
      const uint32_t w0 = b < e ? gd.col[b] : 0xFFFFFFFFu;
      const uint32_t w1 = b + 1 < e ? gd.col[b + 1] : 0xFFFFFFFFu;
      const bool a0 = w0 < gd.n && w0 > v;
      const bool a1 = w1 < gd.n && w1 != v;
      if (!a0 && !a1) continue;
      if (Rem) {
        if (a0) rem_unite(v, w0, p);
        if (a1) rem_unite(v, w1, p);
        continue;
      }
      uint32_t x[3] = {v, a0 ? w0 : v, a1 ? w1 : v};
      bool more = true;
      while (more) {  // the three root walks in lockstep
        more = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const uint32_t px = p[x[k]];
          if (px != x[k]) {
            x[k] = px;
            more = true;
          }
        }
      }
      uint32_t rv = x[0];
      if (a0 && rv != x[1]) {
        unite_roots(rv, x[1], p);
        rv = rv < x[1] ? rv : x[1];  // a member of the merged tree
      }
      if (a1 && rv != x[2]) unite_roots(rv, x[2], p);
    }
  }
}

__global__ __launch_bounds__(kBlock) void kb_compress(const Header *__restrict__ hdr,
                                                      const GraphDesc *__restrict__ graphs,
                                                      const WorkItem *__restrict__ items) {
  FOR_EACH_ITEM {
    ITEM_VARS
    uint32_t *p = gd.labels;
    for (uint32_t v = it.v_begin + threadIdx.x; v < it.v_end; v += kBlock) {
      const uint32_t pv = p[v];
      if (pv == v) continue;
      const uint32_t r = root_of(pv, p);
      if (r != pv) p[v] = r;
    }
  }
}

__device__ __forceinline__ uint32_t sample_index(uint32_t i, uint32_t n) {
  uint64_t z = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)(((z >> 32) * (uint64_t)n) >> 32);
}

// Each graph's most frequent root (ties: smaller root), one workgroup at a
// time per graph.
__global__ __launch_bounds__(kBlock) void kb_sample_giant(const Header *__restrict__ hdr,
                                                          GraphDesc *__restrict__ graphs) {
  constexpr uint32_t kSlots = 2 * kSamples;
  constexpr uint32_t kEmpty = 0xFFFFFFFFu;
  __shared__ uint32_t keys[kSlots];
  __shared__ uint32_t cnts[kSlots];
  __shared__ uint64_t best[kBlock / kWave];
  // grid-stride over the graphs; every barrier below is reached by the
  // whole workgroup (g is uniform across it)
  for (uint32_t g = blockIdx.x; g < hdr->n_graphs; g += gridDim.x) {
    GraphDesc &gd = graphs[g];
    const uint32_t n = gd.n;
    if (n == 0) continue;
    const int t = threadIdx.x;
    for (uint32_t i = t; i < kSlots; i += kBlock) {
      keys[i] = kEmpty;
      cnts[i] = 0;
    }
    __syncthreads();
    for (int i = t; i < kSamples; i += kBlock) {
      const uint32_t x = root_of(sample_index(i, n), gd.labels);
      uint32_t h = (x * 0x9E3779B1u) >> 21;
      for (;;) {
        const uint32_t prev = atomicCAS(&keys[h], kEmpty, x);
        if (prev == kEmpty || prev == x) {
          atomicAdd(&cnts[h], 1u);
          break;
        }
        h = (h + 1) & (kSlots - 1);
      }
    }
    __syncthreads();
    uint64_t mine = 0;
    for (uint32_t i = t; i < kSlots; i += kBlock)
      if (cnts[i]) {
        const uint64_t cand = ((uint64_t)cnts[i] << 32) | (uint32_t)~keys[i];
        if (cand > mine) mine = cand;
      }
    for (int d = kWave / 2; d > 0; d >>= 1) {
      const uint64_t o = __shfl_xor(mine, d, kWave);
      if (o > mine) mine = o;
    }
    if ((t & (kWave - 1)) == 0) best[t / kWave] = mine;
    __syncthreads();
    if (t == 0) {
      uint64_t b = best[0];
      for (int w = 1; w < kBlock / kWave; ++w)
        if (best[w] > b) b = best[w];
      gd.giant = ~(uint32_t)b;
    }
  }
}

__global__ __launch_bounds__(kBlock) void kb_hook_frontier(const Header *__restrict__ hdr,
                                                           const GraphDesc *__restrict__ graphs,
                                                           const WorkItem *__restrict__ items,
                                                           uint32_t k_first) {
  __shared__ uint64_t s_excl[kBlock / kWave][kWave];
  __shared__ uint64_t s_begin[kBlock / kWave][kWave];
  FOR_EACH_ITEM {
    ITEM_VARS
    uint32_t *p = gd.labels;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    for (uint32_t base = it.v_begin + wid * kWave; base < it.v_end; base += kBlock) {
      const uint32_t v = base + lane;
      uint64_t work = 0, begin = 0;
      if (v < it.v_end) {
        const uint32_t pv = p[v];
        bool in_giant = false;
        const uint32_t r = root_via(v, p, gd.giant, in_giant);
        if (r != pv) p[v] = r;
        if (!in_giant) {
          const uint64_t b = gd.row_ptr[v] + k_first, e = gd.row_ptr[v + 1];
          if (e > b) {
            work = e - b;
            begin = b;
          }
        }
      }
      uint64_t incl = work;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t up = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += up;
      }
      const uint64_t total = __shfl(incl, kWave - 1, kWave);
      if (total == 0) continue;
      s_excl[wid][lane] = incl - work;
      s_begin[wid][lane] = begin;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint64_t *ex = s_excl[wid];
      for (uint64_t e = lane; e < total; e += kWave) {
        int j = 0;
#pragma unroll
        for (int step = kWave / 2; step > 0; step >>= 1)
          if (ex[j + step] <= e) j += step;
        const uint32_t w = gd.col[s_begin[wid][j] + (e - ex[j])];
        if (w < gd.n) unite(base + j, w, p);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// Engine knobs, read when an engine is created (its graph is captured then).
int g_batch_one_pass = 1;    // hook both sampled neighbours in one pass
int g_batch_rem = 1;         // ... by Rem's unions (else lockstep root walks)
int g_batch_grid_per_cu = 8;  // workgroups per CU of each launch (0: one per item)

}  // namespace

extern "C" int cc_hip_batch_set(const char *key, int value) {
  const char *k = key ? key : "";
  if (strcmp(k, "one_pass") == 0 && (value == 0 || value == 1)) g_batch_one_pass = value;
  else if (strcmp(k, "rem") == 0 && (value == 0 || value == 1)) g_batch_rem = value;
  else if (strcmp(k, "grid_per_cu") == 0 && value >= 0 && value <= 64) g_batch_grid_per_cu = value;
  else return ccx::set_error(4, "cc_hip_batch_set: bad key or value (%s=%d)", k, value);
  return 0;
}

struct cc_hip_batch_s {
  uint32_t max_items = 0, max_graphs = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // device tables: [Header][GraphDesc x max_graphs][WorkItem x max_items]
  char *d_tables = nullptr;
  char *h_tables = nullptr;  // pinned mirror
  size_t table_bytes = 0;
  size_t off_graphs = 0, off_items = 0;
  bool pending = false;
};

static size_t align_up(size_t x, size_t a) { return (x + a - 1) & ~(a - 1); }

extern "C" int cc_hip_batch_create(uint32_t max_items, uint32_t max_graphs, cc_hip_batch_t *out) {
  if (!ccx::ctx()) return 2;
  if (!out || max_items == 0 || max_graphs == 0)
    return ccx::set_error(4, "cc_hip_batch_create: bad arguments");
  *out = nullptr;
  cc_hip_batch_s *b = new cc_hip_batch_s();
  b->max_items = max_items;
  b->max_graphs = max_graphs;
  b->off_graphs = align_up(sizeof(Header), 64);
  b->off_items = align_up(b->off_graphs + (size_t)max_graphs * sizeof(GraphDesc), 64);
  b->table_bytes = b->off_items + (size_t)max_items * sizeof(WorkItem);
  CCX_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  CCX_TRY(hipMalloc(&b->d_tables, b->table_bytes));
  CCX_TRY(hipHostMalloc(&b->h_tables, b->table_bytes, hipHostMallocDefault));
  memset(b->h_tables, 0, b->table_bytes);
  CCX_TRY(hipMemcpy(b->d_tables, b->h_tables, b->table_bytes, hipMemcpyHostToDevice));

  const Header *hdr = (const Header *)b->d_tables;
  GraphDesc *graphs = (GraphDesc *)(b->d_tables + b->off_graphs);
  const WorkItem *items = (const WorkItem *)(b->d_tables + b->off_items);
  hipStream_t s = b->stream;
  // launch widths: the kernels loop over items (graphs) grid-stride
  const uint32_t cus = (uint32_t)(ccx::ctx()->cu_count > 0 ? ccx::ctx()->cu_count : 256);
  const uint32_t cap = g_batch_grid_per_cu ? cus * (uint32_t)g_batch_grid_per_cu : max_items;
  const uint32_t gi = max_items < cap ? max_items : cap;
  const uint32_t gg = max_graphs < cap ? max_graphs : cap;
  CCX_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipMemcpyAsync(b->d_tables, b->h_tables, b->table_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    kb_init_min<<<gi, kBlock, 0, s>>>(hdr, graphs, items);
    if (g_batch_one_pass && g_batch_rem) {
      kb_hook_pair<true><<<gi, kBlock, 0, s>>>(hdr, graphs, items);
    } else if (g_batch_one_pass) {
      kb_hook_pair<false><<<gi, kBlock, 0, s>>>(hdr, graphs, items);
    } else {
      kb_hook_neighbour<<<gi, kBlock, 0, s>>>(hdr, graphs, items, 0u);
      kb_compress<<<gi, kBlock, 0, s>>>(hdr, graphs, items);
      kb_hook_neighbour<<<gi, kBlock, 0, s>>>(hdr, graphs, items, 1u);
    }
    kb_sample_giant<<<gg, kBlock, 0, s>>>(hdr, graphs);
    kb_hook_frontier<<<gi, kBlock, 0, s>>>(hdr, graphs, items, 2u);
    kb_compress<<<gi, kBlock, 0, s>>>(hdr, graphs, items);
  }
  hipError_t e2 = hipStreamEndCapture(s, &b->graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    delete b;
    return ccx::set_error(1, "cc_hip_batch_create: capture failed (%s)",
                          hipGetErrorString(e != hipSuccess ? e : e2));
  }
  CCX_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  *out = b;
  return 0;
}

extern "C" int cc_hip_batch_destroy(cc_hip_batch_t b) {
  if (!b) return 0;
  if (!ccx::ctx()) return 2;
  (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->d_tables);
  (void)hipHostFree(b->h_tables);
  ccx::release_stream_state(b->stream);  // direct solves on this stream made one
  (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

extern "C" int cc_hip_batch_launch(cc_hip_batch_t b, uint32_t n_graphs,
                                   const uint64_t *const *d_row_ptr,
                                   const uint32_t *const *d_col_idx, const uint32_t *n,
                                   uint32_t *const *d_labels, uint32_t item_vertices,
                                   uint32_t *items_used) {
  if (!ccx::ctx()) return 2;
  if (!b) return ccx::set_error(4, "cc_hip_batch_launch: null batch");
  if (n_graphs > b->max_graphs)
    return ccx::set_error(4, "cc_hip_batch_launch: %u graphs > capacity %u", n_graphs, b->max_graphs);
  if (item_vertices == 0 || item_vertices % kBlock)
    return ccx::set_error(4, "cc_hip_batch_launch: item_vertices must be a positive multiple of %d",
                          kBlock);
  if (n_graphs > 0 && (!d_row_ptr || !d_col_idx || !n || !d_labels))
    return ccx::set_error(4, "cc_hip_batch_launch: null pointer");
  uint64_t need = 0;
  for (uint32_t g = 0; g < n_graphs; ++g) {
    if (n[g] > ccx::kMaxVertices)
      return ccx::set_error(4, "cc_hip_batch_launch: graph %u has %u vertices (limit %u)", g, n[g],
                            ccx::kMaxVertices);
    need += (n[g] + (uint64_t)item_vertices - 1) / item_vertices;
  }
  if (need > b->max_items)
    return ccx::set_error(4, "cc_hip_batch_launch: %llu work items > capacity %u",
                          (unsigned long long)need, b->max_items);
  // the pinned mirror is read by the previous replay's copy node
  if (b->pending) CCX_TRY(hipStreamSynchronize(b->stream));
  Header *hdr = (Header *)b->h_tables;
  GraphDesc *gds = (GraphDesc *)(b->h_tables + b->off_graphs);
  WorkItem *its = (WorkItem *)(b->h_tables + b->off_items);
  uint32_t k = 0;
  for (uint32_t g = 0; g < n_graphs; ++g) {
    gds[g] = GraphDesc{d_row_ptr[g], d_col_idx[g], d_labels[g], n[g], 0u};
    for (uint64_t v = 0; v < n[g]; v += item_vertices) {
      const uint64_t e = v + item_vertices < n[g] ? v + item_vertices : n[g];
      its[k++] = WorkItem{g, (uint32_t)v, (uint32_t)e, 0u};
    }
  }
  hdr->n_items = k;
  hdr->n_graphs = n_graphs;
  if (items_used) *items_used = k;
  ccx::prof_begin("cc_batch_graph", b->stream);
  hipError_t e = hipGraphLaunch(b->exec, b->stream);
  ccx::prof_end(b->stream);
  if (e != hipSuccess) return ccx::set_error(1, "hipGraphLaunch: %s", hipGetErrorString(e));
  b->pending = true;
  return 0;
}

extern "C" int cc_hip_batch_query(cc_hip_batch_t b) {
  if (!b) return 1;
  hipError_t e = hipStreamQuery(b->stream);
  if (e == hipSuccess) {
    b->pending = false;
    return 1;
  }
  if (e == hipErrorNotReady) return 0;
  return -ccx::set_error(1, "cc_hip_batch_query: %s", hipGetErrorString(e));
}

extern "C" int cc_hip_batch_sync(cc_hip_batch_t b) {
  if (!b) return 0;
  if (!ccx::ctx()) return 2;
  CCX_TRY(hipStreamSynchronize(b->stream));
  b->pending = false;
  return 0;
}

extern "C" void *cc_hip_batch_stream(cc_hip_batch_t b) { return b ? (void *)b->stream : nullptr; }
