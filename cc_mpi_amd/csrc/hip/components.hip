// Local connected components on gfx950: labels[v] = smallest vertex id in
// v's component.
//
// Replaces the reference's pthreads label propagation
// (src/cc_pthreads.c:32-229), which lowers every vertex's label to the
// minimum over its neighbourhood round after round until nothing changes.
// The fixed point of that process is unique (the per-component minimum id),
// so any algorithm reaching the same fixed point is bit-exact with it.  This
// one is a parent-pointer forest that keeps the invariant p[v] <= v:
//
//   1. p[v] = min(v, first listed neighbour of v): an own-entry write with
//      no atomics.  With sorted rows (the generators' and the split's) this
//      is v's smallest neighbour (the result needs no row order); the
//      point is that afterwards every vertex already hangs at or below its
//      smallest neighbour, so a low-id vertex that later hooks onto a hub
//      finds the hub's root at or below its own id and lowers its own entry
//      instead of every such vertex hammering the hub's entry.
//   2. Sampled hooking: every vertex is united with its neighbour 0 (only
//      where step 1 did not already link them) and its neighbour 1, in one
//      pass: by Rem's unions (both paths climbed together, stopping where
//      they meet) on sparse graphs and on dense ones below 2^23 vertices,
//      else by walking v's root and both neighbours' roots in lockstep with
//      path halving, reading each row's first entries from the CSR
//      (or, as a tuning choice, a pass per neighbour with pointer jumping
//      between).  After this the forest already holds the giant component
//      of a power-law or random graph in one tree.  On those dense graphs
//      and on sparse graphs of degree 6 and more, a vertex linked
//      under neighbour 0 skips neighbour 1 and takes neighbour 0's parent
//      as its own (one hop): the trees then hang from the local minima
//      (vertices below all their neighbours), which unite with both
//      sampled neighbours, and those links join nearly every tree into the
//      giant one while only the minima walk.  On the dense ones the
//      ids below n/16 are hooked first; a bitmap of which of them then hang
//      under their most frequent root g0 (1 bit per id, small enough to
//      stay in every XCD's L2) lets each other vertex whose neighbour 0 is
//      marked link straight under g0 instead of gathering neighbour 0's
//      parent from HBM.
//   3. The most frequent root among 1024 sampled vertices is taken as the
//      giant root g.
//   4. Frontier sweep: every vertex whose parent is not g hooks with all its
//      remaining neighbours (indices 2..deg-1, or 1.. after skipping).
//      Edges between two vertices already under g are skipped; an edge
//      between a g-vertex and another vertex is still seen from the other
//      end because the graph is symmetric (with the knob symmetric = 0, for
//      a CSR holding one-directional entries, every vertex hooks with its
//      remaining neighbours, which gives the weakly connected components
//      the reference's push-pull propagation computes on such input).  Each wave64 takes 64
//      consecutive vertices, compacts the active ones with a wave prefix
//      sum of their remaining degrees and sweeps the concatenated neighbour
//      lists 64 entries at a time, so the col_idx reads of a wave are
//      contiguous runs regardless of how the degrees are spread over its
//      lanes.
//   5. Pointer jumping to the root: labels[v] = root(v) (skipped when the
//      frontier sweep linked no root: it then wrote every final root).
//
// Hooking: for an edge (a, b) with roots ra != rb, let hi = max, lo = min;
// old = atomicMin(&p[hi], lo).  If old == hi, hi was still a root and is now
// under lo.  Otherwise hi had meanwhile been hooked elsewhere (old < hi) and
// the pair (lo, old) is hooked in the next iteration, which keeps every
// connection the overwritten pointer carried.  max(ra, rb) strictly
// decreases per iteration, so the loop ends.  Because a pointer only ever
// moves to a smaller id, each tree's root is its minimum id, hence the final
// label is the component minimum whatever order the atomics land in.

#include <hip/hip_runtime.h>

#include <string>

#include "ccx_internal.h"
#include "../../../include/cc_hip.h"
#include "../../../include/cc.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kGiantSamples = 1024;

// Tuning knobs.  None changes the result: sampling only decides how much of
// the work the cheap per-vertex rounds do before the frontier sweep.  The
// defaults are the measured plan (notes at plan_for below).
struct Knobs {
  int sampled_neighbours = 2;  // sampled hooking rounds (1..4)
  int blocks_per_cu = 0;       // grid cap of the sampled hook (0 = auto)
  // 0: plain loads; 1: L2-confirmed roots, L2 re-walks; 2: all loads from
  // L2; 3: lossy sampling; 4: L2-confirmed roots, plain re-walks
  int hook_mode = 4;
  int stream_blocks_per_cu = 8;    // grid cap of the contention-free kernels (init, compress)
  int frontier_blocks_per_cu = 8;  // grid cap of the frontier sweep
  int hook_ilp = 0;                // roots walked per thread in the sampled hook (0 = auto)
  int stream_ilp = 1;              // vertices per thread per step in init/compress (1, 2, 4)
  int sample_cache = 1;            // sampled neighbours read once into coalesced arrays
  int fused_pair = -1;             // both sampled rounds in one pass (-1 = auto)
  int halve = -1;                  // path halving in the one-pass hook's walks (-1 = auto)
  int rem = -1;                    // one-pass hook by Rem's unions (-1 = auto)
  int rem_splice = -1;             // splicing in Rem's unions (-1 = auto)
  int giant_early = 1;             // two-pass path: sample the giant after round 0, skip it in round 1
  // non-temporal streaming (bitmask; -1 = auto): 1 the init pass stores the
  // sample arrays non-temporally, 2 it loads the column ids
  // non-temporally, 4 the one-pass hook loads the sample arrays
  // non-temporally
  int nt = -1;
  int init_ilp = 0;  // rows per thread per step of the sampling init pass (0 = auto; 1, 2, 4)
  int frontier_chunks = 0;  // 64-vertex chunks per wave step of the frontier sweep (0 = auto; 1, 2, 4)
  // the one-pass hook reads the first two entries of every row from the
  // CSR itself, after an identity init, instead of from sample arrays
  // written by a sampling init pass (-1 = auto)
  int csr_hook = -1;
  // a vertex linked under its neighbour 0 by its own entry skips the union
  // with neighbour 1 (only local minima and failed links take both); the
  // frontier sweep then starts at entry 1.  2: the link goes to neighbour
  // 0's parent instead (one hop), which keeps the trees flat (-1 = auto)
  int skip_linked = -1;
  // entries of its row a local minimum (neighbour 0 above it) unites with
  // when skipping is on (2..8; 0 = auto)
  int minima_entries = 0;
  // 1: the CSR stores both directions of every edge, so the frontier sweep
  // may skip the giant tree's members; 0: it unites every vertex's entries,
  // which gives the reference's labels (weakly connected components) on a
  // CSR that stores some edges in one direction only
  int symmetric = 1;
  // vertex tiles, the plan for graphs whose ids are local (k_tile_local):
  // -1 a probe on the device decides, for sparse graphs of 2^20+ vertices
  // on the plan that unites by Rem's algorithm from the sample arrays; 0
  // never; 1 always
  int tiles = -1;
  int tile_shift = 0;  // log2 of the vertices per tile (6..15; 0 = auto, 15)
  int tile_pairs = 0;  // pairs per tile segment of the cross list (0 = auto, one per vertex)
  // low-first hook (the skipping one-hop hook from the CSR): the ids below
  // n >> low_shift are hooked first, then a map of which of them hang
  // under their giant root lets the rest link there without reading
  // neighbour 0's parent (-1 = auto, 0 = off, 1..8)
  int low_shift = -1;
  // the low-first hook's passes queue the unions a vertex still owes after
  // its own-entry link per wave and run them 64 at a time (k_hook_queued):
  // 1, 0, or -1 = auto
  int split_walks = -1;
};
// Process-wide settings (cc_hip_components_set / _tune); the entry point
// without allocations always runs with the defaults.
Knobs g_knobs;
const Knobs kDefaultKnobs{};

// Parent read that bypasses the CU's L1 (an agent-scope atomic load), so it
// sees every hook that has reached L2.
__device__ __forceinline__ uint32_t load_l2(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x: follow parents until a fixed point.  Read-only: concurrent
// walkers through a hub would otherwise all store to the hub's entry, and
// same-address stores serialise in L2 while loads are served from L1.
// Trees are flattened by the separate pointer-jumping passes instead.  A
// stale parent read is still an ancestor (pointers only decrease), so the
// walk ends at a root or at a vertex the caller's atomicMin reveals as hooked.
__device__ __forceinline__ uint32_t find_root(uint32_t x, const uint32_t *__restrict__ p) {
  uint32_t px = p[x];
  while (px != x) {
    x = px;
    px = p[x];
  }
  return x;
}

__device__ __forceinline__ uint32_t find_root_l2(uint32_t x, const uint32_t *__restrict__ p) {
  uint32_t px = load_l2(&p[x]);
  while (px != x) {
    x = px;
    px = load_l2(&p[x]);
  }
  return x;
}

// Root walk after a hook found a root changed: L2 reads in modes 1 and 2,
// plain (possibly stale) reads otherwise.  Either way the walk ends at or
// below the ids it starts from, which is all the hook loop needs.
template <int Mode>
__device__ __forceinline__ uint32_t rewalk(uint32_t x, const uint32_t *__restrict__ p) {
  return (Mode == 1 || Mode == 2) ? find_root_l2(x, p) : find_root(x, p);
}

// Union of the trees of a and b.  Many members of one tree often hook
// towards the same other tree in the same launch; with plain loads their
// CUs keep seeing the tree's root as unhooked (stale L1 lines) and every one
// of them pays a same-address atomic.  Modes 1 and 4 confirm, with one L2
// read, that the larger root is still a root before the atomic; mode 1
// follows any change with L2 reads, mode 4 (the default) with plain reads,
// which measured faster on every graph tried; mode 2 walks with L2 reads
// throughout.  Returns true if this call linked a root (the forest's root
// set changed).
template <int Mode>
__device__ __forceinline__ bool hook(uint32_t a, uint32_t b, uint32_t *__restrict__ p) {
  uint32_t ra = Mode == 2 ? find_root_l2(a, p) : find_root(a, p);
  uint32_t rb = Mode == 2 ? find_root_l2(b, p) : find_root(b, p);
  while (ra != rb) {
    const uint32_t hi = ra > rb ? ra : rb;
    const uint32_t lo = ra > rb ? rb : ra;
    if (Mode != 0) {
      const uint32_t cur = load_l2(&p[hi]);
      if (cur != hi) {  // hooked meanwhile: cur < hi, so max(ra, rb) still drops
        ra = rewalk<Mode>(cur, p);
        rb = rewalk<Mode>(lo, p);
        continue;
      }
    }
    const uint32_t old = atomicMin(&p[hi], lo);
    if (old == hi) return true;
    if (old == lo) return false;
    ra = rewalk<Mode>(old, p);
    rb = rewalk<Mode>(lo, p);
  }
  return false;
}

// Lossy union used only by the sampling rounds of mode 3: one
// fire-and-forget atomicMin on the larger root, no confirmation and no
// retry.  p[hi] only ever drops to an id connected to hi, so the forest
// stays valid (every tree inside one component); a link it loses is
// recovered by the frontier sweep, which in mode 3 hooks every neighbour of
// every vertex outside the giant tree.
__device__ __forceinline__ void hook_lossy(uint32_t a, uint32_t b, uint32_t *__restrict__ p) {
  const uint32_t ra = find_root(a, p);
  const uint32_t rb = find_root(b, p);
  if (ra == rb) return;
  (void)__hip_atomic_fetch_min(&p[ra > rb ? ra : rb], ra > rb ? rb : ra, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

#define CCX_GRID_STRIDE(v, n)                                                       \
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < (n);                     \
       v += gridDim.x * kBlock)

__global__ __launch_bounds__(kBlock) void k_init(uint32_t n, uint32_t *__restrict__ p) {
  CCX_GRID_STRIDE(v, n) p[v] = v;
}

// p[v] = min(v, col[row_ptr[v]]): links v to its first listed neighbour when
// that one is smaller.  Writes only v's own entry.
__global__ __launch_bounds__(kBlock) void k_init_min(const uint64_t *__restrict__ row_ptr,
                                                     const uint32_t *__restrict__ col,
                                                     uint32_t n, uint32_t *__restrict__ p) {
  CCX_GRID_STRIDE(v, n) {
    uint32_t x = v;
    const uint64_t b = row_ptr[v];
    if (b < row_ptr[v + 1]) {
      const uint32_t w = col[b];
      if (w < x) x = w;
    }
    p[v] = x;
  }
}

// Hook every vertex with its neighbour number r (if it has one).
template <int Mode>
__global__ __launch_bounds__(kBlock) void k_hook_neighbour(const uint64_t *__restrict__ row_ptr,
                                                           const uint32_t *__restrict__ col,
                                                           uint32_t n, uint32_t r,
                                                           uint32_t *__restrict__ p) {
  CCX_GRID_STRIDE(v, n) {
    const uint64_t b = row_ptr[v] + r;
    if (b >= row_ptr[v + 1]) continue;
    const uint32_t w = col[b];
    // neighbour 0 below v was linked by k_init_min already
    if (w < n && !(r == 0 && w < v)) {
      if (Mode == 3) hook_lossy(v, w, p);
      else hook<Mode>(v, w, p);
    }
  }
}

// Pointer jumping: p[v] = root(v).  Run with no hooks in flight.  With the
// solve's slots (slots[0] the sampled giant root, slots[1] the frontier
// sweep's link flag) it is the last pass: skipped when the flag is 0 (the
// sweep linked no root and has already written every vertex's final
// root), and otherwise an entry equal to the giant root needs no walk when
// that root is still a root (the sweep's write-back pointed nearly every
// vertex at it), which leaves most vertices a single streamed load.
__device__ __forceinline__ uint32_t known_root(const uint32_t *__restrict__ slots,
                                               const uint32_t *__restrict__ p) {
  if (!slots) return 0xFFFFFFFFu;
  const uint32_t g = slots[0];
  return p[g] == g ? g : 0xFFFFFFFFu;
}

// sw: the plan switch word (see k_tile_probe), if both plans are launched:
// the pass runs only when it holds sw_run.
__global__ __launch_bounds__(kBlock) void k_compress(uint32_t n, uint32_t *__restrict__ p,
                                                     const uint32_t *__restrict__ slots,
                                                     const uint32_t *__restrict__ sw = nullptr,
                                                     uint32_t sw_run = 0) {
  if (sw && sw[0] != sw_run) return;
  if (slots && slots[1] == 0) return;
  const uint32_t g = known_root(slots, p);
  CCX_GRID_STRIDE(v, n) {
    const uint32_t pv = p[v];
    if (pv == v || pv == g) continue;
    const uint32_t r = find_root(pv, p);
    if (r != pv) p[v] = r;
  }
}

// ---- ILP variants: each thread carries ILP vertices through every
// dependent-load step together (v, v + S, ..., S = grid size), so ILP
// independent gathers are in flight per step instead of one.  The kernels
// are latency-bound chains (offset -> neighbour -> parent -> parent's
// parent), so this multiplies memory-level parallelism without adding
// waves (more waves would add contention on hot roots).

template <int ILP>
__device__ __forceinline__ void roots_lockstep(uint32_t (&x)[ILP], const bool (&act)[ILP],
                                               const uint32_t *__restrict__ p) {
  bool more = true;
  while (more) {
    more = false;
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      if (!act[k]) continue;
      const uint32_t px = p[x[k]];
      if (px != x[k]) {
        x[k] = px;
        more = true;
      }
    }
  }
}

// hook() continued from two known (possibly stale) roots.
template <int Mode>
__device__ __forceinline__ void hook_roots(uint32_t ra, uint32_t rb, uint32_t *__restrict__ p) {
  while (ra != rb) {
    const uint32_t hi = ra > rb ? ra : rb;
    const uint32_t lo = ra > rb ? rb : ra;
    if (Mode != 0) {
      const uint32_t cur = load_l2(&p[hi]);
      if (cur != hi) {
        ra = rewalk<Mode>(cur, p);
        rb = rewalk<Mode>(lo, p);
        continue;
      }
    }
    const uint32_t old = atomicMin(&p[hi], lo);
    if (old == hi || old == lo) return;
    ra = rewalk<Mode>(old, p);
    rb = rewalk<Mode>(lo, p);
  }
}

// roots_lockstep with path halving: each step moves x to its grandparent
// and lowers p[x] to that grandparent.  The lowering is an atomic min (with
// no return value, so the wave does not wait for it): p[x] only ever moves
// to a smaller ancestor, which keeps the forest valid whatever hooks land
// in between.  A plain store could raise an entry that a concurrent hook
// had just lowered, which lost links in practice.  Halving pays where the
// parent array is far larger than the L2s (n >= 2^24): each walk then costs
// HBM lines, and flatter trees make every later walk, including the
// frontier sweep's, shorter.
template <int ILP>
__device__ __forceinline__ void roots_lockstep_halve(uint32_t (&x)[ILP], const bool (&act)[ILP],
                                                     uint32_t *__restrict__ p) {
  bool more = true;
  while (more) {
    more = false;
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      if (!act[k]) continue;
      const uint32_t px = p[x[k]];
      if (px != x[k]) {
        const uint32_t ppx = p[px];
        if (ppx != px)
          (void)__hip_atomic_fetch_min(&p[x[k]], ppx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x[k] = ppx;
        more = true;
      }
    }
  }
}

template <int ILP>
__global__ __launch_bounds__(kBlock) void k_init_min_ilp(const uint64_t *__restrict__ row_ptr,
                                                         const uint32_t *__restrict__ col,
                                                         uint32_t n, uint32_t *__restrict__ p) {
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint64_t b[ILP], e[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      b[k] = (v < n) ? row_ptr[v] : 0;
      e[k] = (v < n) ? row_ptr[v + 1] : 0;
    }
    uint32_t w[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) w[k] = (b[k] < e[k]) ? col[b[k]] : 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      if (v < n) p[v] = w[k] < (uint32_t)v ? w[k] : (uint32_t)v;
    }
  }
}

template <int Mode, int ILP>
__global__ __launch_bounds__(kBlock) void k_hook_neighbour_ilp(const uint64_t *__restrict__ row_ptr,
                                                               const uint32_t *__restrict__ col,
                                                               uint32_t n, uint32_t r,
                                                               uint32_t *__restrict__ p) {
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint64_t b[ILP], e[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      b[k] = (v < n) ? row_ptr[v] + r : 1;
      e[k] = (v < n) ? row_ptr[v + 1] : 0;
    }
    uint32_t ra[ILP], rb[ILP];
    bool act[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint32_t v = (uint32_t)(v0 + k * S);
      act[k] = b[k] < e[k];
      rb[k] = act[k] ? col[b[k]] : 0;
      // neighbour 0 below v was linked by the init kernel already
      act[k] = act[k] && rb[k] < n && !(r == 0 && rb[k] < v);
      ra[k] = v;
    }
    // both endpoints' roots walked together: 2*ILP gathers per step
    uint32_t x[2 * ILP];
    bool a2[2 * ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      x[k] = ra[k];
      x[ILP + k] = rb[k];
      a2[k] = a2[ILP + k] = act[k];
    }
    roots_lockstep<2 * ILP>(x, a2, p);
#pragma unroll
    for (int k = 0; k < ILP; ++k)
      if (act[k] && x[k] != x[ILP + k]) hook_roots<Mode>(x[k], x[ILP + k], p);
  }
}

template <int ILP>
__global__ __launch_bounds__(kBlock) void k_compress_ilp(uint32_t n, uint32_t *__restrict__ p,
                                                         const uint32_t *__restrict__ slots,
                                                         const uint32_t *__restrict__ sw = nullptr,
                                                         uint32_t sw_run = 0) {
  if (sw && sw[0] != sw_run) return;
  if (slots && slots[1] == 0) return;
  const uint32_t g = known_root(slots, p);
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint32_t x[ILP], first[ILP];
    bool act[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      act[k] = v < n;
      x[k] = act[k] ? p[v] : 0;
      first[k] = x[k];
      act[k] = act[k] && x[k] != (uint32_t)v && x[k] != g;
    }
    roots_lockstep<ILP>(x, act, p);
#pragma unroll
    for (int k = 0; k < ILP; ++k)
      if (act[k] && x[k] != first[k]) p[v0 + k * S] = x[k];
  }
}

// ---- Sampled-neighbour cache.  Each sampled round gathers one column
// entry per vertex: a random 4-byte read that moves a whole cache line from
// HBM.  The init pass reads the first `Rounds` entries of every row once
// (they share that line) into one coalesced array per round, so the rounds
// stream 4 bytes per vertex instead of gathering a line plus an offset.
constexpr uint32_t kNoNeighbour = 0xFFFFFFFFu;
// Low-first hook (Knobs::low_shift) over the ids below n >> 4 on the plan
// that takes it (the skipping one-hop hook from the CSR).  Per solve:
// Chung-Lu 65.6M 2.45 -> 2.27 ms (shifts 1/2/3/5: 2.37/2.30/2.30/2.30),
// Chung-Lu 40M 1.58 -> 1.47-1.51, R-MAT 27 3.84 -> 3.78.
constexpr int kLowShiftAuto = 4;
// Queued owed unions (Knobs::split_walks), on below 2^31 entries
// (kSplitWalksMaxEntries), per solve: R-MAT 24 0.547 -> 0.493 ms, R-MAT 25
// 1.05 -> 0.905, R-MAT 26 2.00 -> 1.785, Chung-Lu 20M 0.749 -> 0.737,
// Chung-Lu 32M 1.26 -> 1.22; above, Chung-Lu 65.6M 2.27 -> 2.32 and
// Chung-Lu 40M 1.47 -> 1.50 lose (R-MAT 27 3.80 -> 3.46 would gain).

// Nt: bit 1 stores the sample arrays non-temporally, bit 2 loads the
// column ids non-temporally (both are streamed once; the parent array
// written here is what the hook gathers from next).
// off: the plan switch word (see k_tile_probe); nonzero means the tile
// plan took the solve, and this pass returns at once.
template <int Rounds, int Nt = 0>
__global__ __launch_bounds__(kBlock) void k_init_sample(const uint64_t *__restrict__ row_ptr,
                                                        const uint32_t *__restrict__ col,
                                                        uint32_t n, uint32_t *__restrict__ p,
                                                        uint32_t *__restrict__ s,
                                                        const uint32_t *__restrict__ off = nullptr) {
  if (off && off[0]) return;
  CCX_GRID_STRIDE(v, n) {
    const uint64_t b = row_ptr[v];
    const uint64_t e = row_ptr[v + 1];
    uint32_t w[Rounds];
#pragma unroll
    for (int r = 0; r < Rounds; ++r)
      w[r] = (b + r < e) ? ((Nt & 2) ? __builtin_nontemporal_load(&col[b + r]) : col[b + r])
                         : kNoNeighbour;
    p[v] = w[0] < v ? w[0] : v;
#pragma unroll
    for (int r = 0; r < Rounds; ++r) {
      if (Nt & 1) __builtin_nontemporal_store(w[r], &s[(size_t)r * n + v]);
      else s[(size_t)r * n + v] = w[r];
    }
  }
}

// k_init_sample<2, Nt> with ILP rows per thread per step (v, v + S, ...,
// S = grid size): every row's offsets, then every row's two column ids,
// are loaded before any is used, so ILP dependent chains overlap.
template <int Nt, int ILP>
__global__ __launch_bounds__(kBlock) void k_init_sample2_ilp(const uint64_t *__restrict__ row_ptr,
                                                             const uint32_t *__restrict__ col,
                                                             uint32_t n, uint32_t *__restrict__ p,
                                                             uint32_t *__restrict__ s) {
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint64_t b[ILP], e[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      b[k] = v < n ? row_ptr[v] : 0;
      e[k] = v < n ? row_ptr[v + 1] : 0;
    }
    uint32_t w0[ILP], w1[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      w0[k] = b[k] < e[k] ? ((Nt & 2) ? __builtin_nontemporal_load(&col[b[k]]) : col[b[k]])
                          : kNoNeighbour;
      w1[k] = b[k] + 1 < e[k]
                  ? ((Nt & 2) ? __builtin_nontemporal_load(&col[b[k] + 1]) : col[b[k] + 1])
                  : kNoNeighbour;
    }
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      if (v >= n) continue;
      p[v] = w0[k] < (uint32_t)v ? w0[k] : (uint32_t)v;
      if (Nt & 1) {
        __builtin_nontemporal_store(w0[k], &s[v]);
        __builtin_nontemporal_store(w1[k], &s[(size_t)n + v]);
      } else {
        s[v] = w0[k];
        s[(size_t)n + v] = w1[k];
      }
    }
  }
}

// Round r of the sampled hook from the cache (s = round r's array).  With
// Skip (rounds after the giant root g was sampled, on a compressed forest,
// so p[v] is v's root), vertices already under g are skipped: an edge
// between two vertices of the giant tree links nothing, and an edge from
// the giant tree to another tree is hooked from that tree's endpoint, which
// visits all of its neighbours (the CSR is symmetric).
template <int Mode, int ILP, bool Skip>
__global__ __launch_bounds__(kBlock) void k_hook_cached(const uint32_t *__restrict__ s,
                                                        uint32_t n, uint32_t r,
                                                        uint32_t *__restrict__ p,
                                                        const uint32_t *__restrict__ giant_slot) {
  const uint32_t g = Skip ? giant_slot[0] : 0u;
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint32_t x[2 * ILP];
    bool act[2 * ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      // Skip: start from p[v] (an ancestor of v) and drop the giant tree
      const uint32_t start = (Skip && v < n) ? p[v] : (uint32_t)v;
      const bool live = v < n && !(Skip && start == g);
      const uint32_t w = live ? s[v] : kNoNeighbour;
      // neighbour 0 below v was linked by the init kernel already
      const bool a = w < n && !(r == 0 && w < (uint32_t)v);
      x[k] = start;
      x[ILP + k] = w;
      act[k] = act[ILP + k] = a;
    }
    roots_lockstep<2 * ILP>(x, act, p);
#pragma unroll
    for (int k = 0; k < ILP; ++k)
      if (act[k] && x[k] != x[ILP + k]) hook_roots<Mode>(x[k], x[ILP + k], p);
  }
}

// Both sampled rounds of a two-round solve in one pass: v's root is walked
// once and hooked with the roots of neighbours 0 and 1 in turn, so the
// pass between the rounds (and its pointer-jumping launch) disappears.
// Halve: the three root walks halve the paths they follow.
// Nt: the sample arrays are loaded non-temporally (streamed once).
//
// Skip1: a vertex the init kernel linked under neighbour 0 skips neighbour
// 1 (2: and its entry is lowered to neighbour 0's parent, one hop); only
// local minima unite with both (the reasoning at k_hook_rem).
template <int Mode, int ILP, bool Halve = false, bool Nt = false, int Skip1 = 0>
__global__ __launch_bounds__(kBlock) void k_hook_pair(const uint32_t *__restrict__ s0,
                                                      const uint32_t *__restrict__ s1,
                                                      uint32_t n, uint32_t *__restrict__ p) {
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint32_t x[3 * ILP];
    bool act[3 * ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      const uint32_t w0 = v < n ? (Nt ? __builtin_nontemporal_load(&s0[v]) : s0[v]) : kNoNeighbour;
      const uint32_t w1 = v < n ? (Nt ? __builtin_nontemporal_load(&s1[v]) : s1[v]) : kNoNeighbour;
      // neighbour 0 below v was linked by the init kernel already
      act[ILP + k] = w0 < n && w0 > (uint32_t)v;
      if (Skip1 == 2 && w0 < (uint32_t)v) {
        const uint32_t q = p[w0];
        if (q < w0) (void)__hip_atomic_fetch_min(&p[v], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      act[2 * ILP + k] = w1 < n && w1 != (uint32_t)v && (!Skip1 || act[ILP + k]);
      act[k] = act[ILP + k] || act[2 * ILP + k];
      x[k] = (uint32_t)v;
      x[ILP + k] = w0;
      x[2 * ILP + k] = w1;
    }
    if (Halve) roots_lockstep_halve<3 * ILP>(x, act, p);
    else roots_lockstep<3 * ILP>(x, act, p);
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      uint32_t rv = x[k];
      const uint32_t r0 = x[ILP + k], r1 = x[2 * ILP + k];
      if (act[ILP + k] && rv != r0) {
        hook_roots<Mode>(rv, r0, p);
        rv = rv < r0 ? rv : r0;  // a member of the merged tree
      }
      if (act[2 * ILP + k] && rv != r1) hook_roots<Mode>(rv, r1, p);
    }
  }
}

// Union of the trees of v and w by climbing both paths together (Rem's
// algorithm, min-root form): always step the side whose parent is larger;
// stop as soon as both sides have the same parent (same tree: no need to
// reach the root) or when the larger side is a root, which is then linked
// under the other side's parent by a CAS from itself (after one L2 read
// confirms it is still a root).  Splicing lowers the entry a step leaves
// to the other side's parent (an atomic min, so an entry never rises), so
// the next walk through it is shorter: Splice 1 on every step, Splice 2
// only on v's own entry (a neighbour's entry may be a hub that many threads
// pass through, where the atomics would queue).  Every link and splice
// points an entry at a smaller id of the same component once this union
// completes, and the union completes in this call, so the forest keeps
// p[v] <= v and each root is its tree's minimum.  Each iteration returns,
// or moves the larger side to its parent, or learns a smaller parent for it
// from a failed CAS, so the loop ends.
template <int Splice>
__device__ __forceinline__ void rem_unite(const uint32_t v, const uint32_t w,
                                          uint32_t *__restrict__ p) {
  uint32_t rx = v, ry = w;
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
    if (rx == px) {  // rx looks like a root; py < rx
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
    if (Splice == 1 || (Splice == 2 && rx == v))
      (void)__hip_atomic_fetch_min(&p[rx], py, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    rx = px;
    px = p[rx];
  }
}

// Both sampled neighbours of every vertex united with it by rem_unite (one
// pass, no lockstep: each union ends as soon as the two paths meet, which
// on graphs whose trees are shallow and hub-centred is well before the
// roots).
//
// Skip1: a vertex that the init kernel linked under neighbour 0 skips
// neighbour 1, and a local minimum (neighbour 0 above it) unites with the
// first `minima_entries` entries of its row (read from the CSR).  Every
// vertex still hangs under neighbour 0's tree; the trees are joined at
// their roots, the local minima, each with up to `minima_entries` links
// leaving it, which on power-law and random graphs joins nearly all of them
// into the giant tree while only the minima pay for random walks.  The
// frontier sweep then starts at entry 1.
template <int Splice, bool Nt = false, int Skip1 = 0>
__global__ __launch_bounds__(kBlock) void k_hook_rem(const uint32_t *__restrict__ s0,
                                                     const uint32_t *__restrict__ s1, uint32_t n,
                                                     uint32_t *__restrict__ p,
                                                     const uint64_t *__restrict__ row_ptr = nullptr,
                                                     const uint32_t *__restrict__ col = nullptr,
                                                     uint32_t minima_entries = 2,
                                                     const uint32_t *__restrict__ off = nullptr) {
  if (off && off[0]) return;
  CCX_GRID_STRIDE(v, n) {
    const uint32_t w0 = Nt ? __builtin_nontemporal_load(&s0[v]) : s0[v];
    const uint32_t w1 = Nt ? __builtin_nontemporal_load(&s1[v]) : s1[v];
    // neighbour 0 below v was linked by the init kernel already
    const bool up0 = w0 < n && w0 > v;
    if (up0) rem_unite<Splice>(v, w0, p);
    if (Skip1 == 2 && w0 < v) {
      // one hop: v's entry (w0, set by the init kernel) lowered to w0's
      // parent, an ancestor of w0 and so a vertex of v's component below v
      const uint32_t q = p[w0];
      if (q < w0) (void)__hip_atomic_fetch_min(&p[v], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (w1 < n && w1 != v && (!Skip1 || up0)) rem_unite<Splice>(v, w1, p);
    if (Skip1 && up0 && minima_entries > 2) {
      const uint64_t b = row_ptr[v] + 2;
      const uint64_t e = row_ptr[v + 1];
      const uint64_t stop = b + (minima_entries - 2) < e ? b + (minima_entries - 2) : e;
      for (uint64_t i = b; i < stop; ++i) {
        const uint32_t w = col[i];
        if (w < n && w != v) rem_unite<Splice>(v, w, p);
      }
    }
  }
}

// The one-pass hooks straight from the CSR, with no sample arrays: p starts
// as the identity (k_init), and each thread reads the first two entries of
// its row itself.  Neighbour 0 below v is linked by a CAS of v's own entry
// from v to it, which succeeds while v is still a root; that is the link
// the sampling init pass makes with a plain store when nothing else runs.
// (Linking a root v under any w0 < v is a union: w0 cannot be in v's tree,
// whose members are all >= v.)  If another thread hooked v first, the CAS
// fails and v is united with neighbour 0 by the walks like neighbour 1.
// Against the init pass + hook from the sample arrays this saves writing
// and reading the arrays (8 bytes per vertex each way) and lets the row
// reads of one vertex overlap the walks of another.
__device__ __forceinline__ void csr_first_two(const uint64_t *__restrict__ row_ptr,
                                              const uint32_t *__restrict__ col, uint64_t v,
                                              uint32_t n, uint32_t &w0, uint32_t &w1) {
  const uint64_t b = v < n ? row_ptr[v] : 0;
  const uint64_t e = v < n ? row_ptr[v + 1] : 0;
  w0 = b < e ? col[b] : kNoNeighbour;
  w1 = b + 1 < e ? col[b + 1] : kNoNeighbour;
}

// True if neighbour 0 still needs a union with v (the CAS link failed, or
// neighbour 0 is above v).
__device__ __forceinline__ bool link_first(uint32_t v, uint32_t w0, uint32_t n,
                                           uint32_t *__restrict__ p) {
  if (w0 >= n || w0 == v) return false;
  if (w0 > v) return true;
  return atomicCAS(&p[v], v, w0) != v;
}

//
// Skip1: a vertex whose own-entry link to neighbour 0 succeeds (Skip1 == 2:
// a link under neighbour 0's parent, one hop) skips neighbour 1; a local
// minimum (neighbour 0 above it) or a vertex whose link failed unites with
// both, and a local minimum also with entries 2 .. minima_entries-1 of its
// row (the reasoning at k_hook_rem).
//
// The launch covers the vertices [v_lo, v_hi) (v_hi clipped to n).  With
// Map, a low-id map (bits; the low-first hook, k_low_bits): a vertex
// whose neighbour 0 lies in the map's range [0, nbits) with its bit set,
// i.e. under the giant root g0 = g0p[0] sampled among the low ids, links
// straight under g0 (g0 <= neighbour 0 < v, as a root is its tree's
// minimum), reading the map, which stays in the L2, instead of neighbour
// 0's parent, a line of the parent array.
template <int Mode, int ILP, bool Halve, int Skip1 = 0, bool Map = false>
__global__ __launch_bounds__(kBlock) void k_hook_pair_csr(const uint64_t *__restrict__ row_ptr,
                                                          const uint32_t *__restrict__ col,
                                                          uint32_t n, uint32_t *__restrict__ p,
                                                          uint32_t minima_entries = 2,
                                                          uint32_t v_lo = 0, uint32_t v_hi = 0xFFFFFFFFu,
                                                          const uint32_t *__restrict__ bits = nullptr,
                                                          uint32_t nbits = 0,
                                                          const uint32_t *__restrict__ g0p = nullptr) {
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  const uint64_t hi = v_hi < n ? v_hi : n;
  const uint32_t g0 = Map ? g0p[0] : 0u;
  for (uint64_t v0 = (uint64_t)v_lo + blockIdx.x * kBlock + threadIdx.x; v0 < hi; v0 += S * ILP) {
    uint32_t w0[ILP], w1[ILP];
    uint64_t rb[ILP], re[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      rb[k] = v < hi ? row_ptr[v] : 0;
      re[k] = v < hi ? row_ptr[v + 1] : 0;
      w0[k] = rb[k] < re[k] ? col[rb[k]] : kNoNeighbour;
      // read with neighbour 0 even where skipping will not use it: a read
      // after the link's CAS, for the vertices that do, measured slower
      // (Chung-Lu 65.6M hook 2.11 -> 2.16 ms), and reading it only for the
      // local minima costs the frontier sweep more than it saves
      w1[k] = rb[k] + 1 < re[k] ? col[rb[k] + 1] : kNoNeighbour;
    }
    uint32_t x[3 * ILP];
    bool act[3 * ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      const bool live = v < hi;
      bool via_g0 = false;
      if (Skip1 == 2 && live && w0[k] < (uint32_t)v) {
        // one hop: link v under neighbour 0's parent (an ancestor of
        // neighbour 0, below v) by the CAS of its own entry; under g0 where
        // the low-id map says neighbour 0 hangs there
        uint32_t q;
        if (Map && w0[k] < nbits && ((bits[w0[k] >> 5] >> (w0[k] & 31)) & 1u)) {
          q = g0;
          via_g0 = true;
        } else {
          q = p[w0[k]];
          q = q < w0[k] ? q : w0[k];
        }
        act[ILP + k] = atomicCAS(&p[v], (uint32_t)v, q) != (uint32_t)v;
      } else {
        act[ILP + k] = live && link_first((uint32_t)v, w0[k], n, p);
      }
      // A link that failed (another union took v's entry first) unites v
      // with both neighbours, walking from g0 where the map marks neighbour
      // 0 (the same tree, a short walk).  Uniting such a vertex with
      // neighbour 0 alone (which the sweep from entry 1 would allow) left
      // more trees apart: Chung-Lu 65.6M 2.27 -> 2.31 ms per solve, R-MAT
      // 27 3.80 -> 4.17 ms, its frontier sweep 582 -> 1,076 us.
      act[2 * ILP + k] = live && w1[k] < n && w1[k] != (uint32_t)v && (!Skip1 || act[ILP + k]);
      act[k] = act[ILP + k] || act[2 * ILP + k];
      x[k] = (uint32_t)v;
      x[ILP + k] = via_g0 ? g0 : w0[k];
      x[2 * ILP + k] = w1[k];
    }
    if (Halve) roots_lockstep_halve<3 * ILP>(x, act, p);
    else roots_lockstep<3 * ILP>(x, act, p);
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      uint32_t rv = x[k];
      const uint32_t r0 = x[ILP + k], r1 = x[2 * ILP + k];
      if (act[ILP + k] && rv != r0) {
        hook_roots<Mode>(rv, r0, p);
        rv = rv < r0 ? rv : r0;  // a member of the merged tree
      }
      if (act[2 * ILP + k] && rv != r1) hook_roots<Mode>(rv, r1, p);
    }
    if (Skip1 && minima_entries > 2) {
#pragma unroll
      for (int k = 0; k < ILP; ++k) {
        const uint64_t v = v0 + k * S;
        if (v >= hi || w0[k] >= n || w0[k] <= (uint32_t)v) continue;
        const uint64_t b = rb[k] + 2;
        const uint64_t stop = b + (minima_entries - 2) < re[k] ? b + (minima_entries - 2) : re[k];
        for (uint64_t i = b; i < stop; ++i) {
          const uint32_t w = col[i];
          if (w < n && w != (uint32_t)v) (void)hook<Mode>((uint32_t)v, w, p);
        }
      }
    }
  }
}

// The unions a vertex of the one-hop skipping hook still owes after its
// own-entry link: with a (neighbour 0, or g0 where the map marks it) and
// with b (neighbour 1, or kNoNeighbour), the three roots walked in lockstep
// with halving, exactly as k_hook_pair_csr does for such a vertex.
__device__ __forceinline__ void unite_owed(uint32_t v, uint32_t a, uint32_t b, uint32_t n,
                                           uint32_t *__restrict__ p) {
  uint32_t x[3] = {v, a, b};
  const bool act[3] = {true, true, b < n && b != v};
  roots_lockstep_halve<3>(x, act, p);
  uint32_t rv = x[0];
  if (rv != x[1]) {
    hook_roots<4>(rv, x[1], p);
    rv = rv < x[1] ? rv : x[1];
  }
  if (act[2] && rv != x[2]) hook_roots<4>(rv, x[2], p);
}

// Queued form of k_hook_pair_csr<4, 1, true, 2, Map> (minima_entries 2).
// Each lane does the own-entry link (row offsets, the row's first two
// entries, the map bit or neighbour 0's parent, the CAS); a vertex that
// still owes unions (a local minimum, or a link that failed) goes to its
// wave's queue in LDS, and whenever 64 are queued the wave's lanes run 64
// of them at once.  In one pass, a wave waits for its slowest lane at
// every step, and nearly every step holds a walker.  The queue needs no
// atomics: waves cover 64 consecutive ids, so the loop, the ballots and
// the queue length are wave-uniform, and a wave's LDS operations complete
// in order (the fences below make the compiler wait for them).
constexpr int kQueue = 2 * kWave;
template <bool Map>
__global__ __launch_bounds__(kBlock) void k_hook_queued(const uint64_t *__restrict__ row_ptr,
                                                        const uint32_t *__restrict__ col,
                                                        uint32_t n, uint32_t *__restrict__ p,
                                                        uint32_t v_lo, uint32_t v_hi,
                                                        const uint32_t *__restrict__ bits,
                                                        uint32_t nbits,
                                                        const uint32_t *__restrict__ g0p) {
  __shared__ uint32_t qv[kBlock / kWave][kQueue];
  __shared__ uint32_t qa[kBlock / kWave][kQueue];
  __shared__ uint32_t qb[kBlock / kWave][kQueue];
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  const uint64_t hi = v_hi < n ? v_hi : n;
  const uint32_t g0 = Map ? g0p[0] : 0u;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t w = threadIdx.x / kWave;
  uint32_t queued = 0;  // wave-uniform, below kWave between steps
  for (uint64_t base = (uint64_t)v_lo + blockIdx.x * kBlock + (threadIdx.x - lane); base < hi;
       base += S) {
    const uint64_t v = base + lane;
    const bool live = v < hi;
    const uint64_t rb = live ? row_ptr[v] : 0;
    const uint64_t re = live ? row_ptr[v + 1] : 0;
    const uint32_t w0 = rb < re ? col[rb] : kNoNeighbour;
    const uint32_t w1 = rb + 1 < re ? col[rb + 1] : kNoNeighbour;
    bool owe = false;
    uint32_t a = w0;
    if (live && w0 < n && w0 != (uint32_t)v) {
      if (w0 < (uint32_t)v) {
        uint32_t q;
        bool marked = false;
        if (Map && w0 < nbits && ((bits[w0 >> 5] >> (w0 & 31)) & 1u)) {
          q = g0;
          marked = true;
        } else {
          q = p[w0];
          q = q < w0 ? q : w0;
        }
        if (atomicCAS(&p[v], (uint32_t)v, q) != (uint32_t)v) {
          owe = true;
          if (marked) a = g0;
        }
      } else {
        owe = true;  // a local minimum
      }
    }
    const uint64_t mask = __ballot(owe);
    if (mask) {
      if (owe) {
        const uint32_t at = queued + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        qv[w][at] = (uint32_t)v;
        qa[w][at] = a;
        qb[w][at] = w1;
      }
      queued += (uint32_t)__popcll(mask);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (queued >= kWave) {
        // 64 owed unions, one per lane; the rest moves to the front
        const uint32_t tv = qv[w][lane], ta = qa[w][lane], tb = qb[w][lane];
        queued -= kWave;
        if (lane < queued) {
          qv[w][lane] = qv[w][kWave + lane];
          qa[w][lane] = qa[w][kWave + lane];
          qb[w][lane] = qb[w][kWave + lane];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        unite_owed(tv, ta, tb, n, p);
      }
    }
  }
  if (lane < queued) unite_owed(qv[w][lane], qa[w][lane], qb[w][lane], n, p);
}

template <int Splice>
__global__ __launch_bounds__(kBlock) void k_hook_rem_csr(const uint64_t *__restrict__ row_ptr,
                                                         const uint32_t *__restrict__ col,
                                                         uint32_t n, uint32_t *__restrict__ p) {
  CCX_GRID_STRIDE(v, n) {
    uint32_t w0, w1;
    csr_first_two(row_ptr, col, v, n, w0, w1);
    if (link_first((uint32_t)v, w0, n, p)) rem_unite<Splice>((uint32_t)v, w0, p);
    if (w1 < n && w1 != (uint32_t)v) rem_unite<Splice>((uint32_t)v, w1, p);
  }
}

// The low-first hook's map over the low ids [0, L): bit u is set iff u's
// root is g0 = g0p[0], the giant root sampled among [0, L) after their
// hook.  Each entry of [0, L) is also pointed at its root, so later walks
// through the low ids take one hop; plain stores suffice, since nothing
// else changes the forest during this pass and a concurrent walker reads
// either value, both ancestors.  Every wave covers 64 consecutive ids
// (256-thread workgroups, grid stride a multiple of 256), so one ballot
// is one 64-bit word of the map.
__global__ __launch_bounds__(kBlock) void k_low_bits(uint32_t L, uint32_t *__restrict__ p,
                                                     const uint32_t *__restrict__ g0p,
                                                     uint64_t *__restrict__ bits) {
  const uint32_t g0 = g0p[0];
  const uint32_t whole = (L + 63u) & ~63u;
  for (uint32_t u = blockIdx.x * kBlock + threadIdx.x; u < whole; u += gridDim.x * kBlock) {
    bool in = false;
    if (u < L) {
      const uint32_t pu = p[u];
      uint32_t r = pu;
      if (pu != u) {
        r = find_root(pu, p);
        if (r != pu) p[u] = r;
      }
      in = r == g0;
    }
    const uint64_t b = __ballot(in);
    if ((threadIdx.x & (kWave - 1)) == 0) bits[u >> 6] = b;
  }
}

__device__ __forceinline__ uint32_t sample_vertex(uint32_t i, uint32_t n) {
  // SplitMix64 finaliser of the sample index, reduced to [0, n).
  uint64_t z = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)(((z >> 32) * (uint64_t)n) >> 32);
}

// One workgroup: the most frequent root among kGiantSamples sampled
// vertices, counted in an LDS open-addressing table; ties go to the smaller
// root so the choice is a function of the forest alone.  One thread per
// sample, so the root walks (chains of dependent gathers) all run at once.
__global__ __launch_bounds__(kGiantSamples) void k_sample_giant(uint32_t n, const uint32_t *__restrict__ p,
                                                         uint32_t *__restrict__ giant,
                                                         const uint32_t *__restrict__ off = nullptr) {
  constexpr uint32_t kSlots = 2 * kGiantSamples;
  constexpr uint32_t kEmpty = 0xFFFFFFFFu;
  __shared__ uint32_t keys[kSlots];
  __shared__ uint32_t cnts[kSlots];
  __shared__ uint64_t best[kGiantSamples / kWave];
  if (off && off[0]) return;
  const int t = threadIdx.x;
  for (uint32_t i = t; i < kSlots; i += kGiantSamples) {
    keys[i] = kEmpty;
    cnts[i] = 0;
  }
  __syncthreads();
  {
    const uint32_t x = find_root(sample_vertex(t, n), p);
    uint32_t h = (x * 0x9E3779B1u) >> 21;  // 11-bit slot
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
  // (count, ~root) packed so the max picks the largest count, then the
  // smallest root.
  uint64_t mine = 0;
  for (uint32_t i = t; i < kSlots; i += kGiantSamples)
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
    for (int w = 1; w < kGiantSamples / kWave; ++w)
      if (best[w] > b) b = best[w];
    giant[0] = ~(uint32_t)b;
    giant[1] = 0;  // the frontier sweep's link flag (see k_hook_frontier)
    giant[2] = 0;  // the plan switch: the sampled path took this solve
  }
}

// Frontier sweep over the vertices not under the giant root g.  Each
// vertex first resolves its root (and writes it back to its own entry,
// which does the pointer jumping the sampled rounds left); a walk that
// passes through g marks the vertex as a giant member and skips it.
// slots[0] = g; slots[1] (cleared by k_sample_giant) is set if any hook
// here linked a root: only then can a written-back root be stale, so the
// last pointer-jumping pass is gated on it.
// Chunks: each wave takes Chunks consecutive chunks of 64 vertices per
// step and walks all their roots in lockstep (Chunks dependent chains per
// lane in flight: on a graph whose sampled forest already holds nearly
// every vertex under g, the sweep is nothing but these walks); then each
// chunk with an active lane sweeps its lanes' remaining neighbours, and a
// chunk with none skips the prefix sums.
template <int Mode, int Chunks>
__global__ __launch_bounds__(kBlock) void k_hook_frontier(const uint64_t *__restrict__ row_ptr,
                                                          const uint32_t *__restrict__ col,
                                                          uint32_t n, uint32_t k_first,
                                                          uint32_t *__restrict__ p,
                                                          uint32_t *__restrict__ slots,
                                                          uint32_t every_vertex = 0,
                                                          const uint32_t *__restrict__ off = nullptr) {
  __shared__ uint64_t s_excl[kBlock / kWave][kWave];
  __shared__ uint64_t s_begin[kBlock / kWave][kWave];
  if (off && off[0]) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const uint32_t giant = slots[0];
  bool linked = false;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t c0 = ((uint64_t)blockIdx.x * (kBlock / kWave) + wid) * Chunks; c0 * kWave < n;
       c0 += waves * Chunks) {
    // root walks of every chunk of this step, in lockstep
    uint32_t cur[Chunks], nxt[Chunks], pv[Chunks];
    bool hit[Chunks];
#pragma unroll
    for (int j = 0; j < Chunks; ++j) {
      const uint64_t v = (c0 + j) * kWave + lane;
      cur[j] = (uint32_t)v;
      pv[j] = v < n ? p[v] : (uint32_t)v;
      nxt[j] = pv[j];
      hit[j] = (v == giant);
    }
    bool more = true;
    while (more) {
      more = false;
#pragma unroll
      for (int j = 0; j < Chunks; ++j) {
        if (nxt[j] == cur[j]) continue;
        cur[j] = nxt[j];
        hit[j] |= (cur[j] == giant);
        nxt[j] = p[cur[j]];
        more = true;
      }
    }
#pragma unroll
    for (int j = 0; j < Chunks; ++j) {
      const uint64_t base = (c0 + j) * kWave;
      const uint64_t v = base + lane;
      if (base >= n) break;
      uint64_t work = 0, begin = 0;
      if (v < n) {
        if (cur[j] != pv[j]) p[v] = cur[j];
        if (!hit[j] || every_vertex) {
          const uint64_t b = row_ptr[v] + k_first;
          const uint64_t e = row_ptr[v + 1];
          if (e > b) {
            work = e - b;
            begin = b;
          }
        }
      }
      if (__ballot(work != 0) == 0) continue;  // wave-uniform
      // Wave-inclusive prefix sum of the remaining degrees.
      uint64_t incl = work;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t up = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += up;
      }
      const uint64_t total = __shfl(incl, kWave - 1, kWave);
      s_excl[wid][lane] = incl - work;
      s_begin[wid][lane] = begin;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint64_t *ex = s_excl[wid];
      for (uint64_t e = lane; e < total; e += kWave) {
        int jj = 0;
#pragma unroll
        for (int step = kWave / 2; step > 0; step >>= 1)
          if (ex[jj + step] <= e) jj += step;
        const uint64_t idx = s_begin[wid][jj] + (e - ex[jj]);
        const uint32_t w = col[idx];
        if (w < n && hook<(Mode == 3 ? 1 : Mode)>((uint32_t)(base + jj), w, p)) linked = true;
      }
      // the LDS rows are rewritten by the next chunk of this wave
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  if (linked) slots[1] = 1;
}

inline unsigned blocks_for(uint64_t n, int cu_count, int per_cu) {
  const uint64_t need = (n + kBlock - 1) / kBlock;
  const uint64_t cap = (uint64_t)(cu_count > 0 ? cu_count : 256) * (uint64_t)per_cu;
  return (unsigned)(need < cap ? need : cap);
}

// ---- Vertex tiles: the plan for graphs whose ids are local ------------------
//
// On a graph whose edges mostly join vertices with nearby ids (a lattice or
// mesh numbered row by row, a road network in a space-filling-curve order),
// the sampled hook builds long parent chains (a row of consecutive ids links
// one vertex under the next) and every union walks them: an 8192^2 lattice
// with 30% of its edges deleted, numbered row by row, took 14.6 ms per solve
// that way against 5.2 ms with its ids permuted; this plan takes 1.20 ms:
//   1. k_tile_local: one workgroup per tile of 2^shift consecutive ids
//      streams the tile's rows (one contiguous slice of col_idx) and unites
//      every entry that stays inside the tile in an LDS array of 16-bit
//      local parents (a
//      vertex still a root links straight under its smaller neighbour, the
//      rest by Rem's unions).
//      Entries that leave the tile go to the tile's segment of a pair list
//      (with a symmetric CSR only those to smaller ids: the other direction
//      is seen from the other end).  Then each vertex's entry is set to its
//      local root (the smallest id of its component inside the tile) and the
//      local roots are marked in a bitmap.
//   2. k_tile_cross: one workgroup per tile unites each pair's two trees
//      (rem_unite<1> on the parent array), once per distinct pair of parents: the
//      first thread to enter a pair in an LDS table hooks it and the others
//      skip it (on a lattice nearly every pair between two tiles joins the
//      same two local roots).  A tile whose pairs did not fit its segment
//      reads its rows again instead.
//   3. k_tile_roots: each local root that was linked takes its root, with
//      path halving by atomicMin, so no store can raise an entry and every
//      local root ends holding exactly its root.
//   4. k_tile_final: every other vertex takes its local root's entry (one
//      hop): after step 1 only local roots are ever parents or change.
// Every link goes from a larger root to a smaller one, in LDS and in HBM, so
// each root is its component's minimum, as on the sampled path.
// When the plan is chosen automatically, k_tile_probe decides on the device
// from the rows of 1024 sampled vertices: tiles when at most 35% of their
// entries leave the vertex's tile.  Its answer is a switch word that
// every kernel of both plans reads; both plans are launched and the kernels
// of the one not taken return at once, so the solve never waits on the host.

constexpr int kTileBlock = 1024;
constexpr uint32_t kTileMaxShift = 16;  // 65,536 vertices: 16-bit local ids, 128 KiB of LDS parents
constexpr uint32_t kTileMinShift = 6;   // a tile covers whole 64-bit words of the root bitmap
// Tiles up to 2^15 vertices take a 64 KiB LDS parent array (two workgroups
// per CU), larger ones 128 KiB (one).  The default is the largest tile: on
// the 8192^2 lattice numbered row by row a 2^16-vertex tile holds eight
// rows and rarely falls apart, while narrower tiles split into many local
// components whose unions between tiles cost more than the extra
// occupancy gains (1.20 ms per solve at 2^16, 1.58 at 2^15, 3.2 at 2^14).
constexpr uint32_t kTileSmallShift = 15;
constexpr uint32_t kTileDefaultShift = 16;
constexpr uint32_t kTileMaxPairs = 1u << 24;
constexpr int kTileProbeSamples = 1024;
constexpr int kTileProbeEntries = 8;  // entries read per sampled row
// The plan runs when at most 35% of the sampled entries leave their tile.
// Measured on 2^24-vertex graphs (sampled plan against tiles, per solve):
// a 256^3 lattice numbered row by row (a third of its entries cross tiles)
// 2.30 against 0.59 ms; random banded graphs of degree 8 whose edges span
// up to 32K / 48K / 64K ids (about 25 / 37 / 50% crossing) 1.13 / 1.11 /
// 1.10 against 1.25 / 1.41 / 1.58 ms; degree 4 and 32K ids 5.12 against
// 0.78 ms.
constexpr uint32_t kTileProbeNum = 7;
constexpr uint32_t kTileProbeDen = 20;
constexpr uint32_t kTileSeenBits = 12;  // pairs entered per tile in the cross pass
constexpr uint32_t kTileSeen = 1u << kTileSeenBits;
constexpr unsigned long long kTileNoPair = ~0ull;  // parents are below 2^32 - 1

// The LDS parents of a tile as 16-bit local ids, two to a 32-bit word:
// vertex i is half i >> (Shift - 1) of word i mod 2^(Shift - 1), so
// consecutive ids, which neighbouring lanes link at the same time, sit in
// different words.  Loads and stores are 16-bit LDS accesses; a CAS goes
// through a 32-bit CAS of the word, retried while only the other half
// changed.
template <uint32_t Shift>
struct TilePar {
  static constexpr uint32_t kHalf = 1u << (Shift - 1);
  uint32_t *w;
  __device__ __forceinline__ uint16_t *at(uint32_t i) const {
    return reinterpret_cast<uint16_t *>(w) + (((i & (kHalf - 1)) << 1) | (i >> (Shift - 1)));
  }
  __device__ __forceinline__ uint32_t ld(uint32_t i) const {
    return __hip_atomic_load(at(i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ void st(uint32_t i, uint32_t v) const {
    __hip_atomic_store(at(i), (uint16_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // Replaces i's parent by v if it is `expect`; returns the parent found.
  __device__ __forceinline__ uint32_t cas(uint32_t i, uint32_t expect, uint32_t v) const {
    uint32_t *word = &w[i & (kHalf - 1)];
    const uint32_t sh = (i >> (Shift - 1)) << 4;
    uint32_t old = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
      const uint32_t cur = (old >> sh) & 0xFFFFu;
      if (cur != expect) return cur;
      const uint32_t got = atomicCAS(word, old, (old & ~(0xFFFFu << sh)) | (v << sh));
      if (got == old) return expect;
      old = got;
    }
  }
};

// Root of x among the LDS parents, halving the path on the way.  Only roots
// are ever linked (by CAS), so a halving store, which gives a non-root its
// grandparent, cannot undo a link.
template <class P>
__device__ __forceinline__ uint32_t tile_find(const P &par, uint32_t x) {
  for (;;) {
    const uint32_t px = par.ld(x);
    if (px == x) return x;
    const uint32_t ppx = par.ld(px);
    if (ppx == px) return px;
    par.st(x, ppx);
    x = ppx;
  }
}

// Union in LDS: both roots found with path halving (plain 16-bit stores),
// then the larger linked under the smaller by a CAS that holds only while it
// is still a root; otherwise the walk resumes from it.  (Rem's unions, which
// lower an entry at every step, cost a 32-bit CAS loop per step on 16-bit
// entries: the local pass took 1,305 us with them on the 8192^2 lattice
// numbered row by row, 2^15-vertex tiles.)
template <class P>
__device__ __forceinline__ void tile_unite(const P &par, uint32_t a, uint32_t b) {
  for (;;) {
    a = tile_find(par, a);
    b = tile_find(par, b);
    if (a == b) return;
    const uint32_t hi = a > b ? a : b;
    const uint32_t lo = a > b ? b : a;
    if (par.cas(hi, hi, lo) == hi) return;
    a = hi;
    b = lo;
  }
}

// f(vl, w) for every entry w of the tile's rows (vl = the row's local
// index), one row per lane: the lanes of a wave take 64 consecutive rows,
// each thread kTileRows rows per step (vl, vl + kTileBlock, ...), and all of
// the step's row offsets, then the first kTileRowEntries entries of each of
// its rows, are loaded before f runs on any of them; a row's end offset is
// the next lane's start offset.  Longer rows continue in a loop, so f runs
// on diverged lanes (its ballots count the lanes that run it).  (Loading a
// step ahead of f measured slower: the unions, not the loads, bound the
// pass.)
constexpr int kTileRows = 4;
constexpr int kTileRowEntries = 4;
template <class F>
__device__ __forceinline__ void tile_rows_lanes(const uint64_t *__restrict__ row_ptr,
                                                const uint32_t *__restrict__ col, uint32_t t0,
                                                uint32_t tn, F &&f) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t wbase = (threadIdx.x / kWave) * kWave;
  for (uint32_t base = wbase; base < tn; base += kTileRows * kTileBlock) {  // wave-uniform
    uint64_t b[kTileRows], e[kTileRows];
#pragma unroll
    for (int r = 0; r < kTileRows; ++r) {
      const uint32_t i = base + r * kTileBlock + lane;
      b[r] = row_ptr[t0 + (i < tn ? i : tn)];
      e[r] = lane == kWave - 1 ? row_ptr[t0 + (i + 1 < tn ? i + 1 : tn)] : 0;
    }
#pragma unroll
    for (int r = 0; r < kTileRows; ++r) {
      const uint64_t nx = __shfl_down(b[r], 1, kWave);
      if (lane != kWave - 1) e[r] = nx;
    }
    uint32_t w[kTileRows][kTileRowEntries];
#pragma unroll
    for (int r = 0; r < kTileRows; ++r)
#pragma unroll
      for (int k = 0; k < kTileRowEntries; ++k)
        w[r][k] = b[r] + k < e[r] ? col[b[r] + k] : 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < kTileRows; ++r) {
      const uint32_t vl = base + r * kTileBlock + lane;
#pragma unroll
      for (int k = 0; k < kTileRowEntries; ++k) f(b[r] + k < e[r], vl, w[r][k]);
      for (uint64_t j = b[r] + kTileRowEntries; j < e[r]; ++j) f(true, vl, col[j]);
    }
  }
}

// Step 1.  Sym: the CSR is symmetric (Knobs::symmetric).  LdsShift: the LDS
// parent array's size, 2^LdsShift 16-bit entries (shift <= LdsShift).  counts[tile] is the number of
// pairs the tile produced, also those past its segment's cap.
template <int Sym, uint32_t LdsShift>
__global__ __launch_bounds__(kTileBlock) void k_tile_local(
    const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col, uint32_t n,
    uint32_t shift, uint32_t *__restrict__ p, uint32_t *__restrict__ bits,
    uint32_t *__restrict__ counts, uint2 *__restrict__ pairs, uint32_t cap,
    const uint32_t *__restrict__ sw) {
  constexpr uint32_t kHalf = TilePar<LdsShift>::kHalf;
  __shared__ uint32_t words[kHalf];
  __shared__ uint32_t s_count;
  if (sw[0] != 1) return;
  const TilePar<LdsShift> par{words};
  const uint32_t t0 = blockIdx.x << shift;
  const uint32_t tn = min(n - t0, 1u << shift);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  for (uint32_t j = threadIdx.x; j < kHalf; j += kTileBlock) words[j] = j | ((j + kHalf) << 16);
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  uint2 *seg = pairs + (size_t)blockIdx.x * cap;
  tile_rows_lanes(row_ptr, col, t0, tn, [&](bool have, uint32_t vl, uint32_t w) {
              bool out = false;
              if (have && w < n) {
                const uint32_t wl = w - t0;  // inside the tile iff below tn
                if (wl < tn) {
                  // A root v links straight under a smaller neighbour's
                  // root (v's tree holds only ids >= v, so w is in another
                  // tree), which keeps v one hop from its root; its other
                  // entries, and a v already linked, take a union.
                  if (wl < vl) {
                    if (par.ld(vl) != vl || par.cas(vl, vl, tile_find(par, wl)) != vl)
                      tile_unite(par, vl, wl);
                  } else if (!Sym && wl != vl) {
                    tile_unite(par, vl, wl);
                  }
                } else {
                  out = !Sym || w < t0;
                }
              }
              const uint64_t bal = __ballot(out);
              if (bal == 0) return;
              const int leader = __builtin_ctzll(bal);
              uint32_t base = 0;
              if (lane == leader) base = atomicAdd(&s_count, (uint32_t)__popcll(bal));
              base = __shfl(base, leader, kWave);  // (the leader runs this call)
              if (out) {
                const uint32_t slot = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1));
                if (slot < cap) seg[slot] = make_uint2(t0 + vl, w);
              }
            });
  __syncthreads();
  for (uint32_t i0 = wid * kWave; i0 < tn; i0 += kTileBlock) {
    const uint32_t i = i0 + lane;
    bool root = false;
    if (i < tn) {
      const uint32_t r = tile_find(par, i);
      p[t0 + i] = t0 + r;
      root = r == i;
    }
    const uint64_t b = __ballot(root);
    const uint32_t word = (t0 + i0) >> 5;
    if (lane == 0) bits[word] = (uint32_t)b;
    if (lane == 1 && i0 + 32 < tn) bits[word + 1] = (uint32_t)(b >> 32);
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = s_count;
}

// Union of the trees of v and w (a pair of v's tile), unless another thread
// of this workgroup entered the same two parents in the table first: that
// thread unites them, so they are united by the time the kernel ends.  A
// slot taken by another pair is overwritten (its later repeats then unite
// again, which is only extra work).  The union is Rem's with splicing from
// the two parents, local roots: it climbs only local roots and writes only
// their entries, with local roots, so every other vertex keeps its local
// root as parent.
__device__ __forceinline__ void tile_cross_unite(uint32_t v, uint32_t w,
                                                 uint32_t *__restrict__ p,
                                                 unsigned long long *seen) {
  const uint32_t a = p[v];
  const uint32_t b = p[w];
  if (a == b) return;
  const unsigned long long key = ((unsigned long long)a << 32) | b;
  const uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - kTileSeenBits));
  const unsigned long long old = atomicCAS(&seen[h], kTileNoPair, key);
  if (old == key) return;
  if (old != kTileNoPair) atomicExch(&seen[h], key);
  rem_unite<1>(a, b, p);
}

// Step 2.
template <int Sym>
__global__ __launch_bounds__(kTileBlock) void k_tile_cross(
    const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col, uint32_t n,
    uint32_t shift, uint32_t *__restrict__ p, const uint32_t *__restrict__ counts,
    const uint2 *__restrict__ pairs, uint32_t cap, const uint32_t *__restrict__ sw) {
  __shared__ unsigned long long seen[kTileSeen];
  if (sw[0] != 1) return;
  for (uint32_t i = threadIdx.x; i < kTileSeen; i += kTileBlock) seen[i] = kTileNoPair;
  __syncthreads();
  const uint32_t t0 = blockIdx.x << shift;
  const uint32_t tn = min(n - t0, 1u << shift);
  const uint32_t cnt = counts[blockIdx.x];
  if (cnt <= cap) {
    const uint2 *seg = pairs + (size_t)blockIdx.x * cap;
    for (uint32_t i = threadIdx.x; i < cnt; i += kTileBlock) {
      const uint2 pr = seg[i];
      tile_cross_unite(pr.x, pr.y, p, seen);
    }
    return;
  }
  tile_rows_lanes(row_ptr, col, t0, tn, [&](bool have, uint32_t vl, uint32_t w) {
    if (have && w < n && w - t0 >= tn && (!Sym || w < t0)) tile_cross_unite(t0 + vl, w, p, seen);
  });
}

// Step 3: every local root (bit set) that was linked takes its root.  No
// links happen here, so the walk ends at the true root; the halving and the
// final store are atomicMin, which never raises an entry, so each local root
// ends holding exactly its root whatever other walks store through it.
__global__ __launch_bounds__(kBlock) void k_tile_roots(uint32_t n, uint32_t *__restrict__ p,
                                                       const uint32_t *__restrict__ bits,
                                                       const uint32_t *__restrict__ sw) {
  if (sw[0] != 1) return;
  const uint32_t words = (uint32_t)(((uint64_t)n + 31) >> 5);
  CCX_GRID_STRIDE(i, words) {
    uint32_t b = bits[i];
    while (b) {
      const uint32_t r = i * 32 + (uint32_t)__builtin_ctz(b);
      b &= b - 1;
      uint32_t x = p[r];
      if (x == r) continue;
      for (;;) {
        const uint32_t px = p[x];
        if (px == x) break;
        const uint32_t ppx = p[px];
        if (ppx == px) {
          x = px;
          break;
        }
        (void)__hip_atomic_fetch_min(&p[x], ppx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = ppx;
      }
      (void)__hip_atomic_fetch_min(&p[r], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Step 4, four vertices per thread per step: p[v] = p[p[v]].  A vertex's
// entry is its local root's (unchanged since step 1), whose entry is its
// root after step 3; local roots themselves already hold their root.
__global__ __launch_bounds__(kBlock) void k_tile_final(uint32_t n, uint32_t *__restrict__ p,
                                                       const uint32_t *__restrict__ sw) {
  constexpr int ILP = 4;
  if (sw[0] != 1) return;
  const uint64_t S = (uint64_t)gridDim.x * kBlock;
  for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < n; v0 += S * ILP) {
    uint32_t x[ILP], y[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      x[k] = v < n ? p[v] : (uint32_t)v;
    }
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = v0 + k * S;
      y[k] = (v < n && x[k] != (uint32_t)v) ? p[x[k]] : x[k];
    }
#pragma unroll
    for (int k = 0; k < ILP; ++k)
      if (y[k] != x[k]) p[v0 + k * S] = y[k];
  }
}

// One workgroup, one sampled vertex per thread: words[2] = 1 if the tile
// plan runs (force, or at most kTileProbeNum / kTileProbeDen of the sampled
// rows' first kTileProbeEntries entries leave their vertex's tile), else 0.
__global__ __launch_bounds__(kTileProbeSamples) void k_tile_probe(
    const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col, uint32_t n,
    uint32_t shift, int force, uint32_t *__restrict__ words) {
  __shared__ uint32_t s_tot, s_out;
  if (threadIdx.x == 0) {
    s_tot = 0;
    s_out = 0;
  }
  __syncthreads();
  if (!force) {
    const uint32_t v = sample_vertex(threadIdx.x, n);
    const uint64_t b = row_ptr[v];
    const uint64_t e = row_ptr[v + 1];
    uint32_t w[kTileProbeEntries];
#pragma unroll
    for (int k = 0; k < kTileProbeEntries; ++k) w[k] = b + k < e ? col[b + k] : 0xFFFFFFFFu;
    uint32_t tot = 0, out = 0;
#pragma unroll
    for (int k = 0; k < kTileProbeEntries; ++k)
      if (w[k] < n && w[k] != v) {
        ++tot;
        out += (w[k] >> shift) != (v >> shift);
      }
    atomicAdd(&s_tot, tot);
    atomicAdd(&s_out, out);
  }
  __syncthreads();
  if (threadIdx.x == 0)
    words[2] = (force || (s_tot > 0 && kTileProbeDen * s_out <= kTileProbeNum * s_tot)) ? 1u : 0u;
}

// Defaults measured on R-MAT scales 20-25, Chung-Lu 8.2M-65.6M vertices,
// ER 4M/32M (degree 8) and 2048^2 / 8192^2 grids (70% of edges kept):
//  * the frontier sweep and init/compress want 8 blocks per CU, one vertex
//    per thread;
//  * reading the sampled neighbours once into coalesced arrays (the cache)
//    never loses much and wins up to 1.6x on the sampled rounds of large
//    graphs, where each column gather is a cache line from HBM;
//  * hooking both sampled neighbours in one pass, with path halving in its
//    root walks once n >= 2^24, beats two passes (with pointer jumping
//    between them) on every graph; without halving the two-pass plan had
//    won on dense graphs beyond 2^30 entries (Chung-Lu 65.6M: 3.70 ms two
//    passes, 3.49 ms one halving pass);
//  * halving loses below 2^24 vertices (R-MAT 22: 235 vs 217 us), where the
//    parent array lives in the L2s and MALL and the extra atomics on hub
//    entries cost more than the shorter walks save;
//  * walking several roots per thread pays on dense graphs (m >= 16 n: 2
//    per thread in the one-pass hook with 8 blocks per CU; 4 per round on
//    the two-pass path), except past 2^31 entries, and loses on sparse ones
//    (1, with 4 blocks per CU);
//  * on the two-pass path, sampling the giant root after round 0 and
//    skipping its members in round 1 saves 5-6% on Chung-Lu 16.4M/65.6M.
constexpr uint64_t kDenseDegree = 16;              // dense: m >= 16 n
constexpr uint64_t kHalveMinVertices = 1ull << 24;  // halving from this many vertices
constexpr uint64_t kWideIlpMaxEntries = 1ull << 31; // dense one-pass ILP 2 below this many entries
constexpr uint64_t kSplitWalksMaxEntries = 1ull << 31;  // queued owed unions below this many entries
constexpr uint64_t kCsrMinVertices = 1ull << 23;       // dense graphs from here: the hook from the CSR
constexpr uint64_t kSkipMinDegree = 6;              // sparse graphs skip from m >= 6 n
constexpr uint64_t kTileMinVertices = 1ull << 20;   // the tile probe runs from this many vertices

struct Plan {
  int cache;     // sampled-neighbour cache in use
  int fused;     // one-pass hook of both sampled neighbours
  int rem;       // the one pass unites by Rem's algorithm (k_hook_rem)
  int splice;    // ... splicing: 0 none, 1 every step, 2 v's own entry only
  int halve;     // path halving in the lockstep one-pass hook (k_hook_pair)
  int hook_ilp;  // roots per thread in the lockstep sampled hook
  int hook_bpc;  // sampled-hook grid cap, workgroups per CU
  int nt;        // non-temporal streaming bits (Knobs::nt)
  int init_ilp;  // rows per thread per step of the sampling init pass
  int frontier_chunks;  // chunks per wave step of the frontier sweep
  int csr;       // one-pass hook from the CSR (k_hook_pair_csr / k_hook_rem_csr), no sample arrays
  int skip1;     // own-entry links skip neighbour 1; the frontier sweep starts at entry 1
  int minima_entries;  // ... and local minima unite with this many entries of their row
  int compress_ilp;    // vertices per thread per step of the pointer-jumping passes
  int tiles;           // vertex tiles: 0 no, 1 always, 2 if the probe finds local ids
  uint32_t tile_shift; // log2 of the vertices per tile
  uint32_t tile_cap;   // pairs per tile segment
  uint32_t low_shift;  // low-first hook over ids below n >> low_shift (0: one pass)
  int split_walks;     // its passes split into links and a walk list
};

inline Plan plan_for(uint32_t n, uint64_t m, const Knobs &K = g_knobs) {
  Plan pl;
  const bool dense = m >= kDenseDegree * (uint64_t)n;
  const bool big = n >= kHalveMinVertices;
  pl.cache = K.sample_cache && (K.hook_mode == 0 || K.hook_mode == 1 || K.hook_mode == 4);
  pl.fused = pl.cache && K.sampled_neighbours == 2 && (K.fused_pair < 0 ? 1 : K.fused_pair);
  // Dense graphs of 2^23+ vertices take the one-pass hook from the CSR
  // (one root per thread, path halving, below), at every size; the
  // lockstep hook from the sample arrays with two roots per thread, which
  // those of 2^24+ vertices below 2^31 entries took before the low-first
  // hook, is what `wide` still gives where knobs pick it, and those below
  // 2^24 vertices took Rem's unions.
  const bool csr_shape = pl.fused && dense && n >= kCsrMinVertices && K.hook_mode == 4;
  pl.rem = pl.fused && K.hook_mode != 0 && (K.rem < 0 ? (!big || !dense) && !csr_shape : K.rem);
  const int splice_auto = big ? 1 : (dense ? 2 : 0);
  pl.splice = pl.rem ? (K.rem_splice < 0 ? splice_auto : K.rem_splice) : 0;
  pl.halve = pl.fused && !pl.rem && K.hook_mode != 0 && (K.halve < 0 ? big || csr_shape : K.halve);
  const int wide = pl.fused ? (m < kWideIlpMaxEntries && !csr_shape ? 2 : 1) : 4;
  pl.hook_ilp = pl.rem ? 1 : (K.hook_ilp ? K.hook_ilp : (dense ? wide : 1));
  const int bpc_auto = pl.rem ? (pl.splice == 1 ? 4 : 8) : (pl.hook_ilp >= 2 ? 8 : 4);
  pl.hook_bpc = K.blocks_per_cu ? K.blocks_per_cu : bpc_auto;
  // From 2^24 vertices the parent array no longer fits the L2s, and the
  // sample arrays streamed through the MALL would evict it: both are
  // stored and read non-temporally (grid 8192^2 5.59 -> 5.15 ms per solve,
  // Chung-Lu 65.6M 3.41 -> 3.36 ms; below that they cost, R-MAT 22 hook
  // 91 -> 96 us).  Dense graphs of that size also take 4 rows per thread
  // in the init pass (Chung-Lu 65.6M 3.36 -> 3.28 ms; the grid gains
  // nothing).
  pl.nt = K.nt >= 0 ? K.nt : (big ? 5 : 0);
  pl.init_ilp = K.init_ilp ? K.init_ilp : (big && dense ? 4 : 1);
  // two 64-vertex chunks per wave step of the frontier sweep: Chung-Lu
  // 65.6M 195 -> 146 us, R-MAT 22 33 -> 29 us, R-MAT 24 80 -> 71 us, ER 32M
  // 103 -> 99 us; four lose on every graph (register pressure)
  pl.frontier_chunks = K.frontier_chunks ? K.frontier_chunks : 2;
  // The lockstep hook with one root per thread (dense graphs of 2^24+
  // vertices) reads the rows itself instead of sample arrays: Chung-Lu
  // 65.6M (3.60e9 entries) 3.30 -> 3.22 ms per solve, Chung-Lu 40M (2.31e9)
  // 2.04 -> 1.98 ms, Chung-Lu 65.6M at gamma 2.3 (3.82e9) 3.76 -> 3.60 ms.
  // With skipping, the low-first map and, below 2^31 entries, the queued
  // walks, it also beats the two-roots-per-thread hook from the sample
  // arrays that dense graphs below 2^31 entries took before: R-MAT 24
  // 0.685 -> 0.493 ms, R-MAT 25 1.06 -> 0.905, R-MAT 26 2.22 -> 1.785,
  // Chung-Lu 20M 1.50 -> 0.737, Chung-Lu 32M 2.22 -> 1.22.  The other plans
  // keep the sampling init pass.  On dense graphs the CSR plan also wins
  // from 2^23 vertices (R-MAT 23 0.313 -> 0.281 ms, Chung-Lu 8.2M 0.413 ->
  // 0.381, 12M 0.609 -> 0.513, 16.4M 0.80 -> 0.68), but it loses below
  // that and on sparse graphs (R-MAT 22 0.160 -> 0.198 ms, ER 32M degree 8
  // 1.76 -> 2.19, 8192^2 grid 5.20 -> 7.99), and Rem's unions from the CSR lose
  // badly (R-MAT 22 hook 93 -> 181 us, 8192^2 grid 3.76 -> 5.67 ms), since
  // the row reads sit in front of every latency-bound union.
  const int csr_auto = csr_shape && !pl.rem && pl.hook_ilp == 1;
  pl.csr = pl.fused && (pl.rem || K.hook_mode == 4) && (K.csr_hook < 0 ? csr_auto : K.csr_hook);
  // Skipping neighbour 1 of linked vertices (Knobs::skip_linked; the
  // lockstep hooks and Rem's unions from the sample arrays) trades the hook's
  // second walk per vertex for a frontier sweep over the few vertices the
  // local minima's links leave outside the giant tree:
  //  * the lockstep hook from the CSR, with one-hop links and 16 workgroups
  //    per CU: Chung-Lu 65.6M 3.21 -> 2.49 ms per solve, Chung-Lu 40M
  //    2.08 -> 1.61, R-MAT 27 (4.2e9 entries) 4.18 -> 3.92 (4.18 being the
  //    plan without skipping at 8-16 workgroups per CU; 5.68 at 4); without
  //    skipping, 16 workgroups per CU lose on Chung-Lu (3.21 -> 3.38);
  //  * Rem's unions on sparse graphs of at least 6 entries per vertex: ER
  //    4M (degree 8) 390 -> 241 us, ER 32M 2.24 -> 1.66 ms;
  //  * it loses where the minima's links do not join the trees: the 8192^2
  //    lattice (2.8 entries per vertex) 5.2 -> 8.8 ms, and R-MAT 22/24/25
  //    (163 -> 204 us, 667 -> 698, 1063 -> 1172), whose trees come out
  //    deeper, so the frontier sweep costs more than the hook saves.
  //    Chung-Lu 16.4M on the sample-array lockstep hook gains (818 -> 687
  //    us) but shares that plan with R-MAT 24/25, which keeps both unions.
  //  * Local minima uniting with more of their row (minima_entries 3, 4)
  //    shrink the frontier but stall their waves: Chung-Lu 65.6M 2.73 ->
  //    3.02 ms, ER 32M 1.66 -> 1.85 ms.
  const bool skip_ok = pl.fused && (pl.rem ? !pl.csr : K.hook_mode == 4);
  const bool skip_auto = (pl.csr && !pl.rem) || (pl.rem && !pl.csr && !dense && m >= kSkipMinDegree * (uint64_t)n);
  pl.skip1 = skip_ok ? (K.skip_linked < 0 ? (skip_auto ? 2 : 0) : K.skip_linked) : 0;
  pl.minima_entries = K.minima_entries ? K.minima_entries : 2;
  // 48 workgroups per CU for the skipping hook from the CSR (16 before the
  // low-first plan): Chung-Lu 65.6M 2.27 -> 2.23 ms per solve, Chung-Lu
  // 32M 1.22 -> 1.18, Chung-Lu 40M 1.48 -> 1.45, R-MAT 26 1.78 -> 1.72
  // (64: 2.22, 1.17, 1.45, 1.75); below 2^24 vertices 16 (R-MAT 23 0.287
  // -> 0.281 ms, Chung-Lu 12M 0.531 -> 0.513, Chung-Lu 8.2M 0.406 -> 0.381)
  if (!K.blocks_per_cu && pl.skip1 && pl.csr) pl.hook_bpc = big ? 48 : 16;
  // the last pointer-jumping pass runs after a skipping hook (the frontier
  // sweep links the vertices left outside the giant tree): 4 vertices per
  // thread, Chung-Lu 65.6M 87 -> 75 us
  pl.compress_ilp = K.stream_ilp > 1 ? K.stream_ilp : (pl.skip1 ? 4 : 1);
  // Vertex tiles (k_tile_local).  The probe runs where the sampled path is
  // the one whose kernels read the plan switch: the sparse plan uniting by
  // Rem's algorithm from the sample arrays.  Below 2^20 vertices a solve
  // is short whichever plan runs.
  const bool tile_shape = pl.cache && pl.fused && pl.rem && !pl.csr && K.hook_mode == 4 &&
                          K.sampled_neighbours == 2 && pl.init_ilp == 1;
  const bool tile_auto = tile_shape && !dense && n >= kTileMinVertices;
  pl.tiles = K.tiles > 0 ? 1 : (K.tiles == 0 ? 0 : (tile_auto ? 2 : 0));
  pl.tile_shift = K.tile_shift ? (uint32_t)K.tile_shift : kTileDefaultShift;
  pl.tile_cap = K.tile_pairs ? (uint32_t)K.tile_pairs : (1u << pl.tile_shift);
  const bool low_ok = pl.csr && !pl.rem && pl.halve && pl.skip1 == 2 && pl.hook_ilp == 1;
  pl.low_shift = low_ok ? (uint32_t)(K.low_shift < 0 ? kLowShiftAuto : K.low_shift) : 0u;
  pl.split_walks = pl.low_shift && pl.minima_entries == 2
                       ? (K.split_walks < 0 ? (m < kSplitWalksMaxEntries ? 1 : 0) : K.split_walks)
                       : 0;
  return pl;
}

// The low-first hook's bound: ids below it are hooked first and mapped.  A
// multiple of 64 (whole map words), at least 64, below n.
inline uint32_t low_bound(uint32_t n, const Plan &pl);
// Words of the low-id map, rounded up to 256 bytes.
inline uint64_t low_map_words(uint32_t n, const Plan &pl) {
  return ((2ull * (((uint64_t)low_bound(n, pl) + 63) / 64)) + 63) & ~63ull;
}
inline uint32_t low_bound(uint32_t n, const Plan &pl) {
  if (!pl.low_shift) return 0;
  uint32_t L = ((n >> pl.low_shift) + 63u) & ~63u;
  if (L < 64u) L = 64u;
  return L >= n ? 0u : L;
}

// Words of the tile plan's workspace: per-tile pair counts and the local
// root bitmap (rounded up to 256 bytes), then the pair segments.
inline uint64_t tile_head_words(uint32_t n, const Plan &pl) {
  const uint64_t nt = ((uint64_t)n + (1u << pl.tile_shift) - 1) >> pl.tile_shift;
  return (nt + ((uint64_t)n + 31) / 32 + 63) & ~63ull;
}
inline uint64_t tile_words(uint32_t n, const Plan &pl) {
  const uint64_t nt = ((uint64_t)n + (1u << pl.tile_shift) - 1) >> pl.tile_shift;
  return tile_head_words(n, pl) + 2 * nt * pl.tile_cap;
}

inline bool ilp_ok(int v) { return v == 1 || v == 2 || v == 4; }

}  // namespace

extern "C" int cc_hip_components_set(const char *key, int value) {
  if (!key) return ccx::set_error(4, "cc_hip_components_set: null key");
  const std::string k(key);
  if (k == "sampled_neighbours" && value >= 1 && value <= 4) g_knobs.sampled_neighbours = value;
  else if (k == "blocks_per_cu" && value >= 0 && value <= 64) g_knobs.blocks_per_cu = value;
  else if (k == "stream_blocks_per_cu" && value >= 1 && value <= 64) g_knobs.stream_blocks_per_cu = value;
  else if (k == "frontier_blocks_per_cu" && value >= 1 && value <= 64) g_knobs.frontier_blocks_per_cu = value;
  else if (k == "hook_mode" && value >= 0 && value <= 4) g_knobs.hook_mode = value;
  else if (k == "halve" && value >= -1 && value <= 1) g_knobs.halve = value;
  else if (k == "rem" && value >= -1 && value <= 1) g_knobs.rem = value;
  else if (k == "rem_splice" && value >= -1 && value <= 2) g_knobs.rem_splice = value;
  else if (k == "hook_ilp" && (value == 0 || ilp_ok(value))) g_knobs.hook_ilp = value;
  else if (k == "stream_ilp" && ilp_ok(value)) g_knobs.stream_ilp = value;
  else if (k == "sample_cache" && (value == 0 || value == 1)) g_knobs.sample_cache = value;
  else if (k == "fused_pair" && value >= -1 && value <= 1) g_knobs.fused_pair = value;
  else if (k == "giant_early" && (value == 0 || value == 1)) g_knobs.giant_early = value;
  else if (k == "nt" && value >= -1 && value <= 7) g_knobs.nt = value;
  else if (k == "init_ilp" && (value == 0 || ilp_ok(value))) g_knobs.init_ilp = value;
  else if (k == "frontier_chunks" && (value == 0 || ilp_ok(value))) g_knobs.frontier_chunks = value;
  else if (k == "csr_hook" && value >= -1 && value <= 1) g_knobs.csr_hook = value;
  else if (k == "skip_linked" && value >= -1 && value <= 2) g_knobs.skip_linked = value;
  else if (k == "minima_entries" && (value == 0 || (value >= 2 && value <= 8))) g_knobs.minima_entries = value;
  else if (k == "symmetric" && (value == 0 || value == 1)) g_knobs.symmetric = value;
  else if (k == "tiles" && value >= -1 && value <= 1) g_knobs.tiles = value;
  else if (k == "tile_shift" && (value == 0 || (value >= (int)kTileMinShift && value <= (int)kTileMaxShift)))
    g_knobs.tile_shift = value;
  else if (k == "tile_pairs" && value >= 0 && value <= (int)kTileMaxPairs) g_knobs.tile_pairs = value;
  else if (k == "low_shift" && value >= -1 && value <= 8) g_knobs.low_shift = value;
  else if (k == "split_walks" && value >= -1 && value <= 1) g_knobs.split_walks = value;
  else return ccx::set_error(4, "cc_hip_components_set: bad key or value (%s=%d)", key, value);
  return 0;
}

extern "C" int cc_hip_components_get(const char *key) {
  const std::string k(key ? key : "");
  if (k == "sampled_neighbours") return g_knobs.sampled_neighbours;
  if (k == "blocks_per_cu") return g_knobs.blocks_per_cu;
  if (k == "stream_blocks_per_cu") return g_knobs.stream_blocks_per_cu;
  if (k == "frontier_blocks_per_cu") return g_knobs.frontier_blocks_per_cu;
  if (k == "hook_mode") return g_knobs.hook_mode;
  if (k == "hook_ilp") return g_knobs.hook_ilp;
  if (k == "stream_ilp") return g_knobs.stream_ilp;
  if (k == "sample_cache") return g_knobs.sample_cache;
  if (k == "fused_pair") return g_knobs.fused_pair;
  if (k == "giant_early") return g_knobs.giant_early;
  if (k == "halve") return g_knobs.halve;
  if (k == "rem") return g_knobs.rem;
  if (k == "rem_splice") return g_knobs.rem_splice;
  if (k == "nt") return g_knobs.nt;
  if (k == "init_ilp") return g_knobs.init_ilp;
  if (k == "frontier_chunks") return g_knobs.frontier_chunks;
  if (k == "csr_hook") return g_knobs.csr_hook;
  if (k == "skip_linked") return g_knobs.skip_linked;
  if (k == "minima_entries") return g_knobs.minima_entries;
  if (k == "symmetric") return g_knobs.symmetric;
  if (k == "tiles") return g_knobs.tiles;
  if (k == "tile_shift") return g_knobs.tile_shift;
  if (k == "tile_pairs") return g_knobs.tile_pairs;
  if (k == "low_shift") return g_knobs.low_shift;
  if (k == "split_walks") return g_knobs.split_walks;
  return -1;
}

extern "C" int cc_hip_components_last_pass_ran(void *stream, int *ran) {
  if (!ccx::ctx()) return 2;
  if (!ran) return ccx::set_error(4, "cc_hip_components_last_pass_ran: null pointer");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  uint32_t flag = 0;
  CCX_TRY(hipMemcpyAsync(&flag, st->d_words + 1, sizeof(flag), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  *ran = flag != 0;
  return 0;
}

extern "C" int cc_hip_components_tiles_ran(void *stream, int *ran) {
  if (!ccx::ctx()) return 2;
  if (!ran) return ccx::set_error(4, "cc_hip_components_tiles_ran: null pointer");
  hipStream_t s = ccx::pick_stream(stream);
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  uint32_t flag = 0;
  CCX_TRY(hipMemcpyAsync(&flag, st->d_words + 2, sizeof(flag), hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  *ran = flag != 0;
  return 0;
}

extern "C" int cc_hip_components_plan_get(uint32_t n, uint64_t m, const char *key) {
  const Plan pl = plan_for(n, m);
  const std::string k(key ? key : "");
  if (k == "cache") return pl.cache;
  if (k == "fused") return pl.fused;
  if (k == "halve") return pl.halve;
  if (k == "rem") return pl.rem;
  if (k == "rem_splice") return pl.splice;
  if (k == "hook_ilp") return pl.hook_ilp;
  if (k == "nt") return pl.nt;
  if (k == "init_ilp") return pl.init_ilp;
  if (k == "frontier_chunks") return pl.frontier_chunks;
  if (k == "csr_hook") return pl.csr;
  if (k == "skip_linked") return pl.skip1;
  if (k == "minima_entries") return pl.minima_entries;
  if (k == "compress_ilp") return pl.compress_ilp;
  if (k == "hook_blocks_per_cu") return pl.hook_bpc;
  if (k == "giant_early") return pl.cache && !pl.fused && g_knobs.giant_early;
  if (k == "hook_mode") return g_knobs.hook_mode;
  if (k == "tiles") return pl.tiles;
  if (k == "tile_shift") return (int)pl.tile_shift;
  if (k == "tile_pairs") return (int)pl.tile_cap;
  if (k == "low_shift") return (int)pl.low_shift;
  if (k == "split_walks") return pl.split_walks;
  return -1;
}

extern "C" int cc_hip_components_plan(uint32_t n, uint64_t m, int *cache, int *fused,
                                      int *hook_ilp, int *hook_blocks_per_cu) {
  const Plan pl = plan_for(n, m);
  if (cache) *cache = pl.cache;
  if (fused) *fused = pl.fused;
  if (hook_ilp) *hook_ilp = pl.hook_ilp;
  if (hook_blocks_per_cu) *hook_blocks_per_cu = pl.hook_bpc;
  return 0;
}

extern "C" int cc_hip_components_tune(int sampled_neighbours, int blocks_per_cu, int hook_mode) {
  if (sampled_neighbours < 0 || sampled_neighbours > 4 || blocks_per_cu < 0 || blocks_per_cu > 64 ||
      hook_mode < -1 || hook_mode > 4)
    return ccx::set_error(4, "cc_hip_components_tune: out of range (%d, %d)", sampled_neighbours,
                          blocks_per_cu);
  if (sampled_neighbours > 0) g_knobs.sampled_neighbours = sampled_neighbours;
  if (blocks_per_cu > 0) g_knobs.blocks_per_cu = blocks_per_cu;
  if (hook_mode >= 0) g_knobs.hook_mode = hook_mode;
  return 0;
}

namespace {

// Launches of the one-pass hooks from the sample arrays, by template
// arguments chosen at run time.
template <int Splice, bool Nt>
int launch_hook_rem(int skip, unsigned nb, hipStream_t s, const uint32_t *s0, const uint32_t *s1,
                    uint32_t n, uint32_t *p, const uint64_t *row_ptr, const uint32_t *col,
                    uint32_t minima_entries, const uint32_t *off) {
  if (skip == 2)
    CCX_LAUNCH("cc_hook_sample", s, k_hook_rem<Splice, Nt, 2><<<nb, kBlock, 0, s>>>(s0, s1, n, p, row_ptr, col, minima_entries, off));
  else if (skip)
    CCX_LAUNCH("cc_hook_sample", s, k_hook_rem<Splice, Nt, 1><<<nb, kBlock, 0, s>>>(s0, s1, n, p, row_ptr, col, minima_entries, off));
  else
    CCX_LAUNCH("cc_hook_sample", s, k_hook_rem<Splice, Nt, 0><<<nb, kBlock, 0, s>>>(s0, s1, n, p, row_ptr, col, minima_entries, off));
  return 0;
}

template <int Mode, int ILP, bool Halve, bool Nt>
int launch_hook_pair(int skip, unsigned nb, hipStream_t s, const uint32_t *s0, const uint32_t *s1,
                     uint32_t n, uint32_t *p) {
  if constexpr (Mode == 4) {
    if (skip == 2) {
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair<Mode, ILP, Halve, Nt, 2><<<nb, kBlock, 0, s>>>(s0, s1, n, p));
      return 0;
    }
    if (skip) {
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair<Mode, ILP, Halve, Nt, 1><<<nb, kBlock, 0, s>>>(s0, s1, n, p));
      return 0;
    }
  }
  CCX_LAUNCH("cc_hook_sample", s, k_hook_pair<Mode, ILP, Halve, Nt><<<nb, kBlock, 0, s>>>(s0, s1, n, p));
  return 0;
}

template <int Mode, bool Halve, bool Nt>
int launch_hook_pair_ilp(int ilp, int skip, unsigned nb, hipStream_t s, const uint32_t *s0,
                         const uint32_t *s1, uint32_t n, uint32_t *p) {
  if (ilp == 4) return launch_hook_pair<Mode, 4, Halve, Nt>(skip, nb, s, s0, s1, n, p);
  if (ilp == 2) return launch_hook_pair<Mode, 2, Halve, Nt>(skip, nb, s, s0, s1, n, p);
  return launch_hook_pair<Mode, 1, Halve, Nt>(skip, nb, s, s0, s1, n, p);
}

// The whole solve on stream s with knobs K.  words[0..2]: the sampled giant
// root, the frontier sweep's link flag and the plan switch (each written
// before it is read); smp: rounds * n words for the sampled-neighbour cache
// when the plan uses it (written by the init pass before the hooks read
// it), or tile_words() for the tile plan, whichever is larger when both are
// launched (only one of them does work).  Allocates nothing itself.
int solve_with(const Knobs &K, const ccx::Context *c, const uint64_t *d_row_ptr,
               const uint32_t *d_col_idx, uint32_t n, uint64_t m, uint32_t *d_labels,
               hipStream_t s, uint32_t *words, uint32_t *smp) {
  uint32_t *giant = words;
  const Plan plan = plan_for(n, m, K);
  const int hook_ilp = plan.hook_ilp;
  const unsigned nb = blocks_for(n, c->cu_count, plan.hook_bpc);
  const unsigned nf = blocks_for(n, c->cu_count, K.frontier_blocks_per_cu);
  const unsigned ns = blocks_for(n, c->cu_count, K.stream_blocks_per_cu);
  const uint32_t rounds = (uint32_t)K.sampled_neighbours;
  const int ilp = K.stream_ilp;
  const uint32_t me = (uint32_t)plan.minima_entries;

  if (m == 0) {
    CCX_LAUNCH("cc_init", s, k_init<<<ns, kBlock, 0, s>>>(n, d_labels));
    return 0;
  }
  // slots: nullptr for a pass between hook rounds, the giant root and the
  // frontier's link flag for the last pass
  // The tile plan: launched first; with the probe deciding, the sampled
  // path's kernels below read the same switch word and return at once when
  // the tiles took the solve.
  uint32_t *sw = words + 2;
  const uint32_t *off = plan.tiles == 2 ? sw : nullptr;
  if (plan.tiles) {
    if (!smp) return ccx::set_error(4, "solve: no tile workspace");
    const uint32_t shift = plan.tile_shift;
    const uint32_t cap = plan.tile_cap;
    const uint32_t nt = (uint32_t)(((uint64_t)n + (1u << shift) - 1) >> shift);
    uint32_t *counts = smp;
    uint32_t *bits = smp + nt;
    uint2 *pairs = (uint2 *)(smp + tile_head_words(n, plan));
    const unsigned nw = blocks_for(((uint64_t)n + 31) / 32, c->cu_count, K.stream_blocks_per_cu);
    CCX_LAUNCH("cc_tile_probe", s, k_tile_probe<<<1, kTileProbeSamples, 0, s>>>(d_row_ptr, d_col_idx, n, shift, plan.tiles == 1, words));
#define CCX_TILE_LOCAL(SYM, LDS)                                                                   \
  CCX_LAUNCH("cc_tile_local", s, k_tile_local<SYM, LDS><<<nt, kTileBlock, 0, s>>>(d_row_ptr, d_col_idx, n, shift, d_labels, bits, counts, pairs, cap, sw))
    if (K.symmetric) {
      if (shift > kTileSmallShift) CCX_TILE_LOCAL(1, kTileMaxShift);
      else CCX_TILE_LOCAL(1, kTileSmallShift);
      CCX_LAUNCH("cc_tile_cross", s, k_tile_cross<1><<<nt, kTileBlock, 0, s>>>(d_row_ptr, d_col_idx, n, shift, d_labels, counts, pairs, cap, sw));
    } else {
      if (shift > kTileSmallShift) CCX_TILE_LOCAL(0, kTileMaxShift);
      else CCX_TILE_LOCAL(0, kTileSmallShift);
      CCX_LAUNCH("cc_tile_cross", s, k_tile_cross<0><<<nt, kTileBlock, 0, s>>>(d_row_ptr, d_col_idx, n, shift, d_labels, counts, pairs, cap, sw));
    }
#undef CCX_TILE_LOCAL
    CCX_LAUNCH("cc_tile_roots", s, k_tile_roots<<<nw, kBlock, 0, s>>>(n, d_labels, bits, sw));
    CCX_LAUNCH("cc_tile_final", s, k_tile_final<<<ns, kBlock, 0, s>>>(n, d_labels, sw));
    if (plan.tiles == 1) return 0;
  }
  auto compress = [&](const uint32_t *gate) -> int {
    const int cilp = plan.compress_ilp;
    if (cilp == 4) CCX_LAUNCH("cc_compress", s, k_compress_ilp<4><<<ns, kBlock, 0, s>>>(n, d_labels, gate, off));
    else if (cilp == 2) CCX_LAUNCH("cc_compress", s, k_compress_ilp<2><<<ns, kBlock, 0, s>>>(n, d_labels, gate, off));
    else CCX_LAUNCH("cc_compress", s, k_compress<<<ns, kBlock, 0, s>>>(n, d_labels, gate, off));
    return 0;
  };
  const int mode = K.hook_mode;
  // The cache holds `rounds` u32 per vertex in this stream's temp buffer;
  // modes 0, 1 and 4 read it (mode 2 walks with L2 reads, mode 3 is lossy).
  const bool cache = plan.cache != 0;
  const bool early_giant = cache && !plan.fused && K.giant_early;
  if (plan.csr) {
    // identity, then the one-pass hook reading the rows itself
    CCX_LAUNCH("cc_init", s, k_init<<<ns, kBlock, 0, s>>>(n, d_labels));
    if (plan.rem && plan.splice == 1)
      CCX_LAUNCH("cc_hook_sample", s, k_hook_rem_csr<1><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels));
    else if (plan.rem && plan.splice == 2)
      CCX_LAUNCH("cc_hook_sample", s, k_hook_rem_csr<2><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels));
    else if (plan.rem)
      CCX_LAUNCH("cc_hook_sample", s, k_hook_rem_csr<0><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels));
#define CCX_HOOK_CSR(H, K1)                                                                          \
  do {                                                                                             \
    if (hook_ilp == 4)                                                                             \
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair_csr<4, 4, H, K1><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, me)); \
    else if (hook_ilp == 2)                                                                        \
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair_csr<4, 2, H, K1><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, me)); \
    else                                                                                           \
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair_csr<4, 1, H, K1><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, me)); \
  } while (0)
    else if (plan.halve && plan.skip1 == 2 && low_bound(n, plan)) {
      // low-first: the ids below L, their giant root g0 (words[4]), the map
      // of which of them hang under g0, then the rest of the ids
      const uint32_t L = low_bound(n, plan);
      if (!smp) return ccx::set_error(4, "solve: no low-id map workspace");
      const unsigned nbl = blocks_for(L, c->cu_count, plan.hook_bpc);
      const unsigned nbh = blocks_for(n - L, c->cu_count, plan.hook_bpc);
      const unsigned nsl = blocks_for(L, c->cu_count, K.stream_blocks_per_cu);
      if (plan.split_walks) {
        // the owed unions queued per wave and run 64 at a time
        CCX_LAUNCH("cc_hook_low", s, k_hook_queued<false><<<nbl, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, 0u, L, nullptr, 0u, nullptr));
        CCX_LAUNCH("cc_sample_low", s, k_sample_giant<<<1, kGiantSamples, 0, s>>>(L, d_labels, words + 4));
        CCX_LAUNCH("cc_low_bits", s, k_low_bits<<<nsl, kBlock, 0, s>>>(L, d_labels, words + 4, (uint64_t *)smp));
        CCX_LAUNCH("cc_hook_sample", s, k_hook_queued<true><<<nbh, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, L, n, smp, L, words + 4));
      } else {
      CCX_LAUNCH("cc_hook_low", s, k_hook_pair_csr<4, 1, true, 2><<<nbl, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, me, 0u, L));
      CCX_LAUNCH("cc_sample_low", s, k_sample_giant<<<1, kGiantSamples, 0, s>>>(L, d_labels, words + 4));
      CCX_LAUNCH("cc_low_bits", s, k_low_bits<<<nsl, kBlock, 0, s>>>(L, d_labels, words + 4, (uint64_t *)smp));
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair_csr<4, 1, true, 2, true><<<nbh, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, me, L, n, smp, L, words + 4));
      }
    }
    else if (plan.halve && plan.skip1 == 2) CCX_HOOK_CSR(true, 2);
    else if (plan.halve && plan.skip1) CCX_HOOK_CSR(true, 1);
    else if (plan.halve) CCX_HOOK_CSR(true, 0);
    else if (plan.skip1 == 2) CCX_HOOK_CSR(false, 2);
    else if (plan.skip1) CCX_HOOK_CSR(false, 1);
    else CCX_HOOK_CSR(false, 0);
#undef CCX_HOOK_CSR
  } else if (cache) {
    if (!smp) return ccx::set_error(4, "solve: no sample buffer");
    switch (rounds) {
      case 1: CCX_LAUNCH("cc_init_min", s, k_init_sample<1><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
      case 2:
        if (plan.init_ilp == 2 || plan.init_ilp == 4) {
#define CCX_INIT2(NT)                                                                              \
  do {                                                                                             \
    if (plan.init_ilp == 4)                                                                        \
      CCX_LAUNCH("cc_init_min", s, k_init_sample2_ilp<NT, 4><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp)); \
    else                                                                                           \
      CCX_LAUNCH("cc_init_min", s, k_init_sample2_ilp<NT, 2><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp)); \
  } while (0)
          switch (plan.nt & 3) {
            case 1: CCX_INIT2(1); break;
            case 2: CCX_INIT2(2); break;
            case 3: CCX_INIT2(3); break;
            default: CCX_INIT2(0); break;
          }
#undef CCX_INIT2
          break;
        }
        switch (plan.nt & 3) {
          case 1: CCX_LAUNCH("cc_init_min", s, k_init_sample<2, 1><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
          case 2: CCX_LAUNCH("cc_init_min", s, k_init_sample<2, 2><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
          case 3: CCX_LAUNCH("cc_init_min", s, k_init_sample<2, 3><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
          default: CCX_LAUNCH("cc_init_min", s, k_init_sample<2><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
        }
        break;
      case 3: CCX_LAUNCH("cc_init_min", s, k_init_sample<3><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
      default: CCX_LAUNCH("cc_init_min", s, k_init_sample<4><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels, smp, off)); break;
    }
  } else if (ilp == 4) {
    CCX_LAUNCH("cc_init_min", s, k_init_min_ilp<4><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels));
  } else if (ilp == 2) {
    CCX_LAUNCH("cc_init_min", s, k_init_min_ilp<2><<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels));
  } else {
    CCX_LAUNCH("cc_init_min", s, k_init_min<<<ns, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, d_labels));
  }
  if (plan.csr) {
    // hooked above
  } else if (plan.fused) {
    const uint32_t *s1 = smp + n;
    const int sk = plan.skip1;
    int rc = 0;
    if (mode == 0) {
      CCX_LAUNCH("cc_hook_sample", s, k_hook_pair<0, 2><<<nb, kBlock, 0, s>>>(smp, s1, n, d_labels));
    } else if (plan.rem) {
      const bool nt = (plan.nt & 4) != 0;
      if (plan.splice == 1)
        rc = nt ? launch_hook_rem<1, true>(sk, nb, s, smp, s1, n, d_labels, d_row_ptr, d_col_idx, me, off)
                : launch_hook_rem<1, false>(sk, nb, s, smp, s1, n, d_labels, d_row_ptr, d_col_idx, me, off);
      else if (plan.splice == 2)
        rc = nt ? launch_hook_rem<2, true>(sk, nb, s, smp, s1, n, d_labels, d_row_ptr, d_col_idx, me, off)
                : launch_hook_rem<2, false>(sk, nb, s, smp, s1, n, d_labels, d_row_ptr, d_col_idx, me, off);
      else
        rc = nt ? launch_hook_rem<0, true>(sk, nb, s, smp, s1, n, d_labels, d_row_ptr, d_col_idx, me, off)
                : launch_hook_rem<0, false>(sk, nb, s, smp, s1, n, d_labels, d_row_ptr, d_col_idx, me, off);
    } else if (mode == 1) {
      rc = plan.halve ? launch_hook_pair_ilp<1, true, false>(hook_ilp, 0, nb, s, smp, s1, n, d_labels)
                      : launch_hook_pair_ilp<1, false, false>(hook_ilp, 0, nb, s, smp, s1, n, d_labels);
    } else if (plan.halve && (plan.nt & 4)) {
      rc = launch_hook_pair_ilp<4, true, true>(hook_ilp, sk, nb, s, smp, s1, n, d_labels);
    } else if (plan.halve) {
      rc = launch_hook_pair_ilp<4, true, false>(hook_ilp, sk, nb, s, smp, s1, n, d_labels);
    } else {
      rc = launch_hook_pair_ilp<4, false, false>(hook_ilp, sk, nb, s, smp, s1, n, d_labels);
    }
    if (rc) return rc;
  } else {
    for (uint32_t r = 0; r < rounds; ++r) {
      if (r > 0)
        if (int rc = compress(nullptr)) return rc;
      if (cache) {
        const uint32_t *sr = smp + (size_t)r * n;
        const bool skip = early_giant && r > 0;
        if (r == 1 && early_giant)
          CCX_LAUNCH("cc_sample_giant", s, k_sample_giant<<<1, kGiantSamples, 0, s>>>(n, d_labels, giant));
#define CCX_HOOK_CACHED(M, I)                                                                      \
  do {                                                                                             \
    if (skip)                                                                                      \
      CCX_LAUNCH("cc_hook_sample", s, k_hook_cached<M, I, true><<<nb, kBlock, 0, s>>>(sr, n, r, d_labels, giant)); \
    else                                                                                           \
      CCX_LAUNCH("cc_hook_sample", s, k_hook_cached<M, I, false><<<nb, kBlock, 0, s>>>(sr, n, r, d_labels, giant)); \
  } while (0)
        if (mode == 0) CCX_HOOK_CACHED(0, 1);
        else if (mode == 1 && hook_ilp == 4) CCX_HOOK_CACHED(1, 4);
        else if (mode == 1 && hook_ilp == 2) CCX_HOOK_CACHED(1, 2);
        else if (mode == 1) CCX_HOOK_CACHED(1, 1);
        else if (hook_ilp == 4) CCX_HOOK_CACHED(4, 4);
        else if (hook_ilp == 2) CCX_HOOK_CACHED(4, 2);
        else CCX_HOOK_CACHED(4, 1);
#undef CCX_HOOK_CACHED
      } else if ((mode == 1 || mode == 4) && hook_ilp == 4) {
        if (mode == 1) CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour_ilp<1, 4><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
        else CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour_ilp<4, 4><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      } else if ((mode == 1 || mode == 4) && hook_ilp == 2) {
        if (mode == 1) CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour_ilp<1, 2><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
        else CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour_ilp<4, 2><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      } else if (mode == 0) {
        CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour<0><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      } else if (mode == 1) {
        CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour<1><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      } else if (mode == 4) {
        CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour<4><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      } else if (mode == 2) {
        CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour<2><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      } else {
        CCX_LAUNCH("cc_hook_sample", s, k_hook_neighbour<3><<<nb, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, r, d_labels));
      }
    }
  }
  // (sampled again after early skipping: the giant tree has since absorbed
  // other trees under smaller roots, whose members would not pass through
  // the early root)
  CCX_LAUNCH("cc_sample_giant", s, k_sample_giant<<<1, kGiantSamples, 0, s>>>(n, d_labels, giant, off));
  // Exact sampling (modes 0-2) links neighbours 0..rounds-1 of every vertex,
  // so the sweep starts after them; lossy sampling (mode 3) links nothing
  // for sure, so the sweep covers every neighbour.
  // With skipping, neighbour 1 of a linked vertex was not united, so the
  // sweep starts at entry 1 (entry 0 is always united).
  // One-directional entries: the giant's members must unite theirs too,
  // from the first entry that not every vertex united (the early giant's
  // members skipped rounds 1.. of the two-pass plan).
  const uint32_t every = K.symmetric ? 0u : 1u;
  const uint32_t first =
      (mode == 3) ? 0u : ((plan.skip1 || (every && early_giant)) ? 1u : rounds);
#define CCX_FRONTIER(M)                                                                            \
  do {                                                                                             \
    if (plan.frontier_chunks == 4)                                                                 \
      CCX_LAUNCH("cc_hook_frontier", s, k_hook_frontier<M, 4><<<nf, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, first, d_labels, giant, every, off)); \
    else if (plan.frontier_chunks == 2)                                                            \
      CCX_LAUNCH("cc_hook_frontier", s, k_hook_frontier<M, 2><<<nf, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, first, d_labels, giant, every, off)); \
    else                                                                                           \
      CCX_LAUNCH("cc_hook_frontier", s, k_hook_frontier<M, 1><<<nf, kBlock, 0, s>>>(d_row_ptr, d_col_idx, n, first, d_labels, giant, every, off)); \
  } while (0)
  if (mode == 0) CCX_FRONTIER(0);
  else if (mode == 1) CCX_FRONTIER(1);
  else if (mode == 2) CCX_FRONTIER(2);
  else if (mode == 4) CCX_FRONTIER(4);
  else CCX_FRONTIER(3);
#undef CCX_FRONTIER
  return compress(giant);
}

}  // namespace

namespace {
// Words of temporary memory the solve needs beyond the control words: the
// sample cache, the tile workspace, or the larger of the two when the probe
// decides between them.
uint64_t work_words(uint32_t n, uint64_t m, const Knobs &K, const Plan &pl) {
  if (m == 0) return 0;
  if (pl.tiles == 1) return tile_words(n, pl);
  if (pl.csr && pl.low_shift) {
    const uint64_t map = low_map_words(n, pl);
    return map;
  }
  const uint64_t cache = (pl.cache && !pl.csr) ? (uint64_t)K.sampled_neighbours * n : 0;
  const uint64_t tiles = pl.tiles ? tile_words(n, pl) : 0;
  return cache > tiles ? cache : tiles;
}
}  // namespace

extern "C" int cc_hip_components_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                                        uint32_t n, uint64_t m, uint32_t *d_labels, void *stream) {
  ccx::Context *c = ccx::ctx();
  if (!c) return 2;
  if (n == 0) return 0;
  if (!d_row_ptr || !d_labels || (m > 0 && !d_col_idx))
    return ccx::set_error(4, "cc_hip_components_device: null pointer");
  if (n > ccx::kMaxVertices)
    return ccx::set_error(4, "cc_hip_components_device: n=%u exceeds the limit of %u vertices", n,
                          ccx::kMaxVertices);
  hipStream_t s = ccx::pick_stream(stream);
  // The sampled giant root and the sample cache live in this stream's own
  // state, so solves on different streams never share them.
  ccx::StreamState *st = ccx::stream_state(s);
  if (!st) return 3;
  const Knobs K = g_knobs;
  uint32_t *smp = nullptr;
  const Plan pl = plan_for(n, m, K);
  const uint64_t words = work_words(n, m, K, pl);
  if (words) {
    smp = (uint32_t *)ccx::stream_tmp(st, words * sizeof(uint32_t));
    if (!smp) return 3;
  }
  return solve_with(K, c, d_row_ptr, d_col_idx, n, m, d_labels, s, st->d_words, smp);
}

// ---- the solve without allocations ------------------------------------------
//
// cc_hip_components_device_noalloc runs the default plan with its scratch in
// the file-scope array `scratch` below instead of the stream's state: 64
// control words (the sampled giant root, the frontier's link flag and the
// plan switch; each written before it is read) followed by the
// sampled-neighbour cache, 2 words per vertex (written by the init pass
// before the hook reads it), or the tile plan's workspace (pair counts and
// root bitmap written by the local pass before they are read).  So nothing
// needs clearing between calls.  Every call on one GPU shares that GPU's
// copy of the array: calls that overlap on one GPU race.

namespace {
constexpr uint32_t kNoAllocMaxVertices = 1u << 22;  // 4,194,304 vertices
constexpr size_t kNoAllocControlWords = 64;
// the larger of the sample cache (2 words per vertex) and the tile plan's
// workspace at the default tile shape, for up to kNoAllocMaxVertices
constexpr size_t kNoAllocWorkWords =
    2 * (size_t)kNoAllocMaxVertices +
    ((((size_t)kNoAllocMaxVertices >> kTileDefaultShift) + kNoAllocMaxVertices / 32 + 63) & ~(size_t)63);
constexpr size_t kScratchWords = kNoAllocControlWords + kNoAllocWorkWords;
constexpr int kMaxDevices = 64;
// Device address of `scratch` on each device, from hipGetSymbolAddress at
// initialisation (the array's name is not a device address in host code).
uint32_t *g_scratch_dev[kMaxDevices];
}  // namespace

// 32.5 MiB, held on every GPU the library runs on until the process ends.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace ccx {
// Called by the library's initialisation once per device, with that device
// current.
int components_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(2, "scratch setup: device %d out of range", device);
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(1, "scratch setup: hipGetSymbolAddress failed on device %d: %s", device,
                     hipGetErrorString(e));
  g_scratch_dev[device] = (uint32_t *)p;
  return 0;
}
}  // namespace ccx

extern "C" int cc_hip_components_device_noalloc(const uint64_t *d_row_ptr,
                                                const uint32_t *d_col_idx, uint32_t n,
                                                uint64_t m, uint32_t *d_labels, void *stream) {
  const ccx::Context *c = ccx::ctx_if_ready();
  if (!c)
    return ccx::set_error(2, "cc_hip_components_device_noalloc: the library is not initialised "
                             "(call cc_hip_init first; this entry point never initialises it)");
  if (n == 0) return 0;
  if (!d_row_ptr || !d_labels || (m > 0 && !d_col_idx))
    return ccx::set_error(4, "cc_hip_components_device_noalloc: null pointer");
  if (n > kNoAllocMaxVertices)
    return ccx::set_error(4, "cc_hip_components_device_noalloc: n=%u exceeds the %u vertices its "
                             "scratch array holds (use cc_hip_components_device)",
                          n, kNoAllocMaxVertices);
  uint32_t *ptr = g_scratch_dev[c->device];
  if (!ptr)
    return ccx::set_error(2, "cc_hip_components_device_noalloc: scratch array not set up");
  const hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (work_words(n, m, kDefaultKnobs, plan_for(n, m, kDefaultKnobs)) > kNoAllocWorkWords)
    return ccx::set_error(4, "cc_hip_components_device_noalloc: the plan for n=%u needs more "
                             "scratch than the array holds", n);
  const ccx::ProfMute mute;  // no timing events either
  return solve_with(kDefaultKnobs, c, d_row_ptr, d_col_idx, n, m, d_labels, s, ptr,
                    ptr + kNoAllocControlWords);
}

extern "C" int cc_hip_components_host(const uint64_t *row_ptr, const uint32_t *col_idx, uint32_t n,
                                      uint64_t m, uint32_t *labels) {
  ccx::Context *c = ccx::ctx();
  if (!c) return 2;
  if (n == 0) return 0;
  if (!row_ptr || !labels || (m > 0 && !col_idx))
    return ccx::set_error(4, "cc_hip_components_host: null pointer");
  if (row_ptr[n] != m)
    return ccx::set_error(4, "cc_hip_components_host: row_ptr[n]=%llu but m=%llu",
                          (unsigned long long)row_ptr[n], (unsigned long long)m);
  const size_t rp_bytes = ((size_t)n + 1) * sizeof(uint64_t);
  const size_t col_bytes = (size_t)m * sizeof(uint32_t);
  const size_t lab_bytes = (size_t)n * sizeof(uint32_t);
  const size_t off_col = (rp_bytes + 255) & ~(size_t)255;
  const size_t off_lab = (off_col + col_bytes + 255) & ~(size_t)255;
  // a host slot of this call's own (stream and workspace): concurrent
  // callers share no device memory
  ccx::SlotGuard g;
  if (!g.slot) return 3;
  hipStream_t s = g.slot->stream;
  char *ws = (char *)ccx::slot_workspace(g.slot, off_lab + lab_bytes);
  if (!ws) return 3;
  uint64_t *d_rp = (uint64_t *)ws;
  uint32_t *d_col = (uint32_t *)(ws + off_col);
  uint32_t *d_lab = (uint32_t *)(ws + off_lab);
  CCX_TRY(hipMemcpyAsync(d_rp, row_ptr, rp_bytes, hipMemcpyHostToDevice, s));
  if (m > 0) CCX_TRY(hipMemcpyAsync(d_col, col_idx, col_bytes, hipMemcpyHostToDevice, s));
  int rc = cc_hip_components_device(d_rp, d_col, n, m, d_lab, s);
  if (rc) return rc;
  CCX_TRY(hipMemcpyAsync(labels, d_lab, lab_bytes, hipMemcpyDeviceToHost, s));
  CCX_TRY(hipStreamSynchronize(s));
  return 0;
}

// Drop-in for include/cc.h:27-28 (reference src/cc_pthreads.c:152-229).
extern "C" void compute_connected_components_pthreads(const CSRGraph *CC_RESTRICT G,
                                                      uint32_t *CC_RESTRICT labels, int num_threads,
                                                      int chunk_size) {
  (void)num_threads;
  (void)chunk_size;
  if (!G || !labels || G->n == 0) return;
  const int rc = cc_hip_components_host(G->row_ptr, G->col_idx, G->n, G->m, labels);
  if (rc != 0) {
    fprintf(stderr, "compute_connected_components_pthreads: %s\n", cc_hip_last_error());
    exit(EXIT_FAILURE);
  }
}
