#ifndef CC_HIP_H
#define CC_HIP_H

/*
 * Device-resident C ABI of the MI355X connected-components library
 * (libccmi355x.so).  Everything here takes plain pointers and sizes; device
 * pointers are named d_*, pinned host pointers h_*.  A `stream` argument is a
 * hipStream_t passed as void*; NULL means the library's own stream.
 *
 * These are the building blocks the reference-compatible entry points
 * (cc.h, boundary_sort.h, exchange_plan.h) are made of, exposed so a caller
 * can keep the CSR, labels and boundary records resident in HBM between
 * calls instead of paying a host round trip each time.
 *
 * Return value convention: 0 on success, nonzero on failure with a message
 * available from cc_hip_last_error().  Nothing here falls back to the CPU.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ---------------------------------------------------------- */

/* Select and initialise the device (idempotent; the other calls initialise
 * device 0 lazily).  Returns nonzero if no usable gfx950 device exists. */
int cc_hip_init(int device);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int cc_hip_device_count(void);
/* Text of the last error on this thread ("" if none). */
const char *cc_hip_last_error(void);
/* Library build identification string (arch and options). */
const char *cc_hip_build_info(void);

int cc_hip_malloc(void **d_ptr, size_t bytes);
int cc_hip_free(void *d_ptr);
int cc_hip_host_alloc(void **h_ptr, size_t bytes); /* pinned, device-visible */
int cc_hip_host_free(void *h_ptr);
int cc_hip_memcpy_h2d(void *d_dst, const void *src, size_t bytes, void *stream);
int cc_hip_memcpy_d2h(void *dst, const void *d_src, size_t bytes, void *stream);
int cc_hip_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream);
/* Device-to-host copy enqueued on `stream` without waiting (h_dst should be
 * pinned for the copy to be asynchronous). */
int cc_hip_memcpy_d2h_async(void *h_dst, const void *d_src, size_t bytes, void *stream);
int cc_hip_stream_create(void **stream);
int cc_hip_stream_destroy(void *stream);
int cc_hip_stream_sync(void *stream);
int cc_hip_device_sync(void);

/* ---- kernel timing ---------------------------------------------------- */

/* When enabled, every kernel launch of the library is bracketed by HIP
 * events on the stream it runs on, and its duration is accumulated per
 * kernel name.  Disabled by default. */
int cc_hip_prof_enable(int on);
int cc_hip_prof_reset(void);
/* Collect pending event pairs (synchronises on them) and return, for kernel
 * name `name`, the number of launches and their total milliseconds. */
int cc_hip_prof_query(const char *name, uint64_t *launches, double *total_ms);
/* Comma-separated list of kernel names seen since the last reset. */
const char *cc_hip_prof_names(void);

/* ---- local connected components --------------------------------------- */

/*
 * d_labels[v] = smallest vertex id in v's component of the undirected graph
 * (d_row_ptr[n+1], d_col_idx[m]) resident in device memory.  The graph must
 * be symmetric (both directions stored), as the reference's local CSR is
 * (src/cc_mpi.c:252-304).  Self-loops, duplicate entries and unsorted rows
 * are allowed; column ids >= n are ignored.  n is at most 0xFF000000 (error
 * 4 above that).  Algorithm: each vertex united with two sampled neighbours
 * on a parent-pointer forest, then every remaining neighbour of the vertices
 * outside the giant tree, then pointer jumping; the result does not depend
 * on the order the atomics land in.  Asynchronous on `stream`.
 */
int cc_hip_components_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                             uint32_t n, uint64_t m, uint32_t *d_labels, void *stream);

/*
 * cc_hip_components_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Runs the default plan (whatever cc_hip_components_set has
 * tuned) with its scratch in a fixed device array of the library (32.5 MiB
 * per GPU, held until the process ends) instead of the stream's state;
 * calls no hipMalloc, hipFree or other allocating routine, records no timing
 * events, and never initialises the library: cc_hip_init must have run
 * (return 2 otherwise).  n <= 4,194,304 (return 4 beyond).  Same labels as
 * cc_hip_components_device.  Every call on one GPU shares that GPU's scratch
 * array, so calls that overlap on one GPU (on different streams, or from
 * different threads) race; calls queued one after another on one stream
 * are fine.  Asynchronous on `stream` (NULL: the library stream).
 */
int cc_hip_components_device_noalloc(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                                     uint32_t n, uint64_t m, uint32_t *d_labels, void *stream);

/* Process-wide tuning of the solve: the number of sampled hooking rounds
 * (1..4, default 2; 0 keeps the current value), the grid cap of the
 * sampled-hook kernel in workgroups per CU (1..64; 0 keeps), and the
 * hooking mode (-1 keeps; 0 plain loads, 1 roots confirmed with an L2 read
 * before each atomic and re-walked with L2 reads, 2 all parent reads from
 * L2, 3 lossy sampling with every neighbour swept outside the giant tree,
 * 4 roots confirmed with an L2 read and re-walked with plain reads
 * (default)).  Results never depend on any of them. */
int cc_hip_components_tune(int sampled_neighbours, int blocks_per_cu, int hook_mode);
/* The same knobs by name:
 *   "sampled_neighbours"      1..4 (default 2)
 *   "hook_mode"               0..4 (default 4)
 *   "sample_cache"            1 (default): the init pass copies each vertex's
 *                             sampled neighbours into coalesced arrays
 *                             (rounds x n u32 in the stream's temp buffer)
 *                             that the sampled rounds read; 0: the rounds
 *                             gather from col_idx.  Modes 0, 1 and 4 only.
 *   "fused_pair"              1 or -1 (default): hook both sampled
 *                             neighbours in one pass (needs the cache and 2
 *                             rounds); 0: one pass per round
 *   "rem"                     the one-pass hook unites each vertex with its
 *                             sampled neighbours by Rem's algorithm (both
 *                             paths climbed together, stopping where they
 *                             meet) instead of walking all three roots in
 *                             lockstep: 1, 0, or -1 (default) = on when
 *                             m < 16 n, or n < 2^23; not in mode 0
 *   "rem_splice"              splicing in those unions (a step lowers the
 *                             entry it leaves to the other side's parent):
 *                             1 on every step, 2 on the vertex's own entry
 *                             only, 0 never, or -1 (default) = 1 when
 *                             n >= 2^24, else 2 when m >= 16 n, else 0
 *   "halve"                   path halving (atomic min of each visited
 *                             entry to its grandparent) in the lockstep
 *                             one-pass hook's root walks: 1, 0, or -1
 *                             (default) = on when n >= 2^24, or n >= 2^23
 *                             and m >= 16 n; not in mode 0
 *   "giant_early"             1 (default): on the two-pass cached path,
 *                             sample the giant root after round 0 and skip
 *                             its members in round 1; 0: sample it once,
 *                             after the last round
 *   "blocks_per_cu"           sampled-hook grid cap; 0 (default) = 8 for
 *                             Rem's unions unless they splice on every step
 *                             (rem_splice 1), and when the lockstep hook
 *                             walks several roots per thread; 48 for the
 *                             skipping hook from the CSR (16 below 2^24
 *                             vertices); else 4
 *   "hook_ilp"                roots walked per thread in lockstep in the
 *                             sampled hook: 1, 2, 4, or 0 (default) = 1 when
 *                             m < 16 n or n >= 2^23, else 2 in the one-pass
 *                             hook, 4 per round
 *   "frontier_blocks_per_cu"  grid cap of the frontier sweep (default 8)
 *   "stream_blocks_per_cu"    grid cap of init/compress (default 8)
 *   "stream_ilp"              vertices per thread per step in init/compress
 *                             without the cache (1 default, 2, 4); with 1 the
 *                             pointer-jumping passes take 4 after a skipping
 *                             hook
 *   "nt"                      non-temporal streaming, a bitmask: 1 the init
 *                             pass stores the sample arrays non-temporally,
 *                             2 it loads the column ids non-temporally, 4 the
 *                             one-pass hook loads the sample arrays
 *                             non-temporally; -1 (default) = 5 when
 *                             n >= 2^24, else 0
 *   "init_ilp"                rows per thread per step of the sampling init
 *                             pass (two sampled neighbours): 1, 2, 4, or 0
 *                             (default) = 4 when n >= 2^24 and m >= 16 n,
 *                             else 1
 *   "frontier_chunks"         64-vertex chunks per wave step of the frontier
 *                             sweep, their root walks in lockstep: 1, 2, 4, or
 *                             0 (default) = 2
 *   "csr_hook"                the one-pass hook reads each row's first two
 *                             entries itself after an identity init, with no
 *                             sample arrays: 1, 0, or -1 (default) = on for
 *                             the lockstep hook with one root per thread
 *                             (n >= 2^23, m >= 16 n)
 *   "skip_linked"             a vertex linked under its neighbour 0 by its
 *                             own entry skips the union with neighbour 1, so
 *                             only the local minima (neighbour 0 above them)
 *                             walk, and the frontier sweep starts at entry 1:
 *                             0 off, 1 on, 2 on with the link going to
 *                             neighbour 0's parent (one hop); -1 (default) =
 *                             2 for the lockstep hook from the CSR and for
 *                             Rem's unions on sparse graphs with m >= 6 n
 *                             (m < 16 n), else 0; applies to the one-pass
 *                             hooks from the sample arrays and to the
 *                             lockstep hook from the CSR
 *   "minima_entries"          entries of its row a local minimum unites with
 *                             when skipping: 2..8, or 0 (default) = 2
 *   "symmetric"               1 (default): the CSR stores both directions of
 *                             every edge, and the frontier sweep skips the
 *                             giant tree's members; 0: every vertex unites
 *                             all its entries, giving weakly connected
 *                             components of a CSR with one-directional
 *                             entries (a union per entry)
 *   "tiles"                   the vertex-tile plan for graphs whose ids are
 *                             local (a lattice or mesh numbered row by row):
 *                             each workgroup unites the entries inside its
 *                             tile of consecutive ids in LDS, then the pairs
 *                             between tiles are united in HBM: 1 always, 0
 *                             never, -1 (default) = a probe on the device
 *                             picks it when at most 35% of the first 8
 *                             entries of 1024 sampled rows leave their
 *                             tile, on
 *                             sparse graphs (m < 16 n) of 2^20+ vertices
 *                             that the default sampled plan would give to
 *                             Rem's unions from the sample arrays; the solve
 *                             never waits for the probe on the host
 *   "tile_shift"              log2 of the vertices per tile: 6..16, or 0
 *                             (default) = 16 (16-bit local parents: a
 *                             128 KiB LDS array; up to 15, 64 KiB)
 *   "tile_pairs"              pairs between tiles kept per tile (a tile with
 *                             more reads its rows again): 1..2^24, or 0
 *                             (default) = one per vertex of the tile; the
 *                             plan's temporary memory is about 8 bytes per
 *                             vertex at the default
 *   "low_shift"               the low-first hook, on the one-hop skipping
 *                             hook from the CSR with one root per thread:
 *                             the ids below L = n >> low_shift (rounded up
 *                             to 64) are hooked first, a bitmap marks which
 *                             of them hang under their most frequent root
 *                             g0, and each other vertex whose neighbour 0
 *                             is marked links under g0 without gathering
 *                             neighbour 0's parent: 1..8, 0 off (one pass),
 *                             or -1 (default) = 4; temporary memory L/8
 *                             bytes
 *   "split_walks"             with the low-first hook: each wave queues, in
 *                             LDS, the vertices that still owe unions after
 *                             their own-entry link (local minima, failed
 *                             links) and runs them 64 at a time, so no wave
 *                             waits on one walker per step: 1, 0, or -1
 *                             (default) = 1 below 2^31 entries, else 0
 *                             (above it the Chung-Lu graphs lose ~2%)
 * Returns nonzero for an unknown key or out-of-range value; _get returns -1
 * for an unknown key. */
int cc_hip_components_set(const char *key, int value);
int cc_hip_components_get(const char *key);

/* The choices a solve of an n-vertex, m-entry graph makes under the current
 * knobs: cache in use, one-pass hook, roots per thread and grid cap of the
 * sampled hook.  Any pointer may be NULL.  Returns 0. */
int cc_hip_components_plan(uint32_t n, uint64_t m, int *cache, int *fused, int *hook_ilp,
                           int *hook_blocks_per_cu);
/* One field of that plan by name: "cache", "fused", "rem", "rem_splice",
 * "halve", "hook_ilp", "hook_blocks_per_cu", "giant_early" (in effect for
 * this graph), "hook_mode", "nt", "init_ilp", "frontier_chunks", "csr_hook",
 * "skip_linked", "minima_entries", "compress_ilp" (vertices per thread of
 * the pointer-jumping passes), "tiles" (0 no, 1 always, 2 the probe
 * decides), "tile_shift", "tile_pairs", "low_shift" (0: no low-first
 * hook), "split_walks".  Returns -1 for an unknown key. */
int cc_hip_components_plan_get(uint32_t n, uint64_t m, const char *key);
/* Whether the last solve on `stream` (NULL = the library stream) ran its
 * final pointer-jumping pass: *ran = 1 if its frontier sweep linked a root,
 * 0 if the pass was skipped because every vertex already held its final
 * root.  Blocking (waits for the stream). */
int cc_hip_components_last_pass_ran(void *stream, int *ran);
/* Whether the last solve with entries on `stream` (NULL = the library
 * stream) ran the vertex-tile plan: *ran = 1 if so, 0 if the sampled plan
 * ran.  Blocking (waits for the stream). */
int cc_hip_components_tiles_ran(void *stream, int *ran);

/* Same result through host buffers (H2D, solve, D2H; blocking). */
int cc_hip_components_host(const uint64_t *row_ptr, const uint32_t *col_idx,
                           uint32_t n, uint64_t m, uint32_t *labels);

/* ---- boundary records (device-resident forms of boundary_sort.h) -------- */

/*
 * Ascending sort of count 64-bit keys.  A BoundaryEdge array is such a key
 * array (see vec_helpers.h).  d_alt is a second buffer of count keys used as
 * the ping-pong target; if result_in_alt is non-NULL it receives 1 when the
 * sorted keys ended in d_alt, else the result is copied back into d_keys.
 * Blocks until the digit plan is read back (one small copy), then enqueues
 * the passes on `stream`.  Replaces src/boundary_sort.c:93-98.
 */
int cc_hip_sort_u64_device(uint64_t *d_keys, uint64_t *d_alt, uint64_t count,
                           int *result_in_alt, void *stream);

/* Rep fill, sort and dedup of a rank's raw boundary records in one call:
 * (remote << 32) | li  ->  ascending unique (remote << 32) | d_comp_of[li],
 * the result of cc_hip_boundary_set_rep_device + cc_hip_sort_u64_device +
 * cc_hip_dedup_u64_device (src/cc_mpi.c:364-394, src/boundary_sort.c:93-119).
 * d_comp_of[0..n_local) must map every local vertex to a root r of its
 * local forest (d_comp_of[r] == r), as cc_hip_components_device leaves it.
 * Reps are sorted by their rank among the roots, which keeps the order and
 * shortens the key to 32 + ceil(log2 roots) bits (fewer digit passes); the
 * records of the most frequent rep (sampled) only set bits of a bitmap over
 * the remotes, and the other records alone are sorted, then merged with the
 * set bits read in order ("boundary_split" of cc_hip_sort_set).  Temporary
 * device memory: 12 bytes per local vertex, about 2 bytes per possible
 * remote id, and for the partitioned bitmap 24 bytes per (bucket of 2^17
 * remote ids, tile of 16,384 records) pair (0.75 bytes per record at 512
 * buckets).  *out_count receives the kept count; the records are in d_alt when
 * *result_in_alt is 1, else in d_records; the other buffer is clobbered.
 * Blocking. */
int cc_hip_boundary_build_device(uint64_t *d_records, uint64_t *d_alt, uint64_t count,
                                 const uint32_t *d_comp_of, uint32_t n_local, uint64_t *out_count,
                                 int *result_in_alt, void *stream);

/* Tuning of the sort's scatter passes: "items_per_thread" (8, 16 (default),
 * 24 or 32 keys per thread) and "sort_threads" (256, 512 (default) with 8 or
 * 16 keys, 1024 with 8; other pairs run 256 threads): a workgroup's tile is
 * threads x keys; of the
 * ordered compactions (dedup, ghost extraction, send lists, delta pack):
 * "compact_rows" (1 default: row-striped tiles, each wave reading 64
 * consecutive items per row; 0: a run of consecutive items per thread,
 * where the next two knobs apply); "dedup_staged" (0 default: each thread
 * stores its own items; 1: outputs staged in LDS and stored as contiguous
 * runs); "dedup_items" (16 or 32 (default) records per thread); of the
 * boundary build: "boundary_split" (the records of the most frequent rep
 * go through a bitmap over the remotes and only the rest are sorted, then
 * the two sorted lists are merged; 2 default: the bitmap is built by
 * partitioning those records' remotes into buckets of 2^17 ids, each set in
 * LDS by one workgroup (remote ids needing more than 1024 buckets take
 * mode 1); 1: each record probes and sets its bit in the bitmap in memory;
 * 0: every record sorted) and "split_dominant_rows" (16 or 32 (default)
 * rows of 64 records per wave step of the probing pass); and of the split:
 * "split_staged" (as dedup_staged; staging measured
 * slower at 854M entries because the LDS costs occupancy).  The result
 * never depends on them.  _set returns nonzero for a bad key or value, _get
 * -1 for an unknown key. */
int cc_hip_sort_set(const char *key, int value);
int cc_hip_sort_get(const char *key);

/* Adjacent-unique of d_in[0..count) into d_out (must not alias d_in);
 * *out_count receives the kept count (blocking).  src/boundary_sort.c:100-119. */
int cc_hip_dedup_u64_device(const uint64_t *d_in, uint64_t count, uint64_t *d_out,
                            uint64_t *out_count, void *stream);

/*
 * Ghost extraction over sorted, deduplicated records (remote << 32) | rep:
 * d_ghost_vertices receives the distinct remotes in order (*ghost_count of
 * them) and d_out_records the records with remote replaced by its ghost
 * index (d_out_records must not alias d_records; blocking).
 * src/cc_mpi.c:431-457.
 */
int cc_hip_ghost_extract_device(const uint64_t *d_records, uint64_t count,
                                uint64_t *d_out_records, uint32_t *d_ghost_vertices,
                                uint64_t *ghost_count, void *stream);

/* ---- delta exchange pack ---------------------------------------------- */

/*
 * The pack half of exchangeplan_exchange_delta (src/exchange_plan.c:325-357)
 * with the plan's send list and the label arrays resident on the device.
 * seg_start (host, nseg+1 ascending entries from 0 to total_send_to) gives
 * each destination's segment of d_send_to_flat in dests order.  For each
 * entry whose current label (comp_label[comp_of[v - v_start]] if v is in
 * [v_start, v_end), else v) is below d_prev_sent, d_prev_sent is lowered and
 * the pair ((idx - segment start) << 32) | label is written to h_sendbuf
 * (pinned host memory from cc_hip_host_alloc, or device memory; room for
 * total_send_to pairs) in ascending idx order: compacted in the stream's
 * HBM temp buffer (which grows to total_send_to pairs), then copied by one
 * DMA of the emitted pairs.  sendcounts/sdispls (nseg ints each) and
 * *total_pairs receive the layout.  Blocking.
 */
int cc_hip_delta_pack_device(const uint32_t *d_send_to_flat, uint64_t total_send_to,
                             const int *seg_start, int nseg,
                             const uint32_t *d_comp_label, const uint32_t *d_comp_of,
                             uint32_t v_start, uint32_t v_end, uint32_t *d_prev_sent,
                             uint64_t *h_sendbuf, int *sendcounts, int *sdispls,
                             uint64_t *total_pairs, void *stream);

/* The owned entries' reps gathered once: d_rep_of_send[i] =
 * d_comp_of[d_send_to_flat[i] - v_start] for every i with the id in
 * [v_start, v_end), 0xFFFFFFFF otherwise.  comp_of does not change during
 * the merge loop (src/cc_mpi.c:541-710), so the list serves every exchange.
 * Asynchronous on `stream`. */
int cc_hip_delta_prepare_device(const uint32_t *d_send_to_flat, uint64_t total_send_to,
                                const uint32_t *d_comp_of, uint32_t v_start, uint32_t v_end,
                                uint32_t *d_rep_of_send, void *stream);

/* cc_hip_delta_pack_device with the reps from cc_hip_delta_prepare_device:
 * lbl = comp_label[rep_of_send[idx]] (the id itself where the rep list holds
 * 0xFFFFFFFF), so a pack streams the rep list and prev_sent and gathers
 * only comp_label at the reps.  Same output, same contract. */
int cc_hip_delta_pack_reps_device(const uint32_t *d_send_to_flat, const uint32_t *d_rep_of_send,
                                  uint64_t total_send_to, const int *seg_start, int nseg,
                                  const uint32_t *d_comp_label, uint32_t *d_prev_sent,
                                  uint64_t *h_sendbuf, int *sendcounts, int *sdispls,
                                  uint64_t *total_pairs, void *stream);

/* Received-pair unpack (src/exchange_plan.c:390-413): pairs (pos << 32) |
 * lbl grouped per source k from rdispls[k] (h_pairs: pinned host or device
 * memory); d_ghost_labels[d_need_from_gidx_flat[need_base[k] + pos]] = lbl.
 * Blocking. */
int cc_hip_delta_unpack_device(const uint64_t *h_pairs, uint64_t total, const int *rdispls,
                               const int *need_base, int nsrc,
                               const uint32_t *d_need_from_gidx_flat, uint32_t *d_ghost_labels,
                               void *stream);

/* ---- per-rank driver building blocks (src/cc_mpi.c) -------------------- */

/* d_out[0..n] = exclusive prefix sums of d_in[0..n) (d_out[n] = total). */
int cc_hip_exclusive_scan_u32_device(const uint32_t *d_in, uint64_t n, uint64_t *d_out,
                                     void *stream);

/*
 * Local-local split of an owned CSR with global column ids
 * (src/cc_mpi.c:252-304, 364-394).  d_row_ptr holds the n_local+1 offsets
 * of the owned rows into d_col_idx (they need not start at 0).  Count
 * phase: the number of local-local entries (columns in [v_start, v_end))
 * and of boundary entries (other columns < n_global), blocking.  Emit
 * phase: d_lrow[0..n_local] = local CSR offsets, d_lcol = rebased local
 * columns, d_boff[0..n_local] = boundary-record offsets per row, d_brec =
 * records (remote << 32) | li, all in row order (one order-preserving
 * compaction over the entries).
 */
int cc_hip_split_count_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                              uint32_t n_local, uint32_t v_start, uint32_t v_end,
                              uint32_t n_global, uint64_t *m_local_local, uint64_t *n_boundary,
                              void *stream);
int cc_hip_split_emit_device(const uint64_t *d_row_ptr, const uint32_t *d_col_idx,
                             uint32_t n_local, uint32_t v_start, uint32_t v_end,
                             uint32_t n_global, uint64_t *d_lrow, uint64_t *d_boff,
                             uint32_t *d_lcol, uint64_t *d_brec, void *stream);

/* records (remote << 32) | li  ->  (remote << 32) | comp_of[li]. */
int cc_hip_boundary_set_rep_device(uint64_t *d_records, uint64_t count,
                                   const uint32_t *d_comp_of, void *stream);

/* d_out[i] = base + i. */
int cc_hip_iota_u32_device(uint32_t *d_out, uint32_t n, uint32_t base, void *stream);

/* Send-list construction for the exchange plan: keys (owner(remote) << 32)
 * | li from raw records (remote << 32) | li, owner by the partition bounds
 * d_bounds[0..comm_size] (src/cc_mpi.c:41-76). */
int cc_hip_owner_keys_device(const uint64_t *d_records, uint64_t count, const uint32_t *d_bounds,
                             int comm_size, uint64_t *d_keys, void *stream);
/* The send lists of the exchange plan in one step, without sorting: for
 * each rank q in [0, comm_size) ascending, the ascending owned local ids li
 * (as li + base) that have a raw record (remote << 32) | li whose remote end
 * q owns under the ascending partition bounds d_bounds[0..comm_size] (the
 * same lists as owner keys -> sort -> dedup -> segments).  h_starts[0..comm_size] (host) receives the
 * segment starts; d_vals needs room for every listed id (at most
 * min(count, n_local * comm_size)).  Records must carry li < n_local; those
 * of one li should be contiguous (as the split emits them) for speed, not
 * for correctness.  comm_size <= 64.  Blocking. */
int cc_hip_send_lists_device(const uint64_t *d_records, uint64_t count, uint32_t n_local,
                             const uint32_t *d_bounds, int comm_size, uint32_t base,
                             int *h_starts, uint32_t *d_vals, void *stream);
/* Segments of sorted unique keys (q << 32) | x for q in [0, nseg):
 * d_starts[q] = first index of segment q (d_starts[nseg] = count) and
 * d_vals[i] = x + base. */
int cc_hip_segments_device(const uint64_t *d_keys, uint64_t count, int nseg, uint32_t base,
                           int *d_starts, uint32_t *d_vals, void *stream);
/* d_out[i] = d_table[d_idx[i]]. */
int cc_hip_gather_u32_device(const uint32_t *d_table, const uint32_t *d_idx, uint64_t count,
                             uint32_t *d_out, void *stream);

/*
 * One merge round (src/cc_mpi.c:547-622) over records (gidx << 32) | rep:
 * comp_min = comp_label on the reps, comp_min[rep] lowered by every
 * record's ghost label, then per rep up to four label hops and the
 * comp_label update.  *changed = 1 if any comp_label dropped (blocking).
 */
int cc_hip_merge_step_device(const uint64_t *d_records, uint64_t count,
                             const uint32_t *d_ghost_vertices, const uint32_t *d_ghost_labels,
                             uint32_t ghost_count, const uint32_t *d_comp_of,
                             uint32_t *d_comp_label, uint32_t *d_comp_min, uint32_t n_local,
                             uint32_t v_start, uint32_t v_end, int *changed, void *stream);

/* The same round with the reps given as a list: d_reps[0..rep_count) holds
 * every r with comp_of[r] == r (cc_hip_local_reps_device), so the per-rep
 * passes touch only those entries instead of scanning comp_of (the merge
 * loop of src/cc_mpi.c:541-710 walks the rep list built at :328-362).
 * A forest over n_local > 0 vertices has at least one rep. */
int cc_hip_merge_step_reps_device(const uint64_t *d_records, uint64_t count,
                                  const uint32_t *d_ghost_vertices, const uint32_t *d_ghost_labels,
                                  uint32_t ghost_count, const uint32_t *d_comp_of,
                                  uint32_t *d_comp_label, uint32_t *d_comp_min, uint32_t n_local,
                                  const uint32_t *d_reps, uint32_t rep_count, uint32_t v_start,
                                  uint32_t v_end, int *changed, void *stream);

/* Once per merge loop, over the records (gidx << 32) | rep of the merge
 * round (sorted, unique; every rep in the ascending rep list d_reps of
 * cc_hip_local_reps_device): the most frequent rep among 1024 records
 * spread evenly over them (the choice affects speed only) is the dominant
 * one, and *dominant receives its index in d_reps; d_dom_bits
 * (ceil(ghost_count / 32) words) gets bit gidx set for every record of the
 * dominant rep; every other record goes, in order, to d_side (room for
 * count) as (gidx << 32) | index of its rep in d_reps, *side_count of them
 * (blocking; error 4 if a record's rep is not in the list).  The merge
 * loop of src/cc_mpi.c:541-710 reads the same records every round; at one
 * rank's share nearly all carry the rep of the rank's giant component. */
int cc_hip_merge_prepare_device(const uint64_t *d_records, uint64_t count, uint32_t ghost_count,
                                const uint32_t *d_reps, uint32_t rep_count, uint32_t *d_dom_bits,
                                uint64_t *d_side, uint64_t *side_count, uint32_t *dominant,
                                void *stream);

/* cc_hip_merge_step_reps_device over the prepared records: the same round
 * (src/cc_mpi.c:547-622), with the hook's per-rep minimum ghost label kept
 * in d_ghost_min (rep_count entries, in rep-list order) across rounds;
 * dominant, d_dom_bits, d_side and side_count as the prepare call left
 * them for the same d_reps.
 * ghosts_changed: 2 on the first round (the minimum is seeded and computed
 * from the dominant rep's bitmap and the side records); 1 when
 * d_ghost_labels has been lowered since the last round (after an exchange
 * that delivered pairs: the minimum is lowered by the records in place);
 * 0 when it is unchanged (the minimum is reused).  Exchanged ghost labels
 * only ever decrease (a rank sends a label only below the one it sent
 * before), which is what makes the in-place update exact; a caller whose
 * ghost labels can rise passes 2 instead of 1. */
int cc_hip_merge_step_prepared_device(
    const uint32_t *d_dom_bits, uint32_t dominant, const uint64_t *d_side, uint64_t side_count,
    const uint32_t *d_ghost_vertices, const uint32_t *d_ghost_labels, uint32_t ghost_count,
    const uint32_t *d_comp_of, uint32_t *d_comp_label, uint32_t *d_comp_min,
    uint32_t *d_ghost_min, uint32_t n_local, const uint32_t *d_reps, uint32_t rep_count,
    uint32_t v_start, uint32_t v_end, int ghosts_changed, int *changed, void *stream);

/* The local reps {r < n_local : comp_of[r] == r}, ascending, into d_reps
 * (room for n_local); *rep_count receives their number (blocking).
 * src/cc_mpi.c:328-362. */
int cc_hip_local_reps_device(const uint32_t *d_comp_of, uint32_t n_local, uint32_t *d_reps,
                             uint32_t *rep_count, void *stream);

/* d_labels[li] = comp_label[comp_of[li]] (if d_labels non-NULL) and *roots =
 * #{li : that label == v_start + li} (src/cc_mpi.c:730-736; blocking). */
int cc_hip_rank_labels_device(const uint32_t *d_comp_label, const uint32_t *d_comp_of,
                              uint32_t n_local, uint32_t v_start, uint32_t *d_labels,
                              uint64_t *roots, void *stream);

/* ---- batched solves (many graphs, one captured HIP graph replay) -------- */

typedef struct cc_hip_batch_s *cc_hip_batch_t;

/* A batch engine with room for max_items work items (vertex ranges of one
 * graph each) and max_graphs graphs per launch: its own stream, device
 * tables, pinned mirrors and an instantiated HIP graph of the whole solve. */
int cc_hip_batch_create(uint32_t max_items, uint32_t max_graphs, cc_hip_batch_t *out);
int cc_hip_batch_destroy(cc_hip_batch_t b);
/* Solve shape captured by engines created afterwards: "one_pass" 1
 * (default) hooks both sampled neighbours in one pass, 0 a pass per
 * neighbour with pointer jumping between; "rem" 1 (default) makes that one
 * pass unite by Rem's algorithm, 0 by lockstep root walks; "grid_per_cu"
 * (default 8, 0..64) workgroups per CU of each launch, which loops over
 * the work items, 0 = one workgroup per item of the engine's capacity.
 * Results never depend on them. */
int cc_hip_batch_set(const char *key, int value);
/*
 * Solve n_graphs resident graphs (symmetric CSRs d_row_ptr[g], d_col_idx[g]
 * of n[g] <= 0xFF000000 vertices) into d_labels[g] (per-graph minimum ids)
 * with one graph replay; each work item holds at most item_vertices (multiple of 256)
 * vertices of one graph.  Asynchronous on the batch's stream; waits for the
 * previous launch of the same batch first.  *items_used = work items.
 */
int cc_hip_batch_launch(cc_hip_batch_t b, uint32_t n_graphs, const uint64_t *const *d_row_ptr,
                        const uint32_t *const *d_col_idx, const uint32_t *n,
                        uint32_t *const *d_labels, uint32_t item_vertices,
                        uint32_t *items_used);
/* 1 when the last launch has finished, 0 while it runs, < 0 on error. */
int cc_hip_batch_query(cc_hip_batch_t b);
int cc_hip_batch_sync(cc_hip_batch_t b);
/* The batch's stream (hipStream_t), for copies ordered after its solves. */
void *cc_hip_batch_stream(cc_hip_batch_t b);

/* ---- device-side construction of the seeded inputs -------------------- */

/*
 * Undirected pairs of a seeded generator, identical pair for pair to the
 * host generators of graphgen.h.  kind 0 = Erdos-Renyi {n, pairs};
 * 1 = R-MAT {scale, edge_factor, ta, tab, tabc} (16-bit thresholds);
 * 2 = grid {rows, cols, keep threshold} (2 slots per vertex);
 * 3 = Chung-Lu {n, pairs, bit pattern of the double cdf total} with d_cdf
 * from ccgen_chung_lu_cdf.  *count receives the number of pairs written.
 */
int cc_hip_gen_pairs_device(int kind, const uint64_t *params, const double *d_cdf,
                            uint64_t seed, uint32_t *d_src, uint32_t *d_dst, uint64_t *count,
                            void *stream);

/*
 * Symmetric simple CSR from `count` undirected pairs: self-loops and
 * out-of-range pairs dropped, both directions kept, duplicates removed,
 * rows sorted; optional renaming d_perm (n entries).  d_keys and d_alt
 * hold 2*count keys each; d_col needs room for 2*count entries (it may be
 * the pair buffer itself, read before it is written); d_row_ptr n+1.
 * *m_out receives the entry count (blocking).
 */
int cc_hip_csr_from_pairs_device(uint32_t n, const uint32_t *d_src, const uint32_t *d_dst,
                                 uint64_t count, const uint32_t *d_perm, uint64_t *d_keys,
                                 uint64_t *d_alt, uint64_t *d_row_ptr, uint32_t *d_col,
                                 uint64_t *m_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CC_HIP_H */
