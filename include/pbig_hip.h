/*
 * pbig_hip.h -- C-ABI of the MI355X (gfx950) pairwise box intersection engine.
 *
 * Plain C types only: pointers, sizes and status codes.  Boxes are 2-D float32
 * records {xmin, xmax, ymin, ymax} (16 bytes, the CoordInfo field order of the
 * reference, src/pbig/rect.h:6-12); a box's id is its index in the array
 * (the reference's implicit contract rid == index, src/pbig/pbig.cu:331,737).
 * Two boxes overlap iff their closed intervals overlap in x and in y
 * (src/pbig/pbig.cu:562-568, 996-1001).  Each overlapping unordered pair is
 * reported exactly once, as {lo, hi} (two uint32, lo < hi); pair order in the
 * output is unspecified (as in the reference, whose order depends on atomics).
 *
 * Every entry point returns PBIG_OK (0) or a negative status; the text of the
 * last error is available from pbig_hip_last_error().  Nothing calls exit().
 */
#ifndef PBIG_HIP_H_
#define PBIG_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PBIG_OK = 0,
  PBIG_E_ARG = -1,        /* bad argument (null pointer, n too large, bad cell size) */
  PBIG_E_INVALID_BOX = -2,/* a box has xmin > xmax, ymin > ymax or a NaN */
  PBIG_E_OVERFLOW = -3,   /* more pairs than the caller's capacity; *n_pairs holds the count */
  PBIG_E_HIP = -4,        /* HIP runtime error */
  PBIG_E_NOMEM = -5,      /* device memory exhausted or over the configured budget */
  PBIG_E_TOO_DENSE = -6   /* one cell holds more boxes than the engine's tile limit */
};

typedef struct pbig_hip_ctx pbig_hip_ctx;

/* Every `stream` argument is a hipStream_t; NULL selects the context's own
 * stream and PBIG_HIP_NULL_STREAM selects HIP's null (default) stream. */
#define PBIG_HIP_NULL_STREAM ((void*)~(uintptr_t)0)

/* Statistics of the last check on a context. */
typedef struct pbig_hip_stats {
  uint32_t cell_size_num;  /* cell side = cell_size_num / cell_size_den */
  uint32_t cell_size_den;
  uint32_t cells_per_row;
  uint32_t reserved0;
  uint64_t n_boxes;
  uint64_t n_cells;
  uint64_t memberships;    /* sum over cells of boxes binned into the cell */
  uint64_t tile_pairs;     /* work items of the per-cell test */
  uint64_t n_pairs;        /* overlapping pairs found */
  uint64_t max_per_cell;   /* general path only */
  int32_t path;            /* PBIG_PATH_GENERAL or PBIG_PATH_BINNED */
  uint32_t bins_per_row;   /* binned path: coarse bins per row */
  uint32_t cells_per_bin_side; /* binned path: fine cells per coarse bin side */
  uint32_t reserved1;
} pbig_hip_stats;

/* Which implementation of the plain path runs.
 *   PBIG_PATH_GENERAL  single-level grid in global memory (any input)
 *   PBIG_PATH_BINNED   two-level grid: coarse bins partitioned in global
 *                      memory, fine cells binned and tested in LDS; falls
 *                      back to GENERAL under AUTO when a bin does not fit
 *   PBIG_PATH_AUTO     BINNED when it applies, else GENERAL (default) */
enum { PBIG_PATH_AUTO = 0, PBIG_PATH_GENERAL = 1, PBIG_PATH_BINNED = 2 };

typedef struct pbig_hip_config {
  int32_t path;          /* PBIG_PATH_* */
  uint32_t bin_target;   /* binned path: boxes per coarse bin to aim for (0 = 2000) */
  /* PBIG_PATH_AUTO, synchronous calls: the smallest instance (in boxes)
   * whose coarse bins beyond LDS, outside mostly dense clusters, are
   * re-binned by levels on the device; smaller ones run on the general path
   * (the levels' fixed cost, about 0.3 ms, exceeds the general path's whole
   * cost below about a million boxes).  Instances of mostly dense clusters
   * take the levels only from 64 times this size (the general path's per-bin
   * cell layout is faster on them at every size measured).  0 = 2^20. */
  uint32_t levels_min_boxes;
  /* Nonzero: coarse bins beyond LDS are always re-binned by the host
   * between levels (synchronous calls; asynchronous launches report them,
   * error bit 3), never by levels planned on the device.  0 (default):
   * synchronous calls settle on device levels where they complete the
   * instance or batch, and later asynchronous launches and captures of it
   * run them too. */
  uint32_t host_levels;
} pbig_hip_config;

/* Lifecycle.  Replaces pbig_init (src/pbig/pbig.h:9, pbig.cu:2139-2217) and
 * pbig_finalize (src/pbig/pbig.h:11, pbig.cu:141): one context per device and
 * caller thread instead of process-global state.  device_id selects the HIP
 * device (the reference's deviceID argument, pbig.cu:2176-2177). */
int pbig_hip_create(pbig_hip_ctx** ctx, int device_id);
int pbig_hip_destroy(pbig_hip_ctx* ctx);
const char* pbig_hip_last_error(const pbig_hip_ctx* ctx);

/* Reference cell-size rule (decideCellSize, src/pbig/pbig.cu:1285-1357) on
 * host boxes: the power-of-two side chosen from the first 1000 boxes. */
uint32_t pbig_hip_reference_cell_size(const float* h_boxes, uint32_t n, double range);

/* Plain path, device-resident boxes.  Replaces the device part of pbig_check
 * (src/pbig/pbig.h:10, pbig.cu:1962-1981: plan, map, schedule, match).
 *   d_boxes     n x {xmin,xmax,ymin,ymax} float32 in device memory (16-byte aligned)
 *   n           number of boxes, < 2^30
 *   range       domain side: cells cover [0, range)^2; coordinates outside
 *               the domain fold into the border cells (any float is accepted)
 *   cell_size   cell side; 0 = the reference rule on the device boxes' first
 *               1000 entries (pbig_init's cs = 0, pbig.cu:1371-1374)
 *   d_pairs     device buffer of pair_capacity x {lo, hi} uint32
 *   n_pairs     out: number of overlapping pairs (also when > pair_capacity,
 *               in which case PBIG_E_OVERFLOW is returned and d_pairs holds
 *               an unspecified subset)
 *   stream      hipStream_t to run on (NULL = the context's stream)
 * Synchronous: returns after the pairs are in d_pairs. */
int pbig_hip_check_device(pbig_hip_ctx* ctx, const float* d_boxes, uint32_t n, double range,
                          float cell_size, uint32_t* d_pairs, uint64_t pair_capacity,
                          uint64_t* n_pairs, void* stream);

/* Asynchronous form of pbig_hip_check_device: enqueues the whole pipeline on
 * `stream` and returns without waiting.  cell_size must be > 0.  After the
 * stream reaches this point, d_result[0] holds the pair count and
 * d_result[1] an error word (0 = ok; bit 0 invalid box, bit 1 workspace too
 * small, bit 2 cell too dense, bit 3 a coarse bin did not fit in LDS, bit 4
 * the instance is mostly dense clusters, which the automatic path choice
 * sends to the general path, bit 5 a coarse bin outgrew its fixed share of
 * the single-pass launch's bucket).  d_result is 2 x uint64 in device memory.
 * The path (and the workspace size) follow the context's last synchronous
 * call; when bit 1, 3, 4 or 5 is set, run the synchronous entry point once
 * (it settles the path for this instance shape) and call again. */
int pbig_hip_check_device_async(pbig_hip_ctx* ctx, const float* d_boxes, uint32_t n, double range,
                                float cell_size, uint32_t* d_pairs, uint64_t pair_capacity,
                                uint64_t* d_result, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the asynchronous plain path on the binned pipeline
 * (same kernels and result words as pbig_hip_check_device_async), with its
 * whole workspace in a fixed 48 MiB device array that the library holds on
 * each device until the process ends.  It calls no hipMalloc, hipFree or
 * memset; the caller passes the boxes, the pair buffer and d_result (2 x
 * uint64 device words: pair count, error word).  n up to about 2^20 boxes
 * (the exact need depends on the plan; a larger call returns PBIG_E_ARG
 * before launching anything); cell_size must be > 0.  The error word's bit
 * 1 (more memberships or pairs per test workgroup than the array holds) or
 * bit 3 (a coarse bin beyond LDS) means the result is incomplete: use the
 * allocating entry points for that input.  Every call on one device shares
 * that device's array and leaves its contents behind (each launch clears
 * what it reads before use); calls that overlap on one device, from any
 * context or stream, race. */
int pbig_hip_check_device_noalloc(pbig_hip_ctx* ctx, const float* d_boxes, uint32_t n, double range,
                                  float cell_size, uint32_t* d_pairs, uint64_t pair_capacity,
                                  uint64_t* d_result, void* stream);

/* Grow the context's workspace for up to n_boxes boxes, n_cells cells and
 * n_memberships cell memberships (used by the async path). */
int pbig_hip_reserve(pbig_hip_ctx* ctx, uint64_t n_boxes, uint64_t n_cells, uint64_t n_memberships);

/* Plain path, host boxes in the reference's record layout: n x {xmin, xmax,
 * ymin, ymax, rid} (20 bytes, RectInfo of src/pbig/rect.h:15-18 with
 * CType = float).  rid must equal the index.  Packs the records into pinned
 * memory over the context's host threads (at most 16), copies in, runs, and
 * copies the pairs back in chunks of 8 Mi pairs through two pinned buffers,
 * each chunk spread into h_pairs (pair_capacity x {lo, hi}) by the host
 * threads while the next one is in flight.  This is the host half of
 * pbig_check (pbig.cu:1401 copy-in, 1850-1909 copy-out). */
int pbig_hip_check_host(pbig_hip_ctx* ctx, const void* h_rects, uint32_t n, double range,
                        float cell_size, uint32_t* h_pairs, uint64_t pair_capacity,
                        uint64_t* n_pairs);

/* The pairs of a host check handed over chunk by chunk: pairs points at n
 * {lo, hi} uint32 in the library's pinned memory, valid during the call. */
typedef void (*pbig_hip_pairs_fn)(void* user, const uint32_t* pairs, uint64_t n);

/* pbig_hip_check_host with the pairs handed to fn(user, chunk, n) chunk by
 * chunk, on the calling thread, while the next chunk is copied: no buffer
 * for all pairs.  The reference's match phase prefetches results on one
 * stream while the host decodes the previous ones into the report
 * (pbig.cu:1848-1866, decode.cc:26-138); pbig_check's report callbacks run
 * here.  *n_pairs = the total. */
int pbig_hip_check_host_each(pbig_hip_ctx* ctx, const void* h_rects, uint32_t n, double range,
                             float cell_size, pbig_hip_pairs_fn fn, void* user, uint64_t* n_pairs);

int pbig_hip_last_stats(const pbig_hip_ctx* ctx, pbig_hip_stats* out);

/* Choose the implementation and its tuning (no counterpart in the reference,
 * whose choices are compile-time flags, src/pbig/config.cmake:1-10). */
int pbig_hip_configure(pbig_hip_ctx* ctx, const pbig_hip_config* cfg);

/* Per-kernel device timing of the asynchronous path (no counterpart in the
 * reference, whose CUDA-event phase timers, src/pbig/cuda_comm.h:39-45, print
 * through LOG).  pbig_hip_timing_start arms HIP events around each of the
 * plain path's kernels for the next max_calls asynchronous calls;
 * pbig_hip_timing_read waits for them and returns the summed milliseconds per
 * kernel in launch order, the number of calls recorded and the path they
 * took (n_kernels >= 7).  GENERAL: clear, count, scan_reduce, scan_apply,
 * scatter, test.  BINNED: hist, scan, starts, scatter, test, pair_scan,
 * compact.  Path 3, the binned path's single-pass launch of one instance:
 * place, test.  Path 4, the two-pass launch of one instance followed by
 * re-binning levels planned on the device: one duration per level (level 0
 * is the two-pass launch).  Batch launches are timed as BINNED. */
int pbig_hip_timing_start(pbig_hip_ctx* ctx, int max_calls);
int pbig_hip_timing_read(pbig_hip_ctx* ctx, double* ms_per_kernel, int n_kernels, int* n_calls, int* path);

/* ------------------------------------------------------------------------
 * Batches: many independent instances in one launch sequence.
 *
 * The reference checks one instance per pbig_check call (src/pbig/pbig.h:10;
 * its example driver loops over inputs, package/main.cc:22-40).  A batch plans a
 * fixed set of instances once; every launch then runs all of them through
 * the binned path's seven kernels, whatever the number of instances.  The
 * instances' boxes are concatenated in one device array; instance i's boxes
 * are d_boxes[box_offset .. box_offset + n_boxes).  Pair ids are indices
 * inside the instance.  The pairs of instance i occupy
 * d_pairs[offset_i .. offset_i + count_i), instances in order.
 * ------------------------------------------------------------------------ */
typedef struct pbig_hip_instance {
  uint64_t box_offset;   /* first box of the instance in the batch's box array */
  uint32_t n_boxes;      /* < 2^30 */
  float cell_size;       /* fine cell side, > 0 */
  double range;          /* domain side (as for pbig_hip_check_device) */
} pbig_hip_instance;

typedef struct pbig_hip_batch pbig_hip_batch;

/* Plan a batch of n_inst instances on ctx's device (tables uploaded, first
 * workspace allocated).  The batch keeps a pointer to ctx: destroy the batch
 * first.  Calls on a batch and on its context must not overlap. */
int pbig_hip_batch_create(pbig_hip_ctx* ctx, const pbig_hip_instance* inst, uint32_t n_inst,
                          pbig_hip_batch** batch);
int pbig_hip_batch_destroy(pbig_hip_batch* batch);

/* Re-plan a batch for a new set of instances, keeping (and only growing) its
 * device workspace: the serving loop's way to reuse a batch slot.  The
 * tables are uploaded asynchronously on `stream` (launches on any stream are
 * ordered after the upload); the batch's earlier work must be complete.
 * Drops a captured graph. */
int pbig_hip_batch_update(pbig_hip_batch* batch, const pbig_hip_instance* inst, uint32_t n_inst, void* stream);

/* Synchronous batch check.  d_boxes holds n_boxes_total boxes (every
 * instance's range must lie inside).  h_inst_pairs (optional, host) receives
 * n_inst x {offset, count}.  Regrows the workspace when needed.  A coarse
 * bin whose boxes exceed LDS is re-binned into smaller bins: by up to three
 * levels planned on the device over every such bin of the batch in one
 * launch sequence (the batch then keeps them: its later launches and
 * captures run them, complete on the device), else (bins the levels cannot
 * split, more of them than a level takes, or host_levels configured) by the
 * host between levels; an instance with a single fine cell beyond LDS runs
 * on the general path.  Either way its pairs are placed in its run.
 * *n_pairs = total; PBIG_E_OVERFLOW when it exceeds pair_capacity (offsets
 * and counts are still returned). */
int pbig_hip_batch_check(pbig_hip_batch* batch, const float* d_boxes, uint64_t n_boxes_total,
                         uint32_t* d_pairs, uint64_t pair_capacity, uint64_t* h_inst_pairs,
                         uint64_t* n_pairs, void* stream);

/* Complete a pbig_hip_batch_launch (same arguments, same stream) whose error
 * word was nonzero: the launch runs again with levels on the device (which
 * the batch then keeps); where they do not apply and only coarse bins beyond
 * LDS were reported, those are re-binned by the host and spliced without
 * running the launch again; otherwise this is pbig_hip_batch_check. */
int pbig_hip_batch_finish(pbig_hip_batch* batch, const float* d_boxes, uint64_t n_boxes_total,
                          uint32_t* d_pairs, uint64_t pair_capacity, uint64_t* h_inst_pairs,
                          uint64_t* n_pairs, void* stream);

/* After a pbig_hip_batch_launch (or replay, on this stream) whose error word
 * holds only the bins-beyond-LDS bit and whose pair total fits the launch's
 * pair_capacity: h_flags[i] = 1 (n_inst bytes) for each instance that had
 * such a bin, *n_flagged their number (after a launch with device levels
 * that completed every instance, none).  Every other instance's (offset,
 * count) entry and pair run from that launch is then complete; a flagged
 * instance's count is short, and pbig_hip_batch_finish, or a call on that
 * instance alone, completes it.  Fails with PBIG_E_ARG when the error word
 * holds anything else, and with PBIG_E_OVERFLOW when the launch found more
 * pairs than its capacity (the pairs past it were dropped, so later runs
 * are incomplete: finish with a larger buffer).  No counterpart in the reference (one instance per call,
 * src/pbig/pbig.h:10); it lets a server hand over the finished instances of
 * a batch without waiting for the others' host work. */
int pbig_hip_batch_unresolved(pbig_hip_batch* batch, uint8_t* h_flags, uint32_t* n_flagged, void* stream);

/* Asynchronous batch launch: enqueues the seven kernels (once a synchronous
 * call settled the batch on device levels, also the levels and the merge of
 * their pairs into the instances' runs) and returns; no host
 * synchronisation, no allocation, no memset, so it may be captured in a HIP
 * graph.  d_inst_pairs (optional, device) receives n_inst x {offset, count};
 * d_result (optional, device, 2 x uint64) the total and the error word of
 * pbig_hip_check_device_async.  When the error word is nonzero, run
 * pbig_hip_batch_check once (it regrows or falls back) and launch again. */
int pbig_hip_batch_launch(pbig_hip_batch* batch, const float* d_boxes, uint64_t n_boxes_total,
                          uint32_t* d_pairs, uint64_t pair_capacity, uint64_t* d_inst_pairs,
                          uint64_t* d_result, void* stream);

/* Capture pbig_hip_batch_launch with these arguments into a HIP graph owned
 * by the batch (replacing an earlier one), and replay it.  Replay returns
 * PBIG_E_ARG when the workspace has moved since the capture (after a
 * pbig_hip_batch_check that regrew it): capture again. */
int pbig_hip_batch_capture(pbig_hip_batch* batch, const float* d_boxes, uint64_t n_boxes_total,
                           uint32_t* d_pairs, uint64_t pair_capacity, uint64_t* d_inst_pairs,
                           uint64_t* d_result);
int pbig_hip_batch_replay(pbig_hip_batch* batch, void* stream);

/* Totals of a batch's plan: chunks, coarse bins, boxes (no counterpart in the
 * reference). */
int pbig_hip_batch_info(const pbig_hip_batch* batch, uint64_t* n_boxes, uint32_t* n_chunks, uint32_t* n_bins);

/* ------------------------------------------------------------------------
 * Reference-compatible unsigned mode: the reference's shipped build
 * (CType = unsigned int, DIMS = 2 or 3; src/pbig/rect.h:4-12,
 * config.cmake:8), restated literally: cells v >> log2(cs), 32-bit cell ids
 * with out-of-grid ids skipped, unsigned <= overlap tests and the
 * reference-cell rule by cell id (pbig.cu:174-221, 562-603).  Boxes are
 * n x 2*dims uint32 {xmin,xmax,ymin,ymax[,zmin,zmax]}; cell_size is a power of
 * two (0 = the reference's rule on the first 1000 boxes).  With
 * pair_capacity 0 the pairs are only counted.
 * ------------------------------------------------------------------------ */
uint32_t pbig_hip_reference_cell_size_u32(const uint32_t* h_boxes, uint32_t n, int dims, uint32_t range);
int pbig_hip_check_u32(pbig_hip_ctx* ctx, const uint32_t* d_boxes, uint32_t n, int dims, uint32_t range,
                       uint32_t cell_size, uint32_t* d_pairs, uint64_t pair_capacity, uint64_t* n_pairs,
                       void* stream);
/* Host records: n x RectInfo {CoordInfo (2*dims uint32), rid} with rid ==
 * index, as pbig_check receives them in the reference's build (packing and
 * the copy back as for pbig_hip_check_host). */
int pbig_hip_check_host_u32(pbig_hip_ctx* ctx, const void* h_rects, uint32_t n, int dims, uint32_t range,
                            uint32_t cell_size, uint32_t* h_pairs, uint64_t pair_capacity, uint64_t* n_pairs);
/* The same, the pairs handed to fn chunk by chunk (pbig_hip_check_host_each). */
int pbig_hip_check_host_u32_each(pbig_hip_ctx* ctx, const void* h_rects, uint32_t n, int dims, uint32_t range,
                                 uint32_t cell_size, pbig_hip_pairs_fn fn, void* user, uint64_t* n_pairs);

/* ------------------------------------------------------------------------
 * Compressed pair stream in the reference's record layout (its default
 * COMPRESS = 2 build: 64-bit words; blocks [0x0fffff11, gid] followed by
 * records {u8 seg1, u8 seg2, u16 l_rid, u16 nlaps, u8 offset[nlaps]} padded to
 * whole words; pbig.cu:849-888, 519-527, decode.cc:26-138), with the cell
 * arrays a reader needs: cell_start (n_cells + 1 u32) and cell_recs (u32 box
 * ids).  Pair (a, b) = (cell_recs[cell_start[gid] + 256 seg1 + l_rid],
 * cell_recs[cell_start[gid] + 256 seg2 + offset]).  Host functions, no
 * context.  The reference's decoder reads the offsets 4 bytes past where its
 * encoder writes them (decode.cc:80 vs pbig.cu:851); this decoder reads them
 * where they are written.
 * ------------------------------------------------------------------------ */
/* Encode n_pairs pairs (lo, hi) among n_boxes boxes.  Sizes are returned in
 * *n_words, *n_cells (cell_start holds n_cells + 1 entries) and *n_recs;
 * PBIG_E_OVERFLOW when a capacity is too small (call with zero capacities to
 * size the buffers). */
int pbig_hip_rice_encode(const uint32_t* pairs, uint64_t n_pairs, uint32_t n_boxes, uint64_t* words,
                         uint64_t word_capacity, uint64_t* n_words, uint32_t* cell_start, uint64_t start_capacity,
                         uint64_t* n_cells, uint32_t* cell_recs, uint64_t recs_capacity, uint64_t* n_recs);
/* Decode a stream into (lo, hi) pairs; *n_pairs = the number of pairs
 * (PBIG_E_OVERFLOW past pair_capacity; PBIG_E_ARG on a malformed stream). */
int pbig_hip_rice_decode(const uint64_t* words, uint64_t n_words, const uint32_t* cell_start, uint64_t n_cells,
                         const uint32_t* cell_recs, uint64_t n_recs, uint32_t* pairs, uint64_t pair_capacity,
                         uint64_t* n_pairs);

#ifdef __cplusplus
}
#endif

#endif /* PBIG_HIP_H_ */
