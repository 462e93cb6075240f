// Definitions shared by the binned kernels and their host launcher.
#ifndef PBIG_BINNED_H_
#define PBIG_BINNED_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_device.h"

namespace pbig {

constexpr uint32_t kMaxBins = 6400;          // coarse bins per instance (LDS cursors of the scatter)
constexpr uint32_t kMaxFine = 4096;          // fine memberships of one strip held in LDS
constexpr uint32_t kMaxCoarse = 4096;        // coarse memberships of one bin (registers + LDS ids)
constexpr uint32_t kNarrowMax = 5120;        // the same under the narrow-box test schedule
constexpr uint32_t kMaxCellsPerBin = 4096;   // k*k fine cells per coarse bin (k <= 64)
constexpr uint32_t kMaxK = 64;
constexpr uint32_t kBinStage = 128;          // staged pairs per wave
constexpr uint32_t kMaxBinsPerGroup = 128;   // bins walked by one persistent test workgroup
constexpr uint32_t kBatch = 4;               // boxes per thread loaded before use (hist)
constexpr uint32_t kHistThreads = 512;       // histogram workgroup: a 4,096-box chunk in two passes, four per CU
constexpr uint32_t kScatterChunk = 4096;     // boxes per histogram/scatter workgroup (at most)
constexpr unsigned long long kErrOverfull = 8ull;  // a coarse bin did not fit in LDS
constexpr unsigned long long kErrDense = 16ull;    // heavy bins hold over a quarter of the memberships
                                                   // (checked only when the plan asks: dense_exit)
constexpr unsigned long long kErrBinCap = 32ull;   // single-pass launch: a bin's memberships exceed its fixed
                                                   // share of the bucket (the two-pass launch handles it)
// Overfull list entries: the bin, with this bit when a single fine cell
// exceeded kMaxFine (re-binning cannot help; the instance needs the general
// path) rather than the bin exceeding kMaxCoarse memberships.
constexpr uint32_t kOverfullCell = 0x80000000u;

// Grid and work layout of one instance (device table, one entry per instance).
// An instance either owns its boxes (box ids = index - box0) or is a heavy
// coarse bin of another launch re-binned on its own ("virtual" instance: its
// boxes are that launch's bin-ordered copies, ids come from the ids array,
// and its fine cells are the parent's cells [fx0, fx0 + span_x) x [fy0, fy0 +
// span_y), in bins of k < parent k cells).
struct InstDesc {
  uint64_t box0;     // first box of the instance in the box array
  uint64_t hist0;    // first entry of its histogram rows (nchunks x hist_pitch(nb))
  uint64_t off0;     // first entry of its offset rows (nchunks x off_pitch(nb))
  uint64_t k_magic;  // ceil(2^32 / k): fine index -> coarse bin by multiply
  uint64_t pairs0;   // pairs of earlier instances found outside the launch
  uint64_t pairs_extra;  // this instance's pairs found outside the launch
  uint32_t n;        // boxes
  uint32_t chunk;    // boxes per histogram/scatter workgroup
  uint32_t chunk0;   // first global chunk
  uint32_t nchunks;
  uint32_t bin0;     // first global coarse bin
  uint32_t nb;       // coarse bins = nbx * nbx
  uint32_t nbx;      // coarse bins per row
  uint32_t k;        // fine cells per coarse bin side (1..kMaxK)
  uint32_t nf;       // fine cells per row of the whole grid
  float inv_f;       // 1 / fine cell side (rounded once on the host)
  uint32_t fx0, fy0; // first fine column / row of the instance's region
  uint32_t span_x;   // fine columns of the region (nf for an owned instance)
  uint32_t span_y;   // fine rows of the region
  uint32_t parent;   // owning instance one level up (virtual instances)
  uint32_t root;     // level-0 instance a virtual instance belongs to (levels planned on the device)
};
static_assert(sizeof(InstDesc) == 112, "InstDesc layout");

// Plan of a virtual instance (host and device): coarse bin lb of instance
// P, holding m memberships at box0 of P's launch's bin-ordered arrays.  Its
// fine cells are P's cells split 2^j ways per side, j chosen so that a cell
// holds about two memberships at the bin's density (dense cells cost
// quadratically): the cell side is P's divided by 2^j, and inv_f * 2^j is
// exact, so a point's cell at this level lies inside its cell at P's level
// and the reference-cell rule claims every pair in the same region as at
// P's level.  j is capped so that the cells stay at most 2^max_refine
// times finer than the instance's own (root_inv_f): past the boxes' own size
// finer cells only multiply the memberships of every box; the bins shrink
// instead.  Bins of about `target` memberships, at most kMaxK cells per side
// and max_bins bins.  False when the bin cannot be split.
__host__ __device__ inline bool plan_virtual_bin(const InstDesc& P, uint32_t parent, uint32_t lb, uint32_t m,
                                                 uint64_t box0, uint32_t target, uint32_t max_bins, InstDesc* v,
                                                 float root_inv_f = 0.0f, uint32_t max_refine = 6) {
  *v = InstDesc{};
  v->n = m;
  v->box0 = box0;
  v->parent = parent;
  const uint32_t fx0 = P.fx0 + (lb % P.nbx) * P.k, fy0 = P.fy0 + (lb / P.nbx) * P.k;
  const uint32_t rx = P.fx0 + P.span_x - fx0, ry = P.fy0 + P.span_y - fy0;
  const uint32_t sx = P.k < rx ? P.k : rx, sy = P.k < ry ? P.k : ry;
  const double per_cell = (double)m / ((double)sx * sy);
  uint32_t done = 0;  // refinement already applied above this level (powers of two: exact)
  if (root_inv_f > 0.0f)
    while (done < 24 && root_inv_f * (float)(1u << done) < P.inv_f) ++done;
  const uint32_t jmax = max_refine > done ? max_refine - done : 0u;
  uint32_t j = 0;
  while (j < 6 && j < jmax && per_cell / (double)(1u << (2 * j)) > 2.0 && ((uint64_t)P.nf << (j + 1)) <= (1u << 21))
    ++j;
  const uint32_t f = 1u << j;
  v->inv_f = P.inv_f * (float)f;
  v->nf = P.nf * f;
  v->fx0 = fx0 * f;
  v->fy0 = fy0 * f;
  v->span_x = sx * f;
  v->span_y = sy * f;
  const uint32_t side = v->span_x > v->span_y ? v->span_x : v->span_y;
  if (side <= 1) return false;
  // bins: about m / target of them, square, k <= kMaxK, at most max_bins
  const double tq = (double)(target > 16u ? target : 16u);
  double nbx_want = ceil(sqrt((double)m / tq));
  if (nbx_want < 2.0) nbx_want = 2.0;
  uint32_t k = (uint32_t)ceil((double)side / nbx_want);
  k = k < 1u ? 1u : (k > kMaxK ? kMaxK : k);
  if (j == 0 && k > side - 1u) k = side - 1u;  // same cells: the bins must shrink
  while (((side + k - 1) / k) * ((side + k - 1) / k) > max_bins) {
    if (k >= kMaxK) return false;
    ++k;
  }
  v->k = k;
  v->k_magic = ((1ull << 32) + k - 1) / k;
  v->nbx = (side + k - 1) / k;
  v->nb = v->nbx * v->nbx;
  uint64_t nw = ((uint64_t)m + kScatterChunk - 1) / kScatterChunk;
  if (nw > 8) nw = (nw + 7) / 8 * 8;
  const uint64_t ch = ((uint64_t)m + nw - 1) / nw;
  v->chunk = (uint32_t)(ch < kScatterChunk ? ch : kScatterChunk);
  v->nchunks = m ? (uint32_t)(((uint64_t)m + v->chunk - 1) / v->chunk) : 0u;
  return true;
}

// Totals of a batch (host side).
// Row pitches of the histogram (nb counts + the chunk's invalid-box word)
// and offset tables, in entries: whole 128-byte lines, so that the scan's
// 32-bin column reads are whole lines.
__host__ __device__ inline uint32_t hist_pitch(uint32_t nb) { return (nb + 1u + 31u) & ~31u; }
__host__ __device__ inline uint32_t off_pitch(uint32_t nb) { return (nb + 31u) & ~31u; }

struct BatchPlan {
  uint32_t n_inst = 0, n_chunks = 0, n_bins = 0, test_blocks = 0;
  uint64_t n_boxes = 0, hist_entries = 0, off_entries = 0;
  uint64_t extra_total = 0;  // pairs found outside the launch (sum of pairs_extra)
  uint32_t nb_max = 0;       // bins of the largest instance (sizes the scatter's LDS)
  uint32_t n_cu = 0;         // compute units of the device (persistent grids)
  uint32_t dense_exit = 0;   // single instance: stop after the bin totals when it is dense (kErrDense)
  uint32_t direct = 0;       // single instance: the test kernel writes the output itself, in no bin order
                             // (no pair scan, no compaction, no per-instance table)
};

struct BinnedWorkspace {
  InstDesc* inst = nullptr;       // per instance
  uint32_t* chunk_inst = nullptr; // instance of each chunk
  uint32_t* bin_inst = nullptr;   // instance of each coarse bin
  uint32_t* hist = nullptr;       // per chunk: bin counts + error word
  uint32_t* off = nullptr;        // per chunk: offset inside each bin
  uint32_t* total = nullptr;      // per bin: memberships
  uint32_t* bin_start = nullptr;  // per bin: exclusive scan of total
  uint32_t* overfull = nullptr;   // bins that did not fit in LDS
  uint32_t* order = nullptr;      // the test kernel's bin order (k_bin_starts; two-pass launches)
  uint32_t* rank = nullptr;       // each bin's entry in that order
  uint2* bin_seg = nullptr;       // per bin: (first pair in its workgroup's region, pairs)
  uint64_t* bin_out = nullptr;    // per bin + 1: output position of its pairs
  uint2* scratch = nullptr;       // test_blocks x cap_wg pair regions
  uint32_t cap_wg = 0;
  uint64_t cap_scratch = 0;
  float4* bbox = nullptr;         // bin-ordered box copies
  uint32_t* bid = nullptr;        // bin-ordered box ids (inside their instance)
  Meta* meta = nullptr;
  uint64_t cap_hist = 0, cap_off = 0, cap_bins = 0, cap_bucket = 0;
  uint64_t cap_inst = 0, cap_chunks = 0, cap_bin_inst = 0;
};

// hist, scan, starts, scatter, test, pair_scan, compact
constexpr int kBinnedKernels = 7;

hipError_t binned_init();

// Enqueue the binned pipeline over the batch described by the workspace's
// tables (uploaded beforehand).  Pairs land in `out` instance by instance;
// inst_out (optional, device) receives (offset, count) per instance.
// events (optional): kBinnedKernels + 1.
// ids_in: null for instances that own their boxes; the ids of the boxes
// array (a parent launch's bin-ordered ids) for virtual instances.
// output = false: stop after the test kernel (launch_binned_output later).
hipError_t launch_binned(const BinnedWorkspace& ws, const float4* boxes, const uint32_t* ids_in,
                         const BatchPlan& bp, uint2* out, uint64_t cap, uint64_t* inst_out, hipStream_t s,
                         hipEvent_t* events = nullptr, bool output = true);

// One run of pairs found outside a launch, to be placed after instance
// inst's own pairs at position `within` of its extra area.
struct SpliceRun {
  const uint2* src;
  uint64_t len;
  uint64_t within;
  uint32_t inst;
  uint32_t pad;
};

// Copy every run to out[off_i + count_i - extra_i + within] (off_i, count_i
// from inst_out, extra_i = the instance's pairs_extra), clipped to cap.
hipError_t launch_splice_copy(const SpliceRun* runs, uint32_t n_runs, const uint64_t* inst_out,
                              const InstDesc* inst, uint2* out, uint64_t cap, hipStream_t s);

// Only the last two kernels (pair scan and output), after the instance table
// changed its pairs0 / pairs_extra fields: the test's regions are reused.
// merge (optional): the batch merge's words (shifts and counts of pairs
// found by device-planned levels, added to the runs).
hipError_t launch_binned_output(const BinnedWorkspace& ws, const BatchPlan& bp, uint2* out, uint64_t cap,
                                uint64_t* inst_out, hipStream_t s, const uint64_t* merge = nullptr);

// The pair count and error word of a call (meta's first two words) copied
// to the caller's two device words by one small kernel.
hipError_t launch_result(const Meta* meta, uint64_t* result, hipStream_t s);
// The count, scan and scatter kernels of launch_binned (events 0-4).
hipError_t launch_binned_front(const BinnedWorkspace& ws, const float4* boxes, const uint32_t* ids_in,
                               const BatchPlan& bp, hipStream_t s, hipEvent_t* ev);
// The general path's tile-pair test over a cell layout built from the bins
// of one instance's two-pass launch (plan bp, with no dense exit): meta
// cleared, count, scan and scatter into ws's bucket, k_fine_count (cell
// starts in gw.cnt, bin ends in gw.start, items, strips; counts and errors
// in meta, the bucket's in ws.meta) and k_fine_place into gw's memberships,
// then k_test (pairs counted in meta).  gw.cnt holds bins x k^2 entries;
// gw.slc bins + bins / 1024 + 1 + bins x k^2, zero before the first
// launch (a bin of more than kCountSlice memberships is counted in slices of
// kCountSlice).
// ev: kSortedKernels + 1 events.
constexpr uint32_t kCountSlice = 8192;
constexpr int kSortedKernels = 8;  // clear, bin_hist, bin_scan, bin_starts, bin_scatter, fine_count, fine_place, test
hipError_t launch_sorted(const BinnedWorkspace& ws, const float4* boxes, const BatchPlan& bp, const Workspace& gw,
                         Meta* meta, uint2* out, uint64_t cap, hipStream_t s, hipEvent_t* ev);

// Single-pass launch of one instance (no histogram or scan pass): every bin
// owns bin_cap consecutive bucket entries (bin b at b * bin_cap); each place
// workgroup counts its boxes' memberships per bin in LDS and reserves its
// share of every bin with one device atomic on fill[b].  The counters and
// result words alternate between two sets from call to call: a call zeroes
// the other set (fill_next, meta_next) for the next one.
struct OnePass {
  uint32_t* fill = nullptr;       // per bin: memberships reserved (this call)
  uint32_t* fill_next = nullptr;  // the next call's counters, zeroed by this one
  uint32_t fill_words = 0;        // entries of each fill array
  uint32_t* bin_start = nullptr;  // per bin: b * bin_cap (written by the place kernel)
  uint32_t bin_cap = 0;
  Meta* meta = nullptr;           // this call's result words
  Meta* meta_next = nullptr;      // the next call's, zeroed by this one
  uint32_t dense_exit = 0;        // a bin over its share: count to the end, and flag kErrDense when the
                                  // instance is mostly dense clusters
};

constexpr int kOnePassKernels = 2;  // place, test

hipError_t launch_binned_onepass(const BinnedWorkspace& ws, const OnePass& op, const float4* boxes,
                                 const BatchPlan& bp, uint2* out, uint64_t cap, hipStream_t s,
                                 hipEvent_t* events = nullptr);

// Workspace bounds of a re-binning level planned on the device (its plan
// kernel flags plan_err when the plan exceeds them).
struct DeepCaps {
  uint32_t inst = 0, chunks = 0, bins = 0, nb_max = 0;
  uint64_t hist = 0, off = 0;
};

// The result words of every level of a launch (level 0 first), for the
// final result kernel.
constexpr uint32_t kDeepLevels = 4;     // level 0 and up to three re-binning levels
constexpr uint32_t kDeepMaxInst = 8192;  // virtual instances of one level (LDS of its plan kernel)
struct DeepMetas {
  const Meta* meta[kDeepLevels];
  uint32_t levels;
};

// One re-binning level: its plan kernel reads the level above's list of bins
// beyond LDS (parent, parent_bins bins; parent_dev: that level was itself
// planned on the device; parent_level: its index, for the bins it cannot
// split, which go to dense_list), then hist, scan, starts, scatter and test run over
// its virtual instances with device counts.  direct: its test kernel writes
// pairs through the shared counter in its meta's out_count (one instance);
// else it leaves per-bin runs in its regions, for the batch merge below.
// max_refine caps the cells' refinement below the root instance's own
// (plan_virtual_bin): the root's inv_f is root_inst[root].inv_f when
// root_inst (the level-0 instance table) is given, else root_inv_f.
hipError_t launch_deep_level(const BinnedWorkspace& parent, uint32_t parent_bins, bool parent_dev,
                             const BinnedWorkspace& ws, const DeepCaps& caps, uint32_t target, float root_inv_f,
                             const InstDesc* root_inst, uint32_t max_refine, uint32_t n_cu, uint32_t direct,
                             uint2* out, uint64_t cap, uint32_t parent_level, uint32_t* dense_list,
                             Meta* dense, hipStream_t s);

// Batch merge of the levels' pairs into the instances' runs.  `merge` holds
// 3 * n_inst + 1 words: per level-0 instance the pairs its levels found, the
// exclusive scan of those (the shift of its run), a cursor, then the total.
//   launch_merge_clear   zero the counts and cursors (first, every launch)
//   launch_level_count   add a level's per-bin pair counts to their roots
//   launch_merge_scan    shifts and total
//   launch_binned_output (with merge) level 0's runs, shifted, and every
//                        instance's (offset, count) including its levels'
//   launch_level_copy    a level's runs after their root's level-0 pairs
hipError_t launch_merge_clear(uint64_t* merge, uint32_t n_inst, hipStream_t s);
hipError_t launch_level_count(const BinnedWorkspace& lv, uint32_t max_bins, uint64_t* merge, hipStream_t s);
hipError_t launch_merge_scan(uint64_t* merge, uint32_t n_inst, hipStream_t s);
hipError_t launch_level_copy(const BinnedWorkspace& lv, uint32_t max_bins, uint32_t lv_test_blocks,
                             const BinnedWorkspace& ws0, const uint64_t* bin_out0, uint64_t* merge, uint32_t n_inst,
                             uint2* out, uint64_t cap, hipStream_t s);
hipError_t launch_deep_result(const DeepMetas& m, const Meta* dense, uint64_t* result, hipStream_t s);

// Bins no level resolves -- bins a level's plan cannot split (a single fine
// cell over LDS at the finest cells) and bins the last level still lists
// beyond LDS -- are tested whole on the device: every pair of their
// memberships, in tiles of kDenseTile x kDenseTile, under the reference-cell
// rule on the bin's cells.  The plans append (level << 30 | bin) to `list`
// (count in the dense meta's dev_inst); the last level's list is read from
// its workspace.  At most kDenseMaxBins bins per launch (else kErrOverfull).
constexpr uint32_t kDenseTile = 256;
constexpr uint32_t kDenseMaxBins = 1024;
struct DenseSrc {
  const float4* bbox[kDeepLevels];
  const uint32_t* bid[kDeepLevels];
  const uint32_t* bin_start[kDeepLevels];
  const uint32_t* total[kDeepLevels];
  const InstDesc* inst[kDeepLevels];
  const uint32_t* bin_inst[kDeepLevels];
  const uint32_t* last_overfull;  // the last level's bins beyond LDS
  const Meta* last_meta;
  const uint32_t* list;           // bins the plans could not split
};
// A run of a batch's dense pairs (batch merge): pairs[off, off + n) belong to
// level-0 instance `root`.
struct DenseChunk {
  uint64_t off;
  uint32_t n;
  uint32_t root;
};
// dense: its meta (zeroed by launch_dense_clear at the start of the launch):
// dev_inst = bins appended by the plans; n_pairs = pairs found (batch);
// items = runs (batch); error = kErrWorkspace when `pairs` or `chunks` was
// short, kErrOverfull when more than kDenseMaxBins bins were listed.
// direct: pairs through out_count into out (one instance); else into
// pairs / chunks for the merge (launch_dense_count, launch_dense_copy).
hipError_t launch_dense_clear(Meta* dense, hipStream_t s);
hipError_t launch_dense_bins(const DenseSrc& src, uint32_t last_level, Meta* dense, uint32_t n_cu,
                             unsigned long long* out_count, uint2* out, uint64_t cap, uint2* pairs,
                             uint64_t cap_pairs, DenseChunk* chunks, uint32_t cap_chunks, hipStream_t s);
hipError_t launch_dense_count(const DenseChunk* chunks, uint32_t cap_chunks, const Meta* dense, uint64_t* merge,
                              uint32_t n_inst, hipStream_t s);
hipError_t launch_dense_copy(const DenseChunk* chunks, uint32_t cap_chunks, const Meta* dense, const uint2* pairs,
                             const BinnedWorkspace& ws0, uint64_t* merge, uint32_t n_inst, uint2* out, uint64_t cap,
                             hipStream_t s);

// Tables of a launch of one instance written by a kernel from its argument:
// the instance record, and zero in the chunk -> instance and bin -> instance
// tables (no copy and no memset, so the sequence may be captured).
hipError_t launch_single_tables(const BinnedWorkspace& ws, const InstDesc& d, hipStream_t s);

}  // namespace pbig

#endif  // PBIG_BINNED_H_
