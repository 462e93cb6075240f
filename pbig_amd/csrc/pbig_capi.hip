// C-ABI of the engine (declared in include/pbig_hip.h): contexts, workspace
// management, the synchronous/asynchronous device entry points and the host
// (RectInfo) entry point.  Status codes instead of exit() (the reference
// exits on every error, src/pbig/log.h:4-19, cuda_comm.h:14-23).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "pbig_binned.h"
#include "pbig_host.h"
#include "pbig_device.h"
#include "pbig_hip.h"
#include "pbig_u32.h"

using pbig::GridParams;
using pbig::Meta;
using pbig::ScanPartial;
using pbig::Workspace;
using pbig::BatchPlan;
using pbig::BinnedWorkspace;
using pbig::InstDesc;

// Tables and workspace of one binned plan: the context's plain path (a batch
// of one) or a prepared batch.
struct BinnedState {
  BinnedWorkspace ws;
  std::vector<InstDesc> inst;
  std::vector<uint32_t> chunk_inst, bin_inst;
  BatchPlan bp;
  uint64_t generation = 0;  // bumped whenever a device buffer moves
  // Pinned staging of the tables: uploads are asynchronous; stage_ev marks
  // the last one (launches on other streams wait for it).
  unsigned char* h_stage = nullptr;
  size_t cap_stage = 0;
  hipEvent_t stage_ev = nullptr;
  // plain-path plan key; capacities of chunk_inst/bin_inst known to be zero
  uint64_t zero_chunks = 0, zero_bins = 0;
  bool key_valid = false;
  uint32_t key_n = 0, key_target = 0;
  double key_range = 0.0;
  float key_cell = 0.0f;
  // Next level: the coarse bins that did not fit in LDS, re-binned as
  // virtual instances; its pairs and (offset, count) table.
  std::unique_ptr<BinnedState> sub;
  uint2* sub_out = nullptr;
  uint64_t cap_sub_out = 0;
  uint64_t* sub_io = nullptr;
  uint64_t cap_sub_io = 0;
  uint64_t* io_buf = nullptr;      // (offset, count) per instance when the caller gives none
  uint64_t cap_io_buf = 0;
  pbig::SpliceRun* runs = nullptr; // device table of the runs to place
  uint64_t cap_runs = 0;
  bool own_meta = false;
  // Single-pass launch of the plain path (no histogram or scan): every bin
  // owns bin_cap bucket entries; counters and result words alternate
  // between two sets by call parity (each call zeroes the other set).
  uint32_t bin_cap = 0;          // 0: this plan runs the two-pass launch
  bool onepass_failed = false;   // this plan's single-pass launch reported trouble: two passes from now on
  bool onepass_dirty = true;     // the counter sets must be cleared before the next single-pass launch
  uint32_t* fill[2] = {nullptr, nullptr};
  uint64_t cap_fill = 0;
  Meta* meta1 = nullptr;         // two result-word sets
  uint32_t* start1 = nullptr;    // per bin: b * bin_cap (written by each place kernel)
  uint64_t cap_start1 = 0;
  uint32_t parity = 0;
  // Re-binning levels planned on the device (levels 1.. of the plain path's
  // two-pass launch), each with its own workspace and bounds.  deep_ok: the
  // plan's last synchronous call completed on them, so asynchronous calls
  // take them too; deep_failed: they cannot complete this plan (bins that do
  // not split, or more levels than there are).
  std::unique_ptr<BinnedState> dl[pbig::kDeepLevels];
  pbig::DeepCaps dcaps[pbig::kDeepLevels];
  bool deep_ok = false, deep_failed = false;
  bool to_general = false;  // the automatic choice sends this plan's instance to the general path
  bool dense_levels = false;  // ... or, mostly dense clusters, to the device levels (level 0 without the dense exit)
  uint64_t* d_res = nullptr;  // result words of a synchronous call on them
  // Bins no level resolves, tested whole: their list, status words, and (a
  // batch) the pairs and runs they produce for the merge.
  Meta* dense_meta = nullptr;
  uint32_t* dense_list = nullptr;
  uint2* dense_pairs = nullptr;
  uint64_t cap_dense_pairs = 0;
  pbig::DenseChunk* dense_chunks = nullptr;
  uint64_t cap_dense_chunks = 0;
};

struct pbig_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  Workspace ws;
  Meta* h_meta = nullptr;          // pinned mirrors of ws.meta and sort_meta
  // Result words of the general path built from the bins (launch_sorted):
  // its fine memberships, items and pairs; the bins' own in ws.meta.
  Meta* sort_meta = nullptr;
  // Host-path staging.
  float4* d_in = nullptr;
  uint64_t cap_in = 0;
  float4* h_in = nullptr;          // pinned
  uint64_t cap_h_in = 0;
  uint2* d_pairs = nullptr;
  uint64_t cap_pairs = 0;
  BinnedState plain;
  // Unsigned (reference-compatible) mode: memberships and host staging.
  pbig::U32Workspace uw;
  uint32_t* d_in32 = nullptr;
  uint64_t cap_in32 = 0;
  uint32_t* h_in32 = nullptr;      // pinned
  uint64_t cap_h_in32 = 0;
  // Host entry points: threads for packing and copies, and two pinned
  // buffers the pairs come back through, chunk by chunk.
  std::unique_ptr<pbig::HostPool> pool;
  uint2* h_out[2] = {nullptr, nullptr};
  uint64_t cap_h_out = 0;
  hipEvent_t out_ev[2] = {nullptr, nullptr};
  // Pinned ring for the plain path's instance record: uploads stay ordered
  // on the stream without a host synchronisation.
  InstDesc* h_ring = nullptr;
  hipEvent_t ring_ev[16] = {};
  uint32_t ring_next = 0;
  unsigned char* noalloc = nullptr;  // this device's address of the no-allocation scratch
  pbig_hip_config cfg{PBIG_PATH_AUTO, 2000, 1u << 20, 0};
  int last_path = 0;
  pbig_hip_stats stats{};
  std::string err;
  // Per-kernel timing (pbig_hip_timing_*): one event set per recorded call.
  std::vector<hipEvent_t> events;
  int timing_cap = 0;     // event sets allocated
  int timing_used = 0;    // event sets recorded since the last read
  int timing_path = 0;
};

namespace {
constexpr int kEventsPerCall = 10;
// Timing path id of the plain path's launch with device re-binning levels:
// one duration per level (level 0 is the two-pass launch).
constexpr int kPathDeep = 4;
// Timing path id of the single-pass launch (place, test), beside
// PBIG_PATH_GENERAL and PBIG_PATH_BINNED.
constexpr int kPathOnePass = 3;
// Timing path id of the general path's tile test on a cell layout built
// from the bins (launch_sorted).
constexpr int kPathSorted = 5;
// Instances of mostly dense clusters take the device levels from this many
// times levels_min_boxes (2^26 boxes by default, past every size measured;
// other instances with bins beyond LDS from levels_min_boxes).
constexpr uint64_t kDenseLevelsFactor = 64;

// Device memory of pbig_hip_check_device_noalloc: the whole workspace of a
// single-instance binned launch, carved per call from this array.  Each
// device that loads the library holds its own copy for the life of the
// process.  In host code `scratch` only names the symbol; its device
// address comes from hipGetSymbolAddress, once per device (noalloc_setup).
constexpr uint64_t kNoAllocBytes = 48ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kMaxDevices] = {};

// The scratch array's address on `device`, which must be current.
hipError_t noalloc_setup(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = (unsigned char*)p;
  }
  *out = g_noalloc_ptr[device];
  return hipSuccess;
}
}

namespace {

// NULL: the context's stream; PBIG_HIP_NULL_STREAM: HIP's null stream.
hipStream_t pick_stream(const pbig_hip_ctx* c, void* stream) {
  if (!stream) return c->stream;
  if (stream == PBIG_HIP_NULL_STREAM) return (hipStream_t)0;
  return (hipStream_t)stream;
}

int fail(pbig_hip_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), e_ == hipErrorOutOfMemory ? PBIG_E_NOMEM : PBIG_E_HIP, "%s: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);                        \
  } while (0)

template <typename T>
int grow(pbig_hip_ctx* c, T** p, uint64_t* cap, uint64_t want) {
  if (*cap >= want && *p) return PBIG_OK;
  uint64_t n = std::max<uint64_t>(want, 1);
  if (*p) HIP_TRY(c, hipFree(*p));
  *p = nullptr;
  *cap = 0;
  HIP_TRY(c, hipMalloc((void**)p, n * sizeof(T)));
  *cap = n;
  return PBIG_OK;
}

// Cells per row for a domain side and cell side.
int grid_of(pbig_hip_ctx* c, double range, float cell, GridParams* gp, uint32_t* n_cells) {
  if (!(cell > 0.0f) || !std::isfinite(cell)) return fail(c, PBIG_E_ARG, "cell size must be > 0");
  if (!(range > 0.0) || !std::isfinite(range)) return fail(c, PBIG_E_ARG, "range must be > 0");
  double nrc = std::ceil(range / (double)cell);
  if (nrc < 1) nrc = 1;
  if (nrc * nrc > 2147483647.0)
    return fail(c, PBIG_E_ARG, "grid of %.0f x %.0f cells is too large (cell size %g, range %g)", nrc,
                nrc, (double)cell, range);
  gp->nrc = (uint32_t)nrc;
  gp->inv = 1.0f / cell;
  *n_cells = gp->nrc * gp->nrc;
  return PBIG_OK;
}

int ensure_cells(pbig_hip_ctx* c, uint64_t n_cells) {
  Workspace& w = c->ws;
  int rc;
  if (w.cap_cells < n_cells || !w.cnt || !w.start) {
    if (w.cnt) HIP_TRY(c, hipFree(w.cnt));
    if (w.start) HIP_TRY(c, hipFree(w.start));
    w.cnt = nullptr;
    w.start = nullptr;
    w.cap_cells = 0;
    uint64_t cap = std::max<uint64_t>(n_cells, 1), d1 = 0, d2 = 0;
    if ((rc = grow(c, &w.cnt, &d1, cap))) return rc;
    if ((rc = grow(c, &w.start, &d2, cap))) return rc;
    w.cap_cells = cap;
  }
  uint64_t parts = (w.cap_cells + pbig::kScanTile - 1) / pbig::kScanTile;
  if ((rc = grow(c, &w.partials, &w.cap_partials, parts))) return rc;
  return PBIG_OK;
}

int ensure_work(pbig_hip_ctx* c, uint64_t memb, uint64_t items) {
  Workspace& w = c->ws;
  int rc;
  if (memb > 0xFFFFFFFFull) return fail(c, PBIG_E_NOMEM, "more than 2^32-1 cell memberships");
  if (w.cap_memb < memb || !w.mbox) {
    uint64_t capm = memb;
    uint64_t d1 = 0, d2 = 0;
    if (w.mbox) HIP_TRY(c, hipFree(w.mbox));
    if (w.mid) HIP_TRY(c, hipFree(w.mid));
    w.mbox = nullptr;
    w.mid = nullptr;
    w.cap_memb = 0;
    if ((rc = grow(c, &w.mbox, &d1, capm))) return rc;
    if ((rc = grow(c, &w.mid, &d2, capm))) return rc;
    w.cap_memb = capm;
  }
  if ((rc = grow(c, &w.items, &w.cap_items, items))) return rc;
  return PBIG_OK;
}

// Workspace sizes for a call: at least the previous call's needs, and a
// first guess of 1.5 memberships per box.
int prepare(pbig_hip_ctx* c, uint32_t n, uint32_t n_cells) {
  int rc;
  if ((rc = ensure_cells(c, n_cells))) return rc;
  uint64_t memb = std::max<uint64_t>(c->ws.cap_memb, (uint64_t)n + n / 2 + 1024);
  uint64_t items = std::max<uint64_t>(c->ws.cap_items, (uint64_t)n_cells / 2 + 1024);
  return ensure_work(c, memb, items);
}

hipEvent_t* next_events(pbig_hip_ctx* c, int path);
int upload_tables(pbig_hip_ctx* c, BinnedState& st, hipStream_t s);
int after_upload(pbig_hip_ctx* c, BinnedState& st, hipStream_t s);

int run_sync(pbig_hip_ctx* c, const float4* d_boxes, uint32_t n, GridParams gp, uint32_t n_cells,
             uint2* d_out, uint64_t cap, hipStream_t s, uint64_t* n_pairs) {
  int rc;
  if ((rc = prepare(c, n, n_cells))) return rc;
  for (int attempt = 0; attempt < 2; ++attempt) {
    HIP_TRY(c, pbig::launch_plain(c->ws, d_boxes, n, gp, n_cells, d_out, cap, s, next_events(c, PBIG_PATH_GENERAL)));
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, c->ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    const Meta m = *c->h_meta;
    if (m.error & pbig::kErrInvalidBox)
      return fail(c, PBIG_E_INVALID_BOX, "input has a box with xmin > xmax, ymin > ymax or a NaN");
    if (m.error & pbig::kErrTooDense)
      return fail(c, PBIG_E_TOO_DENSE, "a cell holds %u boxes, above the %u-tile limit", m.max_per_cell,
                  pbig::kMaxTilesPerCell);
    if (m.error & pbig::kErrWorkspace) {
      if (attempt == 1) return fail(c, PBIG_E_NOMEM, "workspace did not fit after regrowth");
      if ((rc = ensure_work(c, m.memberships + m.memberships / 16 + 1024, m.items + m.items / 16 + 1024)))
        return rc;
      continue;
    }
    c->stats.n_boxes = n;
    c->stats.n_cells = n_cells;
    c->stats.cells_per_row = gp.nrc;
    c->stats.memberships = m.memberships;
    c->stats.tile_pairs = m.items;
    c->stats.n_pairs = m.n_pairs;
    c->stats.max_per_cell = m.max_per_cell;
    *n_pairs = m.n_pairs;
    if (m.n_pairs > cap)
      return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu",
                  (unsigned long long)m.n_pairs, (unsigned long long)cap);
    return PBIG_OK;
  }
  return fail(c, PBIG_E_HIP, "unreachable");
}

// Plan of one instance on the binned path: n boxes, domain side `range`,
// fine cell side `cell`.  K fine cells per coarse bin side (1..kMaxK) such
// that a coarse bin holds about bin_target boxes, at most kMaxBins bins.
// Chunks of at most kScatterChunk boxes (the scatter holds its chunk in
// registers); more than 8 chunks are rounded to a multiple of 8 so that the
// chunks spread evenly over the 8 XCDs.  False when the grid cannot be
// planned (the general path then runs).  With n_cu (a launch of this one
// instance), k may move by one from the target's so that the persistent test
// workgroups' largest share, ceil(bins / n_cu) bins of k^2 cells, is least.
bool plan_instance(uint32_t n, double range, float cell, uint32_t bin_target, InstDesc* d, uint32_t n_cu = 0) {
  *d = InstDesc{};
  d->n = n;
  if (n == 0) return true;  // no chunks, no bins
  const double nf0 = std::max(1.0, std::ceil(range / (double)cell));
  const double target = std::max<double>(16.0, bin_target);
  double kd = std::sqrt(target / std::max<double>(n, 1)) * nf0;
  kd = std::min<double>(std::max<double>(std::round(kd), 1.0), std::min<double>(pbig::kMaxK, nf0));
  uint32_t k = (uint32_t)kd;
  if (n_cu) {
    const uint32_t k0 = k;
    double best = 0.0;
    for (uint32_t kc = k0 > 1 ? k0 - 1 : 1; kc <= k0 + 1 && kc <= pbig::kMaxK && kc <= nf0; ++kc) {
      const double bx = std::ceil(nf0 / kc), nbins = bx * bx;
      if (nbins > pbig::kMaxBins) continue;
      const double share = std::ceil(nbins / n_cu) * (double)kc * kc;
      if (best == 0.0 || share < best) {
        best = share;
        k = kc;
      }
    }
  }
  uint32_t nbx = 0;
  while (true) {
    const double nbxd = std::ceil(nf0 / k);
    if (nbxd * nbxd <= pbig::kMaxBins) { nbx = (uint32_t)nbxd; break; }
    if (k >= pbig::kMaxK) return false;
    ++k;
  }
  d->inv_f = 1.0f / cell;
  d->k = k;
  d->k_magic = ((1ull << 32) + k - 1) / k;
  d->nbx = nbx;
  d->nb = nbx * nbx;
  d->nf = nbx * k;
  uint64_t nw = ((uint64_t)n + pbig::kScatterChunk - 1) / pbig::kScatterChunk;
  if (nw > 8) nw = (nw + 7) / 8 * 8;
  d->chunk = (uint32_t)std::min<uint64_t>(pbig::kScatterChunk, ((uint64_t)n + nw - 1) / nw);
  d->nchunks = (uint32_t)(((uint64_t)n + d->chunk - 1) / d->chunk);
  d->fx0 = d->fy0 = 0;
  d->span_x = d->span_y = d->nf;
  return true;
}

// Plan of a virtual instance (pbig::plan_virtual_bin, shared with the
// device planner of re-binning levels).
bool plan_virtual(const InstDesc& P, uint32_t parent, uint32_t lb, uint32_t m, uint64_t box0, uint32_t target,
                  InstDesc* v) {
  return pbig::plan_virtual_bin(P, parent, lb, m, box0, target, pbig::kMaxBins, v);
}

// Global layout of a planned batch: first chunk, bin, histogram row, offset
// row and earlier pairs of every instance; the chunk -> instance and bin ->
// instance tables.  `skip` (optional) marks instances left to the general
// path: they keep their place in the output (pairs_extra pairs) but get no
// chunks or bins.
bool layout_tables(BinnedState& st, uint32_t n_cu, const std::vector<uint64_t>* skip_pairs) {
  BatchPlan& bp = st.bp;
  bp = BatchPlan{};
  bp.n_inst = (uint32_t)st.inst.size();
  uint64_t chunks = 0, bins = 0, hist = 0, off = 0, pairs0 = 0;
  for (size_t i = 0; i < st.inst.size(); ++i) {
    InstDesc& d = st.inst[i];
    const bool skipped = skip_pairs && (*skip_pairs)[i] != ~0ull;
    if (skipped) {  // no chunks, no bins: the output step sees an empty bin range
      d.nchunks = 0;
      d.nb = 0;
    }
    const uint32_t nchunks = d.nchunks, nb = d.n == 0 ? 0u : d.nb;
    d.chunk0 = (uint32_t)chunks;
    d.bin0 = (uint32_t)bins;
    d.hist0 = hist;
    d.off0 = off;
    d.pairs0 = pairs0;
    d.pairs_extra = skipped ? (*skip_pairs)[i] : 0ull;
    pairs0 += d.pairs_extra;
    chunks += nchunks;
    bins += nb;
    if (nchunks) bp.nb_max = std::max(bp.nb_max, d.nb);
    hist += (uint64_t)nchunks * pbig::hist_pitch(d.nb);
    off += (uint64_t)nchunks * pbig::off_pitch(d.nb);
    bp.n_boxes += skipped ? 0u : d.n;
  }
  if (chunks > 0x7FFFFFFFull || bins > 0x7FFFFFFFull) return false;
  st.chunk_inst.assign(chunks, 0u);
  st.bin_inst.assign(bins, 0u);
  for (size_t i = 0; i < st.inst.size(); ++i) {
    const InstDesc& d = st.inst[i];
    for (uint32_t w = 0; w < d.nchunks; ++w) st.chunk_inst[d.chunk0 + w] = (uint32_t)i;
    if (d.n)
      for (uint32_t b = 0; b < d.nb; ++b) st.bin_inst[d.bin0 + b] = (uint32_t)i;
  }
  bp.n_chunks = (uint32_t)chunks;
  bp.n_bins = (uint32_t)bins;
  bp.hist_entries = hist;
  bp.off_entries = off;
  bp.extra_total = pairs0;
  bp.test_blocks = std::max<uint32_t>(1u, std::min<uint32_t>((uint32_t)n_cu, (uint32_t)std::max<uint64_t>(bins, 1)));
  bp.n_cu = n_cu;
  return true;
}

// Device buffers of a binned state for its current plan: tables, histogram
// rows, per-bin arrays, pair regions (cap_wg each) and the bucket.
int ensure_state_(pbig_hip_ctx* c, BinnedState& st, uint64_t bucket);

int ensure_state(pbig_hip_ctx* c, BinnedState& st, uint64_t bucket) {
  const BinnedWorkspace before = st.ws;
  const int rc = ensure_state_(c, st, bucket);
  const BinnedWorkspace& w = st.ws;
  if (w.inst != before.inst || w.chunk_inst != before.chunk_inst || w.bin_inst != before.bin_inst ||
      w.hist != before.hist || w.off != before.off || w.total != before.total || w.scratch != before.scratch ||
      w.bbox != before.bbox || w.cap_scratch != before.cap_scratch)
    ++st.generation;
  return rc;
}

int ensure_state_(pbig_hip_ctx* c, BinnedState& st, uint64_t bucket) {
  BinnedWorkspace& w = st.ws;
  const BatchPlan& bp = st.bp;
  int rc;
  if ((rc = grow(c, &w.inst, &w.cap_inst, std::max<uint64_t>(bp.n_inst, 1)))) return rc;
  if ((rc = grow(c, &w.chunk_inst, &w.cap_chunks, std::max<uint64_t>(bp.n_chunks, 1)))) return rc;
  if ((rc = grow(c, &w.bin_inst, &w.cap_bin_inst, std::max<uint64_t>(bp.n_bins, 1)))) return rc;
  if ((rc = grow(c, &w.hist, &w.cap_hist, std::max<uint64_t>(bp.hist_entries, 1)))) return rc;
  if ((rc = grow(c, &w.off, &w.cap_off, std::max<uint64_t>(bp.off_entries, 1)))) return rc;
  const uint64_t nb = (uint64_t)bp.n_bins + 1;
  if (w.cap_bins < nb || !w.total) {
    for (void** q : {(void**)&w.total, (void**)&w.bin_start, (void**)&w.overfull, (void**)&w.order,
                     (void**)&w.rank, (void**)&w.bin_seg, (void**)&w.bin_out})
      if (*q) HIP_TRY(c, hipFree(*q));
    w.total = w.bin_start = w.overfull = w.order = w.rank = nullptr;
    w.bin_seg = nullptr;
    w.bin_out = nullptr;
    w.cap_bins = 0;
    uint64_t d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0, d7 = 0;
    if ((rc = grow(c, &w.total, &d1, nb))) return rc;
    if ((rc = grow(c, &w.bin_start, &d2, nb))) return rc;
    if ((rc = grow(c, &w.overfull, &d3, nb))) return rc;
    if ((rc = grow(c, &w.order, &d6, nb))) return rc;
    if ((rc = grow(c, &w.rank, &d7, nb))) return rc;
    if ((rc = grow(c, &w.bin_seg, &d4, nb))) return rc;
    if ((rc = grow(c, &w.bin_out, &d5, nb))) return rc;
    w.cap_bins = nb;
  }
  const uint64_t scratch = (uint64_t)bp.test_blocks * w.cap_wg;
  if ((rc = grow(c, &w.scratch, &w.cap_scratch, std::max<uint64_t>(scratch, 1)))) return rc;
  // Bucket offsets, scatter cursors and test ranges are 32-bit: a launch with
  // more memberships than 2^32-1 exceeds every bucket (the scan flags
  // kErrWorkspace) and its regrowth stops here.
  if (bucket > 0xFFFFFFFFull)
    return fail(c, PBIG_E_NOMEM, "%llu bin memberships exceed the binned path's 2^32-1",
                (unsigned long long)bucket);
  if (w.cap_bucket < bucket || !w.bbox) {
    if (w.bbox) HIP_TRY(c, hipFree(w.bbox));
    if (w.bid) HIP_TRY(c, hipFree(w.bid));
    w.bbox = nullptr;
    w.bid = nullptr;
    w.cap_bucket = 0;
    uint64_t d1 = 0, d2 = 0;
    if ((rc = grow(c, &w.bbox, &d1, std::max<uint64_t>(bucket, 1)))) return rc;
    if ((rc = grow(c, &w.bid, &d2, std::max<uint64_t>(bucket, 1)))) return rc;
    w.cap_bucket = std::max<uint64_t>(bucket, 1);
  }
  return PBIG_OK;
}

// Size a state for its plan (at least the previous need; first guesses of
// 1.25 bucket entries and, per test workgroup, twice its share of one pair
// per box) and upload its tables.  Synchronises the stream: the host tables
// are pageable memory.
int size_state(pbig_hip_ctx* c, BinnedState& st) {
  const uint64_t n = st.bp.n_boxes;
  const uint64_t bucket = std::min<uint64_t>(std::max<uint64_t>(st.ws.cap_bucket, n + n / 4 + 4096), 0xFFFFFFFFull);
  const uint64_t per = std::max<uint64_t>(8192, (n + st.bp.test_blocks - 1) / st.bp.test_blocks * 2);
  const uint32_t cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, std::max<uint64_t>(st.ws.cap_wg, per));
  if (cap_wg != st.ws.cap_wg) ++st.generation;
  st.ws.cap_wg = cap_wg;
  return ensure_state(c, st, bucket);
}

void free_state(BinnedState& st) {
  if (st.stage_ev) (void)hipEventSynchronize(st.stage_ev);
  if (st.h_stage) (void)hipHostFree(st.h_stage);
  if (st.stage_ev) (void)hipEventDestroy(st.stage_ev);
  st.h_stage = nullptr;
  st.stage_ev = nullptr;
  st.cap_stage = 0;
  if (st.sub) {
    free_state(*st.sub);
    st.sub.reset();
  }
  for (void* p : {(void*)st.sub_out, (void*)st.sub_io, (void*)st.io_buf, (void*)st.runs})
    if (p) (void)hipFree(p);
  st.sub_out = nullptr;
  st.sub_io = nullptr;
  st.io_buf = nullptr;
  st.runs = nullptr;
  st.cap_sub_out = st.cap_sub_io = st.cap_io_buf = st.cap_runs = 0;
  const BinnedWorkspace& b = st.ws;
  for (void* p : {(void*)b.inst, (void*)b.chunk_inst, (void*)b.bin_inst, (void*)b.hist, (void*)b.off,
                  (void*)b.total, (void*)b.bin_start, (void*)b.overfull, (void*)b.order, (void*)b.rank,
                  (void*)b.bin_seg, (void*)b.bin_out,
                  (void*)b.scratch, (void*)b.bbox, (void*)b.bid})
    if (p) (void)hipFree(p);
  if (st.own_meta && b.meta) (void)hipFree(b.meta);
  st.ws = BinnedWorkspace{};
  for (void* p : {(void*)st.fill[0], (void*)st.fill[1], (void*)st.meta1, (void*)st.start1, (void*)st.d_res,
                  (void*)st.dense_meta, (void*)st.dense_list, (void*)st.dense_pairs, (void*)st.dense_chunks})
    if (p) (void)hipFree(p);
  st.d_res = nullptr;
  st.dense_meta = nullptr;
  st.dense_list = nullptr;
  st.dense_pairs = nullptr;
  st.dense_chunks = nullptr;
  st.cap_dense_pairs = st.cap_dense_chunks = 0;
  for (auto& lv : st.dl)
    if (lv) {
      free_state(*lv);
      lv.reset();
    }
  for (auto& k : st.dcaps) k = pbig::DeepCaps{};
  st.deep_ok = st.deep_failed = false;
  st.fill[0] = st.fill[1] = st.start1 = nullptr;
  st.meta1 = nullptr;
  st.cap_fill = st.cap_start1 = 0;
  st.onepass_dirty = true;
}

int upload_tables(pbig_hip_ctx* c, BinnedState& st, hipStream_t s) {
  const BatchPlan& bp = st.bp;
  const size_t b_inst = bp.n_inst * sizeof(InstDesc), b_chunks = bp.n_chunks * sizeof(uint32_t),
               b_bins = bp.n_bins * sizeof(uint32_t);
  const size_t bytes = b_inst + b_chunks + b_bins;
  if (!st.stage_ev) HIP_TRY(c, hipEventCreateWithFlags(&st.stage_ev, hipEventDisableTiming));
  else HIP_TRY(c, hipEventSynchronize(st.stage_ev));  // the staging buffer is free again
  if (st.cap_stage < bytes) {
    if (st.h_stage) HIP_TRY(c, hipHostFree(st.h_stage));
    st.h_stage = nullptr;
    st.cap_stage = 0;
    HIP_TRY(c, hipHostMalloc((void**)&st.h_stage, bytes + bytes / 2, hipHostMallocDefault));
    st.cap_stage = bytes + bytes / 2;
  }
  unsigned char* p = st.h_stage;
  if (b_inst) memcpy(p, st.inst.data(), b_inst);
  if (b_chunks) memcpy(p + b_inst, st.chunk_inst.data(), b_chunks);
  if (b_bins) memcpy(p + b_inst + b_chunks, st.bin_inst.data(), b_bins);
  if (b_inst) HIP_TRY(c, hipMemcpyAsync(st.ws.inst, p, b_inst, hipMemcpyHostToDevice, s));
  if (b_chunks) HIP_TRY(c, hipMemcpyAsync(st.ws.chunk_inst, p + b_inst, b_chunks, hipMemcpyHostToDevice, s));
  if (b_bins) HIP_TRY(c, hipMemcpyAsync(st.ws.bin_inst, p + b_inst + b_chunks, b_bins, hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipEventRecord(st.stage_ev, s));
  return PBIG_OK;
}

// Order stream s after the last table upload (which may have gone on another
// stream).
int after_upload(pbig_hip_ctx* c, BinnedState& st, hipStream_t s) {
  if (st.stage_ev) HIP_TRY(c, hipStreamWaitEvent(s, st.stage_ev, 0));
  return PBIG_OK;
}

// Plain path: the context's batch of one, re-planned only when the instance
// shape or the configuration changes.
int plan_plain(pbig_hip_ctx* c, uint32_t n, double range, float cell, hipStream_t s, bool* ok) {
  BinnedState& st = c->plain;
  *ok = true;
  if (st.key_valid && st.key_n == n && st.key_range == range && st.key_cell == cell &&
      st.key_target == c->cfg.bin_target)
    return PBIG_OK;
  st.key_valid = false;
  InstDesc d;
  if (n == 0 || !plan_instance(n, range, cell, c->cfg.bin_target, &d, (uint32_t)c->n_cu)) {
    *ok = false;
    return PBIG_OK;
  }
  st.inst.assign(1, d);
  if (!layout_tables(st, (uint32_t)c->n_cu, nullptr)) {
    *ok = false;
    return PBIG_OK;
  }
  // Single-pass eligibility: a full bin's expected memberships (boxes per
  // bin area, 10% for boxes crossing into neighbouring bins) plus eight
  // standard deviations fit what one test workgroup holds.
  {
    const double nf0 = std::max(1.0, std::ceil(range / (double)cell));
    const double e = 1.1 * (double)n * (double)d.k * d.k / (nf0 * nf0);
    const double want = e + 8.0 * std::sqrt(e) + 64.0;
    st.bin_cap = want <= (double)pbig::kNarrowMax
                     ? std::min<uint32_t>(pbig::kNarrowMax, ((uint32_t)std::ceil(want) + 63u) & ~63u)
                     : 0u;
    st.onepass_failed = false;
    st.deep_ok = st.deep_failed = false;
    st.to_general = st.dense_levels = false;
  }
  st.ws.meta = c->ws.meta;
  int rc;
  if ((rc = size_state(c, st))) return rc;
  // A batch of one: every chunk and bin belongs to instance 0, so the
  // chunk/bin tables only have to be zero; the instance record goes through
  // the pinned ring, ordered on the stream behind earlier launches.
  if (st.zero_chunks != st.ws.cap_chunks || st.zero_bins != st.ws.cap_bin_inst) {
    HIP_TRY(c, hipMemsetAsync(st.ws.chunk_inst, 0, st.ws.cap_chunks * sizeof(uint32_t), s));
    HIP_TRY(c, hipMemsetAsync(st.ws.bin_inst, 0, st.ws.cap_bin_inst * sizeof(uint32_t), s));
    st.zero_chunks = st.ws.cap_chunks;
    st.zero_bins = st.ws.cap_bin_inst;
  }
  const uint32_t slot = c->ring_next++ % 16u;
  HIP_TRY(c, hipEventSynchronize(c->ring_ev[slot]));
  c->h_ring[slot] = st.inst[0];
  HIP_TRY(c, hipMemcpyAsync(st.ws.inst, &c->h_ring[slot], sizeof(InstDesc), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipEventRecord(c->ring_ev[slot], s));
  st.key_valid = true;
  st.key_n = n;
  st.key_range = range;
  st.key_cell = cell;
  st.key_target = c->cfg.bin_target;
  return PBIG_OK;
}

hipEvent_t* next_events(pbig_hip_ctx* c, int path) {
  if (c->timing_used >= c->timing_cap) return nullptr;
  c->timing_path = path;
  return &c->events[(size_t)c->timing_used++ * kEventsPerCall];
}

// The general path's tile test on a cell layout built from the bins of the
// plain state's plan (launch_sorted), synchronously: the bucket, the
// membership array and the item array grow until the launch fits.
int size_sorted(pbig_hip_ctx* c, uint32_t n, hipStream_t s) {
  const InstDesc& d = c->plain.inst[0];
  int rc;
  // cell starts (bins x k^2), and bin ends then the large bins' cell cursors (bins + bins x k^2)
  if ((rc = ensure_cells(c, (uint64_t)d.nb + (uint64_t)d.nb * d.k * d.k + 1))) return rc;
  if ((rc = grow(c, &c->ws.strips, &c->ws.cap_strips,
                 std::max<uint64_t>(c->ws.cap_strips, 2ull * d.nb + (uint64_t)n / 1024 + 1024))))
    return rc;
  // the bins' slice counters, a ticket counter per 1,024 bins, then per bin
  // a row of fine counts that the counting slices of a large bin add to
  // (zero between launches: the kernel takes each sum with an exchange
  // against zero)
  const uint64_t want_slc = (uint64_t)d.nb + (d.nb + 1023u) / 1024u + (uint64_t)d.nb * d.k * d.k;
  if (c->ws.cap_slc < want_slc || !c->ws.slc) {
    if ((rc = grow(c, &c->ws.slc, &c->ws.cap_slc, want_slc))) return rc;
    HIP_TRY(c, hipMemsetAsync(c->ws.slc, 0, c->ws.cap_slc * sizeof(uint32_t), s));
  }
  return ensure_work(c, std::max<uint64_t>(c->ws.cap_memb, (uint64_t)n + n / 2 + 1024),
                     std::max<uint64_t>(c->ws.cap_items, (uint64_t)n / 2 + 1024));
}

pbig::BatchPlan sorted_plan(const BinnedState& st) {
  pbig::BatchPlan bp = st.bp;
  bp.dense_exit = 0u;  // every bin is laid out, dense or not
  bp.direct = 0u;
  return bp;
}

int run_sorted_sync(pbig_hip_ctx* c, const float4* d_boxes, uint32_t n, uint2* d_out, uint64_t cap, hipStream_t s,
                    uint64_t* n_pairs) {
  BinnedState& st = c->plain;
  int rc;
  if ((rc = size_sorted(c, n, s))) return rc;
  if ((rc = after_upload(c, st, s))) return rc;
  const pbig::BatchPlan bp = sorted_plan(st);
  for (int attempt = 0; attempt < 3; ++attempt) {
    HIP_TRY(c, pbig::launch_sorted(st.ws, d_boxes, bp, c->ws, c->sort_meta, d_out, cap, s,
                                   next_events(c, kPathSorted)));
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, st.ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(c->h_meta + 1, c->sort_meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    const Meta mb = c->h_meta[0], m = c->h_meta[1];
    if ((mb.error | m.error) & pbig::kErrInvalidBox)
      return fail(c, PBIG_E_INVALID_BOX, "input has a box with xmin > xmax, ymin > ymax or a NaN");
    if (m.error & pbig::kErrTooDense)
      return fail(c, PBIG_E_TOO_DENSE, "a cell holds %u boxes, above the %u-tile limit", m.max_per_cell,
                  pbig::kMaxTilesPerCell);
    if ((mb.error | m.error) & pbig::kErrWorkspace) {
      if (attempt == 2) return fail(c, PBIG_E_NOMEM, "workspace did not fit after regrowth");
      if (mb.error & pbig::kErrWorkspace) {  // the bucket (the cell layout was not built)
        uint64_t bucket = std::max<uint64_t>(st.ws.cap_bucket, mb.memberships + mb.memberships / 16 + 4096);
        if (mb.memberships <= 0xFFFFFFFFull) bucket = std::min<uint64_t>(bucket, 0xFFFFFFFFull);
        if ((rc = ensure_state(c, st, bucket))) return rc;
      } else {
        if ((rc = ensure_work(c, std::max<uint64_t>(c->ws.cap_memb, m.memberships + m.memberships / 16 + 1024),
                              std::max<uint64_t>(c->ws.cap_items, m.items + m.items / 16 + 1024))))
          return rc;
        if ((rc = grow(c, &c->ws.strips, &c->ws.cap_strips,
                       std::max<uint64_t>(c->ws.cap_strips, (uint64_t)m.dev_chunks + m.dev_chunks / 8 + 1024))))
          return rc;
      }
      continue;
    }
    const InstDesc& d = st.inst[0];
    c->stats.n_boxes = n;
    c->stats.n_cells = (uint64_t)d.nf * d.nf;
    c->stats.cells_per_row = d.nf;
    c->stats.memberships = m.memberships;
    c->stats.tile_pairs = m.items;
    c->stats.n_pairs = m.n_pairs;
    c->stats.max_per_cell = m.max_per_cell;
    *n_pairs = m.n_pairs;
    if (m.n_pairs > cap)
      return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu",
                  (unsigned long long)m.n_pairs, (unsigned long long)cap);
    return PBIG_OK;
  }
  return fail(c, PBIG_E_HIP, "unreachable");
}

// Run a binned state synchronously, regrowing the workspace when the bucket
// or a pair region overflowed.  On success *m holds the result words; the
// caller handles kErrOverfull.
int run_state_sync(pbig_hip_ctx* c, BinnedState& st, const float4* d_boxes, const uint32_t* ids, uint2* d_out,
                   uint64_t cap, uint64_t* d_inst_out, hipStream_t s, Meta* m) {
  int rc;
  if ((rc = after_upload(c, st, s))) return rc;
  for (int attempt = 0; attempt < 3; ++attempt) {
    HIP_TRY(c, pbig::launch_binned(st.ws, d_boxes, ids, st.bp, d_out, cap, d_inst_out, s,
                                   next_events(c, PBIG_PATH_BINNED)));
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, st.ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    *m = *c->h_meta;
    if (m->error & pbig::kErrInvalidBox)
      return fail(c, PBIG_E_INVALID_BOX, "input has a box with xmin > xmax, ymin > ymax or a NaN");
    if (m->error & pbig::kErrDense) return PBIG_OK;  // the caller runs the general path
    if (m->error & pbig::kErrWorkspace) {
      if (attempt == 2) return fail(c, PBIG_E_NOMEM, "bin workspace did not fit after regrowth");
      if (m->items > st.ws.cap_wg)
        st.ws.cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, m->items + m->items / 4 + 1024);
      uint64_t bucket = std::max<uint64_t>(st.ws.cap_bucket, m->memberships + m->memberships / 16 + 4096);
      if (m->memberships <= 0xFFFFFFFFull) bucket = std::min<uint64_t>(bucket, 0xFFFFFFFFull);
      if ((rc = ensure_state(c, st, bucket))) return rc;
      continue;
    }
    return PBIG_OK;
  }
  return fail(c, PBIG_E_HIP, "unreachable");
}

// Buffers of the plain state's single-pass launch for its current plan.
// *ok is false when the plan runs the two-pass launch.
int prepare_onepass(pbig_hip_ctx* c, BinnedState& st, hipStream_t s, bool* ok) {
  *ok = st.bin_cap != 0 && !st.onepass_failed && st.bp.n_inst == 1 && st.bp.n_bins > 0;
  if (!*ok) return PBIG_OK;
  const uint32_t nb = st.bp.n_bins;
  int rc;
  if ((rc = ensure_state(c, st, std::max<uint64_t>(st.ws.cap_bucket, (uint64_t)nb * st.bin_cap)))) return rc;
  if (st.cap_fill < nb || !st.fill[0] || !st.fill[1]) {
    for (int k = 0; k < 2; ++k) {
      if (st.fill[k]) HIP_TRY(c, hipFree(st.fill[k]));
      st.fill[k] = nullptr;
    }
    st.cap_fill = 0;
    const uint64_t want = std::max<uint64_t>(nb, 1024);
    uint64_t d0 = 0, d1 = 0;
    if ((rc = grow(c, &st.fill[0], &d0, want)) || (rc = grow(c, &st.fill[1], &d1, want))) return rc;
    st.cap_fill = want;
    st.onepass_dirty = true;
  }
  if (!st.meta1) {
    HIP_TRY(c, hipMalloc((void**)&st.meta1, 2 * sizeof(Meta)));
    st.onepass_dirty = true;
  }
  if ((rc = grow(c, &st.start1, &st.cap_start1, nb))) return rc;  // filled by the place kernel
  if (st.onepass_dirty) {
    HIP_TRY(c, hipMemsetAsync(st.fill[0], 0, st.cap_fill * sizeof(uint32_t), s));
    HIP_TRY(c, hipMemsetAsync(st.fill[1], 0, st.cap_fill * sizeof(uint32_t), s));
    HIP_TRY(c, hipMemsetAsync(st.meta1, 0, 2 * sizeof(Meta), s));
    st.parity = 0;
    st.onepass_dirty = false;
  }
  return PBIG_OK;
}

// Enqueue the single-pass launch; *dm receives the device address of its
// result words.
int launch_onepass(pbig_hip_ctx* c, BinnedState& st, const float4* d_boxes, uint2* d_out, uint64_t cap,
                   hipStream_t s, Meta** dm) {
  const uint32_t p = st.parity & 1u;
  pbig::OnePass op;
  op.fill = st.fill[p];
  op.fill_next = st.fill[p ^ 1u];
  op.fill_words = (uint32_t)st.cap_fill;
  op.bin_start = st.start1;
  op.bin_cap = st.bin_cap;
  op.meta = st.meta1 + p;
  op.meta_next = st.meta1 + (p ^ 1u);
  op.dense_exit = st.bp.dense_exit;
  const hipError_t e = pbig::launch_binned_onepass(st.ws, op, d_boxes, st.bp, d_out, cap, s,
                                                   next_events(c, kPathOnePass));
  if (e != hipSuccess) {
    st.onepass_dirty = true;  // the counter sets may not have been cleared for the next call
    HIP_TRY(c, e);
  }
  st.parity ^= 1u;
  *dm = op.meta;
  return PBIG_OK;
}

// The single-pass launch, synchronously.  0: done, *m holds the result
// words; 1: the plan runs the two-pass launch instead (not eligible, or the
// launch reported trouble, which also marks the plan); 2: a bin outgrew its
// share and the automatic choice's dense rule holds (the general path runs
// instead); < 0: an error.
int run_onepass_sync(pbig_hip_ctx* c, const float4* d_boxes, uint2* d_out, uint64_t cap, hipStream_t s, Meta* m) {
  BinnedState& st = c->plain;
  int rc;
  for (int attempt = 0; attempt < 3; ++attempt) {
    bool ok = false;
    if ((rc = prepare_onepass(c, st, s, &ok))) return rc;
    if (!ok) return 1;
    Meta* dm = nullptr;
    if ((rc = launch_onepass(c, st, d_boxes, d_out, cap, s, &dm))) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, dm, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    *m = *c->h_meta;
    if (m->error & pbig::kErrInvalidBox)
      return fail(c, PBIG_E_INVALID_BOX, "input has a box with xmin > xmax, ymin > ymax or a NaN");
    if (m->error == pbig::kErrWorkspace && m->items > st.ws.cap_wg) {  // a pair region overflowed
      st.ws.cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, m->items + m->items / 4 + 1024);
      if ((rc = ensure_state(c, st, st.ws.cap_bucket))) return rc;
      continue;
    }
    if (m->error & pbig::kErrDense) {  // (the automatic choice: mostly dense clusters)
      st.onepass_failed = st.to_general = true;
      return 2;
    }
    if (m->error) {
      st.onepass_failed = true;
      return 1;
    }
    return 0;
  }
  st.onepass_failed = true;
  return 1;
}

// ---------------------------------------------------------------------------
// The plain path's two-pass launch followed by re-binning levels planned on
// the device (no host work between levels).

// Workspace of device-planned level l (1..) for its bounds; its result words
// point at level 0's pair counter.
int ensure_deep_level(pbig_hip_ctx* c, BinnedState& st, uint32_t l, uint64_t bucket) {
  if (!st.dl[l]) {
    st.dl[l] = std::make_unique<BinnedState>();
    BinnedState& L = *st.dl[l];
    HIP_TRY(c, hipMalloc((void**)&L.ws.meta, sizeof(Meta)));
    L.own_meta = true;
    Meta h{};
    h.out_count = (unsigned long long*)&st.ws.meta->n_pairs;
    HIP_TRY(c, hipMemcpy(L.ws.meta, &h, sizeof(Meta), hipMemcpyHostToDevice));
  }
  BinnedState& L = *st.dl[l];
  const pbig::DeepCaps& k = st.dcaps[l];
  L.bp = BatchPlan{};
  L.bp.n_inst = k.inst;
  L.bp.n_chunks = k.chunks;
  L.bp.n_bins = k.bins;
  L.bp.hist_entries = k.hist;
  L.bp.off_entries = k.off;
  L.bp.test_blocks = (uint32_t)std::max(c->n_cu, 1);
  L.bp.n_cu = (uint32_t)c->n_cu;
  if (!L.ws.cap_wg) L.ws.cap_wg = (uint32_t)std::max<uint64_t>(8192, 2 * bucket / L.bp.test_blocks);
  const uint64_t gen = L.generation;
  const int rc = ensure_state(c, L, bucket);
  // A graph captured with the levels holds their arrays' addresses: any
  // reallocation invalidates it like one of level 0's.
  if (L.generation != gen) ++st.generation;
  return rc;
}

// Cap on the device levels' cell refinement (powers of two below the
// instance's own cell).  PBIG_DEEP_REFINE overrides it (for measurements).
uint32_t deep_refine() {
  static const uint32_t v = [] {
    const char* e = getenv("PBIG_DEEP_REFINE");
    return e ? (uint32_t)atoi(e) : 2u;
  }();
  return v;
}

// First bounds of the levels, from level 0's plan and bucket; the
// synchronous driver grows them to what the levels report.
void init_deep_caps(BinnedState& st) {
  const uint64_t m0 = std::max<uint64_t>(st.ws.cap_bucket, 4096);
  for (uint32_t l = 1; l < pbig::kDeepLevels; ++l) {
    pbig::DeepCaps& k = st.dcaps[l];
    k.inst = std::min<uint32_t>(2048, std::max<uint32_t>(64, st.bp.n_bins / 4));
    k.chunks = (uint32_t)std::min<uint64_t>(0x7FFFFFFF, m0 / 4096 + k.inst);
    k.bins = k.inst * 16;
    k.nb_max = 1024;
    k.hist = (uint64_t)k.chunks * pbig::hist_pitch(64);
    k.off = (uint64_t)k.chunks * pbig::off_pitch(64);
  }
}

// The dense kernel's view of a state's levels (level 0 and the device levels).
pbig::DenseSrc dense_src(const BinnedState& st) {
  pbig::DenseSrc d{};
  for (uint32_t l = 0; l < pbig::kDeepLevels; ++l) {
    const BinnedWorkspace& w = l ? st.dl[l]->ws : st.ws;
    d.bbox[l] = w.bbox;
    d.bid[l] = w.bid;
    d.bin_start[l] = w.bin_start;
    d.total[l] = w.total;
    d.inst[l] = w.inst;
    d.bin_inst[l] = w.bin_inst;
  }
  const BinnedWorkspace& last = st.dl[pbig::kDeepLevels - 1]->ws;
  d.last_overfull = last.overfull;
  d.last_meta = last.meta;
  d.list = st.dense_list;
  return d;
}

// The dense kernel's status words and list (once), and for a batch its pair
// and run buffers at least `pairs` and `chunks` entries.
int ensure_dense(pbig_hip_ctx* c, BinnedState& st, uint64_t pairs, uint64_t chunks) {
  if (!st.dense_meta) {
    HIP_TRY(c, hipMalloc((void**)&st.dense_meta, sizeof(Meta)));
    HIP_TRY(c, hipMemset(st.dense_meta, 0, sizeof(Meta)));
  }
  uint64_t dl = 0;
  if (!st.dense_list) {
    int rc;
    if ((rc = grow(c, &st.dense_list, &dl, pbig::kDenseMaxBins))) return rc;
  }
  const uint64_t before_p = st.cap_dense_pairs, before_c = st.cap_dense_chunks;
  int rc;
  if (pairs && (rc = grow(c, &st.dense_pairs, &st.cap_dense_pairs, pairs))) return rc;
  if (chunks && (rc = grow(c, &st.dense_chunks, &st.cap_dense_chunks, chunks))) return rc;
  if (st.cap_dense_pairs != before_p || st.cap_dense_chunks != before_c) ++st.generation;  // (captured graphs)
  return PBIG_OK;
}

// Enqueue level 0 (the two-pass launch, direct output) and the levels, then
// the bins no level resolves, then the result words into d_result (optional).
int enqueue_deep(pbig_hip_ctx* c, BinnedState& st, const float4* d_boxes, uint2* d_out, uint64_t cap,
                 uint64_t* d_result, hipStream_t s) {
  hipEvent_t* ev = next_events(c, kPathDeep);
  if (ev) HIP_TRY(c, hipEventRecord(ev[0], s));
  st.bp.direct = 1u;
  // dense_exit as the caller set it (a dense instance stops after level 0's
  // bin totals), unless the automatic choice sent a dense instance here
  pbig::BatchPlan bp0 = st.bp;
  if (st.dense_levels) bp0.dense_exit = 0u;
  HIP_TRY(c, pbig::launch_binned(st.ws, d_boxes, nullptr, bp0, d_out, cap, nullptr, s, nullptr));
  if (ev) HIP_TRY(c, hipEventRecord(ev[1], s));
  HIP_TRY(c, pbig::launch_dense_clear(st.dense_meta, s));
  const BinnedWorkspace* parent = &st.ws;
  uint32_t parent_bins = st.bp.n_bins;
  pbig::DeepMetas dm{};
  dm.meta[0] = st.ws.meta;
  const uint32_t target = std::max<uint32_t>(c->cfg.bin_target / 2, 64);
  for (uint32_t l = 1; l < pbig::kDeepLevels; ++l) {
    const BinnedState& L = *st.dl[l];
    HIP_TRY(c, pbig::launch_deep_level(*parent, parent_bins, l > 1, L.ws, st.dcaps[l], target, st.inst[0].inv_f,
                                       nullptr, deep_refine(), (uint32_t)c->n_cu, 1u, d_out, cap, l - 1,
                                       st.dense_list, st.dense_meta, s));
    if (ev) HIP_TRY(c, hipEventRecord(ev[l + 1], s));
    dm.meta[l] = L.ws.meta;
    parent = &L.ws;
    parent_bins = st.dcaps[l].bins;
  }
  HIP_TRY(c, pbig::launch_dense_bins(dense_src(st), pbig::kDeepLevels - 1, st.dense_meta, (uint32_t)c->n_cu,
                                     (unsigned long long*)&st.ws.meta->n_pairs, d_out, cap, nullptr, 0, nullptr, 0,
                                     s));
  dm.levels = pbig::kDeepLevels;
  if (d_result) HIP_TRY(c, pbig::launch_deep_result(dm, st.dense_meta, d_result, s));
  return PBIG_OK;
}

// A launch with device levels of state st (one instance or a batch),
// synchronously: enqueue(d_res) enqueues it with its result words in d_res.
// 0: done (*m: pair count, level 0's memberships); 1: the levels cannot
// complete this plan (the host re-binning runs instead); 2: the automatic
// choice's dense rule stopped level 0 (the general path runs instead); < 0:
// an error.  Grows every level's workspace to what it reports and runs
// again (a few times at most).
template <class Enqueue>
int run_levels_sync(pbig_hip_ctx* c, BinnedState& st, hipStream_t s, Meta* m, bool batch, Enqueue enqueue) {
  int rc;
  if (!st.d_res) HIP_TRY(c, hipMalloc((void**)&st.d_res, 2 * sizeof(uint64_t)));
  if (!st.dcaps[1].inst) init_deep_caps(st);
  std::vector<uint64_t> bucket(pbig::kDeepLevels, 0);
  for (uint32_t l = 1; l < pbig::kDeepLevels; ++l)
    bucket[l] = st.dl[l] ? st.dl[l]->ws.cap_bucket : 2 * std::max<uint64_t>(st.ws.cap_bucket, 4096);
  // The levels' arrays always match their bounds (dcaps, each level's
  // cap_wg and bucket): a later launch trusts the bounds, so every growth
  // below is allocated before the next attempt or the return.
  uint64_t dense_pairs = std::max<uint64_t>(st.cap_dense_pairs, batch ? 1u << 20 : 0u);
  uint64_t dense_chunks = std::max<uint64_t>(st.cap_dense_chunks, batch ? 1u << 14 : 0u);
  auto allocate = [&]() -> int {
    int r0 = ensure_dense(c, st, dense_pairs, dense_chunks);
    if (r0) {
      st.deep_ok = false;
      st.deep_failed = true;
      return r0;
    }
    for (uint32_t l = 1; l < pbig::kDeepLevels; ++l) {
      const int r = ensure_deep_level(c, st, l, bucket[l]);
      if (r) {  // (a level's arrays may be short of its bounds: no launch may take the levels)
        st.deep_ok = false;
        st.deep_failed = true;
        return r;
      }
    }
    return PBIG_OK;
  };
  if ((rc = allocate())) return rc;
  for (int attempt = 0; attempt < 6; ++attempt) {
    if ((rc = after_upload(c, st, s))) return rc;
    if ((rc = enqueue(st.d_res))) return rc;
    uint64_t res[2];
    std::vector<Meta> lm(pbig::kDeepLevels);
    HIP_TRY(c, hipMemcpyAsync(res, st.d_res, sizeof(res), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&lm[0], st.ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    for (uint32_t l = 1; l < pbig::kDeepLevels; ++l)
      HIP_TRY(c, hipMemcpyAsync(&lm[l], st.dl[l]->ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    Meta dn{};
    HIP_TRY(c, hipMemcpyAsync(&dn, st.dense_meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    if (res[1] == 0) {
      *m = lm[0];
      m->n_pairs = res[0];
      m->error = 0;
      st.deep_ok = true;  // (the bounds stay: instances of other shapes follow on this context)
      return 0;
    }
    if (res[1] & pbig::kErrInvalidBox)
      return fail(c, PBIG_E_INVALID_BOX, "input has a box with xmin > xmax, ymin > ymax or a NaN");
    if (res[1] & pbig::kErrDense) {
      st.to_general = true;
      return 2;
    }
    bool grown = false;
    if (dn.error & pbig::kErrWorkspace) {  // the dense kernel's pairs or runs (a batch)
      dense_pairs = std::max<uint64_t>(dense_pairs, dn.n_pairs + dn.n_pairs / 4 + 4096);
      dense_chunks = std::max<uint64_t>(dense_chunks, dn.items + dn.items / 4 + 1024);
      grown = true;
    }
    // level 0: bucket and pair regions
    if (lm[0].error & pbig::kErrWorkspace) {
      if (lm[0].items > st.ws.cap_wg)
        st.ws.cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, lm[0].items + lm[0].items / 4 + 1024);
      uint64_t b0 = std::max<uint64_t>(st.ws.cap_bucket, lm[0].memberships + lm[0].memberships / 16 + 4096);
      if (lm[0].memberships <= 0xFFFFFFFFull) b0 = std::min<uint64_t>(b0, 0xFFFFFFFFull);
      if ((rc = ensure_state(c, st, b0))) return rc;
      grown = true;
    }
    for (uint32_t l = 1; l < pbig::kDeepLevels; ++l) {
      const Meta& x = lm[l];
      pbig::DeepCaps& k = st.dcaps[l];
      if (x.plan_err & 1u) {
        if (x.dev_inst > pbig::kDeepMaxInst) continue;  // more bins beyond LDS than a level takes
        k.inst = std::max(k.inst, x.dev_inst + x.dev_inst / 4 + 16);
        k.chunks = std::max(k.chunks, x.dev_chunks + x.dev_chunks / 4 + 64);
        k.bins = std::max(k.bins, x.dev_bins + x.dev_bins / 4 + 256);
        k.hist = std::max<uint64_t>(k.hist, x.dev_hist + x.dev_hist / 4 + 4096);
        k.off = std::max<uint64_t>(k.off, x.dev_off + x.dev_off / 4 + 4096);
        grown = true;
      }
      if (x.error & pbig::kErrWorkspace) {
        BinnedState& L = *st.dl[l];
        if (x.items > L.ws.cap_wg) {
          L.ws.cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, x.items + x.items / 4 + 1024);
          grown = true;
        }
        if (x.memberships > L.ws.cap_bucket) {
          if (x.memberships > 0xFFFFFFFFull) continue;
          bucket[l] = std::min<uint64_t>(x.memberships + x.memberships / 8 + 4096, 0xFFFFFFFFull);
          grown = true;
        }
      }
    }
    if ((rc = allocate())) return rc;
    if (!grown) break;  // bins the levels cannot split or finish
  }
  st.deep_failed = true;
  return 1;
}

// The plain path's two-pass launch followed by levels on the device,
// synchronously (run_levels_sync's codes).
int run_deep_sync(pbig_hip_ctx* c, const float4* d_boxes, uint2* d_out, uint64_t cap, hipStream_t s, Meta* m) {
  BinnedState& st = c->plain;
  if (c->cfg.host_levels || st.deep_failed || st.bp.n_inst != 1 || st.bp.n_bins == 0) return 1;
  return run_levels_sync(c, st, s, m, false,
                         [&](uint64_t* res) { return enqueue_deep(c, st, d_boxes, d_out, cap, res, s); });
}

// Pairs of a launch's instances found outside it (by re-binning their
// heavy bins one level down, or by the general path): per instance the
// extra count and the device runs holding those pairs, in order; and the
// instances whose own pairs from the launch are to be discarded.
struct Splice {
  std::vector<uint64_t> extra;
  std::vector<std::vector<std::pair<const uint2*, uint64_t>>> src;
  std::vector<uint8_t> drop;
  explicit Splice(size_t n = 0) : extra(n, 0), src(n), drop(n, 0) {}
  bool empty() const {
    for (size_t i = 0; i < extra.size(); ++i)
      if (extra[i] || drop[i]) return false;
    return true;
  }
};

constexpr int kMaxLevels = 4;  // re-binning depth before an instance goes to the general path
constexpr size_t kFewDense = 8;  // dense instances of a batch that go to the general path rather than re-binning

// Give every instance its extra pairs (pairs0 / pairs_extra), discard the
// dropped instances' bins, rerun the pair scan and output kernels and copy
// the extra pairs after each instance's own run.  *m receives the new result
// words.  The device table is restored to the plan afterwards.
int apply_splice(pbig_hip_ctx* c, BinnedState& st, const Splice& sp, uint2* d_out, uint64_t cap, uint64_t* d_io,
                 hipStream_t s, Meta* m) {
  const size_t n = st.inst.size();
  int rc;
  if (!d_io) {
    if ((rc = grow(c, &st.io_buf, &st.cap_io_buf, 2 * std::max<uint64_t>(n, 1)))) return rc;
    d_io = st.io_buf;
  }
  std::vector<InstDesc> tab = st.inst;
  uint64_t acc = 0;
  std::vector<pbig::SpliceRun> runs;
  for (size_t i = 0; i < n; ++i) {
    tab[i].pairs0 = acc;
    tab[i].pairs_extra = sp.extra[i];
    acc += sp.extra[i];
    uint64_t within = 0;
    for (const auto& r : sp.src[i]) {
      // pieces of at most kPiece pairs, one copy block each
      constexpr uint64_t kPiece = 16384;
      for (uint64_t o = 0; o < r.second; o += kPiece)
        runs.push_back(pbig::SpliceRun{r.first + o, std::min<uint64_t>(kPiece, r.second - o), within + o,
                                       (uint32_t)i, 0u});
      within += r.second;
    }
  }
  BatchPlan bp = st.bp;
  bp.extra_total = acc;
  if ((rc = grow(c, &st.runs, &st.cap_runs, std::max<uint64_t>(runs.size(), 1)))) return rc;
  if (n) HIP_TRY(c, hipMemcpyAsync(st.ws.inst, tab.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, s));
  if (!runs.empty())
    HIP_TRY(c, hipMemcpyAsync(st.runs, runs.data(), runs.size() * sizeof(pbig::SpliceRun), hipMemcpyHostToDevice,
                              s));
  if (bp.direct) {
    // A single instance whose launch wrote its own pairs to d_out[0, n0):
    // the extra pairs go right after them (one instance, never dropped).
    const uint64_t n0 = m->n_pairs;
    const uint64_t hio[2] = {0, n0 + acc};  // alive until the synchronisation below
    HIP_TRY(c, hipMemcpyAsync(d_io, hio, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIP_TRY(c, pbig::launch_splice_copy(st.runs, (uint32_t)runs.size(), d_io, st.ws.inst, d_out, cap, s));
    if (n) HIP_TRY(c, hipMemcpyAsync(st.ws.inst, st.inst.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    m->n_pairs = n0 + acc;
    return PBIG_OK;
  }
  for (size_t i = 0; i < n; ++i)
    if (sp.drop[i] && tab[i].nb && tab[i].n)
      HIP_TRY(c, hipMemsetAsync(st.ws.bin_seg + tab[i].bin0, 0, tab[i].nb * sizeof(uint2), s));
  HIP_TRY(c, pbig::launch_binned_output(st.ws, bp, d_out, cap, d_io, s));
  HIP_TRY(c, pbig::launch_splice_copy(st.runs, (uint32_t)runs.size(), d_io, st.ws.inst, d_out, cap, s));
  HIP_TRY(c, hipMemcpyAsync(c->h_meta, st.ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
  if (n) HIP_TRY(c, hipMemcpyAsync(st.ws.inst, st.inst.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  *m = *c->h_meta;
  return PBIG_OK;
}

// Run a state synchronously and resolve its coarse bins that did not fit in
// LDS: each becomes a virtual instance of the next level (the same kernels
// over this launch's bin-ordered boxes, with smaller bins), recursively; the
// next level's pairs come back as *sp for the caller to apply_splice.
// Instances that cannot be resolved so (a single fine cell over the LDS
// buffer, or kMaxLevels reached) are marked in *failed.
// pass_done: the state's launch already ran with these results in *m (an
// asynchronous launch whose only error is kErrOverfull).
int resolve(pbig_hip_ctx* c, BinnedState& st, const float4* boxes, const uint32_t* ids, uint2* d_out,
            uint64_t cap, uint64_t* d_io, hipStream_t s, int level, Meta* m, Splice* sp,
            std::vector<uint8_t>* failed, bool pass_done = false) {
  const size_t n_inst = st.inst.size();
  failed->assign(n_inst, 0);
  *sp = Splice(n_inst);
  int rc;
  if (!pass_done && (rc = run_state_sync(c, st, boxes, ids, d_out, cap, d_io, s, m))) return rc;
  if (m->error & pbig::kErrDense) {  // a dense single instance: the general path
    failed->assign(n_inst, 1);
    return PBIG_OK;
  }
  if (!(m->error & pbig::kErrOverfull)) return PBIG_OK;
  const uint32_t n_bins = st.bp.n_bins;
  const uint32_t n_over = (uint32_t)std::min<uint64_t>(m->max_per_cell, n_bins);
  std::vector<uint32_t> over(std::max<uint32_t>(n_over, 1)), total(std::max<uint32_t>(n_bins, 1)),
      start(std::max<uint32_t>(n_bins, 1));
  if (n_over) HIP_TRY(c, hipMemcpyAsync(over.data(), st.ws.overfull, n_over * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(total.data(), st.ws.total, n_bins * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(start.data(), st.ws.bin_start, n_bins * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  std::vector<uint32_t> heavy;
  for (uint32_t k = 0; k < n_over; ++k) {
    const uint32_t b = over[k] & ~pbig::kOverfullCell;
    if (b >= n_bins) continue;
    const uint32_t i = st.bin_inst[b];
    if ((over[k] & pbig::kOverfullCell) || level + 1 >= kMaxLevels) (*failed)[i] = 1;
    else heavy.push_back(b);
  }
  std::sort(heavy.begin(), heavy.end());
  if (level == 0 && c->cfg.path == PBIG_PATH_AUTO && !heavy.empty()) {
    // Instances whose heavy bins hold more than a quarter of their
    // memberships are mostly dense clusters.  Re-binning them costs a few
    // host round trips per level (about a millisecond) however many there
    // are; the general path costs two synchronous runs (about 0.1 ms) per
    // instance.  A few such instances go to the general path; many are
    // re-binned together.
    std::vector<uint64_t> dense(n_inst, 0), all(n_inst, 0);
    for (uint32_t b : heavy) dense[st.bin_inst[b]] += total[b];
    for (uint32_t b = 0; b < n_bins; ++b) all[st.bin_inst[b]] += total[b];
    std::vector<uint32_t> clustered;
    for (size_t i = 0; i < n_inst; ++i)
      if (!(*failed)[i] && 4 * dense[i] > all[i]) clustered.push_back((uint32_t)i);
    if (clustered.size() <= kFewDense)
      for (uint32_t i : clustered) (*failed)[i] = 1;
  }
  if (!st.sub) {
    st.sub = std::make_unique<BinnedState>();
    HIP_TRY(c, hipMalloc((void**)&st.sub->ws.meta, sizeof(Meta)));
    HIP_TRY(c, hipMemset(st.sub->ws.meta, 0, sizeof(Meta)));
    st.sub->own_meta = true;
  }
  BinnedState& V = *st.sub;
  V.inst.clear();
  const uint32_t target = std::max<uint32_t>(c->cfg.bin_target / 2, 64);
  for (uint32_t b : heavy) {
    const uint32_t i = st.bin_inst[b];
    const InstDesc& P = st.inst[i];
    InstDesc v;
    if (!plan_virtual(P, i, b - P.bin0, total[b], start[b], target, &v)) (*failed)[i] = 1;
    else V.inst.push_back(v);
  }
  {
    std::vector<InstDesc> keep;
    for (const InstDesc& v : V.inst)
      if (!(*failed)[v.parent]) keep.push_back(v);
    V.inst.swap(keep);
  }
  sp->drop = *failed;
  if (V.inst.empty()) return PBIG_OK;
  if (!layout_tables(V, (uint32_t)c->n_cu, nullptr)) return fail(c, PBIG_E_ARG, "re-binning layout failed");
  if ((rc = size_state(c, V))) return rc;
  if ((rc = upload_tables(c, V, s))) return rc;
  const size_t nv = V.inst.size();
  if ((rc = grow(c, &st.sub_io, &st.cap_sub_io, 2 * nv))) return rc;
  if ((rc = grow(c, &st.sub_out, &st.cap_sub_out, std::max<uint64_t>(4096, 2 * V.bp.n_boxes)))) return rc;
  Meta m2;
  Splice sp2;
  std::vector<uint8_t> f2;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if ((rc = after_upload(c, V, s))) return rc;
    if ((rc = resolve(c, V, st.ws.bbox, st.ws.bid, st.sub_out, st.cap_sub_out, st.sub_io, s, level + 1, &m2, &sp2,
                      &f2)))
      return rc;
    if (!sp2.empty() && (rc = apply_splice(c, V, sp2, st.sub_out, st.cap_sub_out, st.sub_io, s, &m2))) return rc;
    if (m2.n_pairs <= st.cap_sub_out) break;
    if (attempt == 1) return fail(c, PBIG_E_NOMEM, "re-binned pairs did not fit after regrowth");
    if ((rc = grow(c, &st.sub_out, &st.cap_sub_out, m2.n_pairs + m2.n_pairs / 8 + 1024))) return rc;
  }
  std::vector<uint64_t> io(2 * nv);
  HIP_TRY(c, hipMemcpyAsync(io.data(), st.sub_io, 2 * nv * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  for (size_t v = 0; v < nv; ++v)
    if (f2[v]) (*failed)[V.inst[v].parent] = 1;
  for (size_t v = 0; v < nv; ++v) {
    const uint32_t i = V.inst[v].parent;
    if ((*failed)[i]) continue;
    sp->extra[i] += io[2 * v + 1];
    sp->src[i].emplace_back(st.sub_out + io[2 * v], io[2 * v + 1]);
  }
  sp->drop = *failed;
  return PBIG_OK;
}

// Binned plain path, synchronous.  Returns PBIG_OK, an error, or 1 when the
// instance cannot be resolved on the binned path (the caller then runs the
// general path).
int run_binned_sync(pbig_hip_ctx* c, const float4* d_boxes, uint32_t n, uint2* d_out, uint64_t cap,
                    hipStream_t s, uint64_t* n_pairs) {
  BinnedState& st = c->plain;
  const bool auto_choice = st.bp.dense_exit != 0;
  if (st.to_general && auto_choice) return 1;  // (known from this plan's last call)
  // An instance of mostly dense clusters goes to the device levels when it
  // has kDenseLevelsFactor x levels_min_boxes or more (level 0 then runs
  // without the dense exit), else to the general path, whose per-bin cell
  // layout is the faster of the two on every size measured (clustered
  // instances of 2^20, 2^21 and 2^23 boxes: 0.20 vs 0.46 ms, 0.26 vs 0.53 ms
  // and 0.96 vs 1.23 ms).
  const bool dense_to_levels =
      auto_choice && (uint64_t)n >= kDenseLevelsFactor * c->cfg.levels_min_boxes && !c->cfg.host_levels;
  Meta m;
  int rc = run_onepass_sync(c, d_boxes, d_out, cap, s, &m);
  if (rc < 0) return rc;
  if (rc == 2) {
    if (!dense_to_levels) return 1;  // the general path
    st.to_general = false;
    st.dense_levels = true;
    rc = 1;
  }
  if (rc == 1 && auto_choice && !st.deep_ok && !st.dense_levels) {
    // The automatic choice runs the two-pass launch alone first: its bin
    // totals find an instance of mostly dense clusters (which stops the
    // launch before it scatters), and an instance with no bin beyond LDS is
    // done.  Bins beyond LDS go to the device levels in instances of
    // levels_min_boxes or more, else to the general path.
    if ((rc = run_state_sync(c, st, d_boxes, nullptr, d_out, cap, nullptr, s, &m))) return rc;
    if ((m.error & pbig::kErrDense) && dense_to_levels) {
      st.dense_levels = true;
    } else if ((m.error & pbig::kErrDense) || (m.error && n < c->cfg.levels_min_boxes)) {
      st.to_general = true;
      return 1;
    }
    rc = m.error ? 1 : 0;
  }
  if (rc == 1) rc = run_deep_sync(c, d_boxes, d_out, cap, s, &m);  // re-binning levels on the device
  if (rc < 0) return rc;
  if (rc == 2) return 1;  // mostly dense clusters: the general path
  if (rc == 1 && st.dense_levels) {  // the levels cannot complete this plan: the general path
    st.dense_levels = false;
    st.to_general = true;
    return 1;
  }
  if (rc == 1) {  // the two-pass launch, with re-binning of bins beyond LDS on the host
    Splice sp;
    std::vector<uint8_t> failed;
    if ((rc = resolve(c, st, d_boxes, nullptr, d_out, cap, nullptr, s, 0, &m, &sp, &failed))) return rc;
    if (failed[0]) return 1;
    if (!sp.empty() && (rc = apply_splice(c, st, sp, d_out, cap, nullptr, s, &m))) return rc;
  }
  const InstDesc& d = st.inst[0];
  c->last_path = PBIG_PATH_BINNED;
  c->stats.n_boxes = n;
  c->stats.n_cells = (uint64_t)d.nf * d.nf;
  c->stats.cells_per_row = d.nf;
  c->stats.memberships = m.memberships;
  c->stats.tile_pairs = 0;
  c->stats.n_pairs = m.n_pairs;
  c->stats.max_per_cell = 0;
  c->stats.path = PBIG_PATH_BINNED;
  c->stats.bins_per_row = d.nbx;
  c->stats.cells_per_bin_side = d.k;
  *n_pairs = m.n_pairs;
  if (m.n_pairs > cap)
    return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu", (unsigned long long)m.n_pairs,
                (unsigned long long)cap);
  return PBIG_OK;
}

// Either path, synchronous, by the context's configuration.
int run_any_sync(pbig_hip_ctx* c, const float4* d_boxes, uint32_t n, double range, float cell, uint2* d_out,
                 uint64_t cap, hipStream_t s, uint64_t* n_pairs) {
  int rc;
  if (c->cfg.path != PBIG_PATH_GENERAL) {
    bool ok = false;
    if ((rc = plan_plain(c, n, range, cell, s, &ok))) return rc;
    c->plain.bp.dense_exit = c->cfg.path == PBIG_PATH_AUTO ? 1u : 0u;
    c->plain.bp.direct = 1u;
    if (ok) {
      rc = run_binned_sync(c, d_boxes, n, d_out, cap, s, n_pairs);
      if (rc != 1) return rc;
      if (c->cfg.path == PBIG_PATH_BINNED)
        return fail(c, PBIG_E_TOO_DENSE, "a coarse bin holds more than %u fine memberships", pbig::kMaxFine);
    } else if (c->cfg.path == PBIG_PATH_BINNED) {
      return fail(c, PBIG_E_ARG, "the binned path cannot plan this grid (more than %u bins)", pbig::kMaxBins);
    }
  }
  GridParams gp;
  uint32_t n_cells;
  if ((rc = grid_of(c, range, cell, &gp, &n_cells))) return rc;
  // The general path's cells laid out from the bins when the binned grid
  // can be planned, else by one device atomic per membership over the whole
  // grid's cells.
  bool planned = false;
  if ((rc = plan_plain(c, n, range, cell, s, &planned))) return rc;
  rc = planned ? run_sorted_sync(c, d_boxes, n, d_out, cap, s, n_pairs)
               : run_sync(c, d_boxes, n, gp, n_cells, d_out, cap, s, n_pairs);
  if (rc == PBIG_OK || rc == PBIG_E_OVERFLOW) {
    c->last_path = PBIG_PATH_GENERAL;
    c->stats.path = PBIG_PATH_GENERAL;
    c->stats.bins_per_row = planned ? c->plain.inst[0].nbx : 0u;
    c->stats.cells_per_bin_side = planned ? c->plain.inst[0].k : 0u;
  }
  return rc;
}

void set_cell_stats(pbig_hip_ctx* c, float cell) {
  // Record the cell side as an exact fraction when it is a dyadic rational.
  uint32_t den = 1;
  double v = cell;
  while (v != std::floor(v) && den < (1u << 20)) { v *= 2; den *= 2; }
  c->stats.cell_size_num = (uint32_t)v;
  c->stats.cell_size_den = den;
}

// ---------------------------------------------------------------------------
// Host entry points: packing and the copy of the pairs back.

// The context's host threads (created on first use, at most 16).
pbig::HostPool& host_pool(pbig_hip_ctx* c) {
  if (!c->pool) {
    const unsigned hw = std::thread::hardware_concurrency();
    c->pool = std::make_unique<pbig::HostPool>(hw ? std::min(hw, 16u) : 4u);
  }
  return *c->pool;
}

// Runs work(t) over the pool's threads, or on this thread alone for small jobs.
unsigned host_parts(pbig_hip_ctx* c, uint64_t n) { return n >= (1u << 16) ? host_pool(c).size() : 1u; }
void host_run(pbig_hip_ctx* c, unsigned parts, const std::function<void(unsigned)>& work) {
  if (parts == 1) work(0);
  else host_pool(c).run(work);
}

constexpr uint64_t kOutChunk = 8ull << 20;  // pairs per chunk of the copy back (64 MB)

// np pairs of d_pairs back to the host in chunks through two pinned buffers:
// consume(chunk, count, first) runs on each chunk while the next one is in
// flight (the reference's match phase overlaps its prefetch of results with
// the decode of the previous ones the same way, src/pbig/pbig.cu:1848-1866).
// consume returns a status; anything but PBIG_OK stops the copy.
template <class Consume>
int stream_pairs(pbig_hip_ctx* c, const uint2* d_pairs, uint64_t np, hipStream_t s, Consume consume) {
  if (np == 0) return PBIG_OK;
  const uint64_t want = std::min<uint64_t>(np, kOutChunk);
  if (c->cap_h_out < want) {
    for (auto& q : c->h_out) {
      if (q) HIP_TRY(c, hipHostFree(q));
      q = nullptr;
    }
    c->cap_h_out = 0;
    for (auto& q : c->h_out) HIP_TRY(c, hipHostMalloc((void**)&q, want * sizeof(uint2), hipHostMallocDefault));
    c->cap_h_out = want;
  }
  for (auto& e : c->out_ev)
    if (!e) HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const uint64_t chunk = c->cap_h_out, nch = (np + chunk - 1) / chunk;
  auto issue = [&](uint64_t i) -> int {
    const uint64_t first = i * chunk, len = std::min(chunk, np - first);
    HIP_TRY(c, hipMemcpyAsync(c->h_out[i & 1], d_pairs + first, len * sizeof(uint2), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipEventRecord(c->out_ev[i & 1], s));
    return PBIG_OK;
  };
  int rc;
  if ((rc = issue(0))) return rc;
  for (uint64_t i = 0; i < nch; ++i) {
    // the next chunk goes to the other buffer, consumed one iteration ago
    if (i + 1 < nch && (rc = issue(i + 1))) return rc;
    HIP_TRY(c, hipEventSynchronize(c->out_ev[i & 1]));
    const uint64_t first = i * chunk, len = std::min(chunk, np - first);
    if ((rc = consume(c->h_out[i & 1], len, first))) {
      HIP_TRY(c, hipStreamSynchronize(s));  // (the copy in flight lands before its buffer is reused)
      return rc;
    }
  }
  return PBIG_OK;
}

// A chunk of pairs into the caller's memory, split over the host threads.
int copy_out(pbig_hip_ctx* c, const uint2* src, uint64_t len, uint32_t* dst) {
  const unsigned parts = host_parts(c, len);
  host_run(c, parts, [&](unsigned t) {
    uint64_t lo, hi;
    pbig::HostPool::part(len, parts, t, &lo, &hi);
    memcpy(dst + 2 * lo, src + lo, (hi - lo) * sizeof(uint2));
  });
  return PBIG_OK;
}

// RectInfo records -> the pinned staging of the boxes, over the host threads.
// check(i, record) returns 0 or a status for record i; each part stops at its
// first bad record, so the lowest part with one holds the first bad record
// of the whole array, as a loop in order would report it.
template <class Pack>
int pack_records(pbig_hip_ctx* c, uint32_t n, Pack pack, uint64_t* bad_index, int* bad_code) {
  const unsigned parts = host_parts(c, n);
  std::vector<uint64_t> bad(parts, ~0ull);
  std::vector<int> code(parts, 0);
  host_run(c, parts, [&](unsigned t) {
    uint64_t lo, hi;
    pbig::HostPool::part(n, parts, t, &lo, &hi);
    for (uint64_t i = lo; i < hi; ++i) {
      const int e = pack((uint32_t)i);
      if (e) {
        bad[t] = i;
        code[t] = e;
        return;
      }
    }
  });
  for (unsigned t = 0; t < parts; ++t)
    if (code[t]) {
      *bad_index = bad[t];
      *bad_code = code[t];
      return code[t];
    }
  return 0;
}

}  // namespace

extern "C" {

int pbig_hip_create(pbig_hip_ctx** out, int device_id) {
  if (!out) return PBIG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PBIG_E_HIP;
  if (device_id < 0 || device_id >= ndev) return PBIG_E_ARG;
  pbig_hip_ctx* c = new pbig_hip_ctx();
  c->device = device_id;
  if (hipSetDevice(device_id) != hipSuccess) { delete c; return PBIG_E_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_id) != hipSuccess) { delete c; return PBIG_E_HIP; }
  c->n_cu = prop.multiProcessorCount;
  if (pbig::binned_init() != hipSuccess) { delete c; return PBIG_E_HIP; }
  if (noalloc_setup(device_id, &c->noalloc) != hipSuccess) { delete c; return PBIG_E_HIP; }
  // tile test: three times the resident workgroups (eight of four waves per
  // CU), so that later workgroups start as those before them run out of
  // items, which evens out items of very different sizes
  c->ws.test_blocks = (uint32_t)c->n_cu * 24;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return PBIG_E_HIP; }
  if (hipMalloc((void**)&c->ws.meta, sizeof(Meta)) != hipSuccess) { delete c; return PBIG_E_NOMEM; }
  if (hipMemset(c->ws.meta, 0, sizeof(Meta)) != hipSuccess) {
    (void)hipFree(c->ws.meta);
    delete c;
    return PBIG_E_HIP;
  }
  if (hipHostMalloc((void**)&c->h_meta, 2 * sizeof(Meta), hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(c->ws.meta);
    delete c;
    return PBIG_E_NOMEM;
  }
  memset(c->h_meta, 0, 2 * sizeof(Meta));
  if (hipMalloc((void**)&c->sort_meta, sizeof(Meta)) != hipSuccess ||
      hipMemset(c->sort_meta, 0, sizeof(Meta)) != hipSuccess) {
    pbig_hip_destroy(c);
    return PBIG_E_NOMEM;
  }
  bool ok = hipHostMalloc((void**)&c->h_ring, 16 * sizeof(InstDesc), hipHostMallocDefault) == hipSuccess;
  for (hipEvent_t& e : c->ring_ev)
    ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess &&
         hipEventRecord(e, c->stream) == hipSuccess;
  if (!ok) {
    pbig_hip_destroy(c);
    return PBIG_E_HIP;
  }
  *out = c;
  return PBIG_OK;
}

int pbig_hip_destroy(pbig_hip_ctx* c) {
  if (!c) return PBIG_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  Workspace& w = c->ws;
  for (void* p : {(void*)w.cnt, (void*)w.start, (void*)w.partials, (void*)w.items, (void*)w.mbox,
                  (void*)w.mid, (void*)w.strips, (void*)w.slc, (void*)w.meta, (void*)c->sort_meta, (void*)c->d_in,
                  (void*)c->d_pairs})
    if (p) (void)hipFree(p);
  free_state(c->plain);
  for (void* p : {(void*)c->uw.mxy, (void*)c->uw.mz, (void*)c->uw.mid, (void*)c->uw.mcell, (void*)c->d_in32})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)c->h_in32, (void*)c->h_out[0], (void*)c->h_out[1]})
    if (p) (void)hipHostFree(p);
  for (hipEvent_t e : c->out_ev)
    if (e) (void)hipEventDestroy(e);
  c->pool.reset();
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ring_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_ring) (void)hipHostFree(c->h_ring);
  if (c->h_meta) (void)hipHostFree(c->h_meta);
  if (c->h_in) (void)hipHostFree(c->h_in);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PBIG_OK;
}

const char* pbig_hip_last_error(const pbig_hip_ctx* c) { return c ? c->err.c_str() : "null context"; }

// decideCellSize, src/pbig/pbig.cu:1285-1357, for float boxes.  The mean
// extent of the first 1000 boxes is accumulated as the reference's
// thrust::transform_reduce with an int init value does it (pbig.cu:1303-1308):
// int accumulator, (int)((float)acc + extent) per step, integer division by
// the sample count.  Then the side halves from the largest power of two below
// the range while a cell still holds >= 32 boxes and the mean replication
// (1 + x/cs)(1 + y/cs) stays <= 4 (pbig.cu:1280-1281, 1316-1331); the answer is
// twice the side at which the loop stopped.
uint32_t pbig_hip_reference_cell_size(const float* h_boxes, uint32_t n, double range) {
  int64_t cs = 1;
  while ((double)(cs * 2) < range) cs *= 2;
  int samples = n < 1000 ? (int)n : 1000;
  if (!h_boxes || samples == 0) return (uint32_t)cs;
  int ax = 0, ay = 0;
  for (int k = 0; k < samples; ++k) {
    const float* b = h_boxes + 4 * k;
    ax = (int)((float)ax + (b[1] - b[0]));
    ay = (int)((float)ay + (b[3] - b[2]));
  }
  const float x = (float)(ax / samples), y = (float)(ay / samples);
  while (true) {
    int64_t nrc = (int64_t)std::ceil(range / (double)cs);
    int64_t nc = nrc * nrc;
    double nmr = (double)((1 + x / (float)cs) * (1 + y / (float)cs));
    int per_cell = (int)(nmr * (double)n / (double)nc);
    if ((float)cs < x) break;
    if (per_cell < 32 || nmr > 4) break;
    if (cs == 1) break;
    cs /= 2;
  }
  return (uint32_t)(cs * 2);
}

int pbig_hip_reserve(pbig_hip_ctx* c, uint64_t n_boxes, uint64_t n_cells, uint64_t n_memberships) {
  if (!c) return PBIG_E_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure_cells(c, n_cells))) return rc;
  uint64_t items = std::max<uint64_t>(n_cells / 2 + 1024, n_memberships / 32 + 1024);
  return ensure_work(c, std::max<uint64_t>(n_memberships, n_boxes), items);
}

int pbig_hip_check_device(pbig_hip_ctx* c, const float* d_boxes, uint32_t n, double range, float cell_size,
                          uint32_t* d_pairs, uint64_t cap, uint64_t* n_pairs, void* stream) {
  if (!c) return PBIG_E_ARG;
  if (!n_pairs) return fail(c, PBIG_E_ARG, "n_pairs is null");
  if (n > pbig::kMaxBoxes) return fail(c, PBIG_E_ARG, "n = %u exceeds the limit of 2^30-1 boxes", n);
  if (n && !d_boxes) return fail(c, PBIG_E_ARG, "d_boxes is null");
  if (cap && !d_pairs) return fail(c, PBIG_E_ARG, "d_pairs is null with a nonzero capacity");
  if (((uintptr_t)d_boxes & 15) != 0) return fail(c, PBIG_E_ARG, "d_boxes must be 16-byte aligned");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  if (cell_size == 0.0f) {
    std::vector<float> head(4 * (size_t)std::min<uint32_t>(n, 1000));
    if (!head.empty()) {
      HIP_TRY(c, hipMemcpyAsync(head.data(), d_boxes, head.size() * sizeof(float), hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipStreamSynchronize(s));
    }
    cell_size = (float)pbig_hip_reference_cell_size(head.data(), n, range);
  }
  GridParams gp;
  uint32_t n_cells;
  int rc;
  if ((rc = grid_of(c, range, cell_size, &gp, &n_cells))) return rc;
  rc = run_any_sync(c, (const float4*)d_boxes, n, range, cell_size, (uint2*)d_pairs, cap, s, n_pairs);
  set_cell_stats(c, cell_size);
  return rc;
}

int pbig_hip_check_device_async(pbig_hip_ctx* c, const float* d_boxes, uint32_t n, double range,
                                float cell_size, uint32_t* d_pairs, uint64_t cap, uint64_t* d_result,
                                void* stream) {
  if (!c) return PBIG_E_ARG;
  if (n > pbig::kMaxBoxes) return fail(c, PBIG_E_ARG, "n = %u exceeds the limit of 2^30-1 boxes", n);
  if (n && !d_boxes) return fail(c, PBIG_E_ARG, "d_boxes is null");
  if (((uintptr_t)d_boxes & 15) != 0) return fail(c, PBIG_E_ARG, "d_boxes must be 16-byte aligned");
  if (!(cell_size > 0.0f)) return fail(c, PBIG_E_ARG, "the asynchronous path needs cell_size > 0");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  GridParams gp;
  uint32_t n_cells;
  int rc;
  if ((rc = grid_of(c, range, cell_size, &gp, &n_cells))) return rc;
  bool binned = c->cfg.path == PBIG_PATH_BINNED ||
                (c->cfg.path == PBIG_PATH_AUTO && c->last_path != PBIG_PATH_GENERAL);
  if (binned) {
    bool ok = false;
    if ((rc = plan_plain(c, n, range, cell_size, s, &ok))) return rc;
    c->plain.bp.dense_exit = c->cfg.path == PBIG_PATH_AUTO ? 1u : 0u;
    c->plain.bp.direct = 1u;
    if (!ok && c->cfg.path == PBIG_PATH_BINNED)
      return fail(c, PBIG_E_ARG, "the binned path cannot plan this grid (more than %u bins)", pbig::kMaxBins);
    binned = ok;
  }
  if (binned) {
    bool one = false;
    if ((rc = prepare_onepass(c, c->plain, s, &one))) return rc;
    if (one) {
      Meta* dm = nullptr;
      if ((rc = launch_onepass(c, c->plain, (const float4*)d_boxes, (uint2*)d_pairs, cap, s, &dm))) return rc;
      if (d_result) HIP_TRY(c, pbig::launch_result(dm, d_result, s));
      return PBIG_OK;
    }
    if (c->plain.deep_ok && !c->plain.deep_failed && !c->cfg.host_levels) {
      if ((rc = after_upload(c, c->plain, s))) return rc;
      return enqueue_deep(c, c->plain, (const float4*)d_boxes, (uint2*)d_pairs, cap, d_result, s);
    }
    HIP_TRY(c, pbig::launch_binned(c->plain.ws, (const float4*)d_boxes, nullptr, c->plain.bp, (uint2*)d_pairs, cap,
                                   nullptr, s, next_events(c, PBIG_PATH_BINNED)));
  } else {
    bool planned = false;
    if ((rc = plan_plain(c, n, range, cell_size, s, &planned))) return rc;
    if (planned) {
      if ((rc = size_sorted(c, n, s))) return rc;
      if ((rc = after_upload(c, c->plain, s))) return rc;
      HIP_TRY(c, pbig::launch_sorted(c->plain.ws, (const float4*)d_boxes, sorted_plan(c->plain), c->ws,
                                     c->sort_meta, (uint2*)d_pairs, cap, s, next_events(c, kPathSorted)));
      if (d_result) HIP_TRY(c, pbig::launch_result(c->sort_meta, d_result, s));
      return PBIG_OK;
    }
    if ((rc = prepare(c, n, n_cells))) return rc;
    HIP_TRY(c, pbig::launch_plain(c->ws, (const float4*)d_boxes, n, gp, n_cells, (uint2*)d_pairs, cap, s,
                                  next_events(c, PBIG_PATH_GENERAL)));
  }
  if (d_result)
    HIP_TRY(c, pbig::launch_result(c->ws.meta, d_result, s));
  return PBIG_OK;
}

int pbig_hip_check_device_noalloc(pbig_hip_ctx* c, const float* d_boxes, uint32_t n, double range,
                                  float cell_size, uint32_t* d_pairs, uint64_t cap, uint64_t* d_result,
                                  void* stream) {
  if (!c) return PBIG_E_ARG;
  if (n > pbig::kMaxBoxes) return fail(c, PBIG_E_ARG, "n = %u exceeds the limit of 2^30-1 boxes", n);
  if (n && !d_boxes) return fail(c, PBIG_E_ARG, "d_boxes is null");
  if (((uintptr_t)d_boxes & 15) != 0) return fail(c, PBIG_E_ARG, "d_boxes must be 16-byte aligned");
  if (cap && !d_pairs) return fail(c, PBIG_E_ARG, "d_pairs is null with a nonzero capacity");
  if (!d_result) return fail(c, PBIG_E_ARG, "d_result is null");
  if (!(cell_size > 0.0f)) return fail(c, PBIG_E_ARG, "the no-allocation path needs cell_size > 0");
  if (!c->noalloc) return fail(c, PBIG_E_HIP, "no scratch address for device %d", c->device);
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  GridParams gp;
  uint32_t n_cells;
  int rc;
  if ((rc = grid_of(c, range, cell_size, &gp, &n_cells))) return rc;
  InstDesc d;
  if (!plan_instance(n, range, cell_size, c->cfg.bin_target, &d, (uint32_t)c->n_cu))
    return fail(c, PBIG_E_ARG, "the binned path cannot plan this grid (more than %u bins)", pbig::kMaxBins);
  // The plan of a batch of one (what layout_tables computes), with no host
  // tables: k_single_tables writes them on the device.
  BatchPlan bp;
  bp.n_inst = 1;
  bp.n_chunks = d.nchunks;
  bp.n_bins = n ? d.nb : 0u;
  bp.n_boxes = n;
  bp.nb_max = d.nchunks ? d.nb : 0u;
  bp.hist_entries = (uint64_t)d.nchunks * pbig::hist_pitch(d.nb);
  bp.off_entries = (uint64_t)d.nchunks * pbig::off_pitch(d.nb);
  bp.test_blocks = std::max<uint32_t>(1u, std::min<uint32_t>((uint32_t)c->n_cu, std::max<uint32_t>(bp.n_bins, 1)));
  bp.n_cu = (uint32_t)c->n_cu;
  bp.direct = 1u;
  bp.dense_exit = 0u;
  // Carve the workspace: every array at a 256-byte boundary; the pair
  // regions sized as the allocating path sizes them first; the rest of the
  // array is the bucket (20 bytes per membership, at least 1.25 per box).
  uint64_t at = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t p = at;
    at = (at + bytes + 255u) & ~255ull;
    return p;
  };
  const uint64_t nb1 = (uint64_t)bp.n_bins + 1;
  const uint64_t o_inst = take(sizeof(InstDesc)), o_meta = take(sizeof(Meta));
  const uint64_t o_ci = take(4ull * std::max<uint32_t>(bp.n_chunks, 1)), o_bi = take(4ull * nb1);
  const uint64_t o_hist = take(4 * std::max<uint64_t>(bp.hist_entries, 1));
  const uint64_t o_off = take(4 * std::max<uint64_t>(bp.off_entries, 1));
  const uint64_t o_total = take(4 * nb1), o_start = take(4 * nb1), o_over = take(4 * nb1);
  const uint64_t o_order = take(4 * nb1), o_rank = take(4 * nb1);
  const uint64_t o_seg = take(8 * nb1), o_out = take(8 * nb1);
  const uint64_t cap_wg = std::max<uint64_t>(8192, ((uint64_t)n + bp.test_blocks - 1) / bp.test_blocks * 2);
  const uint64_t o_reg = take(8 * cap_wg * bp.test_blocks);
  const uint64_t want_bucket = (uint64_t)n + n / 4 + 4096;
  const uint64_t room = at < kNoAllocBytes ? (kNoAllocBytes - at) / 20u : 0u;
  if (room < want_bucket)
    return fail(c, PBIG_E_ARG,
                "%u boxes need %llu bytes of scratch; the no-allocation entry point holds %llu (about 2^20 boxes)",
                n, (unsigned long long)(at + 20 * want_bucket), (unsigned long long)kNoAllocBytes);
  const uint64_t cap_bucket = std::min<uint64_t>(room & ~63ull, 0xFFFFFFFFull);
  const uint64_t o_bbox = take(16 * cap_bucket), o_bid = take(4 * cap_bucket);
  unsigned char* base = c->noalloc;
  pbig::BinnedWorkspace ws;
  ws.inst = (InstDesc*)(base + o_inst);
  ws.meta = (Meta*)(base + o_meta);
  ws.chunk_inst = (uint32_t*)(base + o_ci);
  ws.bin_inst = (uint32_t*)(base + o_bi);
  ws.hist = (uint32_t*)(base + o_hist);
  ws.off = (uint32_t*)(base + o_off);
  ws.total = (uint32_t*)(base + o_total);
  ws.bin_start = (uint32_t*)(base + o_start);
  ws.overfull = (uint32_t*)(base + o_over);
  ws.order = (uint32_t*)(base + o_order);
  ws.rank = (uint32_t*)(base + o_rank);
  ws.bin_seg = (uint2*)(base + o_seg);
  ws.bin_out = (uint64_t*)(base + o_out);
  ws.scratch = (uint2*)(base + o_reg);
  ws.cap_wg = (uint32_t)cap_wg;
  ws.cap_scratch = cap_wg * bp.test_blocks;
  ws.bbox = (float4*)(base + o_bbox);
  ws.bid = (uint32_t*)(base + o_bid);
  ws.cap_bucket = cap_bucket;
  ws.cap_hist = std::max<uint64_t>(bp.hist_entries, 1);
  ws.cap_off = std::max<uint64_t>(bp.off_entries, 1);
  ws.cap_bins = nb1;
  ws.cap_inst = 1;
  ws.cap_chunks = std::max<uint32_t>(bp.n_chunks, 1);
  ws.cap_bin_inst = nb1;
  d.box0 = 0;
  HIP_TRY(c, pbig::launch_single_tables(ws, d, s));
  HIP_TRY(c, pbig::launch_binned(ws, (const float4*)d_boxes, nullptr, bp, (uint2*)d_pairs, cap, nullptr, s, nullptr));
  HIP_TRY(c, pbig::launch_result(ws.meta, d_result, s));
  return PBIG_OK;
}

// The host entry points' device part (float records): pack RectInfo
// {xmin,xmax,ymin,ymax,rid} into 16-byte boxes in pinned memory over the host
// threads, checking the reference's contract rid == index and the boxes'
// orientation, upload, and check; the pairs are then in c->d_pairs.
static int host_check_float(pbig_hip_ctx* c, const void* h_rects, uint32_t n, double range, float cell_size,
                            uint64_t* n_pairs) {
  if (!c) return PBIG_E_ARG;
  if (!n_pairs) return fail(c, PBIG_E_ARG, "n_pairs is null");
  if (n > pbig::kMaxBoxes) return fail(c, PBIG_E_ARG, "n = %u exceeds the limit of 2^30-1 boxes", n);
  if (n && !h_rects) return fail(c, PBIG_E_ARG, "h_rects is null");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (c->cap_h_in < n || !c->h_in) {
    if (c->h_in) HIP_TRY(c, hipHostFree(c->h_in));
    c->h_in = nullptr;
    c->cap_h_in = 0;
    HIP_TRY(c, hipHostMalloc((void**)&c->h_in, std::max<uint64_t>(n, 1) * sizeof(float4), hipHostMallocDefault));
    c->cap_h_in = std::max<uint64_t>(n, 1);
  }
  const unsigned char* src = (const unsigned char*)h_rects;
  float4* h_in = c->h_in;
  uint64_t bad = 0;
  int code = 0;
  if (pack_records(
          c, n,
          [&](uint32_t i) {
            float f[4];
            uint32_t rid;
            memcpy(f, src + (size_t)i * 20, 16);
            memcpy(&rid, src + (size_t)i * 20 + 16, 4);
            if (rid != i) return 1;
            if (!(f[0] <= f[1]) || !(f[2] <= f[3])) return 2;
            h_in[i] = make_float4(f[0], f[1], f[2], f[3]);
            return 0;
          },
          &bad, &code)) {
    uint32_t rid;
    float f[4];
    memcpy(f, src + bad * 20, 16);
    memcpy(&rid, src + bad * 20 + 16, 4);
    if (code == 1)
      return fail(c, PBIG_E_ARG, "rect %llu has rid %u; rid must equal the index", (unsigned long long)bad, rid);
    return fail(c, PBIG_E_INVALID_BOX, "rect %llu is inverted or NaN: [%g,%g]x[%g,%g]", (unsigned long long)bad, f[0],
                f[1], f[2], f[3]);
  }
  int rc;
  if ((rc = grow(c, &c->d_in, &c->cap_in, n))) return rc;
  if (n) HIP_TRY(c, hipMemcpyAsync(c->d_in, c->h_in, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s));
  if (cell_size == 0.0f) cell_size = (float)pbig_hip_reference_cell_size((const float*)c->h_in, n, range);
  GridParams gp;
  uint32_t n_cells;
  if ((rc = grid_of(c, range, cell_size, &gp, &n_cells))) return rc;
  uint64_t want = std::max<uint64_t>(c->cap_pairs, std::max<uint64_t>(1024, (uint64_t)n));
  if ((rc = grow(c, &c->d_pairs, &c->cap_pairs, want))) return rc;
  uint64_t np = 0;
  rc = run_any_sync(c, c->d_in, n, range, cell_size, c->d_pairs, c->cap_pairs, s, &np);
  if (rc == PBIG_E_OVERFLOW) {
    if ((rc = grow(c, &c->d_pairs, &c->cap_pairs, np + np / 8 + 1024))) return rc;
    rc = run_any_sync(c, c->d_in, n, range, cell_size, c->d_pairs, c->cap_pairs, s, &np);
  }
  set_cell_stats(c, cell_size);
  if (rc) return rc;
  *n_pairs = np;
  return PBIG_OK;
}

int pbig_hip_check_host(pbig_hip_ctx* c, const void* h_rects, uint32_t n, double range, float cell_size,
                        uint32_t* h_pairs, uint64_t cap, uint64_t* n_pairs) {
  if (c && cap && !h_pairs) return fail(c, PBIG_E_ARG, "h_pairs is null with a nonzero capacity");
  int rc;
  if ((rc = host_check_float(c, h_rects, n, range, cell_size, n_pairs))) return rc;
  const uint64_t np = *n_pairs;
  if ((rc = stream_pairs(c, c->d_pairs, std::min<uint64_t>(np, cap), c->stream,
                         [&](const uint2* chunk, uint64_t len, uint64_t first) {
                           return copy_out(c, chunk, len, h_pairs + 2 * first);
                         })))
    return rc;
  if (np > cap)
    return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu", (unsigned long long)np,
                (unsigned long long)cap);
  return PBIG_OK;
}

int pbig_hip_check_host_each(pbig_hip_ctx* c, const void* h_rects, uint32_t n, double range, float cell_size,
                             pbig_hip_pairs_fn fn, void* user, uint64_t* n_pairs) {
  if (c && !fn) return fail(c, PBIG_E_ARG, "fn is null");
  int rc;
  if ((rc = host_check_float(c, h_rects, n, range, cell_size, n_pairs))) return rc;
  return stream_pairs(c, c->d_pairs, *n_pairs, c->stream, [&](const uint2* chunk, uint64_t len, uint64_t) {
    fn(user, (const uint32_t*)chunk, len);
    return PBIG_OK;
  });
}

int pbig_hip_timing_start(pbig_hip_ctx* c, int max_calls) {
  if (!c || max_calls < 0) return PBIG_E_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  for (hipEvent_t e : c->events) HIP_TRY(c, hipEventDestroy(e));
  c->events.assign((size_t)max_calls * kEventsPerCall, nullptr);
  for (auto& e : c->events) HIP_TRY(c, hipEventCreate(&e));
  c->timing_cap = max_calls;
  c->timing_used = 0;
  c->timing_path = 0;
  return PBIG_OK;
}

int pbig_hip_timing_read(pbig_hip_ctx* c, double* ms_per_kernel, int n_kernels, int* n_calls, int* path) {
  if (!c || !ms_per_kernel || n_kernels < pbig::kPlainKernels || !n_calls) return PBIG_E_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  const int nk = c->timing_path == PBIG_PATH_BINNED ? pbig::kBinnedKernels
                 : c->timing_path == kPathOnePass ? pbig::kOnePassKernels
                 : c->timing_path == kPathDeep    ? (int)pbig::kDeepLevels
                 : c->timing_path == kPathSorted  ? pbig::kSortedKernels
                                                   : pbig::kPlainKernels;
  for (int k = 0; k < n_kernels; ++k) ms_per_kernel[k] = 0.0;
  for (int call = 0; call < c->timing_used; ++call) {
    hipEvent_t* ev = &c->events[(size_t)call * kEventsPerCall];
    HIP_TRY(c, hipEventSynchronize(ev[nk]));
    for (int k = 0; k < nk; ++k) {
      float ms = 0.f;
      HIP_TRY(c, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      ms_per_kernel[k] += ms;
    }
  }
  *n_calls = c->timing_used;
  if (path) *path = c->timing_path;
  c->timing_used = 0;
  c->timing_cap = 0;
  return PBIG_OK;
}

int pbig_hip_configure(pbig_hip_ctx* c, const pbig_hip_config* cfg) {
  if (!c || !cfg) return PBIG_E_ARG;
  if (cfg->path < PBIG_PATH_AUTO || cfg->path > PBIG_PATH_BINNED) return fail(c, PBIG_E_ARG, "unknown path %d", cfg->path);
  c->cfg = *cfg;
  if (c->cfg.bin_target == 0) c->cfg.bin_target = 2000;
  if (c->cfg.levels_min_boxes == 0) c->cfg.levels_min_boxes = 1u << 20;
  c->last_path = 0;
  return PBIG_OK;
}

int pbig_hip_last_stats(const pbig_hip_ctx* c, pbig_hip_stats* out) {
  if (!c || !out) return PBIG_E_ARG;
  *out = c->stats;
  return PBIG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches.
struct pbig_hip_batch {
  pbig_hip_ctx* ctx = nullptr;
  std::vector<pbig_hip_instance> desc;
  std::vector<uint8_t> general;     // instances the binned path cannot plan
  BinnedState st;
  uint64_t max_box_end = 0;         // largest box_offset + n_boxes
  uint2* d_tmp = nullptr;           // fallback pairs
  uint64_t cap_tmp = 0;
  uint64_t* d_inst_out = nullptr;   // n_inst x {offset, count}
  uint64_t cap_inst_out = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_generation = 0;
  hipStream_t capture_stream = nullptr;
  uint64_t last_cap = 0;            // pair capacity of the last launch (or of the captured graph)
  uint64_t* d_merge = nullptr;      // merge words of launches with device levels (3 n_inst + 1)
  uint64_t cap_merge = 0;
  bool last_levels = false;         // the last launch (or the captured graph) ran device levels
};

namespace {

int batch_fail(pbig_hip_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  b->ctx->err = buf;
  return code;
}

int batch_check_args(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total) {
  if (b->max_box_end > n_boxes_total)
    return batch_fail(b, PBIG_E_ARG, "an instance reaches box %llu beyond the %llu boxes given",
                      (unsigned long long)b->max_box_end, (unsigned long long)n_boxes_total);
  if (b->max_box_end && !d_boxes) return batch_fail(b, PBIG_E_ARG, "d_boxes is null");
  if (((uintptr_t)d_boxes & 15) != 0) return batch_fail(b, PBIG_E_ARG, "d_boxes must be 16-byte aligned");
  return PBIG_OK;
}

void destroy_graph(pbig_hip_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Instance i through the general path into d_out (capacity cap); the count
// also when it exceeds cap.
int run_general_instance(pbig_hip_batch* b, const float4* d_boxes, uint32_t i, uint2* d_out, uint64_t cap,
                         hipStream_t s, uint64_t* count) {
  pbig_hip_ctx* c = b->ctx;
  const pbig_hip_instance& in = b->desc[i];
  GridParams gp;
  uint32_t n_cells;
  int rc;
  if ((rc = grid_of(c, in.range, in.cell_size, &gp, &n_cells))) return rc;
  rc = run_sync(c, d_boxes + in.box_offset, in.n_boxes, gp, n_cells, d_out, cap, s, count);
  return rc == PBIG_E_OVERFLOW ? PBIG_OK : rc;
}

// Device levels for a batch: every instance planned on the binned path and
// the levels not found unable to complete it.
bool batch_levels_apply(const pbig_hip_batch* b) {
  if (b->ctx->cfg.host_levels || b->st.deep_failed || b->st.bp.n_bins == 0 || b->desc.empty()) return false;
  for (uint8_t g : b->general)
    if (g) return false;
  return true;
}

// Enqueue a batch launch with levels on the device: level 0 (the two-pass
// launch up to its test kernel), the levels over every bin beyond LDS of
// every instance (their pairs kept per bin), then the merge: each
// instance's run holds its level-0 pairs followed by its levels' pairs;
// result words into d_result (optional).  The merge words are grown here.
int enqueue_batch_levels(pbig_hip_batch* b, const float4* boxes, uint2* out, uint64_t cap, uint64_t* d_inst_out,
                         uint64_t* d_result, hipStream_t s) {
  pbig_hip_ctx* c = b->ctx;
  BinnedState& st = b->st;
  const uint32_t n_inst = st.bp.n_inst;
  int rc;
  if ((rc = grow(c, &b->d_merge, &b->cap_merge, 3 * (uint64_t)n_inst + 1))) return rc;
  HIP_TRY(c, pbig::launch_merge_clear(b->d_merge, n_inst, s));
  HIP_TRY(c, pbig::launch_dense_clear(st.dense_meta, s));
  st.bp.direct = 0u;
  st.bp.dense_exit = 0u;
  HIP_TRY(c, pbig::launch_binned(st.ws, boxes, nullptr, st.bp, out, cap, nullptr, s, nullptr, false));
  const BinnedWorkspace* parent = &st.ws;
  uint32_t parent_bins = st.bp.n_bins;
  pbig::DeepMetas dm{};
  dm.meta[0] = st.ws.meta;
  const uint32_t target = std::max<uint32_t>(c->cfg.bin_target / 2, 64);
  const uint32_t n_cu = (uint32_t)std::max(c->n_cu, 1);
  for (uint32_t l = 1; l < pbig::kDeepLevels; ++l) {
    const BinnedState& L = *st.dl[l];
    HIP_TRY(c, pbig::launch_deep_level(*parent, parent_bins, l > 1, L.ws, st.dcaps[l], target, 0.0f, st.ws.inst,
                                       deep_refine(), n_cu, 0u, out, cap, l - 1, st.dense_list, st.dense_meta, s));
    dm.meta[l] = L.ws.meta;
    parent = &L.ws;
    parent_bins = st.dcaps[l].bins;
  }
  HIP_TRY(c, pbig::launch_dense_bins(dense_src(st), pbig::kDeepLevels - 1, st.dense_meta, n_cu, nullptr, out, cap,
                                     st.dense_pairs, st.cap_dense_pairs, st.dense_chunks,
                                     (uint32_t)std::min<uint64_t>(st.cap_dense_chunks, 0xFFFFFFFFull), s));
  for (uint32_t l = 1; l < pbig::kDeepLevels; ++l)
    HIP_TRY(c, pbig::launch_level_count(st.dl[l]->ws, st.dcaps[l].bins, b->d_merge, s));
  const uint32_t cap_chunks = (uint32_t)std::min<uint64_t>(st.cap_dense_chunks, 0xFFFFFFFFull);
  HIP_TRY(c, pbig::launch_dense_count(st.dense_chunks, cap_chunks, st.dense_meta, b->d_merge, n_inst, s));
  HIP_TRY(c, pbig::launch_merge_scan(b->d_merge, n_inst, s));
  HIP_TRY(c, pbig::launch_binned_output(st.ws, st.bp, out, cap, d_inst_out, s, b->d_merge));
  for (uint32_t l = 1; l < pbig::kDeepLevels; ++l)
    HIP_TRY(c, pbig::launch_level_copy(st.dl[l]->ws, st.dcaps[l].bins, n_cu, st.ws, st.ws.bin_out, b->d_merge, n_inst,
                                       out, cap, s));
  HIP_TRY(c, pbig::launch_dense_copy(st.dense_chunks, cap_chunks, st.dense_meta, st.dense_pairs, st.ws, b->d_merge,
                                     n_inst, out, cap, s));
  dm.levels = pbig::kDeepLevels;
  if (d_result) HIP_TRY(c, pbig::launch_deep_result(dm, st.dense_meta, d_result, s));
  return PBIG_OK;
}

// A batch launch with device levels, synchronously (run_levels_sync's
// codes; 1 also when the batch cannot take them).  On success the batch's
// later asynchronous launches and captures take the levels too.
int run_batch_levels_sync(pbig_hip_batch* b, const float4* boxes, uint2* out, uint64_t cap, hipStream_t s,
                          Meta* m) {
  if (!batch_levels_apply(b)) return 1;
  return run_levels_sync(b->ctx, b->st, s, m, true, [&](uint64_t* res) {
    return enqueue_batch_levels(b, boxes, out, cap, b->d_inst_out, res, s);
  });
}

// (Re)plan a batch: validate and plan every instance, lay out the tables,
// grow the workspace (it is never shrunk) and upload the tables on stream s
// (which it synchronises).  A captured graph no longer matches and is dropped.
int plan_batch(pbig_hip_batch* b, const pbig_hip_instance* inst, uint32_t n_inst, hipStream_t s) {
  pbig_hip_ctx* c = b->ctx;
  if (n_inst && !inst) return fail(c, PBIG_E_ARG, "inst is null");
  destroy_graph(b);
  std::vector<InstDesc> plan(n_inst);
  std::vector<uint8_t> general(n_inst, 0);
  uint64_t max_end = 0;
  for (uint32_t i = 0; i < n_inst; ++i) {
    const pbig_hip_instance& in = inst[i];
    GridParams gp;
    uint32_t n_cells;
    int rc;
    if (in.n_boxes > pbig::kMaxBoxes) return fail(c, PBIG_E_ARG, "instance %u: n = %u exceeds 2^30-1 boxes", i, in.n_boxes);
    if ((rc = grid_of(c, in.range, in.cell_size, &gp, &n_cells))) return rc;
    max_end = std::max<uint64_t>(max_end, in.box_offset + in.n_boxes);
    if (!plan_instance(in.n_boxes, in.range, in.cell_size, c->cfg.bin_target, &plan[i])) {
      general[i] = 1;
      plan[i] = InstDesc{};
      plan[i].n = in.n_boxes;
    }
    plan[i].box0 = in.box_offset;
  }
  b->desc.assign(inst, inst + n_inst);
  b->general = general;
  b->st.deep_failed = false;  // (deep_ok and the levels' bounds stay: batches of other shapes follow)
  b->max_box_end = max_end;
  b->st.inst = plan;
  std::vector<uint64_t> skip(n_inst, ~0ull);
  bool any_general = false;
  for (uint32_t i = 0; i < n_inst; ++i)
    if (general[i]) {
      skip[i] = 0;
      any_general = true;
    }
  if (!layout_tables(b->st, (uint32_t)c->n_cu, any_general ? &skip : nullptr))
    return fail(c, PBIG_E_ARG, "batch too large (more than 2^31 chunks or bins)");
  int rc;
  if ((rc = size_state(c, b->st))) return rc;
  if ((rc = grow(c, &b->d_inst_out, &b->cap_inst_out, 2 * (uint64_t)std::max<uint32_t>(n_inst, 1)))) return rc;
  if ((rc = grow(c, &b->d_merge, &b->cap_merge, 3 * (uint64_t)std::max<uint32_t>(n_inst, 1) + 1))) return rc;
  return upload_tables(c, b->st, s);
}

}  // namespace

extern "C" {

int pbig_hip_batch_create(pbig_hip_ctx* c, const pbig_hip_instance* inst, uint32_t n_inst, pbig_hip_batch** out) {
  if (!c || !out) return PBIG_E_ARG;
  *out = nullptr;
  HIP_TRY(c, hipSetDevice(c->device));
  pbig_hip_batch* b = new pbig_hip_batch();
  b->ctx = c;
  if (hipMalloc((void**)&b->st.ws.meta, sizeof(Meta)) != hipSuccess ||
      hipMemset(b->st.ws.meta, 0, sizeof(Meta)) != hipSuccess) {
    if (b->st.ws.meta) (void)hipFree(b->st.ws.meta);
    delete b;
    return fail(c, PBIG_E_NOMEM, "hipMalloc of the batch's result words failed");
  }
  if (hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) {
    pbig_hip_batch_destroy(b);
    return fail(c, PBIG_E_HIP, "stream creation failed");
  }
  int rc;
  if ((rc = plan_batch(b, inst, n_inst, c->stream))) {
    pbig_hip_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PBIG_OK;
}

int pbig_hip_batch_update(pbig_hip_batch* b, const pbig_hip_instance* inst, uint32_t n_inst, void* stream) {
  if (!b) return PBIG_E_ARG;
  pbig_hip_ctx* c = b->ctx;
  HIP_TRY(c, hipSetDevice(c->device));
  return plan_batch(b, inst, n_inst, pick_stream(c, stream));
}

int pbig_hip_batch_destroy(pbig_hip_batch* b) {
  if (!b) return PBIG_E_ARG;
  (void)hipSetDevice(b->ctx->device);
  (void)hipDeviceSynchronize();
  destroy_graph(b);
  free_state(b->st);
  for (void* p : {(void*)b->st.ws.meta, (void*)b->d_tmp, (void*)b->d_inst_out, (void*)b->d_merge})
    if (p) (void)hipFree(p);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  delete b;
  return PBIG_OK;
}

int pbig_hip_batch_info(const pbig_hip_batch* b, uint64_t* n_boxes, uint32_t* n_chunks, uint32_t* n_bins) {
  if (!b) return PBIG_E_ARG;
  if (n_boxes) *n_boxes = b->st.bp.n_boxes;
  if (n_chunks) *n_chunks = b->st.bp.n_chunks;
  if (n_bins) *n_bins = b->st.bp.n_bins;
  return PBIG_OK;
}

int pbig_hip_batch_launch(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total, uint32_t* d_pairs,
                          uint64_t cap, uint64_t* d_inst_pairs, uint64_t* d_result, void* stream) {
  if (!b) return PBIG_E_ARG;
  pbig_hip_ctx* c = b->ctx;
  int rc;
  if ((rc = batch_check_args(b, d_boxes, n_boxes_total))) return rc;
  if (cap && !d_pairs) return fail(c, PBIG_E_ARG, "d_pairs is null with a nonzero capacity");
  for (uint8_t g : b->general)
    if (g) return fail(c, PBIG_E_ARG, "the batch has instances only the synchronous entry point can run");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  if ((rc = after_upload(c, b->st, s))) return rc;
  b->last_cap = cap;
  b->last_levels = b->st.deep_ok && batch_levels_apply(b);
  if (b->last_levels)  // (settled by a synchronous call: levels on the device)
    return enqueue_batch_levels(b, (const float4*)d_boxes, (uint2*)d_pairs, cap, d_inst_pairs, d_result, s);
  HIP_TRY(c, pbig::launch_binned(b->st.ws, (const float4*)d_boxes, nullptr, b->st.bp, (uint2*)d_pairs, cap, d_inst_pairs, s,
                                 next_events(c, PBIG_PATH_BINNED)));
  if (d_result)
    HIP_TRY(c, pbig::launch_result(b->st.ws.meta, d_result, s));
  return PBIG_OK;
}

int pbig_hip_batch_capture(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total, uint32_t* d_pairs,
                           uint64_t cap, uint64_t* d_inst_pairs, uint64_t* d_result) {
  if (!b) return PBIG_E_ARG;
  pbig_hip_ctx* c = b->ctx;
  int rc;
  if ((rc = batch_check_args(b, d_boxes, n_boxes_total))) return rc;
  for (uint8_t g : b->general)
    if (g) return fail(c, PBIG_E_ARG, "the batch has instances only the synchronous entry point can run");
  HIP_TRY(c, hipSetDevice(c->device));
  destroy_graph(b);
  if (b->st.stage_ev) HIP_TRY(c, hipEventSynchronize(b->st.stage_ev));  // tables in place before capture
  const bool levels = b->st.deep_ok && batch_levels_apply(b);
  if (levels && (rc = grow(c, &b->d_merge, &b->cap_merge, 3 * (uint64_t)b->st.bp.n_inst + 1)))
    return rc;  // (no allocation under capture)
  hipStream_t s = b->capture_stream;
  HIP_TRY(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (levels) {
    rc = enqueue_batch_levels(b, (const float4*)d_boxes, (uint2*)d_pairs, cap, d_inst_pairs, d_result, s);
    if (rc != PBIG_OK) e = hipErrorUnknown;
  } else {
    e = pbig::launch_binned(b->st.ws, (const float4*)d_boxes, nullptr, b->st.bp, (uint2*)d_pairs, cap, d_inst_pairs, s,
                            nullptr);
    if (e == hipSuccess && d_result)
      e = pbig::launch_result(b->st.ws.meta, d_result, s);
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    HIP_TRY(c, e != hipSuccess ? e : e2);
  }
  b->graph = g;
  HIP_TRY(c, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_generation = b->st.generation;
  b->last_cap = cap;
  b->last_levels = levels;
  return PBIG_OK;
}

int pbig_hip_batch_replay(pbig_hip_batch* b, void* stream) {
  if (!b) return PBIG_E_ARG;
  pbig_hip_ctx* c = b->ctx;
  if (!b->exec) return fail(c, PBIG_E_ARG, "no graph captured");
  if (b->graph_generation != b->st.generation)
    return fail(c, PBIG_E_ARG, "the batch's workspace moved since the capture; capture again");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipGraphLaunch(b->exec, pick_stream(c, stream)));
  return PBIG_OK;
}

int batch_check_impl(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total, uint32_t* d_pairs,
                     uint64_t cap, uint64_t* h_inst_pairs, uint64_t* n_pairs, void* stream, bool after_launch);

// The end of a synchronous batch check: every instance's (offset, count)
// to the host, the statistics, the total and the capacity check.
int batch_check_done(pbig_hip_batch* b, const Meta& m, uint64_t cap, uint64_t* h_inst_pairs, uint64_t* n_pairs) {
  pbig_hip_ctx* c = b->ctx;
  const uint32_t n_inst = (uint32_t)b->desc.size();
  if (h_inst_pairs && n_inst)
    HIP_TRY(c, hipMemcpy(h_inst_pairs, b->d_inst_out, 2 * (size_t)n_inst * sizeof(uint64_t), hipMemcpyDeviceToHost));
  c->stats = pbig_hip_stats{};
  c->stats.n_boxes = b->st.bp.n_boxes;
  c->stats.memberships = m.memberships;
  c->stats.n_pairs = m.n_pairs;
  c->stats.path = PBIG_PATH_BINNED;
  *n_pairs = m.n_pairs;
  if (m.n_pairs > cap)
    return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu", (unsigned long long)m.n_pairs,
                (unsigned long long)cap);
  return PBIG_OK;
}

int pbig_hip_batch_check(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total, uint32_t* d_pairs,
                         uint64_t cap, uint64_t* h_inst_pairs, uint64_t* n_pairs, void* stream) {
  return batch_check_impl(b, d_boxes, n_boxes_total, d_pairs, cap, h_inst_pairs, n_pairs, stream, false);
}

int pbig_hip_batch_finish(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total, uint32_t* d_pairs,
                          uint64_t cap, uint64_t* h_inst_pairs, uint64_t* n_pairs, void* stream) {
  return batch_check_impl(b, d_boxes, n_boxes_total, d_pairs, cap, h_inst_pairs, n_pairs, stream, true);
}

int pbig_hip_batch_unresolved(pbig_hip_batch* b, uint8_t* h_flags, uint32_t* n_flagged, void* stream) {
  if (!b) return PBIG_E_ARG;
  pbig_hip_ctx* c = b->ctx;
  if (!n_flagged) return fail(c, PBIG_E_ARG, "n_flagged is null");
  const uint32_t n_inst = (uint32_t)b->desc.size();
  if (n_inst && !h_flags) return fail(c, PBIG_E_ARG, "h_flags is null");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  BinnedState& st = b->st;
  HIP_TRY(c, hipMemcpyAsync(c->h_meta, st.ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  const Meta m = *c->h_meta;
  *n_flagged = 0;
  if (m.error & ~(unsigned long long)pbig::kErrOverfull)
    return fail(c, PBIG_E_ARG, "the last launch reported more than bins beyond LDS (error word %#llx)", m.error);
  // Pairs past the launch's capacity were dropped (k_bin_compact keeps only
  // d < cap), so the runs of the instances there are not complete.
  if (m.n_pairs > b->last_cap)
    return fail(c, PBIG_E_OVERFLOW, "the last launch found %llu pairs, more than its capacity of %llu",
                (unsigned long long)m.n_pairs, (unsigned long long)b->last_cap);
  std::fill(h_flags, h_flags + n_inst, (uint8_t)0);
  if (b->last_levels) {
    // Levels on the device took the bins beyond LDS: every instance is
    // complete unless a level reported trouble or still listed such bins
    // (then each instance with a bin beyond LDS at level 0 is flagged).
    std::vector<Meta> lm(pbig::kDeepLevels);
    for (uint32_t l = 1; l < pbig::kDeepLevels; ++l)
      HIP_TRY(c, hipMemcpyAsync(&lm[l], st.dl[l]->ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    Meta dn{};
    HIP_TRY(c, hipMemcpyAsync(&dn, st.dense_meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    bool complete = dn.error == 0;  // (the dense kernel took the bins the levels left)
    for (uint32_t l = 1; l < pbig::kDeepLevels; ++l)
      complete = complete && lm[l].plan_err == 0 && !(lm[l].error & ~(unsigned long long)pbig::kErrOverfull);
    if (complete) return PBIG_OK;
  }
  uint32_t nf = 0;
  const uint32_t n_bins = st.bp.n_bins;
  const uint32_t n_over = (uint32_t)std::min<uint64_t>(m.max_per_cell, n_bins);
  if (n_over) {
    std::vector<uint32_t> over(n_over);
    HIP_TRY(c, hipMemcpy(over.data(), st.ws.overfull, n_over * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t x : over) {
      const uint32_t bin = x & ~pbig::kOverfullCell;
      if (bin >= n_bins) continue;
      const uint32_t i = st.bin_inst[bin];
      if (i < n_inst && !h_flags[i]) {
        h_flags[i] = 1;
        ++nf;
      }
    }
  }
  *n_flagged = nf;
  return PBIG_OK;
}

int batch_check_impl(pbig_hip_batch* b, const float* d_boxes, uint64_t n_boxes_total, uint32_t* d_pairs,
                     uint64_t cap, uint64_t* h_inst_pairs, uint64_t* n_pairs, void* stream, bool after_launch) {
  if (!b) return PBIG_E_ARG;
  pbig_hip_ctx* c = b->ctx;
  if (!n_pairs) return fail(c, PBIG_E_ARG, "n_pairs is null");
  if (cap && !d_pairs) return fail(c, PBIG_E_ARG, "d_pairs is null with a nonzero capacity");
  int rc;
  if ((rc = batch_check_args(b, d_boxes, n_boxes_total))) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  const float4* boxes = (const float4*)d_boxes;
  const uint32_t n_inst = (uint32_t)b->desc.size();
  BinnedState& st = b->st;
  Meta m{};
  // Device levels: the two-pass launch alone first (unless this batch
  // already settled on the levels, or a launch just reported its bins
  // beyond LDS); bins beyond LDS, and nothing else, then run the launch
  // again with levels planned on the device over every such bin.
  bool done = false;
  if (batch_levels_apply(b)) {
    bool levels = st.deep_ok || after_launch;
    if (!levels) {
      if ((rc = run_state_sync(c, st, boxes, nullptr, (uint2*)d_pairs, cap, b->d_inst_out, s, &m))) return rc;
      done = m.error == 0;
      levels = (m.error & ~(unsigned long long)pbig::kErrOverfull) == 0 && (m.error & pbig::kErrOverfull);
    }
    if (!done && levels) {
      rc = run_batch_levels_sync(b, boxes, (uint2*)d_pairs, cap, s, &m);
      if (rc < 0) return rc;
      done = rc == 0;
    }
    after_launch = false;  // (the host path below starts over)
  }
  if (done) return batch_check_done(b, m, cap, h_inst_pairs, n_pairs);
  // The launch, with heavy coarse bins re-binned one level down; instances
  // that stay unresolved (or that the binned path cannot plan) go to the
  // general path, and every instance's pairs are then placed in its run.
  Splice sp;
  std::vector<uint8_t> failed;
  bool pass_done = false;
  if (after_launch) {
    // Results of the last pbig_hip_batch_launch on this stream: continue
    // from them when the only trouble was bins beyond LDS.
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, st.ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    m = *c->h_meta;
    pass_done = (m.error & ~(unsigned long long)pbig::kErrOverfull) == 0 && m.n_pairs <= cap;
  }
  if ((rc = resolve(c, st, boxes, nullptr, (uint2*)d_pairs, cap, b->d_inst_out, s, 0, &m, &sp, &failed,
                    pass_done)))
    return rc;
  std::vector<uint64_t> cnt(n_inst, 0), at(n_inst, 0);
  uint64_t tmp_total = 0;
  bool any_gen = false;
  for (uint32_t i = 0; i < n_inst; ++i) {
    if (!(b->general[i] || failed[i])) continue;
    any_gen = true;
    uint64_t k = 0;
    if (b->desc[i].n_boxes && (rc = run_general_instance(b, boxes, i, nullptr, 0, s, &k))) return rc;
    cnt[i] = k;
    at[i] = tmp_total;
    tmp_total += k;
  }
  if (any_gen) {
    if ((rc = grow(c, &b->d_tmp, &b->cap_tmp, std::max<uint64_t>(tmp_total, 1)))) return rc;
    for (uint32_t i = 0; i < n_inst; ++i) {
      if (!(b->general[i] || failed[i])) continue;
      uint64_t k = 0;
      if (cnt[i] && (rc = run_general_instance(b, boxes, i, b->d_tmp + at[i], cnt[i], s, &k))) return rc;
      sp.drop[i] = 1;
      sp.extra[i] = cnt[i];
      sp.src[i].assign(1, std::make_pair((const uint2*)(b->d_tmp + at[i]), cnt[i]));
    }
  }
  if (!sp.empty() && (rc = apply_splice(c, st, sp, (uint2*)d_pairs, cap, b->d_inst_out, s, &m))) return rc;
  return batch_check_done(b, m, cap, h_inst_pairs, n_pairs);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-compatible unsigned mode.
namespace {

int ensure_u32(pbig_hip_ctx* c, uint64_t memb) {
  pbig::U32Workspace& u = c->uw;
  if (memb > 0xFFFFFFFFull) return fail(c, PBIG_E_NOMEM, "more than 2^32-1 cell memberships");
  if (u.cap_memb >= memb && u.mxy) return PBIG_OK;
  for (void* p : {(void*)u.mxy, (void*)u.mz, (void*)u.mid, (void*)u.mcell})
    if (p) HIP_TRY(c, hipFree(p));
  u = pbig::U32Workspace{};
  uint64_t d1 = 0, d2 = 0, d3 = 0, d4 = 0;
  int rc;
  if ((rc = grow(c, &u.mxy, &d1, memb)) || (rc = grow(c, &u.mz, &d2, memb)) || (rc = grow(c, &u.mid, &d3, memb)) ||
      (rc = grow(c, &u.mcell, &d4, memb)))
    return rc;
  u.cap_memb = memb;
  return PBIG_OK;
}

int check_u32(pbig_hip_ctx* c, const uint32_t* d_boxes, uint32_t n, int dims, uint32_t range, uint32_t cs,
              uint2* d_out, uint64_t cap, uint64_t* n_pairs, hipStream_t s) {
  if (cs == 0 || (cs & (cs - 1)) != 0) return fail(c, PBIG_E_ARG, "cell size %u is not a power of two", cs);
  if (range == 0) return fail(c, PBIG_E_ARG, "range must be > 0");
  pbig::U32Params p;
  p.dims = (uint32_t)dims;
  p.shift = 0;
  while ((1u << p.shift) < cs) ++p.shift;
  p.nrc = (uint32_t)(((uint64_t)range + cs - 1) / cs);
  p.nrc2 = p.nrc * p.nrc;
  const uint64_t cells = dims == 2 ? (uint64_t)p.nrc * p.nrc : (uint64_t)p.nrc * p.nrc * p.nrc;
  if (cells > 0x7FFFFFFFull)
    return fail(c, PBIG_E_ARG, "grid of %llu cells is too large (cell size %u, range %u)",
                (unsigned long long)cells, cs, range);
  p.n_cells = (uint32_t)cells;
  int rc;
  if ((rc = ensure_cells(c, p.n_cells))) return rc;
  if ((rc = grow(c, &c->ws.items, &c->ws.cap_items, std::max<uint64_t>(c->ws.cap_items, p.n_cells / 2 + 1024))))
    return rc;
  if ((rc = ensure_u32(c, std::max<uint64_t>(c->uw.cap_memb, (uint64_t)n * 2 + 1024)))) return rc;
  const bool count_only = cap == 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    HIP_TRY(c, pbig::launch_u32(c->ws, c->uw, d_boxes, n, p, d_out, cap, count_only, s));
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, c->ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    const Meta m = *c->h_meta;
    if (m.error & pbig::kErrTooDense)
      return fail(c, PBIG_E_TOO_DENSE, "a cell holds %u boxes, above the %u-tile limit", m.max_per_cell,
                  pbig::kMaxTilesPerCell);
    if (m.error & pbig::kErrWorkspace) {
      if (attempt == 2) return fail(c, PBIG_E_NOMEM, "workspace did not fit after regrowth");
      if ((rc = ensure_u32(c, m.memberships + m.memberships / 16 + 1024))) return rc;
      if ((rc = grow(c, &c->ws.items, &c->ws.cap_items, m.items + m.items / 16 + 1024))) return rc;
      continue;
    }
    c->stats = pbig_hip_stats{};
    c->stats.cell_size_num = cs;
    c->stats.cell_size_den = 1;
    c->stats.cells_per_row = p.nrc;
    c->stats.n_boxes = n;
    c->stats.n_cells = p.n_cells;
    c->stats.memberships = m.memberships;
    c->stats.tile_pairs = m.items;
    c->stats.n_pairs = m.n_pairs;
    c->stats.max_per_cell = m.max_per_cell;
    c->stats.path = PBIG_PATH_GENERAL;
    *n_pairs = m.n_pairs;
    if (!count_only && m.n_pairs > cap)
      return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu", (unsigned long long)m.n_pairs,
                  (unsigned long long)cap);
    return PBIG_OK;
  }
  return fail(c, PBIG_E_HIP, "unreachable");
}

}  // namespace

extern "C" {

// decideCellSize (pbig.cu:1285-1357) for unsigned boxes of `dims`
// dimensions: the extent xmax - xmin is taken in unsigned arithmetic and
// converted to float (avgX, pbig.cu:1255-1260), accumulated with the int
// init value of the reference's transform_reduce, divided by the sample
// count in integers; then the loop of the float rule with NMR_BOUND = 4
// (2-D) or 8 (3-D) and nc = nrc^dims, the third factor multiplying the double
// nmr (pbig.cu:1280-1325).
uint32_t pbig_hip_reference_cell_size_u32(const uint32_t* h_boxes, uint32_t n, int dims, uint32_t range) {
  int64_t cs = 1;
  while (cs * 2 < (int64_t)range) cs *= 2;
  const int samples = n < 1000 ? (int)n : 1000;
  if (!h_boxes || samples == 0 || (dims != 2 && dims != 3)) return (uint32_t)cs;
  int acc[3] = {0, 0, 0};
  for (int k = 0; k < samples; ++k) {
    const uint32_t* b = h_boxes + (size_t)k * 2 * dims;
    for (int d = 0; d < dims; ++d) acc[d] = (int)((float)acc[d] + (float)(b[2 * d + 1] - b[2 * d]));
  }
  const float x = (float)(acc[0] / samples), y = (float)(acc[1] / samples), z = (float)(acc[2] / samples);
  const double bound = dims == 2 ? 4.0 : 8.0;
  while (true) {
    const int64_t nrc = ((int64_t)range + cs - 1) / cs;
    int64_t nc = nrc * nrc;
    double nmr = (double)((1 + x / (float)cs) * (1 + y / (float)cs));
    if (dims == 3) {
      nc *= nrc;
      nmr *= (double)(1 + z / (float)cs);
    }
    const int per_cell = (int)(nmr * (double)n / (double)nc);
    if ((float)cs < x) break;
    if (per_cell < 32 || nmr > bound) break;
    if (cs == 1) break;
    cs /= 2;
  }
  return (uint32_t)(cs * 2);
}

int pbig_hip_check_u32(pbig_hip_ctx* c, const uint32_t* d_boxes, uint32_t n, int dims, uint32_t range,
                       uint32_t cell_size, uint32_t* d_pairs, uint64_t cap, uint64_t* n_pairs, void* stream) {
  if (!c) return PBIG_E_ARG;
  if (!n_pairs) return fail(c, PBIG_E_ARG, "n_pairs is null");
  if (dims != 2 && dims != 3) return fail(c, PBIG_E_ARG, "dims must be 2 or 3");
  if (n && !d_boxes) return fail(c, PBIG_E_ARG, "d_boxes is null");
  if (cap && !d_pairs) return fail(c, PBIG_E_ARG, "d_pairs is null with a nonzero capacity");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = pick_stream(c, stream);
  if (cell_size == 0) {
    std::vector<uint32_t> head((size_t)std::min<uint32_t>(n, 1000) * 2 * dims);
    if (!head.empty()) {
      HIP_TRY(c, hipMemcpyAsync(head.data(), d_boxes, head.size() * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(c, hipStreamSynchronize(s));
    }
    cell_size = pbig_hip_reference_cell_size_u32(head.data(), n, dims, range);
  }
  return check_u32(c, d_boxes, n, dims, range, cell_size, (uint2*)d_pairs, cap, n_pairs, s);
}

// The unsigned host entry points' device part: RectInfo {CoordInfo (2*dims
// unsigned), rid} packed into pinned memory over the host threads (rid ==
// index checked), uploaded, checked; the pairs are then in c->d_pairs.
static int host_check_u32(pbig_hip_ctx* c, const void* h_rects, uint32_t n, int dims, uint32_t range,
                          uint32_t cell_size, uint64_t* n_pairs) {
  if (!c) return PBIG_E_ARG;
  if (!n_pairs) return fail(c, PBIG_E_ARG, "n_pairs is null");
  if (dims != 2 && dims != 3) return fail(c, PBIG_E_ARG, "dims must be 2 or 3");
  if (n && !h_rects) return fail(c, PBIG_E_ARG, "h_rects is null");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t w = 2 * (size_t)dims;
  const uint64_t words = std::max<uint64_t>((uint64_t)n * w, 1);
  if (c->cap_h_in32 < words || !c->h_in32) {
    if (c->h_in32) HIP_TRY(c, hipHostFree(c->h_in32));
    c->h_in32 = nullptr;
    c->cap_h_in32 = 0;
    HIP_TRY(c, hipHostMalloc((void**)&c->h_in32, words * sizeof(uint32_t), hipHostMallocDefault));
    c->cap_h_in32 = words;
  }
  const uint32_t* src = (const uint32_t*)h_rects;
  uint32_t* coords = c->h_in32;
  uint64_t bad = 0;
  int code = 0;
  if (pack_records(
          c, n,
          [&](uint32_t i) {
            const uint32_t* r = src + (size_t)i * (w + 1);
            if (r[w] != i) return 1;
            memcpy(coords + (size_t)i * w, r, w * 4);
            return 0;
          },
          &bad, &code))
    return fail(c, PBIG_E_ARG, "rect %llu has rid %u; rid must equal the index", (unsigned long long)bad,
                src[bad * (w + 1) + w]);
  int rc;
  if ((rc = grow(c, &c->d_in32, &c->cap_in32, words))) return rc;
  if (n) HIP_TRY(c, hipMemcpyAsync(c->d_in32, coords, (size_t)n * w * 4, hipMemcpyHostToDevice, s));
  if (cell_size == 0) cell_size = pbig_hip_reference_cell_size_u32(coords, n, dims, range);
  uint64_t want = std::max<uint64_t>(c->cap_pairs, std::max<uint64_t>(1024, (uint64_t)n));
  if ((rc = grow(c, &c->d_pairs, &c->cap_pairs, want))) return rc;
  uint64_t np = 0;
  rc = check_u32(c, c->d_in32, n, dims, range, cell_size, c->d_pairs, c->cap_pairs, &np, s);
  if (rc == PBIG_E_OVERFLOW) {
    if ((rc = grow(c, &c->d_pairs, &c->cap_pairs, np + np / 8 + 1024))) return rc;
    rc = check_u32(c, c->d_in32, n, dims, range, cell_size, c->d_pairs, c->cap_pairs, &np, s);
  }
  if (rc) return rc;
  *n_pairs = np;
  return PBIG_OK;
}

int pbig_hip_check_host_u32(pbig_hip_ctx* c, const void* h_rects, uint32_t n, int dims, uint32_t range,
                            uint32_t cell_size, uint32_t* h_pairs, uint64_t cap, uint64_t* n_pairs) {
  if (c && cap && !h_pairs) return fail(c, PBIG_E_ARG, "h_pairs is null with a nonzero capacity");
  int rc;
  if ((rc = host_check_u32(c, h_rects, n, dims, range, cell_size, n_pairs))) return rc;
  const uint64_t np = *n_pairs;
  if ((rc = stream_pairs(c, c->d_pairs, std::min<uint64_t>(np, cap), c->stream,
                         [&](const uint2* chunk, uint64_t len, uint64_t first) {
                           return copy_out(c, chunk, len, h_pairs + 2 * first);
                         })))
    return rc;
  if (np > cap)
    return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu", (unsigned long long)np,
                (unsigned long long)cap);
  return PBIG_OK;
}

int pbig_hip_check_host_u32_each(pbig_hip_ctx* c, const void* h_rects, uint32_t n, int dims, uint32_t range,
                                 uint32_t cell_size, pbig_hip_pairs_fn fn, void* user, uint64_t* n_pairs) {
  if (c && !fn) return fail(c, PBIG_E_ARG, "fn is null");
  int rc;
  if ((rc = host_check_u32(c, h_rects, n, dims, range, cell_size, n_pairs))) return rc;
  return stream_pairs(c, c->d_pairs, *n_pairs, c->stream, [&](const uint2* chunk, uint64_t len, uint64_t) {
    fn(user, (const uint32_t*)chunk, len);
    return PBIG_OK;
  });
}

}  // extern "C"

