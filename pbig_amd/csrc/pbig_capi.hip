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
#include <string>
#include <vector>

#include "pbig_binned.h"
#include "pbig_device.h"
#include "pbig_hip.h"

using pbig::GridParams;
using pbig::Meta;
using pbig::ScanPartial;
using pbig::Workspace;
using pbig::BinnedParams;
using pbig::BinnedWorkspace;

struct pbig_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  Workspace ws;
  Meta* h_meta = nullptr;          // pinned mirror of ws.meta
  // Host-path staging.
  float4* d_in = nullptr;
  uint64_t cap_in = 0;
  float4* h_in = nullptr;          // pinned
  uint64_t cap_h_in = 0;
  uint2* d_pairs = nullptr;
  uint64_t cap_pairs = 0;
  BinnedWorkspace bws;
  pbig_hip_config cfg{PBIG_PATH_AUTO, 3000, 0, 0};
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
constexpr int kEventsPerCall = 8;
}

namespace {

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

// Plan of the binned path for n boxes, domain side `range`, fine cell side
// `cell`: K fine cells per coarse bin side (1..32) such that a coarse bin
// holds about cfg.bin_target boxes, with at most kMaxBins bins.
bool plan_binned(const pbig_hip_ctx* c, uint32_t n, double range, float cell, BinnedParams* bp) {
  const double nf0 = std::max(1.0, std::ceil(range / (double)cell));
  const double target = std::max<double>(16.0, c->cfg.bin_target);
  uint32_t k = (uint32_t)std::lround(std::sqrt(target / std::max<double>(n, 1)) * nf0);
  k = std::min<uint32_t>(pbig::kMaxK, std::max<uint32_t>(1, k));
  uint32_t nbx = 0;
  while (true) {
    const double nbxd = std::ceil(nf0 / k);
    if (nbxd * nbxd <= pbig::kMaxBins) { nbx = (uint32_t)nbxd; break; }
    if (k == pbig::kMaxK) return false;
    ++k;
  }
  bp->inv_f = 1.0f / cell;
  bp->k = k;
  bp->k_magic = ((1ull << 32) + k - 1) / k;
  bp->nbx = nbx;
  bp->nb = nbx * nbx;
  bp->nf = nbx * k;
  uint32_t nw = (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(8, ((uint64_t)n / 16384 + 7) / 8 * 8));
  bp->nw = nw;
  bp->chunk = std::max<uint32_t>(1024, (uint32_t)(((uint64_t)n + nw - 1) / nw));
  bp->test_blocks = std::max<uint32_t>((uint32_t)c->n_cu, (bp->nb + pbig::kMaxBinsPerGroup - 1) / pbig::kMaxBinsPerGroup);
  return true;
}

int ensure_binned(pbig_hip_ctx* c, const BinnedParams& bp, uint64_t bucket) {
  BinnedWorkspace& w = c->bws;
  int rc;
  uint64_t hist = (uint64_t)bp.nw * (bp.nb + 1), off = (uint64_t)bp.nw * bp.nb;
  if ((rc = grow(c, &w.hist, &w.cap_hist, hist))) return rc;
  if ((rc = grow(c, &w.off, &w.cap_off, off))) return rc;
  if (w.cap_bins < bp.nb || !w.total) {
    for (uint32_t** q : {&w.total, &w.bin_start, &w.overfull})
      if (*q) HIP_TRY(c, hipFree(*q));
    w.total = w.bin_start = w.overfull = nullptr;
    w.cap_bins = 0;
    uint64_t d1 = 0, d2 = 0, d3 = 0;
    if ((rc = grow(c, &w.total, &d1, bp.nb))) return rc;
    if ((rc = grow(c, &w.bin_start, &d2, bp.nb))) return rc;
    if ((rc = grow(c, &w.overfull, &d3, bp.nb))) return rc;
    w.cap_bins = bp.nb;
  }
  if ((rc = grow(c, &w.wg_count, &w.cap_wgc, bp.test_blocks))) return rc;
  uint64_t scratch = (uint64_t)bp.test_blocks * w.cap_wg;
  if (w.cap_scratch < scratch || !w.scratch) {
    if (w.scratch) HIP_TRY(c, hipFree(w.scratch));
    w.scratch = nullptr;
    w.cap_scratch = 0;
    uint64_t d = 0;
    if ((rc = grow(c, &w.scratch, &d, scratch))) return rc;
    w.cap_scratch = scratch;
  }
  if (w.cap_bucket < bucket || !w.bbox) {
    if (w.bbox) HIP_TRY(c, hipFree(w.bbox));
    if (w.bid) HIP_TRY(c, hipFree(w.bid));
    w.bbox = nullptr;
    w.bid = nullptr;
    w.cap_bucket = 0;
    uint64_t d1 = 0, d2 = 0;
    if ((rc = grow(c, &w.bbox, &d1, bucket))) return rc;
    if ((rc = grow(c, &w.bid, &d2, bucket))) return rc;
    w.cap_bucket = bucket;
  }
  w.meta = c->ws.meta;
  return PBIG_OK;
}

int prepare_binned(pbig_hip_ctx* c, uint32_t n, const BinnedParams& bp) {
  uint64_t bucket = std::max<uint64_t>(c->bws.cap_bucket, (uint64_t)n + n / 4 + 4096);
  // Pair region per test workgroup: at least the previous need, first guess
  // one pair per box spread over the workgroups, plus slack.
  uint64_t per = std::max<uint64_t>(8192, ((uint64_t)n + bp.test_blocks - 1) / bp.test_blocks * 2);
  c->bws.cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, std::max<uint64_t>(c->bws.cap_wg, per));
  return ensure_binned(c, bp, bucket);
}

hipEvent_t* next_events(pbig_hip_ctx* c, int path) {
  if (c->timing_used >= c->timing_cap) return nullptr;
  c->timing_path = path;
  return &c->events[(size_t)c->timing_used++ * kEventsPerCall];
}

// Binned path, synchronous.  Returns PBIG_OK, an error, or 1 when a coarse
// bin did not fit in LDS (the caller then runs the general path).
int run_binned_sync(pbig_hip_ctx* c, const float4* d_boxes, uint32_t n, const BinnedParams& bp, uint2* d_out,
                    uint64_t cap, hipStream_t s, uint64_t* n_pairs) {
  int rc;
  if ((rc = prepare_binned(c, n, bp))) return rc;
  for (int attempt = 0; attempt < 2; ++attempt) {
    HIP_TRY(c, pbig::launch_binned(c->bws, d_boxes, n, bp, d_out, cap, s, next_events(c, PBIG_PATH_BINNED)));
    HIP_TRY(c, hipMemcpyAsync(c->h_meta, c->ws.meta, sizeof(Meta), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    const Meta m = *c->h_meta;
    if (m.error & pbig::kErrInvalidBox)
      return fail(c, PBIG_E_INVALID_BOX, "input has a box with xmin > xmax, ymin > ymax or a NaN");
    if (m.error & pbig::kErrWorkspace) {
      if (attempt == 1) return fail(c, PBIG_E_NOMEM, "bin workspace did not fit after regrowth");
      if (m.items > c->bws.cap_wg)
        c->bws.cap_wg = (uint32_t)std::min<uint64_t>(0x7FFFFFFFull, m.items + m.items / 4 + 1024);
      if ((rc = ensure_binned(c, bp, std::max<uint64_t>(c->bws.cap_bucket, m.memberships + m.memberships / 16 + 4096))))
        return rc;
      continue;
    }
    if (m.error & pbig::kErrOverfull) return 1;
    c->last_path = PBIG_PATH_BINNED;
    c->stats.n_boxes = n;
    c->stats.n_cells = (uint64_t)bp.nf * bp.nf;
    c->stats.cells_per_row = bp.nf;
    c->stats.memberships = m.memberships;
    c->stats.tile_pairs = 0;
    c->stats.n_pairs = m.n_pairs;
    c->stats.max_per_cell = 0;
    c->stats.path = PBIG_PATH_BINNED;
    c->stats.bins_per_row = bp.nbx;
    c->stats.cells_per_bin_side = bp.k;
    *n_pairs = m.n_pairs;
    if (m.n_pairs > cap)
      return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu",
                  (unsigned long long)m.n_pairs, (unsigned long long)cap);
    return PBIG_OK;
  }
  return fail(c, PBIG_E_HIP, "unreachable");
}

// Either path, synchronous, by the context's configuration.
int run_any_sync(pbig_hip_ctx* c, const float4* d_boxes, uint32_t n, double range, float cell, uint2* d_out,
                 uint64_t cap, hipStream_t s, uint64_t* n_pairs) {
  int rc;
  if (c->cfg.path != PBIG_PATH_GENERAL) {
    BinnedParams bp;
    if (plan_binned(c, n, range, cell, &bp)) {
      rc = run_binned_sync(c, d_boxes, n, bp, d_out, cap, s, n_pairs);
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
  rc = run_sync(c, d_boxes, n, gp, n_cells, d_out, cap, s, n_pairs);
  if (rc == PBIG_OK || rc == PBIG_E_OVERFLOW) {
    c->last_path = PBIG_PATH_GENERAL;
    c->stats.path = PBIG_PATH_GENERAL;
    c->stats.bins_per_row = 0;
    c->stats.cells_per_bin_side = 0;
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
  c->ws.test_blocks = (uint32_t)c->n_cu * 8;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return PBIG_E_HIP; }
  if (hipMalloc((void**)&c->ws.meta, sizeof(Meta)) != hipSuccess) { delete c; return PBIG_E_NOMEM; }
  if (hipHostMalloc((void**)&c->h_meta, sizeof(Meta), hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(c->ws.meta);
    delete c;
    return PBIG_E_NOMEM;
  }
  memset(c->h_meta, 0, sizeof(Meta));
  *out = c;
  return PBIG_OK;
}

int pbig_hip_destroy(pbig_hip_ctx* c) {
  if (!c) return PBIG_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  Workspace& w = c->ws;
  const BinnedWorkspace& b = c->bws;
  for (void* p : {(void*)w.cnt, (void*)w.start, (void*)w.partials, (void*)w.items, (void*)w.mbox,
                  (void*)w.mid, (void*)w.meta, (void*)c->d_in, (void*)c->d_pairs, (void*)b.hist, (void*)b.off,
                  (void*)b.total, (void*)b.bin_start, (void*)b.overfull, (void*)b.bbox, (void*)b.bid,
                  (void*)b.scratch, (void*)b.wg_count})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
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
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
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
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  GridParams gp;
  uint32_t n_cells;
  int rc;
  if ((rc = grid_of(c, range, cell_size, &gp, &n_cells))) return rc;
  BinnedParams bp;
  bool binned = c->cfg.path == PBIG_PATH_BINNED ||
                (c->cfg.path == PBIG_PATH_AUTO && c->last_path != PBIG_PATH_GENERAL);
  if (binned && !plan_binned(c, n, range, cell_size, &bp)) {
    if (c->cfg.path == PBIG_PATH_BINNED)
      return fail(c, PBIG_E_ARG, "the binned path cannot plan this grid (more than %u bins)", pbig::kMaxBins);
    binned = false;
  }
  if (binned) {
    if ((rc = prepare_binned(c, n, bp))) return rc;
    HIP_TRY(c, pbig::launch_binned(c->bws, (const float4*)d_boxes, n, bp, (uint2*)d_pairs, cap, s,
                                   next_events(c, PBIG_PATH_BINNED)));
  } else {
    if ((rc = prepare(c, n, n_cells))) return rc;
    HIP_TRY(c, pbig::launch_plain(c->ws, (const float4*)d_boxes, n, gp, n_cells, (uint2*)d_pairs, cap, s,
                                  next_events(c, PBIG_PATH_GENERAL)));
  }
  if (d_result)
    HIP_TRY(c, hipMemcpyAsync(d_result, c->ws.meta, 2 * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  return PBIG_OK;
}

int pbig_hip_check_host(pbig_hip_ctx* c, const void* h_rects, uint32_t n, double range, float cell_size,
                        uint32_t* h_pairs, uint64_t cap, uint64_t* n_pairs) {
  if (!c) return PBIG_E_ARG;
  if (!n_pairs) return fail(c, PBIG_E_ARG, "n_pairs is null");
  if (n > pbig::kMaxBoxes) return fail(c, PBIG_E_ARG, "n = %u exceeds the limit of 2^30-1 boxes", n);
  if (n && !h_rects) return fail(c, PBIG_E_ARG, "h_rects is null");
  if (cap && !h_pairs) return fail(c, PBIG_E_ARG, "h_pairs is null with a nonzero capacity");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // Pack RectInfo {xmin,xmax,ymin,ymax,rid} into 16-byte boxes, checking the
  // reference's contract rid == index and the box orientation.
  if (c->cap_h_in < n || !c->h_in) {
    if (c->h_in) HIP_TRY(c, hipHostFree(c->h_in));
    c->h_in = nullptr;
    c->cap_h_in = 0;
    HIP_TRY(c, hipHostMalloc((void**)&c->h_in, std::max<uint64_t>(n, 1) * sizeof(float4), hipHostMallocDefault));
    c->cap_h_in = std::max<uint64_t>(n, 1);
  }
  const unsigned char* src = (const unsigned char*)h_rects;
  for (uint32_t i = 0; i < n; ++i) {
    float f[4];
    uint32_t rid;
    memcpy(f, src + (size_t)i * 20, 16);
    memcpy(&rid, src + (size_t)i * 20 + 16, 4);
    if (rid != i) return fail(c, PBIG_E_ARG, "rect %u has rid %u; rid must equal the index", i, rid);
    if (!(f[0] <= f[1]) || !(f[2] <= f[3]))
      return fail(c, PBIG_E_INVALID_BOX, "rect %u is inverted or NaN: [%g,%g]x[%g,%g]", i, f[0], f[1], f[2], f[3]);
    c->h_in[i] = make_float4(f[0], f[1], f[2], f[3]);
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
  uint64_t ncopy = std::min<uint64_t>(np, cap);
  if (ncopy) {
    HIP_TRY(c, hipMemcpyAsync(h_pairs, c->d_pairs, ncopy * sizeof(uint2), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
  }
  if (np > cap)
    return fail(c, PBIG_E_OVERFLOW, "%llu pairs do not fit the capacity of %llu", (unsigned long long)np,
                (unsigned long long)cap);
  return PBIG_OK;
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
  const int nk = c->timing_path == PBIG_PATH_BINNED ? pbig::kBinnedKernels : pbig::kPlainKernels;
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
  if (c->cfg.bin_target == 0) c->cfg.bin_target = 3000;
  c->last_path = 0;
  return PBIG_OK;
}

int pbig_hip_last_stats(const pbig_hip_ctx* c, pbig_hip_stats* out) {
  if (!c || !out) return PBIG_E_ARG;
  *out = c->stats;
  return PBIG_OK;
}

}  // extern "C"
