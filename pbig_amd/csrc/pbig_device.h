// Shared definitions of the device kernels and their host launcher.
#ifndef PBIG_DEVICE_H_
#define PBIG_DEVICE_H_

#include <stdint.h>
#include <hip/hip_runtime.h>

namespace pbig {

constexpr uint32_t kTile = 64;              // boxes per tile: one per lane of a wave64
constexpr uint32_t kMaxTilesPerCell = 4096;  // 262,144 boxes in one cell
constexpr int kScanPerThread = 16;
constexpr uint32_t kScanTile = 256 * kScanPerThread;  // cells per scan block
constexpr uint32_t kStage = 512;            // pairs staged in LDS per wave
constexpr uint32_t kFlagX = 0x80000000u;    // box starts in this cell's column
constexpr uint32_t kFlagY = 0x40000000u;    // box starts in this cell's row
constexpr uint32_t kIdMask = 0x3FFFFFFFu;
constexpr uint32_t kMaxBoxes = 0x3FFFFFFFu;

constexpr unsigned long long kErrInvalidBox = 1ull;
constexpr unsigned long long kErrWorkspace = 2ull;
constexpr unsigned long long kErrTooDense = 4ull;

struct GridParams {
  float inv;      // 1 / cell side, rounded once on the host
  uint32_t nrc;   // cells per row (and per column)
};

// Device-resident result words of one call.
struct Meta {
  unsigned long long n_pairs;
  unsigned long long error;
  unsigned long long memberships;
  unsigned long long items;
  unsigned int max_per_cell;
  // Binned path: the launch's boxes that span more than two fine cells on an
  // axis (any: the scatter places every membership by loops), and its box
  // count (saturated at 2^32 - 1), from which every kernel of the launch
  // picks the same test schedule.
  unsigned int wide;
  // A re-binning level planned on the device: the sizes its plan kernel
  // wrote (also when they exceed the level's workspace, which plan_err then
  // flags), read by the level's other kernels.
  unsigned int dev_inst, dev_chunks, dev_bins, dev_nbmax;
  unsigned long long dev_hist, dev_off;
  // Direct output shared by a launch's levels: the pair counter every
  // level's test kernel reserves its output from (device-planned levels).
  unsigned long long* out_count;
  unsigned int plan_err;  // the level's plan did not fit its workspace
  unsigned int boxes;
  unsigned int pad[2];
};
static_assert(sizeof(Meta) == 96, "Meta layout");

struct ScanPartial {
  unsigned long long memberships;
  unsigned long long items;
  unsigned int max_count;
  unsigned int pad;
};

// Device workspace of one context (all sizes are element counts).
struct Workspace {
  uint32_t* cnt = nullptr;        // per-cell count, then scatter cursor
  uint32_t* start = nullptr;      // per-cell exclusive start
  ScanPartial* partials = nullptr;
  uint4* items = nullptr;         // {start, count, row tile, column tile}
  float4* mbox = nullptr;         // cell-ordered box copies
  uint32_t* mid = nullptr;        // cell-ordered ids | start-cell flags
  uint4* strips = nullptr;        // cells laid out from the bins: {bin, first cell, end cell, 0} per strip
  uint32_t* slc = nullptr;        // cells laid out from the bins: per bin a slice counter, then per
                                  // counting slice of a large bin its fine counts
  Meta* meta = nullptr;
  uint64_t cap_cells = 0, cap_partials = 0, cap_items = 0, cap_memb = 0, cap_strips = 0, cap_slc = 0;
  uint32_t test_blocks = 0;
};

}  // namespace pbig

namespace pbig {
// Number of kernels of the plain path, in launch order:
// clear, count, scan_reduce, scan_apply, scatter, test.
constexpr int kPlainKernels = 6;

// Optional per-kernel timing: when `events` is non-null, events[k] is
// recorded before kernel k and events[kPlainKernels] after the last one.
// The general path's first and scan stages, shared with the unsigned mode:
// clear the cell counters and result words; scan the counts into starts and
// 64 x 64 tile-pair items (cnt is reset to serve as the scatter cursor).
hipError_t launch_clear(const Workspace& ws, uint32_t n_cells, hipStream_t s);
hipError_t launch_scan(const Workspace& ws, uint32_t n_cells, uint64_t memb_cap, hipStream_t s);

hipError_t launch_plain(const Workspace& ws, const float4* boxes, uint32_t n, GridParams gp,
                        uint32_t n_cells, uint2* out, uint64_t cap, hipStream_t s,
                        hipEvent_t* events = nullptr);
// The result words of a call zeroed (the whole Meta), and n_words words at
// `words` (optional).
hipError_t launch_meta_clear(Meta* meta, hipStream_t s, uint32_t* words = nullptr, uint32_t n_words = 0);
// k_test over ws's tile-pair items, with the counts and pair counter in meta.
hipError_t launch_tile_test(const Workspace& ws, Meta* meta, uint2* out, uint64_t cap, hipStream_t s);
}  // namespace pbig

#endif  // PBIG_DEVICE_H_
