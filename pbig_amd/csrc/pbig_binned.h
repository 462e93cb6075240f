// Definitions shared by the binned kernels and their host launcher.
#ifndef PBIG_BINNED_H_
#define PBIG_BINNED_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_device.h"

namespace pbig {

constexpr uint32_t kMaxBins = 8192;          // coarse bins per instance (LDS histogram)
constexpr uint32_t kMaxFine = 4096;          // fine memberships of one strip held in LDS
constexpr uint32_t kMaxCoarse = 4096;        // coarse memberships of one bin (registers + LDS ids)
constexpr uint32_t kMaxCellsPerBin = 4096;   // k*k fine cells per coarse bin (k <= 64)
constexpr uint32_t kMaxK = 64;
constexpr uint32_t kBinStage = 128;          // staged pairs per wave
constexpr uint32_t kMaxBinsPerGroup = 128;   // bins walked by one persistent test workgroup
constexpr uint32_t kBatch = 8;               // boxes per thread loaded before use (hist, scatter)
constexpr unsigned long long kErrOverfull = 8ull;  // a coarse bin did not fit in LDS

struct BinnedParams {
  float inv_f;          // 1 / fine cell side (rounded once on the host)
  uint32_t nf;          // fine cells per row = nbx * k
  uint32_t k;           // fine cells per coarse bin side (1..32)
  uint32_t nbx;         // coarse bins per row
  uint32_t nb;          // coarse bins = nbx * nbx
  uint32_t nw;          // workgroups of the histogram and scatter kernels (multiple of 8)
  uint32_t chunk;       // boxes per such workgroup
  uint32_t test_blocks; // persistent workgroups of the test kernel
  uint64_t k_magic;     // ceil(2^32 / k): fine index -> coarse bin by multiply
};

struct BinnedWorkspace {
  uint32_t* hist = nullptr;       // nw x (nb + 1): per-workgroup bin counts + error word
  uint32_t* off = nullptr;        // nw x nb: workgroup offset inside each bin
  uint32_t* total = nullptr;      // nb: memberships per bin
  uint32_t* bin_start = nullptr;  // nb: exclusive scan of total
  uint32_t* overfull = nullptr;   // nb: bins that did not fit in LDS
  uint2* scratch = nullptr;       // test_blocks x cap_wg pair regions
  uint32_t* wg_count = nullptr;   // pairs found by each test workgroup
  uint32_t cap_wg = 0;
  uint64_t cap_scratch = 0, cap_wgc = 0;
  float4* bbox = nullptr;         // bin-ordered box copies
  uint32_t* bid = nullptr;        // bin-ordered box ids
  Meta* meta = nullptr;
  uint64_t cap_hist = 0, cap_off = 0, cap_bins = 0, cap_bucket = 0;
};

constexpr int kBinnedKernels = 5;  // hist, scan, scatter, test, compact

size_t binned_test_lds_bytes();
hipError_t binned_init();

hipError_t launch_binned(const BinnedWorkspace& ws, const float4* boxes, uint32_t n, const BinnedParams& p,
                         uint2* out, uint64_t cap, hipStream_t s, hipEvent_t* events = nullptr);

}  // namespace pbig

#endif  // PBIG_BINNED_H_
