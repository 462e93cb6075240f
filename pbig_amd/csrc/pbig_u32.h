// Reference-compatible unsigned mode (CType = unsigned int, DIMS = 2 or 3):
// device-side types and launcher.
#ifndef PBIG_U32_H_
#define PBIG_U32_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_device.h"

namespace pbig {

struct U32Params {
  uint32_t shift;    // cs = 2^shift
  uint32_t nrc;      // cells per row = ceil(range / cs)
  uint32_t nrc2;     // nrc * nrc (32-bit, as the reference computes it)
  uint32_t n_cells;  // nrc^dims (32-bit)
  uint32_t dims;     // 2 or 3
};

struct U32Workspace {
  uint4* mxy = nullptr;      // per membership: xmin, xmax, ymin, ymax
  uint2* mz = nullptr;       // per membership: zmin, zmax (3-D)
  uint32_t* mid = nullptr;   // per membership: box index
  uint32_t* mcell = nullptr; // per membership: cell id
  uint64_t cap_memb = 0;
};

// Clear, count, scan (starts + tile-pair items, in ws), scatter, test.
// count_only: the pairs are counted into the result words, not written.
hipError_t launch_u32(const Workspace& ws, const U32Workspace& uw, const uint32_t* boxes, uint32_t n,
                      const U32Params& p, uint2* out, uint64_t cap, bool count_only, hipStream_t s);

}  // namespace pbig

#endif  // PBIG_U32_H_
