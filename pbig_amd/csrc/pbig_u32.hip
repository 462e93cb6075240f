// Reference-compatible unsigned mode, written for gfx950 (wave64): the
// reference's shipped build (CType = unsigned int, DIMS = 2 or 3,
// src/pbig/rect.h:4-12, config.cmake:8), restated literally so that its
// logged answers are reproduced:
//
//   cell of a coordinate   v >> shift, cs = 2^shift (pbig.cu:185-216)
//   cell id                gid = nrc^2 * z + nrc * y + x in 32-bit unsigned
//                          arithmetic, with int loop counters; gid >= nCells
//                          is skipped (pbig.cu:199-214, 398-420)
//   overlap                every dimension's two unsigned <= tests
//                          (calcOverlaps == 2 * DIMS, pbig.cu:562-577)
//   reference cell         the cell id of (max xmin, max ymin, max zmin)
//                          must equal the cell being tested
//                          (calcRightToWrite, pbig.cu:579-603)
//
// Pipeline (the general path's clear / scan-reduce / scan-apply kernels
// build the per-cell starts and the 64 x 64 tile-pair items):
//   k_u32_count    box -> covered cells, one device atomic per membership
//   k_u32_scatter  box -> cell lists: the box's coordinates, its id and the
//                  cell id next to each other
//   k_u32_test     one wave per tile pair: row boxes in VGPRs, column boxes
//                  wave-uniform (scalar loads), the literal reference-cell
//                  test, and either a per-lane count or ballot/mbcnt
//                  compaction of the pairs

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_device.h"
#include "pbig_u32.h"

namespace pbig {

// Cell ids covered by box b (2 * dims coordinates), fn(gid) for each, in the
// reference's loop order and arithmetic.
template <typename Fn>
__device__ __forceinline__ void u32_cells(const uint32_t* __restrict__ b, const U32Params& p, Fn&& fn) {
  const int x0 = (int)(b[0] >> p.shift), x1 = (int)(b[1] >> p.shift);
  const int y0 = (int)(b[2] >> p.shift), y1 = (int)(b[3] >> p.shift);
  int z0 = 0, z1 = 0;
  if (p.dims == 3) {
    z0 = (int)(b[4] >> p.shift);
    z1 = (int)(b[5] >> p.shift);
  }
  for (int k = z0; k <= z1; ++k) {
    const uint32_t product = p.nrc2 * (uint32_t)k;
    for (int j = y0; j <= y1; ++j) {
      const uint32_t m_gid = product + p.nrc * (uint32_t)j;
      for (int i = x0; i <= x1; ++i) {
        const uint32_t gid = m_gid + (uint32_t)i;
        if (gid >= p.n_cells) continue;
        fn(gid);
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_u32_count(const uint32_t* __restrict__ boxes, uint32_t n, U32Params p,
                                                    uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t b[6];
  const uint32_t w = 2 * p.dims;
  for (uint32_t d = 0; d < w; ++d) b[d] = boxes[(uint64_t)i * w + d];
  u32_cells(b, p, [&](uint32_t gid) { atomicAdd(&cnt[gid], 1u); });
}

__global__ __launch_bounds__(256) void k_u32_scatter(const uint32_t* __restrict__ boxes, uint32_t n, U32Params p,
                                                      uint32_t* __restrict__ cursor,
                                                      const uint32_t* __restrict__ start,
                                                      uint4* __restrict__ mxy, uint2* __restrict__ mz,
                                                      uint32_t* __restrict__ mid, uint32_t* __restrict__ mcell,
                                                      uint64_t memb_cap) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t b[6] = {0, 0, 0, 0, 0, 0};
  const uint32_t w = 2 * p.dims;
  for (uint32_t d = 0; d < w; ++d) b[d] = boxes[(uint64_t)i * w + d];
  u32_cells(b, p, [&](uint32_t gid) {
    const uint64_t pos = (uint64_t)start[gid] + atomicAdd(&cursor[gid], 1u);
    if (pos < memb_cap) {
      mxy[pos] = make_uint4(b[0], b[1], b[2], b[3]);
      mz[pos] = make_uint2(b[4], b[5]);
      mid[pos] = i;
      mcell[pos] = gid;
    }
  });
}

// Stage of one wave's pairs, flushed with one reservation on the pair counter.
__device__ __forceinline__ void u32_flush(uint2* stage, uint32_t& stage_n, uint2* __restrict__ out, uint64_t cap,
                                          Meta* __restrict__ meta) {
  const int lane = threadIdx.x & 63;
  uint64_t g = 0;
  if (lane == 0) g = atomicAdd((unsigned long long*)&meta->n_pairs, (unsigned long long)stage_n);
  g = __shfl(g, 0, 64);
  for (uint32_t k = lane; k < stage_n; k += 64) {
    const uint64_t dst = g + k;
    if (dst < cap) out[dst] = stage[k];
  }
  stage_n = 0;
}

__global__ __launch_bounds__(256) void k_u32_test(const uint4* __restrict__ items, const uint4* __restrict__ mxy,
                                                   const uint2* __restrict__ mz, const uint32_t* __restrict__ mid,
                                                   const uint32_t* __restrict__ mcell, U32Params p,
                                                   uint2* __restrict__ out, uint64_t cap, bool count_only,
                                                   Meta* __restrict__ meta) {
  __shared__ uint2 stage_all[4][kStage];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  uint2* stage = stage_all[wib];
  uint32_t stage_n = 0;
  uint64_t found = 0;
  const uint64_t n_items = meta->items;
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  const bool ok = meta->error == 0ull;
  const bool three = p.dims == 3;

  for (uint64_t it = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wib; ok && it < n_items; it += nwaves) {
    const uint4 d = items[it];
    const uint32_t base = __builtin_amdgcn_readfirstlane(d.x);
    const uint32_t n = __builtin_amdgcn_readfirstlane(d.y);
    const uint32_t rt = __builtin_amdgcn_readfirstlane(d.z);
    const uint32_t ct = __builtin_amdgcn_readfirstlane(d.w);
    const uint32_t cell = __builtin_amdgcn_readfirstlane(mcell[base]);

    const uint32_t i = rt * kTile + lane;
    const bool valid = i < n;
    uint4 a = make_uint4(1u, 0u, 1u, 0u);
    uint2 az = make_uint2(0u, 0u);
    uint32_t aid = 0;
    if (valid) {
      a = mxy[base + i];
      az = mz[base + i];
      aid = mid[base + i];
    }
    const uint32_t j0 = ct * kTile;
    const uint32_t j1 = (j0 + kTile < n) ? j0 + kTile : n;
    const bool diag = rt == ct;
    for (uint32_t j = j0; j < j1; ++j) {
      const uint4 b = mxy[base + j];
      const uint2 bz = mz[base + j];
      bool hit = valid && (a.x <= b.y) && (b.x <= a.y) && (a.z <= b.w) && (b.z <= a.w);
      if (three) hit = hit && (az.x <= bz.y) && (bz.x <= az.y);
      // calcRightToWrite: cell id of the maximum minima
      const uint32_t fx = (a.x > b.x ? a.x : b.x) >> p.shift;
      const uint32_t fy = (a.z > b.z ? a.z : b.z) >> p.shift;
      const uint32_t fz = three ? ((az.x > bz.x ? az.x : bz.x) >> p.shift) : 0u;
      hit = hit && (p.nrc2 * fz + p.nrc * fy + fx == cell);
      if (diag) hit = hit && (j > i);
      if (count_only) {
        found += hit ? 1u : 0u;
        continue;
      }
      const uint64_t mask = __ballot(hit);
      if (mask) {
        const uint32_t pos = stage_n + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (hit) {
          const uint32_t bid = mid[base + j];
          stage[pos] = make_uint2(aid < bid ? aid : bid, aid < bid ? bid : aid);
        }
        stage_n += (uint32_t)__popcll(mask);
        if (stage_n > kStage - 64) u32_flush(stage, stage_n, out, cap, meta);
      }
    }
  }
  if (count_only) {
#pragma unroll
    for (int dd = 32; dd > 0; dd >>= 1) found += __shfl_xor(found, dd, 64);
    if (lane == 0 && found) atomicAdd((unsigned long long*)&meta->n_pairs, (unsigned long long)found);
  } else if (stage_n) {
    u32_flush(stage, stage_n, out, cap, meta);
  }
}

hipError_t launch_u32(const Workspace& ws, const U32Workspace& uw, const uint32_t* boxes, uint32_t n,
                      const U32Params& p, uint2* out, uint64_t cap, bool count_only, hipStream_t s) {
  const uint32_t nb_boxes = (n + 255) / 256;
  hipError_t e = launch_clear(ws, p.n_cells, s);
  if (e != hipSuccess) return e;
  if (n) k_u32_count<<<nb_boxes, 256, 0, s>>>(boxes, n, p, ws.cnt);
  e = launch_scan(ws, p.n_cells, uw.cap_memb, s);
  if (e != hipSuccess) return e;
  if (n)
    k_u32_scatter<<<nb_boxes, 256, 0, s>>>(boxes, n, p, ws.cnt, ws.start, uw.mxy, uw.mz, uw.mid, uw.mcell,
                                           uw.cap_memb);
  k_u32_test<<<ws.test_blocks, 256, 0, s>>>(ws.items, uw.mxy, uw.mz, uw.mid, uw.mcell, p, out, cap, count_only,
                                            ws.meta);
  return hipGetLastError();
}

}  // namespace pbig
