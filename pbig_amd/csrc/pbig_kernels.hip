// Device kernels of the plain path, written for gfx950 (CDNA4, wave64).
//
// Pipeline (one instance, all sizes on the device, no host round trip):
//   k_clear        zero the per-cell counters and the result words
//   k_count        box -> covered cells, one atomic increment per membership
//                  (reference d_precounting, src/pbig/pbig.cu:174-221)
//   k_scan_reduce  per-block sums of cell counts and of test work items
//   k_scan_apply   exclusive scan -> cell starts (reference host
//                  thrust::exclusive_scan, pbig.cu:1497-1505) and expansion of
//                  each cell into 64x64 tile-pair work items (the reference's
//                  host sched_phase, pbig.cu:1592-1688, cut cells into
//                  256-box jobs)
//   k_scatter      box -> cell lists, storing the box coordinates next to the
//                  id so the test reads each cell contiguously (reference
//                  d_mapping, pbig.cu:321-427, stores only ids)
//   k_test         persistent waves over the tile-pair items: all-pairs test
//                  with the column box broadcast from scalar registers (a
//                  tile against itself: lanes rotated across the wave), the
//                  reference-cell rule, and ballot/mbcnt compaction into a
//                  per-wave LDS stage flushed with one atomic per flush
//                  (reference d_matching + prefix_scan, pbig.cu:476-543,
//                  642-932)
//
// Bit-exactness: the overlap predicate is the reference's four float32 `<=`
// tests (pbig.cu:562-568).  The cell of a coordinate is
// clamp(floor(v * inv), 0, nrc-1) with inv computed once on the host, the
// same function in every kernel.  Because it is monotone,
// cell(max(a.xmin, b.xmin)) = max(cell(a.xmin), cell(b.xmin)); both boxes lie
// in the current cell, so the reference-cell rule (pbig.cu:579-603) reduces
// to "one of the two boxes starts in this cell's column, and one starts in
// this cell's row".  k_scatter stores those two facts as the top bits of the
// membership id.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_device.h"

namespace pbig {

__device__ __forceinline__ uint32_t cellmap(float v, float inv, uint32_t nrc) {
  float f = floorf(v * inv);
  if (!(f >= 0.0f)) return 0u;
  if (f >= (float)(nrc - 1u)) return nrc - 1u;
  return (uint32_t)f;
}

__device__ __forceinline__ bool box_valid(const float4& b) {
  return (b.x <= b.y) && (b.z <= b.w);  // false for NaN as well
}

__device__ __forceinline__ uint64_t items_of(uint32_t n) {
  if (n < 2) return 0;
  uint64_t nt = (n + kTile - 1) / kTile;
  return nt * (nt + 1) / 2;
}

// ---------------------------------------------------------------------------
__global__ void k_clear(uint32_t* __restrict__ cnt, uint32_t n_cells, Meta* __restrict__ meta) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_cells) cnt[i] = 0u;
  if (i == 0) {
    meta->n_pairs = 0ull;
    meta->error = 0ull;
    meta->memberships = 0ull;
    meta->items = 0ull;
    meta->max_per_cell = 0u;
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_count(const float4* __restrict__ boxes, uint32_t n,
                                                GridParams gp, uint32_t* __restrict__ cnt,
                                                Meta* __restrict__ meta) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 b = boxes[i];
  if (!box_valid(b)) {
    atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrInvalidBox);
    return;
  }
  uint32_t x0 = cellmap(b.x, gp.inv, gp.nrc), x1 = cellmap(b.y, gp.inv, gp.nrc);
  uint32_t y0 = cellmap(b.z, gp.inv, gp.nrc), y1 = cellmap(b.w, gp.inv, gp.nrc);
  for (uint32_t cy = y0; cy <= y1; ++cy) {
    uint32_t row = cy * gp.nrc;
    for (uint32_t cx = x0; cx <= x1; ++cx) atomicAdd(&cnt[row + cx], 1u);
  }
}

// ---------------------------------------------------------------------------
// Scan: block b owns cells [b*kScanTile, (b+1)*kScanTile); 256 threads x 16.

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ cnt, uint32_t n_cells,
                                                      ScanPartial* __restrict__ partials) {
  __shared__ uint64_t sm[4], si[4];
  __shared__ uint32_t smax[4];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
  uint64_t m = 0, it = 0;
  uint32_t mx = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    uint32_t c = base + k;
    uint32_t v = c < n_cells ? cnt[c] : 0u;
    m += v;
    it += items_of(v);
    mx = v > mx ? v : mx;
  }
  m = wave_sum_u64(m);
  it = wave_sum_u64(it);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    uint32_t o = __shfl_xor(mx, d, 64);
    mx = o > mx ? o : mx;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sm[w] = m; si[w] = it; smax[w] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    ScanPartial p;
    p.memberships = sm[0] + sm[1] + sm[2] + sm[3];
    p.items = si[0] + si[1] + si[2] + si[3];
    uint32_t a = smax[0] > smax[1] ? smax[0] : smax[1];
    uint32_t b = smax[2] > smax[3] ? smax[2] : smax[3];
    p.max_count = a > b ? a : b;
    partials[blockIdx.x] = p;
  }
}

__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* __restrict__ cnt, uint32_t n_cells,
                                                     const ScanPartial* __restrict__ partials,
                                                     uint32_t n_partials, uint32_t* __restrict__ start,
                                                     uint4* __restrict__ items, uint64_t item_cap,
                                                     uint64_t memb_cap, Meta* __restrict__ meta) {
  __shared__ uint64_t wm[4], wi[4];
  __shared__ uint32_t wmax[4];
  __shared__ uint64_t off_m, off_i;
  __shared__ uint32_t blk_max;
  // 1. Offset of this block = sum of the partials of the blocks before it.
  uint64_t pm = 0, pi = 0;
  uint32_t pmax = 0;
  const bool last = blockIdx.x == gridDim.x - 1;
  const uint32_t upto = last ? n_partials : blockIdx.x;
  for (uint32_t k = threadIdx.x; k < upto; k += blockDim.x) {
    ScanPartial p = partials[k];
    if (k < blockIdx.x) { pm += p.memberships; pi += p.items; }
    pmax = p.max_count > pmax ? p.max_count : pmax;
  }
  pm = wave_sum_u64(pm);
  pi = wave_sum_u64(pi);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    uint32_t o = __shfl_xor(pmax, d, 64);
    pmax = o > pmax ? o : pmax;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { wm[w] = pm; wi[w] = pi; wmax[w] = pmax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    off_m = wm[0] + wm[1] + wm[2] + wm[3];
    off_i = wi[0] + wi[1] + wi[2] + wi[3];
    uint32_t a = wmax[0] > wmax[1] ? wmax[0] : wmax[1];
    uint32_t b = wmax[2] > wmax[3] ? wmax[2] : wmax[3];
    blk_max = a > b ? a : b;
  }
  __syncthreads();
  // 2. Local values.
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
  uint32_t v[kScanPerThread];
  uint64_t m = 0, it = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    uint32_t c = base + k;
    v[k] = c < n_cells ? cnt[c] : 0u;
    m += v[k];
    it += items_of(v[k]);
  }
  // 3. Block exclusive scan of (m, it).
  uint64_t im = wave_incl_scan_u64(m), ii = wave_incl_scan_u64(it);
  __shared__ uint64_t tm[4], ti[4];
  if (lane == 63) { tm[w] = im; ti[w] = ii; }
  __syncthreads();
  uint64_t em = off_m + im - m, ei = off_i + ii - it;
  for (int k = 0; k < w; ++k) { em += tm[k]; ei += ti[k]; }
  // 4. Starts, cursor reset, item expansion.
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    uint32_t c = base + k;
    if (c < n_cells) {
      start[c] = (uint32_t)(em < memb_cap ? em : memb_cap);
      cnt[c] = 0u;  // reused as the scatter cursor
    }
    const uint32_t nk = v[k];
    if (nk >= 2) {
      const uint32_t nt = (nk + kTile - 1) / kTile;
      const uint64_t ni = items_of(nk);
      if (nt > kMaxTilesPerCell) {
        atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrTooDense);
      } else if (ei + ni <= item_cap && em + nk <= memb_cap) {
        uint64_t e = ei;
        for (uint32_t r = 0; r < nt; ++r)
          for (uint32_t q = r; q < nt; ++q) items[e++] = make_uint4((uint32_t)em, nk, r, q);
      }
      ei += ni;
    }
    em += nk;
  }
  if (last && threadIdx.x == blockDim.x - 1) {
    // The last thread of the last block holds the grand totals.
    meta->memberships = em;
    meta->items = ei;
    meta->max_per_cell = blk_max;
    unsigned long long err = 0;
    if (em > memb_cap) err |= kErrWorkspace;
    if (ei > item_cap) err |= kErrWorkspace;
    if (err) atomicOr((unsigned long long*)&meta->error, err);
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(const float4* __restrict__ boxes, uint32_t n,
                                                  GridParams gp, uint32_t* __restrict__ cursor,
                                                  const uint32_t* __restrict__ start,
                                                  float4* __restrict__ mbox, uint32_t* __restrict__ mid,
                                                  uint64_t memb_cap) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 b = boxes[i];
  if (!box_valid(b)) return;
  uint32_t x0 = cellmap(b.x, gp.inv, gp.nrc), x1 = cellmap(b.y, gp.inv, gp.nrc);
  uint32_t y0 = cellmap(b.z, gp.inv, gp.nrc), y1 = cellmap(b.w, gp.inv, gp.nrc);
  for (uint32_t cy = y0; cy <= y1; ++cy) {
    uint32_t row = cy * gp.nrc;
    uint32_t fy = (cy == y0) ? kFlagY : 0u;
    for (uint32_t cx = x0; cx <= x1; ++cx) {
      uint32_t c = row + cx;
      uint64_t pos = (uint64_t)start[c] + atomicAdd(&cursor[c], 1u);
      if (pos < memb_cap) {
        mbox[pos] = b;
        mid[pos] = i | fy | ((cx == x0) ? kFlagX : 0u);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Per-cell all-pairs test over 64x64 tile pairs.
//
// One wave per item.  Lane l holds row box r*64+l in VGPRs; the column boxes
// are wave-uniform and are read into SGPRs, so every compare is a VALU op with
// a scalar operand and every match mask is a scalar register.  Matches are
// compacted with mbcnt into a per-wave LDS stage of kStage pairs; a full
// stage is flushed with one 64-bit atomic reservation on the pair counter.

// A staged pair as it goes out: both ids without their flag bits, the
// smaller first (the test loops stage the raw words).
__device__ __forceinline__ uint2 pair_out(uint2 v) {
  const uint32_t a = v.x & kIdMask, b = v.y & kIdMask;
  return make_uint2(a < b ? a : b, a < b ? b : a);
}

__device__ __forceinline__ void flush_stage(uint2* stage, uint32_t& stage_n, uint2* __restrict__ out,
                                            uint64_t cap, Meta* __restrict__ meta) {
  const int lane = threadIdx.x & 63;
  uint64_t g = 0;
  if (lane == 0) g = atomicAdd((unsigned long long*)&meta->n_pairs, (unsigned long long)stage_n);
  g = __shfl(g, 0, 64);
  for (uint32_t k = lane; k < stage_n; k += 64) {
    uint64_t dst = g + k;
    if (dst < cap) out[dst] = pair_out(stage[k]);
  }
  stage_n = 0;
}

// An item of two different tiles: rows (lane l holds row box rt*64+l)
// against the column tile broadcast one column at a time.  Every condition is
// a lane mask: the four `<=` tests, the reference-cell rule as (row flags |
// column flags) == both, and the valid rows.
__device__ __forceinline__ void test_columns(const float4& a, uint32_t aid, uint32_t rf, uint64_t vm,
                                             const float4& cb, uint32_t cw, uint32_t ncol, uint2* stage,
                                             uint32_t& stage_n, uint2* __restrict__ out, uint64_t cap,
                                             Meta* __restrict__ meta) {
  for (uint32_t jj = 0; jj < ncol; ++jj) {
    const float bx = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cb.x), (int)jj));
    const float by = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cb.y), (int)jj));
    const float bz = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cb.z), (int)jj));
    const float bwv = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cb.w), (int)jj));
    const uint32_t bw = (uint32_t)__builtin_amdgcn_readlane((int)cw, (int)jj);
    const uint64_t m = __ballot(a.x <= by) & __ballot(bx <= a.y) & __ballot(a.z <= bwv) & __ballot(bz <= a.w) &
                       __ballot((rf | (bw >> 30)) == 3u) & vm;
    if (m) {
      if (__builtin_amdgcn_inverse_ballot_w64(m)) {  // (this lane's bit of m, as the exec mask)
        const uint32_t pos = stage_n + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        stage[pos] = make_uint2(aid, bw);
      }
      stage_n += (uint32_t)__popcll(m);
      if (stage_n > kStage - 64) flush_stage(stage, stage_n, out, cap, meta);
    }
  }
}

// One step of a diagonal item: the partner boxes moved one lane on.
__device__ __forceinline__ void diag_step(const float4& a, uint32_t aw, uint32_t rf, uint64_t vm, int addr, int& px,
                                          int& py, int& pz, int& pw, int& pid, uint2* stage, uint32_t& stage_n,
                                          uint2* __restrict__ out, uint64_t cap, Meta* __restrict__ meta) {
  px = __builtin_amdgcn_ds_bpermute(addr, px);
  py = __builtin_amdgcn_ds_bpermute(addr, py);
  pz = __builtin_amdgcn_ds_bpermute(addr, pz);
  pw = __builtin_amdgcn_ds_bpermute(addr, pw);
  pid = __builtin_amdgcn_ds_bpermute(addr, pid);
  const float bx = __builtin_bit_cast(float, px), by = __builtin_bit_cast(float, py);
  const float bz = __builtin_bit_cast(float, pz), bwv = __builtin_bit_cast(float, pw);
  const uint32_t bw = (uint32_t)pid;
  const uint64_t m = __ballot(a.x <= by) & __ballot(bx <= a.y) & __ballot(a.z <= bwv) & __ballot(bz <= a.w) &
                     __ballot((rf | (bw >> 30)) == 3u) & vm;
  if (m) {
    if (__builtin_amdgcn_inverse_ballot_w64(m)) {  // (this lane's bit of m, as the exec mask)
      const uint32_t pos = stage_n + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      stage[pos] = make_uint2(aw, bw);
    }
    stage_n += (uint32_t)__popcll(m);
    if (stage_n > kStage - 64) flush_stage(stage, stage_n, out, cap, meta);
  }
}

// A diagonal item (a tile against itself, w boxes on lanes 0 .. w-1): at step
// k = 1 .. w/2 lane l tests the box of lane (l + k) mod w, so every
// unordered pair is tested once (at k = w/2 of an even w only by the lanes
// below w/2) in half the steps of a column broadcast.  The partner boxes
// move one lane per step (each lane fetches lane (l + 1) mod w's partner
// across the wave), so a step needs no index arithmetic.
__device__ __forceinline__ void test_diagonal(const float4& a, uint32_t aw, uint64_t vm, uint32_t w,
                                              uint2* stage, uint32_t& stage_n, uint2* __restrict__ out,
                                              uint64_t cap, Meta* __restrict__ meta) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t rf = aw >> 30;
  const int addr = (int)((lane + 1u >= w ? lane + 1u - w : lane + 1u) << 2);
  int px = __builtin_bit_cast(int, a.x), py = __builtin_bit_cast(int, a.y);
  int pz = __builtin_bit_cast(int, a.z), pw = __builtin_bit_cast(int, a.w), pid = (int)aw;
  for (uint32_t k = 1; 2u * k < w; ++k)
    diag_step(a, aw, rf, vm, addr, px, py, pz, pw, pid, stage, stage_n, out, cap, meta);
  if (w >= 2u && (w & 1u) == 0u)  // k = w/2: the lanes below w/2 only
    diag_step(a, aw, rf, vm & ((1ull << (w >> 1)) - 1ull), addr, px, py, pz, pw, pid, stage, stage_n, out, cap,
              meta);
}

__global__ __launch_bounds__(256) void k_test(const uint4* __restrict__ items,
                                               const float4* __restrict__ mbox,
                                               const uint32_t* __restrict__ mid,
                                               uint2* __restrict__ out, uint64_t cap,
                                               Meta* __restrict__ meta) {
  __shared__ uint2 stage_all[4][kStage];
  __shared__ uint32_t tail_n[4];
  __shared__ uint64_t tail_base;
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  uint2* stage = stage_all[wib];
  uint32_t stage_n = 0;

  const uint64_t n_items = __builtin_amdgcn_readfirstlane((uint32_t)meta->items) |
                           ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(meta->items >> 32)) << 32);
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  const bool ok = meta->error == 0ull;

  // The next item's descriptor is loaded while the current one is tested.
  uint64_t it = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wib;
  uint4 d = make_uint4(0u, 0u, 0u, 0u);
  if (ok && it < n_items) d = items[it];
  for (; ok && it < n_items; it += nwaves) {
    const uint32_t base = __builtin_amdgcn_readfirstlane(d.x);
    const uint32_t n = __builtin_amdgcn_readfirstlane(d.y);
    const uint32_t rt = __builtin_amdgcn_readfirstlane(d.z);
    const uint32_t ct = __builtin_amdgcn_readfirstlane(d.w);
    if (it + nwaves < n_items) d = items[it + nwaves];

    const uint32_t i = rt * kTile + lane;
    const bool valid = i < n;
    float4 a = make_float4(0.f, -1.f, 0.f, -1.f);
    uint32_t aid = 0;
    if (valid) {
      a = mbox[base + i];
      aid = mid[base + i];
    }
    const uint64_t vm = __ballot(valid);
    const uint32_t j0 = ct * kTile;
    const uint32_t ncol = n - j0 < kTile ? n - j0 : kTile;
    if (rt == ct) {
      test_diagonal(a, aid, vm, ncol, stage, stage_n, out, cap, meta);
      continue;
    }
    // the column tile: lane l holds column j0 + l
    float4 cb = make_float4(0.f, -1.f, 0.f, -1.f);
    uint32_t cw = 0;
    if (lane < (int)ncol) {
      cb = mbox[base + j0 + lane];
      cw = mid[base + j0 + lane];
    }
    const uint32_t rf = aid >> 30;  // (start column, start row) flags of the row box
    test_columns(a, aid, rf, vm, cb, cw, ncol, stage, stage_n, out, cap, meta);
  }
  // what is left in the four stages goes out under one reservation
  if (lane == 0) tail_n[wib] = stage_n;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t t = (uint64_t)tail_n[0] + tail_n[1] + tail_n[2] + tail_n[3];
    tail_base = t ? atomicAdd((unsigned long long*)&meta->n_pairs, (unsigned long long)t) : 0ull;
  }
  __syncthreads();
  uint64_t g = tail_base;
  for (int k = 0; k < wib; ++k) g += tail_n[k];
  for (uint32_t k = lane; k < stage_n; k += 64) {
    const uint64_t dst = g + k;
    if (dst < cap) out[dst] = pair_out(stage[k]);
  }
}

}  // namespace pbig

// ---------------------------------------------------------------------------
// Host launcher of the plain path: six kernels on one stream, no host sync.
namespace pbig {

hipError_t launch_clear(const Workspace& ws, uint32_t n_cells, hipStream_t s) {
  const uint32_t nb_cells = (n_cells + 255) / 256;
  k_clear<<<nb_cells ? nb_cells : 1, 256, 0, s>>>(ws.cnt, n_cells, ws.meta);
  return hipGetLastError();
}

hipError_t launch_scan(const Workspace& ws, uint32_t n_cells, uint64_t memb_cap, hipStream_t s) {
  const uint32_t n_partials = (n_cells + kScanTile - 1) / kScanTile;
  k_scan_reduce<<<n_partials, 256, 0, s>>>(ws.cnt, n_cells, ws.partials);
  k_scan_apply<<<n_partials, 256, 0, s>>>(ws.cnt, n_cells, ws.partials, n_partials, ws.start, ws.items,
                                          ws.cap_items, memb_cap, ws.meta);
  return hipGetLastError();
}

// Every word of a Meta zeroed.
__global__ __launch_bounds__(256) void k_meta_zero(Meta* __restrict__ meta, uint32_t* __restrict__ words,
                                                    uint32_t n_words) {
  if (blockIdx.x == 0 && threadIdx.x < sizeof(Meta) / sizeof(unsigned long long))
    reinterpret_cast<unsigned long long*>(meta)[threadIdx.x] = 0ull;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += gridDim.x * blockDim.x) words[i] = 0u;
}

hipError_t launch_meta_clear(Meta* meta, hipStream_t s, uint32_t* words, uint32_t n_words) {
  const uint32_t blocks = words && n_words ? (n_words + 255u) / 256u : 1u;
  k_meta_zero<<<blocks < 64u ? blocks : 64u, 256, 0, s>>>(meta, words, words ? n_words : 0u);
  return hipGetLastError();
}

hipError_t launch_tile_test(const Workspace& ws, Meta* meta, uint2* out, uint64_t cap, hipStream_t s) {
  k_test<<<ws.test_blocks, 256, 0, s>>>(ws.items, ws.mbox, ws.mid, out, cap, meta);
  return hipGetLastError();
}

hipError_t launch_plain(const Workspace& ws, const float4* boxes, uint32_t n, GridParams gp,
                        uint32_t n_cells, uint2* out, uint64_t cap, hipStream_t s, hipEvent_t* ev) {
  const uint32_t nb_cells = (n_cells + 255) / 256;
  const uint32_t nb_boxes = (n + 255) / 256;
  const uint32_t n_partials = (n_cells + kScanTile - 1) / kScanTile;
  hipError_t e;
#define PBIG_MARK(k) \
  if (ev && (e = hipEventRecord(ev[k], s)) != hipSuccess) return e;
  PBIG_MARK(0)
  k_clear<<<nb_cells ? nb_cells : 1, 256, 0, s>>>(ws.cnt, n_cells, ws.meta);
  PBIG_MARK(1)
  if (n) k_count<<<nb_boxes, 256, 0, s>>>(boxes, n, gp, ws.cnt, ws.meta);
  PBIG_MARK(2)
  k_scan_reduce<<<n_partials, 256, 0, s>>>(ws.cnt, n_cells, ws.partials);
  PBIG_MARK(3)
  k_scan_apply<<<n_partials, 256, 0, s>>>(ws.cnt, n_cells, ws.partials, n_partials, ws.start,
                                          ws.items, ws.cap_items, ws.cap_memb, ws.meta);
  PBIG_MARK(4)
  if (n) k_scatter<<<nb_boxes, 256, 0, s>>>(boxes, n, gp, ws.cnt, ws.start, ws.mbox, ws.mid, ws.cap_memb);
  PBIG_MARK(5)
  k_test<<<ws.test_blocks, 256, 0, s>>>(ws.items, ws.mbox, ws.mid, out, cap, ws.meta);
  PBIG_MARK(6)
#undef PBIG_MARK
  return hipGetLastError();
}

}  // namespace pbig
