// Binned plain path for gfx950: a two-level grid whose fine cells are binned,
// tested and compacted inside LDS, one workgroup per coarse bin.
//
// The reference bins every box straight into its cells with one global atomic
// per membership (src/pbig/pbig.cu:174-221 count, 321-427 scatter).  On
// MI355X a device-wide atomic costs ~37 ps of whole-GPU throughput (about
// 27 G/s), so two atomics per membership would dominate the path.  Here
// global memory sees no per-membership atomics:
//
//   k_bin_hist     each workgroup owns a contiguous chunk of boxes and counts
//                  its coarse-bin memberships in an LDS histogram, written as
//                  one row of hist[workgroup][bin]
//   k_bin_scan     per bin, exclusive scan of the rows (workgroups ordered
//                  XCD-major, so runs written from one XCD sit side by side)
//   k_bin_scatter  re-reads its chunk, ranks memberships with LDS atomics and
//                  writes each coarse bin's boxes as contiguous runs
//   k_bin_test     one workgroup per coarse bin: fine-cell count, scan and
//                  scatter in LDS; all-pairs test of every fine cell with the
//                  reference-cell rule (pbig.cu:562-603); per-wave LDS staging
//                  of the pairs, one global atomic per flush of 128 pairs
//
// Cell maps.  The fine cell of a coordinate is clamp(floor(v * inv_f), 0,
// nf-1) with nf = nbx * K fine cells per row and inv_f rounded once on the
// host; the coarse bin is the fine index >> log2(K).  Both maps are monotone,
// so the reference-cell rule reduces to per-membership "starts in this
// column/row" flags exactly as in the single-level kernels.
//
// Per-cell test.  Cells here hold a handful of boxes, so a lane owns one
// membership r of a cell with n members and tests it against members
// (r + k) mod n for k = 1..h(r): h = n/2 (r < n/2) or n/2 - 1 (r >= n/2) for
// even n, (n-1)/2 for odd n.  Every unordered pair is tested exactly once;
// this is the reference's BALANCE schedule for the diagonal tile
// (pbig.cu:746-755, 802-809) applied to whole cells.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_binned.h"

namespace pbig {

__device__ __forceinline__ uint32_t fine_of(float v, float inv, uint32_t nf) {
  float f = floorf(v * inv);
  if (!(f >= 0.0f)) return 0u;
  if (f >= (float)(nf - 1u)) return nf - 1u;
  return (uint32_t)f;
}

// Coarse bin range of a box: fine index / K (K fine cells per bin side).
struct BinRange {
  uint32_t x0, x1, y0, y1;
};

// fine / k as a multiply by ceil(2^32 / k): exact for fine < 2^21, k < 2^11.
__device__ __forceinline__ uint32_t div_k(uint32_t fine, uint64_t magic) {
  return (uint32_t)(((uint64_t)fine * magic) >> 32);
}

__device__ __forceinline__ BinRange bin_range(const float4& b, const BinnedParams& p) {
  BinRange r;
  r.x0 = div_k(fine_of(b.x, p.inv_f, p.nf), p.k_magic);
  r.x1 = div_k(fine_of(b.y, p.inv_f, p.nf), p.k_magic);
  r.y0 = div_k(fine_of(b.z, p.inv_f, p.nf), p.k_magic);
  r.y1 = div_k(fine_of(b.w, p.inv_f, p.nf), p.k_magic);
  return r;
}

__global__ __launch_bounds__(1024) void k_bin_hist(const float4* __restrict__ boxes, uint32_t n,
                                                    BinnedParams p, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kMaxBins];
  __shared__ uint32_t err;
  for (uint32_t b = threadIdx.x; b < p.nb; b += blockDim.x) h[b] = 0u;
  if (threadIdx.x == 0) err = 0u;
  __syncthreads();
  const uint32_t start = blockIdx.x * p.chunk;
  const uint32_t end = start + p.chunk < n ? start + p.chunk : n;
  // Issue kBatch independent loads per thread before using any of them.
  for (uint32_t i0 = start; i0 < end; i0 += kBatch * blockDim.x) {
    float4 q[kBatch];
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint32_t i = i0 + threadIdx.x + k * blockDim.x;
      q[k] = i < end ? boxes[i] : make_float4(1.f, 0.f, 1.f, 0.f);
    }
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint32_t i = i0 + threadIdx.x + k * blockDim.x;
      if (i >= end) continue;
      const float4 b = q[k];
      if (!((b.x <= b.y) && (b.z <= b.w))) {
        err = 1u;
        continue;
      }
      const BinRange r = bin_range(b, p);
      for (uint32_t cy = r.y0; cy <= r.y1; ++cy)
        for (uint32_t cx = r.x0; cx <= r.x1; ++cx) atomicAdd(&h[cy * p.nbx + cx], 1u);
    }
  }
  __syncthreads();
  uint32_t* row = hist + (size_t)blockIdx.x * (p.nb + 1);
  for (uint32_t b = threadIdx.x; b < p.nb; b += blockDim.x) row[b] = h[b];
  if (threadIdx.x == 0) row[p.nb] = err;
}

// ---------------------------------------------------------------------------
// Workgroup w' in XCD-major order is hardware workgroup w = (w' % q) * 8 + w' / q
// with q = nw / 8 (workgroup w runs on XCD w % 8).
__device__ __forceinline__ uint32_t xcd_major(uint32_t wp, uint32_t q) { return (wp % q) * 8u + wp / q; }

__global__ __launch_bounds__(256) void k_bin_scan(const uint32_t* __restrict__ hist, BinnedParams p,
                                                   uint32_t* __restrict__ off, uint32_t* __restrict__ total,
                                                   Meta* __restrict__ meta) {
  // Thread t handles bin b0 + (t % 16) over row group t / 16 (rows in
  // XCD-major order).
  __shared__ uint32_t part[16][16];
  const uint32_t bl = threadIdx.x & 15, g = threadIdx.x >> 4;
  const uint32_t bin = blockIdx.x * 16 + bl;
  const uint32_t q = p.nw / 8u;
  const uint32_t per = (p.nw + 15) / 16;
  const uint32_t w0 = g * per, w1 = (w0 + per < p.nw) ? w0 + per : p.nw;
  uint32_t s = 0;
  if (bin < p.nb) {
#pragma unroll 8
    for (uint32_t wp = w0; wp < w1; ++wp) s += hist[(size_t)xcd_major(wp, q) * (p.nb + 1) + bin];
  }
  part[g][bl] = s;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t k = 0; k < g; ++k) base += part[k][bl];
  if (bin < p.nb) {
#pragma unroll 8
    for (uint32_t wp = w0; wp < w1; ++wp) {
      const size_t w = xcd_major(wp, q);
      const uint32_t v = hist[w * (p.nb + 1) + bin];
      off[w * p.nb + bin] = base;
      base += v;
    }
    if (g == 15) total[bin] = base;
  }
  if (blockIdx.x == 0) {
    // Result words of the call: pair counter and the boxes' validity.
    uint32_t e = 0;
    for (uint32_t w = threadIdx.x; w < p.nw; w += blockDim.x) e |= hist[(size_t)w * (p.nb + 1) + p.nb];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) e |= __shfl_xor(e, d, 64);
    __shared__ uint32_t es[4];
    if ((threadIdx.x & 63) == 0) es[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
      meta->n_pairs = 0ull;
      meta->error = (es[0] | es[1] | es[2] | es[3]) ? kErrInvalidBox : 0ull;
      meta->items = 0ull;
      meta->max_per_cell = 0u;
    }
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of total[0..nb) into s[] by one workgroup (blockDim a
// multiple of 64, at most 1024 threads); returns the grand total.
__device__ uint64_t block_exscan_bins(const uint32_t* __restrict__ total, uint32_t nb, uint32_t* s) {
  __shared__ uint64_t wsum[16];
  const uint32_t nt = blockDim.x, nwv = nt >> 6;
  const uint32_t per = (nb + nt - 1) / nt;
  const uint32_t b0 = threadIdx.x * per;
  uint64_t acc = 0;
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t b = b0 + k;
    if (b < nb) acc += total[b];
  }
  uint64_t v = acc;
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t o = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += o;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  uint64_t ex = v - acc, grand = 0;
  for (uint32_t k = 0; k < nwv; ++k) {
    if (k < w) ex += wsum[k];
    grand += wsum[k];
  }
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t b = b0 + k;
    if (b < nb) {
      s[b] = (uint32_t)ex;
      ex += total[b];
    }
  }
  __syncthreads();
  return grand;
}

__global__ __launch_bounds__(1024) void k_bin_scatter(const float4* __restrict__ boxes, uint32_t n, BinnedParams p,
                                                       const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ total,
                                                       uint32_t* __restrict__ bin_start,
                                                       float4* __restrict__ bbox, uint32_t* __restrict__ bid,
                                                       uint64_t cap, Meta* __restrict__ meta) {
  __shared__ uint32_t cur[kMaxBins];
  const uint64_t grand = block_exscan_bins(total, p.nb, cur);
  if (blockIdx.x == 0) {
    for (uint32_t b = threadIdx.x; b < p.nb; b += blockDim.x) bin_start[b] = cur[b];
    if (threadIdx.x == 0) {
      meta->memberships = grand;
      if (grand > cap) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrWorkspace);
    }
  }
  if (grand > cap) return;
  const uint32_t* myoff = off + (size_t)blockIdx.x * p.nb;
  for (uint32_t b = threadIdx.x; b < p.nb; b += blockDim.x) cur[b] += myoff[b];
  __syncthreads();
  const uint32_t start = blockIdx.x * p.chunk;
  const uint32_t end = start + p.chunk < n ? start + p.chunk : n;
  for (uint32_t i0 = start; i0 < end; i0 += kBatch * blockDim.x) {
    float4 q[kBatch];
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint32_t i = i0 + threadIdx.x + k * blockDim.x;
      q[k] = i < end ? boxes[i] : make_float4(1.f, 0.f, 1.f, 0.f);
    }
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint32_t i = i0 + threadIdx.x + k * blockDim.x;
      const float4 b = q[k];
      if (i >= end || !((b.x <= b.y) && (b.z <= b.w))) continue;
      const BinRange r = bin_range(b, p);
      for (uint32_t cy = r.y0; cy <= r.y1; ++cy)
        for (uint32_t cx = r.x0; cx <= r.x1; ++cx) {
          const uint32_t pos = atomicAdd(&cur[cy * p.nbx + cx], 1u);
          bbox[pos] = b;
          bid[pos] = i;
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Fine binning + per-cell test of coarse bins in LDS.
//
// One persistent 1024-thread workgroup per CU walks bins b = blockIdx.x,
// + gridDim.x, ...  A bin's coarse memberships sit in registers (kRegs per
// thread); its fine cells (k x k) are counted and scanned in LDS and then
// processed in strips: row-major ranges of fine cells whose memberships fit
// the LDS buffer.  For each strip the memberships are scattered into
// fine-cell order in LDS and every cell's pairs are tested.  Pairs go to a
// per-wave LDS stage and then to this workgroup's own region of the pair
// scratch through an LDS counter: no global atomics.  k_bin_compact later
// concatenates the regions.  The next bin's registers are loaded after the
// last strip's scatter and land while that strip is tested.

constexpr uint32_t kTestThreads = 1024;
constexpr uint32_t kUnroll = 2;  // partner loads issued together in the test loop
constexpr uint32_t kTestWaves = kTestThreads / 64;
constexpr uint32_t kRegs = kMaxCoarse / kTestThreads;  // coarse memberships per thread in registers

// Position word: coarse-membership index (13 bits) | local fine cell (12 bits)
// | "starts in this row" (bit 30) | "starts in this column" (bit 31).
constexpr uint32_t kWordCellShift = 13;
constexpr uint32_t kWordIdxMask = (1u << kWordCellShift) - 1u;
constexpr uint32_t kWordCellMask = (1u << 12) - 1u;
static_assert(kMaxCoarse <= (1u << kWordCellShift), "membership index field");
static_assert(kMaxCellsPerBin <= (1u << 12), "cell field");

struct TestLds {
  float4 box[kMaxFine];            // strip memberships in fine-cell order
  uint32_t word[kMaxFine];         // position words
  uint32_t id[kMaxCoarse];         // box id of each coarse membership
  uint32_t cnt[kMaxCellsPerBin];   // fine-cell counts of the bin
  uint32_t start[kMaxCellsPerBin + 1];  // exclusive scan of cnt (start[kk] = total)
  uint2 stage[kTestWaves][kBinStage];   // per-wave pair staging
  uint2 binfo[kMaxBinsPerGroup];   // (start, count) of this workgroup's bins
  uint32_t wsum[kTestWaves];
  uint32_t out_n;                  // pairs appended to this workgroup's region
  uint32_t strip_hi;
};

__device__ __forceinline__ void stage_flush(TestLds& L, uint32_t w, uint32_t& sn, uint2* __restrict__ region,
                                            uint32_t cap_wg) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t g = 0;
  if (lane == 0) g = atomicAdd(&L.out_n, sn);
  g = __shfl(g, 0, 64);
  for (uint32_t k = lane; k < sn; k += 64)
    if (g + k < cap_wg) region[g + k] = L.stage[w][k];
  sn = 0;
}

// Fine-cell range of a box clipped to the bin whose first fine cell is (gx, gy).
struct FineRange {
  uint32_t x0, x1, y0, y1, fx0, fy0;
};

__device__ __forceinline__ FineRange fine_range(const float4& q, const BinnedParams& p, uint32_t gx, uint32_t gy) {
  FineRange r;
  r.fx0 = fine_of(q.x, p.inv_f, p.nf);
  r.fy0 = fine_of(q.z, p.inv_f, p.nf);
  const uint32_t fx1 = fine_of(q.y, p.inv_f, p.nf), fy1 = fine_of(q.w, p.inv_f, p.nf);
  r.x0 = r.fx0 > gx ? r.fx0 - gx : 0u;
  r.y0 = r.fy0 > gy ? r.fy0 - gy : 0u;
  r.x1 = fx1 - gx < p.k ? fx1 - gx : p.k - 1;
  r.y1 = fy1 - gy < p.k ? fy1 - gy : p.k - 1;
  return r;
}

// Scatter the fine memberships of one coarse membership that fall in cells
// [c_lo, c_hi) to their positions in the strip (start[] is the cursor).
__device__ __forceinline__ void scatter_strip(TestLds& L, const float4& q, uint32_t j, const BinnedParams& p,
                                              uint32_t gx, uint32_t gy, uint32_t c_lo, uint32_t c_hi,
                                              uint32_t base) {
  const FineRange r = fine_range(q, p, gx, gy);
  const uint32_t row_lo = c_lo / p.k, row_hi = (c_hi - 1) / p.k;
  const uint32_t y0 = r.y0 > row_lo ? r.y0 : row_lo, y1 = r.y1 < row_hi ? r.y1 : row_hi;
  for (uint32_t cy = y0; cy <= y1; ++cy) {
    const uint32_t fy = (gy + cy == r.fy0) ? kFlagY : 0u;
    for (uint32_t cx = r.x0; cx <= r.x1; ++cx) {
      const uint32_t c = cy * p.k + cx;
      if (c < c_lo || c >= c_hi) continue;
      const uint32_t pos = atomicAdd(&L.start[c], 1u) - base;
      L.box[pos] = q;
      L.word[pos] = j | (c << kWordCellShift) | fy | ((gx + cx == r.fx0) ? kFlagX : 0u);
    }
  }
}

__global__ __launch_bounds__(kTestThreads) void k_bin_test(const float4* __restrict__ bbox,
                                                           const uint32_t* __restrict__ bid,
                                                           const uint32_t* __restrict__ bin_start,
                                                           const uint32_t* __restrict__ total, BinnedParams p,
                                                           uint64_t cap_bucket, uint2* __restrict__ scratch,
                                                           uint32_t cap_wg, uint32_t* __restrict__ wg_count,
                                                           uint32_t* __restrict__ overfull,
                                                           Meta* __restrict__ meta) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  TestLds& L = *reinterpret_cast<TestLds*>(lds_raw);
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (meta->error & (kErrInvalidBox | kErrWorkspace)) {
    if (tid == 0) wg_count[blockIdx.x] = 0u;
    return;
  }
  const uint32_t K = p.k, kk = K * K;
  uint2* region = scratch + (size_t)blockIdx.x * cap_wg;
  uint32_t sn = 0;

  const uint32_t my_bins = blockIdx.x < p.nb ? (p.nb - blockIdx.x + gridDim.x - 1) / gridDim.x : 0u;
  for (uint32_t t = tid; t < my_bins; t += kTestThreads) {
    const uint32_t bb = blockIdx.x + t * gridDim.x;
    L.binfo[t] = make_uint2(bin_start[bb], total[bb]);
  }
  if (tid == 0) L.out_n = 0u;
  __syncthreads();

  // Registers: unconditional loads with the index clamped into the bucket.
  const uint64_t last = cap_bucket - 1;
  float4 rb[kRegs];
  uint32_t ri[kRegs];
  if (my_bins) {
    const uint64_t s0 = L.binfo[0].x;
#pragma unroll
    for (uint32_t k = 0; k < kRegs; ++k) {
      const uint64_t at = s0 + tid + k * kTestThreads < last ? s0 + tid + k * kTestThreads : last;
      rb[k] = bbox[at];
      ri[k] = bid[at];
    }
  }
  for (uint32_t t = 0; t < my_bins; ++t) {
    const uint32_t b = blockIdx.x + t * gridDim.x;
    const uint32_t m = L.binfo[t].y;
    const uint64_t sN = t + 1 < my_bins ? L.binfo[t + 1].x : 0ull;
    const uint32_t gx = (b % p.nbx) * K, gy = (b / p.nbx) * K;
    bool prefetched = false;
    if (m >= 2 && m <= kMaxCoarse) {
      // A. fine-cell counts of the whole bin; ids to LDS
      for (uint32_t c = tid; c < kk; c += kTestThreads) L.cnt[c] = 0u;
      __syncthreads();
#pragma unroll
      for (uint32_t k = 0; k < kRegs; ++k) {
        const uint32_t j = tid + k * kTestThreads;
        if (j < m) {
          L.id[j] = ri[k];
          const FineRange r = fine_range(rb[k], p, gx, gy);
          for (uint32_t cy = r.y0; cy <= r.y1; ++cy)
            for (uint32_t cx = r.x0; cx <= r.x1; ++cx) atomicAdd(&L.cnt[cy * K + cx], 1u);
        }
      }
      __syncthreads();
      // B. exclusive scan of the counts (kk <= 4096: 4 per thread)
      {
        uint32_t v[4], acc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t c = tid * 4 + k;
          v[k] = c < kk ? L.cnt[c] : 0u;
          acc += v[k];
        }
        uint32_t inc = acc;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const uint32_t o = __shfl_up(inc, d, 64);
          if (lane >= (uint32_t)d) inc += o;
        }
        if (lane == 63) L.wsum[w] = inc;
        __syncthreads();
        uint32_t ex = inc - acc;
        for (uint32_t k = 0; k < w; ++k) ex += L.wsum[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t c = tid * 4 + k;
          if (c < kk) L.start[c] = ex;
          ex += v[k];
        }
        if (tid == kTestThreads - 1) L.start[kk] = ex;
      }
      __syncthreads();
      // C/D per strip
      uint32_t c_lo = 0;
      while (c_lo < kk) {
        if (tid == 0) {
          // largest c_hi with start[c_hi] - start[c_lo] <= kMaxFine
          const uint32_t lim = L.start[c_lo] + kMaxFine;
          uint32_t lo = c_lo, hi = kk;  // start[lo] <= lim; find max index with start <= lim
          while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (L.start[mid] <= lim) lo = mid; else hi = mid - 1;
          }
          L.strip_hi = lo;
        }
        __syncthreads();
        const uint32_t c_hi = L.strip_hi;
        if (c_hi == c_lo) {  // one fine cell alone exceeds the LDS buffer
          if (tid == 0) {
            overfull[atomicAdd(&meta->max_per_cell, 1u)] = b;
            atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrOverfull);
          }
          break;
        }
        const uint32_t base = L.start[c_lo];
        const uint32_t len = L.start[c_hi] - base;
        __syncthreads();  // every thread has read start[] before it becomes a cursor
#pragma unroll
        for (uint32_t k = 0; k < kRegs; ++k) {
          const uint32_t j = tid + k * kTestThreads;
          if (j < m) scatter_strip(L, rb[k], j, p, gx, gy, c_lo, c_hi, base);
        }
        __syncthreads();
        // restore start[] for the strip's cells (the cursors advanced by cnt)
        for (uint32_t c = c_lo + tid; c < c_hi; c += kTestThreads) L.start[c] -= L.cnt[c];
        const bool last_strip = c_hi == kk;
        if (last_strip && t + 1 < my_bins) {
#pragma unroll
          for (uint32_t k = 0; k < kRegs; ++k) {
            const uint64_t at = sN + tid + k * kTestThreads < last ? sN + tid + k * kTestThreads : last;
            rb[k] = bbox[at];
            ri[k] = bid[at];
          }
          prefetched = true;
        }
        __syncthreads();
        // D. test every pair of each cell of the strip, one position per lane
        for (uint32_t p0 = w * 64; p0 < len; p0 += kTestThreads) {
          const uint32_t pp = p0 + lane;
          uint32_t h = 0, cs = 0, cn = 1, aw = 0;
          float4 a = make_float4(0.f, -1.f, 0.f, -1.f);
          if (pp < len) {
            aw = L.word[pp];
            const uint32_t c = (aw >> kWordCellShift) & kWordCellMask;
            cs = L.start[c] - base;
            cn = L.cnt[c];
            const uint32_t r = pp - cs;
            h = (cn & 1u) ? (cn - 1) >> 1 : ((r < (cn >> 1)) ? (cn >> 1) : (cn >> 1) - 1);
            a = L.box[pp];
          }
          const bool afx = (int32_t)aw < 0, afy = (aw & kFlagY) != 0;
          const uint32_t ja = aw & kWordIdxMask;
          const uint32_t wrap = cs + cn;
          // partners (r + k) mod n for k = 1..h, four at a time; loads first
          for (uint32_t k = 1; __ballot(k <= h); k += kUnroll) {
            float4 o[kUnroll];
            uint32_t ow[kUnroll];
            bool act[kUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
              act[u] = k + u <= h;
              uint32_t qq = pp + k + u;
              qq = qq >= wrap ? qq - cn : qq;
              qq = act[u] ? qq : 0u;
              o[u] = L.box[qq];
              ow[u] = L.word[qq];
            }
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; ++u) {
              const bool hit = act[u] & (a.x <= o[u].y) & (o[u].x <= a.y) & (a.z <= o[u].w) & (o[u].z <= a.w) &
                               (afx | ((int32_t)ow[u] < 0)) & (afy | ((ow[u] & kFlagY) != 0));
              const uint64_t mask = __ballot(hit);
              if (mask) {
                const uint32_t pos = sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                if (hit) {
                  const uint32_t ia = L.id[ja], ib = L.id[ow[u] & kWordIdxMask];
                  L.stage[w][pos] = make_uint2(ia < ib ? ia : ib, ia < ib ? ib : ia);
                }
                sn += (uint32_t)__popcll(mask);
                if (sn > kBinStage - 64) stage_flush(L, w, sn, region, cap_wg);
              }
            }
          }
        }
        __syncthreads();  // the strip buffer is reused
        c_lo = c_hi;
      }
    } else if (m > kMaxCoarse) {
      if (tid == 0) {
        overfull[atomicAdd(&meta->max_per_cell, 1u)] = b;
        atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrOverfull);
      }
    }
    if (!prefetched && t + 1 < my_bins) {
      __syncthreads();
#pragma unroll
      for (uint32_t k = 0; k < kRegs; ++k) {
        const uint64_t at = sN + tid + k * kTestThreads < last ? sN + tid + k * kTestThreads : last;
        rb[k] = bbox[at];
        ri[k] = bid[at];
      }
    }
  }
  if (sn) stage_flush(L, w, sn, region, cap_wg);
  __syncthreads();
  if (tid == 0) {
    wg_count[blockIdx.x] = L.out_n;
    if (L.out_n > cap_wg) {
      atomicMax((unsigned long long*)&meta->items, (unsigned long long)L.out_n);
      atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrWorkspace);
    }
  }
}

// Concatenate the per-workgroup pair regions into the caller's buffer.
// Block (v, part) copies the part-th slice of region v.
constexpr uint32_t kCompactParts = 8;

__global__ __launch_bounds__(256) void k_bin_compact(const uint2* __restrict__ scratch, uint32_t cap_wg,
                                                      const uint32_t* __restrict__ wg_count, uint32_t nwg,
                                                      uint2* __restrict__ out, uint64_t cap,
                                                      Meta* __restrict__ meta) {
  __shared__ uint64_t part[4];
  const uint32_t v = blockIdx.x / kCompactParts, slice = blockIdx.x % kCompactParts;
  uint64_t before = 0;
  for (uint32_t u = threadIdx.x; u < v; u += 256) before += wg_count[u] < cap_wg ? wg_count[u] : cap_wg;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) before += __shfl_xor(before, d, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = before;
  __syncthreads();
  const uint64_t off = part[0] + part[1] + part[2] + part[3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the true count, also when a region overflowed (the host then retries)
    uint64_t t = 0;
    for (uint32_t u = 0; u < nwg; ++u) t += wg_count[u];
    meta->n_pairs = t;
  }
  const uint32_t mine = wg_count[v] < cap_wg ? wg_count[v] : cap_wg;
  const uint32_t per = (mine + kCompactParts - 1) / kCompactParts;
  const uint32_t k0 = slice * per, k1 = k0 + per < mine ? k0 + per : mine;
  const uint2* src = scratch + (size_t)v * cap_wg;
#pragma unroll 4
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const uint64_t dst = off + k;
    if (dst < cap) out[dst] = src[k];
  }
}

// ---------------------------------------------------------------------------
size_t binned_test_lds_bytes() { return sizeof(TestLds); }

hipError_t binned_init() {
  return hipFuncSetAttribute((const void*)k_bin_test, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)sizeof(TestLds));
}

hipError_t launch_binned(const BinnedWorkspace& ws, const float4* boxes, uint32_t n, const BinnedParams& p,
                         uint2* out, uint64_t cap, hipStream_t s, hipEvent_t* ev) {
  hipError_t e;
#define PBIG_MARK(k) \
  if (ev && (e = hipEventRecord(ev[k], s)) != hipSuccess) return e;
  PBIG_MARK(0)
  k_bin_hist<<<p.nw, 1024, 0, s>>>(boxes, n, p, ws.hist);
  PBIG_MARK(1)
  k_bin_scan<<<(p.nb + 15) / 16, 256, 0, s>>>(ws.hist, p, ws.off, ws.total, ws.meta);
  PBIG_MARK(2)
  k_bin_scatter<<<p.nw, 1024, 0, s>>>(boxes, n, p, ws.off, ws.total, ws.bin_start, ws.bbox, ws.bid, ws.cap_bucket,
                                      ws.meta);
  PBIG_MARK(3)
  k_bin_test<<<p.test_blocks, kTestThreads, sizeof(TestLds), s>>>(ws.bbox, ws.bid, ws.bin_start, ws.total, p,
                                                                  ws.cap_bucket, ws.scratch, ws.cap_wg,
                                                                  ws.wg_count, ws.overfull, ws.meta);
  PBIG_MARK(4)
  k_bin_compact<<<p.test_blocks * kCompactParts, 256, 0, s>>>(ws.scratch, ws.cap_wg, ws.wg_count, p.test_blocks,
                                                              out, cap, ws.meta);
  PBIG_MARK(5)
#undef PBIG_MARK
  return hipGetLastError();
}

}  // namespace pbig
