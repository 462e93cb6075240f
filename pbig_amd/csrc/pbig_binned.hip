// Binned path for gfx950: a two-level grid whose fine cells are binned,
// tested and compacted inside LDS.  One launch sequence serves a batch of
// independent instances (a single instance is a batch of one): every
// histogram/scatter chunk and every coarse bin belongs to one instance, and
// the kernels read that instance's grid from a small device table.
//
// The reference bins every box straight into its cells with one global atomic
// per membership (src/pbig/pbig.cu:174-221 count, 321-427 scatter).  On
// MI355X a device-wide atomic costs ~37 ps of whole-GPU throughput (about
// 27 G/s) and same-address atomics serialise at ~12 ns, so here global memory
// sees no per-membership atomics:
//
//   k_bin_hist     each workgroup owns a chunk of one instance's boxes and
//                  counts its coarse-bin memberships in an LDS histogram,
//                  written as one row of that instance's hist[chunk][bin]
//   k_bin_scan     per bin, exclusive scan over the instance's chunks (the
//                  table read once)
//   k_bin_starts   exclusive scan of all bins' totals (bucket offsets)
//   k_bin_scatter  the chunk is held in registers; every membership is
//                  ranked among the chunk's memberships in bin order and
//                  staged in LDS, so stores cover whole runs; each XCD walks
//                  one contiguous range of chunks, so neighbouring runs merge
//                  in its L2
//   k_bin_test     persistent workgroups over all bins, one of two
//                  schedules picked by the histogram's wide flag: when every
//                  box spans at most two fine cells per axis, a half-stencil
//                  over neighbouring min-corner cells; otherwise fine-cell
//                  count, layout and scatter in LDS (in strips that fit) and
//                  an all-pairs test of every fine cell with the
//                  reference-cell rule (pbig.cu:562-603).  Pairs are appended
//                  through an LDS counter to the workgroup's own region,
//                  which a single instance's workgroup then copies to its
//                  share of the output
//   k_bin_pair_scan, k_bin_compact   (batches) every bin's output position,
//                  then the regions' runs placed in bin order, so that each
//                  instance's pairs are contiguous

// Cell maps.  The fine cell of a coordinate is clamp(floor(v * inv_f), 0,
// nf-1) with nf = nbx * k fine cells per row and inv_f rounded once on the
// host; the coarse bin is the fine index / k.  Both maps are monotone, so the
// reference-cell rule reduces to per-membership "starts in this column/row"
// flags.
//
// Per-cell test.  A lane owns one membership r of a cell with n members and
// tests it against members (r + k) mod n for k = 1..h(r): h = n/2 (r < n/2)
// or n/2 - 1 (r >= n/2) for even n, (n-1)/2 for odd n.  Every unordered pair
// is tested exactly once; this is the reference's BALANCE schedule for the
// diagonal tile (pbig.cu:746-755, 802-809) applied to whole cells.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbig_binned.h"

namespace pbig {

// Branch-free: fmaxf returns its other operand for a NaN, so a NaN lands in
// cell 0; the clamped float is an integer in [0, nf-1] (nf < 2^24) and
// converts exactly.
__device__ __forceinline__ uint32_t fine_of(float v, float inv, uint32_t nf) {
  return (uint32_t)fminf(fmaxf(floorf(v * inv), 0.0f), (float)(nf - 1u));
}

// Inclusive prefix sums across lanes by DPP lane moves (GFX9 family): row
// shifts by 1, 2, 4 and 8 within each 16-lane row; then, for the whole
// wave, lane 15 added into rows 1 and 3 and lane 31 into rows 2 and 3.  A
// lane with no source lane adds the `old` operand, zero.
__device__ __forceinline__ uint32_t row_prefix_sum(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  return v;
}

__device__ __forceinline__ uint32_t wave_prefix_sum(uint32_t v) {
  v = row_prefix_sum(v);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// The same for 64-bit values: each lane move takes both halves from the same
// source lane (or zero for both), and the add carries.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, kCtrl, kRowMask, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), kCtrl, kRowMask, 0xf, false);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_prefix_sum64(uint64_t v) {
  v += dpp_u64<0x111, 0xf>(v);  // row_shr:1
  v += dpp_u64<0x112, 0xf>(v);  // row_shr:2
  v += dpp_u64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_u64<0x118, 0xf>(v);  // row_shr:8
  v += dpp_u64<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_u64<0x143, 0xc>(v);  // row_bcast:31
  return v;
}

// Coarse bin range of a box: fine index / K (K fine cells per bin side).
struct BinRange {
  uint32_t x0, x1, y0, y1;
};

// fine / k as a multiply by magic = ceil(2^32 / k): exact for fine < 2^21,
// k < 2^11.  The magic fits 32 bits for every k >= 2; k = 1 (magic = 2^32)
// is the identity.
__device__ __forceinline__ uint32_t div_k(uint32_t fine, uint64_t magic) {
  const uint32_t q = __umulhi(fine, (uint32_t)magic);
  return (magic >> 32) ? fine : q;
}

// Fine index relative to the instance's region, clamped into it (a no-op
// for an owned instance: fx0 = 0, span = nf).  Bins wholly outside a
// non-square region receive no memberships.
__device__ __forceinline__ uint32_t region_of(uint32_t f, uint32_t f0, uint32_t span) {
  return min(max(f, f0), f0 + span - 1u) - f0;
}

// Coarse bin range from the fine cells of a box's corners (f: x0, x1, y0, y1).
__device__ __forceinline__ BinRange bin_range_of(const uint4& f, const InstDesc& p) {
  BinRange r;
  r.x0 = div_k(region_of(f.x, p.fx0, p.span_x), p.k_magic);
  r.x1 = div_k(region_of(f.y, p.fx0, p.span_x), p.k_magic);
  r.y0 = div_k(region_of(f.z, p.fy0, p.span_y), p.k_magic);
  r.y1 = div_k(region_of(f.w, p.fy0, p.span_y), p.k_magic);
  return r;
}

__device__ __forceinline__ uint4 fine_corners(const float4& b, const InstDesc& p) {
  return make_uint4(fine_of(b.x, p.inv_f, p.nf), fine_of(b.y, p.inv_f, p.nf), fine_of(b.z, p.inv_f, p.nf),
                    fine_of(b.w, p.inv_f, p.nf));
}

// The region clamp folded into the float clamp: the region [f0, f0+span-1]
// lies inside [0, nf-1], so clamping floor(v * inv) straight into it gives
// region_of(fine_of(v)) for every v (a NaN lands on f0, as fine_of's 0 does).
__device__ __forceinline__ uint32_t region_bin(float v, float inv, uint32_t f0, uint32_t span, uint64_t magic) {
  const float c = fminf(fmaxf(floorf(v * inv), (float)f0), (float)(f0 + span - 1u));
  return div_k((uint32_t)c - f0, magic);
}

__device__ __forceinline__ BinRange bin_range(const float4& b, const InstDesc& p) {
  BinRange r;
  r.x0 = region_bin(b.x, p.inv_f, p.fx0, p.span_x, p.k_magic);
  r.x1 = region_bin(b.y, p.inv_f, p.fx0, p.span_x, p.k_magic);
  r.y0 = region_bin(b.z, p.inv_f, p.fy0, p.span_y, p.k_magic);
  r.y1 = region_bin(b.w, p.inv_f, p.fy0, p.span_y, p.k_magic);
  return r;
}

// Box range [start, end) of global chunk w.
__device__ __forceinline__ void chunk_range(const InstDesc& I, uint32_t w, uint64_t* start, uint64_t* end) {
  const uint64_t lc = w - I.chunk0;
  const uint64_t inst_end = I.box0 + I.n;
  *start = I.box0 + lc * I.chunk;
  *end = *start + I.chunk < inst_end ? *start + I.chunk : inst_end;
}

// Test schedule of a launch, the same in every kernel that depends on it:
// the narrow one (positions keyed by their min-corner cell; the few wide
// boxes tested against the cells under them) unless more than one box in
// kWideShare spans more than two fine cells on an axis, and then the general
// one (fine replication).  The narrow schedule tests each box against about
// four and a half cells' boxes and the general one each pair of a cell once,
// so the narrow one wins on sparse cells only; wide boxes come with varied
// sizes and dense cells (measured on the clustered 2M instance: the general
// schedule ahead at every wide share above one in a thousand), so a stray
// wide box keeps the narrow schedule and more send the launch to the general
// one.
constexpr uint32_t kWideShare = 1024;
__device__ __forceinline__ bool general_schedule(const Meta* m) {
  return (uint64_t)m->wide * kWideShare > (uint64_t)m->boxes;
}

// kDev (all kernels below): a level planned on the device; its chunk and bin
// counts are read from the level's result words (meta->dev_*), and the grid
// is an upper bound.
template <bool kDev>
__global__ __launch_bounds__(kHistThreads) void k_bin_hist(const float4* __restrict__ boxes,
                                                    const InstDesc* __restrict__ inst,
                                                    const uint32_t* __restrict__ chunk_inst, uint32_t n_chunks,
                                                    uint32_t* __restrict__ hist, const Meta* __restrict__ meta) {
  __shared__ uint32_t h[kMaxBins];
  __shared__ uint32_t err;  // bit 0: an invalid box; bits 2..: the chunk's wide boxes
  if (kDev) n_chunks = meta->plan_err ? 0u : meta->dev_chunks;
  if (blockIdx.x >= n_chunks) return;
  const InstDesc I = inst[chunk_inst[blockIdx.x]];
  for (uint32_t b = threadIdx.x; b < I.nb; b += blockDim.x) h[b] = 0u;
  if (threadIdx.x == 0) err = 0u;
  __syncthreads();
  uint64_t start, end;
  chunk_range(I, blockIdx.x, &start, &end);
  // Issue kBatch independent loads per thread before using any of them.
  for (uint64_t i0 = start; i0 < end; i0 += kBatch * blockDim.x) {
    float4 q[kBatch];
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint64_t i = i0 + threadIdx.x + k * blockDim.x;
      q[k] = i < end ? boxes[i] : make_float4(1.f, 0.f, 1.f, 0.f);
    }
#pragma unroll
    for (uint32_t k = 0; k < kBatch; ++k) {
      const uint64_t i = i0 + threadIdx.x + k * blockDim.x;
      if (i >= end) continue;
      const float4 b = q[k];
      if (!((b.x <= b.y) && (b.z <= b.w))) {
        atomicOr(&err, 1u);
        continue;
      }
      const uint4 f = fine_corners(b, I);
      const BinRange r = bin_range_of(f, I);
      if (f.y - f.x > 1u || f.w - f.z > 1u) {  // spans more than two fine cells
        atomicAdd(&err, 4u);
        for (uint32_t cy = r.y0; cy <= r.y1; ++cy)
          for (uint32_t cx = r.x0; cx <= r.x1; ++cx) atomicAdd(&h[cy * I.nbx + cx], 1u);
      } else {  // at most two bins per axis: the first, then east, north, north-east
        const uint32_t bb = r.y0 * I.nbx + r.x0;
        const bool ex = r.x1 != r.x0, ey = r.y1 != r.y0;
        atomicAdd(&h[bb], 1u);
        if (ex) atomicAdd(&h[bb + 1u], 1u);
        if (ey) atomicAdd(&h[bb + I.nbx], 1u);
        if (ex && ey) atomicAdd(&h[bb + I.nbx + 1u], 1u);
      }
    }
  }
  __syncthreads();
  uint32_t* row = hist + I.hist0 + (uint64_t)(blockIdx.x - I.chunk0) * hist_pitch(I.nb);
  for (uint32_t b = threadIdx.x; b < I.nb; b += blockDim.x) row[b] = h[b];
  if (threadIdx.x == 0) row[I.nb] = err;
}

// ---------------------------------------------------------------------------
// Per bin, exclusive scan of its histogram column over the instance's
// chunks.  A block takes 16 bins (half a 128-byte line of a row; the other
// half is the next block's, an L2 hit) and 64 groups of chunks; each thread
// holds the 16 counts of its chunks in registers, so the table is read
// once, and the offsets are written from the registers after a prefix over
// the groups (one wave per bin, a lane per group).  Instances of more than
// 1,024 chunks take several tiles, the running total carried.
constexpr uint32_t kScanBins = 16, kScanGroups = 64, kScanPer = 16;
static_assert(kScanBins * kScanGroups == 1024 && kScanGroups == 64, "the group prefix is one wave's scan per bin");

template <bool kDev>
__global__ __launch_bounds__(1024) void k_bin_scan(const uint32_t* __restrict__ hist,
                                                    const InstDesc* __restrict__ inst,
                                                    const uint32_t* __restrict__ bin_inst, uint32_t n_bins,
                                                    const uint32_t* __restrict__ chunk_inst, uint32_t n_chunks,
                                                    uint32_t* __restrict__ off, uint32_t* __restrict__ total,
                                                    Meta* __restrict__ meta) {
  __shared__ uint32_t part[kScanGroups][kScanBins + 1];
  __shared__ uint32_t carry[kScanBins];
  __shared__ uint32_t tiles;
  if (kDev) {
    n_bins = meta->plan_err ? 0u : meta->dev_bins;
    n_chunks = meta->plan_err ? 0u : meta->dev_chunks;
  }
  if (blockIdx.x == gridDim.x - 1) {
    // The last block: the result words of the call, the pair counter, the
    // boxes' validity, and the launch's box and wide-box counts, from every
    // chunk's histogram row.
    uint32_t e = 0;
    uint64_t nw = 0, nbox = 0;
    for (uint32_t w = threadIdx.x; w < n_chunks; w += blockDim.x) {
      const InstDesc I = inst[chunk_inst[w]];
      const uint32_t x = hist[I.hist0 + (uint64_t)(w - I.chunk0) * hist_pitch(I.nb) + I.nb];
      e |= x & 1u;
      nw += x >> 2;
      uint64_t c0, c1;
      chunk_range(I, w, &c0, &c1);
      nbox += c1 - c0;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      e |= __shfl_xor(e, d, 64);
      nw += __shfl_xor(nw, d, 64);
      nbox += __shfl_xor(nbox, d, 64);
    }
    __shared__ uint32_t es[16];
    __shared__ uint64_t ws[16], bs[16];
    if ((threadIdx.x & 63) == 0) {
      es[threadIdx.x >> 6] = e;
      ws[threadIdx.x >> 6] = nw;
      bs[threadIdx.x >> 6] = nbox;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t any = 0;
      uint64_t tw = 0, tb = 0;
      for (int k = 0; k < 16; ++k) {
        any |= es[k];
        tw += ws[k];
        tb += bs[k];
      }
      meta->n_pairs = 0ull;
      meta->error = (any & 1u) ? kErrInvalidBox : 0ull;
      meta->items = 0ull;
      meta->max_per_cell = 0u;
      meta->wide = tw < 0xFFFFFFFFull ? (uint32_t)tw : 0xFFFFFFFFu;
      meta->boxes = tb < 0xFFFFFFFFull ? (uint32_t)tb : 0xFFFFFFFFu;
    }
    return;
  }
  const uint32_t bl = threadIdx.x % kScanBins, g = threadIdx.x / kScanBins;
  const uint32_t gb = blockIdx.x * kScanBins + bl;
  uint32_t b = 0, nch = 0, hp = 0, op = 0;
  uint64_t hist0 = 0, off0 = 0;
  if (threadIdx.x == 0) tiles = 0u;
  if (g == 0) carry[bl] = 0u;
  if (gb < n_bins) {
    const InstDesc I = inst[bin_inst[gb]];
    b = gb - I.bin0;
    nch = I.nchunks;
    hp = hist_pitch(I.nb);
    op = off_pitch(I.nb);
    hist0 = I.hist0;
    off0 = I.off0;
  }
  __syncthreads();
  if (g == 0 && nch) atomicMax(&tiles, (nch + kScanGroups * kScanPer - 1) / (kScanGroups * kScanPer));
  __syncthreads();
  const uint32_t nt = tiles;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t c0 = t * kScanGroups * kScanPer + g * kScanPer;
    uint32_t v[kScanPer], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
      v[k] = c0 + k < nch ? hist[hist0 + (uint64_t)(c0 + k) * hp + b] : 0u;
      s += v[k];
    }
    part[g][bl] = s;
    __syncthreads();
    {  // exclusive prefix over the groups, after the earlier tiles: each
       // wave takes one bin, a lane per group
      const uint32_t tb = threadIdx.x / kScanGroups, tg = threadIdx.x % kScanGroups;
      const uint32_t x = part[tg][tb];
      const uint32_t inc = wave_prefix_sum(x);
      const uint32_t c = carry[tb];
      part[tg][tb] = c + inc - x;
      __syncthreads();  // every lane has read carry[tb]
      if (tg == kScanGroups - 1) carry[tb] = c + inc;
    }
    __syncthreads();
    uint32_t base = part[g][bl];
#pragma unroll
    for (uint32_t k = 0; k < kScanPer; ++k) {
      if (c0 + k < nch) off[off0 + (uint64_t)(c0 + k) * op + b] = base;
      base += v[k];
    }
    __syncthreads();  // part[] is rewritten by the next tile
  }
  if (g == 0 && gb < n_bins) total[gb] = carry[bl];
}

// Exclusive scan of n values by one 1024-thread workgroup, in tiles of 4096:
// each thread takes 4 consecutive entries (a wave reads one contiguous 1 KB
// run), wave and block scans, the running total carried between tiles.
// Returns the grand total (every thread).
struct ScanLds {
  uint64_t w[16];
};

template <typename Get, typename Put>
__device__ __forceinline__ uint64_t block_exscan(uint32_t n, ScanLds& L, Get get, Put put) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint64_t carry = 0;
  for (uint32_t base = 0; base < n; base += 4096) {
    const uint32_t i0 = base + tid * 4;
    uint64_t x[4], acc = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      x[k] = i0 + k < n ? get(i0 + k) : 0ull;
      acc += x[k];
    }
    const uint64_t inc = wave_prefix_sum64(acc);
    if (lane == 63) L.w[w] = inc;
    __syncthreads();
    uint64_t ex = carry + inc - acc, tile = 0;
    for (uint32_t k = 0; k < 16; ++k) {
      if (k < w) ex += L.w[k];
      tile += L.w[k];
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      if (i0 + k < n) put(i0 + k, ex);
      ex += x[k];
    }
    carry += tile;
    __syncthreads();  // L.w is reused by the next tile
  }
  return carry;
}

// Bucket offset of every bin (exclusive scan of all bins' totals), the
// grand total and the workspace check.  With dense_exit (a launch of one
// instance), an instance whose heavy bins (over what the launch's test
// schedule holds: kNarrowMax memberships when no box is wide, else
// kMaxCoarse) hold more than a quarter of its memberships is flagged
// kErrDense, and the scatter and test kernels then do nothing.
// With `order`, also the test kernel's bin order: the bins in four tiers by
// memberships (the schedule's overfull bins, which the test only lists, in
// the last), heaviest first, in bin order inside a tier, so that the test
// workgroups, which take the order in snake rounds, each get a share of the
// heavy bins, and every launch of the same input the same share.
__device__ __forceinline__ uint32_t test_tier(uint32_t m, uint32_t limit) {
  return m > limit ? 3u : (m >= 2048u ? 0u : (m >= 512u ? 1u : (m >= 64u ? 2u : 3u)));
}

template <bool kDev>
__global__ __launch_bounds__(1024) void k_bin_starts(const uint32_t* __restrict__ total, uint32_t n_bins,
                                                     uint32_t* __restrict__ bin_start, uint64_t cap,
                                                     uint32_t dense_exit, Meta* __restrict__ meta,
                                                     uint32_t* __restrict__ order, uint32_t* __restrict__ rank) {
  __shared__ ScanLds L;
  __shared__ unsigned long long heavy;
  __shared__ uint32_t tier_w[16][4];
  if (kDev) n_bins = meta->plan_err ? 0u : meta->dev_bins;
  if (threadIdx.x == 0) heavy = 0ull;
  const uint64_t grand = block_exscan(
      n_bins, L, [&](uint32_t i) { return (uint64_t)total[i]; },
      [&](uint32_t i, uint64_t ex) { bin_start[i] = (uint32_t)(ex < cap ? ex : cap); });
  if (dense_exit) {
    const uint32_t limit = general_schedule(meta) ? kMaxCoarse : kNarrowMax;
    uint64_t h = 0;
    for (uint32_t i = threadIdx.x; i < n_bins; i += blockDim.x) {
      const uint32_t t = total[i];
      h += t > limit ? t : 0u;
    }
    h = wave_prefix_sum64(h);
    if ((threadIdx.x & 63) == 63 && h) atomicAdd(&heavy, (unsigned long long)h);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    meta->memberships = grand;
    if (grand > cap) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrWorkspace);
    if (dense_exit && 4ull * heavy > grand)
      atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrDense);
  }
  if (order) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t limit = general_schedule(meta) ? kMaxCoarse : kNarrowMax;
    const uint32_t per = (n_bins + 1023u) / 1024u;
    const uint32_t i0 = tid * per < n_bins ? tid * per : n_bins;
    const uint32_t i1 = i0 + per < n_bins ? i0 + per : n_bins;
    uint32_t c[4] = {0u, 0u, 0u, 0u};
    for (uint32_t i = i0; i < i1; ++i) ++c[test_tier(total[i], limit)];
    uint32_t at[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t inc = wave_prefix_sum(c[k]);
      if (lane == 63) tier_w[w][k] = inc;
      at[k] = inc - c[k];
    }
    __syncthreads();
    uint32_t base = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t before = 0, all = 0;
      for (uint32_t v = 0; v < 16; ++v) {
        before += v < w ? tier_w[v][k] : 0u;
        all += tier_w[v][k];
      }
      at[k] += base + before;
      base += all;
    }
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t r = at[test_tier(total[i], limit)]++;
      order[r] = i;
      rank[i] = r;
    }
  }
}

// The scatter holds its whole chunk in registers (kScatterPer boxes per
// thread) and ranks every membership among the chunk's memberships in bin
// order with LDS atomics on per-bin cursors; the memberships are staged in
// LDS at their rank, then written out in rank order, so a wave's stores
// cover whole runs of each bin.  Every workgroup walks the bins in the same
// order, and the runs that share cache lines (adjacent chunks on one XCD)
// are written within the same short window and merge in that XCD's L2.  A
// chunk with more memberships than the stage writes the rest directly.
constexpr uint32_t kScatterThreads = 1024;
constexpr uint32_t kScatterPer = kScatterChunk / kScatterThreads;
constexpr uint32_t kStageEntries = kScatterChunk + kScatterChunk / 8;  // 1.125 memberships per box
constexpr uint32_t kScatterRowsPer = (kMaxBins + kScatterThreads - 1) / kScatterThreads;  // bins per thread, at most

// Fixed part of the scatter's LDS; after it, sized per launch to the
// batch's largest instance (nb_cap bins), come
//   uint32_t cur[nb_cap]   next destination of each bin (global index)
//   int32_t woff[nb_cap]   chunk-local rank of a destination = dst + woff
struct ScatterLds {
  float4 box[kStageEntries];
  uint32_t id[kStageEntries];
  uint32_t dst[kStageEntries];
  uint32_t wsum[16];
  ScanLds scan;              // the bucket starts of a single instance (fold)
  unsigned long long heavy;  // its memberships in bins over the test schedule's capacity
};

// A single instance with at most kFoldBins bins: the scatter scans the bins'
// totals itself (bucket starts in LDS after the cursors, nb_cap more words),
// and no k_bin_starts launch precedes it.
constexpr uint32_t kFoldBins = 2048;

// kIds: the boxes' ids come from ids_in (virtual instances), else they are
// the index inside the instance.
template <bool kIds, bool kDev>
__global__ __launch_bounds__(kScatterThreads) void k_bin_scatter(const float4* __restrict__ boxes,
                                                                 const uint32_t* __restrict__ ids_in,
                                                                 const InstDesc* __restrict__ inst,
                                                                 const uint32_t* __restrict__ chunk_inst,
                                                                 uint32_t n_chunks,
                                                                 const uint32_t* __restrict__ hist,
                                                                 const uint32_t* __restrict__ off,
                                                                 uint32_t* __restrict__ bin_start,
                                                                 float4* __restrict__ bbox,
                                                                 uint32_t* __restrict__ bid, uint64_t cap,
                                                                 uint32_t nb_cap, const uint32_t* __restrict__ total,
                                                                 uint32_t fold, uint32_t dense_exit,
                                                                 Meta* __restrict__ meta) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  ScatterLds& L = *reinterpret_cast<ScatterLds*>(lds_raw);
  uint32_t* const cur = reinterpret_cast<uint32_t*>(lds_raw + sizeof(ScatterLds));
  int32_t* const woff = reinterpret_cast<int32_t*>(cur + nb_cap);
  uint32_t* const bst = reinterpret_cast<uint32_t*>(woff + nb_cap);  // fold: the bins' bucket starts
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // Persistent workgroups, XCD-aware.  Workgroups go to the 8 XCDs
  // round-robin; the gridDim / 8 workgroups of XCD x walk that XCD's
  // contiguous range of chunks (x * per .. x * per + per - 1) together, so
  // the runs that chunks c and c + 1 write side by side in every bin's
  // bucket come from the same XCD at about the same time and merge in its
  // L2.  The next chunk's boxes and table rows are loaded while the current
  // chunk's staged memberships are written out.
  if (kDev) n_chunks = meta->plan_err ? 0u : meta->dev_chunks;
  if (!fold) {
    if (meta->memberships > cap) return;  // the bucket would overflow; the host regrows it
    if (meta->error & kErrDense) return;  // the host runs the general path
  }
  const uint32_t xper = (n_chunks + 7u) / 8u;
  const uint32_t xnext = gridDim.x / 8u;  // workgroups per XCD (the grid is a multiple of 8)
  uint32_t r = blockIdx.x >> 3;
  uint32_t wc = (blockIdx.x & 7u) * xper + r;
  const uint32_t wend = min((blockIdx.x & 7u) * xper + xper, n_chunks);
  float4 q[kScatterPer];
  // The chunk's table rows for this thread's bins b0 .. b0 + per - 1 (count,
  // offset over the earlier chunks, bucket start), loaded with its boxes.
  uint32_t rh[kScatterRowsPer], ro[kScatterRowsPer], rs[kScatterRowsPer];
  auto load_chunk = [&](uint32_t c) {
    const InstDesc J = inst[chunk_inst[c]];
    uint64_t st, en;
    chunk_range(J, c, &st, &en);
#pragma unroll
    for (uint32_t k = 0; k < kScatterPer; ++k) {
      const uint64_t i = st + tid + k * kScatterThreads;
      q[k] = boxes[i < en ? i : en - 1];
    }
    const uint32_t nbj = J.nb, per = (nbj + kScatterThreads - 1) / kScatterThreads, b0 = tid * per;
    const uint64_t lc = c - J.chunk0;
    const uint32_t* jh = hist + J.hist0 + lc * hist_pitch(nbj);
    const uint32_t* jo = off + J.off0 + lc * off_pitch(nbj);
    const uint32_t* js = bin_start + J.bin0;
#pragma unroll
    for (uint32_t k = 0; k < kScatterRowsPer; ++k) {
      const bool in = k < per && b0 + k < nbj;
      rh[k] = in ? jh[b0 + k] : 0u;
      ro[k] = in ? jo[b0 + k] : 0u;
      rs[k] = in && !fold ? js[b0 + k] : 0u;
    }
  };
  const bool narrow = !meta->wide;  // no box of the launch spans more than two fine cells
  if (wc < wend) load_chunk(wc);
  if (fold) {
    // One instance: every workgroup scans its bins' totals into LDS (the
    // bucket starts) and checks them as k_bin_starts would, while its first
    // chunk loads; workgroup 0 also publishes the starts, the membership
    // count and the error bits.  Every workgroup reaches the same verdict.
    const InstDesc I0 = inst[0];
    const uint32_t limit = general_schedule(meta) ? kMaxCoarse : kNarrowMax;
    if (tid == 0) L.heavy = 0ull;
    uint64_t hv = 0;
    const uint64_t grand = block_exscan(
        I0.nb, L.scan,
        [&](uint32_t i) {
          const uint32_t t = total[I0.bin0 + i];
          hv += t > limit ? t : 0u;
          return (uint64_t)t;
        },
        [&](uint32_t i, uint64_t ex) {
          const uint32_t v = (uint32_t)(ex < cap ? ex : cap);
          bst[i] = v;
          if (blockIdx.x == 0) bin_start[I0.bin0 + i] = v;
        });
    hv = wave_prefix_sum64(hv);
    if (lane == 63 && hv) atomicAdd(&L.heavy, hv);
    __syncthreads();
    const bool dense = dense_exit && 4ull * L.heavy > grand;
    if (blockIdx.x == 0 && tid == 0) {
      meta->memberships = grand;
      if (grand > cap) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrWorkspace);
      if (dense) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrDense);
    }
    if (grand > cap || dense) return;
  }
  for (; wc < wend; wc += xnext) {
  const InstDesc I = inst[chunk_inst[wc]];
  const uint32_t nb = I.nb;
  uint64_t start, end;
  chunk_range(I, wc, &start, &end);
  // 2. per bin: destination cursor, and the offset from a destination to the
  //    membership's rank among this chunk's memberships in bin order
  uint32_t chunk_total = 0;
  __syncthreads();  // the previous chunk's stage and cursors are done with
  {
    const uint32_t per = (nb + kScatterThreads - 1) / kScatterThreads;
    const uint32_t b0 = tid * per;
    uint32_t acc = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScatterRowsPer; ++k) acc += rh[k];  // zero past this thread's bins
    const uint32_t v = wave_prefix_sum(acc);
    if (lane == 63) L.wsum[w] = v;
    __syncthreads();
    uint32_t ex = v - acc;
    for (uint32_t k = 0; k < kScatterThreads / 64; ++k) {
      if (k < w) ex += L.wsum[k];
      chunk_total += L.wsum[k];
    }
#pragma unroll
    for (uint32_t k = 0; k < kScatterRowsPer; ++k) {
      const uint32_t b = b0 + k;
      if (k < per && b < nb) {
        const uint32_t dst0 = (fold ? bst[b] : rs[k]) + ro[k];
        cur[b] = dst0;
        woff[b] = (int32_t)(ex - dst0);
        ex += rh[k];
      }
    }
  }
  __syncthreads();
  // 3. rank every membership; stage it at its rank (or write it directly
  //    past the stage).  When no box of the launch spans more than two fine
  //    cells per axis, none spans more than two bins per axis: a box's bins
  //    are its first one and, when it crosses into them, the next one east,
  //    north and north-east, placed without loops.
  auto place = [&](uint32_t bb, const float4& b, uint32_t i) {
    const uint32_t d = atomicAdd(&cur[bb], 1u);
    const uint32_t at = (uint32_t)((int32_t)d + woff[bb]);
    if (at < kStageEntries) {
      L.box[at] = b;
      L.id[at] = i;
      L.dst[at] = d;
    } else if (d < cap) {
      bbox[d] = b;
      bid[d] = i;
    }
  };
#pragma unroll
  for (uint32_t k = 0; k < kScatterPer; ++k) {
    const uint64_t gi = start + tid + k * kScatterThreads;
    const float4 b = q[k];
    if (!(gi < end && (b.x <= b.y) && (b.z <= b.w))) continue;
    const BinRange rg = bin_range(b, I);
    uint32_t i = (uint32_t)(gi - I.box0);  // id inside the instance
    if (kIds) i = ids_in[gi];
    if (narrow) {
      const uint32_t bb = rg.y0 * I.nbx + rg.x0;
      const bool ex = rg.x1 != rg.x0, ey = rg.y1 != rg.y0;
      place(bb, b, i);
      if (ex) place(bb + 1u, b, i);
      if (ey) place(bb + I.nbx, b, i);
      if (ex && ey) place(bb + I.nbx + 1u, b, i);
    } else {
      for (uint32_t cy = rg.y0; cy <= rg.y1; ++cy)
        for (uint32_t cx = rg.x0; cx <= rg.x1; ++cx) place(cy * I.nbx + cx, b, i);
    }
  }
  __syncthreads();
  if (wc + xnext < wend) load_chunk(wc + xnext);  // in flight during the write-out
  // 4. the staged memberships in rank order
  const uint32_t staged = chunk_total < kStageEntries ? chunk_total : kStageEntries;
  for (uint32_t k = tid; k < staged; k += kScatterThreads) {
    const uint32_t d = L.dst[k];
    if (d < cap) {
      bbox[d] = L.box[k];
      bid[d] = L.id[k];
    }
  }
  }
}

// ---------------------------------------------------------------------------
// Single-pass placement of one instance's boxes into fixed-capacity bins.
//
// One persistent 1024-thread workgroup per CU takes a contiguous range of
// boxes, 16,384 at a time (16 per thread, held in registers):
//   1. counts the round's memberships per coarse bin with LDS atomics: a
//      box's first bin returns the box's rank there (kept, 16 bits, in a
//      register); its other bins (east, north, north-east of a box that
//      crosses a bin edge, or every bin of a wide box) are counted apart;
//   2. reserves the round's share of every bin it touches with one device
//      atomic on fill[b] (about 1,000 per round instead of one per
//      membership): first memberships, then the others.  A share past the
//      bin's capacity flags kErrBinCap, and the workgroup stops placing
//      (with count_all its later rounds only count, so that the counters
//      end with every bin's total);
//   3. writes the other memberships straight to slots handed out by LDS
//      atomics, from a list of the boxes that have any (kept in the stage's
//      space during the count: a few percent of narrow boxes; a round with
//      more than the stage holds fails over like a bin over capacity);
//   4. places the first memberships through an LDS stage, in passes of
//      `stage` entries in the workgroup's bin order (a block scan of the
//      first counts), so that a wave's stores cover whole runs of a bin.
// The boxes are read once (no histogram pass and no scan kernel).  The call
// also zeroes the next call's counters and result words.
constexpr uint32_t kPlaceThreads = 1024;
constexpr uint32_t kPlacePer = 16;
constexpr uint32_t kPlaceRound = kPlaceThreads * kPlacePer;
// A box's first bin and its rank there share one register word.
constexpr uint32_t kRankShift = 13;
static_assert(kMaxBins < (1u << kRankShift) && kPlaceRound <= (1u << (32 - kRankShift)), "bin | rank << 13");
// Register word of a box in the round: its first bin, and bits for "also
// the bin east" / "also the bin north"; kSlotWide: the box spans more than
// two fine cells on an axis (all its bins go the direct way);
// kSlotNone: no box (past the range or invalid).
constexpr uint32_t kSlotEast = 1u << 29, kSlotNorth = 1u << 30, kSlotWide = 1u << 31;
constexpr uint32_t kSlotBin = kSlotEast - 1u;
constexpr uint32_t kSlotNone = 0xFFFFFFFFu;
static_assert(kMaxBins <= kSlotBin, "bin field");

// Dynamic LDS of the place kernel: three words per bin, then the stage.
__host__ __device__ inline uint32_t place_stage_entries(uint32_t nb, uint32_t lds_bytes) {
  const uint32_t fixed = 12u * nb + 64u;
  const uint32_t e = lds_bytes > fixed ? (lds_bytes - fixed) / 24u : 0u;
  return (e > 8192u ? 8192u : e) & ~63u;
}

__global__ __launch_bounds__(kPlaceThreads) void k_bin_place(const float4* __restrict__ boxes,
                                                             const InstDesc* __restrict__ inst,
                                                             uint32_t* __restrict__ fill,
                                                             uint32_t* __restrict__ fill_next, uint32_t fill_words,
                                                             float4* __restrict__ bbox, uint32_t* __restrict__ bid,
                                                             uint32_t bin_cap, uint32_t stage,
                                                             uint32_t* __restrict__ bin_start,
                                                             Meta* __restrict__ meta,
                                                             Meta* __restrict__ meta_next, uint32_t count_all) {
  extern __shared__ __align__(16) unsigned char place_raw[];
  __shared__ uint32_t flags;  // 1 invalid box, 4 bin over capacity (or too many boxes listed)
  __shared__ uint32_t wsum[kPlaceThreads / 64];
  __shared__ uint32_t npend;  // boxes of the round with memberships besides a first one
  __shared__ uint64_t resume;           // first box not counted when a bin outgrew its share
  __shared__ uint32_t nwide;            // wide boxes of the workgroup's range
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const InstDesc I = inst[0];
  const uint32_t nb = I.nb, nbx = I.nbx;
  float4* const sbox = reinterpret_cast<float4*>(place_raw);     // stage: boxes,
  uint32_t* const sid = reinterpret_cast<uint32_t*>(sbox + stage);  // ids,
  uint32_t* const sdst = sid + stage;                              // destinations
  uint32_t* const cnt1 = sdst + stage;  // per bin: first memberships, then their position in the stage order
  uint32_t* const cnt2 = cnt1 + nb;     // per bin: other memberships, then their next bucket slot
  uint32_t* const base1 = cnt2 + nb;    // per bin: bucket slot of the first memberships' run
  for (uint32_t i = blockIdx.x * kPlaceThreads + tid; i < fill_words; i += gridDim.x * kPlaceThreads)
    fill_next[i] = 0u;
  for (uint32_t b = blockIdx.x * kPlaceThreads + tid; b < I.nb; b += gridDim.x * kPlaceThreads)
    bin_start[b] = b * bin_cap;  // every bin's share starts at b * bin_cap (read by the test kernel)
  if (blockIdx.x == 0 && tid < sizeof(Meta) / sizeof(unsigned long long))
    reinterpret_cast<unsigned long long*>(meta_next)[tid] = 0ull;
  if (tid == 0) {
    flags = 0u;
    nwide = 0u;
  }
  const uint64_t n = I.n;
  const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t g0 = (uint64_t)blockIdx.x * per;
  const uint64_t g1 = g0 + per < n ? g0 + per : n;
  const uint32_t bper = (nb + kPlaceThreads - 1) / kPlaceThreads;  // bins per thread in the scan
  for (uint64_t r0 = g0; r0 < g1; r0 += kPlaceRound) {
    const uint64_t r1 = r0 + kPlaceRound < g1 ? r0 + kPlaceRound : g1;
    for (uint32_t b = tid; b < nb; b += kPlaceThreads) {
      cnt1[b] = 0u;
      cnt2[b] = 0u;
    }
    if (tid == 0) npend = 0u;
    float4 q[kPlacePer];
#pragma unroll
    for (uint32_t k = 0; k < kPlacePer; ++k) {
      const uint64_t i = r0 + tid + k * kPlaceThreads;
      q[k] = boxes[I.box0 + (i < r1 ? i : r1 - 1)];
    }
    __syncthreads();  // counters zeroed (and the last round placed)
    // 1. count; a first membership's bin and rank in it are kept
    uint32_t slot[kPlacePer];
    uint32_t fl = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPlacePer; ++k) {
      const uint64_t i = r0 + tid + k * kPlaceThreads;
      const float4 b = q[k];
      slot[k] = kSlotNone;
      uint32_t rk = 0u;
      if (i < r1) {
        if (!((b.x <= b.y) && (b.z <= b.w))) {
          fl |= 1u;
        } else {
          const uint4 f = fine_corners(b, I);
          const BinRange rg = bin_range_of(f, I);
          const uint32_t bb = rg.y0 * nbx + rg.x0;
          uint32_t sw;
          if (f.y - f.x > 1u || f.w - f.z > 1u) {
            atomicAdd(&nwide, 1u);
            sw = bb | kSlotWide;
            for (uint32_t cy = rg.y0; cy <= rg.y1; ++cy)
              for (uint32_t cx = rg.x0; cx <= rg.x1; ++cx) atomicAdd(&cnt2[cy * nbx + cx], 1u);
          } else {
            const bool ex = rg.x1 != rg.x0, ey = rg.y1 != rg.y0;
            sw = bb | (ex ? kSlotEast : 0u) | (ey ? kSlotNorth : 0u);
            rk = atomicAdd(&cnt1[bb], 1u);
            if (ex) atomicAdd(&cnt2[bb + 1u], 1u);
            if (ey) atomicAdd(&cnt2[bb + nbx], 1u);
            if (ex && ey) atomicAdd(&cnt2[bb + nbx + 1u], 1u);
          }
          slot[k] = sw & kSlotWide ? kSlotNone : bb | (rk << kRankShift);  // the staged passes' boxes
          if (sw & (kSlotWide | kSlotEast | kSlotNorth)) {  // listed for the direct writes
            const uint32_t at = atomicAdd(&npend, 1u);
            if (at < stage) {
              sbox[at] = b;
              sid[at] = (uint32_t)i;
              sdst[at] = sw;
            } else {
              fl |= 4u;
            }
          }
        }
      }
    }
    if (fl) atomicOr(&flags, fl);
    __syncthreads();
    // 2. reserve every touched bin's share, and the stage order of the first
    //    memberships (exclusive scan of their counts over the bins)
    uint32_t acc = 0;
    for (uint32_t k = 0; k < bper; ++k) {
      const uint32_t b = tid * bper + k;
      if (b < nb) {
        const uint32_t c1 = cnt1[b], c2 = cnt2[b];
        uint32_t base = 0u;
        if (c1 + c2) {
          base = atomicAdd(&fill[b], c1 + c2);
          if (base + c1 + c2 > bin_cap) atomicOr(&flags, 4u);
        }
        base1[b] = b * bin_cap + base;
        cnt2[b] = b * bin_cap + base + c1;
        acc += c1;
      }
    }
    const uint32_t inc = wave_prefix_sum(acc);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t ex = inc - acc, total1 = 0;
    for (uint32_t k = 0; k < kPlaceThreads / 64; ++k) {
      if (k < w) ex += wsum[k];
      total1 += wsum[k];
    }
    for (uint32_t k = 0; k < bper; ++k) {
      const uint32_t b = tid * bper + k;
      if (b < nb) {
        const uint32_t c1 = cnt1[b];
        cnt1[b] = ex;
        ex += c1;
      }
    }
    __syncthreads();
    if (flags & 4u) {  // the launch fails over to another pipeline
      if (tid == 0) resume = r1;
      break;
    }
    // 3. the listed boxes' other memberships (every one of a wide box)
    {
      const uint32_t np = npend;
      for (uint32_t j = tid; j < np; j += kPlaceThreads) {
        const float4 b = sbox[j];
        const uint32_t id = sid[j], sw = sdst[j];
        auto put = [&](uint32_t bin) {
          const uint32_t d = atomicAdd(&cnt2[bin], 1u);
          bbox[d] = b;
          bid[d] = id;
        };
        if (sw & kSlotWide) {
          const BinRange rg = bin_range_of(fine_corners(b, I), I);
          for (uint32_t cy = rg.y0; cy <= rg.y1; ++cy)
            for (uint32_t cx = rg.x0; cx <= rg.x1; ++cx) put(cy * nbx + cx);
        } else {
          const uint32_t bb = sw & kSlotBin;
          if (sw & kSlotEast) put(bb + 1u);
          if (sw & kSlotNorth) put(bb + nbx);
          if ((sw & kSlotEast) && (sw & kSlotNorth)) put(bb + nbx + 1u);
        }
      }
      __syncthreads();  // the stage is reused
    }
    // 4. first memberships, staged in passes in the workgroup's bin order:
    //    each box's word becomes (bin, position in that order); the writer
    //    finds the bucket slot from the staged bin
    const uint32_t kBinMask = (1u << kRankShift) - 1u;
#pragma unroll
    for (uint32_t k = 0; k < kPlacePer; ++k) {
      const uint32_t sl = slot[k];
      if (sl != kSlotNone) slot[k] = (sl & kBinMask) | ((cnt1[sl & kBinMask] + (sl >> kRankShift)) << kRankShift);
    }
    for (uint32_t v0 = 0; v0 < total1; v0 += stage) {
      const uint32_t lim = total1 - v0 < stage ? total1 - v0 : stage;
#pragma unroll
      for (uint32_t k = 0; k < kPlacePer; ++k) {
        const uint32_t sl = slot[k];
        const uint32_t j = (sl >> kRankShift) - v0;
        if (sl != kSlotNone && j < lim) {
          sbox[j] = q[k];
          sid[j] = (uint32_t)(r0 + tid + k * kPlaceThreads);
          sdst[j] = sl & kBinMask;
        }
      }
      __syncthreads();
      for (uint32_t j = tid; j < lim; j += kPlaceThreads) {
        const uint32_t bb = sdst[j];
        const uint32_t d = base1[bb] + (v0 + j - cnt1[bb]);
        bbox[d] = sbox[j];
        bid[d] = sid[j];
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // A bin over its share, with count_all: the rest of the workgroup's range
  // is only counted (every membership in LDS, one device atomic per touched
  // bin and round), so that the counters end with every bin's total (the
  // test kernel then applies the dense rule to them).
  if ((flags & 4u) && count_all) {
    for (uint64_t r0 = resume; r0 < g1; r0 += kPlaceRound) {
      const uint64_t r1 = r0 + kPlaceRound < g1 ? r0 + kPlaceRound : g1;
      for (uint32_t b = tid; b < nb; b += kPlaceThreads) cnt1[b] = 0u;
      __syncthreads();
      for (uint64_t i = r0 + tid; i < r1; i += kPlaceThreads) {
        const float4 b = boxes[I.box0 + i];
        if (!((b.x <= b.y) && (b.z <= b.w))) continue;  // (flagged by no one: the launch has failed anyway)
        const uint4 f = fine_corners(b, I);
        if (f.y - f.x > 1u || f.w - f.z > 1u) atomicAdd(&nwide, 1u);
        const BinRange rg = bin_range_of(f, I);
        for (uint32_t cy = rg.y0; cy <= rg.y1; ++cy)
          for (uint32_t cx = rg.x0; cx <= rg.x1; ++cx) atomicAdd(&cnt1[cy * nbx + cx], 1u);
      }
      __syncthreads();
      for (uint32_t b = tid; b < nb; b += kPlaceThreads)
        if (cnt1[b]) atomicAdd(&fill[b], cnt1[b]);
      __syncthreads();
    }
  }
  if (tid == 0) {
    if (flags & 1u) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrInvalidBox);
    if (nwide) atomicAdd(&meta->wide, nwide);
    if (flags & 4u) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrBinCap);
    if (blockIdx.x == 0) meta->boxes = n < 0xFFFFFFFFull ? (uint32_t)n : 0xFFFFFFFFu;
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
// fine-cell order in LDS (cells by size class) and every cell's pairs are
// tested.  A box covering more than kHugeCells cells of the bin is listed
// instead and tested against every membership of the bin.  Pairs go to a
// per-wave LDS stage and then to this workgroup's own region of the pair
// scratch through an LDS counter: no global atomics.  k_bin_compact later
// concatenates the regions.  The next bin's registers are loaded after the
// last strip's scatter and land while that strip is tested.

constexpr uint32_t kTestThreads = 1024;
constexpr uint32_t kUnroll = 1;  // partner loads issued together in the test loop
constexpr uint32_t kTestWaves = kTestThreads / 64;
constexpr uint32_t kRegs = kMaxCoarse / kTestThreads;  // coarse memberships per thread in registers

// Position word: coarse-membership index (13 bits) | local fine cell (12 bits)
// | "starts in this row" (bit 30) | "starts in this column" (bit 31).
// A box covering more than kHugeCells fine cells of a bin is not replicated
// into them (one lane would place it cell by cell, the workgroup waiting):
// up to kHugeMax such boxes per bin are listed and tested against every
// coarse membership of the bin, with the reference-cell rule on the bin.
constexpr uint32_t kHugeCells = 16, kHugeMax = 256;
constexpr uint32_t kFrHuge = 1u << 31;  // packed fine range: the box is listed (index in the low bits)
constexpr uint32_t kWordCellShift = 13;
constexpr uint32_t kWordIdxMask = (1u << kWordCellShift) - 1u;
constexpr uint32_t kWordCellMask = (1u << 12) - 1u;
static_assert(kMaxCoarse <= (1u << kWordCellShift), "membership index field");
static_assert(kMaxCellsPerBin <= (1u << 12), "cell field");

// Cells of a strip are laid out by size class (min(count, 63), largest
// first), so that the 64 positions of a wave belong to cells of about the
// same size and the wave's partner loop (its longest lane) wastes few slots.
constexpr uint32_t kSizeClasses = 64;

struct TestLds {
  float4 box[kMaxFine];            // strip memberships in fine-cell order
  uint32_t word[kMaxFine];         // position words
  uint32_t id[kMaxCoarse];         // box id of each coarse membership
  uint32_t cnt[kMaxCellsPerBin];   // fine-cell counts of the bin, then the scatter cursors
  uint32_t start[kMaxCellsPerBin + 1];  // first position of each cell's memberships
  uint16_t ccount[kMaxCellsPerBin];     // memberships of each cell
  uint32_t csum[kSizeClasses];     // memberships of the cells of each size class
  uint32_t cstart[kSizeClasses];   // first position of each size class, then its cursor
  uint32_t nfine;                  // fine memberships of the bin
  uint32_t npair;                  // positions in cells of two or more memberships (sorted layout)
  uint32_t nhuge;                  // the bin's large boxes (listed, not replicated)
  uint2 stage[kTestWaves][kBinStage];   // per-wave staging of hits (coarse membership indices)
  uint2 binfo[kMaxBinsPerGroup];   // (start, count) of this workgroup's bins
  uint4 bingrid[kMaxBinsPerGroup]; // (first fine column, first fine row, kx | ky << 16, nf) of each bin
  float bininv[kMaxBinsPerGroup];  // inv_f of each bin's instance
  uint32_t wsum[kTestWaves];
  uint32_t out_n;                  // pairs appended to this workgroup's region
  uint32_t strip_hi;
  unsigned long long out_base;     // direct output: this workgroup's first output position
  unsigned long long memb;         // kSingleSum: memberships of this workgroup's bins
};
static_assert(sizeof(TestLds) <= 160u * 1024u, "one test workgroup per CU");

// Appends a wave's staged hits to the workgroup's region as (smaller id, larger id).
template <typename Lds>
__device__ __forceinline__ void stage_flush(Lds& L, uint32_t w, uint32_t& sn, uint2* __restrict__ region,
                                            uint32_t cap_wg) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t g = 0;
  if (lane == 0) g = atomicAdd(&L.out_n, sn);
  g = __builtin_amdgcn_readfirstlane(g);
  for (uint32_t k = lane; k < sn; k += 64)
    if (g + k < cap_wg) {
      const uint2 v = L.stage[w][k];
      const uint32_t ia = L.id[v.x], ib = L.id[v.y];
      region[g + k] = make_uint2(ia < ib ? ia : ib, ia < ib ? ib : ia);
    }
  sn = 0;
}

// Fine grid of one bin's instance and the bin's extent in fine cells (bins
// of a virtual instance may be cut short at its region's edge).
struct FineGrid {
  float inv_f;
  uint32_t nf, kx, ky;
};

// Fine-cell range of a box clipped to the bin whose first fine cell is (gx, gy).
struct FineRange {
  uint32_t x0, x1, y0, y1, fx0, fy0;
};

__device__ __forceinline__ FineRange fine_range(const float4& q, const FineGrid& p, uint32_t gx, uint32_t gy) {
  FineRange r;
  r.fx0 = fine_of(q.x, p.inv_f, p.nf);
  r.fy0 = fine_of(q.z, p.inv_f, p.nf);
  const uint32_t fx1 = fine_of(q.y, p.inv_f, p.nf), fy1 = fine_of(q.w, p.inv_f, p.nf);
  r.x0 = r.fx0 > gx ? r.fx0 - gx : 0u;
  r.y0 = r.fy0 > gy ? r.fy0 - gy : 0u;
  r.x1 = fx1 - gx < p.kx ? fx1 - gx : p.kx - 1;
  r.y1 = fy1 - gy < p.ky ? fy1 - gy : p.ky - 1;
  return r;
}

// A coarse membership's fine cells inside its bin, packed once per bin:
// x0 | x1 << 6 | y0 << 12 | y1 << 18 (bin-local, < 64), bit 24 = the box
// starts in this bin's columns (its first column is x0), bit 25 = it starts
// in this bin's rows.
__device__ __forceinline__ uint32_t pack_fine(const FineRange& r, uint32_t gx, uint32_t gy) {
  return r.x0 | (r.x1 << 6) | (r.y0 << 12) | (r.y1 << 18) | ((r.fx0 >= gx ? 1u : 0u) << 24) |
         ((r.fy0 >= gy ? 1u : 0u) << 25);
}

// Scatter the fine memberships of one coarse membership (packed range pr)
// that fall in cells [c_lo, c_hi) to their positions in the strip (cnt[] is
// the cursor, set to start[] by the scan).
__device__ __forceinline__ void scatter_strip(TestLds& L, const float4& q, uint32_t j, uint32_t pr, uint32_t kx,
                                              uint32_t c_lo, uint32_t c_hi, uint32_t base) {
  const uint32_t x0 = pr & 63u, x1 = (pr >> 6) & 63u, ry0 = (pr >> 12) & 63u, ry1 = (pr >> 18) & 63u;
  const uint32_t sx = (pr >> 24) & 1u, sy = (pr >> 25) & 1u;
  const uint32_t row_lo = c_lo / kx, row_hi = (c_hi - 1) / kx;
  const uint32_t y0 = ry0 > row_lo ? ry0 : row_lo, y1 = ry1 < row_hi ? ry1 : row_hi;
  for (uint32_t cy = y0; cy <= y1; ++cy) {
    const uint32_t fy = (sy && cy == ry0) ? kFlagY : 0u;
    for (uint32_t cx = x0; cx <= x1; ++cx) {
      const uint32_t c = cy * kx + cx;
      if (c < c_lo || c >= c_hi) continue;
      const uint32_t pos = atomicAdd(&L.cnt[c], 1u) - base;
      L.box[pos] = q;
      L.word[pos] = j | (c << kWordCellShift) | fy | ((sx && cx == x0) ? kFlagX : 0u);
    }
  }
}

// The single-pass launch's bins' totals are the place kernel's counters:
// each test workgroup adds its bins' totals to the memberships in the result
// words (kSingleSum).  When a bin outgrew its share and the place kernel
// counted to the end (kSingleDense), workgroup 0 applies the
// automatic path choice's dense rule to them: bins beyond what the two-pass
// test schedule holds (kNarrowMax memberships, kMaxCoarse with wide boxes)
// carrying over a quarter of the memberships flag kErrDense, as the
// two-pass launch's bin totals would.
constexpr uint32_t kSingleSum = 1u, kSingleDense = 2u;

__device__ void single_pass_dense_rule(const uint32_t* __restrict__ total, uint32_t n_bins, uint32_t single,
                                       unsigned long long err, Meta* __restrict__ meta) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  if (!(err & kErrBinCap) || !(single & kSingleDense) || blockIdx.x != 0 || (err & kErrInvalidBox)) return;
  const uint32_t limit = general_schedule(meta) ? kMaxCoarse : kNarrowMax;
  unsigned long long h = 0, g = 0;
  for (uint32_t b = tid; b < n_bins; b += blockDim.x) {
    const uint32_t t = total[b];
    g += t;
    h += t > limit ? t : 0u;
  }
  for (uint32_t o = 32; o; o >>= 1) {
    h += __shfl_xor(h, o);
    g += __shfl_xor(g, o);
  }
  __shared__ unsigned long long sh, sg;
  if (tid == 0) sh = sg = 0ull;
  __syncthreads();
  if (lane == 0) {
    atomicAdd(&sh, h);
    atomicAdd(&sg, g);
  }
  __syncthreads();
  if (tid == 0) {
    meta->memberships = sg;
    if (4ull * sh > sg) atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrDense);
  }
}

// The bins of this test workgroup: with an order (k_bin_starts), round t
// takes entry t * grid + (t even: blockIdx.x, odd: grid - 1 - blockIdx.x);
// without, bin blockIdx.x + t * grid.  test_wg: the workgroup (so the pair
// region) of bin gb, from its rank in the order.
__device__ __forceinline__ uint32_t test_bin_count(uint32_t n_bins, const uint32_t* order) {
  const uint32_t G = gridDim.x, w = blockIdx.x;
  if (!order) return w < n_bins ? (n_bins - w + G - 1u) / G : 0u;
  const uint32_t full = n_bins / G, rem = n_bins % G;
  return full + (((full & 1u) ? G - 1u - w : w) < rem ? 1u : 0u);
}
__device__ __forceinline__ uint32_t test_bin(uint32_t t, const uint32_t* order) {
  const uint32_t G = gridDim.x, w = blockIdx.x;
  return order ? order[t * G + ((t & 1u) ? G - 1u - w : w)] : w + t * G;
}
__device__ __forceinline__ uint32_t test_wg(uint32_t gb, uint32_t nwg, const uint32_t* rank) {
  if (!rank) return gb % nwg;
  const uint32_t idx = rank[gb], t = idx / nwg, pos = idx % nwg;
  return (t & 1u) ? nwg - 1u - pos : pos;
}

__device__ __forceinline__ void test_general(unsigned char* lds_raw, const float4* __restrict__ bbox,
                                                           const uint32_t* __restrict__ bid,
                                                           const uint32_t* __restrict__ bin_start,
                                                           const uint32_t* __restrict__ total,
                                                           const InstDesc* __restrict__ inst,
                                                           const uint32_t* __restrict__ bin_inst, uint32_t n_bins,
                                                           uint64_t cap_bucket, uint2* __restrict__ scratch,
                                                           uint32_t cap_wg, uint2* __restrict__ bin_seg,
                                                           uint32_t* __restrict__ overfull,
                                                           uint2* __restrict__ out, uint64_t cap,
                                                           unsigned long long* out_count, Meta* __restrict__ meta,
                                                           uint32_t single, const uint32_t* __restrict__ order) {
  TestLds& L = *reinterpret_cast<TestLds*>(lds_raw);
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t my_bins = test_bin_count(n_bins, order);
  const unsigned long long err = meta->error;
  if (err & (kErrInvalidBox | kErrWorkspace | kErrDense | kErrBinCap)) {
    for (uint32_t t = tid; t < my_bins; t += kTestThreads) bin_seg[test_bin(t, order)] = make_uint2(0u, 0u);
    if (single) single_pass_dense_rule(total, n_bins, single, err, meta);
    return;
  }
  uint2* region = scratch + (size_t)blockIdx.x * cap_wg;
  uint32_t sn = 0;
  uint32_t seg0 = 0;  // region offset of the current bin's first pair (thread 0)
  // cnt[] is zero at the start of every bin: cleared here, then after each
  // bin's last scatter (and on the overflow exit)
  for (uint32_t c = tid; c < kMaxCellsPerBin; c += kTestThreads) L.cnt[c] = 0u;
  if (tid < kSizeClasses) L.csum[tid] = 0u;  // csum and nfine are also cleared after each use
  if (tid == 0) {
    L.out_n = 0u;
    L.nfine = 0u;
    L.nhuge = 0u;
    L.memb = 0ull;
  }

  // Registers: unconditional loads with the index clamped into the bucket.
  const uint64_t last = cap_bucket - 1;
  float4 rb[kRegs];
  uint32_t ri[kRegs];
  for (uint32_t t = 0; t < my_bins; ++t) {
    const uint32_t slot = t % kMaxBinsPerGroup;
    if (slot == 0) {
      // The next group of this workgroup's bins: location in the bucket and
      // their instance's grid.
      __syncthreads();
      for (uint32_t u = tid; u < kMaxBinsPerGroup && t + u < my_bins; u += kTestThreads) {
        const uint32_t gb = test_bin(t + u, order);
        const InstDesc I = inst[bin_inst[gb]];
        const uint32_t lb = gb - I.bin0;
        const uint32_t gx = I.fx0 + (lb % I.nbx) * I.k, gy = I.fy0 + (lb / I.nbx) * I.k;
        const uint32_t ex = I.fx0 + I.span_x - gx, ey = I.fy0 + I.span_y - gy;  // cells left in the region
        const uint32_t kx = I.k < ex ? I.k : ex, ky = I.k < ey ? I.k : ey;
        const uint32_t tt = total[gb];
        L.binfo[u] = make_uint2(bin_start[gb], tt);
        if ((single & kSingleSum) && tt) atomicAdd(&L.memb, (unsigned long long)tt);
        L.bingrid[u] = make_uint4(gx, gy, kx | (ky << 16), I.nf);
        L.bininv[u] = I.inv_f;
      }
      __syncthreads();
      if (t == 0) {
        const uint64_t s0 = L.binfo[0].x;
        const uint32_t j_hi = L.binfo[0].y ? L.binfo[0].y - 1u : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kRegs; ++k) {
          const uint32_t j = tid + k * kTestThreads < j_hi ? tid + k * kTestThreads : j_hi;
          const uint64_t at = s0 + j < last ? s0 + j : last;
          rb[k] = bbox[at];
          ri[k] = bid[at];
        }
      }
    }
    const uint32_t b = test_bin(t, order);
    const uint32_t bN = t + 1 < my_bins ? test_bin(t + 1, order) : 0u;
    const uint32_t m = L.binfo[slot].y;
    // Next bin's registers: entries past its end repeat its last one (same
    // cache line, no extra HBM traffic) so that every load is unconditional.
    const uint64_t sN = t + 1 < my_bins ? bin_start[bN] : 0ull;
    const uint32_t mN = t + 1 < my_bins ? total[bN] : 0u;
    const uint32_t jN = mN ? mN - 1u : 0u;
    const uint4 bg = L.bingrid[slot];
    const uint32_t gx = bg.x, gy = bg.y, Kx = bg.z & 0xFFFFu, Ky = bg.z >> 16, kk = Kx * Ky;
    const FineGrid p{L.bininv[slot], bg.w, Kx, Ky};
    bool prefetched = false;
    if (m >= 2 && m <= kMaxCoarse) {
      // A. fine-cell counts of the whole bin; ids to LDS
      uint32_t fr[kRegs];  // packed fine ranges, reused by every strip's scatter (or kFrHuge | list index)
      uint32_t mine = 0;   // this thread's fine memberships
      // the large boxes' list: the tail of the strip buffer, free until the
      // first strip's scatter
      float4* const hbox = L.box + (kMaxFine - kHugeMax);
      uint32_t* const hidx = L.word + (kMaxFine - kHugeMax);
      uint32_t* const hcx = L.word + (kMaxFine - 2u * kHugeMax);
      uint32_t* const hcy = L.word + (kMaxFine - 3u * kHugeMax);
#pragma unroll
      for (uint32_t k = 0; k < kRegs; ++k) {
        const uint32_t j = tid + k * kTestThreads;
        fr[k] = 0u;
        if (j < m) {
          L.id[j] = ri[k];
          const FineRange r = fine_range(rb[k], p, gx, gy);
          const uint32_t cells = (r.x1 - r.x0 + 1u) * (r.y1 - r.y0 + 1u);
          const uint32_t at = cells > kHugeCells ? atomicAdd(&L.nhuge, 1u) : kHugeMax;
          if (at < kHugeMax) {
            hbox[at] = rb[k];
            hidx[at] = j;
            hcx[at] = r.fx0;
            hcy[at] = r.fy0;
            fr[k] = kFrHuge | at;
          } else {
            fr[k] = pack_fine(r, gx, gy);
            mine += cells;
            for (uint32_t cy = r.y0; cy <= r.y1; ++cy)
              for (uint32_t cx = r.x0; cx <= r.x1; ++cx) atomicAdd(&L.cnt[cy * Kx + cx], 1u);
          }
        }
      }
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) mine += __shfl_xor(mine, d, 64);
      if (lane == 0) atomicAdd(&L.nfine, mine);
      __syncthreads();
      const uint32_t nfine = L.nfine;
      const bool multi = nfine > kMaxFine;  // more than one strip
      // B. each cell's count kept (ccount); with more than one strip, the
      //    row-major exclusive scan of the counts (start[]) places the
      //    strips' boundaries (kk <= 4096: 4 cells per thread)
      {
        uint32_t v[4], acc = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t c = tid * 4 + k;
          v[k] = c < kk ? L.cnt[c] : 0u;
          if (c < kk) L.ccount[c] = (uint16_t)v[k];
          acc += v[k];
        }
        if (multi) {
          const uint32_t inc = wave_prefix_sum(acc);
          if (lane == 63) L.wsum[w] = inc;
          __syncthreads();
          uint32_t ex = inc - acc;
          ex += __builtin_amdgcn_readlane(row_prefix_sum(lane < w ? L.wsum[lane] : 0u), 15);  // waves before mine
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t c = tid * 4 + k;
            if (c < kk) L.start[c] = ex;
            ex += v[k];
          }
          if (tid == kTestThreads - 1) L.start[kk] = ex;
        }
        __syncthreads();
        if (tid == 0) L.nfine = 0u;  // (every thread has read it)
      }
      // L. the listed large boxes against every coarse membership of the bin
      //    (each lane its registers' boxes; a listed box against the entries
      //    after its own): a pair is the bin's when its reference cell R =
      //    (max of the min-corner cells) is in the bin, that is, not west
      //    or south of it (both boxes reach the bin, so R is never east or
      //    north of it).
      const uint32_t nh = L.nhuge < kHugeMax ? L.nhuge : kHugeMax;
      if (nh) {
        uint32_t afx[kRegs], afy[kRegs], hfrom[kRegs];
#pragma unroll
        for (uint32_t k = 0; k < kRegs; ++k) {
          const uint32_t j = tid + k * kTestThreads;
          afx[k] = fine_of(rb[k].x, p.inv_f, p.nf);
          afy[k] = fine_of(rb[k].z, p.inv_f, p.nf);
          hfrom[k] = j >= m ? nh : ((fr[k] & kFrHuge) ? (fr[k] & ~kFrHuge) + 1u : 0u);
        }
        float4 xn = hbox[0];
        uint32_t hxn = hcx[0], hyn = hcy[0], hjn = hidx[0];
        for (uint32_t h = 0; h < nh; ++h) {
          const float4 x = xn;
          const uint32_t hx = hxn, hy = hyn, hj = hjn;
          if (h + 1u < nh) {  // the next listed box, in flight during these tests
            xn = hbox[h + 1u];
            hxn = hcx[h + 1u];
            hyn = hcy[h + 1u];
            hjn = hidx[h + 1u];
          }
#pragma unroll
          for (uint32_t k = 0; k < kRegs; ++k) {
            const float4 a = rb[k];
            const uint64_t mask = __ballot(h >= hfrom[k] && max(afx[k], hx) >= gx && max(afy[k], hy) >= gy &&
                                           a.x <= x.y && x.x <= a.y && a.z <= x.w && x.z <= a.w);
            if (mask) {
              const uint32_t pos = sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
              if (__builtin_amdgcn_inverse_ballot_w64(mask)) L.stage[w][pos] = make_uint2(tid + k * kTestThreads, hj);
              sn = __builtin_amdgcn_readfirstlane(sn + (uint32_t)__popcll(mask));
              if (sn > kBinStage - 64) stage_flush(L, w, sn, region, cap_wg);
            }
          }
        }
        __syncthreads();  // the list is the strip buffer's tail
        if (tid == 0) L.nhuge = 0u;
      }
      // C/D per strip
      // start[] is read-only from here on
      uint32_t c_lo = 0;
      while (c_lo < kk) {
        uint32_t c_hi = kk;  // usually the rest of the bin fits in one strip
        if (multi && L.start[kk] - L.start[c_lo] > kMaxFine) {
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
          c_hi = L.strip_hi;
        }
        if (c_hi == c_lo) {  // one fine cell alone exceeds the LDS buffer
          if (tid == 0) {
            // drop the bin's pairs so far: the host reruns the whole bin
            L.out_n = seg0;
            overfull[atomicAdd(&meta->max_per_cell, 1u)] = b | kOverfullCell;
            atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrOverfull);
          }
          for (uint32_t c = tid; c < kMaxCellsPerBin; c += kTestThreads) L.cnt[c] = 0u;
          __syncthreads();
          break;
        }
        // the strip's cells laid out by size class: class totals, their
        // exclusive scan (largest class first), then each cell takes its run
        // from its class's cursor (start[] of the strip's cells becomes the
        // strip position; the next strip's boundary reads only later cells)
        {
          uint32_t v[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t c = tid * 4 + k;
            v[k] = c >= c_lo && c < c_hi ? (uint32_t)L.ccount[c] : 0u;
            if (v[k]) atomicAdd(&L.csum[v[k] < kSizeClasses ? v[k] : kSizeClasses - 1u], v[k]);
          }
          __syncthreads();
          if (w == 0) {
            const uint32_t cls = kSizeClasses - 1u - lane;  // lane 0: the largest class
            const uint32_t own = L.csum[cls];
            const uint32_t inc = wave_prefix_sum(own);
            L.cstart[cls] = inc - own;
            L.csum[cls] = 0u;  // cleared for the next strip
            if (cls == 1u) L.npair = inc - own;
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t c = tid * 4 + k;
            if (v[k]) {
              const uint32_t st = atomicAdd(&L.cstart[v[k] < kSizeClasses ? v[k] : kSizeClasses - 1u], v[k]);
              L.start[c] = st;
              L.cnt[c] = st;  // scatter cursor
            }
          }
          __syncthreads();
        }
        const uint32_t base = 0u;
        const uint32_t len = L.npair;  // positions in cells of two or more memberships
#pragma unroll
        for (uint32_t k = 0; k < kRegs; ++k) {
          const uint32_t j = tid + k * kTestThreads;
          if (j < m && !(fr[k] & kFrHuge)) scatter_strip(L, rb[k], j, fr[k], Kx, c_lo, c_hi, base);
        }
        __syncthreads();
        const bool last_strip = c_hi == kk;
        if (last_strip)  // the cursors are dead: zero for the next bin
          for (uint32_t c = tid; c < kMaxCellsPerBin; c += kTestThreads) L.cnt[c] = 0u;
        if (last_strip && t + 1 < my_bins) {
#pragma unroll
          for (uint32_t k = 0; k < kRegs; ++k) {
            const uint32_t j = tid + k * kTestThreads < jN ? tid + k * kTestThreads : jN;
            const uint64_t at = sN + j < last ? sN + j : last;
            rb[k] = bbox[at];
            ri[k] = bid[at];
          }
          prefetched = true;
        }
        // D. test every pair of each cell of the strip, one position per lane
        for (uint32_t p0 = w * 64; p0 < len; p0 += kTestThreads) {
          const uint32_t pp = p0 + lane;
          uint32_t h = 0, cs = 0, cn = 1, aw = 0;
          float4 a = make_float4(0.f, -1.f, 0.f, -1.f);
          if (pp < len) {
            aw = L.word[pp];
            const uint32_t c = (aw >> kWordCellShift) & kWordCellMask;
            cs = L.start[c] - base;
            cn = L.ccount[c];
            const uint32_t r = pp - cs;
            // (n-1)/2 partners, plus one for the first half of an even cell
            h = (cn - 1 + (((cn & 1u) == 0u) & (r < (cn >> 1)))) >> 1;
            a = L.box[pp];
          }
          const uint32_t ja = aw & kWordIdxMask;
          const uint32_t wrap = cs + cn;
          // partners (r + k) mod n for k = 1..h, kUnroll at a time; all loads first
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
              // Overlap, and the reference-cell rule: one of the two boxes
              // starts in this column and one starts in this row (bits 31,
              // 30).  Each test is one compare into a wave mask.
              const uint64_t mask = __ballot(act[u]) & __ballot(a.x <= o[u].y) & __ballot(o[u].x <= a.y) &
                                    __ballot(a.z <= o[u].w) & __ballot(o[u].z <= a.w) &
                                    __ballot((aw | ow[u]) >= (kFlagX | kFlagY));
              if (mask) {
                const uint32_t pos = sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                if (__builtin_amdgcn_inverse_ballot_w64(mask)) L.stage[w][pos] = make_uint2(ja, ow[u] & kWordIdxMask);
                sn = __builtin_amdgcn_readfirstlane(sn + (uint32_t)__popcll(mask));
                if (sn > kBinStage - 64) stage_flush(L, w, sn, region, cap_wg);
              }
            }
          }
        }
        // Every strip's pairs are in the region before the barrier, so each
        // bin's pairs form one run of the region.
        if (sn) stage_flush(L, w, sn, region, cap_wg);
        __syncthreads();  // the strip buffer is reused
        c_lo = c_hi;
      }
    } else if (m > kMaxCoarse) {
      if (tid == 0) {
        overfull[atomicAdd(&meta->max_per_cell, 1u)] = b;
        atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrOverfull);
      }
    }
    // The bin's pairs were all appended before the last barrier; the next
    // bin appends only after further barriers.
    if (tid == 0) {
      const uint32_t e = L.out_n;
      bin_seg[b] = make_uint2(seg0, e - seg0);
      seg0 = e;
    }
    if (!prefetched && t + 1 < my_bins) {
      __syncthreads();
#pragma unroll
      for (uint32_t k = 0; k < kRegs; ++k) {
        const uint32_t j = tid + k * kTestThreads < jN ? tid + k * kTestThreads : jN;
        const uint64_t at = sN + j < last ? sN + j : last;
        rb[k] = bbox[at];
        ri[k] = bid[at];
      }
    }
  }
  if (tid == 0 && seg0 > cap_wg) {
    atomicMax((unsigned long long*)&meta->items, (unsigned long long)seg0);
    atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrWorkspace);
  }
  if (out_count) {
    // A single instance needs no bin order: the region (every bin's pairs,
    // dropped ones rewound) goes to a range of the output reserved with one
    // device atomic on the pair counter, which ends as the total.
    __syncthreads();
    const uint32_t n = L.out_n < cap_wg ? L.out_n : cap_wg;
    if (tid == 0) L.out_base = L.out_n ? atomicAdd(out_count, (unsigned long long)L.out_n) : 0ull;
    __syncthreads();
    const uint64_t base = L.out_base;
    for (uint32_t k = tid; k < n; k += kTestThreads)
      if (base + k < cap) out[base + k] = region[k];
  }
  // (after the pair counter's atomic, which the copy above waits for)
  if ((single & kSingleSum) && tid == 0 && L.memb) atomicAdd(&meta->memberships, L.memb);
}

// ---------------------------------------------------------------------------
// Test of a launch whose boxes all span at most two fine cells per axis.
//
// Then two overlapping boxes a, b have min-corner cells A, B at most one
// apart per axis (their cell ranges, of length <= 2, intersect), and the
// reference cell of the pair, R = (max(A.x, B.x), max(A.y, B.y)), is covered
// by both boxes; so the bin holding R holds both.  Each coarse membership is
// keyed once, by its min-corner cell in the bin's grid extended by one halo
// column and row (x, y in [-1, k)); no fine ranges, no replication.  A box in
// cell C = (x, y) tests
//   * if C is inside the bin: the boxes before it in C, and all boxes of W =
//     (x-1, y), SW = (x-1, y-1), S = (x, y-1)      (R = C);
//   * if (x+1, y) is inside the bin: the boxes of SE = (x+1, y-1)
//                                                   (R = (x+1, y)).
// Case by case (B = A, W, SW, S or SE of A, and the mirror cases), every
// overlapping pair with R in this bin is tested exactly once, and only
// there: the overlap test alone decides, with no reference-cell flags.  W's
// positions end where C's begin, and SW, S and SE are adjacent cells of one
// row, so a box's partners are two contiguous position ranges.  The partner
// loop reads only the partner's box; a hit stages the two positions, whose
// ids the flush looks up.
constexpr uint32_t kMaxExt = (kMaxK + 1) * (kMaxK + 1);  // extended cells of a bin
// Coarse memberships of one bin under this schedule: kNarrowMax (a position
// costs 22 bytes of LDS here), kNarrowRegs per thread in registers.
constexpr uint32_t kNarrowRegs = kNarrowMax / kTestThreads;
static_assert(kNarrowMax <= 65535u, "positions fit 16 bits");

struct NarrowLds {
  float4 box[kNarrowMax];          // the bin's coarse memberships in cell order
  uint16_t cell[kNarrowMax];       // extended cell of each position
  uint32_t id[kNarrowMax];         // box id of each position
  uint32_t cnt[kMaxExt + 1];       // counts (each membership's rank in its cell is the count before it);
                                   // the wide boxes count as one more cell after the extended ones
  uint16_t start[kMaxExt + 2];     // first position of each extended cell (and of the wide boxes)
  uint2 stage[kTestWaves][kBinStage];
  uint2 binfo[kMaxBinsPerGroup];
  uint4 bingrid[kMaxBinsPerGroup];
  float bininv[kMaxBinsPerGroup];
  uint32_t wsum[kTestWaves];
  uint32_t dnext;                  // next chunk of positions to claim, past the first 16
  uint32_t wnext;                  // next wide box to claim
  uint32_t out_n;
  unsigned long long out_base;
  unsigned long long memb;         // kSingleSum: memberships of this workgroup's bins
};
static_assert(sizeof(NarrowLds) <= 160u * 1024u, "one test workgroup per CU");
static_assert(kMaxExt + 1u <= 5u * kTestThreads, "the cell scan takes five cells per thread");

__device__ __forceinline__ void test_narrow(unsigned char* lds_raw, const float4* __restrict__ bbox,
                                                                  const uint32_t* __restrict__ bid,
                                                                  const uint32_t* __restrict__ bin_start,
                                                                  const uint32_t* __restrict__ total,
                                                                  const InstDesc* __restrict__ inst,
                                                                  const uint32_t* __restrict__ bin_inst,
                                                                  uint32_t n_bins, uint64_t cap_bucket,
                                                                  uint2* __restrict__ scratch, uint32_t cap_wg,
                                                                  uint2* __restrict__ bin_seg,
                                                                  uint32_t* __restrict__ overfull,
                                                                  uint2* __restrict__ out, uint64_t cap,
                                                                  unsigned long long* out_count,
                                                                  Meta* __restrict__ meta, uint32_t single,
                                                                  const uint32_t* __restrict__ order) {
  NarrowLds& L = *reinterpret_cast<NarrowLds*>(lds_raw);
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint32_t w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, kept scalar
  const uint32_t my_bins = test_bin_count(n_bins, order);
  const unsigned long long err = meta->error;
  if (err & (kErrInvalidBox | kErrWorkspace | kErrDense | kErrBinCap)) {
    for (uint32_t t = tid; t < my_bins; t += kTestThreads) bin_seg[test_bin(t, order)] = make_uint2(0u, 0u);
    if (single) single_pass_dense_rule(total, n_bins, single, err, meta);
    return;
  }
  uint2* region = scratch + (size_t)blockIdx.x * cap_wg;
  uint32_t sn = 0;
  uint32_t seg0 = 0;
  for (uint32_t c = tid; c <= kMaxExt; c += kTestThreads) L.cnt[c] = 0u;  // zero at the start of every bin
  const bool any_wide = meta->wide != 0u;  // (the launch's count: the classification below is skipped without)
  if (tid == 0) {
    L.out_n = 0u;
    L.memb = 0ull;
  }

  const uint64_t last = cap_bucket - 1;
  float4 rb[kNarrowRegs];
  uint32_t ri[kNarrowRegs];
  for (uint32_t t = 0; t < my_bins; ++t) {
    const uint32_t slot = t % kMaxBinsPerGroup;
    if (slot == 0) {
      __syncthreads();
      for (uint32_t u = tid; u < kMaxBinsPerGroup && t + u < my_bins; u += kTestThreads) {
        const uint32_t gb = test_bin(t + u, order);
        const InstDesc I = inst[bin_inst[gb]];
        const uint32_t lb = gb - I.bin0;
        const uint32_t gx = I.fx0 + (lb % I.nbx) * I.k, gy = I.fy0 + (lb / I.nbx) * I.k;
        const uint32_t ex = I.fx0 + I.span_x - gx, ey = I.fy0 + I.span_y - gy;
        const uint32_t kx = I.k < ex ? I.k : ex, ky = I.k < ey ? I.k : ey;
        const uint32_t tt = total[gb];
        L.binfo[u] = make_uint2(bin_start[gb], tt);
        if ((single & kSingleSum) && tt) atomicAdd(&L.memb, (unsigned long long)tt);
        L.bingrid[u] = make_uint4(gx, gy, kx | (ky << 16), I.nf);
        L.bininv[u] = I.inv_f;
      }
      __syncthreads();
      if (t == 0) {
        const uint64_t s0 = L.binfo[0].x;
        const uint32_t j_hi = L.binfo[0].y ? L.binfo[0].y - 1u : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kNarrowRegs; ++k) {
          const uint32_t j = tid + k * kTestThreads < j_hi ? tid + k * kTestThreads : j_hi;
          const uint64_t at = s0 + j < last ? s0 + j : last;
          rb[k] = bbox[at];
          ri[k] = bid[at];
        }
      }
    }
    const uint32_t b = test_bin(t, order);
    const uint32_t bN = t + 1 < my_bins ? test_bin(t + 1, order) : 0u;
    const uint32_t m = L.binfo[slot].y;
    const uint64_t sN = t + 1 < my_bins ? bin_start[bN] : 0ull;
    const uint32_t mN = t + 1 < my_bins ? total[bN] : 0u;
    const uint32_t jN = mN ? mN - 1u : 0u;
    const uint4 bg = L.bingrid[slot];
    const int gx = (int)bg.x, gy = (int)bg.y;
    const uint32_t Kx = bg.z & 0xFFFFu, Ky = bg.z >> 16, ex = Kx + 1u, ne = ex * (Ky + 1u);
    const float inv_f = L.bininv[slot];
    const uint32_t nf = bg.w;
    bool prefetched = false;
    if (m >= 2 && m <= kNarrowMax) {
      // A. each membership's extended cell (a wide box: cell ne, after the
      //    extended cells), and its rank among the cell's memberships (the
      //    count before it)
      uint32_t ek[kNarrowRegs], rk[kNarrowRegs];
#pragma unroll
      for (uint32_t k = 0; k < kNarrowRegs; ++k) {
        const uint32_t j = tid + k * kTestThreads;
        ek[k] = 0u;
        rk[k] = 0u;
        if (j < m) {
          const uint32_t fx = fine_of(rb[k].x, inv_f, nf), fy = fine_of(rb[k].z, inv_f, nf);
          if (any_wide && (fine_of(rb[k].y, inv_f, nf) - fx > 1u || fine_of(rb[k].w, inv_f, nf) - fy > 1u)) {
            ek[k] = ne;
          } else {
            int rx = (int)fx - gx, ry = (int)fy - gy;
            rx = rx < -1 ? -1 : (rx > (int)Kx - 1 ? (int)Kx - 1 : rx);
            ry = ry < -1 ? -1 : (ry > (int)Ky - 1 ? (int)Ky - 1 : ry);
            ek[k] = (uint32_t)(rx + 1) + (uint32_t)(ry + 1) * ex;
          }
          rk[k] = atomicAdd(&L.cnt[ek[k]], 1u);
        }
      }
      __syncthreads();
      // B. exclusive scan of the counts (ne + 1 <= kMaxExt + 1: 5 per thread)
      {
        uint32_t v[5], acc = 0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const uint32_t c = tid * 5 + k;
          v[k] = c <= ne ? L.cnt[c] : 0u;
          acc += v[k];
        }
        const uint32_t inc = wave_prefix_sum(acc);
        if (lane == 63) L.wsum[w] = inc;
        __syncthreads();
        uint32_t exs = inc - acc;
        exs += __builtin_amdgcn_readlane(row_prefix_sum(lane < w ? L.wsum[lane] : 0u), 15);  // waves before mine
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const uint32_t c = tid * 5 + k;
          if (c <= ne) L.start[c] = (uint16_t)exs;
          exs += v[k];
        }
        if (tid == kTestThreads - 1) L.start[ne + 1u] = (uint16_t)exs;
        if (tid == 0) {
          L.dnext = 0u;
          L.wnext = 0u;
        }
      }
      __syncthreads();
      // C. one position per membership, in cell order: box, cell and id
#pragma unroll
      for (uint32_t k = 0; k < kNarrowRegs; ++k) {
        const uint32_t j = tid + k * kTestThreads;
        if (j < m) {
          const uint32_t pos = L.start[ek[k]] + rk[k];
          L.box[pos] = rb[k];
          L.cell[pos] = (uint16_t)ek[k];
          L.id[pos] = ri[k];
        }
      }
      __syncthreads();
      const uint32_t mn = L.start[ne];  // narrow positions; the wide boxes follow
      for (uint32_t c = tid; c <= ne; c += kTestThreads) L.cnt[c] = 0u;  // for the next bin
      if (t + 1 < my_bins) {
#pragma unroll
        for (uint32_t k = 0; k < kNarrowRegs; ++k) {
          const uint32_t j = tid + k * kTestThreads < jN ? tid + k * kTestThreads : jN;
          const uint64_t at = sN + j < last ? sN + j : last;
          rb[k] = bbox[at];
          ri[k] = bid[at];
        }
        prefetched = true;
      }
      // D. one position per lane: partners in W and before me in my cell
      //    (W's positions end where my cell's begin: one range), then
      //    SW..S(..SE) of the row below.  Each pair inside a cell is tested
      //    once, by its later member.
      //    Chunks of 64 positions: a wave's first by its index, the rest from
      //    a shared counter, so that waves whose chunks had fewer partners
      //    take more of them.
      for (uint32_t p0 = w * 64; p0 < mn;) {
        const uint32_t p = p0 + lane;
        uint32_t b1 = 0, n1 = 0, b3 = 0, n3 = 0;
        float4 a = make_float4(0.f, -1.f, 0.f, -1.f);
        if (p < mn) {
          a = L.box[p];
          const uint32_t e = L.cell[p];
          const uint32_t X = e % ex, Y = e / ex;
          const bool inside = X >= 1u && Y >= 1u;
          const bool se = Y >= 1u && X + 1u <= Kx;
          if (inside) {
            b1 = L.start[e - 1u];                  // W, then my cell up to me
            n1 = p - b1;
            b3 = L.start[e - ex - 1u];              // SW, S (and SE)
            n3 = L.start[e - ex + (se ? 2u : 1u)] - b3;
          } else if (se) {
            b3 = L.start[e - ex + 1u];              // SE only
            n3 = L.start[e - ex + 2u] - b3;
          }
        }
        const uint32_t nall = n1 + n3;
        // partner k is position k + o1 (k < n1) or k + o3; a lane past its
        // partners reads a clamped position and its bit is masked off
        const uint32_t o1 = b1, o3 = b3 - n1;
        auto partner = [&](uint32_t k) { return min(k + (k < n1 ? o1 : o3), kNarrowMax - 1u); };
        auto stage_hits = [&](uint64_t mask, uint32_t qq) {
          const uint32_t pos = sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
          // the ballot itself as the lane mask of the hits
          if (__builtin_amdgcn_inverse_ballot_w64(mask))
            L.stage[w][pos] = make_uint2(p, qq);  // positions; ids resolved at the flush
          sn = __builtin_amdgcn_readfirstlane(sn + (uint32_t)__popcll(mask));
          if (sn > kBinStage - 64) stage_flush(L, w, sn, region, cap_wg);
        };
        // two partners per iteration: two independent test chains per wave
        for (uint32_t k = 0;; k += 2) {
          const uint64_t act0 = __ballot(k < nall);
          if (!act0) break;
          const uint64_t act1 = __ballot(k + 1u < nall);
          const uint32_t q0 = partner(k), q1 = partner(k + 1u);
          const float4 x0 = L.box[q0];
          const float4 x1 = L.box[q1];
          const uint64_t m0 = act0 & __ballot(a.x <= x0.y) & __ballot(x0.x <= a.y) & __ballot(a.z <= x0.w) &
                              __ballot(x0.z <= a.w);
          const uint64_t m1 = act1 & __ballot(a.x <= x1.y) & __ballot(x1.x <= a.y) & __ballot(a.z <= x1.w) &
                              __ballot(x1.z <= a.w);
          if (m0) stage_hits(m0, q0);
          if (m1) stage_hits(m1, q1);
        }
        uint32_t claim = 0;
        if (lane == 0) claim = atomicAdd(&L.dnext, 64u);
        p0 = kTestThreads + __builtin_amdgcn_readfirstlane(claim);
      }
      // E. the wide boxes (positions mn..m-1), claimed one at a time by a
      //    wave.  A narrow box B overlapping a wide box A has its min-corner
      //    cell in A's cells or one column west or one row south of them, and
      //    the pair's R = (max of the min-corner cells) is inside the bin
      //    exactly when B's column is >= the bin's first column unless A's
      //    is, and likewise for rows: in each of those rows of the extended
      //    grid the partners are one contiguous run of positions, a lane per
      //    partner.  Then the later wide boxes, R checked.
      if (m > mn) {
        const uint32_t nwb = m - mn;
        for (;;) {
          uint32_t claim = 0;
          if (lane == 0) claim = atomicAdd(&L.wnext, 1u);
          const uint32_t i = __builtin_amdgcn_readfirstlane(claim);
          if (i >= nwb) break;
          const uint32_t pw = mn + i;
          const float4 a = L.box[pw];
          const int ax0 = (int)fine_of(a.x, inv_f, nf), ax1 = (int)fine_of(a.y, inv_f, nf);
          const int ay0 = (int)fine_of(a.z, inv_f, nf), ay1 = (int)fine_of(a.w, inv_f, nf);
          auto stage_pairs = [&](uint64_t mask, uint32_t qq) {
            const uint32_t pos = sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            if (__builtin_amdgcn_inverse_ballot_w64(mask)) L.stage[w][pos] = make_uint2(pw, qq);
            sn = __builtin_amdgcn_readfirstlane(sn + (uint32_t)__popcll(mask));
            if (sn > kBinStage - 64) stage_flush(L, w, sn, region, cap_wg);
          };
          // extended columns X0..X1 and rows Y0..Y1 (bin column c is extended column c + 1)
          const int X0 = (ax0 >= gx ? ax0 - 1 : gx) - gx + 1, X1 = min(ax1, gx + (int)Kx - 1) - gx + 1;
          const int Y0 = (ay0 >= gy ? ay0 - 1 : gy) - gy + 1, Y1 = min(ay1, gy + (int)Ky - 1) - gy + 1;
          if (X0 <= X1 && Y0 <= Y1) {
            for (int Y = Y0; Y <= Y1; ++Y) {
              const uint32_t e0 = (uint32_t)Y * ex;
              const uint32_t s0 = L.start[e0 + (uint32_t)X0], s1 = L.start[e0 + (uint32_t)X1 + 1u];
              for (uint32_t q0 = s0; q0 < s1; q0 += 64) {
                const uint32_t q = q0 + lane < s1 ? q0 + lane : s1 - 1u;
                const float4 x = L.box[q];
                const uint64_t hit =
                    __ballot(q0 + lane < s1 && a.x <= x.y && x.x <= a.y && a.z <= x.w && x.z <= a.w);
                if (hit) stage_pairs(hit, q);
              }
            }
          }
          for (uint32_t j0 = i + 1u; j0 < nwb; j0 += 64) {
            const uint32_t j = j0 + lane;
            const uint32_t q = mn + (j < nwb ? j : nwb - 1u);
            const float4 x = L.box[q];
            const bool r_in = max(ax0, (int)fine_of(x.x, inv_f, nf)) >= gx && max(ay0, (int)fine_of(x.z, inv_f, nf)) >= gy;
            const uint64_t hit =
                __ballot(j < nwb && r_in && a.x <= x.y && x.x <= a.y && a.z <= x.w && x.z <= a.w);
            if (hit) stage_pairs(hit, q);
          }
        }
      }
      if (sn) stage_flush(L, w, sn, region, cap_wg);
      __syncthreads();  // the positions are reused by the next bin
    } else if (m > kNarrowMax) {
      if (tid == 0) {
        overfull[atomicAdd(&meta->max_per_cell, 1u)] = b;
        atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrOverfull);
      }
    }
    if (tid == 0) {
      const uint32_t e = L.out_n;
      bin_seg[b] = make_uint2(seg0, e - seg0);
      seg0 = e;
    }
    if (!prefetched && t + 1 < my_bins) {
      __syncthreads();
#pragma unroll
      for (uint32_t k = 0; k < kNarrowRegs; ++k) {
        const uint32_t j = tid + k * kTestThreads < jN ? tid + k * kTestThreads : jN;
        const uint64_t at = sN + j < last ? sN + j : last;
        rb[k] = bbox[at];
        ri[k] = bid[at];
      }
    }
  }
  if (tid == 0 && seg0 > cap_wg) {
    atomicMax((unsigned long long*)&meta->items, (unsigned long long)seg0);
    atomicOr((unsigned long long*)&meta->error, (unsigned long long)kErrWorkspace);
  }
  if (out_count) {
    __syncthreads();
    const uint32_t n = L.out_n < cap_wg ? L.out_n : cap_wg;
    if (tid == 0) L.out_base = L.out_n ? atomicAdd(out_count, (unsigned long long)L.out_n) : 0ull;
    __syncthreads();
    const uint64_t base = L.out_base;
    for (uint32_t k = tid; k < n; k += kTestThreads)
      if (base + k < cap) out[base + k] = region[k];
  }
  // (after the pair counter's atomic, which the copy above waits for)
  if ((single & kSingleSum) && tid == 0 && L.memb) atomicAdd(&meta->memberships, L.memb);
}

// One launch for both schedules: the histogram's wide flag, uniform over the
// launch, picks the narrow-box schedule or the general one, so no launch is
// spent on a schedule that has nothing to do.
constexpr size_t kTestLdsBytes = sizeof(TestLds) > sizeof(NarrowLds) ? sizeof(TestLds) : sizeof(NarrowLds);

template <bool kDev>
__global__ __launch_bounds__(kTestThreads) void k_bin_test(const float4* __restrict__ bbox,
                                                           const uint32_t* __restrict__ bid,
                                                           const uint32_t* __restrict__ bin_start,
                                                           const uint32_t* __restrict__ total,
                                                           const InstDesc* __restrict__ inst,
                                                           const uint32_t* __restrict__ bin_inst, uint32_t n_bins,
                                                           uint64_t cap_bucket, uint2* __restrict__ scratch,
                                                           uint32_t cap_wg, uint2* __restrict__ bin_seg,
                                                           uint32_t* __restrict__ overfull,
                                                           uint2* __restrict__ out, uint64_t cap, uint32_t direct,
                                                           Meta* __restrict__ meta, uint32_t single,
                                                           const uint32_t* __restrict__ order) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  // direct output: the launch's pair counter (a single instance's own, or
  // the one its device-planned levels share)
  unsigned long long* const out_count =
      direct ? (kDev ? meta->out_count : (unsigned long long*)&meta->n_pairs) : nullptr;
  if (kDev) n_bins = meta->plan_err ? 0u : meta->dev_bins;
  if (general_schedule(meta))
    test_general(lds_raw, bbox, bid, bin_start, total, inst, bin_inst, n_bins, cap_bucket, scratch, cap_wg, bin_seg,
                 overfull, out, cap, out_count, meta, single, order);
  else
    test_narrow(lds_raw, bbox, bid, bin_start, total, inst, bin_inst, n_bins, cap_bucket, scratch, cap_wg, bin_seg,
                overfull, out, cap, out_count, meta, single, order);
}

// The pair count and error word, copied to the caller's two result words
// after the last kernel of a call.
__global__ __launch_bounds__(64) void k_bin_result(const Meta* __restrict__ meta, uint64_t* __restrict__ result) {
  if (threadIdx.x == 0) {
    result[0] = meta->n_pairs;
    result[1] = meta->error;
  }
}

// Output position of every bin's pairs: exclusive scan of the bins' pair
// counts by one 1024-thread workgroup (bin_out[n_bins] = the batch's total).
// Pairs are laid out in bin order, so each instance's pairs are one run.
__global__ __launch_bounds__(1024) void k_bin_pair_scan(const uint2* __restrict__ bin_seg, uint32_t n_bins,
                                                        uint64_t* __restrict__ bin_out, uint64_t extra_total,
                                                        const uint64_t* __restrict__ dev_extra,
                                                        Meta* __restrict__ meta) {
  __shared__ ScanLds L;
  const uint64_t grand = block_exscan(
      n_bins, L, [&](uint32_t i) { return (uint64_t)bin_seg[i].y; }, [&](uint32_t i, uint64_t ex) { bin_out[i] = ex; });
  if (threadIdx.x == 0) {
    bin_out[n_bins] = grand;
    // the true count, also past the capacity
    meta->n_pairs = grand + extra_total + (dev_extra ? *dev_extra : 0ull);
  }
}

// Copy each bin's run of its workgroup's region to the bin's output position
// (one block per bin), and write every instance's (offset, count).
__global__ __launch_bounds__(256) void k_bin_compact(const uint2* __restrict__ scratch, uint32_t cap_wg,
                                                      uint32_t nwg, const uint2* __restrict__ bin_seg,
                                                      const uint64_t* __restrict__ bin_out, uint32_t n_bins,
                                                      const InstDesc* __restrict__ inst,
                                                      const uint32_t* __restrict__ bin_inst, uint32_t n_inst,
                                                      uint2* __restrict__ out, uint64_t cap,
                                                      uint64_t* __restrict__ inst_out,
                                                      const uint64_t* __restrict__ merge,
                                                      const uint32_t* __restrict__ rank) {
  if (inst_out) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n_inst; i += gridDim.x * 256) {
      const InstDesc I = inst[i];
      const uint64_t o0 = bin_out[I.bin0];
      inst_out[2 * (uint64_t)i] = o0 + I.pairs0 + (merge ? merge[n_inst + i] : 0ull);
      inst_out[2 * (uint64_t)i + 1] = bin_out[I.bin0 + I.nb] - o0 + I.pairs_extra + (merge ? merge[i] : 0ull);
    }
  }
  const uint32_t gb = blockIdx.x;
  if (gb >= n_bins) return;
  const uint2 seg = bin_seg[gb];
  const uint32_t gi = bin_inst[gb];
  const uint64_t dst0 = bin_out[gb] + inst[gi].pairs0 + (merge ? merge[n_inst + gi] : 0ull);
  const uint2* src = scratch + (size_t)test_wg(gb, nwg, rank) * cap_wg;
  // a region that overflowed holds only its first cap_wg pairs (the host
  // regrows and runs again)
  const uint32_t n = seg.x < cap_wg ? (seg.y < cap_wg - seg.x ? seg.y : cap_wg - seg.x) : 0u;
  for (uint32_t k = threadIdx.x; k < n; k += 256) {
    const uint64_t d = dst0 + k;
    if (d < cap) out[d] = src[seg.x + k];
  }
}

// ---------------------------------------------------------------------------
// Re-binning levels planned on the device.
//
// After a launch of one instance whose test kernel listed bins beyond LDS,
// each further level turns the list of the level above into virtual
// instances (plan_virtual_bin: the bin's region, finer cells, smaller bins)
// and runs the histogram, scans, scatter and test over them, reading the
// boxes and ids from the level above's bucket; its test kernel writes its
// pairs straight to the output through the launch's shared pair counter and
// lists its own bins beyond LDS for the next level.  No host work between
// levels: the grids are upper bounds and every count comes from the level's
// result words, written by its plan kernel (one workgroup).

constexpr uint32_t kDeepPlanLds = (3u * kDeepMaxInst + 2u) * sizeof(uint32_t);

__global__ __launch_bounds__(1024) void k_deep_plan(const InstDesc* __restrict__ pinst,
                                                    const uint32_t* __restrict__ pbin_inst,
                                                    const uint32_t* __restrict__ ptotal,
                                                    const uint32_t* __restrict__ pstart,
                                                    const uint32_t* __restrict__ poverfull, uint32_t p_bins,
                                                    uint32_t p_dev, const Meta* __restrict__ pmeta, uint32_t target,
                                                    DeepCaps caps, float root_inv_f,
                                                    const InstDesc* __restrict__ root_inst, uint32_t max_refine,
                                                    InstDesc* __restrict__ inst, uint32_t* __restrict__ chunk_inst,
                                                    uint32_t* __restrict__ bin_inst, Meta* __restrict__ meta,
                                                    uint32_t parent_level, uint32_t* __restrict__ dense_list,
                                                    Meta* __restrict__ dense) {
  __shared__ ScanLds L;
  __shared__ uint32_t nbmax;
  // dynamic LDS: first chunk and first bin of every instance (and the
  // totals after them), and the sorted list
  extern __shared__ __align__(16) unsigned char plan_lds[];
  uint32_t* const c0 = reinterpret_cast<uint32_t*>(plan_lds);
  uint32_t* const b0 = c0 + (kDeepMaxInst + 1);
  uint32_t* const ord = b0 + (kDeepMaxInst + 1);
  const uint32_t tid = threadIdx.x;
  // (a level planned on the host has no plan error word of its own)
  uint32_t n_over = p_dev && pmeta->plan_err ? 0u : pmeta->max_per_cell;
  n_over = n_over < p_bins ? n_over : p_bins;
  if (tid == 0) nbmax = 0u;
  __syncthreads();
  const uint32_t n_inst = n_over <= caps.inst && n_over <= kDeepMaxInst ? n_over : 0u;
  // 0. the listed bins in increasing order: the test kernel appends them in
  //    the order its workgroups finish, and the plan (so every bin's
  //    workgroup and every workgroup's pair count) must be the same in each
  //    launch of the same input.  Bitonic sort in LDS, padded to a power of
  //    two with entries that sort last.
  const uint32_t np2 = n_inst <= 1u ? 1u : 1u << (32 - __builtin_clz(n_inst - 1u));
  for (uint32_t i = tid; i < np2; i += 1024) ord[i] = i < n_inst ? poverfull[i] : 0xFFFFFFFFu;
  __syncthreads();
  for (uint32_t size = 2; size <= np2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = tid; t < np2 / 2u; t += 1024) {
        const uint32_t lo = 2u * t - (t & (stride - 1u)), hi = lo + stride;
        const uint32_t a = ord[lo], b = ord[hi];
        if ((a > b) == ((lo & size) == 0u)) {
          ord[lo] = b;
          ord[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  // 1. every listed bin's virtual instance
  for (uint32_t i = tid; i < n_inst; i += 1024) {
    const uint32_t b = ord[i] & ~kOverfullCell;
    InstDesc v{};
    if (b < p_bins) {
      const uint32_t pi = pbin_inst[b];
      const InstDesc P = pinst[pi];
      const uint32_t root = p_dev ? P.root : pi;
      if (!plan_virtual_bin(P, pi, b - P.bin0, ptotal[b], pstart[b], target, caps.nb_max, &v,
                            root_inst ? root_inst[root].inv_f : root_inv_f, max_refine)) {
        // no finer bins: the parent's bin is tested whole (dense kernel)
        v = InstDesc{};
        const uint32_t at = atomicAdd(&dense->dev_inst, 1u);
        if (at < kDenseMaxBins) dense_list[at] = (parent_level << 30) | b;
      }
      v.root = root;
    }
    inst[i] = v;
    atomicMax(&nbmax, v.n ? v.nb : 0u);
  }
  __syncthreads();
  // 2. layout: first chunk, bin, histogram and offset row of each instance
  const uint64_t chunks = block_exscan(
      n_inst, L, [&](uint32_t i) { return (uint64_t)inst[i].nchunks; },
      [&](uint32_t i, uint64_t ex) { inst[i].chunk0 = (uint32_t)ex; c0[i] = (uint32_t)ex; });
  const uint64_t bins = block_exscan(
      n_inst, L, [&](uint32_t i) { return (uint64_t)(inst[i].n ? inst[i].nb : 0u); },
      [&](uint32_t i, uint64_t ex) { inst[i].bin0 = (uint32_t)ex; b0[i] = (uint32_t)ex; });
  const uint64_t hist = block_exscan(
      n_inst, L, [&](uint32_t i) { return (uint64_t)inst[i].nchunks * hist_pitch(inst[i].nb); },
      [&](uint32_t i, uint64_t ex) { inst[i].hist0 = ex; });
  const uint64_t off = block_exscan(
      n_inst, L, [&](uint32_t i) { return (uint64_t)inst[i].nchunks * off_pitch(inst[i].nb); },
      [&](uint32_t i, uint64_t ex) { inst[i].off0 = ex; });
  if (tid == 0) {
    c0[n_inst] = (uint32_t)chunks;
    b0[n_inst] = (uint32_t)bins;
  }
  __syncthreads();
  const bool fits = n_inst == n_over && chunks <= caps.chunks && bins <= caps.bins && hist <= caps.hist &&
                    off <= caps.off && nbmax <= caps.nb_max;
  // 3. chunk -> instance and bin -> instance: a wave per instance writes
  //    its ranges
  if (fits) {
    const uint32_t wv = tid >> 6, ln = tid & 63u;
    for (uint32_t i = wv; i < n_inst; i += 1024u / 64u) {
      for (uint32_t c = c0[i] + ln; c < c0[i + 1u]; c += 64u) chunk_inst[c] = i;
      for (uint32_t b = b0[i] + ln; b < b0[i + 1u]; b += 64u) bin_inst[b] = i;
    }
  }
  if (tid == 0) {
    meta->dev_inst = n_over;
    meta->dev_chunks = (uint32_t)chunks;
    meta->dev_bins = (uint32_t)bins;
    meta->dev_nbmax = nbmax;
    meta->dev_hist = hist;
    meta->dev_off = off;
    meta->plan_err = fits ? 0u : 1u;
  }
}

// The result words of a launch with device-planned levels: the pair count
// (level 0's: the shared counter of one instance, or a batch's merged total),
// and an error word: an invalid box, a workspace that did not fit at any
// level (a level's plan, bucket or pair regions, or the dense kernel's
// buffers), or more bins left to the dense kernel than it takes.
__global__ __launch_bounds__(64) void k_deep_result(DeepMetas m, const Meta* __restrict__ dense,
                                                    uint64_t* __restrict__ result) {
  if (threadIdx.x != 0) return;
  unsigned long long e = 0ull;
  for (uint32_t l = 0; l < m.levels; ++l) {
    const Meta* x = m.meta[l];
    e |= x->error & (kErrInvalidBox | kErrWorkspace | kErrTooDense | kErrDense);
    if (l == 0) continue;  // level 0 is planned on the host
    if (x->plan_err & 1u) e |= kErrWorkspace;
  }
  e |= dense->error & (kErrWorkspace | kErrOverfull);
  result[0] = m.meta[0]->n_pairs;
  result[1] = e;
}

hipError_t launch_deep_level(const BinnedWorkspace& parent, uint32_t parent_bins, bool parent_dev,
                             const BinnedWorkspace& ws, const DeepCaps& caps, uint32_t target, float root_inv_f,
                             const InstDesc* root_inst, uint32_t max_refine, uint32_t n_cu, uint32_t direct,
                             uint2* out, uint64_t cap, uint32_t parent_level, uint32_t* dense_list, Meta* dense,
                             hipStream_t s) {
  k_deep_plan<<<1, 1024, kDeepPlanLds, s>>>(parent.inst, parent.bin_inst, parent.total, parent.bin_start, parent.overfull,
                                 parent_bins, parent_dev ? 1u : 0u, parent.meta, target, caps, root_inv_f, root_inst,
                                 max_refine, ws.inst, ws.chunk_inst, ws.bin_inst, ws.meta, parent_level, dense_list,
                                 dense);
  const uint32_t nc = caps.chunks ? caps.chunks : 1u, nbins = caps.bins ? caps.bins : 1u;
  k_bin_hist<true><<<nc, kHistThreads, 0, s>>>(parent.bbox, ws.inst, ws.chunk_inst, 0u, ws.hist, ws.meta);
  k_bin_scan<true><<<(nbins + kScanBins - 1) / kScanBins + 1, 1024, 0, s>>>(
      ws.hist, ws.inst, ws.bin_inst, 0u, ws.chunk_inst, 0u, ws.off, ws.total, ws.meta);
  k_bin_starts<true><<<1, 1024, 0, s>>>(ws.total, 0u, ws.bin_start, ws.cap_bucket, 0u, ws.meta, ws.order, ws.rank);
  const uint32_t per_xcd = (nc + 7u) / 8u < n_cu / 8u ? (nc + 7u) / 8u : (n_cu >= 8u ? n_cu / 8u : 1u);
  const size_t scatter_lds = sizeof(ScatterLds) + 8u * (caps.nb_max ? caps.nb_max : 1u);
  k_bin_scatter<true, true><<<per_xcd * 8u, kScatterThreads, scatter_lds, s>>>(
      parent.bbox, parent.bid, ws.inst, ws.chunk_inst, 0u, ws.hist, ws.off, ws.bin_start, ws.bbox, ws.bid,
      ws.cap_bucket, caps.nb_max ? caps.nb_max : 1u, ws.total, 0u, 0u, ws.meta);
  k_bin_test<true><<<n_cu ? n_cu : 1u, kTestThreads, kTestLdsBytes, s>>>(
      ws.bbox, ws.bid, ws.bin_start, ws.total, ws.inst, ws.bin_inst, 0u, ws.cap_bucket, ws.scratch, ws.cap_wg,
      ws.bin_seg, ws.overfull, out, cap, direct, ws.meta, 0u, ws.order);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batch merge of the device-planned levels' pairs (see launch_level_count
// and the others in the header): per level-0 instance, the pairs its levels
// found (by the root of each level bin), their exclusive scan (the shift of
// the instance's run: the level-0 compaction leaves room after each
// instance's own pairs) and a cursor that hands each level bin its place in
// that room.

__global__ __launch_bounds__(256) void k_merge_clear(uint64_t* __restrict__ merge, uint32_t n_inst) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n_inst; i += gridDim.x * 256) {
    merge[i] = 0ull;
    merge[2ull * n_inst + i] = 0ull;
  }
}

// Pairs of a level bin's run (a region that overflowed holds only its first
// cap_wg pairs; the result words then ask for regrowth).
__device__ __forceinline__ uint32_t region_run(uint2 seg, uint32_t cap_wg) {
  return seg.x < cap_wg ? (seg.y < cap_wg - seg.x ? seg.y : cap_wg - seg.x) : 0u;
}

__global__ __launch_bounds__(256) void k_level_count(const uint2* __restrict__ bin_seg,
                                                     const InstDesc* __restrict__ inst,
                                                     const uint32_t* __restrict__ bin_inst, uint32_t cap_wg,
                                                     const Meta* __restrict__ meta, uint64_t* __restrict__ merge) {
  const uint32_t n_bins = meta->plan_err ? 0u : meta->dev_bins;
  for (uint32_t b = blockIdx.x * 256 + threadIdx.x; b < n_bins; b += gridDim.x * 256) {
    const uint32_t c = region_run(bin_seg[b], cap_wg);
    if (c) atomicAdd((unsigned long long*)&merge[inst[bin_inst[b]].root], (unsigned long long)c);
  }
}

__global__ __launch_bounds__(1024) void k_merge_scan(uint64_t* __restrict__ merge, uint32_t n_inst) {
  __shared__ ScanLds L;
  const uint64_t total = block_exscan(
      n_inst, L, [&](uint32_t i) { return merge[i]; }, [&](uint32_t i, uint64_t ex) { merge[n_inst + i] = ex; });
  if (threadIdx.x == 0) merge[3ull * n_inst] = total;
}

// A block per level bin in turn: the bin's run goes after its root's level-0
// pairs (and after the runs of the root's level bins placed before it).
__global__ __launch_bounds__(256) void k_level_copy(const uint2* __restrict__ scratch, uint32_t cap_wg, uint32_t nwg,
                                                    const uint2* __restrict__ bin_seg,
                                                    const InstDesc* __restrict__ inst,
                                                    const uint32_t* __restrict__ bin_inst,
                                                    const Meta* __restrict__ meta,
                                                    const InstDesc* __restrict__ inst0,
                                                    const uint64_t* __restrict__ bin_out0,
                                                    uint64_t* __restrict__ merge, uint32_t n_inst,
                                                    uint2* __restrict__ out, uint64_t cap,
                                                    const uint32_t* __restrict__ rank) {
  __shared__ uint64_t dst0;
  const uint32_t n_bins = meta->plan_err ? 0u : meta->dev_bins;
  for (uint32_t gb = blockIdx.x; gb < n_bins; gb += gridDim.x) {
    const uint2 seg = bin_seg[gb];
    const uint32_t n = region_run(seg, cap_wg);
    if (!n) continue;
    __syncthreads();  // dst0 of the block's previous bin has been read
    if (threadIdx.x == 0) {
      const uint32_t r = inst[bin_inst[gb]].root;
      const InstDesc R = inst0[r];
      dst0 = bin_out0[R.bin0 + R.nb] + R.pairs0 + merge[n_inst + r] +
             atomicAdd((unsigned long long*)&merge[2ull * n_inst + r], (unsigned long long)n);
    }
    __syncthreads();
    const uint64_t d0 = dst0;
    const uint2* src = scratch + (size_t)test_wg(gb, nwg, rank) * cap_wg + seg.x;
    for (uint32_t k = threadIdx.x; k < n; k += 256) {
      const uint64_t d = d0 + k;
      if (d < cap) out[d] = src[k];
    }
  }
}

hipError_t launch_merge_clear(uint64_t* merge, uint32_t n_inst, hipStream_t s) {
  const uint32_t g = (n_inst + 255u) / 256u;
  k_merge_clear<<<g ? (g < 1024u ? g : 1024u) : 1u, 256, 0, s>>>(merge, n_inst);
  return hipGetLastError();
}

hipError_t launch_level_count(const BinnedWorkspace& lv, uint32_t max_bins, uint64_t* merge, hipStream_t s) {
  const uint32_t g = (max_bins + 255u) / 256u;
  k_level_count<<<g ? (g < 1024u ? g : 1024u) : 1u, 256, 0, s>>>(lv.bin_seg, lv.inst, lv.bin_inst, lv.cap_wg,
                                                                  lv.meta, merge);
  return hipGetLastError();
}

hipError_t launch_merge_scan(uint64_t* merge, uint32_t n_inst, hipStream_t s) {
  k_merge_scan<<<1, 1024, 0, s>>>(merge, n_inst);
  return hipGetLastError();
}

hipError_t launch_level_copy(const BinnedWorkspace& lv, uint32_t max_bins, uint32_t lv_test_blocks,
                             const BinnedWorkspace& ws0, const uint64_t* bin_out0, uint64_t* merge, uint32_t n_inst,
                             uint2* out, uint64_t cap, hipStream_t s) {
  // a persistent grid: each block walks level bins gb = blockIdx.x, + grid, ...
  const uint32_t grid = max_bins < 2048u ? (max_bins ? max_bins : 1u) : 2048u;
  k_level_copy<<<grid, 256, 0, s>>>(lv.scratch, lv.cap_wg, lv_test_blocks ? lv_test_blocks : 1u,
                                                        lv.bin_seg, lv.inst, lv.bin_inst, lv.meta, ws0.inst, bin_out0,
                                                        merge, n_inst, out, cap, lv.rank);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Bins no level resolves, tested whole (see DenseSrc in the header).
//
// Every workgroup gathers the listed bins (at most kDenseMaxBins) and the
// inclusive prefix of their tile counts in LDS; tiles are then taken
// round-robin over the whole list.  A bin of m memberships has
// t = ceil(m / kDenseTile) row tiles and t (t + 1) / 2 tiles (ti <= tj).  A
// tile stages its column memberships (box, id, first fine column and row)
// in LDS; each thread holds one row membership and tests it against every
// column (only later ones on a diagonal tile).  A pair is the bin's when
// its reference cell -- (max of the first columns, max of the first rows),
// the cell of the lower-left corner of the two boxes' intersection -- lies
// in the bin's cells: the rule the test kernel applies cell by cell.  Hits
// go through a per-wave LDS stage.

constexpr uint32_t kDenseThreads = kDenseTile;
constexpr uint32_t kDenseStage = 512;              // staged pairs per wave
constexpr uint32_t kDenseMaxTiles = 1u << 20;      // per bin (about 370,000 memberships)

__global__ __launch_bounds__(64) void k_dense_clear(Meta* __restrict__ dense) {
  if (threadIdx.x < sizeof(Meta) / sizeof(unsigned long long))
    reinterpret_cast<unsigned long long*>(dense)[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kDenseThreads) void k_dense_bins(DenseSrc src, uint32_t last_level,
                                                              Meta* __restrict__ dense,
                                                              unsigned long long* __restrict__ out_count,
                                                              uint2* __restrict__ out, uint64_t cap,
                                                              uint2* __restrict__ pairs, uint64_t cap_pairs,
                                                              DenseChunk* __restrict__ chunks, uint32_t cap_chunks) {
  __shared__ uint32_t ent[kDenseMaxBins];  // level << 30 | bin
  __shared__ uint32_t pre[kDenseMaxBins];  // tiles, then their inclusive prefix
  __shared__ float4 cb[kDenseTile];
  __shared__ uint32_t cid[kDenseTile], cfx[kDenseTile], cfy[kDenseTile];
  __shared__ uint2 stage[kDenseThreads / 64][kDenseStage];
  __shared__ uint32_t wsum[kDenseThreads / 64];
  __shared__ uint32_t too_big;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const Meta* lm = src.last_meta;
  const uint32_t n_plan = dense->dev_inst;
  const uint32_t n_last = lm->plan_err ? 0u : lm->max_per_cell;
  if ((uint64_t)n_plan + n_last > kDenseMaxBins) {
    if (blockIdx.x == 0 && tid == 0) atomicOr((unsigned long long*)&dense->error, (unsigned long long)kErrOverfull);
    return;
  }
  const uint32_t n = n_plan + n_last;
  if (n == 0) return;
  if (tid == 0) too_big = 0u;
  __syncthreads();
  // 1. the list and each bin's tile count; then the inclusive prefix (four
  //    consecutive entries per thread, wave scans, block carry)
  constexpr uint32_t kPer = kDenseMaxBins / kDenseThreads;
  uint32_t tl[kPer], acc = 0;
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t e = tid * kPer + k;
    tl[k] = 0u;
    if (e < n) {
      const uint32_t x = e < n_plan ? src.list[e] : ((last_level << 30) | (src.last_overfull[e - n_plan] & ~kOverfullCell));
      ent[e] = x;
      const uint32_t m = src.total[x >> 30][x & 0x3FFFFFFFu];
      const uint64_t t = (m + kDenseTile - 1u) / kDenseTile;
      const uint64_t tiles = t * (t + 1u) / 2u;
      if (tiles > kDenseMaxTiles) atomicOr(&too_big, 1u);
      tl[k] = tiles > kDenseMaxTiles ? 0u : (uint32_t)tiles;
    }
    acc += tl[k];
  }
  const uint32_t inc = wave_prefix_sum(acc);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (too_big) {  // (the same in every workgroup)
    if (blockIdx.x == 0 && tid == 0) atomicOr((unsigned long long*)&dense->error, (unsigned long long)kErrOverfull);
    return;
  }
  uint32_t run = inc - acc;
  for (uint32_t k = 0; k < w; ++k) run += wsum[k];
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    run += tl[k];
    if (tid * kPer + k < n) pre[tid * kPer + k] = run;
  }
  __syncthreads();
  const uint32_t total_tiles = pre[n - 1];
  uint32_t sn = 0;  // this wave's staged pairs
  uint32_t root = 0;
  auto flush = [&]() {
    uint64_t base = 0;
    bool ok = true;
    if (out_count) {
      if (lane == 0) base = atomicAdd(out_count, (unsigned long long)sn);
      base = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
             __builtin_amdgcn_readfirstlane((uint32_t)base);
      for (uint32_t k = lane; k < sn; k += 64)
        if (base + k < cap) out[base + k] = stage[w][k];
    } else {
      uint64_t ch = 0;
      if (lane == 0) {
        base = atomicAdd((unsigned long long*)&dense->n_pairs, (unsigned long long)sn);
        ch = atomicAdd((unsigned long long*)&dense->items, 1ull);
      }
      base = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
             __builtin_amdgcn_readfirstlane((uint32_t)base);
      ch = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ch >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)ch);
      ok = base + sn <= cap_pairs && ch < cap_chunks;
      if (ok) {
        for (uint32_t k = lane; k < sn; k += 64) pairs[base + k] = stage[w][k];
        if (lane == 0) {
          DenseChunk c;
          c.off = base;
          c.n = sn;
          c.root = root;
          chunks[ch] = c;
        }
      } else if (lane == 0) {
        atomicOr((unsigned long long*)&dense->error, (unsigned long long)kErrWorkspace);
      }
    }
    sn = 0;
  };
  // 2. the tiles
  for (uint32_t tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
    uint32_t lo = 0, hi = n - 1u;  // first entry whose prefix exceeds the tile
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pre[mid] > tile) hi = mid; else lo = mid + 1u;
    }
    const uint32_t x = ent[lo], lv = x >> 30, b = x & 0x3FFFFFFFu;
    const uint32_t u = tile - (lo ? pre[lo - 1u] : 0u);
    const uint32_t m = src.total[lv][b], s0 = src.bin_start[lv][b];
    const uint32_t t = (m + kDenseTile - 1u) / kDenseTile;
    // row ti holds t - ti tiles: S(ti) = ti t - ti (ti - 1) / 2 tiles before it
    auto S = [&](uint32_t r) { return (uint64_t)r * t - (uint64_t)r * (r - 1u) / 2u; };
    const double tt = 2.0 * t + 1.0;
    uint32_t ti = (uint32_t)fmax(0.0, floor((tt - sqrt(fmax(0.0, tt * tt - 8.0 * u))) / 2.0));
    while (ti > 0u && S(ti) > u) --ti;
    while (ti + 1u < t && S(ti + 1u) <= u) ++ti;
    const uint32_t tj = ti + (uint32_t)(u - S(ti));
    const uint32_t ii = src.bin_inst[lv][b];
    const InstDesc I = src.inst[lv][ii];
    root = lv ? I.root : ii;
    const uint32_t lb = b - I.bin0;
    const uint32_t gx = I.fx0 + (lb % I.nbx) * I.k, gy = I.fy0 + (lb / I.nbx) * I.k;
    const uint32_t ex = I.fx0 + I.span_x - gx, ey = I.fy0 + I.span_y - gy;
    const uint32_t kx = I.k < ex ? I.k : ex, ky = I.k < ey ? I.k : ey;
    const float4* bb = src.bbox[lv] + s0;
    const uint32_t* bi = src.bid[lv] + s0;
    const uint32_t j0 = tj * kDenseTile, nc = m - j0 < kDenseTile ? m - j0 : kDenseTile;
    if (tid < nc) {
      const float4 q = bb[j0 + tid];
      cb[tid] = q;
      cid[tid] = bi[j0 + tid];
      cfx[tid] = fine_of(q.x, I.inv_f, I.nf);
      cfy[tid] = fine_of(q.z, I.inv_f, I.nf);
    }
    const uint32_t i = ti * kDenseTile + tid;
    const bool row = i < m;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t ia = 0, fxa = 0, fya = 0;
    if (row) {
      a = bb[i];
      ia = bi[i];
      fxa = fine_of(a.x, I.inv_f, I.nf);
      fya = fine_of(a.z, I.inv_f, I.nf);
    }
    __syncthreads();
    const uint32_t jmin = ti == tj ? tid + 1u : 0u;  // a diagonal tile: later columns only
    for (uint32_t jj = 0; jj < nc; ++jj) {
      const float4 q = cb[jj];
      const uint32_t rx = fxa > cfx[jj] ? fxa : cfx[jj], ry = fya > cfy[jj] ? fya : cfy[jj];
      const bool hit = row && jj >= jmin && a.x <= q.y && q.x <= a.y && a.z <= q.w && q.z <= a.w &&
                       rx - gx < kx && ry - gy < ky;
      const uint64_t mask = __ballot(hit);
      if (hit) {
        const uint32_t pos = sn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        const uint32_t ib = cid[jj];
        stage[w][pos] = make_uint2(ia < ib ? ia : ib, ia < ib ? ib : ia);
      }
      sn += (uint32_t)__popcll(mask);
      if (sn > kDenseStage - 64u) flush();
    }
    if (sn) flush();
    __syncthreads();  // the columns are reloaded
  }
}

// Batch merge of the dense runs (after launch_dense_bins, whose error word
// voids the runs: the launch then reports it and nothing is merged).
__global__ __launch_bounds__(256) void k_dense_count(const DenseChunk* __restrict__ chunks, uint32_t cap_chunks,
                                                     const Meta* __restrict__ dense, uint64_t* __restrict__ merge,
                                                     uint32_t n_inst) {
  if (dense->error) return;
  const uint64_t nch = dense->items < cap_chunks ? dense->items : cap_chunks;
  for (uint64_t c = blockIdx.x * 256ull + threadIdx.x; c < nch; c += gridDim.x * 256ull) {
    const DenseChunk ch = chunks[c];
    if (ch.root < n_inst) atomicAdd((unsigned long long*)&merge[ch.root], (unsigned long long)ch.n);
  }
}

__global__ __launch_bounds__(256) void k_dense_copy(const DenseChunk* __restrict__ chunks, uint32_t cap_chunks,
                                                    const Meta* __restrict__ dense, const uint2* __restrict__ pairs,
                                                    const InstDesc* __restrict__ inst0,
                                                    const uint64_t* __restrict__ bin_out0,
                                                    uint64_t* __restrict__ merge, uint32_t n_inst,
                                                    uint2* __restrict__ out, uint64_t cap) {
  __shared__ uint64_t dst0;
  if (dense->error) return;
  const uint64_t nch = dense->items < cap_chunks ? dense->items : cap_chunks;
  for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const DenseChunk ch = chunks[c];
    if (ch.root >= n_inst) continue;  // (uniform over the block)
    if (threadIdx.x == 0) {
      const InstDesc R = inst0[ch.root];
      dst0 = bin_out0[R.bin0 + R.nb] + R.pairs0 + merge[n_inst + ch.root] +
             atomicAdd((unsigned long long*)&merge[2ull * n_inst + ch.root], (unsigned long long)ch.n);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < ch.n; k += 256) {
      const uint64_t d = dst0 + k;
      if (d < cap) out[d] = pairs[ch.off + k];
    }
    __syncthreads();  // dst0 is rewritten
  }
}

hipError_t launch_dense_clear(Meta* dense, hipStream_t s) {
  k_dense_clear<<<1, 64, 0, s>>>(dense);
  return hipGetLastError();
}

hipError_t launch_dense_bins(const DenseSrc& src, uint32_t last_level, Meta* dense, uint32_t n_cu,
                             unsigned long long* out_count, uint2* out, uint64_t cap, uint2* pairs,
                             uint64_t cap_pairs, DenseChunk* chunks, uint32_t cap_chunks, hipStream_t s) {
  const uint32_t grid = 4u * (n_cu ? n_cu : 1u);
  k_dense_bins<<<grid, kDenseThreads, 0, s>>>(src, last_level, dense, out_count, out, cap, pairs, cap_pairs, chunks,
                                             cap_chunks);
  return hipGetLastError();
}

hipError_t launch_dense_count(const DenseChunk* chunks, uint32_t cap_chunks, const Meta* dense, uint64_t* merge,
                              uint32_t n_inst, hipStream_t s) {
  const uint32_t g = (cap_chunks + 255u) / 256u;
  k_dense_count<<<g ? (g < 1024u ? g : 1024u) : 1u, 256, 0, s>>>(chunks, cap_chunks, dense, merge, n_inst);
  return hipGetLastError();
}

hipError_t launch_dense_copy(const DenseChunk* chunks, uint32_t cap_chunks, const Meta* dense, const uint2* pairs,
                             const BinnedWorkspace& ws0, uint64_t* merge, uint32_t n_inst, uint2* out, uint64_t cap,
                             hipStream_t s) {
  k_dense_copy<<<1024, 256, 0, s>>>(chunks, cap_chunks, dense, pairs, ws0.inst, ws0.bin_out, merge, n_inst, out, cap);
  return hipGetLastError();
}

hipError_t launch_deep_result(const DeepMetas& m, const Meta* dense, uint64_t* result, hipStream_t s) {
  k_deep_result<<<1, 64, 0, s>>>(m, dense, result);
  return hipGetLastError();
}

// Place runs of pairs found outside the launch (block r copies run r).
__global__ __launch_bounds__(256) void k_splice_copy(const SpliceRun* __restrict__ runs,
                                                      const uint64_t* __restrict__ inst_out,
                                                      const InstDesc* __restrict__ inst, uint2* __restrict__ out,
                                                      uint64_t cap) {
  const SpliceRun r = runs[blockIdx.x];
  const uint64_t dst0 = inst_out[2 * (uint64_t)r.inst] + inst_out[2 * (uint64_t)r.inst + 1] -
                        inst[r.inst].pairs_extra + r.within;
  for (uint64_t k = threadIdx.x; k < r.len; k += 256) {
    const uint64_t d = dst0 + k;
    if (d < cap) out[d] = r.src[k];
  }
}

hipError_t launch_splice_copy(const SpliceRun* runs, uint32_t n_runs, const uint64_t* inst_out,
                              const InstDesc* inst, uint2* out, uint64_t cap, hipStream_t s) {
  if (!n_runs) return hipSuccess;
  k_splice_copy<<<n_runs, 256, 0, s>>>(runs, inst_out, inst, out, cap);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------

// Dynamic LDS of the place kernel (its static part is a few words).
constexpr uint32_t kPlaceLdsBytes = 156u * 1024u;
static_assert(12u * kMaxBins + 64u + 24u * 1024u <= kPlaceLdsBytes, "a stage of 1,024 entries at the most bins");

hipError_t launch_result(const Meta* meta, uint64_t* result, hipStream_t s) {
  k_bin_result<<<1, 64, 0, s>>>(meta, result);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_single_tables(InstDesc d, InstDesc* __restrict__ inst,
                                                       uint32_t* __restrict__ chunk_inst,
                                                       uint32_t* __restrict__ bin_inst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) inst[0] = d;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.nchunks; i += stride) chunk_inst[i] = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.nb; i += stride) bin_inst[i] = 0u;
}

hipError_t launch_single_tables(const BinnedWorkspace& ws, const InstDesc& d, hipStream_t s) {
  const uint32_t m = d.nchunks > d.nb ? d.nchunks : d.nb;
  const uint32_t blocks = m ? (m + 255u) / 256u : 1u;
  k_single_tables<<<blocks < 64u ? blocks : 64u, 256, 0, s>>>(d, ws.inst, ws.chunk_inst, ws.bin_inst);
  return hipGetLastError();
}

// A single instance with at most kFoldBins bins: the scatter scans the bin
// totals itself (no k_bin_starts), and the test walks the bins in plain
// order.
static bool launch_folds(const BatchPlan& bp) { return bp.n_inst == 1u && bp.nb_max <= kFoldBins; }

hipError_t launch_binned_output(const BinnedWorkspace& ws, const BatchPlan& bp, uint2* out, uint64_t cap,
                                uint64_t* inst_out, hipStream_t s, const uint64_t* merge) {
  const uint32_t nbins = bp.n_bins ? bp.n_bins : 1u;
  k_bin_pair_scan<<<1, 1024, 0, s>>>(ws.bin_seg, bp.n_bins, ws.bin_out, bp.extra_total,
                                     merge ? merge + 3ull * bp.n_inst : nullptr, ws.meta);
  k_bin_compact<<<nbins, 256, 0, s>>>(ws.scratch, ws.cap_wg, bp.test_blocks, ws.bin_seg, ws.bin_out, bp.n_bins,
                                      ws.inst, ws.bin_inst, bp.n_inst, out, cap, inst_out, merge,
                                      launch_folds(bp) ? nullptr : ws.rank);
  return hipGetLastError();
}

hipError_t launch_binned_onepass(const BinnedWorkspace& ws, const OnePass& op, const float4* boxes,
                                 const BatchPlan& bp, uint2* out, uint64_t cap, hipStream_t s, hipEvent_t* ev) {
  hipError_t e;
#define PBIG_MARK(k) \
  if (ev && (e = hipEventRecord(ev[k], s)) != hipSuccess) return e;
  PBIG_MARK(0)
  // one workgroup per CU, each with at least 2,048 boxes (device atomics:
  // workgroups x bins touched)
  const uint64_t want = (bp.n_boxes + 2047u) / 2048u;
  uint32_t grid = bp.n_cu ? bp.n_cu : 1u;
  if (want < grid) grid = want ? (uint32_t)want : 1u;
  const uint32_t nbp = bp.nb_max ? bp.nb_max : 1u;
  const uint32_t stage = place_stage_entries(nbp, kPlaceLdsBytes);
  if (stage < 64u) return hipErrorInvalidValue;  // too many bins for the stage (the planner keeps nb <= kMaxBins)
  k_bin_place<<<grid, kPlaceThreads, 24u * stage + 12u * nbp, s>>>(boxes, ws.inst, op.fill, op.fill_next,
                                                                  op.fill_words, ws.bbox, ws.bid, op.bin_cap, stage,
                                                                  op.bin_start, op.meta, op.meta_next,
                                                                  op.dense_exit);
  PBIG_MARK(1)
  k_bin_test<false><<<bp.test_blocks, kTestThreads, kTestLdsBytes, s>>>(
      ws.bbox, ws.bid, op.bin_start, op.fill, ws.inst, ws.bin_inst, bp.n_bins, (uint64_t)bp.n_bins * op.bin_cap,
      ws.scratch, ws.cap_wg, ws.bin_seg, ws.overfull, out, cap, 1u, op.meta,
      kSingleSum | (op.dense_exit ? kSingleDense : 0u), nullptr);
  PBIG_MARK(2)
#undef PBIG_MARK
  return hipGetLastError();
}

// The count, scan and scatter kernels of the two-pass launch: every box's
// coarse memberships in bin order in the bucket (events 0-4 when ev).
hipError_t launch_binned_front(const BinnedWorkspace& ws, const float4* boxes, const uint32_t* ids_in,
                               const BatchPlan& bp, hipStream_t s, hipEvent_t* ev) {
  hipError_t e;
  // Every kernel tolerates an empty batch (no chunks, no bins): grids of at
  // least one block whose blocks find no work.
  const uint32_t nc = bp.n_chunks ? bp.n_chunks : 1u;
  const uint32_t nbins = bp.n_bins ? bp.n_bins : 1u;
#define PBIG_MARK(k) \
  if (ev && (e = hipEventRecord(ev[k], s)) != hipSuccess) return e;
  PBIG_MARK(0)
  k_bin_hist<false><<<nc, kHistThreads, 0, s>>>(boxes, ws.inst, ws.chunk_inst, bp.n_chunks, ws.hist, ws.meta);
  PBIG_MARK(1)
  // one block more than the bins need: the last one makes the call's result words
  k_bin_scan<false><<<(nbins + kScanBins - 1) / kScanBins + 1, 1024, 0, s>>>(ws.hist, ws.inst, ws.bin_inst, bp.n_bins, ws.chunk_inst,
                                               bp.n_chunks, ws.off, ws.total, ws.meta);
  PBIG_MARK(2)
  const uint32_t fold = launch_folds(bp) ? 1u : 0u;
  if (!fold)
    k_bin_starts<false><<<1, 1024, 0, s>>>(ws.total, bp.n_bins, ws.bin_start, ws.cap_bucket, bp.dense_exit, ws.meta,
                                           ws.order, ws.rank);
  PBIG_MARK(3)
  const uint32_t nb_cap = bp.nb_max ? bp.nb_max : 1u;
  const size_t scatter_lds = sizeof(ScatterLds) + (fold ? 12u : 8u) * nb_cap;
  // A single instance's chunks are all alike: persistent workgroups, one per
  // CU (a multiple of 8, one share per XCD).  A batch's chunks differ in
  // size: one workgroup per chunk, balanced by the dispatcher.
  const uint32_t xcd_share = bp.n_inst == 1u && bp.n_cu >= 16u ? bp.n_cu / 8u : (nc + 7u) / 8u;
  const uint32_t sgrid = (nc + 7u) / 8u < xcd_share ? (nc + 7u) / 8u * 8u : xcd_share * 8u;
  if (ids_in)
    k_bin_scatter<true, false><<<sgrid, kScatterThreads, scatter_lds, s>>>(
        boxes, ids_in, ws.inst, ws.chunk_inst, bp.n_chunks, ws.hist, ws.off, ws.bin_start, ws.bbox, ws.bid,
        ws.cap_bucket, nb_cap, ws.total, fold, bp.dense_exit, ws.meta);
  else
    k_bin_scatter<false, false><<<sgrid, kScatterThreads, scatter_lds, s>>>(
        boxes, nullptr, ws.inst, ws.chunk_inst, bp.n_chunks, ws.hist, ws.off, ws.bin_start, ws.bbox, ws.bid,
        ws.cap_bucket, nb_cap, ws.total, fold, bp.dense_exit, ws.meta);
  PBIG_MARK(4)
#undef PBIG_MARK
  return hipGetLastError();
}

hipError_t launch_binned(const BinnedWorkspace& ws, const float4* boxes, const uint32_t* ids_in,
                         const BatchPlan& bp, uint2* out, uint64_t cap, uint64_t* inst_out, hipStream_t s,
                         hipEvent_t* ev, bool output) {
  hipError_t e;
  if ((e = launch_binned_front(ws, boxes, ids_in, bp, s, ev)) != hipSuccess) return e;
  const uint32_t nbins = bp.n_bins ? bp.n_bins : 1u;
  const uint32_t fold = launch_folds(bp) ? 1u : 0u;
#define PBIG_MARK(k) \
  if (ev && (e = hipEventRecord(ev[k], s)) != hipSuccess) return e;
  k_bin_test<false><<<bp.test_blocks, kTestThreads, kTestLdsBytes, s>>>(ws.bbox, ws.bid, ws.bin_start, ws.total,
                                                                   ws.inst, ws.bin_inst, bp.n_bins, ws.cap_bucket,
                                                                   ws.scratch, ws.cap_wg, ws.bin_seg,
                                                                   ws.overfull, out, cap, bp.direct, ws.meta, 0u,
                                                                   fold ? nullptr : ws.order);
  PBIG_MARK(5)
  if (!bp.direct && output)
    k_bin_pair_scan<<<1, 1024, 0, s>>>(ws.bin_seg, bp.n_bins, ws.bin_out, bp.extra_total, nullptr, ws.meta);
  PBIG_MARK(6)
  if (!bp.direct && output)
    k_bin_compact<<<nbins, 256, 0, s>>>(ws.scratch, ws.cap_wg, bp.test_blocks, ws.bin_seg, ws.bin_out, bp.n_bins,
                                        ws.inst, ws.bin_inst, bp.n_inst, out, cap, inst_out, nullptr,
                                        fold ? nullptr : ws.rank);
  PBIG_MARK(7)
#undef PBIG_MARK
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The general path's cell layout built from the bins (after the two-pass
// launch's count, scan and scatter, one instance), for k_test: every fine
// membership of a bin (a box once per fine cell it covers inside the bin)
// in fine-cell order, every cell's run inside the bin's share of the
// membership array, with the box's start-column and start-row flags, and
// the 64 x 64 tile-pair items of the bin's cells.  Two kernels:
//   k_fine_count  persistent workgroups taking the bins by ticket: fine
//                 counts in LDS (a bin of more than kCountSlice memberships
//                 counted in slices spread over the chip, then summed by the
//                 workgroup that counts its last slice), their scan, the
//                 bin's shares of the membership, item and strip arrays
//                 (three device atomics per bin, not one per membership),
//                 the items, every cell's absolute start, and the bin's
//                 strips (row-major runs of cells starting within one window
//                 of kSortStage / 2 memberships) on a work list; a bin of at
//                 most kCountStage memberships is placed right there instead;
//   k_fine_place  persistent workgroups taking the strips of all bins by
//                 ticket: the bin's memberships stream from the bucket, the
//                 strip's ones go to an LDS stage in cell order and out in
//                 one run (a strip past the stage, i.e. a cell of more than
//                 kSortStage / 2 memberships at its end, is written in
//                 place); a bin of more than kSortBigBin memberships is
//                 placed in slices instead, spread over the chip like the
//                 strips.
constexpr uint32_t kSortThreads = 1024;
constexpr uint32_t kSortPer = kMaxCellsPerBin / kSortThreads;  // cells per thread in the scans
static_assert(kSortPer * kSortThreads == kMaxCellsPerBin, "four cells per thread");
constexpr uint32_t kSortStage = 5120;
constexpr uint32_t kSortWindow = kSortStage / 2;
// Memberships a thread loads before it uses any of them (a dense bin's tens
// of thousands of memberships are otherwise one load round trip each).
constexpr uint32_t kSortLoads = 8;
constexpr uint32_t kSortBigBin = 8u * kSortWindow;
// A bin of more than kSortBigBin fine memberships is placed in slices of
// kSortSlice coarse memberships, each by any workgroup: the slice's cells
// counted in LDS, every cell's range reserved with one device atomic on that
// cell's cursor, the memberships written there (strip items carry
// kSortSliceTag in their last word, the slice index in the second).
constexpr uint32_t kSortSlice = 4096;
// A bin of at most kCountStage fine memberships is placed by k_fine_count
// itself, right after its layout (no strips, no second read by k_fine_place).
constexpr uint32_t kCountStage = 5120;
static_assert(kCountStage >= kMaxCellsPerBin, "the stage's id words hold a strip list");
constexpr uint32_t kSortSliceTag = 0xFFFFFFFFu;

// Tile-pair items of a cell of n memberships (none past the tile limit: the
// call then fails with kErrTooDense).
__device__ __forceinline__ uint64_t tile_items(uint32_t n) {
  const uint64_t nt = (n + kTile - 1u) / kTile;
  return n < 2u || nt > kMaxTilesPerCell ? 0ull : nt * (nt + 1u) / 2u;
}

// A bin's fine cells (k x k cells of the instance's grid, an owned instance
// of the plain path) and the fine cells a membership covers inside it.
struct SortBin {
  uint32_t gx, gy, kx, ky;
  __device__ SortBin(const InstDesc& I, uint32_t b) {
    gx = I.fx0 + (b % I.nbx) * I.k;
    gy = I.fy0 + (b / I.nbx) * I.k;
    const uint32_t rx = I.fx0 + I.span_x - gx, ry = I.fy0 + I.span_y - gy;
    kx = I.k < rx ? I.k : rx;
    ky = I.k < ry ? I.k : ry;
  }
  // (every membership of the bin covers at least one of its cells)
  __device__ void cells(const InstDesc& I, const float4& q, uint32_t& x0, uint32_t& x1, uint32_t& y0, uint32_t& y1,
                        uint32_t& fx, uint32_t& fy) const {
    fx = fine_of(q.x, I.inv_f, I.nf);
    fy = fine_of(q.z, I.inv_f, I.nf);
    const uint32_t ux = fine_of(q.y, I.inv_f, I.nf), uy = fine_of(q.w, I.inv_f, I.nf);
    x0 = (fx > gx ? fx : gx) - gx;
    y0 = (fy > gy ? fy : gy) - gy;
    x1 = (ux < gx + kx - 1u ? ux : gx + kx - 1u) - gx;
    y1 = (uy < gy + ky - 1u ? uy : gy + ky - 1u) - gy;
  }
};

struct CountLds {
  uint32_t cnt[kMaxCellsPerBin];
  uint32_t eoff[kMaxCellsPerBin + 1];  // per cell: first membership and first item in the bin
  uint32_t ioff[kMaxCellsPerBin];
  float4 sbox[kCountStage];  // a bin of at most kCountStage memberships, laid out in place
  uint32_t sid[kCountStage];  // (for a bin placed by strips: the first cell of each strip)
  uint32_t soff[kSortThreads + 1];  // a chunk of bins: inclusive scan of their slices
  uint32_t boff[kSortThreads + 1];  // and of their other bins
  uint64_t wm[kSortThreads / 64], wi[kSortThreads / 64];
  uint32_t ws[kSortThreads / 64], wb[kSortThreads / 64];
  unsigned long long fbase, ibase;
  uint32_t bmax, sbase, last, tick;
};
static_assert(sizeof(CountLds) <= 160u * 1024u, "one workgroup per CU");

// A bin's fine counts of memberships j0 .. j1 - 1 of the bucket into L.cnt.
__device__ __forceinline__ void count_fine(CountLds& L, const InstDesc& I, const SortBin& B,
                                           const float4* __restrict__ bbox, uint64_t s0, uint32_t j0,
                                           uint32_t j1) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t i0 = j0 + tid; i0 < j1; i0 += kSortThreads * kSortLoads) {
    float4 q[kSortLoads];
#pragma unroll
    for (uint32_t u = 0; u < kSortLoads; ++u) {
      const uint32_t j = i0 + u * kSortThreads;
      q[u] = bbox[s0 + (j < j1 ? j : j1 - 1u)];
    }
#pragma unroll
    for (uint32_t u = 0; u < kSortLoads; ++u) {
      if (i0 + u * kSortThreads >= j1) break;
      uint32_t x0, x1, y0, y1, fx, fy;
      B.cells(I, q[u], x0, x1, y0, y1, fx, fy);
      for (uint32_t cy = y0; cy <= y1; ++cy)
        for (uint32_t cx = x0; cx <= x1; ++cx) atomicAdd(&L.cnt[cy * B.kx + cx], 1u);
    }
  }
}

// Bin b's layout from its fine counts in L.cnt (cleared again on return).
// A bin of at most kCountStage memberships (and not counted in slices) is
// also placed here, through an LDS stage, instead of by k_fine_place.
__device__ __forceinline__ void layout_bin(CountLds& L, const InstDesc& I, uint32_t b, uint32_t m, bool whole,
                                        const float4* __restrict__ bbox, const uint32_t* __restrict__ bid,
                                        uint64_t s0, uint32_t* __restrict__ cell_start,
                                        uint32_t* __restrict__ bin_end, uint32_t* __restrict__ cell_cur,
                                        uint64_t cap_memb, uint4* __restrict__ items, uint64_t cap_items,
                                        uint4* __restrict__ strips, uint64_t cap_strips, float4* __restrict__ mbox,
                                        uint32_t* __restrict__ mid, Meta* __restrict__ meta) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t k2 = I.k * I.k;
  const SortBin B(I, b);
  const uint32_t kk = B.kx * B.ky;
  // 1. exclusive scans of the cells' memberships and items (four cells
  //    per thread, 64-bit sums: a share past 2^32 - 1 cannot fit and
  //    fails the call with kErrWorkspace)
  uint32_t v[kSortPer], vmax = 0;
  uint64_t acc = 0, iacc = 0;
#pragma unroll
  for (uint32_t k = 0; k < kSortPer; ++k) {
    const uint32_t c = tid * kSortPer + k;
    v[k] = c < kk ? L.cnt[c] : 0u;
    acc += v[k];
    iacc += tile_items(v[k]);
    vmax = v[k] > vmax ? v[k] : vmax;
  }
  const uint64_t inc = wave_prefix_sum64(acc), iinc = wave_prefix_sum64(iacc);
  if (lane == 63u) {
    L.wm[w] = inc;
    L.wi[w] = iinc;
  }
  if (vmax) atomicMax(&L.bmax, vmax);
  __syncthreads();  // (the counts are in registers from here: L.cnt is free until it is cleared)
  uint64_t e = inc - acc, ie = iinc - iacc, F = 0, II = 0;
#pragma unroll
  for (uint32_t k = 0; k < kSortThreads / 64; ++k) {  // (all loads issued before the sums wait)
    const uint64_t a = L.wm[k], c = L.wi[k];
    e += k < w ? a : 0ull;
    ie += k < w ? c : 0ull;
    F += a;
    II += c;
  }
  {
    uint64_t e2 = e, ie2 = ie;
#pragma unroll
    for (uint32_t k = 0; k < kSortPer; ++k) {
      const uint32_t c = tid * kSortPer + k;
      if (c < kk) {
        L.eoff[c] = (uint32_t)e2;  // (a bin past 2^32 - 1 memberships or items fails below)
        L.ioff[c] = (uint32_t)ie2;
      }
      e2 += v[k];
      ie2 += tile_items(v[k]);
    }
    if (tid == 0) L.eoff[kk] = (uint32_t)F;
  }
  // 2. the strips: cell 0 opens one, and so does every cell whose start
  //    lies in a later window of kSortWindow memberships than its
  //    predecessor's (a strip then spans less than two windows unless its
  //    last cell alone is larger); a bin of more than kSortBigBin
  //    memberships is placed in slices instead
  uint32_t win[kSortPer];
  {
    uint64_t e3 = e;
#pragma unroll
    for (uint32_t k = 0; k < kSortPer; ++k) {
      win[k] = (uint32_t)(e3 / kSortWindow);
      e3 += v[k];
    }
  }
  L.cnt[tid] = win[kSortPer - 1];  // the window of this thread's last cell
  __syncthreads();
  const uint32_t before = tid ? L.cnt[tid - 1u] : 0xFFFFFFFFu;
  uint32_t open = 0;  // bit k: this thread's cell k opens a strip
#pragma unroll
  for (uint32_t k = 0; k < kSortPer; ++k) {
    const uint32_t c = tid * kSortPer + k;
    const uint32_t pw = k ? win[k - 1] : before;
    if (c < kk && (c == 0u || win[k] != pw)) open |= 1u << k;
  }
  const uint32_t nopen = (uint32_t)__popc(open);
  const uint32_t sinc = wave_prefix_sum(nopen);
  if (lane == 63u) L.ws[w] = sinc;
  __syncthreads();
  uint32_t sidx = sinc - nopen, nst = 0;
#pragma unroll
  for (uint32_t k = 0; k < kSortThreads / 64; ++k) {
    const uint32_t a = L.ws[k];
    sidx += k < w ? a : 0u;
    nst += a;
  }
#pragma unroll
  for (uint32_t k = 0; k < kSortPer; ++k)
    if (open & (1u << k)) L.sid[sidx++] = tid * kSortPer + k;
  const bool sliced = F > kSortBigBin;
  if (sliced) nst = (m + kSortSlice - 1u) / kSortSlice;
  const bool here = whole && F <= kCountStage;  // placed by this workgroup
  if (here) nst = 0u;
  // 3. the bin's shares of the membership, item and strip arrays: three
  //    device atomics from three waves, in flight together (a share that
  //    does not fit fails the call; the kernels after this one then stop
  //    at the error word before they read a strip)
  if (tid == 0u) L.fbase = atomicAdd((unsigned long long*)&meta->memberships, (unsigned long long)F);
  if (tid == 64u) L.ibase = atomicAdd((unsigned long long*)&meta->items, (unsigned long long)II);
  if (tid == 128u) L.sbase = atomicAdd(&meta->dev_chunks, nst);
  if (tid == 192u) {
    atomicMax(&meta->max_per_cell, L.bmax);
    if (L.bmax > kTile * kMaxTilesPerCell) atomicOr((unsigned long long*)&meta->error, kErrTooDense);
  }
  __syncthreads();
  const uint64_t fb = L.fbase, ib = L.ibase;
  const uint32_t sb = L.sbase;
  const bool fits = F < (1ull << 32) && II < (1ull << 32) && fb + F <= cap_memb && ib + II <= cap_items;
  if (!fits && tid == 0) atomicOr((unsigned long long*)&meta->error, kErrWorkspace);
  // 4. every cell's absolute start, the items (spread over the threads:
  //    item t of the bin is item x of the cell c whose items it falls in, the
  //    tile pair (r, q) at x in row-major order of r <= q < nt), and the
  //    strips
  if (fits) {
    uint64_t e2 = e;
#pragma unroll
    for (uint32_t k = 0; k < kSortPer; ++k) {
      const uint32_t c = tid * kSortPer + k;
      if (c < kk) {
        cell_start[(uint64_t)b * k2 + c] = (uint32_t)(fb + e2);
        if (sliced) cell_cur[(uint64_t)b * k2 + c] = (uint32_t)(fb + e2);
      }
      e2 += v[k];
    }
    for (uint32_t t = tid; t < (uint32_t)II; t += kSortThreads) {
      uint32_t lo = 0, hi = kk;  // the last cell with ioff <= t
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (L.ioff[mid] <= t) lo = mid; else hi = mid;
      }
      const uint32_t n = L.eoff[lo + 1u] - L.eoff[lo];
      const uint64_t nt = (n + kTile - 1u) / kTile;
      const uint64_t x = t - L.ioff[lo];
      // rows before row r hold r nt - r (r - 1) / 2 items
      const double d2 = 2.0 * (double)nt + 1.0;
      uint64_t r = (uint64_t)((d2 - sqrt(d2 * d2 - 8.0 * (double)x)) * 0.5);
      while (r > 0u && r * nt - r * (r - 1u) / 2u > x) --r;
      while ((r + 1u) * nt - (r + 1u) * r / 2u <= x) ++r;
      const uint64_t q = r + (x - (r * nt - r * (r - 1u) / 2u));
      items[ib + t] = make_uint4((uint32_t)(fb + L.eoff[lo]), n, (uint32_t)r, (uint32_t)q);
    }
    for (uint32_t t = tid; t < nst; t += kSortThreads) {
      const uint4 item = sliced ? make_uint4(b, t, 0u, kSortSliceTag)
                                : make_uint4(b, L.sid[t], t + 1u < nst ? L.sid[t + 1u] : kk, 0u);
      if ((uint64_t)sb + t < cap_strips) strips[sb + t] = item;
    }
    if (tid == 0) {
      bin_end[b] = (uint32_t)(fb + F);
      if ((uint64_t)sb + nst > cap_strips) atomicOr((unsigned long long*)&meta->error, kErrWorkspace);
    }
  }
  __syncthreads();
  L.cnt[tid] = 0u;  // (step 2's window word, also where kk is below kSortThreads)
  for (uint32_t c = tid + kSortThreads; c < kk; c += kSortThreads) L.cnt[c] = 0u;
  if (tid == 0) L.bmax = 0u;
  __syncthreads();
  if (here && fits) {
    // 5. the memberships ranked per cell into the stage (L.cnt as the cells'
    //    cursors), then written out in one run
    for (uint32_t j0 = tid; j0 < m; j0 += kSortThreads * kSortLoads) {
      float4 qs[kSortLoads];
      uint32_t ids[kSortLoads];
#pragma unroll
      for (uint32_t u = 0; u < kSortLoads; ++u) {
        const uint32_t j = j0 + u * kSortThreads;
        const uint64_t at = s0 + (j < m ? j : m - 1u);
        qs[u] = bbox[at];
        ids[u] = bid[at];
      }
#pragma unroll
      for (uint32_t u = 0; u < kSortLoads; ++u) {
        if (j0 + u * kSortThreads >= m) break;
        uint32_t x0, x1, y0, y1, fx, fy;
        B.cells(I, qs[u], x0, x1, y0, y1, fx, fy);
        for (uint32_t cy = y0; cy <= y1; ++cy) {
          const uint32_t flag_y = cy + B.gy == fy ? kFlagY : 0u;
          for (uint32_t cx = x0; cx <= x1; ++cx) {
            const uint32_t c = cy * B.kx + cx;
            const uint32_t p = L.eoff[c] + atomicAdd(&L.cnt[c], 1u);
            L.sbox[p] = qs[u];
            L.sid[p] = ids[u] | flag_y | (cx + B.gx == fx ? kFlagX : 0u);
          }
        }
      }
    }
    __syncthreads();
    for (uint32_t i = tid; i < (uint32_t)F; i += kSortThreads) {
      mbox[fb + i] = L.sbox[i];
      mid[fb + i] = L.sid[i];
    }
    for (uint32_t c = tid; c < kk; c += kSortThreads) L.cnt[c] = 0u;
    __syncthreads();
  }
}

// Bins of more than kCountSlice memberships are counted in slices of
// kCountSlice by any workgroup: a slice adds its fine counts to the bin's
// row of slc with device atomics, and the workgroup that counts a bin's last
// slice takes the sums with atomic exchanges (the row is zero again for the
// next launch) and lays the bin out.  Work is taken by ticket, chunk of
// kSortThreads bins by chunk: the chunk's slices first (heaviest work), then
// its other bins, one ticket each.  slc: per bin a slice counter, per chunk
// a ticket counter (both zeroed before the launch), per bin a row of k^2
// counts.
__global__ __launch_bounds__(kSortThreads) void k_fine_count(const float4* __restrict__ bbox,
                                                             const uint32_t* __restrict__ bid,
                                                             const uint32_t* __restrict__ bin_start,
                                                             const uint32_t* __restrict__ total,
                                                             const InstDesc* __restrict__ inst, uint32_t n_bins,
                                                             const Meta* __restrict__ bmeta,
                                                             uint32_t* __restrict__ cell_start,
                                                             uint32_t* __restrict__ bin_end,
                                                             uint32_t* __restrict__ cell_cur, uint64_t cap_memb,
                                                             uint4* __restrict__ items, uint64_t cap_items,
                                                             uint4* __restrict__ strips, uint64_t cap_strips,
                                                             uint32_t* __restrict__ slc, float4* __restrict__ mbox,
                                                             uint32_t* __restrict__ mid, Meta* __restrict__ meta) {
  extern __shared__ __align__(16) unsigned char count_lds[];
  CountLds& L = *reinterpret_cast<CountLds*>(count_lds);
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  // an invalid box or a bucket that did not fit: nothing to lay out
  const unsigned long long berr = bmeta->error & (kErrInvalidBox | kErrWorkspace);
  if (berr) {
    if (blockIdx.x == 0 && tid == 0) atomicOr((unsigned long long*)&meta->error, berr);
    return;
  }
  const InstDesc I = inst[0];
  const uint32_t k2 = I.k * I.k;
  const uint32_t n_chunks = (n_bins + kSortThreads - 1u) / kSortThreads;
  uint32_t* const done = slc;
  uint32_t* const tick = slc + n_bins;
  uint32_t* const rows = slc + n_bins + n_chunks;
  for (uint32_t c = tid; c < kMaxCellsPerBin; c += kSortThreads) L.cnt[c] = 0u;
  if (tid == 0) L.bmax = 0u;
  __syncthreads();
  for (uint32_t chunk = 0; chunk < n_chunks; ++chunk) {
    const uint32_t base = chunk * kSortThreads;
    const uint32_t bb = base + tid;
    const uint32_t mm = bb < n_bins ? total[bb] : 0u;
    const uint32_t ns = mm > kCountSlice ? (mm + kCountSlice - 1u) / kCountSlice : 0u;
    const uint32_t sm = mm > 0u && mm <= kCountSlice ? 1u : 0u;
    const uint32_t inc = wave_prefix_sum(ns), incb = wave_prefix_sum(sm);
    if (lane == 63u) {
      L.ws[w] = inc;
      L.wb[w] = incb;
    }
    if (tid == 0) L.tick = atomicAdd(&tick[chunk], 1u);
    __syncthreads();
    uint32_t pre = inc, preb = incb, T = 0, TB = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSortThreads / 64; ++k) {
      const uint32_t a = L.ws[k], c = L.wb[k];
      pre += k < w ? a : 0u;
      preb += k < w ? c : 0u;
      T += a;
      TB += c;
    }
    L.soff[tid + 1u] = pre;  // inclusive scans: slices, other bins
    L.boff[tid + 1u] = preb;
    if (tid == 0) {
      L.soff[0] = 0u;
      L.boff[0] = 0u;
    }
    __syncthreads();
    for (;;) {
      const uint32_t u = L.tick;
      if (u >= T + TB) break;  // (uniform)
      __syncthreads();
      if (tid == 0) L.tick = atomicAdd(&tick[chunk], 1u);  // the next ticket, in flight meanwhile
      const bool slice = u < T;
      const uint32_t r = slice ? u : u - T;
      const uint32_t* const off = slice ? L.soff : L.boff;
      uint32_t lo = 0, hi = kSortThreads;  // off[lo] <= r < off[lo + 1]
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= r) lo = mid; else hi = mid;
      }
      const uint32_t b = base + lo, m = total[b];
      const SortBin B(I, b);
      if (slice) {
        const uint32_t kk = B.kx * B.ky;
        const uint32_t j = r - L.soff[lo], nsb = L.soff[lo + 1u] - L.soff[lo];
        const uint32_t j0 = j * kCountSlice, j1 = j0 + kCountSlice < m ? j0 + kCountSlice : m;
        count_fine(L, I, B, bbox, bin_start[b], j0, j1);
        __syncthreads();
        uint32_t seen = 0;
        for (uint32_t c = tid; c < kk; c += kSortThreads) {
          if (L.cnt[c]) seen |= atomicAdd(&rows[(uint64_t)b * k2 + c], L.cnt[c]);
          L.cnt[c] = 0u;
        }
        // every add has returned, so it is performed at device scope, before
        // this workgroup counts the slice as done
        asm volatile("" ::"v"(seen));
        __syncthreads();
        if (tid == 0) L.last = atomicAdd(&done[b], 1u) == nsb - 1u;
        __syncthreads();
        if (L.last) {  // every slice of bin b is counted: take the sums
          for (uint32_t c = tid; c < kk; c += kSortThreads)
            L.cnt[c] = atomicExch(&rows[(uint64_t)b * k2 + c], 0u);
          __syncthreads();
          layout_bin(L, I, b, m, false, bbox, bid, bin_start[b], cell_start, bin_end, cell_cur, cap_memb, items,
                     cap_items, strips, cap_strips, mbox, mid, meta);
        }
      } else {
        count_fine(L, I, B, bbox, bin_start[b], 0u, m);
        __syncthreads();
        layout_bin(L, I, b, m, true, bbox, bid, bin_start[b], cell_start, bin_end, cell_cur, cap_memb, items,
                   cap_items, strips, cap_strips, mbox, mid, meta);
      }
      __syncthreads();  // (the next ticket is in L.tick)
    }
    __syncthreads();  // (soff, boff and ws are reused by the next chunk)
  }
}

struct PlaceLds {
  uint32_t cur[kMaxCellsPerBin];  // per cell of the strip: cursor
  uint32_t cst[kMaxCellsPerBin];  // per cell of the strip: start, from the strip's
  float4 sbox[kSortStage];
  uint32_t sid[kSortStage];
  uint32_t tick;  // the workgroup's next strip
};
static_assert(sizeof(PlaceLds) <= 160u * 1024u, "one workgroup per CU");

__global__ __launch_bounds__(kSortThreads) void k_fine_place(const float4* __restrict__ bbox,
                                                             const uint32_t* __restrict__ bid,
                                                             const uint32_t* __restrict__ bin_start,
                                                             const uint32_t* __restrict__ total,
                                                             const InstDesc* __restrict__ inst,
                                                             const uint32_t* __restrict__ cell_start,
                                                             const uint32_t* __restrict__ bin_end,
                                                             uint32_t* __restrict__ cell_cur,
                                                             const uint4* __restrict__ strips, uint64_t cap_strips,
                                                             float4* __restrict__ mbox, uint32_t* __restrict__ mid,
                                                             Meta* __restrict__ meta) {
  extern __shared__ __align__(16) unsigned char place_lds[];
  PlaceLds& L = *reinterpret_cast<PlaceLds*>(place_lds);
  const uint32_t tid = threadIdx.x;
  if (meta->error) return;  // (the layout is incomplete: the call fails or grows and reruns)
  const uint32_t n_strips = (uint64_t)meta->dev_chunks < cap_strips ? meta->dev_chunks : (uint32_t)cap_strips;
  const InstDesc I = inst[0];
  const uint32_t k2 = I.k * I.k;
  // strips taken by ticket (meta->dev_bins, zeroed with the result words)
  if (tid == 0) L.tick = atomicAdd(&meta->dev_bins, 1u);
  __syncthreads();
  for (;;) {
    const uint32_t t = L.tick;
    if (t >= n_strips) break;  // (uniform)
    __syncthreads();
    if (tid == 0) L.tick = atomicAdd(&meta->dev_bins, 1u);  // the next one, in flight meanwhile
    const uint4 st = strips[t];
    const uint32_t b = st.x, lo = st.y, hi = st.z;
    const uint32_t m = total[b];
    const uint64_t s0 = bin_start[b];
    const SortBin B(I, b);
    const uint32_t kk = B.kx * B.ky;
    const uint64_t cb = (uint64_t)b * k2;
    if (st.w == kSortSliceTag) {
      // a slice of a large bin: its memberships j0 .. j1 - 1, counted per
      // cell, each cell's range reserved on the cell's cursor, then written
      const uint32_t j0 = lo * kSortSlice, j1 = j0 + kSortSlice < m ? j0 + kSortSlice : m;
      for (uint32_t c = tid; c < kk; c += kSortThreads) L.cur[c] = 0u;
      __syncthreads();
      for (uint32_t i0 = j0 + tid; i0 < j1; i0 += kSortThreads * kSortLoads) {
        float4 qs[kSortLoads];
#pragma unroll
        for (uint32_t u = 0; u < kSortLoads; ++u) {
          const uint32_t j = i0 + u * kSortThreads;
          qs[u] = bbox[s0 + (j < j1 ? j : j1 - 1u)];
        }
#pragma unroll
        for (uint32_t u = 0; u < kSortLoads; ++u) {
          if (i0 + u * kSortThreads >= j1) break;
          uint32_t x0, x1, y0, y1, fx, fy;
          B.cells(I, qs[u], x0, x1, y0, y1, fx, fy);
          for (uint32_t cy = y0; cy <= y1; ++cy)
            for (uint32_t cx = x0; cx <= x1; ++cx) atomicAdd(&L.cur[cy * B.kx + cx], 1u);
        }
      }
      __syncthreads();
      for (uint32_t c = tid; c < kk; c += kSortThreads) {
        const uint32_t n = L.cur[c];
        L.cst[c] = n ? atomicAdd(&cell_cur[cb + c], n) : 0u;  // the slice's first slot in cell c
      }
      __syncthreads();
      for (uint32_t i0 = j0 + tid; i0 < j1; i0 += kSortThreads * kSortLoads) {
        float4 qs[kSortLoads];
        uint32_t ids[kSortLoads];
#pragma unroll
        for (uint32_t u = 0; u < kSortLoads; ++u) {
          const uint32_t j = i0 + u * kSortThreads;
          const uint64_t at = s0 + (j < j1 ? j : j1 - 1u);
          qs[u] = bbox[at];
          ids[u] = bid[at];
        }
#pragma unroll
        for (uint32_t u = 0; u < kSortLoads; ++u) {
          if (i0 + u * kSortThreads >= j1) break;
          uint32_t x0, x1, y0, y1, fx, fy;
          B.cells(I, qs[u], x0, x1, y0, y1, fx, fy);
          for (uint32_t cy = y0; cy <= y1; ++cy) {
            const uint32_t flag_y = cy + B.gy == fy ? kFlagY : 0u;
            for (uint32_t cx = x0; cx <= x1; ++cx) {
              const uint32_t c = cy * B.kx + cx;
              const uint64_t pos = (uint64_t)atomicAdd(&L.cst[c], 1u);
              mbox[pos] = qs[u];
              mid[pos] = ids[u] | flag_y | (cx + B.gx == fx ? kFlagX : 0u);
            }
          }
        }
      }
      __syncthreads();  // the counters are reused by the next item
      continue;
    }
    const uint32_t base = cell_start[cb + lo];
    const uint32_t span = (hi < kk ? cell_start[cb + hi] : bin_end[b]) - base;
    const bool staged = span <= kSortStage;
    for (uint32_t c = lo + tid; c < hi; c += kSortThreads) {
      L.cur[c - lo] = 0u;
      L.cst[c - lo] = cell_start[cb + c] - base;
    }
    __syncthreads();
    const uint32_t ylo = lo / B.kx, yhi = (hi - 1u) / B.kx;
    for (uint32_t j0 = tid; j0 < m; j0 += kSortThreads * kSortLoads) {
      float4 qs[kSortLoads];
      uint32_t ids[kSortLoads];
#pragma unroll
      for (uint32_t u = 0; u < kSortLoads; ++u) {
        const uint32_t j = j0 + u * kSortThreads;
        const uint64_t at = s0 + (j < m ? j : m - 1u);
        qs[u] = bbox[at];
        ids[u] = bid[at];
      }
#pragma unroll
      for (uint32_t u = 0; u < kSortLoads; ++u) {
        if (j0 + u * kSortThreads >= m) break;
        const float4 q = qs[u];
        uint32_t x0, x1, y0, y1, fx, fy;
        B.cells(I, q, x0, x1, y0, y1, fx, fy);
        if (y1 < ylo || y0 > yhi) continue;
        const uint32_t id = ids[u];
        for (uint32_t cy = y0 > ylo ? y0 : ylo; cy <= (y1 < yhi ? y1 : yhi); ++cy) {
          const uint32_t flag_y = cy + B.gy == fy ? kFlagY : 0u;
          for (uint32_t cx = x0; cx <= x1; ++cx) {
            const uint32_t c = cy * B.kx + cx;
            if (c < lo || c >= hi) continue;
            const uint32_t p = L.cst[c - lo] + atomicAdd(&L.cur[c - lo], 1u);
            const uint32_t word = id | flag_y | (cx + B.gx == fx ? kFlagX : 0u);
            if (staged) {
              L.sbox[p] = q;
              L.sid[p] = word;
            } else {
              mbox[(uint64_t)base + p] = q;
              mid[(uint64_t)base + p] = word;
            }
          }
        }
      }
    }
    __syncthreads();
    if (staged) {
      for (uint32_t i = tid; i < span; i += kSortThreads) {
        mbox[(uint64_t)base + i] = L.sbox[i];
        mid[(uint64_t)base + i] = L.sid[i];
      }
    }
    __syncthreads();  // the stage and the cursors are reused by the next strip
  }
}

hipError_t launch_sorted(const BinnedWorkspace& ws, const float4* boxes, const BatchPlan& bp, const Workspace& gw,
                         Meta* meta, uint2* out, uint64_t cap, hipStream_t s, hipEvent_t* ev) {
  hipError_t e;
  if (ev && (e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
  // (the result words, and the bins' slice counters of k_fine_count)
  const uint32_t n_chunks = (bp.n_bins + kSortThreads - 1u) / kSortThreads;
  if ((e = launch_meta_clear(meta, s, gw.slc, bp.n_bins + n_chunks)) != hipSuccess) return e;
  if ((e = launch_binned_front(ws, boxes, nullptr, bp, s, ev ? ev + 1 : nullptr)) != hipSuccess) return e;
  const uint32_t grid = bp.n_cu ? bp.n_cu : 1u;
  // (cell starts in gw.cnt; bin ends and then the large bins' cell cursors in gw.start)
  uint32_t* const cell_cur = gw.start + bp.n_bins;
  k_fine_count<<<grid, kSortThreads, sizeof(CountLds), s>>>(
      ws.bbox, ws.bid, ws.bin_start, ws.total, ws.inst, bp.n_bins, ws.meta, gw.cnt, gw.start, cell_cur, gw.cap_memb,
      gw.items, gw.cap_items, gw.strips, gw.cap_strips, gw.slc, gw.mbox, gw.mid, meta);
  if (ev && (e = hipEventRecord(ev[6], s)) != hipSuccess) return e;
  k_fine_place<<<grid, kSortThreads, sizeof(PlaceLds), s>>>(ws.bbox, ws.bid, ws.bin_start, ws.total, ws.inst, gw.cnt,
                                                            gw.start, cell_cur, gw.strips, gw.cap_strips, gw.mbox,
                                                            gw.mid, meta);
  if (ev && (e = hipEventRecord(ev[7], s)) != hipSuccess) return e;
  if ((e = launch_tile_test(gw, meta, out, cap, s)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(ev[8], s)) != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t binned_init() {
  {
    const hipError_t e0 = hipFuncSetAttribute((const void*)k_bin_place, hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)kPlaceLdsBytes);
    if (e0 != hipSuccess) return e0;
  }
  hipError_t e = hipFuncSetAttribute((const void*)k_bin_test<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kTestLdsBytes);
  if (e != hipSuccess) return e;
  const int scatter_max = (int)(sizeof(ScatterLds) + 8u * kMaxBins);
  for (const void* f : {(const void*)k_bin_test<true>, (const void*)k_bin_scatter<false, false>,
                        (const void*)k_bin_scatter<true, false>, (const void*)k_bin_scatter<true, true>}) {
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                            f == (const void*)k_bin_test<true> ? (int)kTestLdsBytes : scatter_max);
    if (e != hipSuccess) return e;
  }
  e = hipFuncSetAttribute((const void*)k_deep_plan, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDeepPlanLds);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)k_fine_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)sizeof(CountLds));
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_fine_place, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)sizeof(PlaceLds));
}

}  // namespace pbig
