// CPU oracle for the pairwise box intersection path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (pbig_amd/, include/) links,
// loads or calls this file.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// It restates the reference's algorithm (awatch/pbig, src/pbig/pbig.cu) on the
// CPU, plus two independent algorithms (brute force, sweep-and-prune) that pin
// it.  Every function cites the reference lines it follows.
//
// Coordinate conventions
//   * 2-D float32 boxes are stored as {xmin, xmax, ymin, ymax} (the field order
//     of CoordInfo, src/pbig/rect.h:6-12) and a box's id is its index
//     (src/pbig/pbig.cu:331,737 fetch boxes by rid as an array index).
//   * 3-D unsigned boxes (the reference's own shipped configuration,
//     src/pbig/rect.h:4, config.cmake:8) are {xmin,xmax,ymin,ymax,zmin,zmax}.
//
// Float cell map: cell(v) = clamp(floor(v * inv_cs), 0, nrc-1).  It replaces
// the reference's `(unsigned)v >> shift` (pbig.cu:185-188) for float input.
// It is monotone, so the reference-cell rule (pbig.cu:579-603) still assigns
// every overlapping pair to exactly one cell.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

struct Box2 { float xmin, xmax, ymin, ymax; };

// pbig.cu:996-1001 (CPU checker) == pbig.cu:562-568 (device calcOverlaps):
// closed intervals, all four `<=` must hold.
inline bool overlap2(const Box2& a, const Box2& b) {
  return (a.xmin <= b.xmax) && (b.xmin <= a.xmax) &&
         (a.ymin <= b.ymax) && (b.ymin <= a.ymax);
}

inline uint32_t cellmap(float v, float inv, uint32_t nrc) {
  float f = std::floor(v * inv);
  if (!(f >= 0.0f)) return 0u;                 // negative (and NaN) -> first cell
  if (f >= (float)(nrc - 1u)) return nrc - 1u;  // beyond the domain -> last cell
  return (uint32_t)f;
}

struct PairSink {
  uint32_t* out;
  uint64_t cap;
  uint64_t n;
  void push(uint32_t a, uint32_t b) {
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    if (out && n < cap) { out[2 * n] = lo; out[2 * n + 1] = hi; }
    ++n;
  }
};

// Concatenate per-thread pair lists into the caller's buffer.
uint64_t gather(std::vector<std::vector<uint32_t>>& parts, uint32_t* out, uint64_t cap) {
  uint64_t total = 0;
  for (auto& p : parts) {
    uint64_t np = p.size() / 2;
    if (out) {
      for (uint64_t k = 0; k < np && total + k < cap; ++k) {
        out[2 * (total + k)] = p[2 * k];
        out[2 * (total + k) + 1] = p[2 * k + 1];
      }
    }
    total += np;
  }
  return total;
}

int nthreads(int t) { return t > 0 ? t : omp_get_max_threads(); }

}  // namespace

extern "C" {

// Grid statistics of one planning/binning pass.  Field meaning follows the
// reference's LOG lines (pbig.cu:1545-1558 "AvgRecs", "maxRecs",
// "numMappedRecs"; pbig.cu:1629 "Compared Pairs"; pbig.cu:1631-1637 jobs).
struct orc_grid_stats {
  uint64_t n_cells;        // nCells
  uint64_t memberships;    // numMappedRecs = sum of per-cell counts
  uint64_t max_per_cell;   // maxRecs
  uint64_t max_cell;       // maxRecsCell
  uint64_t compared_pairs; // sum over cells with n>1 of (u64)n*n/2
  uint64_t jobs256;        // sum over cells with n>1 of ceil(n/256)
  uint64_t nonempty_cells; // cells with n>0
  uint64_t pairs;          // overlap pairs found
};

// ---------------------------------------------------------------------------
// Predicate and independent algorithms (brute force, sweep-and-prune).
// ---------------------------------------------------------------------------

int orc_overlap2f(const float* a, const float* b) {
  return overlap2(*(const Box2*)a, *(const Box2*)b) ? 1 : 0;
}

// All-pairs brute force with the reference checker's predicate
// (pbig.cu:988-1023 checkOverlaps, minus its duplicate detector).  O(n^2).
uint64_t orc_brute2f(const float* boxes, uint32_t n, uint32_t* out, uint64_t cap, int threads) {
  const Box2* b = (const Box2*)boxes;
  int T = nthreads(threads);
  std::vector<std::vector<uint32_t>> parts(T);
#pragma omp parallel num_threads(T)
  {
    auto& mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      Box2 a = b[i];
      for (uint32_t j = (uint32_t)i + 1; j < n; ++j)
        if (overlap2(a, b[j])) { mine.push_back((uint32_t)i); mine.push_back(j); }
    }
  }
  return gather(parts, out, cap);
}

// Brute-force overlap count only (no output), used for timing the reference's
// CPU checker predicate on a bounded sample.
uint64_t orc_brute2f_count(const float* boxes, uint32_t n, int threads) {
  const Box2* b = (const Box2*)boxes;
  int T = nthreads(threads);
  uint64_t total = 0;
#pragma omp parallel for num_threads(T) schedule(dynamic, 64) reduction(+ : total)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    Box2 a = b[i];
    uint64_t c = 0;
    for (uint32_t j = (uint32_t)i + 1; j < n; ++j) c += overlap2(a, b[j]);
    total += c;
  }
  return total;
}

// Sweep-and-prune: sort by xmin, then each box scans later boxes while their
// xmin <= its xmax.  Independent of the grid; used to pin the grid path on
// inputs too large for brute force.
uint64_t orc_sweep2f(const float* boxes, uint32_t n, uint32_t* out, uint64_t cap, int threads) {
  const Box2* b = (const Box2*)boxes;
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) {
    if (b[p].xmin != b[q].xmin) return b[p].xmin < b[q].xmin;
    return p < q;
  });
  std::vector<Box2> s(n);
  for (uint32_t k = 0; k < n; ++k) s[k] = b[order[k]];
  int T = nthreads(threads);
  std::vector<std::vector<uint32_t>> parts(T);
#pragma omp parallel num_threads(T)
  {
    auto& mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1024)
    for (int64_t k = 0; k < (int64_t)n; ++k) {
      const Box2 a = s[k];
      for (uint32_t m = (uint32_t)k + 1; m < n && s[m].xmin <= a.xmax; ++m)
        if (overlap2(a, s[m])) { mine.push_back(order[k]); mine.push_back(order[m]); }
    }
  }
  PairSink sink{out, cap, 0};
  for (auto& p : parts)
    for (size_t k = 0; k < p.size(); k += 2) sink.push(p[k], p[k + 1]);
  return sink.n;
}

// ---------------------------------------------------------------------------
// Reference cell-size rule, pbig.cu:1285-1357 (decideCellSize), restated for
// float (dims=2) boxes.  The mean extent is accumulated the way the
// reference's thrust::transform_reduce call does it (pbig.cu:1303-1308): the
// init value is int 0, so the accumulator is an int, each step is
// (int)((float)acc + extent), and the sum is divided by nSamples in integer
// arithmetic before it is stored in a float.  The IS_AVAIL_MEM early exit
// (pbig.cu:1334-1341) is omitted: memory is never the limit here.
// ---------------------------------------------------------------------------
uint32_t orc_decide_cell_size2f(const float* boxes, uint32_t n, double range) {
  const Box2* b = (const Box2*)boxes;
  int64_t cs = 1;
  while ((double)(cs * 2) < range) cs *= 2;
  int nSamples = n < 1000 ? (int)n : 1000;
  if (nSamples == 0) return (uint32_t)cs;
  int ax = 0, ay = 0;
  for (int k = 0; k < nSamples; ++k) {
    ax = (int)((float)ax + (b[k].xmax - b[k].xmin));
    ay = (int)((float)ay + (b[k].ymax - b[k].ymin));
  }
  float x = (float)(ax / nSamples), y = (float)(ay / nSamples);
  while (true) {
    int64_t nrc = (int64_t)std::ceil(range / (double)cs);
    int64_t nc = nrc * nrc;
    double nmr = (double)((1 + x / (float)cs) * (1 + y / (float)cs));
    int nRecsPerCell = (int)(nmr * (double)n / (double)nc);
    if ((float)cs < x) break;
    if (nRecsPerCell < 32 || nmr > 4) break;  // NMR_BOUND = 4 in 2-D (pbig.cu:1280-1281)
    if (cs == 1) break;
    cs /= 2;
  }
  return (uint32_t)(cs * 2);
}

// ---------------------------------------------------------------------------
// Grid restatement for 2-D float boxes:
//   count     pbig.cu:174-221  (d_precounting: inclusive cell range per box)
//   scan      pbig.cu:1497-1505 (exclusive scan of the counts)
//   scatter   pbig.cu:321-427  (d_mapping: rid into the cell's list)
//   test      pbig.cu:642-932  (d_matching: every unordered pair in a cell,
//             reported iff calcOverlaps == 4 (pbig.cu:562-568) and the
//             reference cell, cell(max xmin), cell(max ymin) == this cell
//             (calcRightToWrite, pbig.cu:579-603))
// Cells are row-major, gid = nrc*cy + cx (pbig.cu:186-196).  Scatter is done in
// box-id order, so the cell lists are deterministic here (the reference's
// atomicAdd order is not; only the pair set matters).
// ---------------------------------------------------------------------------
uint64_t orc_grid2f(const float* boxes, uint32_t n, double range, uint32_t cs,
                    uint32_t* out, uint64_t cap, int threads, orc_grid_stats* st) {
  const Box2* b = (const Box2*)boxes;
  const uint32_t nrc = (uint32_t)std::max<double>(1.0, std::ceil(range / (double)cs));
  const uint64_t nCells = (uint64_t)nrc * nrc;
  const float inv = 1.0f / (float)cs;
  std::vector<uint32_t> cnt(nCells, 0);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t x0 = cellmap(b[i].xmin, inv, nrc), x1 = cellmap(b[i].xmax, inv, nrc);
    uint32_t y0 = cellmap(b[i].ymin, inv, nrc), y1 = cellmap(b[i].ymax, inv, nrc);
    for (uint32_t cy = y0; cy <= y1; ++cy)
      for (uint32_t cx = x0; cx <= x1; ++cx) cnt[(uint64_t)nrc * cy + cx]++;
  }
  std::vector<uint64_t> start(nCells + 1, 0);
  for (uint64_t c = 0; c < nCells; ++c) start[c + 1] = start[c] + cnt[c];
  const uint64_t M = start[nCells];
  std::vector<uint32_t> recs(M);
  std::vector<uint64_t> cur(start.begin(), start.end() - 1);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t x0 = cellmap(b[i].xmin, inv, nrc), x1 = cellmap(b[i].xmax, inv, nrc);
    uint32_t y0 = cellmap(b[i].ymin, inv, nrc), y1 = cellmap(b[i].ymax, inv, nrc);
    for (uint32_t cy = y0; cy <= y1; ++cy)
      for (uint32_t cx = x0; cx <= x1; ++cx) recs[cur[(uint64_t)nrc * cy + cx]++] = i;
  }
  int T = nthreads(threads);
  std::vector<std::vector<uint32_t>> parts(T);
#pragma omp parallel num_threads(T)
  {
    auto& mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 16)
    for (int64_t c = 0; c < (int64_t)nCells; ++c) {
      const uint64_t s = start[c], e = start[c + 1];
      if (e - s < 2) continue;
      const uint32_t cx = (uint32_t)(c % nrc), cy = (uint32_t)(c / nrc);
      for (uint64_t p = s; p < e; ++p) {
        const Box2 a = b[recs[p]];
        for (uint64_t q = p + 1; q < e; ++q) {
          const Box2 o = b[recs[q]];
          if (!overlap2(a, o)) continue;
          uint32_t fx = cellmap(std::max(a.xmin, o.xmin), inv, nrc);
          uint32_t fy = cellmap(std::max(a.ymin, o.ymin), inv, nrc);
          if (fx == cx && fy == cy) { mine.push_back(recs[p]); mine.push_back(recs[q]); }
        }
      }
    }
  }
  PairSink sink{out, cap, 0};
  for (auto& p : parts)
    for (size_t k = 0; k < p.size(); k += 2) sink.push(p[k], p[k + 1]);
  if (st) {
    memset(st, 0, sizeof(*st));
    st->n_cells = nCells;
    st->memberships = M;
    for (uint64_t c = 0; c < nCells; ++c) {
      uint64_t k = cnt[c];
      if (k > st->max_per_cell) { st->max_per_cell = k; st->max_cell = c; }
      if (k > 0) st->nonempty_cells++;
      if (k > 1) { st->compared_pairs += k * k / 2; st->jobs256 += (k + 255) / 256; }
    }
    st->pairs = sink.n;
  }
  return sink.n;
}

// Binning statistics only (no pair test): per-cell counts at a given cell size.
void orc_grid2f_stats(const float* boxes, uint32_t n, double range, uint32_t cs, orc_grid_stats* st) {
  const Box2* b = (const Box2*)boxes;
  const uint32_t nrc = (uint32_t)std::max<double>(1.0, std::ceil(range / (double)cs));
  const uint64_t nCells = (uint64_t)nrc * nrc;
  const float inv = 1.0f / (float)cs;
  std::vector<uint32_t> cnt(nCells, 0);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t x0 = cellmap(b[i].xmin, inv, nrc), x1 = cellmap(b[i].xmax, inv, nrc);
    uint32_t y0 = cellmap(b[i].ymin, inv, nrc), y1 = cellmap(b[i].ymax, inv, nrc);
    for (uint32_t cy = y0; cy <= y1; ++cy)
      for (uint32_t cx = x0; cx <= x1; ++cx) cnt[(uint64_t)nrc * cy + cx]++;
  }
  memset(st, 0, sizeof(*st));
  st->n_cells = nCells;
  for (uint64_t c = 0; c < nCells; ++c) {
    uint64_t k = cnt[c];
    st->memberships += k;
    if (k > st->max_per_cell) { st->max_per_cell = k; st->max_cell = c; }
    if (k > 0) st->nonempty_cells++;
    if (k > 1) { st->compared_pairs += k * k / 2; st->jobs256 += (k + 255) / 256; }
  }
}

// Per-cell counts and exclusive starts at a given cell size (binning parity).
void orc_grid2f_counts(const float* boxes, uint32_t n, double range, uint32_t cs, uint32_t* cnt_out) {
  const Box2* b = (const Box2*)boxes;
  const uint32_t nrc = (uint32_t)std::max<double>(1.0, std::ceil(range / (double)cs));
  const uint64_t nCells = (uint64_t)nrc * nrc;
  const float inv = 1.0f / (float)cs;
  memset(cnt_out, 0, nCells * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t x0 = cellmap(b[i].xmin, inv, nrc), x1 = cellmap(b[i].xmax, inv, nrc);
    uint32_t y0 = cellmap(b[i].ymin, inv, nrc), y1 = cellmap(b[i].ymax, inv, nrc);
    for (uint32_t cy = y0; cy <= y1; ++cy)
      for (uint32_t cx = x0; cx <= x1; ++cx) cnt_out[(uint64_t)nrc * cy + cx]++;
  }
}

uint32_t orc_cellmap(float v, float inv, uint32_t nrc) { return cellmap(v, inv, nrc); }

// ---------------------------------------------------------------------------
// Reference-compatible 3-D unsigned mode (the reference's shipped build,
// DIMS=3, CType=unsigned int).  Used only for the reference's logged answers.
// ---------------------------------------------------------------------------

// Generator of src/pbig/main.cc:204-335 (InitTestPerformance): srand(100),
// then per box the three sizes, then the three positions, from glibc rand().
// sim[0] picks the size law ('U': rand()%S, main.cc:204-208; 'N': normal
// (S, 0.2*S), main.cc:210-218), sim[1] the position law ('U': rand()%range,
// 'N': normal(range/2, 0.2*range/2)).  xmax = Inc(pos, size, range)
// (main.cc:223, 313), where the int sum is compared against the unsigned
// range, so a negative sum clamps to range-1.  Output: n x {xmin,xmax,ymin,
// ymax,zmin,zmax} as unsigned (CType, rect.h:4).
static double ref_uniform(unsigned max) { return rand() % max; }
static double ref_normal(unsigned mean, unsigned sd) {
  double u = rand() / (double)RAND_MAX;
  double v = rand() / (double)RAND_MAX;
  return sqrt(-2 * log(u)) * cos(2 * M_PI * v) * sd + mean;
}
void orc_ref_generate(uint32_t n, uint32_t range, uint32_t S, const char* sim, int dims, uint32_t* out) {
  srand(100);
  const double STDRATIO = 0.2;
  for (uint32_t i = 0; i < n; ++i) {
    int sz[3] = {0, 0, 0}, pos[3] = {0, 0, 0};
    for (int d = 0; d < dims; ++d)
      sz[d] = sim[0] == 'U' ? (int)ref_uniform(S) : (int)ref_normal(S, (unsigned)(S * STDRATIO));
    for (int d = 0; d < dims; ++d)
      pos[d] = sim[1] == 'U' ? (int)ref_uniform(range)
                             : (int)ref_normal(range / 2, (unsigned)((range / 2) * STDRATIO));
    for (int d = 0; d < dims; ++d) {
      uint32_t lo = (uint32_t)pos[d];
      uint32_t hi = ((unsigned)(pos[d] + sz[d]) >= range) ? range - 1 : (uint32_t)(pos[d] + sz[d]);
      out[(size_t)i * 2 * dims + 2 * d] = lo;
      out[(size_t)i * 2 * dims + 2 * d + 1] = hi;
    }
  }
}

// decideCellSize (pbig.cu:1285-1357) for unsigned boxes of `dims` dimensions.
// The extent is computed in unsigned then converted to float (avgX,
// pbig.cu:1262-1268); accumulation as in orc_decide_cell_size2f.
uint32_t orc_decide_cell_size_u32(const uint32_t* boxes, uint32_t n, int dims, uint32_t range) {
  int64_t cs = 1;
  while (cs * 2 < (int64_t)range) cs *= 2;
  int nSamples = n < 1000 ? (int)n : 1000;
  int acc[3] = {0, 0, 0};
  for (int k = 0; k < nSamples; ++k)
    for (int d = 0; d < dims; ++d) {
      const uint32_t* bx = boxes + (size_t)k * 2 * dims;
      acc[d] = (int)((float)acc[d] + (float)(bx[2 * d + 1] - bx[2 * d]));
    }
  float m[3];
  for (int d = 0; d < 3; ++d) m[d] = nSamples ? (float)(acc[d] / nSamples) : 0.0f;
  const double bound = dims == 2 ? 4 : 8;  // NMR_BOUND, pbig.cu:1280-1284
  while (true) {
    int64_t nrc = ((int64_t)range + cs - 1) / cs;
    int64_t nc = nrc * nrc;
    // nmr is a double assigned from a float product; the third factor
    // multiplies the double (pbig.cu:1320-1325).
    double nmr = (double)((1 + m[0] / (float)cs) * (1 + m[1] / (float)cs));
    if (dims == 3) { nc *= nrc; nmr *= (double)(1 + m[2] / (float)cs); }
    int nRecsPerCell = (int)(nmr * (double)n / (double)nc);
    if ((float)cs < m[0]) break;
    if (nRecsPerCell < 32 || nmr > bound) break;
    if (cs == 1) break;
    cs /= 2;
  }
  return (uint32_t)(cs * 2);
}

// Grid count for unsigned boxes, restating the reference literally:
// f_l = xmin >> shift ... (pbig.cu:185-216), gid in unsigned 32-bit arithmetic,
// `gid >= nCells` skipped, int loop counters; per-cell test = calcOverlaps ==
// 2*dims and calcRightToWrite (pbig.cu:562-603) with max() of unsigned mins.
// Returns the number of reported pairs; fills stats.
uint64_t orc_grid_u32(const uint32_t* boxes, uint32_t n, int dims, uint32_t range, uint32_t cs,
                      int threads, orc_grid_stats* st) {
  uint32_t shift = 0;
  while ((1u << shift) < cs) ++shift;
  const uint32_t nrc = (range + cs - 1) / cs;
  const uint32_t nrc2 = nrc * nrc;
  const uint32_t nCells = dims == 2 ? nrc2 : nrc2 * nrc;
  const int D2 = 2 * dims;
  auto cells_of = [&](const uint32_t* bx, auto&& fn) {
    uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < dims; ++d) { lo[d] = bx[2 * d] >> shift; hi[d] = bx[2 * d + 1] >> shift; }
    if (dims == 2) {
      for (int j = (int)lo[1]; j <= (int)hi[1]; ++j) {
        uint32_t product = nrc * (uint32_t)j;
        for (int i = (int)lo[0]; i <= (int)hi[0]; ++i) {
          uint32_t gid = product + (uint32_t)i;
          if (gid >= nCells) continue;
          fn(gid);
        }
      }
    } else {
      for (int k = (int)lo[2]; k <= (int)hi[2]; ++k) {
        uint32_t product = nrc2 * (uint32_t)k;
        for (int j = (int)lo[1]; j <= (int)hi[1]; ++j) {
          uint32_t m_gid = product + nrc * (uint32_t)j;
          for (int i = (int)lo[0]; i <= (int)hi[0]; ++i) {
            uint32_t gid = m_gid + (uint32_t)i;
            if (gid >= nCells) continue;
            fn(gid);
          }
        }
      }
    }
  };
  std::vector<uint32_t> cnt(nCells, 0);
  for (uint32_t r = 0; r < n; ++r) cells_of(boxes + (size_t)r * D2, [&](uint32_t g) { cnt[g]++; });
  std::vector<uint64_t> start(nCells + 1, 0);
  for (uint32_t c = 0; c < nCells; ++c) start[c + 1] = start[c] + cnt[c];
  std::vector<uint32_t> recs(start[nCells]);
  std::vector<uint64_t> cur(start.begin(), start.end() - 1);
  for (uint32_t r = 0; r < n; ++r)
    cells_of(boxes + (size_t)r * D2, [&](uint32_t g) { recs[cur[g]++] = r; });
  int T = nthreads(threads);
  uint64_t total = 0;
#pragma omp parallel for num_threads(T) schedule(dynamic, 64) reduction(+ : total)
  for (int64_t c = 0; c < (int64_t)nCells; ++c) {
    const uint64_t s = start[c], e = start[c + 1];
    uint64_t found = 0;
    for (uint64_t p = s; p < e; ++p) {
      const uint32_t* a = boxes + (size_t)recs[p] * D2;
      for (uint64_t q = p + 1; q < e; ++q) {
        const uint32_t* o = boxes + (size_t)recs[q] * D2;
        bool ov = true;
        for (int d = 0; d < dims; ++d) ov = ov && (a[2 * d] <= o[2 * d + 1]) && (o[2 * d] <= a[2 * d + 1]);
        if (!ov) continue;
        uint32_t f[3] = {0, 0, 0};
        for (int d = 0; d < dims; ++d) f[d] = std::max(a[2 * d], o[2 * d]) >> shift;
        uint32_t first = dims == 2 ? nrc * f[1] + f[0] : nrc2 * f[2] + nrc * f[1] + f[0];
        found += (first == (uint32_t)c);
      }
    }
    total += found;
  }
  if (st) {
    memset(st, 0, sizeof(*st));
    st->n_cells = nCells;
    st->memberships = start[nCells];
    for (uint32_t c = 0; c < nCells; ++c) {
      uint64_t k = cnt[c];
      if (k > st->max_per_cell) { st->max_per_cell = k; st->max_cell = c; }
      if (k > 0) st->nonempty_cells++;
      if (k > 1) { st->compared_pairs += k * k / 2; st->jobs256 += (k + 255) / 256; }
    }
    st->pairs = total;
  }
  return total;
}

// All-pairs brute force for unsigned boxes of `dims` dimensions with the
// reference's predicate (every dimension's two unsigned <= tests,
// pbig.cu:562-577); the pair set any correct grid run must report.  O(n^2).
uint64_t orc_brute_u32(const uint32_t* boxes, uint32_t n, int dims, uint32_t* out, uint64_t cap, int threads) {
  const int D2 = 2 * dims;
  int T = nthreads(threads);
  std::vector<std::vector<uint32_t>> parts(T);
#pragma omp parallel num_threads(T)
  {
    auto& mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const uint32_t* a = boxes + (size_t)i * D2;
      for (uint32_t j = (uint32_t)i + 1; j < n; ++j) {
        const uint32_t* o = boxes + (size_t)j * D2;
        bool ov = true;
        for (int d = 0; d < dims; ++d) ov = ov && (a[2 * d] <= o[2 * d + 1]) && (o[2 * d] <= a[2 * d + 1]);
        if (ov) { mine.push_back((uint32_t)i); mine.push_back(j); }
      }
    }
  }
  return gather(parts, out, cap);
}

}  // extern "C"
