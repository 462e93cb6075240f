// Compressed pair stream in the reference's record layout (its default build,
// COMPRESS = 2: RICE_COMPRESSION with 64-bit words, src/pbig/CMakeLists.txt:36-38,
// pbig_internal.h:4-6), host side.
//
// Stream: blocks of [UNUSED, gid] (UNUSED = 0x0fffff11, cuda_comm.h:10;
// prefix_scan, pbig.cu:519-527) each followed by records of cell gid.  A
// record is a byte string padded to whole 64-bit words (pbig.cu:849-888):
//   byte 0     seg1   segment of the row box in the cell's list (256 per segment)
//   byte 1     seg2   segment of the partners
//   bytes 2-3  l_rid  row box's index inside its segment (u16)
//   bytes 4-5  nlaps  number of partners (u16, >= 1)
//   bytes 6..  one byte per partner: its index inside segment seg2
// The pair is (recs[start[gid] + 256 seg1 + l_rid], recs[start[gid] + 256 seg2 + t])
// (decode.cc:52-86).  The decoder here reads the partner bytes from byte 6 of
// the record, where the encoder writes them; the reference's decoder reads
// them from byte 10 (decode.cc:80 vs pbig.cu:851) and so reports wrong
// partners.
//
// The encoder builds cells of its own from a pair list: one cell per box with
// partners (the box, then up to 65,535 of its partners; more partners open
// further cells), so any pair set round-trips.  About one byte per pair in
// the stream plus four per pair in the cell lists.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pbig_hip.h"

namespace {

constexpr uint64_t kUnused = 0x0fffff11ull;
constexpr uint32_t kSeg = 256;              // entries per segment (NTPB = 256)
constexpr uint32_t kCellPartners = 65535;   // 256 segments x 256 entries, minus the row box

}  // namespace

extern "C" {

int pbig_hip_rice_encode(const uint32_t* pairs, uint64_t n_pairs, uint32_t n_boxes, uint64_t* words,
                         uint64_t cap_words, uint64_t* n_words, uint32_t* cell_start, uint64_t cap_cells,
                         uint64_t* n_cells, uint32_t* cell_recs, uint64_t cap_recs, uint64_t* n_recs) {
  if (!n_words || !n_cells || !n_recs) return PBIG_E_ARG;
  if (n_pairs && !pairs) return PBIG_E_ARG;
  // group the partners by the smaller id (counting sort), each list sorted
  std::vector<uint64_t> first((size_t)n_boxes + 1, 0);
  for (uint64_t k = 0; k < n_pairs; ++k) {
    const uint32_t a = std::min(pairs[2 * k], pairs[2 * k + 1]), b = std::max(pairs[2 * k], pairs[2 * k + 1]);
    if (b >= n_boxes || a == b) return PBIG_E_ARG;
    first[a + 1]++;
  }
  for (uint32_t a = 0; a < n_boxes; ++a) first[a + 1] += first[a];
  std::vector<uint32_t> part(n_pairs);
  {
    std::vector<uint64_t> cur(first.begin(), first.end() - 1);
    for (uint64_t k = 0; k < n_pairs; ++k) {
      const uint32_t a = std::min(pairs[2 * k], pairs[2 * k + 1]), b = std::max(pairs[2 * k], pairs[2 * k + 1]);
      part[cur[a]++] = b;
    }
  }
  // sizes
  uint64_t nw = 0, nc = 0, nr = 0;
  for (uint32_t a = 0; a < n_boxes; ++a) {
    const uint64_t np = first[a + 1] - first[a];
    for (uint64_t o = 0; o < np; o += kCellPartners) {
      const uint64_t len = std::min<uint64_t>(kCellPartners, np - o);
      nc++;
      nr += 1 + len;
      nw += 2;  // [UNUSED, gid]
      // positions 1..len in segments of 256: segment s holds positions [256 s, 256 s + 255]
      for (uint64_t s = 0; s * kSeg <= len; ++s) {
        const uint64_t lo = std::max<uint64_t>(1, s * kSeg), hi = std::min<uint64_t>(len, s * kSeg + kSeg - 1);
        if (hi < lo) continue;
        nw += (6 + (hi - lo + 1) + 7) / 8;
      }
    }
  }
  *n_words = nw;
  *n_cells = nc;
  *n_recs = nr;
  if (cap_words < nw || cap_cells < nc + 1 || cap_recs < nr || (nw && !words) || !cell_start || (nr && !cell_recs))
    return PBIG_E_OVERFLOW;
  // write
  uint64_t w = 0, c = 0, r = 0;
  for (uint32_t a = 0; a < n_boxes; ++a) {
    const uint64_t np = first[a + 1] - first[a];
    for (uint64_t o = 0; o < np; o += kCellPartners) {
      const uint64_t len = std::min<uint64_t>(kCellPartners, np - o);
      std::sort(part.begin() + first[a] + o, part.begin() + first[a] + o + len);
      cell_start[c] = (uint32_t)r;
      cell_recs[r] = a;
      for (uint64_t p = 0; p < len; ++p) cell_recs[r + 1 + p] = part[first[a] + o + p];
      words[w++] = kUnused;
      words[w++] = c;
      for (uint64_t s = 0; s * kSeg <= len; ++s) {
        const uint64_t lo = std::max<uint64_t>(1, s * kSeg), hi = std::min<uint64_t>(len, s * kSeg + kSeg - 1);
        if (hi < lo) continue;
        const uint64_t nl = hi - lo + 1, nwr = (6 + nl + 7) / 8;
        unsigned char* b = (unsigned char*)(words + w);
        memset(b, 0, nwr * 8);
        b[0] = 0;                 // seg1: the row box is entry 0 of the cell
        b[1] = (unsigned char)s;  // seg2
        const uint16_t l_rid = 0, nlaps = (uint16_t)nl;
        memcpy(b + 2, &l_rid, 2);
        memcpy(b + 4, &nlaps, 2);
        for (uint64_t q = 0; q < nl; ++q) b[6 + q] = (unsigned char)((lo + q) & (kSeg - 1));
        w += nwr;
      }
      r += 1 + len;
      c++;
    }
  }
  cell_start[c] = (uint32_t)r;
  return PBIG_OK;
}

int pbig_hip_rice_decode(const uint64_t* words, uint64_t n_words, const uint32_t* cell_start, uint64_t n_cells,
                         const uint32_t* cell_recs, uint64_t n_recs, uint32_t* pairs, uint64_t cap,
                         uint64_t* n_pairs) {
  if (!n_pairs) return PBIG_E_ARG;
  if (n_words && (!words || !cell_start || !cell_recs)) return PBIG_E_ARG;
  if (n_words && (n_words < 3 || words[0] != kUnused)) return PBIG_E_ARG;  // a stream opens with a block
  const unsigned char* bytes = (const unsigned char*)words;
  uint64_t found = 0;
  uint64_t i = 2;
  while (i < n_words) {
    if (words[i - 2] != kUnused) return PBIG_E_ARG;  // not at a block header
    const uint64_t gid = words[i - 1];
    if (gid >= n_cells) return PBIG_E_ARG;
    const uint64_t base = cell_start[gid];
    do {
      if (i >= n_words) return PBIG_E_ARG;
      const unsigned char* rec = bytes + 8 * i;
      const uint32_t seg1 = rec[0], seg2 = rec[1];
      uint16_t l_rid, nlaps;
      memcpy(&l_rid, rec + 2, 2);
      memcpy(&nlaps, rec + 4, 2);
      const uint64_t nwr = (6 + (uint64_t)nlaps + 7) / 8;
      if (nlaps == 0 || i + nwr > n_words) return PBIG_E_ARG;
      const uint64_t p1 = base + (uint64_t)seg1 * kSeg + l_rid;
      if (p1 >= n_recs) return PBIG_E_ARG;
      const uint32_t rid1 = cell_recs[p1];
      for (uint32_t k = 0; k < nlaps; ++k) {
        const uint64_t p2 = base + (uint64_t)seg2 * kSeg + rec[6 + k];
        if (p2 >= n_recs) return PBIG_E_ARG;
        const uint32_t rid2 = cell_recs[p2];
        if (found < cap) {
          pairs[2 * found] = rid1 < rid2 ? rid1 : rid2;
          pairs[2 * found + 1] = rid1 < rid2 ? rid2 : rid1;
        }
        ++found;
      }
      i += nwr;
    } while (i < n_words && words[i] != kUnused);
    i += 2;  // past the next block's header
  }
  *n_pairs = found;
  return found > cap ? PBIG_E_OVERFLOW : PBIG_OK;
}

}  // extern "C"
