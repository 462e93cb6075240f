// Reference-compatible C++ API (include/pbig.h, report.h, rect.h) on top of
// the C-ABI of include/pbig_hip.h.
//
//   pbig_init      src/pbig/pbig.cu:2139-2217 (globals, device, pools)
//   pbig_check     src/pbig/pbig.cu:1962-1981 (plan, map, schedule, match,
//                  decode into report callbacks, decode.cc:26-138)
//   pbig_finalize  src/pbig/pbig.cu:141
//   report         src/pbig/report.cc:6-8
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pbig.h"
#include "pbig_hip.h"
#include "rect.h"
#include "report.h"

namespace {

pbig_hip_ctx* g_ctx = nullptr;
unsigned int g_n = 0;
unsigned int g_range = 0;
unsigned int g_cs = 0;
unsigned long long g_last_pairs = 0;  // pairs of the last pbig_check

[[noreturn]] void die(const char* what, int rc) {
  printf("ERROR:%s failed (status %d): %s\n", what, rc, g_ctx ? pbig_hip_last_error(g_ctx) : "no context");
  if (g_ctx) pbig_hip_destroy(g_ctx);
  g_ctx = nullptr;
  exit(-1);
}

}  // namespace

report::report() : sum(0) {}
#if defined(PBIG_COMPAT_U32)
void report::print() { printf("Number of pairs = %d\n", sum); }
#else
void report::print() { printf("Number of pairs = %llu\n", sum); }
#endif
void report::found(const RectInfo* a, const RectInfo* b) {
  (void)a;
  (void)b;
  sum++;
}

void pbig_init(unsigned int num, unsigned int range, unsigned int cs, unsigned int deviceID,
               unsigned int hasProfSteps) {
  (void)hasProfSteps;
  if (g_ctx) pbig_finalize();
  if (cs != 0 && (cs & (cs - 1)) != 0) {
    printf("ERROR:cell size %u is not a power of two\n", cs);
    exit(-1);
  }
  int rc = pbig_hip_create(&g_ctx, (int)deviceID);
  if (rc) die("pbig_init", rc);
  g_n = num;
  g_range = range;
  g_cs = cs;
  rc = pbig_hip_reserve(g_ctx, num, 1, (unsigned long long)num + num / 2);
  if (rc) die("pbig_init", rc);
}

// One host entry point per build of rect.h (the native float 2-D records, or
// the reference's unsigned records, DIMS 2 or 3): the pairs come back chunk by
// chunk to fn while the next chunk is copied.
int check_each(RectInfo* input, pbig_hip_pairs_fn fn, void* user, unsigned long long* np) {
#if defined(PBIG_COMPAT_U32)
  return pbig_hip_check_host_u32_each(g_ctx, input, g_n, DIMS, g_range, g_cs, fn, user, (uint64_t*)np);
#else
  return pbig_hip_check_host_each(g_ctx, input, g_n, (double)g_range, (float)g_cs, fn, user, (uint64_t*)np);
#endif
}

void pbig_check(RectInfo* input, report* r) {
  if (!g_ctx) {
    printf("ERROR:pbig_check called before pbig_init\n");
    exit(-1);
  }
  // the report's callbacks run on each chunk of pairs while the next one is
  // copied back (the reference decodes into the report during its prefetch)
  unsigned long long np = 0;
  const int rc = check_each(
      input,
      [](void* user, const uint32_t* p, uint64_t n) {
        report& rep = *static_cast<report*>(user);
        for (uint64_t k = 0; k < n; ++k) rep(p[2 * k], p[2 * k + 1]);
      },
      r, &np);
  if (rc) die("pbig_check", rc);
  g_last_pairs = np;
}

void pbig_check_stream(RectInfo* input, std::vector<unsigned long long>* words, std::vector<unsigned int>* cell_start,
                       std::vector<unsigned int>* cell_recs) {
  // the pairs of the check, then the encoder (sizes first)
  std::vector<unsigned int> flat;
  {
    unsigned long long np = 0;
    const int rc = check_each(
        input,
        [](void* user, const uint32_t* p, uint64_t n) {
          auto& v = *static_cast<std::vector<unsigned int>*>(user);
          v.insert(v.end(), p, p + 2 * n);
        },
        &flat, &np);
    if (rc) die("pbig_check_stream", rc);
    g_last_pairs = np;
  }
  uint64_t nw = 0, nc = 0, nr = 0;
  const uint64_t np = flat.size() / 2;
  int rc = pbig_hip_rice_encode(flat.data(), np, g_n, nullptr, 0, &nw, nullptr, 0, &nc, nullptr, 0, &nr);
  if (rc != PBIG_OK && rc != PBIG_E_OVERFLOW) die("pbig_check_stream", rc);
  words->assign(nw, 0ull);
  cell_start->assign(nc + 1, 0u);
  cell_recs->assign(nr, 0u);
  rc = pbig_hip_rice_encode(flat.data(), np, g_n, (uint64_t*)words->data(), nw, &nw, cell_start->data(), nc + 1, &nc,
                            cell_recs->data(), nr, &nr);
  if (rc) die("pbig_check_stream", rc);
}

void pbig_decode_stream(const std::vector<unsigned long long>& words, const std::vector<unsigned int>& cell_start,
                        const std::vector<unsigned int>& cell_recs, report* r) {
  uint64_t np = 0;
  const uint64_t n_cells = cell_start.empty() ? 0 : cell_start.size() - 1;
  int rc = pbig_hip_rice_decode((const uint64_t*)words.data(), words.size(), cell_start.data(), n_cells,
                                cell_recs.data(), cell_recs.size(), nullptr, 0, &np);
  if (rc != PBIG_OK && rc != PBIG_E_OVERFLOW) die("pbig_decode_stream", rc);
  std::vector<unsigned int> pairs(2 * np + 2);
  rc = pbig_hip_rice_decode((const uint64_t*)words.data(), words.size(), cell_start.data(), n_cells,
                            cell_recs.data(), cell_recs.size(), pairs.data(), np, &np);
  if (rc) die("pbig_decode_stream", rc);
  for (uint64_t k = 0; k < np; ++k) (*r)(pairs[2 * k], pairs[2 * k + 1]);
}

void pbig_finalize() {
  if (g_ctx) pbig_hip_destroy(g_ctx);
  g_ctx = nullptr;
}
