/*
 * report.h -- result sink of the reference API (src/pbig/report.h:6-16).
 *
 * pbig_check() calls operator()(rid1, rid2) once per overlapping unordered
 * pair, on the caller's thread, after the device has finished; rid1 < rid2.
 *
 * Layout.  With PBIG_COMPAT_U32 (the reference's shipped unsigned build,
 * libpbig_u32.so) the object is the reference's: one int counter, 4 bytes,
 * so a caller compiled against the reference's own report.h can link to
 * that library.  The native float build (libpbig.so) widens the counter to
 * 64 bits, because the reference's int wraps above 2^31-1 pairs
 * (src/pbig/report.h:15); callers of libpbig.so compile against this
 * header.
 */
#ifndef __REPORT_H
#define __REPORT_H

#include "rect.h"

struct report {
 public:
  report();
  void operator()(unsigned int a, unsigned int b) { (void)a; (void)b; sum++; }
  void found(const RectInfo *a, const RectInfo *b);
  void print();
  void reset() { sum = 0; }
#if defined(PBIG_COMPAT_U32)
  unsigned long long count() const { return (unsigned int)sum; }

 private:
  int sum;
#else
  unsigned long long count() const { return sum; }

 private:
  unsigned long long sum;
#endif
};

#if defined(__cplusplus) && __cplusplus >= 201103L
#if defined(PBIG_COMPAT_U32)
static_assert(sizeof(report) == sizeof(int), "report keeps the reference's layout (one int)");
#else
static_assert(sizeof(report) == sizeof(unsigned long long), "report holds one 64-bit counter");
#endif
#endif

#endif
