/*
 * report.h -- result sink of the reference API (src/pbig/report.h:6-16).
 *
 * pbig_check() calls operator()(rid1, rid2) once per overlapping unordered
 * pair, on the caller's thread, after the device has finished; rid1 < rid2.
 * The counter is 64-bit (the reference's int counter wraps above 2^31-1
 * pairs, src/pbig/report.h:15).
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
  unsigned long long count() const { return sum; }

 private:
  unsigned long long sum;
};

#endif
