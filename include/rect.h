/*
 * rect.h -- box record of the reference API (src/pbig/rect.h:4-18), 2-D float32.
 *
 * CType is float (the reference ships unsigned int coordinates; its 2-D mode,
 * DIMS == 2, has the same field order).  RectInfo stays a 20-byte
 * array-of-structures record and rid must equal the record's index, as every
 * reference driver sets it (src/pbig/main.cc:327, package/main.cc:27).
 */
#ifndef __RECT_H_
#define __RECT_H_

typedef float CType;

typedef struct CoordInfo {
  CType xmin, xmax;
  CType ymin, ymax;
} CoordInfo;

typedef struct RectInfo {
  CoordInfo coordinate;
  unsigned int rid;
} RectInfo;

#endif
