/*
 * rect.h -- box record of the reference API (src/pbig/rect.h:4-18).
 *
 * Default: CType is float and boxes are 2-D (the engine's native mode).
 * With PBIG_COMPAT_U32 defined, this is the reference's shipped build:
 * CType = unsigned int and DIMS = 3 unless DIMS is defined as 2
 * (src/pbig/config.cmake:8); link libpbig_u32.so (DIMS = 3) then.
 * RectInfo stays an array-of-structures record and rid must equal the
 * record's index, as every reference driver sets it (src/pbig/main.cc:327,
 * package/main.cc:27).
 */
#ifndef __RECT_H_
#define __RECT_H_

#if defined(PBIG_COMPAT_U32)
#ifndef DIMS
#define DIMS 3
#endif
typedef unsigned int CType;
#else
typedef float CType;
#endif

typedef struct CoordInfo {
  CType xmin, xmax;
  CType ymin, ymax;
#if defined(PBIG_COMPAT_U32) && DIMS == 3
  CType zmin, zmax;
#endif
} CoordInfo;

typedef struct RectInfo {
  CoordInfo coordinate;
  unsigned int rid;
} RectInfo;

#endif
