/*
 * oracle.c -- CPU restatement of the ten open-earth-benchmarks stencils.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of it.
 * Built by oracle/Makefile with -O2 -ffp-contract=off.
 *
 * The stencils are written once, in oracle_body.h, for an element type R.
 * This file instantiates them twice: R = double gives oracle_<name> (the
 * FP64 oracle) and R = float gives oracle_<name>_f32 (the FP32 oracle: the
 * reference's float build, plus the FP64 islands and the compensated fvtp2d
 * divergence described in oracle_body.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

/* Element (i,j,k) of a 3-D descriptor, as Storage::operator() in
 * include/util.h:84-94 of the reference. */
#define AT(s, i, j, k)                                                                             \
  ((s)->alignedPtr[(s)->offset + (int64_t)(i) * (s)->strides[2] + (int64_t)(j) * (s)->strides[1] + \
                   (int64_t)(k) * (s)->strides[0]])
/* Element j of a 1-D descriptor (include/util.h:56-64). */
#define AT1(s, j) ((s)->alignedPtr[(s)->offset + (int64_t)(j) * (s)->strides[0]])

#define FOR_INTERIOR(d)                          \
  for (int64_t k = 0; k < (d)->nz; ++k)          \
    for (int64_t j = 0; j < (d)->ny; ++j)        \
      for (int64_t i = 0; i < (d)->nx; ++i)

/* FP64 instantiation: oracle_<name>. */
#define R double
#define S3 oeb_s3
#define S1 oeb_s1
#define ORC(n) oracle_##n
#define LOC(n) f64_##n
#define OEB_ORACLE_R_IS_FLOAT 0
#include "oracle_body.h"
#undef R
#undef S3
#undef S1
#undef ORC
#undef LOC
#undef OEB_ORACLE_R_IS_FLOAT

/* FP32 instantiation: oracle_<name>_f32. */
#define R float
#define S3 oeb_s3f
#define S1 oeb_s1f
#define ORC(n) oracle_##n##_f32
#define LOC(n) f32_##n
#define OEB_ORACLE_R_IS_FLOAT 1
#include "oracle_body.h"
#undef R
#undef S3
#undef S1
#undef ORC
#undef LOC
#undef OEB_ORACLE_R_IS_FLOAT
