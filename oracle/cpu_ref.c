/*
 * Test infrastructure only: the oracle (see cpu_ref.h).  Instantiates
 * cpu_ref_impl.h for the reference's two kinds and for two wider formats,
 * x87 extended (long double, 64-bit significand) and binary128 (__float128,
 * 113-bit significand, libquadmath), whose only use is to evaluate finite
 * differences of the FP64 map above its roundoff floor.
 */
#include <math.h>
#include <quadmath.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "cpu_ref.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* Rows of the grid loops run in parallel (every cell is computed alone; the
 * dt reduction is a max), which changes no result.  n > 0 sets the thread
 * count; returns the count that calls will use. */
int oracle_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* variable and face indices: serial_ref/m_constants.f90:11-14, 51-54 */
enum { ID = 1, IP = 2, IU = 3, IV = 4 };
enum { FACE_XMIN = 0, FACE_XMAX = 1, FACE_YMIN = 2, FACE_YMAX = 3 };

#define R double
#define S f64
#define SQRT sqrt
#define FABS fabs
#define LIT_F32(x) ((double)(x##f))
#define LIT_F64(x) (x)
#include "cpu_ref_impl.h"
#undef R
#undef S
#undef SQRT
#undef FABS
#undef LIT_F32
#undef LIT_F64

#define R float
#define S f32
#define SQRT sqrtf
#define FABS fabsf
#define LIT_F32(x) (x##f)
#define LIT_F64(x) (x##f)
#include "cpu_ref_impl.h"
#undef R
#undef S
#undef SQRT
#undef FABS
#undef LIT_F32
#undef LIT_F64

/* Extended precision: the same constants as the FP64 build (a default-real
 * literal is the FP32 value, an fp_kind literal the FP64 value), carried in
 * long double, so the map is the FP64 map's function evaluated with a 64-bit
 * significand. */
#define R long double
#define S f80
#define SQRT sqrtl
#define FABS fabsl
#define LIT_F32(x) ((long double)(x##f))
#define LIT_F64(x) ((long double)(x))
#include "cpu_ref_impl.h"
#undef R
#undef S
#undef SQRT
#undef FABS
#undef LIT_F32
#undef LIT_F64

/* binary128, the same constants again */
#define R __float128
#define S f128
#define SQRT sqrtq
#define FABS fabsq
#define LIT_F32(x) ((__float128)(x##f))
#define LIT_F64(x) ((__float128)(x))
#include "cpu_ref_impl.h"
