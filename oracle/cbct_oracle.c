/*
 * CPU oracle for the cone-beam CT operators -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, on the CPU, the algorithms of the reference operators:
 *   forward ray integral   csrc/src_cbct_forwardprojector_cuda.cu:55-339
 *   its adjoint            csrc/src_cbct_forwardprojector_cuda.cu:342-528
 *   weighted backprojector csrc/src_cbct_backprojector_cuda.cu:15-98
 *   its backward           csrc/src_cbct_backprojector_cuda.cu:101-295
 * in two instantiations of one core (oracle_core.inc):
 *   *_f32  the FP32 reference schedule: float arithmetic in a fixed operation
 *          order, t advanced by repeated float addition of the step;
 *   *_f64  the FP64 oracle: double arithmetic, sample k at t_min + k*step.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark load
 * this library, and only as a checker.  The product never links it.
 * Build: oracle/Makefile (gcc, OpenMP).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL float
#define SFX(name) name##_f32
#define RFMA(a, b, c) fmaf((a), (b), (c))
#define RSQRT sqrtf
#define RFLOOR floorf
#define RABS fabsf
#define RMIN fminf
#define BIG FLT_MAX
#define IDEAL_SCHEDULE 0
#include "oracle_core.inc"
#undef REAL
#undef SFX
#undef RFMA
#undef RSQRT
#undef RFLOOR
#undef RABS
#undef RMIN
#undef BIG
#undef IDEAL_SCHEDULE

#define REAL double
#define SFX(name) name##_f64
#define RFMA(a, b, c) fma((a), (b), (c))
#define RSQRT sqrt
#define RFLOOR floor
#define RABS fabs
#define RMIN fmin
#define BIG DBL_MAX
#define IDEAL_SCHEDULE 1
#include "oracle_core.inc"

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
