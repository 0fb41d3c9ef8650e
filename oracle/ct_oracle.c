/*
 * ct_oracle.c -- CPU oracle for the CT projector hot path (TEST INFRASTRUCTURE).
 *
 * Instantiates ct_oracle_impl.h in float ("compat", the FP32 restatement of the
 * reference OpenCL kernels) and double (the FP64 oracle).  Built by
 * oracle/Makefile into oracle/liboracle_ct.so.  Only tests, the smoke check and
 * the benchmark's CPU-baseline leg load it; the product never does.
 */
#include "ct_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CTOR_PI 3.14159265358979323846
#define CTOR_PI_2 1.57079632679489661923

/* GPU.cpp:564-582: if ss is a power of two the subsets are visited in
 * bit-reversed order (ss=16 -> 0,8,4,12,2,10,...), otherwise 0..ss-1. */
void ctor_subset_order(int ss, int *sub)
{
    int bits = 0;
    for (int i = 0; i < 32; i++) bits += (ss >> i) & 1;
    if (bits == 1) {
        int m1 = 0;
        sub[m1++] = 0;
        for (int i = ss, m2 = 1; i > 1; i /= 2, m2 *= 2)
            for (int j = 0; j < m2; j++) sub[m1++] = sub[j] + i / 2;
    } else {
        for (int i = 0; i < ss; i++) sub[i] = i;
    }
}

int ctor_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- float (compat) instantiation ---- */
#define R float
#define S(n) n##_f32
#define RSQRT sqrtf
#define RCOS cosf
#define RSIN sinf
#define RFLOOR floorf
#define RASIN asinf
#define RFABS fabsf
#define REXP expf
#define RFMIN fminf
#include "ct_oracle_impl.h"
#undef R
#undef S
#undef RSQRT
#undef RCOS
#undef RSIN
#undef RFLOOR
#undef RASIN
#undef RFABS
#undef REXP
#undef RFMIN

/* ---- double (FP64 oracle) instantiation ---- */
#define R double
#define S(n) n##_f64
#define RSQRT sqrt
#define RCOS cos
#define RSIN sin
#define RFLOOR floor
#define RASIN asin
#define RFABS fabs
#define REXP exp
#define RFMIN fmin
#include "ct_oracle_impl.h"
