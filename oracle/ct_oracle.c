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
/* Subset order of GPU.cpp:564-582.  For ss = 2^m the list doubles m times:
 * with spacing `half` between the subsets placed so far, every placed entry is
 * followed (in a second copy of the list) by itself plus half the spacing.
 * Any other ss runs 0 .. ss - 1.  (The library forms the same order as the
 * m-bit reversal of the position; the two are checked against each other and
 * against literal lists.) */
void ctor_subset_order(int ss, int *sub)
{
    if (ss & (ss - 1)) {
        for (int i = 0; i < ss; i++) sub[i] = i;
        return;
    }
    int placed = 1;
    sub[0] = 0;
    for (int half = ss / 2; half >= 1; half /= 2) {
        for (int j = 0; j < placed; j++) sub[placed + j] = sub[j] + half;
        placed *= 2;
    }
}

/* setMask's angles, rotCenterSearch.cl:59-60: {a*PI/180, a*PI/180 - PI_2}
 * formed in double (float angle times double literals), stored as float, then
 * cos() in float: e = {cos a_min, cos(a_min - pi/2), cos a_max, cos(a_max - pi/2)}. */
static void ctor_rcs_angles(float ang_min, float ang_max, float e[4])
{
    e[0] = cosf((float)((double)ang_min * CTOR_PI / 180.0));
    e[1] = cosf((float)((double)ang_min * CTOR_PI / 180.0 - CTOR_PI_2));
    e[2] = cosf((float)((double)ang_max * CTOR_PI / 180.0));
    e[3] = cosf((float)((double)ang_max * CTOR_PI / 180.0 - CTOR_PI_2));
}

/* setMask, rotCenterSearch.cl:50-92, in float as the kernel evaluates it, with
 * the grid's own (nx, ny) in place of the IMAGESIZE_X/Y macros. */
typedef struct {
    float rp, rm, pdiv, cx, cy, p1x, p1y, p2x, p2y, p3x, p3y, p4x, p4y;
} ctor_rcs_mask_t;

static ctor_rcs_mask_t ctor_rcs_params(float shift, const float e[4], int nx, int ny)
{
    ctor_rcs_mask_t m;
    m.cx = (float)(nx / 2);
    m.cy = (float)(ny / 2);
    m.rp = fminf((float)(nx / 2) + shift, (float)(nx / 2));
    m.rm = fmaxf((float)(-(nx / 2)) + shift, (float)(-(nx / 2)));
    m.p1x = m.rp * e[2]; m.p1y = m.rp * e[3];   /* rp * cos(a_max) */
    m.p2x = m.rp * e[0]; m.p2y = m.rp * e[1];   /* rp * cos(a_min) */
    m.p3x = m.rm * e[2]; m.p3y = m.rm * e[3];   /* rm * cos(a_max) */
    m.p4x = m.rm * e[0]; m.p4y = m.rm * e[1];   /* rm * cos(a_min) */
    m.pdiv = (m.rp * m.rp < m.rm * m.rm) ? m.p1x : m.p3x;
    return m;
}

/* One mask texel; texels outside the grid read 0 (CLK_ADDRESS_CLAMP). */
static int ctor_rcs_in(const ctor_rcs_mask_t *m, int X, int Y, int nx, int ny)
{
    if (X < 0 || X >= nx || Y < 0 || Y >= ny) return 0;
    const float x = (float)X - m->cx, y = (float)Y - m->cy;
    const float rp2 = m->rp * m->rp, rm2 = m->rm * m->rm;
    const int b1 = (x * m->p1x + y * m->p1y) <= rp2;
    const int b2 = (x * m->p2x + y * m->p2y) <= rp2;
    const int b3 = (x * m->p3x + y * m->p3y) <= rm2;
    const int b4 = (x * m->p4x + y * m->p4y) <= rm2;
    const float r2 = x * x + y * y;
    const int b5 = r2 <= rp2 && x > m->pdiv;
    const int b6 = r2 <= rm2 && x <= m->pdiv;
    return b1 && b2 && b3 && b4 && (b5 || b6);
}

/* The mask of setMask for candidates shift_i = start + step * i (float),
 * mask [n][ny][nx] of 0/1. */
void ctor_rcs_mask(int nx, int ny, int n, float start, float step, float ang_min, float ang_max,
                   unsigned char *mask)
{
    float e[4];
    ctor_rcs_angles(ang_min, ang_max, e);
    for (int i = 0; i < n; i++) {
        const ctor_rcs_mask_t m = ctor_rcs_params(start + step * (float)i, e, nx, ny);
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++)
                mask[((size_t)i * ny + Y) * nx + X] = (unsigned char)ctor_rcs_in(&m, X, Y, nx, ny);
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
#define RLOG logf
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
#undef RLOG

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
#define RLOG log
#include "ct_oracle_impl.h"
