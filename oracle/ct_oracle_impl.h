/*
 * ct_oracle_impl.h -- precision-generic body of the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Included twice by ct_oracle.c with
 *   R      the arithmetic type (float or double),
 *   S(n)   the symbol suffixer (n##_f32 / n##_f64),
 *   RSQRT/RCOS/RSIN/RFLOOR/RASIN/RFABS/REXP/RFMIN  the matching libm calls.
 *
 * The expressions below keep the operand order of the reference OpenCL text so
 * that the float instantiation rounds the way an FP32 OpenCL CPU device does
 * (no contraction: the file is compiled with -ffp-contract=off).  Citations are
 * file:line in the reference repository.
 */

/* ---- OpenCL 1.2 image sampling, restated (OpenCL 1.2 spec sec. 8.2 / 8.3) -- */

/* CLK_ADDRESS_CLAMP: any texel outside [0,w) x [0,h) reads 0 (CL_R border). */
static inline R S(texel)(const R *img, int w, int h, long i, long j)
{
    return (i >= 0 && i < w && j >= 0 && j < h) ? img[(size_t)j * (size_t)w + (size_t)i] : (R)0;
}

/* CLK_FILTER_NEAREST with unnormalised coordinates: texel (floor(u), floor(v)). */
static inline R S(sample_nearest)(const R *img, int w, int h, R u, R v)
{
    R fu = RFLOOR(u), fv = RFLOOR(v);
    if (!(fu > (R)-4 && fu < (R)(w + 4) && fv > (R)-4 && fv < (R)(h + 4))) return (R)0;
    return S(texel)(img, w, h, (long)fu, (long)fv);
}

/* CLK_FILTER_LINEAR: i0 = floor(u-0.5), a = frac(u-0.5) (texel centres at i+0.5),
 * value = (1-a)(1-b)T00 + a(1-b)T10 + (1-a)bT01 + abT11. */
static inline R S(sample_linear)(const R *img, int w, int h, R u, R v)
{
    R uu = u - (R)0.5, vv = v - (R)0.5;
    R fi = RFLOOR(uu), fj = RFLOOR(vv);
    if (!(fi > (R)-4 && fi < (R)(w + 4) && fj > (R)-4 && fj < (R)(h + 4))) return (R)0;
    long i0 = (long)fi, j0 = (long)fj;
    R a = uu - fi, b = vv - fj;
    R t00 = S(texel)(img, w, h, i0, j0);
    R t10 = S(texel)(img, w, h, i0 + 1, j0);
    R t01 = S(texel)(img, w, h, i0, j0 + 1);
    R t11 = S(texel)(img, w, h, i0 + 1, j0 + 1);
    return ((R)1 - a) * ((R)1 - b) * t00 + a * ((R)1 - b) * t10 + ((R)1 - a) * b * t01 + a * b * t11;
}

/* CLK_FILTER_LINEAR with CLK_ADDRESS_CLAMP_TO_EDGE: texel indices clamped to
 * the image, the weights from the unclamped coordinate. */
static inline R S(sample_linear_edge)(const R *img, int w, int h, R u, R v)
{
    R uu = u - (R)0.5, vv = v - (R)0.5;
    R fi = RFLOOR(uu), fj = RFLOOR(vv);
    R a = uu - fi, b = vv - fj;
    if (!(fi >= (R)-1)) fi = (R)-1;
    if (fi > (R)w) fi = (R)w;
    if (!(fj >= (R)-1)) fj = (R)-1;
    if (fj > (R)h) fj = (R)h;
    long i0 = (long)fi, j0 = (long)fj;
    long i1 = i0 + 1, j1 = j0 + 1;
    i0 = i0 < 0 ? 0 : (i0 >= w ? w - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 >= w ? w - 1 : i1);
    j0 = j0 < 0 ? 0 : (j0 >= h ? h - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 >= h ? h - 1 : j1);
    R t00 = img[(size_t)j0 * w + i0], t10 = img[(size_t)j0 * w + i1];
    R t01 = img[(size_t)j1 * w + i0], t11 = img[(size_t)j1 * w + i1];
    return ((R)1 - a) * ((R)1 - b) * t00 + a * ((R)1 - b) * t10 + ((R)1 - a) * b * t01 + a * b * t11;
}

/* angle_pr = angle[th]*PI/180.0f: float angle times a double PI, rounded to the
 * kernel's float (OSEM.cl:334, AART.cl:440, FBP.cl:180, reconstShare.cl:264). */
static inline R S(angle_rad)(float deg)
{
    return (R)((double)deg * CTOR_PI / 180.0);
}

/* ------------------------------------------------------------------------- */
/* Forward projection of one ray, OSEM1 / AART1 form (OSEM.cl:56-62):
 *   u =  (X-P/2)cos + (Y-D/2)sin + Nx/2 + 0.5
 *   v = -(X-P/2)sin + (Y-D/2)cos + Ny/2 + 0.5,  Y in [0, D), linear sampler. */
static R S(ray_sum)(const ctor_geom *g, const R *layer, int X, R c, R s)
{
    const int D = g->nx > g->ny ? g->nx : g->ny;
    R prj = (R)0;
    for (int Y = 0; Y < D; Y++) {
        R x = ((R)X - (R)g->px * (R)0.5) * c + ((R)Y - (R)D * (R)0.5) * s + (R)g->nx * (R)0.5 + (R)0.5;
        R y = -((R)X - (R)g->px * (R)0.5) * s + ((R)Y - (R)D * (R)0.5) * c + (R)g->ny * (R)0.5 + (R)0.5;
        prj += S(sample_linear)(layer, g->nx, g->ny, x, y);
    }
    return prj;
}

/* OSEM1, OSEM.cl:38-68: ratio r = y/(A x) for the rays of subset `sub`. */
void S(ctor_osem1)(const ctor_geom *g, const float *ang, const R *img, const R *prj,
                   R *rprj, int nz, int sub)
{
    const int P = g->px, A = g->pa, SS = g->ss, NA = A / SS;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int j = 0; j < NA; j++) {
            const int th = sub + j * SS;
            const R ap = S(angle_rad)(ang[th]);
            const R c = RCOS(ap), s = RSIN(ap);
            const R *layer = img + (size_t)Z * g->nx * g->ny;
            for (int X = 0; X < P; X++) {
                R p = S(ray_sum)(g, layer, X, c, s);
                /* int coordinates on a linear sampler: taken as the texel itself */
                R y = prj[((size_t)Z * A + th) * P + X];
                R r = (y > (R)1.0e-6 || y != y) ? y : (R)1.0e-6;   /* max(1e-6, y) */
                r = (r != r) ? (R)1.0e-6 : r;                        /* NaN -> 1e-6 */
                r = (p < (R)0.0001) ? r : r / p;
                rprj[((size_t)Z * A + th) * P + X] = r;
            }
        }
}

/* AART1, AART.cl:38-69: residual d = y - A x (NaN y -> 0). */
void S(ctor_aart1)(const ctor_geom *g, const float *ang, const R *img, const R *prj,
                   R *dprj, int nz, int sub)
{
    const int P = g->px, A = g->pa, SS = g->ss, NA = A / SS;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int j = 0; j < NA; j++) {
            const int th = sub + j * SS;
            const R ap = S(angle_rad)(ang[th]);
            const R c = RCOS(ap), s = RSIN(ap);
            const R *layer = img + (size_t)Z * g->nx * g->ny;
            for (int X = 0; X < P; X++) {
                R p = S(ray_sum)(g, layer, X, c, s);
                R d = prj[((size_t)Z * A + th) * P + X];
                d = (d != d) ? (R)0 : d;
                dprj[((size_t)Z * A + th) * P + X] = d - p;
            }
        }
}

/* Nearest-bin backprojection sum over a subset (OSEM.cl:93-99, AART.cl:99-105):
 * bin = floor((X-Nx/2)cos - (Y-Ny/2)sin + P/2 + 0.5), row th. */
static R S(bp_nearest_sum)(const ctor_geom *g, const float *ang, const R *rows_z, int X, int Y,
                           int sub, int nan_to_zero)
{
    R bprj = (R)0;
    for (int th = sub; th < g->pa; th += g->ss) {
        const R ap = S(angle_rad)(ang[th]);
        R x = ((R)X - (R)g->nx * (R)0.5) * RCOS(ap) - ((R)Y - (R)g->ny * (R)0.5) * RSIN(ap)
              + (R)g->px * (R)0.5 + (R)0.5;
        R v = S(sample_nearest)(rows_z, g->px, g->pa, x, (R)th + (R)0.5);
        if (nan_to_zero) v = (v != v) ? (R)0 : v;
        bprj += v;
    }
    return bprj;
}

/* OSEM2, OSEM.cl:71-105: x' = x * sum(r) * SS / A, NaN -> 1e-6. */
void S(ctor_osem2)(const ctor_geom *g, const float *ang, const R *img, const R *rprj,
                   R *img_dst, int nz, int sub)
{
    const int nx = g->nx, ny = g->ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *rows = rprj + (size_t)Z * g->pa * g->px;
                R bprj = S(bp_nearest_sum)(g, ang, rows, X, Y, sub, 0);
                size_t o = ((size_t)Z * ny + Y) * nx + X;
                R v = img[o] * bprj * (R)g->ss / (R)g->pa;
                img_dst[o] = (v != v) ? (R)1.0e-6 : v;
            }
}

/* AART2, AART.cl:73-106: x' = x + sum(d) * alpha * SS / A / D. */
void S(ctor_aart2)(const ctor_geom *g, const float *ang, const R *img, const R *dprj,
                   R *img_dst, int nz, int sub, float alpha)
{
    const int nx = g->nx, ny = g->ny, D = nx > ny ? nx : ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *rows = dprj + (size_t)Z * g->pa * g->px;
                R bprj = S(bp_nearest_sum)(g, ang, rows, X, Y, sub, 0);
                size_t o = ((size_t)Z * ny + Y) * nx + X;
                img_dst[o] = img[o] + bprj * (R)alpha * (R)g->ss / (R)g->pa / (R)D;
            }
}

/* backProjection, reconstShare.cl:245-275: x = sum(y, NaN->0) * SS / A / D,
 * integer half sizes (X - IMAGESIZE_X/2). */
void S(ctor_backprojection)(const ctor_geom *g, const float *ang, const R *prj,
                            R *img_dst, int nz, int sub)
{
    const int nx = g->nx, ny = g->ny, D = nx > ny ? nx : ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *rows = prj + (size_t)Z * g->pa * g->px;
                R bprj = (R)0;
                for (int th = sub; th < g->pa; th += g->ss) {
                    const R ap = S(angle_rad)(ang[th]);
                    R x = (R)(X - nx / 2) * RCOS(ap) - (R)(Y - ny / 2) * RSIN(ap)
                          + (R)(g->px / 2) + (R)0.5;
                    R v = S(sample_nearest)(rows, g->px, g->pa, x, (R)th + (R)0.5);
                    bprj += (v != v) ? (R)0 : v;
                }
                img_dst[((size_t)Z * ny + Y) * nx + X] = bprj * (R)g->ss / (R)g->pa / (R)D;
            }
}

/* sinogramCorrection, reconstShare.cl:35-70 (a = b = 0.5 from the host). */
void S(ctor_sinogram_correction)(const ctor_geom *g, const float *ang, const R *src,
                                 R *dst, int nz, int mode, float a, float b)
{
    const int P = g->px, A = g->pa;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int th = 0; th < A; th++) {
            /* angle[th]/180*PI: float division first, then the double PI */
            const R ap = (R)((double)(R)((R)ang[th] / (R)180) * CTOR_PI);
            const R sinfactor = (R)a + (R)b * RCOS(ap);
            for (int X = 0; X < P; X++) {
                size_t o = ((size_t)Z * A + th) * P + X;
                R v = src[o];
                R t = (R)2.0 * RFABS((R)X - (R)P / (R)2.0) / (R)P;
                switch (mode) {
                case 4: v = (t * t < sinfactor * sinfactor) ? v : (R)0; break;
                case 3: v *= RCOS(RASIN(t)) * RCOS(ap); break;
                case 2: v *= RCOS(ap); break;
                case 1: v *= RCOS(RASIN(t)); break;
                default: break;
                }
                dst[o] = v;
            }
        }
}

/* ---- FBP filter chain, FBP.cl:34-160 with the launch order of GPU.cpp:746-802 */

static unsigned S(bitrev)(unsigned x, unsigned iter)
{
    x = (x & 0x55555555u) << 1 | (x & 0xAAAAAAAAu) >> 1;
    x = (x & 0x33333333u) << 2 | (x & 0xCCCCCCCCu) >> 2;
    x = (x & 0x0F0F0F0Fu) << 4 | (x & 0xF0F0F0F0u) >> 4;
    x = (x & 0x00FF00FFu) << 8 | (x & 0xFF00FF00u) >> 8;
    x = (x & 0x0000FFFFu) << 16 | (x & 0xFFFF0000u) >> 16;
    return x >> (32 - iter);
}

/* One radix-2 DIT transform of length zp in place (bitReverse + `iter` butterfly
 * stages, FBP.cl:60-120).  conj != 0 is the IFFT flag 0x80000000 (twiddle
 * imaginary part negated). */
static void S(fft_radix2)(R *re, R *im, R *tre, R *tim, const R *Wre, const R *Wim,
                          int zp, unsigned iter, int conj)
{
    for (unsigned X1 = 0; X1 < (unsigned)zp; X1++) {
        unsigned X2 = S(bitrev)(X1, iter);
        tre[X2] = re[X1];
        tim[X2] = im[X1];
    }
    for (unsigned it = 1; it <= iter; it++) {
        unsigned bf_size = 1u << (it - 1), gpdist = 1u << it, gpnum = (unsigned)zp >> it;
        for (unsigned X = 0; X < (unsigned)zp / 2; X++) {
            unsigned base = (X >> (it - 1)) * gpdist, off = X & (bf_size - 1);
            unsigned a = base + off, b = a + bf_size, l = gpnum * off;
            R wx = Wre[l], wy = conj ? -Wim[l] : Wim[l];
            R xax = tre[a], xay = tim[a], xbx = tre[b], xby = tim[b];
            /* res_a = xa + xb.xx*W + xb.yy*(-W.y, W.x); res_b = xa - xb.xx*W + xb.yy*(W.y, -W.x) */
            R rax = xax + xbx * wx + xby * (-wy);
            R ray = xay + xbx * wy + xby * wx;
            R rbx = xax - xbx * wx + xby * wy;
            R rby = xay - xbx * wy + xby * (-wx);
            tre[a] = rax; tim[a] = ray; tre[b] = rbx; tim[b] = rby;
        }
    }
    for (int i = 0; i < zp; i++) { re[i] = tre[i]; im[i] = tim[i]; }
}

/* Ramp filtering of every sinogram row: zero padding with (X - P/2) mod ZP index,
 * FFT, H[k] = (pi/ZP) min(k, ZP-k), IFFT, / ZP, real part back at the same index
 * (FBP.cl:34-160; GPU.cpp:746-802). */
void S(ctor_fbp_filter)(const ctor_geom *g, const R *prj, R *fprj, int nz)
{
    const int P = g->px, A = g->pa, zp = g->zp;
    unsigned iter = 0;
    while ((1 << iter) < zp) iter++;
    R *Wre = (R *)malloc(sizeof(R) * zp), *Wim = (R *)malloc(sizeof(R) * zp);
    /* spinFactor, FBP.cl:34-40: d = (float)(2 PI / ZP); W = cos((d*l, d*l + PI/2)) */
    const R d = (R)(2.0 * CTOR_PI / zp);
    for (int l = 0; l < zp; l++) {
        R a0 = d * (R)l;
        R a1 = (R)((double)(d * (R)l) + CTOR_PI_2);
        Wre[l] = RCOS(a0);
        Wim[l] = RCOS(a1);
    }
    const R h = (R)(CTOR_PI / zp);
    #pragma omp parallel
    {
        R *re = (R *)malloc(sizeof(R) * zp), *im = (R *)malloc(sizeof(R) * zp);
        R *tre = (R *)malloc(sizeof(R) * zp), *tim = (R *)malloc(sizeof(R) * zp);
        #pragma omp for collapse(2) schedule(static)
        for (int Z = 0; Z < nz; Z++)
            for (int th = 0; th < A; th++) {
                const R *row = prj + ((size_t)Z * A + th) * P;
                for (int i = 0; i < zp; i++) { re[i] = (R)0; im[i] = (R)0; }
                for (int X = 0; X < P; X++) {
                    int xc = X - P / 2 + zp / 2;
                    int xs = (xc < zp / 2) ? (xc + zp / 2) : (xc - zp / 2);
                    re[xs] = row[X];
                }
                S(fft_radix2)(re, im, tre, tim, Wre, Wim, zp, iter, 0);
                for (int X = 0; X < zp; X++) {       /* filtering, FBP.cl:125-134 */
                    R f = (X < zp / 2) ? ((R)X * h) : ((R)(zp - X) * h);
                    re[X] *= f;
                    im[X] *= f;
                }
                S(fft_radix2)(re, im, tre, tim, Wre, Wim, zp, iter, 1);
                for (int X = 0; X < zp; X++) {       /* normalization, FBP.cl:138-145 */
                    re[X] /= (R)zp;
                    im[X] /= (R)zp;
                }
                R *out = fprj + ((size_t)Z * A + th) * P;
                for (int X = 0; X < P; X++) {         /* outputImage, FBP.cl:149-160 */
                    int xc = X - P / 2 + zp / 2;
                    int xs = (xc < zp / 2) ? (xc + zp / 2) : (xc - zp / 2);
                    out[X] = re[xs];
                }
            }
        free(re); free(im); free(tre); free(tim);
    }
    free(Wre); free(Wim);
}

/* backProjectionFBP, FBP.cl:164-188: linear sampling of the filtered row inside
 * the disk r < Nx/2, divided by 2A. */
void S(ctor_fbp_backprojection)(const ctor_geom *g, const float *ang, const R *fprj, R *img, int nz)
{
    const int nx = g->nx, ny = g->ny, A = g->pa;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *rows = fprj + (size_t)Z * A * g->px;
                R bprj = (R)0;
                R dx = (R)X - (R)nx * (R)0.5, dy = (R)Y - (R)ny * (R)0.5;
                R radius = RSQRT(dx * dx + dy * dy);
                for (int th = 0; th < A; th++) {
                    const R ap = S(angle_rad)(ang[th]);
                    R x = dx * RCOS(ap) - dy * RSIN(ap) + (R)g->px * (R)0.5 + (R)0.5;
                    bprj += (radius < (R)nx * (R)0.5)
                            ? S(sample_linear)(rows, g->px, A, x, (R)th + (R)0.5) : (R)0;
                }
                bprj /= (R)A * (R)2.0;
                img[((size_t)Z * ny + Y) * nx + X] = bprj;
            }
}

/* ---- power iteration, FISTA.cl:50-208 ------------------------------------ */

/* powerIter1, FISTA.cl:50-75: layer s projects subset s into compact row j,
 * integer half sizes and no +0.5 offsets. img and prjc hold ss layers. */
void S(ctor_power_iter1)(const ctor_geom *g, const float *ang, const R *img, R *prjc)
{
    const int P = g->px, SS = g->ss, NA = g->pa / SS, nx = g->nx, ny = g->ny;
    const int D = nx > ny ? nx : ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < SS; Z++)
        for (int j = 0; j < NA; j++) {
            const int th = Z + j * SS;
            const R ap = S(angle_rad)(ang[th]);
            const R c = RCOS(ap), s = RSIN(ap);
            const R *layer = img + (size_t)Z * nx * ny;
            for (int X = 0; X < P; X++) {
                R prj = (R)0;
                for (int Y = 0; Y < D; Y++) {
                    R x = (R)(X - P / 2) * c + (R)(Y - D / 2) * s + (R)(nx / 2);
                    R y = -(R)(X - P / 2) * s + (R)(Y - D / 2) * c + (R)(ny / 2);
                    prj += S(sample_linear)(layer, nx, ny, x, y);
                }
                prjc[((size_t)Z * g->pa + j) * P + X] = prj;
            }
        }
}

/* powerIter2, FISTA.cl:79-110: truncating-bin backprojection of the compact rows. */
void S(ctor_power_iter2)(const ctor_geom *g, const float *ang, const R *prjc, R *cnd)
{
    const int SS = g->ss, NA = g->pa / SS, nx = g->nx, ny = g->ny;
    const int D = nx > ny ? nx : ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < SS; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *rows = prjc + (size_t)Z * g->pa * g->px;
                R bprj = (R)0;
                for (int i = 0; i < NA; i++) {
                    const int th = Z + i * SS;
                    const R ap = S(angle_rad)(ang[th]);
                    R x = (R)(X - nx / 2) * RCOS(ap) - (R)(Y - ny / 2) * RSIN(ap) + (R)(g->px / 2);
                    bprj += S(sample_nearest)(rows, g->px, g->pa, x, (R)i);
                }
                bprj *= (R)SS / (R)g->pa / (R)D;
                cnd[((size_t)Z * ny + Y) * nx + X] = bprj;
            }
}

/* imageL2AbsX + imageL2AbsY, FISTA.cl:133-208: per-layer sum of squares. */
void S(ctor_image_l2)(const ctor_geom *g, const R *img, R *l2, int nz)
{
    const size_t n = (size_t)g->nx * g->ny;
    for (int Z = 0; Z < nz; Z++) {
        double acc = 0.0;   /* reduction order is not part of the contract */
        for (size_t i = 0; i < n; i++) acc += (double)img[(size_t)Z * n + i] * (double)img[(size_t)Z * n + i];
        l2[Z] = (R)acc;
    }
}

/* powerIter3, FISTA.cl:114-128: x = cnd / sqrt(L2[layer]). */
void S(ctor_power_iter3)(const ctor_geom *g, const R *cnd, const R *l2, R *img)
{
    const size_t n = (size_t)g->nx * g->ny;
    for (int Z = 0; Z < g->ss; Z++) {
        R q = RSQRT(l2[Z]);
        for (size_t i = 0; i < n; i++) img[(size_t)Z * n + i] = cnd[(size_t)Z * n + i] / q;
    }
}

/* ---- FISTA helpers, FISTA.cl:211-594 -------------------------------------- */

/* FISTAbackProjection, FISTA.cl:211-246: truncating bin, no +0.5. */
void S(ctor_fista_backprojection)(const ctor_geom *g, const float *ang, const R *img,
                                  const R *dprj, const R *L, R *img_dst, int nz, int sub, float alpha)
{
    const int nx = g->nx, ny = g->ny, D = nx > ny ? nx : ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *rows = dprj + (size_t)Z * g->pa * g->px;
                R bprj = (R)0;
                for (int th = sub; th < g->pa; th += g->ss) {
                    const R ap = S(angle_rad)(ang[th]);
                    R x = ((R)X - (R)nx * (R)0.5) * RCOS(ap) - ((R)Y - (R)ny * (R)0.5) * RSIN(ap)
                          + (R)g->px * (R)0.5;
                    bprj += S(sample_nearest)(rows, g->px, g->pa, x, (R)th);
                }
                size_t o = ((size_t)Z * ny + Y) * nx + X;
                img_dst[o] = img[o] + bprj * (R)alpha * (R)g->ss / (R)g->pa / (R)D / L[sub] * (R)0.5;
            }
}

/* Shared TV-prox core of ISTA3D / FISTA3D (FISTA.cl:429-487, 522-579): the six
 * neighbours are selection-sorted descending, then x is picked from 13
 * candidates by the range scan.  Neighbours read border 0 in-plane; z+ is layer
 * min(Z+1, nz-1) and z- is always layer 0 (fmin(Z-1, 0), FISTA.cl:425). */
static R S(tv_prox)(const ctor_geom *g, const R *v, int nz, int X, int Y, int Z, R lambda)
{
    const int nx = g->nx, ny = g->ny;
    const R *lz = v + (size_t)Z * nx * ny;
    R nb[6];
    nb[0] = S(texel)(lz, nx, ny, X + 1, Y);
    nb[1] = S(texel)(lz, nx, ny, X - 1, Y);
    nb[2] = S(texel)(lz, nx, ny, X, Y + 1);
    nb[3] = S(texel)(lz, nx, ny, X, Y - 1);
    int zp = (Z + 1 < nz - 1) ? Z + 1 : nz - 1;
    nb[4] = v[((size_t)zp * ny + Y) * nx + X];
    nb[5] = v[((size_t)0 * ny + Y) * nx + X];
    R vi = lz[(size_t)Y * nx + X];
    for (int i = 0; i < 6; i++) {
        R img1 = nb[i], img2 = img1;
        int bigger = i;
        for (int j = i + 1; j < 6; j++) {
            bigger = (nb[j] > img2) ? j : bigger;
            img2 = (nb[j] > img2) ? nb[j] : img2;
        }
        nb[i] = img2;
        nb[bigger] = img1;
    }
    R cnd[13], rng[12];
    cnd[0] = vi - (R)6.0 * lambda;  cnd[1] = nb[0];
    cnd[2] = vi - (R)4.0 * lambda;  cnd[3] = nb[1];
    cnd[4] = vi - (R)2.0 * lambda;  cnd[5] = nb[2];
    cnd[6] = vi;                     cnd[7] = nb[3];
    cnd[8] = vi + (R)2.0 * lambda;  cnd[9] = nb[4];
    cnd[10] = vi + (R)4.0 * lambda; cnd[11] = nb[5];
    cnd[12] = vi + (R)6.0 * lambda;
    rng[0] = nb[0] + (R)6.0 * lambda; rng[1] = nb[0] + (R)4.0 * lambda;
    rng[2] = nb[1] + (R)4.0 * lambda; rng[3] = nb[1] + (R)2.0 * lambda;
    rng[4] = nb[2] + (R)2.0 * lambda; rng[5] = nb[2];
    rng[6] = nb[3];                    rng[7] = nb[3] - (R)2.0 * lambda;
    rng[8] = nb[4] - (R)2.0 * lambda; rng[9] = nb[4] - (R)4.0 * lambda;
    rng[10] = nb[5] - (R)4.0 * lambda; rng[11] = nb[5] - (R)6.0 * lambda;
    R x = cnd[0];
    for (int i = 0; i < 12; i++) x = (vi < rng[i]) ? cnd[i + 1] : x;
    return x;
}

/* lambda = LAMBDA_FISTA / L[sub] * 0.5f with LAMBDA_FISTA a double literal. */
static inline R S(prox_lambda)(const R *L, int sub, double lambda_fista)
{
    return (R)(lambda_fista / (double)L[sub] * 0.5);
}

/* ISTA3D, FISTA.cl:409-493: prox, then max(1e-6, x), NaN -> 1e-6. */
void S(ctor_ista3d)(const ctor_geom *g, const R *v, R *x_new, int nz, int sub,
                    const R *L, double lambda_fista)
{
    const int nx = g->nx, ny = g->ny;
    const R lambda = S(prox_lambda)(L, sub, lambda_fista);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                R x = S(tv_prox)(g, v, nz, X, Y, Z, lambda);
                x = (x > (R)1.0e-6 || x != x) ? x : (R)1.0e-6;   /* fmax(1e-6, x) */
                x = (x != x) ? (R)1.0e-6 : x;
                x_new[((size_t)Z * ny + Y) * nx + X] = x;
            }
}

/* Per-slice TV prox of the joint alignment model (the 2-D form of ISTA3D's
 * median prox, FISTA.cl:409-493; no reference kernel): the median of the four
 * in-plane neighbours (border 0) and v - 4 lam, v - 2 lam, v, v + 2 lam,
 * v + 4 lam, found here by sorting all nine values; then max(lower, x),
 * NaN -> lower. */
void S(ctor_tv_prox2d)(const ctor_geom *g, const R *v, R *x_new, int nz, R lambda, R lower)
{
    const int nx = g->nx, ny = g->ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                const R *lz = v + (size_t)Z * nx * ny;
                const R vi = lz[(size_t)Y * nx + X];
                R t[9] = {S(texel)(lz, nx, ny, X + 1, Y), S(texel)(lz, nx, ny, X - 1, Y),
                          S(texel)(lz, nx, ny, X, Y + 1), S(texel)(lz, nx, ny, X, Y - 1),
                          vi - (R)4 * lambda, vi - (R)2 * lambda, vi, vi + (R)2 * lambda, vi + (R)4 * lambda};
                for (int a = 1; a < 9; a++) {          /* insertion sort, ascending */
                    R key = t[a];
                    int b = a - 1;
                    while (b >= 0 && t[b] > key) { t[b + 1] = t[b]; b--; }
                    t[b + 1] = key;
                }
                R x = t[4];
                x = (x > lower) ? x : lower;
                x_new[((size_t)Z * ny + Y) * nx + X] = x;
            }
}

/* FISTA3D, FISTA.cl:497-594: prox (no clamp), beta' = (sqrt(4 beta^2 + 1) + 1)/2,
 * gamma = (beta - 1)/beta', w' = gamma x_new + (1 - gamma) x (NaN -> 1e-6).
 * The x_new write happens before the beta write, so when both outputs are the
 * same buffer (GPU.cpp:1847-1848) the beta value is what remains. */
void S(ctor_fista3d)(const ctor_geom *g, const R *x, const R *v, const R *b,
                     R *w_new, R *x_new, R *b_new, int nz, int sub, const R *L,
                     double lambda_fista)
{
    const int nx = g->nx, ny = g->ny;
    const R lambda = S(prox_lambda)(L, sub, lambda_fista);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                size_t o = ((size_t)Z * ny + Y) * nx + X;
                R xo = x[o], beta = b[o];
                R xn = S(tv_prox)(g, v, nz, X, Y, Z, lambda);
                R bn = RSQRT(beta * beta * (R)4.0 + (R)1.0);
                bn = (bn + (R)1.0) * (R)0.5;
                R gamma = (beta - (R)1.0) / bn;
                R w = gamma * xn + ((R)1.0 - gamma) * xo;
                w = (w != w) ? (R)1.0e-6 : w;
                x_new[o] = xn;
                b_new[o] = bn;
                w_new[o] = w;
            }
}

/* ---- execution drivers (GPU.cpp) ----------------------------------------- */

/* The optional sinogram correction step shared by every driver: the sinogram is
 * copied to a temporary and corrected back into prj (GPU.cpp:175-186). */
static void S(correct_prj)(const ctor_geom *g, const float *ang, R *prj, int nz, int mode)
{
    size_t n = (size_t)nz * g->pa * g->px;
    R *tmp = (R *)malloc(sizeof(R) * n);
    memcpy(tmp, prj, sizeof(R) * n);
    S(ctor_sinogram_correction)(g, ang, tmp, prj, nz, mode, 0.5f, 0.5f);
    free(tmp);
}

/* OSEM_execution, GPU.cpp:125-240. */
void S(ctor_osem_execution)(const ctor_geom *g, const float *ang, const int *sub,
                            R *img, R *prj, int nz, int it, int prj_correction, int correction_mode)
{
    size_t ni = (size_t)nz * g->nx * g->ny, np = (size_t)nz * g->pa * g->px;
    R *dst = (R *)malloc(sizeof(R) * ni), *rprj = (R *)calloc(np, sizeof(R));
    if (prj_correction && correction_mode > 0) S(correct_prj)(g, ang, prj, nz, correction_mode);
    for (int i = 0; i < it; i++)
        for (int k = 0; k < g->ss; k++) {
            S(ctor_osem1)(g, ang, img, prj, rprj, nz, sub[k]);
            S(ctor_osem2)(g, ang, img, rprj, dst, nz, sub[k]);
            memcpy(img, dst, sizeof(R) * ni);
        }
    free(dst); free(rprj);
}

/* FISTA_OSEM_execution, GPU.cpp:243-405.  First sub-iteration: ISTA3D of the
 * pre-update image replaces the OS-EM result; afterwards x <- FISTA3D prox of the
 * OS-EM update.  L is filled with 0.5 (GPU.cpp:326). */
void S(ctor_fista_osem_execution)(const ctor_geom *g, const float *ang, const int *sub,
                                  R *img, R *prj, int nz, int it, int prj_correction,
                                  int correction_mode, double cs_lambda)
{
    size_t ni = (size_t)nz * g->nx * g->ny, np = (size_t)nz * g->pa * g->px;
    R *dst = (R *)malloc(sizeof(R) * ni), *rprj = (R *)calloc(np, sizeof(R));
    R *v = (R *)malloc(sizeof(R) * ni), *w = (R *)malloc(sizeof(R) * ni);
    R *beta = (R *)calloc(ni, sizeof(R)), *beta_dst = (R *)calloc(ni, sizeof(R));
    R *L = (R *)malloc(sizeof(R) * g->ss);
    for (int s = 0; s < g->ss; s++) L[s] = (R)0.5;
    if (prj_correction && correction_mode > 0) S(correct_prj)(g, ang, prj, nz, correction_mode);
    memcpy(w, img, sizeof(R) * ni);
    int first = 1;
    for (int i = 0; i < it; i++)
        for (int k = 0; k < g->ss; k++) {
            S(ctor_osem1)(g, ang, img, prj, rprj, nz, sub[k]);
            S(ctor_osem2)(g, ang, img, rprj, dst, nz, sub[k]);
            if (first) {
                S(ctor_ista3d)(g, img, dst, nz, sub[k], L, cs_lambda);
                first = 0;
            } else {
                memcpy(v, dst, sizeof(R) * ni);
                /* x input and w output are the same image (GPU.cpp:376-377, 336) */
                S(ctor_fista3d)(g, w, v, beta, w, dst, beta_dst, nz, sub[k], L, cs_lambda);
                memcpy(beta, beta_dst, sizeof(R) * ni);
            }
            memcpy(img, dst, sizeof(R) * ni);
        }
    free(dst); free(rprj); free(v); free(w); free(beta); free(beta_dst); free(L);
}

/* AART_execution, GPU.cpp:1088-1216: the image is first replaced by the plain
 * backprojection of subset sub[0] (GPU.cpp:1167-1176). */
void S(ctor_aart_execution)(const ctor_geom *g, const float *ang, const int *sub,
                            R *img, R *prj, int nz, int it, int correction_mode, float alpha)
{
    size_t ni = (size_t)nz * g->nx * g->ny, np = (size_t)nz * g->pa * g->px;
    R *dst = (R *)malloc(sizeof(R) * ni), *dprj = (R *)calloc(np, sizeof(R));
    if (correction_mode > 0) S(correct_prj)(g, ang, prj, nz, correction_mode);
    S(ctor_backprojection)(g, ang, prj, dst, nz, sub[0]);
    memcpy(img, dst, sizeof(R) * ni);
    for (int i = 0; i < it; i++)
        for (int k = 0; k < g->ss; k++) {
            S(ctor_aart1)(g, ang, img, prj, dprj, nz, sub[k]);
            S(ctor_aart2)(g, ang, img, dprj, dst, nz, sub[k], alpha);
            memcpy(img, dst, sizeof(R) * ni);
        }
    free(dst); free(dprj);
}

/* FBP_execution, GPU.cpp:665-819. */
void S(ctor_fbp_execution)(const ctor_geom *g, const float *ang, R *img, R *prj,
                           int nz, int prj_correction, int correction_mode)
{
    size_t np = (size_t)nz * g->pa * g->px;
    R *fprj = (R *)malloc(sizeof(R) * np);
    if (prj_correction && correction_mode > 0) S(correct_prj)(g, ang, prj, nz, correction_mode);
    S(ctor_fbp_filter)(g, prj, fprj, nz);
    S(ctor_fbp_backprojection)(g, ang, fprj, img, nz);
    free(fprj);
}

/* findMinimumX / findMinimumY, reconstShare.cl:114-177: the minimum of each
 * layer (fmin skips NaN); the reference's work-group loop covers the whole
 * layer when its sizes are multiples of the work-group size. */
void S(ctor_layer_min)(const ctor_geom *g, const R *img, R *mins, int nz)
{
    const size_t plane = (size_t)g->nx * g->ny;
    for (int z = 0; z < nz; z++) {
        R m = (R)INFINITY;
        for (size_t i = 0; i < plane; i++) m = RFMIN(m, img[(size_t)z * plane + i]);
        mins[z] = m;
    }
}

/* baseUp, reconstShare.cl:88-110: x -= a * exp(g) with a the layer minimum and
 * g the ln(1 + b) series cut at `order`, b = (x - a) / a. */
void S(ctor_baseup)(const ctor_geom *g, R *img, const R *mins, int nz, int order)
{
    const size_t plane = (size_t)g->nx * g->ny;
    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)(plane * nz); i++) {
        const R a = mins[(size_t)i / plane];
        R x = img[i];
        const R b1 = (a != (R)0) ? (x - a) / a : (R)0;
        R b2 = (R)1, flag = (R)-1, gg = (R)0;
        for (int k = 1; k <= order; k++) {
            b2 *= b1;
            flag *= (R)-1;
            gg += flag * b2 / (R)k;
        }
        x -= a * REXP(gg);
        img[i] = x;
    }
}

/* hybrid_thread, GPU.cpp:1459-1580 (mode 8): FBP_execution with the
 * correction, base-up (FBPcorrection = 1), then OSEM_execution or
 * FISTA_OSEM_execution without a second correction. */
void S(ctor_hybrid_execution)(const ctor_geom *g, const float *ang, const int *sub, R *img, R *prj,
                              int nz, int it, int prj_correction, int correction_mode, int baseup_order,
                              int cs, double cs_lambda)
{
    R *mins = (R *)malloc(sizeof(R) * (size_t)nz);
    S(ctor_fbp_execution)(g, ang, img, prj, nz, prj_correction, correction_mode);
    S(ctor_layer_min)(g, img, mins, nz);
    S(ctor_baseup)(g, img, mins, nz, baseup_order);
    free(mins);
    if (cs) S(ctor_fista_osem_execution)(g, ang, sub, img, prj, nz, it, 0, correction_mode, cs_lambda);
    else S(ctor_osem_execution)(g, ang, sub, img, prj, nz, it, 0, correction_mode);
}

/* powerIteration, GPU.cpp:2079-2188: ss layers filled with 1.0, `trials` (30 in
 * the reference) normalised steps; L[s] = ||M_s x||^2 of the last step. */
void S(ctor_power_iteration)(const ctor_geom *g, const float *ang, R *L, int trials)
{
    size_t ni = (size_t)g->ss * g->nx * g->ny, np = (size_t)g->ss * g->pa * g->px;
    R *x = (R *)malloc(sizeof(R) * ni), *cnd = (R *)malloc(sizeof(R) * ni);
    R *prjc = (R *)calloc(np, sizeof(R));
    for (size_t i = 0; i < ni; i++) x[i] = (R)1;
    for (int t = 0; t < trials; t++) {
        S(ctor_power_iter1)(g, ang, x, prjc);
        S(ctor_power_iter2)(g, ang, prjc, cnd);
        S(ctor_image_l2)(g, cnd, L, g->ss);
        S(ctor_power_iter3)(g, cnd, L, x);
    }
    free(x); free(cnd); free(prjc);
}

/* FISTA_execution, GPU.cpp:1755-1924, with the reference argument bindings:
 * FISTA3D's x_new output is bound to beta_dest_img (GPU.cpp:1847), so the
 * returned image is the first ISTA3D result x1 for any it >= 1. */
void S(ctor_fista_execution)(const ctor_geom *g, const float *ang, const R *L, const int *sub,
                             R *img, R *prj, int nz, int it, int prj_correction,
                             int correction_mode, float alpha, double cs_lambda)
{
    size_t ni = (size_t)nz * g->nx * g->ny, np = (size_t)nz * g->pa * g->px;
    R *v = (R *)malloc(sizeof(R) * ni), *w = (R *)malloc(sizeof(R) * ni);
    R *beta = (R *)calloc(ni, sizeof(R)), *dst = (R *)malloc(sizeof(R) * ni);
    R *beta_dst = (R *)calloc(ni, sizeof(R)), *vprj = (R *)calloc(np, sizeof(R));
    if (prj_correction && correction_mode > 0) S(correct_prj)(g, ang, prj, nz, correction_mode);
    S(ctor_backprojection)(g, ang, prj, dst, nz, sub[0]);
    memcpy(img, dst, sizeof(R) * ni);
    memcpy(w, img, sizeof(R) * ni);
    int first = 1;
    for (int i = 0; i < it; i++)
        for (int k = 0; k < g->ss; k++) {
            const R *src = first ? img : w;
            S(ctor_aart1)(g, ang, src, prj, vprj, nz, sub[k]);
            S(ctor_fista_backprojection)(g, ang, src, vprj, L, v, nz, sub[k], alpha);
            if (first) {
                S(ctor_ista3d)(g, v, dst, nz, sub[k], L, cs_lambda);
                first = 0;
            } else {
                S(ctor_fista3d)(g, img, v, beta, w, beta_dst, beta_dst, nz, sub[k], L, cs_lambda);
                memcpy(beta, beta_dst, sizeof(R) * ni);
            }
            memcpy(img, dst, sizeof(R) * ni);
        }
    free(v); free(w); free(beta); free(dst); free(beta_dst); free(vprj);
}

/* ---- differentiable alignment model (no counterpart in the reference) ----- *
 * The OSEM1 projector over all angles with a per-projection detector shift
 * delta[a] and a rotation-centre offset c moving the detector coordinate:
 *   xi  = (X - P/2) - delta[a] - c,   eta = k - D/2,  k in [0, D)
 *   fu  =  xi cos + eta sin + Nx/2,   fv = -xi sin + eta cos + Ny/2
 *   p[a][X] = sum_k bilinear(img; texel corner at floor(fu), floor(fv)).
 * With delta = c = 0 it equals the OSEM1 forward projection (OSEM.cl:56-62). */
static inline void S(align_coord)(const ctor_geom *g, R c, R s, R xi, int k, R *fu, R *fv)
{
    const int D = g->nx > g->ny ? g->nx : g->ny;
    R eta = (R)k - (R)D * (R)0.5;
    *fu = xi * c + eta * s + (R)g->nx * (R)0.5;
    *fv = -xi * s + eta * c + (R)g->ny * (R)0.5;
}

void S(ctor_align_forward)(const ctor_geom *g, const float *ang, const R *img,
                           const R *shift, R center, R *prj, int nz)
{
    const int P = g->px, A = g->pa, nx = g->nx, ny = g->ny, D = nx > ny ? nx : ny;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int Z = 0; Z < nz; Z++)
        for (int a = 0; a < A; a++) {
            const R ap = S(angle_rad)(ang[a]);
            const R c = RCOS(ap), s = RSIN(ap);
            const R *layer = img + (size_t)Z * nx * ny;
            for (int X = 0; X < P; X++) {
                R xi = (R)X - (R)P * (R)0.5 - (shift ? shift[a] : (R)0) - center;
                R p = (R)0;
                for (int k = 0; k < D; k++) {
                    R fu, fv;
                    S(align_coord)(g, c, s, xi, k, &fu, &fv);
                    p += S(sample_linear)(layer, nx, ny, fu + (R)0.5, fv + (R)0.5);
                }
                prj[((size_t)Z * A + a) * P + X] = p;
            }
        }
}

/* Exact transpose of ctor_align_forward, written as a scatter of every sample's
 * bilinear weights (parallel over layers only, so sums are deterministic). */
void S(ctor_align_adjoint)(const ctor_geom *g, const float *ang, const R *prj,
                           const R *shift, R center, R *img, int nz)
{
    const int P = g->px, A = g->pa, nx = g->nx, ny = g->ny, D = nx > ny ? nx : ny;
    #pragma omp parallel for schedule(static)
    for (int Z = 0; Z < nz; Z++) {
        R *layer = img + (size_t)Z * nx * ny;
        for (size_t i = 0; i < (size_t)nx * ny; i++) layer[i] = (R)0;
        for (int a = 0; a < A; a++) {
            const R ap = S(angle_rad)(ang[a]);
            const R c = RCOS(ap), s = RSIN(ap);
            for (int X = 0; X < P; X++) {
                R gv = prj[((size_t)Z * A + a) * P + X];
                R xi = (R)X - (R)P * (R)0.5 - (shift ? shift[a] : (R)0) - center;
                for (int k = 0; k < D; k++) {
                    R fu, fv;
                    S(align_coord)(g, c, s, xi, k, &fu, &fv);
                    R fi = RFLOOR(fu), fj = RFLOOR(fv);
                    if (!(fi > (R)-4 && fi < (R)(nx + 4) && fj > (R)-4 && fj < (R)(ny + 4))) continue;
                    long i0 = (long)fi, j0 = (long)fj;
                    R wa = fu - fi, wb = fv - fj;
                    R w[4] = {((R)1 - wa) * ((R)1 - wb), wa * ((R)1 - wb), ((R)1 - wa) * wb, wa * wb};
                    long ii[4] = {i0, i0 + 1, i0, i0 + 1}, jj[4] = {j0, j0, j0 + 1, j0 + 1};
                    for (int t = 0; t < 4; t++)
                        if (ii[t] >= 0 && ii[t] < nx && jj[t] >= 0 && jj[t] < ny)
                            layer[(size_t)jj[t] * nx + ii[t]] += w[t] * gv;
                }
            }
        }
    }
}

/* d/d delta[a] of sum_{Z,X} gprj[Z][a][X] p[Z][a][X]; delta is shared by all
 * layers.  d fu/d delta = -cos, d fv/d delta = +sin; the bilinear partials are
 * (1-b)(T10-T00) + b(T11-T01) and (1-a)(T01-T00) + a(T11-T10). */
void S(ctor_align_grad)(const ctor_geom *g, const float *ang, const R *img,
                        const R *gprj, const R *shift, R center, R *gshift, int nz)
{
    const int P = g->px, A = g->pa, nx = g->nx, ny = g->ny, D = nx > ny ? nx : ny;
    #pragma omp parallel for schedule(static)
    for (int a = 0; a < A; a++) {
        const R ap = S(angle_rad)(ang[a]);
        const R c = RCOS(ap), s = RSIN(ap);
        R acc = (R)0;
        for (int Z = 0; Z < nz; Z++) {
            const R *layer = img + (size_t)Z * nx * ny;
            for (int X = 0; X < P; X++) {
                R gv = gprj[((size_t)Z * A + a) * P + X];
                R xi = (R)X - (R)P * (R)0.5 - (shift ? shift[a] : (R)0) - center;
                R sum = (R)0;
                for (int k = 0; k < D; k++) {
                    R fu, fv;
                    S(align_coord)(g, c, s, xi, k, &fu, &fv);
                    R fi = RFLOOR(fu), fj = RFLOOR(fv);
                    if (!(fi > (R)-4 && fi < (R)(nx + 4) && fj > (R)-4 && fj < (R)(ny + 4))) continue;
                    long i0 = (long)fi, j0 = (long)fj;
                    R wa = fu - fi, wb = fv - fj;
                    R t00 = S(texel)(layer, nx, ny, i0, j0), t10 = S(texel)(layer, nx, ny, i0 + 1, j0);
                    R t01 = S(texel)(layer, nx, ny, i0, j0 + 1), t11 = S(texel)(layer, nx, ny, i0 + 1, j0 + 1);
                    R du = ((R)1 - wb) * (t10 - t00) + wb * (t11 - t01);
                    R dv = ((R)1 - wa) * (t01 - t00) + wa * (t11 - t10);
                    sum += -c * du + s * dv;
                }
                acc += gv * sum;
            }
        }
        gshift[a] = acc;
    }
}

/* ---- rotation-centre search (RotationCenterSearch/) ---------------------- */

/* rotCenterShift, rotCenterSearch.cl:35-48, over the candidates of the host
 * loop rotationCenterSearch_ocl.cpp:106-147: out[i] samples the one-layer
 * sinogram linearly at (X + shift_i + 0.5, th + 0.5), shift_i = start + step*i
 * in float (rotationCenterSearch_ocl.cpp:139). */
void S(ctor_rot_center_shift)(const ctor_geom *g, const R *prj, R *out, int n, float start, float step)
{
    const int P = g->px, A = g->pa;
    for (int i = 0; i < n; i++) {
        const float shift = start + step * (float)i;
        for (int th = 0; th < A; th++)
            for (int X = 0; X < P; X++)
                out[((size_t)i * A + th) * P + X] =
                    S(sample_linear)(prj, P, A, ((R)X + (R)shift) + (R)0.5, (R)th + (R)0.5);
    }
}

/* imgFocusIndex, rotCenterSearch.cl:154-227, with setMask's mask of candidate
 * i (ctor_rcs_params): over the pixels whose 3 x 3 block lies in the mask,
 * F = sqrt(mean of sum_8 (a_n - a)^2 / a^2).  The per-pixel terms follow the
 * kernel's order in R; the sums over pixels are kept in double (the kernel's
 * work-item and tree order is not reproduced). */
void S(ctor_focus_index)(const ctor_geom *g, const R *img, R *findex, int n, float start, float step,
                         float ang_min, float ang_max)
{
    const int nx = g->nx, ny = g->ny;
    float e[4];
    ctor_rcs_angles(ang_min, ang_max, e);
    #pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; i++) {
        const ctor_rcs_mask_t m = ctor_rcs_params(start + step * (float)i, e, nx, ny);
        const R *p = img + (size_t)i * nx * ny;
        double sum = 0.0, cnt = 0.0;
        for (int Y = 0; Y < ny; Y++)
            for (int X = 0; X < nx; X++) {
                int all = 1;
                for (int dy = -1; dy <= 1 && all; dy++)
                    for (int dx = -1; dx <= 1 && all; dx++) all = ctor_rcs_in(&m, X + dx, Y + dy, nx, ny);
                if (!all) continue;
                const size_t k = (size_t)Y * nx + X;
                const R a = p[k];
                const R nb[8] = {p[k + nx], p[k - nx], p[k + 1], p[k - 1],
                                 p[k + nx + 1], p[k - nx + 1], p[k + nx - 1], p[k - nx - 1]};
                R sq = (R)0;
                for (int j = 0; j < 8; j++) sq += (nb[j] - a) * (nb[j] - a);
                sq /= a * a;
                sum += (double)sq;
                cnt += 1.0;
            }
        findex[i] = (R)sqrt(sum / cnt);
    }
}

/* rotationCenterSearch from the resliced layer on (rotationCenterSearch_ocl.cpp:
 * 489-551): candidates, reconstRotCntShift_thread (rotationCenterSearch_ocl.cpp:
 * 281-331: fill with init, OSEM_execution(it, prjCorrection)), focus indices. */
void S(ctor_rotation_center_search)(const ctor_geom *g, const float *ang, const int *sub, const R *prj, int n,
                                    float start, float step, float ang_min, float ang_max, int it, float init,
                                    int prj_correction, int correction_mode, R *recon, R *findex)
{
    const size_t np = (size_t)n * g->pa * g->px, ni = (size_t)n * g->nx * g->ny;
    R *sh = (R *)malloc(sizeof(R) * np);
    S(ctor_rot_center_shift)(g, prj, sh, n, start, step);
    for (size_t i = 0; i < ni; i++) recon[i] = (R)init;
    S(ctor_osem_execution)(g, ang, sub, recon, sh, n, it, prj_correction, correction_mode);
    S(ctor_focus_index)(g, recon, findex, n, start, step, ang_min, ang_max);
    free(sh);
}

/* ---- sinogram formation (Reslice/, ImageReg share source/) --------------- */

/* merge_rawhisdata, imageregistration_kernel_src.cl:868-881. */
void S(ctor_merge_frames)(const unsigned short *raw, int n, int W, int H, long long stride, R *out)
{
    const size_t m = (size_t)W * H;
    for (size_t i = 0; i < m; i++) {
        R v = (R)0;
        for (int k = 0; k < n; k++) v += (R)raw[i + (size_t)stride * k];
        out[i] = v / (R)n;
    }
}

/* mt_conversion, imageregistration_kernel_src.cl:161-223 (the mu-t value). */
void S(ctor_mt_conversion)(const R *dark, const R *I0, const unsigned short *It, int n, int W, int H,
                           long long stride, R *mt)
{
    const size_t m = (size_t)W * H;
    for (int z = 0; z < n; z++)
        for (size_t i = 0; i < m; i++) {
            const R it = (R)It[i + (size_t)stride * z];
            const R d = dark[i];
            R tr = (I0[i] - d) / (it - d);
            tr = (tr < (R)1.0e-5f) ? (R)1.0e-5f : tr;
            tr = (tr > (R)1.0e7f) ? (R)1.0e7f : tr;
            tr = isnan(tr) ? (R)1 : tr;
            R v = RLOG(tr);
            v = isnan(v) ? (R)0 : v;
            mt[(size_t)z * m + i] = v;
        }
}

/* xprojection, reslice.cl:72-88: xproj[th][Y] = sum_{j = sx..ex} of the linear
 * sample (border 0) at the texel corner (j, Y), added in order of j to a zero
 * start (the reference's buffer is never cleared, reslice_ocl.cpp:114). */
void S(ctor_xprojection)(const R *mt, int A, int H, int W, int sx, int ex, R *xproj)
{
    #pragma omp parallel for schedule(static)
    for (int th = 0; th < A; th++)
        for (int Y = 0; Y < H; Y++) {
            R acc = (R)0;
            for (int j = sx; j <= ex; j++)
                acc += S(sample_linear)(mt + (size_t)th * H * W, W, H, (R)j, (R)Y);
            xproj[(size_t)th * H + Y] = acc;
        }
}

/* reduction(), reslice.cl:35-48, over 256 work-item partial sums. */
static R S(zc_tree)(R *loc)
{
    for (int s2 = 128; s2 > 0; s2 >>= 1)
        for (int l = 0; l < s2; l++) loc[l] += loc[l + s2];
    return loc[0];
}

/* zcorrection, reslice.cl:90-219, for a 256-item work-group: item l adds rows
 * l, l + 256, ... in order (all H rows), then reduction().  The xprojection
 * image is read with clamp-to-edge at texel corners (Y, angle); the target is
 * the profile at angle A/2.  delta and nyu start at 0 and 2 (unset in the
 * reference before the first trial, reslice.cl:111-114). */
void S(ctor_zcorrection)(const R *xp, int A, int H, int sZ, int eZ, R *yshift)
{
    #pragma omp parallel for schedule(dynamic)
    for (int ang = 0; ang < A; ang++) {
        R *target = (R *)malloc(sizeof(R) * (size_t)H);
        R l1[256], l2[256], l3[256];
        const R vt = (R)(A / 2), va = (R)ang;
        for (int Y = 0; Y < H; Y++) target[Y] = S(sample_linear_edge)(xp, H, A, (R)Y, vt);
        R ys = yshift[ang];
        R lambda = (R)1.0e-3f, nyu = (R)2, delta = (R)0;
        for (int trial = 0; trial < 20; trial++) {
            for (int l = 0; l < 256; l++) {
                R tJJ = (R)0, tJdF = (R)0, chi2 = (R)0;
                for (int Y = l; Y < H; Y += 256) {
                    const R s0 = S(sample_linear_edge)(xp, H, A, (R)Y, va);
                    const R sm1 = S(sample_linear_edge)(xp, H, A, ((R)Y + ys) - (R)1, va);
                    const R sp1 = S(sample_linear_edge)(xp, H, A, ((R)Y + ys) + (R)1, va);
                    const R dFdy = (sp1 - sm1) / (R)2;
                    const R mask = (Y >= sZ && Y <= eZ) ? (R)1 : (R)0;
                    const R dF = target[Y] - s0;
                    tJJ += dFdy * dFdy * mask;
                    tJdF += dFdy * dF * mask;
                    chi2 += dF * dF * mask;
                }
                l1[l] = tJJ; l2[l] = tJdF; l3[l] = chi2;
            }
            const R tJJ = S(zc_tree)(l1), tJdF = S(zc_tree)(l2), chi2_old = S(zc_tree)(l3);
            const R cand = ys + delta;
            for (int l = 0; l < 256; l++) {
                R chi2 = (R)0;
                for (int Y = l; Y < H; Y += 256) {
                    const R sv = S(sample_linear_edge)(xp, H, A, (R)Y + cand, va);
                    const R mask = (Y >= sZ && Y <= eZ) ? (R)1 : (R)0;
                    const R dF = target[Y] - sv;
                    chi2 += dF * dF * mask;
                }
                l3[l] = chi2;
            }
            const R chi2_new = S(zc_tree)(l3);
            delta = tJdF / tJJ / ((R)1 + lambda);
            const R d_rho = delta * (lambda * delta + tJdF);
            const R rho = (chi2_old - chi2_new) / d_rho;
            const R q = (R)2 * rho - (R)1;
            const R lA = (R)1 - q * q * q;
            const R lB = lambda * (lA > (R)0.333f ? lA : (R)0.333f);
            if (rho > (R)0) {
                lambda = lB;
                nyu = (R)2;
                ys += delta;
            } else {
                lambda *= nyu;
                nyu *= (R)2;
            }
        }
        yshift[ang] = ys;
        free(target);
    }
}

/* reslice, reslice.cl:52-70: prj[l][th][X], camera row Y = y0 + l. */
void S(ctor_reslice)(const R *mt, int A, int H, int W, const R *xs, const R *ys, float baseup, int correction,
                     int y0, int nyo, R *prj)
{
    #pragma omp parallel for schedule(static)
    for (int l = 0; l < nyo; l++)
        for (int th = 0; th < A; th++)
            for (int X = 0; X < W; X++) {
                const int Y = y0 + l;
                const R dx = xs ? xs[th] : (R)0, dy = ys ? ys[th] : (R)0;
                R v = S(sample_linear)(mt + (size_t)th * H * W, W, H, ((R)X + dx) + (R)0.5, ((R)Y + dy) + (R)0.5);
                v += (R)baseup;
                if (correction == 1) {
                    const R t = (R)2 * RFABS((R)X - (R)W / (R)2) / (R)W;
                    v *= RCOS(RASIN(t));
                }
                prj[((size_t)l * A + th) * W + X] = v;
            }
}

/* denoiseSinogram, reslice.cl:222-248. */
void S(ctor_denoise_sinogram)(const R *src, R *dst, int nz, int A, int W, float threshold)
{
    const R thr = (R)threshold;
    for (int z = 0; z < nz; z++)
        for (int th = 0; th < A; th++)
            for (int X = 0; X < W; X++) {
                const R *b = src + (size_t)z * A * W;
                const R f1 = b[(size_t)th * W + X];
                const R f2 = b[(size_t)(th + 1 < A ? th + 1 : A - 1) * W + X];
                const R f3 = b[(size_t)(th > 0 ? th - 1 : 0) * W + X];
                const R d1 = f2 - f1, d2 = f1 - f3;
                const R corr = (f2 + f3) / (R)2;
                dst[((size_t)z * A + th) * W + X] = ((d1 > thr && d2 < -thr) || (d1 < -thr && d2 > thr)) ? corr : f1;
            }
}

/* reslice_thread, reslice_ocl.cpp:258-350. */
void S(ctor_reslice_execution)(const R *mt, int A, int H, int W, float baseup, int z_correction, int sx, int ex,
                               int sz, int ez, int denoise, float threshold, int y0, int nyo, R *yshift, R *prj)
{
    for (int a = 0; a < A; a++) yshift[a] = (R)0;
    if (z_correction) {
        R *xp = (R *)malloc(sizeof(R) * (size_t)A * H);
        S(ctor_xprojection)(mt, A, H, W, sx, ex, xp);
        S(ctor_zcorrection)(xp, A, H, sz, ez, yshift);
        free(xp);
    }
    if (denoise) {
        R *tmp = (R *)malloc(sizeof(R) * (size_t)nyo * A * W);
        S(ctor_reslice)(mt, A, H, W, NULL, yshift, baseup, 0, y0, nyo, tmp);
        S(ctor_denoise_sinogram)(tmp, prj, nyo, A, W, threshold);
        free(tmp);
    } else {
        S(ctor_reslice)(mt, A, H, W, NULL, yshift, baseup, 0, y0, nyo, prj);
    }
}
