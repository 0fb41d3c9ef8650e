/*
 * ct_oracle.h -- CPU restatement of the CT_reconstruction projector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it.  Nothing in the product library links or calls it.
 *
 * Every operator is restated from the reference OpenCL C text
 * (CT_reconstruction/{OSEM,AART,FBP,FISTA,reconstShare}.cl) and the host
 * drivers in CT_reconstruction/GPU.cpp.  Each function is compiled twice:
 *   *_f32  -- "compat": float arithmetic in the reference's evaluation order,
 *             OpenCL 1.2 image-sampling semantics emulated in software
 *             (CLK_ADDRESS_CLAMP border 0, texel centres at i+0.5, layer clamp);
 *   *_f64  -- the same operators evaluated in double (the FP64 oracle).
 *
 * Layouts (row-major, x fastest, identical to the reference host buffers):
 *   image stack     [nz][ny][nx]
 *   sinogram stack  [nz][pa][px]
 *
 * Pinning status of each function is recorded in the tests that cover it
 * (tests/test_oracle_pins.py); see the header comment there.
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry: g_nx, g_ny, g_px, g_pa, g_ss, g_zp of CT_reconstruction.hpp:110-124.
 * The output ROI (g_ox, g_oy) always equals (nx, ny) here. */
typedef struct {
    int nx, ny;   /* reconstruction grid */
    int px;       /* detector bins */
    int pa;       /* projection angles */
    int ss;       /* ordered subsets */
    int zp;       /* FFT length (FBP) */
} ctor_geom;

/* GPU.cpp:564-582 -- bit-reversal subset order when ss is a power of two. */
void ctor_subset_order(int ss, int *sub);

/* Per-precision entry points.  R = float for _f32, double for _f64. */
#define CTOR_DECLARE(R, S)                                                              \
void ctor_osem1_##S(const ctor_geom *g, const float *ang, const R *img, const R *prj,   \
                    R *rprj, int nz, int sub);                                           \
void ctor_osem2_##S(const ctor_geom *g, const float *ang, const R *img, const R *rprj,  \
                    R *img_dst, int nz, int sub);                                        \
void ctor_aart1_##S(const ctor_geom *g, const float *ang, const R *img, const R *prj,   \
                    R *dprj, int nz, int sub);                                           \
void ctor_aart2_##S(const ctor_geom *g, const float *ang, const R *img, const R *dprj,  \
                    R *img_dst, int nz, int sub, float alpha);                           \
void ctor_backprojection_##S(const ctor_geom *g, const float *ang, const R *prj,        \
                    R *img_dst, int nz, int sub);                                        \
void ctor_sinogram_correction_##S(const ctor_geom *g, const float *ang, const R *src,   \
                    R *dst, int nz, int mode, float a, float b);                         \
void ctor_fbp_filter_##S(const ctor_geom *g, const R *prj, R *fprj, int nz);            \
void ctor_fbp_backprojection_##S(const ctor_geom *g, const float *ang, const R *fprj,   \
                    R *img, int nz);                                                     \
void ctor_power_iter1_##S(const ctor_geom *g, const float *ang, const R *img, R *prjc); \
void ctor_power_iter2_##S(const ctor_geom *g, const float *ang, const R *prjc, R *cnd); \
void ctor_power_iter3_##S(const ctor_geom *g, const R *cnd, const R *l2, R *img);       \
void ctor_image_l2_##S(const ctor_geom *g, const R *img, R *l2, int nz);                \
void ctor_fista_backprojection_##S(const ctor_geom *g, const float *ang, const R *img,  \
                    const R *dprj, const R *L, R *img_dst, int nz, int sub, float alpha);\
void ctor_ista3d_##S(const ctor_geom *g, const R *v, R *x_new, int nz, int sub,         \
                    const R *L, double lambda_fista);                                    \
void ctor_fista3d_##S(const ctor_geom *g, const R *x, const R *v, const R *b,           \
                    R *w_new, R *x_new, R *b_new, int nz, int sub, const R *L,           \
                    double lambda_fista);                                                \
void ctor_tv_prox2d_##S(const ctor_geom *g, const R *v, R *x_new, int nz, R lambda,      \
                    R lower);                                                            \
void ctor_osem_execution_##S(const ctor_geom *g, const float *ang, const int *sub,      \
                    R *img, R *prj, int nz, int it, int prj_correction,                  \
                    int correction_mode);                                                \
void ctor_fista_osem_execution_##S(const ctor_geom *g, const float *ang, const int *sub,\
                    R *img, R *prj, int nz, int it, int prj_correction,                  \
                    int correction_mode, double cs_lambda);                              \
void ctor_aart_execution_##S(const ctor_geom *g, const float *ang, const int *sub,      \
                    R *img, R *prj, int nz, int it, int correction_mode, float alpha);   \
void ctor_fbp_execution_##S(const ctor_geom *g, const float *ang, R *img, R *prj,       \
                    int nz, int prj_correction, int correction_mode);                    \
void ctor_layer_min_##S(const ctor_geom *g, const R *img, R *mins, int nz);            \
void ctor_baseup_##S(const ctor_geom *g, R *img, const R *mins, int nz, int order);     \
void ctor_hybrid_execution_##S(const ctor_geom *g, const float *ang, const int *sub,    \
                    R *img, R *prj, int nz, int it, int prj_correction,                  \
                    int correction_mode, int baseup_order, int cs, double cs_lambda);    \
void ctor_power_iteration_##S(const ctor_geom *g, const float *ang, R *L, int trials);  \
void ctor_fista_execution_##S(const ctor_geom *g, const float *ang, const R *L,         \
                    const int *sub, R *img, R *prj, int nz, int it, int prj_correction, \
                    int correction_mode, float alpha, double cs_lambda);                 \
void ctor_align_forward_##S(const ctor_geom *g, const float *ang, const R *img,         \
                    const R *shift, R center, R *prj, int nz);                           \
void ctor_align_adjoint_##S(const ctor_geom *g, const float *ang, const R *prj,         \
                    const R *shift, R center, R *img, int nz);                           \
void ctor_align_grad_##S(const ctor_geom *g, const float *ang, const R *img,            \
                    const R *gprj, const R *shift, R center, R *gshift, int nz);         \
void ctor_rot_center_shift_##S(const ctor_geom *g, const R *prj, R *out, int n,        \
                    float start, float step);                                            \
void ctor_focus_index_##S(const ctor_geom *g, const R *img, R *findex, int n,           \
                    float start, float step, float ang_min, float ang_max);              \
void ctor_rotation_center_search_##S(const ctor_geom *g, const float *ang,              \
                    const int *sub, const R *prj, int n, float start, float step,        \
                    float ang_min, float ang_max, int it, float init, int prj_correction,\
                    int correction_mode, R *recon, R *findex);                        \
void ctor_merge_frames_##S(const unsigned short *raw, int n, int W, int H,              \
                    long long stride, R *out);                                            \
void ctor_mt_conversion_##S(const R *dark, const R *I0, const unsigned short *It,       \
                    int n, int W, int H, long long stride, R *mt);                        \
void ctor_xprojection_##S(const R *mt, int A, int H, int W, int sx, int ex, R *xproj);  \
void ctor_zcorrection_##S(const R *xp, int A, int H, int sZ, int eZ, R *yshift);        \
void ctor_reslice_##S(const R *mt, int A, int H, int W, const R *xs, const R *ys,       \
                    float baseup, int correction, int y0, int nyo, R *prj);               \
void ctor_denoise_sinogram_##S(const R *src, R *dst, int nz, int A, int W,              \
                    float threshold);                                                     \
void ctor_reslice_execution_##S(const R *mt, int A, int H, int W, float baseup,         \
                    int z_correction, int sx, int ex, int sz, int ez, int denoise,        \
                    float threshold, int y0, int nyo, R *yshift, R *prj);

CTOR_DECLARE(float, f32)
CTOR_DECLARE(double, f64)

/* setMask (RotationCenterSearch/rotCenterSearch.cl:50-92): mask [n][ny][nx]
 * (0/1) of the candidates shift_i = start + step * i. */
void ctor_rcs_mask(int nx, int ny, int n, float start, float step, float ang_min, float ang_max,
                   unsigned char *mask);

/* Number of OpenMP threads the library will use (1 when built without OpenMP). */
int ctor_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
