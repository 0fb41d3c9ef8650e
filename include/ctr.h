/*
 * ctr.h -- C ABI of the MI355X CT reconstruction hot path.
 *
 * The library (ct-xafs_analysis_amd/libctr_hip.so) replaces the OpenCL queue
 * that CT_reconstruction/GPU.cpp drives: the projector / backprojector kernels
 * of OSEM.cl, AART.cl, FBP.cl, FISTA.cl and reconstShare.cl and the execution
 * drivers that launch them.  Every entry point takes plain C types.
 *
 * Conventions
 *  - Buffers named reconst_img / prj_img / L2norm / img / prj are DEVICE
 *    pointers (hipMalloc'd or any HIP-visible device allocation), borrowed for
 *    the duration of the call and used on the context's stream.
 *  - Image stacks are [dN][ny][nx] float32, sinogram stacks [dN][pa][px]
 *    float32, row-major with x fastest: the layouts of the reference's
 *    Image2DArray objects (GPU.cpp:422-423) and of its raw files.
 *  - Calls are asynchronous on the context stream: the execution drivers and
 *    operators queue their uploads (through a pinned staging ring) and
 *    kernels and return; ctr_synchronize() waits.  These block the host:
 *    ctr_set_geometry and ctr_destroy (they wait for the context's queued
 *    work), ctr_autotune (it times kernels), ctr_osem_batched when its slice
 *    list differs from the previous call's (it rebuilds its plan), any call
 *    that must grow a workspace buffer (the old one is freed, which waits
 *    for the device), ctr_set_schedule and ctr_set_adaptive_rows (when it
 *    changes the setting) after a batched call, ctr_set_sync_each_kernel(1)
 *    mode, and kernel timing.
 *    A context is thread-compatible, not thread-safe: use one context per
 *    host thread.
 *  - Errors are returned as ctr_status; ctr_last_error() gives the message.
 *    The library never exits or aborts the caller's process.
 */
#ifndef CTR_H
#define CTR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ctr_ctx ctr_ctx;

typedef enum ctr_status {
    CTR_OK = 0,
    CTR_INVALID_ARGUMENT = 1,
    CTR_INVALID_GEOMETRY = 2,
    CTR_NO_DEVICE = 3,
    CTR_OUT_OF_MEMORY = 4,
    CTR_HIP_ERROR = 5,
    CTR_IO_ERROR = 6,
    CTR_NOT_SUPPORTED = 7
} ctr_status;

/* Reconstruction geometry and algorithm constants.  Replaces the globals of
 * CT_reconstruction.hpp:110-135 that the reference compiles into each kernel
 * program with -D (GPU.cpp:538-547, 948-954, 1336-1343, 2051-2059). */
typedef struct ctr_geometry {
    int nx, ny;           /* g_nx, g_ny: reconstruction grid */
    int ox, oy;           /* g_ox, g_oy: output ROI, 1..nx x 1..ny; below the grid
                             size for ctr_fbp_execution only */
    int px;               /* g_px: detector bins */
    int pa;               /* g_pa: projection angles */
    int ss;               /* g_ss: ordered subsets; pa % ss == 0 */
    int zp;               /* g_zp: FBP zero-padding length, power of two >= px */
    int mode;             /* g_mode: 3 (ASIRT) selects wt2 as the AART factor */
    int correction_mode;  /* correctionMode: 0 off, 1-4 (reconstShare.cl:53-68) */
    float wt1, wt2;       /* g_wt1, g_wt2: AART / ASIRT relaxation */
    float cs_lambda;      /* CSlambda: the LAMBDA_FISTA build constant */
} ctr_geometry;

/* One slice of an uneven batched stack: square n x n image, n detector bins,
 * pa angles shared by the whole batch (the ctr_set_geometry angle table). */
typedef struct ctr_slice {
    int n;                /* side: nx = ny = px = n */
    float *img;           /* device [n][n], initial image in, result out */
    const float *prj;     /* device [pa][n] sinogram */
} ctr_slice;

/* ---- context ----------------------------------------------------------- */

/* Creates a context on HIP device `device` with its own non-blocking stream.
 * Replaces OCL_platform_device (share source/OCL_platform_device_class.cpp:55-208):
 * one device + one in-order queue. */
ctr_status ctr_create(int device, ctr_ctx **out);
ctr_status ctr_destroy(ctr_ctx *ctx);
const char *ctr_last_error(const ctr_ctx *ctx);
const char *ctr_status_string(ctr_status status);

/* Uses `hip_stream` (a hipStream_t) instead of the context's own stream;
 * NULL restores the context stream (so the null default stream cannot be
 * selected: pass a created stream). */
ctr_status ctr_set_stream(ctr_ctx *ctx, void *hip_stream);
/* Writes the hipStream_t the context's work runs on (its own stream, or the
 * one given to ctr_set_stream), so a caller can order its own streams with it.
 * No reference counterpart (an OpenCL queue is the caller's object there). */
ctr_status ctr_get_stream(const ctr_ctx *ctx, void **hip_stream);
ctr_status ctr_synchronize(ctr_ctx *ctx);

/* 1: synchronise after every kernel launch, as GPU.cpp does with finish()
 * after each enqueue (GPU.cpp:216-227).  0 (default): stream-ordered only. */
ctr_status ctr_set_sync_each_kernel(ctr_ctx *ctx, int enable);

/* Sets the geometry, the host angle table (pa float degrees, the angle log of
 * fileIO.cpp:11-24) and the subset order (ss ints; NULL = the reference order
 * of GPU.cpp:564-582).  Replaces *_programBuild (GPU.cpp:513-559, 922-995,
 * 1311-1354, 2023-2077) and the angle-buffer upload (GPU.cpp:619-624). */
ctr_status ctr_set_geometry(ctr_ctx *ctx, const ctr_geometry *geom,
                            const float *angles_deg, const int *sub);

/* Writes the reference subset order for `ss` subsets (GPU.cpp:564-582). */
ctr_status ctr_subset_order(int ss, int *sub);

/* The device layout of one slice of nx x ny pixels and px bins: the zero frame
 * (pad texels on every side) and row pitches of the padded image and of its
 * transposed copy, and the compact subset rows' pitch with roff zeros before
 * bin 0 (and at least as many after bin px - 1).  Host-only: no device call.
 * No reference counterpart (OpenCL images hide their layout); for callers
 * that size workspace and for layout checks. */
typedef struct ctr_layout {
    int pad, ipitch, tpitch, rpitch, roff;
} ctr_layout;
ctr_status ctr_query_layout(int nx, int ny, int px, ctr_layout *out);

/* ---- execution drivers (drop-ins for GPU.cpp) --------------------------- */

/* OSEM_execution, GPU.cpp:125-240 (declared CT_reconstruction.hpp:77). */
ctr_status ctr_osem_execution(ctr_ctx *ctx, float *reconst_img, float *prj_img,
                              int dN, int it, int prj_correction);

/* Bytes of the per-device scratch array of ctr_osem_execution_noalloc. */
#define CTR_NOALLOC_SCRATCH_BYTES ((unsigned long long)64 << 20)

/* ctr_osem_execution that allocates nothing: no hipMalloc, hipFree or other
 * allocating call, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No reference
 * counterpart.  Its workspace is a file-scope device array of
 * CTR_NOALLOC_SCRATCH_BYTES (64 MiB) that each device the library runs on
 * holds until the process ends: per layer the padded image and its
 * transposed copy, (nx + 200 rounded up to 16) x (ny + 200) floats each, and
 * pa/ss subset rows of px + 2 roff floats (ctr_query_layout), plus a few KB
 * of descriptors -- e.g. up to 5 layers of 1024^2 or 20 of 512^2.  A larger
 * stack returns CTR_INVALID_ARGUMENT before anything runs.  The results equal
 * ctr_osem_execution's bit for bit.  Every call on one device shares that
 * device's array: calls that overlap on one device (two contexts, or two
 * host threads) race and give wrong results; calls on one context are
 * ordered by its stream. */
ctr_status ctr_osem_execution_noalloc(ctr_ctx *ctx, float *reconst_img, float *prj_img,
                                      int dN, int it, int prj_correction);

/* FISTA_OSEM_execution, GPU.cpp:243-405 (CT_reconstruction.hpp:78-81). */
ctr_status ctr_fista_osem_execution(ctr_ctx *ctx, float *reconst_img, float *prj_img,
                                    int dN, int it, int prj_correction);

/* AART_execution, GPU.cpp:1088-1216 (CT_reconstruction.hpp:82). */
ctr_status ctr_aart_execution(ctr_ctx *ctx, float *reconst_img, float *prj_img,
                              int dN, int it);

/* FBP_execution, GPU.cpp:665-819 (log2(zp) stages, GPU.cpp:855).  Writes the
 * whole nx x ny grid of each layer.  With an output ROI (ox < nx or oy < ny)
 * the reference's FBP_thread returns the top-left ox x oy block of it
 * (backProjectionFBP over an ox x oy range in full-grid coordinates,
 * GPU.cpp:704, read back from origin (0, 0), GPU.cpp:848-870).  The iterative
 * drivers reject an ROI (CTR_INVALID_GEOMETRY). */
ctr_status ctr_fbp_execution(ctr_ctx *ctx, float *reconst_img, float *prj_img,
                             int dN, int prj_correction);

/* Mode 8, the FBP + OS-EM hybrid (hybrid_thread, GPU.cpp:1459-1580): FBP of
 * the sinograms (projection correction when prj_correction != 0), a base-up of
 * each FBP layer that maps its minimum to 0 (findMinimumX/Y and baseUp,
 * reconstShare.cl:88-177, series order `baseup_order`, 4 in the reference),
 * then OSEM_execution (FISTA_OSEM_execution when cs != 0) from that image. */
ctr_status ctr_hybrid_execution(ctr_ctx *ctx, float *reconst_img, float *prj_img, int dN, int it,
                                int prj_correction, int baseup_order, int cs);

/* powerIteration, GPU.cpp:2079-2188: writes L2norm[ss] (device). `trials` is
 * 30 in the reference. */
ctr_status ctr_power_iteration(ctr_ctx *ctx, float *L2norm, int trials);

/* FISTA_execution, GPU.cpp:1755-1924 (CT_reconstruction.hpp:83), with the
 * reference's kernel-argument bindings (the returned image is the first
 * ISTA3D step for any it >= 1). */
ctr_status ctr_fista_execution(ctr_ctx *ctx, const float *L2norm, float *reconst_img,
                               float *prj_img, int dN, int it, int prj_correction);

/* ---- sinogram formation (Reslice/, mt conversion of ImageReg share source/)
 * Projection stacks are [A][H][W] float32 (angle, camera row, camera column:
 * the mu-t images the reference reads one per angle); the sinograms they
 * become are [layers][A][W], the sinogram stack layout above with pa = A,
 * px = W, so they feed the reconstruction drivers directly. */

/* merge_rawhisdata, ImageReg share source/imageregistration_kernel_src.cl:
 * 868-881: out [H][W] = mean of the n 16-bit frames raw + k * frame_stride
 * (frame_stride = W*H + 32 for the frames of a .his file,
 * share source/readHisRawFile.cpp:63-95). */
ctr_status ctr_merge_frames(ctr_ctx *ctx, const unsigned short *raw, int n, int W, int H,
                            long long frame_stride, float *out);

/* mt_conversion, imageregistration_kernel_src.cl:161-223, the mu-t value (its
 * registration mask is not formed): mt [n][H][W] = ln((I0 - dark) / (It_k -
 * dark)) with the transmission clamped to [1e-5, 1e7], a NaN transmission read
 * as 1 and a NaN logarithm as 0; frame k is It + k * frame_stride. */
ctr_status ctr_mt_conversion(ctr_ctx *ctx, const float *dark, const float *I0, const unsigned short *It,
                             int n, int W, int H, long long frame_stride, float *mt);

/* xprojection, Reslice/reslice.cl:72-88 (Z_Correction, reslice_ocl.cpp:89-150):
 * xproj [A][H] = for each angle and row, the sum over columns start_x..end_x of
 * the linear sample at the texel corner (X, Y). */
ctr_status ctr_xprojection(ctr_ctx *ctx, const float *mt, int A, int H, int W, int start_x, int end_x,
                           float *xproj);

/* zcorrection, reslice.cl:90-219: for each angle, 20 Levenberg-Marquardt
 * trials (lambda 0.001) on the vertical drift yshift[A] (device, in/out) that
 * matches the angle's x-projection profile to the middle angle's over rows
 * start_z..end_z.  H <= 8192. */
ctr_status ctr_zcorrection(ctr_ctx *ctx, const float *xproj, int A, int H, int start_z, int end_z,
                           float *yshift);

/* reslice, reslice.cl:52-70 (host reslice, reslice_ocl.cpp:153-208):
 * prj [n_layers][A][W], layer l = camera row Y = y0 + l: row th, bin X is the
 * linear sample (border 0) of projection th at (X + xshift[th] + 0.5,
 * Y + yshift[th] + 0.5) plus baseup, times cos(asin(2|X - W/2| / W)) when
 * correction == 1.  xshift / yshift: device [A] or NULL (zero). */
ctr_status ctr_reslice(ctr_ctx *ctx, const float *mt, int A, int H, int W, const float *xshift,
                       const float *yshift, float baseup, int correction, int y0, int n_layers,
                       float *prj);

/* denoiseSinogram, reslice.cl:222-248: a bin whose differences to both theta
 * neighbours exceed `threshold` with opposite signs becomes their mean (rows
 * past the ends clamp to the edge).  src, dst: [nz][A][W], src != dst. */
ctr_status ctr_denoise_sinogram(ctr_ctx *ctx, const float *src, float *dst, int nz, int A, int W,
                                float threshold);

/* The reslice parameters of share source/input_parameter_reslice_class.cpp. */
typedef struct ctr_reslice_params {
    float baseup;            /* "#Image baseup" */
    int z_correction;        /* "#Z-correction" == 1 */
    int start_x, end_x;      /* "#X range of correction evaluation" */
    int start_z, end_z;      /* "#Z range of correction evaluation" */
    int denoise;             /* "#Noise removal" */
    float denoise_threshold; /* "#Noise removal threshold" */
    int y0, n_layers;        /* camera rows to reslice: y0 .. y0 + n_layers - 1 */
} ctr_reslice_params;

/* reslice_thread, reslice_ocl.cpp:258-350: yshift[A] (device out) = 0, then
 * Z_Correction (xprojection + zcorrection) when z_correction, reslice with
 * those vertical shifts (the x shifts stay 0, as in the reference), and
 * denoiseSinogram when denoise.  prj: [n_layers][A][W]. */
ctr_status ctr_reslice_execution(ctr_ctx *ctx, const float *mt, int A, int H, int W,
                                 const ctr_reslice_params *params, float *yshift, float *prj);

/* ---- rotation-centre search (RotationCenterSearch/) --------------------- */

/* rotCenterShift, rotCenterSearch.cl:35-48, over the host loop of
 * rotationCenterSearch_ocl.cpp:106-147: out[i] is the one-layer sinogram prj
 * [pa][px] resampled along the detector at X + shift_start + shift_step * i
 * (linear, border 0), for i < n_shift.  out: [n_shift][pa][px]. */
ctr_status ctr_op_rot_center_shift(ctr_ctx *ctx, const float *prj, float *out, int n_shift,
                                   float shift_start, float shift_step);

/* setMask + imgFocusIndex, rotCenterSearch.cl:50-92, 154-227 (host:
 * imgFocusIndex_estimation, rotationCenterSearch_ocl.cpp:181-217).  For layer
 * i of img [n][ny][nx]: the mask of the region that candidate shift
 * shift_start + shift_step * i keeps on the detector over the angle range
 * [ang_min, ang_max] (degrees), and the focus index
 * sqrt(mean over pixels whose 3 x 3 block lies in the mask of
 * sum_8 (a_n - a)^2 / a^2).  findex: n device floats. */
ctr_status ctr_op_focus_index(ctr_ctx *ctx, const float *img, int n, float shift_start,
                              float shift_step, float ang_min, float ang_max, float *findex);

/* rotationCenterSearch, rotationCenterSearch_ocl.cpp:489-551, from the
 * resliced layer on: the candidate sinograms (ctr_op_rot_center_shift), each
 * reconstructed from the constant `init` (g_ini) by OSEM_execution with `it`
 * iterations and the projection correction when prj_correction != 0
 * (reconstRotCntShift_thread, rotationCenterSearch_ocl.cpp:281-331), and their
 * focus indices (ctr_op_focus_index).  recon: [n_shift][ny][nx] device output;
 * findex: n_shift device floats.  The candidate of the smallest focus index is
 * the estimated rotation-centre shift (rotationCenterSearch_ocl.cpp:521-549). */
ctr_status ctr_rotation_center_search(ctr_ctx *ctx, const float *prj, int n_shift, float shift_start,
                                      float shift_step, float ang_min, float ang_max, int it, float init,
                                      int prj_correction, float *recon, float *findex);

/* ---- single operators (device buffers, dN layers) ----------------------- */

/* OSEM1, OSEM.cl:38-68: writes rows th = sub + j*ss of rprj [dN][pa][px]. */
ctr_status ctr_op_osem1(ctr_ctx *ctx, const float *img, const float *prj, float *rprj,
                        int dN, int sub);
/* OSEM2, OSEM.cl:71-105: img_dst = img * BP(rprj) * ss / pa. */
ctr_status ctr_op_osem2(ctr_ctx *ctx, const float *img, const float *rprj, float *img_dst,
                        int dN, int sub);
/* AART1, AART.cl:38-69: rows of dprj = y - A x. */
ctr_status ctr_op_aart1(ctr_ctx *ctx, const float *img, const float *prj, float *dprj,
                        int dN, int sub);
/* AART2, AART.cl:73-106. */
ctr_status ctr_op_aart2(ctr_ctx *ctx, const float *img, const float *dprj, float *img_dst,
                        int dN, int sub, float alpha);
/* backProjection, reconstShare.cl:245-275. */
ctr_status ctr_op_backprojection(ctr_ctx *ctx, const float *prj, float *img_dst,
                                 int dN, int sub);
/* sinogramCorrection, reconstShare.cl:35-70 (a = b = 0.5 as the host passes). */
ctr_status ctr_op_sinogram_correction(ctr_ctx *ctx, const float *src, float *dst,
                                      int dN, int mode);
/* FBP.cl:34-160 chain (spinFactor .. outputImage): ramp-filtered rows. */
ctr_status ctr_op_fbp_filter(ctr_ctx *ctx, const float *prj, float *fprj, int dN);
/* backProjectionFBP, FBP.cl:164-188. */
ctr_status ctr_op_fbp_backprojection(ctr_ctx *ctx, const float *fprj, float *img, int dN);
/* powerIter1, FISTA.cl:50-75: img and prjc hold ss layers; rows j < pa/ss. */
ctr_status ctr_op_power_iter1(ctr_ctx *ctx, const float *img, float *prjc);
/* powerIter2, FISTA.cl:79-110. */
ctr_status ctr_op_power_iter2(ctr_ctx *ctx, const float *prjc, float *cnd);
/* FISTAbackProjection, FISTA.cl:211-246. */
ctr_status ctr_op_fista_backprojection(ctr_ctx *ctx, const float *img, const float *dprj,
                                       const float *L2norm, float *img_dst, int dN, int sub,
                                       float alpha);
/* ISTA3D, FISTA.cl:409-493 (lambda = cs_lambda / L2norm[sub] * 0.5). */
ctr_status ctr_op_ista3d(ctr_ctx *ctx, const float *v, float *x_new, int dN, int sub,
                         const float *L2norm);
/* FISTA3D, FISTA.cl:497-594.  Any output may be NULL to skip it; when
 * x_new == b_new the beta value is what remains (reference write order). */
ctr_status ctr_op_fista3d(ctr_ctx *ctx, const float *x, const float *v, const float *b,
                          float *w_new, float *x_new, float *b_new, int dN, int sub,
                          const float *L2norm);

/* Per-slice TV prox: x_new = argmin 1/2 (x - v)^2 + lambda sum |x - n_i| over
 * the four in-plane neighbours n_i of each pixel (texels outside the slice
 * read 0), i.e. the median of n_0..n_3 and v - 4 lambda .. v + 4 lambda in
 * steps of 2 lambda; then max(lower, x), NaN -> lower.  v, x_new: device
 * [dN][ny][nx], distinct buffers.  No reference counterpart: the 2-D form of
 * ISTA3D (FISTA.cl:409-493) for the joint alignment model, whose slices are
 * unrelated, so unlike ISTA3D no slice is coupled to another. */
ctr_status ctr_tv_prox_2d(ctr_ctx *ctx, const float *v, float *x_new, int dN, float lambda, float lower);

/* ---- batched uneven stacks (no reference counterpart: the reference runs one
 *      layer batch of equal size per OSEM_execution call) ------------------- */

/* Runs OS-EM (OSEM_execution semantics, no correction) on `count` slices of
 * possibly different sides in one stream-ordered schedule: per subset, one
 * launch per kernel covers every slice.  use_graph != 0 captures one
 * iteration in a hipGraph and replays it.  Workspace and graphs are cached
 * in the context for repeated calls with the same slice list. */
ctr_status ctr_osem_batched(ctr_ctx *ctx, const ctr_slice *slices, int count, int it,
                            int use_graph);

/* A projector / backprojector schedule for one slice side.  No reference
 * counterpart: the reference fixes its work-group shape per compiled program
 * (GPU.cpp:169-172, 538-547). */
typedef struct ctr_schedule {
    int fp_rays;    /* 0: the LDS-tiled projector; 64 / 32 / 16 / 8: the direct-gather one, rays per wave */
    int fp_waves;   /* waves per projection block: 4 for the LDS-tiled projector, 2 or 4 otherwise */
    int fp_tile;    /* samples per LDS tile: 32, 44 or 60 (64 tiles must cover the side) */
    int bp_rows;    /* backprojection rows per thread: 1, 2, 4, 8 or 16 (small stacks may
                       run with fewer: ctr_set_adaptive_rows) */
    float fp_ms, bp_ms;                  /* ctr_autotune: the chosen candidates' launch times */
    float fp_ms_default, bp_ms_default;  /* ... and the default schedule's (tile 60, 8 rows) */
} ctr_schedule;

/* Times the projector schedules (LDS tiles of 60 / 44 / 32 samples, direct
 * gathers of 64 / 32 / 16 / 8 rays per wave with 2 or 4 waves per block) and
 * the backprojection row counts on a batch-sized launch of side n, keeps the
 * fastest for side n and writes it (with its and the default's times) to
 * *best.  Blocks the host.  Batched stacks run each side's schedule. */
ctr_status ctr_autotune(ctr_ctx *ctx, int n, ctr_schedule *best);
/* Sets the schedule of side n (the time fields are ignored). */
ctr_status ctr_set_schedule(ctr_ctx *ctx, int n, const ctr_schedule *schedule);
/* The schedule side n runs (the defaults unless tuned or set). */
ctr_status ctr_get_schedule(ctr_ctx *ctx, int n, ctr_schedule *schedule);
/* enable != 0 (the default): a stack whose backprojection would have fewer
 * than 4 blocks per compute unit at its side's bp_rows runs with bp_rows
 * halved until it has them or reaches 1 (one 512^2 slice: 1 row, 1024
 * blocks).  0: every stack runs its side's bp_rows as set.  The images are
 * the same either way (every row count sums each pixel's angles in the same
 * order); only the launch shape changes.  No reference counterpart. */
ctr_status ctr_set_adaptive_rows(ctr_ctx *ctx, int enable);

/* Average duration (ms) of the last timed forward-projection and
 * backprojection launches recorded with ctr_set_kernel_timing(ctx, 1). */
ctr_status ctr_set_kernel_timing(ctr_ctx *ctx, int enable);
ctr_status ctr_kernel_timing(ctr_ctx *ctx, double *fp_ms, double *bp_ms, int *fp_launches,
                             int *bp_launches);

/* ---- differentiable alignment model (no reference counterpart) ---------- *
 * Forward model: the OSEM1 projector (OSEM.cl:56-62) over all pa angles with
 * the detector coordinate of projection a moved by shifts[a] + center:
 *   xi = (X - px/2) - shifts[a] - center.  shifts may be NULL (all zero). */
ctr_status ctr_align_forward(ctr_ctx *ctx, const float *img, const float *shifts,
                             float center, float *prj, int dN);
/* ctr_align_forward that also writes dprj[z][a][X], the derivative of
 * prj[z][a][X] with respect to shifts[a] (and to center), from the same
 * samples: sum over k of -cos(a) d/du + sin(a) d/dv of the bilinear sample.
 * The shift gradient of a loss with incoming gradient g is then
 * sum over z and X of g[z][a][X] * dprj[z][a][X] (what ctr_align_grad_shift
 * gives by sampling the image again).  dprj has prj's shape. */
ctr_status ctr_align_forward_deriv(ctr_ctx *ctx, const float *img, const float *shifts,
                                   float center, float *prj, float *dprj, int dN);
/* Adjoint of ctr_align_forward with respect to the image: the transpose of
 * the forward's bilinear sampling weights (1 - fract, fract) of each sample's
 * coordinates, exact except in the second image column / row, where one
 * weight can differ by at most 2^-25.  A deterministic per-pixel gather. */
ctr_status ctr_align_adjoint(ctr_ctx *ctx, const float *prj, const float *shifts,
                             float center, float *img, int dN);
/* gshift[a] = d/d shifts[a] of sum(gprj * forward(img)); summed over the dN
 * layers (shifts are shared).  d/d center = sum_a gshift[a]. */
ctr_status ctr_align_grad_shift(ctr_ctx *ctx, const float *img, const float *gprj,
                                const float *shifts, float center, float *gshift, int dN);

#ifdef __cplusplus
}
#endif
#endif
