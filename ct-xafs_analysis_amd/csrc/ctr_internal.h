// ctr_internal.h -- structures shared by the gfx950 kernels and the C ABI layer.
//
// Internal data layout in HBM (one "slice" = one reconstruction layer):
//   pimg   zero-padded image, PAD texels of zeros on every side, row pitch
//          ipitch floats, (ny + 2 PAD) rows.  The padding lets the projectors
//          read without bounds tests: a direct bilinear gather clamps its
//          coordinates into the padded frame, where any clamped sample lands
//          on zeros (exactly CLK_ADDRESS_CLAMP's border value), and the
//          96 x 92 window of an LDS tile, whose box meets the interior and so
//          starts at most 94 texels before it, lies entirely inside the frame.
//   pimgT  the same image transposed (pitch tpitch, nx + 2 PAD rows).  Rays
//          steeper than 45 degrees read it so that the 64 lanes of a wave,
//          which hold 64 neighbouring detector bins, touch consecutive
//          addresses.
//   rows   compact subset rows [pa/ss][rpitch]: entry roff + X holds bin X;
//          every other entry stays zero.  roff >= hypot(nx, ny)/2 - px/2 + 2,
//          so every bin a pixel maps to lands on the row or on its zeros and
//          the nearest-bin backprojectors read it with no clamp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctr {

constexpr int PAD = 100;
// Largest max(nx, ny) the LDS-tiled projector takes (one tile of 60 samples
// per lane of a wave); larger slices use the direct-gather projector.
constexpr int LDS_FP_MAX_SIDE = 60 * 64;

struct SliceDev {
    float *img;          // padded image (pitch ipitch, ny + 2 PAD rows)
    float *imgT;         // padded transposed image (pitch tpitch, nx + 2 PAD rows)
    const float *prj;    // sinogram [pa][px], unpadded
    float *rows;         // compact subset rows [pa/ss][rpitch], bin b at entry roff + b
    const float *src;    // unpadded [ny][nx] source of pad_in
    float *dst;          // unpadded [ny][nx] destination of unpad_out
    float *aux;          // unpadded [ny][nx] output plane of to_aux backprojections
    int nx, ny, px;
    int ipitch, tpitch, rpitch;
    int roff;            // zeros before bin 0 (and at least as many after bin px - 1) in a row
    int layer;           // index in its stack (subset index for the power iteration)
    int blk[6];          // first logical block of this slice for each grid kind (the
                         // G_FP map holds (slice, block) pairs instead)
};

enum GridKind { G_FP = 0, G_BP = 1, G_TR = 2, G_PAD = 3, G_EW = 4, G_SPARE = 5 };

// Forward-projection epilogues.
enum FpVar { FP_RATIO = 0, FP_RESID = 1, FP_RAW = 2 };
// Backprojection variants (kernel + epilogue).
enum BpVar { BP_OSEM2 = 0, BP_AART2 = 1, BP_INIT = 2, BP_FISTA = 3, BP_POWER = 4 };

struct Launch {
    const SliceDev *sl;   // device descriptor array
    const int *bmap;      // device logical-block -> slice map for this grid kind
                          // (G_FP: (slice, block within slice) pairs, slice -1 = empty)
    const float2 *cs;     // device [pa] (cos, sin) of the float angle
    int nblocks;
};

struct FpParams {
    int sub, ss, na, sub_from_layer, geo_power;
    int nblk;   // logical blocks of the launch (set by launch_fp; the grid may be 2-D)
};

struct BpParams {
    int sub, ss, na, pa, sub_from_layer;
    float alpha;
    const float *L;       // L2norm (FISTA backprojection) or null
    int to_aux;           // write the result to aux instead of the padded image
    int nblk;             // logical blocks of the launch (set by launch_bp)
    int write_t;          // also write the transposed padded copy (the next projection reads it)
};

// ---- launchers (ctr_kernels.hip) ----
void launch_fp(int var, int wpb, int rays, int tile, const Launch &L, const FpParams &p, hipStream_t stream);
void launch_bp(int var, int rows, const Launch &L, const BpParams &p, hipStream_t stream);
void launch_transpose(const Launch &L, hipStream_t stream);
void launch_pad_in(const Launch &L, hipStream_t stream);
void launch_unpad_out(const Launch &L, hipStream_t stream);
void launch_fill(float *dst, size_t n, float value, hipStream_t stream);
void launch_copy(const float *src, float *dst, size_t n, hipStream_t stream);
void launch_sino_correction(const float *src, float *dst, const float2 *cs_corr, int nz, int pa,
                            int px, int mode, float a, float b, hipStream_t stream);
void launch_tv(int fista, const float *x, const float *v, const float *b, float *w_new, float *x_new,
               float *b_new, int out_padded, int ipitch, int nz, int nx, int ny, const float *L,
               int sub, double lambda_fista, hipStream_t stream);
void launch_tv2d(const float *v, float *x_new, int nz, int nx, int ny, float lambda, float lower,
                 hipStream_t stream);
void launch_l2(const float *src, int nz, size_t n, float *partial, float *l2, hipStream_t stream);
void launch_layer_min(const float *src, int nz, size_t n, float *partial, float *mins, hipStream_t stream);
void launch_baseup(float *img, const float *mins, int nz, size_t plane, int order, hipStream_t stream);
void launch_rot_shift(const float *prj, float *out, int pa, int px, int nshift, const float *shifts,
                      hipStream_t stream);
void launch_merge_frames(const unsigned short *raw, float *out, int n, size_t m, size_t stride, hipStream_t stream);
void launch_mt_conv(const float *dark, const float *I0, const unsigned short *It, float *mt, int n, size_t m,
                    size_t stride, hipStream_t stream);
void launch_reslice(const float *mt, float *prj, int A, int H, int W, const float *xs, const float *ys,
                    float baseup, int correction, int y0, int nyo, hipStream_t stream);
void launch_denoise(const float *src, float *dst, int nz, int A, int W, float thr, hipStream_t stream);
void launch_xproj(const float *mt, float *xproj, int A, int H, int W, int sx, int ex, hipStream_t stream);
// Largest x-projection length the shift search keeps in LDS (H + 256 floats).
constexpr int ZCORR_MAX_H = 8192;
void launch_zcorr(const float *xproj, float *yshift, int A, int H, int sY, int eY, hipStream_t stream);
// cos(a) and cos(a - pi/2) of the two ends of the angle range (setMask's
// cos(a_min), cos(a_max), rotCenterSearch.cl:59-66).
struct RcsAngles {
    float cmin, smin, cmax, smax;
};
void launch_focus_index(const float *img, int nz, int nx, int ny, const float *shifts, RcsAngles ends,
                        float2 *partial, float *findex, hipStream_t stream);
void launch_power3(const float *cnd, const float *l2, float *pimg, int nz, int nx, int ny, int ipitch,
                   hipStream_t stream);
void launch_fbp_filter(const float *prj, float *frows, int nz, int pa, int px, int rpitch, int roff, int zp,
                       const float2 *tw, hipStream_t stream);
void launch_fbp_bp(const float *frows, float *img, const float2 *cs, int nz, int nx, int ny, int pa,
                   int px, int rpitch, int roff, hipStream_t stream);
void launch_rows_to_full(const float *rows, float *full, int nz, int na, int rpitch, int roff, int pa, int px,
                         int sub, int ss, hipStream_t stream);
void launch_full_to_rows(const float *full, float *rows, int nz, int na, int rpitch, int roff, int pa, int px,
                         int sub, int ss, hipStream_t stream);
// dprj, when not null, also receives each ray's derivative with respect to
// its projection's offset.
void launch_align_forward(const float *pimg, const float *pimgT, int ipitch, int tpitch, const float2 *cs,
                          const float *shifts, float center, float *prj, float *dprj, int nz, int nx, int ny, int pa,
                          int px, hipStream_t stream);
// Zero rays on either side of each projection row in the adjoint's copy of
// the rows, and that copy's size in floats (slice-fastest, in chunks of 1, 4
// or 16 slices).
constexpr int AD_PAD = 2;
size_t align_adj_floats(int nz, int pa, int px);
void launch_align_adjoint(const float *prj, float *gpad, const float2 *cs, const float *shifts, float center,
                          float *img, int nz, int nx, int ny, int pa, int px, hipStream_t stream);
void launch_align_grad(const float *pimg, const float *pimgT, int ipitch, int tpitch, const float2 *cs,
                       const float *gprj, const float *shifts, float center, float *partial, float *gshift,
                       int nz, int nx, int ny, int pa, int px, hipStream_t stream);

// blocks per slice for each grid kind
int fp_blocks(int px, int na, int wpb, int rays);
int bp_blocks(int nx, int ny, int rows);
int tr_blocks(int nx, int ny);
int pad_blocks(int ny, int ipitch);
int ew_blocks(int nx, int ny);
size_t l2_partial_count(int nz);

}  // namespace ctr
