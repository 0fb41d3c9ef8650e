"""Global-memory bounds of the two LDS-staged kernels, checked on the CPU.

The backprojector (k_bp) reads each subset row at entries roff + bin with no
clamp: the roff zeros on either side of a row must cover every bin any pixel
of the slice maps to, at any angle (its threads past the image's last column
or row take the bins of the nearest image pixel).  A bin past that range once
faulted the GPU, so this restates the kernel's float32 bin arithmetic in
numpy and checks every pixel (the image border and a sample of the interior
for large slices) against the layout the library itself reports
(ctr_query_layout), for the four bin rules and 14 geometries.

The projector (k_fp_lds, and the alignment projector k_align_lds, which uses
the same tiles) loads each live tile's bounding box straight from the padded
frame into an LDS window of 96 columns and 92, 82 or 76 rows for tiles of 60,
44 or 32 samples.  This restates the tile list (TileList in ctr_kernels.hip)
and checks, for each tile size, that every live box fits the window, that
every loaded texel lies inside the padded frame, and that every sample's four
taps lie inside its box.
"""
import numpy as np
import pytest

from ctrecon._lib import query_layout

f32 = np.float32
LT, LS, LDS_W = 64, 60, 96
TILE_ROWS = {60: 92, 44: 82, 32: 76}          # samples per tile -> LDS window rows


def _fma(a, b, c):
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(f32)


def _angles(extra_seed):
    rng = np.random.default_rng(extra_seed)
    deg = np.concatenate([np.arange(0.0, 360.0, 0.5), rng.uniform(0.0, 360.0, 200),
                          np.array([44.99, 45.0, 45.01, 89.999, 90.0, 134.99, 135.0, 180.0])])
    ap = (deg.astype(f32).astype(np.float64) * np.pi / 180.0).astype(f32)
    return np.cos(ap).astype(f32), np.sin(ap).astype(f32)


def _pixels(nx, ny, rng):
    if nx * ny <= 300_000:
        X, Y = np.meshgrid(np.arange(nx), np.arange(ny))
        return X.ravel(), Y.ravel()
    border = [(x, 0) for x in range(nx)] + [(x, ny - 1) for x in range(nx)] + \
             [(0, y) for y in range(ny)] + [(nx - 1, y) for y in range(ny)]
    inner = list(zip(rng.integers(0, nx, 20000), rng.integers(0, ny, 20000)))
    p = np.array(border + inner)
    return p[:, 0], p[:, 1]


BP_GEOMS = [(2, 2, 2), (3, 3, 3), (32, 32, 32), (37, 37, 41), (64, 64, 64), (100, 7, 3), (7, 100, 3),
            (80, 48, 64), (128, 128, 256), (256, 256, 128), (1024, 1024, 1024), (2048, 2048, 2048),
            (3856, 3856, 3856), (3900, 40, 3900)]


@pytest.mark.parametrize("nx,ny,px", BP_GEOMS)
def test_backprojection_bins_stay_in_the_padded_row(nx, ny, px):
    l = query_layout(nx, ny, px)
    C, S = _angles(nx + 7 * ny)
    X, Y = _pixels(nx, ny, np.random.default_rng(px))
    for int_half in (False, True):
        if int_half:
            xr = (X - nx // 2).astype(f32)
            yr = (Y - ny // 2).astype(f32)
            ph = f32(px // 2)
        else:
            xr = (X.astype(f32) - f32(nx * 0.5)).astype(f32)
            yr = (Y.astype(f32) - f32(ny * 0.5)).astype(f32)
            ph = f32(px * 0.5)
        for i0 in range(0, C.size, 64):
            c, s = C[i0:i0 + 64, None], S[i0:i0 + 64, None]
            t1 = (xr[None, :] * c).astype(f32)
            bx = ((t1 - (yr[None, :] * s).astype(f32)).astype(f32) + ph).astype(f32)
            for add_half in (False, True):
                b = np.floor(bx.astype(np.float64) + (0.5 if add_half else 0.0))
                lo, hi = b.min(), b.max()
                assert lo + l.roff >= 0, (int_half, add_half, lo, l.roff)
                assert hi + l.roff < l.rpitch, (int_half, add_half, hi, l.roff, l.rpitch)


def _tiles(nx, ny, px, C, S, geo_power=False, offset=0.0, LS=LS):
    """Every (angle, ray chunk, tile) of the LDS-tiled projector: the box
    origin and extent and whether it is live, as TileList computes them."""
    D = max(nx, ny)
    if geo_power:
        xh, dh = f32(px // 2), f32(D // 2)
        cu, cv = f32(f32(nx // 2) - f32(0.5)), f32(f32(ny // 2) - f32(0.5))
    else:
        xh, dh = f32(px * 0.5), f32(D * 0.5)
        cu, cv = f32(nx * 0.5), f32(ny * 0.5)
    nchunk = (px + LT - 1) // LT
    X0 = np.arange(nchunk) * LT
    Xl = np.minimum(X0 + LT - 1, px - 1)
    c, s = C[:, None, None], S[:, None, None]
    tr = np.abs(s) > np.abs(c)
    cdl, rdl = np.where(tr, c, s), np.where(tr, s, c)

    def base(XX):
        xi = (XX.astype(f32) - xh).astype(f32)
        xi = (xi - f32(offset)).astype(f32)
        bu, bv = _fma(xi[None, :, None], c, cu), _fma(-xi[None, :, None], s, cv)
        return np.where(tr, bv, bu), np.where(tr, bu, bv)

    cbA, rbA = base(X0)
    cbB, rbB = base(Xl)
    kt = np.arange(LT) * LS
    kn = np.minimum(LS, D - kt)
    e0 = (kt.astype(f32) - dh).astype(f32)[None, None, :]
    e1 = ((kt + kn - 1).astype(f32) - dh).astype(f32)[None, None, :]
    cs = [_fma(e, cdl, cb) for e in (e0, e1) for cb in (cbA, cbB)]
    rs = [_fma(e, rdl, rb) for e in (e0, e1) for rb in (rbA, rbB)]
    cmn, cmx = np.minimum.reduce(cs), np.maximum.reduce(cs)
    rmn, rmx = np.minimum.reduce(rs), np.maximum.reduce(rs)
    bx0 = (np.floor(cmn).astype(np.int64) - 1) & ~3
    by0 = np.floor(rmn).astype(np.int64) - 1
    bx1 = np.floor(cmx).astype(np.int64) + 3
    by1 = np.floor(rmx).astype(np.int64) + 3
    W = np.where(tr, ny, nx)
    H = np.where(tr, nx, ny)
    live = (kt < D)[None, None, :] & ~((bx1 <= 0) | (bx0 >= W) | (by1 <= 0) | (by0 >= H))
    return dict(bx0=bx0, by0=by0, bw=bx1 - bx0, bh=by1 - by0, live=live, W=W, H=H, tr=tr,
                cbA=cbA, rbA=rbA, cdl=cdl, rdl=rdl, dh=dh, xh=xh, cu=cu, cv=cv, D=D)


FP_GEOMS = [(2, 2, 2), (3, 3, 3), (100, 7, 3), (7, 100, 3), (32, 32, 32), (37, 37, 41), (64, 64, 64),
            (80, 48, 64), (256, 256, 384), (513, 257, 700), (1024, 1024, 1024), (2048, 2048, 2048),
            (3840, 3840, 3840), (3840, 64, 3840)]


@pytest.mark.parametrize("tile", sorted(TILE_ROWS))
@pytest.mark.parametrize("nx,ny,px", FP_GEOMS)
def test_projector_tile_boxes_fit_window_and_frame(nx, ny, px, tile):
    if max(nx, ny) > 64 * tile:
        pytest.skip("64 tiles of this size do not cover the side (the schedule is refused)")
    LDS_H = TILE_ROWS[tile]
    l = query_layout(nx, ny, px)
    C, S = _angles(nx + 3 * ny + px)
    if nx >= 2048:
        C, S = C[::4], S[::4]
    for geo_power, offset in ((False, 0.0), (True, 0.0), (False, 9.7), (False, -13.25)):
        T = _tiles(nx, ny, px, C, S, geo_power, offset, LS=tile)
        live = T["live"]
        if offset != 0.0 and not live.any():
            continue                      # the moved rays miss a tiny slice entirely
        assert live.any()
        bw, bh, bx0, by0 = T["bw"][live], T["bh"][live], T["bx0"][live], T["by0"][live]
        tr = np.broadcast_to(T["tr"], live.shape)[live]
        H = np.broadcast_to(T["H"], live.shape)[live]
        pitch = np.where(tr, l.tpitch, l.ipitch)
        # the window: rows < bh <= 91 of 92, 16-byte slots below bw <= 96
        assert bw.max() <= LDS_W and bh.max() <= LDS_H - 1, (bw.max(), bh.max())
        # every loaded float lies in the padded frame: rows [0, H + 2 PAD), columns [0, pitch)
        assert (by0 + l.pad).min() >= 0 and (bx0 + l.pad).min() >= 0
        assert (by0 + bh - 1 + l.pad <= H + 2 * l.pad - 1).all()
        last_slot = ((bw + 3) // 4) * 4 - 1
        assert (bx0 + last_slot + l.pad <= pitch - 1).all()


@pytest.mark.parametrize("tile", sorted(TILE_ROWS))
@pytest.mark.parametrize("nx,ny,px", [(37, 37, 41), (80, 48, 64), (256, 256, 384), (1024, 1024, 1024)])
def test_projector_samples_tap_inside_their_box(nx, ny, px, tile):
    """Each sample of a live tile reads taps (ic, ir), (ic + 1, ir + 1) of the
    box, ic = trunc(box-local column): all inside [0, bw) x [0, bh)."""
    LS = tile
    C, S = _angles(nx * 5 + px)
    C, S = C[::6], S[::6]
    T = _tiles(nx, ny, px, C, S, LS=LS)
    D = T["D"]
    rng = np.random.default_rng(nx)
    na, nch, nt = T["live"].shape
    for a, ch, ti in zip(*np.nonzero(T["live"])):
        if rng.random() > 0.05:
            continue
        rays = np.arange(ch * LT, min(ch * LT + LT, px))
        k = np.arange(ti * LS, min(ti * LS + LS, D))
        c, s = C[a], S[a]
        tr = abs(s) > abs(c)
        xi = (rays.astype(f32) - T["xh"]).astype(f32)
        bu, bv = _fma(xi, c, T["cu"]), _fma(-xi, s, T["cv"])
        cb, rb = (bv, bu) if tr else (bu, bv)
        cdl, rdl = (c, s) if tr else (s, c)
        eta = (k.astype(f32) - T["dh"]).astype(f32)
        lc = _fma(eta[None, :], cdl, (cb - f32(T["bx0"][a, ch, ti])).astype(f32)[:, None])
        lr = _fma(eta[None, :], rdl, (rb - f32(T["by0"][a, ch, ti])).astype(f32)[:, None])
        ic, ir = np.trunc(lc).astype(np.int64), np.trunc(lr).astype(np.int64)
        assert ic.min() >= 0 and ir.min() >= 0
        assert ic.max() + 1 < T["bw"][a, ch, ti] and ir.max() + 1 < T["bh"][a, ch, ti]


AJ_T, AJ_W, AD_PAD = 16, 32, 2


def _window_start(du0, dv0, c, s, pxh, sh, center, px):
    """k_align_adj_lds's ceil(Xs - 1.45) clamped to [-AD_PAD, px - 1], in float32
    with the kernel's operation order (no fused operations)."""
    Xs = ((((du0 * c).astype(f32) - (dv0 * s).astype(f32)).astype(f32) + pxh).astype(f32) + sh).astype(f32)
    Xs = (Xs + f32(center)).astype(f32)
    return np.clip(np.ceil((Xs - f32(1.45)).astype(f32)), -AD_PAD, px - 1).astype(np.int64)


@pytest.mark.parametrize("nx,ny,px", [(37, 37, 41), (64, 48, 20), (512, 512, 512), (2048, 2048, 2048), (40, 3900, 3900)])
def test_adjoint_tiles_stage_every_window(nx, ny, px):
    """Each thread of a 16 x 16 tile reads rays xa .. xa + 2 of the AJ_W rays
    its block staged from sbase = clip(min over the tile's corners of xa,
    -AD_PAD, px + AD_PAD - AJ_W): all inside [sbase, sbase + AJ_W) and inside
    the padded row [-AD_PAD, px + AD_PAD)."""
    rng = np.random.default_rng(nx * 3 + px)
    C, S = _angles(px)
    C, S = C[::3], S[::3]
    pxh = f32(px * 0.5)
    cu, cv = f32(nx * 0.5), f32(ny * 0.5)
    tx_n, ty_n = (nx + AJ_T - 1) // AJ_T, (ny + AJ_T - 1) // AJ_T
    tiles = [(tx, ty) for tx in range(tx_n) for ty in range(ty_n)]
    if len(tiles) > 40:
        tiles = [tiles[0], tiles[-1]] + [tiles[k] for k in rng.choice(len(tiles), 38, replace=False)]
    for sh, center in ((0.0, 0.0), (6.3, -2.9), (-7.75, 3.0)):
        for tx, ty in tiles:
            i = np.minimum(tx * AJ_T + np.arange(AJ_T), nx - 1)
            j = np.minimum(ty * AJ_T + np.arange(AJ_T), ny - 1)
            du0 = (i.astype(f32) - cu).astype(f32)[None, :, None]
            dv0 = (j.astype(f32) - cv).astype(f32)[None, None, :]
            c, s = C[:, None, None], S[:, None, None]
            xa = _window_start(du0, dv0, c, s, pxh, f32(sh), center, px)
            corners = [_window_start(du0[:, [0, -1][p]:, :][:, :1], dv0[:, :, [0, -1][q]:][:, :, :1], c, s, pxh,
                                     f32(sh), center, px) for p in (0, 1) for q in (0, 1)]
            lo = np.minimum.reduce(corners)[:, 0, 0]
            sbase = np.maximum(np.minimum(lo, px + AD_PAD - AJ_W), -AD_PAD)
            off = xa - sbase[:, None, None]
            assert off.min() >= 0, (tx, ty, sh)
            assert (off + 2).max() <= AJ_W - 1, (tx, ty, sh, (off + 2).max())
            assert (xa + 2).max() <= px + AD_PAD - 1
