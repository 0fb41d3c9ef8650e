"""The backprojector stages, per angle, the bins between the extremes its
block's four (clamped) corner pixels map to, and gathers from that LDS
segment of at most 96 entries (68 with one row per thread).  This restates
the kernel's float32 bin arithmetic in numpy (no fused operations, the same
order) and checks, for sampled blocks and angles of several geometries and
row counts, that every thread's bin lies in [bin(min corner), bin(max
corner)] and that the segment fits."""
import numpy as np
import pytest

f32 = np.float32
SEG = {1: 68, 8: 96, 16: 96}      # entries per angle of the LDS segment, by rows per thread


def _half(i, n, int_half):
    i = np.asarray(i).astype(f32)
    return (i - f32(n // 2)) if int_half else (i - f32(n * 0.5))


@pytest.mark.parametrize("nx,ny,px", [(1024, 1024, 1024), (37, 37, 41), (80, 48, 64), (100, 7, 3),
                                      (3900, 40, 3900), (2048, 2048, 2048)])
def test_block_segments_hold_every_bin(nx, ny, px):
    rng = np.random.default_rng(nx + ny)
    th = np.concatenate([np.linspace(0, np.pi, 181), rng.uniform(0, 2 * np.pi, 60), np.arange(8) * np.pi / 4])
    C = np.cos(th).astype(f32)[:, None, None]
    S = np.sin(th).astype(f32)[:, None, None]
    C[np.abs(C) < 1e-7] = 0
    S[np.abs(S) < 1e-7] = 0
    widest = {R: 0 for R in SEG}
    for R in SEG:
        tiles = [(tx, ty) for tx in range((nx + 63) // 64) for ty in range((ny + 4 * R - 1) // (4 * R))]
        if len(tiles) > 12:
            tiles = [tiles[0], tiles[-1]] + [tiles[i] for i in rng.choice(len(tiles), 10, replace=False)]
        for tx, ty in tiles:
            X = np.minimum(tx * 64 + np.arange(64), nx - 1)
            Y = np.minimum(ty * 4 * R + np.arange(4 * R), ny - 1)
            for int_half in (False, True):
                ph = f32(px // 2) if int_half else f32(px * 0.5)
                xr = _half(X, nx, int_half)[None, :, None]
                yr = _half(Y, ny, int_half)[None, None, :]
                bx = (((xr * C).astype(f32) - (yr * S).astype(f32)).astype(f32) + ph).astype(f32)
                ta = (_half(X[0], nx, int_half) * C).astype(f32)
                tb = (_half(X[-1], nx, int_half) * C).astype(f32)
                qa = (_half(Y[0], ny, int_half) * S).astype(f32)
                qb = (_half(Y[-1], ny, int_half) * S).astype(f32)
                corners = [((t - q).astype(f32) + ph).astype(f32) for t in (ta, tb) for q in (qa, qb)]
                cmin = np.minimum.reduce(corners)
                cmax = np.maximum.reduce(corners)
                for add_half in (False, True):
                    def binf(v):
                        return np.floor(v.astype(np.float64) + 0.5) if add_half else np.floor(v)
                    b = binf(bx)
                    lo, hi = binf(cmin), binf(cmax)
                    assert np.all(b >= lo) and np.all(b <= hi), (R, tx, ty, int_half, add_half)
                    widest[R] = max(widest[R], int((hi - lo + 1).max()))
    for R, w in widest.items():
        assert w <= SEG[R], (R, w)
