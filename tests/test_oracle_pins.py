"""Pins of the CPU oracle (oracle/ct_oracle*.c) without running the reference.

The reference OpenCL kernels cannot execute here (no OpenCL CPU device), and the
reference ships no tests or golden data, so each oracle function is pinned by
something independent of its own formula:

  subset order            literal lists (GPU.cpp:564-582 worked by hand)
  forward projector       theta=0 closed form (column sums), translation
                          equivariance, 90-degree rotation symmetry, FP32 vs FP64
  powerIter1              theta=0 / 90 closed forms of its half-texel
                          geometry (no +0.5): (C[X-1] + C[X]) / 2 at theta=0
  nearest backprojectors  single-angle probes on odd sizes: rounding vs
  (OSEM2, AART2,          truncation, integer vs float halves, and each
  backProjection,         update's closed form (x q, x + q a S/(A D),
  FISTAbackProjection,    q S/(A D), x + q a S/(A D)/L 0.5, q S/A/D); update
  powerIter2)             scales over a two-angle subset with S != A
  FISTA3D momentum        literal beta', gamma and w for beta = 0, 1, 3, 10
  backProjectionFBP       an impulse lands on exactly the line the +0.5
                          offsets predict (theta = 0 and 90 degrees)
  per-slice TV prox       brute-force L1-prox minimisation with four
                          neighbours; slices independent
  alignment adjoint       dot-product test against the independently written
                          scatter adjoint (FP64)
  shift gradient          FP64 central finite differences of the forward model
  nearest backprojection  theta=0 identity, normalisation fixed point of OSEM2
  NaN / clamp rules       crafted inputs (OSEM.cl:65-67, 101; AART.cl:69;
                          reconstShare.cl:268; FISTA.cl:488-489)
  FBP filter              brute-force O(ZP^2) inverse DFT of the ramp in numpy
  FBP amplitude           uniform disk reconstructs to mu/2 (FBP.cl:186)
  TV prox                 brute-force minimisation of 1/2(x-v)^2 + lam sum|x-n_i|
  power iteration         dense eigen-decomposition of the 16^2 system
  driver quirks           FISTA_execution returns x1; AART init replaces the
                          image; FISTA_OSEM discards its first OS-EM update
  phantom                 analytic line integral of a centred disk
  centre-shift resampling integer shifts = index shifts with zero fill, half
                          shifts = neighbour means, np.interp elsewhere
  centre-search mask      full 180-degree range = integer disk of radius
                          nx/2 - |shift|; the 0-90 degree case worked by hand
  focus index             constant image = 0, checkerboard closed form over the
                          eroded integer disk
  centre search           composition of its stages; the minimum of the focus
                          index lies at the detector shift the data were made with
  frame merge / mu-t      integer means; ln((I0 - dark) / (It - dark)) and its
                          clamps at the three limits
  reslice                 zero shifts = transpose (bit-exact), integer shifts =
                          index shifts, fractional shifts = scipy's bilinear
                          map_coordinates, cos(asin t) = sqrt(1 - t^2)
  spike removal           crafted spikes replaced by the theta-neighbour mean,
                          ramps and edge rows kept
  x-projection            constant image: (ex - sx + 1) c inside, half at row 0
  z-correction            equal profiles stay at 0 exactly; on ramp profiles the
                          20 trials end at -2x the drift (two accepted
                          Gauss-Newton steps, every later step raises the misfit)
"""
import numpy as np
import pytest

import oracle as O
from _util import relerr

F32, F64 = np.float32, np.float64


def test_subset_order_literals():
    assert O.subset_order(16).tolist() == [0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15]
    assert O.subset_order(8).tolist() == [0, 4, 2, 6, 1, 5, 3, 7]
    assert O.subset_order(10).tolist() == list(range(10))
    assert O.subset_order(1).tolist() == [0]
    assert O.subset_order(6).tolist() == list(range(6))


def _proj(g, ang, img, dtype):
    """A x through AART1 with y = 0 (residual = -A x)."""
    z = np.zeros((img.shape[0], g.pa, g.px), dtype=dtype)
    return -O.aart1(g, ang, img.astype(dtype), z, 0, dtype=dtype)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_projector_theta0_column_sums(dtype):
    n = 32
    g = O.geom(n, n, 1, ss=1)
    rng = np.random.default_rng(0)
    img = rng.random((1, n, n))
    p = _proj(g, np.array([0.0], np.float32), img, dtype)
    # at theta = 0 with P = N = D the samples hit texel centres: p[X] = sum_y img[y][X]
    assert np.allclose(p[0, 0], img[0].sum(axis=0), rtol=1e-6, atol=1e-5)


def test_projector_translation_and_rotation():
    n = 32
    g = O.geom(n, n, 2, ss=1)
    rng = np.random.default_rng(1)
    img = np.zeros((1, n, n))
    img[0, 8:24, 8:24] = rng.random((16, 16))
    ang = np.array([0.0, 90.0], np.float32)
    p = _proj(g, ang, img, F64)
    # a one-pixel shift in x moves the theta=0 projection one bin
    p_s = _proj(g, ang, np.roll(img, 1, axis=2), F64)
    assert np.allclose(p_s[0, 0, 1:], p[0, 0, :-1], atol=1e-9)
    # theta = 90 projects along x onto the reversed y axis: s = -y
    rows = img[0].sum(axis=1)
    assert np.allclose(p[0, 1, 1:], rows[::-1][:-1], atol=1e-6)


def test_projector_fp32_vs_fp64_tie_free():
    n, pa = 48, 30
    g = O.geom(n, n, pa, ss=1)
    ang = (np.arange(pa) * 6.0 + 0.731).astype(np.float32)
    img = np.random.default_rng(2).random((1, n, n))
    assert relerr(_proj(g, ang, img, F32), _proj(g, ang, img, F64)) < 2e-6


def test_align_adjoint_dot_product_fp64():
    n, pa = 24, 17
    g = O.geom(n, n, pa, ss=1)
    rng = np.random.default_rng(3)
    ang = rng.uniform(0, 180, pa).astype(np.float32)
    sh = rng.normal(0, 1.5, pa)
    for trial in range(3):
        x = rng.normal(size=(1, n, n))
        y = rng.normal(size=(1, pa, n))
        ax = O.align_forward(g, ang, x, sh, 0.7)
        aty = O.align_adjoint(g, ang, y, sh, 0.7)
        lhs, rhs = np.vdot(ax, y), np.vdot(x, aty)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(ax) * np.linalg.norm(y)


def test_align_forward_equals_osem_projector():
    n, pa = 32, 12
    g = O.geom(n, n, pa, ss=1)
    ang = (np.arange(pa) * 15.0 + 0.3).astype(np.float32)
    img = np.random.default_rng(4).random((1, n, n))
    assert relerr(O.align_forward(g, ang, img, None, 0.0), _proj(g, ang, img, F64)) < 1e-12


def test_align_shift_gradient_fd_fp64():
    n, pa = 24, 9
    g = O.geom(n, n, pa, ss=1)
    rng = np.random.default_rng(5)
    ang = rng.uniform(0, 180, pa).astype(np.float32)
    from ctrecon import phantom as ph
    img = ph.rasterize(ph.random_ellipses(4, rng), n, supersample=2)[None]
    sh = rng.normal(0, 1.0, pa)
    gp = rng.normal(size=(1, pa, n))
    grad = O.align_grad(g, ang, img, gp, sh, 0.4)
    h = 1e-6
    fd = np.zeros(pa)
    for a in range(pa):
        sp, sm = sh.copy(), sh.copy()
        sp[a] += h
        sm[a] -= h
        fd[a] = (np.vdot(gp, O.align_forward(g, ang, img, sp, 0.4))
                 - np.vdot(gp, O.align_forward(g, ang, img, sm, 0.4))) / (2 * h)
    assert relerr(grad, fd) < 1e-4


@pytest.mark.parametrize("dtype", [F32, F64])
def test_backprojection_theta0_identity(dtype):
    n = 16
    g = O.geom(n, n, 1, ss=1)
    rows = np.random.default_rng(6).random((1, 1, n))
    img = np.ones((1, n, n))
    # OSEM2 at a single angle theta = 0: x' = x * r[bin(X)] * SS/A with bin(X) = X
    out = O.osem2(g, np.array([0.0], np.float32), img, rows, 0, dtype=dtype)
    assert np.allclose(out[0], np.broadcast_to(rows[0, 0], (n, n)), rtol=1e-6)


def test_osem2_normalisation_fixed_point():
    """With all ratios 1, OSEM2 multiplies x by (#in-range bins) * SS / A, which is
    exactly 1 for pixels whose every subset bin falls on the detector."""
    n, pa, ss = 32, 60, 6
    g = O.geom(n, n, pa, ss=ss)
    ang = (np.arange(pa) * 3.0).astype(np.float32)
    x = np.random.default_rng(7).uniform(0.5, 1.5, (1, n, n))
    out = O.osem2(g, ang, x, np.ones((1, pa, n)), 2, dtype=F64)
    yy, xx = np.mgrid[0:n, 0:n]
    inside = (xx - n / 2) ** 2 + (yy - n / 2) ** 2 < (n / 2 - 1.5) ** 2
    assert np.allclose(out[0][inside], x[0][inside], rtol=1e-12)


def test_nan_and_clamp_rules():
    n, pa = 8, 4
    g = O.geom(n, n, pa, ss=1)
    ang = np.array([0, 45, 90, 135], np.float32)
    img = np.zeros((1, n, n), np.float32)          # projection p = 0 < 1e-4
    prj = np.ones((1, pa, n), np.float32)
    prj[0, 0, 0] = np.nan
    prj[0, 1, 1] = -5.0
    r = O.osem1(g, ang, img, prj, 0)
    assert r[0, 0, 0] == np.float32(1e-6)            # NaN -> 1e-6
    assert r[0, 1, 1] == np.float32(1e-6)            # max(1e-6, y)
    assert r[0, 2, 2] == 1.0                          # p < 1e-4: ratio = y
    d = O.aart1(g, ang, img, prj, 0)
    assert d[0, 0, 0] == 0.0                          # NaN y -> 0
    g1 = O.geom(n, n, 1, ss=1)
    bp = O.backprojection(g1, np.array([0.0], np.float32), prj[:, :1], 0)
    assert np.isfinite(bp).all() and (bp[0, :, 0] == 0.0).all()     # NaN bin -> 0
    assert np.allclose(bp[0, :, 1], 1.0 / n)                           # SS / A / D
    # OSEM2: NaN product -> 1e-6
    x = np.full((1, n, n), np.nan, np.float32)
    out = O.osem2(g1, ang[:1], x, np.ones((1, 1, n), np.float32), 0)
    assert (out == np.float32(1e-6)).all()
    # ISTA3D clamps to >= 1e-6
    L = np.ones(1, np.float32)
    v = np.full((1, n, n), -3.0, np.float32)
    assert (O.ista3d(g, v, L, 0, 1e-3) == np.float32(1e-6)).all()


@pytest.mark.parametrize("px,zp", [(16, 32), (24, 64), (64, 128)])
def test_fbp_filter_closed_form(px, zp):
    """Filtered row = circular convolution with h = IDFT of H[k] = (pi/ZP) min(k, ZP-k),
    the row placed at index (X - P/2) mod ZP (FBP.cl:44-57, 125-160)."""
    g = O.geom(px, px, 3, ss=1, zp=zp)
    rng = np.random.default_rng(8)
    prj = rng.normal(size=(1, 3, px))
    k = np.arange(zp)
    H = (np.pi / zp) * np.minimum(k, zp - k)
    nn = np.arange(zp)
    h = (np.exp(2j * np.pi * np.outer(nn, k) / zp) @ H).real / zp      # brute-force IDFT
    idx = (np.arange(px) - px // 2) % zp
    exp = np.zeros_like(prj)
    for t in range(3):
        row = np.zeros(zp)
        row[idx] = prj[0, t]
        conv = np.array([np.sum(row * h[(m - nn) % zp]) for m in range(zp)])
        exp[0, t] = conv[idx]
    assert relerr(O.fbp_filter(g, prj, dtype=F64), exp) < 1e-12
    assert relerr(O.fbp_filter(g, prj.astype(F32)), exp) < 1e-5


def test_fbp_disk_amplitude_is_half_mu():
    n, pa, mu, rad = 128, 360, 0.5, 40.0
    g = O.geom(n, n, pa, ss=1, zp=256)
    ang = (np.arange(pa) * 0.5).astype(np.float32)
    s = np.arange(n) - n / 2
    row = np.where(np.abs(s) < rad, 2 * mu * np.sqrt(np.maximum(rad ** 2 - s ** 2, 0)), 0.0)
    prj = np.repeat(row[None, None], pa, axis=1)
    img = O.fbp_execution(g, ang, prj, dtype=F64)[0]
    yy, xx = np.mgrid[0:n, 0:n]
    r = np.hypot(xx - n / 2, yy - n / 2)
    centre = img[r < rad - 8].mean()
    assert abs(centre - mu / 2) < 0.03 * mu, centre
    # outside the disk only the small negative offset of a band-limited ramp remains
    assert abs(img[(r > rad + 8) & (r < n / 2 - 4)].mean()) < 0.05 * mu


def test_tv_prox_is_the_l1_prox():
    """ISTA3D (nz = 1: both z-neighbours are the pixel itself) returns
    argmin_x 1/2 (x - v)^2 + lam * sum_i |x - n_i| over its six neighbours."""
    rng = np.random.default_rng(9)
    n = 6
    g = O.geom(n, n, 1, ss=1)
    v = rng.uniform(0.2, 1.0, (1, n, n))
    L = np.array([0.8])
    lam_f = 0.05
    lam = lam_f / L[0] * 0.5
    out = O.ista3d(g, v, L, 0, lam_f, dtype=F64)
    grid = np.linspace(-0.5, 1.6, 210001)
    for (yy, xx) in [(2, 3), (1, 1), (4, 2), (0, 0), (5, 5)]:
        nb = [v[0, yy, xx + 1] if xx + 1 < n else 0.0, v[0, yy, xx - 1] if xx >= 1 else 0.0,
              v[0, yy + 1, xx] if yy + 1 < n else 0.0, v[0, yy - 1, xx] if yy >= 1 else 0.0,
              v[0, yy, xx], v[0, yy, xx]]
        f = 0.5 * (grid - v[0, yy, xx]) ** 2 + lam * sum(np.abs(grid - t) for t in nb)
        assert abs(grid[np.argmin(f)] - max(out[0, yy, xx], 1e-6)) < 2e-5


def test_power_iteration_dense_eigenvalue():
    n, pa, ss = 16, 8, 2
    g = O.geom(n, n, pa, ss=ss)
    ang = (np.arange(pa) * 22.5 + 0.4).astype(np.float32)
    L = O.power_iteration(g, ang, 30, dtype=F64)
    for s in range(ss):
        M = np.zeros((n * n, n * n))
        for i in range(n * n):
            e = np.zeros((ss, n, n))
            e[s].flat[i] = 1.0
            M[:, i] = O.power_iter2(g, ang, O.power_iter1(g, ang, e, dtype=F64), dtype=F64)[s].ravel()
        lam = np.max(np.abs(np.linalg.eigvals(M)))
        assert abs(L[s] - lam ** 2) / lam ** 2 < 1e-3, (s, L[s], lam ** 2)


def _small_case(seed=10, n=24, pa=12, ss=4):
    from ctrecon import phantom as ph
    img, sino, ang = ph.make_slice(n, pa, 5, seed)
    return O.geom(n, n, pa, ss=ss), ang, sino[None]


def test_fista_execution_returns_first_step():
    g, ang, y = _small_case()
    L = O.power_iteration(g, ang, 30)
    x0 = np.ones((1, g.ny, g.nx), np.float32)
    a = O.fista_execution(g, ang, L, x0, y, 1, 1.0, 1e-3)
    b = O.fista_execution(g, ang, L, x0, y, 4, 1.0, 1e-3)
    assert np.array_equal(a, b)
    c = O.fista_execution(g, ang, L, x0, y, 0, 1.0, 1e-3)
    assert np.array_equal(c, O.backprojection(g, ang, y, int(O.subset_order(g.ss)[0])))


def test_aart_init_replaces_image():
    g, ang, y = _small_case()
    a = O.aart_execution(g, ang, np.ones((1, g.ny, g.nx)), y, 1, 1.0)
    b = O.aart_execution(g, ang, np.full((1, g.ny, g.nx), 9.0), y, 1, 1.0)
    assert np.array_equal(a, b)


def test_fista_osem_first_step_discards_osem_update():
    g, ang, y = _small_case()
    x0 = np.random.default_rng(11).uniform(0.5, 1.0, (1, g.ny, g.nx)).astype(np.float32)
    one = O.fista_osem_execution(g, ang, x0, y, 1, 1e-3, sub=np.array([0]) if False else None)
    # a single subset step: the OS-EM update is overwritten by ISTA3D(x0) with L = 0.5
    g1 = O.geom(g.nx, g.px, g.pa, ss=1)
    ista = O.ista3d(g1, x0, np.array([0.5], np.float32), 0, 1e-3)
    g_one = O.geom(g.nx, g.px, g.pa, ss=1)
    out = O.fista_osem_execution(g_one, ang, x0, y, 1, 1e-3)
    assert np.array_equal(out, ista)
    assert one.shape == x0.shape


def test_phantom_disk_line_integral():
    from ctrecon import phantom as ph
    n = 64
    ell = ph.Ellipses(np.array([0.0]), np.array([0.0]), np.array([0.5]), np.array([0.5]), np.array([0.3]),
                      np.array([0.8]))
    ang = np.array([0.0, 33.0, 90.0], np.float32)
    p = ph.line_integrals(ell, n, ang, n)
    s = (np.arange(n) - n / 2) / (n / 2)
    exp = np.where(np.abs(s) < 0.5, 2 * 0.8 * np.sqrt(np.maximum(0.25 - s ** 2, 0)), 0)
    assert np.allclose(p, exp[None], atol=1e-7)   # r^2 = a^2 (cos^2 + sin^2) rounds at the rim


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_sinogram_correction_closed_forms(mode):
    """reconstShare.cl:35-70 with a = b = 0.5: detector weight sqrt(1 - t^2)
    (= cos asin t), angle weight cos(theta), both, or the mask
    t^2 < (a + b cos theta)^2, with t = |2X - P| / P."""
    px, pa = 20, 12
    g = O.geom(16, px, pa)
    ang = (np.arange(pa) * 15.0).astype(np.float32)
    y = np.random.default_rng(11).uniform(0.5, 2.0, (2, pa, px))
    got = O.sinogram_correction(g, ang, y, mode, dtype=F64)
    t = np.abs(2.0 * np.arange(px) - px) / px
    th = np.deg2rad(ang.astype(np.float64))[:, None]
    wx, wt = np.sqrt(1.0 - t * t)[None, :], np.cos(th)
    want = {0: y, 1: y * wx, 2: y * wt, 3: y * wx * wt,
            4: np.where((t * t)[None, :] < (0.5 + 0.5 * wt) ** 2, y, 0.0)}[mode]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_layer_min_and_baseup():
    """findMinimumX/Y = the NaN-skipping minimum of each layer; baseUp maps each
    layer's minimum to exactly 0 for any order, and with a long series
    (ln(1 + b) for |b| < 1) maps every pixel near the minimum to ~0, since
    a * exp(ln(1 + b)) = a + (x - a) = x."""
    rng = np.random.default_rng(3)
    g = O.geom(16, 16, 4)
    img = rng.uniform(-0.2, 1.0, (3, 16, 16))
    img[1, 2, 3] = np.nan
    mins = O.layer_min(g, img, dtype=F64)
    assert np.array_equal(mins, np.fmin.reduce(img.reshape(3, -1), axis=1))
    for order in (1, 4, 9):
        out = O.baseup(g, img, mins, order, dtype=F64)
        for z in range(3):
            k = np.nanargmin(img[z])
            assert out[z].ravel()[k] == 0.0
    near = np.full((1, 16, 16), -0.5) + rng.uniform(0.0, 0.2, (1, 16, 16))   # |b| = |x - a| / |a| < 0.5
    m = O.layer_min(g, near, dtype=F64)
    out = O.baseup(g, near, m, 80, dtype=F64)
    assert np.abs(out).max() < 1e-12


def test_hybrid_is_fbp_baseup_then_osem():
    """hybrid_thread (GPU.cpp:1459-1580): the correction is applied once, to
    the FBP; OS-EM then starts from the based-up FBP image."""
    from ctrecon import phantom as ph
    img, sino, ang = ph.make_slice(32, 24, 5, 7)
    g = O.geom(32, 32, 24, ss=4, zp=64)
    y = sino[None].astype(np.float64)
    got = O.hybrid_execution(g, ang, y, 2, correction_mode=2, dtype=F64)
    yc = y.copy()
    x = O.fbp_execution(g, ang, yc, correction_mode=2, dtype=F64)
    x = O.baseup(g, x, O.layer_min(g, x, dtype=F64), 4, dtype=F64)
    want = O.osem_execution(g, ang, x, O.sinogram_correction(g, ang, y, 2, dtype=F64), 2, prj_correction=False,
                            dtype=F64)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
    assert (got >= 0).all()


# ---- rotation-centre search (RotationCenterSearch/) --------------------------

def _zero_fill_shift(row, k):
    """row moved left by k samples (out[X] = row[X + k]), zeros past the ends."""
    out = np.zeros_like(row)
    px = row.shape[-1]
    lo, hi = max(0, -k), min(px, px - k)
    out[..., lo:hi] = row[..., lo + k:hi + k]
    return out


def test_rot_center_shift_integer_half_and_fractional():
    rng = np.random.default_rng(4)
    pa, px = 7, 19
    g = O.geom(16, px, pa)
    prj = rng.uniform(-1.0, 2.0, (pa, px)).astype(np.float32)
    out = O.rot_center_shift(g, prj, 7, -3.0, 1.0)
    for i in range(7):
        assert np.array_equal(out[i], _zero_fill_shift(prj, i - 3)), i
    half = O.rot_center_shift(g, prj, 1, 0.5, 0.0)[0]
    assert np.array_equal(half, np.float32(0.5) * prj + np.float32(0.5) * _zero_fill_shift(prj, 1))
    xs = np.arange(-1, px + 1, dtype=np.float64)
    for s in (-2.3, 0.77, 4.1):
        got = O.rot_center_shift(g, prj, 1, s, 0.0, dtype=F64)[0]
        for th in range(pa):
            want = np.interp(np.arange(px) + float(np.float32(s)), xs, np.concatenate([[0.0], prj[th], [0.0]]), left=0.0, right=0.0)
            assert np.allclose(got[th], want, rtol=0, atol=1e-12), (s, th)


@pytest.mark.parametrize("n,shift", [(40, 0.0), (40, 3.5), (40, -3.5), (33, -2.5)])
def test_rcs_mask_full_range_is_disk(n, shift):
    """Over 0-180 degrees the four half-planes reduce to |x| <= nx/2 - |shift|
    and the two arcs to one circle of that radius."""
    m = O.rcs_mask(n, n, 1, shift, 0.0, 0.0, 180.0)[0]
    c = n // 2
    y, x = np.mgrid[0:n, 0:n] - c
    r = n // 2 - abs(shift)
    assert np.array_equal(m.astype(bool), 4 * (x * x + y * y) <= int(round(4 * r * r)))


def test_rcs_mask_quarter_range_by_hand():
    """0-90 degrees, shift +2.5 on a 32 grid: rp = 16, rm = -13.5, the split
    line p_div = rm cos(pi/2) is just right of x = 0; x >= 1 keeps the circle of
    radius 16 cut by y >= rm, x <= 0 the circle of radius 13.5."""
    n = 32
    m = O.rcs_mask(n, n, 1, 2.5, 0.0, 0.0, 90.0)[0].astype(bool)
    y, x = np.mgrid[0:n, 0:n] - n // 2
    r2 = x * x + y * y
    want = ((x >= 1) & (r2 <= 256) & (y >= -13)) | ((x <= 0) & (4 * r2 <= 729))
    assert np.array_equal(m, want)


def _eroded_disk(n, r2x4):
    """Pixels whose 3 x 3 block lies in the integer disk 4(x^2+y^2) <= r2x4 and the grid."""
    y, x = np.mgrid[0:n, 0:n] - n // 2
    inside = 4 * (x * x + y * y) <= r2x4
    ok = np.zeros_like(inside)
    ok[1:-1, 1:-1] = True
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            ok[1:-1, 1:-1] &= inside[1 + dy:n - 1 + dy, 1 + dx:n - 1 + dx]
    return ok


@pytest.mark.parametrize("dtype", [F32, F64])
def test_focus_index_closed_forms(dtype):
    n = 24
    g = O.geom(n, n, 4)
    const = np.full((1, n, n), 0.7)
    assert O.focus_index(g, const, 0.0, 0.0, 0.0, 180.0, dtype=dtype)[0] == 0.0
    # checkerboard 1 +- d: 4 edge neighbours differ by 2d, 4 diagonals by 0
    d = 0.25
    y, x = np.mgrid[0:n, 0:n]
    sign = np.where((x + y) % 2 == 0, 1.0, -1.0)
    img = (1.0 + d * sign)[None]
    for shift in (0.0, 2.5):
        ok = _eroded_disk(n, int(round(4 * (n // 2 - shift) ** 2)))
        a = 1.0 + d * sign[ok]
        want = np.sqrt(np.mean(16 * d * d / (a * a)))
        got = O.focus_index(g, img, shift, 0.0, 0.0, 180.0, dtype=dtype)[0]
        assert abs(got - want) <= (1e-6 if dtype == F32 else 1e-14) * want, (shift, got, want)


def _filled_phantom(n, seed):
    """0.5 over the whole grid plus random ellipses inside the central half."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:n] - n / 2 + 0.5
    img = np.full((n, n), 0.5)
    for _ in range(6):
        cx, cy = rng.uniform(-0.25 * n, 0.25 * n, 2)
        a, b = rng.uniform(0.03 * n, 0.12 * n, 2)
        img += np.where(((x - cx) / a) ** 2 + ((y - cy) / b) ** 2 < 1, rng.uniform(0.2, 1.0), 0.0)
    return img


def test_rotation_center_search_composition_and_minimum():
    """The search = shifted candidates, constant start, OS-EM with the
    correction, focus index.  On data made with the rotation axis s0 bins off
    the detector centre (a detector wide enough to lose nothing), the focus
    index over one common mask is smallest exactly at s0; with the per-candidate
    masks of setMask (radius nx/2 - |shift| over 180 degrees), which admit more
    smooth background for smaller |shift|, the minimum the reference picks
    (rotationCenterSearch_ocl.cpp:521-549) lies within one bin of s0."""
    n, px, pa, ss = 48, 80, 72, 8
    ang = (np.arange(pa) * 180.0 / pa).astype(np.float32)
    g = O.geom(n, px, pa, ss=ss)
    sino = O.align_forward(g, ang, _filled_phantom(n, 1)[None])[0]
    for s0 in (2, -3):
        y = _zero_fill_shift(sino, -s0).astype(np.float32)     # axis at px/2 + s0
        rec, f = O.rotation_center_search(g, ang, y, 11, -5.0, 1.0, it=1, init=1.0, correction_mode=1)
        cand = O.rot_center_shift(g, y, 11, -5.0, 1.0)
        x = O.osem_execution(g, ang, np.ones((11, n, n), np.float32), cand, 1, correction_mode=1)
        assert np.array_equal(rec, x)
        assert np.array_equal(f, O.focus_index(g, x, -5.0, 1.0, 0.0, float(ang[-1])))
        assert abs(int(np.argmin(f)) - 5 - s0) <= 1, (s0, f)
        common = O.focus_index(g, x, 0.0, 0.0, 0.0, float(ang[-1]))
        assert int(np.argmin(common)) - 5 == s0, (s0, common)


# ---- sinogram formation (Reslice/, ImageReg share source/) -------------------

def test_merge_frames_integer_mean():
    rng = np.random.default_rng(2)
    W, H, n = 5, 3, 4
    stride = W * H + 32
    raw = rng.integers(0, 65535, stride * n).astype(np.uint16)
    got = O.merge_frames(raw, n, W, H, stride, dtype=F64)
    frames = np.stack([raw[k * stride:k * stride + W * H] for k in range(n)]).astype(np.float64)
    assert np.array_equal(got.ravel(), frames.sum(0) / n)


def test_mt_conversion_closed_forms():
    W, H = 4, 2
    dark = np.full((H, W), 100.0)
    I0 = np.full((H, W), 1100.0)
    It = np.array([[600, 350, 100, 50], [1100, 99, 2000, 101]], np.uint16)
    mt = O.mt_conversion(dark, I0, It, 1, W, H, dtype=F64)[0]
    with np.errstate(divide="ignore"):
        tr = (1100.0 - 100.0) / (It.astype(np.float64) - 100.0)
    want = np.log(np.clip(np.where(np.isinf(tr), 1e7, tr), np.float32(1e-5), np.float32(1e7)))
    assert np.allclose(mt, want, rtol=1e-12, atol=1e-12)
    assert mt[0, 2] == np.log(np.float64(np.float32(1e7)))          # It = dark: +inf clamps to 1e7
    assert mt[0, 3] == np.log(np.float64(np.float32(1e-5)))         # negative transmission clamps to 1e-5
    zero = O.mt_conversion(np.full((1, 1), 5.0), np.full((1, 1), 5.0), np.full(1, 5, np.uint16), 1, 1, 1, dtype=F64)
    assert zero[0, 0, 0] == 0.0                                     # 0/0 -> NaN -> 1 -> ln 1


def test_reslice_transpose_shifts_and_bilinear():
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(3)
    A, H, W = 5, 9, 11
    mt = rng.uniform(-1.0, 2.0, (A, H, W)).astype(np.float32)
    plain = O.reslice(mt)
    assert np.array_equal(plain, mt.transpose(1, 0, 2))
    xs = np.array([0, 1, -2, 3, 0], np.float32)
    ys = np.array([1, 0, -1, 2, -3], np.float32)
    got = O.reslice(mt, xs, ys, baseup=0.25, y0=2, n_layers=5)
    for th in range(A):
        want = np.zeros((H, W), np.float32)
        for Y in range(H):
            for X in range(W):
                y, x = Y + int(ys[th]), X + int(xs[th])
                want[Y, X] = mt[th, y, x] if (0 <= y < H and 0 <= x < W) else 0.0
        assert np.array_equal(got[:, th], want[2:7] + np.float32(0.25)), th
    fx = rng.uniform(-2.5, 2.5, A).astype(np.float32)
    fy = rng.uniform(-2.5, 2.5, A).astype(np.float32)
    got = O.reslice(mt.astype(np.float64), fx, fy, dtype=F64)
    pad = np.pad(mt.astype(np.float64), ((0, 0), (2, 2), (2, 2)))
    Yg, Xg = np.mgrid[0:H, 0:W].astype(np.float64)
    for th in range(A):
        want = map_coordinates(pad[th], [Yg + float(fy[th]) + 2, Xg + float(fx[th]) + 2], order=1, mode="nearest")
        assert np.allclose(got[:, th], want, rtol=0, atol=1e-12), th
    corr = O.reslice(np.ones((1, 2, 8)), correction=1, dtype=F64)[0, 0]
    t = 2 * np.abs(np.arange(8) - 4.0) / 8
    assert np.allclose(corr, np.sqrt(1 - t * t), rtol=0, atol=1e-15)


def test_denoise_sinogram_spikes():
    A, W = 8, 5
    base = np.tile(np.linspace(0.0, 1.0, A)[:, None], (1, W))       # a ramp along theta stays
    src = base.copy()
    src[3, 1] += 5.0                                                 # up spike
    src[5, 4] -= 5.0                                                 # down spike
    src[0, 2] += 5.0                                                 # first row: its theta-1 neighbour is itself
    got = O.denoise_sinogram(src[None], 0.5, dtype=F64)[0]
    want = src.copy()
    want[3, 1] = (src[2, 1] + src[4, 1]) / 2
    want[5, 4] = (src[4, 4] + src[6, 4]) / 2
    assert np.array_equal(got, want)
    assert np.array_equal(O.denoise_sinogram(base[None], 0.5, dtype=F64)[0], base)


def test_xprojection_constant_image():
    A, H, W, c = 3, 6, 10, 0.75
    mt = np.full((A, H, W), c)
    got = O.xprojection(mt, 2, 7, dtype=F64)
    assert np.allclose(got[:, 1:], 6 * c, rtol=0, atol=1e-14)       # corners (j, Y) inside: 4 texels of c
    assert np.allclose(got[:, 0], 3 * c, rtol=0, atol=1e-14)        # row -1 is the zero border
    edge = O.xprojection(mt, 0, 0, dtype=F64)                        # column -1 is the zero border
    assert np.allclose(edge[:, 1:], c / 2, rtol=0, atol=1e-15) and np.allclose(edge[:, 0], c / 4, atol=1e-15)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_zcorrection_fixed_point_and_ramp(dtype):
    A, H = 12, 256
    Y = np.arange(H, dtype=np.float64)
    prof = np.exp(-((Y - 100) / 12) ** 2) + 0.1
    assert np.array_equal(O.zcorrection(np.tile(prof, (A, 1)), 10, 240, dtype=dtype), np.zeros(A))
    rng = np.random.default_rng(0)
    d = rng.uniform(-2.0, 2.0, A)
    d[A // 2 - 1:A // 2 + 1] = 0.0                                   # the target is the mean of these two angles
    xp = np.stack([0.5 * (Y + d[a]) + 3.0 for a in range(A)])
    e = np.array([(d[max(a - 1, 0)] + d[a]) / 2 for a in range(A)])  # each read is the mean of angles a-1, a
    ys = O.zcorrection(xp, 20, H - 20, dtype=dtype)
    k = np.abs(e) > 0
    assert np.all(np.abs(ys[k] / e[k] + 2.0) < 5e-3), ys / np.where(k, e, 1)
    assert np.all(ys[~k] == 0.0)


def test_reslice_execution_composition():
    rng = np.random.default_rng(5)
    A, H, W = 6, 300, 16
    Y = np.arange(H, dtype=np.float64)
    drift = rng.uniform(-2.0, 2.0, A)
    mt = np.stack([np.tile((0.5 * (Y + drift[a]) + 3.0)[:, None], (1, W)) for a in range(A)]).astype(np.float32)
    mt[2, 150, 7] += 9.0
    prj, ys = O.reslice_execution(mt, baseup=0.5, z_correction=True, start_x=1, end_x=W - 2, start_z=20, end_z=H - 20,
                                  denoise=True, threshold=1.0, y0=100, n_layers=80)
    xp = O.xprojection(mt, 1, W - 2)
    assert np.array_equal(ys, O.zcorrection(xp, 20, H - 20))
    want = O.denoise_sinogram(O.reslice(mt, None, ys, baseup=0.5, y0=100, n_layers=80), 1.0)
    assert np.array_equal(prj, want)


# ---- independent pins of the nearest-bin rules, scales and offsets ----------
#
# Each expected value below is computed in FP64 numpy from the reference's
# kernel text at one or two angles, with bins chosen independently of the
# oracle: rounding floor(s + P/2 + 0.5) (OSEM.cl:95, AART.cl:97,
# reconstShare.cl:265) against truncation floor(s + P/2) (FISTA.cl:237, :101),
# integer halves X - Nx/2 (reconstShare.cl:265, FISTA.cl:101) against float
# halves X - Nx*0.5f.  Sizes are odd so the two kinds of halves differ, and
# only pixels whose continuous bin coordinate lies at least 0.02 from an
# integer are compared, so FP32 and FP64 pick the same bin there.

def _probe_bins(nx, ny, px, theta_deg, round_half, int_half):
    """Continuous bin coordinate and its bin for every pixel at one angle."""
    th = np.float64(np.float32(np.float32(theta_deg) * np.float64(np.pi) / 180.0))
    c, s = np.cos(th), np.sin(th)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny))
    if int_half:
        xr, yr, ph = X - nx // 2, Y - ny // 2, float(px // 2)
    else:
        xr, yr, ph = X - nx * 0.5, Y - ny * 0.5, px * 0.5
    t = xr * c - yr * s + ph + (0.5 if round_half else 0.0)
    return t, np.floor(t).astype(np.int64)


def _probe_rows(pa, px):
    # distinct value per (angle, bin): a wrong bin or a wrong row changes it
    return (1.0 + np.arange(px)[None, :] + 100.0 * np.arange(pa)[:, None])[None]


def _gather(rows, th, b, px):
    ok = (b >= 0) & (b < px)
    return np.where(ok, rows[0, th][np.clip(b, 0, px - 1)], 0.0)


def _safe(t):
    f = t - np.floor(t)
    return (f > 0.02) & (f < 0.98)


PROBE_ANGLES = (75.5225, 21.3, 128.7)       # cos 0.25 and two generic angles


@pytest.mark.parametrize("dtype", [F32, F64])
def test_nearest_bin_probes_round_vs_truncate(dtype):
    """Single-angle subsets (pa = ss = 3, subset 1 = angle 1): each
    backprojector's output against its closed form; pixels at bin
    coordinate k + 0.5..1 separate rounding from truncation, and the odd
    sizes separate integer from float halves."""
    nx, ny, px, pa = 15, 13, 21, 3
    D = max(nx, ny)
    g = O.geom(nx, px, pa, ss=pa, ny=ny)
    ang = np.array(PROBE_ANGLES, np.float32)
    rows = _probe_rows(pa, px)
    rng = np.random.default_rng(31)
    x0 = rng.uniform(0.5, 1.5, (1, ny, nx))
    alpha, L = 0.7, np.array([0.9, 0.3, 2.0])
    sub = 1
    cases = {
        # name: (oracle output, round_half, int_half, closed form from the gathered sum q)
        "osem2": (O.osem2(g, ang, x0, rows, sub, dtype=dtype), True, False,
                  lambda q: x0[0] * q * pa / pa),
        "aart2": (O.aart2(g, ang, x0, rows, sub, alpha, dtype=dtype), True, False,
                  lambda q: x0[0] + q * alpha * pa / pa / D),
        "backprojection": (O.backprojection(g, ang, rows, sub, dtype=dtype), True, True,
                           lambda q: q * pa / pa / D),
        "fista_backprojection": (O.fista_backprojection(g, ang, x0, rows, L, sub, alpha, dtype=dtype),
                                 False, False, lambda q: x0[0] + q * alpha * pa / pa / D / L[sub] * 0.5),
    }
    for name, (out, rh, ih, form) in cases.items():
        t, b = _probe_bins(nx, ny, px, ang[sub], rh, ih)
        q = _gather(rows, sub, b, px)
        m = _safe(t)
        assert m.sum() > 0.8 * m.size, name
        want = form(q)
        assert np.allclose(out[0][m], want[m], rtol=2e-6, atol=1e-6), name
        # the test discriminates: the other rounding rule and the other kind
        # of halves each predict different values at many compared pixels
        for rh2, ih2 in ((not rh, ih), (rh, not ih)):
            t2, b2 = _probe_bins(nx, ny, px, ang[sub], rh2, ih2)
            alt = form(_gather(rows, sub, b2, px))
            m2 = m & _safe(t2)
            assert (np.abs(alt[m2] - want[m2]) > 1e-3).sum() > 0.1 * m2.sum(), (name, rh2, ih2)


@pytest.mark.parametrize("dtype", [F32, F64])
def test_power_iter2_probe_truncates_with_integer_halves(dtype):
    """powerIter2 (FISTA.cl:79-110): layer Z gathers its compact row i of
    angle Z + i*SS at floor((X - Nx/2)cos - (Y - Ny/2)sin + P/2), integer
    halves, no +0.5, then * SS / A / D."""
    nx, ny, px, pa = 15, 13, 21, 3
    D = max(nx, ny)
    g = O.geom(nx, px, pa, ss=pa, ny=ny)
    ang = np.array(PROBE_ANGLES, np.float32)
    prjc = np.zeros((pa, pa, px))
    for z in range(pa):
        prjc[z, 0] = 1.0 + np.arange(px) + 100.0 * z    # layer z: compact row 0 = angle z
    out = O.power_iter2(g, ang, prjc, dtype=dtype)
    for z in range(pa):
        t, b = _probe_bins(nx, ny, px, ang[z], False, True)
        ok = (b >= 0) & (b < px)
        want = np.where(ok, prjc[z, 0][np.clip(b, 0, px - 1)], 0.0) * pa / pa / D
        m = _safe(t)
        assert np.allclose(out[z][m], want[m], rtol=2e-6), z
        t2, b2 = _probe_bins(nx, ny, px, ang[z], True, True)
        ok2 = (b2 >= 0) & (b2 < px)
        alt = np.where(ok2, prjc[z, 0][np.clip(b2, 0, px - 1)], 0.0) * pa / pa / D
        m2 = m & _safe(t2)
        assert (np.abs(alt[m2] - want[m2]) > 1e-4).sum() > 0.1 * m2.sum()


def test_update_scales_over_a_two_angle_subset():
    """The update constants with SS != A: pa = 6, ss = 3, subset 2 = angles
    2 and 5.  Rows all equal to 1 make the gathered sum the count of the
    subset's angles whose bin falls on the detector (2 in the centre), so
    OSEM2 multiplies by 2 * 3/6 = 1, AART2 adds 2 * alpha * 3/6/D,
    backProjection gives 2 * 3/6/D and FISTAbackProjection adds
    2 * alpha * 3/6/D/L[2] * 0.5."""
    n, px, pa, ss, sub = 17, 17, 6, 3, 2
    g = O.geom(n, px, pa, ss=ss)
    ang = np.array([0, 30, 60, 90, 120, 150], np.float32)
    ones = np.ones((1, pa, px))
    x0 = np.full((1, n, n), 0.75)
    alpha, L = 0.6, np.array([1.0, 1.0, 0.25])
    yy, xx = np.mgrid[0:n, 0:n]
    centre = (xx - n / 2) ** 2 + (yy - n / 2) ** 2 < (n / 2 - 2) ** 2
    got = {
        "osem2": O.osem2(g, ang, x0, ones, sub, dtype=F64)[0],
        "aart2": O.aart2(g, ang, x0, ones, sub, alpha, dtype=F64)[0],
        "backprojection": O.backprojection(g, ang, ones, sub, dtype=F64)[0],
        "fista_backprojection": O.fista_backprojection(g, ang, x0, ones, L, sub, alpha, dtype=F64)[0],
    }
    want = {
        "osem2": 0.75 * 1.0,
        "aart2": 0.75 + 2 * 0.6 * 0.5 / 17,
        "backprojection": 2 * 0.5 / 17,
        "fista_backprojection": 0.75 + 2 * 0.6 * 0.5 / 17 / 0.25 * 0.5,
    }
    for k in got:
        assert np.allclose(got[k][centre], want[k], rtol=1e-12), k


# ---- powerIter1: the geometry without the +0.5 terms ------------------------

def test_power_iter1_half_texel_closed_forms():
    """powerIter1 (FISTA.cl:66-72) samples at x = (X - P/2)cos + (Y - D/2)sin +
    Nx/2 with integer halves and no +0.5: at theta = 0 every sample sits on a
    texel corner, so each ray is the half-texel average of columns X - 1 and
    X (column -1 reads the border 0), p = (C[X-1] + C[X]) / 2 when every row
    is hit twice (here Ny < D).  OSEM1 at the same angle samples texel
    centres in x: p = C[X], the plain column sums.  At theta = 90 degrees
    powerIter1 averages rows 13 - X and 14 - X over columns -1..15, so the
    last column is counted once: p = (R[13-X] + R[14-X]) / 2 - (T[13-X, 15] +
    T[14-X, 15]) / 4."""
    nx, ny, px = 16, 12, 16
    rng = np.random.default_rng(32)
    img = rng.uniform(0.0, 1.0, (1, ny, nx))
    g = O.geom(nx, px, 2, ss=1, ny=ny)
    ang = np.array([0.0, 90.0], np.float32)
    p1 = O.power_iter1(g, ang, img, dtype=F64)[0]
    C = img[0].sum(axis=0)
    Cm1 = np.concatenate([[0.0], C[:-1]])
    assert np.allclose(p1[0], 0.5 * (Cm1 + C), rtol=1e-12)
    R = img[0].sum(axis=1)
    T = img[0]
    X = np.arange(px)
    r0, r1 = 13 - X, 14 - X

    def row_or0(v, r):
        return np.where((r >= 0) & (r < ny), v[np.clip(r, 0, ny - 1)], 0.0)
    want90 = 0.5 * (row_or0(R, r0) + row_or0(R, r1)) - 0.25 * (row_or0(T[:, 15], r0) + row_or0(T[:, 15], r1))
    assert np.allclose(p1[1], want90, rtol=1e-5, atol=1e-6)      # cos(float 90 deg) = -4.4e-8
    # OSEM1 at theta = 0 on the same grid: column sums (texel centres)
    prj = np.ones((1, 2, px))
    r = O.osem1(g, ang, img, prj, 0, dtype=F64)
    assert np.allclose(1.0 / r[0, 0], C, rtol=1e-12)
    assert not np.allclose(p1[0], C, rtol=1e-3)


# ---- FISTA3D momentum: literal values ----------------------------------------

def test_fista3d_momentum_literals():
    """FISTA.cl:584-592 with lambda = 0 (the prox returns v itself, x_new = v):
    beta' = (sqrt(4 beta^2 + 1) + 1) / 2, gamma = (beta - 1) / beta',
    w = gamma x_new + (1 - gamma) x.  Literal values: beta = 0 -> beta' = 1,
    gamma = -1, w = 2x - v; beta = 1 -> beta' = 1.6180339887 (the golden
    ratio), gamma = 0, w = x; beta = 3 -> beta' = 3.5413812651,
    gamma = 0.5647513923; beta = 10 -> beta' = 10.5124921973, gamma = 0.8561242978."""
    n = 4
    g = O.geom(n, n, 1, ss=1)
    betas = np.array([0.0, 1.0, 3.0, 10.0])
    b = np.repeat(betas[:, None, None], n * n, axis=1).reshape(4, n, n)
    rng = np.random.default_rng(33)
    x = rng.uniform(0.5, 1.0, (4, n, n))
    v = rng.uniform(0.5, 1.0, (4, n, n))
    L = np.array([1.0])
    for dtype, tol in ((F64, 1e-12), (F32, 2e-7)):
        w, xn, bn = O.fista3d(g, x, v, b, L, 0, 0.0, dtype=dtype)
        assert np.allclose(xn, v.astype(dtype), rtol=0, atol=0)
        assert np.allclose(bn[:, 0, 0], [1.0, 1.6180339887498949, 3.5413812651491097, 10.512492197250394],
                           rtol=tol)
        gam = np.array([-1.0, 0.0, 2.0 / 3.5413812651491097, 9.0 / 10.512492197250394])
        assert abs(gam[2] - 0.5647513923) < 1e-9 and abs(gam[3] - 0.8561242978) < 1e-9
        want = gam[:, None, None] * v + (1.0 - gam[:, None, None]) * x
        assert np.allclose(w, want, rtol=4 * tol, atol=4 * tol)
        assert np.allclose(w[0], 2 * x[0] - v[0], rtol=4 * tol) and np.allclose(w[1], x[1], rtol=4 * tol)


# ---- backProjectionFBP: where an impulse lands ------------------------------

def test_fbp_backprojection_impulse_lands_on_its_line():
    """backProjectionFBP (FBP.cl:181-186) samples the filtered row linearly at
    u = s + P/2 + 0.5, i.e. at continuous bin s + P/2 with bin centres at
    integers.  With one angle and an impulse at bin b0, theta = 0 lights
    exactly column X = b0 (weight 1, value 1/(2A)) inside the disk; theta =
    90 degrees lights exactly row Y = (Ny + P)/2 - b0.  A half-bin offset
    error would split each line over two columns (rows) at 1/2."""
    n, px = 32, 32
    g = O.geom(n, px, 1, ss=1)
    b0 = 11
    fprj = np.zeros((1, 1, px))
    fprj[0, 0, b0] = 1.0
    yy, xx = np.mgrid[0:n, 0:n]
    disk = np.hypot(xx - n / 2, yy - n / 2) < n / 2
    for theta, lit in ((0.0, xx == b0), (90.0, yy == (n + px) // 2 - b0)):
        out = O.fbp_backprojection(g, np.array([theta], np.float32), fprj, dtype=F64)[0]
        want = np.where(lit & disk, 1.0 / 2.0, 0.0)
        assert np.allclose(out, want, atol=1e-6), theta
        assert np.abs(out[lit & disk] - 0.5).max() < 1e-6


# ---- per-slice TV prox of the joint alignment model -------------------------

def test_tv_prox2d_is_the_l1_prox_with_four_neighbours():
    """argmin_x 1/2 (x - v)^2 + lam sum_i |x - n_i| over the four in-plane
    neighbours (border 0), by brute force on a fine grid, then max(lower, x)."""
    rng = np.random.default_rng(34)
    n = 7
    g = O.geom(n, n, 1, ss=1)
    v = rng.uniform(0.0, 1.0, (1, n, n))
    v[0, 3, 3] = -0.2
    lam = 0.04
    grid = np.linspace(-0.6, 1.6, 220001)
    for dtype in (F64, F32):
        out = O.tv_prox2d(g, v, lam, lower=-1.0, dtype=dtype)
        for (yy, xx) in [(2, 3), (0, 0), (6, 6), (3, 3), (0, 4), (5, 1)]:
            nb = [v[0, yy, xx + 1] if xx + 1 < n else 0.0, v[0, yy, xx - 1] if xx >= 1 else 0.0,
                  v[0, yy + 1, xx] if yy + 1 < n else 0.0, v[0, yy - 1, xx] if yy >= 1 else 0.0]
            f = 0.5 * (grid - v[0, yy, xx]) ** 2 + lam * sum(np.abs(grid - t) for t in nb)
            assert abs(grid[np.argmin(f)] - out[0, yy, xx]) < 2e-5, (dtype, yy, xx)
    # the floor
    out0 = O.tv_prox2d(g, v, lam, lower=0.0, dtype=F64)
    assert out0.min() >= 0.0 and out0[0, 3, 3] == 0.0


def test_tv_prox2d_slices_are_independent():
    """Each slice's result depends on that slice alone: changing the other
    slices changes nothing (unlike ISTA3D, whose z-neighbours couple every
    slice to the next one and to layer 0)."""
    rng = np.random.default_rng(35)
    g = O.geom(9, 9, 1, ss=1)
    v = rng.uniform(0.0, 1.0, (4, 9, 9))
    a = O.tv_prox2d(g, v, 0.03, dtype=F64)
    w = v.copy()
    w[[0, 1, 3]] = rng.uniform(0.0, 1.0, (3, 9, 9))
    b = O.tv_prox2d(g, w, 0.03, dtype=F64)
    assert np.array_equal(a[2], b[2])
    single = O.tv_prox2d(g, v[2:3], 0.03, dtype=F64)
    assert np.array_equal(a[2], single[0])
    # ISTA3D over the same stack does couple slice 2 to the others
    L = np.array([0.5])
    i3a = O.ista3d(g, v, L, 0, 0.03, dtype=F64)
    i3b = O.ista3d(g, w, L, 0, 0.03, dtype=F64)
    assert not np.array_equal(i3a[2], i3b[2])
