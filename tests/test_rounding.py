"""The backprojector's nearest bin floor(bx + 0.5f) (OSEM.cl:95, AART.cl:97)
is computed with one v_cvt_rpi_i32_f32, which floors the exact sum bx + 0.5.
The reference floors the float-rounded sum.  The two differ only where
rounding bx + 0.5 crosses an integer, which for |bx| <= 65536 happens at the
single value bx = 0.5 - 2^-25; bx = fl(v + px / 2) with px >= 2 never takes
it.  These checks run on the CPU in float32 / float64 numpy arithmetic."""
import numpy as np


def _floats(lo, hi):
    """Every float32 in [lo, hi), 0 < lo < hi."""
    a, b = np.float32(lo).view(np.uint32), np.float32(hi).view(np.uint32)
    return np.arange(a, b, dtype=np.uint32).view(np.float32)


def _two_step(w):
    return np.floor((w + np.float32(0.5)).astype(np.float32))


def _exact(w):
    return np.floor(w.astype(np.float64) + 0.5)


def test_two_step_and_exact_floor_differ_only_below_one_half():
    bad = []
    for lo, hi in ((2.0 ** -3, 1.0), (1.0, 4.0), (4096.0, 8192.0)):
        for sign in (1.0, -1.0):
            w = (_floats(lo, hi) * np.float32(sign)).astype(np.float32)
            d = _two_step(w) != _exact(w)
            bad += w[d].tolist()
    assert bad == [np.float32(0.5) - np.float32(2.0 ** -25)]


def test_bin_coordinate_never_hits_the_exception():
    # bx = fl(v + ph) for every float v within 1.5 of the point where bx = 0.5,
    # for the half-sizes ph of px = 2, 3, 4, 5 and 1024
    target = np.float32(0.5) - np.float32(2.0 ** -25)
    for ph in (1.0, 1.5, 2.0, 2.5, 512.0):
        c = 0.5 - ph
        v = np.concatenate([-_floats(max(-c - 1.5, 2.0 ** -3), -c + 1.5)]).astype(np.float32)
        bx = (v + np.float32(ph)).astype(np.float32)
        assert not np.any(bx == target), ph
        small = (np.abs(v) < 0.5)
        assert np.all(bx[small] >= 0.5)
