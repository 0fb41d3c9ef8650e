"""The oracle pins catch plausible mistakes.  Each case copies the oracle's C
sources to a temporary directory, plants one mistake (a dropped or added
+0.5, swapped integer / float halves, a dropped scale factor, a changed
momentum formula), builds that copy, points the oracle module at it and runs
the pins of the function it changes: at least one must fail.  CPU only; the
committed oracle is never modified."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
import test_oracle_pins as P

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")

MUTANTS = {
    "OSEM2 truncates (no +0.5)": (
        "+ (R)g->px * (R)0.5 + (R)0.5;\n        R v = S(sample_nearest)(rows_z, g->px, g->pa, x, (R)th + (R)0.5);",
        "+ (R)g->px * (R)0.5;\n        R v = S(sample_nearest)(rows_z, g->px, g->pa, x, (R)th + (R)0.5);",
        ["test_nearest_bin_probes_round_vs_truncate"]),
    "FISTAbackProjection rounds (+0.5)": (
        "              + (R)g->px * (R)0.5;\n                    bprj += S(sample_nearest)(rows, g->px, g->pa, x, (R)th);",
        "              + (R)g->px * (R)0.5 + (R)0.5;\n                    bprj += S(sample_nearest)(rows, g->px, g->pa, x, (R)th);",
        ["test_nearest_bin_probes_round_vs_truncate"]),
    "AART2 without / D": (
        "bprj * (R)alpha * (R)g->ss / (R)g->pa / (R)D;",
        "bprj * (R)alpha * (R)g->ss / (R)g->pa;",
        ["test_nearest_bin_probes_round_vs_truncate", "test_update_scales_over_a_two_angle_subset"]),
    "FISTAbackProjection without * 0.5": (
        "/ L[sub] * (R)0.5;", "/ L[sub];",
        ["test_nearest_bin_probes_round_vs_truncate", "test_update_scales_over_a_two_angle_subset"]),
    "backProjection with float halves": (
        "R x = (R)(X - nx / 2) * RCOS(ap) - (R)(Y - ny / 2) * RSIN(ap)\n                          + (R)(g->px / 2) + (R)0.5;",
        "R x = ((R)X - (R)nx * (R)0.5) * RCOS(ap) - ((R)Y - (R)ny * (R)0.5) * RSIN(ap)\n"
        "                          + (R)g->px * (R)0.5 + (R)0.5;",
        ["test_nearest_bin_probes_round_vs_truncate"]),
    "powerIter2 rounds (+0.5)": (
        "+ (R)(g->px / 2);\n                    bprj += S(sample_nearest)(rows, g->px, g->pa, x, (R)i);",
        "+ (R)(g->px / 2) + (R)0.5;\n                    bprj += S(sample_nearest)(rows, g->px, g->pa, x, (R)i);",
        ["test_power_iter2_probe_truncates_with_integer_halves"]),
    "powerIter1 with the +0.5 terms": (
        "+ (R)(nx / 2);", "+ (R)(nx / 2) + (R)0.5;", ["test_power_iter1_half_texel_closed_forms"]),
    "FISTA3D gamma over beta + 1": (
        "R gamma = (beta - (R)1.0) / bn;", "R gamma = (beta - (R)1.0) / (beta + (R)1.0);",
        ["test_fista3d_momentum_literals"]),
    "FISTA3D w with x and x_new swapped": (
        "R w = gamma * xn + ((R)1.0 - gamma) * xo;", "R w = gamma * xo + ((R)1.0 - gamma) * xn;",
        ["test_fista3d_momentum_literals"]),
    "backProjectionFBP without +0.5": (
        "R x = dx * RCOS(ap) - dy * RSIN(ap) + (R)g->px * (R)0.5 + (R)0.5;",
        "R x = dx * RCOS(ap) - dy * RSIN(ap) + (R)g->px * (R)0.5;",
        ["test_fbp_backprojection_impulse_lands_on_its_line"]),
    "OSEM2 without SS / A": (
        "R v = img[o] * bprj * (R)g->ss / (R)g->pa;", "R v = img[o] * bprj;",
        ["test_nearest_bin_probes_round_vs_truncate", "test_update_scales_over_a_two_angle_subset"]),
    "TV prox 2-D with 2 lambda steps halved": (
        "vi - (R)4 * lambda, vi - (R)2 * lambda, vi, vi + (R)2 * lambda, vi + (R)4 * lambda};",
        "vi - (R)2 * lambda, vi - (R)1 * lambda, vi, vi + (R)1 * lambda, vi + (R)2 * lambda};",
        ["test_tv_prox2d_is_the_l1_prox_with_four_neighbours"]),
}


def _run_pin(name):
    fn = getattr(P, name)
    if name in ("test_nearest_bin_probes_round_vs_truncate", "test_power_iter2_probe_truncates_with_integer_halves"):
        for dt in (np.float32, np.float64):
            fn(dt)
    else:
        fn()


@pytest.mark.parametrize("case", sorted(MUTANTS))
def test_pins_fail_on_a_planted_mistake(case, tmp_path):
    old, new, pins = MUTANTS[case]
    for f in ("ct_oracle.c", "ct_oracle.h", "ct_oracle_impl.h"):
        shutil.copy(os.path.join(ORACLE, f), tmp_path / f)
    impl = (tmp_path / "ct_oracle_impl.h").read_text()
    assert impl.count(old) == 1, f"mutation site not found once: {case}"
    (tmp_path / "ct_oracle_impl.h").write_text(impl.replace(old, new))
    so = tmp_path / "liboracle_mutant.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11", "-shared", "-o", str(so),
                    str(tmp_path / "ct_oracle.c"), "-lm"], check=True)
    import ctypes
    saved = O._lib
    O._lib = ctypes.CDLL(str(so))
    try:
        failed = []
        for name in pins:
            try:
                _run_pin(name)
            except AssertionError:
                failed.append(name)
        assert failed, f"no pin caught: {case}"
    finally:
        O._lib = saved
