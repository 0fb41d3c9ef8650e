"""The C++ host drop-in (ct_reconstruction): input-parameter file keys and
defaults (CT_reconstruction/main.cpp:12-41,
CT_reconstruction_input_parameter.cpp:47-179), file checks of main.cpp:69-105
and its error reporting.  No GPU needed."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ct-xafs_analysis_amd", "ct_reconstruction")

# every key of the reference's parameter reader with a non-default value
FULL_INP = [
    ("#Open CL Platform & Device No", "1,3-4"),
    ("#Number of Parallel processing per one device", "7"),
    ("#測定ファイルディレクトリパス", "/data/sino dir"),
    ("#測定ファイル名", "s****.raw"),
    ("#出力ファイルディレクトリパス", "/data/out"),
    ("#出力ファイル名", "rec_****.img"),
    ("#ログファイルパス", "/data/angles.dat"),
    ("#初期画像ファイル名", "/data/first.img"),
    ("#一時ファイルディレクトリパス", "/tmp/x"),
    ("#レイヤー数", "12"),
    ("#開始ナンバー", "5"),
    ("#初期画像の使用 1:Yes 2: No", "1"),
    ("#カメラピクセル数", "384"),
    ("#投影数", "900"),
    ("#再構成画像サイズ X", "400"),
    ("#再構成画像サイズ Y", "300"),
    ("#出力画像サイズ X", "400"),
    ("#出力画像サイズ Y", "300"),
    ("#再構成法", "1"),
    ("#反復回数", "3"),
    ("#AARTファクター", "0.5"),
    ("#サブセット(OSEM)", "15"),
    ("#ゼロパディングサイズ(FBP)", "1024"),
    ("#投影像補正 0:なし 1: x方向のみ 2:θ方向のみ 3:x,θ方向両方", "2"),
    ("#再構成像強度増幅因子", "2.5"),
    ("#Base up 減少速度次数 (Hybrid)", "3"),
    ("#逐次圧縮センシング計算 (FBP適応外)", "1"),
    ("#逐次圧縮センシング計算加算ファクター", "0.002"),
    ("#逐次圧縮センシング計算重複レイヤー数", "2"),
]


def write_inp(path, pairs, crlf=False, bom=False):
    nl = "\r\n" if crlf else "\n"
    text = nl.join(f"{k}{nl}{v}" for k, v in pairs) + nl
    data = text.encode("utf-8")
    if bom:
        data = b"\xef\xbb\xbf" + data
    with open(path, "wb") as f:
        f.write(data)


def params(path):
    r = subprocess.run([CLI, "--print-params", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_cli_built():
    assert os.path.exists(CLI), "build first (__graft_entry__.build())"


@pytest.mark.parametrize("crlf,bom", [(False, False), (True, True)])
def test_every_reference_key(tmp_path, crlf, bom):
    inp = tmp_path / "p.inp"
    write_inp(inp, FULL_INP, crlf=crlf, bom=bom)
    p = params(inp)
    assert p["g_devList"] == "1,3-4" and p["numParallel"] == 7
    assert p["g_d1"] == "/data/sino dir" and p["g_f1"] == "s****.raw"
    assert p["g_d2"] == "/data/out" and p["g_f2"] == "rec_****.img"
    assert p["g_f3"] == "/data/angles.dat" and p["g_f4"] == "/data/first.img" and p["g_d5"] == "/tmp/x"
    assert (p["g_num"], p["g_st"], p["g_fst"]) == (12, 5, 1)
    assert (p["g_px"], p["g_pa"], p["g_nx"], p["g_ny"], p["g_ox"], p["g_oy"]) == (384, 900, 400, 300, 400, 300)
    assert (p["g_mode"], p["g_it"], p["g_ss"], p["g_zp"]) == (1, 3, 15, 1024)
    assert p["g_wt1"] == 0.5 and p["correctionMode"] == 2 and p["amp"] == 2.5 and p["baseupOrder"] == 3
    assert p["CSitBool"] is True and abs(p["CSlambda"] - 0.002) < 1e-9 and p["CSoverlap"] == 2


def test_defaults_of_main_cpp(tmp_path):
    inp = tmp_path / "empty.inp"
    inp.write_text("# nothing known here\n")
    p = params(inp)
    assert (p["g_d1"], p["g_f1"], p["g_d2"], p["g_f2"], p["g_f3"]) == ("input", "i****.raw", "output", "r****.img",
                                                                       "log.dat")
    assert (p["g_fst"], p["g_ini"], p["g_px"], p["g_pa"], p["g_nx"], p["g_ny"]) == (2, 1, 2048, 1600, 2048, 2048)
    assert (p["g_mode"], p["g_it"], p["g_num"], p["g_st"], p["g_ss"], p["g_zp"]) == (6, 1, 1, 1, 20, 4096)
    assert (p["g_wt1"], p["g_wt2"], p["numParallel"], p["correctionMode"]) == (1, 0.01, 10, 0)
    assert p["CSitBool"] is False and abs(p["CSlambda"] - 1e-4) < 1e-12 and p["CSoverlap"] == 1


def test_single_size_key_sets_grid_and_output(tmp_path):
    inp = tmp_path / "s.inp"
    write_inp(inp, [("#再構成画像サイズ", "512")])
    p = params(inp)
    assert (p["g_nx"], p["g_ny"], p["g_ox"], p["g_oy"]) == (512, 512, 512, 512)


@pytest.mark.parametrize("mode,expect", [(1, 20), (2, 1600), (3, 1), (4, 1), (5, 1), (6, 20), (7, 20)])
def test_method_subset_counts(tmp_path, mode, expect):
    inp = tmp_path / "m.inp"
    write_inp(inp, [("#再構成法", str(mode))])
    assert params(inp)["method_subsets"] == expect


def test_bad_value_is_an_error(tmp_path):
    inp = tmp_path / "bad.inp"
    write_inp(inp, [("#投影数", "many")])
    r = subprocess.run([CLI, "--print-params", str(inp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "#投影数" in r.stderr


def test_missing_parameter_file(tmp_path):
    r = subprocess.run([CLI, str(tmp_path / "nope.inp")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not found" in r.stderr


def test_missing_input_directory_and_file(tmp_path):
    inp = tmp_path / "p.inp"
    write_inp(inp, [("#測定ファイルディレクトリパス", str(tmp_path / "absent"))])
    r = subprocess.run([CLI, str(inp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "directory not found" in r.stderr
    (tmp_path / "in").mkdir()
    write_inp(inp, [("#測定ファイルディレクトリパス", str(tmp_path / "in")), ("#開始ナンバー", "3")])
    r = subprocess.run([CLI, str(inp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "i0003.raw" in r.stderr


def _tiny_dataset(tmp_path, extra):
    d = tmp_path / "in"
    d.mkdir()
    np.zeros(16 * 12, np.float32).tofile(d / "i0001.raw")
    (tmp_path / "log.dat").write_text("".join(f"{a}\n" for a in np.arange(12) * 15.0))
    inp = tmp_path / "p.inp"
    write_inp(inp, [("#測定ファイルディレクトリパス", str(d)), ("#出力ファイルディレクトリパス", str(tmp_path / "out")),
                    ("#ログファイルパス", str(tmp_path / "log.dat")), ("#カメラピクセル数", "16"), ("#投影数", "12"),
                    ("#再構成画像サイズ", "16"), ("#サブセット(OSEM)", "4")] + extra)
    return inp


@pytest.mark.parametrize("extra,msg", [
    ([("#再構成法", "9")], "not 1-8"),
    ([("#再構成法", "8"), ("#ゼロパディングサイズ(FBP)", "24")], "power of two"),
    ([("#サブセット(OSEM)", "5")], "not a multiple"),
    ([("#出力画像サイズ X", "8")], "output size"),
    ([("#出力画像サイズ Y", "20")], "within the reconstruction size"),
    ([("#再構成法", "1"), ("#出力画像サイズ Y", "8")], "except for FBP"),
    ([("#再構成法", "7"), ("#ゼロパディングサイズ(FBP)", "24")], "power of two"),
])
def test_rejected_parameter_sets(tmp_path, extra, msg):
    inp = _tiny_dataset(tmp_path, extra)
    r = subprocess.run([CLI, str(inp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and msg in r.stdout + r.stderr


def test_fbp_output_region_passes_validation(tmp_path):
    """FBP takes an output region below the grid size (its top-left block is
    read back, GPU.cpp:848-870); the run then stops only for want of a GPU
    here, or runs to the end where one is present."""
    inp = _tiny_dataset(tmp_path, [("#再構成法", "7"), ("#ゼロパディングサイズ(FBP)", "32"),
                                   ("#出力画像サイズ X", "8"), ("#出力画像サイズ Y", "6")])
    r = subprocess.run([CLI, str(inp)], capture_output=True, text=True, timeout=120)
    assert "output size" not in r.stdout + r.stderr


def test_short_angle_log(tmp_path):
    inp = _tiny_dataset(tmp_path, [("#投影数", "24"), ("#サブセット(OSEM)", "4")])
    r = subprocess.run([CLI, str(inp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "fewer than 24 angles" in r.stdout + r.stderr
