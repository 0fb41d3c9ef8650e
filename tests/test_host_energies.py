"""The multi-energy drop-in (ct_reslice_reconstruction) without a GPU: the
shared parameter keys it reads (share source/input_parameter_class.cpp:318-470,
input_parameter_fitting_class.cpp:599-610), the command-line flags of
Reslice_CTreconst/main.cpp:72-164 and the parameter sets it refuses."""
import json
import os
import subprocess

import numpy as np
import pytest

from test_host_cli import write_inp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI_E = os.path.join(ROOT, "ct-xafs_analysis_amd", "ct_reslice_reconstruction")


def params(path, *flags):
    r = subprocess.run([CLI_E, "--print-params", str(path), *flags], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_cli_built():
    assert os.path.exists(CLI_E), "build first (__graft_entry__.build())"


def test_shared_keys_and_defaults(tmp_path):
    inp = tmp_path / "e.inp"
    write_inp(inp, [("#Input directory path", "/data/mt"), ("#Output directory path", "/data/ct"),
                    ("#Output file base name", "rec_"), ("#Angle number range", "1-1600"),
                    ("#Energy number range", "3-40"), ("#Image pixel size", "2048,1024"),
                    ("#Processing parameter name", "a0,a1,E0"), ("#再構成法", "7")], crlf=True, bom=True)
    p = params(inp)
    assert (p["input_dir"], p["output_dir"], p["output_base"]) == ("/data/mt", "/data/ct", "rec_")
    assert p["angles"] == [1, 1600] and p["energies"] == [3, 40] and p["image"] == [2048, 1024]
    assert p["names"] == ["a0", "a1", "E0"] and p["g_mode"] == 7
    assert abs(p["CSlambda"] - 0.001) < 1e-9          # Reslice_CTreconst/main.cpp:48


def test_flags_override_the_file(tmp_path):
    inp = tmp_path / "e.inp"
    write_inp(inp, [("#Input directory path", "/a"), ("#Angle number range", "1-10")])
    p = params(inp, "-ip", "/b", "-enr", "5-7", "-fpn", "x,y", "-op", "/c", "-ar", "2-9")
    assert (p["input_dir"], p["output_dir"]) == ("/b", "/c")
    assert p["energies"] == [5, 7] and p["names"] == ["x", "y"] and p["angles"] == [2, 9]


def _dataset(tmp_path, extra):
    d = tmp_path / "mt" / "001"
    d.mkdir(parents=True)
    for a in range(1, 13):
        np.zeros((6, 16), np.float32).tofile(d / f"mt_{a:04d}.raw")
    (tmp_path / "log.dat").write_text("".join(f"{a * 15.0}\n" for a in range(12)))
    inp = tmp_path / "e.inp"
    base = [("#Input directory path", str(tmp_path / "mt")), ("#Output directory path", str(tmp_path / "out")),
            ("#Angle number range", "1-12"), ("#Energy number range", "1-1"), ("#Image pixel size", "16,6"),
            ("#ログファイルパス", str(tmp_path / "log.dat")), ("#カメラピクセル数", "16"), ("#投影数", "12"),
            ("#再構成画像サイズ", "16"), ("#サブセット(OSEM)", "4"), ("#開始ナンバー", "1"), ("#レイヤー数", "3")]
    keys = dict(base)
    keys.update(dict(extra))
    write_inp(inp, list(keys.items()))
    return inp


@pytest.mark.parametrize("extra,msg", [
    ([("#再構成法", "1")], "not run by the multi-energy driver"),
    ([("#再構成法", "8")], "not run by the multi-energy driver"),
    ([("#投影数", "24")], "holds 12 projections"),
    ([("#カメラピクセル数", "20")], "differs from"),
    ([("#レイヤー数", "1")], "at least 2"),
    ([("#開始ナンバー", "5"), ("#レイヤー数", "3")], "camera rows"),
    ([("#Energy number range", "-1--1")], "no energy range"),
    ([("#Input directory path", "/nonexistent/dir")], "directory not found"),
])
def test_rejected_parameter_sets(tmp_path, extra, msg):
    inp = _dataset(tmp_path, extra)
    r = subprocess.run([CLI_E, str(inp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and msg in r.stdout + r.stderr, r.stdout + r.stderr
