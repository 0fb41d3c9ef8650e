"""The rotation-centre-search drop-in (ct_rotation_center_search) without a
GPU: its parameter keys (share source/input_parameter_class.cpp:318-470,
input_parameter_reslice_class.cpp:150-240, RotationCenterSearch/main.cpp
defaults), the flags of RotationCenterSearch/main.cpp:84-211 and the
parameter sets it refuses before touching a device."""
import json
import os
import subprocess

import pytest

from test_host_cli import write_inp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI_R = os.path.join(ROOT, "ct-xafs_analysis_amd", "ct_rotation_center_search")


def params(path, *flags):
    r = subprocess.run([CLI_R, "--print-params", str(path), *flags], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_cli_built():
    assert os.path.exists(CLI_R), "build first (__graft_entry__.build())"


def test_keys_defaults_and_flags(tmp_path):
    inp = tmp_path / "r.inp"
    write_inp(inp, [("#Input directory path", "/his"), ("#Output directory path", "/out"),
                    ("#Output file base name", "rc_"), ("#Angle number range", "1-1600"),
                    ("#Image pixel size", "2048,2048"), ("#Target energy number for rotation center search", "12"),
                    ("#Number of scan for I0 and dark image", "20"),
                    ("#Target layer number for rotation center search", "1024"),
                    ("#Start shift for rotation center search", "-10.5"),
                    ("#Number of shift step for rotation center search", "43"),
                    ("#Shift step for rotation center search", "0.5")], crlf=True)
    p = params(inp)
    assert (p["input_dir"], p["output_dir"], p["output_base"]) == ("/his", "/out", "rc_")
    assert p["angles"] == [1, 1600] and p["image"] == [2048, 2048] and p["target_energy"] == 12
    assert (p["scan_n"], p["layer"], p["shift_start"], p["shift_n"], p["shift_step"]) == (20, 1024, -10.5, 43, 0.5)
    assert p["correctionMode"] == 1                      # RotationCenterSearch/main.cpp:43
    q = params(inp, "-re", "3", "-ly", "7", "-rss", "-2", "-rsn", "9", "-rsst", "0.25", "-ob", "x_")
    assert (q["target_energy"], q["layer"], q["shift_start"], q["shift_n"], q["shift_step"], q["output_base"]) == \
        (3, 7, -2.0, 9, 0.25, "x_")
    empty = tmp_path / "empty.inp"
    empty.write_text("\n")
    d = params(empty)
    assert d["scan_n"] == 30 and d["shift_step"] is None and d["layer"] == 0


@pytest.mark.parametrize("extra,msg", [
    ([("#Target energy number for rotation center search", "-1")], "no target energy"),
    ([("#Target layer number for rotation center search", "0")], "target layer"),
    ([("#Number of shift step for rotation center search", "0")], "shift count"),
    ([("#カメラピクセル数", "64")], "differs from the camera width"),
    ([("#サブセット(OSEM)", "7")], "not a multiple"),
    ([("#Input directory path", "/nonexistent/his")], "directory not found"),
])
def test_rejected_parameter_sets(tmp_path, extra, msg):
    (tmp_path / "log.dat").write_text("".join(f"{a * 15.0}\n" for a in range(12)))
    keys = {"#Input directory path": str(tmp_path), "#Output directory path": str(tmp_path / "out"),
            "#Angle number range": "1-12", "#Image pixel size": "32,8",
            "#Target energy number for rotation center search": "1",
            "#Target layer number for rotation center search": "3",
            "#Number of shift step for rotation center search": "5", "#Shift step for rotation center search": "1",
            "#ログファイルパス": str(tmp_path / "log.dat"), "#カメラピクセル数": "32", "#サブセット(OSEM)": "4"}
    keys.update(dict(extra))
    inp = tmp_path / "r.inp"
    write_inp(inp, list(keys.items()))
    r = subprocess.run([CLI_R, str(inp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and msg in r.stdout + r.stderr, r.stdout + r.stderr
