"""End-to-end run of the rotation-centre-search drop-in on the GPU:
synthetic .his files (dark, I0 and It frames laid out as
share source/readHisRawFile.cpp:63-95 reads them) of a phantom whose rotation
axis sits s0 bins off the detector centre; the candidates' images,
result.dat and the chosen shift against the oracle's merge -> mu-t ->
reslice -> centre search."""
import subprocess

import numpy as np
import pytest

import oracle as O
from _util import relerr, say
from test_host_centersearch import CLI_R
from test_host_cli import write_inp
from test_oracle_pins import _filled_phantom, _zero_fill_shift

pytestmark = pytest.mark.gpu


def write_his(path, frames, comment=b"synthetic"):
    """Header of 4 + len(comment) + 64 bytes (16-bit comment length at byte
    2), then every frame followed by 32 zero pixels."""
    n, H, W = frames.shape
    head = bytearray(4 + len(comment) + 64)
    head[0:2] = b"IM"
    head[2:4] = len(comment).to_bytes(2, "little")
    head[4:4 + len(comment)] = comment
    with open(path, "wb") as f:
        f.write(bytes(head))
        gap = np.zeros(32, np.uint16).tobytes()
        for k in range(n):
            f.write(np.ascontiguousarray(frames[k], dtype="<u2").tobytes())
            f.write(gap)


def test_center_search_from_his_files(tmp_path):
    n, W, A, ss, s0, H, layer, scans = 48, 80, 72, 8, 2, 4, 3, 5
    rng = np.random.default_rng(7)
    ang = (np.arange(A) * 180.0 / A).astype(np.float32)
    g = O.geom(n, W, A, ss=ss)
    sino = O.align_forward(g, ang, _filled_phantom(n, 1)[None])[0]
    mu = _zero_fill_shift(sino, -s0) / 40.0                       # axis at W/2 + s0, attenuation ~ 0.5
    dark = rng.integers(95, 105, (scans, H, W)).astype(np.uint16)
    i0 = rng.integers(39000, 41000, (scans, H, W)).astype(np.uint16)
    It = np.empty((A, H, W), np.uint16)
    for a in range(A):
        t = 100.0 + 39900.0 * np.exp(-np.tile(mu[a], (H, 1)) * (1.0 + 0.05 * np.arange(H)[:, None]))
        It[a] = np.round(t).astype(np.uint16)
    d = tmp_path / "his"
    d.mkdir()
    write_his(d / "scan_dark.his", dark)
    write_his(d / "scan_005_I0.his", i0)
    write_his(d / "scan_005.his", It)
    (tmp_path / "log.dat").write_text("".join(f"{a:.9g}\n" for a in ang))
    inp = tmp_path / "r.inp"
    write_inp(inp, [("#Input directory path", str(d)), ("#Output directory path", str(tmp_path / "out")),
                    ("#Output file base name", "rc_"), ("#Angle number range", f"1-{A}"),
                    ("#Image pixel size", f"{W},{H}"), ("#Target energy number for rotation center search", "5"),
                    ("#Number of scan for I0 and dark image", str(scans)),
                    ("#Target layer number for rotation center search", str(layer)),
                    ("#Start shift for rotation center search", "-3"),
                    ("#Number of shift step for rotation center search", "11"),
                    ("#Shift step for rotation center search", "0.5"),
                    ("#ログファイルパス", str(tmp_path / "log.dat")), ("#カメラピクセル数", str(W)),
                    ("#再構成画像サイズ", str(n)), ("#サブセット(OSEM)", str(ss))], crlf=True)
    say("ct_rotation_center_search")
    r = subprocess.run([CLI_R, str(inp)], capture_output=True, text=True, timeout=180)
    print(r.stdout[-2500:], r.stderr[-1000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr

    # the oracle's chain on the full frames
    M = W * H
    dm = O.merge_frames(dark.reshape(-1), scans, W, H)
    im = O.merge_frames(i0.reshape(-1), scans, W, H)
    mt = O.mt_conversion(dm, im, It.reshape(-1), A, W, H)
    y = O.reslice(mt, y0=layer - 1, n_layers=1)[0]
    rec, f = O.rotation_center_search(g, ang, y, 11, -3.0, 0.5, it=1, init=1.0, correction_mode=1)
    got = np.stack([np.fromfile(tmp_path / "out" / f"rc_{i + 1:04d}.raw", np.float32).reshape(n, n)
                    for i in range(11)])
    e = relerr(got, rec)
    lines = (tmp_path / "out" / "result.dat").read_text().splitlines()
    assert lines[0] == "Rotation center shift\tFocus Index\tfilename"
    fi = np.array([float(s.split("\t")[1]) for s in lines[1:]])
    sh = np.array([float(s.split("\t")[0]) for s in lines[1:]])
    print(f"[parity] centre search from .his: images {e:.3e}, focus index {relerr(fi, f):.3e}", flush=True)
    assert e < 1e-5 and relerr(fi, f) < 1e-4
    assert np.allclose(sh, -3.0 + 0.5 * np.arange(11))
    best = -3.0 + 0.5 * int(np.argmin(f))
    assert f"Estimated rotation center shift is {best:g}" in r.stdout
    assert abs(best - s0) <= 1.0
