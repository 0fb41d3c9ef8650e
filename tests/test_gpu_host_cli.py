"""End-to-end runs of the C++ host drop-in (ct_reconstruction) on the GPU:
parameter file + angle log + raw float32 layer files in, raw images out, each
layer compared with the CPU oracle's driver for the same method
(CT_reconstruction/GPU.cpp:23-97 method dispatch; the batch overlap of the
3-D TV pipelines at GPU.cpp:443-457, 634-648)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from _util import phantom_stack, relerr, say
from test_host_cli import CLI, write_inp

pytestmark = pytest.mark.gpu

N, PA = 48, 24


def dataset(tmp_path, nz, extra, seed=5, init=None):
    imgs, sinos, ang = phantom_stack(nz, N, PA, 5, seed, i0=2e4)
    d = tmp_path / "in"
    d.mkdir()
    for z in range(nz):
        sinos[z].astype(np.float32).tofile(d / f"i{z + 1:04d}.raw")
    (tmp_path / "log.dat").write_text("".join(f"{a:.9g}\r\n" for a in ang))
    pairs = [("#測定ファイルディレクトリパス", str(d)), ("#出力ファイルディレクトリパス", str(tmp_path / "out")),
             ("#ログファイルパス", str(tmp_path / "log.dat")), ("#カメラピクセル数", str(N)), ("#投影数", str(PA)),
             ("#再構成画像サイズ", str(N)), ("#レイヤー数", str(nz))]
    if init is not None:
        init.astype(np.float32).tofile(tmp_path / "first.img")
        pairs += [("#初期画像の使用 1:Yes 2: No", "1"), ("#初期画像ファイル名", str(tmp_path / "first.img"))]
    inp = tmp_path / "p.inp"
    write_inp(inp, pairs + extra, crlf=True)
    return inp, sinos, ang


def run_cli(inp, label):
    say("ct_reconstruction", label)
    r = subprocess.run([CLI, str(inp)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def outputs(tmp_path, nz):
    return np.stack([np.fromfile(tmp_path / "out" / f"r{z + 1:04d}.img", np.float32).reshape(N, N)
                     for z in range(nz)])


@pytest.mark.parametrize("mode,ss,it,corr", [(6, 6, 2, 0), (6, 4, 1, 2), (5, 6, 2, 0), (2, 6, 1, 0)])
def test_osem_family(tmp_path, mode, ss, it, corr):
    nz = 5
    inp, sinos, ang = dataset(tmp_path, nz, [("#再構成法", str(mode)), ("#サブセット(OSEM)", str(ss)),
                                             ("#反復回数", str(it)), ("#Number of Parallel processing per one device", "2"),
                                             ("#投影像補正 0:なし 1: x方向のみ 2:θ方向のみ 3:x,θ方向両方", str(corr))])
    out = run_cli(inp, f"mode {mode} ss {ss} it {it} corr {corr}")
    assert "[Layer 0005/5]" in out
    got = outputs(tmp_path, nz)
    ss_eff = {2: PA, 5: 1}.get(mode, ss)
    g = O.geom(N, N, PA, ss=ss_eff)
    want = O.osem_execution(g, ang, np.ones((nz, N, N), np.float32), sinos, it, correction_mode=corr)
    e = relerr(got, want)
    print(f"[parity] CLI mode {mode}: relerr vs FP32 compat {e:.3e}")
    assert e < 1e-5


def test_initial_image_file(tmp_path):
    nz = 2
    x0 = np.random.default_rng(3).uniform(0.5, 1.5, (N, N)).astype(np.float32)
    inp, sinos, ang = dataset(tmp_path, nz, [("#サブセット(OSEM)", "4"), ("#反復回数", "1")], init=x0)
    run_cli(inp, "initial image file")
    want = O.osem_execution(O.geom(N, N, PA, ss=4), ang, np.stack([x0] * nz), sinos, 1)
    assert relerr(outputs(tmp_path, nz), want) < 1e-5


def test_fbp(tmp_path):
    nz = 3
    inp, sinos, ang = dataset(tmp_path, nz, [("#再構成法", "7"), ("#ゼロパディングサイズ(FBP)", "128")])
    run_cli(inp, "FBP")
    want = O.fbp_execution(O.geom(N, N, PA, zp=128), ang, sinos)
    assert relerr(outputs(tmp_path, nz), want) < 1e-5


def test_fbp_output_region(tmp_path):
    """An output region below the grid size: backProjectionFBP runs over an
    ox x oy range in full-grid coordinates and FBP_thread reads that top-left
    block back (GPU.cpp:704, 848-870), so each file holds oy rows of ox
    values of the full-grid image."""
    nz, ox, oy = 2, 30, 20
    inp, sinos, ang = dataset(tmp_path, nz, [("#再構成法", "7"), ("#ゼロパディングサイズ(FBP)", "128"),
                                             ("#出力画像サイズ X", str(ox)), ("#出力画像サイズ Y", str(oy))])
    run_cli(inp, "FBP output region")
    want = O.fbp_execution(O.geom(N, N, PA, zp=128), ang, sinos)[:, :oy, :ox]
    got = [np.fromfile(tmp_path / "out" / f"r{z + 1:04d}.img", np.float32) for z in range(nz)]
    assert all(g.size == ox * oy for g in got)
    e = relerr(np.stack(got).reshape(nz, oy, ox), want)
    print(f"[parity] CLI FBP output region {ox} x {oy}: relerr {e:.3e}", flush=True)
    assert e < 1e-5 and np.abs(want).max() > 0


@pytest.mark.parametrize("mode,alpha", [(1, 1.0), (3, 0.01)])
def test_aart_asirt(tmp_path, mode, alpha):
    nz = 3
    inp, sinos, ang = dataset(tmp_path, nz, [("#再構成法", str(mode)), ("#サブセット(OSEM)", "6"), ("#反復回数", "2"),
                                             ("#Number of Parallel processing per one device", "2")])
    run_cli(inp, f"AART mode {mode}")
    g = O.geom(N, N, PA, ss=6 if mode == 1 else 1)
    want = O.aart_execution(g, ang, np.ones((nz, N, N), np.float32), sinos, 2, alpha)
    e = relerr(outputs(tmp_path, nz), want)
    print(f"[parity] CLI AART mode {mode}: relerr {e:.3e}")
    assert e < 1e-5


@pytest.mark.parametrize("mode", [1, 6])
def test_compressed_sensing_batches_overlap(tmp_path, mode):
    """num=6 layers in batches of 4 overlapping by 2 x 1 layer: [1..4] writes
    1..3, [3..6] writes 4..6; the 3-D TV couples the layers of a batch."""
    nz, lam = 6, 2e-3
    inp, sinos, ang = dataset(tmp_path, nz, [("#再構成法", str(mode)), ("#サブセット(OSEM)", "6"), ("#反復回数", "2"),
                                             ("#Number of Parallel processing per one device", "4"),
                                             ("#逐次圧縮センシング計算 (FBP適応外)", "1"),
                                             ("#逐次圧縮センシング計算加算ファクター", str(lam)),
                                             ("#逐次圧縮センシング計算重複レイヤー数", "1")])
    run_cli(inp, f"compressed sensing mode {mode}")
    got = outputs(tmp_path, nz)
    g = O.geom(N, N, PA, ss=6)
    x0 = np.ones((4, N, N), np.float32)
    want = np.zeros((nz, N, N), np.float32)
    L = O.power_iteration(g, ang, 30) if mode == 1 else None
    for start, keep in ((0, range(0, 3)), (2, range(3, 6))):
        y = sinos[start:start + 4]
        if mode == 1:
            r = O.fista_execution(g, ang, L, x0, y, 2, 1.0, lam)
        else:
            r = O.fista_osem_execution(g, ang, x0, y, 2, lam)
        for z in keep:
            want[z] = r[z - start]
    e = relerr(got, want)
    print(f"[parity] CLI compressed sensing mode {mode}: relerr {e:.3e}")
    assert e < 1e-4


@pytest.mark.parametrize("cs", [False, True])
def test_hybrid_mode_8(tmp_path, cs):
    nz = 3
    extra = [("#再構成法", "8"), ("#サブセット(OSEM)", "6"), ("#反復回数", "2"), ("#ゼロパディングサイズ(FBP)", "128"),
             ("#Base up 減少速度次数 (Hybrid)", "4")]
    if cs:
        extra += [("#逐次圧縮センシング計算 (FBP適応外)", "1"), ("#逐次圧縮センシング計算加算ファクター", "0.002"),
                  ("#Number of Parallel processing per one device", "8")]
    inp, sinos, ang = dataset(tmp_path, nz, extra)
    run_cli(inp, f"hybrid cs={cs}")
    g = O.geom(N, N, PA, ss=6, zp=128)
    want = O.hybrid_execution(g, ang, sinos, 2, baseup_order=4, cs=cs, cs_lambda=0.002)
    e = relerr(outputs(tmp_path, nz), want)
    print(f"[parity] CLI hybrid mode 8 cs={cs}: relerr {e:.3e}")
    assert e < 1e-4
