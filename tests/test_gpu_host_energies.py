"""End-to-end runs of the multi-energy drop-in (ct_reslice_reconstruction) on
the GPU: mu-t projection files of several energies in, one image file per
layer and energy out, each compared with the oracle: the projections'
camera row N resliced into a sinogram (Reslice_CTreconst_ocl.cpp:26-90,
layer N = row N, GPU.cpp:414-416) and reconstructed by the method's driver."""
import subprocess

import numpy as np
import pytest

import oracle as O
from _util import phantom_stack, relerr, say
from test_host_cli import write_inp
from test_host_energies import CLI_E

pytestmark = pytest.mark.gpu

N, PA, H = 48, 24, 10


def energy_dataset(tmp_path, subs, extra, scale=(1.0, 1.35, 0.8)):
    """Projections [A][H][N] per energy: the sinograms of H phantom layers,
    written one file per angle as the reference's mu-t images."""
    _, sinos, ang = phantom_stack(H, N, PA, 5, seed=21)
    stacks = {}
    for k, sub in enumerate(subs):
        s = (sinos * np.float32(scale[k % len(scale)])).astype(np.float32)
        stacks[sub] = s
        d = tmp_path / "mt" / sub
        d.mkdir(parents=True)
        mt = np.ascontiguousarray(s.transpose(1, 0, 2))          # [A][H][W]
        for a in range(PA):
            mt[a].tofile(d / f"mt_{a + 1:04d}.raw")
    (tmp_path / "log.dat").write_text("".join(f"{a:.9g}\n" for a in ang))
    keys = {"#Input directory path": str(tmp_path / "mt"), "#Output directory path": str(tmp_path / "out"),
            "#Output file base name": "rec_", "#Angle number range": f"1-{PA}", "#Image pixel size": f"{N},{H}",
            "#ログファイルパス": str(tmp_path / "log.dat"), "#カメラピクセル数": str(N), "#投影数": str(PA),
            "#再構成画像サイズ": str(N), "#開始ナンバー": "2", "#レイヤー数": "6"}
    keys.update(dict(extra))
    inp = tmp_path / "e.inp"
    write_inp(inp, list(keys.items()), crlf=True)
    return inp, stacks, ang


def run(inp, label, *flags):
    say("ct_reslice_reconstruction", label)
    r = subprocess.run([CLI_E, str(inp), *flags], capture_output=True, text=True, timeout=180)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def layer(tmp_path, sub, n):
    return np.fromfile(tmp_path / "out" / sub / f"rec_{n:04d}.raw", np.float32).reshape(N, N)


@pytest.mark.parametrize("mode,ss,corr", [(6, 6, 0), (6, 4, 2), (5, 6, 0), (2, 6, 1)])
def test_osem_family_over_energies(tmp_path, mode, ss, corr):
    subs = ["001", "002", "003"]
    inp, stacks, ang = energy_dataset(tmp_path, subs, [("#Energy number range", "1-3"), ("#再構成法", str(mode)),
                                                        ("#サブセット(OSEM)", str(ss)), ("#反復回数", "2"),
                                                        ("#投影像補正 0:なし 1: x方向のみ 2:θ方向のみ 3:x,θ方向両方",
                                                         str(corr))])
    run(inp, f"mode {mode} ss {ss} corr {corr}")
    g = O.geom(N, N, PA, ss={2: PA, 5: 1}.get(mode, ss))
    for sub in subs:
        rows = stacks[sub][2:7]                                  # layers 2..6 = camera rows 2..6
        want = O.osem_execution(g, ang, np.ones((5, N, N), np.float32), rows, 2, correction_mode=corr)
        got = np.stack([layer(tmp_path, sub, n) for n in range(2, 7)])
        e = relerr(got, want)
        print(f"[parity] multi-energy mode {mode} energy {sub}: relerr {e:.3e}", flush=True)
        assert e < 1e-5
        assert not (tmp_path / "out" / sub / "rec_0007.raw").exists()   # st + num - 1 is not run


def test_fbp_over_parameter_names(tmp_path):
    subs = ["a0", "a1"]
    inp, stacks, ang = energy_dataset(tmp_path, subs, [("#再構成法", "7"), ("#ゼロパディングサイズ(FBP)", "128")])
    run(inp, "FBP over names", "-fpn", "a0,a1")
    g = O.geom(N, N, PA, zp=128)
    for sub in subs:
        want = O.fbp_execution(g, ang, stacks[sub][2:7])
        got = np.stack([layer(tmp_path, sub, n) for n in range(2, 7)])
        assert relerr(got, want) < 1e-5


def test_fbp_output_region_over_energies(tmp_path):
    """FBP with an output region below the grid: each file is the top-left
    ox x oy block of the layer (FBP_thread, GPU.cpp:848-870)."""
    subs, ox, oy = ["001", "002"], 40, 28
    inp, stacks, ang = energy_dataset(tmp_path, subs, [("#Energy number range", "1-2"), ("#再構成法", "7"),
                                                        ("#ゼロパディングサイズ(FBP)", "128"),
                                                        ("#出力画像サイズ X", str(ox)), ("#出力画像サイズ Y", str(oy))])
    run(inp, "FBP output region")
    g = O.geom(N, N, PA, zp=128)
    for sub in subs:
        want = O.fbp_execution(g, ang, stacks[sub][2:7])[:, :oy, :ox]
        got = np.stack([np.fromfile(tmp_path / "out" / sub / f"rec_{n:04d}.raw", np.float32) for n in range(2, 7)])
        assert got.shape == (5, ox * oy)
        assert relerr(got.reshape(5, oy, ox), want) < 1e-5


def test_fista_osem_batches_and_skipped_layers(tmp_path):
    """FISTA-OS-EM in batches of numParallel = 3 layers without overlap:
    [2..4], [5..6]; OSEM_thread writes a batch's layers minus CSoverlap at
    each inner end (GPU.cpp:453-457), so only layers 2 and 3 are written."""
    subs = ["001"]
    inp, stacks, ang = energy_dataset(tmp_path, subs, [("#Energy number range", "1-1"), ("#サブセット(OSEM)", "6"),
                                                        ("#反復回数", "2"),
                                                        ("#逐次圧縮センシング計算 (FBP適応外)", "1"),
                                                        ("#逐次圧縮センシング計算加算ファクター", "0.002"),
                                                        ("#Number of Parallel processing per one device", "3")])
    run(inp, "FISTA-OS-EM batches")
    g = O.geom(N, N, PA, ss=6)
    want = O.fista_osem_execution(g, ang, np.ones((3, N, N), np.float32), stacks["001"][2:5], 2, 0.002)
    got = np.stack([layer(tmp_path, "001", n) for n in (2, 3)])
    e = relerr(got, want[:2])
    print(f"[parity] multi-energy FISTA-OS-EM: relerr {e:.3e}", flush=True)
    assert e < 1e-4
    for n in (4, 5, 6):
        assert not (tmp_path / "out" / "001" / f"rec_{n:04d}.raw").exists()
