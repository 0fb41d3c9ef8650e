"""Benchmark: batched OS-EM slice-iterations/s on one MI355X.

Default workload: a batch of B slices of 1024 x 1024, 1200 projections over
[0, 180), 1024 bins, 10 ordered subsets (the 1024^2 / 1200-projection case of
the batched OS-EM stack).  A step is one OS-EM iteration (all 10 subsets) over
the whole batch through ctr_osem_batched, with the sinograms and images
already resident in HBM; value = B * steps / time.

--workload reslice measures sinogram formation instead (Reslice/reslice.cl):
A mu-t projections of an H x W camera (default the reference's 2048 x 2048
camera and 1600 projections) resliced into H sinograms [A][W] with
per-projection fractional x / y shifts; value = algorithmic GB/s
(4 bytes read + 4 written per sinogram value).

--workload joint runs the batched joint image + alignment reconstruction
(ctrecon.joint, FISTA-TV with per-projection detector offsets) on 96 uneven
slices with sides from {512, ..., 2048}, 1800 projections, shifts N(0, 2) px
and a centre offset U(-3, 3), each slice reconstructed on a grid 1.45x its
detector width (the phantoms reach beyond the detector's square); a step is
one joint iteration over all slices; value = iterations/s after the set-up
(the power-iteration step sizes).

--workload fbp times plain FBP of one 256 x 256 slice, 360 projections, 256
bins, ZP = 512, a noiseless 10-ellipse phantom; a step is one FBP_execution
call; value = slices/s.  --workload osem times OS-EM of one 512 x 512 slice,
900 projections, 10 subsets, 20 iterations, a 30-ellipse phantom with
Poisson noise at I0 = 1e4; a step is one 20-iteration OSEM_execution call;
value = slice-iterations/s.  Both report the per-slice-launch form (a
synchronisation after every kernel), parity against the oracle and the
oracle's time on the host.  --workload fista runs the differentiable
FISTA-TV with alignment (ctrecon.joint) on one 512 x 512 slice, 720
projections, TV weight 1e-3, per-projection shifts N(0, 1.5) px and a
centre offset U(-3, 3) in a noiseless sinogram, reconstructed on a 742 x
742 grid; value = iterations/s over 100 iterations.

--workload mixed runs the uneven CT-XAFS stack instead: E energies x S slices,
each slice's side seeded from {256, 384, 512, 768, 1024} and shared by its
energies, attenuation with a sigmoid absorption edge over the energies.
value is then in 1024^2-equivalent slice-iterations/s,
sum(n^2 / 1024^2) * steps / time.

Printed beside it:
  per_slice_launch   the same kernels launched one slice at a time with a
                     synchronisation after every kernel, as GPU.cpp does
                     (finish() after each enqueue, GPU.cpp:216-227)
  relerr_*           slice 0 after all the run's iterations (relerr_iterations)
                     against the CPU oracle's FP32 compat restatement and its
                     FP64 version, from the same start
  roofline           the forward-projection kernel (the dominant one): its
                     algorithmic FLOPs per launch over its mean launch time,
                     measured with HIP events on the library stream
  cpu_baseline       the oracle's FP32 restatement (OpenMP, all host threads)
                     timed on one slice-iteration of a 1024^2 slice: 1 warm-up
                     and 5 timed runs, the median, with the thread count,
                     nproc and the CPU model
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ct-xafs_analysis_amd"))


@contextlib.contextmanager
def no_gc():
    """A timed loop without Python's cyclic garbage collector: one collection
    before it and none inside, as timeit does.  With torch loaded a full
    collection stalls the host for tens of milliseconds, which a loop of
    asynchronous calls would count as device time (one such pause in 50 FBP
    calls of 0.08 ms made them read 0.87 ms)."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()

METRIC = "batched OS-EM slice-iters/s vs per-slice launch, 1024² 1200 proj, relerr≤1e-4"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# LDS read rate of the device: conflict-free ds_read2_b32, 275 bytes per ns
# per CU measured on MI355X (128 bytes per clock), over its 256 CUs.  The
# projector reads 16 bytes of LDS per bilinear sample (two ds_read2_b32 of
# two texels each).
LDS_READ_TBS = 275.0 * 256 / 1e3
FP_LDS_BYTES_PER_SAMPLE = 16.0
# Bytes moved past L2 per forward-projection launch, per slice of 1024^2 with
# 120 angles (one subset of 1200): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on
# the 128-slice k_fp_lds<0> launch, scaled by the same counters on a plain
# 256 MiB copy (FETCH x 1.99969, WRITE x 1.0), divided by 128
# (profiles/k_fp_lds_k_bp_osem_1024sq_1200proj_128slices.md).
FP_TRAFFIC_BYTES_PER_SLICE = 44102141.0
FP_TRAFFIC_SHAPE = (1024, 1200, 10)
# Bytes moved past L2 per k_reslice launch for 1600 projections of 2048 x 2048
# (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled the same way; the 9 rows a
# block reads for its 8 output rows make the fetch 9/8 of the input)
# (profiles/k_reslice_2048sq_1600proj.md).
RS_TRAFFIC_BYTES = 57097926552.0
RS_TRAFFIC_SHAPE = (1600, 2048, 2048)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_batch(B, n, pa, seed, device):
    import numpy as np

    from ctrecon import phantom as ph
    ang = ph.uniform_angles(pa)
    ells = [ph.random_ellipses(10, np.random.default_rng(ph.stream_seed(seed, i))) for i in range(B)]
    sinos = []
    chunk = 16
    for c0 in range(0, B, chunk):
        p = ph.line_integrals_torch(ells[c0:c0 + chunk], n, ang, n, device=device)
        y = ph.transmission_noise_torch(p, 1.0e4, ph.stream_seed(seed, c0, 7))
        sinos.extend(y.float().contiguous().unbind(0))
        del p, y
    sinos = [s.contiguous() for s in sinos]
    return ang, sinos


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pa, ss, ang, n_time, sino_time, reps=5):
    """The oracle's FP32 restatement (OpenMP, all host threads) timed on one
    OS-EM slice-iteration of an n_time^2 slice: 1 warm-up run, then `reps`
    timed runs; the median is reported."""
    import statistics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    gt = O.geom(n_time, n_time, pa, ss=ss)
    xt = np.ones((1, n_time, n_time), np.float32)
    O.osem_execution(gt, ang, xt.copy(), sino_time[None].copy(), 1)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.osem_execution(gt, ang, xt.copy(), sino_time[None].copy(), 1)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {
        "value": 1.0 / med, "unit": "slice-iters/s", "cores": O.num_threads(), "nproc": os.cpu_count(),
        "cpu_model": cpu_model(), "kind": "port",
        "sample": f"1 OS-EM slice-iteration ({ss} subsets) of one {n_time}^2 x {pa}-angle slice, FP32 compat "
                  f"restatement of OSEM1/OSEM2 in C with OpenMP; 1 warm-up + {reps} runs, median {med:.3f} s "
                  f"(runs {', '.join(f'{t:.3f}' for t in times)} s)",
    }


def parity_vs_oracle(n, pa, ss, ang, sino0, gpu_slice0, iters):
    """Slice 0 after the run's `iters` OS-EM iterations against the oracle's
    FP32 compat restatement and its FP64 version, from the same start."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    x0 = np.ones((1, n, n), np.float32)
    g = O.geom(n, n, pa, ss=ss)
    c32 = O.osem_execution(g, ang, x0.copy(), sino0[None].copy(), iters)
    c64 = O.osem_execution(g, ang, x0.copy(), sino0[None].copy(), iters, dtype=np.float64)
    rel = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))  # noqa: E731
    return rel(gpu_slice0, c32), rel(gpu_slice0, c64), rel(c32, c64)


def reslice_main(args):
    """Sinogram formation: ctr_reslice over A projections of an n x n camera."""
    import numpy as np
    import torch

    from ctrecon import Context
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    A, H, W = args.pa, args.n, args.n
    dev = torch.device("cuda:0")
    log(f"generating {A} mu-t projections of {H}x{W}")
    g = torch.Generator(device=dev).manual_seed(2026)
    mt = torch.rand((A, H, W), device=dev, generator=g).mul_(2.0)
    xs = (torch.rand((A,), device=dev, generator=g) - 0.5).mul_(4.0)
    ys = (torch.rand((A,), device=dev, generator=g) - 0.5).mul_(4.0)
    prj = torch.empty((H, A, W), device=dev)
    ctx = Context(0)
    st = ctx.library_stream()
    ctx.reslice(mt, xs, ys, baseup=0.1, out=prj)
    ctx.synchronize()
    # kernel time with HIP events on the library stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nk = 3
    e0.record(st)
    for _ in range(nk):
        ctx.reslice(mt, xs, ys, baseup=0.1, out=prj)
    e1.record(st)
    ctx.synchronize()
    k_ms = e0.elapsed_time(e1) / nk
    copy_ms = None
    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib * 262144, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        dst.copy_(src)
        c1.record()
        torch.cuda.synchronize()
        copy_ms = c0.elapsed_time(c1)
        log(f"reference copy of {args.copy_mib} MiB: {copy_ms:.3f} ms, "
            f"{2 * args.copy_mib * 1048576 / (copy_ms * 1e-3) / 1e9:.0f} GB/s read + write")
        del src, dst
    for _ in range(args.warmup):
        ctx.reslice(mt, xs, ys, baseup=0.1, out=prj)
    ctx.synchronize()
    torch.cuda.synchronize()
    with no_gc():
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.reslice(mt, xs, ys, baseup=0.1, out=prj)
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    nbytes = 8.0 * A * H * W
    value = nbytes * args.steps / dt / 1e9
    achieved = nbytes / (k_ms * 1e-3) / 1e9
    log(f"reslice: {value:.1f} GB/s ({dt / args.steps * 1e3:.2f} ms per step, kernel {k_ms:.2f} ms)")
    out = {
        "metric": f"reslice (sinogram formation) throughput, {H}x{W} camera, {A} projections",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: uniform random mu-t projections in [0, 2), per-projection x / y shifts in [-2, 2)",
        "config": {"workload": f"reslice of {A} projections {H}x{W} into {H} sinograms [{A}][{W}] with fractional "
                               f"per-projection shifts and a base-up; step = one ctr_reslice over the stack",
                   "A": A, "H": H, "W": W},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": RS_TRAFFIC_BYTES if (A, H, W) == RS_TRAFFIC_SHAPE else None,
                     "algorithmic_bytes": nbytes,
                     "kernel": "k_reslice", "ms_per_launch": round(k_ms, 4)},
    }
    if copy_ms is not None:
        out["device_copy"] = {"mib": args.copy_mib, "ms": round(copy_ms, 4),
                              "gbs_read_plus_write": round(2 * args.copy_mib * 1048576 / (copy_ms * 1e-3) / 1e9, 1)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        na = 8
        log(f"CPU baseline / oracle check on {na} projections ...")
        mt_s = mt[:na].cpu().numpy()
        xs_s, ys_s = xs[:na].cpu().numpy(), ys[:na].cpu().numpy()
        t1 = time.perf_counter()
        want = O.reslice(mt_s, xs_s, ys_s, baseup=0.1)
        tc = time.perf_counter() - t1
        got = prj[:, :na].cpu().numpy()
        out["bitexact_vs_oracle"] = bool(np.array_equal(got, want))
        out["cpu_baseline"] = {"value": round(8.0 * na * H * W / tc / 1e9, 3), "unit": "GB/s",
                               "cores": O.num_threads(), "kind": "port",
                               "sample": f"{na} of the {A} projections through the oracle's reslice (OpenMP)"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    ctx.close()


def slice_main(args):
    """One slice through one reference driver: FBP (256^2, 360 angles) or
    OS-EM (512^2, 900 angles, 10 subsets, 20 iterations)."""
    import statistics

    import torch

    from ctrecon import Context
    from ctrecon import phantom as ph
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    fbp = args.workload == "fbp"
    if fbp:
        n, pa, ss, it, zp = 256, 360, 1, 1, 512
        _, sino, ang = ph.make_slice(n, pa, 10, 1)
    else:
        n, pa, ss, it, zp = 512, 900, 10, 20, None
        _, sino, ang = ph.make_slice(n, pa, 30, 2, i0=1e4)
    y = torch.from_numpy(sino[None].copy()).cuda()
    x0 = np.zeros((1, n, n), np.float32) if fbp else np.ones((1, n, n), np.float32)

    def run(ctx, x):
        if fbp:
            ctx.fbp_execution(x, y.clone())
        else:
            ctx.osem_execution(x, y.clone(), it)

    def timed(sync_each):
        ctx = Context(0)
        ctx.set_geometry(n, n, pa, ss, zp=zp, angles_deg=ang)
        ctx.set_sync_each_kernel(sync_each)
        xs = [torch.from_numpy(x0.copy()).cuda() for _ in range(args.warmup + args.steps)]
        for k in range(args.warmup):
            run(ctx, xs[k])
        ctx.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            for k in range(args.steps):
                run(ctx, xs[args.warmup + k])
            ctx.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
        out = xs[-1].cpu().numpy()
        ctx.close()
        return dt, out

    dt, got = timed(False)
    dt_each, _ = timed(True)
    g = O.geom(n, n, pa, ss=ss, zp=zp)
    if fbp:
        c32, c64 = O.fbp_execution(g, ang, sino[None]), O.fbp_execution(g, ang, sino[None], dtype=np.float64)
    else:
        c32 = O.osem_execution(g, ang, x0.copy(), sino[None].copy(), it)
        c64 = O.osem_execution(g, ang, x0.copy(), sino[None].copy(), it, dtype=np.float64)
    rel = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))  # noqa: E731
    times = []
    for _ in range(3):
        t1 = time.perf_counter()
        if fbp:
            O.fbp_execution(g, ang, sino[None])
        else:
            O.osem_execution(g, ang, x0.copy(), sino[None].copy(), it)
        times.append(time.perf_counter() - t1)
    t_cpu = statistics.median(times)
    unit = "slices/s" if fbp else "slice-iters/s"
    per = 1 if fbp else it
    if fbp:
        workload = f"FBP_execution of one {n}x{n} slice, {pa} projections over [0,180), {n} bins, ZP={zp}, Ram-Lak"
        data = "synthetic: seeded noiseless 10-ellipse phantom, analytic sinogram"
    else:
        workload = (f"OSEM_execution of one {n}x{n} slice, {pa} projections over [0,180), {n} bins, {ss} subsets, "
                    f"{it} iterations from a uniform 1.0 image")
        data = "synthetic: seeded 30-ellipse phantom, analytic sinogram with Poisson noise at I0=1e4"
    out = {
        "metric": f"{'FBP' if fbp else 'OS-EM'} of one slice ({n}^2, {pa} projections)", "value": round(per / dt, 3),
        "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": data,
        "config": {"workload": workload + "; step = one driver call", "n": n, "pa": pa, "ss": ss, "iterations": it},
        "per_slice_launch": {"value": round(per / dt_each, 3), "unit": unit, "ms_per_step": round(dt_each * 1e3, 3)},
        "relerr_vs_fp32_compat": rel(got, c32), "relerr_vs_fp64": rel(got, c64),
        "relerr_fp32_compat_vs_fp64": rel(c32, c64),
        "roofline": None,
        "cpu_baseline": {"value": round(per / t_cpu, 3), "unit": unit, "cores": O.num_threads(),
                         "nproc": os.cpu_count(), "cpu_model": cpu_model(), "kind": "port",
                         "sample": f"the same call in the oracle's FP32 compat restatement (C, OpenMP): 3 runs after "
                                   f"the parity run, median {t_cpu * 1e3:.1f} ms"},
    }
    print(json.dumps(out))


def fista_main(args):
    """Differentiable FISTA-TV with alignment on one 512^2 slice, 720
    projections (ctrecon.joint)."""
    import torch

    from ctrecon import Context
    from ctrecon import phantom as ph
    from ctrecon.joint import joint_fista_tv, joint_grid
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n, pa, iters = 512, 720, 100
    m = joint_grid(n, args.grid_pad)          # the phantom reaches beyond the detector's square
    rng = np.random.default_rng(2027)
    shifts = rng.normal(0.0, 1.5, pa)
    center = rng.uniform(-3.0, 3.0)
    _, sino, ang = ph.make_slice(n, pa, 10, 4, shifts=shifts, center=center)
    ctx = Context(0)
    ctx.set_geometry(m, n, pa, 1, ny=m, angles_deg=ang)
    y = torch.from_numpy(sino[None].copy()).cuda()
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    marks = [torch.cuda.Event(enable_timing=True)]
    joint_fista_tv(ctx, y, iters=12, tv=1e-3)         # warm-up: every kernel loaded (offsets move from 11)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks[0].record()

    def cb(k, x, off, f):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        marks.append(ev)

    x, s, c, hist = joint_fista_tv(ctx, y, iters=iters, tv=1e-3, callback=cb)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    iter_s = marks[1].elapsed_time(marks[-1]) / 1e3 / (len(marks) - 2)
    e0, e1 = off_err(np.zeros(pa)), off_err((s + c).cpu().numpy().astype(np.float64))
    log(f"FISTA-TV {n}^2 x {pa}: {iter_s * 1e3:.2f} ms per iteration, offset error {e0:.3f} -> {e1:.3f} px")
    out = {
        "metric": f"differentiable FISTA-TV iterations/s, one {n}^2 slice, {pa} projections",
        "value": round(1.0 / iter_s, 3), "unit": "iterations/s", "n_gpus": 1, "steps": iters, "warmup": 12,
        "ms_per_step": round(iter_s * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded noiseless 10-ellipse phantom, analytic sinogram with per-projection shifts "
                "N(0, 1.5) px and a centre offset U(-3, 3)",
        "config": {"workload": f"joint FISTA-TV (TV 1e-3) + detector-offset Gauss-Newton, one {n}x{n} slice, {pa} "
                               f"projections ({n} bins), reconstructed on a {m}x{m} grid"
                               + (" that holds the whole phantom" if m >= 1.4 * n else "")
                               + f", {iters} iterations; step = one iteration",
                   "grid": m, "total_s_with_setup": round(total, 3)},
        "offset_error_px": {"start": round(e0, 4), "end": round(e1, 4), "modulo": "image translation"},
        "data_term": {"first": hist[0], "last": hist[-1]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def joint_main(args):
    """Batched joint alignment over uneven slices (ctrecon.joint)."""
    import numpy as np
    import torch

    from ctrecon import Context
    from ctrecon import phantom as ph
    from ctrecon.joint import joint_fista_tv_groups, joint_grid
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    pa, nsl = args.pa, 96
    rng = np.random.default_rng(2026)
    shifts = rng.normal(0.0, 2.0, pa)
    center = rng.uniform(-3.0, 3.0)
    sides = rng.choice((512, 768, 1024, 1280, 1536, 1792, 2048), nsl)
    t0 = time.perf_counter()
    groups, counts = [], {}
    for n in sorted(set(sides.tolist())):
        cnt = int((sides == n).sum())
        counts[n] = cnt
        ys = []
        for k in range(cnt):
            _, sino, ang = ph.make_slice(n, pa, 10, 900 + n, shifts=shifts, center=center, i0=5e3, slice_index=k)
            ys.append(sino)
        c = Context(0)
        # a grid that holds the whole phantom (it reaches beyond the square
        # the detector width spans)
        m = joint_grid(n, args.grid_pad)
        c.set_geometry(m, n, pa, 1, ny=m, angles_deg=ang)
        groups.append((c, torch.from_numpy(np.stack(ys)).cuda()))
    torch.cuda.synchronize()
    log(f"joint: {nsl} slices {counts}, {pa} projections generated in {time.perf_counter() - t0:.1f} s")
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        # offsets are known up to an image translation: span{cos, sin}
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    # one event per iteration on torch's stream, which every group's work
    # joins; the host waits only every tenth iteration, to log
    marks = [torch.cuda.Event(enable_timing=True)]
    marks[0].record()

    def cb(k, xs, off, f):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        marks.append(ev)
        if k % 10 == 9 or k == 0:
            log(f"joint iteration {k + 1}/{args.steps}: offset error {off_err(off.cpu().numpy().astype(np.float64)):.3f} "
                f"px, data {float(f):.4e}")

    t1 = time.perf_counter()
    xs, s, c, hist = joint_fista_tv_groups(groups, iters=args.steps, tv=1e-3, warmup=10, callback=cb)
    torch.cuda.synchronize()
    total = time.perf_counter() - t1
    # iterations after the first (which includes the first calls' set-up)
    first = 1 if len(marks) > 2 else 0
    iter_s = marks[first].elapsed_time(marks[-1]) / 1e3 / (len(marks) - 1 - first)
    e0, e1 = off_err(np.zeros(pa)), off_err((s + c).cpu().numpy().astype(np.float64))
    out = {
        "metric": f"batched joint alignment iterations/s, {nsl} uneven slices 512-2048^2, {pa} projections",
        "value": round(1.0 / iter_s, 4), "unit": "iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": round(iter_s * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 10-ellipse phantoms, analytic sinograms with per-projection shifts N(0, 2) px, "
                "a centre offset U(-3, 3) and Poisson noise at I0=5e3",
        "config": {"workload": "joint FISTA-TV + detector-offset Gauss-Newton over uneven slices (one context per "
                               f"side, each slice reconstructed on a grid {args.grid_pad}x its detector width"
                               + (", which holds the whole phantom" if args.grid_pad >= 1.4 else "")
                               + "); step = one iteration over all slices", "slices_per_side": counts,
                   "grid_pad": args.grid_pad, "pa": pa, "iterations": args.steps,
                   "total_s_with_setup": round(total, 2)},
        "offset_error_px": {"start": round(e0, 4), "end": round(e1, 4), "modulo": "image translation"},
        "data_term": {"first": hist[0], "last": hist[-1]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("uniform", "mixed", "reslice", "joint", "fbp", "osem", "fista"),
                    default="uniform")
    ap.add_argument("--batch", type=int, default=128, help="uniform: slices in the batch")
    ap.add_argument("--n", type=int, default=None, help="uniform: slice side (1024); reslice: camera side (2048)")
    ap.add_argument("--energies", type=int, default=64, help="mixed: energies per slice")
    ap.add_argument("--slices", type=int, default=128, help="mixed: slice positions")
    ap.add_argument("--pa", type=int, default=None, help="projections (1200; reslice: 1600)")
    ap.add_argument("--ss", type=int, default=10)
    ap.add_argument("--per-slice", type=int, default=4, help="uniform: slices timed in the per-slice-launch baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / oracle check")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--grid-pad", type=float, default=1.45,
                    help="joint / fista: reconstruction grid side over the detector width (1.0: the detector's square)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain device copy of this many MiB (reference size for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); every box here has one GPU.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "reslice":
        args.n = 2048 if args.n is None else args.n
        args.pa = 1600 if args.pa is None else args.pa
        return reslice_main(args)
    if args.workload == "joint":
        args.pa = 1800 if args.pa is None else args.pa
        return joint_main(args)
    if args.workload in ("fbp", "osem"):
        return slice_main(args)
    if args.workload == "fista":
        return fista_main(args)
    args.n = 1024 if args.n is None else args.n
    args.pa = 1200 if args.pa is None else args.pa

    import numpy as np
    import torch

    from ctrecon import Context
    from ctrecon import phantom as ph
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    pa, ss = args.pa, args.ss
    na = pa // ss
    dev = torch.device("cuda:0")
    if args.workload == "uniform":
        log(f"generating {args.batch} phantoms {args.n}^2 x {pa}")
        ang, sinos = make_batch(args.batch, args.n, pa, seed=2026, device=dev)
        sides = [args.n] * args.batch
    else:
        E, S = args.energies, args.slices
        log(f"generating the mixed stack: {E} energies x {S} slices, sides {ph.C3_SIDES}, {pa} angles")
        sides_s, ang, sinos = ph.mixed_energy_stack_torch(E, S, pa, seed=2026, device=dev)
        sides = [sides_s[i // E] for i in range(E * S)]
    B = len(sides)
    imgs = [torch.ones((m, m), dtype=torch.float32, device=dev) for m in sides]
    torch.cuda.synchronize()
    eq = float(sum(m * m for m in sides)) / 1024.0 ** 2        # 1024^2-equivalent slices

    ctx = Context(0)
    ctx.set_geometry(sides[0], sides[0], pa, ss, angles_deg=ang)
    tune, default_ms = None, None
    iters_done = 0
    if not args.no_tune:
        # the stack under the default schedule (LDS tiles of 60 samples, 8
        # backprojection rows): one iteration to build the plan, then as many
        # timed as the tuned run below
        ctx.osem_batched(imgs, sinos, 1, use_graph=True)
        ctx.synchronize()
        with no_gc():
            t0 = time.perf_counter()
            for _ in range(args.steps):
                ctx.osem_batched(imgs, sinos, 1, use_graph=True)
            ctx.synchronize()
            default_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        iters_done += 1 + args.steps
        log(f"default schedule: {default_ms:.1f} ms per step")
        # then every side's own schedule (each side of a batched stack runs its own)
        tune = {m: ctx.autotune(m) for m in sorted(set(sides))}
        for m, t in tune.items():
            log(f"autotune side {m}: rays {t['fp_rays']} waves {t['fp_waves']} tile {t['fp_tile']} rows "
                f"{t['bp_rows']}; FP {t['fp_ms']:.3f} ms (default {t['fp_ms_default']:.3f}), BP {t['bp_ms']:.3f} ms "
                f"(default {t['bp_ms_default']:.3f}) on a batch-sized launch")

    # first call with the tuned schedules: one iteration
    ctx.osem_batched(imgs, sinos, 1, use_graph=True)
    ctx.synchronize()
    iters_done += 1

    # kernel timing pass (eager launches, HIP events on the library stream)
    ctx.set_kernel_timing(True)
    ctx.osem_batched(imgs, sinos, 1, use_graph=False)
    ctx.synchronize()
    iters_done += 1
    fp_launch, bp_launch, nfp, nbp = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    # a subset is one FP and one BP launch per schedule set (slices of sides
    # with different autotuned schedules run as separate launches)
    fp_ms, bp_ms = fp_launch * nfp / ss, bp_launch * nbp / ss
    log(f"FP kernel {fp_launch:.3f} ms/launch, BP kernel {bp_launch:.3f} ms/launch ({nfp}, {nbp} launches); "
        f"per subset FP {fp_ms:.3f} ms, BP {bp_ms:.3f} ms")

    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib * 262144, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"reference copy of {args.copy_mib} MiB done")
        del src, dst

    for _ in range(args.warmup):
        ctx.osem_batched(imgs, sinos, 1, use_graph=True)
    ctx.synchronize()
    torch.cuda.synchronize()
    with no_gc():
        t0 = time.perf_counter()
        for s in range(args.steps):
            ctx.osem_batched(imgs, sinos, 1, use_graph=True)
            if args.workload == "mixed":
                ctx.synchronize()
                log(f"step {s + 1}/{args.steps} done")
        ctx.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    iters_done += args.warmup + args.steps
    gpu_slice0 = imgs[0].cpu().numpy()[None]        # after every iteration of the run
    value = eq * args.steps / dt
    log(f"batched: {value:.2f} (1024^2-equivalent) slice-iters/s ({dt / args.steps * 1e3:.1f} ms per step)")

    # per-slice launch baseline: own kernels, one slice per call, sync after every kernel
    pctx = Context(0)
    pctx.set_sync_each_kernel(True)
    if args.workload == "uniform":
        sample = [i % B for i in range(max(1, args.per_slice))]
    else:
        sample = [sides.index(m) for m in sorted(set(sides))]     # one slice of each side
    t_side = {}
    for m in sorted(set(sides[i] for i in sample)):
        pctx.set_geometry(m, m, pa, ss, angles_deg=ang)
        if tune:
            t = tune[m]
            pctx.set_schedule(m, t["fp_rays"], t["fp_waves"], t["bp_rows"], fp_tile=t["fp_tile"])
        for i in (i for i in sample if sides[i] == m):
            x = torch.ones((1, m, m), dtype=torch.float32, device=dev)
            y = sinos[i].view(1, pa, m)
            pctx.osem_execution(x.clone(), y.clone(), 1)          # warm
            pctx.synchronize()
            yc = y.clone()
            torch.cuda.synchronize()
            with no_gc():
                t1 = time.perf_counter()
                pctx.osem_execution(x, yc, 1)
                pctx.synchronize()
                t_side.setdefault(m, []).append(time.perf_counter() - t1)
    t_mean = {m: sum(v) / len(v) for m, v in t_side.items()}
    per_slice = eq / sum(t_mean[m] for m in sides)
    log(f"per-slice launch: {per_slice:.2f} (1024^2-equivalent) slice-iters/s")

    kinds = sorted({(tune[m]["fp_rays"], tune[m]["fp_tile"]) for m in set(sides)} if tune else {(0, 60)})
    fp_kernel = " + ".join(f"k_fp_lds<FP_RATIO, {tile}>" if rays == 0 else "k_fp<FP_RATIO>" for rays, tile in kinds)
    fp_tile60 = kinds == [(0, 60)]
    fp_flops = sum(14.0 * m * na * m for m in sides)          # 14 FLOP per bilinear sample, D = m samples per ray
    fp_bytes = sum(4 * m * m + 8 * na * m for m in sides)    # image once, measured rows in, ratio rows out
    fp_tflops = fp_flops / (fp_ms * 1e-3) / 1e12
    fp_lds_tbs = sum(FP_LDS_BYTES_PER_SAMPLE * m * na * m for m in sides) / (fp_ms * 1e-3) / 1e12
    step_flops = sum(14.0 * m * pa * m + 2.0 * m * m * pa for m in sides)
    hbm_bytes = sum(12.0 * ss * m * m + 12.0 * m * pa for m in sides)
    uniform_1024 = args.workload == "uniform" and (args.n, pa, ss) == FP_TRAFFIC_SHAPE
    roofline = {
        "bound": "valu", "achieved": round(fp_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": round(fp_tflops / PEAK_FP32_TFLOPS, 4),
        "traffic": (round(FP_TRAFFIC_BYTES_PER_SLICE * B) if uniform_1024 and fp_tile60 else None),
        "algorithmic_bytes": fp_bytes,
        "kernel": f"{fp_kernel} (OSEM1: forward projection + ratio), timed per subset of every slice",
        "fp_ms_per_subset": round(fp_ms, 4), "bp_ms_per_subset": round(bp_ms, 4),
        "launches_per_subset": nfp // ss,
        "whole_step_tflops": round(step_flops * args.steps / dt / 1e12, 3),
        "whole_step_hbm_frac": round(hbm_bytes * args.steps / dt / (PEAK_HBM_GBS * 1e9), 5),
        # the bound the projector meets first: its LDS reads against the
        # device's measured conflict-free read rate
        "lds": {"achieved": round(fp_lds_tbs, 3), "peak": LDS_READ_TBS, "unit": "TB/s",
                "frac": round(fp_lds_tbs / LDS_READ_TBS, 4)},
    }
    if args.workload == "uniform":
        workload = (f"batched OS-EM, {B} slices of {args.n}x{args.n}, {pa} projections over [0,180), {args.n} bins, "
                    f"{ss} subsets; step = one OS-EM iteration over the batch")
        cfg = {"workload": workload, "slices": B, "n": args.n, "pa": pa, "ss": ss, "autotune": tune}
        unit = "slice-iters/s"
        data = f"synthetic: {B} seeded 10-ellipse phantoms, analytic sinograms with Poisson noise at I0=1e4"
    else:
        counts = {m: sides.count(m) for m in sorted(set(sides))}
        workload = (f"batched OS-EM over a mixed CT-XAFS stack: {args.energies} energies x {args.slices} slices "
                    f"(slices per side {counts}), {pa} projections, {ss} subsets; step = one OS-EM iteration over "
                    f"the stack; value in 1024^2-equivalent slice-iterations/s")
        cfg = {"workload": workload, "energies": args.energies, "slices": args.slices, "side_counts": counts,
               "pa": pa, "ss": ss, "autotune": tune}
        unit = "1024^2-equivalent slice-iters/s"
        data = ("synthetic: seeded 10-ellipse phantoms per slice, sigmoid absorption edge over energies, analytic "
                "sinograms with Poisson noise at I0=1e4")
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": data, "config": cfg,
        "per_slice_launch": {"value": round(per_slice, 3), "unit": unit, "slices_timed": len(sample)},
        "batched_over_per_slice": round(value / per_slice, 3),
        "schedule_gain": (None if default_ms is None else
                          {"default_ms_per_step": round(default_ms, 3),
                           "tuned_ms_per_step": round(dt / args.steps * 1e3, 3),
                           "speedup": round(default_ms / (dt / args.steps * 1e3), 4)}),
        "roofline": roofline,
    }
    if not args.no_cpu:
        log(f"oracle check of slice 0 after the run's {iters_done} iterations ...")
        e32, e64, c_64 = parity_vs_oracle(sides[0], pa, ss, ang, sinos[0].cpu().numpy(), gpu_slice0, iters_done)
        log(f"relerr vs FP32 compat {e32:.3e}, vs FP64 {e64:.3e} (FP32 compat vs FP64 {c_64:.3e}); CPU baseline ...")
        i_time = sides.index(1024) if 1024 in sides else 0
        cb = cpu_baseline(pa, ss, ang, sides[i_time], sinos[i_time].cpu().numpy())
        if args.workload == "mixed":
            cb["value"] *= sides[i_time] ** 2 / 1024.0 ** 2
            cb["unit"] = unit
        out["cpu_baseline"] = cb
        out["relerr_vs_fp32_compat"] = e32
        out["relerr_vs_fp64"] = e64
        out["relerr_iterations"] = iters_done
        # the reference's FP32 arithmetic (the oracle's compat restatement) is
        # itself this far from FP64 after the same iterations
        out["relerr_fp32_compat_vs_fp64"] = c_64
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    ctx.close()
    pctx.close()


if __name__ == "__main__":
    main()
