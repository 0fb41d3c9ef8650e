"""The uneven CT-XAFS stack (energies x slices of sides 256..1024, 1200
projections, 10 subsets) through one batched schedule: every slice equals its
own OSEM_execution (CT_reconstruction/GPU.cpp:125-240) bit for bit, and the
first energy of each side matches the oracle's FP32 restatement."""
import numpy as np
import pytest

import oracle as O
from _util import relerr, say

pytestmark = pytest.mark.gpu


def test_mixed_stack_reduced(ctx):
    import torch
    from ctrecon import phantom as ph
    E, pa, ss, it = 2, 1200, 10, 2
    sides_s = list(ph.C3_SIDES)
    say("mixed stack generation", sides_s, E)
    _, ang, sinos = ph.mixed_energy_stack_torch(E, len(sides_s), pa, seed=99, slice_sides=sides_s)
    sides = [sides_s[i // E] for i in range(len(sinos))]
    ctx.set_geometry(sides[0], sides[0], pa, ss, angles_deg=ang)
    for m in sides_s:
        ctx.set_schedule(m, 0, 4, 8)   # one schedule for every side: batched == per-slice bit for bit
    imgs = [torch.ones((m, m), device="cuda") for m in sides]
    say("batched mixed stack", len(imgs), "slices")
    ctx.osem_batched(imgs, sinos, it, use_graph=True)
    ctx.synchronize()
    for i, m in enumerate(sides):
        ctx.set_geometry(m, m, pa, ss, angles_deg=ang)
        x = torch.ones((1, m, m), device="cuda")
        say("per-slice", i, m)
        ctx.osem_execution(x, sinos[i].view(1, pa, m).clone(), it)
        ctx.synchronize()
        assert torch.equal(x[0], imgs[i]), (i, m)
        if i % E == 0:
            want = O.osem_execution(O.geom(m, m, pa, ss=ss), ang, np.ones((1, m, m), np.float32),
                                    sinos[i].cpu().numpy()[None], it)
            e = relerr(imgs[i].cpu().numpy()[None], want)
            print(f"[parity] mixed stack side {m}: relerr vs FP32 compat {e:.3e}", flush=True)
            assert e < 1e-5


def test_multi_row_grids(ctx):
    """40 slices of 1024^2: the forward projection (76800 blocks), padding and
    elementwise grids pass 2^16 blocks and launch as 2-D grids; every slice
    still equals its own OSEM_execution."""
    import torch
    from ctrecon import phantom as ph
    n, pa, ss, it, B = 1024, 1200, 10, 1, 40
    _, ang, sinos = ph.mixed_energy_stack_torch(1, 1, pa, seed=5, slice_sides=[n])
    ctx.set_geometry(n, n, pa, ss, angles_deg=ang)
    ctx.set_schedule(n, 0, 4, 8)
    x = torch.ones((1, n, n), device="cuda")
    say("per-slice 1024")
    ctx.osem_execution(x, sinos[0].view(1, pa, n).clone(), it)
    imgs = [torch.ones((n, n), device="cuda") for _ in range(B)]
    say("batched", B, "x 1024")
    ctx.osem_batched(imgs, [sinos[0]] * B, it, use_graph=False)
    ctx.synchronize()
    for i in range(B):
        assert torch.equal(imgs[i], x[0]), i


def test_mixed_stack_per_side_schedules(ctx):
    """Each side runs its own schedule inside one batched stack (LDS tiles of
    32, 44 and 60 samples, the direct-gather projector, 4 to 16 backprojection
    rows): every slice still equals its own OSEM_execution under that
    schedule bit for bit, and the oracle's FP32 restatement to 1e-5."""
    import torch
    from ctrecon import phantom as ph
    pa, ss, it = 1200, 10, 2
    sched = {256: (0, 4, 4, 32), 384: (0, 4, 16, 44), 512: (0, 4, 8, 60), 768: (64, 2, 8, 60)}
    sides_s = sorted(sched)
    say("mixed stack with per-side schedules", sides_s)
    _, ang, sinos = ph.mixed_energy_stack_torch(1, len(sides_s), pa, seed=98, slice_sides=sides_s)
    ctx.set_geometry(sides_s[0], sides_s[0], pa, ss, angles_deg=ang)
    for m, (rays, waves, rows, tile) in sched.items():
        ctx.set_schedule(m, rays, waves, rows, fp_tile=tile)
        assert ctx.get_schedule(m) == {"fp_rays": rays, "fp_waves": waves, "fp_tile": tile, "bp_rows": rows}
    imgs = [torch.ones((m, m), device="cuda") for m in sides_s]
    ctx.osem_batched(imgs, sinos, it, use_graph=True)
    ctx.synchronize()
    for i, m in enumerate(sides_s):
        ctx.set_geometry(m, m, pa, ss, angles_deg=ang)
        x = torch.ones((1, m, m), device="cuda")
        say("per-slice", m, sched[m])
        ctx.osem_execution(x, sinos[i].view(1, pa, m).clone(), it)
        ctx.synchronize()
        assert torch.equal(x[0], imgs[i]), m
        want = O.osem_execution(O.geom(m, m, pa, ss=ss), ang, np.ones((1, m, m), np.float32),
                                sinos[i].cpu().numpy()[None], it)
        e = relerr(imgs[i].cpu().numpy()[None], want)
        print(f"[parity] per-side schedule {m} {sched[m]}: relerr vs FP32 compat {e:.3e}", flush=True)
        assert e < 1e-5


def test_batched_stack_past_the_lds_projector_side(ctx):
    """A side above 3840 (one LDS tile of 60 samples per lane no longer
    covers a ray) runs the direct-gather projector by default, in its own
    slice set beside the LDS-tiled sides: every slice still equals its own
    OSEM_execution bit for bit."""
    import torch
    from ctrecon import phantom as ph
    pa, ss, it = 12, 3, 1
    sides_s = [64, 3900, 200]
    say("batched stack with a 3900^2 slice")
    _, ang, sinos = ph.mixed_energy_stack_torch(1, len(sides_s), pa, seed=97, slice_sides=sides_s)
    ctx.set_geometry(sides_s[0], sides_s[0], pa, ss, angles_deg=ang)
    for m in sides_s:
        if m <= 3840:
            ctx.set_schedule(m, 0, 4, 8)
        else:
            with pytest.raises(Exception):
                ctx.set_schedule(m, 0, 4, 8)  # no LDS tile covers the ray: refused
    assert ctx.get_schedule(3900)["fp_rays"] > 0      # the default there gathers directly
    imgs = [torch.ones((m, m), device="cuda") for m in sides_s]
    ctx.osem_batched(imgs, sinos, it, use_graph=True)
    ctx.synchronize()
    for i, m in enumerate(sides_s):
        ctx.set_geometry(m, m, pa, ss, angles_deg=ang)
        x = torch.ones((1, m, m), device="cuda")
        say("per-slice", m)
        ctx.osem_execution(x, sinos[i].view(1, pa, m).clone(), it)
        ctx.synchronize()
        assert torch.equal(x[0], imgs[i]), m
        assert torch.isfinite(imgs[i]).all(), m
