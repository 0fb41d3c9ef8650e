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
