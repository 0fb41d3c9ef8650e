"""The execution drivers are asynchronous on the context stream: a call
queues its uploads (through the pinned staging ring) and kernels and returns
before they complete, and two drivers queued back to back both give the
results they give when run one at a time with a synchronisation after every
kernel (GPU.cpp's finish() pattern)."""
import numpy as np
import pytest

from _util import host, say

pytestmark = pytest.mark.gpu


def test_two_drivers_queued_back_to_back():
    import torch
    from ctrecon import Context
    from ctrecon import phantom as ph
    n, pa, ss, nz = 1024, 1200, 10, 8
    dev_ = torch.device("cuda:0")
    ang = ph.uniform_angles(pa)
    ells = [ph.random_ellipses(10, np.random.default_rng(ph.stream_seed(5, i))) for i in range(2 * nz)]
    p = ph.line_integrals_torch(ells, n, ang, n, device=dev_)
    y = ph.transmission_noise_torch(p, 1.0e4, ph.stream_seed(5, 0, 7)).float().contiguous()
    ya, yb = y[:nz].contiguous(), y[nz:].contiguous()
    torch.cuda.synchronize()

    # one at a time, synchronising after every kernel
    ref = Context(0)
    ref.set_geometry(n, n, pa, ss, angles_deg=ang)
    ref.set_sync_each_kernel(True)
    xa_ref = torch.ones((nz, n, n), device=dev_)
    xb_ref = torch.ones((nz, n, n), device=dev_)
    ref.osem_execution(xa_ref, ya.clone(), 3)
    ref.aart_execution(xb_ref, yb.clone(), 1)
    ref.synchronize()

    ctx = Context(0)
    ctx.set_geometry(n, n, pa, ss, angles_deg=ang)
    xa = torch.ones((nz, n, n), device=dev_)
    xb = torch.ones((nz, n, n), device=dev_)
    ya2, yb2 = ya.clone(), yb.clone()
    torch.cuda.synchronize()
    ctx.synchronize()
    say("async: OS-EM then AART queued, 8 x 1024^2 x 1200")
    ctx.osem_execution(xa, ya2, 3)
    done_after_first = ctx.library_stream().query()
    ctx.aart_execution(xb, yb2, 1)
    done_after_second = ctx.library_stream().query()
    print(f"[async] stream idle after the first call: {done_after_first}, after the second: {done_after_second}",
          flush=True)
    assert not done_after_first and not done_after_second
    ctx.synchronize()
    assert torch.equal(xa, xa_ref) and torch.equal(xb, xb_ref)
    assert np.isfinite(host(xa)).all()
    ctx.close()
    ref.close()


def test_drivers_on_a_caller_stream():
    """ctr_set_stream: a context created on torch's current stream (a
    torch.cuda.Stream; the default stream's null handle is refused) and one
    given a separate torch stream give the images of a context on its own
    stream, bit for bit; resetting the stream restores the library's own."""
    import torch
    from ctrecon import Context
    from _util import phantom_stack
    n, pa, ss, nz = 128, 60, 6, 2
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    _, sino, _ = phantom_stack(nz, n, pa, 6, seed=31, i0=1e4)
    y = torch.from_numpy(sino).cuda()

    def run(cx):
        cx.set_geometry(n, n, pa, ss, angles_deg=ang)
        x = torch.ones((nz, n, n), device="cuda")
        cx.osem_execution(x, y.clone(), 2)
        cx.synchronize()
        torch.cuda.synchronize()
        return x

    own = Context(0)
    want = run(own)
    say("contexts on a caller's stream")
    with pytest.raises(ValueError):
        Context(0, use_torch_stream=True)     # under torch's default stream
    ts = torch.cuda.Stream()
    with torch.cuda.stream(ts):
        on_torch = Context(0, use_torch_stream=True)
        assert on_torch.library_stream().cuda_stream == ts.cuda_stream
        got_torch = run(on_torch)
    side = torch.cuda.Stream()
    on_side = Context(0)
    on_side.set_stream(side.cuda_stream)
    assert on_side.library_stream().cuda_stream == side.cuda_stream
    got_side = run(on_side)
    on_side.set_stream(None)                  # back to the context's own stream
    assert on_side.library_stream().cuda_stream not in (side.cuda_stream, torch.cuda.current_stream().cuda_stream)
    got_own_again = run(on_side)
    assert torch.equal(got_torch, want) and torch.equal(got_side, want) and torch.equal(got_own_again, want)
    for cx in (own, on_torch, on_side):
        cx.close()
