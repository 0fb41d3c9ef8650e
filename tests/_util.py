"""Shared helpers for the parity tests."""
import sys

import numpy as np


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def say(*args):
    """Names the case before a GPU step (flushed, so a crash still shows it)."""
    print("[case]", *args, flush=True)
    sys.stdout.flush()


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def random_image(rng, nz, n, lo=0.1, hi=1.0):
    return rng.uniform(lo, hi, (nz, n, n)).astype(np.float32)


def phantom_stack(nz, n, pa, n_ell, seed, i0=None, px=None):
    from ctrecon import phantom as ph
    imgs, sinos = [], []
    ang = None
    for z in range(nz):
        img, sino, ang = ph.make_slice(n, pa, n_ell, seed, px=px, i0=i0, slice_index=z)
        imgs.append(img)
        sinos.append(sino)
    return np.stack(imgs), np.stack(sinos), ang
