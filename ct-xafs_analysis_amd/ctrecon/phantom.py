"""Seeded random-ellipse phantoms and their analytic parallel-beam sinograms.

Geometry follows the reference projector's convention
(CT_reconstruction/OSEM.cl:56-62, OSEM.cl:95): pixel (i, j) of an N x N image
has centre (x, y) = (i - N/2, j - N/2), detector bin X sits at s = X - P/2 and
a ray at angle theta is the line x cos(theta) - y sin(theta) = s.

Ellipses live on the unit disk (radius N/2 pixels).  Centres are uniform in
the disk, semi-axes U(0.05, 0.4), orientations U(0, pi) and attenuations
U(0.1, 1.0) per unit-disk radius, so images hold mu / (N/2) per pixel and
noiseless line integrals are O(1).  The analytic line integral of an ellipse
(Kak & Slaney, Principles of Computerized Tomographic Imaging, eq. 3.8) is
    p = 2 rho a b sqrt(r^2 - s'^2) / r^2,   r^2 = a^2 cos^2(t) + b^2 sin^2(t),
with t the angle between the ray normal and the major axis and s' the signed
distance of the ray from the ellipse centre.

Transmission noise mirrors mt_conversion
(ImageReg share source/imageregistration_kernel_src.cl:208-215): counts are
Poisson(I0 exp(-p)) and y = ln(I0 / counts) with the transmission clamped to
[1e-5, 1e7].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """One SplitMix64 step (Steele, Lea & Flood, OOPSLA 2014)."""
    x = (x + 0x9E3779B97F4A7C15) & _MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return z ^ (z >> 31)


def stream_seed(seed: int, slice_index: int = 0, energy_index: int = 0) -> int:
    """Per-(slice, energy) seed derived by chained SplitMix64."""
    s = splitmix64(seed & _MASK64)
    s = splitmix64(s ^ (slice_index & _MASK64))
    s = splitmix64(s ^ ((energy_index * 0x100000001B3) & _MASK64))
    return s


@dataclass
class Ellipses:
    """Ellipse parameters in unit-disk coordinates."""
    cx: np.ndarray
    cy: np.ndarray
    a: np.ndarray
    b: np.ndarray
    phi: np.ndarray
    mu: np.ndarray

    def __len__(self) -> int:
        return len(self.cx)

    def with_mu(self, mu: np.ndarray) -> "Ellipses":
        return Ellipses(self.cx, self.cy, self.a, self.b, self.phi, np.asarray(mu, dtype=np.float64))


def random_ellipses(n: int, rng: np.random.Generator) -> Ellipses:
    r = np.sqrt(rng.uniform(0.0, 1.0, n))
    t = rng.uniform(0.0, 2.0 * np.pi, n)
    return Ellipses(
        cx=r * np.cos(t), cy=r * np.sin(t),
        a=rng.uniform(0.05, 0.4, n), b=rng.uniform(0.05, 0.4, n),
        phi=rng.uniform(0.0, np.pi, n), mu=rng.uniform(0.1, 1.0, n))


def rasterize(ell: Ellipses, n: int, supersample: int = 1) -> np.ndarray:
    """Image [n][n] of mu per pixel (sum of overlapping ellipses)."""
    half = n / 2.0
    ss = supersample
    offs = (np.arange(ss) + 0.5) / ss - 0.5
    idx = np.arange(n, dtype=np.float64)
    img = np.zeros((n, n), dtype=np.float64)
    for oy in offs:
        for ox in offs:
            x = (idx[None, :] + ox - half) / half
            y = (idx[:, None] + oy - half) / half
            for k in range(len(ell)):
                c, s = np.cos(ell.phi[k]), np.sin(ell.phi[k])
                dx, dy = x - ell.cx[k], y - ell.cy[k]
                u = (dx * c + dy * s) / ell.a[k]
                v = (-dx * s + dy * c) / ell.b[k]
                img += np.where(u * u + v * v <= 1.0, ell.mu[k], 0.0)
    return img / (ss * ss) / half


def line_integrals(ell: Ellipses, n: int, angles_deg: np.ndarray, px: int,
                   shifts: np.ndarray | None = None, center: float = 0.0) -> np.ndarray:
    """Analytic sinogram [A][px] in per-pixel units.

    Bin X samples s = X - px/2 - center - shifts[a] (pixels): a positive
    `center` moves the rotation axis to bin px/2 + center, a positive shift
    moves projection a towards higher bins.
    """
    half = n / 2.0
    th = np.deg2rad(np.asarray(angles_deg, dtype=np.float64))
    X = np.arange(px, dtype=np.float64)
    s = X[None, :] - px / 2.0 - center
    if shifts is not None:
        s = s - np.asarray(shifts, dtype=np.float64)[:, None]
    s = s / half                                      # unit-disk units
    ct, st = np.cos(th)[:, None], np.sin(th)[:, None]
    out = np.zeros((len(th), px), dtype=np.float64)
    for k in range(len(ell)):
        # ray normal n = (cos th, -sin th); angle between n and the major axis
        alpha = -th[:, None] - ell.phi[k]
        r2 = (ell.a[k] * np.cos(alpha)) ** 2 + (ell.b[k] * np.sin(alpha)) ** 2
        sp = s - (ell.cx[k] * ct - ell.cy[k] * st)
        inside = r2 - sp * sp
        out += np.where(inside > 0.0,
                        2.0 * ell.mu[k] * ell.a[k] * ell.b[k] * np.sqrt(np.maximum(inside, 0.0)) / r2,
                        0.0)
    return out


def transmission_noise(p: np.ndarray, i0: float, rng: np.random.Generator) -> np.ndarray:
    """y = ln(I0 / Poisson(I0 exp(-p))) with the mt_conversion clamps."""
    counts = rng.poisson(i0 * np.exp(-p)).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        trans = i0 / counts
    trans = np.where(np.isnan(trans), 1.0, np.clip(trans, 1.0e-5, 1.0e7))
    return np.log(trans)


def uniform_angles(pa: int, span_deg: float = 180.0) -> np.ndarray:
    """pa angles uniformly over [0, span) degrees, as float32 (the angle log)."""
    return (np.arange(pa, dtype=np.float64) * (span_deg / pa)).astype(np.float32)


def sigmoid_edge_mu(base_mu: np.ndarray, n_energy: int, rng: np.random.Generator,
                    width: float = 0.02) -> np.ndarray:
    """Per-ellipse attenuation across energies E in [0, 1]: base + jump * sigmoid((E - edge)/width)
    with edge ~ U(0.4, 0.6) and jump ~ U(0.2, 1.0) per ellipse.  Returns [n_energy][n_ellipses]."""
    m = len(base_mu)
    edge = rng.uniform(0.4, 0.6, m)
    jump = rng.uniform(0.2, 1.0, m)
    e = np.linspace(0.0, 1.0, n_energy)[:, None]
    return base_mu[None, :] + jump[None, :] / (1.0 + np.exp(-(e - edge[None, :]) / width))


def make_slice(n: int, pa: int, n_ellipses: int, seed: int, *, px: int | None = None,
               i0: float | None = None, shifts: np.ndarray | None = None, center: float = 0.0,
               slice_index: int = 0, energy_index: int = 0):
    """One seeded test case: (image, sinogram float32 [pa][px], angles float32)."""
    px = n if px is None else px
    rng = np.random.default_rng(stream_seed(seed, slice_index, energy_index))
    ell = random_ellipses(n_ellipses, rng)
    ang = uniform_angles(pa)
    p = line_integrals(ell, n, ang, px, shifts=shifts, center=center)
    if i0 is not None:
        p = transmission_noise(p, i0, rng)
    img = rasterize(ell, n)
    return img.astype(np.float32), p.astype(np.float32), ang


def line_integrals_torch(ells, n: int, angles_deg, px: int, device="cuda", shifts=None, center: float = 0.0):
    """line_integrals() for a list of Ellipses sets at once, evaluated with torch
    on `device` (float64).  Returns a tensor [len(ells)][A][px]."""
    import torch
    half = n / 2.0
    th = torch.deg2rad(torch.as_tensor(np.asarray(angles_deg, dtype=np.float64), device=device))
    X = torch.arange(px, dtype=torch.float64, device=device)
    s = X[None, :] - px / 2.0 - center
    if shifts is not None:
        s = s - torch.as_tensor(np.asarray(shifts, dtype=np.float64), device=device)[:, None]
    s = (s / half)[None]                                 # [1][A][px]
    ct, st = torch.cos(th)[None, :, None], torch.sin(th)[None, :, None]
    cat = lambda f: torch.as_tensor(np.stack([getattr(e, f) for e in ells]), device=device)  # noqa: E731
    cx, cy, a, b, phi, mu = (cat(f) for f in ("cx", "cy", "a", "b", "phi", "mu"))   # [S][E]
    out = torch.zeros((len(ells), len(th), px), dtype=torch.float64, device=device)
    for k in range(cx.shape[1]):
        alpha = -th[None, :, None] - phi[:, k, None, None]
        r2 = (a[:, k, None, None] * torch.cos(alpha)) ** 2 + (b[:, k, None, None] * torch.sin(alpha)) ** 2
        sp = s - (cx[:, k, None, None] * ct - cy[:, k, None, None] * st)
        inside = r2 - sp * sp
        out += torch.where(inside > 0.0,
                           2.0 * mu[:, k, None, None] * a[:, k, None, None] * b[:, k, None, None]
                           * torch.sqrt(torch.clamp(inside, min=0.0)) / r2,
                           torch.zeros((), dtype=torch.float64, device=device))
    return out


def transmission_noise_torch(p, i0: float, seed: int):
    """transmission_noise() with torch.poisson on p's device (seeded generator)."""
    import torch
    g = torch.Generator(device=p.device)
    g.manual_seed(seed & 0x7FFFFFFFFFFFFFFF)
    counts = torch.poisson(i0 * torch.exp(-p), generator=g)
    trans = torch.where(counts > 0, i0 / counts, torch.full_like(counts, 1.0e7))
    return torch.log(torch.clamp(trans, 1.0e-5, 1.0e7))


C3_SIDES = (256, 384, 512, 768, 1024)


def mixed_stack_sides(n_slices: int, seed: int, sides=C3_SIDES) -> list:
    """Seeded side of each slice of an uneven stack (shared by all energies)."""
    return [int(sides[np.random.default_rng(stream_seed(seed, s, 0)).integers(len(sides))]) for s in range(n_slices)]


def mixed_energy_stack_torch(n_energy: int, n_slices: int, pa: int, seed: int, *, n_ellipses: int = 10,
                             i0: float | None = 1.0e4, sides=C3_SIDES, slice_sides=None, device="cuda"):
    """The batched CT-XAFS stack: slice s has a seeded side from `sides` and a
    seeded ellipse set shared by all energies; each ellipse's attenuation
    follows sigmoid_edge_mu over the energies; Poisson noise per (slice,
    energy).  slice_sides, when given, fixes the side of every slice instead.
    Returns (sides, angles, sinograms) with sinograms a list in slice-major
    order (index s * n_energy + e) of float32 tensors [pa][n]."""
    ang = uniform_angles(pa)
    n_of = list(slice_sides) if slice_sides is not None else mixed_stack_sides(n_slices, seed, sides)
    sinos = [None] * (n_slices * n_energy)
    for s in range(n_slices):
        rng = np.random.default_rng(stream_seed(seed, s, 1))
        ell = random_ellipses(n_ellipses, rng)
        mus = sigmoid_edge_mu(ell.mu, n_energy, rng)
        sets = [ell.with_mu(mus[e]) for e in range(n_energy)]
        p = line_integrals_torch(sets, n_of[s], ang, n_of[s], device=device)
        for e in range(n_energy):
            y = p[e] if i0 is None else transmission_noise_torch(p[e], i0, stream_seed(seed, s, e + 2))
            sinos[s * n_energy + e] = y.float().contiguous()
        del p
    return n_of, ang, sinos
