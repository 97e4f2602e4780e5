"""Synthetic initial conditions and forcing for the benchmark workloads.

* ``sh_field``: seeded random spherical-harmonic field up to degree ``lmax``
  with power spectrum ~ l**-p, evaluated with fully normalised associated
  Legendre functions (stable three-term recurrence) in float64, on the GPU
  when one is given (the 2.6M- and 10.5M-vertex spheres) or on the CPU.
* ``affine_range``: map a field onto [lo, hi].
* ``lognormal_rain``: lognormal precipitation (median, sigma_ln), clipped.
* ``planar_initial``: 0.01 x distance-to-boundary + U[0,1) on a planar mesh.
"""

from __future__ import annotations

import math

import numpy as np


def sh_field(coords, lmax: int, slope: float = 2.0, seed: int = 0, device=None, lmin: int = 1):
    import torch

    dev = torch.device(device) if device is not None else torch.device("cpu")
    xyz = torch.as_tensor(np.asarray(coords, dtype=np.float64), device=dev)
    r = torch.linalg.norm(xyz, dim=1)
    ct = (xyz[:, 2] / r).clamp(-1.0, 1.0)
    st = torch.sqrt(torch.clamp(1.0 - ct * ct, min=0.0))
    phi = torch.atan2(xyz[:, 1], xyz[:, 0])
    rng = np.random.default_rng(seed)
    # coefficients a_lm, b_lm ~ N(0, l^-slope / (2l+1))
    out = torch.zeros_like(ct)
    # P_mm via recurrence, then P_lm for l > m
    pmm = torch.ones_like(ct) * math.sqrt(1.0 / (4.0 * math.pi))
    coef = {}
    for l in range(lmax + 1):
        sd = math.sqrt((l ** -slope) / (2 * l + 1)) if l >= max(lmin, 1) else 0.0
        a = rng.standard_normal(l + 1) * sd
        b = rng.standard_normal(l + 1) * sd
        coef[l] = (a, b)
    for m in range(lmax + 1):
        if m > 0:
            pmm = -math.sqrt((2 * m + 1) / (2.0 * m)) * st * pmm
        p_lm2 = None
        p_lm1 = pmm
        acc_c = torch.zeros_like(ct)
        acc_s = torch.zeros_like(ct)
        for l in range(m, lmax + 1):
            if l == m:
                plm = pmm
            elif l == m + 1:
                plm = math.sqrt(2 * m + 3) * ct * pmm
            else:
                a_ = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                b_ = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                plm = a_ * (ct * p_lm1 - b_ * p_lm2)
            if l != m:
                p_lm2, p_lm1 = p_lm1, plm
            if l >= lmin:
                a, b = coef[l]
                acc_c = acc_c + a[m] * plm
                if m > 0:
                    acc_s = acc_s + b[m] * plm
        if m == 0:
            out = out + acc_c
        else:
            out = out + acc_c * torch.cos(m * phi) + acc_s * torch.sin(m * phi)
    return out.cpu().numpy()


def affine_range(f, lo: float, hi: float):
    f = np.asarray(f, dtype=np.float64)
    fmin, fmax = f.min(), f.max()
    return lo + (f - fmin) * ((hi - lo) / (fmax - fmin))


def lognormal_rain(n: int, median: float = 1.0, sigma: float = 0.6, lo: float = 0.2,
                   hi: float = 3.0, seed: int = 0):
    rng = np.random.default_rng(seed)
    r = median * np.exp(sigma * rng.standard_normal(n))
    return np.clip(r, lo, hi)


def planar_initial(coords, boundary_mask, seed: int = 0):
    """z = 0.01 * distance-to-boundary + U[0,1) inside, 0 on the boundary."""
    rng = np.random.default_rng(seed)
    x, y = coords[:, 0], coords[:, 1]
    dist = np.minimum.reduce([x - x.min(), x.max() - x, y - y.min(), y.max() - y])
    z = 0.01 * dist + rng.uniform(0.0, 1.0, len(x))
    z[boundary_mask] = 0.0
    return z
