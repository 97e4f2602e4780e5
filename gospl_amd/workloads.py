"""The benchmark workloads the hot path is measured and checked on.

``planar_case``  — 100x100 km jittered grid (200x200 at 500 m, +-100 m seeded
                   jitter), fixed base-level borders at 0 m, z = 0.01 x
                   distance-to-boundary + U[0,1), uplift 1 mm/yr, rain 1 m/yr,
                   K=1e-5, m=0.5, n=1, 3 receivers, diffusivity 0.1 m2/yr,
                   dt=1000 yr.
``sphere_case``  — icosahedral sphere of radius 6371 km at a given refinement
                   level (9 -> 2,621,442 vertices), SH field up to degree 64
                   with spectrum ~ l^-2 scaled to [-6000, 5000] m, sea level
                   0, lognormal rain (median 1 m/yr, sigma_ln 0.6, clipped to
                   [0.2, 3]), K=5e-6, m=0.5, n=1, 3 receivers, land and marine
                   diffusivity 0.5 m2/yr, dt=5000 yr.

Each returns ``(mesh, h0, params)`` with ``mesh`` a ``gospl_amd.mesh.Mesh``
(device order), ``h0`` the initial elevation in device order and ``params``
the model parameters.  Sizes can be reduced for tests (``n``/``level``).
"""

from __future__ import annotations

import numpy as np

from . import fields
from . import mesh as _mesh


def planar_case(lib, n: int = 200, dx: float = 500.0, seed: int = 0, order="morton"):
    coords, cells = _mesh.planar_jittered(n, n, dx, 100.0 * dx / 500.0, seed)
    m = _mesh.build_mesh(coords, cells, lib, flat=True, order=order)
    bmask_in = np.zeros(m.N, bool)
    bmask_in[_mesh.boundary_vertices(cells)] = True
    z = fields.planar_initial(coords, bmask_in, seed)
    upsub = np.where(bmask_in, 0.0, 1.0e-3)
    params = dict(K=1.0e-5, m=0.5, d=0.0, flowdir=3, cda=0.1, cdm=0.1, dt=1000.0,
                  rain=1.0, upsub=m.to_device_order(upsub), sealevel=0.0)
    return m, m.to_device_order(z), params


def sphere_case(lib, level: int = 9, seed: int = 0, lmax: int = 64, slope: float = 2.0,
                device=None, order="morton", z=None):
    """The icosahedral-sphere workload; ``z`` (input order) replaces the SH
    surface when given (e.g. one saved by an earlier run)."""
    coords, cells = _mesh.icosphere(level, 6.371e6)
    m = _mesh.build_mesh(coords, cells, lib, flat=False, order=order)
    if z is None:
        z = fields.affine_range(fields.sh_field(coords, lmax, slope, seed=seed, device=device), -6000.0, 5000.0)
    rain = fields.lognormal_rain(m.N, 1.0, 0.6, 0.2, 3.0, seed=seed + 1)
    params = dict(K=5.0e-6, m=0.5, d=0.0, flowdir=3, cda=0.5, cdm=0.5, dt=5000.0,
                  rain=m.to_device_order(rain), upsub=None, sealevel=0.0)
    return m, m.to_device_order(z), params


def adjoint_case(lib, level: int = 8, seed: int = 0, lmax: int = 64, device=None, order="morton"):
    """The inversion workload: the sphere of ``sphere_case``; a target
    surface from a run with per-vertex log10 K = -5.3 plus a seeded smooth
    perturbation (SH degree <= 16, scaled to max |dlogK| = 0.3) and uplift
    0.5 mm/yr; the inverse run starts from uniform log10 K = -5.3.  dt 5000
    yr, 50 steps.  Returns (mesh, h0, params, logk_true, logk0, uplift),
    per-vertex arrays in device order; params exclude K and upsub."""
    m, h0, p = sphere_case(lib, level=level, seed=seed, lmax=lmax, device=device, order=order)
    coords_in = m.to_input_order(m.coords)
    pert = fields.sh_field(coords_in, 16, 2.0, seed=seed + 17, device=device)
    pert = 0.3 * pert / np.abs(pert).max()
    logk_true = m.to_device_order(-5.3 + pert)
    logk0 = np.full(m.N, -5.3)
    uplift = np.full(m.N, 5.0e-4)
    p = {k: v for k, v in p.items() if k not in ("K", "upsub", "d")}
    p["nsteps"] = 50
    return m, h0, p, logk_true, logk0, uplift


def ensemble_case(lib, counts=((6, 64), (7, 64), (8, 16)), seed: int = 0, lmax: int = 64, device=None,
                  order="morton"):
    """The ensemble workload: members on icosahedral spheres of uneven size
    (default 64 at level 6, 64 at level 7, 16 at level 8), each initialised
    as ``sphere_case`` with its own seed, with per-member K log-uniform in
    [1e-6, 1e-5], m uniform in [0.3, 0.6] and a rain scale uniform in
    [0.5, 2]; dt 5000 yr, 50 steps.  Returns [(mesh, h0, params)] (the mesh
    of a level is built once and shared)."""
    rng = np.random.default_rng(seed + 1000)
    members = []
    k = 0
    for level, n in counts:
        coords, cells = _mesh.icosphere(level, 6.371e6)
        m = _mesh.build_mesh(coords, cells, lib, flat=False, order=order)
        for _ in range(n):
            s = seed + 1 + k
            z = fields.affine_range(fields.sh_field(coords, lmax, 2.0, seed=s, device=device), -6000.0, 5000.0)
            rain = fields.lognormal_rain(m.N, 1.0, 0.6, 0.2, 3.0, seed=s + 1)
            p = dict(K=float(10.0 ** rng.uniform(-6.0, -5.0)), m=float(rng.uniform(0.3, 0.6)), d=0.0, flowdir=3,
                     cda=0.5, cdm=0.5, dt=5000.0, rain=m.to_device_order(rain) * float(rng.uniform(0.5, 2.0)),
                     upsub=None, sealevel=0.0)
            members.append((m, m.to_device_order(z), p))
            k += 1
    return members


def oracle_geometry(m: "_mesh.Mesh"):
    """Input-order geometry dict for ``oracle.OracleModel``."""
    ex = m.extras
    return dict(ngb=ex["ngb_input"], nnb=ex["nnb_input"], elgt=ex["elgt_input"],
                vdist=ex["vdist_input"], area=m.to_input_order(m.area),
                outlet=m.to_input_order(m.outlet).astype(bool), flat=m.flat,
                coords=m.to_input_order(m.coords))
