"""Pins of the oracle's marine deposition (oracle/sea.py, goSPL's
SEAMesh.seaChange, gospl/sed/seaplex.py:51-861) -- CPU only.

* epsfill and mfdrcvrs against goSPL's own Fortran (oracle/_ref), on the
  golden meshes with random surfaces, flats included;
* the coastline points against their definition, the edges enumerated from
  the triangles (not the neighbour table): one point per edge straddling the
  level, on that edge, where the linear interpolation of the elevation is the
  level; the distances against a brute-force nearest point;
* the routing of a whole call: every input volume is deposited somewhere or
  leaves through an outlet, no vertex outside the sinks takes more than its
  shelf capacity;
* the coupled marine deposition on a hand-solved chain (each sea vertex keeps
  fDepm of what reaches it) and its flux block's balance on a sphere;
* the marine smoothing (Chebyshev semi-iteration) against a sparse direct
  solve and against plain Jacobi sweeps run to their fixed point.
"""

from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from gospl_amd import mesh as gmesh
from oracle import ref, sea
from tests.helpers import case, oracle_model

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(path):
    return dict(np.load(path, allow_pickle=False))


@pytest.mark.skipif(not ref.available(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(p))
def test_epsfill_and_marine_receivers_match_fortran(path):
    g = _load(path)
    v, c, flat = g["v"], g["c"], bool(g["flat"])
    N = len(v)
    vi = gmesh.voronoi_inputs(v, c, flat)
    ref.fv_setup(v, c, vi["cells_edges"], vi["edges"], vi["circumcenters"], 0 if flat else 1)
    ref.globalngbhs(N, c)
    rng = np.random.default_rng(11)
    gid = np.arange(N, dtype=np.int32)
    for trial in range(4):
        h = rng.uniform(-40.0, 40.0, N)
        if trial >= 2:
            h = np.round(h / 8.0) * 8.0  # wide flats: the epsilon increments decide
        sl = float(np.quantile(h, 0.2))
        f_ref = ref.epsfill(sl, h)
        f_or = sea.epsfill(sl, h, g["ngb"], g["nnb"])
        assert np.array_equal(f_or, f_ref)
        assert (f_ref != h).any()
        r_ref, _, w_ref = ref.mfdrcvrs(12, 1.1, f_ref, sea.BOUNDARY_FLOW_SENTINEL, gid)
        r_or, w_or, sink = sea.marine_receivers(f_ref, g["ngb"], g["nnb"], g["elgt"])
        assert np.array_equal(r_or, r_ref)
        assert np.array_equal(w_or, w_ref)
        assert np.array_equal(sink, np.all(r_ref == np.arange(N)[:, None], axis=1))


def _tri_edges(cells):
    e = np.concatenate([cells[:, [0, 1]], cells[:, [1, 2]], cells[:, [2, 0]]])
    e = np.sort(e, axis=1)
    return np.unique(e, axis=0)


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(p))
def test_contour_points_and_distances_by_definition(path):
    from scipy import spatial

    g = _load(path)
    v, c = g["v"], g["c"]
    N = len(v)
    rng = np.random.default_rng(5)
    z = rng.uniform(-1.0, 1.0, N)
    level = 0.1
    pts, pe = sea.contour_points(z, level, v, g["ngb"], g["nnb"], with_edges=True)
    want = set()
    for a, b in _tri_edges(c):
        lo, hi = (a, b) if z[a] < z[b] else (b, a)
        if z[lo] < level <= z[hi]:
            want.add((int(lo), int(hi)))
    got = [(int(a), int(b)) for a, b in pe]
    assert len(got) == len(set(got)) == len(want) > 0 and set(got) == want
    # each point lies on its edge, where the interpolated elevation is the level
    lo, hi = pe[:, 0], pe[:, 1]
    d = v[hi] - v[lo]
    t = ((pts - v[lo]) * d).sum(axis=1) / (d * d).sum(axis=1)
    assert np.all(t >= -1e-12) and np.all(t <= 1 + 1e-12)
    off = np.linalg.norm(v[lo] + t[:, None] * d - pts, axis=1)
    assert np.all(off <= 1e-9 * np.linalg.norm(d, axis=1))
    assert np.abs(z[lo] + t * (z[hi] - z[lo]) - level).max() <= 1e-9
    # cKDTree distances equal brute force
    q = v[rng.random(N) < 0.3]
    dt, _ = spatial.cKDTree(pts).query(q, k=1)
    bf = np.sqrt(((q[:, None, :] - pts[None, :, :]) ** 2).sum(axis=2)).min(axis=1)
    assert np.allclose(dt, bf, rtol=1e-14, atol=0.0)


def _sea_state(kind, size, **kw):
    m, h0, p = case(kind, size)
    o = oracle_model(m, h0, p, nodep=False, seadepo=True, marineSolver="picard", **kw)
    o.initialise()
    o.step()
    return o


def test_routing_conserves_volume_and_respects_capacity():
    o = _sea_state("sphere", 5)
    vin = o.marineInput
    assert vin.sum() > 0
    dep = o._distOcean(vin)
    # closed sphere, no outlet: all of it is deposited, up to the 1 + O(6e-8)
    # sums of goSPL's single-precision weights (1. / kk), once per pass
    wsum = o.marWgt.sum(axis=1)
    assert np.abs(wsum[~o.is_sink] - 1.0).max() < 1.2e-7
    assert abs(dep.sum() - vin.sum()) <= 1.2e-7 * len(o.route_passes) * o.route_passes[-1] * vin.sum()
    over = (dep > o.maxDepQs * (1 + 1e-12)) & ~o.is_sink
    # only vertices holding the final residual (deposited where it stands) may exceed their capacity
    assert over.sum() <= 0.01 * o.lpoints


def test_coupled_marine_chain_by_hand():
    # 2 -> 1 -> 0 (sink); all below sea level
    rcv = np.array([[0] * 12, [0] + [1] * 11, [1] + [2] * 11], np.int32)
    wgt = np.zeros((3, 12))
    wgt[1, 0] = 1.0
    wgt[2, 0] = 1.0
    h = np.array([-100.0, -50.0, -20.0])
    area = np.array([2.0, 3.0, 5.0])
    vol = np.array([7.0, 11.0, 13.0])
    dt, G = 4.0, 0.3
    hn, q, vd = sea.marine_dep_solve(rcv, wgt, h, vol, G, 0.0, area, dt)
    # each vertex keeps fDepm of the volume reaching it, passes on the rest
    d2 = G * vol[2]
    d1 = G * (vol[1] + (1 - G) * vol[2])
    d0 = G * (vol[0] + (1 - G) * (vol[1] + (1 - G) * vol[2]))
    assert np.allclose(vd, [d0, d1, d2], rtol=1e-13, atol=0)
    assert np.allclose(q, [(1 - G) * (vol[0] + (1 - G) * (vol[1] + (1 - G) * vol[2])) / dt,
                           (1 - G) * (vol[1] + (1 - G) * vol[2]) / dt, (1 - G) * vol[2] / dt], rtol=1e-13)
    # above sea level nothing is kept
    hn2, _, vd2 = sea.marine_dep_solve(rcv, wgt, np.array([-100.0, 5.0, -20.0]), vol, G, 0.0, area, dt)
    assert vd2[1] == 0.0


def test_coupled_marine_flux_balance_on_sphere():
    o = _sea_state("sphere", 5, Gmar=0.4)
    o._matOcean()
    vol = o.marDep.copy()
    hn, q, vd = sea.marine_dep_solve(o.marRcv, o.marWgt, o.h, vol, 0.4, o.sealevel, o.larea, o.dt)
    kept = np.sum((hn - o.h) * o.larea)
    # what each vertex does not pass on leaves the block: all of it at sinks,
    # 1 - sum(w) (single-precision weights) elsewhere
    passed = np.where(o.marRcv != np.arange(o.lpoints)[:, None], o.marWgt, 0.0).sum(axis=1)
    left = np.sum(q * (1.0 - passed)) * o.dt
    assert abs(vol.sum() - kept - left) <= 1e-9 * vol.sum()
    assert vd.sum() > 0


@pytest.mark.parametrize("scale", [1.0, 40.0])
def test_marine_smoothing_against_direct_solve(scale):
    """The Chebyshev smoothing of the oracle (oracle.smooth_solve) against a
    sparse direct solve of the same system (I + L(kd)) x = h and against
    plain Jacobi sweeps run to their fixed point; scale 40 makes the sweep's
    contraction close to 1 (several hundred Jacobi sweeps)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from oracle import oracle as orc

    o = _sea_state("sphere", 5)
    N = o.lpoints
    Cd = np.full(N, sea.MARINE_SMOOTH_N_LAND)
    Cd[o.seaID] = sea.MARINE_SMOOTH_N_SEA
    coeff = orc.sethillslopecoeff(scale * Cd * o.larea, o.ngb, o.nnb, o.larea, o.elgt, o.vdist)
    h = o.h.copy()
    x, nsw = orc.smooth_solve(coeff, h, o.ngb, o.nnb)
    rho, om = orc.chebyshev_plan(coeff[:, 0])
    assert nsw == len(om) and 0.0 < rho < 1.0
    rows, cols, vals = [np.arange(N)], [np.arange(N)], [coeff[:, 0]]
    for p in range(o.ngb.shape[1]):
        ok = (p < o.nnb) & (o.ngb[:, p] >= 0) & (coeff[:, p + 1] != 0.0)
        rows.append(np.nonzero(ok)[0])
        cols.append(o.ngb[ok, p])
        vals.append(coeff[ok, p + 1])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    xd = spla.spsolve(A.tocsc(), h)
    scl = np.abs(h).max()
    assert np.abs(x - xd).max() <= 1e-11 * scl, np.abs(x - xd).max() / scl
    xj, it = orc.hillslope_solve(coeff, h, o.ngb, o.nnb, maxit=200000)
    assert np.abs(x - xj).max() <= 1e-11 * scl
    # the plain sweeps need several times more passes than the semi-iteration
    if scale > 1.0:
        assert it > 3 * nsw, (it, nsw)
