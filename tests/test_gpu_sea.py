"""Marine deposition of river sediment (seadepo True, nodep False) on the GPU
against the oracle's restatement of goSPL's SEAMesh
(gospl/sed/seaplex.py:51-861).

Per phase from identical state, as the sediment-routing test does: the
oracle is handed the GPU's state after the continental sediment routing of
each step (elevation, cumulative erosion/deposition, sediment load, filled
surface), then both run seaChange.  Bars: coastline distances, the smoothed
and epsilon-filled marine surfaces, the marine receivers and weights and the
routed volumes bit for bit (the same operations in the same order: the same
count of Chebyshev steps over the Jacobi sweep with the same weights, the
epsilon flood's unique fixed point, sums per receiver in increasing input
id); the coupled Gmar > 0 solve, the marine
diffusion, elevation, cumulative erosion/deposition and rate within 1e-9
relative (GMRES / Jacobi against sparse direct solves).  Cases: the closed
level-5 sphere with Gmar = 0 and Gmar = 0.5, and a plane whose sea level
floods its draining borders.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import case, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _flooded_plane():
    m, h0, p = case("planar", 40)
    q = dict(p)
    q["sealevel"] = 6.0
    return m, h0, q


CASES = {
    "sphere5": lambda: (case("sphere", 5), {}),
    "sphere5_gmar": lambda: (case("sphere", 5), dict(Gmar=0.5)),
    "plane40_sea": lambda: (_flooded_plane(), {}),
}


@pytest.mark.parametrize("name", ["sphere5", "sphere5_gmar", "plane40_sea"])
def test_sea_change_matches_oracle(name):
    from gospl_amd.model import Model

    (m, h0, p), kw = CASES[name]()
    print("marine deposition:", name, m.N, flush=True)
    g = Model(m, h0, tend=1e12, nodep=False, seadepo=True, marineSolver="picard", **p, **kw)
    o = oracle_model(m, h0, p, nodep=False, seadepo=True, marineSolver="picard", **kw)
    g.initialise()
    o.initialise()
    inp = lambda t: m.to_input_order(t.cpu().numpy())
    moved = 0
    for s in range(2):
        o.h = inp(g.hLocal).copy()
        o.cumED = inp(g.cumED).copy()
        o.tNow = g.tNow
        g.flowAccumulation()
        o.flowAccumulation()
        g.erodepSPL()
        o.erodepSPL()
        o.h, o.cumED, o.Eb = inp(g.hLocal).copy(), inp(g.cumED).copy(), inp(g.Eb).copy()
        g.flowAccumulation()
        o.flowAccumulation()
        g._snapshot_sea()
        g.sedChange()
        o.sedChange()
        torch.cuda.synchronize()
        # identical state before seaChange
        o.h, o.cumED, o.EbLocal = inp(g.hLocal).copy(), inp(g.cumED).copy(), inp(g.EbLocal).copy()
        o.vSedLocal = inp(g.vSedLocal).copy()
        assert np.array_equal(inp(g.lFill), o.lFill), s
        assert np.array_equal(inp(g._seaFlow).astype(bool), np.isin(np.arange(m.N), o.seaID)), s
        g.seaChange()
        o.seaChange()
        torch.cuda.synchronize()
        st = g.sea_stats
        print("  step %d: coast points %d (oracle %d), smooth sweeps %s, epsfill passes %d, routing passes %s "
              "(oracle %s)" % (s, st["coast_points"], len(o.coast_points), st.get("smooth_sweeps"),
                               st["epsfill_passes"], st["route_passes"], o.route_passes[-len(st["route_passes"]):]),
              flush=True)
        assert st["coast_points"] == len(o.coast_points)
        assert np.array_equal(inp(g.coastDist), o.coastDist), (s, rel_err(inp(g.coastDist), o.coastDist))
        sb = g._seabuf
        assert np.array_equal(inp(sb["fillz"]), o.fillz), s
        rcv = m.to_input_order(sb["rcv"].cpu().numpy().T)
        rcv_in = np.where(rcv >= 0, m.perm[np.maximum(rcv, 0)], -1)
        assert np.array_equal(rcv_in, o.marRcv), s
        assert np.array_equal(m.to_input_order(sb["wgt"].cpu().numpy().T), o.marWgt), s
        assert np.array_equal(inp(sb["sink"]).astype(bool), o.is_sink), s
        assert st["route_passes"] == o.route_passes[-len(st["route_passes"]):], s
        mdep = inp(sb["vdep"])
        if kw.get("Gmar"):
            assert st["dep_relres"] <= 1e-13, st
            assert np.abs(mdep - o.marDep).max() <= TOL * max(np.abs(o.marDep).max(), 1.0), s
        else:
            assert np.array_equal(mdep, o.marDep), (s, np.abs(mdep - o.marDep).max())
        assert np.abs(inp(sb["dh"]) - o.marThick).max() <= TOL * max(o.marThick.max(), 1.0), s
        hmax = np.abs(o.h).max()
        assert rel_err(inp(g.hLocal), o.h) < TOL, (s, rel_err(inp(g.hLocal), o.h))
        assert np.abs(inp(g.cumED) - o.cumED).max() <= TOL * hmax, s
        assert np.abs(inp(g.EbLocal) - o.EbLocal).max() <= TOL * hmax / o.dt, s
        moved += int((o.marThick > 0).sum())
        g.getHillslope()
        o.getHillslope()
        g.applyForces()
        o.applyForces()
        g.tNow += g.dt
        o.tNow += o.dt
    if name != "plane40_sea":
        assert moved > 0
    g.close()


def test_coast_distance_through_c_abi():
    """gospl_coast_distance on random surfaces of the level-5 sphere against
    the contour restatement and scipy's cKDTree (all vertices, then a mask)."""
    from scipy import spatial

    from gospl_amd.model import Model, _ptr
    from oracle import sea

    m, h0, p = case("sphere", 5)
    g = Model(m, h0, tend=1e12, **p)
    o = oracle_model(m, h0, p)
    xyz = torch.as_tensor(np.ascontiguousarray(m.coords, np.float64)).to(g.device)
    rng = np.random.default_rng(3)
    for trial in range(3):
        z_in = rng.uniform(-1.0, 1.0, m.N) + (0.0 if trial < 2 else 0.8)
        pts = sea.contour_points(z_in, 0.0, o.coords, o.ngb, o.nnb)
        want = spatial.cKDTree(pts).query(o.coords, k=1)[0]
        mask_in = rng.random(m.N) < 0.5 if trial == 1 else np.ones(m.N, bool)
        want = np.where(mask_in, want, 0.0)
        z = g._dev(m.to_device_order(z_in))
        mask = g._dev(m.to_device_order(mask_in).astype(np.int8), torch.int8)
        out = torch.zeros(m.N, dtype=torch.float64, device=g.device)
        n = ctypes.c_int64(0)
        g._call("gospl_coast_distance", g.H, _ptr(xyz), _ptr(z), 0.0, _ptr(mask), _ptr(out), ctypes.byref(n),
                g.stream)
        got = m.to_input_order(out.cpu().numpy())
        print("coast distance trial %d: %d points" % (trial, n.value), flush=True)
        assert n.value == len(pts)
        assert np.array_equal(got, want), np.abs(got - want).max()
    # no coastline: all distances 0
    z = g._dev(np.full(m.N, 5.0))
    g._call("gospl_coast_distance", g.H, _ptr(xyz), _ptr(z), 0.0, None, _ptr(out), ctypes.byref(n), g.stream)
    assert n.value == 0 and float(out.abs().max()) == 0.0
    g.close()


@pytest.mark.parametrize("scale", [1.0, 40.0])
def test_diffuse_fixed_matches_oracle_bitwise(scale):
    """gospl_diffuse_fixed (the marine smoothing: Chebyshev semi-iteration
    over the Jacobi sweep) through the C ABI against the oracle's
    smooth_solve on the level-5 sphere, bit for bit, with the marine Kd and
    with it scaled by 40 (contraction ~0.9987: ~800 steps)."""
    from gospl_amd.model import Model, _ptr
    from oracle import oracle as orc
    from oracle import sea

    m, h0, p = case("sphere", 5)
    g = Model(m, h0, tend=1e12, **p)
    o = oracle_model(m, h0, p)
    rng = np.random.default_rng(21)
    h_in = m.to_input_order(h0) + rng.uniform(-50.0, 50.0, m.N)
    Cd = np.where(h_in < 0.0, sea.MARINE_SMOOTH_N_SEA, sea.MARINE_SMOOTH_N_LAND)
    kd_in = scale * Cd * o.larea
    coeff = orc.sethillslopecoeff(kd_in, o.ngb, o.nnb, o.larea, o.elgt, o.vdist)
    want, nsw = orc.smooth_solve(coeff, h_in.copy(), o.ngb, o.nnb)
    kd = g._dev(m.to_device_order(kd_in))
    rhs = g._dev(m.to_device_order(h_in))
    out = torch.zeros(m.N, dtype=torch.float64, device=g.device)
    sw = ctypes.c_int32(0)
    g._call("gospl_diffuse_fixed", g.H, _ptr(kd), _ptr(rhs), _ptr(out), 1000000, ctypes.byref(sw), g.stream)
    got = m.to_input_order(out.cpu().numpy())
    print("diffuse_fixed scale %g: %d steps (oracle %d)" % (scale, sw.value, nsw), flush=True)
    assert sw.value == nsw
    assert np.array_equal(got, want), np.abs(got - want).max()
    g.close()


def test_coast_distance_single_island():
    """A coastline of fewer points than one leaf of the search tree: one
    vertex above the level, all others below."""
    from scipy import spatial

    from gospl_amd.model import Model, _ptr
    from oracle import sea

    m, h0, p = case("sphere", 5)
    g = Model(m, h0, tend=1e12, **p)
    o = oracle_model(m, h0, p)
    xyz = torch.as_tensor(np.ascontiguousarray(m.coords, np.float64)).to(g.device)
    z_in = np.full(m.N, -1.0)
    z_in[123] = 1.0
    pts = sea.contour_points(z_in, 0.0, o.coords, o.ngb, o.nnb)
    assert 0 < len(pts) < 8
    want = spatial.cKDTree(pts).query(o.coords, k=1)[0]
    z = g._dev(m.to_device_order(z_in))
    out = torch.zeros(m.N, dtype=torch.float64, device=g.device)
    n = ctypes.c_int64(0)
    g._call("gospl_coast_distance", g.H, _ptr(xyz), _ptr(z), 0.0, None, _ptr(out), ctypes.byref(n), g.stream)
    got = m.to_input_order(out.cpu().numpy())
    assert n.value == len(pts)
    assert np.array_equal(got, want), np.abs(got - want).max()
    g.close()
