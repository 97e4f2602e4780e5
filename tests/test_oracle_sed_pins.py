"""Pins of the oracle's restatement of goSPL's continental sediment routing
(oracle/sed.py) -- CPU only.

* goSPL's own Fortran: the lake-diffusion operator (jacobiancoeff with a zero
  derivative term, functions.F90:1421-1467) bit for bit against the live
  reference build, and its residual form fctcoeff.
* Definitions checked by brute force: bottom-up deposits at full fill and at
  the knots of the depth-volume curves (getpitvol's volumes), the micro-tilt's
  zero per-pit mean, pit inlets over every receiver edge.
* Mathematics: the Picard step solves its linear systems, conserves volume
  before the clip at zero, and keeps a flat surface flat; a whole
  sedChange call conserves the eroded volume (pit deposits + deposits at
  stray sinks + what reaches the sea) up to the excess goSPL drops when a
  cascade pass carries less than half the largest cell area.
"""

from __future__ import annotations

import numpy as np
import pytest

from gospl_amd import _lib, mesh as gmesh, workloads
from oracle import oracle as orc, ref
from oracle.sed import jacobian_rows, rows_matrix
from tests.helpers import oracle_model


def _sphere(level=4, seed=1, **kw):
    lib = _lib.load()
    m, h0, p = workloads.sphere_case(lib, level=level, seed=seed, lmax=min(64, 4 * 2 ** level))
    return m, oracle_model(m, h0, p, nodep=False, **kw)


@pytest.mark.skipif(not ref.available(), reason="reference build oracle/_ref absent")
def test_lake_operator_matches_fortran_jacobiancoeff():
    rng = np.random.default_rng(3)
    for level, flat in ((3, False), (None, True)):
        if flat:
            v, c = gmesh.planar_jittered(14, 11, 300.0, 60.0, 4)
        else:
            v, c = gmesh.icosphere(level, 6.371e6)
        vi = gmesh.voronoi_inputs(v, c, flat)
        ngb, nnb, elgt, vdist, area, narea = ref.fv_setup(v, c, vi["cells_edges"], vi["edges"],
                                                          vi["circumcenters"], 0 if flat else 1)
        N = len(v)
        h = rng.uniform(-100, 100, N)
        kd = np.where(rng.uniform(size=N) < 0.7, rng.uniform(0.5, 20.0, N), 0.0)
        diag, off = jacobian_rows(h, kd, ngb, nnb, area, elgt, vdist)
        want = ref.jacobiancoeff(h, kd, np.zeros(N))
        assert np.array_equal(diag, want[:, 0])
        assert np.array_equal(off, want[:, 1:])
        # fctcoeff is the same operator applied to h (up to summation order)
        lh = rows_matrix(diag, off, ngb) @ h
        assert np.allclose(lh, ref.fctcoeff(h, kd), rtol=1e-12, atol=1e-12 * np.abs(lh).max())


def test_picard_step_conserves_volume_and_solves_its_systems():
    m, o = _sphere(4)
    rng = np.random.default_rng(5)
    N = o.lpoints
    mask = rng.uniform(size=N) < 0.5
    h = o.h + np.where(mask, rng.uniform(0.0, 300.0, N), 0.0)
    Kd = np.where(mask, 7.0 + rng.uniform(0, 3, N), 0.0)
    diag, off = jacobian_rows(h, Kd, o.ngb, o.nnb, o.larea, o.elgt, o.vdist)
    L = rows_matrix(diag, off, o.ngb)
    # area-weighted, the operator is symmetric (fluxes balance across faces)
    A = L.multiply(o.larea[:, None]).tocsr()
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    x = rng.normal(size=N)
    assert abs((o.larea * (L @ x)).sum()) <= 1e-10 * np.abs(o.larea * (L @ x)).sum()
    # a flat surface is a fixed point of the diffusion
    assert np.abs(L @ np.full(N, 3.0)).max() <= 1e-12 * np.abs(diag).max() * 3.0
    # the Picard form: deposit volume kept before the clip, systems solved
    dh = np.where(mask, rng.uniform(0.0, 50.0, N), 0.0)
    o.picardSub, o.picardIts = 3, 2
    out = o._diffuseImplicit(dh.copy(), mask, 10.0)
    assert (out >= 0).all()
    lost = (o.larea * np.minimum(0.0, out)).sum()
    assert lost == 0.0
    # rerun without the clip: the volume is conserved exactly
    zb = o.h.copy()
    hh = dh + zb
    dts = o.dt / 3
    for _ in range(3):
        hs = hh.copy()
        for _ in range(2):
            dd = hh - zb
            dd[dd < 0.1] = 0.0
            Cd = np.where(mask, 1e-4 + 10.0 * (1 - np.exp(-0.05 * dd)), 0.0)
            dg, of = jacobian_rows(hh, Cd, o.ngb, o.nnb, o.larea, o.elgt, o.vdist)
            M = rows_matrix(dts * dg + 1.0, dts * of, o.ngb)
            hh = np.linalg.solve(M.toarray(), hs)
            assert np.abs(M @ hh - hs).max() <= 1e-9 * np.abs(hs).max()
    vol0 = (o.larea * dh).sum()
    assert abs((o.larea * (hh - zb)).sum() - vol0) <= 1e-10 * vol0
    assert np.allclose(np.maximum(hh - zb, 0.0), out, rtol=1e-9, atol=1e-9 * np.abs(out).max())


def _after_sed_fill(o):
    o.initialise()
    o.flowAccumulation()
    o.erodepSPL()
    o.flowAccumulation()
    o._getSedFlux()
    o.fillElevation(sed=True)
    return o.h.copy()


def test_bottom_up_deposits_follow_the_depth_volume_curves():
    m, o = _sphere(4)
    hl = _after_sed_fill(o)
    npit = len(o.pitParams)
    assert npit > 5
    sel = np.zeros(npit, bool)
    sel[1:] = True
    # full fill: every pit cell rises to its filled level, volume = pit volume
    depo = o.pitParams[:, 0].copy()
    delta = o._bottomUpDelta(hl, depo, sel)
    inpit = o.pitIDs > 0
    # (the top of the curve is bottom + 5 x depth / 5, the filled level up to
    # rounding)
    scale = np.abs(o.lFill).max()
    assert np.abs(delta[inpit] - np.maximum(0.0, o.lFill - hl)[inpit]).max() <= 1e-12 * scale
    assert (delta[~inpit] == 0).all()
    vol = np.zeros(npit)
    np.add.at(vol, o.pitIDs[inpit], (delta * o.larea)[inpit])
    assert np.allclose(vol[1:], o.pitParams[1:, 0], rtol=1e-9)
    # at each knot of the curve, the deposit is the water below that level
    for k in range(4):
        depo = o.filled_vol[:, k].copy()
        delta = o._bottomUpDelta(hl, depo, sel)
        lvl = o.filled_lvl[:, k]
        for p in range(1, npit):
            cells = np.nonzero(o.pitIDs == p)[0]
            if depo[p] <= 0 or o.pitParams[p, 0] <= 0:
                # (marine depressions hold no volume above sea level: no curve)
                continue
            brute = sum(max(0.0, lvl[p] - hl[c]) * o.larea[c] for c in cells)
            got = (delta[cells] * o.larea[cells]).sum()
            assert abs(got - brute) <= 1e-12 * max(brute, 1.0), (k, p)


def test_micro_tilt_has_zero_mean_and_points_to_the_spill():
    m, o = _sphere(4)
    hl = _after_sed_fill(o)
    npit = len(o.pitParams)
    sel = np.zeros(npit, bool)
    sel[1:] = True
    depo = 0.6 * o.pitParams[:, 0]
    delta = o._bottomUpDelta(hl, depo, sel)
    tilted = o._addPitMicroTilt(delta.copy(), sel)
    for p in range(1, npit):
        cells = np.nonzero((o.pitIDs == p) & (delta > 0))[0]
        if cells.size < 2:
            continue
        t = tilted[cells] - delta[cells]
        if (tilted[cells] > 0).all():
            assert abs(t.sum()) <= 1e-9 * np.abs(t).sum() + 1e-15
        sp = o.pitInfo[p, 0]
        d = np.array([np.sqrt(((o.coords[c] - o.coords[sp]) ** 2).sum()) for c in cells])
        far, near = cells[np.argmax(d)], cells[np.argmin(d)]
        assert tilted[far] - delta[far] >= tilted[near] - delta[near]


def test_inlets_by_brute_force():
    m, o = _sphere(4)
    _after_sed_fill(o)
    mask = o._identifyPitInlets()
    brute = np.zeros(o.lpoints, bool)
    for i in range(o.lpoints):
        for k in range(o.rcvIDi.shape[1]):
            r = int(o.rcvIDi[i, k])
            if r != i and o.pitIDs[r] >= 0 and o.pitIDs[r] != o.pitIDs[i]:
                brute[r] = True
    assert np.array_equal(mask, brute)


@pytest.mark.parametrize("level,seed", [(4, 1), (5, 3)])
def test_sediment_volume_balance(level, seed):
    """Eroded volume = deposits in pits + deposits at stray sinks (a closed
    sphere keeps them) + what reaches the sea, to within the passes'
    dropped excess (each below half the largest cell area)."""
    m, o = _sphere(level, seed)
    o.initialise()
    o.flowAccumulation()
    o.erodepSPL()
    o.flowAccumulation()
    V0 = (-o.Eb * o.larea).sum() * o.dt
    o.sedChange()
    sea = o.lFill <= o.sealevel
    at_sea = (o.vSedLocal[sea] * o.dt).sum()
    kept = o.sed_depo.sum() + o._closedDepo.sum()
    assert V0 > 0 and kept > 0
    slack = 0.5 * o.maxarea[0] * o.sed_passes + 1e-3 * len(o.pitParams)
    assert abs(V0 - kept - at_sea) <= slack + 1e-9 * V0, (V0, kept, at_sea, slack)
