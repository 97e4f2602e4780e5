"""Pins of the oracle's transport-limited stream-power solve
(oracle.coupled_ed_solve, goSPL's SPL._coupledEDSystem, SPL.py:116-258).

goSPL solves its nested h / q system with PETSc (absent here), so the
restatement is pinned by what the mathematics fixes:

* a three-vertex chain solved by hand: the deposition term enters with the
  right sign and scale (fDep dt / area on the upstream sediment flux);
* fDep = 0 reduces to the detachment-limited sweep of the same rows;
* the flux block: the eroded volume rate of the whole mesh equals the flux
  that leaves it (at vertices whose receivers do not take all their flow);
* the floored-denominator form of fDep equals min(G area / PA, 0.99).
"""

from __future__ import annotations

import numpy as np

from oracle import oracle as orc
from tests.helpers import case, oracle_model


def test_three_vertex_chain_by_hand():
    # 2 -> 1 -> 0, vertex 0 at sea (its own receiver, no erosion, no deposition)
    rcv = np.array([[0], [0], [1]], np.int32)
    wgt = np.array([[0.0], [1.0], [1.0]])
    c1, c2 = 0.7, 1.9  # stream-power coefficients Kbr lim / dist of vertices 1 and 2
    off = np.array([[0.0], [-c1], [-c2]])
    diag = np.array([1.0, 1.0 + c1, 1.0 + c2])
    area = np.array([3.0, 2.0, 5.0])
    dt = 7.0
    f1 = 0.4
    fdep = np.array([0.0, f1, 0.0])
    hold = np.array([-10.0, 20.0, 50.0])
    h, q = orc.coupled_ed_solve(rcv, off, diag, rcv, wgt, fdep, area, dt, hold)
    # vertex 2 erodes towards vertex 1 and its whole eroded volume reaches 1;
    # vertex 1 keeps fDep of the incoming flux:
    #   h1 (1 + c1 + f1 (a2/a1) c2/(1 + c2)) = hOld1 + c1 h0 + f1 (a2/a1) c2/(1 + c2) hOld2
    g = f1 * (area[2] / area[1]) * c2 / (1.0 + c2)
    h1 = (hold[1] + c1 * hold[0] + g * hold[2]) / (1.0 + c1 + g)
    h2 = (hold[2] + c2 * h1) / (1.0 + c2)
    assert abs(h[0] - hold[0]) < 1e-12
    assert abs(h[1] - h1) < 1e-12 * abs(h1)
    assert abs(h[2] - h2) < 1e-12 * abs(h2)
    # flux: what leaves 2 is its eroded volume rate; 1 passes on the rest
    q2 = area[2] * (hold[2] - h2) / dt
    q1 = q2 + area[1] * (hold[1] - h1) / dt
    assert abs(q[2] - q2) < 1e-12 * abs(q2)
    assert abs(q[1] - q1) < 1e-12 * abs(q1)
    # the deposit at 1 is fDep dt / a1 of the incoming flux, beside its erosion
    assert abs((1.0 + c1) * h[1] - c1 * h[0] - hold[1] - f1 * dt / area[1] * q2) < 1e-10


def _state(kind, size):
    m, h0, p = case(kind, size)
    o = oracle_model(m, h0, p)
    o.initialise()
    o.step()
    o.flowAccumulation()
    hold = o.h.copy()
    off, diag = o._eroMats(hold)
    return o, hold, off, diag


def test_zero_deposition_is_the_detachment_limited_sweep():
    o, hold, off, diag = _state("planar", 30)
    h, _ = orc.coupled_ed_solve(o.rcvIDi, off, diag, o.rcvIDi, o.wghtVali, np.zeros(o.lpoints), o.larea, o.dt,
                                hold)
    x, bad = orc.spl_solve(o.rcvIDi, off, diag, hold)
    assert bad == 0
    assert np.abs(h - x).max() <= 1e-11 * np.abs(x).max()


def test_flux_block_volume_balance():
    for kind, size in (("planar", 30), ("sphere", 4)):
        o, hold, off, diag = _state(kind, size)
        fdep = orc.transport_fdep(1.0, o.larea, o.FAL, o.seaID, o.outletIDs if o.flatModel else None)
        assert fdep.max() > 0.5  # the case deposits
        h, q = orc.coupled_ed_solve(o.rcvIDi, off, diag, o.rcvIDi, o.wghtVali, fdep, o.larea, o.dt, hold)
        N = o.lpoints
        passed = np.zeros(N)
        for i in range(N):
            for k in range(o.rcvIDi.shape[1]):
                if o.rcvIDi[i, k] != i:
                    passed[i] += o.wghtVali[i, k]
        eroded = np.sum(o.larea * (hold - h) / o.dt)
        leaving = np.sum(q * (1.0 - passed))
        assert abs(eroded - leaving) <= 1e-9 * np.sum(np.abs(o.larea * (hold - h) / o.dt))
        # something deposits (h rises somewhere off the sea) and erosion stays below the detachment-limited one
        x, _ = orc.spl_solve(o.rcvIDi, off, diag, hold)
        assert (h - x).min() >= -1e-9 * np.abs(hold).max()


def test_fdep_floor_is_the_capped_ratio():
    rng = np.random.default_rng(3)
    area = rng.uniform(1.0, 100.0, 500)
    pa = rng.uniform(0.0, 300.0, 500)
    pa[::17] = 0.0
    sea = np.nonzero(rng.random(500) < 0.1)[0]
    f = orc.transport_fdep(2.0, area, pa, sea)
    for i in range(500):
        want = 0.0 if (pa[i] == 0.0 or i in set(sea.tolist())) else min(2.0 * area[i] / pa[i], 0.99)
        assert abs(f[i] - want) <= 1e-15 * max(want, 1.0)


# ---------------------------------------------------------------------------
# non-linear transport-limited solve (oracle.nl_transport_solve,
# nlSPL._solveNL_ed, nlSPL.py:82-125, 255-320)
# ---------------------------------------------------------------------------

def test_nonlinear_three_vertex_chain_by_nested_roots():
    from scipy.optimize import brentq

    rcv = np.array([[0], [0], [1]], np.int32)
    wgt = np.array([[0.0], [1.0], [1.0]])
    dist = np.array([[1.0], [1.0], [2.0]])
    kbr = np.array([0.0, 3.0, 0.01])
    area = np.array([3.0, 2.0, 5.0])
    dt = 7.0
    f1 = 0.4
    fdep = np.array([0.0, f1, 0.0])
    hold = np.array([-10.0, 20.0, 50.0])
    n = 2.0
    x, its = orc.nl_transport_solve(rcv, wgt, dist, kbr, n, fdep, area, dt, hold)

    def x2_of(x1):  # vertex 2: x2 - hOld2 + kbr2 ((x2 - x1) / d2)^n = 0, x2 in [x1, hOld2]
        return brentq(lambda y: y - hold[2] + kbr[2] * (max(y - x1, 0.0) / dist[2, 0]) ** n, x1, hold[2],
                      xtol=1e-15, rtol=1e-15)

    def F1(x1):  # vertex 1 with the total flux leaving it (its own and vertex 2's eroded volume rate)
        qt = area[1] * (hold[1] - x1) / dt + area[2] * (hold[2] - x2_of(x1)) / dt
        return (x1 - hold[1]) * (1.0 - f1) + kbr[1] * (max(x1 - hold[0], 0.0) / dist[1, 0]) ** n \
            - f1 * dt * max(qt, 0.0) / area[1]

    x1 = brentq(F1, hold[0], hold[1], xtol=1e-15, rtol=1e-15)
    assert abs(x[0] - hold[0]) < 1e-13
    assert abs(x[1] - x1) < 1e-11 * abs(hold).max()
    assert abs(x[2] - x2_of(x1)) < 1e-11 * abs(hold).max()
    assert its <= 20


def test_nonlinear_without_deposition_is_the_vertex_by_vertex_solve():
    """G = 0: the Newton iteration with sparse-LU steps equals the oracle's
    independent vertex-by-vertex root finding (erodepSPLnl)."""
    m, h0, p = case("sphere", 5)
    o = oracle_model(m, h0, p, n=1.5)
    o.initialise()
    o.step()
    o.flowAccumulation()
    hold = o.h.copy()
    PA = o.FAL
    dh = (hold[:, None] - hold[o.rcvIDi]).max(axis=1)
    elim = np.divide(dh, dh + 1.0e-2, out=np.zeros_like(dh), where=dh != 0)
    kbr = o.K * (o.rainVal ** o.coeffd) * (o.dt * (PA ** o.spl_m) * elim)
    kbr[o.seaID] = 0.0
    x, _ = orc.nl_transport_solve(o.rcvIDi, o.wghtVali, o.distRcvi, kbr, 1.5, np.zeros(o.lpoints), o.larea, o.dt,
                                  hold)
    o.erodepSPLnl()
    assert np.abs(x - o.nl_x).max() <= 1e-12 * np.abs(hold).max()
