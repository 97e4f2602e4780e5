"""Transport-limited stream power (G > 0) on the GPU against the oracle.

Bar: per step from identical state (the oracle given the GPU's elevation and
cumulative erosion), elevation, cumulative erosion/deposition and erosion
rate within 1e-9 relative of the oracle's sparse-LU solution of goSPL's
nested h / q system; the sediment flux of the solution (gospl_spl_dep's qs)
within 1e-9 of the oracle's q block; GMRES reaching a relative residual of
1e-13 in a bounded number of iterations.  G = 1 and G = 5 (deposition capped
at 0.99 on many vertices) on a sphere with sea and pits and on the planar
workload with draining borders.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from tests.helpers import case, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _sync_oracle(m, g, o):
    o.h = m.to_input_order(g.hLocal.cpu().numpy()).copy()
    o.cumED = m.to_input_order(g.cumED.cpu().numpy()).copy()
    o.tNow = g.tNow


@pytest.mark.parametrize("kind,size,G", [("sphere", 5, 1.0), ("planar", 40, 1.0), ("planar", 40, 5.0)])
def test_transport_limited_steps_match_oracle(kind, size, G):
    from gospl_amd.model import Model

    m, h0, p = case(kind, size)
    print("transport-limited SPL: %s %d, G = %g" % (kind, size, G), flush=True)
    g = Model(m, h0, tend=1e12, G=G, **p)
    o = oracle_model(m, h0, p, G=G)
    g.initialise()
    o.initialise()
    for s in range(3):
        _sync_oracle(m, g, o)
        g.step()
        o.step()
        assert g.ed_relres <= 1e-13 and 1 <= g.ed_iters <= 200, (g.ed_iters, g.ed_relres)
        h = m.to_input_order(g.hLocal.cpu().numpy())
        assert rel_err(h, o.h) < TOL, (s, rel_err(h, o.h))
        ed = m.to_input_order(g.cumED.cpu().numpy())
        assert np.abs(ed - o.cumED).max() <= TOL * np.abs(o.h).max(), s
        eb = m.to_input_order(g.EbLocal.cpu().numpy())
        assert np.abs(eb - o.EbLocal).max() <= TOL * np.abs(o.h).max() / o.dt, s
        print("  step %d: GMRES iterations %d, residual %.1e, rel err %.2e" %
              (s, g.ed_iters, g.ed_relres, rel_err(h, o.h)), flush=True)
    g.close()


def test_transport_limited_flux_block_through_c_abi():
    """gospl_spl_dep called directly: the q block of its solution against the
    oracle's, and its elevation against the oracle's h block."""
    from gospl_amd.model import Model, _ptr

    m, h0, p = case("sphere", 5)
    G = 2.0
    g = Model(m, h0, tend=1e12, **p)
    g.initialise()
    g.step()
    g.flowAccumulation()
    torch.cuda.synchronize()
    o = oracle_model(m, h0, p)
    o.initialise()
    _sync_oracle(m, g, o)
    o.flowAccumulation()
    hold = o.h.copy()
    off, diag = o._eroMats(hold)
    fdep = orc.transport_fdep(G, o.larea, o.FAL, o.seaID, None)
    hx, qx = orc.coupled_ed_solve(o.rcvIDi, off, diag, o.rcvIDi, o.wghtVali, fdep, o.larea, o.dt, hold)
    N = m.N
    h = g.hLocal.clone()
    cumed = torch.zeros(N, dtype=torch.float64, device=h.device)
    eb = torch.zeros_like(cumed)
    ebl = torch.zeros_like(cumed)
    qs = torch.zeros_like(cumed)
    it = ctypes.c_int32(0)
    rel = ctypes.c_double(0.0)
    g._call("gospl_spl_dep", g.H, g.flowDir, _ptr(g.rcvIDi), _ptr(g.wghtVali), _ptr(g.distRcvi), _ptr(g.FAL),
            _ptr(g._kfac), g.K, g.spl_m, G, g.dt, g.sealevel, _ptr(g.lFill), _ptr(g.lsink), _ptr(h), _ptr(cumed),
            _ptr(eb), _ptr(ebl), 1e-13, 400, ctypes.byref(it), ctypes.byref(rel), _ptr(qs), g.stream)
    torch.cuda.synchronize()
    print("gospl_spl_dep: %d iterations, residual %.1e" % (it.value, rel.value), flush=True)
    assert rel.value <= 1e-13
    # h after the update keeps hOld at lsink cells; compare the solve's h where it is applied
    keep = o.lsink.copy()
    hg = m.to_input_order(h.cpu().numpy())
    assert np.abs(hg[~keep] - hx[~keep]).max() <= TOL * np.abs(hx).max()
    qg = m.to_input_order(qs.cpu().numpy())
    assert np.abs(qg - qx).max() <= TOL * np.abs(qx).max(), np.abs(qg - qx).max() / np.abs(qx).max()
    # argument checks
    with pytest.raises(RuntimeError):
        g._call("gospl_spl_dep", g.H, g.flowDir, _ptr(g.rcvIDi), _ptr(g.wghtVali), _ptr(g.distRcvi), _ptr(g.FAL),
                _ptr(g._kfac), g.K, g.spl_m, 0.0, g.dt, g.sealevel, _ptr(g.lFill), _ptr(g.lsink), _ptr(h),
                _ptr(cumed), _ptr(eb), _ptr(ebl), 1e-13, 400, None, None, None, g.stream)
    g.close()


@pytest.mark.parametrize("kind,size,n,G", [("sphere", 5, 1.5, 1.0), ("planar", 40, 2.0, 2.0)])
def test_nonlinear_transport_limited_steps_match_oracle(kind, size, n, G):
    """n != 1 with G > 0 (nlSPL._solveNL_ed): Newton steps of GMRES solves
    against the oracle's Newton steps of sparse-LU solves of the two-field
    system, per step from identical state, 1e-9."""
    from gospl_amd.model import Model

    m, h0, p = case(kind, size)
    print("non-linear transport-limited SPL: %s %d, n = %g, G = %g" % (kind, size, n, G), flush=True)
    g = Model(m, h0, tend=1e12, n=n, G=G, **p)
    o = oracle_model(m, h0, p, n=n, G=G)
    g.initialise()
    o.initialise()
    for s in range(3):
        _sync_oracle(m, g, o)
        g.step()
        o.step()
        assert 1 <= g.nl_iters <= 30, g.nl_iters
        assert g.ed_iters <= 40 * g.nl_iters, (g.ed_iters, g.nl_iters)
        h = m.to_input_order(g.hLocal.cpu().numpy())
        assert rel_err(h, o.h) < TOL, (s, rel_err(h, o.h))
        ed = m.to_input_order(g.cumED.cpu().numpy())
        assert np.abs(ed - o.cumED).max() <= TOL * np.abs(o.h).max(), s
        print("  step %d: Newton steps %d (oracle %d), GMRES iterations %d, rel err %.2e" %
              (s, g.nl_iters, o.nl_iters, g.ed_iters, rel_err(h, o.h)), flush=True)
    g.close()
