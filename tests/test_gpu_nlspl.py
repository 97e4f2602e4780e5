"""Non-linear stream power (n != 1) on the GPU against the oracle.

Bar: per step from identical state (the oracle given the GPU's elevation and
cumulative erosion), elevation, discharge, cumulative erosion and erosion
rate within 1e-9 relative of the oracle's vertex-by-vertex exact solve, over
4 steps with n = 1.5 and n = 2, on a sphere with sea and pits and on the
planar workload; Newton converges in a few exact sweeps.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import case, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.mark.parametrize("kind,size,n", [("sphere", 5, 1.5), ("planar", 40, 2.0)])
def test_nonlinear_spl_steps_match_oracle(kind, size, n):
    from gospl_amd.model import Model

    m, h0, p = case(kind, size)
    print("non-linear SPL: %s %d, n = %g" % (kind, size, n), flush=True)
    g = Model(m, h0, tend=1e12, n=n, **p)
    o = oracle_model(m, h0, p, n=n)
    g.initialise()
    o.initialise()
    for s in range(4):
        o.h = m.to_input_order(g.hLocal.cpu().numpy()).copy()
        o.cumED = m.to_input_order(g.cumED.cpu().numpy()).copy()
        o.tNow = g.tNow
        g.step()
        o.step()
        assert 1 <= g.nl_iters <= 30, g.nl_iters
        h = m.to_input_order(g.hLocal.cpu().numpy())
        assert rel_err(h, o.h) < TOL, (s, rel_err(h, o.h))
        assert rel_err(m.to_input_order(g.FAL.cpu().numpy()), o.FAL) < TOL, s
        ed = m.to_input_order(g.cumED.cpu().numpy())
        assert np.abs(ed - o.cumED).max() <= TOL * np.abs(o.h).max(), s
        eb = m.to_input_order(g.EbLocal.cpu().numpy())
        assert np.abs(eb - o.EbLocal).max() <= TOL * np.abs(o.h).max() / o.dt, s
        print("  step %d: Newton steps %d, rel err %.2e" % (s, g.nl_iters, rel_err(h, o.h)), flush=True)
