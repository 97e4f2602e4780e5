"""Pins of the oracle's non-linear stream power (n != 1, CPU only).

goSPL solves F(x) = x - hOld + Kbr S(x)^n = 0 with PETSc SNES (absent here).
The oracle solves it vertex by vertex; its answer is checked against the
definition, computed independently and vectorised: the residual, with S the
weighted positive slope of local_spl (functions.F90:1469-1502) and Kbr with
the elimiter of nlSPL.py:330-346, vanishes to rounding, and erosion only
lowers the surface.  For n = 1 the per-vertex roots also solve goSPL's system
as a plain linear system on the slopes that are positive at the solution
(a dense solve).
"""

import numpy as np
import pytest

from tests.helpers import case, oracle_model


def _residual(o, hOld, x):
    rcv, wgt, dist = o.rcvIDi, o.wghtVali, o.distRcvi
    xr = x[rcv]
    on = (wgt > 0.0) & (x[:, None] > xr)
    S = np.where(on, wgt * (x[:, None] - xr) / np.where(dist > 0, dist, 1.0), 0.0).sum(axis=1)
    dh = (hOld[:, None] - hOld[rcv]).max(axis=1)
    el = np.divide(dh, dh + 1.0e-2, out=np.zeros_like(dh), where=dh != 0)
    kb = o.K * (o.rainVal ** o.coeffd) * (o.dt * (o.FAL ** o.spl_m) * el)
    kb[o.seaID] = 0.0
    return x - hOld + kb * np.where(S > 0, S, 0.0) ** o.spl_n, kb, on, S


@pytest.mark.parametrize("kind,size,n", [("planar", 30, 1.5), ("planar", 30, 2.0), ("sphere", 4, 1.5),
                                         ("sphere", 4, 1.0)])
def test_nonlinear_spl_solves_its_definition(kind, size, n):
    m, h0, p = case(kind, size)
    o = oracle_model(m, h0, p, n=n)
    o.initialise()
    o.flowAccumulation()
    hOld = o.h.copy()
    o.erodepSPLnl()
    x = o.nl_x
    F, kb, on, S = _residual(o, hOld, x)
    scale = np.abs(hOld).max()
    assert np.abs(F).max() <= 1e-11 * scale
    assert np.all(x <= hOld)
    assert np.count_nonzero(x < hOld) > 0.2 * len(x)
    if n == 1.0:
        # at the solution: x_i (1 + kb_i sum_on w/d) - kb_i sum_on w/d x_r = hOld_i
        N = len(x)
        A = np.eye(N)
        for k in range(o.rcvIDi.shape[1]):
            d = np.where(o.distRcvi[:, k] > 0, o.distRcvi[:, k], 1.0)
            c = np.where(on[:, k], kb * o.wghtVali[:, k] / d, 0.0)
            np.add.at(A, (np.arange(N), np.arange(N)), c)
            np.add.at(A, (np.arange(N), o.rcvIDi[:, k]), -c)
        xs = np.linalg.solve(A, hOld)
        assert np.abs(xs - x).max() <= 1e-10 * scale
