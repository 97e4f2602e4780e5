"""The CPU baseline in reference-solver mode (goSPL's FGMRES / Richardson +
ILU(0) at rtol 1e-8 instead of exact solves) — CPU only.

Per step from identical state it must agree with the exact oracle to the
accuracy goSPL's tolerance allows (1e-6 relative on elevation and
discharge), every solve must report convergence, and the threaded run must
agree with the single-threaded one to the same bar.
"""

from __future__ import annotations

import numpy as np
import pytest

from gospl_amd import _lib, workloads
from oracle import oracle as orc


def _case(n=40):
    m, h0, p = workloads.planar_case(_lib.load(), n=n, seed=2)
    geom = workloads.oracle_geometry(m)
    conv = {k: (m.to_input_order(v) if isinstance(v, np.ndarray) else v) for k, v in p.items()}
    return geom, m.to_input_order(h0), conv


@pytest.mark.parametrize("threads", [1, 4])
def test_reference_solver_mode_matches_exact_oracle(threads):
    geom, h0, p = _case()
    ex = orc.OracleModel(geom, h0, tend=1e12, **p)
    rs = orc.ReferenceSolverModel(geom, h0, tend=1e12, threads=threads, **p)
    ex.initialise()
    rs.initialise()
    for s in range(3):
        rs.h = ex.h.copy()
        rs.cumED = ex.cumED.copy()
        rs.tNow = ex.tNow
        ex.step()
        rs.step()
        scale = np.abs(ex.h).max()
        assert np.abs(rs.h - ex.h).max() <= 1e-6 * scale, s
        assert np.abs(rs.FAL - ex.FAL).max() <= 1e-6 * np.abs(ex.FAL).max(), s
    kinds = {k for k, _, _ in rs.solver_its}
    assert {"flow", "spl", "hillslope"} <= kinds
    assert all(it >= 0 and rr <= 1e-8 for _, it, rr in rs.solver_its), rs.solver_its
