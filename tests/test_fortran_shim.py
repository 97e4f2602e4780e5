"""gospl_amd.fortran, the drop-in for goSPL's f2py kernels, called the way
goSPL calls gospl._fortran (unstructuredmesh.py:155-188, flowplex.py:498,
hillslope.py:171-180) and compared with the golden vectors written by goSPL's
own Fortran."""

from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from gospl_amd import fortran, mesh as gmesh

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
IDS = [os.path.basename(p)[:-4] for p in GOLD]


def _definetin(g):
    flat = bool(g["flat"])
    vi = gmesh.voronoi_inputs(g["v"], g["c"].astype(np.int32), flat)
    fortran.setcurvedmesh(0 if flat else 1)
    return fortran.definetin(g["v"], g["c"], vi["cells_edges"], vi["edges"], vi["circumcenters"].T)


@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_definetin(path):
    g = dict(np.load(path))
    ngbid, narea = _definetin(g)
    assert ngbid.shape == (len(g["v"]), 12) and ngbid.dtype == np.int32
    assert np.array_equal(ngbid, g["ngb"])
    assert np.array_equal(np.isnan(narea), np.isnan(g["narea"]))
    ok = ~np.isnan(narea)
    assert np.array_equal(narea[ok], g["narea"][ok])


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_mfdreceivers_and_hillslope_coefficients(path):
    g = dict(np.load(path))
    _definetin(g)
    N = len(g["v"])
    if bool(g["flat"]):
        # goSPL swaps in the control volumes when the Voronoi areas are off
        # (unstructuredmesh.py:172-188); the golden coefficients used the
        # stored areas
        fortran.updatearea(g["area"])
    gid = np.arange(N, dtype=np.int32)
    for t in range(2):
        rcv, dist, wgt = fortran.mfdreceivers(3, 1.1, g["mfd%d_h" % t], 0.0, gid)
        assert rcv.shape == (N, 3) and rcv.flags["F_CONTIGUOUS"]
        assert np.array_equal(rcv, g["mfd%d_rcv" % t])
        assert np.array_equal(dist, g["mfd%d_dist" % t])
        ref = g["mfd%d_wgt" % t]
        assert np.abs(wgt - ref).max() <= 4.5e-16 * max(1.0, np.abs(ref).max())
    hc = fortran.sethillslopecoeff(N, g["kd"])
    assert hc.shape == (N, 13)
    assert np.array_equal(hc, g["hillcoeff"])
    fortran.close()
