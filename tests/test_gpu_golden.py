"""HIP path vs goSPL's own Fortran kernels, directly.

tests/golden/*.npz hold inputs and the outputs of the reference Fortran
(mfdreceivers, the one-rank fill chain, label_pits, fill_dir + nghb_dir,
getpitvol, sethillslopecoeff; see make_golden.py).  Here the device entry
points are run on the same inputs and compared with those outputs without
going through the oracle: receivers, distances, filled elevations, pit ids,
flat-routing distances and hillslope coefficients bit for bit; receiver
weights to the last bit or two (the device pow may differ from the host's by
1 ulp); depth-volume sums to 1e-12 (summation order).  One column is the
oracle's: the spill cells (goSPL selects them in numpy, pitfilling.py:
756-801, so make_golden.py writes the oracle's choice, itself checked
against the definition by brute force in test_oracle_glue_pins.py); the
Fortran fill_dir/nghb_dir distances stored beside them start from those
cells.

On a flat mesh goSPL replaces the Voronoi areas by the mesh's control volumes
when they differ by more than 1e2 m2 (unstructuredmesh.py:172-188), and the
30x30 plane does.  There the area-weighted outputs (depth volumes, hillslope
coefficients) are compared with the oracle's restatement of getpitvol and
sethillslopecoeff evaluated on the swapped areas; both restatements equal the
Fortran bit for bit on the stored areas (test_oracle_pins).
"""

from __future__ import annotations

import ctypes
import glob
import os

import numpy as np
import pytest

from tests.helpers import dev_index_to_input, rel_err

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
IDS = [os.path.basename(p)[:-4] for p in GOLD]


def _setup(path):
    import torch

    from gospl_amd import _lib, mesh as gmesh
    from gospl_amd.model import DeviceMesh

    g = dict(np.load(path))
    m = gmesh.build_mesh(g["v"], g["c"].astype(np.int32), _lib.load(), flat=bool(g["flat"]))
    assert np.array_equal(m.to_input_order(m.outlet), g["outlet"])
    dm = DeviceMesh(m, "cuda")
    return g, m, dm, torch


def _rows(m, t):
    a = t.detach().cpu().numpy().T
    return m.to_input_order(a)


@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_receivers_match_fortran(path):
    from gospl_amd.model import _ptr

    g, m, dm, torch = _setup(path)
    lib = dm.lib
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for t in range(2):
        h = torch.as_tensor(m.to_device_order(g["mfd%d_h" % t]), device="cuda")
        rcv = torch.zeros((3, m.N), dtype=torch.int32, device="cuda")
        dist = torch.zeros((3, m.N), dtype=torch.float64, device="cuda")
        wgt = torch.zeros((3, m.N), dtype=torch.float64, device="cuda")
        assert lib.gospl_mfdreceivers(dm.handle, 3, 1.1, _ptr(h), 0.0, _ptr(rcv), _ptr(dist), _ptr(wgt), s) == 0
        torch.cuda.synchronize()
        r = dev_index_to_input(m, _rows(m, rcv))
        assert np.array_equal(r, g["mfd%d_rcv" % t]), np.count_nonzero(r != g["mfd%d_rcv" % t])
        assert np.array_equal(_rows(m, dist), g["mfd%d_dist" % t])
        w = _rows(m, wgt)
        ref = g["mfd%d_wgt" % t]
        assert np.abs(w - ref).max() <= 4.5e-16 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_fill_matches_fortran(path):
    from gospl_amd.model import _ptr

    g, m, dm, torch = _setup(path)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    h = torch.as_tensor(m.to_device_order(g["z"]), device="cuda")
    lf = torch.zeros_like(h)
    level = float(g["level"])
    assert dm.lib.gospl_fill_depressions(dm.handle, _ptr(h), level, -1.0e30, _ptr(lf), s) == 0
    torch.cuda.synchronize()
    got = m.to_input_order(lf.cpu().numpy())
    want = g["fill"] + level  # the Fortran chain fills z - level
    assert np.array_equal(got, want), np.count_nonzero(got != want)


@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_pits_and_flat_routing_match_fortran(path):
    from gospl_amd.model import Model

    g, m, dm, torch = _setup(path)
    mod = Model(dm, m.to_device_order(g["z"]), K=1e-5, sealevel=0.0)
    mod.fillElevation()
    assert np.array_equal(m.to_input_order(mod.lFill.cpu().numpy()), g["lfill"])
    pit = m.to_input_order(mod.pitIDs.cpu().numpy())
    assert np.array_equal(pit, g["pitids"])
    sp = np.unique(m.perm[np.nonzero(mod.lspill.cpu().numpy())[0]])
    assert np.array_equal(sp, np.unique(g["spill"]))
    # fill_dir + nghb_dir, relaxed: the breadth-first distance from the spill cells
    pt = m.to_input_order(mod.ptdir.cpu().numpy())
    assert np.array_equal(pt, g["ptdir_relax"]), np.count_nonzero(pt != g["ptdir_relax"])
    # getpitvol on the first four depth levels
    assert mod.filled_vol.shape[0] == g["pitvol"].shape[0]
    assert np.array_equal(mod.filled_lvl[:, :4], g["pitlvl"])
    area = m.to_input_order(m.area)
    want = g["pitvol"]
    if not np.array_equal(area, g["area"]):
        from oracle import oracle as orc

        want = orc.getpitvol(g["pitlvl"], np.maximum(g["z"], 0.0), g["pitids"],
                             np.ones(m.N, np.int32), area)
    assert rel_err(mod.filled_vol[:, :4], want) < 1e-12


@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_hillslope_coefficients_match_fortran(path):
    from gospl_amd.model import _ptr

    g, m, dm, torch = _setup(path)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    kd = torch.as_tensor(m.to_device_order(g["kd"]), device="cuda")
    out = torch.zeros((13, m.N), dtype=torch.float64, device="cuda")
    assert dm.lib.gospl_hillslope_coeff(dm.handle, _ptr(kd), _ptr(out), s) == 0
    torch.cuda.synchronize()
    got = _rows(m, out)
    want = g["hillcoeff"]
    area = m.to_input_order(m.area)
    if not np.array_equal(area, g["area"]):
        from oracle import oracle as orc

        want = orc.sethillslopecoeff(g["kd"], g["ngb"], g["nnb"], area, g["elgt"], g["vdist"])
    assert np.array_equal(got, want), np.count_nonzero(got != want)


@pytest.mark.parametrize("path", GOLD, ids=IDS)
def test_runs_are_bit_reproducible(path):
    """Two runs from the same state give identical bits (no floating-point
    atomics anywhere on the path)."""
    from gospl_amd.model import Model

    g, m, dm, torch = _setup(path)
    z = m.to_device_order(g["z"])
    p = dict(K=4e-6, m=0.5, flowdir=3, cda=0.01, cdm=0.2, dt=100.0, rain=1.0, sealevel=0.0)
    out = []
    for _ in range(2):
        mod = Model(dm, z, **p, tend=1e12)
        mod.initialise()
        for _ in range(3):
            mod.step()
        out.append([t.cpu().numpy().copy() for t in (mod.hLocal, mod.FAL, mod.cumED, mod.waterFilled)]
                   + [mod.pitParams.copy(), mod.filled_vol.copy()])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
