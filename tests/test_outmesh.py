"""goSPL's output files and restart read-back (gospl_amd.outmesh), CPU only.

The writer runs on a stand-in model holding CPU tensors.  The files are read
back by an independent reader where the image has one (the h5py that goSPL
itself uses, under the Conda Python), and the XMF/XDMF are parsed as XML.
The GPU test of a whole run with outputs and a restart is in
test_gpu_outputs.py."""

from __future__ import annotations

import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest
import torch

from gospl_amd import _lib, h5, mesh as gmesh
from gospl_amd.outmesh import DISCHARGE_FLOOR, WriteMesh

CONDA_PY = "/opt/conda/bin/python3.9"


def _have_hdf5():
    try:
        h5.lib()
        return True
    except RuntimeError:
        return False


pytestmark = pytest.mark.skipif(not _have_hdf5(), reason="no HDF5 C library in this image")


class _Dev:
    def __init__(self, mesh):
        self.mesh = mesh


class Stand(WriteMesh):
    """The attributes WriteMesh reads from goSPL's Model."""

    def __init__(self, mesh, out, rstep=0, tout=2000.0):
        N = mesh.N
        self.dm = _Dev(mesh)
        self.device = torch.device("cpu")
        self.lpoints = N
        self.tStart, self.tEnd, self.dt, self.tNow = 0.0, 10000.0, 1000.0, 0.0
        self.sealevel = 0.5
        self.verbose = False
        rng = np.random.default_rng(3)
        f = lambda: torch.as_tensor(rng.normal(0.0, 1.0, N))
        self.hLocal, self.cumED, self.EbLocal, self.waterFilled = f(), f(), f(), f()
        self.FAL = torch.as_tensor(np.abs(rng.normal(0.0, 1e-7, N)))
        self.fillFAL = torch.as_tensor(np.abs(rng.normal(0.0, 1.0, N)))
        self.vSedLocal = torch.zeros(N, dtype=torch.float64)
        self.lFill = self.hLocal.clone()
        self.upsub = None
        self.rainVal = torch.full((N,), 2.0, dtype=torch.float64)
        self._init_output(outputDir=str(out), tout=tout, rstep=rstep)


def _mesh():
    v, c = gmesh.planar_jittered(9, 7, 500.0, 100.0, seed=5)
    return gmesh.build_mesh(v, c, _lib.load(), flat=True), v, c


def test_output_schedule_files_and_values(tmp_path):
    m, v, c = _mesh()
    s = Stand(m, tmp_path / "out")
    written = []
    for k in range(5):                   # t = 0, 1000, ..., 4000 with tout = 2000
        before = s.outputStep
        s.visModel()
        if s.outputStep != before:
            written.append((s.tNow, before))
        s.tNow += s.dt
    assert written == [(0.0, 0), (2000.0, 1), (4000.0, 2)]
    out = tmp_path / "out"
    assert sorted(os.listdir(out / "h5")) == ["gospl.0.p0.h5", "gospl.1.p0.h5", "gospl.2.p0.h5", "topology.p0.h5"]
    with h5.File(out / "h5" / "topology.p0.h5") as f:
        assert np.array_equal(f.read("coords", np.float32), v.astype(np.float32))
        assert np.array_equal(f.read("cells", np.int32), c + 1)
    inp = m.to_input_order
    with h5.File(out / "h5" / "gospl.2.p0.h5") as f:
        for name, t in (("elev", s.hLocal), ("erodep", s.cumED), ("EDrate", s.EbLocal),
                        ("waterFill", s.waterFilled), ("rain", s.rainVal)):
            assert np.array_equal(f.read(name, np.float32)[:, 0], inp(t.numpy()).astype(np.float32)), name
        fa = inp(s.FAL.numpy()).copy()
        fa[fa <= DISCHARGE_FLOOR] = DISCHARGE_FLOOR
        fa[inp(s.lFill.numpy()) <= s.sealevel] = 1.0       # the sea gets 1 (outmesh.py:346-354)
        assert np.array_equal(f.read("FA", np.float32)[:, 0], fa.astype(np.float32))
        assert np.all(f.read("sedLoad", np.float32) == np.float32(DISCHARGE_FLOOR))
        assert "uplift" not in f and "rain" in f
    # the time series includes every per-output XMF, each pointing at its files
    root = ET.parse(out / "gospl.xdmf").getroot()
    inc = [e.get("href") for e in root.iter("{http://www.w3.org/2001/XInclude}include")]
    assert inc == ["xmf/gospl0.xmf", "xmf/gospl1.xmf", "xmf/gospl2.xmf"]
    x = ET.parse(out / "xmf" / "gospl2.xmf").getroot()
    assert x.find(".//Time").get("Value") == "4000.00"
    items = {a.get("Name"): a.find("DataItem").text for a in x.iter("Attribute") if a.get("Name") != "sea"}
    assert items["Z"] == "h5/gospl.2.p0.h5:/elev" and items["SL"] == "h5/gospl.2.p0.h5:/sedLoad"
    assert x.find(".//Topology").get("NumberOfElements") == str(len(c))


@pytest.mark.skipif(not os.path.exists(CONDA_PY), reason="no Conda Python (h5py) in this image")
def test_files_read_by_h5py(tmp_path):
    """goSPL reads and writes these files with h5py: the same datasets,
    shapes, types, gzip filter and values through it."""
    m, v, c = _mesh()
    s = Stand(m, tmp_path / "out")
    s.visModel()
    code = ("import h5py, json, numpy as np\n"
            "f = h5py.File('%s', 'r'); t = h5py.File('%s', 'r')\n"
            "print(json.dumps({'keys': sorted(f.keys()), 'shape': list(f['elev'].shape), 'dtype': str(f['elev'].dtype),"
            " 'comp': f['elev'].compression, 'elev': f['elev'][:, 0].astype(float).tolist(),"
            " 'cells': t['cells'][:].tolist(), 'cdtype': str(t['cells'].dtype)}))\n"
            % (tmp_path / "out" / "h5" / "gospl.0.p0.h5", tmp_path / "out" / "h5" / "topology.p0.h5"))
    r = subprocess.run([CONDA_PY, "-c", code], capture_output=True, text=True, env={"PATH": "/usr/bin:/bin"})
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("the Conda Python has no h5py")
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["keys"] == ["EDrate", "FA", "elev", "erodep", "fillFA", "rain", "sedLoad", "waterFill"]
    assert d["shape"] == [m.N, 1] and d["dtype"] == "float32" and d["comp"] == "gzip"
    assert np.array_equal(np.array(d["elev"], np.float32), m.to_input_order(s.hLocal.numpy()).astype(np.float32))
    assert d["cdtype"] == "int32" and np.array_equal(np.array(d["cells"]), c + 1)


def test_restart_reads_state_back(tmp_path):
    m, v, c = _mesh()
    s = Stand(m, tmp_path / "out")
    for k in range(3):
        s.visModel()
        s.tNow += s.dt
    r = Stand(m, tmp_path / "out", rstep=1)        # goSPL: tNow = start + rstep * tout, next output = rstep + 1
    assert r.tNow == 2000.0 and r.outputStep == 2 and r.saveTime == 4000.0
    r.readData()
    f32 = lambda t: torch.as_tensor(t.numpy().astype(np.float32).astype(np.float64))
    assert torch.equal(r.hLocal, f32(s.hLocal)) and torch.equal(r.cumED, f32(s.cumED))
    assert torch.equal(r.EbLocal, f32(s.EbLocal))
    with pytest.raises(ValueError, match="Restart file is missing"):
        Stand(m, tmp_path / "out", rstep=7).readData()


def test_makedir_picks_a_free_folder(tmp_path):
    m, v, c = _mesh()
    os.makedirs(tmp_path / "run")
    os.makedirs(tmp_path / "run_0")
    s = Stand.__new__(Stand)
    Stand.__init__(s, m, tmp_path / "x")
    s.outputDir, s.rStep, s.makedir, s.finput = str(tmp_path / "run"), 0, True, None
    s._createOutputDir()
    assert s.outputDir == str(tmp_path / "run_1") and os.path.isdir(tmp_path / "run_1" / "h5")
