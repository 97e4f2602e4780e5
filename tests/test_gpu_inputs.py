"""A model built from goSPL's YAML + npz input equals the same model built
from keyword arguments, bit for bit (the 200x200-vertex planar workload at a
reduced size), and the time-stamped forcing events switch as goSPL's do."""

from __future__ import annotations

import textwrap

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_yaml_model_matches_kwargs_model(tmp_path, monkeypatch):
    import torch

    monkeypatch.chdir(tmp_path)   # goSPL writes its outputs under the working directory

    from gospl_amd import _lib, workloads
    from gospl_amd.model import Model

    from gospl_amd import mesh as gmesh

    lib = _lib.load()
    m, h0, p = workloads.planar_case(lib, n=60, seed=4)
    v, c = gmesh.planar_jittered(60, 60, 500.0, 100.0, 4)  # the triangulation planar_case builds
    z = m.to_input_order(h0)
    up = m.to_input_order(p["upsub"])
    np.savez(tmp_path / "plane.npz", v=v, c=c, z=z)
    np.savez(tmp_path / "uplift.npz", u=up)
    (tmp_path / "plane.yml").write_text(textwrap.dedent("""\
        name: plane
        domain:
          npdata: ['plane', 'v', 'c', 'z']
          flowdir: 3
          bc: 'ffff'
          nodep: True
        time:
          start: 0.
          end: 100000.
          dt: 1000.
        spl:
          K: 1.e-5
          m: 0.5
        diffusion:
          hillslopeKa: 0.1
          hillslopeKm: 0.1
        sea:
          position: 0.
        climate:
          - start: 0.
            uniform: 1.
        tectonics:
          - start: 0.
            end: 100000.
            upsub: ['uplift', 'u']
        """))
    gy = Model(str(tmp_path / "plane.yml"), False, False)
    gk = Model(m, h0, tstart=0.0, tend=100000.0, **p)
    assert gy.dm.N == gk.dm.N and gy.flowDir == 3 and gy.K == 1e-5
    assert np.array_equal(gy.dm.mesh.perm, gk.dm.mesh.perm)
    gy.initialise()
    gk.initialise()
    for s in range(3):
        gy.step()
        gk.step()
        torch.cuda.synchronize()
        assert torch.equal(gy.hLocal, gk.hLocal), s
        assert torch.equal(gy.FAG, gk.FAL), s
    assert torch.equal(gy.seaID, torch.nonzero(gy.lFill <= gy.sealevel).flatten())
    gy.close()
    gk.close()


def test_forcing_events_switch(tmp_path, monkeypatch):
    import torch

    monkeypatch.chdir(tmp_path)

    from gospl_amd import _lib, mesh as gmesh
    from gospl_amd.model import Model

    v, c = gmesh.planar_jittered(20, 20, 500.0, 100.0, seed=2)
    z = np.random.default_rng(1).uniform(5, 15, len(v))
    np.savez(tmp_path / "m.npz", v=v, c=c, z=z)
    np.savez(tmp_path / "f.npz", r=np.full(len(v), 3.0), u=np.full(len(v), 2e-3))
    (tmp_path / "ev.yml").write_text(textwrap.dedent("""\
        domain:
          npdata: ['m', 'v', 'c', 'z']
          flowdir: 1
          bc: 'ffff'
          nodep: True
        time: {start: 0., end: 10000., dt: 1000.}
        spl: {K: 1.e-6}
        climate:
          - {start: 0., uniform: 1.}
          - {start: 2000., map: ['f', 'r']}
        tectonics:
          - {start: 1000., end: 3000., upsub: ['f', 'u']}
        """))
    g = Model(str(tmp_path / "ev.yml"))
    g.initialise()
    seen = []
    for _ in range(4):
        seen.append((g.tNow, float(g.rainVal.max()), None if g.upsub is None else float(g.upsub.max())))
        g.step()
    # forces are applied for the next step at the end of each step (tNow before it advances)
    assert seen[0][1] == 1.0 and seen[0][2] is None
    assert seen[2][1] == 1.0 and seen[2][2] == 2e-3   # applied at t = 1000
    assert seen[3][1] == 3.0 and seen[3][2] == 2e-3   # applied at t = 2000
    g.close()
