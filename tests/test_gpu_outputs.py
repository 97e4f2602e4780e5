"""Model outputs and restart on the GPU (goSPL's visModel / readData,
gospl/tools/outmesh.py:89-104, 256-625; gospl/model.py:284-295, 347-370).

A run from a YAML file writes an output every ``tout``; each file holds the
float32 values of the model state at that time (checked against a keyword
model stepped alongside).  A restart from output 1 (``time.rstep: 1``)
continues bit for bit like a model whose state was set to those float32
values at the same time."""

from __future__ import annotations

import os
import textwrap

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(tmp_path, rstep=0):
    from gospl_amd import _lib, mesh as gmesh, workloads

    lib = _lib.load()
    m, h0, p = workloads.planar_case(lib, n=40, seed=6)
    v, c = gmesh.planar_jittered(40, 40, 500.0, 100.0, 6)
    np.savez(tmp_path / "plane.npz", v=v, c=c, z=m.to_input_order(h0))
    np.savez(tmp_path / "uplift.npz", u=m.to_input_order(p["upsub"]))
    (tmp_path / "run.yml").write_text(textwrap.dedent("""\
        domain:
          npdata: ['plane', 'v', 'c', 'z']
          flowdir: 3
          bc: 'ffff'
          nodep: True
        time: {start: 0., end: 6000., dt: 1000., tout: 2000., rstep: %d}
        spl: {K: 1.e-5, m: 0.5}
        diffusion: {hillslopeKa: 0.1, hillslopeKm: 0.1}
        sea: {position: 0.}
        climate:
          - {start: 0., uniform: 1.}
        tectonics:
          - {start: 0., end: 6000., upsub: ['uplift', 'u']}
        output: {dir: '%s'}
        """ % (rstep, tmp_path / "out")))
    return m, h0, p


def test_outputs_hold_the_state_and_restart_continues(tmp_path):
    import torch

    from gospl_amd import h5
    from gospl_amd.model import Model

    m, h0, p = _case(tmp_path)
    print("outputs: planar 40x40, 6 steps, tout = 2 steps", flush=True)
    ga = Model(str(tmp_path / "run.yml"))
    ga.initialise()
    ga.runProcesses()
    assert ga.outputStep == 4 and ga.tNow == 6000.0
    out = tmp_path / "out"
    assert sorted(os.listdir(out / "h5")) == ["gospl.%d.p0.h5" % k for k in range(4)] + ["topology.p0.h5"]

    # a keyword model stepped alongside: each output holds its state in float32
    gb = Model(m, h0, tstart=0.0, tend=6000.0, outputDir=str(tmp_path / "outb"), **p)
    gb.initialise()
    snaps = {}
    for k in range(7):
        if k % 2 == 0:
            sea = gb._sea_mask()
            snaps[k // 2] = {n: m.to_input_order(t.cpu().numpy()).astype(np.float32) for n, t in
                             (("elev", gb.hLocal), ("erodep", gb.cumED), ("EDrate", gb.EbLocal),
                              ("waterFill", gb.waterFilled))}
            fa = gb.FAL.clone()
            fa[fa <= 1e-8] = 1e-8
            fa[sea] = 1.0
            snaps[k // 2]["FA"] = m.to_input_order(fa.cpu().numpy()).astype(np.float32)
            snaps[k // 2]["state"] = (gb.hLocal.clone(), gb.cumED.clone(), gb.EbLocal.clone())
        if k < 6:
            gb.step()
    for k, want in snaps.items():
        with h5.File(out / "h5" / ("gospl.%d.p0.h5" % k)) as f:
            for name in ("elev", "erodep", "EDrate", "waterFill", "FA"):
                assert np.array_equal(f.read(name, np.float32)[:, 0], want[name]), (k, name)

    # restart from output 1 (t = 2000) against a model set to its float32 state
    _case(tmp_path, rstep=1)
    gc = Model(str(tmp_path / "run.yml"))
    assert gc.tNow == 2000.0 and gc.outputStep == 2
    gc.initialise()
    f32 = lambda t: t.to(torch.float32).to(torch.float64)
    h, ced, eb = snaps[1]["state"]
    assert torch.equal(gc.hLocal, f32(h)) and torch.equal(gc.cumED, f32(ced))
    gd = Model(m, h0, tstart=0.0, tend=6000.0, **p)
    gd.tNow = 2000.0
    gd.applyForces()
    gd.hLocal = f32(h).contiguous()
    gd.cumED = f32(ced)
    gd.EbLocal = f32(eb)
    gd.flowAccumulation()
    gc.runProcesses()
    while gd.tNow < gd.tEnd:
        gd.step()
    torch.cuda.synchronize()
    assert gc.tNow == gd.tNow == 6000.0
    assert torch.equal(gc.hLocal, gd.hLocal) and torch.equal(gc.cumED, gd.cumED)
    with h5.File(out / "h5" / "gospl.3.p0.h5") as f:
        assert np.array_equal(f.read("elev", np.float32)[:, 0],
                              m.to_input_order(gd.hLocal.cpu().numpy()).astype(np.float32))
    for g in (ga, gb, gc, gd):
        g.close()
