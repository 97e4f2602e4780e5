"""Continental sediment routing (nodep False) on the GPU against the oracle's
restatement of goSPL's SEDMesh (gospl/sed/sedplex.py).

Per phase from identical state: goSPL's step is discontinuous at the last
bit of its inputs (exact sea/land and filled-rim comparisons), and with
nodep False the second flowAccumulation reads the eroded surface, which two
correct implementations reach with last-bit differences; the oracle is
therefore given the GPU's eroded surface (and erosion rate) before that
second pass, as it is given the GPU's state at the start of each step.
Elevation, cumulative erosion-deposition, the sediment load and the deposit
then agree to 1e-9 relative (FP64 exact solves on both sides; per-pit sums in another order; the
lake-deposit diffusion solved by Jacobi sweeps to convergence here and by a
sparse direct solve in the oracle).  Cases: the closed sphere (stray-sink
deposits in place, large deep lakes through the non-linear diffusion path)
and a bumpy plane with draining borders (stray-sink sediment leaves), with
the lake-size thresholds lowered so that its small lakes take the diffusion
path too.
"""

from __future__ import annotations

import functools

import numpy as np
import pytest

from tests.helpers import case, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


@functools.lru_cache(maxsize=None)
def bumpy_plane(n=60, seed=2):
    m, h0, p = case("planar", n)
    rng = np.random.default_rng(seed)
    bump = rng.uniform(0.0, 25.0, m.N) * (m.outlet == 0)
    return m, h0 + bump, p


CASES = {
    "sphere5": lambda: (case("sphere", 5), {}),
    "sphere6": lambda: (case("sphere", 6), {}),
    "plane60": lambda: (bumpy_plane(), dict(nlPitVolume=1.0e5, nlPitDepth=2.0)),
}


def _models(name):
    from gospl_amd.model import Model

    (m, h0, p), kw = CASES[name]()
    g = Model(m, h0, tend=1e12, nodep=False, **p, **kw)
    o = oracle_model(m, h0, p, nodep=False, **kw)
    return m, g, o


@pytest.mark.parametrize("name", ["sphere5", "plane60", "sphere6"])
def test_sed_steps_match_oracle(name):
    import torch

    m, g, o = _models(name)
    print("sediment routing steps:", name, m.N, flush=True)
    g.initialise()
    o.initialise()
    deposited = 0
    for s in range(3):
        o.h = m.to_input_order(g.hLocal.cpu().numpy()).copy()
        o.cumED = m.to_input_order(g.cumED.cpu().numpy()).copy()
        o.tNow = g.tNow
        inp = lambda t: m.to_input_order(t.cpu().numpy())
        g.flowAccumulation()
        o.flowAccumulation()
        g.erodepSPL()
        o.erodepSPL()
        assert rel_err(inp(g.hLocal), o.h) < TOL, s
        o.h = inp(g.hLocal).copy()
        o.cumED = inp(g.cumED).copy()
        o.Eb = inp(g.Eb).copy()
        g.flowAccumulation()
        o.flowAccumulation()
        assert rel_err(inp(g.FAL), o.FAL) < TOL, s
        g.sedChange()
        o.sedChange()
        g.getHillslope()
        o.getHillslope()
        g.applyForces()
        o.applyForces()
        g.tNow += g.dt
        o.tNow += o.dt
        torch.cuda.synchronize()
        assert g.sed_passes == o.sed_passes, (s, g.sed_passes, o.sed_passes)
        assert rel_err(inp(g.hLocal), o.h) < TOL, (s, rel_err(inp(g.hLocal), o.h))
        hmax = np.abs(o.h).max()
        assert np.abs(inp(g.cumED) - o.cumED).max() <= TOL * hmax, s
        assert rel_err(inp(g.vSedLocal), o.vSedLocal) < TOL, (s, rel_err(inp(g.vSedLocal), o.vSedLocal))
        dep = inp(g.sed_deposit)
        assert np.abs(dep - o.sed_deposit).max() <= TOL * max(np.abs(o.sed_deposit).max(), 1.0), s
        deposited += int((o.sed_depo > 0).sum())
    assert deposited > 0
    g.close()


def test_nl_diffusion_matches_direct_solve():
    """gospl_nl_diffusion against the oracle's Picard form with sparse direct
    solves: a random deposit on part of the level-5 sphere."""
    import ctypes

    import torch

    from gospl_amd.model import Model, _ptr

    m, h0, p = case("sphere", 5)
    g = Model(m, h0, tend=1e12, nodep=False, **p)
    o = oracle_model(m, h0, p, nodep=False)
    rng = np.random.default_rng(7)
    mask_in = rng.uniform(size=m.N) < 0.6
    dh_in = np.where(mask_in, rng.uniform(0.0, 400.0, m.N), 0.0)
    for K, nsub, npic in ((10.0, 10, 2), (5.0e4, 3, 3)):
        o.picardSub, o.picardIts = nsub, npic
        o.h = m.to_input_order(h0).copy()
        want = o._diffuseImplicit(dh_in.copy(), mask_in, K)
        out = torch.zeros(m.N, dtype=torch.float64, device=g.device)
        sw = ctypes.c_int64(0)
        dh_d = g._dev(m.to_device_order(dh_in))
        mask_d = g._dev(m.to_device_order(mask_in).astype(np.int8), torch.int8)
        g._call("gospl_nl_diffusion", g.H, _ptr(dh_d), _ptr(mask_d), K, 1.0e-4, 0.05, g.dt, nsub,
                npic, _ptr(g.hLocal), _ptr(out), ctypes.byref(sw), g.stream)
        got = m.to_input_order(out.cpu().numpy())
        assert rel_err(got, want) < 1e-10, (K, rel_err(got, want))
        assert sw.value > 0
    g.close()
