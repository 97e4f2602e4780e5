"""Ensembles on one union mesh (gospl_amd.ensemble) against separate runs.

Bar: every member of an EnsembleModel -- spheres of different sizes (levels
3, 4, 5), each with its own seed, K, m and rain scale, plus one whose land
has no pits -- follows its
own single-mesh Model (refined mode) bit for bit: elevation, discharge,
filled surface, cumulative erosion, water and fillFAL after every step, with
pit-overflow cascades running in the members.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 6


def _members():
    from gospl_amd import _lib, workloads

    rng = np.random.default_rng(42)
    out = []
    for level, seed in ((4, 3), (3, 5), (5, 7)):
        m, h0, p = workloads.sphere_case(_lib.load(), level=level, seed=seed, lmax=min(64, 4 * 2 ** level))
        p = dict(p)
        p["K"] = float(10.0 ** rng.uniform(-6.0, -5.0))
        p["m"] = float(rng.uniform(0.3, 0.6))
        p["rain"] = np.asarray(p["rain"]) * float(rng.uniform(0.5, 2.0))
        out.append((m, h0, p))
    # a member with no land pits (a tilted surface: one land cap with its
    # maximum at the pole), so its cascade never runs while the others' do
    m, h0, p = workloads.sphere_case(_lib.load(), level=3, seed=9, lmax=8)
    z = m.coords[:, 2] / np.abs(m.coords[:, 2]).max()
    out.append((m, 2000.0 * z - 300.0, dict(p, K=3.0e-6, m=0.4)))
    return out


def test_members_match_single_runs_bit_for_bit():
    import torch

    from gospl_amd.ensemble import EnsembleModel
    from gospl_amd.model import Model

    members = _members()
    print("ensemble: members of %s vertices, %d steps" % ([m.N for m, _, _ in members], STEPS), flush=True)
    ens = EnsembleModel(members, tend=1e12)
    singles = [Model(m, h0, tend=1e12, refine=True, **p) for m, h0, p in members]
    ens.initialise()
    for s in singles:
        s.initialise()
    fields = ("hLocal", "FAL", "fillFAL", "lFill", "cumED", "waterFilled", "EbLocal")
    casc = 0
    for step in range(STEPS):
        ens.step()
        for s in singles:
            s.step()
        casc += int(ens.member_cascade_passes.sum())
        for k, s in enumerate(singles):
            assert ens.member_cascade_passes[k] == s.cascade_passes, (step, k)
            for f in fields:
                a = ens.member(k, getattr(ens, f))
                b = getattr(s, f)
                if not torch.equal(a, b):
                    d = float((a - b).abs().max())
                    pytest.fail("step %d member %d: %s differs (max |diff| %.3e)" % (step, k, f, d))
        print("  step %d: identical (cascade passes %s)" % (step, list(ens.member_cascade_passes)), flush=True)
    assert casc > 0, "no member ran a pit-overflow cascade: the test would not cover it"
    assert ens.member_cascade_passes[-1] == 0 and ens.member_cascade_passes[:-1].min() > 0


def test_dag_launch_shape_does_not_change_results():
    """The dataflow sweeps' launch shape (the autotuner's knob) changes their
    time only: a level-6 sphere stepped with the default shape and with
    four workgroups per compute unit, pool hand-off at 64 and 128 pollers
    gives the same bits."""
    import torch

    from gospl_amd import _lib, autotune, workloads
    from gospl_amd.model import Model

    m, h0, p = workloads.sphere_case(_lib.load(), level=6, seed=2, lmax=32)
    runs = []
    for shape in (None, "wide"):
        g = Model(m, h0, tend=1e12, refine=True, **p)
        if shape == "wide":
            sh = autotune.get_shape(g)
            autotune.set_shape(g, min(4 * sh["num_cu"], sh["resident"]), 64, 128)
            assert autotune.get_shape(g)["chain_blocks"] > sh["num_cu"]
        g.initialise()
        for _ in range(4):
            g.step()
        runs.append((g.hLocal.clone(), g.FAL.clone()))
    assert torch.equal(runs[0][0], runs[1][0]) and torch.equal(runs[0][1], runs[1][1])
