"""Pins of the CPU reference gradient (oracle/adjoint_ref.py), CPU only.

* its frozen forward, fed the flow data an OracleModel run recorded, repeats
  that run's elevations (it is the oracle's step with the data held);
* its gradient equals central finite differences of its own frozen forward
  at every vertex of small planar and spherical cases (brute force), for
  log10 K and the uplift rate, with sinks, sea cells and marine diffusion on
  the sphere.
"""

import numpy as np
import pytest

from gospl_amd import _lib, workloads
from oracle import adjoint_ref as ar
from oracle import oracle as orc


def _record(om, nsteps):
    """Run an OracleModel for nsteps (after initialise's forcing), recording
    each step's frozen data and the elevations."""
    steps, hs = [], []
    om.applyForces()
    for _ in range(nsteps):
        om.flowAccumulation()
        st = dict(rcv=om.rcvIDi.T.copy(), wgt=om.wghtVali.T.copy(), dist=om.distRcvi.T.copy(),
                  fal=om.FAL.copy(), lfill=om.lFill.copy(), lsink=np.asarray(om.lsink, bool).copy())
        om.erodepSPL()
        st["sea"] = om.h <= om.sealevel
        om.getHillslope()
        st["uplift"] = om.tNow < om.tEnd
        if st["uplift"]:
            om.applyForces()
        om.tNow += om.dt
        steps.append(st)
        hs.append(om.h.copy())
    return steps, hs


def _geo(m):
    g = workloads.oracle_geometry(m)
    return dict(ngb=g["ngb"], nnb=g["nnb"], elgt=g["elgt"], vdist=g["vdist"], area=g["area"],
                outlet=g["outlet"], flat=g["flat"])


def _case(kind):
    lib = _lib.load()
    if kind == "planar":
        m, h0, p = workloads.planar_case(lib, n=12, seed=3)
        up = m.to_input_order(p["upsub"])
        nsteps = 5
    else:
        m, h0, p = workloads.sphere_case(lib, level=3, seed=2, lmax=12)
        up = np.full(m.N, 5.0e-4)
        # diffusivities large enough for the coarse cells (about 800 km) that
        # the operator is far from the identity and land/sea rows differ
        p = dict(p, dt=20000.0, K=2.0e-5, cda=2.0e5, cdm=5.0e5)
        nsteps = 4
    geo = _geo(m)
    hin = m.to_input_order(h0)
    rain = p["rain"] if np.isscalar(p["rain"]) else m.to_input_order(p["rain"])
    om = orc.OracleModel(geo, hin, K=p["K"], m=p["m"], flowdir=3, cda=p["cda"], cdm=p["cdm"], dt=p["dt"],
                         rain=rain, upsub=up, sealevel=p["sealevel"], tend=1e12)
    steps, hs = _record(om, nsteps)
    pp = dict(m=p["m"], dt=p["dt"], cda=p["cda"], cdm=p["cdm"], sealevel=p["sealevel"], init_uplift=True)
    logk = np.full(m.N, np.log10(p["K"]))
    return hin, logk, up, steps, hs, geo, pp


@pytest.mark.parametrize("kind", ["planar", "sphere"])
def test_frozen_forward_repeats_the_oracle_run(kind):
    hin, logk, up, steps, hs, geo, pp = _case(kind)
    h = ar.frozen_forward(hin, logk, up, steps, geo, pp)
    scale = np.abs(hs[-1]).max()
    assert np.abs(h - hs[-1]).max() <= 1e-10 * scale


@pytest.mark.parametrize("kind", ["planar", "sphere"])
def test_reference_gradient_matches_finite_differences(kind):
    hin, logk, up, steps, hs, geo, pp = _case(kind)
    rng = np.random.default_rng(7)
    target = hs[-1] + rng.normal(0.0, 5.0, len(hin))
    hfin = ar.frozen_forward(hin, logk, up, steps, geo, pp)
    g_logk, g_up = ar.frozen_gradient(hin, logk, up, steps, geo, pp, 2.0 * (hfin - target))

    def dloss(a, b):  # L(a) - L(b) without cancelling the large sums
        return float(np.sum((a - b) * (a + b - 2.0 * target)))

    N = len(hin)
    idx = np.arange(N) if N <= 200 else rng.choice(N, 120, replace=False)
    for what, base, g, h in (("logk", logk, g_logk, 1.0e-4), ("uplift", up, g_up, 1.0e-7)):
        fd = np.empty(len(idx))
        for n, i in enumerate(idx):
            xp = base.copy()
            xm = base.copy()
            xp[i] += h
            xm[i] -= h
            a = ar.frozen_forward(hin, logk if what == "uplift" else xp, xp if what == "uplift" else up, steps,
                                  geo, pp)
            b = ar.frozen_forward(hin, logk if what == "uplift" else xm, xm if what == "uplift" else up, steps,
                                  geo, pp)
            fd[n] = dloss(a, b) / (2.0 * h)
        scale = np.abs(g[idx]).max()
        assert scale > 0.0
        err = np.abs(g[idx] - fd).max() / scale
        assert err <= 1e-5, (what, err)
        # the gradient is not trivially zero where it should not be
        assert np.count_nonzero(np.abs(fd) > 1e-6 * scale) >= len(idx) // 4
