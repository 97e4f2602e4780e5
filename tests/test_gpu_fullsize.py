"""The benchmark workload itself: the 2,621,442-vertex sphere.

Per-step parity with the oracle from identical state at full size (the
oracle's exact sweeps take a few seconds per step here), plus properties
that hold whatever the size: the filled surface is never below the surface
and equals it at sea, receivers with weight are strictly lower and weights
sum to one, the flow solve's residual vanishes and its discharge conserves
mass (everything that rains ends in a sink), and two runs give the same
bits.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.helpers import oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.fixture(scope="module")
def full():
    import torch

    from gospl_amd import _lib, workloads

    m, h0, p = workloads.sphere_case(_lib.load(), level=9, seed=0, device=torch.device("cuda"))
    print("[fullsize] mesh N=%d" % m.N, flush=True)
    return m, h0, p


def test_fullsize_steps_match_oracle(full):
    from gospl_amd.model import Model

    m, h0, p = full
    g = Model(m, h0, tend=1e12, **p)
    o = oracle_model(m, h0, p)
    g.initialise()
    o.initialise()
    assert rel_err(m.to_input_order(g.FAL.cpu().numpy()), o.FAL) < TOL
    for s in range(2):
        o.h = m.to_input_order(g.hLocal.cpu().numpy()).copy()
        o.cumED = m.to_input_order(g.cumED.cpu().numpy()).copy()
        o.tNow = g.tNow
        g.step()
        o.step()
        print("[fullsize] step %d: pits %d cascade %d" % (s, g.npits, g.cascade_passes), flush=True)
        assert np.array_equal(m.to_input_order(g.lFill.cpu().numpy()), o.lFill), s
        assert np.array_equal(m.to_input_order(g.pitIDs.cpu().numpy()), o.pitIDs), s
        assert g.cascade_passes == o.cascade_passes, s
        h = m.to_input_order(g.hLocal.cpu().numpy())
        assert rel_err(h, o.h) < TOL, (s, rel_err(h, o.h))
        assert rel_err(m.to_input_order(g.FAL.cpu().numpy()), o.FAL) < TOL, s
        ed = m.to_input_order(g.cumED.cpu().numpy())
        assert np.abs(ed - o.cumED).max() <= TOL * np.abs(o.h).max(), s


def test_fullsize_invariants(full):
    import torch

    from gospl_amd.model import Model, _ptr

    m, h0, p = full
    g = Model(m, h0, tend=1e12, **p)
    g.fillElevation()
    h = g.hLocal
    lf = g.lFill
    assert bool((lf >= h).all())
    sea = lf <= g.sealevel
    assert bool((lf[sea] == h[sea]).all())

    g.flowAccumulation()
    N = m.N
    rcv = g.rcvIDi.long()
    w = g.wghtVali
    own = torch.arange(N, device=h.device).unsqueeze(0).expand_as(rcv)
    out = (w > 0) & (rcv != own)
    assert bool((h.gather(0, rcv.flatten()).view_as(rcv)[out] < h.unsqueeze(0).expand_as(rcv)[out]).all())
    ws = w.sum(0)
    has = out.any(0)
    assert float((ws[has] - 1.0).abs().max()) <= 4e-16 * 3

    # flow accumulation on the same receivers: residual and mass balance
    b = g._b
    q = torch.zeros_like(b)
    lev = ctypes.c_int32(0)
    bad = ctypes.c_int64(0)
    assert g.lib.gospl_flowacc(g.H, 3, _ptr(g.rcvIDi), _ptr(g.wghtVali), _ptr(b), _ptr(q),
                               ctypes.byref(lev), ctypes.byref(bad), g.stream) == 0
    torch.cuda.synchronize()
    assert bad.value == 0
    print("[fullsize] flow DAG depth %d" % lev.value, flush=True)
    wo = torch.where(out, w, torch.zeros_like(w))
    r = q - b
    r.index_add_(0, rcv.flatten(), -(wo * q.unsqueeze(0)).flatten())
    assert float(r.abs().max()) <= 1e-12 * float(q.abs().max())
    sink = (q * (1.0 - wo.sum(0))).sum()
    assert abs(float(sink) - float(b.sum())) <= 1e-9 * float(b.sum())


def test_fullsize_reproducible(full):
    from gospl_amd.model import Model

    m, h0, p = full
    res = []
    for _ in range(2):
        g = Model(m, h0, tend=1e12, **p)
        g.initialise()
        g.step()
        g.step()
        res.append([t.cpu().numpy() for t in (g.hLocal, g.FAL, g.cumED, g.fillFAL)])
        g.dm.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)
