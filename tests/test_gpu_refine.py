"""Refined solves (fixed-point correction sweeps with FP64 residuals) against
the exact FP64 sweeps and the FP64 oracle.

Bars:
* a refined solve stops at ||rhs - A x||_2 <= 1e-12 ||rhs||_2 and its
  solution is within 1e-10 of the exact sweep's (max-norm, relative to the
  largest value), from a cold start and from a warm start;
* it is bit-reproducible (integer accumulation, order independent);
* a receiver cycle makes it fall back to the exact sweep (passes == -1);
* one model step in refined mode from identical state matches the oracle to
  1e-9 relative (the FP64 path's own bar): filled surface, pit ids and
  receivers bit for bit, elevation / discharge / erosion to 1e-9.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.helpers import case, dev_index_to_input, oracle_model, rel_err

pytestmark = pytest.mark.gpu

RTOL = 1e-12
SOL_TOL = 1e-10
TOL = 1e-9


def _model(m, h0, p, **kw):
    from gospl_amd.model import Model

    return Model(m, h0, tend=1e12, **p, **kw)


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _flowacc_exact(g, b):
    import torch

    from gospl_amd import _lib
    from gospl_amd.model import _ptr

    q = torch.zeros_like(b)
    lev, bad = ctypes.c_int32(0), ctypes.c_int64(0)
    _lib.check(_lib.load().gospl_flowacc(g.H, g.flowDir, _ptr(g.rcvIDi), _ptr(g.wghtVali), _ptr(b), _ptr(q),
                                         ctypes.byref(lev), ctypes.byref(bad), _stream()), "gospl_flowacc")
    assert bad.value == 0
    return q


def _flowacc_refine(g, b, q0, rcv=None, wgt=None, nrcv=None):
    from gospl_amd import _lib
    from gospl_amd.model import _ptr

    q = q0.clone()
    npass, rr, bad = ctypes.c_int32(0), ctypes.c_double(0.0), ctypes.c_int64(0)
    rcv = g.rcvIDi if rcv is None else rcv
    wgt = g.wghtVali if wgt is None else wgt
    _lib.check(_lib.load().gospl_flowacc_refine(g.H, g.flowDir if nrcv is None else nrcv, _ptr(rcv), _ptr(wgt),
                                                _ptr(b), _ptr(q), RTOL, 4, ctypes.byref(npass), ctypes.byref(rr),
                                                ctypes.byref(bad), _stream()), "gospl_flowacc_refine")
    return q, int(npass.value), float(rr.value), int(bad.value)


@pytest.mark.parametrize("kind,size", [("planar", 120), ("sphere", 6)])
def test_flowacc_refine_matches_exact_sweep(kind, size):
    import torch

    m, h0, p = case(kind, size)
    g = _model(m, h0, p)
    g.initialise()
    g.step()
    b = g._b.clone()
    qe = _flowacc_exact(g, b)
    # cold start
    q, npass, rr, bad = _flowacc_refine(g, b, torch.zeros_like(b))
    torch.cuda.synchronize()
    print("[refine] %s %d cold: passes %d relres %.3e" % (kind, size, npass, rr), flush=True)
    assert npass >= 1 and bad == 0
    assert rr <= RTOL
    err = float((q - qe).abs().max() / qe.abs().max())
    assert err <= SOL_TOL, err
    # reproducible to the bit (integer sums do not depend on the schedule)
    for _ in range(5):
        q2, npass2, rr2, _ = _flowacc_refine(g, b, torch.zeros_like(b))
        assert torch.equal(q, q2) and npass2 == npass and rr2 == rr
    # warm start: the exact solution of slightly different rain
    rng = np.random.default_rng(5)
    bw = b * torch.as_tensor(rng.uniform(0.99, 1.01, m.N), device=b.device)
    qw = _flowacc_exact(g, bw)
    q, npass, rr, bad = _flowacc_refine(g, b, qw)
    torch.cuda.synchronize()
    print("[refine] %s %d warm: passes %d relres %.3e" % (kind, size, npass, rr), flush=True)
    assert npass >= 1 and rr <= RTOL and bad == 0
    err = float((q - qe).abs().max() / qe.abs().max())
    assert err <= SOL_TOL, err


@pytest.mark.parametrize("kind,size", [("planar", 120), ("sphere", 6)])
def test_spl_refine_matches_exact_sweep(kind, size):
    import torch

    from gospl_amd import _lib
    from gospl_amd.model import _ptr

    m, h0, p = case(kind, size)
    g = _model(m, h0, p)
    g.initialise()
    g.step()
    out = {}
    for name in ("gospl_spl", "gospl_spl_refine"):
        h = g.hLocal.clone()
        ced = g.cumED.clone()
        eb = torch.zeros_like(h)
        ebl = torch.zeros_like(h)
        args = [g.H, g.flowDir, _ptr(g.rcvIDi), _ptr(g.wghtVali), _ptr(g.distRcvi), _ptr(g.FAL), _ptr(g._kfac),
                g.K, g.spl_m, g.dt, g.sealevel, _ptr(g.lFill), _ptr(g.lsink), _ptr(h), _ptr(ced), _ptr(eb),
                _ptr(ebl)]
        if name == "gospl_spl":
            lev = ctypes.c_int32(0)
            args += [ctypes.byref(lev), _stream()]
        else:
            npass, rr = ctypes.c_int32(0), ctypes.c_double(0.0)
            args += [RTOL, 4, ctypes.byref(npass), ctypes.byref(rr), _stream()]
        _lib.check(getattr(_lib.load(), name)(*args), name)
        torch.cuda.synchronize()
        out[name] = (h, eb)
    print("[refine] spl %s %d: passes %d relres %.3e" % (kind, size, npass.value, rr.value), flush=True)
    assert npass.value >= 1 and rr.value <= RTOL
    he, ebe = out["gospl_spl"]
    hr, ebr = out["gospl_spl_refine"]
    assert float((hr - he).abs().max()) <= SOL_TOL * float(he.abs().max())
    assert float((ebr - ebe).abs().max()) * g.dt <= SOL_TOL * float(he.abs().max())


def test_flowacc_refine_cycle_falls_back_to_exact():
    """A draining receiver cycle leaves vertices the fixed-point sweep cannot
    reach: the call falls back to the exact sweep with its cycle pass."""
    import torch

    from gospl_amd.model import DeviceMesh, Model
    from oracle import oracle as orc
    from tests.helpers import make_cycle

    m, h0, p = case("planar", 12)
    o = oracle_model(m, h0, p)
    N = m.N
    z = m.to_input_order(h0)
    rcv, dist, wgt = orc.mfdreceivers(3, 1.1, z, -1e9, np.arange(N, dtype=np.int32), o.ngb, o.nnb, o.elgt)
    rc, wc, _ = make_cycle(rcv, wgt, drain=True)
    b = np.random.default_rng(4).uniform(0.5, 2.0, N) * m.to_input_order(m.area)
    qo, bado = orc.flowacc(rc, wc, b, o.ngb, o.nnb)
    g = Model(DeviceMesh(m, "cuda"), h0, tend=1e12, **p)
    R = torch.as_tensor(np.ascontiguousarray(m.iperm[rc[m.perm]].T.astype(np.int32)), device="cuda")
    Wt = torch.as_tensor(np.ascontiguousarray(wc[m.perm].T), device="cuda")
    B = torch.as_tensor(m.to_device_order(b), device="cuda")
    q, npass, rr, bad = _flowacc_refine(g, B, torch.zeros_like(B), rcv=R, wgt=Wt, nrcv=3)
    torch.cuda.synchronize()
    assert bado == 0 and bad == 0 and npass == -1
    assert rel_err(m.to_input_order(q.cpu().numpy()), qo) < 1e-13


def _sync_oracle(o, g, m):
    o.h = m.to_input_order(g.hLocal.cpu().numpy()).copy()
    o.cumED = m.to_input_order(g.cumED.cpu().numpy()).copy()
    o.tNow = g.tNow


@pytest.mark.parametrize("kind,size", [("planar", 60), ("sphere", 6)])
def test_refined_steps_match_oracle(kind, size):
    """Refined mode, per step from identical state, at the FP64 bar."""
    m, h0, p = case(kind, size)
    g = _model(m, h0, p, refine=True, refine_dag=True)
    g.refine_stats = True
    o = oracle_model(m, h0, p)
    g.initialise()
    o.initialise()
    for s in range(4):
        _sync_oracle(o, g, m)
        g.step()
        o.step()
        assert np.array_equal(m.to_input_order(g.lFill.cpu().numpy()), o.lFill), s
        assert np.array_equal(m.to_input_order(g.pitIDs.cpu().numpy()), o.pitIDs), s
        r_in = dev_index_to_input(m, m.to_input_order(g.rcvIDi.cpu().numpy().T))
        assert np.array_equal(r_in, o.rcvIDi), s
        h = m.to_input_order(g.hLocal.cpu().numpy())
        assert rel_err(h, o.h) < TOL, (s, rel_err(h, o.h))
        assert rel_err(m.to_input_order(g.FAL.cpu().numpy()), o.FAL) < TOL, s
        ed = m.to_input_order(g.cumED.cpu().numpy())
        assert np.abs(ed - o.cumED).max() <= TOL * np.abs(o.h).max(), s
    # every solve converged (or fell back to the exact sweep on a receiver cycle)
    assert all(ps == -1 or rr <= RTOL for _, ps, rr in g.solve_log), g.solve_log


@pytest.mark.parametrize("kind,size", [("planar", 60), ("sphere", 6)])
def test_hillslope_refine_matches_fp64_solve(kind, size):
    """FP32 Jacobi corrections of the FP64 residual reach rtol 1e-12 and the
    FP64 solve's elevation to 1e-12 relative."""
    import torch

    from gospl_amd import _lib
    from gospl_amd.model import _ptr

    m, h0, p = case(kind, size)
    q = dict(p)
    q.update(cda=5.0, cdm=50.0)  # diffusion strong enough to move the surface
    g = _model(m, h0, q)
    g.initialise()
    g.step()
    res = {}
    for mode in ("fp64", "refine"):
        h = g.hLocal.clone()
        ced = torch.zeros_like(h)
        ebl = torch.zeros_like(h)
        if mode == "fp64":
            _lib.check(_lib.load().gospl_hillslope(g.H, g.Cda, g.Cdm, g.dt, g.sealevel, g.hillslope_sweeps(), _ptr(h),
                                                   _ptr(ced), _ptr(ebl), _stream()), "gospl_hillslope")
        else:
            npass, rr = ctypes.c_int32(0), ctypes.c_double(0.0)
            _lib.check(_lib.load().gospl_hillslope_refine(g.H, g.Cda, g.Cdm, g.dt, g.sealevel, g.hillslope_sweeps32(),
                                                          RTOL, 4, _ptr(h), _ptr(ced), _ptr(ebl), ctypes.byref(npass),
                                                          ctypes.byref(rr), _stream()), "gospl_hillslope_refine")
        torch.cuda.synchronize()
        res[mode] = (h, ced)
    print("[refine] hillslope %s %d: sweeps32 %d passes %d relres %.3e" %
          (kind, size, g.hillslope_sweeps32(), npass.value, rr.value), flush=True)
    assert 1 <= npass.value and rr.value <= RTOL
    he, ce = res["fp64"]
    hr, cr = res["refine"]
    assert float((hr - he).abs().max()) <= 1e-12 * float(he.abs().max())
    assert float((cr - ce).abs().max()) <= 1e-12 * float(he.abs().max())
    assert float(ce.abs().max()) > 0.0
