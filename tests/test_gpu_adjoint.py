"""Reverse-mode gradients (gospl_amd.adjoint) on the GPU.

Bars:
* each transposed solve against its definition: the flow transpose and the
  hillslope forward / transpose against sparse LU on the CPU, the
  stream-power reverse step against the CPU reference gradient's step, to
  1e-10 relative (max-norm, relative to the largest value);
* the whole gradient of a 12-step run (two checkpoint segments, the last
  partial) against the CPU FP64 reference gradient (oracle/adjoint_ref.py,
  itself pinned by brute-force finite differences) to 1e-9;
* the inversion workload at icosahedral level 6 (50 steps, checkpoints every
  10): the gradient against central finite differences of the frozen map,
  h = 1e-4 in log10 K and in mm/yr of uplift, on 32 random vertices each, to
  1e-4 relative;
* the forward of Evolution is Model's run bit for bit, the frozen map at the
  recorded inputs repeats it bit for bit, and the gradient does not depend
  on the checkpoint spacing (bit for bit); the autograd wrapper gives the
  same gradient.
"""

from __future__ import annotations

import ctypes
import functools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SOLVE_TOL = 1e-10
REF_TOL = 1e-9
FD_TOL = 1e-4


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@functools.lru_cache(maxsize=None)
def _case(level, nsteps):
    from gospl_amd import _lib, workloads

    m, h0, p, lk_true, lk0, up = workloads.adjoint_case(_lib.load(), level=level, seed=1,
                                                        lmax=min(64, 4 * 2 ** level))
    p = dict(p)
    p["nsteps"] = nsteps
    return m, h0, p, lk_true, lk0, up


def _evolution(level, nsteps, every):
    import torch

    from gospl_amd.adjoint import Evolution

    m, h0, p, lk_true, lk0, up = _case(level, nsteps)
    kw = {k: v for k, v in p.items() if k != "nsteps"}
    evo = Evolution(m, h0, nsteps=p["nsteps"], checkpoint_every=every, **kw)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64)).cuda()
    return evo, t(lk_true), t(lk0), t(up), m, p


def _np_steps(st):
    return [dict(rcv=s.rcv.cpu().numpy(), wgt=s.wgt.cpu().numpy(), dist=s.dist.cpu().numpy(),
                 fal=s.fal.cpu().numpy(), lfill=s.lfill.cpu().numpy(), lsink=s.lsink.cpu().numpy() != 0,
                 sea=s.sea.cpu().numpy() != 0, uplift=s.uplift) for s in st.steps]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def test_transposed_solves_against_sparse_lu():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    from gospl_amd import _lib
    from oracle import adjoint_ref as ar

    print("adjoint transposes: level 5, 3 steps", flush=True)
    evo, lk_true, lk0, up, m, p = _evolution(5, 3, 10)
    _, st = evo.run(lk0, up, record=True)
    fs = st.steps[-1]
    model = evo.model
    lib, H, S = model.lib, model.H, model.stream
    N = m.N
    rng = np.random.default_rng(5)
    lam = torch.as_tensor(rng.normal(size=N)).cuda()
    # flow transpose: (I - W) y = lam
    y = torch.empty_like(lam)
    _lib.check(lib.gospl_flowacc_adjoint(H, 3, _ptr(fs.rcv), _ptr(fs.wgt), _ptr(lam), _ptr(y), S),
               "gospl_flowacc_adjoint")
    rcv = fs.rcv.cpu().numpy().astype(np.int64)
    wgt = fs.wgt.cpu().numpy()
    nodes = np.arange(N)
    keep = rcv != nodes[None, :]
    W = sp.csr_matrix((wgt[keep], (np.broadcast_to(nodes, rcv.shape)[keep], rcv[keep])), shape=(N, N))
    y_ref = spla.spsolve((sp.identity(N, format="csr") - W).tocsc(), lam.cpu().numpy())
    assert _rel(y.cpu().numpy(), y_ref) <= SOLVE_TOL
    # and it is the transpose of the forward flow solve: <lam, q(b)> = <y, b>
    b = torch.as_tensor(rng.uniform(0.5, 2.0, N)).cuda()
    q = torch.empty_like(b)
    _lib.check(lib.gospl_flowacc(H, 3, _ptr(fs.rcv), _ptr(fs.wgt), _ptr(b), _ptr(q), None, None, S),
               "gospl_flowacc")
    lhs = float(torch.dot(lam, q))
    rhs = float(torch.dot(y, b))
    assert abs(lhs - rhs) <= 1e-12 * float(torch.dot(lam.abs(), q.abs()))
    # hillslope forward and transpose of the frozen operator, with
    # diffusivities that put the operator far from the identity on these
    # ~220 km cells (contraction about 0.6) and make land and sea rows differ
    geo = ar.device_geometry(m)
    pp = dict(m=p["m"], dt=p["dt"], cda=p["cda"], cdm=p["cdm"], sealevel=p["sealevel"])
    hp = dict(pp, cda=5.0e5, cdm=2.0e6)
    M = ar.hill_matrix(fs.sea.cpu().numpy() != 0, hp, geo).tocsc()
    s_ = max(hp["cda"], hp["cdm"]) * hp["dt"] * evo.dm.gmax
    rho = s_ / (1.0 + s_)
    assert rho > 0.3
    sweeps = int(np.ceil(np.log(2.0 ** -60) / np.log(rho))) + 1
    hin = fs.h.clone()
    hout = torch.empty_like(hin)
    _lib.check(lib.gospl_hillslope_frozen(H, hp["cda"], hp["cdm"], hp["dt"], _ptr(fs.sea), sweeps, _ptr(hin),
                                          _ptr(hout), S), "gospl_hillslope_frozen")
    assert _rel(hout.cpu().numpy(), spla.spsolve(M, hin.cpu().numpy())) <= SOLVE_TOL
    out = torch.empty_like(lam)
    _lib.check(lib.gospl_hillslope_adjoint(H, hp["cda"], hp["cdm"], hp["dt"], _ptr(fs.sea), sweeps, _ptr(lam),
                                           _ptr(out), S), "gospl_hillslope_adjoint")
    lt = lam.cpu().numpy()
    assert _rel(out.cpu().numpy(), spla.spsolve(M.T.tocsc(), lt)) <= SOLVE_TOL
    assert _rel(spla.spsolve(M, lt), spla.spsolve(M.T.tocsc(), lt)) > 1e-6   # M is not symmetric here
    # stream-power reverse step against the reference's
    kfac = torch.pow(10.0, lk0).contiguous()
    gh = torch.empty_like(lam)
    gk = torch.empty_like(lam)
    gf = torch.empty_like(lam)
    _lib.check(lib.gospl_spl_adjoint(H, 3, _ptr(fs.rcv), _ptr(fs.wgt), _ptr(fs.dist), _ptr(fs.fal), _ptr(kfac),
                                     1.0, p["m"], p["dt"], p["sealevel"], _ptr(fs.lfill), _ptr(fs.lsink),
                                     _ptr(fs.h), None, _ptr(lam), _ptr(gh), _ptr(gk), _ptr(gf), S),
               "gospl_spl_adjoint")
    stn = _np_steps(st)[-1]
    gh_ref, gk_ref = ar.spl_step_gradient(stn, fs.h.cpu().numpy(), lk0.cpu().numpy(), pp, geo,
                                          lam.cpu().numpy())
    assert _rel(gh.cpu().numpy(), gh_ref) <= SOLVE_TOL
    assert _rel(gk.cpu().numpy(), gk_ref) <= SOLVE_TOL
    # dA/dFAL and dA/dlog10 K scale the same coefficients: FAL/m * g_fal = g_logk / ln 10
    fal = fs.fal.cpu().numpy()
    pos = fal > 0
    lhs = (fal * gf.cpu().numpy() / p["m"])[pos]
    rhs = (gk.cpu().numpy() / np.log(10.0))[pos]
    assert _rel(lhs, rhs) <= SOLVE_TOL
    # the reverse step is not vacuous: sea cells have no stream-power term
    sea = fs.lfill.cpu().numpy() <= p["sealevel"]
    assert np.all(gk.cpu().numpy()[sea] == 0.0)
    assert np.count_nonzero(gk.cpu().numpy()[~sea]) > 0.5 * np.count_nonzero(~sea)


def test_gradient_matches_cpu_reference():
    import torch

    from oracle import adjoint_ref as ar

    print("adjoint vs CPU reference: level 5, 12 steps, checkpoints every 5", flush=True)
    evo, lk_true, lk0, up, m, p = _evolution(5, 12, 5)
    target, _ = evo.run(lk_true, up)
    h, st = evo.run(lk0, up, record=True)
    g_logk, g_up = evo.gradient(st, 2.0 * (h - target))
    pp = dict(m=p["m"], dt=p["dt"], cda=p["cda"], cdm=p["cdm"], sealevel=p["sealevel"], init_uplift=True)
    geo = ar.device_geometry(m)
    gk_ref, gu_ref = ar.frozen_gradient(evo.h0_host, lk0.cpu().numpy(), up.cpu().numpy(), _np_steps(st), geo, pp,
                                        (2.0 * (h - target)).cpu().numpy())
    ek = _rel(g_logk.cpu().numpy(), gk_ref)
    eu = _rel(g_up.cpu().numpy(), gu_ref)
    print("  rel err logk %.3e uplift %.3e" % (ek, eu), flush=True)
    assert ek <= REF_TOL and eu <= REF_TOL
    assert float(torch.abs(g_logk).max()) > 0.0


def test_forward_replay_and_checkpoint_independence():
    import torch

    from gospl_amd.adjoint import Evolution, evolve
    from gospl_amd.model import Model

    print("adjoint forward identity: level 5, 12 steps", flush=True)
    evo, lk_true, lk0, up, m, p = _evolution(5, 12, 5)
    h, st = evo.run(lk_true, up, record=True)
    kw = {k: v for k, v in p.items() if k != "nsteps"}
    ref = Model(evo.dm, evo.h0_host, K=1.0, upsub=up.cpu().numpy(), tend=12 * p["dt"], **kw)
    ref._kfac = torch.pow(10.0, lk_true).contiguous()
    ref.initialise()
    ref.runProcesses()
    assert torch.equal(h, ref.hLocal)
    hf = evo.frozen_run(st, lk_true, up)
    print("  frozen map vs run: max |diff| %.3e" % float((hf - h).abs().max()), flush=True)
    assert torch.equal(hf, h)
    target = h + 3.0
    h0, st0 = evo.run(lk0, up)
    ga = evo.gradient(st0, 2.0 * (h0 - target))
    evo3 = Evolution(evo.dm, evo.h0_host, nsteps=12, checkpoint_every=3, **kw)
    h3, st3 = evo3.run(lk0, up)
    gb = evo3.gradient(st3, 2.0 * (h3 - target))
    assert torch.equal(h0, h3)
    assert torch.equal(ga[0], gb[0]) and torch.equal(ga[1], gb[1])
    x = lk0.clone().requires_grad_(True)
    u = up.clone().requires_grad_(True)
    loss = ((evolve(evo, x, u) - target) ** 2).sum()
    loss.backward()
    assert torch.equal(x.grad, ga[0]) and torch.equal(u.grad, ga[1])


def test_finite_differences_level6():
    import torch

    print("adjoint finite differences: level 6, 50 steps, checkpoints every 10", flush=True)
    evo, lk_true, lk0, up, m, p = _evolution(6, 50, 10)
    target, _ = evo.run(lk_true, up)
    h, st = evo.run(lk0, up, record=True)
    g_logk, g_up = evo.gradient(st, 2.0 * (h - target))
    rng = np.random.default_rng(11)

    def dloss(a, b):  # L(a) - L(b) without cancelling the large sums
        return float(torch.sum((a - b) * (a + b - 2.0 * target)))

    worst = {}
    for what, base, g, step in (("logk", lk0, g_logk, 1.0e-4), ("uplift", up, g_up, 1.0e-7)):
        ga = g.cpu().numpy()
        # 32 random vertices among those the parameter acts on (elsewhere the
        # gradient is exactly or nearly zero and a relative error is noise)
        cand = np.nonzero(np.abs(ga) > 1e-3 * np.abs(ga).max())[0]
        assert len(cand) >= 32
        idx = rng.choice(cand, 32, replace=False)
        errs = []
        for i in idx:
            xp = base.clone()
            xm = base.clone()
            xp[i] += step
            xm[i] -= step
            if what == "logk":
                a, b = evo.frozen_run(st, xp, up), evo.frozen_run(st, xm, up)
            else:
                a, b = evo.frozen_run(st, lk0, xp), evo.frozen_run(st, lk0, xm)
            fd = dloss(a, b) / (2.0 * step)
            errs.append(abs(fd - ga[i]) / abs(fd))
        worst[what] = max(errs)
    print("  worst relative error: logk %.3e uplift %.3e" % (worst["logk"], worst["uplift"]), flush=True)
    assert worst["logk"] <= FD_TOL and worst["uplift"] <= FD_TOL
