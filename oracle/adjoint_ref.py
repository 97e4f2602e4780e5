"""CPU FP64 reference of the frozen-step gradient.

TEST INFRASTRUCTURE ONLY: the GPU tests compare ``gospl_amd.adjoint`` with it.

goSPL has no derivatives; this restates, as sparse matrices, the step that
the adjoint differentiates (receivers, weights, distances, discharge, fill,
sink and sea flags held at recorded values):

* stream power, assembled as ``SPL._eroMats`` (gospl/eroder/SPL.py:45-114;
  ``OracleModel._eroMats`` with a per-vertex K = 10**logk):
  A(h) = diag(1 + sum_k c_ik) - sum_k c_ik E(i, r_ik), c_ik = Kbr_i w_ik l(h_i -
  h_rk) / d_ik, l(s) = s / (s + 0.01), Kbr = K dt FAL^m (0 where lfill <= sea
  level, SPL.py:86-96); x = A^-1 h; h' = x except at flat-mesh outlets and
  sinks, which keep h (SPL.py:345-353);
* hillslope diffusion, assembled as ``hillSLP._hillSlope`` (gospl/sed/
  hillslope.py:235-293) from ``sethillslopecoeff`` with Cdm*dt at sea flags,
  Cda*dt elsewhere, identity rows at flat-mesh outlets; h'' = M^-1 h';
* uplift: h_next = h'' + U dt (unstructuredmesh.py:743-794).

Everything is solved with SuperLU (scipy.sparse.linalg.splu), and the
gradient is formed from the matrices, not from per-edge recurrences:
mu = A^-T (1-R) lam, dL/dh = mu - J^T mu + R lam with J = d(A x)/dh at fixed x,
dL/dlog10 K = -ln(10) mu * ((A - I) x), and M^-T for diffusion.  The
frozen forward's central differences pin it (tests/test_adjoint_ref.py).

Steps are dicts of numpy arrays: ``rcv`` (K, N) int, ``wgt``, ``dist`` (K, N),
``fal``, ``lfill`` (N), ``lsink`` (N) bool, ``sea`` (N) bool, ``uplift`` bool.
``geo``: ``ngb`` (N, 12) (-1 padded), ``nnb``, ``elgt``/``vdist`` (N, 12),
``area``, ``outlet`` (N) bool, ``flat``.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as orc

LN10 = math.log(10.0)


def spl_system(st, h, logk, p):
    """A(h) (CSR), the per-receiver coefficients c (K, N) and their
    derivatives dc/d(h_i - h_rk) (K, N)."""
    N = len(h)
    rcv = np.asarray(st["rcv"], np.int64)
    K = rcv.shape[0]
    sea = st["lfill"] <= p["sealevel"]
    kbr = (10.0 ** logk) * (p["dt"] * st["fal"] ** p["m"])
    kbr[sea] = 0.0
    nodes = np.arange(N)
    c = np.zeros((K, N))
    dc = np.zeros((K, N))
    for k in range(K):
        r = rcv[k]
        dh = h - h[r]
        lim = np.divide(dh, dh + 1.0e-2, out=np.zeros(N), where=dh != 0)
        dlim = 1.0e-2 / (dh + 1.0e-2) ** 2
        w = np.where(sea, 0.0, st["wgt"][k])
        d = st["dist"][k]
        ok = (d != 0) & (r != nodes)
        dd = np.where(ok, d, 1.0)
        c[k] = np.where(ok, kbr * lim * w / dd, 0.0)
        dc[k] = np.where(ok, kbr * dlim * w / dd, 0.0)
    rows = np.concatenate([nodes] + [nodes] * K)
    cols = np.concatenate([nodes] + [rcv[k] for k in range(K)])
    vals = np.concatenate([1.0 + c.sum(axis=0)] + [-c[k] for k in range(K)])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    return A, c, dc


def reset_mask(st, geo):
    R = np.asarray(st["lsink"], bool).copy()
    if geo["flat"]:
        R |= np.asarray(geo["outlet"], bool)
    return R


def hill_matrix(sea, p, geo):
    N = len(sea)
    kd = np.where(sea, p["cdm"], p["cda"]) * p["dt"]
    coeff = orc.sethillslopecoeff(kd, geo["ngb"], geo["nnb"], geo["area"], geo["elgt"], geo["vdist"])
    if geo["flat"]:
        o = np.asarray(geo["outlet"], bool)
        coeff[o, 1:] = 0.0
        coeff[o, 0] = 1.0
    diag = coeff[:, 0].copy()
    diag[diag == 0.0] = 1.0       # zero-area rows keep their value
    rows = [np.arange(N)]
    cols = [np.arange(N)]
    vals = [diag]
    ngb = np.asarray(geo["ngb"])
    for q in range(ngb.shape[1]):
        j = ngb[:, q]
        ok = (j >= 0) & (q < np.asarray(geo["nnb"]))
        rows.append(np.nonzero(ok)[0])
        cols.append(j[ok])
        vals.append(coeff[ok, 1 + q])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def _diffuses(p):
    return p["cda"] != 0.0 or p["cdm"] != 0.0


def frozen_forward(h0, logk, U, steps, geo, p, keep=False):
    """Final elevation of the frozen map; keep=True also returns each step's
    (h, x, LU(A), A, J-data, LU(M))."""
    h = np.array(h0, np.float64)
    if p.get("init_uplift", True):
        h = h + U * p["dt"]
    traj = []
    for st in steps:
        A, c, dc = spl_system(st, h, logk, p)
        luA = spla.splu(A.tocsc())
        x = luA.solve(h)
        R = reset_mask(st, geo)
        hp = np.where(R, h, x)
        luM = None
        if _diffuses(p):
            luM = spla.splu(hill_matrix(np.asarray(st["sea"], bool), p, geo).tocsc())
            hpp = luM.solve(hp)
        else:
            hpp = hp
        if keep:
            traj.append((h, x, luA, A, c, dc, R, luM))
        h = hpp + U * p["dt"] if st["uplift"] else hpp
    return (h, traj) if keep else h


def frozen_gradient(h0, logk, U, steps, geo, p, grad_h):
    """(dL/dlog10 K, dL/dU) of the frozen map for dL/dh_final = grad_h."""
    _, traj = frozen_forward(h0, logk, U, steps, geo, p, keep=True)
    N = len(h0)
    lam = np.array(grad_h, np.float64)
    g_logk = np.zeros(N)
    g_U = np.zeros(N)
    for st, (h, x, luA, A, c, dc, R, luM) in zip(reversed(steps), reversed(traj)):
        if st["uplift"]:
            g_U += p["dt"] * lam
        lp = luM.solve(lam, trans="T") if luM is not None else lam
        lam, gk = _spl_reverse(st, x, luA, A, dc, R, lp)
        g_logk += gk
    if p.get("init_uplift", True):
        g_U += p["dt"] * lam
    return g_logk, g_U


def _spl_reverse(st, x, luA, A, dc, R, lp):
    N = len(x)
    mu = luA.solve(np.where(R, 0.0, lp), trans="T")
    rcv = np.asarray(st["rcv"], np.int64)
    nodes = np.arange(N)
    rows, cols, vals = [], [], []
    for k in range(rcv.shape[0]):
        t = dc[k] * (x - x[rcv[k]])
        rows += [nodes, nodes]
        cols += [nodes, rcv[k]]
        vals += [t, -t]
    J = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    g_logk = -LN10 * mu * (A @ x - x)
    g_h = mu - J.T @ mu + np.where(R, lp, 0.0)
    return g_h, g_logk


def spl_step_gradient(st, h, logk, p, geo, lam):
    """One stream-power step's reverse: (dL/dh, dL/dlog10 K) for
    lam = dL/dh' at pre-erosion elevation h."""
    A, c, dc = spl_system(st, np.asarray(h, np.float64), logk, p)
    luA = spla.splu(A.tocsc())
    x = luA.solve(np.asarray(h, np.float64))
    return _spl_reverse(st, x, luA, A, dc, reset_mask(st, geo), np.asarray(lam, np.float64))


def device_geometry(mesh):
    """``geo`` in device order from a ``gospl_amd.mesh.Mesh`` ((W, N) arrays)."""
    W, N = mesh.nbr.shape
    pad = lambda a, v: np.concatenate([np.asarray(a).T, np.full((N, 12 - W), v, np.asarray(a).dtype)], axis=1)
    return dict(ngb=np.ascontiguousarray(pad(mesh.nbr, -1), dtype=np.int32),
                nnb=np.ascontiguousarray(mesh.deg, dtype=np.int32),
                elgt=np.ascontiguousarray(pad(mesh.elgt, 0.0)), vdist=np.ascontiguousarray(pad(mesh.vdist, 0.0)),
                area=np.asarray(mesh.area, np.float64), outlet=np.asarray(mesh.outlet).astype(bool),
                flat=bool(mesh.flat))
