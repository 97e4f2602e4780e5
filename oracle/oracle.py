"""CPU FP64 oracle of goSPL's per-timestep surface-process hot path.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it as the checker; the product
(``gospl_amd``) never imports it.

The serial kernels live in ``oracle.c`` (built into ``liboracle.so``); this
module restates goSPL's Python glue around them, method by method, in input
mesh order:

* ``fillElevation``            gospl/flow/pitfilling.py:849-903 (+ :557-847)
* ``_buildFlowDirection``      gospl/flow/flowplex.py:480-544
* ``flowAccumulation``         gospl/flow/flowplex.py:752-945
* ``_distributeDownstream``    gospl/flow/flowplex.py:576-686
* ``erodepSPL``                gospl/eroder/SPL.py:45-114, 260-366, 402-454
* ``erodepSPLnl`` (n != 1)     gospl/eroder/nlSPL.py:321-456
* ``getHillslope``             gospl/sed/hillslope.py:90-293, 614-649
* ``sedChange`` (nodep False)  gospl/sed/sedplex.py:103-1086 (oracle/sed.py)
* ``seaChange`` (seadepo True) gospl/sed/seaplex.py:51-861 (oracle/sea.py)
* ``applyForces``              gospl/mesher/unstructuredmesh.py:743-794, 903-972
* ``runProcesses``             gospl/model.py:331-466

The linear systems are solved exactly (topological sweeps for the receiver
DAGs, Jacobi to a fixed point for the diffusion operator) where goSPL stops
its Krylov solvers at rtol 1e-8; parity of the HIP path is defined against
these exact solutions.

The parts of this glue that goSPL implements with numpy, pandas,
numpy_indexed or PETSc (absent here, so no golden vector can be written) are
pinned by their definitions, restated with plain loops in
tests/test_oracle_glue_pins.py: spill points and flat routing (brute force,
BFS), per-pit tables (goSPL's own Fortran getpitvol, brute-force maxima), the
cascade bookkeeping (per-pass inflows, capacities, water levels, spills and
the whole call's water balance) and the end of flowAccumulation.

Parity unpinned: the Krylov solutions themselves.  goSPL stops FGMRES /
Richardson at rtol 1e-8 (flowplex.py:38, hillslope.py:362-371); the oracle
solves exactly, so an actual goSPL run agrees with it only to that
tolerance.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

from .sea import SEA
from .sed import SED

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MISSING_DATA_SENTINEL = -1.0e8   # gospl/tools/constants.py:34

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, os.path.join(HERE, "liboracle.so")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(os.path.join(HERE, "oracle.c")):
            build()
        L = ctypes.CDLL(LIB)
        P, I64, I32, F64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.or_mfdreceivers.argtypes = [I32, F64, P, F64, P, I64, P, P, P, P, P, P]
        L.or_fill.argtypes = [I64, P, P, P, P, P, P, P]
        L.or_fill.restype = I32
        L.or_label_pits.argtypes = [F64, P, I64, P, P, P]
        L.or_fill_dir.argtypes = [P, I64, P, P, I64, P, P, P]
        L.or_nghb_dir.argtypes = [P, P, P, I64, P, P, P]
        L.or_nghb_dir.restype = I32
        L.or_getpitvol.argtypes = [P, P, P, P, P, I64, I64, P]
        L.or_sethillslopecoeff.argtypes = [I64, P, P, P, P, P, P, P]
        L.or_flowacc.argtypes = [I64, I32, P, P, P, P, P, P]
        L.or_flowacc.restype = I64
        L.or_spl_solve.argtypes = [I64, I32, P, P, P, P, P]
        L.or_spl_solve.restype = I64
        L.or_hillslope_solve.argtypes = [I64, P, P, P, P, P, I32]
        L.or_hillslope_solve.restype = I32
        L.or_cheb_smooth.argtypes = [I64, P, P, P, P, P, I32, P]
        L.or_cheb_smooth.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --------------------------------------------------------------------------
# kernel wrappers (input order, (N,12) neighbour tables)
# --------------------------------------------------------------------------

def mfdreceivers(nrcv, fexp, elev, sl, gid, ngb, nnb, elgt):
    N = len(elev)
    rcv = np.empty((N, nrcv), np.int32)
    dist = np.empty((N, nrcv))
    wgt = np.empty((N, nrcv))
    elev = _c(elev, np.float64)
    gid = _c(gid, np.int32)
    lib().or_mfdreceivers(int(nrcv), float(fexp), _p(elev), float(sl), _p(gid), N,
                          _p(ngb), _p(nnb), _p(elgt), _p(rcv), _p(dist), _p(wgt))
    return rcv, dist, wgt


def fill(dem, outlet, ngb, nnb):
    """One-rank _performFilling core: (fill, reached, ledge)."""
    N = len(dem)
    dem = _c(dem, np.float64)
    out = np.empty(N)
    reached = np.empty(N, np.int8)
    ledge = np.empty(N, np.int8)
    rc = lib().or_fill(N, _p(ngb), _p(nnb), _p(dem), _p(_c(outlet, np.int8)), _p(out), _p(reached), _p(ledge))
    if rc != 0:
        raise MemoryError("or_fill")
    return out, reached, ledge


def label_pits(lvl, fill_, ngb, nnb):
    N = len(fill_)
    out = np.empty(N, np.int32)
    lib().or_label_pits(float(lvl), _p(_c(fill_, np.float64)), N, _p(ngb), _p(nnb), _p(out))
    return out


def fill_dir(spill, pitids, h, ngb, nnb):
    N = len(h)
    out = np.empty(N, np.int32)
    s = _c(spill, np.int32)
    lib().or_fill_dir(_p(s), len(s), _p(_c(pitids, np.int32)), _p(_c(h, np.float64)), N,
                      _p(ngb), _p(nnb), _p(out))
    return out


def nghb_dir(pitids, h, ptdir, ngb, nnb):
    N = len(h)
    out = np.empty(N, np.int32)
    ch = lib().or_nghb_dir(_p(_c(pitids, np.int32)), _p(_c(h, np.float64)), _p(_c(ptdir, np.int32)),
                           N, _p(ngb), _p(nnb), _p(out))
    return out, ch


def getpitvol(hlvl, elev, pit, ids, area):
    hlvl = _c(hlvl, np.float64)
    vol = np.empty((len(hlvl), 4))
    lib().or_getpitvol(_p(hlvl), _p(_c(elev, np.float64)), _p(_c(pit, np.int32)),
                       _p(_c(ids, np.int32)), _p(_c(area, np.float64)), len(elev), len(hlvl), _p(vol))
    return vol


def sethillslopecoeff(kd, ngb, nnb, area, elgt, vdist):
    N = len(kd)
    out = np.empty((N, 13))
    lib().or_sethillslopecoeff(N, _p(ngb), _p(nnb), _p(_c(area, np.float64)), _p(elgt), _p(vdist),
                               _p(_c(kd, np.float64)), _p(out))
    return out


def flowacc(rcv, wgt, b, ngb, nnb):
    N, K = rcv.shape
    q = np.empty(N)
    bad = lib().or_flowacc(N, K, _p(_c(rcv, np.int32)), _p(_c(wgt, np.float64)), _p(_c(b, np.float64)), _p(q),
                           _p(_c(ngb, np.int32)), _p(_c(nnb, np.int32)))
    return q, int(bad)


def spl_solve(rcv, off, diag, rhs):
    N, K = rcv.shape
    x = np.empty(N)
    bad = lib().or_spl_solve(N, K, _p(_c(rcv, np.int32)), _p(_c(off, np.float64)),
                             _p(_c(diag, np.float64)), _p(_c(rhs, np.float64)), _p(x))
    return x, int(bad)


def flow_matrix(rcv, wgt):
    """goSPL's flow matrix fMat = (I - W)^T of matrixFlow (flowplex.py:429-478):
    W[i, rcv[i, k]] += wgt[i, k], entries at self receivers dropped."""
    import scipy.sparse as sp

    N, K = rcv.shape
    rows = np.repeat(np.arange(N), K)
    cols = np.asarray(rcv, np.int64).ravel()
    data = -np.asarray(wgt, np.float64).ravel()
    data[cols == rows] = 0.0
    return (sp.identity(N, format="csr") + sp.csr_matrix((data, (rows, cols)), shape=(N, N))).T.tocsr()


def ero_matrix(rcv, off, diag):
    """SPL's eMat (SPL.py:73-108): diag on the diagonal, off[:, k] at column
    rcv[:, k] (zero at self receivers)."""
    import scipy.sparse as sp

    N, K = rcv.shape
    rows = np.repeat(np.arange(N), K)
    return (sp.diags(np.asarray(diag, np.float64)) +
            sp.csr_matrix((np.asarray(off, np.float64).ravel(), (rows, np.asarray(rcv, np.int64).ravel())),
                          shape=(N, N))).tocsr()


def coupled_ed_solve(rcv, off, diag, rcv_f, wgt_f, fdep, area, dt, hold):
    """Transport-limited stream power (SPL._coupledEDSystem, SPL.py:116-258):
    goSPL's nested system in elevation h and sediment flux q,

        [eMat - diag(fDep)   -diag(fDep dt / area)] [h]   [(1 - fDep) hOld]
        [diag(area / dt)      fMati               ] [q] = [hOld area / dt ]

    (A00 = eMat - diag(fDep) :155-157, A01 :156, A10 :157, fMati the flow
    matrix of the step's main receivers, right-hand sides :170-174), assembled
    the same way and solved by sparse LU where goSPL runs TFQMR with a Schur
    fieldsplit preconditioner to rtol 1e-8 (:181-206).  Returns (h, q)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    N = len(hold)
    E = ero_matrix(rcv, off, diag)
    A00 = E + sp.diags(-fdep)
    A01 = sp.diags(-fdep * dt / area)
    A10 = sp.diags(area / dt)
    F = flow_matrix(rcv_f, wgt_f)
    M = sp.bmat([[A00, A01], [A10, F]], format="csc")
    rhs = np.concatenate([(1.0 - fdep) * hold, hold * area / dt])
    x = spla.spsolve(M, rhs)
    return x[:N], x[N:]


def transport_fdep(G, area, PA, sea, outlets=None):
    """Dimensionless deposition coefficient of the transport-limited solve
    (SPL.py:316-333): fDep = G area / PA with the denominator floored where
    the ratio reaches the 0.99 cap, 0 where PA = 0, at sea and (flat meshes)
    at outlets, capped at 0.99."""
    num = G * area
    floor = num / 0.99
    fdep = np.divide(num, np.maximum(PA, floor), out=np.zeros_like(PA), where=PA != 0)
    fdep[sea] = 0.0
    fdep[fdep > 0.99] = 0.99
    if outlets is not None:
        fdep[outlets] = 0.0
    return fdep


def nl_transport_solve(rcv, wgt, dist, kbr, n, fdep, area, dt, hold, rcv_f=None, wgt_f=None, tol=1.0e-14,
                       maxit=100):
    """Non-linear transport-limited stream power (nlSPL._solveNL_ed and
    _form_residual_ed, gospl/eroder/nlSPL.py:82-125, 255-320): the root of

        F(h) = (h - hOld)(1 - fDep) + Kbr S(h)^n - fDep dt Qt(h) / area,
        Qt(h) = max(fMati^-1 ((hOld - h) area / dt), 0),

    S the weighted positive slope of local_spl (functions.F90:1469-1502:
    receivers with w > 0 lower than the vertex, S clipped at 0), by Newton
    steps whose linear systems are goSPL's two-field form solved by sparse LU:

        [J_S + diag(1 - fDep)   -diag(fDep dt / area [Qt > 0])] [dh]   [-F]
        [diag(area / dt)         fMati                        ] [dq] = [ 0]

    (J_S the derivative of Kbr S^n, local_spl_coeff :1504-1542), until the
    step is below tol max(max|h|, 1).  goSPL solves it with SNES (L-BFGS, rtol
    / atol 1e-6).  Returns (h, Newton steps)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    N, K = rcv.shape
    rcv_f = rcv if rcv_f is None else rcv_f
    wgt_f = wgt if wgt_f is None else wgt_f
    F = flow_matrix(rcv_f, wgt_f).tocsc()
    Fs = spla.splu(F)
    x = np.array(hold, np.float64)
    rows = np.repeat(np.arange(N), K)
    for it in range(1, maxit + 1):
        Qt = Fs.solve(((hold - x) * area) * (1.0 / dt))
        Qp = np.where(Qt < 0.0, 0.0, Qt)
        on = (wgt > 0.0) & (x[:, None] > x[rcv])
        wd = np.where(on, wgt / np.where(on, dist, 1.0), 0.0)
        S = np.sum(wd * (x[:, None] - x[rcv]), axis=1)
        S = np.where(S < 0.0, 0.0, S)
        Sn = np.where(S > 0.0, np.power(np.where(S > 0.0, S, 1.0), n), 0.0)
        res = (x - hold) * (1.0 - fdep)
        res = res + kbr * Sn
        res = res - (fdep * dt) * Qp / area
        dd = np.where(S > 0.0, n * kbr * np.power(np.where(S > 0.0, S, 1.0), n - 1.0), 0.0)
        T = sp.diags((1.0 - fdep) + dd * wd.sum(axis=1)) + sp.csr_matrix(
            ((-dd[:, None] * wd).ravel(), (rows, np.asarray(rcv, np.int64).ravel())), shape=(N, N))
        dprime = np.where(Qt > 0.0, fdep * dt / area, 0.0)
        M = sp.bmat([[T, sp.diags(-dprime)], [sp.diags(area / dt), F]], format="csc")
        step = spla.spsolve(M, np.concatenate([-res, np.zeros(N)]))[:N]
        x = x + step
        if np.abs(step).max() <= tol * max(np.abs(x).max(), 1.0):
            return x, it
    raise RuntimeError("non-linear transport-limited oracle: Newton did not converge in %d steps" % maxit)


def chebyshev_plan(diag, maxit=1000000):
    """Sweeps and weights of the Chebyshev semi-iteration over the Jacobi
    sweep of a diffusion operator with diagonal `diag`.  rho = max over rows
    of (d - 1) / d bounds the spectrum of the Jacobi iteration matrix (real:
    it is similar to a symmetric matrix); the count is the least k with
    2 sigma^k <= 2^-49 (1 - rho), sigma = rho / (1 + sqrt(1 - rho^2)) (the
    error bound 2 sigma^k in the norm that symmetrises the sweep, with 2^-16
    of room for that norm against the max norm, under the 2^-33 of the
    Jacobi bound rho^k / (1 - rho)); omega[0] = 1 (a plain sweep), omega[1] =
    1 / (1 - rho^2 / 2), omega[k] = 1 / (1 - rho^2 omega[k-1] / 4).  Returns
    (rho, omega) or None when the sweep does not contract."""
    d = np.asarray(diag, np.float64)
    d = d[d > 0.0]
    rho = max(float(((d - 1.0) / d).max()), 0.0) if len(d) else 0.0
    if not rho < 1.0:
        return None
    r2 = rho * rho
    sig = rho / (1.0 + math.sqrt(1.0 - r2))
    tgt = 2.0 ** -49 * (1.0 - rho)
    pw, k = 2.0, 0
    while pw > tgt and k < maxit:
        pw *= sig
        k += 1
    om = [1.0]
    for j in range(1, k):
        om.append(1.0 / (1.0 - 0.5 * r2) if j == 1 else 1.0 / (1.0 - 0.25 * r2 * om[-1]))
    return rho, np.array(om, np.float64)


def smooth_solve(coeff, rhs, ngb, nnb):
    """The marine smoothing: Chebyshev semi-iteration over the Jacobi sweep
    of or_hillslope_solve from x = rhs, for the count of chebyshev_plan.
    Returns (x, sweeps)."""
    plan = chebyshev_plan(coeff[:, 0])
    if plan is None:
        raise RuntimeError("Jacobi sweeps do not contract")
    _, om = plan
    N = len(rhs)
    x = np.empty(N)
    lib().or_cheb_smooth(N, _p(ngb), _p(nnb), _p(_c(coeff, np.float64)), _p(_c(rhs, np.float64)), _p(x),
                         len(om), _p(om))
    return x, len(om)


def hillslope_solve(coeff, rhs, ngb, nnb, maxit=1000000):
    """Jacobi sweeps to the bitwise fixed point; raises when ``maxit`` sweeps
    do not reach it (goSPL's Richardson solve would stop at rtol 1e-8)."""
    N = len(rhs)
    x = np.empty(N)
    it = lib().or_hillslope_solve(N, _p(ngb), _p(nnb), _p(_c(coeff, np.float64)), _p(_c(rhs, np.float64)),
                                  _p(x), int(maxit))
    if it > maxit:
        raise RuntimeError("oracle hillslope Jacobi did not reach its fixed point in %d sweeps" % maxit)
    return x, int(it)


# --------------------------------------------------------------------------
# Model-level restatement
# --------------------------------------------------------------------------

class OracleModel(SED, SEA):
    """goSPL's hot-path state machine in FP64 on the CPU (one rank).

    ``geom`` holds input-order arrays: ``ngb`` (N,12) int32 (-1 pad), ``nnb``,
    ``elgt`` (N,12), ``vdist`` (N,12), ``area`` (N,), ``outlet`` (N,) bool
    (draining borders of a flat mesh) and ``flat``.
    """

    def __init__(self, geom, h, *, K, m=0.5, d=0.0, flowdir=3, cda=0.0, cdm=0.0, dt=1000.0,
                 rain=1.0, upsub=None, sedfac=None, sealevel=0.0, ofill=-6000.0,
                 tstart=0.0, tend=None, n=1.0, G=0.0, nodep=True, **sedkw):
        self.ngb = _c(geom["ngb"], np.int32)
        self.nnb = _c(geom["nnb"], np.int32)
        self.elgt = _c(geom["elgt"], np.float64)
        self.vdist = _c(geom["vdist"], np.float64)
        self.larea = _c(geom["area"], np.float64)
        self.flatModel = bool(geom["flat"])
        N = len(self.larea)
        self.lpoints = N
        outlet = np.asarray(geom.get("outlet", np.zeros(N, bool)), bool)
        self.outletIDs = np.nonzero(outlet)[0] if self.flatModel else np.zeros(0, np.int64)
        self.outlet = np.zeros(N, np.int8)
        self.outlet[self.outletIDs] = 1
        self.borders = -np.ones((N, 2), dtype=int)
        self.borders[self.outletIDs, 0] = self.outletIDs
        self.borders[self.outletIDs, 1] = 1
        self.gid = np.arange(N, dtype=np.int32)
        self.locIDs = self.gid
        self.inIDs = np.ones(N, dtype=int)
        self.maxarea = np.array([self.larea.max()])
        self.K, self.spl_m, self.coeffd = float(K), float(m), float(d)
        self.spl_n = float(n)
        self.fDepa = float(G)
        self.flowDir, self.flowExp = int(flowdir), 1.1
        self.Cda, self.Cdm = float(cda), float(cdm)
        self.dt = float(dt)
        self.sealevel = float(sealevel)
        self.oFill = float(ofill)
        self.rainVal = np.broadcast_to(np.asarray(rain, np.float64), (N,)).copy()
        self.rainVal[self.rainVal < 0] = 0.0
        self.upsub = None if upsub is None else np.broadcast_to(np.asarray(upsub, np.float64), (N,)).copy()
        self.sedfacVal = None if sedfac is None else np.broadcast_to(np.asarray(sedfac, np.float64), (N,)).copy()
        if self.sedfacVal is not None:
            # erodibility factors below 0.1 are raised to 0.1 when goSPL loads the
            # map (_updateEroFactor, gospl/mesher/unstructuredmesh.py:1089)
            self.sedfacVal[self.sedfacVal < 0.1] = 0.1
        self.h = np.array(h, dtype=np.float64)
        self.cumED = np.zeros(N)
        self.EbLocal = np.zeros(N)
        self.tStart = float(tstart)
        self.tNow = float(tstart)
        self.tEnd = float(tend) if tend is not None else float(tstart)
        self._cascade_max_steps = 100
        self._cascade_patience = 3
        self._cascade_rel_improve = 1.0e-3
        self._cascade_resid = 0.0
        self.cascade_passes = 0
        self.tmp = np.zeros(N)
        self.bL = self.rainVal * self.larea
        self.coords = None if geom.get("coords") is None else np.asarray(geom["coords"], np.float64)
        seakw = {k: sedkw.pop(k) for k in ("seadepo", "clinSlp", "offshore", "Gmar") if k in sedkw}
        self._init_sed(nodep=nodep, **sedkw)
        self._init_sea(**seakw)

    # ---------------- pit filling (pitfilling.py) ----------------
    def fillElevation(self, sed=False):
        hl = self.h.copy()
        minh = self.h.min()
        if not self.flatModel:
            minh += 1.0e-3
        level = max(minh, self.sealevel + self.oFill)
        self.level = level
        dem = hl - level
        lFill, reached, ledge = fill(dem, self.outlet, self.ngb, self.nnb)
        if not sed:
            idx = lFill < self.sealevel - level
            lFill[idx] = dem[idx]
        self.lFill = lFill + level
        self._pitInformation(hl, level, sed)

        ids = self.pitParams[:, 0] > 0.0
        dh = np.zeros((len(self.pitInfo), 6), dtype=np.float64)
        dh[ids, 0] = self.pitParams[ids, 1] - self.pitParams[ids, 2]
        dh[ids, 1:] = np.expand_dims(self.pitParams[ids, 2] / 5.0, axis=1)
        self.filled_lvl = np.cumsum(dh, axis=1)[:, 1:]
        self.filled_vol = np.zeros((len(self.pitInfo), 5), dtype=np.float64)
        hl_clip = hl.copy()
        if not sed:
            hl_clip[hl_clip < self.sealevel] = self.sealevel
        self.filled_vol[:, :-1] = getpitvol(self.filled_lvl[:, :-1], hl_clip, self.pitIDs,
                                            self.inIDs, self.larea)
        self.filled_vol[:, -1] = self.pitParams[:, 0]

    def _pitInformation(self, hl, level, sed=False):
        pitIDs = label_pits(level if sed else self.sealevel, self.lFill, self.ngb, self.nnb).astype(np.int64)
        pitIDs[self.outletIDs] = -1
        # _transferIDs at one rank: no cross-rank pairs; consecutive ids (pits_cons)
        pitNb = np.unique(pitIDs[pitIDs >= 0])
        newids = -np.ones(len(pitIDs), np.int64)
        sel = pitIDs >= 0
        newids[sel] = np.searchsorted(pitNb, pitIDs[sel]) + 1
        self.pitIDs = newids
        pitnbs = len(pitNb) + 1

        ngb = self.ngb
        valid = ngb >= 0
        cidx = np.where(valid, ngb, 0)
        hc = self.lFill[cidx]
        hi = self.lFill[:, None]
        lower = valid & (hc < hi)
        eqlvl = valid & (hc == hi) & ((self.borders[cidx, 1] > 0) | (self.pitIDs[cidx] == -1))
        ispit = self.pitIDs > -1
        qual = (lower.any(axis=1) | eqlvl.any(axis=1)) & ispit
        qidx = np.where(qual)[0]
        bestlid = np.full(pitnbs, np.iinfo(np.int64).max, dtype=np.int64)
        if qidx.size:
            np.minimum.at(bestlid, self.pitIDs[qidx], self.locIDs[qidx].astype(np.int64))
        spillIDs = -np.ones(pitnbs, dtype=int)
        rank = -np.ones(pitnbs, dtype=int)
        lspill = np.zeros(self.lpoints, dtype=np.int32)
        if qidx.size:
            star = qidx[self.locIDs[qidx].astype(np.int64) == bestlid[self.pitIDs[qidx]]]
            isA = lower[star].any(axis=1)
            sA = star[isA]
            spillIDs[self.pitIDs[sA]] = sA
            rank[self.pitIDs[sA]] = 0
            lspill[sA] = 1
            for i in star[~isA]:
                knb = ngb[i][eqlvl[i]]
                s = int(knb[np.argmin(self.locIDs[knb])])
                spillIDs[self.pitIDs[i]] = s
                rank[self.pitIDs[i]] = 0
                lspill[s] = 1
        self.lspillIDs = np.where(lspill == 1)[0]
        self.pitInfo = np.zeros((pitnbs, 2), dtype=int)
        self.pitInfo[:, 0] = spillIDs
        self.pitInfo[:, 1] = rank
        self.pitIDs[self.outletIDs] = -1

        self._dirFlats()
        idx = self.lFill <= self.sealevel
        h = hl.copy()
        self.lFill[idx] = hl[idx]
        self.flatDirs[idx] = -1
        if not sed:
            h[h < self.sealevel] = self.sealevel
        self._getPitParams(h, pitnbs)

    def _dirFlats(self):
        if (self.pitIDs > -1).any():
            pdir = fill_dir(self.lspillIDs, self.pitIDs, self.lFill, self.ngb, self.nnb)
        else:
            pdir = -np.ones(self.lpoints, dtype=np.int32)
        stp = 0
        while True:
            prev = pdir
            pdir, _ = nghb_dir(self.pitIDs, self.lFill, pdir, self.ngb, self.nnb)
            stp += 1
            if np.count_nonzero(pdir != prev) == 0 or stp > 10000:
                break
        self.ptdir = pdir
        ngb = self.ngb
        n = self.lpoints
        valid = ngb >= 0
        cidx = np.where(valid, ngb, 0)
        cpd = pdir[cidx]
        ch = self.lFill[cidx]
        clid = self.locIDs[cidx].astype(np.int64)
        cand = (valid & (cpd > -1) & (cpd < pdir[:, None]) & (ch <= self.lFill[:, None])
                & (self.pitIDs[:, None] > -1))
        BIG = np.iinfo(np.int64).max
        nid = np.int64(self.locIDs.max()) + 1
        comp = np.where(cand, cpd.astype(np.int64) * nid + clid, BIG)
        kbest = comp.argmin(axis=1)
        rows = np.arange(n)
        has = comp[rows, kbest] < BIG
        self.flatDirs = np.full(n, -1, dtype=np.int32)
        ispit = self.pitIDs > -1
        self.flatDirs[ispit] = np.where(has[ispit], ngb[rows, kbest][ispit], rows[ispit]).astype(np.int32)
        self.flatDirs[self.lspillIDs] = -1

    def _getPitParams(self, hl, nbpits):
        pid = self.pitIDs
        sel = pid > -1
        totv = np.zeros(nbpits)
        np.add.at(totv, pid[sel], ((self.lFill - hl) * self.larea)[sel])
        hmax = np.full(nbpits, MISSING_DATA_SENTINEL)
        np.maximum.at(hmax, pid[sel], self.lFill[sel])
        diffh = np.zeros(nbpits)
        has = np.zeros(nbpits, bool)
        has[pid[sel]] = True
        dtmp = np.full(nbpits, -np.inf)
        np.maximum.at(dtmp, pid[sel], (self.lFill - hl)[sel])
        diffh[has] = dtmp[has]
        self.pitParams = np.stack([totv, hmax, diffh], axis=1)

    # ---------------- flow (flowplex.py) ----------------
    def _buildFlowDirection(self, h, down=True):
        self.seaID = np.where(self.lFill <= self.sealevel)[0]
        donRcvs, self.distRcv, self.wghtVal = mfdreceivers(self.flowDir, self.flowExp, h, self.sealevel,
                                                          self.gid, self.ngb, self.nnb, self.elgt)
        self.rcvID = donRcvs.copy()
        if down:
            sum_weight = np.sum(self.wghtVal, axis=1)
            ids = ((h == self.lFill) & (self.pitIDs > -1) & (self.flatDirs > -1) & (sum_weight == 0.0))
            ids = ids.nonzero()[0]
            self.rcvID[ids, :] = np.tile(ids, (self.flowDir, 1)).T
            self.rcvID[ids, 0] = self.flatDirs[ids]
            self.wghtVal[ids, :] = 0.0
            self.wghtVal[ids, 0] = 1.0
        if self.flatModel:
            self.rcvID[self.outletIDs, :] = np.tile(self.outletIDs, (self.flowDir, 1)).T
            self.distRcv[self.outletIDs, :] = 0.0
            self.wghtVal[self.outletIDs, :] = 0.0
        sum_weight = np.sum(self.wghtVal, axis=1)
        lsink = sum_weight == 0.0
        lsink[self.outletIDs] = False
        lsink[self.seaID] = False
        self.lsink = (lsink.astype(int) * self.inIDs) == 1

    def _solve_flow(self, b, fatal=True):
        """_solve_KSP (flowplex.py:323-427) on the flow matrix.  A receiver
        cycle that does not drain makes goSPL's solvers fail: the main solve
        (fatal) raises, a cascade solve is zeroed (flowplex.py:304-320)."""
        q, bad = flowacc(self.rcvID, self.wghtVal, b, self.ngb, self.nnb)
        self.undrained = bad
        if bad:
            if fatal:
                raise RuntimeError("flow accumulation did not converge (%d undrained vertices)" % bad)
            q = np.zeros_like(q)
        return q

    def _distributeDownstream(self, pitVol, FA, hl, step):
        excess = False
        FA = np.multiply(FA, self.inIDs)
        inV = np.zeros(len(self.pitParams), dtype=np.float64)
        sel = self.lsink & (self.pitIDs > -1)
        np.add.at(inV, self.pitIDs[sel], FA[sel])
        eV = inV - pitVol
        nFA = None
        if (eV > 0.0).any():
            eIDs = eV > 0.0
            pitVol[eIDs] = 0.0
            spillIDs = self.pitInfo[eIDs, 0]
            localSpill = np.where(self.pitInfo[eIDs, 1] == 0)[0]
            localPts = spillIDs[localSpill]
            nFA = np.zeros(self.lpoints, dtype=np.float64)
            nFA[localPts] = eV[eIDs][localSpill]
            ids = np.isin(self.pitIDs, np.where(eV > 0.0)[0])
            self.waterFilled[ids] = self.lFill[ids]
        if (eV < 0.0).any():
            eIDs = np.where((eV < 0.0) & (inV > 0.0))[0]
            pitVol[eIDs] += eV[eIDs]
            nid = np.absolute(self.filled_vol[eIDs] - pitVol[eIDs][:, None]).argmin(axis=1)
            fill_lvl = self.filled_lvl[eIDs, nid]
            in_eIDs = np.isin(self.pitIDs, eIDs)
            pit_fill_lvl = np.zeros(len(self.pitParams))
            pit_fill_lvl[eIDs] = fill_lvl
            node_lvl = pit_fill_lvl[self.pitIDs]
            mask = in_eIDs & (self.waterFilled <= node_lvl)
            self.waterFilled[mask] = node_lvl[mask]
        if (eV > 1.0e-3).any():
            if step == 100:
                self._buildFlowDirection(self.lFill)
            else:
                self._buildFlowDirection(self.waterFilled)
            rhs = nFA / self.dt
            resid = rhs.sum()
            self._cascade_resid = float(resid)
            if resid > self.maxarea[0]:
                excess = True
                q = self._solve_flow(rhs, fatal=False)
                nFA = q * self.dt
                FA = nFA.copy()
                FA[hl < self.waterFilled] = 0.0
                self.FAL = self.FAL + FA / self.dt
            else:
                nFA = None
        else:
            nFA = None
        return excess, pitVol, nFA

    def flowAccumulation(self):
        self.fillElevation()
        pitVol = self.pitParams[:, 0].copy()
        hl = self.h.copy()
        self._buildFlowDirection(hl, False)
        self.wghtVali = self.wghtVal.copy()
        self.rcvIDi = self.rcvID.copy()
        self.distRcvi = self.distRcv.copy()
        self.lsinki = self.lsink.copy()
        rainA = self.bL.copy()
        rainA[self.seaID] = 0.0
        self.FAL = self._solve_flow(rainA)
        self.waterFilled = hl.copy()
        self.cascade_passes = 0
        if (pitVol > 0.0).any():
            FA = self.FAL.copy() * self.dt
            excess = True
            step = 0
            best_resid = None
            stall = 0
            while excess and step <= self._cascade_max_steps:
                excess, pitVol, FA = self._distributeDownstream(pitVol, FA, hl, step)
                step += 1
                if excess:
                    resid = self._cascade_resid
                    if best_resid is None or resid < best_resid * (1.0 - self._cascade_rel_improve):
                        best_resid = resid
                        stall = 0
                    else:
                        stall += 1
                        if stall >= self._cascade_patience:
                            break
            self.cascade_passes = step
            FA = self.FAL.copy()
            ids = (hl <= self.waterFilled) & (self.pitIDs > -1)
            FA[ids] = 0.0
            self.FAL = FA.copy()
            FA[ids] = self.FAL.max() * 0.1
            self.fillFAL = FA
        else:
            self.fillFAL = self.FAL.copy()
        self.FAL = np.clip(self.FAL, 0.0, None)
        self.fillFAL = np.clip(self.fillFAL, 0.0, None)
        self.waterFilled = self.waterFilled - hl

    # ---------------- erosion (SPL.py) ----------------
    def _eroMats(self, hOld):
        PA = self.FAL
        if self.sedfacVal is not None:
            Kbr = self.K * 1.0 * self.sedfacVal * (self.rainVal ** self.coeffd)
        else:
            Kbr = self.K * 1.0 * (self.rainVal ** self.coeffd)
        Kbr = Kbr * (self.dt * (PA ** self.spl_m))
        Kbr[self.seaID] = 0.0
        N = self.lpoints
        wght = self.wghtVali.copy()
        nodes = np.arange(N)
        wght[self.seaID, :] = 0.0
        diag = np.ones(N)
        off = np.zeros((N, self.flowDir))
        for k in range(self.flowDir):
            dh = hOld - hOld[self.rcvIDi[:, k]]
            limiter = np.divide(dh, dh + 1.0e-2, out=np.zeros_like(dh), where=dh != 0)
            data = np.divide(Kbr * limiter, self.distRcvi[:, k], out=np.zeros_like(PA),
                             where=self.distRcvi[:, k] != 0)
            data = np.multiply(data, -wght[:, k])
            data[self.rcvIDi[:, k] == nodes] = 0.0
            off[:, k] = data
            diag = diag - data
        return off, diag

    def erodepSPL(self):
        hOld = self.h.copy()
        off, diag = self._eroMats(hOld)
        if self.fDepa > 0.0:
            # transport-limited (SPL.py:306-333): the coupled h / q system
            self.fDep = transport_fdep(self.fDepa, self.larea, self.FAL, self.seaID,
                                       self.outletIDs if self.flatModel else None)
            stepED, self.sedFluxED = coupled_ed_solve(self.rcvIDi, off, diag, self.rcvIDi, self.wghtVali,
                                                      self.fDep, self.larea, self.dt, hOld)
            bad = 0
        else:
            stepED, bad = spl_solve(self.rcvIDi, off, diag, hOld)
        if bad:
            raise RuntimeError("SPL receiver graph has a cycle")
        E = (stepED - hOld) / self.dt
        if self.flatModel:
            E[self.outletIDs] = 0.0
        E[self.lsink] = 0.0
        self.Eb = E
        self.tmp = E * self.dt
        self.cumED = self.cumED + self.tmp
        self.h = self.h + self.tmp
        self.EbLocal = self.tmp / self.dt

    # ---------------- non-linear erosion (nlSPL.py) ----------------
    def erodepSPLnl(self):
        """nlSPL._solveNL + erodepSPLnl (gospl/eroder/nlSPL.py:321-456); G > 0
        goes to nl_transport_solve.  For G = 0 goSPL solves F(x) = x - hOld + Kbr S(x)^n = 0 with SNES
        (rtol/atol 1e-6); S_i = sum over receivers with w > 0 and x_i > x_r of
        w (x_i - x_r) / d (local_spl, functions.F90:1469-1502).  F_i depends on
        x_i and its receivers only, so here each vertex is solved exactly in
        order of increasing hOld (a receiver with w > 0 is lower): F_i is
        continuous and increasing in x_i, with a root in [min(hOld_i, min x_r),
        hOld_i], found by safeguarded Newton to the last bits."""
        hOld = self.h.copy()
        PA = self.FAL
        dh = (hOld[:, None] - hOld[self.rcvIDi]).max(axis=1)
        elim = np.divide(dh, dh + 1.0e-2, out=np.zeros_like(dh), where=dh != 0)
        if self.sedfacVal is not None:
            Kbr = self.K * 1.0 * self.sedfacVal * (self.rainVal ** self.coeffd)
        else:
            Kbr = self.K * 1.0 * (self.rainVal ** self.coeffd)
        Kbr = Kbr * (self.dt * (PA ** self.spl_m) * elim)
        Kbr[self.seaID] = 0.0
        n = self.spl_n
        x = hOld.copy()
        rcv, wgt, dist = self.rcvIDi, self.wghtVali, self.distRcvi
        if self.fDepa > 0.0:
            # transport-limited (nlSPL._solveNL_ed, nlSPL.py:255-320)
            self.fDep = transport_fdep(self.fDepa, self.larea, PA, self.seaID,
                                       self.outletIDs if self.flatModel else None)
            x, self.nl_iters = nl_transport_solve(rcv, wgt, dist, Kbr, n, self.fDep, self.larea, self.dt, hOld)
            order = []
        else:
            order = np.argsort(hOld, kind="stable")
        for i in order:
            kb = Kbr[i]
            if kb == 0.0:
                continue
            ks = [k for k in range(rcv.shape[1]) if wgt[i, k] > 0.0]
            if not ks:
                continue
            xr = np.array([x[rcv[i, k]] for k in ks])
            wd = np.array([wgt[i, k] / dist[i, k] for k in ks])

            def f(y):
                on = y > xr
                S = float(np.sum(wd[on] * (y - xr[on])))
                return y - hOld[i] + kb * (S ** n if S > 0.0 else 0.0), S, float(np.sum(wd[on]))

            lo, hi = min(hOld[i], xr.min()), hOld[i]
            y = hi
            for _ in range(400):
                fy, S, c = f(y)
                if fy == 0.0:
                    break
                if fy > 0.0:
                    hi = y
                else:
                    lo = y
                dfy = 1.0 + (n * kb * S ** (n - 1.0) * c if S > 0.0 else 0.0)
                yn = y - fy / dfy
                if not (lo < yn < hi):
                    yn = 0.5 * (lo + hi)
                if abs(yn - y) <= 1.0e-15 * max(abs(y), 1.0) or hi - lo <= 1.0e-15 * max(abs(y), 1.0):
                    y = yn
                    break
                y = yn
            x[i] = y
        self.nl_x = x
        E = (x - hOld) / self.dt
        if self.flatModel:
            E[self.outletIDs] = 0.0
        E[self.lsink] = 0.0
        self.Eb = E
        self.tmp = E * self.dt
        self.cumED = self.cumED + self.tmp
        self.h = self.h + self.tmp
        self.EbLocal = self.tmp / self.dt

    # ---------------- hillslope (hillslope.py) ----------------
    def getHillslope(self):
        h = self.h.copy()
        self.seaID = np.where(h <= self.sealevel)[0]
        if not (self.Cda == 0.0 and self.Cdm == 0.0):
            Cd = np.full(self.lpoints, self.Cda)
            Cd[self.seaID] = self.Cdm
            coeff = sethillslopecoeff(Cd * self.dt, self.ngb, self.nnb, self.larea, self.elgt, self.vdist)
            if self.flatModel:
                coeff[self.outletIDs, 1:] = 0.0
                coeff[self.outletIDs, 0] = 1.0
            hOld = self.h.copy()
            hnew, self.hill_iters = hillslope_solve(coeff, hOld, self.ngb, self.nnb)
            self.h = hnew
            self.tmp = self.h - hOld
            self.cumED = self.cumED + self.tmp
        self.EbLocal = self.EbLocal + self.tmp / self.dt

    # ---------------- forcing / driver ----------------
    def applyForces(self):
        self.bL = self.rainVal * self.larea
        if self.upsub is not None:
            self.h = self.h + self.upsub * self.dt

    def step(self):
        """One iteration of the runProcesses loop body (model.py:379-450)."""
        self.flowAccumulation()
        if self.spl_n == 1.0:
            self.erodepSPL()
        else:
            self.erodepSPLnl()
        if not self.nodep:
            self.flowAccumulation()
            self.sedChange()
            if self.seaDepo:
                self.seaChange()
        self.getHillslope()
        if self.tNow < self.tEnd:
            self.applyForces()
        self.tNow += self.dt

    def initialise(self):
        """model.py:284-295: applyForces then the init-time flowAccumulation."""
        self.applyForces()
        self.flowAccumulation()

    def runProcesses(self):
        while self.tNow <= self.tEnd:
            if self.tNow == self.tEnd:
                break
            self.step()


# --------------------------------------------------------------------------
# Reference-solver mode (CPU baseline): goSPL's solvers instead of exact solves
# --------------------------------------------------------------------------

def _rs_lib():
    L = lib()
    if not getattr(L, "_rs_declared", False):
        P, I64, I32, F64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.rs_flow_solve.argtypes = [I64, I32, P, P, P, P, F64, I32, I32, dp]
        L.rs_flow_solve.restype = I32
        L.rs_spl_solve.argtypes = [I64, I32, P, P, P, P, P, F64, I32, I32, dp]
        L.rs_spl_solve.restype = I32
        L.rs_hill_solve.argtypes = [I64, P, P, P, P, P, F64, I32, I32, dp]
        L.rs_hill_solve.restype = I32
        L._rs_declared = True
    return L


class ReferenceSolverModel(OracleModel):
    """The oracle with goSPL's own solver choices in place of exact solves,
    as goSPL runs on one MPI rank (PETSc block Jacobi = one ILU(0) block):
    FGMRES(30) + ILU(0) at rtol 1e-8 for the flow systems, max_it 5000 (1000
    for cascade solves), warm-started from the previous solution or seeded
    with the right-hand side (flowplex.py:323-427); the same for the
    stream-power system, warm-started from the previous solution
    (SPL.py:260-366); Richardson + ILU(0) at rtol 1e-8 for hillslope
    diffusion, started from h (hillslope.py:347-371).  Fill (heap
    priority-flood) and receivers are the oracle's serial kernels.  `threads`
    OpenMP threads run the sparse products and vector updates."""

    def __init__(self, geom, h, *, threads=1, rtol=1.0e-8, **kw):
        super().__init__(geom, h, **kw)
        self.threads = int(threads)
        self.rtol = float(rtol)
        self._fa_prev = None
        self._cq_prev = None
        self._stepED = None
        self.solver_its = []

    def _solve_flow(self, b, fatal=True):
        L = _rs_lib()
        prev = self._fa_prev if fatal else self._cq_prev
        x = np.array(b, dtype=np.float64, copy=True) if prev is None or not np.any(prev) else prev.copy()
        rr = ctypes.c_double(0.0)
        it = L.rs_flow_solve(self.lpoints, self.flowDir, _p(_c(self.rcvID, np.int32)),
                             _p(_c(self.wghtVal, np.float64)), _p(_c(b, np.float64)), _p(x), self.rtol,
                             5000 if fatal else 1000, self.threads, ctypes.byref(rr))
        self.solver_its.append(("flow", int(it), float(rr.value)))
        self.undrained = 0 if it >= 0 else 1
        if it < 0:
            if fatal:
                raise RuntimeError("FGMRES did not converge on the flow system (relres %.3e)" % rr.value)
            x = np.zeros_like(x)
        if fatal:
            self._fa_prev = x.copy()
        else:
            self._cq_prev = x.copy()
        return x

    def erodepSPL(self):
        L = _rs_lib()
        hOld = self.h.copy()
        off, diag = self._eroMats(hOld)
        x = (self._stepED if self._stepED is not None else hOld).copy()
        rr = ctypes.c_double(0.0)
        it = L.rs_spl_solve(self.lpoints, self.flowDir, _p(_c(self.rcvIDi, np.int32)), _p(_c(off, np.float64)),
                            _p(_c(diag, np.float64)), _p(hOld), _p(x), self.rtol, 5000, self.threads,
                            ctypes.byref(rr))
        self.solver_its.append(("spl", int(it), float(rr.value)))
        if it < 0:
            raise RuntimeError("FGMRES did not converge on the stream-power system")
        self._stepED = x.copy()
        E = (x - hOld) / self.dt
        if self.flatModel:
            E[self.outletIDs] = 0.0
        E[self.lsink] = 0.0
        self.Eb = E
        self.tmp = E * self.dt
        self.cumED = self.cumED + self.tmp
        self.h = self.h + self.tmp
        self.EbLocal = self.tmp / self.dt

    def getHillslope(self):
        L = _rs_lib()
        h = self.h.copy()
        self.seaID = np.where(h <= self.sealevel)[0]
        if not (self.Cda == 0.0 and self.Cdm == 0.0):
            Cd = np.full(self.lpoints, self.Cda)
            Cd[self.seaID] = self.Cdm
            coeff = sethillslopecoeff(Cd * self.dt, self.ngb, self.nnb, self.larea, self.elgt, self.vdist)
            if self.flatModel:
                coeff[self.outletIDs, 1:] = 0.0
                coeff[self.outletIDs, 0] = 1.0
            hOld = self.h.copy()
            x = hOld.copy()
            rr = ctypes.c_double(0.0)
            it = L.rs_hill_solve(self.lpoints, _p(self.ngb), _p(self.nnb), _p(_c(coeff, np.float64)), _p(hOld),
                                 _p(x), self.rtol, 100000, self.threads, ctypes.byref(rr))
            self.solver_its.append(("hillslope", int(it), float(rr.value)))
            if it < 0:
                raise RuntimeError("Richardson did not converge on the diffusion system")
            self.h = x
            self.tmp = self.h - hOld
            self.cumED = self.cumED + self.tmp
        self.EbLocal = self.EbLocal + self.tmp / self.dt
