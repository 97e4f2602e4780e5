"""Reverse-mode gradients of the surface-process path (no counterpart in goSPL).

``Evolution`` runs ``nsteps`` of the hot path (goSPL's runProcesses loop body,
gospl/model.py:379-453: flowAccumulation, erodepSPL, getHillslope,
applyForces) with a per-vertex ``log10 K`` and a per-vertex uplift rate as
inputs, and gives the gradient of a function of the final elevation with
respect to both.

What is differentiated.  Each step's flow data -- receivers, MFD weights and
distances, discharge (FAL), filled elevation, sink flags (flowplex.py:
480-544, 752-945) -- and the hillslope operator's sea flags are held at their
forward values (the "frozen" step); through them the step is smooth:

* stream power (SPL.py:45-114, 402-438): A(h) x = h with
  A = I + sum_k c_ik (e_i - e_rk)^T, c_ik = K_i dt FAL_i^m w_ik / d_ik
  l(h_i - h_rk), l(s) = s / (s + 0.01); h' = x except at flat-mesh outlets and
  sink cells, which keep h;
* hillslope (hillslope.py:235-293, 614-649): h'' = M^-1 h', M the diffusion
  operator of the frozen sea flags;
* uplift (unstructuredmesh.py:743-794): h_next = h'' + U dt.

The reverse sweep is the transposed chain: ``gospl_hillslope_adjoint`` (M^T,
Jacobi through the reverse-neighbour table), ``gospl_spl_adjoint`` (A^T by
one sweep from the sources down, then the local terms of dA/dh and
dA/dlog10 K), uplift's dt.  Memory: the elevation at every
``checkpoint_every``-th step is kept; the reverse pass replays each segment
forward from its checkpoint (the same library calls, so the replay repeats
the forward bit for bit) to record the segment's frozen data, then walks it
backwards.  The forward runs the exact FP64 solves.

``evolve(evo, logk, uplift)`` wraps it as a ``torch.autograd.Function``.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .model import DeviceMesh, Model, _ptr

LN10 = math.log(10.0)


@dataclass
class FrozenStep:
    """One step's forward data that the derivative holds fixed, plus the
    step's pre-erosion elevation."""
    h: torch.Tensor        # elevation entering erodepSPL
    rcv: torch.Tensor      # (K, N) main receivers
    wgt: torch.Tensor      # (K, N)
    dist: torch.Tensor     # (K, N)
    fal: torch.Tensor      # discharge
    lfill: torch.Tensor    # filled elevation
    lsink: torch.Tensor    # int8 sink flags passed to erodepSPL
    sea: torch.Tensor = None     # int8 hillslope sea flags (h' <= sea level)
    uplift: bool = False         # applyForces ran at the end of the step


@dataclass
class RunState:
    """What a forward run leaves for its gradient: the inputs and the
    checkpointed elevations (step index -> elevation entering the step)."""
    logk: torch.Tensor
    uplift: torch.Tensor
    checkpoints: dict
    init_uplift: bool
    steps: list = None           # every step's frozen data (record=True)


class Evolution:
    """nsteps of the surface-process path as a differentiable map
    (log10 K, uplift) -> final elevation.

    ``mesh`` a ``Mesh`` or ``DeviceMesh``; ``h0`` the initial elevation
    (device order; goSPL's initialise adds one uplift step to it before the
    first step, as Model.initialise does).  Other keywords as ``Model``
    (``m, cda, cdm, dt, rain, sealevel, flowdir, ofill``); ``K`` and
    ``upsub`` are the inputs and ``d``/``sedfac`` are not supported.
    """

    def __init__(self, mesh, h0, *, nsteps, dt, checkpoint_every=10, device="cuda", **model_kw):
        for bad in ("K", "upsub", "sedfac", "d", "refine", "refine_dag", "schedule"):
            if bad in model_kw:
                raise TypeError("Evolution: %s is not a keyword here (K and the uplift are inputs; the "
                                "exact FP64 solves are differentiated)" % bad)
        if nsteps < 1 or checkpoint_every < 1:
            raise ValueError("Evolution: nsteps and checkpoint_every must be >= 1")
        self.dm = mesh if isinstance(mesh, DeviceMesh) else DeviceMesh(mesh, device)
        self.device = self.dm.device
        h0 = h0.detach().cpu().numpy() if isinstance(h0, torch.Tensor) else h0
        self.h0_host = np.array(h0, dtype=np.float64)
        self.h0 = torch.as_tensor(self.h0_host).to(self.device)
        self.nsteps = int(nsteps)
        self.dt = float(dt)
        self.every = int(checkpoint_every)
        self.kw = dict(model_kw)
        self.model = None

    # ------------------------------------------------------------------ forward
    def _new_model(self, logk, uplift):
        m = Model(self.dm, self.h0_host, K=1.0, dt=self.dt, upsub=None, tstart=0.0,
                  tend=self.nsteps * self.dt, **self.kw)
        m._kfac = torch.pow(10.0, logk.to(torch.float64)).contiguous()
        m.upsub = uplift.to(torch.float64).contiguous()
        return m

    def _check(self, logk, uplift):
        N = self.dm.N
        for name, t in (("logk", logk), ("uplift", uplift)):
            dev = t.device
            same = dev.type == self.device.type and (self.device.index is None or dev.index == self.device.index)
            if t.shape != (N,) or not same:
                raise ValueError("Evolution: %s must be a (%d,) tensor on %s" % (name, N, self.device))

    def _step(self, m, rec):
        """Model.step with the frozen data of the step recorded in ``rec``."""
        m.flowAccumulation()
        fs = None
        if rec is not None:
            fs = FrozenStep(h=m.hLocal.clone(), rcv=m.rcvIDi.clone(), wgt=m.wghtVali.clone(),
                            dist=m.distRcvi.clone(), fal=m.FAL.clone(), lfill=m.lFill.clone(),
                            lsink=m.lsink.clone())
        m.erodepSPL()
        if fs is not None:
            fs.sea = (m.hLocal <= m.sealevel).to(torch.int8)
        m.getHillslope()
        up = m.tNow < m.tEnd
        if up:
            m.applyForces()
        m.tNow += m.dt
        if fs is not None:
            fs.uplift = up
            rec.append(fs)

    def run(self, logk, uplift, record=False):
        """Forward run; returns (final elevation, RunState).  record=True
        keeps every step's frozen data (for ``frozen_run``)."""
        self._check(logk, uplift)
        m = self._new_model(logk, uplift)
        self.model = m
        m.applyForces()                   # Model.initialise: forcing ...
        init_up = m.upsub is not None
        st = RunState(logk=logk.detach().clone(), uplift=uplift.detach().clone(), checkpoints={},
                      init_uplift=init_up, steps=[] if record else None)
        for n in range(self.nsteps):
            if n % self.every == 0:
                st.checkpoints[n] = m.hLocal.clone()
            self._step(m, st.steps)
        return m.hLocal.clone(), st

    def _replay(self, st, n0, n1):
        """Frozen data of steps n0..n1-1, replayed from the checkpoint at n0."""
        m = self._new_model(st.logk, st.uplift)
        m.hLocal = st.checkpoints[n0].clone()
        m.tNow = n0 * self.dt
        rec = []
        for _ in range(n0, n1):
            self._step(m, rec)
        return m, rec

    # ------------------------------------------------------------------ reverse
    def _hill_sweeps(self, m):
        """The forward's Jacobi count, or, when the forward iterates to
        convergence, the count from the same contraction bound (2^-60)."""
        n = m.hillslope_sweeps()
        if n > 0:
            return n
        s = max(m.Cda, m.Cdm) * m.dt * self.dm.gmax
        rho = s / (1.0 + s)
        return int(math.ceil(math.log(2.0 ** -60) / math.log(rho))) + 1

    def _reverse_step(self, m, fs, lam, g_logk, g_up, kfac, out_h):
        lib, H, S = m.lib, m.H, m.stream
        if fs.uplift:
            g_up.add_(lam, alpha=self.dt)
        if m.Cda != 0.0 or m.Cdm != 0.0:
            lam2 = torch.empty_like(lam)
            _lib.check(lib.gospl_hillslope_adjoint(H, m.Cda, m.Cdm, m.dt, _ptr(fs.sea), self._hill_sweeps(m),
                                                   _ptr(lam), _ptr(lam2), S), "gospl_hillslope_adjoint")
        else:
            lam2 = lam
        gk = torch.empty_like(lam)
        _lib.check(lib.gospl_spl_adjoint(H, m.flowDir, _ptr(fs.rcv), _ptr(fs.wgt), _ptr(fs.dist), _ptr(fs.fal),
                                         _ptr(kfac), 1.0, m.spl_m, m.dt, m.sealevel, _ptr(fs.lfill),
                                         _ptr(fs.lsink), _ptr(fs.h), None, _ptr(lam2), _ptr(out_h), _ptr(gk),
                                         None, S), "gospl_spl_adjoint")
        g_logk.add_(gk)

    def gradient(self, st, grad_h):
        """dL/dlog10 K and dL/duplift given grad_h = dL/d(final elevation)."""
        lam = grad_h.to(torch.float64).contiguous().clone()
        g_logk = torch.zeros_like(lam)
        g_up = torch.zeros_like(lam)
        kfac = torch.pow(10.0, st.logk).contiguous()
        starts = sorted(st.checkpoints)
        for si in range(len(starts) - 1, -1, -1):
            n0 = starts[si]
            n1 = starts[si + 1] if si + 1 < len(starts) else self.nsteps
            m, rec = self._replay(st, n0, n1)
            for fs in reversed(rec):
                nxt = torch.empty_like(lam)
                self._reverse_step(m, fs, lam, g_logk, g_up, kfac, nxt)
                lam = nxt
            del rec
        if st.init_uplift:
            g_up.add_(lam, alpha=self.dt)
        return g_logk, g_up

    # ------------------------------------------------------------------ frozen map
    def frozen_run(self, st, logk, uplift):
        """The frozen map itself: the recorded steps of ``st`` (run with
        record=True) re-run with other inputs, flow data and sea flags held.
        Its central differences are what ``gradient`` must match."""
        if st.steps is None:
            raise ValueError("frozen_run needs a run(..., record=True) state")
        self._check(logk, uplift)
        m = self.model if self.model is not None else self._new_model(logk, uplift)
        lib, H, S = m.lib, m.H, m.stream
        kfac = torch.pow(10.0, logk.to(torch.float64)).contiguous()
        up = uplift.to(torch.float64).contiguous()
        h = self.h0.clone()
        if st.init_uplift:
            h += up * self.dt             # as applyForces rounds it
        scratch = [torch.zeros_like(h) for _ in range(3)]
        hout = torch.empty_like(h)
        sweeps = self._hill_sweeps(m)
        for fs in st.steps:
            _lib.check(lib.gospl_spl(H, m.flowDir, _ptr(fs.rcv), _ptr(fs.wgt), _ptr(fs.dist), _ptr(fs.fal),
                                     _ptr(kfac), 1.0, m.spl_m, m.dt, m.sealevel, _ptr(fs.lfill), _ptr(fs.lsink),
                                     _ptr(h), _ptr(scratch[0]), _ptr(scratch[1]), _ptr(scratch[2]), None, S),
                       "gospl_spl")
            if m.Cda != 0.0 or m.Cdm != 0.0:
                _lib.check(lib.gospl_hillslope_frozen(H, m.Cda, m.Cdm, m.dt, _ptr(fs.sea), sweeps, _ptr(h),
                                                      _ptr(hout), S), "gospl_hillslope_frozen")
                h, hout = hout, h
            if fs.uplift:
                h += up * self.dt
        return h


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, logk, uplift, evo):
        h, st = evo.run(logk.detach(), uplift.detach())
        ctx.evo = evo
        ctx.st = st
        return h

    @staticmethod
    def backward(ctx, grad_h):
        g_logk, g_up = ctx.evo.gradient(ctx.st, grad_h)
        return g_logk, g_up, None


def evolve(evo: Evolution, logk: torch.Tensor, uplift: torch.Tensor) -> torch.Tensor:
    """Final elevation of ``evo``'s run as an autograd node of (logk, uplift)."""
    return _Evolve.apply(logk, uplift, evo)
