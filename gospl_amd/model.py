"""goSPL's hot-path driver on one MI355X.

``Model`` keeps the method names and state attributes of goSPL's mixin
``Model`` (gospl/model.py:179-466) for the per-step surface-process path:

* ``flowAccumulation``      gospl/flow/flowplex.py:752-945
* ``fillElevation``         gospl/flow/pitfilling.py:849-903
* ``_buildFlowDirection``   gospl/flow/flowplex.py:480-544
* ``_distributeDownstream`` gospl/flow/flowplex.py:576-686
* ``erodepSPL``             gospl/eroder/SPL.py:402-454
* ``getHillslope``          gospl/sed/hillslope.py:614-649
* ``applyForces``           gospl/mesher/unstructuredmesh.py:743-794
* ``runProcesses``          gospl/model.py:331-466

State lives in torch tensors on the GPU in device (Morton) order; the
compute is the HIP library behind ``gospl_amd._lib`` (no CPU fallback: a
missing library or GPU is an error).  Per-pit bookkeeping of the overflow
cascade (a few thousand numbers) runs on the host as in goSPL.
"""

from __future__ import annotations

import collections
import ctypes
import math
import os
import time
import warnings

import numpy as np
import torch

from . import _lib
from .mesh import Mesh
from .outmesh import WriteMesh
from .sea import SEAMesh
from .sed import SEDMesh

MISSING_DATA_SENTINEL = -1.0e8   # gospl/tools/constants.py:34


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def contraction_gmax(mesh):
    """max_i sum_p vdist/elgt / area_i: the hillslope operator's contraction
    bound is s/(1+s) with s = Kd*dt*gmax."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(mesh.elgt > 0, mesh.vdist / np.where(mesh.elgt > 0, mesh.elgt, 1.0), 0.0)
        g = np.where(mesh.area > 0, ratio.sum(axis=0) / np.where(mesh.area > 0, mesh.area, 1.0), 0.0)
    return float(np.max(g)) if len(g) else 0.0


def sweeps_for(kd_dt, gmax, bits):
    """Jacobi sweeps that bring the contraction bound below 2^-bits (at least
    2); 0 past 64 sweeps."""
    s = kd_dt * gmax
    if s <= 0.0:
        return 1
    rho = s / (1.0 + s)
    n = int(math.ceil(math.log(2.0 ** -bits) / math.log(rho))) + 1
    return max(2, n) if n <= 64 else 0


class DeviceMesh:
    """A ``Mesh`` uploaded to the GPU plus its library handle."""

    def __init__(self, mesh: Mesh, device="cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("gospl_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.mesh = mesh
        self.device = torch.device(device)
        self.N = mesh.N
        self.W = mesh.width
        self.flat = mesh.flat
        dev = self.device
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
        self.nbr = t(mesh.nbr, torch.int32)
        self.deg = t(mesh.deg, torch.int32)
        self.elgt = t(mesh.elgt, torch.float64)
        self.vdist = t(mesh.vdist, torch.float64)
        self.area = t(mesh.area, torch.float64)
        self.gid = t(mesh.gid, torch.int32)
        self.outlet = t(mesh.outlet, torch.int8)
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.gospl_mesh_create(
            self.N, self.W, 1 if self.flat else 0, _ptr(self.nbr), _ptr(self.deg), _ptr(self.elgt),
            _ptr(self.vdist), _ptr(self.area), _ptr(self.gid), _ptr(self.outlet), ctypes.byref(h)),
            "gospl_mesh_create")
        self.handle = h
        self.gmax = contraction_gmax(mesh)
        self.outletIDs = torch.nonzero(self.outlet > 0).flatten() if self.flat else None

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.gospl_mesh_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model(SEDMesh, WriteMesh, SEAMesh):
    """Per-step hot path of goSPL on one GPU (FP64 state).

    ``Model.sweeps`` counts, process-wide, the exact sweeps launched by every
    model (keys ``"flow"``, ``"spl"``): one launch each of the level-0 and the
    dataflow kernel, so a profile's dispatches can be matched to a window.

    ``mesh`` is a ``gospl_amd.mesh.Mesh`` (or ``DeviceMesh``), ``h0`` the
    initial elevation in device order; keyword parameters follow goSPL's YAML
    keys: ``K, m, d`` (spl), ``flowdir``, ``cda/cdm`` (hillslopeKa/Km),
    ``dt``, ``rain`` (scalar or per-vertex m/yr), ``upsub`` (per-vertex m/yr
    or None), ``sedfac`` (per-vertex K factor or None), ``sealevel``,
    ``ofill`` (oFill), ``tstart``/``tend``.

    ``refine=True`` solves hillslope diffusion by FP32 Jacobi corrections of
    an FP64 residual (``gospl_hillslope_refine``, to ``||r|| <= refine_rtol
    ||h||``); ``refine_dag=True`` also solves flow accumulation and
    stream-power erosion by refinement (``gospl_flowacc_refine`` /
    ``gospl_spl_refine``: fixed-point correction sweeps and FP64 residuals,
    at most ``refine_maxpass`` corrections; the flow solves warm-start from
    their previous solutions as goSPL's KSP does) instead of the exact FP64
    sweeps, which are direct solves.  State, fill, pits and receivers are
    FP64 / exact either way.
    """

    sweeps = {"flow": 0, "spl": 0}

    def __init__(self, mesh, h0=None, *args, K=None, m=0.5, d=0.0, n=1.0, G=0.0, flowdir=3, cda=0.0, cdm=0.0, dt=1000.0,
                 rain=1.0, upsub=None, sedfac=None, sealevel=0.0, ofill=-6000.0, tstart=0.0,
                 tend=None, device="cuda", verbose=False, refine=False, refine_dag=False,
                 refine_rtol=1.0e-12, refine_maxpass=4, showlog=False, profile=False, schedule=None,
                 outputDir=None, makedir=False, tout=None, rstep=0, compression="gzip", finput=None,
                 nodep=True, seadepo=False, nlPitVolume=1.0e9, nlPitDepth=100.0, nlPitK=None, nonlinKm=10.0,
                 pitInletBias=0.10, picardSub=10, picardIts=2, marineSolver="ts", clinSlp=1.0e-6, offshore=100.0e5,
                 Gmar=0.0):
        if isinstance(mesh, (str, os.PathLike)):
            # goSPL's constructor, Model(filename, verbose, showlog, profile)
            # (gospl/model.py:212-304): the YAML file and its npz mesh
            flags = ([h0] if h0 is not None else []) + list(args)
            verbose = bool(flags[0]) if len(flags) > 0 else verbose
            showlog = bool(flags[1]) if len(flags) > 1 else showlog
            profile = bool(flags[2]) if len(flags) > 2 else profile
            (m_, h_), params = Model._from_file(mesh, device)
            self.__init__(m_, h_, device=device, verbose=verbose, showlog=showlog, profile=profile,
                          refine=refine, refine_dag=refine_dag, refine_rtol=refine_rtol,
                          refine_maxpass=refine_maxpass, **params)
            return
        if h0 is None or K is None:
            raise TypeError("Model(mesh, h0, K=..., ...) or Model(filename)")
        self.showlog = bool(showlog)
        self._schedule = schedule or {}
        self.dm = mesh if isinstance(mesh, DeviceMesh) else DeviceMesh(mesh, device)
        self.lib = self.dm.lib
        self.H = self.dm.handle
        dev = self.dm.device
        self.device = dev
        N = self.dm.N
        self.lpoints = N
        self.flatModel = self.dm.flat
        self.K, self.spl_m, self.coeffd = float(K), float(m), float(d)
        self.spl_n = float(n)
        self.fDepa = float(G)
        if self.fDepa < 0.0:
            raise ValueError("G (the deposition coefficient) must be >= 0")
        # transport-limited solve: GMRES to this relative residual, at most this many iterations
        self.ed_rtol, self.ed_maxit = 1.0e-13, 400
        self.ed_iters, self.ed_relres = 0, 0.0
        self.nl_tol, self.nl_maxit = 1.0e-13, 100
        self.nl_iters = 0
        self.flowDir, self.flowExp = int(flowdir), 1.1
        self.Cda, self.Cdm = float(cda), float(cdm)
        self.dt = float(dt)
        self.sealevel = float(sealevel)
        self.oFill = float(ofill)
        self.maxarea = np.array([float(self.dm.mesh.maxarea)])
        self.verbose = verbose
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        i8 = dict(dtype=torch.int8, device=dev)
        vec = lambda v: torch.as_tensor(np.broadcast_to(np.asarray(v, np.float64), (N,)).copy()).to(dev)
        self.rainVal = torch.clamp(vec(rain), min=0.0)
        self.upsub = None if upsub is None else vec(upsub)
        self.sedfacVal = None if sedfac is None else vec(sedfac)
        self._kfac = self._make_kfac()
        self.hLocal = torch.as_tensor(np.asarray(h0, np.float64)).to(dev).contiguous()
        self.cumED = torch.zeros(N, **f64)
        self.EbLocal = torch.zeros(N, **f64)
        self.Eb = torch.zeros(N, **f64)
        self.FAL = torch.zeros(N, **f64)
        self.fillFAL = torch.zeros(N, **f64)
        self.bL = self.rainVal * self.dm.area
        self.lFill = torch.zeros(N, **f64)
        self.waterFilled = torch.zeros(N, **f64)
        self.pitIDs = torch.full((N,), -1, **i32)
        self.flatDirs = torch.full((N,), -1, **i32)
        self.ptdir = torch.full((N,), -1, **i32)
        self.lspill = torch.zeros(N, **i8)
        K_ = self.flowDir
        # two receiver sets: the step's main receivers (goSPL's rcvIDi/wghtVali/
        # distRcvi/lsinki snapshots, written only by the first direction build)
        # and the cascade's rebuilds; rcvID/distRcv/wghtVal/lsink name the set
        # built last, as goSPL's live arrays do
        self._rcv_sets = [(torch.zeros((K_, N), **i32), torch.zeros((K_, N), **f64),
                           torch.zeros((K_, N), **f64), torch.zeros(N, **i8)) for _ in range(2)]
        self.rcvIDi, self.distRcvi, self.wghtVali, self.lsinki = self._rcv_sets[0]
        self.rcvID, self.distRcv, self.wghtVal, self.lsink = self._rcv_sets[0]
        self._b = torch.zeros(N, **f64)
        self._q = torch.zeros(N, **f64)
        self._nfa = torch.zeros(N, **f64)
        cap = N + 1
        self._cap = cap
        self._spill = torch.zeros(cap, **i32)
        self._pitparams = torch.zeros((3, cap), **f64)
        self._flvl = torch.zeros((5, cap), **f64)
        self._fvol = torch.zeros((5, cap), **f64)
        self._inv = torch.zeros(cap, **f64)
        self._rhs = torch.zeros(N, **f64)
        # a cascade pass's host decisions go to the device in one transfer,
        # staged through a pinned buffer (grown on demand)
        self._stage_h = None
        self._stage_d = None
        self._tables_h = None
        self.tStart = float(tstart)
        self.tNow = float(tstart)
        self.tEnd = float(tend) if tend is not None else float(tstart)
        self._cascade_max_steps = 100
        self._cascade_patience = 3
        self._cascade_rel_improve = 1.0e-3
        self._cascade_resid = 0.0
        self.cascade_passes = 0
        self.fa_levels = 0
        self.spl_levels = 0
        self.timers = {}
        self.profile = False
        self._events = []
        if profile:
            self.enable_profiling(True)
        self.refine_dag = bool(refine_dag)
        self.refine = bool(refine) or self.refine_dag
        self.refine_rtol = float(refine_rtol)
        self.refine_maxpass = int(refine_maxpass)
        # warm starts of the refined flow solves (main solve, cascade solves)
        self._ws_main = torch.zeros(N, **f64) if self.refine_dag else None
        self._ws_casc = torch.zeros(N, **f64) if self.refine_dag else None
        # per solve: (entry point, correction passes (-1: exact fallback), relative residual)
        self.solve_log = collections.deque(maxlen=4096)
        self.refine_stats = False  # also log the hillslope passes (one read-back per call)
        self._sched_state = {}
        self.native_flow = True
        self._fbuf = None
        self.vSedLocal = torch.zeros(N, **f64)
        self._init_sea(seadepo=seadepo, clinSlp=clinSlp, offshore=offshore, Gmar=Gmar)
        self._init_sed(nodep=nodep, nlPitVolume=nlPitVolume, nlPitDepth=nlPitDepth, nlPitK=nlPitK,
                       nonlinKm=nonlinKm, pitInletBias=pitInletBias, picardSub=picardSub, picardIts=picardIts,
                       marineSolver=marineSolver)
        self._init_output(outputDir=outputDir, makedir=makedir, tout=tout, rstep=rstep, compression=compression,
                          finput=finput)

    @staticmethod
    def _from_file(filename, device):
        """(mesh, h0, parameters) of goSPL's YAML + npz input
        (gospl_amd.inputs.read_input; the mesh in device order)."""
        from . import _lib as lib_, inputs, mesh as mesh_

        inp = inputs.read_input(filename)
        lib = lib_.load()
        bverts = mesh_.boundary_vertices(inp["cells"])
        flat = len(bverts) > 0
        outlet = inputs.border_outlets(inp["coords"], inp["cells"], inp["bc"], bverts) if flat else None
        m = mesh_.build_mesh(inp["coords"], inp["cells"], lib, flat=flat, outlet_mask=outlet)
        dev = lambda a: m.to_device_order(a) if isinstance(a, np.ndarray) else a
        sched = dict(climate=[(t, dev(v)) for t, v in inp["climate"]],
                     tectonics=[(t0, t1, dev(u) if u is not None else None) for t0, t1, u in inp["tectonics"]],
                     sedfactor=[(t, dev(v)) for t, v in inp["sedfactor"]])
        first = lambda evs, t, dflt: next((v for s, v in reversed(evs) if s <= t), dflt)
        params = dict(K=inp["K"], m=inp["m"], d=inp["d"], n=inp["n"], G=inp["G"], flowdir=inp["flowdir"], cda=inp["cda"],
                      cdm=inp["cdm"],
                      dt=inp["dt"], rain=first(sched["climate"], inp["tstart"], 0.0),
                      sedfac=first(sched["sedfactor"], inp["tstart"], None), sealevel=inp["sealevel"],
                      ofill=inp["ofill"], tstart=inp["tstart"], tend=inp["tend"], schedule=sched,
                      outputDir=inp["outputDir"], makedir=inp["makedir"], tout=inp["tout"], rstep=inp["rstep"],
                      compression=inp["compression"], finput=os.path.abspath(filename), nodep=inp["nodep"],
                      seadepo=inp["seadepo"], **inp["sed"])
        return (m, m.to_device_order(inp["elev"])), params

    def _apply_schedule(self):
        """The forcing events active at tNow (goSPL's _updateRain, the
        tectonics event table and _updateEroFactor, unstructuredmesh.py:
        743-794, 903-972, 1063-1094): the latest climate / erodibility event
        started, the tectonic event with start <= tNow < end."""
        sch = self._schedule
        if not sch:
            return
        N = self.dm.N
        dev = self.device
        vec = lambda v: torch.as_tensor(np.broadcast_to(np.asarray(v, np.float64), (N,)).copy()).to(dev)
        ev = [v for s, v in sch.get("climate", []) if s <= self.tNow]
        key = ("rain", len(ev))
        if ev and self._sched_state.get("rain") != key:
            self.rainVal = torch.clamp(vec(ev[-1]), min=0.0)
            self._sched_state["rain"] = key
            self._kfac = self._make_kfac()
        ev = [v for s, v in sch.get("sedfactor", []) if s <= self.tNow]
        key = ("sed", len(ev))
        if ev and self._sched_state.get("sed") != key:
            self.sedfacVal = vec(ev[-1])
            self._sched_state["sed"] = key
            self._kfac = self._make_kfac()
        tec = [u for s, e, u in sch.get("tectonics", []) if s <= self.tNow < e]
        self.upsub = vec(tec[-1]) if tec and tec[-1] is not None else None

    # -------------------------------------------------------------- helpers
    @property
    def FAG(self):
        """Global discharge (on one rank the local vector, flowplex.py:906-933)."""
        return self.FAL

    @property
    def seaID(self):
        """Vertices whose filled elevation is at or below sea level, as the
        last flow-direction build defines them (flowplex.py:493)."""
        return torch.nonzero(self.lFill <= self.sealevel).flatten()

    @property
    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _call(self, name, *args):
        if self.profile:
            s = torch.cuda.current_stream(self.device)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            _lib.check(getattr(self.lib, name)(*args), name)
            e1.record(s)
            self._events.append((name, e0, e1))
        else:
            _lib.check(getattr(self.lib, name)(*args), name)

    def enable_profiling(self, on=True):
        """Record HIP events (on the launch stream) around every library call."""
        self.profile = on
        self._events = []

    def phase_times(self, reset=True):
        """Milliseconds per library entry point since the last reset."""
        torch.cuda.synchronize(self.device)
        out = {}
        for name, e0, e1 in self._events:
            d = out.setdefault(name, [0.0, 0])
            d[0] += e0.elapsed_time(e1)
            d[1] += 1
        if reset:
            self._events = []
        return out

    def _host_tables(self, n):
        """Pinned host buffer of at least n doubles for the per-pit tables
        (read only after the call that fills it has synchronised)."""
        if self._tables_h is None or self._tables_h.numel() < n:
            self._tables_h = torch.empty(max(4096, 2 * n), dtype=torch.float64, pin_memory=True)
        return self._tables_h

    def _staging(self, n):
        """A numpy view of the pinned staging buffer, at least n doubles.  The
        previous pass's transfer out of it has completed: every cascade pass
        synchronises the stream (its inflow read-back) before writing."""
        if self._stage_h is None or self._stage_h.numel() < n:
            size = max(4096, 2 * n)
            self._stage_h = torch.empty(size, dtype=torch.float64, pin_memory=True)
            self._stage_d = torch.empty(size, dtype=torch.float64, device=self.device)
        return self._stage_h.numpy()

    def _make_kfac(self):
        if self.sedfacVal is None and self.coeffd == 0.0:
            return None
        k = torch.ones_like(self.rainVal)
        if self.sedfacVal is not None:
            k = k * torch.clamp(self.sedfacVal, min=0.1)
        return k * self.rainVal ** self.coeffd

    def hillslope_sweeps(self):
        """Jacobi sweeps that bring the error below 2^-60 of max|h| by the
        contraction bound rho = s/(1+s), s = Kd*dt*gmax; 0 (the library then
        iterates to a device-checked convergence) when that takes more than
        64 sweeps."""
        return sweeps_for(max(self.Cda, self.Cdm) * self.dt, self.dm.gmax, 60)

    def hillslope_sweeps32(self):
        """FP32 Jacobi sweeps per correction of the refined hillslope solve:
        the contraction bound brought below 2^-27 (FP32 rounding of the
        correction); 0 (the converged FP64 solve) past 64 sweeps."""
        return sweeps_for(max(self.Cda, self.Cdm) * self.dt, self.dm.gmax, 27)

    # -------------------------------------------------------------- pits
    def fillElevation(self, sed=False):
        """pitfilling.py:849-903: water filling (sed=False) or sediment
        filling (sed=True: submerged cells keep their filled values, pits are
        labelled above the fill level, volumes are not clipped at sea
        level)."""
        S = self.stream
        npits = ctypes.c_int64(0)
        if sed:
            # the fill level read back: the pits are labelled above it
            minh = float(torch.min(self.hLocal))
            if not self.flatModel:
                minh += 1.0e-3
            level = max(minh, self.sealevel + self.oFill)
            self._call("gospl_fill_depressions", self.H, _ptr(self.hLocal), level, -math.inf, _ptr(self.lFill), S)
            self._call("gospl_pit_info_sed", self.H, _ptr(self.hLocal), _ptr(self.lFill), level, self.sealevel,
                       _ptr(self.pitIDs), _ptr(self.flatDirs), _ptr(self.ptdir), _ptr(self.lspill),
                       self._cap, _ptr(self._spill), _ptr(self._pitparams), _ptr(self._flvl),
                       _ptr(self._fvol), ctypes.byref(npits), S)
        else:
            # level = max(min h (+1e-3 on the sphere), sealevel + oFill), chosen
            # on the device (no read-back)
            self._call("gospl_fill_elevation", self.H, _ptr(self.hLocal), 0 if self.flatModel else 1,
                       self.sealevel + self.oFill, self.sealevel, _ptr(self.lFill), S)
            self._call("gospl_pit_info", self.H, _ptr(self.hLocal), _ptr(self.lFill), self.sealevel,
                       _ptr(self.pitIDs), _ptr(self.flatDirs), _ptr(self.ptdir), _ptr(self.lspill),
                       self._cap, _ptr(self._spill), _ptr(self._pitparams), _ptr(self._flvl),
                       _ptr(self._fvol), ctypes.byref(npits), S)
        n1 = int(npits.value) + 1
        self.npits = n1 - 1
        # the per-pit tables in one transfer (spill ids are exact in f64)
        host = self._host_tables(14 * n1)
        self._call("gospl_pit_tables", self.H, n1 - 1, self._cap, _ptr(self._spill), _ptr(self._pitparams),
                   _ptr(self._flvl), _ptr(self._fvol), host.data_ptr(), S)
        packed = host.numpy()[:14 * n1]
        spill = packed[:n1].astype(np.int64)
        self.pitParams = packed[n1:4 * n1].reshape(3, n1).T.copy()
        self.filled_lvl = packed[4 * n1:9 * n1].reshape(5, n1).T.copy()
        self.filled_vol = packed[9 * n1:14 * n1].reshape(5, n1).T.copy()
        self.pitInfo = np.zeros((n1, 2), dtype=np.int64)
        self.pitInfo[:, 0] = spill
        self.pitInfo[:, 1] = np.where(spill >= 0, 0, -1)

    # -------------------------------------------------------------- flow
    def _buildFlowDirection(self, h, down=True):
        """flowplex.py:480-544.  The step's first build (down=False) fills the
        main set, later builds (the cascade's) the other one."""
        self.rcvID, self.distRcv, self.wghtVal, self.lsink = self._rcv_sets[1 if down else 0]
        self._call("gospl_flow_direction", self.H, self.flowDir, self.flowExp, _ptr(h), self.sealevel,
                   _ptr(self.lFill), 1 if down else 0, _ptr(self.pitIDs), _ptr(self.flatDirs),
                   _ptr(self.rcvID), _ptr(self.distRcv), _ptr(self.wghtVal), _ptr(self.lsink),
                   self.stream)
        self._snapshot_sea()

    def _solve_KSP(self, guess, matrix, vector1, vector2, fatal=False, seed=False):
        """goSPL's flow solve (flowplex.py:323-427) with its signature:
        vector1 the right-hand side, vector2 the solution (and, with guess,
        the initial guess of the refined solve; seed: a zero vector2 starts
        from vector1).  The operator is the one built from the last flow
        directions (rcvID, wghtVal); `matrix` is accepted for the signature
        and must be None.  A receiver cycle that does not drain raises when
        fatal, else zeroes vector2, as goSPL's fallback does."""
        if matrix is not None:
            raise ValueError("_solve_KSP builds its operator from the flow directions: pass matrix=None")
        ws = None
        if self.refine_dag:
            ws = vector2.clone() if guess else torch.zeros_like(vector2)
            if seed and not bool(ws.any()):
                ws.copy_(vector1)
        self._solve_flow(vector1, vector2, fatal=fatal, ws=ws)
        return vector2

    def _solve_flow(self, b, out, fatal=True, ws=None):
        """_solve_KSP (flowplex.py:323-427) on the flow matrix.  A receiver
        cycle that does not drain makes goSPL's solvers give up: the main
        solve (fatal) raises, a cascade solve is zeroed (flowplex.py:304-320).
        Refined mode: ``ws`` holds the warm start and receives the solution."""
        lev = ctypes.c_int32(0)
        bad = ctypes.c_int64(0)
        if self.refine_dag:
            npass = ctypes.c_int32(0)
            rr = ctypes.c_double(0.0)
            out.copy_(ws)
            self._call("gospl_flowacc_refine", self.H, self.flowDir, _ptr(self.rcvID), _ptr(self.wghtVal),
                       _ptr(b), _ptr(out), self.refine_rtol, self.refine_maxpass, ctypes.byref(npass),
                       ctypes.byref(rr), ctypes.byref(bad), self.stream)
            self.solve_log.append(("flowacc", int(npass.value), float(rr.value)))
            if not bad.value:
                ws.copy_(out)
        else:
            self._call("gospl_flowacc", self.H, self.flowDir, _ptr(self.rcvID), _ptr(self.wghtVal), _ptr(b),
                       _ptr(out), ctypes.byref(lev), ctypes.byref(bad), self.stream)
            Model.sweeps["flow"] += 1
        self.undrained = int(bad.value)
        if bad.value:
            if fatal:
                raise RuntimeError("flow accumulation did not converge (%d undrained vertices)" % bad.value)
            out.zero_()
        return int(lev.value)

    def _distributeDownstream(self, pitVol, FA, hl, step):
        excess = False
        npit = len(self.pitParams)
        S = self.stream
        self._call("gospl_pit_inflow", self.H, _ptr(FA), _ptr(self.lsink), _ptr(self.pitIDs), npit,
                   _ptr(self._inv), S)
        inV = self._inv[:npit].cpu().numpy()
        eV = inV - pitVol
        full = np.zeros(npit)
        level = np.full(npit, np.nan)
        u = vv = np.zeros(0)
        nfa_sum = 0.0
        if (eV > 0.0).any():
            eIDs = eV > 0.0
            pitVol[eIDs] = 0.0
            full[eIDs] = 1.0
            spillIDs = self.pitInfo[eIDs, 0]
            localSpill = np.where(self.pitInfo[eIDs, 1] == 0)[0]
            localPts = spillIDs[localSpill]
            vals = eV[eIDs][localSpill]
            if len(localPts):
                # numpy fancy assignment: a repeated index keeps the last value
                u, last = np.unique(localPts[::-1], return_index=True)
                vv = vals[::-1][last]
                nfa_sum = float(np.sum(vv / self.dt))
        if (eV < 0.0).any():
            eIDs = np.where((eV < 0.0) & (inV > 0.0))[0]
            pitVol[eIDs] += eV[eIDs]
            if len(eIDs):
                nid = np.absolute(self.filled_vol[eIDs] - pitVol[eIDs][:, None]).argmin(axis=1)
                level[eIDs] = self.filled_lvl[eIDs, nid]
        go = bool((eV > 1.0e-3).any())
        if go or full.any() or np.isfinite(level).any():
            # water levels of the pits and the next solve's right-hand side
            # (the spilled volumes / dt), in one transfer and one call
            k = len(u)
            st = self._staging(2 * npit + 2 * k)
            st[:npit] = full
            st[npit:2 * npit] = level
            st[2 * npit:2 * npit + k] = u
            st[2 * npit + k:2 * npit + 2 * k] = vv
            n = 2 * npit + 2 * k
            self._stage_d[:n].copy_(self._stage_h[:n], non_blocking=True)
            self._call("gospl_cascade_seed", self.H, _ptr(self._stage_d), npit, k, _ptr(self.pitIDs),
                       _ptr(self.lFill), _ptr(self.waterFilled), self.dt, _ptr(self._rhs), S)
        nFA = None
        if go:
            if step == 100:
                self._buildFlowDirection(self.lFill)
            else:
                self._buildFlowDirection(self.waterFilled)
            # sum of the right-hand side: its nonzero entries are the spill
            # values above (host side, no device synchronisation)
            resid = nfa_sum
            self._cascade_resid = resid
            if resid > self.maxarea[0]:
                excess = True
                self._solve_flow(self._rhs, self._q, fatal=False, ws=self._ws_casc)
                self._call("gospl_cascade_add", self.H, _ptr(self._q), _ptr(hl), _ptr(self.waterFilled),
                           self.dt, _ptr(self._nfa), _ptr(self.FAL), S)
                nFA = self._nfa
        return excess, pitVol, nFA

    def flowAccumulation(self):
        """flowplex.py:752-945.  The exact-sweep modes run it as one native
        call (``gospl_flow_accumulation``: the same entry points and the
        cascade's decisions in C++); ``native_flow = False``, and the
        refined DAG mode (warm-started flow solves), run the Python driver
        below over the same entry points."""
        if self.native_flow and not self.refine_dag:
            return self._flowAccumulation_native()
        return self._flowAccumulation_host()

    def _flow_buffers(self):
        if self._fbuf is None:
            N = self.lpoints
            f64 = dict(dtype=torch.float64, device=self.device)
            self._water = torch.zeros(N, **f64)
            self._hl = torch.zeros(N, **f64)
            self._fa = torch.zeros(N, **f64)
            r0, r1 = self._rcv_sets
            fb = _lib.FlowBuffers()
            for name, t in (("h", self.hLocal), ("rain", self.rainVal), ("lfill", self.lFill),
                            ("pitids", self.pitIDs), ("flatdirs", self.flatDirs), ("ptdir", self.ptdir),
                            ("lspill", self.lspill), ("rcv0", r0[0]), ("dist0", r0[1]), ("wgt0", r0[2]),
                            ("lsink0", r0[3]), ("rcv1", r1[0]), ("dist1", r1[1]), ("wgt1", r1[2]),
                            ("lsink1", r1[3]), ("b", self._b), ("fal", self.FAL), ("fillfal", self.fillFAL),
                            ("water", self._water), ("hl", self._hl), ("q", self._q), ("nfa", self._nfa),
                            ("rhs", self._rhs), ("fa", self._fa), ("spill", self._spill),
                            ("pitparams", self._pitparams), ("flvl", self._flvl), ("fvol", self._fvol),
                            ("inv", self._inv)):
                setattr(fb, name, t.data_ptr())
            fb.cap = self._cap
            self._fbuf = fb
            self._fbuf_keep = (self.hLocal, self.rainVal)
        # the state tensors may have been replaced since (hLocal, rainVal)
        self._fbuf.h = self.hLocal.data_ptr()
        self._fbuf.rain = self.rainVal.data_ptr()
        self._fbuf_keep = (self.hLocal, self.rainVal)
        return self._fbuf

    def _flowAccumulation_native(self):
        t0 = time.perf_counter()
        fb = self._flow_buffers()
        fp = _lib.FlowParams(nrcv=self.flowDir, flowexp=self.flowExp, sealevel=self.sealevel,
                             floor_level=self.sealevel + self.oFill, add_eps=0 if self.flatModel else 1,
                             dt=self.dt, maxarea=float(self.maxarea[0]), max_passes=self._cascade_max_steps,
                             patience=self._cascade_patience, rel_improve=self._cascade_rel_improve)
        res = _lib.FlowResult()
        self._call("gospl_flow_accumulation", self.H, ctypes.byref(fp), ctypes.byref(fb), ctypes.byref(res),
                   self.stream)
        Model.sweeps["flow"] += int(res.flow_solves)
        n1 = int(res.npits) + 1
        self.npits = n1 - 1
        ptr = ctypes.c_void_p()
        rows = ctypes.c_int64(0)
        self._call("gospl_flow_tables", self.H, ctypes.byref(ptr), ctypes.byref(rows))
        packed = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double)), shape=(14 * n1,))
        spill = packed[:n1].astype(np.int64)
        self.pitParams = packed[n1:4 * n1].reshape(3, n1).T.copy()
        self.filled_lvl = packed[4 * n1:9 * n1].reshape(5, n1).T.copy()
        self.filled_vol = packed[9 * n1:14 * n1].reshape(5, n1).T.copy()
        self.pitInfo = np.zeros((n1, 2), dtype=np.int64)
        self.pitInfo[:, 0] = spill
        self.pitInfo[:, 1] = np.where(spill >= 0, 0, -1)
        self.rcvID, self.distRcv, self.wghtVal, self.lsink = self._rcv_sets[int(res.last_set)]
        self.cascade_passes = int(res.cascade_passes)
        self.undrained = int(res.undrained)
        self.fa_levels = 0
        self.waterFilled = self._water
        self.timers["flow"] = time.perf_counter() - t0

    def _flowAccumulation_host(self):
        t0 = time.perf_counter()
        self.fillElevation(sed=False)
        pitVol = self.pitParams[:, 0].copy()
        hl = self.hLocal.clone()
        self._buildFlowDirection(hl, False)  # the main set: rcvIDi/wghtVali/distRcvi/lsinki
        S = self.stream
        self._call("gospl_rain_source", self.H, _ptr(self.rainVal), _ptr(self.lFill), self.sealevel,
                   _ptr(self._b), S)
        self.fa_levels = self._solve_flow(self._b, self.FAL, ws=self._ws_main)
        self.waterFilled = hl.clone()
        self.cascade_passes = 0
        has_pits = bool((pitVol > 0.0).any())
        if has_pits:
            FA = self.FAL * self.dt
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
        self._call("gospl_flow_finalize", self.H, _ptr(hl), _ptr(self.waterFilled), _ptr(self.pitIDs),
                   1 if has_pits else 0, _ptr(self.FAL), _ptr(self.fillFAL), S)
        self.waterFilled = self.waterFilled - hl
        self.timers["flow"] = time.perf_counter() - t0

    # -------------------------------------------------------------- erosion
    def erodepSPLnl(self):
        """nlSPL.erodepSPLnl (gospl/eroder/nlSPL.py:321-456): Newton steps of
        exact erosion-direction sweeps (G = 0) or of GMRES solves of the
        coupled step (G > 0, _solveNL_ed) to max|step| <= 1e-13 max|h|
        (goSPL's SNES stops at 1e-6)."""
        it = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        if self.fDepa > 0.0:
            inner = ctypes.c_int32(0)
            self._call("gospl_spl_nl_dep", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali),
                       _ptr(self.distRcvi), _ptr(self.FAL), _ptr(self._kfac), self.K, self.spl_m, self.spl_n,
                       self.fDepa, self.dt, self.sealevel, _ptr(self.lFill), _ptr(self.lsink), _ptr(self.hLocal),
                       _ptr(self.cumED), _ptr(self.Eb), _ptr(self.EbLocal), self.nl_tol, self.nl_maxit,
                       ctypes.byref(it), ctypes.byref(inner), ctypes.byref(rel), self.stream)
            self.nl_iters = int(it.value)
            self.ed_iters = int(inner.value)
            Model.sweeps["spl"] += 2 * self.ed_iters
            return
        self._call("gospl_spl_nl", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali),
                   _ptr(self.distRcvi), _ptr(self.FAL), _ptr(self._kfac), self.K, self.spl_m, self.spl_n, self.dt,
                   self.sealevel, _ptr(self.lFill), _ptr(self.lsink), _ptr(self.hLocal), _ptr(self.cumED),
                   _ptr(self.Eb), _ptr(self.EbLocal), self.nl_tol, self.nl_maxit, ctypes.byref(it),
                   ctypes.byref(rel), self.stream)
        Model.sweeps["spl"] += int(it.value)
        self.nl_iters = int(it.value)

    def _erodepSPL_ed(self):
        """SPL._getEroDepRate with G > 0 (gospl/eroder/SPL.py:306-333,
        116-258): the coupled elevation / sediment-flux solve; goSPL warns
        and keeps the iterate when its solve does not converge, as here."""
        it = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        self._call("gospl_spl_dep", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali),
                   _ptr(self.distRcvi), _ptr(self.FAL), _ptr(self._kfac), self.K, self.spl_m, self.fDepa, self.dt,
                   self.sealevel, _ptr(self.lFill), _ptr(self.lsink), _ptr(self.hLocal), _ptr(self.cumED),
                   _ptr(self.Eb), _ptr(self.EbLocal), self.ed_rtol, self.ed_maxit, ctypes.byref(it),
                   ctypes.byref(rel), None, self.stream)
        self.ed_iters, self.ed_relres = int(it.value), float(rel.value)
        Model.sweeps["spl"] += 2 * self.ed_iters
        if not self.ed_relres <= self.ed_rtol:
            warnings.warn("Transport-limited SPL solve did not converge after %d iterations (relative residual "
                          "%.3e)" % (self.ed_iters, self.ed_relres))

    def erodepSPL(self):
        if self.spl_n != 1.0:
            # goSPL's runProcesses takes the non-linear solver for n != 1 (model.py:396-401)
            return self.erodepSPLnl()
        if self.fDepa > 0.0:
            return self._erodepSPL_ed()
        lev = ctypes.c_int32(0)
        if self.refine_dag:
            npass = ctypes.c_int32(0)
            rr = ctypes.c_double(0.0)
            self._call("gospl_spl_refine", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali),
                       _ptr(self.distRcvi), _ptr(self.FAL), _ptr(self._kfac), self.K, self.spl_m, self.dt,
                       self.sealevel, _ptr(self.lFill), _ptr(self.lsink), _ptr(self.hLocal), _ptr(self.cumED),
                       _ptr(self.Eb), _ptr(self.EbLocal), self.refine_rtol, self.refine_maxpass,
                       ctypes.byref(npass), ctypes.byref(rr), self.stream)
            self.solve_log.append(("spl", int(npass.value), float(rr.value)))
            self.spl_levels = 0
            return
        Model.sweeps["spl"] += 1
        self._call("gospl_spl", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali),
                   _ptr(self.distRcvi), _ptr(self.FAL), _ptr(self._kfac), self.K, self.spl_m, self.dt,
                   self.sealevel, _ptr(self.lFill), _ptr(self.lsink), _ptr(self.hLocal), _ptr(self.cumED),
                   _ptr(self.Eb), _ptr(self.EbLocal), ctypes.byref(lev), self.stream)
        self.spl_levels = int(lev.value)

    # -------------------------------------------------------------- hillslope
    def getHillslope(self):
        if self.outputDir is not None:
            # goSPL's seaID as the outputs read it (hillslope.py:620)
            self._seaOut = self.hLocal <= self.sealevel
        if self.Cda == 0.0 and self.Cdm == 0.0:
            # goSPL then adds the erosion step's thickness again (hillslope.py:638-643)
            self.EbLocal += self.Eb * self.dt / self.dt
            return
        if self.refine:
            if self.refine_stats:
                npass = ctypes.c_int32(0)
                rr = ctypes.c_double(0.0)
                pp, rp = ctypes.byref(npass), ctypes.byref(rr)
            else:  # no read-back, no synchronisation
                pp = rp = None
            self._call("gospl_hillslope_refine", self.H, self.Cda, self.Cdm, self.dt, self.sealevel,
                       self.hillslope_sweeps32(), self.refine_rtol, self.refine_maxpass, _ptr(self.hLocal),
                       _ptr(self.cumED), _ptr(self.EbLocal), pp, rp, self.stream)
            if self.refine_stats:
                self.solve_log.append(("hillslope", int(npass.value), float(rr.value)))
            return
        self._call("gospl_hillslope", self.H, self.Cda, self.Cdm, self.dt, self.sealevel,
                   self.hillslope_sweeps(), _ptr(self.hLocal), _ptr(self.cumED), _ptr(self.EbLocal),
                   self.stream)

    # -------------------------------------------------------------- driver
    def applyForces(self):
        self._apply_schedule()
        self.bL = self.rainVal * self.dm.area
        if self.upsub is not None:
            self.hLocal += self.upsub * self.dt

    def initialise(self):
        """model.py:284-295: forcing, a restart's state read back (readData),
        then the init-time flow accumulation."""
        self.applyForces()
        if self.rStep > 0:
            self.readData()
        self.flowAccumulation()

    def step(self):
        """Body of the runProcesses loop (model.py:379-453) for this path:
        flow, erosion, (with nodep False) flow again and continental
        sediment routing, hillslope."""
        self.flowAccumulation()
        self.erodepSPL()
        if not self.nodep:
            self.flowAccumulation()
            self._snapshot_sea()
            self.sedChange()
            if self.seaDepo:
                self.seaChange()
        self.getHillslope()
        if self.tNow < self.tEnd:
            self.applyForces()
        self.tNow += self.dt

    def runProcesses(self):
        """model.py:347-453: an output when one is due (visModel), the step."""
        while self.tNow <= self.tEnd:
            self.visModel()
            if self.tNow == self.tEnd:
                break
            self.step()

    def close(self):
        """Release the library handle now (its device and pinned buffers and
        events) instead of at interpreter exit, when the HIP runtime may be
        shutting down already."""
        torch.cuda.synchronize(self.device)
        self.dm.close()

    # -------------------------------------------------------------- views
    def input_order(self, t):
        """Host copy of a per-vertex tensor in input-mesh order."""
        a = t.detach().cpu().numpy()
        return self.dm.mesh.to_input_order(a)
