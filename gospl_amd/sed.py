"""Continental sediment routing and deposition: goSPL's ``SEDMesh``
(gospl/sed/sedplex.py) on the GPU, mixed into ``Model``.

It runs when ``nodep`` is False, after the second flowAccumulation of the
step (gospl/model.py:411-416):

* ``_getSedFlux`` (sedplex.py:103-190): the eroded volume rate -Eb * area
  routed through the step's flow operator (the main receivers of the last
  flowAccumulation, goSPL's ``fMati``): ``vSedLocal`` in m3/yr;
* ``fillElevation(sed=True)`` (pitfilling.py:849-903): depressions of the
  eroded surface, labelled above the fill level, volumes not clipped at sea
  level;
* ``_distributeSediment`` / ``_moveDownstream`` (sedplex.py:199-523): the
  sediment cascade -- each pass adds the volume reaching every depression,
  fills the depressions it overfills (``sedFilled = lFill`` there), sends the
  excess from their spill points down the receivers of the sediment-filled
  surface, and adds what passes to ``vSedLocal``; flow-terminal cells outside
  every depression drop their load (deposited in place on a closed domain);
* ``_updateSinks`` (sedplex.py:982-1059): the deposit in each depression --
  full or small/shallow ones filled bottom-up to the level whose depth-volume
  curve matches the deposited volume (``_bottomUpDelta``) and tilted by 1e-6
  towards the spill point (``_addPitMicroTilt``); large and deep partly
  filled ones seeded along the water depth and at their inlets, spread by
  non-linear diffusion and rescaled to the deposited volume
  (``_diffuseLargePit``, ``_diffuseImplicit``).

Every per-vertex step is a library call (``gospl_sed_*``, the flow sweep,
the pit information, ``gospl_nl_diffusion``); the per-depression decisions
(a few numbers per depression) are taken on the host from the pit tables, as
goSPL takes them.
"""

from __future__ import annotations

import ctypes
import math
import warnings

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _interp_levels(depo, pits, pitParams, filled_vol, filled_lvl):
    """Lake level per pit matching a deposited volume on its depth-volume
    curve (np.interp on [0, filled_vol] -> [bottom, filled_lvl],
    sedplex.py:538-545); -inf for the other pits."""
    target = np.full(len(pitParams), -np.inf)
    for p in pits:
        pit_min = pitParams[p, 1] - pitParams[p, 2]
        vols = np.concatenate(([0.0], filled_vol[p]))
        lvls = np.concatenate(([pit_min], filled_lvl[p]))
        target[p] = np.interp(depo[p], vols, lvls)
    return target


class SEDMesh:
    """sedChange and its helpers for ``Model`` (goSPL's method names)."""

    def _init_sed(self, nodep=True, nlPitVolume=1.0e9, nlPitDepth=100.0, nlPitK=None, nonlinKm=10.0,
                  pitInletBias=0.10, picardSub=10, picardIts=2, marineSolver="ts"):
        """SEDMesh.__init__ (sedplex.py:23-98) and the diffusion keys of
        goSPL's input (inputparser.py:546-597)."""
        self.nodep = bool(nodep)
        self.nlK = float(nonlinKm)
        self.nl_pit_volume = float(nlPitVolume)
        self.nl_pit_depth = float(nlPitDepth)
        self.nl_pit_K = float(nlPitK) if nlPitK is not None else self.nlK
        self.nl_pit_inlet_bias = min(1.0, max(0.0, float(pitInletBias)))
        self.picardSub = max(1, int(picardSub))
        self.picardIts = max(1, int(picardIts))
        ms = str(marineSolver).lower()
        self.marineSolver = ms if ms in ("ts", "picard") else "ts"
        # hillSLP.__init__ (hillslope.py:32-35)
        self.dexp = 0.05
        self.minDiff = 1.0e-4
        self._domainHasOutlet = bool(self.flatModel and int((self.dm.outlet > 0).sum()) > 0)
        self._sedbuf = None
        self.nl_sweeps = 0
        self._warned_ts = False

    def _sed_buffers(self):
        if self._sedbuf is None:
            N = self.lpoints
            f64 = dict(dtype=torch.float64, device=self.device)
            self._sedbuf = dict(b=torch.zeros(N, **f64), q=torch.zeros(N, **f64), vsed=torch.zeros(N, **f64),
                                closed=torch.zeros(N, **f64), delta=torch.zeros(N, **f64),
                                delta2=torch.zeros(N, **f64), wd=torch.zeros(N, **f64),
                                sel=torch.zeros(N, dtype=torch.int8, device=self.device),
                                inlet=torch.zeros(N, dtype=torch.int8, device=self.device),
                                xyz=torch.as_tensor(np.ascontiguousarray(self.dm.mesh.coords, np.float64)).to(
                                    self.device))
        return self._sedbuf

    def _dev(self, a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a)).to(device=self.device, dtype=dtype)

    # ------------------------------------------------------------ fluxes
    def _getSedFlux(self):
        """vSed = (I - W^T)^-1 (-Eb * area) on the main receivers."""
        sb = self._sed_buffers()
        S = self.stream
        self._call("gospl_sed_source", self.H, _ptr(self.Eb), _ptr(sb["b"]), S)
        lev = ctypes.c_int32(0)
        bad = ctypes.c_int64(0)
        self._call("gospl_flowacc", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali), _ptr(sb["b"]),
                   _ptr(self.vSedLocal), ctypes.byref(lev), ctypes.byref(bad), S)
        if bad.value:
            # _solve_KSP(fatal=False) leaves a zero solution (flowplex.py:304-320)
            self.vSedLocal.zero_()

    # ------------------------------------------------------------ cascade
    def _moveDownstream(self, vsed, step):
        """One pass of the sediment cascade (sedplex.py:199-374).  vsed holds
        the volumes to distribute; returns whether a routed excess is left
        (then in the q buffer)."""
        sb = self._sed_buffers()
        S = self.stream
        excess = False
        changed = False
        self._call("gospl_sed_stray", self.H, _ptr(self.lsink), _ptr(self.pitIDs),
                   0 if self._domainHasOutlet else 1, _ptr(vsed), _ptr(sb["closed"]), S)
        npit = len(self.pitParams)
        self._call("gospl_pit_inflow", self.H, _ptr(vsed), _ptr(self.lsink), _ptr(self.pitIDs), npit,
                   _ptr(self._inv), S)
        inV = self._inv[:npit].cpu().numpy()
        eV = inV - self.pitVol
        full = np.zeros(npit)
        u = vv = np.zeros(0)
        if (eV > 0.0).any():
            eIDs = eV > 0.0
            spillIDs = self.pitInfo[eIDs, 0]
            localSpill = np.where(self.pitInfo[eIDs, 1] == 0)[0]
            localPts = spillIDs[localSpill]
            vals = eV[eIDs][localSpill]
            if len(localPts):
                # numpy fancy assignment: a repeated index keeps the last value
                u, last = np.unique(localPts[::-1], return_index=True)
                vv = vals[::-1][last]
            self.pitVol[eIDs] = 0.0
            full[eIDs] = 1.0
            changed = True
        eIDs = eV < 0
        self.pitVol[eIDs] -= inV[eIDs]
        self.pitVol[self.pitVol < 0] = 0.0
        go = bool((eV > 1.0e-3).any())
        if changed or go:
            # sedFilled = lFill on the overfilled pits, and the spilled volumes
            # as the next solve's right-hand side (dt 1: volumes, not rates)
            k = len(u)
            st = self._staging(2 * npit + 2 * k)
            st[:npit] = full
            st[npit:2 * npit] = np.nan
            st[2 * npit:2 * npit + k] = u
            st[2 * npit + k:2 * npit + 2 * k] = vv
            n = 2 * npit + 2 * k
            self._stage_d[:n].copy_(self._stage_h[:n], non_blocking=True)
            self._call("gospl_cascade_seed", self.H, _ptr(self._stage_d), npit, k, _ptr(self.pitIDs),
                       _ptr(self.lFill), _ptr(self.sedFilled), 1.0, _ptr(self._rhs), S)
        if go:
            if changed or step == 0 or step == 100:
                self._buildFlowDirection(self.lFill if step == 100 else self.sedFilled)
            # PETSc's VecSum of the right-hand side: its nonzero entries in
            # ascending vertex order
            total = 0.0
            for x in vv:
                total += float(x)
            if total > 0.5 * self.maxarea[0]:
                excess = True
                # a refined solve starts from the right-hand side (goSPL's seed)
                self._solve_flow(self._rhs, sb["q"], fatal=False, ws=self._rhs.clone() if self.refine_dag else None)
        return excess

    def _distributeSediment(self, hl):
        """The cascade loop (sedplex.py:376-523)."""
        sb = self._sed_buffers()
        S = self.stream
        vsed = sb["vsed"]
        torch.mul(self.vSedLocal, self.dt, out=vsed)
        sb["closed"].zero_()
        step = 0
        excess = True
        while excess:
            if step >= 5000:
                if self.verbose:
                    print("Continental sediment routing did not converge after 5000 iterations; continuing.")
                break
            excess = self._moveDownstream(vsed, step)
            if excess:
                self._call("gospl_sed_add", self.H, _ptr(sb["q"]), _ptr(hl), _ptr(self.sedFilled), self.dt,
                           _ptr(vsed), _ptr(self.vSedLocal), S)
            step += 1
        self.sed_passes = step
        if not self._domainHasOutlet:
            # closed domain: the sediment of stray terminal cells is deposited
            # in place (sedplex.py:475-500)
            self._call("gospl_sed_closed", self.H, _ptr(sb["closed"]), _ptr(self.hLocal), _ptr(self.cumED), S)

    # ------------------------------------------------------------ deposits
    def _bottomUpDelta(self, hl, depo, pit_select):
        sb = self._sed_buffers()
        pits = np.where(pit_select & (depo > 0))[0]
        target = _interp_levels(depo, pits, self.pitParams, self.filled_vol, self.filled_lvl)
        target_d = self._dev(target)  # held until the call is queued: a freed block is reused at once
        self._call("gospl_sed_bottomup", self.H, _ptr(self.pitIDs), _ptr(target_d), _ptr(self.lFill),
                   _ptr(hl), _ptr(sb["delta"]), self.stream)
        return sb["delta"]

    def _spillCoords(self):
        """Coordinates of each pit's spill point (zeros without one)."""
        npit = len(self.pitParams)
        xyz = np.zeros((npit, 3))
        sp = self.pitInfo[:, 0].astype(np.int64)
        ok = (sp >= 0) & (self.pitInfo[:, 1] == 0)
        xyz[ok] = self.dm.mesh.coords[sp[ok]]
        return xyz

    def _addPitMicroTilt(self, delta, pit_select):
        if not pit_select.any():
            return delta
        sb = self._sed_buffers()
        npit = len(self.pitParams)
        mean = torch.zeros(npit, dtype=torch.float64, device=self.device)
        psel = self._dev(pit_select.astype(np.int8), torch.int8)
        spill_xyz = self._dev(self._spillCoords())
        self._call("gospl_sed_tilt", self.H, _ptr(sb["xyz"]), _ptr(self.pitIDs), npit, _ptr(psel), _ptr(spill_xyz),
                   _ptr(delta), _ptr(mean), self.stream)
        return delta

    def _identifyPitInlets(self, psel):
        sb = self._sed_buffers()
        npit = len(self.pitParams)
        count = torch.zeros(npit, dtype=torch.float64, device=self.device)
        self._call("gospl_sed_inlets", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.pitIDs), npit,
                   _ptr(psel), _ptr(sb["inlet"]), _ptr(count), self.stream)
        return sb["inlet"], count

    def _diffuseImplicit(self, dh, mask, Cd_val, out):
        """hillSLP._diffuseImplicit (hillslope.py:829-923) in its
        lagged-diffusivity form (_diffuseImplicitPicard, :925-1076)."""
        if self.marineSolver == "ts" and not self._warned_ts:
            warnings.warn("marineSolver 'ts' (PETSc's adaptive Rosenbrock-W integrator) is served by goSPL's "
                          "Picard scheme (picardSub sub-steps x picardIts updates)", stacklevel=3)
            self._warned_ts = True
        sw = ctypes.c_int64(0)
        self._call("gospl_nl_diffusion", self.H, _ptr(dh), _ptr(mask), float(Cd_val), self.minDiff, self.dexp,
                   self.dt, self.picardSub, self.picardIts, _ptr(self.hLocal), _ptr(out), ctypes.byref(sw),
                   self.stream)
        self.nl_sweeps += int(sw.value)
        return out

    def _diffuseLargePit(self, hl, depo, pit_select):
        sb = self._sed_buffers()
        S = self.stream
        npit = len(self.pitParams)
        active = np.where(pit_select & (depo > 0))[0]
        out = sb["delta2"]
        if active.size == 0:
            out.zero_()
            return out
        ps = np.zeros(npit, np.int8)
        ps[active] = 1
        psel = self._dev(ps, torch.int8)
        inlet, count_d = self._identifyPitInlets(psel)
        count = count_d.cpu().numpy()
        no_inlet = active[count[active] == 0]
        if no_inlet.size:
            # the highest cell of the pit (first in input-mesh order on ties)
            pid = self.pitIDs.cpu().numpy()
            h = hl.cpu().numpy()
            gid = self.dm.mesh.gid
            for p in no_inlet:
                cells = np.where(pid == p)[0]
                if cells.size:
                    cells = cells[np.argsort(gid[cells], kind="stable")]
                    c = int(cells[np.argmax(h[cells])])
                    inlet[c] = 1
                    count[p] = 1.0
            count_d = self._dev(count)
        bias = self.nl_pit_inlet_bias
        target = _interp_levels(depo, active, self.pitParams, self.filled_vol, self.filled_lvl)
        denom_d = torch.zeros(npit, dtype=torch.float64, device=self.device)
        target_d = self._dev(target)
        self._call("gospl_sed_lake_depth", self.H, _ptr(self.pitIDs), npit, _ptr(psel), _ptr(target_d),
                   _ptr(hl), _ptr(sb["sel"]), _ptr(sb["wd"]), _ptr(denom_d), S)
        denom = denom_d.cpu().numpy()
        ca = (1.0 - bias) * depo
        cb = np.where(denom > 0, bias, 1.0) * depo
        ca_d, cb_d = self._dev(ca), self._dev(cb)
        self._call("gospl_sed_lake_init", self.H, _ptr(self.pitIDs), _ptr(sb["sel"]), _ptr(inlet), _ptr(ca_d),
                   _ptr(cb_d), _ptr(denom_d), _ptr(count_d), _ptr(sb["wd"]), _ptr(sb["b"]), S)
        self._diffuseImplicit(sb["b"], sb["sel"], self.nl_pit_K, out)
        vol_d = torch.zeros(npit, dtype=torch.float64, device=self.device)
        self._call("gospl_sed_pit_volume", self.H, _ptr(out), _ptr(sb["sel"]), _ptr(self.pitIDs), npit, _ptr(vol_d), S)
        actual = vol_d.cpu().numpy()
        with np.errstate(divide="ignore", invalid="ignore"):
            scale = np.where(actual > 0, depo / np.maximum(actual, 1e-30), 1.0)
        scale_d = self._dev(scale)
        self._call("gospl_sed_scale", self.H, _ptr(self.pitIDs), _ptr(sb["sel"]), _ptr(scale_d), _ptr(out), S)
        return out

    def _updateSinks(self, hl):
        depo = self.pitParams[:, 0] - self.pitVol
        depo[depo < 0] = 0.0
        pit_volume = self.pitParams[:, 0]
        pit_depth = self.pitParams[:, 2]
        active = depo > 0
        full_fill = active & (depo >= pit_volume - 1.0e-6)
        partial = active & ~full_fill
        is_large = (pit_volume >= self.nl_pit_volume) & (pit_depth >= self.nl_pit_depth)
        diffuse_path = partial & is_large
        bottomup_path = full_fill | (partial & ~is_large)
        delta = self._bottomUpDelta(hl, depo, bottomup_path)
        if bottomup_path.any():
            delta = self._addPitMicroTilt(delta, bottomup_path)
        if diffuse_path.any():
            delta += self._diffuseLargePit(hl, depo, diffuse_path)
        self._call("gospl_sed_apply", self.H, _ptr(delta), _ptr(self.hLocal), _ptr(self.cumED), self.stream)
        self.pitParams[:, 0] = self.pitVol.copy()
        self.sed_deposit = delta

    def sedChange(self):
        """SEDMesh.sedChange (sedplex.py:1061-1086)."""
        self._getSedFlux()
        self.fillElevation(sed=True)
        hl = self.hLocal.clone()
        self.lsink = self.lsinki.clone()
        self.pitVol = self.pitParams[:, 0].copy()
        self.sedFilled = hl.clone()
        self._distributeSediment(hl)
        self._updateSinks(hl)
