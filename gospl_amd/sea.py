"""Marine deposition of river sediment: goSPL's ``SEAMesh``
(gospl/sed/seaplex.py) on the GPU, mixed into ``Model``.

It runs when ``nodep`` is False and ``seadepo`` is True (goSPL's defaults),
after the continental sediment routing of the step (gospl/model.py:411-420):

* the distance of every marine vertex to the coastline (``_distanceCoasts``,
  seaplex.py:51-135): the contour of the surface at sea level and an exact
  nearest-point search on the GPU (``gospl_coast_distance``);
* the shelf capacity below the clinoform ``clinoH = sealevel - 1e-3 -
  distance * clinSlp`` and the volumes the rivers bring to the sea
  (:757-776, ``gospl_marine_prep``);
* the marine flow directions (``_matOcean``, :137-238): the surface smoothed
  by the marine diffusion operator (``_hillSlope(smooth=2)``,
  ``gospl_diffuse_fixed``), cut at ``offshore`` metres from the coast,
  epsilon-filled (``epsfill``, ``gospl_eps_fill``) and split evenly over all
  lower neighbours (``mfdrcvrs``, ``gospl_marine_receivers``);
* the downslope routing of the volumes against the capacities
  (``_distOcean``, :404-619, ``gospl_marine_route``), with ``Gmar > 0`` the
  coupled deposition solve (``_depMarineSystem``, :240-402,
  ``gospl_marine_dep``) and a second routing;
* the deposit thickness (below 1 mm dropped) spread by the non-linear marine
  diffusion (``_diffuseOcean``, hillslope.py:739-786, the same
  ``gospl_nl_diffusion`` as the lake deposits) and added to the elevation,
  the cumulative erosion/deposition and the rate (:836-847).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

BOUNDARY_FLOW_SENTINEL = -1.0e6   # gospl/tools/constants.py:90
DEPOSIT_FLOOR = 1.0e-3            # gospl/tools/constants.py:58
MARINE_SMOOTH_N_LAND = 1.0        # gospl/tools/constants.py:126
MARINE_SMOOTH_N_SEA = 5.0         # gospl/tools/constants.py:127


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class SEAMesh:
    """seaChange and its helpers for ``Model`` (goSPL's method names)."""

    def _init_sea(self, seadepo=False, clinSlp=1.0e-6, offshore=100.0e5, Gmar=0.0):
        """The marine keys of goSPL's input (inputparser.py:546-597)."""
        self.seaDepo = bool(seadepo)
        self.clinSlp = max(1.0e-6, float(clinSlp))
        self.offshore = float(offshore)
        self.Gmar = float(Gmar)
        self._seabuf = None
        self._seaFlow = None
        self.coastDist = None
        # marine coupled solve: GMRES to this relative residual, at most this many iterations
        self.mar_rtol, self.mar_maxit = 1.0e-13, 400
        self.sea_stats = {}

    def _sea_buffers(self):
        if self._seabuf is None:
            N = self.lpoints
            f64 = dict(dtype=torch.float64, device=self.device)
            names = ("cdist", "cap", "vin", "vdep", "vdep2", "kd", "hsmth", "fillz", "dh", "ndepo")
            self._seabuf = {k: torch.zeros(N, **f64) for k in names}
            self._seabuf["rcv"] = torch.zeros((12, N), dtype=torch.int32, device=self.device)
            self._seabuf["wgt"] = torch.zeros((12, N), **f64)
            self._seabuf["sink"] = torch.zeros(N, dtype=torch.int8, device=self.device)
            self._seabuf["xyz"] = torch.as_tensor(np.ascontiguousarray(self.dm.mesh.coords, np.float64)).to(
                self.device)
        return self._seabuf

    def _snapshot_sea(self):
        """goSPL's seaID as the last flow-direction build leaves it
        (flowplex.py:493): filled elevation at or below sea level."""
        if self.seaDepo and not self.nodep:
            self._seaFlow = (self.lFill <= self.sealevel).to(torch.int8)

    # ------------------------------------------------------------ coastline
    def _distanceCoasts(self, data):
        sb = self._sea_buffers()
        n = ctypes.c_int64(0)
        self._call("gospl_coast_distance", self.H, _ptr(sb["xyz"]), _ptr(data), self.sealevel, _ptr(self._seaFlow),
                   _ptr(sb["cdist"]), ctypes.byref(n), self.stream)
        self.coastDist = sb["cdist"]
        self.sea_stats["coast_points"] = int(n.value)

    # ------------------------------------------------------------ directions
    def _matOcean(self):
        sb = self._sea_buffers()
        S = self.stream
        hl = self.hLocal
        minh = float(torch.min(hl)) + 0.1
        minh = max(minh, self.oFill)
        if not self.flatModel:
            # _hillSlope(smooth=2): Kd = N area, N by the sea flags of the flow build
            f64 = dict(dtype=torch.float64, device=self.device)
            cd = torch.where(self._seaFlow.bool(), torch.tensor(MARINE_SMOOTH_N_SEA, **f64),
                             torch.tensor(MARINE_SMOOTH_N_LAND, **f64))
            torch.mul(cd, self.dm.area, out=sb["kd"])
            sw = ctypes.c_int32(0)
            self._call("gospl_diffuse_fixed", self.H, _ptr(sb["kd"]), _ptr(hl), _ptr(sb["hsmth"]), 1000000,
                       ctypes.byref(sw), S)
            self.sea_stats["smooth_sweeps"] = int(sw.value)
            far = self.coastDist > self.offshore
            sb["hsmth"].masked_fill_(far, BOUNDARY_FLOW_SENTINEL)
        else:
            sb["hsmth"].copy_(hl)
            sb["hsmth"].masked_fill_(self.dm.outlet.bool(), BOUNDARY_FLOW_SENTINEL)
        ps = ctypes.c_int32(0)
        self._call("gospl_eps_fill", self.H, _ptr(sb["hsmth"]), minh, _ptr(sb["fillz"]), ctypes.byref(ps), S)
        self.sea_stats["epsfill_passes"] = int(ps.value)
        if not self.flatModel:
            sb["fillz"][far] = hl[far]
        self._call("gospl_marine_receivers", self.H, _ptr(sb["fillz"]), _ptr(sb["rcv"]), _ptr(sb["wgt"]),
                   _ptr(sb["sink"]), S)

    # ------------------------------------------------------------- routing
    def _distOcean(self, vin, vdep):
        sb = self._sea_buffers()
        ps = ctypes.c_int32(0)
        self._call("gospl_marine_route", self.H, _ptr(sb["rcv"]), _ptr(sb["wgt"]), _ptr(sb["sink"]), _ptr(sb["cap"]),
                   _ptr(vin), _ptr(vdep), ctypes.byref(ps), self.stream)
        self.sea_stats.setdefault("route_passes", []).append(int(ps.value))
        return vdep

    def _depMarineSystem(self, marDep, out):
        sb = self._sea_buffers()
        it = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        self._call("gospl_marine_dep", self.H, _ptr(sb["rcv"]), _ptr(sb["wgt"]), _ptr(self.hLocal), _ptr(marDep),
                   self.Gmar, self.sealevel, self.dt, _ptr(out), self.mar_rtol, self.mar_maxit, ctypes.byref(it),
                   ctypes.byref(rel), self.stream)
        self.sea_stats["dep_iters"] = int(it.value)
        self.sea_stats["dep_relres"] = float(rel.value)
        return out

    def seaChange(self):
        """SEAMesh.seaChange (seaplex.py:738-861)."""
        if self._seaFlow is None:
            self._snapshot_sea()
        sb = self._sea_buffers()
        S = self.stream
        self.sea_stats = {}
        self._distanceCoasts(self.hLocal)
        self._call("gospl_marine_prep", self.H, _ptr(self.hLocal), _ptr(sb["cdist"]), _ptr(self.lFill),
                   _ptr(self.vSedLocal), self.sealevel, self.clinSlp, self.dt, _ptr(sb["cap"]), _ptr(sb["vin"]), S)
        self._matOcean()
        marDep = self._distOcean(sb["vin"], sb["vdep"])
        if not self.flatModel and self.Gmar > 0.0:
            vdep = self._depMarineSystem(marDep, sb["vdep2"])
            marDep = self._distOcean(vdep, sb["vdep"])
        self._call("gospl_marine_thickness", self.H, _ptr(marDep), DEPOSIT_FLOOR, _ptr(sb["dh"]), S)
        # _diffuseOcean (hillslope.py:739-786)
        self._diffuseImplicit(sb["dh"], self._seaFlow, self.nlK, sb["ndepo"])
        self._call("gospl_marine_apply", self.H, _ptr(sb["ndepo"]), self.dt, _ptr(self.hLocal), _ptr(self.cumED),
                   _ptr(self.EbLocal), S)
        self.marine_deposit = sb["ndepo"]
