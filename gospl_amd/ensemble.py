"""Ensembles of independent models stepped together on one GPU (no
counterpart in goSPL, which runs one model per process).

The members' meshes are joined into one union mesh -- member k's vertices at
[offsets[k], offsets[k] + N_k), its range padded to whole 256-vertex
workgroups with isolated deep-sea vertices -- so that every kernel of a step
runs once for all members (one launch per kernel, the receiver-DAG sweeps in
one persistent dataflow launch), each member with its own parameters: K and
the rain scale per vertex (kfac, rain), the stream-power exponent m per
vertex (``gospl_spl_batch``).

Every decision that goSPL takes over a whole model is taken per member, so
a member's trajectory is the one its own ``Model`` computes, bit for bit:

* the fill level max(min h + 1e-3, sealevel + oFill) (pitfilling.py:849-870)
  per member (``gospl_fill_elevation_seg``);
* the pit-overflow cascade (flowplex.py:576-686, 849-904): per-member pass
  loop, stall test, maximum-area threshold, receiver rebuilds and the "last
  built" sink flags that erodepSPL reads; the union's cascade solve carries
  only the spills of members that solve in that pass, on receivers rebuilt
  for them (self-receivers elsewhere);
* the discharge maximum and pit flag of flowAccumulation's end
  (flowplex.py:906-933, ``gospl_flow_finalize_seg``);
* the refined hillslope solve's sweep count, residual norm and stop test
  (``gospl_hillslope_refine_seg``);
* the receivers' slope tie-break ids (``gospl_mesh_set_tiebreak``): each
  member's own.

Members share dt, flowdir, the diffusivities, sea level and oFill, and run
the refined (FP32 hillslope) mode.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .mesh import Mesh
from .model import Model, _ptr, contraction_gmax, sweeps_for

BLOCK = 256
PAD_H = -1.0e4            # padding vertices: isolated, below every fill level


def union_mesh(meshes):
    """Join meshes (device order each) into one.  Returns (Mesh, offsets
    (nmem+1,) int64, seg (N,) int32 with -1 on padding, tgid (N,) int32 the
    members' own ids)."""
    if len({m.flat for m in meshes}) != 1:
        raise ValueError("union_mesh: members must all be spheres or all planar")
    sizes = [m.N for m in meshes]
    padded = [((n + BLOCK - 1) // BLOCK) * BLOCK for n in sizes]
    offsets = np.zeros(len(meshes) + 1, np.int64)
    offsets[1:] = np.cumsum(padded)
    N = int(offsets[-1])
    W = max(m.width for m in meshes)
    nbr = np.full((W, N), -1, np.int32)
    elgt = np.zeros((W, N))
    vdist = np.zeros((W, N))
    deg = np.zeros(N, np.int32)
    area = np.ones(N)
    gid = np.arange(N, dtype=np.int32)
    tgid = np.zeros(N, np.int32)
    outlet = np.zeros(N, np.int8)
    coords = np.zeros((N, 3))
    seg = np.full(N, -1, np.int32)
    for k, m in enumerate(meshes):
        o, n, w = int(offsets[k]), m.N, m.width
        nbr[:w, o:o + n] = np.where(m.nbr >= 0, m.nbr + o, -1)
        elgt[:w, o:o + n] = m.elgt
        vdist[:w, o:o + n] = m.vdist
        deg[o:o + n] = m.deg
        area[o:o + n] = m.area
        gid[o:o + n] = m.gid + o          # a permutation of the member's range
        tgid[o:o + n] = m.gid
        outlet[o:o + n] = m.outlet
        coords[o:o + n] = m.coords
        seg[o:o + n] = k
    perm = np.arange(N)
    u = Mesh(N=N, width=W, flat=meshes[0].flat, coords=coords, nbr=nbr, deg=deg, elgt=elgt, vdist=vdist,
             area=area, gid=gid, perm=perm, iperm=perm.copy(), outlet=outlet,
             maxarea=max(m.maxarea for m in meshes), extras={})
    return u, offsets, seg, tgid


class EnsembleModel(Model):
    """Members stepped together.  ``members`` is a list of (Mesh, h0 (device
    order), params) with params as ``Model``'s keywords; ``K``, ``m``,
    ``rain`` and ``upsub`` may differ between members, the other keywords
    must agree.  ``member(k, t)`` is member k's slice of a per-vertex
    tensor (e.g. ``hLocal``).  ``dag_shape="tuned"`` launches the dataflow
    sweeps with the shape recorded by ``gospl_amd.autotune`` for ensembles
    (None: the library's default); results do not depend on it."""

    COMMON = ("dt", "flowdir", "cda", "cdm", "sealevel", "ofill", "tstart")

    def __init__(self, members, *, device="cuda", tend=None, refine_rtol=1.0e-12, refine_maxpass=4,
                 dag_shape="tuned"):
        if not members:
            raise ValueError("EnsembleModel: no members")
        meshes = [m for m, _, _ in members]
        ps = [dict(p) for _, _, p in members]
        for p in ps:
            for key in ("d", "sedfac"):
                if p.get(key) not in (None, 0.0):
                    raise NotImplementedError("EnsembleModel: %s is not supported per member" % key)
            p.setdefault("ofill", -6000.0)
            p.setdefault("tstart", 0.0)
            p.setdefault("flowdir", 3)
        for key in self.COMMON:
            vals = {float(p[key]) for p in ps}
            if len(vals) != 1:
                raise ValueError("EnsembleModel: members must share %s (got %s)" % (key, sorted(vals)))
        u, offsets, seg, tgid = union_mesh(meshes)
        N = u.N
        h0 = np.full(N, PAD_H)
        rain = np.zeros(N)
        kfac = np.ones(N)
        mvec = np.full(N, 0.5)
        up = np.zeros(N)
        any_up = False
        for k, ((mesh, h, _), p) in enumerate(zip(members, ps)):
            o, n = int(offsets[k]), mesh.N
            h0[o:o + n] = np.asarray(h, np.float64)
            rain[o:o + n] = np.broadcast_to(np.asarray(p.get("rain", 1.0), np.float64), (n,))
            kfac[o:o + n] = float(p["K"])
            mvec[o:o + n] = float(p.get("m", 0.5))
            if p.get("upsub") is not None:
                up[o:o + n] = np.broadcast_to(np.asarray(p["upsub"], np.float64), (n,))
                any_up = True
        p0 = ps[0]
        super().__init__(u, h0, K=1.0, m=0.5, flowdir=int(p0["flowdir"]), cda=p0["cda"], cdm=p0["cdm"],
                         dt=p0["dt"], rain=rain, upsub=up if any_up else None, sealevel=p0["sealevel"],
                         ofill=p0["ofill"], tstart=p0["tstart"], tend=tend, device=device, refine=True,
                         refine_rtol=refine_rtol, refine_maxpass=refine_maxpass)
        dev = self.device
        self.nmem = len(members)
        self.offsets = offsets
        self.sizes = np.array([m.N for m in meshes], np.int64)
        self._kfac = torch.as_tensor(kfac).to(dev)
        self._mvec = torch.as_tensor(mvec).to(dev)
        self.seg = torch.as_tensor(seg).to(dev)
        self._segl = torch.as_tensor(np.where(seg >= 0, seg, self.nmem).astype(np.int64)).to(dev)
        self._seg_host = seg
        self._tgid = torch.as_tensor(tgid).to(dev)
        self._call("gospl_mesh_set_tiebreak", self.H, _ptr(self._tgid))
        self.member_maxarea = np.array([m.maxarea for m in meshes])
        kd = max(self.Cda, self.Cdm) * self.dt
        self._sweeps32 = np.array([sweeps_for(kd, contraction_gmax(m), 27) for m in meshes], np.int32)
        if (self.Cda != 0.0 or self.Cdm != 0.0) and (self._sweeps32 <= 0).any():
            raise NotImplementedError("EnsembleModel: a member needs the converged FP64 hillslope solve "
                                      "(more than 64 FP32 sweeps per correction)")
        K_ = self.flowDir
        i32 = dict(dtype=torch.int32, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self._selfrcv = torch.arange(N, **i32).expand(K_, N).contiguous()
        self._tset = (torch.zeros((K_, N), **i32), torch.zeros((K_, N), **f64), torch.zeros((K_, N), **f64),
                      torch.zeros(N, dtype=torch.int8, device=dev))
        self._crcv = torch.zeros((K_, N), **i32)
        self._cwgt = torch.zeros((K_, N), **f64)
        self._lsink_eff = torch.zeros(N, dtype=torch.int8, device=dev)
        self._has_pits = torch.zeros(self.nmem, dtype=torch.int8, device=dev)
        self.member_cascade_passes = np.zeros(self.nmem, np.int64)
        # launch shape of the dataflow sweeps: the autotuner's choice for
        # ensembles (a union's sweeps are throughput-bound), when recorded
        from .autotune import apply_tuned
        self.dag_shape = apply_tuned(self, "ensemble") if dag_shape == "tuned" else None

    def member(self, k, t):
        o = int(self.offsets[k])
        return t[..., o:o + int(self.sizes[k])]

    def _vmask(self, flags):
        """Per-vertex bool mask from per-member flags (padding: False)."""
        f = torch.as_tensor(np.append(np.asarray(flags, bool), False)).to(self.device)
        return f[self._segl]

    # ---------------------------------------------------------------- pits
    def fillElevation(self, sed=False):
        S = self.stream
        self._call("gospl_fill_elevation_seg", self.H, _ptr(self.hLocal), 0 if self.flatModel else 1,
                   self.sealevel + self.oFill, self.sealevel, _ptr(self.seg), self.nmem, _ptr(self.lFill), S)
        npits = ctypes.c_int64(0)
        self._call("gospl_pit_info", self.H, _ptr(self.hLocal), _ptr(self.lFill), self.sealevel,
                   _ptr(self.pitIDs), _ptr(self.flatDirs), _ptr(self.ptdir), _ptr(self.lspill),
                   self._cap, _ptr(self._spill), _ptr(self._pitparams), _ptr(self._flvl),
                   _ptr(self._fvol), ctypes.byref(npits), S)
        n1 = int(npits.value) + 1
        self.npits = n1 - 1
        # each pit's member (all its cells share one); unused slots -1
        pm = torch.full((n1 + 1,), -1, dtype=torch.int32, device=self.device)
        idx = torch.where(self.pitIDs >= 0, self.pitIDs, n1).long()
        pm.scatter_(0, idx, self.seg)
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
        self.pitMember = pm[:n1].cpu().numpy().astype(np.int64)

    # ---------------------------------------------------------------- flow
    def _pass(self, pitVol, FA, hl, active, steps, lsink_cur):
        """One cascade pass (FAMesh._distributeDownstream, flowplex.py:
        576-686) for the active members.  Returns per-member (go, excess,
        resid) and the next pass's FA (or None)."""
        nm = self.nmem
        npit = len(self.pitParams)
        S = self.stream
        self._call("gospl_pit_inflow", self.H, _ptr(FA), _ptr(lsink_cur), _ptr(self.pitIDs), npit,
                   _ptr(self._inv), S)
        inV = self._inv[:npit].cpu().numpy()
        pmem = self.pitMember
        act_p = (pmem >= 0) & active[np.maximum(pmem, 0)]
        eV = inV - pitVol
        full = np.zeros(npit)
        level = np.full(npit, np.nan)
        resid = np.zeros(nm)
        go = np.zeros(nm, bool)
        u_all = np.zeros(0, np.int64)
        v_all = np.zeros(0)
        pos = (eV > 0.0) & act_p
        if pos.any():
            eIDs = pos
            pitVol[eIDs] = 0.0
            full[eIDs] = 1.0
            spillIDs = self.pitInfo[eIDs, 0]
            localSpill = np.where(self.pitInfo[eIDs, 1] == 0)[0]
            localPts = spillIDs[localSpill]
            vals = eV[eIDs][localSpill]
            if len(localPts):
                # numpy fancy assignment keeps the last value of a repeated
                # index; members never share spill vertices
                u_all, last = np.unique(localPts[::-1], return_index=True)
                v_all = vals[::-1][last]
        neg = (eV < 0.0) & (inV > 0.0) & act_p
        if neg.any():
            eIDs = np.where(neg)[0]
            pitVol[eIDs] += eV[eIDs]
            nid = np.absolute(self.filled_vol[eIDs] - pitVol[eIDs][:, None]).argmin(axis=1)
            level[eIDs] = self.filled_lvl[eIDs, nid]
        gp = (eV > 1.0e-3) & act_p
        if gp.any():
            go[np.unique(pmem[gp])] = True
        umem = self._seg_host[u_all] if len(u_all) else np.zeros(0, np.int64)
        for k in np.unique(umem):
            sel = umem == k       # a contiguous run: u_all is sorted, members' ranges are disjoint
            resid[k] = float(np.sum(v_all[sel] / self.dt))
        excess = go & (resid > self.member_maxarea)
        seed = go | np.zeros(nm, bool)
        fm = pmem[(full > 0) | np.isfinite(level)]
        seed[fm[fm >= 0]] = True
        seed &= active
        nFA = None
        if seed.any():
            keep = excess[umem] if len(u_all) else np.zeros(0, bool)
            u, vv = u_all[keep], v_all[keep]
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
        if go.any():
            src = self.waterFilled
            at100 = go & (steps == 100)
            if at100.any():
                src = torch.where(self._vmask(at100), self.lFill, self.waterFilled)
            rT, dT, wT, lT = self._tset
            self._call("gospl_flow_direction", self.H, self.flowDir, self.flowExp, _ptr(src), self.sealevel,
                       _ptr(self.lFill), 1, _ptr(self.pitIDs), _ptr(self.flatDirs), _ptr(rT), _ptr(dT), _ptr(wT),
                       _ptr(lT), S)
            gv = self._vmask(go)
            lsink_cur.copy_(torch.where(gv, lT, lsink_cur))
            if excess.any():
                ev = self._vmask(excess)
                torch.where(ev, rT, self._selfrcv, out=self._crcv)
                torch.where(ev, wT, torch.zeros_like(wT), out=self._cwgt)
                lev = ctypes.c_int32(0)
                bad = ctypes.c_int64(0)
                self._call("gospl_flowacc", self.H, self.flowDir, _ptr(self._crcv), _ptr(self._cwgt),
                           _ptr(self._rhs), _ptr(self._q), ctypes.byref(lev), ctypes.byref(bad), S)
                Model.sweeps["flow"] += 1
                if bad.value:
                    raise NotImplementedError("EnsembleModel: a cascade solve met an undrained receiver cycle "
                                              "(%d vertices); goSPL zeroes that member's solve" % bad.value)
                self._call("gospl_cascade_add", self.H, _ptr(self._q), _ptr(hl), _ptr(self.waterFilled),
                           self.dt, _ptr(self._nfa), _ptr(self.FAL), S)
                nFA = self._nfa
        return go, excess, resid, nFA

    def flowAccumulation(self):
        self.fillElevation(sed=False)
        pitVol = self.pitParams[:, 0].copy()
        hl = self.hLocal.clone()
        self._buildFlowDirection(hl, False)
        S = self.stream
        self._call("gospl_rain_source", self.H, _ptr(self.rainVal), _ptr(self.lFill), self.sealevel,
                   _ptr(self._b), S)
        self.fa_levels = self._solve_flow(self._b, self.FAL)
        self.waterFilled = hl.clone()
        nm = self.nmem
        pmem = self.pitMember
        has = np.zeros(nm, bool)
        sel = (pitVol > 0.0) & (pmem >= 0)
        has[np.unique(pmem[sel])] = True
        lsink_cur = self._lsink_eff
        lsink_cur.copy_(self.lsinki)
        passes = np.zeros(nm, np.int64)
        if has.any():
            FA = self.FAL * self.dt
            active = has.copy()
            best = np.full(nm, np.nan)
            stall = np.zeros(nm, np.int64)
            while active.any():
                go, excess, resid, nFA = self._pass(pitVol, FA, hl, active, passes, lsink_cur)
                passes[active] += 1
                for k in np.nonzero(active)[0]:
                    if not excess[k]:
                        active[k] = False
                        continue
                    if np.isnan(best[k]) or resid[k] < best[k] * (1.0 - self._cascade_rel_improve):
                        best[k] = resid[k]
                        stall[k] = 0
                    else:
                        stall[k] += 1
                        if stall[k] >= self._cascade_patience:
                            active[k] = False
                            continue
                    if passes[k] > self._cascade_max_steps:
                        active[k] = False
                if nFA is not None:
                    FA = nFA
        self.member_cascade_passes = passes
        self.cascade_passes = int(passes.max()) if nm else 0
        self._has_pits.copy_(torch.as_tensor(has.astype(np.int8)))
        self._call("gospl_flow_finalize_seg", self.H, _ptr(hl), _ptr(self.waterFilled), _ptr(self.pitIDs),
                   _ptr(self.seg), nm, _ptr(self._has_pits), _ptr(self.FAL), _ptr(self.fillFAL), S)
        self.waterFilled = self.waterFilled - hl
        self.lsink = lsink_cur

    # ---------------------------------------------------------------- erosion
    def erodepSPL(self):
        lev = ctypes.c_int32(0)
        Model.sweeps["spl"] += 1
        self._call("gospl_spl_batch", self.H, self.flowDir, _ptr(self.rcvIDi), _ptr(self.wghtVali),
                   _ptr(self.distRcvi), _ptr(self.FAL), _ptr(self._kfac), 1.0, _ptr(self._mvec), self.dt,
                   self.sealevel, _ptr(self.lFill), _ptr(self._lsink_eff), _ptr(self.hLocal), _ptr(self.cumED),
                   _ptr(self.Eb), _ptr(self.EbLocal), ctypes.byref(lev), self.stream)
        self.spl_levels = int(lev.value)

    # ---------------------------------------------------------------- hillslope
    def getHillslope(self):
        if self.Cda == 0.0 and self.Cdm == 0.0:
            self.EbLocal += self.Eb * self.dt / self.dt
            return
        off = np.ascontiguousarray(self.offsets, np.int64)
        sw = np.ascontiguousarray(self._sweeps32, np.int32)
        self._call("gospl_hillslope_refine_seg", self.H, self.Cda, self.Cdm, self.dt, self.sealevel,
                   _ptr(self.seg), self.nmem, off.ctypes.data_as(ctypes.c_void_p),
                   sw.ctypes.data_as(ctypes.c_void_p), self.refine_rtol, self.refine_maxpass, _ptr(self.hLocal),
                   _ptr(self.cumED), _ptr(self.EbLocal), None, self.stream)
