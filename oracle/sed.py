"""CPU FP64 restatement of goSPL's continental sediment routing (SEDMesh).

TEST INFRASTRUCTURE ONLY: mixed into ``oracle.OracleModel``; the product
(``gospl_amd``) never imports it.

Restates, method by method, in input-mesh order and one rank (no dual
lithology, provenance or stratigraphy, which are off on this path):

* ``_getSedFlux``          gospl/sed/sedplex.py:103-190
* ``_moveDownstream``      gospl/sed/sedplex.py:199-374
* ``_distributeSediment``  gospl/sed/sedplex.py:376-523
* ``_bottomUpDelta``       gospl/sed/sedplex.py:525-560
* ``_addPitMicroTilt``     gospl/sed/sedplex.py:584-655
* ``_identifyPitInlets``   gospl/sed/sedplex.py:657-697
* ``_diffuseLargePit``     gospl/sed/sedplex.py:699-926
* ``_updateSinks``         gospl/sed/sedplex.py:982-1059
* ``sedChange``            gospl/sed/sedplex.py:1061-1086
* ``_diffuseImplicit``     gospl/sed/hillslope.py:925-1076 (the Picard form,
  goSPL's ``marineSolver: picard``), with the jacobiancoeff operator of
  fortran/functions.F90:1421-1467 (zero derivative term) assembled as a
  sparse matrix and solved directly (SuperLU) where goSPL runs Richardson +
  block Jacobi to rtol 1e-8.

Pins (tests/test_oracle_sed_pins.py): the sediment volume balance of a whole
call, full and knot-volume bottom-up deposits against the depth-volume
curves, the tilt's zero per-pit mean, inlets by brute force over edges, the
Picard operator's volume conservation and its constant-thickness fixed point,
and the implicit-Euler step against the heat kernel's decay on a 1-D chain.

Parity unpinned: goSPL's default ``marineSolver: ts`` (PETSc's adaptive
Rosenbrock-W integrator, atol 5e-3 / rtol 1e-4) is not restated; the Picard
form stands for it.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def jacobian_rows(h, Kd, ngb, nnb, area, elgt, vdist):
    """jacobiancoeff (functions.F90:1421-1467) with Kp = 0: the diagonal
    (sum of v over the faces, in neighbour order) and the off-diagonal -v per
    neighbour slot; a face counts when its Voronoi length is positive and
    both cells diffuse."""
    N, W = ngb.shape
    diag = np.zeros(N)
    off = np.zeros((N, W))
    for p in range(W):
        valid = (p < nnb) & (vdist[:, p] > 0.0) & (area > 0.0) & (ngb[:, p] >= 0)
        n = np.where(valid, ngb[:, p], 0)
        ck = Kd
        cn = Kd[n]
        ok = valid & (ck > 0.0) & (cn > 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            c = 0.5 * (ck + cn) / area
            v = c * vdist[:, p] / elgt[:, p]
        diag[ok] = diag[ok] + v[ok]
        off[ok, p] = -v[ok]
    return diag, off


def rows_matrix(diag, off, ngb):
    N, W = ngb.shape
    r = [np.arange(N)]
    c = [np.arange(N)]
    d = [diag]
    for p in range(W):
        sel = off[:, p] != 0.0
        r.append(np.nonzero(sel)[0])
        c.append(ngb[sel, p])
        d.append(off[sel, p])
    return sp.csr_matrix((np.concatenate(d), (np.concatenate(r), np.concatenate(c))), shape=(N, N))


class SED:
    """goSPL's SEDMesh glue for the oracle model."""

    def _init_sed(self, nodep=True, nlPitVolume=1.0e9, nlPitDepth=100.0, nlPitK=None, nonlinKm=10.0,
                  pitInletBias=0.10, picardSub=10, picardIts=2, marineSolver="ts"):
        self.nodep = bool(nodep)
        self.nlK = float(nonlinKm)
        self.nl_pit_volume = float(nlPitVolume)
        self.nl_pit_depth = float(nlPitDepth)
        self.nl_pit_K = float(nlPitK) if nlPitK is not None else self.nlK
        self.nl_pit_inlet_bias = min(1.0, max(0.0, float(pitInletBias)))
        self.picardSub = max(1, int(picardSub))
        self.picardIts = max(1, int(picardIts))
        self.dexp = 0.05
        self.minDiff = 1.0e-4
        self._domainHasOutlet = bool(len(self.outletIDs) > 0)
        self.vSedLocal = np.zeros(self.lpoints)

    # ---------------------------------------------------------------- flux
    def _getSedFlux(self):
        from .oracle import flowacc
        tmp = -self.Eb * self.larea
        q, bad = flowacc(self.rcvIDi, self.wghtVali, tmp, self.ngb, self.nnb)
        self.vSedLocal = np.zeros_like(q) if bad else q

    # ---------------------------------------------------------------- cascade
    def _moveDownstream(self, vSed, step):
        excess = False
        sedFilled_changed = False
        vSed = np.multiply(vSed, self.inIDs)
        csink = self.lsink & (self.pitIDs < 0)
        if csink.any():
            if not self._domainHasOutlet:
                self._closedDepo[csink] += vSed[csink]
            vSed = vSed.copy()
            vSed[csink] = 0.0
        inV = np.zeros(len(self.pitParams), dtype=np.float64)
        sel = self.lsink & (self.pitIDs > -1)
        np.add.at(inV, self.pitIDs[sel], vSed[sel])
        nvSed = np.zeros(self.lpoints, dtype=np.float64)
        eV = inV - self.pitVol
        if (eV > 0.0).any():
            eIDs = eV > 0.0
            spillIDs = self.pitInfo[eIDs, 0]
            localSpill = np.where(self.pitInfo[eIDs, 1] == 0)[0]
            localPts = spillIDs[localSpill]
            nvSed[localPts] = eV[eIDs][localSpill]
            self.pitVol[eIDs] = 0.0
            ids = np.isin(self.pitIDs, np.where(eV > 0.0)[0])
            self.sedFilled[ids] = self.lFill[ids]
            sedFilled_changed = True
        eIDs = eV < 0
        self.pitVol[eIDs] -= inV[eIDs]
        self.pitVol[self.pitVol < 0] = 0.0
        self._routed = None
        if (eV > 1.0e-3).any():
            if sedFilled_changed or step == 0 or step == 100:
                if step == 100:
                    self._buildFlowDirection(self.lFill)
                else:
                    self._buildFlowDirection(self.sedFilled)
            total = 0.0
            for x in nvSed:          # PETSc's VecSum: a sequential sum
                total += x
            if total > 0.5 * self.maxarea[0]:
                excess = True
                self._routed = self._solve_flow(nvSed, fatal=False)
        return excess

    def _distributeSediment(self, hl):
        vSed = self.vSedLocal.copy() * self.dt
        self._closedDepo = np.zeros(self.lpoints, dtype=np.float64)
        step = 0
        excess = True
        while excess:
            if step >= 5000:
                break
            excess = self._moveDownstream(vSed, step)
            if excess:
                vSed = self._routed.copy()
                nvSed = vSed.copy()
                nvSed[hl < self.sedFilled] = 0.0
                vSed[self.outletIDs] = 0.0
                self.vSedLocal = self.vSedLocal + 1.0 * (nvSed / self.dt)
            step += 1
        self.sed_passes = step
        if self._closedDepo.any():
            delta = np.zeros(self.lpoints)
            nz = self._closedDepo > 0.0
            delta[nz] = self._closedDepo[nz] / self.larea[nz]
            self.cumED = self.cumED + 1.0 * delta
            self.h = self.h + 1.0 * delta

    # ---------------------------------------------------------------- deposits
    def _targetLevels(self, depo, pits):
        target = np.full(len(self.pitParams), -np.inf, dtype=np.float64)
        for p in pits:
            pit_min = self.pitParams[p, 1] - self.pitParams[p, 2]
            vols = np.concatenate(([0.0], self.filled_vol[p]))
            lvls = np.concatenate(([pit_min], self.filled_lvl[p]))
            target[p] = np.interp(depo[p], vols, lvls)
        return target

    def _bottomUpDelta(self, hl, depo, pit_select):
        active = np.where(pit_select & (depo > 0))[0]
        target_lvl = self._targetLevels(depo, active)
        node_target_lvl = np.full(self.lpoints, -np.inf, dtype=np.float64)
        in_pit = self.pitIDs >= 0
        node_target_lvl[in_pit] = target_lvl[self.pitIDs[in_pit]]
        surface = np.minimum(node_target_lvl, self.lFill)
        delta = np.maximum(0.0, surface - hl)
        sel_node = np.isin(self.pitIDs, active) if active.size > 0 else np.zeros(self.lpoints, bool)
        delta[~sel_node] = 0.0
        return delta

    def _spillCoords(self):
        num_pits = len(self.pitParams)
        spill_xyz = np.zeros((num_pits, 3), dtype=np.float64)
        spill_idx = self.pitInfo[:, 0].astype(int)
        my = np.where(self.pitInfo[:, 1].astype(int) == 0)[0]
        mp = my[spill_idx[my] >= 0]
        spill_xyz[mp] = self.coords[spill_idx[mp]]
        return spill_xyz

    def _addPitMicroTilt(self, delta, pit_select):
        num_pits = len(self.pitParams)
        if not pit_select.any():
            return delta
        active_pit = np.where(pit_select)[0]
        sel_node = np.isin(self.pitIDs, active_pit) & (delta > 0)
        spill_xyz = self._spillCoords()
        in_pit = self.pitIDs >= 0
        pit_safe = np.where(in_pit, self.pitIDs, 0)
        dist = np.linalg.norm(self.coords - spill_xyz[pit_safe], axis=1)
        local_sum = np.zeros(num_pits)
        local_cnt = np.zeros(num_pits)
        np.add.at(local_sum, self.pitIDs[sel_node], dist[sel_node])
        np.add.at(local_cnt, self.pitIDs[sel_node], 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            mean_dist = np.where(local_cnt > 0, local_sum / np.maximum(local_cnt, 1.0), 0.0)
        tilt = 1.0e-6 * (dist - mean_dist[pit_safe])
        tilt[~sel_node] = 0.0
        tilted = delta + tilt
        tilted[tilted < 0] = 0.0
        return tilted

    def _identifyPitInlets(self):
        flowDir = self.rcvIDi.shape[1]
        src = np.repeat(np.arange(self.lpoints), flowDir)
        dst = self.rcvIDi.flatten().astype(int)
        good = src != dst
        src = src[good]
        dst = dst[good]
        src_pit = self.pitIDs[src]
        dst_pit = self.pitIDs[dst]
        inlet_edges = (dst_pit >= 0) & (src_pit != dst_pit)
        inlet_mask = np.zeros(self.lpoints, dtype=bool)
        inlet_mask[dst[inlet_edges]] = True
        return inlet_mask

    def _diffuseImplicit(self, dh, mask, Cd_val):
        """_diffuseImplicitPicard (hillslope.py:925-1076) with direct solves."""
        nsub, npic = self.picardSub, self.picardIts
        dt_sub = self.dt / nsub
        Cd0 = np.zeros(self.lpoints)
        Cd0[mask] = Cd_val
        minDiff_vec = np.zeros(self.lpoints)
        minDiff_vec[mask] = self.minDiff
        zb = self.h.copy()
        h = dh + zb
        for _ in range(nsub):
            hstart = h.copy()
            for _ in range(npic):
                dhdep = h - zb
                dhdep[dhdep < 0.1] = 0.0
                Cd = minDiff_vec + np.multiply(Cd0, 1.0 - np.exp(-self.dexp * dhdep))
                diag, off = jacobian_rows(h, Cd, self.ngb, self.nnb, self.larea, self.elgt, self.vdist)
                M = rows_matrix(dt_sub * diag + 1.0, dt_sub * off, self.ngb)
                h = spla.spsolve(M.tocsc(), hstart)
        ndepo = h - zb
        ndepo[ndepo < 0.0] = 0.0
        return ndepo

    def _diffuseLargePit(self, hl, depo, pit_select):
        num_pits = len(self.pitParams)
        active = np.where(pit_select & (depo > 0))[0]
        if active.size == 0:
            return np.zeros(self.lpoints, dtype=np.float64)
        sel_node = np.isin(self.pitIDs, active)
        inlet_mask = self._identifyPitInlets() & sel_node
        inlet_count = np.zeros(num_pits, dtype=np.float64)
        np.add.at(inlet_count, self.pitIDs[inlet_mask], 1.0)
        no_inlet = active[inlet_count[active] == 0]
        for p in no_inlet:
            pit_local = np.where(self.pitIDs == p)[0]
            if pit_local.size > 0:
                k = np.argmax(hl[pit_local])
                inlet_mask[int(pit_local[k])] = True
                inlet_count[p] = 1.0
        bias_frac = self.nl_pit_inlet_bias
        target_lvl = self._targetLevels(depo, active)
        in_pit = self.pitIDs >= 0
        pit_safe = np.where(in_pit, self.pitIDs, 0)
        node_lvl = np.where(in_pit, target_lvl[pit_safe], -np.inf)
        water_depth = np.maximum(0.0, node_lvl - hl)
        water_depth[~sel_node] = 0.0
        denom = np.zeros(num_pits, dtype=np.float64)
        np.add.at(denom, self.pitIDs[sel_node], (water_depth * self.larea)[sel_node])
        delta_init = np.zeros(self.lpoints, dtype=np.float64)
        denom_node = denom[pit_safe]
        ok = (denom_node > 0) & sel_node
        delta_init[ok] = ((1.0 - bias_frac) * depo[pit_safe[ok]] * water_depth[ok]
                          / np.maximum(denom_node[ok], 1e-30))
        bias_at_inlet = np.where(denom > 0, bias_frac, 1.0)
        for c in np.where(inlet_mask)[0]:
            p = self.pitIDs[c]
            if inlet_count[p] > 0:
                delta_init[c] += bias_at_inlet[p] * depo[p] / (inlet_count[p] * self.larea[c])
        delta_smooth = self._diffuseImplicit(delta_init, sel_node, self.nl_pit_K)
        actual_vol = np.zeros(num_pits, dtype=np.float64)
        np.add.at(actual_vol, self.pitIDs[sel_node], (delta_smooth * self.larea)[sel_node])
        with np.errstate(divide="ignore", invalid="ignore"):
            scale = np.where(actual_vol > 0, depo / np.maximum(actual_vol, 1e-30), 1.0)
        node_scale = scale[pit_safe]
        node_scale[~sel_node] = 0.0
        return delta_smooth * node_scale

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
            delta = delta + self._diffuseLargePit(hl, depo, diffuse_path)
        self.sed_deposit = delta
        self.sed_depo = depo
        self.cumED = self.cumED + 1.0 * delta
        self.h = self.h + 1.0 * delta
        self.pitParams[:, 0] = self.pitVol.copy()

    def sedChange(self):
        self._getSedFlux()
        self.fillElevation(sed=True)
        hl = self.h.copy()
        self.lsink = self.lsinki.copy()
        self.pitVol = self.pitParams[:, 0].copy()
        self.sedFilled = hl.copy()
        self._distributeSediment(hl)
        self._updateSinks(hl)
