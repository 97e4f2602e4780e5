"""Write the golden vectors of tests/golden/*.npz from goSPL's own Fortran.

Run in a container that has goSPL's source tree (the reference Fortran is
compiled from it into oracle/_ref by `make -C oracle ref`):

    python tests/golden/make_golden.py

For each mesh (goSPL's test meshes `tests/fixtures/mesh.npz` (3766-vertex
sphere) and `incising.npz` (139-vertex plane), plus a level-3 icosphere and a
30x30 jittered plane generated here) it stores the inputs and what the
reference kernels return for them:
  definetin (neighbours, edge lengths, Voronoi face lengths, areas),
  mfdreceivers on two elevation fields (the second rounded to force ties),
  the one-rank fill chain (edge_tile, fill_tile, graph_nodes, combine_edges,
  fill_edges, fill_depressions driven through pitfilling.py's glue),
  label_pits on the filled surface, sethillslopecoeff for a seeded Kd,
  fill_dir + nghb_dir from the oracle's spill points, and getpitvol.
The files are data (inputs and expected outputs); the reference itself is not
copied.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gospl_amd import _lib, mesh as gmesh  # noqa: E402
from oracle import oracle as orc, ref  # noqa: E402

REF_FIX = "/root/reference/tests/fixtures"


def ref_fill_chain(dem, outlet_ids, N):
    """pitfilling.py:557-702 at one rank, with the Fortran kernels."""
    import pandas as pd

    borders = -np.ones((N, 2), dtype=np.int32)
    borders[outlet_ids, 0] = outlet_ids
    borders[outlet_ids, 1] = 1
    ledges = ref.edge_tile(0.0, borders, dem)
    out = np.where(ledges >= 0)[0]
    local = np.empty((len(out), 2), dtype=np.int32)
    local[:, 0] = out
    local[:, 1] = ledges[out]
    inids = np.ones(N, np.int32)
    inids[ledges == -2] = 0
    outedges = np.zeros(N, np.int32)
    lfill, label, gnb = ref.fill_tile(local, dem, inids)
    lgth = 0
    graph = None
    if gnb > 0:
        graph = ref.graph_nodes(gnb)
        lgth = int(np.amax(graph[:, :2]))
    cgraph, graphnb = ref.combine_edges(lfill, label, local[:, 0], outedges)
    if graphnb > 0 and lgth > 0:
        cg = np.concatenate((graph, cgraph[:graphnb]))
    elif graphnb > 0:
        cg = cgraph[:graphnb]
    elif lgth > 0:
        cg = graph
    else:
        cg = None
    if cg is None:
        gg = -np.ones((int(label.max()) + 2, 5))
    else:
        mg = np.concatenate([cg, np.zeros((len(cg), 1))], axis=1)
        c = (pd.DataFrame(mg, columns=["source", "target", "elev", "spill", "rank"])
             .sort_values("elev").drop_duplicates(["source", "target"], keep="first").values)
        nbg = int(c[:, 1].max()) + 2
        src = c[:, 0].astype(np.int64)
        tgt = c[:, 1].astype(np.int64)
        uniq = np.unique(np.concatenate((src, tgt)))
        if uniq[0] != 0:
            uniq = np.concatenate(([0], uniq))
        nlab = len(uniq)
        cc = c.copy()
        cc[:, 0] = np.searchsorted(uniq, src)
        cc[:, 1] = np.searchsorted(uniq, tgt)
        cmax = np.max(np.bincount(cc[:, :2].astype(int).ravel())) + 1
        elev, _, _, _ = ref.fill_edges(nlab + 1, cc, cmax)
        gg = -np.ones((nbg, 5))
        gg[:, 0] = 1e8
        gg[uniq, 0] = elev[:nlab]
    g0 = gg[:, 0].copy()
    g0[g0 < -1.0e8] = -1.0e8
    g0[g0 > 1.0e7] = -1.0e8
    return ref.fill_depressions(0.0, dem, lfill, label, g0)


def meshes():
    out = []
    for name in ("mesh", "incising"):
        d = np.load(os.path.join(REF_FIX, name + ".npz"))
        v, c, z = d["v"], d["c"].astype(np.int32), d["z"]
        flat = bool(np.allclose(v[:, 2], 0.0))
        out.append((name, v, c, z, flat))
    v, c = gmesh.icosphere(3, 6.371e6)
    rng = np.random.default_rng(11)
    out.append(("ico3", v, c, rng.uniform(-3000, 3000, len(v)), False))
    v, c = gmesh.planar_jittered(30, 30, 500.0, 100.0, seed=5)
    z = rng.uniform(0, 60, len(v))
    out.append(("plane30", v, c, z, True))
    return out


def main():
    lib = _lib.load()
    rng = np.random.default_rng(2024)
    for name, v, c, z, flat in meshes():
        N = len(v)
        vi = gmesh.voronoi_inputs(v, c, flat)
        ngb, nnb, elgt, vdist, area, narea = ref.fv_setup(v, c, vi["cells_edges"], vi["edges"],
                                                          vi["circumcenters"], 0 if flat else 1)
        bv = gmesh.boundary_vertices(c) if flat else np.zeros(0, np.int64)
        outlet = np.zeros(N, np.int8)
        outlet[bv] = 1
        gid = np.arange(N, dtype=np.int32)
        rec = {}
        for t, h in enumerate((z, np.round(z / 50.0) * 50.0)):
            r, d_, w = ref.mfdreceivers(3, 1.1, h, 0.0, gid)
            rec["mfd%d_h" % t] = h
            rec["mfd%d_rcv" % t] = r
            rec["mfd%d_dist" % t] = d_
            rec["mfd%d_wgt" % t] = w
        level = z.min() + (0.0 if flat else 1.0e-3)
        dem = z - level
        fill = ref_fill_chain(dem, bv, N)
        lab = ref.label_pits(0.0, fill + level)
        kd = rng.uniform(0.1, 2.0, N)
        hc = ref.sethillslopecoeff(kd)
        # flat-routing distances from the oracle's spill points on the filled surface
        om = orc.OracleModel(dict(ngb=ngb, nnb=nnb, elgt=elgt, vdist=vdist, area=area,
                                  outlet=outlet.astype(bool), flat=flat), z, K=1e-5, sealevel=0.0)
        om.fillElevation()
        spill = om.lspillIDs.astype(np.int32)
        pit = om.pitIDs.astype(np.int32)
        pd0 = ref.fill_dir(spill, pit, om.lFill) if (pit > -1).any() else -np.ones(N, np.int32)
        pd1 = ref.nghb_dir(pit, om.lFill, pd0)
        npit = len(om.pitParams)
        hlvl = np.ascontiguousarray(om.filled_lvl[:, :4]) if npit else np.zeros((1, 4))
        hclip = np.maximum(z, 0.0)
        pv = ref.getpitvol(hlvl, hclip, pit, np.ones(N, np.int32))
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            v=v, c=c, z=z, flat=flat, outlet=outlet,
            ngb=ngb, nnb=nnb, elgt=elgt, vdist=vdist, area=area, narea=narea,
            level=level, dem=dem, fill=fill, labels=lab, kd=kd, hillcoeff=hc,
            spill=spill, pitids=pit, lfill=om.lFill, ptdir_fill=pd0, ptdir_relax=pd1,
            pitlvl=hlvl, pitvol=pv, **rec)
        print("wrote", name, N, "pits", npit - 1)


if __name__ == "__main__":
    main()
