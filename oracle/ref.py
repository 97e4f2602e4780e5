"""ctypes driver for ``oracle/_ref/libgospl_ref.so`` — goSPL's own Fortran
kernels (fortran/functions.F90) compiled from the reference source.

TEST INFRASTRUCTURE ONLY.  Used by ``tests/`` (and the golden-vector script)
to pin the C oracle and the C-ABI geometry against the reference's compiled
code.  Nothing in the product imports this module.

Calling convention: Fortran symbols carry a trailing underscore, every argument
is passed by reference, arrays are column-major, and the hidden f2py
dimensions (``functions.pyf``) are trailing explicit arguments.  ``definetin``
(or ``fv_setup``) must run first: it fills the module state (``meshparams``)
the other kernels read.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PATH = os.path.join(HERE, "_ref", "libgospl_ref.so")

_lib = None


def available() -> bool:
    return os.path.exists(REF_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libgospl_ref.so is not built (make -C oracle ref)")
        _lib = ctypes.CDLL(REF_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def fv_setup(coords, cells_nodes, cells_edges, edges_nodes, circumcenters, curved):
    """setcurvedmesh + definetin (functions.F90:2715).  Returns the module's
    (ngb (N,12), nnb, elgt (N,12), vdist (N,12), area, narea)."""
    L = lib()
    nb, n, m = len(coords), len(cells_nodes), len(edges_nodes)
    c = np.asfortranarray(coords, dtype=np.float64)
    cn = np.asfortranarray(cells_nodes, dtype=np.int32)
    ce = np.asfortranarray(cells_edges, dtype=np.int32)
    en = np.asfortranarray(edges_nodes, dtype=np.int32)
    cc = np.ascontiguousarray(circumcenters, dtype=np.float64)  # (n,3) C == (3,n) F
    ngb = np.zeros((nb, 12), dtype=np.int32, order="F")
    narea = np.zeros(nb)
    L.setcurvedmesh_(_i(curved))
    L.definetin_(_p(c), _p(cn), _p(ce), _p(en), _p(cc), _p(ngb), _p(narea), _i(n), _i(nb), _i(m))
    nnb = np.zeros(nb, np.int32)
    elgt = np.zeros((nb, 12), order="F")
    vdist = np.zeros((nb, 12), order="F")
    area = np.zeros(nb)
    L.ref_fv_export_(_i(nb), _p(nnb), _p(elgt), _p(vdist), _p(area))
    return (np.ascontiguousarray(ngb), nnb, np.ascontiguousarray(elgt),
            np.ascontiguousarray(vdist), area, narea)


def mfdreceivers(nrcv, exp, elev, sl, gid):
    """functions.F90:1544 — returns (rcv, dist, wgt) as (N, nrcv) arrays."""
    L = lib()
    nb = len(elev)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    g = np.ascontiguousarray(gid, dtype=np.int32)
    rcv = np.zeros((nb, nrcv), np.int32, order="F")
    dist = np.zeros((nb, nrcv), order="F")
    wgt = np.zeros((nb, nrcv), order="F")
    L.mfdreceivers_(_i(nrcv), _d(exp), _p(e), _d(sl), _p(g), _p(rcv), _p(dist), _p(wgt), _i(nb))
    return np.ascontiguousarray(rcv), np.ascontiguousarray(dist), np.ascontiguousarray(wgt)


def sethillslopecoeff(kd):
    """functions.F90:901 — returns dcoeff (N, 13)."""
    L = lib()
    nb = len(kd)
    k = np.ascontiguousarray(kd, dtype=np.float64)
    out = np.zeros((nb, 13), order="F")
    L.sethillslopecoeff_(_i(nb), _p(k), _p(out))
    return np.ascontiguousarray(out)


def edge_tile(lvl, border, elev):
    """functions.F90:2036 — ledge (N,)."""
    L = lib()
    nb = len(elev)
    b = np.asfortranarray(border, dtype=np.int32)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    out = np.zeros(nb, np.int32)
    L.edge_tile_(_d(lvl), _p(b), _p(e), _p(out), _i(nb))
    return out


def fill_tile(edge, elev, inids):
    """functions.F90:2082 — (fillz, labels, graphnb)."""
    L = lib()
    nb = len(elev)
    m = len(edge)
    ed = np.asfortranarray(edge, dtype=np.int32)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    ii = np.ascontiguousarray(inids, dtype=np.int32)
    fz = np.zeros(nb)
    lab = np.zeros(nb, np.int32)
    gnb = ctypes.c_int(0)
    L.fill_tile_(_p(ed), _p(e), _p(ii), _p(fz), _p(lab), ctypes.byref(gnb), _i(m), _i(nb))
    return fz, lab, gnb.value


def graph_nodes(graphnb):
    """functions.F90:2394 — (graphnb, 4) graph rows; empties the module graph."""
    L = lib()
    out = np.zeros((max(graphnb, 1), 4), order="F")
    L.graph_nodes_(_i(graphnb), _p(out))
    return np.ascontiguousarray(out[:graphnb])


def combine_edges(elev, labels, ins, outs):
    """functions.F90:2331 — (newgraph (n*12,4), graphnb)."""
    L = lib()
    m = len(elev)
    n = len(ins)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    i_ = np.ascontiguousarray(ins, dtype=np.int32)
    o_ = np.ascontiguousarray(outs, dtype=np.int32)
    out = np.zeros((max(n, 1) * 12, 4), order="F")
    gnb = ctypes.c_int(0)
    L.combine_edges_(_p(e), _p(lab), _p(i_), _p(o_), _p(out), ctypes.byref(gnb), _i(m), _i(n))
    return np.ascontiguousarray(out), gnb.value


def fill_edges(nb, cgraph, maxnghbs):
    """functions.F90:2180 — (nelev, spillrank, spillnodes, spillid)."""
    L = lib()
    m = len(cgraph)
    g = np.asfortranarray(cgraph, dtype=np.float64)
    nelev = np.zeros(nb)
    srank = np.zeros(nb, np.int32)
    snodes = np.zeros(nb, np.int32)
    sid = np.zeros(nb, np.int32)
    L.fill_edges_(_i(nb), _p(g), _i(maxnghbs), _p(nelev), _p(srank), _p(snodes), _p(sid), _i(m))
    return nelev, srank, snodes, sid


def fill_depressions(lvl, dem, fillp, wsh, ggraph):
    """functions.F90:2290 — elev (N,)."""
    L = lib()
    m = len(dem)
    nb = len(ggraph)
    d = np.ascontiguousarray(dem, dtype=np.float64)
    f = np.ascontiguousarray(fillp, dtype=np.float64)
    w = np.ascontiguousarray(wsh, dtype=np.int32)
    g = np.ascontiguousarray(ggraph, dtype=np.float64)
    out = np.zeros(m)
    L.fill_depressions_(_d(lvl), _p(d), _p(f), _p(w), _p(g), _p(out), _i(m), _i(nb))
    return out


def label_pits(lvl, fill):
    """functions.F90:2420 — labels (N,)."""
    L = lib()
    nb = len(fill)
    f = np.ascontiguousarray(fill, dtype=np.float64)
    out = np.zeros(nb, np.int32)
    L.label_pits_(_d(lvl), _p(f), _p(out), _i(nb))
    return out


def fill_dir(spill, pitids, h):
    """functions.F90:1845 — ptdir (N,)."""
    L = lib()
    n = len(h)
    s = np.ascontiguousarray(spill, dtype=np.int32)
    p = np.ascontiguousarray(pitids, dtype=np.int32)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros(n, np.int32)
    L.fill_dir_(_p(s), _p(p), _p(hh), _p(out), _i(len(s)), _i(n))
    return out


def nghb_dir(pitids, h, ptdir):
    """functions.F90:1901 — relaxed ptdir (N,)."""
    L = lib()
    n = len(h)
    p = np.ascontiguousarray(pitids, dtype=np.int32)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    pd = np.ascontiguousarray(ptdir, dtype=np.int32)
    out = np.zeros(n, np.int32)
    L.nghb_dir_(_p(p), _p(hh), _p(pd), _p(out), _i(n))
    return out


def getpitvol(hlvl, elev, pit, ids):
    """functions.F90:1971 — vol (npit, 4)."""
    L = lib()
    m = len(elev)
    n = len(hlvl)
    hv = np.asfortranarray(hlvl, dtype=np.float64)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    p = np.ascontiguousarray(pit, dtype=np.int32)
    i_ = np.ascontiguousarray(ids, dtype=np.int32)
    out = np.zeros((n, 4), order="F")
    L.getpitvol_(_p(hv), _p(e), _p(p), _p(i_), _p(out), _i(m), _i(n))
    return np.ascontiguousarray(out)


def jacobiancoeff(h, kd, kp):
    """functions.F90:1421 — dcoeff (N, 13): diagonal then one column per
    neighbour slot."""
    L = lib()
    nb = len(h)
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (h, kd, kp)]
    out = np.zeros((nb, 13), order="F")
    L.jacobiancoeff_(_p(a[0]), _p(a[1]), _p(a[2]), _p(out), _i(nb))
    return np.ascontiguousarray(out)


def fctcoeff(h, kd):
    """functions.F90:1382 — the non-linear diffusion residual term (N,)."""
    L = lib()
    nb = len(h)
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (h, kd)]
    out = np.zeros(nb)
    L.fctcoeff_(_p(a[0]), _p(a[1]), _p(out), _i(nb))
    return out


def globalngbhs(nt, cells):
    """functions.F90:2539 — the global neighbour table epsfill reads
    (cells (n, 3), 0-based)."""
    L = lib()
    c = np.asfortranarray(cells, dtype=np.int32)
    L.globalngbhs_(_i(nt), _p(c), _i(len(c)))


def epsfill(sl, elev):
    """functions.F90:2633 — the priority flood with epsilon increments
    (globalngbhs must have run on the same mesh)."""
    L = lib()
    nb = len(elev)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    out = np.zeros(nb)
    L.epsfill_(_d(sl), _p(e), _p(out), _i(nb))
    return out


def mfdrcvrs(nrcv, exp, elev, sl, gid):
    """functions.F90:1673 — marine receivers: (rcv, dist, wgt) as (N, 12)."""
    L = lib()
    nb = len(elev)
    e = np.ascontiguousarray(elev, dtype=np.float64)
    g = np.ascontiguousarray(gid, dtype=np.int32)
    rcv = np.zeros((nb, 12), dtype=np.int32, order="F")
    dist = np.zeros((nb, 12), order="F")
    wgt = np.zeros((nb, 12), order="F")
    L.mfdrcvrs_(_i(nrcv), _d(exp), _p(e), _d(sl), _p(g), _p(rcv), _p(dist), _p(wgt), _i(nb))
    return np.ascontiguousarray(rcv), np.ascontiguousarray(dist), np.ascontiguousarray(wgt)
