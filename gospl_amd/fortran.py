"""Drop-in for the hot-path subset of goSPL's f2py module ``gospl._fortran``.

Same names, argument meaning and output shapes as ``fortran/functions.pyf``,
so goSPL's mixins can import these in place of the Fortran ones:

* ``setcurvedmesh(flag)``                           functions.pyf:44-46
* ``definetin(coords, cells_nodes, cells_edges,
  edges_nodes, circumcenter) -> (ngbid, narea)``    functions.pyf:297-308
* ``updatearea(narea)``                             functions.F90:2699-2713
* ``mfdreceivers(nrcv, exp, elev, sl, gid)
  -> (rcv, dist, wgt)``                             functions.pyf:146-156
* ``sethillslopecoeff(nb, kd) -> dcoeff``           functions.pyf:10-14

Like the Fortran module, it keeps the mesh of the last ``definetin`` call as
module state, in the caller's vertex order.  ``definetin`` is host code
(``gospl_fv_geometry``, bit-identical to the Fortran); the other kernels run
on the GPU through ``libgospl_hip.so``, with numpy arrays in and out as f2py
has them (so each call crosses PCIe: ``gospl_amd.model.Model`` keeps the
whole step on the device instead).  The fill and depression kernels
(``edge_tile`` .. ``getpitvol``) have no one-to-one counterpart: their chain
is replaced as a whole by ``gospl_fill_depressions`` + ``gospl_pit_info``
(``Model.fillElevation``).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .mesh import MAXNB

_state = {"curved": -1}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def setcurvedmesh(flag):
    """1 = sphere (great-circle geometry), 0 = plane (functions.F90:1000-1017)."""
    _state["curved"] = int(flag)


def definetin(coords, cells_nodes, cells_edges, edges_nodes, circumcenter):
    """FV geometry of a triangulation (functions.F90:2715-2940).

    ``circumcenter`` is (3, n) as goSPL passes ``cc.T``.  Returns
    ``ngbid`` (nb, 12) int32 (-1 padded, definetin's neighbour order) and
    ``narea`` (nb) with the Fortran's NaNs for vertices without a complete
    Voronoi cell.  Edge lengths and Voronoi face lengths stay in the module
    state for the kernels, as in the Fortran module.
    """
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    cn = np.ascontiguousarray(cells_nodes, dtype=np.int32)
    ce = np.ascontiguousarray(cells_edges, dtype=np.int32)
    en = np.ascontiguousarray(edges_nodes, dtype=np.int32)
    cc = np.ascontiguousarray(np.asarray(circumcenter, dtype=np.float64).T)
    nb, n, m = len(coords), len(cn), len(en)
    curved = _state["curved"]
    if curved < 0:  # the Fortran's own inference when the flag was never set
        curved = 1 if coords[0, 2] != 0.0 else 0
    ngb = np.empty((nb, MAXNB), np.int32)
    nnb = np.empty(nb, np.int32)
    elgt = np.empty((nb, MAXNB))
    vdist = np.empty((nb, MAXNB))
    area = np.empty(nb)
    lib = _lib.load()
    _lib.check(lib.gospl_fv_geometry(_p(coords), nb, _p(cn), _p(ce), n, _p(en), m, _p(cc), curved,
                                     _p(ngb), _p(nnb), _p(elgt), _p(vdist), _p(area)),
               "gospl_fv_geometry")
    close()
    _state.update(N=nb, ngb=ngb, nnb=nnb, elgt=elgt, vdist=vdist, area=area.copy(),
                  flat=not curved)
    return ngb.copy(), area


def updatearea(narea):
    """Replace the Voronoi areas the kernels use (functions.F90:2699-2713)."""
    a = np.ascontiguousarray(narea, dtype=np.float64)
    if len(a) != _state.get("N", -1):
        raise ValueError("updatearea: %d areas for a %s-vertex mesh" % (len(a), _state.get("N")))
    _state["area"] = a.copy()
    dev = _state.get("dev")
    if dev is not None:
        dev["area"].copy_(_torch().as_tensor(a))
        _lib.check(_lib.load().gospl_mesh_touch(dev["handle"]), "gospl_mesh_touch")


def _torch():
    import torch

    return torch


def _device():
    """Device copy of the module mesh and its library handle (created on the
    first kernel call, so ``definetin`` itself needs no GPU)."""
    if "N" not in _state:
        raise RuntimeError("definetin has not been called")
    dev = _state.get("dev")
    if dev is not None:
        return dev
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("gospl_amd.fortran kernels need a ROCm GPU")
    N = _state["N"]
    nnb = _state["nnb"]
    W = max(int(nnb.max()), 1)
    d = torch.device("cuda")
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=d)
    area = _state["area"]  # NaN areas give zero operator rows, as in the Fortran
    dev = dict(nbr=t(_state["ngb"][:, :W].T, torch.int32), deg=t(nnb, torch.int32),
               elgt=t(_state["elgt"][:, :W].T, torch.float64),
               vdist=t(_state["vdist"][:, :W].T, torch.float64), area=t(area, torch.float64),
               gid=t(np.arange(N), torch.int32), outlet=torch.zeros(N, dtype=torch.int8, device=d))
    P = lambda x: ctypes.c_void_p(x.data_ptr())
    h = ctypes.c_void_p()
    lib = _lib.load()
    _lib.check(lib.gospl_mesh_create(N, W, 1 if _state["flat"] else 0, P(dev["nbr"]), P(dev["deg"]),
                                     P(dev["elgt"]), P(dev["vdist"]), P(dev["area"]), P(dev["gid"]),
                                     P(dev["outlet"]), ctypes.byref(h)), "gospl_mesh_create")
    dev["handle"] = h
    dev["P"] = P
    _state["dev"] = dev
    return dev


def close():
    """Release the device mesh of the module state (if any)."""
    dev = _state.pop("dev", None)
    if dev is not None and dev["handle"].value:
        _lib.load().gospl_mesh_destroy(dev["handle"])


def mfdreceivers(nrcv, exp, elev, sl, gid):
    """Multiple-flow-direction receivers (functions.F90:1544-1671).

    Returns ``rcv`` (nb, nrcv) int32, ``dist`` and ``wgt`` (nb, nrcv) f64 in
    Fortran order, as f2py does.  ``gid`` is the global id used in the slope
    perturbation (the local index on one rank).
    """
    torch = _torch()
    dev = _device()
    N = _state["N"]
    P = dev["P"]
    elev = np.ascontiguousarray(elev, dtype=np.float64).ravel()
    gid = np.ascontiguousarray(gid, dtype=np.int32).ravel()
    # the kernel reads N entries of each: refuse shorter arrays before any launch
    if len(elev) != N or len(gid) != N:
        raise ValueError("mfdreceivers: len(elev)=%d, len(gid)=%d for a %d-vertex mesh" % (len(elev), len(gid), N))
    if not 1 <= int(nrcv) <= 12:
        raise ValueError("mfdreceivers: nrcv=%d outside 1..12" % int(nrcv))
    e = torch.as_tensor(elev, device="cuda")
    dev["gid"].copy_(torch.as_tensor(gid))
    rcv = torch.empty((nrcv, N), dtype=torch.int32, device="cuda")
    dist = torch.empty((nrcv, N), dtype=torch.float64, device="cuda")
    wgt = torch.empty_like(dist)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.load().gospl_mfdreceivers(dev["handle"], int(nrcv), float(exp), P(e), float(sl), P(rcv),
                                              P(dist), P(wgt), s), "gospl_mfdreceivers")
    return (np.asfortranarray(rcv.T.cpu().numpy()), np.asfortranarray(dist.T.cpu().numpy()),
            np.asfortranarray(wgt.T.cpu().numpy()))


def sethillslopecoeff(nb, kd):
    """Hillslope operator rows (functions.F90:901-938): (nb, 13), column 0
    the diagonal, column 1 + p the coefficient of neighbour p."""
    torch = _torch()
    dev = _device()
    N = _state["N"]
    kd = np.ascontiguousarray(kd, dtype=np.float64).ravel()
    if nb != N or len(kd) != N:
        raise ValueError("sethillslopecoeff: nb=%d, len(kd)=%d for a %d-vertex mesh" % (nb, len(kd), N))
    k = torch.as_tensor(kd, device="cuda")
    out = torch.empty((13, N), dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.load().gospl_hillslope_coeff(dev["handle"], dev["P"](k), dev["P"](out), s),
               "gospl_hillslope_coeff")
    return np.asfortranarray(out.T.cpu().numpy())
