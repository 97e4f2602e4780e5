"""Mesh construction for the surface-process hot path (init only, host side).

The hot-path kernels work on a finite-volume (Voronoi) dual of a triangulated
surface: per vertex, an ordered neighbour list, the edge lengths to those
neighbours, the Voronoi face lengths and the Voronoi cell area.  goSPL builds
these in ``UnstMesh._meshStructure`` (gospl/mesher/unstructuredmesh.py:120-202)
from ``VoroBuild`` (gospl/mesher/meshfunc.py:75-408) plus the Fortran
``definetin`` (fortran/functions.F90:2715-2940).

This module restates that pipeline:

* the numpy half (edge list, cell->edge map, circumcentres, control volumes) is
  written with the same numpy operations as ``meshfunc.py`` so the edge order
  (numpy ``unique`` over raw int32 pairs, i.e. byte order) and the circumcentres
  come out identical;
* the Fortran half runs in the C-ABI library (``gospl_fv_geometry``).

It also provides the synthetic meshes the benchmark configurations are quoted
on (icosahedral spheres, jittered planar Delaunay grids) and the space-filling
curve vertex order used on the device.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

MAXNB = 12  # neighbour cap of the FV stencil (fortran/functions.F90:2733)


# ----------------------------------------------------------------------------
# Synthetic meshes
# ----------------------------------------------------------------------------

def icosphere(level: int, radius: float = 6.371e6):
    """Icosahedral sphere refined ``level`` times (10*4**level + 2 vertices).

    Returns ``(coords (N,3) float64, cells (T,3) int32)``.  Each refinement
    splits a triangle into four; new vertices are edge midpoints pushed back
    onto the sphere.
    """
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        ukey, inv = np.unique(key, return_inverse=True)
        a, b = ukey // nv, ukey % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        v = np.concatenate([v, mid])
        nf = len(f)
        m01 = nv + inv[:nf]
        m12 = nv + inv[nf:2 * nf]
        m20 = nv + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1),
            np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1),
            np.stack([m01, m12, m20], 1)])
    return v * radius, f.astype(np.int32)


def planar_jittered(nx: int = 200, ny: int = 200, dx: float = 500.0,
                    jitter: float = 100.0, seed: int = 0):
    """Jittered ``nx``x``ny`` grid (spacing ``dx``) Delaunay-triangulated.

    Interior points move by U(-jitter, +jitter) in x and y (seeded); boundary
    points stay on the domain edges so goSPL's boundary detection finds them.
    Returns ``(coords (N,3) with z=0, cells (T,3) int32)``.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    xs = np.arange(nx, dtype=np.float64) * dx
    ys = np.arange(ny, dtype=np.float64) * dx
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    X = X.ravel()
    Y = Y.ravel()
    jx = rng.uniform(-jitter, jitter, X.shape)
    jy = rng.uniform(-jitter, jitter, Y.shape)
    inner = (X > 0) & (X < xs[-1]) & (Y > 0) & (Y < ys[-1])
    X = np.where(inner, X + jx, X)
    Y = np.where(inner, Y + jy, Y)
    pts = np.stack([X, Y], 1)
    tri = Delaunay(pts)
    coords = np.zeros((len(pts), 3))
    coords[:, :2] = pts
    return coords, tri.simplices.astype(np.int32)


# ----------------------------------------------------------------------------
# numpy half of the FV construction (restates gospl/mesher/meshfunc.py)
# ----------------------------------------------------------------------------

def _unique_rows(a):
    # Same construction as meshfunc.unique_rows (meshfunc.py:20-28): unique over
    # a raw-bytes view, which orders rows by their memory bytes.
    b = np.ascontiguousarray(a).view(np.dtype((np.void, a.dtype.itemsize * a.shape[1])))
    a_unique, inv, cts = np.unique(b, return_inverse=True, return_counts=True)
    a_unique = a_unique.view(a.dtype).reshape(-1, a.shape[1])
    return a_unique, inv.reshape(-1), cts


def voronoi_inputs(coords: np.ndarray, cells: np.ndarray, flat: bool):
    """Edge list, cell->edge map, circumcentres and control volumes.

    Follows ``VoroBuild.initVoronoi`` / ``create_edges`` / ``cell_circumcenters``
    / ``control_volumes`` (meshfunc.py:75-154, 176-198, 295-328, 401-408) and the
    sphere re-projection of the circumcentres (unstructuredmesh.py:145-148).
    ``cells`` are taken as int32 (PETSc's default index type), which fixes the
    byte order that sorts the edges.
    """
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    local_idx = np.array([[1, 2], [2, 0], [0, 1]]).T
    nds = cells.T
    idx_h = nds[local_idx]
    half = coords[idx_h[1]] - coords[idx_h[0]]
    ei_dot_ej = np.einsum("ijk, ijk->ij", half[[1, 2, 0]], half[[2, 0, 1]])
    ei_dot_ei = np.einsum("ijk, ijk->ij", half, half)
    cell_volumes = 0.5 * np.sqrt(
        ei_dot_ej[2] * ei_dot_ej[0] + ei_dot_ej[0] * ei_dot_ej[1] + ei_dot_ej[1] * ei_dot_ej[2])
    ce_ratios = -ei_dot_ej * 0.25 / cell_volumes[None]

    # control volumes
    part = 0.25 * ei_dot_ei * ce_ratios
    vals = np.array([sum([part[i] for i in np.where(local_idx.T == k)[0]]) for k in range(3)])
    cvol = np.zeros(len(coords))
    np.add.at(cvol, nds, vals)

    # edges
    s = idx_h.shape
    a = np.sort(idx_h.reshape(s[0], -1).T)
    edges, inv, cts = _unique_rows(a)
    if not np.all(cts < 3):
        raise ValueError("an edge is shared by more than two cells")
    cells_edges = inv.reshape(3, -1).T
    boundary_edge = cts == 1

    # circumcentres
    alpha = ei_dot_ei * ei_dot_ej
    alpha_sum = alpha[0] + alpha[1] + alpha[2]
    beta = alpha / alpha_sum[None]
    cc = coords[nds] * beta[..., None]
    cc = cc[0] + cc[1] + cc[2]
    if not flat:
        radius = np.linalg.norm(coords[0])
        cc = cc * (radius / np.linalg.norm(cc, axis=1)).reshape((len(cc), 1))

    return dict(edges=np.ascontiguousarray(edges, dtype=np.int32),
                cells_edges=np.ascontiguousarray(cells_edges, dtype=np.int32),
                circumcenters=np.ascontiguousarray(cc),
                control_volumes=cvol,
                boundary_edge=boundary_edge)


def fv_geometry(coords, cells, flat, lib):
    """Full FV geometry in input vertex order.

    Returns ``(ngb (N,12) int32 with -1 padding, nnb (N,), elgt (N,12),
    vdist (N,12), area (N,), vinfo)`` as goSPL's ``definetin`` + area fix-ups
    produce them (unstructuredmesh.py:150-188).
    """
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    vi = voronoi_inputs(coords, cells, flat)
    N, T, M = len(coords), len(cells), len(vi["edges"])
    ngb = np.empty((N, MAXNB), np.int32)
    nnb = np.empty(N, np.int32)
    elgt = np.empty((N, MAXNB))
    vdist = np.empty((N, MAXNB))
    area = np.empty(N)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    rc = lib.gospl_fv_geometry(p(coords), N, p(cells), p(vi["cells_edges"]), T,
                               p(vi["edges"]), M, p(vi["circumcenters"]),
                               0 if flat else 1, p(ngb), p(nnb), p(elgt), p(vdist), p(area))
    if rc != 0:
        raise RuntimeError("gospl_fv_geometry failed (%d)" % rc)
    larea = np.abs(vi["control_volumes"])
    larea[np.isnan(larea)] = 1.0
    area[np.isnan(area)] = 1.0
    if flat and np.max(np.abs(larea - area)) > 1.0e2:
        # unstructuredmesh.py:172-188: fall back to the meshfunc control volumes
        area = larea.copy()
    return ngb, nnb, elgt, vdist, area, vi


def boundary_vertices(cells: np.ndarray, boundary_edge_mask=None, edges=None):
    """Vertices on boundary edges (edges that belong to a single triangle)."""
    if edges is None or boundary_edge_mask is None:
        c = np.asarray(cells, dtype=np.int64)
        e = np.concatenate([c[:, [0, 1]], c[:, [1, 2]], c[:, [2, 0]]])
        e.sort(axis=1)
        u, cnt = np.unique(e, axis=0, return_counts=True)
        be = u[cnt == 1]
    else:
        be = edges[boundary_edge_mask]
    return np.unique(be.ravel()).astype(np.int64)


# ----------------------------------------------------------------------------
# Device-order mesh
# ----------------------------------------------------------------------------

def morton_order(coords: np.ndarray) -> np.ndarray:
    """Permutation sorting vertices along a 3-D Morton (Z-order) curve.

    Consecutive device indices are then spatial neighbours, so the neighbour
    gathers of a wavefront hit few cache lines.
    """
    lo = coords.min(axis=0)
    span = np.maximum(coords.max(axis=0) - lo, 1e-30)
    q = np.clip(((coords - lo) / span * (2 ** 21 - 1)).astype(np.uint64), 0, 2 ** 21 - 1)

    def spread(x):
        x = x & np.uint64(0x1FFFFF)
        x = (x | (x << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        x = (x | (x << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        x = (x | (x << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        x = (x | (x << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        x = (x | (x << np.uint64(2))) & np.uint64(0x1249249249249249)
        return x

    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


@dataclass
class Mesh:
    """Mesh in device (space-filling-curve) order.

    ``perm[d]`` is the input-mesh id of device vertex ``d`` (goSPL's ``locIDs``
    / ``gid``); ``iperm`` is its inverse.  Neighbour tables are ELL,
    column-major: entry ``k`` of vertex ``d`` sits at ``[k, d]``; unused slots
    hold -1 (indices) and 0 (lengths).  Neighbour order within a row is
    goSPL's (``definetin`` edge order).
    """
    N: int
    width: int
    flat: bool
    coords: np.ndarray          # (N,3) device order
    nbr: np.ndarray             # (width, N) int32 device indices
    deg: np.ndarray             # (N,) int32
    elgt: np.ndarray            # (width, N) float64
    vdist: np.ndarray           # (width, N) float64
    area: np.ndarray            # (N,) float64
    gid: np.ndarray             # (N,) int32 input ids
    perm: np.ndarray            # device -> input
    iperm: np.ndarray           # input -> device
    outlet: np.ndarray          # (N,) int8 1 = draining border (flat meshes)
    maxarea: float = 0.0
    extras: dict = field(default_factory=dict)

    def to_device_order(self, a):
        return np.asarray(a)[self.perm]

    def to_input_order(self, a):
        a = np.asarray(a)
        out = np.empty_like(a)
        out[self.perm] = a
        return out


def build_mesh(coords, cells, lib, flat=None, order="morton", outlet_mask=None) -> Mesh:
    """FV geometry + boundary + device order for a triangulation.

    ``flat`` defaults to "the mesh has boundary edges" (goSPL sets
    ``flatModel`` when DMPlex reports boundary vertices,
    unstructuredmesh.py:468-472).  On flat meshes every boundary vertex is a
    draining outlet unless ``outlet_mask`` says otherwise (``bc`` walls).
    """
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    N = len(coords)
    # a closed sphere has no boundary edges: skip the (large) edge count
    bverts = boundary_vertices(cells) if flat is not False else np.zeros(0, np.int64)
    if flat is None:
        flat = len(bverts) > 0
    ngb, nnb, elgt, vdist, area, vi = fv_geometry(coords, cells, flat, lib)
    outlet = np.zeros(N, np.int8)
    if flat:
        outlet[bverts] = 1
        if outlet_mask is not None:
            outlet &= np.asarray(outlet_mask, np.int8)

    if order == "morton":
        perm = morton_order(coords)
    else:
        perm = np.arange(N)
    iperm = np.empty(N, np.int64)
    iperm[perm] = np.arange(N)
    width = int(nnb.max())
    ngb_d = ngb[perm, :width]
    nbr = np.where(ngb_d >= 0, iperm[np.maximum(ngb_d, 0)], -1).astype(np.int32)
    return Mesh(
        N=N, width=width, flat=bool(flat),
        coords=coords[perm],
        nbr=np.ascontiguousarray(nbr.T),
        deg=nnb[perm].astype(np.int32),
        elgt=np.ascontiguousarray(elgt[perm, :width].T),
        vdist=np.ascontiguousarray(vdist[perm, :width].T),
        area=area[perm].copy(),
        gid=perm.astype(np.int32),
        perm=perm, iperm=iperm,
        outlet=outlet[perm].copy(),
        maxarea=float(area.max()),
        extras=dict(ngb_input=ngb, nnb_input=nnb, elgt_input=elgt, vdist_input=vdist,
                    area_input=area, cells=cells),
    )
