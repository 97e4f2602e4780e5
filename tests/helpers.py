"""Shared builders for the parity tests: the same seeded workload built once
for the HIP path (device order) and once for the oracle (input order)."""

from __future__ import annotations

import functools

import numpy as np

from gospl_amd import _lib, mesh as gmesh, workloads
from oracle import oracle as orc


@functools.lru_cache(maxsize=None)
def case(kind: str, size: int, seed: int = 0):
    lib = _lib.load()
    if kind == "planar":
        m, h0, p = workloads.planar_case(lib, n=size, seed=seed)
    elif kind == "sphere":
        m, h0, p = workloads.sphere_case(lib, level=size, seed=seed, lmax=min(64, 4 * 2 ** size))
    elif kind == "fixture":
        d = np.load(size)
        flat = not np.all(d["v"][:, 2] != 0) and np.allclose(d["v"][:, 2], 0)
        m = gmesh.build_mesh(d["v"], d["c"], lib, flat=flat)
        h0 = m.to_device_order(d["z"])
        p = dict(K=4e-6, m=0.5, d=0.0, flowdir=3, cda=0.01, cdm=0.2, dt=100.0, rain=1.0,
                 upsub=None, sealevel=-100.0)
    else:
        raise ValueError(kind)
    return m, h0, p


def oracle_model(m, h0, p, **over):
    q = dict(p)
    q.update(over)
    conv = {k: (m.to_input_order(v) if isinstance(v, np.ndarray) else v) for k, v in q.items()}
    return orc.OracleModel(workloads.oracle_geometry(m), m.to_input_order(h0), tend=1e12, **conv)


def dev_index_to_input(m, idx):
    """Map device indices (-1 kept) to input ids, then reorder rows to input order."""
    idx = np.asarray(idx)
    out = np.where(idx >= 0, m.perm[np.maximum(idx, 0)], -1)
    return out


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def make_cycle(rcv, wgt, drain=True):
    """Receivers (N, K) in input order with a two-vertex cycle added: a vertex
    A with a positive-weight receiver B that itself drains elsewhere gets B
    routing half its flow back to A (a cycle of gain < 1 that drains), or,
    with drain=False, A and B sending everything to each other (a cycle that
    never drains).  Returns new (rcv, wgt) and (A, B)."""
    rcv = rcv.copy()
    wgt = wgt.copy()
    N, K = rcv.shape
    for a in range(N):
        for k in range(K):
            b = rcv[a, k]
            if b == a or wgt[a, k] <= 0.0:
                continue
            out_b = [(rcv[b, j], wgt[b, j]) for j in range(K) if rcv[b, j] != b and wgt[b, j] > 0 and rcv[b, j] != a]
            if not out_b:
                continue
            if drain:
                rcv[b, :] = b
                wgt[b, :] = 0.0
                rcv[b, 0], wgt[b, 0] = a, 0.5
                rcv[b, 1], wgt[b, 1] = out_b[0][0], 0.5
            else:
                rcv[a, :] = a
                wgt[a, :] = 0.0
                rcv[b, :] = b
                wgt[b, :] = 0.0
                rcv[a, 0], wgt[a, 0] = b, 1.0
                rcv[b, 0], wgt[b, 0] = a, 1.0
            return rcv, wgt, (a, b)
    raise ValueError("no vertex pair for a cycle")


def spl_steady_case(n: int = 30, U: float = 1.0e-3, K: float = 1.0e-5):
    """goSPL's stream-power benchmark (benchmarks/test_spl.py:443-577): a
    planar mesh with draining borders, uniform uplift U, single-direction
    routing, no diffusion, run to steady state, where E = U and the
    steepness index ks = S * A^(m/n) = (U/K)^(1/n) everywhere on channels."""
    m, h0, p = case("planar", n, 3)
    q = dict(p)
    interior = m.outlet == 0
    q.update(K=K, m=0.5, d=0.0, flowdir=1, cda=0.0, cdm=0.0, dt=2.0e4, rain=1.0, sealevel=-1.0e4,
             upsub=np.where(interior, U, 0.0))
    return m, h0, q, interior


def spl_steady_metrics(h, rcv, dist, fa, interior, U, K, h_prev, dt_span, dt):
    """E/U over the last span, and on channel nodes (drainage area above 5 %
    of the largest): ks median / (U/K), its relative spread (IQR / median)
    and the fraction of nodes whose ks is within 1 % of U/K."""
    rcv = np.asarray(rcv)
    d = np.asarray(dist)
    # slopes of the eroded surface, before the step's uplift (which lifts a
    # node above a fixed border receiver by U * dt)
    he = h - np.where(interior, U * dt, 0.0)
    S = np.where(d > 0, (he - he[rcv]) / np.where(d > 0, d, 1.0), 0.0)
    E = U - (h - h_prev)[interior].mean() / dt_span
    ch = interior & (fa > 0.05 * fa[interior].max()) & (S > 1e-12)
    ks = S[ch] * fa[ch] ** 0.5
    med = float(np.median(ks))
    iqr = float(np.percentile(ks, 75) - np.percentile(ks, 25)) / med
    within = float(np.mean(np.abs(ks / (U / K) - 1.0) < 0.01))
    return E / U, med / (U / K), iqr, within, int(ch.sum())
