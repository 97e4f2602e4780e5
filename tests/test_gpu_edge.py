"""Edge cases of the hot path, HIP vs the oracle from identical state.

The surfaces here stress the discrete parts of goSPL's step: a mesh that is
entirely below sea level (no runoff, no stream power, marine diffusion
only), a flat plateau (every cell is on a flat: routing is the flat-receiver
field of _dirFlats), one closed bowl (a single depression that fills and
overflows through the cascade), integer elevations (receiver ties on every
slope) and the smallest planar mesh with interior cells.  Each case runs a
few steps; before every step the oracle takes the GPU's state, and the
outputs of the step must agree to 1e-9 relative (pit ids, filled
elevations and receivers bit for bit).
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import case, dev_index_to_input, oracle_model, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _surface(kind, m, h0):
    xy = m.coords
    if kind == "all_sea":
        rng = np.random.default_rng(3)
        return -1000.0 + rng.uniform(0.0, 50.0, m.N)
    if kind == "plateau":
        # flat interior at 10 m, borders at 0: one flat drained through the borders
        h = np.full(m.N, 10.0)
        h[m.outlet > 0] = 0.0
        return h
    if kind == "bowl":
        c = xy[:, :2].mean(axis=0)
        r = np.hypot(xy[:, 0] - c[0], xy[:, 1] - c[1])
        R = r.max()
        h = 200.0 - 150.0 * np.exp(-(r / (0.3 * R)) ** 2) + 0.5 * (r / R)
        h[m.outlet > 0] = 0.0
        return h
    if kind == "ties":
        return np.round(h0)
    raise ValueError(kind)


def _compare_steps(m, h, p, nsteps=3):
    from gospl_amd.model import Model

    g = Model(m, h, tend=1e12, **p)
    o = oracle_model(m, h, p)
    g.initialise()
    o.initialise()
    for s in range(nsteps):
        o.h = m.to_input_order(g.hLocal.cpu().numpy()).copy()
        o.cumED = m.to_input_order(g.cumED.cpu().numpy()).copy()
        o.tNow = g.tNow
        g.step()
        o.step()
        assert np.array_equal(m.to_input_order(g.pitIDs.cpu().numpy()), o.pitIDs), s
        assert np.array_equal(m.to_input_order(g.lFill.cpu().numpy()), o.lFill), s
        r_in = dev_index_to_input(m, m.to_input_order(g.rcvIDi.cpu().numpy().T))
        assert np.array_equal(r_in, o.rcvIDi), s
        assert g.cascade_passes == o.cascade_passes, s
        h_g = m.to_input_order(g.hLocal.cpu().numpy())
        assert rel_err(h_g, o.h) < TOL, (s, rel_err(h_g, o.h))
        fal = m.to_input_order(g.FAL.cpu().numpy())
        if np.abs(o.FAL).max() > 0:
            assert rel_err(fal, o.FAL) < TOL, s
        else:
            assert np.abs(fal).max() == 0.0, s
        ed = m.to_input_order(g.cumED.cpu().numpy())
        assert np.abs(ed - o.cumED).max() <= TOL * max(np.abs(o.h).max(), 1.0), s
    return g, o


@pytest.mark.parametrize("kind", ["plateau", "bowl", "ties"])
def test_planar_surfaces(kind):
    m, h0, p = case("planar", 40)
    h = _surface(kind, m, h0)
    g, o = _compare_steps(m, h, p)
    if kind == "bowl":
        assert g.npits >= 1


def test_all_sea_sphere():
    m, h0, p = case("sphere", 4)
    h = _surface("all_sea", m, h0)
    g, o = _compare_steps(m, h, p)
    assert float(g.FAL.abs().max()) == 0.0
    assert float(g.Eb.abs().max()) == 0.0


def test_smallest_planar_mesh():
    m, h0, p = case("planar", 4)
    _compare_steps(m, h0, p, nsteps=2)


def test_sphere_ties():
    m, h0, p = case("sphere", 5)
    _compare_steps(m, np.round(h0 / 10.0) * 10.0, p)


@pytest.mark.parametrize("drain", [True, False])
def test_flow_receiver_cycle(drain):
    """Receivers with a two-vertex cycle through gospl_flowacc: a cycle that
    drains is solved by the fixed-point pass to the oracle's values; one
    that never drains is reported as unsolved."""
    import ctypes

    import torch

    from gospl_amd import _lib
    from gospl_amd.model import DeviceMesh, _ptr
    from oracle import oracle as orc
    from tests.helpers import make_cycle

    m, h0, p = case("planar", 12)
    o = oracle_model(m, h0, p)
    N = m.N
    z = m.to_input_order(h0)
    rcv, dist, wgt = orc.mfdreceivers(3, 1.1, z, -1e9, np.arange(N, dtype=np.int32), o.ngb, o.nnb, o.elgt)
    rc, wc, _ = make_cycle(rcv, wgt, drain=drain)
    b = np.random.default_rng(4).uniform(0.5, 2.0, N) * m.to_input_order(m.area)
    qo, bado = orc.flowacc(rc, wc, b, o.ngb, o.nnb)
    dm = DeviceMesh(m, "cuda")
    rdev = m.iperm[rc[m.perm]].T.astype(np.int32)  # (K, N) device indices
    wdev = wc[m.perm].T
    R = torch.as_tensor(np.ascontiguousarray(rdev), device="cuda")
    Wt = torch.as_tensor(np.ascontiguousarray(wdev), device="cuda")
    B = torch.as_tensor(m.to_device_order(b), device="cuda")
    Q = torch.zeros_like(B)
    lev = ctypes.c_int32(0)
    bad = ctypes.c_int64(0)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert _lib.load().gospl_flowacc(dm.handle, 3, _ptr(R), _ptr(Wt), _ptr(B), _ptr(Q), ctypes.byref(lev),
                                     ctypes.byref(bad), s) == 0
    torch.cuda.synchronize()
    if drain:
        assert bado == 0 and bad.value == 0
        q = m.to_input_order(Q.cpu().numpy())
        assert rel_err(q, qo) < 1e-13
    else:
        assert bado > 0 and bad.value == bado


@pytest.mark.parametrize("chain", ["1", "0"])
def test_single_long_chain_closed_form(chain):
    """Flow accumulation along one 40,000-vertex path (a snake over the
    200x200 planar grid, unit weights, unit sources): the discharge at the
    k-th vertex of the path is exactly k, for the dataflow and the level
    sweeps (the deepest dependency chain the sweeps can meet)."""
    import ctypes
    import os

    import torch

    from gospl_amd import _lib, workloads
    from gospl_amd.model import DeviceMesh, _ptr

    n = 200
    lib = _lib.load()
    m, h0, p = workloads.planar_case(lib, n=n)
    N = m.N
    order = np.concatenate([np.arange(i * n, (i + 1) * n)[:: 1 if i % 2 == 0 else -1] for i in range(n)])
    nxt = np.empty(N, np.int64)
    nxt[order[:-1]] = order[1:]
    nxt[order[-1]] = order[-1]
    rcv = m.iperm[nxt[m.perm]].astype(np.int32)  # device order
    linked = np.zeros(N, bool)
    for k in range(m.nbr.shape[0]):
        linked |= m.nbr[k] == rcv
    assert linked.sum() == N - 1  # every link is a mesh edge (the last vertex is its own receiver)
    os.environ["GOSPL_DAG_CHAIN"] = chain
    try:
        dm = DeviceMesh(m, "cuda")
    finally:
        del os.environ["GOSPL_DAG_CHAIN"]
    R = torch.as_tensor(rcv[None, :].copy(), device="cuda")
    Wt = torch.ones((1, N), dtype=torch.float64, device="cuda")
    B = torch.ones(N, dtype=torch.float64, device="cuda")
    Q = torch.zeros_like(B)
    lev = ctypes.c_int32(0)
    bad = ctypes.c_int64(0)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.gospl_flowacc(dm.handle, 1, _ptr(R), _ptr(Wt), _ptr(B), _ptr(Q), ctypes.byref(lev),
                             ctypes.byref(bad), s) == 0
    q = m.to_input_order(Q.cpu().numpy())
    assert bad.value == 0
    assert np.array_equal(q[order], np.arange(1, N + 1, dtype=np.float64))
    if chain == "0":
        assert abs(lev.value - N) <= 1  # the depth of the path
    dm.close()


def test_csr_rows_match_packed_records():
    """The general dependency-row layout (any receiver count) and the packed
    per-vertex records (up to three receivers) give the same bits."""
    import os

    from gospl_amd.model import Model

    m, h0, p = case("planar", 60)
    out = []
    for csr in ("0", "1"):
        os.environ["GOSPL_DAG_CSR"] = csr
        try:
            g = Model(m, h0, tend=1e12, **p)
        finally:
            del os.environ["GOSPL_DAG_CSR"]
        g.initialise()
        for _ in range(3):
            g.step()
        out.append([t.cpu().numpy().copy() for t in (g.hLocal, g.FAL, g.cumED, g.fillFAL)])
        g.dm.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("level", [6, 8])
def test_dataflow_sweep_matches_level_sweep(level):
    """The dataflow sweeps (the default for flow accumulation and stream
    power) give the bits of the level-by-level sweeps: with the default
    hand-off (a batch goes to the global queue as soon as a pool holds one),
    with pools that keep 128 vertices, and with small pools."""
    import os

    from gospl_amd.model import Model

    m, h0, p = case("sphere", level)
    out = []
    for chain, pool, share in (("0", "512", "0"), ("1", "512", "0"), ("1", "512", "128"), ("1", "128", "64")):
        os.environ["GOSPL_DAG_CHAIN"] = chain
        os.environ["GOSPL_CHAIN_POOL"] = pool
        os.environ["GOSPL_CHAIN_SHARE"] = share
        try:
            g = Model(m, h0, tend=1e12, **p)
        finally:
            del os.environ["GOSPL_DAG_CHAIN"]
            del os.environ["GOSPL_CHAIN_POOL"]
            del os.environ["GOSPL_CHAIN_SHARE"]
        g.initialise()
        for _ in range(3):
            g.step()
        out.append([t.cpu().numpy().copy() for t in (g.hLocal, g.FAL, g.cumED, g.fillFAL)])
        g.dm.close()
    for res in out[1:]:
        for a, b in zip(out[0], res):
            assert np.array_equal(a, b)


def test_dataflow_bounded_wait_aborts_cleanly():
    """Every wait of the dataflow kernel is bounded.  With the bound forced
    to zero from the host, the sweep gives up at its first wait: every
    wavefront leaves, the call reports the abort as an error, and a sweep on
    a normal handle afterwards is unaffected."""
    import os

    from gospl_amd.model import Model

    m, h0, p = case("sphere", 7)
    os.environ["GOSPL_CHAIN_SPIN_MAX"] = "0"
    try:
        g = Model(m, h0, tend=1e12, **p)
    finally:
        del os.environ["GOSPL_CHAIN_SPIN_MAX"]
    with pytest.raises(RuntimeError, match="dataflow sweep aborted"):
        g.initialise()
    g.dm.close()
    ref = []
    for chain in ("1", "0"):
        os.environ["GOSPL_DAG_CHAIN"] = chain
        try:
            g = Model(m, h0, tend=1e12, **p)
        finally:
            del os.environ["GOSPL_DAG_CHAIN"]
        g.initialise()
        g.step()
        ref.append([t.cpu().numpy().copy() for t in (g.hLocal, g.FAL)])
        g.dm.close()
    for a, b in zip(*ref):
        assert np.array_equal(a, b)


def test_mesh_with_one_sided_neighbour_is_refused():
    """The sweeps rely on mutual neighbour lists; a handle on lists where one
    relation is one-sided is refused with an error instead of giving wrong
    solves."""
    import ctypes

    import torch

    from gospl_amd import _lib
    from gospl_amd.model import _ptr

    m, h0, p = case("planar", 12)
    nbr = m.nbr.copy()
    deg = m.deg.copy()
    i = int(np.argmax(deg))
    nbr[deg[i] - 1, i] = -1  # i forgets its last neighbour; that one still lists i
    deg[i] -= 1
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    args = [t(nbr, torch.int32), t(deg, torch.int32), t(m.elgt, torch.float64), t(m.vdist, torch.float64),
            t(m.area, torch.float64), t(m.gid, torch.int32), t(m.outlet, torch.int8)]
    h = ctypes.c_void_p()
    lib = _lib.load()
    rc = lib.gospl_mesh_create(m.N, nbr.shape[0], 1, *[_ptr(a) for a in args], ctypes.byref(h))
    assert rc != 0
    assert b"not mutual" in lib.gospl_last_error()


@pytest.mark.parametrize("flowdir", [1, 6])
def test_other_receiver_counts(flowdir):
    """Single- and six-receiver routing (the latter on the general row
    layout) against the oracle, step by step."""
    m, h0, p = case("sphere", 5)
    q = dict(p)
    q["flowdir"] = flowdir
    _compare_steps(m, h0, q, nsteps=2)


def test_spl_steady_state_analytic_gpu():
    """The HIP path reaches goSPL's analytic stream-power steady state
    (benchmarks/test_spl.py:443-577): E = U, median ks = U/K within 0.5 %,
    ks spread below 10 %, on a 60x60 plane."""
    from gospl_amd.model import Model
    from tests.helpers import spl_steady_case, spl_steady_metrics

    U, K = 1.0e-3, 1.0e-5
    m, h0, q, interior = spl_steady_case(60, U, K)
    g = Model(m, h0, tend=1e12, **q)
    g.initialise()
    for _ in range(400):
        g.step()
    prev = g.hLocal.cpu().numpy().copy()
    for _ in range(50):
        g.step()
    h = g.hLocal.cpu().numpy()
    rcv = g.rcvIDi.cpu().numpy()[0]
    dist = g.distRcvi.cpu().numpy()[0]
    fa = g.FAL.cpu().numpy()
    eu, ksr, iqr, within, nch = spl_steady_metrics(h, rcv, dist, fa, interior.astype(bool), U, K, prev,
                                                   50 * q["dt"], q["dt"])
    assert nch > 100
    assert abs(eu - 1.0) < 0.05, eu
    assert abs(ksr - 1.0) < 0.005, ksr
    assert iqr < 0.10, iqr


@pytest.mark.parametrize("nspill", [0, 7])
def test_cascade_seed_matches_pit_water_and_scatter(nspill):
    """The one-call cascade update gives bit for bit what the pit-water call
    and the scatter of the spilled volumes (numpy fancy assignment, then a
    true division by dt) give, on random pit ids, levels (NaN = none) and spill points, with
    the right-hand side starting from garbage."""
    import torch

    from gospl_amd.model import DeviceMesh, _ptr

    m, h0, p = case("planar", 30)
    dm = DeviceMesh(m, torch.device("cuda:0"))
    lib = dm.lib
    N, npit, dt = m.N, 50, 1234.5
    rng = np.random.default_rng(11)
    t = lambda a, dtp: torch.as_tensor(np.ascontiguousarray(a), dtype=dtp, device="cuda")
    pit = t(rng.integers(-1, npit, N), torch.int32)
    lfill = t(rng.uniform(0.0, 100.0, N), torch.float64)
    water0 = t(rng.uniform(0.0, 100.0, N), torch.float64)
    full = (rng.random(npit) < 0.3).astype(np.float64)
    level = np.where(rng.random(npit) < 0.5, rng.uniform(0.0, 100.0, npit), np.nan)
    u = rng.choice(N, size=nspill, replace=False).astype(np.float64)
    vv = rng.uniform(1.0, 1.0e6, nspill)
    s = torch.cuda.current_stream().cuda_stream

    w1 = water0.clone()
    assert lib.gospl_pit_water(dm.handle, _ptr(pit), _ptr(t(full, torch.int8)), _ptr(t(level, torch.float64)),
                               _ptr(lfill), _ptr(w1), s) == 0
    # numpy's true division, as goSPL's `nFA / self.dt` (flowplex.py:663);
    # torch divides a tensor by a scalar as a product with its reciprocal
    nfa = np.zeros(N)
    nfa[u.astype(np.int64)] = vv
    rhs1 = nfa / dt

    w2 = water0.clone()
    rhs2 = torch.full((N,), 7.0, dtype=torch.float64, device="cuda")
    stage = t(np.concatenate([full, level, u, vv]), torch.float64)
    assert lib.gospl_cascade_seed(dm.handle, _ptr(stage), npit, nspill, _ptr(pit), _ptr(lfill), _ptr(w2), dt,
                                  _ptr(rhs2), s) == 0
    torch.cuda.synchronize()
    assert torch.equal(w1, w2)
    assert np.array_equal(rhs2.cpu().numpy(), rhs1)
    assert not torch.equal(w1, water0)  # the case moves some water levels
    dm.close()


@pytest.mark.parametrize("kind,size", [("sphere", 5), ("planar", 30)])
def test_fill_elevation_level_on_device(kind, size):
    """The fill with its level chosen on the device equals the fill given
    goSPL's host-side level max(min h (+1e-3 on the sphere), sealevel +
    oFill) (pitfilling.py:862-868), bit for bit, with the floor both below
    and above the minimum."""
    import torch

    from gospl_amd.model import DeviceMesh, _ptr

    m, h0, p = case(kind, size)
    dm = DeviceMesh(m, torch.device("cuda:0"))
    lib = dm.lib
    s = torch.cuda.current_stream().cuda_stream
    h = torch.as_tensor(np.ascontiguousarray(h0), dtype=torch.float64, device="cuda")
    add_eps = 0 if kind == "planar" else 1
    for floor in (float(h0.min()) - 5000.0, float(np.median(h0))):
        minh = float(h0.min()) + (1.0e-3 if add_eps else 0.0)
        level = max(minh, floor)
        a = torch.empty_like(h)
        b = torch.empty_like(h)
        assert lib.gospl_fill_depressions(dm.handle, _ptr(h), level, 0.0, _ptr(a), s) == 0
        assert lib.gospl_fill_elevation(dm.handle, _ptr(h), add_eps, floor, 0.0, _ptr(b), s) == 0
        torch.cuda.synchronize()
        assert torch.equal(a, b), floor
    dm.close()


def test_pit_tables_read_back_matches_device_tables():
    """The packed read-back of the per-pit tables holds rows 0..npits of the
    spill, parameter and depth-volume tables exactly as they are on the
    device."""
    import torch

    from gospl_amd.model import Model

    m, h0, p = case("sphere", 6)
    g = Model(m, h0, tend=1e12, **p)
    g.initialise()
    n1 = g.npits + 1
    assert n1 > 1
    host = torch.empty(14 * n1 + 7, dtype=torch.float64, pin_memory=True)
    host.fill_(123.0)
    assert g.lib.gospl_pit_tables(g.H, g.npits, g._cap, g._spill.data_ptr(), g._pitparams.data_ptr(),
                                  g._flvl.data_ptr(), g._fvol.data_ptr(), host.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream) == 0
    ref = torch.cat([g._spill[:n1].to(torch.float64), g._pitparams[:, :n1].reshape(-1),
                     g._flvl[:, :n1].reshape(-1), g._fvol[:, :n1].reshape(-1)]).cpu()
    assert torch.equal(host[:14 * n1], ref)
    assert torch.all(host[14 * n1:] == 123.0)  # nothing written past the tables
    # a bad pit count is refused
    assert g.lib.gospl_pit_tables(g.H, g._cap, g._cap, g._spill.data_ptr(), g._pitparams.data_ptr(),
                                  g._flvl.data_ptr(), g._fvol.data_ptr(), host.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream) != 0
    g.dm.close()


def test_hillslope_operator_kept_between_calls():
    """A handle that keeps its hillslope rows from an earlier call (other
    sea flags, then other diffusivities) gives bit for bit what a fresh
    handle gives on the same elevation."""
    import torch

    from gospl_amd.model import DeviceMesh, _ptr

    m, h0, p = case("sphere", 5)
    dev = torch.device("cuda:0")
    kept = DeviceMesh(m, dev)
    s = torch.cuda.current_stream().cuda_stream
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")

    def run(dm, h, cda, cdm, sl):
        hh, ce, eb = t(h), torch.zeros(m.N, dtype=torch.float64, device="cuda"), \
            torch.zeros(m.N, dtype=torch.float64, device="cuda")
        assert dm.lib.gospl_hillslope(dm.handle, cda, cdm, 5000.0, sl, 8, _ptr(hh), _ptr(ce), _ptr(eb), s) == 0
        torch.cuda.synchronize()
        return hh.cpu().numpy(), ce.cpu().numpy()

    # first call: sea level 0; second: a higher sea level moves the coastline;
    # third: other diffusivities
    for (cda, cdm, sl) in [(0.5, 0.5, 0.0), (0.5, 2.0, 800.0), (0.5, 2.0, -300.0), (0.1, 3.0, -300.0)]:
        a = run(kept, h0, cda, cdm, sl)
        fresh = DeviceMesh(m, dev)
        b = run(fresh, h0, cda, cdm, sl)
        fresh.close()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (cda, cdm, sl)
    # areas changed in place (goSPL's updatearea) and announced: the kept rows
    # are dropped, as a fresh handle on the new areas shows
    before = run(kept, h0, 0.1, 3.0, -300.0)
    kept.area.mul_(1.5)
    assert kept.lib.gospl_mesh_touch(kept.handle) == 0
    a = run(kept, h0, 0.1, 3.0, -300.0)
    import copy
    m2 = copy.copy(m)
    m2.area = m.area * 1.5
    fresh = DeviceMesh(m2, dev)
    b = run(fresh, h0, 0.1, 3.0, -300.0)
    fresh.close()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], before[0])
    kept.close()


def test_erodibility_factor_and_precipitation_exponent():
    """A per-vertex erodibility factor (goSPL's sedfactor map, clipped at 0.1
    on loading, unstructuredmesh.py:1089; the map config 4's log-K field
    feeds) and a precipitation exponent d (Kbr = K * factor * P^d,
    SPL.py:65-70) over lognormal rain, step by step against the oracle."""
    m, h0, p = case("sphere", 5)
    rng = np.random.default_rng(7)
    q = dict(p)
    q["sedfac"] = m.to_device_order(np.exp(rng.uniform(np.log(0.02), np.log(3.0), m.N)))
    q["d"] = 0.5
    assert (q["sedfac"] < 0.1).any()
    _compare_steps(m, h0, q, nsteps=3)
