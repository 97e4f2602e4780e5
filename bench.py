"""Benchmark of the per-timestep surface-process hot path on one MI355X.

Workload: the 2.6M-vertex icosahedral sphere (refinement level 9, 2,621,442
vertices, radius 6371 km), SH initial surface (degree <= 64, spectrum l^-2,
scaled to [-6000, 5000] m, sea level 0), lognormal rain, K=5e-6, m=0.5, n=1,
3 MFD receivers, diffusivity 0.5 m2/yr, dt = 5000 yr.  ``--level 10`` runs the
10,485,762-vertex sphere (SH degree <= 256, spectrum l^-2.5, m = 0.45,
dt = 2000 yr).  One "step" is one pass of goSPL's runProcesses body over the
whole mesh: flow accumulation (fill, pit information, receivers, flow solve,
overflow cascade), stream-power erosion, hillslope diffusion and forcing.
Inputs are resident in HBM before the timed region.

Three legs of the same workload, each timed over the same window (W untimed
warm-up steps, then K timed steps between synchronisations), after ~2 s of
untimed steps of a throw-away model that bring the GPU out of idle:
* refined (``value``): hillslope diffusion by FP32 Jacobi corrections of an
  FP64 residual (rtol 1e-12); the triangular receiver-DAG systems by their
  exact FP64 sweeps (a direct solve: one pass, nothing left to refine);
* refined_dag: in addition the DAG systems by refinement -- 64-bit
  fixed-point correction sweeps with FP64 residuals (rtol 1e-12);
* fp64: exact FP64 sweeps and FP64 Jacobi hillslope (the plain path).
Then every leg runs ``--acc-steps`` steps from the initial surface, beside
FP64 runs from the surface nudged by +-1 ulp (three seeds), and each refined
trajectory's relative elevation error against the FP64 one is reported next
to the FP64 path's own divergence under that nudge.

A gradient leg (``--adjoint-level``, 8 by default: 655,362 vertices) times
the inversion workload's forward run and its gradient in per-vertex log10 K
and uplift (50 steps, checkpoints every 10; ``gospl_amd.adjoint``), with a
finite-difference spot check of the gradient.  An ensemble leg
(``--ensemble-steps``) steps 144 members of uneven size (64 x level 6,
64 x level 7, 16 x level 8) together on one union mesh and reports
member-steps/s beside running them one at a time.

Prints one JSON line (the driver's contract) with the roofline of the
dominant kernel of the reported leg and the CPU oracle timed on a bounded
sample beside it.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# Algorithmic (compulsory, single-pass) bytes per vertex per step of each
# phase at FP64 state, int32 indices, 6 neighbours, 3 receivers.
ALG_BYTES = {
    "gospl_fill_elevation": 44,
    "gospl_pit_info": 64,
    "gospl_flow_direction": 156,
    "gospl_flowacc": 68,
    "gospl_flow_accumulation": 68,   # its flow sweeps (the native flowAccumulation call)
    "gospl_spl": 116,
    "gospl_hillslope": 116,
    "gospl_flowacc_refine": 68,
    "gospl_spl_refine": 116,
    "gospl_hillslope_refine": 116,
}
METRIC = "steps/s, 2.6M-vertex sphere: FP32+refine vs FP64 plain at ≤1e-6 rel err"
STEP_BYTES = 564  # compulsory bytes per vertex and step (sum over the phases)
# the sweep kernels between the library's timing events, per entry point
SWEEP_KERNELS = {"gospl_flowacc": ("k_rec_build0 (with level 0) + k_chain<0>", 0),
                 "gospl_flow_accumulation": ("k_rec_build0 (with level 0) + k_chain<0>", 0),
                 "gospl_spl": ("k_rec_build1 (with level 0) + k_chain<1>", 1),
                 "gospl_flowacc_refine": ("k_rx_level0<0> + k_rx_chain<0>", 0),
                 "gospl_spl_refine": ("k_rx_level0<1> + k_rx_chain<1>", 1)}
# HBM traffic per sweep (record build with level 0 + dataflow kernel) on the level-9 workload,
# from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per
# dispatch, mean over the passes' sweeps), each scaled by what the counter
# reads for a 256 MiB device copy (`--copy-probe 256`): FETCH_SIZE reads 0.5x
# the bytes moved (scale 2), WRITE_SIZE 1.0x.  The runs and their commands are
# in profiles/*_pmc_traffic.md.
PMC_SCALE = {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0}
PMC_KIB = {
    "gospl_flowacc": {"FETCH_SIZE": 92162.2 + 124407.3, "WRITE_SIZE": 352432.5 + 140066.2},
    "gospl_spl": {"FETCH_SIZE": 96833.1 + 111452.7, "WRITE_SIZE": 353008.5 + 132284.4},
}
PMC_KIB["gospl_flow_accumulation"] = PMC_KIB["gospl_flowacc"]


def pmc_traffic(entry, level):
    """HBM bytes per launch of the entry point's sweep (level-9 sphere only)."""
    if level != 9 or entry not in PMC_KIB:
        return None
    return sum(kib * PMC_SCALE[c] * 1024.0 for c, kib in PMC_KIB[entry].items())


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(level, seed, device, surface=None, save_surface=None):
    from gospl_amd import _lib, workloads

    t0 = time.time()
    z = np.load(surface) if surface else None
    if level >= 10:
        m, h0, p = workloads.sphere_case(_lib.load(), level=level, seed=seed, device=device, z=z, lmax=256,
                                         slope=2.5)
        p.update(m=0.45, dt=2000.0)
    else:
        m, h0, p = workloads.sphere_case(_lib.load(), level=level, seed=seed, device=device, z=z)
    if save_surface:
        np.save(save_surface, m.to_input_order(h0))
    log("[bench] mesh level %d: N=%d built in %.1fs" % (level, m.N, time.time() - t0))
    return m, h0, p


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m, h0, p, steps):
    """The CPU restatement of goSPL's path in reference-solver mode (one rank:
    heap priority-flood fill, the mfdreceivers kernel, FGMRES(30) + ILU(0) at
    rtol 1e-8 for the flow and stream-power systems, Richardson + ILU(0) for
    hillslope, goSPL's cascade glue), timed on this host single-threaded and
    with OpenMP on the host threads available (sparse products and vector
    updates; the ILU factorisation and solves stay sequential, as PETSc's)."""
    from oracle import oracle as orc
    from gospl_amd import workloads

    conv = {k: (m.to_input_order(v) if isinstance(v, np.ndarray) else v) for k, v in p.items()}
    nthr = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    o = orc.ReferenceSolverModel(workloads.oracle_geometry(m), m.to_input_order(h0), tend=1e12, threads=nthr,
                                 **conv)
    t0 = time.perf_counter()
    o.initialise()
    # one untimed step: the first step after the initial flow accumulation
    # meets an unchanged surface, so its warm-started solves have nothing to do
    o.step()
    log("[bench] cpu baseline: initialise + 1 untimed step %.1fs" % (time.perf_counter() - t0))
    res = {}
    for th in sorted({1, nthr}):
        o.threads = th
        o.solver_its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            o.step()
        dt = time.perf_counter() - t0
        res[th] = dict(steps_per_s=steps / dt, s_per_step=dt / steps,
                       solver_iterations=[(k, i) for k, i, _ in o.solver_its])
        log("[bench] cpu baseline: %d step(s) on %d thread(s) in %.1fs" % (steps, th, dt))
    return res, nthr


LEGS = {"refined": dict(refine=True), "refined_dag": dict(refine=True, refine_dag=True), "fp64": {}}


def device_warmup(m, h0, p, dev, seconds):
    """Untimed steps of a throw-away model of the same workload before the
    first leg, for at least `seconds`: a leg run on a GPU that has been idle
    (mesh construction on the host) measured ~1.5x slower over its first
    ~25 steps than over the next ones, the clocks still ramping up.  Logs the
    rate of each 10-step window so the ramp shows in the log."""
    import torch

    from gospl_amd.model import Model

    if seconds <= 0:
        return
    g = Model(m, h0, tend=1e12, device=dev, **p)
    g.initialise()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = 0
    while True:
        t1 = time.perf_counter()
        for _ in range(10):
            g.step()
        torch.cuda.synchronize()
        n += 10
        log("[bench] device warm-up: steps %d-%d at %.1f steps/s" % (n - 10, n - 1, 10 / (time.perf_counter() - t1)))
        if time.perf_counter() - t0 >= seconds:
            break
    g.close()


def run_leg(m, h0, p, dev, name, steps, warmup):
    """W warm-up steps, K timed steps, then K profiled steps (HIP events on the
    launch stream around every entry point and every sweep)."""
    import torch

    from gospl_amd.model import Model

    g = Model(m, h0, tend=1e12, device=dev, **LEGS[name], **p)
    t0 = time.time()
    g.initialise()
    torch.cuda.synchronize()
    log("[bench] %s: initialise %.2fs" % (name, time.time() - t0))
    for i in range(warmup):
        g.step()
        torch.cuda.synchronize()
        log("[bench] %s: warmup %d done (cascade %d, pits %d)" % (name, i, g.cascade_passes, g.npits))
    g.solve_log.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        g.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sps = steps / elapsed
    log("[bench] %s: %d steps in %.3fs -> %.2f steps/s" % (name, steps, elapsed, sps))
    solves = list(g.solve_log)
    lib = g.lib
    f64p, i64p = ctypes.c_double, ctypes.c_int64
    z = [f64p(0), i64p(0), f64p(0), i64p(0)]
    lib.gospl_sweep_timing(g.H, 1, *[ctypes.byref(v) for v in z])
    first = dict(Model.sweeps)  # process-wide sweep launches before the profiled window
    g.enable_profiling(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        g.step()
    torch.cuda.synchronize()
    prof_elapsed = time.perf_counter() - t1
    ph = g.phase_times()
    g.enable_profiling(False)
    window = {k: (first[k], Model.sweeps[k]) for k in first}
    sw = [f64p(0), i64p(0), f64p(0), i64p(0)]
    lib.gospl_sweep_timing(g.H, 0, *[ctypes.byref(v) for v in sw])
    g.close()
    phases = {k: dict(ms_total=v[0], calls=v[1], ms_per_call=v[0] / v[1]) for k, v in ph.items()}
    log("[bench] %s phases (ms per step): " % name + ", ".join(
        "%s=%.3f" % (k, v["ms_total"] / steps) for k, v in sorted(phases.items(), key=lambda x: -x[1]["ms_total"])))
    return dict(sps=sps, elapsed=elapsed, phases=phases, prof_elapsed=prof_elapsed, solves=solves, window=window,
                sweep=dict(fa_ms=sw[0].value, fa_calls=sw[1].value, spl_ms=sw[2].value, spl_calls=sw[3].value))


def default_physics_leg(m, h0, p, dev, steps, warmup, G):
    """goSPL's default step on the same workload (informational, not `value`):
    nodep False and seadepo True -- flow, erosion, flow again, continental
    sediment routing, marine deposition, hillslope -- with the erosion
    transport-limited when G > 0.  W warm-up steps, K timed steps, then K
    profiled steps (HIP events around every entry point)."""
    import torch

    from gospl_amd.model import Model

    g = Model(m, h0, tend=1e12, device=dev, nodep=False, seadepo=True, marineSolver="picard", G=G, **p)
    g.initialise()
    for _ in range(warmup):
        g.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    g.enable_profiling(True)
    for _ in range(steps):
        g.step()
    ph = g.phase_times()
    g.enable_profiling(False)
    st = dict(g.sea_stats)
    it = int(getattr(g, "ed_iters", 0) or 0)
    g.close()
    out = dict(G=G, steps=steps, warmup=warmup, steps_per_s=steps / elapsed, ms_per_step=1e3 * elapsed / steps,
               phases_ms_per_step={k: v[0] / steps for k, v in sorted(ph.items(), key=lambda x: -x[1][0])},
               last_step={"coast_points": st.get("coast_points"), "marine_smoothing_sweeps": st.get("smooth_sweeps"),
                          "marine_routing_passes": st.get("route_passes"),
                          "transport_limited_gmres_iterations": it})
    log("[bench] default physics (G=%g): %.1f ms per step" % (G, out["ms_per_step"]))
    return out


def dominant_kernel(leg, N, level):
    """The entry point with the most time among the per-vertex phases, and its
    sweep kernels' duration per launch (HIP events inside the library)."""
    phases = leg["phases"]
    dom = max((k for k in phases if k in ALG_BYTES), key=lambda k: phases[k]["ms_total"])
    sw = leg["sweep"]
    if dom in SWEEP_KERNELS:
        kname, slot = SWEEP_KERNELS[dom]
        kms_tot, kcalls = (sw["fa_ms"], sw["fa_calls"]) if slot == 0 else (sw["spl_ms"], sw["spl_calls"])
        kms = kms_tot / kcalls if kcalls else phases[dom]["ms_per_call"]
    else:
        kname, kms, kcalls = dom, phases[dom]["ms_per_call"], phases[dom]["calls"]
    achieved = ALG_BYTES[dom] * N / (kms * 1e-3) / 1e9
    traffic = pmc_traffic(dom, level) if dom in SWEEP_KERNELS else None
    return dict(bound="hbm", kernel=kname, entry_point=dom, achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                frac=achieved / PEAK_HBM_GBS, traffic=traffic, alg_bytes_per_vertex=ALG_BYTES[dom],
                ms_per_launch=kms, launches=kcalls)


def accuracy(m, h0, p, dev, nsteps, legs):
    """Relative elevation error of each refined trajectory after nsteps
    against the FP64 one, and the FP64 path's own divergence from surfaces
    nudged by +-1 ulp at every vertex (three seeds: the envelope a change of
    rounding is read against, as goSPL's step is discontinuous at the last bit
    of its inputs)."""
    import torch

    from gospl_amd.model import Model

    cases = [("fp64", h0, {})] + [(n, h0, LEGS[n]) for n in legs if n != "fp64"]
    for seed in (7, 8, 9):
        rng = np.random.default_rng(seed)
        h1 = np.where(rng.random(m.N) < 0.5, np.nextafter(h0, np.inf), np.nextafter(h0, -np.inf))
        cases.append(("fp64_ulp%d" % seed, h1, {}))
    runs = {}
    t0 = time.time()
    for name, h, kw in cases:
        g = Model(m, h, tend=1e12, device=dev, **kw, **p)
        g.refine_stats = bool(kw)
        g.initialise()
        for _ in range(nsteps):
            g.step()
        torch.cuda.synchronize()
        runs[name] = (g.hLocal.cpu().numpy(), list(g.solve_log))
        g.close()
        log("[bench] accuracy: %s %d steps done (%.1fs)" % (name, nsteps, time.time() - t0))
    a = m.area
    hf = runs["fp64"][0]

    def err(h):
        d = h - hf
        return dict(l2_area=float(np.sqrt((a * d * d).sum() / (a * hf * hf).sum())),
                    max_rel=float(np.abs(d).max() / np.abs(hf).max()),
                    frac_vertices_over_1e6=float(np.mean(np.abs(d) > 1e-6 * np.abs(hf).max())),
                    mean_m=float(abs((a * d).sum() / a.sum())))

    out = dict(steps=nsteps)
    ulp = [err(runs[k][0]) for k in runs if k.startswith("fp64_ulp")]
    out["fp64_ulp_vs_fp64"] = {k: max(u[k] for u in ulp) for k in ulp[0]}  # worst of the three seeds
    for n in legs:
        if n == "fp64":
            continue
        stats = {}
        solves = runs[n][1]
        for kind in ("flowacc", "spl", "hillslope"):
            ps = [s[1] for s in solves if s[0] == kind]
            rr = [s[2] for s in solves if s[0] == kind and s[1] >= 0]
            if ps:
                stats[kind] = dict(solves=len(ps), passes_mean=float(np.mean([x for x in ps if x >= 0] or [0])),
                                   passes_max=int(max(ps)), exact_fallbacks=int(sum(1 for x in ps if x < 0)),
                                   relres_max=float(max(rr)) if rr else None)
        out[n + "_vs_fp64"] = err(runs[n][0])
        out[n + "_solves"] = stats
    return out


def adjoint_leg(level, seed, dev, nfd):
    """The inversion workload (sphere of the given level, 50 steps of 5000 yr,
    checkpoints every 10): a target from per-vertex log10 K = -5.3 + smooth
    perturbation and uplift 0.5 mm/yr, then at uniform log10 K = -5.3 the
    loss sum (h - target)^2 and its gradient in log10 K and uplift.  Times
    the forward run and the gradient (replay + reverse sweeps); checks nfd
    random vertices per input against central differences of the frozen map
    (h = 1e-4 in log10 K and in mm/yr)."""
    import torch

    from gospl_amd import _lib, workloads
    from gospl_amd.adjoint import Evolution

    t0 = time.time()
    m, h0, p, lk_true, lk0, up = workloads.adjoint_case(_lib.load(), level=level, seed=seed, device=dev)
    nsteps = p.pop("nsteps")
    log("[bench] adjoint workload level %d: N=%d built in %.1fs" % (level, m.N, time.time() - t0))
    evo = Evolution(m, h0, nsteps=nsteps, checkpoint_every=10, **p)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64)).to(dev)
    lk_true, lk0, up = t(lk_true), t(lk0), t(up)
    target, _ = evo.run(lk_true, up)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    h, st = evo.run(lk0, up)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    loss = float(((h - target) ** 2).sum())
    t0 = time.perf_counter()
    g_logk, g_up = evo.gradient(st, 2.0 * (h - target))
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    log("[bench] adjoint: forward %.3f s, gradient %.3f s (%d steps)" % (t_fwd, t_grad, nsteps))
    out = {"workload": "icosahedral sphere level %d, %d vertices, %d steps of %g yr, checkpoints every 10, "
                       "loss sum (h - target)^2" % (level, m.N, nsteps, p["dt"]),
           "forward_s": t_fwd, "gradient_s": t_grad, "forward_steps_per_s": nsteps / t_fwd,
           "gradient_over_forward": t_grad / t_fwd, "loss": loss,
           "grad_logk_maxabs": float(g_logk.abs().max()), "grad_uplift_maxabs": float(g_up.abs().max())}
    if nfd > 0:
        _, str_ = evo.run(lk0, up, record=True)
        rng = np.random.default_rng(seed + 5)
        worst = {}
        for what, base, g, step in (("logk", lk0, g_logk, 1.0e-4), ("uplift", up, g_up, 1.0e-7)):
            ga = g.cpu().numpy()
            cand = np.nonzero(np.abs(ga) > 1e-3 * np.abs(ga).max())[0]
            errs = []
            for i in rng.choice(cand, min(nfd, len(cand)), replace=False):
                xp, xm = base.clone(), base.clone()
                xp[i] += step
                xm[i] -= step
                a = evo.frozen_run(str_, xp, up) if what == "logk" else evo.frozen_run(str_, lk0, xp)
                b = evo.frozen_run(str_, xm, up) if what == "logk" else evo.frozen_run(str_, lk0, xm)
                fd = float(torch.sum((a - b) * (a + b - 2.0 * target))) / (2.0 * step)
                errs.append(abs(fd - ga[i]) / abs(fd))
            worst[what] = max(errs)
        del str_
        out["fd_check"] = {"vertices_per_input": nfd, "h": "1e-4 (log10 K), 1e-4 mm/yr (uplift)",
                           "worst_rel_err_logk": worst["logk"], "worst_rel_err_uplift": worst["uplift"]}
        log("[bench] adjoint finite-difference check: %s" % json.dumps(out["fd_check"]))
    evo.dm.close()
    return out


def ensemble_leg(dev, steps, warmup, seed):
    """The ensemble workload (64 members at level 6, 64 at level 7, 16 at
    level 8; per-member K, m and rain scale; refined mode) stepped together
    on one union mesh: member-steps/s over ``steps`` timed steps, beside the
    rate of running the members one after another (one member per level
    timed alone, weighted by the member counts)."""
    import torch

    from gospl_amd import _lib, workloads
    from gospl_amd.ensemble import EnsembleModel
    from gospl_amd.model import Model

    t0 = time.time()
    members = workloads.ensemble_case(_lib.load(), seed=seed, device=dev)
    log("[bench] ensemble: %d members built in %.1fs" % (len(members), time.time() - t0))
    t0 = time.time()
    ens = EnsembleModel(members, tend=1e12, device=dev)
    ens.initialise()
    torch.cuda.synchronize()
    log("[bench] ensemble: union of %d vertices (%d with padding), initialised in %.1fs" %
        (int(ens.sizes.sum()), ens.dm.N, time.time() - t0))
    for i in range(warmup):
        ens.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ens.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nm = ens.nmem
    log("[bench] ensemble: %d steps in %.3fs -> %.1f member-steps/s" % (steps, el, nm * steps / el))
    ens.enable_profiling(True)
    for i in range(2):
        ens.step()
    ph = {k: v[0] / 2 for k, v in ens.phase_times().items()}
    ens.enable_profiling(False)
    # the same members one at a time: one member per level, timed alone
    seq = {}
    levels = {}
    for (m, h0, p) in members:
        levels.setdefault(m.N, (m, h0, p, 0))
        mm, hh, pp, c = levels[m.N]
        levels[m.N] = (mm, hh, pp, c + 1)
    total = 0.0
    for n, (m, h0, p, count) in sorted(levels.items()):
        g = Model(m, h0, tend=1e12, device=dev, refine=True, **p)
        g.initialise()
        for _ in range(2):
            g.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(10):
            g.step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) / 10 * 1e3
        seq["%d vertices" % n] = {"members": count, "ms_per_step_alone": ms}
        total += count * ms
        g.close()
    seq_rate = nm / (total / 1e3)
    out = {"workload": "%d members (64 x 40,962, 64 x 163,842, 16 x 655,362 vertices), per-member K, m, rain "
                       "scale, dt 5000 yr, refined mode, one union mesh" % nm,
           "members": nm, "vertices": int(ens.sizes.sum()), "steps": steps, "warmup": warmup,
           "elapsed_s": el, "member_steps_per_s": nm * steps / el, "ms_per_union_step": el / steps * 1e3,
           "vertex_steps_per_s": float(ens.sizes.sum()) * steps / el,
           "phases_ms_per_step": ph,
           "one_member_at_a_time": {"member_steps_per_s": seq_rate, "per_level": seq},
           "speedup_over_one_at_a_time": (nm * steps / el) / seq_rate}
    ens.close()
    return out


def rocpd_summary(db, first, count, mode, N):
    """Markdown summary of a rocprofv3 database (--kernel-trace): kernels by
    total time, and the exact sweep dispatches first .. first + count - 1
    (level 0 + dataflow kernel of `mode`, in launch order) with the roofline
    they give at the workload's algorithmic bytes."""
    import sqlite3

    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name "
                       "order by sum(duration) desc").fetchall()
    out = ["| kernel | calls | total ms | mean us |", "|---|---|---|---|"]
    for name, n, tot, avg in rows[:30]:
        out.append("| `%s` | %d | %.2f | %.1f |" % (name.split("(")[0].replace("void ", "")[:70], n, tot / 1e6,
                                                  avg / 1e3))
    key = {0: ("k_rec_build0", "k_chain<0>"), 1: ("k_rec_build1", "k_chain<1>")}[mode]
    lv = [r[0] for r in cur.execute("select duration from kernels where name like ? order by start",
                                    ("%gospl::" + key[0] + "%",))]
    ch = [r[0] for r in cur.execute("select duration from kernels where name like ? order by start",
                                    ("%gospl::" + key[1] + "(%",))]
    lv, ch = lv[first:first + count], ch[first:first + count]
    if lv and ch:
        ml, mc = np.mean(lv) / 1e3, np.mean(ch) / 1e3
        ab = {0: 68, 1: 116}[mode] * N
        gbs = ab / ((ml + mc) * 1e-6) / 1e9
        out += ["", "Window: sweep dispatches %d-%d (%d launches): `%s` mean %.1f us + `%s` mean %.1f us = "
                "%.1f us; algorithmic bytes %d x %d = %.1f MB -> %.1f GB/s = %.4f of %.0f GB/s." % (
                    first, first + len(ch) - 1, len(ch), key[0], ml, key[1], mc, ml + mc, ab // N, N, ab / 1e6,
                    gbs, gbs / PEAK_HBM_GBS, PEAK_HBM_GBS)]
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps per leg")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--device-warmup", type=float, default=2.0, metavar="SECONDS",
                    help="untimed steps of a throw-away model before the first leg (GPU clock ramp)")
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--legs", default="refined,refined_dag,fp64", help="comma list of " + ", ".join(LEGS))
    ap.add_argument("--acc-steps", type=int, default=100,
                    help="steps of the accuracy comparison (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=1, help="timed CPU-baseline steps per thread count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-level", type=int, default=8,
                    help="sphere level of the gradient (inversion) leg; 0 skips it")
    ap.add_argument("--adjoint-fd", type=int, default=4,
                    help="vertices per input of the leg's finite-difference check")
    ap.add_argument("--ensemble-steps", type=int, default=50,
                    help="timed steps of the 144-member ensemble leg; 0 skips it")
    ap.add_argument("--default-physics-steps", type=int, default=5,
                    help="timed steps of goSPL's default step (sediment routing and marine deposition on), "
                         "with G = 0 and G = 1; 0 skips it")
    ap.add_argument("--surface", default=None, metavar="NPY",
                    help="initial surface (input order) saved by --save-surface instead of the SH field "
                         "(keeps the SH construction's kernels out of a counter run)")
    ap.add_argument("--save-surface", default=None, metavar="NPY")
    ap.add_argument("--rocpd-summary", default=None, metavar="DB",
                    help="print a markdown summary of a rocprofv3 --kernel-trace database and exit")
    ap.add_argument("--window", default="0:0", metavar="FIRST:COUNT",
                    help="sweep dispatches (launch order) whose mean the summary reports")
    ap.add_argument("--sweep-mode", type=int, default=0, help="0: flow sweeps, 1: stream-power sweeps")
    ap.add_argument("--vertices", type=int, default=2621442)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.rocpd_summary:
        f, c = (int(x) for x in args.window.split(":"))
        print(rocpd_summary(args.rocpd_summary, f, c, args.sweep_mode, args.vertices))
        return
    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copy probe: 5 copies of %d MiB" % args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the whole mesh lives on one MI355X and this build has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    m, h0, p = build_workload(args.level, args.seed, dev, args.surface, args.save_surface)
    if args.save_surface and args.steps == 0:
        return
    N = m.N
    device_warmup(m, h0, p, dev, args.device_warmup)
    legs = {}
    for name in [x.strip() for x in args.legs.split(",") if x.strip()]:
        legs[name] = run_leg(m, h0, p, dev, name, args.steps, args.warmup)
    main_leg = "refined" if "refined" in legs else next(iter(legs))
    L = legs[main_leg]
    sps = L["sps"]
    roof = dominant_kernel(L, N, args.level)
    sys_ = "flow" if roof["entry_point"] in ("gospl_flowacc", "gospl_flowacc_refine",
                                             "gospl_flow_accumulation") else "spl"
    a0, a1 = L["window"][sys_]
    roof["window"] = ("steps %d-%d of the %s leg (after %d warm-up and %d timed steps); exact %s sweeps "
                      "%d-%d of the process (0-based dispatch order of the sweep kernels)" % (
                          args.warmup + args.steps + 1, args.warmup + 2 * args.steps, main_leg, args.warmup,
                          args.steps, sys_, a0, a1 - 1))
    roof["note"] = ("latency-bound: ~600 dependent hops per sweep (the longest flow path); achieved = "
                    "algorithmic bytes per launch / mean launch duration (HIP events on the launch stream)")
    log("[bench] dominant kernel %s: %.3f ms per launch over %d launches (%.1f GB/s)" %
        (roof["kernel"], roof["ms_per_launch"], roof["launches"], roof["achieved"]))

    acc = None
    if args.acc_steps > 0 and any(n != "fp64" for n in legs):
        acc = accuracy(m, h0, p, dev, args.acc_steps, list(legs))
        log("[bench] accuracy after %d steps: %s" % (args.acc_steps, json.dumps(
            {k: v for k, v in acc.items() if k.endswith("_vs_fp64")})))

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        res, nthr = cpu_baseline(m, h0, p, args.cpu_steps)
        log("[bench] cpu baseline total %.1fs" % (time.time() - t0))
        one = res[1]
        cpu = {"value": one["steps_per_s"], "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": ("CPU restatement of goSPL's path in reference-solver mode (1 rank: heap "
                          "priority-flood fill, mfdreceivers, FGMRES(30)+ILU(0) rtol 1e-8 for flow and "
                          "stream power, Richardson+ILU(0) rtol 1e-8 for hillslope, goSPL's cascade "
                          "glue), %d step(s) of the same %d-vertex workload after its init flow "
                          "accumulation, single-threaded; 100 steps extrapolate to %.0f s" %
                          (args.cpu_steps, N, 100.0 * one["s_per_step"])),
               "ranks": 1, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "solver_iterations": one["solver_iterations"]}
        if nthr > 1:
            cpu["threaded"] = {"threads": nthr, "value": res[nthr]["steps_per_s"],
                               "solver_iterations": res[nthr]["solver_iterations"]}

    adj = None
    if args.adjoint_level > 0:
        adj = adjoint_leg(args.adjoint_level, args.seed, dev, args.adjoint_fd)
    ens = None
    if args.ensemble_steps > 0:
        ens = ensemble_leg(dev, args.ensemble_steps, 2, args.seed)
    dphys = None
    if args.default_physics_steps > 0:
        dphys = [default_physics_leg(m, h0, p, dev, args.default_physics_steps, 2, G) for G in (0.0, 1.0)]

    leg_cfg = {}
    for name, lg in legs.items():
        leg_cfg[name] = dict(steps_per_s=lg["sps"], ms_per_step=1e3 / lg["sps"],
                             phases_ms_per_step={k: v["ms_total"] / args.steps for k, v in lg["phases"].items()},
                             profiled_pass_s=lg["prof_elapsed"],
                             solves_in_timed_window=len(lg["solves"]),
                             correction_passes_in_timed_window=int(sum(s[1] for s in lg["solves"] if s[1] > 0)))
    cfg = {"workload": "icosahedral sphere level %d, %d vertices, %s leg reported" % (args.level, N, main_leg),
           "vertices": N, "dt_yr": p["dt"], "flowdir": p["flowdir"], "spl_m": p["m"],
           "vertex_steps_per_s": sps * N, "legs": leg_cfg,
           "step_roofline": {"alg_bytes_per_vertex_step": STEP_BYTES,
                             "achieved_GBps": STEP_BYTES * N * sps / 1e9,
                             "frac": STEP_BYTES * N * sps / 1e9 / PEAK_HBM_GBS}}
    if "fp64" in legs:
        cfg["fp64_steps_per_s"] = legs["fp64"]["sps"]
        for n in legs:
            if n != "fp64":
                cfg[n + "_over_fp64"] = legs[n]["sps"] / legs["fp64"]["sps"]
    if acc is not None:
        cfg["accuracy"] = acc
    if adj is not None:
        cfg["adjoint"] = adj
    if ens is not None:
        cfg["ensemble"] = ens
    if dphys is not None:
        cfg["default_physics"] = dphys
    dtype = {"refined": "f64 (DAG sweeps, state, residuals) + f32 (hillslope corrections)",
             "refined_dag": "f64 (state, residuals) + i64 fixed-point (DAG corrections) + f32 (hillslope)",
             "fp64": "f64"}[main_leg]
    out = {
        "metric": METRIC,
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded SH surface, lognormal rain)",
        "config": cfg,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
