"""Benchmark of the per-timestep surface-process hot path on one MI355X.

Workload: the 2.6M-vertex icosahedral sphere (refinement level 9, 2,621,442
vertices, radius 6371 km), SH initial surface (degree <= 64, spectrum l^-2,
scaled to [-6000, 5000] m, sea level 0), lognormal rain, K=5e-6, m=0.5, n=1,
3 MFD receivers, diffusivity 0.5 m2/yr, dt = 5000 yr.  One "step" is one pass
of goSPL's runProcesses body over the whole mesh: flow accumulation (fill,
pit information, receivers, flow solve, overflow cascade), stream-power
erosion, hillslope diffusion and forcing.  Inputs are resident in HBM before
the timed region.

Prints one JSON line (the driver's contract) with the roofline of the
dominant kernel and the CPU oracle timed on a bounded sample beside it.
"""

from __future__ import annotations

import argparse
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
    "gospl_spl": 116,
    "gospl_hillslope": 116,
}
METRIC = "steps/s, 2.6M-vertex sphere: FP32+refine vs FP64 plain at ≤1e-6 rel err"
STEP_BYTES = 564  # sum of ALG_BYTES: compulsory bytes per vertex and step
# the sweep kernels between the library's timing events, per entry point
SWEEP_KERNELS = {"gospl_flowacc": "k_dag_level0<0> + k_chain<0>",
                 "gospl_spl": "k_dag_level0<1> + k_chain<1>"}
# HBM traffic per sweep (level 0 + dataflow kernel) on this workload, from two
# rocprofv3 --pmc passes over `bench.py --steps 2 --warmup 1 --no-cpu
# --surface z9.npy` (KiB per dispatch, mean over the passes' solves), each
# scaled by what the counter reads for a 256 MiB device copy
# (`--copy-probe 256`): FETCH_SIZE reads 0.5x the bytes moved (scale 2),
# WRITE_SIZE 1.0x.  Counter runs need --surface: the SH construction's torch
# kernels crash the counter-collecting profiler.
PMC_SCALE = {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0}
PMC_KIB = {
    "gospl_flowacc": {"FETCH_SIZE": 213481.8 + 50784.1, "WRITE_SIZE": 15297.1 + 139938.3},
    "gospl_spl": {"FETCH_SIZE": 196818.8 + 46179.1, "WRITE_SIZE": 17258.1 + 132079.6},
}


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
    m, h0, p = workloads.sphere_case(_lib.load(), level=level, seed=seed, device=device, z=z)
    if save_surface:
        np.save(save_surface, m.to_input_order(h0))
    log("[bench] mesh level %d: N=%d built in %.1fs" % (level, m.N, time.time() - t0))
    return m, h0, p


def cpu_baseline(m, h0, p, steps):
    """The FP64 oracle (serial C + numpy, exact solves) timed on this host."""
    from oracle import oracle as orc
    from gospl_amd import workloads

    conv = {k: (m.to_input_order(v) if isinstance(v, np.ndarray) else v) for k, v in p.items()}
    o = orc.OracleModel(workloads.oracle_geometry(m), m.to_input_order(h0), tend=1e12, **conv)
    o.initialise()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
    dt = time.perf_counter() - t0
    return steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed steps (the workload is a 100-step run at dt = 5000 yr)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--surface", default=None, metavar="NPY",
                    help="initial surface (input order) saved by --save-surface instead of the SH field "
                         "(keeps the SH construction's kernels out of a counter run)")
    ap.add_argument("--save-surface", default=None, metavar="NPY")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
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

    from gospl_amd.model import Model

    dev = torch.device("cuda:0")
    m, h0, p = build_workload(args.level, args.seed, dev, args.surface, args.save_surface)
    if args.save_surface and args.steps == 0:
        return
    g = Model(m, h0, tend=1e12, device=dev, **p)
    t0 = time.time()
    g.initialise()
    torch.cuda.synchronize()
    log("[bench] initialise %.2fs" % (time.time() - t0))
    for i in range(args.warmup):
        g.step()
        torch.cuda.synchronize()
        log("[bench] warmup %d done (fa levels %d, spl levels %d, cascade %d, pits %d)" %
            (i, g.fa_levels, g.spl_levels, g.cascade_passes, g.npits))

    # timed region: K steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        g.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sps = args.steps / elapsed
    log("[bench] %d steps in %.3fs -> %.2f steps/s" % (args.steps, elapsed, sps))

    # per-entry-point durations (HIP events on the launch stream) over a second
    # timed pass of the same length
    import ctypes

    lib = g.lib
    f64p, i64p = ctypes.c_double, ctypes.c_int64
    z = [f64p(0), i64p(0), f64p(0), i64p(0)]
    lib.gospl_sweep_timing(g.H, 1, *[ctypes.byref(v) for v in z])
    g.enable_profiling(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(args.steps):
        g.step()
    torch.cuda.synchronize()
    prof_elapsed = time.perf_counter() - t1
    ph = g.phase_times()
    g.enable_profiling(False)
    fa_ms, fa_calls, spl_ms, spl_calls = f64p(0), i64p(0), f64p(0), i64p(0)
    lib.gospl_sweep_timing(g.H, 0, ctypes.byref(fa_ms), ctypes.byref(fa_calls), ctypes.byref(spl_ms),
                           ctypes.byref(spl_calls))
    N = m.N
    phases = {k: dict(ms_total=v[0], calls=v[1], ms_per_call=v[0] / v[1]) for k, v in ph.items()}
    dom = max((k for k in phases if k in ALG_BYTES), key=lambda k: phases[k]["ms_total"])
    # the dominant entry point's dominant kernel: the DAG sweep, i.e. its
    # back-to-back kernels (level 0 on the full grid, then the dataflow
    # kernel), timed by HIP events on their launch stream inside the library
    kern = {"gospl_flowacc": (SWEEP_KERNELS["gospl_flowacc"], fa_ms.value, fa_calls.value),
            "gospl_spl": (SWEEP_KERNELS["gospl_spl"], spl_ms.value, spl_calls.value)}.get(dom)
    if kern and kern[2] > 0:
        kname, kms, kcalls = kern[0], kern[1] / kern[2], kern[2]
    else:
        kname, kms, kcalls = dom, phases[dom]["ms_per_call"], phases[dom]["calls"]
    bytes_per_launch = ALG_BYTES[dom] * N
    traffic = pmc_traffic(dom, args.level) if kern and kern[2] > 0 else None
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9
    log("[bench] dominant kernel %s: %.3f ms per launch over %d launches" % (kname, kms, kcalls))
    log("[bench] phases (ms per step): " + ", ".join(
        "%s=%.3f" % (k, v["ms_total"] / args.steps) for k, v in sorted(phases.items(), key=lambda x: -x[1]["ms_total"])))

    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cps, cdt = cpu_baseline(m, h0, p, args.cpu_steps)
        log("[bench] cpu oracle %d steps in %.1fs (%.3f steps/s), total %.1fs" %
            (args.cpu_steps, cdt, cps, time.time() - t0))
        cpu = {"value": cps, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": "FP64 oracle (C + numpy, exact DAG solves), %d steps of the same %d-vertex "
                         "workload after its init flow accumulation, one host core" % (args.cpu_steps, N)}

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
        "dtype": "f64",
        "data": "synthetic (seeded SH surface, lognormal rain)",
        "config": {"workload": "icosahedral sphere level %d, %d vertices, FP64 plain path" % (args.level, N),
                   "vertices": N, "dt_yr": p["dt"], "flowdir": p["flowdir"],
                   "vertex_steps_per_s": sps * N,
                   "fp32_refine": "not implemented",
                   "phases_ms_per_step": {k: v["ms_total"] / args.steps for k, v in phases.items()},
                   "profiled_pass_s": prof_elapsed,
                   "step_roofline": {"alg_bytes_per_vertex_step": STEP_BYTES,
                                     "achieved_GBps": STEP_BYTES * N * sps / 1e9,
                                     "frac": STEP_BYTES * N * sps / 1e9 / PEAK_HBM_GBS}},
        "roofline": {"bound": "hbm", "kernel": kname, "entry_point": dom, "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic, "alg_bytes_per_vertex": ALG_BYTES[dom],
                     "ms_per_launch": kms, "launches": kcalls,
                     "note": "latency-bound: ~600 dependent levels per solve (the longest flow "
                             "path); traffic = HBM bytes per launch from rocprofv3 --pmc "
                             "FETCH_SIZE x2 + WRITE_SIZE x1 (copy-calibrated), measured on this "
                             "workload"},
        "cpu_baseline": cpu,
    }
    g.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
