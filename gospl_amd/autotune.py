"""Launch-shape autotuner for the receiver-DAG dataflow sweeps.

The sweeps' results do not depend on their launch shape (workgroup count,
the pool size from which wavefronts hand work to the global queue, idle
pollers), only their time does: a single mesh's sweeps are latency-bound
(one long dependency chain), a union of many members is throughput-bound.
``tune`` times whole model steps for each candidate shape and keeps the
fastest; ``python -m gospl_amd.autotune --workload ensemble|sphere`` runs it
on a workload and writes the winner, with the table it was chosen from,
into a JSON file that ``lookup`` reads (``TUNED``, committed beside this
module).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned_shapes.json")


def get_shape(model):
    out = (ctypes.c_int32 * 5)()
    model._call("gospl_dag_get_shape", model.H, ctypes.cast(out, ctypes.c_void_p))
    return dict(chain_blocks=out[0], chain_share=out[1], chain_pollers=out[2], resident=out[3], num_cu=out[4])


def set_shape(model, chain_blocks=-1, chain_share=-1, chain_pollers=-1):
    model._call("gospl_dag_set_shape", model.H, int(chain_blocks), int(chain_share), int(chain_pollers))


def candidates(shape):
    """Workgroup counts of 1, 2, 3, 4, 6 and 8 per compute unit (within the
    resident limit) times pool hand-off at 0 or 64 times 32 or 128 pollers."""
    cu = shape["num_cu"]
    blocks = [cu * k for k in (1, 2, 3, 4, 6, 8) if cu * k <= shape["resident"]]
    return [dict(chain_blocks=b, chain_share=s, chain_pollers=p) for b in blocks for s in (0, 64) for p in (32, 128)]


STATE = ("hLocal", "cumED", "EbLocal", "Eb")


def tune(model, cands=None, steps=3, warmup=1, rounds=2, log=None):
    """Time ``steps`` steps of ``model`` per candidate shape, each from the
    same state (restored before every candidate, so all candidates do the
    same work), after ``warmup`` untimed steps; ``rounds`` passes over the
    candidates, keeping each one's best time.  Sets the fastest shape and
    returns (best, table); the model is left at its starting state."""
    import torch

    base = get_shape(model)
    cands = cands or candidates(base)
    snap = {k: getattr(model, k).clone() for k in STATE}
    t_now = model.tNow
    best_ms = [float("inf")] * len(cands)
    for r in range(rounds):
        for i, c in enumerate(cands):
            set_shape(model, **c)
            for k in STATE:
                getattr(model, k).copy_(snap[k])
            model.tNow = t_now
            for _ in range(warmup):
                model.step()
            for k in STATE:
                getattr(model, k).copy_(snap[k])
            model.tNow = t_now
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                model.step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / steps * 1e3
            best_ms[i] = min(best_ms[i], ms)
            if log:
                log("[autotune] round %d %s: %.3f ms/step" % (r, c, ms))
    for k in STATE:
        getattr(model, k).copy_(snap[k])
    model.tNow = t_now
    table = [dict(c, ms_per_step=t) for c, t in zip(cands, best_ms)]
    best = min(table, key=lambda r: r["ms_per_step"])
    set_shape(model, best["chain_blocks"], best["chain_share"], best["chain_pollers"])
    return best, table


def lookup(workload, vertices, path=TUNED):
    """The tuned shape recorded for ``workload`` at the nearest vertex
    count, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        entries = json.load(fh).get("entries", [])
    cand = [e for e in entries if e.get("workload") == workload]
    if not cand:
        return None
    e = min(cand, key=lambda e: abs(np.log(max(e["vertices"], 1)) - np.log(max(vertices, 1))))
    return dict(chain_blocks=e["chain_blocks"], chain_share=e["chain_share"], chain_pollers=e["chain_pollers"])


def apply_tuned(model, workload):
    """Set the tuned shape for ``workload`` when one is recorded and fits."""
    s = lookup(workload, model.dm.N)
    if s is None:
        return None
    if s["chain_blocks"] > get_shape(model)["resident"]:
        return None
    set_shape(model, **s)
    return s


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=("ensemble", "sphere"), default="ensemble")
    ap.add_argument("--level", type=int, default=9, help="sphere level (--workload sphere)")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--out", default=TUNED)
    args = ap.parse_args()
    import torch

    from gospl_amd import _lib, workloads
    from gospl_amd.ensemble import EnsembleModel
    from gospl_amd.model import Model

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    if args.workload == "ensemble":
        model = EnsembleModel(workloads.ensemble_case(_lib.load(), device=dev), tend=1e12, device=dev)
        label = "ensemble"
    else:
        m, h0, p = workloads.sphere_case(_lib.load(), level=args.level, device=dev)
        model = Model(m, h0, tend=1e12, device=dev, refine=True, **p)
        label = "sphere"
    model.initialise()
    best, table = tune(model, steps=args.steps, log=log)
    entry = dict(workload=label, vertices=int(model.dm.N), chain_blocks=best["chain_blocks"],
                 chain_share=best["chain_share"], chain_pollers=best["chain_pollers"],
                 ms_per_step=best["ms_per_step"], device=torch.cuda.get_device_name(0), table=table,
                 command="python -m gospl_amd.autotune --workload %s%s --steps %d" % (
                     args.workload, (" --level %d" % args.level) if args.workload == "sphere" else "", args.steps))
    data = {"entries": []}
    if os.path.exists(args.out):
        with open(args.out) as fh:
            data = json.load(fh)
    data["entries"] = [e for e in data["entries"] if not (e["workload"] == label and e["vertices"] == entry["vertices"])]
    data["entries"].append(entry)
    with open(args.out, "w") as fh:
        json.dump(data, fh, indent=1)
    log("[autotune] best %s -> %s" % (best, args.out))
    print(json.dumps(entry))


if __name__ == "__main__":
    main()
