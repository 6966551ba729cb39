"""Mixed precision against FP64 on the same kicked-Ising circuit.

Runs the Floquet evolution twice on one GPU - FP64 and mixed precision (FP32
MFMA contractions, FP64 messages, truncation, normalisation and ratios) - in
two contexts whose streams overlap, and reports the error of every <Z_i> of
the mixed run against the FP64 run: max and mean absolute error, relative
error over sites with |<Z>| >= 1e-2, and the floor-1 relative error
|m - f| / max(|f|, 1).

    python -m tnbp_amd.precision_report --L 32 --chi 32 --steps 10 [--seed 2024]
        [--schedule checkerboard --accel 4]

``--schedule`` / ``--accel`` set the mixed run's BP schedule and Anderson
mixing depth (the headline path's settings); the FP64 run keeps the
reference's Jacobi loop.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import threading
import time

import numpy as np

from . import driver
from ._lib import HBP_PREC_FP64, HBP_PREC_MIXED
from .tci import Context


def compare(L, chi, steps, h, tol=1e-10, max_bp_iterations=50, sv_min=1e-8, out=sys.stdout,
            schedule=0, accel=0):
    res, errs = {}, {}

    def run(prec, tag):
        mixed = prec == HBP_PREC_MIXED
        ctx = Context(0)
        try:
            def log(rec):
                print(json.dumps({"run": tag, "step": rec["step"], "bp_iterations": rec["bp_iterations"],
                                  "seconds": round(rec["seconds"], 3),
                                  "max_bond": max(rec["bond_dims"])}), file=out, flush=True)
            res[tag] = driver.run_kicked(ctx, L, steps, 0.125, 0.0, h / (2 * math.pi), 1.0,
                                         max_bp_iterations, tol, chi, sv_min, precision=prec,
                                         measure="last", observables=("z",), log=log,
                                         schedule=schedule if mixed else 0,
                                         accel=accel if mixed else 0)
        except BaseException as e:  # reported by the caller
            errs[tag] = e
        finally:
            ctx.close()

    t0 = time.perf_counter()
    th = [threading.Thread(target=run, args=(HBP_PREC_FP64, "fp64")),
          threading.Thread(target=run, args=(HBP_PREC_MIXED, "mixed"))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise next(iter(errs.values()))
    f = np.array(res["fp64"]["steps"][-1]["z"])
    m = np.array(res["mixed"]["steps"][-1]["z"])
    d = np.abs(m - f)
    big = np.abs(f) >= 1e-2
    return {
        "L": L, "chi": chi, "steps": steps, "h": h, "tol": tol, "sites": int(f.size),
        "mixed_schedule": ["jacobi", "checkerboard"][schedule], "mixed_accel": accel,
        "max_abs_err": float(d.max()), "mean_abs_err": float(d.mean()),
        "max_rel_err": float((d[big] / np.abs(f[big])).max()) if big.any() else None,
        "mean_rel_err": float((d[big] / np.abs(f[big])).mean()) if big.any() else None,
        "sites_rel": int(big.sum()),
        "max_rel_err_floor1": float((d / np.maximum(np.abs(f), 1.0)).max()),
        "bp_iterations_fp64": [s["bp_iterations"] for s in res["fp64"]["steps"]],
        "bp_iterations_mixed": [s["bp_iterations"] for s in res["mixed"]["steps"]],
        "seconds_fp64": sum(s["seconds"] for s in res["fp64"]["steps"]),
        "seconds_mixed": sum(s["seconds"] for s in res["mixed"]["steps"]),
        "wall_s": time.perf_counter() - t0,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-bp-iterations", type=int, default=50)
    ap.add_argument("--schedule", choices=["jacobi", "checkerboard"], default="jacobi")
    ap.add_argument("--accel", type=int, default=0)
    a = ap.parse_args()
    h = float(np.random.default_rng(a.seed).uniform(0.1, 1.0))
    r = compare(a.L, a.chi, a.steps, h, a.tol, a.max_bp_iterations,
                schedule=1 if a.schedule == "checkerboard" else 0, accel=a.accel)
    print(json.dumps({"report": r}), flush=True)


if __name__ == "__main__":
    main()
