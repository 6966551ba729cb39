"""Benchmark of the BP-PEPS hot path on one MI355X.

Workload: kicked-Ising Floquet evolution on a 16x16 open square PEPS, chi = 32,
J = pi/4 (RZZ angle), h drawn from a seed in [0.1, 1.0], BP to residual 1e-10
(at most 50 iterations, the kicked driver's default), 256 <Z> + 480 <ZZ>
measured after every step (the driver of apps/kicked_ising_lattice/main.cc
measures every step).  The bond dimension ramps 1 -> 32 over the first six
steps; those and `--warmup` more steps are untimed.  A timed step is one
steady-state Floquet step: AbsorbTPO, the BP loop to tolerance at bond
dimension D = 64, Truncation back to 32, and the 736 expectation values.

`value` is expectation values / s of the headline path (`--precision`,
default mixed: FP32 MFMA contractions, FP64 messages, normalisation,
truncation and ratios).  With `--compare` (default on) the FP64 fused path
runs the same steps afterwards and is reported beside it.  `roofline` is for
the kernel that takes the most time in the timed region, from HIP events on
the stream it runs on; its `traffic` comes from the rocprofv3 FETCH_SIZE /
WRITE_SIZE runs recorded in profiles/traffic.json.

Prints one JSON line.  Runs on one GPU: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}
KERNELS = {0: "k_mode_product", 1: "k_message_partials", 2: "k_product_message"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--mode", choices=["fused", "plain"], default="fused",
                    help="plain: the reference's op sequence one TCI call at a time")
    ap.add_argument("--sweep", choices=["auto", "per-site", "merged"], default="auto",
                    help="BP sweep launch form (grouped needs bond dims <= 32; merged computes "
                         "each message's last mode product inside the message kernel)")
    ap.add_argument("--compare", type=int, default=1,
                    help="also time the FP64 fused path on the same steps")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-bp-iterations", type=int, default=50,
                    help="the kicked driver's default (apps/kicked_ising_lattice/option.h:19)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--max-ramp", type=int, default=12,
                    help="untimed steps allowed for the bond dimension to reach chi")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--kernel-timing", type=int, default=1,
                    help="time the hot kernels with HIP events carried in their dispatches "
                         "(0 for runs under an external profiler, which times them itself)")
    return ap.parse_args()


def cpu_baseline(site_tensor, msgs, iters, args):
    """Time the oracle (numpy + OpenBLAS on the host cores) on a bounded sample
    of the same workload: the outgoing messages of one interior site of the
    steady-state PEPS, with the same message/normalisation sequence as
    bp.h:128-175.  Scaled to expectation values/s with the site census of the
    lattice and the GPU run's BP iteration count, with the reference's two
    sweeps per BP iteration (sweep + condition, main.cc:121-159)."""
    from oracle import tci as ot
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    V = site_tensor
    z = V.ndim - 1
    t0 = time.perf_counter()
    n_msgs = 0
    while True:
        for k in range(z):
            T = V.copy()
            for j in range(z):
                if j != k:
                    la = list(range(V.ndim))
                    la[j] = -1
                    T = ot.contract(T, la, msgs[j], [-1, j], list(range(V.ndim)))
                T, _ = ot.normalize(T)
            idx_t = list(range(-V.ndim, 0))
            idx_a = list(range(-V.ndim, 0))
            idx_t[k], idx_a[k] = 0, 1
            M = ot.contract(T, idx_t, ot.cplx_conj(V), idx_a, [0, 1])
            ot.normalize(M)
            n_msgs += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    dt = time.perf_counter() - t0
    per_msg = dt / n_msgs
    L = args.L
    # messages per sweep weighted by their cost relative to an interior (z=4) message
    D = V.shape[0]
    cost4 = D ** 4 * 2 * D * 4
    cost3 = D ** 3 * 2 * D * 3
    cost2 = D ** 2 * 2 * D * 2
    n4, n3, n2 = (L - 2) ** 2, 4 * (L - 2), 4
    sweep_equiv = (n4 * 4 * cost4 + n3 * 3 * cost3 + n2 * 2 * cost2) / cost4
    iters = max(1.0, iters)
    step_s = 2 * iters * sweep_equiv * per_msg
    n_vals = L * L + 2 * L * (L - 1)
    return {
        "value": n_vals / step_s,
        "unit": "expectation values/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (numpy/OpenBLAS) computing {n_msgs} outgoing BP messages of one interior "
                   f"site of the steady-state {L}x{L} PEPS (D={D}) in {dt:.1f}s; scaled by the site "
                   f"census to one sweep ({sweep_equiv:.0f} interior-message equivalents) and by "
                   f"2 sweeps x {iters:.1f} BP iterations per step (sweep + condition, the GPU run's "
                   f"mean count); truncation and measurement not included"),
        "seconds_per_interior_message": per_msg,
    }


def traffic_table():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def run_path(args, precision, mode, want_cpu_sample=False):
    """Ramp, warm up and time `args.steps` steady-state steps of one path."""
    from tnbp_amd import circuits
    from tnbp_amd._lib import check
    from tnbp_amd.peps import (HBP_PREC_FP64, HBP_PREC_MIXED, AbsorbTPO, InitTensorProductState,
                               Measure, MeasureEdges, Truncation, bp_loop)
    from tnbp_amd.tci import Context

    L, chi = args.L, args.chi
    rng = np.random.default_rng(args.seed)
    h = float(rng.uniform(0.1, 1.0))
    J = math.pi / 4
    edges = circuits.square_lattice(L, L)
    sites = sorted({s for e in edges for s in e})
    # RZZ(J) = exp(-i J/2 ZZ) and RX(h) = exp(-i h/2 X): dt = 1 (in units of pi),
    # Jz = J / (2 pi), hx = h / (2 pi) in the reference's CircuitTPO parametrisation
    tpo = circuits.kicked_ising_tpo(edges, J / (2 * math.pi), 0.0, h / (2 * math.pi), math.pi)
    prec = HBP_PREC_FP64 if precision == "fp64" else HBP_PREC_MIXED
    zop = circuits.op_tensor(circuits.PAULI["Z"])
    zz = circuits.op_tensor(np.kron(circuits.PAULI["Z"], circuits.PAULI["Z"]))
    tag = f"{precision}/{mode}"

    ctx = Context(0)
    peps = InitTensorProductState(ctx, edges, [2] * len(sites), prec)
    check(ctx.lib.hbp_peps_set_mode(peps.h, 0 if mode == "fused" else 1))
    peps.set_sweep_kernel({"auto": 0, "per-site": 1, "merged": 3}[args.sweep])

    def step(record):
        t0 = time.perf_counter()
        AbsorbTPO(ctx, tpo, peps)
        it, res = bp_loop(ctx, peps, args.max_bp_iterations, args.tol,
                          progress=lambda k, r: print(json.dumps(
                              {"path": tag, "bp": k, "residual": r, "t": time.perf_counter() - t0}),
                              file=sys.stderr, flush=True))
        dims, errs = Truncation(ctx, peps, chi, 1e-8)
        z = Measure(ctx, peps, sites, [zop] * len(sites))
        zzv = MeasureEdges(ctx, peps, edges, [zz] * len(edges))
        ctx.synchronize()
        dt = time.perf_counter() - t0
        rec = {"seconds": dt, "bp_iterations": it, "residual": res[-1] if res else None,
               "max_bond": max(dims), "max_trunc_err": max(errs), "n_values": len(z) + len(zzv),
               "z": [v.real for v in z]}
        if record is not None:
            record.append(rec)
        print(json.dumps({"path": tag, "progress": {k: v for k, v in rec.items() if k != "z"}}),
              file=sys.stderr, flush=True)
        return rec

    # ramp: untimed steps until the bond dimension has saturated at chi, so that
    # every timed step absorbs to D = 2 chi and truncates back
    for _ in range(args.max_ramp):
        if step(None)["max_bond"] >= chi:
            break
    for _ in range(args.warmup):
        step(None)
    ctx.synchronize()
    check(ctx.lib.hbp_ctx_profile(ctx.h, 1 if args.kernel_timing else 0))
    timed = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(timed)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    stats = {}
    for kind, name in KERNELS.items():
        n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        check(ctx.lib.hbp_ctx_profile_read(ctx.h, kind, C.byref(n), C.byref(ms), C.byref(fl),
                                           C.byref(by)))
        stats[name] = {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}
    check(ctx.lib.hbp_ctx_profile(ctx.h, 0))
    sample = None
    if want_cpu_sample:
        # a steady-state interior site and its incoming messages (bp.h:147-151),
        # absorbed once more so that D = 2 chi as during BP
        s = (L // 2) * L + L // 2
        inc = [m for m, e in enumerate(edges) if s in e]
        AbsorbTPO(ctx, tpo, peps)
        sample = (peps.get_site(s),
                  [peps.get_message(m + len(edges) if edges[m][0] == s else m) for m in inc])
    peps.close()
    ctx.close()
    n_vals = sum(r["n_values"] for r in timed)
    lat = sorted(r["seconds"] for r in timed)
    return {"h": h, "wall": wall, "timed": timed, "stats": stats, "value": n_vals / wall,
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": lat[-1]},
            "bp_iterations": [r["bp_iterations"] for r in timed], "sample": sample}


def roofline(res, precision):
    """Roofline object of the kernel with the most time in the timed region."""
    st = res["stats"]
    name = max(st, key=lambda k: st[k]["ms"])
    k = st[name]
    peak = PEAK["fp64"] if precision == "fp64" else PEAK["fp32"]
    achieved = k["flops"] / (k["ms"] * 1e-3) if k["ms"] > 0 else 0.0
    mfma = "FP64 MFMA 16x16x4" if precision == "fp64" else "FP32 MFMA 16x16x4"
    tr = traffic_table().get(precision, {}).get(name)
    return {
        "bound": "mfma",
        "kernel": f"{name} ({mfma})",
        "achieved": achieved / 1e12,
        "peak": peak / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": tr["hbm_bytes_per_launch"] if tr else None,
        "traffic_launch": tr["launch"] if tr else None,
        "traffic_algorithmic_bytes": tr["algorithmic_bytes"] if tr else None,
        "launches": k["launches"],
        "avg_launch_ms": k["ms"] / max(1, k["launches"]),
        "algorithmic_flops_per_launch": k["flops"] / max(1, k["launches"]),
        "algorithmic_bytes_per_launch": k["bytes"] / max(1, k["launches"]),
        "achieved_GBps_algorithmic": k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else 0.0,
        "kernel_time_fraction": k["ms"] / (res["wall"] * 1e3),
        "other_kernels": {n: {"TFLOPs": (v["flops"] / (v["ms"] * 1e-3) / 1e12 if v["ms"] > 0 else 0.0),
                              "time_fraction": v["ms"] / (res["wall"] * 1e3)}
                          for n, v in st.items() if n != name},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU path is a design "
                                   "(real-space decomposition of the lattice), not measured"}))
        sys.exit(2)
    head = run_path(args, args.precision, args.mode, want_cpu_sample=bool(args.cpu_baseline))
    comp = None
    if args.compare and not (args.precision == "fp64" and args.mode == "fused"):
        comp = run_path(args, "fp64", "fused")
    cpu = None
    if args.cpu_baseline:
        V, msgs = head["sample"]
        iters = float(np.mean(comp["bp_iterations"] if comp else head["bp_iterations"]))
        cpu = cpu_baseline(V, msgs, iters, args)
    L, chi = args.L, args.chi
    out = {
        "metric": "expectation values/s + p99 latency, 16x16 PEPS chi=32 mixed vs FP64 path",
        "value": head["value"],
        "unit": "expectation values/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-mfma/f64",
        "data": "synthetic (seeded kicked-Ising circuit, no external data)",
        "config": {"workload": f"kicked-Ising {L}x{L} PEPS chi={chi}, complex128 state, J=pi/4, "
                               f"h={head['h']:.6f} (seed {args.seed}), BP tol {args.tol:g} "
                               f"(max {args.max_bp_iterations} iterations), steady-state Floquet "
                               f"step with 256 <Z> + 480 <ZZ>",
                   "precision": args.precision, "mode": args.mode, "sweep": args.sweep,
                   "L": L, "chi": chi},
        "latency_s": head["latency_s"],
        "bp_iterations": head["bp_iterations"],
        "roofline": roofline(head, args.precision),
        "cpu_baseline": cpu,
    }
    if comp:
        zh = np.array([v for r in head["timed"] for v in r["z"]])
        zc = np.array([v for r in comp["timed"] for v in r["z"]])
        out["fp64_fused"] = {"value": comp["value"], "ms_per_step": comp["wall"] / args.steps * 1e3,
                             "latency_s": comp["latency_s"], "bp_iterations": comp["bp_iterations"],
                             "roofline": roofline(comp, "fp64")}
        out["speedup_vs_fp64_fused"] = head["value"] / comp["value"]
        out["max_abs_diff_z_vs_fp64"] = float(np.max(np.abs(zh - zc)))
    from tnbp_amd.tci import autotune_report
    out["autotune"] = autotune_report().splitlines()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
