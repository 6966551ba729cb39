"""Benchmark of the BP-PEPS hot path on one MI355X: 16x16 kicked-Ising jobs.

Workload (one job): kicked-Ising Floquet evolution on a 16x16 open square
PEPS, chi = 32, from |0...0>, 20 Floquet steps, J = pi/4 (RZZ angle), h drawn
from a seed in [0.1, 1.0]; each step is AbsorbTPO, the BP loop to residual
1e-10 (at most 50 iterations, the kicked driver's default), Truncation to
chi, and at the end of the job the 256 <Z> + 480 <ZZ> = 736 expectation
values.  The bond dimension ramps 1 -> 32 over the first six steps; from the
seventh on every BP loop runs at D = chi * t = 64.

A bench step is one Floquet step: `--steps K` runs K steps as consecutive
jobs of 20 steps (the last one shorter when K is not a multiple of 20; every
job ends with its 736 values).  `--warmup W` runs W untimed steps of a
warm-up job first, plus (W > 0) one sweep on a 4x4 lattice with all bonds at
64, so that the kernel autotuner has settled the D = 64 shapes before the
clock starts.

`value` = 736 x jobs / wall seconds of the timed steps on the headline path
(`--precision`, default mixed: FP32 MFMA contractions, FP64 messages,
normalisation, residual (floor 1e-6), truncation and ratios; `--schedule`,
default checkerboard: the messages leaving one colour of the bipartite
lattice are updated before those leaving the other, which reaches the same
BP fixed point in about half the sweeps of the reference's Jacobi schedule;
`--accel`, default depth 4: Anderson mixing of the message vector, the next
message set a combination of the last sweeps' outputs, which again reaches
the same fixed point in fewer sweeps).  During the first timed job the 736 values are also taken, untimed (the clock stops), at
every step, to compare with the FP64 leg.

After the headline line is printed (and flushed) further legs run and a
second, complete line is printed:
  * `fp64_fused`: the same job on the FP64 fused path (Jacobi, the
    reference's schedule) within a time budget
    (`--fp64-budget` s): steps completed, per-step times and BP iterations,
    and the mixed path's error against it on all 736 values at every step
    both completed;
  * `plain_fp64`: one BP iteration (sweep + condition) of the plain FP64 GPU
    path (the reference's op sequence one call at a time) at D = 64, from the
    FP64 leg's state, and the job time it implies;
  * `cpu_baseline` (in both lines): the oracle (numpy + the host BLAS) on the
    outgoing messages of one D = 64 interior site, scaled to a job.

Runs on one GPU: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}
KERNELS = {0: "k_mode_product", 1: "k_message_partials", 2: "k_product_message"}
JOB_STEPS = 20
METRIC = "expectation values/s + p99 latency, 16x16 PEPS χ=32 mixed vs FP64 path"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed Floquet steps (20 = one job)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--sweep", choices=["auto", "per-site", "merged"], default="auto")
    ap.add_argument("--schedule", choices=["jacobi", "checkerboard"], default="checkerboard",
                    help="BP update schedule of the headline path (the FP64 leg stays Jacobi, "
                         "the reference's)")
    ap.add_argument("--accel", type=int, default=4,
                    help="Anderson mixing depth of the headline path's BP loop (0: the plain "
                         "fixed-point iteration; the FP64 leg never mixes)")
    ap.add_argument("--accel-start", type=float, default=1e-2,
                    help="residual below which the mixing starts")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-bp-iterations", type=int, default=50,
                    help="the kicked driver's default (apps/kicked_ising_lattice/option.h:19)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--fp64-budget", type=float, default=170.0,
                    help="seconds for the FP64 fused comparison leg (0: skip)")
    ap.add_argument("--plain-sample", type=int, default=1,
                    help="time one BP iteration of the plain FP64 path at D = 64")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--kernel-timing", type=int, default=1,
                    help="time the hot kernels with HIP events carried in their dispatches "
                         "(0 for runs under an external profiler, which times them itself)")
    return ap.parse_args()


def log(obj):
    print(json.dumps(obj), file=sys.stderr, flush=True)


# ------------------------------------------------------------------ workload
class Job:
    """The 16x16 kicked-Ising job's fixed inputs."""

    def __init__(self, args):
        from tnbp_amd import circuits
        self.L, self.chi = args.L, args.chi
        rng = np.random.default_rng(args.seed)
        self.h = float(rng.uniform(0.1, 1.0))
        self.J = math.pi / 4
        self.edges = circuits.square_lattice(self.L, self.L)
        self.sites = sorted({s for e in self.edges for s in e})
        # RZZ(J) = exp(-i J/2 ZZ) and RX(h) = exp(-i h/2 X): dt = 1 (in units of
        # pi), Jz = J / (2 pi), hx = h / (2 pi) in CircuitTPO's parametrisation
        self.tpo = circuits.kicked_ising_tpo(self.edges, self.J / (2 * math.pi), 0.0,
                                             self.h / (2 * math.pi), math.pi)
        self.zop = circuits.op_tensor(circuits.PAULI["Z"])
        self.zz = circuits.op_tensor(np.kron(circuits.PAULI["Z"], circuits.PAULI["Z"]))
        self.n_values = len(self.sites) + len(self.edges)


class Runner:
    """One PEPS on one context, stepping the job."""

    def __init__(self, ctx, job, precision, sweep, schedule="jacobi", accel=0, accel_start=1e-2):
        from tnbp_amd.peps import HBP_PREC_FP64, HBP_PREC_MIXED, InitTensorProductState
        self.ctx, self.job = ctx, job
        prec = HBP_PREC_FP64 if precision == "fp64" else HBP_PREC_MIXED
        self.peps = InitTensorProductState(ctx, job.edges, [2] * len(job.sites), prec)
        self.peps.set_sweep_kernel({"auto": 0, "per-site": 1, "merged": 3}[sweep])
        self.peps.set_schedule({"jacobi": 0, "checkerboard": 1}[schedule])
        self.peps.set_accel(accel, accel_start)
        self.tag = precision

    def step(self, args):
        from tnbp_amd.peps import AbsorbTPO, Truncation, bp_loop
        AbsorbTPO(self.ctx, self.job.tpo, self.peps)
        it, res = bp_loop(self.ctx, self.peps, args.max_bp_iterations, args.tol)
        dims, errs = Truncation(self.ctx, self.peps, self.job.chi, 1e-8)
        return {"bp_iterations": it, "residual": res[-1] if res else None, "max_bond": max(dims),
                "max_trunc_err": max(errs)}

    def measure(self):
        from tnbp_amd.peps import Measure, MeasureEdges
        j = self.job
        z = Measure(self.ctx, self.peps, j.sites, [j.zop] * len(j.sites))
        zz = MeasureEdges(self.ctx, self.peps, j.edges, [j.zz] * len(j.edges))
        return np.array([v.real for v in z] + [v.real for v in zz])

    def close(self):
        self.peps.close()


def settle_autotuner(ctx, precision):
    """One sweep on a 4x4 lattice with every bond at 64: the interior, edge and
    corner shapes of the D = 64 steady state, so that the autotuner's trial
    launches happen before the clock starts."""
    from tnbp_amd.kernel_probe import build_state
    from tnbp_amd.peps import HBP_PREC_FP64, HBP_PREC_MIXED, BeliefPropagation
    p = build_state(ctx, 4, 64, HBP_PREC_FP64 if precision == "fp64" else HBP_PREC_MIXED)
    BeliefPropagation(ctx, p)
    BeliefPropagation(ctx, p)
    ctx.synchronize()
    p.close()


def read_profile(ctx):
    from tnbp_amd._lib import check
    stats = {}
    for kind, name in KERNELS.items():
        n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        check(ctx.lib.hbp_ctx_profile_read(ctx.h, kind, C.byref(n), C.byref(ms), C.byref(fl),
                                           C.byref(by)))
        stats[name] = {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}
    return stats


def run_headline(ctx, job, args):
    """Warm-up, then `args.steps` timed Floquet steps as jobs of 20 steps."""
    from tnbp_amd._lib import check
    if args.warmup > 0:
        w = Runner(ctx, job, args.precision, args.sweep, args.schedule, args.accel,
                   args.accel_start)
        for k in range(args.warmup):
            rec = w.step(args)
            log({"warmup_step": k + 1, **rec})
        w.close()
        settle_autotuner(ctx, args.precision)
        log({"warmup": "autotuner settled at D=64"})
    ctx.synchronize()
    check(ctx.lib.hbp_ctx_profile(ctx.h, 1 if args.kernel_timing else 0))
    steps, jobs, snapshots = [], [], []
    timed = 0.0
    left = args.steps
    while left > 0:
        n = min(JOB_STEPS, left)
        left -= n
        r = Runner(ctx, job, args.precision, args.sweep, args.schedule, args.accel,
                   args.accel_start)
        ctx.synchronize()
        t_job = time.perf_counter()
        paused = 0.0
        for k in range(n):
            t0 = time.perf_counter()
            rec = r.step(args)
            if k == n - 1:
                vals = r.measure()
                ctx.synchronize()
                rec["seconds"] = time.perf_counter() - t0
            else:
                ctx.synchronize()
                rec["seconds"] = time.perf_counter() - t0
                if not jobs:  # first job: untimed snapshot for the FP64 comparison
                    tp = time.perf_counter()
                    snapshots.append(r.measure())
                    paused += time.perf_counter() - tp
            rec["step"] = k + 1
            steps.append(rec)
            log({"path": args.precision, **rec})
        if not jobs:
            snapshots.append(vals)
        wall = time.perf_counter() - t_job - paused
        timed += wall
        jobs.append({"steps": n, "seconds": wall, "values": vals})
        r.close()
    stats = read_profile(ctx)
    check(ctx.lib.hbp_ctx_profile(ctx.h, 0))
    return {"wall": timed, "steps": steps, "jobs": jobs, "snapshots": snapshots, "stats": stats}


def run_fp64_leg(ctx, job, args, want_plain):
    """The same job on the FP64 fused path until the time budget would be
    exceeded; the 736 values at every completed step (outside the step
    times).  Optionally one plain-path BP iteration from the state reached."""
    from tnbp_amd._lib import check
    from tnbp_amd.peps import HBP_MODE_PLAIN, AbsorbTPO, BPIterate
    r = Runner(ctx, job, "fp64", "auto")
    t_leg = time.perf_counter()
    steps, vals = [], []
    last = 0.0
    while len(steps) < JOB_STEPS:
        el = time.perf_counter() - t_leg
        if steps and el + 1.2 * last > args.fp64_budget:
            break
        t0 = time.perf_counter()
        rec = r.step(args)
        ctx.synchronize()
        rec["seconds"] = last = time.perf_counter() - t0
        rec["step"] = len(steps) + 1
        steps.append(rec)
        log({"path": "fp64", **rec})
        vals.append(r.measure())
    plain = None
    if want_plain and steps and steps[-1]["max_bond"] >= job.chi:
        # the next step's absorption, then one iteration (sweep + condition,
        # main.cc:121-141) of the reference's op sequence at D = 64
        AbsorbTPO(ctx, job.tpo, r.peps)
        check(ctx.lib.hbp_peps_set_mode(r.peps.h, HBP_MODE_PLAIN))
        ctx.synchronize()
        t0 = time.perf_counter()
        it, conv, res = BPIterate(ctx, r.peps, 1, args.tol, True, False)
        ctx.synchronize()
        plain = {"seconds_per_bp_iteration": time.perf_counter() - t0, "residual": res[-1],
                 "at_step": len(steps) + 1}
        log({"path": "plain", **plain})
    r.close()
    return {"steps": steps, "values": vals, "plain": plain}


# --------------------------------------------------------------- reporting
def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def blas_threads():
    try:
        from threadpoolctl import threadpool_info
        n = [p["num_threads"] for p in threadpool_info() if p.get("user_api") == "blas"]
        if n:
            return max(n)
    except Exception:
        pass
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_sample(ctx, job, args):
    """A steady-state interior site at D = 64 and its incoming messages
    (bp.h:147-151), from a fresh FP64 run of the job's first 7 steps."""
    from tnbp_amd.peps import AbsorbTPO
    r = Runner(ctx, job, "fp64", "auto")
    while True:
        if r.step(args)["max_bond"] >= job.chi:
            break
    AbsorbTPO(ctx, job.tpo, r.peps)
    L = job.L
    s = (L // 2) * L + L // 2
    inc = [m for m, e in enumerate(job.edges) if s in e]
    ne = len(job.edges)
    sample = (r.peps.get_site(s), [r.peps.get_message(m + ne if job.edges[m][0] == s else m)
                                   for m in inc])
    r.close()
    return sample


def cpu_baseline(sample, iters_per_step, job, args):
    """Time the oracle (numpy + the host BLAS on the host cores) on a bounded
    sample of the same workload: outgoing BP messages of one D = 64 interior
    site with the message/normalisation sequence of bp.h:128-175.  Scaled to a
    job with the site census, the reference's two sweeps per BP iteration
    (sweep + condition, main.cc:121-159) and the given iteration counts of the
    D = 64 steps; ramp steps, truncation and measurement not counted."""
    from oracle import tci as ot
    V, msgs = sample
    z = V.ndim - 1
    t0 = time.perf_counter()
    n_msgs = 0
    while time.perf_counter() - t0 < args.cpu_seconds:
        k = n_msgs % z
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
    dt = time.perf_counter() - t0
    per_msg = dt / n_msgs
    L, D = job.L, V.shape[0]
    cost = {4: D ** 4 * 2 * D * 4, 3: D ** 3 * 2 * D * 3, 2: D ** 2 * 2 * D * 2}
    n4, n3, n2 = (L - 2) ** 2, 4 * (L - 2), 4
    sweep_equiv = (n4 * 4 * cost[4] + n3 * 3 * cost[3] + n2 * 2 * cost[2]) / cost[4]
    job_s = sum(2 * n * sweep_equiv * per_msg for n in iters_per_step)
    return {
        "value": job.n_values / job_s,
        "unit": "expectation values/s",
        "cores": blas_threads(),
        "cpu_model": cpu_info(),
        "kind": "port",
        "extrapolated": True,
        "sample": (f"oracle (numpy + host BLAS, {blas_threads()} threads) computing {n_msgs} outgoing "
                   f"BP messages of one interior site of the {L}x{L} PEPS at D={D} in {dt:.1f}s "
                   f"({per_msg:.2f} s per message); scaled by the site census to one sweep "
                   f"({sweep_equiv:.0f} interior-message equivalents), by 2 sweeps per BP iteration "
                   f"(sweep + condition) and by {sum(iters_per_step)} BP iterations over the job's "
                   f"{len(iters_per_step)} D=64 steps; ramp, truncation and measurement not counted"),
        "job_seconds": job_s,
    }


def roofline(stats, wall, precision):
    """Roofline object of the kernel with the most time in the timed region."""
    name = max(stats, key=lambda k: stats[k]["ms"])
    k = stats[name]
    peak = PEAK["fp64"] if precision == "fp64" else PEAK["fp32"]
    achieved = k["flops"] / (k["ms"] * 1e-3) if k["ms"] > 0 else 0.0
    mfma = "FP64 MFMA 16x16x4" if precision == "fp64" else "FP32 MFMA 16x16x4"
    tr = {}
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f).get(precision, {}).get(name) or {}
    except (OSError, ValueError):
        pass
    return {
        "bound": "mfma",
        "kernel": f"{name} ({mfma})",
        "achieved": achieved / 1e12,
        "peak": peak / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": tr.get("hbm_bytes_per_launch"),
        "traffic_launch": tr.get("launch"),
        "traffic_algorithmic_bytes": tr.get("algorithmic_bytes"),
        "launches": k["launches"],
        "avg_launch_ms": k["ms"] / max(1, k["launches"]),
        "algorithmic_flops_per_launch": k["flops"] / max(1, k["launches"]),
        "algorithmic_bytes_per_launch": k["bytes"] / max(1, k["launches"]),
        "kernel_time_fraction": k["ms"] / (wall * 1e3),
        "other_kernels": {n: {"TFLOPs": (v["flops"] / (v["ms"] * 1e-3) / 1e12 if v["ms"] > 0 else 0.0),
                              "avg_launch_ms": v["ms"] / max(1, v["launches"]),
                              "time_fraction": v["ms"] / (wall * 1e3)}
                          for n, v in stats.items() if n != name},
    }


def d64_counts(steps, chi):
    """BP iteration counts of the steps whose BP loop ran at D = 2 chi: those
    after a step that truncated to chi."""
    return [s["bp_iterations"] for p, s in zip(steps, steps[1:]) if p["max_bond"] >= chi]


def errors(a, b):
    """Mixed-vs-FP64 errors over the 736 values."""
    d = np.abs(a - b)
    mag = np.abs(b)
    big = mag >= 1e-3
    return {"max_abs": float(d.max()), "mean_abs": float(d.mean()),
            "max_rel": float(np.max(d / np.maximum(mag, 1e-300))),
            "max_rel_where_abs_ge_1e-3": float(np.max(d[big] / mag[big])) if big.any() else None,
            "n_values": int(a.size), "n_abs_ge_1e-3": int(big.sum())}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU path is a design "
                                   "(real-space decomposition of the lattice), not measured"}))
        sys.exit(2)
    from tnbp_amd.tci import Context, autotune_report
    job = Job(args)
    ctx = Context(0)
    head = run_headline(ctx, job, args)
    n_jobs = len(head["jobs"])
    job_lat = sorted(j["seconds"] for j in head["jobs"])
    step_lat = sorted(s["seconds"] for s in head["steps"])
    d64_iters = d64_counts(head["steps"][:JOB_STEPS], job.chi)
    mixed = args.precision == "mixed"
    floor = " (mixed: stops once the residual is below max(tol, 1e-6))" if mixed else ""
    out = {
        "metric": METRIC,
        "value": job.n_values * n_jobs / head["wall"],
        "unit": "expectation values/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (MFMA contractions) + f64" if mixed else "f64",
        "data": "synthetic (seeded kicked-Ising circuit, no external data)",
        "config": {"workload": f"kicked-Ising job: {job.L}x{job.L} PEPS, chi={job.chi}, |0..0>, "
                               f"{JOB_STEPS} Floquet steps, J=pi/4, h={job.h:.6f} (seed {args.seed}), "
                               f"BP tol {args.tol:g}{floor}, at most {args.max_bp_iterations} BP "
                               f"iterations, 256 <Z> + 480 <ZZ> at the end of the job; a bench step "
                               f"is one Floquet step",
                   "precision": args.precision, "sweep": args.sweep, "schedule": args.schedule,
                   "bp_mixing_depth": args.accel, "bp_mixing_start": args.accel_start,
                   "L": job.L, "chi": job.chi,
                   "jobs_timed": n_jobs, "values_per_job": job.n_values},
        "latency_s": {"per_job": {"n": n_jobs, "p50": float(np.percentile(job_lat, 50)),
                                  "p99": float(np.percentile(job_lat, 99)), "max": job_lat[-1]},
                      "per_floquet_step": {"n": len(step_lat), "p50": float(np.percentile(step_lat, 50)),
                                           "p99": float(np.percentile(step_lat, 99)),
                                           "max": step_lat[-1]}},
        "bp_iterations": [s["bp_iterations"] for s in head["steps"]],
        "step_seconds": [round(s["seconds"], 3) for s in head["steps"]],
        "roofline": roofline(head["stats"], head["wall"], args.precision),
        "cpu_baseline": None,
    }
    ctx.trim()
    if args.cpu_baseline:
        sample = cpu_sample(ctx, job, args)
        out["cpu_baseline"] = cpu_baseline(sample, d64_iters, job, args)
        del sample
    out["autotune"] = autotune_report().splitlines()
    print(json.dumps(out), flush=True)

    if args.fp64_budget > 0 and mixed:
        ctx.trim()
        leg = run_fp64_leg(ctx, job, args, bool(args.plain_sample))
        fs = leg["steps"]
        k = len(fs)
        t64 = sum(s["seconds"] for s in fs)
        tmx = sum(s["seconds"] for s in head["steps"][:k])
        per_step_err = [errors(head["snapshots"][i], leg["values"][i])
                        for i in range(min(k, len(head["snapshots"])))]
        fp = {"steps_completed": k, "of": JOB_STEPS, "budget_s": args.fp64_budget,
              "bp_iterations": [s["bp_iterations"] for s in fs],
              "residuals": [s["residual"] for s in fs],
              "step_seconds": [round(s["seconds"], 3) for s in fs],
              "seconds_same_steps": {"fp64": t64, "mixed": tmx},
              "speedup_mixed_vs_fp64_same_steps": t64 / tmx if tmx > 0 else None}
        if k == JOB_STEPS:
            fp["value"] = job.n_values / t64
        fp["mixed_error_vs_fp64"] = {"at_step": k, **per_step_err[-1]} if per_step_err else None
        fp["mixed_max_abs_error_per_step"] = [e["max_abs"] for e in per_step_err]
        out["fp64_fused"] = fp
        pl = leg["plain"]
        if pl:
            # a job of the plain path needs at least the iterations the fused
            # FP64 path needed (its residual is the reference's, which also
            # carries the messages' phases); the mixed path's counts stand in
            # for the steps the FP64 leg did not reach (they are lower still)
            lb = d64_counts(fs, job.chi)
            lb = lb + d64_iters[len(lb):]
            plain_job = sum(lb) * pl["seconds_per_bp_iteration"]
            mixed_job = head["jobs"][0]["seconds"]
            out["plain_fp64"] = {**pl, "job_seconds_lower_bound": plain_job,
                                 "value_upper_bound": job.n_values / plain_job,
                                 "mixed_speedup_lower_bound": plain_job / mixed_job,
                                 "note": "job time = BP iterations of the D=64 steps (fused FP64 "
                                         "counts where measured, mixed counts beyond) x the measured "
                                         "time of one plain iteration; ramp, truncation and "
                                         "measurement not counted (a lower bound)"}
        print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
