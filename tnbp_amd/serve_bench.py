"""Serving stream benchmark: Poisson arrivals of sparse-Pauli estimator jobs.

Each job is a random circuit on an L x L lattice (RZZ(theta ~ U[0, pi]) on a
random matching then U3 on every site, per layer) with a random sparse Pauli
observable (weight 1..w_max on contiguous row runs, coefficients ~ N(0, 1)),
drawn from a seed.  Jobs arrive with exponential inter-arrival times at each
swept rate; the server (continuous batching, footprint admission) reports
p50/p99 latency (arrival -> result) and expectation values/s.  One JSON line
per (path, server, rate).  ``--server`` picks the threaded server (one job
per worker thread and context), the layer-batched one (every resident job's
BP loop in one batched call per circuit layer) or both on the same stream.

    python -m tnbp_amd.serve_bench [--jobs 24] [--rates 2,8,32] [--L 6,10]
        [--chi 8,16] [--depth 3,6] [--terms 1,64] [--wmax 6] [--slots 4] [--plain 1]
        [--server threaded|batched|both]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import serving
from ._lib import HBP_PREC_FP64, HBP_PREC_MIXED


def make_jobs(seed, n, Ls, chis, depths, terms, wmax, precision, mode=0):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        L = int(rng.integers(Ls[0], Ls[1] + 1))
        chi = int(rng.choice(chis))
        depth = int(rng.integers(depths[0], depths[1] + 1))
        nt = int(rng.integers(terms[0], terms[1] + 1))
        j = serving.random_circuit_job(rng, L, depth, chi, nt, w_max=wmax, precision=precision,
                                       name=f"j{i}-L{L}-chi{chi}-d{depth}")
        j.mode = mode
        jobs.append(j)
    return jobs


def run_stream(jobs, rate, slots, seed, server="threaded"):
    rng = np.random.default_rng(seed + 1)
    t0 = time.perf_counter()
    make = serving.BatchedServer if server == "batched" else (lambda: serving.Server(slots=slots))
    with make() as srv:
        futs = serving.poisson_submit(srv, jobs, rate, rng)
        for f in futs:  # one progress line per finished job
            f.add_done_callback(lambda f: print(json.dumps(
                {"done": f.result()["name"], "latency_s": round(f.result()["latency"], 3)})
                if f.exception() is None else json.dumps({"failed": str(f.exception())}),
                file=sys.stderr, flush=True))
        res = [f.result() for f in futs]
        st = srv.stats()
    st["wall_s"] = time.perf_counter() - t0
    st["rate_jobs_per_s"] = rate
    if server == "batched":
        st["rounds"] = len(srv.rounds)
        st["mean_jobs_per_round"] = float(np.mean([r["jobs"] for r in srv.rounds]))
        # share of (job, layer) BP loops that ran in the small-bond batched
        # launch, and in the iteration graph with their own sweeps
        nloops = max(1, sum(r["jobs"] for r in srv.rounds))
        st["batched_fraction"] = sum(r["batched"] for r in srv.rounds) / nloops
        st["graph_fraction"] = sum(r.get("graph", 0) for r in srv.rounds) / nloops
        st["phase_seconds"] = {k: round(sum(r[k] for r in srv.rounds), 3)
                               for k in ("absorb_s", "bp_s", "truncate_s", "measure_s")}
    return st, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=24)
    ap.add_argument("--rates", default="2,8,32")
    ap.add_argument("--L", default="6,10")
    ap.add_argument("--chi", default="8,16")
    ap.add_argument("--depth", default="3,6")
    ap.add_argument("--terms", default="1,64")
    ap.add_argument("--wmax", type=int, default=6)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--plain", type=int, default=1, help="also run the plain FP64 path at the lowest rate")
    ap.add_argument("--server", default="threaded", choices=["threaded", "batched", "both"])
    a = ap.parse_args()
    servers = ["threaded", "batched"] if a.server == "both" else [a.server]
    two = lambda s: [int(x) for x in s.split(",")]
    Ls, chis, depths, terms = two(a.L), two(a.chi), two(a.depth), two(a.terms)
    rates = [float(x) for x in a.rates.split(",")]
    ref = None
    for rate in rates:
        for server in servers:
            jobs = make_jobs(a.seed, a.jobs, Ls, chis, depths, terms, a.wmax, HBP_PREC_MIXED)
            st, res = run_stream(jobs, rate, a.slots, a.seed, server)
            if ref is None:
                ref = res
            print(json.dumps({"path": "mixed/fused", "server": server, **st}), flush=True)
    if a.plain:
        jobs = make_jobs(a.seed, a.jobs, Ls, chis, depths, terms, a.wmax, HBP_PREC_FP64, mode=1)
        st, res = run_stream(jobs, rates[0], a.slots, a.seed)
        # per expectation value <P_i> (relative, floor 1: the values lie in
        # [-1, 1]) and per job sum_i c_i <P_i> relative to sum_i |c_i| (the
        # scale of a weighted sum of values that each carry such an error)
        err_v = max(abs(u - v) / max(1.0, abs(v))
                    for x, y in zip(ref, res) for u, v in zip(x["values"], y["values"]))
        err_j = max(abs(x["value"] - y["value"]) / sum(abs(c) for c, _ in j.terms)
                    for x, y, j in zip(ref, res, jobs))
        print(json.dumps({"path": "fp64/plain", "server": "threaded", **st,
                          "max_err_values_mixed_vs_plain": err_v,
                          "max_err_job_sums_mixed_vs_plain": err_j}), flush=True)


if __name__ == "__main__":
    main()
