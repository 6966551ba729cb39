"""The layer-batched server on the GPU.

* A job's result from the server in a mix of jobs (several sizes, bonds above
  and below the batched sweep's limit of 16, staggered arrivals) equals the
  same job run alone through the batched BP loop (``run_job_batched``): the
  stated tolerance is 1e-12 relative, equality is expected (each site's
  messages come from one workgroup in a fixed order).
* In FP64 the server's results agree with the per-site fused path
  (``run_job``) to 1e-10 relative (floor 1), with the same BP iteration counts.
* A job that arrives while others are resident joins at the next layer
  boundary (a round holds more than one job).
"""
import time

import numpy as np
import pytest

from gpu_util import announce, rel_close

pytestmark = pytest.mark.gpu


def _jobs(precision):
    from tnbp_amd import serving
    rng = np.random.default_rng(23)
    # (L, depth, chi): the chi = 16 jobs absorb to bonds of 32 and run their
    # own fused sweeps inside the batched call
    spec = [(3, 3, 4), (4, 4, 8), (5, 3, 4), (4, 3, 16), (3, 5, 8), (6, 2, 8)]
    jobs = [serving.random_circuit_job(rng, L, depth, chi, 8, w_max=4, precision=precision,
                                       name=f"job{i}-L{L}-d{depth}-chi{chi}")
            for i, (L, depth, chi) in enumerate(spec)]
    for j in jobs:
        j.bp_tolerance = 1e-12 if precision == 0 else 1e-8
        j.max_bp_iterations = 100
    return jobs


def _alone(runner, jobs):
    from tnbp_amd.tci import Context
    ctx = Context(0)
    try:
        return [runner(ctx, j) for j in jobs]
    finally:
        ctx.close()


def _served(jobs, stagger):
    from tnbp_amd import serving
    with serving.BatchedServer() as srv:
        futs = []
        for i, j in enumerate(jobs):
            futs.append(srv.submit(j))
            if i == 2 and stagger == "after-first":
                futs[0].result(timeout=300)
            elif i == 2:
                time.sleep(stagger)
        res = [f.result(timeout=300) for f in futs]
        return res, srv.stats(), list(srv.rounds)


@pytest.mark.parametrize("precision", [1, 0], ids=["mixed", "fp64"])
def test_mix_equals_alone(precision):
    from tnbp_amd import serving
    announce(f"batched server: mix vs alone, precision {precision}")
    alone = _alone(serving.run_job_batched, _jobs(precision))
    mix, st, rounds = _served(_jobs(precision), 0.05)
    announce(f"  stats {st}; jobs per round {[r['jobs'] for r in rounds]}")
    assert max(r["jobs"] for r in rounds) >= 2
    assert any(r["batched"] >= 2 for r in rounds)
    for a, m in zip(alone, mix):
        assert a["bp_iterations"] == m["bp_iterations"], m["name"]
        assert rel_close(m["values"], a["values"], 1e-12) <= 1.0, m["name"]


def test_fp64_server_matches_fused_path():
    from tnbp_amd import serving
    announce("batched server vs run_job, FP64")
    ref = _alone(serving.run_job, _jobs(0))
    got, st, rounds = _served(_jobs(0), "after-first")
    announce(f"  jobs per round {[r['jobs'] for r in rounds]}")
    for a, m in zip(ref, got):
        announce(f"  {m['name']}: iterations {m['bp_iterations']} vs {a['bp_iterations']}")
        assert a["bp_iterations"] == m["bp_iterations"], m["name"]
        assert rel_close(m["values"], a["values"], 1e-10) <= 1.0, m["name"]
