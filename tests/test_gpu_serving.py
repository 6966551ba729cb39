"""Serving on the GPU: every job's result is the same alone and in the mix
(stated tolerance 1e-12 relative; the kernels are deterministic, so equality
is expected), admission by footprint still gives correct results, and the
mixed-precision result of a job is within 1e-5 of its FP64 result."""
import numpy as np
import pytest

from gpu_util import announce, rel_close

pytestmark = pytest.mark.gpu


def _jobs(precision):
    from tnbp_amd import serving
    rng = np.random.default_rng(12)
    spec = [(3, 2, 4), (4, 3, 8), (5, 2, 4), (4, 2, 8), (3, 3, 8)]
    return [serving.random_circuit_job(rng, L, depth, chi, 8, w_max=4, precision=precision,
                                       name=f"job{i}")
            for i, (L, depth, chi) in enumerate(spec)]


def test_alone_equals_mix():
    from tnbp_amd import serving
    from tnbp_amd.tci import Context
    announce("serving alone vs mix")
    alone = []
    ctx = Context(0)
    try:
        for job in _jobs(1):
            alone.append(serving.run_job(ctx, job))
    finally:
        ctx.close()
    with serving.Server(slots=4) as srv:
        futs = [srv.submit(j) for j in _jobs(1)]
        mix = [f.result(timeout=300) for f in futs]
        st = srv.stats()
    announce(f"serving stats {st}")
    assert st["max_concurrent"] >= 2
    for a, m in zip(alone, mix):
        assert a["bp_iterations"] == m["bp_iterations"]
        assert rel_close(m["values"], a["values"], 1e-12) <= 1.0
        assert abs(m["value"] - a["value"]) <= 1e-12 * max(1.0, abs(a["value"]))


def test_admission_one_at_a_time_is_correct():
    from tnbp_amd import serving
    announce("serving tight budget")
    jobs = _jobs(1)
    biggest = max(serving.footprint_bytes(j) for j in jobs)
    with serving.Server(slots=3, memory_budget=biggest) as srv:
        futs = [srv.submit(j) for j in jobs]
        res = [f.result(timeout=300) for f in futs]
        st = srv.stats()
    assert st["jobs"] == len(jobs)
    with serving.Server(slots=1) as srv:
        ref = [srv.submit(j).result(timeout=300) for j in _jobs(1)]
    for a, b in zip(res, ref):
        assert rel_close(a["values"], b["values"], 1e-12) <= 1.0


def test_mixed_within_1e5_of_fp64():
    from tnbp_amd import serving
    from tnbp_amd.tci import Context
    announce("serving mixed vs fp64")
    ctx = Context(0)
    try:
        for jm, jd in zip(_jobs(1)[:3], _jobs(0)[:3]):
            # the same circuit and terms in both precisions, FP64 to the FP64 tolerance
            jd.bp_tolerance = 1e-12
            rm, rd = serving.run_job(ctx, jm), serving.run_job(ctx, jd)
            assert rel_close(rm["values"], rd["values"], 1e-5) <= 1.0, (jm.name,)
    finally:
        ctx.close()


def test_footprint_covers_the_peak_of_a_merged_sweep_job():
    """footprint_bytes (the admission model) is at least the peak device
    memory the job's context reports, for a job whose bonds reach 64 (merged
    sweep with pair-front copies, counted in the footprint) and one at 16."""
    from tnbp_amd import serving
    from tnbp_amd.tci import Context
    rng = np.random.default_rng(31)
    jobs = [serving.kicked_job(3, 32, 7, 0.7, precision=1, name="kicked-3x3-chi32"),
            serving.random_circuit_job(rng, 5, 4, 8, 8, w_max=3, precision=1, name="random-5x5-chi8")]
    assert serving.pair_cache_bytes(jobs[0]) > 0
    for job in jobs:
        ctx = Context(0)
        try:
            ctx.trim()
            serving.run_job(ctx, job)
            _, peak = ctx.memory()
        finally:
            ctx.close()
        fp = serving.footprint_bytes(job)
        announce(f"footprint {job.name}: model {fp / 2**20:.1f} MiB, peak {peak / 2**20:.1f} MiB, "
                 f"pair cache {serving.pair_cache_bytes(job) / 2**20:.1f} MiB")
        assert peak <= fp
