"""Bounded runs of the four large configurations, checked through properties
that hold at full size.

* 16x16 kicked Ising, chi = 32 (J = pi/4, h from seed 2024 in [0.1, 1]): the
  first eight Floquet steps (the last two with BP at D = 64) on the mixed and
  the FP64 path; all 256 <Z> and 480 <ZZ> of the mixed path within 1e-5 of
  the FP64 path (relative, floor 1: every value lies in [-1, 1]).
* 32x32, chi = 32, the same circuit family: eight steps, all 1024 <Z> of the
  mixed path within 1e-5 of an all-FP64 run (lazily absorbed sites).
* eight seeded circuits of the ragged-batch family (L in 6..20, depth 4..16,
  chi in {8, 16, 32}, 32 Pauli strings of weight 1..4 with N(0, 1)
  coefficients): mixed within 1e-5 of FP64 per term.
* a serving stream with chi in {16, 32, 48} (bonds of 96 during BP), 1..256
  terms of weight up to 6: every job's result equal alone and in the mix
  (1e-12 relative, floor 1), all jobs resident together.
"""
import math

import numpy as np
import pytest

from gpu_util import announce, rel_close

pytestmark = pytest.mark.gpu


def _kicked(ctx, L, chi, steps, precision, sites_only=False):
    from tnbp_amd import circuits
    from tnbp_amd.peps import (AbsorbTPO, InitTensorProductState, Measure, MeasureEdges,
                               Truncation, bp_loop)
    rng = np.random.default_rng(2024)
    h = float(rng.uniform(0.1, 1.0))
    edges = circuits.square_lattice(L, L)
    n = L * L
    tpo = circuits.kicked_ising_tpo(edges, 1 / 8, 0.0, h / (2 * math.pi), math.pi)
    p = InitTensorProductState(ctx, edges, [2] * n, precision)
    its = []
    for k in range(steps):
        AbsorbTPO(ctx, tpo, p)
        it, res = bp_loop(ctx, p, 50, 1e-10)
        its.append(it)
        Truncation(ctx, p, chi, 1e-8)
    z = [v.real for v in Measure(ctx, p, list(range(n)), [circuits.op_tensor(circuits.PAULI["Z"])] * n)]
    zz = [] if sites_only else [v.real for v in MeasureEdges(
        ctx, p, edges, [circuits.op_tensor(np.diag([1.0, -1, -1, 1]))] * len(edges))]
    p.close()
    return np.array(z + zz), its


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def test_16x16_chi32_mixed_within_1e5_of_fp64(ctx):
    announce("16x16 chi=32 kicked Ising, 8 steps: fp64")
    v64, it64 = _kicked(ctx, 16, 32, 8, 0)
    announce(f"  fp64 BP iterations {it64}; mixed")
    vmx, itmx = _kicked(ctx, 16, 32, 8, 1)
    d = np.abs(vmx - v64)
    announce(f"  mixed BP iterations {itmx}; 736 values: max |diff| {d.max():.2e}, "
             f"mean {d.mean():.2e}, max relative {np.max(d / np.abs(v64)):.2e}")
    assert len(v64) == 736
    assert max(it64[-2:]) > 1  # the last steps ran real BP loops at D = 64
    assert rel_close(vmx, v64, 1e-5) <= 1.0


def test_32x32_chi32_mixed_within_1e5_of_fp64(ctx):
    announce("32x32 chi=32 kicked Ising, 8 steps: fp64")
    v64, it64 = _kicked(ctx, 32, 32, 8, 0, sites_only=True)
    announce(f"  fp64 BP iterations {it64}; mixed")
    vmx, itmx = _kicked(ctx, 32, 32, 8, 1, sites_only=True)
    d = np.abs(vmx - v64)
    announce(f"  mixed BP iterations {itmx}; 1024 <Z>: max |diff| {d.max():.2e}, "
             f"mean {d.mean():.2e}, max relative {np.max(d / np.abs(v64)):.2e}")
    assert len(v64) == 1024
    assert rel_close(vmx, v64, 1e-5) <= 1.0


def test_ragged_batch_circuits_mixed_within_1e5_of_fp64(ctx):
    from tnbp_amd import serving
    from tnbp_amd.serve_bench import make_jobs
    jm = make_jobs(41, 8, (6, 20), (8, 16, 32), (4, 16), (32, 32), 4, 1)
    jd = make_jobs(41, 8, (6, 20), (8, 16, 32), (4, 16), (32, 32), 4, 0)
    for a, b in zip(jm, jd):
        b.bp_tolerance = 1e-12
        announce(f"ragged-batch job {a.name}")
        ra, rb = serving.run_job(ctx, a), serving.run_job(ctx, b)
        assert len(ra["values"]) == 32
        w = rel_close(ra["values"], rb["values"], 1e-5)
        announce(f"  mixed vs fp64 at {w:.3f} of the tolerance")
        assert w <= 1.0, a.name


def test_serving_stream_chi48_alone_equals_mix():
    from tnbp_amd import serving
    from tnbp_amd.serve_bench import make_jobs
    from tnbp_amd.tci import Context
    jobs = lambda: make_jobs(5, 6, (8, 11), (16, 32, 48), (3, 5), (1, 256), 6, 1)
    chis = [j.chi for j in jobs()]
    assert 48 in chis
    alone = []
    c = Context(0)
    try:
        for j in jobs():
            announce(f"serving job alone {j.name}")
            alone.append(serving.run_job(c, j))
    finally:
        c.close()
    announce("serving stream: all jobs at once")
    with serving.Server(slots=6) as srv:
        futs = [srv.submit(j) for j in jobs()]
        mix = [f.result(timeout=600) for f in futs]
        st = srv.stats()
    announce(f"  stats {st}")
    assert st["max_concurrent"] >= 2
    for a, m in zip(alone, mix):
        assert a["bp_iterations"] == m["bp_iterations"]
        assert rel_close(m["values"], a["values"], 1e-12) <= 1.0
