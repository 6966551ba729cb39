"""The kernels of the 16x16 chi=32 steady state at their real shape (bonds of
64) against the oracle's golden fixture tests/golden/d64_3x3.json
(tests/golden/make_d64.py).

A 3x3 lattice with every bond at D = 64 has one interior site of the
steady-state shape (64, 64, 64, 64, 2), four edge sites (64, 64, 64, 2) and
four corners.  Its sweeps run the merged product-message kernel on whole
64 x 64 tiles (`k_product_message<S, *, true>`), the pair Gram kernel on whole
tiles (`k_pair_gram<S, true>`, first sweep from diagonal messages), the
whole-tile mode products (the FP32 buffer form where the autotuner picks it;
every launch variant is bitwise equal to the others, tests/test_gpu_variants.py)
and, in Truncation(32), the 64 -> 32 projector products.

Tolerances: FP64 messages within 1e-10 of the oracle's (each probe
u^H (M_gpu - M_oracle) v and the norm difference against ||M_oracle||_F), the
residual within 1e-10 relative, Lambda within 1e-10 relative and <Z>, <ZZ>
after Truncation(32) within 1e-10 relative (floor 1); the mixed path within
1e-5 on the same quantities.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from gpu_util import announce, rel_close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    import make_d64  # tests/golden on sys.path below
    with open(os.path.join(GOLDEN, "d64_3x3.json")) as f:
        return json.load(f), make_d64


@pytest.fixture(scope="module")
def gold():
    import sys
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    return _golden()


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def _gpu_state(ctx, mk, precision, diag):
    from tnbp_amd._lib import check, i32
    from tnbp_amd.peps import Peps
    V, E = mk.inputs(diag)
    n = mk.L * mk.L
    p = Peps(ctx, mk.EDGES, [2] * n, precision)
    check(ctx.lib.hbp_peps_set_bond_dims(ctx.h, p.h, i32([mk.D] * mk.NE), i32([2] * n)))
    for s, v in enumerate(V):
        p.set_site(s, v)
    for m, e in enumerate(E):
        p.set_message(m, e)
    return p


def _message_error(mk, M, want):
    """max over the probes and the norm of |difference| / ||M_oracle||_F."""
    got = mk.summary(M)
    nrm = want[0]
    err = abs(got[0] - nrm)
    for k in range(4):
        err = max(err, abs(complex(got[1 + 2 * k], got[2 + 2 * k])
                           - complex(want[1 + 2 * k], want[2 + 2 * k])))
    return err / nrm


CASES = [  # precision, sweep kernel, tolerance
    (0, "auto", 1e-10),
    (0, "per-site", 1e-10),
    (1, "auto", 1e-5),
]


@pytest.mark.parametrize("diag", [True, False], ids=["diag-env", "dense-env"])
@pytest.mark.parametrize("precision,sweep,tol", CASES)
def test_d64_sweeps_truncation_measure_match_oracle(ctx, gold, precision, sweep, tol, diag):
    from tnbp_amd._lib import check
    from tnbp_amd.peps import (HBP_SWEEP_AUTO, HBP_SWEEP_PER_SITE, BeliefPropagation,
                               BeliefPropagationCondition, Measure, MeasureEdges, Truncation)
    g, mk = gold
    want = g["cases"]["diag" if diag else "dense"]
    announce(f"D=64 3x3 precision={precision} sweep={sweep} diag={diag}")
    p = _gpu_state(ctx, mk, precision, diag)
    p.set_sweep_kernel(HBP_SWEEP_AUTO if sweep == "auto" else HBP_SWEEP_PER_SITE)
    counts = []
    for it in range(2):
        check(ctx.lib.hbp_ctx_profile(ctx.h, 1))
        BeliefPropagation(ctx, p)
        c = []
        for k in (0, 1, 2):
            n = ctypes.c_int64()
            check(ctx.lib.hbp_ctx_profile_read(ctx.h, k, ctypes.byref(n), None, None, None))
            c.append(n.value)
        check(ctx.lib.hbp_ctx_profile(ctx.h, 0))
        counts.append(c)
        worst = max(_message_error(mk, p.get_message(m), want["messages"][it][m])
                    for m in range(2 * mk.NE))
        r = BeliefPropagationCondition(ctx, p)
        announce(f"  sweep {it}: worst message error {worst:.2e}, residual {r:.9e} vs "
                 f"{want['residuals'][it]:.9e}, launches (mode product, message, merged) {c}")
        assert worst <= tol, (it, worst)
        assert abs(r - want["residuals"][it]) <= tol * want["residuals"][it]
    if sweep == "auto":
        # the interior and edge sites' messages ran merged once the environment
        # was dense; a diagonal first sweep ran no mode products at all
        assert counts[1][2] >= 4
        if diag:
            assert counts[0][0] == 0 and counts[0][2] == 0
    else:
        assert counts[0][2] == 0 and counts[1][2] == 0
    dims, errs = Truncation(ctx, p, g["chi"], g["sv_min"])
    assert list(dims) == list(want["bond_dims"])
    for m in range(mk.NE):
        lam = np.real(np.diag(p.get_message(m)))
        np.testing.assert_allclose(lam, want["lambda"][m], rtol=tol, atol=tol * want["lambda"][m][0])
    n = mk.L * mk.L
    z = [v.real for v in Measure(ctx, p, list(range(n)), [mk.ZT] * n)]
    zz = [v.real for v in MeasureEdges(ctx, p, mk.EDGES, [mk.ZZT] * mk.NE)]
    wz, wzz = rel_close(z, want["z"], tol), rel_close(zz, want["zz"], tol)
    announce(f"  after Truncation(32): <Z> at {wz:.3f}, <ZZ> at {wzz:.3f} of the tolerance")
    assert wz <= 1.0 and wzz <= 1.0
    p.close()
