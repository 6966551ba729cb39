"""TCI layer on the GPU (hbp_contract & co.) against the oracle's restatement of min-tci.

Tolerance: complex128 results agree with the oracle to 1e-12 relative to the
operands' scale (they differ by summation order only).
"""
import numpy as np
import pytest

from gpu_util import announce
from oracle import tci as ot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from tnbp_amd.tci import Context
    c = Context(0)
    yield c
    c.close()


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _contract(ctx, a, la, b, lb, lc):
    from tnbp_amd import tci as T
    ta, tb = T.Tensor.from_numpy(ctx, a), T.Tensor.from_numpy(ctx, b)
    return T.contract(ctx, ta, la, tb, lb, lc).to_numpy()


CASES = [
    # mode products at every bond position (bp.h:148), square and projector-shaped E
    ((8, 8, 8, 8, 2), [-1, 1, 2, 3, 4], (8, 8), [-1, 0], [0, 1, 2, 3, 4]),
    ((8, 8, 8, 8, 2), [0, -1, 2, 3, 4], (8, 8), [-1, 1], [0, 1, 2, 3, 4]),
    ((16, 16, 16, 2), [0, 1, -1, 3], (16, 16), [-1, 2], [0, 1, 2, 3]),
    ((64, 64, 2), [0, -1, 2], (64, 64), [-1, 1], [0, 1, 2]),
    ((64, 32, 2), [-1, 1, 2], (64, 32), [-1, 0], [0, 1, 2]),          # projector D -> chi
    ((6, 5, 7, 3), [0, 1, -1, 3], (5, 7), [2, -1], [0, 1, 2, 3]),     # shared bond is E's axis 1
    ((3, 4, 5), [0, 1, -1], (5, 100), [-1, 2], [0, 1, 2]),
    ((96, 96, 2), [-1, 1, 2], (96, 96), [-1, 0], [0, 1, 2]),
    # all-but-one message contraction (bp.h:165)
    ((8, 8, 8, 2), [0, -2, -3, -4], (8, 8, 8, 2), [1, -2, -3, -4], [0, 1]),
    ((8, 8, 8, 2), [-1, -2, 0, -4], (8, 8, 8, 2), [-1, -2, 1, -4], [0, 1]),
    ((32, 32, 32, 2), [-1, 0, -3, -4], (32, 32, 32, 2), [-1, 1, -3, -4], [0, 1]),
    ((5, 7, 3), [-1, 0, -3], (5, 7, 3), [-1, 1, -3], [1, 0]),
    # full contraction (measure.h:83)
    ((4, 5, 6, 2), [-1, -2, -3, -4], (4, 5, 6, 2), [-1, -2, -3, -4], []),
    # general: edge environment (measure.h:204), AbsorbTPO interleave (apply.h:64), outer product
    ((3, 4, 5, 2), [-1, 2, -3, 0], (3, 4, 5, 2), [-1, 3, -3, 1], [0, 1, 2, 3]),
    ((2, 3, 2, 2), [1, 3, -1, 5], (4, 5, 2), [0, 2, -1], [0, 1, 2, 3, 5]),
    ((4, 4), [0, 2], (3, 3), [1, 3], [0, 1, 2, 3]),
    ((150, 7), [0, -1], (7, 140), [-1, 1], [0, 1]),
    ((3, 4), [-1, 0], (3, 5), [-1, 1], [1, 0]),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_contract(ctx, case):
    sa, la, sb, lb, lc = case
    announce(f"contract {sa} {la} x {sb} {lb} -> {lc}")
    rng = np.random.default_rng(hash(str(case)) % 2**32)
    a, b = _rand(rng, sa), _rand(rng, sb)
    got = _contract(ctx, a, la, b, lb, lc)
    want = ot.contract(a, la, b, lb, lc)
    scale = np.sqrt(np.sum(np.abs(a) ** 2) * np.sum(np.abs(b) ** 2))
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) <= 1e-12 * scale


def test_contract_errors(ctx):
    from tnbp_amd import tci as T
    a = T.Tensor.from_numpy(ctx, np.ones((2, 2), dtype=complex))
    with pytest.raises(ValueError):
        T.contract(ctx, a, [0, 0], a, [1, 2], [0, 1, 2])
    with pytest.raises(ValueError):
        T.contract(ctx, a, [0, -1], a, [-1, 1], [0, 2])


def test_contract_in_place_alias(ctx):
    from tnbp_amd import tci as T
    rng = np.random.default_rng(8)
    a, e = _rand(rng, (4, 5, 2)), _rand(rng, (5, 5))
    ta, te = T.Tensor.from_numpy(ctx, a), T.Tensor.from_numpy(ctx, e)
    T.contract(ctx, ta, [0, -1, 2], te, [-1, 1], [0, 1, 2], out=ta)
    np.testing.assert_allclose(ta.to_numpy(), ot.contract(a, [0, -1, 2], e, [-1, 1], [0, 1, 2]), atol=1e-12)


def test_elementwise_and_reductions(ctx):
    from tnbp_amd import tci as T
    announce("elementwise")
    rng = np.random.default_rng(9)
    a = _rand(rng, (3, 4, 5))
    t = T.Tensor.from_numpy(ctx, a)
    assert abs(T.norm(ctx, t) - ot.norm(a)) < 1e-12 * ot.norm(a)
    n = T.normalize(ctx, t)
    assert abs(n - ot.norm(a)) < 1e-12 * n
    np.testing.assert_allclose(t.to_numpy(), a / ot.norm(a), atol=1e-14)
    T.cplx_conj(ctx, t)
    np.testing.assert_allclose(t.to_numpy(), np.conj(a) / ot.norm(a), atol=1e-14)
    t2 = T.Tensor.from_numpy(ctx, a)
    T.transpose(ctx, t2, [2, 0, 1])
    np.testing.assert_array_equal(t2.to_numpy(), np.transpose(a, [2, 0, 1]))
    T.reshape(ctx, t2, [5, 12])
    np.testing.assert_array_equal(t2.to_numpy(), ot.reshape(np.transpose(a, [2, 0, 1]), (5, 12)))
    b = _rand(rng, (3, 4, 5))
    lc = T.linear_combine(ctx, [T.Tensor.from_numpy(ctx, a), T.Tensor.from_numpy(ctx, b)], [1.0, -1.0 + 0.5j])
    np.testing.assert_allclose(lc.to_numpy(), a + (-1.0 + 0.5j) * b, atol=1e-13)
    c = _rand(rng, (2, 2, 3, 3))
    tr = T.trace(ctx, T.Tensor.from_numpy(ctx, c), [(0, 1)])
    np.testing.assert_allclose(tr.to_numpy(), ot.trace(c, [(0, 1)]), atol=1e-13)
    v = T.Tensor.from_numpy(ctx, np.array([1.0, 2.0, 3.0]))
    vc = T.to_cplx(ctx, v)
    T.diag(ctx, vc)
    np.testing.assert_allclose(vc.to_numpy(), np.diag([1.0, 2.0, 3.0]), atol=0)
    assert T.get_elem(ctx, T.Tensor.from_numpy(ctx, a), [1, 2, 3]) == a[1, 2, 3]


# (64, 64): W and V in LDS; (96, 96): W in LDS, V in global memory; (128, 100):
# both in global memory
@pytest.mark.parametrize("shape,nrows", [((6, 6), 1), ((4, 3, 5), 2), ((64, 64), 1), ((96, 96), 1),
                                         ((128, 100), 1), ((2, 7), 1), ((9, 4), 1)])
def test_svd(ctx, shape, nrows):
    from tnbp_amd import tci as T
    announce(f"svd {shape} {nrows}")
    rng = np.random.default_rng(10)
    a = _rand(rng, shape)
    u, s, vd = T.svd(ctx, T.Tensor.from_numpy(ctx, a), nrows)
    U, S, Vd = u.to_numpy(), s.to_numpy(), vd.to_numpy()
    _, s_ref, _ = ot.svd(a, nrows)
    np.testing.assert_allclose(S, s_ref, rtol=1e-12, atol=1e-13 * s_ref[0])
    k = S.size
    rec = ot.contract(ot.contract(U, list(range(nrows)) + [-1], np.diag(S).astype(complex), [-1, 100], list(range(nrows)) + [100]),
                      list(range(nrows)) + [-1], Vd, [-1] + list(range(nrows, a.ndim)), list(range(a.ndim)))
    np.testing.assert_allclose(rec, a, atol=1e-12 * s_ref[0])
    Um = U.reshape(-1, k, order="F")
    np.testing.assert_allclose(Um.conj().T @ Um, np.eye(k), atol=1e-12)


def test_svd_rank_deficient(ctx):
    """Rank-deficient matrices, the shape of truncation's Ra^T Rb when the
    messages have low rank: the Jacobi rotations turn some columns into
    rounding noise, and pairs of such columns count as converged; rotating
    them by the relative orthogonality test alone exhausts the sweep limit
    and leaves singular values wrong by up to 1e-5 on such inputs.  Sizes
    cover the kernel's three memory forms (up to 64 x 64 in LDS, 80 x 80 with
    W in LDS, 120 x 120 in global memory)."""
    from tnbp_amd import tci as T
    announce("svd of rank-deficient matrices (all three memory forms)")
    rng = np.random.default_rng(7)
    shapes = ([(n, int(rng.integers(1, n))) for n in rng.integers(3, 9, size=40)]
              + [(80, 20), (72, 5), (120, 30)])
    for i, (n, r) in enumerate(shapes):
        a = _rand(rng, (n, r)) @ _rand(rng, (r, n))
        if i % 2:
            a[rng.integers(0, n), :] = 0
        a /= np.linalg.norm(a)
        u, s, vd = T.svd(ctx, T.Tensor.from_numpy(ctx, a), 1)
        U, S, Vd = u.to_numpy(), s.to_numpy(), vd.to_numpy()
        _, s_ref, _ = ot.svd(a, 1)
        np.testing.assert_allclose(S, s_ref, rtol=0, atol=1e-13, err_msg=f"matrix {i} ({n}x{n}, rank {r})")
        np.testing.assert_allclose(U @ np.diag(S) @ Vd, a, rtol=0, atol=1e-13)
        Ur = U[:, :r]
        np.testing.assert_allclose(Ur.conj().T @ Ur, np.eye(r), atol=1e-12)


def test_trunc_svd_rule(ctx):
    from tnbp_amd import tci as T
    rng = np.random.default_rng(3)
    q1, _ = np.linalg.qr(_rand(rng, (6, 6)))
    q2, _ = np.linalg.qr(_rand(rng, (6, 6)))
    s_true = np.array([1.0, 0.5, 0.25, 1e-3, 1e-9, 0.0])
    m = q1 @ np.diag(s_true) @ q2
    u, s, vd, err = T.trunc_svd(ctx, T.Tensor.from_numpy(ctx, m), 1, 3, 1e-8)
    np.testing.assert_allclose(s.to_numpy(), s_true[:3], rtol=1e-12)
    assert abs(err - (1e-6 + 1e-18) / np.sum(s_true ** 2)) < 1e-12
    u, s, vd, err = T.trunc_svd(ctx, T.Tensor.from_numpy(ctx, m), 1, 10, 1e-8)
    assert s.to_numpy().size == 4


def test_context_block_cache_accounting():
    """Freed blocks are kept by size class and handed out again: hbp_ctx_memory
    counts live blocks only (at their class size: powers of two from 256 B
    below 32 MiB, 2 MiB steps above), the same sizes come back without
    growing the live count, the data of a reused block is the new tensor's,
    and trim returns everything while keeping live tensors intact."""
    from tnbp_amd import tci as T
    announce("context block cache accounting")
    c = T.Context(0)
    try:
        base, _ = c.memory()
        rng = np.random.default_rng(5)
        sizes = [(3,), (100,), (1000, 3), (300, 1000), (2200, 1000)]  # up to 35 MB complex128
        expect = 0
        for shp in sizes:
            n = int(np.prod(shp)) * 16
            expect += 256 if n <= 256 else (1 << (n - 1).bit_length()) if n < (32 << 20) \
                else -(-n // (2 << 20)) * (2 << 20)
        for rnd in range(3):
            ts = [T.Tensor.from_numpy(c, _rand(rng, s)) for s in sizes]
            live, _ = c.memory()
            assert live - base == expect, (rnd, live - base, expect)
            # a reused block holds exactly what was uploaded into it
            for s in sizes:
                a = _rand(rng, s)
                t2 = T.Tensor.from_numpy(c, a)
                np.testing.assert_array_equal(t2.to_numpy(), a)
                del t2
            del ts
            c.synchronize()
            live, _ = c.memory()
            assert live == base
        keep = T.Tensor.from_numpy(c, np.arange(12.0).reshape(3, 4) + 0j)
        c.trim()
        np.testing.assert_array_equal(keep.to_numpy(), np.arange(12.0).reshape(3, 4) + 0j)
        del keep
    finally:
        c.close()


def test_single_precision_and_real_tensors(ctx):
    """Element types of the drivers' USE_SINGLE build (complex<float>, real
    float) and real tensors through contract, svd, trunc_svd, trace and
    linear_combine: the operands are taken exactly to complex128, so each
    result equals the oracle's on the same values up to the final rounding
    (1e-6 relative for single precision, 1e-12 for double); the result type
    follows the operands (complex if either is, single if both are)."""
    from tnbp_amd import tci as T
    from tnbp_amd._lib import HBP_C64, HBP_C128, HBP_F32, HBP_F64
    rng = np.random.default_rng(41)
    a = _rand(rng, (8, 6, 5, 2)).astype(np.complex64)
    e = _rand(rng, (6, 7)).astype(np.complex64)
    ta, te = T.Tensor.from_numpy(ctx, a), T.Tensor.from_numpy(ctx, e)
    assert ta.dtype == HBP_C64
    got = T.contract(ctx, ta, [0, -1, 2, 3], te, [-1, 1], [0, 1, 2, 3])
    assert got.dtype == HBP_C64
    want = ot.contract(a.astype(np.complex128), [0, -1, 2, 3], e.astype(np.complex128), [-1, 1],
                       [0, 1, 2, 3])
    assert np.max(np.abs(got.to_numpy() - want)) <= 1e-6 * np.max(np.abs(want))
    # message-shaped contraction of a single tensor with a double one -> complex128
    b = _rand(rng, (8, 6, 5, 2))
    got = T.contract(ctx, ta, [0, -1, -2, -3], T.Tensor.from_numpy(ctx, b), [1, -1, -2, -3], [0, 1])
    assert got.dtype == HBP_C128
    want = ot.contract(a.astype(np.complex128), [0, -1, -2, -3], b, [1, -1, -2, -3], [0, 1])
    assert np.max(np.abs(got.to_numpy() - want)) <= 1e-12 * np.max(np.abs(want))
    # real x real stays real
    r1, r2 = rng.standard_normal((5, 4)), rng.standard_normal((4, 3))
    got = T.contract(ctx, T.Tensor.from_numpy(ctx, r1), [0, -1], T.Tensor.from_numpy(ctx, r2), [-1, 1],
                     [0, 1])
    assert got.dtype == HBP_F64
    np.testing.assert_allclose(got.to_numpy(), r1 @ r2, atol=1e-13)
    got = T.contract(ctx, T.Tensor.from_numpy(ctx, r1.astype(np.float32)), [0, -1],
                     T.Tensor.from_numpy(ctx, r2.astype(np.float32)), [-1, 1], [0, 1])
    assert got.dtype == HBP_F32
    np.testing.assert_allclose(got.to_numpy(), r1.astype(np.float32).astype(np.float64)
                               @ r2.astype(np.float32).astype(np.float64), rtol=1e-6, atol=1e-6)
    # svd / trunc_svd of a complex64 matrix: complex64 factors, float32 values
    m = _rand(rng, (40, 24)).astype(np.complex64)
    u, s, vd = T.svd(ctx, T.Tensor.from_numpy(ctx, m), 1)
    assert (u.dtype, s.dtype, vd.dtype) == (HBP_C64, HBP_F32, HBP_C64)
    U, S, Vd = u.to_numpy(), s.to_numpy(), vd.to_numpy()
    rec = (U.astype(np.complex128) * S.astype(np.float64)) @ Vd.astype(np.complex128)
    assert np.max(np.abs(rec - m)) <= 1e-5 * np.max(np.abs(m))
    np.testing.assert_allclose(S, np.linalg.svd(m.astype(np.complex128), compute_uv=False),
                               rtol=1e-5)
    u, s, vd, err = T.trunc_svd(ctx, T.Tensor.from_numpy(ctx, m), 1, 10, 0.0)
    assert (u.dtype, s.dtype, vd.dtype) == (HBP_C64, HBP_F32, HBP_C64) and s.shape == (10,)
    # trace and linear_combine keep complex64
    t = _rand(rng, (4, 3, 4)).astype(np.complex64)
    got = T.trace(ctx, T.Tensor.from_numpy(ctx, t), [(0, 2)])
    assert got.dtype == HBP_C64
    np.testing.assert_allclose(got.to_numpy(), np.einsum("iji->j", t.astype(np.complex128)),
                               rtol=1e-6, atol=1e-6)
    x, y = _rand(rng, (3, 5)).astype(np.complex64), _rand(rng, (3, 5)).astype(np.complex64)
    got = T.linear_combine(ctx, [T.Tensor.from_numpy(ctx, x), T.Tensor.from_numpy(ctx, y)],
                           [0.5 - 1j, 2.0])
    assert got.dtype == HBP_C64
    np.testing.assert_allclose(got.to_numpy(), (0.5 - 1j) * x.astype(np.complex128) + 2.0 * y,
                               rtol=1e-6, atol=1e-6)
    announce("single-precision and real tensors through contract / svd / trace / linear_combine")
