"""Pins of the TCI restatement (oracle/tci.py) by brute force and closed forms."""
import itertools

import numpy as np
import pytest

from oracle import tci


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _brute_contract(a, la, b, lb, lc):
    """Nested-loop einsum: every label assignment, multiply-accumulate."""
    dims = {}
    for lab, n in list(zip(la, a.shape)) + list(zip(lb, b.shape)):
        dims[lab] = n
    labels = sorted(dims)
    out = np.zeros([dims[x] for x in lc], dtype=complex)
    for vals in itertools.product(*[range(dims[x]) for x in labels]):
        env = dict(zip(labels, vals))
        out[tuple(env[x] for x in lc)] += a[tuple(env[x] for x in la)] * b[tuple(env[x] for x in lb)]
    return out


@pytest.mark.parametrize("case", [
    # mode-k product of bp.h:148 (T x_k E), order 4, k = 1
    ((2, 3, 4, 2), [0, -1, 2, 3], (3, 5), [-1, 1], [0, 1, 2, 3]),
    # all-but-one message contraction of bp.h:165
    ((2, 3, 4, 2), [-4, 0, -2, -1], (2, 3, 4, 2), [-4, 1, -2, -1], [0, 1]),
    # edge environment of measure.h:204 (phys and target kept)
    ((2, 3, 2), [-3, 2, 0], (2, 3, 2), [-3, 3, 1], [0, 1, 2, 3]),
    # full contraction to order 0 (measure.h:83)
    ((2, 3), [-2, -1], (2, 3), [-2, -1], []),
    # AbsorbTPO label pattern (apply.h:52-66), order-3 site with two bonds
    ((2, 3, 2, 2), [1, 3, -1, 5], (4, 5, 2), [0, 2, -1], [0, 1, 2, 3, 5]),
    # transposed output order
    ((3, 4), [0, -1], (4, 2), [-1, 1], [1, 0]),
])
def test_contract_matches_brute_force(case):
    rng = np.random.default_rng(1)
    sa, la, sb, lb, lc = case
    a, b = _rand(rng, sa), _rand(rng, sb)
    np.testing.assert_allclose(tci.contract(a, la, b, lb, lc),
                               _brute_contract(a, la, b, lb, lc), rtol=1e-13, atol=1e-13)


def test_contract_label_errors():
    a = np.zeros((2, 2))
    with pytest.raises(ValueError):
        tci.contract(a, [0, 0], a, [1, 2], [0, 1, 2])
    with pytest.raises(ValueError):
        tci.contract(a, [0, -1], a, [-1, 1], [0, 2])


def test_reshape_is_column_major():
    a = np.arange(24).reshape((2, 3, 4), order="F")
    # element (i, j, k) sits at address i + 2 j + 6 k
    r = tci.reshape(a, (6, 4))
    assert r[1 + 2 * 2, 3] == a[1, 2, 3]
    assert tci.address_from_coor((2, 3, 4), (1, 2, 3)) == 1 + 2 * 2 + 6 * 3


def test_assign_from_range_row_major_gate():
    u = np.array([[1, 2], [3, 4]], dtype=complex)
    g = tci.assign_from_range([2, 2], u.reshape(-1), lambda c: tci.address_from_coor((2, 2), c))
    # G(in, out) = U[out][in]
    assert g[0, 1] == u[1, 0] and g[1, 0] == u[0, 1]


def test_svd_reconstructs_and_orders():
    rng = np.random.default_rng(2)
    a = _rand(rng, (3, 2, 5))
    u, s, vd = tci.svd(a, 2)
    assert u.shape == (3, 2, 5) and vd.shape == (5, 5) and s.size == 5
    assert np.all(np.diff(s) <= 0)
    rec = tci.contract(tci.contract(u, [0, 1, -1], tci.diag(tci.to_cplx(s)), [-1, 2], [0, 1, 2]),
                       [0, 1, -1], vd, [-1, 2], [0, 1, 2])
    np.testing.assert_allclose(rec, a, atol=1e-12)


def test_trunc_svd_rule():
    rng = np.random.default_rng(3)
    q1, _ = np.linalg.qr(_rand(rng, (6, 6)))
    q2, _ = np.linalg.qr(_rand(rng, (6, 6)))
    s_true = np.array([1.0, 0.5, 0.25, 1e-3, 1e-9, 0.0])
    m = q1 @ np.diag(s_true) @ q2
    u, s, vd, err = tci.trunc_svd(m, 1, 3, 1e-8)
    np.testing.assert_allclose(s, s_true[:3], rtol=1e-12)
    np.testing.assert_allclose(err, (1e-6 + 1e-18) / np.sum(s_true ** 2), rtol=1e-8)
    u, s, vd, err = tci.trunc_svd(m, 1, 10, 1e-8)
    assert s.size == 4  # 1e-9 and 0 are <= s_min


def test_normalize_and_norm():
    a = np.array([[3, 0], [0, 4j]])
    b, n = tci.normalize(a)
    assert n == pytest.approx(5.0)
    assert tci.norm(b) == pytest.approx(1.0)


def test_trace_and_diag():
    rng = np.random.default_rng(4)
    a = _rand(rng, (2, 2, 3, 3))
    np.testing.assert_allclose(tci.trace(a, [(0, 1)]), a[0, 0] + a[1, 1])
    d = tci.diag(np.array([1.0, 2.0, 3.0]))
    assert d.reshape(-1, order="F")[1 + 3 * 1] == 2.0


def test_trace_pairs_transpose_and_small_ops_brute_force():
    """The index bookkeeping of the small TCI operations against nested loops:
    trace over two bond pairs given in either order (axes shift as pairs are
    removed), transpose (axis k of the result is axis new_order[k]),
    linear_combine, cplx_conj, to_cplx, fill and get_elem of an order-0
    tensor."""
    rng = np.random.default_rng(11)
    a = _rand(rng, (2, 3, 2, 3, 4))
    want = np.zeros(4, dtype=complex)
    for i in range(2):
        for j in range(3):
            for k in range(4):
                want[k] += a[i, j, i, j, k]
    np.testing.assert_allclose(tci.trace(a, [(0, 2), (1, 3)]), want, atol=1e-14)
    np.testing.assert_allclose(tci.trace(a, [(1, 3), (0, 2)]), want, atol=1e-14)
    b = _rand(rng, (2, 3, 4))
    order = [2, 0, 1]
    t = tci.transpose(b, order)
    assert t.shape == (4, 2, 3)
    for x in itertools.product(range(4), range(2), range(3)):
        y = [0, 0, 0]
        for k in range(3):
            y[order[k]] = x[k]
        assert t[x] == b[tuple(y)]
    c = _rand(rng, (2, 3, 4))
    np.testing.assert_allclose(tci.linear_combine([b, c], [2.0, -1j]), 2.0 * b - 1j * c, atol=1e-14)
    np.testing.assert_array_equal(tci.cplx_conj(np.array([1 + 2j, -3j])), np.array([1 - 2j, 3j]))
    z = tci.to_cplx(np.array([1.5, -2.0]))
    assert z.dtype == np.complex128 and list(z) == [1.5 + 0j, -2.0 + 0j]
    f = tci.fill((2, 3), 0.5 - 1j)
    assert f.shape == (2, 3) and np.all(f == 0.5 - 1j)
    assert tci.get_elem(np.array(2.5 - 1j), []) == 2.5 - 1j
