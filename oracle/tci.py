"""CPU restatement (numpy) of the Tensor Computing Interface calls on the BP-PEPS path.

TEST INFRASTRUCTURE ONLY.  This module is the parity oracle: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product path (``tnbp_amd``) never
imports it and fails loudly when its HIP library is missing.

A tensor is a numpy ``complex128`` (or ``float64`` for ``real_ten_t``) array
whose axes are the TCI bonds in order.  TCI stores tensors column-major (first
index fastest, min-tci constructor.h:109-110), so every reshape here is
``order='F'``; numpy's own memory order is irrelevant to the results.

The arithmetic of min-tci lives in gqten (GraceQ/tensor-ng-dev, branch
``to-tci-release``, commit unpinned: ``external/tensor-ng-dev`` is an empty
submodule).  Where gqten's semantics are not fixed by min-tci's own text the
choice made here is stated on the function and marked *unpinned*; tests/
pins each choice that the mathematics can pin.
"""
from __future__ import annotations

import numpy as np

CPLX = np.complex128


# ---------------------------------------------------------------- contract
def contract_labels(la, lb, lc):
    """Label rules of ``tci::contract`` (min-tci linalg.h:205-305).

    Labels shared by ``a`` and ``b`` are summed; the survivors appear in the
    output in the order of ``lc``.  Raises ``ValueError`` for the same
    conditions min-tci throws ``std::invalid_argument``/``logic_error``
    (duplicates, survivor set != ``lc``).
    Returns (axes_a, axes_b, perm) for ``np.tensordot`` + ``transpose``.
    """
    la, lb, lc = list(la), list(lb), list(lc)
    for name, lab in (("a", la), ("b", lb), ("c", lc)):
        if len(set(lab)) != len(lab):
            raise ValueError(f"bond_labs_{name} has duplicate label.")
    shared = sorted(set(la) & set(lb))
    survivors = (set(la) | set(lb)) - set(shared)
    if set(lc) != survivors:
        raise ValueError("bond_labs_c does not match the set of survivor labels.")
    axes_a = [la.index(x) for x in shared]
    axes_b = [lb.index(x) for x in shared]
    free = [x for x in la if x not in shared] + [x for x in lb if x not in shared]
    perm = [free.index(x) for x in lc]
    return axes_a, axes_b, perm


def contract(a, la, b, lb, lc):
    """``tci::contract(ctx, a, la, b, lb, c, lc)`` (min-tci linalg.h:311-323)."""
    axes_a, axes_b, perm = contract_labels(la, lb, lc)
    for x, y in zip(axes_a, axes_b):
        if a.shape[x] != b.shape[y]:
            raise ValueError("contracted bond dimensions differ")
    c = np.tensordot(a, b, axes=(axes_a, axes_b))
    if not perm:  # order-0 result (measure.h:83-84)
        return np.asarray(c)
    return np.ascontiguousarray(np.transpose(c, perm))


# ---------------------------------------------------------- constructors
def assign_from_range(shape, data, coors2idx):
    """``tci::assign_from_range`` (min-tci constructor.h:114-163).

    Element at coordinate ``coor`` is ``data[coors2idx(coor)]``; coordinates
    enumerate storage order column-major.
    """
    out = np.zeros(tuple(shape), dtype=CPLX)
    for coor in np.ndindex(*shape):
        out[coor] = data[coors2idx(coor)]
    return out


def address_from_coor(shape, coor):
    """Column-major address (tnbp framework/mpiutility.h:67-77, circuit.h:1-11)."""
    addr = 0
    for k in range(len(shape) - 1, -1, -1):
        addr = addr * shape[k] + coor[k]
    return addr


def fill(shape, value):
    """``tci::fill`` (min-tci constructor.h:82-104)."""
    return np.full(tuple(shape), value, dtype=CPLX)


# ----------------------------------------------------------- manipulators
def reshape(a, shape):
    """``tci::reshape``: metadata-only change in column-major order (manipulator.h:37-55)."""
    return np.reshape(a, tuple(shape), order="F")


def transpose(a, new_order):
    """``tci::transpose``: axis i of the result is axis ``new_order[i]`` of ``a``."""
    return np.ascontiguousarray(np.transpose(a, new_order))


def cplx_conj(a):
    """``tci::cplx_conj`` (manipulator.h:85-101): elementwise complex conjugate."""
    return np.conj(a)


def diag(v):
    """``tci::diag`` (min-tci linalg.h:32-62): vector of n -> n x n with ``praw[i+n*i]``."""
    v = np.asarray(v).reshape(-1, order="F")
    out = np.zeros((v.size, v.size), dtype=v.dtype)
    out[np.arange(v.size), np.arange(v.size)] = v
    return out


def to_cplx(r):
    """``tci::to_cplx`` (miscellaneous.h:107-142)."""
    return np.asarray(r, dtype=CPLX)


# ------------------------------------------------------------ reductions
def norm(a):
    """``tci::norm`` -> gqten ``CalcNorm``; Frobenius norm (unpinned: assumed 2-norm)."""
    return float(np.sqrt(np.sum(np.abs(a) ** 2)))


def normalize(a):
    """``tci::normalize`` -> gqten ``Normalize`` (linalg.h:78-98).

    Returns ``(a / ||a||_F, ||a||_F)``.  Unpinned: the convention (Frobenius
    to unit norm) is assumed; it fixes the BP residual's value.
    """
    n = norm(a)
    if n == 0.0:
        return a.copy(), 0.0
    return a / n, n


def linear_combine(ins, coefs):
    """``tci::linear_combine`` (linalg.h:468-493): sum_i coefs[i] * ins[i]."""
    out = np.zeros_like(ins[0])
    for t, c in zip(ins, coefs):
        out = out + c * t
    return out


def trace(a, pairs):
    """``tci::trace`` (linalg.h:130-149): trace over each (i, j) bond pair."""
    axes = list(range(a.ndim))
    out = a
    removed = []
    for i, j in pairs:
        ii = i - sum(1 for r in removed if r < i)
        jj = j - sum(1 for r in removed if r < j)
        out = np.trace(out, axis1=ii, axis2=jj)
        removed += [i, j]
    del axes
    return np.ascontiguousarray(out)


def get_elem(a, coors):
    """``tci::get_elem`` (getter.h:50-80); an order-0 tensor returns its scalar."""
    if a.ndim == 0:
        return complex(a[()])
    return complex(a[tuple(coors)])


# ------------------------------------------------------------------ SVD
def _matricize(a, nrows):
    shape = a.shape
    m = int(np.prod(shape[:nrows])) if nrows > 0 else 1
    n = int(np.prod(shape[nrows:])) if nrows < len(shape) else 1
    return reshape(a, (m, n)), shape[:nrows], shape[nrows:]


def svd(a, nrows):
    """``tci::svd`` -> gqten ``SVD`` (linalg.h:500-512).

    Thin SVD of the column-major matricisation; k = min(m, n) singular values
    in descending order (unpinned: k for rank-deficient input assumed min(m, n)).
    Returns (u, s, vdag) with u (rows..., k), vdag (k, cols...).
    """
    mat, rshape, cshape = _matricize(a, nrows)
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    k = s.size
    return reshape(u, tuple(rshape) + (k,)), s, reshape(vh, (k,) + tuple(cshape))


def trunc_svd(a, nrows, chi_max, s_min):
    """``tci::trunc_svd(..., chi_max, s_min)`` -> gqten ``TruncSVD`` (linalg.h:552-569).

    Unpinned rule (gqten absent): keep the singular values strictly greater
    than ``s_min``, at most ``chi_max`` of them, in descending order, and at
    least one; ``trunc_err`` = (sum of discarded s^2) / (sum of all s^2).
    """
    mat, rshape, cshape = _matricize(a, nrows)
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    keep = int(np.count_nonzero(s > s_min))
    keep = max(1, min(keep, int(chi_max)))
    tot = float(np.sum(s ** 2))
    err = float(np.sum(s[keep:] ** 2) / tot) if tot > 0 else 0.0
    return (reshape(u[:, :keep], tuple(rshape) + (keep,)), s[:keep].copy(),
            reshape(vh[:keep, :], (keep,) + tuple(cshape)), err)
