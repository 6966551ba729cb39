"""Circuit front end: lattices and tensor-product operators (host work).

The reference builds these on the host from 2x2 / 4x4 gate matrices
(apps/kicked_ising_lattice/circuit.h:14-229, include/tnbp/parser/qasmtotpo.h);
they are a few hundred complex numbers, so they are built here with numpy and
handed to the device by ``AbsorbTPO``.
"""
from __future__ import annotations

import math

import numpy as np


# ------------------------------------------------------------ lattices
def square_lattice(lx, ly):
    """Edge list of ``bond_square_lattice`` (include/tnbp/lattice/square.h:14-34)."""
    out = []
    for iy in range(ly):
        for ix in range(lx):
            i = ix + lx * iy
            if ix + 1 < lx:
                out.append((i, i + 1))
            if iy + 1 < ly:
                out.append((i, i + lx))
    return out


def chain(lx):
    """Open chain of ``bond_oned_lattice`` (include/tnbp/lattice/oned.h:14-36)."""
    return [(i, i + 1) for i in range(lx - 1)]


def incident(edges, site):
    """``GetSurroundingBondIndex`` (graph.h:40-55)."""
    return [m for m, (a, b) in enumerate(edges) if a == site or b == site]


# --------------------------------------------------------------- gates
def u3(th, ph, la):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -np.exp(1j * la) * s], [np.exp(1j * ph) * s, np.exp(1j * (ph + la)) * c]])


def rx(th):
    return np.array([[math.cos(th / 2), -1j * math.sin(th / 2)],
                     [-1j * math.sin(th / 2), math.cos(th / 2)]])


def rz(th):
    return np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)])


def rzz(th):
    """exp(-i th/2 Z Z), qasmtotpo.h:106-114; first qubit = least-significant bit."""
    return np.diag([np.exp(-0.5j * th), np.exp(0.5j * th), np.exp(0.5j * th), np.exp(-0.5j * th)])


PAULI = {
    "I": np.eye(2, dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


def op_tensor(u):
    """Row-major matrix U -> operator tensor O(in..., out...) = U[out][in]."""
    n = int(round(math.log2(u.shape[0])))
    # U[out][in] with little-endian qubit order: reshape reverses bit order
    t = np.asarray(u, dtype=complex).reshape([2] * (2 * n))  # [out_{n-1}..out_0, in_{n-1}..in_0]
    axes = list(range(2 * n - 1, n - 1, -1)) + list(range(n - 1, -1, -1))
    return np.transpose(t, axes)  # -> (in_0..in_{n-1}, out_0..out_{n-1})


def split_two_qubit(u, s_min=1.0e-12, keep_zero=False):
    """Split a 4x4 gate into A(k, a_in, a_out) and B(k, b_in, b_out) with
    sum_k A B = G(a_in, b_in, a_out, b_out), sqrt(S) on both sides
    (circuit.h:55-80, qasmtotpo.h:292-331).  ``keep_zero`` keeps all four
    singular values like circuit.h:62 does (t = 4)."""
    g = op_tensor(u)                                   # (a_in, b_in, a_out, b_out)
    m = np.transpose(g, (0, 2, 1, 3)).reshape(4, 4, order="F")  # rows (a_in, a_out)
    uu, s, vh = np.linalg.svd(m)
    k = 4 if keep_zero else max(1, int(np.count_nonzero(s > s_min)))
    r = np.sqrt(s[:k])
    a = (uu[:, :k] * r).reshape(2, 2, k, order="F")    # (a_in, a_out, k)
    b = (r[:, None] * vh[:k, :]).reshape(k, 2, 2, order="F")  # (k, b_in, b_out)
    return np.transpose(a, (2, 0, 1)), b


def _append_bond_factor(O, k_pos, F):
    """O(..bonds.., in, x) -> O(.., (old, new) at k_pos, .., in, out) with
    F(new, x, out); the old bond index is the faster one (circuit.h:140-159)."""
    z = O.ndim - 2
    res = np.tensordot(O, F, axes=([z + 1], [1]))     # (bonds..., in, new, out)
    # move `new` next to bond k_pos, then fuse (old, new) column-major
    res = np.moveaxis(res, z + 1, k_pos + 1)          # (.., old, new, .., in, out)
    sh = list(O.shape[:z])
    sh[k_pos] = O.shape[k_pos] * F.shape[0]
    return res.reshape(sh + [O.shape[z], F.shape[2]], order="F")


def layer_tpo(edges, gates, two_site_keep_zero=False):
    """Per-site operators O_s(t_1..t_z, d_in, d_out) for one circuit slice.

    ``gates``: list of (U, qubits) applied in order; one-qubit gates act on a
    site, two-qubit gates on the two ends of an edge (nearest neighbours).
    """
    sites = sorted({s for e in edges for s in e})
    inc = {s: incident(edges, s) for s in sites}
    O = {s: np.eye(2, dtype=complex).reshape([1] * len(inc[s]) + [2, 2]) for s in sites}
    for u, q in gates:
        u = np.asarray(u, dtype=complex)
        if len(q) == 1:
            s = q[0]
            O[s] = np.tensordot(O[s], op_tensor(u), axes=([O[s].ndim - 1], [0]))
            continue
        a, b = q
        m = [e for e in inc[a] if set(edges[e]) == {a, b}]
        if not m:
            raise ValueError("two-qubit gates must act on nearest neighbours")
        A, B = split_two_qubit(u, keep_zero=two_site_keep_zero)
        O[a] = _append_bond_factor(O[a], inc[a].index(m[0]), A)
        O[b] = _append_bond_factor(O[b], inc[b].index(m[0]), B)
    return [O[s] for s in sites]


def kicked_ising_tpo(edges, Jz, hz, hx, dt, two_site_keep_zero=False):
    """``CircuitTPO`` (circuit.h:14-229): exp(-i dt Jz ZZ) on every edge in edge
    order, then exp(-i dt hz Z) and exp(-i dt hx X) on every site (dt in radians)."""
    a = dt * Jz
    zz = np.diag([np.exp(-1j * a), np.exp(1j * a), np.exp(1j * a), np.exp(-1j * a)])
    sites = sorted({s for e in edges for s in e})
    gates = [(zz, e) for e in edges]
    gates += [(np.diag([np.exp(-1j * dt * hz), np.exp(1j * dt * hz)]), (s,)) for s in sites]
    g = dt * hx
    ux = np.array([[math.cos(g), -1j * math.sin(g)], [-1j * math.sin(g), math.cos(g)]])
    gates += [(ux, (s,)) for s in sites]
    return layer_tpo(edges, gates, two_site_keep_zero)


def random_layer(rng, edges, n_sites):
    """One layer of the ragged-batch family: RZZ(theta ~ U[0, pi]) on a random
    greedy matching, then U3 with angles ~ U[0, 2 pi] on every site."""
    used, gates = set(), []
    for m in rng.permutation(len(edges)):
        a, b = edges[m]
        if a in used or b in used:
            continue
        used.update((a, b))
        gates.append((rzz(rng.uniform(0, math.pi)), (a, b)))
    for s in range(n_sites):
        gates.append((u3(*rng.uniform(0, 2 * math.pi, 3)), (s,)))
    return gates


def opt_tpo_by_svd(edges, O, eps=1.0e-12):
    """Compress every TPO bond by SVD (``OptTPObySVD``, optsvdtpo.h:87-214):
    per edge in order, the two site operators contracted over the shared bond
    are refactored as U sqrt(S) | sqrt(S) V keeping singular values > eps (at
    most the old dimension, at least one).  E.g. cx . u1 . cx on an edge
    (t = 4 when stacked gate by gate) compresses to t = 2."""
    sites = sorted({s for e in edges for s in e})
    pos = {s: i for i, s in enumerate(sites)}
    inc = {s: incident(edges, s) for s in sites}
    O = [np.asarray(o, dtype=complex) for o in O]
    for m, (a, b) in enumerate(edges):
        ia, ib = pos[a], pos[b]
        ka, kb = inc[a].index(m), inc[b].index(m)
        A = np.moveaxis(O[ia], ka, -1)                     # (others_a..., t)
        B = np.moveaxis(O[ib], kb, 0)                      # (t, others_b...)
        t = A.shape[-1]
        sa, sb = A.shape[:-1], B.shape[1:]
        C = A.reshape(-1, t, order="F") @ B.reshape(t, -1, order="F")
        u, s, vh = np.linalg.svd(C, full_matrices=False)
        k = max(1, min(t, int(np.count_nonzero(s > eps))))
        r = np.sqrt(s[:k])
        O[ia] = np.moveaxis((u[:, :k] * r).reshape(sa + (k,), order="F"), -1, ka)
        O[ib] = np.moveaxis((r[:, None] * vh[:k]).reshape((k,) + sb, order="F"), 0, kb)
    return O
