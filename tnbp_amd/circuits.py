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


def _bond_between(edges, inc, s, t):
    """Position of the bond joining sites s and t among s's bonds."""
    for pos, e in enumerate(inc[s]):
        if set(edges[e]) == {s, t}:
            return pos
    raise ValueError(f"sites {s} and {t} are not neighbours")


def _append_path_factor(O, k_prev, k_next, W):
    """O(..bonds.., in, x) with W(l, r, x, out) on a path through the site: l
    fuses into bond k_prev and r into bond k_next (either may be None, with
    that dimension 1), the old bond index the faster one (qasmtotpo.h:333-436)."""
    z = O.ndim - 2
    res = np.tensordot(O, W, axes=([z + 1], [2]))      # (bonds..., in, l, r, out)
    res = np.moveaxis(res, z + 3, z + 2)               # (bonds..., in, l, out, r)
    if k_prev is None:
        res = res.reshape(res.shape[: z + 1] + res.shape[z + 2:], order="F")  # drop l = 1
        r_axis = z + 2
    else:
        r_axis = z + 3
    if k_next is None:
        res = np.take(res, 0, axis=r_axis)             # r = 1
    # now (bonds..., in, [l], out, [r]); fuse l into k_prev and r into k_next
    axes = list(range(z))
    shape = list(O.shape[:z])
    cur = z + 1
    extra = {}
    if k_prev is not None:
        extra[k_prev] = cur
        cur += 1
    out_axis = cur
    cur += 1
    if k_next is not None:
        extra[k_next] = cur
    order = []
    for k in range(z):
        order.append(k)
        if k in extra:
            order.append(extra[k])
            shape[k] *= res.shape[extra[k]]
    order += [z, out_axis]
    res = np.transpose(res, order)
    return res.reshape(shape + [O.shape[z], W.shape[3]], order="F")


def _shortest_path(edges, a, b):
    """BFS shortest path a..b over the lattice (graph.h:57-108)."""
    from collections import deque
    adj = {}
    for u, v in edges:
        adj.setdefault(u, []).append(v)
        adj.setdefault(v, []).append(u)
    prev = {a: None}
    q = deque([a])
    while q:
        u = q.popleft()
        if u == b:
            break
        for v in adj.get(u, []):
            if v not in prev:
                prev[v] = u
                q.append(v)
    if b not in prev:
        raise ValueError(f"qubits {a} and {b} are not connected by the lattice")
    path, cur = [], b
    while cur is not None:
        path.append(cur)
        cur = prev[cur]
    return path[::-1]


def split_three_qubit(u, s_min=1.0e-12):
    """Split an 8x8 gate on (x, y, w) into W_x(1, k1, in, out),
    W_y(k1, k2, in, out), W_w(k2, 1, in, out): SVD of the (x_in x_out | rest)
    matricisation, then of (k1 y_in y_out | w_in w_out), sqrt(S) on both
    sides (qasmtotpo.h:497-590); singular values <= s_min * the largest are
    dropped (exact zeros of controlled gates)."""
    g = op_tensor(u)                                               # (x, y, w in; x, y, w out)
    g = np.transpose(g, (0, 3, 1, 4, 2, 5))                        # (xi, xo, yi, yo, wi, wo)
    m = g.reshape(4, 16, order="F")
    uu, s, vh = np.linalg.svd(m, full_matrices=False)
    k1 = max(1, int(np.count_nonzero(s > s_min * s[0])))
    r = np.sqrt(s[:k1])
    wx = (uu[:, :k1] * r).reshape(2, 2, k1, order="F")            # (xi, xo, k1)
    rest = (r[:, None] * vh[:k1, :]).reshape(k1 * 4, 4, order="F")  # (k1 yi yo | wi wo)
    uu, s, vh = np.linalg.svd(rest, full_matrices=False)
    k2 = max(1, int(np.count_nonzero(s > s_min * s[0])))
    r = np.sqrt(s[:k2])
    wy = (uu[:, :k2] * r).reshape(k1, 2, 2, k2, order="F")        # (k1, yi, yo, k2)
    ww = (r[:, None] * vh[:k2, :]).reshape(k2, 2, 2, order="F")   # (k2, wi, wo)
    return (np.transpose(wx, (2, 0, 1))[None],                     # (1, k1, xi, xo)
            np.transpose(wy, (0, 3, 1, 2)),                        # (k1, k2, yi, yo)
            ww[:, None])                                           # (k2, 1, wi, wo)


def _path_gate(O, edges, inc, path, factors):
    """A gate as an operator string along `path`: factors {site: W(l, r, in,
    out)}, identity (delta(l, r) x I) on the other path sites."""
    bond = 1
    for i, p in enumerate(path):
        k_prev = _bond_between(edges, inc, p, path[i - 1]) if i > 0 else None
        k_next = _bond_between(edges, inc, p, path[i + 1]) if i + 1 < len(path) else None
        if p in factors:
            W = factors[p]
            bond = W.shape[1]
        else:
            W = np.einsum("lr,io->lrio", np.eye(bond), np.eye(2)).astype(complex)
        O[p] = _append_path_factor(O[p], k_prev, k_next, W)


def _three_qubit_route(edges, q):
    """Of the routes a-b-c, b-c-a, c-a-b through the lattice, the shortest
    (first on ties): (rotation, path) with the middle qubit of the rotation
    in between (qasmtotpo.h:440-490, whose index arithmetic there runs past
    its arrays; this keeps its intent: path(x, y) + path(y, w) minus the
    repeated y)."""
    a, b, c = q
    rots = [(a, b, c), (b, c, a), (c, a, b)]
    best = None
    for r, (x, y, w) in enumerate(rots):
        p = _shortest_path(edges, x, y) + _shortest_path(edges, y, w)[1:]
        if best is None or len(p) < len(best[1]):
            best = (r, p)
    r, path = best
    if len(set(path)) != len(path):
        raise ValueError(f"three-qubit gate on {q}: its route through the lattice repeats a site")
    return r, path


def layer_tpo(edges, gates, two_site_keep_zero=False):
    """Per-site operators O_s(t_1..t_z, d_in, d_out) for one circuit slice.

    ``gates``: list of (U, qubits) applied in order, each after the ones
    before it; one-qubit gates act on a site, two-qubit gates on the two ends
    of an edge, or along the shortest lattice path with identities in between
    (qasmtotpo.h:333-436); three-qubit gates (8 x 8) along the shortest route
    through their middle qubit.
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
        if len(q) == 2:
            a, b = q
            m = [e for e in inc[a] if set(edges[e]) == {a, b}]
            A, B = split_two_qubit(u, keep_zero=two_site_keep_zero)
            if m:
                O[a] = _append_bond_factor(O[a], inc[a].index(m[0]), A)
                O[b] = _append_bond_factor(O[b], inc[b].index(m[0]), B)
            else:
                _path_gate(O, edges, inc, _shortest_path(edges, a, b), {a: A[None], b: B[:, None]})
            continue
        if len(q) == 3:
            if u.shape != (8, 8):
                raise ValueError("three-qubit gates take an 8 x 8 matrix")
            r, path = _three_qubit_route(edges, q)
            x, y, w = [(q[0], q[1], q[2]), (q[1], q[2], q[0]), (q[2], q[0], q[1])][r]
            # the matrix with its qubits in the rotated order (x low bit)
            perm = [(0, 1, 2), (1, 2, 0), (2, 0, 1)][r]
            t = op_tensor(u)                              # (in a b c, out a b c)
            t = np.transpose(t, list(perm) + [3 + p for p in perm])
            n = t.reshape(8, 8, order="F").T              # back to U[out][in], x the low bit
            wx, wy, ww = split_three_qubit(n)
            _path_gate(O, edges, inc, path, {x: wx, y: wy, w: ww})
            continue
        raise ValueError(f"{len(q)}-qubit gates are not supported")
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
