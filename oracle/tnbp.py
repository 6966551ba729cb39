"""CPU restatement (numpy) of tnbp's belief-propagation PEPS/TPS path.

TEST INFRASTRUCTURE ONLY (the parity oracle): importable from ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg, never from
the product path.

Every function follows the reference's operation sequence through the TCI
restatement in ``oracle/tci.py`` - the same contractions with the same labels,
normalisations at the same points, messages stored at the same addresses - so
that rounding is the only difference to the reference's min-tci/CPU run.  One
MPI rank is assumed throughout: every ``type``/``mpi_type`` branch of the
reference takes its "both sites local" arm, Send/Recv never fire and
Allreduce is the identity.

Parity status of each function (what pins it) is listed in tests/test_oracle_*.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np

from . import tci

CPLX = np.complex128


# =============================================================== graph
def site_index_from_bond(edges):
    """``GetSiteIndexFromBond`` (include/tnbp/framework/graph.h:23-35): sorted unique sites."""
    return sorted({s for e in edges for s in e})


def surrounding_bond_index(site, edges):
    """``GetSurroundingBondIndex`` (graph.h:40-55): incident edge ids, ascending."""
    res = []
    for m, (i, j) in enumerate(edges):
        if i == site:
            res.append(m)
        if j == site:
            res.append(m)
    return res


def greedy_edge_layering(edges):
    """``GreedyEdgeLayering`` (graph.h:138-170): first-fit edge colouring."""
    used = [False] * len(edges)
    layers = []
    while True:
        vs, layer = set(), []
        for i, (u, v) in enumerate(edges):
            if used[i] or u in vs or v in vs:
                continue
            layer.append((u, v))
            vs.update((u, v))
            used[i] = True
        if not layer:
            break
        layers.append(layer)
    return layers


def find_shortest_path(edges, src, dst):
    """``FindShortestPath`` (graph.h:57-108): BFS path src..dst (inclusive)."""
    sites = site_index_from_bond(edges)
    if src not in sites or dst not in sites:
        return []
    adj = {s: [] for s in sites}
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    prev = {src: None}
    q = deque([src])
    while q:
        u = q.popleft()
        if u == dst:
            break
        for v in adj[u]:
            if v not in prev:
                prev[v] = u
                q.append(v)
    if dst not in prev:
        return []
    path, cur = [], dst
    while cur is not None:
        path.append(cur)
        cur = prev[cur]
    return path[::-1]


def bond_square_lattice(lx, ly):
    """``bond_square_lattice`` (include/tnbp/lattice/square.h:14-34): (i,i+1) then (i,i+Lx)."""
    res = []
    for iy in range(ly):
        for ix in range(lx):
            i = ix + lx * iy
            if ix + 1 < lx:
                res.append((i, ix + 1 + lx * iy))
            if iy + 1 < ly:
                res.append((i, ix + lx * (iy + 1)))
    return res


def bond_oned_lattice(lx, px=0):
    """``bond_oned_lattice`` (include/tnbp/lattice/oned.h:14-36)."""
    res = []
    for i in range(lx):
        j = (i + 1) % lx if px != 0 else i + 1
        if j < lx:
            res.append((i, j))
    return res


def _locate(edges, bonds, a, b):
    """Position k of edge {a,b} in ``bonds`` and its global id (bp.h:107-121)."""
    for k, m in enumerate(bonds):
        if (edges[m][0] == a and edges[m][1] == b) or (edges[m][0] == b and edges[m][1] == a):
            return k, m
    raise ValueError("edge not found")


# ================================================================ init
class State:
    """The per-rank data the reference threads through every call:
    ``I`` (edges), ``V`` (site tensors, in ``SiteIdx`` order), ``SiteIdx``,
    ``E`` (2*n_e messages; ``E[m]`` first->second, ``E[m+n_e]`` second->first)
    and ``EdgeIdx`` (bp.h:65-75)."""

    def __init__(self, edges, V, site_idx, E, edge_idx):
        self.I = list(edges)
        self.V = V
        self.SiteIdx = site_idx
        self.E = E
        self.EdgeIdx = edge_idx

    def copy(self):
        return State(self.I, [v.copy() for v in self.V], list(self.SiteIdx),
                     [e.copy() for e in self.E], list(self.EdgeIdx))


def init_tensor_product_state(edges, pdim):
    """``InitTensorProductState`` (include/tnbp/ptns/function/init.h:17-95), one rank.

    V_s = |0> with all virtual bonds of dimension 1 (:57-60); messages are 1x1
    ones (:72-78); EdgeIdx = sorted incident edges of the local sites (:67-69).
    """
    sites = site_index_from_bond(edges)
    V, edge_idx = [], []
    for addr, s in enumerate(sites):
        bonds = surrounding_bond_index(s, edges)
        edge_idx += bonds
        shape = [1] * len(bonds) + [pdim[addr]]
        data = np.zeros(pdim[addr], dtype=CPLX)
        data[0] = 1.0
        V.append(tci.assign_from_range(shape, data,
                                       lambda c, sh=shape: tci.address_from_coor(sh, c)))
    edge_idx = sorted(set(edge_idx))
    E = [tci.fill((1, 1), 1.0) for _ in range(2 * len(edge_idx))]
    return State(edges, V, sites, E, edge_idx)


# ========================================================== absorb TPO
def absorb_tpo(st, O):
    """``AbsorbTPO`` (include/tnbp/ptns/function/apply.h:14-177).

    V'[(v1 o1),...,(vz oz), p'] = sum_p O[o1..oz, p, p'] V[v1..vz, p] with the
    fused bond index v fastest (:52-66); every message becomes E (x) I_t with
    the same fusion (:110-128).
    """
    I = st.I
    gsites = site_index_from_bond(I)
    for addr, site in enumerate(st.SiteIdx):
        o = O[gsites.index(site)]
        v = st.V[addr]
        nvb = v.ndim - 1
        label_v = [2 * k for k in range(nvb)] + [-1]
        label_o = [2 * k + 1 for k in range(nvb)] + [-1, 2 * nvb + 1]
        label_r = list(range(2 * nvb)) + [2 * nvb + 1]
        new_shape = [v.shape[k] * o.shape[k] for k in range(nvb)] + [o.shape[nvb + 1]]
        r = tci.contract(o, label_o, v, label_v, label_r)
        st.V[addr] = tci.reshape(r, new_shape)
    ne = len(st.EdgeIdx)
    for addr in range(ne):
        a, b = I[st.EdgeIdx[addr]]
        site = a
        bonds = surrounding_bond_index(site, I)
        kt, _ = _locate(I, bonds, a, b)
        t = O[gsites.index(site)].shape[kt]
        eye = np.eye(t, dtype=CPLX)
        for idx in (addr, addr + ne):
            e = st.E[idx]
            tt = tci.contract(e, [0, 2], eye, [1, 3], [0, 1, 2, 3])
            st.E[idx] = tci.reshape(tt, (t * e.shape[0], t * e.shape[0]))


# ================================================================== BP
def _absorb_messages(T, I, site, bonds, skip_k, E, edge_idx, normalize_every=True):
    """The mode-k absorption loop of bp.h:137-154 (and bp.h:295-315, measure.h:54-70)."""
    ne = len(edge_idx)
    order = T.ndim
    for k, m in enumerate(bonds):
        if k != skip_k:
            mp = edge_idx.index(m)
            msg = E[mp + ne] if I[m][0] == site else E[mp]
            la = list(range(order))
            la[k] = -1
            T = tci.contract(T, la, msg, [-1, k], list(range(order)))
        if normalize_every:
            T, _ = tci.normalize(T)
    return T


def _message_from(T, V, k):
    """``contract(T, IdxT, conj V, IdxA, M, {0,1})`` of bp.h:155-165: all modes but k."""
    order = V.ndim
    idx_t = list(range(-order, 0))
    idx_a = list(range(-order, 0))
    idx_t[k] = 0
    idx_a[k] = 1
    return tci.contract(T, idx_t, tci.cplx_conj(V), idx_a, [0, 1])


def belief_propagation(st, J, F):
    """``BeliefPropagation`` (include/tnbp/ptns/function/bp.h:65-261) for one layer ``J``.

    For each edge (A,B) of the layer and each end, the message toward the other
    end is computed from ``st.E`` and written into ``F`` (a list of 2*n_e).
    """
    I, E, edge_idx = st.I, st.E, st.EdgeIdx
    ne = len(edge_idx)
    for (sa, sb) in J:
        bonds_a = surrounding_bond_index(sa, I)
        bonds_b = surrounding_bond_index(sb, I)
        ka, tgt = _locate(I, bonds_a, sa, sb)
        kb = bonds_b.index(tgt)
        mpt = edge_idx.index(tgt)
        for site, bonds, k in ((sa, bonds_a, ka), (sb, bonds_b, kb)):
            v = st.V[st.SiteIdx.index(site)]
            T = _absorb_messages(v.copy(), I, site, bonds, k, E, edge_idx)
            M, _ = tci.normalize(_message_from(T, v, k))
            if I[tgt][0] == site:
                F[mpt] = M
            else:
                F[mpt + ne] = M


def belief_propagation_condition(st):
    """``BeliefPropagationCondition`` (bp.h:266-355): mean Frobenius distance
    between every recomputed normalised message and the stored one."""
    I, E, edge_idx = st.I, st.E, st.EdgeIdx
    ne = len(edge_idx)
    vol = 1.0 / (2.0 * len(I))
    result = 0.0
    for addr, site in enumerate(st.SiteIdx):
        bonds = surrounding_bond_index(site, I)
        v = st.V[addr]
        for k in range(len(bonds)):
            T = v.copy()
            for l, m in enumerate(bonds):
                if l != k:
                    ea = edge_idx.index(m)
                    if I[m][0] == site:
                        ea += ne
                    la = list(range(T.ndim))
                    la[l] = -1
                    T = tci.contract(T, la, E[ea], [-1, l], list(range(T.ndim)))
                    T, _ = tci.normalize(T)
            T, _ = tci.normalize(_message_from(T, v, k))
            ea = edge_idx.index(bonds[k])
            if I[bonds[k]][1] == site:
                ea += ne
            Fm, _ = tci.normalize(E[ea])
            D = tci.linear_combine([T, Fm], [1.0, -1.0])
            result += tci.norm(D) * vol
    return result


def bp_loop(st, layers, max_iter, tol, log=None):
    """Driver BP loop (apps/kicked_ising_lattice/main.cc:121-159).

    Returns (iterations run, list of residuals).  E keeps the last sweep's
    messages, not the condition's recomputed ones (main.cc:138-141).
    """
    F = [None] * (2 * len(st.EdgeIdx))
    res = []
    for k in range(max_iter):
        for J in layers:
            belief_propagation(st, J, F)
        st.E = [f.copy() for f in F]
        r = belief_propagation_condition(st)
        res.append(r)
        if log is not None:
            log(k, r)
        if r < tol:
            return k + 1, res
    return max_iter, res


# ========================================================== truncation
def square_root_and_inverse(M, sv_min):
    """``SquareRootAndInverse`` (include/tnbp/framework/root.h:65-148).

    With M = U diag(e) Vd: R = U diag(sqrt e [e > sv_min]) Vd and
    S = Vd^H diag(1/sqrt e [sqrt e > sqrt sv_min]) U^H.
    """
    U, e, Vd = tci.svd(M, 1)
    d = np.where(np.abs(e) > sv_min, np.sqrt(np.abs(e)), 0.0)
    safe = np.where(d == 0, 1.0, d)
    f = np.where(np.abs(d) > math.sqrt(sv_min), safe / (safe * safe), 0.0)
    D = tci.diag(tci.to_cplx(d))
    Fm = tci.diag(tci.to_cplx(f))
    W = tci.contract(U, [0, -1], D, [-1, 1], [0, 1])
    R = tci.contract(W, [0, -1], Vd, [-1, 1], [0, 1])
    Uc, Vc = tci.cplx_conj(U), tci.cplx_conj(Vd)
    W = tci.contract(Fm, [1, -1], Vc, [-1, 0], [0, 1])
    S = tci.contract(W, [0, -1], Uc, [1, -1], [0, 1])
    return R, S


def truncation(st, max_dim, eps):
    """``Truncation`` (include/tnbp/ptns/function/truncation.h:17-222), one rank.

    Per edge in global order: T = Ra^T Rb (normalised), (X, L, Y) =
    trunc_svd(T, chi=max_dim, s_min=eps); both messages become diag(L); the
    bond of A gets Sa^T X sqrt(L), that of B sqrt(L) Y Sb; both sites are
    normalised.  Returns (bond dims, truncation errors) indexed like EdgeIdx.
    """
    I, E, edge_idx = st.I, st.E, st.EdgeIdx
    ne = len(edge_idx)
    dims = [0] * ne
    errs = [0.0] * ne
    for g, (sa, sb) in enumerate(I):
        bonds_a = surrounding_bond_index(sa, I)
        bonds_b = surrounding_bond_index(sb, I)
        direction = 0
        for k, m in enumerate(bonds_a):
            if I[m][0] == sa and I[m][1] == sb:
                ka, tgt, direction = k, m, 0
                break
            if I[m][0] == sb and I[m][1] == sa:
                ka, tgt, direction = k, m, 1
                break
        kb = bonds_b.index(tgt)
        ea = edge_idx.index(tgt)
        if direction == 0:
            Ra, Sa = square_root_and_inverse(E[ea], eps)
            Rb, Sb = square_root_and_inverse(E[ea + ne], eps)
        else:
            Ra, Sa = square_root_and_inverse(E[ea + ne], eps)
            Rb, Sb = square_root_and_inverse(E[ea], eps)
        T = tci.contract(Ra, [-1, 0], Rb, [-1, 1], [0, 1])
        T, _ = tci.normalize(T)
        X, s, Y, err = tci.trunc_svd(T, 1, max_dim, eps)
        dims[ea] = s.size
        errs[ea] = err
        Z = tci.diag(tci.to_cplx(s))
        P = tci.diag(tci.to_cplx(np.sqrt(s)))
        E[ea] = Z.copy()
        E[ea + ne] = Z.copy()
        # site A (truncation.h:156-186)
        Tm = tci.contract(Sa, [-1, 0], X, [-1, 1], [0, 1])
        Tm = tci.contract(Tm, [0, -1], P, [-1, 1], [0, 1])
        ia = st.SiteIdx.index(sa)
        va = st.V[ia]
        la = list(range(va.ndim))
        la[ka] = -1
        va = tci.contract(va, la, Tm, [-1, ka], list(range(va.ndim)))
        st.V[ia], _ = tci.normalize(va)
        # site B (truncation.h:188-217)
        Tm = tci.contract(Y, [0, -1], Sb, [-1, 1], [0, 1])
        Tm = tci.contract(P, [0, -1], Tm, [-1, 1], [0, 1])
        ib = st.SiteIdx.index(sb)
        vb = st.V[ib]
        lb = list(range(vb.ndim))
        lb[kb] = -1
        vb = tci.contract(Tm, [kb, -1], vb, lb, list(range(vb.ndim)))
        st.V[ib], _ = tci.normalize(vb)
    return dims, errs


# ========================================================= measurement
def measure_sites(st, sites, ops):
    """Single-site ``Measure`` (include/tnbp/ptns/function/measure.h:14-94).

    <O> = <W O, conj V> / <W, conj V> with W = V absorbed with every incoming
    message, unnormalised (:54-70, :79-89).
    """
    I, E, edge_idx = st.I, st.E, st.EdgeIdx
    ne = len(edge_idx)
    out = []
    for s, o in zip(sites, ops):
        v = st.V[st.SiteIdx.index(s)]
        W = v.copy()
        bonds = surrounding_bond_index(s, I)
        for m_k, m in enumerate(bonds):
            ea = edge_idx.index(m)
            if I[m][0] == s:
                ea += ne
            la = list(range(W.ndim))
            la[m_k] = -1
            W = tci.contract(W, la, E[ea], [-1, m_k], list(range(W.ndim)))
        n = W.ndim
        lw = list(range(n))
        lw[n - 1] = -1
        R = tci.contract(W, lw, o, [-1, n - 1], list(range(n)))
        full = list(range(-n, 0))
        wd = tci.cplx_conj(v)
        w = tci.get_elem(tci.contract(W, full, wd, full, []), [])
        r = tci.get_elem(tci.contract(R, full, wd, full, []), [])
        out.append(r / w)
    return out


def _edge_env(st, site, bonds, kt):
    """AdagA[p, p', i, j] of measure.h:168-205 (and :207-245 for B)."""
    I, E, edge_idx = st.I, st.E, st.EdgeIdx
    ne = len(edge_idx)
    v = st.V[st.SiteIdx.index(site)]
    A = v.copy()
    n = A.ndim
    for k, m in enumerate(bonds):
        if k != kt:
            ea = edge_idx.index(m)
            if I[m][0] == site:
                ea += ne
            la = list(range(n))
            la[k] = -1
            A = tci.contract(A, la, E[ea], [-1, k], list(range(n)))
            A, _ = tci.normalize(A)
    ia = list(range(-n, 0))
    ic = list(range(-n, 0))
    ia[n - 1], ic[n - 1] = 0, 1
    ia[kt], ic[kt] = 2, 3
    return tci.contract(A, ia, tci.cplx_conj(v), ic, [0, 1, 2, 3])


def measure_edges(st, edge_list, ops):
    """Two-site ``Measure`` (measure.h:99-300): <O> = R / S with
    R = <AdagA . O, BdagB> and S = <Tr AdagA, Tr BdagB>.  The operator index
    order is (a_in, b_in, a_out, b_out).  Results are indexed by position in
    ``edge_list`` (the reference sizes its vector |I| but indexes by term)."""
    I = st.I
    out = []
    for (sa, sb), o in zip(edge_list, ops):
        bonds_a = surrounding_bond_index(sa, I)
        bonds_b = surrounding_bond_index(sb, I)
        ka, tgt = _locate(I, bonds_a, sa, sb)
        kb = bonds_b.index(tgt)
        AA = _edge_env(st, sa, bonds_a, ka)
        BB = _edge_env(st, sb, bonds_b, kb)
        AIA = tci.trace(AA, [(0, 1)])
        BIB = tci.trace(BB, [(0, 1)])
        AO = tci.contract(AA, [-1, -2, 2, 3], o, [-1, 0, -2, 1], [0, 1, 2, 3])
        full = [-4, -3, -2, -1]
        R = tci.get_elem(tci.contract(AO, full, BB, full, []), [])
        S = tci.get_elem(tci.contract(AIA, [-2, -1], BIB, [-2, -1], []), [])
        out.append(R / S)
    return out


def measure_paths(st, paths, ops):
    """Product operators O_1 ... O_k on simple paths of neighbouring sites
    under the BP environment - the extension of ``Measure`` (measure.h:14-300,
    which stops at two sites) that hbp_peps_measure_paths implements; no
    reference counterpart beyond weight 2.

    Stated as one network: every path site's ket layer W_s (V_s with the
    incoming message absorbed on each bond that leaves the path, as in
    measure.h:54-70) with O_s on its physical index, the conjugate bra layer,
    path bonds open between neighbours in both layers; <O> = N / S with S the
    same network without operators.  Contracted in one ``np.einsum`` call,
    independent of the transfer-matrix order the library uses.
    """
    I, E, edge_idx = st.I, st.E, st.EdgeIdx
    ne = len(edge_idx)
    out = []
    for path, o in zip(paths, ops):
        on_path = set()
        for a, b in zip(path, path[1:]):
            on_path.add(_locate(I, surrounding_bond_index(a, I), a, b)[1])
        letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
        ket_lab, bra_lab = {}, {}
        terms_n, terms_s, ops_n, ops_s = [], [], [], []
        for s, op in zip(path, o):
            v = st.V[st.SiteIdx.index(s)]
            bonds = surrounding_bond_index(s, I)
            W = v.copy()
            lk, lb = [], []
            for k, m in enumerate(bonds):
                if m in on_path:
                    if m not in ket_lab:
                        ket_lab[m], bra_lab[m] = next(letters), next(letters)
                    lk.append(ket_lab[m])
                    lb.append(bra_lab[m])
                else:
                    ea = edge_idx.index(m) + (ne if I[m][0] == s else 0)
                    la = list(range(W.ndim))
                    la[k] = -1
                    W = tci.contract(W, la, E[ea], [-1, k], list(range(W.ndim)))
                    c = next(letters)
                    lk.append(c)
                    lb.append(c)
            p, q = next(letters), next(letters)
            terms_n += ["".join(lk) + p, p + q, "".join(lb) + q]
            ops_n += [W, np.asarray(op, dtype=CPLX), np.conj(v)]
            terms_s += ["".join(lk) + p, "".join(lb) + p]
            ops_s += [W, np.conj(v)]
        N = np.einsum(",".join(terms_n) + "->", *ops_n, optimize="greedy")
        S = np.einsum(",".join(terms_s) + "->", *ops_s, optimize="greedy")
        out.append(complex(N / S))
    return out
