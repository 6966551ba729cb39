"""CPU restatement (numpy) of the circuit front ends and the kicked-Ising driver.

TEST INFRASTRUCTURE ONLY (the parity oracle).  See ``oracle/tci.py``.
"""
from __future__ import annotations

import math

import numpy as np

from . import tci
from . import tnbp

CPLX = np.complex128


# ===================================================== gate matrices
def gate_matrix(name, params=()):
    """``InstructionMatrix`` (include/tnbp/parser/qasmtotpo.h:19-144), row-major U.

    Two-qubit gates: the first listed qubit is the least-significant bit of the
    4x4 row/column index.
    """
    c = lambda x: complex(x)
    e = lambda x: complex(math.cos(x), math.sin(x))
    if name == "u3":
        th, ph, la = params
        return np.array([[math.cos(th / 2), -e(la) * math.sin(th / 2)],
                         [e(ph) * math.sin(th / 2), e(ph + la) * math.cos(th / 2)]], dtype=CPLX)
    if name == "rx":
        th = params[0]
        return np.array([[math.cos(th / 2), c(-1j * math.sin(th / 2))],
                         [c(-1j * math.sin(th / 2)), math.cos(th / 2)]], dtype=CPLX)
    if name == "ry":
        th = params[0]
        return np.array([[math.cos(th / 2), -math.sin(th / 2)],
                         [math.sin(th / 2), math.cos(th / 2)]], dtype=CPLX)
    if name == "rz":
        th = params[0]
        return np.array([[e(-th / 2), 0], [0, e(th / 2)]], dtype=CPLX)
    if name == "h":
        s = 1 / math.sqrt(2.0)
        return np.array([[s, s], [s, -s]], dtype=CPLX)
    if name == "x":
        return np.array([[0, 1], [1, 0]], dtype=CPLX)
    if name == "y":
        return np.array([[0, -1j], [1j, 0]], dtype=CPLX)
    if name == "z":
        return np.array([[1, 0], [0, -1]], dtype=CPLX)
    if name == "rzz":
        th = params[0]
        return np.diag([e(-th / 2), e(th / 2), e(th / 2), e(-th / 2)]).astype(CPLX)
    if name == "cx":
        m = np.zeros((4, 4), dtype=CPLX)
        m[0, 0] = m[2, 2] = m[1, 3] = m[3, 1] = 1
        return m
    if name == "cz":
        return np.diag([1, 1, 1, -1]).astype(CPLX)
    raise ValueError(f"gate {name} not in the restated subset")


def gate_tensor(u):
    """``InstructionTensor`` (qasmtotpo.h:174-190): row-major U stored through
    ``address_from_coor`` so G(in..., out...) = U[out][in]."""
    nq = int(round(math.log2(u.shape[0])))
    shape = [2] * (2 * nq)
    data = u.reshape(-1)  # row-major flattening: data[r * n + c] = U[r][c]
    return tci.assign_from_range(shape, data, lambda co: tci.address_from_coor(shape, co))


# ============================================ kicked-Ising TPO and ops
def circuit_tpo(edges, Jz, hz, hx, dt, two_site_svd="full"):
    """``CircuitTPO`` (apps/kicked_ising_lattice/circuit.h:14-229).

    Per-site operator O_s(t_1..t_z, d_in, d_out) for one Floquet step
    Ux . Uz . prod_edges exp(-i dt Jz ZZ).  ``two_site_svd="full"`` is the
    reference as written (``tci::svd``, circuit.h:62: t = 4 with two zero
    singular values); ``"trunc"`` drops singular values <= 1e-12 the way the
    QASM path does (qasmtotpo.h:303-307), giving t = 2.
    """
    sites = tnbp.site_index_from_bond(edges)
    a = dt * Jz
    dj = [complex(math.cos(a), -math.sin(a)), 0, 0, 0,
          0, complex(math.cos(a), math.sin(a)), 0, 0,
          0, 0, complex(math.cos(a), math.sin(a)), 0,
          0, 0, 0, complex(math.cos(a), -math.sin(a))]
    Uj = tci.assign_from_range([2, 2, 2, 2], dj,
                               lambda c: c[0] + c[1] * 2 + c[2] * 4 + c[3] * 8)
    Uj = tci.transpose(Uj, [0, 2, 1, 3])
    if two_site_svd == "full":
        Ua, S, Ub = tci.svd(Uj, 2)
    else:
        Ua, S, Ub, _ = tci.trunc_svd(Uj, 2, 4, 1.0e-12)
    D = tci.diag(tci.to_cplx(np.sqrt(S)))
    Va = tci.contract(Ua, [0, 1, -1], D, [-1, 2], [2, 0, 1])
    Vb = tci.contract(D, [0, -1], Ub, [-1, 1, 2], [0, 1, 2])
    b = dt * hz
    Uz = tci.assign_from_range([2, 2], [complex(math.cos(b), -math.sin(b)), 0, 0,
                                        complex(math.cos(b), math.sin(b))],
                               lambda c: c[0] + c[1] * 2)
    g = dt * hx
    Ux = tci.assign_from_range([2, 2], [math.cos(g), complex(0, -math.sin(g)),
                                        complex(0, -math.sin(g)), math.cos(g)],
                               lambda c: c[0] + c[1] * 2)
    V = [None] * len(sites)
    for s in sites:
        nb = len(tnbp.surrounding_bond_index(s, edges))
        shape = [1] * nb + [2, 2]
        V[s] = tci.assign_from_range(shape, [1, 0, 0, 1],
                                     lambda c, sh=shape: tci.address_from_coor(sh, c))
    for m, (sa, sb) in enumerate(edges):
        for site, G in ((sa, Va), (sb, Vb)):
            T = V[site]
            n = T.ndim
            bonds = tnbp.surrounding_bond_index(site, edges)
            lab, lt, lg = 0, [0] * n, [0, 0, 0]
            shape_w = list(T.shape)
            for k, bm in enumerate(bonds):
                lt[k] = lab
                lab += 1
                if bm == m:
                    lg[0] = lab
                    lab += 1
                    shape_w[k] = T.shape[k] * G.shape[0]
            lt[n - 2] = lab
            lab += 1
            lt[n - 1] = -1
            lg[1] = -1
            lg[2] = lab
            lab += 1
            W = tci.contract(T, lt, G, lg, list(range(n + 1)))
            V[site] = tci.reshape(W, shape_w)
    for U in (Uz, Ux):
        for i in range(len(sites)):
            n = V[i].ndim
            lv = list(range(n))
            lv[n - 1] = -1
            V[i] = tci.contract(V[i], lv, U, [-1, n - 1], list(range(n)))
    return V


def measops(edges):
    """``measops`` (apps/kicked_ising_lattice/measop.h:1-73): X and Z on every
    site, ZZ on every edge, stored column-major from row-major data."""
    sites = tnbp.site_index_from_bond(edges)
    Z = tci.assign_from_range([2, 2], [1, 0, 0, -1], lambda c: c[0] + c[1] * 2)
    X = tci.assign_from_range([2, 2], [0, 1, 1, 0], lambda c: c[0] + c[1] * 2)
    dj = [1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1]
    J = tci.assign_from_range([2, 2, 2, 2], dj, lambda c: c[0] + c[1] * 2 + c[2] * 4 + c[3] * 8)
    return (list(sites), [X.copy() for _ in sites], [Z.copy() for _ in sites],
            list(edges), [J.copy() for _ in edges])


# ======================================================= QASM-style TPO
def init_qasm_tpo(edges):
    """``InitQasmParseTPO`` (qasmtotpo.h:200-231): identity operator, bonds of dim 1."""
    sites = tnbp.site_index_from_bond(edges)
    T = []
    for s in sites:
        nb = len(tnbp.surrounding_bond_index(s, edges))
        shape = [1] * nb + [2, 2]
        T.append(tci.assign_from_range(shape, [1, 0, 0, 1],
                                       lambda c, sh=shape: tci.address_from_coor(sh, c)))
    return T


def append_gate_to_tpo(T, u, qubits, edges):
    """``AppendGateToTPO`` (qasmtotpo.h:242-437) for one- and two-qubit gates.

    A one-qubit gate acts after the operator already stored (:276-285).  A
    two-qubit gate is split by trunc_svd(chi <= 4, s_min = 1e-12) into
    A(a_in, a_out, k) sqrt(S) and sqrt(S) B(k, b_in, b_out) (:292-331) and laid
    along the BFS shortest path, identity on intermediate sites (:333-436).
    """
    sites = tnbp.site_index_from_bond(edges)
    G = gate_tensor(u)
    if G.ndim == 2:
        s = qubits[0]
        ia = sites.index(s)
        nv = T[ia].ndim - 2
        it = list(range(nv + 2))
        it[nv + 1] = -1
        T[ia] = tci.contract(T[ia], it, G, [-1, nv + 1], list(range(nv + 2)))
        return
    if G.ndim != 4:
        raise ValueError("only one- and two-qubit gates are restated")
    G = tci.transpose(G, [0, 2, 1, 3])
    A, S, B, _ = tci.trunc_svd(G, 2, 4, 1.0e-12)
    S = np.where(np.abs(S) > 1.0e-12, np.sqrt(np.abs(S)), 0.0)
    D = tci.diag(tci.to_cplx(S))
    A = tci.contract(A, [1, 2, -1], D, [-1, 0], [0, 1, 2])
    B = tci.contract(D, [0, -1], B, [-1, 1, 2], [0, 1, 2])
    vdim = D.shape[0]
    sa, sb = qubits
    path = tnbp.find_shortest_path(edges, sa, sb)
    for i, p in enumerate(path):
        if 0 < i < len(path) - 1:
            X = tci.diag(np.ones(2 * vdim, dtype=CPLX))
            X = tci.reshape(X, (vdim, 2, vdim, 2))
            X = tci.transpose(X, [0, 2, 1, 3])
        elif i == 0:
            X = A.copy()
        else:
            X = B.copy()
        vb = tnbp.surrounding_bond_index(p, edges)
        ia = sites.index(p)
        TA = T[ia]
        nb = len(vb)
        tbm = nb + 2
        if i > 0:
            tbm = 0
            for k in vb:
                if set(edges[k]) == {p, path[i - 1]}:
                    break
                tbm += 1
        tbp = nb + 2
        if i < len(path) - 1:
            tbp = 0
            for k in vb:
                if set(edges[k]) == {p, path[i + 1]}:
                    break
                tbp += 1
        ka = kx = 0
        ia_l = [0] * TA.ndim
        ix_l = [0] * X.ndim
        new_shape = [0] * TA.ndim
        for bnd in range(nb):
            if bnd == tbm or bnd == tbp:
                ia_l[bnd] = ka
                ka += 1
                ix_l[kx] = ka
                kx += 1
                ka += 1
                new_shape[bnd] = vdim * TA.shape[bnd]
            else:
                ia_l[bnd] = ka
                ka += 1
                new_shape[bnd] = TA.shape[bnd]
        ia_l[nb] = ka
        ka += 1
        ia_l[nb + 1] = -1
        ix_l[kx] = -1
        kx += 1
        ix_l[kx] = ka
        ka += 1
        new_shape[nb] = TA.shape[nb]
        new_shape[nb + 1] = TA.shape[nb + 1]
        R = tci.contract(X, ix_l, TA, ia_l, list(range(X.ndim + TA.ndim - 2)))
        T[ia] = tci.reshape(R, new_shape)


# ================================================== kicked-Ising driver
def run_kicked(L, num_steps, Jz, hz, hx, dt, max_bp_iterations, bp_tolerance,
               max_bond_dim, sv_min, lattice="square", two_site_svd="trunc",
               measure_every_step=True, log=None):
    """Main loop of ``kicked_ising_exvalue`` (apps/kicked_ising_lattice/main.cc:21-238).

    ``dt`` is in units of pi as on the reference's command line (option.h:62).
    Returns a dict of per-step records: BP iterations, residuals, bond dims,
    truncation errors and the X, Z (per site) and ZZ (per edge) values.
    """
    if lattice == "square":
        edges = tnbp.bond_square_lattice(L, L)
    elif lattice == "oned":
        edges = tnbp.bond_oned_lattice(L, 0)
    else:
        raise ValueError(lattice)
    layers = tnbp.greedy_edge_layering(edges)
    dt_rad = dt * math.pi
    tpo = circuit_tpo(edges, Jz, hz, hx, dt_rad, two_site_svd)
    qubit = tnbp.site_index_from_bond(edges)
    st = tnbp.init_tensor_product_state(edges, [2] * len(qubit))
    msites, Ox, Oz, medges, Oj = measops(edges)
    steps = []
    for step in range(num_steps):
        tnbp.absorb_tpo(st, tpo)
        it, res = tnbp.bp_loop(st, layers, max_bp_iterations, bp_tolerance)
        dims, errs = tnbp.truncation(st, max_bond_dim, sv_min)
        rec = {"step": step, "bp_iterations": it, "residuals": res,
               "bond_dims": dims, "trunc_errs": errs}
        if measure_every_step or step == num_steps - 1:
            rec["x"] = [v.real for v in tnbp.measure_sites(st, msites, Ox)]
            rec["z"] = [v.real for v in tnbp.measure_sites(st, msites, Oz)]
            rec["zz"] = [v.real for v in tnbp.measure_edges(st, medges, Oj)]
        steps.append(rec)
        if log is not None:
            log(rec)
    return {"edges": edges, "steps": steps, "state": st}


# ================================================ layered-circuit driver
def layer_tpo(edges, gates):
    """One TPO slice built gate by gate (``QasmToTPO`` per barrier slice,
    qasmtotpo.h:709-767).  ``gates`` is a list of (U, qubits)."""
    T = init_qasm_tpo(edges)
    for u, q in gates:
        append_gate_to_tpo(T, u, list(q), edges)
    return T


def run_tpo_layers(edges, tpos, max_bp_iterations, bp_tolerance, max_bond_dim, sv_min):
    """Estimator step loop (apps/estimator/main.cc:119-178): per TPO slice
    AbsorbTPO, the BP loop to tolerance, then Truncation."""
    layers = tnbp.greedy_edge_layering(edges)
    sites = tnbp.site_index_from_bond(edges)
    st = tnbp.init_tensor_product_state(edges, [2] * len(sites))
    recs = []
    for tpo in tpos:
        tnbp.absorb_tpo(st, tpo)
        it, res = tnbp.bp_loop(st, layers, max_bp_iterations, bp_tolerance)
        dims, errs = tnbp.truncation(st, max_bond_dim, sv_min)
        recs.append({"bp_iterations": it, "residuals": res, "bond_dims": dims,
                     "trunc_errs": errs})
    return st, recs


def random_layer_gates(rng, edges, n_sites):
    """One layer of the ragged-batch circuit family: RZZ(theta ~ U[0, pi]) on a
    greedy random matching of the edges, then U3 with angles ~ U[0, 2 pi] on
    every site."""
    order = rng.permutation(len(edges))
    used, gates = set(), []
    for m in order:
        a, b = edges[m]
        if a in used or b in used:
            continue
        used.update((a, b))
        gates.append((gate_matrix("rzz", (rng.uniform(0, math.pi),)), (a, b)))
    for s in range(n_sites):
        gates.append((gate_matrix("u3", tuple(rng.uniform(0, 2 * math.pi, 3))), (s,)))
    return gates


def opt_tpo_by_svd(edges, O, eps):
    """``OptTPObySVD`` (include/tnbp/ptns/function/optsvdtpo.h:87-214): for
    every edge in order, contract the two site operators over their shared
    bond, trunc_svd with rows = the first site's other indices (chi_max = the
    bond's dimension, s_min = eps), and put sqrt(S) on both sides, the new
    bond back at its position.  Returns the new list (order of
    ``site_index_from_bond``)."""
    sites = tnbp.site_index_from_bond(edges)
    O = [o.copy() for o in O]
    for m, (a, b) in enumerate(edges):
        ia, ib = sites.index(a), sites.index(b)
        ka = tnbp.surrounding_bond_index(a, edges).index(m)
        kb = tnbp.surrounding_bond_index(b, edges).index(m)
        na, nb = O[ia].ndim, O[ib].ndim
        la, lb, lc = [0] * na, [0] * nb, []
        u = 0
        for k in range(na):
            if k == ka:
                la[k] = -1
            else:
                la[k] = u
                lc.append(u)
                u += 1
        for k in range(nb):
            if k == kb:
                lb[k] = -1
            else:
                lb[k] = u
                lc.append(u)
                u += 1
        C = tci.contract(O[ia], la, O[ib], lb, lc)
        U, S, V, _ = tci.trunc_svd(C, na - 1, O[ia].shape[ka], eps)
        D = np.diag(np.sqrt(S)).astype(np.complex128)
        lu = [0] * na
        for k in range(na):
            if k == ka:
                lu[na - 1] = -1
            elif k < ka:
                lu[k] = k
            else:
                lu[k - 1] = k
        O[ia] = tci.contract(U, lu, D, [-1, ka], list(range(na)))
        lv = [0] * nb
        for k in range(nb):
            if k == kb:
                lv[0] = -1
            elif k < kb:
                lv[k + 1] = k
            else:
                lv[k] = k
        O[ib] = tci.contract(D, [kb, -1], V, lv, list(range(nb)))
    return O
