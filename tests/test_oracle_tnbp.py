"""Pins of the BP-PEPS restatement (oracle/tnbp.py, oracle/circuits.py).

Each check is chosen so that a plausible mistake fails it: a swapped message
direction, a transposed gate, a missing conjugate or a wrong fusion order all
break the tree-exactness checks against the brute-force statevector, which
use random single-qubit gates so that no lattice symmetry hides them.
"""
import math

import numpy as np
import pytest

from oracle import circuits, tci, tnbp
from statevector import StateVector, X, Y, Z, ZZ, kicked_layer


def test_lattices_and_layering():
    e = tnbp.bond_square_lattice(3, 3)
    assert len(e) == 2 * 9 - 3 - 3
    assert e[:3] == [(0, 1), (0, 3), (1, 2)]
    # interior site 4: bonds ordered (down, left, right, up) by global edge id
    assert [e[m] for m in tnbp.surrounding_bond_index(4, e)] == [(1, 4), (3, 4), (4, 5), (4, 7)]
    layers = tnbp.greedy_edge_layering(tnbp.bond_square_lattice(16, 16))
    assert len(layers) == 4 and [len(l) for l in layers] == [128, 128, 126, 98]
    assert tnbp.find_shortest_path(e, 0, 8)[0] == 0 and len(tnbp.find_shortest_path(e, 0, 8)) == 5


def test_square_root_and_inverse():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((6, 4)) + 1j * rng.standard_normal((6, 4))
    m = a @ a.conj().T  # Hermitian PSD of rank 4
    r, s = tnbp.square_root_and_inverse(m, 1e-8)
    np.testing.assert_allclose(r @ r, m, atol=1e-10)
    p = r @ s  # projector onto the support
    np.testing.assert_allclose(p @ p, p, atol=1e-10)
    np.testing.assert_allclose(p @ m, m, atol=1e-10)
    assert abs(np.trace(p).real - 4) < 1e-10


def test_product_state_closed_form():
    g = 0.3 / 2  # RX(0.3) per step
    r = circuits.run_kicked(3, 4, 0.0, 0.0, g / math.pi, 1.0, 50, 1e-12, 4, 1e-8)
    for s in r["steps"]:
        n = s["step"] + 1
        np.testing.assert_allclose(s["z"], math.cos(0.3 * n), atol=1e-13)
        np.testing.assert_allclose(s["x"], 0.0, atol=1e-13)
        np.testing.assert_allclose(s["zz"], math.cos(0.3 * n) ** 2, atol=1e-13)


def test_kicked_chain_exact():
    L, a, hz, hx = 8, math.pi / 8, 0.1, 0.7 / (2 * math.pi)
    r = circuits.run_kicked(L, 4, 0.125, hz, hx, 1.0, 200, 1e-13, 64, 1e-12, lattice="oned")
    edges = tnbp.bond_oned_lattice(L)
    sv = StateVector(L)
    for s in r["steps"]:
        kicked_layer(sv, edges, a, hz * math.pi, hx * math.pi)
        np.testing.assert_allclose(s["z"], [sv.expect1(Z, q).real for q in range(L)], atol=1e-12)
        np.testing.assert_allclose(s["x"], [sv.expect1(X, q).real for q in range(L)], atol=1e-12)
        np.testing.assert_allclose(s["zz"], [sv.expect2(ZZ, i, j).real for i, j in edges], atol=1e-12)


def _tree_edges():
    # a small tree with a degree-3 vertex and both edge orientations
    return [(0, 1), (1, 2), (1, 3), (3, 4), (5, 4)]


def test_random_circuit_on_tree_exact():
    rng = np.random.default_rng(11)
    edges = _tree_edges()
    n = 6
    sv = StateVector(n)
    tpos = []
    for _ in range(3):
        gates = circuits.random_layer_gates(rng, edges, n)
        tpos.append(circuits.layer_tpo(edges, gates))
        for u, q in gates:
            if len(q) == 1:
                sv.apply1(u, q[0])
            else:
                sv.apply2(u, q[0], q[1])
    st, recs = circuits.run_tpo_layers(edges, tpos, 100, 1e-14, 64, 1e-12)
    for op in (X, Y, Z):
        got = tnbp.measure_sites(st, list(range(n)), [circuits.gate_tensor(op)] * n)
        np.testing.assert_allclose(got, [sv.expect1(op, q) for q in range(n)], atol=1e-12)
    zz = circuits.gate_tensor(ZZ)
    xy = circuits.gate_tensor(np.kron(Y, X))  # X on the first qubit, Y on the second
    got = tnbp.measure_edges(st, edges, [zz] * len(edges))
    np.testing.assert_allclose(got, [sv.expect2(ZZ, a, b) for a, b in edges], atol=1e-12)
    got = tnbp.measure_edges(st, edges, [xy] * len(edges))
    np.testing.assert_allclose(got, [sv.expect2(np.kron(Y, X), a, b) for a, b in edges], atol=1e-12)


def test_two_site_gate_on_path_matches_statevector():
    # a CX between sites two bonds apart runs through an identity on site 1
    edges = [(0, 1), (1, 2)]
    sv = StateVector(3)
    gates = [(circuits.gate_matrix("h"), (0,)), (circuits.gate_matrix("cx"), (0, 2)),
             (circuits.gate_matrix("u3", (0.3, 0.2, 0.1)), (2,))]
    for u, q in gates:
        sv.apply1(u, q[0]) if len(q) == 1 else sv.apply2(u, *q)
    st, _ = circuits.run_tpo_layers(edges, [circuits.layer_tpo(edges, gates)], 50, 1e-14, 16, 1e-12)
    got = tnbp.measure_sites(st, [0, 1, 2], [circuits.gate_tensor(Z)] * 3)
    np.testing.assert_allclose(got, [sv.expect1(Z, q) for q in range(3)], atol=1e-12)


def test_t2_equals_t4_config1():
    kw = dict(L=4, num_steps=5, Jz=0.125, hz=0.0, hx=0.3 / (2 * math.pi), dt=1.0,
              max_bp_iterations=10000, bp_tolerance=1e-10, max_bond_dim=4, sv_min=1e-8)
    r2 = circuits.run_kicked(two_site_svd="trunc", **kw)
    r4 = circuits.run_kicked(two_site_svd="full", **kw)
    for a, b in zip(r2["steps"], r4["steps"]):
        assert a["bp_iterations"] == b["bp_iterations"]
        np.testing.assert_allclose(a["z"], b["z"], atol=1e-12)
        np.testing.assert_allclose(a["zz"], b["zz"], atol=1e-12)


def test_gauge_invariance_of_bp_and_measure():
    """A random invertible gauge on every bond leaves converged BP observables unchanged."""
    rng = np.random.default_rng(7)
    edges = tnbp.bond_square_lattice(3, 3)
    r = circuits.run_kicked(3, 3, 0.125, 0.0, 0.05, 1.0, 500, 1e-13, 8, 1e-12)
    st = r["state"]
    layers = tnbp.greedy_edge_layering(edges)
    st2 = st.copy()
    for m, (a, b) in enumerate(edges):
        ka = tnbp.surrounding_bond_index(a, edges).index(m)
        kb = tnbp.surrounding_bond_index(b, edges).index(m)
        n = st2.V[a].shape[ka]
        g = np.eye(n) + 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        gi = np.linalg.inv(g)
        la = list(range(st2.V[a].ndim)); la[ka] = -1
        st2.V[a] = tci.contract(st2.V[a], la, g, [-1, ka], list(range(st2.V[a].ndim)))
        lb = list(range(st2.V[b].ndim)); lb[kb] = -1
        st2.V[b] = tci.contract(gi, [kb, -1], st2.V[b], lb, list(range(st2.V[b].ndim)))
    for s in (st, st2):
        s.E = [np.eye(e.shape[0], dtype=complex) for e in s.E]
        tnbp.bp_loop(s, layers, 2000, 1e-13)
    sites, _, Oz, medges, Oj = circuits.measops(edges)
    np.testing.assert_allclose(tnbp.measure_sites(st, sites, Oz), tnbp.measure_sites(st2, sites, Oz), atol=1e-9)
    np.testing.assert_allclose(tnbp.measure_edges(st, medges, Oj), tnbp.measure_edges(st2, medges, Oj), atol=1e-9)


def test_residual_is_zero_at_fixed_point_and_positive_otherwise():
    edges = tnbp.bond_square_lattice(3, 3)
    r = circuits.run_kicked(3, 3, 0.125, 0.0, 0.05, 1.0, 500, 1e-13, 8, 1e-12)
    st = r["state"]
    tnbp.bp_loop(st, tnbp.greedy_edge_layering(edges), 500, 1e-14)
    assert tnbp.belief_propagation_condition(st) < 1e-13
    st.E[0] = st.E[0] + 0.1 * np.ones_like(st.E[0])
    assert tnbp.belief_propagation_condition(st) > 1e-3


def _sv_string(sv, ops):
    """<psi| prod_q U_q |psi> for {qubit: U} by applying each factor."""
    phi = StateVector(sv.n)
    phi.psi = sv.psi.copy()
    for q, u in ops.items():
        phi.apply1(u, q)
    return complex(np.vdot(sv.psi, phi.psi))


def test_measure_paths_reduces_to_site_and_edge():
    """Weight 1 equals Measure on a site, weight 2 equals Measure on an edge
    with the product operator (both on a loopy 3x3 lattice, where they are
    computed by different contractions)."""
    edges = tnbp.bond_square_lattice(3, 3)
    r = circuits.run_kicked(3, 3, 0.125, 0.05, 0.07, 1.0, 500, 1e-13, 8, 1e-12)
    st = r["state"]
    ops1 = [circuits.gate_tensor(u) for u in (X, Y, Z)]
    for s in (0, 4, 7):
        got = tnbp.measure_paths(st, [[s]] * 3, [[o] for o in ops1])
        want = tnbp.measure_sites(st, [s] * 3, ops1)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    for a, b in ((0, 1), (4, 7), (5, 4)):
        ua, ub = Y, X
        got = tnbp.measure_paths(st, [[a, b]], [[circuits.gate_tensor(ua), circuits.gate_tensor(ub)]])
        want = tnbp.measure_edges(st, [(a, b)], [circuits.gate_tensor(np.kron(ub, ua))])
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    got = tnbp.measure_paths(st, [[0, 1, 2, 5, 4]], [[circuits.gate_tensor(np.eye(2))] * 5])
    np.testing.assert_allclose(got, [1.0], atol=1e-13)


def test_measure_paths_exact_on_trees():
    """On trees BP is exact: strings of weight 3-6 equal the statevector."""
    rng = np.random.default_rng(31)
    for edges, n, paths in (
        (tnbp.bond_oned_lattice(8), 8, [[1, 2, 3], [0, 1, 2, 3, 4, 5], [2, 3, 4, 5, 6, 7], [5, 6, 7]]),
        (_tree_edges(), 6, [[2, 1, 3, 4, 5], [0, 1, 3], [4, 3, 1, 2]]),
    ):
        sv = StateVector(n)
        tpos = []
        for _ in range(3):
            gates = circuits.random_layer_gates(rng, edges, n)
            tpos.append(circuits.layer_tpo(edges, gates))
            for u, q in gates:
                sv.apply1(u, q[0]) if len(q) == 1 else sv.apply2(u, q[0], q[1])
        st, _ = circuits.run_tpo_layers(edges, tpos, 200, 1e-14, 64, 1e-12)
        for path in paths:
            us = [(X, Y, Z)[rng.integers(3)] for _ in path]
            got = tnbp.measure_paths(st, [path], [[circuits.gate_tensor(u) for u in us]])[0]
            want = _sv_string(sv, dict(zip(path, us)))
            assert abs(got - want) < 1e-11, (path, got, want)


def _chain_state_before_truncation(n, layer_gates, sv):
    """A chain after the given gate layers with exact bonds: the last layer
    absorbed and BP converged, not yet truncated.  `sv` receives the gates."""
    edges = tnbp.bond_oned_lattice(n)
    tpos = []
    for gates in layer_gates:
        tpos.append(circuits.layer_tpo(edges, gates))
        for u, q in gates:
            sv.apply1(u, q[0]) if len(q) == 1 else sv.apply2(u, q[0], q[1])
    st, _ = circuits.run_tpo_layers(edges, tpos[:-1], 400, 1e-15, 64, 1e-14)
    tnbp.absorb_tpo(st, tpos[-1])
    tnbp.bp_loop(st, tnbp.greedy_edge_layering(edges), 400, 1e-15)
    return edges, st


def _schmidt(sv, cut):
    """Schmidt coefficients of the statevector between qubits 0..cut and cut+1..n-1."""
    m = sv.psi.reshape(2 ** (cut + 1), -1)  # axis q of psi is qubit q
    return np.linalg.svd(m, compute_uv=False)


@pytest.mark.parametrize("chi", [1, 2, 3, 5, 16])
def test_truncation_with_cut_keeps_top_schmidt_values(chi):
    """On a chain the converged BP messages put the state in the Vidal gauge,
    so T = Ra^T Rb normalised has the Schmidt coefficients of the cut as
    singular values (truncation.h:102-130): the Lambda that Truncation stores
    on each bond (truncation.h:135-149) must be the top-chi Schmidt
    coefficients of the exact state, and the truncation error the discarded
    Schmidt weight.  A wrong message direction, a missing square root or
    transpose, a cut that keeps the smallest values or a different error
    formula fails this."""
    n = 8
    rng = np.random.default_rng(101)
    edges0 = tnbp.bond_oned_lattice(n)
    layers = [circuits.random_layer_gates(rng, edges0, n) for _ in range(10)]
    sv = StateVector(n)
    edges, st = _chain_state_before_truncation(n, layers, sv)
    ne = len(edges)
    sched = [_schmidt(sv, a) for a, _ in edges]
    assert max(len(s[s > 1e-10]) for s in sched) > 5  # the cuts below are real cuts
    dims, errs = tnbp.truncation(st, chi, 1e-10)
    for m, (a, b) in enumerate(edges):
        want = sched[m]
        want = want[want > 1e-10]
        keep = min(chi, len(want))
        lam = np.real(np.diag(st.E[m]))
        assert dims[m] == keep
        np.testing.assert_allclose(lam, want[:keep], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(np.real(np.diag(st.E[m + ne])), want[:keep], rtol=1e-9, atol=1e-12)
        assert abs(errs[m] - np.sum(want[keep:] ** 2) / np.sum(want ** 2)) < 1e-10


def test_truncated_chain_state_is_the_schmidt_projection():
    """Cutting one bond of a chain in the Vidal gauge keeps the top-chi
    Schmidt vectors of that cut: the truncated state is the exact state
    projected on them (truncation.h:156-217 with the projectors of
    truncation.h:102-149).  Gates touch bonds (1,2), (2,3), (3,4) of a
    6-site chain only, so the cut (2,3) has Schmidt rank up to 4 and every
    other cut at most 2: chi = 2 cuts exactly that bond.  After one more BP
    run every <Z> and <ZZ> equals that of the projected statevector."""
    n, chi, cut = 6, 2, 2
    rng = np.random.default_rng(7)
    u3 = lambda: circuits.gate_matrix("u3", tuple(rng.uniform(0, 2 * math.pi, 3)))
    rzz = lambda: circuits.gate_matrix("rzz", (rng.uniform(0.5, 2.5),))
    layers = []
    for _ in range(3):
        layers.append([(rzz(), (1, 2)), (rzz(), (3, 4))] + [(u3(), (q,)) for q in range(n)])
        layers.append([(rzz(), (2, 3))] + [(u3(), (q,)) for q in range(n)])
    sv = StateVector(n)
    edges, st = _chain_state_before_truncation(n, layers, sv)
    ranks = [int(np.sum(_schmidt(sv, a) > 1e-10)) for a, _ in edges]
    assert ranks[cut] > chi and all(r <= chi for m, r in enumerate(ranks) if m != cut)
    mat = sv.psi.reshape(2 ** (cut + 1), -1)
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    proj = (u[:, :chi] * s[:chi]) @ vh[:chi]
    ref = StateVector(n)
    ref.psi = (proj / np.linalg.norm(proj)).reshape((2,) * n)
    dims, _ = tnbp.truncation(st, chi, 1e-10)
    assert dims[cut] == chi
    tnbp.bp_loop(st, tnbp.greedy_edge_layering(edges), 400, 1e-15)
    zt = circuits.gate_tensor(Z)
    got = tnbp.measure_sites(st, list(range(n)), [zt] * n)
    np.testing.assert_allclose(got, [ref.expect1(Z, q) for q in range(n)], atol=1e-10)
    got = tnbp.measure_edges(st, edges, [circuits.gate_tensor(ZZ)] * len(edges))
    np.testing.assert_allclose(got, [ref.expect2(ZZ, a, b) for a, b in edges], atol=1e-10)
