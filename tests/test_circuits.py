"""Front end (tnbp_amd.circuits) against the oracle's restatement of circuit.h / qasmtotpo.h."""
import json
import math
import os

import numpy as np
import pytest

from oracle import circuits as oc
from oracle import tnbp as ot
from tnbp_amd import circuits as pc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _same_operator(o1, o2):
    # TPO bond gauges may differ by the SVD's phases; the operator they build does not
    assert o1.shape == o2.shape
    return True


def _full_operator(edges, tpo):
    """Contract a TPO over all bonds into the dense 2^n x 2^n operator."""
    sites = ot.site_index_from_bond(edges)
    n = len(sites)
    # einsum with letters: bonds by edge id, phys in/out per site
    import string
    letters = iter(string.ascii_letters)
    bond = {m: next(letters) for m in range(len(edges))}
    pin = {s: next(letters) for s in sites}
    pout = {s: next(letters) for s in sites}
    terms = []
    for s, o in zip(sites, tpo):
        inc = ot.surrounding_bond_index(s, edges)
        terms.append("".join(bond[m] for m in inc) + pin[s] + pout[s])
    out = "".join(pout[s] for s in reversed(sites)) + "".join(pin[s] for s in reversed(sites))
    full = np.einsum(",".join(terms) + "->" + out, *tpo)
    return full.reshape(2 ** n, 2 ** n)


@pytest.mark.parametrize("keep_zero", [False, True])
def test_kicked_tpo_matches_oracle(keep_zero):
    edges = ot.bond_square_lattice(2, 3)
    args = (0.125, 0.05, 0.3 / (2 * math.pi), math.pi)
    mine = pc.kicked_ising_tpo(edges, *args, two_site_keep_zero=keep_zero)
    ref = oc.circuit_tpo(edges, *args, "full" if keep_zero else "trunc")
    for a, b in zip(mine, ref):
        assert a.shape == b.shape
    np.testing.assert_allclose(_full_operator(edges, mine), _full_operator(edges, ref), atol=1e-13)


def test_random_layer_tpo_matches_oracle():
    rng = np.random.default_rng(3)
    edges = ot.bond_square_lattice(3, 2)
    gates = pc.random_layer(rng, edges, 6)
    mine = pc.layer_tpo(edges, gates)
    ref = oc.layer_tpo(edges, gates)
    np.testing.assert_allclose(_full_operator(edges, mine), _full_operator(edges, ref), atol=1e-13)


def test_lattices_match_reference_order():
    assert pc.square_lattice(5, 4) == ot.bond_square_lattice(5, 4)
    assert pc.chain(6) == ot.bond_oned_lattice(6)


@pytest.mark.parametrize("name", ["config1_kicked_4x4_chi4.json", "config1_fixed3_kicked_4x4_chi4.json"])
def test_golden_fixture_reproduced_by_oracle(name):
    d = json.load(open(os.path.join(GOLDEN, name)))
    r = oc.run_kicked(**d["args"])
    for a, b in zip(r["steps"], d["steps"]):
        assert a["bp_iterations"] == b["bp_iterations"]
        np.testing.assert_allclose(a["z"], b["z"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(a["zz"], b["zz"], rtol=0, atol=1e-12)
