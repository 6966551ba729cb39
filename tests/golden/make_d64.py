"""Golden fixture of the BP kernels at bond dimension 64 (the 16x16 chi=32
steady-state shape) from the oracle.

A 3x3 square lattice with every bond at D = 64, seeded random site tensors
and messages either diagonal (the state after AbsorbTPO / Truncation) or
random Hermitian positive definite.  The oracle runs two BP sweeps (bp.h:65-261)
with the condition after each (bp.h:266-355), then Truncation(32, 1e-8)
(truncation.h:17-222) and <Z> on every site, <ZZ> on every edge
(measure.h:14-300).  A full message is 64 x 64 complex; the fixture keeps per
message its Frobenius norm and its bilinear forms u_k^H M v_k with four fixed
unit probe vectors (a difference X between two messages gives
|u^H X v| <= ||X||_F, and random probes see any X != 0), so that the file
stays small.  The inputs are regenerated from the seeds by `inputs()`.

Run from the repo root (about 3 minutes on 8 cores):
    python tests/golden/make_d64.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import circuits as oc  # noqa: E402
from oracle import tnbp as ot  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "d64_3x3.json")
L, D, SEED, PROBE_SEED = 3, 64, 64, 99
EDGES = ot.bond_square_lattice(L, L)
NE = len(EDGES)


def inputs(diag):
    """Seeded site tensors and messages (the same numbers on every machine
    with numpy's PCG64 default generator)."""
    rng = np.random.default_rng(SEED)
    bonds = [ot.surrounding_bond_index(s, EDGES) for s in range(L * L)]
    V = []
    for s in range(L * L):
        shape = [D] * len(bonds[s]) + [2]
        v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        V.append(v / np.linalg.norm(v))
    E = []
    for _ in range(2 * NE):
        if diag:
            e = np.diag(rng.uniform(0.2, 1.0, D)).astype(complex)
        else:
            a = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
            e = a @ a.conj().T + D * np.eye(D)
        E.append(e / np.linalg.norm(e))
    return V, E


def probes():
    rng = np.random.default_rng(PROBE_SEED)
    u = rng.standard_normal((4, D)) + 1j * rng.standard_normal((4, D))
    v = rng.standard_normal((4, D)) + 1j * rng.standard_normal((4, D))
    return (u / np.linalg.norm(u, axis=1, keepdims=True),
            v / np.linalg.norm(v, axis=1, keepdims=True))


def summary(M):
    """[||M||_F, re/im of u_k^H M v_k for k = 0..3]."""
    u, v = probes()
    f = [float(np.linalg.norm(M))]
    for k in range(4):
        x = complex(u[k].conj() @ M @ v[k])
        f += [x.real, x.imag]
    return f


ZT = oc.gate_tensor(np.diag([1.0, -1.0]).astype(complex))
ZZT = oc.gate_tensor(np.diag([1.0, -1.0, -1.0, 1.0]).astype(complex))


def run(diag):
    V, E = inputs(diag)
    st = ot.init_tensor_product_state(EDGES, [2] * (L * L))
    st.V = [v.copy() for v in V]
    st.E = [e.copy() for e in E]
    layers = ot.greedy_edge_layering(EDGES)
    out = {"messages": [], "residuals": []}
    for _ in range(2):
        F = [None] * (2 * NE)
        for J in layers:
            ot.belief_propagation(st, J, F)
        st.E = F
        out["messages"].append([summary(f) for f in F])
        out["residuals"].append(ot.belief_propagation_condition(st))
    dims, errs = ot.truncation(st, 32, 1e-8)
    out["bond_dims"] = dims
    out["trunc_errs"] = errs
    out["lambda"] = [np.real(np.diag(st.E[m])).tolist() for m in range(NE)]
    out["z"] = [complex(x).real for x in ot.measure_sites(st, list(range(L * L)), [ZT] * (L * L))]
    out["zz"] = [complex(x).real for x in ot.measure_edges(st, EDGES, [ZZT] * NE)]
    return out


def main():
    res = {"L": L, "D": D, "seed": SEED, "probe_seed": PROBE_SEED, "chi": 32, "sv_min": 1e-8,
           "edges": EDGES, "cases": {}}
    for diag in (True, False):
        res["cases"]["diag" if diag else "dense"] = run(diag)
        print("case done", diag, flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
