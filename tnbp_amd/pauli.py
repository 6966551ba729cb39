"""Sparse Pauli observables on a PEPS.

Reads the reference's text format (include/pauli/sparse_pauli.h:47-105: lines
"re im LABEL" or "re LABEL", '#' and '//' comments, qubit 0 = rightmost
character, pauli_string.h:27-51) and builds the Pauli matrices of
paulitotpo.h:23-36.

The reference measures only weight-1 terms and weight-2 terms on lattice
edges and drops the rest and the coefficients (paulitotpo.h:98-153,
apps/estimator/main.cc:181-240).  Here every term is measured: those two kinds
as the reference does, every other term on a simple path of the lattice
through its support (shortest paths between the support sites in ascending
order, as graph.h:FindShortestPath would give, padded with identities) by
``MeasurePaths`` under the BP environment; the observable value
sum_i c_i <P_i> is returned beside the per-term values.
"""
from __future__ import annotations

import re
from collections import deque

import numpy as np

from . import circuits
from .peps import Measure, MeasureEdges, MeasurePaths

PAULI = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def op_tensor(u):
    """Operator tensor G(in, out) = U[out][in] (qasmtotpo.h:174-190)."""
    return np.ascontiguousarray(np.asarray(u, dtype=np.complex128).T)


def parse_sparse_pauli(text):
    """Terms [(coefficient, label)] from the text of a sparse Pauli file
    (sparse_pauli.h:53-105)."""
    terms = []
    for lineno, raw in enumerate(text.splitlines(), 1):
        line = raw.strip()
        if not line or line.startswith("#") or line.startswith("//"):
            continue
        parts = line.split()
        try:
            if len(parts) == 3:
                coef = complex(float(parts[0]), float(parts[1]))
                label = parts[2]
            elif len(parts) == 2:
                coef = complex(float(parts[0]), 0.0)
                label = parts[1]
            else:
                raise ValueError
        except ValueError:
            raise ValueError(f"parse error at line {lineno}: {raw!r}") from None
        label = label.upper()
        if any(c not in "IXYZ" for c in label):
            raise ValueError("Invalid Pauli label: only I/X/Y/Z are allowed.")
        terms.append((coef, label))
    return terms


def load_sparse_pauli(path):
    if not str(path).endswith(".txt"):
        raise ValueError("only .txt sparse Pauli files are supported")
    with open(path) as f:
        return parse_sparse_pauli(f.read())


_NON_IDENTITY = re.compile("[^I]")


def local_ops(label):
    """Non-identity (qubit, op) pairs, ascending qubit, qubit 0 = last character."""
    n = len(label)
    return [(n - 1 - m.start(), m.group()) for m in _NON_IDENTITY.finditer(label)][::-1]


def _adjacency(edges):
    adj = {}
    for a, b in edges:
        adj.setdefault(a, []).append(b)
        adj.setdefault(b, []).append(a)
    for v in adj.values():
        v.sort()
    return adj


def _shortest(adj, src, dst):
    prev = {src: None}
    q = deque([src])
    while q:
        u = q.popleft()
        if u == dst:
            break
        for w in adj.get(u, []):
            if w not in prev:
                prev[w] = u
                q.append(w)
    if dst not in prev:
        raise ValueError(f"sites {src} and {dst} are not connected")
    path = [dst]
    while path[-1] != src:
        path.append(prev[path[-1]])
    return path[::-1]


def support_path(edges, qubits, adj=None):
    """A simple lattice path through the given sites (ascending order), joined
    by shortest paths; raises ValueError if it would revisit a site.  ``adj``:
    the lattice's adjacency (``_adjacency(edges)``) when the caller has it."""
    if adj is None:
        adj = _adjacency(edges)
    qs = sorted(qubits)
    path = [qs[0]]
    for q in qs[1:]:
        seg = _shortest(adj, path[-1], q)
        path += seg[1:]
    if len(set(path)) != len(path):
        raise ValueError(f"support {qs} does not lie on a simple path of the lattice")
    return path


def term_path_ops(edges, label, adj=None):
    """(path, operators) of one Pauli term; empty for the identity."""
    ops = dict(local_ops(label))
    if not ops:
        return [], []
    path = support_path(edges, list(ops), adj)
    return path, [op_tensor(PAULI[ops.get(s, "I")]) for s in path]


def classify(edges, terms):
    """How ``estimate`` measures each term (host work only, so a server can do
    it when a job arrives): ((indices, sites, operators) of the weight-1
    terms, (indices, site pairs, operators) of the weight-2 terms on lattice
    edges, (indices, paths, operator lists) of the others).

    As ClassifyOps does (paulitotpo.h:98-153), weight-1 terms go to the site
    measurement and weight-2 terms on a lattice edge to the edge measurement
    (operator index order (a_in, b_in, a_out, b_out), qubits ascending,
    paulitotpo.h:62-92); every other term is measured on its support path."""
    edge_set = {frozenset(e) for e in edges}
    adj = _adjacency(edges)
    one, two, paths = ([], [], []), ([], [], []), ([], [], [])
    for i, (_, label) in enumerate(terms):
        lops = local_ops(label)
        if len(lops) == 1:
            q, o = lops[0]
            one[0].append(i), one[1].append(q), one[2].append(op_tensor(PAULI[o]))
        elif len(lops) == 2 and frozenset((lops[0][0], lops[1][0])) in edge_set:
            (qa, oa), (qb, ob) = lops
            u = np.kron(PAULI[ob], PAULI[oa])  # qubit qa is the low bit
            two[0].append(i), two[1].append((qa, qb)), two[2].append(circuits.op_tensor(u))
        elif lops:
            path, o = term_path_ops(edges, label, adj)
            paths[0].append(i), paths[1].append(path), paths[2].append(o)
    return one, two, paths


def estimate(ctx, peps, edges, terms, plan=None):
    """(sum_i c_i <P_i>, [<P_i>]) for terms [(c_i, label_i)] on the PEPS;
    ``plan``: ``classify(edges, terms)`` when the caller already has it."""
    one, two, paths = plan if plan is not None else classify(edges, terms)
    vals = np.ones(len(terms), dtype=np.complex128)
    if one[0]:
        vals[one[0]] = Measure(ctx, peps, one[1], one[2])
    if two[0]:
        vals[two[0]] = MeasureEdges(ctx, peps, two[1], two[2])
    if paths[0]:
        vals[paths[0]] = MeasurePaths(ctx, peps, paths[1], paths[2])
    total = complex(sum(c * v for (c, _), v in zip(terms, vals)))
    return total, list(vals)
