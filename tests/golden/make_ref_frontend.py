"""Golden fixture of the reference's front end, from its own sources.

oracle/_ref/tnbp_ref (built by oracle/Makefile from the reference's
gqten-free headers: lattices, GreedyEdgeLayering, FindShortestPath, the
OpenQASM 2/3 parsers, the sparse-Pauli reader) is run on the seeded inputs of
`cases()`; the fixture keeps its outputs, so the tests pin this build's
restatements against the reference's code also where the binary is absent.

Run from the repo root after `make -C oracle`:
    python tests/golden/make_ref_frontend.py
"""
import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref", "tnbp_ref")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_frontend.json")

ONE = ["u3", "u", "u2", "u1", "p", "rx", "ry", "rz", "h", "x", "y", "z", "s", "sdg", "t", "tdg", "id"]
TWO = ["cx", "cnot", "cz", "swap", "rzz"]
THREE = ["ccx", "toffoli", "cswap"]
NPAR = {"u3": 3, "u": 3, "u2": 2, "u1": 1, "p": 1, "rx": 1, "ry": 1, "rz": 1, "rzz": 1}
EXPR = ["pi/2", "-pi/4", "0.3", "2*pi/3", "tau/8", "-(0.25+pi)", "1.5e-1", "PI/7", "3*(pi-1)/2"]


def run(*args, stdin=None):
    r = subprocess.run([REF, *args], input=stdin, capture_output=True, text=True, check=True)
    return json.loads(r.stdout)


def random_qasm(rng, version, n):
    """A seeded program: gates with expression parameters, register broadcast,
    (version 3) slices and aliases, comments, barriers."""
    lines = ["// generated", f"OPENQASM {'3.0' if version == 3 else '2.0'};",
             'include "stdgates.inc";' if version == 3 else 'include "qelib1.inc";']
    lines.append(f"qubit[{n}] q;" if version == 3 else f"qreg q[{n}];")
    lines.append(f"bit[{n}] c;" if version == 3 else f"creg c[{n}];")
    for _ in range(int(rng.integers(20, 40))):
        r = rng.uniform()
        if r < 0.08:
            lines.append("barrier q;")
            continue
        if r < 0.12:
            g = ONE[int(rng.integers(len(ONE)))]
            par = "(" + ",".join(EXPR[int(rng.integers(len(EXPR)))] for _ in range(NPAR.get(g, 0))) + ")" \
                if NPAR.get(g, 0) else ""
            lines.append(f"{g}{par} q;")
            continue
        if version == 3 and r < 0.16:
            lo = int(rng.integers(0, n - 1))
            lines.append(f"h q[{lo}:{n}];")
            continue
        pool = ONE if r < 0.6 else TWO if r < 0.9 else THREE
        g = pool[int(rng.integers(len(pool)))]
        k = 1 if pool is ONE else 2 if pool is TWO else 3
        qs = rng.choice(n, size=k, replace=False)
        par = "(" + ",".join(EXPR[int(rng.integers(len(EXPR)))] for _ in range(NPAR.get(g, 0))) + ")" \
            if NPAR.get(g, 0) else ""
        name = g.upper() if rng.uniform() < 0.1 else g
        lines.append(f"{name}{par} " + ", ".join(f"q[{int(x)}]" for x in qs) + ";  // gate")
    return "\n".join(lines) + "\n"


def random_pauli(rng, n):
    lines = ["# generated terms", "// and a comment"]
    for _ in range(12):
        lab = "".join("IXYZ"[int(rng.integers(4))] for _ in range(n))
        if rng.uniform() < 0.3:
            lines.append(f"{rng.normal():.6f} {lab.lower()}")
        else:
            lines.append(f"{rng.normal():.6f} {rng.normal():.6f} {lab}")
    return "\n".join(lines) + "\n"


def random_edges(rng, n, m):
    e = set()
    while len(e) < m:
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        e.add((a, b) if rng.uniform() < 0.5 else (b, a))
    return sorted(e, key=lambda x: rng.uniform())


def cases():
    rng = np.random.default_rng(2025)
    c = {"square": [[lx, ly] for lx in (1, 2, 3, 5) for ly in (1, 2, 4)] + [[16, 16], [32, 32], [7, 3]],
         "oned": [[5, 0], [5, 1], [8, 0], [8, 1]],
         "graphs": [random_edges(rng, int(rng.integers(5, 12)), int(rng.integers(6, 20))) for _ in range(6)],
         "qasm": [random_qasm(rng, v, int(rng.integers(4, 8))) for v in (2, 3) for _ in range(4)],
         "pauli": [random_pauli(rng, int(rng.integers(3, 9))) for _ in range(3)]}
    return c


def main():
    c = cases()
    out = {"square": [run("square", str(a), str(b)) for a, b in c["square"]],
           "oned": [run("oned", str(a), str(b)) for a, b in c["oned"]]}
    lat = [out["square"][c["square"].index([16, 16])], out["square"][c["square"].index([7, 3])]]
    graphs = [[list(e) for e in g] for g in c["graphs"]] + lat
    out["layers"] = [run("layers", stdin="".join(f"{u} {v}\n" for u, v in g)) for g in graphs]
    out["paths"] = []
    rng = np.random.default_rng(5)
    for g in graphs:
        sites = sorted({s for e in g for s in e})
        for _ in range(6):
            a, b = (int(x) for x in rng.choice(sites, size=2))
            out["paths"].append([a, b, run("path", str(a), str(b),
                                           stdin="".join(f"{u} {v}\n" for u, v in g))])
    with tempfile.TemporaryDirectory() as d:
        out["qasm"] = []
        for k, text in enumerate(c["qasm"]):
            path = os.path.join(d, f"p{k}.qasm")
            open(path, "w").write(text)
            out["qasm"].append(run("qasm", path))
        out["pauli"] = []
        for k, text in enumerate(c["pauli"]):
            path = os.path.join(d, f"h{k}.txt")
            open(path, "w").write(text)
            out["pauli"].append(run("pauli", path))
    with open(OUT, "w") as f:
        json.dump(out, f)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
