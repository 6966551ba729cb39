"""Generate the golden fixtures of the BP-PEPS path from the oracle.

The reference's arithmetic (gqten) cannot be built or imported here, so the
fixtures are the oracle's outputs (oracle/ is pinned by tests/test_oracle_*).
Run from the repo root:  python tests/golden/make_golden.py
"""
import json
import math
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import circuits  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def kicked(name, **kw):
    r = circuits.run_kicked(**kw)
    steps = [{k: v for k, v in s.items()} for s in r["steps"]]
    with open(os.path.join(OUT, name), "w") as f:
        json.dump({"args": kw, "edges": r["edges"], "steps": steps}, f, indent=0)
    print("wrote", name)


if __name__ == "__main__":
    c1 = dict(L=4, num_steps=5, Jz=0.125, hz=0.0, hx=0.3 / (2 * math.pi), dt=1.0,
              max_bp_iterations=10000, bp_tolerance=1e-10, max_bond_dim=4, sv_min=1e-8,
              two_site_svd="trunc")
    kicked("config1_kicked_4x4_chi4.json", **c1)
    # fixed iteration count (tolerance 0) so that parity does not hinge on the stop rule
    c1f = dict(c1, max_bp_iterations=3, bp_tolerance=0.0)
    kicked("config1_fixed3_kicked_4x4_chi4.json", **c1f)
    # the reference driver as written: t = 4 two-site split (circuit.h:62)
    c1t4 = dict(c1, two_site_svd="full", num_steps=3)
    kicked("config1_t4_kicked_4x4_chi4.json", **c1t4)
