"""Short fixed-shape run of the BP sweep kernels, for hardware-counter runs.

Builds an L x L square-lattice PEPS with every bond at dimension D and
physical dimension 2, seeded random site tensors and random Hermitian
positive messages, and runs a few fused BP sweeps on it: at D = 64 every
interior site takes the shared half-products (`k_mode_product`) and the
merged product-message kernel (`k_product_message`) at the shapes of the
16x16 chi=32 steady state (bonds chi * t = 64), without the ramp of a full
circuit.  Meant for `rocprofv3 --pmc ... -- python3 -m tnbp_amd.kernel_probe`
and `rocprofv3 --kernel-trace --stats`, whose runs must stay short.

    python -m tnbp_amd.kernel_probe [--L 4] [--D 64] [--precision mixed] [--sweeps 3]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from ._lib import check, i32
from .circuits import square_lattice
from .peps import (HBP_PREC_FP64, HBP_PREC_MIXED, BeliefPropagation, Peps)
from .tci import Context


def build_state(ctx, L, D, precision, seed=1):
    edges = square_lattice(L, L)
    n = L * L
    peps = Peps(ctx, edges, [2] * n, precision)
    check(ctx.lib.hbp_peps_set_bond_dims(ctx.h, peps.h, i32([D] * len(edges)), i32([2] * n)))
    rng = np.random.default_rng(seed)
    for s in range(n):
        shape = peps.site_shape(s)
        v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        peps.set_site(s, v / np.linalg.norm(v))
    for m in range(2 * len(edges)):
        a = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))
        e = a @ a.conj().T + D * np.eye(D)
        peps.set_message(m, e / np.linalg.norm(e))
    return peps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=4)
    ap.add_argument("--D", type=int, default=64)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--sweeps", type=int, default=3)
    a = ap.parse_args()
    prec = HBP_PREC_MIXED if a.precision == "mixed" else HBP_PREC_FP64
    ctx = Context(0)
    peps = build_state(ctx, a.L, a.D, prec)
    print(json.dumps({"probe": "state ready", "L": a.L, "D": a.D, "precision": a.precision}),
          flush=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(a.sweeps):
        BeliefPropagation(ctx, peps)
        ctx.synchronize()
        print(json.dumps({"sweep": i, "seconds": time.perf_counter() - t0}), flush=True)
    peps.close()
    ctx.close()


if __name__ == "__main__":
    main()
