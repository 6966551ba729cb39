"""Kernel launch variants compute bitwise the same result.

The autotuner picks, per shape class, the fastest of several launch variants
of the mode-product and message kernels (rows per wave, loads in flight,
occupancy).  Each variant accumulates every output element over k in the same
order, so the choice must not change a single bit: here each variant is forced
for the whole process (HBP_FORCE_VARIANT) on 4x4 kicked-Ising PEPS whose
absorbed bonds reach 40 (partial column blocks) and 64 (whole 64 -> 64 tiles,
where the FP32 mode product has its bounds-free variant), and the messages
and expectation values of every run are compared bitwise with variant 0.
"""
import hashlib
import json
import os
import subprocess
import sys

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import hashlib, json, math, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from tnbp_amd import circuits as pc
from tnbp_amd.peps import AbsorbTPO, InitTensorProductState, Measure, Truncation, bp_loop
from tnbp_amd.tci import Context
out = {}
with Context(0) as ctx:
    for prec in (0, 1):
        for chi in (20, 32):
            edges = pc.square_lattice(4, 4)
            tpo = pc.kicked_ising_tpo(edges, 0.125, 0.0, 0.11, math.pi)
            p = InitTensorProductState(ctx, edges, [2] * 16, prec)
            for step in range(7):
                AbsorbTPO(ctx, tpo, p)
                bp_loop(ctx, p, 4, 0.0)
                if step < 6:
                    Truncation(ctx, p, chi, 1e-12)
            h = hashlib.sha256()
            for m in range(2 * len(edges)):
                h.update(np.ascontiguousarray(p.get_message(m)).tobytes())
            z = Measure(ctx, p, list(range(16)), [pc.op_tensor(pc.PAULI["Z"])] * 16)
            h.update(np.asarray(z).tobytes())
            out[f"{prec}/{chi}"] = (h.hexdigest(), max(max(p.site_shape(s)[:-1]) for s in range(16)))
            p.close()
print(json.dumps(out))
"""


def _run(variant, **extra):
    env = dict(os.environ)
    env["HBP_FORCE_VARIANT"] = str(variant)
    env.update(extra)
    r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], capture_output=True, text=True,
                       env=env, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_forced_variants_bitwise_equal():
    announce("launch variants 0..6 forced, bitwise comparison")
    ref = _run(0)
    # the absorbed bonds reached 2 chi in both lattices
    assert [ref[k][1] for k in ("0/20", "0/32", "1/20", "1/32")] == [40, 64, 40, 64]
    for v in range(1, 7):
        got = _run(v)
        assert got == ref, (v, got, ref)


def test_batched_reductions_bitwise_equal():
    """A sweep defers its split-K reductions and runs up to eight per launch;
    each keeps its fixed summation order, so the messages equal those of one
    reduction per message bit for bit (same lattices: bonds 1 .. 64, degrees
    2 .. 4, diagonal and dense environments, merged and separate messages)."""
    announce("batched vs separate split-K reductions, bitwise comparison")
    assert _run(0, HBP_REDUCE_BATCH="0") == _run(0)
