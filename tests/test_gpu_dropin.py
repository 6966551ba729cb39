"""The reference's own kicked-Ising driver, compiled unchanged against this
backend's TCI header (tnbp_amd/dropin), runs config 1 on the GPU and prints
the oracle's values (tolerance 1e-10 relative, floor 1).

The binary is built by `make -C tnbp_amd/dropin` wherever the reference
sources are present; the test is skipped when it has not been built.
"""
import json
import math
import os
import re
import subprocess

import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tnbp_amd", "dropin", "_bin", "kicked_ising_exvalue")


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in driver not built")
def test_reference_kicked_driver_config1():
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "config1_kicked_4x4_chi4.json")))
    a = g["args"]
    cmd = [BIN, "--lattice", "square", "--L", str(a["L"]), "--num_steps", str(a["num_steps"]),
           "--Jz", repr(a["Jz"]), "--hz", repr(a["hz"]), "--hx", repr(a["hx"]), "--dt", repr(a["dt"]),
           "--max_bp_iterations", str(a["max_bp_iterations"]), "--bp_tolerance", repr(a["bp_tolerance"]),
           "--max_bond_dim", str(a["max_bond_dim"]), "--sv_min", repr(a["sv_min"])]
    announce("drop-in " + " ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    vals = {}
    for m in re.finditer(r"measurement of (X|Z) at site (\d+) time step (\d+) = (\S+)", r.stdout):
        vals[(m.group(1), int(m.group(2)), int(m.group(3)))] = float(m.group(4))
    zz = {}
    for m in re.finditer(r"measurement of ZZ at sites \((\d+),(\d+)\) time step (\d+) = (\S+)", r.stdout):
        zz[(int(m.group(1)), int(m.group(2)), int(m.group(3)))] = float(m.group(4))
    iters = [len(re.findall(rf"error from belief propagation condition at time step {s} =", r.stdout))
             for s in range(a["num_steps"])]
    assert iters == [s["bp_iterations"] for s in g["steps"]]
    worst = 0.0
    for s in g["steps"]:
        for i, want in enumerate(s["z"]):
            got = vals[("Z", i, s["step"])]
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
        for i, want in enumerate(s["x"]):
            got = vals[("X", i, s["step"])]
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
        for (u, v), want in zip(g["edges"], s["zz"]):
            got = zz[(u, v, s["step"])]
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
    announce(f"drop-in worst relative error {worst:.3e}; <Z_5> = {vals[('Z', 5, a['num_steps'] - 1)]!r}")
    assert worst <= 1e-10
