"""Drivers: the kicked-Ising Floquet loop and the layered-circuit estimator loop.

``run_kicked`` follows apps/kicked_ising_lattice/main.cc:21-238 (options of
option.h:27-69, ``dt`` in units of pi); ``run_layers`` follows the estimator's
per-slice loop (apps/estimator/main.cc:119-178).  Both run on the fused PEPS
layer; every step is AbsorbTPO, the BP loop to tolerance, Truncation and the
requested measurements.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import circuits
from .peps import (HBP_MODE_FUSED, HBP_PREC_FP64, AbsorbTPO, InitTensorProductState, Measure, MeasureEdges,
                   Truncation, bp_loop)

Z = circuits.PAULI["Z"]
X = circuits.PAULI["X"]
ZZ = np.kron(Z, Z)


def op2(u):
    """Two-site operator tensor (a_in, b_in, a_out, b_out) from a row-major 4x4."""
    return circuits.op_tensor(u)


def run_kicked(ctx, L, num_steps, Jz, hz, hx, dt, max_bp_iterations, bp_tolerance,
               max_bond_dim, sv_min, lattice="square", two_site_keep_zero=False,
               precision=HBP_PREC_FP64, measure="every", observables=("x", "z", "zz"),
               log=None, mode=HBP_MODE_FUSED, schedule=0, accel=0):
    """Kicked-Ising Floquet evolution; returns per-step records like the oracle's.
    ``mode`` HBP_MODE_PLAIN runs the reference's op sequence one TCI-level
    call at a time (bp.h:128-175, :290-350) instead of the fused kernels;
    ``schedule`` 1 the checkerboard BP schedule, ``accel`` > 0 Anderson
    mixing of that depth (defaults: the reference's Jacobi loop)."""
    edges = circuits.square_lattice(L, L) if lattice == "square" else circuits.chain(L)
    sites = sorted({s for e in edges for s in e})
    tpo = circuits.kicked_ising_tpo(edges, Jz, hz, hx, dt * math.pi, two_site_keep_zero)
    peps = InitTensorProductState(ctx, edges, [2] * len(sites), precision)
    peps.set_mode(mode)
    peps.set_schedule(schedule)
    peps.set_accel(accel)
    steps = []
    try:
        for step in range(num_steps):
            t0 = time.perf_counter()
            AbsorbTPO(ctx, tpo, peps)
            it, res = bp_loop(ctx, peps, max_bp_iterations, bp_tolerance)
            dims, errs = Truncation(ctx, peps, max_bond_dim, sv_min)
            rec = {"step": step, "bp_iterations": it, "residuals": res, "bond_dims": dims,
                   "trunc_errs": errs}
            if measure == "every" or step == num_steps - 1:
                if "x" in observables:
                    rec["x"] = [v.real for v in Measure(ctx, peps, sites, [circuits.op_tensor(X)] * len(sites))]
                if "z" in observables:
                    rec["z"] = [v.real for v in Measure(ctx, peps, sites, [circuits.op_tensor(Z)] * len(sites))]
                if "zz" in observables:
                    rec["zz"] = [v.real for v in MeasureEdges(ctx, peps, edges, [op2(ZZ)] * len(edges))]
            ctx.synchronize()
            rec["seconds"] = time.perf_counter() - t0
            steps.append(rec)
            if log is not None:
                log(rec)
    finally:
        peps.close()
    return {"edges": edges, "steps": steps}


def run_layers(ctx, edges, layers, max_bp_iterations, bp_tolerance, max_bond_dim, sv_min,
               precision=HBP_PREC_FP64, keep=False):
    """Estimator loop: one TPO slice per layer of (U, qubits) gates."""
    sites = sorted({s for e in edges for s in e})
    peps = InitTensorProductState(ctx, edges, [2] * len(sites), precision)
    recs = []
    for gates in layers:
        AbsorbTPO(ctx, circuits.layer_tpo(edges, gates), peps)
        it, res = bp_loop(ctx, peps, max_bp_iterations, bp_tolerance)
        dims, errs = Truncation(ctx, peps, max_bond_dim, sv_min)
        recs.append({"bp_iterations": it, "residuals": res, "bond_dims": dims, "trunc_errs": errs})
    if keep:
        return peps, recs
    peps.close()
    return None, recs
