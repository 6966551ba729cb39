"""Host mirror of tnbp's PEPS algorithm functions over the fused hbp PEPS layer.

Names follow include/tnbp/ptns/function/*.h: ``InitTensorProductState``,
``AbsorbTPO``, ``BeliefPropagation``, ``BeliefPropagationCondition``,
``Truncation`` and ``Measure``.  The state (site tensors and messages) stays
resident on the device between calls; only operators and results cross the
host boundary.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import HBP_MAX_ORDER, HBP_PREC_FP64, HBP_PREC_MIXED, check, i32
from .tci import Context

__all__ = ["Peps", "InitTensorProductState", "AbsorbTPO", "BeliefPropagation",
           "BeliefPropagationCondition", "bp_loop", "Truncation", "Measure", "MeasureEdges",
           "HBP_PREC_FP64", "HBP_PREC_MIXED"]


class Peps:
    """Device-resident tensor-product state with its BP messages.

    ``edges`` is the global edge list (``I`` in the reference); site tensors
    are indexed by position in the sorted site list; message ``m`` runs
    first->second and ``m + n_e`` second->first (bp.h:170-174).
    """

    def __init__(self, ctx: Context, edges, pdim=None, precision=HBP_PREC_FP64):
        self.ctx = ctx
        self.edges = [tuple(map(int, e)) for e in edges]
        self.sites = sorted({s for e in self.edges for s in e})
        flat = [x for e in self.edges for x in e]
        pd = list(pdim) if pdim is not None else [2] * len(self.sites)
        h = C.c_void_p()
        check(ctx.lib.hbp_peps_create(ctx.h, len(self.edges), i32(flat), i32(pd), precision,
                                      C.byref(h)))
        self.h = h
        self.precision = precision

    def close(self):
        # a closed context has already released the device pool holding the state
        if self.h and self.ctx.h:
            check(self.ctx.lib.hbp_peps_destroy(self.ctx.h, self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_edges(self):
        return len(self.edges)

    def site_shape(self, s):
        order = C.c_int()
        sh = (C.c_int32 * HBP_MAX_ORDER)()
        check(self.ctx.lib.hbp_peps_site_shape(self.h, s, C.byref(order), sh))
        return tuple(sh[: order.value])

    def get_site(self, s):
        shape = self.site_shape(s)
        out = np.empty(int(np.prod(shape)), dtype=np.complex128)
        check(self.ctx.lib.hbp_peps_get_site(self.ctx.h, self.h, s,
                                             out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(shape, order="F")

    def set_site(self, s, a):
        a = np.asarray(a, dtype=np.complex128)
        buf = np.ascontiguousarray(a.reshape(-1, order="F"))
        check(self.ctx.lib.hbp_peps_set_site(self.ctx.h, self.h, s, a.ndim, i32(a.shape),
                                             buf.ctypes.data_as(C.POINTER(C.c_double))))

    def get_message(self, m):
        d = C.c_int32()
        check(self.ctx.lib.hbp_peps_get_message(self.ctx.h, self.h, m, C.byref(d), None))
        out = np.empty(d.value * d.value, dtype=np.complex128)
        check(self.ctx.lib.hbp_peps_get_message(self.ctx.h, self.h, m, C.byref(d),
                                                out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape((d.value, d.value), order="F")

    def set_message(self, m, a):
        a = np.asarray(a, dtype=np.complex128)
        buf = np.ascontiguousarray(a.reshape(-1, order="F"))
        check(self.ctx.lib.hbp_peps_set_message(self.ctx.h, self.h, m, a.shape[0],
                                                buf.ctypes.data_as(C.POINTER(C.c_double))))


def InitTensorProductState(ctx, edges, pdim=None, precision=HBP_PREC_FP64) -> Peps:
    """``InitTensorProductState`` (include/tnbp/ptns/function/init.h:17-95)."""
    return Peps(ctx, edges, pdim, precision)


def AbsorbTPO(ctx, O, peps: Peps) -> None:
    """``AbsorbTPO`` (apply.h:14-177); ``O[s]`` is site s's operator
    (t_1..t_z, d_in, d_out) as a numpy complex array."""
    arrs = [np.ascontiguousarray(np.asarray(o, dtype=np.complex128).reshape(-1, order="F"))
            for o in O]
    ptrs = (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    shapes = []
    for o in O:
        sh = list(np.shape(o)) + [0] * (HBP_MAX_ORDER - np.ndim(o))
        shapes += sh
    check(ctx.lib.hbp_peps_absorb_tpo(ctx.h, peps.h, ptrs, i32(shapes)))


def BeliefPropagation(ctx, peps: Peps) -> None:
    """``BeliefPropagation`` over every layer followed by E <- F (bp.h:65-261,
    apps/kicked_ising_lattice/main.cc:128-141)."""
    check(ctx.lib.hbp_peps_bp_sweep(ctx.h, peps.h))


def BeliefPropagationCondition(ctx, peps: Peps) -> float:
    """``BeliefPropagationCondition`` (bp.h:266-355)."""
    r = C.c_double()
    check(ctx.lib.hbp_peps_bp_condition(ctx.h, peps.h, C.byref(r)))
    return r.value


def bp_loop(ctx, peps: Peps, max_iter: int, tol: float, progress=None):
    """The drivers' BP loop (main.cc:121-159) -> (iterations, residuals).

    With ``progress`` the loop runs one iteration per call into the library
    and calls ``progress(iteration, residual)`` after each."""
    if progress is None:
        it = C.c_int()
        res = np.zeros(max_iter, dtype=np.float64)
        check(ctx.lib.hbp_peps_bp_run(ctx.h, peps.h, max_iter, tol, C.byref(it),
                                      res.ctypes.data_as(C.POINTER(C.c_double))))
        return it.value, list(res[: it.value])
    residuals = []
    for k in range(max_iter):
        it, conv, r = BPIterate(ctx, peps, 1, tol, first=(k == 0), last=(k == max_iter - 1))
        residuals += r
        progress(len(residuals), r[-1])
        if conv:
            break
    return len(residuals), residuals


def BPIterate(ctx, peps: Peps, n: int, tol: float, first: bool, last: bool):
    """Up to n iterations of the BP loop -> (iterations, converged, residuals)."""
    it, conv = C.c_int(), C.c_int()
    res = np.zeros(n, dtype=np.float64)
    check(ctx.lib.hbp_peps_bp_iterate(ctx.h, peps.h, n, tol, int(first), int(last), C.byref(it),
                                      res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(conv)))
    return it.value, bool(conv.value), list(res[: it.value])


def Truncation(ctx, peps: Peps, max_dim: int, eps: float):
    """``Truncation`` (truncation.h:17-222) -> (bond dims, truncation errors) per edge."""
    dims = np.zeros(peps.n_edges, dtype=np.int32)
    errs = np.zeros(peps.n_edges, dtype=np.float64)
    check(ctx.lib.hbp_peps_truncate(ctx.h, peps.h, max_dim, eps,
                                    dims.ctypes.data_as(C.POINTER(C.c_int32)),
                                    errs.ctypes.data_as(C.POINTER(C.c_double))))
    return list(map(int, dims)), list(map(float, errs))


def Measure(ctx, peps: Peps, sites, ops):
    """Single-site ``Measure`` (measure.h:14-94) -> list of complex values."""
    n = len(sites)
    if n == 0:
        return []
    buf = np.ascontiguousarray(np.stack([np.asarray(o, dtype=np.complex128).reshape(-1, order="F")
                                         for o in ops]))
    out = np.zeros(n, dtype=np.complex128)
    check(ctx.lib.hbp_peps_measure_sites(ctx.h, peps.h, n, i32(sites),
                                         buf.ctypes.data_as(C.POINTER(C.c_double)),
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
    return list(out)


def MeasureEdges(ctx, peps: Peps, pairs, ops):
    """Nearest-neighbour ``Measure`` (measure.h:99-300) -> list of complex values."""
    n = len(pairs)
    if n == 0:
        return []
    flat = [x for p in pairs for x in p]
    buf = np.ascontiguousarray(np.stack([np.asarray(o, dtype=np.complex128).reshape(-1, order="F")
                                         for o in ops]))
    out = np.zeros(n, dtype=np.complex128)
    check(ctx.lib.hbp_peps_measure_edges(ctx.h, peps.h, n, i32(flat),
                                         buf.ctypes.data_as(C.POINTER(C.c_double)),
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
    return list(out)


def MeasurePaths(ctx, peps: Peps, paths, ops):
    """<O_1 ... O_k> for product operators on simple paths of neighbouring
    sites (extends measure.h:14-300 beyond two sites) -> list of complex values.

    paths[t] is a list of site labels (consecutive ones adjacent); ops[t] the
    list of (d, d) operators (in, out) on those sites, in the same order."""
    n = len(paths)
    if n == 0:
        return []
    offsets = [0]
    flat_sites, flat_ops = [], []
    for path, o in zip(paths, ops):
        if len(path) != len(o):
            raise ValueError("MeasurePaths: one operator per path site")
        flat_sites += list(path)
        flat_ops += [np.asarray(x, dtype=np.complex128).reshape(-1, order="F") for x in o]
        offsets.append(len(flat_sites))
    buf = np.ascontiguousarray(np.stack(flat_ops))
    out = np.zeros(n, dtype=np.complex128)
    check(ctx.lib.hbp_peps_measure_paths(ctx.h, peps.h, n, i32(offsets), i32(flat_sites),
                                         buf.ctypes.data_as(C.POINTER(C.c_double)),
                                         out.ctypes.data_as(C.POINTER(C.c_double))))
    return list(out)
