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
from ._lib import (HBP_ABSORB_AUTO, HBP_ABSORB_EAGER, HBP_ABSORB_LAZY, HBP_BATCH_ALONE,
                   HBP_BATCH_GRAPH, HBP_BATCH_SMALL, HBP_MAX_ORDER,
                   HBP_MODE_FUSED, HBP_MODE_PLAIN, HBP_SCHEDULE_CHECKERBOARD,
                   HBP_SCHEDULE_JACOBI,
                   HBP_PREC_FP64, HBP_PREC_MIXED, HBP_SWEEP_AUTO, HBP_SWEEP_GROUPED,
                   HBP_SWEEP_MERGED, HBP_SWEEP_PER_SITE, check, i32)
from .tci import Context

__all__ = ["Peps", "InitTensorProductState", "AbsorbTPO", "BeliefPropagation",
           "BeliefPropagationCondition", "bp_loop", "bp_loop_batched", "Truncation", "AbsorbTPOBatched",
           "TruncationBatched", "Measure", "MeasureEdges",
           "HBP_PREC_FP64", "HBP_PREC_MIXED", "HBP_MODE_FUSED", "HBP_MODE_PLAIN",
           "HBP_SCHEDULE_JACOBI", "HBP_SCHEDULE_CHECKERBOARD", "HBP_BATCH_ALONE", "HBP_BATCH_SMALL",
           "HBP_BATCH_GRAPH"]


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

    def set_mode(self, mode):
        """HBP_MODE_FUSED (0) or HBP_MODE_PLAIN (1, the reference's op sequence)."""
        check(self.ctx.lib.hbp_peps_set_mode(self.h, int(mode)))

    def set_absorb_mode(self, mode):
        """HBP_ABSORB_AUTO (0), HBP_ABSORB_LAZY (1) or HBP_ABSORB_EAGER (2)."""
        check(self.ctx.lib.hbp_peps_set_absorb_mode(self.h, int(mode)))

    def set_sweep_kernel(self, kind):
        """HBP_SWEEP_AUTO (0), HBP_SWEEP_PER_SITE (1) or HBP_SWEEP_MERGED (3);
        HBP_SWEEP_GROUPED (2) is reserved and raises NotImplementedError."""
        check(self.ctx.lib.hbp_peps_set_sweep_kernel(self.h, int(kind)))

    def set_schedule(self, schedule):
        """HBP_SCHEDULE_JACOBI (0, the reference's) or HBP_SCHEDULE_CHECKERBOARD (1)."""
        check(self.ctx.lib.hbp_peps_set_schedule(self.h, int(schedule)))

    def set_pair_cache(self, policy):
        """-1 by free memory (default), 0 never, 1 always (memory budgeted by the caller)."""
        check(self.ctx.lib.hbp_peps_set_pair_cache(self.h, int(policy)))

    def set_accel(self, depth, start=1e-2):
        """Anderson mixing of the BP loop (depth 0: off, the default): once the
        residual is below ``start`` the next message set combines the last
        ``depth + 1`` sweep outputs; same fixed point and stopping test."""
        check(self.ctx.lib.hbp_peps_set_accel(self.h, int(depth), float(start)))

    def set_residual_floor(self, floor):
        """Mixed precision: BP also stops once the residual is below this."""
        check(self.ctx.lib.hbp_peps_set_residual_floor(self.h, float(floor)))

    def set_phase_gauge(self, on):
        check(self.ctx.lib.hbp_peps_set_phase_gauge(self.h, int(bool(on))))

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


# ------------------------------------------------------------ checkpoints
CHECKPOINT_MAGIC = b"HBPP"
CHECKPOINT_VERSION = 1


def save_peps(peps: Peps, path):
    """Write the state (edges, site tensors, messages) to a versioned binary
    file: "HBPP", u32 version, u32 precision, u32 n_sites, u32 n_edges,
    int32 edges (2 n_edges), then per site u32 order, int32 shape, complex128
    column-major data, then per message int32 D, complex128 D x D.  The
    reference's restart stores V and E with gqten's stream format
    (iotps.h:37-105), which is not available to match."""
    n = len(peps.sites)
    with open(path, "wb") as f:
        f.write(CHECKPOINT_MAGIC)
        np.array([CHECKPOINT_VERSION, peps.precision, n, len(peps.edges)], dtype="<u4").tofile(f)
        np.array([x for e in peps.edges for x in e], dtype="<i4").tofile(f)
        for s in range(n):
            v = peps.get_site(s)
            np.array([v.ndim], dtype="<u4").tofile(f)
            np.array(v.shape, dtype="<i4").tofile(f)
            np.ascontiguousarray(v.reshape(-1, order="F")).astype("<c16").tofile(f)
        for m in range(2 * len(peps.edges)):
            e = peps.get_message(m)
            np.array([e.shape[0]], dtype="<i4").tofile(f)
            np.ascontiguousarray(e.reshape(-1, order="F")).astype("<c16").tofile(f)


def load_peps(ctx, path) -> Peps:
    """Read a state written by ``save_peps`` into a new PEPS on ctx."""
    with open(path, "rb") as f:
        if f.read(4) != CHECKPOINT_MAGIC:
            raise ValueError(f"{path}: not an hbp PEPS checkpoint")
        ver, prec, n, ne = (int(x) for x in np.fromfile(f, dtype="<u4", count=4))
        if ver != CHECKPOINT_VERSION:
            raise ValueError(f"{path}: checkpoint version {ver} (expected {CHECKPOINT_VERSION})")
        flat = np.fromfile(f, dtype="<i4", count=2 * ne)
        edges = [(int(flat[2 * i]), int(flat[2 * i + 1])) for i in range(ne)]
        sites = []
        for _ in range(n):
            order = int(np.fromfile(f, dtype="<u4", count=1)[0])
            shape = tuple(int(x) for x in np.fromfile(f, dtype="<i4", count=order))
            data = np.fromfile(f, dtype="<c16", count=int(np.prod(shape)))
            sites.append(data.reshape(shape, order="F"))
        msgs = []
        for _ in range(2 * ne):
            d = int(np.fromfile(f, dtype="<i4", count=1)[0])
            msgs.append(np.fromfile(f, dtype="<c16", count=d * d).reshape((d, d), order="F"))
    peps = Peps(ctx, edges, [v.shape[-1] for v in sites], prec)
    dims = [msgs[m].shape[0] for m in range(ne)]
    check(ctx.lib.hbp_peps_set_bond_dims(ctx.h, peps.h, i32(dims), i32([v.shape[-1] for v in sites])))
    for s, v in enumerate(sites):
        peps.set_site(s, v)
    for m, e in enumerate(msgs):
        peps.set_message(m, e)
    return peps


def InitTensorProductState(ctx, edges, pdim=None, precision=HBP_PREC_FP64) -> Peps:
    """``InitTensorProductState`` (include/tnbp/ptns/function/init.h:17-95)."""
    return Peps(ctx, edges, pdim, precision)


class PackedTPO:
    """A TPO layer in the C-ABI's form, built once: every site's operator
    (column-major complex128) in one contiguous buffer, the array of pointers
    to them and the flat shape array (n x HBP_MAX_ORDER).  ``AbsorbTPO`` and
    ``AbsorbTPOBatched`` take it in place of the operator list, so a layer
    absorbed again (or by a server, packed when the job arrives) costs no
    per-site host work."""

    __slots__ = ("buf", "ptrs", "shapes", "n")

    def __init__(self, O):
        flats = [np.asarray(o, dtype=np.complex128).ravel(order="F") for o in O]
        n = len(flats)
        self.n = n
        sizes = np.fromiter((f.size for f in flats), dtype=np.int64, count=n)
        self.buf = np.concatenate(flats) if n else np.zeros(1, dtype=np.complex128)
        offs = np.zeros(n, dtype=np.int64)
        if n > 1:
            np.cumsum(sizes[:-1], out=offs[1:])
        addr = (self.buf.ctypes.data + 16 * offs).astype(np.uint64)
        self.ptrs = (C.POINTER(C.c_double) * max(n, 1))()
        C.memmove(self.ptrs, addr.ctypes.data, n * C.sizeof(C.c_void_p))
        shp = np.zeros((max(n, 1), HBP_MAX_ORDER), dtype=np.int32)
        for i, o in enumerate(O):
            nd = np.ndim(o)
            if nd > HBP_MAX_ORDER:
                raise ValueError("TPO operator of too high an order")
            shp[i, :nd] = np.shape(o)
        self.shapes = (C.c_int32 * shp.size).from_buffer_copy(shp.tobytes())


def _tpo_args(O):
    """(keep-alive object, pointer array, flat shapes) of a TPO for the C-ABI;
    ``O`` an operator list or a ``PackedTPO``."""
    pk = O if isinstance(O, PackedTPO) else PackedTPO(O)
    return pk, pk.ptrs, pk.shapes


def AbsorbTPO(ctx, O, peps: Peps) -> None:
    """``AbsorbTPO`` (apply.h:14-177); ``O[s]`` is site s's operator
    (t_1..t_z, d_in, d_out) as a numpy complex array (or ``O`` a
    ``PackedTPO`` of such a list)."""
    _, ptrs, shapes = _tpo_args(O)
    check(ctx.lib.hbp_peps_absorb_tpo(ctx.h, peps.h, ptrs, shapes))


def AbsorbTPOBatched(ctx, Os, peps_list) -> None:
    """``AbsorbTPO`` of several PEPS of one context at once
    (``hbp_peps_absorb_tpo_batched``); ``Os[i]`` is PEPS i's operator list.
    Each PEPS ends bitwise as ``AbsorbTPO`` leaves it alone."""
    n = len(peps_list)
    if n == 0:
        return
    args = [_tpo_args(O) for O in Os]
    hs = (C.c_void_p * n)(*[p.h.value if isinstance(p.h, C.c_void_p) else p.h for p in peps_list])
    ops = (C.POINTER(C.POINTER(C.c_double)) * n)(*[C.cast(a[1], C.POINTER(C.POINTER(C.c_double)))
                                                   for a in args])
    shp = (C.POINTER(C.c_int32) * n)(*[C.cast(a[2], C.POINTER(C.c_int32)) for a in args])
    check(ctx.lib.hbp_peps_absorb_tpo_batched(ctx.h, n, hs, ops, shp))


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


def bp_loop_batched(ctx, peps_list, max_iter: int, tol: float):
    """The BP loop of several PEPS of one context at once
    (``hbp_peps_bp_run_batched``): PEPS in fused mode with the Jacobi schedule
    and no mixing iterate in lockstep, one replay of a HIP graph per iteration
    for all of them (those with bonds <= 16 in one persistent launch, the
    others with their own sweeps); each gets what ``bp_loop`` gives it alone.
    Returns a list of (iterations, residuals, converged, batched) per PEPS,
    batched = HBP_BATCH_SMALL, HBP_BATCH_GRAPH or 0 (run on its own)."""
    n = len(peps_list)
    if n == 0:
        return []
    hs = (C.c_void_p * n)(*[p.h.value if isinstance(p.h, C.c_void_p) else p.h for p in peps_list])
    it = (C.c_int * n)()
    conv = (C.c_int * n)()
    bat = (C.c_int * n)()
    res = np.zeros((n, max_iter), dtype=np.float64)
    check(ctx.lib.hbp_peps_bp_run_batched(ctx.h, n, hs, max_iter, tol, it, res.ctypes.data_as(
        C.POINTER(C.c_double)), conv, bat))
    return [(it[i], list(res[i, : it[i]]), bool(conv[i]), int(bat[i])) for i in range(n)]


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


def TruncationBatched(ctx, peps_list, max_dims, epss):
    """``Truncation`` of several PEPS of one context at once
    (``hbp_peps_truncate_batched``: one launch per stage for all of them) ->
    list of (bond dims, truncation errors) per PEPS, each bitwise what
    ``Truncation`` gives that PEPS alone."""
    n = len(peps_list)
    if n == 0:
        return []
    hs = (C.c_void_p * n)(*[p.h.value if isinstance(p.h, C.c_void_p) else p.h for p in peps_list])
    ne = [p.n_edges for p in peps_list]
    dims = np.zeros(max(sum(ne), 1), dtype=np.int32)
    errs = np.zeros(max(sum(ne), 1), dtype=np.float64)
    check(ctx.lib.hbp_peps_truncate_batched(
        ctx.h, n, hs, i32([int(x) for x in max_dims]),
        np.ascontiguousarray(epss, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double)),
        dims.ctypes.data_as(C.POINTER(C.c_int32)), errs.ctypes.data_as(C.POINTER(C.c_double))))
    out, o = [], 0
    for k in ne:
        out.append((list(map(int, dims[o:o + k])), list(map(float, errs[o:o + k]))))
        o += k
    return out


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
