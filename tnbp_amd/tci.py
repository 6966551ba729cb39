"""Host-side mirror of the Tensor Computing Interface over the hbp C-ABI.

Function names, argument meaning and errors follow min-tci
(external/min-tci/include/tci/gqten/*.h): ``contract(ctx, a, la, b, lb, lc)``
sums shared labels and orders survivors by ``lc``; ``normalize`` returns the
norm; ``svd``/``trunc_svd`` split at ``nrows`` row bonds.  Tensors are
device-resident ``Tensor`` handles with value semantics on ``copy``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import HBP_C64, HBP_C128, HBP_F32, HBP_F64, HBP_MAX_ORDER, check, i32


class Context:
    """``tci::create_context``: one device, one in-order HIP stream."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.hbp_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(self.lib.hbp_ctx_destroy(self.h))
            self.h = None

    def synchronize(self):
        check(self.lib.hbp_ctx_synchronize(self.h))

    def memory(self):
        a, b = C.c_size_t(), C.c_size_t()
        check(self.lib.hbp_ctx_memory(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def trim(self):
        """Give the pool's unused memory back to the device."""
        check(self.lib.hbp_ctx_trim(self.h))

    def graph_stats(self):
        """(recorded, replays, dropped) HIP graphs of batched BP loops so far."""
        r, n, d = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.hbp_ctx_graph_stats(self.h, C.byref(r), C.byref(n), C.byref(d)))
        return r.value, n.value, d.value

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def autotune_report():
    """The kernel autotuner's choices so far (one line per shape class)."""
    lib = _lib.load()
    need = C.c_size_t()
    check(lib.hbp_autotune_report(None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value + 1)
    check(lib.hbp_autotune_report(buf, need.value + 1, C.byref(need)))
    return buf.value.decode()


def device_memory(device: int = 0):
    """(free, total) bytes of a device."""
    a, b = C.c_size_t(), C.c_size_t()
    check(_lib.load().hbp_device_memory(device, C.byref(a), C.byref(b)))
    return a.value, b.value


class Tensor:
    """A device tensor (column-major); ``dtype`` HBP_C128, HBP_F64, HBP_C64 or HBP_F32."""

    def __init__(self, ctx: Context, shape=(), dtype=HBP_C128):
        self.ctx = ctx
        h = C.c_void_p()
        shape = list(shape)
        check(ctx.lib.hbp_tensor_create(ctx.h, dtype, len(shape), i32(shape), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                self.ctx.lib.hbp_tensor_destroy(self.ctx.h, self.h)
        except Exception:
            pass
        self.h = None

    # -- getters (min-tci getter.h)
    def info(self):
        dt, order, sz = C.c_int(), C.c_int(), C.c_size_t()
        sh = (C.c_int32 * HBP_MAX_ORDER)()
        check(self.ctx.lib.hbp_tensor_info(self.h, C.byref(dt), C.byref(order), sh, C.byref(sz)))
        return dt.value, order.value, tuple(sh[: order.value]), sz.value

    @property
    def shape(self):
        return self.info()[2]

    @property
    def order(self):
        return self.info()[1]

    @property
    def dtype(self):
        return self.info()[0]

    def to_numpy(self):
        dt, order, shape, size = self.info()
        npdt = {HBP_C128: np.complex128, HBP_F64: np.float64, HBP_C64: np.complex64,
                HBP_F32: np.float32}[dt]
        out = np.empty(size, dtype=npdt)
        check(self.ctx.lib.hbp_tensor_download(self.ctx.h, self.h, out.ctypes.data_as(C.c_void_p)))
        return out.reshape(shape, order="F") if order else out.reshape(())

    @staticmethod
    def from_numpy(ctx: Context, a):
        """complex64 / float32 arrays give single-precision tensors, every
        other array complex128 or float64."""
        a = np.asarray(a)
        real = not np.iscomplexobj(a)
        single = a.dtype in (np.complex64, np.float32)
        dt, npdt = {(True, False): (HBP_F64, np.float64), (False, False): (HBP_C128, np.complex128),
                    (True, True): (HBP_F32, np.float32), (False, True): (HBP_C64, np.complex64)}[
                        (real, single)]
        t = Tensor(ctx, a.shape, dt)
        buf = np.ascontiguousarray(a.reshape(-1, order="F"), dtype=npdt)
        check(ctx.lib.hbp_tensor_upload(ctx.h, t.h, buf.ctypes.data_as(C.c_void_p)))
        return t


# ----------------------------------------------------------------- functions
def copy(ctx, a: Tensor) -> Tensor:
    """``tci::copy`` (constructor.h:192-204)."""
    t = Tensor(ctx, (), a.dtype)
    check(ctx.lib.hbp_tensor_copy(ctx.h, a.h, t.h))
    return t


def reshape(ctx, a: Tensor, shape) -> None:
    """``tci::reshape`` in place (manipulator.h:37-44)."""
    check(ctx.lib.hbp_tensor_reshape(ctx.h, a.h, len(shape), i32(shape)))


def transpose(ctx, a: Tensor, new_order) -> None:
    """``tci::transpose`` in place (manipulator.h:61-68)."""
    check(ctx.lib.hbp_transpose(ctx.h, a.h, i32(new_order), a.h))


def cplx_conj(ctx, a: Tensor) -> None:
    """``tci::cplx_conj`` in place (manipulator.h:85-92)."""
    check(ctx.lib.hbp_conj(ctx.h, a.h))


def contract(ctx, a: Tensor, la, b: Tensor, lb, lc, out: Tensor | None = None) -> Tensor:
    """``tci::contract`` (linalg.h:311-323); ``out`` may be ``a`` or ``b``."""
    c = out if out is not None else Tensor(ctx, (), HBP_C128)
    check(ctx.lib.hbp_contract(ctx.h, a.h, i32(la), b.h, i32(lb), c.h, i32(lc), len(lc)))
    return c


def norm(ctx, a: Tensor) -> float:
    """``tci::norm`` (linalg.h:67-72)."""
    v = C.c_double()
    check(ctx.lib.hbp_norm(ctx.h, a.h, C.byref(v)))
    return v.value


def normalize(ctx, a: Tensor) -> float:
    """``tci::normalize`` in place, returning the norm (linalg.h:77-84)."""
    v = C.c_double()
    check(ctx.lib.hbp_normalize(ctx.h, a.h, C.byref(v)))
    return v.value


def linear_combine(ctx, ins, coefs) -> Tensor:
    """``tci::linear_combine`` (linalg.h:485-493)."""
    arr = (C.c_void_p * len(ins))(*[t.h for t in ins])
    cf = (C.c_double * (2 * len(coefs)))(*[x for c in coefs for x in (complex(c).real, complex(c).imag)])
    out = Tensor(ctx, (), HBP_C128)
    check(ctx.lib.hbp_linear_combine(ctx.h, len(ins), arr, cf, out.h))
    return out


def trace(ctx, a: Tensor, pairs) -> Tensor:
    """``tci::trace`` (linalg.h:142-149)."""
    flat = [x for p in pairs for x in p]
    out = Tensor(ctx, (), HBP_C128)
    check(ctx.lib.hbp_trace(ctx.h, a.h, len(pairs), i32(flat), out.h))
    return out


def diag(ctx, a: Tensor) -> None:
    """``tci::diag`` in place (linalg.h:31-44); a real vector becomes a complex
    matrix only through ``to_cplx`` first, as in the reference."""
    check(ctx.lib.hbp_diag(ctx.h, a.h, a.h))


def to_cplx(ctx, a: Tensor) -> Tensor:
    """``tci::to_cplx`` (miscellaneous.h:107-142)."""
    return Tensor.from_numpy(ctx, a.to_numpy().astype(np.complex128))


def get_elem(ctx, a: Tensor, coors) -> complex:
    """``tci::get_elem`` (getter.h:50-80)."""
    re, im = C.c_double(), C.c_double()
    check(ctx.lib.hbp_get_elem(ctx.h, a.h, i32(coors), C.byref(re), C.byref(im)))
    return complex(re.value, im.value)


def svd(ctx, a: Tensor, nrows: int):
    """``tci::svd`` (linalg.h:500-512) -> (u, s, vdag)."""
    u, s, vd = Tensor(ctx, ()), Tensor(ctx, (), HBP_F64), Tensor(ctx, ())
    check(ctx.lib.hbp_svd(ctx.h, a.h, nrows, u.h, s.h, vd.h))
    return u, s, vd


def trunc_svd(ctx, a: Tensor, nrows: int, chi_max: int, s_min: float):
    """``tci::trunc_svd`` (linalg.h:552-569) -> (u, s, vdag, trunc_err)."""
    u, s, vd = Tensor(ctx, ()), Tensor(ctx, (), HBP_F64), Tensor(ctx, ())
    err = C.c_double()
    check(ctx.lib.hbp_trunc_svd(ctx.h, a.h, nrows, chi_max, s_min, u.h, s.h, vd.h, C.byref(err)))
    return u, s, vd, err.value
