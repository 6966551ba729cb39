"""hbp_svd_noalloc (allocation-free SVD on caller-owned device buffers, work
arrays in the library's fixed device scratch) against hbp_svd on the same
matrices, through the C-ABI: two calls in a row, the second on a smaller
matrix, so that a stale scratch from the first call would show.  Both run
the same Jacobi kernel on the same input, so the factors must agree bit for
bit; also checked against numpy's singular values (1e-12 relative) and the
reconstruction U diag(s) V^H = A (1e-12)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu


def _dptr(ctx, t):
    p = C.c_void_p()
    from tnbp_amd._lib import check
    check(ctx.lib.hbp_tensor_data(t.h, C.byref(p)))
    return p


def test_svd_noalloc_equals_svd_twice_shrinking():
    from tnbp_amd._lib import HBP_F64, check
    from tnbp_amd.tci import Context, Tensor, svd
    rng = np.random.default_rng(77)
    with Context(0) as ctx:
        for m, n in ((96, 64), (40, 24)):
            announce(f"svd_noalloc m={m} n={n}")
            a = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
            A = Tensor.from_numpy(ctx, a)
            u_ref, s_ref, vd_ref = svd(ctx, A, 1)
            U, S, Vd = Tensor(ctx, (m, n)), Tensor(ctx, (n,), HBP_F64), Tensor(ctx, (n, n))
            check(ctx.lib.hbp_svd_noalloc(ctx.h, _dptr(ctx, A), m, n, _dptr(ctx, U),
                                          C.cast(_dptr(ctx, S), C.POINTER(C.c_double)),
                                          _dptr(ctx, Vd)))
            ctx.synchronize()
            u, s, vd = U.to_numpy(), S.to_numpy(), Vd.to_numpy()
            np.testing.assert_array_equal(u, u_ref.to_numpy())
            np.testing.assert_array_equal(s, s_ref.to_numpy())
            np.testing.assert_array_equal(vd, vd_ref.to_numpy())
            np.testing.assert_allclose(s, np.linalg.svd(a, compute_uv=False), rtol=1e-12)
            np.testing.assert_allclose((u * s) @ vd, a, atol=1e-12 * np.abs(a).max() * n)


def test_svd_noalloc_rejects_what_it_cannot_hold():
    from tnbp_amd._lib import check
    from tnbp_amd.tci import Context
    with Context(0) as ctx:
        dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
        with pytest.raises(NotImplementedError, match="16 MiB"):
            check(ctx.lib.hbp_svd_noalloc(ctx.h, dummy, 4096, 512, dummy,
                                          C.cast(dummy, C.POINTER(C.c_double)), dummy))
        with pytest.raises(NotImplementedError, match="rows >= columns"):
            check(ctx.lib.hbp_svd_noalloc(ctx.h, dummy, 8, 16, dummy,
                                          C.cast(dummy, C.POINTER(C.c_double)), dummy))
