"""The allocation-free inner product (edk_vec_dot_noalloc) against
edk_vec_dot through the C ABI: the same kernels and reduction order, so the
results are bitwise equal, on two calls in a row (the second with a smaller
input, whose partials overwrite only part of the shared scratch array), in
ComplexF64 and ComplexF32."""
import ctypes as C

import numpy as np
import pytest

from tests.models import rand_state

pytestmark = pytest.mark.gpu


def _dot(fn, ctx, code, x, y, n):
    re, im = C.c_double(), C.c_double()
    rc = fn(ctx.h, code, x.ptr, y.ptr, n, C.byref(re), C.byref(im))
    assert rc == 0, rc
    return re.value, im.value


@pytest.mark.parametrize("dt,code", [(np.complex128, 0), (np.complex64, 1)])
def test_noalloc_dot_matches_dot(gpu, dt, code):
    import edkit_hip as edk
    from edkit_hip._lib import lib
    ctx = edk.default_context()
    for n in (3_000_001, 70_001, 1):
        x = edk.DeviceArray.from_host(rand_state(n, 1, dt))
        y = edk.DeviceArray.from_host(rand_state(n, 2, dt))
        a = _dot(lib().edk_vec_dot, ctx, code, x, y, n)
        b = _dot(lib().edk_vec_dot_noalloc, ctx, code, x, y, n)
        assert a == b, (n, a, b)
        ref = np.vdot(x.to_host().astype(np.complex128), y.to_host().astype(np.complex128))
        assert abs(complex(*b) - ref) <= 1e-12 * max(1.0, abs(ref))
    # the norm form (y = x) and an empty vector
    x = edk.DeviceArray.from_host(rand_state(1000, 3, dt))
    assert _dot(lib().edk_vec_dot_noalloc, ctx, code, x, x, 1000) == _dot(lib().edk_vec_dot, ctx, code, x, x, 1000)
    assert _dot(lib().edk_vec_dot_noalloc, ctx, code, x, x, 0) == (0.0, 0.0)


def test_noalloc_dot_rejects_bad_arguments(gpu):
    import edkit_hip as edk
    from edkit_hip._lib import lib
    ctx = edk.default_context()
    x = edk.DeviceArray((4,), np.complex128)
    re, im = C.c_double(), C.c_double()
    assert lib().edk_vec_dot_noalloc(ctx.h, 0, x.ptr, x.ptr, -1, C.byref(re), C.byref(im)) != 0
    assert lib().edk_vec_dot_noalloc(ctx.h, 7, x.ptr, x.ptr, 4, C.byref(re), C.byref(im)) != 0
