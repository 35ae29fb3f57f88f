"""omhd_set_dt_noalloc (the allocation-free form of omhd_set_dt) through the
C ABI: the same vmax and dt, bit for bit, as omhd_set_dt on the same device
arrays, on two calls in a row where the second grid is smaller and slower
than the first (so partial maxima the first call left in the shared scratch
array would show), in 2D and 3D; and its argument checks."""
import ctypes as C

import numpy as np
import pytest

from omhd_testlib import prim_to_cons, random_prim

pytestmark = pytest.mark.gpu

D = C.c_double
P = C.c_void_p
I = C.c_int32
SIG = [P, P, C.POINTER(D), C.POINTER(D), D, D, I, I, I, I]


def lib():
    from openmhd_amd import LIB_PATH

    L = C.CDLL(str(LIB_PATH))
    for name in ("omhd_set_dt", "omhd_set_dt_noalloc"):
        getattr(L, name).argtypes = SIG
        getattr(L, name).restype = C.c_int
    return L


def device_state(rng, ix, jx, kx, vscale):
    """Fortran-ordered U(ix,jx,kx,9) and V(ix,jx,kx,4) as flat device tensors."""
    import torch

    W = random_prim(rng, (ix, jx, kx), vscale=vscale)
    U = np.asfortranarray(prim_to_cons(W))
    V = np.asfortranarray(W[..., :4])
    return (torch.tensor(U.ravel(order="F"), device="cuda"), torch.tensor(V.ravel(order="F"), device="cuda"))


def call(fn, U, V, ix, jx, kx, nd, dx=0.01, cfl=0.4):
    vmax, dt = D(), D()
    rc = fn(P(U.data_ptr()), P(V.data_ptr()), C.byref(vmax), C.byref(dt), dx, cfl, ix, jx, kx, nd)
    return rc, vmax.value, dt.value


def bits(x):
    return np.float64(x).view(np.int64)


@pytest.mark.parametrize("nd", (2, 3))
def test_noalloc_set_dt_equals_set_dt(nd):
    L = lib()
    rng = np.random.default_rng(7 + nd)
    # first a large, fast grid (many partial slots, large maxima), then a
    # small, slow one: the second call must not see the first call's slots
    cases = ([(1030, 770, 1, 4.0), (67, 45, 1, 0.5)] if nd == 2 else
             [(130, 98, 66, 4.0), (18, 12, 10, 0.5)])
    last = None
    for ix, jx, kx, vs in cases:
        U, V = device_state(rng, ix, jx, kx, vs)
        rc0, vm0, dt0 = call(L.omhd_set_dt, U, V, ix, jx, kx, nd)
        rc1, vm1, dt1 = call(L.omhd_set_dt_noalloc, U, V, ix, jx, kx, nd)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1)
        assert bits(vm1) == bits(vm0) and bits(dt1) == bits(dt0), (ix, jx, kx, vm0, vm1)
        if last is not None:
            assert vm1 < last  # the second grid really is slower
        last = vm1


def test_noalloc_set_dt_argument_checks():
    import torch

    L = lib()
    rng = np.random.default_rng(1)
    U, V = device_state(rng, 12, 10, 1, 1.0)
    # host arrays are refused (staging them would allocate)
    Uh, Vh = U.cpu().numpy(), V.cpu().numpy()
    vmax, dt = D(), D()
    rc = L.omhd_set_dt_noalloc(P(Uh.ctypes.data), P(Vh.ctypes.data), C.byref(vmax), C.byref(dt), 0.1, 0.4, 12, 10, 1,
                               2)
    assert rc == 8  # OMHD_ERR_ARG
    # a grid above the scratch array's limit is refused before any device work
    rc, _, _ = call(L.omhd_set_dt_noalloc, U, V, 8194, 8194, 5, 3)  # 8192 x 8192 x 3 interior cells
    assert rc == 6  # OMHD_ERR_EXTENT
    # and the entry point still works afterwards
    rc0, vm0, dt0 = call(L.omhd_set_dt, U, V, 12, 10, 1, 2)
    rc1, vm1, dt1 = call(L.omhd_set_dt_noalloc, U, V, 12, 10, 1, 2)
    assert rc0 == rc1 == 0 and bits(vm0) == bits(vm1) and bits(dt0) == bits(dt1)
    torch.cuda.synchronize()
