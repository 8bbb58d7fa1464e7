"""dm_apply_noalloc: the allocation-free one-vector product (device pointers,
scratch in a per-GPU device array) against dm_apply, which allocates, on the
same stored matrices: every stored form, both op(Z) = Z and Z^H, two matrices
in a row (the second smaller), x == y in place, and the size limit."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
K0 = 2 * math.pi / 0.1


def _hip():
    """The HIP runtime the library already loaded (device buffers for the test)."""
    try:
        h = C.CDLL("libamdhip64.so.7")
    except OSError:
        h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _noalloc(dm, Z, form, adjoint, x_host, inplace=False):
    hip = _hip()
    x_host = np.ascontiguousarray(x_host, dtype=np.complex128)
    nb = x_host.nbytes
    dx, dy = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dx), nb) == 0
    assert hip.hipMalloc(C.byref(dy), nb) == 0
    try:
        assert hip.hipMemcpy(dx, x_host.ctypes.data_as(C.c_void_p), nb, 1) == 0  # host to device
        out = dx if inplace else dy
        dm._check(dm.lib().dm_apply_noalloc(Z.ctx.h, Z.h, form, int(adjoint), dx, out))
        Z.ctx.sync()
        y = np.empty_like(x_host)
        assert hip.hipMemcpy(y.ctypes.data_as(C.c_void_p), out, nb, 2) == 0  # device to host
        return y
    finally:
        hip.hipFree(dx)
        hip.hipFree(dy)


@pytest.mark.parametrize("dims", [(0.3, 0.3, 30, 28), (0.12, 0.1, 10, 9)], ids=["larger", "smaller"])
def test_noalloc_product_matches_dm_apply(dm, dims):
    m = dm.make_rect_plate(*dims)
    r = dm.build_rwg(m)
    forms = dm.DM_Z64 | dm.DM_Z32 | dm.DM_Z64P | dm.DM_Z32P
    Z = dm.assemble_Z_efie(m, r, K0, forms=forms, device=True)
    rng = np.random.default_rng(r.nedges)
    x = rng.standard_normal(r.nedges) + 1j * rng.standard_normal(r.nedges)
    for form in (dm.DM_Z64, dm.DM_Z32, dm.DM_Z64P, dm.DM_Z32P):
        for adjoint in (False, True):
            y_ref = Z.matmul(x, form=form, adjoint=adjoint)
            y = _noalloc(dm, Z, form, adjoint, x)
            err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
            print(f"N={r.nedges} form {form} adjoint {adjoint}: rel {err:.2e}", flush=True)
            assert err <= 1e-13, (form, adjoint, err)
    y = _noalloc(dm, Z, dm.DM_Z64P, False, x, inplace=True)
    assert np.linalg.norm(y - Z.matmul(x, form=dm.DM_Z64P)) <= 1e-13 * np.linalg.norm(y)


def test_noalloc_refuses_sizes_beyond_its_scratch(dm):
    m = dm.make_rect_plate(0.7, 0.7, 70, 70)
    r = dm.build_rwg(m)
    assert r.nedges > 11392
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z32P, device=True, max_output_bytes=None)
    x = np.ones(r.nedges, np.complex128)
    with pytest.raises(ValueError, match="scratch"):
        _noalloc(dm, Z, dm.DM_Z32P, False, x)
    with pytest.raises(ValueError):
        _noalloc(dm, Z, dm.DM_Z64, False, x)  # form not assembled
