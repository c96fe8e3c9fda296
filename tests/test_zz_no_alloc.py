"""st_objective_noalloc (no allocation, scratch in a device array of the library)
against st_objective on the same inputs, through the C ABI: two calls in a row per
dimension, the second on a smaller grid, and the size limit."""
import ctypes

import numpy as np
import pytest

from helpers import DeviceArray, random_dofs, setup, solver_for
from stokes_to_amd import _lib
from stokes_to_amd._lib import ST_EINVAL

pytestmark = pytest.mark.gpu


def _inputs(make_solver, dim, n, seed):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    fixed, al = s.mg_level_desc(0)          # the flags and face alpha the context holds
    nc = n ** dim
    pen = np.concatenate([np.zeros(nc), al])
    x = random_dofs(P, seed)
    return s, fixed, pen, x


@pytest.mark.parametrize("dim,sizes", [(3, (32, 16)), (2, (64, 32))])
def test_objective_noalloc_matches_st_objective(make_solver, dim, sizes):
    lib = _lib.load()
    hip = DeviceArray.hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    for call, n in enumerate(sizes):
        s, fixed, pen, x = _inputs(make_solver, dim, n, 40 + call)
        J_ref = s.compute_energy(x)
        df, dp, dx = DeviceArray(fixed), DeviceArray(pen), DeviceArray(x)
        dJ = DeviceArray(np.full(1, np.nan))
        rc = lib.st_objective_noalloc(dim, n, df.ptr, dp.ptr, dx.ptr, dJ.ptr, stream)
        assert rc == 0, lib.st_last_error(None)
        assert hip.hipStreamSynchronize(stream) == 0
        J = float(dJ.get()[0])
        for a in (df, dp, dx, dJ):
            a.free()
        print(f"noalloc dim {dim} n {n}: J {J:.15g} vs st_objective {J_ref:.15g}", flush=True)
        # the same sum in another order: rounding only
        assert abs(J - J_ref) <= 1e-12 * abs(J_ref), (n, J, J_ref)
    hip.hipStreamDestroy(stream)


def test_objective_noalloc_rejects_sizes_beyond_its_limit():
    lib = _lib.load()
    one = ctypes.c_void_p(1)
    rc = lib.st_objective_noalloc(3, 2048, one, one, one, one, None)
    assert rc == ST_EINVAL and b"n must lie" in lib.st_last_error(None)
    rc = lib.st_objective_noalloc(4, 16, one, one, one, one, None)
    assert rc == ST_EINVAL and b"dim" in lib.st_last_error(None)
