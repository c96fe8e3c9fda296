"""Failure modes and context state through the C ABI (on the GPU).

The reference returns silently after its iteration cap (CX/GMRES.cpp:46-95), ignores
the LU `info` of its coarsest solve (ST/StokeLike.cu:640) and has no NaN checks; this
library returns ST_ENOCONV, ST_ESINGULAR and ST_ENAN with a message instead.  Each
case below forces the condition from the host (an iteration cap of one, a closed box
without its pressure gauge, a NaN in an input) -- every kernel still runs to its end --
and then checks that the same context solves the valid problem afterwards.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel
from helpers import DeviceArray, setup, solver_for
from stokes_to_amd import _lib
from stokes_to_amd._lib import ST_ENAN, ST_ENOCONV, ST_ESINGULAR, StokesError, dptr
from stokes_to_amd.problems import channel_problem, reference_test0

pytestmark = pytest.mark.gpu


def _err(s):
    return s.lib.st_last_error(s.h).decode()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_iteration_cap_returns_enoconv_and_context_stays_usable(make_solver, prec):
    P, rho = setup(3, 16)
    s = solver_for(make_solver, P, rho, rtol=1e-10, precision=prec, max_iter=1)
    with pytest.raises(StokesError) as e:
        s.solve()
    assert e.value.code == ST_ENOCONV and "ENOCONV" in str(e.value)
    x, st = s.solve(allow_noconv=True)              # the capped result is still returned
    assert st["iters"] == 1 and not st["converged"] and st["relres"] < st["relres0"]
    ref = solver_for(make_solver, P, rho, rtol=1e-10, precision=prec)
    xr, sr = ref.solve()
    # continue from the capped iterate with the cap lifted: same solution
    p = _lib.StParams()
    s.lib.st_get_params(s.h, ctypes.byref(p))
    assert p.max_iter == 1
    s2 = solver_for(make_solver, P, rho, rtol=1e-10, precision=prec)
    x2, st2 = s2.solve(init="host", x0=x)
    assert st2["converged"] and rel(x2, xr) < 1e-8


def test_closed_box_without_gauge_is_singular_then_solves_with_it(make_solver):
    """The reference's test 0 is a closed box: without the all-ones pressure gauge the
    coarsest matrix is singular (ST_ESINGULAR); st_set_geometry(..., closed = 1) on the
    same context rebuilds the coarsest inverse and the solve converges."""
    P = reference_test0(3, 32)
    a = P.alpha
    s = make_solver(3, 32, alpha_max=a.alpha_max, alpha_min=a.alpha_min, q=a.q, power=a.power,
                    alpha_scale=a.scale, filter_radius=1.0, rtol=1e-10)
    s.init_boundary(P.fixed, closed=False)
    s.update_b(body=P.body)
    s.set_design(np.full(32 ** 3, 0.3))
    with pytest.raises(StokesError) as e:
        s.solve()
    assert e.value.code == ST_ESINGULAR and "singular" in str(e.value)
    s.init_boundary(P.fixed, closed=True)
    x, st = s.solve()
    assert st["converged"] and np.isfinite(x).all()


def test_nan_inputs_return_enan_and_context_stays_usable(make_solver):
    P, rho = setup(3, 16)
    s = solver_for(make_solver, P, rho, rtol=1e-10)
    x0, st0 = s.solve()
    # NaN in the design (host array): rejected before anything reaches the device
    bad = rho.copy()
    bad[7] = np.nan
    with pytest.raises(StokesError) as e:
        s.set_design(bad)
    assert e.value.code == ST_ENAN and "design" in str(e.value)
    # NaN in the body force: rejected by st_set_rhs
    body = np.zeros(P.fixed.size)
    body[100] = np.nan
    with pytest.raises(StokesError) as e:
        s.update_b(body=body, dirichlet=P.dirichlet)
    assert e.value.code == ST_ENAN and "body" in str(e.value)
    # NaN in a device design: only the solve sees it (non-finite residual)
    d = DeviceArray(bad)
    ptr = ctypes.cast(d.ptr, ctypes.POINTER(ctypes.c_double))
    assert s.lib.st_set_design_device(s.h, ptr) == 0
    d.free()
    with pytest.raises(StokesError) as e:
        s.solve()
    assert e.value.code == ST_ENAN and "non-finite" in str(e.value)
    # the same context with valid inputs again gives the first solution, bit for bit
    s.set_design(rho)
    s.update_b(dirichlet=P.dirichlet)
    x1, st1 = s.solve()
    assert st1["converged"] and st1["iters"] == st0["iters"] and np.array_equal(x1, x0)


def test_component_entry_points_leave_the_held_solutions_alone(make_solver):
    """st_apply, st_apply_device and st_mg_smooth(_mp) write through scratch buffers: a
    TO step that continues from the held forward and adjoint solutions (init keep)
    needs the same iterations and gives the same bits with those calls in between."""
    dim, n, vf = 3, 16, 0.25
    P = channel_problem(dim, n)
    rho0 = np.full(n ** dim, vf)
    runs = []
    for interleave in (False, True):
        s = solver_for(make_solver, P, rho0, rtol=1e-10, precision="fp32")
        out = [s.to_step(vf, init="keep") for _ in range(2)]
        if interleave:
            rng = np.random.default_rng(3)
            x = rng.standard_normal(P.fixed.size)
            s.applyMapping(x)
            dx, dy = DeviceArray(x), DeviceArray(np.zeros_like(x))
            assert s.lib.st_apply_device(s.h, dx.ptr, dy.ptr) == 0
            assert rel(dy.get(), s.applyMapping(x)) == 0.0
            dx.free()
            dy.free()
            s.mg_smooth(0, 0, x, x)
            s.mg_smooth(0, 1, x, x, mixed=True)
        out.append(s.to_step(vf, init="keep"))
        runs.append(out)
    a, b = runs
    for ra, rb in zip(a, b):
        assert ra["fwd"]["iters"] == rb["fwd"]["iters"] and ra["adj"]["iters"] == rb["adj"]["iters"]
        assert ra["J"] == rb["J"]


def test_fp32_solve_is_unchanged_by_fp64_work_on_shared_buffers(make_solver):
    """128^3 (row pitch 144, padded rows): FP64 component calls write the scratch
    buffers that the FP32 Krylov vectors reuse; a solve afterwards gives the bits of a
    fresh context (the tiled kernels write the row padding of every vector they
    produce, so no stale bytes from another precision enter an inner product)."""
    n = 128
    P, rho = setup(3, n)
    fresh = solver_for(make_solver, P, rho, rtol=1e-8, precision="fp32")
    x1, st1 = fresh.solve()
    fresh.close()
    s = solver_for(make_solver, P, rho, rtol=1e-8, precision="fp32")
    rng = np.random.default_rng(4)
    v = rng.standard_normal(P.fixed.size) * 1e3
    s.applyMapping(v)
    s.mg_smooth(0, 0, v, v)
    s.mg_smooth(0, 1, v, v, mixed=True)
    s.mg_vcycle(v)
    x2, st2 = s.solve()
    assert st1["converged"] and st2["iters"] == st1["iters"]
    assert np.array_equal(x1, x2)


def test_failed_solve_in_a_warm_started_sequence(make_solver):
    """A sequence of warm-started (extrapolated) solves, a solve that fails on a NaN
    design in the middle, then the sequence goes on: the next solve converges to the
    solution of a fresh context."""
    P, rho = setup(3, 16)
    s = solver_for(make_solver, P, rho, rtol=1e-10, precision="fp32")
    s.solve()
    for f in (0.95, 0.9, 0.85):
        s.set_design(rho * f)
        _, st = s.solve(init="keep")
        assert st["converged"]
    bad = rho.copy()
    bad[3] = np.nan
    d = DeviceArray(bad)
    assert s.lib.st_set_design_device(s.h, ctypes.cast(d.ptr, ctypes.POINTER(ctypes.c_double))) == 0
    d.free()
    with pytest.raises(StokesError) as e:
        s.solve(init="keep")
    assert e.value.code == ST_ENAN
    s.set_design(rho * 0.8)
    x, st = s.solve(init="keep")
    ref = solver_for(make_solver, P, rho * 0.8, rtol=1e-10, precision="fp32")
    xr, _ = ref.solve()
    assert st["converged"] and rel(x, xr) < 1e-7
