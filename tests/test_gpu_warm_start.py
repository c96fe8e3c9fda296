"""The extrapolated start of a sequence of solves (st_params.warm_order).

A solve with ST_INIT_KEEP starts from the quadratic (or linear) extrapolation of the
sequence's previous solutions when its true FP64 residual is below that of the held
solution (orders 1, 2), or from the combination of the held solution and the last
differences of solutions with the least residual (order 3, the FP32 path).  The
contract: the start is never worse than the held solution, the solutions and the
optimisation's designs do not change beyond the solver tolerance, and along an
optimisation run the starts are better.
"""
import numpy as np
import pytest

from conftest import rel
from helpers import oracle_desc, setup, solver_for
from stokes_to_amd.problems import channel_problem

pytestmark = pytest.mark.gpu


def _sequence(n, k, seed=7):
    """k designs moving smoothly from one GRF to another, then a jump to a third."""
    _, a = setup(3, n, mean=0.3, seed=seed)
    _, b = setup(3, n, mean=0.3, seed=seed + 1)
    _, c = setup(3, n, mean=0.4, seed=seed + 2)
    return [(1 - t) * a + t * b for t in np.linspace(0.0, 0.3, k)] + [c]


@pytest.mark.parametrize("order,n", [(1, 32), (2, 32), (3, 64)])
def test_extrapolated_start_is_never_worse_than_the_held_solution(make_solver, order, n):
    """(order 3 at 64^3: its path needs the tiled finest level; smaller grids fall back
    to order 2)"""
    P = channel_problem(3, n)
    designs = _sequence(n, 7)
    runs = {}
    for wo in (0, order):
        s = solver_for(make_solver, P, None, precision="fp32", rtol=1e-8, warm_order=wo)
        out = []
        for k, d in enumerate(designs):
            s.set_design(d)
            x, st = s.solve(init="zero" if k == 0 else "keep")
            assert st["converged"] and st["relres"] <= 1e-8
            out.append((x, st))
        runs[wo] = out
    for k in range(1, len(designs)):
        (x0, s0), (x1, s1) = runs[0][k], runs[order][k]
        # the held solutions of the two runs differ at the solver tolerance only
        assert s1["relres0"] <= s0["relres0"] + 1e-7, (k, s1["relres0"], s0["relres0"])
        assert rel(x1, x0) < 1e-6
    # along the smooth part the extrapolation is taken and starts closer
    assert sum(runs[order][k][1]["relres0"] for k in range(3, 7)) < 0.5 * sum(runs[0][k][1]["relres0"] for k in range(3, 7))


def test_minimal_residual_start_is_never_worse_than_the_extrapolation(make_solver):
    """warm_order 3 minimises the residual over the held solution plus the last three
    differences of solutions, a span that holds both polynomial extrapolations: along the
    smooth part of a sequence its start's residual is at most the quadratic
    extrapolation's (up to the solver tolerance by which the held solutions of the two
    runs differ), and the solutions agree.  64^3: the tiled finest level."""
    n = 64
    P = channel_problem(3, n)
    designs = _sequence(n, 8)
    runs = {}
    for wo in (2, 3):
        s = solver_for(make_solver, P, None, precision="fp32", rtol=1e-8, warm_order=wo)
        out = []
        for k, d in enumerate(designs):
            s.set_design(d)
            x, st = s.solve(init="zero" if k == 0 else "keep")
            assert st["converged"] and st["relres"] <= 1e-8
            out.append((x, st))
        runs[wo] = out
    for k in range(3, len(designs) - 1):
        (x2, s2), (x3, s3) = runs[2][k], runs[3][k]
        assert s3["relres0"] <= 1.05 * s2["relres0"] + 1e-8, (k, s3["relres0"], s2["relres0"])
        assert rel(x3, x2) < 1e-6
    # the jump to an unrelated design: never worse than the held solution either
    assert runs[3][-1][1]["relres0"] <= runs[2][-1][1]["relres0"] * 1.05 + 1e-8


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_order_3_on_other_paths_is_the_quadratic_extrapolation(make_solver, prec):
    """The minimal-residual start belongs to the FP32 path; on the FP64 and BF16 paths
    order 3 is order 2: the same starts and solutions, bit for bit."""
    n = 64
    P = channel_problem(3, n)
    designs = _sequence(n, 5)
    runs = {}
    for wo in (2, 3):
        s = solver_for(make_solver, P, None, precision=prec, rtol=1e-8, warm_order=wo)
        out = []
        for k, d in enumerate(designs):
            s.set_design(d)
            x, st = s.solve(init="zero" if k == 0 else "keep")
            assert st["converged"]
            out.append((x, st))
        runs[wo] = out
    for (x2, s2), (x3, s3) in zip(runs[2], runs[3]):
        assert s2["relres0"] == s3["relres0"] and s2["iters"] == s3["iters"]
        assert np.array_equal(x2, x3)


def test_minimal_residual_start_is_the_least_squares_optimum(make_solver, orc):
    """The start of order 3 against an independent restatement: with the solutions the
    GPU returned, the oracle's FP64 operator and right-hand side of the new design, and
    numpy's least squares over x1 + span{x1 - x2, x2 - x3, x3 - x4}, the optimal residual
    equals the solve's initial residual (relres0) to the FP32 rounding of the
    differences and their products (bar 1e-3 relative)."""
    n = 64
    P = channel_problem(3, n)
    designs = _sequence(n, 7)[:-1]  # the smooth part
    s = solver_for(make_solver, P, None, precision="fp32", rtol=1e-8, warm_order=3)
    xs = []
    checked = 0
    for k, d in enumerate(designs):
        s.set_design(d)
        x, st = s.solve(init="zero" if k == 0 else "keep")
        assert st["converged"]
        if k >= 4:
            fixed, vol, pen = oracle_desc(orc, P, d)
            A = lambda v: orc.apply(3, P.shape, fixed, vol, pen, v)
            b = orc.lift_rhs(3, P.shape, fixed, P.dirichlet)
            x1 = xs[-1]
            D = [xs[-1] - xs[-2], xs[-2] - xs[-3], xs[-3] - xs[-4]]
            r1 = b - A(x1)
            Y = np.stack([A(v) for v in D], axis=1)
            c, *_ = np.linalg.lstsq(Y, r1, rcond=None)
            r_opt = np.linalg.norm(r1 - Y @ c) / np.linalg.norm(b)
            r_held = np.linalg.norm(r1) / np.linalg.norm(b)
            print(f"step {k}: relres0 {st['relres0']:.6e}, least-squares optimum {r_opt:.6e}, "
                  f"held solution {r_held:.3e}", flush=True)
            assert r_opt < 0.9 * r_held
            assert abs(st["relres0"] - r_opt) <= 1e-3 * r_opt
            checked += 1
        xs.append(x)
    assert checked == 3


def test_optimisation_run_with_extrapolated_starts(make_solver):
    """A 64^3 OC run (24 iterations, forward and adjoint warm-started): the same
    objectives and final design as with the held solution as start, fewer iterations
    with the quadratic extrapolation and fewer still with the minimal-residual start."""
    n, vf, its = 64, 0.25, 24
    P = channel_problem(3, n)
    res = {}
    for wo in (0, 2, 3):
        s = solver_for(make_solver, P, np.full(n ** 3, vf), precision="fp32", rtol=1e-8, warm_order=wo)
        J, fi, ai = [], 0, 0
        for it in range(its):
            o = s.to_step(vf, init="keep")
            J.append(o["J"])
            if it >= 3:
                fi += o["fwd"]["iters"]
                ai += o["adj"]["iters"]
        res[wo] = (np.array(J), s.get_design(), fi, ai)
    J0, d0, f0, a0 = res[0]
    J2, d2, f2, a2 = res[2]
    J3, d3, f3, a3 = res[3]
    print(f"forward iterations {f0} -> {f2} -> {f3}, adjoint {a0} -> {a2} -> {a3}", flush=True)
    for J, d in ((J2, d2), (J3, d3)):
        assert np.max(np.abs(J - J0) / np.abs(J0)) < 1e-6
        assert rel(d, d0) < 1e-4
    assert f2 < f0 and a2 <= a0
    assert f3 < f2 and a3 <= a2
