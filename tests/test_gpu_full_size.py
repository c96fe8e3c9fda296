"""Parity at the full size of the headline measurement (3D 256³, 67.3 M DOFs).

The oracle cannot solve at this size in test time, but everything except the solve is
linear-time and finishes in seconds: the operator, the lifted right-hand side, the
objective and its adjoint right-hand side, the sensitivity and the filter are compared
with the GPU at the same bars as on the small grids, and the GPU's solutions are
checked by the oracle's own FP64 residual (the property a solve must have whatever
path produced it), plus the symmetry of A.  Design: the seeded GRF of the benchmark's
single-solve workload (mean 0.2, sigma 8 cells).
"""
import numpy as np
import pytest

from conftest import rel
from helpers import oracle_desc, random_dofs, setup, solver_for

pytestmark = pytest.mark.gpu

N = 256


@pytest.fixture(scope="module")
def big(orc):
    P, rho = setup(3, N, mean=0.2, sigma=8.0)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    return P, rho, fixed, vol, pen


def test_full_size_operator_objective_sensitivity_filter_match_oracle(make_solver, orc, big):
    P, rho, fixed, vol, pen = big
    s = solver_for(make_solver, P, rho)
    nc = orc.sizes(3, *P.shape)[1][0]
    x = random_dofs(P, 11)
    y = random_dofs(P, 12, dirichlet=False)
    Ax, Ay = s.applyMapping(x), s.applyMapping(y)
    d = rel(Ax, orc.apply(3, P.shape, fixed, vol, pen, x))
    print(f"256^3 apply vs oracle: {d:.2e}", flush=True)
    assert d < 1e-12
    # A is symmetric: y.Ax = x.Ay to rounding
    a, b = float(y @ Ax), float(x @ Ay)
    assert abs(a - b) <= 1e-11 * np.linalg.norm(y) * np.linalg.norm(Ax)
    J = s.compute_energy(x)
    Jo = orc.objective(3, P.shape, fixed, vol, pen, x)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert rel(s.compute_gradient_x(x), orc.energy_op(3, P.shape, fixed, vol, pen, x)) < 1e-12
    lam = random_dofs(P, 13, dirichlet=False)
    dj = s.sensitivity(rho, x, lam)
    djo = orc.sensitivity(3, P.shape, rho, P.alpha.array(), P.fixed[nc:], x[nc:], lam[nc:])
    assert rel(dj, djo) < 1e-13
    v = np.random.default_rng(14).random(rho.size)
    for tr in (False, True):
        assert rel(s.filter(v, tr), orc.filt(3, P.shape, 1.5, v, tr)) < 1e-14


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_full_size_solve_reaches_the_oracle_residual(make_solver, orc, big, prec):
    """The GPU solution's residual recomputed by the oracle in FP64 (its own lifting and
    operator) is at the solver's 1e-8 tolerance, as the library reports it."""
    P, rho, fixed, vol, pen = big
    s = solver_for(make_solver, P, rho, precision=prec, rtol=1e-8)
    x, st = s.solve()
    b = orc.lift_rhs(3, P.shape, fixed, P.dirichlet)
    r = np.linalg.norm(b - orc.apply(3, P.shape, fixed, vol, pen, x)) / np.linalg.norm(b)
    print(f"256^3 {prec}: {st['iters']} it, {st['ms']:.1f} ms, library relres {st['relres']:.3e}, "
          f"oracle relres {r:.3e}", flush=True)
    assert st["converged"] and st["relres"] <= 1e-8
    assert r <= 1e-8 * (1 + 1e-3)
    assert abs(r - st["relres"]) <= 1e-3 * st["relres"]
