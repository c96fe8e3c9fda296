"""GPU parity: every HIP kernel path against the CPU FP64 oracle, through the C ABI.

Tolerances (FP64 paths): the stencils sum the same terms in a different order than
the oracle's separate DEC passes, so results agree to rounding: 1e-12 relative.
Solves are compared after both reach a 1e-10 true relative residual.
"""
import numpy as np
import pytest

from helpers import oracle_desc, random_dofs, setup, solver_for
from conftest import rel

pytestmark = pytest.mark.gpu

CASES = [(3, 8), (3, 16), (2, 16), (2, 32)]


@pytest.mark.parametrize("dim,n", CASES)
def test_apply_matches_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    for seed in range(2):
        x = random_dofs(P, seed)
        y = s.applyMapping(x)
        yo = orc.apply(dim, P.shape, fixed, vol, pen, x)
        assert rel(y, yo) < 1e-12


@pytest.mark.parametrize("dim,n", CASES)
def test_design_alpha_matches_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    f0, a0 = s.mg_level_desc(0)
    nc = orc.sizes(dim, *P.shape)[1][0]
    assert np.array_equal(f0, P.fixed)
    assert rel(a0, pen[nc:]) < 1e-14


@pytest.mark.parametrize("dim,n", CASES)
def test_objective_and_adjoint_rhs(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x = random_dofs(P, 3)
    J = s.compute_energy(x)
    Jo = orc.objective(dim, P.shape, fixed, vol, pen, x)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    g = s.compute_gradient_x(x)
    go = orc.energy_op(dim, P.shape, fixed, vol, pen, x)
    assert rel(g, go) < 1e-12


@pytest.mark.parametrize("dim,n", CASES)
def test_sensitivity_matches_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    nc = orc.sizes(dim, *P.shape)[1][0]
    x = random_dofs(P, 4)
    lam = random_dofs(P, 5, dirichlet=False)
    dj = s.sensitivity(rho, x, lam)
    djo = orc.sensitivity(dim, P.shape, rho, P.alpha.array(), P.fixed[nc:], x[nc:], lam[nc:])
    assert rel(dj, djo) < 1e-13


@pytest.mark.parametrize("dim,n", CASES)
def test_filter_matches_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    v = np.random.default_rng(1).random(rho.size)
    for tr in (False, True):
        assert rel(s.filter(v, tr), orc.filt(dim, P.shape, 1.5, v, tr)) < 1e-14


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 32)])
def test_oc_matches_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    dc = -np.random.default_rng(2).random(rho.size) * 1e-2
    r1, c1 = s.oc_update(rho, dc, 0.3)
    r2, c2 = orc.oc_update(dim, P.shape, 1.5, rho, dc, 0.3)
    assert np.max(np.abs(r1 - r2)) < 1e-9
    assert abs(c1 - c2) < 1e-9
    # later updates start their multiplier search around the previous multiplier: the
    # same result as the oracle's bisection from its full bracket, whether the new
    # multiplier lies inside that start (a similar step) or far outside it (a scaled
    # sensitivity moves the multiplier by 1e4)
    rng = np.random.default_rng(3)
    for scale in (1.0, 1e4, 1e-4):
        rho_k = np.clip(r1 + 0.01 * rng.standard_normal(rho.size), 1e-3, 1.0)
        dc_k = dc * (1.0 + 0.1 * rng.random(rho.size)) * scale
        ra, ca = s.oc_update(rho_k, dc_k, 0.3)
        rb, cb = orc.oc_update(dim, P.shape, 1.5, rho_k, dc_k, 0.3)
        assert np.max(np.abs(ra - rb)) < 1e-9, scale
        assert abs(ca - cb) < 1e-9, scale


@pytest.mark.parametrize("dim,n", [(3, 16), (2, 32)])
def test_mg_components_match_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    mg = orc.MG(dim, P.shape, fixed, pen, nu_pre=2, nu_post=2, omega=1.0, tang=0)
    assert s.mg_levels() == mg.levels
    rng = np.random.default_rng(7)
    for l in range(mg.levels):
        fo, po = mg.level_desc(l)
        f, a = s.mg_level_desc(l)
        nl = mg.level_dims(l)[0]
        nc = orc.sizes(dim, nl, nl, nl if dim == 3 else 1)[1][0]
        assert np.array_equal(f, fo), f"flags level {l}"
        assert rel(a, po[nc:]) < 1e-13, f"alpha level {l}"
        shp = (nl, nl, nl if dim == 3 else 1)
        b = rng.standard_normal(fo.size)
        xin = rng.standard_normal(fo.size)
        vol_l = np.where(np.arange(fo.size) >= nc, 1.0 - fo, 0.0)
        for color in (0, 1):
            xo = orc.smooth_color(dim, shp, fo, po, b, xin, color, 1.0)
            xg = s.mg_smooth(l, color, b, xin)
            assert rel(xg, xo) < 1e-12, f"smooth level {l} colour {color}"
        if l + 1 < mg.levels:
            fc, pc = mg.level_desc(l + 1)
            bo = orc.restrict(dim, shp, fc, 0, b)
            assert rel(s.mg_restrict(l, b), bo) < 1e-13
            ec = rng.standard_normal(fc.size)
            xo = orc.prolong_add(dim, shp, fo, 0, ec, xin)
            assert rel(s.mg_prolong(l, ec, xin), xo) < 1e-13


@pytest.mark.parametrize("nu", [(1, 2), (2, 2)])
@pytest.mark.parametrize("dim,n", [(3, 16), (2, 32), (3, 64)])
def test_vcycle_matches_oracle(make_solver, orc, dim, n, nu):
    """At 64^3 the fine level runs the tiled kernels, including the smoother pass from
    x = 0 that starts the V-cycle.  V(1,2) is the default, V(2,2) the other sweep
    counts' representative."""
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho, nu_pre=nu[0], nu_post=nu[1])
    fixed, vol, pen = oracle_desc(orc, P, rho)
    mg = orc.MG(dim, P.shape, fixed, pen, nu_pre=nu[0], nu_post=nu[1], omega=1.0, tang=0)
    b = orc.lift_rhs(dim, P.shape, fixed, P.dirichlet)
    z = s.mg_vcycle(b)
    zo = mg.vcycle(b)
    assert rel(z, zo) < 1e-10


@pytest.mark.parametrize("dim,n", [(3, 16), (2, 32)])
def test_solve_matches_oracle(make_solver, orc, dim, n):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho, rtol=1e-10)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x, st = s.solve()
    assert st["converged"] and st["relres"] <= 1e-10
    b = orc.lift_rhs(dim, P.shape, fixed, P.dirichlet)
    # the true residual recomputed by the oracle's operator
    r = b - orc.apply(dim, P.shape, fixed, vol, pen, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 2e-10
    mg = orc.MG(dim, P.shape, fixed, pen, nu_pre=2, nu_post=2, omega=1.0, tang=0)
    xo, it, rr, _ = mg.fgmres(b, rtol=1e-12, restart=30, maxit=100)
    assert rr < 1e-12
    assert rel(x, xo) < 1e-7


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("dim,n", [(3, 16), (2, 32)])
def test_mixed_precision_solve(make_solver, orc, dim, n, prec):
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho, rtol=1e-8, precision=prec)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x, st = s.solve()
    assert st["converged"] and st["relres"] <= 1e-8
    b = orc.lift_rhs(dim, P.shape, fixed, P.dirichlet)
    r = b - orc.apply(dim, P.shape, fixed, vol, pen, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1.5e-8


@pytest.mark.parametrize("dim,n", [(3, 64)])
def test_fine_level_tiled_kernels_match_oracle(make_solver, orc, dim, n):
    """64^3 exercises the 2.5D tiled (marching) operator and smoother kernels."""
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x = random_dofs(P, 11)
    assert rel(s.applyMapping(x), orc.apply(dim, P.shape, fixed, vol, pen, x)) < 1e-12
    b = random_dofs(P, 12, dirichlet=False)
    for color in (0, 1):
        xo = orc.smooth_color(dim, P.shape, fixed, pen, b, x, color, 1.0)
        assert rel(s.mg_smooth(0, color, b, x), xo) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-6), ("bf16", 1e-2)])
def test_mixed_precision_smoother_matches_oracle(make_solver, orc, prec, tol):
    """The FP32 / BF16 smoother colour step against the FP64 oracle: level 0 of 64^3
    runs the tiled (marching) kernel, level 1 (32^3) the per-cell kernel.  The bar is
    the storage rounding of b, x and alpha (measured 2e-7 in FP32, 2.2e-3 in BF16)."""
    dim, n = 3, 64
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho, precision=prec)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    mg = orc.MG(dim, P.shape, fixed, pen, nu_pre=2, nu_post=2, omega=1.0, tang=0)
    rng = np.random.default_rng(20)
    for l in (0, 1):
        fo, po = mg.level_desc(l)
        nl = mg.level_dims(l)[0]
        b = rng.standard_normal(fo.size)
        x = rng.standard_normal(fo.size)
        for color in (0, 1):
            xo = orc.smooth_color(dim, (nl,) * dim, fo, po, b, x, color, 1.0)
            assert rel(s.mg_smooth(l, color, b, x, mixed=True), xo) < tol, (l, color)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("dim,n", [(3, 64)])
def test_fine_level_solve_64(make_solver, orc, dim, n, prec):
    """Tiled kernels in a whole solve; the mixed-precision paths keep their Krylov
    vectors in FP32 and apply the FP64 operator to FP32 vectors (march_apply_mixed),
    the true FP64 residual must still reach rtol."""
    P, rho = setup(dim, n)
    s = solver_for(make_solver, P, rho, rtol=1e-8, precision=prec)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x, st = s.solve()
    assert st["converged"] and st["iters"] <= 15
    b = orc.lift_rhs(dim, P.shape, fixed, P.dirichlet)
    r = b - orc.apply(dim, P.shape, fixed, vol, pen, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1.5e-8


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_solution_is_linear_in_the_boundary_data(make_solver, prec):
    """Dirichlet data scaled by 1e14 must give 1e14 times the solution.  The Krylov basis
    is stored unnormalised, so its vectors reach norms far beyond 1e12 here: the rescale
    that keeps them in range (and, in FP32, inside the float range) must not change the
    result.  64^3 runs the tiled kernels, including the FP32 restart residual."""
    n = 64
    P, rho = setup(3, n)
    a = solver_for(make_solver, P, rho, rtol=1e-10, precision=prec)
    x1, st1 = a.solve()
    b = make_solver(3, n, rtol=1e-10, precision=prec)
    b.init_boundary(P.fixed)
    b.update_b(dirichlet=P.dirichlet * 1e14)
    b.set_design(rho)
    x2, st2 = b.solve()
    assert st1["converged"] and st2["converged"]
    assert rel(x2, 1e14 * x1) < 1e-8
    assert abs(st2["iters"] - st1["iters"]) <= 1


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_solve_is_bitwise_reproducible(make_solver, prec):
    """Every reduction (inner products, residual norms) sums fixed partials in a fixed
    order and no kernel accumulates with atomics, so the same solve gives the same bits."""
    P, rho = setup(3, 64)
    s = solver_for(make_solver, P, rho, rtol=1e-8, precision=prec)
    x1, st1 = s.solve()
    s.reset_solution()
    x2, st2 = s.solve()
    assert st1["iters"] == st2["iters"]
    assert np.array_equal(x1, x2)



def test_design_penalty_follows_a_new_geometry(make_solver):
    """A penalty set from a design is recomputed when st_set_geometry changes the free
    faces afterwards: faces freed by the new geometry get their penalty (a design set
    first and the geometry changed later solves bit for bit as the geometry set first)."""
    n = 16
    P, rho = setup(3, n)
    nc = n ** 3
    idx = nc + np.flatnonzero(P.fixed[nc:] == 0)[::5]  # free faces of the channel
    more = P.fixed.copy()
    more[idx] = 1
    a = make_solver(3, n, rtol=1e-10)
    a.init_boundary(more)
    a.set_design(rho)
    a.init_boundary(P.fixed)  # frees idx again
    a.update_b(dirichlet=P.dirichlet)
    b = solver_for(make_solver, P, rho, rtol=1e-10)
    xa, sa = a.solve()
    xb, sb = b.solve()
    assert sa["converged"] and sb["converged"]
    assert sa["iters"] == sb["iters"]
    assert np.array_equal(xa, xb)


@pytest.mark.parametrize("dim,n,prec,tol", [(3, 16, "fp64", 1e-12), (2, 32, "fp64", 1e-12), (3, 64, "fp64", 1e-12),
                                            (3, 64, "fp32", 1e-6)])
def test_smoother_with_fixed_cells_matches_oracle(make_solver, orc, dim, n, prec, tol):
    """Fixed cells (a definite-fix pressure row) change the smoother block: the
    curl-curl couplings between the cell's faces no longer cancel and the block is
    solved by Sherman-Morrison (pinned against dense probed blocks in test_oracle.py).
    16^3 / 2D: the per-cell kernel; 64^3: the tiled kernel (FP64 and FP32 storage)."""
    P, rho = setup(dim, n)
    nc = n ** dim
    rng = np.random.default_rng(31)
    P.fixed[rng.choice(nc, nc // 16, replace=False)] = 1
    s = solver_for(make_solver, P, rho, precision=prec)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x = random_dofs(P, 13)
    assert rel(s.applyMapping(x), orc.apply(dim, P.shape, fixed, vol, pen, x)) < 1e-12
    b = random_dofs(P, 14, dirichlet=False)
    for color in (0, 1):
        xo = orc.smooth_color(dim, P.shape, fixed, pen, b, x, color, 1.0)
        assert rel(s.mg_smooth(0, color, b, x, mixed=(prec != "fp64")), xo) < tol, color
    if prec == "fp64":
        x, st = s.solve()
        assert st["converged"]
        r = orc.lift_rhs(dim, P.shape, fixed, P.dirichlet) - orc.apply(dim, P.shape, fixed, vol, pen, x)
        assert np.linalg.norm(r) <= 1.5e-8 * np.linalg.norm(orc.lift_rhs(dim, P.shape, fixed, P.dirichlet))


@pytest.mark.parametrize("n", [16, 64])
def test_2d_objective_is_the_reference_face_laplacian(make_solver, orc, n):
    """2D objective and adjoint right-hand side against the literal restatement of the
    reference's 2D FluidEnergyGrad (face 5-point Laplacian of the unmasked velocity,
    FT/FluidEnergyGrad.cu:30-61), with Dirichlet inflow values on fixed faces."""
    P, rho = setup(2, n)
    s = solver_for(make_solver, P, rho)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    for seed in (21, 22):
        x = random_dofs(P, seed)
        go = orc.energy_op_ref2d(P.shape, fixed, pen, x)
        assert rel(s.compute_gradient_x(x), go) < 1e-12
        Jo = 0.5 * x @ go
        assert abs(s.compute_energy(x) - Jo) <= 1e-12 * abs(Jo)
