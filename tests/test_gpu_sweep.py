"""The one-pass two-colour smoothing sweep (k_sweep, st_sweep.hip) against two colour
passes (k_march SMOOTH) and the oracle.

The sweep runs the first colour on a ring of cells around each tile (recomputed by the
tiles that need it) and the second colour on the tile, with the same block arithmetic
as the colour passes; in FP32 storage the intermediate values are the same FP32
numbers either way, so V-cycles with and without it agree to rounding of the
operation order (bar 1e-6 relative, FP32 storage rounding of the colour passes).
"""
import numpy as np
import pytest

from conftest import rel
from helpers import oracle_desc, setup, solver_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_vcycle_with_one_pass_sweeps_matches_colour_passes(make_solver, orc, n, prec):
    P, rho = setup(3, n)
    a = solver_for(make_solver, P, rho, precision=prec, fused_sweep=1, nu_pre=2, nu_post=2)
    b = solver_for(make_solver, P, rho, precision=prec, fused_sweep=0, nu_pre=2, nu_post=2)
    rng = np.random.default_rng(n)
    v = rng.standard_normal(P.fixed.size)
    za, zb = a.mg_vcycle(v), b.mg_vcycle(v)
    d = rel(za, zb)
    print(f"{prec} {n}^3 V-cycle, one-pass sweeps vs colour passes: rel diff {d:.2e}", flush=True)
    # BF16: the colour passes round the intermediate state to BF16, the sweep keeps it in FP32
    assert d < (1e-6 if prec == "fp32" else 2e-2)
    if n == 64 and prec == "fp32":
        fixed, vol, pen = oracle_desc(orc, P, rho)
        mg = orc.MG(3, P.shape, fixed, pen, nu_pre=2, nu_post=2, omega=1.0, tang=0)
        zo = mg.vcycle(v)
        assert rel(za, zo) < 1e-5


@pytest.mark.parametrize("n", [64, 128])
def test_solve_with_one_pass_sweeps(make_solver, orc, n):
    """Whole FP32 solves: same iterations (+-1) and the true FP64 residual at rtol; the
    solution matches the solve with colour passes to the solver tolerance."""
    P, rho = setup(3, n)
    a = solver_for(make_solver, P, rho, precision="fp32", rtol=1e-8, fused_sweep=1)
    b = solver_for(make_solver, P, rho, precision="fp32", rtol=1e-8, fused_sweep=0)
    xa, sa = a.solve()
    xb, sb = b.solve()
    print(f"{n}^3 FP32 solve: one-pass sweeps {sa['iters']} it {sa['ms']:.1f} ms, colour passes {sb['iters']} it "
          f"{sb['ms']:.1f} ms", flush=True)
    assert sa["converged"] and sa["relres"] <= 1e-8
    assert abs(sa["iters"] - sb["iters"]) <= 1
    assert rel(xa, xb) < 1e-6
