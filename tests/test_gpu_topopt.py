"""GPU topology-optimization loop against the oracle's loop (same inputs).

Each design iteration runs filter -> design->alpha -> forward solve -> objective ->
adjoint solve -> sensitivity -> filter transpose -> OC on the device (st_to_step);
the oracle composes the same steps on the CPU.  Bar: objective within 1e-6
relative per iteration and final designs within 1e-4 relative L2.
"""
import numpy as np
import pytest

from conftest import rel
from helpers import solver_for
from stokes_to_amd.problems import channel_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n,iters,volfrac", [(3, 16, 6, 0.25), (2, 32, 10, 0.15)])
def test_to_loop_matches_oracle(make_solver, orc, dim, n, iters, volfrac):
    P = channel_problem(dim, n)
    rho0 = np.full(n ** dim, volfrac)
    s = solver_for(make_solver, P, rho0, rtol=1e-10)
    rho_o = rho0.copy()
    for it in range(iters):
        out = s.to_step(volfrac)
        rho_o, J_o, ch_o, _, _ = orc.to_step(P, rho_o, volfrac=volfrac, rtol=1e-10)
        assert out["fwd"]["converged"] and out["adj"]["converged"]
        assert abs(out["J"] - J_o) <= 1e-6 * abs(J_o), (it, out["J"], J_o)
    rho_g = s.get_design()
    assert rel(rho_g, rho_o) < 1e-4
    assert abs(out["vol"] - volfrac) < 1e-8


@pytest.mark.parametrize("n", [4096, 262144])
def test_mma_update_matches_oracle(make_solver, orc, n):
    """st_mma_update (device MMA, one constraint) against the oracle's MMA, which is
    bit-identical to the reference's MMASolver (test_oracle.py).  The dual
    interior-point iteration stops at a barrier level epsi of about 1e-9 sqrt(n)
    (1e-6 at n = 64^3), so two correct implementations whose sums differ in order
    agree on x to about that level after one update, and the asymptote updates carry
    the difference into later ones: bars 1e-7 for the first 5 updates and 1e-4 for
    the next 10 (x in [0, 1]), Newton steps within 1."""
    N = round(n ** (1 / 3))
    s = make_solver(3, N)
    rng = np.random.default_rng(5)
    af = rng.uniform(0.5, 2.0, n)
    bf = rng.uniform(-1.0, 1.0, n)
    grad = lambda x: 2.0 * af * (x - 0.3) + bf * np.cos(3.0 * x)
    xo = np.full(n, 0.5)
    xg = xo.copy()
    M = orc.MMA(n)
    for it in range(15):
        g, dg = orc.volume_constraint(xo, 0.4)
        lo, hi = orc.mma_bounds(xo, 0.2, 0.01, 1.0)
        steps_o = M.update(xo, grad(xo), g, dg, lo, hi)
        g2, dg2 = orc.volume_constraint(xg, 0.4)
        lo2, hi2 = orc.mma_bounds(xg, 0.2, 0.01, 1.0)
        xg, steps_g = s.mma_update(xg, grad(xg), g2, dg2, lo2, hi2)
        assert np.abs(xg - xo).max() < (1e-7 if it < 5 else 1e-4), (it, np.abs(xg - xo).max())
        assert abs(steps_g - steps_o) <= 1


@pytest.mark.parametrize("radius", [1.5, 1.0])
def test_to_loop_mma_matches_oracle(make_solver, orc, radius):
    """TO loop with the reference's optimizer (params.optimizer = MMA) against the
    oracle's loop with its MMA; radius 1.0 is the reference's unfiltered setting."""
    dim, n, iters, volfrac = 3, 16, 6, 0.25
    P = channel_problem(dim, n)
    rho0 = np.full(n ** dim, volfrac)
    s = solver_for(make_solver, P, rho0, rtol=1e-10, optimizer="mma", filter_radius=radius)
    rho_o = rho0.copy()
    M = orc.MMA(rho0.size)
    for it in range(iters):
        out = s.to_step(volfrac, move=0.2, rho_min=0.01)
        rho_o, J_o, ch_o, _, _ = orc.to_step(P, rho_o, radius=radius, volfrac=volfrac, move=0.2, rho_min=0.01,
                                             rtol=1e-10, mma=M)
        assert out["fwd"]["converged"] and out["adj"]["converged"]
        assert abs(out["J"] - J_o) <= 1e-6 * abs(J_o), (it, out["J"], J_o)
        assert abs(out["change"] - ch_o) <= 1e-6 + 1e-4 * ch_o
    assert rel(s.get_design(), rho_o) < 1e-4


def test_learned_mp_run_reaches_the_plain_design(make_solver):
    """Two independent TO runs on the same problem: the plain FP64 method (zero
    starts) and learned + FP32 multigrid (online training on the first iterations,
    learned starts).  Both solve to 1e-8 true FP64 residual, so they must reach the
    same design: final designs within 1e-4 relative L2 and objectives within 1e-6."""
    dim, n, iters, vf = 3, 32, 12, 0.25
    P = channel_problem(dim, n)
    rho0 = np.full(n ** dim, vf)
    a = solver_for(make_solver, P, rho0, rtol=1e-8)
    b = solver_for(make_solver, P, rho0, rtol=1e-8, precision="fp32")
    for it in range(iters):
        ra = a.to_step(vf, init="zero")
        rb = b.to_step(vf, init="learned")
        assert ra["fwd"]["converged"] and rb["fwd"]["converged"]
        if it < 5:
            b.record_sample()
            b.net_train(20, 1e-3)
    assert abs(ra["J"] - rb["J"]) <= 1e-6 * abs(ra["J"])
    assert rel(b.get_design(), a.get_design()) < 1e-4


def test_config_128_learned_to_run_matches_plain_fp64():
    """The 3D 128^3 configuration: channel problem, volume fraction 0.25, 60 OC
    iterations, learned warm start trained online on (rho, u, p) pairs of the run's own
    first 15 iterations, FP32 multigrid inside FP64 refinement, against an independent
    plain FP64 run (zero starts, the reference's choice).  Every solve meets the 1e-8
    true FP64 residual, so the runs agree: objectives within 1e-6 relative at every
    iteration and final designs within 1e-4 relative L2."""
    from stokes_to_amd.steady import FluidSteadyState
    dim, n, iters, vf, train = 3, 128, 60, 0.25, 15
    P = channel_problem(dim, n)
    rho0 = np.full(n ** dim, vf)
    runs = []
    try:
        for prec in ("fp64", "fp32"):
            s = FluidSteadyState(dim, n, rtol=1e-8, precision=prec)
            s.init_boundary(P.fixed)
            s.set_design(rho0)
            s.update_b(dirichlet=P.dirichlet)
            runs.append(s)
        a, b = runs
        its_a, its_b = [], []
        for it in range(iters):
            ra = a.to_step(vf, init="zero")
            rb = b.to_step(vf, init="learned")
            assert ra["fwd"]["converged"] and rb["fwd"]["converged"] and ra["adj"]["converged"]
            assert ra["fwd"]["relres"] <= 1e-8 and rb["fwd"]["relres"] <= 1e-8
            assert abs(ra["J"] - rb["J"]) <= 1e-6 * abs(ra["J"]), (it, ra["J"], rb["J"])
            its_a.append(ra["fwd"]["iters"])
            its_b.append(rb["fwd"]["iters"])
            if it < train:
                b.record_sample()
                b.net_train(3, 1e-3)
            if it % 10 == 0:
                print(f"128^3 it {it}: J {ra['J']:.9g} / {rb['J']:.9g}, forward iterations {its_a[-1]} / {its_b[-1]}",
                      flush=True)
        assert b.num_samples() == train
        d = rel(b.get_design(), a.get_design())
        print(f"128^3 final design rel L2 {d:.2e}; forward iterations plain median {np.median(its_a)}, "
              f"learned+fp32 median {np.median(its_b[train:])}", flush=True)
        assert d < 1e-4
    finally:
        for s in runs:
            s.close()
