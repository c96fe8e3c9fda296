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
