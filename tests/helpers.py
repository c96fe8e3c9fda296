"""Shared test setup: problems, oracle descriptors, random vectors."""
import numpy as np

from stokes_to_amd.problems import channel_problem, grf_density, dof_sizes


def oracle_desc(orc, P, rho):
    """fixed / vol / pen descriptor arrays of the oracle for problem P and design rho
    (vol = free mask, pen = [0 cells | face alpha])."""
    n, s = orc.sizes(P.dim, *P.shape)
    nc = s[0]
    fa = orc.design_to_alpha(P.dim, P.shape, rho, P.alpha.array(), P.fixed[nc:])
    pen = np.zeros(n)
    pen[nc:] = fa
    vol = np.zeros(n)
    vol[nc:] = 1.0 - P.fixed[nc:]
    return P.fixed.copy(), vol, pen


def setup(dim, n, mean=0.3, seed=12345, sigma=None):
    P = channel_problem(dim, n)
    if sigma is None:
        sigma = max(1.0, 8.0 * n / 64)
    rho = grf_density(dim, n, mean=mean, sigma=sigma, seed=seed)
    return P, rho


def random_dofs(P, seed=0, dirichlet=True):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(P.fixed.size)
    if dirichlet:
        # fixed faces carry the Dirichlet data, as they do in a solve
        x[P.fixed.astype(bool)] = P.dirichlet[P.fixed.astype(bool)]
    return x


def solver_for(make_solver, P, rho=None, **kw):
    s = make_solver(P.dim, P.n, **kw)
    s.init_boundary(P.fixed)
    if rho is not None:
        s.set_design(rho)
    s.update_b(dirichlet=P.dirichlet)
    return s


def sizes(dim, n):
    return dof_sizes(dim, n)
