"""What a predictor of the coarse state can gain at best (CPU, oracle).

The learned part predicts the solution on multigrid level 2 (n/4 cells per axis) and
enters the fine level through two prolongations (the initial guess of
ST_INIT_LEARNED).  Here the network is replaced by the exact answer -- the converged
fine solution restricted to level L the way the training targets are formed (velocity:
mean of the coincident child faces, pressure: 2 x the child mean) -- and the FGMRES +
V(2,2) iterations to a 1e-8 true residual are counted.  This bounds what any
level-L predictor can save, whatever its training:
  * cold start (no previous solution): the exact level-2 state saves at most one
    iteration against the Dirichlet start, and its residual is larger (the prolonged
    state jumps across coarse cells), so a residual-based choice never takes it and a
    minimal-residual coefficient on it is ~0;
  * warm start inside a design sequence (the previous design's solution at hand):
    the exact level-2 *change* saves nothing.
The measured numbers are printed.
"""
import numpy as np

from helpers import oracle_desc, setup
from stokes_to_amd.problems import channel_problem


def _split(v, m):
    nc, nf = m ** 3, (m + 1) * m * m
    shapes = [(m + 1, m, m), (m, m + 1, m), (m, m, m + 1)]
    return v[:nc].reshape(m, m, m), [v[nc + d * nf: nc + (d + 1) * nf].reshape(shapes[d]) for d in range(3)]


def _join(p, fs):
    return np.concatenate([p.reshape(-1)] + [f.reshape(-1) for f in fs])


def _state_restrict(v, m):
    """fine state (m^3) -> coarse state ((m/2)^3), as the network's targets
    (st_net.hip k_state_restrict)."""
    p, (fx, fy, fz) = _split(v, m)
    c = m // 2
    pc = 2.0 * p.reshape(c, 2, c, 2, c, 2).mean(axis=(1, 3, 5))
    return _join(pc, [fx[0::2].reshape(c + 1, c, 2, c, 2).mean(axis=(2, 4)),
                      fy[:, 0::2].reshape(c, 2, c + 1, c, 2).mean(axis=(1, 4)),
                      fz[:, :, 0::2].reshape(c, 2, c, 2, c + 1).mean(axis=(1, 3))])


def _coarse_correction(orc, mg, d, n, L):
    """P^L R_s^L d: the part of d a level-L predictor can represent."""
    c, m = d, n
    for _ in range(L):
        c = _state_restrict(c, m)
        m //= 2
    for l in reversed(range(L)):
        ff = mg.level_desc(l)[0]
        c = orc.prolong_add(3, (2 * m,) * 3, ff, 0, c, np.zeros(ff.size))
        m *= 2
    return c


def test_exact_coarse_state_saves_at_most_one_iteration_from_a_cold_start(orc):
    n = 32
    P, rho = setup(3, n, mean=0.3)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    b = orc.lift_rhs(3, P.shape, fixed, P.dirichlet)
    mg = orc.MG(3, P.shape, fixed, pen)
    xs, _, rr, _ = mg.fgmres(b, x0=P.dirichlet, rtol=1e-13)
    A = lambda v: orc.apply(3, P.shape, fixed, vol, pen, v)
    its = lambda x0: mg.fgmres(b, x0=x0, rtol=1e-8)[1]
    res = lambda x: np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    it_d = its(P.dirichlet)
    out = {}
    for L in (1, 2):
        d = _coarse_correction(orc, mg, xs - P.dirichlet, n, L)
        x1 = P.dirichlet + d
        r0 = b - A(P.dirichlet)
        Ad = A(d)
        cmr = (r0 @ Ad) / (Ad @ Ad)
        out[L] = (its(x1), res(x1), cmr)
        print(f"{n}^3 cold start: x_D {it_d} it (r0 {res(P.dirichlet):.2e}); exact level-{L} state {out[L][0]} it "
              f"(r0 {out[L][1]:.2e}), minimal-residual coefficient {cmr:.3f}")
    it2, r2, c2 = out[2]
    assert it_d - 2 < it2 <= it_d              # at most one iteration saved
    assert r2 > res(P.dirichlet)               # its residual is larger than the Dirichlet start's
    assert abs(c2) < 0.05                      # and the residual sees almost nothing of it
    assert out[1][0] >= it_d - 2               # one level finer: at most two


def test_exact_coarse_change_saves_nothing_in_a_design_sequence(orc):
    n, vf = 32, 0.25
    P = channel_problem(3, n)
    rho = np.full(n ** 3, vf)
    x_prev = None
    saved = []
    for it in range(6):
        rn, J, ch, x, lam = orc.to_step(P, rho, volfrac=vf, rtol=1e-12, x_prev=x_prev)
        if x_prev is not None and it >= 3:
            # the system of this iteration (design rho) from the previous solution, and
            # from the previous solution plus the exact level-2 change
            rhof = orc.filt(3, P.shape, 1.5, rho)
            fixed, vol, pen = oracle_desc(orc, P, rhof)
            b = orc.lift_rhs(3, P.shape, fixed, P.dirichlet)
            mg = orc.MG(3, P.shape, fixed, pen)
            its = lambda x0: mg.fgmres(b, x0=x0, rtol=1e-8)[1]
            d = _coarse_correction(orc, mg, x - x_prev, n, 2)
            a, c = its(x_prev), its(x_prev + d)
            saved.append(a - c)
            print(f"{n}^3 design iteration {it}: previous solution {a} it, plus the exact level-2 change {c} it")
        x_prev, rho = x, rn
    assert max(saved) <= 1 and sum(saved) <= 1
