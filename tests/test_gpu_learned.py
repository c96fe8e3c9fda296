"""Learned coarse-state predictor on the GPU against the oracle.

The network (3-level 3D U-Net, BF16 weights and activations, FP32 accumulation on
MFMA) is checked layer by layer and as a whole:
  * each convolution kind (3x3x3 and 1x1x1; forward, input gradient, weight and
    bias gradient) against the oracle's FP64 convolution on BF16-representable
    inputs, so the only difference left is FP32 accumulation order (bar 2e-5);
  * the gradient of the training loss w.r.t. every parameter against the oracle
    network evaluated with the GPU's BF16 storage points (bar 1e-3 relative L2),
    and against the plain FP64 oracle network (bar 0.3: BF16 storage noise);
  * training on solver samples lowers the loss;
  * what the trained network gains: a lower initial residual than the zero start on
    an unseen design, and never a worse start than the previous solution in a
    design sequence (the residual gate).
The oracle network itself is pinned by finite differences in test_oracle.py.
"""
import numpy as np
import pytest

import oracle.net as N
from conftest import rel
from helpers import setup, solver_for
from stokes_to_amd.steady import conv_test

pytestmark = pytest.mark.gpu


def bf16(a):
    return N.bf16_round(a)


@pytest.mark.parametrize("R,ntap,cin,cout,relu", [
    (8, 27, 16, 16, 1), (8, 27, 16, 32, 0), (4, 27, 32, 64, 1), (4, 27, 64, 64, 1),
    (8, 27, 96, 32, 1), (8, 27, 48, 16, 1), (6, 27, 32, 32, 0), (8, 1, 16, 16, 0), (16, 27, 16, 32, 1),
    # grids where four 16-voxel blocks share each weight load (>= 512 workgroups)
    (64, 27, 16, 16, 1), (32, 27, 32, 64, 0),
])
def test_conv_layer_matches_oracle(orc, R, ntap, cin, cout, relu):
    rng = np.random.default_rng(R * 1000 + cin + cout)
    nv = R ** 3
    x = bf16(rng.standard_normal(nv * cin))
    w = bf16(rng.standard_normal(ntap * cin * cout) * (2.0 / (ntap * cin)) ** 0.5)
    b = rng.standard_normal(cout).astype(np.float32)
    g = bf16(rng.standard_normal(nv * cout))
    y = conv_test(R, ntap, cin, cout, relu, 0, x, w, b, g)
    yr = orc.conv3d(R, ntap, cin, cout, x, w, b.astype(np.float64), bool(relu))
    assert np.abs(y - yr).max() <= 2e-5 * np.abs(yr).max()
    gi = conv_test(R, ntap, cin, cout, 0, 1, x, w, b, g)
    gir = orc.conv3d_bwd_data(R, ntap, cin, cout, g, w)
    assert np.abs(gi - gir).max() <= 2e-5 * np.abs(gir).max()
    gw, gb = conv_test(R, ntap, cin, cout, 0, 2, x, w, b, g)
    gwr, gbr = orc.conv3d_bwd_weight(R, ntap, cin, cout, x, g)
    assert np.abs(gw - gwr).max() <= 2e-5 * np.abs(gwr).max()
    assert np.abs(gb - gbr).max() <= 2e-5 * np.abs(gbr).max()


def _net_solver(make_solver, n=32):
    P, rho = setup(3, n, mean=0.3)
    s = solver_for(make_solver, P, rho, rtol=1e-8)
    return s, P, rho


def test_network_gradient_matches_oracle(make_solver, orc):
    s, P, rho = _net_solver(make_solver, 32)
    R = 8
    p = s.net_params()
    assert p.size == sum(nt * ci * co + co for ci, co, nt, _, _ in N.SPEC)
    rng = np.random.default_rng(7)
    feat = rng.standard_normal(R ** 3 * 4).astype(np.float32)
    tgt = rng.standard_normal(R ** 3 * 4).astype(np.float32)
    L, g, out = s.net_grad(feat, tgt)
    Lq, gq, outq = N.loss_and_grad(p.astype(np.float64), feat.astype(np.float64), tgt.astype(np.float64), R,
                                   q=N.bf16_round)
    assert abs(L - Lq) <= 1e-4 * Lq
    assert rel(out, outq) < 1e-3
    assert rel(g, gq) < 1e-3
    # every layer's weight and bias gradient, not only the total
    o = 0
    for ci, co, nt, _, _ in N.SPEC:
        nw = nt * ci * co
        assert rel(g[o:o + nw], gq[o:o + nw]) < 2e-3, (ci, co, nt)
        assert rel(g[o + nw:o + nw + co], gq[o + nw:o + nw + co]) < 2e-3, (ci, co, nt)
        o += nw + co
    L64, g64, _ = N.loss_and_grad(p.astype(np.float64), feat.astype(np.float64), tgt.astype(np.float64), R)
    assert abs(L - L64) <= 1e-2 * L64
    assert rel(g, g64) < 0.3


def test_params_roundtrip(make_solver):
    s, P, rho = _net_solver(make_solver, 32)
    p = s.net_params()
    q = (p * 0.5).astype(np.float32)
    s.net_set_params(q)
    assert np.array_equal(s.net_params(), q)


def test_training_lowers_loss(make_solver):
    s, P, rho = _net_solver(make_solver, 32)
    for seed in (1, 2, 3):
        _, r = setup(3, 32, mean=0.3, seed=seed)
        s.set_design(r)
        _, st = s.solve()
        assert st["converged"]
        s.record_sample()
    assert s.num_samples() == 3
    first = s.net_train(1, 1e-3)
    last = s.net_train(150, 1e-3)
    assert np.isfinite(first) and np.isfinite(last)
    assert last < 0.5 * first, (first, last)


def test_trained_network_gains_over_the_zero_start(make_solver):
    """What the trained network gains over the method without it, on a design it was
    not trained on.  The plain method's initial guess carries no coarse-scale
    information (Dirichlet values, zero elsewhere); the network predicts the coarse
    state from the density: its loss on the unseen design must be well below that of
    the zero prediction.  At the fine level the prediction is only used if its true
    FP64 residual beats the plain start (the residual gate), so the learned solve
    never needs more iterations than the plain one."""
    n = 32
    P, _ = setup(3, n, mean=0.3)
    s = solver_for(make_solver, P, None, rtol=1e-8)
    for seed in range(1, 7):
        _, r = setup(3, n, mean=0.3, seed=seed)
        s.set_design(r)
        _, st = s.solve()
        s.record_sample()
    s.net_train(300, 1e-3)
    _, r_new = setup(3, n, mean=0.3, seed=99)
    s.set_design(r_new)
    s.reset_solution()
    _, st_zero = s.solve(init="zero")
    loss, loss_zero = s.net_eval()
    assert loss < 0.5 * loss_zero, (loss, loss_zero)
    s.reset_solution()
    _, st_learned = s.solve(init="learned")
    assert st_learned["converged"]
    assert st_learned["relres0"] <= st_zero["relres0"] * (1 + 1e-12)
    assert st_learned["iters"] <= st_zero["iters"], (st_learned, st_zero)


def test_residual_gate_never_loses_to_the_warm_start(make_solver):
    """In a topology-optimization sequence the learned candidate (previous solution
    plus the predicted coarse change) is taken only when its true residual beats the
    previous solution's: the learned start is never worse than the plain warm start.
    (The two contexts hold different solutions of the previous design, each within
    rtol, hence the small allowance.)"""
    n, vf = 32, 0.25
    from stokes_to_amd.problems import channel_problem
    Pc = channel_problem(3, n)
    a = make_solver(3, n, rtol=1e-8)
    k = make_solver(3, n, rtol=1e-8)
    for s in (a, k):
        s.init_boundary(Pc.fixed)
        s.update_b(dirichlet=Pc.dirichlet)
    a.set_design(np.full(n ** 3, vf))
    for it in range(8):
        rho = a.filter(a.get_design())
        k.set_design(rho)
        _, sk = k.solve(init="keep")
        out = a.to_step(vf, init="learned")
        f = out["fwd"]
        if it >= 1:
            assert f["relres0"] <= sk["relres0"] * (1 + 1e-3) + 1e-7, (it, f["relres0"], sk["relres0"])
        a.record_sample()
        a.net_train(20, 1e-3)
