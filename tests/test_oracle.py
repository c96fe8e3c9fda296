"""CPU pins of the oracle (no GPU).

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned by what the mathematics fixes: adjoint identities of the DEC
operators read off CX/grid3D.cu, the vector-Laplacian identity, symmetry, dense
solves, closed forms, finite differences (the reference's own check,
FT/FluidTopoOpt3D.cpp:206-223) and invariants.
"""
import numpy as np
import pytest

from conftest import rel
from helpers import oracle_desc, setup
from stokes_to_amd.problems import channel_problem, grf_density


@pytest.mark.parametrize("dim,shape", [(3, (4, 5, 3)), (2, (5, 4, 1))])
def test_dec_adjoints_and_exactness(orc, dim, shape):
    rng = np.random.default_rng(0)
    n, s = orc.sizes(dim, *shape)
    nc = s[0]
    c = rng.standard_normal(nc)
    f = rng.standard_normal(n - nc)
    e = rng.standard_normal(orc.edge_count(dim, *shape))
    # Cod0 = -D2^T  (CX/grid3D.cu:401-456 vs :330-390)
    assert abs(c @ orc.d2(dim, shape, f) + orc.cod0(dim, shape, c) @ f) < 1e-12 * np.abs(f).sum()
    # D1 = Cod1^T   (CX/grid3D.cu:219-318 vs :469-568)
    assert abs(e @ orc.cod1(dim, shape, f) - orc.d1(dim, shape, e) @ f) < 1e-12 * np.abs(f).sum()
    # div curl = 0, curl grad = 0 exactly under zero extension
    assert np.abs(orc.d2(dim, shape, orc.d1(dim, shape, e))).max() < 1e-13
    assert np.abs(orc.cod1(dim, shape, orc.cod0(dim, shape, c))).max() < 1e-13


def test_divergence_gradient_closed_forms(orc):
    # u_x = i on x-faces -> div = 1 in every cell; p = i -> grad_x = 1 inside
    shape = (4, 3, 2)
    n, s = orc.sizes(3, *shape)
    nc = s[0]
    f = np.zeros(n - nc)
    fx = f[: s[1]].reshape(5, 3, 2)
    fx[...] = np.arange(5)[:, None, None]
    assert np.allclose(orc.d2(3, shape, f), 1.0)
    p = np.repeat(np.arange(4.0), 6)
    g = orc.cod0(3, shape, p)[: s[1]].reshape(5, 3, 2)
    assert np.allclose(g[1:4], 1.0) and np.allclose(g[0], 0.0) and np.allclose(g[4], -3.0)


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_operator_symmetric_and_fixed_rows(orc, dim, n):
    P, rho = setup(dim, n)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(fixed.size), rng.standard_normal(fixed.size)
    Ax = orc.apply(dim, P.shape, fixed, vol, pen, x)
    Ay = orc.apply(dim, P.shape, fixed, vol, pen, y)
    assert abs(x @ Ay - y @ Ax) < 1e-12 * abs(x @ Ay)
    fx = fixed.astype(bool)
    assert np.array_equal(Ax[fx], x[fx])  # definite fix (ST/StokeFlowMapping3D.cu:301-305)


def test_viscous_block_is_vector_laplacian(orc):
    # grad-div + curl-curl = -Laplacian (7-point) on interior faces, all DOFs free
    dim, shape = 3, (6, 6, 6)
    n, s = orc.sizes(dim, *shape)
    nc = s[0]
    fixed = np.zeros(n, np.uint8)
    vol = np.ones(n)
    vol[:nc] = 0
    pen = np.zeros(n)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(n)
    x[:nc] = 0
    y = orc.energy_op(dim, shape, fixed, vol, pen, x)
    ux = x[nc: nc + s[1]].reshape(7, 6, 6)
    yx = y[nc: nc + s[1]].reshape(7, 6, 6)
    lap = 6 * ux[1:-1, 1:-1, 1:-1] - (ux[2:, 1:-1, 1:-1] + ux[:-2, 1:-1, 1:-1] + ux[1:-1, 2:, 1:-1]
                                       + ux[1:-1, :-2, 1:-1] + ux[1:-1, 1:-1, 2:] + ux[1:-1, 1:-1, :-2])
    assert np.allclose(yx[1:-1, 1:-1, 1:-1], lap, atol=1e-12)


def test_2d_reference_literal_agrees_on_free_subspace(orc):
    # ST/StokeFlowMapping.cu:30-139 (y-sign flips, unmasked curl input, no fix)
    P, rho = setup(2, 16)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    x = np.random.default_rng(3).standard_normal(fixed.size)
    fx = fixed.astype(bool)
    x[fx] = 0
    a = orc.apply(2, P.shape, fixed, vol, pen, x)
    r = orc.apply_ref2d(P.shape, fixed, vol, pen, x)
    assert np.allclose(a[~fx], r[~fx], atol=1e-12)
    assert np.allclose(r[fx], 0.0)


def _dense(orc, dim, shape, fixed, vol, pen):
    n = fixed.size
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        A[:, j] = orc.apply(dim, shape, fixed, vol, pen, e)
    return A


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 8)])
def test_mg_fgmres_matches_dense_solve(orc, dim, n):
    P, rho = setup(dim, n, sigma=1.0)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    A = _dense(orc, dim, P.shape, fixed, vol, pen)
    b = orc.lift_rhs(dim, P.shape, fixed, P.dirichlet)
    xd = np.linalg.solve(A, b)
    mg = orc.MG(dim, P.shape, fixed, pen)
    x, it, rr, _ = mg.fgmres(b, rtol=1e-13)
    assert rel(x, xd) < 1e-10
    # the lifted rhs reproduces the Dirichlet data on fixed faces
    fx = fixed.astype(bool)
    assert np.allclose(x[fx], P.dirichlet[fx])


def test_lift_equals_unmasked_operator(orc):
    # b_free = -(A with unmasked columns) xd: build it from a copy of the problem in which
    # the Dirichlet faces are free, and compare on the rows that are free in both
    P, rho = setup(3, 6)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    nd, s = orc.sizes(3, *P.shape)
    xd = P.dirichlet
    b = orc.lift_rhs(3, P.shape, fixed, xd)
    fixed2 = fixed.copy()
    fx = fixed.astype(bool) & (xd != 0)
    fixed2[fx] = 0
    vol2 = np.where(np.arange(nd) >= s[0], 1.0 - fixed2, 0.0)
    y = orc.apply(3, P.shape, fixed2, vol2, np.zeros(nd), xd)
    free = ~fixed.astype(bool)
    assert np.allclose(b[free], -y[free], atol=1e-12)


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 8)])
def test_sensitivity_matches_finite_differences(orc, dim, n):
    P = channel_problem(dim, n)
    rho = grf_density(dim, n, mean=0.4, sigma=1.0, seed=5)
    am = P.alpha.array()
    nd, s = orc.sizes(dim, *P.shape)
    nc = s[0]
    b = orc.lift_rhs(dim, P.shape, P.fixed, P.dirichlet)

    def J_of(r):
        fixed, vol, pen = oracle_desc(orc, P, r)
        mg = orc.MG(dim, P.shape, fixed, pen)
        x, it, rr, _ = mg.fgmres(b, rtol=1e-14, maxit=300)
        return orc.objective(dim, P.shape, fixed, vol, pen, x), x, mg, (fixed, vol, pen)

    J0, x, mg, (fixed, vol, pen) = J_of(rho)
    gx = orc.energy_op(dim, P.shape, fixed, vol, pen, x)
    if dim == 2:
        # the reference's 2D G (face Laplacian of the unmasked velocity, masked rows,
        # FT/FluidEnergyGrad.cu:30-61) is not symmetric where a free face neighbours a
        # Dirichlet face, so dJ/dx is (G + G^T) x / 2, and G^T x = G (masked x) on free
        # rows.  The reference's adjoint right-hand side G x (FT/FluidTopoOpt.cpp:170-199)
        # differs from it by L u_D / 2 on those rows only.
        xm = np.where(fixed.astype(bool), 0.0, x)
        gxm = orc.energy_op(dim, P.shape, fixed, vol, pen, xm)
        free = ~fixed.astype(bool)
        gs = np.where(free, 0.5 * (gx + gxm), gx)
        d = gx - gs
        nearD = _faces_next_to_dirichlet(P)
        assert np.abs(d[free & ~nearD]).max() < 1e-13 * np.abs(gx).max()
        assert np.abs(d[free & nearD]).max() > 1e-6 * np.abs(gx).max()
        gx = gs
    lam, _, _, _ = mg.fgmres(gx, rtol=1e-14, maxit=300)  # A symmetric: adjoint = forward operator
    dj = orc.sensitivity(dim, P.shape, rho, am, P.fixed[nc:], x[nc:], lam[nc:])
    rng = np.random.default_rng(0)
    for c in rng.choice(rho.size, 4, replace=False):
        h = 1e-6 * max(rho[c], 1e-3)
        rp, rm = rho.copy(), rho.copy()
        rp[c] += h
        rm[c] -= h
        fd = (J_of(rp)[0] - J_of(rm)[0]) / (2 * h)
        assert abs(fd - dj[c]) <= 1e-5 * max(abs(fd), 1e-8 * abs(J0)), (c, fd, dj[c])
    # Taylor remainder is second order along a random direction
    d = rng.standard_normal(rho.size)
    errs = []
    for eps in (1e-3, 5e-4):
        errs.append(abs(J_of(rho + eps * d)[0] - J0 - eps * dj @ d))
    assert errs[1] < 0.3 * errs[0]


def _faces_next_to_dirichlet(P):
    """2D: free faces whose 5-point stencil (same component) touches a fixed face with a
    nonzero Dirichlet value."""
    n = P.n
    nc = n * n
    near = np.zeros(P.fixed.size, bool)
    o = nc
    for A, B in ((n + 1, n), (n, n + 1)):
        m = A * B
        dv = (P.dirichlet[o:o + m] != 0).reshape(A, B)
        hit = np.zeros((A, B), bool)
        hit[1:] |= dv[:-1]
        hit[:-1] |= dv[1:]
        hit[:, 1:] |= dv[:, :-1]
        hit[:, :-1] |= dv[:, 1:]
        near[o:o + m] = hit.reshape(-1)
        o += m
    return near


def _tridiag(m):
    return 2.0 * np.eye(m) - np.eye(m, k=1) - np.eye(m, k=-1)


def test_2d_objective_is_the_kronecker_face_laplacian(orc):
    """The reference's 2D objective operator (FT/FluidEnergyGrad.cu:30-61): per velocity
    component the 5-point Laplacian with zero extension, i.e. the Kronecker sum
    T_A (x) I_B + I_A (x) T_B of 1D second differences over the face array (A x B),
    applied to the unmasked velocity, rows of fixed faces zeroed, plus pen * x."""
    n = 6
    P, rho = setup(2, n, sigma=1.0)
    fixed, vol, pen = oracle_desc(orc, P, rho)
    pen[: n * n] = 0.0
    nd = fixed.size
    G = np.zeros((nd, nd))
    e = np.zeros(nd)
    for j in range(nd):
        e[:] = 0
        e[j] = 1
        G[:, j] = orc.energy_op_ref2d(P.shape, fixed, pen, e)
    assert np.array_equal(G, np.column_stack([orc.energy_op(2, P.shape, fixed, vol, pen, c) for c in np.eye(nd)]))
    ref = np.zeros((nd, nd))
    o = n * n
    for A, B in ((n + 1, n), (n, n + 1)):
        m = A * B
        ref[o:o + m, o:o + m] = np.kron(_tridiag(A), np.eye(B)) + np.kron(np.eye(A), _tridiag(B))
        o += m
    ref[fixed.astype(bool), :] = 0.0       # masked rows, unmasked columns
    ref += np.diag(pen)
    assert np.abs(G - ref).max() < 1e-14
    # a fixed face with a value feeds its free neighbours (the columns are not masked)
    fx = np.flatnonzero(fixed[n * n:]) + n * n
    assert np.abs(G[:, fx]).sum() > 0 and np.all(G[fx][:, ~fixed.astype(bool)] == 0)


def test_2d_objective_quadratic_field_has_constant_laplacian(orc):
    """u_x(i, j) = a i^2 + b j^2 + c i j + d: 4u - sum of neighbours = -(2a + 2b) at every
    interior free face (second differences of a quadratic are constant); u_y likewise."""
    n = 8
    P = channel_problem(2, n)
    nc = n * n
    fixed = np.zeros(P.fixed.size, np.uint8)
    pen = np.zeros(P.fixed.size)
    x = np.zeros(P.fixed.size)
    a, b, c, d = 0.7, -1.3, 0.4, 2.0
    I, J = np.meshgrid(np.arange(n + 1), np.arange(n), indexing="ij")
    x[nc:nc + (n + 1) * n] = (a * I ** 2 + b * J ** 2 + c * I * J + d).reshape(-1)
    I2, J2 = np.meshgrid(np.arange(n), np.arange(n + 1), indexing="ij")
    x[nc + (n + 1) * n:] = (b * I2 ** 2 + a * J2 ** 2 - c * I2 * J2).reshape(-1)
    y = orc.energy_op_ref2d(P.shape, fixed, pen, x)
    yx = y[nc:nc + (n + 1) * n].reshape(n + 1, n)
    yy = y[nc + (n + 1) * n:].reshape(n, n + 1)
    assert np.allclose(yx[1:-1, 1:-1], -(2 * a + 2 * b), atol=1e-12)
    assert np.allclose(yy[1:-1, 1:-1], -(2 * b + 2 * a), atol=1e-12)
    # J = x.Gx/2 >= 0 for a velocity that vanishes on fixed faces (G is then a Gram form)
    P2, rho = setup(2, n, sigma=1.0)
    f2, v2, p2 = oracle_desc(orc, P2, rho)
    z = np.random.default_rng(9).standard_normal(f2.size) * (1 - f2)
    assert orc.objective(2, P2.shape, f2, v2, p2, z) > 0


def _cell_block(dim, n, i, j, k):
    """ABI indices of a cell's block: its pressure and its 2*dim faces."""
    nz = n if dim == 3 else 1
    nc = n * n * nz
    dims = [(n + 1, n, nz), (n, n + 1, nz), (n, n, nz + 1)][:dim]
    offs = [nc]
    for sh in dims[:-1]:
        offs.append(offs[-1] + int(np.prod(sh)))
    idx = [(i * n + j) * nz + k]
    for d, (A, B, C) in enumerate(dims):
        for s in (0, 1):
            f = [i, j, k]
            f[d] += s
            idx.append(offs[d] + (f[0] * B + f[1]) * C + f[2])
    return idx


@pytest.mark.parametrize("dim,n", [(3, 4), (3, 8), (2, 8), (2, 16)])
def test_smoother_colour_step_is_the_dense_probed_block_solve(orc, dim, n):
    """The reference smooths by probing A for each colour cell's block [p_c + its faces]
    and solving it with batched LU (ST/StokeLike.cu:41-96, CX/BatchInverse.cuh:49-102).
    Here the blocks are extracted from the dense operator (unit-vector probes), solved
    with numpy.linalg.solve, and one colour step x += omega A_BB^-1 (b - A x)_B is
    compared with the oracle's closed-form step.  The grid has fixed walls, the inflow,
    fixed cells, free boundary faces (outflow) and penalties from 0 to 1e4."""
    P = channel_problem(dim, n)
    nz = n if dim == 3 else 1
    nc = n * n * nz
    fixed = P.fixed.copy()
    rng = np.random.default_rng(n + dim)
    fixed[rng.choice(nc, max(2, nc // 16), replace=False)] = 1           # fixed cells
    nd = fixed.size
    vol = np.where(np.arange(nd) >= nc, 1.0 - fixed, 0.0)
    pen = np.zeros(nd)
    pen[nc:] = 10.0 ** rng.uniform(-2, 4, nd - nc) * rng.integers(0, 2, nd - nc)
    pen[fixed.astype(bool)] = 0.0
    A = np.zeros((nd, nd))
    e = np.zeros(nd)
    for j in range(nd):
        e[:] = 0
        e[j] = 1
        A[:, j] = orc.apply(dim, P.shape, fixed, vol, pen, e)
    b = rng.standard_normal(nd)
    xin = rng.standard_normal(nd)
    r = b - A @ xin
    for color in (0, 1):
        for omega in (1.0, 0.7):
            got = orc.smooth_color(dim, P.shape, fixed, pen, b, xin, color, omega)
            want = xin.copy()
            for i in range(n):
                for j in range(n):
                    for k in range(nz):
                        if (i + j + k) % 2 != color:
                            continue
                        B = _cell_block(dim, n, i, j, k)
                        want[B] += omega * np.linalg.solve(A[np.ix_(B, B)], r[B])
            assert np.abs(got - want).max() < 1e-11 * max(1.0, np.abs(want).max()), (color, omega)


@pytest.mark.parametrize("R", [1.0, 1.5])
def test_oc_exponent_closed_form(orc, R):
    """OC (Andreassen et al. 2011): rho_new = rho (-dc / (lambda dV))^eta where no bound is
    active, so (rho_new / rho) / (-dc / dV)^eta is one constant (lambda^-eta) over the
    cells; a wrong exponent or a dropped dV leaves a spread of the size of the dc spread.
    dV = H^T (1/(n W)) is the gradient of the filtered mean (uniform 1/n without a filter)."""
    dim, shape = 3, (6, 6, 6)
    n = 216
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.4, 0.6, n)
    dc = -rng.uniform(0.8, 1.2, n)
    dv = orc.filt(dim, shape, R, np.full(n, 1.0 / n), transpose=True)
    if R == 1.0:
        assert np.allclose(dv, 1.0 / n, rtol=1e-15)
    vf = orc.filt(dim, shape, R, rho).mean()
    for eta in (0.5, 0.3, 1.0):
        rn, ch = orc.oc_update(dim, shape, R, rho, dc, vf, move=0.2, eta=eta, rho_min=1e-3)
        assert np.all(rn > rho - 0.2) and np.all(rn < np.minimum(rho + 0.2, 1.0))   # no bound active
        c = (rn / rho) / (-dc / dv) ** eta
        assert np.ptp(c) < 1e-10 * c.mean(), eta
        # the multiplier meets the filtered volume
        assert abs(orc.filt(dim, shape, R, rn).mean() - vf) < 1e-10


@pytest.mark.parametrize("dim,shape", [(3, (5, 4, 6)), (2, (7, 5, 1))])
def test_filter_brute_force_and_adjoint(orc, dim, shape):
    R = 1.5
    n = int(np.prod(shape))
    coords = np.array(np.unravel_index(np.arange(n), shape)).T.astype(float)
    d = np.sqrt(((coords[:, None, :] - coords[None, :, :]) ** 2).sum(-1))
    H = np.maximum(0.0, R - d)
    F = H / H.sum(1, keepdims=True)
    x = np.random.default_rng(4).random(n)
    assert np.allclose(orc.filt(dim, shape, R, x), F @ x, atol=1e-14)
    assert np.allclose(orc.filt(dim, shape, R, x, transpose=True), F.T @ x, atol=1e-14)
    assert np.allclose(orc.filt(dim, shape, R, np.ones(n)), 1.0)


def test_oc_volume_and_bounds(orc):
    dim, shape = 3, (8, 8, 8)
    rng = np.random.default_rng(6)
    rho = rng.uniform(0.1, 0.9, 512)
    dc = -rng.random(512)
    rn, ch = orc.oc_update(dim, shape, 1.5, rho, dc, volfrac=0.4, move=0.2)
    vol = orc.filt(dim, shape, 1.5, rn).mean()
    assert abs(vol - 0.4) < 1e-10
    assert np.all(rn >= np.maximum(1e-3, rho - 0.2) - 1e-15) and np.all(rn <= np.minimum(1.0, rho + 0.2) + 1e-15)
    assert abs(ch - np.abs(rn - rho).max()) < 1e-15


@pytest.mark.parametrize("dim", [3, 2])
def test_restriction_is_scaled_prolongation_transpose(orc, dim):
    n = 8
    shape = (n, n, n if dim == 3 else 1)
    P = channel_problem(dim, n)
    ffix = P.fixed
    cfix, _ = orc.coarsen(dim, shape, ffix, np.zeros(ffix.size))
    rng = np.random.default_rng(7)
    r = rng.standard_normal(ffix.size) * (1 - ffix)
    e = rng.standard_normal(cfix.size) * (1 - cfix)
    nf, sf = orc.sizes(dim, *shape)
    ncs = orc.sizes(dim, n // 2, n // 2, n // 2 if dim == 3 else 1)[1]
    Rr = orc.restrict(dim, shape, cfix, 0, r)
    Pe = orc.prolong_add(dim, shape, ffix, 0, e, np.zeros(ffix.size))
    cu, cp = 4.0 / 2 ** dim, 2.0 / 2 ** dim
    # velocity part: <R r, e> = cu <r, P e>; pressure: R_p = cp P_p^T, prolong = P_p / 2
    assert abs(Rr[ncs[0]:] @ e[ncs[0]:] - cu * (r[sf[0]:] @ Pe[sf[0]:])) < 1e-11
    assert abs(Rr[: ncs[0]] @ e[: ncs[0]] - 2 * cp * (r[: sf[0]] @ Pe[: sf[0]])) < 1e-11


def test_mg_is_h_independent(orc):
    its = []
    for n in (8, 16):
        P, rho = setup(3, n)
        fixed, vol, pen = oracle_desc(orc, P, rho)
        b = orc.lift_rhs(3, P.shape, fixed, P.dirichlet)
        mg = orc.MG(3, P.shape, fixed, pen)
        x, it, rr, _ = mg.fgmres(b, rtol=1e-8)
        assert rr <= 1e-8
        its.append(it)
    assert max(its) <= 12 and abs(its[1] - its[0]) <= 3


def test_grf_density_seeded_mean():
    a = grf_density(3, 16, mean=0.3, sigma=2.0, seed=12345)
    b = grf_density(3, 16, mean=0.3, sigma=2.0, seed=12345)
    assert np.array_equal(a, b)
    assert abs(a.mean() - 0.3) < 1e-6 and a.min() >= 1e-3 and a.max() <= 1.0


# ---------------------------------------------------------------- learned part

@pytest.mark.parametrize("R,ntap,cin,cout", [(5, 27, 3, 4), (4, 1, 5, 2)])
def test_conv3d_matches_library_correlation(orc, R, ntap, cin, cout):
    """Forward convolution against scipy.ndimage.correlate (zero padding); the
    input-gradient is its adjoint; the weight-gradient is the derivative of the
    (linear-in-W) pairing <conv(x; W), g>; the bias gradient is the sum of g."""
    from scipy.ndimage import correlate
    rng = np.random.default_rng(R + cin)
    nv = R ** 3
    x = rng.standard_normal(nv * cin)
    w = rng.standard_normal(ntap * cin * cout)
    b = rng.standard_normal(cout)
    y = orc.conv3d(R, ntap, cin, cout, x, w, b).reshape(R, R, R, cout)
    xs = x.reshape(R, R, R, cin)
    k = 3 if ntap == 27 else 1
    W = w.reshape(k, k, k, cin, cout)   # tap t = (a+1)*9 + (b+1)*3 + (c+1) <-> offset (a, b, c)
    ref = np.zeros((R, R, R, cout))
    for co in range(cout):
        ref[..., co] = b[co] + sum(correlate(xs[..., ci], W[..., ci, co], mode="constant", cval=0.0)
                                   for ci in range(cin))
    assert np.abs(y - ref).max() < 1e-12 * np.abs(ref).max()
    yr = orc.conv3d(R, ntap, cin, cout, x, w, b, relu=True).reshape(y.shape)
    assert np.array_equal(yr, np.maximum(y, 0.0))
    g = rng.standard_normal(nv * cout)
    y0 = orc.conv3d(R, ntap, cin, cout, x, w, None)
    gi = orc.conv3d_bwd_data(R, ntap, cin, cout, g, w)
    assert abs(y0 @ g - x @ gi) < 1e-12 * np.abs(g).sum() * np.abs(y0).max()
    gw, gb = orc.conv3d_bwd_weight(R, ntap, cin, cout, x, g)
    w2 = rng.standard_normal(w.size)
    assert abs(orc.conv3d(R, ntap, cin, cout, x, w2, None) @ g - w2 @ gw) < 1e-11 * np.abs(gw).sum()
    assert np.allclose(gb, g.reshape(nv, cout).sum(axis=0), rtol=1e-13, atol=1e-12)


def test_network_pool_and_upsample_adjoints():
    import oracle.net as N
    rng = np.random.default_rng(4)
    R, C = 4, 3
    x = rng.standard_normal(R ** 3 * C)
    g = rng.standard_normal((R // 2) ** 3 * C)
    assert abs(N.pool(x, R, C) @ g - x @ N.pool_bwd(g, R, C)) < 1e-12
    assert abs(N.up(g, R // 2, C) @ x - g @ N.up_bwd(x, R, C)) < 1e-12
    # pooling a constant field keeps it; upsampling then pooling is the identity
    assert np.allclose(N.pool(np.ones(R ** 3 * C), R, C), 1.0)
    assert np.allclose(N.pool(N.up(g, R // 2, C), R, C), g)


def test_network_gradient_matches_finite_differences():
    """Loss gradient of the oracle network (plain FP64) by central differences,
    coordinate-wise on a sample of parameters and along a random direction."""
    import oracle.net as N
    R = 4
    rng = np.random.default_rng(0)
    npar = sum(nt * ci * co + co for ci, co, nt, _, _ in N.SPEC)
    p = rng.standard_normal(npar) * 0.15
    feat = rng.standard_normal(R ** 3 * 4)
    tgt = rng.standard_normal(R ** 3 * 4)
    L, g, _ = N.loss_and_grad(p, feat, tgt, R)
    h = 1e-5
    f = lambda q: N.loss_and_grad(q, feat, tgt, R)[0]
    # one coordinate from every layer's weights and biases
    idx, o = [], 0
    for ci, co, nt, _, _ in N.SPEC:
        idx += [o + int(rng.integers(nt * ci * co)), o + nt * ci * co + int(rng.integers(co))]
        o += nt * ci * co + co
    for i in idx:
        e = np.zeros(npar)
        e[i] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g[i]), 1e-3 * np.abs(g).max()), i
    d = rng.standard_normal(npar)
    d /= np.linalg.norm(d)
    fd = (f(p + h * d) - f(p - h * d)) / (2 * h)
    assert abs(fd - g @ d) <= 1e-6 * abs(g @ d)


def test_bf16_rounding_known_values():
    import oracle.net as N
    u = 2.0 ** -8
    v = np.array([1.0, 1 + u, 1 + 3 * u, 1 + u + u / 16, -2.5, 3.0e38, 1e-40])
    r = N.bf16_round(v)
    # ties go to the even mantissa; above the tie rounds up; sign and range kept
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 4 * u and r[3] == 1 + 2 * u
    assert r[4] == -2.5 and abs(r[5] / 3.0e38 - 1) < 2 ** -8
    # BF16 keeps FP32's exponent range: 1e-40 is an FP32 subnormal that BF16 keeps approximately
    assert abs(r[6] / 1e-40 - 1) < 2 ** -2
    # idempotent
    assert np.array_equal(N.bf16_round(r), r)


# ---------------------------------------------------------------- MMA

def _mma_problem(n, seed):
    rng = np.random.default_rng(seed)
    af = rng.uniform(0.5, 2.0, n)
    bf = rng.uniform(-1.0, 1.0, n)
    return lambda x: 2.0 * af * (x - 0.3) + bf * np.cos(3.0 * x)


@pytest.fixture(scope="session")
def mma_ref(orc):
    """The reference's own MMASolver.cpp, compiled where it lies (oracle/Makefile `ref`)."""
    import os
    import subprocess
    if not orc.mma_ref_available():
        if not os.path.isdir("/root/reference/simplex/ext/mma"):
            pytest.skip("reference sources not present (MMA compared against the compiled reference here only)")
        subprocess.check_call(["make", "-s", "-C", os.path.dirname(orc.__file__), "ref"])
    return orc.MMARef


@pytest.mark.parametrize("n,seed", [(500, 0), (4096, 1)])
def test_mma_restatement_matches_compiled_reference(orc, mma_ref, n, seed):
    """The oracle's MMA against the reference's MMASolver (one volume constraint,
    outer move box of OptimizerMma.h:157-164): identical iterates, bit for bit."""
    grad = _mma_problem(n, seed)
    xa = np.full(n, 0.5)
    xb = xa.copy()
    A, B = orc.MMA(n), mma_ref(n)
    for it in range(25):
        for x, M in ((xa, A), (xb, B)):
            g, dg = orc.volume_constraint(x, 0.4)
            lo, hi = orc.mma_bounds(x, 0.2, 0.01, 1.0)
            M.update(x, grad(x), g, dg, lo, hi)
        assert np.array_equal(xa, xb), (it, np.abs(xa - xb).max())


def test_mma_step_is_feasible_and_inside_the_box(orc):
    """Properties fixed by the method: every iterate stays inside the move box and
    the (linear) volume constraint ends up satisfied once it is active."""
    n = 1000
    grad = lambda x: -np.ones(n) - 0.1 * x      # objective prefers large x -> constraint active
    x = np.full(n, 0.2)
    M = orc.MMA(n)
    for it in range(30):
        g, dg = orc.volume_constraint(x, 0.4)
        lo, hi = orc.mma_bounds(x, 0.2, 0.01, 1.0)
        xo = x.copy()
        M.update(x, grad(x), g, dg, lo, hi)
        assert np.all(x >= lo - 1e-15) and np.all(x <= hi + 1e-15)
        assert np.abs(x - xo).max() <= 0.2 + 1e-12
    assert abs(x.mean() - 0.4) < 1e-6


def test_cpu_configuration_2d_64_to_run(orc):
    """The CPU-runnable configuration end to end on the oracle: 2D 64^2 channel,
    initial density 0.15, volume fraction 0.15, filter radius 1.5, 10 OC iterations.
    Invariants: the (filtered) volume is met at every step, designs stay in
    [rho_min, 1], the change respects the move limit, and the dissipation drops."""
    n, vf = 64, 0.15
    P = channel_problem(2, n)
    rho = np.full(n * n, vf)
    Js = []
    for it in range(10):
        rn, J, ch, x, lam = orc.to_step(P, rho, volfrac=vf, rtol=1e-10)
        rf = orc.filt(2, P.shape, 1.5, rn)
        assert abs(rf.mean() - vf) < 1e-8
        assert rn.min() >= 1e-3 - 1e-15 and rn.max() <= 1.0 + 1e-15
        assert ch <= 0.2 + 1e-12 and abs(ch - np.abs(rn - rho).max()) < 1e-15
        Js.append(J)
        rho = rn
    assert Js[-1] < 0.5 * Js[0]


def test_coarse_alpha_is_the_midpoint_of_the_wiener_bounds(orc):
    """Coarse face penalty = 4 x (weighted arithmetic + weighted harmonic mean)/2 of the
    child faces: a uniform fine alpha a gives 4a; any child with alpha 0 (fluid) makes
    the harmonic mean 0, so the coarse value is half the scaled arithmetic mean."""
    dim, shape = 3, (8, 8, 8)
    n, s = orc.sizes(dim, *shape)
    nc = s[0]
    fixed = np.zeros(n, np.uint8)
    pen = np.zeros(n)
    pen[nc:] = 0.7
    cfix, cpen = orc.coarsen(dim, shape, fixed, pen)
    ncc = orc.sizes(dim, 4, 4, 4)[1][0]
    assert np.allclose(cpen[ncc:], 4 * 0.7, rtol=1e-14)
    rng = np.random.default_rng(3)
    pen[nc:] = rng.uniform(0.1, 1.0, n - nc)
    pen[nc + 5] = 0.0                                   # one fluid child face
    _, cpen2 = orc.coarsen(dim, shape, fixed, pen)
    pen_nz = pen.copy()
    pen_nz[nc + 5] = 1e-300
    _, cpen3 = orc.coarsen(dim, shape, fixed, pen_nz)  # harmonic mean -> ~0 as well
    assert np.allclose(cpen2, cpen3, rtol=1e-12, atol=1e-300)
    # one coarse x-face by hand: children (2I+o, 2J+u, 2K+v), o in {-1,0,1} with weights
    # 1/2, 1, 1/2, u, v in {0,1}
    I, J, K = 1, 2, 1
    fx = pen[nc:nc + 9 * 8 * 8].reshape(9, 8, 8)
    w, a = [], []
    for o, wo in ((-1, 0.5), (0, 1.0), (1, 0.5)):
        for u in (0, 1):
            for v in (0, 1):
                w.append(wo)
                a.append(fx[2 * I + o, 2 * J + u, 2 * K + v])
    w, a = np.array(w), np.array(a)
    expect = 4.0 * 0.5 * ((w * a).sum() / w.sum() + w.sum() / (w / a).sum())
    got = cpen2[ncc:ncc + 5 * 4 * 4].reshape(5, 4, 4)[I, J, K]
    assert abs(got - expect) < 1e-13 * expect
