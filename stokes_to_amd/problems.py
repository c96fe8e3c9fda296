"""Problem setup on the host: boundary data, material model, synthetic densities.

This mirrors what the reference's drivers do in `init_boundary` / `init_penalty`
(FT/FluidTopoOptDriver3D.cpp:99-167, FT/FluidTopoOptDriver.cpp:104-166) and the
design-to-penalty parameters of FT/FluidTopoOpt3D.h:42-46.  All arrays use the
ABI layout: C-order (i, j, k) with k fastest; a dof vector is
[p | u_x | u_y | u_z]; a descriptor array (fixed flags) uses the same order.
"""
from dataclasses import dataclass

import numpy as np


def dof_sizes(dim, n):
    """(cells, faces_x, faces_y, faces_z, total) for an n^dim grid."""
    if dim == 3:
        nc = n ** 3
        nf = (n + 1) * n * n
        return nc, nf, nf, nf, nc + 3 * nf
    nc = n * n
    nf = (n + 1) * n
    return nc, nf, nf, 0, nc + 2 * nf


def face_shapes(dim, n):
    if dim == 3:
        return [(n + 1, n, n), (n, n + 1, n), (n, n, n + 1)]
    return [(n + 1, n, 1), (n, n + 1, 1)]


def cell_shape(dim, n):
    return (n, n, n) if dim == 3 else (n, n, 1)


@dataclass
class AlphaModel:
    """alpha(r) = alpha_max + (alpha_min - alpha_max) * r (1+q)/(r+q), applied to rho^power,
    times `scale` (FT/FluidTopoOpt3D.cpp:110-123).  alpha_min=0 gives the
    alpha_max*q(1-r)/(q+r) form."""
    alpha_max: float = 1e4
    alpha_min: float = 0.0
    q: float = 0.1
    power: float = 1.0
    scale: float = 1.0

    def array(self):
        return np.array([self.alpha_max, self.alpha_min, self.q, self.power, self.scale])

    def alpha(self, r):
        r = np.asarray(r, dtype=np.float64) ** self.power
        return self.scale * (self.alpha_max + (self.alpha_min - self.alpha_max) * r * (1 + self.q) / (r + self.q))


@dataclass
class Problem:
    dim: int
    n: int
    fixed: np.ndarray      # uint8 [cells | faces]
    dirichlet: np.ndarray  # float64 dof vector, nonzero only on fixed faces
    body: np.ndarray       # float64 dof vector of source terms (cell_b | face_b)
    alpha: AlphaModel
    closed: bool = False
    name: str = ""

    @property
    def shape(self):
        return (self.n, self.n, self.n) if self.dim == 3 else (self.n, self.n, 1)

    @property
    def h(self):
        return 1.0 / self.n


def _split(dim, n, vec):
    nc, nfx, nfy, nfz, tot = dof_sizes(dim, n)
    cs = cell_shape(dim, n)
    fs = face_shapes(dim, n)
    out = [vec[:nc].reshape(cs)]
    o = nc
    for s in fs:
        m = int(np.prod(s))
        out.append(vec[o:o + m].reshape(s))
        o += m
    return out


def channel_problem(dim, n, patch=0.25, alpha_max=1e4, q=0.1, peak=1.0):
    """Unit square/cube, parabolic inflow (peak `peak`) on the centred patch of
    the x=0 side, zero-traction outflow patch of the same size on x=1, no-slip
    elsewhere; Brinkman alpha = alpha_max*q(1-r)/(q+r) in the unit-spacing
    system (scale h^2)."""
    nc, nfx, nfy, nfz, tot = dof_sizes(dim, n)
    fixed = np.zeros(tot, np.uint8)
    xd = np.zeros(tot)
    fx, fy, *rest = _split(dim, n, fixed)[1:]
    dx = _split(dim, n, xd)[1]
    c = (np.arange(n) + 0.5) / n
    lo, hi = 0.5 - patch / 2, 0.5 + patch / 2
    inside = (c > lo) & (c < hi)
    prof = np.where(inside, 1.0 - (2.0 * (c - 0.5) / patch) ** 2, 0.0)
    # wall-normal faces are fixed
    fx[0] = 1
    fx[n] = 1
    fy[:, 0] = 1
    fy[:, n] = 1
    if dim == 3:
        fz = rest[0]
        fz[:, :, 0] = 1
        fz[:, :, n] = 1
        out_mask = inside[:, None] & inside[None, :]
        fx[n][out_mask] = 0
        dx[0] = peak * prof[:, None] * prof[None, :]
    else:
        fx[n][inside, 0] = 0
        dx[0][:, 0] = peak * prof
    am = AlphaModel(alpha_max=alpha_max, alpha_min=0.0, q=q, power=1.0, scale=1.0 / n ** 2)
    return Problem(dim=dim, n=n, fixed=fixed, dirichlet=xd, body=np.zeros(tot), alpha=am,
                   closed=False, name=f"channel{dim}d_{n}")


def reference_test0(dim, n):
    """The reference's own demo case, test 0 (`fluid_topo -driver 0|1 -test 0`): a
    closed box (every boundary-normal face fixed, pressure gauge on) driven by
    source/sink terms, with the reference's material model alpha_os = 500 (32/n),
    alpha_us = 1e-3, q = 0.014 (2D) / 0.01 (3D) on rho^3, unit spacing.

    3D (FT/FluidTopoOptDriver3D.cpp:99-167): a source of 600 on the cells (0, j, k)
    and x-faces (1, j, k), j, k in [28, 36); three sinks of 200 on the x = n-1 side
    (j in [10, 18), k in [34, 42)), the z = n-1 side (i in [28, 36), j in [22, 30))
    and the z = 0 side (i in [22, 30), j in [28, 36)).  The reference uses these
    absolute indices at every size (they fit from n = 42); below n = 64 they are
    scaled by n/64 here so that small grids keep the same layout.
    2D (FT/FluidTopoOptDriver.cpp:104-166): inflow e = 60 on the cells (0, i) and
    x-faces (1, i) with |i - n/2| <= n/32, and 16 outflows of strength
    r = e (2 n/32 + 1) / (2 n/64 + 1) / 16 centred at n/32 (2o + 1) on the x = n-1 side.
    """
    nc, nfx, nfy, nfz, tot = dof_sizes(dim, n)
    fixed = np.zeros(tot, np.uint8)
    body = np.zeros(tot)
    cf, fx, fy, *fzr = _split(dim, n, fixed)
    cb, bx, by, *bzr = _split(dim, n, body)
    fx[0] = 1
    fx[n] = 1
    fy[:, 0] = 1
    fy[:, n] = 1
    if dim == 3:
        fz, bz = fzr[0], bzr[0]
        fz[:, :, 0] = 1
        fz[:, :, n] = 1
        sc = (lambda v: v) if n >= 64 else (lambda v: v * n // 64)
        r = lambda a, b: slice(sc(a), sc(b))
        cb[0, r(28, 36), r(28, 36)] = 600.0
        bx[1, r(28, 36), r(28, 36)] = 600.0
        cb[n - 1, r(10, 18), r(34, 42)] = -200.0
        bx[n - 1, r(10, 18), r(34, 42)] = 200.0
        cb[r(28, 36), r(22, 30), n - 1] = -200.0
        bz[r(28, 36), r(22, 30), n - 1] = 200.0
        cb[r(22, 30), r(28, 36), 0] = -200.0
        bz[r(22, 30), r(28, 36), 1] = -200.0
        am = AlphaModel(alpha_max=500.0 / (n / 32.0), alpha_min=1e-3, q=0.01, power=3.0, scale=1.0)
    else:
        O = 16
        hin, hout = n // 32, n // (4 * O)
        e = 60.0
        rr = e * (2 * hin + 1) / (2 * hout + 1) / O
        i = np.arange(n)
        inflow = np.abs(i - n // 2) <= hin
        cb[0, inflow, 0] = e
        bx[1, inflow, 0] = e
        for o in range(O):
            out = np.abs(i - n // (2 * O) * (2 * o + 1)) <= hout
            cb[n - 1, out, 0] = -rr
            bx[n - 1, out, 0] = rr
        am = AlphaModel(alpha_max=500.0 / (n / 32.0), alpha_min=1e-3, q=0.014, power=3.0, scale=1.0)
    return Problem(dim=dim, n=n, fixed=fixed, dirichlet=np.zeros(tot), body=body, alpha=am, closed=True,
                   name=f"test0_{dim}d_{n}")


def grf_density(dim, n, mean=0.3, sigma=8.0, seed=12345, rho_min=1e-3):
    """Seeded Gaussian random field: white noise N(0,1) convolved (FFT) with a
    Gaussian of std `sigma` cells, normalised to unit variance, logistic-squashed
    with a shift found by bisection so the mean is `mean`, clipped at rho_min."""
    shape = (n,) * dim
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for _ in range(dim)]
    kk = np.zeros(shape)
    for a in range(dim):
        sh = [1] * dim
        sh[a] = n
        kk = kk + (k[a].reshape(sh)) ** 2
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-2.0 * (np.pi ** 2) * (sigma ** 2) * kk)))
    g = (g - g.mean()) / (g.std() + 1e-300)
    lo, hi = -20.0, 20.0
    for _ in range(80):
        s = 0.5 * (lo + hi)
        m = np.mean(1.0 / (1.0 + np.exp(-(g + s))))
        if m > mean:
            hi = s
        else:
            lo = s
    rho = 1.0 / (1.0 + np.exp(-(g + 0.5 * (lo + hi))))
    rho = np.maximum(rho, rho_min)
    return rho.reshape(-1).astype(np.float64)
