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
