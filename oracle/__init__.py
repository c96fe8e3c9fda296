"""CPU FP64 oracle (ctypes wrapper over oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference results tests compare against and
as the CPU baseline.  The product
package (stokes_to_amd) never imports it.

Arrays use the ABI layout: C-order (i, j, k), k fastest; a dof vector is
[p | u_x | u_y | u_z] and a descriptor array uses the same [cells | faces] order
(ST/StokeFlowDescriptor3D.cu:6-50).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")
_lib = None

D = ctypes.POINTER(ctypes.c_double)
U8 = ctypes.POINTER(ctypes.c_uint8)
I = ctypes.c_int


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        L = _lib
        L.orc_dof.restype = ctypes.c_long
        L.orc_dof.argtypes = [I, I, I, I, ctypes.POINTER(ctypes.c_long)]
        L.orc_edge_count.restype = ctypes.c_long
        for nm in ("orc_d2", "orc_cod0", "orc_cod1", "orc_d1"):
            getattr(L, nm).argtypes = [I, I, I, I, D, D]
        for nm in ("orc_apply", "orc_energy_op"):
            getattr(L, nm).argtypes = [I, I, I, I, U8, D, D, D, D]
        L.orc_apply_ref2d.argtypes = [I, I, U8, D, D, D, D]
        L.orc_energy_op_ref2d.argtypes = [I, I, U8, D, D, D]
        L.orc_lift_rhs.argtypes = [I, I, I, I, U8, D, D]
        L.orc_alpha.restype = ctypes.c_double
        L.orc_alpha.argtypes = [ctypes.c_double, D]
        L.orc_dalpha.restype = ctypes.c_double
        L.orc_dalpha.argtypes = [ctypes.c_double, D]
        L.orc_design_to_alpha.argtypes = [I, I, I, I, D, D, U8, D]
        L.orc_sensitivity.argtypes = [I, I, I, I, D, D, U8, D, D, D]
        L.orc_objective.restype = ctypes.c_double
        L.orc_objective.argtypes = [I, I, I, I, U8, D, D, D]
        L.orc_filter.argtypes = [I, I, I, I, ctypes.c_double, D, D, I]
        L.orc_oc_update.restype = I
        L.orc_oc_update.argtypes = [I, I, I, I, ctypes.c_double, D, D, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D]
        L.orc_mg_create.restype = ctypes.c_void_p
        L.orc_mg_create.argtypes = [I, I, I, I, U8, D, D]
        L.orc_mg_free.argtypes = [ctypes.c_void_p]
        L.orc_mg_levels.restype = I
        L.orc_mg_levels.argtypes = [ctypes.c_void_p]
        L.orc_mg_level_dims.argtypes = [ctypes.c_void_p, I, ctypes.POINTER(I)]
        L.orc_mg_level_desc.argtypes = [ctypes.c_void_p, I, U8, D]
        L.orc_mg_vcycle.argtypes = [ctypes.c_void_p, D, D]
        L.orc_mg_apply.argtypes = [ctypes.c_void_p, I, D, D]
        L.orc_smooth_color.argtypes = [I, I, I, I, U8, D, D, D, D, I, ctypes.c_double]
        L.orc_restrict.argtypes = [I, I, I, I, U8, I, D, D]
        L.orc_prolong_add.argtypes = [I, I, I, I, U8, I, D, D]
        L.orc_coarsen.argtypes = [I, I, I, I, U8, D, U8, D]
        L.orc_conv3d.argtypes = [I, I, I, I, D, D, D, D, I]
        L.orc_conv3d_bwd_data.argtypes = [I, I, I, I, D, D, D]
        L.orc_conv3d_bwd_weight.argtypes = [I, I, I, I, D, D, D, D]
        L.orc_mma_create.restype = ctypes.c_void_p
        L.orc_mma_create.argtypes = [ctypes.c_long]
        L.orc_mma_free.argtypes = [ctypes.c_void_p]
        L.orc_mma_update.restype = I
        L.orc_mma_update.argtypes = [ctypes.c_void_p, D, D, ctypes.c_double, D, D, D]
        L.orc_mma_state.argtypes = [ctypes.c_void_p, D, D, D]
        L.orc_fgmres.restype = I
        L.orc_fgmres.argtypes = [ctypes.c_void_p, D, D, ctypes.c_double, I, I, D, D]
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def _u(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(U8)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _b(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def sizes(dim, nx, ny, nz):
    s = (ctypes.c_long * 4)()
    n = lib().orc_dof(dim, nx, ny, nz, s)
    return n, list(s)


def edge_count(dim, nx, ny, nz):
    return lib().orc_edge_count(dim, nx, ny, nz)


def d2(dim, shape, faces):
    nx, ny, nz = shape
    n, s = sizes(dim, nx, ny, nz)
    out = np.zeros(s[0])
    lib().orc_d2(dim, nx, ny, nz, _d(_f64(faces)), _d(out))
    return out


def cod0(dim, shape, cells):
    nx, ny, nz = shape
    n, s = sizes(dim, nx, ny, nz)
    out = np.zeros(n - s[0])
    lib().orc_cod0(dim, nx, ny, nz, _d(_f64(cells)), _d(out))
    return out


def cod1(dim, shape, faces):
    nx, ny, nz = shape
    out = np.zeros(edge_count(dim, nx, ny, nz))
    lib().orc_cod1(dim, nx, ny, nz, _d(_f64(faces)), _d(out))
    return out


def d1(dim, shape, edges):
    nx, ny, nz = shape
    n, s = sizes(dim, nx, ny, nz)
    out = np.zeros(n - s[0])
    lib().orc_d1(dim, nx, ny, nz, _d(_f64(edges)), _d(out))
    return out


def apply(dim, shape, fixed, vol, pen, x):
    nx, ny, nz = shape
    y = np.zeros_like(_f64(x))
    lib().orc_apply(dim, nx, ny, nz, _u(_b(fixed)), _d(_f64(vol)), _d(_f64(pen)), _d(_f64(x)), _d(y))
    return y


def apply_ref2d(shape, fixed, vol, pen, x):
    nx, ny = shape[:2]
    y = np.zeros_like(_f64(x))
    lib().orc_apply_ref2d(nx, ny, _u(_b(fixed)), _d(_f64(vol)), _d(_f64(pen)), _d(_f64(x)), _d(y))
    return y


def energy_op(dim, shape, fixed, vol, pen, x):
    nx, ny, nz = shape
    y = np.zeros_like(_f64(x))
    lib().orc_energy_op(dim, nx, ny, nz, _u(_b(fixed)), _d(_f64(vol)), _d(_f64(pen)), _d(_f64(x)), _d(y))
    return y


def energy_op_ref2d(shape, fixed, pen, x):
    nx, ny = shape[:2]
    y = np.zeros_like(_f64(x))
    lib().orc_energy_op_ref2d(nx, ny, _u(_b(fixed)), _d(_f64(pen)), _d(_f64(x)), _d(y))
    return y


def objective(dim, shape, fixed, vol, pen, x):
    nx, ny, nz = shape
    return lib().orc_objective(dim, nx, ny, nz, _u(_b(fixed)), _d(_f64(vol)), _d(_f64(pen)), _d(_f64(x)))


def lift_rhs(dim, shape, fixed, xd):
    nx, ny, nz = shape
    b = np.zeros_like(_f64(xd))
    lib().orc_lift_rhs(dim, nx, ny, nz, _u(_b(fixed)), _d(_f64(xd)), _d(b))
    return b


def alpha_params(alpha_max, alpha_min, q, power, scale):
    return np.array([alpha_max, alpha_min, q, power, scale], dtype=np.float64)


def alpha(rho, am):
    return lib().orc_alpha(float(rho), _d(_f64(am)))


def dalpha(rho, am):
    return lib().orc_dalpha(float(rho), _d(_f64(am)))


def design_to_alpha(dim, shape, rho, am, face_fixed):
    nx, ny, nz = shape
    n, s = sizes(dim, nx, ny, nz)
    out = np.zeros(n - s[0])
    lib().orc_design_to_alpha(dim, nx, ny, nz, _d(_f64(rho)), _d(_f64(am)), _u(_b(face_fixed)), _d(out))
    return out


def sensitivity(dim, shape, rho, am, face_fixed, u, lam):
    nx, ny, nz = shape
    n, s = sizes(dim, nx, ny, nz)
    out = np.zeros(s[0])
    lib().orc_sensitivity(dim, nx, ny, nz, _d(_f64(rho)), _d(_f64(am)), _u(_b(face_fixed)),
                          _d(_f64(u)), _d(_f64(lam)), _d(out))
    return out


def filt(dim, shape, radius, x, transpose=False):
    nx, ny, nz = shape
    out = np.zeros_like(_f64(x))
    lib().orc_filter(dim, nx, ny, nz, float(radius), _d(_f64(x)), _d(out), int(transpose))
    return out


def oc_update(dim, shape, radius, rho, dc, volfrac, move=0.2, eta=0.5, rho_min=1e-3):
    nx, ny, nz = shape
    out = np.zeros_like(_f64(rho))
    ch = ctypes.c_double(0)
    lib().orc_oc_update(dim, nx, ny, nz, float(radius), _d(_f64(rho)), _d(_f64(dc)), float(volfrac),
                        float(move), float(eta), float(rho_min), _d(out), ctypes.byref(ch))
    return out, ch.value


def conv3d(R, ntap, cin, cout, x, w, b=None, relu=False):
    """FP64 convolution, channels-last (R^3 x cin) -> (R^3 x cout), zero padding."""
    out = np.zeros(R ** 3 * cout)
    bb = _f64(b) if b is not None else np.zeros(cout)
    lib().orc_conv3d(R, ntap, cin, cout, _d(_f64(x)), _d(_f64(w)), _d(bb), _d(out), int(relu))
    return out


def conv3d_bwd_data(R, ntap, cin, cout, gout, w):
    out = np.zeros(R ** 3 * cin)
    lib().orc_conv3d_bwd_data(R, ntap, cin, cout, _d(_f64(gout)), _d(_f64(w)), _d(out))
    return out


def conv3d_bwd_weight(R, ntap, cin, cout, x, gout):
    gw = np.zeros(ntap * cin * cout)
    gb = np.zeros(cout)
    lib().orc_conv3d_bwd_weight(R, ntap, cin, cout, _d(_f64(x)), _d(_f64(gout)), _d(gw), _d(gb))
    return gw, gb


def smooth_color(dim, shape, fixed, pen, b, xin, color, omega):
    nx, ny, nz = shape
    out = np.zeros_like(_f64(xin))
    lib().orc_smooth_color(dim, nx, ny, nz, _u(_b(fixed)), _d(_f64(pen)), _d(_f64(b)), _d(_f64(xin)),
                           _d(out), int(color), float(omega))
    return out


def restrict(dim, fshape, cfixed, tang, rf):
    nx, ny, nz = fshape
    cs = (nx // 2, ny // 2, nz // 2 if dim == 3 else 1)
    n, _ = sizes(dim, *cs)
    out = np.zeros(n)
    lib().orc_restrict(dim, nx, ny, nz, _u(_b(cfixed)), int(tang), _d(_f64(rf)), _d(out))
    return out


def prolong_add(dim, fshape, ffixed, tang, ec, xf):
    nx, ny, nz = fshape
    xf = _f64(xf).copy()
    lib().orc_prolong_add(dim, nx, ny, nz, _u(_b(ffixed)), int(tang), _d(_f64(ec)), _d(xf))
    return xf


def coarsen(dim, fshape, ffixed, fpen):
    nx, ny, nz = fshape
    cs = (nx // 2, ny // 2, nz // 2 if dim == 3 else 1)
    n, _ = sizes(dim, *cs)
    cf = np.zeros(n, np.uint8)
    cp = np.zeros(n)
    lib().orc_coarsen(dim, nx, ny, nz, _u(_b(ffixed)), _d(_f64(fpen)), _u(cf), _d(cp))
    return cf, cp


class MG:
    """The multigrid preconditioner and FGMRES driver, restated on the CPU."""

    def __init__(self, dim, shape, fixed, pen, nu_pre=2, nu_post=2, omega=1.0, tang=0, closed=0,
                 coarsest=2):
        nx, ny, nz = shape
        self.dim, self.shape = dim, shape
        self._fixed = _b(fixed)
        self._pen = _f64(pen)
        prm = np.array([nu_pre, nu_post, omega, tang, closed, coarsest, 0.0])
        self.h = lib().orc_mg_create(dim, nx, ny, nz, _u(self._fixed), _d(self._pen), _d(prm))
        self.n = self._pen.size

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_free(self.h)
            self.h = None

    @property
    def levels(self):
        return lib().orc_mg_levels(self.h)

    def level_dims(self, l):
        d = (ctypes.c_int * 3)()
        lib().orc_mg_level_dims(self.h, l, d)
        return tuple(d)

    def level_desc(self, l):
        n, _ = sizes(self.dim, *self.level_dims(l))
        f = np.zeros(n, np.uint8)
        p = np.zeros(n)
        lib().orc_mg_level_desc(self.h, l, _u(f), _d(p))
        return f, p

    def vcycle(self, b):
        x = np.zeros(self.n)
        lib().orc_mg_vcycle(self.h, _d(_f64(b)), _d(x))
        return x

    def apply(self, x, level=0):
        y = np.zeros_like(_f64(x))
        lib().orc_mg_apply(self.h, level, _d(_f64(x)), _d(y))
        return y

    def fgmres(self, b, x0=None, rtol=1e-10, restart=30, maxit=200):
        x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
        hist = np.zeros(maxit + 2)
        rr = ctypes.c_double(0)
        it = lib().orc_fgmres(self.h, _d(_f64(b)), _d(x), float(rtol), int(restart), int(maxit),
                              ctypes.byref(rr), _d(hist))
        return x, it, rr.value, hist[: it + 1]


def to_step(P, rho, radius=1.5, volfrac=0.3, move=0.2, eta=0.5, rho_min=1e-3, rtol=1e-10, x_prev=None, mma=None):
    """One density-filtered design iteration, composed from the oracle pieces in the
    order of FT/FluidTopoOpt3D.cpp:4-17 (objective -> gradient via adjoint -> update).
    The update is OC, or with `mma` (an oracle MMA object) the MMA step with the
    reference's volume constraint 1e4 (mean(H rho) - volfrac) <= 0 and move box.
    Returns (rho_new, J, change, x, lam)."""
    dim, shape = P.dim, P.shape
    n, s = sizes(dim, *shape)
    nc = s[0]
    rhof = filt(dim, shape, radius, rho) if radius > 1.0 else rho.copy()
    am = P.alpha.array()
    fa = design_to_alpha(dim, shape, rhof, am, P.fixed[nc:])
    pen = np.zeros(n)
    pen[nc:] = fa
    vol = np.zeros(n)
    vol[nc:] = 1.0 - P.fixed[nc:]
    b = lift_rhs(dim, shape, P.fixed, P.dirichlet)
    mg = MG(dim, shape, P.fixed, pen)
    x, _, rr, _ = mg.fgmres(b, x0=P.dirichlet if x_prev is None else x_prev, rtol=rtol)
    J = objective(dim, shape, P.fixed, vol, pen, x)
    gx = energy_op(dim, shape, P.fixed, vol, pen, x)
    lam, _, _, _ = mg.fgmres(gx, rtol=rtol)
    dj = sensitivity(dim, shape, rhof, am, P.fixed[nc:], x[nc:], lam[nc:])
    dc = filt(dim, shape, radius, dj, transpose=True) if radius > 1.0 else dj
    if mma is None:
        rn, ch = oc_update(dim, shape, radius, rho, dc, volfrac, move, eta, rho_min)
        return rn, J, ch, x, lam
    nc_ = rho.size
    dv = np.full(nc_, 1.0 / nc_)
    if radius > 1.0:
        dv = filt(dim, shape, radius, dv, transpose=True)
    coef = 1e4
    lo, hi = mma_bounds(rho, move, rho_min, 1.0)
    rn = _f64(rho).copy()
    mma.update(rn, dc, coef * (rho @ dv - volfrac), coef * dv, lo, hi)
    return rn, J, np.abs(rn - rho).max(), x, lam


# ------------------------------------------------------------------ MMA

class MMA:
    """The oracle's MMA (one constraint); `update` changes x in place like the
    reference's MMASolver::Update (EXT/mma/MMASolver.cpp:74-89)."""

    def __init__(self, n):
        self.n = n
        self.h = lib().orc_mma_create(n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mma_free(self.h)
            self.h = None

    def update(self, x, df, g, dg, xmin, xmax):
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == self.n
        return lib().orc_mma_update(self.h, _d(x), _d(_f64(df)), float(g), _d(_f64(dg)), _d(_f64(xmin)),
                                    _d(_f64(xmax)))

    def state(self):
        lm = np.zeros(2)
        lo = np.zeros(self.n)
        up = np.zeros(self.n)
        lib().orc_mma_state(self.h, _d(lm), _d(lo), _d(up))
        return lm, lo, up


_REF_MMA = os.path.join(_HERE, "_ref", "libmma_ref.so")


def mma_ref_available():
    return os.path.exists(_REF_MMA)


class MMARef:
    """The reference's own MMASolver (compiled from its source by `make ref`)."""
    _l = None

    def __init__(self, n, m=1):
        if MMARef._l is None:
            L = ctypes.CDLL(_REF_MMA)
            L.mmaref_create.restype = ctypes.c_void_p
            L.mmaref_create.argtypes = [I, I]
            L.mmaref_free.argtypes = [ctypes.c_void_p]
            L.mmaref_update.argtypes = [ctypes.c_void_p, D, D, D, D, D, D]
            MMARef._l = L
        self.n, self.m = n, m
        self.h = MMARef._l.mmaref_create(n, m)

    def __del__(self):
        if getattr(self, "h", None):
            MMARef._l.mmaref_free(self.h)
            self.h = None

    def update(self, x, df, g, dg, xmin, xmax):
        gg = np.atleast_1d(_f64(g))
        MMARef._l.mmaref_update(self.h, _d(x), _d(_f64(df)), _d(gg), _d(_f64(dg)), _d(_f64(xmin)), _d(_f64(xmax)))


def mma_bounds(x, move, lb, ub):
    """Outer move limits of OptimizerMma::Optimize_MMA (SX/optimization/OptimizerMma.h:157-164)."""
    return np.maximum(lb, x - move), np.minimum(ub, x + move)


def volume_constraint(x, frac, coef=1e4):
    """g = coef (mean x - frac), dg/dx = coef / n (FT/FluidTopoOpt3D.cpp:19-39)."""
    n = x.size
    return coef * (x.sum() / n - frac), np.full(n, coef / n)
