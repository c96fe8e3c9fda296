"""Host-side mirror of the reference's Stokes solver facade and TO hooks.

`FluidSteadyState` follows ST/FluidSteadyState3D.h:11-39 (init / init_boundary /
update_b / update_penalty / solve, results in `vel` and `p`) and exposes the
operator through `applyMapping` like CX/LinearMapping.h:4-14.  Every call goes
through the C ABI of libstokes_hip.so; nothing is computed on the host.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ST_OK, ST_ENOCONV, ST_INIT_ZERO, ST_INIT_KEEP, ST_INIT_HOST, ST_INIT_LEARNED,
                   StParams, StStats, StokesError, dptr, fptr, u8ptr)
from .problems import dof_sizes

_INIT = {"zero": ST_INIT_ZERO, "keep": ST_INIT_KEEP, "host": ST_INIT_HOST, "learned": ST_INIT_LEARNED}
_PREC = {"fp64": 0, "fp32": 1, "bf16": 2}
_OPT = {"oc": 0, "mma": 1}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class FluidSteadyState:
    """One solver context on the GPU (one per thread; not thread-safe)."""

    def __init__(self, dim, n, lib_path=None, **params):
        self.lib = _lib.load(lib_path)
        p = StParams()
        self.lib.st_default_params(ctypes.byref(p), dim, n)
        for k, v in params.items():
            if k == "precision" and isinstance(v, str):
                v = _PREC[v]
            if k == "optimizer" and isinstance(v, str):
                v = _OPT[v]
            if not hasattr(p, k):
                raise TypeError(f"unknown parameter {k}")
            setattr(p, k, v)
        self.params = p
        h = ctypes.c_void_p()
        rc = self.lib.st_create(ctypes.byref(h), ctypes.byref(p))
        if rc != ST_OK:
            raise StokesError(rc, "st_create failed (is a GPU visible?)")
        self.h = h
        self.dim, self.n = dim, n
        self.ndof = self.lib.st_num_dofs(h)
        self.ncell = self.lib.st_num_cells(h)
        self.vel = None
        self.p = None
        self.stats = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.st_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, allow=()):
        if rc != ST_OK and rc not in allow:
            raise StokesError(rc, self.lib.st_last_error(self.h).decode())
        return rc

    # ---- ST/FluidSteadyState3D.cpp:92-150 ----
    def init_boundary(self, fixed, closed=None):
        """fixed[cells | faces] (1 = fixed); closed: pressure gauge of a domain whose
        boundary is fixed everywhere (the reference's init_boundary(..., _closed));
        None keeps the context's value."""
        f = np.ascontiguousarray(fixed, dtype=np.uint8)
        assert f.size == self.ndof
        c = -1 if closed is None else int(bool(closed))
        self._check(self.lib.st_set_geometry(self.h, u8ptr(f), c))

    def update_penalty(self, pen):
        pen = _f64(pen)
        assert pen.size == self.ndof
        self._check(self.lib.st_set_penalty(self.h, dptr(pen)))

    def set_design(self, rho):
        rho = _f64(rho)
        assert rho.size == self.ncell
        self._check(self.lib.st_set_design(self.h, dptr(rho)))

    def get_design(self):
        out = np.zeros(self.ncell)
        self._check(self.lib.st_get_design(self.h, dptr(out)))
        return out

    def update_b(self, body=None, dirichlet=None):
        b = None if body is None else _f64(body)
        d = None if dirichlet is None else _f64(dirichlet)
        self._check(self.lib.st_set_rhs(self.h, dptr(b) if b is not None else None,
                                        dptr(d) if d is not None else None))

    # ---- ST/FluidSteadyState3D.cpp:152-170 ----
    def solve(self, init="zero", x0=None, fetch=True, allow_noconv=False):
        st = StStats()
        mode = _INIT[init] if isinstance(init, str) else int(init)
        x0p = None
        if mode == ST_INIT_HOST:
            x0 = _f64(x0)
            x0p = dptr(x0)
        out = np.zeros(self.ndof) if fetch else None
        rc = self.lib.st_solve(self.h, mode, x0p, dptr(out) if fetch else None, ctypes.byref(st))
        self._check(rc, allow=(ST_ENOCONV,) if allow_noconv else ())
        self.stats = st.as_dict()
        if fetch:
            nc = self.ncell
            self.p = out[:nc]
            self.vel = out[nc:]
        return out, self.stats

    def gradient(self):
        """dJ/drho of the last to_step (filter chain rule applied)."""
        out = np.zeros(self.ncell)
        self._check(self.lib.st_get_gradient(self.h, dptr(out)))
        return out

    def reset_solution(self):
        """Forget the held solutions (start of a new run)."""
        self._check(self.lib.st_reset_solution(self.h))

    def solution(self):
        out = np.zeros(self.ndof)
        self._check(self.lib.st_get_solution(self.h, dptr(out)))
        return out

    # ---- CX/LinearMapping.h:4-14 ----
    def applyMapping(self, x):
        x = _f64(x)
        y = np.zeros(self.ndof)
        self._check(self.lib.st_apply(self.h, dptr(x), dptr(y)))
        return y

    # ---- FT/FluidEnergy3D.cpp:71-109 ----
    def compute_energy(self, x):
        x = _f64(x)
        J = ctypes.c_double()
        self._check(self.lib.st_objective(self.h, dptr(x), ctypes.byref(J)))
        return J.value

    def compute_gradient_x(self, x):
        x = _f64(x)
        y = np.zeros(self.ndof)
        self._check(self.lib.st_energy_grad(self.h, dptr(x), dptr(y)))
        return y

    # ---- FT/FluidTopoOpt3D.cpp:170-198 ----
    def sensitivity(self, rho, x, lam):
        out = np.zeros(self.ncell)
        self._check(self.lib.st_sensitivity(self.h, dptr(_f64(rho)), dptr(_f64(x)), dptr(_f64(lam)), dptr(out)))
        return out

    def filter(self, v, transpose=False):
        out = np.zeros(self.ncell)
        self._check(self.lib.st_filter(self.h, dptr(_f64(v)), dptr(out), int(transpose)))
        return out

    def oc_update(self, rho, dc, volfrac, move=0.2, eta=0.5, rho_min=1e-3):
        out = np.zeros(self.ncell)
        ch = ctypes.c_double()
        self._check(self.lib.st_oc_update(self.h, dptr(_f64(rho)), dptr(_f64(dc)), float(volfrac), float(move),
                                          float(eta), float(rho_min), dptr(out), ctypes.byref(ch)))
        return out, ch.value

    # ---- EXT/mma/MMASolver.cpp:74-89 ----
    def mma_update(self, x, dfdx, g, dgdx, xmin, xmax):
        """One MMA step of x (returned, a new array) with one constraint g(x) <= 0."""
        x = _f64(x).copy()
        steps = ctypes.c_int(0)
        self._check(self.lib.st_mma_update(self.h, dptr(x), dptr(_f64(dfdx)), float(g), dptr(_f64(dgdx)),
                                           dptr(_f64(xmin)), dptr(_f64(xmax)), ctypes.byref(steps)))
        return x, steps.value

    def mma_reset(self):
        self._check(self.lib.st_mma_reset(self.h))

    def to_step(self, volfrac, move=0.2, eta=0.5, rho_min=1e-3, init="zero"):
        """One design iteration on the device (FT/FluidTopoOpt3D.cpp:4-17); init is
        'zero' (both solves from zero), 'keep' (previous solutions) or 'learned'
        (previous solution + learned coarse-grid correction)."""
        J, vol, ch = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        f, a = StStats(), StStats()
        mode = _INIT[init] if isinstance(init, str) else int(init)
        self._check(self.lib.st_to_step(self.h, mode, float(volfrac), float(move), float(eta), float(rho_min),
                                        ctypes.byref(J), ctypes.byref(vol), ctypes.byref(ch), ctypes.byref(f),
                                        ctypes.byref(a)))
        return {"J": J.value, "vol": vol.value, "change": ch.value, "fwd": f.as_dict(), "adj": a.as_dict()}

    # ---- learned predictor ----
    def record_sample(self):
        self._check(self.lib.st_record_sample(self.h))

    def net_eval(self):
        """(loss of the network, loss of the zero prediction) on the current solve."""
        l, z = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.st_net_eval(self.h, ctypes.byref(l), ctypes.byref(z)))
        return l.value, z.value

    def num_samples(self):
        return self.lib.st_num_samples(self.h)

    def net_train(self, epochs, lr=1e-3):
        loss = ctypes.c_double()
        self._check(self.lib.st_net_train(self.h, int(epochs), float(lr), ctypes.byref(loss)))
        return loss.value

    def net_params(self):
        n = self.lib.st_net_num_params(self.h)
        if n < 0:
            self._check(-1)
        p = np.zeros(n, np.float32)
        self._check(self.lib.st_net_get_params(self.h, fptr(p)))
        return p

    def net_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.lib.st_net_set_params(self.h, fptr(p)))

    def net_grad(self, feat, tgt):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        tgt = np.ascontiguousarray(tgt, dtype=np.float32)
        n = self.lib.st_net_num_params(self.h)
        g = np.zeros(n, np.float32)
        out = np.zeros(feat.size // 4 * 16, np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.st_net_grad(self.h, fptr(feat), fptr(tgt), ctypes.byref(loss), fptr(g), fptr(out)))
        return loss.value, g, out

    # ---- multigrid internals (component tests) ----
    def mg_levels(self):
        return self.lib.st_mg_levels(self.h)

    def mg_level_n(self, l):
        return self.lib.st_mg_level_n(self.h, l)

    def mg_level_desc(self, l):
        nl = self.mg_level_n(l)
        nd = dof_sizes(self.dim, nl)[-1]
        nc = dof_sizes(self.dim, nl)[0]
        fixed = np.zeros(nd, np.uint8)
        al = np.zeros(nd - nc)
        self._check(self.lib.st_mg_level_desc(self.h, l, u8ptr(fixed), dptr(al)))
        return fixed, al

    def mg_vcycle(self, b):
        z = np.zeros(self.ndof)
        self._check(self.lib.st_mg_vcycle(self.h, dptr(_f64(b)), dptr(z)))
        return z

    def mg_smooth(self, l, color, b, xin, mixed=False):
        """One smoother colour step on level l; mixed=True runs the context's
        mixed-precision (FP32/BF16 storage) smoother."""
        nd = dof_sizes(self.dim, self.mg_level_n(l))[-1]
        out = np.zeros(nd)
        fn = self.lib.st_mg_smooth_mp if mixed else self.lib.st_mg_smooth
        self._check(fn(self.h, l, color, dptr(_f64(b)), dptr(_f64(xin)), dptr(out)))
        return out

    def mg_restrict(self, l, rf):
        nd = dof_sizes(self.dim, self.mg_level_n(l + 1))[-1]
        out = np.zeros(nd)
        self._check(self.lib.st_mg_restrict(self.h, l, dptr(_f64(rf)), dptr(out)))
        return out

    def mg_prolong(self, l, ec, xf):
        out = np.zeros_like(_f64(xf))
        self._check(self.lib.st_mg_prolong(self.h, l, dptr(_f64(ec)), dptr(_f64(xf)), dptr(out)))
        return out


def conv_test(R, ntap, cin, cout, relu, mode, x=None, w=None, b=None, gout=None, lib_path=None):
    """One network convolution layer on the GPU through st_net_conv_test (host arrays, FP32
    in/out, BF16 operands and FP32 accumulation inside)."""
    lib = _lib.load(lib_path)
    nv = R ** 3
    f = lambda a, n: np.ascontiguousarray(a if a is not None else np.zeros(n), dtype=np.float32)
    x, w, b, gout = f(x, nv * cin), f(w, ntap * cin * cout), f(b, cout), f(gout, nv * cout)
    out = np.zeros(nv * max(cin, cout), np.float32)
    gw = np.zeros(ntap * cin * cout, np.float32)
    gb = np.zeros(cout, np.float32)
    rc = lib.st_net_conv_test(R, ntap, cin, cout, int(relu), int(mode), fptr(x), fptr(w), fptr(b), fptr(gout),
                              fptr(out), fptr(gw), fptr(gb))
    if rc != ST_OK:
        raise StokesError(rc, "st_net_conv_test failed")
    if mode == 0:
        return out[: nv * cout]
    if mode == 1:
        return out[: nv * cin]
    return gw, gb
