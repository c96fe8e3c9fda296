"""Shared test setup: problems, oracle descriptors, random vectors, device buffers."""
import ctypes

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


class DeviceArray:
    """A device buffer allocated through the HIP runtime that libstokes_hip.so uses
    (the process's libamdhip64.so.7), for the entry points that take device pointers.
    torch is not used for this: it ships its own HIP runtime under the same soname, and
    a process can only hold one of them."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            from stokes_to_amd import _lib
            _lib.load()
            h = ctypes.CDLL("libamdhip64.so.7")
            vp = ctypes.c_void_p
            h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
            h.hipFree.argtypes = [vp]
            h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
            h.hipStreamSynchronize.argtypes = [vp]
            h.hipStreamDestroy.argtypes = [vp]
            for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                      "hipStreamSynchronize", "hipStreamDestroy"):
                getattr(h, f).restype = ctypes.c_int
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.dtype, self.size = host.dtype, host.size
        self.nbytes = max(host.nbytes, 1)
        self.ptr = ctypes.c_void_p()
        h = self.hip()
        assert h.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        assert h.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1) == 0   # host to device

    def get(self):
        out = np.empty(self.size, self.dtype)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()
