"""The reference's test 0 on the GPU: closed-box solves against the oracle, and the
fluid_topo driver end to end (frames in the reference's format)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel
from stokes_to_amd import output
from stokes_to_amd.problems import reference_test0

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "stokes_to_amd", "fluid_topo")


def _test0_solver(make_solver, P, **kw):
    a = P.alpha
    s = make_solver(P.dim, P.n, alpha_max=a.alpha_max, alpha_min=a.alpha_min, q=a.q, power=a.power,
                    alpha_scale=a.scale, filter_radius=1.0, **kw)
    # the gauge comes with the geometry, as in init_boundary(..., closed=true)
    s.init_boundary(P.fixed, closed=P.closed)
    s.update_b(body=P.body)
    return s


@pytest.mark.parametrize("dim,n", [(3, 32), (2, 64)])
def test_closed_box_solve_matches_oracle(make_solver, orc, dim, n):
    """Closed box (pressure defined up to a constant): the velocity and the pressure
    modulo its mean agree with the oracle's solve, and the true residual meets rtol."""
    P = reference_test0(dim, n)
    rho = np.full(n ** dim, 0.3)
    s = _test0_solver(make_solver, P, rtol=1e-10)
    s.set_design(rho)
    x, st = s.solve()
    assert st["converged"]
    nn, sz = orc.sizes(dim, *P.shape)
    nc = sz[0]
    fa = orc.design_to_alpha(dim, P.shape, rho, P.alpha.array(), P.fixed[nc:])
    pen = np.zeros(nn)
    pen[nc:] = fa
    vol = np.zeros(nn)
    vol[nc:] = 1.0 - P.fixed[nc:]
    b = orc.lift_rhs(dim, P.shape, P.fixed, P.dirichlet) + np.where(P.fixed.astype(bool), 0.0, P.body)
    r = b - orc.apply(dim, P.shape, P.fixed, vol, pen, x)
    assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
    mg = orc.MG(dim, P.shape, P.fixed, pen, closed=1)
    xo, it, rr, _ = mg.fgmres(b, rtol=1e-12)
    assert rel(x[nc:], xo[nc:]) < 1e-8
    po, pg = xo[:nc] - xo[:nc].mean(), x[:nc] - x[:nc].mean()
    assert rel(pg, po) < 1e-7


@pytest.mark.parametrize("driver,n", [(1, 32), (0, 64)])
def test_fluid_topo_driver_end_to_end(tmp_path, driver, n):
    out = tmp_path / "out"
    frac, lf = 0.3, 3
    r = subprocess.run([DRIVER, "-driver", str(driver), "-test", "0", "-s", str(n), "-o", str(out), "-frac", str(frac),
                        "-lf", str(lf)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("iteration")]
    assert len(lines) == lf
    J = [float(l.split("objective")[1].split()[0]) for l in lines]
    assert all(np.isfinite(J)) and J[-1] < J[0]           # MMA lowers the dissipation
    dim = 3 if driver == 1 else 2
    nz = n if dim == 3 else 1
    counts, dx, dmin = output.read_grid(str(out / "0" / "grid"))
    assert list(counts) == [n, n, n if dim == 3 else 0] and dx == 1.0
    for k in range(lf):
        c, x = output.read_field(str(out / str(k) / "x"))
        assert list(c) == [n, n, nz]
        c, v = output.read_velocity(str(out / str(k) / "velocity"))
        assert v.shape == (n, n, nz, 3) and np.isfinite(v).all() and np.abs(v).max() > 0
    _, x0 = output.read_field(str(out / "0" / "x"))
    assert np.allclose(x0, frac)                           # frame 0 is the initial design
    prev = x0
    for k in range(1, lf):
        _, xk = output.read_field(str(out / str(k) / "x"))
        # MMA box of OptimizerMma.h: move 0.2, bounds [0.01, 1].  (The volume may
        # exceed frac in early iterations: the constraint is enforced through the
        # subproblem's penalty c = 1000, as in the reference.)
        assert np.abs(xk - prev).max() <= 0.2 + 1e-12
        assert xk.min() >= 0.01 - 1e-12 and xk.max() <= 1.0 + 1e-12
        prev = xk
    nfix = 6 * n * n if dim == 3 else 4 * n
    assert output.read_points(str(out / "0" / "psi_N")).shape == (nfix, 3)
    assert output.read_points(str(out / "0" / "psi_D")).shape == (0, 3)
    assert (out / "0" / "last_frame.txt").read_text() == str(lf - 1)


BINDING = os.path.join(ROOT, "stokes_to_amd", "binding_test0")


def test_reference_call_sequence_through_the_cpp_binding(tmp_path, orc):
    """The C++ binding a maintainer adds to the reference (stokes_to_amd/driver/
    hip_steady_state3d.h, shown in INTEGRATION.md) run in the reference's order --
    init, init_boundary(..., closed = true), update_penalty, update_b, solve -- on the
    reference's test 0 at 32^3, a closed box: the velocity and the pressure modulo its
    mean agree with the oracle's solve."""
    n = 32
    P = reference_test0(3, n)
    rho = np.random.default_rng(2).uniform(0.05, 1.0, n ** 3)
    nn, sz = orc.sizes(3, *P.shape)
    nc = sz[0]
    fa = orc.design_to_alpha(3, P.shape, rho, P.alpha.array(), P.fixed[nc:])
    d = tmp_path
    P.fixed[:nc].astype(np.int32).tofile(d / "cell_fixed")
    P.fixed[nc:].astype(np.int32).tofile(d / "face_fixed")
    P.body[:nc].tofile(d / "cell_b")
    P.body[nc:].tofile(d / "face_b")
    np.zeros(nc).tofile(d / "cell_pen")
    fa.tofile(d / "face_pen")
    (d / "closed").write_text("1")
    print(f"binding_test0 {n} (test 0, closed box)", flush=True)
    r = subprocess.run([BINDING, str(n), str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    iters, relres = r.stdout.split()
    assert float(relres) <= 1e-10
    x = np.concatenate([np.fromfile(d / "p"), np.fromfile(d / "vel")])
    pen = np.zeros(nn)
    pen[nc:] = fa
    b = np.where(P.fixed.astype(bool), 0.0, P.body)
    mg = orc.MG(3, P.shape, P.fixed, pen, closed=1)
    xo, it, rr, _ = mg.fgmres(b, rtol=1e-12)
    assert rel(x[nc:], xo[nc:]) < 1e-8
    assert rel(x[:nc] - x[:nc].mean(), xo[:nc] - xo[:nc].mean()) < 1e-7
    # without the gauge the same sequence stops with the library's error, not silently
    (d / "closed").write_text("0")
    r = subprocess.run([BINDING, str(n), str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "singular" in r.stderr
