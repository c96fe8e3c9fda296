"""The fluid_topo command-line driver and the frame format (CPU only).

The driver's problem setup (-dump) must equal problems.reference_test0, whose
entries are checked against the values written in the reference's drivers
(FT/FluidTopoOptDriver3D.cpp:99-167, FT/FluidTopoOptDriver.cpp:104-166).
"""
import os
import subprocess

import numpy as np
import pytest

from stokes_to_amd import output
from stokes_to_amd.problems import _split, reference_test0

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "stokes_to_amd", "fluid_topo")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        from stokes_to_amd.build import build_driver
        build_driver()
    return DRIVER


def test_test0_entries_match_the_reference_drivers():
    P = reference_test0(3, 64)
    c, fx, fy, fz = _split(3, 64, P.body)
    assert c[0, 28, 28] == 600 and c[0, 35, 35] == 600 and c[0, 36, 28] == 0 and c[0, 27, 28] == 0
    assert fx[1, 28, 35] == 600 and fx[0, 28, 35] == 0
    assert c[63, 10, 34] == -200 and c[63, 17, 41] == -200 and fx[63, 17, 41] == 200 and c[63, 18, 41] == 0
    assert c[28, 22, 63] == -200 and fz[35, 29, 63] == 200 and fz[35, 29, 64] == 0
    assert c[22, 28, 0] == -200 and fz[29, 35, 1] == -200 and fz[29, 35, 0] == 0
    assert abs(c.sum()) < 1e-9                       # sources balance the sinks (closed box)
    assert P.fixed.sum() == 6 * 64 * 64 and P.closed
    f = _split(3, 64, P.fixed)
    assert f[1][0].all() and f[1][64].all() and not f[1][1:64].any()
    assert abs(P.alpha.alpha_max - 250.0) < 1e-12 and P.alpha.power == 3 and P.alpha.q == 0.01
    Q = reference_test0(2, 256)
    c2, fx2, fy2 = _split(2, 256, Q.body)
    inflow = np.nonzero(c2[0, :, 0])[0]
    assert inflow.min() == 120 and inflow.max() == 136 and np.all(c2[0, inflow, 0] == 60)
    assert np.all(fx2[1, inflow, 0] == 60)
    out = np.nonzero(c2[255, :, 0])[0]
    assert len(out) == 16 * 9 and out.min() == 4 and out.max() == 252
    assert abs(c2.sum()) < 1e-9 and Q.alpha.q == 0.014


@pytest.mark.parametrize("dim,n", [(3, 64), (3, 32), (2, 256), (2, 64)])
def test_driver_problem_equals_python_setup(driver, tmp_path, dim, n):
    f = tmp_path / "p.bin"
    subprocess.check_call([driver, "-driver", str(dim - 2), "-test", "0", "-s", str(n), "-dump", str(f)])
    P = reference_test0(dim, n)
    raw = f.read_bytes()
    tot = P.fixed.size
    assert len(raw) == tot + 8 * tot
    assert np.array_equal(np.frombuffer(raw[:tot], np.uint8), P.fixed)
    assert np.array_equal(np.frombuffer(raw[tot:], np.float64), P.body)


@pytest.mark.parametrize("args", [["-test", "3"], ["-driver", "2"], ["-s", "48"], ["-frac", "1.5"], ["-bogus", "1"]])
def test_driver_rejects_bad_arguments(driver, args):
    base = {"-driver": "1", "-test": "0", "-s": "16", "-frac": "0.3", "-lf": "1"}
    for k, v in zip(args[::2], args[1::2]):
        base[k] = v
    cmd = [driver] + [x for kv in base.items() for x in kv]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode != 0 and "fluid_topo:" in r.stderr


def test_frame_format_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    for dim, n in ((3, 8), (2, 16)):
        P = reference_test0(dim, n)
        rho = rng.random(n ** dim)
        x = rng.standard_normal(P.fixed.size)
        d = tmp_path / f"o{dim}"
        output.write_frame(str(d), 0, dim, n, rho, x, np.zeros((0, 3)), np.ones((5, 3)))
        counts, dx, dmin = output.read_grid(str(d / "0" / "grid"))
        assert list(counts) == [n, n, n if dim == 3 else 0] and dx == 1.0 and not dmin.any()
        c, a = output.read_field(str(d / "0" / "x"))
        assert list(c) == [n, n, n if dim == 3 else 1] and np.array_equal(a.ravel(), rho)
        c, v = output.read_velocity(str(d / "0" / "velocity"))
        assert np.array_equal(v, output.cell_velocity(dim, n, x))
        assert output.read_points(str(d / "0" / "psi_N")).shape == (5, 3)
        assert output.read_points(str(d / "0" / "psi_D")).shape == (0, 3)
        assert (d / "0" / "last_frame.txt").read_text() == "0"
