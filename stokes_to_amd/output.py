"""Frames in the reference's on-disk format (read and write).

The reference's drivers write, per design iteration k, DIR/k/{x, velocity, psi_D,
psi_N}, plus DIR/0/grid and DIR/0/last_frame.txt (SX/common/Driver.h:57-70,
FT/FluidTopoOptDriver3D.cpp:35-92).  The binary layouts follow the Write_Binary
functions of SX/reservoir (Field.cpp:75-79, FaceField.cpp:108-117,
Grid.cpp:240-245, Particles.cpp:257-267), little-endian:
  grid      int32 cell_counts[3], double dx, double domain_min[3]
  field     int32 counts[3], then one value (or double[3]) per cell, C order
  points    int32 n, then X, V, F (double[3] each), M, C (double) and I (int32)
            arrays of n entries
2D data is written as 3D: counts (n, n, 1), grid counts (n, n, 0), zero z parts.
"""
import os

import numpy as np


def _i32(f, n):
    return np.frombuffer(f.read(4 * n), dtype="<i4")


def _f64(f, n):
    return np.frombuffer(f.read(8 * n), dtype="<f8")


def read_grid(path):
    with open(path, "rb") as f:
        counts = _i32(f, 3).copy()
        dx = float(_f64(f, 1)[0])
        dmin = _f64(f, 3).copy()
    return counts, dx, dmin


def read_field(path, ncomp=1):
    """(counts, array of shape counts[+ (ncomp,)])."""
    with open(path, "rb") as f:
        counts = _i32(f, 3).copy()
        n = int(np.prod(counts))
        a = _f64(f, n * ncomp).copy()
    shape = tuple(int(c) for c in counts) + ((ncomp,) if ncomp > 1 else ())
    return counts, a.reshape(shape)


def read_velocity(path):
    return read_field(path, ncomp=3)


def read_points(path):
    """Positions (n, 3) of a psi_D / psi_N point set."""
    with open(path, "rb") as f:
        n = int(_i32(f, 1)[0])
        X = _f64(f, 3 * n).reshape(n, 3).copy() if n > 0 else np.zeros((0, 3))
    return X


def cell_velocity(dim, n, x):
    """Face velocities of a dof vector -> per-cell vectors (mean of the two faces per
    axis), the conversion of FaceField::Write_To_File_3d."""
    from .problems import dof_sizes
    nc, nfx, nfy, nfz, _ = dof_sizes(dim, n)
    nz = n if dim == 3 else 1
    ux = x[nc:nc + nfx].reshape(n + 1, n, nz)
    uy = x[nc + nfx:nc + nfx + nfy].reshape(n, n + 1, nz)
    v = np.zeros((n, n, nz, 3))
    v[..., 0] = 0.5 * (ux[:-1] + ux[1:])
    v[..., 1] = 0.5 * (uy[:, :-1] + uy[:, 1:])
    if dim == 3:
        uz = x[nc + nfx + nfy:].reshape(n, n, n + 1)
        v[..., 2] = 0.5 * (uz[:, :, :-1] + uz[:, :, 1:])
    return v


def write_frame(out_dir, frame, dim, n, rho, x, fixed_cells, fixed_face_centres):
    """Write one frame like the reference's drivers (design rho, dof vector x)."""
    fd = os.path.join(out_dir, str(frame))
    os.makedirs(fd, exist_ok=True)
    os.makedirs(os.path.join(out_dir, "0"), exist_ok=True)
    nz = n if dim == 3 else 1
    counts = np.array([n, n, nz], "<i4")
    if frame == 0:
        with open(os.path.join(fd, "grid"), "wb") as f:
            f.write(np.array([n, n, n if dim == 3 else 0], "<i4").tobytes())
            f.write(np.array([1.0, 0.0, 0.0, 0.0], "<f8").tobytes())
    with open(os.path.join(fd, "x"), "wb") as f:
        f.write(counts.tobytes())
        f.write(np.ascontiguousarray(rho, "<f8").tobytes())
    with open(os.path.join(fd, "velocity"), "wb") as f:
        f.write(counts.tobytes())
        f.write(np.ascontiguousarray(cell_velocity(dim, n, x), "<f8").tobytes())
    for name, pts in (("psi_D", fixed_cells), ("psi_N", fixed_face_centres)):
        pts = np.asarray(pts, "<f8").reshape(-1, 3)
        m = pts.shape[0]
        with open(os.path.join(fd, name), "wb") as f:
            f.write(np.array([m], "<i4").tobytes())
            if m:
                f.write(pts.tobytes())
                f.write(np.zeros(6 * m + 2 * m, "<f8").tobytes())
                f.write(np.zeros(m, "<i4").tobytes())
    with open(os.path.join(out_dir, "0", "last_frame.txt"), "w") as f:
        f.write(str(frame))
