"""Builds libstokes_hip.so in-tree with hipcc for gfx950.

Run as `python -m stokes_to_amd.build` or through __graft_entry__.build().
The library is plain HIP + C++ behind the C ABI of include/stokes_to.h; no
torch extension is involved.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstokes_hip.so")
DRIVER = os.path.join(HERE, "fluid_topo")
SOURCES = ["st_kernels.hip", "st_march.hip", "st_solver.hip", "st_net.hip", "st_mma.hip", "st_noalloc.hip", "st_sweep.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ST_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep packed-FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "warning: unknown" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(jobs=4, verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out + "\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_driver()
    return LIB


BINDING_TEST = os.path.join(HERE, "binding_test0")


def _host_program(src, out):
    """Host C++ on the C ABI only (g++), linked against the library next to it."""
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src, "-o", tmp, f"-L{HERE}",
          "-l:libstokes_hip.so", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_driver():
    """The fluid_topo command-line driver and the program that runs the reference's
    FluidSteadyState3D call sequence through the C++ binding (hip_steady_state3d.h)."""
    _host_program(os.path.join(HERE, "driver", "fluid_topo.cpp"), DRIVER)
    _host_program(os.path.join(HERE, "driver", "binding_test0.cpp"), BINDING_TEST)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
