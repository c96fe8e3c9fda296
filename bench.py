#!/usr/bin/env python3
"""Benchmark: ms per Stokes solve to a 1e-8 true FP64 relative residual at 3D 256^3,
plain FP64 GMG-Krylov vs the mixed-precision (and learned) solve, on one MI355X.

A "step" is one solve of the 3D 256^3 channel problem (parabolic inflow on a
centred 1/4 patch of x=0, zero-traction outflow patch on x=1, no-slip elsewhere,
Brinkman alpha = 1e4*0.1(1-r)/(0.1+r)) on a seeded Gaussian-random-field density
of mean 0.2, including the design->alpha map and the multigrid hierarchy setup
that every design iteration repeats.  Inputs are resident in HBM before timing.

Prints one JSON line (see the keys below).  --gpus N with N > 1 is rejected:
this workload runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--mean", type=float, default=0.2)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--mp", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=1)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (HBM counter calibration)")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def run_path(args, prob, rho_dev, precision, profile):
    import torch
    from stokes_to_amd import _lib
    from stokes_to_amd.steady import FluidSteadyState

    s = FluidSteadyState(3, args.n, precision=precision, rtol=args.rtol)
    s.init_boundary(prob.fixed)
    s.update_b(dirichlet=prob.dirichlet)
    lib = s.lib
    ptr = ctypes.c_void_p(rho_dev.data_ptr())

    def step():
        rc = lib.st_set_design_device(s.h, ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double)))
        s._check(rc)
        _, st = s.solve(init="zero", fetch=False)
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if profile:
        lib.st_profile(s.h, 1)
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = {}
    if profile:
        lib.st_profile(s.h, 0)
        names = {0: "smooth", 1: "apply", 2: "residual"}
        for k, nm in names.items():
            tot, cnt, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
            lib.st_profile_read(s.h, k, ctypes.byref(tot), ctypes.byref(cnt), ctypes.byref(by))
            prof[nm] = {"total_ms": tot.value, "launches": cnt.value, "bytes": by.value}
    s.close()
    wall_ms = (t1 - t0) * 1e3 / args.steps
    return {
        "wall_ms": wall_ms,
        "event_ms": float(np.median([st["ms"] for st in stats])),
        "setup_ms": float(np.median([st["setup_ms"] for st in stats])),
        "iters": int(np.median([st["iters"] for st in stats])),
        "relres": float(max(st["relres"] for st in stats)),
        "converged": all(st["converged"] for st in stats),
        "prof": prof,
    }


def cpu_baseline(args, prob, rho, gpu_iters):
    """The oracle (CPU FP64 port of the same FGMRES + V-cycle) on the host cores:
    hierarchy setup + `cpu_iters` FGMRES iterations of the same 256^3 solve, scaled
    to a whole solve with the iteration count the FP64 GPU solve needed."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oracle.build()
    n, s = oracle.sizes(3, *prob.shape)
    nc = s[0]
    fa = oracle.design_to_alpha(3, prob.shape, rho, prob.alpha.array(), prob.fixed[nc:])
    pen = np.zeros(n)
    pen[nc:] = fa
    b = oracle.lift_rhs(3, prob.shape, prob.fixed, prob.dirichlet)
    t0 = time.perf_counter()
    mg = oracle.MG(3, prob.shape, prob.fixed, pen, nu_pre=2, nu_post=2, omega=1.0, tang=0)
    t1 = time.perf_counter()
    x0 = prob.dirichlet.copy()
    _, it, rr, _ = mg.fgmres(b, x0=x0, rtol=1e-30, restart=max(1, args.cpu_iters), maxit=args.cpu_iters)
    t2 = time.perf_counter()
    per_it = (t2 - t1) / max(it, 1)
    est = (t1 - t0) + per_it * gpu_iters
    return {
        "value": est * 1e3, "unit": "ms/solve", "cores": threads, "kind": "port",
        "sample": f"oracle FGMRES+V(2,2) FP64 on the same {args.n}^3 problem: hierarchy setup "
                  f"({t1 - t0:.2f} s) + {it} iteration(s) ({per_it:.2f} s each), scaled to the "
                  f"{gpu_iters} iterations of the FP64 GPU solve",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this workload runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    import torch
    from stokes_to_amd.problems import channel_problem, grf_density

    t = time.perf_counter()
    prob = channel_problem(3, args.n)
    rho = grf_density(3, args.n, mean=args.mean, sigma=8.0, seed=12345)
    log(args, f"[bench] problem {args.n}^3 ready in {time.perf_counter() - t:.1f}s")
    rho_dev = torch.from_numpy(rho).to("cuda")
    torch.cuda.synchronize()
    if args.copy_probe_mb > 0:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters calibrate
        # the counters read for the solver kernels in the same profiled process
        src = torch.ones(args.copy_probe_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    plain = run_path(args, prob, rho_dev, "fp64", profile=True)
    log(args, f"[bench] fp64: {plain['event_ms']:.2f} ms/solve, {plain['iters']} it, relres {plain['relres']:.2e}")
    mp = run_path(args, prob, rho_dev, args.mp, profile=True)
    log(args, f"[bench] {args.mp}: {mp['event_ms']:.2f} ms/solve, {mp['iters']} it, relres {mp['relres']:.2e}")

    head = mp if (mp["converged"] and mp["wall_ms"] < plain["wall_ms"]) else plain
    pr = plain["prof"]["smooth"]
    avg_ms = pr["total_ms"] / max(pr["launches"], 1)
    achieved = pr["bytes"] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    out = {
        "metric": "ms per Stokes solve to 1e-8 rel residual at 3D 256³: plain FP64 MG vs learned+MP",
        "value": round(head["wall_ms"], 3),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall_ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head is plain else f"f64+{args.mp}",
        "data": "synthetic",
        "config": {
            "workload": f"3D {args.n}^3 single Stokes/Brinkman solve (channel BCs, GRF density mean {args.mean}, "
                        f"seed 12345), rtol {args.rtol} true FP64 residual",
            "n": args.n, "dofs": int(prob.fixed.size), "rtol": args.rtol,
            "headline_path": "fp64" if head is plain else args.mp,
        },
        "plain_fp64": {k: plain[k] for k in ("wall_ms", "event_ms", "setup_ms", "iters", "relres", "converged")},
        "mixed": {k: mp[k] for k in ("wall_ms", "event_ms", "setup_ms", "iters", "relres", "converged")},
        "speedup_vs_plain": round(plain["wall_ms"] / head["wall_ms"], 3),
        "roofline": {
            "bound": "hbm", "kernel": "k_smooth<double,3> (finest-level colour pass, FP64 path)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": round(avg_ms, 4), "launches": pr["launches"],
            "algorithmic_bytes_per_launch": pr["bytes"],
        },
        "kernels": {"fp64": plain["prof"], args.mp: mp["prof"]},
    }
    if not args.no_cpu_baseline:
        log(args, "[bench] cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(args, prob, rho, plain["iters"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
