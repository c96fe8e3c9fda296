#!/usr/bin/env python3
"""Benchmark: ms per Stokes solve to a 1e-8 true FP64 relative residual at 3D 256^3,
plain FP64 GMG-Krylov vs the learned + mixed-precision solve, on one MI355X.

Default workload ("to"): a density-filtered topology-optimization run on the 3D
256^3 channel problem (parabolic inflow on a centred 1/4 patch of x=0,
zero-traction outflow patch on x=1, no-slip elsewhere; Brinkman
alpha = 1e4*0.1(1-r)/(0.1+r); volume fraction 0.2, initial density 0.2, filter
radius 1.5, OC move 0.2).  The design iterations are driven by the learned+MP
context (FP32 or BF16 multigrid inside FP64 FGMRES, forward solve from the learned
initial guess, adjoint from the previous adjoint).  At every iteration the same
filtered design is also solved by
  * the plain path: FP64 multigrid, zero initial guess (the reference's GMRES
    starts from zero), and
  * the warm-start path: the same mixed precision, previous solution as initial
    guess, no network (what the method gains without the network).
The first --warmup iterations (15) are the online training window: after each,
the solver pair (design, coarse state) is recorded and the network trained for
--epochs epochs (training time reported separately).  The next --steps
iterations (65; 80 design iterations in all, as in the 256^3 configuration) are
timed with the network frozen.  A "step" is one forward solve;
the time of a solve is HIP events around hierarchy setup, initial guess
(network inference and its residual check) and FGMRES on the solver stream,
with every input resident in HBM.

Workload "grf": a single solve of a seeded Gaussian-random-field density (mean
0.2) repeated --steps times, FP64 vs the mixed-precision path (kernel profile).

Prints one JSON line.  --gpus N with N > 1 is rejected: this workload runs on
one GPU.
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
# measured HBM traffic per launch of the 256^3 smoother colour pass: rocprofv3
# FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB read / write of known size
# (profiles/k_march_smooth256_counters.md)
SMOOTH256_TRAFFIC_BYTES = {"fp64": 2.761e9, "fp32": 1.472e9}
METRIC = "ms per Stokes solve to 1e-8 rel residual at 3D 256³: plain FP64 MG vs learned+MP"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=65, help="timed design iterations (solves)")
    ap.add_argument("--warmup", type=int, default=15, help="untimed iterations (online training window)")
    ap.add_argument("--workload", default="to", choices=["to", "grf", "to-pair"])
    ap.add_argument("--iters", type=int, default=80, help="design iterations of the to-pair workload")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--volfrac", type=float, default=0.2)
    ap.add_argument("--mean", type=float, default=0.2, help="GRF mean (grf workload)")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--mp", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=1)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (HBM counter calibration)")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def new_ctx(args, prob, precision):
    from stokes_to_amd.steady import FluidSteadyState
    s = FluidSteadyState(3, args.n, precision=precision, rtol=args.rtol)
    s.init_boundary(prob.fixed)
    s.update_b(dirichlet=prob.dirichlet)
    return s


def read_profile(s):
    prof = {}
    names = {0: "smooth", 1: "apply", 2: "residual"}
    for k, nm in names.items():
        tot, cnt, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        s.lib.st_profile_read(s.h, k, ctypes.byref(tot), ctypes.byref(cnt), ctypes.byref(by))
        prof[nm] = {"total_ms": tot.value, "launches": cnt.value, "bytes": by.value}
    return prof


def summary(stats):
    ms = [st["ms"] for st in stats]
    return {
        "median_ms": float(np.median(ms)), "mean_ms": float(np.mean(ms)),
        "iters": [int(st["iters"]) for st in stats],
        "relres_max": float(max(st["relres"] for st in stats)),
        "converged": all(bool(st["converged"]) for st in stats),
        "setup_ms_median": float(np.median([st["setup_ms"] for st in stats])),
    }


def roofline_of(prof, n, prec):
    """Roofline of the finest-level smoother colour pass (`k_march<prec, SMOOTH>`) from
    the solver's own HIP-event launch timings."""
    pr = prof["smooth"]
    avg_ms = pr["total_ms"] / max(pr["launches"], 1)
    achieved = pr["bytes"] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    ctype = {"fp64": "double", "fp32": "float", "bf16": "bf16"}[prec]
    return {
        "bound": "hbm", "kernel": f"k_march<{ctype},SMOOTH> ({n}^3 finest-level colour pass, {prec} storage)",
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": SMOOTH256_TRAFFIC_BYTES.get(prec) if n == 256 else None,
        "avg_launch_ms": round(avg_ms, 4), "launches": pr["launches"],
        "algorithmic_bytes_per_launch": pr["bytes"],
    }


def run_to(args, prob):
    """The topology-optimization workload; returns (plain, learned, keep, extra)."""
    import torch
    L = new_ctx(args, prob, args.mp)
    Z = new_ctx(args, prob, "fp64")
    K = new_ctx(args, prob, args.mp)
    rho0 = np.full(args.n ** 3, args.volfrac)
    L.set_design(rho0)
    plain, learned, keep, train_ms, inf_ms, used, lossr = [], [], [], [], [], [], []
    total = args.warmup + args.steps
    t_region = None
    for it in range(total):
        timed = it >= args.warmup
        if it == args.warmup:
            torch.cuda.synchronize()
            Z.lib.st_profile(Z.h, 1)
            K.lib.st_profile(K.h, 1)   # the warm-start context runs the same MP kernels as L
            t_region = time.perf_counter()
        rho = L.get_design()
        rhof = L.filter(rho)
        Z.set_design(rhof)
        K.set_design(rhof)
        _, sz = Z.solve(init="zero", fetch=False)
        _, sk = K.solve(init="keep", fetch=False)
        r = L.to_step(args.volfrac, init="learned")
        sl = r["fwd"]
        if timed:
            plain.append(sz)
            keep.append(sk)
            learned.append(sl)
            inf_ms.append(sl["init_ms"])
            used.append(int(sl["learned_used"]))
            lo, lz = L.net_eval()  # untimed: the network's loss on this unseen pair vs predicting zero
            lossr.append(lo / lz if lz > 0 else float("nan"))
        else:
            t0 = time.perf_counter()
            L.record_sample()
            L.net_train(args.epochs, args.lr)
            torch.cuda.synchronize()
            train_ms.append((time.perf_counter() - t0) * 1e3)
        log(args, f"[bench] it {it:2d} {'timed' if timed else 'train'} change {r['change']:.3f} J {r['J']:.6g} | "
                  f"plain {sz['iters']} it {sz['ms']:.1f} ms | keep {sk['iters']} it {sk['ms']:.1f} ms "
                  f"r0 {sk['relres0']:.1e} | learned {sl['iters']} it {sl['ms']:.1f} ms r0 {sl['relres0']:.1e} "
                  f"used {sl['learned_used']}")
    torch.cuda.synchronize()
    region_ms = (time.perf_counter() - t_region) * 1e3 if t_region is not None else 0.0
    # untimed: the last design solved from a cold start, without and with the network
    # (no previous solution held: the learned start is x_D + P P net(rho))
    cold = {}
    for mode in ("zero", "learned"):
        L.reset_solution()
        _, sc = L.solve(init=mode, fetch=False)
        cold[mode] = {"iters": int(sc["iters"]), "ms": round(sc["ms"], 3), "relres0": sc["relres0"],
                      "learned_used": int(sc["learned_used"])}
    log(args, f"[bench] cold start on the last design: zero {cold['zero']['iters']} it "
              f"(r0 {cold['zero']['relres0']:.2e}), learned {cold['learned']['iters']} it "
              f"(r0 {cold['learned']['relres0']:.2e}, taken {cold['learned']['learned_used']})")
    Z.lib.st_profile(Z.h, 0)
    K.lib.st_profile(K.h, 0)
    prof = read_profile(Z)
    prof_mp = read_profile(K)
    last_rhof = L.filter(L.get_design())
    extra = {"train_ms_per_iter": train_ms, "train_ms_total": float(sum(train_ms)),
             "inference_ms_median": float(np.median(inf_ms)) if inf_ms else 0.0,
             "learned_taken": int(sum(used)), "prof": prof, "prof_mp": prof_mp, "last_design": last_rhof,
             "loss_vs_zero_median": float(np.median(lossr)) if lossr else None, "cold_start": cold,
             "final_J": r["J"], "final_vol": r["vol"], "region_ms": region_ms}
    for s in (L, Z, K):
        s.close()
    return summary(plain), summary(learned), summary(keep), extra


def run_to_pair(args, prob):
    """Two independent TO runs of --iters iterations: plain FP64 (zero starts) and
    learned + MP (online training on the first --warmup iterations); per-iteration
    forward-solve times and the L2 difference of the final designs."""
    import torch
    P = new_ctx(args, prob, "fp64")
    L = new_ctx(args, prob, args.mp)
    rho0 = np.full(args.n ** 3, args.volfrac)
    P.set_design(rho0)
    L.set_design(rho0)
    tp, tl = [], []
    for it in range(args.iters):
        rp = P.to_step(args.volfrac, init="zero")
        rl = L.to_step(args.volfrac, init="learned")
        tp.append(rp["fwd"]["ms"])
        tl.append(rl["fwd"]["ms"])
        if it < args.warmup:
            L.record_sample()
            L.net_train(args.epochs, args.lr)
        log(args, f"[bench] it {it:2d} plain {rp['fwd']['iters']} it {rp['fwd']['ms']:.1f} ms J {rp['J']:.8g} | "
                  f"learned+{args.mp} {rl['fwd']['iters']} it {rl['fwd']['ms']:.1f} ms J {rl['J']:.8g}")
    torch.cuda.synchronize()
    dp, dl = P.get_design(), L.get_design()
    diff = float(np.linalg.norm(dp - dl) / np.linalg.norm(dp))
    P.close()
    L.close()
    return {"plain_ms": tp, "learned_ms": tl, "final_design_rel_l2": diff,
            "final_J_plain": rp["J"], "final_J_learned": rl["J"]}


def run_grf(args, prob):
    import torch
    from stokes_to_amd.problems import grf_density
    rho = grf_density(3, args.n, mean=args.mean, sigma=8.0, seed=12345)
    rho_dev = torch.from_numpy(rho).to("cuda")
    out = {}
    for prec in ("fp64", args.mp):
        s = new_ctx(args, prob, prec)
        ptr = ctypes.cast(ctypes.c_void_p(rho_dev.data_ptr()), ctypes.POINTER(ctypes.c_double))
        stats = []
        for i in range(1 + args.steps):
            if i == 1:
                torch.cuda.synchronize()
                s.lib.st_profile(s.h, 1)
            s._check(s.lib.st_set_design_device(s.h, ptr))
            _, st = s.solve(init="zero", fetch=False)
            log(args, f"[bench] grf {prec} solve {i}: {st['iters']} it, {st['ms']:.1f} ms, relres {st['relres']:.2e}")
            if i >= 1:
                stats.append(st)
        torch.cuda.synchronize()
        s.lib.st_profile(s.h, 0)
        out[prec] = (summary(stats), read_profile(s))
        s.close()
    return out, rho


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
        "sample": f"oracle FGMRES+V(2,2) FP64 on the same {args.n}^3 system (the plain path's last design): "
                  f"hierarchy setup ({t1 - t0:.2f} s) + {it} iteration(s) ({per_it:.2f} s each), scaled to "
                  f"the {gpu_iters} iterations of the FP64 GPU solve",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this workload runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    import torch
    from stokes_to_amd.problems import channel_problem

    prob = channel_problem(3, args.n)
    torch.zeros(1, device="cuda")
    if args.copy_probe_mb > 0:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters calibrate
        # the counters read for the solver kernels in the same profiled process
        src = torch.ones(args.copy_probe_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        float(src.sum())          # a pure read of the same size
        torch.cuda.synchronize()
        del src, dst

    out = {"metric": METRIC, "unit": "ms/solve", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    if args.workload == "to-pair":
        r = run_to_pair(args, prob)
        tail = slice(args.warmup, None)
        value = float(np.median(r["learned_ms"][tail]))
        out.update({
            "value": round(value, 3), "ms_per_step": round(value, 3), "dtype": f"f64+{args.mp}",
            "config": {"workload": f"two independent 3D {args.n}^3 TO runs of {args.iters} iterations (plain FP64 "
                                   f"zero start vs learned+{args.mp}), volume fraction {args.volfrac}",
                       "n": args.n, "iters": args.iters},
            "plain_median_ms": float(np.median(r["plain_ms"][tail])),
            "speedup_vs_plain": round(float(np.median(r["plain_ms"][tail])) / value, 3),
            "final_design_rel_l2": r["final_design_rel_l2"],
            "final_J_plain": r["final_J_plain"], "final_J_learned": r["final_J_learned"],
            "plain_ms": r["plain_ms"], "learned_ms": r["learned_ms"], "roofline": None, "cpu_baseline": None,
        })
        print(json.dumps(out), flush=True)
        return 0
    if args.workload == "to":
        plain, learned, keep, extra = run_to(args, prob)
        value = learned["median_ms"]
        out.update({
            "value": round(value, 3), "ms_per_step": round(value, 3), "dtype": f"f64+{args.mp}",
            "config": {
                "workload": f"3D {args.n}^3 topology optimization (channel BCs, volume fraction {args.volfrac}, "
                            f"filter radius 1.5, OC): forward solves of design iterations "
                            f"{args.warmup}..{args.warmup + args.steps - 1}, rtol {args.rtol} true FP64 residual",
                "n": args.n, "dofs": int(prob.fixed.size), "rtol": args.rtol, "mp": args.mp,
                "train_iterations": args.warmup, "epochs_per_iteration": args.epochs,
            },
            "plain_fp64": plain, "learned_mp": learned, "keep_mp": keep,
            "speedup_vs_plain": round(plain["median_ms"] / value, 3),
            "speedup_keep_vs_plain": round(plain["median_ms"] / keep["median_ms"], 3),
            "learned": {k: extra[k] for k in ("train_ms_total", "train_ms_per_iter", "inference_ms_median",
                                              "learned_taken", "loss_vs_zero_median", "cold_start")},
            "final_objective": extra["final_J"], "final_volume": extra["final_vol"],
            "timed_region_wall_ms": round(extra["region_ms"], 1),
            "timed_region_note": "wall time of the timed iterations, each solving the design three times "
                                 "(plain, warm start, learned) plus the adjoint, OC and host copies",
            "roofline": roofline_of(extra["prof_mp"], args.n, args.mp),
            "roofline_plain_fp64": roofline_of(extra["prof"], args.n, "fp64"),
            "kernels_plain_fp64": extra["prof"], f"kernels_{args.mp}": extra["prof_mp"],
        })
        base_design, base_iters = extra["last_design"], int(np.median(plain["iters"]))
        log(args, f"[bench] plain {plain['median_ms']:.1f} ms, keep+{args.mp} {keep['median_ms']:.1f} ms, "
                  f"learned+{args.mp} {learned['median_ms']:.1f} ms (median of {args.steps})")
    else:
        res, rho = run_grf(args, prob)
        p, pp = res["fp64"]
        m, mp = res[args.mp]
        value = m["median_ms"]
        out.update({
            "value": round(value, 3), "ms_per_step": round(value, 3), "dtype": f"f64+{args.mp}",
            "config": {"workload": f"3D {args.n}^3 single Stokes/Brinkman solve (channel BCs, GRF density mean "
                                   f"{args.mean}, seed 12345), rtol {args.rtol} true FP64 residual",
                       "n": args.n, "dofs": int(prob.fixed.size), "rtol": args.rtol, "mp": args.mp},
            "plain_fp64": p, "mixed": m, "speedup_vs_plain": round(p["median_ms"] / value, 3),
            "roofline": roofline_of(mp, args.n, args.mp), "roofline_plain_fp64": roofline_of(pp, args.n, "fp64"),
            "kernels": {"fp64": pp, args.mp: mp},
        })
        base_design, base_iters = rho, int(np.median(p["iters"]))
        log(args, f"[bench] fp64 {p['median_ms']:.1f} ms, {args.mp} {m['median_ms']:.1f} ms")
    if not args.no_cpu_baseline:
        log(args, "[bench] cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(args, prob, base_design, base_iters)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
