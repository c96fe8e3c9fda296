#!/usr/bin/env python3
"""Benchmark: ms per Stokes solve to a 1e-8 true FP64 relative residual at 3D 256^3,
plain FP64 GMG-Krylov vs the learned + mixed-precision solve, on one MI355X.

Default workload ("to", the 256^3 topology-optimization configuration): the 3D 256^3
channel problem (parabolic inflow on a centred 1/4 patch of x=0, zero-traction
outflow patch on x=1, no-slip elsewhere; Brinkman alpha = 1e4*0.1(1-r)/(0.1+r);
volume fraction 0.2, initial density 0.2, filter radius 1.5, OC move 0.2), run as
three independent TO runs of 80 design iterations each:
  * plain: FP64 multigrid, forward and adjoint solves from zero (the reference's
    GMRES starts from zero);
  * keep: FP32 (or BF16) multigrid inside FP64 refinement, solves warm-started from
    the run's previous solutions, no network (what the method gains without it);
  * learned: the same mixed precision with the learned initial guess; the network is
    trained online on (design, solution) pairs of the run's own first 15 iterations
    (training time reported separately).
The run length (80) and the training window (15) are fixed by the configuration.
A "step" is one forward solve; --steps K timed steps are K design iterations spread
evenly over the post-training window (all 65 by default), every earlier iteration is
untimed (at least --warmup of them).  The time of a solve is HIP events around
hierarchy setup, initial guess (network inference included) and FGMRES on the
solver stream, with every input resident in HBM.  `value` is the median learned+MP
forward solve of the timed steps; the final designs of the three runs are compared.
Kernel timings for `roofline` are taken after the runs on the plain and keep
contexts (the timed solves carry no profiling events).

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
# (profiles/solver_kernels256_hbm_counters.md)
SMOOTH256_TRAFFIC_BYTES = {"fp64": 2.314e9, "fp32": 1.227e9}
# the same for the one-pass two-colour sweep (filled in from its counter runs)
SWEEP256_TRAFFIC_BYTES = {}
METRIC = "ms per Stokes solve to 1e-8 rel residual at 3D 256³: plain FP64 MG vs learned+MP"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=65, help="timed design iterations (solves)")
    ap.add_argument("--warmup", type=int, default=15, help="untimed iterations (online training window)")
    ap.add_argument("--workload", default="to", choices=["to", "grf", "net", "to-warm"])
    ap.add_argument("--to-iters", type=int, default=30, help="design iterations of the to-warm workload")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--volfrac", type=float, default=0.2)
    ap.add_argument("--mean", type=float, default=0.2, help="GRF mean (grf workload)")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--mp", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--nu", type=int, nargs=2, default=None, metavar=("PRE", "POST"),
                    help="smoothing sweeps of the V-cycle (all contexts; default: the library's)")
    ap.add_argument("--warm-order", type=int, default=None, choices=[0, 1, 2, 3],
                    help="start of a sequence of solves (st_params.warm_order; default: the library's)")
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3, help="FGMRES iterations of the 256^3 CPU sample")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (HBM counter calibration)")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def sweeps(args):
    """(nu_pre, nu_post) of every context: --nu, else the library's default."""
    if args.nu is not None:
        return tuple(args.nu)
    import ctypes as ct
    from stokes_to_amd import _lib
    p = _lib.StParams()
    _lib.load().st_default_params(ct.byref(p), 3, args.n)
    return (p.nu_pre, p.nu_post)


def new_ctx(args, prob, precision):
    from stokes_to_amd.steady import FluidSteadyState
    kw = {} if args.warm_order is None else {"warm_order": args.warm_order}
    nu = sweeps(args)
    s = FluidSteadyState(3, args.n, precision=precision, rtol=args.rtol, nu_pre=nu[0], nu_post=nu[1], **kw)
    s.init_boundary(prob.fixed)
    s.update_b(dirichlet=prob.dirichlet)
    return s


def read_profile(s):
    prof = {}
    names = {0: "smooth", 1: "apply", 2: "residual", 3: "smooth0", 4: "sweep"}
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
        "init_ms_median": float(np.median([st["init_ms"] for st in stats])),
    }


def roofline_of(prof, n, prec):
    """Roofline of the finest-level smoother from the solver's own HIP-event launch
    timings: the one-pass two-colour sweep (`k_sweep`) where the path runs it, else the
    colour pass (`k_march<prec, SMOOTH>`)."""
    ctype = {"fp64": "double", "fp32": "float", "bf16": "bf16"}[prec]
    if prof.get("sweep", {}).get("launches", 0) > 0:
        pr, kname = prof["sweep"], f"k_sweep<{ctype}> ({n}^3 finest-level two-colour sweep, {prec} storage)"
        traffic = SWEEP256_TRAFFIC_BYTES.get(prec) if n == 256 else None
    else:
        pr, kname = prof["smooth"], f"k_march<{ctype},SMOOTH> ({n}^3 finest-level colour pass, {prec} storage)"
        traffic = SMOOTH256_TRAFFIC_BYTES.get(prec) if n == 256 else None
    avg_ms = pr["total_ms"] / max(pr["launches"], 1)
    achieved = pr["bytes"] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {
        "bound": "hbm", "kernel": kname,
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "avg_launch_ms": round(avg_ms, 4), "launches": pr["launches"],
        "algorithmic_bytes_per_launch": pr["bytes"],
    }


TO_ITERS, TO_TRAIN = 80, 15   # the 256^3 configuration: 80 OC iterations, online training on the first 15


def timed_iterations(args):
    """The K = --steps timed design iterations, spread evenly over the post-training
    window [max(15, --warmup), 80) so that the headline covers the whole run whatever K
    is; every iteration before the first timed one is untimed (>= --warmup of them)."""
    lo = max(TO_TRAIN, args.warmup)
    span = TO_ITERS - lo
    k = max(1, min(args.steps, span))
    if k == 1:
        return [lo]
    return sorted({lo + round(j * (span - 1) / (k - 1)) for j in range(k)})


def run_to(args, prob):
    """The 256^3 topology-optimization configuration as three independent runs of 80
    design iterations each (volume fraction 0.2, filter 1.5, OC):
      plain    FP64 multigrid, forward and adjoint solves from zero (the reference's
               GMRES starts from zero);
      keep     FP32/BF16 multigrid in FP64 refinement, solves warm-started from the
               run's previous solutions, no network;
      learned  the same mixed precision with the learned initial guess, the network
               trained online on the run's own first 15 iterations.
    Returns per-run statistics of the timed iterations, the final designs' differences
    and the profiled kernel timings (taken after the runs, on contexts not timed)."""
    import torch
    # the learned context is created before the warm-start one: the two do the same
    # solver work, and the context created last runs it a few per cent faster, so this
    # order gives the learned start no placement advantage
    P = new_ctx(args, prob, "fp64")
    L = new_ctx(args, prob, args.mp)
    K = new_ctx(args, prob, args.mp)
    rho0 = np.full(args.n ** 3, args.volfrac)
    for c in (P, K, L):
        c.set_design(rho0)
    timed = timed_iterations(args)
    tset = set(timed)
    st = {"plain": [], "keep": [], "learned": []}
    win = {"plain": [], "keep": [], "learned": []}
    adj = {"plain": [], "keep": [], "learned": []}
    train_ms, inf_ms, used, lossr, jdiff = [], [], [], [], []
    t_region = t_end = None
    for it in range(TO_ITERS):
        if it == timed[0]:
            torch.cuda.synchronize()
            t_region = time.perf_counter()
        rp = P.to_step(args.volfrac, init="zero")
        # the warm-start and learned runs alternate their order, so that neither is
        # always the one that follows the plain run on the GPU
        if it % 2 == 0:
            rk = K.to_step(args.volfrac, init="keep")
            rl = L.to_step(args.volfrac, init="learned")
        else:
            rl = L.to_step(args.volfrac, init="learned")
            rk = K.to_step(args.volfrac, init="keep")
        jdiff.append(abs(rl["J"] - rp["J"]) / abs(rp["J"]))
        for nm, r in (("plain", rp), ("keep", rk), ("learned", rl)):
            if it >= TO_TRAIN:
                win[nm].append(r["fwd"])
                adj[nm].append(r["adj"])
            if it in tset:
                st[nm].append(r["fwd"])
        if it in tset:
            inf_ms.append(rl["fwd"]["init_ms"])
            used.append(int(rl["fwd"]["learned_used"]))
        if it == timed[-1]:
            torch.cuda.synchronize()
            t_end = time.perf_counter()
        if it < TO_TRAIN:
            t0 = time.perf_counter()
            L.record_sample()
            L.net_train(args.epochs, args.lr)
            torch.cuda.synchronize()
            train_ms.append((time.perf_counter() - t0) * 1e3)
        elif it in tset:
            lo, lz = L.net_eval()  # untimed: the network's loss on this unseen pair vs predicting zero
            lossr.append(lo / lz if lz > 0 else float("nan"))
        log(args, f"[bench] it {it:2d} {'timed' if it in tset else 'train' if it < TO_TRAIN else '     '} "
                  f"J {rp['J']:.9g} | plain {rp['fwd']['iters']} it {rp['fwd']['ms']:.1f} ms | "
                  f"keep {rk['fwd']['iters']} it {rk['fwd']['ms']:.1f} ms r0 {rk['fwd']['relres0']:.1e} | "
                  f"learned {rl['fwd']['iters']} it {rl['fwd']['ms']:.1f} ms r0 {rl['fwd']['relres0']:.1e} "
                  f"used {rl['fwd']['learned_used']}")
    region_ms = (t_end - t_region) * 1e3
    dp, dk, dl = P.get_design(), K.get_design(), L.get_design()
    final = {
        "final_design_rel_l2_learned_vs_plain": float(np.linalg.norm(dl - dp) / np.linalg.norm(dp)),
        "final_design_rel_l2_keep_vs_plain": float(np.linalg.norm(dk - dp) / np.linalg.norm(dp)),
        "final_J": {"plain": rp["J"], "keep": rk["J"], "learned": rl["J"]},
        "objective_max_rel_diff_learned_vs_plain": float(max(jdiff)),
    }
    # untimed: the last design from a cold start, without and with the network
    cold = {}
    for mode in ("zero", "learned"):
        L.reset_solution()
        _, sc = L.solve(init=mode, fetch=False)
        cold[mode] = {"iters": int(sc["iters"]), "ms": round(sc["ms"], 3), "relres0": sc["relres0"],
                      "learned_used": int(sc["learned_used"])}
    log(args, f"[bench] cold start on the last design: zero {cold['zero']['iters']} it, learned "
              f"{cold['learned']['iters']} it (taken {cold['learned']['learned_used']})")
    # kernel timings (HIP events around every finest-level launch) on the plain and the
    # warm-start contexts after the runs: two solves of their last systems from zero
    prof = {}
    for nm, c in (("fp64", P), (args.mp, K)):
        torch.cuda.synchronize()
        c.lib.st_profile(c.h, 1)
        for _ in range(2):
            c.reset_solution()
            c.solve(init="zero", fetch=False)
        c.lib.st_profile(c.h, 0)
        prof[nm] = read_profile(c)
    last_rhof = L.filter(dl)
    for c in (P, K, L):
        c.close()
    extra = {"train_ms_per_iter": train_ms, "train_ms_total": float(sum(train_ms)),
             "inference_ms_median": float(np.median(inf_ms)) if inf_ms else 0.0,
             "learned_taken": int(sum(used)), "loss_vs_zero_median": float(np.median(lossr)) if lossr else None,
             "cold_start": cold, "region_ms": region_ms, "prof": prof, "last_design": last_rhof,
             "window": {nm: summary(v) for nm, v in win.items()},
             "adjoint_window": {nm: summary(v) for nm, v in adj.items()}, "timed": timed}
    extra.update(final)
    return {nm: summary(v) for nm, v in st.items()}, extra


def run_to_warm(args, prob):
    """One TO run with warm starts only (FP32/BF16 multigrid, extrapolated starts, no
    network) for --to-iters design iterations: the workload for kernel traces of whole
    design steps (forward solve, objective, adjoint solve, sensitivity, filter, OC)."""
    s = new_ctx(args, prob, args.mp)
    s.set_design(np.full(args.n ** 3, args.volfrac))
    rows = []
    for it in range(args.to_iters):
        t0 = time.perf_counter()
        o = s.to_step(args.volfrac, init="keep")
        step_ms = (time.perf_counter() - t0) * 1e3
        f, a = o["fwd"], o["adj"]
        rows.append({"it": it, "step_ms": round(step_ms, 2), "fwd_iters": int(f["iters"]), "fwd_ms": round(f["ms"], 2),
                     "adj_iters": int(a["iters"]), "adj_ms": round(a["ms"], 2), "J": o["J"]})
        log(args, f"[bench] it {it:2d} step {step_ms:.1f} ms | fwd {f['iters']} it {f['ms']:.1f} ms | "
                  f"adj {a['iters']} it {a['ms']:.1f} ms | J {o['J']:.9g}")
    s.close()
    return rows


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


def run_net(args, prob):
    """The learned part alone at n^3 (kernel profiles of the convolutions): two solves
    of seeded GRF designs recorded as samples, --epochs training epochs, then the loss
    on a third, unseen design and one prediction (timed with torch events)."""
    import torch
    from stokes_to_amd.problems import grf_density
    s = new_ctx(args, prob, args.mp)
    for seed in (1, 2):
        s.set_design(grf_density(3, args.n, mean=args.mean, sigma=8.0, seed=seed))
        s.solve(init="zero", fetch=False)
        s.record_sample()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = s.net_train(args.epochs, args.lr)
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t0) * 1e3
    s.set_design(grf_density(3, args.n, mean=args.mean, sigma=8.0, seed=3))
    _, st0 = s.solve(init="zero", fetch=False)
    lo, lz = s.net_eval()
    s.reset_solution()
    _, stl = s.solve(init="learned", fetch=False)
    s.close()
    return {"train_ms": t_train, "epochs": args.epochs, "samples": 2, "train_loss": loss, "eval_loss": lo,
            "eval_loss_zero": lz, "cold_zero_iters": st0["iters"], "cold_learned_iters": stl["iters"],
            "inference_and_check_ms": stl["init_ms"]}


def host_cores():
    """Cores this process may use (affinity / cgroup share, capped by OMP_NUM_THREADS
    when the host sets it), and the CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if omp > 0:
        n = min(n, omp)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, os.cpu_count(), model


def cpu_baseline(args, prob, rho, gpu_iters):
    """The oracle (the CPU FP64 restatement of the same FGMRES + V-cycle, OpenMP) on the
    host cores.  Headline: hierarchy setup + --cpu-iters FGMRES iterations of the same
    256^3 system, scaled to a whole solve with the iteration count of the FP64 GPU
    solve.  Beside it, whole CPU runs of the smaller configurations: the 2D 64^2 TO run
    (10 OC iterations) and single 64^3 / 128^3 GRF solves (mean 0.3) to 1e-8."""
    import oracle
    from stokes_to_amd.problems import channel_problem, grf_density
    threads, ncpu, model = host_cores()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oracle.build()

    def system(P, r):
        nn, sz = oracle.sizes(P.dim, *P.shape)
        fa = oracle.design_to_alpha(P.dim, P.shape, r, P.alpha.array(), P.fixed[sz[0]:])
        pen = np.zeros(nn)
        pen[sz[0]:] = fa
        return pen, oracle.lift_rhs(P.dim, P.shape, P.fixed, P.dirichlet)

    pen, b = system(prob, rho)
    t0 = time.perf_counter()
    nu = sweeps(args)
    mg = oracle.MG(3, prob.shape, prob.fixed, pen, nu_pre=nu[0], nu_post=nu[1], omega=1.0, tang=0)
    t1 = time.perf_counter()
    k = max(1, args.cpu_iters)
    _, it, rr, _ = mg.fgmres(b, x0=prob.dirichlet.copy(), rtol=1e-30, restart=k, maxit=k)
    t2 = time.perf_counter()
    del mg
    setup_s = t1 - t0
    per_it = (t2 - t1) / max(it, 1)
    est = setup_s + per_it * gpu_iters
    log(args, f"[bench] cpu {args.n}^3: setup {setup_s:.2f} s, {it} iterations {per_it:.2f} s each")
    small = {}
    for n in (64, 128):
        P = channel_problem(3, n)
        pen_n, b_n = system(P, grf_density(3, n, mean=0.3, sigma=8.0 * n / 64, seed=12345))
        t0 = time.perf_counter()
        m = oracle.MG(3, P.shape, P.fixed, pen_n, nu_pre=nu[0], nu_post=nu[1], omega=1.0, tang=0)
        _, itn, rrn, _ = m.fgmres(b_n, x0=P.dirichlet.copy(), rtol=1e-8)
        small[f"3d_{n}_grf_solve_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        small[f"3d_{n}_grf_iters"] = int(itn)
        log(args, f"[bench] cpu {n}^3 solve: {small[f'3d_{n}_grf_solve_ms']:.0f} ms, {itn} iterations")
    P2 = channel_problem(2, 64)
    r2 = np.full(64 * 64, 0.15)
    t0 = time.perf_counter()
    for _ in range(10):
        r2, J2, _, _, _ = oracle.to_step(P2, r2, volfrac=0.15, rtol=1e-8)
    small["2d_64_to10_s"] = round(time.perf_counter() - t0, 3)
    small["2d_64_to10_final_J"] = J2
    log(args, f"[bench] cpu 2D 64^2 TO, 10 iterations: {small['2d_64_to10_s']:.2f} s")
    return {
        "value": round(est * 1e3, 1), "unit": "ms/solve", "cores": threads, "kind": "port",
        "cpu_model": model, "host_cpu_count": ncpu,
        "sample": f"oracle FGMRES+V({nu[0]},{nu[1]}) FP64, OpenMP on {threads} threads (the cores this process may use), on "
                  f"the same {args.n}^3 system (the plain run's last design): hierarchy setup ({setup_s:.2f} s) "
                  f"+ {it} FGMRES iterations ({per_it:.2f} s each), scaled to the {gpu_iters} iterations of the "
                  f"FP64 GPU solve",
        "other_configurations": small,
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
    if args.workload == "to":
        st, extra = run_to(args, prob)
        plain, keep, learned = st["plain"], st["keep"], st["learned"]
        value = learned["median_ms"]
        timed = extra["timed"]
        out.update({
            "value": round(value, 3), "ms_per_step": round(value, 3), "dtype": f"f64+{args.mp}",
            "config": {
                "workload": f"3D {args.n}^3 topology optimization, 80 OC iterations (channel BCs, volume fraction "
                            f"{args.volfrac}, filter radius 1.5): forward solves of {len(timed)} design iterations "
                            f"spread over iterations {timed[0]}..{timed[-1]} (after the 15-iteration training "
                            f"window), rtol {args.rtol} true FP64 residual",
                "n": args.n, "dofs": int(prob.fixed.size), "rtol": args.rtol, "mp": args.mp,
                "design_iterations": TO_ITERS, "train_iterations": TO_TRAIN, "epochs_per_iteration": args.epochs,
                "timed_iterations": timed,
            },
            "plain_fp64": plain, "learned_mp": learned, "keep_mp": keep,
            "speedup_vs_plain": round(plain["median_ms"] / value, 3),
            "speedup_keep_vs_plain": round(plain["median_ms"] / keep["median_ms"], 3),
            "learned_vs_keep": round(keep["median_ms"] / value, 3),
            "window_all_post_training": extra["window"],
            "adjoint_window": extra["adjoint_window"],
            "final_design_rel_l2_learned_vs_plain": extra["final_design_rel_l2_learned_vs_plain"],
            "final_design_rel_l2_keep_vs_plain": extra["final_design_rel_l2_keep_vs_plain"],
            "objective_max_rel_diff_learned_vs_plain": extra["objective_max_rel_diff_learned_vs_plain"],
            "final_objective": extra["final_J"],
            "learned": {k: extra[k] for k in ("train_ms_total", "train_ms_per_iter", "inference_ms_median",
                                              "learned_taken", "loss_vs_zero_median", "cold_start")},
            "timed_region_wall_ms": round(extra["region_ms"], 1),
            "timed_region_note": "wall time from the first to the last timed iteration; every iteration runs "
                                 "the three TO runs (forward and adjoint solves, OC, host copies)",
            "roofline": roofline_of(extra["prof"][args.mp], args.n, args.mp),
            "roofline_plain_fp64": roofline_of(extra["prof"]["fp64"], args.n, "fp64"),
            "kernels_plain_fp64": extra["prof"]["fp64"], f"kernels_{args.mp}": extra["prof"][args.mp],
        })
        base_design, base_iters = extra["last_design"], int(np.median(plain["iters"]))
        log(args, f"[bench] plain {plain['median_ms']:.1f} ms, keep+{args.mp} {keep['median_ms']:.1f} ms, "
                  f"learned+{args.mp} {learned['median_ms']:.1f} ms (median of {len(timed)}); final designs "
                  f"{extra['final_design_rel_l2_learned_vs_plain']:.2e} (learned vs plain)")
    elif args.workload == "to-warm":
        rows = run_to_warm(args, prob)
        late = rows[len(rows) // 2:]
        value = float(np.median([r["step_ms"] for r in late]))
        out.update({"metric": "ms per warm-started design step (forward + adjoint solve, objective, sensitivity, "
                              "filter, OC)", "value": round(value, 3), "ms_per_step": round(value, 3),
                    "dtype": f"f64+{args.mp}", "config": {"workload": f"3D {args.n}^3 TO, {args.to_iters} OC "
                                                                    f"iterations, warm starts", "n": args.n},
                    "iterations": rows, "roofline": None, "cpu_baseline": None})
        print(json.dumps(out), flush=True)
        return 0
    elif args.workload == "net":
        r = run_net(args, prob)
        value = r["train_ms"] / max(args.epochs, 1) / 2
        out.update({"metric": "ms per training sample-epoch of the learned predictor", "value": round(value, 3),
                    "ms_per_step": round(value, 3), "dtype": "bf16+f32", "higher_is_better": False,
                    "config": {"workload": f"learned predictor at {args.n}^3 (network grid {args.n // 4}^3)",
                               "n": args.n}, "net": r, "roofline": None, "cpu_baseline": None})
        print(json.dumps(out), flush=True)
        return 0
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
