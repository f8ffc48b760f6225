"""Benchmark: Munknowns/s to relative residual 1e-10 on one MI355X.

Workload (default): the unit-square Poisson problem on 8193 x 8193 nodes with
64 seeded Gaussian bumps as right-hand side and g = sin(3x) cosh(3y) on the
boundary.  One step = one complete solve from the zero interior guess until
||r||_2 <= 1e-10 ||r_0||_2, with the problem already resident in HBM (the
iterate is restored on the device between steps, outside the timed region).

Reported: the mixed-precision solver (FP32 V(1,3) cycles inside FP64
iterative refinement, at most two FP32 cycles per refinement step; --cycle
picks another V(nu1,nu2), --inner the cap) as ``value``; the
plain FP64 multigrid with the same cycle on the same
problem beside it (``fp64``); the same two solves with config 3's literal
V(2,2) cycle (``v22``); the dominant kernel's roofline from hipEvent timings
of a profiled solve; the HBM traffic of one whole solve of each precision
(``solve_traffic``, from the committed rocprofv3 PMC summary) against this
run's solve times; and the oracle's CPU multigrid on the same workload on
every CPU the job may use.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--seg 8192]
    python bench.py --only mixed|fp64 --steps 1 --warmup 1 --no-cpu --no-profile --no-v22
        (one precision only: the run rocprofv3 --pmc profiles for solve_traffic)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
PMC_SUMMARY = os.path.join(ROOT, "profiles", "poisson8193_level0_pmc.md")
SOLVE_PMC_SUMMARY = os.path.join(ROOT, "profiles", "poisson8193_solve_pmc.md")
# launch classes that run on the FP32 hierarchy in a mixed-precision solve
FP32_CLASSES_IN_MIXED = {"down_fused", "down_fused_resnorm64", "up_fused", "up_fused_refine", "up_down_fused", "up_lds",
                         "coarse_vcycle", "rb_half", "residual", "restrict", "prolong", "fill"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solve_times(scheme, steps, warmup, rtol):
    for _ in range(warmup):
        scheme.restore_iterate()
        scheme.solve(rtol, 100)
    wall, dev, cycles, rel, vcyc = [], [], [], [], []
    for _ in range(steps):
        scheme.restore_iterate()  # device copy, before the clock starts
        t0 = time.perf_counter()
        st = scheme.solve(rtol, 100)  # returns after the final stream sync
        wall.append(time.perf_counter() - t0)
        dev.append(st["seconds"])
        cycles.append(st["cycles"])
        rel.append(st["rel_residual"])
        vcyc.append(st["vcycles"])
    solve_times.vcycles = vcyc
    return wall, dev, cycles, rel


def profile_solve(p, mixed, rtol, nu, inner=1):
    """One solve with every launch bracketed by hipEvents; returns per
    (kernel, level) launches / ms / algorithmic bytes."""
    from vlp2d_amd import MultiGridSchemeHIP, Options
    from vlp2d_amd._lib import FLAG_PROFILE
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f,
                           options=Options.performance(mixed, nu1=nu[0], nu2=nu[1], flags=FLAG_PROFILE,
                                                       max_inner=inner if mixed else 1))
    s.un0[:] = p.initial("zero")
    s.init_after_boundaries_and_initial_iteration_inited()
    s.save_iterate()
    s.solve(rtol, 100)  # warm
    s.restore_iterate()
    s.reset_kernel_times()
    s.solve(rtol, 100)
    kt = s.kernel_times()
    s.cleanup()
    return kt


def cpu_baseline(seg, rtol, nu):
    """Oracle (C++ restatement of the reference CPU multigrid, OpenMP over
    rows) on the same formulas: FP64 V(nu1,nu2) to rtol.  Returns the
    baseline record and the oracle's solution (the FP64 oracle the GPU result
    is compared with)."""
    from oracle import Oracle
    from vlp2d_amd import problems
    p = problems.config(3, seg)
    host = host_cpus()
    cores = host["usable"]
    o = Oracle(p.seg_x, p.seg_y, p.hx, p.hy, p.f, "double", nthreads=cores)
    o.set_u(p.initial("zero"))
    t0 = time.perf_counter()
    k, hist = o.plain_solve(nu[0], nu[1], 3, rtol, 100)
    dt = time.perf_counter() - t0
    u = o.get_u()
    o.close()
    rec = {"value": round(p.unknowns / dt / 1e6, 2), "unit": "Munknowns/s", "cores": cores, "kind": "port",
           "sample": f"oracle FP64 V({nu[0]},{nu[1]}) RB-GS multigrid on the {seg + 1}x{seg + 1} workload, "
                     f"{k} cycles to rel. residual {hist[-1]:.2e} in {dt:.2f} s, one OpenMP thread per usable CPU",
           "nproc": host["nproc"], "affinity": host["affinity"], "cgroup_cpus": host["cgroup"],
           "cpu_model": host["model"]}
    return rec, u


def host_cpus():
    """CPUs of the host and the share this job may use: the affinity mask,
    capped by the cgroup CPU quota when one is set (a quota below the mask
    makes more threads time-share the same CPU time)."""
    nproc = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) // int(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = max(1, min(aff, quota) if quota else aff)
    return {"nproc": nproc, "affinity": aff, "cgroup": quota, "model": model, "usable": usable}


def stream_ceiling(n):
    """Bandwidth the GPU's HBM delivers to plain streaming kernels (torch's
    elementwise copy and add, and a sum) over FP64 arrays of n elements each:
    the practical ceiling for a kernel with the same read:write mix, below the
    published 8 TB/s.  Measured in a child process (torch's HIP runtime and
    the solver library's are separate libraries, kept out of one process)."""
    import subprocess
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--ceiling-only", str(n)], check=True,
                         capture_output=True, text=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def _stream_ceiling_child(n, reps=20):
    """Best of reps launches of each streaming kernel, timed with events on
    torch's stream (where these kernels run)."""
    import torch
    dev = torch.device("cuda", 0)
    a = torch.rand(n, dtype=torch.float64, device=dev)
    b = torch.rand(n, dtype=torch.float64, device=dev)
    c = torch.empty_like(a)
    nb = 8.0 * n
    cases = {"copy_1r1w": (lambda: c.copy_(a), 2 * nb), "add_2r1w": (lambda: torch.add(a, b, out=c), 3 * nb),
             "sum_1r": (lambda: a.sum(), nb)}
    out = {}
    for key, (fn, nbytes) in cases.items():
        for _ in range(3):
            fn()
        best = None
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        out[key] = round(nbytes / (best * 1e-3) / 1e9, 1)
    del a, b, c
    torch.cuda.empty_cache()
    out["unit"] = "GB/s"
    out["arrays"] = f"{n} FP64 elements each ({8 * n / 1e6:.0f} MB)"
    return out


def dd128_check(nu, seg=512, rtol=1e-10, inner=1):
    """Errors of the GPU mixed-precision and FP64 solutions (to rtol) of the
    513^2 replica against the oracle's double-double multigrid converged to
    its floor: (mixed, fp64)."""
    import numpy as np
    from oracle import Oracle
    from vlp2d_amd import MultiGridSchemeHIP, Options, problems
    p = problems.config(3, seg)
    o = Oracle(p.seg_x, p.seg_y, p.hx, p.hy, p.f, "dd")
    o.set_u(p.initial("zero"))
    for _ in range(60):
        if o.vcycle(3, 3, 3) < 1e-26:
            break
    hi, lo = o.get_u(with_lo=True)
    o.close()
    err = []
    for mixed in (True, False):
        s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f,
                               options=Options.performance(mixed, nu1=nu[0], nu2=nu[1],
                                                           max_inner=inner if mixed else 1))
        s.un0[:] = p.initial("zero")
        s.init_after_boundaries_and_initial_iteration_inited()
        s.solve(rtol, 100)
        u = s.get_array()
        s.cleanup()
        err.append(float(np.abs((u - hi) - lo).max() / np.abs(hi).max()))
    return tuple(err)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seg", type=int, default=8192)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--cpu-seg", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--cycle", default="1,3",
                    help="nu1,nu2 of the V-cycle (V(1,3): fewest cycles to 1e-10 on this workload, measured)")
    ap.add_argument("--inner", type=int, default=2,
                    help="mixed: at most this many FP32 V-cycles per FP64 refinement step")
    ap.add_argument("--only", choices=("mixed", "fp64"),
                    help="time one precision only (the runs rocprofv3 --pmc profiles)")
    ap.add_argument("--no-v22", action="store_true", help="skip the V(2,2) sub-record")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the plain-stream bandwidth ceiling")
    ap.add_argument("--ceiling-only", type=int, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.ceiling_only:
        print(json.dumps(_stream_ceiling_child(args.ceiling_only)), flush=True)
        return
    nu = tuple(int(x) for x in args.cycle.split(","))
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multigrid path runs on one MI355X)", file=sys.stderr)
        sys.exit(2)

    from vlp2d_amd import MultiGridSchemeHIP, Options, problems

    t0 = time.perf_counter()
    p = problems.config(3, args.seg)
    log(f"problem {p.name}: {p.shape} nodes, {p.unknowns} unknowns ({time.perf_counter() - t0:.1f}s to generate)")

    def timed(mixed, cyc_nu, keep_u=True):
        s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f,
                               options=Options.performance(mixed, nu1=cyc_nu[0], nu2=cyc_nu[1],
                                                           max_inner=args.inner if mixed else 1))
        s.un0[:] = p.initial("zero")
        s.init_after_boundaries_and_initial_iteration_inited()
        s.save_iterate()
        wall, dev, cyc, rel = solve_times(s, args.steps, args.warmup, args.rtol)
        r = {"wall": wall, "dev": dev, "cycles": cyc, "rel": rel, "info": s.info(),
             "vcycles": solve_times.vcycles[0], "u": s.get_array().copy() if keep_u else None}
        log(f"{'mixed' if mixed else 'fp64 '} V({cyc_nu[0]},{cyc_nu[1]}): steps {cyc}, V-cycles "
            f"{solve_times.vcycles[0]}, wall ms {[round(w * 1e3, 3) for w in wall]}, rel {max(rel):.2e}")
        s.cleanup()
        return r

    if args.only:
        r = timed(args.only == "mixed", nu, keep_u=False)
        print(json.dumps({"only": args.only, "ms_per_step": round(1e3 * float(np.mean(r["wall"])), 4),
                          "cycles": r["cycles"][0], "vcycles": r["vcycles"]}), flush=True)
        return

    res = {mixed: timed(mixed, nu) for mixed in (True, False)}

    mixed_ms = 1e3 * float(np.mean(res[True]["wall"]))
    fp64_ms = 1e3 * float(np.mean(res[False]["wall"]))
    value = p.unknowns / (mixed_ms * 1e-3) / 1e6

    roofline = None
    kernels = None
    if not args.no_profile:
        kt = profile_solve(p, True, args.rtol, nu, args.inner)
        tot = sum(e[3] for e in kt)
        dom = max(kt, key=lambda e: e[3])
        name, lvl, launches, ms, nbytes = dom
        per_launch_bytes = nbytes / launches
        avg_s = ms * 1e-3 / launches
        achieved = per_launch_bytes / avg_s / 1e9
        key = f"{name}@level{lvl}"
        prec = "f32" if name in FP32_CLASSES_IN_MIXED else "f64"
        traffic = measured_traffic(f"{key}/{prec}")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": f"{key}/{prec}", "avg_launch_us": round(avg_s * 1e6, 3),
                    "alg_bytes_per_launch": per_launch_bytes, "share_of_solve": round(ms / tot, 3)}
        if traffic:
            roofline["traffic_over_alg"] = round(traffic / per_launch_bytes, 4)
            roofline["measured_GBps"] = round(traffic / avg_s / 1e9, 1)
        if not args.no_ceiling:
            # the dominant kernel reads about twice what it writes: compare
            # with a plain stream of the same mix on this GPU
            ceil = stream_ceiling(p.shape[0] * p.shape[1])
            roofline["stream_ceiling"] = ceil
            roofline["frac_of_2r1w_ceiling"] = round(achieved / ceil["add_2r1w"], 4)
            if traffic:
                roofline["measured_frac_of_2r1w_ceiling"] = round(traffic / avg_s / 1e9 / ceil["add_2r1w"], 4)
        kernels = sorted(([e[0], e[1], e[2], round(e[3], 4), e[4]] for e in kt), key=lambda e: -e[3])[:12]
    # whole solve: the HBM bytes one solve of each precision moves (rocprofv3
    # FETCH_SIZE + WRITE_SIZE over all its dispatches, committed summary of a
    # run of the same command) over this run's solve time
    solve_traffic = None
    if args.seg == 8192:
        cyc = f"V({nu[0]},{nu[1]})"
        solve_traffic = {"source": os.path.relpath(SOLVE_PMC_SUMMARY, ROOT), "cycle": cyc}
        for key, ms in (("mixed", mixed_ms), ("fp64", fp64_ms)):
            b = solve_traffic_bytes(f"{key}{cyc}")
            solve_traffic[key] = None if b is None else {
                "bytes": b, "GBps": round(b / (ms * 1e-3) / 1e9, 1), "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    # config 3's literal cycle, V(2,2), for both precisions
    v22 = None
    if not args.no_v22 and nu != (2, 2):
        r22 = {m: timed(m, (2, 2), keep_u=False) for m in (True, False)}
        m22, f22 = (1e3 * float(np.mean(r22[m]["wall"])) for m in (True, False))
        v22 = {"mixed": {"value": round(p.unknowns / (m22 * 1e-3) / 1e6, 2), "ms_per_step": round(m22, 4),
                         "steps": r22[True]["cycles"][0], "vcycles": r22[True]["vcycles"],
                         "rel_residual": max(r22[True]["rel"])},
               "fp64": {"value": round(p.unknowns / (f22 * 1e-3) / 1e6, 2), "ms_per_step": round(f22, 4),
                        "cycles": r22[False]["cycles"][0], "rel_residual": max(r22[False]["rel"])},
               "speedup_mixed_vs_fp64": round(f22 / m22, 3)}

    accuracy = {"mixed_vs_fp64_gpu_rel": float(np.abs(res[True]["u"] - res[False]["u"]).max()
                                               / np.abs(res[False]["u"]).max())}
    cpu = None
    if not args.no_cpu:
        cpu, u_oracle = cpu_baseline(args.cpu_seg, args.rtol, nu)
        if args.cpu_seg == args.seg:
            scale = np.abs(u_oracle).max()
            accuracy["mixed_vs_fp64_oracle_rel"] = float(np.abs(res[True]["u"] - u_oracle).max() / scale)
            accuracy["fp64_gpu_vs_fp64_oracle_rel"] = float(np.abs(res[False]["u"] - u_oracle).max() / scale)
        em, e64 = dd128_check(nu, inner=args.inner)
        accuracy["mixed_513_vs_dd128_oracle_rel"] = em
        accuracy["fp64_513_vs_dd128_oracle_rel"] = e64

    line = {
        "metric": "Munknowns/s to rel. residual 1e-10: FP32+FP64-refine vs FP64 MG; % HBM roofline",
        "value": round(value, 2),
        "unit": "Munknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded Gaussian bumps, SplitMix64)",
        "config": {"workload": f"poisson_unit_square_{args.seg + 1}x{args.seg + 1}_gaussian_bumps64_"
                               f"mixed_fp32_vcycles_fp64_refinement_to_rel_residual_{args.rtol:g}",
                   "nodes": [p.seg_x + 1, p.seg_y + 1], "unknowns": p.unknowns,
                   "cycle": f"V({nu[0]},{nu[1]}) RB-GS, half-weighting restriction", "parallelism": "single GPU"},
        "fp64": {"value": round(p.unknowns / (fp64_ms * 1e-3) / 1e6, 2), "ms_per_step": round(fp64_ms, 4),
                 "cycles": res[False]["cycles"][0]},
        "mixed_cycles": res[True]["cycles"][0],
        "mixed_vcycles": res[True]["vcycles"],
        "mixed_max_inner": args.inner,
        "speedup_mixed_vs_fp64": round(fp64_ms / mixed_ms, 3),
        "roofline": roofline,
        "solve_traffic": solve_traffic,
        "v22": v22,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
        "kernels_top": kernels,
    }
    print(json.dumps(line), flush=True)


def measured_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (lines 'TRAFFIC <class>@level0/<f32|f64> <bytes>' written by
    vlp2d_amd.rocprof_summary), or None when no measurement of that kernel
    is recorded."""
    path = PMC_SUMMARY
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel_key:
            return float(parts[2])
    return None


def solve_traffic_bytes(label):
    """HBM bytes of one whole solve from the committed PMC summary (lines
    'SOLVE_TRAFFIC <label> <bytes> <dispatches>' written by
    vlp2d_amd.rocprof_summary solve), or None."""
    if not os.path.exists(SOLVE_PMC_SUMMARY):
        return None
    for line in open(SOLVE_PMC_SUMMARY):
        parts = line.split()
        if len(parts) == 4 and parts[0] == "SOLVE_TRAFFIC" and parts[1] == label:
            return float(parts[2])
    return None


if __name__ == "__main__":
    main()
