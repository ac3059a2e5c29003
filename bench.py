#!/usr/bin/env python
"""CCSD iterations/s and TFLOP/s on one MI355X: mixed precision vs plain FP64 at <= 1 uEh.

Workload (default): seeded synthetic closed-shell Hamiltonian, o = 48, v = 384 (nmo = 432,
naux = 864, seed 3), integrals generated on the GPU from density-fitting factors. One "step" is
one iteration of the CCSD solve loop (residual, Jacobi update, energy, DIIS), exactly as
``solve_cc`` runs it.

Legs (each in its own context, run one after the other):
  mixed  -- precision ``--precision`` (default MIXED_F16X3): W warmup + K timed iterations,
            then iterated (untimed) to convergence;
  fp64   -- precision DP, the same; its converged energy is the FP64 reference at this size. The
            iteration and time at which it first meets the mixed leg's thresholds give the time to
            solution at equal thresholds;
  others -- ``--secondary`` (default MIXED, MIXED_BF16X3: the FP32 form and the BF16x3-ladder form),
            the same as the mixed leg, reported under ``other_mixed_forms``.
``value`` is the mixed leg's iterations/s; the FP64 leg, the speed-up and the energy / amplitude
errors of the mixed solutions against the FP64 one are reported beside it. Each leg reports its
reference-equivalent rate (pycc's contraction count F_ref per second), its executed rate (the flops
its kernels run, each product counted once) and the fraction of the peak-bound time (each kernel
class's flops at the MFMA peak of the precision it runs in).

The roofline object describes the dominant kernel family of the mixed leg (the ring GEMMs), timed
with HIP events on the residual's stream by the library over the timed iterations; ``traffic``
comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed under profiles/ (null if absent).
The CPU baseline times the numpy oracle (a restatement of pycc's numpy path) on a bounded sample:
iterations at o = 16, v = 128, scaled to this workload by the ratio of reference flop counts.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# ring GEMM precision peak (FP16x3: three dense FP16 MFMA products per algorithmic product)
PEAK = {"DP": 78.6e12, "MIXED": 157.3e12, "MIXED_BF16X3": 157.3e12, "MIXED_F16X3": 2.5e15 / 3}
PEAK_FP64, PEAK_FP32, PEAK_BF16 = 78.6e12, 157.3e12, 2.5e15
HBM_PEAK = 8.0e12
DTYPES = {"MIXED": "f32 (ladder+rings+hole-hole), f64 (rest)",
          "MIXED_BF16X3": "bf16x3 ladder, f32 rings and hole-hole, f64 rest",
          "MIXED_F16X3": "f16x3 (ladder, rings and the other large contractions; scaled hi+lo, f32 accumulation), "
                         "f64 rest",
          "DP": "f64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _ms(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0), out


def backward_pass(cc, store_vvvv=False):
    """One backward of the converged solve (FP64 in a DP context; in a MIXED_F16X3 context its large
    contractions as FP16x3 split products, HBAR held in FP64): HBAR at (F, t1, t2), then the gradient of the CCSD
    Lagrangian E + l1.r1 + 1/2 l2.r2 at the guess l1 = 2 t1, l2 = 2(2 t2 - t2^T) (the Lambda
    residual), and a second Lambda residual on the same HBAR (the cost of one Lambda iteration)."""
    import torch
    import pycc_amd
    l1 = (2.0 * cc.t1).contiguous()
    l2 = (2.0 * (2.0 * cc.t2 - cc.t2.swapaxes(2, 3))).contiguous()
    hbar_ms, hbar = _ms(lambda: pycc_amd.cchbar(cc, store_vvvv=store_vvvv))
    res_ms, (r1, r2) = _ms(lambda: cc._lambda_residual(l1, l2, with_source=True))
    res2_ms, _ = _ms(lambda: cc._lambda_residual(l1, l2, with_source=True))
    out = {"store_vvvv": bool(store_vvvv), "hbar_build_ms": hbar_ms, "lambda_residual_ms": res2_ms,
           "first_residual_ms": res_ms, "total_ms": hbar_ms + res_ms, "device_GB": cc.device_bytes() / 1e9,
           "gradient_norm": float(torch.sqrt((r1 * r1).sum() + (r2 * r2).sum()))}
    del hbar, r1, r2, l1, l2
    return out


def free_device():
    """Release every context no longer referenced (its device memory goes with it) before the next
    leg builds its own."""
    import gc
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def run_leg(sd, no, precision, warmup, steps, converge, maxiter, e_conv, r_conv, backward=False, mark=None,
            refine_from=None):
    """One precision leg: W warm-up and K timed iterations of the solve loop, then (untimed) on to
    convergence. ``mark = (e_conv2, r_conv2)`` records the iteration and elapsed solve time at which
    the loop first meets that second pair of thresholds (its convergence test, ccwfn.py:295).
    ``refine_from = (t1, t2)``: after the converged solve (and the backward), the same context solves
    again from those amplitudes to the same thresholds (FP64 refinement of a mixed solution)."""
    import torch
    import pycc_amd
    t0 = time.time()
    cc = pycc_amd.CCwfn(sd, no=no, precision=precision, verbose=False)
    torch.cuda.synchronize()
    setup = time.time() - t0
    log("[%s] setup %.1f s, device memory %.1f GB" % (precision, setup, cc.device_bytes() / 1e9))
    marked = {}

    def step():
        r = next(it, None)
        if r is not None and mark and not marked and abs(r["ediff"]) < mark[0] and r["rms"] < mark[1]:
            torch.cuda.synchronize()
            marked.update(iterations=r["niter"], time_s=time.perf_counter() - t_solve,
                          e_conv=mark[0], r_conv=mark[1])
        return r

    it = cc.iterate(e_conv, r_conv, maxiter, 8, 1)
    last = None
    done = False
    t_solve = time.perf_counter()
    for _ in range(warmup):
        last = step()
        if last is None or last["converged"]:
            done = True
            break
    torch.cuda.synchronize()
    cc.timing_start()
    t1 = time.perf_counter()
    n = 0
    while n < steps and not done:
        last = step()
        if last is None:
            done = True
            break
        n += 1
        if last["converged"]:
            done = True
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    timing = cc.timing_read()
    log("[%s] %d timed iterations: %.3f s/iter (E = %.12f, rms %.2e)" % (
        precision, n, wall / max(n, 1), last["ecc"] if last else float("nan"), last["rms"] if last else 0))
    if converge:
        while not done:
            last = step()
            if last is None:
                break
            if last["niter"] % 5 == 0:
                log("[%s] iter %d E = %.12f dE = %.2e rms = %.2e" % (
                    precision, last["niter"], last["ecc"], last["ediff"], last["rms"]))
            if last["converged"]:
                done = True
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t_solve
    out = {
        "precision": precision, "setup_s": setup, "timed_steps": n, "wall_s": wall,
        "ms_per_step": 1e3 * wall / max(n, 1), "iters_per_s": n / wall if wall > 0 else 0.0,
        "timing": timing, "converged": bool(last and last["converged"]),
        "iterations": cc.niter, "energy": cc.ecc if (last and last["converged"]) else (last["ecc"] if last else None),
        "device_GB": cc.device_bytes() / 1e9, "solve_s": solve_s if converge else None,
        "e_conv": e_conv, "r_conv": r_conv, "at_mark": marked or None,
    }
    if backward and out["converged"]:
        out["backward"] = backward_pass(cc)
        log("[%s] backward: HBAR %.0f ms, Lambda residual %.0f ms, %.1f GB" % (
            precision, out["backward"]["hbar_build_ms"], out["backward"]["lambda_residual_ms"],
            out["backward"]["device_GB"]))
    t1c, t2c = cc.t1.clone(), cc.t2.clone()
    if refine_from is not None and out["converged"]:
        cc.t1.copy_(refine_from[0].to(cc.t1.device))
        cc.t2.copy_(refine_from[1].to(cc.t2.device))
        torch.cuda.synchronize()
        t_ref = time.perf_counter()
        e_ref = cc.solve_cc(e_conv, r_conv, maxiter)
        torch.cuda.synchronize()
        ref_s = time.perf_counter() - t_ref
        out["refine"] = {"iterations": cc.niter, "time_s": ref_s, "converged": e_ref is not None,
                         "energy": e_ref, "e_conv": e_conv, "r_conv": r_conv,
                         "energy_minus_solve_uEh": 1e6 * (e_ref - out["energy"]) if e_ref is not None else None,
                         "max_abs_dT2_vs_solve": float((cc.t2 - t2c).abs().max())}
        log("[%s] refinement from the mixed amplitudes: %d iterations, %.1f s, E - E(solve) = %.2e uEh" % (
            precision, cc.niter, ref_s, out["refine"]["energy_minus_solve_uEh"] or float("nan")))
    del cc, it
    free_device()
    return out, t1c, t2c


def cpu_info():
    """CPU model, logical CPU counts and the BLAS numpy calls (threadpoolctl), for the CPU baseline."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = None
    try:
        from threadpoolctl import threadpool_info
        for lib in threadpool_info():
            if lib.get("user_api") == "blas":
                blas = "%s %s (%s, %s)" % (lib.get("internal_api"), lib.get("version"),
                                           os.path.basename(lib.get("filepath", "")), lib.get("threading_layer"))
                break
    except Exception:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count()
    return {"model": model, "os_cpu_count": os.cpu_count(), "affinity": affinity, "blas": blas}


def cpu_threads():
    """Threads for the CPU baseline: the CPU share this process is given. On a shared GPU host that
    share is OMP_NUM_THREADS (one GPU's cores), which can be far below os.cpu_count(); otherwise
    the CPUs this process may run on."""
    info = cpu_info()
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        return min(int(share), info["affinity"] or int(share))
    return info["affinity"] or 1


def cpu_baseline(threads):
    """numpy oracle iterations at o=16, v=128 (seed 0), scaled by F_ref(workload)/F_ref(sample)."""
    from threadpoolctl import threadpool_limits
    from oracle.ccsd_ref import OracleCCSD, OracleDIIS
    from pycc_amd.synth import synthetic_df, DenseHamiltonian, dirac_eri
    so, sv = 16, 128
    sd = synthetic_df(so, sv, seed=0)
    H = DenseHamiltonian(sd.fock, dirac_eri(sd.B))
    with threadpool_limits(threads):
        cc = OracleCCSD(H, so)
        diis = OracleDIIS(cc.t1, cc.t2, 8)
        times = []
        for k in range(4):
            t0 = time.perf_counter()
            r1, r2 = cc.residuals(H.F, cc.t1, cc.t2)
            cc.t1 = cc.t1 + r1 / cc.Dia
            cc.t2 = cc.t2 + r2 / cc.Dijab
            cc.energy()
            diis.add_error_vector(cc.t1, cc.t2)
            cc.t1, cc.t2 = diis.extrapolate(cc.t1, cc.t2)
            times.append(time.perf_counter() - t0)
    s_per_iter = sorted(times[1:])[len(times[1:]) // 2]      # median after one warm-up
    return so, sv, s_per_iter


# Peaks by kernel class and precision mode. FP16x3 / BF16x3 products count once against a third of
# the dense 16-bit MFMA rate (three products per algorithmic product).
P_X3 = PEAK_BF16 / 3
CLASS_PEAKS = {
    "DP": {"ladder": PEAK_FP64, "ring": PEAK_FP64, "hole_hole": PEAK_FP64, "rest": PEAK_FP64},
    "MIXED": {"ladder": PEAK_FP32, "ring": PEAK_FP32, "hole_hole": PEAK_FP32, "rest": PEAK_FP64},
    "MIXED_BF16X3": {"ladder": P_X3, "ring": PEAK_FP32, "hole_hole": PEAK_FP32, "rest": PEAK_FP64},
    "MIXED_F16X3": {"ladder": P_X3, "ring": P_X3, "hole_hole": P_X3, "rest": P_X3},
}


def leg_rates(precision, o, v, ms_per_step):
    """Reference-equivalent and executed TFLOP/s of one leg, and the fraction of the peak-bound
    time: sum over kernel classes of (class flops / class peak), divided by the measured time."""
    from pycc_amd.flops import f_ref, f_executed, f_ladder_kernel, f_ring_kernel, f_rest, f_hole_hole
    t = ms_per_step * 1e-3
    parts = {"ladder": f_ladder_kernel(o, v), "ring": f_ring_kernel(o, v), "hole_hole": f_hole_hole(o, v),
             "rest": f_rest(o, v) - f_hole_hole(o, v)}
    t_peak = sum(parts[k] / CLASS_PEAKS[precision][k] for k in parts)
    return {"tflops_reference_equivalent": f_ref(o, v) / t / 1e12,
            "tflops_executed": f_executed(o, v) / t / 1e12,
            "peak_bound_ms": 1e3 * t_peak, "frac_of_peak": t_peak / t}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no", type=int, default=48)
    ap.add_argument("--nv", type=int, default=384)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--precision", default="MIXED_F16X3", choices=["MIXED", "MIXED_BF16X3", "MIXED_F16X3", "DP"])
    ap.add_argument("--maxiter", type=int, default=60)
    ap.add_argument("--e-conv", type=float, default=1e-9, help="energy threshold of the FP64 leg")
    ap.add_argument("--e-conv-mixed", type=float, default=1e-7,
                    help="energy threshold of the mixed leg (its iteration-to-iteration energy noise is "
                         "~3e-8 Eh at o=48, v=384, so a tighter threshold is met only by chance)")
    ap.add_argument("--r-conv-dp", type=float, default=1e-8)
    ap.add_argument("--r-conv-mixed", type=float, default=1e-6)
    ap.add_argument("--no-converge", action="store_true", help="skip the untimed convergence legs")
    ap.add_argument("--no-dp", action="store_true", help="skip the FP64 leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-backward", action="store_true", help="skip the backward passes (FP64 and mixed)")
    ap.add_argument("--secondary", default="MIXED,MIXED_BF16X3",
                    help="comma-separated further mixed forms timed and converged after the FP64 leg "
                         "(empty: none)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only copy a GIB-sized FP64 device buffer 3 times (calibrates HBM byte counters)")
    ap.add_argument("--backward", action="store_true",
                    help="measure the backward instead: o=32 v=256 seed 2 FP64 (override with --no/--nv/"
                         "--seed), HBAR stored vs factored, finite-difference checks, Lambda solve")
    ap.add_argument("--triples", nargs="?", const="32,256", default=None, metavar="O,V",
                    help="measure the (T) correction instead: FP64 CCSD converged at o=32 v=256 (or O,V) "
                         "seed 3 (--seed), E(T) timed in FP64 and in MIXED_F16X3")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: the library runs on one GPU; --gpus %d is not supported (multi-GPU sharding "
              "is designed, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.backward:
        return backward_bench(args)
    if args.triples:
        return triples_bench(args)
    if args.copy_probe > 0:
        import torch
        n = int(args.copy_probe * (1 << 30)) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 3}))
        return

    import torch
    from pycc_amd.flops import f_ref, f_ladder_kernel, f_executed
    from pycc_amd.synth import synthetic_df

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    o, v = args.no, args.nv
    t0 = time.time()
    sd = synthetic_df(o, v, seed=args.seed)
    log("synthetic Hamiltonian o=%d v=%d naux=%d seed=%d generated in %.1f s" % (o, v, sd.naux, args.seed, time.time() - t0))

    converge = not args.no_converge
    mixed_thr = (args.e_conv_mixed, args.r_conv_mixed)
    mixed, mt1, mt2 = run_leg(sd, o, args.precision, args.warmup, args.steps, converge, args.maxiter,
                              *(mixed_thr if args.precision != "DP" else (args.e_conv, args.r_conv_dp)),
                              backward=not args.no_backward and args.precision == "MIXED_F16X3")
    mt1, mt2 = mt1.cpu(), mt2.cpu()
    dp = None
    if not args.no_dp and args.precision != "DP":
        free_device()
        dp, dt1, dt2 = run_leg(sd, o, "DP", args.warmup, args.steps, converge, args.maxiter,
                               args.e_conv, args.r_conv_dp, backward=not args.no_backward, mark=mixed_thr,
                               refine_from=(mt1, mt2) if converge and mixed["converged"] else None)
        dt1, dt2 = dt1.cpu(), dt2.cpu()
    # the other mixed forms, each in its own context after the ones above are freed
    others = {}
    for prec in [p for p in args.secondary.split(",") if p and p != args.precision and p != "DP"]:
        free_device()
        leg, t1x, t2x = run_leg(sd, o, prec, args.warmup, args.steps, converge, args.maxiter, *mixed_thr)
        others[prec] = (leg, t1x.cpu(), t2x.cpu())
        del t1x, t2x

    def versus_fp64(leg, t1x, t2x):
        out = {}
        if dp and dp["iters_per_s"]:
            out["speedup_vs_fp64"] = leg["iters_per_s"] / dp["iters_per_s"]
        if dp and leg["converged"] and dp["converged"]:
            out["energy_error_uEh"] = 1e6 * (leg["energy"] - dp["energy"])
            out["max_abs_dT1"] = float((t1x - dt1).abs().max())
            out["max_abs_dT2"] = float((t2x - dt2).abs().max())
            mk = dp.get("at_mark")
            if mk and leg["solve_s"]:
                out["time_to_solution_vs_fp64_same_thresholds"] = mk["time_s"] / leg["solve_s"]
        return out

    F = f_ref(o, v)
    res = {}
    res["metric"] = "CCSD iters/sec + TFLOP/s, mixed vs FP64 plain at <=1 uEh error (o=48,v=384)"
    res["value"] = mixed["iters_per_s"]
    res["unit"] = "iters/s"
    res["n_gpus"] = 1
    res["steps"] = mixed["timed_steps"]
    res["warmup"] = args.warmup
    res["ms_per_step"] = mixed["ms_per_step"]
    res["higher_is_better"] = True
    res["scaling"] = "weak"
    res["vs_baseline"] = None
    res["dtype"] = DTYPES[args.precision]
    res["data"] = "synthetic (seeded DF Hamiltonian, no external data)"
    res["config"] = {"workload": "closed-shell CCSD iteration, synthetic DF Hamiltonian",
                     "no": o, "nv": v, "nmo": o + v, "naux": sd.naux, "seed": args.seed,
                     "precision": args.precision, "diis": 8}
    res.update(leg_rates(args.precision, o, v, mixed["ms_per_step"]))
    res["f_ref_per_iter"] = F
    res["f_executed_per_iter"] = f_executed(o, v)
    if dp:
        res["fp64"] = {"iters_per_s": dp["iters_per_s"], "ms_per_step": dp["ms_per_step"],
                       "iterations": dp["iterations"], "energy": dp["energy"], "converged": dp["converged"],
                       "time_to_solution_s": dp["solve_s"], "e_conv": dp["e_conv"], "r_conv": dp["r_conv"],
                       "at_mixed_thresholds": dp["at_mark"]}
        res["fp64"].update(leg_rates("DP", o, v, dp["ms_per_step"]))
        if "refine" in dp:
            # the FP64 solution reached by a mixed solve then FP64 iterations from its amplitudes
            r = dict(dp["refine"])
            if mixed["solve_s"] and dp["solve_s"] and r["converged"]:
                r["mixed_plus_refine_s"] = mixed["solve_s"] + r["time_s"]
                r["speedup_vs_fp64_solve"] = dp["solve_s"] / r["mixed_plus_refine_s"]
            res["fp64"]["refined_from_mixed"] = r
        res.update(versus_fp64(mixed, mt1, mt2))
    if dp and "backward" in dp:
        b = dp["backward"]
        b["workload"] = ("FP64 Lagrangian backward (factored HBAR + Lambda residual = vJp with the energy "
                         "source) at the converged FP64 amplitudes, o=%d v=%d" % (o, v))
        res["backward"] = b
    if "backward" in mixed:
        b = mixed["backward"]
        b["workload"] = ("MIXED_F16X3 Lagrangian backward (factored HBAR + Lambda residual, the contractions "
                         "of >= 1e11 flops and the ladder as FP16x3 split products) at the converged mixed "
                         "amplitudes, o=%d v=%d" % (o, v))
        if "backward" in res:
            b["speedup_vs_fp64"] = {k: res["backward"][k] / b[k] for k in ("hbar_build_ms", "lambda_residual_ms")}
        res["backward_mixed"] = b
    res["mixed"] = {"iterations": mixed["iterations"], "energy": mixed["energy"], "converged": mixed["converged"],
                    "time_to_solution_s": mixed["solve_s"], "e_conv": mixed["e_conv"], "r_conv": mixed["r_conv"],
                    "setup_s": mixed["setup_s"], "device_GB": mixed["device_GB"]}
    for prec, (leg, t1x, t2x) in others.items():
        n_r, ms_r, fl_r = leg["timing"]["ring"]
        n_l, ms_l, fl_l = leg["timing"]["ladder"]
        entry = {"dtype": DTYPES[prec], "iters_per_s": leg["iters_per_s"], "ms_per_step": leg["ms_per_step"],
                 "iterations": leg["iterations"], "energy": leg["energy"], "converged": leg["converged"],
                 "time_to_solution_s": leg["solve_s"], "e_conv": leg["e_conv"], "r_conv": leg["r_conv"],
                 "ring_tflops": fl_r / (ms_r * 1e-3) / 1e12 if ms_r > 0 else None,
                 "ladder_tflops": fl_l / (ms_l * 1e-3) / 1e12 if ms_l > 0 else None,
                 "device_GB": leg["device_GB"]}
        entry.update(leg_rates(prec, o, v, leg["ms_per_step"]))
        entry.update(versus_fp64(leg, t1x, t2x))
        res.setdefault("other_mixed_forms", {})[prec] = entry

    # roofline of the dominant kernel family (ring GEMMs) over the timed iterations
    n_r, ms_r, fl_r = mixed["timing"]["ring"]
    n_l, ms_l, fl_l = mixed["timing"]["ladder"]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ring_gemm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(args.precision)
    ring_prec = args.precision if args.precision in ("DP", "MIXED_F16X3") else "MIXED"
    achieved = fl_r / (ms_r * 1e-3) if ms_r > 0 else 0.0
    ring_kernel = {"DP": "plain_gemm_kernel (ring GEMMs, f64)", "MIXED": "plain_gemm_kernel (ring GEMMs, f32)",
                   "MIXED_F16X3": "x3_gemm_dma_kernel<_Float16, 320> (ring GEMMs, f16x3)"}[ring_prec]
    res["roofline"] = {"kernel": ring_kernel,
                       "bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[ring_prec] / 1e12,
                       "unit": "TFLOP/s", "frac": achieved / PEAK[ring_prec], "traffic": traffic,
                       "launches": n_r, "avg_launch_ms": ms_r / max(n_r, 1),
                       "flops_per_launch": fl_r / max(n_r, 1)}
    lad_peak = PEAK_BF16 / 3 if args.precision in ("MIXED_BF16X3", "MIXED_F16X3") else PEAK[args.precision]
    res["ladder"] = {"achieved_tflops": (fl_l / (ms_l * 1e-3) / 1e12) if ms_l > 0 else 0.0,
                     "peak_tflops": lad_peak / 1e12, "ms_per_step": ms_l / max(mixed["timed_steps"], 1),
                     "flops_per_step": f_ladder_kernel(o, v)}
    res["ring_ms_per_step"] = ms_r / max(mixed["timed_steps"], 1)

    if not args.no_cpu:
        threads = cpu_threads()
        info = cpu_info()
        log("[cpu] numpy oracle on %d threads (%s; %s)" % (threads, info["model"], info["blas"]))
        so, sv, s_it = cpu_baseline(threads)
        scale = f_ref(o, v) / f_ref(so, sv)
        res["cpu_baseline"] = {"value": 1.0 / (s_it * scale), "unit": "iters/s", "cores": threads,
                               "kind": "port", "cpu": info,
                               "sample": "numpy oracle (restated pycc numpy path), median of 3 iterations at "
                                         "o=16,v=128 seed 0 = %.3f s/iter, scaled by F_ref ratio %.1f" % (s_it, scale)}
    print(json.dumps(res), flush=True)


def _gamma_element(g, no, idx):
    """Element of the four-fold symmetrised Dirac-order Gamma (ccdensity.py:315-334) from density
    blocks held on the device: dL/d<pq|rs> for one orbital index quadruple."""
    table = {"oooo": (0.5, "Doooo", False), "vvvv": (0.5, "Dvvvv", False), "ovov": (1.0, "Dovov", False),
             "oovv": (0.25, "Doovv", False), "vvoo": (0.25, "Doovv", True), "ooov": (0.5, "Dooov", False),
             "ovoo": (0.5, "Dooov", True), "vvvo": (0.5, "Dvvvo", False), "vovv": (0.5, "Dvvvo", True)}

    def g0(a, b, c, d):
        pat = "".join("o" if x < no else "v" for x in (a, b, c, d))
        if pat not in table:
            return 0.0
        coef, name, swap = table[pat]
        k = tuple(x if x < no else x - no for x in (a, b, c, d))
        if swap:
            k = (k[2], k[3], k[0], k[1])
        return coef * float(g[name][k])

    a, b, c, d = idx
    return 0.25 * (g0(a, b, c, d) + g0(b, a, d, c) + g0(c, d, a, b) + g0(d, c, b, a))


def integral_element_checks(cc, sd, t1, t2, c1, c2, rng, n_fock=16, n_eri=32, h=1e-5):
    """The integral half of the vJp of c.r(t) at fixed amplitudes against difference quotients.
    Fock elements: autograd's dF vs central differences through the residual's F argument. ERI
    elements <pq|rs> (8-fold orbits, seeded): the two-particle density vs one-sided difference
    quotients over DF contexts with one extra auxiliary vector sqrt(h) u (c.r is linear in the
    integrals at fixed amplitudes, and a DF context can only add positive rank-one terms), using
    d(orbit) = q(u_P + u_Q) - q(u_P) - q(u_Q) for P = (p, r), Q = (q, s)."""
    import types
    import torch
    import pycc_amd
    from pycc_amd import autograd as ccad
    o, v = cc.no, cc.nv
    nmo = o + v
    F = cc.H.F

    def cr(ctx, Fx):
        r1, r2 = ctx.residuals(Fx, t1, t2)
        return float((c1 * r1).sum() + (c2 * r2).sum())

    t0 = time.time()
    Fg = F.clone().requires_grad_(True)
    gF, = torch.autograd.grad(ccad.residual(cc, t1, t2, F=Fg), (Fg,), (c1, c2))
    worst_f = 0.0
    for _ in range(n_fock):
        p, q = (int(x) for x in rng.integers(0, nmo, 2))
        dF = torch.zeros_like(F)
        dF[p, q] = 1.0
        fd = (cr(cc, F + h * dF) - cr(cc, F - h * dF)) / (2 * h)
        worst_f = max(worst_f, abs(fd - float(gF[p, q])) / max(1.0, abs(fd)))
    log("[backward] %d Fock elements: max rel error %.2e" % (n_fock, worst_f))
    g = cc.integral_vjp(t1, t2, c1, c2, vvvv=True)
    B = torch.as_tensor(sd.B, device="cuda")
    base = cr(cc, F)
    cache = {}

    def quotient(pairs):
        key = tuple(sorted(set(pairs)))
        if key not in cache:
            u = torch.zeros((1, nmo, nmo), dtype=torch.float64, device="cuda")
            for (a, b) in key:
                u[0, a, b] = 1.0
                u[0, b, a] = 1.0
            Hx = types.SimpleNamespace(B=torch.cat([B, (h ** 0.5) * u]), fock=sd.fock, no=o)
            ctx = pycc_amd.CCwfn(Hx, no=o, precision="DP", verbose=False)
            cache[key] = (cr(ctx, F) - base) / h
            del ctx, Hx
        return cache[key]

    worst_e, largest = 0.0, 0.0
    for k in range(n_eri):
        if k % 8 == 0:
            log("[backward] ERI elements %d-%d of %d (%.0f s)" % (k, min(n_eri, k + 8) - 1, n_eri, time.time() - t0))
        p, q, r, s = (int(x) for x in rng.integers(0, nmo, 4))
        orbit = {(p, q, r, s), (r, q, p, s), (p, s, r, q), (r, s, p, q),
                 (q, p, s, r), (s, p, q, r), (q, r, s, p), (s, r, q, p)}
        an = sum(_gamma_element(g, o, idx) for idx in orbit)
        P, Q = (min(p, r), max(p, r)), (min(q, s), max(q, s))
        fd = quotient([P]) if P == Q else quotient([P, Q]) - quotient([P]) - quotient([Q])
        worst_e = max(worst_e, abs(fd - an) / max(1.0, abs(an)))
        largest = max(largest, abs(an))
    del g, B
    log("[backward] %d ERI elements: max rel error %.2e (largest |dL/dERI| %.3e), %d DF contexts, %.0f s" % (
        n_eri, worst_e, largest, len(cache), time.time() - t0))
    return {"fock_elements": n_fock, "fock_fd_step": h, "fock_fd_max_rel_error": worst_f,
            "eri_elements": n_eri, "eri_fd_step": h, "eri_fd": "one-sided, DF rank-one perturbations",
            "eri_fd_max_rel_error": worst_e, "eri_max_abs_gradient": largest,
            "seconds": time.time() - t0}


def triples_bench(args):
    """The (T) correction (pycc/cctriples.py:177-240) at o=32, v=256 by default (seed 3): converge CCSD in FP64,
    time E(T) once in FP64, then twice in MIXED_F16X3 (FP16x3 slot GEMMs) on the same amplitudes,
    each in ranges of 8 values of i (ccx_triples_range) with a progress line per range;
    `value` is the MIXED_F16X3 time and E_T_error its difference from the FP64 E(T). Algorithmic
    work: the reference's twelve t3c_ijk contractions per triple i >= j >= k, 12 v^3 (v + o) flops,
    priced at the FP16x3 peak (dense FP16 MFMA / 3).
    The CPU baseline runs the oracle's t3c_ijk and t3d_ijk and one v^3 reduction on the first 4
    triples and scales by the triple count."""
    import numpy as np
    import torch
    import pycc_amd
    from pycc_amd.synth import synthetic_df
    o, v = (int(x) for x in args.triples.split(","))
    sd = synthetic_df(o, v, seed=args.seed)
    prec = "MIXED_F16X3" if args.no_dp else "DP"     # --no-dp: no FP64 leg (amplitudes from MIXED_F16X3)
    cc = pycc_amd.CCwfn(sd, no=o, precision=prec, verbose=False)
    solve_ms, ecc = _ms(lambda: cc.solve_cc(1e-10 if prec == "DP" else 1e-7, 1e-10 if prec == "DP" else 1e-6, 100))
    log("[triples] %s CCSD o=%d v=%d: E = %.12f in %d iterations (%.1f s)" % (prec, o, v, ecc, cc.niter, solve_ms / 1e3))
    dp_ms, et_dp = None, None
    if prec == "DP":
        dp_ms, et_dp = _ms(lambda: cc.triples_energy(
            i_chunk=8, progress=lambda d, n: log("[triples] FP64 E(T): i < %d of %d done" % (d, n))))
        log("[triples] FP64 E(T) = %.15f in %.1f ms" % (et_dp, dp_ms))
    t1, t2 = cc.t1.clone(), cc.t2.clone()
    del cc
    torch.cuda.empty_cache()
    mx = pycc_amd.CCwfn(sd, no=o, precision="MIXED_F16X3", verbose=False)
    runs = []
    for rep in range(2):
        ms, et = _ms(lambda: mx.triples_energy(
            t1, t2, i_chunk=8, progress=lambda d, n: log("[triples] MIXED_F16X3 E(T): i < %d of %d done" % (d, n))))
        runs.append(ms)
        log("[triples] MIXED_F16X3 E(T) = %.15f in %.1f ms (error %s Eh)" % (
            et, ms, "%.2e" % (et - et_dp) if et_dp is not None else "n/a"))
    ntri = o * (o + 1) * (o + 2) // 6
    flops = ntri * 12.0 * v ** 3 * (v + o)
    best = min(runs)
    # CPU sample: the oracle on the first 4 triples (i, j, k) = (0,0,0), (1,0,0), (1,1,0), (1,1,1)
    log("[triples] CPU sample: oracle on 4 triples")
    from oracle.triples_ref import t3c_ijk, t3d_ijk
    Fh = mx.H.F.cpu().numpy()
    t1 = t1.cpu().numpy()
    t2 = t2.cpu().numpy()
    B = np.asarray(sd.B)                      # (pq|rs) = sum_Q B[Q,p,q] B[Q,r,s]; <pq|rs> = (pr|qs)
    Boo, Bvv, Bov, Bvo = B[:, :o, :o], B[:, o:, o:], B[:, :o, o:], B[:, o:, :o]
    Wvvvo = np.einsum('Qac,Qbi->abci', Bvv, Bvo, optimize=True)     # <ab|ci> = (ac|bi)
    Wovoo = np.einsum('Qmj,Qck->mcjk', Boo, Bvo, optimize=True)     # <mc|jk> = (mj|ck)
    Woovv = np.einsum('Qia,Qjb->ijab', Bov, Bov, optimize=True)     # <ij|ab> = (ia|jb)
    t0 = time.perf_counter()
    sample = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)]
    for (i, j, k) in sample:
        W3 = t3c_ijk(i, j, k, t2, Wvvvo, Wovoo)
        V3 = W3 + t3d_ijk(i, j, k, t1, t2, Woovv, Fh[:o, o:])
        float((W3 * V3).sum())
    cpu_s = (time.perf_counter() - t0) / len(sample) * ntri
    out = {"metric": "(T) correction wall time", "value": best / 1e3, "unit": "s", "higher_is_better": False,
           "dtype": "f16x3", "data": "synthetic",
           "config": {"workload": "CCSD(T) (T) correction, synthetic DF Hamiltonian", "no": o, "nv": v,
                      "seed": args.seed, "triples": ntri},
           "E_T": et, "E_T_fp64": et_dp, "E_T_error": (et - et_dp) if et_dp is not None else None,
           "E_CCSD": ecc, "ccsd_precision": prec, "runs_ms": runs,
           "fp64_s": dp_ms / 1e3 if dp_ms else None,
           "fp64_tflops": flops / (dp_ms / 1e3) / 1e12 if dp_ms else None,
           "speedup_vs_fp64": dp_ms / best if dp_ms else None,
           "roofline": {"bound": "mfma", "achieved": flops / (best / 1e3) / 1e12, "peak": PEAK["MIXED_F16X3"] / 1e12,
                        "unit": "TFLOP/s", "frac": flops / (best / 1e3) / PEAK["MIXED_F16X3"], "traffic": None},
           "cpu_baseline": {"value": cpu_s, "unit": "s", "cores": int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or None,
                            "kind": "port", "sample": "oracle t3c_ijk + t3d_ijk on 4 triples, scaled to %d" % ntri}}
    print(json.dumps(out))


def backward_bench(args):
    """Differentiable CCSD at o=32, v=256 (seed 2, FP64): converge, then the backward through the
    torch.autograd.Function with two cotangents (the energy Lagrangian and a seeded N(0,1) one),
    32 seeded t2 elements against central differences (step 1e-5), 16 Fock and 32 ERI elements of
    the integral half against difference quotients, timings and device memory with the v^4 HBAR
    block stored and factored, and the Lambda solve."""
    import numpy as np
    import torch
    import pycc_amd
    from pycc_amd import autograd as ccad
    from pycc_amd.synth import synthetic_df
    o = 32 if args.no == 48 else args.no
    v = 256 if args.nv == 384 else args.nv
    seed = 2 if args.seed == 3 else args.seed
    sd = synthetic_df(o, v, seed=seed)
    cc = pycc_amd.CCwfn(sd, no=o, precision="DP", verbose=False)
    solve_ms, ecc = _ms(lambda: cc.solve_cc(1e-10, 1e-10, 100))
    log("[backward] FP64 CCSD o=%d v=%d: E = %.12f in %d iterations (%.1f s)" % (o, v, ecc, cc.niter, solve_ms / 1e3))
    res = {"metric": "CCSD backward (vJp) time and HBM, FP64, v^4 stored vs factored", "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic (seeded DF Hamiltonian)",
           "config": {"workload": "differentiable CCSD backward", "no": o, "nv": v, "seed": seed},
           "forward": {"energy": ecc, "iterations": cc.niter, "solve_ms": solve_ms}}
    base_GB = cc.device_bytes() / 1e9

    def one_pass(stored):
        b = backward_pass(cc, store_vvvv=stored)
        b["device_GB_before"] = base_GB
        res["stored" if stored else "factored"] = b
        log("[backward] %s: HBAR %.0f ms, Lambda residual %.0f ms, device %.1f GB" % (
            "stored" if stored else "factored", b["hbar_build_ms"], b["lambda_residual_ms"], b["device_GB"]))

    one_pass(False)
    # autograd with a seeded N(0,1) cotangent vs central differences on 32 seeded t2 elements
    rng = np.random.default_rng(seed)
    c1 = torch.tensor(rng.standard_normal((o, v)), device="cuda")
    c2 = torch.tensor(rng.standard_normal((o, o, v, v)), device="cuda")
    T1 = cc.t1.clone().requires_grad_(True)
    T2 = cc.t2.clone().requires_grad_(True)
    vjp_ms, grads = _ms(lambda: torch.autograd.grad(ccad.residual(cc, T1, T2), (T1, T2), (c1, c2)))
    g1, g2 = grads
    F = cc.H.F

    def f(a1, a2):
        r1, r2 = cc.residuals(F, a1, a2)
        return float((c1 * r1).sum() + (c2 * r2).sum())

    h, worst = 1e-5, 0.0
    t1d, t2d = T1.detach(), T2.detach()
    for _ in range(32):
        i, j, a, b = (int(x) for x in rng.integers(0, [o, o, v, v]))
        d2 = torch.zeros_like(t2d)
        d2[i, j, a, b] = 1.0
        d2[j, i, b, a] = 1.0
        mult = 1.0 if (i, a) == (j, b) else 2.0
        fd = (f(t1d, t2d + h * d2) - f(t1d, t2d - h * d2)) / (2 * h)
        an = mult * float(g2[i, j, a, b])
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    res["random_cotangent"] = {"vjp_ms_incl_forward_and_hbar": vjp_ms, "fd_t2_elements": 32,
                               "fd_step": h, "fd_max_rel_error": worst}
    log("[backward] random cotangent: %.0f ms, FD max rel error %.2e over 32 t2 elements" % (vjp_ms, worst))
    res["integral_elements"] = integral_element_checks(cc, sd, t1d, t2d, c1, c2, rng)
    one_pass(True)            # after the checks above: the stored v^4 HBAR stays in the context
    # Lambda solve (factored HBAR)
    lam = pycc_amd.cclambda(cc, pycc_amd.cchbar(cc))
    lsolve_ms, lecc = _ms(lambda: lam.solve_lambda(1e-10, 1e-10, 100))
    res["lambda_solve"] = {"pseudoenergy": lecc, "iterations": len(lam.history) - 1, "ms": lsolve_ms,
                           "ms_per_iter": lsolve_ms / max(1, len(lam.history) - 1)}
    log("[backward] Lambda solve: %s in %d iterations (%.1f s)" % (lecc, len(lam.history) - 1, lsolve_ms / 1e3))
    res["value"] = res["factored"]["total_ms"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
