"""H.psi throughput and sweep-step times on one MI355X, on real superblocks.

Default workload: the YC6 6x24 t1-t2-J1-J2 cylinder (144 sites, 132
electrons, t2 = 0.3, J2 = 0.036) at M = 6000 multiplets. A DMRG warm-up
grows the blocks to the centre (kept_max 1000, its last seven steps 6000),
then the first half-sweep's first three centre steps run twice from the
same blocks:
mixed precision (BF16x3 H.psi, FP64 Lanczos and refinement) and FP64. At
the first centre step, before its solve, the H.psi of that real superblock
is timed: a step is one application of the grouped two-stage contraction
with psi and every operator resident in HBM, replayed from a hipGraph, in
FP64 and in the mixed path; its relative error against FP64 is reported.
The same-path energy error of the mixed solves is measured against FP64
solves of the same superblocks from the same start vectors. The isolated
H.psi superblock (400 sectors x 24 terms, d in [8, 1536]) runs first: the
grouped kernels in FP64 and BF16x3 against the per-product ZGEMM loop.
The CPU baseline is the numpy (OpenBLAS) execution of a bounded sample of
the real superblock's plan on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0}
HBM_GBS = 8000.0

# HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of `bench.py --workload yc6-m6000-hist
# --steps 2 --warmup 1 --no-cpu --no-tune --calibrate-copy-mib 1024`, each
# counter scaled by its reading for the 1 GiB copy of the same run (x2.0 and
# x1.0); keyed by (workload, precision, stage). The work-unit rules give the
# same units for every autotuner candidate on this superblock. Not measured
# on the default workload: the counter runs need its warm-up, whose
# concurrent eigensolves the profiler did not survive.
MEASURED_TRAFFIC = {
    ("yc6-m6000-hist", "bf16x3", 2): 344.68e9 + 1.63e9,
    ("yc6-m6000-hist", "f64", 2): 1076.5e9 + 1.79e9,
}
TRAFFIC_NOTE = ("not measured on this workload; the same kernel on the yc6-m6000-hist superblock (sector "
                "lists of real M=6000 centre blocks, seeded values, 30.0M states) moves {:.1f} GB per launch")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="yc6-m6000",
                    choices=["yc6-m6000", "yc6-m6000-hist", "yc6-m12000-hist", "synthetic400", "synthetic-small"],
                    help="yc6-m6000: real DMRG superblocks (warm-up, centre window); *-hist: the real centre "
                         "sector histogram (x2 per sector for M=12000) with seeded random block values, H.psi only")
    ap.add_argument("--prec", default="bf16x3", choices=["bf16x3", "f32", "f64"])
    ap.add_argument("--window-steps", type=int, default=3, help="centre steps timed per precision")
    ap.add_argument("--ramp", default="1000,7", help="warm-up kept_max and the number of final steps at M "
                                                      "(at least the window steps + 4, so that every block the "
                                                      "window reads was grown from an M-block)")
    ap.add_argument("--warmup-lanzero", type=float, default=1e-6)
    ap.add_argument("--window-lanzero", type=float, default=1e-8)
    ap.add_argument("--refine", type=int, default=2, help="steps of the refinement pass after each mixed solve "
                                                           "(its first H.psi FP64)")
    ap.add_argument("--no-synthetic", action="store_true", help="skip the isolated 400x24 H.psi superblock")
    ap.add_argument("--no-window", action="store_true", help="real run: time H.psi only, no centre window")
    ap.add_argument("--no-tune", action="store_true", help="time H.psi with the default work units (no autotuning)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the measurement, copy a buffer of this many MiB three times (a kernel of known "
                         "HBM traffic, to scale hardware byte counters)")
    return ap.parse_args()


def stage_flops(c, o):
    """8 * m * n * k summed over the GEMM contributions of one stage (8 * m * n
    for elementwise ones)."""
    if len(o) == 0:
        return 0.0
    cnt = (o["cend"] - o["cbeg"]).astype(np.int64)
    idx = np.concatenate([np.arange(b, e) for b, e in zip(o["cbeg"], o["cend"])]) if cnt.sum() else np.zeros(0, int)
    mn = np.repeat(o["m"].astype(np.float64) * o["n"].astype(np.float64), cnt)
    ident = (c["a_op"][idx] == 4) | (c["b_op"][idx] == 4)
    k = np.where(ident, 1.0, c["k"][idx].astype(np.float64))
    return float(8.0 * np.sum(mn * k))


def time_gpu(ctx, plan, vin, vout, prec, steps, warmup):
    """Per-stage kernel times (HIP events on the plan's stream) and the wall
    time per application of `steps` graph replays."""
    for _ in range(max(1, warmup)):
        plan.apply(vin, vout, prec)
    ctx.sync()
    ev1, ev2 = [], []
    for _ in range(steps):
        plan.apply(vin, vout, prec)
        a, b = plan.last_times()
        ev1.append(a)
        ev2.append(b)
    plan.apply_graph(vin, vout, prec)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.apply_graph(vin, vout, prec)
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(ev1)), float(np.mean(ev2))


def blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"] or [1])
    except Exception:
        return os.cpu_count() or 1


def hpsi_measure(ctx, plan, vin, args, prec, arrays, flops_pairmin, workload):
    """Tuned work units, FP64 and mixed H.psi times, relative error, and the
    dominant kernel's roofline, on one plan."""
    from tsc_tri_t1t2j1j2_amd import capi
    c1, o1, c2, o2 = arrays
    f1, f2 = stage_flops(c1, o1), stage_flops(c2, o2)
    F = f1 + f2
    vout = ctx.vec(vin.n)
    vref = ctx.vec(vin.n)
    try:
        t_tune = time.perf_counter()
        if args.no_tune:
            tuned = {"f64": 2048.0, args.prec: 2048.0}
        else:
            tuned = {"f64": plan.autotune(vin, vout, capi.PREC_F64)}
            if prec != capi.PREC_F64:
                tuned[args.prec] = plan.autotune(vin, vout, prec)
        tuned["seconds"] = round(time.perf_counter() - t_tune, 2)
        w64, a64, b64 = time_gpu(ctx, plan, vin, vref, capi.PREC_F64, args.steps, args.warmup)
        print(f"# f64: {w64 * 1e3:.2f} ms/H.psi, {F / w64 / 1e12:.2f} TFLOP/s (stage 1 {a64:.2f} ms, stage 2 "
              f"{b64:.2f} ms)", file=sys.stderr, flush=True)
        if prec != capi.PREC_F64:
            wall, k1, k2 = time_gpu(ctx, plan, vin, vout, prec, args.steps, args.warmup)
            ref, out = vref.download(), vout.download()
            rel = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
        else:
            wall, k1, k2, rel = w64, a64, b64, 0.0
        print(f"# {args.prec}: {wall * 1e3:.2f} ms/H.psi, {F / wall / 1e12:.2f} TFLOP/s, rel err {rel:.2e} "
              f"(stage 1 {k1:.2f} ms, stage 2 {k2:.2f} ms)", file=sys.stderr, flush=True)
    finally:
        vout.close()
        vref.close()
    if k1 >= k2:
        stage, kms, kf = 1, k1, f1
    else:
        stage, kms, kf = 2, k2, f2
    achieved = kf / (kms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.prec]
    traffic = MEASURED_TRAFFIC.get((workload, args.prec, stage))
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 5), "traffic": traffic, "kernel": f"grouped stage {stage}",
            "kernel_ms": round(kms, 3), "flops_per_launch": kf}
    if args.prec == "bf16x3":
        # three bf16 MFMA products per real product (hi*hi + hi*lo + lo*hi):
        # the path's own ceiling is a third of the bf16 peak
        roof["mfma_products_per_product"] = 3
        roof["frac_of_bf16x3_ceiling"] = round(3.0 * achieved / peak, 5)
    if traffic:
        roof["traffic_tb_per_s"] = round(traffic / (kms * 1e-3) / 1e12, 3)
    elif workload == "yc6-m6000" and ("yc6-m6000-hist", args.prec, stage) in MEASURED_TRAFFIC:
        roof["traffic_note"] = TRAFFIC_NOTE.format(MEASURED_TRAFFIC[("yc6-m6000-hist", args.prec, stage)] / 1e9)
    res = {"algorithmic_flops_per_step": F, "flops_pairmin_per_step": flops_pairmin,
           "flops_note": "algorithmic_flops_per_step: the products as planned (O_L psi or psi O_R^T first per "
                         "source window, whichever is cheaper; each intermediate formed once); "
                         "flops_pairmin_per_step: every product at its own cheaper order, nothing shared. The "
                         "rates use the first (the smaller of the two here)",
           "stage_flops": [f1, f2], "precision": args.prec, "ms_per_step": round(wall * 1e3, 4),
           "tflops": round(F / wall / 1e12, 4), "f64_ms_per_step": round(w64 * 1e3, 4),
           "f64_tflops": round(F / w64 / 1e12, 4), "speedup_vs_f64": round(w64 / wall, 3), "rel_err_vs_f64": rel,
           "f64_stage_ms": [round(a64, 3), round(b64, 3)], "stage_ms": [round(k1, 3), round(k2, 3)],
           "autotune_units": tuned}
    return res, roof, wall


def cpu_sample(ops_host, arrays, mid_len, psi_host, seconds):
    from oracle import plan_exec
    c1, o1, c2, o2 = arrays
    fl, el, nw = plan_exec.timed_sample(ops_host, c1, o1, c2, o2, mid_len, psi_host, seconds)
    return {"value": fl / el / 1e12, "unit": "TFLOP/s", "cores": int(blas_threads()), "kind": "port",
            "sample": f"numpy/OpenBLAS execution of {nw} random output windows of the same H.psi plan (with the "
                      f"intermediates they read), {fl / 1e12:.3f} TFLOP in {el:.1f} s"}


def synthetic_leg(ctx, args, nsec=400, nterms=24, dmin=8, dmax=1536):
    """The isolated H.psi superblock: grouped FP64 / BF16x3 kernels against
    the per-product ZGEMM loop, all resident."""
    from tsc_tri_t1t2j1j2_amd import capi
    from tsc_tri_t1t2j1j2_amd.synthetic import SyntheticSuperblock
    t0 = time.perf_counter()
    sb = SyntheticSuperblock(nsec=nsec, nterms=nterms, dmin=dmin, dmax=dmax)
    plan = sb.make_plan(ctx, None)
    vin, vout, vref = ctx.vec(sb.psi_len), ctx.vec(sb.psi_len), ctx.vec(sb.psi_len)
    try:
        vin.fill_gaussian(sb.seed, 0xC0FFEE, 1.0)
        vin.axpby(0.0, vin, 1.0 / vin.norm())
        ctx.sync()
        setup = time.perf_counter() - t0
        F = sb.algorithmic_flops()
        out = {"synthetic_workload": f"{nsec} sectors x {nterms} terms, d_L, d_R log-uniform in [{dmin}, {dmax}], "
                                     f"seed {sb.seed}; psi {sb.psi_len} complex, operators {sb.ops_len} complex",
               "synthetic_flops_per_step": F, "synthetic_setup_s": round(setup, 2)}
        n = max(3, args.steps // 2)
        for name, p in (("f64", capi.PREC_F64), ("bf16x3", capi.PREC_BF16X3)):
            plan.autotune(vin, vout, p)
            w, _, _ = time_gpu(ctx, plan, vin, vout if p != capi.PREC_F64 else vref, p, n, 1)
            out[f"synthetic_{name}_ms"] = round(w * 1e3, 3)
            out[f"synthetic_{name}_tflops"] = round(F / w / 1e12, 3)
        ref, got = vref.download(), vout.download()
        out["synthetic_bf16x3_rel_err"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        plan.apply_loop(vin, vout)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(n):
            plan.apply_loop(vin, vout)
        ctx.sync()
        w = (time.perf_counter() - t1) / n
        got = vout.download()
        out["synthetic_loop_f64_ms"] = round(w * 1e3, 3)
        out["synthetic_loop_f64_tflops"] = round(F / w / 1e12, 3)
        out["synthetic_loop_rel_diff"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        out["synthetic_grouped_vs_loop_f64"] = round(w / (out["synthetic_f64_ms"] * 1e-3), 3)
        print(f"# synthetic 400x24: f64 {out['synthetic_f64_ms']} ms, bf16x3 {out['synthetic_bf16x3_ms']} ms, "
              f"per-product loop {out['synthetic_loop_f64_ms']} ms", file=sys.stderr, flush=True)
        return out
    finally:
        for v in (vin, vout, vref):
            v.close()
        plan.close()


class RealRun:
    """The real-superblock measurements: warm-up, H.psi timing on the first
    centre superblock (hook), and the centre window per precision."""

    def __init__(self, ctx, args, prec):
        self.ctx, self.args, self.prec = ctx, args, prec
        self.hpsi = None
        self.roof = None
        self.cpu = None
        self.wall = None
        self.hbm_peak = 0

    def hook(self, d, sb, plan, psi, tasks, step):
        from tsc_tri_t1t2j1j2_amd import capi
        """on_plan hook of the DMRG: time H.psi once, on the first centre
        superblock of the mixed window (psi is its start vector)."""
        if self.hpsi is not None or not getattr(d, "_measure_next", False):
            return
        d._measure_next = False
        arrays = tasks.arrays()
        _, pm = tasks.flops()
        self.hpsi, self.roof, self.wall = hpsi_measure(self.ctx, plan, psi, self.args, self.prec, arrays, pm,
                                                       "yc6-m6000")
        # the window's steps run with the default work units in both
        # precisions: put this plan's back
        scratch = self.ctx.vec(psi.n)
        for p in {self.prec, capi.PREC_F64}:
            if p != capi.PREC_F32:
                plan.autotune(psi, scratch, p, target=2048.0)
        scratch.close()
        self.hpsi.update({"psi_multiplet_dim": int(sb.super.dim), "sectors": int(sb.super.num),
                          "operator_arena_complex": int(step.len)})
        free, _ = self.ctx.mem_info()
        self.hbm_peak = max(self.hbm_peak, self.free0 - free)
        if not self.args.no_cpu:
            ops_host = step.vec.download()
            self.cpu = cpu_sample(ops_host, arrays, tasks.mid_len, psi.download(), self.args.cpu_seconds)
            del ops_host

    def run(self):
        from tsc_tri_t1t2j1j2_amd import capi
        from tsc_tri_t1t2j1j2_amd.dmrg import DMRG
        args, ctx = self.args, self.ctx
        self.free0, _ = ctx.mem_info()
        m0, nr = (int(x) for x in args.ramp.split(","))
        k = max(1, args.window_steps)
        t0 = time.perf_counter()
        d = DMRG(ctx, 24, 6, t2=0.3, J2=0.036, kept_min=110, kept_max=6000, sweeps=0, lanzero=args.warmup_lanzero,
                 prec=self.prec, refine=args.refine, seed=1234, retain=k + 1, warmup_ramp=(m0, nr),
                 on_plan=self.hook,
                 log=lambda m: print(f"# [{time.perf_counter() - t0:7.1f}s] {m}", file=sys.stderr, flush=True))
        d.warm_up()
        t_warm = time.perf_counter() - t0
        l0 = d.N // 2 - 1
        out = {"warmup_s": round(t_warm, 1),
               "warmup_schedule": f"kept_max {m0}, the last {nr} steps 6000; lanzero {args.warmup_lanzero}",
               "centre_sys_sectors": len(d.sys_blocks[l0].sectors),
               "centre_sys_multiplets": int(sum(x[2] for x in d.sys_blocks[l0].sectors))}
        if d.precision_record:
            out["warmup_mixed_precision"] = d.precision_by_phase().get("warmup")
        tol = args.window_lanzero * 15   # the first half-sweep's Lanczos tolerance (main.f90:124-131)
        terr = d.trun_err * 15
        windows = {}
        for name, p in (("mixed", self.prec), ("f64", capi.PREC_F64)):
            if args.no_window and name == "f64":
                break
            d.prec, d.refine = p, (args.refine if p != capi.PREC_F64 else 0)
            d._measure_next = name == "mixed"
            # the guessed steps of the mixed window are also solved in FP64
            # from the same start vectors (same-path error); the first step's
            # FP64 solve is the FP64 window's first step (same start vector)
            d.shadow, d.shadow_f64, d.precision_record = [], False, []
            recs = d.centre_window(1 if args.no_window else k, tol, 6000, terr, tag=f"window-{name}",
                                   shadow_from=1 if p != capi.PREC_F64 else None)
            windows[name] = {"steps": recs, "shadow": list(d.shadow), "refine": list(d.precision_record)}
            free, _ = ctx.mem_info()
            self.hbm_peak = max(self.hbm_peak, self.free0 - free)
        out["window_tol"] = tol
        mx = windows["mixed"]["steps"]
        if windows["mixed"]["refine"]:
            # from each mixed solve's refinement pass (its first product FP64): the mixed Ritz value
            # against the FP64 <psi|H psi> of its own Ritz vector, per site, and that vector's FP64 residual
            rr = windows["mixed"]["refine"]
            out["mixed_ritz_dE_per_site_vs_f64_rayleigh"] = [float(f"{abs(em - ea) / d.N:.3e}")
                                                             for (_, em, ea, _, _, _) in rr]
            out["refined_dE_est_per_site"] = [float(f"{wr * abs(em - ea) / d.N:.3e}")
                                              for (_, em, ea, _, _, wr) in rr]
            out["mixed_ritz_resid_f64"] = [float(f"{res:.3e}") for (_, _, _, res, _, _) in rr]
            out["refine_shift_per_site"] = [float(f"{(ea - er) / d.N:.3e}") for (_, _, ea, _, er, _) in rr]
        out["window_steps_mixed"] = [{kk: (round(v, 3) if isinstance(v, float) and kk == "seconds" else v)
                                      for kk, v in r.items()} for r in mx]
        if "f64" in windows:
            fx = windows["f64"]["steps"]
            out["window_steps_f64"] = [{kk: (round(v, 3) if isinstance(v, float) and kk == "seconds" else v)
                                        for kk, v in r.items()} for r in fx]
            sm, sf = sum(r["seconds"] for r in mx), sum(r["seconds"] for r in fx)
            out["window_s_mixed"], out["window_s_f64"] = round(sm, 2), round(sf, 2)
            out["window_speedup_mixed_vs_f64"] = round(sf / sm, 3)
            if len(mx) > 1:
                gm = sum(r["seconds"] for r in mx[1:])
                gf = sum(r["seconds"] for r in fx[1:])
                out["guessed_steps_speedup_mixed_vs_f64"] = round(gf / gm, 3)
                # a sweep is two random-start steps and 2(N-3)-2 guessed ones
                nsw = 2 * (d.N - 3)
                out["sweep_s_mixed_projected"] = round(2 * mx[0]["seconds"] + (nsw - 2) * gm / (len(mx) - 1), 1)
                out["sweep_s_f64_projected"] = round(2 * fx[0]["seconds"] + (nsw - 2) * gf / (len(fx) - 1), 1)
                out["sweep_projection_note"] = (f"{nsw} steps per sweep priced at the measured centre steps "
                                                "(random-start and guessed); the steps away from the centre are "
                                                "smaller, so this bounds the sweep from above")
            # the rho / eigensolve / truncation stage against HBM (its algorithmic bytes over its wall time;
            # the eigensolves are latency-bound, so this is far below the bandwidth)
            g = mx[-1]
            if g.get("truncate_gbs"):
                out["rho_truncation_roofline"] = {"bound": "hbm", "achieved": g["truncate_gbs"], "peak": HBM_GBS,
                                                  "unit": "GB/s", "frac": round(g["truncate_gbs"] / HBM_GBS, 6),
                                                  "bytes": g["truncate_bytes"], "step": g["l"]}
            N = d.N
            de = [abs(mx[0]["E"] - fx[0]["E"]) / N]
            de += [abs(em - e64) / N for (_, em, e64) in windows["mixed"]["shadow"]]
            out["same_path_dE_per_site"] = [float(f"{x:.3e}") for x in de]
            out["same_path_max_dE_per_site"] = max(de)
        out["hbm_peak_gib"] = round(self.hbm_peak / 2 ** 30, 2)
        return out


def hist_leg(ctx, args, prec, scale):
    """H.psi of a centre superblock with the real sector histogram (scaled)
    and seeded random block values."""
    from tsc_tri_t1t2j1j2_amd import hamiltonian as ham
    from tsc_tri_t1t2j1j2_amd import native
    from tsc_tri_t1t2j1j2_amd import resident
    from tsc_tri_t1t2j1j2_amd import workloads
    free0, _ = ctx.mem_info()
    t0 = time.perf_counter()
    ss, es, _ = workloads.real_centre_sectors(scale)
    sysb, envb, s1, s2, bonds, ntot = workloads.centre_superblock(24, 6, int(6000 * scale), t2=0.3, J2=0.036,
                                                                  sectors=(ss, es))
    sysd, envd = resident.model_to_device(ctx, sysb), resident.model_to_device(ctx, envb)
    del sysb, envb
    sb = ham.Superblock(sysd, envd, s1, s2, bonds, ntot, 0)
    terms = sb.terms()
    step = resident.StepArena(ctx, [sysd, envd])
    step.add_merges(terms)
    step.finalize()
    tasks = native.NativeTasks(sb, step=step, terms=terms)
    plan = tasks.plan(ctx)
    arrays = tasks.arrays()
    _, pm = tasks.flops()
    vin = ctx.vec(sb.super.dim)
    vin.fill_gaussian(7, 0xC0FFEE, 1.0)
    vin.axpby(0.0, vin, 1.0 / vin.norm())
    setup = time.perf_counter() - t0
    print(f"# setup {setup:.1f} s, psi {sb.super.dim} complex, step arena {step.len} complex", file=sys.stderr,
          flush=True)
    wl = "yc6-m6000-hist" if scale == 1.0 else "yc6-m12000-hist"
    res, roof, wall = hpsi_measure(ctx, plan, vin, args, prec, arrays, pm, wl)
    free, _ = ctx.mem_info()
    res.update({"psi_multiplet_dim": int(sb.super.dim), "sectors": int(sb.super.num),
                "operator_arena_complex": int(step.len), "setup_s": round(setup, 1),
                "hbm_in_use_gib": round((free0 - free) / 2 ** 30, 2)})
    cpu = None
    if not args.no_cpu:
        cpu = cpu_sample(step.vec.download(), arrays, tasks.mid_len, vin.download(), args.cpu_seconds)
    vin.close()
    plan.close()
    tasks.close()
    step.close()
    for m in (sysd, envd):
        m.arena.close()
    return res, roof, wall, cpu


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        sys.exit(2)
    from tsc_tri_t1t2j1j2_amd import capi
    prec = capi.PREC_NAMES[args.prec]
    ctx = capi.Context(0)
    _, total = ctx.mem_info()
    extra, cpu = {}, None
    if args.workload in ("synthetic400", "synthetic-small"):
        small = args.workload == "synthetic-small"
        res = synthetic_leg(ctx, args, *((60, 12, 8, 400) if small else (400, 24, 8, 1536)))
        line_cfg = dict(res, workload=res["synthetic_workload"])
        value, wall = res["synthetic_bf16x3_tflops"], res["synthetic_bf16x3_ms"] * 1e-3
        roof = {"bound": "mfma", "achieved": value, "peak": PEAK_TFLOPS["bf16x3"], "unit": "TFLOP/s",
                "frac": round(value / PEAK_TFLOPS["bf16x3"], 5), "traffic": None,
                "kernel": "grouped stages 1+2 (wall)"}
        data = "synthetic (seeded counter-based Gaussian operators and psi)"
    else:
        if not args.no_synthetic and args.workload == "yc6-m6000":
            extra.update(synthetic_leg(ctx, args))
        if args.workload == "yc6-m6000":
            rr = RealRun(ctx, args, prec)
            extra.update(rr.run())
            res, roof, wall, cpu = rr.hpsi, rr.roof, rr.wall, rr.cpu
            desc = ("YC6 6x24 (144 sites, 132 electrons, S=0) t1-t2-J1-J2 t-J cylinder, t1=1 t2=0.3 J1=0.4 "
                    "J2=0.036, M=6000 SU(2) multiplets: centre superblock [71 sites + 2 + 71 sites] of a DMRG "
                    "warm-up, its own blocks and ground-state start vector")
            data = "real DMRG superblocks (warm-up blocks of this run); seeded random start vectors"
        else:
            scale = 1.0 if args.workload == "yc6-m6000-hist" else 2.0
            res, roof, wall, cpu = hist_leg(ctx, args, prec, scale)
            desc = (f"YC6 6x24 centre superblock with the sector histogram of the real M=6000 centre blocks"
                    f"{' (every sector x2: M=12000)' if scale != 1.0 else ''}, seeded random block values")
            data = "synthetic values on a real sector histogram"
        value = res["tflops"]
        line_cfg = dict(res, workload=desc, parallelism="1 GPU")
    line = {
        "metric": "sweep time + H·ψ TFLOP/s, YC6 M=6000: mixed vs FP64 GPU at |ΔE|/site≤1e-8",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (fp32 accumulate)", "f32": "complex f32", "f64": "complex f64"}[args.prec],
        "data": data,
        "config": line_cfg,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line["config"].update(extra)
    line["config"]["hbm_total_gib"] = round(total / 2 ** 30, 1)
    print(json.dumps(line), flush=True)
    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib * (1 << 20) // 16
        a, b = ctx.vec(n), ctx.vec(n)
        a.fill_gaussian(1, 2, 1.0)
        for _ in range(3):
            b.copy_from(a)
        ctx.sync()
        print(f"# copy calibration: {n * 16} bytes read + {n * 16} bytes written per copy", file=sys.stderr,
              flush=True)


if __name__ == "__main__":
    main()
