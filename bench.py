"""H.psi throughput bench on one MI355X.

A step is one superblock H.psi application of the grouped two-stage
contraction (stage 1: O_L psi or psi O_R^T into intermediates; stage 2:
per target sector sum of c * intermediate x operator), inputs resident in
HBM, replayed from a hipGraph. The headline precision is the mixed path
(BF16x3 products, FP32 accumulation); the FP64 path is timed beside it
and the relative ||H psi|| error of the mixed path against FP64 is
reported. The CPU baseline is the numpy (OpenBLAS) oracle on a bounded
sample of the same superblock.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--workload synthetic400] [--prec bf16x3|f32|f64]
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


# HBM bytes per launch of the dominant kernel (stage 2 of H.psi), YC6 M=6000
# centre superblock, from two rocprofv3 runs of
#   bench.py --steps 2 --warmup 1 --no-cpu --no-lanczos --calibrate-copy-mib 1024
# with --pmc FETCH_SIZE and --pmc WRITE_SIZE; each counter is scaled by its
# reading for the 1 GiB copy in the same run (FETCH_SIZE x 2.0, WRITE_SIZE
# x 1.0). Per-dispatch KiB of the BF16x3 pk_gemm_kernel for each work-unit
# target the autotuner may keep: 2048: 14007535 fetched, 622965 written;
# 4096: 13638127 / 826403; 8192: 12844715 / 1239763.
MEASURED_TRAFFIC = {
    ("yc6-m6000", "bf16x3", 2048.0): (14007535 * 2.0 + 622965 * 1.0) * 1024,
    ("yc6-m6000", "bf16x3", 4096.0): (13638127 * 2.0 + 826403 * 1.0) * 1024,
    ("yc6-m6000", "bf16x3", 8192.0): (12844715 * 2.0 + 1239763 * 1.0) * 1024,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="yc6-m6000",
                    choices=["yc6-m6000", "yc6-m12000", "synthetic400", "synthetic-small"])
    ap.add_argument("--no-lanczos", action="store_true", help="H.psi only (no Lanczos, truncation or DMRG check)")
    ap.add_argument("--lanczos-iters", type=int, default=40)
    ap.add_argument("--refine", type=int, default=2, help="steps of the refinement Lanczos pass after the mixed one (its first H.psi FP64)")
    ap.add_argument("--no-dmrg", action="store_true", help="skip the FP64-vs-mixed DMRG energy check")
    ap.add_argument("--no-full-step", action="store_true", help="skip the resident full-step timing")
    ap.add_argument("--dmrg-sweeps", type=int, default=6)
    ap.add_argument("--prec", default="bf16x3", choices=["bf16x3", "f32", "f64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the measurement, copy a buffer of this many MiB three times (a kernel of known "
                         "HBM traffic, to scale hardware byte counters)")
    return ap.parse_args()


class PhysicsWorkload:
    """Centre superblock of a YC cylinder at M multiplets: the real term
    list and sector structure of one DMRG step, seeded random blocks."""

    def __init__(self, Lx, Ny, M, t2=0.2, seed=7):
        from tsc_tri_t1t2j1j2_amd import hamiltonian as ham
        from tsc_tri_t1t2j1j2_amd import workloads
        t0 = time.perf_counter()
        sysb, envb, s1, s2, bonds, ntot = workloads.centre_superblock(Lx, Ny, M, t2=t2, seed=seed)
        self.t2 = t2
        self.blocks = (sysb, envb, s1, s2, bonds, ntot)
        self.Lx, self.Ny, self.M = Lx, Ny, M
        self.sb = ham.Superblock(sysb, envb, s1, s2, bonds, ntot, 0)
        self.pb = self.sb.plan_builder()
        self.c1, self.o1, self.c2, self.o2 = self.pb.arrays()
        self.ops = self.pb.arena_array()
        self.psi_len = self.sb.super.dim
        self.ops_len = self.ops.size
        self.N = Lx * Ny
        self.plan_seconds = time.perf_counter() - t0

    def stage_flops(self):
        def fl(c, o):
            f = 0.0
            for ob in o:
                m, n = float(ob["m"]), float(ob["n"])
                for x in c[ob["cbeg"]:ob["cend"]]:
                    f += 8.0 * m * n * (x["k"] if x["a_op"] != 4 and x["b_op"] != 4 else 1.0)
            return f
        return fl(self.c1, self.o1), fl(self.c2, self.o2)

    def make_plan(self, ctx):
        from tsc_tri_t1t2j1j2_amd import capi
        return capi.Plan(ctx, self.ops, self.ops.size, self.pb.mid_len, self.c1, self.o1, self.c2, self.o2,
                         self.psi_len, self.psi_len)


def build_workload(name):
    from tsc_tri_t1t2j1j2_amd.synthetic import SyntheticSuperblock
    if name == "yc6-m6000":
        wl = PhysicsWorkload(24, 6, 6000, t2=0.3)
        desc = {"workload": "YC6 6x24 (144 sites, 132 electrons, S=0) centre superblock at M=6000 SU(2) "
                            "multiplets: real t1-t2-J1-J2 term list and sector structure, seeded random blocks",
                "model": "t1=1 t2=0.3 J1=0.4 J2=0.036", "psi_multiplet_dim": wl.psi_len,
                "sectors": wl.sb.super.num, "plan_build_s": round(wl.plan_seconds, 2)}
        return wl, desc
    if name == "yc6-m12000":
        wl = PhysicsWorkload(24, 6, 12000, t2=0.2)
        desc = {"workload": "YC6 6x24 (144 sites, 132 electrons, S=0) centre superblock at M=12000 SU(2) "
                            "multiplets: real t1-t2-J1-J2 term list and sector structure, seeded random blocks",
                "model": "t1=1 t2=0.2 J1=0.4 J2=0.016", "psi_multiplet_dim": wl.psi_len,
                "sectors": wl.sb.super.num, "plan_build_s": round(wl.plan_seconds, 2)}
        return wl, desc
    if name == "synthetic400":
        sb = SyntheticSuperblock(nsec=400, nterms=24, dmin=8, dmax=1536)
        desc = {"workload": "synthetic superblock H.psi: 400 SU(2) sectors, d_L,d_R log-uniform [8,1536], "
                            "24 terms O_L psi O_R^T, seed 20261203"}
    elif name == "synthetic-small":
        sb = SyntheticSuperblock(nsec=60, nterms=12, dmin=8, dmax=400)
        desc = {"workload": "synthetic superblock H.psi: 60 sectors, d in [8,400], 12 terms"}
    else:
        raise SystemExit(f"unknown workload {name}")
    return sb, desc


def stage_flops(sb):
    """Algorithmic flops of stage 1 and stage 2 (each (t,s) at its cheaper order)."""
    f1 = f2 = 0.0
    for t in range(sb.T):
        for s in range(sb.S):
            g = sb.target[t, s]
            dl, dr, dl2, dr2 = (int(x) for x in (sb.dL[s], sb.dR[s], sb.dL[g], sb.dR[g]))
            lf, rf = sb.order_costs(t, s)
            if lf <= rf:
                f1 += 8.0 * dl2 * dl * dr
                f2 += 8.0 * dl2 * dr * dr2
            else:
                f1 += 8.0 * dl * dr * dr2
                f2 += 8.0 * dl2 * dl * dr2
    return f1, f2


def time_gpu(ctx, plan, vin, vout, prec, steps, warmup):
    from tsc_tri_t1t2j1j2_amd import capi
    # per-stage kernel durations from HIP events on the plan's stream
    s1 = s2 = 0.0
    for _ in range(max(1, warmup)):
        plan.apply(vin, vout, prec)
        ctx.sync()
        a, b = plan.last_times()
        s1, s2 = a, b
    ev1, ev2 = [], []
    for _ in range(steps):
        plan.apply(vin, vout, prec)
        a, b = plan.last_times()
        ev1.append(a)
        ev2.append(b)
    # timed region: graph replays, synchronised on both sides
    plan.apply_graph(vin, vout, prec)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.apply_graph(vin, vout, prec)
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    del s1, s2, capi
    return wall, float(np.mean(ev1)), float(np.mean(ev2))


def lanczos_and_step(ctx, plan, wl, vin, prec, args, hbm_mark=None):
    """Device time of one DMRG step's hot path on the centre superblock in
    FP64 and in mixed precision: Lanczos at a fixed Krylov budget (FP64: `it`
    H.psi; mixed: `it` mixed H.psi, then a refinement pass of `refine` steps
    restarted from the mixed Ritz vector, whose first H.psi is FP64 and the
    later ones mixed), then rho / eigensolve / truncation. The sweep time is
    extrapolated from it. The blocks are random, so these Lanczos energies
    are not converged and their difference measures convergence, not
    precision; the precision check is dmrg_precision()."""
    from tsc_tri_t1t2j1j2_amd import capi
    it, refine = args.lanczos_iters, args.refine
    psi64 = ctx.vec(wl.psi_len)
    psi64.copy_from(vin)
    ctx.sync()
    t0 = time.perf_counter()
    e64, (mv64, _) = plan.lanczos(psi64, tol=1e-13, max_inner=it, max_outer=0, prec=capi.PREC_F64)
    t64 = time.perf_counter() - t0
    psim = ctx.vec(wl.psi_len)
    psim.copy_from(vin)
    ctx.sync()
    t0 = time.perf_counter()
    em, (mvm, _) = plan.lanczos(psim, tol=1e-13, max_inner=it, max_outer=0, prec=prec, refine_iters=refine)
    tm = time.perf_counter() - t0
    if hbm_mark is not None:
        hbm_mark()
    print(f"# lanczos f64 {t64:.2f}s ({mv64} H.psi) E={e64:.12f}; mixed {tm:.2f}s ({mvm} H.psi) E={em:.12f}",
          file=sys.stderr, flush=True)
    sbk = wl.sb
    sd = [dx for (_, _, dx, _) in sbk.super.sub]
    ed = [dy for (_, _, _, dy) in sbk.super.sub]
    twos = [sbk.sys_bs.sectors[x][1] for (x, _, _, _) in sbk.super.sub]
    # the first call also creates the solver handles and streams: time the second
    capi.rho_truncate(ctx, psim, sd, ed, twos, 0, 110, wl.M, 1e-15)
    ctx.sync()
    t0 = time.perf_counter()
    tr = capi.rho_truncate(ctx, psim, sd, ed, twos, 0, 110, wl.M, 1e-15)
    ttr = time.perf_counter() - t0
    s64, smx = t64 + ttr, tm + ttr
    return {"lanczos_budget": f"FP64 {it} H.psi; mixed {it} + a {max(2, refine)}-step refinement pass (its first H.psi FP64)",
            "lanczos_f64_s": round(t64, 3), "lanczos_mixed_s": round(tm, 3), "rho_truncate_s": round(ttr, 3),
            "kept": int(tr["num_state"]), "lanczos_rho_s_f64": round(s64, 3), "lanczos_rho_s_mixed": round(smx, 3)}


def full_step(ctx, wl, prec, args):
    """Wall time of one whole sweep step on the centre superblock with the
    block operators resident on the device, as dmrg.DMRG runs it: step arena
    + merged operators, superblock plan, Lanczos at the fixed budget of
    lanczos_and_step(), rho / eigensolve / truncation of the system side, and
    the renormalised (enlarged + truncated) system block written to a new
    device arena. The first upload of the two blocks is timed apart: in a
    sweep they are already resident."""
    from tsc_tri_t1t2j1j2_amd import capi
    from tsc_tri_t1t2j1j2_amd.dmrg import DMRG
    sysb, envb, s1, s2, bonds, ntot = wl.blocks
    out = {}
    for name, p, rf in (("f64", capi.PREC_F64, 0), ("mixed", prec, args.refine)):
        # the step's plan has the shape bench tuned above: the cached unit cut is applied
        d = DMRG(ctx, wl.Lx, wl.Ny, t2=wl.t2, kept_min=110, kept_max=wl.M, prec=p, refine=rf, resident=True,
                 lanczos_budget=(args.lanczos_iters, 0), autotune_flops=1e11)
        t0 = time.perf_counter()
        sysd, envd = d._dev(sysb), d._dev(envb)
        ctx.sync()
        tup = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = d.step(sysd, envd, s1, s2, bonds, ntot, 1e-13, "sys", wl.M, 1e-15)
        ctx.sync()
        # the work-unit tuning of the first plan of a shape is cached for the
        # rest of a sweep: report it apart from the step
        tstep = time.perf_counter() - t0 - d.timers.get("autotune", 0.0)
        tm = {k: round(v, 3) for k, v in d.timers.items()}
        rfl, rms = d.renorm_device
        tr_s = d.timers["truncate"]
        side = {"renorm_uou": {"bound": "mfma", "achieved": round(rfl / (rms * 1e-3) / 1e12, 3) if rms else None,
                               "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s", "flops": rfl,
                               "device_ms": round(rms, 3)},
                "rho_truncate": {"bound": "hbm", "achieved": round(d.trunc_bytes / tr_s / 1e9, 3) if tr_s else None,
                                 "peak": 8000.0, "unit": "GB/s", "bytes": d.trunc_bytes,
                                 "wall_ms": round(tr_s * 1e3, 2)}}
        print(f"# full step {name}: {tstep:.2f}s {tm} kept {r['sys_kept']} E={r['E']:.10f} "
              f"(upload of both blocks {tup:.2f}s)", file=sys.stderr, flush=True)
        out[name] = (tstep, tm, r["sys_kept"], tup, side)
        del r, d, sysd, envd
    nsteps = 2 * (wl.N - 3)
    s64, smx = out["f64"][0], out["mixed"][0]
    return {"full_step_s_f64": round(s64, 3), "full_step_s_mixed": round(smx, 3),
            "full_step_parts_f64": out["f64"][1], "full_step_parts_mixed": out["mixed"][1],
            "full_step_kept": int(out["mixed"][2]), "block_upload_s": round(out["mixed"][3], 3),
            "full_step_other_rooflines_f64": out["f64"][4],
            "sweep_s_f64_estimate": round(s64 * nsteps, 1), "sweep_s_mixed_estimate": round(smx * nsteps, 1),
            "sweep_speedup_mixed_vs_f64": round(s64 / smx, 3),
            "sweep_estimate_note": f"{nsteps} centre-sized steps x one measured full step (step arena and merges, "
                                   "plan, Lanczos at the fixed budget from a random start, rho/truncation, "
                                   "renormalisation); in a sweep the wave-function guess lowers the H.psi count per "
                                   "step (measured_dmrg runs whole sweeps at a smaller size)"}


def measured_dmrg(ctx, refine, Lx=8, Ny=4, M=1000, sweeps=2):
    """Whole DMRG runs (warm-up, two half-sweeps, `sweeps` sweeps; resident
    operators, wave-function guess) on YC4 Lx x 4 at kept_max M, FP64 and
    mixed: measured wall time, per-phase times and H.psi counts. The mixed
    run solves unguessed starts in BF16x3 + FP64 refinement and guessed ones
    in FP64 (a guessed start converges in a few iterations)."""
    from tsc_tri_t1t2j1j2_amd import capi
    from tsc_tri_t1t2j1j2_amd.dmrg import DMRG
    out = {}
    for name, prec, rf in (("f64", capi.PREC_F64, 0), ("mixed", capi.PREC_BF16X3, refine)):
        mv = []
        t0 = time.perf_counter()
        d = DMRG(ctx, Lx, Ny, kept_min=110, kept_max=M, sweeps=sweeps, lanzero=1e-10, prec=prec, refine=rf,
                 guessed_f64=True)
        orig = d.step

        def step(*a, **k):
            r = orig(*a, **k)
            mv.append(r["matvecs"])
            return r

        d.step = step
        E = d.run()
        el = time.perf_counter() - t0
        out[name] = {"wall_s": round(el, 2), "E_per_site": E / d.N, "steps": d.nstep, "guessed_steps": d.guessed,
                     "hpsi_calls": int(sum(mv)), "parts_s": {k: round(v, 2) for k, v in d.timers.items()}}
        print(f"# dmrg YC4 {Lx}x{Ny} M={M} {name}: E={E:.12f} wall {el:.1f}s steps {d.nstep} "
              f"guessed {d.guessed} H.psi {sum(mv)} {out[name]['parts_s']}", file=sys.stderr, flush=True)
    return {"measured_dmrg": f"YC4 {Lx}x{Ny} ({Lx * Ny} sites), kept_max {M}, warm-up + 2 half-sweeps + {sweeps} "
                             f"sweeps, lanzero 1e-10, wave-function guess",
            "measured_dmrg_f64": out["f64"], "measured_dmrg_mixed": out["mixed"],
            "measured_dmrg_speedup_mixed_vs_f64": round(out["f64"]["wall_s"] / out["mixed"]["wall_s"], 3)}


def dmrg_precision(ctx, sweeps, refine):
    """|dE|/site of the mixed precision (BF16x3, then per solve a refinement
    pass of `refine` steps whose first H.psi is FP64) on YC4 4x6 (24 sites, 22 electrons, kept_max 256), two ways:
    (1) two whole runs, FP64 and mixed: their difference also contains the
    effect of different truncation choices along the two paths; (2) along the
    mixed run, every superblock is also solved in FP64 from the same start
    vector: the largest per-step difference is the precision of the mixed
    solve itself."""
    from tsc_tri_t1t2j1j2_amd import capi
    from tsc_tri_t1t2j1j2_amd.dmrg import DMRG
    out = {}
    shadow = []
    by_phase = {}
    for name, prec, rf in (("f64", capi.PREC_F64, 0), ("bf16x3", capi.PREC_BF16X3, refine)):
        t0 = time.perf_counter()
        d = DMRG(ctx, 6, 4, kept_min=110, kept_max=256, sweeps=sweeps, lanzero=1e-10, prec=prec, refine=rf,
                 seed=1234, shadow_f64=(prec != capi.PREC_F64))
        E = d.run()
        out[name] = (E, time.perf_counter() - t0, d.N)
        shadow = shadow or d.shadow
        by_phase = by_phase or d.shadow_by_phase()
        print(f"# dmrg YC4 6x4 M=256 {name}: E={E:.12f} in {out[name][1]:.1f}s", file=sys.stderr, flush=True)
    N = out["f64"][2]
    per_step = [abs(em - e64) / N for (_, em, e64) in shadow]
    return {"dmrg_check": f"YC4 4x6 (24 sites, 22 electrons) kept_max 256, {sweeps} sweeps, lanzero 1e-10",
            "dmrg_E_per_site_f64": out["f64"][0] / N, "dmrg_E_per_site_mixed": out["bf16x3"][0] / N,
            "dmrg_dE_per_site_mixed_vs_f64": abs(out["f64"][0] - out["bf16x3"][0]) / N, "dmrg_refine": refine,
            "dmrg_same_path_steps": len(per_step),
            "dmrg_same_path_max_dE_per_site": max(per_step) if per_step else None,
            "dmrg_same_path_last_dE_per_site": per_step[-1] if per_step else None,
            "dmrg_same_path_max_dE_per_site_by_phase": {k: float(f"{v:.3e}") for k, v in by_phase.items()},
            "dmrg_wall_s_f64": round(out["f64"][1], 2),
            "dmrg_wall_s_mixed_with_fp64_shadow": round(out["bf16x3"][1], 2)}


def cpu_baseline(sb, psi_host, seconds):
    """numpy oracle on the first (t, s) pairs until ~`seconds` of work."""
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    pairs = [(t, s) for s in range(sb.S) for t in range(sb.T)]
    rng = np.random.default_rng(0)
    rng.shuffle(pairs)
    flops = 0.0
    elapsed = 0.0
    n = 0
    for t, s in pairs:
        g = int(sb.target[t, s])
        OL, OR = sb.op_L(t, s), sb.op_R(t, s)
        x = sb.psi_sector(psi_host, s)
        lf, rf = sb.order_costs(t, s)
        t0 = time.perf_counter()
        if lf <= rf:
            y = (OL @ x) @ OR.T
        else:
            y = OL @ (x @ OR.T)
        y *= sb.coef[t, s]
        elapsed += time.perf_counter() - t0
        flops += 8.0 * min(lf, rf)
        n += 1
        del g
        if elapsed >= seconds:
            break
    return {"value": flops / elapsed / 1e12, "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
            "sample": f"numpy/OpenBLAS oracle on {n} random (term, sector) pairs of the same superblock, "
                      f"{flops / 1e12:.3f} TFLOP in {elapsed:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only)"}))
        sys.exit(2)
    from tsc_tri_t1t2j1j2_amd import capi
    prec = capi.PREC_NAMES[args.prec]
    sb, desc = build_workload(args.workload)
    t_setup = time.perf_counter()
    ctx = capi.Context(0)
    free0, total = ctx.mem_info()
    hbm = {"peak": 0}

    def hbm_mark():
        # device memory in use beyond what was in use when the context was made
        free, _ = ctx.mem_info()
        hbm["peak"] = max(hbm["peak"], free0 - free)

    plan = sb.make_plan(ctx) if isinstance(sb, PhysicsWorkload) else sb.make_plan(ctx, None)
    vin = ctx.vec(sb.psi_len)
    vin.fill_gaussian(getattr(sb, "seed", 20261203), 0xC0FFEE, 1.0)
    nrm = vin.norm()
    tmp = ctx.vec(sb.psi_len)
    tmp.axpby(1.0 / nrm, vin, 0.0)
    vin, tmp = tmp, vin
    vout = ctx.vec(sb.psi_len)
    vref = ctx.vec(sb.psi_len)
    print(f"# setup {time.perf_counter() - t_setup:.1f} s, psi {sb.psi_len} complex, ops {sb.ops_len} complex",
          file=sys.stderr, flush=True)
    if isinstance(sb, PhysicsWorkload):
        f1, f2 = sb.stage_flops()
        F = f1 + f2
    else:
        F = sb.algorithmic_flops()
        f1, f2 = stage_flops(sb)

    # FP64 reference application and timing
    hbm_mark()
    t_tune = time.perf_counter()
    tuned = {"f64": plan.autotune(vin, vout, capi.PREC_F64)}
    if prec != capi.PREC_F64:
        tuned[args.prec] = plan.autotune(vin, vout, prec)
    tuned["seconds"] = round(time.perf_counter() - t_tune, 2)
    print(f"# work-unit targets after tuning: {tuned}", file=sys.stderr, flush=True)
    w64, k1_64, k2_64 = time_gpu(ctx, plan, vin, vref, capi.PREC_F64, args.steps, args.warmup)
    hbm_mark()
    print(f"# f64: {w64 * 1e3:.2f} ms/step, {F / w64 / 1e12:.2f} TFLOP/s (stage 1 {k1_64:.2f} ms, stage 2 "
          f"{k2_64:.2f} ms)", file=sys.stderr, flush=True)
    ref = vref.download()
    if prec != capi.PREC_F64:
        wall, k1, k2 = time_gpu(ctx, plan, vin, vout, prec, args.steps, args.warmup)
        hbm_mark()
        out = vout.download()
        rel = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    else:
        wall, k1, k2, rel = w64, k1_64, k2_64, 0.0
    print(f"# {args.prec}: {wall * 1e3:.2f} ms/step, {F / wall / 1e12:.2f} TFLOP/s, rel err {rel:.2e} (stage 1 "
          f"{k1:.2f} ms, stage 2 {k2:.2f} ms)", file=sys.stderr, flush=True)

    # dominant kernel roofline (MFMA-bound)
    if k1 >= k2:
        dom, kms, kf = "grouped stage 1", k1, f1
    else:
        dom, kms, kf = "grouped stage 2", k2, f2
    achieved = kf / (kms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.prec]
    traffic = (MEASURED_TRAFFIC.get((args.workload, args.prec, tuned.get(args.prec)))
               if dom == "grouped stage 2" else None)
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 5), "traffic": traffic, "kernel": dom,
            "kernel_ms": round(kms, 3), "flops_per_launch": kf}
    if args.prec == "bf16x3":
        # every real product is three bf16 MFMA products (hi*hi + hi*lo + lo*hi),
        # so the path's own ceiling is a third of the bf16 peak
        roof["mfma_products_per_product"] = 3
        roof["frac_of_bf16x3_ceiling"] = round(3.0 * achieved / peak, 5)
    if traffic:
        roof["traffic_tb_per_s"] = round(traffic / (kms * 1e-3) / 1e12, 3)

    extra = {}
    if isinstance(sb, PhysicsWorkload) and not args.no_lanczos:
        extra = lanczos_and_step(ctx, plan, sb, vin, prec, args, hbm_mark)
        if not args.no_full_step:
            plan.close()
            extra.update(full_step(ctx, sb, prec, args))
    if isinstance(sb, PhysicsWorkload) and not (args.no_lanczos or args.no_dmrg) and prec != capi.PREC_F64:
        extra.update(dmrg_precision(ctx, args.dmrg_sweeps, args.refine))
        extra.update(measured_dmrg(ctx, args.refine))
    cpu = None
    if not args.no_cpu:
        psi_host = vin.download()
        if isinstance(sb, PhysicsWorkload):
            from oracle import plan_exec
            try:
                from threadpoolctl import threadpool_info
                threads = max([d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"]
                              or [1])
            except Exception:
                threads = os.cpu_count() or 1
            fl, el, nw = plan_exec.timed_sample(sb.ops, sb.c1, sb.o1, sb.c2, sb.o2, sb.pb.mid_len, psi_host,
                                                args.cpu_seconds)
            cpu = {"value": fl / el / 1e12, "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
                   "sample": f"numpy/OpenBLAS execution of {nw} random output windows of the same H.psi plan "
                             f"(with their intermediates), {fl / 1e12:.3f} TFLOP in {el:.1f} s"}
        else:
            cpu = cpu_baseline(sb, psi_host, args.cpu_seconds)

    line = {
        "metric": "sweep time + H\u00b7\u03c8 TFLOP/s, YC6 M=6000: mixed vs FP64 GPU at |\u0394E|/site\u22641e-8",
        "value": round(F / wall / 1e12, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (fp32 accumulate)", "f32": "complex f32", "f64": "complex f64"}[args.prec],
        "data": ("synthetic: seeded random block operators with the sector structure and term list of a real "
                 "DMRG step; random unit psi" if isinstance(sb, PhysicsWorkload)
                 else "synthetic (seeded counter-based Gaussian operators and psi)"),
        "config": dict(desc, algorithmic_flops_per_step=F, precision=args.prec,
                       f64_ms_per_step=round(w64 * 1e3, 4), f64_tflops=round(F / w64 / 1e12, 4),
                       speedup_vs_f64=round(w64 / wall, 3), rel_err_vs_f64=rel, parallelism="1 GPU",
                       autotune_units=tuned),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line["config"].update(extra)
    line["config"]["hbm_peak_gib"] = round(hbm["peak"] / 2 ** 30, 2)
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
