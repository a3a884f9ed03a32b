"""Run the SU(2) DMRG for the t1-t2-J1-J2 t-J model on a YC cylinder on
one GPU and print a JSON summary (energy per site, truncation error,
timings). One line of progress per DMRG step goes to stderr.

    python -m tsc_tri_t1t2j1j2_amd.run_dmrg --Lx 6 --Ny 4 --kept-max 256 \
        --sweeps 6 --lanzero 1e-10 --prec f64
"""
import argparse
import json
import sys
import time

from . import capi
from .dmrg import DMRG


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--Lx", type=int, default=6)
    ap.add_argument("--Ny", type=int, default=4)
    ap.add_argument("--t1", type=float, default=1.0)
    ap.add_argument("--t2", type=float, default=0.2)
    ap.add_argument("--J1", type=float, default=0.4)
    ap.add_argument("--J2", type=float, default=None)
    ap.add_argument("--delta", type=int, default=12)
    ap.add_argument("--kept-min", type=int, default=110)
    ap.add_argument("--kept-max", type=int, default=256)
    ap.add_argument("--kept-maxf", type=int, default=None)
    ap.add_argument("--sweeps", type=int, default=6)
    ap.add_argument("--lanzero", type=float, default=1e-10)
    ap.add_argument("--trun-err", type=float, default=1e-15)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "bf16x3"])
    ap.add_argument("--refine", type=int, default=2)
    ap.add_argument("--renorm-prec", default="f64", choices=["f64", "f32", "bf16x3"],
                    help="precision of the enlargement + U^H O U plans (default f64: the renormalised operators "
                         "enter every later step)")
    ap.add_argument("--no-guess", action="store_true", help="start every Lanczos from a random vector")
    ap.add_argument("--guessed-f64", action="store_true",
                    help="with a mixed precision, solve the steps that start from the guess in FP64")
    ap.add_argument("--shadow-f64", action="store_true",
                    help="with a mixed precision, also solve every superblock in FP64 (same start) and report "
                         "the per-step energy difference")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--measure", default=None, metavar="DIR",
                    help="after the sweeps, sweep left to right to 5N/6 and write the reference's measurement "
                         "files (densities, spin / hopping / density correlators, singlet pair correlators) to DIR")
    ap.add_argument("--checkpoint", default=None, metavar="DIR",
                    help="write every block, truncation, basis and the current ground state to DIR in the "
                         "reference's per-length files")
    ap.add_argument("--restart", default=None, metavar="FILE",
                    help="restart.dat of the reference ('restart1 sys_len_start' / 'restart2'): 1 or 2 resume "
                         "from the blocks in --checkpoint, 4 only measures the state stored there")
    ap.add_argument("--log-dir", default=None, metavar="DIR",
                    help="append Energy_All.dat, Kept_State.dat and entropy1.dat lines per step to DIR")
    ap.add_argument("--window", type=int, default=0, metavar="K",
                    help="run only the warm-up to the centre, then K left-to-right sweep steps from the centre "
                         "in each precision of --window-precs, every one from the same warm-up blocks; print "
                         "per-step records (blocks of older lengths are freed during the warm-up)")
    ap.add_argument("--window-precs", default="bf16x3,f64",
                    help="precisions of the centre window; a mixed one is also solved in FP64 from the same "
                         "start vectors (same-path energy error)")
    ap.add_argument("--window-lanzero", type=float, default=None,
                    help="Lanczos tolerance base of the centre window (default: --lanzero); the window solves to "
                         "15 times it, as the first half-sweep does")
    ap.add_argument("--warmup-ramp", default=None, metavar="M0,N",
                    help="warm-up keeping at most M0 multiplets except in its last N steps (which keep "
                         "--kept-max)")
    ap.add_argument("--dump-centre", default=None, metavar="FILE",
                    help="with --window: write the sector lists (N, 2S, multiplets) of the two centre blocks "
                         "after the warm-up, and the window records, to FILE (JSON)")
    a = ap.parse_args(argv)
    ctx = capi.Context(0)
    if a.window:
        return centre_window(ctx, a)
    restart = None
    if a.restart:
        from . import persist
        restart = persist.read_restart(a.restart)
        if restart[0] == 4:
            from . import measure as meas
            if not (a.checkpoint and a.measure):
                ap.error("restart1 = 4 measures the state in --checkpoint and needs --measure DIR")
            J2 = a.J1 * (a.t2 / a.t1) ** 2 if a.J2 is None else a.J2
            t1 = time.perf_counter()
            res = meas.measure_from_disk(ctx, a.checkpoint, a.Lx, a.Ny)
            files = meas.write_reference_files(res, a.measure, a.Lx, a.Ny, J2, a.kept_min, a.kept_max,
                                               int(round(sum(res["density"].values()))))
            print(json.dumps({"measure_s": round(time.perf_counter() - t1, 2), "measure_stats": res["stats"],
                              "measure_files": files}), flush=True)
            return
    t0 = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - t0:8.1f}s] {msg}", file=sys.stderr, flush=True)

    prec = capi.PREC_NAMES[a.prec]
    d = DMRG(ctx, a.Lx, a.Ny, t1=a.t1, t2=a.t2, J1=a.J1, J2=a.J2, delta=a.delta, kept_min=a.kept_min,
             kept_max=a.kept_max, kept_maxf=a.kept_maxf, sweeps=a.sweeps, lanzero=a.lanzero, trun_err=a.trun_err,
             prec=prec, refine=a.refine if prec != capi.PREC_F64 else 0, seed=a.seed, log=log,
             wave_guess=not a.no_guess, guessed_f64=a.guessed_f64, shadow_f64=a.shadow_f64,
             renorm_prec=capi.PREC_NAMES[a.renorm_prec])
    E = d.run(checkpoint=a.checkpoint, restart=restart, log_dir=a.log_dir)
    wall = time.perf_counter() - t0
    sweep_rows = [h for h in d.history if h[0].startswith("sweep")]
    last = [h for h in sweep_rows if h[0] == f"sweep{a.sweeps}"]
    out = {
        "lattice": f"YC{a.Ny} {a.Ny}x{a.Lx}", "sites": d.N, "electrons": d.Ne, "kept_max": d.kept_maxf,
        "precision": a.prec, "E": E, "E_per_site": E / d.N,
        "E_min_last_sweep": min(h[2] for h in last) if last else E,
        "max_trun_err_last_sweep": max(h[3] for h in last) if last else None,
        "wall_s": round(wall, 2), "timers_s": {k: round(v, 2) for k, v in d.timers.items()},
        "hpsi_flops_total": d.hpsi_flops, "hpsi_calls": d.hpsi_calls, "steps": len(d.history),
        "guessed_steps": d.guessed, "lanzero": a.lanzero,
    }
    if d.precision_record:
        # every mixed solve, from its refinement pass: |E_mixed - <psi|H psi>_FP64| / site and the FP64
        # residual, per phase, against the phase's truncation error
        out["mixed_precision_by_phase"] = d.precision_by_phase()
    if d.shadow:
        de = [abs(em - e64) / d.N for (_, em, e64) in d.shadow]
        out["same_path_max_dE_per_site"] = max(de)
        out["same_path_last_dE_per_site"] = de[-1]
        out["same_path_max_dE_per_site_by_phase"] = d.shadow_by_phase()
    if a.measure:
        from . import measure as meas
        t1 = time.perf_counter()
        res = d.measure()
        out["measure_s"] = round(time.perf_counter() - t1, 2)
        out["measure_stats"] = res["stats"]
        out["measure_files"] = meas.write_reference_files(res, a.measure, a.Lx, a.Ny, d.J2, d.kept_min,
                                                          d.kept_maxf, d.Ne)
        out["E_measured_state"] = res["energy"]
    print(json.dumps(out), flush=True)


def centre_window(ctx, a):
    """Warm-up to the centre, then the same K centre steps per precision."""
    t0 = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - t0:8.1f}s] {msg}", file=sys.stderr, flush=True)

    precs = [p.strip() for p in a.window_precs.split(",") if p.strip()]
    p0 = capi.PREC_NAMES[precs[0]]
    d = DMRG(ctx, a.Lx, a.Ny, t1=a.t1, t2=a.t2, J1=a.J1, J2=a.J2, delta=a.delta, kept_min=a.kept_min,
             kept_max=a.kept_max, sweeps=0, lanzero=a.lanzero, trun_err=a.trun_err, prec=p0,
             refine=a.refine if p0 != capi.PREC_F64 else 0, seed=a.seed, log=log, retain=a.window + 1,
             warmup_ramp=tuple(int(x) for x in a.warmup_ramp.split(",")) if a.warmup_ramp else None)
    d.warm_up()
    t_warm = time.perf_counter() - t0
    l0 = d.N // 2 - 1
    out = {"lattice": f"YC{a.Ny} {a.Ny}x{a.Lx}", "sites": d.N, "electrons": d.Ne, "kept_max": a.kept_max,
           "couplings": {"t1": d.t1, "t2": d.t2, "J1": d.J1, "J2": d.J2}, "warmup_precision": precs[0],
           "warmup_s": round(t_warm, 2), "warmup_timers_s": {k: round(v, 2) for k, v in d.timers.items()},
           "centre_sys_len": l0, "centre_sys_sectors": [list(map(int, x)) for x in d.sys_blocks[l0].sectors],
           "centre_env_sectors": [list(map(int, x)) for x in d.env_blocks[l0].sectors],
           "centre_sys_sites_kept": sorted(int(i) for i in d.sys_blocks[l0].ops),
           "centre_env_sites_kept": sorted(int(i) for i in d.env_blocks[l0].ops), "windows": {}}
    if a.dump_centre:
        with open(a.dump_centre, "w") as fh:
            json.dump(out, fh)
    lanerr = a.lanzero if a.window_lanzero is None else a.window_lanzero
    tol, terr = lanerr * 15, a.trun_err * 15   # the first half-sweep's tolerances (main.f90:124-131)
    out["window_tol"] = tol
    for name in precs:
        p = capi.PREC_NAMES[name]
        d.prec, d.refine = p, (a.refine if p != capi.PREC_F64 else 0)
        d.shadow_f64, d.shadow = (p != capi.PREC_F64), []
        t1 = dict(d.timers)
        recs = d.centre_window(a.window, tol, a.kept_max, terr, tag=f"window-{name}")
        w = {"steps": recs, "seconds": round(sum(r["seconds"] for r in recs), 3),
             "timers_s": {k: round(v - t1.get(k, 0.0), 3) for k, v in d.timers.items()}}
        if d.shadow:
            de = [abs(em - e64) / d.N for (_, em, e64) in d.shadow]
            w["same_path_dE_per_site"] = de
            w["same_path_max_dE_per_site"] = max(de)
        out["windows"][name] = w
        log(f"window {name}: {w['seconds']:.2f}s {w['timers_s']}")
        if a.dump_centre:
            with open(a.dump_centre, "w") as fh:
                json.dump(out, fh)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
