"""Two-site SU(2) DMRG driver on the device hot path.

Reproduces the step order of hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/dmrg_sub.f90 (warm_up_point :974-1147, left_to_right :1151-1302,
right_to_left :1306-1465) and the schedule of DMRG_code/main.f90:114-209:
per step basis -> super basis -> start vector -> Lanczos -> rho ->
truncation -> renormalisation. Every numerical stage runs through the
C-ABI (grouped contraction plans, device Lanczos, device rho/eigensolve,
renormalisation plans); this module only orders the steps and keeps the
blocks. Inside a half-sweep each Lanczos starts from the previous ground
state carried into the new superblock (wave_trans_LR/RL,
dmrg_sub.f90:1471-1789; wavetrans.py); the first step of a half-sweep and
the warm-up start from seeded uniform random numbers (wave_initialize,
block.f90:1154-1180, made reproducible). After run(), measure() does the
final half-sweep and the measurement (Measurement, main.f90:213-354);
run(checkpoint=..., restart=...) writes and reads the per-length files of
the reference (persist.py).
"""
import json
import os
import threading
import time

import numpy as np

from types import SimpleNamespace

from . import capi
from . import resident
from . import hamiltonian as ham
from . import lattice
from . import native
from . import wavetrans
from .blocks import SITE_SECTORS, TotalBlock, TotalModel, site_operators


def fortran_e(x, w, d):
    """Fortran Ew.d: 0.dddE+xx right-aligned in w characters."""
    if x == 0.0:
        m, e = 0.0, 0
    else:
        e = int(np.floor(np.log10(abs(x)))) + 1
        m = x / 10.0 ** e
        if round(abs(m), d) >= 1.0:
            m, e = m / 10.0, e + 1
    body = ("-" if m < 0 else "") + f"{abs(m):.{d}f}"
    return f"{body}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(w)


def site_model(i):
    """Block of one site (Get_hamiltonian_site, dmrg_sub.f90:913-942)."""
    H = TotalBlock(0, 0, {(0, 0): np.zeros((1, 1), complex), (1, 1): np.zeros((1, 1), complex)})
    return TotalModel([i], SITE_SECTORS, H, {i: site_operators()})


def start_vector(super_basis, seed):
    """Uniform [0,1) real and imaginary parts per sector, normalised."""
    rng = np.random.default_rng(seed)
    parts = []
    for (x, y, dx, dy) in super_basis.sub:
        re = rng.random((dx, dy))
        im = rng.random((dx, dy))
        parts.append((re + 1j * im).ravel(order="F"))
    v = np.concatenate(parts) if parts else np.zeros(0, complex)
    return v / np.linalg.norm(v)


class DMRG:
    def __init__(self, ctx, Lx, Ny, t1=1.0, t2=0.2, J1=0.4, J2=None, Ne=None, delta=12, S2=0,
                 kept_min=110, kept_max=300, kept_maxf=None, sweeps=6, lanzero=1e-6, trun_err=1e-15,
                 prec=capi.PREC_F64, refine=None, seed=1234, log=None, native_planner=True, resident=True,
                 lanczos_budget=(100, 4), shadow_f64=False, wave_guess=True, guessed_f64=False,
                 autotune_flops=None, retain=None, warmup_ramp=None, on_plan=None, renorm_prec=capi.PREC_F64):
        self.ctx = ctx
        self.Lx, self.Ny = Lx, Ny
        self.N = Lx * Ny
        self.t1, self.t2, self.J1 = t1, t2, J1
        self.J2 = J1 * (t2 / t1) ** 2 if J2 is None else J2
        self.Ne = lattice.target_electrons(self.N, delta) if Ne is None else Ne
        self.S2 = S2
        self.kept_min, self.kept_max = kept_min, kept_max
        self.kept_maxf = kept_max if kept_maxf is None else kept_maxf
        self.sweeps = sweeps
        self.lanzero = lanzero
        self.trun_err = trun_err
        # refine: steps of the refinement pass after each mixed-precision
        # solve (its first H.psi FP64; tsc_lanczos refine_iters); None gives
        # 2 for a mixed precision and 0 for FP64, 0 turns the pass off
        self.prec = prec
        self.refine = (2 if prec != capi.PREC_F64 else 0) if refine is None else int(refine)
        # (inner iterations, restarts) of each solve: lanczos_memo_new's loop_in=100, loop_out=4 (eigen.f90:14)
        self.lanczos_budget = tuple(int(v) for v in lanczos_budget)
        # shadow_f64: with a mixed precision, also solve every superblock in
        # FP64 from the same start vector (the sweep follows the mixed
        # solution); self.shadow holds (step, E_mixed, E_f64)
        self.shadow_f64 = shadow_f64
        self.shadow = []
        # every mixed solve with a refinement pass: (step, Ritz value of the
        # mixed pass, FP64 <psi|H psi> and FP64 residual norm of its Ritz
        # vector, refined energy, weight of the refined vector outside the
        # mixed one), from the pass's first (FP64) product
        self.precision_record = []
        self.budget_cuts = 0   # solves whose Krylov space was cut to fit in HBM (_budget)
        # wave_guess: inside a half-sweep, start each Lanczos from the previous
        # ground state carried into the new superblock (wave_trans_LR/RL)
        self.wave_guess = wave_guess
        # guessed_f64: solve the steps that start from the guess in FP64 even
        # when prec is a mixed one (a guessed start converges in a few
        # iterations, which the FP64 refinement would repeat anyway)
        self.guessed_f64 = guessed_f64
        self._prec_now = prec
        # renorm_prec: precision of the enlargement + U^H O U plan of the
        # resident path (update_trun_*, Get_Hamiltonian_Sys/Env); FP64 by
        # default, since the renormalised operators enter every later step
        self.renorm_prec = renorm_prec
        self.rho_parts = {}
        # superblocks with at least this many flops per H.psi get their work
        # units tuned (tsc_plan_autotune; the choice is cached per plan shape);
        # None: off (along a sweep the plan shapes change from step to step,
        # so tuning costs more than it saves at M=2000)
        self.autotune_flops = autotune_flops
        self.tuned = {}
        # retain: during the warm-up keep only the blocks of the last `retain`
        # lengths on each side (None: keep all, as the sweeps need them); for
        # runs that stop at the centre, whose stored blocks would not fit in
        # HBM at large M
        self.retain = retain
        # warmup_ramp = (m0, n): the warm-up keeps at most m0 multiplets except
        # in its last n steps, which keep kept_max (reaches the centre blocks
        # of a large kept_max at a fraction of the cost; None: kept_max
        # throughout, as the reference's warm-up)
        self.warmup_ramp = None if warmup_ramp is None else (int(warmup_ramp[0]), int(warmup_ramp[1]))
        # on_plan(dmrg, sb, plan, psi, tasks, step): called in a resident step
        # after the superblock plan and the start vector exist, before the
        # solve (measurement hook; must leave psi unchanged)
        self.on_plan = on_plan
        self.sys_trun, self.env_trun = {}, {}   # block length -> wavetrans.Trunc that built it
        self.seed = seed
        self.native_planner = native_planner
        # block operators stay on the device between steps (resident.py)
        self.resident = resident
        self._dev_cache = {}
        self.bonds = lattice.bonds(Lx, Ny, t1, t2, J1, self.J2)
        self.nb = {}
        for (a, b, t, J) in self.bonds:
            self.nb.setdefault(a, []).append((b, t, J))
            self.nb.setdefault(b, []).append((a, t, J))
        self.sys_blocks = {}
        self.env_blocks = {}
        self.history = []
        self.log = log
        self.nstep = 0
        self.hpsi_flops = 0.0
        self.hpsi_calls = 0
        self.timers = {"plan": 0.0, "guess": 0.0, "lanczos": 0.0, "truncate": 0.0, "renorm": 0.0}
        self.guessed = 0
        # U^dagger O U contractions: algorithmic flops, device ms; rho and
        # truncation: algorithmic HBM bytes (psi read, rho written and
        # factorised in place, U copied out)
        self.renorm_device = [0.0, 0.0]
        self.trunc_bytes = 0.0
        # checkpoint directory: every new block, its truncation and basis,
        # and each step's ground state are written there in the reference's
        # per-length files (persist.py), as model_to_disk / truns_to_disk /
        # basis_to_disk / wave_to_disk do at each step of the reference
        self.checkpoint = None
        # log_dir: Energy_All.dat, Kept_State.dat and entropy1.dat lines per
        # step in the reference's formats (eigen.f90:166-192,
        # den_mat.f90:342-352)
        self.log_dir = None

    # ---- helpers ---------------------------------------------------------
    def _say(self, msg):
        if self.log:
            self.log(msg)

    def _keep_sites(self, sites_in_block, extra_partners=(), side=None):
        """Sites whose operators the grown block carries: those with a bond
        or a partner outside it, and, when the run writes block files, also
        every site of the block's operator slots (slot_sites), which the
        reference's model files hold."""
        inside = set(sites_in_block)
        slots = set(self.slot_sites(side, len(sites_in_block))) if (side and self.checkpoint) else set()
        keep = []
        for i in sites_in_block:
            partners = [j for (j, _, _) in self.nb.get(i, [])] + [j for (a, j) in extra_partners if a == i]
            if i in slots or any(j not in inside for j in partners):
                keep.append(i)
        return keep

    def slot_sites(self, side, L):
        """Chain sites of the operator slots of a `side` ("sys"/"env") block
        of L sites, in slot order (lattice.op_slots; an env block counts its
        sites from the chain's far end)."""
        local = lattice.op_slots(self.bonds, L)
        return local if side == "sys" else [self.N + 1 - i for i in local]

    def _truncate(self, sb, psi_vec, side, kept_min, kept_max, trun_err, identity):
        """U per enlarged sector and the new sector order of the grown block."""
        basis = sb.sys_bs if side == 0 else sb.env_bs
        if identity:
            U = {x: np.eye(d, dtype=complex) for x, (_, _, d, _) in enumerate(basis.sectors)}
            return U, list(range(basis.num)), 0.0, basis.dim
        if self.S2 != 0:
            raise NotImplementedError("truncation implemented for total spin 0 targets")
        sd = [dx for (_, _, dx, _) in sb.super.sub]
        ed = [dy for (_, _, _, dy) in sb.super.sub]
        xs = [x if side == 0 else y for (x, y, _, _) in sb.super.sub]
        twos = [basis.sectors[x][1] for x in xs]
        t0 = time.perf_counter()
        res = capi.rho_truncate(self.ctx, psi_vec, sd, ed, twos, side, kept_min, kept_max, trun_err)
        self.timers["truncate"] += time.perf_counter() - t0
        # the library's own split of that call: rho build, eigensolves,
        # selection, kept eigenvectors to the host
        for name, ms in zip(("rho_build", "rho_eig", "rho_select", "rho_u"), capi.rho_last_times(self.ctx)):
            self.rho_parts[name] = self.rho_parts.get(name, 0.0) + ms / 1e3
        dd = np.asarray(sd if side == 0 else ed, np.float64)
        kk = np.asarray(res["kept"], np.float64)
        self.trunc_bytes += 16.0 * (psi_vec.n + 3.0 * float(np.sum(dd * dd)) + float(np.sum(dd * kk)))
        U, order = {}, []
        for x, u, k in zip(xs, res["U"], res["kept"]):
            if k > 0:
                # env operators act on psi as psi.O^T here, so the env reduced
                # density matrix is conj(psi^H psi): its eigenvectors are the
                # conjugates of those of the reference's rho (den_mat.f90:89)
                U[x] = u if side == 0 else np.conj(u)
                order.append(x)
        if self.log_dir:
            n_enl = len((sb.sys if side == 0 else sb.env).sites) + 1
            self._log_truncation(n_enl if side == 0 else self.N - n_enl, res)
        return U, order, res["trun_err"], res["num_state"]

    def _log(self, name, line):
        with open(os.path.join(self.log_dir, name), "a") as fh:
            fh.write(line + "\n")

    def _log_truncation(self, sys_len, res):
        self._log("entropy1.dat", f"{sys_len:8d}{float(res['entropy1']):20.10f}{self.N:9d}{self.Ne:9d}{self.S2:9d}")
        self._log("Kept_State.dat", f"Sys_len={sys_len:3d}  Num_State={int(res['num_state']):5d}  "
                                    f"Err={fortran_e(float(res['trun_err']), 18, 8)}")

    def _log_energy(self, sys_len, E, iters, seconds):
        if self.log_dir:
            self._log("Energy_All.dat", f"Sys_len={sys_len:3d}  E0=  {E:18.12f}  Iter={iters:4d}  time=  {seconds:12.6f}")

    def _initial(self, sb, guess):
        """Lanczos start: the transformed previous ground state when the
        sweep provides one (wave_trans_LR/RL), else the seeded random vector
        (wave_initialize)."""
        self._prec_now = self.prec
        if guess is not None:
            t0 = time.perf_counter()
            v = guess(sb)
            self.timers["guess"] += time.perf_counter() - t0
            if v is not None:
                self.guessed += 1
                if self.guessed_f64:
                    self._prec_now = capi.PREC_F64
                return v
        return self.ctx.vec_from(start_vector(sb.super, self.seed + self.nstep))

    def _autotune(self, plan, psi):
        """Tune the plan's work units for the precisions this solve uses."""
        if self.autotune_flops is None or plan.flops() < self.autotune_flops:
            return 0.0
        t0 = time.perf_counter()
        scratch = self.ctx.vec(psi.n)
        try:
            precs = {self._prec_now}
            if self._prec_now != capi.PREC_F64 and (self.refine or self.shadow_f64):
                precs.add(capi.PREC_F64)
            for p in sorted(precs):
                self.tuned[p] = plan.autotune(psi, scratch, p)
        finally:
            scratch.close()
        dt = time.perf_counter() - t0
        self.timers["autotune"] = self.timers.get("autotune", 0.0) + dt
        return dt

    def _shadow_start(self, psi):
        if not self.shadow_f64 or self._prec_now == capi.PREC_F64:
            return None
        v = self.ctx.vec(psi.n)
        v.copy_from(psi)
        return v

    def _shadow(self, psi0, plan, tol, E):
        if psi0 is None:
            return
        t0 = time.perf_counter()
        inner, outer = self._budget(psi0.n)
        E64, _ = plan.lanczos(psi0, tol=tol, max_inner=inner, max_outer=outer,
                              prec=capi.PREC_F64)
        psi0.close()
        self.shadow.append((self.nstep, E, E64))
        self.timers["shadow"] = self.timers.get("shadow", 0.0) + time.perf_counter() - t0

    def _budget(self, n):
        """(max_inner, max_outer) of a solve on n-element vectors: the
        configured budget, with the Krylov space cut to what free HBM holds
        (its max_inner + 1 vectors are allocated by the solve, less 4 GiB
        kept free) and the restarts raised so that the total number of
        products stays at least the configured one. Large superblocks (M of
        12000 and more) then run a restarted Lanczos instead of failing to
        allocate."""
        inner, outer = self.lanczos_budget
        mem = getattr(self.ctx, "mem_info", None)
        if mem is None:
            return inner, outer
        free, _ = mem()
        fit = int((free - (4 << 30)) // (16 * max(int(n), 1))) - 1
        if fit >= inner:
            return inner, outer
        fit = max(fit, 8)
        total = inner * (outer + 1)
        self.budget_cuts += 1
        return fit, max(outer, -(-total // fit) - 1)

    def _record_refinement(self, plan):
        if self._prec_now == capi.PREC_F64 or not self.refine:
            return
        r = plan.lanczos_refinement()
        self.precision_record.append((self.nstep, r["e_mixed"], r["e_rayleigh_f64"], r["resid_f64"],
                                      r["e_refined"], r["w_rest"]))

    def precision_by_phase(self):
        """Per phase of the run (warm-up, the half-sweeps, each sweep), from
        the refinement passes of the mixed solves (no extra work):
        max_mixed_ritz_dE_per_site, the largest |E_mixed - <psi|H psi>_FP64|
        / site (the Ritz-energy error of the mixed arithmetic at its own
        solution, before refinement); max_refined_dE_est_per_site, the
        largest w_rest * |E_mixed - <psi|H psi>_FP64| / site (the refined
        energy meets the pass's mixed products only through the refined
        vector's weight w_rest outside psi: a first-order estimate of its
        mixed-arithmetic error); max_refine_shift_per_site, the largest
        (<psi|H psi>_FP64 - E_refined) / site; max_resid_f64, the largest
        FP64 residual norm of the mixed Ritz vectors; the largest truncation
        error of the phase and whether the refined estimate stays below it."""
        trunc = {}
        for h in self.history:
            trunc[h[0]] = max(trunc.get(h[0], 0.0), float(h[3]))
        out = {}
        for (k, em, ea, res, er, wr) in self.precision_record:
            tag = self.history[k][0] if k < len(self.history) else "final"
            o = out.setdefault(tag, {"max_mixed_ritz_dE_per_site": 0.0, "max_refined_dE_est_per_site": 0.0,
                                     "max_refine_shift_per_site": 0.0, "max_resid_f64": 0.0, "solves": 0})
            raw = abs(em - ea) / self.N
            o["max_mixed_ritz_dE_per_site"] = max(o["max_mixed_ritz_dE_per_site"], raw)
            o["max_refined_dE_est_per_site"] = max(o["max_refined_dE_est_per_site"], float(wr) * raw)
            o["max_refine_shift_per_site"] = max(o["max_refine_shift_per_site"], (ea - er) / self.N)
            o["max_resid_f64"] = max(o["max_resid_f64"], float(res))
            o["solves"] += 1
        for tag, o in out.items():
            o["max_trun_err"] = trunc.get(tag, 0.0)
            o["dE_below_trun_err"] = bool(o["max_refined_dE_est_per_site"] <= o["max_trun_err"])
        return out

    def shadow_by_phase(self):
        """Largest same-superblock |E_mixed - E_FP64| / site per phase of the
        run (warm-up, the half-sweeps, each sweep), from the FP64 shadow
        solves: the Ritz-energy error of the mixed precision, per sweep."""
        out = {}
        for (k, em, e64) in self.shadow:
            tag = self.history[k][0] if k < len(self.history) else "final"
            out[tag] = max(out.get(tag, 0.0), abs(em - e64) / self.N)
        return out

    def _dev(self, model):
        """The block with its operators on the device (uploaded once)."""
        if getattr(model, "arena", None) is not None:
            return model
        d = self._dev_cache.get(id(model))
        if d is None or d[0] is not model:
            d = (model, resident.model_to_device(self.ctx, model))
            self._dev_cache[id(model)] = d
        return d[1]

    def step(self, sys, env, s1, s2, bonds, ntot, tol, grow, kept_max, trun_err, identity_sys=False,
             identity_env=False, partners=(), guess=None):
        """One DMRG step; grow: 'sys', 'env' or 'both'. Returns the energy,
        the grown block(s) with the truncations that built them, and the
        ground state (device vector) with its superblock bases. guess(sb):
        Lanczos start vector for superblock sb, or None. out["seconds"] is
        the step's wall time without the FP64 shadow solve, the work-unit
        tuning and the on_plan hook (each kept in its own timer)."""
        aside = ("shadow", "autotune", "hook")
        t_step, sh0 = time.perf_counter(), sum(self.timers.get(k, 0.0) for k in aside)
        if self.resident:
            out = self._step_resident(sys, env, s1, s2, bonds, ntot, tol, grow, kept_max, trun_err,
                                      identity_sys, identity_env, partners, guess)
        else:
            out = self._step_host(sys, env, s1, s2, bonds, ntot, tol, grow, kept_max, trun_err, identity_sys,
                                  identity_env, partners, guess)
        out["seconds"] = time.perf_counter() - t_step - (sum(self.timers.get(k, 0.0) for k in aside) - sh0)
        return out

    def _step_host(self, sys, env, s1, s2, bonds, ntot, tol, grow, kept_max, trun_err, identity_sys,
                   identity_env, partners, guess):
        """step() with the block operators uploaded with each plan."""
        t0 = time.perf_counter()
        sb = ham.Superblock(sys, env, s1, s2, bonds, ntot, self.S2)
        if self.native_planner:
            tasks = native.NativeTasks(sb)
            plan = tasks.plan(self.ctx)
            tasks.close()
        else:
            plan = sb.plan_builder().build(self.ctx, sb.super.dim, sb.super.dim)
        self.timers["plan"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        psi = self._initial(sb, guess)
        t0 += self._autotune(plan, psi)   # reported under "autotune", not "lanczos"
        psi0 = self._shadow_start(psi)
        t1 = time.perf_counter()
        inner, outer = self._budget(psi.n)
        E, (nmv, _) = plan.lanczos(psi, tol=tol, max_inner=inner, max_outer=outer, prec=self._prec_now,
                                   refine_iters=self.refine if self._prec_now != capi.PREC_F64 else 0)
        self.timers["lanczos"] += time.perf_counter() - t0
        self._log_energy(len(sys.sites) + 1, E, nmv, time.perf_counter() - t1)
        self._record_refinement(plan)
        self._shadow(psi0, plan, tol, E)
        self.hpsi_flops += nmv * plan.flops()
        self.hpsi_calls += nmv
        out = {"E": E, "matvecs": nmv, "dim": sb.super.dim, "nsec": sb.super.num}
        if grow in ("sys", "both"):
            U, order, err, ns = self._truncate(sb, psi, 0, self.kept_min, kept_max, trun_err, identity_sys)
            t0 = time.perf_counter()
            to_new = [(i, t, J) for (i, j, t, J) in self._pairs_with(s1, sys.sites, bonds)]
            keep = self._keep_sites(sys.sites + [s1], partners, "sys")
            out["sys"] = ham.renormalize(self.ctx, sys, s1, to_new, keep, U, order)
            out["sys_trun"] = wavetrans.Trunc(sb.sys_bs, U, order)
            self.timers["renorm"] += time.perf_counter() - t0
            out["sys_err"], out["sys_kept"] = err, ns
        if grow in ("env", "both"):
            U, order, err, ns = self._truncate(sb, psi, 1, self.kept_min, kept_max, trun_err, identity_env)
            t0 = time.perf_counter()
            to_new = [(i, t, J) for (i, j, t, J) in self._pairs_with(s2, env.sites, bonds)]
            keep = self._keep_sites(env.sites + [s2], partners, "env")
            out["env"] = ham.renormalize(self.ctx, env, s2, to_new, keep, U, order)
            out["env_trun"] = wavetrans.Trunc(sb.env_bs, U, order)
            self.timers["renorm"] += time.perf_counter() - t0
            out["env_err"], out["env_kept"] = err, ns
        plan.close()
        out["psi"] = psi
        out["bases"] = SimpleNamespace(super=sb.super, sys_bs=sb.sys_bs, env_bs=sb.env_bs)
        self.nstep += 1
        return out

    def _step_resident(self, sys, env, s1, s2, bonds, ntot, tol, grow, kept_max, trun_err, identity_sys,
                       identity_env, partners, guess=None):
        """step() with every operator block on the device: the superblock
        plan, the merged operators and the renormalisation all read one
        StepArena, and the grown block is written to a new device arena."""
        ctx = self.ctx
        t0 = time.perf_counter()
        sysd, envd = self._dev(sys), self._dev(env)
        sb = ham.Superblock(sysd, envd, s1, s2, bonds, ntot, self.S2)
        terms = sb.terms()
        grows = {"sys": grow in ("sys", "both"), "env": grow in ("env", "both")}
        spec = {}
        for side, blk, site in (("sys", sysd, s1), ("env", envd, s2)):
            if grows[side]:
                to_new = [(i, t, J) for (i, j, t, J) in self._pairs_with(site, blk.sites, bonds)]
                keep = self._keep_sites(blk.sites + [site], partners, side)
                spec[side] = (blk, site, to_new, keep, ham.renorm_pairs(blk, to_new))
        step = resident.StepArena(ctx, [sysd, envd])
        made = {}   # device objects of this step, recorded as they appear
        try:
            out = self._solve_and_grow(ctx, sb, step, terms, spec, guess, tol, kept_max, trun_err, identity_sys,
                                       identity_env, sys, t0, made)
        except BaseException:
            # a failed step releases what it allocated on the device
            for key in ("psi0", "psi", "sys", "env"):
                obj = made.get(key)
                obj = getattr(obj, "arena", obj)
                if obj is not None:
                    obj.close()
            raise
        finally:
            if made.get("plan") is not None:
                made["plan"].close()
            step.close()
        self.nstep += 1
        return out

    def _solve_and_grow(self, ctx, sb, step, terms, spec, guess, tol, kept_max, trun_err, identity_sys,
                        identity_env, sys, t0, made):
        """The resident step once its StepArena exists: merges, plan, start
        vector, Lanczos, truncation and renormalisation. Device objects go
        into `made` (plan, psi0, psi, sys, env) as they are created."""
        step.add_merges(terms)
        for v in spec.values():
            step.add_merges(v[4])
        step.finalize()
        tasks = None
        if self.native_planner:
            tasks = native.NativeTasks(sb, step=step, terms=terms)
            plan = tasks.plan(ctx)
        else:
            plan = sb.plan_builder(resolver=step.offset, terms=terms).build_dev(ctx, step.vec, sb.super.dim,
                                                                                sb.super.dim)
        made["plan"] = plan
        self.timers["plan"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        psi = made["psi"] = self._initial(sb, guess)
        if self.on_plan is not None:
            th = time.perf_counter()
            self.on_plan(self, sb, plan, psi, tasks, step)
            dt = time.perf_counter() - th
            t0 += dt   # the hook's time is not the solve's
            self.timers["hook"] = self.timers.get("hook", 0.0) + dt
        if tasks is not None:
            tasks.close()
        t0 += self._autotune(plan, psi)   # reported under "autotune", not "lanczos"
        psi0 = made["psi0"] = self._shadow_start(psi)
        # the truncation-independent half of the renormalisation plans is host
        # bookkeeping: a second thread builds it while the solve runs in the
        # library (ctypes releases the interpreter lock for the call)
        prepared, failure = {}, []

        def prepare():
            try:
                for side, (blk, site, _, keep, pairs) in spec.items():
                    prepared[side] = ham.renorm_transitions(blk, site, keep, pairs)
            except BaseException as exc:   # re-raised by the stepping thread
                failure.append(exc)

        worker = threading.Thread(target=prepare, daemon=True)
        worker.start()
        try:
            t1 = time.perf_counter()
            inner, outer = self._budget(psi.n)
            E, (nmv, _) = plan.lanczos(psi, tol=tol, max_inner=inner, max_outer=outer, prec=self._prec_now,
                                       refine_iters=self.refine if self._prec_now != capi.PREC_F64 else 0)
            self.timers["lanczos"] += time.perf_counter() - t0
            self._log_energy(len(sys.sites) + 1, E, nmv, time.perf_counter() - t1)
            self._record_refinement(plan)
            made["psi0"] = None   # _shadow closes it
            self._shadow(psi0, plan, tol, E)
        finally:
            worker.join()
        if failure:
            raise failure[0]
        self.hpsi_flops += nmv * plan.flops()
        self.hpsi_calls += nmv
        out = {"E": E, "matvecs": nmv, "dim": sb.super.dim, "nsec": sb.super.num, "psi": psi}
        for side, idx, ident in (("sys", 0, identity_sys), ("env", 1, identity_env)):
            if side not in spec:
                continue
            blk, site, to_new, keep, pairs = spec[side]
            U, order, err, ns = self._truncate(sb, psi, idx, self.kept_min, kept_max, trun_err, ident)
            t0 = time.perf_counter()
            rp = ham.renormalize_plan(blk, site, to_new, keep, U, order, pairs=pairs, resolver=step.offset,
                                      prepared=prepared[side])
            out[side] = made[side] = resident.renormalize(ctx, step, rp, self.renorm_prec)
            fl, ms = getattr(out[side], "renorm_stats", (0.0, 0.0))
            self.renorm_device[0] += fl
            self.renorm_device[1] += ms
            out[side + "_trun"] = wavetrans.Trunc(sb.sys_bs if idx == 0 else sb.env_bs, U, order)
            self.timers["renorm"] += time.perf_counter() - t0
            out[side + "_err"], out[side + "_kept"] = err, ns
        out["bases"] = SimpleNamespace(super=sb.super, sys_bs=sb.sys_bs, env_bs=sb.env_bs)
        return out

    @staticmethod
    def _pairs_with(site, block_sites, bonds):
        bs = set(block_sites)
        out = []
        for (a, b, t, J) in bonds:
            if a == site and b in bs:
                out.append((b, site, t, J))
            elif b == site and a in bs:
                out.append((a, site, t, J))
        return out

    # ---- warm-up (infinite algorithm) ---------------------------------
    def warm_up(self, end_len=None):
        N, Ny = self.N, self.Ny
        end_len = N // 2 if end_len is None else end_len
        sys, env = site_model(1), site_model(N)
        self.sys_blocks[1], self.env_blocks[1] = sys, env
        label = 1
        tol = self.lanzero * 1e3
        while label + 1 <= end_len:
            l = label
            nn = 2 * l + 2
            nxc = (nn - 1) // Ny + 1
            if nxc == 1:
                nxc = 2
            tot = Ny * nxc
            present_small = set(range(1, l + 2)) | set(range(tot - l, tot + 1))
            small = lattice.bonds(nxc, Ny, self.t1, self.t2, self.J1, self.J2, present=present_small)

            def full(k):
                return k if k <= l + 1 else N - tot + k

            bonds = [(min(full(a), full(b)), max(full(a), full(b)), t, J) for (a, b, t, J) in small]
            partners = [(full(a), full(b)) for (a, b, _, _) in small] + [(full(b), full(a)) for (a, b, _, _) in small]
            ntot = int((2 * l + 2) / N * self.Ne + 0.001)
            ntot -= ntot % 2
            ntot = min(ntot, self.Ne)
            s1, s2 = l + 1, N - l
            basis_dim_sys = sum(d for (_, _, d) in sys.sectors) * 2
            basis_dim_env = sum(d for (_, _, d) in env.sectors) * 2
            from .blocks import TotalBasis
            kmax = self.kept_max
            if self.warmup_ramp is not None and label + 1 <= end_len - self.warmup_ramp[1]:
                kmax = self.warmup_ramp[0]
            ident_s = TotalBasis(sys.sectors).dim <= kmax
            ident_e = TotalBasis(env.sectors).dim <= kmax
            del basis_dim_sys, basis_dim_env
            r = self.step(sys, env, s1, s2, bonds, ntot, tol, "both", kmax, self.trun_err * 1e2,
                          identity_sys=ident_s, identity_env=ident_e, partners=partners)
            sys, env = r["sys"], r["env"]
            self.sys_blocks[l + 1], self.env_blocks[l + 1] = sys, env
            self.sys_trun[l + 1], self.env_trun[l + 1] = r["sys_trun"], r["env_trun"]
            self._save_block("sys", l + 1)
            self._save_block("env", l + 1)
            self._save_wave(r, l + 1, l + 1)
            r["psi"].close()
            self.history.append(("warmup", l, r["E"], r.get("sys_err", 0.0), r.get("sys_kept", 0), r["dim"]))
            self._save_state()
            self._say(f"warm-up l={l} N_e={ntot} E={r['E']:.12f} dim={r['dim']} nsec={r['nsec']} "
                      f"mv={r['matvecs']} {r['seconds']:.2f}s")
            if self.retain:
                self._drop_blocks(l + 2 - self.retain)
            label += 1

    def _drop_blocks(self, below):
        """Free the stored blocks (and their truncations) of lengths 2 ..
        below-1 on both sides."""
        for blocks, truns in ((self.sys_blocks, self.sys_trun), (self.env_blocks, self.env_trun)):
            for L in [L for L in blocks if 1 < L < below]:
                arena = getattr(blocks.pop(L), "arena", None)
                if arena is not None:
                    arena.close()
                truns.pop(L, None)

    def centre_window(self, nsteps, tol, kept_max, trun_err, tag="window", shadow_from=None):
        """After warm_up(): `nsteps` left-to-right sweep steps from the centre
        (l = N/2-1, ...), as the first half-sweep runs them (the first from
        a random start, the others from the wave-function guess). The blocks
        they grow are dropped afterwards, so the window can be run again from
        the same warm-up blocks in another precision. Returns one record per
        step: energy, wall seconds (without FP64 shadow solves), H.psi calls,
        superblock dimension and sectors, kept states, discarded weight.
        Every call starts from the same seeded random vector. shadow_from =
        k: the steps from the k-th (0-based) on are also solved in FP64 from
        their start vectors (self.shadow); None leaves shadow_f64 as set."""
        N = self.N
        l0 = N // 2 - 1
        if not hasattr(self, "_window_nstep"):
            self._window_nstep = self.nstep
        self.nstep = self._window_nstep
        shadow_setting = self.shadow_f64
        if self.sys_blocks.get(l0) is None or any(self.env_blocks.get(N - l - 2) is None
                                                   for l in range(l0, l0 + nsteps)):
            raise ValueError("the centre window needs the warm-up blocks of lengths N/2-1 and below")
        saved = {L: (self.sys_blocks.get(L), self.sys_trun.get(L)) for L in range(l0 + 1, l0 + nsteps + 1)}
        recs = []
        prev = None
        try:
            for l in range(l0, l0 + nsteps):
                if shadow_from is not None:
                    self.shadow_f64 = l - l0 >= shadow_from
                t_before = dict(self.timers)
                r_before = dict(self.rho_parts)
                b_before = self.trunc_bytes
                guess = self._guess("LR", prev, self.sys_trun.get(l), self.env_trun.get(N - l - 1))
                sys, env = self.sys_blocks[l], self.env_blocks[N - l - 2]
                r = self.step(sys, env, l + 1, l + 2, self.bonds, self.Ne, tol, "sys", kept_max, trun_err,
                              guess=guess)
                if l + 1 in self.sys_blocks and self.sys_blocks[l + 1] is not saved[l + 1][0]:
                    old = getattr(self.sys_blocks[l + 1], "arena", None)
                    if old is not None:
                        old.close()
                self.sys_blocks[l + 1], self.sys_trun[l + 1] = r["sys"], r["sys_trun"]
                if prev is not None:
                    prev[0].close()
                prev = (r["psi"], r["bases"])
                parts = {k: round(v - t_before.get(k, 0.0), 3) for k, v in self.timers.items()
                         if v - t_before.get(k, 0.0) > 0.0}
                truncate_s = {k: round(v - r_before.get(k, 0.0), 4) for k, v in self.rho_parts.items()}
                # algorithmic HBM bytes of the rho / eigensolve / truncation call (psi read, rho written and
                # read, eigenvectors written, kept U read) and the rate over its wall time
                tb = self.trunc_bytes - b_before
                tsec = self.timers["truncate"] - t_before.get("truncate", 0.0)
                recs.append({"l": l, "E": r["E"], "seconds": r["seconds"], "matvecs": r["matvecs"],
                             "dim": r["dim"], "nsec": r["nsec"], "kept": r["sys_kept"], "trun_err": r["sys_err"],
                             "parts_s": parts, "truncate_parts_s": truncate_s, "truncate_bytes": tb,
                             "truncate_gbs": round(tb / tsec / 1e9, 2) if tsec > 0 else None})
                self.history.append((tag, l, r["E"], r["sys_err"], r["sys_kept"], r["dim"]))
                self._say(f"{tag} l={l} E={r['E']:.12f} err={r['sys_err']:.3e} kept={r['sys_kept']} "
                          f"dim={r['dim']} nsec={r['nsec']} mv={r['matvecs']} {r['seconds']:.2f}s")
        finally:
            self.shadow_f64 = shadow_setting
            if prev is not None:
                prev[0].close()
            for L, (blk, tr) in saved.items():
                cur = self.sys_blocks.get(L)
                if cur is not None and cur is not blk:
                    arena = getattr(cur, "arena", None)
                    if arena is not None:
                        arena.close()
                if blk is None:
                    self.sys_blocks.pop(L, None)
                    self.sys_trun.pop(L, None)
                else:
                    self.sys_blocks[L], self.sys_trun[L] = blk, tr
        return recs

    # ---- finite sweeps -------------------------------------------------
    def _sweep_step_sys(self, l, tol, kept_max, trun_err, guess=None):
        N = self.N
        sys, env = self.sys_blocks[l], self.env_blocks[N - l - 2]
        r = self.step(sys, env, l + 1, l + 2, self.bonds, self.Ne, tol, "sys", kept_max, trun_err, guess=guess)
        self.sys_blocks[l + 1] = r["sys"]
        self.sys_trun[l + 1] = r["sys_trun"]
        self._save_block("sys", l + 1)
        self._save_wave(r, l + 1, N - l - 1)
        return r

    def _sweep_step_env(self, l, tol, kept_max, trun_err, guess=None):
        N = self.N
        sys, env = self.sys_blocks[l], self.env_blocks[N - l - 2]
        r = self.step(sys, env, l + 1, l + 2, self.bonds, self.Ne, tol, "env", kept_max, trun_err, guess=guess)
        self.env_blocks[N - l - 1] = r["env"]
        self.env_trun[N - l - 1] = r["env_trun"]
        self._save_block("env", N - l - 1)
        self._save_wave(r, l + 1, N - l - 1)
        return r

    def _guess(self, direction, prev, sys_trun, env_trun):
        """Lanczos start for the next step of a half-sweep, from the previous
        step's ground state (None: seeded random start)."""
        if not self.wave_guess or prev is None or sys_trun is None or env_trun is None:
            return None
        psi_old, bases_old = prev
        ctx = self.ctx
        return lambda sb: wavetrans.transform(ctx, direction, psi_old, bases_old, sb, sys_trun, env_trun)

    def left_to_right(self, l_from, l_to, tol, kept_max, trun_err, tag, keep=False):
        """keep: return the last step's (ground state, bases) instead of
        releasing the state."""
        N = self.N
        prev = None
        for l in range(l_from, l_to + 1):
            guess = self._guess("LR", prev, self.sys_trun.get(l), self.env_trun.get(N - l - 1))
            r = self._sweep_step_sys(l, tol, kept_max, trun_err, guess)
            if prev is not None:
                prev[0].close()
            prev = (r["psi"], r["bases"])
            self.history.append((tag, l, r["E"], r["sys_err"], r["sys_kept"], r["dim"]))
            self._save_state()
            self._say(f"{tag} L->R l={l} E={r['E']:.12f} err={r['sys_err']:.3e} kept={r['sys_kept']} "
                      f"dim={r['dim']} mv={r['matvecs']}")
        if keep:
            return prev
        if prev is not None:
            prev[0].close()

    def right_to_left(self, l_from, l_to, tol, kept_max, trun_err, tag):
        N = self.N
        prev = None
        for l in range(l_from, l_to - 1, -1):
            guess = self._guess("RL", prev, self.sys_trun.get(l + 1), self.env_trun.get(N - l - 2))
            r = self._sweep_step_env(l, tol, kept_max, trun_err, guess)
            if prev is not None:
                prev[0].close()
            prev = (r["psi"], r["bases"])
            self.history.append((tag, l, r["E"], r["env_err"], r["env_kept"], r["dim"]))
            self._save_state()
            self._say(f"{tag} R->L l={l} E={r['E']:.12f} err={r['env_err']:.3e} kept={r['env_kept']} "
                      f"dim={r['dim']} mv={r['matvecs']}")
        if prev is not None:
            prev[0].close()

    # ---- checkpoint / restart (persist.py) ---------------------------------
    def _save_block(self, side, length):
        if not self.checkpoint:
            return
        from . import persist
        blocks, truns = (self.sys_blocks, self.sys_trun) if side == "sys" else (self.env_blocks, self.env_trun)
        d = self.checkpoint
        persist.model_to_disk(os.path.join(d, persist.block_name(side, length)), blocks[length], side,
                              self.slot_sites(side, length))
        tr = truns[length]
        persist.truns_to_disk(os.path.join(d, persist.trun_name(side, length)), tr, length, env=(side == "env"))
        persist.basis_to_disk(os.path.join(d, persist.basis_name(side, length)), tr.basis, length)

    def _save_wave(self, r, sys_len, env_len):
        if not self.checkpoint:
            return
        from . import persist
        b = r["bases"]
        persist.wave_to_disk(os.path.join(self.checkpoint, persist.wave_name(1)), b.super, b.sys_bs, b.env_bs,
                             r["psi"].download(), sys_len, env_len)

    def _save_state(self):
        """Step count and energy history after each step (state.json)."""
        if not self.checkpoint:
            return
        with open(os.path.join(self.checkpoint, "state.json"), "w") as fh:
            json.dump({"N": self.N, "Ne": self.Ne, "nstep": self.nstep, "guessed": self.guessed,
                       "history": [[h[0], int(h[1]), float(h[2]), float(h[3]), int(h[4]), int(h[5])]
                                   for h in self.history]}, fh)

    def load_checkpoint(self, path):
        """Blocks, truncations and step count written by a checkpointed run
        (block_from_disk / truns_from_disk for every stored length)."""
        from . import persist
        from .blocks import TotalBasis
        with open(os.path.join(path, "state.json")) as fh:
            st = json.load(fh)
        if (st["N"], st["Ne"]) != (self.N, self.Ne):
            raise ValueError(f"checkpoint of N={st['N']}, N_e={st['Ne']}, not N={self.N}, N_e={self.Ne}")
        self.nstep, self.guessed = st["nstep"], st["guessed"]
        self.history = [tuple(h) for h in st["history"]]
        self.sys_blocks = {1: site_model(1)}
        self.env_blocks = {1: site_model(self.N)}
        for side, blocks, truns in (("sys", self.sys_blocks, self.sys_trun), ("env", self.env_blocks, self.env_trun)):
            for L in range(2, self.N):
                f = os.path.join(path, persist.block_name(side, L))
                if not os.path.exists(f):
                    continue
                sites = list(range(1, L + 1)) if side == "sys" else list(range(self.N, self.N - L, -1))
                blocks[L] = persist.model_from_disk(f, side, sites, self.slot_sites(side, L))
                prev = blocks.get(L - 1)
                if prev is not None:
                    # a block that was built from an earlier block of length
                    # L - 1 (the pass stopped before rebuilding it) keeps no
                    # truncation: only the measurement and the guess use them
                    try:
                        _, truns[L] = persist.truns_from_disk(os.path.join(path, persist.trun_name(side, L)),
                                                              TotalBasis(prev.sectors), env=(side == "env"))
                    except (KeyError, ValueError):
                        truns.pop(L, None)

    def run(self, checkpoint=None, restart=None, log_dir=None):
        """Get_Ground_State (main.f90:114-209) without the final measurement
        half-sweep; returns the last energy. checkpoint: directory for the
        per-length files; restart: (restart1, sys_len_start, restart2) of
        restart.dat (main.f90:36-39): 1 skips the warm-up and starts the
        first left-to-right half-sweep at sys_len_start (0: its usual start),
        2 also skips that half-sweep and starts the right-to-left one at
        sys_len_start - 1; restart2 != 0 uses the final tolerances at once
        (main.f90:152-155). Blocks come from the checkpoint directory.
        log_dir: where the per-step log files go."""
        N = self.N
        self.checkpoint = checkpoint
        for d in (checkpoint, log_dir):
            if d:
                os.makedirs(d, exist_ok=True)
        self.log_dir = log_dir
        r1, start, r2 = restart if restart else (0, 0, 0)
        if r1 not in (0, 1, 2):
            raise ValueError(f"restart1 = {r1} is not 0, 1 or 2")
        if r1 == 0:
            self.warm_up()
        else:
            if not checkpoint:
                raise ValueError("a restart reads the blocks of a checkpoint directory")
            self.load_checkpoint(checkpoint)
        lanerr = self.lanzero
        tol, terr = lanerr * 15, self.trun_err * 15
        if r1 != 0:
            tol, terr = lanerr * 2, self.trun_err * 2
        if r2 != 0:
            tol, terr = lanerr, self.trun_err
        if r1 <= 1:
            self.left_to_right(start if (r1 == 1 and start) else N // 2 - 1, N - 3, tol, self.kept_max, terr, "half")
        self.right_to_left(start - 1 if (r1 == 2 and start) else N - 3, 1, tol, self.kept_max, terr, "half")
        kept_max = self.kept_max
        for it in range(1, self.sweeps + 1):
            if it == 1:
                kept_max = (self.kept_max + self.kept_maxf) // 2
            if it == 2:
                kept_max = self.kept_maxf
            tol = lanerr * (0.5 if it >= 5 else 1.0)
            self.left_to_right(1, N - 3, tol, kept_max, self.trun_err, f"sweep{it}")
            self.right_to_left(N - 3, 1, tol, kept_max, self.trun_err, f"sweep{it}")
        self._final = (tol, kept_max, self.trun_err if self.sweeps else terr)
        return self.history[-1][2]

    def measure(self, i0=None, js=None, kinds=("SS", "nn", "cc"), l_end=None, density=True, pair_bonds=(1, 3, 5)):
        """After run(): the final left-to-right half-sweep to l_end (5N/6 - 1,
        main.f90:205-206) and the measurement of its last ground state
        (Measurement, main.f90:213-354): densities of every site, the
        correlators `kinds` of site i0 with the sites js (default: the
        reference's site xi1 and the sites of its leg) and, for each bond
        type of pair_bonds, the singlet pair correlators the reference
        writes to Sc_cor_y / _x / _xy (main.f90:259-263)."""
        from . import measure as meas
        N = self.N
        l_end = 5 * N // 6 - 1 if l_end is None else int(l_end)
        if not 1 <= l_end <= N - 3:
            raise ValueError(f"final half-sweep end {l_end} outside 1..{N - 3}")
        d_i0, d_js = meas.reference_sites(self.Lx, self.Ny)
        i0 = d_i0 if i0 is None else int(i0)
        js = [j for j in (d_js if js is None else js) if j > i0]
        tol, kept_max, terr = getattr(self, "_final", (self.lanzero, self.kept_maxf, self.trun_err))
        psi, bases = self.left_to_right(1, l_end, tol, kept_max, terr, "final", keep=True)
        t0 = time.perf_counter()
        try:
            out = meas.measure_state(self.ctx, psi, bases, l_end + 1, N, self.sys_trun, self.env_trun, i0, js,
                                     kinds=kinds, density=density)
            out["pairs"] = {}
            for ij1 in pair_bonds:
                bond, partners = meas.reference_pair_partners(self.Lx, self.Ny, l_end + 1, ij1)
                vals = {}
                if partners:
                    got = meas.measure_pairs(self.ctx, psi, bases, l_end + 1, N, self.sys_trun, self.env_trun, bond,
                                             [p for (_, _, p) in partners])
                    vals = {(sid3, t): got[tuple(sorted(p))] for (sid3, t, p) in partners}
                out["pairs"][ij1] = {"bond": bond, "values": vals}
        finally:
            psi.close()
        out["energy"] = self.history[-1][2]
        out["i0"], out["js"], out["l_end"] = i0, js, l_end
        self.timers["measure"] = time.perf_counter() - t0
        return out
