"""The workloads of the headline configurations, on the device at their
full sizes, checked against the oracle where it finishes in seconds and by
size-independent properties elsewhere.

* YC4 4x6, M=256, 6 sweeps (the CPU-runnable configuration): whole DMRG in
  FP64 and in mixed precision with FP64 shadow solves; same-path energy
  error, the measured state's sum rules, variational order in M.
* YC4 4x24, M=2000, warm-up and the first half-sweep: same-path energy
  error and the sum rules of the state at the end of the half-sweep.
* The isolated H.psi superblock (400 sectors x 24 terms, d in [8, 1536]):
  operators filled on the device, sampled target sectors against the
  oracle's direct product, every precision against FP64, and the
  per-product ZGEMM loop against the grouped kernels.
"""
import numpy as np
import pytest

from oracle import synth as osynth
from tsc_tri_t1t2j1j2_amd import capi
from tsc_tri_t1t2j1j2_amd import measure as meas
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG
from tsc_tri_t1t2j1j2_amd.synthetic import SyntheticSuperblock

pytestmark = pytest.mark.gpu


def _sum_rules(ctx, d, psi, bases, s1):
    """<n> over all sites = N_e and sum_{j>1} <S_1.S_j> = -3/4 <n_1> (a
    singlet without double occupancy: S_1.S_tot = 0, S_1^2 = 3/4 n_1)."""
    N = d.N
    out = meas.measure_state(ctx, psi, bases, s1, N, d.sys_trun, d.env_trun, 1, list(range(2, N + 1)),
                             kinds=("SS",), density=True)
    ntot = sum(out["density"].values())
    ss = sum(out["SS"].values())
    return ntot, ss, out["density"][1]


@pytest.mark.timeout(900)
def test_yc4_4x6_m256_six_sweeps(ctx):
    """YC4 4x6 (24 sites, 22 electrons), t2 = 0.2, J1 = 0.4, J2 = 0.016,
    kept_max 256, 6 sweeps, lanzero 1e-10."""
    kw = dict(kept_min=110, sweeps=6, lanzero=1e-10, seed=1234)
    f64 = DMRG(ctx, 6, 4, kept_max=256, prec=capi.PREC_F64, **kw)
    e64 = f64.run()
    mix = DMRG(ctx, 6, 4, kept_max=256, prec=capi.PREC_BF16X3, refine=2, shadow_f64=True, **kw)
    emx = mix.run()
    N = f64.N
    same_path = max(abs(em - e) / N for (_, em, e) in mix.shadow)
    print(f"E/site f64 {e64 / N:.12f} mixed {emx / N:.12f}; same-path max |dE|/site {same_path:.2e}; "
          f"two runs {abs(emx - e64) / N:.2e}")
    # the precision gate: every mixed solve against FP64 from the same start
    assert same_path <= 1e-8
    # two whole runs follow different truncation choices; they still agree
    # far below the truncation error's effect on the energy
    assert abs(emx - e64) / N <= 1e-7
    # truncation errors are reported for every sweep step and stay small
    last = [h for h in f64.history if h[0] == "sweep6"]
    assert len(last) == 2 * (N - 3)
    # (a fully kept spectrum leaves a rounding remainder of either sign)
    assert all(-1e-12 <= h[3] < 1e-4 for h in last)
    assert all(h[4] <= 257 for h in last)
    # variational order in the kept multiplets
    small = DMRG(ctx, 6, 4, kept_max=64, prec=capi.PREC_F64, **kw)
    e_small = small.run()
    assert e_small > e64 + 1e-8
    # the measured state (final half-sweep to 5N/6): sum rules
    res = f64.measure(i0=1, js=list(range(2, N + 1)), kinds=("SS",))
    ntot = sum(res["density"].values())
    n1 = res["density"][1]
    ss = sum(res["SS"].values())
    print(f"sum n = {ntot:.12f}; sum S1.Sj = {ss:.12f}; -3/4 n1 = {-0.75 * n1:.12f}")
    assert abs(ntot - f64.Ne) < 1e-9
    assert abs(ss + 0.75 * n1) < 1e-9


@pytest.mark.timeout(900)
def test_yc4_4x24_m2000_warmup_and_half_sweep(ctx):
    """YC4 4x24 (96 sites, 88 electrons), t2 = 0.2, J1 = 0.4, J2 = 0.016,
    kept_max 2000: warm-up and the first half-sweep in mixed precision
    with FP64 shadow solves, then the sum rules of the state at the end of
    the half-sweep."""
    d = DMRG(ctx, 24, 4, kept_min=110, kept_max=2000, sweeps=0, lanzero=1e-10, prec=capi.PREC_BF16X3, refine=2,
             shadow_f64=True, seed=1234)
    d.warm_up()
    e_warm = d.history[-1][2]
    N = d.N
    psi, bases = d.left_to_right(N // 2 - 1, N - 3, d.lanzero * 15, 2000, d.trun_err * 15, "half", keep=True)
    try:
        same_path = max(abs(em - e) / N for (_, em, e) in d.shadow)
        half = [h for h in d.history if h[0] == "half"]
        e_half = min(h[2] for h in half)
        print(f"warm-up E/site {e_warm / N:.10f}; half-sweep min E/site {e_half / N:.10f}; "
              f"same-path max |dE|/site {same_path:.2e}; largest superblock {max(h[5] for h in d.history)}")
        assert same_path <= 1e-8
        assert e_half <= e_warm + 1e-9
        assert max(h[4] for h in half) >= 2000
        ntot, ss, n1 = _sum_rules(ctx, d, psi, bases, N - 2)
        print(f"sum n = {ntot:.12f}; sum S1.Sj = {ss:.12f}; -3/4 n1 = {-0.75 * n1:.12f}")
        assert abs(ntot - d.Ne) < 1e-9
        assert abs(ss + 0.75 * n1) < 1e-9
    finally:
        psi.close()


def _oracle_targets(sb, rng, budget_flops=1.5e11, nrand=5):
    """Target sectors whose oracle evaluation stays within the flop budget:
    the costliest such one, the one with the largest block, and random ones."""
    cost = np.zeros(sb.S)
    for t in range(sb.T):
        for s in range(sb.S):
            g = sb.target[t, s]
            dl, dr, dl2, dr2 = (float(x) for x in (sb.dL[s], sb.dR[s], sb.dL[g], sb.dR[g]))
            cost[g] += 8.0 * (dl2 * dl * dr + dl2 * dr * dr2)
    ok = np.flatnonzero(cost <= budget_flops)
    pick = {int(ok[np.argmax(cost[ok])]), int(ok[np.argmax((sb.dL * sb.dR)[ok])])}
    pick |= set(int(x) for x in rng.choice(ok, size=nrand, replace=False))
    return sorted(pick)


@pytest.mark.timeout(900)
def test_isolated_hpsi_full_size(ctx):
    """400 sectors x 24 terms, d_L, d_R log-uniform in [8, 1536], seed
    20261203, operators filled on the device (2.97e9 complex values)."""
    sb = SyntheticSuperblock(nsec=400, nterms=24, dmin=8, dmax=1536)
    plan = sb.make_plan(ctx, None)
    vin = ctx.vec(sb.psi_len)
    vin.fill_gaussian(sb.seed, 0xC0FFEE, 1.0)
    tmp = ctx.vec(sb.psi_len)
    tmp.axpby(1.0 / vin.norm(), vin, 0.0)
    vin.close()
    vin = tmp
    vout = ctx.vec(sb.psi_len)
    try:
        plan.apply(vin, vout, capi.PREC_F64)
        ref = vout.download()
        psi = vin.download()
        assert abs(np.linalg.norm(psi) - 1.0) < 1e-12
        # sampled targets against the direct product with host-regenerated operators
        rng = np.random.default_rng(5)
        for g in _oracle_targets(sb, rng):
            pairs = [(t, s) for t in range(sb.T) for s in range(sb.S) if sb.target[t, s] == g]
            want = osynth.hpsi(sb, psi, pairs)
            o, n = int(sb.psi_off[g]), int(sb.dL[g] * sb.dR[g])
            err = np.linalg.norm(ref[o:o + n] - want[o:o + n]) / np.linalg.norm(want[o:o + n])
            print(f"target {g} ({sb.dL[g]}x{sb.dR[g]}, {len(pairs)} products): rel err {err:.2e}")
            assert err <= 1e-13
        # the per-product ZGEMM loop (the reference's structure) gives the same vector
        plan.apply_loop(vin, vout)
        loop = vout.download()
        e_loop = np.linalg.norm(loop - ref) / np.linalg.norm(ref)
        print(f"per-product loop vs grouped FP64: {e_loop:.2e}")
        assert e_loop <= 1e-13
        for prec, tol in ((capi.PREC_BF16X3, 2e-5), (capi.PREC_F32, 2e-5)):
            plan.apply(vin, vout, prec)
            got = vout.download()
            e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(f"precision {prec}: rel err vs FP64 {e:.2e}")
            assert e <= tol
    finally:
        for v in (vin, vout):
            v.close()
        plan.close()


def _hist_superblock(ctx, scale):
    """Centre superblock of YC6 6x24 with the real M=6000 sector histogram
    (every sector's multiplets times `scale`) and seeded random blocks:
    blocks on the device, merged operators, native plan."""
    from tsc_tri_t1t2j1j2_amd import hamiltonian as ham
    from tsc_tri_t1t2j1j2_amd import native, resident, workloads
    ss, es, _ = workloads.real_centre_sectors(scale)
    sysb, envb, s1, s2, bonds, ntot = workloads.centre_superblock(24, 6, int(6000 * scale), t2=0.3, J2=0.036,
                                                                  sectors=(ss, es))
    sysd, envd = resident.model_to_device(ctx, sysb), resident.model_to_device(ctx, envb)
    sb = ham.Superblock(sysd, envd, s1, s2, bonds, ntot, 0)
    terms = sb.terms()
    step = resident.StepArena(ctx, [sysd, envd])
    step.add_merges(terms)
    step.finalize()
    tasks = native.NativeTasks(sb, step=step, terms=terms)
    plan = tasks.plan(ctx)
    return sb, (sysd, envd), step, tasks, plan


@pytest.mark.timeout(900)
def test_yc6_m6000_centre_superblock(ctx):
    """The YC6 M=6000 centre superblock at its real size (psi 30,009,747
    multiplet states in 25 sectors): H is Hermitian (<phi|H psi> = <H phi|psi>*),
    sampled output windows equal the host execution of the same plan, the
    contraction orders agree, and BF16x3 stays within its error bound."""
    sb, blocks, step, tasks, plan = _hist_superblock(ctx, 1.0)
    n = sb.super.dim
    assert n == 30009747 and sb.super.num == 25
    vs = [ctx.vec(n) for _ in range(4)]
    try:
        psi, phi, hpsi, hphi = vs
        psi.fill_gaussian(11, 1, 1.0)
        phi.fill_gaussian(11, 2, 1.0)
        plan.apply(psi, hpsi, capi.PREC_F64)
        plan.apply(phi, hphi, capi.PREC_F64)
        x, y, hx, hy = (v.download() for v in vs)
        a, b = np.vdot(y, hx), np.vdot(hy, x)
        print(f"<phi|H psi> = {a:.12e}, <H phi|psi> = {b:.12e}")
        assert abs(a - b) <= 1e-12 * np.linalg.norm(hx) * np.linalg.norm(y)
        # sampled stage-2 windows on the host (numpy over the same arena)
        from oracle import plan_exec
        c1, o1, c2, o2 = tasks.arrays()
        arena = step.vec.download()
        rng = np.random.default_rng(3)
        sizes = o2["m"].astype(np.int64) * o2["n"]
        pick = [int(w) for w in rng.choice(np.flatnonzero(sizes > 0), size=12, replace=False)]
        got = plan_exec.execute_windows(arena, c1, o1, c2, o2, tasks.mid_len, x, pick)
        for w, want in got.items():
            ob = o2[w]
            o, ldc, m, nn = int(ob["c_off"]), int(ob["ldc"]), int(ob["m"]), int(ob["n"])
            idx = (o + np.arange(m)[:, None] + ldc * np.arange(nn)[None, :]).ravel(order="F")
            err = np.linalg.norm(hx[idx] - want.ravel(order="F")) / max(np.linalg.norm(want), 1e-300)
            assert err <= 1e-12, (w, m, nn, err)
        del arena
        # mixed precision against FP64
        plan.apply(psi, hphi, capi.PREC_BF16X3)
        e = np.linalg.norm(hphi.download() - hx) / np.linalg.norm(hx)
        print(f"BF16x3 rel err vs FP64 at YC6 M=6000: {e:.2e}")
        assert e <= 3e-5
    finally:
        for v in vs:
            v.close()
        plan.close()
        tasks.close()
        step.close()
        for m in blocks:
            m.arena.close()


@pytest.mark.timeout(900)
def test_yc6_m12000_centre_superblock(ctx):
    """The YC6 centre superblock with every sector of the real M=6000 centre
    blocks doubled (M=12000; psi 120,038,988 multiplet states in 25 sectors,
    an operator arena of 3.9e9 complex values) fits one MI355X with FP64 and
    BF16x3 state together: H is Hermitian (<phi|H psi> = <H phi|psi>*) and
    BF16x3 stays within its error bound."""
    sb, blocks, step, tasks, plan = _hist_superblock(ctx, 2.0)
    n = sb.super.dim
    assert n == 120038988 and sb.super.num == 25
    vs = [ctx.vec(n) for _ in range(4)]
    try:
        psi, phi, hpsi, hphi = vs
        psi.fill_gaussian(12, 1, 1.0)
        phi.fill_gaussian(12, 2, 1.0)
        plan.apply(psi, hpsi, capi.PREC_F64)
        plan.apply(phi, hphi, capi.PREC_F64)
        x, y, hx, hy = (v.download() for v in vs)
        a, b = np.vdot(y, hx), np.vdot(hy, x)
        print(f"<phi|H psi> = {a:.12e}, <H phi|psi> = {b:.12e}")
        assert abs(a - b) <= 1e-12 * np.linalg.norm(hx) * np.linalg.norm(y)
        del y, hy
        plan.apply(psi, hphi, capi.PREC_BF16X3)
        e = np.linalg.norm(hphi.download() - hx) / np.linalg.norm(hx)
        free, total = ctx.mem_info()
        print(f"BF16x3 rel err vs FP64 at YC6 M=12000: {e:.2e}; HBM in use {(total - free) / 2 ** 30:.1f} GiB")
        assert e <= 3e-5
    finally:
        for v in vs:
            v.close()
        plan.close()
        tasks.close()
        step.close()
        for m in blocks:
            m.arena.close()
