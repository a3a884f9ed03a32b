"""Device parity of the SU(2) physics path: model_wave and block
renormalisation plans against the oracle's dense superblock, a truncating
DMRG step against the oracle's step, and whole DMRG runs against exact
diagonalisation."""
import numpy as np
import pytest

from oracle import dmrg_exact, su2_dense, tj_ed, tj_model
from oracle import lanczos as olanczos
from oracle import truncation as otrunc
from tests.model_convert import block_to_product, oracle_superblock
from tsc_tri_t1t2j1j2_amd import capi
from tsc_tri_t1t2j1j2_amd import hamiltonian as ham
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG, start_vector

pytestmark = pytest.mark.gpu


def _setup(Lx, Ny, Ne, S2, l):
    N = Lx * Ny
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    sys_blk, env_blk, nb = dmrg_exact.build_chain_blocks(N, b, l)
    osb = oracle_superblock(sys_blk, env_blk, l + 1, l + 2, b, Ne, S2)
    psb = ham.Superblock(block_to_product(sys_blk), block_to_product(env_blk), l + 1, l + 2, b, Ne, S2)
    return osb, psb, sys_blk, env_blk, b, nb


@pytest.mark.parametrize("prec,tol", [(capi.PREC_F64, 1e-12), (capi.PREC_F32, 1e-5), (capi.PREC_BF16X3, 3e-5)])
@pytest.mark.parametrize("case", [(2, 4, 6, 0, 3), (3, 4, 10, 0, 4), (2, 4, 7, 1, 1)])
def test_model_wave_device_matches_oracle(ctx, case, prec, tol):
    osb, psb, *_ = _setup(*case)
    plan = psb.plan_builder().build(ctx, psb.super.dim, psb.super.dim)
    rng = np.random.default_rng(5)
    v = rng.normal(size=osb.n) + 1j * rng.normal(size=osb.n)
    got = plan.apply_host(v, prec)
    want = osb.matvec(v)
    assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)


def test_lanczos_on_superblock_equals_ed(ctx):
    Lx, Ny, Ne, S2, l = 3, 3, 8, 0, 3
    osb, psb, *_ = _setup(Lx, Ny, Ne, S2, l)
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b, twoS=S2)
    plan = psb.plan_builder().build(ctx, psb.super.dim, psb.super.dim)
    psi = ctx.vec_from(start_vector(psb.super, 7))
    e, _ = plan.lanczos(psi, tol=1e-12)
    n = Lx * Ny
    assert abs(e - e_ed) / n < 1e-10
    psi32 = ctx.vec_from(start_vector(psb.super, 7))
    e32, _ = plan.lanczos(psi32, tol=1e-12, prec=capi.PREC_BF16X3, refine_iters=40)
    assert abs(e32 - e_ed) / n < 1e-10
    # the default refinement pass (two steps: one FP64 product, one BF16x3)
    psi2 = ctx.vec_from(start_vector(psb.super, 7))
    e2, _ = plan.lanczos(psi2, tol=1e-12, prec=capi.PREC_BF16X3, refine_iters=2)
    assert abs(e2 - e_ed) / n < 1e-10


def test_truncating_step_matches_oracle(ctx):
    """Same blocks and start vector: energy, kept split and the spectrum of
    the renormalised block Hamiltonian agree with the oracle's step."""
    Lx, Ny, Ne, l = 3, 4, 10, 4
    osb, psb, sys_blk, env_blk, b, nb = _setup(Lx, Ny, Ne, 0, l)
    psi0 = start_vector(psb.super, 11)
    plan = psb.plan_builder().build(ctx, psb.super.dim, psb.super.dim)
    v = ctx.vec_from(psi0)
    e, _ = plan.lanczos(v, tol=1e-12)
    e_or, psi_or, _ = olanczos.lanczos(osb.matvec, psi0, tol=1e-12)
    assert abs(e - e_or) < 1e-10
    kmin, kmax, terr = 12, 30, 1e-7
    sd = [dx for (_, _, dx, _) in psb.super.sub]
    ed = [dy for (_, _, _, dy) in psb.super.sub]
    xs = [x for (x, _, _, _) in psb.super.sub]
    twos = [psb.sys_bs.sectors[x][1] for x in xs]
    got = capi.rho_truncate(ctx, v, sd, ed, twos, 0, kmin, kmax, terr)
    secs = [osb.sector(psi_or, x, y) for (x, y, _, _) in osb.sb]
    want = otrunc.truncation_operator(otrunc.density_matrix(secs, 0), twos, kmin, kmax, terr)
    assert list(got["kept"]) == list(want["kept"])
    assert abs(got["trun_err"] - want["trun_err"]) < 1e-10
    # renormalise the sys block with both U's; compare gauge-invariant spectra
    s1 = l + 1
    to_new = [(i, t, J) for (i, t, J) in nb[s1] if i in sys_blk.sites]
    keep = [i for i in sys_blk.sites + [s1] if any(j > s1 for (j, _, _) in nb.get(i, []))]
    U, order = {}, []
    for x, u, k in zip(xs, got["U"], got["kept"]):
        if k:
            U[x] = u
            order.append(x)
    newb = ham.renormalize(ctx, block_to_product(sys_blk), s1, to_new, keep, U, order)
    eb = su2_dense.enlarge(sys_blk, s1, to_new, keep)
    Uo = [None] * len(eb.sectors)
    for x, u in zip(xs, want["U"]):
        Uo[x] = u
    ob = su2_dense.truncate(eb, Uo)
    ospec = sorted(np.round(np.linalg.eigvalsh(m), 9).tolist() for m in ob.H.blocks.values())
    gspec = sorted(np.round(np.linalg.eigvalsh(m), 9).tolist() for (bk, kk), m in newb.ham.blocks.items() if bk == kk)
    assert ospec == gspec


@pytest.mark.parametrize("resident", [True, False])
@pytest.mark.parametrize("Lx,Ny,Ne", [(3, 3, 8), (2, 4, 6), (2, 6, 10)])
def test_dmrg_untruncated_equals_ed(ctx, Lx, Ny, Ne, resident):
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-10, resident=resident)
    e = d.run()
    assert abs(e - e_ed) / (Lx * Ny) < 1e-10, (e, e_ed)


def test_mixed_dmrg_six_leg_equals_ed(ctx):
    """BF16x3 H.psi with FP64 refinement on the six-leg cylinder (2 columns,
    every state kept): the energy is the exact one within 1e-10 per site."""
    Lx, Ny, Ne = 2, 6, 10
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-10,
             prec=capi.PREC_BF16X3, refine=2)
    e = d.run()
    assert abs(e - e_ed) / (Lx * Ny) < 1e-10, (e, e_ed)
    # every solve left its precision record: the mixed Ritz value is off by
    # the BF16x3 rounding, the refined energy sits at or below the FP64
    # <psi|H psi> of the mixed vector, and the refinement estimate is tiny
    assert len(d.precision_record) == len(d.history)
    for (_, em, ea, res, er, wr) in d.precision_record:
        assert abs(em - ea) < 1e-3 and res >= 0.0 and 0.0 <= wr <= 1.0
        assert er <= ea + 1e-12
    for o in d.precision_by_phase().values():
        assert o["max_refined_dE_est_per_site"] < 1e-11


def test_resident_dmrg_matches_host_operators(ctx):
    """Operators kept on the device (merged and renormalised there) give the
    energies of the host-operator driver, step by step, under truncation."""
    runs = []
    for resident in (True, False):
        d = DMRG(ctx, 3, 4, Ne=10, kept_min=20, kept_max=40, sweeps=2, lanzero=1e-10, trun_err=1e-12,
                 resident=resident)
        runs.append((d.run(), [h[2] for h in d.history]))
    (e_dev, steps_dev), (e_host, steps_host) = runs
    assert len(steps_dev) == len(steps_host) > 0
    np.testing.assert_allclose(steps_dev, steps_host, rtol=0, atol=1e-9)
    assert abs(e_dev - e_host) < 1e-9, (e_dev, e_host)


def test_dmrg_truncated_is_variational_and_converges_in_M(ctx):
    Lx, Ny, Ne = 3, 4, 10
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    es = []
    for kmax in (30, 60, 200):
        d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=min(20, kmax), kept_max=kmax, sweeps=3, lanzero=1e-10,
                 trun_err=1e-12)
        es.append(d.run())
    assert all(e >= e_ed - 1e-9 for e in es), (es, e_ed)
    assert es[0] >= es[1] - 1e-8 >= es[2] - 2e-8, es
    assert (es[1] - e_ed) / (Lx * Ny) < 1e-3
    assert (es[2] - e_ed) / (Lx * Ny) < 1e-7, (es[2], e_ed)


def test_wave_guess_on_device(ctx, monkeypatch):
    """The transformed previous ground state (wave_trans_LR/RL) is the new
    superblock's ground state when nothing is truncated, and the guessed
    solves need only a few H.psi."""
    starts, finals, mvs = [], [], []
    orig_initial = DMRG._initial

    def initial(self, sb, guess):
        v = orig_initial(self, sb, guess)
        starts.append((guess is not None, v.download()))
        return v

    monkeypatch.setattr(DMRG, "_initial", initial)
    orig_step = DMRG.step

    def step(self, *a, **k):
        r = orig_step(self, *a, **k)
        finals.append(r["psi"].download())
        mvs.append(r["matvecs"])
        return r

    monkeypatch.setattr(DMRG, "step", step)
    d = DMRG(ctx, 3, 3, Ne=8, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-11)
    E = d.run()
    ovs = [abs(np.vdot(f, v)) / (np.linalg.norm(f) * np.linalg.norm(v)) for (g, v), f in zip(starts, finals) if g]
    assert d.guessed == len(ovs) > 8
    assert min(ovs) > 1 - 1e-9, ovs
    guessed_mv = [m for (g, _), m in zip(starts, mvs) if g]
    assert max(guessed_mv) <= 6, guessed_mv
    e_ed = tj_ed.ground_energy(9, 8, tj_model.bonds(3, 3, 1.0, 0.2, 0.4, 0.016))
    assert abs(E - e_ed) / 9 < 1e-10
