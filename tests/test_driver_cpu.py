"""The DMRG driver's step logic on the CPU (device objects replaced by the
oracle through tests/fake_device.py), against exact diagonalisation."""
import numpy as np
import pytest

from oracle import tj_ed, tj_model
from tests import fake_device
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG


@pytest.mark.parametrize("kmax,resident", [(4000, False), (40, False), (4000, True), (40, True)])
def test_driver_on_fake_device(monkeypatch, kmax, resident):
    ctx = fake_device.install(monkeypatch)
    Lx, Ny, Ne = 3, 3, 8
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    log = []
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=min(20, kmax), kept_max=kmax, sweeps=2, lanzero=1e-10, trun_err=1e-12,
             log=log.append, native_planner=False, resident=resident)
    e = d.run()
    print("\n".join(log))
    assert e >= e_ed - 1e-9
    assert (e - e_ed) / (Lx * Ny) < (1e-9 if kmax > 1000 else 1e-4), (e, e_ed)


def test_driver_yc6_untruncated_equals_ed(monkeypatch):
    """Six-leg cylinder (the headline lattice's width, 2 columns, 12 sites,
    y-periodic bonds of types 1, 3, 5 and the next-nearest ones): with every
    state kept, warm-up and sweep reproduce the exact ground energy."""
    ctx = fake_device.install(monkeypatch)
    Lx, Ny, Ne = 2, 6, 10
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-10, trun_err=1e-12,
             native_planner=False, resident=True)
    e = d.run()
    assert abs(e - e_ed) / (Lx * Ny) < 1e-10, (e, e_ed)


def test_resident_driver_matches_host_driver(monkeypatch):
    """Operators kept in (fake) device arenas, merged and renormalised there,
    give the same DMRG trajectory as host-side blocks."""
    ctx = fake_device.install(monkeypatch)
    runs = []
    for resident in (False, True):
        d = DMRG(ctx, 3, 3, Ne=8, kept_min=20, kept_max=40, sweeps=1, lanzero=1e-11, trun_err=1e-12,
                 native_planner=False, resident=resident)
        d.run()
        runs.append([h[2] for h in d.history])
    assert len(runs[0]) == len(runs[1])
    assert np.allclose(runs[0], runs[1], atol=1e-9, rtol=0)


def test_shadow_error_by_phase(monkeypatch):
    """With FP64 shadow solves the per-phase maximum of |E_mixed - E_FP64| /
    site is reported for the warm-up, both half-sweeps and every sweep."""
    import fake_device
    from tsc_tri_t1t2j1j2_amd import capi
    from tsc_tri_t1t2j1j2_amd.dmrg import DMRG
    ctx = fake_device.install(monkeypatch)
    d = DMRG(ctx, 3, 3, Ne=8, kept_min=4, kept_max=12, sweeps=2, lanzero=1e-10, trun_err=1e-14,
             native_planner=False, prec=capi.PREC_BF16X3, shadow_f64=True)
    d.run()
    ph = d.shadow_by_phase()
    assert set(ph) == {"warmup", "half", "sweep1", "sweep2"}
    assert len(d.shadow) == len(d.history)
    # the fake device solves every precision in FP64: both solves agree
    assert max(ph.values()) < 1e-12


def test_mixed_precision_record_by_phase(monkeypatch):
    """Every mixed solve leaves a refinement record; the per-phase summary
    covers the warm-up, both half-sweeps and every sweep, and compares the
    mixed Ritz-energy error with the phase's truncation error. An FP64 run
    records nothing."""
    from tsc_tri_t1t2j1j2_amd import capi
    ctx = fake_device.install(monkeypatch)
    d = DMRG(ctx, 3, 3, Ne=8, kept_min=4, kept_max=12, sweeps=2, lanzero=1e-10, trun_err=1e-14,
             native_planner=False, prec=capi.PREC_BF16X3)
    d.run()
    assert len(d.precision_record) == len(d.history)
    ph = d.precision_by_phase()
    assert set(ph) == {"warmup", "half", "sweep1", "sweep2"}
    for tag, o in ph.items():
        # the fake device solves in FP64: <psi|H psi> of its vector is its energy
        assert o["max_mixed_ritz_dE_per_site"] < 1e-12 and abs(o["max_refine_shift_per_site"]) < 1e-12
        assert o["max_refined_dE_est_per_site"] == 0.0
        # (an energy-change stop leaves a residual of about the square root of the tolerance)
        assert 0.0 <= o["max_resid_f64"] < 1e-2
        assert o["max_trun_err"] == max(h[3] for h in d.history if h[0] == tag)
        assert o["dE_below_trun_err"] == (o["max_refined_dE_est_per_site"] <= o["max_trun_err"])
    assert sum(o["solves"] for o in ph.values()) == len(d.history)
    d64 = DMRG(ctx, 3, 3, Ne=8, kept_min=4, kept_max=12, sweeps=1, lanzero=1e-10, trun_err=1e-14,
               native_planner=False, prec=capi.PREC_F64)
    d64.run()
    assert d64.precision_record == [] and d64.precision_by_phase() == {}


def test_precision_summary_arithmetic():
    """precision_by_phase on planted records: per-phase maxima of
    |E_mixed - E_rayleigh| / site and of the residual, the phase's largest
    truncation error, and the comparison of the two."""
    d = DMRG.__new__(DMRG)
    d.N = 10
    d.history = [("warmup", 2, -1.0, 1e-9, 4, 10), ("warmup", 3, -2.0, 3e-9, 4, 10), ("half", 4, -3.0, 2e-12, 4, 10)]
    d.precision_record = [(0, -1.0 + 2e-8, -1.0, 1e-5, -1.0 - 4e-9, 1e-2),
                          (1, -2.0 - 1e-8, -2.0, 3e-5, -2.0 - 1e-9, 0.5),
                          (2, -3.0 + 5e-10, -3.0, 2e-6, -3.0, 1e-1)]
    ph = d.precision_by_phase()
    assert ph["warmup"]["solves"] == 2 and ph["half"]["solves"] == 1
    assert abs(ph["warmup"]["max_mixed_ritz_dE_per_site"] - 2e-9) < 1e-15
    # w_rest * raw: 1e-2 * 2e-9 and 0.5 * 1e-9
    assert abs(ph["warmup"]["max_refined_dE_est_per_site"] - 5e-10) < 1e-16
    assert abs(ph["warmup"]["max_refine_shift_per_site"] - 4e-10) < 1e-15
    assert ph["warmup"]["max_resid_f64"] == 3e-5
    assert ph["warmup"]["max_trun_err"] == 3e-9 and ph["warmup"]["dE_below_trun_err"]
    assert abs(ph["half"]["max_mixed_ritz_dE_per_site"] - 5e-11) < 1e-16
    assert abs(ph["half"]["max_refined_dE_est_per_site"] - 5e-12) < 1e-17
    assert ph["half"]["max_trun_err"] == 2e-12 and not ph["half"]["dE_below_trun_err"]


def test_krylov_budget_fits_free_memory():
    """The Krylov space of a solve is cut to what free device memory holds
    (less 4 GiB), with more restarts so that the product budget stays."""
    from types import SimpleNamespace
    d = DMRG.__new__(DMRG)
    d.lanczos_budget, d.budget_cuts = (100, 4), 0
    n = 1_000_000
    d.ctx = SimpleNamespace(mem_info=lambda: ((4 << 30) + 20 * 16 * n, 1 << 40))
    assert d._budget(n) == (19, 26)   # ceil(500 / 19) - 1
    d.ctx = SimpleNamespace(mem_info=lambda: (1 << 40, 1 << 40))
    assert d._budget(n) == (100, 4)
    d.ctx = SimpleNamespace(mem_info=lambda: (1 << 20, 1 << 40))
    assert d._budget(n) == (8, 62)
    assert d.budget_cuts == 2
    d.ctx = SimpleNamespace()   # no memory query: the configured budget
    assert d._budget(n) == (100, 4)


def test_driver_with_a_cut_krylov_space_equals_ed(monkeypatch):
    """With free memory for only a few Krylov vectors the solves restart
    and the untruncated run still reaches the exact energy."""
    ctx = fake_device.install(monkeypatch)
    Lx, Ny, Ne = 3, 3, 8
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    cur = {}
    ctx.mem_info = lambda: ((4 << 30) + 13 * 16 * cur.get("n", 1), 1 << 40)
    real_budget = DMRG._budget

    def spy(self, n):
        cur["n"] = n
        out = real_budget(self, n)
        assert out[0] == 12
        return out
    monkeypatch.setattr(DMRG, "_budget", spy)
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-10, trun_err=1e-12,
             native_planner=False)
    e = d.run()
    assert d.budget_cuts > 0
    assert abs(e - e_ed) / (Lx * Ny) < 1e-10, (e, e_ed)


def test_renorm_preparation_failure_reaches_the_caller(monkeypatch):
    """The renormalisation bookkeeping built beside the solve: an error
    raised there ends the step with that error, after the solve."""
    from tsc_tri_t1t2j1j2_amd import hamiltonian as ham
    ctx = fake_device.install(monkeypatch)

    def broken(*args, **kwargs):
        raise RuntimeError("renorm bookkeeping failed")

    monkeypatch.setattr(ham, "renorm_transitions", broken)
    d = DMRG(ctx, 3, 3, Ne=8, kept_min=4, kept_max=12, sweeps=1, lanzero=1e-10, native_planner=False)
    with pytest.raises(RuntimeError, match="renorm bookkeeping failed"):
        d.run()
    # the failed step freed its plan, step arena and state vectors: only the
    # arenas of the two site blocks it was given remain
    site_arenas = {id(m.arena.vec) for (_, m) in d._dev_cache.values()}
    assert ctx.live == site_arenas


def test_failed_truncation_frees_the_step(monkeypatch):
    """A failure after the solve (in the truncation) also frees the plan,
    the step arena, the ground state and the FP64 shadow's start vector."""
    from tsc_tri_t1t2j1j2_amd import capi
    ctx = fake_device.install(monkeypatch)

    def broken(*args, **kwargs):
        raise RuntimeError("eigensolver failed")

    monkeypatch.setattr(capi, "rho_truncate", broken)
    d = DMRG(ctx, 3, 3, Ne=8, kept_min=1, kept_max=2, sweeps=1, lanzero=1e-10, native_planner=False,
             prec=capi.PREC_BF16X3, shadow_f64=True)
    with pytest.raises(RuntimeError, match="eigensolver failed"):
        d.run()
    site_arenas = {id(m.arena.vec) for (_, m) in d._dev_cache.values()}
    assert ctx.live == site_arenas


def test_centre_window_repeats_from_the_warm_up_blocks(monkeypatch):
    """centre_window() runs the first half-sweep's centre steps from the
    warm-up blocks; a second call from the same blocks (another precision on
    a device) follows the same steps, and a full run's first half-sweep
    steps give the same energies. The warm-up with `retain` keeps only the
    lengths the window needs."""
    ctx = fake_device.install(monkeypatch)
    kw = dict(Ne=10, kept_min=10, kept_max=30, sweeps=0, lanzero=1e-11, trun_err=1e-12, native_planner=False)
    d = DMRG(ctx, 4, 3, retain=3, **kw)
    d.warm_up()
    l0 = d.N // 2 - 1
    assert sorted(L for L in d.sys_blocks if L > 1) == [l0 - 1, l0, l0 + 1]
    w1 = d.centre_window(2, 1e-10, 30, 1e-12)
    w2 = d.centre_window(2, 1e-10, 30, 1e-12)
    assert [r["l"] for r in w1] == [l0, l0 + 1]
    for a, b in zip(w1, w2):
        assert abs(a["E"] - b["E"]) < 1e-8 and a["dim"] == b["dim"] and a["kept"] == b["kept"]
    full = DMRG(ctx, 4, 3, **kw)
    full.warm_up()
    full.left_to_right(l0, l0 + 1, 1e-10, 30, 1e-12, "half")
    for a, h in zip(w1, full.history[-2:]):
        assert abs(a["E"] - h[2]) < 1e-8 and a["dim"] == h[5]
