"""Ground-state measurement on the device (Measurement, DMRG_code/main.f90:
213-354; Get_density_cor / Get_mSCOP_Operator_Cor0, DMRG_code/measure.f90:
2525-2834): operator transport and expectation plans through the C-ABI.

Untruncated runs must reproduce the exact-diagonalisation correlators of the
oracle; a truncated run on a 4x4 cylinder must keep the sum rules that hold
for any state of the singlet N_e sector."""
import pytest

from oracle import tj_ed, tj_model
from tsc_tri_t1t2j1j2_amd import measure as meas
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG

pytestmark = pytest.mark.gpu

ED_TERMS = {"SS": tj_ed.spin_terms, "nn": tj_ed.density_density_terms, "cc": tj_ed.hopping_terms}


def _ed(Lx, Ny, Ne):
    N = Lx * Ny
    E, v, st, gap = tj_ed.ground_state(N, Ne, tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016))
    assert gap > 1e-3
    return v, st


@pytest.mark.parametrize("l_end,i0", [(3, 2), (3, 4), (None, 1)])
def test_device_correlators_equal_ed(ctx, l_end, i0):
    Lx, Ny, Ne = 3, 3, 8
    N = Lx * Ny
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-13, trun_err=1e-14)
    d.run()
    out = d.measure(i0=i0, js=range(i0 + 1, N + 1), l_end=l_end, pair_bonds=())
    v, st = _ed(Lx, Ny, Ne)
    for j, val in out["density"].items():
        assert abs(val - tj_ed.expect(v, st, N, tj_ed.density_terms(j)).real) < 1e-7, j
    for kind, fn in ED_TERMS.items():
        assert len(out[kind]) == N - i0
        for (a, b), val in out[kind].items():
            ref = tj_ed.expect(v, st, N, fn(a, b)).real
            assert abs(val - ref) < 1e-7, (kind, a, b, val, ref)
    assert out["stats"]["transport_device_ms"] > 0.0


def test_device_pair_correlators_equal_ed(ctx):
    Lx, Ny, Ne, l_end = 2, 4, 6, 4
    N = Lx * Ny
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=4000, kept_max=4000, sweeps=1, lanzero=1e-13, trun_err=1e-14)
    d.run()
    tol, kept_max, terr = d._final
    psi, bases = d.left_to_right(1, l_end, tol, kept_max, terr, "final", keep=True)
    v, st = _ed(Lx, Ny, Ne)
    try:
        got = meas.measure_pairs(ctx, psi, bases, l_end + 1, N, d.sys_trun, d.env_trun, (1, 2), [(3, 4), (7, 8)])
        got.update(meas.measure_pairs(ctx, psi, bases, l_end + 1, N, d.sys_trun, d.env_trun, (1, 4), [(7, 8)]))
    finally:
        psi.close()
    refs = {(3, 4): (1, 2), (7, 8): (1, 4)}
    for (c, e), val in got.items():
        a, b = refs[(c, e)]
        ref = tj_ed.expect(v, st, N, tj_ed.pair_pair_terms(a, b, c, e)).real
        assert abs(ref) > 1e-4 and abs(val - ref) < 1e-7, ((a, b), (c, e), val, ref)


def test_device_measurement_sum_rules_under_truncation(ctx):
    Lx, Ny, Ne = 4, 4, 14
    N = Lx * Ny
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=20, kept_max=60, sweeps=2, lanzero=1e-10, trun_err=1e-12)
    d.run()
    out = d.measure(i0=1, js=range(2, N + 1), pair_bonds=(1, 3, 5))
    n = out["density"]
    assert abs(sum(n.values()) - Ne) < 1e-9
    assert abs(sum(out["SS"].values()) + 0.75 * n[1]) < 1e-9
    assert abs(sum(out["nn"].values()) - (Ne - 1) * n[1]) < 1e-9
    assert all(0.0 < x < 1.0 for x in n.values())
    assert set(out["pairs"]) == {1, 3, 5}


def test_device_checkpoint_and_measurement_replay(ctx, tmp_path):
    """The per-length files written during a device run reproduce its
    measurement when replayed (restart1 = 4) and restart it (restart1 = 2)."""
    d = DMRG(ctx, 4, 3, Ne=10, kept_min=20, kept_max=40, sweeps=1, lanzero=1e-10, trun_err=1e-12)
    e = d.run(checkpoint=str(tmp_path))
    out = d.measure()
    rep = meas.measure_from_disk(ctx, str(tmp_path), 4, 3)
    for kind in ("SS", "nn", "cc"):
        for key, v in out[kind].items():
            assert abs(rep[kind][key] - v) < 1e-12, (kind, key)
    for j, v in out["density"].items():
        assert abs(rep["density"][j] - v) < 1e-12
    again = DMRG(ctx, 4, 3, Ne=10, kept_min=20, kept_max=40, sweeps=1, lanzero=1e-10, trun_err=1e-12)
    e2 = again.run(checkpoint=str(tmp_path), restart=(2, 0, 0))
    assert abs(e2 - e) < 1e-6, (e2, e)


def test_device_block_files_are_the_references(ctx, tmp_path):
    """The resident device path (enlargement and U^H O U on the GPU) writes
    block, truncation and wave-function files whose every matrix equals the
    reduced matrix of the physical operator between the reference's block
    states (oracle/ref_blocks.py): a 4 x 3 warm-up with truncation, both
    sides, and the ground state of a 2 x 3 warm-up."""
    import ref_file_checks as C
    pytest.importorskip("sympy")
    cpl = (1.0, 0.5, 0.4, 0.2)
    d = DMRG(ctx, 4, 3, t1=cpl[0], t2=cpl[1], J1=cpl[2], J2=cpl[3], Ne=10, kept_min=4, kept_max=7, sweeps=1,
             lanzero=1e-10, trun_err=1e-14)
    d.checkpoint = str(tmp_path / "a")
    import os
    os.makedirs(d.checkpoint)
    d.warm_up()
    assert C.check_block_files(d.checkpoint, "sys", 4, 3, cpl, d.N // 2) == 0
    assert C.check_block_files(d.checkpoint, "env", 4, 3, cpl, d.N // 2) > 0
    w = DMRG(ctx, 2, 3, t1=cpl[0], t2=cpl[1], J1=cpl[2], J2=cpl[3], Ne=4, kept_min=4, kept_max=40, sweeps=1,
             lanzero=1e-12, trun_err=1e-14)
    w.checkpoint = str(tmp_path / "b")
    os.makedirs(w.checkpoint)
    w.warm_up()
    nrm, Eh, res, res_bad = C.wave_check(w.checkpoint, 6, 2, 3, cpl, w.history[-1][2])
    assert abs(nrm - 1.0) < 1e-10 and abs(Eh - w.history[-1][2]) < 1e-9
    assert res < 1e-6 and res_bad > 1e-3
