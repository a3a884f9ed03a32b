"""Ground-state measurement (Measurement, DMRG_code/main.f90:213-354;
Get_density_cor and Get_mSCOP_Operator_Cor0, DMRG_code/measure.f90:
2525-2834) on the fake device (numpy plans).

Without truncation the measured state is the exact ground state, so every
density, spin, density-density, hopping and singlet-pair correlator must
equal brute-force values from exact diagonalisation in the Fock basis
(oracle/tj_ed.py, whose correlators are pinned in test_oracle_physics.py).
This covers every place the two sites can sit in the final superblock
[A, s1 | B, s2]: both in A, A-s1, A-s2, A-B, s1-s2, s1-B, and pairs inside
A and across A-B. With truncation, sum rules that hold for any state of the
target sector still hold exactly."""
import numpy as np
import pytest

import fake_device
from oracle import tj_ed, tj_model
from tsc_tri_t1t2j1j2_amd import measure as meas
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG

ED_TERMS = {"SS": tj_ed.spin_terms, "nn": tj_ed.density_density_terms, "cc": tj_ed.hopping_terms}


def _dmrg(ctx, Lx, Ny, Ne, kmax=4000, lanzero=1e-15, kmin=20):
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=kmin, kept_max=kmax, sweeps=1, lanzero=lanzero, trun_err=1e-14,
             native_planner=False)
    d.run()
    return d


def _ed(Lx, Ny, Ne):
    N = Lx * Ny
    E, v, st, gap = tj_ed.ground_state(N, Ne, tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016))
    assert gap > 1e-3      # a unique ground state fixes the correlators
    return N, v, st


@pytest.mark.parametrize("Lx,Ny,Ne,l_end,i0", [
    (3, 3, 8, 3, 2),       # i0 in A; j in A, s1, s2 and a 4-site B
    (3, 3, 8, 3, 4),       # i0 = s1; j = s2 and in B
    (3, 3, 8, None, 1),    # final half-sweep to 5N/6 - 1 as the reference
    (2, 4, 6, 3, 1),
])
def test_correlators_equal_exact_diagonalisation(monkeypatch, Lx, Ny, Ne, l_end, i0):
    ctx = fake_device.install(monkeypatch)
    d = _dmrg(ctx, Lx, Ny, Ne)
    N, v, st = _ed(Lx, Ny, Ne)
    out = d.measure(i0=i0, js=range(i0 + 1, N + 1), l_end=l_end, pair_bonds=())
    assert sorted(out["density"]) == list(range(1, N + 1))
    for j, val in out["density"].items():
        assert abs(val - tj_ed.expect(v, st, N, tj_ed.density_terms(j)).real) < 1e-8, j
    for kind, fn in ED_TERMS.items():
        assert sorted(out[kind]) == [(i0, j) for j in range(i0 + 1, N + 1)]
        for (a, b), val in out[kind].items():
            ref = tj_ed.expect(v, st, N, fn(a, b)).real
            assert abs(val - ref) < 1e-8, (kind, a, b, val, ref)


def test_pair_correlators_equal_exact_diagonalisation(monkeypatch):
    ctx = fake_device.install(monkeypatch)
    Lx, Ny, Ne, l_end = 2, 4, 6, 4
    d = _dmrg(ctx, Lx, Ny, Ne)
    N, v, st = _ed(Lx, Ny, Ne)
    tol, kept_max, terr = d._final
    psi, bases = d.left_to_right(1, l_end, tol, kept_max, terr, "final", keep=True)
    seen = 0
    for bond, partners in (((1, 2), [(3, 4), (7, 8)]), ((1, 3), [(7, 8)]), ((2, 1), [(8, 7), (4, 3)]),
                           ((1, 4), [(7, 8)])):
        got = meas.measure_pairs(ctx, psi, bases, l_end + 1, N, d.sys_trun, d.env_trun, bond, partners)
        for (c, e), val in got.items():
            ref = tj_ed.expect(v, st, N, tj_ed.pair_pair_terms(bond[0], bond[1], c, e))
            assert abs(ref) > 1e-4       # the test compares non-zero values
            assert abs(val - ref.real) < 1e-8, (bond, c, e, val, ref)
            seen += 1
    assert seen == 6
    with pytest.raises(ValueError):
        meas.measure_pairs(ctx, psi, bases, l_end + 1, N, d.sys_trun, d.env_trun, (1, 2), [(2, 3)])


def test_sum_rules_hold_with_truncation(monkeypatch):
    """Truncation keeps N and S exact: sum_j <n_j> = N_e, sum_{j>1} <S_1.S_j>
    = -3/4 <n_1> (singlet) and sum_{j>1} <n_1 n_j> = (N_e - 1) <n_1>."""
    ctx = fake_device.install(monkeypatch)
    Lx, Ny, Ne = 4, 2, 6
    d = _dmrg(ctx, Lx, Ny, Ne, kmax=6, lanzero=1e-10, kmin=2)
    assert max(h[3] for h in d.history if h[0] == "sweep1") > 1e-4     # truncation is active
    N = Lx * Ny
    out = d.measure(i0=1, js=range(2, N + 1), pair_bonds=())
    n = out["density"]
    assert abs(sum(n.values()) - Ne) < 1e-10
    assert abs(sum(out["SS"].values()) + 0.75 * n[1]) < 1e-10
    assert abs(sum(out["nn"].values()) - (Ne - 1) * n[1]) < 1e-10
    assert -0.75 * n[1] <= out["SS"][(1, 2)] <= 0.0


def test_reference_site_choice():
    """xi1 = Nx/4 Ny + 2 - Ny and the same-leg sites up to N - N/6 + Ny
    (main.f90:255; measure.f90:2582-2610), the bond partners of
    Get_mSCOP_Operator_Cor0 (measure.f90:2750-2830)."""
    i0, js = meas.reference_sites(24, 4)
    assert i0 == 22 and js == list(range(26, 83, 4))
    bond, partners = meas.reference_pair_partners(24, 4, 80, 1)
    assert bond == (22, 23)
    assert partners[0] == (26, 1, (26, 27)) and partners[-1] == (74, 5, (74, 77))
    assert len(partners) == 13 * 3
    assert meas.reference_pair_partners(24, 4, 80, 3)[0] == (22, 26)
    assert meas.reference_pair_partners(24, 4, 80, 5)[0] == (22, 25)


def test_default_measurement_and_files(monkeypatch, tmp_path):
    ctx = fake_device.install(monkeypatch)
    d = _dmrg(ctx, 4, 2, 6, lanzero=1e-10)
    out = d.measure()
    assert out["l_end"] == 5 * 8 // 6 - 1
    assert out["i0"] == 2 and out["js"] == [4, 6, 8]
    assert set(out["pairs"]) == {1, 3, 5}
    assert abs(sum(out["density"].values()) - 6) < 1e-9
    paths = meas.write_reference_files(out, str(tmp_path), 4, 2, d.J2, d.kept_min, d.kept_maxf, d.Ne)
    names = sorted(p.rsplit("/", 1)[1] for p in paths)
    assert names == ["Sc_cor_x.dat", "Sc_cor_xy.dat", "Sc_cor_y.dat", "density_cor.dat", "elec_cor.dat",
                     "elec_density.dat", "spin_cor.dat"]
    # elec_density: (2i6, 3f14.8) = ix, iy, <n>, running sum
    rows = (tmp_path / "elec_density.dat").read_text().splitlines()[1:9]
    assert [(int(r[:6]), int(r[6:12])) for r in rows] == [(ix, iy) for ix in range(1, 5) for iy in (1, 2)]
    assert abs(float(rows[-1][26:40]) - 6.0) < 1e-7
    # spin_cor: (3i8, 3f21.14) = i, j, column distance, Re, Im
    rows = (tmp_path / "spin_cor.dat").read_text().splitlines()[1:]
    assert [(int(r[:8]), int(r[8:16]), int(r[16:24])) for r in rows] == [(2, 4, 1), (2, 6, 2), (2, 8, 3)]
    assert abs(float(rows[0][24:45]) - out["SS"][(2, 4)]) < 1e-13


def test_reference_pair_partners_in_the_env_block():
    """Partners of Get_mSCOP_Operator_Cor0's env loop (measure.f90:2804-2827)
    on an 8 x 2 cylinder measured at s1 = 7: xi1 = 4, bond (4, 6) of type 3;
    the leg sites 10, 12, 14 <= xj1 = 14 lie in the env block (8 sites), the
    type-1 neighbour does not exist on the last leg of a 2-leg cylinder."""
    bond, partners = meas.reference_pair_partners(8, 2, 7, 3)
    assert bond == (4, 6)
    assert partners == [(10, 3, (10, 12)), (10, 5, (10, 11)), (12, 3, (12, 14)), (12, 5, (12, 13)),
                        (14, 3, (14, 16)), (14, 5, (14, 15))]
