"""The quantities in the reference's measurement files, fixed from its source.

The reference's coupling rule (block_mul_block_ndia, DMRG_code/block.f90:
376-635), its Wigner-Eckart convention, its site reduced elements
(dmrg_sub.f90:733-860) and the factors Get_density_cor applies
(measure.f90:2586-2587, 2662-2663) are restated in oracle/ref_measure.py and
evaluated as explicit matrices on small Fock spaces. These tests show that
every column the reference writes is a physical correlator: spin_cor.dat
holds <S_i.S_j>, density_cor.dat <n_i n_j>, elec_cor.dat <sum_s c+_is c_js>,
Sc_cor_*.dat |<D+_ab D_cd>| and its phase, for the in-block pairs
(Get_sys_two / Get_sys_two1, measure.f90:2837-2935) and for the pairs across
the two blocks (sys_block_env_block_cor_ndia followed by dconjg,
measure.f90:2804-2812, 4398-4664). measure.py writes exactly these physical
values, pinned against exact diagonalisation in test_measure_cpu.py and
test_gpu_measure.py.

Each check is built so that a plausible slip fails it: the standard
Clebsch-Gordan phase instead of the reference's flips the pair operator's
sign, dropping the sqrt(1 + down_dif) factor or the sign for down_dif = 2
breaks the on-site identities, and a transposed site element breaks the
hopping identity.
"""
from math import sqrt

import numpy as np
import pytest

pytest.importorskip("sympy")
from oracle import ref_measure as R  # noqa: E402


def _ratio(X, Y):
    """s with X = s Y (asserting proportionality)."""
    x, y = X.ravel(), Y.ravel()
    s = np.vdot(y, x) / np.vdot(y, y)
    assert np.abs(x - s * y).max() < 1e-12
    return float(np.real(s))


def _file_factor(rank2):
    """Get_density_cor's factor: sqrt(1 + down_dif), negated for down_dif = 2."""
    f = sqrt(1.0 + rank2)
    return -f if rank2 == 2 else f


def test_site_operators_are_the_standard_tensors():
    # c+ has components c+_up (q = +1/2) and c+_dn; st_elec_down is c~ with
    # c~_{+1/2} = -c_dn and c~_{-1/2} = c_up; S_{+1} = -S+/sqrt2
    _, cp = R.site_components("c+")
    assert cp[1][1, 0] == pytest.approx(1.0) and cp[-1][2, 0] == pytest.approx(1.0)
    _, ct = R.site_components("c~")
    assert ct[-1][0, 1] == pytest.approx(1.0) and ct[1][0, 2] == pytest.approx(-1.0)
    _, S = R.site_components("S")
    assert S[0][1, 1] == pytest.approx(0.5) and S[0][2, 2] == pytest.approx(-0.5)
    assert S[2][1, 2] == pytest.approx(-1.0 / sqrt(2.0))


def test_on_site_identities_fix_the_file_factors():
    F = R.Fock(1)
    n = R.physical(F, "n", 0)
    S, cp, ct, nn = (F.site_op(0, x) for x in ("S", "c+", "c~", "n"))
    # S.S = 3/4 n, n n = n, sum_s c+_s c_s = n on one site
    assert _ratio(_file_factor(2) * R.scalar(R.couple(S, S, 0)), n) == pytest.approx(0.75)
    assert _ratio(_file_factor(0) * R.scalar(R.couple(nn, nn, 0)), n) == pytest.approx(1.0)
    assert _ratio(_file_factor(1) * R.scalar(R.couple(cp, ct, 0)), n) == pytest.approx(1.0)


@pytest.mark.parametrize("name,ops,rank2", [("SS", ("S", "S"), 2), ("nn", ("n", "n"), 0), ("hop", ("c+", "c~"), 1)])
def test_two_site_correlators_are_physical(name, ops, rank2):
    F = R.Fock(3)
    for i, j in ((0, 1), (0, 2), (1, 2)):
        X = R.couple(F.site_op(i, ops[0]), F.site_op(j, ops[1]), 0)
        assert _ratio(_file_factor(rank2) * R.scalar(X), R.physical(F, name, i, j)) == pytest.approx(1.0)


def test_transposed_site_element_breaks_the_hopping_identity():
    # using c+ where the reference uses its annihilation partner gives a pair
    # creator, not a hopping operator: the identity is sensitive to it
    F = R.Fock(2)
    X = R.scalar(R.couple(F.site_op(0, "c+"), F.site_op(1, "c+"), 0))
    Y = R.physical(F, "hop", 0, 1)
    assert abs(np.vdot(Y.ravel(), X.ravel())) < 1e-12


def test_pair_operators_of_the_reference():
    F = R.Fock(4)
    cp = [F.site_op(i, "c+") for i in range(4)]
    ct = [F.site_op(i, "c~") for i in range(4)]
    Dab = R.physical(F, "Ddag", 0, 1)
    Dcd = R.physical(F, "Ddag", 2, 3).T
    # Get_sys_two: [c+_a x c+_b]^0 = D+_ab
    P = R.couple(cp[0], cp[1], 0)
    assert _ratio(R.scalar(P), Dab) == pytest.approx(1.0)
    # Get_sys_two1: [[D+_ab x c~_c]^(1/2) x c~_d]^0 = + D+_ab D_cd with the
    # reference's coupling phase ...
    X = R.couple(R.couple(P, ct[2], 1), ct[3], 0)
    assert _ratio(R.scalar(X), Dab @ Dcd) == pytest.approx(1.0)
    # ... and - D+_ab D_cd with the standard Clebsch-Gordan phase
    Ps = R.couple(cp[0], cp[1], 0, R.std_cg)
    Xs = R.couple(R.couple(Ps, ct[2], 1, R.std_cg), ct[3], 0, R.std_cg)
    assert _ratio(R.scalar(Xs), Dab @ Dcd) == pytest.approx(-1.0)


def test_cross_block_scalar_coefficient():
    # sys_block_env_block_cor_ndia with two scalars: the coefficient is
    # 1/sqrt((2 j_A + 1)(2 j_B + 1)), i.e. the reduced elements of both
    # scalar block operators turned into their m-independent values; the
    # value <D+_ab psi | D+_cd psi> = <D_ab D+_cd> is then conjugated
    # (dconjg, measure.f90:2806), giving <D+_ab D_cd> for disjoint pairs
    seen = 0
    for J in range(0, 9):
        for ja in range(0, 9):
            for sa in (0, 1):
                for jb in range(0, 9):
                    for sb in (0, 1):
                        c = R.cross_scalar_coef(J, ja, sa, J, jb, sb)
                        if c == 0.0:
                            continue
                        seen += 1
                        assert c == pytest.approx(1.0 / sqrt((ja + 1.0) * (jb + 1.0)), rel=1e-12)
    assert seen > 20
    F = R.Fock(4)
    Dab = R.physical(F, "Ddag", 0, 1)
    Dcd = R.physical(F, "Ddag", 2, 3)
    rng = np.random.default_rng(5)
    psi = rng.normal(size=len(F.states)) + 1j * rng.normal(size=len(F.states))
    value = np.vdot(Dab @ psi, Dcd @ psi)
    assert np.conj(value) == pytest.approx(np.vdot(psi, Dab @ Dcd.T @ psi), abs=1e-12)
