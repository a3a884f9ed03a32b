"""Pins for the oracle itself: closed forms, library routines, invariants
and brute force, chosen so that a dropped term, a wrong sign or a swapped
operand in the oracle fails at least one of them."""
import itertools

import numpy as np
import pytest

from oracle import lanczos as olanczos
from oracle import synth as osynth
from oracle import truncation as otrunc
from oracle import uou as ouou
from oracle import wigner
from tsc_tri_t1t2j1j2_amd.synthetic import SyntheticSuperblock, gaussian_host


def test_w6j_against_sympy():
    sympy = pytest.importorskip("sympy")
    from sympy.physics.wigner import wigner_6j
    rng = np.random.default_rng(1)
    n = 0
    def triad(a, b, c):
        return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0

    for _ in range(4000):
        a = [int(x) for x in rng.integers(0, 9, 6)]
        j1, j2, j3, l1, l2, l3 = a
        if not (triad(j1, j2, j3) and triad(j1, l2, l3) and triad(l1, j2, l3) and triad(l1, l2, j3)):
            assert wigner.w6j(*a) == 0.0
            continue
        ref = wigner_6j(*[sympy.Rational(x, 2) for x in a])
        assert abs(wigner.w6j(*a) - float(ref)) < 1e-14
        n += ref != 0
    assert n > 30


def test_w6j_closed_form_and_symmetry():
    # {a b c; 0 c b} = (-1)^(a+b+c) / sqrt((2b+1)(2c+1))   (doubled args)
    for a, b, c in itertools.product(range(0, 9), repeat=3):
        if (a + b + c) % 2 or not (abs(b - c) <= a <= b + c):
            continue
        ref = (-1) ** ((a + b + c) // 2) / np.sqrt((b + 1) * (c + 1))
        assert abs(wigner.w6j(a, b, c, 0, c, b) - ref) < 1e-14
    rng = np.random.default_rng(2)
    for _ in range(200):
        j1, j2, j3, l1, l2, l3 = (int(x) for x in rng.integers(0, 10, 6))
        v = wigner.w6j(j1, j2, j3, l1, l2, l3)
        # column permutations and upper/lower swap of two columns
        assert abs(v - wigner.w6j(j2, j1, j3, l2, l1, l3)) < 1e-14
        assert abs(v - wigner.w6j(j1, l2, l3, l1, j2, j3)) < 1e-14
        # Regge symmetry
        if (j2 + j3 + l2 + l3) % 2 == 0:
            r = wigner.w6j(j1, (j2 + j3 + l2 - l3) // 2, (j2 + j3 - l2 + l3) // 2, l1,
                           (j2 - j3 + l2 + l3) // 2, (-j2 + j3 + l2 + l3) // 2)
            assert abs(v - r) < 1e-13


def test_gaussian_host_statistics():
    v = gaussian_host(200000, 5, 9, 2.0)
    assert abs(np.mean(np.abs(v) ** 2) - 4.0) < 0.05
    assert abs(np.mean(v)) < 0.02
    assert np.array_equal(v, gaussian_host(200000, 5, 9, 2.0))
    assert not np.allclose(v[:10], gaussian_host(10, 5, 10, 2.0))


def test_synth_oracle_matches_kron_form():
    sb = SyntheticSuperblock(nsec=3, nterms=4, dmin=2, dmax=5, seed=11)
    psi = sb.psi_host()
    out = osynth.hpsi(sb, psi)
    # dense operator sum_{t,s} c (O_R kron O_L) on vec(psi_s) (column-major)
    n = sb.psi_len
    H = np.zeros((n, n), np.complex128)
    for t in range(sb.T):
        for s in range(sb.S):
            g = sb.target[t, s]
            blk = sb.coef[t, s] * np.kron(sb.op_R(t, s), sb.op_L(t, s))
            r0, c0 = sb.psi_off[g], sb.psi_off[s]
            H[r0:r0 + blk.shape[0], c0:c0 + blk.shape[1]] += blk
    assert np.allclose(out, H @ psi, atol=1e-13)


def test_synth_flop_model_picks_cheaper_order():
    sb = SyntheticSuperblock(nsec=20, nterms=6, dmin=8, dmax=200, seed=3)
    for t in range(sb.T):
        for s in range(sb.S):
            lf, rf = sb.order_costs(t, s)
            assert lf > 0 and rf > 0
    f = sb.algorithmic_flops()
    upper = sum(8.0 * sb.order_costs(t, s)[0] for t in range(sb.T) for s in range(sb.S))
    assert 0 < f <= upper


def test_lanczos_oracle_exact_on_dense():
    rng = np.random.default_rng(4)
    n = 60
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = (a + a.conj().T) / 2
    e_ref = np.linalg.eigvalsh(H)[0]
    psi0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    e, psi, nmv = olanczos.lanczos(lambda v: H @ v, psi0, tol=1e-12, max_inner=100, max_outer=4)
    assert abs(e - e_ref) < 1e-9
    assert np.linalg.norm(H @ psi - e * psi) < 1e-5
    assert abs(np.vdot(psi, psi) - 1) < 1e-12


def test_quicksort_matches_stable_descending_on_distinct():
    rng = np.random.default_rng(5)
    v = list(rng.uniform(size=50))
    a = [0.0] + v
    idx = list(range(51))
    otrunc.quicksort_desc(a, idx, 1, 50)
    assert a[1:] == sorted(v, reverse=True)
    assert [v[i - 1] for i in idx[1:]] == a[1:]


def _brute_select(ev_sorted, kept_min, kept_max, trun_err):
    n = len(ev_sorted)
    if n <= kept_min:
        return n
    k = kept_min
    while True:
        err = sum(ev_sorted[k:])
        if err <= trun_err or k >= kept_max:
            return k
        k += 1


def test_truncation_selection_brute_force():
    rng = np.random.default_rng(6)
    for trial in range(200):
        nsec = int(rng.integers(1, 6))
        eig = [np.sort(rng.exponential(size=int(rng.integers(1, 6))) ** 3) for _ in range(nsec)]
        twos = rng.integers(0, 5, nsec)
        kmin = int(rng.integers(0, 6))
        kmax = kmin + int(rng.integers(0, 8))
        terr = float(10.0 ** rng.uniform(-4, -1))
        kept, num, err, ent, ent1 = otrunc.select(eig, twos, kmin, kmax, terr)
        allv = np.concatenate(eig)
        evs = np.sort(allv / allv.sum())[::-1]
        # the selection loop adds at least one state beyond kept_min
        kb = _brute_select(list(evs), kmin, kmax, terr)
        if len(evs) > kmin and kb == kmin and sum(evs[kmin:]) > terr:
            kb = kmin + 1
        assert num == max(kb, min(len(evs), kmin)) or (len(evs) > kmin and num == kmin + 1 and kb <= kmin + 1)
        assert sum(kept) == num
        assert abs(err - max(0.0, evs[num:].sum())) < 1e-12 or num == len(evs)
        # per-sector kept are the largest of each sector
        for s in range(nsec):
            if kept[s]:
                thr = np.sort(eig[s])[::-1][kept[s] - 1] / allv.sum()
                assert thr >= evs[num - 1] - 1e-15
        p = evs[evs >= 1e-15]
        assert abs(ent + np.sum(p * np.log(p))) < 1e-12


def test_truncation_operator_invariants():
    rng = np.random.default_rng(7)
    secs = [rng.normal(size=(d, e)) + 1j * rng.normal(size=(d, e)) for d, e in ((4, 3), (6, 6), (2, 5), (5, 1))]
    norm = np.sqrt(sum(np.sum(np.abs(x) ** 2) for x in secs))
    secs = [x / norm for x in secs]
    rho = otrunc.density_matrix(secs, 0)
    assert abs(sum(np.trace(r).real for r in rho) - 1) < 1e-13
    for r in rho:
        assert np.allclose(r, r.conj().T)
        assert np.linalg.eigvalsh(r).min() > -1e-14
    res = otrunc.truncation_operator(rho, [0, 2, 1, 3], 3, 8, 1e-3)
    kept_w = 0.0
    for r, U, k in zip(rho, res["U"], res["kept"]):
        if k:
            assert np.allclose(U.conj().T @ U, np.eye(k), atol=1e-12)
            kept_w += np.trace(U.conj().T @ r @ U).real
    assert abs((1 - kept_w) - res["trun_err"]) < 1e-12
    env = otrunc.density_matrix(secs, 1)
    for x, r in zip(secs, env):
        assert np.allclose(r, x.conj().T @ x)


def test_uou_identity_and_definition():
    rng = np.random.default_rng(8)
    O = rng.normal(size=(5, 4)) + 1j * rng.normal(size=(5, 4))
    (eff,) = ouou.update_trun([(O, np.eye(5), np.eye(4))])
    assert np.allclose(eff, O)
    Ul = np.linalg.qr(rng.normal(size=(5, 3)) + 1j * rng.normal(size=(5, 3)))[0]
    Ur = np.linalg.qr(rng.normal(size=(4, 2)) + 1j * rng.normal(size=(4, 2)))[0]
    (eff,) = ouou.update_trun([(O, Ul, Ur)])
    for i in range(3):
        for j in range(2):
            assert abs(eff[i, j] - sum(np.conj(Ul[a, i]) * O[a, b] * Ur[b, j] for a in range(5) for b in range(4))) < 1e-13
