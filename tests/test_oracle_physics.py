"""Pins of the oracle's t-J physics: ED against closed forms, the lattice
against triangular geometry, and the dense SU(2) superblock against ED."""
import numpy as np
import pytest

from oracle import dmrg_exact, su2_dense, tj_ed, tj_model


def test_ed_closed_forms():
    ring = [(1, 2, 0.0, 1.0), (2, 3, 0.0, 1.0), (3, 4, 0.0, 1.0), (1, 4, 0.0, 1.0)]
    assert abs(tj_ed.ground_energy(4, 4, ring) - (-3.0)) < 1e-12          # Heisenberg ring, S=0
    assert abs(tj_ed.ground_energy(2, 2, [(1, 2, 0.0, 1.0)]) + 1.0) < 1e-12      # singlet J(-3/4-1/4)
    assert abs(tj_ed.ground_energy(2, 2, [(1, 2, 0.0, 1.0)], twoS=2)) < 1e-12    # triplet 0
    tri = [(1, 2, 1.0, 0.0), (2, 3, 1.0, 0.0), (1, 3, 1.0, 0.0)]
    # one electron on a triangle: band -2t, t, t -> sign of t matters
    assert abs(tj_ed.ground_energy(3, 1, tri, twoS=1) + 2.0) < 1e-12
    tri_m = [(a, b, -t, J) for (a, b, t, J) in tri]
    assert abs(tj_ed.ground_energy(3, 1, tri_m, twoS=1) + 1.0) < 1e-12


def test_lattice_is_triangular():
    Lx, Ny = 5, 4
    b = tj_model.bonds(Lx, Ny, 1.0, 0.5, 1.0, 0.25)
    pos = {}
    for ix in range(1, Lx + 1):
        for iy in range(1, Ny + 1):
            pos[iy + (ix - 1) * Ny] = (ix, iy)

    def dist(i, j):
        (x1, y1), (x2, y2) = pos[i], pos[j]
        best = 1e9
        for w in (-Ny, 0, Ny):     # periodic y, minimum image
            dx, dy = x2 - x1, y2 - y1 + w
            best = min(best, np.hypot(dx + 0.5 * dy, np.sqrt(3) / 2 * dy))
        return best

    nn = [(a, c) for (a, c, t, J) in b if t == 1.0]
    nnn = [(a, c) for (a, c, t, J) in b if t == 0.5]
    for a, c in nn:
        assert abs(dist(a, c) - 1.0) < 1e-12
    for a, c in nnn:
        assert abs(dist(a, c) - np.sqrt(3)) < 1e-12
    # bulk coordination: 6 NN and 6 NNN around an interior site
    inner = 2 + 2 * Ny
    assert sum(1 for a, c in nn if inner in (a, c)) == 6
    assert sum(1 for a, c in nnn if inner in (a, c)) == 6
    # open x: 3*Ny*Lx - 2*Ny ... count every NN pair once
    assert len(nn) == Ny * Lx + 2 * Ny * (Lx - 1)


@pytest.mark.parametrize("Lx,Ny,Ne,S2,l", [(2, 3, 4, 0, 2), (2, 3, 5, 1, 2), (2, 4, 6, 0, 3), (3, 3, 8, 0, 3),
                                           (2, 4, 7, 1, 1)])
def test_su2_superblock_equals_ed(Lx, Ny, Ne, S2, l):
    N = Lx * Ny
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(N, Ne, b, twoS=S2)
    e, H = dmrg_exact.ground_energy(N, Ne, S2, b, l=l)
    assert np.allclose(H, H.conj().T, atol=1e-12)
    assert abs(e - e_ed) < 1e-10


def test_get_basis_order_and_dims():
    blk = [(2, 0, 3), (1, 1, 2), (2, 2, 1), (0, 0, 1)]
    site = [(1, 1, 1), (0, 0, 1)]
    bs = su2_dense.get_basis(blk, site)
    qns = [(N, J) for (N, J, _, _) in bs]
    # 2S descending, then N descending, unique
    assert qns == sorted(set(qns), key=lambda q: (-q[1], -q[0]))
    # every (block, site, J) combination appears once; dims add up
    tot = sum(d for (_, _, d, _) in bs)
    expect = sum(bd * sd * len(range(abs(bs_ - ss), bs_ + ss + 1, 2)) for (_, bs_, bd) in blk for (_, ss, sd) in site)
    assert tot == expect
    for (N, J, d, subs) in bs:
        assert [p for (_, _, p, _) in subs] == list(np.cumsum([0] + [x for (_, _, _, x) in subs])[:-1])
        assert [(a, s) for (a, s, _, _) in subs] == sorted((a, s) for (a, s, _, _) in subs)


@pytest.mark.parametrize("Lx,Ny,Ne", [(3, 3, 8), (2, 4, 6)])
def test_ed_correlators_pinned(Lx, Ny, Ne):
    """The ED correlators against facts that do not reuse their strings:
    the bond sum of <hopping>, <S.S> and <n n> rebuilds the ground energy
    (the Hamiltonian assembled independently in tj_ed.hamiltonian),
    sum_i <n_i> = N_e, sum_ij <S_i.S_j> = S(S+1) = 0, and a pair
    correlator equals the overlap <D_ij psi | D_kl psi> built by applying
    the annihilators alone."""
    N = Lx * Ny
    bonds = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    E, v, st, gap = tj_ed.ground_state(N, Ne, bonds)
    assert abs(E - tj_ed.ground_energy(N, Ne, bonds)) < 1e-9
    e = 0.0
    for (i, j, t, J) in bonds:
        hop = tj_ed.expect(v, st, N, tj_ed.hopping_terms(i, j)).real
        ss = tj_ed.expect(v, st, N, tj_ed.spin_terms(i, j)).real
        nn = tj_ed.expect(v, st, N, tj_ed.density_density_terms(i, j)).real
        e += -2.0 * t * hop + J * (ss - 0.25 * nn)
    assert abs(e - E) < 1e-10
    dens = [tj_ed.expect(v, st, N, tj_ed.density_terms(i)).real for i in range(1, N + 1)]
    assert abs(sum(dens) - Ne) < 1e-10
    tot = sum(tj_ed.expect(v, st, N, tj_ed.spin_terms(i, j)).real for i in range(1, N + 1) for j in range(1, N + 1))
    assert abs(tot) < 1e-10
    if Ne == N - 2:
        # annihilate a singlet pair: states of N_e - 2 electrons
        st2 = tj_ed._states(N, Ne - 2, 0)
        idx = {s: k for k, s in enumerate(st2)}

        def ann(i, j):
            out = np.zeros(len(st2), complex)
            for (c, string) in tj_ed.pair_creation(i, j):
                ops = [(s, sp, not dg) for (s, sp, dg) in reversed(string)]
                for k, (up, dn) in enumerate(st):
                    u, d, sgn, ok = up, dn, 1.0, True
                    for (site, spin, dag) in reversed(ops):
                        bits = tj_ed._occ_modes(u, d, N)
                        m = tj_ed._mode(site, spin)
                        if dag == bool(bits >> m & 1):
                            ok = False
                            break
                        if bin(bits & ((1 << m) - 1)).count("1") % 2:
                            sgn = -sgn
                        if spin == 0:
                            u ^= 1 << site
                        else:
                            d ^= 1 << site
                    if ok and (u, d) in idx:
                        out[idx[(u, d)]] += np.conj(c) * sgn * v[k]
            return out

        for (a, b, c, d) in ((1, 2, 5, 6), (1, 4, 7, 8), (2, 3, 6, 7)):
            ref = np.vdot(ann(a, b), ann(c, d))
            got = tj_ed.expect(v, st, N, tj_ed.pair_pair_terms(a, b, c, d))
            assert abs(got - ref) < 1e-12 and abs(ref) > 1e-4


def test_ed_dimer_pair():
    """Two electrons on a t-J dimer: the ground state is the singlet
    D+_12 |0>, so <D+_12 D_12> = 1, and with both sites occupied only the
    exchange acts: E = J (S.S - 1/4) = -J (J = 0.4)."""
    E, v, st, gap = tj_ed.ground_state(2, 2, [(1, 2, 1.0, 0.4)])
    assert abs(E + 0.4) < 1e-12
    assert abs(tj_ed.expect(v, st, 2, tj_ed.pair_pair_terms(1, 2, 1, 2)) - 1.0) < 1e-12
