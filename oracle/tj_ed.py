"""Exact diagonalisation of the t-J model in the U(1) Fock basis (oracle
pin for the SU(2) DMRG on small clusters).

Modes are ordered (site 1 up, site 1 down, site 2 up, ...); a basis state
is a pair of occupation bit masks with no doubly occupied site. The lowest
total-spin-S_target state is isolated with a penalty lambda * S^2 in the
S_z = 0 (or 1/2) sector.
"""
import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _states(L, Ne, twoSz):
    nup = (Ne + twoSz) // 2
    ndn = Ne - nup
    out = []
    for occ in itertools.combinations(range(L), Ne):
        for ups in itertools.combinations(occ, nup):
            up = sum(1 << i for i in ups)
            dn = sum(1 << i for i in occ if i not in ups)
            out.append((up, dn))
    return out


def _mode(site, spin):
    return 2 * site + spin


def _occ_modes(up, dn, L):
    bits = 0
    for i in range(L):
        if up >> i & 1:
            bits |= 1 << _mode(i, 0)
        if dn >> i & 1:
            bits |= 1 << _mode(i, 1)
    return bits


def _sign_between(bits, m1, m2):
    lo, hi = min(m1, m2), max(m1, m2)
    mask = ((1 << hi) - 1) & ~((1 << (lo + 1)) - 1)
    return -1.0 if bin(bits & mask).count("1") % 2 else 1.0


def hamiltonian(L, Ne, bond_list, twoSz=0, spin_penalty=0.0):
    """Sparse H (+ penalty * S^2) on sites 1..L (bond_list 1-based)."""
    states = _states(L, Ne, twoSz)
    index = {s: k for k, s in enumerate(states)}
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r)
        cols.append(c)
        vals.append(v)

    for k, (up, dn) in enumerate(states):
        bits = _occ_modes(up, dn, L)
        diag = 0.0
        for (a, b, t, J) in bond_list:
            i, j = a - 1, b - 1
            ni_u, ni_d = up >> i & 1, dn >> i & 1
            nj_u, nj_d = up >> j & 1, dn >> j & 1
            ni, nj = ni_u + ni_d, nj_u + nj_d
            szi, szj = 0.5 * (ni_u - ni_d), 0.5 * (nj_u - nj_d)
            diag += J * (szi * szj - 0.25 * ni * nj)
            # spin flip-flop: 1/2 (S+_i S-_j + S-_i S+_j)
            if ni_d and nj_u:
                nu, nd = up | (1 << i), dn & ~(1 << i)
                nu, nd = nu & ~(1 << j), nd | (1 << j)
                add(index[(nu, nd)], k, 0.5 * J)
            if ni_u and nj_d:
                nu, nd = up & ~(1 << i), dn | (1 << i)
                nu, nd = nu | (1 << j), nd & ~(1 << j)
                add(index[(nu, nd)], k, 0.5 * J)
            # hopping -t (c+_is c_js + c+_js c_is)
            for spin, mask in ((0, up), (1, dn)):
                for src, dst in ((j, i), (i, j)):
                    if (mask >> src & 1) and not ((up | dn) >> dst & 1):
                        sgn = _sign_between(bits, _mode(src, spin), _mode(dst, spin))
                        m2 = (mask & ~(1 << src)) | (1 << dst)
                        nu, nd = (m2, dn) if spin == 0 else (up, m2)
                        add(index[(nu, nd)], k, -t * sgn)
        add(k, k, diag)
        if spin_penalty:
            # S^2 diagonal: Sz_tot^2 + sum_i (S_i^2 - Sz_i^2) = Sz^2 + n_occ/2; flips below
            sz = 0.5 * (bin(up).count("1") - bin(dn).count("1"))
            nocc = bin(up | dn).count("1")
            add(k, k, spin_penalty * (sz * sz + 0.5 * nocc))
            for i in range(L):
                for j in range(L):
                    if i != j and (dn >> i & 1) and (up >> j & 1):
                        nu, nd = (up | (1 << i)) & ~(1 << j), (dn & ~(1 << i)) | (1 << j)
                        add(index[(nu, nd)], k, spin_penalty * 1.0)
            # the S+S- + S-S+ halves sum to one full flip amplitude per ordered pair
    H = sp.csr_matrix((vals, (rows, cols)), shape=(len(states), len(states)))
    return H, states


def ground_energy(L, Ne, bond_list, twoS=0, penalty=10.0):
    """Lowest energy with total spin S = twoS/2 (S_z = S sector + S^2 penalty)."""
    H, _ = hamiltonian(L, Ne, bond_list, twoSz=twoS, spin_penalty=penalty)
    n = H.shape[0]
    if n <= 2000:
        w = np.linalg.eigvalsh(H.toarray())
    else:
        w = spla.eigsh(H, k=4, which="SA", tol=1e-12)[0]
    s = twoS / 2.0
    return float(np.min(w)) - penalty * s * (s + 1)


# ---------------------------------------------------------------------
# ground state and correlators (pins the measurement of Measurement,
# DMRG_code/main.f90:213-354, by brute force)

def ground_state(L, Ne, bond_list, twoS=0, penalty=10.0):
    """(E, psi, states, gap): lowest total-spin-S state of the cluster, with
    the gap to the next level of the penalised S_z sector (a correlator is
    only fixed by the energy when the ground state is unique)."""
    H, states = hamiltonian(L, Ne, bond_list, twoSz=twoS, spin_penalty=penalty)
    n = H.shape[0]
    if n <= 3000:
        w, v = np.linalg.eigh(H.toarray())
    else:
        w, v = spla.eigsh(H, k=3, which="SA", tol=1e-13)
        o = np.argsort(w)
        w, v = w[o], v[:, o]
    s = twoS / 2.0
    return float(w[0]) - penalty * s * (s + 1), v[:, 0], states, float(w[1] - w[0])


def _apply_string(psi, states, string, L):
    """sum over basis states of (prod of mode operators) psi, operators
    applied right to left; string = [(site, spin, dagger)], 0-based sites.
    Projected t-J operators: a doubly occupied site gives zero."""
    index = {s: k for k, s in enumerate(states)}
    out = np.zeros_like(psi)
    for k, (up, dn) in enumerate(states):
        amp = psi[k]
        if amp == 0:
            continue
        u, d, sgn = up, dn, 1.0
        ok = True
        for (site, spin, dag) in reversed(string):
            bits = _occ_modes(u, d, L)
            m = _mode(site, spin)
            occ = bits >> m & 1
            if dag == bool(occ):
                ok = False
                break
            if bin(bits & ((1 << m) - 1)).count("1") % 2:
                sgn = -sgn
            if spin == 0:
                u ^= 1 << site
            else:
                d ^= 1 << site
            if u & d:
                ok = False
                break
        if not ok:
            continue
        t = index.get((u, d))
        if t is not None:
            out[t] += sgn * amp
    return out


def expect(psi, states, L, terms):
    """<psi| sum_k c_k string_k |psi> for terms [(c_k, string_k)]."""
    acc = 0.0 + 0.0j
    for (c, string) in terms:
        acc += c * np.vdot(psi, _apply_string(psi, states, string, L))
    return acc / np.vdot(psi, psi)


def density_terms(i):
    """n_i (1-based site)."""
    a = i - 1
    return [(1.0, [(a, 0, True), (a, 0, False)]), (1.0, [(a, 1, True), (a, 1, False)])]


def spin_terms(i, j):
    """S_i . S_j = Sz Sz + (S+_i S-_j + S-_i S+_j) / 2, S+ = c+_up c_dn."""
    a, b = i - 1, j - 1
    out = []
    for si in (0, 1):
        for sj in (0, 1):
            c = 0.25 * (1 if si == sj else -1)
            out.append((c, [(a, si, True), (a, si, False), (b, sj, True), (b, sj, False)]))
    out.append((0.5, [(a, 0, True), (a, 1, False), (b, 1, True), (b, 0, False)]))
    out.append((0.5, [(a, 1, True), (a, 0, False), (b, 0, True), (b, 1, False)]))
    return out


def density_density_terms(i, j):
    return [(ci * cj, si + sj) for (ci, si) in density_terms(i) for (cj, sj) in density_terms(j)]


def hopping_terms(i, j):
    """(sum_s c+_is c_js + h.c.) / 2: the real part of <sum_s c+_is c_js>."""
    a, b = i - 1, j - 1
    out = []
    for s in (0, 1):
        out.append((0.5, [(a, s, True), (b, s, False)]))
        out.append((0.5, [(b, s, True), (a, s, False)]))
    return out


def pair_creation(i, j):
    """Singlet pair D+_ij = (c+_i,up c+_j,dn - c+_i,dn c+_j,up) / sqrt(2)."""
    a, b = i - 1, j - 1
    r = 1.0 / np.sqrt(2.0)
    return [(r, [(a, 0, True), (b, 1, True)]), (-r, [(a, 1, True), (b, 0, True)])]


def pair_pair_terms(i, j, k, l):
    """D+_ij D_kl with D_kl = (D+_kl)^dagger."""
    ann = []
    for (c, string) in pair_creation(k, l):
        ann.append((np.conj(c), [(s, sp, not dg) for (s, sp, dg) in reversed(string)]))
    return [(c1 * c2, s1 + s2) for (c1, s1) in pair_creation(i, j) for (c2, s2) in ann]
