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
