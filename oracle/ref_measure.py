"""What the reference's measurement files hold, in physical terms (oracle /
test infrastructure only; nothing in the product imports it).

The reference couples two reduced (SU(2)) operators with
block_mul_block_ndia (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/block.f90:376-635):
the reduced element of the coupled operator [A^k1 x B^k2]^k is rebuilt from
the m-components of A and B,

    <J'||[A x B]^k||J> (J' k J; -J' q J)
        = sum_Mmid <J' J'|A_q1|Jmid Mmid> <Jmid Mmid|B_q2|J J> C(k1 q1 k2 q2|k q)

with the Wigner-Eckart convention <J'M'|T^k_q|JM> = (-1)^(J'-M')
(J' k J; -M' q M) <J'||T||J> (block.f90:537-575, the phase
(-1)**((j11+j44)/2) on the second factor) and the expansion coefficient

    C(k1 q1 k2 q2|k q) = (-1)^(-k1+k2+q) sqrt(2k+1) (k1 k2 k; q1 q2 -q)

(block.f90:566-569). For half-integer k1 - k2 this is minus the standard
Clebsch-Gordan coefficient (-1)^(k1-k2+q) sqrt(2k+1) (...), so coupling a
scalar with a rank-1/2 operator flips the sign. measure_sys_block
(measure.f90:778-840) divides a scalar block operator by sqrt(2S+1) of its
sector, i.e. returns <psi|X|psi> of the physical scalar X, and
Get_density_cor multiplies by sqrt(1+down_dif) and negates for down_dif = 2
(measure.f90:2586-2587,2662-2663).

This module builds those coupled operators as explicit matrices on a small
Fock space from the reference's site reduced elements (Get_site_operator_su2,
dmrg_sub.f90:733-860: c+ = -sqrt2 on hole -> electron, its annihilation
partner st_elec_down = -sqrt2 on electron -> hole, S = sqrt(3/2), n = sqrt2),
so tests can compare them with the physical operators (S_i.S_j, n_i n_j,
sum_s c+_is c_js, singlet pairs) and thereby fix the factor and sign of
every file column. Angular momenta are doubled integers throughout.
"""
import itertools
from math import sqrt

import numpy as np
from sympy import Rational
from sympy.physics.wigner import wigner_3j


def w3j(j1, j2, j3, m1, m2, m3):
    """3j symbol, doubled arguments (the reference's w3js)."""
    if (m1 + m2 + m3) != 0:
        return 0.0
    return float(wigner_3j(Rational(j1, 2), Rational(j2, 2), Rational(j3, 2),
                           Rational(m1, 2), Rational(m2, 2), Rational(m3, 2)))


def _pm1(n2):
    """(-1)^(n2/2) for an even doubled exponent."""
    assert n2 % 2 == 0
    return -1.0 if (n2 // 2) % 2 else 1.0


def ref_cg(k1, q1, k2, q2, k, q):
    """The reference's expansion coefficient of [A^k1 x B^k2]^k_q in
    A_q1 B_q2 (block.f90:566-569), doubled arguments."""
    if q1 + q2 != q:
        return 0.0
    return w3j(k1, k2, k, q1, q2, -q) * _pm1(-k1 + k2 + q) * sqrt(k + 1.0)


def std_cg(k1, q1, k2, q2, k, q):
    """Standard Clebsch-Gordan coefficient <k1 q1 k2 q2|k q> (Condon-Shortley)."""
    if q1 + q2 != q:
        return 0.0
    return w3j(k1, k2, k, q1, q2, -q) * _pm1(k1 - k2 + q) * sqrt(k + 1.0)


# single t-J site: states (2j, 2m) in the order hole, up, down
SITE_STATES = [(0, 0), (1, 1), (1, -1)]
# the reference's site reduced elements: (rank 2k, {(2J bra, 2J ket): value})
REF_SITE_OPS = {
    "c+": (1, {(1, 0): -sqrt(2.0)}),       # st_elec_up, dmrg_sub.f90:759-765
    "c~": (1, {(0, 1): -sqrt(2.0)}),       # st_elec_down, dmrg_sub.f90:785-791
    "S": (2, {(1, 1): sqrt(1.5)}),          # spin_sd, dmrg_sub.f90:859
    "n": (0, {(1, 1): sqrt(2.0)}),          # num_elec, dmrg_sub.f90:809
}


def site_components(name):
    """m-components T_q (dict 2q -> 3x3 matrix on SITE_STATES) of a site
    operator from its reduced element, with the reference's Wigner-Eckart
    convention."""
    k, red = REF_SITE_OPS[name]
    comps = {}
    for q in range(-k, k + 1, 2):
        M = np.zeros((3, 3))
        for a, (jb, mb) in enumerate(SITE_STATES):
            for b, (jk, mk) in enumerate(SITE_STATES):
                r = red.get((jb, jk))
                if r is None:
                    continue
                M[a, b] = _pm1(jb - mb) * w3j(jb, k, jk, -mb, q, mk) * r
        comps[q] = M
    return k, comps


class Fock:
    """The projected (no double occupancy) Fock space of L sites, fermion
    modes ordered (site 1 up, site 1 down, site 2 up, ...)."""

    def __init__(self, L):
        self.L = L
        self.states = list(itertools.product(range(3), repeat=L))  # per site 0 hole, 1 up, 2 down
        self.index = {s: i for i, s in enumerate(self.states)}

    def _occ_before(self, st, site, spin):
        n = 0
        for i in range(site):
            n += st[i] != 0
        if spin == 1 and st[site] == 1:
            n += 1
        return n

    def cdag(self, site, spin):
        """c+_{site, spin} (spin 0 up, 1 down), projected."""
        D = len(self.states)
        M = np.zeros((D, D))
        for b, st in enumerate(self.states):
            if st[site] != 0:
                continue
            new = list(st)
            new[site] = 1 + spin
            sign = -1.0 if self._occ_before(st, site, spin) % 2 else 1.0
            M[self.index[tuple(new)], b] = sign
        return M

    def site_op(self, site, name):
        """m-components of a reference site operator on `site`, embedded with
        the Jordan-Wigner string for the fermionic ones."""
        k, comps = site_components(name)
        out = {}
        for q, m in comps.items():
            D = len(self.states)
            M = np.zeros((D, D))
            for b, st in enumerate(self.states):
                for a_loc in range(3):
                    v = m[a_loc, st[site]]
                    if v == 0.0:
                        continue
                    new = list(st)
                    new[site] = a_loc
                    sign = 1.0
                    if k % 2 == 1:  # odd fermion parity: string over earlier sites
                        sign = -1.0 if sum(x != 0 for x in st[:site]) % 2 else 1.0
                    M[self.index[tuple(new)], b] += sign * v
            out[q] = M
        return k, out


def couple(A, B, k, cg=ref_cg):
    """[A x B]^k as m-components from m-components (A, B given as (2k, comps))."""
    ka, ca = A
    kb, cb = B
    out = {}
    for q in range(-k, k + 1, 2):
        M = 0.0
        for q1, m1 in ca.items():
            q2 = q - q1
            if q2 in cb:
                c = cg(ka, q1, kb, q2, k, q)
                if c != 0.0:
                    M = M + c * (m1 @ cb[q2])
        out[q] = M
    return k, out


def scalar(X):
    """The single component of a rank-0 operator."""
    k, comps = X
    assert k == 0
    return comps[0]


def physical(F, name, *sites):
    """Physical operators on the Fock space (sites 0-based)."""
    cd = lambda i, s: F.cdag(i, s)
    c = lambda i, s: F.cdag(i, s).T
    if name == "n":
        (i,) = sites
        return sum(cd(i, s) @ c(i, s) for s in (0, 1))
    if name == "SS":
        i, j = sites
        Sz = lambda x: 0.5 * (cd(x, 0) @ c(x, 0) - cd(x, 1) @ c(x, 1))
        Sp = lambda x: cd(x, 0) @ c(x, 1)
        return Sz(i) @ Sz(j) + 0.5 * (Sp(i) @ Sp(j).T + Sp(i).T @ Sp(j))
    if name == "nn":
        i, j = sites
        return physical(F, "n", i) @ physical(F, "n", j)
    if name == "hop":
        i, j = sites
        return sum(cd(i, s) @ c(j, s) for s in (0, 1))
    if name == "Ddag":
        i, j = sites
        return (cd(i, 0) @ cd(j, 1) - cd(i, 1) @ cd(j, 0)) / sqrt(2.0)
    raise KeyError(name)


def cross_scalar_coef(J, ja, sa, Je, jb, sb):
    """The coefficient sys_block_env_block_cor_ndia (measure.f90:4580-4605)
    gives a sys-block scalar times an env-block scalar (down_dif1 = 0):
    6j{J 0 J; ja sa ja} 6j{Je 0 Je; jb sb jb} sqrt((Je+1)(Je+1)) with its
    sign rule. ja, jb: block spins; sa, sb: site spins; J, Je: enlarged."""
    from oracle.wigner import w6j
    c = w6j(J, 0, J, ja, sa, ja)
    if c == 0.0:
        return 0.0
    c *= w6j(Je, 0, Je, jb, sb, jb)
    if c == 0.0:
        return 0.0
    c *= sqrt((1.0 + Je) * (1.0 + Je) / 1.0)
    if ((ja + sa + J + 0) // 2 + (jb + sb + Je + 0) // 2) % 2 == 1:
        c = -c
    return c
