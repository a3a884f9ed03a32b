"""The reference's superblock H.psi kernels, restated (oracle / test
infrastructure only; nothing in the product imports it).

model_wave (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:6915-7393)
applies the superblock Hamiltonian as a sum of calls to 14 kernels
(get_sys.f90:1783-5204), each the product of an operator on one part of the
superblock (sys block, sys site, env site, env block) with an operator on
another, coupled to a scalar. Every kernel carries its own 6j coefficients,
phase and fermion signs. They are restated here loop by loop from the
Fortran text, with the operators in the reference's Total_Block form:
sub-blocks keyed by (ket N, ket 2S, 2S_bra - 2S_ket) holding bra x ket
matrices, env-side operators as the reference stores them (the complex
conjugate of the physical reduced matrix, which its kernels undo with a
conjugate transpose), and block_transfer_trans (block.f90:33-64) where
model_wave passes one.

Applied to wave functions read in the reference's conventions
(oracle/ref_blocks.py: Fock-space block states coupled (block, site), sys
then env), each kernel must act as one constant times a physical operator
(sum_s c+_is c_js + h.c., S_i.S_j, n_i n_j, the block Hamiltonians); the
tests check that, and that the constant is the one model_wave's
coefficients assume (jt = -sqrt2 t, jd = -sqrt3 J, jn = -J/4,
main.f90:61-73). Angular momenta are doubled integers.
"""
import numpy as np

from .wigner import w6j

SITE = [(1, 1), (0, 0)]      # site sectors: electron doublet, hole


def _pm(x2):
    assert x2 % 2 == 0
    return -1.0 if (x2 // 2) % 2 else 1.0


class ROp:
    """A Total_Block: particle change up, rank k (down_dif), subs[(ket N,
    ket 2S, dd)] = bra x ket matrix."""

    def __init__(self, up, k, subs):
        self.up, self.k, self.subs = up, k, dict(subs)

    def get(self, n, j2, dd):
        return self.subs.get((n, j2, dd))


def trans(op):
    """block_transfer_trans (block.f90:33-64): up_dif negated, each sub-block
    re-keyed by its bra, down_dif negated, matrix conjugate-transposed."""
    out = {}
    for (n, j2, dd), m in op.subs.items():
        out[(n + op.up, j2 + dd, -dd)] = np.conj(m).T
    return ROp(-op.up, op.k, out)


def site_op(name):
    """Get_site_operator_su2 (dmrg_sub.f90:733-908) for the t-J site."""
    r2 = np.sqrt(2.0)
    if name == "st_elec_up":
        return ROp(1, 1, {(0, 0, 1): np.array([[-r2]])})
    if name == "st_elec_down":
        return ROp(-1, 1, {(1, 1, -1): np.array([[-r2]])})
    if name == "st_sd":
        return ROp(0, 2, {(1, 1, 0): np.array([[np.sqrt(1.5)]])})
    if name == "num_elec":
        return ROp(0, 0, {(1, 1, 0): np.array([[r2]]), (0, 0, 0): np.array([[0.0]])})
    raise KeyError(name)


def from_reduced(red, sectors, up, k, env=False):
    """An ROp from {(bra sector, ket sector): reduced matrix} over a block's
    sectors [(N, 2S, dim)]; env: stored as the reference stores env
    operators (complex conjugate)."""
    subs = {}
    for (b, kk), m in red.items():
        nk, jk, _ = sectors[kk]
        subs[(nk, jk, sectors[b][1] - jk)] = np.conj(m) if env else np.array(m)
    return ROp(up, k, subs)


# ---------------------------------------------------------------------
# bookkeeping on an enlarged basis (BlockStates.enlarge of ref_blocks)

def _subs(eb, x):
    """[(bl (N, 2S), st (N, 2S), spos, sdim)] of enlarged sector x."""
    return [(tuple(eb.block_sectors[a][:2]), SITE[s], p, sd) for (a, s, p, sd) in eb.basis[x][3]]


def _sector(eb, n, j2):
    for x, b in enumerate(eb.basis):
        if (b[0], b[1]) == (n, j2):
            return x
    return None


def _find(eb, x, bl, st):
    if x is None:
        return None
    for (b, s, p, sd) in _subs(eb, x):
        if b == bl and s == st:
            return p, sd
    return None


def _qn(eb, x):
    return eb.basis[x][0], eb.basis[x][1]


def _new_psi(psi):
    return {key: np.zeros_like(m, dtype=complex) for key, m in psi.items()}


def _parity(n):
    return -1.0 if n % 2 else 1.0


# ---------------------------------------------------------------------
# operators inside one enlarged part

def block_site_vec_ndia(sys_bl, flag_bl, sys_st, flag_st, sys_eb, psi, coef):
    """get_sys.f90:2578-2840 (the live branch, up to 'go to 110')."""
    out = _new_psi(psi)
    up, k = sys_bl.up, sys_bl.k
    for (x, y), v in psi.items():
        _ns, Js = _qn(sys_eb, x)
        for (bl, st, p, sd) in _subs(sys_eb, x):
            for dd in range(-k, k + 1, 2):
                blm = sys_bl.get(bl[0], bl[1], dd)
                for sdd in range(-k, k + 1, 2):
                    new_st = (st[0] - up, st[1] - sdd)
                    stm = sys_st.get(new_st[0], new_st[1], sdd)
                    new_bl = (bl[0] + up, bl[1] + dd)
                    if blm is None or stm is None:
                        continue
                    f = _find(sys_eb, x, new_bl, new_st)
                    if f is None:
                        continue
                    q, nd = f
                    sign_st = _parity(bl[0]) if flag_st == "F" else 1.0
                    c = w6j(new_bl[1], bl[1], k, st[1], new_st[1], Js)
                    if c == 0.0:
                        continue
                    c *= _pm(bl[1] + new_st[1] + k + Js) / np.sqrt(1.0 + k)
                    c = sign_st * c * coef * stm[0, 0]
                    out[(x, y)][q:q + nd, :] += c * (blm @ v[p:p + sd, :])
                    if up != 0:
                        out[(x, y)][p:p + sd, :] += np.conj(c) * (np.conj(blm).T @ v[q:q + nd, :])
    return out


def vec_block_site_ndia(env_bl, flag_bl, env_st, flag_st, env_eb, sys_eb, psi, coef):
    """get_sys.f90:2842-3015 (the live branch)."""
    out = _new_psi(psi)
    up, k = env_bl.up, env_bl.k
    for (x, y), v in psi.items():
        ns, _Js = _qn(sys_eb, x)
        _ne, Je = _qn(env_eb, y)
        for (bl, st, p, sd) in _subs(env_eb, y):
            for bdd in range(-k, k + 1, 2):
                blm = env_bl.get(bl[0], bl[1], bdd)
                for sdd in range(-k, k + 1, 2):
                    new_st = (st[0] - up, st[1] - sdd)
                    stm = env_st.get(new_st[0], new_st[1], sdd)
                    new_bl = (bl[0] + up, bl[1] + bdd)
                    if blm is None or stm is None:
                        continue
                    f = _find(env_eb, y, new_bl, new_st)
                    if f is None:
                        continue
                    q, nd = f
                    sign_bl = _parity(ns + new_st[0]) if flag_bl == "F" else 1.0
                    sign_st = _parity(ns) if flag_st == "F" else 1.0
                    c = w6j(new_bl[1], bl[1], k, st[1], new_st[1], Je)
                    if c == 0.0:
                        continue
                    c *= _pm(bl[1] + new_st[1] + k + Je) / np.sqrt(1.0 + k)
                    c = sign_bl * sign_st * c * coef * stm[0, 0]
                    out[(x, y)][:, q:q + nd] += c * (v[:, p:p + sd] @ np.conj(blm).T)
                    if up != 0:
                        out[(x, y)][:, p:p + sd] += np.conj(c) * (v[:, q:q + nd] @ blm)
    return out


# ---------------------------------------------------------------------
# operators on both parts

def _target(sys_eb, env_eb, psi, ns, ne, Jn):
    xn, yn = _sector(sys_eb, ns, Jn), _sector(env_eb, ne, Jn)
    if xn is None or yn is None or (xn, yn) not in psi:
        return None
    return xn, yn


def block_vec_site_ndia(sys_bl, flag_sl, env_st, flag_et, sys_eb, env_eb, psi, coef):
    """get_sys.f90:3970-4180 (the live branch, case <1>)."""
    out = _new_psi(psi)
    up, k = sys_bl.up, sys_bl.k
    for (x, y), v in psi.items():
        ns, Js = _qn(sys_eb, x)
        ne, Je = _qn(env_eb, y)
        for (sbl, sst, p, sd) in _subs(sys_eb, x):
            for (ebl, est, pe, sde) in _subs(env_eb, y):
                for bdd in range(-k, k + 1, 2):
                    blm = sys_bl.get(sbl[0], sbl[1], bdd)
                    new_sbl = (sbl[0] + up, sbl[1] + bdd)
                    for sdd in range(-k, k + 1, 2):
                        new_est = (est[0] - up, est[1] - sdd)
                        stm = env_st.get(new_est[0], new_est[1], sdd)
                        if blm is None or stm is None:
                            continue
                        for Jn in range(abs(Js - k), Js + k + 1, 2):
                            t = _target(sys_eb, env_eb, psi, ns + up, ne - up, Jn)
                            if t is None:
                                continue
                            fs, fe = _find(sys_eb, t[0], new_sbl, sst), _find(env_eb, t[1], ebl, new_est)
                            if fs is None or fe is None:
                                continue
                            sign_et = _parity(ns) if flag_et == "F" else 1.0
                            j1, j2, j3, j4, j5, j6 = Jn, k, Js, sbl[1], sst[1], new_sbl[1]
                            j11, j22, j33, j44, j55, j66 = Jn, k, Je, est[1], ebl[1], new_est[1]
                            c = w6j(j1, j2, j3, j4, j5, j6) * w6j(j11, j22, j33, j44, j55, j66)
                            if c == 0.0:
                                continue
                            c *= np.sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + k))
                            c *= _pm(j6 + j5 + j3 + j2) * _pm(j55 + j44 + j11 + j22)
                            c = sign_et * c * coef * stm[0, 0]
                            (qs, _), (qe, _) = fs, fe
                            out[t][qs:qs + blm.shape[0], qe:qe + sde] += c * (blm @ v[p:p + sd, pe:pe + sde])
                            if up != 0:
                                out[(x, y)][p:p + sd, pe:pe + sde] += np.conj(c) * (
                                    np.conj(blm).T @ psi[t][qs:qs + blm.shape[0], qe:qe + sde])
    return out


def site_vec_block_ndia(sys_st, flag_st, env_bl, flag_el, sys_eb, env_eb, psi, coef):
    """get_sys.f90:4349-4600 (the live branch)."""
    out = _new_psi(psi)
    up, k = sys_st.up, sys_st.k
    for (x, y), v in psi.items():
        ns, Js = _qn(sys_eb, x)
        ne, Je = _qn(env_eb, y)
        for (sbl, sst, p, sd) in _subs(sys_eb, x):
            for (ebl, est, pe, sde) in _subs(env_eb, y):
                for sdd in range(-k, k + 1, 2):
                    stm = sys_st.get(sst[0], sst[1], sdd)
                    new_sst = (sst[0] + up, sst[1] + sdd)
                    for bdd in range(-k, k + 1, 2):
                        new_ebl = (ebl[0] - up, ebl[1] - bdd)
                        blm = env_bl.get(new_ebl[0], new_ebl[1], bdd)
                        if stm is None or blm is None:
                            continue
                        for Jn in range(abs(Js - k), Js + k + 1, 2):
                            t = _target(sys_eb, env_eb, psi, ns + up, ne - up, Jn)
                            if t is None:
                                continue
                            fs, fe = _find(sys_eb, t[0], sbl, new_sst), _find(env_eb, t[1], new_ebl, est)
                            if fs is None or fe is None:
                                continue
                            j1, j2, j3, j4, j5, j6 = Jn, k, Js, sst[1], sbl[1], new_sst[1]
                            j11, j22, j33, j44, j55, j66 = Jn, k, Je, ebl[1], est[1], new_ebl[1]
                            c = w6j(j1, j2, j3, j4, j5, j6) * w6j(j11, j22, j33, j44, j55, j66)
                            if c == 0.0:
                                continue
                            c *= np.sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + k))
                            c *= _pm(j66 + j55 + j33 + j22) * _pm(j5 + j4 + j1 + j2)
                            sign_st = _parity(sbl[0]) if flag_st == "F" else 1.0
                            sign_el = _parity(ns + est[0]) if flag_el == "F" else 1.0
                            c = sign_st * sign_el * c * coef * stm[0, 0]
                            (qs, _), (qe, nde) = fs, fe
                            out[t][qs:qs + sd, qe:qe + nde] += c * (v[p:p + sd, pe:pe + sde] @ blm)
                            if up != 0:
                                out[(x, y)][p:p + sd, pe:pe + sde] += np.conj(c) * (
                                    psi[t][qs:qs + sd, qe:qe + nde] @ np.conj(blm).T)
    return out


def block_vec_block_ndia(sys_bl, flag_sl, env_bl, flag_el, sys_eb, env_eb, psi, coef):
    """get_sys.f90:3131-3572 (the live branch from label 1051)."""
    out = _new_psi(psi)
    up, k = sys_bl.up, sys_bl.k
    for (x, y), v in psi.items():
        ns, Js = _qn(sys_eb, x)
        ne, Je = _qn(env_eb, y)
        for (sbl, sst, p, sd) in _subs(sys_eb, x):
            for (ebl, est, pe, sde) in _subs(env_eb, y):
                for dd in range(-k, k + 1, 2):
                    sm = sys_bl.get(sbl[0], sbl[1], dd)
                    new_sbl = (sbl[0] + up, sbl[1] + dd)
                    for edd in range(-k, k + 1, 2):
                        new_ebl = (ebl[0] - up, ebl[1] - edd)
                        em = env_bl.get(new_ebl[0], new_ebl[1], edd)
                        if sm is None or em is None:
                            continue
                        for Jn in range(abs(Js - k), Js + k + 1, 2):
                            t = _target(sys_eb, env_eb, psi, ns + up, ne - up, Jn)
                            if t is None:
                                continue
                            fs, fe = _find(sys_eb, t[0], new_sbl, sst), _find(env_eb, t[1], new_ebl, est)
                            if fs is None or fe is None:
                                continue
                            sign_el = _parity(ns + est[0]) if flag_el == "F" else 1.0
                            j1, j2, j3, j4, j5, j6 = Jn, k, Js, sbl[1], sst[1], new_sbl[1]
                            j11, j22, j33, j44, j55, j66 = Jn, k, Je, ebl[1], est[1], new_ebl[1]
                            c = w6j(j1, j2, j3, j4, j5, j6) * w6j(j11, j22, j33, j44, j55, j66)
                            if c == 0.0:
                                continue
                            c *= np.sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + k))
                            c *= _pm(j6 + j5 + j3 + j2) * _pm(j66 + j55 + j33 + j22)
                            c = sign_el * c * coef
                            (qs, nds), (qe, nde) = fs, fe
                            out[t][qs:qs + nds, qe:qe + nde] += c * (sm @ v[p:p + sd, pe:pe + sde] @ em)
                            if up != 0:
                                out[(x, y)][p:p + sd, pe:pe + sde] += np.conj(c) * (
                                    np.conj(sm).T @ psi[t][qs:qs + nds, qe:qe + nde] @ np.conj(em).T)
    return out


def site_vec_site_ndia(sys_st, flag_st, env_st, flag_et, sys_eb, env_eb, psi, coef):
    """get_sys.f90:4816-5211 (the live branch)."""
    out = _new_psi(psi)
    up, k = sys_st.up, sys_st.k
    for (x, y), v in psi.items():
        ns, Js = _qn(sys_eb, x)
        ne, Je = _qn(env_eb, y)
        for (sbl, sst, p, sd) in _subs(sys_eb, x):
            for (ebl, est, pe, sde) in _subs(env_eb, y):
                for sdd1 in range(-k, k + 1, 2):
                    sm = sys_st.get(sst[0], sst[1], sdd1)
                    new_sst = (sst[0] + up, sst[1] + sdd1)
                    for sdd2 in range(-k, k + 1, 2):
                        new_est = (est[0] - up, est[1] - sdd2)
                        em = env_st.get(new_est[0], new_est[1], sdd2)
                        if sm is None or em is None:
                            continue
                        for Jn in range(abs(Js - k), Js + k + 1, 2):
                            t = _target(sys_eb, env_eb, psi, ns + up, ne - up, Jn)
                            if t is None:
                                continue
                            fs, fe = _find(sys_eb, t[0], sbl, new_sst), _find(env_eb, t[1], ebl, new_est)
                            if fs is None or fe is None:
                                continue
                            sign_st = _parity(sbl[0]) if flag_st == "F" else 1.0
                            sign_et = _parity(ns) if flag_et == "F" else 1.0
                            j1, j2, j3, j4, j5, j6 = Jn, k, Js, sst[1], sbl[1], new_sst[1]
                            j11, j22, j33, j44, j55, j66 = Jn, k, Je, est[1], ebl[1], new_est[1]
                            c = w6j(j1, j2, j3, j4, j5, j6) * w6j(j11, j22, j33, j44, j55, j66)
                            if c == 0.0:
                                continue
                            c *= np.sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + k))
                            c *= _pm(j55 + j44 + j11 + j22) * _pm(j5 + j4 + j1 + j2)
                            c = sign_st * sign_et * c * coef * sm[0, 0] * em[0, 0]
                            (qs, _), (qe, _) = fs, fe
                            out[t][qs:qs + sd, qe:qe + sde] += c * v[p:p + sd, pe:pe + sde]
                            if up != 0:
                                out[(x, y)][p:p + sd, pe:pe + sde] += np.conj(c) * psi[t][qs:qs + sd, qe:qe + sde]
    return out


# ---------------------------------------------------------------------
# scalar ('dia') kernels: operators stored as reduced elements, divided
# by sqrt(2S+1) of their multiplet to act on psi

def block_vec_dia(sys_bl, sys_eb, psi, coef):
    """get_sys.f90:1783-1835: the sys block's Hamiltonian (a scalar over
    the block's sectors) on every block part of the enlarged sys basis,
    divided by sqrt(2S+1) of the block multiplet."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        out[(x, y)] += coef * (_scalar_on_part(sys_eb, x, sys_bl, "bl") @ v)
    return out


def vec_block_dia(env_bl, env_eb, psi, coef):
    """get_sys.f90:1836-1888 (psi times the conjugate transpose)."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        out[(x, y)] += coef * (v @ np.conj(_scalar_on_part(env_eb, y, env_bl, "bl")).T)
    return out


def _scalar_on_part(eb, x, op, part):
    """The enlarged-sector matrix of a block ('bl') or site ('st') scalar
    with its reduced elements divided by sqrt(2S+1) of the part's multiplet
    (the diagonal blocks block_site_vec_dia and its kin apply)."""
    d = eb.basis[x][2]
    M = np.zeros((d, d), complex)
    for (bl, st, p, sd) in _subs(eb, x):
        qn = bl if part == "bl" else st
        m = op.get(qn[0], qn[1], 0)
        if m is None:
            continue
        if part == "bl":
            M[p:p + sd, p:p + sd] += m / np.sqrt(1.0 + qn[1])
        else:
            M[p:p + sd, p:p + sd] += m[0, 0] / np.sqrt(1.0 + qn[1]) * np.eye(sd)
    return M


def block_site_vec_dia(sys_bl, sys_st, sys_eb, psi, coef):
    """get_sys.f90:2006-2080: block scalar times site scalar, both on the
    enlarged sys part (coefficient coef * st / sqrt((2S_bl+1)(2S_st+1)))."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        A = _scalar_on_part(sys_eb, x, sys_bl, "bl")
        B = _scalar_on_part(sys_eb, x, sys_st, "st")
        out[(x, y)] += coef * (A @ B @ v)
    return out


def vec_block_site_dia(env_bl, env_st, env_eb, psi, coef):
    """get_sys.f90:2081-2157."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        A = _scalar_on_part(env_eb, y, env_bl, "bl")
        B = _scalar_on_part(env_eb, y, env_st, "st")
        out[(x, y)] += coef * (v @ np.conj(A @ B).T)
    return out


def site_vec_site_dia(sys_st, env_st, sys_eb, env_eb, psi, coef):
    """get_sys.f90:2158-2257."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        A = _scalar_on_part(sys_eb, x, sys_st, "st")
        B = _scalar_on_part(env_eb, y, env_st, "st")
        out[(x, y)] += coef * (A @ v @ B.T)
    return out


def block_vec_site_dia(sys_bl, env_st, sys_eb, env_eb, psi, coef):
    """get_sys.f90:2258-2360."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        A = _scalar_on_part(sys_eb, x, sys_bl, "bl")
        B = _scalar_on_part(env_eb, y, env_st, "st")
        out[(x, y)] += coef * (A @ v @ B.T)
    return out


def site_vec_block_dia(sys_st, env_bl, sys_eb, env_eb, psi, coef):
    """get_sys.f90:2361-2463."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        A = _scalar_on_part(sys_eb, x, sys_st, "st")
        B = _scalar_on_part(env_eb, y, env_bl, "bl")
        out[(x, y)] += coef * (A @ v @ np.conj(B).T)
    return out


def block_vec_block_dia(sys_bl, env_bl, sys_eb, env_eb, psi, coef):
    """get_sys.f90:2464-2577."""
    out = _new_psi(psi)
    for (x, y), v in psi.items():
        A = _scalar_on_part(sys_eb, x, sys_bl, "bl")
        B = _scalar_on_part(env_eb, y, env_bl, "bl")
        out[(x, y)] += coef * (A @ v @ np.conj(B).T)
    return out


# ---------------------------------------------------------------------
# the wave-function guess between steps (dmrg_sub.f90:1471-1789)

def wave_trans_LR(sys_eb, env_eb, sys_trun, env_trun, old_eb, old_psi, new_keys):
    """wave_trans_LR (dmrg_sub.f90:1471-1633): psi of the superblock
    (old_eb = (old enlarged sys, old enlarged env)) carried to the next one
    of a left-to-right pass: the old enlarged sys truncated by sys_trun
    ({(N, 2S): U}) becomes the new sys block, the old env site becomes the
    new sys site, the old env block is expanded back to its enlarged basis
    by env_trun (the truncation that built it), applied conjugate-transposed.
    sys_eb / env_eb: the new enlarged bases; new_keys: the new psi sectors."""
    o_sys, o_env = old_eb
    new = {key: np.zeros((sys_eb.basis[key[0]][2], env_eb.basis[key[1]][2]), complex) for key in new_keys}
    for (x, y), v in old_psi.items():
        ns, js = _qn(o_sys, x)
        ne, je = _qn(o_env, y)
        us = sys_trun.get((ns, js))
        if us is None:
            continue
        for (ebl, est, p, sd) in _subs(o_env, y):
            ue = env_trun.get(ebl)
            if ue is None:
                continue
            for jn in range(abs(js - est[1]), js + est[1] + 1, 2):
                xs = _sector(sys_eb, ns + est[0], jn)
                ye = _sector(env_eb, ebl[0], ebl[1])
                f = _find(sys_eb, xs, (ns, js), est)
                if f is None or ye is None or (xs, ye) not in new:
                    continue
                q, nd = f
                j12, j3, j4, j123, j34 = js, est[1], ebl[1], jn, je
                c = _pm(j12 + j3 + j123) / np.sqrt((1.0 + j12) * (1.0 + j123))
                c *= np.sqrt((1.0 + j123) * (1.0 + j34))
                c *= _pm(j12 + j3 + j4) * _pm(j3 + j4 - j34)
                sub = c * (np.conj(us).T @ v[:, p:p + sd])
                new[(xs, ye)][q:q + nd, :] += sub @ np.conj(ue).T
    nrm = np.sqrt(sum(np.vdot(m, m).real for m in new.values()))
    return {k: m / nrm for k, m in new.items()}


def wave_trans_RL(sys_eb, env_eb, sys_trun, env_trun, old_eb, old_psi, new_keys):
    """wave_trans_RL (dmrg_sub.f90:1637-1789): the right-to-left move; the
    old enlarged env truncated by env_trun becomes the new env block, the
    old sys site the new env site, and the old sys block is expanded back
    by sys_trun; both truncations applied without conjugation."""
    o_sys, o_env = old_eb
    new = {key: np.zeros((sys_eb.basis[key[0]][2], env_eb.basis[key[1]][2]), complex) for key in new_keys}
    for (x, y), v in old_psi.items():
        ns, js = _qn(o_sys, x)
        ne, je = _qn(o_env, y)
        ue = env_trun.get((ne, je))
        if ue is None:
            continue
        for (sbl, sst, p, sd) in _subs(o_sys, x):
            us = sys_trun.get(sbl)
            if us is None:
                continue
            for jn in range(abs(je - sst[1]), je + sst[1] + 1, 2):
                ye = _sector(env_eb, ne + sst[0], jn)
                xs = _sector(sys_eb, sbl[0], sbl[1])
                f = _find(env_eb, ye, (ne, je), sst)
                if f is None or xs is None or (xs, ye) not in new:
                    continue
                q, nd = f
                j12, j3, j4, j123, j34 = je, sst[1], sbl[1], jn, js
                c = _pm(j34 + j4 + j3) / np.sqrt((1.0 + j12) * (1.0 + j4))
                c *= np.sqrt((1.0 + j123) * (1.0 + j34))
                c *= _pm(j4 + j3 + j12) * _pm(j12 + j3 - j123)
                sub = c * (us @ v[p:p + sd, :])
                new[(xs, ye)][:, q:q + nd] += sub @ ue
    nrm = np.sqrt(sum(np.vdot(m, m).real for m in new.values()))
    return {k: m / nrm for k, m in new.items()}


# ---------------------------------------------------------------------
# enlargement: an operator of the block or of the new site carried into
# the enlarged basis (block.f90:2495-3459)

def _enlarged_op(eb, up, k, place):
    """Total_Block skeleton of an enlarged-basis operator: every (ket
    sector, dd) whose bra sector exists (update_block_ndia_sys's first two
    loops), with zero matrices; place(ket x, bra x, dd, M) fills one."""
    subs = {}
    for x, (n, j, d, _s) in enumerate(eb.basis):
        for dd in range(-k, k + 1, 2):
            xb = _sector(eb, n + up, j + dd)
            if xb is None:
                continue
            M = np.zeros((eb.basis[xb][2], d), complex)
            place(x, xb, dd, M)
            subs[(n, j, dd)] = M
    return ROp(up, k, subs)


def update_block_ndia(block, eb, flag, env):
    """update_block_ndia_sys (block.f90:2624-2820) and, env = True,
    update_block_ndia_env (block.f90:2828-3036): a block operator in the
    enlarged basis; the env routine signs fermionic operators by
    (-1)^N_site of the ket's site."""
    k, up = block.k, block.up

    def place(x, xb, dd, M):
        _n, J, _d, _s = eb.basis[x]
        Jb = eb.basis[xb][1]
        for (bl, st, p, sd) in _subs(eb, x):
            for sdd in range(-k, k + 1, 2):
                m = block.get(bl[0], bl[1], sdd)
                if m is None:
                    continue
                new_bl = (bl[0] + up, bl[1] + sdd)
                f = _find(eb, xb, new_bl, st)
                if f is None:
                    continue
                q, nd = f
                c = w6j(Jb, k, J, bl[1], st[1], new_bl[1]) * np.sqrt((1.0 + Jb) * (1.0 + J))
                c *= _pm(new_bl[1] + st[1] + J + k)
                if env and flag == "F":
                    c *= _parity(st[0])
                M[q:q + nd, p:p + sd] += c * m

    return _enlarged_op(eb, up, k, place)


def update_site_ndia(site, eb, flag, env):
    """update_site_ndia_sys (block.f90:3047-3260), signing fermionic site
    operators by (-1)^N_block, and update_site_ndia_env (block.f90:
    3269-3459), which signs nothing."""
    k, up = site.k, site.up

    def place(x, xb, dd, M):
        _n, J, _d, _s = eb.basis[x]
        Jb = eb.basis[xb][1]
        for (bl, st, p, sd) in _subs(eb, x):
            for sdd in range(-k, k + 1, 2):
                m = site.get(st[0], st[1], sdd)
                if m is None:
                    continue
                new_st = (st[0] + up, st[1] + sdd)
                f = _find(eb, xb, bl, new_st)
                if f is None:
                    continue
                q, _nd = f
                c = w6j(Jb, k, J, st[1], bl[1], new_st[1]) * _pm(bl[1] + st[1] + Jb + k)
                c *= np.sqrt((1.0 + Jb) * (1.0 + J))
                if not env and flag == "F":
                    c *= _parity(bl[0])
                M[q:q + sd, p:p + sd] += c * m[0, 0] * np.eye(sd)

    return _enlarged_op(eb, up, k, place)


def update_block_dia(block, eb):
    """update_block_dia (block.f90:2495-2548): a block scalar (reduced
    elements) in the enlarged basis, times sqrt((2J+1)/(2j_block+1))."""
    def place(x, xb, dd, M):
        J = eb.basis[x][1]
        for (bl, st, p, sd) in _subs(eb, x):
            m = block.get(bl[0], bl[1], 0)
            if m is not None:
                M[p:p + sd, p:p + sd] += np.sqrt((1.0 + J) / (1.0 + bl[1])) * m

    return _enlarged_op(eb, 0, 0, place)


def update_site_dia(site, eb):
    """update_site_dia (block.f90:2554-2615)."""
    def place(x, xb, dd, M):
        J = eb.basis[x][1]
        for (bl, st, p, sd) in _subs(eb, x):
            m = site.get(st[0], st[1], 0)
            if m is not None:
                M[p:p + sd, p:p + sd] += np.sqrt((1.0 + J) / (1.0 + st[1])) * m[0, 0] * np.eye(sd)

    return _enlarged_op(eb, 0, 0, place)


# ---------------------------------------------------------------------
# the enlarged block's Hamiltonian (Get_Hamiltonian_Sys/Env,
# get_sys.f90:4-592)

def get_hamiltonian_ndia(block, site, eb, coef, flag_sign, flags):
    """Get_hamiltonian_ndia (get_sys.f90:865-1056): the scalar coupling of
    a block operator and a site operator (site data read as <st|.|new st>,
    new st = st - up_dif of the site data) added to the enlarged block's
    Hamiltonian, with its Hermitian conjugate when the particle number
    changes; flags 'S' (sys: fermion sign (-1)^N of the ket's block) or
    'E' (env: -1 when the ket's site holds an even number). Returns
    {enlarged sector: matrix}."""
    H = {x: np.zeros((d, d), complex) for x, (_n, _j, d, _s) in enumerate(eb.basis)}
    up, k = block.up, block.k
    for x, (_n, J, _d, _s) in enumerate(eb.basis):
        for (bl, st, p, sd) in _subs(eb, x):
            for sdd in range(-k, k + 1, 2):
                m = block.get(bl[0], bl[1], sdd)
                if m is None:
                    continue
                new_bl = (bl[0] + up, bl[1] + sdd)
                for jn in range(abs(st[1] - site.k), st[1] + site.k + 1, 2):
                    new_st = (st[0] - site.up, jn)
                    sm = site.get(new_st[0], new_st[1], st[1] - jn)
                    if sm is None:
                        continue
                    f = _find(eb, x, new_bl, new_st)
                    if f is None:
                        continue
                    q, nd = f
                    c = w6j(new_bl[1], bl[1], k, st[1], new_st[1], J)
                    c *= np.sqrt((1.0 + J) / (1.0 + k)) * _pm(bl[1] + new_st[1] + k + J) * sm[0, 0]
                    sign = 1.0
                    if flag_sign == "F":
                        if flags == "S":
                            sign = _parity(bl[0])
                        elif st[0] % 2 == 0:
                            sign = -1.0
                    c1 = c * coef * sign
                    if c1 == 0.0:
                        continue
                    H[x][q:q + nd, p:p + sd] += c1 * m
                    if site.up != 0:
                        H[x][p:p + sd, q:q + nd] += np.conj(c1) * np.conj(m).T
    return H


def get_hamiltonian_dia(A, B, eb, coef):
    """Get_hamiltonian_dia (get_sys.f90:776-854) with flags 'N', 'N' and no
    conjugate part: coef / sqrt(2J+1) A B per enlarged sector (A, B scalars
    in the enlarged basis, as update_block_dia / update_site_dia give)."""
    H = {}
    for x, (n, J, d, _s) in enumerate(eb.basis):
        a, b = A.get(n, J, 0), B.get(n, J, 0)
        H[x] = np.zeros((d, d), complex) if a is None or b is None else coef / np.sqrt(1.0 + J) * (a @ b)
    return H
