"""The reference's block bases as explicit Fock-space states, and its
operator-slot table (oracle / test infrastructure only; nothing in the
product imports it).

A block of the reference (hyx15903/TSC_Tri_t1t2J1J2, DMRG_code/) is grown
one site at a time: the enlarged basis is Get_basis's list of sectors
(block.f90:2087-2245; su2_dense.get_basis restates it), each enlarged state
the SU(2) coupling of a block multiplet (part 1) and a site multiplet
(part 2), and the grown block keeps the columns of the truncation matrix
stored in systruns<L> / envtruns<L>. Its operator updates fix what those
states are as Fock-space vectors:

* the block's operators are embedded with Edmonds' 6j formula for a
  tensor acting on part 1 or part 2 of a coupled pair
  (update_block_ndia_sys / update_site_ndia_sys / update_block_ndia_env /
  update_site_ndia_env, block.f90:2624-3459, the same formula on both
  sides), i.e. the enlarged states are coupled with the standard
  Clebsch-Gordan coefficients in the order (block, site);
* on the sys side a site operator takes the sign (-1)^N_block and a block
  operator none, so the product state is (block creators)(site creators)
  |0>; on the env side a block operator takes (-1)^N_site and a site
  operator none, so the product is (site creators)(block creators)|0>.
  Both are the natural order of chain sites (the sys block holds sites
  1..L-1 and grows by site L; the env block holds sites above its new
  site), so every product state is a plain tensor product of occupation
  strings in chain order.

BlockStates rebuilds those vectors (every multiplet, every M) from a run's
truncation files; reduced() takes the reduced matrix of any operator
between two such bases with the reference's Wigner-Eckart convention
<J'M'|T^k_q|JM> = (-1)^(J'-M') (J' k J; -M' q M) <J'||T||J>. A file that
holds the reference's numbers then agrees with the reduced elements of the
physical operators (c+, c~, S, n, the block Hamiltonian) computed here.

The operator slots restate dmrg_setup0 (dmrg_sub.f90:304-331): INB1(i, ns)
numbers, in increasing i, the sites i <= ns with a neighbour (any of the 18
bond types that dmrg_setup keeps) numbered ns or more; nleg11(ns) counts
them.
"""
from functools import lru_cache

import numpy as np

from . import ref_measure as R
from .su2_dense import get_basis
from .tj_model import neighbour_table
from .wigner import w6j

SITE_SECTORS = [(1, 1, 1), (0, 0, 1)]      # electron doublet, hole


def inb1(Lx, Ny, t2, J2):
    """{ns: [sites in slot order]} for ns = 1 .. N-2 (dmrg_setup0). Bond
    types 7-12 are zeroed by dmrg_setup when all their couplings vanish
    (dmrg_sub.f90:634-643), i.e. when t2 = J2 = 0."""
    neib = neighbour_table(Lx, Ny)
    if t2 == 0 and J2 == 0:
        neib = {k: v for k, v in neib.items() if k[1] <= 6}
    N = Lx * Ny
    out = {}
    for ns in range(1, N - 1):
        out[ns] = [i for i in range(1, ns + 1) if any(neib.get((i, typ), 0) >= ns for typ in range(1, 19))]
    return out


@lru_cache(maxsize=None)
def std_cg(j1, m1, j2, m2, J, M):
    return R.std_cg(j1, m1, j2, m2, J, M)


@lru_cache(maxsize=None)
def we_factor(jb, mb, k, q, jk, mk):
    """(-1)^(J'-M') (J' k J; -M' q M), doubled arguments."""
    if -mb + q + mk != 0:
        return 0.0
    return R._pm1(jb - mb) * R.w3j(jb, k, jk, -mb, q, mk)


class BlockStates:
    """Fock vectors of a block basis: sectors [(N, 2S, dim)]; vecs[x][r] =
    {2M: vector over the block's 3^L occupation strings, chain order}."""

    def __init__(self, L, sectors, vecs):
        self.L = L
        self.sectors = list(sectors)
        self.vecs = vecs
        self.index = {(n, j): x for x, (n, j, _d) in enumerate(self.sectors)}

    @staticmethod
    def site():
        e = np.eye(3)   # occupation 0 hole, 1 up, 2 down
        return BlockStates(1, SITE_SECTORS, [[{1: e[1], -1: e[2]}], [{0: e[0]}]])

    def enlarge(self, site_first):
        """The enlarged basis (block (x) site, coupled in that order), as
        product states in chain order: the new site after the block (sys)
        or before it (env, site_first)."""
        site = BlockStates.site()
        basis = get_basis(self.sectors, SITE_SECTORS)
        vecs = []
        for (_N, J, d, subs) in basis:
            rows = [None] * d
            for (a, s, p, sd) in subs:
                ja, js = self.sectors[a][1], SITE_SECTORS[s][1]
                for i in range(sd):
                    st = {}
                    for M in range(-J, J + 1, 2):
                        v = 0.0
                        for ma, va in self.vecs[a][i].items():
                            ms = M - ma
                            if ms not in site.vecs[s][0]:
                                continue
                            c = std_cg(ja, ma, js, ms, J, M)
                            if c == 0.0:
                                continue
                            vs = site.vecs[s][0][ms]
                            v = v + c * (np.kron(vs, va) if site_first else np.kron(va, vs))
                        st[M] = v
                    rows[p + i] = st
            vecs.append(rows)
        eb = BlockStates(self.L + 1, [(N, J, d) for (N, J, d, _) in basis], vecs)
        eb.basis = basis
        eb.block_sectors = list(self.sectors)
        return eb

    def truncate(self, U):
        """U: [(enlarged sector index, d x kept matrix)] in the new block's
        sector order."""
        sectors, vecs = [], []
        for (x, u) in U:
            n, j, d = self.sectors[x]
            assert u.shape[0] == d
            cols = []
            for c in range(u.shape[1]):
                cols.append({M: sum(u[r, c] * self.vecs[x][r][M] for r in range(d)) for M in range(-j, j + 1, 2)})
            sectors.append((n, j, u.shape[1]))
            vecs.append(cols)
        return BlockStates(self.L, sectors, vecs)


def reduced(bra, ket, k, comps, dN):
    """Reduced matrices {(bra sector, ket sector): matrix} of the operator
    with m-components comps (2q -> Fock matrix), rank 2k = k, particle
    change dN, between two bases of one block; also returns the largest
    deviation of any component from the Wigner-Eckart form (a check that
    the operator is a tensor of that rank in these bases)."""
    out, worst = {}, 0.0
    for xk, (nk, jk, dk) in enumerate(ket.sectors):
        for xb, (nb, jb, db) in enumerate(bra.sectors):
            if nb != nk + dN or abs(jb - jk) > k or jb + jk < k:
                continue
            num = np.zeros((db, dk), complex)
            den = 0.0
            samples = []
            for mk in range(-jk, jk + 1, 2):
                for q, Oq in comps.items():
                    mb = mk + q
                    if abs(mb) > jb:
                        continue
                    w = we_factor(jb, mb, k, q, jk, mk)
                    K = np.array([Oq @ ket.vecs[xk][c][mk] for c in range(dk)]).T
                    B = np.array([bra.vecs[xb][r][mb] for r in range(db)]).conj()
                    E = B @ K
                    samples.append((w, E))
                    num += w * E
                    den += w * w
            if den == 0.0:
                continue
            red = num / den
            for (w, E) in samples:
                worst = max(worst, float(np.abs(E - w * red).max()) if E.size else 0.0)
            out[(xb, xk)] = red
    return out, worst


def fock_ops(L):
    """m-components of the reference's site tensors on every site of an
    L-site block (chain order), and the bond operators, on Fock(L)."""
    F = R.Fock(L)
    ops = {i: {name: F.site_op(i, name) for name in ("c+", "c~", "S", "n")} for i in range(L)}
    return F, ops


def block_hamiltonian(F, bonds_local):
    """sum over bonds (i, j, t, J) of -t sum_s (c+_is c_js + h.c.) +
    J (S_i.S_j - n_i n_j / 4), sites 0-based in the block."""
    H = 0.0
    for (i, j, t, J) in bonds_local:
        hop = R.physical(F, "hop", i, j)
        H = H - t * (hop + hop.T) + J * (R.physical(F, "SS", i, j) - 0.25 * R.physical(F, "nn", i, j))
    return {0: H}


def super_state(sys_states, env_states, subs):
    """The Fock vector (sites in chain order, sys then env) of a total-spin-0
    wave function given as [(sys sector, env sector, matrix)] over an
    enlarged sys and an enlarged env basis, coupled (sys, env) with the
    standard Clebsch-Gordan coefficients."""
    psi = 0.0
    for (x, y, m) in subs:
        js, je = sys_states.sectors[x][1], env_states.sectors[y][1]
        assert js == je
        for M in range(-js, js + 1, 2):
            c = std_cg(js, M, je, -M, 0, 0)
            S = np.array([sys_states.vecs[x][r][M] for r in range(m.shape[0])])
            E = np.array([env_states.vecs[y][cc][-M] for cc in range(m.shape[1])])
            psi = psi + c * np.einsum("rc,ra,cb->ab", m, S, E).ravel()
    return psi


def _pm(x2):
    """(-1)^(x2/2) for an even doubled exponent."""
    assert x2 % 2 == 0
    return -1.0 if (x2 // 2) % 2 else 1.0


# the site operators both hopping routines receive for the t-J model:
# st_elec_up for the block / sys site and block_transfer_trans(st_elec_down)
# for the other side (get_sys.f90:56-57, 6963-6965); as data both are one
# sub-block, ket hole (0, 0), down_dif +1, value -sqrt2 (dmrg_sub.f90:759-791,
# block.f90:33-64 transposes and conjugates st_elec_down's single element)
_SITE_HOP = -np.sqrt(2.0)


def ref_block_site_hop(eb, cplus, coef):
    """Get_hamiltonian_ndia(block = elec_up of a block site, site =
    trans(st_elec_down), basis, coef, ham, 'F', 'S') (get_sys.f90:865-1056):
    the hopping between a sys block site and the new site added to the
    enlarged block's Hamiltonian, as {(x, x): reduced matrix} over the
    enlarged basis eb (BlockStates.enlarge, sys side). cplus: {(bra, ket):
    reduced matrix} of the block site's c+ between the block's sectors
    (eb.block_sectors)."""
    bsec = eb.block_sectors
    H = {}
    for x, (N, J, d, subs) in enumerate(eb.basis):
        M = np.zeros((d, d), complex)
        for (a, st, p, sd) in subs:                     # ket: block a, site st
            nb, jb, _ = bsec[a]
            n_st, j_st, _ = SITE_SECTORS[st]
            for (ab, ak), blk in cplus.items():
                if ak != a:
                    continue
                nb2, jb2, _ = bsec[ab]
                # the site loses its electron: ket site electron, bra hole
                if n_st != 1:
                    continue
                j_new = 0
                for (a2, st2, p2, sd2) in subs:        # bra sub in the same sector
                    if a2 != ab or SITE_SECTORS[st2][1] != j_new or SITE_SECTORS[st2][0] != 0:
                        continue
                    k = 1
                    c11 = w6j(jb2, jb, k, j_st, j_new, J)
                    c11 *= np.sqrt((1.0 + J) / (1.0 + k)) * _pm(jb + j_new + k + J) * _SITE_HOP
                    sign = -1.0 if nb % 2 == 1 else 1.0     # (-1)^N_block of the ket, 'S'
                    c1 = c11 * coef * sign
                    if c1 == 0.0:
                        continue
                    M[p2:p2 + sd2, p:p + sd] += c1 * blk
                    M[p:p + sd, p2:p2 + sd2] += np.conj(c1) * np.conj(blk).T
        H[(x, x)] = M
    return H


def ref_site_site_hop(sys_eb, env_eb, psi, coef):
    """site_vec_site_ndia(st_elec_up, 'F', trans(st_elec_down), 'F',
    sys_bs, env_bs, invec, outvec, coef) (get_sys.f90:4816-5211, called
    from model_wave, get_sys.f90:6963-6966): the hopping between the sys
    site and the env site applied to a total-spin-0 wave function psi =
    {(sys sector, env sector): dx x dy matrix} over the enlarged bases."""
    out = {key: np.zeros_like(m) for key, m in psi.items()}
    k = 1
    for (x, y), vin in psi.items():
        Ns, Js, _dx, subs_s = sys_eb.basis[x]
        Ne, Je, _dy, subs_e = env_eb.basis[y]
        assert Js == Je
        for (a, st, p, sd) in subs_s:
            n_bl, j_bl, _ = sys_eb.block_sectors[a]
            n_st, j_st, _ = SITE_SECTORS[st]
            if n_st != 0:                      # st_elec_up acts on an empty site
                continue
            nsn, jsn = 1, 1                    # new sys site: the electron
            for (ae, ste, pe, sde) in subs_e:
                n_ebl, j_ebl, _ = env_eb.block_sectors[ae]
                n_est, j_est, _ = SITE_SECTORS[ste]
                if n_est != 1:                 # the env site gives up its electron
                    continue
                nen, jen = 0, 0
                for Jn in range(abs(Js - k), Js + k + 1, 2):
                    xs = [i for i, b in enumerate(sys_eb.basis) if (b[0], b[1]) == (Ns + 1, Jn)]
                    ye = [i for i, b in enumerate(env_eb.basis) if (b[0], b[1]) == (Ne - 1, Jn)]
                    if not xs or not ye or (xs[0], ye[0]) not in psi:
                        continue
                    xn, yn = xs[0], ye[0]
                    ps = [q for (a2, s2, q, _d) in sys_eb.basis[xn][3]
                          if a2 == a and SITE_SECTORS[s2][:2] == (nsn, jsn)]
                    pe2 = [q for (a2, s2, q, _d) in env_eb.basis[yn][3]
                           if a2 == ae and SITE_SECTORS[s2][:2] == (nen, jen)]
                    if not ps or not pe2:
                        continue
                    sign_st = -1.0 if n_bl % 2 else 1.0        # sys site passes the sys block
                    sign_et = -1.0 if Ns % 2 else 1.0          # env site passes the whole sys part
                    j1, j2, j3, j4, j5, j6 = Jn, k, Js, j_st, j_bl, jsn
                    j11, j22, j33, j44, j55, j66 = Jn, k, Je, j_est, j_ebl, jen
                    c = w6j(j1, j2, j3, j4, j5, j6) * w6j(j11, j22, j33, j44, j55, j66)
                    if c == 0.0:
                        continue
                    c *= np.sqrt((1.0 + j11) * (1.0 + j33) / (1.0 + k))
                    if ((j55 + j44 + j11 + j22) // 2 + (j5 + j4 + j1 + j2) // 2) % 2:
                        c = -c
                    c *= sign_st * sign_et * coef * _SITE_HOP * _SITE_HOP
                    q_s, q_e = ps[0], pe2[0]
                    out[(xn, yn)][q_s:q_s + sd, q_e:q_e + sde] += c * vin[p:p + sd, pe:pe + sde]
                    out[(x, y)][p:p + sd, pe:pe + sde] += np.conj(c) * psi[(xn, yn)][q_s:q_s + sd, q_e:q_e + sde]
    return out


__all__ = ["inb1", "BlockStates", "reduced", "fock_ops", "block_hamiltonian", "super_state", "ref_block_site_hop",
           "ref_site_site_hop"]
