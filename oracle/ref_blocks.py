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


__all__ = ["inb1", "BlockStates", "reduced", "fock_ops", "block_hamiltonian", "super_state"]
