"""SU(2) block algebra of the two-site DMRG, dense per enlarged sector
(oracle for model_wave, the block enlargement and Get_Hamiltonian_Sys/Env
of hyx15903/TSC_Tri_t1t2J1J2: DMRG_code/get_sys.f90:4-592, 776-1057,
1783-5204, 6915-7626; DMRG_code/block.f90:2087-2245, 2495-3464).

Conventions (Wigner-Eckart in Edmonds' form; quantum numbers (N, 2S)):
  site basis: electron (1, 1) first, then hole (0, 0)
  <e||c+||h> = -sqrt(2), <e||S||e> = sqrt(3/2), <e||n||e> = sqrt(2)
  scalar operators (H) are stored as reduced elements, i.e. with the
  factor sqrt(2S+1) of their multiplet
  c~ (the rank-1/2 annihilation tensor) from c+: <j'||c~||j> =
  (-1)^(j-j'+1/2) conj(<j||c+||j'>)
  bond terms: -t sum_s (c+_i c_j + h.c.) = t sqrt(2) ([c+_P x c~_Q]^0 +
  [c~_P x c+_Q]^0) for P before Q in the fermion order; J S_i.S_j =
  -sqrt(3) J [S x S]^0; -J/4 n_i n_j = -J/4 [n x n]^0.
The enlarged block is (block (x) site) coupled in that order; the fermion
order of a superblock is [sys block, sys site, env block, env site].

Everything here is dense per enlarged sector and written for clarity, not
speed; it is the checker for the product's sub-block task plans.
"""
import numpy as np

from .wigner import w6j

SITE_SECTORS = [(1, 1), (0, 0)]          # electron, hole
SQ2 = np.sqrt(2.0)


def _par(x2):
    """(-1)^(x2/2) for an even doubled exponent x2."""
    assert x2 % 2 == 0, x2
    return -1.0 if (x2 // 2) % 2 else 1.0


class Op:
    """Tensor operator: rank 2k (k2), particle change dN, fermion parity,
    blocks[(bra, ket)] = reduced matrix between sector indices."""

    def __init__(self, k2, dN, blocks=None):
        self.k2, self.dN = k2, dN
        self.odd = dN % 2 != 0
        self.blocks = {} if blocks is None else blocks

    def scaled(self, c):
        return Op(self.k2, self.dN, {k: c * v for k, v in self.blocks.items()})

    def add(self, other, c=1.0):
        assert other.k2 == self.k2 and other.dN == self.dN
        for k, v in other.blocks.items():
            self.blocks[k] = self.blocks[k] + c * v if k in self.blocks else c * v
        return self


class Block:
    """A (possibly enlarged or truncated) block: sectors [(N, 2S, dim)],
    scalar Hamiltonian H (Op of rank 0), and per-site operator sets
    ops[site] = {'c': c+, 'S': S, 'n': n}."""

    def __init__(self, sectors, H, ops, sites):
        self.sectors = list(sectors)
        self.index = {(n, s): i for i, (n, s, _) in enumerate(self.sectors)}
        self.H = H
        self.ops = ops
        self.sites = list(sites)   # chain indices in fermion order

    def dim(self, i):
        return self.sectors[i][2]


def site_block(site):
    sec = [(1, 1, 1), (0, 0, 1)]
    c = Op(1, 1, {(0, 1): np.array([[-SQ2]], complex)})
    S = Op(2, 0, {(0, 0): np.array([[np.sqrt(1.5)]], complex)})
    n = Op(0, 0, {(0, 0): np.array([[SQ2]], complex)})
    H = Op(0, 0, {(0, 0): np.zeros((1, 1), complex), (1, 1): np.zeros((1, 1), complex)})
    return Block(sec, H, {site: {"c": c, "S": S, "n": n}}, [site])


def c_tilde(blk, c):
    """Annihilation tensor from the creation tensor (adjoint with phase)."""
    out = Op(1, -1)
    for (b, k), m in c.blocks.items():
        j_b, j_k = blk.sectors[b][1], blk.sectors[k][1]
        # c~ : ket = b (spin j_b), bra = k ; phase (-1)^(j_b - j_k + 1/2)
        out.blocks[(k, b)] = _par(j_b - j_k + 1) * m.conj().T
    return out


# ---------------------------------------------------------------------
# bookkeeping: Get_basis / Get_super_basis (block.f90:2087-2245)

def sort_qn_desc(qns):
    """QuickSort_Integer (block.f90:2316-2348): 2S descending, then N
    descending; restated as the same two exchange passes."""
    arr = [list(q) for q in qns]
    n = len(arr)
    for i in range(n):
        for x in range(i + 1, n):
            if arr[i][1] < arr[x][1]:
                arr[i], arr[x] = arr[x], arr[i]
    for i in range(n):
        for x in range(i + 1, n):
            if arr[i][1] == arr[x][1] and arr[i][0] < arr[x][0]:
                arr[i], arr[x] = arr[x], arr[i]
    return [tuple(a) for a in arr]


def get_basis(block_sectors, site_sectors):
    """Enlarged sectors [(N, 2J, dim, [(a, s, spos, sdim)])] in the
    reference's order: unique (N, 2J) sorted 2S desc then N desc; the
    sub-blocks of each in (block sector, site sector) loop order."""
    pre = []
    for (bn, bs, _) in block_sectors:
        for (sn, ss, _) in site_sectors:
            for j in range(abs(bs - ss), bs + ss + 1, 2):
                pre.append((bn + sn, j))
    srt = sort_qn_desc(pre)
    uniq = []
    for q in srt:
        if not uniq or uniq[-1] != q:
            uniq.append(q)
    out = []
    for (N, J) in uniq:
        subs, pos = [], 0
        for a, (bn, bs, bd) in enumerate(block_sectors):
            for s, (sn, ss, sd) in enumerate(site_sectors):
                if bn + sn == N and abs(bs - ss) <= J <= bs + ss:
                    subs.append((a, s, pos, bd * sd))
                    pos += bd * sd
        out.append((N, J, pos, subs))
    return out


def get_super_basis(sys_bs, env_bs, Ntot, S2tot):
    out = []
    for x, (nx, jx, dx, _) in enumerate(sys_bs):
        for y, (ny, jy, dy, _) in enumerate(env_bs):
            if ny == Ntot - nx and abs(jx - jy) <= S2tot <= jx + jy:
                out.append((x, y, dx, dy))
    return out


# ---------------------------------------------------------------------
# embedding into the enlarged basis (Edmonds 7.1.7 / 7.1.8 / 7.1.6)

def embed_part1(T, blk, site, basis):
    """T acting on the block (part 1) of block (x) site, enlarged basis."""
    out = Op(T.k2, T.dN)
    k2 = T.k2
    for bi, (Nb, Jb, db, subs_b) in enumerate(basis):
        for ki, (Nk, Jk, dk, subs_k) in enumerate(basis):
            if Nb != Nk + T.dN:
                continue
            M = None
            for (a2, s2, p2, d2) in subs_b:
                for (a1, s1, p1, d1) in subs_k:
                    if s2 != s1 or (a2, a1) not in T.blocks:
                        continue
                    ja2, ja1, js = blk.sectors[a2][1], blk.sectors[a1][1], site.sectors[s1][1]
                    f = _par(ja2 + js + Jk + k2) * np.sqrt((Jk + 1) * (Jb + 1)) * w6j(ja2, Jb, js, Jk, ja1, k2)
                    if f == 0.0:
                        continue
                    if M is None:
                        M = np.zeros((db, dk), complex)
                    M[p2:p2 + d2, p1:p1 + d1] += f * T.blocks[(a2, a1)]
            if M is not None:
                out.blocks[(bi, ki)] = M
    return out


def embed_part2(U, blk, site, basis):
    """U acting on the site (part 2) of block (x) site; fermion sign
    (-1)^N_block for odd U."""
    out = Op(U.k2, U.dN)
    k2 = U.k2
    for bi, (Nb, Jb, db, subs_b) in enumerate(basis):
        for ki, (Nk, Jk, dk, subs_k) in enumerate(basis):
            if Nb != Nk + U.dN:
                continue
            M = None
            for (a2, s2, p2, d2) in subs_b:
                for (a1, s1, p1, d1) in subs_k:
                    if a2 != a1 or (s2, s1) not in U.blocks:
                        continue
                    ja, js2, js1 = blk.sectors[a1][1], site.sectors[s2][1], site.sectors[s1][1]
                    f = _par(ja + js1 + Jb + k2) * np.sqrt((Jk + 1) * (Jb + 1)) * w6j(js2, Jb, ja, Jk, js1, k2)
                    if U.odd and blk.sectors[a1][0] % 2:
                        f = -f
                    if f == 0.0:
                        continue
                    if M is None:
                        M = np.zeros((db, dk), complex)
                    M[p2:p2 + d2, p1:p1 + d1] += f * U.blocks[(s2, s1)][0, 0] * np.eye(d1)
            if M is not None:
                out.blocks[(bi, ki)] = M
    return out


def scalar_pair(T, U, blk, site, basis):
    """[T(block) x U(site)]^0 in the enlarged basis (reduced elements)."""
    assert T.k2 == U.k2
    k2 = T.k2
    out = Op(0, T.dN + U.dN)
    for bi, (Nb, Jb, db, subs_b) in enumerate(basis):
        for ki, (Nk, Jk, dk, subs_k) in enumerate(basis):
            if Jb != Jk or Nb != Nk + T.dN + U.dN:
                continue
            M = None
            for (a2, s2, p2, d2) in subs_b:
                for (a1, s1, p1, d1) in subs_k:
                    if (a2, a1) not in T.blocks or (s2, s1) not in U.blocks:
                        continue
                    ja2, ja1 = blk.sectors[a2][1], blk.sectors[a1][1]
                    js2, js1 = site.sectors[s2][1], site.sectors[s1][1]
                    f = np.sqrt(Jk + 1) * _par(ja1 + js2 + k2 + Jk) / np.sqrt(k2 + 1) * w6j(ja2, ja1, k2, js1, js2, Jk)
                    if U.odd and blk.sectors[a1][0] % 2:
                        f = -f
                    if f == 0.0:
                        continue
                    if M is None:
                        M = np.zeros((db, dk), complex)
                    M[p2:p2 + d2, p1:p1 + d1] += f * U.blocks[(s2, s1)][0, 0] * T.blocks[(a2, a1)]
            if M is not None:
                out.blocks[(bi, ki)] = M
    return out


def bond_pair_ops(blk_P, ops_P, blk_Q, ops_Q, t, J):
    """The four (coef, T on P, U on Q) products of one bond, P before Q."""
    terms = []
    if t != 0.0:
        terms.append((t * SQ2, ops_P["c"], c_tilde(blk_Q, ops_Q["c"])))
        terms.append((t * SQ2, c_tilde(blk_P, ops_P["c"]), ops_Q["c"]))
    if J != 0.0:
        terms.append((-np.sqrt(3.0) * J, ops_P["S"], ops_Q["S"]))
        terms.append((-0.25 * J, ops_P["n"], ops_Q["n"]))
    return terms


def enlarge(blk, new_site, bonds_to_new, keep_sites):
    """Block (x) site with its Hamiltonian (block H + bonds block-site) and
    the operators of keep_sites embedded. bonds_to_new: [(i, t, J)] with i
    a site of blk."""
    site = site_block(new_site)
    basis = get_basis(blk.sectors, site.sectors)
    sectors = [(N, J, d) for (N, J, d, _) in basis]
    H = embed_part1(blk.H, blk, site, basis)
    for (i, t, J) in bonds_to_new:
        for coef, T, U in bond_pair_ops(blk, blk.ops[i], site, site.ops[new_site], t, J):
            H.add(scalar_pair(T, U, blk, site, basis), coef)
    ops = {}
    for i in keep_sites:
        if i == new_site:
            ops[i] = {k: embed_part2(v, blk, site, basis) for k, v in site.ops[new_site].items()}
        else:
            ops[i] = {k: embed_part1(v, blk, site, basis) for k, v in blk.ops[i].items()}
    eb = Block(sectors, H, ops, blk.sites + [new_site])
    eb.basis = basis
    return eb


def empty_block():
    return Block([(0, 0, 1)], Op(0, 0, {(0, 0): np.zeros((1, 1), complex)}), {}, [])


def truncate(eb, U_by_sector):
    """U^H O U for H and every kept operator (update_trun_dia/ndia);
    U_by_sector[i] = (d_i x kept_i) matrix or None (sector dropped)."""
    keep = [i for i, u in enumerate(U_by_sector) if u is not None and u.shape[1] > 0]
    new_index = {i: n for n, i in enumerate(keep)}
    sectors = [(eb.sectors[i][0], eb.sectors[i][1], U_by_sector[i].shape[1]) for i in keep]

    def tr(op):
        out = Op(op.k2, op.dN)
        for (b, k), m in op.blocks.items():
            if b in new_index and k in new_index:
                out.blocks[(new_index[b], new_index[k])] = U_by_sector[b].conj().T @ m @ U_by_sector[k]
        return out

    ops = {i: {k: tr(v) for k, v in d.items()} for i, d in eb.ops.items()}
    return Block(sectors, tr(eb.H), ops, eb.sites)


# ---------------------------------------------------------------------
# superblock H.psi (model_wave, get_sys.f90:6915-7393) on enlarged blocks

class Superblock:
    """L = enlarged sys, R = enlarged env; psi sectors (x, y) over the
    super basis; bonds_LR = [(i in L sites, j in R sites, t, J)]."""

    def __init__(self, L, R, Ntot, S2tot, bonds_LR):
        self.L, self.R = L, R
        self.S2 = S2tot
        self.sb = [(x, y, L.sectors[x][2], R.sectors[y][2]) for x, (nx, jx, dx) in enumerate(L.sectors)
                   for y, (ny, jy, dy) in enumerate(R.sectors)
                   if ny == Ntot - nx and abs(jx - jy) <= S2tot <= jx + jy]
        self.pos = {}
        off = 0
        for idx, (x, y, dx, dy) in enumerate(self.sb):
            self.pos[(x, y)] = (idx, off, dx, dy)
            off += dx * dy
        self.n = off
        self.terms = []
        for (i, j, t, J) in bonds_LR:
            self.terms.extend(bond_pair_ops(L, L.ops[i], R, R.ops[j], t, J))

    def sector(self, v, x, y):
        _, o, dx, dy = self.pos[(x, y)]
        return v[o:o + dx * dy].reshape((dx, dy), order="F")

    def matvec(self, v):
        out = np.zeros(self.n, complex)
        L, R, S2 = self.L, self.R, self.S2

        def add(x, y, m):
            _, o, dx, dy = self.pos[(x, y)]
            out[o:o + dx * dy] += m.ravel(order="F")

        for (x, y, dx, dy) in self.sb:
            psi = self.sector(v, x, y)
            jx, jy = L.sectors[x][1], R.sectors[y][1]
            acc = np.zeros((dx, dy), complex)
            if (x, x) in L.H.blocks:
                acc += L.H.blocks[(x, x)] @ psi / np.sqrt(jx + 1)
            if (y, y) in R.H.blocks:
                acc += psi @ R.H.blocks[(y, y)].T / np.sqrt(jy + 1)
            add(x, y, acc)
        for coef, T, U in self.terms:
            k2 = T.k2
            for (x2, x1), mt in T.blocks.items():
                for (y2, y1), mu in U.blocks.items():
                    if (x1, y1) not in self.pos or (x2, y2) not in self.pos:
                        continue
                    jx2, jx1 = L.sectors[x2][1], L.sectors[x1][1]
                    jy2, jy1 = R.sectors[y2][1], R.sectors[y1][1]
                    f = _par(jx1 + jy2 + k2 + S2) / np.sqrt(k2 + 1) * w6j(jx2, jx1, k2, jy1, jy2, S2)
                    if U.odd and L.sectors[x1][0] % 2:
                        f = -f
                    if f == 0.0:
                        continue
                    add(x2, y2, coef * f * (mt @ self.sector(v, x1, y1) @ mu.T))
        return out
