"""Sector-blocked data of the DMRG hot path, host side.

Mirrors the derived types of hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/common.f90:84-174 (Total_Basis, Super_Basis, Wavefunction,
Total_Block, Total_Model) and the bookkeeping of DMRG_code/block.f90:
2087-2245 (Get_basis, Get_super_basis, Get_Unique_QN, QuickSort_Integer),
which is reproduced exactly: sector order and sub-block offsets decide the
memory layout of psi and of every operator.
"""
import numpy as np

SITE_SECTORS = ((1, 1, 1), (0, 0, 1))   # electron (N=1, 2S=1) first, then hole


class TotalBlock:
    """A tensor operator on a block: particle change up_dif, rank 2k = k2,
    blocks[(bra, ket)] = reduced matrix (Fortran order) between sector
    indices of the block it acts on."""

    __slots__ = ("up_dif", "k2", "blocks")

    def __init__(self, up_dif, k2, blocks=None):
        self.up_dif = int(up_dif)
        self.k2 = int(k2)
        self.blocks = {} if blocks is None else blocks

    @property
    def odd(self):
        return self.up_dif % 2 != 0

    def copy(self):
        return TotalBlock(self.up_dif, self.k2, {k: v.copy() for k, v in self.blocks.items()})

    def axpy(self, c, other):
        """self += c * other (block_add_block_two, block.f90:177-215)."""
        assert other.up_dif == self.up_dif and other.k2 == self.k2
        for k, v in other.blocks.items():
            if k in self.blocks:
                self.blocks[k] = self.blocks[k] + c * v
            else:
                self.blocks[k] = c * v
        return self


class TotalModel:
    """Block Hamiltonian plus the operators (c+, S, n) of the sites that
    still have bonds outside the block (Total_Model%ham / %sub)."""

    def __init__(self, sites, sectors, ham, ops):
        self.sites = list(sites)            # fermion order inside the block
        self.sectors = [tuple(int(v) for v in s) for s in sectors]   # (N, 2S, dim)
        self.ham = ham
        self.ops = ops                       # {site: {"c": TotalBlock, "S": ..., "n": ...}}
        self.index = {(s[0], s[1]): i for i, s in enumerate(self.sectors)}

    @property
    def len(self):
        return len(self.sites)

    @property
    def dim(self):
        return sum(s[2] for s in self.sectors)


def empty_model():
    return TotalModel([], [(0, 0, 1)], TotalBlock(0, 0, {(0, 0): np.zeros((1, 1), complex)}), {})


def site_operators():
    """Reduced site operators of the t-J site (Get_site_operator_su2,
    dmrg_sub.f90:733-908): c+ = -sqrt(2) hole->electron, S = sqrt(3/2),
    n = sqrt(2) on the electron. Indices refer to SITE_SECTORS."""
    c = TotalBlock(1, 1, {(0, 1): np.array([[-np.sqrt(2.0)]], complex)})
    S = TotalBlock(0, 2, {(0, 0): np.array([[np.sqrt(1.5)]], complex)})
    n = TotalBlock(0, 0, {(0, 0): np.array([[np.sqrt(2.0)]], complex)})
    return {"c": c, "S": S, "n": n}


def sort_qn(pairs):
    """QuickSort_Integer (block.f90:2316-2348): by 2S descending, then N
    descending among equal 2S, as two exchange passes."""
    a = [list(p) for p in pairs]
    n = len(a)
    for i in range(n):
        for x in range(i + 1, n):
            if a[i][1] < a[x][1]:
                a[i], a[x] = a[x], a[i]
    for i in range(n):
        for x in range(i + 1, n):
            if a[i][1] == a[x][1] and a[i][0] < a[x][0]:
                a[i], a[x] = a[x], a[i]
    return [tuple(p) for p in a]


def unique_qn(sorted_pairs):
    """Get_Unique_QN (block.f90:2249-2280) on a sorted list."""
    out = []
    for p in sorted_pairs:
        if not out or out[-1] != p:
            out.append(p)
    return out


class TotalBasis:
    """Enlarged (block x site) basis (Get_basis, block.f90:2087-2189).
    sectors[x] = (N, 2J, dim, subs) with subs[(a, s, spos, sdim)] in
    (block sector, site sector) loop order."""

    def __init__(self, block_sectors, site_sectors=SITE_SECTORS):
        self.block_sectors = [tuple(b) for b in block_sectors]
        self.site_sectors = [tuple(s) for s in site_sectors]
        pre = []
        for (bn, bs, _) in self.block_sectors:
            for (sn, ss, _) in self.site_sectors:
                for j in range(abs(bs - ss), bs + ss + 1, 2):
                    pre.append((bn + sn, j))
        qns = unique_qn(sort_qn(pre))
        self.sectors = []
        for (N, J) in qns:
            subs, pos = [], 0
            for a, (bn, bs, bd) in enumerate(self.block_sectors):
                for s, (sn, ss, sd) in enumerate(self.site_sectors):
                    if bn + sn == N and abs(bs - ss) <= J <= bs + ss:
                        subs.append((a, s, pos, bd * sd))
                        pos += bd * sd
            self.sectors.append((N, J, pos, subs))
        self.index = {(N, J): x for x, (N, J, _, _) in enumerate(self.sectors)}
        # (a, s) -> [(x, spos)] : every enlarged sector holding that sub-block
        self.where = {}
        for x, (_, _, _, subs) in enumerate(self.sectors):
            for (a, s, p, d) in subs:
                self.where.setdefault((a, s), []).append((x, p))

    @property
    def num(self):
        return len(self.sectors)

    @property
    def dim(self):
        return sum(s[2] for s in self.sectors)

    def plain_sectors(self):
        return [(N, J, d) for (N, J, d, _) in self.sectors]


class SuperBasis:
    """Get_super_basis (block.f90:2194-2245): sys-outer, env-inner pairs
    with N_sys + N_env = N_tot and the triangle (2J_sys, 2J_env, 2S_tot)."""

    def __init__(self, sys_bs, env_bs, ntot, s2tot):
        self.ntot, self.s2tot = int(ntot), int(s2tot)
        self.sub = []
        for x, (nx, jx, dx, _) in enumerate(sys_bs.sectors):
            for y, (ny, jy, dy, _) in enumerate(env_bs.sectors):
                if ny == ntot - nx and abs(jx - jy) <= s2tot <= jx + jy:
                    self.sub.append((x, y, dx, dy))
        self.offset = []
        off = 0
        for (_, _, dx, dy) in self.sub:
            self.offset.append(off)
            off += dx * dy
        self.dim = off
        self.index = {(x, y): i for i, (x, y, _, _) in enumerate(self.sub)}

    @property
    def num(self):
        return len(self.sub)

    def sector(self, vec, i):
        x, y, dx, dy = self.sub[i]
        o = self.offset[i]
        return vec[o:o + dx * dy].reshape((dx, dy), order="F")

    def split(self, vec):
        return [self.sector(vec, i) for i in range(self.num)]
