"""Untruncated SU(2) superblock built with the dense block algebra: the
two-site DMRG at a kept dimension above the full multiplet count is exact,
so its ground energy must equal exact diagonalisation (oracle pin)."""
import numpy as np

from . import su2_dense as su


def build_chain_blocks(N, bond_list, l):
    """Sys block of sites 1..l and env block of sites N..l+3 (fermion order
    of the env block: N, N-1, ..., l+3), untruncated, with operators of all
    sites that still have bonds outside the block."""
    nb = {}
    for (a, b, t, J) in bond_list:
        nb.setdefault(a, []).append((b, t, J))
        nb.setdefault(b, []).append((a, t, J))

    def grow(order):
        blk = su.empty_block()
        inside = set()
        for s in order:
            to_new = [(i, t, J) for (i, t, J) in nb.get(s, []) if i in inside]
            inside.add(s)
            keep = [i for i in inside if any(j not in inside for (j, _, _) in nb.get(i, []))]
            blk = su.enlarge(blk, s, to_new, keep)
        return blk, inside

    sys_blk, sys_in = grow(list(range(1, l + 1)))
    env_blk, env_in = grow(list(range(N, l + 2, -1)))
    return sys_blk, env_blk, nb


def superblock(N, Ne, S2, bond_list, l):
    sys_blk, env_blk, nb = build_chain_blocks(N, bond_list, l)
    Lsites = set(range(1, l + 2))
    Rsites = set(range(l + 2, N + 1))

    def keep_for(block_sites, other):
        return [i for i in block_sites if any(j in other for (j, _, _) in nb.get(i, []))]

    L = su.enlarge(sys_blk, l + 1, [(i, t, J) for (i, t, J) in nb.get(l + 1, []) if i < l + 1],
                   keep_for(Lsites, Rsites))
    R = su.enlarge(env_blk, l + 2, [(i, t, J) for (i, t, J) in nb.get(l + 2, []) if i > l + 2],
                   keep_for(Rsites, Lsites))
    bonds_LR = [(a, b, t, J) for (a, b, t, J) in bond_list if a in Lsites and b in Rsites]
    return su.Superblock(L, R, Ne, S2, bonds_LR)


def ground_energy(N, Ne, S2, bond_list, l=None):
    l = (N - 2) // 2 if l is None else l
    sb = superblock(N, Ne, S2, bond_list, l)
    eye = np.eye(sb.n, dtype=complex)
    H = np.stack([sb.matvec(eye[:, i]) for i in range(sb.n)], axis=1)
    return float(np.linalg.eigvalsh((H + H.conj().T) / 2)[0]), H
