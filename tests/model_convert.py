"""Test helper: the oracle's dense SU(2) blocks as the product's TotalModel
(same sector order and reduced matrices, different container)."""
from oracle import su2_dense as su
from tsc_tri_t1t2j1j2_amd.blocks import TotalBlock, TotalModel


def op_to_product(op):
    return TotalBlock(op.dN, op.k2, {k: v.copy() for k, v in op.blocks.items()})


def block_to_product(blk):
    ops = {i: {name: op_to_product(o) for name, o in d.items()} for i, d in blk.ops.items()}
    return TotalModel(blk.sites, blk.sectors, op_to_product(blk.H), ops)


def oracle_superblock(sys_blk, env_blk, s1, s2, bond_list, ntot, s2tot):
    """The oracle's superblock for the same blocks and centre sites."""
    nb = {}
    for (a, b, t, J) in bond_list:
        nb.setdefault(a, []).append((b, t, J))
        nb.setdefault(b, []).append((a, t, J))
    Lsites = set(sys_blk.sites) | {s1}
    Rsites = set(env_blk.sites) | {s2}

    def keep_for(block_sites, other):
        return [i for i in block_sites if any(j in other for (j, _, _) in nb.get(i, []))]

    L = su.enlarge(sys_blk, s1, [(i, t, J) for (i, t, J) in nb.get(s1, []) if i in sys_blk.sites],
                   keep_for(Lsites, Rsites))
    R = su.enlarge(env_blk, s2, [(i, t, J) for (i, t, J) in nb.get(s2, []) if i in env_blk.sites],
                   keep_for(Rsites, Lsites))
    bonds_LR = []
    for (a, b, t, J) in bond_list:
        if a in Lsites and b in Rsites:
            bonds_LR.append((a, b, t, J))
        elif b in Lsites and a in Rsites:
            bonds_LR.append((b, a, t, J))
    return su.Superblock(L, R, ntot, s2tot, bonds_LR)
