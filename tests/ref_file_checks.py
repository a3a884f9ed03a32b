"""Checks of stored DMRG files against the reference's conventions, made
explicit as Fock-space block states (oracle/ref_blocks.py). Shared by the
CPU tests (driver on a host stand-in of the device) and the GPU tests
(resident device path)."""
import os

import numpy as np

from oracle import ref_blocks as RB
from oracle import tj_model
from tsc_tri_t1t2j1j2_amd import persist
from tsc_tri_t1t2j1j2_amd.blocks import SITE_SECTORS


def raw_total_block(r):
    """A Total_Block record group as written: header and [(N, 2S, dd, matrix)]."""
    head = [r.int1() for _ in range(5)]
    subs = []
    for _ in range(head[1]):
        r.empty()
        n, j2, dd = r.ints(3)
        rows, cols = r.ints(2)
        subs.append((n, j2, dd, r.matrix(rows, cols)))
    return head, subs


def as_blocks(subs, states, up_dif):
    out = {}
    for (n, j2, dd, m) in subs:
        k = states.index[(n, j2)]
        b = states.index[(n + up_dif, j2 + dd)]
        out[(b, k)] = m
    return out


def agree(stored, want):
    keys = set(stored) | set(want)
    worst = 0.0
    for key in keys:
        a = stored.get(key)
        b = want.get(key)
        if a is None:
            a = np.zeros_like(b)
        if b is None:
            b = np.zeros_like(a)
        worst = max(worst, float(np.abs(a - b).max()) if a.size else 0.0)
    return worst


def grow_side(path, side, upto):
    """{L: BlockStates} of one side's blocks from its truncation files."""
    states = {1: RB.BlockStates.site()}
    for L in range(2, upto + 1):
        eb = states[L - 1].enlarge(site_first=(side == "env"))
        r = persist.RecordReader(os.path.join(path, persist.trun_name(side, L)))
        head, subs = raw_total_block(r)
        assert head[0] == L
        # the reference's env block states are the conjugated columns of its
        # env truncation (wave_trans_LR applies it conjugate-transposed,
        # wave_trans_RL plainly; its kernels conjugate env operators)
        U = [(eb.index[(n, j2)], np.conj(m) if side == "env" else m) for (n, j2, _dd, m) in subs]
        states[L] = eb.truncate(U)
    return states


def check_block_files(path, side, Lx, Ny, couplings, upto):
    """Every block file of one side, lengths 2..upto, against the reduced
    matrices of the physical operators between the reference's block
    states rebuilt from the truncation files; returns the number of
    enlarged env sub-blocks with an odd block and an odd site seen."""
    t1, t2, J1, J2 = couplings
    N = Lx * Ny
    bl = tj_model.bonds(Lx, Ny, t1, t2, J1, J2)
    slots = RB.inb1(Lx, Ny, t2, J2)
    states = grow_side(path, side, upto)
    seen_odd = 0
    for L in range(2, upto + 1):
        st = states[L]
        first = 1 if side == "sys" else N - L + 1
        local = {i: i - first for i in range(first, first + L)}
        F, ops = RB.fock_ops(L)
        r = persist.RecordReader(os.path.join(path, persist.block_name(side, L)))
        assert r.int1() == L
        head, hsubs = raw_total_block(r)
        assert head[0] == L and head[3:] == [0, 0]
        assert [(n, j2, m.shape[0]) for (n, j2, _dd, m) in hsubs] == st.sectors
        inner = [(local[a], local[b], t, J) for (a, b, t, J) in bl if a in local and b in local]
        want, tens = RB.reduced(st, st, 0, RB.block_hamiltonian(F, inner), 0)
        assert tens < 1e-10
        # env operators are stored as the complex conjugates of the physical
        # reduced matrices (the kernels apply them conjugate-transposed)
        conj = (lambda d: {k: np.conj(v) for k, v in d.items()}) if side == "env" else (lambda d: d)
        assert agree(as_blocks(hsubs, st, 0), conj(want)) < 1e-10, ("ham", side, L)
        for II in slots[L]:
            site = II if side == "sys" else N + 1 - II
            loc = local[site]
            for name, (k, dN) in (("c+", (1, 1)), ("c~", (1, -1)), ("S", (2, 0)), ("n", (0, 0))):
                head, subs = raw_total_block(r)
                assert head[0] == L and head[3:] == [dN, k], (name, head)
                want, tens = RB.reduced(st, st, k, ops[loc][name][1], dN)
                assert tens < 1e-10, (name, L)
                assert agree(as_blocks(subs, st, dN), conj(want)) < 1e-10, (side, L, site, name)
        assert r.pos == len(r.data)
        if side == "env" and L >= 3:
            eb = states[L - 1].enlarge(site_first=True)
            seen_odd += sum(1 for (_N, _J, _d, subs) in eb.basis for (a, s, _p, _sd) in subs
                            if states[L - 1].sectors[a][0] % 2 and SITE_SECTORS[s][0] == 1)
    return seen_odd


def wave_check(path, N, Lx, Ny, couplings, E):
    """eigvec1 read in the reference's conventions as a Fock vector over the
    whole lattice (the superblock must be the whole lattice): returns
    (norm^2, <H>, |H psi - E psi|, the same residual with the env sign
    dropped)."""
    t1, t2, J1, J2 = couplings
    r = persist.RecordReader(os.path.join(path, persist.wave_name(1)))
    sys_len, env_len, num, _dim = (r.int1() for _ in range(4))
    assert sys_len + env_len == N
    sys_st = grow_side(path, "sys", sys_len - 1)[sys_len - 1].enlarge(site_first=False)
    env_st = grow_side(path, "env", env_len - 1)[env_len - 1].enlarge(site_first=True)
    subs = []
    for _ in range(num):
        r.empty()
        sq, eq = tuple(r.ints(2)), tuple(r.ints(2))
        dx, dy = r.ints(2)
        subs.append((sys_st.index[sq], env_st.index[eq], r.matrix(dx, dy)))
    psi = RB.super_state(sys_st, env_st, subs)
    F = RB.R.Fock(N)
    bl = [(a - 1, b - 1, t, J) for (a, b, t, J) in tj_model.bonds(Lx, Ny, t1, t2, J1, J2)]
    H = RB.block_hamiltonian(F, bl)[0]
    flipped = [(x, y, m * odd_odd_sign(env_st, y)[None, :]) for (x, y, m) in subs]
    bad = RB.super_state(sys_st, env_st, flipped)
    return (np.vdot(psi, psi).real, np.vdot(psi, H @ psi).real, np.linalg.norm(H @ psi - E * psi),
            np.linalg.norm(H @ bad - E * bad))


def odd_odd_sign(enlarged, y):
    """-1 on the rows of enlarged sector y whose block part holds an odd
    number of electrons and whose site holds one."""
    _N, _J, d, subs = enlarged.basis[y]
    v = np.ones(d)
    for (a, s, p, sd) in subs:
        if enlarged.block_sectors[a][0] % 2 and SITE_SECTORS[s][0] == 1:
            v[p:p + sd] = -1.0
    return v
