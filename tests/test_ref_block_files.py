"""The block, truncation and wave-function files hold the reference's numbers.

The reference's files are defined by its conventions (sector order of
Get_basis, SU(2) coupling of block and site, fermion order, the
Wigner-Eckart form of reduced elements, the operator slots of
dmrg_setup0); oracle/ref_blocks.py turns those conventions into explicit
Fock-space states, built level by level from the stored truncation
matrices. A file holds the reference's numbers when each stored block
matrix is the reduced matrix of the physical operator between those
states: the block Hamiltonian (sum of the block's bonds), c+ (elec_up),
c~ (elec_down), S (spin_sd) and n (num_sn) of every slot site, on both
sides, and, for eigvec1, a Fock vector that is the superblock's ground
state. The checks are exact Fock-space algebra, independent of this
build's transition generator; a missing env sign (-1)^(N_block N_site), a
missing conjugation, a wrong elec_down phase, a slot out of order or a
scalar stored without its sqrt(2S+1) each breaks one of them.

Ny = 2 cylinders are avoided: there a pair can be both a first and a
second neighbour, which the oracle's lattice and the product's assign
differently.
"""
import os

import numpy as np
import pytest

pytest.importorskip("sympy")
import fake_device  # noqa: E402
from oracle import ref_blocks as RB  # noqa: E402
from oracle import tj_model  # noqa: E402
from tsc_tri_t1t2j1j2_amd import lattice, persist  # noqa: E402
from tsc_tri_t1t2j1j2_amd.blocks import SITE_SECTORS, site_operators  # noqa: E402
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG  # noqa: E402


@pytest.mark.parametrize("Lx,Ny,t2,J2", [(4, 3, 0.5, 0.1), (6, 4, 0.5, 0.1), (8, 6, 0.5, 0.1), (6, 4, 0.0, 0.0),
                                          (5, 6, 0.0, 0.0)])
def test_slot_table_is_dmrg_setup0(Lx, Ny, t2, J2):
    bl = lattice.bonds(Lx, Ny, 3.0, t2, 1.0, J2)
    ref = RB.inb1(Lx, Ny, t2, J2)
    N = Lx * Ny
    for ns in range(1, N - 1):
        assert lattice.op_slots(bl, ns) == ref[ns], ns
    # the env side reuses the table through i -> N + 1 - i: the cylinder is
    # symmetric under that map
    assert sorted((N + 1 - b, N + 1 - a) for (a, b, _t, _J) in bl) == sorted((a, b) for (a, b, _t, _J) in bl)
    # the slots are the sites with bonds leaving the block plus those whose
    # last neighbour is the block's last site
    reach = {}
    for (a, b, _t, _J) in bl:
        reach[a] = max(reach.get(a, 0), b)
    for ns in range(2, N - 1):
        out = {i for i in range(1, ns + 1) if reach.get(i, 0) > ns}
        assert set(ref[ns]) == out | {i for i in range(1, ns) if reach.get(i, 0) == ns}


def test_elec_down_of_the_site_is_st_elec_down():
    # st_elec_down (dmrg_sub.f90:779-791): up_dif -1, ket electron, bra hole, -sqrt2
    c = site_operators()["c"]
    d = persist.elec_down(c, list(SITE_SECTORS))
    assert (d.up_dif, d.k2) == (-1, 1) and list(d.blocks) == [(1, 0)]
    assert d.blocks[(1, 0)][0, 0] == pytest.approx(-np.sqrt(2.0))


def _raw_total_block(r):
    """A Total_Block record group as written: header and [(N, 2S, dd, matrix)]."""
    head = [r.int1() for _ in range(5)]
    subs = []
    for _ in range(head[1]):
        r.empty()
        n, j2, dd = r.ints(3)
        rows, cols = r.ints(2)
        subs.append((n, j2, dd, r.matrix(rows, cols)))
    return head, subs


def _as_blocks(subs, states, up_dif):
    out = {}
    for (n, j2, dd, m) in subs:
        k = states.index[(n, j2)]
        b = states.index[(n + up_dif, j2 + dd)]
        out[(b, k)] = m
    return out


def _agree(stored, want, scale):
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
    return worst / scale


def _grow_side(path, side, N, upto):
    """{L: BlockStates} of one side's blocks from its truncation files."""
    states = {1: RB.BlockStates.site()}
    for L in range(2, upto + 1):
        eb = states[L - 1].enlarge(site_first=(side == "env"))
        r = persist.RecordReader(os.path.join(path, persist.trun_name(side, L)))
        head, subs = _raw_total_block(r)
        assert head[0] == L
        U = [(eb.index[(n, j2)], m) for (n, j2, _dd, m) in subs]
        states[L] = eb.truncate(U)
    return states


@pytest.fixture(scope="module")
def warm_run(tmp_path_factory):
    """A 4 x 3 warm-up (blocks of up to 6 sites) with truncation from
    3 sites on, writing its files."""
    mp = pytest.MonkeyPatch()
    ctx = fake_device.install(mp)
    path = str(tmp_path_factory.mktemp("ref_files"))
    d = DMRG(ctx, 4, 3, t1=1.0, t2=0.5, J1=0.4, J2=0.2, Ne=10, kept_min=4, kept_max=7, sweeps=1,
             lanzero=1e-10, trun_err=1e-14, native_planner=False)
    d.checkpoint = path
    d.warm_up()
    yield d, path
    mp.undo()


@pytest.mark.parametrize("side", ["sys", "env"])
def test_block_files_are_the_references(warm_run, side):
    d, path = warm_run
    N = d.N
    bl = tj_model.bonds(4, 3, 1.0, 0.5, 0.4, 0.2)
    slots = RB.inb1(4, 3, 0.5, 0.2)
    states = _grow_side(path, side, N, N // 2)
    seen_odd = 0
    for L in range(2, N // 2 + 1):
        st = states[L]
        first = 1 if side == "sys" else N - L + 1
        local = {i: i - first for i in range(first, first + L)}
        F, ops = RB.fock_ops(L)
        r = persist.RecordReader(os.path.join(path, persist.block_name(side, L)))
        assert r.int1() == L
        head, hsubs = _raw_total_block(r)
        assert head[0] == L and head[3:] == [0, 0]
        assert [(n, j2, m.shape[0]) for (n, j2, _dd, m) in hsubs] == st.sectors
        inner = [(local[a], local[b], t, J) for (a, b, t, J) in bl if a in local and b in local]
        want, tens = RB.reduced(st, st, 0, RB.block_hamiltonian(F, inner), 0)
        assert tens < 1e-10
        assert _agree(_as_blocks(hsubs, st, 0), want, 1.0) < 1e-10, ("ham", L)
        for II in slots[L]:
            site = II if side == "sys" else N + 1 - II
            loc = local[site]
            for name, (k, dN) in (("c+", (1, 1)), ("c~", (1, -1)), ("S", (2, 0)), ("n", (0, 0))):
                head, subs = _raw_total_block(r)
                assert head[0] == L and head[3:] == [dN, k], (name, head)
                want, tens = RB.reduced(st, st, k, ops[loc][name][1], dN)
                assert tens < 1e-10, (name, L)
                assert _agree(_as_blocks(subs, st, dN), want, 1.0) < 1e-10, (side, L, site, name)
        assert r.pos == len(r.data)
        # the enlarged env basis has product states with an odd block and an
        # odd site, where the env sign acts
        if side == "env" and L >= 3:
            eb = states[L - 1].enlarge(site_first=True)
            seen_odd += sum(1 for (_N, _J, _d, subs) in eb.basis for (a, s, _p, _sd) in subs
                            if states[L - 1].sectors[a][0] % 2 and s == 0)
    if side == "env":
        assert seen_odd > 0


def test_reading_the_files_back_gives_the_run_blocks(warm_run):
    d, path = warm_run
    from tsc_tri_t1t2j1j2_amd.resident import model_to_host
    for side, blocks in (("sys", d.sys_blocks), ("env", d.env_blocks)):
        for L in range(2, d.N // 2 + 1):
            sites = list(range(1, L + 1)) if side == "sys" else list(range(d.N, d.N - L, -1))
            m = persist.model_from_disk(os.path.join(path, persist.block_name(side, L)), side, sites,
                                        d.slot_sites(side, L))
            mine = model_to_host(blocks[L])
            assert m.sectors == mine.sectors and m.sites == mine.sites
            assert set(m.ops) == set(mine.ops)
            for i in m.ops:
                for name in ("c", "S", "n"):
                    for key, v in mine.ops[i][name].blocks.items():
                        assert np.array_equal(m.ops[i][name].blocks[key], v)
            for key, v in mine.ham.blocks.items():
                assert np.array_equal(m.ham.blocks[key], v)


def test_wave_file_is_the_ground_state():
    """eigvec1 of the last warm-up step of a 2 x 3 cylinder (the superblock
    is the whole lattice, nothing truncated): the stored sectors, read in
    the reference's conventions, are the Fock-space ground state."""
    mp = pytest.MonkeyPatch()
    try:
        ctx = fake_device.install(mp)
        import tempfile
        with tempfile.TemporaryDirectory() as path:
            d = DMRG(ctx, 2, 3, t1=1.0, t2=0.5, J1=0.4, J2=0.2, Ne=4, kept_min=4, kept_max=40, sweeps=1,
                     lanzero=1e-12, trun_err=1e-14, native_planner=False)
            d.checkpoint = path
            d.warm_up()
            E = d.history[-1][2]
            r = persist.RecordReader(os.path.join(path, persist.wave_name(1)))
            sys_len, env_len, num, _dim = (r.int1() for _ in range(4))
            assert (sys_len, env_len) == (3, 3)
            sys_st = _grow_side(path, "sys", 6, sys_len - 1)[sys_len - 1].enlarge(site_first=False)
            env_st = _grow_side(path, "env", 6, env_len - 1)[env_len - 1].enlarge(site_first=True)
            subs = []
            for _ in range(num):
                r.empty()
                sq, eq = tuple(r.ints(2)), tuple(r.ints(2))
                dx, dy = r.ints(2)
                subs.append((sys_st.index[sq], env_st.index[eq], r.matrix(dx, dy)))
    finally:
        mp.undo()
    psi = RB.super_state(sys_st, env_st, subs)
    F = RB.R.Fock(6)
    bl = [(a - 1, b - 1, t, J) for (a, b, t, J) in tj_model.bonds(2, 3, 1.0, 0.5, 0.4, 0.2)]
    H = RB.block_hamiltonian(F, bl)[0]
    assert np.vdot(psi, psi).real == pytest.approx(1.0, abs=1e-10)
    assert np.vdot(psi, H @ psi).real == pytest.approx(E, abs=1e-9)
    assert np.linalg.norm(H @ psi - E * psi) < 1e-6
    # read without the env sign (as if the file kept this build's own fermion
    # order) the stored vector is no eigenvector
    flipped = [(x, y, m * _odd_odd_sign(env_st, y)[None, :]) for (x, y, m) in subs]
    bad = RB.super_state(sys_st, env_st, flipped)
    assert np.linalg.norm(H @ bad - E * bad) > 1e-3


def _odd_odd_sign(enlarged, y):
    """-1 on the rows of enlarged sector y whose block part holds an odd
    number of electrons and whose site holds one."""
    _N, _J, d, subs = enlarged.basis[y]
    v = np.ones(d)
    for (a, s, p, sd) in subs:
        if enlarged.block_sectors[a][0] % 2 and SITE_SECTORS[s][0] == 1:
            v[p:p + sd] = -1.0
    return v
