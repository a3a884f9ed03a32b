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
import ref_file_checks as C  # noqa: E402
from oracle import ref_blocks as RB  # noqa: E402
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
    seen_odd = C.check_block_files(path, side, 4, 3, (1.0, 0.5, 0.4, 0.2), d.N // 2)
    if side == "env":
        # the env sign has something to act on
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


def test_wave_file_is_the_ground_state(tmp_path):
    """eigvec1 of the last warm-up step of a 2 x 3 cylinder (the superblock
    is the whole lattice, nothing truncated): the stored sectors, read in
    the reference's conventions, are the Fock-space ground state; read
    without the env sign (as if the file kept this build's own fermion
    order) they are not an eigenvector."""
    mp = pytest.MonkeyPatch()
    try:
        ctx = fake_device.install(mp)
        d = DMRG(ctx, 2, 3, t1=1.0, t2=0.5, J1=0.4, J2=0.2, Ne=4, kept_min=4, kept_max=40, sweeps=1,
                 lanzero=1e-12, trun_err=1e-14, native_planner=False)
        d.checkpoint = str(tmp_path)
        d.warm_up()
    finally:
        mp.undo()
    E = d.history[-1][2]
    nrm, Eh, res, res_bad = C.wave_check(str(tmp_path), 6, 2, 3, (1.0, 0.5, 0.4, 0.2), E)
    assert nrm == pytest.approx(1.0, abs=1e-10)
    assert Eh == pytest.approx(E, abs=1e-9)
    assert res < 1e-6
    assert res_bad > 1e-3
