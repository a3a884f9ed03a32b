"""The product's sub-block task plans (model_wave, renormalisation)
evaluated on the host against the oracle's dense SU(2) superblock, on
identical blocks. CPU only: the device kernels are checked elsewhere."""
import numpy as np
import pytest

from oracle import dmrg_exact, su2_dense, tj_model
from tests.model_convert import block_to_product, oracle_superblock
from oracle import plan_exec
from tests.plan_eval import apply_numpy
from tsc_tri_t1t2j1j2_amd import hamiltonian as ham
from tsc_tri_t1t2j1j2_amd import native

CASES = [(2, 3, 4, 0, 2), (2, 3, 5, 1, 2), (2, 4, 6, 0, 3), (2, 4, 6, 0, 2), (3, 3, 8, 0, 3), (2, 4, 7, 1, 1),
         (3, 4, 10, 0, 4)]


def _setup(Lx, Ny, Ne, S2, l, t2=0.2, J2=0.016):
    N = Lx * Ny
    b = tj_model.bonds(Lx, Ny, 1.0, t2, 0.4, J2)
    sys_blk, env_blk, _ = dmrg_exact.build_chain_blocks(N, b, l)
    osb = oracle_superblock(sys_blk, env_blk, l + 1, l + 2, b, Ne, S2)
    psb = ham.Superblock(block_to_product(sys_blk), block_to_product(env_blk), l + 1, l + 2, b, Ne, S2)
    return osb, psb


@pytest.mark.parametrize("Lx,Ny,Ne,S2,l", CASES)
def test_model_wave_plan_matches_oracle(Lx, Ny, Ne, S2, l):
    osb, psb = _setup(Lx, Ny, Ne, S2, l)
    assert psb.super.dim == osb.n
    assert [(x, y) for (x, y, _, _) in psb.super.sub] == [(x, y) for (x, y, _, _) in osb.sb]
    rng = np.random.default_rng(Lx * 100 + Ny * 10 + l)
    v = rng.normal(size=osb.n) + 1j * rng.normal(size=osb.n)
    pb = psb.plan_builder()
    got = apply_numpy(pb, v, osb.n)
    want = osb.matvec(v)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


def test_renormalize_plan_matches_oracle():
    Lx, Ny = 3, 4
    N = Lx * Ny
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    sys_blk, _, nb = dmrg_exact.build_chain_blocks(N, b, 5)
    new_site = 6
    to_new = [(i, t, J) for (i, t, J) in nb[new_site] if i in sys_blk.sites]
    keep = [i for i in sys_blk.sites + [new_site] if any(j > new_site for (j, _, _) in nb.get(i, []))]
    eb = su2_dense.enlarge(sys_blk, new_site, to_new, keep)
    rng = np.random.default_rng(3)
    U, order = {}, []
    for x, (Nx, Jx, dx) in enumerate(eb.sectors):
        if x % 3 == 1:
            continue     # drop a few sectors
        k = max(1, dx // 2)
        q = np.linalg.qr(rng.normal(size=(dx, k)) + 1j * rng.normal(size=(dx, k)))[0]
        U[x] = q
        order.append(x)
    order = order[::-1]  # a non-trivial new sector order
    want = su2_dense.truncate(eb, [U.get(x) for x in range(len(eb.sectors))])
    rp = ham.renormalize_plan(block_to_product(sys_blk), new_site, to_new, keep, U, order)
    res = apply_numpy(rp.builder, rp.uin, rp.out_len)
    got = ham.renormalize_result(rp, res)
    # oracle sectors are in enlarged order; map ours back
    pos = {x: n for n, x in enumerate(order)}
    okeep = [x for x in range(len(eb.sectors)) if x in U]
    oidx = {x: n for n, x in enumerate(okeep)}
    assert sorted(got.sectors) == sorted(want.sectors)

    def check(gop, wop):
        for (b2, k2), m in wop.blocks.items():
            gb = gop.blocks.get((pos[okeep[b2]], pos[okeep[k2]]))
            if gb is None:
                assert np.abs(m).max() < 1e-13
            else:
                assert np.allclose(gb, m, atol=1e-12)
        for (b2, k2), m in gop.blocks.items():
            key = (oidx[order[b2]], oidx[order[k2]])
            if key not in wop.blocks:
                assert np.abs(m).max() < 1e-13

    check(got.ham, want.H)
    for i in keep:
        for name in ("c", "S", "n"):
            check(got.ops[i][name], want.ops[i][name])


@pytest.mark.parametrize("order", ["auto", "left", "right"])
@pytest.mark.parametrize("Lx,Ny,Ne,S2,l", CASES)
def test_native_planner_matches_oracle(Lx, Ny, Ne, S2, l, order, monkeypatch):
    """Every contraction order of the two-sided products (O_L psi first,
    psi O_R^T first, or the cheaper per source window) gives the oracle's
    H psi; the planned flops never exceed either fixed order's, and the
    per-product minimum (nothing shared) does not depend on the order."""
    osb, psb = _setup(Lx, Ny, Ne, S2, l)
    monkeypatch.setenv("TSC_PLAN_ORDER", order)
    nt = native.NativeTasks(psb)
    if order == "auto":
        fl = {}
        for o in ("left", "right"):
            monkeypatch.setenv("TSC_PLAN_ORDER", o)
            fl[o] = native.NativeTasks(psb).flops()
        ex, pm = nt.flops()
        assert ex <= min(fl["left"][0], fl["right"][0]) + 1e-6
        assert fl["left"][1] == fl["right"][1] == pm
    assert nt.psi_len == osb.n
    c1, o1, c2, o2 = nt.arrays()
    rng = np.random.default_rng(Lx * 1000 + Ny * 10 + l)
    v = rng.normal(size=osb.n) + 1j * rng.normal(size=osb.n)
    got = plan_exec.execute(nt.arena, c1, o1, c2, o2, nt.mid_len, v, osb.n)
    want = osb.matvec(v)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    # the window-by-window evaluation (sampled checks of large plans) agrees
    which = list(range(0, len(o2), 3))
    for w, acc in plan_exec.execute_windows(nt.arena, c1, o1, c2, o2, nt.mid_len, v, which).items():
        ob = o2[w]
        o, ldc, m, n = int(ob["c_off"]), int(ob["ldc"]), int(ob["m"]), int(ob["n"])
        idx = (o + np.arange(m)[:, None] + ldc * np.arange(n)[None, :]).ravel(order="F")
        assert np.allclose(acc.ravel(order="F"), got[idx], rtol=0, atol=1e-12 * max(1.0, np.abs(got).max()))


@pytest.mark.parametrize("Lx,Ny,Ne,S2,l", CASES)
def test_library_model_tasks_match_oracle(Lx, Ny, Ne, S2, l):
    """tsc_model_tasks (bond classes, merged partner operators and the term
    list in the library, the host logic of model_wave and hmerge) gives the
    oracle's H psi from the two block models and the lattice bonds."""
    osb, psb = _setup(Lx, Ny, Ne, S2, l)
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    nt = native.ModelTasks(psb.sys, psb.env, l + 1, l + 2, b, Ne, S2)
    assert nt.psi_len == osb.n
    c1, o1, c2, o2 = nt.arrays()
    rng = np.random.default_rng(Lx * 1000 + Ny * 10 + l + 7)
    v = rng.normal(size=osb.n) + 1j * rng.normal(size=osb.n)
    got = plan_exec.execute(nt.arena, c1, o1, c2, o2, nt.mid_len, v, osb.n)
    want = osb.matvec(v)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    nt.close()


def test_library_renorm_tasks_match_oracle():
    """tsc_renorm_tasks (Get_Hamiltonian_Sys + update_trun in the library)
    gives the oracle's enlarged-and-truncated block: Hamiltonian and the
    c+, S, n of every kept site, in the new sector order."""
    Lx, Ny = 3, 4
    N = Lx * Ny
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    sys_blk, _, nb = dmrg_exact.build_chain_blocks(N, b, 5)
    new_site = 6
    to_new = [(i, t, J) for (i, t, J) in nb[new_site] if i in sys_blk.sites]
    keep = [i for i in sys_blk.sites + [new_site] if any(j > new_site for (j, _, _) in nb.get(i, []))]
    eb = su2_dense.enlarge(sys_blk, new_site, to_new, keep)
    rng = np.random.default_rng(5)
    U, order = {}, []
    for x, (Nx, Jx, dx) in enumerate(eb.sectors):
        if x % 3 == 1:
            continue
        k = max(1, dx // 2)
        U[x] = np.linalg.qr(rng.normal(size=(dx, k)) + 1j * rng.normal(size=(dx, k)))[0]
        order.append(x)
    order = order[::-1]
    want = su2_dense.truncate(eb, [U.get(x) for x in range(len(eb.sectors))])
    bonds = [(i, new_site, t, J) for (i, t, J) in to_new]
    rt = native.RenormTasks(block_to_product(sys_blk), new_site, bonds, keep, order,
                            [U[x].shape[1] for x in order])
    assert rt.enlarged_sectors == [tuple(s) for s in eb.sectors]
    uin = np.concatenate([U[x].ravel(order="F") for x in order])
    assert rt.in_len == uin.size
    c1, o1, c2, o2 = rt.arrays()
    res = plan_exec.execute(rt.arena, c1, o1, c2, o2, rt.mid_len, uin, rt.out_len)
    okeep = [x for x in range(len(eb.sectors)) if x in U]
    oidx = {x: n for n, x in enumerate(okeep)}
    got = {}
    for (key, b2, k2), off in zip(rt.layout, rt.offsets):
        kb, kk = rt.new_sectors[b2][2], rt.new_sectors[k2][2]
        got[(key, oidx[order[b2]], oidx[order[k2]])] = res[off:off + kb * kk].reshape((kb, kk), order="F")
    ops = [("H", want.H)] + [((i, name), want.ops[i][name]) for i in keep for name in ("c", "S", "n")]
    for key, (name, wop) in zip([-1] + list(range(3 * len(keep))), ops):
        for (b2, k2), m in wop.blocks.items():
            g = got.get((key, b2, k2))
            if g is None:
                assert np.abs(m).max() < 1e-13, (name, b2, k2)
            else:
                assert np.allclose(g, m, atol=1e-12), (name, b2, k2)
        for (kk, b2, k2), m in got.items():
            if kk == key and (b2, k2) not in wop.blocks:
                assert np.abs(m).max() < 1e-13
    rt.close()
