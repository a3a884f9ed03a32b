"""The reference's 14 superblock kernels act as physical operators on the
wave functions this build stores.

Each kernel of model_wave (get_sys.f90:1783-5204), restated loop by loop in
oracle/ref_kernels.py with the reference's own 6j coefficients, phases and
fermion signs, is applied to every basis wave function of a 6-site
superblock (sys block 1-2, sys site 3, env site 4, env block 5-6; N = 4,
S = 0) built in the reference's conventions (oracle/ref_blocks.py), with
block operators given as the reference stores them. Each must equal one
constant times the physical operator of its term, with the constant that
model_wave's couplings assume (main.f90:61-73): jt = -sqrt2 t needs
1/sqrt2 for sum_s (c+_is c_js + h.c.), jd = -sqrt3 J needs -1/sqrt3 for
S_i.S_j, jn = -J/4 needs 1 for n_i n_j, and the block Hamiltonians enter
with 1. This pins the reference's per-kernel phases and the superblock
conventions of eigvec1 together: a sign slip in any kernel, or a wave
function coupled or ordered differently, leaves a residual or changes a
constant. The product's H.psi is pinned to the same physical Hamiltonian
by exact diagonalisation (test_gpu_physics.py, test_physics_plan.py), so on
the same blocks it applies what model_wave applies.
"""
import numpy as np
import pytest

pytest.importorskip("sympy")
from oracle import ref_blocks as RB  # noqa: E402
from oracle import ref_kernels as K  # noqa: E402

LAM_HOP, LAM_SPIN, LAM_N = 1.0 / np.sqrt(2.0), -1.0 / np.sqrt(3.0), 1.0
NAMES = {"elec_up": ("c+", 1, 1), "elec_down": ("c~", -1, 1), "spin_sd": ("S", 0, 2), "num_sn": ("n", 0, 0)}


@pytest.fixture(scope="module")
def sb():
    site = RB.BlockStates.site()

    def whole(eb):
        return eb.truncate([(x, np.eye(d)) for x, (_, _, d) in enumerate(eb.sectors)])

    sys2, env2 = whole(site.enlarge(False)), whole(site.enlarge(True))
    sys_eb, env_eb = sys2.enlarge(False), env2.enlarge(True)
    _F2, o2 = RB.fock_ops(2)

    def ops(states, loc, env):
        out = {}
        for name, (fock, up, k) in NAMES.items():
            red, dev = RB.reduced(states, states, k, o2[loc][fock][1], up)
            assert dev < 1e-12
            out[name] = K.from_reduced(red, states.sectors, up, k, env)
        return out

    t, J = 0.7, 0.3
    hams = []
    for states, env in ((sys2, False), (env2, True)):
        red, dev = RB.reduced(states, states, 0, RB.block_hamiltonian(RB.fock_ops(2)[0], [(0, 1, t, J)]), 0)
        hams.append(K.from_reduced(red, states.sectors, 0, 0, env))
    keys = [(x, y) for x, bx in enumerate(sys_eb.basis) for y, by in enumerate(env_eb.basis)
            if bx[0] + by[0] == 4 and bx[1] == by[1]]
    idx = [(x, y, r, c) for (x, y) in keys for r in range(sys_eb.basis[x][2]) for c in range(env_eb.basis[y][2])]
    P = []
    for (x, y, r, c) in idx:
        m = np.zeros((sys_eb.basis[x][2], env_eb.basis[y][2]))
        m[r, c] = 1.0
        P.append(RB.super_state(sys_eb, env_eb, [(x, y, m)]))
    P = np.array(P)
    assert np.abs(P.conj() @ P.T - np.eye(len(idx))).max() < 1e-12
    return {"sys_eb": sys_eb, "env_eb": env_eb, "keys": keys, "idx": idx, "P": P, "F": RB.R.Fock(6),
            "sys": {1: ops(sys2, 0, False), 2: ops(sys2, 1, False)},
            "env": {5: ops(env2, 0, True), 6: ops(env2, 1, True)}, "hams": hams, "tJ": (t, J)}


def _matrix(sb, fn):
    pos = {k: i for i, k in enumerate(sb["idx"])}
    M = np.zeros((len(sb["idx"]),) * 2, complex)
    for b, (x, y, r, c) in enumerate(sb["idx"]):
        one = {k: np.zeros((sb["sys_eb"].basis[k[0]][2], sb["env_eb"].basis[k[1]][2]), complex) for k in sb["keys"]}
        one[(x, y)][r, c] = 1.0
        for k, m in fn(one).items():
            for (rr, cc) in zip(*np.nonzero(m)):
                M[pos[(k[0], k[1], rr, cc)], b] = m[rr, cc]
    return M


def _phys(sb, kind, a, b):
    X = RB.R.physical(sb["F"], kind, a - 1, b - 1)
    return X + X.T if kind == "hop" else X


def _constant(sb, fn, X):
    """c with kernel = c X on the wave functions, asserting the residual."""
    M = _matrix(sb, fn)
    XM = sb["P"].conj() @ X @ sb["P"].T
    assert np.abs(XM).max() > 0.1
    lam = (np.vdot(XM, M) / np.vdot(XM, XM)).real
    assert np.abs(M - lam * XM).max() < 1e-12
    return lam


def _cases(sb):
    """(name, kernel applied to psi, physical operator, constant)."""
    se, ee = sb["sys_eb"], sb["env_eb"]
    up, dn, sd, num = (K.site_op(n) for n in ("st_elec_up", "st_elec_down", "st_sd", "num_elec"))
    T = K.trans
    S, E = sb["sys"], sb["env"]
    out = [("site_vec_site_ndia hop 3-4", lambda p: K.site_vec_site_ndia(up, "F", T(dn), "F", se, ee, p, 1.0),
            ("hop", 3, 4), LAM_HOP),
           ("site_vec_site_ndia S.S 3-4", lambda p: K.site_vec_site_ndia(sd, "B", sd, "B", se, ee, p, 1.0),
            ("SS", 3, 4), LAM_SPIN),
           ("site_vec_site_dia n n 3-4", lambda p: K.site_vec_site_dia(num, num, se, ee, p, 1.0), ("nn", 3, 4), LAM_N)]
    for i in (1, 2):
        o = S[i]
        out += [(f"block_site_vec_ndia hop {i}-3",
                 lambda p, o=o: K.block_site_vec_ndia(o["elec_up"], "F", T(dn), "F", se, p, 1.0), ("hop", i, 3), LAM_HOP),
                (f"block_site_vec_ndia S.S {i}-3",
                 lambda p, o=o: K.block_site_vec_ndia(o["spin_sd"], "B", sd, "B", se, p, 1.0), ("SS", i, 3), LAM_SPIN),
                (f"block_site_vec_dia n n {i}-3",
                 lambda p, o=o: K.block_site_vec_dia(o["num_sn"], num, se, p, 1.0), ("nn", i, 3), LAM_N),
                (f"block_vec_site_ndia hop {i}-4",
                 lambda p, o=o: K.block_vec_site_ndia(o["elec_up"], "F", T(dn), "F", se, ee, p, 1.0), ("hop", i, 4),
                 LAM_HOP),
                (f"block_vec_site_ndia S.S {i}-4",
                 lambda p, o=o: K.block_vec_site_ndia(o["spin_sd"], "B", sd, "B", se, ee, p, 1.0), ("SS", i, 4),
                 LAM_SPIN),
                (f"block_vec_site_dia n n {i}-4",
                 lambda p, o=o: K.block_vec_site_dia(o["num_sn"], num, se, ee, p, 1.0), ("nn", i, 4), LAM_N)]
        for j in (5, 6):
            q = E[j]
            out += [(f"block_vec_block_ndia hop {i}-{j}",
                     lambda p, o=o, q=q: K.block_vec_block_ndia(o["elec_up"], "F", T(q["elec_down"]), "F", se, ee, p, 1.0),
                     ("hop", i, j), LAM_HOP),
                    (f"block_vec_block_ndia S.S {i}-{j}",
                     lambda p, o=o, q=q: K.block_vec_block_ndia(o["spin_sd"], "B", T(q["spin_sd"]), "B", se, ee, p, 1.0),
                     ("SS", i, j), LAM_SPIN),
                    (f"block_vec_block_dia n n {i}-{j}",
                     lambda p, o=o, q=q: K.block_vec_block_dia(o["num_sn"], q["num_sn"], se, ee, p, 1.0), ("nn", i, j),
                     LAM_N)]
    for j in (5, 6):
        q = E[j]
        out += [(f"vec_block_site_ndia hop {j}-4",
                 lambda p, q=q: K.vec_block_site_ndia(q["elec_up"], "F", T(dn), "F", ee, se, p, 1.0), ("hop", j, 4),
                 LAM_HOP),
                (f"vec_block_site_ndia S.S {j}-4",
                 lambda p, q=q: K.vec_block_site_ndia(q["spin_sd"], "B", sd, "B", ee, se, p, 1.0), ("SS", j, 4),
                 LAM_SPIN),
                (f"vec_block_site_dia n n {j}-4",
                 lambda p, q=q: K.vec_block_site_dia(q["num_sn"], num, ee, p, 1.0), ("nn", j, 4), LAM_N),
                (f"site_vec_block_ndia hop 3-{j}",
                 lambda p, q=q: K.site_vec_block_ndia(up, "F", T(q["elec_down"]), "F", se, ee, p, 1.0), ("hop", 3, j),
                 LAM_HOP),
                (f"site_vec_block_ndia S.S 3-{j}",
                 lambda p, q=q: K.site_vec_block_ndia(sd, "B", T(q["spin_sd"]), "B", se, ee, p, 1.0), ("SS", 3, j),
                 LAM_SPIN),
                (f"site_vec_block_dia n n 3-{j}",
                 lambda p, q=q: K.site_vec_block_dia(num, q["num_sn"], se, ee, p, 1.0), ("nn", 3, j), LAM_N)]
    return out


def test_every_kernel_is_its_physical_term(sb):
    seen = set()
    for name, fn, (kind, a, b), want in _cases(sb):
        lam = _constant(sb, fn, _phys(sb, kind, a, b))
        assert lam == pytest.approx(want, rel=1e-12), name
        seen.add(name.split()[0])
    assert len(seen) == 12


def test_block_hamiltonian_kernels(sb):
    t, J = sb["tJ"]
    F = sb["F"]
    Hs = RB.block_hamiltonian(F, [(0, 1, t, J)])[0]
    He = RB.block_hamiltonian(F, [(4, 5, t, J)])[0]
    se, ee = sb["sys_eb"], sb["env_eb"]
    assert _constant(sb, lambda p: K.block_vec_dia(sb["hams"][0], se, p, 1.0), Hs) == pytest.approx(1.0, rel=1e-12)
    assert _constant(sb, lambda p: K.vec_block_dia(sb["hams"][1], ee, p, 1.0), He) == pytest.approx(1.0, rel=1e-12)


def test_a_sign_slip_is_seen(sb):
    """The check is sensitive: the env block operator passed without
    block_transfer_trans, or the env block sign (-1)^(N_sys + N_site)
    dropped, no longer gives a multiple of the physical hopping."""
    se, ee = sb["sys_eb"], sb["env_eb"]
    o, q = sb["sys"][2], sb["env"][5]
    X = _phys(sb, "hop", 2, 5)
    XM = sb["P"].conj() @ X @ sb["P"].T
    bad = [lambda p: K.block_vec_block_ndia(o["elec_up"], "F", q["elec_down"], "F", se, ee, p, 1.0),
           lambda p: K.block_vec_block_ndia(o["elec_up"], "F", K.trans(q["elec_down"]), "B", se, ee, p, 1.0)]
    for fn in bad:
        M = _matrix(sb, fn)
        lam = (np.vdot(XM, M) / np.vdot(XM, XM)).real
        assert np.abs(M - lam * XM).max() > 1e-3


def test_wave_trans_keeps_the_state():
    """wave_trans_LR / wave_trans_RL (dmrg_sub.f90:1471-1789), restated, carry
    a wave function to the next superblock of a sweep without changing the
    physical state when the truncations are complete, here random complex
    unitaries per sector: the Fock vector before and after agree to 1e-15.
    This fixes the recoupling phases of both routines and that the
    reference's env block states are the conjugated columns of its env
    truncation matrices (reading them as plain columns fails)."""
    rng = np.random.default_rng(7)

    def unitaries(eb):
        out = {}
        for (n, j, d) in eb.sectors:
            q, _ = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
            out[(n, j)] = q
        return out

    def apply(eb, U, env):
        return eb.truncate([(eb.index[k], np.conj(u) if env else u) for k, u in U.items()])

    def keys_of(se, ee):
        return [(x, y) for x, bx in enumerate(se.basis) for y, by in enumerate(ee.basis)
                if bx[0] + by[0] == 4 and bx[1] == by[1]]

    def state(se, ee, psi):
        return RB.super_state(se, ee, [(x, y, m) for (x, y), m in psi.items()])

    site = RB.BlockStates.site()
    s12, e56 = site.enlarge(False), site.enlarge(True)
    Us2, Ue2 = unitaries(s12), unitaries(e56)
    so, eo = apply(s12, Us2, False).enlarge(False), apply(e56, Ue2, True).enlarge(True)   # 1-3 | 4-6
    psi = {k: rng.normal(size=(so.basis[k[0]][2], eo.basis[k[1]][2]))
           + 1j * rng.normal(size=(so.basis[k[0]][2], eo.basis[k[1]][2])) for k in keys_of(so, eo)}
    nrm = np.sqrt(sum(np.vdot(m, m).real for m in psi.values()))
    psi = {k: m / nrm for k, m in psi.items()}
    before = state(so, eo, psi)
    # left to right: sys 1-3 truncated (complete), new sys site 4, env 5-6 expanded
    Us3 = unitaries(so)
    sn = apply(so, Us3, False).enlarge(False)
    new = K.wave_trans_LR(sn, e56, Us3, Ue2, (so, eo), psi, keys_of(sn, e56))
    assert np.linalg.norm(state(sn, e56, new) - before) < 1e-13
    # right to left: env 4-6 truncated, new env site 3, sys 1-2 expanded
    Ue3 = unitaries(eo)
    en = apply(eo, Ue3, True).enlarge(True)
    new = K.wave_trans_RL(s12, en, Us2, Ue3, (so, eo), psi, keys_of(s12, en))
    assert np.linalg.norm(state(s12, en, new) - before) < 1e-13
    wrong = apply(eo, Ue3, False).enlarge(True)
    assert np.linalg.norm(state(s12, wrong, new) - before) > 0.1


@pytest.mark.parametrize("side", ["sys", "env"])
def test_enlargement_routines_give_the_physical_operators(side):
    """update_block_ndia_sys/env, update_site_ndia_sys/env, update_block_dia
    and update_site_dia (block.f90:2495-3459), restated, carry the block's
    and the new site's c+, c~, S and n into the enlarged basis as exactly
    the reduced matrices of the physical operators between the enlarged
    states (sys: block then site in chain order; env: site then block), the
    6j coefficients and both sides' sign rules included; swapping the sign
    rules of the two sides breaks it."""
    env = side == "env"
    site = RB.BlockStates.site()

    def whole(eb):
        return eb.truncate([(x, np.eye(d)) for x, (_, _, d) in enumerate(eb.sectors)])

    blk = whole(site.enlarge(env))                       # two sites
    eb = blk.enlarge(env)                                # three sites
    _F2, o2 = RB.fock_ops(2)
    _F3, o3 = RB.fock_ops(3)
    # chain-order positions in the 3-site Fock space: sys block 0, 1 + site 2;
    # env site 0 + block 1, 2
    blk_loc = {0: 0, 1: 1} if not env else {0: 1, 1: 2}
    new_loc = 2 if not env else 0
    checked = 0
    for name, (fock, up, k) in NAMES.items():
        flag = "F" if up else "B"
        for b, loc in blk_loc.items():
            red, dev = RB.reduced(blk, blk, k, o2[b][fock][1], up)
            assert dev < 1e-12
            op = K.from_reduced(red, blk.sectors, up, k)
            got = K.update_block_dia(op, eb) if k == 0 else K.update_block_ndia(op, eb, flag, env)
            want, dev = RB.reduced(eb, eb, k, o3[loc][fock][1], up)
            assert dev < 1e-12
            want = K.from_reduced(want, eb.sectors, up, k)
            for key in set(got.subs) | set(want.subs):
                a, w = got.subs.get(key), want.subs.get(key)
                a = np.zeros_like(w) if a is None else a
                w = np.zeros_like(a) if w is None else w
                assert np.abs(a - w).max() < 1e-12, (side, name, b, key)
            checked += 1
        sop = K.site_op({"c+": "st_elec_up", "c~": "st_elec_down", "S": "st_sd", "n": "num_elec"}[fock])
        got = K.update_site_dia(sop, eb) if k == 0 else K.update_site_ndia(sop, eb, flag, env)
        want, dev = RB.reduced(eb, eb, k, o3[new_loc][fock][1], up)
        want = K.from_reduced(want, eb.sectors, up, k)
        for key in set(got.subs) | set(want.subs):
            a, w = got.subs.get(key), want.subs.get(key)
            a = np.zeros_like(w) if a is None else a
            w = np.zeros_like(a) if w is None else w
            assert np.abs(a - w).max() < 1e-12, (side, name, "site", key)
        # the other side's sign rule is wrong here
        if flag == "F":
            bad = K.update_site_ndia(sop, eb, flag, not env)
            assert max(np.abs(bad.subs[key] - want.subs.get(key, 0 * bad.subs[key])).max() for key in bad.subs) > 0.1
        checked += 1
    assert checked == 12


@pytest.mark.parametrize("side", ["sys", "env"])
def test_block_site_bond_of_the_enlarged_hamiltonian(side):
    """Get_Hamiltonian_Sys/Env's bond terms (get_sys.f90:40-180, 331-460):
    Get_hamiltonian_ndia with jt on (elec_up, trans(st_elec_down)) for the
    sys side or (elec_down, trans(st_elec_up)) for the env side, with jd on
    (spin_sd, st_sd), and Get_hamiltonian_dia with jn on the enlarged n's,
    restated, sum to the reduced matrix of the physical bond
    -t sum_s (c+ c + h.c.) + J (S.S - n n / 4) between a block site and the
    new site, for the couplings main.f90:61-73 sets (jt = -sqrt2 t,
    jd = -sqrt3 J, jn = -J/4)."""
    env = side == "env"
    site = RB.BlockStates.site()

    def whole(eb):
        return eb.truncate([(x, np.eye(d)) for x, (_, _, d) in enumerate(eb.sectors)])

    blk = whole(site.enlarge(env))
    eb = blk.enlarge(env)
    _F2, o2 = RB.fock_ops(2)
    F3 = RB.R.Fock(3)
    new_loc = 0 if env else 2
    T = K.trans
    for b in (0, 1):
        loc = b + 1 if env else b

        def op(fock, up, k):
            red, dev = RB.reduced(blk, blk, k, o2[b][fock][1], up)
            assert dev < 1e-12
            return K.from_reduced(red, blk.sectors, up, k)

        for (t, J) in ((1.0, 0.0), (0.0, 1.0), (0.7, 0.3)):
            jt, jd, jn = -np.sqrt(2.0) * t, -np.sqrt(3.0) * J, -0.25 * J
            if env:
                hop = K.get_hamiltonian_ndia(op("c~", -1, 1), T(K.site_op("st_elec_up")), eb, jt, "F", "E")
            else:
                hop = K.get_hamiltonian_ndia(op("c+", 1, 1), T(K.site_op("st_elec_down")), eb, jt, "F", "S")
            spin = K.get_hamiltonian_ndia(op("S", 0, 2), K.site_op("st_sd"), eb, jd, "B", "S" if not env else "E")
            nn = K.get_hamiltonian_dia(K.update_block_dia(op("n", 0, 0), eb), K.update_site_dia(K.site_op("num_elec"), eb),
                                       eb, jn)
            a, c_ = min(loc, new_loc), max(loc, new_loc)
            X = RB.block_hamiltonian(F3, [(a, c_, t, J)])
            want, dev = RB.reduced(eb, eb, 0, X, 0)
            assert dev < 1e-12
            for x in range(len(eb.basis)):
                got = hop[x] + spin[x] + nn[x]
                assert np.abs(got - want.get((x, x), 0 * got)).max() < 1e-12, (side, b, t, J, x)
