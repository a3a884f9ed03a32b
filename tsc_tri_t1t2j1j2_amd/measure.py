"""Ground-state measurement: electron density, spin, density and hopping
correlators of the final wave function (Measurement,
hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/main.f90:213-354; measure_operator_dia
and Get_density_cor, DMRG_code/measure.f90:2525-2681).

The state measured is the ground state of the last step of a final
left-to-right half-sweep (the reference sweeps to 5N/6 before measuring,
main.f90:205-206): psi of the superblock [A, s1 | B, s2] whose sys blocks
were built by that half-sweep's truncations and whose env blocks were built
by the last right-to-left half-sweep.

A site operator is carried from the length at which its site joins a block
up to the measured block through the stored truncations, U^H (O x 1) U per
enlarged sector (the reference transports them the same way:
Get_sys_oper_ndia / update_trun_*, get_sys.f90:1147-1782). A two-site
operator of sites i < j that both lie in A is formed when j joins the block,
as [O_i x O_j]^0, and carried on as a scalar. All carried operators of one
side stay on the device; one grouped plan per block length transports all
of them at once. An expectation value is <psi|X psi> / <psi|psi> with X psi
from one contraction plan over the final superblock and the dot product on
the device.

Normalisation (physical operators, no SU(2) factors left):
  n_i;   S_i.S_j = -sqrt3 [S_i x S_j]^0;   n_i n_j = [n_i x n_j]^0;
  (sum_s c+_is c_js + h.c.) / 2 = -([c+_i x c~_j]^0 + [c~_i x c+_j]^0) / sqrt2
with i before j in the fermion order A, s1, B, s2. These are the values the
reference's files hold: its coupling rule (block.f90:376-635) with its site
elements and Get_density_cor's factor sqrt(1+down_dif), negated for
down_dif = 2 (measure.f90:2586-2587), gives <S_i.S_j>, <n_i n_j> and
<sum_s c+_is c_js>, and its pair values (measure.f90:2762, 2806) are
<D+_ab D_cd>. For the real Hamiltonian the hopping value is real; the
imaginary columns the reference writes are rounding noise, written as 0.
"""
import time
from math import sqrt

import numpy as np

from . import capi
from .blocks import SITE_SECTORS, TotalBlock, site_operators
from . import su2
from .hamiltonian import _entries, _renorm_tasks, cross_terms, scalar_transitions, side_transitions
from .planner import IDENT, Operand, PlanBuilder
from .resident import DevArena, DevMat

SQ2 = sqrt(2.0)
SQ3 = sqrt(3.0)

# (coupling, operator of the earlier site, its tilde, operator of the later site, its tilde)
CORRELATORS = {
    "SS": ((-SQ3, "S", False, "S", False),),
    "nn": ((1.0, "n", False, "n", False),),
    "cc": ((-1.0 / SQ2, "c", False, "c", True), (-1.0 / SQ2, "c", True, "c", False)),
}


def _needed(kinds):
    """Site operator names the correlators of `kinds` use."""
    out = {"n"}
    for k in kinds:
        for (_, pn, _, qn, _) in CORRELATORS[k]:
            out.add(pn)
            out.add(qn)
    return sorted(out)


class Carrier:
    """Operators carried along one side's blocks.

    truns[L]: the truncation (wavetrans.Trunc) that built the block of
    length L from the block of length L-1 and the site site_at(L); length 1
    is the bare site site_at(1). ops: {key: TotalBlock} on the current
    block, every matrix a DevMat of one device arena."""

    def __init__(self, ctx, truns, site_at, prec=capi.PREC_F64):
        self.ctx, self.truns, self.site_at, self.prec = ctx, truns, site_at, prec
        self.length = 1
        self.sectors = [tuple(s) for s in SITE_SECTORS]
        self.ops = {}
        self.arena = None
        self.flops = 0.0
        self.ms = 0.0
        self.seconds = 0.0

    def start(self, names):
        """Operators of the first site itself (the block of length 1)."""
        so = site_operators()
        vals, keys = [], []
        for name in names:
            for key, m in so[name].blocks.items():
                keys.append((name, key, len(vals)))
                vals.append(complex(m[0, 0]))
        arena = DevArena(self.ctx.vec_from(np.array(vals, np.complex128)))
        site = self.site_at(1)
        for name in names:
            self.ops[("op", site, name)] = TotalBlock(so[name].up_dif, so[name].k2)
        for (name, key, o) in keys:
            self.ops[("op", site, name)].blocks[key] = DevMat(arena, o, 1, 1)
        self.arena = arena

    def grow(self, new_terms=None):
        """Advance one length: carry every operator through the next
        truncation and add the operators new_terms(basis, sectors, ops)
        returns ({key: (transitions, (up_dif, k2))}) on the enlarged block."""
        t0 = time.perf_counter()
        ctx = self.ctx
        L = self.length + 1
        tr = self.truns[L]
        basis = tr.basis
        if [tuple(s) for s in basis.block_sectors] != self.sectors:
            raise ValueError(f"stored truncation of length {L} does not match the carried block")
        trans, meta = {}, {}
        for key, tb in self.ops.items():
            trans[key] = side_transitions(basis, self.sectors, ("blk", tb, False))
            meta[key] = (tb.up_dif, tb.k2)
        if new_terms is not None:
            for key, (t, m) in new_terms(basis, self.sectors, self.ops).items():
                trans[key] = t
                meta[key] = m
        U, order = tr.U, tr.order
        uoff, parts, off = {}, [], 0
        for x in order:
            u = np.asarray(U[x], np.complex128)
            uoff[x] = off
            parts.append(u.ravel(order="F"))
            off += u.size
        uin = np.concatenate(parts) if parts else np.zeros(1, np.complex128)
        plan, windows, out_len = _renorm_tasks(trans, basis, U, order, uoff,
                                               lambda m: (m.off, max(m.shape[0], 1)))
        dummy = None
        if self.arena is None:
            dummy = ctx.vec(1)
            ops_vec = dummy
        else:
            ops_vec = self.arena.vec
        pl = plan.build_dev(ctx, ops_vec, uin.size, max(out_len, 1))
        uvec = ctx.vec_from(uin)
        out = ctx.vec(max(out_len, 1))
        pl.apply(uvec, out, self.prec)
        ctx.sync()
        self.flops += pl.flops()
        self.ms += sum(pl.last_times())
        pl.close()
        uvec.close()
        if dummy is not None:
            dummy.close()
        arena = DevArena(out)
        new_index = {x: n for n, x in enumerate(order)}
        ops = {key: TotalBlock(*meta[key]) for key in trans}
        for (opkey, x2, x), (_, o) in windows.items():
            ops[opkey].blocks[(new_index[x2], new_index[x])] = DevMat(arena, o, U[x2].shape[1], U[x].shape[1])
        if self.arena is not None:
            self.arena.close()
        self.arena = arena
        self.ops = ops
        self.sectors = [(basis.sectors[x][0], basis.sectors[x][1], U[x].shape[1]) for x in order]
        self.length = L
        self.seconds += time.perf_counter() - t0

    def close(self):
        if self.arena is not None:
            self.arena.close()
        self.arena = None
        self.ops = {}


def new_site_terms(site, names, pair_specs=()):
    """new_terms for Carrier.grow: the operators `names` of the joining site
    and, for each (key, products, partner) of pair_specs, the scalar
    sum_k g_k [O_partner,k x O_site,k]^0 (products as in CORRELATORS)."""
    so = site_operators()

    def terms(basis, sectors, ops):
        out = {}
        for name in names:
            out[("op", site, name)] = (side_transitions(basis, sectors, ("site", so[name], False)),
                                       (so[name].up_dif, so[name].k2))
        for (key, products, partner) in pair_specs:
            pairs = [(g, ops[("op", partner, pn)], ptil, so[qn], qtil) for (g, pn, ptil, qn, qtil) in products]
            out[key] = (scalar_transitions(basis, sectors, pairs=pairs, reduced=True), (0, 0))
        return out

    return terms


def pair0_transitions(basis, blk_sectors, g, T, ttil, U, util, reduced=True):
    """g [T(block) x U(site)]^0 for operators that may change the particle
    number (a singlet pair creator, for one): the target enlarged sector has
    the same spin and N + dN. f: reduced element of the enlarged block
    (reduced=True) or the m-independent element."""
    k2 = T.k2
    tent = _entries(T, blk_sectors, ttil)
    uent = _entries(U, SITE_SECTORS, util)
    pos = [{(a, s): (p, d) for (a, s, p, d) in sec[3]} for sec in basis.sectors]
    out = []
    for x, (_, J, _, subs) in enumerate(basis.sectors):
        scale = sqrt(J + 1) if reduced else 1.0
        for (a, s, p, d) in subs:
            Na, ja = blk_sectors[a][0], blk_sectors[a][1]
            js = SITE_SECTORS[s][1]
            # U passes the block's fermions of the ket
            sgn = -1.0 if (U.odd and Na % 2) else 1.0
            for (a2, a_, stored, opc, ph) in tent:
                if a_ != a:
                    continue
                for (s2, s_, ust, uopc, uph) in uent:
                    if s_ != s:
                        continue
                    x2 = basis.index.get((blk_sectors[a2][0] + SITE_SECTORS[s2][0], J))
                    if x2 is None or (a2, s2) not in pos[x2]:
                        continue
                    p2, d2 = pos[x2][(a2, s2)]
                    elem = ust[0, 0] if uopc == capi.OP_N else np.conj(ust[0, 0])
                    f = su2.scalar_pair(blk_sectors[a2][1], ja, SITE_SECTORS[s2][1], js, J, k2)
                    f *= g * elem * ph * uph * sgn * scale
                    if f != 0.0:
                        out.append((x, p, d, x2, p2, d2, f, ((id(T), a2, a, opc), stored, opc)))
    return out


def scalar_times_site(basis, blk_sectors, P, U, util):
    """Reduced transitions of P(block) U(site): P a spin scalar of the block
    (any particle change), U a site operator of rank k; the product has
    rank k. P's m-independent element is its reduced one / sqrt(2j+1)."""
    k2 = U.k2
    pent = _entries(P, blk_sectors, False)
    uent = _entries(U, SITE_SECTORS, util)
    out = []
    for x, (_, J, _, subs) in enumerate(basis.sectors):
        for (a, s, p, d) in subs:
            Na, ja = blk_sectors[a][0], blk_sectors[a][1]
            js = SITE_SECTORS[s][1]
            sgn = -1.0 if (U.odd and Na % 2) else 1.0
            for (a2, a_, stored, opc, ph) in pent:
                if a_ != a:
                    continue
                d2 = blk_sectors[a2][2]
                for (s2, s_, ust, uopc, uph) in uent:
                    if s_ != s:
                        continue
                    elem = ust[0, 0] if uopc == capi.OP_N else np.conj(ust[0, 0])
                    js2 = SITE_SECTORS[s2][1]
                    for (x2, p2) in basis.where.get((a2, s2), ()):
                        f = su2.embed_site(js2, js, ja, basis.sectors[x2][1], J, k2) / sqrt(ja + 1)
                        f *= elem * ph * uph * sgn
                        if f != 0.0:
                            out.append((x, p, d, x2, p2, d2, f, ((id(P), a2, a, opc), stored, opc)))
    return out


class SuperblockExpectation:
    """<psi| X |psi> / <psi|psi> for operators X of the final superblock,
    built from the operators two carriers hold on A and B."""

    def __init__(self, ctx, psi, bases, sysc, envc, prec=capi.PREC_F64):
        self.ctx, self.psi, self.prec = ctx, psi, prec
        self.L, self.R, self.SB = bases.sys_bs, bases.env_bs, bases.super
        if [tuple(s) for s in self.L.block_sectors] != sysc.sectors:
            raise ValueError("sys carrier does not end on the measured sys block")
        if [tuple(s) for s in self.R.block_sectors] != envc.sectors:
            raise ValueError("env carrier does not end on the measured env block")
        self.A_sec, self.B_sec = sysc.sectors, envc.sectors
        nA = sysc.arena.n if sysc.arena is not None else 0
        nB = envc.arena.n if envc.arena is not None else 0
        self.vec = ctx.vec(max(nA + nB, 1))
        self.base = {}
        if nA:
            self.vec.copy_range(0, sysc.arena.vec, 0, nA)
            self.base[id(sysc.arena)] = 0
        if nB:
            self.vec.copy_range(nA, envc.arena.vec, 0, nB)
            self.base[id(envc.arena)] = nA
        ctx.sync()
        self.phi = ctx.vec(self.SB.dim)
        self.norm2 = psi.dot_re(psi)
        self.by_x, self.by_y = {}, {}
        for i, (x, y, _, _) in enumerate(self.SB.sub):
            self.by_x.setdefault(x, []).append(i)
            self.by_y.setdefault(y, []).append(i)
        self.plans = 0
        self.seconds = 0.0

    def _resolve(self, m):
        return self.base[id(m.arena)] + m.off, max(m.shape[0], 1)

    def _value(self, pb):
        t0 = time.perf_counter()
        plan = pb.build_dev(self.ctx, self.vec, self.SB.dim, self.SB.dim)
        plan.apply(self.psi, self.phi, self.prec)
        self.ctx.sync()
        v = self.psi.dot_re(self.phi) / self.norm2
        plan.close()
        self.plans += 1
        self.seconds += time.perf_counter() - t0
        return v

    # ---- one-side scalars: X_L x 1 (row bands) or 1 x X_R (column bands)
    def left_scalar(self, trans):
        """trans: J-preserving transitions on the sys enlarged basis, with
        m-independent factors."""
        SB, L = self.SB, self.L
        pb = PlanBuilder(resolver=self._resolve)
        win = {}
        for i, (x, y, dx, dy) in enumerate(SB.sub):
            for (a, s, p, d) in L.sectors[x][3]:
                win[(i, p)] = pb.out_window(("l", i, p), SB.offset[i] + p, dx, d, dy)
        for (x, p, d, x2, p2, d2, f, mat) in trans:
            if x2 != x:
                raise ValueError("left_scalar needs a J-preserving operator")
            for i in self.by_x.get(x, ()):
                dx = SB.sub[i][2]
                src = Operand(capi.BUF_IN, SB.offset[i] + p, dx, capi.OP_N)
                if mat is None:
                    pb.add(win[(i, p2)], src, IDENT, 0, f)
                else:
                    mo, mld = pb.op(mat[0], mat[1])
                    pb.add(win[(i, p2)], Operand(capi.BUF_OPS, mo, mld, mat[2]), src, d, f)
        return self._value(pb)

    def right_scalar(self, trans):
        SB, R = self.SB, self.R
        pb = PlanBuilder(resolver=self._resolve)
        win = {}
        for i, (x, y, dx, dy) in enumerate(SB.sub):
            for (b, t, q, e) in R.sectors[y][3]:
                win[(i, q)] = pb.out_window(("r", i, q), SB.offset[i] + q * dx, dx, dx, e)
        for (y, q, e, y2, q2, e2, f, mat) in trans:
            if y2 != y:
                raise ValueError("right_scalar needs a J-preserving operator")
            for i in self.by_y.get(y, ()):
                dx = SB.sub[i][2]
                src = Operand(capi.BUF_IN, SB.offset[i] + q * dx, dx, capi.OP_N)
                if mat is None:
                    pb.add(win[(i, q2)], src, IDENT, 0, f)
                else:
                    mo, mld = pb.op(mat[0], mat[1])
                    op = capi.OP_T if mat[2] == capi.OP_N else capi.OP_CONJ
                    pb.add(win[(i, q2)], src, Operand(capi.BUF_OPS, mo, mld, op), e, f)
        return self._value(pb)

    def cross(self, terms):
        """terms: [(g, TL, TR, k2, y_odd)], rank-k2 parts of the two sides
        coupled to the total spin."""
        SB, L, R = self.SB, self.L, self.R
        pb = PlanBuilder(resolver=self._resolve)
        win = {}
        for i, (x, y, dx, dy) in enumerate(SB.sub):
            for (a, s, p, d) in L.sectors[x][3]:
                for (b, t, q, e) in R.sectors[y][3]:
                    win[(i, p, q)] = pb.out_window((i, p, q), SB.offset[i] + p + q * dx, dx, d, e)
        for (g, TL, TR, k2, y_odd) in terms:
            cross_terms(pb, win, L, R, SB, TL, TR, g, True, k2, y_odd)
        return self._value(pb)

    def close(self):
        for v in (self.vec, self.phi):
            if v is not None:
                v.close()
        self.vec = self.phi = None


def _site_scalar(basis, blk_sectors, op):
    """m-independent transitions of a scalar site operator on an enlarged
    basis (the reduced element divided by sqrt(2J+1))."""
    out = []
    for (x, p, d, x2, p2, d2, f, mat) in side_transitions(basis, blk_sectors, ("site", op, False)):
        out.append((x, p, d, x2, p2, d2, f / sqrt(basis.sectors[x][1] + 1), mat))
    return out


def measure_state(ctx, psi, bases, s1, N, sys_truns, env_truns, i0, js, kinds=("SS", "nn", "cc"),
                  density=True, prec=capi.PREC_F64):
    """Measurements of psi on the superblock [A = 1..s1-1, s1 | B, s2 = s1+1].

    sys_truns[L] built the sys block of length L (sites 1..L), env_truns[m]
    the env block of length m (sites N..N-m+1). i0: reference site in A or
    s1; js: sites j > i0. Returns {"density": {site: <n>}, kind: {(i0, j):
    value}, "stats": {...}}."""
    s2 = s1 + 1
    Lb, Le = s1 - 1, N - s2
    if not 1 <= i0 <= s1:
        raise ValueError(f"reference site {i0} must lie in the sys block or be its site ({s1})")
    js = sorted(set(int(j) for j in js))
    if any(j <= i0 or j > N for j in js):
        raise ValueError("correlated sites must follow the reference site")
    if Lb < 1 or Le < 1:
        raise ValueError("the measured superblock needs non-empty blocks")
    names = _needed(kinds)
    so = site_operators()
    t0 = time.perf_counter()
    # ---- sys side: sites 1..Lb
    sysc = Carrier(ctx, sys_truns, lambda L: L, prec)
    envc = Carrier(ctx, env_truns, lambda m: N - m + 1, prec)
    try:
        sys_names = {j: (names if j == i0 else (["n"] if density else [])) for j in range(1, Lb + 1)}
        sysc.start(sys_names[1])
        for L in range(2, Lb + 1):
            pair_specs = []
            if i0 < L and L in js:
                for kind in kinds:
                    pair_specs.append((("corr", kind, L), CORRELATORS[kind], i0))
            sysc.grow(new_site_terms(L, sys_names[L], pair_specs))
        # ---- env side: sites N..s2+1
        env_sites = {N - m + 1 for m in range(1, Le + 1)}
        env_names = {}
        for j in env_sites:
            nm = set(["n"] if density else [])
            if j in js:
                nm |= set(names)
            env_names[j] = sorted(nm)
        envc.start(env_names[N])
        for m in range(2, Le + 1):
            j = N - m + 1
            envc.grow(new_site_terms(j, env_names[j]))
        transport_s = time.perf_counter() - t0
        ev = SuperblockExpectation(ctx, psi, bases, sysc, envc, prec)
        L, R = ev.L, ev.R
        A, B = sysc.ops, envc.ops
        out = {"density": {}}
        try:
            if density:
                for j in range(1, N + 1):
                    if j <= Lb:
                        v = ev.left_scalar(scalar_transitions(L, ev.A_sec, ham=A[("op", j, "n")]))
                    elif j == s1:
                        v = ev.left_scalar(_site_scalar(L, ev.A_sec, so["n"]))
                    elif j == s2:
                        v = ev.right_scalar(_site_scalar(R, ev.B_sec, so["n"]))
                    else:
                        v = ev.right_scalar(scalar_transitions(R, ev.B_sec, ham=B[("op", j, "n")]))
                    out["density"][j] = v
            for kind in kinds:
                prods = CORRELATORS[kind]
                res = {}
                for j in js:
                    if j <= Lb:
                        v = ev.left_scalar(scalar_transitions(L, ev.A_sec, ham=A[("corr", kind, j)]))
                    elif j == s1:
                        pairs = [(g, A[("op", i0, pn)], ptil, so[qn], qtil) for (g, pn, ptil, qn, qtil) in prods]
                        v = ev.left_scalar(scalar_transitions(L, ev.A_sec, pairs=pairs))
                    else:
                        terms = []
                        for (g, pn, ptil, qn, qtil) in prods:
                            if i0 <= Lb:
                                TL = side_transitions(L, ev.A_sec, ("blk", A[("op", i0, pn)], ptil))
                            else:
                                TL = side_transitions(L, ev.A_sec, ("site", so[pn], ptil))
                            if j == s2:
                                TR = side_transitions(R, ev.B_sec, ("site", so[qn], qtil))
                            else:
                                TR = side_transitions(R, ev.B_sec, ("blk", B[("op", j, qn)], qtil))
                            terms.append((g, TL, TR, so[pn].k2, so[qn].odd))
                        v = ev.cross(terms)
                    res[(i0, j)] = v
                out[kind] = res
            out["stats"] = {"transport_s": transport_s, "transport_flops": sysc.flops + envc.flops,
                            "transport_device_ms": sysc.ms + envc.ms, "expect_plans": ev.plans,
                            "expect_s": ev.seconds}
        finally:
            ev.close()
    finally:
        sysc.close()
        envc.close()
    return out


def reference_sites(Lx, Ny):
    """Reference site xi1 and the sites the reference correlates with it
    (main.f90:255-256; measure.f90:2582-2595 and :2610: the sites of the
    same leg, i0 + k Ny, up to N - N/6 + Ny)."""
    N = Lx * Ny
    i0 = max(Lx // 4 * Ny + 2 - Ny, 1)
    last = min(N - N // 6 + Ny, N)
    return i0, list(range(i0 + Ny, last + 1, Ny))


# ---------------------------------------------------------------------
# singlet pair correlators (Get_mSCOP_Operator_Cor0, measure.f90:2683-2834)
#
# D+_ab = (c+_a,up c+_b,dn - c+_a,dn c+_b,up) / sqrt2 = [c+_a x c+_b]^0 and
# D_cd = -[c~_c x c~_d]^0, both symmetric in their two sites, so the
# coupling may follow whichever site joins a block first.

def measure_pairs(ctx, psi, bases, s1, N, sys_truns, env_truns, bond, partners, prec=capi.PREC_F64):
    """<D+_bond D_cd> for the bonds (c, d) of `partners`; bond and every
    partner lie in the sys block A = 1..s1-1, or bond in A and the partner
    in the env block B = N..s1+2 (the cases the reference measures:
    measure.f90:2762 and :2804). Returns {(c, d): value}."""
    s2 = s1 + 1
    Lb, Le = s1 - 1, N - s2
    a, b = sorted(int(v) for v in bond)
    if b > Lb:
        raise ValueError(f"pair ({a}, {b}) must lie in the sys block 1..{Lb}")
    in_a, in_b = [], []
    for (c, d) in partners:
        c, d = sorted((int(c), int(d)))
        if d <= Lb and c > b:
            in_a.append((c, d))
        elif c > s2 and d <= N:
            in_b.append((c, d))
        else:
            raise ValueError(f"partner bond ({c}, {d}) must follow ({a}, {b}) inside one block")
    so = site_operators()
    cdag = so["c"]
    sysc = Carrier(ctx, sys_truns, lambda L: L, prec)
    envc = Carrier(ctx, env_truns, lambda m: N - m + 1, prec)
    out = {}
    try:
        firsts = {c for (c, d) in in_a}
        seconds = {d for (c, d) in in_a}
        # sys side: c+_a carried, D+_ab formed when b joins, D+_ab c~_c when
        # c joins, [D+_ab c~_c x c~_d]^0 when d joins
        sysc.start(["c"] if a == 1 else [])
        for L in range(2, Lb + 1):
            def terms(basis, sectors, ops, L=L):
                t = {}
                if L == a:
                    t[("op", L, "c")] = (side_transitions(basis, sectors, ("site", cdag, False)), (1, 1))
                if L == b:
                    t[("P",)] = (pair0_transitions(basis, sectors, 1.0, ops[("op", a, "c")], False, cdag, False), (2, 0))
                if L in firsts:
                    t[("Y", L)] = (scalar_times_site(basis, sectors, ops[("P",)], cdag, True), (1, 1))
                if L in seconds:
                    for (c, d) in in_a:
                        if d == L:
                            t[("Z", c, d)] = (pair0_transitions(basis, sectors, -1.0, ops[("Y", c)], False, cdag, True),
                                              (0, 0))
                return t
            sysc.grow(terms)
        # env side: c~ pairs [c~_d x c~_c]^0 formed when the lower site joins
        env_first = {d for (c, d) in in_b}
        env_second = {c for (c, d) in in_b}
        envc.start(["c"] if N in env_first else [])
        for m in range(2, Le + 1):
            j = N - m + 1

            def eterms(basis, sectors, ops, j=j):
                t = {}
                if j in env_first:
                    t[("op", j, "c")] = (side_transitions(basis, sectors, ("site", cdag, False)), (1, 1))
                if j in env_second:
                    for (c, d) in in_b:
                        if c == j:
                            t[("Q", c, d)] = (pair0_transitions(basis, sectors, 1.0, ops[("op", d, "c")], True, cdag,
                                                                True), (-2, 0))
                return t
            envc.grow(eterms)
        ev = SuperblockExpectation(ctx, psi, bases, sysc, envc, prec)
        try:
            L, R = ev.L, ev.R
            for (c, d) in in_a:
                out[(c, d)] = ev.left_scalar(scalar_transitions(L, ev.A_sec, ham=sysc.ops[("Z", c, d)]))
            for (c, d) in in_b:
                TL = side_transitions(L, ev.A_sec, ("blk", sysc.ops[("P",)], False))
                TR = side_transitions(R, ev.B_sec, ("blk", envc.ops[("Q", c, d)], False))
                out[(c, d)] = ev.cross([(-1.0, TL, TR, 0, False)])
        finally:
            ev.close()
    finally:
        sysc.close()
        envc.close()
    return out


def reference_pair_partners(Lx, Ny, s1, ij1, ij=0):
    """The bond (xi1, neib(xi1, ij1)) and the partner bonds (sid3,
    neib(sid3, t)), t = 1, 3, 5, that Get_mSCOP_Operator_Cor0 measures with
    it (measure.f90:2750-2830) on the superblock whose sys site is s1:
    partners inside the sys block after the bond, and partners of the legs'
    sites up to xj1 inside the env block. Returns (bond, [(sid3, t, (sid3,
    sid4))])."""
    from .lattice import neighbour
    N = Lx * Ny
    sys_len, env_len = s1, N - s1
    xc1 = max(Lx // 4 * Ny + 2 - Ny, 1)
    xc2 = Lx * 3 // 4 * Ny + 2
    sid1 = xc1
    sid2 = neighbour(sid1, ij1, Lx, Ny)
    if not sid2:
        return (sid1, sid2), []
    hi = max(sid1, sid2)
    out = []
    for sid3 in range(sid1 + Ny + ij, sys_len, Ny):
        if sid3 <= hi or sid3 > xc2:
            continue
        for t in (1, 3, 5):
            sid4 = neighbour(sid3, t, Lx, Ny)
            if not sid4 or min(sid3, sid4) <= hi:
                continue
            if max(sid3, sid4) <= sys_len - 1:
                out.append((sid3, t, (sid3, sid4)))
    start = sys_len + 2 - (sys_len + 2) % Ny + (sid1 - 1) % Ny + 1 + ij
    for sid3 in range(start, xc2 + 1, Ny):
        if N - sid3 + 1 > env_len - 1:
            continue
        for t in (1, 3, 5):
            sid4 = neighbour(sid3, t, Lx, Ny)
            if sid4 and max(N - sid3 + 1, N - sid4 + 1) <= env_len - 1:
                out.append((sid3, t, (sid3, sid4)))
    return (sid1, sid2), out


# ---------------------------------------------------------------------
# measurement files in the reference's text layout

def _f(v, w, d):
    return f"{v:{w}.{d}f}"


def write_reference_files(out, dirpath, Lx, Ny, J2, kept_min, kept_max, n_electrons):
    """elec_density.dat, spin_cor.dat, elec_cor.dat, density_cor.dat and
    Sc_cor_y/_x/_xy.dat with the columns and Fortran formats of the
    reference (measure.f90:2599-2601 format (2i6, 3f14.8); main.f90:276-308
    format (3i8, 3f21.14); measure.f90:2769 format (4i8, 6f21.12, 2i8)).
    The header lines carry the same fields in free format. Returns the
    paths written."""
    import os
    os.makedirs(dirpath, exist_ok=True)
    N = Lx * Ny
    paths = []

    def put(name, lines):
        p = os.path.join(dirpath, name)
        with open(p, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        paths.append(p)

    dens = out.get("density", {})
    if dens:
        lines = [f" Nx= {Lx} Ny= {Ny} N_up= {n_electrons} N_down= 0 M1= {kept_min} M2= {kept_max}"]
        acc = 0.0
        for i in range(1, N + 1):
            acc += dens[i]
            lines.append(f"{(i - 1) // Ny + 1:6d}{(i - 1) % Ny + 1:6d}{_f(dens[i], 14, 8)}{_f(acc, 14, 8)}")
        lines.append("")
        put("elec_density.dat", lines)
    head = f" Nx= {Lx} Ny= {Ny} J2= {J2} kept_max= {kept_max}"
    for kind, name in (("SS", "spin_cor.dat"), ("cc", "elec_cor.dat"), ("nn", "density_cor.dat")):
        vals = out.get(kind)
        if vals is None:
            continue
        lines = [head]
        for (i, j), v in sorted(vals.items()):
            if v == 0.0:
                continue
            cols = f"{i:8d}{j:8d}{(j - 1) // Ny - (i - 1) // Ny:8d}{_f(v, 21, 14)}"
            if kind == "nn":
                cols += _f(dens.get(i, 0.0), 21, 14) + _f(dens.get(j, 0.0), 21, 14)
            else:
                cols += _f(0.0, 21, 14)
            lines.append(cols)
        put(name, lines)
    from .lattice import neighbour
    for ij1, name in ((1, "Sc_cor_y.dat"), (3, "Sc_cor_x.dat"), (5, "Sc_cor_xy.dat")):
        rec = out.get("pairs", {}).get(ij1)
        if rec is None:
            continue
        sid1, sid2 = rec["bond"]
        lines = [f" Nx= {Lx} Ny= {Ny} Jd= 0 t= 1 Jchi= 0 N_up= {n_electrons} N_down= 0 {kept_min} {kept_max}"
                 f" bond {ij1}"]
        for sid3 in sorted({s for (s, _) in rec["values"]}):
            v = [rec["values"].get((sid3, t), 0.0) for t in (1, 3, 5)]
            sid4 = neighbour(sid3, 5, Lx, Ny)
            cols = f"{sid1:8d}{sid3:8d}{(sid3 - sid1 + N) % Ny:8d}{(sid3 - 1) // Ny - (sid1 - 1) // Ny:8d}"
            cols += "".join(_f(abs(x), 21, 12) for x in v) + "".join(_f(float(np.angle(x)), 21, 12) for x in v)
            cols += f"{sid2:8d}{sid4:8d}"
            lines.append(cols)
        put(name, lines)
    return paths


def measure_from_disk(ctx, path, Lx, Ny, i0=None, js=None, kinds=("SS", "nn", "cc"), density=True,
                      pair_bonds=(1, 3, 5), prec=capi.PREC_F64):
    """Measurement replay (restart1 = 4 of the reference, main.f90:102-107):
    the stored ground state eigvec1 and the stored truncations of both sides
    (systruns<L>, envtruns<L>, persist.py layouts) are read, the blocks'
    sector structure is rebuilt from the truncations, and the state is
    measured as measure_state / measure_pairs do."""
    import os
    from types import SimpleNamespace
    from . import persist
    from .blocks import SuperBasis, TotalBasis
    N = Lx * Ny
    sys_len, env_len, subs = persist.wave_from_disk(os.path.join(path, persist.wave_name(1)))
    Lb, Le = sys_len - 1, env_len - 1
    if Lb + Le + 2 != N:
        raise ValueError(f"stored wave function covers {Lb + Le + 2} sites, the lattice {N}")

    def chain(side, Lend):
        sectors = [tuple(s) for s in SITE_SECTORS]
        truns = {}
        for L in range(2, Lend + 1):
            basis = TotalBasis(sectors)
            _, tr = persist.truns_from_disk(os.path.join(path, persist.trun_name(side, L)), basis, env=(side == "env"))
            truns[L] = tr
            sectors = [(basis.sectors[x][0], basis.sectors[x][1], tr.U[x].shape[1]) for x in tr.order]
        return truns, sectors

    sys_truns, A_sec = chain("sys", Lb)
    env_truns, B_sec = chain("env", Le)
    sys_bs, env_bs = TotalBasis(A_sec), TotalBasis(B_sec)
    (sq, eq, _) = subs[0]
    sb = SuperBasis(sys_bs, env_bs, sq[0] + eq[0], 0)
    psi = ctx.vec_from(persist.psi_in_super_basis(subs, sb, sys_bs, env_bs))
    bases = SimpleNamespace(super=sb, sys_bs=sys_bs, env_bs=env_bs)
    d_i0, d_js = reference_sites(Lx, Ny)
    i0 = d_i0 if i0 is None else int(i0)
    js = [j for j in (d_js if js is None else js) if j > i0]
    try:
        out = measure_state(ctx, psi, bases, Lb + 1, N, sys_truns, env_truns, i0, js, kinds=kinds, density=density,
                            prec=prec)
        out["pairs"] = {}
        for ij1 in pair_bonds:
            bond, partners = reference_pair_partners(Lx, Ny, Lb + 1, ij1)
            vals = {}
            if partners:
                got = measure_pairs(ctx, psi, bases, Lb + 1, N, sys_truns, env_truns, bond,
                                    [p for (_, _, p) in partners], prec=prec)
                vals = {(sid3, t): got[tuple(sorted(p))] for (sid3, t, p) in partners}
            out["pairs"][ij1] = {"bond": bond, "values": vals}
    finally:
        psi.close()
    out["i0"], out["js"], out["l_end"] = i0, js, Lb
    return out
