"""Superblock H.psi, operator merging and block renormalisation as grouped
contraction plans.

Replaces, for the live t-J branch (add_op = 1, no chiral ring term):
  model_wave          hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:6915-7393
  hmerge              get_sys.f90:7398-7626
  the 14 H.psi kernels (block_vec_dia ... site_vec_site_ndia) get_sys.f90:1783-5204
  Get_Hamiltonian_Sys / _Env with the enlargement routines and
  update_trun_dia / _ndia   get_sys.f90:4-592, 776-1057; block.f90:2495-3687

Superblock = [sys block A, sys site s1 | env block B, env site s2]; the
fermion order is A, s1, B, s2 and the enlarged blocks are L = (A x s1),
R = (B x s2), each coupled block first. psi sectors follow the super basis;
a psi sector (x, y) is a d_L(x) x d_R(y) column-major matrix whose rows are
the sub-blocks (a, s) of L sector x and columns the sub-blocks (b, t) of R
sector y, in Get_basis order.

A bond between parts P before Q contributes
  -t sum_s (c+_i c_j + h.c.) = t sqrt2 ([c+_P x c~_Q]^0 + [c~_P x c+_Q]^0)
   J S_i.S_j = -sqrt3 J [S_P x S_Q]^0,   -J/4 n_i n_j = -J/4 [n_P x n_Q]^0.
Operators of a block that couple to several partners are merged first
(sum_j coupling_j O_j), as hmerge does, giving one term per boundary site.
"""
from math import sqrt

import numpy as np

from . import capi
from . import su2
from .blocks import SITE_SECTORS, SuperBasis, TotalBasis, TotalBlock, TotalModel, site_operators
from .planner import IDENT, Operand, PlanBuilder
from .resident import DevMat

SQ2 = sqrt(2.0)
SQ3 = sqrt(3.0)


# ---------------------------------------------------------------------
# operator specs: ('blk', TotalBlock, tilde) on the block part,
# ('site', TotalBlock, tilde) on the site part, None = identity.

def _entries(op, sectors, tilde):
    """[(bra, ket, stored, opcode, phase)] of op, or of its c~ when tilde."""
    if not tilde:
        return [(b, k, m, capi.OP_N, 1.0) for (b, k), m in op.blocks.items()]
    out = []
    for (b, k), m in op.blocks.items():
        # c~ : ket = b, bra = k, matrix = phase * m^H
        out.append((k, b, m, capi.OP_C, su2.ctilde_phase(sectors[b][1], sectors[k][1])))
    return out


def _rank(spec):
    kind, op, tilde = spec
    return op.k2


def _odd(spec):
    return spec is not None and spec[1].odd


def side_transitions(basis, blk_sectors, spec, reduced=False):
    """Transitions of one side for an operator acting on its block or site
    part: [(x, p, d, x2, p2, d2, f, mat)], mat = (key, stored, opcode) or
    None for the identity on the block part; f holds the embedding factor
    (times the site element, c~ phase and fermion sign)."""
    kind, op, tilde = spec
    k2 = op.k2
    out = []
    if kind == "blk":
        for (a2, a, stored, opc, ph) in _entries(op, blk_sectors, tilde):
            ja2, ja = blk_sectors[a2][1], blk_sectors[a][1]
            d2, d = blk_sectors[a2][2], blk_sectors[a][2]
            key = (id(op), a2, a, opc)
            for s, (_, js, _) in enumerate(SITE_SECTORS):
                for (x, p) in basis.where.get((a, s), ()):
                    Jx = basis.sectors[x][1]
                    for (x2, p2) in basis.where.get((a2, s), ()):
                        f = su2.embed_block(ja2, ja, js, basis.sectors[x2][1], Jx, k2) * ph
                        if f != 0.0:
                            out.append((x, p, d, x2, p2, d2, f, (key, stored, opc)))
    else:
        for (s2, s, stored, opc, ph) in _entries(op, SITE_SECTORS, tilde):
            elem = stored[0, 0]
            if opc == capi.OP_C:
                elem = np.conj(elem)
            js2, js = SITE_SECTORS[s2][1], SITE_SECTORS[s][1]
            for a, (Na, ja, da) in enumerate(blk_sectors):
                sgn = -1.0 if (op.odd and Na % 2) else 1.0
                for (x, p) in basis.where.get((a, s), ()):
                    Jx = basis.sectors[x][1]
                    for (x2, p2) in basis.where.get((a, s2), ()):
                        f = su2.embed_site(js2, js, ja, basis.sectors[x2][1], Jx, k2) * elem * ph * sgn
                        if f != 0.0:
                            out.append((x, p, da, x2, p2, da, f, None))
    return out


def scalar_transitions(basis, blk_sectors, ham=None, pairs=(), reduced=False):
    """Scalar (J-preserving) operators of one side: the block Hamiltonian
    (stored with sqrt(2j+1)) and [T(block) x U(site)]^0 pairs with their
    couplings. f is the m-independent element (the superblock coefficient);
    with reduced=True it is the reduced element of the enlarged block."""
    out = []
    if ham is not None:
        for x, (_, J, _, subs) in enumerate(basis.sectors):
            scale = sqrt(J + 1) if reduced else 1.0
            for (a, s, p, d) in subs:
                if (a, a) in ham.blocks:
                    f = scale / sqrt(blk_sectors[a][1] + 1)
                    out.append((x, p, d, x, p, d, f, ((id(ham), a, a, capi.OP_N), ham.blocks[(a, a)], capi.OP_N)))
    for (g, T, ttil, U, util) in pairs:
        k2 = T.k2
        tent = _entries(T, blk_sectors, ttil)
        uent = _entries(U, SITE_SECTORS, util)
        for x, (_, J, _, subs) in enumerate(basis.sectors):
            scale = sqrt(J + 1) if reduced else 1.0
            pos = {(a, s): (p, d) for (a, s, p, d) in subs}
            for (a, s, p, d) in subs:
                Na, ja = blk_sectors[a][0], blk_sectors[a][1]
                js = SITE_SECTORS[s][1]
                sgn = -1.0 if (U.odd and Na % 2) else 1.0
                for (a2, a_, stored, opc, ph) in tent:
                    if a_ != a:
                        continue
                    for (s2, s_, ust, uopc, uph) in uent:
                        if s_ != s or (a2, s2) not in pos:
                            continue
                        elem = ust[0, 0] if uopc == capi.OP_N else np.conj(ust[0, 0])
                        f = su2.scalar_pair(blk_sectors[a2][1], ja, SITE_SECTORS[s2][1], js, J, k2)
                        f *= g * elem * ph * uph * sgn * scale
                        if f != 0.0:
                            p2, d2 = pos[(a2, s2)]
                            out.append((x, p, d, x, p2, d2, f, ((id(T), a2, a, opc), stored, opc)))
    return out


def identity_transitions(basis):
    return [(x, p, d, x, p, d, 1.0, None) for x, (_, _, _, subs) in enumerate(basis.sectors)
            for (a, s, p, d) in subs]


# ---------------------------------------------------------------------
# bonds and merging (hmerge)

def bond_products(t, J):
    """(coupling, P operator name, P tilde, Q name, Q tilde) of one bond."""
    out = []
    if t != 0.0:
        out.append((t * SQ2, "c", False, "c", True))
        out.append((t * SQ2, "c", True, "c", False))
    if J != 0.0:
        out.append((-SQ3 * J, "S", False, "S", False))
        out.append((-0.25 * J, "n", False, "n", False))
    return out


def merged(ops_by_site, weights, name):
    """sum_i w_i O_i[name] (block_add_block_two over the partner sites).
    For blocks resident on the device the sum is left pending; the step's
    StepArena forms it on the device."""
    from .resident import merged_resident, resident
    if any(resident(m) for i, _ in weights for m in ops_by_site[i][name].blocks.values()):
        return merged_resident(ops_by_site, weights, name)
    acc = None
    for i, w in weights:
        o = ops_by_site[i][name]
        if acc is None:
            acc = TotalBlock(o.up_dif, o.k2)
        acc.axpy(w, o)
    return acc


class Superblock:
    """One DMRG step's superblock: blocks, bases, super basis and terms.

    sys/env: TotalModel of A and B; s1, s2: site indices of the two centre
    sites; bonds: [(i, j, t, J)] among the superblock's sites."""

    def __init__(self, sys, env, s1, s2, bonds, ntot, s2tot):
        self.sys, self.env, self.s1, self.s2 = sys, env, s1, s2
        self.sys_bs = TotalBasis(sys.sectors)
        self.env_bs = TotalBasis(env.sectors)
        self.super = SuperBasis(self.sys_bs, self.env_bs, ntot, s2tot)
        self.s2tot = s2tot
        part = {}
        for i in sys.sites:
            part[i] = "A"
        for i in env.sites:
            part[i] = "B"
        part[s1] = "s1"
        part[s2] = "s2"
        order = {"A": 0, "s1": 1, "B": 2, "s2": 3}
        self.classes = {}
        for (i, j, t, J) in bonds:
            if i not in part or j not in part:
                continue
            pi, pj = part[i], part[j]
            if pi == pj:
                continue
            if order[pi] > order[pj]:
                i, j, pi, pj = j, i, pj, pi
            self.classes.setdefault((pi, pj), []).append((i, j, t, J))
        self.site_ops = site_operators()

    # ---- term list ---------------------------------------------------
    def terms(self):
        """[(g, Lspec_or_scalar, Rspec, cut, k2)] where scalar side terms are
        given as ('scalarL', transitions) / ('scalarR', transitions)."""
        A, B = self.sys, self.env
        so = self.site_ops
        L_scalar_pairs, R_scalar_pairs = [], []
        lr = []
        # (A, s1) and (B, s2): scalar within one side, merged over block sites
        for cls, blk, target in ((("A", "s1"), A, L_scalar_pairs), (("B", "s2"), B, R_scalar_pairs)):
            bl = self.classes.get(cls, [])
            if not bl:
                continue
            for name in ("c+c~", "c~c+", "SS", "nn"):
                ws = []
                for (i, j, t, J) in bl:
                    for (g, pn, ptil, qn, qtil) in bond_products(t, J):
                        tag = {("c", False): "c+c~", ("c", True): "c~c+"}.get((pn, ptil), pn + qn)
                        if tag == name:
                            ws.append((i, g, pn, ptil, qn, qtil))
                if not ws:
                    continue
                _, _, pn, ptil, qn, qtil = ws[0]
                M = merged(blk.ops, [(i, g) for (i, g, *_r) in ws], pn)
                target.append((1.0, M, ptil, so[qn], qtil))
        # L-R classes
        for cls in (("A", "B"), ("A", "s2"), ("s1", "B"), ("s1", "s2")):
            bl = self.classes.get(cls, [])
            if not bl:
                continue
            P, Q = cls
            if cls == ("A", "B"):
                # one term per A boundary site with the partners merged in B
                by_i = {}
                for (i, j, t, J) in bl:
                    by_i.setdefault(i, []).append((j, t, J))
                for i, partners in sorted(by_i.items()):
                    for idx in range(4):
                        ws = []
                        spec_p = None
                        for (j, t, J) in partners:
                            prods = bond_products(t, J)
                            for (g, pn, ptil, qn, qtil) in prods:
                                tag = {("c", False): 0, ("c", True): 1}.get((pn, ptil), 2 if pn == "S" else 3)
                                if tag == idx:
                                    ws.append((j, g))
                                    spec_p = (pn, ptil, qn, qtil)
                        if not ws:
                            continue
                        pn, ptil, qn, qtil = spec_p
                        MB = merged(B.ops, ws, qn)
                        lr.append((1.0, ("blk", A.ops[i][pn], ptil), ("blk", MB, qtil)))
                continue
            for idx in range(4):
                ws = []
                spec_p = None
                for (i, j, t, J) in bl:
                    for (g, pn, ptil, qn, qtil) in bond_products(t, J):
                        tag = {("c", False): 0, ("c", True): 1}.get((pn, ptil), 2 if pn == "S" else 3)
                        if tag == idx:
                            ws.append((i, j, g))
                            spec_p = (pn, ptil, qn, qtil)
                if not ws:
                    continue
                pn, ptil, qn, qtil = spec_p
                if P == "A":       # (A, s2): merge over A sites
                    M = merged(A.ops, [(i, g) for (i, j, g) in ws], pn)
                    lr.append((1.0, ("blk", M, ptil), ("site", so[qn], qtil)))
                elif Q == "B":     # (s1, B): merge over B sites
                    M = merged(B.ops, [(j, g) for (i, j, g) in ws], qn)
                    lr.append((1.0, ("site", so[pn], ptil), ("blk", M, qtil)))
                else:              # (s1, s2)
                    g = sum(g for (_, _, g) in ws)
                    lr.append((g, ("site", so[pn], ptil), ("site", so[qn], qtil)))
        return L_scalar_pairs, R_scalar_pairs, lr

    # ---- plan ----------------------------------------------------------
    def plan_builder(self, resolver=None, terms=None):
        """PlanBuilder of model_wave for this superblock (psi in -> H psi).
        resolver: (offset, ld) of a matrix in a device arena the plan will
        borrow (resident blocks); terms: a precomputed self.terms()."""
        L, R, SB = self.sys_bs, self.env_bs, self.super
        A, B = self.sys, self.env
        pb = PlanBuilder(resolver=resolver)
        # every target sub-block window, written exactly once
        win = {}
        for i, (x, y, dx, dy) in enumerate(SB.sub):
            off = SB.offset[i]
            for (a, s, p, d) in L.sectors[x][3]:
                for (b, t, q, e) in R.sectors[y][3]:
                    win[(i, p, q)] = pb.out_window((i, p, q), off + p + q * dx, dx, d, e)
        Lps, Rps, lr = self.terms() if terms is None else terms
        idL = identity_transitions(L)
        idR = identity_transitions(R)
        scalarL = scalar_transitions(L, A.sectors, ham=A.ham, pairs=Lps)
        scalarR = scalar_transitions(R, B.sectors, ham=B.ham, pairs=Rps)
        self._cross(pb, win, scalarL, idR, 1.0, False, 0, False)
        self._cross(pb, win, idL, scalarR, 1.0, False, 0, False)
        for (g, lspec, rspec) in lr:
            TL = side_transitions(L, A.sectors, lspec)
            TR = side_transitions(R, B.sectors, rspec)
            self._cross(pb, win, TL, TR, g, True, lspec[1].k2, rspec[1].odd)
        return pb

    def _cross(self, pb, win, TL, TR, g, cut, k2, y_odd):
        cross_terms(pb, win, self.sys_bs, self.env_bs, self.super, TL, TR, g, cut, k2, y_odd)


def cross_terms(pb, win, L, R, SB, TL, TR, g, cut, k2, y_odd):
    """Contributions of an L-side x R-side product to the windows win[(i, p,
    q)] of psi: TL / TR are transitions of the two sides (identity, scalar or
    rank-k2 parts); cut couples two rank-k2 parts into the total spin;
    y_odd: the R part is fermionic (sign of the L sector's particles)."""
    S2 = SB.s2tot
    byx, byy = {}, {}
    for t in TL:
        byx.setdefault(t[0], []).append(t)
    for t in TR:
        byy.setdefault(t[0], []).append(t)
    for i, (x, y, dx, dy) in enumerate(SB.sub):
        tl, tr = byx.get(x), byy.get(y)
        if not tl or not tr:
            continue
        off = SB.offset[i]
        Jx, Jy = L.sectors[x][1], R.sectors[y][1]
        fsgn = -1.0 if (y_odd and L.sectors[x][0] % 2) else 1.0
        for (_, p, d, x2, p2, d2, fL, Lm) in tl:
            for (_, q, e, y2, q2, e2, fR, Rm) in tr:
                i2 = SB.index.get((x2, y2))
                if i2 is None:
                    continue
                f = g * fL * fR
                if cut:
                    f *= su2.cut_pair(L.sectors[x2][1], Jx, R.sectors[y2][1], Jy, S2, k2) * fsgn
                if f == 0.0:
                    continue
                out = win[(i2, p2, q2)]
                src = Operand(capi.BUF_IN, off + p + q * dx, dx, capi.OP_N)
                if Lm is not None:
                    lo, lld = pb.op(Lm[0], Lm[1])
                    lop = Operand(capi.BUF_OPS, lo, lld, Lm[2])
                if Rm is not None:
                    ro, rld = pb.op(Rm[0], Rm[1])
                    rop = Operand(capi.BUF_OPS, ro, rld, capi.OP_T if Rm[2] == capi.OP_N else capi.OP_CONJ)
                if Lm is not None and Rm is not None:
                    Z = pb.intermediate((Lm[0], i, p, q), lop, src, d2, e, d)
                    pb.add(out, Z, rop, e, f)
                elif Lm is not None:
                    pb.add(out, lop, src, d, f)
                elif Rm is not None:
                    pb.add(out, src, rop, e, f)
                else:
                    pb.add(out, src, IDENT, 0, f)


# ---------------------------------------------------------------------
# block renormalisation (Get_Hamiltonian_Sys/Env + update_trun_*)

class RenormPlan:
    """Plan of one block renormalisation; see renormalize()."""


def renorm_pairs(blk, bonds_to_new):
    """[T(block) x U(site)]^0 pairs of the enlarged block's Hamiltonian: the
    block-site bonds, partner operators merged per operator type."""
    so = site_operators()
    pairs = []
    for name in range(4):
        ws, spec = [], None
        for (i, t, J) in bonds_to_new:
            for (g, pn, ptil, qn, qtil) in bond_products(t, J):
                tag = {("c", False): 0, ("c", True): 1}.get((pn, ptil), 2 if pn == "S" else 3)
                if tag == name:
                    ws.append((i, g))
                    spec = (pn, ptil, qn, qtil)
        if ws:
            pn, ptil, qn, qtil = spec
            pairs.append((1.0, merged(blk.ops, ws, pn), ptil, so[qn], qtil))
    return pairs


def renorm_transitions(blk, new_site, keep_sites, pairs):
    """The part of renormalize_plan that does not depend on the truncation:
    enlarged basis, transitions (embedding factors) of the enlarged block's
    H and of every carried operator, and their (up_dif, k2). Host-only
    bookkeeping on the block's structure (no device calls), so a driver
    may compute it while the superblock is being solved."""
    basis = TotalBasis(blk.sectors)
    so = site_operators()
    # H of the enlarged block: block H + merged block-site bonds
    trans = {"H": scalar_transitions(basis, blk.sectors, ham=blk.ham, pairs=pairs, reduced=True)}
    meta = {"H": (0, 0)}
    for i in keep_sites:
        for name in ("c", "S", "n"):
            if i == new_site:
                spec = ("site", so[name], False)
                op = so[name]
            else:
                spec = ("blk", blk.ops[i][name], False)
                op = blk.ops[i][name]
            trans[(i, name)] = side_transitions(basis, blk.sectors, spec)
            meta[(i, name)] = (op.up_dif, op.k2)
    return basis, trans, meta


def renormalize_plan(blk, new_site, bonds_to_new, keep_sites, U, order, pairs=None, resolver=None, prepared=None):
    """Task arrays computing every operator of blk (x) new_site truncated by U.

    bonds_to_new: [(i, t, J)] with i in blk; keep_sites: sites whose
    operators the new block carries; U: {enlarged sector x: d_x x kept_x};
    order: enlarged sectors of the new block, in their new order; pairs: a
    precomputed renorm_pairs(blk, bonds_to_new); resolver: as in
    Superblock.plan_builder (resident blocks); prepared: the result of
    renorm_transitions for the same block, site, kept sites and pairs."""
    if pairs is None:
        pairs = renorm_pairs(blk, bonds_to_new)
    if prepared is None:
        prepared = renorm_transitions(blk, new_site, keep_sites, pairs)
    basis, trans, meta = prepared
    new_index = {x: n for n, x in enumerate(order)}
    # IN buffer: U_x laid end to end (order of the new block)
    uoff, parts, off = {}, [], 0
    for x in order:
        u = np.asarray(U[x], np.complex128)
        uoff[x] = off
        parts.append(u.ravel(order="F"))
        off += u.size
    uin = np.concatenate(parts) if parts else np.zeros(1, np.complex128)
    pb, windows, out_off = _renorm_tasks(trans, basis, U, order, uoff, resolver)
    # diagonal H blocks no term reaches stay zero (untouched output windows)
    for x in order:
        wk = ("H", x, x)
        if wk not in windows:
            kx = U[x].shape[1]
            windows[wk] = (None, out_off)
            out_off += kx * kx
    rp = RenormPlan()
    rp.builder, rp.uin, rp.out_len, rp.windows, rp.meta = pb, uin, max(out_off, 1), windows, meta
    rp.basis, rp.new_index, rp.order, rp.U = basis, new_index, order, U
    rp.blk, rp.new_site, rp.keep_sites = blk, new_site, keep_sites
    rp.keepalive = (pairs, trans)
    return rp


class ArrayPlan:
    """Task arrays built directly (the PlanBuilder interface the callers use:
    arrays, arena_array, mid_len, build, build_dev)."""

    def __init__(self, c1, o1, c2, o2, mid_len, arena):
        self._arr = (c1, o1, c2, o2)
        self.mid_len = int(mid_len)
        self._arena = arena

    def arrays(self):
        return self._arr

    def arena_array(self):
        return self._arena if self._arena is not None else np.zeros(1, np.complex128)

    def build(self, ctx, in_len, out_len):
        c1, o1, c2, o2 = self._arr
        ops = self.arena_array()
        return capi.Plan(ctx, ops, ops.size, self.mid_len, c1, o1, c2, o2, int(in_len), int(out_len))

    def build_dev(self, ctx, ops_vec, in_len, out_len):
        c1, o1, c2, o2 = self._arr
        return capi.Plan.create_dev(ctx, ops_vec, self.mid_len, c1, o1, c2, o2, int(in_len), int(out_len))


def _renorm_tasks(trans, basis, U, order, uoff, resolver):
    """Task arrays of U_x2^H . T . U_x for every transition, vectorised.

    Output windows (opkey, x2, x) are laid out in order of first appearance,
    as are the stage-1 intermediates T U_x[rows] (one per (matrix, x, p)).
    Returns (plan, windows {(opkey, x2, x): (window id, offset)}, length)."""
    nsec = len(basis.sectors)
    kept = np.full(nsec, -1, np.int64)
    for x in order:
        kept[x] = U[x].shape[1]
    dims = np.array([sec[2] for sec in basis.sectors], np.int64)
    uo = np.zeros(nsec, np.int64)
    for x, o in uoff.items():
        uo[x] = o
    # flatten the transitions; matrix keys to ids, stored matrices to offsets
    opkeys = list(trans.keys())
    rows, mkeys, mats, mid_of = [], {}, [], []
    for ki, opkey in enumerate(opkeys):
        for (x, p, d, x2, p2, d2, f, mat) in trans[opkey]:
            if mat is None:
                mi = -1
            else:
                mi = mkeys.get(mat[0])
                if mi is None:
                    mi = mkeys[mat[0]] = len(mats)
                    mats.append(mat)
            rows.append((ki, x, p, d, x2, p2, d2, mi, f))
    if not rows:
        z = np.zeros(0, capi.CONTRIB_DTYPE)
        return ArrayPlan(z, np.zeros(0, capi.OUTBLOCK_DTYPE), z, np.zeros(0, capi.OUTBLOCK_DTYPE), 0, None), {}, 0
    R = np.array([r[:8] for r in rows], np.int64)
    F = np.array([r[8] for r in rows], np.complex128)
    ki, x, p, d, x2, p2, d2, mi = R.T
    sel = (kept[x] >= 0) & (kept[x2] >= 0) & (F != 0)
    ki, x, p, d, x2, p2, d2, mi, F = ki[sel], x[sel], p[sel], d[sel], x2[sel], p2[sel], d2[sel], mi[sel], F[sel]
    kx, kx2, dx, dx2 = kept[x], kept[x2], dims[x], dims[x2]
    # windows in first-appearance order
    wkey = (ki * nsec + x2) * nsec + x
    uniq, first, inv = np.unique(wkey, return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    wid = rank[inv]
    nw = len(uniq)
    wx = np.empty(nw, np.int64)
    wx2 = np.empty(nw, np.int64)
    wk = np.empty(nw, np.int64)
    wx[wid], wx2[wid], wk[wid] = x, x2, ki
    wsize = kept[wx2] * kept[wx]
    woff = np.concatenate(([0], np.cumsum(wsize)[:-1])) if nw else np.zeros(0, np.int64)
    windows = {(opkeys[int(wk[w])], int(wx2[w]), int(wx[w])): (int(w), int(woff[w])) for w in range(nw)}
    out_len = int(wsize.sum()) if nw else 0
    # matrix operands: arena offsets (resolver) or a plan-owned host arena
    mo = np.zeros(max(len(mats), 1), np.int64)
    mld = np.ones(max(len(mats), 1), np.int64)
    mop = np.zeros(max(len(mats), 1), np.int64)
    arena, alen = [], 0
    for i, (key, stored, opc) in enumerate(mats):
        if resolver is not None:
            o, l = resolver(stored)
        else:
            m = np.asarray(stored, np.complex128)
            o, l = alen, max(m.shape[0], 1)
            arena.append(m.ravel(order="F"))
            alen += m.size
        mo[i], mld[i], mop[i] = o, l, opc
    # stage-1 intermediates T U_x[rows] per (matrix, x, p), first appearance
    hm = mi >= 0
    ikey = (mi[hm] * nsec + x[hm]) * (int(p.max()) + 1 if len(p) else 1) + p[hm]
    iu, ifirst, iinv = np.unique(ikey, return_index=True, return_inverse=True)
    irank = np.empty(len(iu), np.int64)
    irank[np.argsort(ifirst, kind="stable")] = np.arange(len(iu))
    iid = irank[iinv]
    ni = len(iu)
    hidx = np.nonzero(hm)[0]
    rep = np.empty(ni, np.int64)
    rep[iid] = hidx                       # a transition that uses each intermediate
    isize = d2[rep] * kx[rep]
    ioff = np.concatenate(([0], np.cumsum(isize)[:-1])) if ni else np.zeros(0, np.int64)
    mid_len = int(isize.sum()) if ni else 0
    c1 = np.zeros(ni, capi.CONTRIB_DTYPE)
    o1 = np.zeros(ni, capi.OUTBLOCK_DTYPE)
    if ni:
        r = rep
        c1["a_off"], c1["lda"], c1["a_buf"], c1["a_op"] = mo[mi[r]], mld[mi[r]], capi.BUF_OPS, mop[mi[r]]
        c1["b_off"], c1["ldb"], c1["b_buf"], c1["b_op"] = uo[x[r]] + p[r], np.maximum(dx[r], 1), capi.BUF_IN, capi.OP_N
        c1["k"], c1["coef_re"] = d[r], 1.0
        o1["c_off"], o1["ldc"], o1["m"], o1["n"] = ioff, np.maximum(d2[r], 1), d2[r], kx[r]
        o1["c_buf"] = capi.BUF_MID
        o1["cbeg"] = np.arange(ni)
        o1["cend"] = np.arange(ni) + 1
    # stage 2, grouped by window (stable: transition order within a window)
    order2 = np.argsort(wid, kind="stable")
    c2 = np.zeros(len(order2), capi.CONTRIB_DTYPE)
    t = order2
    c2["a_off"], c2["lda"], c2["a_buf"], c2["a_op"] = uo[x2[t]] + p2[t], np.maximum(dx2[t], 1), capi.BUF_IN, capi.OP_C
    mid_row = np.full(len(x), -1, np.int64)
    mid_row[hidx] = ioff[iid]
    withm = hm[t]
    c2["b_off"] = np.where(withm, mid_row[t], uo[x[t]] + p[t])
    c2["ldb"] = np.where(withm, np.maximum(d2[t], 1), np.maximum(dx[t], 1))
    c2["b_buf"] = np.where(withm, capi.BUF_MID, capi.BUF_IN)
    c2["b_op"] = capi.OP_N
    c2["k"] = np.where(withm, d2[t], d[t])
    c2["coef_re"], c2["coef_im"] = F[t].real, F[t].imag
    o2 = np.zeros(nw, capi.OUTBLOCK_DTYPE)
    if nw:
        counts = np.bincount(wid, minlength=nw)
        ends = np.cumsum(counts)
        o2["c_off"], o2["ldc"], o2["m"], o2["n"] = woff, np.maximum(kept[wx2], 1), kept[wx2], kept[wx]
        o2["c_buf"] = capi.BUF_OUT
        o2["cbeg"] = ends - counts
        o2["cend"] = ends
    plan = ArrayPlan(c1, o1, c2, o2, mid_len, np.concatenate(arena) if arena else None)
    return plan, windows, out_len


def renormalize_result(rp, res, arena=None):
    """TotalModel from the output buffer of a renormalisation plan: host
    copies of `res`, or DevMat views of `arena` (the output vector left on
    the device)."""
    basis, new_index, order, U = rp.basis, rp.new_index, rp.order, rp.U
    windows, meta, blk, new_site, keep_sites = rp.windows, rp.meta, rp.blk, rp.new_site, rp.keep_sites
    sectors = [(basis.sectors[x][0], basis.sectors[x][1], U[x].shape[1]) for x in order]
    ops_new = {}
    H = None
    for wk, (wid, o) in windows.items():
        opkey, x2, x = wk
        kx, kx2 = U[x].shape[1], U[x2].shape[1]
        if arena is None:
            m = res[o:o + kx2 * kx].reshape((kx2, kx), order="F").copy()
        else:
            m = DevMat(arena, o, kx2, kx)
        if opkey == "H":
            if H is None:
                H = TotalBlock(0, 0)
            H.blocks[(new_index[x2], new_index[x])] = m
        else:
            i, name = opkey
            d = ops_new.setdefault(i, {})
            if name not in d:
                d[name] = TotalBlock(*meta[opkey])
            d[name].blocks[(new_index[x2], new_index[x])] = m
    if H is None:
        H = TotalBlock(0, 0)
    for n, x in enumerate(order):
        if (n, n) not in H.blocks:
            H.blocks[(n, n)] = np.zeros((U[x].shape[1],) * 2, complex)
    for i in keep_sites:
        d = ops_new.setdefault(i, {})
        for name in ("c", "S", "n"):
            if name not in d:
                d[name] = TotalBlock(*meta[(i, name)])
    return TotalModel(blk.sites + [new_site], sectors, H, ops_new)


def renormalize(ctx, blk, new_site, bonds_to_new, keep_sites, U, order, prec=capi.PREC_F64):
    """Get_Hamiltonian_Sys/Env + update_trun_*: the new block on the device."""
    rp = renormalize_plan(blk, new_site, bonds_to_new, keep_sites, U, order)
    plan = rp.builder.build(ctx, rp.uin.size, rp.out_len)
    res = plan.apply_host(rp.uin, prec)
    plan.close()
    return renormalize_result(rp, res)
