"""Host packing for the native SU(2) planner (tsc_superblock_tasks).

The Python side keeps the physics bookkeeping that is cheap (bond list,
hmerge-style merging of partner operators); the library enumerates the
sub-block transitions, 6j factors and the contribution lists in C++.
"""
import ctypes

import numpy as np

from . import capi

c_int32, c_int64, c_double, P = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.POINTER


class Sector(ctypes.Structure):
    _fields_ = [("n", c_int32), ("twos", c_int32), ("dim", c_int32)]


class OpBlock(ctypes.Structure):
    _fields_ = [("bra", c_int32), ("ket", c_int32), ("off", c_int64)]


class OpDesc(ctypes.Structure):
    _fields_ = [("side", c_int32), ("up_dif", c_int32), ("k2", c_int32), ("nblk", c_int32), ("blk", P(OpBlock))]


class PairDesc(ctypes.Structure):
    _fields_ = [("coef", c_double), ("block_op", c_int32), ("block_tilde", c_int32), ("site_op", c_int32),
                ("site_tilde", c_int32)]


class TermDesc(ctypes.Structure):
    _fields_ = [("coef", c_double), ("l_kind", c_int32), ("l_op", c_int32), ("l_tilde", c_int32),
                ("r_kind", c_int32), ("r_op", c_int32), ("r_tilde", c_int32)]


class SbDesc(ctypes.Structure):
    _fields_ = [("nsys", c_int32), ("nenv", c_int32), ("sys", P(Sector)), ("env", P(Sector)),
                ("arena_len", c_int64), ("nops", c_int32), ("ops", P(OpDesc)), ("sys_ham", P(OpDesc)),
                ("env_ham", P(OpDesc)), ("nlpairs", c_int32), ("lpairs", P(PairDesc)), ("nrpairs", c_int32),
                ("rpairs", P(PairDesc)), ("nterms", c_int32), ("terms", P(TermDesc)), ("ntot", c_int32),
                ("s2tot", c_int32)]


class SiteOps(ctypes.Structure):
    _fields_ = [("site", c_int32), ("c", P(OpDesc)), ("S", P(OpDesc)), ("n", P(OpDesc))]


class ModelDesc(ctypes.Structure):
    _fields_ = [("nsec", c_int32), ("sec", P(Sector)), ("nsites", c_int32), ("sites", P(c_int32)),
                ("ham", P(OpDesc)), ("nkeep", c_int32), ("keep", P(SiteOps))]


class Bond(ctypes.Structure):
    _fields_ = [("i", c_int32), ("j", c_int32), ("t", c_double), ("J", c_double)]


_SITE_ID = {"c": 0, "S": 1, "n": 2}


def _bind():
    L = capi.lib()
    if getattr(L, "_tsc_native_bound", False):
        return L
    L.tsc_superblock_tasks.argtypes = [P(SbDesc), P(ctypes.c_void_p)]
    L.tsc_superblock_tasks.restype = ctypes.c_int
    L.tsc_tasks_info.argtypes = [ctypes.c_void_p, P(c_int64)]
    L.tsc_tasks_info.restype = ctypes.c_int
    L.tsc_model_tasks.argtypes = [P(ModelDesc), P(ModelDesc), c_int32, c_int32, c_int32, P(Bond), c_int32,
                                  c_int32, ctypes.c_void_p, c_int64, P(ctypes.c_void_p)]
    L.tsc_model_tasks.restype = ctypes.c_int
    L.tsc_renorm_tasks.argtypes = [P(ModelDesc), c_int32, c_int32, P(Bond), c_int32, P(c_int32), c_int32,
                                   P(c_int32), P(c_int32), ctypes.c_void_p, c_int64, P(ctypes.c_void_p)]
    L.tsc_renorm_tasks.restype = ctypes.c_int
    L.tsc_tasks_layout.argtypes = [ctypes.c_void_p, P(c_int64)] + [ctypes.c_void_p] * 4
    L.tsc_tasks_layout.restype = ctypes.c_int
    L.tsc_tasks_lengths.argtypes = [ctypes.c_void_p, P(c_int64)]
    L.tsc_tasks_lengths.restype = ctypes.c_int
    L.tsc_tasks_copy_arena.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.tsc_tasks_copy_arena.restype = ctypes.c_int
    L.tsc_tasks_flops.argtypes = [ctypes.c_void_p, P(c_double)]
    L.tsc_tasks_flops.restype = ctypes.c_int
    L.tsc_tasks_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 7
    L.tsc_tasks_copy.restype = ctypes.c_int
    L.tsc_tasks_plan.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_int64, P(ctypes.c_void_p)]
    L.tsc_tasks_plan.restype = ctypes.c_int
    L.tsc_tasks_plan_dev.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_void_p)]
    L.tsc_tasks_plan_dev.restype = ctypes.c_int
    L.tsc_tasks_destroy.argtypes = [ctypes.c_void_p]
    L.tsc_tasks_destroy.restype = ctypes.c_int
    L._tsc_native_bound = True
    return L


class NativeTasks:
    """Task arrays of one superblock built by the library's planner.

    step: a finalized resident.StepArena holding every operator block (the
    descriptors then carry its offsets and the plan borrows its device
    vector; nothing is packed on the host); terms: a precomputed sb.terms()."""

    def __init__(self, sb, step=None, terms=None):
        L = _bind()
        Lps, Rps, lr = sb.terms() if terms is None else terms
        site_id = {id(o): _SITE_ID[k] for k, o in sb.site_ops.items()}
        arena, self._keep = [], []
        self.step = step
        self.arena_len = step.len if step is not None else 0
        op_index = {}
        descs = []

        def add_op(op, side):
            key = (id(op), side)
            if key in op_index:
                return op_index[key]
            items = list(op.blocks.items())
            blks = (OpBlock * max(len(items), 1))()
            for n, ((b, k), m) in enumerate(items):
                if step is not None:
                    blks[n].bra, blks[n].ket, blks[n].off = b, k, step.offset(m)[0]
                    continue
                m = np.asarray(m, np.complex128)
                blks[n].bra, blks[n].ket, blks[n].off = b, k, self.arena_len
                arena.append(m.ravel(order="F"))
                self.arena_len += m.size
            self._keep.append(blks)
            descs.append(OpDesc(side, op.up_dif, op.k2, len(items), blks))
            op_index[key] = len(descs) - 1
            return op_index[key]

        ham_sys = add_op(sb.sys.ham, 0)
        ham_env = add_op(sb.env.ham, 1)

        def pairs(lst, side):
            arr = (PairDesc * max(len(lst), 1))()
            for n, (g, T, ttil, U, util) in enumerate(lst):
                arr[n] = PairDesc(g, add_op(T, side), int(ttil), site_id[id(U)], int(util))
            return arr, len(lst)

        lp, nlp = pairs(Lps, 0)
        rp, nrp = pairs(Rps, 1)
        terms = (TermDesc * max(len(lr), 1))()
        for n, (g, ls, rs) in enumerate(lr):
            lk, lo = (1, add_op(ls[1], 0)) if ls[0] == "blk" else (2, site_id[id(ls[1])])
            rk, ro = (1, add_op(rs[1], 1)) if rs[0] == "blk" else (2, site_id[id(rs[1])])
            terms[n] = TermDesc(g, lk, lo, int(ls[2]), rk, ro, int(rs[2]))
        ops = (OpDesc * len(descs))(*descs)
        sys_s = (Sector * len(sb.sys.sectors))(*[Sector(*s) for s in sb.sys.sectors])
        env_s = (Sector * len(sb.env.sectors))(*[Sector(*s) for s in sb.env.sectors])
        d = SbDesc(len(sb.sys.sectors), len(sb.env.sectors), sys_s, env_s, self.arena_len, len(descs), ops,
                   ctypes.pointer(ops[ham_sys]), ctypes.pointer(ops[ham_env]), nlp, lp, nrp, rp, len(lr), terms,
                   sb.super.ntot, sb.super.s2tot)
        h = ctypes.c_void_p()
        capi.check(L.tsc_superblock_tasks(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.arena = np.concatenate(arena) if arena else np.zeros(1, np.complex128)
        if step is not None:
            self.arena = None
        cnt = (c_int64 * 9)()
        capi.check(L.tsc_tasks_info(h, cnt))
        self.counts = list(cnt)
        self.mid_len, self.psi_len = cnt[4], cnt[5]

    def flops(self):
        """(flops as planned, per-product minimum without sharing)."""
        out = (c_double * 2)()
        capi.check(_bind().tsc_tasks_flops(self.h, out))
        return out[0], out[1]

    def arrays(self):
        L = _bind()
        nc1, no1, nc2, no2, _, _, ns, nl, nr = self.counts
        c1 = np.zeros(nc1, capi.CONTRIB_DTYPE)
        o1 = np.zeros(no1, capi.OUTBLOCK_DTYPE)
        c2 = np.zeros(nc2, capi.CONTRIB_DTYPE)
        o2 = np.zeros(no2, capi.OUTBLOCK_DTYPE)
        sxy = np.zeros(4 * max(ns, 1), np.int32)
        ls = np.zeros(3 * max(nl, 1), np.int32)
        rs = np.zeros(3 * max(nr, 1), np.int32)
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        capi.check(L.tsc_tasks_copy(self.h, ptr(c1), ptr(o1), ptr(c2), ptr(o2), ptr(sxy), ptr(ls), ptr(rs)))
        return c1, o1, c2, o2

    def plan(self, ctx):
        L = _bind()
        h = ctypes.c_void_p()
        pl = capi.Plan.__new__(capi.Plan)
        pl.ops_vec = None
        if self.step is not None:
            capi.check(L.tsc_tasks_plan_dev(ctx.h, self.h, self.step.vec.h, ctypes.byref(h)))
            pl.ops_vec = self.step.vec
            pl.ops_elems = int(self.step.vec.n)
        else:
            capi.check(L.tsc_tasks_plan(ctx.h, self.h, self.arena.ctypes.data_as(ctypes.c_void_p), self.arena.size,
                                        ctypes.byref(h)))
            pl.ops_elems = int(self.arena.size)
        pl.ctx, pl.h = ctx, h
        pl.in_elems = pl.out_elems = int(self.psi_len)
        return pl

    def close(self):
        if getattr(self, "h", None):
            _bind().tsc_tasks_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ModelPack:
    """A TotalModel with host blocks packed for the library: one arena
    (shared by several models), sector list, operator descriptors."""

    def __init__(self, model, arena, side):
        self._keep = []
        self.sec = (Sector * max(len(model.sectors), 1))(*[Sector(*s) for s in model.sectors])
        self.sites = (c_int32 * max(len(model.sites), 1))(*model.sites)

        def op(tb):
            items = list(tb.blocks.items())
            blks = (OpBlock * max(len(items), 1))()
            for n, ((b, k), m) in enumerate(items):
                m = np.asarray(m, np.complex128)
                blks[n].bra, blks[n].ket, blks[n].off = b, k, arena.add(m)
            self._keep.append(blks)
            d = OpDesc(side, tb.up_dif, tb.k2, len(items), blks)
            self._keep.append(d)
            return ctypes.pointer(d)

        self.ham = op(model.ham)
        keep = sorted(model.ops)
        self.keep = (SiteOps * max(len(keep), 1))()
        for n, i in enumerate(keep):
            o = model.ops[i]
            self.keep[n] = SiteOps(i, op(o["c"]), op(o["S"]), op(o["n"]))
        self.desc = ModelDesc(len(model.sectors), self.sec, len(model.sites), self.sites, self.ham, len(keep),
                              self.keep)


class _Arena:
    def __init__(self):
        self.parts, self.n = [], 0

    def add(self, m):
        off = self.n
        self.parts.append(m.ravel(order="F"))
        self.n += m.size
        return off

    def array(self):
        return np.concatenate(self.parts) if self.parts else np.zeros(1, np.complex128)


def _bonds(bonds):
    arr = (Bond * max(len(bonds), 1))()
    for n, (i, j, t, J) in enumerate(bonds):
        arr[n] = Bond(int(i), int(j), float(t), float(J))
    return arr, len(bonds)


class ModelTasks(NativeTasks):
    """Task arrays of a superblock built by the library from the two block
    models and the lattice bonds (tsc_model_tasks: term list and merged
    partner operators in the library)."""

    def __init__(self, sys, env, s1, s2, bonds, ntot, s2tot):
        L = _bind()
        ar = _Arena()
        ps, pe = _ModelPack(sys, ar, 0), _ModelPack(env, ar, 1)
        arena = ar.array()
        b, nb = _bonds(bonds)
        h = ctypes.c_void_p()
        capi.check(L.tsc_model_tasks(ctypes.byref(ps.desc), ctypes.byref(pe.desc), int(s1), int(s2), nb, b,
                                     int(ntot), int(s2tot), arena.ctypes.data_as(ctypes.c_void_p), arena.size,
                                     ctypes.byref(h)))
        self.h = h
        self.step = None
        lens = (c_int64 * 3)()
        capi.check(L.tsc_tasks_lengths(h, lens))
        self.arena = np.empty(max(lens[2], 1), np.complex128)
        capi.check(L.tsc_tasks_copy_arena(h, self.arena.ctypes.data_as(ctypes.c_void_p)))
        self.arena_len = int(lens[2])
        cnt = (c_int64 * 9)()
        capi.check(L.tsc_tasks_info(h, cnt))
        self.counts = list(cnt)
        self.mid_len, self.psi_len = cnt[4], cnt[5]


class RenormTasks(NativeTasks):
    """Enlargement + truncation of a block built by the library
    (tsc_renorm_tasks); in_len / out_len and the output layout."""

    def __init__(self, blk, new_site, bonds, keep_sites, order, kept):
        L = _bind()
        ar = _Arena()
        pb = _ModelPack(blk, ar, 0)
        arena = ar.array()
        b, nb = _bonds(bonds)
        ks = (c_int32 * max(len(keep_sites), 1))(*keep_sites)
        od = (c_int32 * max(len(order), 1))(*order)
        kp = (c_int32 * max(len(kept), 1))(*kept)
        h = ctypes.c_void_p()
        capi.check(L.tsc_renorm_tasks(ctypes.byref(pb.desc), int(new_site), nb, b, len(keep_sites), ks, len(order),
                                      od, kp, arena.ctypes.data_as(ctypes.c_void_p), arena.size, ctypes.byref(h)))
        self.h = h
        self.step = None
        lens = (c_int64 * 3)()
        capi.check(L.tsc_tasks_lengths(h, lens))
        self.in_len, self.out_len = int(lens[0]), int(lens[1])
        self.arena = np.empty(max(lens[2], 1), np.complex128)
        capi.check(L.tsc_tasks_copy_arena(h, self.arena.ctypes.data_as(ctypes.c_void_p)))
        self.arena_len = int(lens[2])
        cnt = (c_int64 * 9)()
        capi.check(L.tsc_tasks_info(h, cnt))
        self.counts = list(cnt)
        self.mid_len = cnt[4]
        c3 = (c_int64 * 3)()
        capi.check(L.tsc_tasks_layout(h, c3, None, None, None, None))
        nw, ns, ne = c3
        lay = np.zeros(3 * max(nw, 1), np.int32)
        off = np.zeros(max(nw, 1), np.int64)
        nsec = np.zeros(3 * max(ns, 1), np.int32)
        esec = np.zeros(3 * max(ne, 1), np.int32)
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        capi.check(L.tsc_tasks_layout(h, c3, ptr(lay), ptr(off), ptr(nsec), ptr(esec)))
        self.layout = [tuple(int(v) for v in lay[3 * w:3 * w + 3]) for w in range(nw)]
        self.offsets = [int(v) for v in off[:nw]]
        self.new_sectors = [tuple(int(v) for v in nsec[3 * q:3 * q + 3]) for q in range(ns)]
        self.enlarged_sectors = [tuple(int(v) for v in esec[3 * q:3 * q + 3]) for q in range(ne)]
