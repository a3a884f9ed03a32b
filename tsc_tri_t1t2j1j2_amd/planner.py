"""Assembly of grouped two-stage contraction plans.

A plan replaces the reference's per-call loops over sectors, sub-blocks and
quantum-number searches (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:
3175-3296) by flat task arrays built once per superblock:
  * operator matrices are laid end to end in one arena (deduplicated);
  * stage-1 intermediates (op x source window) are registered once and
    shared by every target that needs them;
  * every output window collects all of its contributions, so the device
    kernel writes it exactly once (no atomics).
"""
import numpy as np

from . import capi


class Operand:
    """(buffer, element offset, leading dimension, op code)."""

    __slots__ = ("buf", "off", "ld", "op")

    def __init__(self, buf, off, ld, op):
        self.buf, self.off, self.ld, self.op = int(buf), int(off), int(ld), int(op)


IDENT = None


class PlanBuilder:
    def __init__(self, resolver=None):
        # resolver(mat) -> (offset, ld) in a device arena the plan borrows;
        # None: matrices are copied into the plan's own host arena
        self.resolver = resolver
        self.arena = []
        self.arena_len = 0
        self.arena_index = {}
        self.mid_len = 0
        self.mid_index = {}
        self.c1 = []
        self.o1 = []
        self.out_keys = {}
        self.out_list = []      # [c_off, ldc, m, n, [contribs]]

    # ---- operator arena ------------------------------------------------
    def op(self, key, mat):
        """Arena offset and leading dimension of a stored matrix."""
        if key in self.arena_index:
            return self.arena_index[key]
        if self.resolver is not None:
            res = self.resolver(mat)
            self.arena_index[key] = res
            return res
        m = np.asarray(mat, dtype=np.complex128)
        off = self.arena_len
        self.arena.append(m.ravel(order="F"))
        self.arena_len += m.size
        res = (off, max(m.shape[0], 1))
        self.arena_index[key] = res
        return res

    def arena_array(self):
        if not self.arena:
            return np.zeros(1, np.complex128)
        return np.concatenate(self.arena)

    # ---- outputs -------------------------------------------------------
    def out_window(self, key, c_off, ldc, m, n):
        if key not in self.out_keys:
            self.out_keys[key] = len(self.out_list)
            self.out_list.append([int(c_off), max(int(ldc), 1), int(m), int(n), []])
        return self.out_keys[key]

    def add(self, out_id, a, b, k, coef):
        """C[out] += coef * op_a(A) * op_b(B); a or b may be IDENT."""
        if coef == 0.0:
            return
        if a is IDENT:
            rec = (0, b.off, 1, b.ld, 0, 0, b.buf, capi.OP_I, b.op)
        elif b is IDENT:
            rec = (a.off, 0, a.ld, 1, 0, a.buf, 0, a.op, capi.OP_I)
        else:
            rec = (a.off, b.off, a.ld, b.ld, int(k), a.buf, b.buf, a.op, b.op)
        self.out_list[out_id][4].append(rec + (complex(coef),))

    # ---- stage-1 intermediates ------------------------------------------
    def intermediate(self, key, a, b, m, n, k):
        """Register Z = op_a(A) op_b(B) (m x n) once; returns its operand."""
        if key in self.mid_index:
            return self.mid_index[key]
        off = self.mid_len
        self.mid_len += m * n
        self.c1.append((a.off, b.off, a.ld, b.ld, int(k), a.buf, b.buf, a.op, b.op, 1.0 + 0j))
        self.o1.append((off, max(m, 1), m, n))
        opnd = Operand(capi.BUF_MID, off, max(m, 1), capi.OP_N)
        self.mid_index[key] = opnd
        return opnd

    # ---- finalisation ----------------------------------------------------
    @staticmethod
    def _contrib_array(recs):
        arr = np.zeros(len(recs), dtype=capi.CONTRIB_DTYPE)
        if not recs:
            return arr
        cols = list(zip(*recs))
        for name, col in zip(("a_off", "b_off", "lda", "ldb", "k", "a_buf", "b_buf", "a_op", "b_op"), cols[:9]):
            arr[name] = col
        coef = np.array(cols[9], dtype=np.complex128)
        arr["coef_re"] = coef.real
        arr["coef_im"] = coef.imag
        return arr

    def arrays(self):
        c1 = self._contrib_array(self.c1)
        o1 = np.zeros(len(self.o1), dtype=capi.OUTBLOCK_DTYPE)
        if self.o1:
            off, ld, m, n = (np.array(v) for v in zip(*self.o1))
            o1["c_off"], o1["ldc"], o1["m"], o1["n"] = off, ld, m, n
            o1["c_buf"] = capi.BUF_MID
            o1["cbeg"] = np.arange(len(self.o1))
            o1["cend"] = np.arange(len(self.o1)) + 1
        recs2 = []
        o2 = np.zeros(len(self.out_list), dtype=capi.OUTBLOCK_DTYPE)
        for i, (c_off, ldc, m, n, cl) in enumerate(self.out_list):
            o2[i]["c_off"], o2[i]["ldc"], o2[i]["m"], o2[i]["n"] = c_off, ldc, m, n
            o2[i]["c_buf"] = capi.BUF_OUT
            o2[i]["cbeg"] = len(recs2)
            recs2.extend(cl)
            o2[i]["cend"] = len(recs2)
        return c1, o1, self._contrib_array(recs2), o2

    def build(self, ctx, in_len, out_len):
        c1, o1, c2, o2 = self.arrays()
        ops = self.arena_array()
        return capi.Plan(ctx, ops, ops.size, self.mid_len, c1, o1, c2, o2, int(in_len), int(out_len))

    def build_dev(self, ctx, ops_vec, in_len, out_len):
        """Plan over a device arena (the resolver's offsets)."""
        c1, o1, c2, o2 = self.arrays()
        return capi.Plan.create_dev(ctx, ops_vec, self.mid_len, c1, o1, c2, o2, int(in_len), int(out_len))
