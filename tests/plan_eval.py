"""Host evaluation of a PlanBuilder's task arrays (test helper: checks the
plan builder independently of the device kernels)."""
import numpy as np

from tsc_tri_t1t2j1j2_amd import capi


def apply_numpy(builder, vin, out_len):
    """Host evaluation of the same task arrays (used by tests to check the
    plan builder independently of the device kernels)."""
    ops = builder.arena_array()
    c1, o1, c2, o2 = builder.arrays()
    mid = np.zeros(max(builder.mid_len, 1), np.complex128)
    out = np.zeros(max(out_len, 1), np.complex128)
    bufs = {capi.BUF_OPS: ops, capi.BUF_IN: vin, capi.BUF_MID: mid, capi.BUF_OUT: out}

    def operand(buf, off, ld, op, rows, cols):
        tr = op in (capi.OP_T, capi.OP_C)
        sr, sc = (cols, rows) if tr else (rows, cols)
        x = bufs[buf][off:off + ld * (sc - 1) + sr] if sr and sc else np.zeros(0, complex)
        M = np.zeros((sr, sc), complex)
        for j in range(sc):
            M[:, j] = x[j * ld:j * ld + sr]
        if op == capi.OP_T:
            M = M.T
        elif op == capi.OP_C:
            M = M.conj().T
        elif op == capi.OP_CONJ:
            M = M.conj()
        return M

    for stage, (cs, os_) in enumerate(((c1, o1), (c2, o2))):
        for ob in os_:
            m, n = int(ob["m"]), int(ob["n"])
            acc = np.zeros((m, n), complex)
            for c in cs[ob["cbeg"]:ob["cend"]]:
                coef = c["coef_re"] + 1j * c["coef_im"]
                if c["b_op"] == capi.OP_I:
                    acc += coef * operand(c["a_buf"], c["a_off"], c["lda"], c["a_op"], m, n)
                elif c["a_op"] == capi.OP_I:
                    acc += coef * operand(c["b_buf"], c["b_off"], c["ldb"], c["b_op"], m, n)
                else:
                    k = int(c["k"])
                    acc += coef * (operand(c["a_buf"], c["a_off"], c["lda"], c["a_op"], m, k)
                                   @ operand(c["b_buf"], c["b_off"], c["ldb"], c["b_op"], k, n))
            tgt = bufs[ob["c_buf"]]
            off, ldc = int(ob["c_off"]), int(ob["ldc"])
            for j in range(n):
                tgt[off + j * ldc:off + j * ldc + m] = acc[:, j]
    return out[:out_len]
