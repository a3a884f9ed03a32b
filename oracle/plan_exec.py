"""CPU execution of grouped contraction task arrays (numpy/BLAS).

Evaluates stage-1 intermediates and stage-2 output windows exactly as the
task semantics define them (C[window] = sum coef * op_a(A) op_b(B)); used
to check plan builders independently of the device and, on a bounded
sample of windows, as the CPU baseline of the H.psi benchmark.
"""
import time

import numpy as np

OP_N, OP_T, OP_C, OP_CONJ, OP_I = 0, 1, 2, 3, 4


def _operand(buf, off, ld, op, rows, cols):
    tr = op in (OP_T, OP_C)
    sr, sc = (cols, rows) if tr else (rows, cols)
    if sr == 0 or sc == 0:
        M = np.zeros((sr, sc), complex)
    else:
        M = np.lib.stride_tricks.as_strided(buf[off:], shape=(sr, sc), strides=(buf.itemsize, buf.itemsize * ld))
    if op == OP_T:
        return M.T
    if op == OP_C:
        return M.conj().T
    if op == OP_CONJ:
        return M.conj()
    return M


def _window(bufs, cs, ob):
    m, n = int(ob["m"]), int(ob["n"])
    acc = np.zeros((m, n), complex)
    flops = 0.0
    for c in cs[int(ob["cbeg"]):int(ob["cend"])]:
        coef = complex(c["coef_re"], c["coef_im"])
        if c["b_op"] == OP_I:
            acc += coef * _operand(bufs[c["a_buf"]], int(c["a_off"]), int(c["lda"]), int(c["a_op"]), m, n)
            flops += 8.0 * m * n
        elif c["a_op"] == OP_I:
            acc += coef * _operand(bufs[c["b_buf"]], int(c["b_off"]), int(c["ldb"]), int(c["b_op"]), m, n)
            flops += 8.0 * m * n
        else:
            k = int(c["k"])
            a = _operand(bufs[c["a_buf"]], int(c["a_off"]), int(c["lda"]), int(c["a_op"]), m, k)
            b = _operand(bufs[c["b_buf"]], int(c["b_off"]), int(c["ldb"]), int(c["b_op"]), k, n)
            acc += coef * (a @ b)
            flops += 8.0 * m * n * k
    return acc, flops


def _store(buf, ob, acc):
    off, ldc = int(ob["c_off"]), int(ob["ldc"])
    m, n = acc.shape
    for j in range(n):
        buf[off + j * ldc:off + j * ldc + m] = acc[:, j]


def execute(ops, c1, o1, c2, o2, mid_len, vin, out_len):
    """Full two-stage evaluation; returns the output vector."""
    mid = np.zeros(max(mid_len, 1), complex)
    out = np.zeros(max(out_len, 1), complex)
    bufs = {0: ops, 1: vin, 2: mid, 3: out}
    for ob in o1:
        acc, _ = _window(bufs, c1, ob)
        _store(mid, ob, acc)
    for ob in o2:
        acc, _ = _window(bufs, c2, ob)
        _store(out, ob, acc)
    return out[:out_len]


def timed_sample(ops, c1, o1, c2, o2, mid_len, vin, seconds, seed=0):
    """Run random stage-2 windows (with the stage-1 intermediates they read)
    until `seconds` of compute; returns (flops, elapsed, windows)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(o2))
    mid = np.zeros(max(mid_len, 1), complex)
    scratch = np.zeros(max(int(max((int(o["c_off"]) + int(o["ldc"]) * int(o["n"]) for o in o2), default=1)), 1),
                       complex)
    bufs = {0: ops, 1: vin, 2: mid, 3: scratch}
    mid_start = o1["c_off"] if len(o1) else np.zeros(0, np.int64)
    done = set()
    flops = elapsed = 0.0
    nwin = 0
    for w in order:
        ob = o2[w]
        t0 = time.perf_counter()
        for c in c2[int(ob["cbeg"]):int(ob["cend"])]:
            for buf, off in ((c["a_buf"], c["a_off"]), (c["b_buf"], c["b_off"])):
                if buf == 2 and int(off) not in done:
                    j = int(np.searchsorted(mid_start, off))
                    acc, f = _window(bufs, c1, o1[j])
                    _store(mid, o1[j], acc)
                    flops += f
                    done.add(int(off))
        acc, f = _window(bufs, c2, ob)
        _store(scratch, ob, acc)
        flops += f
        elapsed += time.perf_counter() - t0
        nwin += 1
        if elapsed >= seconds:
            break
    return flops, elapsed, nwin


def execute_windows(ops, c1, o1, c2, o2, mid_len, vin, which):
    """Stage-2 output windows o2[which] only, with the stage-1 intermediates
    they read; returns {window index: m x n array}."""
    mid = np.zeros(max(mid_len, 1), complex)
    bufs = {0: ops, 1: vin, 2: mid, 3: None}
    mid_start = o1["c_off"] if len(o1) else np.zeros(0, np.int64)
    done = set()
    out = {}
    for w in which:
        ob = o2[int(w)]
        for c in c2[int(ob["cbeg"]):int(ob["cend"])]:
            for buf, off in ((c["a_buf"], c["a_off"]), (c["b_buf"], c["b_off"])):
                if buf == 2 and int(off) not in done:
                    j = int(np.searchsorted(mid_start, off))
                    acc, _ = _window(bufs, c1, o1[j])
                    _store(mid, o1[j], acc)
                    done.add(int(off))
        out[int(w)], _ = _window(bufs, c2, ob)
    return out
