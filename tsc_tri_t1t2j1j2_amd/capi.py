"""ctypes binding of libtsc_hotpath.so (the C-ABI in include/tsc_hotpath.h).

The library is the product: every call here goes to the HIP implementation.
Loading fails loudly when the shared library has not been built; there is
no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsc_hotpath.so")

PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
PREC_NAMES = {"f64": PREC_F64, "f32": PREC_F32, "bf16x3": PREC_BF16X3}
# size of the static device buffer of tsc_hpsi_noalloc (TSC_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20
OP_N, OP_T, OP_C, OP_CONJ, OP_I = 0, 1, 2, 3, 4
BUF_OPS, BUF_IN, BUF_MID, BUF_OUT = 0, 1, 2, 3

# numpy mirrors of the C structs (layout checked by tests against the header)
CONTRIB_DTYPE = np.dtype({
    "names": ["a_off", "b_off", "lda", "ldb", "k", "a_buf", "b_buf", "a_op", "b_op", "coef_re", "coef_im"],
    "formats": [np.int64, np.int64, np.int32, np.int32, np.int32, np.int8, np.int8, np.int8, np.int8,
                np.float64, np.float64],
    "offsets": [0, 8, 16, 20, 24, 28, 29, 30, 31, 32, 40],
    "itemsize": 48,
})
OUTBLOCK_DTYPE = np.dtype({
    "names": ["c_off", "ldc", "m", "n", "c_buf", "cbeg", "cend", "mode"],
    "formats": [np.int64, np.int32, np.int32, np.int32, np.int32, np.int32, np.int32, np.int32],
    "offsets": [0, 8, 12, 16, 20, 24, 28, 32],
    "itemsize": 40,
})

# every symbol declared in include/tsc_hotpath.h
EXPORTS = [
    "tsc_last_error", "tsc_version", "tsc_ctx_create", "tsc_ctx_destroy", "tsc_ctx_sync", "tsc_ctx_mfma_layout",
    "tsc_ctx_mem_info",
    "tsc_vec_create", "tsc_vec_destroy", "tsc_vec_upload", "tsc_vec_download", "tsc_vec_dot_re",
    "tsc_vec_norm", "tsc_vec_dptr", "tsc_vec_fill_gaussian", "tsc_vec_axpby", "tsc_vec_copy",
    "tsc_vec_copy_range", "tsc_vec_upload_range", "tsc_vec_download_range",
    "tsc_plan_create", "tsc_plan_create_dev", "tsc_plan_destroy", "tsc_plan_update_ops", "tsc_plan_fill_ops_gaussian",
    "tsc_plan_flops", "tsc_hpsi", "tsc_hpsi_host", "tsc_hpsi_loop", "tsc_hpsi_noalloc", "tsc_hpsi_graph",
    "tsc_plan_last_times",
    "tsc_plan_autotune",
    "tsc_lanczos", "tsc_lanczos_replays", "tsc_lanczos_refinement", "tsc_density_matrix", "tsc_truncation_operator", "tsc_rho_truncate", "tsc_rho_last_times", "tsc_select_states",
    "tsc_update_trun", "tsc_w6j", "tsc_superblock_tasks", "tsc_tasks_info", "tsc_tasks_flops", "tsc_tasks_copy",
    "tsc_model_tasks", "tsc_renorm_tasks", "tsc_tasks_layout", "tsc_tasks_lengths", "tsc_tasks_copy_arena",
    "tsc_tasks_plan", "tsc_tasks_plan_dev", "tsc_tasks_destroy",
]


class TscError(RuntimeError):
    """Nonzero status from the C-ABI."""


_lib = None

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_uint64 = ctypes.c_uint64
c_float = ctypes.c_float
P = ctypes.POINTER


def lib():
    """Load the shared library once (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TscError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "tsc_last_error": (ctypes.c_char_p, []),
        "tsc_version": (c_int, []),
        "tsc_ctx_create": (c_int, [c_int, P(c_void_p)]),
        "tsc_ctx_destroy": (c_int, [c_void_p]),
        "tsc_ctx_sync": (c_int, [c_void_p]),
        "tsc_ctx_mfma_layout": (c_int, [c_void_p, P(c_int32)]),
        "tsc_ctx_mem_info": (c_int, [c_void_p, P(ctypes.c_uint64), P(ctypes.c_uint64)]),
        "tsc_vec_create": (c_int, [c_void_p, c_int64, P(c_void_p)]),
        "tsc_vec_destroy": (c_int, [c_void_p]),
        "tsc_vec_upload": (c_int, [c_void_p, c_void_p]),
        "tsc_vec_download": (c_int, [c_void_p, c_void_p]),
        "tsc_vec_dot_re": (c_int, [c_void_p, c_void_p, P(c_double)]),
        "tsc_vec_norm": (c_int, [c_void_p, P(c_double)]),
        "tsc_vec_dptr": (c_void_p, [c_void_p]),
        "tsc_vec_fill_gaussian": (c_int, [c_void_p, c_uint64, c_uint64, c_double]),
        "tsc_vec_axpby": (c_int, [c_void_p, c_double, c_void_p, c_double]),
        "tsc_vec_copy": (c_int, [c_void_p, c_void_p]),
        "tsc_vec_copy_range": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_int64]),
        "tsc_vec_upload_range": (c_int, [c_void_p, c_int64, c_void_p, c_int64]),
        "tsc_vec_download_range": (c_int, [c_void_p, c_int64, c_void_p, c_int64]),
        "tsc_plan_create_dev": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_int32,
                                        c_void_p, c_int32, c_void_p, c_int32, c_int64, c_int64, P(c_void_p)]),
        "tsc_plan_create": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_int32, c_void_p, c_int32,
                                    c_void_p, c_int32, c_void_p, c_int32, c_int64, c_int64, P(c_void_p)]),
        "tsc_plan_destroy": (c_int, [c_void_p]),
        "tsc_plan_update_ops": (c_int, [c_void_p, c_void_p]),
        "tsc_plan_fill_ops_gaussian": (c_int, [c_void_p, c_int64, c_int64, c_uint64, c_uint64, c_double]),
        "tsc_plan_flops": (c_int, [c_void_p, P(c_double)]),
        "tsc_hpsi": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
        "tsc_hpsi_host": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
        "tsc_hpsi_graph": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
        "tsc_hpsi_loop": (c_int, [c_void_p, c_void_p, c_void_p]),
        "tsc_hpsi_noalloc": (c_int, [c_void_p, c_void_p, c_void_p]),
        "tsc_plan_last_times": (c_int, [c_void_p, P(c_float), P(c_float)]),
        "tsc_plan_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, P(c_double)]),
        "tsc_lanczos": (c_int, [c_void_p, c_void_p, c_double, c_int, c_int, c_int, c_int, P(c_double), P(c_int32)]),
        "tsc_lanczos_replays": (c_int, [c_void_p, P(c_int32)]),
        "tsc_lanczos_refinement": (c_int, [c_void_p, P(c_double)]),
        "tsc_density_matrix": (c_int, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int, c_void_p]),
        "tsc_truncation_operator": (c_int, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_int32,
                                            c_double, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
        "tsc_rho_last_times": (c_int, [c_void_p, P(c_float)]),
        "tsc_rho_truncate": (c_int, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int, c_int32,
                                     c_int32, c_double, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
        "tsc_select_states": (c_int, [c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_double, c_void_p,
                                      c_void_p]),
        "tsc_update_trun": (c_int, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64,
                                    c_int]),
        "tsc_w6j": (c_double, [c_int32] * 6),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().tsc_last_error().decode(errors="replace")
        raise TscError(f"tsc status {rc}: {msg}")


def _ptr(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


class Context:
    """A device context (one HIP stream on one GPU)."""

    def __init__(self, device=0):
        h = c_void_p()
        check(lib().tsc_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        """Destroy the context. Vectors and plans made from it must not be
        used afterwards; their own destructors then skip the device free."""
        if self.h:
            lib().tsc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mfma_layout(self):
        m = (c_int32 * 3)()
        check(lib().tsc_ctx_mfma_layout(self.h, m))
        return list(m)

    def sync(self):
        check(lib().tsc_ctx_sync(self.h))

    def mem_info(self):
        """(free, total) device memory in bytes."""
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().tsc_ctx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def vec(self, n):
        return Vec(self, n)

    def vec_from(self, x):
        x = _c128(x).ravel(order="F")
        v = Vec(self, x.size)
        v.upload(x)
        return v


class Vec:
    """A resident complex128 vector on the device."""

    def __init__(self, ctx, n):
        self.ctx = ctx
        self.n = int(n)
        h = c_void_p()
        check(lib().tsc_vec_create(ctx.h, self.n, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h and self.ctx.h:
            lib().tsc_vec_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, x):
        x = _c128(x).ravel(order="K")
        assert x.size == self.n, (x.size, self.n)
        check(lib().tsc_vec_upload(self.h, _ptr(x)))

    def download(self):
        out = np.empty(self.n, dtype=np.complex128)
        check(lib().tsc_vec_download(self.h, _ptr(out)))
        return out

    def dptr(self):
        """Device address of the vector's data (complex128)."""
        return lib().tsc_vec_dptr(self.h)

    def dot_re(self, other):
        r = c_double()
        check(lib().tsc_vec_dot_re(self.h, other.h, ctypes.byref(r)))
        return r.value

    def norm(self):
        r = c_double()
        check(lib().tsc_vec_norm(self.h, ctypes.byref(r)))
        return r.value

    def fill_gaussian(self, seed, key, sigma=1.0):
        check(lib().tsc_vec_fill_gaussian(self.h, seed, key, sigma))

    def axpby(self, a, x, b):
        check(lib().tsc_vec_axpby(self.h, a, x.h, b))

    def copy_from(self, x):
        check(lib().tsc_vec_copy(self.h, x.h))

    def copy_range(self, off, src, src_off, n):
        """self[off:off+n] = src[src_off:src_off+n] on the device (asynchronous)."""
        check(lib().tsc_vec_copy_range(self.h, int(off), src.h, int(src_off), int(n)))

    def upload_range(self, off, x):
        """self[off:off+len(x)] = x; synchronous with respect to x."""
        x = _c128(x).ravel(order="K")
        check(lib().tsc_vec_upload_range(self.h, int(off), _ptr(x), x.size))
        self.ctx.sync()

    def download_range(self, off, n):
        out = np.empty(int(n), dtype=np.complex128)
        check(lib().tsc_vec_download_range(self.h, int(off), _ptr(out), int(n)))
        self.ctx.sync()
        return out


class Plan:
    """A grouped two-stage contraction plan (H.psi, U^H O U, rho)."""

    def __init__(self, ctx, ops, ops_elems, mid_elems, c1, o1, c2, o2, in_elems, out_elems):
        self.ctx = ctx
        c1 = np.ascontiguousarray(c1, dtype=CONTRIB_DTYPE)
        o1 = np.ascontiguousarray(o1, dtype=OUTBLOCK_DTYPE)
        c2 = np.ascontiguousarray(c2, dtype=CONTRIB_DTYPE)
        o2 = np.ascontiguousarray(o2, dtype=OUTBLOCK_DTYPE)
        ops_arr = None if ops is None else _c128(ops)
        h = c_void_p()
        check(lib().tsc_plan_create(ctx.h, _ptr(ops_arr), int(ops_elems), int(mid_elems), _ptr(c1), len(c1),
                                    _ptr(o1), len(o1), _ptr(c2), len(c2), _ptr(o2), len(o2), int(in_elems),
                                    int(out_elems), ctypes.byref(h)))
        self.h = h
        self.in_elems = int(in_elems)
        self.out_elems = int(out_elems)
        self.ops_elems = int(ops_elems)
        self.ops_vec = None

    @classmethod
    def create_dev(cls, ctx, ops_vec, mid_elems, c1, o1, c2, o2, in_elems, out_elems):
        """Plan over an operator arena already on the device (borrowed: the
        plan keeps a reference to ops_vec)."""
        c1 = np.ascontiguousarray(c1, dtype=CONTRIB_DTYPE)
        o1 = np.ascontiguousarray(o1, dtype=OUTBLOCK_DTYPE)
        c2 = np.ascontiguousarray(c2, dtype=CONTRIB_DTYPE)
        o2 = np.ascontiguousarray(o2, dtype=OUTBLOCK_DTYPE)
        h = c_void_p()
        check(lib().tsc_plan_create_dev(ctx.h, ops_vec.h, int(mid_elems), _ptr(c1), len(c1), _ptr(o1), len(o1),
                                        _ptr(c2), len(c2), _ptr(o2), len(o2), int(in_elems), int(out_elems),
                                        ctypes.byref(h)))
        pl = cls.__new__(cls)
        pl.ctx, pl.h = ctx, h
        pl.in_elems, pl.out_elems, pl.ops_elems = int(in_elems), int(out_elems), ops_vec.n
        pl.ops_vec = ops_vec
        return pl

    def close(self):
        if self.h and self.ctx.h:
            lib().tsc_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def flops(self):
        r = c_double()
        check(lib().tsc_plan_flops(self.h, ctypes.byref(r)))
        return r.value

    def update_ops(self, ops):
        ops = _c128(ops)
        assert ops.size == self.ops_elems
        check(lib().tsc_plan_update_ops(self.h, _ptr(ops)))

    def fill_ops_gaussian(self, off, count, seed, key, sigma):
        check(lib().tsc_plan_fill_ops_gaussian(self.h, int(off), int(count), seed, key, float(sigma)))

    def apply(self, vin, vout, prec=PREC_F64):
        check(lib().tsc_hpsi(self.h, vin.h, vout.h, int(prec)))

    def apply_loop(self, vin, vout):
        """FP64 H.psi as one ZGEMM call per contribution (the per-product
        baseline; plans of GEMM contributions only)."""
        check(lib().tsc_hpsi_loop(self.h, vin.h, vout.h))

    def apply_noalloc(self, vin, vout):
        """FP64 H.psi that allocates nothing: intermediates and partial tiles
        in the library's static scratch buffer (NOALLOC_SCRATCH_BYTES per
        device); enqueued on the context stream."""
        check(lib().tsc_hpsi_noalloc(self.h, vin.dptr(), vout.dptr()))

    def apply_graph(self, vin, vout, prec=PREC_F64):
        check(lib().tsc_hpsi_graph(self.h, vin.h, vout.h, int(prec)))

    def apply_host(self, x, prec=PREC_F64):
        x = _c128(x)
        assert x.size == self.in_elems
        out = np.empty(self.out_elems, dtype=np.complex128)
        check(lib().tsc_hpsi_host(self.h, _ptr(x), _ptr(out), int(prec)))
        return out

    def last_times(self):
        a, b = c_float(), c_float()
        check(lib().tsc_plan_last_times(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, vin, vout, prec, force=False, target=None):
        """Choose the work-unit target of this plan's `prec` path by timing
        H.psi (vin -> vout, vout overwritten); returns the target kept.
        target: cut the units for this target instead, without timing."""
        r = c_double(0.0 if target is None else float(target))
        flags = (1 if force else 0) | (2 if target is not None else 0)
        check(lib().tsc_plan_autotune(self.h, vin.h, vout.h, int(prec), flags, ctypes.byref(r)))
        return r.value

    def lanczos_replays(self):
        """H.psi products of the last lanczos() that ran as graph replays."""
        r = c_int32()
        check(lib().tsc_lanczos_replays(self.h, ctypes.byref(r)))
        return r.value

    def lanczos_refinement(self):
        """Precision record of the last lanczos(): dict with e_mixed (the
        working-precision Ritz value), e_rayleigh_f64 and resid_f64 (FP64
        <psi|H psi> and |H psi - <psi|H psi> psi| of its normalised Ritz
        vector, from the refinement pass's first product), e_refined (the
        returned energy) and w_rest (the weight of the refined vector
        outside the mixed one); all but e_mixed are NaN without refinement."""
        r = (c_double * 5)()
        check(lib().tsc_lanczos_refinement(self.h, r))
        return dict(e_mixed=r[0], e_rayleigh_f64=r[1], resid_f64=r[2], e_refined=r[3], w_rest=r[4])

    def lanczos(self, psi, tol=1e-10, max_inner=100, max_outer=4, prec=PREC_F64, refine_iters=0):
        e = c_double()
        info = (c_int32 * 2)()
        check(lib().tsc_lanczos(self.h, psi.h, float(tol), int(max_inner), int(max_outer), int(prec),
                                int(refine_iters), ctypes.byref(e), info))
        return e.value, (info[0], info[1])


def density_matrix(ctx, psi_vec, sys_dim, env_dim, side):
    sys_dim = np.ascontiguousarray(sys_dim, dtype=np.int32)
    env_dim = np.ascontiguousarray(env_dim, dtype=np.int32)
    d = sys_dim if side == 0 else env_dim
    out = np.empty(int(np.sum(d.astype(np.int64) ** 2)), dtype=np.complex128)
    check(lib().tsc_density_matrix(ctx.h, psi_vec.h, len(sys_dim), _ptr(sys_dim), _ptr(env_dim), int(side),
                                   _ptr(out)))
    return out


def _trunc_outputs(dim):
    nsec = len(dim)
    eig = np.empty(int(np.sum(dim)), dtype=np.float64)
    kept = np.empty(nsec, dtype=np.int32)
    # a sector keeps at most all of its states (the kept rule can pass
    # kept_max by one and keeps everything when n <= kept_min)
    ucap = int(np.sum(dim.astype(np.int64) ** 2))
    u = np.empty(max(ucap, 1), dtype=np.complex128)
    scal = np.empty(4, dtype=np.float64)
    return eig, kept, u, scal


def _split_u(dim, kept, u):
    out, off = [], 0
    for d, k in zip(dim, kept):
        n = int(d) * int(k)
        out.append(u[off:off + n].reshape((int(d), int(k)), order="F") if k > 0 else None)
        off += n
    return out


def truncation_operator(ctx, dim, twos, rho, kept_min, kept_max, trun_err):
    """Get_truncation_operator on host rho blocks laid end to end."""
    dim = np.ascontiguousarray(dim, dtype=np.int32)
    twos = np.ascontiguousarray(twos, dtype=np.int32)
    rho = _c128(rho)
    eig, kept, u, scal = _trunc_outputs(dim)
    check(lib().tsc_truncation_operator(ctx.h, len(dim), _ptr(dim), _ptr(twos), _ptr(rho), int(kept_min),
                                        int(kept_max), float(trun_err), _ptr(eig), _ptr(kept), _ptr(u), u.size,
                                        _ptr(scal)))
    return dict(eigval=eig, kept=kept, U=_split_u(dim, kept, u), num_state=int(scal[0]), trun_err=scal[1],
                entropy=scal[2], entropy1=scal[3])


def rho_truncate(ctx, psi_vec, sys_dim, env_dim, twos, side, kept_min, kept_max, trun_err):
    sys_dim = np.ascontiguousarray(sys_dim, dtype=np.int32)
    env_dim = np.ascontiguousarray(env_dim, dtype=np.int32)
    twos = np.ascontiguousarray(twos, dtype=np.int32)
    dim = sys_dim if side == 0 else env_dim
    eig, kept, u, scal = _trunc_outputs(dim)
    check(lib().tsc_rho_truncate(ctx.h, psi_vec.h, len(dim), _ptr(sys_dim), _ptr(env_dim), _ptr(twos), int(side),
                                 int(kept_min), int(kept_max), float(trun_err), _ptr(eig), _ptr(kept), _ptr(u),
                                 u.size, _ptr(scal)))
    return dict(eigval=eig, kept=kept, U=_split_u(dim, kept, u), num_state=int(scal[0]), trun_err=scal[1],
                entropy=scal[2], entropy1=scal[3])


def rho_last_times(ctx):
    """Wall ms of the last rho / eigensolve / truncation call on ctx: rho
    build, eigensolves, eigenvalues to the host + selection, U to the host."""
    ms = (c_float * 4)()
    check(lib().tsc_rho_last_times(ctx.h, ms))
    return [float(x) for x in ms]


def select_states(dim, twos, eigval, kept_min, kept_max, trun_err):
    """The kept-state rule alone (host code of the library, no device)."""
    dim = np.ascontiguousarray(dim, dtype=np.int32)
    twos = np.ascontiguousarray(twos, dtype=np.int32)
    eigval = np.ascontiguousarray(eigval, dtype=np.float64)
    assert eigval.size == int(np.sum(dim))
    kept = np.empty(len(dim), dtype=np.int32)
    scal = np.empty(4, dtype=np.float64)
    check(lib().tsc_select_states(len(dim), _ptr(dim), _ptr(twos), _ptr(eigval), int(kept_min), int(kept_max),
                                  float(trun_err), _ptr(kept), _ptr(scal)))
    return dict(kept=kept, num_state=int(scal[0]), trun_err=scal[1], entropy=scal[2], entropy1=scal[3])


def update_trun(ctx, blocks, prec=PREC_F64):
    """eff_i = U_l^H O_i U_r for a list of (O, U_l, U_r) complex matrices."""
    nb = len(blocks)
    row = np.array([b[0].shape[0] for b in blocks], dtype=np.int32)
    col = np.array([b[0].shape[1] for b in blocks], dtype=np.int32)
    kl = np.array([b[1].shape[1] for b in blocks], dtype=np.int32)
    kr = np.array([b[2].shape[1] for b in blocks], dtype=np.int32)
    ori_off = np.zeros(nb, np.int64)
    ul_off = np.zeros(nb, np.int64)
    ur_off = np.zeros(nb, np.int64)
    eff_off = np.zeros(nb, np.int64)
    ori_parts, u_parts = [], []
    oo = uo = eo = 0
    for i, (O, Ul, Ur) in enumerate(blocks):
        assert Ul.shape[0] == O.shape[0] and Ur.shape[0] == O.shape[1]
        ori_off[i] = oo
        ori_parts.append(np.asarray(O, np.complex128).ravel(order="F"))
        oo += O.size
        ul_off[i] = uo
        u_parts.append(np.asarray(Ul, np.complex128).ravel(order="F"))
        uo += Ul.size
        ur_off[i] = uo
        u_parts.append(np.asarray(Ur, np.complex128).ravel(order="F"))
        uo += Ur.size
        eff_off[i] = eo
        eo += Ul.shape[1] * Ur.shape[1]
    ori = np.concatenate(ori_parts) if ori_parts else np.zeros(1, np.complex128)
    u = np.concatenate(u_parts) if u_parts else np.zeros(1, np.complex128)
    eff = np.empty(max(eo, 1), dtype=np.complex128)
    check(lib().tsc_update_trun(ctx.h, nb, _ptr(row), _ptr(col), _ptr(kl), _ptr(kr), _ptr(ori_off), _ptr(ul_off),
                                _ptr(ur_off), _ptr(eff_off), _ptr(ori), int(oo), _ptr(u), int(uo), _ptr(eff),
                                int(eo), int(prec)))
    out = []
    for i in range(nb):
        n = int(kl[i]) * int(kr[i])
        out.append(eff[eff_off[i]:eff_off[i] + n].reshape((int(kl[i]), int(kr[i])), order="F"))
    return out


def w6j(j1, j2, j3, l1, l2, l3):
    """6j symbol with doubled integer arguments (host code in the library)."""
    return lib().tsc_w6j(int(j1), int(j2), int(j3), int(l1), int(l2), int(l3))
