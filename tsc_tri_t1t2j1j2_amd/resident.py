"""Block operators kept on the device between DMRG steps.

The reference keeps every block's operators as host matrices
(Total_Model%ham / %sub(:)%mat, hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/common.f90:84-174), merges partner operators on the host (hmerge,
get_sys.f90:7398-7626; block_add_block_two, block.f90:177-215) and builds
the enlarged, truncated operators on the host (Get_Hamiltonian_Sys/Env,
get_sys.f90:4-592; update_trun_dia/ndia, block.f90:3465-3683). Here:

* a block's operators live in one device vector (DevArena); a sub-block is a
  DevMat (arena, element offset, shape), column-major as in the reference;
* hmerge's sums are PendingMerge values, formed on the device by one plan of
  elementwise terms per step;
* each step's plans read one StepArena: the arenas of the blocks in use
  copied end to end, followed by the merged operators;
* renormalisation writes the new block's operators straight into a new
  DevArena (the plan's output vector).

The host keeps only quantum-number bookkeeping.
"""
import numpy as np

from . import capi
from .blocks import TotalBlock, TotalModel


class DevArena:
    """A device vector of complex128 values holding operator sub-blocks."""

    def __init__(self, vec):
        self.vec = vec
        self.n = vec.n

    def close(self):
        if self.vec is not None:
            self.vec.close()
        self.vec = None


class DevMat:
    """A column-major matrix at element offset `off` of a DevArena."""

    __slots__ = ("arena", "off", "shape")

    def __init__(self, arena, off, rows, cols):
        self.arena = arena
        self.off = int(off)
        self.shape = (int(rows), int(cols))

    @property
    def size(self):
        return self.shape[0] * self.shape[1]


class PendingMerge:
    """sum_j w_j M_j over matrices of one shape, formed on the device by
    StepArena.finalize; `dev` is the resulting DevMat."""

    __slots__ = ("parts", "shape", "dev")

    def __init__(self, shape):
        self.parts = []
        self.shape = tuple(int(v) for v in shape)
        self.dev = None

    @property
    def size(self):
        return self.shape[0] * self.shape[1]


def resident(m):
    return isinstance(m, (DevMat, PendingMerge))


def model_blocks(model):
    """Every TotalBlock of a model (Hamiltonian first)."""
    out = [model.ham]
    for i in sorted(model.ops):
        for name in sorted(model.ops[i]):
            out.append(model.ops[i][name])
    return out


def model_to_device(ctx, model):
    """The same block with all its matrices in one DevArena."""
    mats = []
    for tb in model_blocks(model):
        for key in sorted(tb.blocks):
            mats.append(np.asarray(tb.blocks[key], np.complex128))
    n = sum(m.size for m in mats)
    arena = DevArena(ctx.vec(max(n, 1)))
    host = np.zeros(max(n, 1), np.complex128)
    off = 0
    placed = {}
    for tb in model_blocks(model):
        for key in sorted(tb.blocks):
            m = np.asarray(tb.blocks[key], np.complex128)
            host[off:off + m.size] = m.ravel(order="F")
            placed[(id(tb), key)] = DevMat(arena, off, m.shape[0], m.shape[1])
            off += m.size
    arena.vec.upload(host)

    def conv(tb):
        return TotalBlock(tb.up_dif, tb.k2, {k: placed[(id(tb), k)] for k in tb.blocks})

    ops = {i: {name: conv(tb) for name, tb in d.items()} for i, d in model.ops.items()}
    out = TotalModel(model.sites, model.sectors, conv(model.ham), ops)
    out.arena = arena
    return out


def model_to_host(model):
    """A host copy (numpy blocks) of a resident block."""
    cache = {}

    def fetch(m):
        if not isinstance(m, DevMat):
            return np.asarray(m)
        a = id(m.arena)
        if a not in cache:
            cache[a] = m.arena.vec.download()
        return cache[a][m.off:m.off + m.size].reshape(m.shape, order="F").copy()

    def conv(tb):
        return TotalBlock(tb.up_dif, tb.k2, {k: fetch(v) for k, v in tb.blocks.items()})

    ops = {i: {name: conv(tb) for name, tb in d.items()} for i, d in model.ops.items()}
    return TotalModel(model.sites, model.sectors, conv(model.ham), ops)


def merged_resident(ops_by_site, weights, name):
    """hmerge's sum_i w_i O_i[name] as pending device merges."""
    acc = None
    for i, w in weights:
        o = ops_by_site[i][name]
        if acc is None:
            acc = TotalBlock(o.up_dif, o.k2)
        for key, m in o.blocks.items():
            pm = acc.blocks.get(key)
            if pm is None:
                pm = acc.blocks[key] = PendingMerge(m.shape)
            pm.parts.append((float(w), m))
    return acc


def pending_in(obj, out=None):
    """PendingMerge values reachable from a term list / TotalBlock / tuple."""
    out = [] if out is None else out
    if isinstance(obj, TotalBlock):
        for m in obj.blocks.values():
            if isinstance(m, PendingMerge):
                out.append(m)
    elif isinstance(obj, (list, tuple)):
        for x in obj:
            pending_in(x, out)
    return out


class StepArena:
    """Operator arena of one DMRG step on the device.

    Layout: the DevArenas of the given blocks end to end, then the merged
    operators. offset(m) gives (element offset, leading dimension) of any
    matrix the step's plans read."""

    def __init__(self, ctx, models):
        self.ctx = ctx
        self.base = {}
        self.arenas = []
        self.len = 0
        for mdl in models:
            for tb in model_blocks(mdl):
                for m in tb.blocks.values():
                    if isinstance(m, DevMat) and id(m.arena) not in self.base:
                        self.base[id(m.arena)] = self.len
                        self.arenas.append(m.arena)
                        self.len += m.arena.n
        self.merges = []
        self._merge_off = {}
        self.vec = None
        self.as_arena = None

    def add_merges(self, obj):
        for pm in pending_in(obj):
            if id(pm) in self._merge_off:
                continue
            for _, m in pm.parts:
                if not isinstance(m, DevMat) or id(m.arena) not in self.base:
                    raise ValueError("merge part outside the step's block arenas")
            self._merge_off[id(pm)] = self.len
            self.merges.append(pm)
            self.len += pm.size

    def finalize(self):
        """Allocate, copy the block arenas in and form the merged operators."""
        ctx = self.ctx
        self.vec = ctx.vec(max(self.len, 1))
        for a in self.arenas:
            self.vec.copy_range(self.base[id(a)], a.vec, 0, a.n)
        self.as_arena = DevArena(self.vec)
        if self.merges:
            # one output window per merged operator, one elementwise term per part
            ob_rows, c_rows = [], []
            for pm in self.merges:
                rows, cols = pm.shape
                cbeg = len(c_rows)
                for w, m in pm.parts:
                    c_rows.append((self.base[id(m.arena)] + m.off, max(rows, 1), w))
                ob_rows.append((self._merge_off[id(pm)], max(rows, 1), rows, cols, cbeg, len(c_rows)))
            c2 = np.zeros(len(c_rows), capi.CONTRIB_DTYPE)
            ca = np.array(c_rows, dtype=np.float64)
            c2["a_off"], c2["lda"], c2["coef_re"] = ca[:, 0].astype(np.int64), ca[:, 1].astype(np.int32), ca[:, 2]
            c2["a_buf"], c2["a_op"], c2["b_op"] = capi.BUF_OPS, capi.OP_N, capi.OP_I
            o2 = np.zeros(len(ob_rows), capi.OUTBLOCK_DTYPE)
            oa = np.array(ob_rows, dtype=np.int64)
            o2["c_off"], o2["ldc"], o2["m"], o2["n"], o2["cbeg"], o2["cend"] = oa.T
            o2["c_buf"] = capi.BUF_OUT
            plan = capi.Plan.create_dev(ctx, self.vec, 0, np.zeros(0, capi.CONTRIB_DTYPE),
                                        np.zeros(0, capi.OUTBLOCK_DTYPE), c2, o2, 0, self.vec.n)
            dummy = ctx.vec(1)
            plan.apply(dummy, self.vec, capi.PREC_F64)
            ctx.sync()
            plan.close()
            dummy.close()
        for pm in self.merges:
            pm.dev = DevMat(self.as_arena, self._merge_off[id(pm)], *pm.shape)
        return self

    def offset(self, m):
        """(element offset in this arena, leading dimension) of a matrix."""
        if isinstance(m, PendingMerge):
            if m.dev is None:
                raise ValueError("merged operator used before StepArena.finalize")
            m = m.dev
        if not isinstance(m, DevMat):
            raise TypeError("host matrix in a device-resident step")
        if m.arena is self.as_arena:
            off = m.off
        else:
            off = self.base[id(m.arena)] + m.off
        return off, max(m.shape[0], 1)

    def close(self):
        if self.vec is not None:
            self.vec.close()
        self.vec = None


def renormalize(ctx, step, rp, prec=capi.PREC_F64):
    """Run a renormalisation plan (hamiltonian.renormalize_plan built with
    step.offset) over the step arena; the new block's operators stay on the
    device in the plan's output vector."""
    from .hamiltonian import renormalize_result
    plan = rp.builder.build_dev(ctx, step.vec, rp.uin.size, rp.out_len)
    uin = ctx.vec_from(rp.uin)
    out = ctx.vec(rp.out_len)
    plan.apply(uin, out, prec)
    ctx.sync()
    # algorithmic flops and device time of the contraction (both stages)
    stats = (plan.flops(), sum(plan.last_times()))
    plan.close()
    uin.close()
    arena = DevArena(out)
    model = renormalize_result(rp, None, arena=arena)
    model.arena = arena
    model.renorm_stats = stats
    return model
