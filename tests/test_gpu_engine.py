"""Parity of the HIP grouped-contraction engine, Lanczos, density matrix /
truncation and U^H O U against the oracle (runs through the C-ABI)."""
import numpy as np
import pytest

from oracle import lanczos as olanczos
from oracle import synth as osynth
from oracle import truncation as otrunc
from oracle import uou as ouou
from tests.hermitian_fixture import HermitianSuperblock
from tsc_tri_t1t2j1j2_amd import capi
from tsc_tri_t1t2j1j2_amd.synthetic import SyntheticSuperblock

pytestmark = pytest.mark.gpu

# tolerances (relative Frobenius error of H psi against FP64 numpy)
TOL = {capi.PREC_F64: 1e-13, capi.PREC_F32: 1e-5, capi.PREC_BF16X3: 3e-5}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _op(M, op):
    return {capi.OP_N: M, capi.OP_T: M.T, capi.OP_C: M.conj().T, capi.OP_CONJ: M.conj()}[op]


def _random_case(rng, nblocks=10):
    """Random single-stage problem covering every op pair and identities."""
    ops, ooff = [], 0
    vin = []
    ioff = 0
    c2, o2 = [], []
    out_off = 0
    expected = []
    for b in range(nblocks):
        m, n = (int(x) for x in rng.integers(1, 140, 2))
        ob = np.zeros((), capi.OUTBLOCK_DTYPE)
        ob["c_off"], ob["ldc"], ob["m"], ob["n"], ob["c_buf"] = out_off, m + 3, m, n, capi.BUF_OUT
        ob["cbeg"] = len(c2)
        acc = np.zeros((m, n), np.complex128)
        for _ in range(int(rng.integers(1, 5))):
            kind = rng.integers(0, 3)
            coef = complex(rng.normal(), rng.normal())
            e = np.zeros((), capi.CONTRIB_DTYPE)
            e["coef_re"], e["coef_im"] = coef.real, coef.imag
            k = int(rng.integers(0, 100))
            aop, bop = int(rng.integers(0, 4)), int(rng.integers(0, 4))
            if kind == 1:
                bop = capi.OP_I
            elif kind == 2:
                aop = capi.OP_I
            # A from the operator arena, B from the input vector
            if aop != capi.OP_I:
                rows, cols = (m, k) if bop != capi.OP_I else (m, n)
                st = (rows, cols) if aop in (capi.OP_N, capi.OP_CONJ) else (cols, rows)
                lda = st[0] + int(rng.integers(0, 3))
                A = rng.normal(size=(lda, st[1])) + 1j * rng.normal(size=(lda, st[1]))
                ops.append(A.ravel(order="F"))
                e["a_off"], e["lda"], e["a_buf"], e["a_op"] = ooff, max(lda, 1), capi.BUF_OPS, aop
                ooff += A.size
                Aop = _op(A[:st[0], :], aop)
            else:
                e["a_op"] = capi.OP_I
            if bop != capi.OP_I:
                rows, cols = (k, n) if aop != capi.OP_I else (m, n)
                st = (rows, cols) if bop in (capi.OP_N, capi.OP_CONJ) else (cols, rows)
                ldb = st[0] + int(rng.integers(0, 3))
                B = rng.normal(size=(ldb, st[1])) + 1j * rng.normal(size=(ldb, st[1]))
                vin.append(B.ravel(order="F"))
                e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = ioff, max(ldb, 1), capi.BUF_IN, bop
                ioff += B.size
                Bop = _op(B[:st[0], :], bop)
            else:
                e["b_op"] = capi.OP_I
            e["k"] = k
            if aop == capi.OP_I:
                acc += coef * Bop
            elif bop == capi.OP_I:
                acc += coef * Aop
            else:
                acc += coef * (Aop @ Bop)
            c2.append(e)
        ob["cend"] = len(c2)
        o2.append(ob)
        expected.append((out_off, m + 3, acc))
        out_off += (m + 3) * n
    return (np.concatenate(ops) if ops else np.zeros(1, np.complex128), ooff,
            np.concatenate(vin) if vin else np.zeros(1, np.complex128), ioff, c2, o2, out_off, expected)


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_F32, capi.PREC_BF16X3])
def test_grouped_contraction_all_op_modes(ctx, prec):
    rng = np.random.default_rng(100 + prec)
    ops, nops, vin, nin, c2, o2, nout, expected = _random_case(rng)
    plan = capi.Plan(ctx, ops, max(nops, 1), 0, np.zeros(0, capi.CONTRIB_DTYPE), np.zeros(0, capi.OUTBLOCK_DTYPE),
                     np.array(c2, capi.CONTRIB_DTYPE), np.array(o2, capi.OUTBLOCK_DTYPE), max(nin, 1), nout)
    x = np.zeros(max(nin, 1), np.complex128)
    x[:nin] = vin[:nin]
    out = plan.apply_host(x, prec)
    for off, ldc, acc in expected:
        m, n = acc.shape
        got = out[off:off + ldc * n].reshape((ldc, n), order="F")[:m, :]
        assert _rel(got, acc) < TOL[prec] * 10, (prec, m, n)


def _two_stage_case(rng, nout=8):
    """Random two-stage problem: stage 1 builds intermediates as sums of
    arena x psi products (long sums, so work units are split), stage 2 reads
    each intermediate as a row operand or as a transposed operand (some in
    both roles) with random ops and complex coefficients."""
    ops, vin = [], []
    nops = nin = 0
    c1, o1, c2, o2 = [], [], [], []
    mids = []  # (offset, ld, M)
    mid_off = 0

    def rnd(r, c):
        return rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c))

    def arena(M):
        nonlocal nops
        ops.append(M.ravel(order="F"))
        nops += M.size
        return nops - M.size

    def psi(M):
        nonlocal nin
        vin.append(M.ravel(order="F"))
        nin += M.size
        return nin - M.size

    def new_mid(mb, nb):
        nonlocal mid_off
        ld = mb + int(rng.integers(0, 2))
        ob = np.zeros((), capi.OUTBLOCK_DTYPE)
        ob["c_off"], ob["ldc"], ob["m"], ob["n"], ob["c_buf"] = mid_off, ld, mb, nb, capi.BUF_MID
        ob["cbeg"] = len(c1)
        acc = np.zeros((mb, nb), np.complex128)
        for t in range(int(rng.integers(1, 7))):
            e = np.zeros((), capi.CONTRIB_DTYPE)
            coef = complex(rng.normal(), rng.normal()) if t % 3 else 1.0
            e["coef_re"], e["coef_im"] = coef.real, coef.imag
            if t == 5:  # elementwise psi term
                X = rnd(mb, nb)
                e["a_off"], e["lda"], e["a_buf"], e["a_op"], e["b_op"] = psi(X), mb, capi.BUF_IN, capi.OP_N, capi.OP_I
                acc += coef * X
            else:
                k = int(rng.integers(1, 90))
                aop, bop = int(rng.integers(0, 4)), int(rng.integers(0, 4))
                A = rnd(*((mb, k) if aop in (capi.OP_N, capi.OP_CONJ) else (k, mb)))
                B = rnd(*((k, nb) if bop in (capi.OP_N, capi.OP_CONJ) else (nb, k)))
                e["a_off"], e["lda"], e["a_buf"], e["a_op"] = arena(A), A.shape[0], capi.BUF_OPS, aop
                e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = psi(B), B.shape[0], capi.BUF_IN, bop
                e["k"] = k
                acc += coef * (_op(A, aop) @ _op(B, bop))
            c1.append(e)
        ob["cend"] = len(c1)
        o1.append(ob)
        mids.append((mid_off, ld, acc))
        mid_off += ld * nb
        return len(mids) - 1

    expected, out_off = [], 0
    for b in range(nout):
        m, n = (int(x) for x in rng.integers(1, 150, 2))
        if b > 0 and rng.random() < 0.5:
            # reuse an intermediate in the other role: its dims fix (m or n)
            idx = int(rng.integers(0, len(mids)))
            mb, nb = mids[idx][2].shape
            m = nb if rng.random() < 0.5 else mb
        ob = np.zeros((), capi.OUTBLOCK_DTYPE)
        ob["c_off"], ob["ldc"], ob["m"], ob["n"], ob["c_buf"] = out_off, m, m, n, capi.BUF_OUT
        ob["cbeg"] = len(c2)
        acc = np.zeros((m, n), np.complex128)
        for t in range(int(rng.integers(1, 6))):
            e = np.zeros((), capi.CONTRIB_DTYPE)
            coef = complex(rng.choice([0.5, -1.25, 2.0]), 0.0) if t % 2 else complex(rng.normal(), rng.normal())
            e["coef_re"], e["coef_im"] = coef.real, coef.imag
            op = int(rng.integers(0, 4))
            # find or make an intermediate usable as A (rows m) or as B (columns n)
            cand = [(i, r) for i, (_, _, M) in enumerate(mids) for r in ("A", "B")
                    if (r == "A" and (M.shape[0] if op in (capi.OP_N, capi.OP_CONJ) else M.shape[1]) == m)
                    or (r == "B" and (M.shape[1] if op in (capi.OP_N, capi.OP_CONJ) else M.shape[0]) == n)]
            if cand and rng.random() < 0.6:
                idx, role = cand[int(rng.integers(0, len(cand)))]
            else:
                role = "A" if rng.random() < 0.5 else "B"
                k = int(rng.integers(1, 120))
                if role == "A":
                    dims = (m, k) if op in (capi.OP_N, capi.OP_CONJ) else (k, m)
                else:
                    dims = (k, n) if op in (capi.OP_N, capi.OP_CONJ) else (n, k)
                idx = new_mid(*dims)
            moff, mld, M = mids[idx]
            Mop = _op(M, op)
            if role == "A":
                k = Mop.shape[1]
                oop = int(rng.integers(0, 4))
                X = rnd(*((k, n) if oop in (capi.OP_N, capi.OP_CONJ) else (n, k)))
                e["a_off"], e["lda"], e["a_buf"], e["a_op"] = moff, mld, capi.BUF_MID, op
                e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = arena(X), X.shape[0], capi.BUF_OPS, oop
                acc += coef * (Mop @ _op(X, oop))
            else:
                k = Mop.shape[0]
                oop = int(rng.integers(0, 4))
                X = rnd(*((m, k) if oop in (capi.OP_N, capi.OP_CONJ) else (k, m)))
                e["a_off"], e["lda"], e["a_buf"], e["a_op"] = arena(X), X.shape[0], capi.BUF_OPS, oop
                e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = moff, mld, capi.BUF_MID, op
                acc += coef * (_op(X, oop) @ Mop)
            e["k"] = k
            c2.append(e)
        ob["cend"] = len(c2)
        o2.append(ob)
        expected.append((out_off, m, acc))
        out_off += m * n
    return (np.concatenate(ops), nops, np.concatenate(vin), nin, mid_off, c1, o1, c2, o2, out_off, expected)


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_F32, capi.PREC_BF16X3])
def test_two_stage_intermediates_in_both_roles(ctx, prec):
    rng = np.random.default_rng(500 + prec)
    ops, nops, vin, nin, nmid, c1, o1, c2, o2, nout, expected = _two_stage_case(rng)
    plan = capi.Plan(ctx, ops, nops, nmid, np.array(c1, capi.CONTRIB_DTYPE), np.array(o1, capi.OUTBLOCK_DTYPE),
                     np.array(c2, capi.CONTRIB_DTYPE), np.array(o2, capi.OUTBLOCK_DTYPE), nin, nout)
    out = plan.apply_host(vin, prec)
    for off, ldc, acc in expected:
        m, n = acc.shape
        got = out[off:off + ldc * n].reshape((ldc, n), order="F")[:m, :]
        assert _rel(got, acc) < TOL[prec] * 10, (prec, m, n)
    # a second call on the same plan (psi panels repacked, intermediates rewritten)
    out2 = plan.apply_host(2.0 * vin, prec)
    assert _rel(out2, 2.0 * out) < TOL[prec] * 10


@pytest.mark.parametrize("case", ["op_modes", "two_stage"])
def test_fp64_tile_shapes_agree(ctx, case):
    """The FP64 path's two tile shapes the autotuner chooses between (128x64
    LDS units, 64x64 grouped tiles, selected with target 1) give the same
    products for every operand mode and for intermediates in both roles."""
    if case == "op_modes":
        rng = np.random.default_rng(100)
        ops, nops, vin, nin, c2, o2, nout, expected = _random_case(rng)
        nmid, c1, o1 = 0, [], []
        nops, nin = max(nops, 1), max(nin, 1)
        x = np.zeros(nin, np.complex128)
        x[:len(vin)] = vin[:nin]
    else:
        rng = np.random.default_rng(500)
        ops, nops, vin, nin, nmid, c1, o1, c2, o2, nout, expected = _two_stage_case(rng)
        x = vin
    plan = capi.Plan(ctx, ops, nops, nmid, np.array(c1, capi.CONTRIB_DTYPE), np.array(o1, capi.OUTBLOCK_DTYPE),
                     np.array(c2, capi.CONTRIB_DTYPE), np.array(o2, capi.OUTBLOCK_DTYPE), nin, nout)
    ref = plan.apply_host(x, capi.PREC_F64)
    vi, vo = ctx.vec_from(x), ctx.vec(nout)
    assert plan.autotune(vi, vo, capi.PREC_F64, target=1.0) == 1.0
    got = plan.apply_host(x, capi.PREC_F64)
    assert _rel(got, ref) < 1e-13
    for off, ldc, acc in expected:
        m, n = acc.shape
        g = got[off:off + ldc * n].reshape((ldc, n), order="F")[:m, :]
        assert _rel(g, acc) < TOL[capi.PREC_F64] * 10
    # back to the 128-row units
    plan.autotune(vi, vo, capi.PREC_F64, target=2048.0)
    assert _rel(plan.apply_host(x, capi.PREC_F64), ref) < 1e-13
    vi.close()
    vo.close()
    plan.close()


def test_plan_rejects_out_of_bounds_operand(ctx):
    e = np.zeros(1, capi.CONTRIB_DTYPE)
    e["a_off"], e["lda"], e["a_buf"], e["a_op"] = 0, 4, capi.BUF_OPS, capi.OP_N
    e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = 0, 4, capi.BUF_IN, capi.OP_N
    e["k"] = 4
    o = np.zeros(1, capi.OUTBLOCK_DTYPE)
    o["ldc"], o["m"], o["n"], o["c_buf"], o["cend"] = 4, 4, 4, capi.BUF_OUT, 1
    with pytest.raises(capi.TscError, match="outside"):
        capi.Plan(ctx, np.zeros(15, np.complex128), 15, 0, np.zeros(0, capi.CONTRIB_DTYPE),
                  np.zeros(0, capi.OUTBLOCK_DTYPE), e, o, 16, 16)


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_F32, capi.PREC_BF16X3])
def test_synthetic_hpsi_matches_oracle(ctx, prec):
    sb = SyntheticSuperblock(nsec=12, nterms=6, dmin=8, dmax=150, seed=77)
    psi = sb.psi_host()
    ref = osynth.hpsi(sb, psi)
    plan = sb.make_plan(ctx, sb.ops_host())
    got = plan.apply_host(psi, prec)
    assert _rel(got, ref) < TOL[prec]
    assert abs(plan.flops() - 0) > 0


def test_device_filled_operators_match_host_generator(ctx):
    sb = SyntheticSuperblock(nsec=8, nterms=4, dmin=8, dmax=100, seed=5)
    psi = sb.psi_host()
    p_host = sb.make_plan(ctx, sb.ops_host())
    p_dev = sb.make_plan(ctx, None)
    assert _rel(p_dev.apply_host(psi), p_host.apply_host(psi)) < 1e-12


def test_resident_vectors_and_graph_replay(ctx):
    sb = SyntheticSuperblock(nsec=10, nterms=4, dmin=8, dmax=120, seed=9)
    psi = sb.psi_host()
    plan = sb.make_plan(ctx, sb.ops_host())
    vin = ctx.vec_from(psi)
    vout = ctx.vec(sb.psi_len)
    plan.apply(vin, vout)
    eager = vout.download()
    assert _rel(eager, osynth.hpsi(sb, psi)) < 1e-13
    v2 = ctx.vec(sb.psi_len)
    for r in range(4):
        vout.fill_gaussian(1, r, 1.0)        # scramble between replays
        plan.apply_graph(vin, vout)
        ctx.sync()
        assert np.array_equal(vout.download(), eager), r
        plan.apply(vin, v2)                  # eager work between replays
        ctx.sync()
    assert abs(vin.dot_re(vin) - 1.0) < 1e-12
    assert abs(vin.norm() - 1.0) < 1e-12
    y = ctx.vec_from(psi)
    y.axpby(2.0, vin, -1.0)
    assert np.allclose(y.download(), psi)


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_F32])
def test_lanczos_matches_exact_and_oracle(ctx, prec):
    hs = HermitianSuperblock(dL=[5, 9, 3, 12], dR=[7, 4, 6, 8], npairs=6, seed=3)
    H = hs.dense()
    assert np.allclose(H, H.conj().T, atol=1e-12)
    e_exact = np.linalg.eigvalsh(H)[0]
    rng = np.random.default_rng(4)
    psi0 = rng.normal(size=hs.n) + 1j * rng.normal(size=hs.n)
    e_or, psi_or, _ = olanczos.lanczos(hs.matvec, psi0, tol=1e-12, max_inner=100, max_outer=4)
    plan = hs.plan(ctx)
    assert _rel(plan.apply_host(psi0), hs.matvec(psi0)) < 1e-13
    v = ctx.vec_from(psi0)
    e, (nmv, nrs) = plan.lanczos(v, tol=1e-12, max_inner=100, max_outer=4, prec=prec,
                                 refine_iters=30 if prec != capi.PREC_F64 else 0)
    assert abs(e - e_exact) < 1e-9
    assert abs(e - e_or) < 1e-9
    psi = v.download()
    assert abs(np.vdot(psi, psi) - 1.0) < 1e-10
    assert np.linalg.norm(H @ psi - e * psi) < 1e-4


@pytest.mark.parametrize("prec", [capi.PREC_F32, capi.PREC_BF16X3])
def test_lanczos_refinement_record(ctx, prec):
    """The precision record of a mixed solve: its working-precision Ritz
    value, and the FP64 <psi|H psi> and residual norm of that Ritz vector,
    equal what the dense matrix gives for the vector a solve without
    refinement returns from the same start; the refined energy is the one
    returned and lies at or below <psi|H psi> (Rayleigh-Ritz over a space
    that holds psi). Without a refinement pass only the Ritz value is set."""
    hs = HermitianSuperblock(dL=[5, 9, 3, 12], dR=[7, 4, 6, 8], npairs=6, seed=3)
    H = hs.dense()
    scale = float(np.abs(np.linalg.eigvalsh(H)).max())
    plan = hs.plan(ctx)
    rng = np.random.default_rng(5)
    psi0 = rng.normal(size=hs.n) + 1j * rng.normal(size=hs.n)
    # one pass of 8 products: the mixed solution is not converged
    kw = dict(tol=1e-12, max_inner=8, max_outer=0, prec=prec)
    v = ctx.vec_from(psi0)
    e_m, _ = plan.lanczos(v, refine_iters=0, **kw)
    psi_m = v.download()
    v.close()
    r0 = plan.lanczos_refinement()
    assert r0["e_mixed"] == e_m
    assert all(np.isnan(r0[k]) for k in ("e_rayleigh_f64", "resid_f64", "e_refined", "w_rest"))
    v = ctx.vec_from(psi0)
    e_r, _ = plan.lanczos(v, refine_iters=3, **kw)
    psi_r = v.download()
    v.close()
    rec = plan.lanczos_refinement()
    assert rec["e_mixed"] == e_m   # the same start vector and products
    x = psi_m / np.linalg.norm(psi_m)
    hx = H @ x
    ray = float(np.vdot(x, hx).real)
    res = float(np.linalg.norm(hx - ray * x))
    assert res > 1e-4 * scale   # not converged: the residual check is not vacuous
    assert abs(rec["e_rayleigh_f64"] - ray) < 1e-12 * scale
    assert abs(rec["resid_f64"] - res) < 1e-9 * scale
    assert rec["e_refined"] == e_r
    assert e_r <= rec["e_rayleigh_f64"] + 1e-12 * scale
    # the weight of the refined vector outside the mixed one
    assert 0.0 < rec["w_rest"] < 1.0
    assert abs(rec["w_rest"] - (1.0 - abs(np.vdot(x, psi_r)) ** 2)) < 1e-9
    # the mixed arithmetic's energy error at its own solution: the
    # working precision's size, not zero
    assert 0.0 < abs(rec["e_mixed"] - rec["e_rayleigh_f64"]) < 1e-4 * scale
    v = ctx.vec_from(psi0)
    e64, _ = plan.lanczos(v, tol=1e-12, max_inner=8, max_outer=0, prec=capi.PREC_F64, refine_iters=3)
    v.close()
    r64 = plan.lanczos_refinement()
    assert r64["e_mixed"] == e64 and np.isnan(r64["e_rayleigh_f64"]) and np.isnan(r64["e_refined"])


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_BF16X3])
def test_lanczos_graph_replays_equal_eager_launches(ctx, prec, monkeypatch):
    """A small plan's Lanczos runs every H.psi as a replay of its captured
    graph (with the loop's eager vector work and host synchronisations
    between replays); the energy and Ritz vector equal those of eager
    launches, on a first call (capture) and on a second one (reuse)."""
    hs = HermitianSuperblock(dL=[40, 72, 24, 96], dR=[56, 32, 48, 64], npairs=6, seed=8)
    plan = hs.plan(ctx)
    rng = np.random.default_rng(9)
    psi0 = rng.normal(size=hs.n) + 1j * rng.normal(size=hs.n)
    runs = []
    for eager in (False, False, True):
        if eager:
            monkeypatch.setenv("TSC_LANCZOS_EAGER", "1")
        v = ctx.vec_from(psi0)
        e, (nmv, _) = plan.lanczos(v, tol=1e-12, max_inner=60, max_outer=2, prec=prec,
                                   refine_iters=4 if prec != capi.PREC_F64 else 0)
        runs.append((e, v.download(), nmv, plan.lanczos_replays()))
        v.close()
    (e1, x1, n1, r1), (e2, x2, n2, r2), (e3, x3, n3, r3) = runs
    assert r1 == n1 and r2 == n2 and r3 == 0
    assert n1 == n2 == n3
    assert abs(e1 - e3) < 1e-12 and abs(e2 - e3) < 1e-12
    assert np.linalg.norm(x1 - x3) < 1e-10 and np.linalg.norm(x2 - x3) < 1e-10


def _sectors(rng, dims):
    secs = [rng.normal(size=(d, e)) + 1j * rng.normal(size=(d, e)) for d, e in dims]
    nrm = np.sqrt(sum(np.sum(np.abs(x) ** 2) for x in secs))
    return [x / nrm for x in secs]


@pytest.mark.parametrize("side", [0, 1])
def test_density_matrix_and_truncation(ctx, side):
    rng = np.random.default_rng(20 + side)
    dims = [(37, 20), (1, 5), (70, 66), (0, 3), (12, 90), (5, 5)]
    secs = _sectors(rng, dims)
    psi = np.concatenate([x.ravel(order="F") for x in secs])
    v = ctx.vec_from(psi)
    sd = np.array([d for d, _ in dims], np.int32)
    ed = np.array([e for _, e in dims], np.int32)
    rho = capi.density_matrix(ctx, v, sd, ed, side)
    ref = otrunc.density_matrix(secs, side)
    off = 0
    for r in ref:
        n = r.shape[0]
        assert np.allclose(rho[off:off + n * n].reshape((n, n), order="F"), r, atol=1e-14)
        off += n * n
    twos = np.array([0, 2, 1, 3, 4, 0], np.int32)
    dim = sd if side == 0 else ed
    for kmin, kmax, terr in ((10, 40, 1e-4), (3, 500, 1e-12), (200, 400, 1e-3)):
        got = capi.rho_truncate(ctx, v, sd, ed, twos, side, kmin, kmax, terr)
        want = otrunc.truncation_operator(ref, twos, kmin, kmax, terr)
        assert list(got["kept"]) == list(want["kept"])
        assert got["num_state"] == want["num_state"]
        assert abs(got["trun_err"] - want["trun_err"]) < 1e-12
        assert abs(got["entropy"] - want["entropy"]) < 1e-11
        assert abs(got["entropy1"] - want["entropy1"]) < 1e-11
        assert np.allclose(got["eigval"], np.concatenate(want["eigval"]), atol=1e-13)
        for Ug, Uw, d in zip(got["U"], want["U"], dim):
            if Uw is None:
                assert Ug is None
                continue
            # same kept subspace (eigenvector phases are free)
            assert np.allclose(Ug @ Ug.conj().T, Uw @ Uw.conj().T, atol=1e-9)
        # host-rho entry point agrees with the device one
        got2 = capi.truncation_operator(ctx, dim, twos, rho, kmin, kmax, terr)
        assert list(got2["kept"]) == list(got["kept"])


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_F32])
def test_update_trun_matches_oracle(ctx, prec):
    rng = np.random.default_rng(30)
    blocks = []
    for r, c, kl, kr in ((40, 40, 12, 12), (70, 33, 20, 9), (5, 130, 5, 64), (1, 1, 1, 1), (90, 90, 0, 3)):
        O = rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c))
        Ul = np.linalg.qr(rng.normal(size=(r, max(kl, 1))) + 1j * rng.normal(size=(r, max(kl, 1))))[0][:, :kl]
        Ur = np.linalg.qr(rng.normal(size=(c, max(kr, 1))) + 1j * rng.normal(size=(c, max(kr, 1))))[0][:, :kr]
        blocks.append((O, Ul, Ur))
    got = capi.update_trun(ctx, blocks, prec)
    want = ouou.update_trun(blocks)
    for g, w in zip(got, want):
        assert g.shape == w.shape
        if w.size:
            assert _rel(g, w) < (1e-13 if prec == capi.PREC_F64 else 1e-5)


@pytest.mark.parametrize("prec", [capi.PREC_F64, capi.PREC_BF16X3])
def test_work_unit_targets_and_autotuner(ctx, prec):
    """Every work-unit cut the autotuner may choose (long contribution lists
    split into units whose partial tiles the reduce kernel sums) gives the
    oracle's H.psi, graph replays after a re-cut use the new units, and a
    plan of a tuned shape takes the cached choice."""
    sb = SyntheticSuperblock(nsec=30, nterms=10, dmin=8, dmax=300, seed=31)
    psi = sb.psi_host()
    ref = osynth.hpsi(sb, psi)
    plan = sb.make_plan(ctx, sb.ops_host())
    vin = ctx.vec_from(psi)
    vout = ctx.vec(sb.psi_len)
    plan.apply_graph(vin, vout)             # a graph of the first units
    ctx.sync()
    targets = (512.0, 2048.0, 4096.0, 8192.0, 32768.0) + ((1.0, 2048.0) if prec == capi.PREC_F64 else ())
    for target in targets:
        plan.autotune(vin, vout, prec, target=target)
        vout.fill_gaussian(3, 1, 1.0)
        plan.apply(vin, vout, prec)
        ctx.sync()
        eager = vout.download()
        assert _rel(eager, ref) < TOL[prec], target
        vout.fill_gaussian(3, 2, 1.0)
        plan.apply_graph(vin, vout, prec)
        ctx.sync()
        assert _rel(vout.download(), eager) < 1e-15, target
    chosen = plan.autotune(vin, vout, prec, force=True)
    # FP64 also weighs its 64x64 tile shape (1)
    assert chosen in ((2048.0, 4096.0, 8192.0, 1.0) if prec == capi.PREC_F64 else (2048.0, 4096.0, 8192.0))
    plan2 = sb.make_plan(ctx, sb.ops_host())
    assert plan2.autotune(vin, vout, prec) == chosen
    got = plan2.apply_host(psi, prec)
    assert _rel(got, ref) < TOL[prec]
    assert capi.Plan.autotune(plan, vin, vout, capi.PREC_F32) == 0.0
