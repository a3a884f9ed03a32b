"""tsc_hpsi_noalloc, the FP64 H.psi that allocates nothing (intermediates and
partial tiles in the library's static scratch buffer), against tsc_hpsi and
the oracle on the same plans: two calls in a row on one device, the second
on a smaller plan (the buffer still holds the first call's contents), and
the rejection of a plan whose intermediates do not fit. Runs through the
C-ABI."""
import numpy as np
import pytest

from oracle import synth as osynth
from tsc_tri_t1t2j1j2_amd import capi
from tsc_tri_t1t2j1j2_amd.synthetic import SyntheticSuperblock

pytestmark = pytest.mark.gpu

# FP64 bar of the H.psi parity tests (relative Frobenius error)
TOL_F64 = 1e-13


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_noalloc_hpsi_equals_hpsi_on_two_calls(ctx):
    cases = [SyntheticSuperblock(nsec=12, nterms=6, dmin=8, dmax=150, seed=77),
             SyntheticSuperblock(nsec=5, nterms=4, dmin=8, dmax=60, seed=78)]
    assert cases[1].psi_len < cases[0].psi_len
    for sb in cases:
        psi = sb.psi_host()
        plan = sb.make_plan(ctx, sb.ops_host())
        vin = ctx.vec_from(psi)
        ref_v, got_v = ctx.vec(sb.psi_len), ctx.vec(sb.psi_len)
        # the output starts as something other than the result
        got_v.upload(np.full(sb.psi_len, 7.0 - 3.0j))
        plan.apply(vin, ref_v, capi.PREC_F64)
        plan.apply_noalloc(vin, got_v)
        ctx.sync()
        ref, got = ref_v.download(), got_v.download()
        print(f"[noalloc] psi {sb.psi_len}: rel. difference to tsc_hpsi {_rel(got, ref):.2e}", flush=True)
        assert _rel(got, ref) < TOL_F64
        assert _rel(got, osynth.hpsi(sb, psi)) < TOL_F64
        for v in (vin, ref_v, got_v):
            v.close()
        plan.close()


def test_noalloc_rejects_a_plan_beyond_the_buffer(ctx):
    # 16 sectors of 400 x 400 and 4 terms: about 10M complex intermediates
    sb = SyntheticSuperblock(nsec=16, nterms=4, dmin=400, dmax=400, seed=79)
    plan = sb.make_plan(ctx, None)
    vin, vout = ctx.vec(sb.psi_len), ctx.vec(sb.psi_len)
    marker = np.full(sb.psi_len, 1.5 + 0.5j)
    vout.upload(marker)
    with pytest.raises(capi.TscError, match="scratch buffer holds"):
        plan.apply_noalloc(vin, vout)
    ctx.sync()
    assert np.array_equal(vout.download(), marker)
    for v in (vin, vout):
        v.close()
    plan.close()
