"""Test helper: a host stand-in for the device objects the DMRG driver
uses (Plan, Vec, rho_truncate), built on the oracle. Lets the driver's
step logic be exercised on the CPU; it is never used by the product."""
import numpy as np

from oracle import lanczos as olanczos
from oracle import plan_exec
from oracle import truncation as otrunc
from tsc_tri_t1t2j1j2_amd import capi


class FakeVec:
    def __init__(self, x):
        self.x = np.array(x, dtype=complex)
        self.n = self.x.size

    def download(self):
        return self.x.copy()

    def close(self):
        pass


class FakeCtx:
    def vec_from(self, x):
        return FakeVec(np.asarray(x).ravel(order="F"))


class FakePlan:
    def __init__(self, ctx, ops, ops_elems, mid_elems, c1, o1, c2, o2, in_elems, out_elems):
        self.args = (np.asarray(ops, complex) if ops is not None else np.zeros(max(ops_elems, 1), complex),
                     c1, o1, c2, o2, mid_elems)
        self.in_elems, self.out_elems = in_elems, out_elems

    def _mv(self, x):
        ops, c1, o1, c2, o2, mid = self.args
        return plan_exec.execute(ops, c1, o1, c2, o2, mid, x, self.out_elems)

    def apply_host(self, x, prec=0):
        return self._mv(np.asarray(x, complex))

    def lanczos(self, psi, tol=1e-10, max_inner=100, max_outer=4, prec=0, refine_iters=0):
        e, v, nmv = olanczos.lanczos(self._mv, psi.x, tol=tol, max_inner=max_inner, max_outer=max_outer)
        psi.x = v
        return e, (nmv, 0)

    def flops(self):
        return 0.0

    def close(self):
        pass


def fake_rho_truncate(ctx, psi_vec, sd, ed, twos, side, kmin, kmax, terr):
    secs, off = [], 0
    for d, e in zip(sd, ed):
        secs.append(psi_vec.x[off:off + d * e].reshape((d, e), order="F"))
        off += d * e
    res = otrunc.truncation_operator(otrunc.density_matrix(secs, side), twos, kmin, kmax, terr)
    return dict(kept=np.array(res["kept"]), U=res["U"], num_state=res["num_state"], trun_err=res["trun_err"],
                eigval=np.concatenate(res["eigval"]), entropy=res["entropy"], entropy1=res["entropy1"])


def install(monkeypatch):
    monkeypatch.setattr(capi, "Plan", FakePlan)
    monkeypatch.setattr(capi, "rho_truncate", fake_rho_truncate)
    return FakeCtx()
