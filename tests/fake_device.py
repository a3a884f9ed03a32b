"""Test helper: a host stand-in for the device objects the DMRG driver
uses (Plan, Vec, rho_truncate), built on the oracle. Lets the driver's
step logic be exercised on the CPU; it is never used by the product."""
import numpy as np

from oracle import lanczos as olanczos
from oracle import plan_exec
from oracle import truncation as otrunc
from tsc_tri_t1t2j1j2_amd import capi


class FakeVec:
    def __init__(self, x, live=None):
        self.x = np.array(x, dtype=complex)
        self.n = self.x.size
        self.live = live
        if live is not None:
            live.add(id(self))

    def download(self):
        return self.x.copy()

    def upload(self, x):
        self.x[:] = np.asarray(x, complex).ravel(order="K")

    def copy_range(self, off, src, src_off, n):
        self.x[off:off + n] = src.x[src_off:src_off + n]

    def upload_range(self, off, x):
        x = np.asarray(x, complex).ravel(order="K")
        self.x[off:off + x.size] = x

    def download_range(self, off, n):
        return self.x[off:off + n].copy()

    def norm(self):
        return float(np.linalg.norm(self.x))

    def dot_re(self, other):
        return float(np.vdot(self.x, other.x).real)

    def axpby(self, a, x, b):
        self.x[:] = a * x.x + b * self.x

    def copy_from(self, x):
        self.x[:] = x.x

    def close(self):
        if self.live is not None:
            self.live.discard(id(self))


class FakeCtx:
    """Records the vectors and plans that are open (`live`: their ids)."""

    def __init__(self):
        self.live = set()

    def vec_from(self, x):
        return FakeVec(np.asarray(x).ravel(order="F"), self.live)

    def vec(self, n):
        return FakeVec(np.zeros(int(n), complex), self.live)

    def sync(self):
        pass


class FakePlan:
    def __init__(self, ctx, ops, ops_elems, mid_elems, c1, o1, c2, o2, in_elems, out_elems):
        self.args = (np.asarray(ops, complex) if ops is not None else np.zeros(max(ops_elems, 1), complex),
                     c1, o1, c2, o2, mid_elems)
        self.in_elems, self.out_elems = in_elems, out_elems
        self.ops_vec = None
        self.live = getattr(ctx, "live", None)
        if self.live is not None:
            self.live.add(id(self))

    @classmethod
    def create_dev(cls, ctx, ops_vec, mid_elems, c1, o1, c2, o2, in_elems, out_elems):
        pl = cls(ctx, None, 1, mid_elems, c1, o1, c2, o2, in_elems, out_elems)
        pl.ops_vec = ops_vec   # read at each application, as the device borrows it
        return pl

    def _mv(self, x):
        ops, c1, o1, c2, o2, mid = self.args
        if self.ops_vec is not None:
            ops = self.ops_vec.x
        return plan_exec.execute(ops, c1, o1, c2, o2, mid, x, self.out_elems)

    def apply_host(self, x, prec=0):
        return self._mv(np.asarray(x, complex))

    def apply(self, vin, vout, prec=0):
        """Writes the output windows only; the rest of vout is untouched."""
        res = self._mv(vin.x)
        o2 = self.args[4]
        for ob in o2:
            ldc, m, n, c = int(ob["ldc"]), int(ob["m"]), int(ob["n"]), int(ob["c_off"])
            for j in range(n):
                vout.x[c + j * ldc:c + j * ldc + m] = res[c + j * ldc:c + j * ldc + m]

    def lanczos(self, psi, tol=1e-10, max_inner=100, max_outer=4, prec=0, refine_iters=0):
        e, v, nmv = olanczos.lanczos(self._mv, psi.x, tol=tol, max_inner=max_inner, max_outer=max_outer)
        psi.x = v
        # every precision is solved in FP64 here: a refinement pass would find
        # <psi|H psi> = e; its residual is that of the returned vector
        nan = float("nan")
        if prec != capi.PREC_F64 and refine_iters > 0:
            hv = self._mv(v)
            self._refine = dict(e_mixed=e, e_rayleigh_f64=float(np.vdot(v, hv).real),
                                resid_f64=float(np.linalg.norm(hv - np.vdot(v, hv).real * v)), e_refined=e,
                                w_rest=0.0)
        else:
            self._refine = dict(e_mixed=e, e_rayleigh_f64=nan, resid_f64=nan, e_refined=nan, w_rest=nan)
        return e, (nmv, 0)

    def lanczos_refinement(self):
        return dict(self._refine)

    def flops(self):
        return 0.0

    def autotune(self, vin, vout, prec, force=False, target=None):
        return 0.0

    def last_times(self):
        return 0.0, 0.0

    def close(self):
        if self.live is not None:
            self.live.discard(id(self))


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
    monkeypatch.setattr(capi, "rho_last_times", lambda ctx: [0.0, 0.0, 0.0, 0.0])
    return FakeCtx()
