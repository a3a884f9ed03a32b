"""Random Hermitian block-sparse superblock operators for engine tests.

H acts on psi = (psi_s)_s, psi_s in C^{dL_s x dR_s} (column-major), as
  (H psi)_s = A_s psi_s + psi_s B_s^T                  (A_s, B_s Hermitian)
            + sum_pairs  c  L psi_u R^T   into sector g  (pair u -> g)
            +            c  L^H psi_g conj(R)  into sector u  (its adjoint)
so H is Hermitian by construction; the dense matrix is available for
exact diagonalisation.
"""
import numpy as np

from tsc_tri_t1t2j1j2_amd import capi


def _herm(rng, n):
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (a + a.conj().T) / (2 * np.sqrt(n))


class HermitianSuperblock:
    def __init__(self, dL, dR, npairs, seed):
        rng = np.random.default_rng(seed)
        self.dL = np.asarray(dL, np.int64)
        self.dR = np.asarray(dR, np.int64)
        S = len(dL)
        self.S = S
        sizes = self.dL * self.dR
        self.off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.n = int(sizes.sum())
        self.A = [_herm(rng, int(d)) for d in self.dL]
        self.B = [_herm(rng, int(d)) for d in self.dR]
        self.pairs = []
        for _ in range(npairs):
            u, g = rng.integers(0, S, 2)
            L = (rng.normal(size=(self.dL[g], self.dL[u])) + 1j * rng.normal(size=(self.dL[g], self.dL[u]))) / 4
            R = (rng.normal(size=(self.dR[g], self.dR[u])) + 1j * rng.normal(size=(self.dR[g], self.dR[u]))) / 4
            c = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
            self.pairs.append((int(u), int(g), L, R, c))

    def sector(self, v, s):
        o = int(self.off[s])
        return v[o:o + int(self.dL[s] * self.dR[s])].reshape((int(self.dL[s]), int(self.dR[s])), order="F")

    def matvec(self, v):
        out = np.zeros(self.n, np.complex128)

        def add(s, y):
            o = int(self.off[s])
            out[o:o + y.size] += y.ravel(order="F")

        for s in range(self.S):
            x = self.sector(v, s)
            add(s, self.A[s] @ x + x @ self.B[s].T)
        for u, g, L, R, c in self.pairs:
            add(g, c * (L @ self.sector(v, u) @ R.T))
            add(u, np.conj(c) * (L.conj().T @ self.sector(v, g) @ R.conj()))
        return out

    def dense(self):
        eye = np.eye(self.n, dtype=np.complex128)
        return np.stack([self.matvec(eye[:, i]) for i in range(self.n)], axis=1)

    def plan(self, ctx):
        """Two-stage plan: stage 1 T = L psi_u (or L^H psi_g); stage 2 sums
        A psi, psi B^T (as A.psi and psi.B^T GEMMs) and c T R^T terms."""
        ops = []
        ooff = [0]

        def put(M):
            M = np.asarray(M, np.complex128)
            o = ooff[0]
            ops.append(M.ravel(order="F"))
            ooff[0] += M.size
            return o

        c1, o1 = [], []
        per_target = [[] for _ in range(self.S)]
        mid = 0

        def contrib(a_off, lda, a_buf, a_op, b_off, ldb, b_buf, b_op, k, coef=1.0):
            e = np.zeros((), capi.CONTRIB_DTYPE)
            e["a_off"], e["lda"], e["a_buf"], e["a_op"] = a_off, lda, a_buf, a_op
            e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = b_off, ldb, b_buf, b_op
            e["k"] = k
            e["coef_re"], e["coef_im"] = np.real(coef), np.imag(coef)
            return e

        for s in range(self.S):
            dl, dr = int(self.dL[s]), int(self.dR[s])
            a = put(self.A[s])
            b = put(self.B[s])
            per_target[s].append(contrib(a, dl, capi.BUF_OPS, capi.OP_N, self.off[s], dl, capi.BUF_IN, capi.OP_N, dl))
            per_target[s].append(contrib(self.off[s], dl, capi.BUF_IN, capi.OP_N, b, dr, capi.BUF_OPS, capi.OP_T, dr))
        for u, g, L, R, c in self.pairs:
            lo, ro = put(L), put(R)
            dlu, dru, dlg, drg = (int(x) for x in (self.dL[u], self.dR[u], self.dL[g], self.dR[g]))
            # forward: T = L psi_u (dlg x dru) ; out_g += c T R^T
            c1.append(contrib(lo, dlg, capi.BUF_OPS, capi.OP_N, self.off[u], dlu, capi.BUF_IN, capi.OP_N, dlu))
            ob = np.zeros((), capi.OUTBLOCK_DTYPE)
            ob["c_off"], ob["ldc"], ob["m"], ob["n"], ob["c_buf"] = mid, dlg, dlg, dru, capi.BUF_MID
            ob["cbeg"], ob["cend"] = len(c1) - 1, len(c1)
            o1.append(ob)
            per_target[g].append(contrib(mid, dlg, capi.BUF_MID, capi.OP_N, ro, drg, capi.BUF_OPS, capi.OP_T, dru, c))
            mid += dlg * dru
            # adjoint: T = L^H psi_g (dlu x drg) ; out_u += conj(c) T conj(R)
            c1.append(contrib(lo, dlg, capi.BUF_OPS, capi.OP_C, self.off[g], dlg, capi.BUF_IN, capi.OP_N, dlg))
            ob = np.zeros((), capi.OUTBLOCK_DTYPE)
            ob["c_off"], ob["ldc"], ob["m"], ob["n"], ob["c_buf"] = mid, dlu, dlu, drg, capi.BUF_MID
            ob["cbeg"], ob["cend"] = len(c1) - 1, len(c1)
            o1.append(ob)
            per_target[u].append(contrib(mid, dlu, capi.BUF_MID, capi.OP_N, ro, drg, capi.BUF_OPS, capi.OP_CONJ, drg,
                                         np.conj(c)))
            mid += dlu * drg
        c2, o2 = [], []
        for s in range(self.S):
            ob = np.zeros((), capi.OUTBLOCK_DTYPE)
            ob["c_off"], ob["ldc"], ob["m"], ob["n"], ob["c_buf"] = self.off[s], self.dL[s], self.dL[s], self.dR[s], \
                capi.BUF_OUT
            ob["cbeg"] = len(c2)
            c2.extend(per_target[s])
            ob["cend"] = len(c2)
            o2.append(ob)
        ops_arr = np.concatenate(ops)
        return capi.Plan(ctx, ops_arr, ops_arr.size, mid, np.array(c1, capi.CONTRIB_DTYPE),
                         np.array(o1, capi.OUTBLOCK_DTYPE), np.array(c2, capi.CONTRIB_DTYPE),
                         np.array(o2, capi.OUTBLOCK_DTYPE), self.n, self.n)
