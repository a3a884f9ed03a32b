"""Seeded synthetic superblocks for the isolated H.psi benchmark.

Workload (the reference's H.psi abstracted to sectors x terms): S symmetry
sectors with left/right block dimensions d_L(s), d_R(s) drawn log-uniform in
[dmin, dmax]; T Hamiltonian terms. Term t maps source sector s to target
sigma_t(s) = s for even t and (s + 1 + t//2) mod S for odd t, and adds

    out[sigma_t(s)] += c[t, s] * O_L[t, s] . psi[s] . O_R[t, s]^T

with O_L in C^{d_L(sigma) x d_L(s)}, O_R in C^{d_R(sigma) x d_R(s)} complex
Gaussian (E|z|^2 = 1 / inner dimension) and c uniform in [-1, 1]. This is
the shape of the block-block terms of model_wave (hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/get_sys.f90:3131-3569: T = O_L psi_i, out_k += c T O_R^T).

Operator values come from a counter-based generator (splitmix64 + Box-Muller)
keyed by (seed, 2*(t*S+s) [+1 for O_R]) so the device can fill a full-size
arena without a host copy, and the host can regenerate any sample of it.
"""
import numpy as np

from . import capi

SEED = 20261203
MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & MASK
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & MASK
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & MASK
    return x ^ (x >> np.uint64(31))


def gaussian_host(n, seed, key, sigma):
    """Host copy of the device generator (tsc_vec_fill_gaussian)."""
    with np.errstate(over="ignore"):
        h = _splitmix64(np.uint64(seed) ^ _splitmix64(np.uint64(key)))
        i = np.arange(n, dtype=np.uint64)
        r1 = _splitmix64(h + np.uint64(2) * i)
        r2 = _splitmix64(h + np.uint64(2) * i + np.uint64(1))
    u1 = ((r1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (r2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    rad = np.sqrt(-2.0 * np.log(u1)) * (sigma / np.sqrt(2.0))
    return rad * np.cos(2 * np.pi * u2) + 1j * rad * np.sin(2 * np.pi * u2)


class SyntheticSuperblock:
    """Sector/term structure, offsets and the grouped-contraction plan arrays."""

    def __init__(self, nsec=400, nterms=24, dmin=8, dmax=1536, seed=SEED):
        rng = np.random.default_rng(seed)
        lo, hi = np.log(dmin), np.log(dmax)
        self.S, self.T, self.seed = nsec, nterms, seed
        self.dL = np.rint(np.exp(rng.uniform(lo, hi, nsec))).astype(np.int64)
        self.dR = np.rint(np.exp(rng.uniform(lo, hi, nsec))).astype(np.int64)
        self.coef = rng.uniform(-1.0, 1.0, (nterms, nsec))
        self.target = np.empty((nterms, nsec), dtype=np.int64)
        for t in range(nterms):
            for s in range(nsec):
                self.target[t, s] = s if t % 2 == 0 else (s + 1 + t // 2) % nsec
        sizes = self.dL * self.dR
        self.psi_off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        self.psi_len = int(sizes.sum())
        # operator arena: O_L[t,s] then O_R[t,s], in (t, s) order
        self.ol_off = np.zeros((nterms, nsec), np.int64)
        self.or_off = np.zeros((nterms, nsec), np.int64)
        off = 0
        for t in range(nterms):
            for s in range(nsec):
                g = self.target[t, s]
                self.ol_off[t, s] = off
                off += int(self.dL[g] * self.dL[s])
                self.or_off[t, s] = off
                off += int(self.dR[g] * self.dR[s])
        self.ops_len = off

    # --- values --------------------------------------------------------
    def key_L(self, t, s):
        return 2 * (t * self.S + s)

    def key_R(self, t, s):
        return 2 * (t * self.S + s) + 1

    def op_L(self, t, s):
        g = self.target[t, s]
        n = int(self.dL[g] * self.dL[s])
        v = gaussian_host(n, self.seed, self.key_L(t, s), 1.0 / np.sqrt(self.dL[s]))
        return v.reshape((int(self.dL[g]), int(self.dL[s])), order="F")

    def op_R(self, t, s):
        g = self.target[t, s]
        n = int(self.dR[g] * self.dR[s])
        v = gaussian_host(n, self.seed, self.key_R(t, s), 1.0 / np.sqrt(self.dR[s]))
        return v.reshape((int(self.dR[g]), int(self.dR[s])), order="F")

    def ops_host(self):
        out = np.empty(self.ops_len, dtype=np.complex128)
        for t in range(self.T):
            for s in range(self.S):
                a, b = self.op_L(t, s), self.op_R(t, s)
                out[self.ol_off[t, s]:self.ol_off[t, s] + a.size] = a.ravel(order="F")
                out[self.or_off[t, s]:self.or_off[t, s] + b.size] = b.ravel(order="F")
        return out

    def fill_plan_ops(self, plan):
        """Fill the plan's operator arena on the device, block by block."""
        for t in range(self.T):
            for s in range(self.S):
                g = self.target[t, s]
                plan.fill_ops_gaussian(self.ol_off[t, s], self.dL[g] * self.dL[s], self.seed, self.key_L(t, s),
                                       1.0 / np.sqrt(self.dL[s]))
                plan.fill_ops_gaussian(self.or_off[t, s], self.dR[g] * self.dR[s], self.seed, self.key_R(t, s),
                                       1.0 / np.sqrt(self.dR[s]))

    def psi_host(self, key=0xC0FFEE):
        v = gaussian_host(self.psi_len, self.seed, key, 1.0)
        return v / np.linalg.norm(v)

    def psi_sector(self, psi, s):
        o = int(self.psi_off[s])
        return psi[o:o + int(self.dL[s] * self.dR[s])].reshape((int(self.dL[s]), int(self.dR[s])), order="F")

    # --- work model ------------------------------------------------------
    def order_costs(self, t, s):
        g = self.target[t, s]
        dl, dr, dl2, dr2 = (int(x) for x in (self.dL[s], self.dR[s], self.dL[g], self.dR[g]))
        left_first = dl2 * dl * dr + dl2 * dr * dr2
        right_first = dl * dr * dr2 + dl2 * dl * dr2
        return left_first, right_first

    def algorithmic_flops(self):
        """F = sum_{t,s} 8 * min(left-first, right-first) complex MACs."""
        f = 0.0
        for t in range(self.T):
            for s in range(self.S):
                f += 8.0 * min(self.order_costs(t, s))
        return f

    def algorithmic_bytes(self):
        """FP64 bytes: every operator block once, psi read and written once."""
        return 16.0 * self.ops_len + 16.0 * 2 * self.psi_len

    # --- plan ----------------------------------------------------------
    def plan_arrays(self):
        """Stage-1 / stage-2 contribution and output-block arrays."""
        T, S = self.T, self.S
        c1 = np.zeros(T * S, dtype=capi.CONTRIB_DTYPE)
        o1 = np.zeros(T * S, dtype=capi.OUTBLOCK_DTYPE)
        contribs2 = [[] for _ in range(S)]
        mid = 0
        i = 0
        for t in range(T):
            for s in range(S):
                g = int(self.target[t, s])
                dl, dr, dl2, dr2 = (int(x) for x in (self.dL[s], self.dR[s], self.dL[g], self.dR[g]))
                lf, rf = self.order_costs(t, s)
                e = c1[i]
                e["coef_re"] = 1.0
                b = o1[i]
                b["c_buf"] = capi.BUF_MID
                b["c_off"] = mid
                b["cbeg"] = i
                b["cend"] = i + 1
                c2 = np.zeros((), dtype=capi.CONTRIB_DTYPE)
                c2["coef_re"] = self.coef[t, s]
                if lf <= rf:
                    # Y = O_L psi_s (dl2 x dr); out += c Y O_R^T
                    e["a_off"], e["lda"], e["a_buf"], e["a_op"] = self.ol_off[t, s], dl2, capi.BUF_OPS, capi.OP_N
                    e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = self.psi_off[s], dl, capi.BUF_IN, capi.OP_N
                    e["k"] = dl
                    b["m"], b["n"], b["ldc"] = dl2, dr, dl2
                    c2["a_off"], c2["lda"], c2["a_buf"], c2["a_op"] = mid, dl2, capi.BUF_MID, capi.OP_N
                    c2["b_off"], c2["ldb"], c2["b_buf"], c2["b_op"] = self.or_off[t, s], dr2, capi.BUF_OPS, capi.OP_T
                    c2["k"] = dr
                    mid += dl2 * dr
                else:
                    # Z = psi_s O_R^T (dl x dr2); out += c O_L Z
                    e["a_off"], e["lda"], e["a_buf"], e["a_op"] = self.psi_off[s], dl, capi.BUF_IN, capi.OP_N
                    e["b_off"], e["ldb"], e["b_buf"], e["b_op"] = self.or_off[t, s], dr2, capi.BUF_OPS, capi.OP_T
                    e["k"] = dr
                    b["m"], b["n"], b["ldc"] = dl, dr2, dl
                    c2["a_off"], c2["lda"], c2["a_buf"], c2["a_op"] = self.ol_off[t, s], dl2, capi.BUF_OPS, capi.OP_N
                    c2["b_off"], c2["ldb"], c2["b_buf"], c2["b_op"] = mid, dl, capi.BUF_MID, capi.OP_N
                    c2["k"] = dl
                    mid += dl * dr2
                contribs2[g].append(c2)
                i += 1
        c2_all = []
        o2 = np.zeros(S, dtype=capi.OUTBLOCK_DTYPE)
        for g in range(S):
            o2[g]["c_off"] = self.psi_off[g]
            o2[g]["ldc"] = self.dL[g]
            o2[g]["m"] = self.dL[g]
            o2[g]["n"] = self.dR[g]
            o2[g]["c_buf"] = capi.BUF_OUT
            o2[g]["cbeg"] = len(c2_all)
            c2_all.extend(contribs2[g])
            o2[g]["cend"] = len(c2_all)
        c2 = np.array(c2_all, dtype=capi.CONTRIB_DTYPE) if c2_all else np.zeros(0, capi.CONTRIB_DTYPE)
        return c1, o1, c2, o2, mid

    def make_plan(self, ctx, ops=None):
        """Build the device plan; ops=None fills the arena on the device."""
        c1, o1, c2, o2, mid = self.plan_arrays()
        plan = capi.Plan(ctx, ops, self.ops_len, mid, c1, o1, c2, o2, self.psi_len, self.psi_len)
        if ops is None:
            self.fill_plan_ops(plan)
        return plan
