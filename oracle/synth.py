"""Direct H.psi of the synthetic superblock (oracle for the grouped
contraction engine on the isolated H.psi workload).

out[sigma_t(s)] += c[t,s] * O_L[t,s] . psi[s] . O_R[t,s]^T, evaluated term
by term in the order written, with no contraction-order choice: this is
the product definition that model_wave's block-block kernel applies per
sector (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:3385-3393).
"""
import numpy as np


def hpsi(sb, psi, pairs=None):
    """Full (or partial: an iterable of (t, s)) application; returns out."""
    out = np.zeros(sb.psi_len, dtype=np.complex128)
    it = pairs if pairs is not None else ((t, s) for t in range(sb.T) for s in range(sb.S))
    for t, s in it:
        g = int(sb.target[t, s])
        x = sb.psi_sector(psi, s)
        y = sb.coef[t, s] * (sb.op_L(t, s) @ x @ sb.op_R(t, s).T)
        o = int(sb.psi_off[g])
        out[o:o + y.size] += y.ravel(order="F")
    return out
