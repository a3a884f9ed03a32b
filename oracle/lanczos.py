"""Lanczos with full re-orthogonalisation (oracle for lanczos_memo_new,
hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/eigen.f90:4-199, levels = 1).

Restated step by step: v1 = psi/|psi|; alpha_k = Re<v_k|H v_k>
(wave_product keeps the real part, block.f90:1092-1108); w = H v_k -
alpha_k v_k - beta_{k-1} v_{k-1}; beta_k = |w|, stop if beta_k <= tol;
w /= beta_k; classical Gram-Schmidt of w against all v_j with real overlaps
and renormalisation (Gram_Schmidt_process, block.f90:1251-1289); lowest
eigenvalue of the tridiagonal matrix after each step (DSTEV 'N', :99);
converged when it moves by <= tol (:101) or the Krylov space is full on the
last restart; Ritz vector sum_i z_i v_i, normalised (:146-154); up to
max_outer+1 passes restarted from the Ritz vector (:36, :114-134).
"""
import numpy as np
from scipy.linalg import eigh_tridiagonal


def _re_dot(a, b):
    return float(np.real(np.vdot(a, b)))


def lanczos(matvec, psi0, tol=1e-10, max_inner=100, max_outer=4):
    psi = np.array(psi0, dtype=np.complex128)
    lastmin = 1024.0
    iter_out = 0
    matvecs = 0
    energy = None
    done = False
    while iter_out <= max_outer and not done:
        iter_out += 1
        alpha = np.zeros(max_inner + 1)
        beta = np.zeros(max_inner + 1)
        V = [psi / np.sqrt(_re_dot(psi, psi))]
        w = matvec(V[0])
        matvecs += 1
        alpha[1] = _re_dot(V[0], w)
        ritzmin = alpha[1]
        last = 1
        while last < max_inner:
            w = w - alpha[last] * V[last - 1]
            if last >= 2:
                w = w - beta[last - 1] * V[last - 2]
            beta[last] = np.sqrt(_re_dot(w, w))
            if abs(beta[last]) <= tol:
                done = True
                break
            w = w / beta[last]
            uv = [_re_dot(v, w) for v in V]
            uu = [_re_dot(v, v) for v in V]
            w = w - sum((a / b) * v for a, b, v in zip(uv, uu, V))
            w = w / np.sqrt(_re_dot(w, w))
            V.append(w)
            last += 1
            w = matvec(V[-1])
            matvecs += 1
            alpha[last] = _re_dot(V[-1], w)
            e0 = eigh_tridiagonal(alpha[1:last + 1], beta[1:last], eigvals_only=True)[0]
            if abs(e0 - ritzmin) <= tol or (last == max_inner and iter_out == max_outer):
                lastmin = e0
                done = True
                break
            ritzmin = e0
            lastmin = ritzmin
        d = alpha[1:last + 1]
        e = beta[1:last]
        if last == 1:
            z = np.array([1.0])
            e0 = d[0]
        else:
            vals, vecs = eigh_tridiagonal(d, e)
            e0, z = vals[0], vecs[:, 0]
        psi = sum(zi * v for zi, v in zip(z, V[:last]))
        psi = psi / np.sqrt(_re_dot(psi, psi))
        energy = e0
        lastmin = e0
    return energy, psi, matvecs
