"""Density matrix and truncation (oracle for Get_density_matrix_sys/_env
and Get_truncation_operator, hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/den_mat.f90).
"""
import numpy as np


def density_matrix(psi_sectors, side):
    """rho_s = psi_s psi_s^H (side 0, den_mat.f90:41-43) or psi_s^H psi_s
    (side 1, :89-91)."""
    return [x @ x.conj().T if side == 0 else x.conj().T @ x for x in psi_sectors]


def quicksort_desc(a, idx, lower, upper):
    """QuickSort_Double (block.f90:2394-2436), 1-based, in place."""
    i, j = lower, upper
    piv = a[int((lower + upper) / 2.0)]
    while i <= j:
        while a[i] > piv:
            i += 1
        while a[j] < piv:
            j -= 1
        if i <= j:
            a[i], a[j] = a[j], a[i]
            idx[i], idx[j] = idx[j], idx[i]
            i += 1
            j -= 1
    if lower < j:
        quicksort_desc(a, idx, lower, j)
    if i < upper:
        quicksort_desc(a, idx, i, upper)


def select(eig_sectors, twos, kept_min, kept_max, trun_err):
    """Kept-state choice of Get_truncation_operator (den_mat.f90:255-364).
    eig_sectors: ascending eigenvalues per sector. Returns kept per sector,
    num_state, discarded weight, entropy (Entropys) and entropy1."""
    ev = [0.0]
    owner = [0]
    wts = [0]
    for s, e in enumerate(eig_sectors):
        for v in e:
            ev.append(float(v))
            owner.append(s)
            wts.append(int(twos[s]) + 1)
    n = len(ev) - 1
    idx = list(range(n + 1))
    entropy1 = 0.0
    for i in range(1, n + 1):
        if ev[i] >= 1.0e-15:
            v = ev[i] / wts[i]
            entropy1 -= v * np.log(v) * wts[i]
    if n > 0:
        quicksort_desc(ev, idx, 1, n)
    tot = sum(ev[1:])
    ev = [0.0] + [v / tot for v in ev[1:]]
    err = 0.0
    if n <= kept_min:
        num_state = n
    else:
        for i in range(kept_min + 1, n + 1):
            err += ev[i]
        num_state = kept_min
        # the reference's loop has no bound at the end of the spectrum (with
        # trun_err = 0 a rounding remainder could step past it); stop at n
        while err > trun_err and num_state < n:
            num_state += 1
            err -= ev[num_state]
            if err <= trun_err or num_state >= kept_max:
                break
    ent = 0.0
    for i in range(1, n + 1):
        if ev[i] >= 1.0e-15:
            ent -= ev[i] * np.log(ev[i])
    kept = [0] * len(eig_sectors)
    for k in range(1, num_state + 1):
        kept[owner[idx[k]]] += 1
    return kept, num_state, err, ent, entropy1


def truncation_operator(rho_sectors, twos, kept_min, kept_max, trun_err):
    """Eigensolve each sector (ascending, like ZHEEV 'V'), select, and return
    U_s = the last kept_s eigenvectors (den_mat.f90:392-398)."""
    eig, vec = [], []
    for r in rho_sectors:
        w, v = np.linalg.eigh(r)
        eig.append(w)
        vec.append(v)
    kept, num_state, err, ent, ent1 = select(eig, twos, kept_min, kept_max, trun_err)
    U = [v[:, v.shape[1] - k:] if k > 0 else None for v, k in zip(vec, kept)]
    return dict(eigval=eig, kept=kept, U=U, num_state=num_state, trun_err=err, entropy=ent, entropy1=ent1)
