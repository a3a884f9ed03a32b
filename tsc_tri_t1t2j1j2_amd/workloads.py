"""DMRG superblocks with the structure of a real step and random values.

A superblock of the YC cylinder at its centre: the sys block holds sites
1..l, the env block sites l+3..N, and the centre sites are l+1, l+2. The
block sectors (N, 2S) and their multiplet dimensions are those of the
centre blocks of a real DMRG run (real_centre_sectors: the YC6 6x24 blocks
at M=6000, stored in data/yc6_24x6_m6000_centre.json, optionally with every
sector's multiplets scaled), or, for tests of small sizes, a Gaussian
profile around the mean filling (sector_profile). Every operator block
allowed by the quantum numbers is filled with seeded random numbers, so the
superblock has the sector structure, term list and GEMM shapes of a DMRG
step, with values that carry no physics.

Hermiticity is kept: block Hamiltonians are Hermitian per sector, n is
Hermitian, and the S blocks obey the adjoint relation of a Hermitian rank-1
tensor, <j'||S||j> = (-1)^(j'-j) conj(<j||S||j'>), so H is Hermitian and
Lanczos is meaningful (used to compare mixed precision against FP64).
"""
import json
import os

import numpy as np

from . import lattice
from .blocks import TotalBlock, TotalModel


def sector_profile(M, N0, sigma_n=3.0, sigma_s=2.2, max_twos=16, n_span=9):
    """[(N, 2S, dim)] with sum(dim) ~ M, ordered 2S descending then N
    descending (the order a truncated block inherits from Get_basis)."""
    raw = []
    for N in range(N0 - n_span, N0 + n_span + 1):
        for twos in range(N % 2, max_twos + 1, 2):
            S = twos / 2.0
            w = np.exp(-((N - N0) ** 2) / (2 * sigma_n ** 2)) * np.exp(-S * (S + 1) / (2 * sigma_s ** 2))
            raw.append((N, twos, w))
    tot = sum(w for (_, _, w) in raw)
    secs = []
    for (N, twos, w) in raw:
        d = int(round(M * w / tot))
        if d >= 1:
            secs.append((N, twos, d))
    secs.sort(key=lambda s: (-s[1], -s[0]))
    return secs


def _rand(rng, m, n):
    return (rng.normal(size=(m, n)) + 1j * rng.normal(size=(m, n))) / np.sqrt(max(n, 1))


def random_block(sites, keep, sectors, seed):
    """TotalModel with random Hermitian H and random c+, S, n for `keep`."""
    rng = np.random.default_rng(seed)
    idx = {(N, s): i for i, (N, s, _) in enumerate(sectors)}
    H = TotalBlock(0, 0)
    for i, (N, s, d) in enumerate(sectors):
        a = _rand(rng, d, d)
        H.blocks[(i, i)] = (a + a.conj().T) * np.sqrt(s + 1)
    ops = {}
    for site in keep:
        c = TotalBlock(1, 1)
        S = TotalBlock(0, 2)
        n = TotalBlock(0, 0)
        for k, (N, s, d) in enumerate(sectors):
            for ds in (-1, 1):
                b = idx.get((N + 1, s + ds))
                if b is not None:
                    c.blocks[(b, k)] = 0.3 * _rand(rng, sectors[b][2], d)
            for ds in (0, 2):
                b = idx.get((N, s + ds))
                if b is None or (ds == 0 and s == 0):
                    continue
                m = 0.3 * _rand(rng, sectors[b][2], d)
                if ds == 0:
                    m = (m + m.conj().T) / 2
                S.blocks[(b, k)] = m
                if ds == 2:
                    # <j'||S||j> = (-1)^(j'-j) conj(<j||S||j'>)  with 2j' = s, 2j = s+2
                    S.blocks[(k, b)] = -m.conj().T
            a = _rand(rng, d, d)
            n.blocks[(k, k)] = 0.3 * (a + a.conj().T)
        ops[site] = {"c": c, "S": S, "n": n}
    return TotalModel(sites, sectors, H, ops)


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def real_centre_sectors(scale=1.0):
    """(sys sectors, env sectors, description) of the YC6 6x24 centre blocks
    at M=6000 from a real run, every sector's multiplets times `scale`
    (rounded, at least 1)."""
    with open(os.path.join(DATA, "yc6_24x6_m6000_centre.json")) as fh:
        d = json.load(fh)

    def sc(secs):
        return [(int(n), int(s), max(1, int(round(m * scale)))) for (n, s, m) in secs]

    return sc(d["sys_sectors"]), sc(d["env_sectors"]), d


def centre_superblock(Lx, Ny, M, t1=1.0, t2=0.2, J1=0.4, J2=None, delta=12, seed=7, sectors=None):
    """(sys, env, s1, s2, bonds, Ntot) for the centre step at M multiplets;
    sectors = (sys sectors, env sectors) replaces the Gaussian profile."""
    N = Lx * Ny
    J2 = J1 * (t2 / t1) ** 2 if J2 is None else J2
    bonds = lattice.bonds(Lx, Ny, t1, t2, J1, J2)
    ntot = lattice.target_electrons(N, delta)
    l = N // 2 - 1
    s1, s2 = l + 1, l + 2
    sys_sites = list(range(1, l + 1))
    env_sites = list(range(N, l + 2, -1))
    nb = {}
    for (a, b, t, J) in bonds:
        nb.setdefault(a, []).append(b)
        nb.setdefault(b, []).append(a)
    keep_sys = [i for i in sys_sites if any(j > l for j in nb.get(i, []))]
    keep_env = [i for i in env_sites if any(j < l + 3 for j in nb.get(i, []))]
    if sectors is None:
        n0 = int(round(l * ntot / N))
        sys_secs = env_secs = sector_profile(M, n0)
    else:
        sys_secs, env_secs = sectors
    sys = random_block(sys_sites, keep_sys, sys_secs, seed)
    env = random_block(env_sites, keep_env, env_secs, seed + 1)
    return sys, env, s1, s2, bonds, ntot
