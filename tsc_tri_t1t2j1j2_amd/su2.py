"""SU(2) recoupling coefficients of the two-site superblock.

Quantum numbers are (N, 2S) with spins doubled, as num_up / num_down of the
reference (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/common.f90:12-13). Operators
are stored as reduced matrix elements (Wigner-Eckart in Edmonds' form); a
scalar operator carries the factor sqrt(2S+1) of its multiplet
(DMRG_code/get_sys.f90:1811 divides it out again).

Coefficients used to place a block or site operator into an enlarged
(block x site) basis and to couple an L operator with an R operator into a
scalar across the cut (Edmonds 7.1.5-7.1.8); 6j symbols come from the
library's Racah routine (tsc_w6j, the role of W6JS, block.f90:3744-3795).
"""
from functools import lru_cache
from math import sqrt

from . import capi


@lru_cache(maxsize=None)
def w6j(a, b, c, d, e, f):
    return capi.w6j(a, b, c, d, e, f)


def par(x2):
    """(-1)^(x2/2), x2 even."""
    if x2 % 2:
        raise ValueError(f"odd doubled exponent {x2}")
    return -1.0 if (x2 // 2) % 2 else 1.0


@lru_cache(maxsize=None)
def embed_block(ja2, ja1, js, J2, J1, k2):
    """<(a' s) J'|| T(block) ||(a s) J> / <a'||T||a>."""
    w = w6j(ja2, J2, js, J1, ja1, k2)
    if w == 0.0:
        return 0.0
    return par(ja2 + js + J1 + k2) * sqrt((J1 + 1) * (J2 + 1)) * w


@lru_cache(maxsize=None)
def embed_site(js2, js1, ja, J2, J1, k2):
    """<(a s') J'|| U(site) ||(a s) J> / <s'||U||s> (fermion sign apart)."""
    w = w6j(js2, J2, ja, J1, js1, k2)
    if w == 0.0:
        return 0.0
    return par(ja + js1 + J2 + k2) * sqrt((J1 + 1) * (J2 + 1)) * w


@lru_cache(maxsize=None)
def scalar_pair(ja2, ja1, js2, js1, J, k2):
    """m-independent element of [T(block) x U(site)]^0 within one enlarged
    sector J, per unit reduced elements (fermion sign apart)."""
    w = w6j(ja2, ja1, k2, js1, js2, J)
    if w == 0.0:
        return 0.0
    return par(ja1 + js2 + k2 + J) / sqrt(k2 + 1) * w


@lru_cache(maxsize=None)
def cut_pair(jx2, jx1, jy2, jy1, S2, k2):
    """m-independent element of [X(L) x Y(R)]^0 between superblock sectors
    (x1, y1) -> (x2, y2) of total spin S2, per unit reduced elements."""
    w = w6j(jx2, jx1, k2, jy1, jy2, S2)
    if w == 0.0:
        return 0.0
    return par(jx1 + jy2 + k2 + S2) / sqrt(k2 + 1) * w


def ctilde_phase(j_ket, j_bra):
    """Phase of the annihilation tensor: c~[bra, ket] = phase * c+[ket, bra]^H
    with (-1)^(j_ket - j_bra + 1/2)."""
    return par(j_ket - j_bra + 1)
