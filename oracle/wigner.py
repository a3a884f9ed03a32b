"""Wigner 6j symbol, exact (oracle for W6JS, hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/block.f90:3744-3795).

Racah's single-sum formula with doubled integer arguments, evaluated in
exact rational arithmetic (the square root is taken last), so it serves as
the reference value for the library's double-precision routine.
"""
from fractions import Fraction
from math import factorial, sqrt


def _triad(a, b, c):
    return (a >= 0 and b >= 0 and c >= 0 and abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0)


def _delta2(a, b, c):
    # Delta(abc)^2 as an exact fraction (doubled arguments)
    return Fraction(factorial((a + b - c) // 2) * factorial((a - b + c) // 2) * factorial((-a + b + c) // 2),
                    factorial((a + b + c) // 2 + 1))


def w6j(j1, j2, j3, l1, l2, l3):
    """{j1 j2 j3; l1 l2 l3} with all arguments doubled (2j)."""
    if not (_triad(j1, j2, j3) and _triad(j1, l2, l3) and _triad(l1, j2, l3) and _triad(l1, l2, j3)):
        return 0.0
    a1, a2 = (j1 + j2 + j3) // 2, (j1 + l2 + l3) // 2
    a3, a4 = (l1 + j2 + l3) // 2, (l1 + l2 + j3) // 2
    b1, b2, b3 = (j1 + j2 + l1 + l2) // 2, (j2 + j3 + l2 + l3) // 2, (j3 + j1 + l3 + l1) // 2
    s = Fraction(0)
    for t in range(max(a1, a2, a3, a4), min(b1, b2, b3) + 1):
        term = Fraction(factorial(t + 1),
                        factorial(t - a1) * factorial(t - a2) * factorial(t - a3) * factorial(t - a4)
                        * factorial(b1 - t) * factorial(b2 - t) * factorial(b3 - t))
        s += term if t % 2 == 0 else -term
    d2 = _delta2(j1, j2, j3) * _delta2(j1, l2, l3) * _delta2(l1, j2, l3) * _delta2(l1, l2, j3)
    # value = s * sqrt(d2); do it as sign * sqrt(s^2 d2) in exact arithmetic until the root
    v = s * s * d2
    return (1.0 if s >= 0 else -1.0) * sqrt(v.numerator) / sqrt(v.denominator)
