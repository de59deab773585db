"""Real regular solid harmonics as Cartesian polynomials.

``S_lm(x, y, z) = sqrt(4 pi / (2l + 1)) r^l Y_lm(theta, phi)`` with the real spherical
harmonics of the usual table (the normalisation of ``_compute_S_l_m``,
atomic_orbital.py:2352-2584: ``S_10 = z``, ``S_11 = x``, ``S_1-1 = y``).  Built from the
closed form of the regular solid harmonics:

    Pi_l^m(z, r^2) = sqrt((l-m)!/(l+m)!) sum_k (-1)^k 2^-l C(l,k) C(2l-2k,l) (l-2k)!/(l-2k-m)! r^2k z^(l-2k-m)
    A_m = Re (x + i y)^m,  B_m = Im (x + i y)^m
    S_l0 = Pi_l^0,  S_lm = sqrt(2) Pi_l^m A_m,  S_l-m = sqrt(2) Pi_l^m B_m  (m > 0)

so that a spherical AO becomes a short sum of Cartesian monomials sharing one radial part,
which the gfx950 AO kernel evaluates like a Cartesian AO.
"""

from __future__ import annotations

from functools import lru_cache
from math import comb, factorial, sqrt


def _poly_mul(p, q):
    out = {}
    for (a, b, c), u in p.items():
        for (d, e, f), v in q.items():
            k = (a + d, b + e, c + f)
            out[k] = out.get(k, 0.0) + u * v
    return out


def _poly_pow(p, n):
    out = {(0, 0, 0): 1.0}
    for _ in range(n):
        out = _poly_mul(out, p)
    return out


@lru_cache(maxsize=None)
def solid_harmonic_terms(l: int, m: int):
    """Tuple of ((nx, ny, nz), coefficient) with S_lm = sum coef x^nx y^ny z^nz."""
    if l < 0 or abs(m) > l:
        raise ValueError("need |m| <= l")
    am = abs(m)
    r2 = {(2, 0, 0): 1.0, (0, 2, 0): 1.0, (0, 0, 2): 1.0}
    pi = {}
    pref = sqrt(factorial(l - am) / factorial(l + am))
    for k in range((l - am) // 2 + 1):
        c = (-1) ** k * 2.0**-l * comb(l, k) * comb(2 * l - 2 * k, l) * factorial(l - 2 * k) / factorial(l - 2 * k - am)
        term = _poly_mul(_poly_pow(r2, k), {(0, 0, l - 2 * k - am): 1.0})
        for key, v in term.items():
            pi[key] = pi.get(key, 0.0) + pref * c * v
    if m == 0:
        poly = pi
    else:
        # (x + i y)^|m| = sum_p C(m,p) x^p (i y)^(m-p); real part keeps even (m-p), imaginary odd
        ab = {}
        for p in range(am + 1):
            q = am - p
            if (q % 2 == 0) == (m > 0):
                sign = (-1) ** (q // 2) if m > 0 else (-1) ** ((q - 1) // 2)
                ab[(p, q, 0)] = ab.get((p, q, 0), 0.0) + sign * comb(am, p)
        poly = {k: sqrt(2.0) * v for k, v in _poly_mul(pi, ab).items()}
    return tuple(sorted((k, v) for k, v in poly.items() if v != 0.0))


def spherical_prim_normalisation(l: int, Z):
    """N = sqrt(2^(2l+3) (l+1)! (2Z)^(l+3/2) / ((2l+2)! sqrt(pi))) (atomic_orbital.py:2315-2320)."""
    import numpy as np

    Z = np.asarray(Z, dtype=np.float64)
    return np.sqrt(2.0 ** (2 * l + 3) * factorial(l + 1) * (2.0 * Z) ** (l + 1.5) / (factorial(2 * l + 2) * np.sqrt(np.pi)))
