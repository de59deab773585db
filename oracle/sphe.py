"""Real-solid-harmonic AOs (test infrastructure only).

``_compute_AOs_sphe`` (atomic_orbital.py:2293-2349): AO = sum_p N_p(l) sqrt((2l+1)/4pi) c_p
exp(-Z_p r^2) S_lm(r - R), with S_lm = sqrt(4pi/(2l+1)) r^l Y_lm (real).  The oracle obtains
the Cartesian coefficients of S_lm independently of the product: it solves for them from
``scipy.special`` complex spherical harmonics (Condon-Shortley phase) at sample points,
``Y_lm(real) = sqrt(2) (-1)^m Re Y_l^m`` (m > 0), ``Y_l^0`` (m = 0),
``sqrt(2) (-1)^m Im Y_l^|m|`` (m < 0).
"""

from __future__ import annotations

from functools import lru_cache
from math import factorial

import numpy as np


def _sph_complex(l, m, polar, azimuth):
    import scipy.special as sp

    if hasattr(sp, "sph_harm_y"):
        return sp.sph_harm_y(l, m, polar, azimuth)
    return sp.sph_harm(m, l, azimuth, polar)


def real_Y(l, m, pts):
    """Real spherical harmonic at unit directions pts (n, 3)."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    polar = np.arccos(np.clip(z, -1.0, 1.0))
    azimuth = np.arctan2(y, x)
    if m == 0:
        return np.real(_sph_complex(l, 0, polar, azimuth))
    Y = _sph_complex(l, abs(m), polar, azimuth)
    if m > 0:
        return np.sqrt(2.0) * (-1) ** m * np.real(Y)
    return np.sqrt(2.0) * (-1) ** m * np.imag(Y)


@lru_cache(maxsize=None)
def solid_harmonic_monomials(l: int, m: int):
    """Monomial exponents and coefficients of S_lm, solved from scipy values (exact up to roundoff)."""
    mons = [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]
    rng = np.random.default_rng(1234 + 31 * l + m)
    pts = rng.normal(size=(4 * len(mons) + 8, 3))
    r = np.linalg.norm(pts, axis=1)
    unit = pts / r[:, None]
    target = np.sqrt(4.0 * np.pi / (2 * l + 1)) * r**l * real_Y(l, m, unit)
    M = np.stack([pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c for a, b, c in mons], axis=1)
    coef = np.linalg.lstsq(M, target, rcond=None)[0]
    coef[np.abs(coef) < 1e-12] = 0.0
    return mons, coef


def _sph_norm(l, Z):
    return np.sqrt(2.0 ** (2 * l + 3) * factorial(l + 1) * (2.0 * Z) ** (l + 1.5) / (factorial(2 * l + 2) * np.sqrt(np.pi)))


def _poly_vgl(l, m, d):
    """S, grad S, lap S (= 0 for a harmonic polynomial, computed anyway) at displacements d (n, 3)."""
    mons, coef = solid_harmonic_monomials(l, m)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]

    def pw(b, n):
        return b**n if n >= 0 else np.zeros_like(b)

    S = np.zeros(len(d))
    g = np.zeros((3, len(d)))
    lap = np.zeros(len(d))
    for (a, b, c), k in zip(mons, coef):
        if k == 0.0:
            continue
        S += k * pw(x, a) * pw(y, b) * pw(z, c)
        g[0] += k * a * pw(x, a - 1) * pw(y, b) * pw(z, c)
        g[1] += k * b * pw(x, a) * pw(y, b - 1) * pw(z, c)
        g[2] += k * c * pw(x, a) * pw(y, b) * pw(z, c - 1)
        lap += k * (a * (a - 1) * pw(x, a - 2) * pw(y, b) * pw(z, c) + b * (b - 1) * pw(x, a) * pw(y, b - 2) * pw(z, c)
                    + c * (c - 1) * pw(x, a) * pw(y, b) * pw(z, c - 2))
    return S, g, lap


def _radial_moments(aos, r):
    n_ao = aos.num_ao
    NR = np.zeros((n_ao, len(r)))
    ZNR = np.zeros_like(NR)
    Z2NR = np.zeros_like(NR)
    centers = aos._atomic_center_carts_np
    for p, a in enumerate(aos.orbital_indices):
        l = aos.angular_momentums[a]
        Z = aos.exponents[p]
        d = r - centers[a]
        r2 = np.sum(d * d, axis=1)
        pref = _sph_norm(l, Z) * np.sqrt((2 * l + 1) / (4.0 * np.pi)) * aos.coefficients[p] * np.exp(-Z * r2)
        NR[a] += pref
        ZNR[a] += Z * pref
        Z2NR[a] += Z * Z * pref
    return NR, ZNR, Z2NR


def aos_sphe_value(aos, r_carts):
    r = np.asarray(r_carts, dtype=np.float64).reshape(-1, 3)
    NR, _, _ = _radial_moments(aos, r)
    out = np.zeros_like(NR)
    centers = aos._atomic_center_carts_np
    for a in range(aos.num_ao):
        S, _, _ = _poly_vgl(aos.angular_momentums[a], aos.magnetic_quantum_numbers[a], r - centers[a])
        out[a] = NR[a] * S
    return out


def aos_sphe_vgl(aos, r_carts):
    """Value, gradient, Laplacian (the radial-moment algebra of atomic_orbital.py:3510-3591)."""
    r = np.asarray(r_carts, dtype=np.float64).reshape(-1, 3)
    NR, ZNR, Z2NR = _radial_moments(aos, r)
    n = NR.shape
    v, gx, gy, gz, lap = (np.zeros(n) for _ in range(5))
    centers = aos._atomic_center_carts_np
    for a in range(aos.num_ao):
        d = r - centers[a]
        S, g, lS = _poly_vgl(aos.angular_momentums[a], aos.magnetic_quantum_numbers[a], d)
        r2 = np.sum(d * d, axis=1)
        rdg = d[:, 0] * g[0] + d[:, 1] * g[1] + d[:, 2] * g[2]
        v[a] = NR[a] * S
        gx[a] = NR[a] * g[0] - 2.0 * d[:, 0] * ZNR[a] * S
        gy[a] = NR[a] * g[1] - 2.0 * d[:, 1] * ZNR[a] * S
        gz[a] = NR[a] * g[2] - 2.0 * d[:, 2] * ZNR[a] * S
        lap[a] = NR[a] * lS - 4.0 * ZNR[a] * rdg + (4.0 * r2 * Z2NR[a] - 6.0 * ZNR[a]) * S
    return v, gx, gy, gz, lap
