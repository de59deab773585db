"""Counter-based random stream shared by the GPU kernels and the oracle.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11).  The reference draws from the JAX threefry key stream
(jqmc_mcmc.py:4322-4500), which cannot be generated without JAX; the MI355X path
instead feeds the same draw *sequence* from this stream to GPU and CPU alike.

Counter layout: ``(walker, move_lo, move_hi, (stream << 16) | sub)``, key = 64-bit seed.
Streams: 0 = VMC proposal, 1 = measurement rotation, 2 = LRDMC projection, 3 = LRDMC rotation.
"""

from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

STREAM_VMC = 0
STREAM_ROT = 1
STREAM_LRDMC = 2
STREAM_LRDMC_ROT = 3


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10; inputs are arrays (or scalars) of 32-bit values."""
    c0, c1, c2, c3 = (np.asarray(v, dtype=np.uint64) & MASK for v in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=np.uint64) & MASK
    k1 = np.asarray(k1, dtype=np.uint64) & MASK
    for r in range(10):
        if r > 0:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


def _key(seed: int):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def draw_words(seed: int, walker, move, stream: int, sub: int):
    """Four 32-bit words for (walker, move, stream, sub)."""
    k0, k1 = _key(seed)
    move = np.asarray(move, dtype=np.uint64)
    return philox4x32_10(walker, move & MASK, move >> np.uint64(32), (stream << 16) | sub, k0, k1)


def u53(hi, lo):
    """53-bit integer from two 32-bit words."""
    return (np.asarray(hi, dtype=np.uint64) << np.uint64(21)) | (np.asarray(lo, dtype=np.uint64) >> np.uint64(11))


def uniform01(hi, lo):
    """Uniform double in [0, 1)."""
    return u53(hi, lo).astype(np.float64) * 2.0**-53


def uniform01_open0(hi, lo):
    """Uniform double in (0, 1]."""
    return (u53(hi, lo).astype(np.float64) + 1.0) * 2.0**-53


def randint(word, n: int):
    """Integer in [0, n) from one 32-bit word by multiply-shift."""
    return ((np.asarray(word, dtype=np.uint64) * np.uint64(n)) >> np.uint64(32)).astype(np.int64)


def vmc_draws(seed: int, walker, move, n_up: int, n_dn: int):
    """The six draws of one Metropolis proposal, in the reference's order (jqmc_mcmc.py:4322-4500):
    rand_num in [0,N) -> up_index -> dn_index -> normal -> axis in [0,3) -> uniform."""
    w0, w1, w2, w3 = draw_words(seed, walker, move, STREAM_VMC, 0)
    w4, w5, w6, w7 = draw_words(seed, walker, move, STREAM_VMC, 1)
    w8, w9, _, _ = draw_words(seed, walker, move, STREAM_VMC, 2)
    rand_num = randint(w0, n_up + n_dn)
    up_index = randint(w1, max(n_up, 1))
    dn_index = randint(w2, max(n_dn, 1))
    axis = randint(w3, 3)
    u1 = uniform01_open0(w4, w5)
    u2 = uniform01(w6, w7)
    normal = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    b = uniform01(w8, w9)
    return rand_num, up_index, dn_index, normal, axis, b


def rotation_angles(seed: int, walker, step, stream: int = STREAM_ROT):
    """Three Euler angles in U(-2 pi, 2 pi) (jqmc_mcmc.py:4229-4233)."""
    a0, a1, a2, a3 = draw_words(seed, walker, step, stream, 0)
    b0, b1, _, _ = draw_words(seed, walker, step, stream, 1)
    lo, hi = -2.0 * np.pi, 2.0 * np.pi
    return [lo + (hi - lo) * uniform01(x, y) for x, y in ((a0, a1), (a2, a3), (b0, b1))]


def rotation_matrix_T(alpha, beta, gamma) -> np.ndarray:
    """``R.T`` for Euler angles, the matrix of ``_generate_rotation_matrix`` (jqmc_mcmc.py:4234-4245)."""
    ca, sa = np.cos(alpha), np.sin(alpha)
    cb, sb = np.cos(beta), np.sin(beta)
    cg, sg = np.cos(gamma), np.sin(gamma)
    R = np.array(
        [
            [cb * cg, cg * sa * sb - ca * sg, sa * sg + ca * cg * sb],
            [cb * sg, ca * cg + sa * sb * sg, ca * sb * sg - cg * sa],
            [-sb, cb * sa, ca * cb],
        ]
    )
    return R.T
