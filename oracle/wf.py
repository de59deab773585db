"""FP64 NumPy restatement of the reference wavefunction (test infrastructure only).

Per-walker functions with the reference's shapes: AO/MO matrices are ``(n_orb, N_e)``,
the geminal is ``(N_up, N_up)``.  File:line citations are into kousuke-nakano/jQMC.
"""

from __future__ import annotations

import numpy as np

from jqmc_amd.atomic_orbital import AOs_cart_data, AOs_sphe_data
from jqmc_amd.molecular_orbital import MOs_data

from . import sphe

# --------------------------------------------------------------------------- AOs


def _prim_tables(aos: AOs_cart_data):
    ao_of_prim = np.asarray(aos.orbital_indices, dtype=np.int64)
    centers_prim = aos.structure_data.positions[np.asarray(aos.nucleus_index)[ao_of_prim]]
    return ao_of_prim, centers_prim


def aos_cart_value(aos: AOs_cart_data, r_carts: np.ndarray) -> np.ndarray:
    """``_compute_AOs_cart`` (atomic_orbital.py:2197-2290): reduce N c exp(-Z r^2) to AOs, times x^nx y^ny z^nz."""
    r = np.asarray(r_carts, dtype=np.float64).reshape(-1, 3)
    ao_of_prim, centers_prim = _prim_tables(aos)
    diff_p = r[None, :, :] - centers_prim[:, None, :]
    r2 = np.sum(diff_p**2, axis=-1)
    NR = aos._normalized_coefficients_np[:, None] * np.exp(-aos.exponents[:, None] * r2)
    NR_ao = np.zeros((aos.num_ao, r.shape[0]))
    np.add.at(NR_ao, ao_of_prim, NR)
    diff = r[None, :, :] - aos._atomic_center_carts_np[:, None, :]
    nx = np.asarray(aos.polynominal_order_x)[:, None]
    ny = np.asarray(aos.polynominal_order_y)[:, None]
    nz = np.asarray(aos.polynominal_order_z)[:, None]
    P = diff[..., 0] ** nx * diff[..., 1] ** ny * diff[..., 2] ** nz
    return NR_ao * P


def aos_cart_value_debug(aos: AOs_cart_data, r_carts: np.ndarray) -> np.ndarray:
    """Loop formulation of ``_compute_AOs_cart_debug`` (atomic_orbital.py:1984-2033), normalisation
    recomputed from ``scipy.special.factorial`` per primitive."""
    from scipy.special import factorial

    out = np.zeros((aos.num_ao, len(r_carts)))
    for a in range(aos.num_ao):
        R = aos._atomic_center_carts_np[a]
        l = aos.angular_momentums[a]
        nx, ny, nz = aos.polynominal_order_x[a], aos.polynominal_order_y[a], aos.polynominal_order_z[a]
        prims = [i for i, v in enumerate(aos.orbital_indices) if v == a]
        for k, rc in enumerate(r_carts):
            d = np.asarray(rc) - R
            val = 0.0
            for p in prims:
                Z = aos.exponents[p]
                N = np.sqrt(
                    (2.0 * Z / np.pi) ** 1.5 * (8.0 * Z) ** l * factorial(nx) * factorial(ny) * factorial(nz)
                    / (factorial(2 * nx) * factorial(2 * ny) * factorial(2 * nz))
                )
                val += N * aos.coefficients[p] * np.exp(-Z * np.dot(d, d))
            out[a, k] = val * d[0] ** nx * d[1] ** ny * d[2] ** nz
    return out


def aos_cart_vgl(aos: AOs_cart_data, r_carts: np.ndarray):
    """``_compute_AOs_value_grad_lap_cart`` (atomic_orbital.py:3377-3507): three radial moments
    NR, ZNR, Z^2NR reduced to AO rank, then polynomial algebra at AO rank."""
    r = np.asarray(r_carts, dtype=np.float64).reshape(-1, 3)
    ao_of_prim, centers_prim = _prim_tables(aos)
    diff_p = r[None, :, :] - centers_prim[:, None, :]
    r2p = np.sum(diff_p**2, axis=-1)
    Z = aos.exponents[:, None]
    pref = aos._normalized_coefficients_np[:, None] * np.exp(-Z * r2p)
    n_ao, n_e = aos.num_ao, r.shape[0]
    NR = np.zeros((n_ao, n_e))
    ZNR = np.zeros((n_ao, n_e))
    Z2NR = np.zeros((n_ao, n_e))
    np.add.at(NR, ao_of_prim, pref)
    np.add.at(ZNR, ao_of_prim, Z * pref)
    np.add.at(Z2NR, ao_of_prim, Z * Z * pref)
    diff = r[None, :, :] - aos._atomic_center_carts_np[:, None, :]
    x, y, z = diff[..., 0], diff[..., 1], diff[..., 2]
    r2 = x * x + y * y + z * z
    nx = np.asarray(aos.polynominal_order_x, dtype=np.float64)[:, None]
    ny = np.asarray(aos.polynominal_order_y, dtype=np.float64)[:, None]
    nz = np.asarray(aos.polynominal_order_z, dtype=np.float64)[:, None]

    def pw(b, n):
        return np.where(n <= 0, 1.0, b ** np.maximum(n, 0))

    px, py, pz = pw(x, nx), pw(y, ny), pw(z, nz)
    qx, qy, qz = pw(x, nx - 1), pw(y, ny - 1), pw(z, nz - 1)
    qqx, qqy, qqz = pw(x, nx - 2), pw(y, ny - 2), pw(z, nz - 2)
    P = px * py * pz
    val = NR * P
    gx = NR * (nx * qx * py * pz) - 2.0 * x * ZNR * P
    gy = NR * (px * ny * qy * pz) - 2.0 * y * ZNR * P
    gz = NR * (px * py * nz * qz) - 2.0 * z * ZNR * P
    lapP = nx * (nx - 1.0) * qqx * py * pz + px * ny * (ny - 1.0) * qqy * pz + px * py * nz * (nz - 1.0) * qqz
    rdg = x * nx * qx * py * pz + y * px * ny * qy * pz + z * px * py * nz * qz
    lap = NR * lapP - 4.0 * ZNR * rdg + (4.0 * r2 * Z2NR - 6.0 * ZNR) * P
    return val, gx, gy, gz, lap


def aos_value(aos, r_carts):
    if isinstance(aos, AOs_cart_data):
        return aos_cart_value(aos, r_carts)
    if isinstance(aos, AOs_sphe_data):
        return sphe.aos_sphe_value(aos, r_carts)
    raise NotImplementedError


def aos_vgl(aos, r_carts):
    if isinstance(aos, AOs_cart_data):
        return aos_cart_vgl(aos, r_carts)
    if isinstance(aos, AOs_sphe_data):
        return sphe.aos_sphe_vgl(aos, r_carts)
    raise NotImplementedError


# --------------------------------------------------------------------------- orbitals


def orb_value(orb, r_carts):
    """``compute_orb_api`` dispatch: AOs or ``C . AO`` (molecular_orbital.py:239-261)."""
    if isinstance(orb, MOs_data):
        return orb.mo_coefficients @ aos_value(orb.aos_data, r_carts)
    return aos_value(orb, r_carts)


def orb_vgl(orb, r_carts):
    """``compute_MOs_value_grad_lap`` (molecular_orbital.py:375-415) or the AO vgl."""
    v = aos_vgl(orb.aos_data if isinstance(orb, MOs_data) else orb, r_carts)
    if isinstance(orb, MOs_data):
        C = orb.mo_coefficients
        return tuple(C @ t for t in v)
    return v


# --------------------------------------------------------------------------- geminal


def geminal(gem, r_up, r_dn) -> np.ndarray:
    """``_compute_geminal_all_elements`` (determinant.py:1379-1402)."""
    lp, lu = gem.lambda_paired, gem.lambda_unpaired
    Pu = orb_value(gem.orb_data_up_spin, r_up)
    Pd = orb_value(gem.orb_data_dn_spin, r_dn)
    return np.hstack([Pu.T @ (lp @ Pd), Pu.T @ lu])


def geminal_up_one_row(gem, r_up_one, r_dn) -> np.ndarray:
    """``compute_geminal_up_one_row_elements`` (determinant.py:1429-1478)."""
    pu = orb_value(gem.orb_data_up_spin, np.reshape(r_up_one, (1, 3))).reshape(-1)
    Pd = orb_value(gem.orb_data_dn_spin, r_dn)
    return np.hstack([pu @ (gem.lambda_paired @ Pd), pu @ gem.lambda_unpaired])


def geminal_dn_one_column(gem, r_up, r_dn_one) -> np.ndarray:
    """``compute_geminal_dn_one_column_elements`` (determinant.py:1481-1526)."""
    Pu = orb_value(gem.orb_data_up_spin, r_up)
    pd = orb_value(gem.orb_data_dn_spin, np.reshape(r_dn_one, (1, 3))).reshape(-1)
    return Pu.T @ (gem.lambda_paired @ pd)


def as_regularization_fast(G: np.ndarray, Ginv: np.ndarray) -> float:
    """``compute_AS_regularization_factor_fast_update`` (determinant.py:1223-1260)."""
    F = np.sum(Ginv**2)
    S = min(np.min(np.sum(G**2, axis=1)), np.min(np.sum(G**2, axis=0)))
    SF = S * F
    return SF ** (-3.0 / 8.0) if SF > 0.0 else 0.0


def geminal_inverse(G: np.ndarray) -> np.ndarray:
    """SVD pseudo-inverse with rcond 1e-20 (jqmc_mcmc.py:4248-4261, _setting.py:136)."""
    U, s, Vt = np.linalg.svd(G, full_matrices=False)
    s_inv = np.where(s > 1e-20 * s[0], 1.0 / s, 0.0)
    return (Vt.T * s_inv[None, :]) @ U.T


def grads_laplacian_ln_det(gem, r_up, r_dn, Ginv):
    """``compute_grads_and_laplacian_ln_Det_fast`` (determinant.py:2140-2250)."""
    lp, lu = gem.lambda_paired, gem.lambda_unpaired
    n_up, n_dn = gem.num_electron_up, gem.num_electron_dn
    vu, gxu, gyu, gzu, lu_ = orb_vgl(gem.orb_data_up_spin, r_up)
    vd, gxd, gyd, gzd, ld_ = orb_vgl(gem.orb_data_dn_spin, r_dn)
    paired_dn = lp @ vd
    grad_up = np.zeros((n_up, 3))
    grad_dn = np.zeros((n_dn, 3))
    for k, gu in enumerate((gxu, gyu, gzu)):
        Gg = np.hstack([gu.T @ paired_dn, gu.T @ lu])  # d G[i,:] / d r_i
        grad_up[:, k] = np.einsum("ij,ji->i", Gg, Ginv)
    for k, gd in enumerate((gxd, gyd, gzd)):
        Gg = np.hstack([vu.T @ (lp @ gd), np.zeros((n_up, n_up - n_dn))])  # d G[:,j] / d r_j
        grad_dn[:, k] = np.einsum("ij,ji->i", Ginv, Gg)[:n_dn]
    lapG_up = np.hstack([lu_.T @ paired_dn, lu_.T @ lu])
    lapG_dn = np.hstack([vu.T @ (lp @ ld_), np.zeros((n_up, n_up - n_dn))])
    lap_up = -np.sum(grad_up**2, axis=1) + np.einsum("ij,ji->i", lapG_up, Ginv)
    lap_dn = -np.sum(grad_dn**2, axis=1) + np.einsum("ij,ji->i", Ginv, lapG_dn)[:n_dn]
    return grad_up, grad_dn, lap_up, lap_dn


# --------------------------------------------------------------------------- Jastrow


def jastrow_one_body(j1, r_up, r_dn) -> float:
    """``compute_Jastrow_one_body`` (jastrow_factor.py:648-726)."""
    R = j1.structure_data.positions
    Z = j1.effective_charges
    a = j1.jastrow_1b_param
    c = (2.0 * Z) ** 0.25
    A = (2.0 * Z) ** 0.75
    r = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    d = np.linalg.norm(r[:, None, :] - R[None, :, :], axis=-1)
    if j1.jastrow_1b_type == "exp":
        f = 1.0 / (2.0 * a) * (1.0 - np.exp(-a * c[None, :] * d))
        return float(np.sum(-A[None, :] * f))
    return float(np.sum(-A[None, :] * d / (2.0 * (1.0 + a * c[None, :] * d))))


def _j2_u(j2, r):
    a = j2.jastrow_2b_param
    if j2.jastrow_2b_type == "pade":
        return r / 2.0 * (1.0 + a * r) ** (-1.0)
    return 1.0 / (2.0 * a) * (1.0 - np.exp(-a * r))


def jastrow_two_body(j2, r_up, r_dn) -> float:
    """``compute_Jastrow_two_body`` (jastrow_factor.py:1180-1251): all pairs, same u for both spin pairs."""
    r = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    n = r.shape[0]
    iu, ju = np.triu_indices(n, k=1)
    return float(np.sum(_j2_u(j2, np.linalg.norm(r[iu] - r[ju], axis=-1))))


def jastrow_three_body(j3, r_up, r_dn) -> float:
    """``compute_Jastrow_three_body`` (jastrow_factor.py:1744-1801)."""
    J = j3.j_matrix[:, :-1]
    j1 = j3.j_matrix[:, -1]
    Au = orb_value(j3.orb_data, r_up)
    Ad = orb_value(j3.orb_data, r_dn)
    nu, nd = Au.shape[1], Ad.shape[1]
    Ku = np.tril(np.ones((nu, nu)), k=-1)
    Kd = np.tril(np.ones((nd, nd)), k=-1)
    return float(
        j1 @ Au @ np.ones(nu)
        + j1 @ Ad @ np.ones(nd)
        + np.trace(Au.T @ J @ Au @ Ku)
        + np.trace(Ad.T @ J @ Ad @ Kd)
        + np.ones(nu) @ Au.T @ J @ Ad @ np.ones(nd)
    )


def jastrow_part(jas, r_up, r_dn) -> float:
    """``compute_Jastrow_part`` (jastrow_factor.py:2127-2178)."""
    J = 0.0
    if jas.jastrow_one_body_data is not None:
        J += jastrow_one_body(jas.jastrow_one_body_data, r_up, r_dn)
    if jas.jastrow_two_body_data is not None:
        J += jastrow_two_body(jas.jastrow_two_body_data, r_up, r_dn)
    if jas.jastrow_three_body_data is not None:
        J += jastrow_three_body(jas.jastrow_three_body_data, r_up, r_dn)
    return J


def grads_laplacian_jastrow(jas, r_up, r_dn):
    """``compute_grads_and_laplacian_Jastrow_part`` (jastrow_factor.py:2982-3101) with the analytic
    J1 (960-1034), J2 (3434-3558) and J3 (4076-4156) pieces."""
    r_up = np.reshape(r_up, (-1, 3))
    r_dn = np.reshape(r_dn, (-1, 3))
    nu, nd = len(r_up), len(r_dn)
    gu, gd = np.zeros((nu, 3)), np.zeros((nd, 3))
    lu, ld = np.zeros(nu), np.zeros(nd)
    eps = 1.0e-12
    j1 = jas.jastrow_one_body_data
    if j1 is not None:
        R = j1.structure_data.positions
        Z = j1.effective_charges
        a = j1.jastrow_1b_param
        c = (2.0 * Z) ** 0.25
        A = (2.0 * Z) ** 0.75
        for rr, g, lap in ((r_up, gu, lu), (r_dn, gd, ld)):
            diff = rr[:, None, :] - R[None, :, :]
            r = np.maximum(np.linalg.norm(diff, axis=-1), eps)
            if j1.jastrow_1b_type == "exp":
                e = np.exp(-a * c[None, :] * r)
                fp = -A[None, :] * (c[None, :] / 2.0) * e
                fs = A[None, :] * (a * c[None, :] * c[None, :] / 2.0) * e
                lp = fs - A[None, :] * c[None, :] * e / r
            else:
                den = 1.0 + a * c[None, :] * r
                fp = -A[None, :] / (2.0 * den * den)
                fs = A[None, :] * a * c[None, :] / (den * den * den)
                lp = fs + 2.0 * fp / r
            g += np.sum(fp[..., None] * diff / r[..., None], axis=1)
            lap += np.sum(lp, axis=1)
    j2 = jas.jastrow_two_body_data
    if j2 is not None:
        a = j2.jastrow_2b_param

        def pair(diff):
            r = np.maximum(np.sqrt(np.sum(diff * diff, axis=-1)), eps)
            if j2.jastrow_2b_type == "pade":
                den = 1.0 + a * r
                fp = 0.5 / (den * den)
                lap = -a / (den * den * den) + 2.0 * fp / r
            else:
                e = np.exp(-a * r)
                fp = 0.5 * e
                lap = -(a / 2.0) * e + 2.0 * fp / r
            return (fp / r)[..., None] * diff, lap

        for rr, g, lap in ((r_up, gu, lu), (r_dn, gd, ld)):
            n = len(rr)
            if n > 1:
                diff = rr[:, None, :] - rr[None, :, :] + np.eye(n)[..., None] * np.array([1.0, 0.0, 0.0])
                gp, lpp = pair(diff)
                mask = 1.0 - np.eye(n)
                g += np.sum(gp * mask[..., None], axis=1)
                lap += np.sum(lpp * mask, axis=1)
        if nu > 0 and nd > 0:
            gp, lpp = pair(r_up[:, None, :] - r_dn[None, :, :])
            gu += np.sum(gp, axis=1)
            gd -= np.sum(gp, axis=0)
            lu += np.sum(lpp, axis=1)
            ld += np.sum(lpp, axis=0)
    j3 = jas.jastrow_three_body_data
    if j3 is not None:
        vu, gxu, gyu, gzu, lpu = orb_vgl(j3.orb_data, r_up)
        vd, gxd, gyd, gzd, lpd = orb_vgl(j3.orb_data, r_dn)
        j1v = j3.j_matrix[:, -1]
        J = j3.j_matrix[:, :-1]
        g_up = (
            j1v[:, None]
            + (J @ vu) @ np.tril(np.ones((nu, nu)), k=-1)
            + (J.T @ vu) @ np.triu(np.ones((nu, nu)), k=1)
            + (J @ vd) @ np.ones((nd, 1))
        )
        g_dn = (
            j1v[:, None]
            + (J @ vd) @ np.tril(np.ones((nd, nd)), k=-1)
            + (J.T @ vd) @ np.triu(np.ones((nd, nd)), k=1)
            + (J.T @ vu) @ np.ones((nu, 1))
        )
        gu += np.stack([np.sum(g_up * t, axis=0) for t in (gxu, gyu, gzu)], axis=1)
        gd += np.stack([np.sum(g_dn * t, axis=0) for t in (gxd, gyd, gzd)], axis=1)
        lu += np.sum(g_up * lpu, axis=0)
        ld += np.sum(g_dn * lpd, axis=0)
    return gu, gd, lu, ld


# --------------------------------------------------------------------------- wavefunction


def ln_wavefunction(wf, r_up, r_dn) -> float:
    """``evaluate_ln_wavefunction`` (wavefunction.py:677-720): J + ln|det G|."""
    G = geminal(wf.geminal_data, r_up, r_dn)
    return jastrow_part(wf.jastrow_data, r_up, r_dn) + float(np.log(np.abs(np.linalg.det(G))))


def wavefunction(wf, r_up, r_dn) -> float:
    """``evaluate_wavefunction`` (wavefunction.py:778-814): exp(J) det G."""
    G = geminal(wf.geminal_data, r_up, r_dn)
    return float(np.exp(jastrow_part(wf.jastrow_data, r_up, r_dn)) * np.linalg.det(G))


def kinetic_all_elements_fast(wf, r_up, r_dn, Ginv):
    """``compute_kinetic_energy_all_elements_fast_update`` (wavefunction.py:1141-1207)."""
    gJu, gJd, lJu, lJd = grads_laplacian_jastrow(wf.jastrow_data, r_up, r_dn)
    gDu, gDd, lDu, lDd = grads_laplacian_ln_det(wf.geminal_data, r_up, r_dn, Ginv)
    gu, gd = gJu + gDu, gJd + gDd
    return -0.5 * (lJu + lDu + np.sum(gu**2, axis=1)), -0.5 * (lJd + lDd + np.sum(gd**2, axis=1))


def kinetic_energy(wf, r_up, r_dn) -> float:
    """``compute_kinetic_energy`` (wavefunction.py:871-937) with the SVD inverse of G."""
    Ginv = geminal_inverse(geminal(wf.geminal_data, r_up, r_dn))
    tu, td = kinetic_all_elements_fast(wf, r_up, r_dn, Ginv)
    return float(np.sum(tu) + np.sum(td))
