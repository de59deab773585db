"""Position derivatives of E_L and ln|Psi| and the space-warp force terms (test infrastructure only).

The reference obtains dE_L/dr, dE_L/dR (``grad(compute_local_energy, argnums=(0, 1, 2))``,
jqmc_mcmc.py:744-780, hamiltonians.py:179-222) and d ln|Psi|/d(r, R)
(``grad(evaluate_ln_wavefunction, argnums=(0, 1, 2))``, jqmc_mcmc.py:4746) by JAX reverse-mode
differentiation, and assembles per walker (jqmc_mcmc.py:796-852)

    F_HF = dE_L/dR + sum_i omega_ai dE_L/dr_i
    F_PP = d ln|Psi|/dR + sum_i omega_ai d ln|Psi|/dr_i + 1/2 sum_i d omega_ai/dr_i

with the space-warp weights omega_ai = r_ai^-4 / sum_b r_bi^-4 (swct.py:63-150).

This module restates the same derivatives as an explicit adjoint (reverse sweep) of the local
energy: the kinetic energy written as T = -1/2 sum_e [L_e + 2 D1_e . gJ_e + lJ_e + |gJ_e|^2]
(L_e, D1_e the Laplacian and gradient of the changed geminal row/column contracted with G^-1,
gJ_e, lJ_e the Jastrow gradient and Laplacian), the ECP non-local term as
sum W(r_e - R_a) rho_e(p) with rho_e(p) = (phi(p) . q_e) exp(Delta J), and every AO evaluated
at a point x contributing to the point's adjoint and, with the opposite sign, to its centre's.
It is pinned against central finite differences of ``oracle.coulomb.local_energy`` and
``oracle.wf.ln_wavefunction`` (tests/test_oracle_forces.py).
"""

from __future__ import annotations

import numpy as np

from jqmc_amd.atomic_orbital import AOs_cart_data, AOs_sphe_data
from jqmc_amd.coulomb_potential import get_mesh
from jqmc_amd.molecular_orbital import MOs_data

from . import sphe
from . import wf as owf
from .coulomb import _nonlocal_table, legendre_tabulated

# --------------------------------------------------------------------------- AO derivatives


def _poly1d(n, Z, x):
    """x^n e^{-Z x^2} and its first three derivatives (without the exponential factor), for
    per-row orders n and exponents Z (broadcast against x)."""

    def pw(k):
        return np.where(k >= 0, x ** np.maximum(k, 0), 0.0)

    d0 = pw(n)
    d1 = n * pw(n - 1) - 2.0 * Z * pw(n + 1)
    d2 = n * (n - 1) * pw(n - 2) - 2.0 * Z * (2 * n + 1) * pw(n) + 4.0 * Z * Z * pw(n + 2)
    d3 = (n * (n - 1) * (n - 2) * pw(n - 3) - 6.0 * Z * n * n * pw(n - 1) + 12.0 * Z * Z * (n + 1) * pw(n + 1)
          - 8.0 * Z**3 * pw(n + 3))
    return d0, d1, d2, d3


def _ao_terms(aos, a, coef=None):
    """(primitive (exponent, coefficient) list, monomial ((nx, ny, nz), coefficient) list) of AO a."""
    prims = [p for p, v in enumerate(aos.orbital_indices) if v == a]
    if isinstance(aos, AOs_cart_data):
        coef = aos._normalized_coefficients_np if coef is None else coef
        pr = [(aos.exponents[p], coef[p]) for p in prims]
        mons = [((aos.polynominal_order_x[a], aos.polynominal_order_y[a], aos.polynominal_order_z[a]), 1.0)]
        return pr, mons
    if isinstance(aos, AOs_sphe_data):
        l = aos.angular_momentums[a]
        pr = [(aos.exponents[p], sphe._sph_norm(l, aos.exponents[p]) * np.sqrt((2 * l + 1) / (4.0 * np.pi))
               * aos.coefficients[p]) for p in prims]
        m, c = sphe.solid_harmonic_monomials(l, aos.magnetic_quantum_numbers[a])
        return pr, [(mm, cc) for mm, cc in zip(m, c) if cc != 0.0]
    raise NotImplementedError


_TERMS = {}


def _term_table(aos):
    """Flat table of every (AO, primitive, monomial) product term of the basis: AO index, exponent,
    coefficient (primitive x monomial) and the three monomial orders. Cached per basis object."""
    hit = _TERMS.get(id(aos))
    if hit is not None and hit[0] is aos:
        return hit[1]
    coef = aos._normalized_coefficients_np if isinstance(aos, AOs_cart_data) else None
    rows = []
    for a in range(aos.num_ao):
        prims, mons = _ao_terms(aos, a, coef)
        for Z, c in prims:
            for (nx, ny, nz), k in mons:
                rows.append((a, Z, c * k, nx, ny, nz))
    t = np.array(rows, dtype=np.float64)
    tab = (t[:, 0].astype(np.int64), t[:, 1], t[:, 2], t[:, 3:6].astype(np.int64))
    _TERMS[id(aos)] = (aos, tab)
    return tab


def aos_d3(aos, r_carts):
    """Value (n_ao, n), gradient (3, n_ao, n), Hessian (3, 3, n_ao, n) and gradient of the
    Laplacian (3, n_ao, n) of every AO, from the per-direction factorisation
    x^nx e^{-Z x^2} y^ny e^{-Z y^2} z^nz e^{-Z z^2} of each primitive-monomial product."""
    r = np.asarray(r_carts, dtype=np.float64).reshape(-1, 3)
    n_ao = aos.num_ao
    ao, Z, c, nn = _term_table(aos)
    d = r[None, :, :] - np.asarray(aos._atomic_center_carts_np)[ao][:, None, :]  # (T, n, 3)
    e = c[:, None] * np.exp(-Z[:, None] * np.sum(d * d, axis=2))  # (T, n)
    D = [np.stack(_poly1d(nn[:, i, None], Z[:, None], d[:, :, i])) for i in range(3)]  # 3 x (4, T, n)

    def prod(o):
        return e * D[0][o[0]] * D[1][o[1]] * D[2][o[2]]

    def acc(x):
        out = np.zeros((n_ao, x.shape[1]))
        np.add.at(out, ao, x)
        return out

    orders = np.eye(3, dtype=int)
    v = acc(prod((0, 0, 0)))
    g = np.stack([acc(prod(orders[i])) for i in range(3)])
    H = np.stack([np.stack([acc(prod(orders[i] + orders[j])) for j in range(3)]) for i in range(3)])
    gl = np.stack([acc(sum(prod(orders[i] + 2 * orders[j]) for j in range(3))) for i in range(3)])
    return v, g, H, gl


def _orb_basis(orb):
    """(AO basis, coefficient matrix C) with orbitals = C @ AOs (identity for an AO set)."""
    if isinstance(orb, MOs_data):
        return orb.aos_data, np.asarray(orb.mo_coefficients, dtype=np.float64)
    return orb, np.eye(orb.num_ao)


# --------------------------------------------------------------------------- radial pair functions


def _radial_derivs(r, f1, f2, f3, u, dJ, lap_coef):
    """Adjoint of (dJ . grad t + lap_coef * lap t) for t(u) = f(|u|) w.r.t. u."""
    uh = u / r
    hess_dJ = f2 * np.dot(uh, dJ) * uh + (f1 / r) * (dJ - np.dot(uh, dJ) * uh)
    grad_lap = (f3 + 2.0 * f2 / r - 2.0 * f1 / (r * r)) * uh
    return hess_dJ + lap_coef * grad_lap


def _j1_f(j1, Zeff, r):
    """f, f', f'', f''' of the one-body term for one nucleus (jastrow_factor.py:672-711)."""
    a = j1.jastrow_1b_param
    c = (2.0 * Zeff) ** 0.25
    A = (2.0 * Zeff) ** 0.75
    if j1.jastrow_1b_type == "exp":
        e = np.exp(-a * c * r)
        return (-A / (2.0 * a) * (1.0 - e), -A * c / 2.0 * e, A * a * c * c / 2.0 * e, -A * a * a * c**3 / 2.0 * e)
    den = 1.0 + a * c * r
    return (-A * r / (2.0 * den), -A / (2.0 * den**2), A * a * c / den**3, -3.0 * A * a * a * c * c / den**4)


def _j2_f(j2, r):
    """f, f', f'', f''' of the two-body pair function (jastrow_factor.py:1207-1230)."""
    a = j2.jastrow_2b_param
    if j2.jastrow_2b_type == "pade":
        den = 1.0 + a * r
        return r / (2.0 * den), 0.5 / den**2, -a / den**3, 3.0 * a * a / den**4
    e = np.exp(-a * r)
    return (1.0 - e) / (2.0 * a), 0.5 * e, -a / 2.0 * e, a * a / 2.0 * e


# --------------------------------------------------------------------------- SWCT


def swct_omega(R, r):
    """``evaluate_swct_omega`` (swct.py:63-101): omega (n_nuc, n_e)."""
    d = np.linalg.norm(np.asarray(r)[None, :, :] - np.asarray(R)[:, None, :], axis=-1)
    k = d**-4.0
    return k / np.sum(k, axis=0, keepdims=True)


def swct_domega(R, r):
    """``evaluate_swct_domega`` (swct.py:132-150): sum_i grad_{r_i} omega_ai, (n_nuc, 3)."""
    R = np.asarray(R)
    r = np.asarray(r)
    u = r[None, :, :] - R[:, None, :]  # (n_nuc, n_e, 3)
    d2 = np.sum(u * u, axis=-1)
    k = d2**-2.0
    dk = -4.0 * (k / d2)[..., None] * u  # grad_r k_ai
    S = np.sum(k, axis=0)  # (n_e)
    dS = np.sum(dk, axis=0)  # (n_e, 3)
    dom = dk / S[None, :, None] - (k / S**2)[..., None] * dS[None, :, :]
    return np.sum(dom, axis=1)


# --------------------------------------------------------------------------- the adjoint


def _atom_of(aos):
    return np.asarray(aos.nucleus_index, dtype=np.int64)


def local_energy_position_grads(ham, r_up, r_dn, RT, NN: int = 1, Nv: int = 6) -> dict:
    """E_L and its gradients with respect to electron and nuclear positions, and the gradients of
    ln|Psi|: {'e_L', 'dE_dr' (N, 3), 'dE_dR' (n_nuc, 3), 'dlnpsi_dr' (N, 3), 'dlnpsi_dR' (n_nuc, 3)}.
    Electrons are ordered up block then dn block."""
    wf = ham.wavefunction_data
    cp = ham.coulomb_potential_data
    gem = wf.geminal_data
    jas = wf.jastrow_data
    r_up = np.reshape(np.asarray(r_up, dtype=np.float64), (-1, 3))
    r_dn = np.reshape(np.asarray(r_dn, dtype=np.float64), (-1, 3))
    nu, nd = len(r_up), len(r_dn)
    N = nu + nd
    X = np.vstack([r_up, r_dn])
    Rn = np.asarray(cp.structure_data.positions, dtype=np.float64)
    n_nuc = len(Rn)
    rbar = np.zeros((N, 3))
    Rbar = np.zeros((n_nuc, 3))
    lnR = np.zeros((n_nuc, 3))  # d ln|Psi| / dR

    # ---- determinant part (forward)
    aos_u, Cu = _orb_basis(gem.orb_data_up_spin)
    aos_d, Cd = _orb_basis(gem.orb_data_dn_spin)
    Au = aos_d3(aos_u, r_up)
    Ad = aos_d3(aos_d, r_dn)
    Lp, Lu = np.asarray(gem.lambda_paired), np.asarray(gem.lambda_unpaired)
    phu, phd = Cu @ Au[0], Cd @ Ad[0]  # (n_orb, n)
    gphu = np.stack([Cu @ Au[1][c] for c in range(3)])
    gphd = np.stack([Cd @ Ad[1][c] for c in range(3)])
    lphu = Cu @ (Au[2][0, 0] + Au[2][1, 1] + Au[2][2, 2])
    lphd = Cd @ (Ad[2][0, 0] + Ad[2][1, 1] + Ad[2][2, 2])
    G = np.hstack([phu.T @ Lp @ phd, phu.T @ Lu])
    Gi = np.linalg.inv(G)
    Qu = Lp @ phd @ Gi[:nd, :] + Lu @ Gi[nd:, :]  # (n_orb_u, nu): q_i of each up electron
    Qd = Lp.T @ phu @ Gi[:nd, :].T  # (n_orb_d, nd)
    D1 = np.vstack([np.einsum("coi,oi->ic", gphu, Qu), np.einsum("coj,oj->jc", gphd, Qd)])
    Lap = np.concatenate([np.sum(lphu * Qu, axis=0), np.sum(lphd * Qd, axis=0)])

    # ---- Jastrow (forward)
    gJu, gJd, lJu, lJd = owf.grads_laplacian_jastrow(jas, r_up, r_dn)
    gJ = np.vstack([gJu, gJd])
    lJ = np.concatenate([lJu, lJd])
    j1 = jas.jastrow_one_body_data
    j2 = jas.jastrow_two_body_data
    j3 = jas.jastrow_three_body_data
    if j3 is not None:
        aos_j, Cj = _orb_basis(j3.orb_data)
        Bj = aos_d3(aos_j, X)
        beta = Cj @ Bj[0]  # (nJ, N)
        gbeta = np.stack([Cj @ Bj[1][c] for c in range(3)])
        lbeta = Cj @ (Bj[2][0, 0] + Bj[2][1, 1] + Bj[2][2, 2])
        Jm = np.asarray(j3.j_matrix)[:, :-1]
        jv = np.asarray(j3.j_matrix)[:, -1]
        gfield = np.zeros_like(beta)
        for e in range(N):
            gfield[:, e] = jv
            for f in range(N):
                if f == e:
                    continue
                # within a spin block pairs f > e use J beta_f, f < e use J^T beta_f; across the blocks
                # up-dn pairs are beta_up^T J beta_dn (jastrow_factor.py:1744-1801)
                gfield[:, e] += (Jm @ beta[:, f]) if f > e else (Jm.T @ beta[:, f])
    Zeff = np.asarray(cp._effective_charges, dtype=np.float64)
    Zj1 = np.asarray(j1.effective_charges, dtype=np.float64) if j1 is not None else None

    # ---- kinetic energy
    T = -0.5 * np.sum(Lap + 2.0 * np.sum(D1 * gJ, axis=1) + lJ + np.sum(gJ * gJ, axis=1))
    dJ = -(D1 + gJ)  # adjoint of gJ_e
    Qbar_u = np.zeros_like(Qu)
    Qbar_d = np.zeros_like(Qd)
    phbar_u = [np.zeros_like(phu), np.zeros_like(gphu), np.zeros_like(lphu)]
    phbar_d = [np.zeros_like(phd), np.zeros_like(gphd), np.zeros_like(lphd)]
    for i in range(nu):
        phbar_u[1][:, :, i] += -gJ[i][:, None] * Qu[:, i][None, :]
        Qbar_u[:, i] += -np.einsum("c,co->o", gJ[i], gphu[:, :, i]) - 0.5 * lphu[:, i]
        phbar_u[2][:, i] += -0.5 * Qu[:, i]
    for j in range(nd):
        e = nu + j
        phbar_d[1][:, :, j] += -gJ[e][:, None] * Qd[:, j][None, :]
        Qbar_d[:, j] += -np.einsum("c,co->o", gJ[e], gphd[:, :, j]) - 0.5 * lphd[:, j]
        phbar_d[2][:, j] += -0.5 * Qd[:, j]
    if j3 is not None:
        bbar = [np.zeros_like(beta), np.zeros_like(gbeta), np.zeros_like(lbeta)]
        gbar = np.zeros_like(gfield)
        for e in range(N):
            bbar[1][:, :, e] += dJ[e][:, None] * gfield[:, e][None, :]
            gbar[:, e] += np.einsum("c,co->o", dJ[e], gbeta[:, :, e]) - 0.5 * lbeta[:, e]
            bbar[2][:, e] += -0.5 * gfield[:, e]
    if j1 is not None:
        for e in range(N):
            for a in range(n_nuc):
                u = X[e] - Rn[a]
                r = np.linalg.norm(u)
                _, f1, f2, f3 = _j1_f(j1, Zj1[a], r)
                v = _radial_derivs(r, f1, f2, f3, u, dJ[e], -0.5)
                rbar[e] += v
                Rbar[a] -= v
                lnR[a] -= f1 * u / r
    if j2 is not None:
        for e in range(N):
            for f in range(N):
                if f == e:
                    continue
                u = X[e] - X[f]
                r = np.linalg.norm(u)
                _, f1, f2, f3 = _j2_f(j2, r)
                v = _radial_derivs(r, f1, f2, f3, u, dJ[e], -0.5)
                rbar[e] += v
                rbar[f] -= v

    # ---- bare Coulomb and ECP local
    for a in range(n_nuc):
        for b in range(a + 1, n_nuc):
            u = Rn[a] - Rn[b]
            v = -Zeff[a] * Zeff[b] * u / np.linalg.norm(u) ** 3
            Rbar[a] += v
            Rbar[b] -= v
    for e in range(N):
        for a in range(n_nuc):
            u = X[e] - Rn[a]
            v = Zeff[a] * u / np.linalg.norm(u) ** 3
            rbar[e] += v
            Rbar[a] -= v
        for f in range(e + 1, N):
            u = X[e] - X[f]
            v = -u / np.linalg.norm(u) ** 3
            rbar[e] += v
            rbar[f] -= v
    V_loc = 0.0
    if cp.ecp_flag:
        for a, ex, c, p in cp._local_terms():
            for e in range(N):
                u = X[e] - Rn[a]
                d = np.linalg.norm(u)
                ee = np.exp(-ex * d * d)
                V_loc += c * d ** (p - 2.0) * ee
                fp = c * ((p - 2.0) * d ** (p - 3.0) - 2.0 * ex * d ** (p - 1.0)) * ee
                rbar[e] += fp * u / d
                Rbar[a] -= fp * u / d

    # ---- ECP non-local
    V_nl = 0.0
    if cp.ecp_flag:
        mesh = get_mesh(Nv)
        gp = mesh.grid_points @ RT
        gun = gp / np.linalg.norm(gp, axis=1, keepdims=True)
        tab = _nonlocal_table(cp)
        for e in range(N):
            up = e < nu
            k_e = e if up else e - nu
            d_all = np.linalg.norm(Rn - X[e], axis=1)
            for a in np.argsort(d_all, kind="stable")[:NN]:
                a = int(a)
                terms = tab.get(a, [])
                if not terms:
                    continue
                u = X[e] - Rn[a]
                d = np.linalg.norm(u)
                uh = u / d
                aos_s, C_s = (aos_u, Cu) if up else (aos_d, Cd)
                q = Qu[:, k_e] if up else Qd[:, k_e]
                for k in range(Nv):
                    cos_t = float(uh @ gun[k])
                    W, dW = 0.0, np.zeros(3)
                    for l, ex, c, pw in terms:
                        ee = np.exp(-ex * d * d)
                        vl = c * d ** (pw - 2.0) * ee
                        vlp = c * ((pw - 2.0) * d ** (pw - 3.0) - 2.0 * ex * d ** (pw - 1.0)) * ee
                        Pl = float(legendre_tabulated(l, cos_t))
                        dPl = _legendre_deriv(l, cos_t)
                        W += vl * (2 * l + 1) * Pl * mesh.weights[k]
                        dW += (2 * l + 1) * mesh.weights[k] * (vlp * Pl * uh + vl * dPl * (gun[k] - cos_t * uh) / d)
                    if W == 0.0:
                        continue
                    p = Rn[a] + d * gp[k]
                    Ap = aos_d3(aos_s, p[None, :])
                    php = C_s @ Ap[0][:, 0]
                    gphp = np.stack([C_s @ Ap[1][c][:, 0] for c in range(3)])
                    D = float(php @ q)
                    dJv = 0.0
                    pbar_J = np.zeros(3)  # d(Delta J)/dp
                    if j1 is not None:
                        for b in range(n_nuc):
                            up_ = p - Rn[b]
                            ue = X[e] - Rn[b]
                            rp_, re_ = np.linalg.norm(up_), np.linalg.norm(ue)
                            f0p, f1p, _, _ = _j1_f(j1, Zj1[b], rp_)
                            f0e, f1e, _, _ = _j1_f(j1, Zj1[b], re_)
                            dJv += f0p - f0e
                    if j2 is not None:
                        for f in range(N):
                            if f == e:
                                continue
                            f0p = _j2_f(j2, np.linalg.norm(p - X[f]))[0]
                            f0e = _j2_f(j2, np.linalg.norm(X[e] - X[f]))[0]
                            dJv += f0p - f0e
                    if j3 is not None:
                        Bp = aos_d3(aos_j, p[None, :])
                        bp = Cj @ Bp[0][:, 0]
                        gbp = np.stack([Cj @ Bp[1][c][:, 0] for c in range(3)])
                        dJv += float((bp - beta[:, e]) @ gfield[:, e])
                    E = np.exp(dJv)
                    rho = D * E
                    V_nl += W * rho
                    s = W * rho
                    # geometry of the projector weight
                    rbar[e] += rho * dW
                    Rbar[a] -= rho * dW
                    # determinant ratio phi(p) . q_e
                    pbar = W * E * (gphp @ q)
                    cdet = C_s.T @ (W * E * q)
                    Rbar_p = np.zeros((n_nuc, 3))
                    np.add.at(Rbar_p, _atom_of(aos_s), -cdet[:, None] * Ap[1][:, :, 0].T)
                    if up:
                        Qbar_u[:, k_e] += W * E * php
                    else:
                        Qbar_d[:, k_e] += W * E * php
                    # Jastrow difference
                    if j1 is not None:
                        for b in range(n_nuc):
                            up_ = p - Rn[b]
                            ue = X[e] - Rn[b]
                            rp_, re_ = np.linalg.norm(up_), np.linalg.norm(ue)
                            f1p = _j1_f(j1, Zj1[b], rp_)[1]
                            f1e = _j1_f(j1, Zj1[b], re_)[1]
                            pbar += s * f1p * up_ / rp_
                            Rbar_p[b] -= s * f1p * up_ / rp_
                            rbar[e] -= s * f1e * ue / re_
                            Rbar[b] += s * f1e * ue / re_
                    if j2 is not None:
                        for f in range(N):
                            if f == e:
                                continue
                            up_ = p - X[f]
                            ue = X[e] - X[f]
                            rp_, re_ = np.linalg.norm(up_), np.linalg.norm(ue)
                            f1p = _j2_f(j2, rp_)[1]
                            f1e = _j2_f(j2, re_)[1]
                            pbar += s * f1p * up_ / rp_
                            rbar[f] -= s * f1p * up_ / rp_
                            rbar[e] -= s * f1e * ue / re_
                            rbar[f] += s * f1e * ue / re_
                    if j3 is not None:
                        pbar += s * (gbp @ gfield[:, e])
                        cj = Cj.T @ (s * gfield[:, e])
                        np.add.at(Rbar_p, _atom_of(aos_j), -cj[:, None] * Bp[1][:, :, 0].T)
                        bbar[0][:, e] -= s * gfield[:, e]
                        gbar[:, e] += s * (bp - beta[:, e])
                    # p = R_a + |r_e - R_a| gp_k
                    Rbar += Rbar_p
                    rbar[e] += uh * float(gp[k] @ pbar)
                    Rbar[a] += pbar - uh * float(gp[k] @ pbar)

    # ---- back through q (G^-1 and the other spin's orbitals)
    Gibar = np.zeros_like(Gi)
    Gibar[:nd, :] += phd.T @ Lp.T @ Qbar_u
    Gibar[nd:, :] += Lu.T @ Qbar_u
    phbar_d[0] += Lp.T @ Qbar_u @ Gi[:nd, :].T
    Gibar[:nd, :] += (phu.T @ Lp @ Qbar_d).T
    phbar_u[0] += Lp @ Qbar_d @ Gi[:nd, :]
    Gbar = -Gi.T @ Gibar @ Gi.T
    phbar_u[0] += Lp @ phd @ Gbar[:, :nd].T + Lu @ Gbar[:, nd:].T
    phbar_d[0] += Lp.T @ phu @ Gbar[:, :nd]

    # ---- back through the J3 field
    if j3 is not None:
        for f in range(N):
            for e in range(N):
                if e == f:
                    continue
                # g_e holds J beta_f for f > e and J^T beta_f for f < e
                bbar[0][:, f] += (Jm.T @ gbar[:, e]) if f > e else (Jm @ gbar[:, e])

    # ---- AO adjoints to positions and centres
    def to_positions(Aset, aos, C, bars, idx):
        v, g, H, gl = Aset
        av = C.T @ bars[0]
        ag = np.stack([C.T @ bars[1][c] for c in range(3)])
        al = C.T @ bars[2]
        atom = _atom_of(aos)
        for col, e in enumerate(idx):
            contrib = (av[:, col][None, :] * g[:, :, col] + np.einsum("ca,cda->da", ag[:, :, col], H[:, :, :, col])
                       + al[:, col][None, :] * gl[:, :, col])  # (3, n_ao)
            rbar[e] += np.sum(contrib, axis=1)
            np.add.at(Rbar, atom, -contrib.T)

    to_positions(Au, aos_u, Cu, phbar_u, range(nu))
    to_positions(Ad, aos_d, Cd, phbar_d, range(nu, N))
    if j3 is not None:
        to_positions(Bj, aos_j, Cj, bbar, range(N))

    # ---- d ln|Psi| / dR through the AO centres
    hu = Cu.T @ Qu
    hd = Cd.T @ Qd
    np.add.at(lnR, _atom_of(aos_u), -np.einsum("an,can->ac", hu, Au[1]))
    np.add.at(lnR, _atom_of(aos_d), -np.einsum("an,can->ac", hd, Ad[1]))
    if j3 is not None:
        hj = Cj.T @ gfield
        np.add.at(lnR, _atom_of(aos_j), -np.einsum("an,can->ac", hj, Bj[1]))

    V_bare = _bare(Rn, Zeff, X)
    e_L = T + V_bare + V_loc + V_nl
    return {"e_L": float(e_L), "dE_dr": rbar, "dE_dR": Rbar, "dlnpsi_dr": D1 + gJ, "dlnpsi_dR": lnR}


def _bare(Rn, Z, X) -> float:
    v = 0.0
    for a in range(len(Rn)):
        for b in range(a + 1, len(Rn)):
            v += Z[a] * Z[b] / np.linalg.norm(Rn[a] - Rn[b])
        v -= np.sum(Z[a] / np.linalg.norm(X - Rn[a], axis=1))
    iu, ju = np.triu_indices(len(X), k=1)
    return float(v + np.sum(1.0 / np.linalg.norm(X[iu] - X[ju], axis=1)))


def _legendre_deriv(n: int, x: float) -> float:
    """d P_n / dx for the tabulated P_0..P_6."""
    table = [
        lambda x: 0.0,
        lambda x: 1.0,
        lambda x: 3.0 * x,
        lambda x: 0.5 * (15.0 * x * x - 3.0),
        lambda x: 0.125 * (140.0 * x**3 - 60.0 * x),
        lambda x: 0.125 * (315.0 * x**4 - 210.0 * x * x + 15.0),
        lambda x: (1386.0 * x**5 - 1260.0 * x**3 + 210.0 * x) / 16.0,
    ]
    return float(table[n](x))


def force_terms(ham, r_up, r_dn, RT, NN: int = 1, Nv: int = 6, swct: bool = True) -> dict:
    """Per-walker force products of jqmc_mcmc.py:796-852: {'e_L', 'force_HF' (n_nuc, 3),
    'force_PP' (n_nuc, 3), 'E_L_force_PP' (n_nuc, 3)}."""
    out = local_energy_position_grads(ham, r_up, r_dn, RT, NN, Nv)
    R = ham.coulomb_potential_data.structure_data.positions
    X = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    nu = np.reshape(r_up, (-1, 3)).shape[0]
    if swct:
        om = np.hstack([swct_omega(R, X[:nu]), swct_omega(R, X[nu:])])
        dom = swct_domega(R, X[:nu]) + swct_domega(R, X[nu:])
    else:
        om = np.zeros((len(R), len(X)))
        dom = np.zeros((len(R), 3))
    F_HF = out["dE_dR"] + om @ out["dE_dr"]
    F_PP = out["dlnpsi_dR"] + om @ out["dlnpsi_dr"] + 0.5 * dom
    return {"e_L": out["e_L"], "force_HF": F_HF, "force_PP": F_PP, "E_L_force_PP": out["e_L"] * F_PP}
