"""FP64 restatement of the per-walker log-derivatives O_k = d ln|Psi| / d p_k for the
variational blocks the reference optimises (test infrastructure only).

Blocks, in the order of ``Wavefunction_data.get_variational_blocks`` (wavefunction.py:515-640):
``j1_param`` (scalar), ``j2_param`` (scalar), ``j3_matrix`` (n_ao x (n_ao + 1)), ``lambda_matrix``
(n_orb x (n_orb + n_up - n_dn)).  The reference obtains them with
``grad(evaluate_ln_wavefunction_fast, argnums=0)`` (jqmc_mcmc.py:854-873); here they are the
closed forms:

* d lnPsi / d J1 b = sum_{i,n} d/db j1(b; r_in);  d / d J2 b = sum_{i<j} d/db u(b; r_ij)
  (jastrow_factor.py:672-711, 1207-1230);
* d / d J[a, b] = sum_{e<f} a_e[a] a_f[b] over the electron order up block then down block,
  d / d j[a] (last column) = sum_e a_e[a] (jastrow_factor.py:1744-1801);
* d ln|det G| / d Lambda_p = Phi_up G^-1[:n_dn, :]^T Phi_dn^T and
  d / d Lambda_u = Phi_up G^-1[n_dn:, :]^T, from G = Phi_up^T [Lambda_p Phi_dn | Lambda_u]
  (determinant.py:1379-1402).

Pinned against central finite differences of ``wf.ln_wavefunction`` in tests/test_oracle_derivs.py.
"""

from __future__ import annotations

import numpy as np

from . import wf as owf


def d_j1_param(j1, r_up, r_dn) -> float:
    R = j1.structure_data.positions
    Z = np.asarray(j1.effective_charges, dtype=float)
    b = j1.jastrow_1b_param
    c = (2.0 * Z) ** 0.25
    A = (2.0 * Z) ** 0.75
    r = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    d = np.linalg.norm(r[:, None, :] - R[None, :, :], axis=-1)
    cd = c[None, :] * d
    if j1.jastrow_1b_type == "exp":
        e = np.exp(-b * cd)
        dfdb = -A[None, :] * (-(1.0 - e) / (2.0 * b * b) + cd * e / (2.0 * b))
    else:
        dfdb = A[None, :] * d * cd / (2.0 * (1.0 + b * cd) ** 2)
    return float(np.sum(dfdb))


def d_j2_param(j2, r_up, r_dn) -> float:
    b = j2.jastrow_2b_param
    r = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    iu, ju = np.triu_indices(len(r), k=1)
    x = np.linalg.norm(r[iu] - r[ju], axis=-1)
    if j2.jastrow_2b_type == "pade":
        return float(np.sum(-x * x / (2.0 * (1.0 + b * x) ** 2)))
    e = np.exp(-b * x)
    return float(np.sum(-(1.0 - e) / (2.0 * b * b) + x * e / (2.0 * b)))


def d_j3_matrix(j3, r_up, r_dn) -> np.ndarray:
    A = np.hstack([owf.orb_value(j3.orb_data, np.reshape(r_up, (-1, 3))),
                   owf.orb_value(j3.orb_data, np.reshape(r_dn, (-1, 3)))])  # (n_ao, N)
    n = A.shape[1]
    out = np.zeros((A.shape[0], A.shape[0] + 1))
    for e in range(n):
        later = A[:, e + 1:].sum(axis=1)
        out[:, :-1] += np.outer(A[:, e], later)
    out[:, -1] = A.sum(axis=1)
    return out


def d_lambda(gem, r_up, r_dn) -> np.ndarray:
    Pu = owf.orb_value(gem.orb_data_up_spin, np.reshape(r_up, (-1, 3)))  # (n_orb, n_up)
    Pd = owf.orb_value(gem.orb_data_dn_spin, np.reshape(r_dn, (-1, 3)))  # (n_orb, n_dn)
    G = owf.geminal(gem, r_up, r_dn)
    Gi = np.linalg.inv(G)
    nd = Pd.shape[1]
    dp = Pu @ Gi[:nd, :].T @ Pd.T
    du = Pu @ Gi[nd:, :].T
    return np.hstack([dp, du])


def dln_psi_dparams(wf, r_up, r_dn) -> dict:
    """{'j1_param', 'j2_param', 'j3_matrix', 'lambda_matrix'} for the blocks present."""
    jas = wf.jastrow_data
    out = {}
    if jas.jastrow_one_body_data is not None:
        out["j1_param"] = d_j1_param(jas.jastrow_one_body_data, r_up, r_dn)
    if jas.jastrow_two_body_data is not None:
        out["j2_param"] = d_j2_param(jas.jastrow_two_body_data, r_up, r_dn)
    if jas.jastrow_three_body_data is not None:
        out["j3_matrix"] = d_j3_matrix(jas.jastrow_three_body_data, r_up, r_dn)
    out["lambda_matrix"] = d_lambda(wf.geminal_data, r_up, r_dn)
    return out
