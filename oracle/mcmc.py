"""FP64 restatement of the VMC Metropolis kernel (test infrastructure only).

``metropolis_walker`` follows ``_update_electron_positions`` (jqmc_mcmc.py:4278-4533) move by
move: electron choice, nearest-nucleus step size f_l (4339-4361), single-axis Gaussian move,
asymmetric proposal ratio T_ratio (4398-4401), Jastrow ratio, Sherman-Morrison determinant
ratio from recomputed geminal rows/columns (4412-4482), Attaccalite-Sorella factors
(4484-4491) and ``b < min(1, R^2 T)`` acceptance (4493-4520).  Random draws come from the
shared Philox stream (``oracle.rng.vmc_draws``) instead of the JAX key stream.

``metropolis_walker_debug`` is the brute-force twin (the reference's ``_MCMC_debug`` strategy,
jqmc_mcmc.py:4968-5115): full wavefunction values and a fresh inverse at every move.
"""

from __future__ import annotations

import numpy as np

from jqmc_amd.structure import _find_nearest_index_np

from . import rng as orng
from . import wf as owf


def proposal_charges(ham) -> np.ndarray:
    """Z (all-electron) or Z - z_core (ECP) used by the step-size rule (jqmc_mcmc.py:4342-4350)."""
    cp = ham.coulomb_potential_data
    return cp._effective_charges


def _f_l(Z, d):
    return 1.0 / Z**2 * (1.0 + Z**2 * d) / (1.0 + d)


def propose(ham, r_up, r_dn, draws, Dt):
    """One proposal: returns (is_up, index, old, new, T_ratio)."""
    rand_num, up_index, dn_index, normal, axis, _ = draws
    n_up = len(r_up)
    is_up = bool(rand_num < n_up)
    idx = int(up_index if is_up else dn_index)
    old = (r_up if is_up else r_dn)[idx].copy()
    st = ham.structure_data
    Zs = proposal_charges(ham)
    a = _find_nearest_index_np(st, old)
    f_l = _f_l(Zs[a], np.linalg.norm(old - st.positions[a]))
    g = normal * (f_l * Dt)
    new = old.copy()
    new[int(axis)] += g
    a2 = _find_nearest_index_np(st, new)
    f_p = _f_l(Zs[a2], np.linalg.norm(new - st.positions[a2]))
    T_ratio = (f_l / f_p) * np.exp(
        -(np.linalg.norm(new - old) ** 2) * (1.0 / (2.0 * f_p**2 * Dt**2) - 1.0 / (2.0 * f_l**2 * Dt**2))
    )
    return is_up, idx, old, new, T_ratio


def metropolis_walker(ham, r_up, r_dn, Ginv, G, seed, walker, move0, n_moves, Dt, epsilon_AS, record=None):
    """Run ``n_moves`` proposals for one walker; returns (acc, rej, r_up, r_dn, Ginv, G)."""
    wf = ham.wavefunction_data
    gem = wf.geminal_data
    r_up = np.array(r_up, dtype=np.float64)
    r_dn = np.array(r_dn, dtype=np.float64)
    Ginv = np.array(Ginv, dtype=np.float64)
    G = np.array(G, dtype=np.float64)
    n_up, n_dn = len(r_up), len(r_dn)
    acc = rej = 0
    for m in range(n_moves):
        draws = orng.vmc_draws(seed, walker, move0 + m, n_up, n_dn)
        is_up, idx, old, new, T_ratio = propose(ham, r_up, r_dn, draws, Dt)
        p_up, p_dn = r_up.copy(), r_dn.copy()
        if is_up:
            p_up[idx] = new
        else:
            p_dn[idx] = new
        J_ratio = np.exp(owf.jastrow_part(wf.jastrow_data, p_up, p_dn) - owf.jastrow_part(wf.jastrow_data, r_up, r_dn))
        if is_up:
            v = owf.geminal_up_one_row(gem, new, r_dn) - owf.geminal_up_one_row(gem, old, r_dn)
            u = np.zeros(n_up)
            u[idx] = 1.0
        else:
            v = np.zeros(n_up)
            v[idx] = 1.0
            u = owf.geminal_dn_one_column(gem, r_up, new) - owf.geminal_dn_one_column(gem, r_up, old)
        Ainv_u = Ginv @ u
        vT_Ainv = v @ Ginv
        det_ratio = 1.0 + v @ Ainv_u
        Ginv_new = Ginv - np.outer(Ainv_u, vT_Ainv) / det_ratio
        G_new = G.copy()
        if is_up:
            G_new[idx, :] += v
        else:
            G_new[:, idx] += u
        R_p = owf.as_regularization_fast(G_new, Ginv_new)
        R_o = owf.as_regularization_fast(G, Ginv)
        R_AS_ratio = (max(R_p, epsilon_AS) / R_p) / (max(R_o, epsilon_AS) / R_o)
        R_ratio = (R_AS_ratio * J_ratio * det_ratio) ** 2.0
        acceptance = min(1.0, R_ratio * T_ratio)
        b = float(draws[5])
        accepted = b < acceptance
        if record is not None:
            record.append((bool(is_up), idx, bool(accepted), float(det_ratio), float(J_ratio)))
        if accepted:
            acc += 1
            r_up, r_dn, Ginv, G = p_up, p_dn, Ginv_new, G_new
        else:
            rej += 1
    return acc, rej, r_up, r_dn, Ginv, G


def metropolis_walker_debug(ham, r_up, r_dn, seed, walker, move0, n_moves, Dt, epsilon_AS, record=None):
    """Brute-force twin: full determinants and SVD inverses at every move (jqmc_mcmc.py:4968-5115)."""
    wf = ham.wavefunction_data
    gem = wf.geminal_data
    r_up = np.array(r_up, dtype=np.float64)
    r_dn = np.array(r_dn, dtype=np.float64)
    n_up, n_dn = len(r_up), len(r_dn)
    acc = rej = 0
    for m in range(n_moves):
        draws = orng.vmc_draws(seed, walker, move0 + m, n_up, n_dn)
        is_up, idx, old, new, T_ratio = propose(ham, r_up, r_dn, draws, Dt)
        p_up, p_dn = r_up.copy(), r_dn.copy()
        if is_up:
            p_up[idx] = new
        else:
            p_dn[idx] = new
        G_o = owf.geminal(gem, r_up, r_dn)
        G_p = owf.geminal(gem, p_up, p_dn)
        det_ratio = np.linalg.det(G_p) / np.linalg.det(G_o)
        J_ratio = np.exp(owf.jastrow_part(wf.jastrow_data, p_up, p_dn) - owf.jastrow_part(wf.jastrow_data, r_up, r_dn))
        R_p = owf.as_regularization_fast(G_p, owf.geminal_inverse(G_p))
        R_o = owf.as_regularization_fast(G_o, owf.geminal_inverse(G_o))
        R_AS_ratio = (max(R_p, epsilon_AS) / R_p) / (max(R_o, epsilon_AS) / R_o)
        acceptance = min(1.0, (R_AS_ratio * J_ratio * det_ratio) ** 2.0 * T_ratio)
        accepted = float(draws[5]) < acceptance
        if record is not None:
            record.append((bool(is_up), idx, bool(accepted), float(det_ratio), float(J_ratio)))
        if accepted:
            acc += 1
            r_up, r_dn = p_up, p_dn
        else:
            rej += 1
    return acc, rej, r_up, r_dn


def init_walker(ham, r_up, r_dn):
    """G and its pseudo-inverse at the start of a run (jqmc_mcmc.py:658-662 -> 4248-4261)."""
    G = owf.geminal(ham.wavefunction_data.geminal_data, r_up, r_dn)
    return G, owf.geminal_inverse(G)


def rotation_RT(seed, walker, step):
    """Measurement rotation ``R.T`` for (walker, step) from the rotation stream."""
    a, b, c = orng.rotation_angles(seed, walker, step)
    return orng.rotation_matrix_T(float(a), float(b), float(c))
