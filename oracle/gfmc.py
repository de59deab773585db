"""FP64 restatement of the LRDMC projection step (GFMC_n and GFMC_t) and of the comb
branching.  Test infrastructure only: used by tests/, smoke() and the CPU baseline.

Pins (tests/test_oracle_lrdmc.py):
  * lattice ratios vs full wavefunction evaluations at the shifted configurations (the route of
    _compute_discretized_kinetic_energy_debug, wavefunction.py:1563-1616);
  * -1/(2a^2) sum_k (ratio_k - 1) -> T_e as a -> 0 (second-order finite-difference Laplacian);
  * with no kinetic sign flip and t-move ECP, V_diag + V_nondiag equals the continuum local
    energy exactly (the 1/(4a^2) bookkeeping cancels), and the local energy is itself pinned on
    the TurboRVB golden values;
  * DLA t-move equals t-move when the Jastrow is off;
  * the Sherman-Morrison inverse after a move equals the inverse of the rebuilt geminal;
  * the comb gives walker i floor(W p_i) or ceil(W p_i) copies and the identity for equal weights.
"""

from __future__ import annotations

import numpy as np

from . import coulomb as oc
from . import rng as orng
from . import wf as owf

_DIRS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)


def step_draws(seed, walker, step):
    """Draws of one projection: three rotation angles in U(-2pi, 2pi) (sub-streams 0-1), the
    move uniform and the waiting-time uniform xi (sub-stream 2) of the LRDMC stream."""
    angles = orng.rotation_angles(seed, walker, step, orng.STREAM_LRDMC)
    w0, w1, w2, w3 = orng.draw_words(seed, walker, step, orng.STREAM_LRDMC, 2)
    return [float(a) for a in angles], float(orng.uniform01(w0, w1)), float(orng.uniform01(w2, w3))


def step_RT(seed, walker, step, random_mesh=True):
    """R^T of ``_generate_rotation_matrix_n`` (jqmc_gfmc.py:4721-4736, 4810-4818)."""
    if not random_mesh:
        return np.eye(3)
    angles, _, _ = step_draws(seed, walker, step)
    return orng.rotation_matrix_T(*angles)


def _el_ion(cp, r, floor=None):
    """``compute_bare_coulomb_potential_el_ion_element_wise`` and its discretized twin with the
    distance floored at ``floor`` (coulomb_potential.py:2284-2378)."""
    R = cp.structure_data.positions
    Z = cp._effective_charges
    d = np.linalg.norm(r[:, None, :] - R[None, :, :], axis=-1)
    if floor is not None:
        d = np.maximum(d, floor)
    return np.sum(-Z[None, :] / d, axis=1)


def _el_el(r):
    n = len(r)
    v = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            v += 1.0 / np.linalg.norm(r[i] - r[j])
    return v


def _ion_ion(cp):
    R = cp.structure_data.positions
    Z = cp._effective_charges
    v = 0.0
    for a in range(len(Z)):
        for b in range(a + 1, len(Z)):
            v += Z[a] * Z[b] / np.linalg.norm(R[a] - R[b])
    return v


def det_ratio_single_electron(gem, r_up, r_dn, spin, i, r_new, Ginv):
    """Determinant-only ratio of one moved electron (determinant.py:1665-1783)."""
    if spin == 0:
        return float(owf.geminal_up_one_row(gem, r_new, r_dn) @ Ginv[:, i])
    return float(Ginv[i, :] @ owf.geminal_dn_one_column(gem, r_up, r_new))


def lattice_elements(wf, r_up, r_dn, Ginv, RT, alat):
    """``compute_discretized_kinetic_energy_fast_update`` (wavefunction.py:1739-1860): elements
    -Psi(r_i + s_k)/Psi / (2 a^2) for shifts s = a [+x, -x, +y, -y, +z, -z] R^T, electron-major with
    the up block first.  Returns (elements (6N,), points [(spin, index, r')])."""
    shifts = (alat * _DIRS) @ RT
    pts, elem = [], []
    for spin, rr in ((0, r_up), (1, r_dn)):
        for i in range(len(rr)):
            for k in range(6):
                rp = rr[i] + shifts[k]
                elem.append(-1.0 / (2.0 * alat**2) * oc.ratio_single_electron(wf, r_up, r_dn, spin, i, rp, Ginv))
                pts.append((spin, i, rp))
    return np.asarray(elem), pts


def hamiltonian_elements(ham, r_up, r_dn, Ginv, RT, alat, non_local_move="tmove", Nv=6, NN=1):
    """Diagonal and off-diagonal parts of the lattice-regularised Hamiltonian at one configuration
    (_body_step_core up to the weight update, jqmc_gfmc.py:4807-5045; identical to
    _compute_V_elements_n, 5360-5628).  Returns (diag, nondiag, p_list, points) with points[k] =
    (spin, index, r') of candidate move k: 6 lattice moves per electron (up block first, order
    +x, -x, +y, -y, +z, -z of the rotated axes) then the ECP mesh (electron, nucleus, point)."""
    wf = ham.wavefunction_data
    gem = wf.geminal_data
    cp = ham.coulomb_potential_data
    r_up = np.reshape(np.asarray(r_up, dtype=np.float64), (-1, 3))
    r_dn = np.reshape(np.asarray(r_dn, dtype=np.float64), (-1, 3))
    n = len(r_up) + len(r_dn)
    tu, td = owf.kinetic_all_elements_fast(wf, r_up, r_dn, Ginv)
    t_cont = np.concatenate([tu, td])
    elem, pts = lattice_elements(wf, r_up, r_dn, Ginv, RT, alat)
    FN = np.minimum(elem, 0.0)
    SP = float(np.sum(np.maximum(elem, 0.0)))
    per_e = elem.reshape(n, 6)
    flags = np.any(per_e >= 0, axis=1)
    nd_e = np.sum(per_e + 1.0 / (4.0 * alat**2), axis=1)
    r_all = np.vstack([r_up, r_dn])
    v_ei = _el_ion(cp, r_all)
    zv = v_ei + t_cont - nd_e
    ei = v_ei if cp.ecp_flag else _el_ion(cp, r_all, floor=alat)
    opt = np.where(flags, np.maximum(zv, ei), zv)
    diag = 3.0 / (2.0 * alat**2) * n + (_el_el(r_all) + _ion_ion(cp) + float(np.sum(opt))) + SP
    nondiag = float(np.sum(FN))
    p_list = list(FN)
    if cp.ecp_flag:
        diag += oc.ecp_local(cp, r_up, r_dn)
        V, Jr = [], []
        for spin, i, rp, wsum in oc.ecp_nonlocal_mesh(cp, r_up, r_dn, RT, NN, Nv):
            if non_local_move == "tmove":
                V.append(wsum * oc.ratio_single_electron(wf, r_up, r_dn, spin, i, rp, Ginv))
                Jr.append(1.0)
            elif non_local_move == "dltmove":
                V.append(wsum * det_ratio_single_electron(gem, r_up, r_dn, spin, i, rp, Ginv))
                nu_, nd_ = r_up.copy(), r_dn.copy()
                (nu_ if spin == 0 else nd_)[i] = rp
                Jr.append(np.exp(owf.jastrow_part(wf.jastrow_data, nu_, nd_) - owf.jastrow_part(wf.jastrow_data, r_up, r_dn)))
            else:
                raise NotImplementedError(non_local_move)
            pts.append((spin, i, rp))
        V = np.asarray(V)
        diag += float(np.sum(np.maximum(V, 0.0)))
        V_FN = np.minimum(V, 0.0) * np.asarray(Jr)
        nondiag += float(np.sum(V_FN))
        p_list += list(V_FN)
    return diag, nondiag, np.asarray(p_list), pts


def choose_move(p_list, u):
    """k = searchsorted(cumsum(p / sum p), u) (jqmc_gfmc.py:5057-5062); an index past the end is
    clamped to the last candidate as JAX's gather does."""
    cdf = np.cumsum(p_list / p_list.sum())
    return min(int(np.searchsorted(cdf, u, side="left")), len(p_list) - 1)


def apply_move(gem, r_up, r_dn, Ginv, G, spin, i, rp):
    """Moves one electron and updates G and its inverse by Sherman-Morrison
    (_update_inv_up_n / _update_inv_dn_n, jqmc_gfmc.py:5083-5141)."""
    r_up = r_up.copy()
    r_dn = r_dn.copy()
    nu = len(r_up)
    G = G.copy()
    if spin == 0:
        v = owf.geminal_up_one_row(gem, rp, r_dn) - owf.geminal_up_one_row(gem, r_up[i], r_dn)
        u = np.zeros(nu)
        u[i] = 1.0
        r_up[i] = rp
        G[i, :] += v
    else:
        u = owf.geminal_dn_one_column(gem, r_up, rp) - owf.geminal_dn_one_column(gem, r_up, r_dn[i])
        v = np.zeros(nu)
        v[i] = 1.0
        r_dn[i] = rp
        G[:, i] += u
    Ainv_u = Ginv @ u
    vT_Ainv = v @ Ginv
    Ginv = Ginv - np.outer(Ainv_u, vT_Ainv) / (1.0 + v @ Ainv_u)
    return r_up, r_dn, Ginv, G


def projection_step(ham, r_up, r_dn, Ginv, G, w, E_scf, alat, seed, walker, step, non_local_move="tmove",
                    random_mesh=True, Nv=6, NN=1):
    """One GFMC_n projection (_body_step_core, jqmc_gfmc.py:4779-5166).
    Returns dict(w, r_up, r_dn, Ginv, G, diag, nondiag, k, RT)."""
    RT = step_RT(seed, walker, step, random_mesh)
    _, u, _ = step_draws(seed, walker, step)
    r_up = np.reshape(np.asarray(r_up, dtype=np.float64), (-1, 3))
    r_dn = np.reshape(np.asarray(r_dn, dtype=np.float64), (-1, 3))
    diag, nondiag, p_list, pts = hamiltonian_elements(ham, r_up, r_dn, Ginv, RT, alat, non_local_move, Nv, NN)
    b_x = 1.0 / (diag - E_scf) * (-nondiag)
    w = w * b_x
    k = choose_move(p_list, u)
    spin, i, rp = pts[k]
    r_up, r_dn, Ginv, G = apply_move(ham.wavefunction_data.geminal_data, r_up, r_dn, Ginv, G, spin, i, rp)
    return dict(w=w, r_up=r_up, r_dn=r_dn, Ginv=Ginv, G=G, diag=diag, nondiag=nondiag, k=k, RT=RT)


def projection_t_step(ham, r_up, r_dn, Ginv, G, w, tau_left, alat, seed, walker, step, non_local_move="tmove",
                      random_mesh=True, Nv=6, NN=1):
    """One GFMC_t projection (_projection_t_core, jqmc_gfmc.py:724-1107): waiting time
    dtau = min(tau_left, ln(1 - xi) / sum offdiag), w <- w exp(-dtau e_L), no move once
    tau_left <= 0.  Returns dict(w, tau_left, r_up, r_dn, Ginv, G, e_L, k, moved, RT)."""
    RT = step_RT(seed, walker, step, random_mesh)
    _, u, xi = step_draws(seed, walker, step)
    r_up = np.reshape(np.asarray(r_up, dtype=np.float64), (-1, 3))
    r_dn = np.reshape(np.asarray(r_dn, dtype=np.float64), (-1, 3))
    diag, nondiag, p_list, pts = hamiltonian_elements(ham, r_up, r_dn, Ginv, RT, alat, non_local_move, Nv, NN)
    e_L = diag + nondiag
    dtau = min(tau_left, np.log(1.0 - xi) / nondiag)
    w = w * np.exp(-dtau * e_L)
    tau_left = tau_left - dtau
    k = choose_move(p_list, u)
    moved = tau_left > 0.0
    if moved:
        spin, i, rp = pts[k]
        r_up, r_dn, Ginv, G = apply_move(ham.wavefunction_data.geminal_data, r_up, r_dn, Ginv, G, spin, i, rp)
    return dict(w=w, tau_left=tau_left, r_up=r_up, r_dn=r_dn, Ginv=Ginv, G=G, e_L=e_L, k=k, moved=moved, RT=RT)


def comb_branching(w, zeta):
    """Comb reconfiguration (jqmc_gfmc.py:6070-6130, one rank): walker j becomes a copy of
    walker searchsorted(cumsum(w / sum w), (j + zeta) / W)."""
    w = np.asarray(w, dtype=np.float64)
    W = len(w)
    cum = np.cumsum(w / np.sum(w))
    z = (np.arange(W) + zeta) / W
    return np.minimum(np.searchsorted(cum, z), W - 1).astype(np.int64)


def weighted_sums(w, V_diag, e_L, E_scf):
    """Per-step sums of jqmc_gfmc.py:5957-5964: (sum w, sum w/(V_diag - E), sum w e_L/(V_diag - E),
    sum w e_L^2/(V_diag - E))."""
    w = np.asarray(w, dtype=np.float64)
    f = w / (np.asarray(V_diag) - E_scf)
    e = np.asarray(e_L)
    return np.array([np.sum(w), np.sum(f), np.sum(f * e), np.sum(f * e**2)])
