"""FP64 NumPy restatement of the reference Coulomb/ECP potential (test infrastructure only)."""

from __future__ import annotations

import numpy as np

from jqmc_amd.coulomb_potential import get_mesh

from . import wf as owf


def legendre_tabulated(n: int, x):
    """``_legendre_tablated`` P_0..P_6 (_function_collections.py:47-66)."""
    table = [
        lambda x: np.ones_like(x),
        lambda x: x,
        lambda x: 0.5 * (3.0 * x**2 - 1.0),
        lambda x: 0.5 * (5.0 * x**3 - 3.0 * x),
        lambda x: 0.125 * (35.0 * x**4 - 30.0 * x**2 + 3.0),
        lambda x: 0.125 * (63.0 * x**5 - 70.0 * x**3 + 15.0 * x),
        lambda x: (231.0 * x**6 - 315.0 * x**4 + 105.0 * x**2 - 5.0) / 16.0,
    ]
    return table[n](np.asarray(x, dtype=np.float64))


def bare_coulomb(cp, r_up, r_dn) -> float:
    """``compute_bare_coulomb_potential`` (coulomb_potential.py:2252-2281): ion-ion + el-ion + el-el."""
    R = cp.structure_data.positions
    Z = cp._effective_charges
    r = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    v = 0.0
    n_nuc = len(Z)
    for a in range(n_nuc):
        for b in range(a + 1, n_nuc):
            v += Z[a] * Z[b] / np.linalg.norm(R[a] - R[b])
    v += np.sum(-Z[None, :] / np.linalg.norm(r[:, None, :] - R[None, :, :], axis=-1))
    iu, ju = np.triu_indices(len(r), k=1)
    v += np.sum(1.0 / np.linalg.norm(r[iu] - r[ju], axis=-1))
    return float(v)


def ecp_local(cp, r_up, r_dn) -> float:
    """``compute_ecp_local_parts_all_pairs`` (coulomb_potential.py:1144-1246): sum c r^(n-2) e^(-a r^2)."""
    if not cp.ecp_flag:
        return 0.0
    r = np.vstack([np.reshape(r_up, (-1, 3)), np.reshape(r_dn, (-1, 3))])
    R = cp.structure_data.positions
    v = 0.0
    for a, ex, c, p in cp._local_terms():
        d = np.linalg.norm(R[a][None, :] - r, axis=-1)
        v += np.sum(d**-2.0 * c * d**p * np.exp(-ex * d**2))
    return float(v)


def _nonlocal_table(cp):
    """Per-atom list of (l, exponent, coefficient, power) for projected terms."""
    tab = {}
    for a, l, ex, c, p in cp._non_local_terms():
        tab.setdefault(a, []).append((l, ex, c, p))
    return tab


def ecp_nonlocal_mesh(cp, r_up, r_dn, RT, NN: int = 1, Nv: int = 6):
    """Mesh of the non-local projector (coulomb_potential.py:1546-1625): for every electron (up
    block then dn block), nearest nuclei, Nv rotated points r' = r + (R - r) + |R - r| g~.
    Returns a list of (spin, electron, r_prime, weight_factor) with weight_factor =
    sum_l V_l(|R-r|) (2l+1) P_l(cos theta) w_k."""
    mesh = get_mesh(Nv)
    gp = mesh.grid_points @ RT
    gnorm = np.linalg.norm(gp, axis=1, keepdims=True)
    gunit = gp / gnorm
    R = cp.structure_data.positions
    tab = _nonlocal_table(cp)
    out = []
    for spin, rr in ((0, np.reshape(r_up, (-1, 3))), (1, np.reshape(r_dn, (-1, 3)))):
        for i, r in enumerate(rr):
            d_all = np.sqrt(np.sum((R - r) ** 2, axis=1))
            for a in np.argsort(d_all, kind="stable")[:NN]:
                rel = R[a] - r
                rn = np.linalg.norm(rel)
                terms = tab.get(int(a), [])
                cos_t = (-rel / rn) @ gunit.T
                for k in range(Nv):
                    rp = r + rel + rn * gp[k]
                    wsum = 0.0
                    for l, ex, c, p in terms:
                        V = rn**-2.0 * c * rn**p * np.exp(-ex * rn**2)
                        wsum += V * (2 * l + 1) * legendre_tabulated(l, cos_t[k]) * mesh.weights[k]
                    out.append((spin, i, rp, wsum))
    return out


def ratio_single_electron(wf, r_up, r_dn, spin: int, i: int, r_new, Ginv) -> float:
    """Psi(r')/Psi(r) for one moved electron: split-spin determinant ratio with G^-1
    (determinant.py:1665-1783) times exp(J(r') - J(r)) (jastrow_factor.py:2696-2960)."""
    gem = wf.geminal_data
    r_up = np.reshape(r_up, (-1, 3))
    r_dn = np.reshape(r_dn, (-1, 3))
    if spin == 0:
        row = owf.geminal_up_one_row(gem, r_new, r_dn)
        det_ratio = float(row @ Ginv[:, i])
        nu = r_up.copy()
        nu[i] = r_new
        nd = r_dn
    else:
        col = owf.geminal_dn_one_column(gem, r_up, r_new)
        det_ratio = float(Ginv[i, :] @ col)
        nu = r_up
        nd = r_dn.copy()
        nd[i] = r_new
    jr = np.exp(owf.jastrow_part(wf.jastrow_data, nu, nd) - owf.jastrow_part(wf.jastrow_data, r_up, r_dn))
    return det_ratio * float(jr)


def ecp_nonlocal_fast(cp, wf, r_up, r_dn, RT, Ginv, NN: int = 1, Nv: int = 6) -> float:
    """``compute_ecp_non_local_parts_nearest_neighbors_fast_update`` (coulomb_potential.py:1477-1712)."""
    if not cp.ecp_flag:
        return 0.0
    total = 0.0
    for spin, i, rp, wsum in ecp_nonlocal_mesh(cp, r_up, r_dn, RT, NN, Nv):
        if wsum == 0.0:
            continue
        total += wsum * ratio_single_electron(wf, r_up, r_dn, spin, i, rp, Ginv)
    return float(total)


def coulomb_potential_fast(cp, wf, r_up, r_dn, RT, Ginv, NN: int = 1, Nv: int = 6) -> float:
    """``compute_coulomb_potential_fast`` (coulomb_potential.py:2695-2761)."""
    v = bare_coulomb(cp, r_up, r_dn)
    if cp.ecp_flag:
        v += ecp_local(cp, r_up, r_dn) + ecp_nonlocal_fast(cp, wf, r_up, r_dn, RT, Ginv, NN, Nv)
    return v


def local_energy_fast(ham, r_up, r_dn, RT, Ginv, NN: int = 1, Nv: int = 6) -> float:
    """``compute_local_energy_fast`` (hamiltonians.py:225-290): T + V with the caller's G^-1."""
    tu, td = owf.kinetic_all_elements_fast(ham.wavefunction_data, r_up, r_dn, Ginv)
    v = coulomb_potential_fast(ham.coulomb_potential_data, ham.wavefunction_data, r_up, r_dn, RT, Ginv, NN, Nv)
    return float(np.sum(tu) + np.sum(td) + v)


def local_energy(ham, r_up, r_dn, RT, NN: int = 1, Nv: int = 6) -> float:
    """``compute_local_energy`` (hamiltonians.py:179-222) with the inverse rebuilt from G."""
    Ginv = owf.geminal_inverse(owf.geminal(ham.wavefunction_data.geminal_data, r_up, r_dn))
    return local_energy_fast(ham, r_up, r_dn, RT, Ginv, NN, Nv)
