"""TREXIO (HDF5 back end) -> Hamiltonian data classes, without h5py or the trexio library.

``read_trexio_file`` keeps the signature and return tuple of the reference
(trexio_wrapper.py:63-564): ``(structure_data, aos_data, mos_data_up, mos_data_dn,
geminal_data, coulomb_potential_data)``.  Conversion rules (same numbers as the reference):

* Cartesian shells: components in ``itertools.combinations_with_replacement("xyz", l)``
  order; coefficients ``shell_factor * prim_factor / sqrt(N_Z N_fact) * coefficient *
  ao_normalization`` (trexio_wrapper.py:276-310), so that the AO normalisation the
  evaluator re-applies cancels.
* Spherical shells: ``m = 0, +1, -1, +2, -2, ...``; coefficients ``shell_factor *
  prim_factor * sqrt(4 pi) / sqrt(2l+1) / N_p * coefficient * ao_normalization``
  (trexio_wrapper.py:352-404).
* MOs: orbitals with occupation >= 1e-6 and an identity-padded pairing matrix.  Spin-independent
  files (every ``mo_spin`` 0) give the same coefficients to both spins (trexio_wrapper.py:451-478);
  spin-dependent (UHF) files with N_up != N_dn take the spin-0 orbitals for up and the spin-1
  orbitals for down, both at the positions of the occupied up orbitals (trexio_wrapper.py:480-510);
  spin-dependent files with N_up == N_dn are refused, as in the reference.
* ECP: TREXIO power + 2 (trexio_wrapper.py:539).
"""

from __future__ import annotations

import itertools
from math import factorial

import numpy as np

from . import _hdf5
from ._solid_harmonics import spherical_prim_normalisation
from .atomic_orbital import AOs_cart_data, AOs_sphe_data
from .coulomb_potential import Coulomb_potential_data
from .determinant import Geminal_data
from .molecular_orbital import MOs_data
from .structure import Structure_data

_ELEMENTS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn"
).split()


def _atomic_number(label: str) -> int:
    sym = "".join(ch for ch in label if ch.isalpha())
    sym = sym[0].upper() + sym[1:].lower() if sym else sym
    if sym in _ELEMENTS:
        return _ELEMENTS.index(sym)
    if sym[:1] in _ELEMENTS:
        return _ELEMENTS.index(sym[:1])
    raise ValueError(f"unknown element label {label}")


def read_trexio_file(trexio_file: str, store_tuple: bool = True):
    f = _hdf5.File(trexio_file)
    if int(f["pbc"].attrs.get("pbc_periodic", 0)) != 0:
        raise NotImplementedError("periodic TREXIO files are outside the MI355X hot path")
    num_up = int(f["electron"].attrs["electron_up_num"])
    num_dn = int(f["electron"].attrs["electron_dn_num"])
    labels = [s.strip() for s in f["nucleus/nucleus_label"].read()]
    coords = np.asarray(f["nucleus/nucleus_coord"].read(), dtype=np.float64).reshape(-1, 3)
    b = f["basis"]
    shell_num = int(b.attrs["basis_shell_num"])
    shell_index = np.asarray(b["basis_shell_index"].read(), dtype=np.int64)
    nucleus_index = np.asarray(b["basis_nucleus_index"].read(), dtype=np.int64)
    shell_l = np.asarray(b["basis_shell_ang_mom"].read(), dtype=np.int64)
    shell_factor = np.asarray(b["basis_shell_factor"].read(), dtype=np.float64)
    exponent = np.asarray(b["basis_exponent"].read(), dtype=np.float64)
    coefficient = np.asarray(b["basis_coefficient"].read(), dtype=np.float64)
    prim_factor = np.asarray(b["basis_prim_factor"].read(), dtype=np.float64)
    ao_cartesian = int(f["ao"].attrs["ao_cartesian"])
    ao_num = int(f["ao"].attrs["ao_num"])
    ao_norm = np.asarray(f["ao/ao_normalization"].read(), dtype=np.float64)
    mo_coef = np.asarray(f["mo/mo_coefficient"].read(), dtype=np.float64)
    mo_occ = np.asarray(f["mo/mo_occupation"].read(), dtype=np.float64)
    mo_spin = np.asarray(f["mo/mo_spin"].read(), dtype=np.int64)
    ecp_group = f["ecp"]
    ecp_flag = "ecp_num" in ecp_group.attrs

    structure = Structure_data(
        positions=coords,
        pbc_flag=False,
        atomic_numbers=[_atomic_number(x) for x in labels],
        element_symbols=labels,
        atomic_labels=labels,
    )

    nuc, orb_idx, exps, coefs, angs = [], [], [], [], []
    px, py, pz, mags = [], [], [], []
    ao = 0
    for s in range(shell_num):
        l = int(shell_l[s])
        prims = np.nonzero(shell_index == s)[0]
        if ao_cartesian:
            comps = ["".join(p) for p in itertools.combinations_with_replacement("xyz", l)]
            powers = [(c.count("x"), c.count("y"), c.count("z")) for c in comps]
            for nx, ny, nz in powers:
                fact = factorial(nx) * factorial(ny) * factorial(nz) / (factorial(2 * nx) * factorial(2 * ny) * factorial(2 * nz))
                for k in prims:
                    NZ = (2.0 * exponent[k] / np.pi) ** 1.5 * (8.0 * exponent[k]) ** l
                    c = shell_factor[s] * prim_factor[k] / np.sqrt(NZ * fact) * coefficient[k]
                    coefs.append(c * ao_norm[ao])
                    exps.append(exponent[k])
                    orb_idx.append(ao)
                nuc.append(int(nucleus_index[s]))
                angs.append(l)
                px.append(nx)
                py.append(ny)
                pz.append(nz)
                ao += 1
        else:
            ms = [0] + [i * (-1) ** j for i in range(1, l + 1) for j in range(2)]
            norms = spherical_prim_normalisation(l, exponent[prims])
            base = [
                shell_factor[s] * prim_factor[k] * np.sqrt(4 * np.pi) / np.sqrt(2 * l + 1) / norms[i] * coefficient[k]
                for i, k in enumerate(prims)
            ]
            for m in ms:
                for i, k in enumerate(prims):
                    coefs.append(base[i] * ao_norm[ao])
                    exps.append(exponent[k])
                    orb_idx.append(ao)
                nuc.append(int(nucleus_index[s]))
                angs.append(l)
                mags.append(m)
                ao += 1
    if ao != ao_num:
        raise ValueError(f"ao count {ao} is inconsistent with ao_num {ao_num}")
    if ao_cartesian:
        aos = AOs_cart_data(structure, nuc, ao, len(exps), orb_idx, np.array(exps), np.array(coefs), angs, px, py, pz)
    else:
        aos = AOs_sphe_data(structure, nuc, ao, len(exps), orb_idx, np.array(exps), np.array(coefs), angs, mags)

    thr = 1.0e-6
    if not np.any(mo_spin != 0):
        occ = [i for i, v in enumerate(mo_occ) if v >= thr]
        C_up = C_dn = mo_coef[occ]
    elif num_up != num_dn:
        up, dn = np.flatnonzero(mo_spin == 0), np.flatnonzero(mo_spin == 1)
        occ_up = [i for i, v in enumerate(mo_occ[up]) if v >= thr]
        occ_dn = [i for i, v in enumerate(mo_occ[dn]) if v >= thr]
        if len(occ_up) < len(occ_dn):
            raise ValueError(f"the number of occupied up orbitals ({len(occ_up)}) must not be smaller than "
                             f"that of down orbitals ({len(occ_dn)})")
        C_up = mo_coef[up][occ_up]
        C_dn = mo_coef[dn][occ_up]
    else:
        raise NotImplementedError("spin-dependent MOs with as many up as down electrons")
    n_mo = C_up.shape[0]
    mos_up = MOs_data(num_mo=n_mo, aos_data=aos, mo_coefficients=C_up)
    mos_dn = MOs_data(num_mo=n_mo, aos_data=aos, mo_coefficients=C_dn)
    diff = num_up - num_dn
    lam_p = np.pad(np.eye(num_dn), ((0, n_mo - num_dn), (0, n_mo - num_dn)))
    lam_u = np.pad(np.eye(diff), ((num_dn, n_mo - num_dn - diff), (0, 0)))
    geminal = Geminal_data(num_up, num_dn, mos_up, mos_dn, np.hstack([lam_p, lam_u]))

    if ecp_flag:
        e = ecp_group
        cp = Coulomb_potential_data(
            structure_data=structure,
            ecp_flag=True,
            z_cores=list(np.asarray(e["ecp_z_core"].read()).ravel()),
            max_ang_mom_plus_1=list(np.asarray(e["ecp_max_ang_mom_plus_1"].read()).ravel()),
            num_ecps=int(e.attrs["ecp_num"]),
            ang_moms=list(np.asarray(e["ecp_ang_mom"].read()).ravel()),
            nucleus_index=list(np.asarray(e["ecp_nucleus_index"].read()).ravel()),
            exponents=list(np.asarray(e["ecp_exponent"].read()).ravel()),
            coefficients=list(np.asarray(e["ecp_coefficient"].read()).ravel()),
            powers=list(np.asarray(e["ecp_power"].read()).ravel() + 2),
        )
    else:
        cp = Coulomb_potential_data(structure_data=structure, ecp_flag=False)
    return structure, aos, mos_up, mos_dn, geminal, cp
