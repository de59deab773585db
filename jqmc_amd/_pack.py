"""Flatten ``Hamiltonian_data`` into the C-ABI system descriptor and upload it.

The pack follows the reference's own preprocessing: Cartesian primitives grouped per AO
(the contraction-depth grouping of ``_build_prim_groups_by_K``, atomic_orbital.py:2036-2087,
becomes a CSR row per AO), the primitive normalisation folded into the coefficients
(atomic_orbital.py:2237-2251), the geminal ``lambda`` split into paired/unpaired blocks
(determinant.py:1390-1394) and the ECP terms split into local and projected tables padded
per atom (``_padded_parameters_tuple``, coulomb_potential.py:482-598).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._solid_harmonics import solid_harmonic_terms, spherical_prim_normalisation
from .atomic_orbital import AOs_cart_data, AOs_sphe_data
from .coulomb_potential import get_mesh
from .hamiltonians import Hamiltonian_data
from .molecular_orbital import MOs_data

_JTYPE = {"exp": 1, "pade": 2}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if a is not None else None


def _flatten_basis(bases):
    """Concatenate AO bases into device tables: per AO its nucleus, primitive range (exponent and
    coefficient with the normalisation folded in) and its angular polynomial as monomial terms."""
    ao_nuc, prim_exp, prim_coef, pp, tp, tcoef, tpow = [], [], [], [0], [0], [], []
    for aos in bases:
        idx = np.asarray(aos.orbital_indices)
        order = np.argsort(idx, kind="stable")
        counts = np.bincount(idx, minlength=aos.num_ao)
        if isinstance(aos, AOs_cart_data):
            coef = aos._normalized_coefficients_np
        elif isinstance(aos, AOs_sphe_data):
            # N_p(l) sqrt((2l+1)/4pi) c_p (atomic_orbital.py:2315-2327)
            ls = np.asarray([aos.angular_momentums[a] for a in idx])
            N = np.array([spherical_prim_normalisation(int(l), z) for l, z in zip(ls, aos.exponents)])
            coef = N * np.sqrt((2 * ls + 1) / (4.0 * np.pi)) * aos.coefficients
        else:
            raise NotImplementedError(f"unsupported basis {type(aos).__name__}")
        prim_exp.append(aos.exponents[order])
        prim_coef.append(coef[order])
        for a in range(aos.num_ao):
            ao_nuc.append(aos.nucleus_index[a])
            pp.append(pp[-1] + int(counts[a]))
            if isinstance(aos, AOs_cart_data):
                terms = (((aos.polynominal_order_x[a], aos.polynominal_order_y[a], aos.polynominal_order_z[a]), 1.0),)
            else:
                terms = solid_harmonic_terms(aos.angular_momentums[a], aos.magnetic_quantum_numbers[a])
            for pw, c in terms:
                tpow.append(pw)
                tcoef.append(c)
            tp.append(len(tcoef))
    return {
        "ao_nuc": np.asarray(ao_nuc),
        "prim_exp": np.concatenate(prim_exp),
        "prim_coef": np.concatenate(prim_coef),
        "ao_pp": np.asarray(pp),
        "ao_tp": np.asarray(tp),
        "term_coef": np.asarray(tcoef, dtype=np.float64),
        "term_pow": np.asarray(tpow, dtype=np.int64).reshape(-1, 3),
    }


def _same_basis(a, b) -> bool:
    if a is b:
        return True
    if type(a) is not type(b) or a.num_ao != b.num_ao or a.num_ao_prim != b.num_ao_prim:
        return False
    return (np.array_equal(a.exponents, b.exponents) and np.array_equal(a.coefficients, b.coefficients)
            and tuple(a.orbital_indices) == tuple(b.orbital_indices) and tuple(a.nucleus_index) == tuple(b.nucleus_index)
            and np.array_equal(a.structure_data.positions, b.structure_data.positions))


class SystemPack:
    """Host arrays of one system (kept alive for the descriptor) and its device handle.

    The device sees one AO set: the geminal's basis, followed by the J3 basis when that is a
    different basis (the orbital coefficients and J3 matrix are zero-padded over the other
    block, so each contraction only picks up its own basis)."""

    def __init__(self, ham: Hamiltonian_data, Nv: int = 6, NN: int = 1, precision: str = "fp64", device: bool = True):
        """``device=False`` builds the host descriptor only (no device system, ``handle`` stays NULL)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        ham.sanity_check()
        wf = ham.wavefunction_data
        gem = wf.geminal_data
        jas = wf.jastrow_data
        cp = ham.coulomb_potential_data
        st = ham.structure_data
        aos = gem.aos_data
        j3 = jas.jastrow_three_body_data
        if j3 is not None and isinstance(j3.orb_data, MOs_data):
            raise NotImplementedError("J3 over molecular orbitals is not supported; use an AO basis")
        jas_separate = j3 is not None and not _same_basis(j3.orb_data, aos)
        bases = [aos] + ([j3.orb_data] if jas_separate else [])
        flat = _flatten_basis(bases)
        n_det = aos.num_ao
        n_tot = len(flat["ao_nuc"])
        self.n_det_ao = n_det
        self.jas_separate = jas_separate
        self.n_up = gem.num_electron_up
        self.n_dn = gem.num_electron_dn
        self.n_el = self.n_up + self.n_dn
        self.n_ao = n_tot
        self.nv = Nv
        self.nn = NN
        self.precision = precision
        self.ecp_on = bool(cp.ecp_flag)
        self.j3_on = j3 is not None
        ao_orbitals = not isinstance(gem.orb_data_up_spin, MOs_data)
        # orbitals = AOs of a padded union basis are expressed as MOs with C = [I | 0]
        self.orb_is_mo = (not ao_orbitals) or jas_separate
        self.n_orb = gem.orb_num_up
        keep = {}

        def f64(name, x):
            keep[name] = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            return keep[name]

        def i32(name, x):
            keep[name] = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
            return keep[name]

        def pad_cols(M):
            out = np.zeros((M.shape[0], n_tot))
            out[:, :n_det] = M
            return out

        d = _lib.SystemDesc()
        d.n_nuc = st.natom
        d.nuc_pos = _dptr(f64("nuc_pos", st.positions))
        d.nuc_charge = _dptr(f64("nuc_charge", cp._effective_charges))
        d.n_ao = n_tot
        d.n_prim = len(flat["prim_exp"])
        d.ao_nuc = _iptr(i32("ao_nuc", flat["ao_nuc"]))
        d.ao_prim_ptr = _iptr(i32("ao_pp", flat["ao_pp"]))
        d.prim_exp = _dptr(f64("prim_exp", flat["prim_exp"]))
        d.prim_coef = _dptr(f64("prim_coef", flat["prim_coef"]))
        d.n_terms = len(flat["term_coef"])
        d.ao_term_ptr = _iptr(i32("ao_tp", flat["ao_tp"]))
        d.term_coef = _dptr(f64("term_coef", flat["term_coef"]))
        d.term_pow = _iptr(i32("term_pow", flat["term_pow"]))
        d.n_up = self.n_up
        d.n_dn = self.n_dn
        d.orb_is_mo = int(self.orb_is_mo)
        d.n_orb = self.n_orb
        if self.orb_is_mo:
            if ao_orbitals:
                Cu = Cd = np.eye(n_det)
            else:
                Cu, Cd = gem.orb_data_up_spin.mo_coefficients, gem.orb_data_dn_spin.mo_coefficients
            d.mo_coef_up = _dptr(f64("mo_up", pad_cols(Cu)))
            d.mo_coef_dn = _dptr(f64("mo_dn", pad_cols(Cd)))
        d.lambda_matrix = _dptr(f64("lambda", gem.lambda_matrix))
        j1 = jas.jastrow_one_body_data
        if j1 is not None:
            d.j1_type = _JTYPE[j1.jastrow_1b_type]
            d.j1_param = j1.jastrow_1b_param
            d.j1_zeff = _dptr(f64("j1_zeff", j1.effective_charges))
        j2 = jas.jastrow_two_body_data
        if j2 is not None:
            d.j2_type = _JTYPE[j2.jastrow_2b_type]
            d.j2_param = j2.jastrow_2b_param
        if j3 is not None:
            d.j3_on = 1
            jm = j3.j_matrix
            if jas_separate:
                nj = j3.orb_num
                full = np.zeros((n_tot, n_tot + 1))
                full[n_det:, n_det:n_tot] = jm[:, :nj]
                full[n_det:, n_tot] = jm[:, nj]
                jm = full
            d.j3_matrix = _dptr(f64("j3", jm))
        if self.ecp_on:
            n = st.natom
            nl_tab = [[] for _ in range(n)]
            for a, l, ex, c, p in cp._non_local_terms():
                nl_tab[a].append((l, ex, c, p))
            lo_tab = [[] for _ in range(n)]
            for a, ex, c, p in cp._local_terms():
                lo_tab[a].append((ex, c, p))
            mt = max(1, max(len(t) for t in nl_tab))
            ml = max(1, max(len(t) for t in lo_tab))
            L = np.zeros((n, mt), np.int32)
            E = np.zeros((n, mt))
            C = np.zeros((n, mt))
            P = np.zeros((n, mt))
            gmax = cp._global_max_ang_mom_plus_1
            nt = np.zeros(n, np.int32)
            for a, terms in enumerate(nl_tab):
                # only channels l < global max(l_max + 1) are projected (coulomb_potential.py:1604-1606)
                terms = [t for t in terms if t[0] < gmax]
                nt[a] = len(terms)
                for k, (l, ex, c, p) in enumerate(terms):
                    L[a, k], E[a, k], C[a, k], P[a, k] = l, ex, c, p
            LE = np.zeros((n, ml))
            LC = np.zeros((n, ml))
            LP = np.zeros((n, ml))
            lnt = np.zeros(n, np.int32)
            for a, terms in enumerate(lo_tab):
                lnt[a] = len(terms)
                for k, (ex, c, p) in enumerate(terms):
                    LE[a, k], LC[a, k], LP[a, k] = ex, c, p
            mesh = get_mesh(Nv)
            d.ecp_on = 1
            d.ecp_lmax = gmax
            d.ecp_max_terms = mt
            d.ecp_n_terms = _iptr(i32("ecp_nt", nt))
            d.ecp_term_l = _iptr(i32("ecp_l", L))
            d.ecp_term_exp = _dptr(f64("ecp_e", E))
            d.ecp_term_coef = _dptr(f64("ecp_c", C))
            d.ecp_term_pow = _dptr(f64("ecp_p", P))
            d.ecp_max_local = ml
            d.ecpl_n_terms = _iptr(i32("ecpl_nt", lnt))
            d.ecpl_exp = _dptr(f64("ecpl_e", LE))
            d.ecpl_coef = _dptr(f64("ecpl_c", LC))
            d.ecpl_pow = _dptr(f64("ecpl_p", LP))
            d.nv = Nv
            d.mesh_points = _dptr(f64("mesh_p", mesh.grid_points))
            d.mesh_weights = _dptr(f64("mesh_w", mesh.weights))
            d.nn = NN
        d.precision = _lib.PRECISION_MIXED_FP32 if precision == "mixed" else _lib.PRECISION_PLAIN_FP64
        self._keep = keep
        self.desc = d
        self.handle = ctypes.c_void_p()
        if device:
            _lib.call("jqmc_system_create", ctypes.byref(d), ctypes.byref(self.handle))

    def close(self):
        if self.handle:
            _lib.load().jqmc_system_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_CACHE: dict = {}


def system_for(ham: Hamiltonian_data, Nv: int = 6, NN: int = 1, precision: str = "fp64") -> SystemPack:
    """Device system pack for ``ham`` (cached per object identity, Nv, NN and precision)."""
    key = (id(ham), Nv, NN, precision)
    hit = _CACHE.get(key)
    if hit is not None and hit[0] is ham:
        return hit[1]
    pack = SystemPack(ham, Nv=Nv, NN=NN, precision=precision)
    _CACHE[key] = (ham, pack)
    return pack
