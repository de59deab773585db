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
from .atomic_orbital import AOs_cart_data
from .coulomb_potential import get_mesh
from .hamiltonians import Hamiltonian_data
from .molecular_orbital import MOs_data

_JTYPE = {"exp": 1, "pade": 2}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if a is not None else None


class SystemPack:
    """Host arrays of one system (kept alive for the descriptor) and its device handle."""

    def __init__(self, ham: Hamiltonian_data, Nv: int = 6, NN: int = 1):
        ham.sanity_check()
        wf = ham.wavefunction_data
        gem = wf.geminal_data
        jas = wf.jastrow_data
        cp = ham.coulomb_potential_data
        st = ham.structure_data
        aos = gem.aos_data
        if not isinstance(aos, AOs_cart_data):
            raise NotImplementedError("the gfx950 path evaluates Cartesian AO bases")
        j3 = jas.jastrow_three_body_data
        if j3 is not None and j3.orb_data is not aos and not (
            isinstance(j3.orb_data, AOs_cart_data) and j3.orb_data.num_ao == aos.num_ao
        ):
            raise NotImplementedError("J3 must be defined on the geminal's AO basis")
        self.n_up = gem.num_electron_up
        self.n_dn = gem.num_electron_dn
        self.n_el = self.n_up + self.n_dn
        self.n_ao = aos.num_ao
        self.nv = Nv
        self.nn = NN
        self.ecp_on = bool(cp.ecp_flag)
        self.j3_on = j3 is not None
        self.orb_is_mo = isinstance(gem.orb_data_up_spin, MOs_data)
        self.n_orb = gem.orb_num_up
        keep = {}

        def f64(name, x):
            keep[name] = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            return keep[name]

        def i32(name, x):
            keep[name] = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
            return keep[name]

        d = _lib.SystemDesc()
        d.n_nuc = st.natom
        d.nuc_pos = _dptr(f64("nuc_pos", st.positions))
        d.nuc_charge = _dptr(f64("nuc_charge", cp._effective_charges))
        # basis: primitives sorted by AO (stable), CSR offsets
        order = np.argsort(np.asarray(aos.orbital_indices), kind="stable")
        counts = np.bincount(np.asarray(aos.orbital_indices), minlength=aos.num_ao)
        d.n_ao = aos.num_ao
        d.n_prim = aos.num_ao_prim
        d.ao_nuc = _iptr(i32("ao_nuc", aos.nucleus_index))
        d.ao_pow = _iptr(
            i32("ao_pow", np.stack([aos.polynominal_order_x, aos.polynominal_order_y, aos.polynominal_order_z], axis=1))
        )
        d.ao_prim_ptr = _iptr(i32("ao_pp", np.concatenate([[0], np.cumsum(counts)])))
        d.prim_exp = _dptr(f64("prim_exp", aos.exponents[order]))
        d.prim_coef = _dptr(f64("prim_coef", aos._normalized_coefficients_np[order]))
        d.n_up = self.n_up
        d.n_dn = self.n_dn
        d.orb_is_mo = int(self.orb_is_mo)
        d.n_orb = self.n_orb
        if self.orb_is_mo:
            d.mo_coef_up = _dptr(f64("mo_up", gem.orb_data_up_spin.mo_coefficients))
            d.mo_coef_dn = _dptr(f64("mo_dn", gem.orb_data_dn_spin.mo_coefficients))
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
            d.j3_matrix = _dptr(f64("j3", j3.j_matrix))
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
        self._keep = keep
        self.desc = d
        self.handle = ctypes.c_void_p()
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


def system_for(ham: Hamiltonian_data, Nv: int = 6, NN: int = 1) -> SystemPack:
    """Device system pack for ``ham`` (cached per object identity, Nv and NN)."""
    key = (id(ham), Nv, NN)
    hit = _CACHE.get(key)
    if hit is not None and hit[0] is ham:
        return hit[1]
    pack = SystemPack(ham, Nv=Nv, NN=NN)
    _CACHE[key] = (ham, pack)
    return pack
