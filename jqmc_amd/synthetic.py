"""Seeded synthetic systems for the benchmark and parity workloads.

Every system is built from a ``numpy.random.default_rng(seed)`` stream so the GPU path,
the CPU oracle and the benchmark see identical inputs.  Basis sets are Cartesian
contracted GTOs with ``n_prim`` primitives per shell (exponents log-uniform in
``[0.1, 50]``, contraction coefficients ``U(0.1, 1)``), duplicated per Cartesian
component in the order used by the reference (``xx, xy, xz, yy, yz, zz`` for d,
atomic_orbital.py:155-180).  MOs are orthonormal rows from a QR factorisation of an
``N(0,1)`` matrix; the JSD pairing matrix is the identity-padded ``lambda`` of
trexio_wrapper.py:470-478.

Workloads:

* ``water_ae``: O (Z=8) at the origin, H at (+-1.43, 1.11, 0); O 3s2p1d + H 2s1p = 25 AOs;
  10 electrons (5 up / 5 dn); J1 exp (a=1), J2 Pade (a=1), J3 ~ N(0, 0.01).
* ``benzene_ecp``: 6 C on a ring of radius 2.64 bohr, 6 H at radius 4.68 bohr (z=0);
  synthetic ECP on C (z_core=2; local l=2 terms plus projected l=0,1 terms);
  C 4s4p2d + H 4s2p = 228 AOs; 30 valence electrons (15/15).
* ``water4_ecp``: four waters with O on a 5.5-bohr square, seeded orientations, ECP on O;
  O 5s4p3d1f + H 3s2p1d = 300 AOs; 32 valence electrons (16/16).
* ``chain_ecp(n_heavy, ...)``: linear/ring molecules for uneven LRDMC batching.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .atomic_orbital import AOs_cart_data
from .coulomb_potential import Coulomb_potential_data
from .determinant import Geminal_data
from .hamiltonians import Hamiltonian_data
from .jastrow_factor import Jastrow_data, Jastrow_one_body_data, Jastrow_three_body_data, Jastrow_two_body_data
from .molecular_orbital import MOs_data
from .structure import Structure_data
from .wavefunction import Wavefunction_data

_L_OF = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}
_SYMBOL_Z = {"H": 1, "C": 6, "N": 7, "O": 8}


def cartesian_components(l: int):
    """(nx, ny, nz) with nx+ny+nz=l, nx descending then ny descending (xx, xy, xz, yy, yz, zz for l=2)."""
    out = []
    for nx in range(l, -1, -1):
        for ny in range(l - nx, -1, -1):
            out.append((nx, ny, l - nx - ny))
    return out


def parse_shells(spec: str):
    """'4s3p2d' -> [0,0,0,0,1,1,1,2,2]."""
    shells = []
    num = ""
    for ch in spec:
        if ch.isdigit():
            num += ch
        else:
            shells += [_L_OF[ch]] * int(num)
            num = ""
    return shells


def build_cart_basis(structure: Structure_data, shell_spec: dict, rng, n_prim: int = 3, exp_range=(0.1, 50.0)):
    """Cartesian AO basis with seeded exponents/coefficients, one shell list per element."""
    nucleus_index, orbital_indices, exps, coefs, ang, px, py, pz = [], [], [], [], [], [], [], []
    ao = 0
    lo, hi = np.log(exp_range[0]), np.log(exp_range[1])
    for ia, sym in enumerate(structure.element_symbols):
        for l in parse_shells(shell_spec[sym]):
            z = np.exp(rng.uniform(lo, hi, size=n_prim))
            c = rng.uniform(0.1, 1.0, size=n_prim)
            for nx, ny, nz in cartesian_components(l):
                nucleus_index.append(ia)
                ang.append(l)
                px.append(nx)
                py.append(ny)
                pz.append(nz)
                for p in range(n_prim):
                    orbital_indices.append(ao)
                    exps.append(z[p])
                    coefs.append(c[p])
                ao += 1
    return AOs_cart_data(
        structure_data=structure,
        nucleus_index=nucleus_index,
        num_ao=ao,
        num_ao_prim=len(exps),
        orbital_indices=orbital_indices,
        exponents=np.array(exps),
        coefficients=np.array(coefs),
        angular_momentums=ang,
        polynominal_order_x=px,
        polynominal_order_y=py,
        polynominal_order_z=pz,
    )


def random_orthonormal_rows(rng, n_rows: int, n_cols: int) -> np.ndarray:
    """Rows of Q^T from QR of an N(0,1) (n_cols x n_rows) matrix: orthonormal MO coefficients."""
    q, _ = np.linalg.qr(rng.standard_normal((n_cols, n_rows)))
    return np.ascontiguousarray(q.T)


def jsd_lambda(n_mo: int, n_up: int, n_dn: int) -> np.ndarray:
    """Identity-padded pairing matrix of trexio_wrapper.py:470-478."""
    paired = np.pad(np.eye(n_dn), ((0, n_mo - n_dn), (0, n_mo - n_dn)))
    diff = n_up - n_dn
    unpaired = np.pad(np.eye(diff), ((n_dn, n_mo - n_dn - diff), (0, 0)))
    return np.hstack([paired, unpaired])


def synthetic_ecp(structure: Structure_data, ecp_elements: dict, rng):
    """Seeded ECP: per ECP element, local (l = lmax+1 = 2) terms c r^(n-2) e^(-a r^2) with
    n = 1, 3, 2, plus projected l = 0 and l = 1 terms with n = 2 (ccECP-like shapes)."""
    n = structure.natom
    z_cores, lmax1 = [], []
    ang, nuc, ex, co, pw = [], [], [], [], []
    params = {}
    for sym, zc in ecp_elements.items():
        zeff = _SYMBOL_Z[sym] - zc
        params[sym] = [
            (2, rng.uniform(6.0, 10.0), float(zeff), 1.0),
            (2, rng.uniform(3.0, 6.0), rng.uniform(20.0, 40.0), 3.0),
            (2, rng.uniform(3.0, 6.0), -rng.uniform(10.0, 25.0), 2.0),
            (0, rng.uniform(3.0, 8.0), rng.uniform(5.0, 25.0), 2.0),
            (1, rng.uniform(3.0, 8.0), rng.uniform(2.0, 10.0), 2.0),
        ]
    for ia, sym in enumerate(structure.element_symbols):
        if sym in ecp_elements:
            z_cores.append(float(ecp_elements[sym]))
            lmax1.append(2)
            for l, a, c, p in params[sym]:
                ang.append(l)
                nuc.append(ia)
                ex.append(a)
                co.append(c)
                pw.append(p)
        else:
            z_cores.append(0.0)
            lmax1.append(0)
    return Coulomb_potential_data(
        structure_data=structure,
        ecp_flag=True,
        z_cores=z_cores,
        max_ang_mom_plus_1=lmax1,
        num_ecps=len(ang),
        ang_moms=ang,
        nucleus_index=nuc,
        exponents=ex,
        coefficients=co,
        powers=pw,
    )


@dataclass
class SyntheticSystem:
    name: str
    hamiltonian_data: Hamiltonian_data
    electron_owner_up: np.ndarray  # nucleus index each up electron is initialised around
    electron_owner_dn: np.ndarray
    seed: int

    @property
    def n_up(self) -> int:
        return self.hamiltonian_data.wavefunction_data.geminal_data.num_electron_up

    @property
    def n_dn(self) -> int:
        return self.hamiltonian_data.wavefunction_data.geminal_data.num_electron_dn

    def initial_walkers(self, n_walkers: int, seed: int | None = None, sigma: float = 1.0):
        """Electrons drawn from N(owner nucleus, sigma^2) per Cartesian axis; shapes (W,N_up,3), (W,N_dn,3)."""
        rng = np.random.default_rng(self.seed + 1 if seed is None else seed)
        pos = self.hamiltonian_data.structure_data.positions
        r_up = pos[self.electron_owner_up][None] + sigma * rng.standard_normal((n_walkers, self.n_up, 3))
        r_dn = pos[self.electron_owner_dn][None] + sigma * rng.standard_normal((n_walkers, self.n_dn, 3))
        return np.ascontiguousarray(r_up), np.ascontiguousarray(r_dn)


def _assign_electrons(structure: Structure_data, valence: list):
    """Alternate spins over the valence electrons of each nucleus, balancing odd counts globally."""
    up, dn = [], []
    flip = 0
    for ia, nv in enumerate(valence):
        nv = int(round(nv))
        for k in range(nv):
            if k < nv - nv % 2:
                (up if k % 2 == 0 else dn).append(ia)
            else:
                (up if flip % 2 == 0 else dn).append(ia)
                flip += 1
    return np.array(up, dtype=np.int64), np.array(dn, dtype=np.int64)


def _assemble(name, structure, shells, seed, ecp_elements=None, jagp=False, j3_sigma=0.01, lam_sigma=0.05, lam_rank=16,
              j1=("exp", 1.0), j2=("pade", 1.0), use_j3=True, j3_symmetric=False):
    rng = np.random.default_rng(seed)
    aos = build_cart_basis(structure, shells, rng)
    if ecp_elements:
        cp = synthetic_ecp(structure, ecp_elements, rng)
        core = list(cp.z_cores)
    else:
        cp = Coulomb_potential_data(structure_data=structure, ecp_flag=False)
        core = [0.0] * structure.natom
    valence = [z - c for z, c in zip(structure.atomic_numbers, core)]
    owners_up, owners_dn = _assign_electrons(structure, valence)
    n_up, n_dn = len(owners_up), len(owners_dn)
    if n_up < n_dn:
        owners_up, owners_dn = owners_dn, owners_up
        n_up, n_dn = n_dn, n_up
    if jagp:
        lam_p = rng.normal(0.0, lam_sigma, size=(aos.num_ao, aos.num_ao))
        lam_p = 0.5 * (lam_p + lam_p.T)
        q = random_orthonormal_rows(rng, lam_rank, aos.num_ao)
        lam_p = lam_p + q.T @ np.diag(np.linspace(2.0, 1.0, lam_rank)) @ q
        lam_u = rng.normal(0.0, lam_sigma, size=(aos.num_ao, n_up - n_dn))
        geminal = Geminal_data(n_up, n_dn, aos, aos, np.hstack([lam_p, lam_u]))
    else:
        C = random_orthonormal_rows(rng, n_up, aos.num_ao)
        mos = MOs_data(num_mo=n_up, aos_data=aos, mo_coefficients=C)
        geminal = Geminal_data(n_up, n_dn, mos, mos, jsd_lambda(n_up, n_up, n_dn))
    j1d = Jastrow_one_body_data(jastrow_1b_param=j1[1], jastrow_1b_type=j1[0], structure_data=structure,
                                core_electrons=core) if j1 else None
    j2d = Jastrow_two_body_data(jastrow_2b_param=j2[1], jastrow_2b_type=j2[0]) if j2 else None
    j3d = None
    if use_j3:
        jm = rng.normal(0.0, j3_sigma, size=(aos.num_ao, aos.num_ao + 1))
        if j3_symmetric:
            # the symmetric pair matrix the reference keeps under optimisation (jastrow_factor.py:2040-2086)
            jm[:, :-1] = 0.5 * (jm[:, :-1] + jm[:, :-1].T)
        j3d = Jastrow_three_body_data(orb_data=aos, j_matrix=jm)
    jas = Jastrow_data(j1d, j2d, j3d)
    wf = Wavefunction_data(jastrow_data=jas, geminal_data=geminal)
    ham = Hamiltonian_data(structure_data=structure, coulomb_potential_data=cp, wavefunction_data=wf)
    ham.sanity_check()
    return SyntheticSystem(name, ham, owners_up, owners_dn, seed)


def _structure(symbols, positions):
    return Structure_data(
        positions=np.asarray(positions, dtype=np.float64),
        atomic_numbers=[_SYMBOL_Z[s] for s in symbols],
        element_symbols=list(symbols),
        atomic_labels=[f"{s}{i}" for i, s in enumerate(symbols)],
    )


def water_ae(seed: int = 0, **kw) -> SyntheticSystem:
    """All-electron water: O 3s2p1d, H 2s1p (25 Cartesian AOs), 5 up / 5 dn."""
    st = _structure(["O", "H", "H"], [[0.0, 0.0, 0.0], [1.43, 1.11, 0.0], [-1.43, 1.11, 0.0]])
    return _assemble("water_ae", st, {"O": "3s2p1d", "H": "2s1p"}, seed, **kw)


def benzene_ecp(seed: int = 0, **kw) -> SyntheticSystem:
    """Benzene-like ring with ECP on C: C 4s4p2d + H 4s2p (228 AOs), 15 up / 15 dn."""
    ang = np.arange(6) * np.pi / 3.0
    pos = [[2.64 * np.cos(a), 2.64 * np.sin(a), 0.0] for a in ang] + [[4.68 * np.cos(a), 4.68 * np.sin(a), 0.0] for a in ang]
    st = _structure(["C"] * 6 + ["H"] * 6, pos)
    return _assemble("benzene_ecp", st, {"C": "4s4p2d", "H": "4s2p"}, seed, ecp_elements={"C": 2}, **kw)


def water4_ecp(seed: int = 0, jagp: bool = False, **kw) -> SyntheticSystem:
    """Four waters, O on a 5.5-bohr square, seeded orientations, ECP on O: 300 AOs, 16/16."""
    rng = np.random.default_rng(seed + 1000)
    symbols, pos = [], []
    base_h = np.array([[1.43, 1.11, 0.0], [-1.43, 1.11, 0.0]])
    for cx, cy in [(-2.75, -2.75), (2.75, -2.75), (2.75, 2.75), (-2.75, 2.75)]:
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))
        o = np.array([cx, cy, 0.0])
        symbols += ["O", "H", "H"]
        pos += [o, o + base_h[0] @ q.T, o + base_h[1] @ q.T]
    st = _structure(symbols, pos)
    name = "water4_jagp" if jagp else "water4_ecp"
    return _assemble(name, st, {"O": "5s4p3d1f", "H": "3s2p1d"}, seed, ecp_elements={"O": 2}, jagp=jagp, **kw)


def chain_ecp(n_valence: int, seed: int = 0, **kw) -> SyntheticSystem:
    """Uneven-size molecules for batched LRDMC (bond 2.6 bohr, ECP on heavy atoms).

    10 e: N2 (N 3s3p, 24 AOs); 30 e: C6H6 ring (C 3s2p1d, H 2s1p); 56 e: C14 ring (C 3s3p2d);
    96 e: C24 ring (C 3s2p2d).
    """
    bond = 2.6
    if n_valence == 10:
        st = _structure(["N", "N"], [[0, 0, -bond / 2], [0, 0, bond / 2]])
        return _assemble("n2_ecp", st, {"N": "3s3p"}, seed, ecp_elements={"N": 2}, **kw)
    if n_valence == 30:
        ang = np.arange(6) * np.pi / 3.0
        pos = [[bond * np.cos(a), bond * np.sin(a), 0] for a in ang] + [
            [(bond + 2.0) * np.cos(a), (bond + 2.0) * np.sin(a), 0] for a in ang]
        st = _structure(["C"] * 6 + ["H"] * 6, pos)
        return _assemble("c6h6_ecp", st, {"C": "3s2p1d", "H": "2s1p"}, seed, ecp_elements={"C": 2}, **kw)
    n_c = {56: 14, 96: 24}.get(n_valence)
    if n_c is None:
        raise ValueError("n_valence must be one of 10, 30, 56, 96")
    rad = bond / (2.0 * np.sin(np.pi / n_c))
    ang = np.arange(n_c) * 2.0 * np.pi / n_c
    st = _structure(["C"] * n_c, [[rad * np.cos(a), rad * np.sin(a), 0.0] for a in ang])
    shells = {"C": "3s3p2d"} if n_c == 14 else {"C": "3s2p2d"}
    return _assemble(f"c{n_c}_ecp", st, shells, seed, ecp_elements={"C": 2}, **kw)


WORKLOADS = {
    "water_ae": water_ae,
    "benzene_ecp": benzene_ecp,
    "water4_ecp": water4_ecp,
}
