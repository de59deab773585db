"""Bare Coulomb and ECP potential data (``jqmc/coulomb_potential.py``).

``Coulomb_potential_data`` mirrors coulomb_potential.py:187-598 (same field names).
The ECP radial term is ``V(r) = c r^(n-2) exp(-alpha r^2)`` with ``n`` the stored power
(TREXIO power + 2, trexio_wrapper.py:539).  A term is *local* when its ``ang_mom`` equals the
atom's ``max_ang_mom_plus_1`` (coulomb_potential.py:354-369); the others are projected on
``P_l`` with the quadrature meshes ``_Mesh`` (Nv = 4, 6, 12, 18; coulomb_potential.py:96-184).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np

from .structure import Structure_data


class _Mesh(NamedTuple):
    Nv: int
    weights: list
    grid_points: np.ndarray


def _tetrahedron_mesh() -> _Mesh:
    q = 1.0 / np.sqrt(3.0)
    pts = np.array([[q, q, q], [q, -q, -q], [-q, q, -q], [-q, -q, q]])
    return _Mesh(4, [0.25] * 4, pts)


def _octahedron_mesh() -> _Mesh:
    pts = np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, -1.0, 0], [0, 0, 1.0], [0, 0, -1.0]])
    return _Mesh(6, [1.0 / 6.0] * 6, pts)


def _icosahedron_mesh() -> _Mesh:
    """Two poles plus two staggered rings of five at polar angle atan(2) (coulomb_potential.py:126-153)."""
    t = np.arctan(2.0)
    sph = [(0.0, 0.0), (np.pi, 0.0)]
    sph += [(t, 2.0 * k * np.pi / 5.0) for k in range(5)]
    sph += [(np.pi - t, (2.0 * k + 1.0) * np.pi / 5.0) for k in range(5)]
    sph = np.array(sph)
    theta, phi = sph[:, 0], sph[:, 1]
    pts = np.vstack((np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta))).T
    return _Mesh(12, [1.0 / 12.0] * 12, pts)


def _octahedron18_mesh() -> _Mesh:
    p = 1.0 / np.sqrt(2.0)
    pts = [[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, -1.0, 0], [0, 0, 1.0], [0, 0, -1.0]]
    pts += [[p, p, 0], [p, -p, 0], [-p, p, 0], [-p, -p, 0]]
    pts += [[p, 0, p], [p, 0, -p], [-p, 0, p], [-p, 0, -p]]
    pts += [[0, p, p], [0, -p, p], [0, p, -p], [0, -p, -p]]
    return _Mesh(18, [1.0 / 6.0] * 6 + [1.0 / 15.0] * 12, np.array(pts))


def get_mesh(Nv: int) -> _Mesh:
    meshes = {4: _tetrahedron_mesh, 6: _octahedron_mesh, 12: _icosahedron_mesh, 18: _octahedron18_mesh}
    if Nv not in meshes:
        raise NotImplementedError(f"Nv={Nv} quadrature is not available (4, 6, 12, 18)")
    return meshes[Nv]()


@dataclass
class Coulomb_potential_data:
    structure_data: Structure_data = field(default_factory=Structure_data)
    ecp_flag: bool = False
    z_cores: tuple = ()
    max_ang_mom_plus_1: tuple = ()
    num_ecps: int = 0
    ang_moms: tuple = ()
    nucleus_index: tuple = ()
    exponents: tuple = ()
    coefficients: tuple = ()
    powers: tuple = ()

    def __post_init__(self):
        self.z_cores = tuple(float(z) for z in self.z_cores)
        self.max_ang_mom_plus_1 = tuple(int(v) for v in self.max_ang_mom_plus_1)
        self.ang_moms = tuple(int(v) for v in self.ang_moms)
        self.nucleus_index = tuple(int(v) for v in self.nucleus_index)
        self.exponents = tuple(float(v) for v in self.exponents)
        self.coefficients = tuple(float(v) for v in self.coefficients)
        self.powers = tuple(float(v) for v in self.powers)

    def sanity_check(self) -> None:
        """Same checks as coulomb_potential.py:246-291."""
        if self.ecp_flag:
            n = self.structure_data.natom
            if len(self.z_cores) != n or len(self.max_ang_mom_plus_1) != n:
                raise ValueError("dim. of z_cores / max_ang_mom_plus_1 is wrong")
            for seq in (self.ang_moms, self.nucleus_index, self.exponents, self.coefficients, self.powers):
                if len(seq) != self.num_ecps:
                    raise ValueError("dim. of ECP term arrays is wrong")
        else:
            for seq in (self.z_cores, self.max_ang_mom_plus_1, self.ang_moms, self.nucleus_index):
                if len(seq) != 0:
                    raise ValueError("ECP arrays must be empty when ecp_flag is False")
        self.structure_data.sanity_check()

    @property
    def _effective_charges(self) -> np.ndarray:
        """Z (all-electron) or Z - z_core (ECP), coulomb_potential.py:299-309."""
        Z = np.asarray(self.structure_data.atomic_numbers, dtype=np.float64)
        if self.ecp_flag:
            return Z - np.asarray(self.z_cores, dtype=np.float64)
        return Z

    @property
    def _global_max_ang_mom_plus_1(self) -> int:
        return int(np.max(self.max_ang_mom_plus_1)) if self.ecp_flag else 0

    def _local_terms(self):
        """(atom, exponent, coefficient, power) of the local terms, coulomb_potential.py:339-353."""
        out = []
        for i, a in enumerate(self.nucleus_index):
            if self.ang_moms[i] == self.max_ang_mom_plus_1[a]:
                out.append((a, self.exponents[i], self.coefficients[i], self.powers[i]))
        return out

    def _non_local_terms(self):
        """(atom, l, exponent, coefficient, power) of the projected terms, coulomb_potential.py:355-369."""
        out = []
        for i, a in enumerate(self.nucleus_index):
            if self.ang_moms[i] != self.max_ang_mom_plus_1[a]:
                out.append((a, self.ang_moms[i], self.exponents[i], self.coefficients[i], self.powers[i]))
        return out


def compute_coulomb_potential(
    coulomb_potential_data, r_up_carts, r_dn_carts, RT, NN: int = 1, Nv: int = 6, wavefunction_data=None
):
    """Bare Coulomb + ECP (local + non-local) per walker (coulomb_potential.py:2643-2692).

    The GPU path evaluates the non-local ratios with the LU inverse of G built in the same call.
    """
    from ._api import coulomb_potential

    return coulomb_potential(coulomb_potential_data, wavefunction_data, r_up_carts, r_dn_carts, RT, NN, Nv)
