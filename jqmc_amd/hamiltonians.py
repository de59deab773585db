"""Hamiltonian container and local energy (``jqmc/hamiltonians.py``).

``Hamiltonian_data`` mirrors hamiltonians.py:82-177.  ``compute_local_energy``
(179-222) and ``compute_local_energy_fast`` (225-290) keep the reference signatures:
``e_L = T + V`` with ``V`` = bare Coulomb + ECP local + ECP non-local (NN nearest nuclei,
Nv-point quadrature rotated by ``RT``).  Both run the fused gfx950 local-energy kernel.
"""

from __future__ import annotations

from dataclasses import dataclass, field

from .coulomb_potential import Coulomb_potential_data
from .structure import Structure_data
from .wavefunction import Wavefunction_data


@dataclass
class Hamiltonian_data:
    structure_data: Structure_data = field(default_factory=Structure_data)
    coulomb_potential_data: Coulomb_potential_data = field(default_factory=Coulomb_potential_data)
    wavefunction_data: Wavefunction_data = field(default_factory=Wavefunction_data)

    def sanity_check(self) -> None:
        self.structure_data.sanity_check()
        self.coulomb_potential_data.sanity_check()
        self.wavefunction_data.sanity_check()


def compute_local_energy(hamiltonian_data: Hamiltonian_data, r_up_carts, r_dn_carts, RT, NN: int = 1, Nv: int = 6):
    """``e_L`` with a fresh LU inverse of G (hamiltonians.py:179-222)."""
    from ._api import local_energy

    return local_energy(hamiltonian_data, r_up_carts, r_dn_carts, RT, None, NN, Nv)


def compute_local_energy_fast(
    hamiltonian_data: Hamiltonian_data, r_up_carts, r_dn_carts, RT, geminal_inverse, NN: int = 1, Nv: int = 6
):
    """``e_L`` reusing the caller's ``G^-1`` (hamiltonians.py:225-290)."""
    from ._api import local_energy

    return local_energy(hamiltonian_data, r_up_carts, r_dn_carts, RT, geminal_inverse, NN, Nv)
