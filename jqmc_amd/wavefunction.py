"""Trial wavefunction ``Psi = exp(J) det G`` (``jqmc/wavefunction.py``).

``Wavefunction_data`` mirrors wavefunction.py:247-675.  ``evaluate_ln_wavefunction``
(677-720) and ``compute_kinetic_energy_all_elements_fast_update`` (1141-1207) keep the
reference signatures; the kinetic term is ``T_i = -1/2 (lap_i ln Psi + |grad_i ln Psi|^2)``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

from .determinant import Geminal_data
from .jastrow_factor import Jastrow_data


@dataclass
class Wavefunction_data:
    jastrow_data: Jastrow_data = field(default_factory=Jastrow_data)
    geminal_data: Geminal_data = field(default_factory=Geminal_data)

    def sanity_check(self) -> None:
        self.jastrow_data.sanity_check()
        self.geminal_data.sanity_check()


def evaluate_ln_wavefunction(wavefunction_data: Wavefunction_data, r_up_carts, r_dn_carts):
    """``J + ln|det G|`` per walker (wavefunction.py:677-720)."""
    from ._api import ln_wavefunction

    return ln_wavefunction(wavefunction_data, r_up_carts, r_dn_carts)


def compute_kinetic_energy_all_elements_fast_update(
    wavefunction_data: Wavefunction_data, r_up_carts, r_dn_carts, geminal_inverse
):
    """Per-electron kinetic energies ``(T_up, T_dn)`` given ``G^-1`` (wavefunction.py:1141-1207)."""
    from ._api import kinetic_all_elements

    return kinetic_all_elements(wavefunction_data, r_up_carts, r_dn_carts, geminal_inverse)
