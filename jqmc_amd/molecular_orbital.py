"""Molecular orbitals ``Phi = C . AO`` (``jqmc/molecular_orbital.py``).

``MOs_data`` mirrors molecular_orbital.py:85-236; ``compute_MOs`` (239-261) and
``compute_MOs_value_grad_lap`` (375-415) keep the reference signatures and run on the
gfx950 AO kernel followed by the AO->MO contraction kernel.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .atomic_orbital import AOs_cart_data


@dataclass
class MOs_data:
    """MO coefficient matrix ``(num_mo, num_ao)`` over an AO basis (molecular_orbital.py:85-140)."""

    num_mo: int = 0
    aos_data: object = field(default_factory=AOs_cart_data)
    mo_coefficients: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))

    def __post_init__(self):
        self.mo_coefficients = np.asarray(self.mo_coefficients, dtype=np.float64)

    def sanity_check(self) -> None:
        if self.mo_coefficients.shape != (self.num_mo, self.aos_data.num_ao):
            raise ValueError(
                f"dim. of mo_coefficients = {self.mo_coefficients.shape} is wrong; "
                f"expected ({self.num_mo}, {self.aos_data.num_ao})."
            )
        self.aos_data.sanity_check()

    @property
    def _num_orb(self) -> int:
        return self.num_mo


def compute_MOs(mos_data: MOs_data, r_carts):
    """MO values ``(num_mo, N_e)`` (molecular_orbital.py:239-261)."""
    from ._api import mos_value_grad_lap

    return mos_value_grad_lap(mos_data, r_carts, value_only=True)[0]


def compute_MOs_value_grad_lap(mos_data: MOs_data, r_carts):
    """``(val, gx, gy, gz, lap)`` of the MOs, each ``(num_mo, N_e)`` (molecular_orbital.py:375-415)."""
    from ._api import mos_value_grad_lap

    return mos_value_grad_lap(mos_data, r_carts, value_only=False)
