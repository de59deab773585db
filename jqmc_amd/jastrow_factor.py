"""One-, two- and three-body Jastrow factors (``jqmc/jastrow_factor.py``).

* ``Jastrow_one_body_data`` (jastrow_factor.py:560-645): ``exp`` form
  ``-(2Z)^(3/4) (1 - exp(-a (2Z)^(1/4) r)) / (2a)`` or Pade form
  ``-(2Z)^(3/4) r / (2 (1 + a (2Z)^(1/4) r))``, ``Z = Z_atom - core``.
* ``Jastrow_two_body_data`` (1110-1177): Pade ``r / (2 (1 + a r))`` or
  ``exp`` ``(1 - exp(-a r)) / (2a)``, the same form for parallel and antiparallel spins.
* ``Jastrow_three_body_data`` (1316-1531): ``j_matrix`` of shape ``(n_orb, n_orb + 1)``,
  square block ``J3`` plus last column ``j1``;
  ``J3 = sum_i j1.a_i + sum_{i<j same spin} a_i^T J3 a_j + sum_{i up, j dn} a_i^T J3 a_j``.

``compute_Jastrow_part`` (2127-2178) returns ``J`` without the exponential.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .atomic_orbital import AOs_cart_data
from .structure import Structure_data


@dataclass
class Jastrow_one_body_data:
    jastrow_1b_param: float = 1.0
    jastrow_1b_type: str = "exp"
    structure_data: Structure_data = field(default_factory=Structure_data)
    core_electrons: tuple = ()

    def __post_init__(self):
        self.jastrow_1b_param = float(self.jastrow_1b_param)
        self.core_electrons = tuple(float(c) for c in self.core_electrons)

    def sanity_check(self) -> None:
        if self.jastrow_1b_param < 0.0:
            raise ValueError("jastrow_1b_param must be non-negative.")
        if self.jastrow_1b_type not in ("exp", "pade"):
            raise ValueError("jastrow_1b_type must be 'exp' or 'pade'.")
        if len(self.core_electrons) != self.structure_data.natom:
            raise ValueError("len(core_electrons) must equal natom.")

    @property
    def effective_charges(self) -> np.ndarray:
        return np.asarray(self.structure_data.atomic_numbers, dtype=np.float64) - np.asarray(
            self.core_electrons, dtype=np.float64
        )


@dataclass
class Jastrow_two_body_data:
    jastrow_2b_param: float = 1.0
    jastrow_2b_type: str = "pade"

    def __post_init__(self):
        self.jastrow_2b_param = float(self.jastrow_2b_param)

    def sanity_check(self) -> None:
        if self.jastrow_2b_param < 0.0:
            raise ValueError("jastrow_2b_param must be non-negative.")
        if self.jastrow_2b_type not in ("exp", "pade"):
            raise ValueError("jastrow_2b_type must be 'exp' or 'pade'.")


@dataclass
class Jastrow_three_body_data:
    orb_data: object = field(default_factory=AOs_cart_data)
    j_matrix: np.ndarray = field(default_factory=lambda: np.zeros((0, 1)))

    def __post_init__(self):
        self.j_matrix = np.asarray(self.j_matrix, dtype=np.float64)

    @property
    def orb_num(self) -> int:
        return self.orb_data._num_orb

    def sanity_check(self) -> None:
        if self.j_matrix.shape != (self.orb_num, self.orb_num + 1):
            raise ValueError(f"dim. of j_matrix = {self.j_matrix.shape} is wrong.")


@dataclass
class Jastrow_data:
    jastrow_one_body_data: Jastrow_one_body_data | None = None
    jastrow_two_body_data: Jastrow_two_body_data | None = None
    jastrow_three_body_data: Jastrow_three_body_data | None = None
    jastrow_nn_data: object | None = None

    def sanity_check(self) -> None:
        if self.jastrow_nn_data is not None:
            raise NotImplementedError("the NN Jastrow is outside the MI355X hot path")
        for part in (self.jastrow_one_body_data, self.jastrow_two_body_data, self.jastrow_three_body_data):
            if part is not None:
                part.sanity_check()


def compute_Jastrow_part(jastrow_data: Jastrow_data, r_up_carts, r_dn_carts, geminal_data=None):
    """Total ``J = J1 + J2 + J3`` per walker (jastrow_factor.py:2127-2178)."""
    from ._api import jastrow_value

    return jastrow_value(jastrow_data, r_up_carts, r_dn_carts)


def _compute_ratio_Jastrow_part_rank1_update(jastrow_data: Jastrow_data, old_r_up_carts, old_r_dn_carts,
                                             new_r_up_carts_arr, new_r_dn_carts_arr, j3_state=None):
    """exp(J(r') - J(r)) for configurations that each move one electron (jastrow_factor.py:2230-2693),
    shapes as _compute_ratio_determinant_part_rank1_update.  The J3 fields of the reference
    configuration are the walker caches built from ``old_r_*_carts`` (``j3_state`` is not needed)."""
    from ._api import jastrow_ratios

    return jastrow_ratios(jastrow_data, old_r_up_carts, old_r_dn_carts, new_r_up_carts_arr, new_r_dn_carts_arr)


def _compute_ratio_Jastrow_part_split_spin(jastrow_data: Jastrow_data, old_r_up_carts, old_r_dn_carts,
                                           new_r_up_shifted, new_r_dn_shifted, j3_state=None):
    """Jastrow ratios of the block-structured mesh (jastrow_factor.py:2696-2960), up block first."""
    from ._api import jastrow_ratios, split_spin_configs

    ou, od, nu, nd = split_spin_configs(old_r_up_carts, old_r_dn_carts, new_r_up_shifted, new_r_dn_shifted)
    return jastrow_ratios(jastrow_data, ou, od, nu, nd)
