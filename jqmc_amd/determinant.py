"""Geminal (AGP / Slater) determinant part (``jqmc/determinant.py``).

``Geminal_data`` mirrors determinant.py:93-200: ``lambda_matrix`` has shape
``(orb_num_up, orb_num_dn + N_up - N_dn)`` = ``[Lambda_paired | Lambda_unpaired]`` and the
geminal matrix is ``G = [Phi_up^T Lambda_p Phi_dn | Phi_up^T Lambda_u]`` (N_up x N_up),
determinant.py:1379-1402.  The GPU-backed functions keep the reference names:
``compute_geminal_all_elements`` (1336-1376), ``compute_det_geminal_all_elements``
(1185-1204), ``compute_ln_det_geminal_all_elements`` (1026-1048) and
``compute_AS_regularization_factor`` (1262-1333, evaluated with the LU inverse).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .atomic_orbital import AOs_cart_data, AOs_sphe_data
from .molecular_orbital import MOs_data


@dataclass
class Geminal_data:
    """Pairing matrix and orbital references for the two spin channels."""

    num_electron_up: int = 0
    num_electron_dn: int = 0
    orb_data_up_spin: object = field(default_factory=AOs_cart_data)
    orb_data_dn_spin: object = field(default_factory=AOs_cart_data)
    lambda_matrix: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))

    def __post_init__(self):
        self.lambda_matrix = np.asarray(self.lambda_matrix, dtype=np.float64)

    @property
    def orb_num_up(self) -> int:
        return self.orb_data_up_spin._num_orb

    @property
    def orb_num_dn(self) -> int:
        return self.orb_data_dn_spin._num_orb

    @property
    def aos_data(self):
        """The AO basis beneath the orbitals (the same object for both spins)."""
        o = self.orb_data_up_spin
        return o.aos_data if isinstance(o, MOs_data) else o

    def sanity_check(self) -> None:
        """Same checks as determinant.py:122-149."""
        if self.orb_num_up != self.orb_num_dn:
            raise ValueError("The number of up and down orbitals should be the same such that the lambda_matrix is square.")
        expected = (self.orb_num_up, self.orb_num_dn + (self.num_electron_up - self.num_electron_dn))
        if self.lambda_matrix.shape != expected:
            raise ValueError(f"dim. of lambda_matrix = {self.lambda_matrix.shape} is imcompatible with {expected}.")
        if self.num_electron_up < self.num_electron_dn:
            raise ValueError("N_up must be >= N_dn.")
        if type(self.orb_data_up_spin) is not type(self.orb_data_dn_spin):
            raise ValueError("orbital data of both spins must be of the same kind")
        self.orb_data_up_spin.sanity_check()
        self.orb_data_dn_spin.sanity_check()

    @property
    def lambda_paired(self) -> np.ndarray:
        return self.lambda_matrix[:, : self.orb_num_dn]

    @property
    def lambda_unpaired(self) -> np.ndarray:
        return self.lambda_matrix[:, self.orb_num_dn :]

    @property
    def is_ao_based(self) -> bool:
        return isinstance(self.orb_data_up_spin, (AOs_cart_data, AOs_sphe_data))


def compute_geminal_all_elements(geminal_data: Geminal_data, r_up_carts, r_dn_carts):
    """Geminal matrix ``(N_up, N_up)`` (or batched ``(W, N_up, N_up)``), determinant.py:1336-1402."""
    from ._api import geminal_matrix

    return geminal_matrix(geminal_data, r_up_carts, r_dn_carts)[0]


def compute_det_geminal_all_elements(geminal_data: Geminal_data, r_up_carts, r_dn_carts):
    """``det G`` via a pivoted FP64 LU on the GPU (determinant.py:1185-1204)."""
    from ._api import geminal_matrix

    _, sign, logabs, _ = geminal_matrix(geminal_data, r_up_carts, r_dn_carts, with_inverse=True)
    import torch

    return sign * torch.exp(logabs)


def compute_ln_det_geminal_all_elements(geminal_data: Geminal_data, r_up_carts, r_dn_carts):
    """``ln |det G|`` (determinant.py:1026-1048)."""
    from ._api import geminal_matrix

    return geminal_matrix(geminal_data, r_up_carts, r_dn_carts, with_inverse=True)[2]


def compute_AS_regularization_factor(geminal_data: Geminal_data, r_up_carts, r_dn_carts):
    """Attaccalite-Sorella factor ``(S * ||G^-1||_F^2)^(-3/8)`` (determinant.py:1262-1333)."""
    from ._api import geminal_matrix

    return geminal_matrix(geminal_data, r_up_carts, r_dn_carts, with_inverse=True)[3]


def _compute_ratio_determinant_part_rank1_update(geminal_data: Geminal_data, A_old_inv, old_r_up_carts,
                                                 old_r_dn_carts, new_r_up_carts_arr, new_r_dn_carts_arr):
    """det G(r') / det G(r) for configurations that each move one electron (determinant.py:1529-1662):
    ``new_r_*_carts_arr`` of shape ``(N_grid, N_sigma, 3)`` (or ``(W, N_grid, N_sigma, 3)`` with a walker
    axis on every argument), the moved electron found by the first changed coordinate as in the
    reference, the ratio from ``A_old_inv`` and the rank-1 row / column (``jqmc_point_ratios``)."""
    from ._api import determinant_ratios

    return determinant_ratios(geminal_data, A_old_inv, old_r_up_carts, old_r_dn_carts, new_r_up_carts_arr,
                              new_r_dn_carts_arr)


def _compute_ratio_determinant_part_split_spin(geminal_data: Geminal_data, A_old_inv, old_r_up_carts,
                                               old_r_dn_carts, new_r_up_shifted, new_r_dn_shifted):
    """Ratios of the block-structured mesh (determinant.py:1665-1783): ``new_r_up_shifted``
    ``(G_up, N_up, 3)`` moves one up electron per configuration, ``new_r_dn_shifted`` ``(G_dn, N_dn, 3)``
    one dn electron; returns the ``(G_up + G_dn,)`` ratios, up block first."""
    from ._api import determinant_ratios, split_spin_configs

    ou, od, nu, nd = split_spin_configs(old_r_up_carts, old_r_dn_carts, new_r_up_shifted, new_r_dn_shifted)
    return determinant_ratios(geminal_data, A_old_inv, ou, od, nu, nd)
