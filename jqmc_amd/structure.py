"""Molecular structure container and nearest-nucleus helpers.

Mirrors ``jqmc/structure.py``: ``Structure_data`` (structure.py:65-357) with the same
field names and meaning (positions in bohr, fp64), and the nearest-nucleus search
``_find_nearest_nucleus_indices_np`` (structure.py:392-408) plus the relative vector
``_get_min_dist_rel_R_cart_np`` (structure.py:429-452), which returns ``R - r``.
Periodic boundary conditions are outside the hot path covered here; constructing a
structure with ``pbc_flag=True`` raises ``NotImplementedError``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Structure_data:
    """Atomic positions (bohr), atomic numbers and labels (structure.py:65-114)."""

    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    pbc_flag: bool = False
    vec_a: tuple = ()
    vec_b: tuple = ()
    vec_c: tuple = ()
    atomic_numbers: tuple = ()
    element_symbols: tuple = ()
    atomic_labels: tuple = ()

    def __post_init__(self):
        self.positions = np.asarray(self.positions, dtype=np.float64).reshape(-1, 3)
        self.atomic_numbers = tuple(int(z) for z in self.atomic_numbers)
        self.element_symbols = tuple(self.element_symbols)
        self.atomic_labels = tuple(self.atomic_labels)
        if self.pbc_flag:
            raise NotImplementedError("periodic systems are not supported by the MI355X hot path")

    def sanity_check(self) -> None:
        """Same checks as structure.py:116-155 (PBC branch excluded)."""
        n = len(self.atomic_numbers)
        if len(self.element_symbols) != n:
            raise ValueError("The length of element_symbols and atomic_numbers must be the same.")
        if len(self.atomic_labels) != n:
            raise ValueError("The length of atomic_labels and atomic_numbers must be the same.")
        if self.positions.shape != (n, 3):
            raise ValueError("positions must have shape (natom, 3).")

    @property
    def natom(self) -> int:
        return len(self.atomic_numbers)

    @property
    def _positions_cart_np(self) -> np.ndarray:
        return self.positions


def _find_nearest_nucleus_indices_np(structure_data: Structure_data, r_cart, N: int) -> np.ndarray:
    """Indices of the N nearest nuclei (stable argsort of distances), structure.py:392-408."""
    diffs = structure_data.positions - np.asarray(r_cart, dtype=np.float64)
    distances = np.sqrt(np.sum(diffs**2, axis=1))
    return np.argsort(distances, kind="stable")[:N]


def _find_nearest_index_np(structure_data: Structure_data, r_cart) -> int:
    """Index of the nearest nucleus, structure.py:360-373."""
    return int(_find_nearest_nucleus_indices_np(structure_data, r_cart, 1)[0])


def _get_min_dist_rel_R_cart_np(structure_data: Structure_data, r_cart, i_atom: int) -> np.ndarray:
    """``R_i_atom - r`` (note the sign), structure.py:429-452."""
    return structure_data.positions[i_atom] - np.asarray(r_cart, dtype=np.float64)
