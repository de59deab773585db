"""Contracted Gaussian atomic orbitals: data classes and GPU-backed evaluators.

Data classes keep the field names of ``jqmc/atomic_orbital.py``:
``AOs_cart_data`` (atomic_orbital.py:87-777) and ``AOs_sphe_data`` (780-1347).
The evaluators ``compute_AOs`` (1886-1916) and ``compute_AOs_value_grad_lap``
(3594-3640) keep the reference's signature ``(aos_data, r_carts) -> (num_ao, N_e)``;
``r_carts`` may also carry a leading walker axis ``(W, N_e, 3)``, in which case the
outputs carry it too.  Both run the hand-written gfx950 kernel
``jqmc_aos_value_grad_lap`` through the C ABI and raise if the HIP library is absent.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from math import factorial

import numpy as np

from .structure import Structure_data


def _tuple_int(x):
    return tuple(int(v) for v in x)


@dataclass
class AOs_cart_data:
    """Cartesian contracted GTOs, one primitive list per AO (atomic_orbital.py:87-258).

    ``phi_a(r) = x^nx y^ny z^nz * sum_p N_p c_p exp(-Z_p |r-R|^2)`` with
    ``N_p = sqrt((2Z/pi)^(3/2) (8Z)^l nx! ny! nz! / ((2nx)! (2ny)! (2nz)!))``
    (atomic_orbital.py:2237-2246, 736-756).
    """

    structure_data: Structure_data = field(default_factory=Structure_data)
    nucleus_index: tuple = ()
    num_ao: int = 0
    num_ao_prim: int = 0
    orbital_indices: tuple = ()
    exponents: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coefficients: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angular_momentums: tuple = ()
    polynominal_order_x: tuple = ()
    polynominal_order_y: tuple = ()
    polynominal_order_z: tuple = ()

    def __post_init__(self):
        self.nucleus_index = _tuple_int(self.nucleus_index)
        self.orbital_indices = _tuple_int(self.orbital_indices)
        self.exponents = np.asarray(self.exponents, dtype=np.float64)
        self.coefficients = np.asarray(self.coefficients, dtype=np.float64)
        self.angular_momentums = _tuple_int(self.angular_momentums)
        self.polynominal_order_x = _tuple_int(self.polynominal_order_x)
        self.polynominal_order_y = _tuple_int(self.polynominal_order_y)
        self.polynominal_order_z = _tuple_int(self.polynominal_order_z)

    def sanity_check(self) -> None:
        if len(self.nucleus_index) != self.num_ao:
            raise ValueError("dim. of nucleus_index is wrong")
        if len(self.orbital_indices) != self.num_ao_prim:
            raise ValueError("dim. of orbital_indices is wrong")
        if len(self.exponents) != self.num_ao_prim or len(self.coefficients) != self.num_ao_prim:
            raise ValueError("dim. of exponents/coefficients is wrong")
        for seq in (self.angular_momentums, self.polynominal_order_x, self.polynominal_order_y, self.polynominal_order_z):
            if len(seq) != self.num_ao:
                raise ValueError("dim. of angular data is wrong")
        for l, nx, ny, nz in zip(
            self.angular_momentums, self.polynominal_order_x, self.polynominal_order_y, self.polynominal_order_z
        ):
            if nx + ny + nz != l:
                raise ValueError("nx + ny + nz must equal l for every Cartesian AO")
        self.structure_data.sanity_check()

    @property
    def _num_orb(self) -> int:
        return self.num_ao

    @property
    def _atomic_center_carts_np(self) -> np.ndarray:
        return self.structure_data.positions[np.asarray(self.nucleus_index, dtype=np.int64)]

    @property
    def _normalization_factorial_ratio_prim_np(self) -> np.ndarray:
        """nx! ny! nz! / ((2nx)! (2ny)! (2nz)!) per primitive, exact integers (atomic_orbital.py:736-756)."""
        out = np.empty(self.num_ao_prim)
        for p, ao in enumerate(self.orbital_indices):
            nx, ny, nz = self.polynominal_order_x[ao], self.polynominal_order_y[ao], self.polynominal_order_z[ao]
            num = factorial(nx) * factorial(ny) * factorial(nz)
            den = factorial(2 * nx) * factorial(2 * ny) * factorial(2 * nz)
            out[p] = np.float64(num / den)
        return out

    @property
    def _normalized_coefficients_np(self) -> np.ndarray:
        """N_p * c_p per primitive, evaluated in the reference's operand order (atomic_orbital.py:2244-2251)."""
        Z = self.exponents
        l = np.asarray([self.angular_momentums[ao] for ao in self.orbital_indices], dtype=np.int64)
        NZ = (2.0 * Z / np.pi) ** (3.0 / 2.0) * np.array([(8.0 * z) ** int(li) for z, li in zip(Z, l)])
        N = np.sqrt(NZ * self._normalization_factorial_ratio_prim_np)
        return N * self.coefficients


@dataclass
class AOs_sphe_data:
    """Real-solid-harmonic contracted GTOs (atomic_orbital.py:780-1347)."""

    structure_data: Structure_data = field(default_factory=Structure_data)
    nucleus_index: tuple = ()
    num_ao: int = 0
    num_ao_prim: int = 0
    orbital_indices: tuple = ()
    exponents: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coefficients: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angular_momentums: tuple = ()
    magnetic_quantum_numbers: tuple = ()

    def __post_init__(self):
        self.nucleus_index = _tuple_int(self.nucleus_index)
        self.orbital_indices = _tuple_int(self.orbital_indices)
        self.exponents = np.asarray(self.exponents, dtype=np.float64)
        self.coefficients = np.asarray(self.coefficients, dtype=np.float64)
        self.angular_momentums = _tuple_int(self.angular_momentums)
        self.magnetic_quantum_numbers = _tuple_int(self.magnetic_quantum_numbers)

    def sanity_check(self) -> None:
        if len(self.nucleus_index) != self.num_ao:
            raise ValueError("dim. of nucleus_index is wrong")
        if len(self.orbital_indices) != self.num_ao_prim:
            raise ValueError("dim. of orbital_indices is wrong")
        if len(self.angular_momentums) != self.num_ao or len(self.magnetic_quantum_numbers) != self.num_ao:
            raise ValueError("dim. of angular data is wrong")
        if max(self.angular_momentums, default=0) > 6:
            raise NotImplementedError("real solid harmonics are tabulated up to l=6")
        self.structure_data.sanity_check()

    @property
    def _num_orb(self) -> int:
        return self.num_ao

    @property
    def _atomic_center_carts_np(self) -> np.ndarray:
        return self.structure_data.positions[np.asarray(self.nucleus_index, dtype=np.int64)]


def compute_AOs(aos_data, r_carts):
    """AO values, shape ``(num_ao, N_e)`` (or ``(W, num_ao, N_e)``); atomic_orbital.py:1886-1916."""
    from ._api import aos_value_grad_lap

    return aos_value_grad_lap(aos_data, r_carts, value_only=True)[0]


def compute_AOs_value_grad_lap(aos_data, r_carts):
    """Fused value, gradient and Laplacian of the AOs; atomic_orbital.py:3594-3640.

    Returns ``(val, gx, gy, gz, lap)``, each ``(num_ao, N_e)`` (leading ``W`` if given).
    """
    from ._api import aos_value_grad_lap

    return aos_value_grad_lap(aos_data, r_carts, value_only=False)
