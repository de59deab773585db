"""MI355X-native (gfx950) implementation of jQMC's per-walker VMC/LRDMC hot path.

The data classes keep the reference's names and fields; the compute functions run the
hand-written HIP kernels of ``libjqmc_hip.so`` through its C ABI (``include/jqmc_hip.h``).
"""

from .structure import Structure_data  # noqa: F401
from .atomic_orbital import AOs_cart_data, AOs_sphe_data  # noqa: F401
from .molecular_orbital import MOs_data  # noqa: F401
from .determinant import Geminal_data  # noqa: F401
from .jastrow_factor import (  # noqa: F401
    Jastrow_data,
    Jastrow_one_body_data,
    Jastrow_three_body_data,
    Jastrow_two_body_data,
)
from .coulomb_potential import Coulomb_potential_data  # noqa: F401
from .wavefunction import Wavefunction_data  # noqa: F401
from .hamiltonians import Hamiltonian_data  # noqa: F401

__version__ = "0.1.0"
