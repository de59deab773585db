"""Hamiltonian container and local energy (``jqmc/hamiltonians.py``).

``Hamiltonian_data`` mirrors hamiltonians.py:82-177.  ``compute_local_energy``
(179-222) and ``compute_local_energy_fast`` (225-290) keep the reference signatures:
``e_L = T + V`` with ``V`` = bare Coulomb + ECP local + ECP non-local (NN nearest nuclei,
Nv-point quadrature rotated by ``RT``).  Both run the fused gfx950 local-energy kernel.

``save_to_hdf5`` / ``load_from_hdf5`` keep the reference's file layout (hamiltonians.py:156-176,
339-573): one group per data class with attributes ``_class_name`` and ``_module_name``, arrays
and lists of numbers or strings as datasets named after the field, empty lists as empty
datasets, lists of arrays as groups flagged ``_is_list`` with datasets "0", "1", ..., and
scalars (int, float, bool, str) as attributes of the group.  The file is written by
``_hdf5_write`` and read by ``_hdf5`` (neither needs h5py); classes are found by
``_class_name``, fields missing from the file keep their defaults and fields unknown here are
ignored, as the reference's loader does.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

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

    def save_to_hdf5(self, filepath: str = "jqmc.h5") -> None:
        """Write this Hamiltonian in the reference's HDF5 layout (hamiltonians.py:156-163)."""
        from . import _hdf5_write

        _hdf5_write.write(filepath, _to_group(self))

    @staticmethod
    def load_from_hdf5(filepath: str = "jqmc.h5") -> "Hamiltonian_data":
        """Read a Hamiltonian written in the reference's HDF5 layout (hamiltonians.py:165-176)."""
        from . import _hdf5

        return _load_dataclass(Hamiltonian_data, _hdf5.File(filepath).root)


# ---------------------------------------------------------------- HDF5 layout


def _classes() -> dict:
    from .atomic_orbital import AOs_cart_data, AOs_sphe_data
    from .determinant import Geminal_data
    from .jastrow_factor import Jastrow_data, Jastrow_one_body_data, Jastrow_three_body_data, Jastrow_two_body_data
    from .molecular_orbital import MOs_data

    mods = {
        Structure_data: "jqmc.structure", AOs_cart_data: "jqmc.atomic_orbital", AOs_sphe_data: "jqmc.atomic_orbital",
        MOs_data: "jqmc.molecular_orbital", Geminal_data: "jqmc.determinant",
        Jastrow_one_body_data: "jqmc.jastrow_factor", Jastrow_two_body_data: "jqmc.jastrow_factor",
        Jastrow_three_body_data: "jqmc.jastrow_factor", Jastrow_data: "jqmc.jastrow_factor",
        Coulomb_potential_data: "jqmc.coulomb_potential", Wavefunction_data: "jqmc.wavefunction",
        Hamiltonian_data: "jqmc.hamiltonians",
    }
    return {c.__name__: (c, m) for c, m in mods.items()}


def _is_scalar(v) -> bool:
    return isinstance(v, (bool, int, float, str, np.number, np.bool_))


def _put(group, name: str, v) -> None:
    from ._hdf5_write import Group

    if v is None:
        return
    if dataclasses.is_dataclass(v):
        group.children[name] = _to_group(v)
    elif isinstance(v, np.ndarray):
        group.children[name] = v
    elif isinstance(v, (list, tuple)):
        if len(v) == 0:
            group.children[name] = np.array([])
        elif all(isinstance(x, str) for x in v):
            group.children[name] = [str(x) for x in v]
        elif all(_is_scalar(x) for x in v):
            group.children[name] = np.asarray(v)
        else:
            sub = Group()
            sub.attrs["_is_list"] = True
            for i, x in enumerate(v):
                _put(sub, str(i), x)
            group.children[name] = sub
    elif _is_scalar(v):
        group.attrs[name] = v
    else:
        raise TypeError(f"cannot store field {name} of type {type(v).__name__}")


def _to_group(obj):
    from ._hdf5_write import Group

    g = Group()
    g.attrs["_class_name"] = type(obj).__name__
    g.attrs["_module_name"] = _classes()[type(obj).__name__][1]
    for f in dataclasses.fields(obj):
        _put(g, f.name, getattr(obj, f.name))
    return g


def _load_item(node):
    if not node.is_group:
        return node.read()
    if node.attrs.get("_is_list"):
        keys = sorted(set(node.keys()) | {k for k in node.attrs if k.isdigit()}, key=int)
        return [_load_item(node[k]) if k in node.keys() else node.attrs[k] for k in keys]
    name = node.attrs.get("_class_name")
    table = _classes()
    if name in table:
        return _load_dataclass(table[name][0], node)
    return {k: _load_item(node[k]) for k in node.keys()}


def _convert(value, ftype: str):
    if "tuple" in ftype or "list" in ftype:
        if isinstance(value, np.ndarray):
            return tuple(value.tolist())
        if isinstance(value, (list, tuple)):
            return tuple(value)
    if "ndarray" in ftype:
        return np.asarray(value)
    if ftype == "bool":
        return bool(value)
    if ftype == "int":
        return int(value)
    if ftype == "float":
        return float(value)
    return value


def _load_dataclass(cls, node):
    kwargs = {}
    children = set(node.keys())
    for f in dataclasses.fields(cls):
        ftype = f.type if isinstance(f.type, str) else getattr(f.type, "__name__", str(f.type))
        if f.name in children:
            kwargs[f.name] = _convert(_load_item(node[f.name]), ftype)
        elif f.name in node.attrs:
            kwargs[f.name] = _convert(node.attrs[f.name], ftype)
    return cls(**kwargs)


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
