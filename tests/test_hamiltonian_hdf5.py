"""Hamiltonian_data.save_to_hdf5 / load_from_hdf5 in the reference's layout (hamiltonians.py:156-176,
339-573), written by jqmc_amd._hdf5_write and read by jqmc_amd._hdf5 without h5py.

Pins: (1) the writer's datatype encodings are byte for byte those the HDF5 library wrote into the
reference's TREXIO fixture (IEEE float64, int64, and the 1-byte base of variable-length strings);
the reader is the one that reads those library-written fixtures; (2) round trips of four
Hamiltonians (ECP benzene JSD, 4 x H2O JAGP, the TREXIO water with a TurboRVB Jastrow, the UHF N
atom) return equal data classes, and the oracle's local energy of the reloaded Hamiltonian is
the original's bit for bit; (3) the group / dataset / attribute layout is the reference's:
class-named groups, field-named members, scalars as attributes, empty lists as empty datasets.
No file written by the reference itself is available to read (no h5py here): that direction is
pinned only through (1) and the layout rules of (3)."""

import dataclasses
import os

import numpy as np
import pytest

from jqmc_amd import _hdf5
from jqmc_amd import _hdf5_write as W
from jqmc_amd import synthetic
from jqmc_amd.hamiltonians import Hamiltonian_data
from oracle import coulomb as oc
from oracle import wf as owf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _datatype_messages(f):
    out = {}

    def walk(node):
        for k in node.keys():
            ch = node[k]
            if ch.is_group:
                walk(ch)
            else:
                for mt, data in ch.messages:
                    if mt == 0x0003:
                        out.setdefault(bytes(data[:8]), bytes(data))

    walk(f.root)
    return out


def test_datatype_encodings_match_the_hdf5_library():
    msgs = _datatype_messages(_hdf5.File(os.path.join(GOLDEN, "water_ccecp_ccpvqz.h5")))
    f64, i64, vl = W._dt_float64(), W._dt_int64(), W._dt_vlen_str()
    assert msgs[f64[:8]][: len(f64)] == f64
    assert msgs[i64[:8]][: len(i64)] == i64
    lib_vlen = [d for k, d in msgs.items() if k[0] & 0x0F == 9]
    assert lib_vlen and lib_vlen[0][8:20] == vl[8:20]  # the 1-byte character base type


def _cases():
    from test_golden_turborvb import build_case
    from jqmc_amd.jastrow_factor import Jastrow_data, Jastrow_two_body_data
    from jqmc_amd.trexio_wrapper import read_trexio_file
    from jqmc_amd.wavefunction import Wavefunction_data

    st, cp, wf, _, _, _ = build_case("test_comparison_with_TurboRVB_w_2b_3b_Jastrow_w_ecp")
    st2, _, _, _, gem2, cp2 = read_trexio_file(os.path.join(GOLDEN, "N_ae_ccpvdz_cart.h5"))
    return {
        "benzene": synthetic.benzene_ecp(0).hamiltonian_data,
        "water4_jagp": synthetic.water4_ecp(0, jagp=True).hamiltonian_data,
        "turborvb_water": Hamiltonian_data(st, cp, wf),
        "uhf_N": Hamiltonian_data(st2, cp2, Wavefunction_data(Jastrow_data(None, Jastrow_two_body_data(1.0), None), gem2)),
    }


def _equal(a, b, path=""):
    if dataclasses.is_dataclass(a):
        assert type(a) is type(b), path
        for f in dataclasses.fields(a):
            _equal(getattr(a, f.name), getattr(b, f.name), f"{path}.{f.name}")
    elif isinstance(a, np.ndarray):
        b = np.asarray(b)
        assert a.shape == b.shape and a.dtype.kind == b.dtype.kind and np.array_equal(a, b), path
    elif isinstance(a, (list, tuple)):
        assert type(a) is type(b) and len(a) == len(b), path
        for i, (x, y) in enumerate(zip(a, b)):
            _equal(x, y, f"{path}[{i}]")
    else:
        assert type(a) is type(b) and a == b, (path, a, b)


@pytest.mark.parametrize("name", ["benzene", "water4_jagp", "turborvb_water", "uhf_N"])
def test_round_trip(name, tmp_path):
    ham = _cases()[name]
    path = str(tmp_path / f"{name}.h5")
    ham.save_to_hdf5(path)
    back = Hamiltonian_data.load_from_hdf5(path)
    _equal(ham, back)
    st = ham.structure_data
    rng = np.random.default_rng(3)
    n_up, n_dn = ham.wavefunction_data.geminal_data.num_electron_up, ham.wavefunction_data.geminal_data.num_electron_dn
    r_up = st.positions[np.arange(n_up) % len(st.positions)] + rng.normal(0, 0.7, (n_up, 3))
    r_dn = st.positions[np.arange(n_dn) % len(st.positions)] + rng.normal(0, 0.7, (n_dn, 3))
    for h in (ham, back):
        G = owf.geminal(h.wavefunction_data.geminal_data, r_up, r_dn)
        e = oc.local_energy_fast(h, r_up, r_dn, np.eye(3), np.linalg.inv(G), NN=1, Nv=6)
        if h is ham:
            e0 = e
    assert e == e0


def test_layout_is_the_reference_layout(tmp_path):
    ham = _cases()["benzene"]
    path = str(tmp_path / "h.h5")
    ham.save_to_hdf5(path)
    root = _hdf5.File(path).root
    assert root.attrs["_class_name"] == "Hamiltonian_data" and root.attrs["_module_name"] == "jqmc.hamiltonians"
    assert sorted(root.keys()) == ["coulomb_potential_data", "structure_data", "wavefunction_data"]
    st = root["structure_data"]
    assert st.attrs["_class_name"] == "Structure_data" and st.attrs["pbc_flag"] == 0  # bool: enum attribute
    assert st["element_symbols"].read() == list(ham.structure_data.element_symbols)  # strings: vlen dataset
    assert st["vec_a"].read().shape == (0,)  # empty list: empty dataset
    gem = root["wavefunction_data/geminal_data"]
    assert gem.attrs["num_electron_up"] == 15 and "lambda_matrix" in gem.keys()
    assert gem["orb_data_up_spin"].attrs["_class_name"] == "MOs_data"
    assert gem["orb_data_up_spin/aos_data"].attrs["_class_name"] == "AOs_cart_data"
    jas = root["wavefunction_data/jastrow_data"]
    assert "jastrow_nn_data" not in jas.keys()  # None fields are not written
    assert jas["jastrow_two_body_data"].attrs["jastrow_2b_type"] == "pade"


def test_list_groups_and_many_members(tmp_path):
    g = W.Group()
    g.attrs["_class_name"] = "Something"
    lst = W.Group()
    lst.attrs["_is_list"] = True
    for i in range(12):  # more members than one symbol-table node holds
        lst.children[str(i)] = np.arange(i + 1, dtype=np.float64)
    g.children["arrays"] = lst
    for i in range(20):
        g.children[f"x{i:02d}"] = np.full((2, 3), float(i))
    path = str(tmp_path / "l.h5")
    W.write(path, g)
    root = _hdf5.File(path).root
    from jqmc_amd.hamiltonians import _load_item

    arrays = _load_item(root["arrays"])
    assert [len(a) for a in arrays] == list(range(1, 13))
    assert sorted(root.keys())[:3] == ["arrays", "x00", "x01"] and len(root.keys()) == 21
    assert np.array_equal(root["x19"].read(), np.full((2, 3), 19.0))
