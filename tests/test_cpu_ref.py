"""The C++/OpenMP restatement of the VMC step (oracle/cpu_ref.cpp, the CPU baseline of bench.py)
against the NumPy oracle: identical accept/reject counts and positions on the shared Philox stream,
local energies to 1e-9 relative, on the all-electron water and the benzene-like ECP system; the
thread count changes no result bit."""

import os

import numpy as np
import pytest

from jqmc_amd import synthetic
from jqmc_amd._pack import SystemPack
from oracle import coulomb as oc
from oracle import cpu_ref
from oracle import mcmc as om

pytestmark = pytest.mark.skipif(not os.path.exists(cpu_ref.LIB), reason="oracle/libcpuref.so not built")


@pytest.mark.parametrize("name,W,n_moves", [("water", 4, 30), ("benzene", 2, 30)])
def test_cpu_ref_matches_oracle(name, W, n_moves):
    system = synthetic.water_ae(0) if name == "water" else synthetic.benzene_ecp(0)
    ham = system.hamiltonian_data
    pack = SystemPack(ham, Nv=12, device=False)
    r_up, r_dn = system.initial_walkers(W, seed=4)
    seed = 99
    e, ou, od, acc, _ = cpu_ref.vmc_step(pack.desc, r_up, r_dn, n_moves, 0.5, 1e-4, seed, move0=0, rot_step=3)
    for w in range(W):
        G, Gi = om.init_walker(ham, r_up[w], r_dn[w])
        a, _, ru, rd, Gi2, _ = om.metropolis_walker(ham, r_up[w], r_dn[w], Gi, G, seed, w, 0, n_moves, 0.5, 1e-4)
        assert acc[w] == a
        np.testing.assert_allclose(ou[w], ru, rtol=0, atol=1e-10)
        np.testing.assert_allclose(od[w], rd, rtol=0, atol=1e-10)
        ref = oc.local_energy_fast(ham, ru, rd, om.rotation_RT(seed, w, 3), Gi2, NN=1, Nv=12)
        assert abs(e[w] - ref) <= 1e-9 * max(1.0, abs(ref)), (w, e[w], ref)


def test_cpu_ref_threads_bit_identical():
    system = synthetic.water_ae(0)
    pack = SystemPack(system.hamiltonian_data, Nv=6, device=False)
    r_up, r_dn = system.initial_walkers(8, seed=5)
    one = cpu_ref.vmc_step(pack.desc, r_up, r_dn, 20, 0.5, 1e-4, 3, n_threads=1)
    many = cpu_ref.vmc_step(pack.desc, r_up, r_dn, 20, 0.5, 1e-4, 3, n_threads=4)
    for a, b in zip(one[:4], many[:4]):
        np.testing.assert_array_equal(a, b)
