"""Mixed-precision form against the FP64 oracle on the benzene-like ECP system and the 4 x H2O
cluster (same seeded inputs).

The mixed form follows the reference's precision zones (jqmc/_precision.py:383-411): AO values
(ao_eval) and the Jastrow work (jastrow_eval, jastrow_ratio) in FP32, the orbital/pair caches,
G, G^-1, every determinant ratio and field, the AO derivatives at the electrons and the energy
sums in FP64.  Per result, the error against the FP64 oracle is bounded here:

* ln|Psi| at set-up: ln|det G| is built from FP64 AO values, so only the FP32 J3 sum
  (1/2 sum_e a_e . (g_e + j)) differs: |d ln|Psi|| <= 2e-5.
* single-electron ratios along a mixed Metropolis sweep: the accept/reject decisions driven by
  the oracle's Philox stream are identical and the positions equal the oracle's (a ratio error
  of ~1e-6 relative moves a decision only when the uniform draw lies that close to it).
* E_L at the same positions, after the FP64 orbital refresh: kinetic determinant part exact,
  FP32 J3 fields and FP32 AO values at the quadrature points: |dE_L| <= 2e-5 * max(1, |E_L|).
* LRDMC V_diag / V_nondiag (lattice and quadrature ratios): <= 2e-5 of their scale.
"""

import numpy as np
import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import WalkerState
from jqmc_amd._pack import system_for
from oracle import coulomb as oc
from oracle import gfmc as og
from oracle import mcmc as om
from oracle import rng as orng
from oracle import wf as owf

pytestmark = pytest.mark.gpu

LNPSI_ATOL = 2e-5
EL_RTOL = 2e-5
LR_RTOL = 2e-5


@pytest.fixture(scope="module")
def systems():
    return {"benzene": synthetic.benzene_ecp(0), "water4": synthetic.water4_ecp(0)}


@pytest.mark.parametrize("name", ["benzene", "water4"])
def test_mixed_init_and_local_energy_vs_oracle(name, systems):
    system = systems[name]
    ham = system.hamiltonian_data
    W = 4
    r_up, r_dn = system.initial_walkers(W, seed=5)
    ws = WalkerState(system_for(ham, Nv=12, precision="mixed"), W).init(r_up, r_dn)
    RT = ws.rotations(seed=9, step=0)
    e_L = ws.local_energy(RT)
    torch.cuda.synchronize()
    wf = ham.wavefunction_data
    d_ln, d_e = [], []
    for w in range(W):
        G = owf.geminal(wf.geminal_data, r_up[w], r_dn[w])
        sign, logabs = np.linalg.slogdet(G)
        assert ws.sign[w].item() == sign
        assert abs(ws.logabsdet[w].item() - logabs) <= 1e-10 * max(1.0, abs(logabs))
        lnpsi = owf.ln_wavefunction(wf, r_up[w], r_dn[w])
        d_ln.append(abs(ws.jastrow[w].item() + ws.logabsdet[w].item() - lnpsi))
        e_ref = oc.local_energy_fast(ham, r_up[w], r_dn[w], RT[w].cpu().numpy(), np.linalg.inv(G), NN=1, Nv=12)
        d_e.append(abs(e_L[w].item() - e_ref) / max(1.0, abs(e_ref)))
    print(f"\n[{name}] mixed set-up: max |d ln Psi| = {max(d_ln):.2e}, max rel |d E_L| = {max(d_e):.2e}")
    assert max(d_ln) <= LNPSI_ATOL
    assert max(d_e) <= EL_RTOL


@pytest.mark.parametrize("name,W,n_moves", [("benzene", 3, 30), ("water4", 3, 32)])
def test_mixed_sweep_decisions_and_energy_vs_oracle(name, W, n_moves, systems):
    system = systems[name]
    ham = system.hamiltonian_data
    seed, Dt, eps = 1234, 0.5, 1e-4
    r_up, r_dn = system.initial_walkers(W, seed=21)
    ws = WalkerState(system_for(ham, Nv=12, precision="mixed"), W).init(r_up, r_dn)
    ws.sweep(n_moves, Dt, eps, seed, 0)
    ws.refresh_orbitals()
    RT = ws.rotations(seed=3, step=1)
    e_L = ws.local_energy(RT)
    torch.cuda.synchronize()
    d_e = []
    for w in range(W):
        G, Ginv = om.init_walker(ham, r_up[w], r_dn[w])
        acc, rej, ou, od, oGinv, _ = om.metropolis_walker(ham, r_up[w], r_dn[w], Ginv, G, seed, w, 0, n_moves, Dt, eps)
        assert ws.accepted[w].item() == acc and ws.rejected[w].item() == rej
        np.testing.assert_allclose(ws.r_up[w].cpu().numpy(), ou, rtol=0, atol=1e-10)
        np.testing.assert_allclose(ws.r_dn[w].cpu().numpy(), od, rtol=0, atol=1e-10)
        Gf = owf.geminal(ham.wavefunction_data.geminal_data, ou, od)
        e_ref = oc.local_energy_fast(ham, ou, od, RT[w].cpu().numpy(), np.linalg.inv(Gf), NN=1, Nv=12)
        d_e.append(abs(e_L[w].item() - e_ref) / max(1.0, abs(e_ref)))
    print(f"\n[{name}] mixed after {n_moves} moves: max rel |d E_L| = {max(d_e):.2e}")
    assert max(d_e) <= EL_RTOL


@pytest.mark.parametrize("name", ["benzene", "water4"])
def test_mixed_lrdmc_elements_vs_oracle(name, systems):
    system = systems[name]
    ham = system.hamiltonian_data
    W = 3
    r_up, r_dn = system.initial_walkers(W, seed=7)
    ws = WalkerState(system_for(ham, Nv=6, NN=1, precision="mixed"), W).init(r_up, r_dn)
    RT = np.stack([orng.rotation_matrix_T(*orng.rotation_angles(9, w, 2)) for w in range(W)])
    alat = 0.3
    Vd, Vn = ws.lrdmc_local_energy(WalkerState.lrdmc_params(alat), torch.as_tensor(RT))
    torch.cuda.synchronize()
    Vd, Vn = Vd.cpu().numpy(), Vn.cpu().numpy()
    errs = []
    for w in range(W):
        G = owf.geminal(ham.wavefunction_data.geminal_data, r_up[w], r_dn[w])
        d, n, _, _ = og.hamiltonian_elements(ham, r_up[w], r_dn[w], np.linalg.inv(G), RT[w], alat, "tmove", Nv=6,
                                             NN=1)
        scale = max(1.0, abs(d), abs(n))
        errs.append(max(abs(Vd[w] - d), abs(Vn[w] - n)) / scale)
    print(f"\n[{name}] mixed LRDMC elements: max scaled error = {max(errs):.2e}")
    assert max(errs) <= LR_RTOL


@pytest.mark.parametrize("name", ["benzene", "water4"])
def test_guard_recomputes_flagged_walkers_in_fp64(name, systems):
    """With a guard width that flags every walker, the mixed E_L call returns the FP64 local
    energy recomputed from the positions by k_local_energy_guard: equal to the plain FP64 form at the
    same positions and rotations (FP64 sums in another order: four waves share the walker) within
    1e-10 of |T| + |V|, the scale of the terms that cancel in e_L, kinetic and potential parts
    alike, and to the oracle at 1e-10; 300 walkers, more than the guard's 256 blocks, so that blocks
    take a second walker."""
    from jqmc_amd import _lib

    system = systems[name]
    ham = system.hamiltonian_data
    W = 300
    r_up, r_dn = system.initial_walkers(W, seed=13)
    pm = system_for(ham, Nv=12, precision="mixed")
    wm = WalkerState(pm, W).init(r_up, r_dn)
    wd = WalkerState(system_for(ham, Nv=12), W).init(r_up, r_dn)
    RT = wd.rotations(seed=4, step=2)
    try:
        _lib.call("jqmc_system_set_mixed_guard", pm.handle, 0.0, 1e-300)
        em, km, vm, tm = wm.local_energy(RT, with_parts=True)
    finally:
        _lib.call("jqmc_system_set_mixed_guard", pm.handle, 0.0, 1e4)
    ed, kd, vd, td = wd.local_energy(RT, with_parts=True)
    torch.cuda.synchronize()
    scale = np.maximum(1.0, np.abs(kd.cpu().numpy()) + np.abs(vd.cpu().numpy()))
    for got, ref in ((em, ed), (km, kd), (vm, vd)):
        g, r = got.cpu().numpy(), ref.cpu().numpy()
        assert np.all(np.abs(g - r) <= 1e-10 * scale), float(np.max(np.abs(g - r) / scale))
    np.testing.assert_allclose(tm.cpu().numpy(), td.cpu().numpy(), rtol=1e-9, atol=1e-9)
    wf = ham.wavefunction_data
    for w in (0, 150, 299):
        G = owf.geminal(wf.geminal_data, r_up[w], r_dn[w])
        e_ref = oc.local_energy_fast(ham, r_up[w], r_dn[w], RT[w].cpu().numpy(), np.linalg.inv(G), NN=1, Nv=12)
        assert abs(em[w].item() - e_ref) <= 1e-10 * scale[w]
