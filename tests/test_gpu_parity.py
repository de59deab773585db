"""GPU parity of the gfx950 kernels against the FP64 CPU oracle (same seeded inputs).

Tolerances: FP64 results must agree to 1e-10 relative (log|Psi|, ratios, E_L), the
tolerance of the plain FP64 form; Metropolis accept/reject decisions driven by the same
Philox stream must be identical.
"""

import numpy as np
import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import WalkerState, aos_value_grad_lap, mos_value_grad_lap
from jqmc_amd._pack import system_for
from oracle import coulomb as oc
from oracle import mcmc as om
from oracle import wf as owf

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _scale_close(a, b, rtol=RTOL, floor=1.0):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(floor, float(np.max(np.abs(b))))
    err = float(np.max(np.abs(a - b))) / scale
    assert err <= rtol, f"max scaled error {err:.3e} > {rtol:.1e}"


@pytest.fixture(scope="module")
def water():
    return synthetic.water_ae(0)


@pytest.fixture(scope="module")
def benzene():
    return synthetic.benzene_ecp(0)


@pytest.fixture(scope="module")
def water_sym():
    return synthetic.water_ae(1, j3_symmetric=True)


@pytest.fixture(scope="module")
def water4():
    return synthetic.water4_ecp(0)


def _pick(name, water, benzene, water_sym=None, water4=None):
    return {"water": water, "benzene": benzene, "water_sym": water_sym, "water4": water4}[name]


@pytest.mark.parametrize("name", ["water", "benzene"])
def test_aos_value_grad_lap(name, water, benzene):
    system = water if name == "water" else benzene
    aos = system.hamiltonian_data.wavefunction_data.geminal_data.aos_data
    rng = np.random.default_rng(3)
    r = aos.structure_data.positions[rng.integers(0, aos.structure_data.natom, 40)] + rng.normal(0, 1.0, (40, 3))
    got = [t.cpu().numpy() for t in aos_value_grad_lap(aos, r, value_only=False)]
    ref = owf.aos_cart_vgl(aos, r)
    for g, f in zip(got, ref):
        _scale_close(g, f, rtol=1e-12)
    val = aos_value_grad_lap(aos, r, value_only=True)[0].cpu().numpy()
    _scale_close(val, owf.aos_cart_value(aos, r), rtol=1e-12)


def test_mos_value_grad_lap(benzene):
    mos = benzene.hamiltonian_data.wavefunction_data.geminal_data.orb_data_up_spin
    rng = np.random.default_rng(4)
    r = rng.normal(0, 2.5, (3, 17, 3))  # batched (W, N, 3)
    got = [t.cpu().numpy() for t in mos_value_grad_lap(mos, r, value_only=False)]
    for w in range(3):
        ref = owf.orb_vgl(mos, r[w])
        for g, f in zip(got, ref):
            _scale_close(g[w], f, rtol=1e-12)


@pytest.mark.parametrize("name", ["water", "benzene", "water4"])
def test_walkers_init_matches_oracle(name, water, benzene, water4):
    system = _pick(name, water, benzene, water4=water4)
    ham = system.hamiltonian_data
    W = 6
    r_up, r_dn = system.initial_walkers(W, seed=11)
    ws = WalkerState(system_for(ham), W).init(r_up, r_dn)
    torch.cuda.synchronize()
    wf = ham.wavefunction_data
    for w in range(W):
        G = owf.geminal(wf.geminal_data, r_up[w], r_dn[w])
        _scale_close(ws.G[w].cpu().numpy(), G, rtol=1e-12)
        sign, logabs = np.linalg.slogdet(G)
        assert abs(ws.logabsdet[w].item() - logabs) <= RTOL * max(1.0, abs(logabs))
        assert ws.sign[w].item() == sign
        _scale_close(ws.Ginv[w].cpu().numpy(), np.linalg.inv(G), rtol=1e-9)
        J = owf.jastrow_part(wf.jastrow_data, r_up[w], r_dn[w])
        assert abs(ws.jastrow[w].item() - J) <= RTOL * max(1.0, abs(J))
        lnpsi = owf.ln_wavefunction(wf, r_up[w], r_dn[w])
        assert abs(ws.jastrow[w].item() + ws.logabsdet[w].item() - lnpsi) <= RTOL * max(1.0, abs(lnpsi))
        assert abs(ws.as_factor[w].item() - owf.as_regularization_fast(G, np.linalg.inv(G))) <= 1e-9 * ws.as_factor[
            w
        ].item()


@pytest.mark.parametrize("name", ["water", "benzene", "water4"])
def test_local_energy_matches_oracle(name, water, benzene, water4):
    system = _pick(name, water, benzene, water4=water4)
    ham = system.hamiltonian_data
    W = 3
    r_up, r_dn = system.initial_walkers(W, seed=5)
    pack = system_for(ham, Nv=12)
    ws = WalkerState(pack, W).init(r_up, r_dn)
    RT = ws.rotations(seed=9, step=0)
    e_L, kin, pot, tel = ws.local_energy(RT, with_parts=True)
    torch.cuda.synchronize()
    for w in range(W):
        G = owf.geminal(ham.wavefunction_data.geminal_data, r_up[w], r_dn[w])
        Ginv = np.linalg.inv(G)
        tu, td = owf.kinetic_all_elements_fast(ham.wavefunction_data, r_up[w], r_dn[w], Ginv)
        _scale_close(tel[w].cpu().numpy(), np.concatenate([tu, td]), rtol=1e-9)
        e_ref = oc.local_energy_fast(ham, r_up[w], r_dn[w], RT[w].cpu().numpy(), Ginv, NN=1, Nv=12)
        assert abs(e_L[w].item() - e_ref) <= RTOL * max(1.0, abs(e_ref)), (w, e_L[w].item(), e_ref)


@pytest.mark.parametrize("name,W,n_moves", [("water", 16, 100), ("water_sym", 9, 80), ("benzene", 3, 40),
                                             ("water4", 3, 32)])
def test_vmc_sweep_decisions_and_state(name, W, n_moves, water, benzene, water_sym, water4):
    system = _pick(name, water, benzene, water_sym, water4)
    ham = system.hamiltonian_data
    seed, Dt, eps = 1234, 0.5, 1e-4
    r_up, r_dn = system.initial_walkers(W, seed=21)
    ws = WalkerState(system_for(ham), W).init(r_up, r_dn)
    # two launches back to back: the stream continues at move0 = n_moves // 2
    half = n_moves // 2
    ws.sweep(half, Dt, eps, seed, 0)
    ws.sweep(n_moves - half, Dt, eps, seed, half)
    torch.cuda.synchronize()
    for w in range(W):
        G, Ginv = om.init_walker(ham, r_up[w], r_dn[w])
        acc, rej, ou, od, oGinv, oG = om.metropolis_walker(ham, r_up[w], r_dn[w], Ginv, G, seed, w, 0, n_moves, Dt,
                                                           eps)
        assert ws.accepted[w].item() == acc and ws.rejected[w].item() == rej
        np.testing.assert_allclose(ws.r_up[w].cpu().numpy(), ou, rtol=0, atol=1e-10)
        np.testing.assert_allclose(ws.r_dn[w].cpu().numpy(), od, rtol=0, atol=1e-10)
        _scale_close(ws.Ginv[w].cpu().numpy(), oGinv, rtol=1e-8)
        _scale_close(ws.G[w].cpu().numpy(), oG, rtol=1e-10)
    # streaming caches equal a fresh rebuild at the final positions
    ao, jf = ws.ao.clone(), ws.jfield.clone() if ws.jfield is not None else None
    Ginv_sm = ws.Ginv.clone()
    ws.refresh()
    torch.cuda.synchronize()
    _scale_close(ao.cpu().numpy(), ws.ao.cpu().numpy(), rtol=1e-12)
    if jf is not None:
        _scale_close(jf.cpu().numpy(), ws.jfield.cpu().numpy(), rtol=1e-10)
    _scale_close(Ginv_sm.cpu().numpy(), ws.Ginv.cpu().numpy(), rtol=1e-8)
