"""Host drivers with the reference's MCMC / GFMC_n / GFMC_t interfaces, run on the GPU kernels.

Checks the step structure rather than statistics: energies finite and below the bare
electron-ion floor, AS weights in (0, 1], comb branching keeps the population, the drivers'
per-step energy equals the weighted sums of the kernels' per-walker values, and GFMC_t
uses its projection time exactly."""

import numpy as np
import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import WalkerState
from jqmc_amd.jqmc_gfmc import GFMC_n, GFMC_t
from jqmc_amd.jqmc_mcmc import MCMC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    return synthetic.water_ae(0)


def test_mcmc_driver_steps(water):
    ham = water.hamiltonian_data
    r_up, r_dn = water.initial_walkers(128, seed=5)
    m = MCMC(ham, mcmc_seed=3, num_walkers=128, num_mcmc_per_measurement=20, Dt=0.5, epsilon_AS=1e-4,
             init_r_up_carts=r_up, init_r_dn_carts=r_dn)
    m.run(12)
    assert m.e_L.shape == (12, 128) and m.w_L.shape == (12, 128)
    assert np.all(np.isfinite(m.e_L))
    assert np.all((m.w_L > 0.0) & (m.w_L <= 1.0))
    acc = m.accepted_moves / (m.accepted_moves + m.rejected_moves)
    assert 0.1 < acc < 0.99
    E, E_std, V, V_std = m.get_E(num_mcmc_warmup_steps=2, num_mcmc_bin_blocks=5)
    assert np.isfinite(E) and E_std >= 0.0 and V > 0.0


def test_mcmc_mixed_driver_refreshes(water):
    ham = water.hamiltonian_data
    r_up, r_dn = water.initial_walkers(64, seed=6)
    m = MCMC(ham, mcmc_seed=3, num_walkers=64, num_mcmc_per_measurement=20, Dt=0.5, epsilon_AS=1e-4,
             precision="mixed", refresh_interval=32, init_r_up_carts=r_up, init_r_dn_carts=r_dn)
    m.run(4)
    assert m._move == 80 and m._since_refresh == 80 % 32
    assert np.all(np.isfinite(m.e_L))


def test_gfmc_n_driver_branching(water):
    ham = water.hamiltonian_data
    W = 96
    r_up, r_dn = water.initial_walkers(W, seed=7)
    g = GFMC_n(ham, num_walkers=W, num_mcmc_per_measurement=5, mcmc_seed=11, E_scf=-70.0, alat=0.3,
               init_r_up_carts=r_up, init_r_dn_carts=r_dn)
    g.run(6)
    assert g.mcmc_counter == 6 and len(g.e_L) == 6
    assert np.all(np.isfinite(g.e_L)) and np.all(g.w_L > 0.0)
    assert g.num_survived_walkers + g.num_killed_walkers == 6 * W
    # after a reconfiguration every cache equals a rebuild from the copied positions
    Gi = g.state.Ginv.clone()
    g.state.refresh()
    scale = float(Gi.abs().max())
    assert float((Gi - g.state.Ginv).abs().max()) <= 1e-9 * scale


def test_gfmc_n_step_energy_is_weighted_mean(water):
    """e_L of a branching step = sum w e_L / (V_diag - E) over sum w / (V_diag - E) (jqmc_gfmc.py:5957-5990)."""
    ham = water.hamiltonian_data
    W = 64
    r_up, r_dn = water.initial_walkers(W, seed=8)
    g = GFMC_n(ham, num_walkers=W, num_mcmc_per_measurement=3, mcmc_seed=2, E_scf=-70.0, alat=0.3,
               init_r_up_carts=r_up, init_r_dn_carts=r_dn)
    ws = g.state
    p = WalkerState.lrdmc_params(0.3, E_scf=-70.0, n_steps=3)
    out = ws.lrdmc_projection(p, 2, 0)
    Vd, Vn = ws.lrdmc_local_energy(p, out["RT"])
    w, vd, e = out["w_L"].cpu().numpy(), Vd.cpu().numpy(), (Vd + Vn).cpu().numpy()
    f = w / (vd + 70.0)
    want = np.sum(f * e) / np.sum(f)
    g._branch(out["w_L"], Vd, Vd + Vn)
    assert g.e_L[-1] == pytest.approx(want, rel=1e-12)


def test_gfmc_t_driver_uses_tau(water):
    ham = water.hamiltonian_data
    W = 48
    r_up, r_dn = water.initial_walkers(W, seed=9)
    g = GFMC_t(ham, num_walkers=W, mcmc_seed=4, E_scf=-70.0, tau=0.02, alat=0.3, init_r_up_carts=r_up,
               init_r_dn_carts=r_dn)
    g.run(3)
    assert len(g.e_L) == 3 and np.all(np.isfinite(g.e_L)) and np.all(g.w_L > 0.0)
