"""Run control of the drivers (jqmc_amd.checkpoint): a run stopped at max_time, and a run continued
from a checkpoint equals the uninterrupted run (same Philox stream position, same comb generator
state, walker caches rebuilt from the saved positions)."""

import numpy as np
import pytest

from jqmc_amd import synthetic
from jqmc_amd.jqmc_gfmc import GFMC_n
from jqmc_amd.jqmc_mcmc import MCMC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    return synthetic.water_ae(0)


def test_mcmc_checkpoint_continues_the_chain(water, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ham = water.hamiltonian_data
    a = MCMC(ham, mcmc_seed=11, num_walkers=64, num_mcmc_per_measurement=8, Dt=0.5)
    a.run(3)
    a.save_checkpoint(str(tmp_path / "ck.npz"))
    a.run(3)
    b = MCMC.load_checkpoint(str(tmp_path / "ck.npz"), ham)
    assert b.mcmc_counter == 3 and b.e_L.shape == (3, 64)
    b.run(3)
    np.testing.assert_allclose(b.e_L, a.e_L, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(b.w_L, a.w_L, rtol=1e-9)  # G^-1 rebuilt by LU instead of streamed
    assert b.accepted_moves == a.accepted_moves


def test_gfmc_n_checkpoint_continues_the_chain(water, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    ham = water.hamiltonian_data
    kw = dict(num_walkers=64, num_mcmc_per_measurement=4, mcmc_seed=5, E_scf=-70.0, alat=0.3)
    a = GFMC_n(ham, **kw)
    a.run(3)
    a.save_checkpoint(str(tmp_path / "ck.npz"))
    a.run(3)
    b = GFMC_n.load_checkpoint(str(tmp_path / "ck.npz"), ham)
    b.run(3)
    np.testing.assert_allclose(b.e_L, a.e_L, rtol=1e-9)
    np.testing.assert_allclose(b.w_L, a.w_L, rtol=1e-9)
    assert b.E_scf == a.E_scf


def test_max_time_stops_after_one_step(water, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    d = MCMC(water.hamiltonian_data, mcmc_seed=3, num_walkers=16, num_mcmc_per_measurement=4)
    d.run(50, max_time=0.0)
    assert d.mcmc_counter == 1
    assert not (tmp_path / "external_control_mcmc.toml").exists()
