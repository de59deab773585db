"""A captured VMC step (jqmc_vmc_sweep_counted + jqmc_rotation_matrices_counted + jqmc_local_energy
+ jqmc_counter_add in one HIP graph) replays exactly the uncaptured calls with advancing Philox
counters: every replay is compared with the eager step on an identical walker set, with eager work
and synchronisations between the replays."""

import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import CapturedStep, WalkerState
from jqmc_amd._pack import system_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,precision", [("water", "fp64"), ("n2", "fp64"), ("n2", "mixed")])
def test_captured_step_replays_match_eager(name, precision):
    system = synthetic.water_ae(0) if name == "water" else synthetic.chain_ecp(10, seed=1)
    pack = system_for(system.hamiltonian_data, Nv=6, precision=precision)
    W, n_moves, seed, move0, step0 = 96, 10, 7, 100, 3
    r_up, r_dn = system.initial_walkers(W, seed=5)
    eager = WalkerState(pack, W).init(r_up, r_dn)
    graphed = WalkerState(pack, W).init(r_up, r_dn)
    step = CapturedStep(graphed, n_moves, 0.5, 1e-4, seed, move0=move0, step0=step0)
    for k in range(4):
        eager.sweep(n_moves, 0.5, 1e-4, seed, move0 + k * n_moves)
        e_eager = eager.local_energy(eager.rotations(seed, step0 + k))
        e_graph = step.replay().clone()
        torch.cuda.synchronize()
        assert torch.equal(e_graph, e_eager), k
        assert torch.equal(step.RT, eager.rotations(seed, step0 + k))
        _ = (eager.r_up * 2.0).sum().item()  # eager work and a synchronisation between replays
    for a, b in ((graphed.r_up, eager.r_up), (graphed.r_dn, eager.r_dn), (graphed.Ginv, eager.Ginv),
                 (graphed.accepted, eager.accepted), (graphed.ao, eager.ao)):
        assert torch.equal(a, b)
    assert step.ctr.tolist() == [move0 + 4 * n_moves, step0 + 4]
