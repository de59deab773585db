"""The move-synchronous sweep (jqmc_vmc_sweep_batched: one launch per move over all walkers plus
the tiled J3 field product of the accepting walkers) against the per-walker sweep kernel
(jqmc_vmc_sweep) from identical states on the same Philox stream: identical decisions, hence
identical positions and accept counts, G^-1 within 1e-10 of its scale (its Sherman-Morrison
updates amplify last-bit differences of the AO values); the J3 fields, summed in a different order, within
1e-12 of their scale (FP64) or 1e-5 (mixed form), the AO cache within 1e-14 (1e-6)."""

import numpy as np
import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import WalkerState
from jqmc_amd._pack import system_for

pytestmark = pytest.mark.gpu

CASES = {
    "water_jsd": (lambda: synthetic.water_ae(0), "fp64"),
    "water_jagp": (lambda: synthetic.water_ae(3, jagp=True, j1=("pade", 0.9), j2=("exp", 1.1)), "fp64"),
    "n2_ecp": (lambda: synthetic.chain_ecp(10, seed=1), "fp64"),
    "n2_sym_j3": (lambda: synthetic.chain_ecp(10, seed=2, j3_symmetric=True), "fp64"),
    "n2_mixed": (lambda: synthetic.chain_ecp(10, seed=1), "mixed"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_batched_sweep_matches_resident_sweep(name):
    mk, prec = CASES[name]
    system = mk()
    pack = system_for(system.hamiltonian_data, Nv=6, precision=prec)
    W = 67  # not a multiple of the 32-walker J3 tile
    r_up, r_dn = system.initial_walkers(W, seed=9)
    a = WalkerState(pack, W).init(r_up, r_dn)
    b = WalkerState(pack, W).init(r_up, r_dn)
    for k in range(3):
        a.sweep(20, 0.5, 1e-4, 31, 20 * k, batched=False)
        b.sweep(20, 0.5, 1e-4, 31, 20 * k, batched=True)
    torch.cuda.synchronize()
    assert torch.equal(a.accepted, b.accepted) and torch.equal(a.rejected, b.rejected)
    assert torch.equal(a.r_up, b.r_up) and torch.equal(a.r_dn, b.r_dn)
    np.testing.assert_allclose(b.Ginv.cpu().numpy(), a.Ginv.cpu().numpy(), rtol=0,
                               atol=1e-10 * float(a.Ginv.abs().max()))
    np.testing.assert_allclose(b.as_factor.cpu().numpy(), a.as_factor.cpu().numpy(), rtol=1e-10)
    if a.jfield is not None:
        tol = 1e-12 if prec == "fp64" else 1e-5
        ja, jb = a.jfield.double().cpu().numpy(), b.jfield.double().cpu().numpy()
        assert np.max(np.abs(ja - jb)) <= tol * max(1.0, np.max(np.abs(ja)))
    # AO values of the move-synchronous sweep are evaluated shell by shell (scale x shared radial
    # x polynomial): equal to the per-AO evaluation up to the last bits
    aa, ab = a.ao.double().cpu().numpy(), b.ao.double().cpu().numpy()
    assert np.max(np.abs(aa - ab)) <= (1e-14 if prec == "fp64" else 1e-6) * np.max(np.abs(aa))
