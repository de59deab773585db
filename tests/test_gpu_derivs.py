"""GPU parity of the parameter log-derivatives O_k (jqmc_dln_psi_dparams) with the FP64 oracle
(oracle/derivs.py, itself pinned on finite differences): 1e-10 relative to each block's scale
in FP64; the autograd wrapper's vector-Jacobian product equals the walker sum of O_k."""

import numpy as np
import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import WalkerState
from jqmc_amd._pack import SystemPack
from jqmc_amd.derivatives import ln_psi_params
from oracle import derivs as od

from test_golden_turborvb import build_case

pytestmark = pytest.mark.gpu


def _close(a, b, rtol):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = max(1e-300, float(np.max(np.abs(b))))
    err = float(np.max(np.abs(a - b))) / scale
    assert err <= rtol, f"max scaled error {err:.3e} > {rtol:.1e}"


CASES = {
    "water_jsd": lambda: synthetic.water_ae(0),
    "water_jagp": lambda: synthetic.water_ae(3, jagp=True, j1=("pade", 0.9), j2=("exp", 1.1)),
    "benzene": lambda: synthetic.benzene_ecp(0),
    # the 4 x H2O JAGP cluster: J3 (300 x 301) and Lambda (300 x 300) blocks
    "water4_jagp": lambda: synthetic.water4_ecp(0, jagp=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_dln_psi_dparams_match_oracle(name):
    system = CASES[name]()
    ham = system.hamiltonian_data
    W = 3
    r_up, r_dn = system.initial_walkers(W, seed=2)
    ws = WalkerState(SystemPack(ham), W).init(r_up, r_dn)
    ws.sweep(system.n_up + system.n_dn, 0.5, 1e-4, 3, 0)  # derivatives from streamed caches
    O = {k: v.cpu().numpy() for k, v in ws.dln_psi_dparams().items()}
    ru, rd = ws.r_up.cpu().numpy(), ws.r_dn.cpu().numpy()
    for w in range(W):
        ref = od.dln_psi_dparams(ham.wavefunction_data, ru[w], rd[w])
        assert set(ref) == set(O)
        for k in ref:
            _close(O[k][w], ref[k], rtol=1e-9)


def test_dln_psi_dparams_separate_j3_basis():
    st, cp, wf, _, (nu, nd), _ = build_case("test_comparison_with_TurboRVB_w_2b_3b_Jastrow_w_ecp")
    from jqmc_amd.hamiltonians import Hamiltonian_data

    ham = Hamiltonian_data(st, cp, wf)
    pack = SystemPack(ham, Nv=6)
    assert pack.jas_separate
    ws = WalkerState(pack, 1).init(np.asarray(nu)[None], np.asarray(nd)[None])
    O = {k: v.cpu().numpy() for k, v in ws.dln_psi_dparams().items()}
    ref = od.dln_psi_dparams(wf, nu, nd)
    for k in ref:
        assert O[k][0].shape == np.shape(ref[k])
        _close(O[k][0], ref[k], rtol=1e-9)


def test_autograd_vjp_is_walker_sum():
    system = CASES["water_jagp"]()
    ham = system.hamiltonian_data
    W = 5
    r_up, r_dn = system.initial_walkers(W, seed=8)
    ws = WalkerState(SystemPack(ham), W).init(r_up, r_dn)
    lnpsi, params = ln_psi_params(ws, ham)
    g = torch.linspace(0.5, 1.5, W, dtype=torch.float64, device=lnpsi.device)
    (lnpsi * g).sum().backward()
    O = ws.dln_psi_dparams()
    for k, t in params.items():
        want = torch.tensordot(g, O[k], dims=([0], [0]))
        _close(t.grad.cpu().numpy(), want.cpu().numpy(), rtol=1e-12)
