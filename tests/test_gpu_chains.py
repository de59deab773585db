"""Batched LRDMC of uneven systems: the chain / ring molecules with 10, 30, 56 and 96 valence
electrons (jqmc_amd.synthetic.chain_ecp, ECP on the heavy atoms, lattice spacing a = 0.2).

* 30-electron member: lattice-Hamiltonian elements and a projection against the FP64 oracle run
  here; 56- and 96-electron members (N_up = 28 and 48, the largest wavefront LU of the set) against
  oracle values stored by tests/golden/make_chain_lrdmc.py (the oracle takes about a minute per
  96-electron projection).  FP64: 1e-10 relative; mixed form: 2e-5 of the elements' scale, and the
  same move.
* One launch of all four systems equals their single-system launches, FP64 and mixed (the same
  kernel arithmetic; system k draws from seed + k).
* Full size, 2048 walkers per system in one launch, 20 projections and a comb reconfiguration:
  finite positive weights, the population kept, and the reconfigured caches equal a fresh set-up.
"""

import os

import numpy as np
import pytest
import torch

from jqmc_amd import synthetic
from jqmc_amd._api import WalkerState, lrdmc_branching, lrdmc_projection_multi
from jqmc_amd._pack import system_for
from oracle import gfmc as og
from oracle import rng as orng
from oracle import wf as owf

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "chain_lrdmc.npz")
ALAT = 0.2


@pytest.fixture(scope="module")
def chains():
    return {n: synthetic.chain_ecp(n, seed=1) for n in (10, 30, 56, 96)}


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(GOLDEN))


def _reference(n, chains, golden):
    """Start positions, rotation and oracle values of one walker of system n."""
    RT = orng.rotation_matrix_T(*orng.rotation_angles(9, 0, 2))
    if n in (56, 96):
        g = {k[len(f"n{n}_"):]: v for k, v in golden.items() if k.startswith(f"n{n}_")}
        np.testing.assert_array_equal(golden["RT"], RT)
        return g["r_up"], g["r_dn"], RT, g
    s = chains[n]
    ham = s.hamiltonian_data
    ru, rd = s.initial_walkers(1, seed=5)
    G = owf.geminal(ham.wavefunction_data.geminal_data, ru[0], rd[0])
    Gi = np.linalg.inv(G)
    d, nd, _, _ = og.hamiltonian_elements(ham, ru[0], rd[0], Gi, RT, ALAT, "tmove", Nv=6, NN=1)
    p = og.projection_step(ham, ru[0], rd[0], Gi, G, 1.0, -50.0, ALAT, 7, 0, 0, "tmove")
    return ru[0], rd[0], RT, {"V_diag": d, "V_nondiag": nd, "proj_w": p["w"], "proj_r_up": p["r_up"],
                              "proj_r_dn": p["r_dn"], "proj_diag": p["diag"], "proj_nondiag": p["nondiag"]}


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [30, 56, 96])
def test_chain_lrdmc_vs_oracle(n, prec, chains, golden):
    ru, rd, RT, ref = _reference(n, chains, golden)
    ham = chains[n].hamiltonian_data
    tol = 1e-10 if prec == "fp64" else 2e-5
    pack = system_for(ham, Nv=6, NN=1, precision=prec)
    ws = WalkerState(pack, 1).init(ru[None], rd[None])
    Vd, Vn = ws.lrdmc_local_energy(WalkerState.lrdmc_params(ALAT), torch.as_tensor(RT[None]))
    scale = max(abs(float(ref["V_diag"])), abs(float(ref["V_nondiag"])))
    assert abs(Vd.item() - float(ref["V_diag"])) <= tol * scale
    assert abs(Vn.item() - float(ref["V_nondiag"])) <= tol * scale
    out = ws.lrdmc_projection(WalkerState.lrdmc_params(ALAT, E_scf=-50.0, n_steps=1), 7, 0)
    torch.cuda.synchronize()
    scale = max(abs(float(ref["proj_diag"])), abs(float(ref["proj_nondiag"])))
    assert abs(out["V_diag"].item() - float(ref["proj_diag"])) <= tol * scale
    assert abs(out["V_nondiag"].item() - float(ref["proj_nondiag"])) <= tol * scale
    assert abs(out["w_L"].item() - float(ref["proj_w"])) <= tol * abs(float(ref["proj_w"])) * 10
    np.testing.assert_allclose(ws.r_up[0].cpu().numpy(), ref["proj_r_up"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(ws.r_dn[0].cpu().numpy(), ref["proj_r_dn"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_multi_launch_equals_single_launches(prec, chains):
    sizes = (10, 30, 56, 96)
    Ws = {10: 6, 30: 5, 56: 3, 96: 2}
    params = [WalkerState.lrdmc_params(ALAT, E_scf=-40.0, n_steps=3, non_local_move=m)
              for m in ("tmove", "dltmove", "tmove", "tmove")]
    single, multi = [], []
    for n in sizes:
        pack = system_for(chains[n].hamiltonian_data, Nv=6, NN=1, precision=prec)
        ru, rd = chains[n].initial_walkers(Ws[n], seed=31)
        single.append(WalkerState(pack, Ws[n]).init(ru, rd))
        multi.append(WalkerState(pack, Ws[n]).init(ru, rd))
    outs_s = [ws.lrdmc_projection(p, 17 + k, 5 * k) for k, (ws, p) in enumerate(zip(single, params))]
    outs_m = lrdmc_projection_multi(multi, params, 17, [5 * k for k in range(4)])
    torch.cuda.synchronize()
    for a, b, os_, om in zip(single, multi, outs_s, outs_m):
        assert torch.equal(os_["n_proj"], om["n_proj"])
        np.testing.assert_allclose(b.r_up.cpu().numpy(), a.r_up.cpu().numpy(), rtol=0, atol=1e-12)
        np.testing.assert_allclose(b.r_dn.cpu().numpy(), a.r_dn.cpu().numpy(), rtol=0, atol=1e-12)
        for key in ("w_L", "V_diag", "V_nondiag"):
            np.testing.assert_allclose(om[key].cpu().numpy(), os_[key].cpu().numpy(), rtol=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_size_batch_properties(prec, chains):
    W = 2048
    states, params, w_L = [], [], []
    for n in (10, 30, 56, 96):
        pack = system_for(chains[n].hamiltonian_data, Nv=6, NN=1, precision=prec)
        ru, rd = chains[n].initial_walkers(W, seed=2)
        ws = WalkerState(pack, W).init(ru, rd)
        E0 = float(ws.local_energy(None).median())
        states.append(ws)
        params.append(WalkerState.lrdmc_params(ALAT, E_scf=E0, n_steps=1))
        w_L.append(torch.ones(W, dtype=torch.float64, device="cuda"))
    rng = np.random.RandomState(4)
    for k in range(20):
        outs = lrdmc_projection_multi(states, params, 23, [k] * 4, w_Ls=w_L)
        w_L = [o["w_L"] for o in outs]
    torch.cuda.synchronize()
    for n, ws, w in zip((10, 30, 56, 96), states, w_L):
        assert bool(torch.isfinite(w).all()) and bool((w > 0).all()), n
        assert bool(torch.isfinite(ws.r_up).all()) and bool(torch.isfinite(ws.Ginv).all()), n
        chosen, _ = lrdmc_branching(w, float(rng.random_sample()))
        ch = chosen.cpu().numpy()
        assert ch.shape == (W,) and ch.min() >= 0 and ch.max() < W and np.all(np.diff(ch) >= 0), n
        ws.reconfigure(chosen)
        fresh = WalkerState(ws.pack, W).init(ws.r_up, ws.r_dn)
        torch.cuda.synchronize()
        tol = 1e-9 if prec == "fp64" else 1e-4
        for key in ("orb", "G", "Ginv", "logabsdet"):
            a, b = getattr(ws, key), getattr(fresh, key)
            err = float((a - b).abs().max()) / max(1.0, float(b.abs().max()))
            assert err <= tol, (n, key, err)
