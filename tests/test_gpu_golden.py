"""The reference's TurboRVB golden cases run through the gfx950 path (spherical AOs, separate
J3 basis, ECP with Nv = 6 and no rotation): wavefunction ratio from ln|Psi| of the walker
initialisation, per-walker kinetic and potential energy from jqmc_local_energy.  Checked
against the golden values at the reference's tolerances (decimal 6 / 5) and against the FP64
oracle at 1e-10 relative."""

import numpy as np
import pytest
import torch

from jqmc_amd._api import WalkerState
from jqmc_amd._pack import SystemPack
from jqmc_amd.hamiltonians import Hamiltonian_data
from oracle import coulomb as oc
from oracle import wf as owf

from test_golden_turborvb import CASES, build_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_reproduces_turborvb(name):
    st, cp, wf, (ou, od), (nu, nd), v = build_case(name)
    ham = Hamiltonian_data(st, cp, wf)
    pack = SystemPack(ham, Nv=6, NN=1)
    ws = WalkerState(pack, 2).init(np.stack([ou, nu]), np.stack([od, nd]))
    e_L, kin, pot, tel = ws.local_energy(torch.eye(3, dtype=torch.float64).expand(2, 3, 3), with_parts=True)
    torch.cuda.synchronize()
    lnpsi = (ws.jastrow + ws.logabsdet).cpu().numpy()
    ratio = float(np.exp(2.0 * (lnpsi[1] - lnpsi[0])))
    # FP64 parity with the oracle
    ref_ratio = (owf.wavefunction(wf, nu, nd) / owf.wavefunction(wf, ou, od)) ** 2
    assert abs(ratio - ref_ratio) <= 1e-10 * abs(ref_ratio)
    ref_kin = owf.kinetic_energy(wf, nu, nd)
    assert abs(kin[1].item() - ref_kin) <= 1e-10 * max(1.0, abs(ref_kin))
    Ginv = np.linalg.inv(owf.geminal(wf.geminal_data, nu, nd))
    ref_pot = oc.bare_coulomb(cp, nu, nd)
    if cp.ecp_flag:
        ref_pot += oc.ecp_local(cp, nu, nd) + oc.ecp_nonlocal_fast(cp, wf, nu, nd, np.eye(3), Ginv, NN=1, Nv=6)
    assert abs(pot[1].item() - ref_pot) <= 1e-10 * max(1.0, abs(ref_pot))
    # golden values at the reference's tolerances
    if not name.startswith("test_full_comparison"):
        np.testing.assert_almost_equal(ratio, v["WF_ratio_ref_turborvb"], decimal=6)
    else:
        np.testing.assert_almost_equal(ws.G[0].cpu().numpy(), np.array(v["geminal_old_turborvb"]), decimal=6)
        np.testing.assert_almost_equal(ws.G[1].cpu().numpy(), np.array(v["geminal_new_turborvb"]), decimal=6)
        np.testing.assert_allclose(ws.Ginv[1].cpu().numpy(), np.array(v["geminal_new_inv_turborvb"]), rtol=1e-6)
    np.testing.assert_almost_equal(kin[1].item(), v["kinc_ref_turborvb"], decimal=6)
    if cp.ecp_flag:
        np.testing.assert_almost_equal(pot[1].item(), v["vpot_ref_turborvb"] + v["vpotoff_ref_turborvb"], decimal=5)
    else:
        np.testing.assert_almost_equal(pot[1].item(), v["vpot_ref_turborvb"], decimal=3)


def test_gpu_spherical_aos_match_oracle():
    from jqmc_amd._api import aos_value_grad_lap

    st, cp, wf, _, _, _ = build_case("test_comparison_with_TurboRVB_w_2b_3b_Jastrow_w_ecp")
    for aos in (wf.geminal_data.aos_data, wf.jastrow_data.jastrow_three_body_data.orb_data):
        r = st.positions[np.arange(24) % st.natom] + np.random.default_rng(2).normal(0, 0.8, (24, 3))
        got = [t.cpu().numpy() for t in aos_value_grad_lap(aos, r, value_only=False)]
        ref = owf.aos_vgl(aos, r)
        for g, f in zip(got, ref):
            assert np.max(np.abs(g - f)) <= 1e-12 * max(1.0, np.max(np.abs(f)))


def test_gpu_uhf_trexio_matches_oracle():
    """Spin-dependent MOs (N_ae_ccpvdz_cart.h5, 5 up / 2 down, separate up and down coefficient
    rows) with a two-body Jastrow: ln|Psi| of the set-up, kinetic and potential energy of the GPU
    path against the FP64 oracle at 1e-10 relative, and the energies again after 20 Metropolis
    moves (streamed caches).  The kinetic energy of
    this all-electron atom sums terms up to ~10^3 Ha (Laplacian and squared gradient of ln|det|
    near the nucleus) into a few Ha, so its tolerance is 1e-10 of that term scale, sum_e |lap_e
    ln|det|| + |grad_e ln|det||^2 (oracle), the condition number of the sum times its value."""
    import os

    from jqmc_amd.jastrow_factor import Jastrow_data, Jastrow_two_body_data
    from jqmc_amd.trexio_wrapper import read_trexio_file
    from jqmc_amd.wavefunction import Wavefunction_data

    h5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "N_ae_ccpvdz_cart.h5")
    st, aos, mu, md, gem, cp = read_trexio_file(h5)
    wf = Wavefunction_data(Jastrow_data(None, Jastrow_two_body_data(1.0), None), gem)
    ham = Hamiltonian_data(st, cp, wf)
    pack = SystemPack(ham, Nv=6, NN=1)
    rng = np.random.default_rng(9)
    W = 8
    r_up = st.positions[0] + rng.normal(0.0, 1.0, (W, 5, 3))
    r_dn = st.positions[0] + rng.normal(0.0, 1.0, (W, 2, 3))
    ws = WalkerState(pack, W).init(r_up, r_dn)
    for n_moves in (0, 20):
        if n_moves:
            ws.sweep(n_moves, 0.5, 1e-4, 3, 0)
        eye = torch.eye(3, dtype=torch.float64).expand(W, 3, 3)
        e_L, kin, pot, tel = ws.local_energy(eye, with_parts=True)
        torch.cuda.synchronize()
        ru, rd = ws.r_up.cpu().numpy(), ws.r_dn.cpu().numpy()
        lnpsi = (ws.jastrow + ws.logabsdet).cpu().numpy()
        for w in range(W):
            if n_moves == 0:  # the walker caches hold ln|Psi| of the set-up
                ref_ln = owf.ln_wavefunction(wf, ru[w], rd[w])
                assert abs(lnpsi[w] - ref_ln) <= 1e-10 * max(1.0, abs(ref_ln))
            ref_kin = owf.kinetic_energy(wf, ru[w], rd[w])
            Gi = owf.geminal_inverse(owf.geminal(gem, ru[w], rd[w]))
            gu, gd, lu, ld = owf.grads_laplacian_ln_det(gem, ru[w], rd[w], Gi)
            scale = np.abs(lu).sum() + np.abs(ld).sum() + (gu**2).sum() + (gd**2).sum()
            assert abs(kin[w].item() - ref_kin) <= 1e-10 * max(1.0, abs(ref_kin), scale), (n_moves, w)
            ref_pot = oc.bare_coulomb(cp, ru[w], rd[w])
            assert abs(pot[w].item() - ref_pot) <= 1e-10 * max(1.0, abs(ref_pot))
