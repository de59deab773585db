"""Pins of the oracle's position derivatives (oracle/forces.py).

* AO Hessians and gradients of the Laplacian against central differences of the value/gradient/
  Laplacian routine, and value/gradient/Laplacian against it exactly;
* dE_L/dr, dE_L/dR against five-point central differences (h = 1e-3 and 5e-4, Richardson-combined) of the
  independent E_L of ``oracle.coulomb.local_energy``, with the nuclei moved in every data object
  that holds them (AO centres, J1, Coulomb/ECP);
* d ln|Psi|/d(r, R) the same way from ``oracle.wf.ln_wavefunction``;
* translation invariance sum_a dE/dR_a + sum_i dE/dr_i = 0, and the same for ln|Psi|;
* the space-warp weights: sum over nuclei = 1, sum_i grad omega against differences.
"""

import copy
import dataclasses

import numpy as np
import pytest

from jqmc_amd import synthetic
from jqmc_amd.structure import Structure_data
from oracle import coulomb as oc
from oracle import forces as of
from oracle import rng as orng
from oracle import wf as owf

H = 1e-3


def _stencil(f):
    """Five-point central difference at h and h/2, Richardson-combined (error O(h^6))."""

    def d(h):
        return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)

    return (16.0 * d(H / 2) - d(H)) / 15.0


def _structures(obj, seen=None):
    """Every Structure_data reachable from obj (dataclass fields, lists, tuples)."""
    seen = {} if seen is None else seen
    if isinstance(obj, Structure_data):
        seen[id(obj)] = obj
    elif dataclasses.is_dataclass(obj) and not isinstance(obj, type):
        for f in dataclasses.fields(obj):
            _structures(getattr(obj, f.name), seen)
    elif isinstance(obj, (list, tuple)):
        for x in obj:
            _structures(x, seen)
    return seen


def _moved(ham, a, c, delta):
    h = copy.deepcopy(ham)
    for st in _structures(h).values():
        st.positions = st.positions.copy()
        st.positions[a, c] += delta
    return h


def _scaled_err(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / max(1.0, float(np.max(np.abs(b))))


CASES = {
    "water_jsd": lambda: synthetic.water_ae(0),
    "water_jagp_pade": lambda: synthetic.water_ae(3, jagp=True, j1=("pade", 0.9), j2=("exp", 1.1)),
    "n2_ecp": lambda: synthetic.chain_ecp(10, seed=1),
}


def test_ao_third_derivatives_match_differences():
    system = synthetic.chain_ecp(10, seed=1)
    aos = system.hamiltonian_data.wavefunction_data.jastrow_data.jastrow_three_body_data.orb_data
    rng = np.random.default_rng(0)
    pts = rng.normal(0.0, 0.9, size=(4, 3))
    v, g, Hs, gl = of.aos_d3(aos, pts)
    v0, gx, gy, gz, lap = owf.aos_vgl(aos, pts)
    np.testing.assert_allclose(v, v0, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(g, np.stack([gx, gy, gz]), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(Hs[0, 0] + Hs[1, 1] + Hs[2, 2], lap, rtol=1e-11, atol=1e-12)
    h = 1e-5
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        _, gxp, gyp, gzp, lp = owf.aos_vgl(aos, pts + e)
        _, gxm, gym, gzm, lm = owf.aos_vgl(aos, pts - e)
        dg = (np.stack([gxp, gyp, gzp]) - np.stack([gxm, gym, gzm])) / (2 * h)
        assert _scaled_err(Hs[:, c], dg) < 1e-7
        assert _scaled_err(gl[c], (lp - lm) / (2 * h)) < 1e-7


def test_spherical_ao_third_derivatives_match_differences():
    from test_golden_turborvb import build_case

    st, cp, wf, _, _, _ = build_case("test_comparison_with_TurboRVB_w_2b_3b_Jastrow_w_ecp")
    aos = wf.geminal_data.orb_data_up_spin
    aos = getattr(aos, "aos_data", aos)
    rng = np.random.default_rng(1)
    pts = st.positions[0] + rng.normal(0.0, 0.7, size=(3, 3))
    v, g, Hs, gl = of.aos_d3(aos, pts)
    v0, gx, gy, gz, lap = owf.aos_vgl(aos, pts)
    np.testing.assert_allclose(v, v0, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(Hs[0, 0] + Hs[1, 1] + Hs[2, 2], lap, rtol=1e-10, atol=1e-11)
    h = 1e-5
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        lp = owf.aos_vgl(aos, pts + e)[4]
        lm = owf.aos_vgl(aos, pts - e)[4]
        assert _scaled_err(gl[c], (lp - lm) / (2 * h)) < 1e-7


def _config(system, seed=4):
    r_up, r_dn = system.initial_walkers(1, seed=seed, sigma=0.8)
    RT = orng.rotation_matrix_T(*orng.rotation_angles(3, 0, 2))
    return r_up[0], r_dn[0], RT


@pytest.mark.parametrize("name", sorted(CASES))
def test_local_energy_gradients_match_differences(name):
    system = CASES[name]()
    ham = system.hamiltonian_data
    ru, rd, RT = _config(system)
    out = of.local_energy_position_grads(ham, ru, rd, RT)
    assert out["e_L"] == pytest.approx(oc.local_energy(ham, ru, rd, RT), rel=1e-11, abs=1e-11)
    nu = len(ru)
    X = np.vstack([ru, rd])
    fd_r = np.zeros_like(X)
    for e in range(len(X)):
        for c in range(3):
            def f(d, e=e, c=c):
                Y = X.copy()
                Y[e, c] += d
                return oc.local_energy(ham, Y[:nu], Y[nu:], RT)

            fd_r[e, c] = _stencil(f)
    assert _scaled_err(out["dE_dr"], fd_r) < 1e-6
    n_nuc = ham.structure_data.positions.shape[0]
    fd_R = np.zeros((n_nuc, 3))
    for a in range(n_nuc):
        for c in range(3):
            fd_R[a, c] = _stencil(lambda d, a=a, c=c: oc.local_energy(_moved(ham, a, c, d), ru, rd, RT))
    assert _scaled_err(out["dE_dR"], fd_R) < 1e-6
    # translation invariance
    tot = out["dE_dR"].sum(axis=0) + out["dE_dr"].sum(axis=0)
    assert np.max(np.abs(tot)) < 1e-9 * max(1.0, np.max(np.abs(out["dE_dr"])))


@pytest.mark.parametrize("name", ["water_jagp_pade", "n2_ecp"])
def test_ln_psi_gradients_match_differences(name):
    system = CASES[name]()
    ham = system.hamiltonian_data
    wf = ham.wavefunction_data
    ru, rd, RT = _config(system, seed=6)
    out = of.local_energy_position_grads(ham, ru, rd, RT)
    nu = len(ru)
    X = np.vstack([ru, rd])
    fd_r = np.zeros_like(X)
    for e in range(len(X)):
        for c in range(3):
            def f(d, e=e, c=c):
                Y = X.copy()
                Y[e, c] += d
                return owf.ln_wavefunction(wf, Y[:nu], Y[nu:])

            fd_r[e, c] = _stencil(f)
    assert _scaled_err(out["dlnpsi_dr"], fd_r) < 1e-8
    n_nuc = ham.structure_data.positions.shape[0]
    fd_R = np.zeros((n_nuc, 3))
    for a in range(n_nuc):
        for c in range(3):
            fd_R[a, c] = _stencil(
                lambda d, a=a, c=c: owf.ln_wavefunction(_moved(ham, a, c, d).wavefunction_data, ru, rd))
    assert _scaled_err(out["dlnpsi_dR"], fd_R) < 1e-8
    tot = out["dlnpsi_dR"].sum(axis=0) + out["dlnpsi_dr"].sum(axis=0)
    assert np.max(np.abs(tot)) < 1e-10 * max(1.0, np.max(np.abs(out["dlnpsi_dr"])))


def test_swct_weights():
    rng = np.random.default_rng(2)
    R = rng.normal(0.0, 2.0, size=(4, 3))
    r = rng.normal(0.0, 2.0, size=(7, 3))
    om = of.swct_omega(R, r)
    np.testing.assert_allclose(om.sum(axis=0), 1.0, rtol=1e-14)
    dom = of.swct_domega(R, r)
    fd = np.zeros_like(dom)
    for i in range(len(r)):
        for c in range(3):
            def f(d, i=i, c=c):
                y = r.copy()
                y[i, c] += d
                return of.swct_omega(R, y)[:, i]

            fd[:, c] += _stencil(f)
    assert _scaled_err(dom, fd) < 1e-9
    # the weights are translation invariant: sum over nuclei of the gradient sums vanishes
    np.testing.assert_allclose(dom.sum(axis=0), 0.0, atol=1e-12)

