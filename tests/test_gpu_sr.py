"""GPU parity of the SR kernels (jqmc_sr_colstats, jqmc_sr_gram on the FP64 matrix cores,
jqmc_sr_gram_f32 on the FP32 matrix cores, jqmc_sr_project) and of jqmc_amd.sr.natural_gradient
against the FP64 oracle (oracle/sr.py, pinned in test_oracle_sr.py), both branches, with the Gram
from the library GEMM and from the fused kernels; and the MCMC driver's O_k sampling feeding it.

Tolerances: the Gram, column statistics and projections within 1e-12 of their largest element
(FP64 sums over up to 10^4 terms in a different order); theta within 1e-9 relative to its largest
element (one FP64 Cholesky solve of a matrix with condition number up to ~10^6).  The FP32 Gram
within the elementwise rounding bound (m + 4) u (|Y| |Y|^T)_ij, u = 2^-24, of rounding Y to FP32
and accumulating m products in FP32; theta with the FP32 Gram as the preconditioner of an FP64
conjugate-gradient solve within 1e-8 of the oracle's.
"""

import numpy as np
import pytest
import torch

from jqmc_amd import _lib
from jqmc_amd._api import _ptr, stream_handle
from jqmc_amd.sr import natural_gradient
from oracle import sr as osr

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b))) / max(1e-300, float(np.max(np.abs(b))))


@pytest.mark.parametrize("n,m,shift", [(64, 16, "col"), (100, 37, "row"), (257, 300, "both"), (5, 3, "none")])
def test_gram_matches_numpy(n, m, shift):
    rng = np.random.default_rng(n + m)
    M = rng.normal(size=(n, m))
    a = rng.normal(size=n) if shift in ("row", "both") else None
    b = rng.normal(size=m) if shift in ("col", "both") else None
    u, v = rng.uniform(0.5, 2.0, size=n), rng.uniform(0.5, 2.0, size=m)
    Y = (M - (a[:, None] if a is not None else 0.0) - (b[None, :] if b is not None else 0.0)) * u[:, None] * v[None, :]
    ref = Y @ Y.T
    Md, ad, bd, ud, vd = _dev(M), _dev(a) if a is not None else None, _dev(b) if b is not None else None, _dev(u), _dev(v)
    G = torch.full((n, n), np.nan, dtype=torch.float64, device="cuda")
    _lib.call("jqmc_sr_gram", _ptr(Md), n, m, m, _ptr(ad), _ptr(bd), _ptr(ud), _ptr(vd), _ptr(G), n, stream_handle())
    g = G.cpu().numpy()
    assert np.all(np.isfinite(g)), "every element of G is written"
    assert _rel(g, ref) < 1e-12
    assert np.array_equal(g, g.T)


@pytest.mark.parametrize("n,m,shift", [(64, 16, "col"), (100, 37, "row"), (257, 300, "both"), (5, 3, "none"),
                                       (300, 1000, "both")])
def test_gram_f32_within_rounding_bound(n, m, shift):
    rng = np.random.default_rng(7 * n + m)
    M = rng.normal(size=(n, m)) + 3.0
    a = rng.normal(size=n) if shift in ("row", "both") else None
    b = rng.normal(size=m) + 3.0 if shift in ("col", "both") else None
    u, v = rng.uniform(0.5, 2.0, size=n), rng.uniform(0.5, 2.0, size=m)
    Y = (M - (a[:, None] if a is not None else 0.0) - (b[None, :] if b is not None else 0.0)) * u[:, None] * v[None, :]
    ref = Y @ Y.T
    bound = (m + 4) * 2.0**-24 * (np.abs(Y) @ np.abs(Y).T) + 1e-300
    Md, ad, bd, ud, vd = _dev(M), _dev(a) if a is not None else None, _dev(b) if b is not None else None, _dev(u), _dev(v)
    G = torch.full((n, n), np.nan, dtype=torch.float64, device="cuda")
    _lib.call("jqmc_sr_gram_f32", _ptr(Md), n, m, m, _ptr(ad), _ptr(bd), _ptr(ud), _ptr(vd), _ptr(G), n,
              stream_handle())
    g = G.cpu().numpy()
    assert np.all(np.isfinite(g)), "every element of G is written"
    assert np.all(np.abs(g - ref) <= bound), float(np.max(np.abs(g - ref) / bound))
    assert np.array_equal(g, g.T)
    # FP32 accumulation is visible: the result is not the FP64 Gram
    if m >= 16:
        assert not np.array_equal(g, ref)


def test_colstats_and_project_match_numpy():
    rng = np.random.default_rng(1)
    S, P = 300, 517
    O = rng.normal(size=(S, P)) + rng.normal(size=P)
    wn = rng.uniform(0.2, 1.0, size=S)
    wn /= wn.sum()
    mean = torch.empty(P, dtype=torch.float64, device="cuda")
    var = torch.empty_like(mean)
    Od, wd = _dev(O), _dev(wn)
    _lib.call("jqmc_sr_colstats", _ptr(Od), S, P, P, _ptr(wd), _ptr(mean), _ptr(var), stream_handle())
    m_ref = wn @ O
    v_ref = wn @ (O - m_ref) ** 2
    assert _rel(mean.cpu().numpy(), m_ref) < 1e-13
    assert _rel(var.cpu().numpy(), v_ref) < 1e-12
    rs, cs, x = rng.uniform(0.5, 1.5, size=S), rng.uniform(0.5, 1.5, size=P), rng.normal(size=S)
    out = torch.empty(P, dtype=torch.float64, device="cuda")
    args = [_dev(t) for t in (m_ref, rs, cs, x)]  # held until the kernel has run
    _lib.call("jqmc_sr_project", _ptr(Od), S, P, P, *[_ptr(t) for t in args], _ptr(out), stream_handle())
    torch.cuda.synchronize()
    ref = cs * (((O - m_ref) * (rs * x)[:, None]).sum(axis=0))
    assert _rel(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("S,P", [(200, 70), (150, 600), (96, 96)])
@pytest.mark.parametrize("branch", ["auto", "primal", "push_through"])
@pytest.mark.parametrize("gram_kernel", ["blas", "fused"])
def test_natural_gradient_matches_oracle(S, P, branch, gram_kernel):
    rng = np.random.default_rng(S * P)
    O = rng.normal(size=(S, P)) @ rng.normal(size=(P, P)) * (0.2 / np.sqrt(P)) + rng.normal(size=P)
    O[:, 5] = 0.75  # a frozen parameter
    w = rng.uniform(0.5, 1.5, size=S)
    e = rng.normal(-30.0, 2.0, size=S) + 0.3 * O[:, 1]
    ref = osr.natural_gradient(O, w, e, epsilon=1e-3, branch=branch)
    out = natural_gradient(O, w, e, epsilon=1e-3, branch=branch, gram_kernel=gram_kernel)
    assert out["branch"] == ref["branch"]
    assert out["frozen"].cpu().numpy().tolist() == ref["frozen"].tolist()
    assert _rel(out["O_bar"].cpu().numpy(), ref["O_bar"]) < 1e-13
    assert _rel(out["diag_S"].cpu().numpy(), ref["diag_S"]) < 1e-11
    assert _rel(out["theta"].cpu().numpy(), ref["theta"]) < 1e-9


@pytest.mark.parametrize("S,P", [(200, 70), (150, 600)])
def test_natural_gradient_cg_matches_oracle(S, P):
    """The conjugate-gradient path (the reference's default cg_flag) against the oracle's direct
    solve, both branches: theta within 1e-7 of its scale at ||r|| <= 1e-12; a warm start from the
    previous solution converges in no step."""
    rng = np.random.default_rng(S + 3 * P)
    O = rng.normal(size=(S, P)) @ rng.normal(size=(P, P)) * (0.2 / np.sqrt(P)) + rng.normal(size=P)
    O[:, 5] = 0.75
    w = rng.uniform(0.5, 1.5, size=S)
    e = rng.normal(-30.0, 2.0, size=S) + 0.3 * O[:, 1]
    for branch in ("primal", "push_through"):
        ref = osr.natural_gradient(O, w, e, epsilon=1e-3, branch=branch)
        out = natural_gradient(O, w, e, epsilon=1e-3, branch=branch, cg=True, cg_tol=1e-12)
        assert out["solver"] == "cg" and out["cg_residual"] <= 1e-12
        assert out["frozen"].cpu().numpy().tolist() == ref["frozen"].tolist()
        assert _rel(out["theta"].cpu().numpy(), ref["theta"]) < 1e-7, branch
        again = natural_gradient(O, w, e, epsilon=1e-3, branch=branch, cg=True, cg_tol=1e-10, x0=out["cg_solution"])
        assert again["cg_steps"] == 0


@pytest.mark.parametrize("S,P", [(200, 70), (150, 600), (300, 3000)])
@pytest.mark.parametrize("gram_kernel", ["blas", "fused"])
def test_natural_gradient_fp32_gram(S, P, gram_kernel):
    """gram='float32' on both branches: the FP32 Gram preconditions an FP64 conjugate-gradient solve
    stopped at ||r|| <= 1e-13 ||b||, so theta equals the oracle's FP64 theta within kappa * 1e-13
    relative (kappa of the shifted Gram of these uncorrelated samples, <= 1e4): 1e-8 is asserted."""
    rng = np.random.default_rng(11 * S + P)
    O = rng.normal(size=(S, P)) + rng.normal(size=P)
    O[:, 5] = 0.75
    w = rng.uniform(0.5, 1.5, size=S)
    e = rng.normal(-30.0, 2.0, size=S) + 0.3 * O[:, 1]
    for branch in ("primal", "push_through"):
        ref = osr.natural_gradient(O, w, e, epsilon=1e-3, branch=branch)
        out = natural_gradient(O, w, e, epsilon=1e-3, branch=branch, gram="float32", gram_kernel=gram_kernel)
        assert out["solver"] == "pcg_fp32_gram" and out["pcg_steps"] > 0
        assert out["preconditioner_shift"] >= 1e-3
        assert out["frozen"].cpu().numpy().tolist() == ref["frozen"].tolist()
        err = _rel(out["theta"].cpu().numpy(), ref["theta"])
        assert err < 1e-8, (branch, err, out["pcg_steps"])
    with pytest.raises(ValueError):
        natural_gradient(O, w, e, gram="bfloat16")
    with pytest.raises(ValueError):
        natural_gradient(O, w, e, gram_kernel="tensile")


def test_mcmc_param_derivs_feed_sr():
    from jqmc_amd import synthetic
    from jqmc_amd.jqmc_mcmc import MCMC

    ham = synthetic.water_ae(0).hamiltonian_data
    mc = MCMC(hamiltonian_data=ham, mcmc_seed=5, num_walkers=16, num_mcmc_per_measurement=8, Dt=0.5,
              epsilon_AS=1e-4, comput_param_deriv=True)
    mc.run(6)
    O = mc.get_dln_WF(num_mcmc_warmup_steps=1)
    assert O.shape[:2] == (5, 16) and np.all(np.isfinite(O))
    out = mc.get_natural_gradient(num_mcmc_warmup_steps=1)
    ref = osr.natural_gradient(O.reshape(-1, O.shape[2]), mc.w_L[1:].ravel(), mc.e_L[1:].ravel())
    assert out["branch"] == ref["branch"] == "push_through"
    assert _rel(out["theta"].cpu().numpy(), ref["theta"]) < 1e-8
