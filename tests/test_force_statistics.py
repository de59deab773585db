"""Host-side force statistics (jqmc_amd.jqmc_mcmc.jackknife_forces) against a direct
leave-one-bin-out loop of the formula in get_aF (jqmc_mcmc.py:1191-1370)."""

import numpy as np

from jqmc_amd.jqmc_mcmc import jackknife_forces


def test_jackknife_forces_matches_loop():
    rng = np.random.default_rng(0)
    M, W, A, nb = 23, 3, 2, 5
    w = rng.random((M, W)) + 0.5
    e = rng.normal(-10.0, 1.0, (M, W))
    fhf = rng.normal(size=(M, W, A, 3))
    fpp = rng.normal(size=(M, W, A, 3))
    efp = e[..., None, None] * fpp
    mean, std = jackknife_forces(w, e, fhf, fpp, efp, nb)
    bins = np.array_split(np.arange(M), nb)
    est = []
    for b in bins:
        for k in range(W):
            keep = np.ones((M, W), dtype=bool)
            keep[b, k] = False
            ww = w * keep
            S = ww.sum()
            E = (ww * e).sum() / S
            HF = np.einsum("mw,mwac->ac", ww, fhf) / S
            PP = np.einsum("mw,mwac->ac", ww, fpp) / S
            EPP = np.einsum("mw,mwac->ac", ww, efp) / S
            est.append(-(HF + 2.0 * (EPP - E * PP)))
    est = np.asarray(est)
    np.testing.assert_allclose(mean, est.mean(axis=0), rtol=1e-12, atol=1e-14)
    n = len(est)
    np.testing.assert_allclose(std, np.sqrt((n - 1) / n * np.sum((est - est.mean(axis=0)) ** 2, axis=0)),
                               rtol=1e-10)
