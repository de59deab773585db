"""Pins of the SR oracle (oracle/sr.py, jqmc_mcmc.py:2952-3362).

* the two branches agree: X (X^T X + eps)^-1 F = (X X^T + eps)^-1 X F (push-through identity),
  for P < S and P > S;
* against the definitions written with explicit loops on a tiny case: weighted means, the
  covariance S_kl = <(O_k - O_bar_k)(O_l - O_bar_l)>_w, the force f_k = -2 <(e - e_bar)(O_k - O_bar_k)>_w,
  theta = D^-1/2 (D^-1/2 S D^-1/2 + eps)^-1 D^-1/2 f with D = diag S;
* invariances the normalisation fixes: theta_k scales as 1/c when column k of O is scaled by c;
  constant shifts of a column and a common factor of the weights leave theta unchanged;
* a zero-variance parameter is frozen (theta = 0) and does not disturb the others.
"""

import numpy as np
import pytest

from oracle import sr


def _case(S, P, seed=0):
    rng = np.random.default_rng(seed)
    O = rng.normal(size=(S, P)) @ rng.normal(size=(P, P)) * 0.3 + rng.normal(size=P)
    w = rng.uniform(0.5, 1.5, size=S)
    e = rng.normal(-10.0, 1.0, size=S) + O[:, 0] * 0.2
    return O, w, e


@pytest.mark.parametrize("S,P", [(40, 12), (12, 40), (30, 30)])
def test_branches_agree(S, P):
    O, w, e = _case(S, P, seed=S + P)
    a = sr.natural_gradient(O, w, e, epsilon=1e-3, branch="primal")["theta"]
    b = sr.natural_gradient(O, w, e, epsilon=1e-3, branch="push_through")["theta"]
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.max(np.abs(a)))


def test_against_definitions():
    S, P, eps = 9, 4, 1e-2
    O, w, e = _case(S, P, seed=3)
    tw = sum(w)
    Ob = [sum(w[i] * O[i, k] for i in range(S)) / tw for k in range(P)]
    eb = sum(w[i] * e[i] for i in range(S)) / tw
    Sm = np.array([[sum(w[i] * (O[i, k] - Ob[k]) * (O[i, l] - Ob[l]) for i in range(S)) / tw for l in range(P)]
                   for k in range(P)])
    f = np.array([-2.0 * sum(w[i] * (e[i] - eb) * (O[i, k] - Ob[k]) for i in range(S)) / tw for k in range(P)])
    d = np.diag(Sm)
    Dh = np.diag(1.0 / np.sqrt(d))
    theta = Dh @ np.linalg.solve(Dh @ Sm @ Dh + eps * np.eye(P), Dh @ f)
    out = sr.natural_gradient(O, w, e, epsilon=eps)
    np.testing.assert_allclose(out["O_bar"], Ob, rtol=1e-13)
    assert out["e_bar"] == pytest.approx(eb, rel=1e-14)
    np.testing.assert_allclose(out["diag_S"], d, rtol=1e-12)
    np.testing.assert_allclose(out["X_F"], f, rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(out["theta"], theta, rtol=1e-10)


def test_invariances():
    S, P = 30, 8
    O, w, e = _case(S, P, seed=5)
    t0 = sr.natural_gradient(O, w, e)["theta"]
    O2 = O.copy()
    O2[:, 3] *= 7.0
    np.testing.assert_allclose(sr.natural_gradient(O2, w, e)["theta"], t0 / np.r_[1, 1, 1, 7.0, 1, 1, 1, 1],
                               rtol=1e-10)
    O3 = O + np.arange(P) * 2.5
    np.testing.assert_allclose(sr.natural_gradient(O3, w, e)["theta"], t0, rtol=1e-9)
    np.testing.assert_allclose(sr.natural_gradient(O, 3.0 * w, e)["theta"], t0, rtol=1e-11)


def test_zero_variance_parameter_is_frozen():
    S, P = 25, 6
    O, w, e = _case(S, P, seed=8)
    O[:, 2] = 1.25
    out = sr.natural_gradient(O, w, e)
    assert out["frozen"].tolist() == [False, False, True, False, False, False]
    assert out["theta"][2] == 0.0
    keep = [0, 1, 3, 4, 5]
    ref = sr.natural_gradient(O[:, keep], w, e)["theta"]
    # the frozen column is exactly centred to zero, so it adds nothing to the Gram matrix
    np.testing.assert_allclose(out["theta"][keep], ref, rtol=1e-9)
