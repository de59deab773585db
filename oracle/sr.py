"""Stochastic-reconfiguration step of MCMC.run_optimize (jqmc_mcmc.py:2952-3350), one process.

Test infrastructure only (see oracle/__init__.py): the FP64 NumPy restatement the GPU SR kernels
are checked against.

Given per-sample log-derivatives O (S x P), weights w (S) and local energies e_L (S):

* O_bar = sum_i w_i O_i / sum w, e_bar = sum_i w_i e_i / sum w           (2974-2987)
* X = ((O - O_bar) sqrt(w) / sqrt(sum w))^T (P x S), F = -2 sqrt(w) (e_L - e_bar) / sqrt(sum w)
                                                                         (2996-3003)
* diag_S = sum_i X_ki^2; parameters with diag_S not finite or <= min_S_diag_abs are frozen and
  their diag_S set to the floor min_S_diag_abs; X <- X / sqrt(diag_S)   (3062-3112)
* P < S: theta = (X X^T + eps I)^-1 X F  (direct solve, 3114-3146)
  P >= S: theta = X (X^T X + eps I)^-1 F (push-through, 3198-3300)
* theta <- theta / sqrt(diag_S); theta[frozen] = 0                       (3354-3362)

The two branches give the same theta (push-through identity), which tests/test_oracle_sr.py uses
as one of the pins.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

MIN_S_DIAG_ABS = 1.0e-10  # _setting.py:149


def sr_inputs(O, w, e_L):
    """(X (P x S), F (S), O_bar (P), e_bar) of jqmc_mcmc.py:2974-3003."""
    O = np.asarray(O, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    e_L = np.asarray(e_L, dtype=np.float64)
    tw = np.sum(w)
    O_bar = np.einsum("i,ij->j", w, O) / tw
    e_bar = float(np.dot(w, e_L) / tw)
    X = ((O - O_bar) * np.sqrt(w)[:, None] / np.sqrt(tw)).T
    F = -2.0 * np.sqrt(w) * (e_L - e_bar) / np.sqrt(tw)
    return X, F, O_bar, e_bar


def natural_gradient(O, w, e_L, epsilon: float = 1.0e-3, min_S_diag_abs: float = MIN_S_DIAG_ABS,
                     branch: str = "auto") -> dict:
    """theta = (S + eps I)^-1 f in the scale-invariant form, as the reference computes it.
    branch: 'auto' (the reference's choice by P < S), 'primal' (P x P) or 'push_through' (S x S)."""
    X, F, O_bar, e_bar = sr_inputs(O, w, e_L)
    P, S = X.shape
    diag_S = np.einsum("jk,kj->j", X, X.T)
    frozen = ~(np.isfinite(diag_S) & (diag_S > min_S_diag_abs))
    diag_S = np.where(np.isfinite(diag_S) & (diag_S > min_S_diag_abs), diag_S, min_S_diag_abs)
    Xn = X / np.sqrt(diag_S)[:, None]
    if branch == "auto":
        branch = "primal" if P < S else "push_through"
    if branch == "primal":
        A = Xn @ Xn.T
        A[np.diag_indices_from(A)] += epsilon
        theta = scipy.linalg.solve(A, Xn @ F, assume_a="sym")
    else:
        A = Xn.T @ Xn
        A[np.diag_indices_from(A)] += epsilon
        theta = Xn @ scipy.linalg.solve(A, F, assume_a="sym")
    theta = theta / np.sqrt(diag_S)
    theta[frozen] = 0.0
    return {"theta": theta, "frozen": frozen, "diag_S": diag_S, "O_bar": O_bar, "e_bar": e_bar,
            "X_F": X @ F, "branch": branch}
