"""Stochastic-reconfiguration (natural-gradient) step on the GPU: the SR branch of
``MCMC.run_optimize`` (jqmc_mcmc.py:2952-3350) for one process.

The (S x P) matrix of per-sample log-derivatives O stays on the device.  Column means and diag S
come from ``jqmc_sr_colstats``, the Gram matrix of the normalised X from ``jqmc_sr_gram`` (FP64
matrix cores), X F and the push-through theta = X x from ``jqmc_sr_project``; only the S x S (or
P x P) solve is a library FP64 Cholesky.  Same normalisation, freezing and branch choice as the
reference; the result equals the reference's theta to rounding.
"""

from __future__ import annotations

import torch

from . import _lib
from ._api import F64, _ptr, as_dev, stream_handle

MIN_S_DIAG_ABS = 1.0e-10  # _setting.py:149


def _gram(M, n, m, a, b, u, v):
    G = torch.empty((n, n), dtype=F64, device=M.device)
    _lib.call("jqmc_sr_gram", _ptr(M), n, m, M.stride(0), _ptr(a), _ptr(b), _ptr(u), _ptr(v), _ptr(G), n,
              stream_handle())
    return G


def _project(O, mean, rs, cs, x):
    S, P = O.shape
    out = torch.empty(P, dtype=F64, device=O.device)
    _lib.call("jqmc_sr_project", _ptr(O), S, P, O.stride(0), _ptr(mean), _ptr(rs), _ptr(cs), _ptr(x), _ptr(out),
              stream_handle())
    return out


def _spd_solve(A, rhs):
    L = torch.linalg.cholesky(A)
    return torch.cholesky_solve(rhs[:, None], L)[:, 0]


def _cg(b, apply_A, x0, max_iter: int, tol: float):
    """Conjugate gradient for the SPD system A x = b (_conjugate_gradient_numpy, jqmc_mcmc.py:2564-2612):
    stop when ||r|| <= tol, on a non-finite or vanishing p.Ap, or after max_iter steps.  Returns
    (x, ||r||, steps)."""
    x = x0.clone()
    r = b - apply_A(x)
    p = r.clone()
    rs_old = float(torch.dot(r, r))
    if not torch.isfinite(torch.tensor(rs_old)):
        raise FloatingPointError("Non-finite initial residual encountered in SR-CG.")
    if rs_old**0.5 <= tol:
        return x, rs_old**0.5, 0
    tiny = torch.finfo(F64).tiny
    steps = 0
    for i in range(int(max_iter)):
        Ap = apply_A(p)
        denom = float(torch.dot(p, Ap))
        if not (abs(denom) > tiny) or denom != denom or abs(denom) == float("inf"):
            break
        alpha = rs_old / denom
        x = x + alpha * p
        r = r - alpha * Ap
        rs_new = float(torch.dot(r, r))
        steps = i + 1
        if rs_new != rs_new or rs_new == float("inf"):
            raise FloatingPointError("Non-finite residual encountered in SR-CG.")
        if rs_new**0.5 <= tol:
            rs_old = rs_new
            break
        p = r + (rs_new / rs_old) * p
        rs_old = rs_new
    return x, rs_old**0.5, steps


def natural_gradient(O, w_L, e_L, epsilon: float = 1.0e-3, min_S_diag_abs: float = MIN_S_DIAG_ABS,
                     branch: str = "auto", cg: bool = False, cg_max_iter: int = 1_000_000, cg_tol: float = 1.0e-8,
                     x0=None) -> dict:
    """theta = (S + epsilon I)^-1 f of jqmc_mcmc.py:2952-3362 in its scale-invariant form.

    O: (S, P) log-derivatives per sample (or (num_mcmc, W, P), flattened like the reference),
    w_L, e_L: (S,) weights and local energies.  branch: 'auto' (push-through when P >= S, as the
    reference), 'primal' or 'push_through'.  cg=True solves the branch's system with conjugate
    gradients (the reference's cg_flag path, 3160-3196 and 3300-3350): matrix-free products with the
    normalised X from library GEMVs over O, stopped at ||r|| <= cg_tol, started from x0 (a warm start
    of the branch's shape) or zero.  Returns device tensors: theta (P,), frozen (P,) bool, diag_S
    (P,) after the floor, O_bar (P,), and e_bar (float); with cg also cg_residual, cg_steps and
    cg_solution (the warm start for the next call).
    """
    O = as_dev(O)
    if O.dim() == 3:
        O = O.reshape(-1, O.shape[-1])
    O = O.contiguous()
    w = as_dev(w_L).reshape(-1).contiguous()
    e = as_dev(e_L).reshape(-1).contiguous()
    S, P = O.shape
    if w.numel() != S or e.numel() != S:
        raise ValueError("w_L and e_L need one entry per row of O")
    wn = (w / w.sum()).contiguous()
    mean = torch.empty(P, dtype=F64, device=O.device)
    var = torch.empty(P, dtype=F64, device=O.device)
    _lib.call("jqmc_sr_colstats", _ptr(O), S, P, O.stride(0), _ptr(wn), _ptr(mean), _ptr(var), stream_handle())
    e_bar = (wn * e).sum()
    rs = wn.sqrt().contiguous()                       # sqrt(w_i / sum w)
    F = (-2.0 * rs * (e - e_bar)).contiguous()        # jqmc_mcmc.py:3001-3003
    ok = torch.isfinite(var) & (var > min_S_diag_abs)
    frozen = ~ok
    diag_S = torch.where(ok, var, torch.full_like(var, min_S_diag_abs))
    cs = diag_S.rsqrt().contiguous()                  # X <- X / sqrt(diag_S)
    if branch == "auto":
        branch = "primal" if P < S else "push_through"
    if cg:
        # X (P x S): X[k, i] = (O[i, k] - O_bar[k]) rs[i] cs[k]
        def X_times(v):  # (S,) -> (P,)
            u = rs * v
            return cs * (O.t() @ u - mean * u.sum())

        def XT_times(v):  # (P,) -> (S,)
            u = cs * v
            return rs * (O @ u - torch.dot(mean, u))

        if branch == "push_through":
            b = F
            apply_A = lambda v: XT_times(X_times(v)) + epsilon * v  # noqa: E731
        elif branch == "primal":
            b = X_times(F)
            apply_A = lambda v: X_times(XT_times(v)) + epsilon * v  # noqa: E731
        else:
            raise ValueError(f"unknown branch {branch!r}")
        start = as_dev(x0).reshape(-1) if x0 is not None and as_dev(x0).numel() == b.numel() else torch.zeros_like(b)
        sol, res, steps = _cg(b, apply_A, start, cg_max_iter, cg_tol)
        theta = X_times(sol) if branch == "push_through" else sol
        theta = torch.where(frozen, torch.zeros_like(theta), theta * cs)
        return {"theta": theta, "frozen": frozen, "diag_S": diag_S, "O_bar": mean, "e_bar": float(e_bar),
                "branch": branch, "solver": "cg", "cg_residual": res, "cg_steps": steps, "cg_solution": sol}
    if branch == "push_through":
        G = _gram(O, S, P, None, mean, rs, cs)        # X^T X (S x S)
        G.diagonal().add_(epsilon)
        x = _spd_solve(G, F)
        theta = _project(O, mean, rs, cs, x)          # X (X^T X + eps)^-1 F
    elif branch == "primal":
        OT = O.t().contiguous()                       # rows = parameters
        G = _gram(OT, P, S, mean, None, cs, rs)       # X X^T (P x P)
        G.diagonal().add_(epsilon)
        theta = _spd_solve(G, _project(O, mean, rs, cs, F))
    else:
        raise ValueError(f"unknown branch {branch!r}")
    theta = theta * cs                                # back to the original scale
    theta = torch.where(frozen, torch.zeros_like(theta), theta)
    return {"theta": theta, "frozen": frozen, "diag_S": diag_S, "O_bar": mean, "e_bar": float(e_bar),
            "branch": branch, "solver": "direct"}
