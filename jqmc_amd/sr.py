"""Stochastic-reconfiguration (natural-gradient) step on the GPU: the SR branch of
``MCMC.run_optimize`` (jqmc_mcmc.py:2952-3350) for one process.

The (S x P) matrix of per-sample log-derivatives O stays on the device.  Column means and diag S
come from ``jqmc_sr_colstats``, X F and the push-through theta = X x from ``jqmc_sr_project``, and
the S x S (or P x P) solve is a library FP64 Cholesky.  The Gram matrix of the normalised X is a
plain GEMM once X is formed: by default (``gram_kernel="blas"``) X is written out and the product
runs in the ROCm BLAS library on the matrix cores (measured at 87 % of the FP64 matrix peak at
S = 4096, P = 52 439); ``gram_kernel="fused"`` uses this library's own kernels ``jqmc_sr_gram`` /
``jqmc_sr_gram_f32``, which apply the normalisation on the way into LDS and need no copy of X.
Same normalisation, freezing and branch choice as the reference; the result equals the
reference's theta to rounding.  ``gram="float32"`` forms the Gram in FP32 on the matrix cores and
uses its Cholesky factor (epsilon shift in FP64, enlarged tenfold until the factor exists: the FP32
rounding of a Gram of 10^4-10^5 terms can exceed epsilon) as the preconditioner of an FP64
conjugate-gradient solve of the reference's system with matrix-free FP64 products, so that theta
is the FP64 one to the stated tolerance while the O(S^2 P) work runs in FP32.
"""

from __future__ import annotations

import torch

from . import _lib
from ._api import F64, _ptr, as_dev, stream_handle

MIN_S_DIAG_ABS = 1.0e-10  # _setting.py:149


def _gram(M, n, m, a, b, u, v, gram="float64"):
    if gram not in ("float64", "float32"):
        raise ValueError(f"gram must be 'float64' or 'float32', not {gram!r}")
    G = torch.empty((n, n), dtype=F64, device=M.device)
    name = "jqmc_sr_gram" if gram == "float64" else "jqmc_sr_gram_f32"
    _lib.call(name, _ptr(M), n, m, M.stride(0), _ptr(a), _ptr(b), _ptr(u), _ptr(v), _ptr(G), n, stream_handle())
    return G


def _gram_blas(O, mean, rs, cs, branch: str, gram: str = "float64"):
    """X^T X (push-through, S x S) or X X^T (primal, P x P) of X_ki = (O_ik - mean_k) rs_i cs_k by a
    library GEMM on the written-out Y = X^T (S x P), in FP64 or with Y rounded to FP32."""
    Y = O - mean
    Y.mul_(rs[:, None]).mul_(cs[None, :])
    if gram == "float32":
        Y = Y.to(torch.float32)
    G = Y @ Y.t() if branch == "push_through" else Y.t() @ Y
    return G.to(F64)


def _project(O, mean, rs, cs, x):
    S, P = O.shape
    out = torch.empty(P, dtype=F64, device=O.device)
    _lib.call("jqmc_sr_project", _ptr(O), S, P, O.stride(0), _ptr(mean), _ptr(rs), _ptr(cs), _ptr(x), _ptr(out),
              stream_handle())
    return out


def _spd_solve(A, rhs):
    L = torch.linalg.cholesky(A)
    return torch.cholesky_solve(rhs[:, None], L)[:, 0]


def _shifted_cholesky(G, epsilon: float):
    """Cholesky factor of G + s I for the smallest s in epsilon * 10^k (k = 0, 1, ...) that has one;
    returns (factor, s)."""
    n = G.shape[0]
    eye = torch.eye(n, dtype=G.dtype, device=G.device)
    s = float(epsilon)
    for _ in range(12):
        L, info = torch.linalg.cholesky_ex(G + s * eye)
        if int(info) == 0:
            return L, s
        s *= 10.0
    raise FloatingPointError("the FP32 Gram has no Cholesky factor for any shift up to 1e12 epsilon")


def _pcg(b, apply_A, precond, max_iter: int, tol: float):
    """Preconditioned conjugate gradients for the SPD system A x = b from x = 0, preconditioner
    z = precond(r): stop when ||r|| <= tol or after max_iter steps.  Returns (x, ||r||, steps)."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = precond(r)
    p = z.clone()
    rz = float(torch.dot(r, z))
    res = float(torch.linalg.vector_norm(r))
    steps = 0
    while res > tol and steps < max_iter:
        Ap = apply_A(p)
        alpha = rz / float(torch.dot(p, Ap))
        x = x + alpha * p
        r = r - alpha * Ap
        res = float(torch.linalg.vector_norm(r))
        steps += 1
        if res != res:
            raise FloatingPointError("Non-finite residual encountered in the preconditioned SR solve.")
        z = precond(r)
        rz_new = float(torch.dot(r, z))
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, res, steps


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
                     x0=None, gram: str = "float64", gram_kernel: str = "blas") -> dict:
    """theta = (S + epsilon I)^-1 f of jqmc_mcmc.py:2952-3362 in its scale-invariant form.

    O: (S, P) log-derivatives per sample (or (num_mcmc, W, P), flattened like the reference),
    w_L, e_L: (S,) weights and local energies.  branch: 'auto' (push-through when P >= S, as the
    reference), 'primal' or 'push_through'.  cg=True solves the branch's system with conjugate
    gradients (the reference's cg_flag path, 3160-3196 and 3300-3350): matrix-free products with the
    normalised X from library GEMVs over O, stopped at ||r|| <= cg_tol, started from x0 (a warm start
    of the branch's shape) or zero.  gram: 'float64' (the reference's direct solve) or 'float32' (the
    FP32 matrix-core Gram as the preconditioner of an FP64 conjugate-gradient solve, stopped at
    ||r|| <= 1e-13 ||b|| or after 500 steps; the result also carries pcg_steps, pcg_residual and the
    preconditioner's shift).  gram_kernel: 'blas' (X written out, library GEMM) or 'fused' (this
    library's Gram kernels, no copy of X).  Returns device tensors: theta (P,), frozen (P,) bool, diag_S
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
    if gram not in ("float64", "float32"):
        raise ValueError(f"gram must be 'float64' or 'float32', not {gram!r}")
    if gram_kernel not in ("blas", "fused"):
        raise ValueError(f"gram_kernel must be 'blas' or 'fused', not {gram_kernel!r}")

    def gram_matrix(prec):
        if gram_kernel == "blas":
            return _gram_blas(O, mean, rs, cs, branch, prec)
        if branch == "push_through":
            return _gram(O, S, P, None, mean, rs, cs, prec)
        return _gram(O.t().contiguous(), P, S, mean, None, cs, rs, prec)

    # X (P x S): X[k, i] = (O[i, k] - O_bar[k]) rs[i] cs[k]
    def X_times(v):  # (S,) -> (P,)
        u = rs * v
        return cs * (O.t() @ u - mean * u.sum())

    def XT_times(v):  # (P,) -> (S,)
        u = cs * v
        return rs * (O @ u - torch.dot(mean, u))

    if gram == "float32" and not cg:
        if branch == "push_through":
            b = F
            apply_A = lambda v: XT_times(X_times(v)) + epsilon * v  # noqa: E731
        elif branch == "primal":
            b = X_times(F)
            apply_A = lambda v: X_times(XT_times(v)) + epsilon * v  # noqa: E731
        else:
            raise ValueError(f"unknown branch {branch!r}")
        G = gram_matrix("float32")
        L, shift = _shifted_cholesky(G, epsilon)
        del G
        precond = lambda r: torch.cholesky_solve(r[:, None], L)[:, 0]  # noqa: E731
        tol = 1e-13 * float(torch.linalg.vector_norm(b))
        sol, res, steps = _pcg(b, apply_A, precond, 500, tol)
        theta = X_times(sol) if branch == "push_through" else sol
        theta = torch.where(frozen, torch.zeros_like(theta), theta * cs)
        return {"theta": theta, "frozen": frozen, "diag_S": diag_S, "O_bar": mean, "e_bar": float(e_bar),
                "branch": branch, "solver": "pcg_fp32_gram", "pcg_steps": steps, "pcg_residual": res,
                "preconditioner_shift": shift}
    if cg:
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
    if branch not in ("push_through", "primal"):
        raise ValueError(f"unknown branch {branch!r}")
    G = gram_matrix("float64")                        # X^T X (S x S) or X X^T (P x P)
    G.diagonal().add_(epsilon)
    if branch == "push_through":
        x = _spd_solve(G, F)
        theta = _project(O, mean, rs, cs, x)          # X (X^T X + eps)^-1 F
    else:
        theta = _spd_solve(G, _project(O, mean, rs, cs, F))
    theta = theta * cs                                # back to the original scale
    theta = torch.where(frozen, torch.zeros_like(theta), theta)
    return {"theta": theta, "frozen": frozen, "diag_S": diag_S, "O_bar": mean, "e_bar": float(e_bar),
            "branch": branch, "solver": "direct"}
