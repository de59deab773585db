"""ln|Psi| as a differentiable function of the variational parameters.

``ln_psi_params(ws)`` returns per-walker ln|Psi| = ln|det G| + J (from the walker state) and the
parameter tensors it depends on; backpropagating through it calls the hand-written O_k kernel
(``jqmc_dln_psi_dparams``): the vector-Jacobian product is sum_w g_w O_k[w].  This is the
gradient the reference takes with ``grad(evaluate_ln_wavefunction_fast, argnums=0)``
(jqmc_mcmc.py:854-873), e.g. for the energy gradient 2 <(E_L - E) O_k>.  The per-walker O_k
matrix itself (the SR input) is ``WalkerState.dln_psi_dparams``.
"""

from __future__ import annotations

import torch

from ._api import WalkerState

BLOCKS = ("j1_param", "j2_param", "j3_matrix", "lambda_matrix")


class LnPsiParams(torch.autograd.Function):
    """forward(ws, names, *params) -> ln|Psi| (W,); backward -> sum_w g_w O_k[w] per block."""

    @staticmethod
    def forward(ctx, ws: WalkerState, names: tuple, *params):
        ctx.ws = ws
        ctx.names = names
        return (ws.logabsdet + ws.jastrow).clone()

    @staticmethod
    def backward(ctx, g):
        O = ctx.ws.dln_psi_dparams(ctx.names)
        grads = []
        for name in ctx.names:
            blk = O[name]
            grads.append(torch.tensordot(g.to(blk.dtype), blk, dims=([0], [0])))
        return (None, None, *grads)


def parameter_tensors(ham, device=None) -> dict:
    """Leaf tensors of the variational blocks present in ``ham`` (requires_grad set)."""
    wf = ham.wavefunction_data
    jas = wf.jastrow_data
    out = {}
    kw = dict(dtype=torch.float64, device=device)
    if jas.jastrow_one_body_data is not None:
        out["j1_param"] = torch.tensor(jas.jastrow_one_body_data.jastrow_1b_param, **kw)
    if jas.jastrow_two_body_data is not None:
        out["j2_param"] = torch.tensor(jas.jastrow_two_body_data.jastrow_2b_param, **kw)
    if jas.jastrow_three_body_data is not None:
        out["j3_matrix"] = torch.tensor(jas.jastrow_three_body_data.j_matrix, **kw)
    out["lambda_matrix"] = torch.tensor(wf.geminal_data.lambda_matrix, **kw)
    for t in out.values():
        t.requires_grad_(True)
    return out


def ln_psi_params(ws: WalkerState, ham) -> tuple[torch.Tensor, dict]:
    """(ln|Psi| per walker, parameter leaf tensors) with the O_k kernel as backward."""
    params = parameter_tensors(ham, device=ws.logabsdet.device)
    names = tuple(n for n in BLOCKS if n in params)
    lnpsi = LnPsiParams.apply(ws, names, *(params[n] for n in names))
    return lnpsi, params
