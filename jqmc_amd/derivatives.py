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


class LocalEnergyPositions(torch.autograd.Function):
    """forward(ws, RT, R, r) -> e_L (W,) at the walkers' positions; backward -> (sum_w g_w dE_L/dR_w,
    g_w dE_L/dr_w) from the position-derivative kernel (jqmc_position_grads), the gradients the
    reference takes with ``grad(compute_local_energy, argnums=(0, 1, 2))`` (jqmc_mcmc.py:744-780).
    R (n_nuc, 3) and r (W, N, 3) stand for the nuclear and electron positions the kernel reads from
    the system and the walker state; their values are not read."""

    @staticmethod
    def forward(ctx, ws: WalkerState, RT, R, r):
        out = ws.position_grads(RT, swct=False)
        ctx.save_for_backward(out["dE_dR"], out["dE_dr"])
        return out["e_L"]

    @staticmethod
    def backward(ctx, g):
        dR, dr = ctx.saved_tensors
        g = g.to(dR.dtype)
        return None, None, torch.tensordot(g, dR, dims=([0], [0])), g[:, None, None] * dr


def local_energy_positions(ws: WalkerState, RT=None) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(e_L per walker, nuclear positions leaf (n_nuc, 3), electron positions leaf (W, N, 3)) with the
    adjoint kernel as backward: ``e_L.sum().backward()`` leaves -sum of the forces in ``R.grad``."""
    R = torch.as_tensor(ws.pack._keep["nuc_pos"], dtype=torch.float64, device=ws.r_up.device).clone().requires_grad_(True)
    r = torch.cat([ws.r_up, ws.r_dn], dim=1).clone().requires_grad_(True)
    e_L = LocalEnergyPositions.apply(ws, RT, R, r)
    return e_L, R, r


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
