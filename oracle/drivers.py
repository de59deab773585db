"""FP64 restatement of the driver loops of ``MCMC.run`` and ``GFMC_n.run`` (test infrastructure
only): the per-walker kernels of ``oracle.mcmc`` / ``oracle.gfmc`` stepped walker by walker in
plain Python, with the per-step bookkeeping of the reference drivers, so that a GPU driver run on
the same Philox stream can be compared measurement by measurement.

* ``mcmc_run`` follows MCMC.run (jqmc_mcmc.py:664-745): ``num_mcmc_per_measurement`` proposals
  per walker (move counter step * n), the measurement rotation of (walker, step), the local energy
  with the streamed G^-1, and the AS weight w_L = (R_AS / max(R_AS, epsilon_AS))^2 (741-747).
* ``jackknife_energy`` restates MCMC.get_E (jqmc_mcmc.py:1064-1189) as an explicit leave-one-bin-out
  loop.
* ``gfmc_n_run`` follows GFMC_n.run (jqmc_gfmc.py:5780-6380): n projections per walker from w = 1,
  V_diag / V_nondiag at the last rotation, the weighted sums and the comb with zeta from the legacy
  NumPy generator seeded with mcmc_seed, positions copied from the parents and G^-1 rebuilt
  (6314-6319), and the on-the-fly E_scf update (6327-6380).
* ``gfmc_t_run`` follows GFMC_t.run (jqmc_gfmc.py:1700-2391): each walker projects from w = 1 until
  its time tau is used up (waiting times from the projection draws, counter (step << 32) + m), e_L
  at the final configuration, the energy sums weighted by w alone (V_diag - E_scf = 1,
  1907-1925), then the same comb, copy and E_scf bookkeeping as ``gfmc_n_run``.
"""

from __future__ import annotations

import numpy as np

from . import coulomb as oc
from . import gfmc as og
from . import mcmc as om
from . import wf as owf

# on-the-fly E_scf estimator constants (_setting.py:59-61)
WARMUP_STEPS = 20
COLLECT_STEPS = 10
BIN_BLOCKS = 10


def mcmc_run(ham, r_up, r_dn, seed, n_steps, n_per_meas, Dt, epsilon_AS, Nv=6, random_mesh=True):
    """Returns e_L (n_steps, W), w_L (n_steps, W), accepted and rejected move totals."""
    W = len(r_up)
    states = []
    for w in range(W):
        G, Gi = om.init_walker(ham, r_up[w], r_dn[w])
        states.append([np.array(r_up[w]), np.array(r_dn[w]), Gi, G])
    e_L = np.zeros((n_steps, W))
    w_L = np.zeros((n_steps, W))
    acc_tot = rej_tot = 0
    for k in range(n_steps):
        for w in range(W):
            ru, rd, Gi, G = states[w]
            acc, rej, ru, rd, Gi, G = om.metropolis_walker(ham, ru, rd, Gi, G, seed, w, k * n_per_meas, n_per_meas,
                                                           Dt, epsilon_AS)
            acc_tot += acc
            rej_tot += rej
            states[w] = [ru, rd, Gi, G]
            RT = om.rotation_RT(seed, w, k) if random_mesh else np.eye(3)
            e_L[k, w] = oc.local_energy_fast(ham, ru, rd, RT, Gi, NN=1, Nv=Nv)
            R = owf.as_regularization_fast(G, Gi)
            w_L[k, w] = (R / max(R, epsilon_AS)) ** 2
    return e_L, w_L, acc_tot, rej_tot


def jackknife_energy(e_L, w_L, num_mcmc_warmup_steps, num_mcmc_bin_blocks):
    """(E, E_std, Var, Var_std): bins of consecutive steps per walker, each leave-one-bin-out
    estimate E_j = sum_{b != j} sum w e / sum_{b != j} sum w."""
    e = np.asarray(e_L)[num_mcmc_warmup_steps:]
    w = np.asarray(w_L)[num_mcmc_warmup_steps:]
    bins_w, bins_we, bins_we2 = [], [], []
    for part_w, part_e in zip(np.array_split(w, num_mcmc_bin_blocks, axis=0),
                              np.array_split(e, num_mcmc_bin_blocks, axis=0)):
        for walker in range(w.shape[1]):
            ww, ee = part_w[:, walker], part_e[:, walker]
            bins_w.append(float(np.sum(ww)))
            bins_we.append(float(np.sum(ww * ee)))
            bins_we2.append(float(np.sum(ww * ee * ee)))
    M = len(bins_w)
    E_j, V_j = [], []
    for j in range(M):
        sw = sum(bins_w[i] for i in range(M) if i != j)
        se = sum(bins_we[i] for i in range(M) if i != j)
        se2 = sum(bins_we2[i] for i in range(M) if i != j)
        E_j.append(se / sw)
        V_j.append(se2 / sw - (se / sw) ** 2)
    E_j, V_j = np.array(E_j), np.array(V_j)
    E, V = E_j.mean(), V_j.mean()
    return E, np.sqrt((M - 1) * np.mean((E_j - E) ** 2)), V, np.sqrt((M - 1) * np.mean((V_j - V) ** 2))


def gfmc_n_run(ham, r_up, r_dn, seed, n_steps, n_proj, alat, E_scf, non_local_move="tmove", Nv=6,
               random_mesh=True):
    """Returns a dict of per-step lists: e_L, e_L2, w_L (mean weight), zeta, chosen (W,) parent
    indices, E_scf (the value used by that step's projections), and the final positions."""
    W = len(r_up)
    gem = ham.wavefunction_data.geminal_data
    rng = np.random.RandomState(seed)
    ru = [np.array(x) for x in r_up]
    rd = [np.array(x) for x in r_dn]
    G = [owf.geminal(gem, ru[w], rd[w]) for w in range(W)]
    Gi = [np.linalg.inv(g) for g in G]
    out = {"e_L": [], "e_L2": [], "w_L": [], "zeta": [], "chosen": [], "E_scf": []}
    G_L, G_e_L = [], []
    interval = int(max(n_steps / 100, 1))
    for i in range(n_steps):
        out["E_scf"].append(E_scf)
        w_fin = np.zeros(W)
        Vd = np.zeros(W)
        Vn = np.zeros(W)
        for w in range(W):
            wgt, last = 1.0, None
            for k in range(n_proj):
                last = og.projection_step(ham, ru[w], rd[w], Gi[w], G[w], wgt, E_scf, alat, seed, w, i * n_proj + k,
                                          non_local_move, Nv=Nv, random_mesh=random_mesh)
                ru[w], rd[w], Gi[w], G[w], wgt = last["r_up"], last["r_dn"], last["Ginv"], last["G"], last["w"]
            w_fin[w] = wgt
            Vd[w], Vn[w], _, _ = og.hamiltonian_elements(ham, ru[w], rd[w], Gi[w], last["RT"], alat, non_local_move,
                                                         Nv=Nv, NN=1)
        e = Vd + Vn
        zeta = float(rng.random_sample())
        s = og.weighted_sums(w_fin, Vd, e, E_scf)
        out["w_L"].append(s[0] / W)
        out["e_L"].append(s[2] / s[1])
        out["e_L2"].append(s[3] / s[1])
        out["zeta"].append(zeta)
        chosen = og.comb_branching(w_fin, zeta)
        out["chosen"].append(chosen)
        ru = [ru[c].copy() for c in chosen]
        rd = [rd[c].copy() for c in chosen]
        G = [owf.geminal(gem, ru[w], rd[w]) for w in range(W)]
        Gi = [np.linalg.inv(g) for g in G]
        # on-the-fly E_scf (jqmc_gfmc.py:6327-6380)
        if i >= COLLECT_STEPS:
            g_l = float(np.prod(out["w_L"][i - COLLECT_STEPS:i]))
            G_L.append(g_l)
            G_e_L.append(g_l * out["e_L"][i])
        if (i + 1) % interval == 0 and i > WARMUP_STEPS:
            n_warm = min(WARMUP_STEPS, i - WARMUP_STEPS)
            Gw = np.array(G_L[n_warm:])
            Gew = np.array(G_e_L[n_warm:])
            Gb = np.array([np.sum(x) for x in np.array_split(Gw, BIN_BLOCKS)])
            Geb = np.array([np.sum(x) for x in np.array_split(Gew, BIN_BLOCKS)])
            E_scf = float(np.average([(Geb.sum() - Geb[m]) / (Gb.sum() - Gb[m]) for m in range(BIN_BLOCKS)]))
    out["r_up"], out["r_dn"], out["E_scf_final"] = np.array(ru), np.array(rd), E_scf
    return out


def _e_scf_update(out, G_L, G_e_L, i, interval, E_scf):
    if i >= COLLECT_STEPS:
        g_l = float(np.prod(out["w_L"][i - COLLECT_STEPS:i]))
        G_L.append(g_l)
        G_e_L.append(g_l * out["e_L"][i])
    if (i + 1) % interval == 0 and i > WARMUP_STEPS:
        n_warm = min(WARMUP_STEPS, i - WARMUP_STEPS)
        Gw = np.array(G_L[n_warm:])
        Gew = np.array(G_e_L[n_warm:])
        Gb = np.array([np.sum(x) for x in np.array_split(Gw, BIN_BLOCKS)])
        Geb = np.array([np.sum(x) for x in np.array_split(Gew, BIN_BLOCKS)])
        E_scf = float(np.average([(Geb.sum() - Geb[m]) / (Gb.sum() - Gb[m]) for m in range(BIN_BLOCKS)]))
    return E_scf


def gfmc_t_run(ham, r_up, r_dn, seed, n_steps, tau, alat, E_scf, non_local_move="tmove", Nv=6,
               random_mesh=True, max_projections=10000):
    """Returns a dict of per-step lists: e_L, e_L2, w_L (mean weight), n_proj (projections per walker),
    zeta, chosen (W,) parent indices, and the final positions and E_scf."""
    W = len(r_up)
    gem = ham.wavefunction_data.geminal_data
    rng = np.random.RandomState(seed)
    ru = [np.array(x) for x in r_up]
    rd = [np.array(x) for x in r_dn]
    G = [owf.geminal(gem, ru[w], rd[w]) for w in range(W)]
    Gi = [np.linalg.inv(g) for g in G]
    out = {"e_L": [], "e_L2": [], "w_L": [], "n_proj": [], "zeta": [], "chosen": []}
    G_L, G_e_L = [], []
    interval = int(max(n_steps / 100, 1))
    for i in range(n_steps):
        w_fin = np.zeros(W)
        e = np.zeros(W)
        n_proj = np.zeros(W, dtype=np.int64)
        for w in range(W):
            wgt, tau_left, m = 1.0, float(tau), 0
            while True:
                last = og.projection_t_step(ham, ru[w], rd[w], Gi[w], G[w], wgt, tau_left, alat, seed, w,
                                            (i << 32) + m, non_local_move, random_mesh=random_mesh, Nv=Nv)
                ru[w], rd[w], Gi[w], G[w] = last["r_up"], last["r_dn"], last["Ginv"], last["G"]
                wgt, tau_left, m = last["w"], last["tau_left"], m + 1
                if not last["moved"] or m >= max_projections:
                    break
            w_fin[w], e[w], n_proj[w] = wgt, last["e_L"], m
        zeta = float(rng.random_sample())
        s = og.weighted_sums(w_fin, np.full(W, E_scf + 1.0), e, E_scf)
        out["w_L"].append(s[0] / W)
        out["e_L"].append(s[2] / s[1])
        out["e_L2"].append(s[3] / s[1])
        out["n_proj"].append(n_proj)
        out["zeta"].append(zeta)
        chosen = og.comb_branching(w_fin, zeta)
        out["chosen"].append(chosen)
        ru = [ru[c].copy() for c in chosen]
        rd = [rd[c].copy() for c in chosen]
        G = [owf.geminal(gem, ru[w], rd[w]) for w in range(W)]
        Gi = [np.linalg.inv(g) for g in G]
        E_scf = _e_scf_update(out, G_L, G_e_L, i, interval, E_scf)
    out["r_up"], out["r_dn"], out["E_scf_final"] = np.array(ru), np.array(rd), E_scf
    return out
