"""Throughput benchmark of the gfx950 VMC/LRDMC hot path on one MI355X.

Workload: the benzene-like ECP system (6 C + 6 H, 228 Cartesian AOs, 30 valence electrons,
JSD with 1b+2b+3b Jastrow, 12-point ECP quadrature).  One step = one Metropolis sweep
(N = 30 single-electron proposals per walker, the move-synchronous ``jqmc_vmc_sweep_batched``) +
the measurement rotation + one local energy per walker (``jqmc_local_energy``), i.e. the per-measurement structure of
``MCMC.run`` (jqmc_mcmc.py:664-745).  Inputs are resident in HBM before timing.

Prints ONE JSON line.  ``value`` = plain-FP64 walker-steps/s (single-electron proposals per
second); ``el_evals_per_s`` = local-energy evaluations per second.  ``mixed`` repeats the same
steps on the same Philox stream in the mixed-precision form (FP32 AO / orbital / J3 caches, FP64
determinant inverse refreshed every step) and reports its rate, the speedup and the energy
difference against the FP64 run.  ``lrdmc`` times LRDMC projections (GFMC_n, t-move) and
``dlnpsi`` the parameter log-derivative kernel and ``forces`` the position-derivative (force)
kernel on the same system.  The ``roofline`` object
covers the kernel with the largest share of the timed region, timed with HIP events on the
stream the kernels are launched on.  ``cpu_baseline`` times the reference algorithm restated in
C++ with OpenMP (oracle/cpu_ref.cpp, FP64) on a bounded sample of the same workload on this host,
with one thread and with every core the process may use.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3


def ecp_electrons(pack, r_up, r_dn):
    """Mean number of electrons per walker whose nearest nucleus carries non-local ECP terms
    (the electrons whose quadrature points the E_L kernel evaluates)."""
    if not pack.ecp_on:
        return 0.0
    R = pack._keep["nuc_pos"].reshape(-1, 3)
    has = pack._keep["ecp_nt"] > 0
    X = np.concatenate([r_up, r_dn], axis=1)  # (W, N, 3)
    d = np.linalg.norm(X[:, :, None, :] - R[None, None, :, :], axis=-1)
    return float(np.mean(np.sum(has[np.argmin(d, axis=2)], axis=1)))


def kernel_work(pack, W, n_moves, accepted, n_ecp_el=None):
    """Algorithmic FP64 flops and compulsory HBM bytes per launch of the sweep and the E_L kernels.

    Conventions: FMA = 2 flops, exp/log/sqrt/div = 1 op.  Bytes = the walker state each kernel must
    touch once per launch: the sweep reads and writes positions, AO and J3-field caches (N x n_ao
    each), orbital and pair caches (N x n_orb each), G and G^-1; the E_L kernel reads positions,
    the J3 field, orbital and pair caches and G^-1, and writes one energy per walker.
    """
    na, ne, no, nu, nprim = pack.n_ao, pack.n_el, pack.n_orb, pack.n_up, int(pack.desc.n_prim)
    d = 8.0
    # ---- sweep: per proposal and per accepted move
    ao_eval = nprim * 3 + na * 6                       # exp + fma per primitive, polynomial per AO
    prop_flops = ao_eval + 2 * na + 2 * no * na + 2 * no * nu + 8 * nu * nu + 60
    acc_flops = 4 * na * na + (ne - 1) * na + 2 * no * no + 2 * nu * nu  # J3 and J3^T times the AO change
    state = 3 * ne + 2 * ne * na + 2 * ne * no + 2 * nu * nu
    sweep = {"flops": W * n_moves * prop_flops + accepted * acc_flops, "bytes": W * 2 * d * state}
    # ---- local energy: kinetic per electron, bare Coulomb, ECP points
    # quadrature points per walker: Nv for every electron whose nearest nucleus has an ECP
    n_pts = (ne if n_ecp_el is None else n_ecp_el) * pack.nv if pack.ecp_on else 0
    kin = ne * (nprim * 5 + na * 40 + 2 * no * na + 16 * na) + ne * (no * ne * 2)
    ecp = n_pts * (nprim * 3 + na * 6 + 4 * na + 30 * ne)
    el_flops = kin + ecp + 10 * ne * ne
    el_bytes = d * (3 * ne + ne * na + 2 * ne * no + nu * nu + 1)
    el = {"flops": W * el_flops, "bytes": W * el_bytes, "quadrature_points_per_walker": n_pts}
    return sweep, el


# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (each scaled by what the
# counter reads for a plain 1 GiB device copy: x2.00 and x1.00) for the default configuration.
MEASURED_TRAFFIC = {
    ("benzene_ecp", 16384, "sweep"): (24.2e9, "profiles/vmc_step_benzene_ecp_w16384_fp64.md"),
    ("benzene_ecp", 16384, "k_local_energy_pts"): (4.70e9, "profiles/vmc_step_benzene_ecp_w16384_fp64.md"),
    ("benzene_ecp", 16384, "k_local_energy_team<double, 2>"): (4.46e9, "profiles/vmc_step_benzene_ecp_w16384_fp64.md"),
}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(system, n_walkers: int, n_moves: int, Nv: int, seed: int):
    """The reference algorithm restated in C++ with OpenMP over walkers (oracle/cpu_ref.cpp, checked
    against the NumPy oracle by tests/test_cpu_ref.py): one VMC step (a sweep of n_moves proposals
    and one local energy per walker) on ``n_walkers`` walkers with one thread, and on 4x as many with
    every core this process may use; walker set-up is outside the timed region."""
    from jqmc_amd._pack import SystemPack
    from oracle import cpu_ref

    pack = SystemPack(system.hamiltonian_data, Nv=Nv, device=False)
    cores = len(os.sched_getaffinity(0))
    out = {}
    for label, threads, W in (("1_thread", 1, n_walkers), ("all_cores", cores, 4 * n_walkers)):
        r_up, r_dn = system.initial_walkers(W, seed=seed)
        _, _, _, _, sec = cpu_ref.vmc_step(pack.desc, r_up, r_dn, n_moves, 0.5, 1e-4, seed, n_threads=threads)
        out[label] = {"threads": threads, "walkers": W, "seconds": sec, "walker_steps_per_s": W * n_moves / sec,
                      "el_evals_per_s": W / sec}
    return {
        "value": out["all_cores"]["walker_steps_per_s"],
        "unit": "walker-steps/s",
        "cores": cores,
        "kind": "port",
        "label": "reference algorithm, CPU restatement (C++/OpenMP, FP64)",
        "cpu_model": cpu_model(),
        "sample": f"{n_walkers} walkers (1 thread) and {4 * n_walkers} walkers ({cores} threads) x (1 sweep of "
                  f"{n_moves} proposals + 1 E_L), same system and Philox stream as the GPU",
        **out,
    }


def same_config_errors(wm, wd, RT, wm_fresh=None):
    """E_L of a mixed walker set against a plain FP64 set at the same positions, and ln|Psi| of a
    mixed set-up at those positions (``wm_fresh``, default ``wm``: the walker caches keep ln|Psi|
    from set-up only) against the FP64 one."""
    import torch

    de = wm.local_energy(RT) - wd.local_energy(RT)
    wf = wm if wm_fresh is None else wm_fresh
    dl = ((wf.logabsdet + wf.jastrow) - (wd.logabsdet + wd.jastrow)).abs()
    W = de.numel()
    return {"e_L_bias_mHa": 1e3 * float(de.mean()), "e_L_bias_stderr_mHa": 1e3 * float(de.std()) / W**0.5,
            "e_L_abs_median_Ha": float(de.abs().median()), "e_L_abs_p99_Ha": float(torch.quantile(de.abs(), 0.99)),
            "e_L_abs_max_Ha": float(de.abs().max()), "ln_psi_abs_max": float(dl.max()), "walkers": W}


def _timed(fn, reps=1):
    import torch

    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        out = fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps, out


def lrdmc_multi_bench(args):
    """Four chain/ring molecules with 10, 30, 56 and 96 valence electrons stepped together by
    jqmc_lrdmc_projection_multi (one projection per step), comb branching every 20 steps; the same
    steps again with one launch per system."""
    import torch

    from jqmc_amd import synthetic
    from jqmc_amd._api import WalkerState, lrdmc_branching, lrdmc_projection_multi
    from jqmc_amd._pack import system_for

    Wm, n_steps, every = args.multi_walkers, 60, 20
    sizes = (10, 30, 56, 96)
    res = {"walkers_per_system": Wm, "steps": n_steps, "branching_every": every, "alat": 0.2,
           "precision": "mixed", "systems": {}}
    for mode in ("one_launch", "launch_per_system"):
        states, params, w_L = [], [], []
        for n in sizes:
            sysd = synthetic.chain_ecp(n, seed=1)
            pk = system_for(sysd.hamiltonian_data, Nv=6, precision="mixed")
            ru, rd = sysd.initial_walkers(Wm, seed=args.seed)
            ws = WalkerState(pk, Wm).init(ru, rd)
            E0 = float(ws.local_energy(None).mean().item())
            states.append(ws)
            params.append(WalkerState.lrdmc_params(0.2, E_scf=E0, n_steps=1, non_local_move="tmove"))
            w_L.append(torch.ones(Wm, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        lat = []
        rng = np.random.RandomState(args.seed)
        t0 = time.perf_counter()
        for k in range(n_steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            if mode == "one_launch":
                outs = lrdmc_projection_multi(states, params, args.seed, [k] * len(sizes), w_Ls=w_L)
            else:
                outs = [ws.lrdmc_projection(p, args.seed, k, w_L=w) for ws, p, w in zip(states, params, w_L)]
            for i, o in enumerate(outs):
                w_L[i] = o["w_L"]
            if (k + 1) % every == 0:
                for i, (ws, o) in enumerate(zip(states, outs)):
                    chosen, _ = lrdmc_branching(w_L[i], float(rng.random_sample()))
                    ws.reconfigure(chosen)
                    w_L[i] = torch.ones_like(w_L[i])
            b.record()
            torch.cuda.synchronize()
            lat.append(a.elapsed_time(b))
        total = time.perf_counter() - t0
        lat = np.asarray(lat)
        res[mode] = {"total_s": total, "p50_step_ms": float(np.percentile(lat, 50)),
                     "p99_step_ms": float(np.percentile(lat, 99)),
                     "walker_steps_per_s_per_system": Wm * n_steps / total}
        for n, ws in zip(sizes, states):
            res["systems"][str(n)] = {"n_ao": ws.pack.n_ao, "n_el": ws.pack.n_el}
        del states
    return res


def extras(system, ham, args, W, n_moves, e_steps_fp64):
    """Mixed-precision rerun of the timed steps, LRDMC projections and the O_k kernel."""
    import torch

    from jqmc_amd._api import WalkerState
    from jqmc_amd._pack import system_for

    out = {}
    r_up, r_dn = system.initial_walkers(W, seed=args.seed)
    # ---- mixed precision, same Philox stream and steps as the FP64 run
    pack_m = system_for(ham, Nv=args.nv, precision="mixed")
    wm = WalkerState(pack_m, W).init(r_up, r_dn)
    pack_d = system_for(ham, Nv=args.nv)
    wd = WalkerState(pack_d, W).init(r_up, r_dn)
    # errors of the mixed form against the plain FP64 form on the unequilibrated start positions
    RT0 = wd.rotations(args.seed, 10**6)
    out_start = same_config_errors(wm, wd, RT0)
    torch.cuda.synchronize()
    move = 0
    for k in range(args.warmup):
        wm.sweep(n_moves, 0.5, 1e-4, args.seed, move)
        move += n_moves
        wm.local_energy(wm.rotations(args.seed, k))
    torch.cuda.synchronize()
    e_m = []

    def mixed_steps():
        nonlocal move
        for k in range(args.steps):
            # the caches stream through every step, as in the reference's run (no rebuild); the
            # local energy recomputes the far-off walkers in FP64 (guard)
            wm.sweep(n_moves, 0.5, 1e-4, args.seed, move)
            move += n_moves
            e_m.append(wm.local_energy(wm.rotations(args.seed, args.warmup + k)))

    ms, _ = _timed(mixed_steps)
    ed = torch.stack(e_steps_fp64)  # (steps, W)
    em = torch.stack(e_m)
    dtraj = (em - ed).mean(dim=0)  # per-walker time average of the difference of the two runs
    out["mixed"] = {
        "value": W * n_moves * args.steps / (ms / 1e3),
        "unit": "walker-steps/s",
        "ms_per_step": ms / args.steps,
        "el_evals_per_s": W * args.steps / (ms / 1e3),
        "speedup_vs_fp64": None,
        "dE_vmc_Ha": float((em - ed).mean()),
        "dE_vmc_stderr_Ha": float(dtraj.std()) / W**0.5,
        "E_vmc_fp64": float(ed.mean()),
        "E_vmc_mixed": float(em.mean()),
        "precision": "reference zones (jqmc/_precision.py:383-411): FP32 AO values and Jastrow (J3 fields, "
                     "J3 products on FP32 matrix cores); FP64 AO->MO and Lambda products, orbital/pair caches, G, "
                     "G^-1, determinant ratios and fields, AO grad/lap at electrons, energies; caches streamed through "
                     "the run (no rebuild); walkers with |e_L| > 1e4 Ha recomputed in FP64 (guard)",
        "start_config": out_start,
    }
    # Bias of the mixed form on identical configurations: the mixed walkers as they stand after
    # the timed steps (J3 fields streamed in FP32 through the sweeps, determinant side refreshed)
    # against the FP64 form built afresh at the same positions, same rotations.  dE_vmc_Ha above
    # compares two trajectories that part ways after the first differing decision, so it carries
    # the statistical error of both runs; this difference does not.
    wd_eq = WalkerState(pack_d, W).init(wm.r_up, wm.r_dn)
    wm_eq = WalkerState(pack_m, W).init(wm.r_up, wm.r_dn)
    RT_eq = wd_eq.rotations(args.seed, 10**6 + 1)
    out["mixed"]["same_config"] = same_config_errors(wm, wd_eq, RT_eq, wm_fresh=wm_eq)
    del wm, wd_eq, wm_eq
    # ---- LRDMC projections (GFMC_n, t-move, a = 0.3)
    WL = min(args.lrdmc_walkers, W)
    lr = {}
    for prec in ("fp64", "mixed"):
        pk = system_for(ham, Nv=args.nv, precision=prec)
        wl = WalkerState(pk, WL).init(r_up[:WL], r_dn[:WL])
        p = WalkerState.lrdmc_params(0.3, E_scf=-40.0, n_steps=args.lrdmc_steps)
        wl.lrdmc_projection(p, args.seed, 0)
        torch.cuda.synchronize()
        # median of three launches (distinct projection counters): one host stall between the
        # start event and the launch must not set the reported rate
        reps = sorted(_timed(lambda i=i: wl.lrdmc_projection(p, args.seed, 10**4 + 100 * i))[0] for i in range(3))
        ms = reps[1]
        lr[prec] = {"walker_steps_per_s": WL * args.lrdmc_steps / (ms / 1e3), "ms_per_projection": ms / args.lrdmc_steps,
                    "launch_ms": reps}
        del wl
    out["lrdmc"] = {"walkers": WL, "projections_per_launch": args.lrdmc_steps, "alat": 0.3,
                    "non_local_move": "tmove", **lr,
                    "mixed_speedup": lr["mixed"]["walker_steps_per_s"] / lr["fp64"]["walker_steps_per_s"]}
    # ---- parameter log-derivatives O_k (J1, J2, J3, Lambda)
    wd.dln_psi_dparams()
    torch.cuda.synchronize()
    ms, O = _timed(lambda: wd.dln_psi_dparams(), reps=3)
    n_par = sum(int(v[0].numel()) for v in O.values())
    out["dlnpsi"] = {"walkers": W, "params_per_walker": n_par, "ms": ms,
                     "evals_per_s": W / (ms / 1e3),
                     "GB_written_per_s": W * n_par * 8 / (ms / 1e3) / 1e9}
    # ---- SR step on those O_k: S samples x P parameters, push-through branch (P >= S)
    from jqmc_amd import sr as jsr

    S_sr = min(args.sr_samples, W)
    Osr = torch.cat([v[:S_sr].reshape(S_sr, -1) for v in O.values()], dim=1).contiguous()
    del O
    w_sr = torch.ones(S_sr, dtype=torch.float64, device=Osr.device)
    e_sr = wd.local_energy(wd.rotations(args.seed, 7))[:S_sr].contiguous()
    th = jsr.natural_gradient(Osr, w_sr, e_sr)
    torch.cuda.synchronize()
    ms_ng, th = _timed(lambda: jsr.natural_gradient(Osr, w_sr, e_sr), reps=2)
    mean, rs, cs = th["O_bar"], torch.full((S_sr,), S_sr ** -0.5, dtype=torch.float64, device=Osr.device), \
        th["diag_S"].rsqrt().contiguous()
    ms_g, G64 = _timed(lambda: jsr._gram(Osr, S_sr, n_par, None, mean, rs, cs), reps=3)
    gram_flops = float(S_sr) * (S_sr + 1) * n_par  # S(S+1)/2 dot products of length P, 2 flops per term
    full_flops = 2.0 * S_sr * S_sr * n_par          # the full S x S product a GEMM computes
    ms_gb, _ = _timed(lambda: jsr._gram_blas(Osr, mean, rs, cs, "push_through"), reps=3)
    ms_gb32, _ = _timed(lambda: jsr._gram_blas(Osr, mean, rs, cs, "push_through", "float32"), reps=3)
    ms_ngf, _ = _timed(lambda: jsr.natural_gradient(Osr, w_sr, e_sr, gram_kernel="fused"), reps=2)
    ms_g32, G32 = _timed(lambda: jsr._gram(Osr, S_sr, n_par, None, mean, rs, cs, "float32"), reps=3)
    ms_ng32, th32 = _timed(lambda: jsr.natural_gradient(Osr, w_sr, e_sr, gram="float32"), reps=2)
    out["sr"] = {"samples": S_sr, "params": n_par, "branch": th["branch"], "natural_gradient_ms": ms_ng,
                 "gram": "library GEMM (ROCm BLAS) on the written-out X, FP64",
                 "gram_blas_ms": ms_gb, "gram_blas_TFLOP_per_s": full_flops / (ms_gb / 1e3) / 1e12,
                 "gram_blas_frac_of_fp64_matrix_peak": full_flops / (ms_gb / 1e3) / 1e12 / FP64_PEAK_TFLOPS,
                 "gram_blas_fp32_ms": ms_gb32,
                 "natural_gradient_fused_gram_ms": ms_ngf,
                 "gram_ms": ms_g, "gram_TFLOP_per_s": gram_flops / (ms_g / 1e3) / 1e12,
                 "gram_frac_of_fp64_matrix_peak": gram_flops / (ms_g / 1e3) / 1e12 / FP64_PEAK_TFLOPS,
                 "frozen_params": int(th["frozen"].sum()),
                 "kernel": "k_sr_gram (fused normalisation, v_mfma_f64_16x16x4f64, 64x64 upper-triangle tiles)",
                 "fp32_gram": {
                     "kernel": "k_sr_gram_f32 (v_mfma_f32_16x16x4f32, 128x128 upper-triangle tiles, FP64 shift "
                               "and solve)",
                     "gram_ms": ms_g32, "gram_TFLOP_per_s": gram_flops / (ms_g32 / 1e3) / 1e12,
                     "gram_frac_of_fp32_matrix_peak": gram_flops / (ms_g32 / 1e3) / 1e12 / FP32_MFMA_PEAK_TFLOPS,
                     "natural_gradient_ms": ms_ng32,
                     "solver": "FP32 Gram as the Cholesky preconditioner of an FP64 conjugate-gradient solve",
                     "pcg_steps": th32["pcg_steps"], "preconditioner_shift": th32["preconditioner_shift"],
                     "gram_max_abs_diff_rel": float((G32 - G64).abs().max() / G64.abs().max()),
                     "theta_max_abs_diff_rel": float((th32["theta"] - th["theta"]).abs().max()
                                                     / th["theta"].abs().max())}}
    del Osr, th, th32, G32, G64
    # ---- the same step captured in a HIP graph (device-side Philox counters) vs eager launches
    from jqmc_amd._api import CapturedStep

    gr = {}
    for Wg in (256, W):
        ws_e = WalkerState(pack_d, Wg).init(r_up[:Wg], r_dn[:Wg])
        ws_g = WalkerState(pack_d, Wg).init(r_up[:Wg], r_dn[:Wg])
        n_rep = 20

        def eager_steps():
            for k in range(n_rep):
                ws_e.sweep(n_moves, 0.5, 1e-4, args.seed, 10**7 + k * n_moves)
                ws_e.local_energy(ws_e.rotations(args.seed, 10**5 + k))

        eager_steps()
        ms_e, _ = _timed(eager_steps)
        cap = CapturedStep(ws_g, n_moves, 0.5, 1e-4, args.seed, move0=10**7, step0=10**5)
        cap.replay()
        torch.cuda.synchronize()

        def graph_steps():
            for _ in range(n_rep):
                cap.replay()

        ms_g, _ = _timed(graph_steps)
        gr[str(Wg)] = {"eager_ms_per_step": ms_e / n_rep, "graph_ms_per_step": ms_g / n_rep,
                       "graph_speedup": ms_e / ms_g}
        del ws_e, ws_g, cap
    out["graph"] = gr
    # ---- batched LRDMC of uneven systems in one launch (N = 10/30/56/96, a = 0.2, mixed precision,
    #      comb reconfiguration every 20 steps): per-system walker-steps/s and step latency
    if args.multi_walkers > 0:
        out["lrdmc_multi"] = lrdmc_multi_bench(args)
    # ---- forces: dE_L/dr, dE_L/dR, d ln|Psi|/d(r, R) and the space-warp products per walker
    WF = min(args.force_walkers, W)
    wf = WalkerState(pack_d, WF).init(r_up[:WF], r_dn[:WF])
    RTf = wf.rotations(args.seed, 7)
    wf.position_grads(RTf)
    torch.cuda.synchronize()
    ms, _ = _timed(lambda: wf.position_grads(RTf), reps=3)
    out["forces"] = {"walkers": WF, "ms": ms, "evals_per_s": WF / (ms / 1e3),
                     "outputs": "e_L, dE_L/dr, dE_L/dR, dln|Psi|/dr, dln|Psi|/dR, F_HF, F_PP (SWCT)",
                     "n_nuclei": int(pack_d.desc.n_nuc)}
    return out


def config_lines(args):
    """The two other single-GPU configurations at their stated sizes: the 4 x H2O cluster (ECP on O,
    32 valence electrons) in plain FP64 and mixed precision at 8192 walkers, and its JAGP form at
    4096 walkers (parameter log-derivatives O_k, the SR step on them, forces on 12 nuclei)."""
    import torch

    from jqmc_amd import synthetic
    from jqmc_amd import sr as jsr
    from jqmc_amd._api import WalkerState
    from jqmc_amd._pack import system_for

    out = {}
    # ---- 4 x H2O, FP64 and mixed, same start positions and Philox stream
    sysw = synthetic.water4_ecp(0)
    ham = sysw.hamiltonian_data
    W = args.water_walkers
    r_up, r_dn = sysw.initial_walkers(W, seed=args.seed)
    wss = {prec: WalkerState(system_for(ham, Nv=args.nv, precision=prec), W).init(r_up, r_dn)
           for prec in ("fp64", "mixed")}
    RT0 = wss["fp64"].rotations(args.seed, 10**6)
    line = {"workload": "water4_ecp: 4 x H2O on a 5.5 bohr square, ECP on O, 32 valence e-, "
                        f"{wss['fp64'].pack.n_ao} AOs, JSD + J1/J2/J3",
            "walkers": W, "start_config": same_config_errors(wss["mixed"], wss["fp64"], RT0)}
    es = {}
    for prec, ws in wss.items():
        n_moves = ws.pack.n_el
        move = 0
        for k in range(2):
            ws.sweep(n_moves, 0.5, 1e-4, args.seed, move)
            move += n_moves
            ws.local_energy(ws.rotations(args.seed, k))
        torch.cuda.synchronize()
        e_list = []

        def steps():
            nonlocal move
            for k in range(args.config_steps):
                ws.sweep(n_moves, 0.5, 1e-4, args.seed, move)
                move += n_moves
                e_list.append(ws.local_energy(ws.rotations(args.seed, 2 + k)))

        ms, _ = _timed(steps)
        es[prec] = torch.stack(e_list)
        line[prec] = {"walker_steps_per_s": W * n_moves * args.config_steps / (ms / 1e3),
                      "ms_per_step": ms / args.config_steps}
    d = (es["mixed"] - es["fp64"]).mean(dim=0)
    line["mixed_speedup"] = line["mixed"]["walker_steps_per_s"] / line["fp64"]["walker_steps_per_s"]
    line["dE_vmc_Ha"] = float(d.mean())
    line["dE_vmc_stderr_Ha"] = float(d.std()) / W**0.5
    line["steps"] = args.config_steps
    out["water4_mixed"] = line
    del wss, es
    # ---- 4 x H2O JAGP: O_k, the SR step on them (FP64 and FP32 Gram), forces
    sysj = synthetic.water4_ecp(0, jagp=True)
    WJ = args.jagp_walkers
    r_up, r_dn = sysj.initial_walkers(WJ, seed=args.seed)
    ws = WalkerState(system_for(sysj.hamiltonian_data, Nv=args.nv), WJ).init(r_up, r_dn)
    n_moves = ws.pack.n_el
    for k in range(2):
        ws.sweep(n_moves, 0.5, 1e-4, args.seed, k * n_moves)
    torch.cuda.synchronize()
    ws.dln_psi_dparams()
    torch.cuda.synchronize()
    ms_o, O = _timed(lambda: ws.dln_psi_dparams(), reps=2)
    n_par = sum(int(v[0].numel()) for v in O.values())
    Osr = torch.cat([v.reshape(WJ, -1) for v in O.values()], dim=1).contiguous()
    del O
    w_sr = torch.ones(WJ, dtype=torch.float64, device=Osr.device)
    RT = ws.rotations(args.seed, 7)
    e_sr = ws.local_energy(RT).contiguous()
    sr = {}
    for gram in ("float64", "float32"):
        jsr.natural_gradient(Osr, w_sr, e_sr, gram=gram)
        torch.cuda.synchronize()
        ms, th = _timed(lambda: jsr.natural_gradient(Osr, w_sr, e_sr, gram=gram), reps=2)
        sr[gram] = {"natural_gradient_ms": ms, "branch": th["branch"], "solver": th["solver"], "theta": th["theta"]}
        if "pcg_steps" in th:
            sr[gram]["pcg_steps"] = th["pcg_steps"]
    t64, t32 = sr["float64"].pop("theta"), sr["float32"].pop("theta")
    sr["theta_fp32_gram_max_abs_diff_rel"] = float((t32 - t64).abs().max() / t64.abs().max())
    del Osr
    ws.position_grads(RT)
    torch.cuda.synchronize()
    ms_f, _ = _timed(lambda: ws.position_grads(RT), reps=2)
    out["water4_jagp"] = {
        "workload": "water4_ecp(jagp=True): symmetric geminal Lambda, 1b+2b+3b Jastrow",
        "walkers": WJ, "params_per_walker": n_par,
        "dlnpsi": {"ms": ms_o, "evals_per_s": WJ / (ms_o / 1e3)},
        "sr": {"samples": WJ, **sr},
        "forces": {"ms": ms_f, "evals_per_s": WJ / (ms_f / 1e3), "n_nuclei": int(ws.pack.desc.n_nuc)},
    }
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--walkers", type=int, default=16384)
    ap.add_argument("--workload", default="benzene_ecp")
    ap.add_argument("--nv", type=int, default=12)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-walkers", type=int, default=2048,
                    help="walkers of the 1-thread CPU-baseline sample (4x as many on all cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the mixed / LRDMC / O_k measurements")
    ap.add_argument("--lrdmc-walkers", type=int, default=4096)
    ap.add_argument("--sr-samples", type=int, default=4096, help="samples (rows of O) of the SR extra")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default launch shapes")
    ap.add_argument("--lrdmc-steps", type=int, default=10)
    ap.add_argument("--force-walkers", type=int, default=4096)
    ap.add_argument("--multi-walkers", type=int, default=2048, help="walkers per system in the uneven LRDMC batch")
    ap.add_argument("--water-walkers", type=int, default=8192, help="walkers of the 4 x H2O FP64 / mixed line")
    ap.add_argument("--jagp-walkers", type=int, default=4096, help="walkers of the 4 x H2O JAGP derivative line")
    ap.add_argument("--config-steps", type=int, default=5, help="timed steps of the 4 x H2O line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale factor for the HBM byte counters)")
    args = ap.parse_args(argv)

    if args.copy_calibration:
        import torch

        n = (1 << 30) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                          "dispatches": 3}))
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(walkers shard across GPUs with no per-step exchange in VMC; see DESIGN.md).", file=sys.stderr)
        return 2

    import torch

    from jqmc_amd import synthetic
    from jqmc_amd._api import WalkerState
    from jqmc_amd._pack import system_for

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    system = synthetic.WORKLOADS[args.workload](0)
    ham = system.hamiltonian_data
    pack = system_for(ham, Nv=args.nv)
    W = args.walkers
    n_moves = pack.n_el
    r_up, r_dn = system.initial_walkers(W, seed=args.seed)
    ws = WalkerState(pack, W).init(r_up, r_dn)
    tuning = None
    if not args.no_autotune:
        # launch shapes for this system and walker count (set-up, outside the timed region; the
        # walkers-per-block knobs change no result bit, a team of waves per walker changes e_L in
        # the last bits)
        from jqmc_amd.autotune import autotune

        tuning = autotune(ws, n_moves=n_moves)
    torch.cuda.synchronize()

    move = 0

    def step(k, ev=None):
        nonlocal move
        if ev is not None:
            ev[0].record()
        ws.sweep(n_moves, 0.5, 1e-4, args.seed, move)
        if ev is not None:
            ev[1].record()
        move += n_moves
        RT = ws.rotations(args.seed, k)
        if ev is not None:
            ev[2].record()
        e = ws.local_energy(RT)
        if ev is not None:
            ev[3].record()
        return e

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    acc0 = int(ws.accepted.sum().item())

    # timed region: the steps alone, events only at its ends
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record()
    e_sum = torch.zeros((), dtype=torch.float64, device="cuda")
    e_steps = []
    for k in range(args.steps):
        e = step(args.warmup + k)
        e_sum += e.sum()
        e_steps.append(e)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    total_ms = start.elapsed_time(end)
    accepted = (int(ws.accepted.sum().item()) - acc0) / args.steps
    # stage breakdown in a separate pass after the timed region (events between the stages, on the
    # stream the kernels run on): sweep, rotations, local energy
    n_break = 3
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(n_break)]
    for k in range(n_break):
        step(args.warmup + args.steps + k, evs[k])
    torch.cuda.synchronize()
    sweep_ms = sum(ev[0].elapsed_time(ev[1]) for ev in evs) / n_break
    rot_ms = sum(ev[1].elapsed_time(ev[2]) for ev in evs) / n_break
    el_ms = sum(ev[2].elapsed_time(ev[3]) for ev in evs) / n_break
    step_ev_ms = sum(ev[0].elapsed_time(ev[3]) for ev in evs) / n_break
    ms_per_step = total_ms / args.steps
    walker_steps = W * n_moves * args.steps
    value = walker_steps / (total_ms / 1e3)
    el_rate = W * args.steps / (total_ms / 1e3)
    e_mean = float(e_sum.item()) / (W * args.steps)
    if not np.isfinite(e_mean):
        print("non-finite local energy in the timed region", file=sys.stderr)
        return 1

    n_ecp_el = ecp_electrons(pack, ws.r_up.cpu().numpy(), ws.r_dn.cpu().numpy())
    sweep_w, el_w = kernel_work(pack, W, n_moves, accepted, n_ecp_el)
    # the roofline kernel: the single kernel with the largest time per step, the local energy
    # (k_local_energy_pts, or k_local_energy_team when the autotuner picked a team of waves per
    # walker; one launch per step; the sweep is 62 launches of four kernels)
    team = tuning["best"].get("el_waves_per_walker", 1) if tuning and "best" in tuning else 1
    name = f"k_local_energy_team<double, {team}>" if team > 1 else "k_local_energy_pts"
    ms, work = el_ms, el_w
    intensity = work["flops"] / work["bytes"]
    traffic = MEASURED_TRAFFIC.get((args.workload, W, name))
    achieved = work["flops"] / (ms / 1e3) / 1e12
    roof = {"kernel": name, "bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "note": "FP64 vector ALU kernel (no matrix instruction) far above the machine balance "
                    f"({intensity:.0f} flop/B of compulsory traffic); peak = MI355X FP64 vector 78.6 TFLOP/s; "
                    "exp/sqrt/div count one op each"}
    roof["traffic"] = traffic[0] if traffic else None
    if traffic:
        roof["traffic_source"] = traffic[1]
    roof["avg_ms"] = ms
    roof["algorithmic"] = {"flops": work["flops"], "bytes": work["bytes"], "flop_per_byte": intensity,
                           "flops_per_walker": work["flops"] / W,
                           "quadrature_points_per_walker": work["quadrature_points_per_walker"]}
    roof["other_kernel"] = {"kernel": "sweep (k_sweep_begin, n_moves x [k_sweep_propose + k_sweep_jfield_mfma], "
                                      "k_sweep_flush)",
                            "avg_ms": sweep_ms, "TFLOP_per_s": sweep_w["flops"] / (sweep_ms / 1e3) / 1e12,
                            "GB_per_s": sweep_w["bytes"] / (sweep_ms / 1e3) / 1e9,
                            "note": "J3 field products on the FP64 matrix cores, proposals on the vector ALUs"}
    sweep_traffic = MEASURED_TRAFFIC.get((args.workload, W, "sweep"))
    if sweep_traffic:
        roof["other_kernel"]["traffic"] = sweep_traffic[0]
        roof["other_kernel"]["traffic_source"] = sweep_traffic[1]

    extra = {}
    if not args.no_extras:
        extra = extras(system, ham, args, W, n_moves, e_steps)
        extra["configs"] = config_lines(args)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(system, args.cpu_walkers, n_moves, args.nv, args.seed)

    line = {
        "metric": "walker-steps/s (VMC single-electron proposals) + E_L evals/s, 1 GPU, plain FP64",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: 6 C (ECP) + 6 H ring, {pack.n_ao} Cartesian AOs, {pack.n_el} valence e-, "
                        f"JSD + J1/J2/J3, ECP Nv={args.nv}, NN=1",
            "walkers": W,
            "moves_per_step": n_moves,
            "el_evals_per_step": W,
            "precision": "plain FP64",
        },
        "el_evals_per_s": el_rate,
        "sweep_ms": sweep_ms,
        "local_energy_ms": el_ms,
        "stages_ms": {"sweep": sweep_ms, "rotations": rot_ms, "local_energy": el_ms,
                      "sum": sweep_ms + rot_ms + el_ms, "step_between_events": step_ev_ms,
                      "timed_step": ms_per_step, "unaccounted": ms_per_step - (sweep_ms + rot_ms + el_ms),
                      "note": "stages timed with events in a separate pass of 3 steps after the timed region"},
        "acceptance": accepted / (W * n_moves),
        "mean_e_L": e_mean,
        "wall_s": wall,
        "roofline": roof,
        "cpu_baseline": cpu,
        "autotune": tuning,
    }
    line.update(extra)
    if "mixed" in line:
        line["mixed"]["speedup_vs_fp64"] = line["mixed"]["value"] / value
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
