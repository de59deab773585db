"""VMC driver with the reference's ``MCMC`` interface (jqmc_mcmc.py:120-1370).

One measurement step (``MCMC.run``, jqmc_mcmc.py:448-1062) is ``num_mcmc_per_measurement``
single-electron Metropolis proposals per walker (``jqmc_vmc_sweep``), the quadrature rotations
(``_generate_rotation_matrix``, 4228-4245), the local energy of every walker
(``jqmc_local_energy``) and the AS weight w_L = (R_AS / max(R_AS, epsilon_AS))^2 (741-747).
All of it runs on the GPU; only the per-step (W,) energies and weights come back to the host.

Draws come from the Philox4x32-10 stream (seed, walker, move / step) instead of JAX's
threefry keys, so trajectories are statistically, not bitwise, those of the reference.
In the mixed-precision form the caches stream through the whole run, as the reference's never
rebuilds its inverse inside ``run`` (jqmc_mcmc.py:658); the local energy recomputes in FP64 the
walkers whose mixed e_L is far off (jqmc_system_set_mixed_guard).  With ``refresh_interval`` > 0
the determinant side of every walker (AO values at the electrons, orbitals, G, G^-1) is also
recomputed in FP64 from the positions every ``refresh_interval`` proposals and before the next
local energy.
"""

from __future__ import annotations

import numpy as np
import torch

from ._api import WalkerState, device
from ._pack import system_for
from .checkpoint import STOP_FILE_MCMC, RunControl, load_checkpoint, save_checkpoint
from .hamiltonians import Hamiltonian_data


def _initial_positions(ham: Hamiltonian_data, num_walkers: int, rng: np.random.Generator):
    """Electrons around the nuclei that own them (valence charges, alternating spins),
    Gaussian spread of 1 bohr per axis; shapes (W, N_up, 3), (W, N_dn, 3)."""
    st = ham.structure_data
    cp = ham.coulomb_potential_data
    gem = ham.wavefunction_data.geminal_data
    charges = np.asarray(st.atomic_numbers, dtype=float)
    if cp.ecp_flag:
        charges = charges - np.asarray(cp.z_cores, dtype=float)
    owners = np.repeat(np.arange(st.natom), np.maximum(np.round(charges).astype(int), 0))
    owners = np.resize(owners, gem.num_electron_up + gem.num_electron_dn)
    up, dn = owners[0::2][: gem.num_electron_up], owners[1::2][: gem.num_electron_dn]
    up = np.resize(up, gem.num_electron_up)
    dn = np.resize(dn, gem.num_electron_dn)
    pos = st.positions
    r_up = pos[up][None] + rng.standard_normal((num_walkers, gem.num_electron_up, 3))
    r_dn = pos[dn][None] + rng.standard_normal((num_walkers, gem.num_electron_dn, 3))
    return r_up, r_dn


class MCMC:
    """Variational Monte Carlo on the GPU with the reference's constructor and ``run``."""

    def __init__(
        self,
        hamiltonian_data: Hamiltonian_data = None,
        mcmc_seed: int = 34467,
        num_walkers: int = 40,
        num_mcmc_per_measurement: int = 16,
        Dt: float = 2.0,
        epsilon_AS: float = 1e-1,
        comput_position_deriv: bool = False,
        comput_param_deriv: bool = False,
        random_discretized_mesh: bool = True,
        use_swct: bool = True,
        Nv: int = 6,
        precision: str = "fp64",
        refresh_interval: int = 0,
        init_r_up_carts=None,
        init_r_dn_carts=None,
    ) -> None:
        hamiltonian_data.sanity_check()
        self.hamiltonian_data = hamiltonian_data
        self.mcmc_seed = int(mcmc_seed)
        self.num_walkers = int(num_walkers)
        self.num_mcmc_per_measurement = int(num_mcmc_per_measurement)
        self.Dt = float(Dt)
        self.epsilon_AS = float(epsilon_AS)
        self.random_discretized_mesh = bool(random_discretized_mesh)
        self.comput_position_deriv = bool(comput_position_deriv)
        self.comput_param_deriv = bool(comput_param_deriv)
        self.use_swct = bool(use_swct)
        if self.comput_position_deriv and precision != "fp64":
            raise NotImplementedError("position derivatives are computed with the plain FP64 form only")
        self.refresh_interval = int(refresh_interval)
        self.pack = system_for(hamiltonian_data, Nv=Nv, NN=1, precision=precision)
        if init_r_up_carts is None:
            init_r_up_carts, init_r_dn_carts = _initial_positions(
                hamiltonian_data, self.num_walkers, np.random.default_rng(self.mcmc_seed))
        self.state = WalkerState(self.pack, self.num_walkers).init(init_r_up_carts, init_r_dn_carts)
        self.mcmc_counter = 0
        self._move = 0
        self._since_refresh = 0
        self._e_L: list[np.ndarray] = []
        self._w_L: list[np.ndarray] = []
        self._force_HF: list[np.ndarray] = []
        self._force_PP: list[np.ndarray] = []
        self._E_L_force_PP: list[np.ndarray] = []
        self._dln_WF: list[torch.Tensor] = []  # (W, P) per measurement, in pinned host memory

    def run(self, num_mcmc_steps: int = 0, max_time: float = 86400) -> None:
        """``num_mcmc_steps`` measurement steps (jqmc_mcmc.py:664-745); stops early once ``max_time``
        seconds have passed or ``stop = true`` is set in ``external_control_mcmc.toml``
        (jqmc_mcmc.py:942-967; jqmc_amd.checkpoint.RunControl)."""
        ctl = RunControl(STOP_FILE_MCMC, max_time)
        try:
            self._run(int(num_mcmc_steps), ctl)
        finally:
            ctl.close()

    def save_checkpoint(self, path: str) -> None:
        """Configuration, counters, stream position, walkers and observables to ``path`` (.npz)."""
        save_checkpoint(self, path)

    @classmethod
    def load_checkpoint(cls, path: str, hamiltonian_data: Hamiltonian_data) -> "MCMC":
        """The driver saved at ``path``; ``run`` continues its chain."""
        return load_checkpoint(path, hamiltonian_data)

    def _run(self, num_mcmc_steps: int, ctl) -> None:
        ws = self.state
        mixed = self.pack.precision == "mixed"
        refresh = mixed and self.refresh_interval > 0
        for _ in range(num_mcmc_steps):
            n = self.num_mcmc_per_measurement
            while n > 0:
                chunk = min(n, self.refresh_interval - self._since_refresh) if refresh else n
                ws.sweep(chunk, self.Dt, self.epsilon_AS, self.mcmc_seed, self._move)
                self._move += chunk
                n -= chunk
                if refresh:
                    self._since_refresh += chunk
                    if self._since_refresh >= self.refresh_interval:
                        ws.refresh_orbitals()
                        self._since_refresh = 0
            if refresh and self._since_refresh > 0:
                # the local energy reads the determinant side at FP64 accuracy
                ws.refresh_orbitals()
                self._since_refresh = 0
            if self.random_discretized_mesh:
                RT = ws.rotations(self.mcmc_seed, self.mcmc_counter)
            else:
                RT = torch.eye(3, dtype=torch.float64, device=device()).expand(self.num_walkers, 3, 3)
            if self.comput_position_deriv:
                # per-walker force products (jqmc_mcmc.py:749-852) from the adjoint kernel; its e_L is
                # the same local energy
                fr = ws.position_grads(RT, swct=self.use_swct)
                e_L = fr["e_L"]
                self._force_HF.append(fr["force_HF"].cpu().numpy())
                self._force_PP.append(fr["force_PP"].cpu().numpy())
                self._E_L_force_PP.append((e_L[:, None, None] * fr["force_PP"]).cpu().numpy())
            else:
                e_L = ws.local_energy(RT)
            if self.comput_param_deriv:
                # O_k of every parameter block, flattened in block order (jqmc_mcmc.py:854-873)
                # (the samples live in host memory, as the reference keeps them: W x P doubles per step)
                d = ws.dln_psi_dparams()
                O = torch.cat([v.reshape(self.num_walkers, -1) for v in d.values()], dim=1)
                host = torch.empty(O.shape, dtype=O.dtype, pin_memory=True)
                host.copy_(O)
                self._dln_WF.append(host)
            R = ws.as_factor
            w_L = (R / torch.clamp(R, min=self.epsilon_AS)) ** 2
            self._e_L.append(e_L.cpu().numpy())
            self._w_L.append(w_L.cpu().numpy())
            self.mcmc_counter += 1
            if ctl.should_stop():
                break

    @property
    def e_L(self) -> np.ndarray:
        return np.asarray(self._e_L).reshape(-1, self.num_walkers)

    @property
    def e_L2(self) -> np.ndarray:
        return self.e_L**2

    @property
    def w_L(self) -> np.ndarray:
        return np.asarray(self._w_L).reshape(-1, self.num_walkers)

    @property
    def accepted_moves(self) -> int:
        return int(self.state.accepted.sum().item())

    @property
    def rejected_moves(self) -> int:
        return int(self.state.rejected.sum().item())

    def get_dln_WF(self, num_mcmc_warmup_steps: int = 50) -> np.ndarray:
        """O_k samples (num_mcmc - warmup, num_walker, num_param) as jqmc_mcmc.py:1372-1514 returns
        them (blocks j1, j2, J3 matrix, lambda matrix, each flattened)."""
        if not self.comput_param_deriv:
            raise RuntimeError("MCMC was built with comput_param_deriv=False")
        return torch.stack(self._dln_WF[num_mcmc_warmup_steps:]).cpu().numpy()

    def get_natural_gradient(self, num_mcmc_warmup_steps: int = 50, epsilon: float = 1.0e-3, branch: str = "auto"):
        """SR step theta = (S + epsilon I)^-1 f from the stored samples (the SR branch of
        run_optimize, jqmc_mcmc.py:2952-3362), computed on the device by jqmc_amd.sr."""
        from .sr import natural_gradient

        if not self.comput_param_deriv:
            raise RuntimeError("MCMC was built with comput_param_deriv=False")
        O = torch.stack(self._dln_WF[num_mcmc_warmup_steps:]).to(device())
        w = self.w_L[num_mcmc_warmup_steps:]
        e = self.e_L[num_mcmc_warmup_steps:]
        return natural_gradient(O, w, e, epsilon=epsilon, branch=branch)

    def get_aF(self, num_mcmc_warmup_steps: int = 50, num_mcmc_bin_blocks: int = 10):
        """(force_mean, force_std), (n_atom, 3) in Ha/bohr: jackknife over binned, weighted samples of
        F = -(<F_HF> + 2 (<E_L F_PP> - <E_L> <F_PP>)) (jqmc_mcmc.py:1191-1370), one process."""
        if not self.comput_position_deriv:
            raise RuntimeError("construct MCMC with comput_position_deriv=True to sample forces")
        if self.mcmc_counter - num_mcmc_warmup_steps < num_mcmc_bin_blocks:
            raise ValueError("(mcmc_counter - num_mcmc_warmup_steps) should be larger than num_mcmc_bin_blocks.")
        k = num_mcmc_warmup_steps
        return jackknife_forces(self.w_L[k:], self.e_L[k:], np.asarray(self._force_HF[k:]),
                                np.asarray(self._force_PP[k:]), np.asarray(self._E_L_force_PP[k:]),
                                num_mcmc_bin_blocks)

    def get_E(self, num_mcmc_warmup_steps: int = 50, num_mcmc_bin_blocks: int = 10):
        """(E_mean, E_std, Var_mean, Var_std) by jackknife over binned, weighted samples
        (jqmc_mcmc.py:1064-1189), one process."""
        if self.mcmc_counter - num_mcmc_warmup_steps < num_mcmc_bin_blocks:
            raise ValueError("(mcmc_counter - num_mcmc_warmup_steps) should be larger than num_mcmc_bin_blocks.")
        e = self.e_L[num_mcmc_warmup_steps:]
        w = self.w_L[num_mcmc_warmup_steps:]
        wb = np.ravel([np.sum(a, axis=0) for a in np.array_split(w, num_mcmc_bin_blocks, axis=0)])
        web = np.ravel([np.sum(a, axis=0) for a in np.array_split(w * e, num_mcmc_bin_blocks, axis=0)])
        we2b = np.ravel([np.sum(a, axis=0) for a in np.array_split(w * e * e, num_mcmc_bin_blocks, axis=0)])
        E_jk = (web.sum() - web) / (wb.sum() - wb)
        E2_jk = (we2b.sum() - we2b) / (wb.sum() - wb)
        V_jk = E2_jk - E_jk**2
        M = E_jk.size
        E_mean = E_jk.mean()
        E_std = np.sqrt((M - 1) * np.mean((E_jk - E_mean) ** 2))
        V_mean = V_jk.mean()
        V_std = np.sqrt((M - 1) * np.mean((V_jk - V_mean) ** 2))
        return float(E_mean), float(E_std), float(V_mean), float(V_std)


def jackknife_forces(w_L, e_L, force_HF, force_PP, E_L_force_PP, num_mcmc_bin_blocks: int):
    """Jackknife force estimate of get_aF (jqmc_mcmc.py:1229-1370) from per-step, per-walker samples
    w_L, e_L (M, W) and force products (M, W, n_atom, 3): bins of consecutive steps per walker, then
    leave-one-bin-out estimates F_j = -(<F_HF>_j + 2 (<E_L F_PP>_j - <E_L>_j <F_PP>_j)).
    Returns (mean, std) over the M_binned = num_mcmc_bin_blocks * W estimates."""
    nw = w_L.shape[1]

    def binned(x):
        parts = np.array_split(x, num_mcmc_bin_blocks, axis=0)
        return np.concatenate([np.sum(p, axis=0)[None] for p in parts], axis=0).reshape(
            (num_mcmc_bin_blocks * nw,) + x.shape[2:])

    w = np.asarray(w_L)
    wb = binned(w)
    web = binned(w * np.asarray(e_L))
    whf = binned(w[..., None, None] * force_HF)
    wpp = binned(w[..., None, None] * force_PP)
    wefp = binned(w[..., None, None] * E_L_force_PP)
    W_tot = wb.sum() - wb
    d = W_tot[:, None, None]
    f_hf = -(whf.sum(axis=0) - whf) / d
    f_pulay = -2.0 * ((wefp.sum(axis=0) - wefp) / d
                      - ((web.sum() - web) / W_tot)[:, None, None] * (wpp.sum(axis=0) - wpp) / d)
    f_jk = f_hf + f_pulay
    M = f_jk.shape[0]
    mean = f_jk.mean(axis=0)
    std = np.sqrt((M - 1) * np.mean((f_jk - mean) ** 2, axis=0))
    return mean, std


def philox_key(seed: int, counter: int = 0) -> np.ndarray:
    """The key of this build's counter-based stream, standing in for the reference's JAX PRNG key:
    (seed, counter) as uint64.  Walker w of a batched call draws from (seed, w, counter + m) for its
    m-th proposal, so a key advances by the number of draws taken, never by splitting."""
    return np.array([int(seed) & 0xFFFFFFFFFFFFFFFF, int(counter)], dtype=np.uint64)


def _key_parts(key):
    k = np.asarray(key, dtype=np.uint64)
    if k.ndim == 2:
        if not np.all(k == k[0]):
            raise ValueError("a batched call draws walker w's numbers from (seed, w, counter): every row of the key "
                             "must hold the same (seed, counter)")
        k = k[0]
    if k.shape != (2,):
        raise ValueError("key must be philox_key(seed, counter) of shape (2,) or (W, 2)")
    return int(k[0]), int(k[1])


def _with_key(key, seed, counter):
    k = np.asarray(key)
    out = philox_key(seed, counter)
    return np.broadcast_to(out, k.shape).copy() if k.ndim == 2 else out


def _update_electron_positions(init_r_up_carts, init_r_dn_carts, jax_PRNG_key, num_mcmc_per_measurement,
                               hamiltonian_data, Dt, epsilon_AS, geminal_inv_init, geminal_init):
    """The Metropolis step with the reference's carry (jqmc_mcmc.py:4278-4533): returns
    ``(accepted_moves, rejected_moves, r_up_carts, r_dn_carts, key, geminal_inv, geminal)`` after
    ``num_mcmc_per_measurement`` single-electron proposals.  Positions ``(N_sigma, 3)`` or, for a batch
    of walkers (the reference vmaps this function), ``(W, N_sigma, 3)``; ``geminal_inv_init`` and
    ``geminal_init`` likewise with ``(N_up, N_up)``.  The key is ``philox_key(seed, counter)``; the
    returned key has its counter advanced by the proposals taken.  Each call rebuilds the walker caches
    from the positions (use ``WalkerState`` to keep them resident across steps)."""
    from ._api import _batched, as_dev

    seed, counter = _key_parts(jax_PRNG_key)
    ru, rd, batched = _batched(init_r_up_carts, init_r_dn_carts)
    pack = system_for(hamiltonian_data, NN=1)
    ws = WalkerState(pack, ru.shape[0]).init(ru, rd)
    if geminal_inv_init is not None:
        ws.Ginv.copy_(as_dev(geminal_inv_init).reshape(ws.Ginv.shape))
    if geminal_init is not None:
        ws.G.copy_(as_dev(geminal_init).reshape(ws.G.shape))
    n = int(num_mcmc_per_measurement)
    ws.sweep(n, float(Dt), float(epsilon_AS), seed, counter)
    out = (ws.accepted.clone(), ws.rejected.clone(), ws.r_up.clone(), ws.r_dn.clone(),
           _with_key(jax_PRNG_key, seed, counter + n), ws.Ginv.clone(), ws.G.clone())
    if batched:
        return out
    return tuple(x[0] if isinstance(x, torch.Tensor) else x for x in out)
