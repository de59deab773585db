"""Run control of the drivers: wall-clock budget, external stop file and checkpoints.

* ``max_time`` (seconds): ``run`` stops after the step in which the budget is used up
  (jqmc_mcmc.py:942-948, jqmc_gfmc.py:6388-6395).
* External stop: ``run`` writes ``external_control_mcmc.toml`` (``external_control_gfmc.toml`` for
  the LRDMC drivers) in the working directory with ``[external_control] stop = false``; setting
  ``stop = true`` there ends the run after the current step, and the file is removed when ``run``
  returns (jqmc_mcmc.py:473-487, 950-967, 1007-1012).
* Checkpoints: ``save_checkpoint(driver, path)`` writes the driver's configuration, counters, the
  position of its random streams (the Philox move / step counters and, for LRDMC, the state of the
  comb's NumPy generator), the walker positions and the accumulated observables to one NumPy
  ``.npz`` file; ``load_checkpoint(path, hamiltonian_data)`` rebuilds the driver with the walker
  caches set up afresh from those positions, so that ``run`` continues the same chain.  The
  reference writes the same groups (driver_config, rng_state, walker_state, observables) per MPI
  rank to HDF5 (_checkpoint.py:146-388) and stores the Hamiltonian with them; here the caller passes
  the Hamiltonian back in.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

STOP_FILE_MCMC = "external_control_mcmc.toml"
STOP_FILE_GFMC = "external_control_gfmc.toml"


class RunControl:
    """The wall-clock budget and the external stop file of one ``run`` call."""

    def __init__(self, stop_file: str, max_time: float):
        self.stop_file = stop_file
        self.max_time = float(max_time)
        self.t0 = time.perf_counter()
        self.reason: str | None = None
        with open(self.stop_file, "w") as f:
            f.write("[external_control]\nstop = false\n")

    def should_stop(self) -> bool:
        """True once the budget is used up or the stop file says stop (read after every step)."""
        if time.perf_counter() - self.t0 > self.max_time:
            self.reason = f"max_time = {self.max_time} s exceeded"
            return True
        if os.path.isfile(self.stop_file):
            import tomli

            try:
                with open(self.stop_file, "rb") as f:
                    stop = bool(tomli.load(f).get("external_control", {}).get("stop", False))
            except (OSError, ValueError):
                stop = False
            if stop:
                self.reason = f"stop = true in {self.stop_file}"
                return True
        return False

    def close(self) -> None:
        if os.path.isfile(self.stop_file):
            os.remove(self.stop_file)


def _pack(arrays: dict, obj: dict) -> None:
    for k, v in obj.items():
        arrays[k] = np.asarray(v)


def save_checkpoint(driver, path: str) -> None:
    """Write ``driver`` (an ``MCMC``, ``GFMC_n`` or ``GFMC_t``) to ``path`` (a ``.npz`` file)."""
    from .jqmc_gfmc import GFMC_n, GFMC_t
    from .jqmc_mcmc import MCMC

    ws = driver.state
    arrays: dict = {
        "r_up": ws.r_up.cpu().numpy(),
        "r_dn": ws.r_dn.cpu().numpy(),
    }
    if isinstance(driver, MCMC):
        kind = "MCMC"
        config = {k: getattr(driver, k) for k in (
            "mcmc_seed", "num_walkers", "num_mcmc_per_measurement", "Dt", "epsilon_AS", "comput_position_deriv",
            "comput_param_deriv", "random_discretized_mesh", "use_swct", "refresh_interval")}
        counters = {"mcmc_counter": driver.mcmc_counter, "move": driver._move, "since_refresh": driver._since_refresh}
        arrays["accepted"] = ws.accepted.cpu().numpy()
        arrays["rejected"] = ws.rejected.cpu().numpy()
        n = driver.mcmc_counter
        arrays["e_L"] = np.asarray(driver._e_L).reshape(n, -1) if n else np.zeros((0, driver.num_walkers))
        arrays["w_L"] = np.asarray(driver._w_L).reshape(n, -1) if n else np.zeros((0, driver.num_walkers))
        for key in ("force_HF", "force_PP", "E_L_force_PP"):
            seq = getattr(driver, f"_{key}")
            if seq:
                arrays[key] = np.asarray(seq)
        if driver._dln_WF:
            arrays["dln_WF"] = np.stack([t.numpy() for t in driver._dln_WF])
    elif isinstance(driver, (GFMC_n, GFMC_t)):
        kind = type(driver).__name__
        keys = ["num_walkers", "mcmc_seed", "alat", "random_discretized_mesh", "non_local_move",
                "num_gfmc_collect_steps"]
        keys += ["num_mcmc_per_measurement"] if kind == "GFMC_n" else ["tau", "max_projections"]
        config = {k: getattr(driver, k) for k in keys}
        counters = {"mcmc_counter": driver.mcmc_counter, "E_scf": driver.E_scf,
                    "num_survived_walkers": driver.num_survived_walkers,
                    "num_killed_walkers": driver.num_killed_walkers}
        _pack(arrays, {"e_L": driver._e_L, "e_L2": driver._e_L2, "w_L": driver._w_L, "G_L": driver._G_L,
                       "G_e_L": driver._G_e_L})
        name, keys_mt, pos, has_gauss, cached = driver._np_rng.get_state()
        arrays["mt_keys"] = keys_mt
        counters.update({"mt_name": name, "mt_pos": int(pos), "mt_has_gauss": int(has_gauss),
                         "mt_cached_gaussian": float(cached)})
    else:
        raise TypeError(f"cannot checkpoint a {type(driver).__name__}")
    config.update({"Nv": driver.pack.nv, "precision": driver.pack.precision})
    meta = {"driver_type": kind, "driver_config": config, "counters": counters}
    np.savez(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **arrays)


def load_checkpoint(path: str, hamiltonian_data):
    """The driver saved by ``save_checkpoint`` at ``path``, for ``hamiltonian_data``."""
    from .jqmc_gfmc import GFMC_n, GFMC_t
    from .jqmc_mcmc import MCMC

    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    meta = json.loads(bytes(data.pop("meta")).decode())
    kind, cfg, cnt = meta["driver_type"], meta["driver_config"], meta["counters"]
    if kind == "MCMC":
        d = MCMC(hamiltonian_data, init_r_up_carts=data["r_up"], init_r_dn_carts=data["r_dn"], **cfg)
        d.mcmc_counter = int(cnt["mcmc_counter"])
        d._move = int(cnt["move"])
        d._since_refresh = 0  # the caches are fresh
        import torch

        d.state.accepted.copy_(torch.as_tensor(data["accepted"]))
        d.state.rejected.copy_(torch.as_tensor(data["rejected"]))
        d._e_L = list(data["e_L"])
        d._w_L = list(data["w_L"])
        for key in ("force_HF", "force_PP", "E_L_force_PP"):
            if key in data:
                setattr(d, f"_{key}", list(data[key]))
        if "dln_WF" in data:
            d._dln_WF = [torch.from_numpy(np.ascontiguousarray(x)).pin_memory() for x in data["dln_WF"]]
        return d
    cls = {"GFMC_n": GFMC_n, "GFMC_t": GFMC_t}[kind]
    d = cls(hamiltonian_data, init_r_up_carts=data["r_up"], init_r_dn_carts=data["r_dn"], E_scf=cnt["E_scf"], **cfg)
    d.mcmc_counter = int(cnt["mcmc_counter"])
    d.num_survived_walkers = int(cnt["num_survived_walkers"])
    d.num_killed_walkers = int(cnt["num_killed_walkers"])
    for key in ("e_L", "e_L2", "w_L", "G_L", "G_e_L"):
        setattr(d, f"_{key}", [float(x) for x in data[key]])
    d._np_rng.set_state((cnt["mt_name"], data["mt_keys"], cnt["mt_pos"], cnt["mt_has_gauss"],
                         cnt["mt_cached_gaussian"]))
    return d
