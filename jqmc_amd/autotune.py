"""Launch-shape autotuner for a system on this GPU.

The walkers per block of the E_L and move-synchronous proposal kernels, and the waves per walker of
the E_L kernel, are chosen per system by timing every candidate on the caller's walker count, on
scratch walkers built at the caller's positions, so the caller's chain does not move.  The
walkers-per-block knobs (jqmc_system_set_tuning) change no result bit; a team of two waves per walker
changes e_L in the last bits, as a fixed function of the choice (tests/test_gpu_autotune.py checks
both).  The reference has no counterpart (XLA picks its own launch shapes).
"""

from __future__ import annotations

import ctypes
import json
import os

import torch

from . import _lib

KNOBS = {"el_walkers_per_block": 2, "sweep_walkers_per_block": 3, "el_waves_per_walker": 4}
CANDIDATES = {"el_walkers_per_block": (1, 2, 4), "sweep_walkers_per_block": (1, 2, 4), "el_waves_per_walker": (1, 2)}


def set_tuning(pack, **values) -> None:
    """Set knobs by name (0 restores the library default)."""
    for name, v in values.items():
        _lib.call("jqmc_system_set_tuning", pack.handle, KNOBS[name], int(v))


def get_tuning(pack) -> dict:
    out = {}
    for name, k in KNOBS.items():
        v = ctypes.c_int32()
        _lib.call("jqmc_system_get_tuning", pack.handle, k, ctypes.byref(v))
        out[name] = v.value
    return out


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _key(pack, W):
    return f"n_ao={pack.n_ao},n_el={pack.n_el},n_orb={pack.n_orb},nv={pack.nv},{pack.precision},W={W}"


def autotune(ws, n_moves: int | None = None, reps: int = 3, seed: int = 1, cache_path: str | None = None) -> dict:
    """Time every candidate walkers-per-block of the E_L kernel and of the proposal kernel for the
    walkers ``ws`` (a WalkerState), apply the fastest to ``ws.pack`` and return {'best': {...},
    'el_ms': {...}, 'sweep_ms': {...}}.  With ``cache_path`` the choice is stored in (and, when
    present, read from) a JSON file keyed by the system shape, precision and walker count."""
    from ._api import WalkerState

    pack, W = ws.pack, ws.W
    key = _key(pack, W)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cached = json.load(f).get(key)
        if cached:
            set_tuning(pack, **cached)
            return {"best": cached, "cached": True}
    scratch = WalkerState(pack, W).init(ws.r_up, ws.r_dn)
    RT = scratch.rotations(seed, 0)
    el = {}
    for wpb in CANDIDATES["el_walkers_per_block"]:
        set_tuning(pack, el_walkers_per_block=wpb, el_waves_per_walker=1)
        el[f"wpb{wpb}"] = ((wpb, 1), _time(lambda: scratch.local_energy(RT), reps))
    for team in CANDIDATES["el_waves_per_walker"]:
        if team == 1:  # the one-wave shapes above
            continue
        set_tuning(pack, el_walkers_per_block=1, el_waves_per_walker=team)
        el[f"team{team}"] = ((1, team), _time(lambda: scratch.local_energy(RT), reps))
    wpb_el, team_el = min(el.values(), key=lambda v: v[1])[0]
    n_moves = n_moves or pack.n_el
    sw = {}
    mv = [0]

    def sweep():
        scratch.sweep(n_moves, 0.5, 1e-4, seed, 10**9 + mv[0])
        mv[0] += n_moves

    for wpb in CANDIDATES["sweep_walkers_per_block"]:
        set_tuning(pack, sweep_walkers_per_block=wpb)
        sw[wpb] = _time(sweep, reps)
    wpb_sw = min(sw, key=sw.get)
    best = {"el_walkers_per_block": wpb_el, "sweep_walkers_per_block": wpb_sw, "el_waves_per_walker": team_el}
    set_tuning(pack, **best)
    del scratch
    if cache_path:
        data = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                data = json.load(f)
        data[key] = best
        with open(cache_path, "w") as f:
            json.dump(data, f, indent=1)
    return {"best": best, "el_ms": {name: t for name, (_, t) in el.items()},
            "sweep_ms": {f"wpb{w}": t for w, t in sw.items()}, "cached": False}
