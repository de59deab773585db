"""Launch-shape knobs and the autotuner (jqmc_amd/autotune.py): every E_L walkers-per-block setting
gives bit-identical local energies, every proposal walkers-per-block setting takes identical
decisions (bit-identical positions), bad knob values are refused, and autotune() applies a valid
choice and round-trips through its JSON cache."""

import json

import pytest
import torch

from jqmc_amd import _lib, synthetic
from jqmc_amd._api import WalkerState
from jqmc_amd._pack import system_for
from jqmc_amd.autotune import CANDIDATES, autotune, get_tuning, set_tuning

pytestmark = pytest.mark.gpu


def _state(prec="fp64", W=96):
    s = synthetic.chain_ecp(10, seed=1)
    pack = system_for(s.hamiltonian_data, Nv=6, precision=prec)
    ru, rd = s.initial_walkers(W, seed=3)
    return pack, ru, rd


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_knobs_change_no_result_bit(prec):
    pack, ru, rd = _state(prec)
    ws = WalkerState(pack, len(ru)).init(ru, rd)
    RT = ws.rotations(2, 0)
    set_tuning(pack, el_walkers_per_block=0)
    ref = ws.local_energy(RT).clone()
    for w in CANDIDATES["el_walkers_per_block"]:
        set_tuning(pack, el_walkers_per_block=w)
        assert torch.equal(ws.local_energy(RT), ref), w
    set_tuning(pack, el_walkers_per_block=0)
    pos = {}
    for w in (0,) + CANDIDATES["sweep_walkers_per_block"]:
        set_tuning(pack, sweep_walkers_per_block=w)
        wk = WalkerState(pack, len(ru)).init(ru, rd)
        wk.sweep(12, 0.5, 1e-4, 7, 0)
        pos[w] = (wk.r_up.clone(), wk.r_dn.clone(), wk.accepted.clone())
    for w in CANDIDATES["sweep_walkers_per_block"]:
        assert all(torch.equal(a, b) for a, b in zip(pos[w], pos[0])), w


@pytest.mark.parametrize("prec,rtol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_el_team_matches_one_wave_within_rounding(prec, rtol):
    """A team of 2 or 4 waves per walker adds its partial sums in wave order: e_L equals the one-wave
    result up to rounding (FP64: 1e-11 of max(1, |e_L|); mixed: 1e-5, the FP32 J3 sums regrouped),
    and repeats bit for bit."""
    pack, ru, rd = _state(prec)
    ws = WalkerState(pack, len(ru)).init(ru, rd)
    RT = ws.rotations(4, 0)
    set_tuning(pack, el_walkers_per_block=0, el_waves_per_walker=0)
    ref = ws.local_energy(RT).clone()
    for team in (2, 4):
        set_tuning(pack, el_waves_per_walker=team)
        assert get_tuning(pack)["el_waves_per_walker"] == team
        e = ws.local_energy(RT).clone()
        assert torch.equal(ws.local_energy(RT), e), team
        err = ((e - ref).abs() / ref.abs().clamp(min=1.0)).max().item()
        assert err <= rtol, (team, err)
    set_tuning(pack, el_waves_per_walker=0)


def test_bad_knob_values_are_refused():
    pack, _, _ = _state()
    with pytest.raises(_lib.JqmcError):
        set_tuning(pack, el_waves_per_walker=3)
    with pytest.raises(_lib.JqmcError):
        set_tuning(pack, el_walkers_per_block=3)
    with pytest.raises(_lib.JqmcError):
        set_tuning(pack, sweep_walkers_per_block=3)
    with pytest.raises(_lib.JqmcError):
        _lib.call("jqmc_system_set_tuning", pack.handle, 99, 1)


def test_autotune_applies_and_caches(tmp_path):
    pack, ru, rd = _state(W=128)
    ws = WalkerState(pack, len(ru)).init(ru, rd)
    r0 = ws.r_up.clone()
    cache = tmp_path / "tune.json"
    out = autotune(ws, reps=1, cache_path=str(cache))
    assert not out["cached"]
    assert get_tuning(pack) == out["best"]
    for k, v in out["best"].items():
        assert v in CANDIDATES[k], (k, v)
    assert torch.equal(ws.r_up, r0), "the caller's walkers do not move"
    assert len(out["el_ms"]) == len(CANDIDATES["el_walkers_per_block"]) + len(CANDIDATES["el_waves_per_walker"]) - 1
    assert len(out["sweep_ms"]) == 3
    set_tuning(pack, el_walkers_per_block=0, sweep_walkers_per_block=0)
    again = autotune(ws, cache_path=str(cache))
    assert again["cached"] and get_tuning(pack) == out["best"]
    assert list(json.loads(cache.read_text()).values())[0] == out["best"]
