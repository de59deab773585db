"""ctypes binding of ``libjqmc_hip.so`` (C ABI declared in ``include/jqmc_hip.h``).

The library is built in-tree by ``__graft_entry__.build()`` (``hipcc --offload-arch=gfx950``).
Loading never falls back to another implementation: a missing or stale library raises
``JqmcLibraryError``.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_int64, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjqmc_hip.so")
ABI_VERSION = 6

_dp = POINTER(c_double)
_ip = POINTER(c_int32)
_lp = POINTER(c_int64)


class JqmcLibraryError(RuntimeError):
    """The HIP library is missing, stale or failed to load."""


class JqmcError(RuntimeError):
    """A C-ABI call returned a non-zero jqmc_status."""

    def __init__(self, fn: str, code: int, message: str):
        super().__init__(f"{fn} failed with status {code}: {message}")
        self.code = code


class SystemDesc(ctypes.Structure):
    _fields_ = [
        ("n_nuc", c_int32),
        ("nuc_pos", _dp),
        ("nuc_charge", _dp),
        ("n_ao", c_int32),
        ("n_prim", c_int32),
        ("ao_nuc", _ip),
        ("ao_pow", _ip),
        ("ao_prim_ptr", _ip),
        ("prim_exp", _dp),
        ("prim_coef", _dp),
        ("n_terms", c_int32),
        ("ao_term_ptr", _ip),
        ("term_coef", _dp),
        ("term_pow", _ip),
        ("n_up", c_int32),
        ("n_dn", c_int32),
        ("orb_is_mo", c_int32),
        ("n_orb", c_int32),
        ("mo_coef_up", _dp),
        ("mo_coef_dn", _dp),
        ("lambda_matrix", _dp),
        ("j1_type", c_int32),
        ("j1_param", c_double),
        ("j1_zeff", _dp),
        ("j2_type", c_int32),
        ("j2_param", c_double),
        ("j3_on", c_int32),
        ("j3_matrix", _dp),
        ("ecp_on", c_int32),
        ("ecp_lmax", c_int32),
        ("ecp_max_terms", c_int32),
        ("ecp_n_terms", _ip),
        ("ecp_term_l", _ip),
        ("ecp_term_exp", _dp),
        ("ecp_term_coef", _dp),
        ("ecp_term_pow", _dp),
        ("ecp_max_local", c_int32),
        ("ecpl_n_terms", _ip),
        ("ecpl_exp", _dp),
        ("ecpl_coef", _dp),
        ("ecpl_pow", _dp),
        ("nv", c_int32),
        ("mesh_points", _dp),
        ("mesh_weights", _dp),
        ("nn", c_int32),
        ("precision", c_int32),
    ]


PRECISION_PLAIN_FP64, PRECISION_MIXED_FP32 = 0, 1


class LrdmcParams(ctypes.Structure):
    _fields_ = [
        ("mode", c_int32),
        ("n_steps", c_int32),
        ("alat", c_double),
        ("E_scf", c_double),
        ("tau", c_double),
        ("non_local_move", c_int32),
        ("random_mesh", c_int32),
    ]


TMOVE, DLTMOVE = 0, 1
PARAM_J1, PARAM_J2, PARAM_J3, PARAM_LAMBDA = 1, 2, 4, 8
LRDMC_FIXED_STEPS, LRDMC_FIXED_TAU = 0, 1


class Walkers(ctypes.Structure):
    _fields_ = [
        ("r_up", c_void_p),
        ("r_dn", c_void_p),
        ("ao", c_void_p),
        ("orb", c_void_p),
        ("pair", c_void_p),
        ("jfield", c_void_p),
        ("G", c_void_p),
        ("Ginv", c_void_p),
        ("logabsdet", c_void_p),
        ("sign", c_void_p),
        ("jastrow", c_void_p),
        ("as_factor", c_void_p),
        ("accepted", c_void_p),
        ("rejected", c_void_p),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    "jqmc_abi_version": (c_int32, []),
    "jqmc_last_error": (ctypes.c_char_p, []),
    "jqmc_system_create": (c_int32, [POINTER(SystemDesc), POINTER(c_void_p)]),
    "jqmc_system_destroy": (c_int32, [c_void_p]),
    "jqmc_system_precision": (c_int32, [c_void_p, POINTER(c_int32)]),
    "jqmc_aos_value_grad_lap": (
        c_int32,
        [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "jqmc_orbs_value_grad_lap": (
        c_int32,
        [c_void_p, c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "jqmc_walkers_init_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_walkers_init": (c_int32, [c_void_p, c_int64, POINTER(Walkers), c_void_p, c_int64, c_void_p]),
    "jqmc_walkers_refresh_inverse": (c_int32, [c_void_p, c_int64, POINTER(Walkers), c_void_p]),
    "jqmc_walkers_refresh_orbitals": (c_int32, [c_void_p, c_int64, POINTER(Walkers), c_void_p]),
    "jqmc_system_set_mixed_guard": (c_int32, [c_void_p, c_double, c_double]),
    "jqmc_lrdmc_branching_static": (c_int32, [c_int64, c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p,
                                              c_void_p, c_void_p]),
    "jqmc_point_ratios": (c_int32, [c_void_p, c_int64, POINTER(Walkers), c_int64, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    "jqmc_vmc_sweep": (
        c_int32,
        [c_void_p, c_int64, POINTER(Walkers), c_int32, c_double, c_double, c_uint64, c_int64, c_void_p],
    ),
    "jqmc_rotation_matrices": (c_int32, [c_int64, c_uint64, c_int64, c_int32, c_void_p, c_void_p]),
    "jqmc_local_energy_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_local_energy": (
        c_int32,
        [c_void_p, c_int64, POINTER(Walkers), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
         c_void_p],
    ),
    "jqmc_lrdmc_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_lrdmc_projection": (
        c_int32,
        [c_void_p, c_int64, POINTER(Walkers), POINTER(LrdmcParams), c_uint64, c_int64, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    ),
    "jqmc_lrdmc_local_energy": (
        c_int32,
        [c_void_p, c_int64, POINTER(Walkers), POINTER(LrdmcParams), c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
         c_void_p],
    ),
    "jqmc_lrdmc_projection_multi": (
        c_int32,
        [c_int32, POINTER(c_void_p), _lp, POINTER(Walkers), POINTER(LrdmcParams), c_uint64, _lp,
         POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p), POINTER(c_void_p),
         POINTER(c_void_p), POINTER(c_void_p), _lp, c_void_p],
    ),
    "jqmc_lrdmc_branching_workspace_bytes": (c_int32, [c_int64, _lp]),
    "jqmc_lrdmc_branching": (
        c_int32,
        [c_int64, c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    ),
    "jqmc_param_layout": (c_int32, [c_void_p, c_int32, _lp, _lp]),
    "jqmc_dln_psi_dparams_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_dln_psi_dparams": (
        c_int32, [c_void_p, c_int64, POINTER(Walkers), c_int32, c_void_p, c_int64, c_void_p, c_int64, c_void_p]),
    "jqmc_vmc_sweep_batched_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_vmc_sweep_batched": (
        c_int32, [c_void_p, c_int64, POINTER(Walkers), c_int32, c_double, c_double, c_uint64, c_int64, c_void_p,
                  c_void_p, c_int64, c_void_p]),
    "jqmc_vmc_sweep_counted": (
        c_int32, [c_void_p, c_int64, POINTER(Walkers), c_int32, c_double, c_double, c_uint64, c_int64, c_void_p,
                  c_void_p]),
    "jqmc_rotation_matrices_counted": (c_int32, [c_int64, c_uint64, c_int64, c_void_p, c_int32, c_void_p, c_void_p]),
    "jqmc_counter_add": (c_int32, [c_void_p, c_int64, c_void_p]),
    "jqmc_position_grads_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_position_grads": (
        c_int32, [c_void_p, c_int64, POINTER(Walkers), c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                  c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "jqmc_walkers_reconfigure_workspace_bytes": (c_int32, [c_void_p, c_int64, _lp]),
    "jqmc_walkers_reconfigure": (c_int32, [c_void_p, c_int64, POINTER(Walkers), c_void_p, c_void_p, c_int64, c_void_p]),
    "jqmc_system_set_tuning": (c_int32, [c_void_p, c_int32, c_int32]),
    "jqmc_system_get_tuning": (c_int32, [c_void_p, c_int32, POINTER(c_int32)]),
    "jqmc_sr_colstats": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "jqmc_sr_gram": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_int64, c_void_p]),
    "jqmc_sr_gram_f32": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_int64, c_void_p]),
    "jqmc_sr_project": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None


def load(path: str | None = None):
    """Load the HIP library once (``path`` defaults to the in-tree build) and bind every signature."""
    global _lib
    if _lib is not None:
        return _lib
    path = LIB_PATH if path is None else path
    # torch ships its own HIP runtime; it must be the one in the process before this library
    # resolves libamdhip64, or torch's device initialisation fails later.
    import torch  # noqa: F401

    if not os.path.exists(path):
        raise JqmcLibraryError(
            f"{path} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    try:
        lib = ctypes.CDLL(path)
    except OSError as exc:  # pragma: no cover - depends on the loader
        raise JqmcLibraryError(f"cannot load {path}: {exc}") from exc
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.jqmc_abi_version() != ABI_VERSION:
        raise JqmcLibraryError("libjqmc_hip.so ABI version mismatch; rebuild the library")
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise ``JqmcError`` on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise JqmcError(name, rc, lib.jqmc_last_error().decode(errors="replace"))
