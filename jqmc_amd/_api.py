"""Batched GPU implementations behind the reference-named API functions.

Walker arrays are ``torch.float64`` CUDA (ROCm) tensors; every computation runs in the
gfx950 kernels of ``libjqmc_hip.so`` on torch's current stream.  No CPU fallback exists:
without a GPU or without the built library these functions raise.
"""

from __future__ import annotations

import ctypes
from dataclasses import replace

import numpy as np
import torch

from . import _lib
from ._pack import SystemPack, system_for
from .atomic_orbital import AOs_cart_data
from .coulomb_potential import Coulomb_potential_data
from .determinant import Geminal_data
from .hamiltonians import Hamiltonian_data
from .jastrow_factor import Jastrow_data
from .molecular_orbital import MOs_data
from .wavefunction import Wavefunction_data

F64 = torch.float64


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jqmc_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def as_dev(x, dtype=F64) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device(), dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _batched(r_up, r_dn):
    """Return (W, N_up, 3), (W, N_dn, 3) device tensors and whether the input was batched."""
    ru, rd = as_dev(r_up), as_dev(r_dn)
    batched = ru.dim() == 3
    if not batched:
        ru, rd = ru.reshape(1, -1, 3), rd.reshape(1, -1, 3)
    return ru.contiguous(), rd.contiguous(), batched


class WalkerState:
    """Device-resident walker state (mirror of ``jqmc_walkers``) for ``W`` walkers."""

    def __init__(self, pack: SystemPack, W: int):
        dev = device()
        self.pack = pack
        self.W = W
        ne, na, no, nu = pack.n_el, pack.n_ao, pack.n_orb, pack.n_up
        z = lambda *s: torch.zeros(s, dtype=F64, device=dev)
        # AO and J3-field caches in the system's arithmetic type (FP32 in the mixed form); the
        # orbital and pair caches feed the determinant and stay FP64 in both forms
        ct = torch.float32 if pack.precision == "mixed" else F64
        zc = lambda *s: torch.zeros(s, dtype=ct, device=dev)
        self.r_up = z(W, pack.n_up, 3)
        self.r_dn = z(W, pack.n_dn, 3)
        self.ao = zc(W, ne, na)
        self.orb = z(W, ne, no)
        self.pair = z(W, ne, no)
        self.jfield = zc(W, ne, na) if pack.j3_on else None
        self.G = z(W, nu, nu)
        self.Ginv = z(W, nu, nu)
        self.logabsdet = z(W)
        self.sign = z(W)
        self.jastrow = z(W)
        self.as_factor = z(W)
        self.accepted = torch.zeros(W, dtype=torch.int64, device=dev)
        self.rejected = torch.zeros(W, dtype=torch.int64, device=dev)
        self._ws_el = None
        self._ws_init = None
        self._ws_dp = None
        self._ws_fr = None
        self._ws_sw = None
        self._ws_rc = None
        self._ws_lr = None

    def struct(self) -> _lib.Walkers:
        s = _lib.Walkers()
        for name in ("r_up", "r_dn", "ao", "orb", "pair", "jfield", "G", "Ginv", "logabsdet", "sign", "jastrow",
                     "as_factor", "accepted", "rejected"):
            t = getattr(self, name)
            setattr(s, name, t.data_ptr() if t is not None and t.numel() > 0 else None)
        return s

    def _workspace(self, attr, fn):
        nbytes = ctypes.c_int64()
        _lib.call(fn, self.pack.handle, self.W, ctypes.byref(nbytes))
        ws = getattr(self, attr)
        if nbytes.value > 0 and (ws is None or ws.numel() * 8 < nbytes.value):
            ws = torch.empty((nbytes.value + 7) // 8, dtype=F64, device=device())
            setattr(self, attr, ws)
        return ws, nbytes.value

    def init(self, r_up, r_dn) -> "WalkerState":
        """Positions -> every cache, G, G^-1, ln|det G|, J (jqmc_walkers_init)."""
        self.r_up.copy_(as_dev(r_up).reshape(self.r_up.shape))
        self.r_dn.copy_(as_dev(r_dn).reshape(self.r_dn.shape))
        return self.refresh()

    def refresh(self) -> "WalkerState":
        """Rebuild every cache and G^-1 from the current positions (LU refresh)."""
        ws, nb = self._workspace("_ws_init", "jqmc_walkers_init_workspace_bytes")
        st = self.struct()
        _lib.call("jqmc_walkers_init", self.pack.handle, self.W, ctypes.byref(st), _ptr(ws), nb, stream_handle())
        return self

    def refresh_inverse(self) -> "WalkerState":
        """G from the cached orbitals and a fresh FP64 inverse (jqmc_walkers_refresh_inverse)."""
        st = self.struct()
        _lib.call("jqmc_walkers_refresh_inverse", self.pack.handle, self.W, ctypes.byref(st), stream_handle())
        return self

    def refresh_orbitals(self) -> "WalkerState":
        """AO values at the electrons in FP64, orbital and pair rows, G and a fresh FP64 inverse from
        the current positions (jqmc_walkers_refresh_orbitals); J3 fields and J are kept."""
        st = self.struct()
        _lib.call("jqmc_walkers_refresh_orbitals", self.pack.handle, self.W, ctypes.byref(st), stream_handle())
        return self

    def point_ratios(self, elec: torch.Tensor, points: torch.Tensor):
        """(det ratio, Jastrow ratio) per walker and configuration: configuration p of walker w moves
        electron elec[w, p] (up block first) to points[w, p] (jqmc_point_ratios).  Shapes (W, P)."""
        el = elec.to(device=device(), dtype=torch.int32).contiguous()
        pts = as_dev(points).reshape(self.W, -1, 3).contiguous()
        P = pts.shape[1]
        if el.shape != (self.W, P):
            raise ValueError(f"elec must be (W, P) = ({self.W}, {P}), got {tuple(el.shape)}")
        det = torch.empty((self.W, P), dtype=F64, device=device())
        jas = torch.empty_like(det)
        st = self.struct()
        _lib.call("jqmc_point_ratios", self.pack.handle, self.W, ctypes.byref(st), P, _ptr(el), _ptr(pts), _ptr(det),
                  _ptr(jas), stream_handle())
        return det, jas

    def sweep(self, n_moves: int, Dt: float, epsilon_AS: float, seed: int, move0: int, batched: bool = True) -> None:
        """``n_moves`` Metropolis proposals per walker: move-synchronous over all walkers
        (jqmc_vmc_sweep_batched, default) or every walker's moves inside one launch (jqmc_vmc_sweep);
        both take the same decisions."""
        st = self.struct()
        if batched:
            ws, nb = self._workspace("_ws_sw", "jqmc_vmc_sweep_batched_workspace_bytes")
            _lib.call("jqmc_vmc_sweep_batched", self.pack.handle, self.W, ctypes.byref(st), int(n_moves), float(Dt),
                      float(epsilon_AS), int(seed) & 0xFFFFFFFFFFFFFFFF, int(move0), None, _ptr(ws), nb,
                      stream_handle())
        else:
            _lib.call("jqmc_vmc_sweep", self.pack.handle, self.W, ctypes.byref(st), int(n_moves), float(Dt),
                      float(epsilon_AS), int(seed) & 0xFFFFFFFFFFFFFFFF, int(move0), stream_handle())

    def rotations(self, seed: int, step: int, stream_id: int = 1) -> torch.Tensor:
        RT = torch.empty((self.W, 3, 3), dtype=F64, device=device())
        _lib.call("jqmc_rotation_matrices", self.W, int(seed) & 0xFFFFFFFFFFFFFFFF, int(step), int(stream_id),
                  _ptr(RT), stream_handle())
        return RT

    def local_energy(self, RT=None, with_parts: bool = False):
        """e_L per walker with the current G^-1 (jqmc_local_energy)."""
        ws, nb = self._workspace("_ws_el", "jqmc_local_energy_workspace_bytes")
        e_L = torch.empty(self.W, dtype=F64, device=device())
        kin = torch.empty_like(e_L) if with_parts else None
        pot = torch.empty_like(e_L) if with_parts else None
        tel = torch.empty((self.W, self.pack.n_el), dtype=F64, device=device()) if with_parts else None
        RTd = as_dev(RT).reshape(self.W, 3, 3) if RT is not None else None
        st = self.struct()
        _lib.call("jqmc_local_energy", self.pack.handle, self.W, ctypes.byref(st), _ptr(RTd), _ptr(e_L), _ptr(kin),
                  _ptr(pot), _ptr(tel), _ptr(ws), nb, stream_handle())
        if with_parts:
            return e_L, kin, pot, tel
        return e_L

    def position_grads(self, RT=None, swct: bool = True) -> dict:
        """Position derivatives per walker (jqmc_position_grads): {'e_L' (W,), 'dE_dr' (W, N, 3),
        'dE_dR' (W, n_nuc, 3), 'dlnpsi_dr' (W, N, 3), 'dlnpsi_dR' (W, n_nuc, 3), 'force_HF' and
        'force_PP' (W, n_nuc, 3)}, the per-walker quantities of jqmc_mcmc.py:744-852.  Electrons
        are ordered up then dn."""
        ws, nb = self._workspace("_ws_fr", "jqmc_position_grads_workspace_bytes")
        pk = self.pack
        n_nuc = int(pk.desc.n_nuc)
        dev = device()
        out = {"e_L": torch.empty(self.W, dtype=F64, device=dev)}
        for k in ("dE_dr", "dlnpsi_dr"):
            out[k] = torch.empty((self.W, pk.n_el, 3), dtype=F64, device=dev)
        for k in ("dE_dR", "dlnpsi_dR", "force_HF", "force_PP"):
            out[k] = torch.empty((self.W, n_nuc, 3), dtype=F64, device=dev)
        RTd = as_dev(RT).reshape(self.W, 3, 3) if RT is not None else None
        st = self.struct()
        _lib.call("jqmc_position_grads", pk.handle, self.W, ctypes.byref(st), _ptr(RTd), int(bool(swct)),
                  _ptr(out["e_L"]), _ptr(out["dE_dr"]), _ptr(out["dE_dR"]), _ptr(out["dlnpsi_dr"]),
                  _ptr(out["dlnpsi_dR"]), _ptr(out["force_HF"]), _ptr(out["force_PP"]), _ptr(ws), nb,
                  stream_handle())
        return out

    # ------------------------------------------------------------- LRDMC

    @staticmethod
    def lrdmc_params(alat: float, E_scf: float = 0.0, n_steps: int = 1, non_local_move: str = "tmove",
                     random_mesh: bool = True, tau: float = 0.0, mode: str = "fixed_steps") -> _lib.LrdmcParams:
        if non_local_move not in ("tmove", "dltmove"):
            raise NotImplementedError(f"non_local_move = {non_local_move} is not implemented")
        p = _lib.LrdmcParams()
        p.mode = {"fixed_steps": _lib.LRDMC_FIXED_STEPS, "fixed_tau": _lib.LRDMC_FIXED_TAU}[mode]
        p.n_steps = int(n_steps)
        p.alat = float(alat)
        p.E_scf = float(E_scf)
        p.tau = float(tau)
        p.non_local_move = _lib.TMOVE if non_local_move == "tmove" else _lib.DLTMOVE
        p.random_mesh = int(bool(random_mesh))
        return p

    def lrdmc_projection(self, params: _lib.LrdmcParams, seed: int, step0: int, w_L: torch.Tensor | None = None):
        """Projections per walker (jqmc_lrdmc_projection).  Returns dict of device tensors
        w_L, RT (last rotation), V_diag, V_nondiag (last projection), tau_left, n_proj."""
        ws, nb = self._workspace("_ws_lr", "jqmc_lrdmc_workspace_bytes")
        dev = device()
        w = torch.ones(self.W, dtype=F64, device=dev) if w_L is None else w_L
        out = {
            "w_L": w,
            "RT": torch.empty((self.W, 3, 3), dtype=F64, device=dev),
            "V_diag": torch.empty(self.W, dtype=F64, device=dev),
            "V_nondiag": torch.empty(self.W, dtype=F64, device=dev),
            "tau_left": torch.empty(self.W, dtype=F64, device=dev),
            "n_proj": torch.empty(self.W, dtype=torch.int64, device=dev),
        }
        st = self.struct()
        _lib.call("jqmc_lrdmc_projection", self.pack.handle, self.W, ctypes.byref(st), ctypes.byref(params),
                  int(seed) & 0xFFFFFFFFFFFFFFFF, int(step0), _ptr(w), _ptr(out["RT"]), _ptr(out["V_diag"]),
                  _ptr(out["V_nondiag"]), _ptr(out["tau_left"]), _ptr(out["n_proj"]), _ptr(ws), nb, stream_handle())
        return out

    def lrdmc_local_energy(self, params: _lib.LrdmcParams, RT=None):
        """(V_diag, V_nondiag) per walker (jqmc_lrdmc_local_energy); e_L = V_diag + V_nondiag."""
        ws, nb = self._workspace("_ws_lr", "jqmc_lrdmc_workspace_bytes")
        Vd = torch.empty(self.W, dtype=F64, device=device())
        Vn = torch.empty_like(Vd)
        RTd = as_dev(RT).reshape(self.W, 3, 3) if RT is not None else None
        st = self.struct()
        _lib.call("jqmc_lrdmc_local_energy", self.pack.handle, self.W, ctypes.byref(st), ctypes.byref(params),
                  _ptr(RTd), _ptr(Vd), _ptr(Vn), _ptr(ws), nb, stream_handle())
        return Vd, Vn

    # ------------------------------------------------------------- parameter derivatives

    _BLOCK_BITS = {"j1_param": _lib.PARAM_J1, "j2_param": _lib.PARAM_J2, "j3_matrix": _lib.PARAM_J3,
                   "lambda_matrix": _lib.PARAM_LAMBDA}

    def dln_psi_dparams(self, blocks=("j1_param", "j2_param", "j3_matrix", "lambda_matrix")):
        """O_k = d ln|Psi| / d p_k per walker (jqmc_dln_psi_dparams) for the requested blocks the
        system has, shaped like the reference's parameters: {'j1_param': (W,), 'j2_param': (W,),
        'j3_matrix': (W, n_j3, n_j3 + 1), 'lambda_matrix': (W, n_orb, n_orb + n_up - n_dn)}."""
        bits = 0
        for b in blocks:
            bits |= self._BLOCK_BITS[b]
        offs = (ctypes.c_int64 * 4)()
        P = ctypes.c_int64()
        _lib.call("jqmc_param_layout", self.pack.handle, bits, offs, ctypes.byref(P))
        ws, nb = self._workspace("_ws_dp", "jqmc_dln_psi_dparams_workspace_bytes")
        O = torch.empty((self.W, max(P.value, 1)), dtype=F64, device=device())
        st = self.struct()
        _lib.call("jqmc_dln_psi_dparams", self.pack.handle, self.W, ctypes.byref(st), bits, _ptr(O), O.shape[1],
                  _ptr(ws), nb, stream_handle())
        pk = self.pack
        na, no, unp = pk.n_ao, pk.n_orb, pk.n_up - pk.n_dn
        out = {}
        if offs[0] >= 0:
            out["j1_param"] = O[:, offs[0]]
        if offs[1] >= 0:
            out["j2_param"] = O[:, offs[1]]
        if offs[2] >= 0:
            J = O[:, offs[2]:offs[2] + na * (na + 1)].reshape(self.W, na, na + 1)
            if pk.jas_separate:  # the J3 basis follows the geminal basis in the device AO set
                n0 = pk.n_det_ao
                J = torch.cat([J[:, n0:, n0:na], J[:, n0:, na:]], dim=2)
            out["j3_matrix"] = J
        if offs[3] >= 0:
            out["lambda_matrix"] = O[:, offs[3]:offs[3] + no * (no + unp)].reshape(self.W, no, no + unp)
        return out

    def reconfigure(self, chosen: torch.Tensor) -> "WalkerState":
        """Walker j takes the state of walker chosen[j] (positions and every cache copied), then G^-1
        is rebuilt from the positions (jqmc_walkers_reconfigure)."""
        ws, nb = self._workspace("_ws_rc", "jqmc_walkers_reconfigure_workspace_bytes")
        ch = chosen.to(device=device(), dtype=torch.int64).contiguous()
        st = self.struct()
        _lib.call("jqmc_walkers_reconfigure", self.pack.handle, self.W, ctypes.byref(st), _ptr(ch), _ptr(ws), nb,
                  stream_handle())
        return self


class CapturedStep:
    """One VMC measurement step (n_moves proposals per walker, the quadrature rotation, e_L) recorded
    once in a HIP graph and replayed: the per-step structure of MCMC.run (jqmc_mcmc.py:664-745)
    without per-launch host overhead.  The Philox counters live on the device (ctr[0] = move,
    ctr[1] = step) and advance inside the graph, so replay k draws what the uncaptured calls
    sweep(n_moves, ..., move0 + k n_moves) and rotations(seed, step0 + k) draw.  The graph reads
    and writes the walker state in place; ``e_L`` and ``RT`` are overwritten on every replay."""

    def __init__(self, ws: WalkerState, n_moves: int, Dt: float, epsilon_AS: float, seed: int, move0: int = 0,
                 step0: int = 0):
        self.ws = ws
        self.n_moves = int(n_moves)
        dev = device()
        self.ctr = torch.tensor([int(move0), int(step0)], dtype=torch.int64, device=dev)
        self.RT = torch.empty((ws.W, 3, 3), dtype=F64, device=dev)
        self.e_L = torch.empty(ws.W, dtype=F64, device=dev)
        self._ws, self._nb = ws._workspace("_ws_el", "jqmc_local_energy_workspace_bytes")
        self._sw, self._sw_nb = ws._workspace("_ws_sw", "jqmc_vmc_sweep_batched_workspace_bytes")
        self._args = (float(Dt), float(epsilon_AS), int(seed) & 0xFFFFFFFFFFFFFFFF)
        self._st = ws.struct()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._record()

    def _record(self) -> None:
        ws, (Dt, eps, seed), h = self.ws, self._args, stream_handle()
        c0 = self.ctr.data_ptr()
        _lib.call("jqmc_vmc_sweep_batched", ws.pack.handle, ws.W, ctypes.byref(self._st), self.n_moves, Dt, eps, seed,
                  0, c0, _ptr(self._sw), self._sw_nb, h)
        _lib.call("jqmc_rotation_matrices_counted", ws.W, seed, 0, c0 + 8, 1, _ptr(self.RT), h)
        _lib.call("jqmc_local_energy", ws.pack.handle, ws.W, ctypes.byref(self._st), _ptr(self.RT), _ptr(self.e_L),
                  None, None, None, _ptr(self._ws), self._nb, h)
        _lib.call("jqmc_counter_add", c0, self.n_moves, h)
        _lib.call("jqmc_counter_add", c0 + 8, 1, h)

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.e_L


def lrdmc_projection_multi(states, params, seed: int, step0s, w_Ls=None):
    """Projections of several walker sets of different systems in one launch
    (jqmc_lrdmc_projection_multi).  Returns one output dict per system as
    ``WalkerState.lrdmc_projection``."""
    n = len(states)
    dev = device()
    outs = []
    for k, ws in enumerate(states):
        w = torch.ones(ws.W, dtype=F64, device=dev) if w_Ls is None else w_Ls[k]
        outs.append({
            "w_L": w,
            "RT": torch.empty((ws.W, 3, 3), dtype=F64, device=dev),
            "V_diag": torch.empty(ws.W, dtype=F64, device=dev),
            "V_nondiag": torch.empty(ws.W, dtype=F64, device=dev),
            "tau_left": torch.empty(ws.W, dtype=F64, device=dev),
            "n_proj": torch.empty(ws.W, dtype=torch.int64, device=dev),
        })
    VP = ctypes.c_void_p * n
    handles = VP(*[ws.pack.handle.value for ws in states])
    Ws = (ctypes.c_int64 * n)(*[ws.W for ws in states])
    wks = (_lib.Walkers * n)(*[ws.struct() for ws in states])
    prm = (_lib.LrdmcParams * n)(*params)
    st0 = (ctypes.c_int64 * n)(*[int(x) for x in step0s])
    arr = lambda key: VP(*[o[key].data_ptr() for o in outs])
    wsp = [ws._workspace("_ws_lr", "jqmc_lrdmc_workspace_bytes") for ws in states]
    wsa = VP(*[(w.data_ptr() if w is not None else None) for w, _ in wsp])
    wsb = (ctypes.c_int64 * n)(*[nb for _, nb in wsp])
    _lib.call("jqmc_lrdmc_projection_multi", n, handles, Ws, wks, prm, int(seed) & 0xFFFFFFFFFFFFFFFF, st0,
              arr("w_L"), arr("RT"), arr("V_diag"), arr("V_nondiag"), arr("tau_left"), arr("n_proj"), wsa, wsb,
              stream_handle())
    return outs


def lrdmc_branching(w_L: torch.Tensor, zeta: float, V_diag=None, e_L=None, E_scf: float = 0.0):
    """Comb reconfiguration indices and weighted sums (jqmc_lrdmc_branching).
    Returns (chosen (W,) int64, sums (4,) float64 or None)."""
    W = int(w_L.numel())
    dev = device()
    chosen = torch.empty(W, dtype=torch.int64, device=dev)
    sums = torch.empty(4, dtype=F64, device=dev) if V_diag is not None else None
    ws = torch.empty(max(W, 1), dtype=F64, device=dev)
    _lib.call("jqmc_lrdmc_branching", W, _ptr(w_L.contiguous()), _ptr(V_diag), _ptr(e_L), float(E_scf), float(zeta),
              _ptr(chosen), _ptr(sums), _ptr(ws), ws.numel() * 8, stream_handle())
    return chosen, sums


# ----------------------------------------------------------------- component systems


def _basis_hamiltonian(orb) -> Hamiltonian_data:
    """A Hamiltonian that carries only an orbital basis (one up electron, AO/MO orbitals)."""
    aos = orb.aos_data if isinstance(orb, MOs_data) else orb
    st = aos.structure_data
    n_orb = orb._num_orb
    gem = Geminal_data(1, 0, orb, orb, np.eye(n_orb, n_orb + 1))
    cp = Coulomb_potential_data(structure_data=st, ecp_flag=False)
    return Hamiltonian_data(st, cp, Wavefunction_data(Jastrow_data(), gem))


_BASIS_CACHE: dict = {}


def _basis_pack(orb) -> SystemPack:
    key = id(orb)
    hit = _BASIS_CACHE.get(key)
    if hit is not None and hit[0] is orb:
        return hit[1]
    pack = SystemPack(_basis_hamiltonian(orb))
    _BASIS_CACHE[key] = (orb, pack)
    return pack


def aos_value_grad_lap(aos, r_carts, value_only: bool):
    pack = _basis_pack(aos)
    r = as_dev(r_carts)
    lead = r.shape[:-2]
    n_e = r.shape[-2]
    pts = r.reshape(-1, 3).contiguous()
    P = pts.shape[0]
    outs = [torch.empty((aos.num_ao, P), dtype=F64, device=r.device) for _ in range(1 if value_only else 5)]
    args = [_ptr(o) for o in outs] + [None] * (5 - len(outs))
    _lib.call("jqmc_aos_value_grad_lap", pack.handle, P, _ptr(pts), *args, stream_handle())
    res = [o.reshape(aos.num_ao, *lead, n_e).movedim(0, -2) for o in outs]
    return res


def mos_value_grad_lap(mos: MOs_data, r_carts, value_only: bool):
    pack = _basis_pack(mos)
    r = as_dev(r_carts)
    lead = r.shape[:-2]
    n_e = r.shape[-2]
    pts = r.reshape(-1, 3).contiguous()
    P = pts.shape[0]
    outs = [torch.empty((mos.num_mo, P), dtype=F64, device=r.device) for _ in range(1 if value_only else 5)]
    ws = torch.empty(5 * mos.aos_data.num_ao * max(P, 1), dtype=F64, device=r.device)
    args = [_ptr(o) for o in outs] + [None] * (5 - len(outs))
    _lib.call("jqmc_orbs_value_grad_lap", pack.handle, 0, P, _ptr(pts), *args, _ptr(ws), stream_handle())
    return [o.reshape(mos.num_mo, *lead, n_e).movedim(0, -2) for o in outs]


def _wavefunction_hamiltonian(wf: Wavefunction_data, cp: Coulomb_potential_data | None = None) -> Hamiltonian_data:
    st = wf.geminal_data.aos_data.structure_data
    if cp is None:
        cp = Coulomb_potential_data(structure_data=st, ecp_flag=False)
    return Hamiltonian_data(st, cp, wf)


_WF_CACHE: dict = {}


def _wf_pack(wf: Wavefunction_data, cp=None, Nv: int = 6, NN: int = 1) -> SystemPack:
    key = (id(wf), id(cp), Nv, NN)
    hit = _WF_CACHE.get(key)
    if hit is not None and hit[0] is wf and hit[1] is cp:
        return hit[2]
    pack = SystemPack(_wavefunction_hamiltonian(wf, cp), Nv=Nv, NN=NN)
    _WF_CACHE[key] = (wf, cp, pack)
    return pack


def _state(pack: SystemPack, r_up, r_dn):
    ru, rd, batched = _batched(r_up, r_dn)
    if ru.shape[1] != pack.n_up or rd.shape[1] != pack.n_dn:
        raise ValueError(
            f"Number of up and dn electrons ({ru.shape[1]}, {rd.shape[1]}) are not consistent with "
            f"({pack.n_up}, {pack.n_dn})"
        )
    ws = WalkerState(pack, ru.shape[0]).init(ru, rd)
    return ws, batched


def _unbatch(t, batched):
    return t if batched else t[0]


def geminal_matrix(gem: Geminal_data, r_up, r_dn, with_inverse: bool = False):
    wf = Wavefunction_data(Jastrow_data(), gem)
    key = ("gem", id(gem))
    hit = _WF_CACHE.get(key)
    if hit is None or hit[0] is not gem:
        hit = (gem, None, SystemPack(_wavefunction_hamiltonian(wf)))
        _WF_CACHE[key] = hit
    ws, b = _state(hit[2], r_up, r_dn)
    return (_unbatch(ws.G, b), _unbatch(ws.sign, b), _unbatch(ws.logabsdet, b), _unbatch(ws.as_factor, b))


def jastrow_value(jas: Jastrow_data, r_up, r_dn, geminal_data: Geminal_data | None = None):
    if geminal_data is None:
        j3 = jas.jastrow_three_body_data
        if j3 is None:
            raise ValueError("pass geminal_data when the Jastrow has no three-body part")
        n_up = np.asarray(r_up).shape[-2] if not isinstance(r_up, torch.Tensor) else r_up.shape[-2]
        n_dn = np.asarray(r_dn).shape[-2] if not isinstance(r_dn, torch.Tensor) else r_dn.shape[-2]
        aos = j3.orb_data
        n = aos.num_ao
        lam = np.hstack([np.eye(n), np.eye(n, n_up - n_dn, k=-n_dn)])
        geminal_data = Geminal_data(n_up, n_dn, aos, aos, lam)
    wf = Wavefunction_data(jas, geminal_data)
    key = ("jas", id(jas), id(geminal_data))
    hit = _WF_CACHE.get(key)
    if hit is None or hit[0] is not jas:
        hit = (jas, geminal_data, SystemPack(_wavefunction_hamiltonian(wf)))
        _WF_CACHE[key] = hit
    ws, b = _state(hit[2], r_up, r_dn)
    return _unbatch(ws.jastrow, b)


def ln_wavefunction(wf: Wavefunction_data, r_up, r_dn):
    ws, b = _state(_wf_pack(wf), r_up, r_dn)
    return _unbatch(ws.jastrow + ws.logabsdet, b)


def kinetic_all_elements(wf: Wavefunction_data, r_up, r_dn, geminal_inverse):
    pack = _wf_pack(wf)
    ws, b = _state(pack, r_up, r_dn)
    if geminal_inverse is not None:
        ws.Ginv.copy_(as_dev(geminal_inverse).reshape(ws.Ginv.shape))
    _, _, _, tel = ws.local_energy(None, with_parts=True)
    t_up, t_dn = tel[:, : pack.n_up], tel[:, pack.n_up :]
    return _unbatch(t_up, b), _unbatch(t_dn, b)


def local_energy(ham: Hamiltonian_data, r_up, r_dn, RT, geminal_inverse, NN: int, Nv: int):
    pack = system_for(ham, Nv=Nv, NN=NN)
    ws, b = _state(pack, r_up, r_dn)
    if geminal_inverse is not None:
        ws.Ginv.copy_(as_dev(geminal_inverse).reshape(ws.Ginv.shape))
    RTd = None
    if RT is not None:
        RTd = as_dev(RT)
        if RTd.dim() == 2:
            RTd = RTd.expand(ws.W, 3, 3).contiguous()
    return _unbatch(ws.local_energy(RTd), b)


def coulomb_potential(cp, wf, r_up, r_dn, RT, NN: int, Nv: int):
    if wf is None:
        raise ValueError("wavefunction_data is required")
    st = cp.structure_data
    ham = Hamiltonian_data(st, cp, wf)
    pack = SystemPack(ham, Nv=Nv, NN=NN)
    ws, b = _state(pack, r_up, r_dn)
    RTd = None
    if RT is not None:
        RTd = as_dev(RT)
        if RTd.dim() == 2:
            RTd = RTd.expand(ws.W, 3, 3).contiguous()
    _, _, pot, _ = ws.local_energy(RTd, with_parts=True)
    return _unbatch(pot, b)


# ----------------------------------------------------------------- single-electron ratio entry points


def _moved_electrons(old_up, old_dn, new_up, new_dn):
    """Electron index and new position of each one-electron-moved configuration: the first electron
    whose coordinates differ (argmax of the change mask, as determinant.py:1570-1600), up block first;
    a configuration equal to the old one reports electron 0 at its own position (ratio 1)."""
    up_mask = (new_up != old_up[:, None]).any(dim=-1)  # (W, G, N_up)
    dn_mask = (new_dn != old_dn[:, None]).any(dim=-1)
    n_up = old_up.shape[1]
    up_moved = up_mask.any(dim=-1)
    i_up = up_mask.to(torch.int64).argmax(dim=-1)
    i_dn = dn_mask.to(torch.int64).argmax(dim=-1)
    elec = torch.where(up_moved | ~dn_mask.any(dim=-1), i_up, n_up + i_dn)
    pos_up = torch.gather(new_up, 2, i_up[..., None, None].expand(-1, -1, 1, 3))[:, :, 0]
    if new_dn.shape[2] > 0:
        pos_dn = torch.gather(new_dn, 2, i_dn[..., None, None].expand(-1, -1, 1, 3))[:, :, 0]
    else:
        pos_dn = pos_up
    pts = torch.where((elec < n_up)[..., None], pos_up, pos_dn)
    return elec, pts


def _placeholder_basis(jas: Jastrow_data, n_up: int) -> AOs_cart_data:
    """n_up s-type AOs of distinct exponents on the first nucleus of the one-body Jastrow's structure
    (a single nucleus at the origin without one): the determinant a Jastrow-only system needs to be
    built; the Jastrow ratios do not depend on it."""
    from .structure import Structure_data

    j1 = jas.jastrow_one_body_data
    st = j1.structure_data if j1 is not None else Structure_data(
        positions=np.zeros((1, 3)), atomic_numbers=(1,), element_symbols=("H",), atomic_labels=("H",))
    n = max(int(n_up), 1)
    return AOs_cart_data(structure_data=st, nucleus_index=(0,) * n, num_ao=n, num_ao_prim=n,
                         orbital_indices=tuple(range(n)), exponents=np.arange(1, n + 1, dtype=np.float64),
                         coefficients=np.ones(n), angular_momentums=(0,) * n, polynominal_order_x=(0,) * n,
                         polynominal_order_y=(0,) * n, polynominal_order_z=(0,) * n)


def _ratio_state(pack, old_up, old_dn, A_old_inv):
    ws, batched = _state(pack, old_up, old_dn)
    if A_old_inv is not None:
        ws.Ginv.copy_(as_dev(A_old_inv).reshape(ws.Ginv.shape))
    return ws, batched


def _as_configs(old_up, old_dn, new_up, new_dn):
    ou, od, batched = _batched(old_up, old_dn)
    nu, nd = as_dev(new_up), as_dev(new_dn)
    if not batched:
        nu, nd = nu.reshape(1, -1, ou.shape[1], 3), nd.reshape(1, -1, od.shape[1], 3)
    return ou, od, nu.contiguous(), nd.contiguous(), batched


def determinant_ratios(gem: Geminal_data, A_old_inv, old_up, old_dn, new_up, new_dn):
    ou, od, nu, nd, batched = _as_configs(old_up, old_dn, new_up, new_dn)
    wf = Wavefunction_data(Jastrow_data(), gem)
    key = ("gem", id(gem))
    hit = _WF_CACHE.get(key)
    if hit is None or hit[0] is not gem:
        hit = (gem, None, SystemPack(_wavefunction_hamiltonian(wf)))
        _WF_CACHE[key] = hit
    ws, _ = _ratio_state(hit[2], ou, od, A_old_inv)
    elec, pts = _moved_electrons(ou, od, nu, nd)
    det, _ = ws.point_ratios(elec, pts)
    return _unbatch(det, batched)


def jastrow_ratios(jas: Jastrow_data, old_up, old_dn, new_up, new_dn):
    ou, od, nu, nd, batched = _as_configs(old_up, old_dn, new_up, new_dn)
    n_up, n_dn = ou.shape[1], od.shape[1]
    j3 = jas.jastrow_three_body_data
    aos = j3.orb_data if j3 is not None else _placeholder_basis(jas, n_up)
    n = aos.num_ao
    lam = np.hstack([np.eye(n), np.eye(n, n_up - n_dn, k=-n_dn)])
    gem = Geminal_data(n_up, n_dn, aos, aos, lam)
    key = ("jas", id(jas), n_up, n_dn)
    hit = _WF_CACHE.get(key)
    if hit is None or hit[0] is not jas:
        hit = (jas, gem, SystemPack(_wavefunction_hamiltonian(Wavefunction_data(jas, gem))))
        _WF_CACHE[key] = hit
    ws, _ = _ratio_state(hit[2], ou, od, None)
    elec, pts = _moved_electrons(ou, od, nu, nd)
    _, jr = ws.point_ratios(elec, pts)
    return _unbatch(jr, batched)


def split_spin_configs(old_up, old_dn, new_up_shifted, new_dn_shifted):
    """Full one-electron-moved configurations from the block-structured arrays of the _split_spin
    ratio functions: the up block moves up electrons with the dn electrons in place, the dn block the
    reverse; returned in the order up block, dn block (the concatenation the reference returns)."""
    ou, od, batched = _batched(old_up, old_dn)
    su, sd = as_dev(new_up_shifted), as_dev(new_dn_shifted)
    if not batched:
        su, sd = su.reshape(1, -1, ou.shape[1], 3), sd.reshape(1, -1, od.shape[1], 3)
    Wn, Gu, Gd = ou.shape[0], su.shape[1], sd.shape[1]
    new_up = torch.cat([su, ou[:, None].expand(Wn, Gd, -1, 3)], dim=1)
    new_dn = torch.cat([od[:, None].expand(Wn, Gu, -1, 3), sd], dim=1)
    if not batched:
        return ou[0], od[0], new_up[0], new_dn[0]
    return ou, od, new_up, new_dn
