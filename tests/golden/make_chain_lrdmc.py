"""Writes tests/golden/chain_lrdmc.npz: FP64 oracle LRDMC values of the uneven chain/ring systems with
56 and 96 valence electrons (jqmc_amd.synthetic.chain_ecp, seed 1), one walker each, for the GPU
parity tests of the batched-LRDMC configuration.  The oracle takes about a minute per 96-electron
projection, too long to run inside a GPU test, so its results are stored here.

Per system: the start positions (initial_walkers(1, seed=5)), the lattice-Hamiltonian elements
V_diag / V_nondiag at the rotation of (seed 9, walker 0, step 2) with a = 0.2 and the t-move
(oracle.gfmc.hamiltonian_elements), and one GFMC_n projection (oracle.gfmc.projection_step, Philox
seed 7, walker 0, step 0, E_scf = -50, a = 0.2, t-move): weight, move index and new positions.

Run from the repository root: python tests/golden/make_chain_lrdmc.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from jqmc_amd import synthetic  # noqa: E402
from oracle import gfmc as og  # noqa: E402
from oracle import rng as orng  # noqa: E402
from oracle import wf as owf  # noqa: E402


def main():
    out = {}
    RT = orng.rotation_matrix_T(*orng.rotation_angles(9, 0, 2))
    for n in (56, 96):
        s = synthetic.chain_ecp(n, seed=1)
        ham = s.hamiltonian_data
        ru, rd = s.initial_walkers(1, seed=5)
        gem = ham.wavefunction_data.geminal_data
        G = owf.geminal(gem, ru[0], rd[0])
        Gi = np.linalg.inv(G)
        d, nd, _, _ = og.hamiltonian_elements(ham, ru[0], rd[0], Gi, RT, 0.2, "tmove", Nv=6, NN=1)
        p = og.projection_step(ham, ru[0], rd[0], Gi, G, 1.0, -50.0, 0.2, 7, 0, 0, "tmove")
        out[f"n{n}_r_up"], out[f"n{n}_r_dn"] = ru[0], rd[0]
        out[f"n{n}_V_diag"], out[f"n{n}_V_nondiag"] = d, nd
        out[f"n{n}_proj_w"], out[f"n{n}_proj_k"] = p["w"], p["k"]
        out[f"n{n}_proj_r_up"], out[f"n{n}_proj_r_dn"] = p["r_up"], p["r_dn"]
        out[f"n{n}_proj_diag"], out[f"n{n}_proj_nondiag"] = p["diag"], p["nondiag"]
        print(n, d, nd, p["w"], p["k"], flush=True)
    out["RT"] = RT
    np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "chain_lrdmc.npz"), **out)


if __name__ == "__main__":
    main()
