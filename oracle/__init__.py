"""FP64 CPU oracle for the jQMC per-walker VMC/LRDMC hot path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy restatement of the reference
algorithm (kousuke-nakano/jQMC, JAX) used as the parity checker for the gfx950 HIP
path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only to check or time the reference algorithm; the
``jqmc_amd`` product never imports it and fails loudly when its HIP library is absent.

Every function cites the reference file:line it follows.  How each function is pinned:

* ``rng``: Philox4x32-10 known-answer vectors (Salmon et al., SC'11 / Random123).
* ``aos``: analytic overlap (normalised primitives have unit self-overlap), central finite
  differences for grad/Laplacian, the reference's debug loop formulation, TurboRVB golden
  values through the wavefunction ratio and kinetic energy (H2 all-electron, water ECP).
* ``wf``: brute-force determinants/inverses (``numpy.linalg``), ``det G = det Phi_up det Phi_dn``
  for identity pairing, Sherman-Morrison ratio vs full determinant ratio, finite-difference
  gradients/Laplacians of ``ln Psi``, hydrogen 1s eigenstate (``e_L = -1/2`` everywhere),
  TurboRVB golden values.
* ``coulomb``: closed-form pair sums, quadrature exactness of the meshes for ``P_l``,
  rotation invariance for an s-only projector, TurboRVB golden potential values.
* ``mcmc``: detailed balance of the proposal kernel, decision parity between the
  fast-update loop and a brute-force loop that recomputes full determinants each move
  (the reference's ``_MCMC_debug`` strategy, jqmc_mcmc.py:4968-5115).
* ``gfmc``: brute-force ratios per candidate move and the comb reconfiguration identity.

The JAX PRNG bit stream itself cannot be reproduced without JAX; decision parity is
defined on the shared Philox stream (parity unpinned against the JAX key stream).
"""
