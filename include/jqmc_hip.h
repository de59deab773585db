/*
 * jqmc_hip.h -- C ABI of the MI355X (gfx950) per-walker VMC/LRDMC hot path.
 *
 * Plain C: no C++ or torch types cross this boundary.  Device arrays are raw
 * pointers allocated and owned by the caller; every call is ordered on the
 * caller's hipStream_t (passed as void*), allocates no device memory and does
 * not synchronise the host.  Return value: JQMC_OK (0) or a jqmc_status code;
 * jqmc_last_error() returns a message for the calling thread's last failure.
 *
 * The reference (kousuke-nakano/jQMC) is pure JAX and exposes these operations
 * as Python functions over flax dataclasses; each entry point below names the
 * function(s) it replaces.  Per-walker shapes follow the reference with a
 * leading walker axis W (the reference's vmap axis, jqmc_mcmc.py:4728-4748).
 */
#ifndef JQMC_HIP_H
#define JQMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JQMC_ABI_VERSION 6

typedef enum {
  JQMC_OK = 0,
  JQMC_ERR_INVALID_ARGUMENT = 1,
  JQMC_ERR_SHAPE = 2,
  JQMC_ERR_UNSUPPORTED = 3,
  JQMC_ERR_HIP = 4,
  JQMC_ERR_OUT_OF_MEMORY = 5,
  JQMC_ERR_WORKSPACE = 6
} jqmc_status;

/* Jastrow functional forms (jastrow_factor.py:560-645, 1110-1177). */
enum { JQMC_JASTROW_OFF = 0, JQMC_JASTROW_EXP = 1, JQMC_JASTROW_PADE = 2 };

/*
 * Arithmetic precision of a system's AO / J3 work, following the precision zones of the
 * reference's mixed mode (jqmc/_precision.py:383-411).  PLAIN_FP64: FP64 everywhere.
 * MIXED_FP32: AO values at electrons, proposed moves, lattice and quadrature points (ao_eval) and
 * the Jastrow work (J3 fields and their products, jastrow_eval / jastrow_ratio) are stored and
 * computed in FP32; the AO->orbital and Lambda contractions, the orbital and pair caches, G, G^-1
 * and its Sherman-Morrison updates, every determinant ratio and field (mo_eval, det_eval,
 * det_ratio, det_grad_lap), the AO gradients and Laplacians at the electrons (ao_grad_lap),
 * ln|det G|, the LU refresh and every energy and weight sum stay FP64.
 */
enum { JQMC_PRECISION_PLAIN_FP64 = 0, JQMC_PRECISION_MIXED_FP32 = 1 };

/*
 * Flattened Hamiltonian_data (hamiltonians.py:82-177) on the host; copied to the
 * device by jqmc_system_create.  Basis: Cartesian contracted GTOs
 * (AOs_cart_data, atomic_orbital.py:87-258) with the primitive normalisation
 * N_p folded into prim_coef (atomic_orbital.py:2237-2251).
 */
typedef struct {
  int32_t n_nuc;
  const double* nuc_pos;        /* (n_nuc, 3) bohr                                        */
  const double* nuc_charge;     /* (n_nuc) Z, or Z - z_core with ECP (coulomb_potential.py:299-309) */

  int32_t n_ao;
  int32_t n_prim;
  const int32_t* ao_nuc;        /* (n_ao) nucleus of each AO                               */
  const int32_t* ao_pow;        /* (n_ao, 3) nx, ny, nz of Cartesian AOs (used when ao_term_ptr is NULL) */
  const int32_t* ao_prim_ptr;   /* (n_ao + 1) CSR offsets of each AO's primitives          */
  const double* prim_exp;       /* (n_prim) Gaussian exponents                             */
  const double* prim_coef;      /* (n_prim) normalisation * contraction coefficient        */
  /* Optional general angular parts: AO a is radial(r) * sum_t term_coef[t] x^nx y^ny z^nz over
     t in [ao_term_ptr[a], ao_term_ptr[a+1]); spherical AOs (AOs_sphe_data, atomic_orbital.py:
     780-1347) are passed this way, their primitive normalisation folded into prim_coef. */
  int32_t n_terms;
  const int32_t* ao_term_ptr;   /* (n_ao + 1) or NULL                                      */
  const double* term_coef;      /* (n_terms)                                               */
  const int32_t* term_pow;      /* (n_terms, 3)                                            */

  int32_t n_up, n_dn;           /* electrons; n_up >= n_dn                                 */

  /* Geminal_data (determinant.py:93-200): orbitals are MOs C.AO or the AOs themselves. */
  int32_t orb_is_mo;
  int32_t n_orb;
  const double* mo_coef_up;     /* (n_orb, n_ao) when orb_is_mo                            */
  const double* mo_coef_dn;     /* (n_orb, n_ao) when orb_is_mo                            */
  const double* lambda_matrix;  /* (n_orb, n_orb + n_up - n_dn) = [Lambda_p | Lambda_u]    */

  /* Jastrow_data (jastrow_factor.py:1864-2125). */
  int32_t j1_type;
  double j1_param;
  const double* j1_zeff;        /* (n_nuc) Z - core_electrons                              */
  int32_t j2_type;
  double j2_param;
  int32_t j3_on;
  const double* j3_matrix;      /* (n_ao, n_ao + 1), J3 over the AO basis                  */

  /* Coulomb_potential_data ECP tables (coulomb_potential.py:187-598). */
  int32_t ecp_on;
  int32_t ecp_lmax;             /* projected channels l = 0 .. ecp_lmax-1                  */
  int32_t ecp_max_terms;        /* padded non-local terms per nucleus                      */
  const int32_t* ecp_n_terms;   /* (n_nuc)                                                 */
  const int32_t* ecp_term_l;    /* (n_nuc, ecp_max_terms)                                  */
  const double* ecp_term_exp;   /* (n_nuc, ecp_max_terms)                                  */
  const double* ecp_term_coef;
  const double* ecp_term_pow;   /* stored power n of c r^(n-2) exp(-a r^2)                 */
  int32_t ecp_max_local;        /* padded local terms per nucleus                          */
  const int32_t* ecpl_n_terms;  /* (n_nuc)                                                 */
  const double* ecpl_exp;       /* (n_nuc, ecp_max_local)                                  */
  const double* ecpl_coef;
  const double* ecpl_pow;
  int32_t nv;                   /* quadrature points (coulomb_potential.py:96-184)         */
  const double* mesh_points;    /* (nv, 3)                                                 */
  const double* mesh_weights;   /* (nv)                                                    */
  int32_t nn;                   /* nearest nuclei for the projector (1)                    */

  int32_t precision;            /* JQMC_PRECISION_*                                        */
} jqmc_system_desc;

typedef struct jqmc_system jqmc_system;

/*
 * Per-walker state, all device pointers with a leading walker axis W.  Electrons
 * are ordered up block then down block (N = n_up + n_dn).  The reference's MCMC
 * carry is (accepted, rejected, r_up, r_dn, key, G^-1, G) (jqmc_mcmc.py:4317-4533);
 * ao/orb/pair/jfield are the single-electron streaming caches that replace its
 * per-move recomputation of full orbital matrices (determinant.py:1429-1526).
 */
typedef struct {
  double* r_up;       /* (W, n_up, 3)                                                 */
  double* r_dn;       /* (W, n_dn, 3)                                                 */
  /* ao and jfield: double for JQMC_PRECISION_PLAIN_FP64, float for JQMC_PRECISION_MIXED_FP32;
     orb and pair: double in both */
  void* ao;           /* (W, N, n_ao)   AO values at each electron                   */
  void* orb;          /* (W, N, n_orb)  geminal orbital values (double)               */
  void* pair;         /* (W, N, n_orb)  Lambda_p^T phi (up) / Lambda_p phi (dn) (double) */
  void* jfield;       /* (W, N, n_ao)   dJ3/da_e (NULL when J3 is off)                 */
  double* G;          /* (W, n_up, n_up) geminal matrix                               */
  double* Ginv;       /* (W, n_up, n_up) its inverse                                  */
  double* logabsdet;  /* (W) ln|det G|                                                */
  double* sign;       /* (W) sign of det G                                            */
  double* jastrow;    /* (W) J = J1 + J2 + J3 (no exponential)                        */
  double* as_factor;  /* (W) Attaccalite-Sorella factor (may be NULL)                 */
  int64_t* accepted;  /* (W) running accepted-move counters                           */
  int64_t* rejected;  /* (W) running rejected-move counters                           */
} jqmc_walkers;

int jqmc_abi_version(void);
const char* jqmc_last_error(void);

/* Builds the device-resident system pack from Hamiltonian_data (hamiltonians.py:82-177). */
int jqmc_system_create(const jqmc_system_desc* desc, jqmc_system** out);
int jqmc_system_destroy(jqmc_system* sys);

/*
 * Launch-shape knobs of a system, set by the autotuner (jqmc_amd/autotune.py) or by hand; value 0
 * restores the library's default.  The walkers-per-block knobs change no result bit: they choose how
 * many walkers share a block in the E_L kernel and the move-synchronous proposal kernel (1, 2 or 4,
 * capped by the LDS budget).  JQMC_TUNE_EL_WAVES_PER_WALKER (1, 2 or 4; one walker per block when
 * above 1 and its LDS fits) gives the E_L kernel a team of waves per walker, whose partial sums are
 * added in wave order: e_L then differs from the one-wave result in the last bits, as a fixed
 * function of the value.  No counterpart in the reference (XLA picks its own launch shapes).
 */
#define JQMC_TUNE_EL_WALKERS_PER_BLOCK 2
#define JQMC_TUNE_SWEEP_WALKERS_PER_BLOCK 3
#define JQMC_TUNE_EL_WAVES_PER_WALKER 4
int jqmc_system_set_tuning(jqmc_system* sys, int32_t knob, int32_t value);
int jqmc_system_get_tuning(const jqmc_system* sys, int32_t knob, int32_t* value);
/*
 * FP64 guard of the mixed form's local energy (no counterpart in the reference, whose mixed mode has
 * none): walkers whose mixed-precision e_L is not finite or lies farther than `width` from `e_ref`
 * (walkers near a node or a nucleus, where E_L is large and the FP32 rounding of the AO values and
 * J3 fields is amplified) are recomputed entirely in FP64 from their positions by the same
 * jqmc_local_energy call.  Default e_ref = 0, width = 1e4 Ha; width <= 0 turns the guard off.  No
 * effect on plain FP64 systems.
 */
int jqmc_system_set_mixed_guard(jqmc_system* sys, double e_ref, double width);
/* The system's JQMC_PRECISION_* (element type of the walker caches). */
int jqmc_system_precision(const jqmc_system* sys, int32_t* precision);

/*
 * AO value / gradient / Laplacian at n_pts points r (n_pts, 3); outputs (n_ao, n_pts).
 * Replaces compute_AOs (atomic_orbital.py:1886-1916) when gx == NULL and
 * compute_AOs_value_grad_lap (atomic_orbital.py:3594-3640) otherwise.
 */
int jqmc_aos_value_grad_lap(const jqmc_system* sys, int64_t n_pts, const double* r, double* val, double* gx,
                            double* gy, double* gz, double* lap, void* stream);

/*
 * Orbital (MO) value / gradient / Laplacian for spin 0 (up) or 1 (dn) coefficients;
 * outputs (n_orb, n_pts).  Replaces compute_MOs (molecular_orbital.py:239-261) and
 * compute_MOs_value_grad_lap (molecular_orbital.py:375-415).  workspace must hold
 * 5 * n_ao * n_pts doubles.
 */
int jqmc_orbs_value_grad_lap(const jqmc_system* sys, int32_t spin, int64_t n_pts, const double* r, double* val,
                             double* gx, double* gy, double* gz, double* lap, double* workspace, void* stream);

/*
 * Builds every cache of wk from r_up/r_dn: AO and orbital values, pair fields, the
 * geminal G (compute_geminal_all_elements, determinant.py:1336-1402), G^-1, ln|det G| and
 * its sign by pivoted FP64 LU (compute_ln_det_geminal_all_elements, determinant.py:1026-1048;
 * _geminal_inv_batched, jqmc_mcmc.py:4248-4275), the J3 fields, J (compute_Jastrow_part,
 * jastrow_factor.py:2127-2178) and the AS factor (determinant.py:1223-1260).
 * workspace: jqmc_walkers_init_workspace_bytes() bytes.
 */
int jqmc_walkers_init_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_walkers_init(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, void* workspace,
                      int64_t workspace_bytes, void* stream);

/*
 * Rebuilds G from the cached orbital and pair fields and replaces G^-1, ln|det G|, its sign and
 * the AS factor by a fresh pivoted FP64 inverse: the periodic refresh of the Sherman-Morrison
 * updates (the reference's SVD rebuild, jqmc_gfmc.py:4706-4717, and the mixed-precision refresh
 * of doc/notes/mixed_precision.md).  No workspace.
 */
int jqmc_walkers_refresh_inverse(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, void* stream);

/*
 * Recomputes, from the current positions, the AO values at the electrons in FP64 (stored in the
 * AO cache's type), the orbital and pair rows, G, G^-1, ln|det G|, its sign and the AS factor; the
 * J3 fields and J are left as they are.  In the mixed form this resets the determinant side to
 * what the plain FP64 form holds at the same positions (the FP64 refresh of the mixed mode,
 * doc/notes/mixed_precision.md; the reference recomputes A^-1 from positions after branching,
 * jqmc_gfmc.py:6314-6319).  No workspace.
 */
int jqmc_walkers_refresh_orbitals(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, void* stream);

/*
 * Single-electron ratios at n_points proposed configurations per walker: configuration p of walker
 * w moves electron elec[w][p] (0 <= elec < N, up block first) to points[w][p][0..2] and leaves the
 * others where wk has them.  det_ratio[w][p] = det G' / det G from the walker's G^-1 and caches
 * (replaces _compute_ratio_determinant_part_rank1_update and _split_spin, determinant.py:1529-1783);
 * jastrow_ratio[w][p] = exp(J' - J) (replaces _compute_ratio_Jastrow_part_rank1_update and
 * _split_spin, jastrow_factor.py:2230-2960).  Consecutive configurations of one electron share its
 * determinant field.  An electron index outside [0, N) yields NaN for that configuration.
 */
int jqmc_point_ratios(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int64_t n_points,
                      const int32_t* elec, const double* points, double* det_ratio, double* jastrow_ratio,
                      void* stream);

/*
 * n_moves single-electron Metropolis proposals per walker (the lax.fori_loop of
 * _update_electron_positions, jqmc_mcmc.py:4278-4533) in one launch, one wavefront per
 * walker.  Draws come from the Philox4x32-10 stream (seed, walker, move0 + m).
 * Updates r, the caches, G, G^-1 and the counters in place.
 */
int jqmc_vmc_sweep(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                   double epsilon_AS, uint64_t seed, int64_t move0, void* stream);

/*
 * Measurement rotations R^T per walker (_generate_rotation_matrix, jqmc_mcmc.py:4228-4245)
 * from the Philox stream (seed, walker, step, stream_id); RT is (W, 3, 3).
 */
int jqmc_rotation_matrices(int64_t W, uint64_t seed, int64_t step, int32_t stream_id, double* RT, void* stream);

/*
 * Local energy with the walker's current G^-1 (compute_local_energy_fast,
 * hamiltonians.py:225-290): kinetic (compute_kinetic_energy_all_elements_fast_update,
 * wavefunction.py:1141-1207) + bare Coulomb (coulomb_potential.py:2252-2281) + ECP local
 * (1144-1246) + ECP non-local with NN=1 and the rotated nv-point mesh (1477-1712).
 * Outputs (W): e_L, kinetic, v_total (may be NULL); t_elec (W, N) per-electron kinetic
 * (may be NULL).  workspace: jqmc_local_energy_workspace_bytes() bytes (the FP64 determinant
 * field rows h_e of every electron, W x N x n_ao doubles, the W x N x n_orb q_e rows when
 * N x n_orb doubles exceed 16 KiB, and in the mixed form the list and FP64 scratch of the guard,
 * jqmc_system_set_mixed_guard).
 */
int jqmc_local_energy_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_local_energy(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const double* RT, double* e_L,
                      double* kinetic, double* v_total, double* t_elec, void* workspace, int64_t workspace_bytes,
                      void* stream);

/*
 * The same sweep taken move-synchronously: each of the n_moves moves is one launch over all
 * walkers (one wavefront per walker) followed by one tiled product [J3 | J3^T] x dA over every
 * walker that accepted, with the J3 field updates fused in.  Same draws, proposals and decisions
 * as jqmc_vmc_sweep (jqmc_mcmc.py:4316-4531); move counter move0 (+ *move_ctr when non-NULL).
 * workspace: jqmc_vmc_sweep_batched_workspace_bytes() bytes; no allocation, memset or
 * synchronisation, so it can be captured in a graph.
 */
int jqmc_vmc_sweep_batched_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_vmc_sweep_batched(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                           double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, void* workspace,
                           int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ graph-capturable step */

/*
 * Variants of jqmc_vmc_sweep and jqmc_rotation_matrices whose Philox move / step counter is
 * move0 + *move_ctr (step + *step_ctr) read from device memory at run time, and
 * jqmc_counter_add (*ctr += inc on the stream).  A step built from them (sweep, rotations,
 * local energy, counter advances) can be captured once in a hipGraph and replayed: every replay
 * draws the numbers the uncaptured calls with the advanced counters would draw.  The reference's
 * equivalent is the jit-compiled per-step function of MCMC.run (jqmc_mcmc.py:664-745).
 * None of them allocates, memsets or synchronises.
 */
int jqmc_vmc_sweep_counted(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                           double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, void* stream);
int jqmc_rotation_matrices_counted(int64_t W, uint64_t seed, int64_t step, const int64_t* step_ctr,
                                   int32_t stream_id, double* RT, void* stream);
int jqmc_counter_add(int64_t* ctr, int64_t inc, void* stream);

/* ------------------------------------------------------------------ position derivatives */

/*
 * Position derivatives of the local energy and of ln|Psi| per walker, and the space-warp force
 * products, replacing the reverse-mode gradients of the reference's MCMC.run:
 *   _jit_vmap_grad_e_L_r / _jit_vmap_grad_e_L_h (jqmc_mcmc.py:744-745, 752-780): dE_dr, dE_dR;
 *   _jit_vmap_grad_ln_psi (jqmc_mcmc.py:746, 782-792): dlnpsi_dr, dlnpsi_dR;
 *   evaluate_swct_omega / evaluate_swct_domega (swct.py:63-150) and the force products
 *   force_HF = dE_dR + omega . dE_dr, force_PP = dlnpsi_dR + omega . dlnpsi_dr + 1/2 sum grad omega
 *   (jqmc_mcmc.py:796-852); with swct = 0 the omega terms are zero.
 * Electrons are ordered up then dn; RT (W, 3, 3) is the quadrature rotation (NULL = identity).
 * Outputs: e_L (W), dE_dr and dlnpsi_dr (W, N, 3), dE_dR, dlnpsi_dR, force_HF, force_PP
 * (W, n_nuc, 3); any may be NULL.  Plain FP64 systems only (JQMC_ERR_UNSUPPORTED otherwise).
 * workspace: jqmc_position_grads_workspace_bytes() bytes, contents need not be initialised.
 */
int jqmc_position_grads_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_position_grads(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const double* RT, int32_t swct,
                        double* e_L, double* dE_dr, double* dE_dR, double* dlnpsi_dr, double* dlnpsi_dR,
                        double* force_HF, double* force_PP, void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ LRDMC */

/* Treatment of the ECP non-local part in LRDMC (jqmc_gfmc.py:4943-5021). */
enum { JQMC_TMOVE = 0, JQMC_DLTMOVE = 1 };
/* GFMC_n: a fixed number of projections per call (jqmc_gfmc.py:4191-5358);
   GFMC_t: projections until the time tau is used up (jqmc_gfmc.py:173-1107). */
enum { JQMC_LRDMC_FIXED_STEPS = 0, JQMC_LRDMC_FIXED_TAU = 1 };

typedef struct {
  int32_t mode;            /* JQMC_LRDMC_FIXED_STEPS or JQMC_LRDMC_FIXED_TAU                 */
  int32_t n_steps;         /* projections (FIXED_STEPS) / per-walker cap (FIXED_TAU)          */
  double alat;             /* lattice spacing a (bohr)                                        */
  double E_scf;            /* trial energy of the weight update (FIXED_STEPS)                 */
  double tau;              /* projection time (FIXED_TAU)                                     */
  int32_t non_local_move;  /* JQMC_TMOVE or JQMC_DLTMOVE                                      */
  int32_t random_mesh;     /* random lattice rotation every projection (random_discretized_mesh) */
} jqmc_lrdmc_params;

/*
 * LRDMC projections in one launch, one wavefront per walker (_projection_n / _body_step_core,
 * jqmc_gfmc.py:4738-5166; _projection_t_core, 724-1107).  Per projection: continuum kinetic
 * energies, 6N lattice moves r_i +- a R e_x,y,z with ratios from the walker caches
 * (compute_discretized_kinetic_energy_fast_update, wavefunction.py:1739-1860), fixed-node /
 * sign-flip split and regularised electron-ion diagonal, ECP local + t-move / DLA t-move
 * elements, the weight update (FIXED_STEPS: w *= -sum offdiag / (sum diag - E_scf); FIXED_TAU:
 * dtau = min(tau_left, ln(1 - xi) / sum offdiag), w *= exp(-dtau e_L)), the move chosen by
 * cumulative probabilities and the single-electron update of r, caches, G and G^-1.  Draws of
 * projection step0 + i come from the Philox stream (seed, walker, step0 + i, LRDMC).
 * w_L (W) is read and updated; RT_last (W, 3, 3) gets the last rotation; V_diag / V_nondiag (W)
 * the diagonal and off-diagonal sums of the last projection; tau_left (W) and n_proj (W) the
 * remaining time and the projection count (all four may be NULL).  The walkers' jastrow,
 * logabsdet, sign and as_factor are not updated.  workspace: jqmc_lrdmc_workspace_bytes() bytes (the
 * FP64 determinant field rows h_e of every walker, W x N x n_ao doubles, the W x N x n_orb q_e
 * rows when N x n_orb doubles exceed 16 KiB, and the per-walker state between a projection's
 * launches: candidates, sums, the moved electron and its AO change).
 */
int jqmc_lrdmc_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_lrdmc_projection(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const jqmc_lrdmc_params* params,
                          uint64_t seed, int64_t step0, double* w_L, double* RT_last, double* V_diag,
                          double* V_nondiag, double* tau_left, int64_t* n_proj, void* workspace,
                          int64_t workspace_bytes, void* stream);

/*
 * The projections of jqmc_lrdmc_projection for n_sys (<= 8) systems of uneven size in ONE launch
 * (batched LRDMC of several molecules): system k has its own jqmc_system (all of one precision),
 * W[k] walkers wks[k], params[k] and step0[k]; per-system output arrays as in
 * jqmc_lrdmc_projection (the pointer arrays RT_last ... n_proj may be NULL, their entries too).
 * Blocks are assigned to systems in order, each with its own walkers-per-block.  System k draws
 * from the Philox key seed + k, so that its walkers' draws are those of a jqmc_lrdmc_projection
 * call with that seed and are independent of the other systems' walkers of the same index.
 * workspaces[k]: jqmc_lrdmc_workspace_bytes(systems[k], W[k]) bytes, sizes in workspace_bytes[k].
 */
int jqmc_lrdmc_projection_multi(int32_t n_sys, const jqmc_system* const* systems, const int64_t* W,
                                const jqmc_walkers* wks, const jqmc_lrdmc_params* params, uint64_t seed,
                                const int64_t* step0, double* const* w_L, double* const* RT_last,
                                double* const* V_diag, double* const* V_nondiag, double* const* tau_left,
                                int64_t* const* n_proj, void* const* workspaces, const int64_t* workspace_bytes,
                                void* stream);

/*
 * V_diag and V_nondiag at the walkers' configurations with rotations RT (W, 3, 3) or the identity
 * when RT is NULL (_compute_V_elements_n, jqmc_gfmc.py:5360-5628); the LRDMC local energy is
 * V_diag + V_nondiag.  Only alat and non_local_move of params are used.  workspace:
 * jqmc_lrdmc_workspace_bytes() bytes.
 */
int jqmc_lrdmc_local_energy(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const jqmc_lrdmc_params* params,
                            const double* RT, double* V_diag, double* V_nondiag, void* workspace,
                            int64_t workspace_bytes, void* stream);

/*
 * Comb reconfiguration of W walkers (jqmc_gfmc.py:6070-6130 on one process): walker j takes
 * parent chosen[j] = searchsorted(cumsum(w / sum w), (j + zeta) / W), zeta in [0, 1).  When
 * V_diag and e_L are given, sums (4) gets sum w, sum w/(V_diag - E_scf), sum w e_L/(V_diag - E_scf)
 * and sum w e_L^2/(V_diag - E_scf) (jqmc_gfmc.py:5957-5964).  workspace: W doubles.
 */
int jqmc_lrdmc_branching_workspace_bytes(int64_t W, int64_t* bytes);
int jqmc_lrdmc_branching(int64_t W, const double* w_L, const double* V_diag, const double* e_L, double E_scf,
                         double zeta, int64_t* chosen, double* sums, void* workspace, int64_t workspace_bytes,
                         void* stream);

/*
 * jqmc_lrdmc_branching without a workspace argument and without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The comb's W-double buffer is a device array inside the
 * library, one per GPU (JQMC_BRANCHING_STATIC_MAX_W doubles = 8 MiB, held until the process ends);
 * W above that maximum is refused with JQMC_ERR_SHAPE.  The array keeps the last call's contents, and
 * calls that overlap on one GPU race on it: serialise them (one stream, or events).
 */
#define JQMC_BRANCHING_STATIC_MAX_W (1 << 20)
int jqmc_lrdmc_branching_static(int64_t W, const double* w_L, const double* V_diag, const double* e_L, double E_scf,
                                double zeta, int64_t* chosen, double* sums, void* stream);

/*
 * Walker j takes the state of walker chosen[j] (the comb's copy, jqmc_gfmc.py:6086-6130): positions,
 * AO / orbital / pair / J3-field caches, G, G^-1, ln|det G|, sign, J and the AS factor are copied
 * from the parent (walkers with chosen[j] == j are untouched), then G^-1 is rebuilt from the
 * positions as by jqmc_walkers_refresh_orbitals (the reference recomputes G^-1 after branching,
 * jqmc_gfmc.py:6314-6319) and J is recomputed from the copied caches.  The accept/reject counters stay with the slot.  workspace: one staging
 * record per walker, jqmc_walkers_reconfigure_workspace_bytes() bytes.
 */
int jqmc_walkers_reconfigure_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_walkers_reconfigure(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const int64_t* chosen,
                             void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ parameter log-derivatives */

/* Variational blocks (Wavefunction_data.get_variational_blocks, wavefunction.py:515-640). */
enum { JQMC_PARAM_J1 = 1, JQMC_PARAM_J2 = 2, JQMC_PARAM_J3 = 4, JQMC_PARAM_LAMBDA = 8 };

/*
 * Layout of the per-walker vector O_k = d ln|Psi| / d p_k for the blocks selected in `blocks` that
 * the system has, in the order J1 b, J2 b, J3 matrix (n_ao x (n_ao + 1) row-major over the device
 * AO basis, the one-body column last), Lambda (n_orb x (n_orb + n_up - n_dn) row-major):
 * offsets[0..3] get each block's offset (-1 when absent), *n_params the length.
 */
int jqmc_param_layout(const jqmc_system* sys, int32_t blocks, int64_t* offsets, int64_t* n_params);

/*
 * O (W, ld) row per walker, ld >= n_params, from the walkers' current caches and G^-1: the SR
 * inputs of MCMC.run (grad of evaluate_ln_wavefunction_fast w.r.t. the parameters,
 * jqmc_mcmc.py:854-873).  J3 and Lambda_p blocks are rank-N / rank-n_dn outer products
 * (sum_{e<f} a_e a_f^T and (G^-1 Phi_up)[:n_dn]^T Phi_dn) evaluated in 64 x 64 tiles.
 * workspace: jqmc_dln_psi_dparams_workspace_bytes() bytes.
 */
int jqmc_dln_psi_dparams_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes);
int jqmc_dln_psi_dparams(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t blocks, double* O,
                         int64_t ld, void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------ stochastic reconfiguration */

/*
 * The SR step of MCMC.run_optimize (jqmc_mcmc.py:2952-3350) works on the (S x P) matrix O of
 * per-sample log-derivatives (S = samples, P = parameters), row-major with leading dimension ld.
 * These three calls are its device parts; the S x S (or P x P) solve is a plain FP64 Cholesky.
 *
 * jqmc_sr_colstats: mean_k = sum_i wn_i O_ik and var_k = sum_i wn_i (O_ik - mean_k)^2 for weights
 * wn that sum to one: O_bar and diag S of jqmc_mcmc.py:2980-2990 and 3062-3064.
 */
int jqmc_sr_colstats(const double* O, int64_t S, int64_t P, int64_t ld, const double* wn, double* mean, double* var,
                     void* stream);

/*
 * jqmc_sr_gram: G (n x n, leading dimension ldg) = Y Y^T, Y_rc = (M_rc - a_r - b_c) u_r v_c for the
 * row-major (n x m) matrix M (leading dimension ldm); a and b may be NULL.  With M = O, b = O_bar,
 * u = sqrt(w / sum w) and v = diag_S^-1/2 this is X^T X of the push-through branch
 * (jqmc_mcmc.py:3198-3350); with M = O^T, a = O_bar, u = diag_S^-1/2 and v = sqrt(w / sum w) it is
 * the X X^T of the P < S branch (3114-3196).  FP64 on the matrix cores.
 */
int jqmc_sr_gram(const double* M, int64_t n, int64_t m, int64_t ldm, const double* a, const double* b, const double* u,
                 const double* v, double* G, int64_t ldg, void* stream);

/*
 * jqmc_sr_gram_f32: the same G with Y rounded to FP32 after its FP64 transform and the products
 * accumulated in FP32 on the matrix cores; G is written in FP64, where the caller adds the FP64
 * diagonal shift.  The reference forms this product in FP64 only (jqmc_mcmc.py:3198-3350); this is
 * the reduced-precision option of the same call.
 */
int jqmc_sr_gram_f32(const double* M, int64_t n, int64_t m, int64_t ldm, const double* a, const double* b,
                     const double* u, const double* v, double* G, int64_t ldg, void* stream);

/*
 * jqmc_sr_project: out_k = cs_k sum_i (O_ik - mean_k) rs_i x_i, i.e. X x with
 * X_ki = (O_ik - mean_k) rs_i cs_k: the X F right-hand side and the push-through theta = X x
 * (jqmc_mcmc.py:3005-3010, 3290-3300).
 */
int jqmc_sr_project(const double* O, int64_t S, int64_t P, int64_t ld, const double* mean, const double* rs,
                    const double* cs, const double* x, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* JQMC_HIP_H */
