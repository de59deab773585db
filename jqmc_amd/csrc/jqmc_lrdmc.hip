// gfx950 kernels of the LRDMC (lattice-regularised diffusion Monte Carlo) path.  One projection of
// every walker is a sequence of launches:
//
//   k_lrdmc_ham[_team|_multi]  lattice Hamiltonian elements at the current configuration (one
//                  wavefront, or a team of waves, per walker); also the measurement of V_diag /
//                  V_nondiag at a given rotation
//   k_lrdmc_move[_multi]  weight update, choice of one off-diagonal move and its rank-1 state
//                  update, one wavefront per walker
//   k_sweep_jfield_mfma<T, true>  J3 fields of the other electrons of every moved walker, batched
//                  over walkers on the matrix cores (the sweep's J3 product kernel with a direct
//                  epilogue: every field row takes its increment)
//
// repeated n_steps times (GFMC_n) or until the projection time is used up (GFMC_t).  Then:
//   k_comb         comb reconfiguration of the walker population (one block).
//   k_state_gather / k_gather_rows  walker rows copied from the chosen parents.
//
// Walker caches are the ones the VMC sweep keeps (AO values, orbitals, pair fields, J3 fields,
// G, G^-1).  Like the other kernels, these are templates over the cache / AO arithmetic type T;
// the weights, the Hamiltonian sums, G and G^-1 are FP64 in both forms.
#pragma once

#include "jqmc_el.hip"
#include "jqmc_sweep_batched.hip"

namespace jqmc {

// Modes of a projection call.
constexpr int LR_FIXED_STEPS = 0;  // GFMC_n: n_steps projections, w <- w b_x
constexpr int LR_FIXED_TAU = 1;    // GFMC_t: projections until tau is used, w <- w exp(-dtau e_L)
constexpr int LR_MEASURE = 2;      // V_diag / V_nondiag at the given rotation, no move

// ---- one projection in two launches: the Hamiltonian of every walker, then its move

// Per-walker state of a projection call between its launches, in the workspace after the field
// rows: the candidate lists, the rotation, the diagonal / off-diagonal sums, the time left, the
// projection count, the AO change of the last move and its electron (the J3 field launch's input),
// the candidate count and the electrons with projected ECP terms.
struct LrState {
  double* cand;    // (W, n_cmax)
  double* RT;      // (W, 9)
  double* diag;    // (W)
  double* nondiag; // (W)
  double* tau;     // (W)
  int64_t* nproj;  // (W)
  double* dA;      // (W, n_ao) values of the cache type T (T* cast)
  int* ncand;      // (W)
  int* elist;      // (W, N)
  int* mv;         // (W) electron moved by the last move launch, -1 when none
  int* act;        // (2) walkers with time left after the projections it, it + 2, ... (it & 1) (GFMC_t)
  int n_cmax;
};

// Doubles of the field rows, the q rows when they do not fit LDS and the state above, per walker.
__host__ __device__ inline int64_t lr_cmax(const Sys& s) { return 6 * (int64_t)s.n_el + (int64_t)s.n_el * (s.nv > 0 ? s.nv : 0); }
__host__ __device__ inline int64_t lr_ws_doubles(const Sys& s) {
  return (int64_t)s.n_el * s.n_ao + (el2_q_in_lds(s) ? 0 : (int64_t)s.n_el * s.n_orb) + lr_cmax(s) + 9 + 4 +
         s.n_ao + (2 + (int64_t)s.n_el + 1) / 2;
}
inline int64_t lr_ws_bytes(const Sys& s, int64_t W) { return W * lr_ws_doubles(s) * (int64_t)sizeof(double) + 16; }

inline LrState lr_state(const Sys& s, int64_t W, double* base) {
  LrState S;
  double* p = base + W * ((int64_t)s.n_el * s.n_ao + (el2_q_in_lds(s) ? 0 : (int64_t)s.n_el * s.n_orb));
  S.n_cmax = (int)lr_cmax(s);
  S.cand = p;
  p += W * S.n_cmax;
  S.RT = p;
  p += 9 * W;
  S.diag = p;
  p += W;
  S.nondiag = p;
  p += W;
  S.tau = p;
  p += W;
  S.nproj = reinterpret_cast<int64_t*>(p);
  p += W;
  S.dA = p;
  p += W * (int64_t)s.n_ao;
  S.ncand = reinterpret_cast<int*>(p);
  S.elist = S.ncand + W;
  S.mv = S.elist + W * (int64_t)s.n_el;
  S.act = reinterpret_cast<int*>(base + W * lr_ws_doubles(s));
  return S;
}

// LDS of one walker of the Hamiltonian launch (a team of nw waves): the local-energy block and the
// candidate list.
template <typename T>
struct LrHamLds {
  El2Lds E;
  double *rr, *RT, *cand;
  uintptr_t end;
  __host__ __device__ LrHamLds(char* base, const Sys& s, int nw = 1) : E(base, s, nw) {
    rr = E.rr;
    RT = E.RT;
    Carve c{base + (E.end - reinterpret_cast<uintptr_t>(base))};
    cand = c.take<double>(lr_cmax(s) > 0 ? lr_cmax(s) : 1);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

// LDS of one wavefront of the move launch.
template <typename T>
struct LrMoveLds {
  double *rr, *Gi, *RT, *gk, *Vt, *vv, *xx, *yy, *phin, *dphi;
  T *anew, *dA;
  uintptr_t end;
  __host__ __device__ LrMoveLds(char* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up, na = s.n_ao, no = s.n_orb, nv = s.nv > 0 ? s.nv : 1;
    Carve c{base};
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(nu * nu);
    RT = c.take<double>(9);
    gk = c.take<double>(3 * nv);
    Vt = c.take<double>(s.ecp_max_terms > 0 ? s.ecp_max_terms : 1);
    vv = c.take<double>(nu);
    xx = c.take<double>(nu);
    yy = c.take<double>(nu);
    phin = c.take<double>(no);
    dphi = c.take<double>(no);
    anew = c.take<T>(na);
    dA = c.take<T>(na);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

template <typename T>
inline int64_t lr_ham_lds_bytes(const Sys& s, int nw = 1) {
  return (int64_t)LrHamLds<T>(0, s, nw).end;
}
template <typename T>
inline int64_t lr_move_lds_bytes(const Sys& s) {
  return (int64_t)LrMoveLds<T>(0, s).end;
}

struct LrArgs {
  double alat, E_scf, tau;
  int n_steps, mode, dlt, random_mesh;
  uint64_t seed;
  int64_t step0;
  double* w;          // (W) weights, in/out
  double* RT;         // (W, 9) rotation: written by the projection, read by the measurement
  double* V_diag;     // (W) out, may be null
  double* V_nondiag;  // (W) out, may be null
  double* tau_left;   // (W) out, may be null
  int64_t* n_proj;    // (W) out, may be null
  double* hws;        // workspace of lr_ws_bytes(s, W) bytes: field rows, then the state
  LrState st;         // lr_state() of that workspace
};

// Diagonal and off-diagonal sums of the lattice-regularised Hamiltonian at the current
// configuration (_body_step_core, jqmc_gfmc.py:4807-5045; _compute_V_elements_n, 5360-5628), with the
// phases of the local energy: determinant fields, continuum kinetic energies T_e (lanes own
// electrons), the per-electron ECP data, then every candidate point with the lanes owning points --
// the 6 lattice moves r_e + s_k, s = a [+x, -x, +y, -y, +z, -z] R^T (wavefunction.py:1779-1860) of
// every electron and the nv quadrature points of the electrons whose nearest nucleus carries
// projected ECP terms.  Leaves in L.cand the candidate list of the move choice: min(element, 0) of
// the lattice moves of every electron (up block first), then the fixed-node ECP elements of those
// electrons' quadrature points in electron order (L.E.elist); *n_cand is its length.  Hw / Qg: the
// walker's field rows and workspace q rows (used when the q rows do not fit LDS).
template <typename T, int NW, bool kWide>
__device__ void lattice_hamiltonian(const Sys& s, const LrHamLds<T>& L, const double* Qw, double* Qg, double* Hw,
                                    const T* gw, double alat, int dlt, const Team<NW>& tm, double& diag,
                                    double& nondiag, int& n_cand) {
  const int ne = s.n_el, lane = tm.lane;
  const int n_lat = 6 * ne;
  const bool nonlocal = s.ecp_on && s.nv > 0;
  const double kfac = -1.0 / (2.0 * (alat * alat));
  const El2Lds& E = L.E;
  if (nonlocal) rotate_mesh(s, L.RT, E.gk, tm.tid, tm.N);
  el_det_fields<NW, kWide>(s, E, Qw, Qg, Hw, tm);
  el_kinetic<T>(s, E, Hw, gw, tm, nullptr);
  const int n_ecp = el_ecp_electrons(s, E, tm, true);
  tm.sync();
  const int nv = nonlocal ? s.nv : 0;
  const int P = n_lat + n_ecp * nv;
  double sp_fn[2] = {0.0, 0.0};  // ECP: sum max(V, 0), sum fixed-node elements
  for (int base = tm.wid * WAVE; base < P; base += tm.N) {
    const int idx = base + lane;
    const bool valid = idx < P;
    const int id = valid ? idx : 0;
    const bool lat = id < n_lat;
    int e;
    double px, py, pz, pwt = 0.0;
    if (lat) {
      e = id / 6;
      const int q = id - 6 * e, axis = q >> 1;
      const double sa = (q & 1) ? -alat : alat;
      px = L.rr[3 * e] + sa * L.RT[3 * axis];
      py = L.rr[3 * e + 1] + sa * L.RT[3 * axis + 1];
      pz = L.rr[3 * e + 2] + sa * L.RT[3 * axis + 2];
    } else {
      const int r = id - n_lat, ie = r / nv, k = r - ie * nv;
      e = E.elist[ie];
      el_ecp_point(s, E, e, k, px, py, pz, pwt);
    }
    double det, dj;
    el_point<T, T>(s, E, e, px, py, pz, Hw, gw, det, dj);
    const double jr = exp(dj);
    if (valid) {
      if (lat) {
        L.cand[id] = kfac * (det * jr);
      } else {
        // t-move: V = w_P Psi(r')/Psi(r); DLA t-move: V = w_P det ratio and the Jastrow ratio multiplies
        // the fixed-node part only (jqmc_gfmc.py:4956-5021)
        const double V = dlt ? pwt * det : pwt * (det * jr);
        double fn = fmin(V, 0.0);
        if (dlt) fn = fn * jr;
        sp_fn[0] += fmax(V, 0.0);
        sp_fn[1] += fn;
        L.cand[id] = fn;
      }
    }
  }
  tm.sync();

  // ---- per electron: sign-flip flag, regularised electron-ion diagonal (jqmc_gfmc.py:4833-4939)
  const double quarter = 1.0 / (4.0 * (alat * alat));
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, sp_fn[0], sp_fn[1]};  // sum opt, SP, FN, ECP local, el-el, ECP SP, ECP FN
  for (int e = tm.tid; e < ne; e += tm.N) {
    bool flag = false;
    double nd = 0.0;
    for (int k = 0; k < 6; ++k) {
      const double x = L.cand[6 * e + k];
      flag |= x >= 0.0;
      nd += x + quarter;
      acc[1] += fmax(x, 0.0);
      const double fn = fmin(x, 0.0);
      acc[2] += fn;
      L.cand[6 * e + k] = fn;
    }
    const double ex = L.rr[3 * e], ey = L.rr[3 * e + 1], ez = L.rr[3 * e + 2];
    double v_ei = 0.0, v_fl = 0.0;
    for (int n = 0; n < s.n_nuc; ++n) {
      const double dx = ex - s.nuc_pos[3 * n], dy = ey - s.nuc_pos[3 * n + 1], dz = ez - s.nuc_pos[3 * n + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      v_ei += -s.nuc_charge[n] / d;
      v_fl += -s.nuc_charge[n] / fmax(d, alat);
    }
    const double zv = (v_ei + E.te[e]) - nd;
    const double ei = s.ecp_on ? v_ei : v_fl;
    acc[0] += flag ? fmax(zv, ei) : zv;
  }
  if (s.ecp_on) {
    // ECP local part (coulomb_potential.py:1144-1246), lanes over electron-nucleus pairs
    for (int t = tm.tid; t < ne * s.n_nuc; t += tm.N) {
      const int e = t / s.n_nuc, n = t - e * s.n_nuc;
      const double dx = L.rr[3 * e] - s.nuc_pos[3 * n], dy = L.rr[3 * e + 1] - s.nuc_pos[3 * n + 1],
                   dz = L.rr[3 * e + 2] - s.nuc_pos[3 * n + 2];
      acc[3] += ecp_local_pair(s, n, sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  acc[4] = el_el_partial(s, L.rr, tm.tid, tm.N);
  tm.template sum_k<7>(acc);
  diag = 3.0 / (2.0 * (alat * alat)) * ne + ((acc[4] + s.ion_ion) + acc[0]);
  nondiag = acc[2];
  n_cand = P;
  if (s.ecp_on) {
    diag = ((diag + acc[3]) + acc[1]) + acc[5];
    nondiag += acc[6];
  } else {
    diag += acc[1];
  }
  tm.sync();
}

// k = first index with cumsum(cand / psum)[k] >= u (searchsorted, jqmc_gfmc.py:5057-5062), the
// last candidate when none qualifies (JAX clamps the gather index).
__device__ int choose_candidate(const double* cand, int n, double psum, double u, int lane) {
  double carry = 0.0;
  for (int base = 0; base < n; base += WAVE) {
    const int p = base + lane;
    double x = p < n ? cand[p] / psum : 0.0;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
      const double y = __shfl_up(x, off, WAVE);
      if (lane >= off) x += y;
    }
    const double c = carry + x;
    const uint64_t hit = __ballot(p < n && c >= u);
    if (hit) return base + (int)__builtin_ctzll(hit);
    carry = __shfl(c, WAVE - 1, WAVE);
  }
  return n - 1;
}

// Electron e moves to (nx, ny, nz): AO, orbital and pair rows, G and G^-1 by Sherman-Morrison
// (_update_inv_up_n / _update_inv_dn_n, jqmc_gfmc.py:5083-5141); the AO change goes to dAw (null
// without J3) for the J3 field update of the other electrons.
template <typename T, typename LDS>
__device__ void lrdmc_apply_move(const Sys& s, const LDS& L, T* Aw, double* Pw, double* Qw, T* dAw, double* Gw, int e,
                                 double nx, double ny, double nz, int lane) {
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  const RealTabs<T>& R = tabs<T>(s);
  const bool is_up = e < nu;
  const int idx = is_up ? e : e - nu;
  T* Ae = Aw + (int64_t)e * na;
  for (int a = lane; a < na; a += WAVE) {
    const T v = ao_value<T>(s, a, nx, ny, nz);
    L.anew[a] = v;
    L.dA[a] = v - Ae[a];
  }
  wave_sync();
  // the J3 fields of the other electrons follow in k_sweep_jfield_mfma<T, true>, batched over walkers
  if (dAw)
    for (int a = lane; a < na; a += WAVE) dAw[a] = L.dA[a];
  const double* Pe = Pw + (int64_t)e * no;
  if (s.orb_is_mo) {
    wave_matvec_T(is_up ? s.td.mo_upT : s.td.mo_dnT, L.anew, na, no, lane, L.phin);
    wave_sync();
  } else {
    for (int k = lane; k < no; k += WAVE) L.phin[k] = (double)L.anew[k];
  }
  for (int k = lane; k < no; k += WAVE) L.dphi[k] = L.phin[k] - Pe[k];
  wave_sync();
  // changed geminal row (up) or column (dn), 1 + v^T G^-1 u, x = G^-1 u, y = v^T G^-1
  double part = 0.0;
  if (is_up) {
    for (int j = lane; j < nu; j += WAVE) {
      double v = 0.0;
      if (j < nd) {
        const double* pr = Qw + (int64_t)(nu + j) * no;
        for (int k = 0; k < no; ++k) v += L.dphi[k] * pr[k];
      } else {
        for (int k = 0; k < no; ++k) v += L.dphi[k] * s.td.lam_u[(int64_t)k * s.n_unp + (j - nd)];
      }
      L.vv[j] = v;
      part += v * L.Gi[j * nu + idx];
    }
  } else {
    for (int i = lane; i < nu; i += WAVE) {
      const double* pr = Qw + (int64_t)i * no;
      double u = 0.0;
      for (int k = 0; k < no; ++k) u += pr[k] * L.dphi[k];
      L.vv[i] = u;
      part += L.Gi[idx * nu + i] * u;
    }
  }
  const double ratio = 1.0 + wave_sum(part);
  wave_sync();
  if (is_up) {
    for (int p = lane; p < nu; p += WAVE) {
      L.xx[p] = L.Gi[p * nu + idx];
      double y = 0.0;
      for (int j = 0; j < nu; ++j) y += L.vv[j] * L.Gi[j * nu + p];
      L.yy[p] = y;
    }
  } else {
    for (int p = lane; p < nu; p += WAVE) {
      double x = 0.0;
      for (int i = 0; i < nu; ++i) x += L.Gi[p * nu + i] * L.vv[i];
      L.xx[p] = x;
      L.yy[p] = L.Gi[idx * nu + p];
    }
  }
  wave_sync();
  for (int t = lane; t < nu * nu; t += WAVE) {
    const int p = t / nu, q = t - p * nu;
    L.Gi[t] -= L.xx[p] * L.yy[q] / ratio;
  }
  for (int p = lane; p < nu; p += WAVE) {
    if (is_up) Gw[(int64_t)idx * nu + p] += L.vv[p];
    else Gw[(int64_t)p * nu + idx] += L.vv[p];
  }
  for (int a = lane; a < na; a += WAVE) Ae[a] = L.anew[a];
  double* Pem = Pw + (int64_t)e * no;
  double* Qem = Qw + (int64_t)e * no;
  const double* Lm = is_up ? s.td.lam_p : s.td.lam_pT;
  for (int k = lane; k < no; k += WAVE) {
    Pem[k] = L.phin[k];
    double qv = 0.0;
    for (int j = 0; j < no; ++j) qv += Lm[(int64_t)j * no + k] * L.phin[j];
    Qem[k] = qv;
  }
  if (lane == 0) {
    L.rr[3 * e] = nx;
    L.rr[3 * e + 1] = ny;
    L.rr[3 * e + 2] = nz;
  }
  wave_sync();
}

// Projection `it` of a call (GFMC_n _projection_n, jqmc_gfmc.py:4738-5358; GFMC_t _projection_t_core,
// 724-1107), part 1, one wavefront per walker: the rotation of projection step0 + it (from the LRDMC
// Philox stream, sub-streams 0-1, when random_mesh), then the lattice-regularised Hamiltonian; the
// candidate list, sums and rotation go to the state for the move launch.  LR_MEASURE writes V_diag /
// V_nondiag at the given rotations instead.  Walkers whose time is used up (GFMC_t) sit out.
template <typename T, int NW, bool kWide>
__device__ __forceinline__ void lr_ham_walker(const Sys& s, const WalkT<T>& wk, const LrArgs& A, int64_t W, int64_t w, int it,
                              const LrHamLds<T>& L, const Team<NW>& tm) {
  const int i = tm.tid;
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  const LrState& S = A.st;
  const bool measure = A.mode == LR_MEASURE;
  if (!measure) {
    if (it >= A.n_steps) return;
    if (A.mode == LR_FIXED_TAU && w == 0 && i == 0) S.act[it & 1] = 0;  // counted by this round's moves
    if (it == 0) {
      if (i == 0) {
        S.tau[w] = A.tau;
        S.nproj[w] = 0;
      }
    } else if (A.mode == LR_FIXED_TAU && !(S.tau[w] > 0.0)) {
      return;
    }
  }
  for (int k = i; k < 3 * nu; k += tm.N) L.rr[k] = wk.r_up[w * 3 * nu + k];
  for (int k = i; k < 3 * nd; k += tm.N) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
  for (int k = i; k < nu * nu; k += tm.N) L.E.Gi[k] = wk.Ginv[w * nu * nu + k];
  if (measure) {
    if (i < 9) L.RT[i] = A.RT ? A.RT[9 * w + i] : ((i % 4) == 0 ? 1.0 : 0.0);
  } else if (A.random_mesh) {
    double Rm[9];
    rotation_RT(A.seed, (uint32_t)w, (uint64_t)(A.step0 + it), STREAM_LRDMC, Rm);
    if (i == 0)
      for (int k = 0; k < 9; ++k) L.RT[k] = Rm[k];
  } else if (i < 9) {
    L.RT[i] = (i % 4) == 0 ? 1.0 : 0.0;
  }
  tm.sync();
  double* Hw = A.hws + w * (int64_t)ne * na;
  double* Qg = A.hws + W * (int64_t)ne * na + w * (int64_t)ne * no;  // q rows when they do not fit LDS
  const T* gw = s.j3_on ? wk.jfield + w * ne * na : nullptr;
  double diag, nondiag;
  int n_cand;
  lattice_hamiltonian<T, NW, kWide>(s, L, wk.pair + w * ne * no, Qg, Hw, gw, A.alat, A.dlt, tm, diag, nondiag, n_cand);
  if (measure) {
    if (i == 0) {
      A.V_diag[w] = diag;
      A.V_nondiag[w] = nondiag;
    }
    return;
  }
  double* cw = S.cand + w * S.n_cmax;
  for (int k = i; k < n_cand; k += tm.N) cw[k] = L.cand[k];
  const int n_ecp = s.nv > 0 ? (n_cand - 6 * ne) / s.nv : 0;
  for (int k = i; k < n_ecp; k += tm.N) S.elist[w * ne + k] = L.E.elist[k];
  if (i < 9) {
    S.RT[9 * w + i] = L.RT[i];
    if (A.RT) A.RT[9 * w + i] = L.RT[i];
  }
  if (i == 0) {
    S.diag[w] = diag;
    S.nondiag[w] = nondiag;
    S.ncand[w] = n_cand;
  }
}

// Part 2: the weight update (GFMC_n: w <- w b_x; GFMC_t: dtau = min(tau_left, ln(1 - xi) / sum offdiag),
// w <- w exp(-dtau e_L)), the move chosen by cumulative probabilities (move and waiting-time uniforms
// from sub-stream 2) and the single-electron update of r, caches, G and G^-1; the call's outputs of
// the walker are rewritten after each of its projections.
template <typename T>
__device__ void lr_move_walker(const Sys& s, const WalkT<T>& wk, const LrArgs& A, int64_t w, int it,
                               const LrMoveLds<T>& L, int lane) {
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  const LrState& S = A.st;
  if (s.j3_on && lane == 0) S.mv[w] = -1;
  if (it >= A.n_steps) return;
  double tau_left = S.tau[w];
  if (A.mode == LR_FIXED_TAU && !(tau_left > 0.0)) return;
  const double diag = S.diag[w], nondiag = S.nondiag[w];
  const int n_cand = S.ncand[w];
  double wgt = A.w ? A.w[w] : 1.0;
  const U4 d2 = draw_words(A.seed, (uint32_t)w, (uint64_t)(A.step0 + it), STREAM_LRDMC, 2);
  const double u = uniform01(d2.x, d2.y);
  const int64_t nproj = S.nproj[w] + 1;
  bool move = true;
  if (A.mode == LR_FIXED_STEPS) {
    // b_x = -(sum offdiag) / (sum diag - E_scf) (jqmc_gfmc.py:5047-5053)
    const double bx = 1.0 / (diag - A.E_scf) * (-nondiag);
    wgt = wgt * bx;
  } else {
    // dtau = min(tau_left, ln(1 - xi) / sum offdiag), w <- w exp(-dtau e_L) (jqmc_gfmc.py:988-1012)
    const double xi = uniform01(d2.z, d2.w);
    const double e_loc = diag + nondiag;
    const double t = log(1.0 - xi) / nondiag;
    const double dtau = t < tau_left ? t : tau_left;
    wgt = wgt * exp(-dtau * e_loc);
    tau_left = tau_left - dtau;
    move = tau_left > 0.0;
  }
  if (move) {
    const int k = choose_candidate(S.cand + w * S.n_cmax, n_cand, nondiag, u, lane);
    for (int q = lane; q < 3 * nu; q += WAVE) L.rr[q] = wk.r_up[w * 3 * nu + q];
    for (int q = lane; q < 3 * nd; q += WAVE) L.rr[3 * nu + q] = wk.r_dn[w * 3 * nd + q];
    for (int q = lane; q < nu * nu; q += WAVE) L.Gi[q] = wk.Ginv[w * nu * nu + q];
    if (lane < 9) L.RT[lane] = S.RT[9 * w + lane];
    wave_sync();
    int e;
    double nx, ny, nz;
    if (k < 6 * ne) {
      e = k / 6;
      const int dir = k - 6 * e, axis = dir >> 1;
      const double sa = (dir & 1) ? -A.alat : A.alat;
      nx = L.rr[3 * e] + sa * L.RT[3 * axis];
      ny = L.rr[3 * e + 1] + sa * L.RT[3 * axis + 1];
      nz = L.rr[3 * e + 2] + sa * L.RT[3 * axis + 2];
    } else {
      const int p = k - 6 * ne;
      e = S.elist[w * ne + p / s.nv];
      rotate_mesh(s, L.RT, L.gk, lane);
      const EcpElec E = ecp_electron(s, L.rr, e, L.Vt, lane);
      wave_sync();
      double wsum;
      ecp_point(s, E, L.Vt, L.rr, L.gk, e, p % s.nv, nx, ny, nz, wsum);
    }
    wave_sync();
    lrdmc_apply_move<T>(s, L, wk.ao + w * ne * na, wk.orb + w * ne * no, wk.pair + w * ne * no,
                        s.j3_on ? reinterpret_cast<T*>(S.dA) + w * na : nullptr, wk.G + w * nu * nu, e, nx, ny, nz,
                        lane);
    if (s.j3_on && lane == 0) S.mv[w] = e;
    for (int q = lane; q < 3 * nu; q += WAVE) wk.r_up[w * 3 * nu + q] = L.rr[q];
    for (int q = lane; q < 3 * nd; q += WAVE) wk.r_dn[w * 3 * nd + q] = L.rr[3 * nu + q];
    for (int q = lane; q < nu * nu; q += WAVE) wk.Ginv[w * nu * nu + q] = L.Gi[q];
  }
  if (lane == 0) {
    if (A.mode == LR_FIXED_TAU && tau_left > 0.0) atomicAdd(S.act + (it & 1), 1);
    if (A.w) A.w[w] = wgt;
    S.tau[w] = tau_left;
    S.nproj[w] = nproj;
    if (A.V_diag) A.V_diag[w] = diag;
    if (A.V_nondiag) A.V_nondiag[w] = nondiag;
    if (A.tau_left) A.tau_left[w] = tau_left;
    if (A.n_proj) A.n_proj[w] = nproj;
  }
}

// One wavefront per walker, wpb walkers per block.  kWide: the determinant fields of systems with
// more than EL_C_REGS orbitals (lr_wide).
template <typename T, bool kWide>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_lrdmc_ham(Sys s, int64_t W, WalkT<T> wk, LrArgs A, int it, int wpb,
                                                   int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  LrHamLds<T> L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  lr_ham_walker<T, 1, kWide>(s, wk, A, W, w, it, L, wave_team(lane));
}

// The Hamiltonian part with LR_TEAM waves per walker (one walker per block): for walker counts
// that leave the one-wave launch with few waves per SIMD, the walker's shells and candidate points
// are shared by the team.
constexpr int LR_TEAM = 4;

// Systems whose Hamiltonian launch takes the kWide determinant fields.
inline bool lr_wide(const Sys& s) { return s.orb_is_mo && s.n_orb > EL_C_REGS; }
template <typename T, bool kWide>
__global__ void __launch_bounds__(LR_TEAM * 64) k_lrdmc_ham_team(Sys s, int64_t W, WalkT<T> wk, LrArgs A, int it) {
  extern __shared__ double smem[];
  const int64_t w = blockIdx.x;
  if (w >= W) return;  // uniform across the block
  LrHamLds<T> L(reinterpret_cast<char*>(smem), s, LR_TEAM);
  const Team<LR_TEAM> tm{(int)threadIdx.x, wave_id(), lane_id(), L.E.red};
  lr_ham_walker<T, LR_TEAM, kWide>(s, wk, A, W, w, it, L, tm);
}

template <typename T>
__global__ void __launch_bounds__(256) k_lrdmc_move(Sys s, int64_t W, WalkT<T> wk, LrArgs A, int it, int wpb,
                                                    int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  LrMoveLds<T> L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  lr_move_walker<T>(s, wk, A, w, it, L, lane);
}

// Several systems of uneven size in one launch: block b belongs to the system k with
// block_start[k] <= b < block_start[k + 1] and carries one of its walkers (in the Hamiltonian launch
// with a team of NW waves, all systems of the launch alike); every system has its own device
// Hamiltonian, walker arrays and arguments (weights, draws, outputs, workspace).
struct LrSystem {
  const Sys* sys;
  int64_t W;
  int64_t lds_ham, lds_move;
  int64_t block_start;
};

constexpr int LR_MAX_SYSTEMS = 8;

// Launch table passed by value in the kernel arguments.
template <typename T>
struct LrMulti {
  int n_sys;
  LrSystem sys[LR_MAX_SYSTEMS];
  WalkT<T> wk[LR_MAX_SYSTEMS];
  LrArgs args[LR_MAX_SYSTEMS];
};

template <typename T>
__device__ __forceinline__ int lr_multi_system(const LrMulti<T>& M, int64_t b) {
  int k = 0;
  while (k + 1 < M.n_sys && M.sys[k + 1].block_start <= b) ++k;
  return k;
}

template <typename T, int NW, bool kWide>
__global__ void __launch_bounds__(NW * 64) k_lrdmc_ham_multi(LrMulti<T> M, int it) {
  extern __shared__ double smem[];
  const int64_t b = blockIdx.x;
  const int k = lr_multi_system(M, b);
  const LrSystem& Tk = M.sys[k];
  const int64_t w = b - Tk.block_start;
  if (w >= Tk.W) return;  // uniform across the block
  const Sys& s = *Tk.sys;
  LrHamLds<T> L(reinterpret_cast<char*>(smem), s, NW);
  const Team<NW> tm{(int)threadIdx.x, wave_id(), lane_id(), L.E.red};
  lr_ham_walker<T, NW, kWide>(s, M.wk[k], M.args[k], Tk.W, w, it, L, tm);
}

template <typename T>
__global__ void __launch_bounds__(64) k_lrdmc_move_multi(LrMulti<T> M, int it) {
  extern __shared__ double smem[];
  const int64_t b = blockIdx.x;
  const int k = lr_multi_system(M, b);
  const LrSystem& Tk = M.sys[k];
  const int64_t w = b - Tk.block_start;
  if (w >= Tk.W) return;
  const Sys& s = *Tk.sys;
  LrMoveLds<T> L(reinterpret_cast<char*>(smem), s);
  lr_move_walker<T>(s, M.wk[k], M.args[k], w, it, L, lane_id());
}

// ============================================================ branching

// Comb reconfiguration of one population (jqmc_gfmc.py:6070-6130): p_i = w_i / sum w,
// cum = cumsum(p) (sequential, as numpy), walker j takes parent searchsorted(cum, (j + zeta) / W)
// clamped to W - 1.  Also the per-step weighted sums of jqmc_gfmc.py:5957-5964
// (sum w, sum w/(V_diag - E), sum w e_L/(V_diag - E), sum w e_L^2/(V_diag - E)).
// One block of 1024 threads; cum is a W-double workspace.
constexpr int COMB_THREADS = 1024;
constexpr int COMB_TILE = 4096;

__global__ void __launch_bounds__(COMB_THREADS) k_comb(int64_t W, const double* __restrict__ w,
                                                       const double* __restrict__ Vd, const double* __restrict__ eL,
                                                       double E_scf, double zeta, int64_t* __restrict__ chosen,
                                                       double* __restrict__ sums, double* __restrict__ cum) {
  __shared__ double red[COMB_THREADS / WAVE][4];
  __shared__ double tile[COMB_TILE];
  __shared__ double total_sh;
  const int t = threadIdx.x, lane = lane_id(), wave = wave_id();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = t; i < W; i += COMB_THREADS) {
    const double wi = w[i];
    acc[0] += wi;
    if (Vd && eL) {
      const double f = wi / (Vd[i] - E_scf);
      acc[1] += f;
      acc[2] += f * eL[i];
      acc[3] += f * (eL[i] * eL[i]);
    }
  }
  wave_sum_k<4>(acc);
  if (lane == 0)
    for (int k = 0; k < 4; ++k) red[wave][k] = acc[k];
  __syncthreads();
  if (t == 0) {
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < COMB_THREADS / WAVE; ++v)
      for (int k = 0; k < 4; ++k) tot[k] += red[v][k];
    total_sh = tot[0];
    if (sums)
      for (int k = 0; k < 4; ++k) sums[k] = tot[k];
  }
  __syncthreads();
  const double total = total_sh;
  // sequential cumulative sum, staged through LDS one tile at a time
  double running = 0.0;
  for (int64_t b0 = 0; b0 < W; b0 += COMB_TILE) {
    const int n = (int)((W - b0) < COMB_TILE ? (W - b0) : COMB_TILE);
    for (int i = t; i < n; i += COMB_THREADS) tile[i] = w[b0 + i] / total;
    __syncthreads();
    if (t == 0) {
      double r = running;
      for (int i = 0; i < n; ++i) {
        r += tile[i];
        tile[i] = r;
      }
      running = r;
    }
    __syncthreads();
    for (int i = t; i < n; i += COMB_THREADS) cum[b0 + i] = tile[i];
    __syncthreads();
  }
  __threadfence_block();
  __syncthreads();
  for (int64_t j = t; j < W; j += COMB_THREADS) {
    const double z = ((double)j + zeta) / (double)W;
    int64_t lo = 0, hi = W;  // first index with cum >= z
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo) >> 1);
      if (cum[mid] < z) lo = mid + 1;
      else hi = mid;
    }
    chosen[j] = lo < W ? lo : W - 1;
  }
}

// Per-walker rows of the walker state that reconfiguration copies from parent to child: every
// array of jqmc_walkers with a leading walker axis, with its row size in bytes and its offset in the
// staging buffer (rows of all arrays of one walker are staged together).
constexpr int GATHER_MAX_ARRAYS = 12;
struct StateRows {
  int n;
  char* base[GATHER_MAX_ARRAYS];
  int64_t bytes[GATHER_MAX_ARRAYS];  // row size, a multiple of 4
  int64_t toff[GATHER_MAX_ARRAYS];   // offset of the row inside one walker's staging record
  int64_t rec;                       // bytes of one walker's staging record (multiple of 16)
};

// Rows of walker j whose parent c = chosen[j] differs from j: dir 0 copies the parent's rows into
// j's staging record, dir 1 copies the staging record back into j's rows.  Two launches in stream
// order, so a parent's rows are read before any child overwrites them.  One block per walker, 4-byte
// words (every row size and base is 4-byte aligned), consecutive threads on consecutive words.
__global__ void __launch_bounds__(256) k_state_gather(int64_t W, StateRows R, const int64_t* __restrict__ chosen,
                                                      char* __restrict__ stage, int dir) {
  const int64_t j = blockIdx.x;
  if (j >= W) return;
  int64_t c = chosen[j];
  c = c < 0 ? 0 : (c >= W ? W - 1 : c);
  if (c == j) return;
  char* rec = stage + j * R.rec;
  for (int k = 0; k < R.n; ++k) {
    const int64_t nw = R.bytes[k] >> 2;
    uint32_t* row = reinterpret_cast<uint32_t*>(R.base[k] + (dir == 0 ? c : j) * R.bytes[k]);
    uint32_t* st = reinterpret_cast<uint32_t*>(rec + R.toff[k]);
    if (dir == 0) {
      for (int64_t t = threadIdx.x; t < nw; t += blockDim.x) st[t] = row[t];
    } else {
      for (int64_t t = threadIdx.x; t < nw; t += blockDim.x) row[t] = st[t];
    }
  }
}

// dst[j] = src[chosen[j]] for per-walker blocks of `per` doubles (parent indices clamped to [0, W)).
__global__ void k_gather_rows(int64_t W, int64_t per, const int64_t* __restrict__ chosen,
                              const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * per) return;
  const int64_t j = idx / per, k = idx - j * per;
  int64_t c = chosen[j];
  c = c < 0 ? 0 : (c >= W ? W - 1 : c);
  dst[idx] = src[c * per + k];
}

}  // namespace jqmc
