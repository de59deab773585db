// gfx950 kernels of the LRDMC (lattice-regularised diffusion Monte Carlo) path:
//
//   k_lrdmc        fused projection loop, one wavefront per walker: lattice Hamiltonian
//                  elements at the current configuration, weight update, choice of one
//                  off-diagonal move and its rank-1 state update, repeated n_steps times
//                  (GFMC_n) or until the projection time is used up (GFMC_t); also the
//                  measurement of V_diag / V_nondiag at a given rotation.
//   k_comb         comb reconfiguration of the walker population (one block).
//   k_gather_rows  walker positions copied from the chosen parents.
//
// Walker caches are the ones the VMC sweep keeps (AO values, orbitals, pair fields, J3 fields,
// G, G^-1); the determinant field h_e of the electron being processed lives in LDS.  Like the
// other kernels, k_lrdmc is a template over the cache / AO arithmetic type T; the weights, the
// Hamiltonian sums, G and G^-1 are FP64 in both forms.
#pragma once

#include "jqmc_kernels.hip"

namespace jqmc {

// Modes of k_lrdmc.
constexpr int LR_FIXED_STEPS = 0;  // GFMC_n: n_steps projections, w <- w b_x
constexpr int LR_FIXED_TAU = 1;    // GFMC_t: projections until tau is used, w <- w exp(-dtau e_L)
constexpr int LR_MEASURE = 2;      // V_diag / V_nondiag at the given rotation, no move

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
};

// LDS layout of one wavefront.
template <typename T>
struct LrLds {
  double *rr, *Gi, *RT, *gk, *te, *pts, *pw, *det, *dj, *Vt, *cand, *vv, *xx, *yy, *q, *h, *phin, *dphi;
  T *anew, *dA, *rad;
  int* plist;
  uintptr_t end;
  __host__ __device__ LrLds(char* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up, na = s.n_ao, no = s.n_orb, nv = s.nv > 0 ? s.nv : 1, P = 6 + nv;
    Carve c{base};
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(nu * nu);
    RT = c.take<double>(9);
    gk = c.take<double>(3 * nv);
    te = c.take<double>(ne);
    pts = c.take<double>(3 * P);
    pw = c.take<double>(P);
    det = c.take<double>(P);
    dj = c.take<double>(P);
    Vt = c.take<double>(s.ecp_max_terms > 0 ? s.ecp_max_terms : 1);
    cand = c.take<double>(6 * ne + ne * nv);
    vv = c.take<double>(nu);
    xx = c.take<double>(nu);
    yy = c.take<double>(nu);
    q = c.take<double>(no);
    h = c.take<double>(na);
    anew = c.take<T>(na);
    dA = c.take<T>(na);
    phin = c.take<double>(no);
    dphi = c.take<double>(no);
    rad = c.take<T>((int64_t)s.n_shell * PTS_CHUNK);
    plist = c.take<int>(ne);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

template <typename T>
inline int64_t lr_lds_bytes(const Sys& s) {
  return (int64_t)LrLds<T>(0, s).end;
}

// Diagonal and off-diagonal sums of the lattice-regularised Hamiltonian at the current
// configuration (_body_step_core, jqmc_gfmc.py:4807-5045; _compute_V_elements_n, 5360-5628).
// Electron by electron: continuum kinetic energy, the 6 lattice moves r_e + s_k with
// s = a [+x, -x, +y, -y, +z, -z] R^T (wavefunction.py:1779-1860) and, when e's nearest nucleus
// carries projected ECP terms, its nv quadrature points.  Leaves in L.cand the candidate list of
// the move choice: min(element, 0) of the lattice moves of every electron (up block first), then
// the fixed-node ECP elements of the active electrons' quadrature points (L.plist lists those
// electrons in order); *n_cand is its length.
template <typename T>
__device__ void lattice_hamiltonian(const Sys& s, const LrLds<T>& L, const T* Aw, const double* Qw, const T* gw,
                                    double alat, int dlt, int lane, double& diag, double& nondiag, int& n_cand) {
  const int ne = s.n_el, na = s.n_ao;
  const int n_lat = 6 * ne;
  const bool nonlocal = s.ecp_on && s.nv > 0;
  const double kfac = -1.0 / (2.0 * (alat * alat));
  if (nonlocal) rotate_mesh(s, L.RT, L.gk, lane);
  double sp_fn[2] = {0.0, 0.0};  // ECP: sum max(V, 0), sum fixed-node elements
  int rank = 0;
  for (int e = 0; e < ne; ++e) {
    const double Te = electron_fields<T>(s, L.rr, L.Gi, L.q, L.h, Qw, gw, e, lane);
    if (lane == 0) L.te[e] = Te;
    EcpElec E{0, 0, 0.0, 0.0, 0.0, 1.0};
    if (nonlocal) E = ecp_electron(s, L.rr, e, L.Vt, lane);
    const bool act = nonlocal && E.nt > 0;
    const int P = 6 + (act ? s.nv : 0);
    wave_sync();
    for (int q = lane; q < P; q += WAVE) {
      if (q < 6) {
        const int axis = q >> 1;
        const double sa = (q & 1) ? -alat : alat;
        L.pts[3 * q] = L.rr[3 * e] + sa * L.RT[3 * axis];
        L.pts[3 * q + 1] = L.rr[3 * e + 1] + sa * L.RT[3 * axis + 1];
        L.pts[3 * q + 2] = L.rr[3 * e + 2] + sa * L.RT[3 * axis + 2];
      } else {
        ecp_point(s, E, L.Vt, L.rr, L.gk, e, q - 6, L.pts[3 * q], L.pts[3 * q + 1], L.pts[3 * q + 2], L.pw[q]);
      }
    }
    wave_sync();
    electron_points_ratio<T>(s, L.rr, e, L.h, Aw + (int64_t)e * na, gw ? gw + (int64_t)e * na : nullptr, L.pts, P,
                             L.det, L.dj, L.rad, lane);
    for (int q = lane; q < P; q += WAVE) {
      const double jr = exp(L.dj[q]);
      if (q < 6) {
        L.cand[6 * e + q] = kfac * (L.det[q] * jr);
      } else {
        // t-move: V = w_P Psi(r')/Psi(r); DLA t-move: V = w_P det ratio and the Jastrow ratio multiplies
        // the fixed-node part only (jqmc_gfmc.py:4956-5021)
        const double V = dlt ? L.pw[q] * L.det[q] : L.pw[q] * (L.det[q] * jr);
        double fn = fmin(V, 0.0);
        if (dlt) fn = fn * jr;
        sp_fn[0] += fmax(V, 0.0);
        sp_fn[1] += fn;
        L.cand[n_lat + rank * s.nv + (q - 6)] = fn;
      }
    }
    if (act) {
      if (lane == 0) L.plist[rank] = e;
      ++rank;
    }
    wave_sync();
  }

  // ---- per electron: sign-flip flag, regularised electron-ion diagonal (jqmc_gfmc.py:4833-4939)
  const double quarter = 1.0 / (4.0 * (alat * alat));
  double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, sp_fn[0], sp_fn[1]};  // sum opt, SP, FN, ECP local, el-el, ECP SP, ECP FN
  for (int e = lane; e < ne; e += WAVE) {
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
    const double zv = (v_ei + L.te[e]) - nd;
    const double ei = s.ecp_on ? v_ei : v_fl;
    acc[0] += flag ? fmax(zv, ei) : zv;
  }
  if (s.ecp_on) {
    // ECP local part (coulomb_potential.py:1144-1246), lanes over electron-nucleus pairs
    for (int t = lane; t < ne * s.n_nuc; t += WAVE) {
      const int e = t / s.n_nuc, n = t - e * s.n_nuc;
      const double dx = L.rr[3 * e] - s.nuc_pos[3 * n], dy = L.rr[3 * e + 1] - s.nuc_pos[3 * n + 1],
                   dz = L.rr[3 * e + 2] - s.nuc_pos[3 * n + 2];
      acc[3] += ecp_local_pair(s, n, sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  acc[4] = el_el_partial(s, L.rr, lane);
  wave_sum_k<7>(acc);
  diag = 3.0 / (2.0 * (alat * alat)) * ne + ((acc[4] + s.ion_ion) + acc[0]);
  nondiag = acc[2];
  n_cand = n_lat + rank * (s.nv > 0 ? s.nv : 0);
  if (s.ecp_on) {
    diag = ((diag + acc[3]) + acc[1]) + acc[5];
    nondiag += acc[6];
  } else {
    diag += acc[1];
  }
  wave_sync();
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

// Electron e moves to (nx, ny, nz): AO, orbital and pair rows, J3 fields of the other electrons,
// G and G^-1 by Sherman-Morrison (_update_inv_up_n / _update_inv_dn_n, jqmc_gfmc.py:5083-5141).
template <typename T>
__device__ void lrdmc_apply_move(const Sys& s, const LrLds<T>& L, T* Aw, double* Pw, double* Qw, T* gw, double* Gw, int e,
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
  if (gw) {
    // J3 fields: electrons before e gain J3 dA, electrons after e gain J3^T dA (jastrow_factor.py:4121-4135)
    for (int a = lane; a < na; a += WAVE) {
      T t1 = 0, t2 = 0;
      if (s.j3_sym) {
        for (int b = 0; b < na; ++b) t2 += R.j3[(int64_t)b * na + a] * L.dA[b];
        t1 = t2;
      } else {
        for (int b = 0; b < na; ++b) {
          const T db = L.dA[b];
          t1 += R.j3T[(int64_t)b * na + a] * db;
          t2 += R.j3[(int64_t)b * na + a] * db;
        }
      }
      for (int f = 0; f < ne; ++f) {
        if (f == e) continue;
        gw[(int64_t)f * na + a] += f < e ? t1 : t2;
      }
    }
  }
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

// LRDMC projection loop (GFMC_n _projection_n, jqmc_gfmc.py:4738-5358; GFMC_t
// _projection_t_core, 724-1107), one wavefront per walker, no block-wide barriers.
// Draws of projection `step` come from the LRDMC Philox stream: rotation angles (sub-streams
// 0-1), move uniform and waiting-time uniform (sub-stream 2).
template <typename T>
__device__ void lrdmc_walker(const Sys& s, const WalkT<T>& wk, const LrArgs& A, int64_t w, const LrLds<T>& L,
                             int lane) {
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  T* Aw = wk.ao + w * ne * na;
  double* Pw = wk.orb + w * ne * no;
  double* Qw = wk.pair + w * ne * no;
  T* gw = s.j3_on ? wk.jfield + w * ne * na : nullptr;
  double* Gw = wk.G + w * nu * nu;
  for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = wk.r_up[w * 3 * nu + k];
  for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
  for (int k = lane; k < nu * nu; k += WAVE) L.Gi[k] = wk.Ginv[w * nu * nu + k];
  wave_sync();

  const bool measure = A.mode == LR_MEASURE;
  double wgt = (!measure && A.w) ? A.w[w] : 1.0;
  double tau_left = A.tau;
  double diag = 0.0, nondiag = 0.0;
  int64_t nproj = 0;
  bool moved_any = false;
  const int cap = measure ? 1 : A.n_steps;
  for (int it = 0; it < cap; ++it) {
    if (A.mode == LR_FIXED_TAU && !(tau_left > 0.0)) break;
    const uint64_t step = (uint64_t)(A.step0 + it);
    if (measure) {
      if (lane < 9) L.RT[lane] = A.RT ? A.RT[9 * w + lane] : ((lane % 4) == 0 ? 1.0 : 0.0);
    } else if (A.random_mesh) {
      double Rm[9];
      rotation_RT(A.seed, (uint32_t)w, step, STREAM_LRDMC, Rm);
      if (lane == 0)
        for (int k = 0; k < 9; ++k) L.RT[k] = Rm[k];
    } else if (lane < 9) {
      L.RT[lane] = (lane % 4) == 0 ? 1.0 : 0.0;
    }
    wave_sync();
    int n_cand = 0;
    lattice_hamiltonian<T>(s, L, Aw, Qw, gw, A.alat, A.dlt, lane, diag, nondiag, n_cand);
    if (measure) break;
    const U4 d2 = draw_words(A.seed, (uint32_t)w, step, STREAM_LRDMC, 2);
    const double u = uniform01(d2.x, d2.y);
    ++nproj;
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
    const int k = choose_candidate(L.cand, n_cand, nondiag, u, lane);
    if (move) {
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
        e = L.plist[p / s.nv];
        const EcpElec E = ecp_electron(s, L.rr, e, L.Vt, lane);
        wave_sync();
        double wsum;
        ecp_point(s, E, L.Vt, L.rr, L.gk, e, p % s.nv, nx, ny, nz, wsum);
      }
      wave_sync();
      lrdmc_apply_move<T>(s, L, Aw, Pw, Qw, gw, Gw, e, nx, ny, nz, lane);
      moved_any = true;
    }
  }

  if (moved_any) {
    for (int k = lane; k < 3 * nu; k += WAVE) wk.r_up[w * 3 * nu + k] = L.rr[k];
    for (int k = lane; k < 3 * nd; k += WAVE) wk.r_dn[w * 3 * nd + k] = L.rr[3 * nu + k];
    for (int k = lane; k < nu * nu; k += WAVE) wk.Ginv[w * nu * nu + k] = L.Gi[k];
  }
  if (!measure && A.RT && lane < 9) A.RT[9 * w + lane] = L.RT[lane];
  if (lane == 0) {
    if (!measure && A.w) A.w[w] = wgt;
    if (A.V_diag) A.V_diag[w] = diag;
    if (A.V_nondiag) A.V_nondiag[w] = nondiag;
    if (A.tau_left) A.tau_left[w] = tau_left;
    if (A.n_proj) A.n_proj[w] = nproj;
  }
}

// Registers capped at 256 (2 waves per SIMD, ~0.5 KB of spills per lane): latency bound like the
// E_L kernel, it runs faster with two waves per SIMD than with one unspilled wave.
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_lrdmc(Sys s, int64_t W, WalkT<T> wk, LrArgs A, int wpb,
                                               int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  LrLds<T> L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  lrdmc_walker<T>(s, wk, A, w, L, lane);
}

// Several systems of uneven size in one launch: block b belongs to the system k with
// block_start[k] <= b < block_start[k + 1] and carries wpb[k] of its walkers; every system has
// its own device Hamiltonian, walker arrays and arguments (weights, draws, outputs).
struct LrSystem {
  const Sys* sys;
  int64_t W;
  int wpb;
  int64_t lds_per_wave;
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
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_lrdmc_multi(LrMulti<T> M) {
  extern __shared__ double smem[];
  const int64_t b = blockIdx.x;
  int k = 0;
  while (k + 1 < M.n_sys && M.sys[k + 1].block_start <= b) ++k;
  const LrSystem& Tk = M.sys[k];
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  if (wave >= Tk.wpb) return;
  const int64_t w = (b - Tk.block_start) * Tk.wpb + wave;
  if (w >= Tk.W) return;
  const Sys& s = *Tk.sys;
  LrLds<T> L(reinterpret_cast<char*>(smem) + wave * Tk.lds_per_wave, s);
  lrdmc_walker<T>(s, M.wk[k], M.args[k], w, L, lane);
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
  const int t = threadIdx.x, lane = lane_id(), wave = t / WAVE;
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
