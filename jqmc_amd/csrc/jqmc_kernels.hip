// gfx950 kernels of the jQMC per-walker hot path.
//
// Layout in HBM: every per-walker array is walker-major and contiguous per walker
// (W, ...), per-electron vectors are electron-major with the AO / orbital index
// innermost, so the lanes of a wavefront that own consecutive AO indices read and
// write contiguous segments.  A lane owns AO/orbital index k = lane + 64 j for every
// per-electron vector it touches; cross-lane traffic goes through LDS.
//
// Every kernel that touches the streamed caches is a template over T, the arithmetic and
// storage type of the AO / orbital / pair / J3 work: double for the plain FP64 form, float for
// the mixed form.  G, G^-1, positions, log-determinants and every energy or weight sum stay FP64.
//
// Kernels:
//   k_aos_vgl          AO value/grad/lap at arbitrary points (thread per AO x point)
//   k_orbs_contract    orbital = C . AO for the five AO outputs (thread per orbital x point)
//   k_init_*           walker-state construction (G, LU inverse, pair and J3 fields)
//   k_vmc_sweep        fused Metropolis loop, one wavefront per walker
//   k_rotations        per-walker quadrature rotations
#pragma once

#include "jqmc_device.hpp"

namespace jqmc {

// LDS carving: consecutive arrays, each 8-byte aligned.  The same walk runs on the host with a
// zero base to size the per-wavefront region.
struct Carve {
  // Pointer arithmetic only (no integer round trip), so that pointers carved from a __shared__
  // base keep the LDS address space and compile to ds_* instructions rather than flat ones.
  char* p;
  template <typename U>
  __host__ __device__ U* take(int64_t n) {
    U* r = reinterpret_cast<U*>(p);
    p += ((n * (int64_t)sizeof(U)) + 7) & ~(int64_t)7;
    return r;
  }
};

// y[k] = sum_a CT[a, k] x[a] for k < no, all 64 lanes busy: lanes split into groups of
// lg >= no lanes, group g sums a = g, g + G, ...; groups combine by butterfly shuffles.  The sums
// run in TY (FP64 for the orbital rows, whatever the type of the AO values x).
template <typename TC, typename TX, typename TY>
__device__ inline void wave_matvec_T(const TC* __restrict__ CT, const TX* x, int na, int no, int lane, TY* y) {
  int lg = 64;
  while (lg / 2 >= no && lg > 1) lg >>= 1;
  const int G = WAVE / lg, k = lane % lg, grp = lane / lg;
  for (int kb = 0; kb < no; kb += lg) {
    const int kk = kb + k;
    TY acc = 0;
    if (kk < no) {
#pragma unroll 4
      for (int a = grp; a < na; a += G) acc += (TY)CT[(int64_t)a * no + kk] * (TY)x[a];
    }
    for (int off = lg; off < WAVE; off <<= 1) acc += __shfl_xor(acc, off, WAVE);
    if (grp == 0 && kk < no) y[kk] = acc;
  }
}

// ============================================================ standalone evaluators

__global__ void k_aos_vgl(Sys s, int64_t n_pts, const double* __restrict__ r, double* __restrict__ val,
                          double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz,
                          double* __restrict__ lap) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)s.n_ao * n_pts) return;
  const int a = (int)(idx / n_pts);
  const int64_t p = idx - (int64_t)a * n_pts;
  const double x = r[3 * p], y = r[3 * p + 1], z = r[3 * p + 2];
  if (gx == nullptr) {
    val[idx] = ao_value<double>(s, a, x, y, z);
    return;
  }
  double v, g0, g1, g2, l;
  ao_vgl<double>(s, a, x, y, z, v, g0, g1, g2, l);
  val[idx] = v;
  gx[idx] = g0;
  gy[idx] = g1;
  gz[idx] = g2;
  lap[idx] = l;
}

// out_k[m, p] = sum_a C[m, a] in_k[a, p] for k < n_out (molecular_orbital.py:405-413).
__global__ void k_orbs_contract(int n_orb, int n_ao, int64_t n_pts, const double* __restrict__ C,
                                const double* __restrict__ in, double* o0, double* o1, double* o2, double* o3,
                                double* o4, int n_out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_orb * n_pts) return;
  const int m = (int)(idx / n_pts);
  const int64_t p = idx - (int64_t)m * n_pts;
  double* outs[5] = {o0, o1, o2, o3, o4};
  const int64_t plane = (int64_t)n_ao * n_pts;
  for (int k = 0; k < n_out; ++k) {
    double acc = 0.0;
    const double* src = in + k * plane + p;
    for (int a = 0; a < n_ao; ++a) acc += C[(int64_t)m * n_ao + a] * src[(int64_t)a * n_pts];
    outs[k][idx] = acc;
  }
}

__global__ void k_rotations(int64_t W, uint64_t seed, int64_t step, const int64_t* __restrict__ step_ctr,
                            uint32_t stream, double* __restrict__ RT) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  double R[9];
  rotation_RT(seed, (uint32_t)w, (uint64_t)(step + (step_ctr ? *step_ctr : 0)), stream, R);
  for (int k = 0; k < 9; ++k) RT[9 * w + k] = R[k];
}

// ctr[0] += inc (one thread): advances a device-resident move or step counter, so that a
// captured graph of a step draws fresh random numbers on every replay.
__global__ void k_counter_add(int64_t* ctr, int64_t inc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr[0] += inc;
}

// ============================================================ walker initialisation

template <typename T>
__device__ __forceinline__ const double* elec_pos(const Sys& s, const WalkT<T>& wk, int64_t w, int e) {
  return e < s.n_up ? wk.r_up + (w * s.n_up + e) * 3 : wk.r_dn + (w * s.n_dn + (e - s.n_up)) * 3;
}

// Per (walker, electron) wavefront: AO values at r_e in FP64 (shell radial factors once per shell),
// stored in the AO cache in T; the orbital row phi_e = C_spin a_e (or the AOs themselves) and the
// pair row Lambda_p^T phi_e (up) or Lambda_p phi_e (dn) from the FP64 values, in FP64
// (compute_MOs, molecular_orbital.py:239-261; the pre-contraction of determinant.py:1616-1646).
// LDS per wavefront: n_ao + n_orb doubles.
template <typename T>
__global__ void __launch_bounds__(256) k_init_rows(Sys s, int64_t W, WalkT<T> wk, int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t we = (int64_t)blockIdx.x * 4 + wave;
  if (we >= W * s.n_el) return;
  const int64_t w = we / s.n_el;
  const int e = (int)(we - w * s.n_el);
  const int na = s.n_ao, no = s.n_orb;
  Carve c{reinterpret_cast<char*>(smem) + wave * lds_per_wave};
  double* av = c.take<double>(na);
  double* ph = c.take<double>(no);
  const double* r = elec_pos(s, wk, w, e);
  const double x = r[0], y = r[1], z = r[2];
  for (int sh = lane; sh < s.n_shell; sh += WAVE) shell_values<double>(s, sh, x, y, z, av);
  wave_sync();
  T* Ae = wk.ao + we * na;
  for (int a = lane; a < na; a += WAVE) Ae[a] = (T)av[a];
  double* Pe = wk.orb + we * no;
  const bool up = e < s.n_up;
  if (s.orb_is_mo) {
    wave_matvec_T<double, double, double>(up ? s.td.mo_upT : s.td.mo_dnT, av, na, no, lane, ph);
  } else {
    for (int k = lane; k < no; k += WAVE) ph[k] = av[k];
  }
  wave_sync();
  const double* Lm = up ? s.td.lam_p : s.td.lam_pT;  // column k of Lambda_p (up) or of its transpose
  double* Qe = wk.pair + we * no;
  for (int k = lane; k < no; k += WAVE) {
    double q = 0.0;
    for (int j = 0; j < no; ++j) q += Lm[(int64_t)j * no + k] * ph[j];
    Pe[k] = ph[k];
    Qe[k] = q;
  }
}

inline int64_t init_rows_lds_bytes(const Sys& s) {
  Carve c{nullptr};
  c.take<double>(s.n_ao);
  c.take<double>(s.n_orb);
  return (int64_t)reinterpret_cast<uintptr_t>(c.p);
}

// G[w, i, j] = phi_i . pair_dn_j (j < n_dn), phi_i . Lambda_u[:, j - n_dn] otherwise
// (_compute_geminal_all_elements, determinant.py:1379-1402), summed in FP64.
template <typename T>
__global__ void k_init_geminal(Sys s, int64_t W, WalkT<T> wk) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = s.n_up;
  const int64_t per_w = (int64_t)nu * nu;
  if (idx >= W * per_w) return;
  const int64_t w = idx / per_w;
  const int rem = (int)(idx - w * per_w);
  const int i = rem / nu, j = rem - i * nu;
  const double* phi = wk.orb + (w * s.n_el + i) * s.n_orb;
  double acc = 0.0;
  if (j < s.n_dn) {
    const double* pr = wk.pair + (w * s.n_el + nu + j) * s.n_orb;
    for (int m = 0; m < s.n_orb; ++m) acc += phi[m] * pr[m];
  } else {
    const int k = j - s.n_dn;
    for (int m = 0; m < s.n_orb; ++m) acc += phi[m] * s.td.lam_u[(int64_t)m * s.n_unp + k];
  }
  wk.G[idx] = acc;
}

// Gauss-Jordan elimination with partial pivoting in FP64 by one wavefront: M (n x n, a copy of G)
// is reduced to the identity while I (the identity on entry) becomes G^-1; ln|det G| and its sign
// come from the pivots.  Pivot = largest |entry| in the column, lowest row on ties (LAPACK getrf rule
// used by jnp.linalg.det, determinant.py:1203).  M and I may be in LDS or global memory.
__device__ inline void wave_gauss_jordan(double* M, double* I, int n, int lane, double& logabs_out, double& sgn_out) {
  double logabs = 0.0, sgn = 1.0;
  for (int k = 0; k < n; ++k) {
    // pivot search: lane i holds |M[i][k]| for i >= k
    double best = -1.0;
    int bi = n;
    for (int i = k + lane; i < n; i += WAVE) {
      const double v = fabs(M[i * n + k]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, WAVE);
      const int oi = __shfl_xor(bi, off, WAVE);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (bi != k) {
      for (int c = lane; c < n; c += WAVE) {
        double t = M[k * n + c];
        M[k * n + c] = M[bi * n + c];
        M[bi * n + c] = t;
        t = I[k * n + c];
        I[k * n + c] = I[bi * n + c];
        I[bi * n + c] = t;
      }
      sgn = -sgn;
    }
    wave_sync();
    const double piv = M[k * n + k];
    logabs += log(fabs(piv));
    if (piv < 0.0) sgn = -sgn;
    const double inv = 1.0 / piv;
    for (int c = lane; c < n; c += WAVE) {
      M[k * n + c] *= inv;
      I[k * n + c] *= inv;
    }
    wave_sync();
    // eliminate column k from every other row
    for (int t = lane; t < n * n; t += WAVE) {
      const int i = t / n, c = t - i * n;
      if (i == k) continue;
      const double f = M[i * n + k];
      if (c != k) M[i * n + c] -= f * M[k * n + c];
      I[i * n + c] -= f * I[k * n + c];
    }
    wave_sync();
    for (int i = lane; i < n; i += WAVE)
      if (i != k) M[i * n + k] = 0.0;
    wave_sync();
  }
  logabs_out = logabs;
  sgn_out = sgn;
}

// Gauss-Jordan inverse with partial pivoting in FP64, one wavefront per walker (n_up <= 64), and
// the AS factor (determinant.py:1223-1260).
__global__ void k_init_inverse(const double* __restrict__ Gall, double* __restrict__ Ginv_all, double* logabsdet,
                               double* sign, double* as_factor, int n, int64_t W, int wpb) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  double* M = smem + (size_t)wave * 2 * n * n;
  double* I = M + n * n;
  const double* G = Gall + w * n * n;
  for (int k = lane; k < n * n; k += WAVE) {
    M[k] = G[k];
    I[k] = (k / n == k % n) ? 1.0 : 0.0;
  }
  wave_sync();
  double logabs, sgn;
  wave_gauss_jordan(M, I, n, lane, logabs, sgn);
  double* Ginv = Ginv_all + w * n * n;
  double fro = 0.0;
  for (int k = lane; k < n * n; k += WAVE) {
    Ginv[k] = I[k];
    fro += I[k] * I[k];
  }
  fro = wave_sum(fro);
  // S = min(min row norm^2, min column norm^2) of G
  double smin = 1.0e300;
  for (int i = lane; i < n; i += WAVE) {
    double rs = 0.0, cs = 0.0;
    for (int j = 0; j < n; ++j) {
      rs += G[i * n + j] * G[i * n + j];
      cs += G[j * n + i] * G[j * n + i];
    }
    smin = fmin(smin, fmin(rs, cs));
  }
  smin = wave_min(smin);
  if (lane == 0) {
    if (logabsdet) logabsdet[w] = logabs;
    if (sign) sign[w] = sgn;
    if (as_factor) {
      const double sf = smin * fro;
      as_factor[w] = sf > 0.0 ? pow(sf, -3.0 / 8.0) : 0.0;
    }
  }
}

// Prefix sums of AO columns feeding the J3 fields: for electron e,
// S1[e] = sum of a_j over j in S1(e), S2[e] likewise, with
//   up e:  S1 = {up j > e} U {dn},  S2 = {up j < e}
//   dn e:  S1 = {dn j > e},         S2 = {dn j < e} U {up}
// so that g_e = j1 + J3 S1[e] + J3^T S2[e] (jastrow_factor.py:4121-4135).  Sums in FP64.
template <typename T>
__global__ void k_init_jsums(Sys s, int64_t W, const WalkT<T> wk, T* __restrict__ S1, T* __restrict__ S2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * s.n_ao) return;
  const int64_t w = idx / s.n_ao;
  const int a = (int)(idx - w * s.n_ao);
  const int nu = s.n_up, ne = s.n_el, na = s.n_ao;
  const T* A = wk.ao + w * ne * na;
  double tot_up = 0.0, tot_dn = 0.0;
  for (int e = 0; e < nu; ++e) tot_up += (double)A[e * na + a];
  for (int e = nu; e < ne; ++e) tot_dn += (double)A[e * na + a];
  double before = 0.0;
  for (int e = 0; e < nu; ++e) {
    const double ae = A[e * na + a];
    S2[(w * ne + e) * na + a] = (T)before;
    S1[(w * ne + e) * na + a] = (T)((tot_up - before - ae) + tot_dn);
    before += ae;
  }
  before = 0.0;
  for (int e = nu; e < ne; ++e) {
    const double ae = A[e * na + a];
    S2[(w * ne + e) * na + a] = (T)(before + tot_up);
    S1[(w * ne + e) * na + a] = (T)(tot_dn - before - ae);
    before += ae;
  }
}

// g[w, e, a] = j1[a] + sum_b J[a, b] S1[e, b] + sum_b J[b, a] S2[e, b].
template <typename T>
__global__ void k_init_jfield(Sys s, int64_t W, WalkT<T> wk, const T* __restrict__ S1, const T* __restrict__ S2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int na = s.n_ao;
  const int64_t per_w = (int64_t)s.n_el * na;
  if (idx >= W * per_w) return;
  const int64_t we = idx / na;
  const int a = (int)(idx - we * na);
  const T* s1 = S1 + we * na;
  const T* s2 = S2 + we * na;
  const RealTabs<T>& R = tabs<T>(s);
  T acc = R.j3v[a];
  for (int b = 0; b < na; ++b) acc += R.j3T[(int64_t)b * na + a] * s1[b] + R.j3[(int64_t)b * na + a] * s2[b];
  wk.jfield[idx] = acc;
}

// J = J1 + J2 + J3 with J3 = 1/2 sum_e a_e . (g_e + j1), one wavefront per walker, FP64 sums.
template <typename T>
__global__ void k_init_jastrow(Sys s, int64_t W, WalkT<T> wk, int wpb) {
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  double acc = 0.0;
  const int ne = s.n_el;
  for (int e = lane; e < ne; e += WAVE) {
    const double* r = elec_pos(s, wk, w, e);
    if (s.j1_type) acc += j1_electron(s, r[0], r[1], r[2]);
    if (s.j2_type) {
      for (int f = e + 1; f < ne; ++f) {
        const double* q = elec_pos(s, wk, w, f);
        const double dx = r[0] - q[0], dy = r[1] - q[1], dz = r[2] - q[2];
        acc += j2_u(s.j2_type, s.j2_param, sqrt(dx * dx + dy * dy + dz * dz));
      }
    }
  }
  if (s.j3_on) {
    const int na = s.n_ao;
    const T* jv = tabs<T>(s).j3v;
    for (int t = lane; t < ne * na; t += WAVE) {
      const int a = t % na;
      acc += 0.5 * (double)wk.ao[w * ne * na + t] * ((double)wk.jfield[w * ne * na + t] + (double)jv[a]);
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) wk.jastrow[w] = acc;
}

// ============================================================ VMC sweep

// LDS layout of one wavefront in the sweep.
template <typename T>
struct SweepLds {
  double *rr, *Gi, *Gm, *vv, *xx, *yy, *phin, *dphi;
  T *anew, *dA, *t1, *t2, *Oup, *Odn;
  uintptr_t end;
  __host__ __device__ SweepLds(char* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up, na = s.n_ao, no = s.n_orb;
    Carve c{base};
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(nu * nu);
    Gm = c.take<double>(nu * nu);
    vv = c.take<double>(nu);
    xx = c.take<double>(nu);
    yy = c.take<double>(nu);
    anew = c.take<T>(na);
    dA = c.take<T>(na);
    t1 = c.take<T>(na);
    t2 = c.take<T>(na);
    Oup = c.take<T>(na);
    Odn = c.take<T>(na);
    phin = c.take<double>(no);
    dphi = c.take<double>(no);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

template <typename T>
inline int64_t sweep_lds_bytes(const Sys& s) {
  return (int64_t)SweepLds<T>(0, s).end;
}

__device__ __forceinline__ double step_factor(double Z, double d) { return 1.0 / (Z * Z) * (1.0 + Z * Z * d) / (1.0 + d); }

// Attaccalite-Sorella factor of (G, G^-1) after a candidate rank-1 change held in LDS:
// mode 0 = current matrices, 1 = row idx += vv (up move), 2 = column idx += vv (dn move);
// the candidate inverse is G^-1 - xx yy^T / ratio (determinant.py:1223-1260).
template <typename LdsT>
__device__ inline double sweep_as_factor(const LdsT& L, int nu, int lane, int mode, int idx, double ratio) {
  double fro = 0.0;
  for (int t = lane; t < nu * nu; t += WAVE) {
    const int p = t / nu, q = t - p * nu;
    double gi = L.Gi[t];
    if (mode) gi -= L.xx[p] * L.yy[q] / ratio;
    fro += gi * gi;
  }
  fro = wave_sum(fro);
  double smin = 1.0e300;
  for (int p = lane; p < nu; p += WAVE) {
    double rs = 0.0, cs = 0.0;
    for (int q = 0; q < nu; ++q) {
      double gr = L.Gm[p * nu + q], gc = L.Gm[q * nu + p];
      if (mode == 1) {
        if (p == idx) gr += L.vv[q];
        if (q == idx) gc += L.vv[p];
      } else if (mode == 2) {
        if (q == idx) gr += L.vv[p];
        if (p == idx) gc += L.vv[q];
      }
      rs += gr * gr;
      cs += gc * gc;
    }
    smin = fmin(smin, fmin(rs, cs));
  }
  smin = wave_min(smin);
  const double sf = smin * fro;
  return sf > 0.0 ? pow(sf, -3.0 / 8.0) : 0.0;
}

// Metropolis sweep (_update_electron_positions, jqmc_mcmc.py:4316-4531): n_moves proposals per
// walker, one wavefront per walker, WPB walkers per block moving in lockstep so that one pass
// over J3 / J3^T per move serves every walker of the block that accepted.
template <int WPB, typename T>
__global__ void __launch_bounds__(WPB * 64) k_vmc_sweep(Sys s, int64_t W, WalkT<T> wk,
                                                         int n_moves, double Dt, double eps_AS, uint64_t seed,
                                                         int64_t move_base, const int64_t* __restrict__ move_ctr,
                                                         int64_t lds_per_wave) {
  const int64_t move0 = move_base + (move_ctr ? *move_ctr : 0);
  extern __shared__ double smem[];
  __shared__ int acc_flag[WPB];
  const RealTabs<T>& R = tabs<T>(s);
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * WPB + wave;
  const bool live = w < W;  // dead waves still take part in the block barriers
  char* const base = reinterpret_cast<char*>(smem);
  SweepLds<T> L(base + wave * lds_per_wave, s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  const int64_t wc = live ? w : 0;
  T* Aw = wk.ao + wc * ne * na;
  double* Pw = wk.orb + wc * ne * no;
  double* Qw = wk.pair + wc * ne * no;
  T* gw = s.j3_on ? wk.jfield + wc * ne * na : nullptr;

  if (live) {
    for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = wk.r_up[w * 3 * nu + k];
    for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
    for (int k = lane; k < nu * nu; k += WAVE) {
      L.Gi[k] = wk.Ginv[w * nu * nu + k];
      L.Gm[k] = wk.G[w * nu * nu + k];
    }
    for (int a = lane; a < na; a += WAVE) L.Oup[a] = L.Odn[a] = 0;
  }
  wave_sync();
  double R_o = live ? sweep_as_factor(L, nu, lane, 0, 0, 1.0) : 0.0;

  int64_t acc = 0, rej = 0;
  for (int m = 0; m < n_moves; ++m) {
    bool accepted = false;
    bool is_up = true;
    int idx = 0, e = 0;
    double det_ratio = 1.0, R_p = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
    if (live) {
      const uint64_t mv = (uint64_t)(move0 + m);
      // ---- draws in the reference order (jqmc_mcmc.py:4322-4500)
      const U4 d0 = draw_words(seed, (uint32_t)w, mv, STREAM_VMC, 0);
      const U4 d1 = draw_words(seed, (uint32_t)w, mv, STREAM_VMC, 1);
      const U4 d2 = draw_words(seed, (uint32_t)w, mv, STREAM_VMC, 2);
      const int rand_num = randint(d0.x, ne);
      const int up_index = randint(d0.y, nu > 0 ? nu : 1);
      const int dn_index = randint(d0.z, nd > 0 ? nd : 1);
      const int axis = randint(d0.w, 3);
      const double u1 = uniform01_open0(d1.x, d1.y), u2 = uniform01(d1.z, d1.w);
      const double gauss = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
      const double bdraw = uniform01(d2.x, d2.y);
      is_up = rand_num < nu;
      idx = is_up ? up_index : dn_index;
      e = is_up ? idx : nu + idx;

      // ---- proposal with nearest-nucleus step size (jqmc_mcmc.py:4339-4401)
      const double ox = L.rr[3 * e], oy = L.rr[3 * e + 1], oz = L.rr[3 * e + 2];
      const int na0 = nearest_nucleus(s, ox, oy, oz);
      const double Z0 = s.nuc_charge[na0];
      double dxn = ox - s.nuc_pos[3 * na0], dyn = oy - s.nuc_pos[3 * na0 + 1], dzn = oz - s.nuc_pos[3 * na0 + 2];
      const double f_l = step_factor(Z0, sqrt(dxn * dxn + dyn * dyn + dzn * dzn));
      const double g = gauss * (f_l * Dt);
      nx = ox;
      ny = oy;
      nz = oz;
      if (axis == 0) nx += g;
      else if (axis == 1) ny += g;
      else nz += g;
      const int na1 = nearest_nucleus(s, nx, ny, nz);
      const double Z1 = s.nuc_charge[na1];
      dxn = nx - s.nuc_pos[3 * na1];
      dyn = ny - s.nuc_pos[3 * na1 + 1];
      dzn = nz - s.nuc_pos[3 * na1 + 2];
      const double f_p = step_factor(Z1, sqrt(dxn * dxn + dyn * dyn + dzn * dzn));
      const double ddx = nx - ox, ddy = ny - oy, ddz = nz - oz;
      const double dist2 = ddx * ddx + ddy * ddy + ddz * ddz;
      const double T_ratio =
          (f_l / f_p) * exp(-dist2 * (1.0 / (2.0 * f_p * f_p * Dt * Dt) - 1.0 / (2.0 * f_l * f_l * Dt * Dt)));

      // ---- AO values at the new position, Jastrow log-ratio
      const T* Ae = Aw + (int64_t)e * na;
      const T* ge = gw ? gw + (int64_t)e * na : nullptr;
      double jpart = 0.0;
      const T* Oe = is_up ? L.Oup : L.Odn;
      for (int a = lane; a < na; a += WAVE) {
        const T v = ao_value<T>(s, a, nx, ny, nz);
        const T da = v - Ae[a];
        L.anew[a] = v;
        L.dA[a] = da;
        if (ge) jpart += (double)(da * (ge[a] + Oe[a]));
      }
      if (s.j1_type) {
        for (int n = lane; n < s.n_nuc; n += WAVE) {
          const double* Rn = s.nuc_pos + 3 * n;
          const double d1x = nx - Rn[0], d1y = ny - Rn[1], d1z = nz - Rn[2];
          const double d0x = ox - Rn[0], d0y = oy - Rn[1], d0z = oz - Rn[2];
          jpart += j1_term(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], sqrt(d1x * d1x + d1y * d1y + d1z * d1z)) -
                   j1_term(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], sqrt(d0x * d0x + d0y * d0y + d0z * d0z));
        }
      }
      if (s.j2_type) {
        for (int f = lane; f < ne; f += WAVE) {
          if (f == e) continue;
          const double* q = L.rr + 3 * f;
          const double d1x = nx - q[0], d1y = ny - q[1], d1z = nz - q[2];
          const double d0x = ox - q[0], d0y = oy - q[1], d0z = oz - q[2];
          jpart += j2_u(s.j2_type, s.j2_param, sqrt(d1x * d1x + d1y * d1y + d1z * d1z)) -
                   j2_u(s.j2_type, s.j2_param, sqrt(d0x * d0x + d0y * d0y + d0z * d0z));
        }
      }
      const double J_ratio = exp(wave_sum(jpart));
      wave_sync();

      // ---- orbitals at the new position (FP64 sums of the AO values) and their change
      const double* Pe = Pw + (int64_t)e * no;
      if (s.orb_is_mo) {
        wave_matvec_T(is_up ? s.td.mo_upT : s.td.mo_dnT, L.anew, na, no, lane, L.phin);
        wave_sync();
      } else {
        for (int k = lane; k < no; k += WAVE) L.phin[k] = (double)L.anew[k];
      }
      for (int k = lane; k < no; k += WAVE) L.dphi[k] = L.phin[k] - Pe[k];
      wave_sync();

      // ---- changed geminal row (up) or column (dn) and the determinant ratio
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
      det_ratio = 1.0 + wave_sum(part);
      wave_sync();
      // x = G^-1 u, y = v^T G^-1
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

      // ---- AS-regularised acceptance (jqmc_mcmc.py:4484-4520)
      R_p = sweep_as_factor(L, nu, lane, is_up ? 1 : 2, idx, det_ratio);
      const double R_AS_ratio = (fmax(R_p, eps_AS) / R_p) / (fmax(R_o, eps_AS) / R_o);
      const double wr = R_AS_ratio * J_ratio * det_ratio;
      accepted = bdraw < fmin(1.0, wr * wr * T_ratio);
      if (lane == 0) acc_flag[wave] = accepted ? 1 : 0;
    } else if (lane == 0) {
      acc_flag[wave] = 0;
    }

    // ---- J3 field increments t1 = J3 dA, t2 = J3^T dA for every accepting walker of the block,
    //      one coalesced pass over J3 and J3^T
    if (s.j3_on) {
      __syncthreads();
      int any = 0;
#pragma unroll
      for (int v = 0; v < WPB; ++v) any |= acc_flag[v];
      if (any) {
        const T* dAv[WPB];
        T* T1v[WPB];
        T* T2v[WPB];
#pragma unroll
        for (int v = 0; v < WPB; ++v) {
          const SweepLds<T> Lv(base + v * lds_per_wave, s);
          dAv[v] = Lv.dA;
          T1v[v] = Lv.t1;
          T2v[v] = Lv.t2;
        }
        for (int a = threadIdx.x; a < na; a += WPB * WAVE) {
          T t1[WPB], t2[WPB];
#pragma unroll
          for (int v = 0; v < WPB; ++v) t1[v] = t2[v] = 0;
          if (s.j3_sym) {
#pragma unroll 4
            for (int b = 0; b < na; ++b) {
              const T jj = R.j3[(int64_t)b * na + a];
#pragma unroll
              for (int v = 0; v < WPB; ++v)
                if (acc_flag[v]) t2[v] += jj * dAv[v][b];
            }
#pragma unroll
            for (int v = 0; v < WPB; ++v) t1[v] = t2[v];
          } else {
#pragma unroll 4
            for (int b = 0; b < na; ++b) {
              const T jt = R.j3T[(int64_t)b * na + a];
              const T jj = R.j3[(int64_t)b * na + a];
#pragma unroll
              for (int v = 0; v < WPB; ++v) {
                if (acc_flag[v]) {
                  const T db = dAv[v][b];
                  t1[v] += jt * db;
                  t2[v] += jj * db;
                }
              }
            }
          }
#pragma unroll
          for (int v = 0; v < WPB; ++v) {
            if (acc_flag[v]) {
              T1v[v][a] = t1[v];
              T2v[v][a] = t2[v];
            }
          }
        }
      }
      __syncthreads();
    }

    if (live) {
      if (accepted) {
        ++acc;
        if (lane == 0) {
          L.rr[3 * e] = nx;
          L.rr[3 * e + 1] = ny;
          L.rr[3 * e + 2] = nz;
        }
        T* Aem = Aw + (int64_t)e * na;
        for (int a = lane; a < na; a += WAVE) Aem[a] = L.anew[a];
        double* Pem = Pw + (int64_t)e * no;
        double* Qem = Qw + (int64_t)e * no;
        const double* Lm = is_up ? s.td.lam_p : s.td.lam_pT;
        for (int k = lane; k < no; k += WAVE) {
          Pem[k] = L.phin[k];
          double q = 0.0;
          for (int j = 0; j < no; ++j) q += Lm[(int64_t)j * no + k] * L.phin[j];
          Qem[k] = q;
        }
        if (gw) {
          // Field increments (jastrow_factor.py:4121-4135): up move e gives t1 to up f < e and t2 to
          // every other electron; dn move e gives t1 to all up and to dn f < e, t2 to dn f > e.
          // The part shared by a whole spin block goes into the offsets O_up / O_dn (flushed at
          // the end of the sweep), the mover's own stored field is compensated, and only part of
          // the mover's spin block takes an individual correction (none for symmetric J3).
          // The same-spin block may instead take t1 as its offset, with t2 - t1 for f > e: the side
          // with fewer electrons is chosen.
          const int f0 = is_up ? 0 : nu, f1 = is_up ? nu : ne;
          const bool low = s.j3_sym || (e - f0) <= (f1 - 1 - e);
          const int c0 = low ? f0 : e + 1, c1 = low ? e : f1;
          for (int a = lane; a < na; a += WAVE) {
            const T t1 = L.t1[a], t2 = L.t2[a];
            const T own = low ? t2 : t1;  // offset of the mover's block
            T* gcol = gw + a;
            gcol[(int64_t)e * na] -= own;
            if (is_up) {
              L.Oup[a] += own;
              L.Odn[a] += t2;
            } else {
              L.Oup[a] += t1;
              L.Odn[a] += own;
            }
            if (!s.j3_sym) {
              const T d = low ? t1 - t2 : t2 - t1;
#pragma unroll 4
              for (int f = c0; f < c1; ++f) gcol[(int64_t)f * na] += d;
            }
          }
        }
        for (int t = lane; t < nu * nu; t += WAVE) {
          const int p = t / nu, q = t - p * nu;
          L.Gi[t] -= L.xx[p] * L.yy[q] / det_ratio;
          if (is_up && p == idx) L.Gm[t] += L.vv[q];
          if (!is_up && q == idx) L.Gm[t] += L.vv[p];
        }
        R_o = R_p;
      } else {
        ++rej;
      }
      wave_sync();
    }
  }

  if (live) {
    if (gw) {
      for (int a = lane; a < na; a += WAVE) {
        const T ou = L.Oup[a], od = L.Odn[a];
        for (int f = 0; f < ne; ++f) gw[(int64_t)f * na + a] += f < nu ? ou : od;
      }
    }
    for (int k = lane; k < 3 * nu; k += WAVE) wk.r_up[w * 3 * nu + k] = L.rr[k];
    for (int k = lane; k < 3 * nd; k += WAVE) wk.r_dn[w * 3 * nd + k] = L.rr[3 * nu + k];
    for (int k = lane; k < nu * nu; k += WAVE) {
      wk.Ginv[w * nu * nu + k] = L.Gi[k];
      wk.G[w * nu * nu + k] = L.Gm[k];
    }
    if (lane == 0) {
      wk.accepted[w] += acc;
      wk.rejected[w] += rej;
      if (wk.as_factor) wk.as_factor[w] = R_o;
    }
  }
}

// ============================================================ single-electron point ratios

// Points evaluated together per pass of the AO loop of electron_points_ratio: the LDS radial table
// holds up to PTS_CHUNK; the forces kernel uses EL_PTS_CHUNK (fewer live registers).
constexpr int PTS_CHUNK = 6;
constexpr int EL_PTS_CHUNK = 4;

// Determinant field of electron e into h (LDS, n_ao): q_e = [M | Lambda_u] G^-1[:, e] (up) or
// M' G^-1[e, :]^T (dn) into q, then h_e = C^T q_e, in FP64; the single-electron determinant ratio at
// r' is a(r') . h_e (determinant.py:1529-1783).  Ends with a wave_sync.
__device__ inline void det_field(const Sys& s, const double* Gi, double* q, double* h, const double* Qw, int e,
                                 int lane) {
  const int nu = s.n_up, nd = s.n_dn, na = s.n_ao, no = s.n_orb;
  const bool up = e < nu;
  for (int k = lane; k < no; k += WAVE) {
    double acc = 0.0;
    if (up) {
      for (int j = 0; j < nd; ++j) acc += Qw[(int64_t)(nu + j) * no + k] * Gi[j * nu + e];
      for (int u = 0; u < s.n_unp; ++u) acc += s.td.lam_u[(int64_t)k * s.n_unp + u] * Gi[(nd + u) * nu + e];
    } else {
      const int j = e - nu;
      for (int i = 0; i < nu; ++i) acc += Qw[(int64_t)i * no + k] * Gi[j * nu + i];
    }
    q[k] = acc;
  }
  wave_sync();
  const double* C = up ? s.td.mo_up : s.td.mo_dn;
  for (int a = lane; a < na; a += WAVE) {
    double hv;
    if (s.orb_is_mo) {
      hv = 0.0;
      for (int k = 0; k < no; ++k) hv += C[(int64_t)k * na + a] * q[k];
    } else {
      hv = q[a];
    }
    h[a] = hv;
  }
  wave_sync();
}

// det[q] = a(p_q) . h and dj[q] = ln J(p_q) - ln J(r_e) for the n points p_q (pts, LDS) of
// electron e: the J3 part (a(p_q) - a(r_e)) . g_e plus the one- and two-body changes
// (_compute_ratio_Jastrow_part_rank1_update, jastrow_factor.py:2696-2960).  Lanes own AOs for the
// basis sums and own (other electron | nucleus) pairs for the J1/J2 sums; PTS_CHUNK points per
// pass, one butterfly per chunk.  Basis sums in T, Jastrow pair sums and the results in FP64.
template <typename T, int PC = PTS_CHUNK>
__device__ void electron_points_ratio(const Sys& s, const double* rr, int e, const double* h, const T* Ae, const T* ge,
                                      const double* pts, int n, double* det_out, double* dj_out, T* rad, int lane) {
  static_assert(PC <= PTS_CHUNK, "radial table holds PTS_CHUNK points per shell");
  const int na = s.n_ao, nsh = s.n_shell;
  const RealTabs<T>& R = tabs<T>(s);
  const int n_j2 = s.j2_type ? s.n_el : 0;
  const int n_j = n_j2 + (s.j1_type ? s.n_nuc : 0);
  const double ex = rr[3 * e], ey = rr[3 * e + 1], ez = rr[3 * e + 2];
  double old = 0.0;  // lane part of ln J1 + ln J2 at r_e
  for (int i = lane; i < n_j; i += WAVE) {
    if (i < n_j2) {
      if (i == e) continue;
      const double dx = ex - rr[3 * i], dy = ey - rr[3 * i + 1], dz = ez - rr[3 * i + 2];
      old += j2_u(s.j2_type, s.j2_param, sqrt(dx * dx + dy * dy + dz * dz));
    } else {
      const int m = i - n_j2;
      const double dx = ex - s.nuc_pos[3 * m], dy = ey - s.nuc_pos[3 * m + 1], dz = ez - s.nuc_pos[3 * m + 2];
      old += j1_term(s.j1_type, s.j1_param, s.j1_c[m], s.j1_A[m], sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  double old_sum = 0.0;
  for (int c0 = 0; c0 < n; c0 += PC) {
    const int nq = (n - c0) < PC ? (n - c0) : PC;
    double px[PC], py[PC], pz[PC];
#pragma unroll
    for (int q = 0; q < PC; ++q) {
      const int qq = q < nq ? c0 + q : c0;
      px[q] = wave_uniform(pts[3 * qq]);
      py[q] = wave_uniform(pts[3 * qq + 1]);
      pz[q] = wave_uniform(pts[3 * qq + 2]);
    }
    // radial factor of every shell at the chunk's points (lanes own shells) ...
    for (int sh = lane; sh < nsh; sh += WAVE) {
      const int a0s = s.sh_ao0[sh];
      const double cx = s.ao_ctr[3 * a0s], cy = s.ao_ctr[3 * a0s + 1], cz = s.ao_ctr[3 * a0s + 2];
      T r2[PC], rd[PC];
#pragma unroll
      for (int q = 0; q < PC; ++q) {
        const T dx = (T)(px[q] - cx), dy = (T)(py[q] - cy), dz = (T)(pz[q] - cz);
        r2[q] = dx * dx + dy * dy + dz * dz;
        rd[q] = 0;
      }
      const int p1 = s.ao_pp[a0s + 1];
      for (int p = s.ao_pp[a0s]; p < p1; ++p) {
        const T Z = R.prim_exp[p], c = R.prim_coef[p];
#pragma unroll
        for (int q = 0; q < PC; ++q) rd[q] += c * exp_nonpos(-Z * r2[q]);
      }
#pragma unroll
      for (int q = 0; q < PC; ++q) rad[sh * PC + q] = rd[q];
    }
    wave_sync();
    // ... then the AO sums (lanes own AOs): a(p) . h in FP64 (determinant ratio) and
    // (a(p) - a_e) . g in T (Jastrow ratio)
    double bdet[PC];
    T bj3[PC];
#pragma unroll
    for (int k = 0; k < PC; ++k) {
      bdet[k] = 0.0;
      bj3[k] = 0;
    }
    for (int a = lane; a < na; a += WAVE) {
      const double cx = s.ao_ctr[3 * a], cy = s.ao_ctr[3 * a + 1], cz = s.ao_ctr[3 * a + 2];
      const T* rdp = rad + s.ao_sh[a] * PC;
      T dx[PC], dy[PC], dz[PC], pol[PC];
#pragma unroll
      for (int q = 0; q < PC; ++q) {
        dx[q] = (T)(px[q] - cx);
        dy[q] = (T)(py[q] - cy);
        dz[q] = (T)(pz[q] - cz);
        pol[q] = 0;
      }
      const int t1 = s.ao_tp[a + 1];
      for (int t = s.ao_tp[a]; t < t1; ++t) {
        const T tc = R.term_coef[t];
        const int nx = s.term_pow[3 * t], ny = s.term_pow[3 * t + 1], nz = s.term_pow[3 * t + 2];
#pragma unroll
        for (int q = 0; q < PC; ++q) pol[q] += tc * (ipow(dx[q], nx) * ipow(dy[q], ny) * ipow(dz[q], nz));
      }
      const T sc = R.ao_scale[a];
      const double hv = h[a];
      const T a0 = Ae[a], gv = ge ? ge[a] : (T)0;
#pragma unroll
      for (int q = 0; q < PC; ++q) {
        const T v = sc * (rdp[q] * pol[q]);
        bdet[q] += (double)v * hv;
        bj3[q] += (v - a0) * gv;
      }
    }
    double acc[2 * PC + 1];
#pragma unroll
    for (int k = 0; k < PC; ++k) {
      acc[k] = bdet[k];
      acc[PC + k] = (double)bj3[k];
    }
    acc[2 * PC] = c0 == 0 ? old : 0.0;
    for (int i = lane; i < n_j; i += WAVE) {
      if (i < n_j2) {
        if (i == e) continue;
        const double fx = rr[3 * i], fy = rr[3 * i + 1], fz = rr[3 * i + 2];
#pragma unroll
        for (int q = 0; q < PC; ++q) {
          if (q < nq) {
            const double dx = px[q] - fx, dy = py[q] - fy, dz = pz[q] - fz;
            acc[PC + q] += j2_u(s.j2_type, s.j2_param, sqrt(dx * dx + dy * dy + dz * dz));
          }
        }
      } else {
        const int m = i - n_j2;
        const double Rx = s.nuc_pos[3 * m], Ry = s.nuc_pos[3 * m + 1], Rz = s.nuc_pos[3 * m + 2];
        const double jc = s.j1_c[m], jA = s.j1_A[m];
#pragma unroll
        for (int q = 0; q < PC; ++q) {
          if (q < nq) {
            const double dx = px[q] - Rx, dy = py[q] - Ry, dz = pz[q] - Rz;
            acc[PC + q] += j1_term(s.j1_type, s.j1_param, jc, jA, sqrt(dx * dx + dy * dy + dz * dz));
          }
        }
      }
    }
    wave_sum_k<2 * PC + 1>(acc);
    if (c0 == 0) old_sum = acc[2 * PC];
    if (lane == 0) {
      for (int q = 0; q < nq; ++q) {
        det_out[c0 + q] = acc[q];
        dj_out[c0 + q] = acc[PC + q] - old_sum;
      }
    }
    wave_sync();  // rad is rewritten by the next chunk
  }
}

// Rotated quadrature points g~_k = g_k RT into gk (nv, 3) (coulomb_potential.py:1546-1547).
// Threads first, first + stride, ... own the points (one wavefront: lane, WAVE).
__device__ __forceinline__ void rotate_mesh(const Sys& s, const double* RT, double* gk, int first, int stride = WAVE) {
  for (int k = first; k < s.nv; k += stride) {
    for (int c = 0; c < 3; ++c)
      gk[3 * k + c] =
          s.mesh_pts[3 * k] * RT[c] + s.mesh_pts[3 * k + 1] * RT[3 + c] + s.mesh_pts[3 * k + 2] * RT[6 + c];
  }
}

// LDS of one wavefront of k_point_ratios.
constexpr int RATIO_RUN = 32;  // points of one electron staged per pass
template <typename T>
struct RatioLds {
  double *rr, *Gi, *q, *h, *pts, *det, *dj;
  T* rad;
  uintptr_t end;
  __host__ __device__ RatioLds(char* base, const Sys& s) {
    Carve c{base};
    rr = c.take<double>(3 * s.n_el);
    Gi = c.take<double>(s.n_up * s.n_up);
    q = c.take<double>(s.n_orb);
    h = c.take<double>(s.n_ao);
    pts = c.take<double>(3 * RATIO_RUN);
    det = c.take<double>(RATIO_RUN);
    dj = c.take<double>(RATIO_RUN);
    rad = c.take<T>((int64_t)s.n_shell * PTS_CHUNK);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

template <typename T>
inline int64_t ratio_lds_bytes(const Sys& s) {
  return (int64_t)RatioLds<T>(0, s).end;
}

// Single-electron ratios of walker w at P proposed configurations: configuration p moves electron
// elec[w, p] (up block first) to pts[w, p] and leaves the others; det_out[w, p] = det G' / det G
// (_compute_ratio_determinant_part_rank1_update, determinant.py:1529-1662, with the walker's G^-1)
// and jas_out[w, p] = exp(J' - J) (_compute_ratio_Jastrow_part_rank1_update,
// jastrow_factor.py:2230-2693).  One wavefront per walker; consecutive configurations of one
// electron share its determinant field and are evaluated RATIO_RUN at a time.  An electron index
// outside [0, N) gives NaN for that configuration.
template <typename T>
__global__ void __launch_bounds__(64) k_point_ratios(Sys s, int64_t W, WalkT<T> wk, int64_t P,
                                                     const int32_t* __restrict__ elec, const double* __restrict__ pts,
                                                     double* det_out, double* jas_out, int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int lane = lane_id();
  const int64_t w = blockIdx.x;
  if (w >= W) return;
  RatioLds<T> L(reinterpret_cast<char*>(smem), s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = wk.r_up[w * 3 * nu + k];
  for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
  for (int k = lane; k < nu * nu; k += WAVE) L.Gi[k] = wk.Ginv[w * nu * nu + k];
  wave_sync();
  const T* Aw = wk.ao + w * ne * na;
  const double* Qw = wk.pair + w * ne * no;
  const T* gw = s.j3_on ? wk.jfield + w * ne * na : nullptr;
  const int32_t* ew = elec + w * P;
  int e_cur = -1;
  int64_t p = 0;
  while (p < P) {
    const int e = ew[p];
    int n = 1;
    while (p + n < P && n < RATIO_RUN && ew[p + n] == e) ++n;
    if (e < 0 || e >= ne) {
      for (int k = lane; k < n; k += WAVE) {
        det_out[w * P + p + k] = __builtin_nan("");
        jas_out[w * P + p + k] = __builtin_nan("");
      }
      p += n;
      continue;
    }
    if (e != e_cur) {
      det_field(s, L.Gi, L.q, L.h, Qw, e, lane);
      e_cur = e;
    }
    for (int k = lane; k < 3 * n; k += WAVE) L.pts[k] = pts[(w * P + p) * 3 + k];
    wave_sync();
    electron_points_ratio<T>(s, L.rr, e, L.h, Aw + (int64_t)e * na, gw ? gw + (int64_t)e * na : nullptr, L.pts, n,
                             L.det, L.dj, L.rad, lane);
    for (int k = lane; k < n; k += WAVE) {
      det_out[w * P + p + k] = L.det[k];
      jas_out[w * P + p + k] = exp(L.dj[k]);
    }
    wave_sync();
    p += n;
  }
}

// Per-electron part of the non-local projector (NN = 1): nearest nucleus R, R - r_e, |R - r_e|
// and the radial factors V_t = |R - r|^-2 c_t |R - r|^n_t exp(-a_t |R - r|^2) of its projected
// terms into Vt (LDS) (coulomb_potential.py:1546-1625).  Every lane gets the same values.
struct EcpElec {
  int n, nt;
  double rx, ry, rz, rn;
};

__device__ __forceinline__ EcpElec ecp_electron(const Sys& s, const double* rr, int e, double* Vt, int lane) {
  EcpElec E;
  const double ex = rr[3 * e], ey = rr[3 * e + 1], ez = rr[3 * e + 2];
  E.n = nearest_nucleus(s, ex, ey, ez);
  E.nt = s.ecp_nt[E.n];
  E.rx = s.nuc_pos[3 * E.n] - ex;
  E.ry = s.nuc_pos[3 * E.n + 1] - ey;
  E.rz = s.nuc_pos[3 * E.n + 2] - ez;
  E.rn = sqrt(E.rx * E.rx + E.ry * E.ry + E.rz * E.rz);
  for (int t = lane; t < E.nt; t += WAVE) {
    const int o = E.n * s.ecp_max_terms + t;
    Vt[t] = (1.0 / (E.rn * E.rn)) * s.ecp_coef[o] * rpow(E.rn, s.ecp_pow[o]) * exp_nonpos(-s.ecp_exp[o] * (E.rn * E.rn));
  }
  return E;
}

// Quadrature point k of the electron, r' = r + (R - r) + |R - r| g~_k, and the projector factor
// sum_t V_t (2 l_t + 1) P_l(cos theta) w_k (coulomb_potential.py:1608-1705).
__device__ __forceinline__ void ecp_point(const Sys& s, const EcpElec& E, const double* Vt, const double* rr,
                                          const double* gk, int e, int k, double& px, double& py, double& pz,
                                          double& wsum) {
  const double gx = gk[3 * k], gy = gk[3 * k + 1], gz = gk[3 * k + 2];
  const double gn = sqrt(gx * gx + gy * gy + gz * gz);
  px = rr[3 * e] + (E.rx + E.rn * gx);
  py = rr[3 * e + 1] + (E.ry + E.rn * gy);
  pz = rr[3 * e + 2] + (E.rz + E.rn * gz);
  const double cth = (-E.rx / E.rn) * (gx / gn) + (-E.ry / E.rn) * (gy / gn) + (-E.rz / E.rn) * (gz / gn);
  wsum = 0.0;
  for (int t = 0; t < E.nt; ++t) {
    const int l = s.ecp_l[E.n * s.ecp_max_terms + t];
    wsum += Vt[t] * ((2 * l + 1) * legendre(l, cth) * s.mesh_w[k]);
  }
}

// Local ECP part of one electron-nucleus distance d (coulomb_potential.py:1144-1246).
__device__ __forceinline__ double ecp_local_pair(const Sys& s, int n, double d) {
  double v = 0.0;
  for (int k = 0; k < s.ecpl_nt[n]; ++k) {
    const int o = n * s.ecp_max_local + k;
    v += (1.0 / (d * d)) * s.ecpl_coef[o] * rpow(d, s.ecpl_pow[o]) * exp_nonpos(-s.ecpl_exp[o] * (d * d));
  }
  return v;
}

// Lane-partial electron-electron repulsion sum_{e<f} 1/r_ef (coulomb_potential.py:2464-2518).
// Threads first, first + stride, ... own the electrons e (one wavefront: lane, WAVE).
__device__ __forceinline__ double el_el_partial(const Sys& s, const double* rr, int first, int stride = WAVE) {
  double v = 0.0;
  const int ne = s.n_el;
  for (int e = first; e < ne; e += stride) {
    for (int f = e + 1; f < ne; ++f) {
      const double dx = rr[3 * e] - rr[3 * f], dy = rr[3 * e + 1] - rr[3 * f + 1], dz = rr[3 * e + 2] - rr[3 * f + 2];
      v += 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  return v;
}

}  // namespace jqmc
