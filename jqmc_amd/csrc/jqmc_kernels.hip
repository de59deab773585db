// gfx950 kernels of the jQMC per-walker hot path.
//
// Layout in HBM: every per-walker array is walker-major and contiguous per walker
// (W, ...), per-electron vectors are electron-major with the AO / orbital index
// innermost, so the lanes of a wavefront that own consecutive AO indices read and
// write contiguous 512-byte segments.  A lane owns AO/orbital index k = lane + 64 j
// for every per-electron vector it touches in the sweep; cross-lane traffic goes
// through LDS.
//
// Kernels:
//   k_aos_vgl          AO value/grad/lap at arbitrary points (thread per AO x point)
//   k_orbs_contract    orbital = C . AO for the five AO outputs (thread per orbital x point)
//   k_init_*           walker-state construction (G, LU inverse, pair and J3 fields)
//   k_vmc_sweep        fused Metropolis loop, one wavefront per walker
//   k_local_energy     fused kinetic + Coulomb + ECP local energy, one wavefront per walker
//   k_rotations        per-walker quadrature rotations
#include "jqmc_device.hpp"

namespace jqmc {

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
    val[idx] = ao_value(s, a, x, y, z);
    return;
  }
  double v, g0, g1, g2, l;
  ao_vgl(s, a, x, y, z, v, g0, g1, g2, l);
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

__global__ void k_rotations(int64_t W, uint64_t seed, int64_t step, uint32_t stream, double* __restrict__ RT) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  double R[9];
  rotation_RT(seed, (uint32_t)w, (uint64_t)step, stream, R);
  for (int k = 0; k < 9; ++k) RT[9 * w + k] = R[k];
}

// ============================================================ walker initialisation

__device__ __forceinline__ const double* elec_pos(const Sys& s, const Walk& wk, int64_t w, int e) {
  return e < s.n_up ? wk.r_up + (w * s.n_up + e) * 3 : wk.r_dn + (w * s.n_dn + (e - s.n_up)) * 3;
}

// ao[w, e, a] = AO a at electron e.
__global__ void k_init_ao(Sys s, int64_t W, Walk wk) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_w = (int64_t)s.n_el * s.n_ao;
  if (idx >= W * per_w) return;
  const int64_t w = idx / per_w;
  const int rem = (int)(idx - w * per_w);
  const int e = rem / s.n_ao, a = rem - e * s.n_ao;
  const double* r = elec_pos(s, wk, w, e);
  wk.ao[idx] = ao_value(s, a, r[0], r[1], r[2]);
}

// orb[w, e, m] = (C_spin . ao[w, e])[m] or the AO itself.
__global__ void k_init_orb(Sys s, int64_t W, Walk wk) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_w = (int64_t)s.n_el * s.n_orb;
  if (idx >= W * per_w) return;
  const int64_t w = idx / per_w;
  const int rem = (int)(idx - w * per_w);
  const int e = rem / s.n_orb, m = rem - e * s.n_orb;
  const double* a_e = wk.ao + (w * s.n_el + e) * s.n_ao;
  if (!s.orb_is_mo) {
    wk.orb[idx] = a_e[m];
    return;
  }
  const double* C = (e < s.n_up ? s.mo_up : s.mo_dn) + (int64_t)m * s.n_ao;
  double acc = 0.0;
  for (int a = 0; a < s.n_ao; ++a) acc += C[a] * a_e[a];
  wk.orb[idx] = acc;
}

// pair[w, e] = Lambda_p^T phi_e (up) or Lambda_p phi_e (dn).
__global__ void k_init_pair(Sys s, int64_t W, Walk wk) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_w = (int64_t)s.n_el * s.n_orb;
  if (idx >= W * per_w) return;
  const int64_t w = idx / per_w;
  const int rem = (int)(idx - w * per_w);
  const int e = rem / s.n_orb, m = rem - e * s.n_orb;
  const double* phi = wk.orb + (w * s.n_el + e) * s.n_orb;
  const double* L = e < s.n_up ? s.lam_p : s.lam_pT;  // column m of L^T or of L, read as rows
  double acc = 0.0;
  for (int k = 0; k < s.n_orb; ++k) acc += L[(int64_t)k * s.n_orb + m] * phi[k];
  wk.pair[idx] = acc;
}

// G[w, i, j] = phi_i . pair_dn_j (j < n_dn), phi_i . Lambda_u[:, j - n_dn] otherwise
// (_compute_geminal_all_elements, determinant.py:1379-1402).
__global__ void k_init_geminal(Sys s, int64_t W, Walk wk) {
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
    for (int m = 0; m < s.n_orb; ++m) acc += phi[m] * s.lam_u[(int64_t)m * s.n_unp + k];
  }
  wk.G[idx] = acc;
}

// Gauss-Jordan inverse with partial pivoting, one wavefront per walker (n_up <= 64).
// Pivot = largest |entry| in the column, lowest row on ties (LAPACK getrf rule used by
// jnp.linalg.det, determinant.py:1203).  Also the AS factor (determinant.py:1223-1260).
__global__ void k_init_inverse(Sys s, int64_t W, Walk wk, int wpb) {
  extern __shared__ double smem[];
  const int n = s.n_up;
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  double* M = smem + (size_t)wave * 2 * n * n;
  double* I = M + n * n;
  const double* G = wk.G + w * n * n;
  for (int k = lane; k < n * n; k += WAVE) {
    M[k] = G[k];
    I[k] = (k / n == k % n) ? 1.0 : 0.0;
  }
  wave_sync();
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
  double* Ginv = wk.Ginv + w * n * n;
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
    wk.logabsdet[w] = logabs;
    wk.sign[w] = sgn;
    if (wk.as_factor) {
      const double sf = smin * fro;
      wk.as_factor[w] = sf > 0.0 ? pow(sf, -3.0 / 8.0) : 0.0;
    }
  }
}

// Prefix sums of AO columns feeding the J3 fields: for electron e,
// S1[e] = sum of a_j over j in S1(e), S2[e] likewise, with
//   up e:  S1 = {up j > e} U {dn},  S2 = {up j < e}
//   dn e:  S1 = {dn j > e},         S2 = {dn j < e} U {up}
// so that g_e = j1 + J3 S1[e] + J3^T S2[e] (jastrow_factor.py:4121-4135).
__global__ void k_init_jsums(Sys s, int64_t W, const Walk wk, double* __restrict__ S1, double* __restrict__ S2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * s.n_ao) return;
  const int64_t w = idx / s.n_ao;
  const int a = (int)(idx - w * s.n_ao);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao;
  const double* A = wk.ao + w * ne * na;
  double tot_up = 0.0, tot_dn = 0.0;
  for (int e = 0; e < nu; ++e) tot_up += A[e * na + a];
  for (int e = nu; e < ne; ++e) tot_dn += A[e * na + a];
  double before = 0.0;
  for (int e = 0; e < nu; ++e) {
    const double ae = A[e * na + a];
    S2[(w * ne + e) * na + a] = before;
    S1[(w * ne + e) * na + a] = (tot_up - before - ae) + tot_dn;
    before += ae;
  }
  before = 0.0;
  for (int e = nu; e < ne; ++e) {
    const double ae = A[e * na + a];
    S2[(w * ne + e) * na + a] = before + tot_up;
    S1[(w * ne + e) * na + a] = tot_dn - before - ae;
    before += ae;
  }
  (void)nd;
}

// g[w, e, a] = j1[a] + sum_b J[a, b] S1[e, b] + sum_b J[b, a] S2[e, b].
__global__ void k_init_jfield(Sys s, int64_t W, Walk wk, const double* __restrict__ S1,
                              const double* __restrict__ S2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int na = s.n_ao;
  const int64_t per_w = (int64_t)s.n_el * na;
  if (idx >= W * per_w) return;
  const int64_t we = idx / na;
  const int a = (int)(idx - we * na);
  const double* s1 = S1 + we * na;
  const double* s2 = S2 + we * na;
  double acc = s.j3v[a];
  for (int b = 0; b < na; ++b) acc += s.j3T[(int64_t)b * na + a] * s1[b] + s.j3[(int64_t)b * na + a] * s2[b];
  wk.jfield[idx] = acc;
}

// J = J1 + J2 + J3 with J3 = 1/2 sum_e a_e . (g_e + j1), one wavefront per walker.
__global__ void k_init_jastrow(Sys s, int64_t W, Walk wk, int wpb) {
  const int wave = threadIdx.x / WAVE, lane = lane_id();
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
    for (int t = lane; t < ne * na; t += WAVE) {
      const int a = t % na;
      acc += 0.5 * wk.ao[w * ne * na + t] * (wk.jfield[w * ne * na + t] + s.j3v[a]);
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) wk.jastrow[w] = acc;
}

// ============================================================ VMC sweep

// LDS layout of one wavefront in the sweep (doubles).
struct SweepLds {
  double *rr, *Gi, *Gm, *anew, *dA, *phin, *dphi, *vv, *xx, *yy, *t1, *t2, *Oup, *Odn;
  __device__ SweepLds(double* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up;
    rr = base;
    Gi = rr + 3 * ne;
    Gm = Gi + nu * nu;
    anew = Gm + nu * nu;
    dA = anew + s.n_ao;
    t1 = dA + s.n_ao;
    t2 = t1 + s.n_ao;
    Oup = t2 + s.n_ao;
    Odn = Oup + s.n_ao;
    phin = Odn + s.n_ao;
    dphi = phin + s.n_orb;
    vv = dphi + s.n_orb;
    xx = vv + nu;
    yy = xx + nu;
  }
};

__host__ __device__ inline int64_t sweep_lds_doubles(int n_el, int n_up, int n_ao, int n_orb) {
  return 3 * n_el + 2 * n_up * n_up + 6 * n_ao + 2 * n_orb + 3 * n_up;
}

__device__ __forceinline__ double step_factor(double Z, double d) { return 1.0 / (Z * Z) * (1.0 + Z * Z * d) / (1.0 + d); }

// Attaccalite-Sorella factor of (G, G^-1) after a candidate rank-1 change held in LDS:
// mode 0 = current matrices, 1 = row idx += vv (up move), 2 = column idx += vv (dn move);
// the candidate inverse is G^-1 - xx yy^T / ratio (determinant.py:1223-1260).
__device__ inline double sweep_as_factor(const SweepLds& L, int nu, int lane, int mode, int idx, double ratio) {
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

// y[k] = sum_a CT[a, k] x[a] for k < no, all 64 lanes busy: lanes split into groups of
// lg >= no lanes, group g sums a = g, g + G, ...; groups combine by butterfly shuffles.
__device__ inline void wave_matvec_T(const double* __restrict__ CT, const double* x, int na, int no, int lane,
                                     double* y) {
  int lg = 64;
  while (lg / 2 >= no && lg > 1) lg >>= 1;
  const int G = WAVE / lg, k = lane % lg, grp = lane / lg;
  for (int kb = 0; kb < no; kb += lg) {
    const int kk = kb + k;
    double acc = 0.0;
    if (kk < no) {
#pragma unroll 4
      for (int a = grp; a < na; a += G) acc += CT[(int64_t)a * no + kk] * x[a];
    }
    for (int off = lg; off < WAVE; off <<= 1) acc += __shfl_xor(acc, off, WAVE);
    if (grp == 0 && kk < no) y[kk] = acc;
  }
}

// Metropolis sweep (_update_electron_positions, jqmc_mcmc.py:4316-4531): n_moves proposals per
// walker, one wavefront per walker, WPB walkers per block moving in lockstep so that one pass
// over J3 / J3^T per move serves every walker of the block that accepted.
template <int WPB>
__global__ void __launch_bounds__(WPB * 64) k_vmc_sweep(const Sys* __restrict__ sp, int64_t W, Walk wk, int n_moves,
                                                         double Dt, double eps_AS, uint64_t seed, int64_t move0,
                                                         int64_t lds_per_wave) {
  extern __shared__ double smem[];
  __shared__ int acc_flag[WPB];
  const Sys& s = *sp;
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * WPB + wave;
  const bool live = w < W;  // dead waves still take part in the block barriers
  SweepLds L(smem + (size_t)wave * lds_per_wave, s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  const int64_t wc = live ? w : 0;
  double* Aw = wk.ao + wc * ne * na;
  double* Pw = wk.orb + wc * ne * no;
  double* Qw = wk.pair + wc * ne * no;
  double* gw = s.j3_on ? wk.jfield + wc * ne * na : nullptr;

  if (live) {
    for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = wk.r_up[w * 3 * nu + k];
    for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
    for (int k = lane; k < nu * nu; k += WAVE) {
      L.Gi[k] = wk.Ginv[w * nu * nu + k];
      L.Gm[k] = wk.G[w * nu * nu + k];
    }
    for (int a = lane; a < na; a += WAVE) L.Oup[a] = L.Odn[a] = 0.0;
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
      const double* Ae = Aw + (int64_t)e * na;
      const double* ge = gw ? gw + (int64_t)e * na : nullptr;
      double jpart = 0.0;
      const double* Oe = is_up ? L.Oup : L.Odn;
      for (int a = lane; a < na; a += WAVE) {
        const double v = ao_value(s, a, nx, ny, nz);
        const double da = v - Ae[a];
        L.anew[a] = v;
        L.dA[a] = da;
        if (ge) jpart += da * (ge[a] + Oe[a]);
      }
      if (s.j1_type) {
        for (int n = lane; n < s.n_nuc; n += WAVE) {
          const double* R = s.nuc_pos + 3 * n;
          const double d1x = nx - R[0], d1y = ny - R[1], d1z = nz - R[2];
          const double d0x = ox - R[0], d0y = oy - R[1], d0z = oz - R[2];
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

      // ---- orbitals at the new position and their change
      const double* Pe = Pw + (int64_t)e * no;
      if (s.orb_is_mo) {
        wave_matvec_T(is_up ? s.mo_upT : s.mo_dnT, L.anew, na, no, lane, L.phin);
        wave_sync();
        for (int k = lane; k < no; k += WAVE) L.dphi[k] = L.phin[k] - Pe[k];
      } else {
        for (int k = lane; k < no; k += WAVE) {
          L.phin[k] = L.anew[k];
          L.dphi[k] = L.dA[k];
        }
      }
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
            for (int k = 0; k < no; ++k) v += L.dphi[k] * s.lam_u[(int64_t)k * s.n_unp + (j - nd)];
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
        for (int a = threadIdx.x; a < na; a += WPB * WAVE) {
          double t1[WPB], t2[WPB];
#pragma unroll
          for (int v = 0; v < WPB; ++v) t1[v] = t2[v] = 0.0;
          const double* dAv[WPB];
#pragma unroll
          for (int v = 0; v < WPB; ++v) dAv[v] = smem + (size_t)v * lds_per_wave + 3 * ne + 2 * nu * nu + na;
          if (s.j3_sym) {
#pragma unroll 4
            for (int b = 0; b < na; ++b) {
              const double jj = s.j3[(int64_t)b * na + a];
#pragma unroll
              for (int v = 0; v < WPB; ++v)
                if (acc_flag[v]) t2[v] += jj * dAv[v][b];
            }
#pragma unroll
            for (int v = 0; v < WPB; ++v) t1[v] = t2[v];
          } else {
#pragma unroll 4
            for (int b = 0; b < na; ++b) {
              const double jt = s.j3T[(int64_t)b * na + a];
              const double jj = s.j3[(int64_t)b * na + a];
#pragma unroll
              for (int v = 0; v < WPB; ++v) {
                if (acc_flag[v]) {
                  const double db = dAv[v][b];
                  t1[v] += jt * db;
                  t2[v] += jj * db;
                }
              }
            }
          }
#pragma unroll
          for (int v = 0; v < WPB; ++v) {
            if (acc_flag[v]) {
              double* T1 = smem + (size_t)v * lds_per_wave + 3 * ne + 2 * nu * nu + 2 * na;
              T1[a] = t1[v];
              T1[na + a] = t2[v];
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
        double* Aem = Aw + (int64_t)e * na;
        for (int a = lane; a < na; a += WAVE) Aem[a] = L.anew[a];
        double* Pem = Pw + (int64_t)e * no;
        double* Qem = Qw + (int64_t)e * no;
        const double* Lm = is_up ? s.lam_p : s.lam_pT;
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
          // the end of the sweep), the mover's own stored field is compensated, and only
          // same-spin f < e take the individual correction t1 - t2 (zero for symmetric J3).
          const int f0 = is_up ? 0 : nu;
          for (int a = lane; a < na; a += WAVE) {
            const double t1 = L.t1[a], t2 = L.t2[a];
            double* gcol = gw + a;
            gcol[(int64_t)e * na] -= t2;
            L.Oup[a] += is_up ? t2 : t1;
            L.Odn[a] += t2;
            if (!s.j3_sym) {
              const double d = t1 - t2;
#pragma unroll 4
              for (int f = f0; f < e; ++f) gcol[(int64_t)f * na] += d;
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
        const double ou = L.Oup[a], od = L.Odn[a];
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

// ============================================================ local energy

struct ElLds {
  double *rr, *Gi, *q, *RT, *gk;
  int* plist;
  __device__ ElLds(double* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up;
    rr = base;
    Gi = rr + 3 * ne;
    q = Gi + nu * nu;
    RT = q + (s.n_orb > s.n_ao ? s.n_orb : s.n_ao);
    gk = RT + 9;                                  // rotated mesh points (nv, 3)
    plist = reinterpret_cast<int*>(gk + 3 * (s.nv > 0 ? s.nv : 1));  // active electrons (ne)
  }
};

__host__ __device__ inline int64_t el_lds_doubles(int n_el, int n_up, int n_ao, int n_orb, int nv) {
  return 3 * n_el + n_up * n_up + (n_orb > n_ao ? n_orb : n_ao) + 9 + 3 * (nv > 0 ? nv : 1) + (n_el + 1) / 2 + 1;
}

// Determinant fields and continuum kinetic energies of every electron, one wavefront.
// Per electron e the determinant enters through h_e = C^T q_e with
// q_e = [M | Lambda_u] G^-1[:, e] (up) or M' G^-1[e, :]^T (dn), so that
// grad_e ln det = grad a(r_e) . h_e and the Laplacian term is lap a(r_e) . h_e
// (compute_grads_and_laplacian_ln_Det_fast, determinant.py:2140-2250); h_e is stored in Hw for
// the single-electron ratios a(r') . h_e (determinant.py:1665-1783).  T_e =
// -1/2 (lap ln Psi + |grad ln Psi|^2) (compute_kinetic_energy_all_elements_fast_update,
// wavefunction.py:1141-1207) goes to te[e] / t_elec[e] when given; returns sum_e T_e.
__device__ double kinetic_fields(const Sys& s, const double* rr, const double* Gi, double* q, const double* Aw,
                                 const double* Qw, const double* gw, double* Hw, double* te, double* t_elec,
                                 int lane) {
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  double kinetic = 0.0;
  for (int e = 0; e < ne; ++e) {
    const bool up = e < nu;
    for (int k = lane; k < no; k += WAVE) {
      double acc = 0.0;
      if (up) {
        for (int j = 0; j < nd; ++j) acc += Qw[(int64_t)(nu + j) * no + k] * Gi[j * nu + e];
        for (int u = 0; u < s.n_unp; ++u) acc += s.lam_u[(int64_t)k * s.n_unp + u] * Gi[(nd + u) * nu + e];
      } else {
        const int j = e - nu;
        for (int i = 0; i < nu; ++i) acc += Qw[(int64_t)i * no + k] * Gi[j * nu + i];
      }
      q[k] = acc;
    }
    wave_sync();
    const double* C = up ? s.mo_up : s.mo_dn;
    const double ex = rr[3 * e], ey = rr[3 * e + 1], ez = rr[3 * e + 2];
    double acc8[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // gJ(3), lapJ, gD(3), lapD'
    for (int a = lane; a < na; a += WAVE) {
      double h;
      if (s.orb_is_mo) {
        h = 0.0;
        for (int k = 0; k < no; ++k) h += C[(int64_t)k * na + a] * q[k];
      } else {
        h = q[a];
      }
      Hw[(int64_t)e * na + a] = h;
      double v, g0, g1, g2, lp;
      ao_vgl(s, a, ex, ey, ez, v, g0, g1, g2, lp);
      if (gw) {
        const double gj = gw[(int64_t)e * na + a];
        acc8[0] += gj * g0;
        acc8[1] += gj * g1;
        acc8[2] += gj * g2;
        acc8[3] += gj * lp;
      }
      acc8[4] += h * g0;
      acc8[5] += h * g1;
      acc8[6] += h * g2;
      acc8[7] += h * lp;
    }
    if (s.j1_type) {
      for (int n = lane; n < s.n_nuc; n += WAVE)
        j1_grad_lap(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], ex - s.nuc_pos[3 * n], ey - s.nuc_pos[3 * n + 1],
                    ez - s.nuc_pos[3 * n + 2], acc8[0], acc8[1], acc8[2], acc8[3]);
    }
    if (s.j2_type) {
      for (int f = lane; f < ne; f += WAVE) {
        if (f == e) continue;
        j2_grad_lap(s.j2_type, s.j2_param, ex - rr[3 * f], ey - rr[3 * f + 1], ez - rr[3 * f + 2], acc8[0],
                    acc8[1], acc8[2], acc8[3]);
      }
    }
    wave_sum_k<8>(acc8);
    const double gDx = acc8[4], gDy = acc8[5], gDz = acc8[6];
    const double lapD = acc8[7] - (gDx * gDx + gDy * gDy + gDz * gDz);
    const double tx = acc8[0] + gDx, ty = acc8[1] + gDy, tz = acc8[2] + gDz;
    const double Te = -0.5 * (acc8[3] + lapD + (tx * tx + ty * ty + tz * tz));
    kinetic += Te;
    if (lane == 0) {
      if (te) te[e] = Te;
      if (t_elec) t_elec[e] = Te;
    }
    wave_sync();
  }
  return kinetic;
}

// Single-electron ratio pieces for electron e moved to p, evaluated by ONE lane: the
// determinant ratio a(p) . h_e and the Jastrow log-ratio dJ = (a(p) - a(r_e)) . g_e + dJ1 + dJ2
// (_compute_ratio_determinant_part_split_spin, determinant.py:1665-1783;
// _compute_ratio_Jastrow_part_rank1_update, jastrow_factor.py:2696-2960).
__device__ __forceinline__ void move_ratio(const Sys& s, const double* rr, const double* Ae, const double* He,
                                           const double* ge, int e, double px, double py, double pz, double& det,
                                           double& dj) {
  const int ne = s.n_el, na = s.n_ao;
  const double ex = rr[3 * e], ey = rr[3 * e + 1], ez = rr[3 * e + 2];
  det = 0.0;
  dj = 0.0;
  for (int a = 0; a < na; ++a) {
    const double v = ao_value(s, a, px, py, pz);
    det += v * He[a];
    if (ge) dj += (v - Ae[a]) * ge[a];
  }
  if (s.j1_type) dj += j1_electron(s, px, py, pz) - j1_electron(s, ex, ey, ez);
  if (s.j2_type) {
    for (int f = 0; f < ne; ++f) {
      if (f == e) continue;
      const double* q = rr + 3 * f;
      const double d1x = px - q[0], d1y = py - q[1], d1z = pz - q[2];
      const double d0x = ex - q[0], d0y = ey - q[1], d0z = ez - q[2];
      dj += j2_u(s.j2_type, s.j2_param, sqrt(d1x * d1x + d1y * d1y + d1z * d1z)) -
            j2_u(s.j2_type, s.j2_param, sqrt(d0x * d0x + d0y * d0y + d0z * d0z));
    }
  }
}

// Rotated quadrature points g~_k = g_k RT into gk (nv, 3) (coulomb_potential.py:1546-1547).
__device__ __forceinline__ void rotate_mesh(const Sys& s, const double* RT, double* gk, int lane) {
  for (int k = lane; k < s.nv; k += WAVE) {
    for (int c = 0; c < 3; ++c)
      gk[3 * k + c] =
          s.mesh_pts[3 * k] * RT[c] + s.mesh_pts[3 * k + 1] * RT[3 + c] + s.mesh_pts[3 * k + 2] * RT[6 + c];
  }
}

// Electrons whose nearest nucleus carries projected ECP terms, compacted in order into plist;
// returns their count (identical on every lane).
__device__ __forceinline__ int ecp_active_electrons(const Sys& s, const double* rr, int* plist, int lane) {
  const int ne = s.n_el;
  int n_act = 0;
  for (int base = 0; base < ne; base += WAVE) {
    const int e = base + lane;
    bool act = false;
    if (e < ne) {
      const int n = nearest_nucleus(s, rr[3 * e], rr[3 * e + 1], rr[3 * e + 2]);
      act = s.ecp_nt[n] > 0;
    }
    const uint64_t mask = __ballot(act);
    if (act) {
      const int pos = n_act + __popcll(mask & ((1ull << lane) - 1ull));
      plist[pos] = e;
    }
    n_act += __popcll(mask);
  }
  return n_act;
}

// Quadrature point k of electron e around its nearest nucleus, r' = r + (R - r) + |R - r| g~_k, and
// the projector factor sum_l V_l(|R - r|) (2l + 1) P_l(cos theta) w_k (coulomb_potential.py:1546-1705).
__device__ __forceinline__ void ecp_point(const Sys& s, const double* rr, const double* gk, int e, int k, double& px,
                                          double& py, double& pz, double& wsum) {
  const double ex = rr[3 * e], ey = rr[3 * e + 1], ez = rr[3 * e + 2];
  const int n = nearest_nucleus(s, ex, ey, ez);
  const double rx = s.nuc_pos[3 * n] - ex, ry = s.nuc_pos[3 * n + 1] - ey, rz = s.nuc_pos[3 * n + 2] - ez;
  const double rn = sqrt(rx * rx + ry * ry + rz * rz);
  const double gx = gk[3 * k], gy = gk[3 * k + 1], gz = gk[3 * k + 2];
  const double gn = sqrt(gx * gx + gy * gy + gz * gz);
  px = ex + (rx + rn * gx);
  py = ey + (ry + rn * gy);
  pz = ez + (rz + rn * gz);
  const double cth = (-rx / rn) * (gx / gn) + (-ry / rn) * (gy / gn) + (-rz / rn) * (gz / gn);
  wsum = 0.0;
  for (int t = 0; t < s.ecp_nt[n]; ++t) {
    const int o = n * s.ecp_max_terms + t;
    const int l = s.ecp_l[o];
    const double V = pow(rn, -2.0) * s.ecp_coef[o] * pow(rn, s.ecp_pow[o]) * exp(-s.ecp_exp[o] * (rn * rn));
    wsum += V * ((2 * l + 1) * legendre(l, cth) * s.mesh_w[k]);
  }
}

// Local ECP part of one electron-nucleus distance d (coulomb_potential.py:1144-1246).
__device__ __forceinline__ double ecp_local_pair(const Sys& s, int n, double d) {
  double v = 0.0;
  for (int k = 0; k < s.ecpl_nt[n]; ++k) {
    const int o = n * s.ecp_max_local + k;
    v += pow(d, -2.0) * s.ecpl_coef[o] * pow(d, s.ecpl_pow[o]) * exp(-s.ecpl_exp[o] * (d * d));
  }
  return v;
}

// Lane-partial electron-electron repulsion sum_{e<f} 1/r_ef (coulomb_potential.py:2464-2518).
__device__ __forceinline__ double el_el_partial(const Sys& s, const double* rr, int lane) {
  double v = 0.0;
  const int ne = s.n_el;
  for (int e = lane; e < ne; e += WAVE) {
    for (int f = e + 1; f < ne; ++f) {
      const double dx = rr[3 * e] - rr[3 * f], dy = rr[3 * e + 1] - rr[3 * f + 1], dz = rr[3 * e + 2] - rr[3 * f + 2];
      v += 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  return v;
}

// e_L = T + V per walker, one wavefront per walker (compute_local_energy_fast,
// hamiltonians.py:225-290).
__global__ void __launch_bounds__(256) k_local_energy(Sys s, int64_t W, Walk wk, const double* __restrict__ RTall,
                                                      double* e_L, double* kin_out, double* v_out, double* t_elec,
                                                      double* __restrict__ Hws, int wpb, int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  ElLds L(smem + (size_t)wave * lds_per_wave, s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  const double* Aw = wk.ao + w * ne * na;
  const double* Qw = wk.pair + w * ne * no;
  const double* gw = s.j3_on ? wk.jfield + w * ne * na : nullptr;
  double* Hw = Hws + w * ne * na;

  for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = wk.r_up[w * 3 * nu + k];
  for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
  for (int k = lane; k < nu * nu; k += WAVE) L.Gi[k] = wk.Ginv[w * nu * nu + k];
  if (lane < 9) L.RT[lane] = RTall ? RTall[9 * w + lane] : ((lane % 4) == 0 ? 1.0 : 0.0);
  wave_sync();

  const double kinetic =
      kinetic_fields(s, L.rr, L.Gi, L.q, Aw, Qw, gw, Hw, nullptr, t_elec ? t_elec + w * ne : nullptr, lane);

  // ------------------------------------------------ bare Coulomb + ECP local
  double pot = 0.0;
  for (int t = lane; t < ne * s.n_nuc; t += WAVE) {
    const int e = t / s.n_nuc, n = t - e * s.n_nuc;
    const double dx = L.rr[3 * e] - s.nuc_pos[3 * n], dy = L.rr[3 * e + 1] - s.nuc_pos[3 * n + 1],
                 dz = L.rr[3 * e + 2] - s.nuc_pos[3 * n + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    pot += -s.nuc_charge[n] / d;
    if (s.ecp_on) pot += ecp_local_pair(s, n, d);
  }
  pot += el_el_partial(s, L.rr, lane);
  pot = wave_sum(pot) + s.ion_ion;

  // ------------------------------------------------ ECP non-local (NN = 1)
  if (s.ecp_on && s.nv > 0) {
    rotate_mesh(s, L.RT, L.gk, lane);
    const int n_act = ecp_active_electrons(s, L.rr, L.plist, lane);
    wave_sync();
    const int n_pts = n_act * s.nv;
    double vnl = 0.0;
    for (int base = 0; base < n_pts; base += WAVE) {
      const int p = base + lane;
      const bool active = p < n_pts;
      const int e = active ? L.plist[p / s.nv] : L.plist[0];
      const int k = active ? p % s.nv : 0;
      double px, py, pz, wsum, det, dj;
      ecp_point(s, L.rr, L.gk, e, k, px, py, pz, wsum);
      move_ratio(s, L.rr, Aw + (int64_t)e * na, Hw + (int64_t)e * na, gw ? gw + (int64_t)e * na : nullptr, e, px, py,
                 pz, det, dj);
      if (active) vnl += wsum * (det * exp(dj));
    }
    pot += wave_sum(vnl);
  }

  if (lane == 0) {
    e_L[w] = kinetic + pot;
    if (kin_out) kin_out[w] = kinetic;
    if (v_out) v_out[w] = pot;
  }
}

}  // namespace jqmc
