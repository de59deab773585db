// Move-synchronous Metropolis sweep: the n_moves proposals of a sweep are taken one move at a
// time across all walkers, so that the per-walker dependency chain of a move runs in a launch
// with one wavefront per walker (throughput, not latency, bound), and the J3 field products of
// every walker that accepted (t1 = J3 dA, t2 = J3^T dA) become one tiled GEMM per move with the
// field updates fused into its epilogue.  Per sweep:
//
//   k_sweep_begin     R_AS of the current (G, G^-1) per walker; spin-block field offsets zeroed
//   n_moves times:
//     k_sweep_propose one wavefront per walker: draws, proposal, AO values, Jastrow / determinant
//                     ratios, AS factor, acceptance; on accept the walker state (position, AO /
//                     orbital / pair rows, G, G^-1) is updated in place and dA is left in the
//                     workspace with the moved electron's index
//     k_sweep_jfield_mfma  [T1 | T2] = [J3 | J3^T] x dA on the matrix cores, one block per 64 AO
//                     rows x 32 walkers (16 x 64 tiles of J3 staged in LDS, the next tile prefetched
//                     into registers), then the J3 field increments of jastrow_factor.py:4121-4135
//                     on those rows
//   k_sweep_flush     spin-block offsets added to every field row; R_AS stored
//
// The draws, proposal and acceptance are those of k_vmc_sweep (jqmc_mcmc.py:4316-4531) on the
// same Philox counters, so both kernels take the same decisions.
#pragma once

#include "jqmc_kernels.hip"
#include "jqmc_sr.hip"

namespace jqmc {

// Per-walker workspace of the move-synchronous sweep.
template <typename T>
struct SweepWs {
  T* dA;       // (W, n_ao) AO change of the last proposal
  T* Oup;      // (W, n_ao) field offset of every up electron
  T* Odn;      // (W, n_ao) field offset of every dn electron
  double* Ro;  // (W) AS factor of the current state
  int* mv;     // (W) electron moved by the last move, -1 when rejected
  __host__ __device__ static int64_t bytes(const Sys& s, int64_t W) {
    const int64_t na = s.n_ao;
    int64_t b = 3 * W * na * (int64_t)sizeof(T);
    b = (b + 7) & ~(int64_t)7;
    return b + W * (int64_t)sizeof(double) + W * (int64_t)sizeof(int);
  }
  SweepWs() = default;
  __host__ __device__ SweepWs(void* base, const Sys& s, int64_t W) {
    const int64_t na = s.n_ao;
    char* p = static_cast<char*>(base);
    dA = reinterpret_cast<T*>(p);
    Oup = dA + W * na;
    Odn = Oup + W * na;
    int64_t off = (3 * W * na * (int64_t)sizeof(T) + 7) & ~(int64_t)7;
    Ro = reinterpret_cast<double*>(p + off);
    mv = reinterpret_cast<int*>(p + off + W * (int64_t)sizeof(double));
  }
};

// dst[i] = src[i] (+ src2[i]) for i < n by one wavefront: each lane issues its U loads before any
// of its LDS stores, so the loads of a row are in flight together (LDS is reached through generic
// pointers here, which the compiler may not reorder around).
template <typename T, int U = 4>
__device__ __forceinline__ void stage_row(T* dst, const T* __restrict__ src, const T* __restrict__ src2, int n,
                                          int lane) {
  for (int base = 0; base < n; base += WAVE * U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + lane + WAVE * u;
      v[u] = i < n ? (src2 ? src[i] + src2[i] : src[i]) : (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + lane + WAVE * u;
      if (i < n) dst[i] = v[u];
    }
  }
}

// Positions (up then dn) and G^-1 into LDS by one wavefront, with every load of a U x 64 chunk of
// both issued before its stores: one memory round trip per chunk, one chunk for systems up to
// U x 64 elements per array.
template <int U = 4>
__device__ __forceinline__ void stage_walker(double* rr, const double* __restrict__ rup, int n_rup,
                                             const double* __restrict__ rdn, int n_rdn, double* gi,
                                             const double* __restrict__ Gi, int n_g, int lane) {
  const int n_r = n_rup + n_rdn;
  const int n_max = n_r > n_g ? n_r : n_g;
  for (int base = 0; base < n_max; base += WAVE * U) {
    double vr[U], vi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + lane + WAVE * u;
      vr[u] = i < n_rup ? rup[i] : (i < n_r ? rdn[i - n_rup] : 0.0);
      vi[u] = i < n_g ? Gi[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + lane + WAVE * u;
      if (i < n_r) rr[i] = vr[u];
      if (i < n_g) gi[i] = vi[u];
    }
  }
}

// Elements of the mover's AO row and J3 field held in registers per lane (a = lane + 64 u); rows
// longer than 64 x PROP_RU are read from memory for the remaining entries.
constexpr int PROP_RU = 4;

// LDS of one wavefront in k_sweep_begin / k_sweep_propose (member names as SweepLds for the
// AS-factor helper): positions and G^-1, staged once per move, and the new AO and orbital rows.  G
// (Gm) and the pair rows of the other spin (pr) are read where they lie in the walker state (the
// kernels set the two pointers), which leaves room in LDS for more walkers per CU.  The moving
// electron's AO row and J3 field stay in registers.
template <typename T>
struct PropLds {
  double *rr, *Gi, *Gm, *vv, *xx, *yy, *phin, *dphi, *pr;
  T* anew;
  uintptr_t end;
  __host__ __device__ PropLds(char* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up, na = s.n_ao, no = s.n_orb;
    Carve c{base};
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(nu * nu);
    Gm = nullptr;
    vv = c.take<double>(nu);
    xx = c.take<double>(nu);
    yy = c.take<double>(nu);
    anew = c.take<T>(na);
    phin = c.take<double>(no);
    dphi = c.take<double>(no);
    pr = nullptr;
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

template <typename T>
inline int64_t prop_lds_bytes(const Sys& s) {
  return (int64_t)PropLds<T>(0, s).end;
}

template <typename T>
__global__ void __launch_bounds__(256) k_sweep_begin(Sys s, int64_t W, WalkT<T> wk,
                                                     SweepWs<T> ws, int wpb, int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  PropLds<T> L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  const int nu = s.n_up, na = s.n_ao;
  L.Gm = wk.G + w * nu * nu;
  for (int k = lane; k < nu * nu; k += WAVE) L.Gi[k] = wk.Ginv[w * nu * nu + k];
  for (int a = lane; a < na; a += WAVE) ws.Oup[w * na + a] = ws.Odn[w * na + a] = 0;
  wave_sync();
  const double R = sweep_as_factor(L, nu, lane, 0, 0, 1.0);
  if (lane == 0) {
    ws.Ro[w] = R;
    ws.mv[w] = -1;
  }
}

// Move m of the sweep for every walker (the body of k_vmc_sweep's move loop).
// Registers capped at 128 (4 waves per SIMD, a few bytes of spills): the per-walker chain is
// latency bound, and the LDS of 4 walkers per block allows 16 waves per CU.
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_sweep_propose(Sys s, int64_t W, WalkT<T> wk,
                                                       SweepWs<T> ws, int m, double Dt, double eps_AS, uint64_t seed,
                                                       int64_t move_base, const int64_t* __restrict__ move_ctr,
                                                       int wpb, int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const RealTabs<T>& R = tabs<T>(s);
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  PropLds<T> L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  T* Aw = wk.ao + w * ne * na;
  double* Pw = wk.orb + w * ne * no;
  double* Qw = wk.pair + w * ne * no;
  const T* gw = s.j3_on ? wk.jfield + w * ne * na : nullptr;
  T* dAw = ws.dA + w * na;
  const uint64_t mvc = (uint64_t)(move_base + (move_ctr ? *move_ctr : 0) + m);

  // ---- draws in the reference order (jqmc_mcmc.py:4322-4500), before any walker load
  const U4 d0 = draw_words(seed, (uint32_t)w, mvc, STREAM_VMC, 0);
  const U4 d1 = draw_words(seed, (uint32_t)w, mvc, STREAM_VMC, 1);
  const U4 d2 = draw_words(seed, (uint32_t)w, mvc, STREAM_VMC, 2);
  const int rand_num = randint(d0.x, ne);
  const int up_index = randint(d0.y, nu > 0 ? nu : 1);
  const int dn_index = randint(d0.z, nd > 0 ? nd : 1);
  const int axis = randint(d0.w, 3);
  const double u1 = uniform01_open0(d1.x, d1.y), u2 = uniform01(d1.z, d1.w);
  const double gauss = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  const double bdraw = uniform01(d2.x, d2.y);
  const bool is_up = rand_num < nu;
  const int idx = is_up ? up_index : dn_index;
  const int e = is_up ? idx : nu + idx;

  // ---- load the walker data this move reads: the mover's AO row and J3 field (with its block
  // offset) into registers, positions and G^-1 into LDS, all loads in flight together; G and the
  // pair rows of the other spin (dn rows for an up move, up rows for a dn move) are read in place
  const T* Ae = Aw + (int64_t)e * na;
  const T* ge = gw ? gw + (int64_t)e * na : nullptr;
  const T* Oe = (is_up ? ws.Oup : ws.Odn) + w * na;
  T ae_r[PROP_RU], ge_r[PROP_RU];
#pragma unroll
  for (int u = 0; u < PROP_RU; ++u) {
    const int a = lane + WAVE * u;
    ae_r[u] = a < na ? Ae[a] : (T)0;
    ge_r[u] = (gw && a < na) ? ge[a] + Oe[a] : (T)0;
  }
  L.Gm = wk.G + w * nu * nu;
  L.pr = Qw + (is_up ? (int64_t)nu * no : 0);
  stage_walker(L.rr, wk.r_up + w * 3 * nu, 3 * nu, wk.r_dn + w * 3 * nd, 3 * nd, L.Gi, wk.Ginv + w * nu * nu,
               nu * nu, lane);
  const double R_o = ws.Ro[w];
  wave_sync();

  // ---- proposal with nearest-nucleus step size (jqmc_mcmc.py:4339-4401)
  const double ox = L.rr[3 * e], oy = L.rr[3 * e + 1], oz = L.rr[3 * e + 2];
  const int na0 = nearest_nucleus(s, ox, oy, oz);
  const double Z0 = s.nuc_charge[na0];
  double dxn = ox - s.nuc_pos[3 * na0], dyn = oy - s.nuc_pos[3 * na0 + 1], dzn = oz - s.nuc_pos[3 * na0 + 2];
  const double f_l = step_factor(Z0, sqrt(dxn * dxn + dyn * dyn + dzn * dzn));
  const double g = gauss * (f_l * Dt);
  double nx = ox, ny = oy, nz = oz;
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
  for (int sh = lane; sh < s.n_shell; sh += WAVE) shell_values<T>(s, sh, nx, ny, nz, L.anew);
  wave_sync();
  double jpart = 0.0;
#pragma unroll
  for (int u = 0; u < PROP_RU; ++u) {
    const int a = lane + WAVE * u;
    if (a < na) {
      const T da = L.anew[a] - ae_r[u];
      dAw[a] = da;
      if (gw) jpart += (double)(da * ge_r[u]);
    }
  }
  for (int a = lane + WAVE * PROP_RU; a < na; a += WAVE) {
    const T da = L.anew[a] - Ae[a];
    dAw[a] = da;
    if (gw) jpart += (double)(da * (ge[a] + Oe[a]));
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
        const double* pr = L.pr + (int64_t)j * no;
        for (int k = 0; k < no; ++k) v += L.dphi[k] * pr[k];
      } else {
        for (int k = 0; k < no; ++k) v += L.dphi[k] * s.td.lam_u[(int64_t)k * s.n_unp + (j - nd)];
      }
      L.vv[j] = v;
      part += v * L.Gi[j * nu + idx];
    }
  } else {
    for (int i = lane; i < nu; i += WAVE) {
      const double* pr = L.pr + (int64_t)i * no;
      double u = 0.0;
      for (int k = 0; k < no; ++k) u += pr[k] * L.dphi[k];
      L.vv[i] = u;
      part += L.Gi[idx * nu + i] * u;
    }
  }
  const double det_ratio = 1.0 + wave_sum(part);
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

  // ---- AS-regularised acceptance (jqmc_mcmc.py:4484-4520)
  const double R_p = sweep_as_factor(L, nu, lane, is_up ? 1 : 2, idx, det_ratio);
  const double R_AS_ratio = (fmax(R_p, eps_AS) / R_p) / (fmax(R_o, eps_AS) / R_o);
  const double wr = R_AS_ratio * J_ratio * det_ratio;
  const bool accepted = bdraw < fmin(1.0, wr * wr * T_ratio);

  if (accepted) {
    if (lane == 0) {
      double* rp = is_up ? wk.r_up + w * 3 * nu + 3 * idx : wk.r_dn + w * 3 * nd + 3 * idx;
      rp[0] = nx;
      rp[1] = ny;
      rp[2] = nz;
      ws.Ro[w] = R_p;
      ws.mv[w] = e;
      wk.accepted[w] += 1;
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
    for (int t = lane; t < nu * nu; t += WAVE) {
      const int p = t / nu, q = t - p * nu;
      wk.Ginv[w * nu * nu + t] = L.Gi[t] - L.xx[p] * L.yy[q] / det_ratio;
      double gm = L.Gm[t];
      if (is_up && p == idx) gm += L.vv[q];
      if (!is_up && q == idx) gm += L.vv[p];
      wk.G[w * nu * nu + t] = gm;
    }
  } else if (lane == 0) {
    ws.mv[w] = -1;
    wk.rejected[w] += 1;
  }
}

constexpr int JF_WALKERS = 32;  // walkers per block of k_sweep_jfield_mfma
constexpr int JF_ROWS = 64;     // AO rows per tile
constexpr int JF_KC = 16;       // AO columns per staged J3 tile
// Row strides of the LDS tiles, padded so that the four 16-lane groups of a wave, which read four
// consecutive tile rows, start 16 banks apart (64 banks of 4 bytes): 72 and 40 doubles.
constexpr int JF_ROWS_LDS = JF_ROWS + 8;
constexpr int JF_WALKERS_LDS = JF_WALKERS + 8;

// Matrix-core step of the J3 field products: D(16 x 16) += A(16 x 4) B(4 x 16) with A(i, k) from lane
// i + 16 k and B(k, j) from lane j + 16 k; row(q, ih) is the tile row that element q of lane
// il + 16 ih holds (column il).  FP64: v_mfma_f64_16x16x4f64, D(i, j) in lane j + 16 (i % 4), element
// i / 4.  FP32 (mixed form): v_mfma_f32_16x16x4f32, D(i, j) in lane j + 16 (i / 4), element i % 4.
template <typename T>
struct JfMfma;
template <>
struct JfMfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int ih) { return 4 * q + ih; }
};
template <>
struct JfMfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int ih) { return 4 * ih + q; }
};

// [T1 | T2] = [J3 dA | J3^T dA] on AO rows a0..a0+63 (a0 = 64 blockIdx.x) for the 32 walkers
// blockIdx.y of the block that accepted, then the field increments on those rows: an up move of
// e gives t1 to up f < e and t2 to every other electron; a dn move gives t1 to every up electron
// and to dn f < e, t2 to dn f > e.  The part shared by a spin block goes to its offset, the mover's
// own row is compensated, and the side of the mover's block with fewer electrons takes the
// individual correction (none for a symmetric J3).  With kDirect (the LRDMC move, whose fields are
// read by the next Hamiltonian launch), every row f != e of a moved walker takes its increment
// instead: t1 for f < e, t2 for f > e, no offsets.  The products run on the matrix cores as
// D(walker, a) = sum_b dA(walker, b) J(b, a): A(i, k) = dA of walker 16 t + i at AO
// b0 + k (LDS da_s), B(k, j) = J3^T or J3 at (b0 + k, AO a0 + 16 wave + j) (LDS jt_s / jj_s).  Wave w
// owns the 16 AOs a0 + 16 w .. + 15 of all 32 walkers (two 16-walker tiles, T1 and T2 each), so in
// the epilogue 16 consecutive lanes touch 16 consecutive AOs of one walker's row.
template <typename T, bool kDirect = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_sweep_jfield_mfma(Sys s, int64_t W, WalkT<T> wk, SweepWs<T> ws) {
  using M = JfMfma<T>;
  typedef typename M::v4 v4;
  const RealTabs<T>& R = tabs<T>(s);
  __shared__ __align__(16) T jt_s[JF_KC][JF_ROWS_LDS];
  __shared__ __align__(16) T jj_s[JF_KC][JF_ROWS_LDS];
  __shared__ __align__(16) T da_s[JF_KC][JF_WALKERS_LDS];
  __shared__ int mv_s[JF_WALKERS];
  __shared__ int any_s;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int a0 = (int)blockIdx.x * JF_ROWS;
  const int64_t w0 = (int64_t)blockIdx.y * JF_WALKERS;
  const int nu = s.n_up, ne = s.n_el, na = s.n_ao;
  if (t == 0) any_s = 0;
  __syncthreads();
  if (t < JF_WALKERS) {
    const int64_t w = w0 + t;
    const int e = w < W ? ws.mv[w] : -1;
    mv_s[t] = e;
    if (e >= 0) atomicOr(&any_s, 1);
  }
  __syncthreads();
  if (!any_s) return;  // uniform across the block
  const bool sym = s.j3_sym != 0;
  const T* __restrict__ JT = R.j3T;
  const T* __restrict__ JJ = R.j3;
  const T* __restrict__ DA = ws.dA;
  T pjt[4], pjj[4], pda[2];
  auto fetch = [&](int b0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = t + 256 * r;
      const int bb = k / JF_ROWS, aa = k - bb * JF_ROWS;
      const int b = b0 + bb, a = a0 + aa;
      const bool ok = b < na && a < na;
      pjt[r] = ok ? JT[(int64_t)b * na + a] : (T)0;
      pjj[r] = (ok && !sym) ? JJ[(int64_t)b * na + a] : (T)0;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = t + 256 * r;
      const int ww = k / JF_KC, bb = k - ww * JF_KC;
      const int b = b0 + bb;
      pda[r] = (mv_s[ww] >= 0 && b < na) ? DA[(w0 + ww) * na + b] : (T)0;
    }
  };
  v4 acc1[2], acc2[2];
#pragma unroll
  for (int x = 0; x < 2; ++x) acc1[x] = acc2[x] = v4{0, 0, 0, 0};
  const int il = lane & 15, ih = lane >> 4;
  fetch(0);
  for (int b0 = 0; b0 < na; b0 += JF_KC) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = t + 256 * r;
      const int bb = k / JF_ROWS, aa = k - bb * JF_ROWS;
      jt_s[bb][aa] = pjt[r];
      jj_s[bb][aa] = pjj[r];
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = t + 256 * r;
      const int ww = k / JF_KC, bb = k - ww * JF_KC;
      da_s[bb][ww] = pda[r];
    }
    __syncthreads();
    if (b0 + JF_KC < na) fetch(b0 + JF_KC);  // next stage in flight during this one
#pragma unroll
    for (int kk = 0; kk < JF_KC; kk += 4) {
      const int kl = kk + ih;
      const T fj = jt_s[kl][16 * wave + il];
      const T fjj = sym ? fj : jj_s[kl][16 * wave + il];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const T fa = da_s[kl][16 * x + il];
        acc1[x] = M::mma(fa, fj, acc1[x]);
        acc2[x] = M::mma(fa, fjj, acc2[x]);
      }
    }
    __syncthreads();
  }
  // epilogue: pair (x, q) of this lane is walker 16 x + row(q, ih) of the block and AO a0 + 16 wave + il
  const int a = a0 + 16 * wave + il;
  const bool aok = a < na;
  if constexpr (kDirect) {
    if (!aok) return;  // no barrier follows
    int ee[8];
    T t1[8], t2[8];
    T* col[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int x = p >> 2, q = p & 3, r = 16 * x + M::row(q, ih);
      ee[p] = mv_s[r];
      t1[p] = acc1[x][q];
      t2[p] = acc2[x][q];
      col[p] = ee[p] >= 0 ? wk.jfield + (w0 + r) * ne * na + a : wk.jfield;
    }
    // two rows of the 8 walkers per pass, all 16 loads in flight together
    for (int f0 = 0; f0 < ne; f0 += 2) {
      T v[2][8];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int p = 0; p < 8; ++p) {
          const int f = f0 + h;
          v[h][p] = (f < ne && ee[p] >= 0 && f != ee[p]) ? col[p][(int64_t)f * na] : (T)0;
        }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int p = 0; p < 8; ++p) {
          const int f = f0 + h;
          if (f < ne && ee[p] >= 0 && f != ee[p]) col[p][(int64_t)f * na] = v[h][p] + (f < ee[p] ? t1[p] : t2[p]);
        }
    }
    return;
  }
  int ee[8], cs[8], cn[8];
  bool upm[8], lw[8];
  T d8[8], t1[8], t2[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int x = p >> 2, q = p & 3;
    const int e = mv_s[16 * x + M::row(q, ih)];
    ee[p] = e;
    upm[p] = e < nu;
    const int f0 = upm[p] ? 0 : nu, f1 = upm[p] ? nu : ne;
    lw[p] = sym || (e - f0) <= (f1 - 1 - e);
    cs[p] = 0;
    cn[p] = 0;
    if (e >= 0 && !sym) {
      cs[p] = lw[p] ? f0 : e + 1;
      cn[p] = (lw[p] ? e : f1) - cs[p];
    }
    t1[p] = acc1[x][q];
    t2[p] = acc2[x][q];
    d8[p] = lw[p] ? t1[p] - t2[p] : t2[p] - t1[p];
  }
  {
    T go[8], ou[8], od[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int64_t w = w0 + 16 * (p >> 2) + M::row(p & 3, ih);
      const bool on = ee[p] >= 0 && aok;
      const int64_t row = w * ne * na + (int64_t)(ee[p] > 0 ? ee[p] : 0) * na;
      go[p] = on ? wk.jfield[row + a] : (T)0;
      ou[p] = on ? ws.Oup[w * na + a] : (T)0;
      od[p] = on ? ws.Odn[w * na + a] : (T)0;
    }
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      if (ee[p] < 0 || !aok) continue;
      const int64_t w = w0 + 16 * (p >> 2) + M::row(p & 3, ih);
      const int64_t row = w * ne * na + (int64_t)ee[p] * na;
      const T own = lw[p] ? t2[p] : t1[p];
      wk.jfield[row + a] = go[p] - own;
      ws.Oup[w * na + a] = ou[p] + (upm[p] ? own : t1[p]);
      ws.Odn[w * na + a] = od[p] + (upm[p] ? t2[p] : own);
    }
  }
  // individual corrections: rows cs .. cs + cn - 1 of each walker get d, two rows per pass with
  // all their loads in flight together
  int nmax = 0;
#pragma unroll
  for (int p = 0; p < 8; ++p) nmax = cn[p] > nmax ? cn[p] : nmax;
  for (int r = 0; r < nmax; r += 2) {
    T v[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        const bool on = r + h < cn[p] && aok;
        const int64_t w = w0 + 16 * (p >> 2) + M::row(p & 3, ih);
        v[h][p] = on ? wk.jfield[w * ne * na + (int64_t)(cs[p] + r + h) * na + a] : (T)0;
      }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        if (r + h < cn[p] && aok) {
          const int64_t w = w0 + 16 * (p >> 2) + M::row(p & 3, ih);
          wk.jfield[w * ne * na + (int64_t)(cs[p] + r + h) * na + a] = v[h][p] + d8[p];
        }
      }
  }
}

// Offsets into every field row and the AS factor into the walker state (one thread per element).
template <typename T>
__global__ void k_sweep_flush(int64_t W, WalkT<T> wk, SweepWs<T> ws, int n_el,
                              int n_ao, int n_up) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_w = (int64_t)n_el * n_ao;
  if (idx < W * per_w && wk.jfield) {
    const int64_t w = idx / per_w;
    const int64_t r = idx - w * per_w;
    const int f = (int)(r / n_ao), a = (int)(r - (int64_t)f * n_ao);
    wk.jfield[idx] += f < n_up ? ws.Oup[w * n_ao + a] : ws.Odn[w * n_ao + a];
  }
  if (idx < W && wk.as_factor) wk.as_factor[idx] = ws.Ro[idx];
}

}  // namespace jqmc
