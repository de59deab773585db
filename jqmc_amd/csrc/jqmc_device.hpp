// Device-side helpers for the gfx950 jQMC hot path: system pack, wave64
// reductions, Philox4x32-10, contracted Cartesian GTO evaluation and the
// Jastrow / ECP pair functions.  Every kernel maps one walker to one 64-lane
// wavefront; lanes cooperate through LDS and cross-lane shuffles only, never
// through block-wide barriers, so walkers diverge freely.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jqmc {

constexpr int WAVE = 64;

// Real-valued tables in the arithmetic type T of the AO / orbital / J3 work: FP64 always, an FP32
// copy as well when the system runs in mixed precision.
template <typename T>
struct RealTabs {
  const T* prim_exp;   // (n_prim)
  const T* prim_coef;  // (n_prim)
  const T* term_coef;  // (n_terms) polynomial coefficient of each monomial
  const T* mo_up;      // (n_orb, n_ao)
  const T* mo_dn;      // (n_orb, n_ao)
  const T* mo_upT;     // (n_ao, n_orb)
  const T* mo_dnT;     // (n_ao, n_orb)
  const T* lam_p;      // (n_orb, n_orb)
  const T* lam_pT;     // (n_orb, n_orb)
  const T* lam_u;      // (n_orb, n_unp)
  const T* j3;         // (n_ao, n_ao) square block
  const T* j3T;        // (n_ao, n_ao) its transpose
  const T* j3v;        // (n_ao) last column of j_matrix
  const T* ao_scale;   // (n_ao) AO coefficients / those of the first AO of its shell
};

// Device copy of the flattened Hamiltonian (host layout in include/jqmc_hip.h).
// Shell records: SHREC doubles per shell, read at a wave-uniform shell index with a few wide scalar
// loads and one wait: [0..2] centre, [3] n_prim | l << 4 | fast << 8 | first AO << 9 (an exact
// integer), then SHR_PMAX (exponent, coefficient) pairs and SHR_AMAX AO scales, zero-padded.  fast: at
// most SHR_PMAX primitives and the Cartesian AOs of l <= 2 in canonical order (1; x, y, z; xx, xy,
// xz, yy, yz, zz) as unit monomials; the other shells keep only the centre and the info word and are
// evaluated from the per-AO tables.
constexpr int SHR_PMAX = 4, SHR_AMAX = 6;
constexpr int SHR_P = 4, SHR_SC = SHR_P + 2 * SHR_PMAX, SHREC = 20;

struct Sys {
  int n_nuc, n_ao, n_prim, n_up, n_dn, n_el, n_orb, n_unp;
  int orb_is_mo, j1_type, j2_type, j3_on, j3_sym, precision;
  int ecp_on, ecp_lmax, ecp_max_terms, ecp_max_local, nv;
  double j1_param, j2_param, ion_ion;
  const double* nuc_pos;     // (n_nuc, 3)
  const double* nuc_charge;  // (n_nuc)
  const double* j1_c;        // (n_nuc) (2 Z_eff)^(1/4)
  const double* j1_A;        // (n_nuc) (2 Z_eff)^(3/4)
  const double* ao_ctr;      // (n_ao, 3)
  const int* ao_tp;          // (n_ao + 1) CSR offsets into the monomial terms
  const int* term_pow;       // (n_terms, 3) nx, ny, nz
  const int* ao_pp;          // (n_ao + 1)
  const int* ao_nuc;         // (n_ao) nucleus of each AO
  int n_shell;
  const int* sh_ao0;         // (n_shell + 1) AOs of shell k are sh_ao0[k] .. sh_ao0[k + 1] - 1
  const int* ao_sh;          // (n_ao) shell of each AO
  const int* ao_meta;        // (n_ao) shell | nx << 16 | ny << 20 | nz << 24 | AO_UNIT_TERM
  const double* shrec;       // (n_shell, SHREC) shell records (below)
  RealTabs<double> td;
  RealTabs<float> tf;        // valid when precision is mixed
  const int* ecp_nt;         // (n_nuc)
  const int* ecp_l;          // (n_nuc, ecp_max_terms)
  const double* ecp_exp;
  const double* ecp_coef;
  const double* ecp_pow;
  const int* ecpl_nt;        // (n_nuc)
  const double* ecpl_exp;    // (n_nuc, ecp_max_local)
  const double* ecpl_coef;
  const double* ecpl_pow;
  const double* mesh_pts;    // (nv, 3)
  const double* mesh_w;      // (nv)
};

template <typename T>
__host__ __device__ __forceinline__ const RealTabs<T>& tabs(const Sys& s);
template <>
__host__ __device__ __forceinline__ const RealTabs<double>& tabs<double>(const Sys& s) {
  return s.td;
}
template <>
__host__ __device__ __forceinline__ const RealTabs<float>& tabs<float>(const Sys& s) {
  return s.tf;
}

// Walker state pointers (mirror of jqmc_walkers in include/jqmc_hip.h).  The AO and J3-field
// caches are in the system's arithmetic type T (FP32 in the mixed form); the orbital and pair
// caches, which feed G, its inverse and every determinant ratio, are FP64 in both forms.
template <typename T>
struct WalkT {
  double* r_up;
  double* r_dn;
  T* ao;
  double* orb;
  double* pair;
  T* jfield;
  double* G;
  double* Ginv;
  double* logabsdet;
  double* sign;
  double* jastrow;
  double* as_factor;
  int64_t* accepted;
  int64_t* rejected;
};

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

// Index of the calling wavefront in its block, in a scalar register: the work derived from it
// (walker index, per-walker addresses, the Philox counters of a walker) is then scalar ALU work
// instead of the same computation repeated in 64 lanes.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE); }

// Makes this wavefront's earlier LDS and global writes visible to all of its lanes:
// release/acquire fences at workgroup scope (outstanding memory operations complete,
// no compiler reordering across the point) around the wavefront barrier.  Only lanes
// of the calling wavefront synchronise, so walkers in one block never wait on each other.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Butterfly sum: every lane ends with the bitwise-identical total.
// A value every lane holds alike (an LDS read at a wave-uniform address), moved to scalar registers
// so that it occupies no vector registers in the loops that use it.
__device__ __forceinline__ double wave_uniform(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b), hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, WAVE));
  return v;
}

// Sum of K values per lane, reduced across the wavefront (all lanes get the totals).
template <int K>
__device__ __forceinline__ void wave_sum_k(double* v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], off, WAVE);
  }
}

// ---------------------------------------------------------------- Philox4x32-10

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

// Counter layout (walker, move_lo, move_hi, stream << 16 | sub); key = 64-bit seed.
__device__ __forceinline__ U4 draw_words(uint64_t seed, uint32_t walker, uint64_t move, uint32_t stream,
                                         uint32_t sub) {
  U4 c{walker, (uint32_t)move, (uint32_t)(move >> 32), (stream << 16) | sub};
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ __forceinline__ uint64_t u53(uint32_t hi, uint32_t lo) {
  return ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
}
__device__ __forceinline__ double uniform01(uint32_t hi, uint32_t lo) {
  return (double)u53(hi, lo) * 0x1.0p-53;
}
__device__ __forceinline__ double uniform01_open0(uint32_t hi, uint32_t lo) {
  return ((double)u53(hi, lo) + 1.0) * 0x1.0p-53;
}
__device__ __forceinline__ int randint(uint32_t w, int n) { return (int)(((uint64_t)w * (uint64_t)n) >> 32); }

constexpr uint32_t STREAM_VMC = 0, STREAM_ROT = 1, STREAM_LRDMC = 2, STREAM_LRDMC_ROT = 3;

// R^T of _generate_rotation_matrix (jqmc_mcmc.py:4228-4245) from three angles in U(-2pi, 2pi).
__device__ inline void rotation_RT(uint64_t seed, uint32_t walker, uint64_t step, uint32_t stream, double* RT) {
  const U4 a = draw_words(seed, walker, step, stream, 0);
  const U4 b = draw_words(seed, walker, step, stream, 1);
  const double lo = -2.0 * M_PI, span = 4.0 * M_PI;
  const double al = lo + span * uniform01(a.x, a.y);
  const double be = lo + span * uniform01(a.z, a.w);
  const double ga = lo + span * uniform01(b.x, b.y);
  const double ca = cos(al), sa = sin(al), cb = cos(be), sb = sin(be), cg = cos(ga), sg = sin(ga);
  const double R[9] = {cb * cg, cg * sa * sb - ca * sg, sa * sg + ca * cg * sb,
                       cb * sg, ca * cg + sa * sb * sg, ca * sb * sg - cg * sa,
                       -sb,     cb * sa,                ca * cb};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) RT[3 * i + j] = R[3 * j + i];
}

// ---------------------------------------------------------------- e^x for x <= 0

// a b + c by v_fma_f64 with three vector operands.  Used where c is a loop-invariant constant: the
// compiler otherwise selects the two-address v_fmac_f64 and copies the constant into the accumulator
// before every step.
__device__ __forceinline__ double fma_vvv(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
}

// e^x for x <= 0 (the Gaussian, ECP and Jastrow exponentials): Cody-Waite reduction
// x = k ln2 + r, |r| <= ln2 / 2, with ln2 split into its double L1 and the remainder L2 (each
// product taken by one fma), then e^r by its Taylor polynomial of degree 12 in Horner form
// (truncation below 2e-16 relative), scaled by 2^k with ldexp; arguments below -800 give 0.
// Within 2 ulp of the correctly rounded value, in 18 instructions.
__device__ __forceinline__ double exp_nonpos(double x) {
  x = fmax(x, -800.0);
  const double k = rint(x * 1.4426950408889634);  // 1 / ln2
  double r = fma(-k, 0.6931471805599453, x);       // L1
  r = fma(-k, 2.3190468138462996e-17, r);          // L2
  double p = 2.08767569878681e-09;                 // 1 / 12!
  p = fma_vvv(p, r, 2.505210838544172e-08);
  p = fma_vvv(p, r, 2.755731922398589e-07);
  p = fma_vvv(p, r, 2.7557319223985893e-06);
  p = fma_vvv(p, r, 2.48015873015873e-05);
  p = fma_vvv(p, r, 0.0001984126984126984);
  p = fma_vvv(p, r, 0.001388888888888889);
  p = fma_vvv(p, r, 0.008333333333333333);
  p = fma_vvv(p, r, 0.041666666666666664);
  p = fma_vvv(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}
// FP32: the hardware exponential
__device__ __forceinline__ float exp_nonpos(float x) { return exp(x); }

// ---------------------------------------------------------------- AOs

// x^n by left-to-right multiplication (the reference's unrolled select tree,
// atomic_orbital.py:2155-2194), x^0 = 1.
template <typename T>
__device__ __forceinline__ T ipow(T x, int n) {
  T r = 1;
  if (n >= 1) {
    r = x;
    for (int k = 2; k <= n; ++k) r *= x;
  }
  return r;
}

// Value of AO a at (x, y, z): (sum_p N_p c_p e^{-Z_p r^2}) * P(x, y, z) with P a short sum of
// monomials: one x^nx y^ny z^nz for a Cartesian AO (_compute_AOs_cart, atomic_orbital.py:2197-2290),
// the real solid harmonic S_lm for a spherical AO (_compute_AOs_sphe, atomic_orbital.py:2293-2349).
// The offset from the centre is formed in FP64, the rest in T.
template <typename T>
__device__ __forceinline__ T ao_value(const Sys& s, int a, double x, double y, double z) {
  const RealTabs<T>& R = tabs<T>(s);
  const T dx = (T)(x - s.ao_ctr[3 * a]), dy = (T)(y - s.ao_ctr[3 * a + 1]), dz = (T)(z - s.ao_ctr[3 * a + 2]);
  const T r2 = dx * dx + dy * dy + dz * dz;
  T nr = 0;
  const int p1 = s.ao_pp[a + 1];
  for (int p = s.ao_pp[a]; p < p1; ++p) nr += R.prim_coef[p] * exp_nonpos(-R.prim_exp[p] * r2);
  T P = 0;
  const int t1 = s.ao_tp[a + 1];
  for (int t = s.ao_tp[a]; t < t1; ++t) {
    const int* pw = s.term_pow + 3 * t;
    P += R.term_coef[t] * (ipow(dx, pw[0]) * ipow(dy, pw[1]) * ipow(dz, pw[2]));
  }
  return nr * P;
}

// Value, gradient and Laplacian of AO a (_compute_AOs_value_grad_lap_cart / _sphe,
// atomic_orbital.py:3377-3591): radial moments NR, ZNR, Z^2NR then polynomial algebra summed
// over the monomials of P.
template <typename T>
__device__ __forceinline__ void ao_vgl(const Sys& s, int a, double x, double y, double z, T& v, T& gx, T& gy, T& gz,
                                       T& lap) {
  const RealTabs<T>& R = tabs<T>(s);
  const T dx = (T)(x - s.ao_ctr[3 * a]), dy = (T)(y - s.ao_ctr[3 * a + 1]), dz = (T)(z - s.ao_ctr[3 * a + 2]);
  const T r2 = dx * dx + dy * dy + dz * dz;
  T nr = 0, znr = 0, z2nr = 0;
  const int p1 = s.ao_pp[a + 1];
  for (int p = s.ao_pp[a]; p < p1; ++p) {
    const T Z = R.prim_exp[p];
    const T pref = R.prim_coef[p] * exp_nonpos(-Z * r2);
    nr += pref;
    znr += Z * pref;
    z2nr += Z * Z * pref;
  }
  T P = 0, Px = 0, Py = 0, Pz = 0, lapP = 0, rdg = 0;
  const int t1 = s.ao_tp[a + 1];
  for (int t = s.ao_tp[a]; t < t1; ++t) {
    const T c = R.term_coef[t];
    const int nx = s.term_pow[3 * t], ny = s.term_pow[3 * t + 1], nz = s.term_pow[3 * t + 2];
    const T px = ipow(dx, nx), py = ipow(dy, ny), pz = ipow(dz, nz);
    const T qx = ipow(dx, nx > 0 ? nx - 1 : 0), qy = ipow(dy, ny > 0 ? ny - 1 : 0), qz = ipow(dz, nz > 0 ? nz - 1 : 0);
    const T qqx = ipow(dx, nx > 1 ? nx - 2 : 0), qqy = ipow(dy, ny > 1 ? ny - 2 : 0), qqz = ipow(dz, nz > 1 ? nz - 2 : 0);
    const T fnx = (T)nx, fny = (T)ny, fnz = (T)nz;
    const T dPx = (fnx * qx) * py * pz, dPy = px * (fny * qy) * pz, dPz = px * py * (fnz * qz);
    P += c * (px * py * pz);
    Px += c * dPx;
    Py += c * dPy;
    Pz += c * dPz;
    lapP += c * ((fnx * (fnx - (T)1) * qqx) * py * pz + px * (fny * (fny - (T)1) * qqy) * pz +
                 px * py * (fnz * (fnz - (T)1) * qqz));
    rdg += c * (dx * dPx + dy * dPy + dz * dPz);
  }
  v = nr * P;
  gx = nr * Px - (T)2 * dx * znr * P;
  gy = nr * Py - (T)2 * dy * znr * P;
  gz = nr * Pz - (T)2 * dz * znr * P;
  lap = nr * lapP - (T)4 * znr * rdg + ((T)4 * r2 * z2nr - (T)6 * znr) * P;
}

// Shells: consecutive AOs on one nucleus whose contraction coefficients are proportional (the
// components of a Cartesian or spherical shell, atomic_orbital.py:2237-2251, 2315-2327) share the
// radial factor sum_p c_p e^{-Z_p r^2} of the shell's first AO; AO a of the shell is
// ao_scale[a] * radial * P_a(r - R).  Evaluating per shell takes one exponential per primitive
// and point instead of one per primitive, AO and point.

// Polynomial P_a(dx, dy, dz) of AO a.
template <typename T>
__device__ __forceinline__ T ao_poly(const Sys& s, int a, T dx, T dy, T dz);

// ao_meta packing: the AO's shell in the low 16 bits, its monomial powers in three 4-bit fields
// and AO_UNIT_TERM when its angular part is that one monomial with coefficient 1 (every Cartesian
// AO), so the value needs no lookup of the term tables.
constexpr int AO_UNIT_TERM = 1 << 28;
__device__ __forceinline__ int meta_shell(int m) { return m & 0xffff; }

// Angular part of AO a from its packed meta word: the unit monomial directly, else the term sum.
// Equal to ao_poly bit for bit (a unit term adds 1 x the monomial to zero).
template <typename T>
__device__ __forceinline__ T ao_poly_meta(const Sys& s, int a, int m, T dx, T dy, T dz) {
  if (m & AO_UNIT_TERM) return ipow(dx, (m >> 16) & 15) * ipow(dy, (m >> 20) & 15) * ipow(dz, (m >> 24) & 15);
  return ao_poly<T>(s, a, dx, dy, dz);
}

template <typename T>
__device__ __forceinline__ T ao_poly(const Sys& s, int a, T dx, T dy, T dz) {
  const RealTabs<T>& R = tabs<T>(s);
  T P = 0;
  const int t1 = s.ao_tp[a + 1];
  for (int t = s.ao_tp[a]; t < t1; ++t) {
    const int* pw = s.term_pow + 3 * t;
    P += R.term_coef[t] * (ipow(dx, pw[0]) * ipow(dy, pw[1]) * ipow(dz, pw[2]));
  }
  return P;
}

// Values of every AO of shell sh at (x, y, z) into out[a].
template <typename T>
__device__ __forceinline__ void shell_values(const Sys& s, int sh, double x, double y, double z, T* out) {
  const RealTabs<T>& R = tabs<T>(s);
  const int a0 = s.sh_ao0[sh], a1 = s.sh_ao0[sh + 1];
  const T dx = (T)(x - s.ao_ctr[3 * a0]), dy = (T)(y - s.ao_ctr[3 * a0 + 1]), dz = (T)(z - s.ao_ctr[3 * a0 + 2]);
  const T r2 = dx * dx + dy * dy + dz * dz;
  T nr = 0;
  const int p1 = s.ao_pp[a0 + 1];
  for (int p = s.ao_pp[a0]; p < p1; ++p) nr += R.prim_coef[p] * exp_nonpos(-R.prim_exp[p] * r2);
  for (int a = a0; a < a1; ++a) out[a] = R.ao_scale[a] * nr * ao_poly_meta<T>(s, a, s.ao_meta[a], dx, dy, dz);
}

// Value, gradient and Laplacian of every AO of shell sh at (x, y, z) into v, g (3 per AO) and l:
// the radial moments NR, ZNR, Z^2NR once per shell, the polynomial algebra of ao_vgl per AO.
template <typename T>
__device__ __forceinline__ void shell_vgl(const Sys& s, int sh, double x, double y, double z, T* v, T* g, T* l) {
  const RealTabs<T>& R = tabs<T>(s);
  const int a0 = s.sh_ao0[sh], a1 = s.sh_ao0[sh + 1];
  const T dx = (T)(x - s.ao_ctr[3 * a0]), dy = (T)(y - s.ao_ctr[3 * a0 + 1]), dz = (T)(z - s.ao_ctr[3 * a0 + 2]);
  const T r2 = dx * dx + dy * dy + dz * dz;
  T nr = 0, znr = 0, z2nr = 0;
  const int p1 = s.ao_pp[a0 + 1];
  for (int p = s.ao_pp[a0]; p < p1; ++p) {
    const T Z = R.prim_exp[p];
    const T pref = R.prim_coef[p] * exp_nonpos(-Z * r2);
    nr += pref;
    znr += Z * pref;
    z2nr += Z * Z * pref;
  }
  for (int a = a0; a < a1; ++a) {
    T P = 0, Px = 0, Py = 0, Pz = 0, lapP = 0, rdg = 0;
    const int t1 = s.ao_tp[a + 1];
    for (int t = s.ao_tp[a]; t < t1; ++t) {
      const T c = R.term_coef[t];
      const int nx = s.term_pow[3 * t], ny = s.term_pow[3 * t + 1], nz = s.term_pow[3 * t + 2];
      const T px = ipow(dx, nx), py = ipow(dy, ny), pz = ipow(dz, nz);
      const T qx = ipow(dx, nx > 0 ? nx - 1 : 0), qy = ipow(dy, ny > 0 ? ny - 1 : 0), qz = ipow(dz, nz > 0 ? nz - 1 : 0);
      const T qqx = ipow(dx, nx > 1 ? nx - 2 : 0), qqy = ipow(dy, ny > 1 ? ny - 2 : 0), qqz = ipow(dz, nz > 1 ? nz - 2 : 0);
      const T fnx = (T)nx, fny = (T)ny, fnz = (T)nz;
      const T dPx = (fnx * qx) * py * pz, dPy = px * (fny * qy) * pz, dPz = px * py * (fnz * qz);
      P += c * (px * py * pz);
      Px += c * dPx;
      Py += c * dPy;
      Pz += c * dPz;
      lapP += c * ((fnx * (fnx - (T)1) * qqx) * py * pz + px * (fny * (fny - (T)1) * qqy) * pz +
                   px * py * (fnz * (fnz - (T)1) * qqz));
      rdg += c * (dx * dPx + dy * dPy + dz * dPz);
    }
    const T sc = R.ao_scale[a];
    v[a] = sc * (nr * P);
    g[3 * a] = sc * (nr * Px - (T)2 * dx * znr * P);
    g[3 * a + 1] = sc * (nr * Py - (T)2 * dy * znr * P);
    g[3 * a + 2] = sc * (nr * Pz - (T)2 * dz * znr * P);
    l[a] = sc * (nr * lapP - (T)4 * znr * rdg + ((T)4 * r2 * z2nr - (T)6 * znr) * P);
  }
}

// ---------------------------------------------------------------- Jastrow pair terms

// One-body term of one electron-nucleus pair (jastrow_factor.py:672-711).
__device__ __forceinline__ double j1_term(int type, double a, double c, double A, double d) {
  if (type == 1) return -A * (1.0 / (2.0 * a) * (1.0 - exp_nonpos(-a * c * d)));
  return -A * d / (2.0 * (1.0 + a * c * d));
}

// Sum of the one-body term over all nuclei for one electron at (x, y, z).
__device__ __forceinline__ double j1_electron(const Sys& s, double x, double y, double z) {
  double acc = 0.0;
  for (int n = 0; n < s.n_nuc; ++n) {
    const double dx = x - s.nuc_pos[3 * n], dy = y - s.nuc_pos[3 * n + 1], dz = z - s.nuc_pos[3 * n + 2];
    acc += j1_term(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], sqrt(dx * dx + dy * dy + dz * dz));
  }
  return acc;
}

// Gradient and Laplacian of the one-body term for one pair (jastrow_factor.py:995-1031).
__device__ __forceinline__ void j1_grad_lap(int type, double a, double c, double A, double dx, double dy, double dz,
                                            double& g0, double& g1, double& g2, double& lap) {
  const double r = fmax(sqrt(dx * dx + dy * dy + dz * dz), 1.0e-12);
  double fp, lp;
  if (type == 1) {
    const double e = exp_nonpos(-a * c * r);
    fp = -A * (c / 2.0) * e;
    const double fs = A * (a * c * c / 2.0) * e;
    lp = fs - A * c * e / r;
  } else {
    const double den = 1.0 + a * c * r;
    fp = -A / (2.0 * den * den);
    const double fs = A * a * c / (den * den * den);
    lp = fs + 2.0 * fp / r;
  }
  g0 += fp * dx / r;
  g1 += fp * dy / r;
  g2 += fp * dz / r;
  lap += lp;
}

// Two-body pair function u(r) (jastrow_factor.py:1207-1230).
__device__ __forceinline__ double j2_u(int type, double a, double r) {
  if (type == 2) return r / 2.0 * (1.0 / (1.0 + a * r));
  return 1.0 / (2.0 * a) * (1.0 - exp_nonpos(-a * r));
}

// Gradient (w.r.t. the first electron) and Laplacian of one pair (jastrow_factor.py:3490-3512).
__device__ __forceinline__ void j2_grad_lap(int type, double a, double dx, double dy, double dz, double& g0,
                                            double& g1, double& g2, double& lap) {
  const double r = fmax(sqrt(dx * dx + dy * dy + dz * dz), 1.0e-12);
  double fp, lp;
  if (type == 2) {
    const double den = 1.0 + a * r;
    fp = 0.5 / (den * den);
    lp = -a / (den * den * den) + 2.0 * fp / r;
  } else {
    const double e = exp_nonpos(-a * r);
    fp = 0.5 * e;
    lp = -(a / 2.0) * e + 2.0 * fp / r;
  }
  const double gc = fp / r;
  g0 += gc * dx;
  g1 += gc * dy;
  g2 += gc * dz;
  lap += lp;
}

// ---------------------------------------------------------------- ECP

// Tabulated Legendre P_0..P_6 (_function_collections.py:47-66).
__device__ __forceinline__ double legendre(int n, double x) {
  switch (n) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * x * x - 1.0);
    case 3: return 0.5 * (5.0 * x * x * x - 3.0 * x);
    case 4: return 0.125 * (35.0 * x * x * x * x - 30.0 * x * x + 3.0);
    case 5: return 0.125 * (63.0 * x * x * x * x * x - 70.0 * x * x * x + 15.0 * x);
    case 6: return (231.0 * x * x * x * x * x * x - 315.0 * x * x * x * x + 105.0 * x * x - 5.0) / 16.0;
    default: return __builtin_nan("");
  }
}

// r^p for the ECP radial powers (stored as doubles, integral in practice): repeated products when
// p is a small integer, pow otherwise.
__device__ __forceinline__ double rpow(double r, double p) {
  const int n = (int)p;
  if ((double)n == p && n >= -8 && n <= 8) {
    const double b = n < 0 ? 1.0 / r : r;
    const int m = n < 0 ? -n : n;
    double v = 1.0;
    for (int k = 0; k < m; ++k) v *= b;
    return v;
  }
  return pow(r, p);
}

// Index of the nearest nucleus (lowest index on ties, a stable argsort, structure.py:411-426).
__device__ __forceinline__ int nearest_nucleus(const Sys& s, double x, double y, double z) {
  int best = 0;
  double bd = 1.0e300;
  for (int n = 0; n < s.n_nuc; ++n) {
    const double dx = s.nuc_pos[3 * n] - x, dy = s.nuc_pos[3 * n + 1] - y, dz = s.nuc_pos[3 * n + 2] - z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (d < bd) {
      bd = d;
      best = n;
    }
  }
  return best;
}

}  // namespace jqmc
