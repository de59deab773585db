// Local energy with the points in the lanes (the default E_L kernel).
//
// compute_local_energy_fast (hamiltonians.py:225-290) per walker, one wavefront per walker, in
// phases whose lanes own the items that are independent:
//
//   1. positions, G^-1 and the rotated quadrature mesh into LDS;
//   2. q_e = [M | Lambda_u] G^-1[:, e] (up) or M' G^-1[e, :]^T (dn) for every electron (lanes own
//      (electron, orbital) pairs) into LDS;
//   3. the determinant fields h_e = C^T q_e of every electron (lanes own AOs) into the per-walker
//      workspace rows hws[w][e][:] (FP64);
//   4. the continuum kinetic energy T_e (wavefunction.py:1141-1207) with the lanes owning electrons:
//      the AO value / gradient / Laplacian at r_e shell after shell, the shell and AO tables read at
//      wave-uniform indices (scalar loads), dotted with h_e and the J3 field g_e, plus the J1 / J2
//      gradients and Laplacians; also c_e = a(r_e) . g_e;
//   5. per electron the nearest nucleus, the projector radial factors V_t and the J1 + J2 value at
//      r_e; the electrons whose nearest nucleus carries non-local terms form a list;
//   6. the non-local ECP (NN = 1, coulomb_potential.py:1477-1712) with the lanes owning the
//      quadrature points of all listed electrons, 64 at a time: each lane builds its point and
//      projector weight, the J1 / J2 change, and runs the AO basis shell after shell with uniform
//      table reads, accumulating a(r') . h_e (determinant ratio, FP64) and a(r') . g_e (in T), whose
//      difference from c_e is the J3 log-ratio;
//   7. bare Coulomb and the local ECP part (coulomb_potential.py:1144-1246, 2252-2281).
//
// No lane idles on a per-AO table lookup: the AO tables sit behind a constant-address-space view,
// so their loads at uniform indices are scalar, and the per-lane loop bodies are streams of
// exponentials and fused multiply-adds over rows that are read in order.
#pragma once

#include <type_traits>

#include "jqmc_kernels.hip"

namespace jqmc {

// Read-only view of a system table in the constant address space: loads at wave-uniform indices
// through it compile to scalar (SMEM) loads.
template <typename U>
using CPtr = const __attribute__((address_space(4))) U*;
template <typename U>
__device__ __forceinline__ CPtr<U> cview(const U* p) {
  return (CPtr<U>)p;
}

// The threads that evaluate one walker: NW wavefronts.  NW = 1: one wavefront of a block that may
// hold several walkers (wave_sync is its barrier); NW > 1: the whole block (__syncthreads).  Loops
// over independent items start at tid and stride by N; sum() and sum_k() reduce over the team in a
// fixed order (the butterfly of each wave, then the waves in order through the LDS words red,
// NW * 8 doubles), so every thread gets the bitwise-identical total.  With NW = 1 every phase below
// runs the same operations in the same order as a plain one-wavefront loop.
template <int NW>
struct Team {
  static constexpr int N = WAVE * NW;
  int tid, wid, lane;
  double* red;
  __device__ __forceinline__ void sync() const {
    if constexpr (NW == 1)
      wave_sync();
    else
      __syncthreads();
  }
  template <int K>
  __device__ __forceinline__ void sum_k(double* v) const {
    wave_sum_k<K>(v);
    if constexpr (NW > 1) {
      if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[wid * 8 + k] = v[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        double t = red[k];
        for (int q = 1; q < NW; ++q) t += red[q * 8 + k];
        v[k] = t;
      }
      __syncthreads();
    }
  }
  __device__ __forceinline__ double sum(double x) const {
    double v[1] = {x};
    sum_k<1>(v);
    return v[0];
  }
};

// The one-wavefront team of lane `lane`.
__device__ __forceinline__ Team<1> wave_team(int lane) { return Team<1>{lane, 0, lane, nullptr}; }

// The q_e rows of all electrons stay in LDS when they take at most EL2_Q_LDS bytes, otherwise they
// go to the workspace after the h_e rows (large AO-orbital geminals).
constexpr int64_t EL2_Q_LDS = 16 * 1024;
__host__ __device__ inline bool el2_q_in_lds(const Sys& s) {
  return (int64_t)s.n_el * s.n_orb * (int64_t)sizeof(double) <= EL2_Q_LDS;
}

// LDS of one walker of k_local_energy_pts (one wavefront; nw > 1: a team of nw wavefronts, which
// also needs the reduction words red and the per-wave kinetic partials kred).
struct El2Lds {
  double *rr, *Gi, *RT, *gk, *Q, *en, *Vt, *jold, *te, *jc, *red, *kred;
  int* elist;
  uintptr_t end;
  __host__ __device__ El2Lds(char* base, const Sys& s, int nw = 1) {
    const int ne = s.n_el, nv = s.nv > 0 ? s.nv : 1, mt = s.ecp_max_terms > 0 ? s.ecp_max_terms : 1;
    Carve c{base};
    red = c.take<double>(nw > 1 ? 8 * nw : 0);
    kred = c.take<double>(nw > 1 ? (int64_t)9 * WAVE * nw : 0);
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(s.n_up * s.n_up);
    RT = c.take<double>(9);
    gk = c.take<double>(3 * nv);
    Q = c.take<double>(el2_q_in_lds(s) ? (int64_t)ne * s.n_orb : 0);
    en = c.take<double>(5 * ne);  // nearest ECP nucleus of each electron: n, R - r (3), |R - r|
    Vt = c.take<double>((int64_t)ne * mt);
    jold = c.take<double>(ne);
    te = c.take<double>(ne);  // kinetic energy T_e
    jc = c.take<double>(ne);  // c_e = a(r_e) . g_e
    elist = c.take<int>(ne);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

inline int64_t el2_lds_bytes(const Sys& s, int nw = 1) { return (int64_t)El2Lds(0, s, nw).end; }

// Powers of AO i of a canonical Cartesian shell of angular momentum L (1; x, y, z; xx, xy, xz, yy, yz,
// zz), and its angular part with gradient, Laplacian and r . grad, in the arithmetic of the
// per-AO path (x^0 = 1, d/dx x^n = n x^(n-1), d2/dx2 x^2 = 2).
template <int L>
struct Cart {
  static constexpr int N = L == 0 ? 1 : (L == 1 ? 3 : 6);
  static __host__ __device__ constexpr int nx(int i) { return L == 0 ? 0 : (L == 1 ? (i == 0) : (i == 0 ? 2 : (i < 3 ? 1 : 0))); }
  static __host__ __device__ constexpr int ny(int i) {
    return L == 0 ? 0 : (L == 1 ? (i == 1) : (i == 1 || i == 4 ? 1 : (i == 3 ? 2 : 0)));
  }
  static __host__ __device__ constexpr int nz(int i) { return L - nx(i) - ny(i); }
};

template <typename U>
__device__ __forceinline__ U cpow(int n, U x) {
  return n == 0 ? (U)1 : (n == 1 ? x : x * x);
}

template <int L, int I>
__device__ __forceinline__ void cart_vgl(double dx, double dy, double dz, double& P, double& Px, double& Py, double& Pz,
                                         double& lapP, double& rdg) {
  constexpr int nx = Cart<L>::nx(I), ny = Cart<L>::ny(I), nz = Cart<L>::nz(I);
  const double px = cpow(nx, dx), py = cpow(ny, dy), pzz = cpow(nz, dz);
  const double qx = nx == 0 ? 0.0 : (nx == 1 ? 1.0 : nx * dx), qy = ny == 0 ? 0.0 : (ny == 1 ? 1.0 : ny * dy),
               qz = nz == 0 ? 0.0 : (nz == 1 ? 1.0 : nz * dz);
  const double qqx = nx == 2 ? 2.0 : 0.0, qqy = ny == 2 ? 2.0 : 0.0, qqz = nz == 2 ? 2.0 : 0.0;
  P = px * py * pzz;
  Px = qx * py * pzz;
  Py = px * qy * pzz;
  Pz = px * py * qz;
  lapP = qqx * py * pzz + px * qqy * pzz + px * py * qqz;
  rdg = dx * Px + dy * Py + dz * Pz;
}

// Value of the monomial or term sum of AO a (meta word m) at (dx, dy, dz), tables read at the
// wave-uniform index a.
template <typename U>
__device__ __forceinline__ U poly_uniform(const Sys& s, int a, int m, U dx, U dy, U dz) {
  if (m & AO_UNIT_TERM) return ipow(dx, (m >> 16) & 15) * ipow(dy, (m >> 20) & 15) * ipow(dz, (m >> 24) & 15);
  const CPtr<int> tp = cview(s.ao_tp);
  const CPtr<int> pw = cview(s.term_pow);
  const CPtr<U> tc = cview(tabs<U>(s).term_coef);
  U P = 0;
  const int t1 = tp[a + 1];
  for (int t = tp[a]; t < t1; ++t) P += tc[t] * (ipow(dx, pw[3 * t]) * ipow(dy, pw[3 * t + 1]) * ipow(dz, pw[3 * t + 2]));
  return P;
}

// ---- the phases, each run by the whole wavefront of one walker

// 1. positions, G^-1 and the rotation RTw (NULL: identity) into LDS; the rotated mesh when the system
// has non-local terms (read after the next wave_sync).
template <int NW>
__device__ __forceinline__ void el_load_walker(const Sys& s, const double* __restrict__ r_up,
                                               const double* __restrict__ r_dn, const double* __restrict__ Giw,
                                               const double* __restrict__ RTw, const El2Lds& L, const Team<NW>& tm) {
  const int nu = s.n_up, nd = s.n_dn, i = tm.tid;
  for (int k = i; k < 3 * nu; k += tm.N) L.rr[k] = r_up[k];
  for (int k = i; k < 3 * nd; k += tm.N) L.rr[3 * nu + k] = r_dn[k];
  for (int k = i; k < nu * nu; k += tm.N) L.Gi[k] = Giw[k];
  if (i < 9) L.RT[i] = RTw ? RTw[i] : ((i % 4) == 0 ? 1.0 : 0.0);
  tm.sync();
  if (s.ecp_on && s.nv > 0) rotate_mesh(s, L.RT, L.gk, i, tm.N);
}

// Orbitals of a spin block whose C column a lane holds in registers in the determinant fields; systems
// with more take the kWide form.
constexpr int EL_C_REGS = 16;

// 2-3. q_e = [M | Lambda_u] G^-1[:, e] (up) or M' G^-1[e, :]^T (dn) of every electron into Qr, then the
// determinant fields h_e = C^T q_e into the rows Hw[e][:] (FP64), every sum over the orbitals k in
// order.  Qw: the walker's pair rows.  kWide (n_orb > EL_C_REGS): the sums of 16 electrons stay in
// registers while the lane's AO column of C streams past 8 orbitals at a time, instead of the
// column's orbitals beyond the first 16 being read again for every electron.
template <int NW, bool kWide = false>
__device__ __forceinline__ void el_det_fields_rows(const Sys& s, const El2Lds& L, const double* __restrict__ Qw,
                                                   double* Qr, double* Hw, const Team<NW>& tm) {
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  for (int t = tm.tid; t < ne * no; t += tm.N) {
    const int e = t / no, k = t - e * no;
    double acc = 0.0;
    if (e < nu) {
      for (int j = 0; j < nd; ++j) acc += Qw[(int64_t)(nu + j) * no + k] * L.Gi[j * nu + e];
      for (int u = 0; u < s.n_unp; ++u) acc += s.td.lam_u[(int64_t)k * s.n_unp + u] * L.Gi[(nd + u) * nu + e];
    } else {
      const int j = e - nu;
      for (int i = 0; i < nu; ++i) acc += Qw[(int64_t)i * no + k] * L.Gi[j * nu + i];
    }
    Qr[t] = acc;
  }
  tm.sync();

  if constexpr (kWide) {
    constexpr int KW = 8, EC = 16;
    for (int a = tm.tid; a < na; a += tm.N) {
      if (!s.orb_is_mo) {
        for (int e = 0; e < ne; ++e) Hw[(int64_t)e * na + a] = Qr[e * no + a];
        continue;
      }
      for (int blk = 0; blk < 2; ++blk) {
        const double* C = blk == 0 ? s.td.mo_up : s.td.mo_dn;
        const int e0 = blk == 0 ? 0 : nu, e1 = blk == 0 ? nu : ne;
        for (int eb = e0; eb < e1; eb += EC) {
          double acc[EC];
#pragma unroll
          for (int i = 0; i < EC; ++i) acc[i] = 0.0;
          for (int kb = 0; kb < no; kb += KW) {
            double c[KW];
#pragma unroll
            for (int k = 0; k < KW; ++k) c[k] = kb + k < no ? C[(int64_t)(kb + k) * na + a] : 0.0;
#pragma unroll
            for (int i = 0; i < EC; ++i) {
              if (eb + i >= e1) continue;
              const double* q = Qr + (eb + i) * no + kb;
#pragma unroll
              for (int k = 0; k < KW; ++k)
                if (kb + k < no) acc[i] += c[k] * q[k];
            }
          }
#pragma unroll
          for (int i = 0; i < EC; ++i)
            if (eb + i < e1) Hw[(int64_t)(eb + i) * na + a] = acc[i];
        }
      }
    }
    tm.sync();
    return;
  }
  // the lane's AO column of C held in registers for the first KC orbitals, up block with C_up, dn
  // block with C_dn
  constexpr int KC = EL_C_REGS;
  for (int a = tm.tid; a < na; a += tm.N) {
    if (!s.orb_is_mo) {
      for (int e = 0; e < ne; ++e) Hw[(int64_t)e * na + a] = Qr[e * no + a];
      continue;
    }
    const int nk = no < KC ? no : KC;
    for (int blk = 0; blk < 2; ++blk) {
      const double* C = blk == 0 ? s.td.mo_up : s.td.mo_dn;
      double c[KC];
#pragma unroll
      for (int k = 0; k < KC; ++k) c[k] = k < nk ? C[(int64_t)k * na + a] : 0.0;
      const int e0 = blk == 0 ? 0 : nu, e1 = blk == 0 ? nu : ne;
      for (int e = e0; e < e1; ++e) {
        const double* q = Qr + e * no;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KC; ++k)
          if (k < nk) acc += c[k] * q[k];
        for (int k = KC; k < no; ++k) acc += C[(int64_t)k * na + a] * q[k];
        Hw[(int64_t)e * na + a] = acc;
      }
    }
  }
  tm.sync();
}

// The q rows in LDS when they fit (el2_q_in_lds), else in the workspace rows Qg; each branch keeps
// its pointer's address space, so that the q loads compile to LDS or global instructions.
template <int NW, bool kWide = false>
__device__ __forceinline__ void el_det_fields(const Sys& s, const El2Lds& L, const double* __restrict__ Qw, double* Qg,
                                              double* Hw, const Team<NW>& tm) {
  if (el2_q_in_lds(s))
    el_det_fields_rows<NW, kWide>(s, L, Qw, L.Q, Hw, tm);
  else
    el_det_fields_rows<NW, kWide>(s, L, Qw, Qg, Hw, tm);
}

// 4. the continuum kinetic energy T_e (wavefunction.py:1141-1207) with the lanes owning electrons:
// AO value / gradient / Laplacian at r_e shell after shell, dotted with h_e (Hw) and the J3 field
// g_e (gw, NULL without J3), plus the J1 / J2 gradients and Laplacians.  T_e goes to L.te and to
// t_elec_w when given; c_e = a(r_e) . g_e (FP64 AO values) to L.jc, the J3 term that the point
// ratios subtract.  Returns the walker's kinetic energy on every thread.  A team of NW > 1 waves
// splits the shells (wave w takes shells w, w + NW, ...) and the J2 partners among its waves, all
// waves holding the same electrons in their lanes; the nine partial sums per electron are added in
// wave order through L.kred, and wave 0 forms T_e.
template <typename TC, int NW>
__device__ double el_kinetic(const Sys& s, const El2Lds& L, const double* __restrict__ Hw, const TC* gw,
                             const Team<NW>& tm, double* t_elec_w) {
  const int lane = tm.lane;
  const int ne = s.n_el, na = s.n_ao, nsh = s.n_shell;
  const CPtr<int> sh0 = cview(s.sh_ao0);
  const CPtr<int> app = cview(s.ao_pp);
  const CPtr<int> meta = cview(s.ao_meta);
  const CPtr<double> shr = cview(s.shrec);
  const CPtr<double> pz = cview(s.td.prim_exp);
  const CPtr<double> pc = cview(s.td.prim_coef);
  const CPtr<double> sc = cview(s.td.ao_scale);
  double kinetic = 0.0;
  for (int e0 = 0; e0 < ne; e0 += WAVE) {
    const int e = e0 + lane;
    const bool act = e < ne;
    const int ec = act ? e : 0;
    const double ex = L.rr[3 * ec], ey = L.rr[3 * ec + 1], ez = L.rr[3 * ec + 2];
    const double* hrow = Hw + (int64_t)ec * na;
    const TC* grow = gw ? gw + (int64_t)ec * na : nullptr;
    double acc8[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // gJ(3), lapJ, gD(3), lapD'
    double jc = 0.0;
    // per AO: value P, gradient and Laplacian of the angular part, the contributions to the fields
    auto add_ao = [&](int a, double f, double dx, double dy, double dz, double r2, double nr, double znr, double z2nr,
                      double P, double Px, double Py, double Pz, double lapP, double rdg) {
      const double g0 = f * (nr * Px - 2.0 * dx * znr * P), g1 = f * (nr * Py - 2.0 * dy * znr * P),
                   g2 = f * (nr * Pz - 2.0 * dz * znr * P);
      const double lp = f * (nr * lapP - 4.0 * znr * rdg + (4.0 * r2 * z2nr - 6.0 * znr) * P);
      if (grow) {
        const double gj = (double)grow[a];
        acc8[0] += gj * g0;
        acc8[1] += gj * g1;
        acc8[2] += gj * g2;
        acc8[3] += gj * lp;
        jc += gj * (f * (nr * P));
      }
      const double hd = hrow[a];
      acc8[4] += hd * g0;
      acc8[5] += hd * g1;
      acc8[6] += hd * g2;
      acc8[7] += hd * lp;
    };
    for (int sh = tm.wid; sh < nsh; sh += NW) {
      const CPtr<double> rec = shr + (int64_t)sh * SHREC;
      const int info = (int)rec[3];
      const int a0 = info >> 9;
      const double dx = ex - rec[0], dy = ey - rec[1], dz = ez - rec[2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      double nr = 0.0, znr = 0.0, z2nr = 0.0;
      if (info & 256) {
        // fast shell: every table word from the record
        const int np = info & 15, l = (info >> 4) & 15;
        double w[2 * SHR_PMAX], scl[SHR_AMAX];
#pragma unroll
        for (int k = 0; k < 2 * SHR_PMAX; ++k) w[k] = rec[SHR_P + k];
#pragma unroll
        for (int k = 0; k < SHR_AMAX; ++k) scl[k] = rec[SHR_SC + k];
#pragma unroll
        for (int q = 0; q < SHR_PMAX; ++q) {
          if (q < np) {
            const double Z = w[2 * q];
            const double pref = w[2 * q + 1] * exp_nonpos(-Z * r2);
            nr += pref;
            znr += Z * pref;
            z2nr += Z * Z * pref;
          }
        }
        auto shell = [&](auto Lc) {
          constexpr int L = decltype(Lc)::value;
          double P, Px, Py, Pz, lapP, rdg;
          cart_vgl<L, 0>(dx, dy, dz, P, Px, Py, Pz, lapP, rdg);
          add_ao(a0, scl[0], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
          if constexpr (L >= 1) {
            cart_vgl<L, 1>(dx, dy, dz, P, Px, Py, Pz, lapP, rdg);
            add_ao(a0 + 1, scl[1], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
            cart_vgl<L, 2>(dx, dy, dz, P, Px, Py, Pz, lapP, rdg);
            add_ao(a0 + 2, scl[2], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
          }
          if constexpr (L == 2) {
            cart_vgl<L, 3>(dx, dy, dz, P, Px, Py, Pz, lapP, rdg);
            add_ao(a0 + 3, scl[3], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
            cart_vgl<L, 4>(dx, dy, dz, P, Px, Py, Pz, lapP, rdg);
            add_ao(a0 + 4, scl[4], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
            cart_vgl<L, 5>(dx, dy, dz, P, Px, Py, Pz, lapP, rdg);
            add_ao(a0 + 5, scl[5], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
          }
        };
        if (l == 0)
          shell(std::integral_constant<int, 0>{});
        else if (l == 1)
          shell(std::integral_constant<int, 1>{});
        else
          shell(std::integral_constant<int, 2>{});
        continue;
      }
      const int a1 = sh0[sh + 1];
      const int p1 = app[a0 + 1];
      for (int p = app[a0]; p < p1; ++p) {
        const double Z = pz[p];
        const double pref = pc[p] * exp_nonpos(-Z * r2);
        nr += pref;
        znr += Z * pref;
        z2nr += Z * Z * pref;
      }
      for (int a = a0; a < a1; ++a) {
        const int m = meta[a];
        double P, Px, Py, Pz, lapP, rdg;
        if (m & AO_UNIT_TERM) {
          const int nx = (m >> 16) & 15, ny = (m >> 20) & 15, nz = (m >> 24) & 15;
          const double px = ipow(dx, nx), py = ipow(dy, ny), pzz = ipow(dz, nz);
          const double qx = nx > 0 ? nx * ipow(dx, nx - 1) : 0.0, qy = ny > 0 ? ny * ipow(dy, ny - 1) : 0.0,
                       qz = nz > 0 ? nz * ipow(dz, nz - 1) : 0.0;
          const double qqx = nx > 1 ? nx * (nx - 1.0) * ipow(dx, nx - 2) : 0.0,
                       qqy = ny > 1 ? ny * (ny - 1.0) * ipow(dy, ny - 2) : 0.0,
                       qqz = nz > 1 ? nz * (nz - 1.0) * ipow(dz, nz - 2) : 0.0;
          P = px * py * pzz;
          Px = qx * py * pzz;
          Py = px * qy * pzz;
          Pz = px * py * qz;
          lapP = qqx * py * pzz + px * qqy * pzz + px * py * qqz;
          rdg = dx * Px + dy * Py + dz * Pz;
        } else {
          const CPtr<int> tp = cview(s.ao_tp);
          const CPtr<int> pw = cview(s.term_pow);
          const CPtr<double> tc = cview(s.td.term_coef);
          P = Px = Py = Pz = lapP = rdg = 0.0;
          const int t1 = tp[a + 1];
          for (int t = tp[a]; t < t1; ++t) {
            const double c = tc[t];
            const int nx = pw[3 * t], ny = pw[3 * t + 1], nz = pw[3 * t + 2];
            const double px = ipow(dx, nx), py = ipow(dy, ny), pzz = ipow(dz, nz);
            const double qx = nx > 0 ? nx * ipow(dx, nx - 1) : 0.0, qy = ny > 0 ? ny * ipow(dy, ny - 1) : 0.0,
                         qz = nz > 0 ? nz * ipow(dz, nz - 1) : 0.0;
            const double qqx = nx > 1 ? nx * (nx - 1.0) * ipow(dx, nx - 2) : 0.0,
                         qqy = ny > 1 ? ny * (ny - 1.0) * ipow(dy, ny - 2) : 0.0,
                         qqz = nz > 1 ? nz * (nz - 1.0) * ipow(dz, nz - 2) : 0.0;
            const double dPx = qx * py * pzz, dPy = px * qy * pzz, dPz = px * py * qz;
            P += c * (px * py * pzz);
            Px += c * dPx;
            Py += c * dPy;
            Pz += c * dPz;
            lapP += c * (qqx * py * pzz + px * qqy * pzz + px * py * qqz);
            rdg += c * (dx * dPx + dy * dPy + dz * dPz);
          }
        }
        add_ao(a, sc[a], dx, dy, dz, r2, nr, znr, z2nr, P, Px, Py, Pz, lapP, rdg);
      }
    }
    if (s.j1_type && tm.wid == 0) {
      for (int n = 0; n < s.n_nuc; ++n)
        j1_grad_lap(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], ex - s.nuc_pos[3 * n], ey - s.nuc_pos[3 * n + 1],
                    ez - s.nuc_pos[3 * n + 2], acc8[0], acc8[1], acc8[2], acc8[3]);
    }
    if (s.j2_type) {
      for (int f = tm.wid; f < ne; f += NW) {
        if (f == ec) continue;
        j2_grad_lap(s.j2_type, s.j2_param, ex - L.rr[3 * f], ey - L.rr[3 * f + 1], ez - L.rr[3 * f + 2], acc8[0],
                    acc8[1], acc8[2], acc8[3]);
      }
    }
    if constexpr (NW > 1) {
#pragma unroll
      for (int k = 0; k < 8; ++k) L.kred[(tm.wid * 9 + k) * WAVE + lane] = acc8[k];
      L.kred[(tm.wid * 9 + 8) * WAVE + lane] = jc;
      __syncthreads();
      if (tm.wid == 0) {
        for (int q = 1; q < NW; ++q) {
#pragma unroll
          for (int k = 0; k < 8; ++k) acc8[k] += L.kred[(q * 9 + k) * WAVE + lane];
          jc += L.kred[(q * 9 + 8) * WAVE + lane];
        }
      }
      __syncthreads();
    }
    const double gDx = acc8[4], gDy = acc8[5], gDz = acc8[6];
    const double lapD = acc8[7] - (gDx * gDx + gDy * gDy + gDz * gDz);
    const double tx = acc8[0] + gDx, ty = acc8[1] + gDy, tz = acc8[2] + gDz;
    const double Te = -0.5 * (acc8[3] + lapD + (tx * tx + ty * ty + tz * tz));
    if (act && tm.wid == 0) {
      kinetic += Te;
      L.te[e] = Te;
      L.jc[e] = jc;
      if (t_elec_w) t_elec_w[e] = Te;
    }
  }
  return tm.sum(kinetic);
}

// 5. per electron the nearest nucleus n, R - r_e and |R - r_e| (L.en) and the projector radial
// factors V_t of its projected terms (L.Vt) when the system has non-local terms, and the J1 + J2
// value at r_e (L.jold) for every electron when all_jold, else for the electrons with projected
// terms.  Those electrons are listed in electron order in L.elist; returns their count on every
// thread (the list is valid after the next team sync).
template <int NW>
__device__ int el_ecp_electrons(const Sys& s, const El2Lds& L, const Team<NW>& tm, bool all_jold) {
  const int ne = s.n_el, lane = tm.lane;
  const bool nonlocal = s.ecp_on && s.nv > 0;
  const int mt = s.ecp_max_terms > 0 ? s.ecp_max_terms : 1;
  int n_ecp = 0;
  for (int e0 = 0; e0 < ne; e0 += tm.N) {
    const int e = e0 + tm.tid;
    bool has = false;
    if (e < ne) {
      const double x = L.rr[3 * e], y = L.rr[3 * e + 1], z = L.rr[3 * e + 2];
      if (nonlocal) {
        EcpElec E;
        E.n = nearest_nucleus(s, x, y, z);
        E.nt = s.ecp_nt[E.n];
        E.rx = s.nuc_pos[3 * E.n] - x;
        E.ry = s.nuc_pos[3 * E.n + 1] - y;
        E.rz = s.nuc_pos[3 * E.n + 2] - z;
        E.rn = sqrt(E.rx * E.rx + E.ry * E.ry + E.rz * E.rz);
        has = E.nt > 0;
        L.en[5 * e] = (double)E.n;
        L.en[5 * e + 1] = E.rx;
        L.en[5 * e + 2] = E.ry;
        L.en[5 * e + 3] = E.rz;
        L.en[5 * e + 4] = E.rn;
        for (int t = 0; t < E.nt; ++t) {
          const int o = E.n * s.ecp_max_terms + t;
          L.Vt[e * mt + t] =
              (1.0 / (E.rn * E.rn)) * s.ecp_coef[o] * rpow(E.rn, s.ecp_pow[o]) * exp_nonpos(-s.ecp_exp[o] * (E.rn * E.rn));
        }
      }
      if (has || all_jold) {
        double old = 0.0;
        if (s.j2_type) {
          for (int f = 0; f < ne; ++f) {
            if (f == e) continue;
            const double dx = x - L.rr[3 * f], dy = y - L.rr[3 * f + 1], dz = z - L.rr[3 * f + 2];
            old += j2_u(s.j2_type, s.j2_param, sqrt(dx * dx + dy * dy + dz * dz));
          }
        }
        if (s.j1_type) old += j1_electron(s, x, y, z);
        L.jold[e] = old;
      }
    }
    const uint64_t mask = __ballot(has);
    int before = 0, total = __popcll(mask);
    if constexpr (NW > 1) {
      // the waves' counts in order: this wave's offset and the round's total
      if (lane == 0) L.red[tm.wid] = (double)total;
      __syncthreads();
      total = 0;
      for (int q = 0; q < NW; ++q) {
        const int c = (int)L.red[q];
        before += q < tm.wid ? c : 0;
        total += c;
      }
      __syncthreads();
    }
    const int rank = n_ecp + before + __popcll(mask & ((1ull << lane) - 1ull));
    if (has) L.elist[rank] = e;
    n_ecp += total;
  }
  return n_ecp;
}

// Quadrature point k of the listed electron e (r' = r_e + (R - r_e) + |R - r_e| g~_k) and its
// projector weight sum_t V_t (2 l_t + 1) P_l(cos theta) w_k (coulomb_potential.py:1608-1705), from
// the per-electron data of el_ecp_electrons.
__device__ __forceinline__ void el_ecp_point(const Sys& s, const El2Lds& L, int e, int k, double& px, double& py,
                                             double& pz, double& pwt) {
  const int mt = s.ecp_max_terms > 0 ? s.ecp_max_terms : 1;
  const double* En = L.en + 5 * e;
  const int n = (int)En[0];
  const double rx = En[1], ry = En[2], rz = En[3], rn = En[4];
  const double gx = L.gk[3 * k], gy = L.gk[3 * k + 1], gz = L.gk[3 * k + 2];
  const double gn = sqrt(gx * gx + gy * gy + gz * gz);
  const double ex = L.rr[3 * e], ey = L.rr[3 * e + 1], ez = L.rr[3 * e + 2];
  px = ex + (rx + rn * gx);
  py = ey + (ry + rn * gy);
  pz = ez + (rz + rn * gz);
  const double cth = (-rx / rn) * (gx / gn) + (-ry / rn) * (gy / gn) + (-rz / rn) * (gz / gn);
  pwt = 0.0;
  const int nt = s.ecp_nt[n];
  for (int t = 0; t < nt; ++t) {
    const int l = s.ecp_l[n * s.ecp_max_terms + t];
    pwt += L.Vt[e * mt + t] * ((2 * l + 1) * legendre(l, cth) * s.mesh_w[k]);
  }
}

// One fast shell (np primitives, canonical Cartesian AOs of angular momentum L) at a point (dx, dy, dz
// from its centre, r2 = |.|^2): its AO values dotted into det (with the field row h, FP64) and dj3 (with
// the J3 field row g, in T).  The record words are read unconditionally, so that they arrive with a
// few wide loads and one wait.
template <typename T, typename TC, int L>
__device__ __forceinline__ void point_shell(const CPtr<double> rec, int np, T dx, T dy, T dz, T r2,
                                            const double* __restrict__ h, const TC* g, double& det, T& dj3) {
  constexpr int NA = Cart<L>::N;
  double w[2 * SHR_PMAX], sc[NA], hv[NA];
  TC gv[NA];
#pragma unroll
  for (int k = 0; k < 2 * SHR_PMAX; ++k) w[k] = rec[SHR_P + k];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    sc[k] = rec[SHR_SC + k];
    hv[k] = h[k];
    gv[k] = g ? g[k] : (TC)0;
  }
  T nr = 0;
#pragma unroll
  for (int q = 0; q < SHR_PMAX; ++q)
    if (q < np) nr += (T)w[2 * q + 1] * exp_nonpos(-(T)w[2 * q] * r2);
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    const T P = cpow(Cart<L>::nx(i), dx) * cpow(Cart<L>::ny(i), dy) * cpow(Cart<L>::nz(i), dz);
    const T v = (T)sc[i] * (nr * P);
    det += (double)v * hv[i];
    if (g) dj3 += v * (T)gv[i];
  }
}

// Electron e moved to (px, py, pz), every other electron in place: the determinant ratio
// a(r') . h_e (FP64) and the log Jastrow ratio a(r') . g_e - c_e (J3, the AO values at r' in T) +
// J1 + J2 at r' - L.jold[e].  Each lane has its own point; the AO basis runs shell after shell with
// the tables read at wave-uniform indices.
template <typename T, typename TC>
__device__ __forceinline__ void el_point(const Sys& s, const El2Lds& L, int e, double px, double py, double pz,
                                         const double* __restrict__ Hw, const TC* gw, double& det_out,
                                         double& dj_out) {
  const int ne = s.n_el, na = s.n_ao, nsh = s.n_shell;
  double jn = 0.0;
  if (s.j2_type) {
    for (int f = 0; f < ne; ++f) {
      const double dx = px - L.rr[3 * f], dy = py - L.rr[3 * f + 1], dz = pz - L.rr[3 * f + 2];
      const double u = j2_u(s.j2_type, s.j2_param, sqrt(dx * dx + dy * dy + dz * dz));
      jn += f == e ? 0.0 : u;
    }
  }
  if (s.j1_type) jn += j1_electron(s, px, py, pz);
  const CPtr<int> sh0 = cview(s.sh_ao0);
  const CPtr<int> app = cview(s.ao_pp);
  const CPtr<int> meta = cview(s.ao_meta);
  const RealTabs<T>& R = tabs<T>(s);
  const CPtr<T> pz_ = cview(R.prim_exp);
  const CPtr<T> pc = cview(R.prim_coef);
  const CPtr<T> sc = cview(R.ao_scale);
  const double* hrow = Hw + (int64_t)e * na;
  const TC* grow = gw ? gw + (int64_t)e * na : nullptr;
  const CPtr<double> shr = cview(s.shrec);
  double det = 0.0;
  T dj3 = 0;
  for (int sh = 0; sh < nsh; ++sh) {
    const CPtr<double> rec = shr + (int64_t)sh * SHREC;
    const int info = (int)rec[3];
    const int a0 = info >> 9;
    const T dx = (T)(px - rec[0]), dy = (T)(py - rec[1]), dz = (T)(pz - rec[2]);
    const T r2 = dx * dx + dy * dy + dz * dz;
    T nr = 0;
    if (info & 256) {
      // fast shell: every table word from the record
      const int np = info & 15, l = (info >> 4) & 15;
      const TC* gs = grow ? grow + a0 : nullptr;
      if (l == 0)
        point_shell<T, TC, 0>(rec, np, dx, dy, dz, r2, hrow + a0, gs, det, dj3);
      else if (l == 1)
        point_shell<T, TC, 1>(rec, np, dx, dy, dz, r2, hrow + a0, gs, det, dj3);
      else
        point_shell<T, TC, 2>(rec, np, dx, dy, dz, r2, hrow + a0, gs, det, dj3);
      continue;
    }
    const int a1 = sh0[sh + 1];
    const int p1 = app[a0 + 1];
    for (int p = app[a0]; p < p1; ++p) nr += pc[p] * exp_nonpos(-pz_[p] * r2);
    for (int a = a0; a < a1; ++a) {
      const T v = sc[a] * (nr * poly_uniform<T>(s, a, meta[a], dx, dy, dz));
      det += (double)v * hrow[a];
      if (grow) dj3 += v * (T)grow[a];
    }
  }
  det_out = det;
  dj_out = (grow ? (double)dj3 - L.jc[e] : 0.0) + (jn - L.jold[e]);
}

// 6. the non-local ECP (NN = 1, coulomb_potential.py:1477-1712) with the lanes owning the quadrature
// points of the n_ecp listed electrons, 64 at a time.  Returns the lane's partial sum.
template <typename T, typename TC, int NW>
__device__ double el_nonlocal(const Sys& s, const El2Lds& L, int n_ecp, const double* __restrict__ Hw, const TC* gw,
                              const Team<NW>& tm) {
  const int nv = s.nv, lane = tm.lane;
  const int P = n_ecp * nv;
  double vnl = 0.0;
  for (int base = tm.wid * WAVE; base < P; base += tm.N) {
    const int idx = base + lane;
    const bool valid = idx < P;
    const int ie = valid ? idx / nv : 0, k = valid ? idx - ie * nv : 0;
    const int e = L.elist[ie];
    double px, py, pz, pwt, det, dj;
    el_ecp_point(s, L, e, k, px, py, pz, pwt);
    el_point<T, TC>(s, L, e, px, py, pz, Hw, gw, det, dj);
    vnl += valid ? pwt * (det * exp(dj)) : 0.0;
  }
  return vnl;
}

// 7. bare Coulomb and the local ECP part (coulomb_potential.py:1144-1246, 2252-2281): the lane's
// partial sum (without the ion-ion term).
template <int NW>
__device__ __forceinline__ double el_potential_partial(const Sys& s, const El2Lds& L, const Team<NW>& tm) {
  const int ne = s.n_el;
  double pot = 0.0;
  for (int t = tm.tid; t < ne * s.n_nuc; t += tm.N) {
    const int e = t / s.n_nuc, n = t - e * s.n_nuc;
    const double dx = L.rr[3 * e] - s.nuc_pos[3 * n], dy = L.rr[3 * e + 1] - s.nuc_pos[3 * n + 1],
                 dz = L.rr[3 * e + 2] - s.nuc_pos[3 * n + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    pot += -s.nuc_charge[n] / d;
    if (s.ecp_on) pot += ecp_local_pair(s, n, d);
  }
  return pot + el_el_partial(s, L.rr, tm.tid, tm.N);
}

// The local energy of one walker by one team (phases 1-7 above).  Inputs: the walker's positions
// r_up / r_dn, G^-1 (Giw), pair rows (Qw), J3 fields (gw, NULL without J3) in the cache type TC, its
// rotation RTw (NULL: identity); workspace rows Hw (N x n_ao doubles) and, when the q rows do not fit
// LDS, Qg (N x n_orb doubles).  T is the arithmetic type of the AO values at the quadrature points.
// Returns kinetic and potential energy on every thread; the per-electron kinetic energies go to
// t_elec_w when given.  A team of NW waves shares the walker's shells (kinetic phase) and its
// quadrature points (non-local phase, 64-point rounds dealt to the waves in turn).
template <typename T, typename TC, int NW>
__device__ void el_pts_walker(const Sys& s, const double* __restrict__ r_up, const double* __restrict__ r_dn,
                              const double* __restrict__ Giw, const double* __restrict__ Qw, const TC* gw,
                              const double* __restrict__ RTw, double* Hw, double* Qg, const El2Lds& L,
                              const Team<NW>& tm, double* t_elec_w, double& kinetic_out, double& pot_out) {
  el_load_walker(s, r_up, r_dn, Giw, RTw, L, tm);
  el_det_fields(s, L, Qw, Qg, Hw, tm);
  kinetic_out = el_kinetic<TC>(s, L, Hw, gw, tm, t_elec_w);
  double vnl = 0.0;
  if (s.ecp_on && s.nv > 0) {
    const int n_ecp = el_ecp_electrons(s, L, tm, false);
    tm.sync();
    vnl = el_nonlocal<T, TC>(s, L, n_ecp, Hw, gw, tm);
  }
  double red[2] = {el_potential_partial(s, L, tm), vnl};
  tm.template sum_k<2>(red);
  pot_out = (red[0] + s.ion_ion) + red[1];
}

// Walkers of the mixed form whose local energy is non-finite or lies farther than `width` from
// e_ref: their indices are appended to list (count: *n) for the FP64 recomputation of
// k_local_energy_guard.
struct ElGuard {
  int* list;   // (W) flagged walkers, NULL: no guard
  int* n;      // flagged count, zeroed before the launch
  double e_ref, width;
};

template <typename T, int WPB>
__global__ void __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(3))) k_local_energy_pts(Sys s, int64_t W, WalkT<T> wk,
                                                               const double* __restrict__ RTall, double* e_L,
                                                               double* kin_out, double* v_out, double* t_elec,
                                                               double* __restrict__ hws, ElGuard guard,
                                                               int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * WPB + wave;
  if (w >= W) return;
  El2Lds L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  double kinetic, pot;
  el_pts_walker<T, T>(s, wk.r_up + w * 3 * nu, wk.r_dn + w * 3 * nd, wk.Ginv + w * nu * nu, wk.pair + w * ne * no,
                      s.j3_on ? wk.jfield + w * ne * na : nullptr, RTall ? RTall + 9 * w : nullptr,
                      hws + w * (int64_t)ne * na, hws + W * (int64_t)ne * na + w * (int64_t)ne * no, L,
                      wave_team(lane), t_elec ? t_elec + w * ne : nullptr, kinetic, pot);
  if (lane == 0) {
    const double e = kinetic + pot;
    e_L[w] = e;
    if (kin_out) kin_out[w] = kinetic;
    if (v_out) v_out[w] = pot;
    if (guard.list && !(fabs(e - guard.e_ref) <= guard.width)) guard.list[atomicAdd(guard.n, 1)] = (int)w;
  }
}

// The same local energy with a team of NW waves per walker (one walker per block): the team deals
// the walker's shells, J2 partners and points to its waves and adds their partial sums in wave order
// (Team<NW>), so e_L is a fixed function of NW that differs from the one-wave result in the last
// bits.  Chosen by the JQMC_TUNE_EL_WAVES_PER_WALKER knob.
template <typename T, int NW>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(3))) k_local_energy_team(Sys s, int64_t W, WalkT<T> wk,
    const double* __restrict__ RTall, double* e_L, double* kin_out, double* v_out, double* t_elec,
    double* __restrict__ hws, ElGuard guard) {
  extern __shared__ double smem[];
  const int64_t w = blockIdx.x;
  if (w >= W) return;  // uniform across the block
  El2Lds L(reinterpret_cast<char*>(smem), s, NW);
  const Team<NW> tm{(int)threadIdx.x, wave_id(), lane_id(), L.red};
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  double kinetic, pot;
  el_pts_walker<T, T>(s, wk.r_up + w * 3 * nu, wk.r_dn + w * 3 * nd, wk.Ginv + w * nu * nu, wk.pair + w * ne * no,
                      s.j3_on ? wk.jfield + w * ne * na : nullptr, RTall ? RTall + 9 * w : nullptr,
                      hws + w * (int64_t)ne * na, hws + W * (int64_t)ne * na + w * (int64_t)ne * no, L,
                      tm, t_elec ? t_elec + w * ne : nullptr, kinetic, pot);
  if (threadIdx.x == 0) {
    const double e = kinetic + pot;
    e_L[w] = e;
    if (kin_out) kin_out[w] = kinetic;
    if (v_out) v_out[w] = pot;
    if (guard.list && !(fabs(e - guard.e_ref) <= guard.width)) guard.list[atomicAdd(guard.n, 1)] = (int)w;
  }
}

}  // namespace jqmc

namespace jqmc {

// FP64 scratch of one guard slot: AO rows, J3 fields, orbital and pair rows, the Gauss-Jordan
// matrices M and I (I ends as G^-1) and the products U = A J3 (one row per electron).
struct GuardRec {
  double *A, *g, *orb, *pair, *M, *I, *U;
  __host__ __device__ static int64_t doubles(const Sys& s) {
    const int64_t ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up;
    return 3 * ne * na + 2 * ne * no + 2 * nu * nu;
  }
  __host__ __device__ GuardRec(double* base, const Sys& s) {
    const int64_t ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up;
    A = base;
    g = A + ne * na;
    orb = g + ne * na;
    pair = orb + ne * no;
    M = pair + ne * no;
    I = M + nu * nu;
    U = I + nu * nu;
  }
};

constexpr int GUARD_SLOTS = 256;  // blocks of the guard pass; each strides over the flagged list
constexpr int GUARD_ROWS = 8;  // electron rows of A per pass of the J3 products

// LDS of a guard block: the local-energy block of its team of nw waves.
inline int64_t guard_lds_bytes(const Sys& s, int nw) { return el2_lds_bytes(s, nw); }

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// The walkers listed by k_local_energy_pts<float> recomputed entirely in FP64 from their positions:
// AO values at the electrons, orbital and pair rows, G and G^-1 (wave_gauss_jordan), the J3 fields
// (jastrow_factor.py:4121-4135) from the products A J3 and A J3^T, then el_pts_walker<double, double>;
// e_L (and the optional parts) of those walkers are overwritten.  The GT / 64 waves of a block share
// the AO, orbital, pair, G and J3-field work and, as one team, the local energy; the first wave runs
// the inversion.  The pass takes as long as one walker's recomputation, so the host launches the
// largest team whose LDS fits (8 waves, else 4).
// Blocks stride over the list and every block exits after the list's end.
template <int GT>
__global__ void __launch_bounds__(GT) k_local_energy_guard(Sys s, int64_t W, WalkT<float> wk,
                                                                    const double* __restrict__ RTall, double* e_L,
                                                                    double* kin_out, double* v_out, double* t_elec,
                                                                    double* __restrict__ hws,
                                                                    const int* __restrict__ list,
                                                                    const int* __restrict__ count,
                                                                    double* __restrict__ scratch, int64_t rec_doubles) {
  extern __shared__ double smem[];
  constexpr int GUARD_THREADS = GT;
  constexpr int n_waves = GT / WAVE;
  El2Lds L(reinterpret_cast<char*>(smem), s, n_waves);
  const int t = threadIdx.x, lane = lane_id(), wave = wave_id();
  const Team<n_waves> tm{t, wave, lane, L.red};
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb, nsh = s.n_shell;
  GuardRec R(scratch + (int64_t)blockIdx.x * rec_doubles, s);
  const int n_flag = *count;
  for (int i = blockIdx.x; i < n_flag; i += gridDim.x) {
    const int64_t w = list[i];
    const double* ru = wk.r_up + w * 3 * nu;
    const double* rd = wk.r_dn + w * 3 * nd;
    // AO values at every electron (shell radial factors once per shell), electrons over the waves
    for (int e = wave; e < ne; e += n_waves) {
      const double* r = e < nu ? ru + 3 * e : rd + 3 * (e - nu);
      const double x = r[0], y = r[1], z = r[2];
      for (int sh = lane; sh < nsh; sh += WAVE) shell_values<double>(s, sh, x, y, z, R.A + (int64_t)e * na);
    }
    __syncthreads();
    for (int e = wave; e < ne; e += n_waves) {
      double* ph = R.orb + (int64_t)e * no;
      if (s.orb_is_mo) {
        wave_matvec_T<double, double, double>(e < nu ? s.td.mo_upT : s.td.mo_dnT, R.A + (int64_t)e * na, na, no,
                                              lane, ph);
      } else {
        for (int k = lane; k < no; k += WAVE) ph[k] = R.A[(int64_t)e * na + k];
      }
    }
    __syncthreads();
    for (int q = t; q < ne * no; q += GUARD_THREADS) {
      const int e = q / no, k = q - e * no;
      const double* Lm = e < nu ? s.td.lam_p : s.td.lam_pT;
      double acc = 0.0;
      for (int j = 0; j < no; ++j) acc += Lm[(int64_t)j * no + k] * R.orb[(int64_t)e * no + j];
      R.pair[q] = acc;
    }
    __syncthreads();
    // G = [Phi_up^T Lambda_p Phi_dn | Phi_up^T Lambda_u] and, by the first wave, its inverse
    for (int q = t; q < nu * nu; q += GUARD_THREADS) {
      const int r = q / nu, cc = q - r * nu;
      double acc = 0.0;
      if (cc < nd) {
        for (int m = 0; m < no; ++m) acc += R.orb[(int64_t)r * no + m] * R.pair[(int64_t)(nu + cc) * no + m];
      } else {
        for (int m = 0; m < no; ++m) acc += R.orb[(int64_t)r * no + m] * s.td.lam_u[(int64_t)m * s.n_unp + (cc - nd)];
      }
      R.M[q] = acc;
      R.I[q] = r == cc ? 1.0 : 0.0;
    }
    __syncthreads();
    if (wave == 0) {
      double logabs, sgn;
      wave_gauss_jordan(R.M, R.I, nu, lane, logabs, sgn);
    }
    // J3 fields (jastrow_factor.py:4121-4135): g_e = j + J3^T sum_{f > e} a_f + J3 sum_{f < e} a_f (for an
    // up e the electrons after it are the later up and all dn ones; for a dn e, the later dn ones).
    // With V = A J3^T and U = A J3 (rows f): g_e = j + sum_{f > e} V_f + sum_{f < e} U_f, a suffix and
    // a prefix sum over the rows.  Threads own AO columns; V goes to g, U to the scratch rows.
    if (s.j3_on) {
      for (int a = t; a < na; a += GUARD_THREADS) {
        for (int f0 = 0; f0 < ne; f0 += GUARD_ROWS) {
          const int nf = ne - f0 < GUARD_ROWS ? ne - f0 : GUARD_ROWS;
          double v[GUARD_ROWS], u[GUARD_ROWS];
#pragma unroll
          for (int q = 0; q < GUARD_ROWS; ++q) v[q] = u[q] = 0.0;
          for (int b = 0; b < na; ++b) {
            const double jt = s.td.j3T[(int64_t)b * na + a], jj = s.td.j3[(int64_t)b * na + a];
#pragma unroll
            for (int q = 0; q < GUARD_ROWS; ++q) {
              const double ab = q < nf ? R.A[(int64_t)(f0 + q) * na + b] : 0.0;
              v[q] += jt * ab;
              u[q] += jj * ab;
            }
          }
#pragma unroll
          for (int q = 0; q < GUARD_ROWS; ++q)
            if (q < nf) {
              R.g[(int64_t)(f0 + q) * na + a] = v[q];
              R.U[(int64_t)(f0 + q) * na + a] = u[q];
            }
        }
        double suf = 0.0;
        for (int f = ne - 1; f >= 0; --f) {
          const double x = R.g[(int64_t)f * na + a];
          R.g[(int64_t)f * na + a] = s.td.j3v[a] + suf;
          suf += x;
        }
        double pre = 0.0;
        for (int f = 0; f < ne; ++f) {
          R.g[(int64_t)f * na + a] += pre;
          pre += R.U[(int64_t)f * na + a];
        }
      }
    }
    __syncthreads();
    double kin, pot;
    el_pts_walker<double, double>(s, ru, rd, R.I, R.pair, s.j3_on ? R.g : nullptr, RTall ? RTall + 9 * w : nullptr,
                                  hws + w * (int64_t)ne * na, hws + W * (int64_t)ne * na + w * (int64_t)ne * no, L, tm,
                                  t_elec ? t_elec + w * ne : nullptr, kin, pot);
    if (t == 0) {
      e_L[w] = kin + pot;
      if (kin_out) kin_out[w] = kin;
      if (v_out) v_out[w] = pot;
    }
    __syncthreads();
  }
}

}  // namespace jqmc
