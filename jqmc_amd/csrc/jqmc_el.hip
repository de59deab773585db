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
//      gradients and Laplacians;
//   5. per electron the nearest nucleus, the projector radial factors V_t and the J1 + J2 value at
//      r_e; the electrons whose nearest nucleus carries non-local terms form a list;
//   6. the non-local ECP (NN = 1, coulomb_potential.py:1477-1712) with the lanes owning the
//      quadrature points of all listed electrons, 64 at a time: each lane builds its point and
//      projector weight, the J1 / J2 change, and runs the AO basis shell after shell with uniform
//      table reads, accumulating a(r') . h_e (determinant ratio, FP64) and (a(r') - a_e) . g_e (J3
//      log-ratio, in T);
//   7. bare Coulomb and the local ECP part (coulomb_potential.py:1144-1246, 2252-2281).
//
// No lane idles on a per-AO table lookup: the AO tables sit behind a constant-address-space view,
// so their loads at uniform indices are scalar, and the per-lane loop bodies are streams of
// exponentials and fused multiply-adds over rows that are read in order.
#pragma once

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

// The q_e rows of all electrons stay in LDS when they take at most EL2_Q_LDS bytes, otherwise they
// go to the workspace after the h_e rows (large AO-orbital geminals).
constexpr int64_t EL2_Q_LDS = 16 * 1024;
__host__ __device__ inline bool el2_q_in_lds(const Sys& s) {
  return (int64_t)s.n_el * s.n_orb * (int64_t)sizeof(double) <= EL2_Q_LDS;
}

// LDS of one wavefront of k_local_energy_pts.
struct El2Lds {
  double *rr, *Gi, *RT, *gk, *Q, *en, *Vt, *jold;
  int* elist;
  uintptr_t end;
  __host__ __device__ El2Lds(char* base, const Sys& s) {
    const int ne = s.n_el, nv = s.nv > 0 ? s.nv : 1, mt = s.ecp_max_terms > 0 ? s.ecp_max_terms : 1;
    Carve c{base};
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(s.n_up * s.n_up);
    RT = c.take<double>(9);
    gk = c.take<double>(3 * nv);
    Q = c.take<double>(el2_q_in_lds(s) ? (int64_t)ne * s.n_orb : 0);
    en = c.take<double>(5 * ne);  // nearest ECP nucleus of each electron: n, R - r (3), |R - r|
    Vt = c.take<double>((int64_t)ne * mt);
    jold = c.take<double>(ne);
    elist = c.take<int>(ne);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

inline int64_t el2_lds_bytes(const Sys& s) { return (int64_t)El2Lds(0, s).end; }

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

// The local energy of one walker by one wavefront (phases 1-7 above).  Inputs: the walker's
// positions r_up / r_dn, G^-1 (Giw), pair rows (Qw), AO rows (Aw) and J3 fields (gw, NULL without J3)
// in the cache type TC, its rotation RTw (NULL: identity); workspace rows Hw (N x n_ao doubles) and,
// when the q rows do not fit LDS, Qg (N x n_orb doubles).  T is the arithmetic type of the AO values
// at the quadrature points.  Returns kinetic and potential energy on every lane; lane 0 writes the
// per-electron kinetic energies to t_elec_w when given.
template <typename T, typename TC>
__device__ void el_pts_walker(const Sys& s, const double* __restrict__ r_up, const double* __restrict__ r_dn,
                              const double* __restrict__ Giw, const double* __restrict__ Qw, const TC* Aw, const TC* gw,
                              const double* __restrict__ RTw, double* Hw, double* Qg, const El2Lds& L, int lane,
                              double* t_elec_w, double& kinetic_out, double& pot_out) {
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb, nsh = s.n_shell;
  // q_e rows: LDS, or the workspace
  double* const Qr = el2_q_in_lds(s) ? L.Q : Qg;

  // ---- 1. walker data into LDS
  for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = r_up[k];
  for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = r_dn[k];
  for (int k = lane; k < nu * nu; k += WAVE) L.Gi[k] = Giw[k];
  if (lane < 9) L.RT[lane] = RTw ? RTw[lane] : ((lane % 4) == 0 ? 1.0 : 0.0);
  wave_sync();
  const bool nonlocal = s.ecp_on && s.nv > 0;
  if (nonlocal) rotate_mesh(s, L.RT, L.gk, lane);

  // ---- 2. q_e for every electron
  for (int t = lane; t < ne * no; t += WAVE) {
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
  wave_sync();

  // ---- 3. h_e = C^T q_e into the workspace rows (the lane's AO column of C held in registers for the
  //      first KC orbitals, up block with C_up, dn block with C_dn)
  constexpr int KC = 16;
  for (int a = lane; a < na; a += WAVE) {
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
  wave_sync();

  const CPtr<int> sh0 = cview(s.sh_ao0);
  const CPtr<int> app = cview(s.ao_pp);
  const CPtr<int> meta = cview(s.ao_meta);
  const CPtr<double> ctr = cview(s.ao_ctr);

  // ---- 4. continuum kinetic energy, lanes own electrons
  double kinetic = 0.0;
  {
    const CPtr<double> pz = cview(s.td.prim_exp);
    const CPtr<double> pc = cview(s.td.prim_coef);
    const CPtr<double> sc = cview(s.td.ao_scale);
    for (int e0 = 0; e0 < ne; e0 += WAVE) {
      const int e = e0 + lane;
      const bool act = e < ne;
      const int ec = act ? e : 0;
      const double ex = L.rr[3 * ec], ey = L.rr[3 * ec + 1], ez = L.rr[3 * ec + 2];
      const double* hrow = Hw + (int64_t)ec * na;
      const TC* grow = gw ? gw + (int64_t)ec * na : nullptr;
      double acc8[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // gJ(3), lapJ, gD(3), lapD'
      for (int sh = 0; sh < nsh; ++sh) {
        const int a0 = sh0[sh], a1 = sh0[sh + 1];
        const double dx = ex - ctr[3 * a0], dy = ey - ctr[3 * a0 + 1], dz = ez - ctr[3 * a0 + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        double nr = 0.0, znr = 0.0, z2nr = 0.0;
        const int p1 = app[a0 + 1];
        for (int p = app[a0]; p < p1; ++p) {
          const double Z = pz[p];
          const double pref = pc[p] * exp(-Z * r2);
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
          const double f = sc[a];
          const double g0 = f * (nr * Px - 2.0 * dx * znr * P), g1 = f * (nr * Py - 2.0 * dy * znr * P),
                       g2 = f * (nr * Pz - 2.0 * dz * znr * P);
          const double lp = f * (nr * lapP - 4.0 * znr * rdg + (4.0 * r2 * z2nr - 6.0 * znr) * P);
          if (grow) {
            const double gj = (double)grow[a];
            acc8[0] += gj * g0;
            acc8[1] += gj * g1;
            acc8[2] += gj * g2;
            acc8[3] += gj * lp;
          }
          const double hd = hrow[a];
          acc8[4] += hd * g0;
          acc8[5] += hd * g1;
          acc8[6] += hd * g2;
          acc8[7] += hd * lp;
        }
      }
      if (s.j1_type) {
        for (int n = 0; n < s.n_nuc; ++n)
          j1_grad_lap(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], ex - s.nuc_pos[3 * n], ey - s.nuc_pos[3 * n + 1],
                      ez - s.nuc_pos[3 * n + 2], acc8[0], acc8[1], acc8[2], acc8[3]);
      }
      if (s.j2_type) {
        for (int f = 0; f < ne; ++f) {
          if (f == ec) continue;
          j2_grad_lap(s.j2_type, s.j2_param, ex - L.rr[3 * f], ey - L.rr[3 * f + 1], ez - L.rr[3 * f + 2], acc8[0],
                      acc8[1], acc8[2], acc8[3]);
        }
      }
      const double gDx = acc8[4], gDy = acc8[5], gDz = acc8[6];
      const double lapD = acc8[7] - (gDx * gDx + gDy * gDy + gDz * gDz);
      const double tx = acc8[0] + gDx, ty = acc8[1] + gDy, tz = acc8[2] + gDz;
      const double Te = -0.5 * (acc8[3] + lapD + (tx * tx + ty * ty + tz * tz));
      if (act) {
        kinetic += Te;
        if (t_elec_w) t_elec_w[e] = Te;
      }
    }
    kinetic = wave_sum(kinetic);
  }

  // ---- 5. nearest nucleus, projector radial factors and J1 + J2 at r_e per electron
  double vnl = 0.0;
  if (nonlocal) {
    const int mt = s.ecp_max_terms > 0 ? s.ecp_max_terms : 1;
    int n_ecp = 0;
    for (int e0 = 0; e0 < ne; e0 += WAVE) {
      const int e = e0 + lane;
      bool has = false;
      if (e < ne) {
        EcpElec E;
        {
          const double x = L.rr[3 * e], y = L.rr[3 * e + 1], z = L.rr[3 * e + 2];
          E.n = nearest_nucleus(s, x, y, z);
          E.nt = s.ecp_nt[E.n];
          E.rx = s.nuc_pos[3 * E.n] - x;
          E.ry = s.nuc_pos[3 * E.n + 1] - y;
          E.rz = s.nuc_pos[3 * E.n + 2] - z;
          E.rn = sqrt(E.rx * E.rx + E.ry * E.ry + E.rz * E.rz);
        }
        has = E.nt > 0;
        L.en[5 * e] = (double)E.n;
        L.en[5 * e + 1] = E.rx;
        L.en[5 * e + 2] = E.ry;
        L.en[5 * e + 3] = E.rz;
        L.en[5 * e + 4] = E.rn;
        for (int t = 0; t < E.nt; ++t) {
          const int o = E.n * s.ecp_max_terms + t;
          L.Vt[e * mt + t] =
              (1.0 / (E.rn * E.rn)) * s.ecp_coef[o] * rpow(E.rn, s.ecp_pow[o]) * exp(-s.ecp_exp[o] * (E.rn * E.rn));
        }
        if (has) {
          double old = 0.0;
          const double x = L.rr[3 * e], y = L.rr[3 * e + 1], z = L.rr[3 * e + 2];
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
      const int rank = n_ecp + __popcll(mask & ((1ull << lane) - 1ull));
      if (has) L.elist[rank] = e;
      n_ecp += __popcll(mask);
    }
    wave_sync();

    // ---- 6. non-local ECP, lanes own quadrature points
    const RealTabs<T>& R = tabs<T>(s);
    const CPtr<T> pz = cview(R.prim_exp);
    const CPtr<T> pc = cview(R.prim_coef);
    const CPtr<T> sc = cview(R.ao_scale);
    const int nv = s.nv;
    const int P = n_ecp * nv;
    for (int base = 0; base < P; base += WAVE) {
      const int idx = base + lane;
      const bool valid = idx < P;
      const int ie = valid ? idx / nv : 0, k = valid ? idx - ie * nv : 0;
      const int e = L.elist[ie];
      const double* En = L.en + 5 * e;
      const int n = (int)En[0];
      const double rx = En[1], ry = En[2], rz = En[3], rn = En[4];
      const double gx = L.gk[3 * k], gy = L.gk[3 * k + 1], gz = L.gk[3 * k + 2];
      const double gn = sqrt(gx * gx + gy * gy + gz * gz);
      const double ex = L.rr[3 * e], ey = L.rr[3 * e + 1], ez = L.rr[3 * e + 2];
      const double px = ex + (rx + rn * gx), py = ey + (ry + rn * gy), pzp = ez + (rz + rn * gz);
      const double cth = (-rx / rn) * (gx / gn) + (-ry / rn) * (gy / gn) + (-rz / rn) * (gz / gn);
      double pwt = 0.0;
      const int nt = s.ecp_nt[n];
      for (int t = 0; t < nt; ++t) {
        const int l = s.ecp_l[n * s.ecp_max_terms + t];
        pwt += L.Vt[e * mt + t] * ((2 * l + 1) * legendre(l, cth) * s.mesh_w[k]);
      }
      // J1 + J2 at the point
      double jn = 0.0;
      if (s.j2_type) {
        for (int f = 0; f < ne; ++f) {
          const double dx = px - L.rr[3 * f], dy = py - L.rr[3 * f + 1], dz = pzp - L.rr[3 * f + 2];
          const double u = j2_u(s.j2_type, s.j2_param, sqrt(dx * dx + dy * dy + dz * dz));
          jn += f == e ? 0.0 : u;
        }
      }
      if (s.j1_type) jn += j1_electron(s, px, py, pzp);
      // AO basis at the point: determinant ratio a(r') . h_e and J3 change (a(r') - a_e) . g_e
      const double* hrow = Hw + (int64_t)e * na;
      const TC* arow = Aw + (int64_t)e * na;
      const TC* grow = gw ? gw + (int64_t)e * na : nullptr;
      double det = 0.0;
      T dj3 = 0;
      for (int sh = 0; sh < nsh; ++sh) {
        const int a0 = sh0[sh], a1 = sh0[sh + 1];
        const T dx = (T)(px - ctr[3 * a0]), dy = (T)(py - ctr[3 * a0 + 1]), dz = (T)(pzp - ctr[3 * a0 + 2]);
        const T r2 = dx * dx + dy * dy + dz * dz;
        T nr = 0;
        const int p1 = app[a0 + 1];
        for (int p = app[a0]; p < p1; ++p) nr += pc[p] * exp(-pz[p] * r2);
        for (int a = a0; a < a1; ++a) {
          const T v = sc[a] * (nr * poly_uniform<T>(s, a, meta[a], dx, dy, dz));
          det += (double)v * hrow[a];
          if (grow) dj3 += (v - (T)arow[a]) * (T)grow[a];
        }
      }
      const double ratio = det * exp((double)dj3 + (jn - L.jold[e]));
      vnl += valid ? pwt * ratio : 0.0;
    }
  }

  // ---- 7. bare Coulomb + ECP local
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
  double red[2] = {pot, vnl};
  wave_sum_k<2>(red);
  kinetic_out = kinetic;
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
__global__ void __launch_bounds__(WPB * 64) k_local_energy_pts(Sys s, int64_t W, WalkT<T> wk,
                                                               const double* __restrict__ RTall, double* e_L,
                                                               double* kin_out, double* v_out, double* t_elec,
                                                               double* __restrict__ hws, ElGuard guard,
                                                               int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * WPB + wave;
  if (w >= W) return;
  El2Lds L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb;
  double kinetic, pot;
  el_pts_walker<T, T>(s, wk.r_up + w * 3 * nu, wk.r_dn + w * 3 * nd, wk.Ginv + w * nu * nu, wk.pair + w * ne * no,
                      wk.ao + w * ne * na, s.j3_on ? wk.jfield + w * ne * na : nullptr, RTall ? RTall + 9 * w : nullptr,
                      hws + w * (int64_t)ne * na, hws + W * (int64_t)ne * na + w * (int64_t)ne * no, L, lane,
                      t_elec ? t_elec + w * ne : nullptr, kinetic, pot);
  if (lane == 0) {
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
// matrices M and I (I ends as G^-1) and the prefix sums S1, S2 of one electron.
struct GuardRec {
  double *A, *g, *orb, *pair, *M, *I, *S1, *S2;
  __host__ __device__ static int64_t doubles(const Sys& s) {
    const int64_t ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up;
    return 2 * ne * na + 2 * ne * no + 2 * nu * nu + 2 * na;
  }
  __host__ __device__ GuardRec(double* base, const Sys& s) {
    const int64_t ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up;
    A = base;
    g = A + ne * na;
    orb = g + ne * na;
    pair = orb + ne * no;
    M = pair + ne * no;
    I = M + nu * nu;
    S1 = I + nu * nu;
    S2 = S1 + na;
  }
};

constexpr int GUARD_SLOTS = 128;  // blocks of the guard pass; each strides over the flagged list

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// The walkers listed by k_local_energy_pts<float> recomputed entirely in FP64 from their positions:
// AO values at the electrons, orbital and pair rows, G and G^-1 (wave_gauss_jordan), the J3 fields
// g_e = j + J3 S1[e] + J3^T S2[e] (jastrow_factor.py:4121-4135), then el_pts_walker<double, double>;
// e_L (and the optional parts) of those walkers are overwritten.  One wavefront per block, blocks
// stride over the list, every block exits after the list's end.
__global__ void __launch_bounds__(64) k_local_energy_guard(Sys s, int64_t W, WalkT<float> wk,
                                                           const double* __restrict__ RTall, double* e_L,
                                                           double* kin_out, double* v_out, double* t_elec,
                                                           double* __restrict__ hws, const int* __restrict__ list,
                                                           const int* __restrict__ count, double* __restrict__ scratch,
                                                           int64_t rec_doubles) {
  extern __shared__ double smem[];
  El2Lds L(reinterpret_cast<char*>(smem), s);
  const int lane = lane_id();
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb, nsh = s.n_shell;
  GuardRec R(scratch + (int64_t)blockIdx.x * rec_doubles, s);
  const int n_flag = *count;
  for (int i = blockIdx.x; i < n_flag; i += gridDim.x) {
    const int64_t w = list[i];
    const double* ru = wk.r_up + w * 3 * nu;
    const double* rd = wk.r_dn + w * 3 * nd;
    // AO values at every electron (shell radial factors once per shell), orbital and pair rows
    for (int e = 0; e < ne; ++e) {
      const double* r = e < nu ? ru + 3 * e : rd + 3 * (e - nu);
      const double x = r[0], y = r[1], z = r[2];
      for (int sh = lane; sh < nsh; sh += WAVE) shell_values<double>(s, sh, x, y, z, R.A + (int64_t)e * na);
    }
    wave_sync();
    for (int e = 0; e < ne; ++e) {
      double* ph = R.orb + (int64_t)e * no;
      if (s.orb_is_mo) {
        wave_matvec_T<double, double, double>(e < nu ? s.td.mo_upT : s.td.mo_dnT, R.A + (int64_t)e * na, na, no,
                                              lane, ph);
      } else {
        for (int k = lane; k < no; k += WAVE) ph[k] = R.A[(int64_t)e * na + k];
      }
    }
    wave_sync();
    for (int t = lane; t < ne * no; t += WAVE) {
      const int e = t / no, k = t - e * no;
      const double* Lm = e < nu ? s.td.lam_p : s.td.lam_pT;
      double q = 0.0;
      for (int j = 0; j < no; ++j) q += Lm[(int64_t)j * no + k] * R.orb[(int64_t)e * no + j];
      R.pair[t] = q;
    }
    wave_sync();
    // G = [Phi_up^T Lambda_p Phi_dn | Phi_up^T Lambda_u] and its inverse
    for (int t = lane; t < nu * nu; t += WAVE) {
      const int r = t / nu, c = t - r * nu;
      double acc = 0.0;
      if (c < nd) {
        for (int m = 0; m < no; ++m) acc += R.orb[(int64_t)r * no + m] * R.pair[(int64_t)(nu + c) * no + m];
      } else {
        for (int m = 0; m < no; ++m) acc += R.orb[(int64_t)r * no + m] * s.td.lam_u[(int64_t)m * s.n_unp + (c - nd)];
      }
      R.M[t] = acc;
      R.I[t] = r == c ? 1.0 : 0.0;
    }
    wave_sync();
    double logabs, sgn;
    wave_gauss_jordan(R.M, R.I, nu, lane, logabs, sgn);
    // J3 fields: S1 = sum over {up f > e} U {dn} (up e) or {dn f > e} (dn e), S2 the rest
    if (s.j3_on) {
      for (int e = 0; e < ne; ++e) {
        for (int a = lane; a < na; a += WAVE) {
          double s1 = 0.0, s2 = 0.0;
          for (int f = 0; f < ne; ++f) {
            if (f == e) continue;
            const bool in1 = e < nu ? (f > e) : (f >= nu && f > e);
            const double v = R.A[(int64_t)f * na + a];
            if (in1) s1 += v;
            else s2 += v;
          }
          R.S1[a] = s1;
          R.S2[a] = s2;
        }
        wave_sync();
        for (int a = lane; a < na; a += WAVE) {
          double acc = s.td.j3v[a];
          for (int b = 0; b < na; ++b) acc += s.td.j3T[(int64_t)b * na + a] * R.S1[b] + s.td.j3[(int64_t)b * na + a] * R.S2[b];
          R.g[(int64_t)e * na + a] = acc;
        }
        wave_sync();
      }
    }
    double kin, pot;
    el_pts_walker<double, double>(s, ru, rd, R.I, R.pair, R.A, s.j3_on ? R.g : nullptr, RTall ? RTall + 9 * w : nullptr,
                                  hws + w * (int64_t)ne * na, hws + W * (int64_t)ne * na + w * (int64_t)ne * no, L, lane,
                                  t_elec ? t_elec + w * ne : nullptr, kin, pot);
    if (lane == 0) {
      e_L[w] = kin + pot;
      if (kin_out) kin_out[w] = kin;
      if (v_out) v_out[w] = pot;
    }
    wave_sync();
  }
}

}  // namespace jqmc
