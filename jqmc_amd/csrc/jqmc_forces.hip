// gfx950 kernel of the position derivatives of the local energy and of ln|Psi| per walker, and
// of the space-warp force terms (jqmc_mcmc.py:744-852, swct.py:63-150).  The reference takes
// them by JAX reverse-mode differentiation of compute_local_energy and evaluate_ln_wavefunction;
// here the same derivatives come from a hand-written adjoint of the streaming E_L evaluation,
// one wavefront per walker, AO values recomputed (never stored) in every pass:
//
//   pass 1, per electron e: q_e, h_e = C^T q_e, AO value/gradient/Laplacian at r_e, the Jastrow
//     gradient gJ_e and Laplacian, T_e; the ECP quadrature ratios rho_k (electron_points_ratio),
//     then a second AO sweep over the points for the adjoints of the points, of the AO centres
//     and of the J3 field; the adjoint qbar_e = C w_e of q_e, the adjoint of g_e, and the J1/J2,
//     Coulomb and ECP-local pair terms.
//   pass 2: Gbar^-1 from the qbar_e and the pair rows, Gbar = -G^-T Gbar^-1 G^-T, and the
//     orbital-value adjoints phibar_e of every electron.
//   pass 3, per electron e: AO value/gradient/Hessian/grad-Laplacian at r_e contracted with
//     the AO-space adjoints (C^T phibar_e plus the J3 value adjoint J^T sum_{f<e} gbar_f +
//     J sum_{f>e} gbar_f) into dE/dr_e and, with the opposite sign, into the AO centres.
//   pass 4: per-nucleus sums of the centre adjoints, ion-ion forces, d ln|Psi|/dR, and the
//     space-warp products F_HF, F_PP.
//
// The kinetic energy is written T = -1/2 sum_e [L_e + 2 D1_e . gJ_e + lJ_e + |gJ_e|^2] with
// D1_e = grad a(r_e) . h_e and L_e = lap a(r_e) . h_e, so its adjoints are local to each
// electron's row/column of G plus the dependence of q_e on G^-1 and on the other spin's
// orbitals.  Plain FP64 systems only.
#pragma once

#include "jqmc_derivs.hip"

namespace jqmc {

// x^n e^{-Z x^2} / e^{-Z x^2} and its first three derivatives divided by the same exponential.
__device__ __forceinline__ void poly1d(int n, double Z, double x, double* d) {
  const double p0 = ipow(x, n), p1 = p0 * x, p2 = p1 * x, p3 = p2 * x;
  const double m1 = n >= 1 ? ipow(x, n - 1) : 0.0, m2 = n >= 2 ? ipow(x, n - 2) : 0.0,
               m3 = n >= 3 ? ipow(x, n - 3) : 0.0;
  const double fn = (double)n;
  d[0] = p0;
  d[1] = fn * m1 - 2.0 * Z * p1;
  d[2] = fn * (fn - 1.0) * m2 - 2.0 * Z * (2.0 * fn + 1.0) * p0 + 4.0 * Z * Z * p2;
  d[3] = fn * (fn - 1.0) * (fn - 2.0) * m3 - 6.0 * Z * fn * fn * m1 + 12.0 * Z * Z * (fn + 1.0) * p1 -
         8.0 * Z * Z * Z * p3;
}

// Value, gradient, Hessian (xx, yy, zz, xy, xz, yz) and gradient of the Laplacian of AO a at
// (x, y, z): each primitive-monomial product factorises into x^nx e^{-Z x^2} y^ny e^{-Z y^2}
// z^nz e^{-Z z^2}, so every derivative is a product of one-dimensional ones.
__device__ inline void ao_d3(const Sys& s, int a, double x, double y, double z, double& v, double* g, double* H,
                             double* gl) {
  const RealTabs<double>& R = s.td;
  const double dx = x - s.ao_ctr[3 * a], dy = y - s.ao_ctr[3 * a + 1], dz = z - s.ao_ctr[3 * a + 2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  v = 0.0;
  for (int c = 0; c < 3; ++c) g[c] = gl[c] = 0.0;
  for (int c = 0; c < 6; ++c) H[c] = 0.0;
  const int p1 = s.ao_pp[a + 1], t0 = s.ao_tp[a], t1 = s.ao_tp[a + 1];
  for (int p = s.ao_pp[a]; p < p1; ++p) {
    const double Z = R.prim_exp[p];
    const double e = R.prim_coef[p] * exp_nonpos(-Z * r2);
    for (int t = t0; t < t1; ++t) {
      const double k = R.term_coef[t] * e;
      double X[4], Y[4], W[4];
      poly1d(s.term_pow[3 * t], Z, dx, X);
      poly1d(s.term_pow[3 * t + 1], Z, dy, Y);
      poly1d(s.term_pow[3 * t + 2], Z, dz, W);
      v += k * (X[0] * Y[0] * W[0]);
      g[0] += k * (X[1] * Y[0] * W[0]);
      g[1] += k * (X[0] * Y[1] * W[0]);
      g[2] += k * (X[0] * Y[0] * W[1]);
      H[0] += k * (X[2] * Y[0] * W[0]);
      H[1] += k * (X[0] * Y[2] * W[0]);
      H[2] += k * (X[0] * Y[0] * W[2]);
      H[3] += k * (X[1] * Y[1] * W[0]);
      H[4] += k * (X[1] * Y[0] * W[1]);
      H[5] += k * (X[0] * Y[1] * W[1]);
      gl[0] += k * (X[3] * Y[0] * W[0] + X[1] * Y[2] * W[0] + X[1] * Y[0] * W[2]);
      gl[1] += k * (X[2] * Y[1] * W[0] + X[0] * Y[3] * W[0] + X[0] * Y[1] * W[2]);
      gl[2] += k * (X[2] * Y[0] * W[1] + X[0] * Y[2] * W[1] + X[0] * Y[0] * W[3]);
    }
  }
}

// f', f'', f''' of the one-body term of one electron-nucleus pair (jastrow_factor.py:672-711).
__device__ __forceinline__ void j1_f123(int type, double a, double c, double A, double r, double& f1, double& f2,
                                        double& f3) {
  if (type == 1) {
    const double e = exp_nonpos(-a * c * r);
    f1 = -A * c / 2.0 * e;
    f2 = A * a * c * c / 2.0 * e;
    f3 = -A * a * a * c * c * c / 2.0 * e;
  } else {
    const double den = 1.0 + a * c * r;
    f1 = -A / (2.0 * den * den);
    f2 = A * a * c / (den * den * den);
    f3 = -3.0 * A * a * a * c * c / (den * den * den * den);
  }
}

// f', f'', f''' of the two-body pair function (jastrow_factor.py:1207-1230).
__device__ __forceinline__ void j2_f123(int type, double a, double r, double& f1, double& f2, double& f3) {
  if (type == 2) {
    const double den = 1.0 + a * r;
    f1 = 0.5 / (den * den);
    f2 = -a / (den * den * den);
    f3 = 3.0 * a * a / (den * den * den * den);
  } else {
    const double e = exp_nonpos(-a * r);
    f1 = 0.5 * e;
    f2 = -a / 2.0 * e;
    f3 = a * a / 2.0 * e;
  }
}

// d/du [ dJ . grad f(|u|) + lc lap f(|u|) ] = (f'' (u^.dJ) - f'/r (u^.dJ) + lc (f''' + 2f''/r - 2f'/r^2)) u^
// + f'/r dJ, into out (added).
__device__ __forceinline__ void radial_adjoint(double ux, double uy, double uz, double f1, double f2, double f3,
                                               const double* dJ, double lc, double* out) {
  const double r = fmax(sqrt(ux * ux + uy * uy + uz * uz), 1.0e-12);
  const double hx = ux / r, hy = uy / r, hz = uz / r;
  const double ud = hx * dJ[0] + hy * dJ[1] + hz * dJ[2];
  const double c = f2 * ud - f1 / r * ud + lc * (f3 + 2.0 * f2 / r - 2.0 * f1 / (r * r));
  out[0] += c * hx + f1 / r * dJ[0];
  out[1] += c * hy + f1 / r * dJ[1];
  out[2] += c * hz + f1 / r * dJ[2];
}

// dP_n/dx of the tabulated Legendre polynomials P_0..P_6.
__device__ __forceinline__ double legendre_d(int n, double x) {
  switch (n) {
    case 0: return 0.0;
    case 1: return 1.0;
    case 2: return 3.0 * x;
    case 3: return 0.5 * (15.0 * x * x - 3.0);
    case 4: return 0.125 * (140.0 * x * x * x - 60.0 * x);
    case 5: return 0.125 * (315.0 * x * x * x * x - 210.0 * x * x + 15.0);
    case 6: return (1386.0 * x * x * x * x * x - 1260.0 * x * x * x + 210.0 * x) / 16.0;
    default: return __builtin_nan("");
  }
}

// Gradient w.r.t. u = r_e - R of the projector factor W_k = sum_t V_t(d) (2l+1) P_l(u^ . g^_k) w_k.
__device__ __forceinline__ void ecp_weight_grad(const Sys& s, const EcpElec& E, const double* Vt, const double* gk,
                                                int k, double* dW) {
  const double gx = gk[3 * k], gy = gk[3 * k + 1], gz = gk[3 * k + 2];
  const double gn = sqrt(gx * gx + gy * gy + gz * gz);
  const double hx = -E.rx / E.rn, hy = -E.ry / E.rn, hz = -E.rz / E.rn;  // u^ = (r - R)/|r - R|
  const double cth = hx * (gx / gn) + hy * (gy / gn) + hz * (gz / gn);
  double rad = 0.0, ang = 0.0;
  for (int t = 0; t < E.nt; ++t) {
    const int o = E.n * s.ecp_max_terms + t;
    const int l = s.ecp_l[o];
    const double dV = Vt[t] * ((s.ecp_pow[o] - 2.0) / E.rn - 2.0 * s.ecp_exp[o] * E.rn);
    rad += dV * ((2 * l + 1) * legendre(l, cth) * s.mesh_w[k]);
    ang += Vt[t] * ((2 * l + 1) * legendre_d(l, cth) * s.mesh_w[k]);
  }
  const double ai = ang / E.rn;
  dW[0] = rad * hx + ai * (gx / gn - cth * hx);
  dW[1] = rad * hy + ai * (gy / gn - cth * hy);
  dW[2] = rad * hz + ai * (gz / gn - cth * hz);
}

// LDS layout of one wavefront of the forces kernel.
struct FrLds {
  double *rr, *Gi, *RT, *gk, *pts, *pw, *det, *dj, *Vt, *cw, *sk, *pb;
  double *q, *h, *ga, *la, *wv, *plt, *gtot, *t, *rad;
  double *rbar, *Rbar, *lnR, *el;
  uintptr_t end;
  static constexpr int EL = 10;  // per electron: gJ(3), dJ(3), D1(3), S_e
  __host__ __device__ FrLds(char* base, const Sys& s) {
    const int ne = s.n_el, nu = s.n_up, nv = s.nv > 0 ? s.nv : 1, na = s.n_ao, no = s.n_orb;
    Carve c{base};
    rr = c.take<double>(3 * ne);
    Gi = c.take<double>(nu * nu);
    RT = c.take<double>(9);
    gk = c.take<double>(3 * nv);
    pts = c.take<double>(3 * nv);
    pw = c.take<double>(nv);
    det = c.take<double>(nv);
    dj = c.take<double>(nv);
    Vt = c.take<double>(s.ecp_max_terms > 0 ? s.ecp_max_terms : 1);
    cw = c.take<double>(nv);
    sk = c.take<double>(nv);
    pb = c.take<double>(3 * nv);
    q = c.take<double>(no);
    h = c.take<double>(na);
    ga = c.take<double>(3 * na);
    la = c.take<double>(na);
    wv = c.take<double>(na);
    plt = c.take<double>(na);
    gtot = c.take<double>(na);
    t = c.take<double>(no > nu ? no : nu);
    rbar = c.take<double>(3 * ne);
    Rbar = c.take<double>(3 * s.n_nuc);
    lnR = c.take<double>(3 * s.n_nuc);
    el = c.take<double>(EL * ne);
    rad = c.take<double>((int64_t)s.n_shell * PTS_CHUNK);
    end = reinterpret_cast<uintptr_t>(c.p);
  }
};

inline int64_t forces_lds_bytes(const Sys& s) { return (int64_t)FrLds(0, s).end; }

// Per-walker global workspace (doubles).
struct FrWs {
  double *qbar, *phibar, *gbar, *RA, *RL, *Gb, *Gb2;
  static __host__ __device__ int64_t per_walker(const Sys& s) {
    const int ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up;
    return 2 * (int64_t)ne * no + (s.j3_on ? (int64_t)ne * na : 0) + 6 * (int64_t)na + 2 * (int64_t)nu * nu;
  }
  __host__ __device__ FrWs(double* base, const Sys& s, int64_t w) {
    const int ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up;
    double* p = base + w * per_walker(s);
    qbar = p;
    p += (int64_t)ne * no;
    phibar = p;
    p += (int64_t)ne * no;
    gbar = s.j3_on ? p : nullptr;
    if (s.j3_on) p += (int64_t)ne * na;
    RA = p;
    p += 3 * (int64_t)na;
    RL = p;
    p += 3 * (int64_t)na;
    Gb = p;
    p += (int64_t)nu * nu;
    Gb2 = p;
  }
};

// q_e of electron e into q (orbital space) and h = C^T q into h (AO space), as in electron_fields.
__device__ inline void electron_q_h(const Sys& s, const double* Gi, const double* Qw, int e, double* q, double* h,
                                    int lane) {
  const int nu = s.n_up, nd = s.n_dn, na = s.n_ao, no = s.n_orb;
  const RealTabs<double>& R = s.td;
  const bool up = e < nu;
  for (int k = lane; k < no; k += WAVE) {
    double acc = 0.0;
    if (up) {
      for (int j = 0; j < nd; ++j) acc += Qw[(int64_t)(nu + j) * no + k] * Gi[j * nu + e];
      for (int u = 0; u < s.n_unp; ++u) acc += R.lam_u[(int64_t)k * s.n_unp + u] * Gi[(nd + u) * nu + e];
    } else {
      const int j = e - nu;
      for (int i = 0; i < nu; ++i) acc += Qw[(int64_t)i * no + k] * Gi[j * nu + i];
    }
    q[k] = acc;
  }
  wave_sync();
  const double* C = up ? R.mo_up : R.mo_dn;
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

// One wavefront per walker; outputs (any may be NULL except the ones the caller asked for):
// e_L (W), dE_dr (W, N, 3), dE_dR (W, n_nuc, 3), dlnpsi_dr (W, N, 3), dlnpsi_dR (W, n_nuc, 3),
// force_HF / force_PP (W, n_nuc, 3) with the space-warp weights when swct != 0.
__global__ void __launch_bounds__(256) k_forces(Sys s, int64_t W, WalkT<double> wk, const double* __restrict__ RTall,
                                                double* ws_base, double* e_L, double* dE_dr, double* dE_dR,
                                                double* dlnpsi_dr, double* dlnpsi_dR, double* force_HF,
                                                double* force_PP, int swct, int wpb, int64_t lds_per_wave) {
  extern __shared__ double smem[];
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  FrLds L(reinterpret_cast<char*>(smem) + wave * lds_per_wave, s);
  FrWs X(ws_base, s, w);
  const RealTabs<double>& R = s.td;
  const int nu = s.n_up, nd = s.n_dn, ne = s.n_el, na = s.n_ao, no = s.n_orb, nn = s.n_nuc;
  const double* Aw = wk.ao + w * ne * na;
  const double* Qw = wk.pair + w * ne * no;
  const double* gw = s.j3_on ? wk.jfield + w * ne * na : nullptr;
  constexpr int EL = FrLds::EL;

  for (int k = lane; k < 3 * nu; k += WAVE) L.rr[k] = wk.r_up[w * 3 * nu + k];
  for (int k = lane; k < 3 * nd; k += WAVE) L.rr[3 * nu + k] = wk.r_dn[w * 3 * nd + k];
  for (int k = lane; k < nu * nu; k += WAVE) L.Gi[k] = wk.Ginv[w * nu * nu + k];
  if (lane < 9) L.RT[lane] = RTall ? RTall[9 * w + lane] : ((lane % 4) == 0 ? 1.0 : 0.0);
  for (int k = lane; k < 3 * ne; k += WAVE) L.rbar[k] = 0.0;
  for (int k = lane; k < 3 * nn; k += WAVE) L.Rbar[k] = L.lnR[k] = 0.0;
  for (int k = lane; k < 3 * na; k += WAVE) X.RA[k] = X.RL[k] = 0.0;
  for (int a = lane; a < na; a += WAVE) L.plt[a] = L.gtot[a] = 0.0;
  wave_sync();
  const bool nonlocal = s.ecp_on && s.nv > 0;
  if (nonlocal) rotate_mesh(s, L.RT, L.gk, lane);
  wave_sync();

  // ---------------------------------------------------------------- pass 1
  double kinetic = 0.0, vnl = 0.0;
  for (int e = 0; e < ne; ++e) {
    const bool up = e < nu;
    const double ex = L.rr[3 * e], ey = L.rr[3 * e + 1], ez = L.rr[3 * e + 2];
    electron_q_h(s, L.Gi, Qw, e, L.q, L.h, lane);
    const double* ge = gw ? gw + (int64_t)e * na : nullptr;
    // fields at r_e: gJ (total Jastrow gradient), gJ12 (J1 + J2 part), lJ, D1, L_e
    double acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = lane; a < na; a += WAVE) {
      double v, g0, g1, g2, lp;
      ao_vgl<double>(s, a, ex, ey, ez, v, g0, g1, g2, lp);
      L.ga[3 * a] = g0;
      L.ga[3 * a + 1] = g1;
      L.ga[3 * a + 2] = g2;
      L.la[a] = lp;
      L.wv[a] = 0.0;
      const double hv = L.h[a], gj = ge ? ge[a] : 0.0;
      acc[0] += gj * g0;
      acc[1] += gj * g1;
      acc[2] += gj * g2;
      acc[3] += gj * lp;
      acc[4] += hv * g0;
      acc[5] += hv * g1;
      acc[6] += hv * g2;
      acc[7] += hv * lp;
      const double c = hv + gj;  // d ln|Psi| / d a_a(r_e) in the union basis
      X.RL[3 * a] += c * g0;
      X.RL[3 * a + 1] += c * g1;
      X.RL[3 * a + 2] += c * g2;
    }
    if (s.j1_type) {
      for (int n = lane; n < nn; n += WAVE) {
        double g[3] = {0, 0, 0}, l = 0.0;
        j1_grad_lap(s.j1_type, s.j1_param, s.j1_c[n], s.j1_A[n], ex - s.nuc_pos[3 * n], ey - s.nuc_pos[3 * n + 1],
                    ez - s.nuc_pos[3 * n + 2], g[0], g[1], g[2], l);
        for (int c = 0; c < 3; ++c) {
          acc[c] += g[c];
          acc[8 + c] += g[c];
          L.lnR[3 * n + c] -= g[c];  // d J1 / d R_n = - d J1 / d r_e for this pair
        }
        acc[3] += l;
      }
    }
    if (s.j2_type) {
      for (int f = lane; f < ne; f += WAVE) {
        if (f == e) continue;
        double g[3] = {0, 0, 0}, l = 0.0;
        j2_grad_lap(s.j2_type, s.j2_param, ex - L.rr[3 * f], ey - L.rr[3 * f + 1], ez - L.rr[3 * f + 2], g[0], g[1],
                    g[2], l);
        for (int c = 0; c < 3; ++c) {
          acc[c] += g[c];
          acc[8 + c] += g[c];
        }
        acc[3] += l;
      }
    }
    wave_sum_k<11>(acc);
    const double gJ[3] = {acc[0], acc[1], acc[2]};
    const double D1[3] = {acc[4], acc[5], acc[6]};
    const double dJ[3] = {-(D1[0] + gJ[0]), -(D1[1] + gJ[1]), -(D1[2] + gJ[2])};
    kinetic += -0.5 * (acc[7] + 2.0 * (D1[0] * gJ[0] + D1[1] * gJ[1] + D1[2] * gJ[2]) + acc[3] +
                       (gJ[0] * gJ[0] + gJ[1] * gJ[1] + gJ[2] * gJ[2]));
    if (lane == 0) {
      for (int c = 0; c < 3; ++c) {
        L.el[EL * e + c] = gJ[c];
        L.el[EL * e + 3 + c] = dJ[c];
        L.el[EL * e + 6 + c] = D1[c];
      }
      L.el[EL * e + 9] = 0.0;
    }
    double re[3] = {0.0, 0.0, 0.0};  // lane part of dE/dr_e
    bool ecp_done = false;           // this electron has quadrature points (wave-uniform)

    // ---- ECP non-local: ratios, then the adjoint sweep over the quadrature points
    if (nonlocal) {
      const EcpElec E = ecp_electron(s, L.rr, e, L.Vt, lane);
      if (E.nt > 0) {
        wave_sync();
        for (int k = lane; k < s.nv; k += WAVE)
          ecp_point(s, E, L.Vt, L.rr, L.gk, e, k, L.pts[3 * k], L.pts[3 * k + 1], L.pts[3 * k + 2], L.pw[k]);
        wave_sync();
        electron_points_ratio<double, EL_PTS_CHUNK>(s, L.rr, e, L.h, Aw + (int64_t)e * na, ge, L.pts, s.nv, L.det,
                                                    L.dj, L.rad, lane);
        double S = 0.0, wgeo[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < s.nv; ++k) {
          const double Ek = exp(L.dj[k]);
          const double rho = L.det[k] * Ek;
          S += L.pw[k] * rho;
          if (lane == k) {
            L.cw[k] = L.pw[k] * Ek;
            L.sk[k] = L.pw[k] * rho;
            // projector geometry: rho_k d W_k / d u with u = r_e - R_a
            double dW[3];
            ecp_weight_grad(s, E, L.Vt, L.gk, k, dW);
            for (int c = 0; c < 3; ++c) wgeo[c] += rho * dW[c];
          }
        }
        wave_sum_k<3>(wgeo);
        ecp_done = true;
        if (lane == 0) {
          for (int c = 0; c < 3; ++c) {
            re[c] += wgeo[c];
            L.Rbar[3 * E.n + c] -= wgeo[c];
          }
        }
        vnl += S;
        if (lane == 0) L.el[EL * e + 9] = S;
        for (int k = lane; k < 3 * s.nv; k += WAVE) L.pb[k] = 0.0;
        wave_sync();
        // AO sweep: point adjoints, centre adjoints, w_e (ECP part) and the J3 field adjoint
        double* gb = X.gbar ? X.gbar + (int64_t)e * na : nullptr;
        for (int c0 = 0; c0 < s.nv; c0 += PTS_CHUNK) {
          const int nq = (s.nv - c0) < PTS_CHUNK ? (s.nv - c0) : PTS_CHUNK;
          double pbl[3 * PTS_CHUNK];
#pragma unroll
          for (int k = 0; k < 3 * PTS_CHUNK; ++k) pbl[k] = 0.0;
          for (int a = lane; a < na; a += WAVE) {
            const double hv = L.h[a], gj = ge ? ge[a] : 0.0, a0 = Aw[(int64_t)e * na + a];
            double wsum = 0.0, gsum = 0.0, ra[3] = {0.0, 0.0, 0.0};
            for (int qi = 0; qi < nq; ++qi) {
              const int k = c0 + qi;
              double v, g0, g1, g2, lp;
              ao_vgl<double>(s, a, L.pts[3 * k], L.pts[3 * k + 1], L.pts[3 * k + 2], v, g0, g1, g2, lp);
              const double cf = L.cw[k] * hv + L.sk[k] * gj;
              pbl[3 * qi] += cf * g0;
              pbl[3 * qi + 1] += cf * g1;
              pbl[3 * qi + 2] += cf * g2;
              ra[0] += cf * g0;
              ra[1] += cf * g1;
              ra[2] += cf * g2;
              wsum += L.cw[k] * v;
              gsum += L.sk[k] * (v - a0);
            }
            X.RA[3 * a] -= ra[0];
            X.RA[3 * a + 1] -= ra[1];
            X.RA[3 * a + 2] -= ra[2];
            L.wv[a] += wsum;
            if (gb) gb[a] = (c0 == 0 ? 0.0 : gb[a]) + gsum;
          }
          // J1 / J2 at the points: lanes own nuclei / other electrons
          for (int i = lane; i < ne + nn; i += WAVE) {
            const bool is_el = i < ne;
            if (is_el && (!s.j2_type || i == e)) continue;
            if (!is_el && !s.j1_type) continue;
            const int m = is_el ? i : i - ne;
            const double Px = is_el ? L.rr[3 * m] : s.nuc_pos[3 * m];
            const double Py = is_el ? L.rr[3 * m + 1] : s.nuc_pos[3 * m + 1];
            const double Pz = is_el ? L.rr[3 * m + 2] : s.nuc_pos[3 * m + 2];
            double own[3] = {0.0, 0.0, 0.0};
            for (int qi = 0; qi < nq; ++qi) {
              const int k = c0 + qi;
              const double ux = L.pts[3 * k] - Px, uy = L.pts[3 * k + 1] - Py, uz = L.pts[3 * k + 2] - Pz;
              const double r = fmax(sqrt(ux * ux + uy * uy + uz * uz), 1.0e-12);
              double f1, f2, f3;
              if (is_el)
                j2_f123(s.j2_type, s.j2_param, r, f1, f2, f3);
              else
                j1_f123(s.j1_type, s.j1_param, s.j1_c[m], s.j1_A[m], r, f1, f2, f3);
              const double c = L.sk[k] * f1 / r;
              pbl[3 * qi] += c * ux;
              pbl[3 * qi + 1] += c * uy;
              pbl[3 * qi + 2] += c * uz;
              own[0] -= c * ux;
              own[1] -= c * uy;
              own[2] -= c * uz;
            }
            double* dst = is_el ? L.rbar + 3 * m : L.Rbar + 3 * m;
            dst[0] += own[0];
            dst[1] += own[1];
            dst[2] += own[2];
          }
          wave_sum_k<3 * PTS_CHUNK>(pbl);
          wave_sync();
          if (lane == 0) {
            const double hx = -E.rx / E.rn, hy = -E.ry / E.rn, hz = -E.rz / E.rn;
            for (int qi = 0; qi < nq; ++qi) {
              const int k = c0 + qi;
              const double gp = L.gk[3 * k] * pbl[3 * qi] + L.gk[3 * k + 1] * pbl[3 * qi + 1] +
                                L.gk[3 * k + 2] * pbl[3 * qi + 2];
              // p = R_a + |r_e - R_a| g~_k
              re[0] += hx * gp;
              re[1] += hy * gp;
              re[2] += hz * gp;
              L.Rbar[3 * E.n] += pbl[3 * qi] - hx * gp;
              L.Rbar[3 * E.n + 1] += pbl[3 * qi + 1] - hy * gp;
              L.Rbar[3 * E.n + 2] += pbl[3 * qi + 2] - hz * gp;
            }
          }
          wave_sync();
        }
        // - S_e (J1 + J2)(r_e): the denominator of every ratio
        for (int i = lane; i < ne + nn; i += WAVE) {
          const bool is_el = i < ne;
          if (is_el && (!s.j2_type || i == e)) continue;
          if (!is_el && !s.j1_type) continue;
          const int m = is_el ? i : i - ne;
          const double ux = ex - (is_el ? L.rr[3 * m] : s.nuc_pos[3 * m]);
          const double uy = ey - (is_el ? L.rr[3 * m + 1] : s.nuc_pos[3 * m + 1]);
          const double uz = ez - (is_el ? L.rr[3 * m + 2] : s.nuc_pos[3 * m + 2]);
          const double r = fmax(sqrt(ux * ux + uy * uy + uz * uz), 1.0e-12);
          double f1, f2, f3;
          if (is_el)
            j2_f123(s.j2_type, s.j2_param, r, f1, f2, f3);
          else
            j1_f123(s.j1_type, s.j1_param, s.j1_c[m], s.j1_A[m], r, f1, f2, f3);
          const double c = S * f1 / r;
          re[0] -= c * ux;
          re[1] -= c * uy;
          re[2] -= c * uz;
          double* dst = is_el ? L.rbar + 3 * m : L.Rbar + 3 * m;
          dst[0] += c * ux;
          dst[1] += c * uy;
          dst[2] += c * uz;
        }
      }
    }
    wave_sync();

    // ---- kinetic adjoints at r_e: J3 field adjoint, w_e, qbar_e = C w_e
    {
      double* gb = X.gbar ? X.gbar + (int64_t)e * na : nullptr;
      for (int a = lane; a < na; a += WAVE) {
        const double g0 = L.ga[3 * a], g1 = L.ga[3 * a + 1], g2 = L.ga[3 * a + 2], lp = L.la[a];
        if (gb) gb[a] = (ecp_done ? gb[a] : 0.0) + (dJ[0] * g0 + dJ[1] * g1 + dJ[2] * g2 - 0.5 * lp);
        L.wv[a] += -(gJ[0] * g0 + gJ[1] * g1 + gJ[2] * g2) - 0.5 * lp;
      }
      wave_sync();
      const double* C = up ? R.mo_up : R.mo_dn;
      double* qb = X.qbar + (int64_t)e * no;
      for (int k = lane; k < no; k += WAVE) {
        double acc2 = 0.0;
        if (s.orb_is_mo) {
          for (int a = 0; a < na; ++a) acc2 += C[(int64_t)k * na + a] * L.wv[a];
        } else {
          acc2 = L.wv[k];
        }
        qb[k] = acc2;
      }
      if (gb) {
        for (int a = lane; a < na; a += WAVE) L.gtot[a] += gb[a];
      }
    }

    // ---- pair terms: J1 / J2 kinetic adjoints, bare Coulomb, ECP local
    for (int i = lane; i < ne + nn; i += WAVE) {
      const bool is_el = i < ne;
      if (is_el && i == e) continue;
      const int m = is_el ? i : i - ne;
      const double ux = ex - (is_el ? L.rr[3 * m] : s.nuc_pos[3 * m]);
      const double uy = ey - (is_el ? L.rr[3 * m + 1] : s.nuc_pos[3 * m + 1]);
      const double uz = ez - (is_el ? L.rr[3 * m + 2] : s.nuc_pos[3 * m + 2]);
      const double d2 = ux * ux + uy * uy + uz * uz, d = sqrt(d2);
      double v[3] = {0.0, 0.0, 0.0};
      double f1, f2, f3;
      if (is_el) {
        if (s.j2_type) {
          j2_f123(s.j2_type, s.j2_param, fmax(d, 1.0e-12), f1, f2, f3);
          radial_adjoint(ux, uy, uz, f1, f2, f3, dJ, -0.5, v);
        }
        if (m > e) {  // e-e repulsion, each pair once
          const double c = -1.0 / (d2 * d);
          v[0] += c * ux;
          v[1] += c * uy;
          v[2] += c * uz;
        }
      } else {
        if (s.j1_type) {
          j1_f123(s.j1_type, s.j1_param, s.j1_c[m], s.j1_A[m], fmax(d, 1.0e-12), f1, f2, f3);
          radial_adjoint(ux, uy, uz, f1, f2, f3, dJ, -0.5, v);
        }
        double c = s.nuc_charge[m] / (d2 * d);  // -Z / d
        if (s.ecp_on) {
          for (int t = 0; t < s.ecpl_nt[m]; ++t) {
            const int o = m * s.ecp_max_local + t;
            const double Vl = (1.0 / d2) * s.ecpl_coef[o] * rpow(d, s.ecpl_pow[o]) * exp_nonpos(-s.ecpl_exp[o] * d2);
            c += Vl * ((s.ecpl_pow[o] - 2.0) / d - 2.0 * s.ecpl_exp[o] * d) / d;
          }
        }
        v[0] += c * ux;
        v[1] += c * uy;
        v[2] += c * uz;
      }
      re[0] += v[0];
      re[1] += v[1];
      re[2] += v[2];
      double* dst = is_el ? L.rbar + 3 * m : L.Rbar + 3 * m;
      dst[0] -= v[0];
      dst[1] -= v[1];
      dst[2] -= v[2];
    }
    wave_sum_k<3>(re);
    wave_sync();
    if (lane == 0) {
      L.rbar[3 * e] += re[0];
      L.rbar[3 * e + 1] += re[1];
      L.rbar[3 * e + 2] += re[2];
    }
    wave_sync();
  }

  // ---------------------------------------------------------------- pass 2
  // Gbar^-1[j][c]: row j < n_dn: pair_dn_j . qbar_c (up c) + pair_up_c . qbar_{n_up + j};
  // rows j >= n_dn: Lambda_u[:, j - n_dn] . qbar_c
  for (int t = lane; t < nu * nu; t += WAVE) {
    const int j = t / nu, c = t - j * nu;
    double acc2 = 0.0;
    const double* qc = X.qbar + (int64_t)c * no;
    if (j < nd) {
      const double* pj = Qw + (int64_t)(nu + j) * no;
      const double* pc = Qw + (int64_t)c * no;
      const double* qj = X.qbar + (int64_t)(nu + j) * no;
      for (int k = 0; k < no; ++k) acc2 += pj[k] * qc[k] + pc[k] * qj[k];
    } else {
      const int u = j - nd;
      for (int k = 0; k < no; ++k) acc2 += R.lam_u[(int64_t)k * s.n_unp + u] * qc[k];
    }
    X.Gb[t] = acc2;
  }
  wave_sync();
  // Gbar = -G^-T Gbar^-1 G^-T: first Y = Gbar^-1 G^-T (Y[j][c] = sum_m Gb[j][m] Gi[c][m]) ...
  for (int t = lane; t < nu * nu; t += WAVE) {
    const int j = t / nu, c = t - j * nu;
    double acc2 = 0.0;
    for (int m = 0; m < nu; ++m) acc2 += X.Gb[j * nu + m] * L.Gi[c * nu + m];
    X.Gb2[t] = acc2;
  }
  wave_sync();
  // ... then Gbar[i][c] = -sum_j Gi[j][i] Y[j][c], into Gb
  for (int t = lane; t < nu * nu; t += WAVE) {
    const int i = t / nu, c = t - i * nu;
    double acc2 = 0.0;
    for (int j = 0; j < nu; ++j) acc2 += L.Gi[j * nu + i] * X.Gb2[j * nu + c];
    X.Gb[t] = -acc2;
  }
  wave_sync();
  // phibar of every electron (orbital space)
  for (int e = 0; e < ne; ++e) {
    const bool up = e < nu;
    // t = sum over the other spin's electrons f of G^-1[..] qbar_f
    for (int l = lane; l < no; l += WAVE) {
      double acc2 = 0.0;
      if (up) {
        for (int j = 0; j < nd; ++j) acc2 += L.Gi[j * nu + e] * X.qbar[(int64_t)(nu + j) * no + l];
      } else {
        const int j = e - nu;
        for (int c = 0; c < nu; ++c) acc2 += L.Gi[j * nu + c] * X.qbar[(int64_t)c * no + l];
      }
      L.t[l] = acc2;
    }
    wave_sync();
    double* pb = X.phibar + (int64_t)e * no;
    for (int k = lane; k < no; k += WAVE) {
      double acc2 = 0.0;
      if (up) {
        for (int j = 0; j < nd; ++j) acc2 += X.Gb[e * nu + j] * Qw[(int64_t)(nu + j) * no + k];
        for (int u = 0; u < s.n_unp; ++u) acc2 += X.Gb[e * nu + nd + u] * R.lam_u[(int64_t)k * s.n_unp + u];
        for (int l = 0; l < no; ++l) acc2 += R.lam_pT[(int64_t)l * no + k] * L.t[l];  // Lambda_p t
      } else {
        const int j = e - nu;
        for (int i = 0; i < nu; ++i) acc2 += X.Gb[i * nu + j] * Qw[(int64_t)i * no + k];
        for (int l = 0; l < no; ++l) acc2 += R.lam_p[(int64_t)l * no + k] * L.t[l];  // Lambda_p^T t
      }
      pb[k] = acc2;
    }
    wave_sync();
  }

  // ---------------------------------------------------------------- pass 3
  for (int e = 0; e < ne; ++e) {
    const bool up = e < nu;
    const double ex = L.rr[3 * e], ey = L.rr[3 * e + 1], ez = L.rr[3 * e + 2];
    electron_q_h(s, L.Gi, Qw, e, L.q, L.h, lane);
    const double* ge = gw ? gw + (int64_t)e * na : nullptr;
    const double* gbe = X.gbar ? X.gbar + (int64_t)e * na : nullptr;
    const double gJ[3] = {L.el[EL * e], L.el[EL * e + 1], L.el[EL * e + 2]};
    const double dJ[3] = {L.el[EL * e + 3], L.el[EL * e + 4], L.el[EL * e + 5]};
    const double S = L.el[EL * e + 9];
    const double* C = up ? R.mo_up : R.mo_dn;
    const double* pb = X.phibar + (int64_t)e * no;
    double re[3] = {0.0, 0.0, 0.0};
    for (int a = lane; a < na; a += WAVE) {
      double a0 = 0.0;
      if (s.orb_is_mo) {
        for (int k = 0; k < no; ++k) a0 += C[(int64_t)k * na + a] * pb[k];
      } else {
        a0 = pb[a];
      }
      double gj = 0.0;
      if (ge) {
        gj = ge[a];
        // J3 value adjoint: J^T sum_{f<e} gbar_f + J sum_{f>e} gbar_f - S_e g_e
        double jb = -S * gj;
        for (int b = 0; b < na; ++b) {
          const double lt = L.plt[b], gt = L.gtot[b] - lt - gbe[b];
          jb += R.j3[(int64_t)b * na + a] * lt + R.j3T[(int64_t)b * na + a] * gt;
        }
        a0 += jb;
      }
      const double hv = L.h[a];
      const double a1[3] = {-gJ[0] * hv + dJ[0] * gj, -gJ[1] * hv + dJ[1] * gj, -gJ[2] * hv + dJ[2] * gj};
      const double a2 = -0.5 * (hv + gj);
      double v, g[3], H[6], gl[3];
      ao_d3(s, a, ex, ey, ez, v, g, H, gl);
      // H rows: x = (xx, xy, xz), y = (xy, yy, yz), z = (xz, yz, zz)
      const double vx = a0 * g[0] + a1[0] * H[0] + a1[1] * H[3] + a1[2] * H[4] + a2 * gl[0];
      const double vy = a0 * g[1] + a1[0] * H[3] + a1[1] * H[1] + a1[2] * H[5] + a2 * gl[1];
      const double vz = a0 * g[2] + a1[0] * H[4] + a1[1] * H[5] + a1[2] * H[2] + a2 * gl[2];
      re[0] += vx;
      re[1] += vy;
      re[2] += vz;
      X.RA[3 * a] -= vx;
      X.RA[3 * a + 1] -= vy;
      X.RA[3 * a + 2] -= vz;
    }
    wave_sum_k<3>(re);
    wave_sync();
    if (gbe) {
      for (int a = lane; a < na; a += WAVE) L.plt[a] += gbe[a];
    }
    if (lane == 0) {
      L.rbar[3 * e] += re[0];
      L.rbar[3 * e + 1] += re[1];
      L.rbar[3 * e + 2] += re[2];
    }
    wave_sync();
  }

  // ---------------------------------------------------------------- pass 4
  for (int n = lane; n < nn; n += WAVE) {
    double acc3[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < na; ++a) {
      if (s.ao_nuc[a] != n) continue;
      for (int c = 0; c < 3; ++c) {
        acc3[c] += X.RA[3 * a + c];
        acc3[3 + c] += X.RL[3 * a + c];
      }
    }
    for (int m = 0; m < nn; ++m) {  // ion-ion repulsion
      if (m == n) continue;
      const double ux = s.nuc_pos[3 * n] - s.nuc_pos[3 * m], uy = s.nuc_pos[3 * n + 1] - s.nuc_pos[3 * m + 1],
                   uz = s.nuc_pos[3 * n + 2] - s.nuc_pos[3 * m + 2];
      const double d2 = ux * ux + uy * uy + uz * uz;
      const double c = -s.nuc_charge[n] * s.nuc_charge[m] / (d2 * sqrt(d2));
      acc3[0] += c * ux;
      acc3[1] += c * uy;
      acc3[2] += c * uz;
    }
    for (int c = 0; c < 3; ++c) {
      L.Rbar[3 * n + c] += acc3[c];
      L.lnR[3 * n + c] -= acc3[3 + c];
    }
  }
  // potential energy for e_L (as in k_local_energy)
  double pot = 0.0;
  for (int t = lane; t < ne * nn; t += WAVE) {
    const int e = t / nn, n = t - e * nn;
    const double dx = L.rr[3 * e] - s.nuc_pos[3 * n], dy = L.rr[3 * e + 1] - s.nuc_pos[3 * n + 1],
                 dz = L.rr[3 * e + 2] - s.nuc_pos[3 * n + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    pot += -s.nuc_charge[n] / d;
    if (s.ecp_on) pot += ecp_local_pair(s, n, d);
  }
  pot += el_el_partial(s, L.rr, lane);
  pot = wave_sum(pot);  // kinetic and vnl are already the same on every lane
  wave_sync();

  if (lane == 0 && e_L) e_L[w] = kinetic + (pot + s.ion_ion) + vnl;
  for (int k = lane; k < 3 * ne; k += WAVE) {
    const int e = k / 3, c = k - 3 * e;
    if (dE_dr) dE_dr[w * 3 * ne + k] = L.rbar[k];
    if (dlnpsi_dr) dlnpsi_dr[w * 3 * ne + k] = L.el[EL * e + 6 + c] + L.el[EL * e + c];
  }
  for (int k = lane; k < 3 * nn; k += WAVE) {
    if (dE_dR) dE_dR[w * 3 * nn + k] = L.Rbar[k];
    if (dlnpsi_dR) dlnpsi_dR[w * 3 * nn + k] = L.lnR[k];
  }
  if (force_HF || force_PP) {
    // space-warp weights omega_ne = |r_e - R_n|^-4 / sum_m |r_e - R_m|^-4 (swct.py:63-150)
    for (int n = lane; n < nn; n += WAVE) {
      double hf[3] = {L.Rbar[3 * n], L.Rbar[3 * n + 1], L.Rbar[3 * n + 2]};
      double pp[3] = {L.lnR[3 * n], L.lnR[3 * n + 1], L.lnR[3 * n + 2]};
      if (swct) {
        for (int e = 0; e < ne; ++e) {
          double Ssum = 0.0, dS[3] = {0.0, 0.0, 0.0}, kn = 0.0, dk[3] = {0.0, 0.0, 0.0};
          for (int m = 0; m < nn; ++m) {
            const double ux = L.rr[3 * e] - s.nuc_pos[3 * m], uy = L.rr[3 * e + 1] - s.nuc_pos[3 * m + 1],
                         uz = L.rr[3 * e + 2] - s.nuc_pos[3 * m + 2];
            const double d2 = ux * ux + uy * uy + uz * uz;
            const double km = 1.0 / (d2 * d2);
            const double c = -4.0 * km / d2;
            Ssum += km;
            dS[0] += c * ux;
            dS[1] += c * uy;
            dS[2] += c * uz;
            if (m == n) {
              kn = km;
              dk[0] = c * ux;
              dk[1] = c * uy;
              dk[2] = c * uz;
            }
          }
          const double om = kn / Ssum;
          for (int c = 0; c < 3; ++c) {
            hf[c] += om * L.rbar[3 * e + c];
            pp[c] += om * (L.el[EL * e + 6 + c] + L.el[EL * e + c]) +
                     0.5 * (dk[c] / Ssum - kn * dS[c] / (Ssum * Ssum));
          }
        }
      }
      for (int c = 0; c < 3; ++c) {
        if (force_HF) force_HF[w * 3 * nn + 3 * n + c] = hf[c];
        if (force_PP) force_PP[w * 3 * nn + 3 * n + c] = pp[c];
      }
    }
  }
}

}  // namespace jqmc
