// gfx950 kernels of the per-walker log-derivatives O_k = d ln|Psi| / d p_k (the SR inputs of
// jqmc_mcmc.py:854-873) for the variational blocks J1 b, J2 b, J3 matrix and Lambda:
//
//   k_ginv_phi       Z = G^-1 Phi_up per walker (n_up x n_orb, FP64)
//   k_rank_outer     out[w][m][n] = sum_k L[w][k][m] R[w][k][n]: the J3 block
//                    sum_{e<f} a_e a_f^T = sum_e a_e (sum_{f>e} a_f)^T and the Lambda_p block
//                    Z[:n_dn]^T Phi_dn, 64 x 64 output tiles per block staged through LDS
//   k_dlnpsi_small   J1 / J2 parameter derivatives, the J3 one-body column sum_e a_e and the
//                    Lambda_u block Z[n_dn:]^T, one wavefront per walker
#pragma once

#include "jqmc_lrdmc.hip"

namespace jqmc {

// Z[w] = G^-1[w] Phi_up[w]: Z[j][k] = sum_i G^-1[j][i] phi_k(r_i) (thread per element).
template <typename T>
__global__ void k_ginv_phi(Sys s, int64_t W, WalkT<T> wk, double* __restrict__ Z) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = s.n_up, no = s.n_orb;
  const int64_t per_w = (int64_t)nu * no;
  if (idx >= W * per_w) return;
  const int64_t w = idx / per_w;
  const int rem = (int)(idx - w * per_w);
  const int j = rem / no, k = rem - j * no;
  const double* Gi = wk.Ginv + w * nu * nu + (int64_t)j * nu;
  const double* P = wk.orb + w * s.n_el * no + k;
  double acc = 0.0;
  for (int i = 0; i < nu; ++i) acc += Gi[i] * P[(int64_t)i * no];
  Z[idx] = acc;
}

constexpr int RO_TILE = 64;
constexpr int RO_KC = 16;

// out[w * out_ld + out_off + m * out_rs + n] = sum_{k < K} L[w * l_ld + k * l_rs + m] R[w * r_ld + k * r_rs + n]
// for m < M, n < N.  Grid (ceil(N/64), ceil(M/64), W); 256 threads each own a 4 x 4 sub-tile
// (rows ty + 16 i, columns tx + 16 j) so that neighbouring threads store neighbouring columns.
template <typename TL, typename TR>
__global__ void __launch_bounds__(256) k_rank_outer(int K, int M, int N, const TL* __restrict__ L, int64_t l_ld,
                                                    int64_t l_rs, const TR* __restrict__ R, int64_t r_ld, int64_t r_rs,
                                                    double* __restrict__ out, int64_t out_ld, int64_t out_rs,
                                                    int64_t out_off) {
  __shared__ double sl[RO_KC][RO_TILE];
  __shared__ double sr[RO_KC][RO_TILE];
  const int64_t w = blockIdx.z;
  const int n0 = blockIdx.x * RO_TILE, m0 = blockIdx.y * RO_TILE;
  const int t = threadIdx.x, tx = t % 16, ty = t / 16;
  const TL* Lw = L + w * l_ld;
  const TR* Rw = R + w * r_ld;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += RO_KC) {
    for (int q = t; q < RO_KC * RO_TILE; q += 256) {
      const int kk = q / RO_TILE, c = q - kk * RO_TILE;
      const int k = k0 + kk;
      sl[kk][c] = (k < K && m0 + c < M) ? (double)Lw[(int64_t)k * l_rs + m0 + c] : 0.0;
      sr[kk][c] = (k < K && n0 + c < N) ? (double)Rw[(int64_t)k * r_rs + n0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < RO_KC; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sl[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sr[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
  double* ow = out + w * out_ld + out_off;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty + 16 * i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx + 16 * j;
      if (n < N) ow[(int64_t)m * out_rs + n] = acc[i][j];
    }
  }
}

// Offsets (in doubles, per walker) of the O_k blocks; -1 when a block is absent.
struct ParamLayout {
  int64_t j1, j2, j3, lam, total;
};

// d/db of the one- and two-body Jastrow sums (jastrow_factor.py:672-711, 1207-1230), the J3
// one-body column d/dj[a] = sum_e a_e[a] and the unpaired Lambda block
// d/dLambda_u[k][u] = Z[n_dn + u][k] (determinant.py:1379-1402); one wavefront per walker.
template <typename T>
__global__ void __launch_bounds__(256) k_dlnpsi_small(Sys s, int64_t W, WalkT<T> wk, const double* __restrict__ Z,
                                                      ParamLayout lay, double* __restrict__ out, int64_t out_ld,
                                                      int wpb) {
  const int wave = wave_id(), lane = lane_id();
  const int64_t w = (int64_t)blockIdx.x * wpb + wave;
  if (w >= W) return;
  const int ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up, nd = s.n_dn;
  double* ow = out + w * out_ld;
  double acc[2] = {0.0, 0.0};
  if (lay.j1 >= 0) {
    const double b = s.j1_param;
    for (int t = lane; t < ne * s.n_nuc; t += WAVE) {
      const int e = t / s.n_nuc, n = t - e * s.n_nuc;
      const double* r = elec_pos(s, wk, w, e);
      const double dx = r[0] - s.nuc_pos[3 * n], dy = r[1] - s.nuc_pos[3 * n + 1], dz = r[2] - s.nuc_pos[3 * n + 2];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double c = s.j1_c[n], A = s.j1_A[n], cd = c * d;
      if (s.j1_type == 1) {
        const double ex = exp_nonpos(-b * cd);
        acc[0] += -A * (-(1.0 - ex) / (2.0 * b * b) + cd * ex / (2.0 * b));
      } else {
        const double den = 1.0 + b * cd;
        acc[0] += A * d * cd / (2.0 * den * den);
      }
    }
  }
  if (lay.j2 >= 0) {
    const double b = s.j2_param;
    for (int e = lane; e < ne; e += WAVE) {
      const double* r = elec_pos(s, wk, w, e);
      for (int f = e + 1; f < ne; ++f) {
        const double* q = elec_pos(s, wk, w, f);
        const double dx = r[0] - q[0], dy = r[1] - q[1], dz = r[2] - q[2];
        const double x = sqrt(dx * dx + dy * dy + dz * dz);
        if (s.j2_type == 2) {
          const double den = 1.0 + b * x;
          acc[1] += -x * x / (2.0 * den * den);
        } else {
          const double ex = exp_nonpos(-b * x);
          acc[1] += -(1.0 - ex) / (2.0 * b * b) + x * ex / (2.0 * b);
        }
      }
    }
  }
  wave_sum_k<2>(acc);
  if (lane == 0) {
    if (lay.j1 >= 0) ow[lay.j1] = acc[0];
    if (lay.j2 >= 0) ow[lay.j2] = acc[1];
  }
  if (lay.j3 >= 0) {
    const T* A = wk.ao + w * ne * na;
    for (int a = lane; a < na; a += WAVE) {
      double sa = 0.0;
      for (int e = 0; e < ne; ++e) sa += (double)A[(int64_t)e * na + a];
      ow[lay.j3 + (int64_t)a * (na + 1) + na] = sa;
    }
  }
  if (lay.lam >= 0 && nu > nd) {
    const int unp = nu - nd;
    const double* Zw = Z + w * nu * no;
    for (int t = lane; t < no * unp; t += WAVE) {
      const int k = t / unp, u = t - k * unp;
      ow[lay.lam + (int64_t)k * (no + unp) + no + u] = Zw[(int64_t)(nd + u) * no + k];
    }
  }
}

}  // namespace jqmc
