// Stochastic-reconfiguration (SR) kernels: the consumer of the per-walker log-derivatives O_k
// (run_optimize, jqmc_mcmc.py:2952-3350).  With samples i (S of them, weights w_i) and
// parameters k (P of them):
//
//   O_bar_k = sum_i w_i O_ik / sum w,   X_ki = (O_ik - O_bar_k) sqrt(w_i / sum w)     (2996-3003)
//   diag_S_k = sum_i X_ki^2, floored at min_S_diag_abs                               (3062-3112)
//   X <- X / sqrt(diag_S)
//   P <  S: theta = (X X^T + eps I)^-1 X F          (a P x P Gram, 3114-3196)
//   P >= S: theta = X (X^T X + eps I)^-1 F          (push-through, an S x S Gram, 3198-3350)
//   theta <- theta / sqrt(diag_S)
//
// Both Grams are "row Grams" G = Y Y^T of one matrix Y whose rows are the samples (push-through)
// or the parameters (P < S, read from O^T): Y_rc = (M_rc - a_r - b_c) u_r v_c, with the shift and
// scale vectors given per row and per column.  k_sr_gram forms G on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), 64 x 64 output tiles of the upper triangle, 16-column slices of Y
// staged through LDS with the next slice prefetched into registers, the transform applied on the
// way into LDS, and each tile written to both triangles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace jqmc {

typedef double sr_double4 __attribute__((ext_vector_type(4)));

constexpr int SRG_TILE = 64;  // output rows/columns per block
constexpr int SRG_KC = 16;    // columns of Y per LDS slice
constexpr int SRG_PAD = 1;    // LDS row padding (doubles): rows 136 bytes apart, no bank conflicts

// Column statistics of a row-major (S x P) matrix with leading dimension ld: mean_k =
// sum_i wn_i O_ik and var_k = sum_i wn_i (O_ik - mean_k)^2 for normalised weights wn (sum 1), in
// two passes as the reference (O_bar first, then the centred squares).  One thread per column.
__global__ void k_sr_colstats(const double* __restrict__ O, int64_t S, int64_t P, int64_t ld,
                              const double* __restrict__ wn, double* __restrict__ mean, double* __restrict__ var) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  double m = 0.0;
  for (int64_t i = 0; i < S; ++i) m += wn[i] * O[i * ld + k];
  double v = 0.0;
  for (int64_t i = 0; i < S; ++i) {
    const double d = O[i * ld + k] - m;
    v += wn[i] * (d * d);
  }
  mean[k] = m;
  var[k] = v;
}

// out_k = cs_k * sum_i (O_ik - mean_k) rs_i x_i for the row-major (S x P) O: theta = X x (and
// X F) with X_ki = (O_ik - mean_k) rs_i cs_k.  One thread per column, rows streamed.
__global__ void k_sr_project(const double* __restrict__ O, int64_t S, int64_t P, int64_t ld,
                             const double* __restrict__ mean, const double* __restrict__ rs,
                             const double* __restrict__ cs, const double* __restrict__ x, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  const double mk = mean[k];
  double acc = 0.0;
  for (int64_t i = 0; i < S; ++i) acc += (O[i * ld + k] - mk) * (rs[i] * x[i]);
  out[k] = cs[k] * acc;
}

// G (n x n, leading dimension ldg) = Y Y^T with Y_rc = (M_rc - a_r - b_c) u_r v_c for the row-major
// (n x m) M (leading dimension ldm); a, b may be NULL (zero shift).  Grid: one block per 64 x 64
// tile of the upper triangle, blockIdx.x = tile index in row-major order of the triangle.  256
// threads = 4 waves, wave (wr, wc) owns the 32 x 32 quadrant (wr, wc): 2 x 2 MFMA tiles of 16 x 16.
// MFMA operand layout (v_mfma_f64_16x16x4f64): A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k,
// D(i, j) in lane j + 16 (i % 4), element i / 4.
__global__ void __launch_bounds__(256) k_sr_gram(const double* __restrict__ M, int64_t n, int64_t m, int64_t ldm,
                                                 const double* __restrict__ a, const double* __restrict__ b,
                                                 const double* __restrict__ u, const double* __restrict__ v,
                                                 double* __restrict__ G, int64_t ldg, int64_t n_tiles) {
  __shared__ double ys_r[SRG_TILE][SRG_KC + SRG_PAD];  // rows of the tile's row block
  __shared__ double ys_c[SRG_TILE][SRG_KC + SRG_PAD];  // rows of the tile's column block
  // tile (ti, tj), ti <= tj, from the triangle index t = ti * n_tiles - ti (ti - 1) / 2 + (tj - ti)
  int64_t t = blockIdx.x, ti = 0;
  while (t >= n_tiles - ti) {
    t -= n_tiles - ti;
    ++ti;
  }
  const int64_t tj = ti + t;
  const int64_t r0 = ti * SRG_TILE, c0 = tj * SRG_TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  // staging: thread tid loads row (tid >> 2) of each block, columns 4 (tid & 3) .. +3 of the slice
  const int sr = tid >> 2, sk = 4 * (tid & 3);
  const int64_t gr = r0 + sr, gc = c0 + sr;
  const bool okr = gr < n, okc = gc < n;
  const double ar = (okr && a) ? a[gr] : 0.0, ac = (okc && a) ? a[gc] : 0.0;
  const double ur = okr ? u[gr] : 0.0, uc = okc ? u[gc] : 0.0;
  double pr[4], pc[4];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t k = k0 + sk + q;
      const bool okk = k < m;
      const double bk = (okk && b) ? b[k] : 0.0, vk = okk ? v[k] : 0.0;
      pr[q] = (okr && okk) ? (M[gr * ldm + k] - ar - bk) * ur * vk : 0.0;
      pc[q] = (okc && okk) ? (M[gc * ldm + k] - ac - bk) * uc * vk : 0.0;
    }
  };
  sr_double4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = sr_double4{0.0, 0.0, 0.0, 0.0};
  fetch(0);
  for (int64_t k0 = 0; k0 < m; k0 += SRG_KC) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ys_r[sr][sk + q] = pr[q];
      ys_c[sr][sk + q] = pc[q];
    }
    __syncthreads();
    if (k0 + SRG_KC < m) fetch(k0 + SRG_KC);  // next slice in flight during this one
#pragma unroll
    for (int kk = 0; kk < SRG_KC; kk += 4) {
      const int kl = kk + (lane >> 4), il = lane & 15;
      double fa[2], fb[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) fa[x] = ys_r[32 * wr + 16 * x + il][kl];
#pragma unroll
      for (int y = 0; y < 2; ++y) fb[y] = ys_c[32 * wc + 16 * y + il][kl];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: element e of lane l in tile (x, y) is G(r0 + 32 wr + 16 x + 4 e + l / 16,
  // c0 + 32 wc + 16 y + l % 16); off-diagonal tiles also write the mirrored element, a diagonal
  // tile holds both triangles itself
  const bool mirror = ti != tj;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t i = r0 + 32 * wr + 16 * x + 4 * e + (lane >> 4);
        const int64_t j = c0 + 32 * wc + 16 * y + (lane & 15);
        if (i < n && j < n) {
          G[i * ldg + j] = acc[x][y][e];
          if (mirror) G[j * ldg + i] = acc[x][y][e];
        }
      }
}

// The same Gram on the FP32 matrix cores (v_mfma_f32_16x16x4f32): the transform Y_rc = (M_rc - a_r -
// b_c) u_r v_c is formed in FP64 and rounded once to FP32, the products are accumulated in FP32, and
// G is written in FP64 (the caller adds its diagonal shift there).  128 x 128 output tiles of the
// upper triangle, 256 threads = 4 waves, wave (wr, wc) owns the 64 x 64 quadrant (wr, wc): 4 x 4
// MFMA tiles.  A 16-column slice of Y is staged per step; the next slice's raw FP64 words are
// loaded into registers while the matrix cores work on the current one.
//
// Within a slice the sum over k is taken in the order k = 4 q + s, s = MFMA step, q = lane / 16, so
// that each lane reads its four k of one row with a single 16-byte LDS load.  MFMA operand layout:
// A(i, q) in lane i + 16 q, B(q, j) in lane j + 16 q, D(i, j) in lane j + 16 (i / 4), element i % 4.
constexpr int SRF_TILE = 128;
constexpr int SRF_KC = 16;
constexpr int SRF_LD = SRF_KC + 4;  // floats per LDS row: 16-byte aligned, rows 80 bytes apart

typedef float sr_float4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_sr_gram_f32(const double* __restrict__ M, int64_t n, int64_t m,
                                                     int64_t ldm, const double* __restrict__ a,
                                                     const double* __restrict__ b, const double* __restrict__ u,
                                                     const double* __restrict__ v, double* __restrict__ G,
                                                     int64_t ldg, int64_t n_tiles) {
  __shared__ __attribute__((aligned(16))) float ys_r[SRF_TILE * SRF_LD];
  __shared__ __attribute__((aligned(16))) float ys_c[SRF_TILE * SRF_LD];
  int64_t t = blockIdx.x, ti = 0;
  while (t >= n_tiles - ti) {
    t -= n_tiles - ti;
    ++ti;
  }
  const int64_t tj = ti + t;
  const int64_t r0 = ti * SRF_TILE, c0 = tj * SRF_TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  // staging: thread tid loads row (tid >> 1) of each block, columns 8 (tid & 1) .. +7 of the slice
  const int sr = tid >> 1, sk = 8 * (tid & 1);
  const int64_t gr = r0 + sr, gc = c0 + sr;
  const bool okr = gr < n, okc = gc < n;
  const double ar = (okr && a) ? a[gr] : 0.0, ac = (okc && a) ? a[gc] : 0.0;
  const double ur = okr ? u[gr] : 0.0, uc = okc ? u[gc] : 0.0;
  double pr[8], pc[8], pb[8], pv[8];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t k = k0 + sk + q;
      const bool okk = k < m;
      pb[q] = (okk && b) ? b[k] : 0.0;
      pv[q] = okk ? v[k] : 0.0;
      pr[q] = (okr && okk) ? M[gr * ldm + k] : 0.0;
      pc[q] = (okc && okk) ? M[gc * ldm + k] : 0.0;
    }
  };
  sr_float4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = sr_float4{0.f, 0.f, 0.f, 0.f};
  fetch(0);
  const int il = lane & 15, q4 = 4 * (lane >> 4);
  for (int64_t k0 = 0; k0 < m; k0 += SRF_KC) {
    sr_float4 wr0, wr1, wc0, wc1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      wr0[q] = (float)((pr[q] - ar - pb[q]) * ur * pv[q]);
      wr1[q] = (float)((pr[q + 4] - ar - pb[q + 4]) * ur * pv[q + 4]);
      wc0[q] = (float)((pc[q] - ac - pb[q]) * uc * pv[q]);
      wc1[q] = (float)((pc[q + 4] - ac - pb[q + 4]) * uc * pv[q + 4]);
    }
    *(sr_float4*)&ys_r[sr * SRF_LD + sk] = wr0;
    *(sr_float4*)&ys_r[sr * SRF_LD + sk + 4] = wr1;
    *(sr_float4*)&ys_c[sr * SRF_LD + sk] = wc0;
    *(sr_float4*)&ys_c[sr * SRF_LD + sk + 4] = wc1;
    __syncthreads();
    if (k0 + SRF_KC < m) fetch(k0 + SRF_KC);  // next slice in flight during this one
    sr_float4 fa[4], fb[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) fa[x] = *(const sr_float4*)&ys_r[(64 * wr + 16 * x + il) * SRF_LD + q4];
#pragma unroll
    for (int y = 0; y < 4; ++y) fb[y] = *(const sr_float4*)&ys_c[(64 * wc + 16 * y + il) * SRF_LD + q4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[x][s], fb[y][s], acc[x][y], 0, 0, 0);
    __syncthreads();
  }
  // epilogue: element e of lane l in tile (x, y) is G(r0 + 64 wr + 16 x + 4 (l / 16) + e,
  // c0 + 64 wc + 16 y + l % 16)
  const bool mirror = ti != tj;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t i = r0 + 64 * wr + 16 * x + q4 + e;
        const int64_t j = c0 + 64 * wc + 16 * y + il;
        if (i < n && j < n) {
          const double g = (double)acc[x][y][e];
          G[i * ldg + j] = g;
          if (mirror) G[j * ldg + i] = g;
        }
      }
}

}  // namespace jqmc
