// C ABI of libjqmc_hip.so (declarations and contracts in include/jqmc_hip.h).
// One translation unit: the kernels are included here so the library is a
// single gfx950 code object plus host stubs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/jqmc_hip.h"
#include "jqmc_el.hip"
#include "jqmc_forces.hip"
#include "jqmc_sr.hip"
#include "jqmc_sweep_batched.hip"

using namespace jqmc;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) return fail(JQMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

constexpr int64_t kMaxLdsPerBlock = 64 * 1024;

}  // namespace

struct jqmc_system {
  Sys s;                 // host copy (kernels taking the pack by value)
  const Sys* d_sys;      // device copy (kernels reading the pack through a pointer)
  std::vector<void*> allocations;
  int device;
  // launch shapes chosen by jqmc_system_set_tuning (0 = the library's default); none of them
  // changes a result bit
  int tune_el_wpb = 0;
  int tune_el_team = 0;
  int tune_sweep_wpb = 0;
  // FP64 guard of the mixed form's local energy (jqmc_system_set_mixed_guard)
  double guard_e_ref = 0.0;
  double guard_width = 1.0e4;
  // pinned host words the GFMC_t projections read their active-walker counts into
  int* host_flags = nullptr;
};

namespace {

template <typename T>
int upload(jqmc_system* sys, const T* host, size_t count, const T** dev_out) {
  *dev_out = nullptr;
  if (count == 0) return JQMC_OK;
  if (host == nullptr) return fail(JQMC_ERR_INVALID_ARGUMENT, "null host array for a non-empty table");
  void* d = nullptr;
  HIP_TRY(hipMalloc(&d, count * sizeof(T)));
  sys->allocations.push_back(d);
  HIP_TRY(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
  *dev_out = static_cast<const T*>(d);
  return JQMC_OK;
}

// FP32 copy of an FP64 table.
int upload_f32(jqmc_system* sys, const double* host, size_t count, const float** dev_out) {
  std::vector<float> f(count);
  for (size_t i = 0; i < count; ++i) f[i] = (float)host[i];
  return upload(sys, f.data(), count, dev_out);
}

#define UPLOAD(host, count, dst)                                  \
  do {                                                            \
    int _rc = upload(sys, (host), (size_t)(count), &(dst));       \
    if (_rc != JQMC_OK) {                                         \
      jqmc_system_destroy(sys);                                   \
      return _rc;                                                 \
    }                                                             \
  } while (0)

// FP64 table into s.td.<field> and, in mixed precision, its FP32 copy into s.tf.<field>.
#define UPLOAD_REAL(host, count, field)                                                          \
  do {                                                                                           \
    UPLOAD((host), (count), s.td.field);                                                         \
    if (s.precision == JQMC_PRECISION_MIXED_FP32) {                                              \
      int _rc = upload_f32(sys, (host), (size_t)(count), &s.tf.field);                           \
      if (_rc != JQMC_OK) {                                                                      \
        jqmc_system_destroy(sys);                                                                \
        return _rc;                                                                              \
      }                                                                                          \
    }                                                                                            \
  } while (0)

int walkers_per_block(int64_t lds_bytes_per_wave) {
  int wpb = 4;
  while (wpb > 1 && wpb * lds_bytes_per_wave > kMaxLdsPerBlock) wpb /= 2;
  return wpb;
}

// Walkers per block: the most that fit the LDS budget, or fewer when tuned so.
int tuned_wpb(int64_t per_wave, int tune) {
  const int wpb = walkers_per_block(per_wave);
  return tune > 0 && tune < wpb ? tune : wpb;
}

int check_walkers(const Sys& s, const jqmc_walkers* wk, bool need_state) {
  if (wk == nullptr) return fail(JQMC_ERR_INVALID_ARGUMENT, "walkers is NULL");
  if ((s.n_up > 0 && !wk->r_up) || (s.n_dn > 0 && !wk->r_dn))
    return fail(JQMC_ERR_INVALID_ARGUMENT, "electron position arrays are NULL");
  if (need_state) {
    if (!wk->ao || !wk->orb || !wk->pair || !wk->G || !wk->Ginv)
      return fail(JQMC_ERR_INVALID_ARGUMENT, "walker state arrays (ao, orb, pair, G, Ginv) are required");
    if (s.j3_on && !wk->jfield) return fail(JQMC_ERR_INVALID_ARGUMENT, "jfield is required when J3 is on");
  }
  return JQMC_OK;
}

template <typename T>
WalkT<T> to_walk(const jqmc_walkers* w) {
  WalkT<T> k;
  k.r_up = w->r_up;
  k.r_dn = w->r_dn;
  k.ao = static_cast<T*>(w->ao);
  k.orb = static_cast<double*>(w->orb);
  k.pair = static_cast<double*>(w->pair);
  k.jfield = static_cast<T*>(w->jfield);
  k.G = w->G;
  k.Ginv = w->Ginv;
  k.logabsdet = w->logabsdet;
  k.sign = w->sign;
  k.jastrow = w->jastrow;
  k.as_factor = w->as_factor;
  k.accepted = w->accepted;
  k.rejected = w->rejected;
  return k;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

inline int64_t cache_bytes(const Sys& s) {
  return s.precision == JQMC_PRECISION_MIXED_FP32 ? (int64_t)sizeof(float) : (int64_t)sizeof(double);
}

// ---------------------------------------------------------------- launchers over T

template <typename T>
int launch_init(const Sys& s, int64_t W, const jqmc_walkers* wk, void* workspace, hipStream_t st) {
  WalkT<T> k = to_walk<T>(wk);
  const int B = 256;
  const int64_t rows_lds = init_rows_lds_bytes(s);
  k_init_rows<T><<<grid_for(W * s.n_el, 4), 4 * WAVE, 4 * rows_lds, st>>>(s, W, k, rows_lds);
  k_init_geminal<T><<<grid_for(W * s.n_up * s.n_up, B), B, 0, st>>>(s, W, k);
  {
    const int64_t per_wave = 2 * (int64_t)s.n_up * s.n_up * 8;
    const int wpb = walkers_per_block(per_wave);
    k_init_inverse<<<grid_for(W, wpb), wpb * WAVE, wpb * per_wave, st>>>(k.G, k.Ginv, k.logabsdet, k.sign,
                                                                         k.as_factor, s.n_up, W, wpb);
  }
  if (s.j3_on) {
    T* S1 = static_cast<T*>(workspace);
    T* S2 = S1 + W * s.n_el * s.n_ao;
    k_init_jsums<T><<<grid_for(W * s.n_ao, B), B, 0, st>>>(s, W, k, S1, S2);
    k_init_jfield<T><<<grid_for(W * s.n_el * s.n_ao, B), B, 0, st>>>(s, W, k, S1, S2);
  }
  k_init_jastrow<T><<<grid_for(W, 4), 4 * WAVE, 0, st>>>(s, W, k, 4);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

template <typename T>
int launch_refresh(const Sys& s, int64_t W, const jqmc_walkers* wk, hipStream_t st, bool rows) {
  WalkT<T> k = to_walk<T>(wk);
  if (rows) {
    const int64_t rows_lds = init_rows_lds_bytes(s);
    k_init_rows<T><<<grid_for(W * s.n_el, 4), 4 * WAVE, 4 * rows_lds, st>>>(s, W, k, rows_lds);
  }
  k_init_geminal<T><<<grid_for(W * s.n_up * s.n_up, 256), 256, 0, st>>>(s, W, k);
  const int64_t per_wave = 2 * (int64_t)s.n_up * s.n_up * 8;
  const int wpb = walkers_per_block(per_wave);
  k_init_inverse<<<grid_for(W, wpb), wpb * WAVE, wpb * per_wave, st>>>(k.G, k.Ginv, k.logabsdet, k.sign, k.as_factor,
                                                                       s.n_up, W, wpb);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

template <typename T>
int launch_sweep(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                 double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, hipStream_t st) {
  const Sys& s = sys->s;
  const int64_t per_wave = sweep_lds_bytes<T>(s);
  if (per_wave > kMaxLdsPerBlock) return fail(JQMC_ERR_UNSUPPORTED, "system too large for the sweep LDS budget");
  const int wpb = walkers_per_block(per_wave);
  const size_t lds = (size_t)wpb * per_wave;
  const WalkT<T> k = to_walk<T>(wk);
  switch (wpb) {
    case 4:
      k_vmc_sweep<4, T><<<grid_for(W, 4), 4 * WAVE, lds, st>>>(sys->s, W, k, n_moves, Dt, epsilon_AS, seed, move0,
                                                               move_ctr, per_wave);
      break;
    case 2:
      k_vmc_sweep<2, T><<<grid_for(W, 2), 2 * WAVE, lds, st>>>(sys->s, W, k, n_moves, Dt, epsilon_AS, seed, move0,
                                                               move_ctr, per_wave);
      break;
    default:
      k_vmc_sweep<1, T><<<grid_for(W, 1), WAVE, lds, st>>>(sys->s, W, k, n_moves, Dt, epsilon_AS, seed, move0,
                                                           move_ctr, per_wave);
  }
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

template <typename T>
int launch_sweep_batched(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                         double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, void* workspace,
                         hipStream_t st) {
  const Sys& s = sys->s;
  if (grid_for(W, JF_WALKERS) > 65535u) return fail(JQMC_ERR_SHAPE, "too many walkers for one batched sweep");
  const int64_t per_wave = prop_lds_bytes<T>(s);
  if (per_wave > kMaxLdsPerBlock) return fail(JQMC_ERR_UNSUPPORTED, "system too large for the sweep LDS budget");
  const int wpb = tuned_wpb(per_wave, sys->tune_sweep_wpb);
  const size_t lds = (size_t)wpb * per_wave;
  const WalkT<T> k = to_walk<T>(wk);
  SweepWs<T> ws(workspace, s, W);
  k_sweep_begin<T><<<grid_for(W, wpb), wpb * WAVE, lds, st>>>(sys->s, W, k, ws, wpb, per_wave);
  for (int m = 0; m < n_moves; ++m) {
    k_sweep_propose<T><<<grid_for(W, wpb), wpb * WAVE, lds, st>>>(sys->s, W, k, ws, m, Dt, epsilon_AS, seed,
                                                                  move0, move_ctr, wpb, per_wave);
    if (s.j3_on) {
      const dim3 jg(grid_for(s.n_ao, JF_ROWS), grid_for(W, JF_WALKERS));
      k_sweep_jfield_mfma<T><<<jg, 256, 0, st>>>(sys->s, W, k, ws);
    }
  }
  const int64_t n_flush = std::max<int64_t>(s.j3_on ? W * s.n_el * s.n_ao : 0, W);
  k_sweep_flush<T><<<grid_for(n_flush, 256), 256, 0, st>>>(W, k, ws, s.n_el, s.n_ao, s.n_up);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

// Layout of the local-energy workspace: the h rows (and q rows) of every walker, then for the mixed
// form's FP64 guard the flagged count, the flagged list and GUARD_SLOTS scratch records.
struct ElWs {
  double* hws;
  int* count;
  int* list;
  double* scratch;
  int64_t rec;  // doubles per guard record
};

inline int64_t el_ws_bytes(const Sys& s, int64_t W) {
  const int64_t rows = s.n_ao + (el2_q_in_lds(s) ? 0 : s.n_orb);
  int64_t b = W * (int64_t)s.n_el * rows * (int64_t)sizeof(double);
  if (s.precision == JQMC_PRECISION_MIXED_FP32) {
    b += 16 + ((W * (int64_t)sizeof(int) + 15) & ~(int64_t)15);
    b += GUARD_SLOTS * GuardRec::doubles(s) * (int64_t)sizeof(double);
  }
  return b;
}

inline ElWs el_ws(const Sys& s, int64_t W, void* base) {
  ElWs x{};
  char* p = static_cast<char*>(base);
  x.hws = reinterpret_cast<double*>(p);
  const int64_t rows = s.n_ao + (el2_q_in_lds(s) ? 0 : s.n_orb);
  p += W * (int64_t)s.n_el * rows * (int64_t)sizeof(double);
  if (s.precision == JQMC_PRECISION_MIXED_FP32) {
    x.count = reinterpret_cast<int*>(p);
    p += 16;
    x.list = reinterpret_cast<int*>(p);
    p += (W * (int64_t)sizeof(int) + 15) & ~(int64_t)15;
    x.scratch = reinterpret_cast<double*>(p);
    x.rec = GuardRec::doubles(s);
  }
  return x;
}

template <typename T>
int launch_local_energy(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const double* RT, double* e_L,
                        double* kinetic, double* v_total, double* t_elec, void* workspace, hipStream_t st) {
  const Sys& s = sys->s;
  const int64_t per_wave = el2_lds_bytes(s);
  if (per_wave > kMaxLdsPerBlock) return fail(JQMC_ERR_UNSUPPORTED, "system too large for the E_L LDS budget");
  const int wpb = tuned_wpb(per_wave, sys->tune_el_wpb);
  const dim3 grid(grid_for(W, wpb)), block(wpb * WAVE);
  const size_t lds = (size_t)wpb * per_wave;
  const WalkT<T> k = to_walk<T>(wk);
  const ElWs ws = el_ws(s, W, workspace);
  ElGuard guard{nullptr, nullptr, 0.0, 0.0};
  const bool guarded = std::is_same<T, float>::value && sys->guard_width > 0.0;
  if (guarded && guard_lds_bytes(s, 4) > kMaxLdsPerBlock)
    return fail(JQMC_ERR_UNSUPPORTED, "system too large for the E_L guard LDS budget");
  if (guarded) {
    guard = ElGuard{ws.list, ws.count, sys->guard_e_ref, sys->guard_width};
    k_zero_int<<<1, 64, 0, st>>>(ws.count);
  }
  const int team = sys->tune_el_team;
  if ((team == 2 || team == 4) && el2_lds_bytes(s, team) <= kMaxLdsPerBlock) {
    if (W > 0x7FFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for one block per walker");
    if (team == 2)
      k_local_energy_team<T, 2><<<(unsigned)W, 2 * WAVE, el2_lds_bytes(s, 2), st>>>(s, W, k, RT, e_L, kinetic, v_total,
                                                                                  t_elec, ws.hws, guard);
    else
      k_local_energy_team<T, 4><<<(unsigned)W, 4 * WAVE, el2_lds_bytes(s, 4), st>>>(s, W, k, RT, e_L, kinetic, v_total,
                                                                                  t_elec, ws.hws, guard);
  } else if (wpb == 4)
    k_local_energy_pts<T, 4><<<grid, block, lds, st>>>(s, W, k, RT, e_L, kinetic, v_total, t_elec, ws.hws, guard,
                                                       per_wave);
  else if (wpb == 2)
    k_local_energy_pts<T, 2><<<grid, block, lds, st>>>(s, W, k, RT, e_L, kinetic, v_total, t_elec, ws.hws, guard,
                                                       per_wave);
  else
    k_local_energy_pts<T, 1><<<grid, block, lds, st>>>(s, W, k, RT, e_L, kinetic, v_total, t_elec, ws.hws, guard,
                                                       per_wave);
  if constexpr (std::is_same<T, float>::value) {
    if (guarded) {
      const unsigned slots = (unsigned)std::min<int64_t>(GUARD_SLOTS, W);
      if (guard_lds_bytes(s, 8) <= kMaxLdsPerBlock)
        k_local_energy_guard<512><<<slots, 512, guard_lds_bytes(s, 8), st>>>(
            s, W, k, RT, e_L, kinetic, v_total, t_elec, ws.hws, ws.list, ws.count, ws.scratch, ws.rec);
      else
        k_local_energy_guard<256><<<slots, 256, guard_lds_bytes(s, 4), st>>>(
            s, W, k, RT, e_L, kinetic, v_total, t_elec, ws.hws, ws.list, ws.count, ws.scratch, ws.rec);
    }
  }
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

// Per-walker rows of the walker state copied by reconfiguration (wk may be NULL to size the staging
// buffer): positions, AO / orbital / pair / J3-field caches, G, G^-1 and the per-walker scalars.
StateRows state_rows(const Sys& s, const jqmc_walkers* wk) {
  StateRows R{};
  const int64_t cb = s.precision == JQMC_PRECISION_MIXED_FP32 ? 4 : 8;
  int64_t off = 0;
  auto add = [&](void* base, int64_t bytes) {
    if (bytes <= 0) return;
    R.base[R.n] = static_cast<char*>(base);
    R.bytes[R.n] = bytes;
    R.toff[R.n] = off;
    off += (bytes + 15) & ~(int64_t)15;
    ++R.n;
  };
  const int64_t ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up, nd = s.n_dn;
  add(wk ? wk->r_up : nullptr, 3 * nu * 8);
  add(wk ? wk->r_dn : nullptr, 3 * nd * 8);
  add(wk ? wk->ao : nullptr, ne * na * cb);
  add(wk ? wk->orb : nullptr, ne * no * 8);
  add(wk ? wk->pair : nullptr, ne * no * 8);
  if (s.j3_on) add(wk ? wk->jfield : nullptr, ne * na * cb);
  add(wk ? wk->G : nullptr, nu * nu * 8);
  add(wk ? wk->Ginv : nullptr, nu * nu * 8);
  add(wk ? wk->logabsdet : nullptr, 8);
  add(wk ? wk->sign : nullptr, 8);
  add(wk ? wk->jastrow : nullptr, 8);
  if (!wk || wk->as_factor) add(wk ? wk->as_factor : nullptr, 8);
  R.rec = off;
  return R;
}

template <typename T>
int launch_point_ratios(const Sys& s, int64_t W, const jqmc_walkers* wk, int64_t P, const int32_t* elec,
                        const double* pts, double* det, double* jas, hipStream_t st) {
  const int64_t lds = ratio_lds_bytes<T>(s);
  if (lds > kMaxLdsPerBlock) return fail(JQMC_ERR_UNSUPPORTED, "system too large for the ratio LDS budget");
  k_point_ratios<T><<<(unsigned)W, WAVE, lds, st>>>(s, W, to_walk<T>(wk), P, elec, pts, det, jas, lds);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

// GFMC_t: every LR_POLL projections (and after the last) the counts of walkers with time left after
// projection it are read back (slot it & 1 of each counter, one per system); *done when all are zero.
constexpr int LR_POLL = 8;
int lr_poll(int* const* act, int n, int it, int* host, hipStream_t st, bool* done) {
  for (int k = 0; k < n; ++k)
    HIP_TRY(hipMemcpyAsync(host + k, act[k] + (it & 1), sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int total = 0;
  for (int k = 0; k < n; ++k) total += host[k];
  *done = total == 0;
  return JQMC_OK;
}

// Waves per walker of the LRDMC Hamiltonian launch: a team of LR_TEAM waves when W one-wave walkers
// would leave fewer than 8 waves per SIMD of the 1024 (W < 8192), the walker has work enough to
// share (n_el n_ao >= 4096: measured, the team loses on the 10- and 30-electron chains with 24 and
// 120 AOs and wins from benzene's 30 x 228 up) and the team's LDS fits; else one wave.  A function
// of the system and W only, so that the single and the multi-system launch take the same
// arithmetic.
template <typename T>
int lr_team(const Sys& s, int64_t W) {
  return (W < 8192 && (int64_t)s.n_el * s.n_ao >= 4096 && lr_ham_lds_bytes<T>(s, LR_TEAM) <= kMaxLdsPerBlock)
             ? LR_TEAM
             : 1;
}

// The LRDMC move's dA rows and movers in the form the J3 product kernel takes (no offsets).
template <typename T>
SweepWs<T> lr_jfield_ws(const LrState& S) {
  SweepWs<T> ws{};
  ws.dA = reinterpret_cast<T*>(S.dA);
  ws.mv = S.mv;
  return ws;
}

// One Hamiltonian launch, one move launch and (with J3) one J3 field launch per projection
// (LR_MEASURE: the Hamiltonian only).
template <typename T>
int launch_lrdmc(const Sys& s, int64_t W, const jqmc_walkers* wk, const LrArgs& A, int* host_flags, hipStream_t st) {
  const int64_t per_ham = lr_ham_lds_bytes<T>(s), per_move = lr_move_lds_bytes<T>(s);
  if (std::max(per_ham, per_move) > kMaxLdsPerBlock)
    return fail(JQMC_ERR_UNSUPPORTED, "system too large for the LRDMC LDS budget");
  const int wpb_h = walkers_per_block(per_ham), wpb_m = walkers_per_block(per_move);
  const int team = lr_team<T>(s, W);
  const int64_t team_ham = lr_ham_lds_bytes<T>(s, LR_TEAM);
  const WalkT<T> k = to_walk<T>(wk);
  const int rounds = A.mode == LR_MEASURE ? 1 : A.n_steps;
  for (int it = 0; it < rounds; ++it) {
    if (team == LR_TEAM)
      (lr_wide(s) ? k_lrdmc_ham_team<T, true> : k_lrdmc_ham_team<T, false>)<<<(unsigned)W, LR_TEAM * WAVE, team_ham,
                                                                              st>>>(s, W, k, A, it);
    else
      (lr_wide(s) ? k_lrdmc_ham<T, true> : k_lrdmc_ham<T, false>)<<<grid_for(W, wpb_h), wpb_h * WAVE, wpb_h * per_ham,
                                                                    st>>>(s, W, k, A, it, wpb_h, per_ham);
    if (A.mode != LR_MEASURE) {
      k_lrdmc_move<T><<<grid_for(W, wpb_m), wpb_m * WAVE, wpb_m * per_move, st>>>(s, W, k, A, it, wpb_m, per_move);
      if (s.j3_on) {
        const dim3 jg(grid_for(s.n_ao, JF_ROWS), grid_for(W, JF_WALKERS));
        k_sweep_jfield_mfma<T, true><<<jg, 256, 0, st>>>(s, W, k, lr_jfield_ws<T>(A.st));
      }
    }
    if (A.mode == LR_FIXED_TAU && (it % LR_POLL == LR_POLL - 1 || it == rounds - 1)) {
      HIP_TRY(hipGetLastError());
      bool done = false;
      int* act = A.st.act;
      const int rc = lr_poll(&act, 1, it, host_flags, st, &done);
      if (rc != JQMC_OK) return rc;
      if (done) break;
    }
  }
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

// The J3 field launch of every system of a multi-system move launch that has J3 and is still
// projecting at round it (one launch per system: the products tile one system's walkers).
template <typename T>
void lr_multi_jfield(const LrMulti<T>& M, const Sys* const* hsys, int it, hipStream_t st) {
  for (int slot = 0; slot < M.n_sys; ++slot) {
    const Sys& s = *hsys[slot];
    const int64_t W = M.sys[slot].W;
    if (!s.j3_on || W == 0 || it >= M.args[slot].n_steps) continue;
    const dim3 jg(grid_for(s.n_ao, JF_ROWS), grid_for(W, JF_WALKERS));
    k_sweep_jfield_mfma<T, true><<<jg, 256, 0, st>>>(s, W, M.wk[slot], lr_jfield_ws<T>(M.args[slot].st));
  }
}

}  // namespace

extern "C" {

int jqmc_abi_version(void) { return JQMC_ABI_VERSION; }

int jqmc_system_set_tuning(jqmc_system* sys, int32_t knob, int32_t value) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  switch (knob) {
    case JQMC_TUNE_EL_WALKERS_PER_BLOCK:
    case JQMC_TUNE_SWEEP_WALKERS_PER_BLOCK:
      if (value != 0 && value != 1 && value != 2 && value != 4)
        return fail(JQMC_ERR_INVALID_ARGUMENT, "walkers per block must be 0 (default), 1, 2 or 4");
      (knob == JQMC_TUNE_EL_WALKERS_PER_BLOCK ? sys->tune_el_wpb : sys->tune_sweep_wpb) = value;
      return JQMC_OK;
    case JQMC_TUNE_EL_WAVES_PER_WALKER:
      if (value != 0 && value != 1 && value != 2 && value != 4)
        return fail(JQMC_ERR_INVALID_ARGUMENT, "waves per walker must be 0 (default), 1, 2 or 4");
      sys->tune_el_team = value;
      return JQMC_OK;
    default:
      return fail(JQMC_ERR_INVALID_ARGUMENT, "unknown tuning knob %d", (int)knob);
  }
}

int jqmc_system_set_mixed_guard(jqmc_system* sys, double e_ref, double width) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  if (!std::isfinite(e_ref) || std::isnan(width)) return fail(JQMC_ERR_INVALID_ARGUMENT, "e_ref and width must be numbers");
  sys->guard_e_ref = e_ref;
  sys->guard_width = width;
  return JQMC_OK;
}

int jqmc_system_get_tuning(const jqmc_system* sys, int32_t knob, int32_t* value) {
  if (!sys || !value) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  switch (knob) {
    case JQMC_TUNE_EL_WALKERS_PER_BLOCK: *value = sys->tune_el_wpb; return JQMC_OK;
    case JQMC_TUNE_SWEEP_WALKERS_PER_BLOCK: *value = sys->tune_sweep_wpb; return JQMC_OK;
    case JQMC_TUNE_EL_WAVES_PER_WALKER: *value = sys->tune_el_team; return JQMC_OK;
    default: return fail(JQMC_ERR_INVALID_ARGUMENT, "unknown tuning knob %d", (int)knob);
  }
}

const char* jqmc_last_error(void) { return g_last_error.c_str(); }

int jqmc_system_create(const jqmc_system_desc* d, jqmc_system** out) {
  if (!d || !out) return fail(JQMC_ERR_INVALID_ARGUMENT, "desc/out is NULL");
  *out = nullptr;
  if (d->n_nuc <= 0 || d->n_ao <= 0 || d->n_prim <= 0) return fail(JQMC_ERR_SHAPE, "empty nuclei or basis");
  if (d->n_up < d->n_dn || d->n_dn < 0 || d->n_up <= 0) return fail(JQMC_ERR_SHAPE, "need n_up >= n_dn >= 0, n_up > 0");
  if (d->n_up > 64) return fail(JQMC_ERR_UNSUPPORTED, "n_up > 64 is not supported by the wavefront LU");
  const int n_orb = d->orb_is_mo ? d->n_orb : d->n_ao;
  if (n_orb < d->n_up) return fail(JQMC_ERR_SHAPE, "fewer orbitals than up electrons");
  if (d->ecp_on && d->nn != 1) return fail(JQMC_ERR_UNSUPPORTED, "only NN = 1 nearest nucleus is supported");
  if (d->ecp_on && d->ecp_lmax > 7) return fail(JQMC_ERR_UNSUPPORTED, "projectors above l = 6 are not tabulated");
  if (d->j1_type < 0 || d->j1_type > 2 || d->j2_type < 0 || d->j2_type > 2)
    return fail(JQMC_ERR_INVALID_ARGUMENT, "unknown Jastrow type");
  if (d->precision != JQMC_PRECISION_PLAIN_FP64 && d->precision != JQMC_PRECISION_MIXED_FP32)
    return fail(JQMC_ERR_INVALID_ARGUMENT, "unknown precision %d", (int)d->precision);
  for (int a = 0; a < d->n_ao; ++a) {
    if (d->ao_prim_ptr[a + 1] < d->ao_prim_ptr[a]) return fail(JQMC_ERR_SHAPE, "ao_prim_ptr not monotone");
    if (d->ao_nuc[a] < 0 || d->ao_nuc[a] >= d->n_nuc) return fail(JQMC_ERR_SHAPE, "ao_nuc out of range");
  }
  if (d->ao_prim_ptr[0] != 0 || d->ao_prim_ptr[d->n_ao] != d->n_prim) return fail(JQMC_ERR_SHAPE, "ao_prim_ptr ends");

  jqmc_system* sys = new jqmc_system();
  HIP_TRY(hipGetDevice(&sys->device));
  Sys& s = sys->s;
  std::memset(&s, 0, sizeof(Sys));
  s.n_nuc = d->n_nuc;
  s.n_ao = d->n_ao;
  s.n_prim = d->n_prim;
  s.n_up = d->n_up;
  s.n_dn = d->n_dn;
  s.n_el = d->n_up + d->n_dn;
  s.n_orb = n_orb;
  s.n_unp = d->n_up - d->n_dn;
  s.orb_is_mo = d->orb_is_mo ? 1 : 0;
  s.j1_type = d->j1_type;
  s.j2_type = d->j2_type;
  s.j3_on = d->j3_on ? 1 : 0;
  s.precision = d->precision;
  s.j1_param = d->j1_param;
  s.j2_param = d->j2_param;
  s.ecp_on = d->ecp_on ? 1 : 0;

  // nuclei and ion-ion energy
  double ion_ion = 0.0;
  for (int a = 0; a < d->n_nuc; ++a)
    for (int b = a + 1; b < d->n_nuc; ++b) {
      const double dx = d->nuc_pos[3 * a] - d->nuc_pos[3 * b], dy = d->nuc_pos[3 * a + 1] - d->nuc_pos[3 * b + 1],
                   dz = d->nuc_pos[3 * a + 2] - d->nuc_pos[3 * b + 2];
      ion_ion += d->nuc_charge[a] * d->nuc_charge[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  s.ion_ion = ion_ion;
  UPLOAD(d->nuc_pos, 3 * d->n_nuc, s.nuc_pos);
  UPLOAD(d->nuc_charge, d->n_nuc, s.nuc_charge);
  if (d->j1_type) {
    std::vector<double> c(d->n_nuc), A(d->n_nuc);
    for (int n = 0; n < d->n_nuc; ++n) {
      c[n] = std::pow(2.0 * d->j1_zeff[n], 1.0 / 4.0);
      A[n] = std::pow(2.0 * d->j1_zeff[n], 3.0 / 4.0);
    }
    UPLOAD(c.data(), d->n_nuc, s.j1_c);
    UPLOAD(A.data(), d->n_nuc, s.j1_A);
  }

  // basis
  std::vector<double> ctr(3 * d->n_ao);
  for (int a = 0; a < d->n_ao; ++a)
    for (int c = 0; c < 3; ++c) ctr[3 * a + c] = d->nuc_pos[3 * d->ao_nuc[a] + c];
  UPLOAD(ctr.data(), ctr.size(), s.ao_ctr);
  {
    // angular parts as monomial terms (one unit term per Cartesian AO)
    std::vector<int32_t> tp(d->n_ao + 1), tpow;
    std::vector<double> tc;
    if (d->ao_term_ptr) {
      if (d->n_terms <= 0 || !d->term_coef || !d->term_pow || d->ao_term_ptr[0] != 0 ||
          d->ao_term_ptr[d->n_ao] != d->n_terms) {
        jqmc_system_destroy(sys);
        return fail(JQMC_ERR_SHAPE, "inconsistent AO term arrays");
      }
      for (int a = 0; a <= d->n_ao; ++a) tp[a] = d->ao_term_ptr[a];
      tc.assign(d->term_coef, d->term_coef + d->n_terms);
      tpow.assign(d->term_pow, d->term_pow + 3 * d->n_terms);
    } else {
      if (!d->ao_pow) {
        jqmc_system_destroy(sys);
        return fail(JQMC_ERR_INVALID_ARGUMENT, "ao_pow or ao_term_ptr is required");
      }
      for (int a = 0; a < d->n_ao; ++a) {
        tp[a] = a;
        tc.push_back(1.0);
        for (int c = 0; c < 3; ++c) tpow.push_back(d->ao_pow[3 * a + c]);
      }
      tp[d->n_ao] = d->n_ao;
    }
    UPLOAD(tp.data(), tp.size(), s.ao_tp);
    UPLOAD_REAL(tc.data(), tc.size(), term_coef);
    UPLOAD(tpow.data(), tpow.size(), s.term_pow);
  }
  UPLOAD(d->ao_prim_ptr, d->n_ao + 1, s.ao_pp);
  UPLOAD(d->ao_nuc, d->n_ao, s.ao_nuc);
  {
    // shells: consecutive AOs on one nucleus with the same exponents and proportional coefficients
    std::vector<int32_t> sh0{0};
    std::vector<double> scale(d->n_ao, 1.0);
    int first = 0;
    for (int a = 1; a < d->n_ao; ++a) {
      const int p0 = d->ao_prim_ptr[first], np0 = d->ao_prim_ptr[first + 1] - p0;
      const int q0 = d->ao_prim_ptr[a], nq = d->ao_prim_ptr[a + 1] - q0;
      bool same = d->ao_nuc[a] == d->ao_nuc[first] && nq == np0 && np0 > 0;
      double ratio = 0.0;
      for (int k = 0; same && k < np0; ++k) {
        if (d->prim_exp[q0 + k] != d->prim_exp[p0 + k] || d->prim_coef[p0 + k] == 0.0) {
          same = false;
          break;
        }
        const double r = d->prim_coef[q0 + k] / d->prim_coef[p0 + k];
        if (k == 0) ratio = r;
        else if (std::fabs(r - ratio) > 1e-12 * std::fabs(ratio)) same = false;
      }
      if (same) {
        scale[a] = ratio;
      } else {
        sh0.push_back(a);
        first = a;
      }
    }
    sh0.push_back(d->n_ao);
    s.n_shell = (int)sh0.size() - 1;
    if (s.n_shell > 0xffff) {  // the packed AO meta word holds the shell in 16 bits
      jqmc_system_destroy(sys);
      return fail(JQMC_ERR_UNSUPPORTED, "more than 65535 shells");
    }
    std::vector<int32_t> ash(d->n_ao);
    for (int k = 0; k < s.n_shell; ++k)
      for (int a = sh0[k]; a < sh0[k + 1]; ++a) ash[a] = k;
    UPLOAD(sh0.data(), sh0.size(), s.sh_ao0);
    UPLOAD(ash.data(), ash.size(), s.ao_sh);
    // packed shell | powers | unit-term flag per AO (the flag only where the powers fit 4 bits)
    std::vector<int32_t> meta(d->n_ao);
    std::vector<int32_t> tpa(d->n_ao + 1);
    if (d->ao_term_ptr) {
      for (int a = 0; a <= d->n_ao; ++a) tpa[a] = d->ao_term_ptr[a];
    } else {
      for (int a = 0; a <= d->n_ao; ++a) tpa[a] = a;
    }
    for (int a = 0; a < d->n_ao; ++a) {
      int m = ash[a] & 0xffff;
      const int t = tpa[a];
      const int* pw = d->ao_term_ptr ? d->term_pow + 3 * t : d->ao_pow + 3 * a;
      const double c = d->ao_term_ptr ? d->term_coef[t] : 1.0;
      const bool fits = pw[0] >= 0 && pw[0] < 16 && pw[1] >= 0 && pw[1] < 16 && pw[2] >= 0 &&
                        pw[2] < 16;
      if (tpa[a + 1] - t == 1 && c == 1.0 && fits) m |= (pw[0] << 16) | (pw[1] << 20) | (pw[2] << 24) | AO_UNIT_TERM;
      meta[a] = m;
    }
    UPLOAD(meta.data(), meta.size(), s.ao_meta);
    UPLOAD_REAL(scale.data(), scale.size(), ao_scale);
    // shell records (Sys::shrec)
    std::vector<double> rec(static_cast<size_t>(s.n_shell) * SHREC, 0.0);
    for (int k = 0; k < s.n_shell; ++k) {
      double* r = rec.data() + static_cast<size_t>(k) * SHREC;
      const int a0 = sh0[k], a1 = sh0[k + 1];
      const int p0 = d->ao_prim_ptr[a0], np = d->ao_prim_ptr[a0 + 1] - p0, nao = a1 - a0;
      for (int c = 0; c < 3; ++c) r[c] = ctr[3 * a0 + c];
      const int l = nao == 1 ? 0 : (nao == 3 ? 1 : (nao == 6 ? 2 : -1));
      bool fast = np >= 1 && np <= SHR_PMAX && l >= 0;
      if (fast) {
        // canonical Cartesian order: nx descending, then ny descending
        int i = 0;
        for (int nx = l; nx >= 0 && fast; --nx)
          for (int ny = l - nx; ny >= 0 && fast; --ny, ++i) {
            const int m = meta[a0 + i], nz = l - nx - ny;
            fast = (m & AO_UNIT_TERM) && ((m >> 16) & 15) == nx && ((m >> 20) & 15) == ny && ((m >> 24) & 15) == nz;
          }
      }
      r[3] = (double)(fast ? (np | (l << 4) | (1 << 8)) : 0) + 512.0 * a0;
      if (fast) {
        for (int q = 0; q < np; ++q) {
          r[SHR_P + 2 * q] = d->prim_exp[p0 + q];
          r[SHR_P + 2 * q + 1] = d->prim_coef[p0 + q];
        }
        for (int i = 0; i < nao; ++i) r[SHR_SC + i] = scale[a0 + i];
      }
    }
    UPLOAD(rec.data(), rec.size(), s.shrec);
  }
  UPLOAD_REAL(d->prim_exp, d->n_prim, prim_exp);
  UPLOAD_REAL(d->prim_coef, d->n_prim, prim_coef);

  // orbitals and pairing matrix
  const int nu = d->n_up, nd = d->n_dn, unp = nu - nd;
  if (d->orb_is_mo) {
    std::vector<double> t(static_cast<size_t>(n_orb) * d->n_ao);
    UPLOAD_REAL(d->mo_coef_up, (size_t)n_orb * d->n_ao, mo_up);
    UPLOAD_REAL(d->mo_coef_dn, (size_t)n_orb * d->n_ao, mo_dn);
    for (int m = 0; m < n_orb; ++m)
      for (int a = 0; a < d->n_ao; ++a) t[(size_t)a * n_orb + m] = d->mo_coef_up[(size_t)m * d->n_ao + a];
    UPLOAD_REAL(t.data(), t.size(), mo_upT);
    for (int m = 0; m < n_orb; ++m)
      for (int a = 0; a < d->n_ao; ++a) t[(size_t)a * n_orb + m] = d->mo_coef_dn[(size_t)m * d->n_ao + a];
    UPLOAD_REAL(t.data(), t.size(), mo_dnT);
  }
  {
    const int ncol = n_orb + unp;
    std::vector<double> lp((size_t)n_orb * n_orb), lpt((size_t)n_orb * n_orb), lu((size_t)n_orb * (unp > 0 ? unp : 1));
    for (int i = 0; i < n_orb; ++i) {
      for (int j = 0; j < n_orb; ++j) {
        lp[(size_t)i * n_orb + j] = d->lambda_matrix[(size_t)i * ncol + j];
        lpt[(size_t)j * n_orb + i] = d->lambda_matrix[(size_t)i * ncol + j];
      }
      for (int k = 0; k < unp; ++k) lu[(size_t)i * unp + k] = d->lambda_matrix[(size_t)i * ncol + n_orb + k];
    }
    UPLOAD_REAL(lp.data(), lp.size(), lam_p);
    UPLOAD_REAL(lpt.data(), lpt.size(), lam_pT);
    if (unp > 0) UPLOAD_REAL(lu.data(), (size_t)n_orb * unp, lam_u);
  }

  // J3 over the AO basis
  if (d->j3_on) {
    const int na = d->n_ao;
    std::vector<double> J((size_t)na * na), JT((size_t)na * na), jv(na);
    for (int a = 0; a < na; ++a) {
      for (int b = 0; b < na; ++b) {
        J[(size_t)a * na + b] = d->j3_matrix[(size_t)a * (na + 1) + b];
        JT[(size_t)b * na + a] = d->j3_matrix[(size_t)a * (na + 1) + b];
      }
      jv[a] = d->j3_matrix[(size_t)a * (na + 1) + na];
    }
    s.j3_sym = (J == JT) ? 1 : 0;  // exact symmetry only: then J3 dA and J3^T dA coincide bitwise
    UPLOAD_REAL(J.data(), J.size(), j3);
    UPLOAD_REAL(JT.data(), JT.size(), j3T);
    UPLOAD_REAL(jv.data(), jv.size(), j3v);
  }

  // ECP tables
  if (d->ecp_on) {
    s.ecp_lmax = d->ecp_lmax;
    s.ecp_max_terms = d->ecp_max_terms;
    s.ecp_max_local = d->ecp_max_local;
    s.nv = d->nv;
    const size_t nt = (size_t)d->n_nuc * (d->ecp_max_terms > 0 ? d->ecp_max_terms : 0);
    const size_t nl = (size_t)d->n_nuc * (d->ecp_max_local > 0 ? d->ecp_max_local : 0);
    UPLOAD(d->ecp_n_terms, d->n_nuc, s.ecp_nt);
    UPLOAD(d->ecp_term_l, nt, s.ecp_l);
    UPLOAD(d->ecp_term_exp, nt, s.ecp_exp);
    UPLOAD(d->ecp_term_coef, nt, s.ecp_coef);
    UPLOAD(d->ecp_term_pow, nt, s.ecp_pow);
    UPLOAD(d->ecpl_n_terms, d->n_nuc, s.ecpl_nt);
    UPLOAD(d->ecpl_exp, nl, s.ecpl_exp);
    UPLOAD(d->ecpl_coef, nl, s.ecpl_coef);
    UPLOAD(d->ecpl_pow, nl, s.ecpl_pow);
    UPLOAD(d->mesh_points, 3 * (size_t)d->nv, s.mesh_pts);
    UPLOAD(d->mesh_weights, (size_t)d->nv, s.mesh_w);
  }
  UPLOAD(&s, 1, sys->d_sys);
  {
    void* h = nullptr;
    HIP_TRY(hipHostMalloc(&h, LR_MAX_SYSTEMS * sizeof(int), hipHostMallocDefault));
    sys->host_flags = static_cast<int*>(h);
  }
  *out = sys;
  return JQMC_OK;
}

int jqmc_system_destroy(jqmc_system* sys) {
  if (!sys) return JQMC_OK;
  int rc = JQMC_OK;
  for (void* p : sys->allocations)
    if (hipFree(p) != hipSuccess) rc = fail(JQMC_ERR_HIP, "hipFree failed");
  if (sys->host_flags && hipHostFree(sys->host_flags) != hipSuccess) rc = fail(JQMC_ERR_HIP, "hipHostFree failed");
  delete sys;
  return rc;
}

int jqmc_system_precision(const jqmc_system* sys, int32_t* precision) {
  if (!sys || !precision) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  *precision = sys->s.precision;
  return JQMC_OK;
}

int jqmc_aos_value_grad_lap(const jqmc_system* sys, int64_t n_pts, const double* r, double* val, double* gx,
                            double* gy, double* gz, double* lap, void* stream) {
  if (!sys || !r || !val) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n_pts < 0) return fail(JQMC_ERR_SHAPE, "n_pts < 0");
  if (gx && (!gy || !gz || !lap)) return fail(JQMC_ERR_INVALID_ARGUMENT, "grad outputs must all be given");
  const int64_t total = (int64_t)sys->s.n_ao * n_pts;
  if (total == 0) return JQMC_OK;
  k_aos_vgl<<<grid_for(total, 256), 256, 0, (hipStream_t)stream>>>(sys->s, n_pts, r, val, gx, gy, gz, lap);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_orbs_value_grad_lap(const jqmc_system* sys, int32_t spin, int64_t n_pts, const double* r, double* val,
                             double* gx, double* gy, double* gz, double* lap, double* workspace, void* stream) {
  if (!sys || !r || !val || !workspace) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!sys->s.orb_is_mo) return fail(JQMC_ERR_UNSUPPORTED, "orbitals are AOs; use jqmc_aos_value_grad_lap");
  if (spin != 0 && spin != 1) return fail(JQMC_ERR_INVALID_ARGUMENT, "spin must be 0 or 1");
  if (n_pts == 0) return JQMC_OK;
  const Sys& s = sys->s;
  const int64_t plane = (int64_t)s.n_ao * n_pts;
  const bool full = gx != nullptr;
  int rc = jqmc_aos_value_grad_lap(sys, n_pts, r, workspace, full ? workspace + plane : nullptr,
                                   full ? workspace + 2 * plane : nullptr, full ? workspace + 3 * plane : nullptr,
                                   full ? workspace + 4 * plane : nullptr, stream);
  if (rc != JQMC_OK) return rc;
  const int64_t total = (int64_t)s.n_orb * n_pts;
  k_orbs_contract<<<grid_for(total, 256), 256, 0, (hipStream_t)stream>>>(
      s.n_orb, s.n_ao, n_pts, spin == 0 ? s.td.mo_up : s.td.mo_dn, workspace, val, gx, gy, gz, lap, full ? 5 : 1);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_walkers_init_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  *bytes = s.j3_on ? 2 * W * s.n_el * s.n_ao * cache_bytes(s) : 0;
  return JQMC_OK;
}

int jqmc_walkers_init(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, void* workspace,
                      int64_t workspace_bytes, void* stream) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (!wk->logabsdet || !wk->sign || !wk->jastrow) return fail(JQMC_ERR_INVALID_ARGUMENT, "scalar outputs are NULL");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  int64_t need = 0;
  jqmc_walkers_init_workspace_bytes(sys, W, &need);
  if (need > 0 && (!workspace || workspace_bytes < need))
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)need);
  hipStream_t st = (hipStream_t)stream;
  rc = s.precision == JQMC_PRECISION_MIXED_FP32 ? launch_init<float>(s, W, wk, workspace, st)
                                                : launch_init<double>(s, W, wk, workspace, st);
  if (rc != JQMC_OK) return rc;
  if (wk->accepted) HIP_TRY(hipMemsetAsync(wk->accepted, 0, W * sizeof(int64_t), st));
  if (wk->rejected) HIP_TRY(hipMemsetAsync(wk->rejected, 0, W * sizeof(int64_t), st));
  return JQMC_OK;
}

int jqmc_walkers_refresh_inverse(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, void* stream) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (!wk->logabsdet || !wk->sign) return fail(JQMC_ERR_INVALID_ARGUMENT, "scalar outputs are NULL");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32 ? launch_refresh<float>(s, W, wk, st, false)
                                                  : launch_refresh<double>(s, W, wk, st, false);
}

int jqmc_point_ratios(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int64_t n_points,
                      const int32_t* elec, const double* points, double* det_ratio, double* jastrow_ratio,
                      void* stream) {
  if (!sys || !elec || !points || !det_ratio || !jastrow_ratio)
    return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (W < 0 || n_points < 0) return fail(JQMC_ERR_SHAPE, "negative size");
  if (W > 0x7FFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for one launch");
  if (W == 0 || n_points == 0) return JQMC_OK;
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32
             ? launch_point_ratios<float>(s, W, wk, n_points, elec, points, det_ratio, jastrow_ratio, st)
             : launch_point_ratios<double>(s, W, wk, n_points, elec, points, det_ratio, jastrow_ratio, st);
}

int jqmc_walkers_refresh_orbitals(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, void* stream) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (!wk->logabsdet || !wk->sign) return fail(JQMC_ERR_INVALID_ARGUMENT, "scalar outputs are NULL");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32 ? launch_refresh<float>(s, W, wk, st, true)
                                                  : launch_refresh<double>(s, W, wk, st, true);
}

namespace {

int vmc_sweep_impl(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                   double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, void* stream) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (!wk->accepted || !wk->rejected) return fail(JQMC_ERR_INVALID_ARGUMENT, "accept/reject counters are NULL");
  if (n_moves < 0 || W < 0) return fail(JQMC_ERR_SHAPE, "negative size");
  if (!(Dt > 0.0)) return fail(JQMC_ERR_INVALID_ARGUMENT, "Dt must be positive");
  if (W > 0xFFFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for the 32-bit walker counter");
  if (W == 0 || n_moves == 0) return JQMC_OK;
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32
             ? launch_sweep<float>(sys, W, wk, n_moves, Dt, epsilon_AS, seed, move0, move_ctr, st)
             : launch_sweep<double>(sys, W, wk, n_moves, Dt, epsilon_AS, seed, move0, move_ctr, st);
}

}  // namespace

int jqmc_vmc_sweep(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                   double epsilon_AS, uint64_t seed, int64_t move0, void* stream) {
  return vmc_sweep_impl(sys, W, wk, n_moves, Dt, epsilon_AS, seed, move0, nullptr, stream);
}

int jqmc_vmc_sweep_counted(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                           double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, void* stream) {
  if (!move_ctr) return fail(JQMC_ERR_INVALID_ARGUMENT, "move_ctr is NULL");
  return vmc_sweep_impl(sys, W, wk, n_moves, Dt, epsilon_AS, seed, move0, move_ctr, stream);
}

int jqmc_vmc_sweep_batched_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W < 0) return fail(JQMC_ERR_SHAPE, "W < 0");
  *bytes = sys->s.precision == JQMC_PRECISION_MIXED_FP32 ? SweepWs<float>::bytes(sys->s, W)
                                                         : SweepWs<double>::bytes(sys->s, W);
  return JQMC_OK;
}

int jqmc_vmc_sweep_batched(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t n_moves, double Dt,
                           double epsilon_AS, uint64_t seed, int64_t move0, const int64_t* move_ctr, void* workspace,
                           int64_t workspace_bytes, void* stream) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "sys is NULL");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (!wk->accepted || !wk->rejected) return fail(JQMC_ERR_INVALID_ARGUMENT, "accept/reject counters are NULL");
  if (n_moves < 0 || W < 0) return fail(JQMC_ERR_SHAPE, "negative size");
  if (!(Dt > 0.0)) return fail(JQMC_ERR_INVALID_ARGUMENT, "Dt must be positive");
  if (W > 0xFFFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for the 32-bit walker counter");
  if (W == 0 || n_moves == 0) return JQMC_OK;
  int64_t need = 0;
  jqmc_vmc_sweep_batched_workspace_bytes(sys, W, &need);
  if (!workspace || workspace_bytes < need)
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)need);
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32
             ? launch_sweep_batched<float>(sys, W, wk, n_moves, Dt, epsilon_AS, seed, move0, move_ctr, workspace, st)
             : launch_sweep_batched<double>(sys, W, wk, n_moves, Dt, epsilon_AS, seed, move0, move_ctr, workspace, st);
}

int jqmc_rotation_matrices(int64_t W, uint64_t seed, int64_t step, int32_t stream_id, double* RT, void* stream) {
  if (!RT) return fail(JQMC_ERR_INVALID_ARGUMENT, "RT is NULL");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  k_rotations<<<grid_for(W, 256), 256, 0, (hipStream_t)stream>>>(W, seed, step, nullptr, (uint32_t)stream_id, RT);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_rotation_matrices_counted(int64_t W, uint64_t seed, int64_t step, const int64_t* step_ctr,
                                   int32_t stream_id, double* RT, void* stream) {
  if (!RT || !step_ctr) return fail(JQMC_ERR_INVALID_ARGUMENT, "RT or step_ctr is NULL");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  k_rotations<<<grid_for(W, 256), 256, 0, (hipStream_t)stream>>>(W, seed, step, step_ctr, (uint32_t)stream_id, RT);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_counter_add(int64_t* ctr, int64_t inc, void* stream) {
  if (!ctr) return fail(JQMC_ERR_INVALID_ARGUMENT, "ctr is NULL");
  k_counter_add<<<1, 64, 0, (hipStream_t)stream>>>(ctr, inc);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_local_energy_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W < 0) return fail(JQMC_ERR_SHAPE, "W < 0");
  // the determinant field rows h_e of every electron (FP64), the q_e rows when they do not fit LDS,
  // and in the mixed form the FP64 guard's list and scratch
  *bytes = el_ws_bytes(sys->s, W);
  return JQMC_OK;
}

int jqmc_local_energy(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const double* RT, double* e_L,
                      double* kinetic, double* v_total, double* t_elec, void* workspace, int64_t workspace_bytes,
                      void* stream) {
  if (!sys || !e_L) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  if (W > 0x7FFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for one launch");
  int64_t need = 0;
  jqmc_local_energy_workspace_bytes(sys, W, &need);
  if (!workspace || workspace_bytes < need)
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)need);
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32
             ? launch_local_energy<float>(sys, W, wk, RT, e_L, kinetic, v_total, t_elec, workspace, st)
             : launch_local_energy<double>(sys, W, wk, RT, e_L, kinetic, v_total, t_elec, workspace, st);
}


int jqmc_position_grads_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W < 0) return fail(JQMC_ERR_SHAPE, "W < 0");
  *bytes = W * FrWs::per_walker(sys->s) * (int64_t)sizeof(double);
  return JQMC_OK;
}

int jqmc_position_grads(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const double* RT, int32_t swct,
                        double* e_L, double* dE_dr, double* dE_dR, double* dlnpsi_dr, double* dlnpsi_dR,
                        double* force_HF, double* force_PP, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!sys) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL system");
  const Sys& s = sys->s;
  if (s.precision != JQMC_PRECISION_PLAIN_FP64)
    return fail(JQMC_ERR_UNSUPPORTED, "position derivatives are computed for plain FP64 systems only");
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  int64_t need = 0;
  jqmc_position_grads_workspace_bytes(sys, W, &need);
  if (!workspace || workspace_bytes < need)
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)need);
  const int64_t per_wave = forces_lds_bytes(s);
  if (per_wave > kMaxLdsPerBlock) return fail(JQMC_ERR_UNSUPPORTED, "system too large for the forces LDS budget");
  const int wpb = walkers_per_block(per_wave);
  hipStream_t st = (hipStream_t)stream;
  k_forces<<<grid_for(W, wpb), wpb * WAVE, wpb * per_wave, st>>>(
      s, W, to_walk<double>(wk), RT, static_cast<double*>(workspace), e_L, dE_dr, dE_dR, dlnpsi_dr, dlnpsi_dR,
      force_HF, force_PP, swct ? 1 : 0, wpb, per_wave);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ LRDMC

namespace {

int lrdmc_check(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const jqmc_lrdmc_params* prm) {
  if (!sys || !prm) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (W < 0) return fail(JQMC_ERR_SHAPE, "W < 0");
  if (W > 0xFFFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for the 32-bit walker counter");
  if (!(prm->alat > 0.0)) return fail(JQMC_ERR_INVALID_ARGUMENT, "alat must be positive");
  if (prm->non_local_move != JQMC_TMOVE && prm->non_local_move != JQMC_DLTMOVE)
    return fail(JQMC_ERR_INVALID_ARGUMENT, "non_local_move must be JQMC_TMOVE or JQMC_DLTMOVE");
  if (prm->mode != JQMC_LRDMC_FIXED_STEPS && prm->mode != JQMC_LRDMC_FIXED_TAU)
    return fail(JQMC_ERR_INVALID_ARGUMENT, "unknown LRDMC mode %d", (int)prm->mode);
  if (prm->n_steps < 0) return fail(JQMC_ERR_SHAPE, "n_steps < 0");
  if (prm->mode == JQMC_LRDMC_FIXED_TAU && !(prm->tau > 0.0))
    return fail(JQMC_ERR_INVALID_ARGUMENT, "tau must be positive for JQMC_LRDMC_FIXED_TAU");
  return JQMC_OK;
}

int lrdmc_launch(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const LrArgs& A, void* stream) {
  const Sys& s = sys->s;
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32 ? launch_lrdmc<float>(s, W, wk, A, sys->host_flags, st)
                                                  : launch_lrdmc<double>(s, W, wk, A, sys->host_flags, st);
}

LrArgs lr_args(const jqmc_lrdmc_params* prm) {
  LrArgs A{};
  A.alat = prm->alat;
  A.E_scf = prm->E_scf;
  A.tau = prm->tau;
  A.n_steps = prm->n_steps;
  A.mode = prm->mode == JQMC_LRDMC_FIXED_TAU ? LR_FIXED_TAU : LR_FIXED_STEPS;
  A.dlt = prm->non_local_move == JQMC_DLTMOVE;
  A.random_mesh = prm->random_mesh != 0;
  return A;
}

}  // namespace

extern "C" {

int jqmc_lrdmc_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W < 0) return fail(JQMC_ERR_SHAPE, "W < 0");
  *bytes = lr_ws_bytes(sys->s, W);  // determinant field rows (and q rows) of every walker
  return JQMC_OK;
}

int jqmc_lrdmc_projection(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const jqmc_lrdmc_params* prm,
                          uint64_t seed, int64_t step0, double* w_L, double* RT_last, double* V_diag,
                          double* V_nondiag, double* tau_left, int64_t* n_proj, void* workspace,
                          int64_t workspace_bytes, void* stream) {
  int rc = lrdmc_check(sys, W, wk, prm);
  if (rc != JQMC_OK) return rc;
  if (!w_L) return fail(JQMC_ERR_INVALID_ARGUMENT, "w_L is NULL");
  if (W == 0) return JQMC_OK;
  if (!workspace || workspace_bytes < lr_ws_bytes(sys->s, W))
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)lr_ws_bytes(sys->s, W));
  LrArgs A = lr_args(prm);
  A.hws = static_cast<double*>(workspace);
  A.st = lr_state(sys->s, W, A.hws);
  A.seed = seed;
  A.step0 = step0;
  A.w = w_L;
  A.RT = RT_last;
  A.V_diag = V_diag;
  A.V_nondiag = V_nondiag;
  A.tau_left = tau_left;
  A.n_proj = n_proj;
  return lrdmc_launch(sys, W, wk, A, stream);
}

int jqmc_lrdmc_local_energy(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const jqmc_lrdmc_params* prm,
                            const double* RT, double* V_diag, double* V_nondiag, void* workspace,
                            int64_t workspace_bytes, void* stream) {
  int rc = lrdmc_check(sys, W, wk, prm);
  if (rc != JQMC_OK) return rc;
  if (!V_diag || !V_nondiag) return fail(JQMC_ERR_INVALID_ARGUMENT, "V_diag / V_nondiag are NULL");
  if (W == 0) return JQMC_OK;
  if (!workspace || workspace_bytes < lr_ws_bytes(sys->s, W))
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)lr_ws_bytes(sys->s, W));
  LrArgs A = lr_args(prm);
  A.hws = static_cast<double*>(workspace);
  A.st = lr_state(sys->s, W, A.hws);
  A.mode = LR_MEASURE;
  A.RT = const_cast<double*>(RT);
  A.V_diag = V_diag;
  A.V_nondiag = V_nondiag;
  return lrdmc_launch(sys, W, wk, A, stream);
}

int jqmc_lrdmc_projection_multi(int32_t n_sys, const jqmc_system* const* systems, const int64_t* W,
                                const jqmc_walkers* wks, const jqmc_lrdmc_params* params, uint64_t seed,
                                const int64_t* step0, double* const* w_L, double* const* RT_last,
                                double* const* V_diag, double* const* V_nondiag, double* const* tau_left,
                                int64_t* const* n_proj, void* const* workspaces, const int64_t* workspace_bytes,
                                void* stream) {
  if (n_sys <= 0 || n_sys > LR_MAX_SYSTEMS)
    return fail(JQMC_ERR_UNSUPPORTED, "n_sys must lie in 1..%d", LR_MAX_SYSTEMS);
  if (!systems || !W || !wks || !params || !step0 || !w_L || !workspaces || !workspace_bytes)
    return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const int prec = systems[0] ? systems[0]->s.precision : -1;
  // One walker per 64-thread block, so that a system whose walkers need a whole LDS / register
  // budget does not leave the other SIMDs of a CU idle; the systems with the most work per walker
  // take the first blocks (longest first), the smallest fill the tail.
  std::vector<int> order(n_sys);
  std::vector<double> cost(n_sys, 0.0);
  for (int k = 0; k < n_sys; ++k) {
    order[k] = k;
    if (systems[k]) {
      const Sys& sk = systems[k]->s;
      cost[k] = (double)sk.n_el * ((double)sk.n_ao + (double)sk.n_el * (double)(sk.nv > 0 ? sk.nv : 1)) +
                (double)sk.n_up * sk.n_up * sk.n_el;
    }
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  int64_t blocks = 0, lds_move = 0;
  int rounds = 0, team[LR_MAX_SYSTEMS];
  LrMulti<double> Md{};
  LrMulti<float> Mf{};
  Md.n_sys = Mf.n_sys = n_sys;
  for (int slot = 0; slot < n_sys; ++slot) {
    const int k = order[slot];
    const jqmc_system* sys = systems[k];
    int rc = lrdmc_check(sys, W[k], &wks[k], &params[k]);
    if (rc != JQMC_OK) return rc;
    if (sys->s.precision != prec) return fail(JQMC_ERR_UNSUPPORTED, "all systems of one launch need one precision");
    if (!w_L[k]) return fail(JQMC_ERR_INVALID_ARGUMENT, "w_L[%d] is NULL", k);
    if (W[k] > 0 && (!workspaces[k] || workspace_bytes[k] < lr_ws_bytes(sys->s, W[k])))
      return fail(JQMC_ERR_WORKSPACE, "workspace %d: %lld bytes required", k, (long long)lr_ws_bytes(sys->s, W[k]));
    const bool mixed = prec == JQMC_PRECISION_MIXED_FP32;
    const int64_t per_ham = mixed ? lr_ham_lds_bytes<float>(sys->s) : lr_ham_lds_bytes<double>(sys->s);
    const int64_t per_move = mixed ? lr_move_lds_bytes<float>(sys->s) : lr_move_lds_bytes<double>(sys->s);
    if (std::max(per_ham, per_move) > kMaxLdsPerBlock)
      return fail(JQMC_ERR_UNSUPPORTED, "system %d too large for the LRDMC LDS budget", k);
    team[slot] = mixed ? lr_team<float>(sys->s, W[k]) : lr_team<double>(sys->s, W[k]);
    const int64_t ham_lds = team[slot] == LR_TEAM ? (mixed ? lr_ham_lds_bytes<float>(sys->s, LR_TEAM)
                                                           : lr_ham_lds_bytes<double>(sys->s, LR_TEAM))
                                                  : per_ham;
    LrSystem e{sys->d_sys, W[k], ham_lds, per_move, blocks};
    blocks += W[k];
    lds_move = std::max<int64_t>(lds_move, per_move);
    rounds = std::max(rounds, params[k].n_steps);
    LrArgs A = lr_args(&params[k]);
    A.seed = seed + (uint64_t)k;  // system k draws from its own Philox key
    A.step0 = step0[k];
    A.w = w_L[k];
    A.RT = RT_last ? RT_last[k] : nullptr;
    A.V_diag = V_diag ? V_diag[k] : nullptr;
    A.V_nondiag = V_nondiag ? V_nondiag[k] : nullptr;
    A.tau_left = tau_left ? tau_left[k] : nullptr;
    A.n_proj = n_proj ? n_proj[k] : nullptr;
    A.hws = static_cast<double*>(workspaces[k]);
    A.st = lr_state(sys->s, W[k], A.hws);
    Md.sys[slot] = Mf.sys[slot] = e;
    Md.args[slot] = Mf.args[slot] = A;
    Md.wk[slot] = to_walk<double>(&wks[k]);
    Mf.wk[slot] = to_walk<float>(&wks[k]);
  }
  if (blocks == 0) return JQMC_OK;
  if (blocks > 0x7FFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers in one launch");
  // the Hamiltonian launches: one per (team size, determinant-field form) present, each over its
  // systems' walkers; group g = (one wave ? 1 : 0) + (lr_wide ? 2 : 0)
  LrMulti<double> Hd[4]{};
  LrMulti<float> Hf[4]{};
  int64_t h_blocks[4] = {0, 0, 0, 0}, h_lds[4] = {0, 0, 0, 0};
  for (int slot = 0; slot < n_sys; ++slot) {
    const int g = (team[slot] == LR_TEAM ? 0 : 1) + (lr_wide(systems[order[slot]]->s) ? 2 : 0);
    const int j = Hd[g].n_sys++;
    Hf[g].n_sys = Hd[g].n_sys;
    Hd[g].sys[j] = Md.sys[slot];
    Hd[g].sys[j].block_start = h_blocks[g];
    Hf[g].sys[j] = Hd[g].sys[j];
    Hd[g].wk[j] = Md.wk[slot];
    Hf[g].wk[j] = Mf.wk[slot];
    Hd[g].args[j] = Md.args[slot];
    Hf[g].args[j] = Mf.args[slot];
    h_blocks[g] += Md.sys[slot].W;
    h_lds[g] = std::max(h_lds[g], Md.sys[slot].lds_ham);
  }
  const Sys* hsys[LR_MAX_SYSTEMS];  // host copies of the systems, in slot order
  for (int slot = 0; slot < n_sys; ++slot) hsys[slot] = &systems[order[slot]]->s;
  hipStream_t st = (hipStream_t)stream;
  bool any_tau = false;
  int* acts[LR_MAX_SYSTEMS];
  for (int slot = 0; slot < n_sys; ++slot) {
    acts[slot] = Md.args[slot].st.act;
    any_tau = any_tau || Md.args[slot].mode == LR_FIXED_TAU;
  }
  for (int it = 0; it < rounds; ++it) {
    if (prec == JQMC_PRECISION_MIXED_FP32) {
      if (h_blocks[0]) k_lrdmc_ham_multi<float, LR_TEAM, false><<<(unsigned)h_blocks[0], LR_TEAM * WAVE, h_lds[0], st>>>(Hf[0], it);
      if (h_blocks[1]) k_lrdmc_ham_multi<float, 1, false><<<(unsigned)h_blocks[1], WAVE, h_lds[1], st>>>(Hf[1], it);
      if (h_blocks[2]) k_lrdmc_ham_multi<float, LR_TEAM, true><<<(unsigned)h_blocks[2], LR_TEAM * WAVE, h_lds[2], st>>>(Hf[2], it);
      if (h_blocks[3]) k_lrdmc_ham_multi<float, 1, true><<<(unsigned)h_blocks[3], WAVE, h_lds[3], st>>>(Hf[3], it);
      k_lrdmc_move_multi<float><<<(unsigned)blocks, WAVE, lds_move, st>>>(Mf, it);
      lr_multi_jfield<float>(Mf, hsys, it, st);
    } else {
      if (h_blocks[0]) k_lrdmc_ham_multi<double, LR_TEAM, false><<<(unsigned)h_blocks[0], LR_TEAM * WAVE, h_lds[0], st>>>(Hd[0], it);
      if (h_blocks[1]) k_lrdmc_ham_multi<double, 1, false><<<(unsigned)h_blocks[1], WAVE, h_lds[1], st>>>(Hd[1], it);
      if (h_blocks[2]) k_lrdmc_ham_multi<double, LR_TEAM, true><<<(unsigned)h_blocks[2], LR_TEAM * WAVE, h_lds[2], st>>>(Hd[2], it);
      if (h_blocks[3]) k_lrdmc_ham_multi<double, 1, true><<<(unsigned)h_blocks[3], WAVE, h_lds[3], st>>>(Hd[3], it);
      k_lrdmc_move_multi<double><<<(unsigned)blocks, WAVE, lds_move, st>>>(Md, it);
      lr_multi_jfield<double>(Md, hsys, it, st);
    }
    // systems in GFMC_n mode count no walkers: the poll stops when the GFMC_t ones are done and
    // every GFMC_n system has had its n_steps
    if (any_tau && (it % LR_POLL == LR_POLL - 1 || it == rounds - 1)) {
      HIP_TRY(hipGetLastError());
      int fixed_left = 0;
      for (int slot = 0; slot < n_sys; ++slot)
        if (Md.args[slot].mode != LR_FIXED_TAU && Md.args[slot].n_steps > it + 1) fixed_left = 1;
      bool done = false;
      int* tau_acts[LR_MAX_SYSTEMS];
      int n_tau = 0;
      for (int slot = 0; slot < n_sys; ++slot)
        if (Md.args[slot].mode == LR_FIXED_TAU) tau_acts[n_tau++] = acts[slot];
      const int rc = lr_poll(tau_acts, n_tau, it, systems[0]->host_flags, st, &done);
      if (rc != JQMC_OK) return rc;
      if (done && !fixed_left) break;
    }
  }
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_lrdmc_branching_workspace_bytes(int64_t W, int64_t* bytes) {
  if (!bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  *bytes = W * (int64_t)sizeof(double);
  return JQMC_OK;
}

int jqmc_lrdmc_branching(int64_t W, const double* w_L, const double* V_diag, const double* e_L, double E_scf,
                         double zeta, int64_t* chosen, double* sums, void* workspace, int64_t workspace_bytes,
                         void* stream) {
  if (!w_L || !chosen) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  if (!(zeta >= 0.0 && zeta < 1.0)) return fail(JQMC_ERR_INVALID_ARGUMENT, "zeta must lie in [0, 1)");
  if ((V_diag == nullptr) != (e_L == nullptr))
    return fail(JQMC_ERR_INVALID_ARGUMENT, "V_diag and e_L must be given together");
  if (!workspace || workspace_bytes < W * (int64_t)sizeof(double))
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)(W * sizeof(double)));
  k_comb<<<1, COMB_THREADS, 0, (hipStream_t)stream>>>(W, w_L, V_diag, e_L, E_scf, zeta, chosen, sums,
                                                      static_cast<double*>(workspace));
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

}  // extern "C"

// Scratch of jqmc_lrdmc_branching_static: the comb's W-double cumulative-sum buffer as a file-scope
// device array (8 MiB on every GPU the library runs on, held until the process ends).  Host code
// never passes the array's name to a kernel: each device's address is taken once with
// hipGetSymbolAddress, with that device current, and kept in g_scratch_ptr.
static __device__ __align__(256) double scratch[JQMC_BRANCHING_STATIC_MAX_W];

namespace {
constexpr int64_t kBranchStaticMaxW = JQMC_BRANCHING_STATIC_MAX_W;
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;

int scratch_on_current_device(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return JQMC_OK;
}
}  // namespace

extern "C" {

int jqmc_lrdmc_branching_static(int64_t W, const double* w_L, const double* V_diag, const double* e_L, double E_scf,
                                double zeta, int64_t* chosen, double* sums, void* stream) {
  if (!w_L || !chosen) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  if (W > kBranchStaticMaxW)
    return fail(JQMC_ERR_SHAPE, "W = %lld exceeds the static branching scratch of %lld walkers", (long long)W,
                (long long)kBranchStaticMaxW);
  if (!(zeta >= 0.0 && zeta < 1.0)) return fail(JQMC_ERR_INVALID_ARGUMENT, "zeta must lie in [0, 1)");
  if ((V_diag == nullptr) != (e_L == nullptr))
    return fail(JQMC_ERR_INVALID_ARGUMENT, "V_diag and e_L must be given together");
  double* cum = nullptr;
  const int rc = scratch_on_current_device(&cum);
  if (rc != JQMC_OK) return rc;
  // every entry of cum the comb reads is written by the same launch first: nothing to clear
  k_comb<<<1, COMB_THREADS, 0, (hipStream_t)stream>>>(W, w_L, V_diag, e_L, E_scf, zeta, chosen, sums, cum);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_walkers_reconfigure_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (W < 0) return fail(JQMC_ERR_SHAPE, "W < 0");
  *bytes = W * state_rows(sys->s, nullptr).rec;
  return JQMC_OK;
}

int jqmc_walkers_reconfigure(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, const int64_t* chosen,
                             void* workspace, int64_t workspace_bytes, void* stream) {
  if (!sys || !chosen) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (!wk->logabsdet || !wk->sign || !wk->jastrow) return fail(JQMC_ERR_INVALID_ARGUMENT, "scalar outputs are NULL");
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  if (W > 0x7FFFFFFFll) return fail(JQMC_ERR_SHAPE, "too many walkers for one reconfiguration launch");
  int64_t need = 0;
  jqmc_walkers_reconfigure_workspace_bytes(sys, W, &need);
  if (!workspace || workspace_bytes < need)
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)need);
  hipStream_t st = (hipStream_t)stream;
  const StateRows R = state_rows(s, wk);
  char* stage = static_cast<char*>(workspace);
  k_state_gather<<<(unsigned)W, 256, 0, st>>>(W, R, chosen, stage, 0);
  k_state_gather<<<(unsigned)W, 256, 0, st>>>(W, R, chosen, stage, 1);
  HIP_TRY(hipGetLastError());
  // G^-1 rebuilt from the positions, as the reference does after branching (jqmc_gfmc.py:6314-6319),
  // and J from the copied caches (the projections move electrons without updating J)
  rc = jqmc_walkers_refresh_orbitals(sys, W, wk, stream);
  if (rc != JQMC_OK) return rc;
  if (s.precision == JQMC_PRECISION_MIXED_FP32)
    k_init_jastrow<float><<<grid_for(W, 4), 4 * WAVE, 0, st>>>(s, W, to_walk<float>(wk), 4);
  else
    k_init_jastrow<double><<<grid_for(W, 4), 4 * WAVE, 0, st>>>(s, W, to_walk<double>(wk), 4);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ parameter log-derivatives

namespace {

ParamLayout param_layout(const Sys& s, int32_t blocks) {
  ParamLayout L{-1, -1, -1, -1, 0};
  if ((blocks & JQMC_PARAM_J1) && s.j1_type) L.j1 = L.total++;
  if ((blocks & JQMC_PARAM_J2) && s.j2_type) L.j2 = L.total++;
  if ((blocks & JQMC_PARAM_J3) && s.j3_on) {
    L.j3 = L.total;
    L.total += (int64_t)s.n_ao * (s.n_ao + 1);
  }
  if (blocks & JQMC_PARAM_LAMBDA) {
    L.lam = L.total;
    L.total += (int64_t)s.n_orb * (s.n_orb + s.n_unp);
  }
  return L;
}

template <typename T>
int launch_dlnpsi(const Sys& s, int64_t W, const jqmc_walkers* wk, const ParamLayout& lay, double* O, int64_t ld,
                  void* workspace, hipStream_t st) {
  WalkT<T> k = to_walk<T>(wk);
  const int ne = s.n_el, na = s.n_ao, no = s.n_orb, nu = s.n_up, nd = s.n_dn;
  double* Z = static_cast<double*>(workspace);
  T* S1 = reinterpret_cast<T*>(Z + W * nu * no);
  T* S2 = S1 + W * ne * na;
  constexpr int64_t kChunk = 32768;  // walkers per launch along grid z
  if (lay.lam >= 0) {
    k_ginv_phi<T><<<grid_for(W * nu * no, 256), 256, 0, st>>>(s, W, k, Z);
    if (nd > 0) {
      for (int64_t w0 = 0; w0 < W; w0 += kChunk) {
        const int64_t nw = (W - w0) < kChunk ? (W - w0) : kChunk;
        dim3 grid(grid_for(no, RO_TILE), grid_for(no, RO_TILE), (unsigned)nw);
        k_rank_outer<double, double><<<grid, 256, 0, st>>>(nd, no, no, Z + w0 * nu * no, (int64_t)nu * no, no,
                                                      k.orb + w0 * ne * no + (int64_t)nu * no, (int64_t)ne * no, no,
                                                      O + w0 * ld, ld, no + (nu - nd), lay.lam);
      }
    }
  }
  if (lay.j3 >= 0) {
    k_init_jsums<T><<<grid_for(W * na, 256), 256, 0, st>>>(s, W, k, S1, S2);
    for (int64_t w0 = 0; w0 < W; w0 += kChunk) {
      const int64_t nw = (W - w0) < kChunk ? (W - w0) : kChunk;
      dim3 grid(grid_for(na, RO_TILE), grid_for(na, RO_TILE), (unsigned)nw);
      k_rank_outer<T, T><<<grid, 256, 0, st>>>(ne, na, na, k.ao + w0 * ne * na, (int64_t)ne * na, na,
                                               S1 + w0 * ne * na, (int64_t)ne * na, na, O + w0 * ld, ld, na + 1,
                                               lay.j3);
    }
  }
  k_dlnpsi_small<T><<<grid_for(W, 4), 4 * WAVE, 0, st>>>(s, W, k, Z, lay, O, ld, 4);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

}  // namespace

extern "C" {

int jqmc_param_layout(const jqmc_system* sys, int32_t blocks, int64_t* offsets, int64_t* n_params) {
  if (!sys || !offsets || !n_params) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const ParamLayout L = param_layout(sys->s, blocks);
  offsets[0] = L.j1;
  offsets[1] = L.j2;
  offsets[2] = L.j3;
  offsets[3] = L.lam;
  *n_params = L.total;
  return JQMC_OK;
}

int jqmc_dln_psi_dparams_workspace_bytes(const jqmc_system* sys, int64_t W, int64_t* bytes) {
  if (!sys || !bytes) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  *bytes = W * (int64_t)s.n_up * s.n_orb * (int64_t)sizeof(double) + 2 * W * (int64_t)s.n_el * s.n_ao * cache_bytes(s);
  return JQMC_OK;
}

int jqmc_dln_psi_dparams(const jqmc_system* sys, int64_t W, const jqmc_walkers* wk, int32_t blocks, double* O,
                         int64_t ld, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!sys || !O) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  const Sys& s = sys->s;
  int rc = check_walkers(s, wk, true);
  if (rc != JQMC_OK) return rc;
  if (W <= 0) return W == 0 ? JQMC_OK : fail(JQMC_ERR_SHAPE, "W < 0");
  const ParamLayout lay = param_layout(s, blocks);
  if (ld < lay.total) return fail(JQMC_ERR_SHAPE, "ld (%lld) < number of parameters (%lld)", (long long)ld,
                                  (long long)lay.total);
  int64_t need = 0;
  jqmc_dln_psi_dparams_workspace_bytes(sys, W, &need);
  if (!workspace || workspace_bytes < need)
    return fail(JQMC_ERR_WORKSPACE, "workspace of %lld bytes required", (long long)need);
  if (lay.total == 0) return JQMC_OK;
  hipStream_t st = (hipStream_t)stream;
  return s.precision == JQMC_PRECISION_MIXED_FP32 ? launch_dlnpsi<float>(s, W, wk, lay, O, ld, workspace, st)
                                                  : launch_dlnpsi<double>(s, W, wk, lay, O, ld, workspace, st);
}

// ------------------------------------------------------------------ stochastic reconfiguration

int jqmc_sr_colstats(const double* O, int64_t S, int64_t P, int64_t ld, const double* wn, double* mean, double* var,
                     void* stream) {
  if (!O || !wn || !mean || !var) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (S <= 0 || P <= 0 || ld < P) return fail(JQMC_ERR_SHAPE, "need S > 0, P > 0, ld >= P");
  k_sr_colstats<<<grid_for(P, 256), 256, 0, (hipStream_t)stream>>>(O, S, P, ld, wn, mean, var);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_sr_project(const double* O, int64_t S, int64_t P, int64_t ld, const double* mean, const double* rs,
                    const double* cs, const double* x, double* out, void* stream) {
  if (!O || !mean || !rs || !cs || !x || !out) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (S <= 0 || P <= 0 || ld < P) return fail(JQMC_ERR_SHAPE, "need S > 0, P > 0, ld >= P");
  k_sr_project<<<grid_for(P, 256), 256, 0, (hipStream_t)stream>>>(O, S, P, ld, mean, rs, cs, x, out);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_sr_gram(const double* M, int64_t n, int64_t m, int64_t ldm, const double* a, const double* b, const double* u,
                 const double* v, double* G, int64_t ldg, void* stream) {
  if (!M || !u || !v || !G) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n <= 0 || m <= 0 || ldm < m || ldg < n) return fail(JQMC_ERR_SHAPE, "need n > 0, m > 0, ldm >= m, ldg >= n");
  const int64_t nt = (n + SRG_TILE - 1) / SRG_TILE;
  const int64_t blocks = nt * (nt + 1) / 2;
  if (blocks > 0x7fffffff) return fail(JQMC_ERR_SHAPE, "Gram too large for one launch");
  k_sr_gram<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(M, n, m, ldm, a, b, u, v, G, ldg, nt);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

int jqmc_sr_gram_f32(const double* M, int64_t n, int64_t m, int64_t ldm, const double* a, const double* b,
                     const double* u, const double* v, double* G, int64_t ldg, void* stream) {
  if (!M || !u || !v || !G) return fail(JQMC_ERR_INVALID_ARGUMENT, "NULL argument");
  if (n <= 0 || m <= 0 || ldm < m || ldg < n) return fail(JQMC_ERR_SHAPE, "need n > 0, m > 0, ldm >= m, ldg >= n");
  const int64_t nt = (n + SRF_TILE - 1) / SRF_TILE;
  const int64_t blocks = nt * (nt + 1) / 2;
  if (blocks > 0x7fffffff) return fail(JQMC_ERR_SHAPE, "Gram too large for one launch");
  k_sr_gram_f32<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(M, n, m, ldm, a, b, u, v, G, ldg, nt);
  HIP_TRY(hipGetLastError());
  return JQMC_OK;
}

}  // extern "C"
