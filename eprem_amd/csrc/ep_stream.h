// ep_stream.h -- the streaming kernels of the EP update (DiffuseStreamData,
// src/energeticParticles.c:2237-2490) and the focusing row update they
// share with the focusing kernels, as templates: ep_kernels.hip uses the
// focusing macro and the DPP helpers, and the four ep_stream_m*.hip
// translation units instantiate the passes for their range of mu counts.
#pragma once
#include <float.h>
#include <math.h>

#include "ep_fastmath.h"
#include "ep_internal.h"
#include "ep_launch.h"

#ifndef EP_CHECK_LAUNCH
#define EP_CHECK_LAUNCH() \
  do {                     \
    hipError_t e_ = hipGetLastError(); \
    if (e_ != hipSuccess) return e_;   \
  } while (0)
#endif

namespace ep {

// Per-mu and per-energy tables (coefficients, grids) read the same on every
// lane and are fixed while a kernel runs.  Read through the constant address
// space they become scalar loads even after an atomic earlier in the kernel
// (the slab queue), which otherwise makes the compiler fetch them with
// vector loads inside the subcycle loop.
typedef const __attribute__((address_space(4))) double cdouble;
__device__ __forceinline__ const cdouble *as_const_space(const double *p) { return (const cdouble *)p; }

// The same pointer, re-derived where the compiler cannot see through it:
// taken once per subcycle, it keeps the coefficient loads inside the
// subcycle loop (scalar loads into the same few registers) instead of
// letting them be hoisted out of it into 30 live register pairs.
__device__ __forceinline__ const cdouble *opaque(const cdouble *p) {
  uint64_t v = (uint64_t)p;
  asm volatile("" : "+s"(v));
  return (const cdouble *)v;
}

// ------------------------------------------------------------------------
// Cross-lane shifts along the stream (DPP wave_shr:1 / wave_shl:1 move a
// 32-bit value one lane up/down the whole 64-lane wave; lanes without a
// source keep `old`).
// ------------------------------------------------------------------------
__device__ __forceinline__ double dpp_shr1(double v, double old) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32), 0x138,
                                             0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double dpp_shl1(double v, double old) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32), 0x130,
                                             0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ float dpp_shr1(float v, float old) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, v), 0x138, 0xF,
                                                               0xF, false));
}

__device__ __forceinline__ float dpp_shl1(float v, float old) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, v), 0x130, 0xF,
                                                               0xF, false));
}

// The subcycles of AdiabaticFocusing on one mu row f[M] (the upwind
// operator, src/energeticParticles.c:1806-1896).  The half-mesh velocity at
// the face below cell m is split by the side the flux is taken from
// (split_face: vh_ = min(va, 0), vl_ = va - vh_, so one of them is an
// exact zero).  An interior face then carries fma(vh_, f[m], vl_ * f[m-1]): one
// product plus an exact zero, which is the reference's v 0.5 ((1 - c) f[m]
// + (1 + c) f[m-1]) (:1830-1836): equal to v * f_upwind for finite f and
// NaN when either cell is NaN or Inf, with no select.  The wall rows
// (cells 0 and M-1) take the upwind cell alone (:1850-1885).
// FASTF (the fused precision modes): the face velocities are scaled by
// dt_sub / dmu once, so that a subcycle is f[m] -= F_hi - F_lo (the
// reference forms (F_hi - F_lo) / dmu and then f - dt_sub * that).
#define EP_FOCUS_ROW(T, M, D, a, b, f, nsub, dtsub, bad, bad_inf, FASTF)           \
  do {                                                                             \
    const T dmu_i_ = (T)(1.0 / (D).dmu);                                           \
    const T dts_ = (T)(dtsub);                                                     \
    const T sc_ = (FASTF) ? dts_ * dmu_i_ : (T)1;                                  \
    T vl_[(M) + 1], vh_[(M) + 1];                                                  \
    T vw0_ = 0, vw1_ = 0; /* face velocities of the wall rows: faces 1, M-1 */     \
    double vprev_ = 0.0;                                                           \
    _Pragma("unroll") for (int m = 0; m < (M); ++m) {                              \
      const double mu = as_const_space((D).mugrid)[m];                             \
      const double vel = 0.5 * (1.0 - mu * mu) * ((a) + mu * (b));                 \
      if (m > 0) {                                                                 \
        T va_ = (T)(0.5 * (vel + vprev_));                                         \
        if (FASTF) va_ = va_ * sc_;                                                \
        split_face(va_, vl_[m], vh_[m]);                                           \
        if (m == 1) vw0_ = va_;                                                    \
        if (m == (M) - 1) vw1_ = va_;                                              \
      }                                                                            \
      vprev_ = vel;                                                                \
    }                                                                              \
    for (int it_ = 0; it_ < (nsub); ++it_) {                                       \
      T flux_lo = 0; /* zero-flux wall at mu = -1 */                               \
      _Pragma("unroll") for (int m = 0; m < (M); ++m) {                            \
        T flux_hi = 0, next_lo = 0; /* zero-flux wall at mu = +1 */                \
        if (m + 1 < (M)) {                                                         \
          const T p_ = fma(vh_[m + 1], f[m + 1], vl_[m + 1] * f[m]);              \
          flux_hi = p_;                                                            \
          next_lo = p_;                                                            \
          if (m == 0) flux_hi = vw0_ * (signbit(vw0_) ? f[1] : f[0]);             \
          if (m + 2 == (M))                                                        \
            next_lo = vw1_ * (signbit(vw1_) ? f[(M) - 1] : f[(M) > 1 ? (M) - 2 : 0]); \
        }                                                                          \
        if (FASTF) {                                                               \
          f[m] = f[m] - (flux_hi - flux_lo);                                       \
        } else {                                                                   \
          const T dfl_ = (flux_hi - flux_lo) * dmu_i_;                             \
          f[m] = f[m] - dts_ * dfl_;                                               \
        }                                                                          \
        flux_lo = next_lo;                                                         \
      }                                                                            \
    }                                                                              \
    _Pragma("unroll") for (int m = 0; m < (M); ++m) {                              \
      if ((bad) < 0 && !isfinite(f[m])) {                                          \
        (bad) = m;                                                                 \
        (bad_inf) = isinf(f[m]) ? 1 : 0;                                           \
      }                                                                            \
    }                                                                              \
  } while (0)

// ------------------------------------------------------------------------
// k_stream_pass: DiffuseStreamData (src/energeticParticles.c:2352-2430).
//
// Block = one chunk of one (stream, energy) slab; a lane owns one kept
// node and keeps its M mu values in registers.  Each launch advances the
// slab by up to `halo` Euler subcycles: the chunk loads `halo` extra nodes
// on each side, the upwind neighbour comes from the adjacent lane (DPP) or,
// across waves, through a small double-buffered LDS edge array, and after
// the subcycles only the interior lanes (whose dependency cone stayed
// inside the chunk) are written, to the other f buffer.
// MPOS = first mu index with mugrid >= 0 (upwind direction).
// ------------------------------------------------------------------------
// FAST: fused multiply-adds, and the relaxation as q*f + (1-q)*iso with
// (1-q)*iso formed once per lane and subcycle: 6 instead of 9 VALU
// operations per cell and subcycle, at a rounding difference from the
// reference's operation order (not bit-identical; ~1e-15 relative).
template <int M, int MPOS, typename T, bool FAST, bool FUSE>
__global__ __launch_bounds__(512) void k_stream_pass(Dev D, int pass, int flip, int do_seed, int substep) {
  if (substep_off(D)) return;
  const int c = blockIdx.x, e = blockIdx.y, s = blockIdx.z;
  const int task = s * D.E + e;
  const int nsteps = D.nsteps[task];
  const int h = D.halo[task];
  const int done = pass * h;
  if (done >= nsteps) return;
  const int nsub = min(h, nsteps - done);
  const bool last = (done + nsub == nsteps);
  const int nk = D.nkept[s];
  const int W = blockDim.x;
  // a stream that fits in one chunk needs no halo (its ends are held):
  // all its subcycles run in one pass over the whole stream
  const int hs = (D.stream_whole && nk <= W) ? 0 : h;
  const int I = W - 2 * hs;
  if (c * I >= nk) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = W >> 6;
  const int k = c * I - hs + tid;
  const bool valid = (k >= 0) && (k < nk);
  const size_t on = (size_t)s * D.Npad;
  const int node = valid ? (D.thin[s] ? D.list[on + k] : k) : 0;
  // flip = pass & 1 in a step; the tile-size search runs every pass with
  // flip = 0 (read the current slab, write the spare one), which leaves the
  // state untouched
  const int cur = D.loc[task] ^ flip;
  constexpr uint32_t B = sizeof(T);
  const uint32_t slab_bytes = (uint32_t)(M * D.Npad * B), row = (uint32_t)D.Npad * B;
  const rsrc_t src = slab_rsrc((const double *)fslab<T>(D, cur, task), slab_bytes);
  const rsrc_t dst = slab_rsrc((const double *)fslab<T>(D, cur ^ 1, task), slab_bytes);
  const uint32_t voff = (uint32_t)node * B;

  T f[M];
  T sd = 0;
  if (valid && do_seed) sd = (T)seed_value(D, D.radial[on + node], e);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    T v = 0;
    if (valid) v = buf_ld(src, voff, m * row, T());
    if (do_seed && v < sd) v = sd;
    f[m] = v;
  }
  const size_t ok = (size_t)task * D.Npad;
  const double dsm = valid ? D.dsm[ok + k] : 0.0;
  // the relaxation (the pitch-angle part) runs in FP64 also when f is
  // stored as float
  const double q = valid ? D.q[ok + k] : 1.0;
  const double omq = valid ? 1.0 - D.q[ok + k] : 0.0;
  // Stream ends are held: kept node 0 for mu >= 0, the last one for mu < 0.
  const double dsm_pos = (k == 0) ? 0.0 : dsm;
  const double dsm_neg = (k == nk - 1) ? 0.0 : dsm;
  const bool bc = D.useBC && valid && (k < 3);
  const T bcv = bc ? (T)D.seed3[(size_t)task * 3 + k] : (T)0;
  const double invM = 1.0 / M;
  // Edge exchange between the waves of the block, double-buffered by
  // subcycle parity.  up[r] holds the mu >= 0 values lane 0 of wave r
  // takes as its upstream neighbour: lane 63 of wave r-1 writes row r, and
  // lane 0 of wave 0 writes its own values into row 0 (the chunk's first
  // lane keeps its value, as DPP leaves a lane without a source alone).
  // dn[r] holds the mu < 0 values lane 63 of wave r-1 takes as its
  // downstream neighbour: lane 0 of wave r writes row r, and lane 63 of the
  // last wave writes its own values into row nw.
  constexpr int NP = M - MPOS;
  __shared__ T up[2][17][NP > 0 ? NP : 1];
  __shared__ T dn[2][17][MPOS > 0 ? MPOS : 1];
  // coef[m] = ds_i_multiplier * del_fac[m]: the product the reference forms
  // every subcycle is loop-invariant, so it is formed once per pass.
  T coef[M];
  {
    const cdouble *del = as_const_space(D.del_fac + (size_t)task * M);
#pragma unroll
    for (int m = 0; m < M; ++m) coef[m] = (T)(((m >= MPOS) ? dsm_pos : dsm_neg) * del[m]);
  }
  const int up_w = (lane == 63) ? w + 1 : ((lane == 0 && w == 0) ? 0 : -1);
  const int dn_w = (lane == 0) ? w : ((lane == 63 && w == nw - 1) ? nw : -1);

  for (int j = 0; j < nsub; ++j) {
    if (bc) {
#pragma unroll
      for (int m = 0; m < M; ++m) f[m] = bcv;
    }
    const int b = j & 1;
    if (up_w >= 0) {
#pragma unroll
      for (int m = MPOS; m < M; ++m) up[b][up_w][m - MPOS] = f[m];
    }
    if (dn_w >= 0) {
#pragma unroll
      for (int m = 0; m < MPOS; ++m) dn[b][dn_w][m] = f[m];
    }
    __syncthreads();
    // Only lane 0 (mu >= 0) and lane 63 (mu < 0) use these values; the
    // other lanes read the same rows harmlessly.
    T eu[NP > 0 ? NP : 1], ed[MPOS > 0 ? MPOS : 1];
#pragma unroll
    for (int m = MPOS; m < M; ++m) eu[m - MPOS] = up[b][w][m - MPOS];
#pragma unroll
    for (int m = 0; m < MPOS; ++m) ed[m] = dn[b][w + 1][m];
    double sum = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const T fm = f[m];
      T fn;
      if (m >= MPOS) {
        const T nb = dpp_shr1(fm, eu[m - MPOS]);
        fn = FAST ? fma(coef[m], nb - fm, fm) : fm + coef[m] * (nb - fm);
      } else {
        const T nb = dpp_shl1(fm, ed[m]);
        fn = FAST ? fma(coef[m], fm - nb, fm) : fm + coef[m] * (fm - nb);
      }
      f[m] = fn;
      sum = sum + (double)fn;
    }
    const double iso = invM * sum;
    if (FAST) {
      const double r = omq * iso;
#pragma unroll
      for (int m = 0; m < M; ++m) f[m] = (T)fma(q, (double)f[m], r);
    } else {
#pragma unroll
      for (int m = 0; m < M; ++m) f[m] = (T)(iso + ((double)f[m] - iso) * q);
    }
  }

  if (valid && tid >= hs && tid < hs + I) {
    if (last) {
      int bad = -1, bad_inf = 0;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (bad < 0 && !isfinite(f[m])) {
          bad = m;
          bad_inf = isinf(f[m]) ? 1 : 0;
        }
        if (f[m] < FloorT<T>::v) f[m] = FloorT<T>::v;
      }
      if (bad >= 0) note_error(D, err_key(substep, 0, s, 1, e, bad, node, bad_inf));
      // AdiabaticFocusing of the node, which in the reference follows the
      // stream's DiffuseStreamData, on the values still in registers (full
      // streams only: thinned nodes are filled after the pass; nodes with
      // many subcycles go to k_focus_heavy)
      if (FUSE && !D.thin[s] && node >= 1) {
        const size_t i = on + node;
        const int nf = D.nsub_f[i];
        if (nf < kHeavyFocus) {
          const double a = -D.vgrid[e] * D.f_dlnbds[i] - (2.0 / D.vgrid[e]) * D.f_dupar[i];
          const double b2 = D.f_b[i];
          int fbad = -1, finf = 0;
          EP_FOCUS_ROW(T, M, D, a, b2, f, nf, D.dtsub_f[i], fbad, finf, FAST);
#pragma unroll
          for (int m = 0; m < M; ++m)
            if (f[m] < FloorT<T>::v) f[m] = FloorT<T>::v;
          if (fbad >= 0) note_error(D, err_key(substep, 1, s, 0, node, e, fbad, finf));
        }
      }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) buf_st(f[m], dst, voff, m * row);
  }
}

// ------------------------------------------------------------------------
// k_stream_pass2: the same pass with two consecutive kept nodes per lane,
// so a block of L lanes carries a chunk of 2L nodes.  At the same halo
// fraction a chunk twice as wide takes twice the halo, so a slab needs half
// as many passes (half the HBM traffic), and the barrier and edge exchange
// of each subcycle are shared by twice the cells.  Only one of the two
// upwind neighbours of a lane's pair crosses lanes (DPP); the other is the
// lane's own other node.  The streaming coefficient ds_i_multiplier *
// del_fac[m] is the same product as above, formed each subcycle from the
// two factors when a second register copy per node would not fit (FP64,
// M > 16), else once per pass.
// ------------------------------------------------------------------------
template <int M, int MPOS, typename T, bool FAST, bool FUSE>
__global__ __launch_bounds__(512) void k_stream_pass2(Dev D, int pass, int flip, int do_seed, int substep) {
  if (substep_off(D)) return;
  const int c = blockIdx.x, e = blockIdx.y, s = blockIdx.z;
  const int task = s * D.E + e;
  const int nsteps = D.nsteps[task];
  const int h = D.halo[task];
  const int done = pass * h;
  if (done >= nsteps) return;
  const int nsub = min(h, nsteps - done);
  const bool last = (done + nsub == nsteps);
  const int nk = D.nkept[s];
  const int L = blockDim.x, W = 2 * L;
  const int hs = (D.stream_whole && nk <= W) ? 0 : h;  // one chunk: no halo
  const int I = W - 2 * hs;
  if (c * I >= nk) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = L >> 6;
  const int j0 = 2 * tid, j1 = j0 + 1;  // positions in the chunk
  const int k0 = c * I - hs + j0, k1 = k0 + 1;
  const bool v0 = (k0 >= 0) && (k0 < nk), v1 = (k1 >= 0) && (k1 < nk);
  const size_t on = (size_t)s * D.Npad;
  const bool thin = D.thin[s] != 0;
  const int node0 = v0 ? (thin ? D.list[on + k0] : k0) : 0;
  const int node1 = v1 ? (thin ? D.list[on + k1] : k1) : 0;
  const int cur = D.loc[task] ^ flip;
  constexpr uint32_t B = sizeof(T);
  const uint32_t slab_bytes = (uint32_t)(M * D.Npad * B), row = (uint32_t)D.Npad * B;
  const rsrc_t src = slab_rsrc((const double *)fslab<T>(D, cur, task), slab_bytes);
  const rsrc_t dst = slab_rsrc((const double *)fslab<T>(D, cur ^ 1, task), slab_bytes);
  const uint32_t vo0 = (uint32_t)node0 * B, vo1 = (uint32_t)node1 * B;

  T f0[M], f1[M];
  T sd0 = 0, sd1 = 0;
  if (do_seed) {
    if (v0) sd0 = (T)seed_value(D, D.radial[on + node0], e);
    if (v1) sd1 = (T)seed_value(D, D.radial[on + node1], e);
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    T a = 0, b = 0;
    if (v0) a = buf_ld(src, vo0, m * row, T());
    if (v1) b = buf_ld(src, vo1, m * row, T());
    if (do_seed && a < sd0) a = sd0;
    if (do_seed && b < sd1) b = sd1;
    f0[m] = a;
    f1[m] = b;
  }
  const size_t ok = (size_t)task * D.Npad;
  const double dsm0 = v0 ? D.dsm[ok + k0] : 0.0, dsm1 = v1 ? D.dsm[ok + k1] : 0.0;
  // the relaxation runs in FP64 also when f is stored as float
  const double q0 = v0 ? D.q[ok + k0] : 1.0, q1 = v1 ? D.q[ok + k1] : 1.0;
  const double omq0 = v0 ? 1.0 - D.q[ok + k0] : 0.0, omq1 = v1 ? 1.0 - D.q[ok + k1] : 0.0;
  // stream ends are held: kept node 0 for mu >= 0, the last one for mu < 0
  const double pos0 = (k0 == 0) ? 0.0 : dsm0, neg0 = (k0 == nk - 1) ? 0.0 : dsm0;
  const double pos1 = (k1 == 0) ? 0.0 : dsm1, neg1 = (k1 == nk - 1) ? 0.0 : dsm1;
  const bool bc0 = D.useBC && v0 && (k0 < 3), bc1 = D.useBC && v1 && (k1 < 3);
  const T bcv0 = bc0 ? (T)D.seed3[(size_t)task * 3 + k0] : (T)0;
  const T bcv1 = bc1 ? (T)D.seed3[(size_t)task * 3 + (bc1 ? k1 : 0)] : (T)0;
  const double invM = 1.0 / M;
  const double *__restrict__ del = D.del_fac + (size_t)task * M;
  constexpr bool kCoefRegs = (sizeof(T) == 4) || (M <= 16);
  T cf0[kCoefRegs ? M : 1], cf1[kCoefRegs ? M : 1];
  if (kCoefRegs) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      cf0[m] = (T)(((m >= MPOS) ? pos0 : neg0) * del[m]);
      cf1[m] = (T)(((m >= MPOS) ? pos1 : neg1) * del[m]);
    }
  }
  // Edge rows between the waves of the block, double-buffered by subcycle
  // parity, as in k_stream_pass: up[r] = the mu >= 0 values the first node
  // of lane 0 of wave r takes from its left (node 1 of lane 63 of wave
  // r-1; wave 0 lane 0 writes its own node 0), dn[r] = the mu < 0 values
  // node 1 of lane 63 of wave r-1 takes from its right (node 0 of lane 0
  // of wave r; the last wave's lane 63 writes its own node 1).
  constexpr int NP = M - MPOS;
  __shared__ T up[2][9][NP > 0 ? NP : 1];
  __shared__ T dn[2][9][MPOS > 0 ? MPOS : 1];
  for (int jj = 0; jj < nsub; ++jj) {
    if (bc0) {
#pragma unroll
      for (int m = 0; m < M; ++m) f0[m] = bcv0;
    }
    if (bc1) {
#pragma unroll
      for (int m = 0; m < M; ++m) f1[m] = bcv1;
    }
    const int b = jj & 1;
    // one writer per row, each from a fixed register set (no selects)
    if (lane == 63) {
#pragma unroll
      for (int m = MPOS; m < M; ++m) up[b][w + 1][m - MPOS] = f1[m];
    }
    if (w == 0 && lane == 0) {
#pragma unroll
      for (int m = MPOS; m < M; ++m) up[b][0][m - MPOS] = f0[m];
    }
    if (lane == 0) {
#pragma unroll
      for (int m = 0; m < MPOS; ++m) dn[b][w][m] = f0[m];
    }
    if (w == nw - 1 && lane == 63) {
#pragma unroll
      for (int m = 0; m < MPOS; ++m) dn[b][nw][m] = f1[m];
    }
    __syncthreads();
    double sum0 = 0, sum1 = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const T a0 = f0[m], a1 = f1[m];
      const T c0 = kCoefRegs ? cf0[m] : (T)(((m >= MPOS) ? pos0 : neg0) * del[m]);
      const T c1 = kCoefRegs ? cf1[m] : (T)(((m >= MPOS) ? pos1 : neg1) * del[m]);
      T n0, n1;
      if (m >= MPOS) {
        const T nb = dpp_shr1(a1, up[b][w][m - MPOS]);  // node 1 of the lane to the left
        n0 = FAST ? fma(c0, nb - a0, a0) : a0 + c0 * (nb - a0);
        n1 = FAST ? fma(c1, a0 - a1, a1) : a1 + c1 * (a0 - a1);
      } else {
        const T nb = dpp_shl1(a0, dn[b][w + 1][m]);  // node 0 of the lane to the right
        n0 = FAST ? fma(c0, a0 - a1, a0) : a0 + c0 * (a0 - a1);
        n1 = FAST ? fma(c1, a1 - nb, a1) : a1 + c1 * (a1 - nb);
      }
      f0[m] = n0;
      f1[m] = n1;
      sum0 = (m == 0) ? (double)n0 : sum0 + (double)n0;
      sum1 = (m == 0) ? (double)n1 : sum1 + (double)n1;
    }
    const double iso0 = invM * sum0, iso1 = invM * sum1;
    if (FAST) {
      const double r0 = omq0 * iso0, r1 = omq1 * iso1;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        f0[m] = (T)fma(q0, (double)f0[m], r0);
        f1[m] = (T)fma(q1, (double)f1[m], r1);
      }
    } else {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        f0[m] = (T)(iso0 + ((double)f0[m] - iso0) * q0);
        f1[m] = (T)(iso1 + ((double)f1[m] - iso1) * q1);
      }
    }
  }

  const bool st0 = v0 && j0 >= hs && j0 < hs + I, st1 = v1 && j1 >= hs && j1 < hs + I;
  if (last) {
    int bad0 = -1, inf0 = 0, bad1 = -1, inf1 = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      if (bad0 < 0 && !isfinite(f0[m])) {
        bad0 = m;
        inf0 = isinf(f0[m]) ? 1 : 0;
      }
      if (bad1 < 0 && !isfinite(f1[m])) {
        bad1 = m;
        inf1 = isinf(f1[m]) ? 1 : 0;
      }
      if (f0[m] < FloorT<T>::v) f0[m] = FloorT<T>::v;
      if (f1[m] < FloorT<T>::v) f1[m] = FloorT<T>::v;
    }
    if (st0 && bad0 >= 0) note_error(D, err_key(substep, 0, s, 1, e, bad0, node0, inf0));
    if (st1 && bad1 >= 0) note_error(D, err_key(substep, 0, s, 1, e, bad1, node1, inf1));
    // AdiabaticFocusing of both nodes, as in k_stream_pass; lanes without
    // a node to focus run zero subcycles (no divergent row updates)
    if (FUSE && !thin) {
      const double v = D.vgrid[e], two_v = 2.0 / v;
      int nf0 = 0, nf1 = 0;
      double a0 = 0, b0 = 0, a1 = 0, b1 = 0, dt0 = 0, dt1 = 0;
      const bool fz0 = st0 && node0 >= 1, fz1 = st1 && node1 >= 1;
      if (fz0) {
        const size_t i = on + node0;
        nf0 = D.nsub_f[i];
        a0 = -v * D.f_dlnbds[i] - two_v * D.f_dupar[i];
        b0 = D.f_b[i];
        dt0 = D.dtsub_f[i];
      }
      if (fz1) {
        const size_t i = on + node1;
        nf1 = D.nsub_f[i];
        a1 = -v * D.f_dlnbds[i] - two_v * D.f_dupar[i];
        b1 = D.f_b[i];
        dt1 = D.dtsub_f[i];
      }
      const bool h0 = nf0 >= kHeavyFocus, h1 = nf1 >= kHeavyFocus;  // k_focus_heavy
      if (h0) nf0 = 0;
      if (h1) nf1 = 0;
      int fb0 = -1, fi0 = 0, fb1 = -1, fi1 = 0;
      EP_FOCUS_ROW(T, M, D, a0, b0, f0, nf0, dt0, fb0, fi0, FAST);
      EP_FOCUS_ROW(T, M, D, a1, b1, f1, nf1, dt1, fb1, fi1, FAST);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (f0[m] < FloorT<T>::v) f0[m] = FloorT<T>::v;
        if (f1[m] < FloorT<T>::v) f1[m] = FloorT<T>::v;
      }
      if (fz0 && !h0 && fb0 >= 0) note_error(D, err_key(substep, 1, s, 0, node0, e, fb0, fi0));
      if (fz1 && !h1 && fb1 >= 0) note_error(D, err_key(substep, 1, s, 0, node1, e, fb1, fi1));
    }
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    if (st0) buf_st(f0[m], dst, vo0, m * row);
    if (st1) buf_st(f1[m], dst, vo1, m * row);
  }
}

// ------------------------------------------------------------------------
// k_stream_passR: the same pass with R consecutive kept nodes per lane
// (instantiated for R = 3), a block of L lanes carrying a chunk of R L
// nodes.  Against two nodes per lane, one DPP move serves three cells
// instead of two and the chunk is wider at the same halo, for fewer
// registers of headroom (occupancy 2 instead of 3 at M = 30 in FP64); the
// tile-size search decides.  Same arithmetic per cell as k_stream_pass2.
// ------------------------------------------------------------------------
//
// stream_chunk: the work of one block on chunk c of slab (s, e) in pass
// `pass`, reading buffer src_buf and writing buffer dst_buf; shared by the
// one-launch-per-pass kernel and the persistent slab kernel below.
template <int M, int MPOS, typename T, bool FAST, bool FUSE, int R, bool LOG = false>
__device__ __forceinline__ void stream_chunk(const Dev &D, int c, int e, int s, int pass, int src_buf, int dst_buf,
                                             int do_seed, int substep) {
  const int task = s * D.E + e;
  const int nsteps = D.nsteps[task];
  const int h = D.halo[task];
  const int done = pass * h;
  if (done >= nsteps) return;
  const int nsub = min(h, nsteps - done);
  const bool last = (done + nsub == nsteps);
  const int nk = D.nkept[s];
  const int L = blockDim.x, W = R * L;
  const int hs = (D.stream_whole && nk <= W) ? 0 : h;  // one chunk: no halo
  const int I = W - 2 * hs;
  if (c * I >= nk) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = L >> 6;
  const size_t on = (size_t)s * D.Npad;
  const bool thin = D.thin[s] != 0;
  constexpr uint32_t B = sizeof(T);
  const uint32_t slab_bytes = (uint32_t)(M * D.Npad * B), row = (uint32_t)D.Npad * B;
  const rsrc_t src = slab_rsrc((const double *)fslab<T>(D, src_buf, task), slab_bytes);
  const rsrc_t dst = slab_rsrc((const double *)fslab<T>(D, dst_buf, task), slab_bytes);
  int jj_[R], kk[R], nd[R];
  bool vv[R];
  uint32_t vo[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    jj_[r] = R * tid + r;  // position in the chunk
    kk[r] = c * I - hs + jj_[r];
    vv[r] = (kk[r] >= 0) && (kk[r] < nk);
    nd[r] = vv[r] ? (thin ? D.list[on + kk[r]] : kk[r]) : 0;
    vo[r] = (uint32_t)nd[r] * B;
  }
  T f[R][M];
  T sd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) sd[r] = (do_seed && vv[r]) ? (T)seed_value(D, D.radial[on + nd[r]], e) : (T)0;
#pragma unroll
  for (int m = 0; m < M; ++m)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      T a = 0;
      if (vv[r]) a = f_decL<LOG>(buf_ld(src, vo[r], m * row, T()));
      if (do_seed && a < sd[r]) a = sd[r];
      f[r][m] = a;
    }
  const size_t ok = (size_t)task * D.Npad;
  double pos[R], neg[R];
  double q[R];  // the relaxation runs in FP64 also when f is stored as float
  bool bc[R];
  // boundary values of the stream's first three kept nodes (the lanes that
  // hold them read them each subcycle; LDS rather than a register that
  // every lane would carry)
  __shared__ T bcs[3];
  if (tid < 3) bcs[tid] = D.useBC ? (T)D.seed3[(size_t)task * 3 + tid] : (T)0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double dsm = vv[r] ? D.dsm[ok + kk[r]] : 0.0;
    q[r] = vv[r] ? D.q[ok + kk[r]] : 1.0;
    // stream ends are held: kept node 0 for mu >= 0, the last one for mu < 0
    pos[r] = (kk[r] == 0) ? 0.0 : dsm;
    neg[r] = (kk[r] == nk - 1) ? 0.0 : dsm;
    bc[r] = D.useBC && vv[r] && (kk[r] < 3);
  }
  __syncthreads();
  const double invM = 1.0 / M;
  const cdouble *del = as_const_space(D.del_fac + (size_t)task * M);
  // edge rows between the waves of the block, double-buffered by subcycle
  // parity: up[r] = the mu >= 0 values the first node of lane 0 of wave r
  // takes from its left (last node of lane 63 of wave r-1; wave 0 lane 0
  // writes its own first node), dn[r] = the mu < 0 values the last node of
  // lane 63 of wave r-1 takes from its right (first node of lane 0 of wave
  // r; the last wave's lane 63 writes its own last node)
  constexpr int NP = M - MPOS;
  __shared__ T up[2][9][NP > 0 ? NP : 1];
  __shared__ T dn[2][9][MPOS > 0 ? MPOS : 1];
  // The streaming coefficients ds_i_multiplier * del_fac[m] of the lane's
  // nodes are formed once per pass where they fit beside the rows: up to 40
  // values (two nodes of up to 20 mu cells), and only in the form
  // with focusing in the last pass, whose focusing sets its register count
  // anyway (without it the copy would cost an occupancy step).  Else they
  // are formed each subcycle; the same product either way.
  constexpr bool kCoefRegs = FUSE && R * M <= 40;
  T coef[kCoefRegs ? R : 1][kCoefRegs ? M : 1];
  if constexpr (kCoefRegs) {
    const double *delg = D.del_fac + (size_t)task * M;
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
      for (int r = 0; r < R; ++r) coef[r][m] = (T)(((m >= MPOS) ? pos[r] : neg[r]) * delg[m]);
  }
  for (int jj = 0; jj < nsub; ++jj) {
    const cdouble *delj = opaque(del);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (bc[r]) {
        const T bv = bcs[kk[r]];
#pragma unroll
        for (int m = 0; m < M; ++m) f[r][m] = bv;
      }
    const int b = jj & 1;
    if (lane == 63) {
#pragma unroll
      for (int m = MPOS; m < M; ++m) up[b][w + 1][m - MPOS] = f[R - 1][m];
    }
    if (w == 0 && lane == 0) {
#pragma unroll
      for (int m = MPOS; m < M; ++m) up[b][0][m - MPOS] = f[0][m];
    }
    if (lane == 0) {
#pragma unroll
      for (int m = 0; m < MPOS; ++m) dn[b][w][m] = f[0][m];
    }
    if (w == nw - 1 && lane == 63) {
#pragma unroll
      for (int m = 0; m < MPOS; ++m) dn[b][nw][m] = f[R - 1][m];
    }
    __syncthreads();
    double sum[R];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      T a[R], cf[R], nn[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r] = f[r][m];
        if constexpr (kCoefRegs)
          cf[r] = coef[r][m];
        else
          cf[r] = (T)(((m >= MPOS) ? pos[r] : neg[r]) * delj[m]);
      }
      if (m >= MPOS) {
        const T nb = dpp_shr1(a[R - 1], up[b][w][m - MPOS]);  // last node of the lane to the left
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const T left = (r == 0) ? nb : a[r > 0 ? r - 1 : 0];
          nn[r] = FAST ? fma(cf[r], left - a[r], a[r]) : a[r] + cf[r] * (left - a[r]);
        }
      } else {
        const T nb = dpp_shl1(a[0], dn[b][w + 1][m]);  // first node of the lane to the right
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const T right = (r == R - 1) ? nb : a[r < R - 1 ? r + 1 : 0];
          nn[r] = FAST ? fma(cf[r], a[r] - right, a[r]) : a[r] + cf[r] * (a[r] - right);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        f[r][m] = nn[r];
        sum[r] = (m == 0) ? (double)nn[r] : sum[r] + (double)nn[r];
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double iso = invM * sum[r];
      if (FAST) {
        const double rr = (1.0 - q[r]) * iso;  // 1 - q: 0 where q = 1 (nodes outside the stream)
#pragma unroll
        for (int m = 0; m < M; ++m) f[r][m] = (T)fma(q[r], (double)f[r][m], rr);
      } else {
#pragma unroll
        for (int m = 0; m < M; ++m) f[r][m] = (T)(iso + ((double)f[r][m] - iso) * q[r]);
      }
    }
  }

  bool st[R];
#pragma unroll
  for (int r = 0; r < R; ++r) st[r] = vv[r] && jj_[r] >= hs && jj_[r] < hs + I;
  if (last) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int bad = -1, inf = 0;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (bad < 0 && !isfinite(f[r][m])) {
          bad = m;
          inf = isinf(f[r][m]) ? 1 : 0;
        }
        if (f[r][m] < FloorT<T>::v) f[r][m] = FloorT<T>::v;
      }
      if (st[r] && bad >= 0) note_error(D, err_key(substep, 0, s, 1, e, bad, nd[r], inf));
    }
    // AdiabaticFocusing of the lane's nodes, as in k_stream_pass2
    if (FUSE && !thin) {
      const double v = D.vgrid[e], two_v = 2.0 / v;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int nf = 0;
        double a = 0, bq = 0, dt = 0;
        const bool fz = st[r] && nd[r] >= 1;
        if (fz) {
          const size_t i = on + nd[r];
          nf = D.nsub_f[i];
          a = -v * D.f_dlnbds[i] - two_v * D.f_dupar[i];
          bq = D.f_b[i];
          dt = D.dtsub_f[i];
        }
        const bool hv = nf >= kHeavyFocus;  // k_focus_heavy
        if (hv) nf = 0;
        int fb = -1, fi = 0;
        EP_FOCUS_ROW(T, M, D, a, bq, f[r], nf, dt, fb, fi, FAST);
#pragma unroll
        for (int m = 0; m < M; ++m)
          if (f[r][m] < FloorT<T>::v) f[r][m] = FloorT<T>::v;
        if (fz && !hv && fb >= 0) note_error(D, err_key(substep, 1, s, 0, nd[r], e, fb, fi));
      }
    }
  }
#pragma unroll
  for (int m = 0; m < M; ++m)
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (st[r]) buf_st(f_encL<LOG>(f[r][m]), dst, vo[r], m * row);
}

template <int M, int MPOS, typename T, bool FAST, bool FUSE, int R>
__global__ __launch_bounds__(512) void k_stream_passR(Dev D, int pass, int flip, int do_seed, int substep) {
  if (substep_off(D)) return;
  const int cur = D.loc[blockIdx.z * D.E + blockIdx.y] ^ flip;
  stream_chunk<M, MPOS, T, FAST, FUSE, R>(D, blockIdx.x, blockIdx.y, blockIdx.z, pass, cur, cur ^ 1, do_seed,
                                          substep);
}

// ------------------------------------------------------------------------
// k_stream_slab: DiffuseStreamData of whole (stream, energy) slabs in one
// persistent launch per EP substep.  Each block takes the next slab from a
// device queue (highest energies first: they carry the most subcycles, so
// the long slabs start early and the short ones fill in behind them) and
// runs all of the slab's passes and chunks itself, ping-ponging between the
// slab's two buffers; the passes of a slab depend only on that slab, so
// the block's own barrier orders them and no other block is waited for.
// The launch does not depend on the pass or chunk counts, which the
// preparation kernels leave on the device.
// tune = 1 (tile-size search): pass 0 reads the current slab and every pass
// writes the spare one, so the work is that of a step and the state is
// left as it was.  one_stream >= 0: the unbatched form, one launch per
// stream whose block b takes energy E-1-b of that stream.
// ------------------------------------------------------------------------
template <int M, int MPOS, typename T, bool FAST, bool FUSE, int R, bool LOG = false>
__global__ __launch_bounds__(512) void k_stream_slab(Dev D, int do_seed, int substep, int tune, int one_stream) {
  if (substep_off(D)) return;
  __shared__ int item_sh;
  const int S = D.S, E = D.E;
  for (;;) {
    int e, s;
    if (one_stream >= 0) {
      if ((int)blockIdx.x >= E) return;
      e = E - 1 - (int)blockIdx.x;
      s = one_stream;
    } else {
      if (threadIdx.x == 0) item_sh = (int)atomicAdd(&D.diag->queue, 1u);
      __syncthreads();
      const int item = item_sh;
      __syncthreads();
      if (item >= S * E) return;
      e = E - 1 - item / S;
      s = item % S;
    }
    const int task = s * E + e;
    const int passes = D.passes[task], chunks = D.chunks[task], loc = D.loc[task];
    for (int p = 0; p < passes; ++p) {
      const int src = tune ? (p == 0 ? loc : loc ^ 1) : (loc ^ (p & 1));
      const int dst = tune ? (loc ^ 1) : (src ^ 1);
      for (int c = 0; c < chunks; ++c) {
        stream_chunk<M, MPOS, T, FAST, FUSE, R, LOG>(D, c, e, s, p, src, dst, do_seed && p == 0, substep);
        // the next chunk reuses the edge rows in LDS, and the next pass
        // reads what this one stored (the block's own stores)
        __syncthreads();
      }
    }
    if (one_stream >= 0) return;
  }
}

// Resident blocks of a slab-kernel instantiation for a block size (cache:
// the instantiation's own table, one entry per block size).
template <typename K>
inline int resident_blocks(K kernel, int block, int *cache) {
  const int slot = block / 64;
  if (slot < 0 || slot > 8) return 0;
  if (!cache[slot]) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) per_cu = 1;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    cache[slot] = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
  }
  return cache[slot];
}

template <int M, int MPOS, typename T, bool FAST, bool FUSE, int R, bool LOG = false>
inline void launch_slab_kernel(const Dev &D, int do_seed, int substep, int tune, int one_stream, hipStream_t st) {
  const int L = D.stream_block / R;
  auto k = k_stream_slab<M, MPOS, T, FAST, FUSE, R, LOG>;
  static int cache[9] = {0};
  int grid;
  if (one_stream >= 0) {
    grid = D.E;
  } else {
    grid = resident_blocks(k, L, cache);
    if (grid > D.S * D.E) grid = D.S * D.E;
    if (grid < 1) grid = 1;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(L), 0, st, D, do_seed, substep, tune, one_stream);
}

template <int M, int MPOS, bool FUSE, int R>
inline void stream_slab_tfr(const Dev &D, int do_seed, int substep, int tune, int one_stream, hipStream_t st) {
  if constexpr (R == 2 && !FUSE) {
    if (D.logf) {
      launch_slab_kernel<M, MPOS, float, true, false, 2, true>(D, do_seed, substep, tune, one_stream, st);
      return;
    }
  }
  if (D.fp32)
    launch_slab_kernel<M, MPOS, float, true, FUSE, R>(D, do_seed, substep, tune, one_stream, st);
  else if (D.fast)
    launch_slab_kernel<M, MPOS, double, true, FUSE, R>(D, do_seed, substep, tune, one_stream, st);
  else
    launch_slab_kernel<M, MPOS, double, false, FUSE, R>(D, do_seed, substep, tune, one_stream, st);
}

// Slab kernels exist for one, two and three nodes per lane; with the
// log-scaled storage, for two nodes per lane without fused focusing.
template <int M, int MPOS>
inline hipError_t stream_slab_t(const Dev &D, int do_seed, int substep, int fuse, int tune, int one_stream,
                                hipStream_t st) {
  if (D.stream_r < 1 || D.stream_r > 3) return hipErrorInvalidValue;
  if (D.logf && (D.stream_r != 2 || fuse)) return hipErrorInvalidValue;
  if (D.fp32 && D.stream_r == 1) return hipErrorInvalidValue;  // switched off with FP32 storage
  if (D.stream_r == 3) {
    if (fuse)
      stream_slab_tfr<M, MPOS, true, 3>(D, do_seed, substep, tune, one_stream, st);
    else
      stream_slab_tfr<M, MPOS, false, 3>(D, do_seed, substep, tune, one_stream, st);
  } else if (D.stream_r == 1) {
    if (fuse)
      stream_slab_tfr<M, MPOS, true, 1>(D, do_seed, substep, tune, one_stream, st);
    else
      stream_slab_tfr<M, MPOS, false, 1>(D, do_seed, substep, tune, one_stream, st);
  } else {
    if (fuse)
      stream_slab_tfr<M, MPOS, true, 2>(D, do_seed, substep, tune, one_stream, st);
    else
      stream_slab_tfr<M, MPOS, false, 2>(D, do_seed, substep, tune, one_stream, st);
  }
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

#define EP_SLAB_CASE(MM)                                                                          \
  case MM:                                                                                        \
    if (D.mpos == (MM) / 2) return stream_slab_t<MM, (MM) / 2>(D, do_seed, sub, fuse, tune, one, st); \
    if (D.mpos == ((MM) + 1) / 2)                                                                 \
      return stream_slab_t<MM, ((MM) + 1) / 2>(D, do_seed, sub, fuse, tune, one, st);            \
    return hipErrorInvalidValue;

template <int M, int MPOS, bool FUSE>
inline hipError_t stream_pass_tf(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks,
                                 hipStream_t st) {
  if (D.logf) return hipErrorInvalidValue;  // the log-scaled storage runs the slab kernel only
  if (D.fp32 && D.stream_r == 1) return hipErrorInvalidValue;  // switched off with FP32 storage
  const dim3 grid(chunks, D.E, D.S), block(D.stream_block / D.stream_r);
  if (D.stream_r == 3) {
    if (D.fp32)
      hipLaunchKernelGGL((k_stream_passR<M, MPOS, float, true, FUSE, 3>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
    else if (D.fast)
      hipLaunchKernelGGL((k_stream_passR<M, MPOS, double, true, FUSE, 3>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
    else
      hipLaunchKernelGGL((k_stream_passR<M, MPOS, double, false, FUSE, 3>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
  } else if (D.stream_r == 2) {
    if (D.fp32)
      hipLaunchKernelGGL((k_stream_pass2<M, MPOS, float, true, FUSE>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
    else if (D.fast)
      hipLaunchKernelGGL((k_stream_pass2<M, MPOS, double, true, FUSE>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
    else
      hipLaunchKernelGGL((k_stream_pass2<M, MPOS, double, false, FUSE>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
  } else {
    if (D.fp32)
      hipLaunchKernelGGL((k_stream_pass<M, MPOS, float, true, FUSE>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
    else if (D.fast)
      hipLaunchKernelGGL((k_stream_pass<M, MPOS, double, true, FUSE>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
    else
      hipLaunchKernelGGL((k_stream_pass<M, MPOS, double, false, FUSE>), grid, block, 0, st, D, pass, flip,
                         do_seed, substep);
  }
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

template <int M, int MPOS>
inline hipError_t stream_pass_t(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks, int fuse,
                                hipStream_t st) {
  return fuse ? stream_pass_tf<M, MPOS, true>(D, pass, flip, do_seed, substep, chunks, st)
              : stream_pass_tf<M, MPOS, false>(D, pass, flip, do_seed, substep, chunks, st);
}


#define EP_PASS_CASE(MM)                                                                 \
  case MM:                                                                               \
    if (D.mpos == (MM) / 2)                                                              \
      return stream_pass_t<MM, (MM) / 2>(D, pass, flip, do_seed, sub, chunks, fuse, st); \
    if (D.mpos == ((MM) + 1) / 2)                                                        \
      return stream_pass_t<MM, ((MM) + 1) / 2>(D, pass, flip, do_seed, sub, chunks, fuse, st); \
    return hipErrorInvalidValue;

}  // namespace ep
