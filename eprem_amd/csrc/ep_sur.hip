// ep_sur.hip -- the learned-surrogate form of the EP update.
//
// With the surrogate on (ep_set_surrogate), each EP substep of ep_step
// advances a (stream, energy) slab that needs at least K streaming
// subcycles (nsteps >= K, full-resolution stream) by A = nsteps / K
// network applications before the plain path runs.  One application
// replaces K subcycles of streaming with pitch-angle relaxation
// (DiffuseStreamData, src/energeticParticles.c:2352-2430) and the
// adiabatic focusing over the same K dtProp (AdiabaticFocusing,
// :1683-1762, :1806-1896) of every kept node k = 1..n-2 from the rows of
// nodes k-1, k, k+1; node 0 keeps its boundary row (the reference resets
// nodes 0-2 to the seed before every subcycle) and the last node takes the
// relaxation-only large step.  The network reads, per node and mu cell,
// ln f(k-1) - ln f(k), ln f(k) minus its mean over mu and ln f(k+1) -
// ln f(k), and six scalars (sur_scalars, ep_sur.h); it returns the
// correction ln f - ln f_L
// to the large step f_L = iso + (f - iso) q^K (the exact relaxation over
// K dtProp, no transport), which is formed in FP64: the network's bf16
// arithmetic only touches the correction.
//
// Residual: the plain path judges the network.  After every application,
// at sampled kept nodes c of each slab (about one per `stride` nodes) one
// wave runs the plain path over the nodes c-K .. c+K (the dependence
// window of node c over K subcycles: upwind streaming moves information
// one node per subcycle and relaxation couples the mu cells of a node
// only), so its node c after K subcycles, then the focusing over K dtProp,
// is the plain result itself; the residual of the slab is the mean over
// its sampled nodes of the relative error of the network's <f>_mu (the
// omnidirectional flux 2 E <f>_mu up to its factor).  A stream with any
// slab above the tolerance, or a non-finite network output, is rejected
// for the substep: its slabs are recomputed by the plain path from the
// state before the substep (the network writes only the spare buffer and
// fC).
//
// After the applications the accepted slabs run the remaining
// nsteps - A K plain subcycles and are focused over that fraction of the
// substep; rejected streams and slabs that did not qualify run the plain
// path in full.  The adiabatic change is always the plain one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ep_internal.h"
#include "ep_launch.h"
#include "ep_mlp_core.h"
#include "ep_state.h"
#include "ep_stream.h"
#include "ep_sur.h"

namespace ep {

#define HIPCHK(x, what)                                \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) return hip_fail(e_, what);   \
  } while (0)

// samples per wave of k_sur_apply (32-sample tiles: kSurNT) and waves
// per block
constexpr int kSurNT = 1;
constexpr int kSurWaves = 8;
constexpr int kSurSeg = 32 * kSurNT;
// one sampled node of the plain check per this many kept nodes
constexpr int kSurStride = 256;

struct SurDev {
  double *fC;
  int32_t *A, *reject;
  double *errsum;  // [S*E] sum of the sampled relative errors of one application
  unsigned long long *res;
  int32_t *nrem, *hrem, *prem, *crem;
  double *phi;
  float *scal;  // [S*E][6][Npad] the scalar inputs of every kept node
  SurStats *stats;
  int K, G, stride, NS;  // G: apply items per slab; NS: most sampled nodes per slab
  double tol;
};

static SurDev make_sur() {
  SurDev S;
  S.fC = g.sur.fC;
  S.A = g.sur.A;
  S.reject = g.sur.reject;
  S.errsum = g.sur.errsum;
  S.res = g.sur.res;
  S.nrem = g.sur.nrem;
  S.hrem = g.sur.hrem;
  S.prem = g.sur.prem;
  S.crem = g.sur.crem;
  S.phi = g.sur.phi;
  S.scal = g.sur.scal;
  S.stats = g.sur.stats;
  S.K = g.sur.K;
  S.G = g.sur.G;
  S.stride = g.sur.stride;
  S.NS = g.sur.NS;
  S.tol = g.sur.tol;
  return S;
}

// f slab of `task` in buffer w (0: A, 1: B, 2: the surrogate's third)
__device__ inline double *sur_slab(const Dev &D, const SurDev &S, int w, int task, int head) {
  double *base = (w == 2) ? S.fC : (w ? D.fB : D.fA);
  return base + (size_t)task * D.M * D.Npad + head;
}

// Buffers of application `level` of a slab with A applications starting
// in buffer `cur`: the last one writes the spare buffer, the ones before
// alternate with the third buffer, and the current state is never written.
__device__ inline void sur_buffers(int A, int level, int cur, int *src, int *dst) {
  const int spare = cur ^ 1;
  *src = (level == 0) ? cur : (((A - level) & 1) ? 2 : spare);
  *dst = (((A - 1 - level) & 1) == 0) ? spare : 2;
}

// ln f of a positive normal double: exponent plus the FP32 log2 of the
// mantissa in [1, 2) (absolute error ~1e-7, far below the bf16 rounding of
// the network's inputs)
__device__ inline double sur_ln(double f) {
  const long long b = __double_as_longlong(f);
  const int ex = (int)((b >> 52) & 0x7ff) - 1023;
  const double m = __longlong_as_double((b & 0x000fffffffffffffLL) | 0x3ff0000000000000LL);
  return ((double)ex + (double)__log2f((float)m)) * 0.6931471805599453;
}

// exp(z) for the network's output: 2^n times the FP32 exp2 of the fraction
// (relative error ~1e-7); below the FP64 range it gives 0 (floored to
// DBL_MIN afterwards), above it Inf
__device__ inline double sur_exp(double z) {
  double t = z * 1.4426950408889634;
  t = fmin(fmax(t, -1100.0), 1100.0);
  const double n = floor(t);
  const double m = (double)__builtin_amdgcn_exp2f((float)(t - n));
  return ldexp(m, (int)n);
}

// ---------------------------------------------------------------------------
// k_sur_plan: per (stream, energy), the applications of this substep.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sur_plan(Dev D, SurDev S, int substep) {
  if (substep_off(D)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (substep == 0 && t == 0) {
    SurStats *st = S.stats;
    st->network_slabs = st->fallback_slabs = st->plain_slabs = st->network_cell_subcycles = 0ull;
    st->max_res_bits = 0ull;
    st->rejected_stream_substeps = 0u;
  }
  if (t >= D.S * D.E) return;
  const int s = t / D.E, e = t % D.E;
  const int nk = D.nkept[s], ns = D.nsteps[t];
  S.A[t] = (!D.thin[s] && nk >= 8 && ns >= S.K) ? ns / S.K : 0;
  S.res[t] = 0ull;
  S.errsum[t] = 0.0;
  if (e == 0) S.reject[s] = 0;
}

// ---------------------------------------------------------------------------
// k_sur_scalars: the six scalar inputs (sur_scalars) of every kept node of
// every qualifying slab, once per global step (they depend on the nodes and
// the step's subcycle counts only).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sur_scalars(Dev D, SurDev S) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int task = blockIdx.y;
  const int s = task / D.E, e = task % D.E;
  const int nk = D.nkept[s];
  if (k < 1 || k > nk - 2 || D.thin[s] || D.nsteps[task] < S.K) return;
  double x[6];
  sur_scalars(D, s, e, k, S.K, x);
  float *o = S.scal + (size_t)task * 6 * D.Npad + k;
#pragma unroll
  for (int j = 0; j < 6; ++j) o[(size_t)j * D.Npad] = (float)x[j];
}

// ---------------------------------------------------------------------------
// k_sur_apply: one network application (level `level`) of every slab with
// more than `level` applications.  A persistent grid of kSurWaves-wave
// blocks holding the network's weights in LDS; a wave takes kSurSeg
// consecutive kept nodes of one slab (kSurNT 32-sample tiles), gathers
// their inputs from f, evaluates the network on the matrix cores
// (ep_mlp_core.h) and writes the new rows.
// ---------------------------------------------------------------------------
template <int KS0>
__global__ __launch_bounds__(64 * kSurWaves) void k_sur_apply(Dev D, SurDev S, int level,
                                                              const uint4 *__restrict__ packed, int packed16,
                                                              int bias16, int nhid) {
  if (substep_off(D)) return;
  constexpr int NT = kSurNT;
  extern __shared__ uint4 lds_w[];
  for (int t = threadIdx.x; t < packed16; t += blockDim.x) lds_w[t] = packed[t];
  __syncthreads();
  const float *lbias = (const float *)(lds_w + bias16);
  const bf16x8_t *frag0 = (const bf16x8_t *)lds_w;
  float *lds_f = (float *)(lds_w + packed16);  // per wave: ln f rows of its node window
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kSurWaves + (threadIdx.x >> 6));
  const int nwave = gridDim.x * kSurWaves;
  const int M = D.M, E = D.E, G = S.G, K = S.K, Np = D.Npad;
  const int nin = 3 * M;
  const int head = D.fhead[0];
  const int total = D.S * E * G;
  for (int item = wave; item < total; item += nwave) {
    const int task = item / G, g = item % G;
    const int A = S.A[task];
    if (A <= level) continue;
    const int s = task / E;
    // a stream rejected at an earlier application recomputes its substep on
    // the plain path: the rest of its applications are not needed
    if (level > 0 && S.reject[s]) continue;
    const int nk = D.nkept[s];
    const int k0 = 1 + kSurSeg * g;
    if (k0 > nk - 2) continue;
    int srcw, dstw;
    sur_buffers(A, level, D.loc[task], &srcw, &dstw);
    const double *src = sur_slab(D, S, srcw, task, head);
    double *dst = sur_slab(D, S, dstw, task, head);
    const size_t ok = (size_t)task * Np;
    const float *scal = S.scal + (size_t)task * 6 * Np;
    // ---- inputs.  ln f of the window of kept nodes k0-1 .. k0+kSurSeg
    // (every sample's three nodes) goes to this wave's LDS rows [m][node];
    // then each lane reads its samples' inputs back from LDS.
    // (stored relative to ln f of the window's node k0-1, mu 0, so that
    // FP32 keeps the small differences between neighbouring nodes)
    constexpr int WN = kSurSeg + 2;
    float *lw = lds_f + (threadIdx.x >> 6) * (32 * WN);
    const double lref = sur_ln(src[(size_t)(k0 - 1)]);
    for (int idx = lane; idx < M * WN; idx += 64) {
      const int m = idx / WN, j = idx - m * WN;
      const int node = k0 - 1 + j;
      float lf = 0.0f;
      if (node <= nk - 1) {
        lf = (float)(sur_ln(src[(size_t)m * Np + node]) - lref);
      }
      lw[idx] = lf;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bool valid[NT];
    int kk[NT];
    bf16x8_t b0[NT][KS0];
    // the lane half, opaque to the compiler inside the loop: otherwise it
    // hoists the index arithmetic of all 8 KS0 inputs out of the loop and
    // keeps it live (in registers and scratch) across the whole kernel
    int hh = h;
    asm volatile("" : "+v"(hh));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int jc = 32 * t + col + 1;  // window column of node k
      const int k = k0 - 1 + jc;
      kk[t] = k;
      valid[t] = k <= nk - 2;
      // mean ln f at node k: each half of the wave sums half of the mu rows
      double csum = 0.0;
      const int mh = (M + 1) / 2;
      for (int m = hh * mh; m < min(M, (hh + 1) * mh); ++m) csum += (double)lw[m * WN + jc];
      csum += __shfl_xor(csum, 32);
      const double c = csum / (double)M;
      const float cf = (float)c;
#pragma unroll
      for (int ks = 0; ks < KS0; ++ks) {
        bf16x8_t v;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int phi = 16 * ks + 8 * hh + q;
          float r = 0.0f;
          if (valid[t]) {
            if (phi < nin) {
              // ln f(k-1) - ln f(k), ln f(k) - mean, ln f(k+1) - ln f(k)
              const int off = (phi >= M) + (phi >= 2 * M);
              const int m = phi - off * M;
              const float own = lw[m * WN + jc];
              r = (off == 1) ? own - cf : lw[m * WN + jc - 1 + off] - own;
            } else if (phi < nin + 6) {
              r = scal[(size_t)(phi - nin) * Np + k];
            }
          }
          v[q] = (__bf16)r;
        }
        b0[t][ks] = v;
      }
    }
    // ---- network
    f32x16_mlp_t acc[NT][kFH / 32];
    mlp_layer_lean<KS0, kFH / 32, NT>(frag0, b0, acc, lane);
    bf16x8_t b[NT][kFH / 16];
    mlp_relu_to_b<NT>(acc, lbias, h, b);
    const bf16x8_t *fr = frag0 + KS0 * (kFH / 32) * 64;
    for (int l = 1; l < nhid; ++l) {
      mlp_layer_lean<kFH / 16, kFH / 32, NT>(fr, b, acc, lane);
      fr += (kFH / 16) * (kFH / 32) * 64;
      mlp_relu_to_b<NT>(acc, lbias + l * kFH, h, b);
    }
    f32x16_mlp_t o[NT][1];
    mlp_layer_lean<kFH / 16, 1, NT>(fr, b, o, lane);
    const float *ob = lbias + nhid * kFH;
    // ---- new rows: the large step (exact relaxation over K dtProp,
    // f -> iso + (f - iso) q^K, in FP64) times exp of the network's
    // correction
    bool bad = false;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double fk[16];
      double part = 0.0;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 8 * (r >> 2) + 4 * h + (r & 3);
        fk[r] = (valid[t] && j < M) ? src[(size_t)j * Np + kk[t]] : 0.0;
        part += fk[r];
      }
      part += __shfl_xor(part, 32);
      if (!valid[t]) continue;
      const double iso = part / (double)M;
      const double qK = pow(D.q[ok + kk[t]], (double)K);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 8 * (r >> 2) + 4 * h + (r & 3);
        if (j < M) {
          const double base = iso + (fk[r] - iso) * qK;
          double fv = base * sur_exp((double)(o[t][0][r] + ob[j]));
          bad = bad || !isfinite(fv);
          fv = fmax(fv, kDblMin);
          dst[(size_t)j * Np + kk[t]] = fv;
        }
      }
    }
    // ---- end nodes: node 0 keeps its row; the last node takes the
    // relaxation-only large step f -> iso + (f - iso) q^K
    if (g == 0 && lane < M) dst[(size_t)lane * Np] = src[(size_t)lane * Np];
    if (k0 + kSurSeg - 1 >= nk - 2) {
      const int kl = nk - 1;
      const double fv = (lane < M) ? src[(size_t)lane * Np + kl] : 0.0;
      double sum = fv;
#pragma unroll
      for (int sh = 32; sh >= 1; sh >>= 1) sum += __shfl_xor(sum, sh);
      const double iso = sum / (double)M;
      const double qK = pow(D.q[ok + kl], (double)K);
      if (lane < M) dst[(size_t)lane * Np + kl] = fmax(iso + (fv - iso) * qK, kDblMin);
    }
    if (__any(bad ? 1 : 0) && lane == 0) atomicOr(&S.reject[s], 1);
  }
}

// ---------------------------------------------------------------------------
// k_sur_check: the plain path at the sampled nodes of every slab with more
// than `level` applications.  A wave per sampled node c: lane l holds node
// c - K + l and its M cells; K subcycles of DiffuseStreamData
// (:2366-2426: seed rows of nodes 0-2, upwind step with the stream ends
// held, relaxation), neighbours by DPP lane shifts; lanes near the window
// edges go wrong one node per subcycle, which after K subcycles has not
// reached lane K.  Lane K then takes the focusing over K dtProp
// (focus = 1) and adds the relative error of the network's <f>_mu at c to
// the slab's sum.
// ---------------------------------------------------------------------------
__device__ inline int sur_samples(const SurDev &S, int nk) {
  const int n = (nk - 2) / S.stride;
  return n > 1 ? n : 1;
}

__global__ __launch_bounds__(256) void k_sur_check(Dev D, SurDev S, int level, int focus) {
  if (substep_off(D)) return;
  const int lane = threadIdx.x & 63;
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (item >= D.S * D.E * S.NS) return;
  const int task = item / S.NS, j = item % S.NS;
  const int A = S.A[task];
  if (A <= level) return;
  const int s = task / D.E, e = task % D.E;
  if (level > 0 && S.reject[s]) return;
  const int nk = D.nkept[s];
  const int ns = sur_samples(S, nk);
  if (j >= ns) return;
  const int K = S.K, M = D.M, Np = D.Npad;
  const int c = 1 + ((2 * j + 1) * (nk - 2)) / (2 * ns);
  int srcw, dstw;
  sur_buffers(A, level, D.loc[task], &srcw, &dstw);
  const int head = D.fhead[0];
  const double *src = sur_slab(D, S, srcw, task, head);
  const double *dst = sur_slab(D, S, dstw, task, head);
  const size_t ok = (size_t)task * Np;
  const int node = c - K + lane;
  const bool act = (lane <= 2 * K) && node >= 0 && node <= nk - 1;
  double f[32];
#pragma unroll
  for (int m = 0; m < 32; ++m) f[m] = (act && m < M) ? src[(size_t)m * Np + node] : 0.0;
  const double dsm = act ? D.dsm[ok + node] : 0.0;
  const double q = act ? D.q[ok + node] : 1.0;
  const double pos = (node == 0) ? 0.0 : dsm, neg = (node == nk - 1) ? 0.0 : dsm;
  const bool bc = D.useBC && act && node < 3;
  const double sd = bc ? D.seed3[(size_t)task * 3 + node] : 0.0;
  const double *del = D.del_fac + (size_t)task * M;
  const int mpos = D.mpos;
  const double inv_m = 1.0 / (double)M;
  for (int it = 0; it < K; ++it) {
    if (bc) {
#pragma unroll
      for (int m = 0; m < 32; ++m) f[m] = (m < M) ? sd : 0.0;
    }
    double iso = 0.0;
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      if (m < M) {
        if (m >= mpos) {
          const double left = dpp_shr1(f[m], f[m]);
          f[m] = f[m] + pos * del[m] * (left - f[m]);
        } else {
          const double right = dpp_shl1(f[m], f[m]);
          f[m] = f[m] + neg * del[m] * (f[m] - right);
        }
        iso += f[m];
      }
    }
    iso = inv_m * iso;
#pragma unroll
    for (int m = 0; m < 32; ++m)
      if (m < M) f[m] = iso + (f[m] - iso) * q;
  }
  if (lane == K) {
    if (focus) sur_focus_row<32>(D, (size_t)s * Np + c, e, (double)K * D.dtprop[task], K, D.nsteps[task], f);
    double pm = 0.0, nm = 0.0;
#pragma unroll
    for (int m = 0; m < 32; ++m)
      if (m < M) {
        pm += f[m];
        nm += dst[(size_t)m * Np + c];
      }
    const double err = fabs(nm - pm) / pm;
    atomicAdd(&S.errsum[task], isfinite(err) ? err : 1e300);
  }
}

// k_sur_judge: per (stream, energy), the residual of this application
// (mean sampled error), the stream's rejection above the tolerance.
__global__ __launch_bounds__(256) void k_sur_judge(Dev D, SurDev S, int level) {
  if (substep_off(D)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.S * D.E) return;
  if (S.A[t] <= level) return;
  const int s = t / D.E;
  const double r = S.errsum[t] / (double)sur_samples(S, D.nkept[s]);
  S.errsum[t] = 0.0;
  if (!(r <= S.tol)) atomicOr(&S.reject[s], 1);
  if (r <= 1e300) atomicMax(&S.res[t], (unsigned long long)__double_as_longlong(r));
}

// ---------------------------------------------------------------------------
// k_sur_commit: per (stream, energy), accept the network's result (the
// slab's current buffer becomes the spare one) or fall back, and set the
// plain remainder: subcycles, tiling and focusing fraction.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sur_commit(Dev D, SurDev S) {
  if (substep_off(D)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.S * D.E) return;
  const int s = t / D.E, e = t % D.E;
  const int A = S.A[t], ns = D.nsteps[t], nk = D.nkept[s];
  const bool rej = S.reject[s] != 0;
  const bool acc = A > 0 && !rej;
  if (acc) D.loc[t] ^= (uint8_t)1;
  const int rem = acc ? ns - A * S.K : ns;
  int h, passes, chunks;
  task_tiling(D, nk, rem, &h, &passes, &chunks);
  S.nrem[t] = rem;
  S.hrem[t] = h;
  S.prem[t] = passes;
  S.crem[t] = chunks;
  S.phi[t] = acc ? (double)rem / (double)ns : 1.0;
  SurStats *st = S.stats;
  if (acc) {
    atomicAdd(&st->network_slabs, 1ull);
    atomicAdd(&st->network_cell_subcycles, (unsigned long long)A * S.K * nk * D.M);
  } else if (A > 0) {
    atomicAdd(&st->fallback_slabs, 1ull);
  } else {
    atomicAdd(&st->plain_slabs, 1ull);
  }
  if (A > 0) atomicMax(&st->max_res_bits, S.res[t]);
  if (e == 0 && rej) atomicAdd(&st->rejected_stream_substeps, 1u);
}

// One application per qualifying slab into the spare buffer, whose slabs
// then become current (the probe of ep_surrogate_probe, mode 0).
__global__ __launch_bounds__(256) void k_sur_flip_applied(Dev D, SurDev S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.S * D.E) return;
  if (S.A[t] > 0) D.loc[t] ^= (uint8_t)1;
}

// A = 1 for the slabs with at least K subcycles (the probes)
__global__ __launch_bounds__(256) void k_sur_one(Dev D, SurDev S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.S * D.E) return;
  const int s = t / D.E;
  S.A[t] = (!D.thin[s] && D.nkept[s] >= 8 && D.nsteps[t] >= S.K) ? 1 : 0;
  S.res[t] = 0ull;
  S.errsum[t] = 0.0;
  if (t % D.E == 0) S.reject[s] = 0;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
template <int KS0>
static hipError_t launch_apply_t(const Dev &D, const SurDev &S, int level, hipStream_t st) {
  const uint4 *packed = nullptr;
  int p16 = 0, b16 = 0;
  std::vector<int> sizes;
  hipError_t e = mlp_fused_weights(st, &packed, &p16, &b16, &sizes);
  if (e != hipSuccess) return e;
  const size_t lds = (size_t)p16 * 16 + (size_t)kSurWaves * 32 * (kSurSeg + 2) * sizeof(float);
  e = hipFuncSetAttribute((const void *)k_sur_apply<KS0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  hipLaunchKernelGGL(k_sur_apply<KS0>, dim3(cus), dim3(64 * kSurWaves), lds, st, D, S, level, packed, p16, b16,
                     (int)sizes.size() - 2);
  return hipGetLastError();
}

static hipError_t launch_apply(const Dev &D, const SurDev &S, int level, hipStream_t st) {
  switch ((3 * D.M + 6 + 15) / 16) {
    case 1: return launch_apply_t<1>(D, S, level, st);
    case 2: return launch_apply_t<2>(D, S, level, st);
    case 3: return launch_apply_t<3>(D, S, level, st);
    case 4: return launch_apply_t<4>(D, S, level, st);
    case 5: return launch_apply_t<5>(D, S, level, st);
    case 6: return launch_apply_t<6>(D, S, level, st);
    case 7: return launch_apply_t<7>(D, S, level, st);
    case 8: return launch_apply_t<8>(D, S, level, st);
    default: return hipErrorInvalidValue;
  }
}

static hipError_t launch_check(const Dev &D, const SurDev &S, int level, int focus, hipStream_t st) {
  const int SE = D.S * D.E, items = SE * S.NS;
  hipLaunchKernelGGL(k_sur_check, dim3((items + 3) / 4), dim3(256), 0, st, D, S, level, focus);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sur_judge, dim3((SE + 255) / 256), dim3(256), 0, st, D, S, level);
  return hipGetLastError();
}

hipError_t launch_sur_substep(const Dev &D, int substep, int levels, hipStream_t st) {
  const SurDev S = make_sur();
  const int SE = D.S * D.E;
  hipLaunchKernelGGL(k_sur_plan, dim3((SE + 255) / 256), dim3(256), 0, st, D, S, substep);
  hipError_t e = hipGetLastError();
  for (int j = 0; j < levels && e == hipSuccess; j++) {
    e = launch_apply(D, S, j, st);
    if (e == hipSuccess) e = launch_check(D, S, j, 1, st);
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sur_commit, dim3((SE + 255) / 256), dim3(256), 0, st, D, S);
  return hipGetLastError();
}

hipError_t launch_sur_prep(const Dev &D, hipStream_t st) {
  const SurDev S = make_sur();
  hipLaunchKernelGGL(k_sur_scalars, dim3((D.N + 255) / 256, D.S * D.E), dim3(256), 0, st, D, S);
  return hipGetLastError();
}

Dev sur_stream_dev(const Dev &D) {
  Dev R = D;
  R.nsteps = g.sur.nrem;
  R.halo = g.sur.hrem;
  R.passes = g.sur.prem;
  R.chunks = g.sur.crem;
  return R;
}

Dev sur_node_dev(const Dev &D) {
  Dev R = D;
  R.fphi = g.sur.phi;
  return R;
}

// The network's inputs must be ln f of 3 M cells plus six scalars and its
// output M cells, in the fused kernel's shape.
static int check_network() {
  const uint4 *packed = nullptr;
  int p16 = 0, b16 = 0;
  std::vector<int> sizes;
  hipError_t e = mlp_fused_weights(g.st, &packed, &p16, &b16, &sizes);
  if (e != hipSuccess)
    return fail(EP_ERR_INVALID_ARG, "surrogate: no network, or not of the fused shape (ep_mlp_init with "
                                    "input <= 128, hidden layers of 128, output <= 32)");
  const int M = g.d.num_mu;
  if (sizes.front() != 3 * M + 6 || sizes.back() != M)
    return fail(EP_ERR_INVALID_ARG, "surrogate: the network must map 3*num_mu+6 = %d inputs to num_mu = %d outputs",
                3 * M + 6, M);
  if (g.sur.packed != (const void *)packed) {
    g.sur.packed = packed;
    g.drop_graph();
  }
  return EP_OK;
}

int sur_check_step() {
  if (g.sur.K <= 0) return EP_OK;
  if (g.fp32) return fail(EP_ERR_UNSUPPORTED, "surrogate: needs FP64 storage (precision mode 0 or 2)");
  if (g.p.adiabaticFocusAlg != 1 || g.p.adiabaticChangeAlg != 1)
    return fail(EP_ERR_UNSUPPORTED, "surrogate: needs the upwind focusing and change operators");
  return check_network();
}

}  // namespace ep

using namespace ep;

extern "C" {

int ep_set_surrogate(int32_t K, double tol) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (K == 0) {
    g.sur.K = 0;
    g.drop_graph();
    return EP_OK;
  }
  // the plain check holds a node's 2K + 1 dependence window in one wave
  if (K < 1 || K > 31 || !(tol >= 0.0)) return fail(EP_ERR_INVALID_ARG, "surrogate: K in 1..31, tol >= 0");
  if (g.fp32) return fail(EP_ERR_UNSUPPORTED, "surrogate: needs FP64 storage (precision mode 0 or 2)");
  if (g.p.adiabaticFocusAlg != 1 || g.p.adiabaticChangeAlg != 1)
    return fail(EP_ERR_UNSUPPORTED, "surrogate: needs the upwind focusing and change operators");
  int rc = check_network();
  if (rc) return rc;
  if (!g.sur.fC) {
    const size_t S = g.d.num_streams, E = g.d.num_energies, M = g.d.num_mu, NP = g.Npad;
    const int G = (int)((NP + kSurSeg - 1) / kSurSeg);
    const size_t SE = S * E;
    auto al = [&](void **p, size_t bytes) -> hipError_t {
      hipError_t e = hipMalloc(p, bytes + 64);
      if (e == hipSuccess) g.allocs.push_back(*p);
      return e;
    };
    hipError_t e = al((void **)&g.sur.fC, SE * M * NP * 8);
    if (e == hipSuccess) e = al((void **)&g.sur.A, SE * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.reject, S * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.errsum, SE * 8);
    if (e == hipSuccess) e = al((void **)&g.sur.res, SE * 8);
    if (e == hipSuccess) e = al((void **)&g.sur.nrem, SE * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.hrem, SE * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.prem, SE * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.crem, SE * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.phi, SE * 8);
    if (e == hipSuccess) e = al((void **)&g.sur.scal, SE * 6 * NP * 4);
    if (e == hipSuccess) e = al((void **)&g.sur.stats, sizeof(SurStats));
    if (e == hipSuccess) e = hipMemset(g.sur.stats, 0, sizeof(SurStats));
    if (e == hipSuccess) e = hipMemset(g.sur.errsum, 0, SE * 8);
    if (e != hipSuccess) {
      // what was allocated stays listed in g.allocs and is released with the rest
      g.sur = SurState();
      return hip_fail(e, "surrogate buffers");
    }
    g.sur.G = G;
    g.sur.stride = kSurStride;
    g.sur.NS = std::max(1, (g.d.num_nodes - 2) / kSurStride);
  }
  g.sur.K = K;
  g.sur.tol = tol;
  g.drop_graph();
  return EP_OK;
}

int ep_get_surrogate_stats(ep_surrogate_stats *out) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!out) return fail(EP_ERR_INVALID_ARG, "null stats");
  memset(out, 0, sizeof(*out));
  if (!g.sur.stats) return EP_OK;
  SurStats h;
  HIPCHK(hipMemcpy(&h, g.sur.stats, sizeof(h), hipMemcpyDeviceToHost), "surrogate stats");
  out->network_slabs = (int64_t)h.network_slabs;
  out->fallback_slabs = (int64_t)h.fallback_slabs;
  out->plain_slabs = (int64_t)h.plain_slabs;
  out->network_cell_subcycles = (int64_t)h.network_cell_subcycles;
  out->rejected_stream_substeps = (int32_t)h.rejected_stream_substeps;
  double r;
  memcpy(&r, &h.max_res_bits, 8);
  out->max_residual = r;
  return EP_OK;
}

int ep_surrogate_probe(int32_t mode, double *max_residual) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  if (g.sur.K <= 0) return fail(EP_ERR_INVALID_ARG, "ep_set_surrogate first");
  if (mode != 0 && mode != 1) return fail(EP_ERR_INVALID_ARG, "probe mode 0 or 1");
  int rc = sur_check_step();
  if (rc) return rc;
  Dev D = make_dev();
  const SurDev S = make_sur();
  hipStream_t st = g.st;
  const int SE = D.S * D.E;
  HIPCHK(launch_reset_diag(g.diag, st), "reset_diag");
  HIPCHK(launch_stream_prep(D, st), "stream_prep");
  HIPCHK(launch_node_prep(D, st), "node_prep");
  DevDiag hd;
  HIPCHK(hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, st), "diag copy");
  HIPCHK(hipStreamSynchronize(st), "prep sync");
  if (hd.prep_err) return fail(EP_ERR_INVALID_ARG, "the step's preparation reported an error");
  HIPCHK(launch_sur_prep(D, st), "surrogate scalars");
  hipLaunchKernelGGL(k_sur_one, dim3((SE + 255) / 256), dim3(256), 0, st, D, S);
  HIPCHK(hipGetLastError(), "sur_one");
  if (mode == 0) {
    HIPCHK(launch_apply(D, S, 0, st), "surrogate apply");
  } else {
    // one plain streaming pass of exactly K subcycles into the spare
    // buffer (halo K, one node per lane), as the sample generator runs it
    const int keep_b = g.stream_block, keep_h = g.stream_halo, keep_r = g.stream_r, keep_w = g.stream_whole;
    g.stream_block = 256;
    g.stream_halo = S.K;
    g.stream_r = 1;
    g.stream_whole = 0;
    Dev P = make_dev();
    hipError_t e = launch_reset_diag(g.diag, st);
    if (e == hipSuccess) e = launch_stream_prep(P, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = launch_stream_pass(P, 0, 0, 0, 0, hd.max_chunks, 0, st);
    g.stream_block = keep_b;
    g.stream_halo = keep_h;
    g.stream_r = keep_r;
    g.stream_whole = keep_w;
    g.drop_graph();
    HIPCHK(e, "plain pass");
  }
  // streaming alone in mode 1: the check leaves out the focusing
  HIPCHK(launch_check(D, S, 0, mode == 0 ? 1 : 0, st), "plain check");
  if (mode == 0) {
    hipLaunchKernelGGL(k_sur_flip_applied, dim3((SE + 255) / 256), dim3(256), 0, st, D, S);
    HIPCHK(hipGetLastError(), "flip");
  }
  std::vector<unsigned long long> res(SE);
  HIPCHK(hipMemcpyAsync(res.data(), S.res, SE * 8, hipMemcpyDeviceToHost, st), "residuals");
  HIPCHK(hipStreamSynchronize(st), "probe sync");
  double mx = 0.0;
  for (unsigned long long b : res) {
    double r;
    memcpy(&r, &b, 8);
    if (r > mx) mx = r;
  }
  if (max_residual) *max_residual = mx;
  // restore the task tables of the current tiling
  Dev R = make_dev();
  HIPCHK(launch_stream_prep(R, st), "stream_prep");
  HIPCHK(launch_reset_diag(g.diag, st), "reset_diag");
  HIPCHK(hipStreamSynchronize(st), "probe sync");
  return EP_OK;
}

}  // extern "C"
