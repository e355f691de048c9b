// ep_internal.h -- device-side state and helpers of the MI355X EP update.
//
// Device layout (HBM):
//   f            [stream][energy][mu][node_pad]   FP64, node fastest, two
//                buffers A/B; loc[stream*E+energy] says which one holds the
//                current distribution of that (stream, energy) slab.
//   node arrays  [stream][node_pad]               FP64 SoA
//   kept arrays  [stream][energy][kept_index]     FP64 (streaming coefs)
//   task arrays  [stream][energy]                 per (stream, energy)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eprem_hip.h"

namespace ep {

constexpr double kDblMin = 2.2250738585072014e-308;  // DBL_MIN
constexpr double kPI = 3.1415926535897932;
constexpr double kTwoPI = 6.283185307179586;
constexpr double kMP = 1.6726e-24;
constexpr double kMEV = 1.6022e-6;
constexpr double kGEV = 1.6022e-3;
constexpr double kC = 2.99792458e10;
constexpr double kAU = 1.495978707e13;
constexpr double kRSUN = 6.96e10;
constexpr double kRSAU = kRSUN / kAU;
constexpr double kMhdDensityNorm = 1.0;

// Largest subcycle count the library accepts for one operator on one node
// or one (stream, energy).  The reference casts ceil(dt/dt_stable) to int
// without a check (src/energeticParticles.c:940, :1665, :2316); counts past
// this bound only arise from degenerate inputs and are reported as errors.
constexpr int kMaxSubcycles = 1 << 22;
constexpr int kMaxE = 64;  // energies per stream (checked by ep_init)

// Nodes whose focusing needs at least this many subcycles (the innermost
// ones, where dlnB/ds is steep) are integrated with one lane per energy
// instead of one lane per node: a node's subcycle count is the same for
// all its energies, so those waves do not idle behind one slow lane.
constexpr int kHeavyFocus = 32;

// Error key.  The reference panics at the first bad cell in its own loop
// order; every failing cell here computes a key that sorts in that order,
// and the smallest key wins (one 64-bit atomicMin).  Layout, high to low:
//   [substep 7][group 4][stream 14][op 2][a 14][b 8][c 14][inf 1]
// group 0 = stream stage (op 0: prep error, a = kind; op 1: streaming,
// a = energy, b = mu, c = node), group 1 = node stage (op 0: focusing,
// op 1: change; a = node, b = energy, c = mu).  inf = 1 for Inf, 0 NaN.
enum PrepErr { kErrDsUnderflow = 1, kErrStreamShort = 2, kErrSubcycles = 3 };

__host__ __device__ inline unsigned long long err_key(int sub, int group, int s, int op, int a,
                                                      int b, int c, int inf) {
  unsigned long long k = 0;
  k |= (unsigned long long)(sub > 127 ? 127 : sub) << 57;
  k |= (unsigned long long)(group & 0xF) << 53;
  k |= (unsigned long long)(s & 0x3FFF) << 39;
  k |= (unsigned long long)(op & 0x3) << 37;
  k |= (unsigned long long)(a & 0x3FFF) << 23;
  k |= (unsigned long long)(b & 0xFF) << 15;
  k |= (unsigned long long)(c & 0x3FFF) << 1;
  k |= (unsigned long long)(inf & 1);
  return k;
}

// Per-step scalar diagnostics kept on the device.
struct DevDiag {
  int max_sub_focus, max_sub_change, max_nsteps, max_passes, max_chunks, any_thin;
  unsigned long long min_tau_bits;  // positive doubles order like their bits
  unsigned long long err_key;       // ~0ull = no error
  unsigned long long stream_sweeps, focus_sweeps, change_sweeps;  // per substep
  int focus_heavy;  // nodes in the heavy-focusing list
  unsigned queue;   // next slab of the persistent streaming launch
  int prep_err;     // the step's preparation found an error (ds underflow, short stream, subcycles)
  int pad_;
};

// Everything the kernels need, passed by value.
struct Dev {
  // dims
  int S, N, Npad, E, M;
  // params (code units)
  double rScale, dt, dt_full;
  int numEpSteps, mfpType, useBC, stream_block, stream_halo;
  int stream_r;     // kept nodes per lane in the streaming passes (1, 2 or 3)
  int stream_fuse;  // focusing of full streams done in their last streaming pass
  int stream_whole; // a stream that fits in one chunk runs all its subcycles in one pass
  double mfpRadialPower, mfpBPower, mhdBsAu, dsh_min, dampCooling;
  double normJ0, normRadius, xi, beta, dlnp, dmu;
  int mpos;  // first mu index with mugrid >= 0
  // grids [E] / [M]
  const double *egrid, *vgrid, *vgrid_i, *rig_lamo, *plt, *ext, *mugrid;
  const double *mu_c1;  // [M] 0.5 * (1 - mu^2), the focusing velocity's factor
  const double *fm_tab; // tables of fm_exp / fm_log (ep_fastmath.h)
  // f
  double *fA, *fB;
  uint8_t *loc;  // [S*E]
  int32_t *fhead;  // [1] the slot of node 0 in every f row
  // nodes [S*Npad]
  const double *x, *y, *z, *rmag, *bmag, *dlnb, *dlnn, *dupar;
  const int32_t *n_active;  // [S]
  double *ds, *bplus, *bminus, *mfp_r, *radial;
  // thinning [S*Npad] / [S]
  int32_t *list, *ref, *nkept, *thin;
  double *dsil, *dsmin;
  // streaming task tables [S*E]
  int32_t *nsteps, *halo, *passes, *chunks;
  double *dtprop;
  double *del_fac;  // [S*E*M]
  double *seed3;    // [S*E*3]
  double *dsm, *q;  // [S*E*Npad] by kept index
  // node-stage coefficients [S*Npad]
  int32_t *nsub_f, *nsub_c;
  int32_t *heavy;  // [S*Npad] node indices (s*Npad + node) with nsub_f >= kHeavyFocus
  double *dtsub_f, *dtsub_c, *f_dlnbds, *f_dupar, *f_b;
  double *c_dupar, *c_dlnn, *c_dlnb, *c_damp, *c_bc0, *c_bc1;
  DevDiag *diag;
  int fp32;  // f stored as float (mixed-precision form)
  int fast;  // streaming with fused multiply-adds (not bit-identical)
  int focus_alg, change_alg;  // 1 upwind Euler, 2 RK3 upwind, 3 RK3 WENO3
  int gated;      // launched as part of an EP substep (see substep_off)
  int node_gate;  // node-stage path: 0 always, 1 / 2 when every / not every node needs <= 1 change subcycle
  // fraction of the substep each (stream, energy) row is focused over
  // [S*E]; null = the whole substep.  The learned-surrogate form sets it
  // for rows whose first subcycles the network has already advanced; the
  // focusing kernels that take it are k_focus and k_focus_heavy (the
  // surrogate form runs the two-kernel node stage).
  const double *fphi;
  int logf;  // f stored as FP32 ln f (precision mode 3); see f_dec / f_enc
};

// Storage codec of f.  FP64 and plain FP32 storage hold f itself; in
// precision mode 3 the FP32 storage holds ln f (the log-scaled variant of
// the mixed-precision form): kernels decode after loading and encode
// before storing, and compute on f in FP32 as in mode 1.  The hot kernels
// take the codec as a template argument (LOG: their mode-3 instantiations,
// so that mode 1 does not carry it); kernels off the hot path test the
// flag at run time.
template <bool LOG> __device__ __forceinline__ double f_decL(double v) { return v; }
template <bool LOG> __device__ __forceinline__ float f_decL(float v) { return LOG ? expf(v) : v; }
template <bool LOG> __device__ __forceinline__ double f_encL(double v) { return v; }
template <bool LOG> __device__ __forceinline__ float f_encL(float v) { return LOG ? logf(v) : v; }
__device__ __forceinline__ double f_dec(const Dev &, double v) { return v; }
__device__ __forceinline__ float f_dec(const Dev &D, float v) { return D.logf ? expf(v) : v; }
__device__ __forceinline__ double f_enc(const Dev &, double v) { return v; }
__device__ __forceinline__ float f_enc(const Dev &D, float v) { return D.logf ? logf(v) : v; }

// Temporal-blocking tiling of the streaming passes of one (stream, energy)
// slab with `size` kept nodes and `nsteps` subcycles: halo h (subcycles
// per pass), passes and chunks.  A stream that fits in one chunk takes all
// its subcycles in one pass without a halo (the streaming kernels hold its
// two ends); a longer one is cut into chunks whose halos cover the
// subcycles of a pass.  nsteps = 0 (nothing left to stream): no pass.
__device__ inline void task_tiling(const Dev &D, int size, int nsteps, int *h_out, int *passes_out,
                                   int *chunks_out) {
  if (nsteps <= 0) {
    *h_out = 1;
    *passes_out = 0;
    *chunks_out = 1;
    return;
  }
  const bool whole = D.stream_whole && D.stream_block >= size;
  const int H = whole ? nsteps : D.stream_halo, W = D.stream_block;
  const int h = nsteps < H ? nsteps : H;
  const int interior = whole ? W : W - 2 * h;
  *h_out = h;
  *passes_out = (nsteps + h - 1) / h;
  *chunks_out = (size + interior - 1) / interior;
}

// Focusing subcycle length of row (task) of a node whose subcycle is dtsub.
__device__ inline double focus_dtsub(const Dev &D, int task, double dtsub) {
  return D.fphi ? dtsub * D.fphi[task] : dtsub;
}

__device__ inline void note_error(const Dev &D, unsigned long long key) {
  atomicMin(&D.diag->err_key, key);
}

// An error of the once-per-step preparation: the reference panics there,
// before any substep, so the substeps of the step are skipped.
__device__ inline void note_prep_error(const Dev &D, unsigned long long key) {
  atomicMin(&D.diag->err_key, key);
  atomicOr(&D.diag->prep_err, 1);
}

// Whether a kernel of an EP substep has nothing to do: the step's
// preparation reported an error, or the kernel belongs to the node-stage
// path the preparation did not choose (node_gate 1: the one-pass focusing
// and change, taken when every node needs at most one change subcycle;
// 2: the two-kernel path otherwise).  The whole global step is then one
// sequence of launches that the host never has to inspect midway.
__device__ inline bool substep_off(const Dev &D) {
  if (!D.gated) return false;
  // a plain load: the values were written by earlier launches.  (A volatile
  // access here counts as a possible clobber of all memory, and the
  // streaming kernels then read their per-mu coefficients with vector
  // instead of scalar loads inside the subcycle loop: 14% slower.)
  const DevDiag *dg = D.diag;
  if (dg->prep_err) return true;
  if (D.node_gate == 1) return dg->max_sub_change > 1;
  if (D.node_gate == 2) return dg->max_sub_change <= 1;
  return false;
}

// sepSeedFunction (src/energeticParticlesBoundary.c:67-116) with the
// node-only factor radial = pow(r/normRadius, -beta) and the energy-only
// factors plt[e] = pow(E/E1, -gamma), ext[e] = exp(-E/E0) precomputed; the
// product keeps the reference's association order.
__device__ inline double seed_value(const Dev &D, double radial, int e) {
  double normJHe = D.normJ0 * radial * D.plt[e] * D.ext[e];
  double normJ = normJHe / D.xi;
  double normf = normJ / (2.0 * D.egrid[e]);
  if (normf < kDblMin) normf = kDblMin;
  return normf;
}

// f rows: each (stream, energy, mu) row has Npad >= N + 64 slots and holds
// the stream's nodes 0..N-1 from slot *D.fhead on.  A shell spawn moves the
// head back one slot and writes the new node 0 there instead of moving
// every node outward; only when the head reaches slot 0 are the rows
// copied back to the far end (once every Npad - N spawns, ep_nodes.hip).
__device__ inline size_t fidx(const Dev &D, int s, int e, int m, int node) {
  return (((size_t)s * D.E + e) * D.M + m) * (size_t)D.Npad + D.fhead[0] + node;
}

__device__ inline double *fbuf(const Dev &D, int which) { return which ? D.fB : D.fA; }

// f is stored as T: double on the parity path, float in the mixed-precision
// form (the same two allocations, used at 4 bytes per cell).
template <typename T>
__device__ inline T *fbufT(const Dev &D, int which) {
  return reinterpret_cast<T *>(which ? D.fB : D.fA);
}

// The f slab of task t in buffer `which`, from the head: node k of mu row m
// at [m * Npad + k].
template <typename T>
__device__ inline T *fslab(const Dev &D, int which, int task) {
  return fbufT<T>(D, which) + (size_t)task * D.M * D.Npad + D.fhead[0];
}
__device__ inline double *fslab(const Dev &D, int which, int task) { return fslab<double>(D, which, task); }

// FP64 scratch slab of one (stream, energy) when f is stored as float: the
// upper halves of the two f allocations, which float storage leaves unused
// (task parity picks the allocation; each has one spare slab of room).
__device__ inline double *spare_slab(const Dev &D, int task) {
  const size_t slab = (size_t)D.M * D.Npad;
  const size_t off = ((size_t)D.S * D.E * slab * 4 + 7) / 8;  // doubles
  return ((task & 1) ? D.fB : D.fA) + off + (size_t)(task >> 1) * slab;
}

// The reference's interior interface flux v 0.5 ((1 - c) f[up] + (1 + c)
// f[other]) with c = copysign(1, v) (src/energeticParticles.c:1830-1836,
// :1103-1110) equals v * f[up] for finite f, but a non-finite value in the
// other cell reaches it too (0 * Inf and 0 * NaN are NaN).  ep_poison(other,
// u) is u exactly when `other` is finite (up to the sign of a zero flux,
// which the update cannot see) and NaN otherwise, so that the first
// non-finite cell is the reference's.  The wall rows (mu = 0, M-1;
// E = 0, E-1) use the upwind cell alone, as in the reference.
template <typename T>
__device__ __forceinline__ T ep_poison(T other, T u) {
  return fma(other, (T)0, u);
}

// Smallest normal value of T: the floor the reference applies as DBL_MIN.
template <typename T>
struct FloorT;
template <>
struct FloorT<double> {
  static constexpr double v = 2.2250738585072014e-308;
};
template <>
struct FloorT<float> {
  static constexpr float v = 1.17549435e-38f;
};
// max(x, fl) as a single v_max: fmax adds a NaN-quieting v_max on every
// value the compiler cannot prove quiet (a loaded one, say).  A NaN x
// gives fl, as fmax does; the kernels report non-finite values before
// they floor them.
__device__ __forceinline__ double floor_max(double x, double fl) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(x), "s"(fl));
  return r;
}
__device__ __forceinline__ float floor_max(float x, float fl) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(x), "s"(fl));
  return r;
}

// Cross-lane shifts within each 16-lane row of a wave (DPP row_shr:1,
// CTRL 0x111: lane i takes lane i-1 of its row; row_shl:1, CTRL 0x101:
// lane i takes lane i+1); the row's first (last) lane gets 0.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
constexpr int kRowShr1 = 0x111, kRowShl1 = 0x101;

// A face velocity split by the side its focusing flux is taken from:
// vh = min(va, 0), vl = va - vh, so vl = va for va >= 0 and +0 otherwise,
// and vl + vh = va exactly.  Two instructions where sign-bit selects take
// six.  A face then carries fma(vh, f_above, vl * f_below).
template <typename T>
__device__ __forceinline__ void split_face(T va, T &vl, T &vh) {
  vh = fmin(va, (T)0);
  vl = va - vh;
}

// ln FLT_MAX and ln of the FP32 floor: the bounds of ln f in the
// log-scaled storage (f decodes to a finite FP32 value at or above the
// floor).
constexpr double kLnFltMax = 88.72283905206835;
constexpr double kLnFloorF = -87.33654475055310;

// Raw buffer access to one wave-uniform slab: the base address lives in
// scalar registers (the buffer descriptor) and the per-mu offset in a
// scalar offset, so a lane needs one 32-bit VGPR for its address instead
// of a 64-bit pointer per mu row.  Word 3 = 32-bit data format (CDNA).
typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t slab_rsrc(const double *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), 0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ double buf_ld(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ void buf_st(double v, rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)) b64_t;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64_t, v), r, (int)voff, (int)soff, 0);
}

__device__ __forceinline__ float buf_ld(rsrc_t r, uint32_t voff, uint32_t soff, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ double buf_ld(rsrc_t r, uint32_t voff, uint32_t soff, double) {
  return buf_ld(r, voff, soff);
}

__device__ __forceinline__ void buf_st(float v, rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, 0);
}

}  // namespace ep
