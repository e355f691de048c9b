// ep_kernels.hip -- CDNA4 kernels of the EP update (FP64 parity path).
//
// Arithmetic follows the reference's operation order and is compiled with
// -ffp-contract=off, so results agree with the FP64 oracle to rounding of
// the device libm (exp/log/pow); division and sqrt are IEEE-correct.
//
// Per global step (node positions are fixed inside it, so everything that
// depends only on nodes is computed once):
//   k_stream_prep   updateStreamValues, GetStreamList, FindSegmentLengths,
//                   mean-free-path and seed radial factors, min tau
//   k_task_prep     per (stream, energy): nsteps, dtProp, tiling, del_fac
//   k_stream_coef   per kept node and energy: ds_i multiplier, exp(-dt/tau)
//   k_node_prep     per node: focusing/change coefficients and subcycles
// Per EP substep:
//   k_stream_pass*  DiffuseStreamData on register-resident node chunks,
//                   temporal blocking with halos (one launch per pass;
//                   the kernels are in ep_stream.h)
//   k_stream_finish alias copies of thinned nodes, buffer bookkeeping
//   k_focus_heavy   AdiabaticFocusing of nodes with many subcycles, a lane
//                   per energy
//   k_focus_change4 / k_focus_change (ep_node.hip) AdiabaticFocusing and
//                   AdiabaticChange of a node in one pass over f, or
//   k_focus         AdiabaticFocusing, a lane per node, mu in registers
//   k_change        AdiabaticChange, a lane per node, energy in registers
#include <float.h>
#include <math.h>

#include "ep_fastmath.h"
#include "ep_internal.h"
#include "ep_launch.h"
#include "ep_stream.h"

namespace ep {

// Lane-layout probe used by the tests: out[i] = in[i-1] (lane 0: -1),
// out2[i] = in[i+1] (lane 63: -2) for one wave.
__global__ void k_dpp_probe(const double *in, double *out, double *out2) {
  const int i = threadIdx.x;
  const double v = in[i];
  out[i] = dpp_shr1(v, -1.0);
  out2[i] = dpp_shl1(v, -2.0);
}

__device__ __forceinline__ double dist_scaled(const Dev &D, size_t a, size_t b) {
  const double rs = D.rScale;
  const double ax = D.x[a] * rs, ay = D.y[a] * rs, az = D.z[a] * rs;
  const double bx = D.x[b] * rs, by = D.y[b] * rs, bz = D.z[b] * rs;
  return sqrt((ax - bx) * (ax - bx) + (ay - by) * (ay - by) + (az - bz) * (az - bz));
}

__device__ __forceinline__ void atomic_min_pos_double(unsigned long long *addr, double v) {
  atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

// Block-wide minimum (256 threads) of non-NaN values.
__device__ double block_min_256(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      const double o = sh[t + st];
      if (o < sh[t]) sh[t] = o;
    }
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------------
// k_stream_prep: one 256-thread block per stream.
//   updateStreamValues      src/simCore.c:325-435
//   GetStreamList           src/energeticParticles.c:2147-2234
//   FindSegmentLengths      src/energeticParticles.c:2493-2601
//   meanFreePath (node part) src/meanFreePath.c:13-50
//   min_tau diagnostic      src/energeticParticles.c:149-156
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stream_prep(Dev D) {
  const int s = blockIdx.x, tid = threadIdx.x;
  const int n = D.n_active[s];
  const size_t o = (size_t)s * D.Npad;
  __shared__ int s_thin, s_size;
  __shared__ double sh[256];
  if (tid == 0) s_thin = 0;
  __syncthreads();
  for (int k = tid; k < n; k += blockDim.x) {
    const size_t i = o + k;
    if (k == 0) {
      D.ds[i] = dist_scaled(D, i + 1, i);
      D.bminus[i] = D.bmag[i];
      D.bplus[i] = D.bmag[i + 1];
    } else if (k == n - 1) {
      D.ds[i] = dist_scaled(D, i, i - 1);
      D.bminus[i] = D.bmag[i - 1];
      D.bplus[i] = D.bmag[i];
    } else {
      const double dsp = dist_scaled(D, i + 1, i);
      const double dsm = dist_scaled(D, i, i - 1);
      D.ds[i] = 0.5 * (dsp + dsm);
      D.bminus[i] = D.bmag[i - 1];
      D.bplus[i] = D.bmag[i + 1];
    }
    D.mfp_r[i] = (D.mfpType == 2) ? pow(D.mhdBsAu / D.bmag[i], D.mfpBPower)
                                  : pow(D.rmag[i] * D.rScale, D.mfpRadialPower);
    D.radial[i] = pow(D.rmag[i] / D.normRadius, -D.beta);
    // Thinning only happens if some neighbour spacing is <= dsh_min: with
    // every node kept, GetStreamList's test for shell k measures exactly
    // |r[k+1] - r[k]|.
    if (k < n - 1) {
      const double d = dist_scaled(D, i + 1, i);
      if (!(d > D.dsh_min)) s_thin = 1;
    }
  }
  __syncthreads();
  const int thin = s_thin;
  if (!thin) {
    for (int k = tid; k < n; k += blockDim.x) {
      D.list[o + k] = k;
      D.ref[o + k] = k;
    }
    if (tid == 0) s_size = n;
  } else if (tid == 0) {
    // Sequential GetStreamList (src/energeticParticles.c:2180-2228).
    int size = 0, down = 0;
    for (int shell = 0; shell < n; shell++) {
      int up;
      if (shell < n - 1) {
        up = shell + 1;
      } else {
        up = shell;
        down = shell - 1;
      }
      const double dsh = dist_scaled(D, o + up, o + down);
      if ((dsh > D.dsh_min) || (shell == n - 1)) {
        D.list[o + size] = shell;
        size++;
        D.ref[o + shell] = shell;
        down = shell + 1;
      } else {
        D.ref[o + shell] = shell + 1;
        int sp = shell - 1;
        while ((sp >= 0) && (D.ref[o + sp] != sp)) {
          D.ref[o + sp] = D.ref[o + sp + 1];
          sp--;
        }
      }
    }
    s_size = size;
    atomicMax(&D.diag->any_thin, 1);
  }
  __syncthreads();
  const int size = s_size;
  if (tid == 0) {
    D.thin[s] = thin;
    D.nkept[s] = size;
  }
  // FindSegmentLengths over the kept list.
  double lmin = 10000.0 * D.rScale;
  if (size >= 2) {
    for (int sl = tid; sl < size; sl += blockDim.x) {
      const size_t ts = o + D.list[o + sl];
      double dsl;
      if (sl == 0) {
        const double d = dist_scaled(D, o + D.list[o + sl + 1], ts);
        if (d < lmin) lmin = d;
        dsl = d;
      } else if (sl == size - 1) {
        const double d = dist_scaled(D, ts, o + D.list[o + sl - 1]);
        if (d < lmin) lmin = d;
        dsl = d;
      } else {
        const double d1 = dist_scaled(D, o + D.list[o + sl + 1], ts);
        const double d9 = dist_scaled(D, ts, o + D.list[o + sl - 1]);
        if (d1 < lmin) lmin = d1;
        if (d9 < lmin) lmin = d9;
        dsl = 0.5 * (d1 + d9);
      }
      D.dsil[o + sl] = 1.0 / dsl;
    }
  }
  const double dsMin = block_min_256(lmin, sh);
  if (tid == 0) {
    D.dsmin[s] = dsMin;
    if (size < 2 || (size < 3 && D.useBC))
      note_prep_error(D, err_key(0, 0, s, 0, kErrStreamShort, 0, 0, 0));
    else if (dsMin <= 0.0)
      note_prep_error(D, err_key(0, 0, s, 0, kErrDsUnderflow, 0, 0, 0));
  }
  // min_tau over the nodes the shell stage processes (1..n-1).
  double tmin = DBL_MAX;
  for (int k = 1 + tid; k < n; k += blockDim.x) {
    const double mr = D.mfp_r[o + k];
    for (int e = 0; e < D.E; e++) {
      const double tau = (D.rig_lamo[e] * mr) * D.vgrid_i[e];
      if (tau < tmin) tmin = tau;
    }
  }
  const double tm = block_min_256(tmin, sh);
  if (tid == 0 && tm < DBL_MAX) atomic_min_pos_double(&D.diag->min_tau_bits, tm);
}

// ------------------------------------------------------------------------
// k_task_prep: one thread per (stream, energy).  DiffuseStreamData's
// per-energy constants (src/energeticParticles.c:2301-2350) and the
// temporal-blocking tiling of the streaming passes.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_task_prep(Dev D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= D.S * D.E) return;
  const int s = t / D.E, e = t % D.E;
  const int size = D.nkept[s];
  const double dsMin = D.dsmin[s];
  int nsteps = 1;
  double dtProp = D.dt;
  if (dsMin > 0.0 && size >= 2) {
    const double dtMin = 0.4 * dsMin * D.vgrid_i[e];
    const double xs = D.dt / dtMin + 1.0;
    if (!(xs < (double)kMaxSubcycles)) {
      note_prep_error(D, err_key(0, 0, s, 0, kErrSubcycles, 0, 0, 0));
    } else {
      nsteps = (int)floor(xs);
      dtProp = D.dt / (1.0 * nsteps);
    }
  }
  int h, passes, chunks;
  task_tiling(D, size, nsteps, &h, &passes, &chunks);
  D.nsteps[t] = nsteps;
  D.halo[t] = h;
  D.passes[t] = passes;
  D.chunks[t] = chunks;
  D.dtprop[t] = dtProp;
  atomicMax(&D.diag->max_nsteps, nsteps);
  atomicMax(&D.diag->max_passes, passes);
  atomicMax(&D.diag->max_chunks, chunks);
  atomicAdd(&D.diag->stream_sweeps, (unsigned long long)nsteps * size * D.M);
  const double v = D.vgrid[e];
  for (int m = 0; m < D.M; m++) D.del_fac[(size_t)t * D.M + m] = v * dtProp * D.mugrid[m];
  const size_t o = (size_t)s * D.Npad;
  for (int j = 0; j < 3; j++)
    D.seed3[(size_t)t * 3 + j] = (j < size) ? seed_value(D, D.radial[o + D.list[o + j]], e) : 0.0;
}

// ------------------------------------------------------------------------
// k_stream_coef: per (stream, energy, kept index).  ds_i multiplier and
// the relaxation factor exp(-dtProp/tau) (src/energeticParticles.c:2320-2337).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stream_coef(Dev D) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.nkept[s]) return;
  const int t = s * D.E + e;
  const size_t o = (size_t)s * D.Npad;
  const int node = D.list[o + k];
  const double tau = D.vgrid_i[e] * (D.rig_lamo[e] * D.mfp_r[o + node]);
  const double dtProp = D.dtprop[t];
  const double dtProp_i = 1.0 / dtProp;
  const double dsi = D.dsil[o + k];
  const double dsm = (dtProp > tau) ? dsi * tau * dtProp_i : dsi * 1.0;
  D.dsm[(size_t)t * D.Npad + k] = dsm;
  D.q[(size_t)t * D.Npad + k] = exp(-dtProp / tau);
}

// ------------------------------------------------------------------------
// k_node_prep: one thread per node 1..n-1 (innerComputeShell skips node 0,
// src/mpiInit.c:105-109).  Focusing coefficients and subcycle count
// (src/energeticParticles.c:1581-1674) and adiabatic-change coefficients,
// subcycle count and inflow boundary values (src/energeticParticles.c:
// 859-949, :1137-1138).
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_node_prep(Dev D) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  const int n = D.n_active[s];
  if (k < 1 || k >= n) return;
  const size_t i = (size_t)s * D.Npad + k;
  const double dt = D.dt, dt_full = D.dt_full;
  const double half = 0.5, one = 1.0, two = 2.0, three = 3.0;
  // ---- focusing
  {
    const double DuParDt = D.dupar[i] / dt_full;
    const double b = two * (D.dlnn[i] / dt_full) - three * (D.dlnb[i] / dt_full);
    const double bp = D.bplus[i], bm = D.bminus[i];
    double dlnBds;
    if ((bp == 0.0) || (bm == 0.0))
      dlnBds = 0.0;
    else
      dlnBds = (log(bp) - log(bm)) / (two * D.ds[i] + kDblMin);
    double vmax = 0.0;
    for (int e = 0; e < D.E; e++) {
      const double v = D.vgrid[e];
      const double a = -v * dlnBds - (two / v) * DuParDt;
      double vprev = 0.0;
      for (int m = 0; m < D.M; m++) {
        const double mu = D.mugrid[m];
        const double vel = half * (one - mu * mu) * (a + mu * b);
        if (m > 0) {
          const double va = half * (vel + vprev);
          if (fabs(va) > vmax) vmax = fabs(va);
        }
        vprev = vel;
      }
    }
    const double dt_stable = 0.9 * D.dmu / vmax;
    const double xs = ceil(dt / dt_stable);
    int nsub = 0;
    if (!(xs < (double)kMaxSubcycles)) {
      note_prep_error(D, err_key(0, 0, s, 0, kErrSubcycles, 0, 0, 0));
    } else {
      nsub = (int)xs;
    }
    D.nsub_f[i] = nsub;
    D.dtsub_f[i] = nsub > 0 ? dt / nsub : 0.0;
    D.f_dlnbds[i] = dlnBds;
    D.f_dupar[i] = DuParDt;
    D.f_b[i] = b;
    atomicMax(&D.diag->max_sub_focus, nsub);
    atomicAdd(&D.diag->focus_sweeps, (unsigned long long)nsub * D.E * D.M);
    if (nsub >= kHeavyFocus) D.heavy[atomicAdd(&D.diag->focus_heavy, 1)] = (int32_t)i;
  }
  // ---- adiabatic change
  {
    const double DlnBDt = D.dlnb[i] / dt_full;
    const double DlnNDt = D.dlnn[i] / dt_full;
    const double damp = (DlnNDt < 0) ? 1.0 - D.dampCooling : 1.0;
    const double DuParDt = D.dupar[i] / dt_full;
    double vmax = 0.0;
    for (int m = 0; m < D.M; m++) {
      const double mu = D.mugrid[m];
      const double a = -mu * DuParDt;
      const double b = mu * mu * (DlnNDt - DlnBDt) + half * (one - mu * mu) * DlnBDt;
      for (int e = 0; e < D.E; e++) {
        const double vel = damp * (a * D.vgrid_i[e] + b);
        if (fabs(vel) > vmax) vmax = fabs(vel);
      }
    }
    const double dt_stable = 0.9 * D.dlnp / vmax;
    const double xs = ceil(dt / dt_stable);
    int nsub = 0;
    if (!(xs < (double)kMaxSubcycles)) {
      note_prep_error(D, err_key(0, 0, s, 0, kErrSubcycles, 0, 0, 0));
    } else {
      nsub = (int)xs;
    }
    D.nsub_c[i] = nsub;
    D.dtsub_c[i] = nsub > 0 ? dt / nsub : 0.0;
    D.c_dupar[i] = DuParDt;
    D.c_dlnn[i] = DlnNDt;
    D.c_dlnb[i] = DlnBDt;
    D.c_damp[i] = damp;
    const double rad = D.radial[i];
    D.c_bc0[i] = log(seed_value(D, rad, 0));
    D.c_bc1[i] = log(seed_value(D, rad, D.E - 1));
    atomicMax(&D.diag->max_sub_change, nsub);
    atomicAdd(&D.diag->change_sweeps, (unsigned long long)nsub * D.E * D.M);
  }
}

// ------------------------------------------------------------------------
// k_seed_check: seedPopulationCheck (src/energeticParticlesBoundary.c:
// 28-60) when streaming is off (otherwise it is fused into pass 0).
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_seed_check(Dev D) {
  if (substep_off(D)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.n_active[s]) return;
  const int t = s * D.E + e;
  const T sd = (T)seed_value(D, D.radial[(size_t)s * D.Npad + k], e);
  T *f = fslab<T>(D, D.loc[t], t) + k;
  for (int m = 0; m < D.M; m++) {
    const T v = f_dec(D, f[(size_t)m * D.Npad]);
    if (v < sd) f[(size_t)m * D.Npad] = f_enc(D, sd);
  }
}

// ------------------------------------------------------------------------
// AdiabaticFocusing of one mu row (src/energeticParticles.c:1676-1803,
// :1806-1896), shared by k_focus and the streaming passes.
// ------------------------------------------------------------------------
// The focusing update of one (node, energy) mu row held in registers:
// half-mesh velocities (src/energeticParticles.c:1613-1653), then nsub
// upwind flux subcycles with zero-flux walls, then the NaN/Inf check
// (first bad mu) -- written out in each kernel so that the row and the
// velocities stay in registers.
// ------------------------------------------------------------------------
// k_stream_finish: one block per (stream, energy).  Records which buffer
// now holds the slab and copies thinned nodes from their reference node,
// with the NaN/Inf check and DBL_MIN floor (src/energeticParticles.c:
// 2437-2471).
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_stream_finish(Dev D, int substep) {
  if (substep_off(D)) return;
  const int task = blockIdx.x;
  const int s = task / D.E, e = task % D.E;
  const int loc_new = D.loc[task] ^ (D.passes[task] & 1);
  __syncthreads();
  if (threadIdx.x == 0) D.loc[task] = (uint8_t)loc_new;
  // the slab queue of the substep's streaming launch, ready for the next
  if (task == 0 && threadIdx.x == 0) D.diag->queue = 0u;
  if (!D.thin[s]) return;
  const int n = D.n_active[s];
  const size_t on = (size_t)s * D.Npad;
  T *f = fslab<T>(D, loc_new, task);
  for (int idx = threadIdx.x; idx < n * D.M; idx += blockDim.x) {
    const int m = idx / n, shell = idx % n;
    const int r = D.ref[on + shell];
    if (r != shell) {
      T v = f_dec(D, f[(size_t)m * D.Npad + r]);
      if (isnan(v))
        note_error(D, err_key(substep, 0, s, 1, e, m, shell, 0));
      else if (isinf(v))
        note_error(D, err_key(substep, 0, s, 1, e, m, shell, 1));
      if (v < FloorT<T>::v) v = FloorT<T>::v;
      f[(size_t)m * D.Npad + shell] = f_enc(D, v);
    }
  }
}

// ------------------------------------------------------------------------
// k_focus: AdiabaticFocusing with the upwind operator
// (src/energeticParticles.c:1676-1803, :1806-1896).  A lane per node, a
// block per 256 nodes of one (stream, energy); the mu row of one (node,
// energy) and its half-mesh velocities stay in registers for all
// subcycles.  The upwind flux v*0.5*((1-cce)f[m]+(1+cce)f[m-1]) equals
// v*f_upwind exactly for finite f, and the zero-flux walls reproduce the
// operator's boundary rows.
// ------------------------------------------------------------------------
template <int M, typename T, bool FASTF, bool LOG = false>
__global__ __launch_bounds__(256) void k_focus(Dev D, int substep, int fused, int spare) {
  if (substep_off(D)) return;
  // blockIdx.x = ((stream, energy), chunk of 256 nodes), chunk fastest:
  // the block reads 2 KB runs of each mu row, consecutive blocks the next.
  // Blocks go round-robin to the 8 XCDs, so the chunk order is rotated by
  // (stream, energy): the inner chunk (most subcycles) then lands on every
  // XCD in turn instead of always on the same one.
  const int nch = (D.N + 255) >> 8;
  const int se = blockIdx.x / nch;
  const int s = se / D.E, e = se % D.E;
  const int chunk = (int)((blockIdx.x % nch + (unsigned)se) % nch);
  const int node = chunk * 256 + threadIdx.x;
  const int n = D.n_active[s];
  if (node < 1 || node >= n) return;
  if (fused && !D.thin[s]) return;  // done in the stream's last streaming pass
  const size_t i = (size_t)s * D.Npad + node;
  const int nsub = D.nsub_f[i];
  if (nsub >= kHeavyFocus) return;  // k_focus_heavy
  const double dtsub = D.dtsub_f[i];
  const double dlnBds = D.f_dlnbds[i], DuParDt = D.f_dupar[i], b = D.f_b[i];
  const double two = 2.0;
  const int task = s * D.E + e;
  constexpr uint32_t B = sizeof(T);
  // spare = 1 (tile-size search only): write the spare buffer, report nothing
  const int cur = D.loc[task];
  const rsrc_t slab = slab_rsrc((const double *)fslab<T>(D, cur, task), (uint32_t)(M * D.Npad * B));
  const rsrc_t out = slab_rsrc((const double *)fslab<T>(D, cur ^ spare, task), (uint32_t)(M * D.Npad * B));
  const uint32_t voff = (uint32_t)node * B, row = (uint32_t)D.Npad * B;
  const double v = D.vgrid[e];
  const double a = -v * dlnBds - (two / v) * DuParDt;
  T f[M];
#pragma unroll
  for (int m = 0; m < M; ++m) f[m] = f_decL<LOG>(buf_ld(slab, voff, m * row, T()));
  int bad = -1, bad_inf = 0;
  EP_FOCUS_ROW(T, M, D, a, b, f, nsub, focus_dtsub(D, task, dtsub), bad, bad_inf, FASTF);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const T val = (f[m] < FloorT<T>::v) ? FloorT<T>::v : f[m];
    if (nsub > 0 || val != f[m]) buf_st(f_encL<LOG>(val), out, voff, m * row);
  }
  if (bad >= 0 && !spare) note_error(D, err_key(substep, 1, s, 0, node, e, bad, bad_inf));
}

// Blocks of the heavy-focusing walk: enough waves to fill the chip, a
// multiple of the 8 XCDs (k_focus_heavy gives each XCD a range of the list).
constexpr int kHeavyGrid = 4096;
static_assert(kHeavyGrid % 8 == 0, "k_focus_heavy splits its grid over 8 XCDs");
// Longest heavy-node list k_focus_heavy4 takes (longer lists go to
// k_focus_heavy), and its grid: up to 4 waves per node.
constexpr int kHeavy4Max = 2048;
constexpr int kHeavy4Grid = 2048;
static_assert(kHeavy4Grid % 8 == 0, "k_focus_heavy4 splits its grid over 8 XCDs");

// k_focus_heavy: the nodes k_focus leaves out (nsub_f >= kHeavyFocus), one
// wave per node and one lane per energy, so every lane runs the node's
// subcycle count.  A fixed grid walks the list built by k_node_prep, whose
// entries come in runs of consecutive nodes of one stream.  A lane's loads
// are 8 B of a different 128-B line each (f is node-fastest), and 16
// consecutive nodes share those lines, so the list is split into one
// contiguous range per XCD (blocks are dispatched round-robin over the 8
// XCDs): neighbouring nodes then run at the same time on the same XCD and
// meet their lines in its L2 instead of every XCD fetching each line.
// The row and its split face velocities take ~250 VGPRs at M = 30; capped
// at 256 (two waves per SIMD, a few spilled scalars) the walk hides more of
// its load latency than with one wave per SIMD (4% faster on C2).
template <int M, typename T, bool FASTF, bool LOG = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_focus_heavy(Dev D, int substep) {
  if (substep_off(D)) return;
  const int count = D.diag->focus_heavy;
  if (M >= 12 && count <= kHeavy4Max) return;  // k_focus_heavy4
  const int e = threadIdx.x;
  const double two = 2.0;
  constexpr int kXcd = 8;
  const int per = gridDim.x / kXcd;  // the grid is a multiple of 8 blocks
  const int xcd = blockIdx.x % kXcd, local = blockIdx.x / kXcd;
  const int seg = (count + kXcd - 1) / kXcd;
  const int end = min(count, (xcd + 1) * seg);
  for (int h = xcd * seg + local; h < end; h += per) {
    const size_t i = (size_t)D.heavy[h];
    const int s = (int)(i / D.Npad), node = (int)(i % D.Npad);
    if (e >= D.E) continue;
    const int nsub = D.nsub_f[i];
    const double dtsub = D.dtsub_f[i];
    const double dlnBds = D.f_dlnbds[i], DuParDt = D.f_dupar[i], b = D.f_b[i];
    const int task = s * D.E + e;
    T *fp = fslab<T>(D, D.loc[task], task) + node;
    const double v = D.vgrid[e];
    const double a = -v * dlnBds - (two / v) * DuParDt;
    T f[M];
#pragma unroll
    for (int m = 0; m < M; ++m) f[m] = f_decL<LOG>(fp[(size_t)m * D.Npad]);
    int bad = -1, bad_inf = 0;
    EP_FOCUS_ROW(T, M, D, a, b, f, nsub, focus_dtsub(D, task, dtsub), bad, bad_inf, FASTF);
#pragma unroll
    for (int m = 0; m < M; ++m) fp[(size_t)m * D.Npad] = f_encL<LOG>((f[m] < FloorT<T>::v) ? FloorT<T>::v : f[m]);
    if (bad >= 0) note_error(D, err_key(substep, 1, s, 0, node, e, bad, bad_inf));
  }
}

// k_focus_heavy4: the same walk with four lanes per (node, energy) row
// (for M >= 12): lanes 4k .. 4k+3 of a wave hold mu cells q MG .. q MG +
// MG - 1 (q = l mod 4, MG = ceil(M/4), the last group padded) of energy
// eg 16 + k, and the wave takes the node's energies 16 at a time.  A
// subcycle's chain is then MG cells and two DPP row shifts instead of M
// cells, which is what bounds a node with hundreds of
// subcycles; a lane needs about a quarter of the registers.  The
// arithmetic is k_focus_change4's (face velocities split by split_face,
// interior faces fma(vh, f_above, vl * f_below), the wall rows' faces the
// upwind cell alone), so the values are those of k_focus_heavy and k_focus
// bit for bit.  Units (node, group of 16 energies) are walked in list
// order, one contiguous range per XCD, as in k_focus_heavy.
template <int M, typename T, bool FASTF, bool LOG = false>
__global__ __launch_bounds__(64) void k_focus_heavy4(Dev D, int substep) {
  if (substep_off(D)) return;
  constexpr int G = 4, W = 64 / G;
  constexpr int MG = (M + G - 1) / G;
  static_assert(M >= 12, "four lanes per row from 12 mu cells");
  constexpr int QL = (M - 1) / MG;       // group of cell M-1
  constexpr int JL = (M - 1) - QL * MG;  // its slot
  static_assert(QL == G - 1, "cell M-1 in the last group");
  __shared__ double smu[G * MG + 2], shc[G * MG + 2], s2v[kMaxE], sv[kMaxE];
  const int E = D.E;
  if (threadIdx.x < G * MG + 2) {
    const int m = threadIdx.x;
    const double mu = (m < M) ? D.mugrid[m] : 0.0;
    smu[m] = mu;
    shc[m] = 0.5 * (1.0 - mu * mu);
  }
  if (threadIdx.x < E) {
    sv[threadIdx.x] = D.vgrid[threadIdx.x];
    s2v[threadIdx.x] = 2.0 / D.vgrid[threadIdx.x];
  }
  __syncthreads();
  const int lane = threadIdx.x;
  const int q = lane % G, m0 = q * MG;
  const int EW = (E + W - 1) / W;  // waves of 16 rows per node
  if (D.diag->focus_heavy > kHeavy4Max) return;  // k_focus_heavy
  const int count = D.diag->focus_heavy * EW;
  constexpr int kXcd = 8;
  const int per = gridDim.x / kXcd;  // the grid is a multiple of 8 blocks
  const int xcd = blockIdx.x % kXcd, local = blockIdx.x / kXcd;
  const int seg = (count + kXcd - 1) / kXcd;
  const int end = min(count, (xcd + 1) * seg);
  auto live = [&](int j) { return (j <= JL) || (q < G - 1); };
  const double half = 0.5;
  const T dmu_i = (T)(1.0 / D.dmu);
  for (int u = xcd * seg + local; u < end; u += per) {
    const size_t i = (size_t)D.heavy[u / EW];
    const int e = (u % EW) * W + lane / G;
    const bool ok = e < E;  // the four lanes of a row agree
    const int s = (int)(i / D.Npad), node = (int)(i % D.Npad);
    const int nsub = D.nsub_f[i];  // the node's: the same in every lane
    const double dlnBds = D.f_dlnbds[i], DuParDt = D.f_dupar[i], fb = D.f_b[i];
    const int ec = ok ? e : 0;
    const int task = s * E + ec;
    T *fp = fslab<T>(D, D.loc[task], task) + node + (size_t)m0 * D.Npad;
    T x[MG];
#pragma unroll
    for (int j = 0; j < MG; ++j) x[j] = (ok && live(j)) ? f_decL<LOG>(fp[(size_t)j * D.Npad]) : T(1);
    const double v = sv[ec];
    const double a = -v * dlnBds - s2v[ec] * DuParDt;
    const T dts = (T)focus_dtsub(D, task, D.dtsub_f[i]);
    const T hsc = (T)half * (dts * dmu_i);
    T vl[MG + 1], vh[MG + 1];  // face below slot j (j = MG: above the last slot)
    double vp = (m0 > 0) ? shc[m0 - 1] * (a + smu[m0 - 1] * fb) : 0.0;
#pragma unroll
    for (int j = 0; j <= MG; ++j) {
      const int m = m0 + j;
      const double vel = shc[m] * (a + smu[m] * fb);
      // fused modes: velocities in units of dmu / dt_sub; (vel + vp) (sc/2)
      // rounds to the same value as ((vel + vp)/2) sc
      const T va = FASTF ? (T)(vel + vp) * hsc : (T)(half * (vel + vp));
      vp = vel;
      split_face(va, vl[j], vh[j]);
    }
    for (int it = 0; it < nsub; ++it) {
      // cell m0-1 (last of the group below), cell m0+MG (first above)
      const T below = dpp_row<kRowShr1>(x[MG - 1]);
      const T above = dpp_row<kRowShl1>(x[0]);
      // face below slot 0: zero-flux wall under cell 0, else interior (or,
      // when cell M-1 is slot 0, that row's upwind-only face)
      T flux_lo = fma(vh[0], x[0], vl[0] * below);
      if (JL == 0) {
        const T ub = (vl[0] + vh[0]) * (signbit(vh[0]) ? x[0] : below);
        flux_lo = (q == QL) ? ub : flux_lo;
      }
      flux_lo = (q == 0) ? (T)0 : flux_lo;
      const T ptop = fma(vh[MG], above, vl[MG] * x[MG - 1]);
#pragma unroll
      for (int j = 0; j < MG; ++j) {
        T flux_hi, next_lo;
        if (j + 1 < MG) {
          const T pf = fma(vh[j + 1], x[j + 1], vl[j + 1] * x[j]);
          flux_hi = pf;
          next_lo = pf;
          if (j == 0 || j + 1 == JL) {
            const T uf = (vl[j + 1] + vh[j + 1]) * (signbit(vh[j + 1]) ? x[j + 1] : x[j]);
            if (j == 0) flux_hi = (q == 0) ? uf : pf;         // row 0's upper face
            if (j + 1 == JL) next_lo = (q == QL) ? uf : pf;   // row M-1's lower face
          }
        } else {
          flux_hi = ptop;
          next_lo = (T)0;
        }
        if (j >= JL) flux_hi = (q == QL) ? (T)0 : flux_hi;  // zero-flux wall above cell M-1
        if (FASTF) {
          x[j] = x[j] - (flux_hi - flux_lo);
        } else {
          const T dfl = (flux_hi - flux_lo) * dmu_i;
          x[j] = x[j] - dts * dfl;
        }
        flux_lo = next_lo;
      }
    }
    if (!ok) continue;
    int bad = -1, bad_inf = 0;
#pragma unroll
    for (int j = 0; j < MG; ++j)
      if (live(j)) {
        if (bad < 0 && !isfinite(x[j])) {
          bad = m0 + j;
          bad_inf = isinf(x[j]) ? 1 : 0;
        }
        fp[(size_t)j * D.Npad] = f_encL<LOG>((x[j] < FloorT<T>::v) ? FloorT<T>::v : x[j]);
      }
    if (bad >= 0) note_error(D, err_key(substep, 1, s, 0, node, e, bad, bad_inf));
  }
}

// The heavy-node walk.  From 12 mu cells both kernels are launched and the
// list length, known only on the device, picks one: a short list (the
// step's time is then the chain of its slowest node) runs four lanes per
// row, a long one (throughput: four lanes per row issue about twice the
// instructions per cell) a lane per row.
template <int M, typename T, bool FASTF, bool LOG = false>
inline void launch_heavy_t(const Dev &D, int sub, hipStream_t st) {
  if constexpr (M >= 12)
    hipLaunchKernelGGL((k_focus_heavy4<M, T, FASTF, LOG>), dim3(kHeavy4Grid), dim3(64), 0, st, D, sub);
  hipLaunchKernelGGL((k_focus_heavy<M, T, FASTF, LOG>), dim3(kHeavyGrid), dim3(64), 0, st, D, sub);
}

// ------------------------------------------------------------------------
// k_change: AdiabaticChange with the upwind operator, advecting ln f in
// ln p (src/energeticParticles.c:859-1067, :1081-1176).  A lane per node,
// a block per 256 nodes of one (stream, mu).  Each subcycle is one
// sweep up the energy column with a sliding window (the new ln f at e
// needs the old values at e-1, e, e+1), reading ahead kPf energies to keep
// loads in flight; intermediate ln f is kept in place in the f slab, the
// first sweep takes the log and the last one writes exp back.  A lane
// needs a handful of registers, so the kernel runs at full occupancy.
// ------------------------------------------------------------------------
// FM: ln f and exp through fm_log / fm_exp (tables in LDS).  LOG: the
// float storage holds ln f, which the sweep reads and writes as it is.
template <typename T, bool FM, bool LOG = false>
__global__ __launch_bounds__(256) void k_change(Dev D, int substep) {
  if (substep_off(D)) return;
  // blockIdx.x = ((stream, mu), chunk of 256 nodes), chunk fastest
  const int nch = (D.N + 255) >> 8;
  const int sm = blockIdx.x / nch;
  const int s = sm / D.M, m = sm % D.M;
  const int node = (blockIdx.x % nch) * 256 + threadIdx.x;
  const int n = D.n_active[s];
  const int E = D.E;
  // the buffer each energy's slab sits in, read once per block (the slab
  // addresses are then known ahead and the loads below can be issued early)
  __shared__ uint8_t sloc[kMaxE];
  __shared__ double fmt[FM ? kFmTable : 1];
  if (threadIdx.x < E) sloc[threadIdx.x] = D.loc[s * E + threadIdx.x];
  if (FM)
    for (int q = threadIdx.x; q < kFmTable; q += blockDim.x) fmt[q] = D.fm_tab[q];
  __syncthreads();
  if (node < 1 || node >= n) return;
  const size_t i = (size_t)s * D.Npad + node;
  const int nsub = D.nsub_c[i];
  const double dtsub = D.dtsub_c[i];
  const double DuParDt = D.c_dupar[i], DlnNDt = D.c_dlnn[i], DlnBDt = D.c_dlnb[i];
  const double damp = D.c_damp[i], bc0 = D.c_bc0[i], bc1 = D.c_bc1[i];
  const double dlnp_i = 1.0 / D.dlnp;
  const double half = 0.5, one = 1.0;
  const size_t slab = (size_t)D.M * D.Npad;
  constexpr uint32_t B = sizeof(T);
  const uint32_t slab_bytes = (uint32_t)(slab * B);
  const uint32_t voff = (uint32_t)(((size_t)m * D.Npad + node) * B);
  const double mu = D.mugrid[m];
  const double a = -mu * DuParDt;
  const double bb = mu * mu * (DlnNDt - DlnBDt) + half * (one - mu * mu) * DlnBDt;
  const double *__restrict__ vgi = D.vgrid_i;
  const bool upd = nsub > 0;
  const int niter = upd ? nsub : 1;  // nsub = 0 still takes exp(log f)
  auto slab_of = [&](int e) {
    const int task = s * E + e;
    const int lc = __builtin_amdgcn_readfirstlane((int)sloc[e]);
    return slab_rsrc((const double *)fslab<T>(D, lc, task), slab_bytes);
  };
  // With float storage, ln f between subcycles is kept in FP64 in the part
  // of the f allocations that float storage leaves free.
  constexpr bool kSpare = sizeof(T) != sizeof(double);
  const uint32_t voff8 = (uint32_t)(((size_t)m * D.Npad + node) * 8);
  auto spare_of = [&](int e) { return slab_rsrc(spare_slab(D, s * E + e), (uint32_t)(slab * 8)); };
  int bad = -1, bad_inf = 0;
  // One sweep up the column; first: take ln f of the stored values, fin:
  // write exp back (both per lane: per-node subcycle counts differ).  The
  // slab descriptors are formed outside any branch on them, so that they
  // stay in scalar registers.
  auto sweep = [&](const bool first, const bool fin) __attribute__((always_inline)) {
    // the column is advanced in FP64 whatever the storage type
    auto ld = [&](int e) {
      const rsrc_t rs = slab_of(e);
      if (kSpare) {
        const rsrc_t rp = spare_of(e);
        const double a = buf_ld(rp, voff8, 0);
        const double b = (double)buf_ld(rs, voff, 0, T());
        return first ? b : a;
      }
      return (double)buf_ld(rs, voff, 0, T());
    };
    auto cvt = [&](double x) { return (first && !LOG) ? (FM ? fm_log(x, fmt) : log(x)) : x; };
    double r0 = cvt(ld(0));
    double r1 = cvt(ld(1));
    // read-ahead: kPf energies beyond the window are in flight
    constexpr int kPf = 3;
    double pf[kPf];
#pragma unroll
    for (int q = 0; q < kPf; ++q) pf[q] = (2 + q < E) ? ld(2 + q) : 0.0;
    double old_prev = 0.0;  // ln f[e-1] before this subcycle
    double vel = damp * (a * vgi[0] + bb);
    double vel_prev = 0.0;
    double flux_e = 0.0;  // flux at interface e (between e-1 and e)
    for (int e = 0; e < E; ++e) {
      const double cur = r0;
      const double nxt = (e + 1 < E) ? r1 : 0.0;
      const double vel1 = (e + 1 < E) ? damp * (a * vgi[e + 1] + bb) : 0.0;
      double f1;
      if (e == 0) {
        const double va1 = half * (vel1 + vel);
        if (va1 <= 0)
          f1 = va1 * (nxt - cur) * dlnp_i;
        else
          f1 = va1 * (cur - bc0) * dlnp_i;
        flux_e = ep_poison(signbit(va1) ? cur : nxt, va1 * (signbit(va1) ? nxt : cur));  // interface 1
      } else if (e == E - 1) {
        const double vr = half * (vel + vel_prev);
        if (vr >= 0)
          f1 = vr * (cur - old_prev) * dlnp_i;
        else
          f1 = vr * (bc1 - cur) * dlnp_i;
      } else {
        const double va1 = half * (vel1 + vel);
        const double flux_hi = ep_poison(signbit(va1) ? cur : nxt, va1 * (signbit(va1) ? nxt : cur));
        f1 = (flux_hi - flux_e) * dlnp_i;
        flux_e = flux_hi;
      }
      const double nv = upd ? cur - dtsub * f1 : cur;
      const rsrc_t rs = slab_of(e);
      if (fin && LOG) {
        // f = exp(nv) is finite in FP32 below ln FLT_MAX; the floor in ln f
        if (bad < 0 && !(nv < kLnFltMax)) {
          bad = e;
          bad_inf = isnan(nv) ? 0 : 1;
        }
        buf_st((T)fmax(nv, kLnFloorF), rs, voff, 0);
      } else if (fin) {
        T val = (T)(FM ? fm_exp(nv, fmt) : exp(nv));
        if (bad < 0 && !isfinite(val)) {
          bad = e;
          bad_inf = isinf(val) ? 1 : 0;
        }
        if (val < FloorT<T>::v) val = FloorT<T>::v;
        buf_st(val, rs, voff, 0);
      } else if (kSpare) {
        const rsrc_t rp = spare_of(e);
        buf_st(nv, rp, voff8, 0);
      } else {
        buf_st((T)nv, rs, voff, 0);
      }
      old_prev = cur;
      vel_prev = vel;
      vel = vel1;
      r0 = r1;
      r1 = (e + 2 < E) ? cvt(pf[0]) : 0.0;
#pragma unroll
      for (int q = 0; q + 1 < kPf; ++q) pf[q] = pf[q + 1];
      pf[kPf - 1] = (e + 2 + kPf < E) ? ld(e + 2 + kPf) : 0.0;
    }
  };
  if (__all(niter == 1)) {
    sweep(true, true);  // the usual case: one sweep per substep everywhere in the wave
  } else {
    for (int it = 0; it < niter; ++it) sweep(it == 0, it == niter - 1);
  }
  if (bad >= 0) note_error(D, err_key(substep, 1, s, 1, node, bad, m, bad_inf));
}

// ------------------------------------------------------------------------
// Layout conversion and initialisation.
// ------------------------------------------------------------------------
// host order [s][node][e][m] (staged in `stage`) -> device [s][e][m][node]
template <typename T>
__global__ __launch_bounds__(256) void k_from_reference(Dev D, const double *stage, T *dstbuf, int s0) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = s0 + blockIdx.z;
  if (node >= D.N) return;
  const size_t src = (((size_t)blockIdx.z * D.N + node) * D.E + e) * D.M;
  for (int m = 0; m < D.M; m++)
    dstbuf[fidx(D, s, e, m, node)] = (sizeof(T) == 4 && D.logf) ? (T)log(stage[src + m]) : (T)stage[src + m];
}

template <typename T>
__global__ __launch_bounds__(256) void k_to_reference(Dev D, const T *srcbuf, double *stage, int s0) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = s0 + blockIdx.z;
  if (node >= D.N) return;
  const size_t dst = (((size_t)blockIdx.z * D.N + node) * D.E + e) * D.M;
  // decoded in FP64: ln f in mode 3 may stand for values below FLT_MIN
  for (int m = 0; m < D.M; m++) {
    const T raw = srcbuf[fidx(D, s, e, m, node)];
    stage[dst + m] = (sizeof(T) == 4 && D.logf) ? exp((double)raw) : (double)raw;
  }
}

// Move every slab whose current copy sits in buffer B back to A.
template <typename T>
__global__ __launch_bounds__(256) void k_normalize(Dev D) {
  const int task = blockIdx.x;
  if (D.loc[task] == 0) return;
  const size_t off = (size_t)task * D.M * D.Npad;
  T *a = fbufT<T>(D, 0);
  const T *b = fbufT<T>(D, 1);
  for (size_t i = threadIdx.x; i < (size_t)D.M * D.Npad; i += blockDim.x) a[off + i] = b[off + i];
  __syncthreads();
  if (threadIdx.x == 0) D.loc[task] = 0;
}

__global__ __launch_bounds__(256) void k_fill(double *p, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// initEnergeticParticles (src/energeticParticlesInit.c:106-147).
template <typename T>
__global__ __launch_bounds__(256) void k_init_seed(Dev D) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.n_active[s]) return;
  const size_t i = (size_t)s * D.Npad + k;
  const double rad = pow(D.rmag[i] / D.normRadius, -D.beta);
  const T sd = (T)seed_value(D, rad, e);
  const int t = s * D.E + e;
  T *f = fslab<T>(D, D.loc[t], t) + k;
  for (int m = 0; m < D.M; m++) f[(size_t)m * D.Npad] = f_enc(D, sd);
}

// Storage-type change of every slab: the current copy in buffer loc is
// converted into the other buffer, which becomes current.
// logf: the float side holds ln f (precision mode 3); the FP64 side
// always holds f.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_convert(Dev D, int logf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.Npad) return;
  const int t = s * D.E + e;
  const TI *src = fbufT<TI>(D, D.loc[t]) + (size_t)t * D.M * D.Npad;
  TO *dst = fbufT<TO>(D, D.loc[t] ^ 1) + (size_t)t * D.M * D.Npad;
  for (int m = 0; m < D.M; m++) {
    double v = (double)src[(size_t)m * D.Npad + k];
    if (logf) v = (sizeof(TI) == 4) ? exp(v) : log(v);
    dst[(size_t)m * D.Npad + k] = (TO)v;
  }
}

__global__ __launch_bounds__(256) void k_flip_all(Dev D) {
  if (substep_off(D)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < D.S * D.E) D.loc[t] ^= 1;
}

// ------------------------------------------------------------------------
// Host-side launchers.
// ------------------------------------------------------------------------

hipError_t launch_task_prep(const Dev &D, hipStream_t st) {
  hipLaunchKernelGGL(k_task_prep, dim3((D.S * D.E + 255) / 256), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_stream_prep(const Dev &D, hipStream_t st) {
  hipLaunchKernelGGL(k_stream_prep, dim3(D.S), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_task_prep, dim3((D.S * D.E + 255) / 256), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_stream_coef, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_node_prep(const Dev &D, hipStream_t st) {
  hipLaunchKernelGGL(k_node_prep, dim3((D.N + 255) / 256, D.S), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_seed_check(const Dev &D, hipStream_t st) {
  if (D.fp32)
    hipLaunchKernelGGL(k_seed_check<float>, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, st, D);
  else
    hipLaunchKernelGGL(k_seed_check<double>, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

// The streaming passes are compiled in four translation units by mu count
// (ep_stream_m*.hip) so that the build runs them in parallel.
hipError_t launch_stream_pass(const Dev &D, int pass, int flip, int do_seed, int sub, int chunks, int fuse,
                              hipStream_t st) {
  if (D.M < 2 || D.M > 32) return hipErrorInvalidValue;
  if (D.M <= 9) return launch_stream_pass_m2_9(D, pass, flip, do_seed, sub, chunks, fuse, st);
  if (D.M <= 17) return launch_stream_pass_m10_17(D, pass, flip, do_seed, sub, chunks, fuse, st);
  if (D.M <= 25) return launch_stream_pass_m18_25(D, pass, flip, do_seed, sub, chunks, fuse, st);
  return launch_stream_pass_m26_32(D, pass, flip, do_seed, sub, chunks, fuse, st);
}

hipError_t launch_stream_slab(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st) {
  if (D.M < 2 || D.M > 32) return hipErrorInvalidValue;
  if (D.M <= 9) return launch_stream_slab_m2_9(D, do_seed, sub, fuse, tune, one, st);
  if (D.M <= 17) return launch_stream_slab_m10_17(D, do_seed, sub, fuse, tune, one, st);
  if (D.M <= 25) return launch_stream_slab_m18_25(D, do_seed, sub, fuse, tune, one, st);
  return launch_stream_slab_m26_32(D, do_seed, sub, fuse, tune, one, st);
}

__global__ void k_reset_queue(DevDiag *d) { d->queue = 0u; }

hipError_t launch_reset_queue(const Dev &D, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(1), 0, st, D.diag);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_stream_finish(const Dev &D, int sub, hipStream_t st) {
  if (D.fp32)
    hipLaunchKernelGGL(k_stream_finish<float>, dim3(D.S * D.E), dim3(256), 0, st, D, sub);
  else
    hipLaunchKernelGGL(k_stream_finish<double>, dim3(D.S * D.E), dim3(256), 0, st, D, sub);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

#define EP_FOCUS_CASE(MM)                                                                     \
  case MM:                                                                                    \
    if (D.logf) {                                                                             \
      hipLaunchKernelGGL((k_focus<MM, float, true, true>), dim3((D.N + 255) / 256 * D.S * D.E), dim3(256), \
                         0, st, D, sub, fused, spare);                                        \
      if (!spare && heavy)                                                                    \
        launch_heavy_t<MM, float, true, true>(D, sub, st); \
    } else if (D.fp32) {                                                                      \
      hipLaunchKernelGGL((k_focus<MM, float, true>), dim3((D.N + 255) / 256 * D.S * D.E), dim3(256), 0, \
                         st, D, sub, fused, spare);                                           \
      if (!spare && heavy)                                                                    \
        launch_heavy_t<MM, float, true>(D, sub, st); \
    } else if (D.fast) {                                                                      \
      hipLaunchKernelGGL((k_focus<MM, double, true>), dim3((D.N + 255) / 256 * D.S * D.E), dim3(256), 0, \
                         st, D, sub, fused, spare);                                           \
      if (!spare && heavy)                                                                    \
        launch_heavy_t<MM, double, true>(D, sub, st); \
    } else {                                                                                  \
      hipLaunchKernelGGL((k_focus<MM, double, false>), dim3((D.N + 255) / 256 * D.S * D.E), dim3(256), 0, \
                         st, D, sub, fused, spare);                                           \
      if (!spare && heavy)                                                                    \
        launch_heavy_t<MM, double, false>(D, sub, st); \
    }                                                                                         \
    break;


hipError_t launch_focus(const Dev &D, int sub, int fused, hipStream_t st, int spare, int heavy) {
  switch (D.M) {
    EP_FOCUS_CASE(2) EP_FOCUS_CASE(3) EP_FOCUS_CASE(4) EP_FOCUS_CASE(5) EP_FOCUS_CASE(6)
    EP_FOCUS_CASE(7) EP_FOCUS_CASE(8) EP_FOCUS_CASE(9) EP_FOCUS_CASE(10) EP_FOCUS_CASE(11)
    EP_FOCUS_CASE(12) EP_FOCUS_CASE(13) EP_FOCUS_CASE(14) EP_FOCUS_CASE(15) EP_FOCUS_CASE(16)
    EP_FOCUS_CASE(17) EP_FOCUS_CASE(18) EP_FOCUS_CASE(19) EP_FOCUS_CASE(20) EP_FOCUS_CASE(21)
    EP_FOCUS_CASE(22) EP_FOCUS_CASE(23) EP_FOCUS_CASE(24) EP_FOCUS_CASE(25) EP_FOCUS_CASE(26)
    EP_FOCUS_CASE(27) EP_FOCUS_CASE(28) EP_FOCUS_CASE(29) EP_FOCUS_CASE(30) EP_FOCUS_CASE(31)
    EP_FOCUS_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

#define EP_HEAVY_CASE(MM)                                                                     \
  case MM:                                                                                    \
    if (D.logf)                                                                               \
      launch_heavy_t<MM, float, true, true>(D, sub, st); \
    else if (D.fp32)                                                                          \
      launch_heavy_t<MM, float, true>(D, sub, st); \
    else if (D.fast)                                                                          \
      launch_heavy_t<MM, double, true>(D, sub, st); \
    else                                                                                      \
      launch_heavy_t<MM, double, false>(D, sub, st); \
    break;

hipError_t launch_focus_heavy(const Dev &D, int sub, hipStream_t st) {
  switch (D.M) {
    EP_HEAVY_CASE(2) EP_HEAVY_CASE(3) EP_HEAVY_CASE(4) EP_HEAVY_CASE(5) EP_HEAVY_CASE(6)
    EP_HEAVY_CASE(7) EP_HEAVY_CASE(8) EP_HEAVY_CASE(9) EP_HEAVY_CASE(10) EP_HEAVY_CASE(11)
    EP_HEAVY_CASE(12) EP_HEAVY_CASE(13) EP_HEAVY_CASE(14) EP_HEAVY_CASE(15) EP_HEAVY_CASE(16)
    EP_HEAVY_CASE(17) EP_HEAVY_CASE(18) EP_HEAVY_CASE(19) EP_HEAVY_CASE(20) EP_HEAVY_CASE(21)
    EP_HEAVY_CASE(22) EP_HEAVY_CASE(23) EP_HEAVY_CASE(24) EP_HEAVY_CASE(25) EP_HEAVY_CASE(26)
    EP_HEAVY_CASE(27) EP_HEAVY_CASE(28) EP_HEAVY_CASE(29) EP_HEAVY_CASE(30) EP_HEAVY_CASE(31)
    EP_HEAVY_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

// AdiabaticChange over every (stream, mu, node); the fused precision
// modes (fast, or float storage) use the table-driven exp/log.
hipError_t launch_change(const Dev &D, int sub, hipStream_t st) {
  const dim3 grid((D.N + 255) / 256 * D.S * D.M), block(256);
  if (D.logf)
    hipLaunchKernelGGL((k_change<float, true, true>), grid, block, 0, st, D, sub);
  else if (D.fp32)
    hipLaunchKernelGGL((k_change<float, true>), grid, block, 0, st, D, sub);
  else if (D.fast)
    hipLaunchKernelGGL((k_change<double, true>), grid, block, 0, st, D, sub);
  else
    hipLaunchKernelGGL((k_change<double, false>), grid, block, 0, st, D, sub);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_from_reference(const Dev &D, const double *stage, double *dstbuf, hipStream_t st, int s0,
                                 int count) {
  if (count < 0) count = D.S - s0;
  const dim3 grid((D.N + 255) / 256, D.E, count);
  if (D.fp32)
    hipLaunchKernelGGL(k_from_reference<float>, grid, dim3(256), 0, st, D, stage, (float *)dstbuf, s0);
  else
    hipLaunchKernelGGL(k_from_reference<double>, grid, dim3(256), 0, st, D, stage, dstbuf, s0);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_to_reference(const Dev &D, const double *srcbuf, double *stage, hipStream_t st, int s0,
                               int count) {
  if (count < 0) count = D.S - s0;
  const dim3 grid((D.N + 255) / 256, D.E, count);
  if (D.fp32)
    hipLaunchKernelGGL(k_to_reference<float>, grid, dim3(256), 0, st, D, (const float *)srcbuf, stage, s0);
  else
    hipLaunchKernelGGL(k_to_reference<double>, grid, dim3(256), 0, st, D, srcbuf, stage, s0);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_normalize(const Dev &D, hipStream_t st) {
  if (D.fp32)
    hipLaunchKernelGGL(k_normalize<float>, dim3(D.S * D.E), dim3(256), 0, st, D);
  else
    hipLaunchKernelGGL(k_normalize<double>, dim3(D.S * D.E), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_flip_all(const Dev &D, hipStream_t st) {
  hipLaunchKernelGGL(k_flip_all, dim3((D.S * D.E + 255) / 256), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_convert(const Dev &D, int to_fp32, int logf, hipStream_t st) {
  const dim3 grid((D.Npad + 255) / 256, D.E, D.S);
  if (to_fp32)
    hipLaunchKernelGGL((k_convert<double, float>), grid, dim3(256), 0, st, D, logf);
  else
    hipLaunchKernelGGL((k_convert<float, double>), grid, dim3(256), 0, st, D, logf);
  EP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_flip_all, dim3((D.S * D.E + 255) / 256), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, st, p, n, v);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_init_seed(const Dev &D, hipStream_t st) {
  if (D.fp32)
    hipLaunchKernelGGL(k_init_seed<float>, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, st, D);
  else
    hipLaunchKernelGGL(k_init_seed<double>, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, st, D);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

// Plain streaming copy, 16 bytes per lane and iteration: the reference
// point for the HBM byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const double2 *src, double2 *dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, st, (const double2 *)src, (double2 *)dst,
                     bytes / 16);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

__global__ __launch_bounds__(256) void k_fastmath_probe(const double *tab, const double *x, int64_t n,
                                                        double *lg, double *ex) {
  __shared__ double t[kFmTable];
  for (int i = threadIdx.x; i < kFmTable; i += blockDim.x) t[i] = tab[i];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lg[i] = fm_log(x[i], t);
  ex[i] = fm_exp(x[i], t);
}

hipError_t launch_fastmath_probe(const double *tab, const double *x, int64_t n, double *lg, double *ex,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_fastmath_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tab, x, n, lg, ex);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_dpp_probe(const double *in, double *out, double *out2, hipStream_t st) {
  hipLaunchKernelGGL(k_dpp_probe, dim3(1), dim3(64), 0, st, in, out, out2);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}

}  // namespace ep

namespace ep {
__global__ void k_reset_diag(DevDiag *d) {
  d->max_sub_focus = 0;
  d->max_sub_change = 0;
  d->max_nsteps = 0;
  d->max_passes = 0;
  d->max_chunks = 0;
  d->any_thin = 0;
  d->min_tau_bits = (unsigned long long)__double_as_longlong(DBL_MAX);
  d->err_key = ~0ull;
  d->stream_sweeps = 0;
  d->focus_sweeps = 0;
  d->change_sweeps = 0;
  d->focus_heavy = 0;
  d->queue = 0u;
  d->prep_err = 0;
}

hipError_t launch_reset_diag(DevDiag *d, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_diag, dim3(1), dim3(1), 0, st, d);
  EP_CHECK_LAUNCH();
  return hipSuccess;
}
}  // namespace ep
