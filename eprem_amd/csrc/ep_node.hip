// ep_node.hip -- AdiabaticFocusing and AdiabaticChange of a node in one
// pass over f.
//
// The reference applies both operators node by node: for each stream,
// AdiabaticFocusing (src/energeticParticles.c:1533-1803, upwind operator
// :1806-1896) and then AdiabaticChange (:805-1077, upwind operator
// :1081-1176), each reading and writing the node's energy x mu tile.  Both
// are local to a node, so this kernel runs them back to back on each node
// while its tile streams through registers: f is read once and written once
// per substep instead of twice each.
//
// Layout below 12 mu cells (k_focus_change): a lane pair per node.  Lanes l
// and l+32 of a wave hold the lower and upper halves of the mu row
// (MH = ceil(M/2) cells each) of node (l mod 32); a 256-lane block covers
// 128 consecutive nodes of one stream.  From 12 mu cells k_focus_change4
// (further down) spreads a node over four lanes instead; the walk over the
// energies is the same.
// Energies are walked upwards.  For each energy the kernel loads the row,
// runs the node's focusing subcycles (the one interface between the halves
// is exchanged across the lane pair each subcycle), takes ln f, and then
// finishes the adiabatic change of the energy below it: with one change
// subcycle (the usual case) the new ln f at e needs the focused ln f at
// e-1, e and e+1 and the flux through the interface below e, which is all a
// lane keeps between energies (ln f and flux of the previous energy,
// 2 * MH doubles).
//
// Arithmetic is operation for operation the one of k_focus and k_change
// (same association, same casts), so the results are bit-identical to the
// two-kernel form.  The kernel is used when every node of the step needs at
// most one change subcycle; nodes with kHeavyFocus or more focusing
// subcycles are focused beforehand by k_focus_heavy and only changed here.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "ep_fastmath.h"
#include "ep_internal.h"
#include "ep_launch.h"

namespace ep {

template <typename T>
__device__ __forceinline__ T pair_swap(T v) {
  return __shfl_xor(v, 32, 64);
}

// FM: ln f and exp through fm_log / fm_exp (tables in LDS).
// fused: focusing of full-resolution streams was done by their last
// streaming pass.  spare: read the current buffers and write the other one
// (tile-size search; f and the step state are left as they were).
template <int M, typename T, bool FM>
__global__ __launch_bounds__(256) void k_focus_change(Dev D, int substep, int fused, int spare) {
  if (substep_off(D)) return;
  constexpr int MH = (M + 1) / 2;
  // blockIdx.x = (stream, chunk of 128 nodes), chunk fastest; the chunk
  // order is rotated by stream so that the inner chunks (most focusing
  // subcycles) spread over the XCDs instead of landing on the same one
  const int nch = (D.N + 127) >> 7;
  const int s = blockIdx.x / nch;
  const int chunk = (int)((blockIdx.x % nch + (unsigned)s) % nch);
  const int lane = threadIdx.x & 63;
  const int hf = lane >> 5;
  const int node = chunk * 128 + (threadIdx.x >> 6) * 32 + (lane & 31);
  const int E = D.E;
  __shared__ uint8_t sloc[kMaxE];
  // per mu (padded to 2 MH + 1 with zeros): mu, mu^2, 0.5 (1 - mu^2)
  __shared__ double smu[2 * MH + 2], smu2[2 * MH + 2], shc[2 * MH + 2];
  __shared__ double fmt[FM ? kFmTable : 1];
  // per energy: v and 2/v (the same correctly rounded quotient each lane
  // would form)
  __shared__ double sv[kMaxE], s2v[kMaxE];
  if (threadIdx.x < E) {
    sloc[threadIdx.x] = D.loc[s * E + threadIdx.x];
    sv[threadIdx.x] = D.vgrid[threadIdx.x];
    s2v[threadIdx.x] = 2.0 / D.vgrid[threadIdx.x];
  }
  if (threadIdx.x < 2 * MH + 2) {
    const int m = threadIdx.x;
    const double mu = (m < M) ? D.mugrid[m] : 0.0;
    smu[m] = mu;
    smu2[m] = mu * mu;
    shc[m] = 0.5 * (1.0 - mu * mu);
  }
  if (FM)
    for (int q = threadIdx.x; q < kFmTable; q += blockDim.x) fmt[q] = D.fm_tab[q];
  __syncthreads();
  const int n = D.n_active[s];
  if (node < 1 || node >= n) return;  // both lanes of a pair leave together
  const size_t i = (size_t)s * D.Npad + node;
  const int m0 = hf * MH;
  // focusing coefficients (k_node_prep)
  const int nsub_f = D.nsub_f[i];
  const bool foc = !(fused && !D.thin[s]) && nsub_f < kHeavyFocus;
  const double dtsub_f = D.dtsub_f[i];
  const double dlnBds = D.f_dlnbds[i], fDuPar = D.f_dupar[i], fb = D.f_b[i];
  // change coefficients
  const int nsub_c = D.nsub_c[i];
  const double dtsub_c = D.dtsub_c[i];
  const double cDuPar = D.c_dupar[i], DlnNDt = D.c_dlnn[i], DlnBDt = D.c_dlnb[i];
  const double damp = D.c_damp[i], bc0 = D.c_bc0[i], bc1 = D.c_bc1[i];
  const double nmb = DlnNDt - DlnBDt;
  const bool upd = nsub_c > 0;
  const double dlnp_i = 1.0 / D.dlnp;
  const double half = 0.5;
  const double *__restrict__ vgi = D.vgrid_i;
  constexpr uint32_t B = sizeof(T);
  const uint32_t slab_bytes = (uint32_t)((size_t)M * D.Npad * B);
  const uint32_t row = (uint32_t)D.Npad * B;
  const uint32_t voff = (uint32_t)(((size_t)m0 * D.Npad + node) * B);
  auto slab_of = [&](int e, int flip) {
    const int task = s * E + e;
    const int lc = __builtin_amdgcn_readfirstlane((int)sloc[e]) ^ flip;
    return slab_rsrc((const double *)fslab<T>(D, lc, task), slab_bytes);
  };
  // slot j of this lane is mu cell m0 + j; the upper half's last slot is
  // padding when M is odd
  auto live = [&](int j) { return (M % 2 == 0) || (j + 1 < MH) || (hf == 0); };
  // velocity of the change in ln p at (energy e, slot j), as in k_change:
  // damp * (a * (1/v) + bb), a = -mu DuPar/dt, bb = mu^2 (DlnN - DlnB)/dt
  // + 0.5 (1 - mu^2) DlnB/dt
  auto cvel = [&](int e, int j) {
    const int m = m0 + j;
    const double a = -smu[m] * cDuPar;
    const double bb = smu2[m] * nmb + shc[m] * DlnBDt;
    return damp * (a * vgi[e] + bb);
  };
  double lprev[MH], flux[MH];  // ln f at e-1 after focusing; flux below e-1
  int fbad_e = -1, fbad_m = 0, fbad_inf = 0;
  int cbad_e = -1, cbad_m = 0, cbad_inf = 0;
  // z + 0 * x[j] over the live slots: NaN iff one of them is NaN or Inf
  // (one operation per cell, no branch)
  auto nonfinite = [&](const T *x) {
    T z = 0;
#pragma unroll
    for (int j = 0; j < MH; ++j) z = live(j) ? fma(x[j], (T)0, z) : z;
    return !isfinite(z);
  };
  // exp back, NaN/Inf check, floor and store of one energy's row; the
  // floor is fmax, equal to the reference's `f < DBL_MIN ? DBL_MIN : f` for
  // every non-NaN value (a NaN has been reported by then)
  auto put_row = [&](int e, const double *nv, const rsrc_t &out) {
    T val[MH];
#pragma unroll
    for (int j = 0; j < MH; ++j) val[j] = (T)(FM ? fm_exp(nv[j], fmt) : exp(nv[j]));
    if (nonfinite(val) && cbad_e < 0) {
      for (int j = 0; j < MH; ++j)
        if (live(j) && !isfinite(val[j])) {
          cbad_e = e;
          cbad_m = m0 + j;
          cbad_inf = isinf(val[j]) ? 1 : 0;
          break;
        }
    }
#pragma unroll
    for (int j = 0; j < MH; ++j)
      if (live(j)) buf_st(floor_max(val[j], FloorT<T>::v), out, voff, j * row);
  };
  for (int e = 0; e < E; ++e) {
    const rsrc_t in = slab_of(e, 0);
    T x[MH];
#pragma unroll
    for (int j = 0; j < MH; ++j) x[j] = live(j) ? buf_ld(in, voff, j * row, T()) : T(1);
    if (foc) {
      // ---- AdiabaticFocusing of (node, e).  Half-mesh velocity va at the
      // interface below slot j (j = MH: above the last slot), split by the
      // side the flux is taken from (split_face: vh = min(va, 0), vl = va -
      // vh, one of them an exact zero).  An interior face then
      // carries fma(vh, f_above, vl * f_below): one product plus an exact
      // zero, which is the reference's v 0.5 ((1 - c) f_above + (1 + c)
      // f_below) (src/energeticParticles.c:1830-1836): equal to v * f_upwind
      // for finite f and NaN when either cell is NaN or Inf.  The wall rows
      // (mu cells 0 and M-1) take the upwind cell alone, as in the reference
      // (:1850-1885): their faces are interface 1 (in the lower half when
      // MH >= 2) and interface M-1 (in the upper half, above slot JWU, when
      // M-1 > MH); otherwise they are the interface between the halves.
      constexpr bool LW = MH >= 2;
      constexpr bool UW = M - 1 > MH;
      constexpr int JWU = UW ? M - 2 - MH : 0;
      const double v = sv[e];
      const double a = -v * dlnBds - s2v[e] * fDuPar;
      const T dmu_i = (T)(1.0 / D.dmu);
      const T dts = (T)dtsub_f;
      const T sc = dts * dmu_i;
      const T hsc = (T)half * sc;
      T vl[MH + 1], vh[MH + 1];
      T vw = 0;  // this lane's wall face velocity
      double vp = (m0 > 0) ? shc[m0 - 1] * (a + smu[m0 - 1] * fb) : 0.0;
#pragma unroll
      for (int j = 0; j <= MH; ++j) {
        const int m = m0 + j;
        const double vel = shc[m] * (a + smu[m] * fb);
        // fused modes: velocities in units of dmu / dt_sub; (vel + vp) (sc/2)
        // rounds to the same value as ((vel + vp)/2) sc
        const T va = FM ? (T)(vel + vp) * hsc : (T)(half * (vel + vp));
        vp = vel;
        split_face(va, vl[j], vh[j]);
        if (LW && j == 1) vw = hf ? vw : va;
        if (UW && j == JWU + 1) vw = hf ? va : vw;
      }
      // the face between the halves (interface MH), seen the same by both
      const T vmid = hf ? vl[0] + vh[0] : vl[MH] + vh[MH];
      const T vlm = hf ? vl[0] : vl[MH], vhm = hf ? vh[0] : vh[MH];
      for (int it = 0; it < nsub_f; ++it) {
        const T recv = pair_swap(hf ? x[0] : x[MH - 1]);
        const T flo = hf ? recv : x[MH - 1];  // cell MH-1
        const T fhi = hf ? x[0] : recv;       // cell MH
        const T pmid = fma(vhm, fhi, vlm * flo);
        T umid = pmid;
        if (MH == 1 || MH == M - 1) umid = vmid * (signbit(vmid) ? fhi : flo);
        // the wall face inside this lane's half (lower: interface 1 between
        // slots 0 and 1; upper: interface M-1 above slot JWU)
        T uw = 0;
        if (LW || UW) {
          const T xa = hf ? x[JWU] : x[0], xb = hf ? x[JWU + 1 < MH ? JWU + 1 : 0] : x[MH > 1 ? 1 : 0];
          uw = vw * (signbit(vw) ? xb : xa);
        }
        T flux_lo = hf ? ((MH == M - 1) ? umid : pmid) : (T)0;  // zero-flux wall at mu = -1
#pragma unroll
        for (int j = 0; j < MH; ++j) {
          T flux_hi, next_lo;
          if (j + 1 < MH) {
            const T pf = fma(vh[j + 1], x[j + 1], vl[j + 1] * x[j]);
            flux_hi = pf;
            if (LW && j == 0) flux_hi = hf ? pf : uw;                     // row 0's upper face
            if (j + 2 == MH && M % 2 == 1) flux_hi = hf ? (T)0 : flux_hi;  // above cell M-1 (padding)
            next_lo = pf;
            if (UW && j == JWU) next_lo = hf ? uw : pf;                    // row M-1's lower face
          } else {
            flux_hi = hf ? (T)0 : ((MH == 1) ? umid : pmid);  // zero-flux wall at mu = +1
            next_lo = (T)0;
          }
          if (FM) {
            x[j] = x[j] - (flux_hi - flux_lo);
          } else {
            const T dfl = (flux_hi - flux_lo) * dmu_i;
            x[j] = x[j] - dts * dfl;
          }
          flux_lo = next_lo;
        }
      }
      if (nonfinite(x) && fbad_e < 0) {
        for (int j = 0; j < MH; ++j)
          if (live(j) && !isfinite(x[j])) {
            fbad_e = e;
            fbad_m = m0 + j;
            fbad_inf = isinf(x[j]) ? 1 : 0;
            break;
          }
      }
#pragma unroll
      for (int j = 0; j < MH; ++j) x[j] = floor_max(x[j], FloorT<T>::v);
    }
    // ---- AdiabaticChange, one sweep: finish energy e-1 (and e = E-1).
    // The energy flux takes the upwind cell alone: the reference's form
    // would also pass on a non-finite value of the other cell, but every
    // value reaching this point has passed a NaN/Inf check whose error
    // sorts first (streaming, or focusing of the same node, op 0 < op 1),
    // so the first reported cell cannot differ.
    double lx[MH];
#pragma unroll
    for (int j = 0; j < MH; ++j) lx[j] = FM ? fm_log((double)x[j], fmt) : log((double)x[j]);
    if (e >= 1) {
      const int ep = e - 1;
      const rsrc_t out = slab_of(ep, spare);
      double nv[MH];
#pragma unroll
      for (int j = 0; j < MH; ++j) {
        const double cur = lprev[j], nxt = lx[j];
        const double va1 = half * (cvel(e, j) + cvel(ep, j));
        double f1;
        if (ep == 0) {
          if (va1 <= 0)
            f1 = va1 * (nxt - cur) * dlnp_i;
          else
            f1 = va1 * (cur - bc0) * dlnp_i;
          flux[j] = va1 * (signbit(va1) ? nxt : cur);
        } else {
          const double flux_hi = va1 * (signbit(va1) ? nxt : cur);
          f1 = (flux_hi - flux[j]) * dlnp_i;
          flux[j] = flux_hi;
        }
        nv[j] = upd ? cur - dtsub_c * f1 : cur;
      }
      put_row(ep, nv, out);
      if (e == E - 1) {
        const rsrc_t out1 = slab_of(e, spare);
#pragma unroll
        for (int j = 0; j < MH; ++j) {
          const double cur = lx[j], old_prev = lprev[j];
          const double vr = half * (cvel(e, j) + cvel(ep, j));
          double f1;
          if (vr >= 0)
            f1 = vr * (cur - old_prev) * dlnp_i;
          else
            f1 = vr * (bc1 - cur) * dlnp_i;
          nv[j] = upd ? cur - dtsub_c * f1 : cur;
        }
        put_row(e, nv, out1);
      }
    }
#pragma unroll
    for (int j = 0; j < MH; ++j) lprev[j] = lx[j];
  }
  if (!spare) {
    if (fbad_e >= 0) note_error(D, err_key(substep, 1, s, 0, node, fbad_e, fbad_m, fbad_inf));
    if (cbad_e >= 0) note_error(D, err_key(substep, 1, s, 1, node, cbad_e, cbad_m, cbad_inf));
  }
}

// ------------------------------------------------------------------------
// k_focus_change4: the same pass with G = 4 lanes per node (for M >= 12).
// Lanes l, l+W, l+2W, .. (W = 64/G) of a wave hold mu cells q MG .. q MG +
// MG - 1 (q = l / W, MG = ceil(M/G), the last group padded) of node
// (l mod W); a 256-lane block covers 4W consecutive nodes.  Each focusing
// subcycle moves the boundary cells of the neighbouring groups in by two
// cross-lane shifts of W.  A lane keeps about half the state of the
// pair layout, so the kernel runs at a higher occupancy; arithmetic and
// results are those of k_focus_change.
// ------------------------------------------------------------------------
// G = 8 (eight lanes per node, 32 nodes per block) serves workloads with
// too few nodes to give every SIMD a wave at four lanes per node: a wave's
// chain over the energies is then about half as long.  Groups after the
// one holding cell M-1 are padding.
template <int M, typename T, bool FM, int G = 4>
__global__ __launch_bounds__(256) void k_focus_change4(Dev D, int substep, int fused, int spare) {
  if (substep_off(D)) return;
  constexpr int W = 64 / G, NB = 4 * W;
  constexpr int MG = (M + G - 1) / G;
  static_assert(G == 4 || G == 8, "four or eight lanes per node");
  static_assert(MG >= 2, "at least two mu cells per lane");
  constexpr int QL = (M - 1) / MG;       // group of cell M-1
  constexpr int JL = (M - 1) - QL * MG;  // its slot
  const int nch = (D.N + NB - 1) / NB;
  const int s = blockIdx.x / nch;
  const int chunk = (int)((blockIdx.x % nch + (unsigned)s) % nch);
  const int lane = threadIdx.x & 63;
  const int q = lane / W;
  const int node = chunk * NB + (threadIdx.x >> 6) * W + (lane % W);
  const int E = D.E;
  __shared__ uint8_t sloc[kMaxE];
  // per mu (padded with zeros): mu, mu^2, 0.5 (1 - mu^2)
  __shared__ double smu[G * MG + 2], smu2[G * MG + 2], shc[G * MG + 2];
  __shared__ double fmt[FM ? kFmTable : 1];
  // per energy: v and 2/v (the same correctly rounded quotient each lane
  // would form)
  __shared__ double sv[kMaxE], s2v[kMaxE];
  if (threadIdx.x < E) {
    sloc[threadIdx.x] = D.loc[s * E + threadIdx.x];
    sv[threadIdx.x] = D.vgrid[threadIdx.x];
    s2v[threadIdx.x] = 2.0 / D.vgrid[threadIdx.x];
  }
  if (threadIdx.x < G * MG + 2) {
    const int m = threadIdx.x;
    const double mu = (m < M) ? D.mugrid[m] : 0.0;
    smu[m] = mu;
    smu2[m] = mu * mu;
    shc[m] = 0.5 * (1.0 - mu * mu);
  }
  if (FM)
    for (int t = threadIdx.x; t < kFmTable; t += blockDim.x) fmt[t] = D.fm_tab[t];
  __syncthreads();
  const int n = D.n_active[s];
  if (node < 1 || node >= n) return;  // the four lanes of a node leave together
  const size_t i = (size_t)s * D.Npad + node;
  const int m0 = q * MG;
  const int nsub_f = D.nsub_f[i];
  const bool foc = !(fused && !D.thin[s]) && nsub_f < kHeavyFocus;
  const double dtsub_f = D.dtsub_f[i];
  const double dlnBds = D.f_dlnbds[i], fDuPar = D.f_dupar[i], fb = D.f_b[i];
  const int nsub_c = D.nsub_c[i];
  const double dtsub_c = D.dtsub_c[i];
  const double cDuPar = D.c_dupar[i], DlnNDt = D.c_dlnn[i], DlnBDt = D.c_dlnb[i];
  const double damp = D.c_damp[i], bc0 = D.c_bc0[i], bc1 = D.c_bc1[i];
  const double nmb = DlnNDt - DlnBDt;
  const bool upd = nsub_c > 0;
  const double dlnp_i = 1.0 / D.dlnp;
  const double half = 0.5;
  const double *__restrict__ vgi = D.vgrid_i;
  constexpr uint32_t B = sizeof(T);
  const uint32_t slab_bytes = (uint32_t)((size_t)M * D.Npad * B);
  const uint32_t row = (uint32_t)D.Npad * B;
  const uint32_t voff = (uint32_t)(((size_t)m0 * D.Npad + node) * B);
  auto slab_of = [&](int e, int flip) {
    const int task = s * E + e;
    const int lc = __builtin_amdgcn_readfirstlane((int)sloc[e]) ^ flip;
    return slab_rsrc((const double *)fslab<T>(D, lc, task), slab_bytes);
  };
  // slot j is mu cell m0 + j; slots past cell M-1 are padding
  auto live = [&](int j) { return (q < QL) || (q == QL && j <= JL); };
  auto cvel = [&](int e, int j) {
    const int m = m0 + j;
    const double a = -smu[m] * cDuPar;
    const double bb = smu2[m] * nmb + shc[m] * DlnBDt;
    return damp * (a * vgi[e] + bb);
  };
  double lprev[MG], flux[MG];  // ln f at e-1 after focusing; flux below e-1
  double cvp[MG];              // change velocity at e-1 (this energy's, next time)
  int fbad_e = -1, fbad_m = 0, fbad_inf = 0;
  int cbad_e = -1, cbad_m = 0, cbad_inf = 0;
  auto nonfinite = [&](const T *x) {
    T z = 0;
#pragma unroll
    for (int j = 0; j < MG; ++j) z = live(j) ? fma(x[j], (T)0, z) : z;
    return !isfinite(z);
  };
  auto put_row = [&](int e, const double *nv, const rsrc_t &out) {
    T val[MG];
#pragma unroll
    for (int j = 0; j < MG; ++j) val[j] = (T)(FM ? fm_exp(nv[j], fmt) : exp(nv[j]));
    if (nonfinite(val) && cbad_e < 0) {
      for (int j = 0; j < MG; ++j)
        if (live(j) && !isfinite(val[j])) {
          cbad_e = e;
          cbad_m = m0 + j;
          cbad_inf = isinf(val[j]) ? 1 : 0;
          break;
        }
    }
#pragma unroll
    for (int j = 0; j < MG; ++j)
      if (live(j)) buf_st(floor_max(val[j], FloorT<T>::v), out, voff, j * row);
  };
  for (int e = 0; e < E; ++e) {
    const rsrc_t in = slab_of(e, 0);
    T x[MG];
#pragma unroll
    for (int j = 0; j < MG; ++j) x[j] = live(j) ? buf_ld(in, voff, j * row, T()) : T(1);
    if (foc) {
      // ---- AdiabaticFocusing of (node, e), as in k_focus_change: face
      // velocities split by upwind side (vl, vh), interior faces
      // fma(vh, f_above, vl * f_below), the wall rows' faces (interface 1
      // above cell 0, interface M-1 below cell M-1) the upwind cell alone.
      const double v = sv[e];
      const double a = -v * dlnBds - s2v[e] * fDuPar;
      const T dmu_i = (T)(1.0 / D.dmu);
      const T dts = (T)dtsub_f;
      const T sc = dts * dmu_i;
      const T hsc = (T)half * sc;
      T vl[MG + 1], vh[MG + 1];  // face below slot j (j = MG: above the last slot)
      double vp = (m0 > 0) ? shc[m0 - 1] * (a + smu[m0 - 1] * fb) : 0.0;
#pragma unroll
      for (int j = 0; j <= MG; ++j) {
        const int m = m0 + j;
        const double vel = shc[m] * (a + smu[m] * fb);
        // fused modes: velocities in units of dmu / dt_sub; (vel + vp) (sc/2)
        // rounds to the same value as ((vel + vp)/2) sc
        const T va = FM ? (T)(vel + vp) * hsc : (T)(half * (vel + vp));
        vp = vel;
        split_face(va, vl[j], vh[j]);
      }
      for (int it = 0; it < nsub_f; ++it) {
        // cell m0-1 (last of the group below), cell m0+MG (first above)
        const T below = __shfl_up(x[MG - 1], W, 64);
        const T above = __shfl_down(x[0], W, 64);
        // face below slot 0: zero-flux wall under cell 0, else interior
        // (or, when cell M-1 is slot 0, that row's upwind-only face)
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
          if (FM) {
            x[j] = x[j] - (flux_hi - flux_lo);
          } else {
            const T dfl = (flux_hi - flux_lo) * dmu_i;
            x[j] = x[j] - dts * dfl;
          }
          flux_lo = next_lo;
        }
      }
      if (nonfinite(x) && fbad_e < 0) {
        for (int j = 0; j < MG; ++j)
          if (live(j) && !isfinite(x[j])) {
            fbad_e = e;
            fbad_m = m0 + j;
            fbad_inf = isinf(x[j]) ? 1 : 0;
            break;
          }
      }
#pragma unroll
      for (int j = 0; j < MG; ++j) x[j] = floor_max(x[j], FloorT<T>::v);
    }
    // ---- AdiabaticChange, one sweep: finish energy e-1 (and e = E-1),
    // as in k_focus_change
    double lx[MG];
#pragma unroll
    for (int j = 0; j < MG; ++j) lx[j] = FM ? fm_log((double)x[j], fmt) : log((double)x[j]);
    if (e >= 1) {
      const int ep = e - 1;
      const rsrc_t out = slab_of(ep, spare);
      double nv[MG];
#pragma unroll
      for (int j = 0; j < MG; ++j) {
        const double cur = lprev[j], nxt = lx[j];
        const double cve = cvel(e, j);
        const double va1 = half * (cve + cvp[j]);
        cvp[j] = cve;
        double f1;
        if (ep == 0) {
          if (va1 <= 0)
            f1 = va1 * (nxt - cur) * dlnp_i;
          else
            f1 = va1 * (cur - bc0) * dlnp_i;
          flux[j] = va1 * (signbit(va1) ? nxt : cur);
        } else {
          const double flux_hi = va1 * (signbit(va1) ? nxt : cur);
          f1 = (flux_hi - flux[j]) * dlnp_i;
          flux[j] = flux_hi;
        }
        nv[j] = upd ? cur - dtsub_c * f1 : cur;
      }
      put_row(ep, nv, out);
      if (e == E - 1) {
        const rsrc_t out1 = slab_of(e, spare);
#pragma unroll
        for (int j = 0; j < MG; ++j) {
          const double cur = lx[j], old_prev = lprev[j];
          const double vr = half * (cvp[j] + cvel(ep, j));
          double f1;
          if (vr >= 0)
            f1 = vr * (cur - old_prev) * dlnp_i;
          else
            f1 = vr * (bc1 - cur) * dlnp_i;
          nv[j] = upd ? cur - dtsub_c * f1 : cur;
        }
        put_row(e, nv, out1);
      }
    }
#pragma unroll
    for (int j = 0; j < MG; ++j) {
      lprev[j] = lx[j];
      if (e == 0) cvp[j] = cvel(0, j);
    }
  }
  if (!spare) {
    if (fbad_e >= 0) note_error(D, err_key(substep, 1, s, 0, node, fbad_e, fbad_m, fbad_inf));
    if (cbad_e >= 0) note_error(D, err_key(substep, 1, s, 1, node, cbad_e, cbad_m, cbad_inf));
  }
}

// Waves of the four-lanes-per-node kernel below which eight lanes per node
// are used (fewer than two per SIMD of the 1024).
constexpr int64_t kFc8Waves = 2048;

template <int M, int G>
static void focus_change_g(const Dev &D, int sub, int fused, int spare, hipStream_t st) {
  const int nb = 4 * (64 / G);  // nodes per 256-lane block
  const dim3 grid((D.N + nb - 1) / nb * D.S), block(256);
  if (D.fp32)
    hipLaunchKernelGGL((k_focus_change4<M, float, true, G>), grid, block, 0, st, D, sub, fused, spare);
  else if (D.fast)
    hipLaunchKernelGGL((k_focus_change4<M, double, true, G>), grid, block, 0, st, D, sub, fused, spare);
  else
    hipLaunchKernelGGL((k_focus_change4<M, double, false, G>), grid, block, 0, st, D, sub, fused, spare);
}

template <int M>
static void focus_change_t(const Dev &D, int sub, int fused, int spare, hipStream_t st) {
  if constexpr (M >= 12) {
    // four lanes per node, or eight when that leaves SIMDs without a wave
    const int64_t waves4 = (int64_t)D.S * ((D.N + 15) / 16);
    if (waves4 < kFc8Waves)
      focus_change_g<M, 8>(D, sub, fused, spare, st);
    else
      focus_change_g<M, 4>(D, sub, fused, spare, st);
  } else {
    const dim3 grid((D.N + 127) / 128 * D.S), block(256);
    if (D.fp32)
      hipLaunchKernelGGL((k_focus_change<M, float, true>), grid, block, 0, st, D, sub, fused, spare);
    else if (D.fast)
      hipLaunchKernelGGL((k_focus_change<M, double, true>), grid, block, 0, st, D, sub, fused, spare);
    else
      hipLaunchKernelGGL((k_focus_change<M, double, false>), grid, block, 0, st, D, sub, fused, spare);
  }
}

#define EP_FC_CASE(MM)                         \
  case MM:                                     \
    focus_change_t<MM>(D, sub, fused, spare, st); \
    break;

hipError_t launch_focus_change(const Dev &D, int sub, int fused, int spare, hipStream_t st) {
  if (D.logf) return hipErrorInvalidValue;  // the log-scaled storage runs the two-kernel node stage
  switch (D.M) {
    EP_FC_CASE(2) EP_FC_CASE(3) EP_FC_CASE(4) EP_FC_CASE(5) EP_FC_CASE(6)
    EP_FC_CASE(7) EP_FC_CASE(8) EP_FC_CASE(9) EP_FC_CASE(10) EP_FC_CASE(11)
    EP_FC_CASE(12) EP_FC_CASE(13) EP_FC_CASE(14) EP_FC_CASE(15) EP_FC_CASE(16)
    EP_FC_CASE(17) EP_FC_CASE(18) EP_FC_CASE(19) EP_FC_CASE(20) EP_FC_CASE(21)
    EP_FC_CASE(22) EP_FC_CASE(23) EP_FC_CASE(24) EP_FC_CASE(25) EP_FC_CASE(26)
    EP_FC_CASE(27) EP_FC_CASE(28) EP_FC_CASE(29) EP_FC_CASE(30) EP_FC_CASE(31)
    EP_FC_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ep
