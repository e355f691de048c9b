// ep_rk.hip -- AdiabaticFocusing and AdiabaticChange with the three-stage
// Runge-Kutta integrator and the upwind (alg 2) or WENO3 (alg 3) operator
// (src/energeticParticles.c:952-1018, :1179-1427, :1677-1743, :1899-2144).
//
// The same subcycle counts as the upwind path (k_node_prep), FP64 storage
// only, operations in the reference's order.
//   k_focus_rk  a lane per node, a block per 64 nodes of one (stream,
//               energy): the mu row, the RK stage and the half-mesh
//               velocities in registers.
//   k_change_rk a lane per (node, mu) column: ln f is advanced in place in
//               the current buffer, the RK stage lives in the spare buffer
//               (free during the node stage), and every stage is one sweep
//               up the energy column with a window of old values, so a
//               stage may overwrite its own input.
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>

#include "ep_internal.h"
#include "ep_launch.h"

namespace ep {

// The WENO3 flux of one face from the split fluxes around it (lnm2, lnm1,
// ln0: LN at the cells two left, one left and right of the face; lpm1, lp0,
// lpp1: LP at the cells left, right and two right of it), as
// AdiabaticFocusing_Operator_WENO3 (:2070-2103) and
// AdiabaticChange_Operator_WENO3 (:1321-1354) form it.
__device__ __forceinline__ double weno3_face(double lnm2, double lnm1, double ln0, double lpm1, double lp0,
                                             double lpp1) {
  const double D_C_CPt = 1 / 2.0, D_C_MCt = 1 / 2.0, D_M_Tt = 1 / 3.0, D_CP_Tt = 2 / 3.0, D_P_Tt = 1 / 3.0,
               D_MC_Tt = 2 / 3.0;
  const double weno_eps = 10.0 * 1.4916681462400413e-154;  // 10 sqrt(DBL_MIN)
  const double p0m = (1.0 + D_C_MCt) * lnm1 - D_C_MCt * lnm2;
  const double p1m = D_C_MCt * lnm1 + D_C_CPt * ln0;
  const double p0p = (1.0 + D_C_CPt) * lp0 - D_C_CPt * lpp1;
  const double p1p = D_C_CPt * lp0 + D_C_MCt * lpm1;
  const double d0m = D_C_MCt * (lnm1 - lnm2), d1m = D_C_CPt * (ln0 - lnm1);
  const double d0p = D_C_CPt * (lpp1 - lp0), d1p = D_C_MCt * (lp0 - lpm1);
  const double B0m = 4.0 * (d0m * d0m), B1m = 4.0 * (d1m * d1m);
  const double B0p = 4.0 * (d0p * d0p), B1p = 4.0 * (d1p * d1p);
  const double q0m = weno_eps + B0m, q1m = weno_eps + B1m, q0p = weno_eps + B0p, q1p = weno_eps + B1p;
  const double w0m = D_P_Tt / (q0m * q0m), w1m = D_MC_Tt / (q1m * q1m);
  const double w0p = D_M_Tt / (q0p * q0p), w1p = D_CP_Tt / (q1p * q1p);
  const double wm_sum = w0m + w1m, wp_sum = w0p + w1p;
  const double um = (w0m / wm_sum) * p0m + (w1m / wm_sum) * p1m;
  const double up = (w0p / wp_sum) * p0p + (w1p / wp_sum) * p1p;
  return up + um;
}

// The reference's upwind face flux v 0.5 ((1 - c) f[right] + (1 + c)
// f[left]), c = copysign(1, v).
__device__ __forceinline__ double upwind_face(double va, double fl, double fr) {
  const double cce = copysign(1.0, va);
  return va * 0.5 * ((1.0 - cce) * fr + (1.0 + cce) * fl);
}

// alpha of cell i of a line of n cells: the largest |va| over the faces
// i-2..i+3 that exist, in the reference's order (:1961-2017, :1218-1274).
template <typename VA>
__device__ __forceinline__ double weno3_alpha(VA va, int i, int n) {
  double v = fabs(va(i));
  auto take = [&](int j) {
    const double x = fabs(va(j));
    if (x > v) v = x;
  };
  if (i >= 2 && i < n - 2) {
    take(i - 1); take(i - 2); take(i + 1); take(i + 2); take(i + 3);
  } else if (i == 0) {
    take(i + 1); take(i + 2); take(i + 3);
  } else if (i == 1) {
    take(i - 1); take(i + 1); take(i + 2); take(i + 3);
  } else if (i == n - 2) {
    take(i + 1); take(i + 2); take(i - 1); take(i - 2);
  } else {
    take(i + 1); take(i - 1); take(i - 2);
  }
  return v;
}

// ------------------------------------------------------------------------
// Focusing: one stage of the RK3 subcycle on a mu row.  f1 = d(v f)/dmu of
// `in` (AdiabaticFocusing_Operator_Upwind :1806-1896 for ALG 2, _WENO3
// :1899-2144 for ALG 3) is formed cell by cell with a window of the old
// values of `in`, so `out` may be `in`; base is f at the start of the
// subcycle.  alpha is taken from the velocities where a cell needs it.
// ------------------------------------------------------------------------
template <int M, int ALG>
__device__ __forceinline__ void focus_stage(int stage, const double (&base)[M], const double (&in)[M],
                                            double (&out)[M], const double (&va)[M + 1], double dmu, double dt) {
  const double dmu_i = 1.0 / dmu;
  auto vaf = [&](int j) { return va[j]; };
  auto al = [&](int c) { return weno3_alpha(vaf, c, M); };
  auto lp = [&](int c, double x) { return 0.5 * x * va[c + 1] - al(c); };
  auto ln = [&](int c, double x) { return 0.5 * x * va[c] + al(c); };
  // w0..w3 = old in[m-1], in[m], in[m+1], in[m+2]
  double w0 = 0.0, w1 = in[0], w2 = in[1], w3 = (M > 2) ? in[2] : 0.0;
  double flo = 0.0;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    double f1;
    if (m == 0) {
      const double v1 = va[1];
      if (ALG == 3)
        f1 = (v1 <= 0) ? v1 * w2 / dmu : v1 * w1 / dmu;
      else
        f1 = (v1 <= 0) ? v1 * w2 * dmu_i : v1 * w1 * dmu_i;
      flo = upwind_face(v1, w1, w2);  // face 1
    } else if (m == M - 1) {
      const double vr = va[M - 1];
      if (ALG == 3)
        f1 = (vr >= 0) ? vr * (-w0) / dmu : vr * (-w1) / dmu;
      else
        f1 = (vr >= 0) ? vr * (-w0) * dmu_i : vr * (-w1) * dmu_i;
    } else {
      const int j = m + 1;  // the face above cell m
      double fhi;
      if (ALG == 3 && j <= M - 2) {
        fhi = weno3_face(ln(j - 2, w0), ln(j - 1, w1), ln(j, w2), lp(j - 1, w1), lp(j, w2), lp(j + 1, w3));
        f1 = (fhi - flo) / dmu;
      } else {
        fhi = upwind_face(va[j], w1, w2);
        f1 = (ALG == 3) ? (fhi - flo) / dmu : (fhi - flo) * dmu_i;
      }
      flo = fhi;
    }
    double r;
    if (stage == 0)
      r = w1 - dt * f1;
    else if (stage == 1)
      r = (3.0 / 4.0) * base[m] + (1.0 / 4.0) * w1 - (1.0 / 4.0) * dt * f1;
    else
      r = (1.0 / 3.0) * base[m] + (2.0 / 3.0) * w1 - (2.0 / 3.0) * dt * f1;
    const double nx = (m + 3 < M) ? in[m + 3 < M ? m + 3 : 0] : 0.0;
    out[m] = r;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = nx;
  }
}

template <int M, int ALG>
__global__ __launch_bounds__(64) void k_focus_rk(Dev D, int substep) {
  static_assert(ALG == 2 || M >= 4, "WENO3 needs four cells per line");
  if (substep_off(D)) return;
  const int nch = (D.N + 63) >> 6;
  const int se = blockIdx.x / nch;
  const int s = se / D.E, e = se % D.E;
  const int node = (blockIdx.x % nch) * 64 + threadIdx.x;
  const int n = D.n_active[s];
  if (node < 1 || node >= n) return;
  const size_t i = (size_t)s * D.Npad + node;
  const int nsub = D.nsub_f[i];
  const double dtsub = D.dtsub_f[i];
  const double dlnBds = D.f_dlnbds[i], DuParDt = D.f_dupar[i], b = D.f_b[i];
  const double half = 0.5, one = 1.0, two = 2.0;
  const int task = s * D.E + e;
  double *fp = fslab(D, D.loc[task], task) + node;
  const double v = D.vgrid[e];
  const double a = -v * dlnBds - (two / v) * DuParDt;
  double f[M], s1[M], va[M + 1];
  double vprev = 0.0;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    f[m] = fp[(size_t)m * D.Npad];
    const double mu = D.mugrid[m];
    const double vel = half * (one - mu * mu) * (a + mu * b);
    if (m > 0) va[m] = half * (vel + vprev);
    vprev = vel;
  }
  va[0] = va[1];
  va[M] = va[M - 1];
  for (int it = 0; it < nsub; ++it) {
    // the velocities pass through an empty asm each subcycle, so that the
    // compiler does not hoist per-face terms out of the loop into (spilled)
    // registers of their own
#pragma unroll
    for (int m = 0; m <= M; ++m) asm volatile("" : "+v"(va[m]));
    focus_stage<M, ALG>(0, f, f, s1, va, D.dmu, dtsub);
    focus_stage<M, ALG>(1, f, s1, s1, va, D.dmu, dtsub);
    focus_stage<M, ALG>(2, f, s1, f, va, D.dmu, dtsub);
  }
  int bad = -1, bad_inf = 0;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    if (bad < 0 && !isfinite(f[m])) {
      bad = m;
      bad_inf = isinf(f[m]) ? 1 : 0;
    }
    fp[(size_t)m * D.Npad] = (f[m] < kDblMin) ? kDblMin : f[m];
  }
  if (bad >= 0) note_error(D, err_key(substep, 1, s, 0, node, e, bad, bad_inf));
}

// ------------------------------------------------------------------------
// Adiabatic change: one stage of the RK3 subcycle on a column, in a sweep
// up the energies.  in: the stage's operand (ln f, or the previous stage in
// the spare buffer), out: where the stage's result goes (may be `in`: the
// window keeps the old values still needed), base: ln f at the start of
// the subcycle (the current buffer, which the last stage overwrites).
// ------------------------------------------------------------------------
template <int ALG, typename CI, typename CO, typename CB, typename VA>
__device__ __forceinline__ void change_stage(int stage, int E, double dlnp, double dt, double bc0, double bc1,
                                             CI in, CO out, CB base, VA va) {
  const double dlnp_i = 1.0 / dlnp;
  auto al = [&](int c) { return weno3_alpha(va, c, E); };
  auto lp = [&](int c, double x) { return 0.5 * x * va(c + 1) - al(c); };
  auto ln = [&](int c, double x) { return 0.5 * x * va(c) + al(c); };
  // window: w0 = in[e-1], w1 = in[e], w2 = in[e+1], w3 = in[e+2] (old values)
  double w0 = 0.0, w1 = *in(0), w2 = *in(1), w3 = (E > 2) ? *in(2) : 0.0;
  double flo = 0.0;
  for (int e = 0; e < E; ++e) {
    double f1;
    if (e == 0) {
      const double v1 = va(1);
      if (ALG == 3)
        f1 = (v1 <= 0) ? v1 * (w2 - w1) / dlnp : v1 * (w1 - bc0) / dlnp;
      else
        f1 = (v1 <= 0) ? v1 * (w2 - w1) * dlnp_i : v1 * (w1 - bc0) * dlnp_i;
      flo = upwind_face(v1, w1, w2);  // face 1
    } else if (e == E - 1) {
      const double vr = va(E - 1);
      if (ALG == 3)
        f1 = (vr >= 0) ? vr * (w1 - w0) / dlnp : vr * (bc1 - w1) / dlnp;
      else
        f1 = (vr >= 0) ? vr * (w1 - w0) * dlnp_i : vr * (bc1 - w1) * dlnp_i;
    } else {
      const int j = e + 1;  // the face above cell e
      double fhi;
      if (ALG == 3 && j <= E - 2) {
        // needs in[j-2..j+1] = in[e-1..e+2]
        fhi = weno3_face(ln(j - 2, w0), ln(j - 1, w1), ln(j, w2), lp(j - 1, w1), lp(j, w2), lp(j + 1, w3));
        f1 = (fhi - flo) / dlnp;
      } else {
        fhi = upwind_face(va(j), w1, w2);
        f1 = (ALG == 3) ? (fhi - flo) / dlnp : (fhi - flo) * dlnp_i;
      }
      flo = fhi;
    }
    double r;
    if (stage == 0)
      r = w1 - dt * f1;
    else if (stage == 1)
      r = (3.0 / 4.0) * *base(e) + (1.0 / 4.0) * w1 - (1.0 / 4.0) * dt * f1;
    else
      r = (1.0 / 3.0) * *base(e) + (2.0 / 3.0) * w1 - (2.0 / 3.0) * dt * f1;
    // read in[e+3] before out[e] can overwrite anything the window needs
    const double nx = (e + 3 < E) ? *in(e + 3) : 0.0;
    *out(e) = r;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = nx;
  }
}

template <int ALG>
__global__ __launch_bounds__(256) void k_change_rk(Dev D, int substep) {
  if (substep_off(D)) return;
  const int nch = (D.N + 255) >> 8;
  const int sm = blockIdx.x / nch;
  const int s = sm / D.M, m = sm % D.M;
  const int node = (blockIdx.x % nch) * 256 + threadIdx.x;
  const int n = D.n_active[s];
  const int E = D.E;
  if (node < 1 || node >= n) return;
  const size_t i = (size_t)s * D.Npad + node;
  const int nsub = D.nsub_c[i];
  const double dtsub = D.dtsub_c[i];
  const double DuParDt = D.c_dupar[i], DlnNDt = D.c_dlnn[i], DlnBDt = D.c_dlnb[i];
  const double damp = D.c_damp[i], bc0 = D.c_bc0[i], bc1 = D.c_bc1[i];
  const double half = 0.5, one = 1.0;
  const double mu = D.mugrid[m];
  const double a = -mu * DuParDt;
  const double bb = mu * mu * (DlnNDt - DlnBDt) + half * (one - mu * mu) * DlnBDt;
  const double *vgi = D.vgrid_i;
  const size_t cell = (size_t)m * D.Npad + node;
  auto cur = [&](int e) {
    const int t = s * E + e;
    return fslab(D, D.loc[t], t) + cell;
  };
  auto spare = [&](int e) {
    const int t = s * E + e;
    return fslab(D, D.loc[t] ^ 1, t) + cell;
  };
  auto vel = [&](int e) { return damp * (a * vgi[e] + bb); };
  auto va = [&](int j) {
    const int jj = (j <= 0) ? 1 : (j >= E ? E - 1 : j);  // copied ends
    return half * (vel(jj) + vel(jj - 1));
  };
  for (int e = 0; e < E; ++e) *cur(e) = log(*cur(e));
  for (int it = 0; it < nsub; ++it) {
    change_stage<ALG>(0, E, D.dlnp, dtsub, bc0, bc1, cur, spare, cur, va);
    change_stage<ALG>(1, E, D.dlnp, dtsub, bc0, bc1, spare, spare, cur, va);
    change_stage<ALG>(2, E, D.dlnp, dtsub, bc0, bc1, spare, cur, cur, va);
  }
  int bad_e = -1, bad_inf = 0;
  for (int e = 0; e < E; ++e) {
    double *p = cur(e);
    const double x = exp(*p);
    if (bad_e < 0 && !isfinite(x)) {
      bad_e = e;
      bad_inf = isinf(x) ? 1 : 0;
    }
    *p = (x < kDblMin) ? kDblMin : x;
  }
  if (bad_e >= 0) note_error(D, err_key(substep, 1, s, 1, node, bad_e, m, bad_inf));
}

template <int M>
static hipError_t focus_rk_m(const Dev &D, int alg, int sub, hipStream_t st) {
  const dim3 grid((D.N + 63) / 64 * D.S * D.E);
  if constexpr (M >= 4) {
    if (alg == 3) {
      hipLaunchKernelGGL((k_focus_rk<M, 3>), grid, dim3(64), 0, st, D, sub);
      return hipGetLastError();
    }
  }
  if (alg == 3) return hipErrorInvalidValue;  // refused on the host before
  hipLaunchKernelGGL((k_focus_rk<M, 2>), grid, dim3(64), 0, st, D, sub);
  return hipGetLastError();
}

#define EP_RK_CASE(MM) \
  case MM:             \
    return focus_rk_m<MM>(D, alg, sub, st);

hipError_t launch_focus_rk(const Dev &D, int alg, int sub, hipStream_t st) {
  switch (D.M) {
    EP_RK_CASE(2) EP_RK_CASE(3) EP_RK_CASE(4) EP_RK_CASE(5) EP_RK_CASE(6) EP_RK_CASE(7) EP_RK_CASE(8)
    EP_RK_CASE(9) EP_RK_CASE(10) EP_RK_CASE(11) EP_RK_CASE(12) EP_RK_CASE(13) EP_RK_CASE(14)
    EP_RK_CASE(15) EP_RK_CASE(16) EP_RK_CASE(17) EP_RK_CASE(18) EP_RK_CASE(19) EP_RK_CASE(20)
    EP_RK_CASE(21) EP_RK_CASE(22) EP_RK_CASE(23) EP_RK_CASE(24) EP_RK_CASE(25) EP_RK_CASE(26)
    EP_RK_CASE(27) EP_RK_CASE(28) EP_RK_CASE(29) EP_RK_CASE(30) EP_RK_CASE(31) EP_RK_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_change_rk(const Dev &D, int alg, int sub, hipStream_t st) {
  const dim3 grid((D.N + 255) / 256 * D.S * D.M);
  if (alg == 3)
    hipLaunchKernelGGL(k_change_rk<3>, grid, dim3(256), 0, st, D, sub);
  else
    hipLaunchKernelGGL(k_change_rk<2>, grid, dim3(256), 0, st, D, sub);
  return hipGetLastError();
}

}  // namespace ep
