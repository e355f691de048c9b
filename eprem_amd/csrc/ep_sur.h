// ep_sur.h -- inputs of the learned large-step surrogate, shared by the
// sample generator (ep_surrogate.hip) and the surrogate EP update
// (ep_sur.hip), so that both compute them with the same operations.
#pragma once
#include "ep_internal.h"

namespace ep {

// The six scalar inputs of the sample at kept node k (1 <= k <= nk-2) of
// (stream s, energy e) of a full-resolution stream, for K subcycles, i.e.
// T = K dtProp of physical time:
//   0  T v / ds_k              cells travelled at mu = 1
//   1  T / tau_k               relaxation e-folds (tau = lambda / v)
//   2  ln(ds_k / ds_{k-1})     grid stretching behind the node
//   3  ln(ds_{k+1} / ds_k)     and ahead of it
//   4  asinh(a T / dmu)        focusing, velocity 0.5 (1 - mu^2) (a + mu b)
//   5  asinh(b T / dmu)        (src/energeticParticles.c:1613-1623)
// with a = -v dlnB/ds - (2/v) DuPar/dt and b from k_node_prep.
__device__ inline void sur_scalars(const Dev &D, int s, int e, int k, int K, double *x) {
  const size_t on = (size_t)s * D.Npad;
  const int task = s * D.E + e;
  const double v = D.vgrid[e], T = (double)K * D.dtprop[task];
  const double tau = D.vgrid_i[e] * (D.rig_lamo[e] * D.mfp_r[on + k]);
  const double dsi = D.dsil[on + k], dsim = D.dsil[on + k - 1], dsip = D.dsil[on + k + 1];
  const size_t i = on + k;
  const double a = -v * D.f_dlnbds[i] - (2.0 / v) * D.f_dupar[i];
  const double b = D.f_b[i];
  x[0] = T * v * dsi;
  x[1] = T / tau;
  x[2] = log(dsim / dsi);
  x[3] = log(dsi / dsip);
  x[4] = asinh(a * T / D.dmu);
  x[5] = asinh(b * T / D.dmu);
}

// Upwind focusing of the mu row f (register array, M <= NM cells) of node
// i, energy e over time T = K dtProp, with the node's own focusing
// subcycle (k_node_prep: nsub_f per substep dt) rescaled to divide T:
// half-mesh velocities, v f_upwind on the interior faces, zero-flux walls,
// DBL_MIN floor (src/energeticParticles.c:1613-1653, :1806-1896).
template <int NM>
__device__ inline void sur_focus_row(const Dev &D, size_t i, int e, double T, int K, int nsteps, double (&f)[NM]) {
  const int nf = D.nsub_f[i];
  if (nf <= 0) return;
  const int M = D.M;
  const int nq = (K * nf + nsteps - 1) / nsteps;
  const double dtq = T / (double)(nq > 0 ? nq : 1);
  const double v = D.vgrid[e];
  const double a = -v * D.f_dlnbds[i] - (2.0 / v) * D.f_dupar[i];
  const double b = D.f_b[i];
  double va[NM];  // va[m]: interface between cells m-1 and m (m >= 1)
  double vp = 0.0;
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const double mu = (m < M) ? D.mugrid[m] : 0.0;
    const double vel = 0.5 * (1.0 - mu * mu) * (a + mu * b);
    va[m] = (m > 0 && m < M) ? 0.5 * (vel + vp) : 0.0;
    vp = vel;
  }
  for (int it = 0; it < nq; ++it) {
    double flo = 0.0;  // zero-flux wall at mu = -1
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if (m < M) {
        double fhi = 0.0;  // zero-flux wall at mu = +1 above the last cell
        if (m + 1 < NM && m + 1 < M) fhi = va[m + 1] * (va[m + 1] >= 0.0 ? f[m] : f[m + 1]);
        const double x = f[m] - dtq * ((fhi - flo) / D.dmu);
        f[m] = x < kDblMin ? kDblMin : x;
        flo = fhi;
      }
    }
  }
}

}  // namespace ep
