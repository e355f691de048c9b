// ep_shell.hip -- the cross-stream operators: perpendicular diffusion
// (ShellData + DiffuseShellData, src/energeticParticles.c:247-555) and
// drift (DriftShellData with driftVelocity, :208-245, :560-801) between the
// streams of one shell.
//
// k_shell_links   once per global step: per node and link to the four
//                 neighbour streams on the same node index, the link length
//                 and its part perpendicular to B (B made Cartesian at the
//                 node from the device MHD state, src/flow.c:339-349), and
//                 for drift curl(B/B^2) of the analytic field
//                 (curlBoverB2, src/flow.c:479-615).
// k_shell_drift   per substep after diffusion: moves del = vdp dt / dl of a
//                 stream's f to each neighbour its drift velocity points to.
// k_shell_diffuse per substep, after focusing and change: the reference
//                 scatters del * f from every stream to its neighbours; here
//                 each stream gathers the same terms in the same order (its
//                 sources sorted by stream index, the stream's own loss at
//                 its own index), writing the spare buffer, so the result is
//                 bit-identical to the scatter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ep_internal.h"
#include "ep_launch.h"
#include "ep_state.h"

namespace ep {

constexpr double kThresh = 0.025;      // THRESH, src/baseTypes.h:106
constexpr double kVerySmall = 1.0e-33;  // VERYSMALL, src/baseTypes.h:50

struct ShellDev {
  const int32_t *nbr;  // [S][4]
  const int32_t *src;  // [S][kMaxIn]: source stream, -1 = unused
  const int32_t *dir;  // [S][kMaxIn]: link of the source (0..3), 4 = own loss
  double *dlper;       // [S*Npad][4]
  double *dl;          // [S*Npad][4] link lengths (AU)
  double *curl;        // [S*Npad][3] curl(B/B^2), drift only
  const double *factor;  // [E] drift velocity per unit curl (driftVelocity)
  const double *state; // MHD state buffer (B at fields 3..5, V at 6..8)
  double kperxkpar;
  int drift;
  double mhdBsAu, omega;  // analytic field for curlBoverB2
};

// parkerB (src/flow.c:440-470) without the ideal shock
__device__ __forceinline__ void parker_b(double r, double theta, double vr, double bs, double om, double *b) {
  b[0] = bs / (r * r);
  b[1] = 0.0;
  b[2] = -1.0 * r * b[0] * (om / (vr + kVerySmall)) * sin(theta);
}
constexpr int kMaxIn = 9;

__global__ __launch_bounds__(256) void k_shell_links(Dev D, ShellDev H) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (k < 1 || k >= D.N) return;
  const size_t SN = (size_t)D.S * D.Npad;
  const size_t i = (size_t)s * D.Npad + k;
  const double theta = acos(D.z[i] / D.rmag[i]);
  const double phi = atan2(D.y[i], D.x[i]);
  const double br = H.state[3 * SN + i], bt = H.state[4 * SN + i], bp = H.state[5 * SN + i];
  const double bx = br * sin(theta) * cos(phi) + bt * cos(theta) * cos(phi) - bp * sin(phi);
  const double by = br * sin(theta) * sin(phi) + bt * cos(theta) * sin(phi) + bp * cos(phi);
  const double bz = br * cos(theta) - bt * sin(theta);
  const double bmag = sqrt(br * br + bt * bt + bp * bp);
  const double ebx = bx / bmag, eby = by / bmag, ebz = bz / bmag;
  const double rx = D.x[i] * D.rScale, ry = D.y[i] * D.rScale, rz = D.z[i] * D.rScale;
  for (int d = 0; d < 4; d++) {
    const size_t j = (size_t)H.nbr[s * 4 + d] * D.Npad + k;
    double ex = D.x[j] * D.rScale - rx, ey = D.y[j] * D.rScale - ry, ez = D.z[j] * D.rScale - rz;
    const double dl = sqrt(ex * ex + ey * ey + ez * ez);
    H.dl[i * 4 + d] = dl;
    ex /= dl;
    ey /= dl;
    ez /= dl;
    const double dot = ex * ebx + ey * eby + ez * ebz;
    H.dlper[i * 4 + d] = dl * sqrt(1.0 - dot * dot);
  }
  if (!H.drift) return;
  // curlBoverB2: centred differences of B/B^2 of the analytic field
  const double rw = 0.001, tw = kPI / 100.0, pw = kPI / 100.0;
  const double r = D.rScale * sqrt(D.x[i] * D.x[i] + D.y[i] * D.y[i] + D.z[i] * D.z[i]);
  const double th = acos(D.z[i] * D.rScale / r);
  const double ph = atan2(D.y[i], D.x[i]);
  const double vr = H.state[6 * SN + i];
  double tp1 = th + tw, tm1 = th - tw;
  if (tp1 > kPI) tp1 = 2.0 * kPI - tp1;
  if (tm1 < 0.0) tm1 *= -1.0;
  (void)ph;  // the analytic field does not depend on phi
  double Brm[3], Brp[3], Btm[3], Btp[3], Bpm[3], Bpp[3];
  parker_b(r - rw, th, vr, H.mhdBsAu, H.omega, Brm);
  parker_b(r + rw, th, vr, H.mhdBsAu, H.omega, Brp);
  parker_b(r, tm1, vr, H.mhdBsAu, H.omega, Btm);
  parker_b(r, tp1, vr, H.mhdBsAu, H.omega, Btp);
  parker_b(r, th, vr, H.mhdBsAu, H.omega, Bpm);
  parker_b(r, th, vr, H.mhdBsAu, H.omega, Bpp);
  auto m2 = [](const double *b) {
    const double m = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    return m * m;
  };
  const double mrm = m2(Brm), mrp = m2(Brp), mtm = m2(Btm), mtp = m2(Btp), mpm = m2(Bpm), mpp = m2(Bpp);
  double A = (0.5 / tw) * (Btp[2] * sin(tp1) / mtp - Btm[2] * sin(tm1) / mtm);
  double B = (0.5 / pw) * (Bpp[1] / mpp - Bpm[1] / mpm);
  H.curl[i * 3 + 0] = (1.0 / (r * sin(th))) * (A - B);
  A = (1.0 / sin(th)) * (0.5 / pw) * (Bpp[0] / mpp - Bpm[0] / mpm);
  B = (0.5 / rw) * ((r + rw) * Brp[2] / mrp - (r - rw) * Brm[2] / mrm);
  H.curl[i * 3 + 1] = (1.0 / r) * (A - B);
  A = (0.5 / rw) * ((r + rw) * Brp[1] / mrp - (r - rw) * Brm[1] / mrm);
  B = (0.5 / tw) * (Btp[0] / mtp - Btm[0] / mtm);
  H.curl[i * 3 + 2] = (1.0 / r) * (A - B);
}

// del of link d of stream s at node k and energy e
__device__ __forceinline__ double shell_del(const Dev &D, const ShellDev &H, int s, int k, int e, int d) {
  const size_t i = (size_t)s * D.Npad + k;
  const double mfp = D.rig_lamo[e] * D.mfp_r[i];
  const double dt_kper = D.dt * 0.33333333 * mfp * D.vgrid[e] * H.kperxkpar;
  const double dp = H.dlper[i * 4 + d];
  double del = dt_kper / (dp * dp + kVerySmall);
  if (del > kThresh) del = kThresh;
  return del;
}

template <typename T>
__global__ __launch_bounds__(256) void k_shell_diffuse(Dev D, ShellDev H, int substep) {
  if (substep_off(D)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.N) return;
  const int task = s * D.E + e;
  const T *fs = fslab<T>(D, D.loc[task], task);
  T *out = fslab<T>(D, D.loc[task] ^ 1, task);
  if (k == 0) {  // not a compute shell: carried over
    for (int m = 0; m < D.M; m++) out[(size_t)m * D.Npad] = fs[(size_t)m * D.Npad];
    return;
  }
  double w[kMaxIn];
  const T *srcf[kMaxIn];
  int nin = 0;
  for (int q = 0; q < kMaxIn; q++) {
    const int j = H.src[s * kMaxIn + q];
    if (j < 0) break;
    const int d = H.dir[s * kMaxIn + q];
    const int tj = j * D.E + e;
    srcf[q] = fslab<T>(D, D.loc[tj], tj) + k;
    if (d < 4) {
      w[q] = shell_del(D, H, j, k, e, d);
    } else {
      const double d0 = shell_del(D, H, s, k, e, 0), d1 = shell_del(D, H, s, k, e, 1);
      const double d2 = shell_del(D, H, s, k, e, 2), d3 = shell_del(D, H, s, k, e, 3);
      w[q] = -(d0 + d1 + d2 + d3);  // the loss term, subtracted below
    }
    nin++;
  }
  int bad = -1, bad_inf = 0;
  for (int m = 0; m < D.M; m++) {
    double delta = 0.0;
    for (int q = 0; q < nin; q++) {
      const double fj = (double)srcf[q][(size_t)m * D.Npad];
      if (w[q] < 0.0)
        delta -= (-w[q]) * fj;
      else
        delta += w[q] * fj;
    }
    const T v = (T)((double)fs[(size_t)m * D.Npad + k] + delta);
    if (bad < 0 && !isfinite(v)) {
      bad = m;
      bad_inf = isinf(v) ? 1 : 0;
    }
    out[(size_t)m * D.Npad + k] = v;
  }
  if (bad >= 0)
    note_error(D, err_key(substep, 2, k, 0, (e << 7) | (s >> 7), s & 0x7F, bad, bad_inf));
}

// vdp (drift speed along link d, AU per code time) and del of stream s
// at node k for energy e; returns del, 0 when the drift points away.
__device__ __forceinline__ double drift_del(const Dev &D, const ShellDev &H, int s, int k, int e, int d) {
  const size_t i = (size_t)s * D.Npad + k;
  const double factor = H.factor[e];
  const double vr = factor * H.curl[i * 3 + 0], vt = factor * H.curl[i * 3 + 1], vp = factor * H.curl[i * 3 + 2];
  // sphToCartVector (src/geometry.c:89-125) at the node position
  const double x = D.x[i], y = D.y[i], z = D.z[i];
  const double rmag = sqrt(x * x + y * y + z * z);
  double vx, vy;
  if ((rmag - fabs(z)) < kVerySmall) {
    vx = 0.0;
    vy = 0.0;
  } else {
    vx = x / rmag * vr + x * z / (rmag * sqrt(rmag * rmag - z * z)) * vt - y / sqrt(rmag * rmag - z * z) * vp;
    vy = y / rmag * vr + y * z / (rmag * sqrt(rmag * rmag - z * z)) * vt + x / sqrt(rmag * rmag - z * z) * vp;
  }
  const double vz = z / rmag * vr - sqrt(rmag * rmag - z * z) / rmag * vt;
  const size_t j = (size_t)H.nbr[s * 4 + d] * D.Npad + k;
  const double ex = D.x[j] - x, ey = D.y[j] - y, ez = D.z[j] - z;
  const double dl = H.dl[i * 4 + d];
  const double vdp = (vx * ex + vy * ey + vz * ez) * D.rScale / dl;
  if (!(vdp > 0.0)) return 0.0;
  double del = vdp * D.dt / dl;
  if (del > kThresh) del = kThresh;
  return del;
}

// DriftShellData as a gather: incoming moves from the sources in stream
// order, the stream's own four outgoing moves (north, east, west, south)
// subtracted one by one at its own position.
template <typename T>
__global__ __launch_bounds__(256) void k_shell_drift(Dev D, ShellDev H, int substep) {
  if (substep_off(D)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.N) return;
  const int task = s * D.E + e;
  const T *fs = fslab<T>(D, D.loc[task], task);
  T *out = fslab<T>(D, D.loc[task] ^ 1, task);
  if (k == 0) {
    for (int m = 0; m < D.M; m++) out[(size_t)m * D.Npad] = fs[(size_t)m * D.Npad];
    return;
  }
  double w[kMaxIn + 3];
  const T *srcf[kMaxIn + 3];
  int sign[kMaxIn + 3];
  int nin = 0;
  for (int q = 0; q < kMaxIn; q++) {
    const int j = H.src[s * kMaxIn + q];
    if (j < 0) break;
    const int d = H.dir[s * kMaxIn + q];
    const int tj = j * D.E + e;
    const T *fp = fslab<T>(D, D.loc[tj], tj) + k;
    if (d < 4) {
      const double del = drift_del(D, H, j, k, e, d);
      if (del > 0.0) {
        w[nin] = del;
        srcf[nin] = fp;
        sign[nin++] = 1;
      }
    } else {
      for (int dd = 0; dd < 4; dd++) {
        const double del = drift_del(D, H, s, k, e, dd);
        if (del > 0.0) {
          w[nin] = del;
          srcf[nin] = fp;
          sign[nin++] = -1;
        }
      }
    }
  }
  int bad = -1, bad_inf = 0;
  for (int m = 0; m < D.M; m++) {
    double delta = 0.0;
    for (int q = 0; q < nin; q++) {
      const double fj = (double)srcf[q][(size_t)m * D.Npad];
      if (sign[q] > 0)
        delta += w[q] * fj;
      else
        delta -= w[q] * fj;
    }
    const T v = (T)((double)fs[(size_t)m * D.Npad + k] + delta);
    if (bad < 0 && !isfinite(v)) {
      bad = m;
      bad_inf = isinf(v) ? 1 : 0;
    }
    out[(size_t)m * D.Npad + k] = v;
  }
  if (bad >= 0)
    note_error(D, err_key(substep, 3, k, 0, (e << 7) | (s >> 7), s & 0x7F, bad, bad_inf));
}

}  // namespace ep

using namespace ep;

static ShellDev shell_dev() {
  ShellDev H;
  H.nbr = g.shell_nbr;
  H.src = g.shell_src;
  H.dir = g.shell_dir;
  H.dlper = g.shell_dlper;
  H.dl = g.shell_dl;
  H.curl = g.shell_curl;
  H.factor = g.shell_factor;
  H.state = g.mas_prev;
  H.kperxkpar = g.kperxkpar;
  H.drift = g.shell_drift;
  H.mhdBsAu = g.mhd_bsau;
  H.omega = g.mhd_omega;
  return H;
}

namespace ep {
hipError_t launch_shell_links(const Dev &D, hipStream_t st) {
  hipLaunchKernelGGL(k_shell_links, dim3((D.N + 255) / 256, D.S), dim3(256), 0, st, D, shell_dev());
  return hipGetLastError();
}

hipError_t launch_shell_diffuse(const Dev &D, int sub, hipStream_t st) {
  const dim3 grid((D.N + 255) / 256, D.E, D.S);
  if (D.fp32)
    hipLaunchKernelGGL(k_shell_diffuse<float>, grid, dim3(256), 0, st, D, shell_dev(), sub);
  else
    hipLaunchKernelGGL(k_shell_diffuse<double>, grid, dim3(256), 0, st, D, shell_dev(), sub);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_flip_all(D, st);
}

hipError_t launch_shell_drift(const Dev &D, int sub, hipStream_t st) {
  const dim3 grid((D.N + 255) / 256, D.E, D.S);
  if (D.fp32)
    hipLaunchKernelGGL(k_shell_drift<float>, grid, dim3(256), 0, st, D, shell_dev(), sub);
  else
    hipLaunchKernelGGL(k_shell_drift<double>, grid, dim3(256), 0, st, D, shell_dev(), sub);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_flip_all(D, st);
}
}  // namespace ep

extern "C" {

// The cross-stream operators: diffusion on while kperxkpar > 0, drift on
// while use_drift != 0.  nbr[S][4]: the north, east, west and south
// neighbour stream of every stream (initNEWS, src/cubeShellInit.c:95-520).
// Both need all streams full; diffusion needs the device MHD state for B,
// drift the analytic-wind field of the last ep_update_mhd (mode 0).
int ep_shell_operators(const int32_t *nbr, double kperxkpar, int32_t use_drift) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!(kperxkpar >= 0) || !std::isfinite(kperxkpar)) return fail(EP_ERR_INVALID_ARG, "kperxkpar must be >= 0");
  g.drop_graph();
  if (kperxkpar == 0.0 && !use_drift) {
    g.kperxkpar = 0.0;
    g.shell_drift = 0;
    return EP_OK;
  }
  if (!nbr) return fail(EP_ERR_INVALID_ARG, "null neighbour table");
  const int S = g.d.num_streams, E = g.d.num_energies;
  std::vector<std::vector<std::pair<int, int>>> in(S);
  for (int s = 0; s < S; s++)
    for (int d = 0; d < 4; d++) {
      const int j = nbr[s * 4 + d];
      if (j < 0 || j >= S || j == s) return fail(EP_ERR_INVALID_ARG, "neighbour %d of stream %d out of range", d, s);
      in[j].push_back({s, d});
    }
  std::vector<int32_t> src((size_t)S * kMaxIn, -1), dir((size_t)S * kMaxIn, -1);
  for (int s = 0; s < S; s++) {
    in[s].push_back({s, 4});  // own terms, at the stream's own position in the source order
    std::stable_sort(in[s].begin(), in[s].end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    if ((int)in[s].size() > kMaxIn) return fail(EP_ERR_INVALID_ARG, "stream %d has too many incoming links", s);
    for (size_t q = 0; q < in[s].size(); q++) {
      src[(size_t)s * kMaxIn + q] = in[s][q].first;
      dir[(size_t)s * kMaxIn + q] = in[s][q].second;
    }
  }
  // driftVelocity's factor (1/om1) (p v / 3), om1 = charge OM / mass
  // (src/energeticParticles.c:217-226, OM src/baseTypes.h:98)
  const double Q = 4.80320425e-10, MP = 1.6726e-24, C = 2.99792458e10, AU = 1.495978707e13;
  const double MHD_B_NORM = std::sqrt(MP * 1.0) * C;
  const double OM = (MHD_B_NORM * Q * AU / (MP * C * C));
  const double om1 = g.p.charge * OM / g.p.mass;
  std::vector<double> factor(E);
  for (int e = 0; e < E; e++) factor[e] = (1.0 / om1) * (g.pgrid[e] * g.vgrid[e] / 3.0);
  if (!g.shell_nbr) {
    hipError_t e = hipMalloc(&g.shell_nbr, (size_t)S * 4 * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&g.shell_src, (size_t)S * kMaxIn * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&g.shell_dir, (size_t)S * kMaxIn * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&g.shell_dlper, (size_t)S * g.Npad * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&g.shell_dl, (size_t)S * g.Npad * 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&g.shell_curl, (size_t)S * g.Npad * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&g.shell_factor, (size_t)E * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "shell operator tables");
  }
  hipError_t e = hipMemcpy(g.shell_nbr, nbr, (size_t)S * 4 * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g.shell_src, src.data(), src.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g.shell_dir, dir.data(), dir.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g.shell_factor, factor.data(), E * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "shell operator tables");
  g.kperxkpar = kperxkpar;
  g.shell_drift = use_drift ? 1 : 0;
  return EP_OK;
}

}  // extern "C"
