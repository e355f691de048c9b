// ep_shell.hip -- shell diffusion: perpendicular diffusion between the
// streams of one shell (ShellData + DiffuseShellData,
// src/energeticParticles.c:247-555), the only cross-stream operator.
//
// k_shell_links   once per global step: per node and link to the four
//                 neighbour streams on the same node index, the link length
//                 perpendicular to B (B made Cartesian at the node from the
//                 device MHD state, src/flow.c:339-349).
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
  const double *state; // MHD state buffer (B at fields 3..5)
  double kperxkpar;
};
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
    ex /= dl;
    ey /= dl;
    ez /= dl;
    const double dot = ex * ebx + ey * eby + ez * ebz;
    H.dlper[i * 4 + d] = dl * sqrt(1.0 - dot * dot);
  }
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
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.N) return;
  const int task = s * D.E + e;
  const T *fs = fbufT<T>(D, D.loc[task]) + (size_t)task * D.M * D.Npad;
  T *out = fbufT<T>(D, D.loc[task] ^ 1) + (size_t)task * D.M * D.Npad;
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
    srcf[q] = fbufT<T>(D, D.loc[tj]) + (size_t)tj * D.M * D.Npad + k;
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

}  // namespace ep

using namespace ep;

static ShellDev shell_dev() {
  ShellDev H;
  H.nbr = g.shell_nbr;
  H.src = g.shell_src;
  H.dir = g.shell_dir;
  H.dlper = g.shell_dlper;
  H.state = g.mas_prev;
  H.kperxkpar = g.kperxkpar;
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
}  // namespace ep

extern "C" {

// Shell diffusion on (kperxkpar > 0) or off (0).  nbr[S][4]: the north,
// east, west and south neighbour stream of every stream (initNEWS,
// src/cubeShellInit.c:95-520); every link must be returned by the
// neighbour, and all streams must hold the same number of active nodes.
int ep_shell_diffusion(const int32_t *nbr, double kperxkpar) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!(kperxkpar >= 0) || !std::isfinite(kperxkpar)) return fail(EP_ERR_INVALID_ARG, "kperxkpar must be >= 0");
  g.drop_graph();
  if (kperxkpar == 0.0) {
    g.kperxkpar = 0.0;
    return EP_OK;
  }
  if (!nbr) return fail(EP_ERR_INVALID_ARG, "null neighbour table");
  const int S = g.d.num_streams;
  std::vector<std::vector<std::pair<int, int>>> in(S);
  for (int s = 0; s < S; s++)
    for (int d = 0; d < 4; d++) {
      const int j = nbr[s * 4 + d];
      if (j < 0 || j >= S || j == s) return fail(EP_ERR_INVALID_ARG, "neighbour %d of stream %d out of range", d, s);
      in[j].push_back({s, d});
    }
  std::vector<int32_t> src((size_t)S * kMaxIn, -1), dir((size_t)S * kMaxIn, -1);
  for (int s = 0; s < S; s++) {
    in[s].push_back({s, 4});  // own loss, at the stream's own position in the source order
    std::stable_sort(in[s].begin(), in[s].end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    if ((int)in[s].size() > kMaxIn) return fail(EP_ERR_INVALID_ARG, "stream %d has too many incoming links", s);
    for (size_t q = 0; q < in[s].size(); q++) {
      src[(size_t)s * kMaxIn + q] = in[s][q].first;
      dir[(size_t)s * kMaxIn + q] = in[s][q].second;
    }
  }
  if (!g.shell_nbr) {
    hipError_t e = hipMalloc(&g.shell_nbr, (size_t)S * 4 * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&g.shell_src, (size_t)S * kMaxIn * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&g.shell_dir, (size_t)S * kMaxIn * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc(&g.shell_dlper, (size_t)S * g.Npad * 4 * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "shell diffusion tables");
  }
  hipError_t e = hipMemcpy(g.shell_nbr, nbr, (size_t)S * 4 * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g.shell_src, src.data(), src.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g.shell_dir, dir.data(), dir.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "shell diffusion tables");
  g.kperxkpar = kperxkpar;
  return EP_OK;
}

}  // extern "C"
