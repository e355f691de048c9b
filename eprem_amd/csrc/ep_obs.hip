// ep_obs.hip -- point observers and omnidirectional flux on the device.
//
// k_obs_project: one block per (stream, observer).  Finds the shells that
//   bracket the observer sphere, intersects the segment with the sphere
//   (pointObsProjectedStreamData + findIntersection, src/search.c:113-268),
//   interpolates the stream's f there and forms the inverse-distance weight
//   (src/unifiedOutput.c:880-898).
// k_obs_reduce: one block per observer.  Sums weight * projection over the
//   streams in stream order (the reference's loop order, so the rounding
//   matches), divides by the weight sum and forms 2 E <f>_mu
//   (computeFlux, src/unifiedOutput.c:153-180).
// k_stream_flux: computeFlux for every node.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ep_internal.h"
#include "ep_state.h"

namespace ep {

struct ObsDev {
  int num;
  double idw_p;
  const double *r, *theta, *phi;  // [num]
  double *proj;                   // [num][S][E*M]
  double *weight;                 // [num][S]
  double *dist, *flux;            // [num][E*M], [num][E]
};

template <typename T>
__global__ __launch_bounds__(256) void k_obs_project(Dev D, ObsDev O) {
  const int s = blockIdx.x, o = blockIdx.y;
  const int n = D.n_active[s];
  const size_t on = (size_t)s * D.Npad;
  __shared__ int s_in, s_out;
  __shared__ double s_fac;
  const double rSph = O.r[o] / D.rScale;
  if (threadIdx.x == 0) {
    const int ss0 = 0, ss1 = n - 1;
    int in = ss0, out = ss1;
    while ((D.rmag[on + ((in < ss1) ? (in + 1) : ss1)] < rSph) && (in < ss1)) in++;
    while ((D.rmag[on + ((out > ss0) ? (out - 1) : ss0)] >= rSph) && (out > ss0)) out--;
    const double x0 = D.x[on + in], y0 = D.y[on + in], z0 = D.z[on + in];
    const double x1 = D.x[on + out], y1 = D.y[on + out], z1 = D.z[on + out];
    double sf = 0.0;
    if (out > in) {
      const double dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
      const double r0square = x0 * x0 + y0 * y0 + z0 * z0;
      const double dx2 = (dx * dx + dy * dy + dz * dz);
      const double b = (x0 * dx + y0 * dy + z0 * dz) / dx2;
      const double c = (rSph * rSph - r0square) / dx2;
      sf = -b + sqrt(b * b + c);
    }
    const double px = x0 * (1.0 - sf) + x1 * sf;
    const double py = y0 * (1.0 - sf) + y1 * sf;
    const double pz = z0 * (1.0 - sf) + z1 * sf;
    double ph = O.phi[o];
    const double th = O.theta[o];
    if (ph < 0.0) ph += 2.0 * kPI;
    if (ph > (2.0 * kPI)) ph -= 2.0 * kPI;
    const double rs = O.r[o] / D.rScale;
    const double cx = rs * sin(th) * cos(ph), cy = rs * sin(th) * sin(ph), cz = rs * cos(th);
    const double distance = sqrt((cx - px) * (cx - px) + (cy - py) * (cy - py) + (cz - pz) * (cz - pz));
    O.weight[(size_t)o * D.S + s] = pow(distance, -1.0 * O.idw_p);
    s_in = in;
    s_out = out;
    s_fac = sf;
  }
  __syncthreads();
  const int in = s_in, out = s_out;
  const double sf = s_fac;
  const int EM = D.E * D.M;
  double *pr = O.proj + ((size_t)o * D.S + s) * EM;
  for (int c = threadIdx.x; c < EM; c += blockDim.x) {
    const int e = c / D.M, m = c % D.M;
    const int task = s * D.E + e;
    const T *fb = fslab<T>(D, D.loc[task], task) + (size_t)m * D.Npad;
    const double fi = (double)f_dec(D, fb[in]), fo = (double)f_dec(D, fb[out]);
    double v;
    if ((sf >= 0.0) && (sf <= 1.0))
      v = fi * (1.0 - sf) + fo * sf;
    else if (sf < 0.0)
      v = fi;
    else
      v = fo;
    pr[c] = v;
  }
}

__global__ __launch_bounds__(256) void k_obs_reduce(Dev D, ObsDev O) {
  const int o = blockIdx.x;
  const int EM = D.E * D.M;
  const double *w = O.weight + (size_t)o * D.S;
  double wsum = 0.0;
  for (int s = 0; s < D.S; s++) wsum += w[s];
  for (int c = threadIdx.x; c < EM; c += blockDim.x) {
    double acc = 0.0;
    for (int s = 0; s < D.S; s++) acc += (w[s] * O.proj[((size_t)o * D.S + s) * EM + c]);
    O.dist[(size_t)o * EM + c] = acc / wsum;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < D.E; e += blockDim.x) {
    double iso = 0.0;
    for (int m = 0; m < D.M; m++) iso += O.dist[(size_t)o * EM + e * D.M + m];
    iso /= D.M;
    O.flux[(size_t)o * D.E + e] = 2.0 * D.egrid[e] * iso;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_stream_flux(Dev D, double *flux) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k >= D.N) return;
  const int task = s * D.E + e;
  const T *fb = fslab<T>(D, D.loc[task], task) + k;
  double iso = 0.0;
  for (int m = 0; m < D.M; m++) iso += (double)f_dec(D, fb[(size_t)m * D.Npad]);
  iso /= D.M;
  flux[((size_t)s * D.N + k) * D.E + e] = 2.0 * D.egrid[e] * iso;
}

}  // namespace ep

using namespace ep;

#define HIPCHK_OBS(x, what)                              \
  do {                                                   \
    hipError_t e_ = (x);                                 \
    if (e_ != hipSuccess) return hip_fail(e_, what);     \
  } while (0)

extern "C" {

}  // extern "C"

namespace ep {

// Doubles of device memory ep_observe needs for O observers: their
// coordinates, the per-stream projections and weights, the blended
// distribution and the flux.
static size_t observe_need(int O, int S, int E, int M) {
  const size_t EM = (size_t)E * M;
  return 3 * (size_t)O + (size_t)O * S * EM + (size_t)O * S + (size_t)O * EM + (size_t)O * E;
}

static int check_observers(const ep_observers *obs) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  if (!obs || obs->num < 1 || obs->num > 256 || !obs->r_au || !obs->theta || !obs->phi)
    return fail(EP_ERR_INVALID_ARG, "bad observer list (1..256 observers)");
  for (int o = 0; o < obs->num; o++)
    if (!(obs->r_au[o] > 0) || !std::isfinite(obs->theta[o]) || !std::isfinite(obs->phi[o]))
      return fail(EP_ERR_INVALID_ARG, "observer %d: need r > 0 and finite angles", o);
  return EP_OK;
}

// The observer kernels on the device buffer b (observe_need doubles);
// every word of b they read they have written first.
static int observe_on(const ep_observers *obs, double *dist, double *flux, double *b) {
  const int O = obs->num, S = g.d.num_streams, E = g.d.num_energies, M = g.d.num_mu;
  const size_t EM = (size_t)E * M;
  ObsDev od;
  od.num = O;
  od.idw_p = obs->idw_p;
  double *rr = b, *tt = b + O, *pp = b + 2 * O;
  od.r = rr;
  od.theta = tt;
  od.phi = pp;
  od.proj = b + 3 * O;
  od.weight = od.proj + (size_t)O * S * EM;
  od.dist = od.weight + (size_t)O * S;
  od.flux = od.dist + (size_t)O * EM;
  hipError_t e = hipMemcpy(rr, obs->r_au, O * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tt, obs->theta, O * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pp, obs->phi, O * 8, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "observer upload");
  Dev D = make_dev();
  if (D.fp32)
    hipLaunchKernelGGL(k_obs_project<float>, dim3(S, O), dim3(256), 0, g.st, D, od);
  else
    hipLaunchKernelGGL(k_obs_project<double>, dim3(S, O), dim3(256), 0, g.st, D, od);
  hipLaunchKernelGGL(k_obs_reduce, dim3(O), dim3(256), 0, g.st, D, od);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e != hipSuccess) return hip_fail(e, "observer kernels");
  if (dist) {
    e = hipMemcpy(dist, od.dist, (size_t)O * EM * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "observer download");
  }
  if (flux) {
    e = hipMemcpy(flux, od.flux, (size_t)O * E * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "observer download");
  }
  return EP_OK;
}

// Fixed scratch of ep_observe_noalloc: 32 MiB that each GPU the library
// runs on holds from the module's load until the process ends.  In host
// code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, once per device (scratch_device), and only that
// address is handed to kernels and copies.
constexpr size_t kScratchDoubles = (size_t)4 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *g_scratch_dev[kMaxDevices];

// The device address of the current device's scratch: looked up with
// hipGetSymbolAddress once per device, with that device current (ep_init
// does it for its device; a later call on another device does it there).
int scratch_address(double **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(EP_ERR_NO_DEVICE, "device %d out of range", dev);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = (double *)p;
  }
  *ptr = g_scratch_dev[dev];
  return EP_OK;
}

}  // namespace ep

extern "C" {

int ep_observe(const ep_observers *obs, double *dist, double *flux) {
  int rc = check_observers(obs);
  if (rc) return rc;
  const size_t need = observe_need(obs->num, g.d.num_streams, g.d.num_energies, g.d.num_mu);
  if (need > g.obs_cap) {
    if (g.obs_buf) (void)hipFree(g.obs_buf);
    g.obs_buf = nullptr;
    g.obs_cap = 0;
    hipError_t e = hipMalloc(&g.obs_buf, need * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "observer scratch");
    g.obs_cap = need;
  }
  return observe_on(obs, dist, flux, g.obs_buf);
}

int ep_observe_noalloc(const ep_observers *obs, double *dist, double *flux) {
  int rc = check_observers(obs);
  if (rc) return rc;
  const size_t need = observe_need(obs->num, g.d.num_streams, g.d.num_energies, g.d.num_mu);
  if (need > kScratchDoubles)
    return fail(EP_ERR_INVALID_ARG,
                "ep_observe_noalloc: %d observers x %d streams x %d x %d cells need %zu bytes of scratch, "
                "more than its %zu",
                obs->num, g.d.num_streams, g.d.num_energies, g.d.num_mu, need * sizeof(double),
                kScratchDoubles * sizeof(double));
  double *b = nullptr;
  if ((rc = scratch_address(&b))) return rc;
  return observe_on(obs, dist, flux, b);
}

int ep_stream_flux(double *flux) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!flux) return fail(EP_ERR_INVALID_ARG, "null flux");
  const size_t cnt = (size_t)g.d.num_streams * g.d.num_nodes * g.d.num_energies;
  hipError_t e;
  if (cnt > g.flux_cap) {
    if (g.flux_buf) (void)hipFree(g.flux_buf);
    g.flux_buf = nullptr;
    g.flux_cap = 0;
    e = hipMalloc(&g.flux_buf, cnt * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "flux buffer");
    g.flux_cap = cnt;
  }
  double *out = g.flux_buf;
  Dev D = make_dev();
  if (D.fp32)
    hipLaunchKernelGGL(k_stream_flux<float>, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, g.st, D, out);
  else
    hipLaunchKernelGGL(k_stream_flux<double>, dim3((D.N + 255) / 256, D.E, D.S), dim3(256), 0, g.st, D, out);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e == hipSuccess) e = hipMemcpy(flux, out, cnt * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "stream flux");
  return EP_OK;
}

}  // extern "C"
