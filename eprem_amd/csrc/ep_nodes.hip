// ep_nodes.hip -- node motion on the device: shell spawn (the ripple of
// every node record one shell outward) and the analytic-wind node move, so
// that a global step needs no node data from the host.
//
// k_ripple_nodes  rippleShellsOut + spawnNewShell (src/simCore.c:629-728):
//                 node k takes node k-1's record, node 0 the new shell.
//                 One block per stream shifts its arrays in place, from the
//                 outer end inward, a block-sized chunk at a time.
// k_spawn_f       the same shift of f without moving it: every f row holds
//                 the nodes from a head slot on (fidx), so the new node 0
//                 goes into the slot before the head and the head moves
//                 back by one; the outermost node of a full stream simply
//                 falls off the end.
// k_ripple_f      when the head has reached slot 0: the shift as a copy of
//                 every slab into the spare buffer, to the far end of its
//                 rows (k_flip_loc then makes it current) -- once every
//                 Npad - N spawns.
// k_move_nodes    moveNodes with delrFlow for the analytic wind
//                 (src/simCore.c:736-785, src/flow.c:169-178): r += tDel V/rScale
//                 with V made Cartesian at the node (src/flow.c:355-403).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ep_internal.h"
#include "ep_launch.h"
#include "ep_state.h"

namespace ep {

// Node arrays that move with a node record.  The MHD state (old position,
// B, V, density) lives in the MHD buffer kept by ep_update_mhd.
constexpr int kRippleFields = 18;

struct RippleArgs {
  double *field[kRippleFields];
  const double *spawn;  // [4][S]: x, y, z, rmag of the new innermost node
  int32_t *n_active;
  int rold;             // index of the first of the three old-position fields
};

__global__ __launch_bounds__(256) void k_ripple_nodes(Dev D, RippleArgs A) {
  const int s = blockIdx.x;
  const int n_old = A.n_active[s];
  const int n_new = n_old < D.N ? n_old + 1 : D.N;
  const size_t o = (size_t)s * D.Npad;
  // Shift k-1 -> k for k = n_new-1 .. 1, chunks from the outer end: every
  // read of a chunk happens before any write that could overwrite it.
  for (int top = n_new - 1; top >= 1; top -= blockDim.x) {
    const int k = top - (int)threadIdx.x;
    double v[kRippleFields];
    if (k >= 1) {
#pragma unroll
      for (int q = 0; q < kRippleFields; ++q) v[q] = A.field[q][o + k - 1];
    }
    __syncthreads();
    if (k >= 1) {
#pragma unroll
      for (int q = 0; q < kRippleFields; ++q) A.field[q][o + k] = v[q];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // the new innermost node; its MHD state is not read before
    // ep_update_mhd recomputes it (derivatives are zero on node 0)
    for (int q = 0; q < 4; ++q) A.field[q][o] = A.spawn[(size_t)q * D.S + s];
    for (int q = 0; q < 3; ++q) A.field[A.rold + q][o] = A.spawn[(size_t)q * D.S + s];
  }
  __syncthreads();
  if (threadIdx.x == 0) A.n_active[s] = n_new;
}

// The distribution of the new node 0 of stream s, energy e, mu m:
// spawn_f[s][e][m], or the seed spectrum at the new node's radius when
// spawn_f is null.
template <typename T>
__device__ inline T spawn_value(const Dev &D, const double *spawn_f, int s, int e, int m) {
  if (spawn_f) return (T)spawn_f[((size_t)s * D.E + e) * D.M + m];
  const double radial = pow(D.rmag[(size_t)s * D.Npad] / D.normRadius, -D.beta);
  return (T)seed_value(D, radial, e);
}

// New head slot new_head = old head - 1 >= 0: node 0 of every row written
// there, in the current buffer; the head itself is advanced by the first
// lane (no lane here reads it).
template <typename T>
__global__ __launch_bounds__(256) void k_spawn_f(Dev D, const double *spawn_f, int new_head) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;  // (energy, mu) of stream blockIdx.y
  const int s = blockIdx.y;
  if (q == 0 && s == 0) D.fhead[0] = new_head;
  if (q >= D.E * D.M) return;
  const int e = q / D.M, m = q % D.M;
  const int task = s * D.E + e;
  T *row = fbufT<T>(D, D.loc[task]) + ((size_t)task * D.M + m) * D.Npad;
  row[new_head] = f_enc(D, spawn_value<T>(D, spawn_f, s, e, m));
}

// Head at slot 0: dst[m][new_head + k] = src[m][k - 1] (old head 0) into
// the spare buffer, node 0 from the spawn; new_head = Npad - N - 1, so a
// full stream's nodes end at slot Npad - 2.  Source and copy use the same
// element type, so a slab occupies the same offsets in both buffers.
template <typename T>
__global__ __launch_bounds__(256) void k_ripple_f(Dev D, const double *spawn_f, int new_head) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y, s = blockIdx.z;
  if (k == 0 && e == 0 && s == 0) D.fhead[0] = new_head;
  if (k >= D.n_active[s]) return;  // n_active already advanced
  const int task = s * D.E + e;
  const int cur = D.loc[task];
  const T *src = fbufT<T>(D, cur) + (size_t)task * D.M * D.Npad;
  T *dst = fbufT<T>(D, cur ^ 1) + (size_t)task * D.M * D.Npad + new_head;
  if (k > 0) {
    for (int m = 0; m < D.M; m++) dst[(size_t)m * D.Npad + k] = src[(size_t)m * D.Npad + k - 1];
  } else {
    for (int m = 0; m < D.M; m++) dst[(size_t)m * D.Npad] = f_enc(D, spawn_value<T>(D, spawn_f, s, e, m));
  }
}

__global__ __launch_bounds__(256) void k_flip_loc(Dev D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < D.S * D.E) D.loc[t] ^= 1;
}

// state: the MHD buffer, fields [0..2] old position, [3..5] B, [6..8] V,
// [9] density, each [S*Npad].
__global__ __launch_bounds__(256) void k_move_nodes(Dev D, double *x, double *y, double *z,
                                                    double *rmag, double *state, double tdel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (k >= D.n_active[s]) return;
  const size_t i = (size_t)s * D.Npad + k;
  const size_t SN = (size_t)D.S * D.Npad;
  double rx = x[i], ry = y[i], rz = z[i];
  state[i] = rx;
  state[SN + i] = ry;
  state[2 * SN + i] = rz;
  if (k == 0) return;  // the freshly spawned shell does not move this step
  const double vr = state[6 * SN + i], vt = state[7 * SN + i], vp = state[8 * SN + i];
  const double theta = acos(rz / rmag[i]);
  const double phi = atan2(ry, rx);
  const double vvx = vr * sin(theta) * cos(phi) + vt * cos(theta) * cos(phi) - vp * sin(phi);
  const double vvy = vr * sin(theta) * sin(phi) + vt * cos(theta) * sin(phi) + vp * cos(phi);
  const double vvz = vr * cos(theta) - vt * sin(theta);
  const double dx = tdel * vvx / D.rScale;
  const double dy = tdel * vvy / D.rScale;
  const double dz = tdel * vvz / D.rScale;
  rx += dx;
  ry += dy;
  rz += dz;
  x[i] = rx;
  y[i] = ry;
  z[i] = rz;
  rmag[i] = sqrt((rx * rx) + (ry * ry) + (rz * rz));
}

}  // namespace ep

using namespace ep;

int ensure_mhd_state();  // ep_mas.hip

extern "C" {

int ep_spawn_shell(const ep_spawn *sp) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set && !g.positions_set) return fail(EP_ERR_INVALID_ARG, "set node positions first");
  if (!sp || !sp->x || !sp->y || !sp->z || !sp->rmag)
    return fail(EP_ERR_INVALID_ARG, "spawn positions are required");
  const int S = g.d.num_streams, E = g.d.num_energies, M = g.d.num_mu;
  for (int s = 0; s < S; s++)
    if (!(sp->rmag[s] > 0)) return fail(EP_ERR_INVALID_ARG, "spawn radius of stream %d must be > 0", s);
  int rc = ensure_mhd_state();
  if (rc) return rc;
  // staging: 4*S positions, then optionally S*E*M distribution values
  const size_t need = 4 * (size_t)S + (sp->f ? (size_t)S * E * M : 0);
  if (need > g.spawn_cap) {
    if (g.spawn_buf) (void)hipFree(g.spawn_buf);
    g.spawn_buf = nullptr;
    g.spawn_cap = 0;
    hipError_t e = hipMalloc(&g.spawn_buf, need * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "spawn staging");
    g.spawn_cap = need;
  }
  const double *pos[4] = {sp->x, sp->y, sp->z, sp->rmag};
  for (int q = 0; q < 4; q++) {
    hipError_t e = hipMemcpyAsync(g.spawn_buf + (size_t)q * S, pos[q], S * 8, hipMemcpyHostToDevice, g.st);
    if (e != hipSuccess) return hip_fail(e, "spawn upload");
  }
  const double *spawn_f = nullptr;
  if (sp->f) {
    hipError_t e = hipMemcpyAsync(g.spawn_buf + 4 * (size_t)S, sp->f, (size_t)S * E * M * 8,
                                  hipMemcpyHostToDevice, g.st);
    if (e != hipSuccess) return hip_fail(e, "spawn upload");
    spawn_f = g.spawn_buf + 4 * (size_t)S;
  }
  const size_t SN = (size_t)S * g.Npad;
  RippleArgs A;
  double *fields[kRippleFields] = {g.nx, g.ny, g.nz, g.nrmag, g.nbmag, g.ndlnb, g.ndlnn, g.ndupar};
  for (int q = 0; q < 10; q++) fields[8 + q] = g.mas_prev + q * SN;
  for (int q = 0; q < kRippleFields; q++) A.field[q] = fields[q];
  A.spawn = g.spawn_buf;
  A.n_active = g.n_active;
  A.rold = 8;
  Dev D = make_dev();
  hipLaunchKernelGGL(k_ripple_nodes, dim3(S), dim3(256), 0, g.st, D, A);
  if (g.fhead_h > 0) {
    // the new node 0 goes into the slot before the head
    const int nh = g.fhead_h - 1;
    const dim3 grid((E * M + 255) / 256, S);
    if (D.fp32)
      hipLaunchKernelGGL(k_spawn_f<float>, grid, dim3(256), 0, g.st, D, spawn_f, nh);
    else
      hipLaunchKernelGGL(k_spawn_f<double>, grid, dim3(256), 0, g.st, D, spawn_f, nh);
    g.fhead_h = nh;
  } else {
    // no slot left before the rows: copy them to the far end
    const int nh = g.Npad - g.d.num_nodes - 1;
    const dim3 grid((g.d.num_nodes + 255) / 256, E, S);
    if (D.fp32)
      hipLaunchKernelGGL(k_ripple_f<float>, grid, dim3(256), 0, g.st, D, spawn_f, nh);
    else
      hipLaunchKernelGGL(k_ripple_f<double>, grid, dim3(256), 0, g.st, D, spawn_f, nh);
    hipLaunchKernelGGL(k_flip_loc, dim3((S * E + 255) / 256), dim3(256), 0, g.st, D);
    g.fhead_h = nh;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e != hipSuccess) return hip_fail(e, "spawn kernels");
  for (int s = 0; s < S; s++)
    if (g.h_nactive[s] < g.d.num_nodes) g.h_nactive[s]++;
  return EP_OK;
}

int ep_move_nodes(double tdel) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.mhd_state) return fail(EP_ERR_INVALID_ARG, "ep_update_mhd first (node velocities)");
  if (!(tdel >= 0) || !std::isfinite(tdel)) return fail(EP_ERR_INVALID_ARG, "tdel must be finite and >= 0");
  Dev D = make_dev();
  hipLaunchKernelGGL(k_move_nodes, dim3((g.d.num_nodes + 255) / 256, g.d.num_streams), dim3(256), 0,
                     g.st, D, g.nx, g.ny, g.nz, g.nrmag, g.mas_prev, tdel);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e != hipSuccess) return hip_fail(e, "move nodes");
  return EP_OK;
}

int ep_get_positions(double *x, double *y, double *z, double *rmag, int32_t *n_active) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  const size_t S = g.d.num_streams, N = g.d.num_nodes;
  double *dst[4] = {x, y, z, rmag};
  const double *src[4] = {g.nx, g.ny, g.nz, g.nrmag};
  for (int q = 0; q < 4; q++)
    if (dst[q]) {
      hipError_t e = hipMemcpy2D(dst[q], N * 8, src[q], g.Npad * 8, N * 8, S, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "position download");
    }
  if (n_active) memcpy(n_active, g.h_nactive.data(), S * 4);
  return EP_OK;
}

}  // extern "C"
