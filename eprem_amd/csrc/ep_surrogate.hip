// ep_surrogate.hip -- training pairs for the learned large-step surrogate.
//
// The surrogate maps the distribution around one kept node of one
// (stream, energy) slab to the distribution at that node after K
// streaming and relaxation subcycles (DiffuseStreamData,
// src/energeticParticles.c:2352-2430) followed by the adiabatic focusing
// of the same K dtProp (AdiabaticFocusing with the upwind operator,
// :1683-1762, :1806-1896).  Pairs come from the plain solver: one
// streaming pass of exactly K subcycles (halo = K; the result goes to the
// spare f buffer, so f and the step state are unchanged), then the
// focusing of the sampled row.
//
// Inputs of a sample (3M + 6), per mu cell: ln f(k-1) - ln f(k), then
// ln f(k) minus its mean over mu, then ln f(k+1) - ln f(k) (neighbours as
// differences, which bf16 keeps to its relative precision), then the six
// scalars of sur_scalars (ep_sur.h).  Target (M): ln f at node k after
// the K subcycles and the focusing, minus the same mean.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ep_internal.h"
#include "ep_launch.h"
#include "ep_state.h"
#include "ep_sur.h"

namespace ep {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One thread per sample; rows whose random (stream, energy, node) does not
// qualify (too few subcycles, thinned or short stream, end node) get
// valid = 0.
__global__ __launch_bounds__(256) void k_samples(Dev D, int K, int n, uint64_t seed, float *X, float *Y,
                                                 int32_t *valid, int32_t *where) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int M = D.M, nin = 3 * M + 6;
  const uint64_t h = mix64(seed ^ (uint64_t)r * 0xD1B54A32D192ED03ull);
  const int s = (int)(h % (uint64_t)D.S);
  const int e = (int)((h >> 20) % (uint64_t)D.E);
  const int task = s * D.E + e;
  const int nk = D.nkept[s];
  const int ok = (D.thin[s] == 0) && (D.nsteps[task] >= K) && (nk >= 8);
  const int kk = ok ? 1 + (int)((h >> 40) % (uint64_t)(nk - 2)) : 1;
  valid[r] = ok;
  where[(size_t)r * 3] = s;
  where[(size_t)r * 3 + 1] = e;
  where[(size_t)r * 3 + 2] = kk;
  if (!ok) return;
  const size_t on = (size_t)s * D.Npad;
  const double *cur = fslab(D, D.loc[task], task);
  const double *aft = fslab(D, D.loc[task] ^ 1, task);
  float *x = X + (size_t)r * nin;
  double c = 0.0;
  for (int m = 0; m < M; m++) c += log(cur[(size_t)m * D.Npad + kk]);
  c /= M;
  for (int m = 0; m < M; m++) {
    const double lk = log(cur[(size_t)m * D.Npad + kk]);
    x[m] = (float)(log(cur[(size_t)m * D.Npad + kk - 1]) - lk);
    x[M + m] = (float)(lk - c);
    x[2 * M + m] = (float)(log(cur[(size_t)m * D.Npad + kk + 1]) - lk);
  }
  double xs[6];
  sur_scalars(D, s, e, kk, K, xs);
  for (int j = 0; j < 6; j++) x[3 * M + j] = (float)xs[j];
  double fr[32];
#pragma unroll
  for (int m = 0; m < 32; m++) fr[m] = (m < M) ? aft[(size_t)m * D.Npad + kk] : 1.0;
  sur_focus_row<32>(D, on + kk, e, (double)K * D.dtprop[task], K, D.nsteps[task], fr);
  float *y = Y + (size_t)r * M;
  for (int m = 0; m < M; m++) y[m] = (float)(log(fr[m]) - c);
}

}  // namespace ep

using namespace ep;

extern "C" {

// n candidate samples of K subcycles of streaming and focusing on the current nodes and f
// (FP64 storage); X [n][3M+6], Y [n][M] on the host, valid[n] marks the
// rows that qualified, where[n][3] (optional) their (stream, energy, node).  f and the step state are unchanged.
int ep_surrogate_samples(int32_t K, int32_t n, uint64_t seed, float *X, float *Y, int32_t *valid,
                         int32_t *where) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  if (g.fp32) return fail(EP_ERR_UNSUPPORTED, "samples need FP64 storage");
  if (K < 1 || K > 64 || n < 1 || !X || !Y || !valid) return fail(EP_ERR_INVALID_ARG, "bad sample request");
  const int M = g.d.num_mu, nin = 3 * M + 6;
  const int keep_b = g.stream_block, keep_h = g.stream_halo, keep_r = g.stream_r, keep_w = g.stream_whole;
  g.stream_block = 256;
  g.stream_whole = 0;  // every slab advances exactly K subcycles, short streams too
  g.stream_halo = K;
  g.stream_r = 1;
  Dev D = make_dev();
  float *dx = nullptr, *dy = nullptr;
  int32_t *dv = nullptr, *dw = nullptr;
  hipError_t e = hipMalloc(&dx, (size_t)n * nin * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dy, (size_t)n * M * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dv, (size_t)n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&dw, (size_t)n * 3 * sizeof(int32_t));
  DevDiag hd;
  if (e == hipSuccess) e = launch_reset_diag(g.diag, g.st);
  if (e == hipSuccess) e = launch_stream_prep(D, g.st);
  if (e == hipSuccess) e = launch_node_prep(D, g.st);
  if (e == hipSuccess) e = hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, g.st);
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  // pass 0 with halo K: slabs with nsteps >= K advance exactly K subcycles
  if (e == hipSuccess) e = launch_stream_pass(D, 0, 0, 0, 0, hd.max_chunks, 0, g.st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_samples, dim3((n + 255) / 256), dim3(256), 0, g.st, D, K, n, seed, dx, dy, dv, dw);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e == hipSuccess) e = hipMemcpy(X, dx, (size_t)n * nin * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(Y, dy, (size_t)n * M * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(valid, dv, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess && where) e = hipMemcpy(where, dw, (size_t)n * 3 * sizeof(int32_t), hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipFree(dv);
  (void)hipFree(dw);
  g.stream_block = keep_b;
  g.stream_halo = keep_h;
  g.stream_r = keep_r;
  g.stream_whole = keep_w;
  g.drop_graph();
  if (e != hipSuccess) return hip_fail(e, "surrogate samples");
  if (hd.err_key != ~0ull) return fail(EP_ERR_INVALID_ARG, "node preparation reported an error");
  return EP_OK;
}

}  // extern "C"
