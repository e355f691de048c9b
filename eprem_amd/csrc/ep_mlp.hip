// ep_mlp.hip -- a small multilayer perceptron on the matrix cores: forward,
// backward and Adam, for the learned large-step surrogate of streaming.
//
// Every layer is Z = X W + b (ReLU on hidden layers).  Products run on
// MFMA v_mfma_f32_32x32x8_bf16: operands rounded to bf16 (round to nearest
// even), fp32 accumulation; master weights, activations and gradients are
// fp32.  One wave computes a 32 x 32 tile of C = op(A) op(B); the weight
// gradient, whose inner dimension is the batch, is split over the batch
// with atomic accumulation.
//
// Parameters are one flat fp32 vector: for each layer W [in][out]
// (row-major) then b [out].
//
// Inference of the surrogate's shape (input <= 128, hidden layers of 128,
// output <= 32) runs in one kernel, k_mlp_fwd: every layer's weights sit in
// LDS as bf16 MFMA fragments (v_mfma_f32_32x32x16_bf16), and the network
// is evaluated transposed, C^T = W^T X^T, so that each layer's accumulators
// (lane = sample, registers = output features) are already the next
// layer's B operand once bias, ReLU and the bf16 rounding are applied: the
// activations never leave registers.  The permutation of the inner index
// this implies is folded into how the weights are packed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ep_internal.h"
#include "ep_mlp_core.h"
#include "ep_state.h"

namespace ep {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

struct GemmArgs {
  const float *A, *B;
  float *C;
  const float *bias;
  int M, N, K;
  int lda, ldb, ldc;
  int transA, transB;  // element (i,k) of A at A[k*lda+i] when transA, ...
  int relu, atomic;    // epilogue: bias + ReLU, or atomic accumulation
  int kchunk;          // inner range per blockIdx.z (multiple of 8)
};

// MFMA 32x32x8 bf16 operand layout: lane l holds A[l%32][4*(l/32) + q] and
// B[4*(l/32) + q][l%32], q = 0..3; result element r of lane l is
// C[8*(r/4) + 4*(l/32) + r%4][l%32].
__global__ __launch_bounds__(64) void k_gemm(GemmArgs g) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const int li = lane & 31, lk = (lane >> 5) * 4;
  const int i = i0 + li, j = j0 + li;
  f32x16_t acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += 8) {
    bf16x4_t a, b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + lk + q;
      float av = 0.0f, bv = 0.0f;
      if (k < kend) {
        if (i < g.M) av = g.transA ? g.A[(size_t)k * g.lda + i] : g.A[(size_t)i * g.lda + k];
        if (j < g.N) bv = g.transB ? g.B[(size_t)j * g.ldb + k] : g.B[(size_t)k * g.ldb + j];
      }
      a[q] = to_bf16(av);
      b[q] = to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int ii = i0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int jj = j0 + (lane & 31);
    if (ii < g.M && jj < g.N) {
      float v = acc[r];
      if (g.atomic) {
        atomicAdd(&g.C[(size_t)ii * g.ldc + jj], v);
      } else {
        if (g.bias) v += g.bias[jj];
        if (g.relu) v = v > 0.0f ? v : 0.0f;
        g.C[(size_t)ii * g.ldc + jj] = v;
      }
    }
  }
}

// out - y and the mean-squared-error gradient 2 (out - y) / (n * m); the
// loss sum goes to *loss (double, atomically).
__global__ __launch_bounds__(256) void k_mse(const float *out, const float *y, float *dout, size_t count,
                                             float scale, double *loss) {
  __shared__ double part[256];
  double acc = 0.0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (size_t)gridDim.x * blockDim.x) {
    const float d = out[t] - y[t];
    dout[t] = scale * d;
    acc += (double)d * (double)d;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// dZ = dA where the layer's ReLU output is positive, else 0 (in place).
__global__ __launch_bounds__(256) void k_relu_back(const float *act, float *d, size_t count) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (size_t)gridDim.x * blockDim.x)
    if (!(act[t] > 0.0f)) d[t] = 0.0f;
}

// db[j] += sum_i dZ[i][j]
__global__ __launch_bounds__(256) void k_bias_grad(const float *dz, float *db, int n, int m) {
  const int j = blockIdx.x;
  float acc = 0.0f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += dz[(size_t)i * m + j];
  __shared__ float part[256];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] += part[0];
}

__global__ __launch_bounds__(256) void k_zero(float *p, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    p[t] = 0.0f;
}

// Adam (Kingma & Ba 2015) with bias correction factors c1 = 1 - beta1^t,
// c2 = 1 - beta2^t.
__global__ __launch_bounds__(256) void k_adam(float *w, const float *gr, float *m, float *v, size_t n, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const float g = gr[t];
    const float mt = b1 * m[t] + (1.0f - b1) * g;
    const float vt = b2 * v[t] + (1.0f - b2) * g * g;
    m[t] = mt;
    v[t] = vt;
    w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  }
}

// Gather a minibatch: rows idx[0..n) of X and Y.
__global__ __launch_bounds__(256) void k_gather_rows(const float *src, const int32_t *idx, float *dst, int n,
                                                     int width) {
  const size_t total = (size_t)n * width;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t r = t / width, c = t % width;
    dst[t] = src[(size_t)idx[r] * width + c];
  }
}

// ---- fused forward -------------------------------------------------------
// Fragment (ks, ot) of a layer: 64 lanes x 8 bf16.  Lane (i, h) = (l % 32,
// l / 32) holds A[i][8h + q] = W[phi(ks, h, q)][32 ot + i], q = 0..7, where
// phi maps the MFMA's inner index to the layer's input feature:
//   first layer (input rows from memory): phi = 16 ks + 8 h + q;
//   later layers (input = previous accumulators, whose lane-half h holds
//   features 8(r/4) + 4h + r%4 of each 32-feature tile in register r):
//   phi = 16 ks + 8 (q/4) + 4 h + q%4.
// Inputs past `in` and outputs past `out` are zero.
__global__ __launch_bounds__(256) void k_mlp_pack(const float *W, int in, int out, int KS, int OT, int first,
                                                  __bf16 *dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = KS * OT * 64 * 8;
  if (t >= total) return;
  const int q = t & 7, lane = (t >> 3) & 63, f = t >> 9;
  const int ks = f / OT, ot = f % OT;
  const int i = lane & 31, h = lane >> 5;
  const int k = first ? 16 * ks + 8 * h + q : 16 * ks + 8 * (q >> 2) + 4 * h + (q & 3);
  const int j = 32 * ot + i;
  const float v = (k < in && j < out) ? W[(size_t)k * out + j] : 0.0f;
  dst[t] = (__bf16)v;
}

// Y[n][out] = network(X[n][in]); a persistent grid of kFWaves-wave blocks,
// each wave taking kFNT x 32 rows at a time.  LDS: the packed fragments of
// every layer, then the biases (nhid x 128 + 32 floats).
template <int KS0>
__global__ __launch_bounds__(64 * kFWaves) void k_mlp_fwd(const float *__restrict__ X, float *__restrict__ Y,
                                                          int n, int in, int out, int nhid,
                                                          const uint4 *__restrict__ packed, int packed16,
                                                          int bias16) {
  extern __shared__ uint4 lds_w[];
  for (int t = threadIdx.x; t < packed16; t += blockDim.x) lds_w[t] = packed[t];
  __syncthreads();
  const float *lbias = (const float *)(lds_w + bias16);
  const bf16x8_t *frag0 = (const bf16x8_t *)lds_w;
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
  const int wave = blockIdx.x * kFWaves + (threadIdx.x >> 6);
  const int nwave = gridDim.x * kFWaves;
  const bool vec = (in & 3) == 0;
  for (int base = wave * 32 * kFNT; base < n; base += nwave * 32 * kFNT) {
    bf16x8_t b0[kFNT][KS0];
#pragma unroll
    for (int t = 0; t < kFNT; ++t) {
      const int row = base + 32 * t + col;
      const float *xr = X + (size_t)row * in;
#pragma unroll
      for (int ks = 0; ks < KS0; ++ks) {
        const int k0 = 16 * ks + 8 * h;
        float v[8];
        if (row < n && vec && k0 + 8 <= in) {
          const float4 lo = *(const float4 *)(xr + k0), hi = *(const float4 *)(xr + k0 + 4);
          v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = lo.w;
          v[4] = hi.x; v[5] = hi.y; v[6] = hi.z; v[7] = hi.w;
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = (row < n && k0 + q < in) ? xr[k0 + q] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) b0[t][ks][q] = (__bf16)v[q];
      }
    }
    f32x16_t acc[kFNT][kFH / 32];
    mlp_layer<KS0, kFH / 32>(frag0, b0, acc, lane);
    bf16x8_t b[kFNT][kFH / 16];
    mlp_relu_to_b(acc, lbias, h, b);
    const bf16x8_t *fr = frag0 + KS0 * (kFH / 32) * 64;
    for (int l = 1; l < nhid; ++l) {
      mlp_layer<kFH / 16, kFH / 32>(fr, b, acc, lane);
      fr += (kFH / 16) * (kFH / 32) * 64;
      mlp_relu_to_b(acc, lbias + l * kFH, h, b);
    }
    f32x16_t o[kFNT][1];
    mlp_layer<kFH / 16, 1>(fr, b, o, lane);
    const float *ob = lbias + nhid * kFH;
#pragma unroll
    for (int t = 0; t < kFNT; ++t) {
      const int row = base + 32 * t + col;
      if (row >= n) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 8 * (r >> 2) + 4 * h + (r & 3);
        if (j < out) Y[(size_t)row * out + j] = o[t][0][r] + ob[j];
      }
    }
  }
}

struct Mlp {
  bool init = false;
  std::vector<int> sizes;       // layer widths, input first
  std::vector<size_t> woff, boff;
  size_t nparam = 0;
  float *param = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  long step = 0;
  // work buffers for up to `cap` rows
  int cap = 0;
  std::vector<float *> act;     // act[0] = input rows, act[l] = layer l output
  std::vector<float *> dact;    // gradients w.r.t. act[l]
  float *ybuf = nullptr;
  double *loss = nullptr;
  // fused forward: packed bf16 fragments + biases, rebuilt when the
  // parameters change
  uint4 *packed = nullptr;
  int packed16 = 0, bias16 = 0;
  bool packed_ok = false;
  bool fusable() const {
    const int L = (int)sizes.size();
    if (L < 3 || L > 6 || sizes[0] > 128 || sizes[L - 1] > 32) return false;
    for (int l = 1; l < L - 1; l++)
      if (sizes[l] != kFH) return false;
    return true;
  }
  void release() {
    for (float *p : act) (void)hipFree(p);
    for (float *p : dact) (void)hipFree(p);
    act.clear();
    dact.clear();
    if (ybuf) (void)hipFree(ybuf);
    ybuf = nullptr;
    cap = 0;
  }
  void free_all() {
    release();
    float *ps[4] = {param, grad, adam_m, adam_v};
    for (float *p : ps)
      if (p) (void)hipFree(p);
    param = grad = adam_m = adam_v = nullptr;
    if (loss) (void)hipFree(loss);
    loss = nullptr;
    if (packed) (void)hipFree(packed);
    packed = nullptr;
    packed16 = bias16 = 0;
    packed_ok = false;
    init = false;
  }
};

static Mlp mlp;

static int mlp_reserve(int rows) {
  if (rows <= mlp.cap) return EP_OK;
  mlp.release();
  const int L = (int)mlp.sizes.size();
  mlp.act.assign(L, nullptr);
  mlp.dact.assign(L, nullptr);
  for (int l = 0; l < L; l++) {
    hipError_t e = hipMalloc(&mlp.act[l], (size_t)rows * mlp.sizes[l] * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&mlp.dact[l], (size_t)rows * mlp.sizes[l] * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "MLP work buffers");
  }
  hipError_t e = hipMalloc(&mlp.ybuf, (size_t)rows * mlp.sizes.back() * sizeof(float));
  if (e != hipSuccess) return hip_fail(e, "MLP work buffers");
  mlp.cap = rows;
  return EP_OK;
}

static hipError_t gemm(const float *A, const float *B, float *C, const float *bias, int M, int N, int K, int lda,
                       int ldb, int ldc, int transA, int transB, int relu, int atomic, int splitk,
                       hipStream_t st) {
  GemmArgs g{A, B, C, bias, M, N, K, lda, ldb, ldc, transA, transB, relu, atomic, 0};
  int chunk = (K + splitk - 1) / splitk;
  chunk = (chunk + 7) / 8 * 8;
  g.kchunk = chunk;
  const dim3 grid((M + 31) / 32, (N + 31) / 32, (K + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_gemm, grid, dim3(64), 0, st, g);
  return hipGetLastError();
}

// Packs the parameters for k_mlp_fwd (once per parameter change).
static hipError_t mlp_pack(hipStream_t st) {
  const int L = (int)mlp.sizes.size();
  const int in = mlp.sizes[0], nhid = L - 2;
  const int KS0 = (in + 15) / 16;
  const int frags = KS0 * (kFH / 32) + (nhid - 1) * (kFH / 16) * (kFH / 32) + (kFH / 16);
  const int bias_floats = nhid * kFH + 32;
  const int w16 = frags * 64;  // 16-byte units
  const int total16 = w16 + bias_floats / 4;
  hipError_t e = hipSuccess;
  if (!mlp.packed) {  // sized for this network; ep_mlp_init frees it
    e = hipMalloc(&mlp.packed, (size_t)total16 * 16);
    if (e != hipSuccess) return e;
  }
  __bf16 *dst = (__bf16 *)mlp.packed;
  float *bdst = (float *)(mlp.packed + w16);
  for (int l = 1; l < L; l++) {
    const int lin = mlp.sizes[l - 1], lout = mlp.sizes[l];
    const int KS = (l == 1) ? KS0 : kFH / 16;
    const int OT = (l == L - 1) ? 1 : kFH / 32;
    const int cnt = KS * OT * 64 * 8;
    hipLaunchKernelGGL(k_mlp_pack, dim3((cnt + 255) / 256), dim3(256), 0, st, mlp.param + mlp.woff[l - 1], lin,
                       lout, KS, OT, l == 1 ? 1 : 0, dst);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    dst += cnt;
    e = hipMemcpyAsync(bdst + (l - 1) * kFH, mlp.param + mlp.boff[l - 1], lout * sizeof(float),
                       hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  mlp.packed16 = total16;
  mlp.bias16 = w16;
  mlp.packed_ok = true;
  return hipSuccess;
}

hipError_t mlp_fused_weights(hipStream_t st, const uint4 **packed, int *packed16, int *bias16,
                             std::vector<int> *sizes) {
  if (!mlp.init || !mlp.fusable()) return hipErrorInvalidValue;
  if (!mlp.packed_ok) {
    hipError_t e = mlp_pack(st);
    if (e != hipSuccess) return e;
  }
  *packed = (const uint4 *)mlp.packed;
  *packed16 = mlp.packed16;
  *bias16 = mlp.bias16;
  *sizes = mlp.sizes;
  return hipSuccess;
}

template <int KS0>
static hipError_t fwd_fused_t(const float *X, float *Y, int n, hipStream_t st) {
  const int L = (int)mlp.sizes.size();
  const size_t lds = (size_t)mlp.packed16 * 16;  // <= 99 KB of the 160 KB per CU
  hipError_t e = hipFuncSetAttribute((const void *)k_mlp_fwd<KS0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int rows_per_block = kFWaves * 32 * kFNT;
  const int blocks = std::max(1, std::min(cus, (n + rows_per_block - 1) / rows_per_block));
  hipLaunchKernelGGL(k_mlp_fwd<KS0>, dim3(blocks), dim3(64 * kFWaves), lds, st, X, Y, n, mlp.sizes[0],
                     mlp.sizes[L - 1], L - 2, (const uint4 *)mlp.packed, mlp.packed16, mlp.bias16);
  return hipGetLastError();
}

// Forward through the fused kernel: X [n][in] -> Y [n][out], device memory.
static hipError_t forward_fused(const float *X, float *Y, int n, hipStream_t st) {
  if (!mlp.packed_ok) {
    hipError_t e = mlp_pack(st);
    if (e != hipSuccess) return e;
  }
  switch ((mlp.sizes[0] + 15) / 16) {
    case 1: return fwd_fused_t<1>(X, Y, n, st);
    case 2: return fwd_fused_t<2>(X, Y, n, st);
    case 3: return fwd_fused_t<3>(X, Y, n, st);
    case 4: return fwd_fused_t<4>(X, Y, n, st);
    case 5: return fwd_fused_t<5>(X, Y, n, st);
    case 6: return fwd_fused_t<6>(X, Y, n, st);
    case 7: return fwd_fused_t<7>(X, Y, n, st);
    case 8: return fwd_fused_t<8>(X, Y, n, st);
    default: return hipErrorInvalidValue;
  }
}

// Forward pass over n rows already in act[0].
static hipError_t forward(int n, hipStream_t st) {
  const int L = (int)mlp.sizes.size();
  for (int l = 1; l < L; l++) {
    const int in = mlp.sizes[l - 1], out = mlp.sizes[l];
    hipError_t e = gemm(mlp.act[l - 1], mlp.param + mlp.woff[l - 1], mlp.act[l], mlp.param + mlp.boff[l - 1], n,
                        out, in, in, out, out, 0, 0, (l < L - 1) ? 1 : 0, 0, 1, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Backward pass from the targets in ybuf; accumulates into grad (zeroed
// here) and adds the summed squared error to *loss.
static hipError_t backward(int n, hipStream_t st) {
  const int L = (int)mlp.sizes.size();
  const int m = mlp.sizes.back();
  hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, st, mlp.grad, mlp.nparam);
  const size_t cnt = (size_t)n * m;
  hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, st, mlp.act[L - 1], mlp.ybuf, mlp.dact[L - 1], cnt,
                     2.0f / (float)cnt, mlp.loss);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int l = L - 1; l >= 1; l--) {
    const int in = mlp.sizes[l - 1], out = mlp.sizes[l];
    if (l < L - 1) {
      hipLaunchKernelGGL(k_relu_back, dim3(1024), dim3(256), 0, st, mlp.act[l], mlp.dact[l], (size_t)n * out);
    }
    // dW = act[l-1]^T dZ   (in x out, inner dimension n)
    e = gemm(mlp.act[l - 1], mlp.dact[l], mlp.grad + mlp.woff[l - 1], nullptr, in, out, n, in, out, out, 1, 0, 0,
             1, (n + 511) / 512, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bias_grad, dim3(out), dim3(256), 0, st, mlp.dact[l], mlp.grad + mlp.boff[l - 1], n,
                       out);
    // dX = dZ W^T   (n x in)
    if (l > 1) {
      e = gemm(mlp.dact[l], mlp.param + mlp.woff[l - 1], mlp.dact[l - 1], nullptr, n, in, out, out, out, in, 0, 1,
               0, 0, 1, st);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

}  // namespace ep

using namespace ep;

extern "C" {

int ep_mlp_init(const int32_t *sizes, int32_t nsizes, uint64_t seed) {
  if (!sizes || nsizes < 2 || nsizes > 8) return fail(EP_ERR_INVALID_ARG, "2..8 layer sizes");
  for (int i = 0; i < nsizes; i++)
    if (sizes[i] < 1 || sizes[i] > 4096) return fail(EP_ERR_INVALID_ARG, "layer width outside 1..4096");
  mlp.free_all();
  mlp.sizes.assign(sizes, sizes + nsizes);
  mlp.woff.clear();
  mlp.boff.clear();
  size_t off = 0;
  for (int l = 1; l < nsizes; l++) {
    mlp.woff.push_back(off);
    off += (size_t)sizes[l - 1] * sizes[l];
    mlp.boff.push_back(off);
    off += sizes[l];
  }
  mlp.nparam = off;
  // He-uniform weights from a splitmix64 stream, zero biases
  std::vector<float> h(off, 0.0f);
  uint64_t x = seed;
  auto next = [&]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 1; l < nsizes; l++) {
    const double lim = std::sqrt(6.0 / sizes[l - 1]);
    for (size_t t = 0; t < (size_t)sizes[l - 1] * sizes[l]; t++)
      h[mlp.woff[l - 1] + t] = (float)((2.0 * next() - 1.0) * lim);
  }
  float **bufs[4] = {&mlp.param, &mlp.grad, &mlp.adam_m, &mlp.adam_v};
  for (float **b : bufs) {
    hipError_t e = hipMalloc(b, off * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "MLP parameters");
    e = hipMemset(*b, 0, off * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "MLP parameters");
  }
  hipError_t e = hipMalloc(&mlp.loss, sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(mlp.param, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "MLP init");
  mlp.step = 0;
  mlp.init = true;
  return EP_OK;
}

int ep_mlp_num_params(int64_t *count) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  if (count) *count = (int64_t)mlp.nparam;
  return EP_OK;
}

int ep_mlp_get_params(float *p) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  hipError_t e = hipMemcpy(p, mlp.param, mlp.nparam * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "MLP params download");
  return EP_OK;
}

int ep_mlp_set_params(const float *p) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  hipError_t e = hipMemcpy(mlp.param, p, mlp.nparam * sizeof(float), hipMemcpyHostToDevice);
  mlp.packed_ok = false;
  if (e != hipSuccess) return hip_fail(e, "MLP params upload");
  return EP_OK;
}

int ep_mlp_forward(const float *X, int32_t n, float *Y) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  if (!X || !Y || n < 1) return fail(EP_ERR_INVALID_ARG, "bad rows");
  int rc = mlp_reserve(n);
  if (rc) return rc;
  const int L = (int)mlp.sizes.size();
  hipError_t e = hipMemcpy(mlp.act[0], X, (size_t)n * mlp.sizes[0] * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mlp.fusable() ? forward_fused(mlp.act[0], mlp.act[L - 1], n, 0) : forward(n, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(Y, mlp.act[L - 1], (size_t)n * mlp.sizes[L - 1] * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "MLP forward");
  return EP_OK;
}

// Loss (mean squared error over n rows and all outputs) and its gradient
// with respect to the flat parameter vector.
int ep_mlp_loss_grad(const float *X, const float *Y, int32_t n, double *loss, float *grad) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  if (!X || !Y || n < 1) return fail(EP_ERR_INVALID_ARG, "bad rows");
  int rc = mlp_reserve(n);
  if (rc) return rc;
  const size_t m = mlp.sizes.back();
  hipError_t e = hipMemcpy(mlp.act[0], X, (size_t)n * mlp.sizes[0] * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(mlp.ybuf, Y, (size_t)n * m * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(mlp.loss, 0, sizeof(double));
  if (e == hipSuccess) e = forward(n, 0);
  if (e == hipSuccess) e = backward(n, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  double ls = 0;
  if (e == hipSuccess) e = hipMemcpy(&ls, mlp.loss, sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && grad) e = hipMemcpy(grad, mlp.grad, mlp.nparam * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "MLP loss/gradient");
  if (loss) *loss = ls / ((double)n * m);
  return EP_OK;
}

// Minibatch training with Adam on n rows (X [n][in], Y [n][out] on the
// host, copied to the device once); loss_per_epoch[epochs] receives the
// mean training loss of each epoch.
int ep_mlp_train(const float *X, const float *Y, int32_t n, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float *loss_per_epoch) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  if (!X || !Y || n < 1 || epochs < 1 || batch < 1) return fail(EP_ERR_INVALID_ARG, "bad training arguments");
  const int in = mlp.sizes[0], out = mlp.sizes.back();
  int rc = mlp_reserve(batch);
  if (rc) return rc;
  float *dx = nullptr, *dy = nullptr;
  int32_t *didx = nullptr;
  hipError_t e = hipMalloc(&dx, (size_t)n * in * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dy, (size_t)n * out * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&didx, (size_t)n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemcpy(dx, X, (size_t)n * in * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dy, Y, (size_t)n * out * sizeof(float), hipMemcpyHostToDevice);
  std::vector<int32_t> perm(n);
  uint64_t x = seed ^ 0x5DEECE66Dull;
  auto rnd = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep_i = 0; ep_i < epochs && e == hipSuccess; ep_i++) {
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
    e = hipMemcpy(didx, perm.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(mlp.loss, 0, sizeof(double));
    int batches = 0;
    for (int b0 = 0; b0 + batch <= n && e == hipSuccess; b0 += batch, batches++) {
      hipLaunchKernelGGL(k_gather_rows, dim3(1024), dim3(256), 0, 0, dx, didx + b0, mlp.act[0], batch, in);
      hipLaunchKernelGGL(k_gather_rows, dim3(1024), dim3(256), 0, 0, dy, didx + b0, mlp.ybuf, batch, out);
      e = forward(batch, 0);
      if (e == hipSuccess) e = backward(batch, 0);
      mlp.step++;
      mlp.packed_ok = false;
      const float c1 = 1.0f - powf(b1, (float)mlp.step), c2 = 1.0f - powf(b2, (float)mlp.step);
      hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, 0, mlp.param, mlp.grad, mlp.adam_m, mlp.adam_v,
                         mlp.nparam, lr, b1, b2, c1, c2, eps);
      if (e == hipSuccess) e = hipGetLastError();
    }
    double ls = 0;
    if (e == hipSuccess) e = hipMemcpy(&ls, mlp.loss, sizeof(double), hipMemcpyDeviceToHost);
    if (loss_per_epoch) loss_per_epoch[ep_i] = (float)(ls / ((double)batches * batch * out));
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipFree(didx);
  if (e != hipSuccess) return hip_fail(e, "MLP training");
  return EP_OK;
}

// Forward pass over n rows already resident on the device (zeros), reps
// times; *ms = mean time per pass.  Measures the network alone, without
// host transfers: the fused kernel when the network has its shape, or
// (layerwise = 1) the one-product-per-layer kernels.
int ep_mlp_bench(int32_t n, int32_t reps, int32_t layerwise, double *ms) {
  if (!mlp.init) return fail(EP_ERR_NOT_INIT, "ep_mlp_init first");
  if (n < 1 || reps < 1) return fail(EP_ERR_INVALID_ARG, "bad rows");
  int rc = mlp_reserve(n);
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, 0, mlp.act[0], (size_t)n * mlp.sizes[0]);
  const int L = (int)mlp.sizes.size();
  auto fwd = [&]() {
    return (mlp.fusable() && !layerwise) ? forward_fused(mlp.act[0], mlp.act[L - 1], n, 0) : forward(n, 0);
  };
  hipError_t e = fwd();
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = fwd();
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(e, "MLP bench");
  if (ms) *ms = t / reps;
  return EP_OK;
}

int ep_mlp_finalize(void) {
  mlp.free_all();
  return EP_OK;
}

}  // extern "C"
