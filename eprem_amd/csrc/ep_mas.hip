// ep_mas.hip -- MAS trilinear interpolation of B, V, rho onto the nodes.
//
// One thread per node.  Inside masRadialMax it runs masTriLinear
// (src/masInterp.c:539-741): position to (r [Rs], theta, phi - phiOffset),
// for each of the 7 fields and both time levels a binary search on each of
// the field's own FP32 meshes (masTriLinearBinarySearch, :117-171), an
// 8-corner trilinear stencil in FP64 on FP32 data (masInterpolate,
// :177-200), the (1-s) / s time blend and the unit conversion; outside it
// applies masWind (:1007-1065).  Fields stay resident in HBM as FP32
// (phi, theta, r) arrays with r fastest.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ep_internal.h"
#include "ep_state.h"

namespace ep {

struct MasDev {
  const float *field[2][7];
  const float *mesh[2][21];
  int dims[2][21];
  // uq[lv][3k + a]: the first (level, slot) of axis a whose mesh has the
  // same points (levels and fields usually share meshes): a search is
  // reused while the mesh it ran on stays the same
  int uq[2][21];
  // the mixed-precision form: search fractions, trilinear stencil and time
  // blend in FP32 (the unit conversions and the node state stay FP64)
  int fp32;
};

template <typename R>
__device__ __forceinline__ R mas_bsearch(const float *__restrict__ A, R key, int *imin, int *imax, int lower,
                                         int upper) {
  while (upper != (lower + 1)) {
    const int imid = (lower + upper) / 2;
    if (A[imid] < key)
      lower = imid;
    else
      upper = imid;
  }
  *imin = lower;
  *imax = upper;
  if (A[lower] > key) return (R)0;
  if (A[upper] < key) return (R)1;
  // the denominator is an FP32 difference, as in the reference
  return ((key - (R)A[lower]) / (R)(A[upper] - A[lower]));
}

// The bracketing indices and fraction of the last search on each axis, and
// the mesh it ran on (MasDev::uq).
struct MasSearch {
  int u[3], i0[3], i1[3];
  double fr[3];
};

// masInterpolate (src/masInterp.c:177-200) of field k at level lv, the
// three searches (masTriLinearBinarySearch) reused from `c` when the mesh
// of an axis is the one searched last.
// R: the arithmetic (double; float in the mixed-precision form).
template <typename R>
__device__ __forceinline__ R mas_field(const MasDev &M, int lv, int k, double rr, double th, double ph,
                                       MasSearch &c) {
  const int nr = M.dims[lv][3 * k], nt = M.dims[lv][3 * k + 1];
  const R key[3] = {(R)rr, (R)th, (R)ph};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int u = M.uq[lv][3 * k + a];
    if (u != c.u[a]) {
      c.fr[a] = mas_bsearch<R>(M.mesh[lv][3 * k + a], key[a], &c.i0[a], &c.i1[a], 0, M.dims[lv][3 * k + a] - 1);
      c.u[a] = u;
    }
  }
  const int r0 = c.i0[0], r1 = c.i1[0], t0 = c.i0[1], t1 = c.i1[1], p0 = c.i0[2], p1 = c.i1[2];
  const R rd = (R)c.fr[0], td = (R)c.fr[1], pd = (R)c.fr[2];
  const R one = 1;
  const float *__restrict__ V = M.field[lv][k];
  const size_t snr = nr, snrt = (size_t)nr * nt;
#define VI(r, t, p) ((R)V[(size_t)(r) + snr * (t) + snrt * (p)])
  const R c00 = VI(r0, t0, p0) * (one - rd) + VI(r1, t0, p0) * rd;
  const R c10 = VI(r0, t0, p1) * (one - rd) + VI(r1, t0, p1) * rd;
  const R c01 = VI(r0, t1, p0) * (one - rd) + VI(r1, t1, p0) * rd;
  const R c11 = VI(r0, t1, p1) * (one - rd) + VI(r1, t1, p1) * rd;
#undef VI
  const R c0 = c00 * (one - pd) + c10 * pd;
  const R c1 = c01 * (one - pd) + c11 * pd;
  return (c0 * (one - td) + c1 * td);
}

// The 7 fields at both levels, blended in time: fields in the order of
// MasLevel (Bp, Bt, Br, Vp, Vt, Vr, rho), before the unit conversions.
template <typename R>
__device__ __forceinline__ void mas_blend(const MasDev &M, double rr, double th, double ph, double s,
                                          double out[7]) {
  MasSearch c;
  c.u[0] = c.u[1] = c.u[2] = -1;
  const R sr = (R)s, one = 1;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const R a = mas_field<R>(M, 0, k, rr, th, ph, c);
    const R b = mas_field<R>(M, 1, k, rr, th, ph, c);
    out[k] = (double)((one - sr) * a + sr * b);
  }
}

// MAS unit conversions (src/baseTypes.h:107-115).  The reference's macros
// are not parenthesised, so `x * MAS_B_CONVERT` multiplies and divides x
// by each factor in turn; these helpers keep that order.
__device__ __forceinline__ double mas_b_apply(double x) {
  return x * sqrt(4.0 * kPI * 1.6726e-16) * 6.96e10 / 1445.87003080685 /
         (sqrt(1.6726e-24 * 1.0) * 2.99792458e10);
}
__device__ __forceinline__ double mas_v_apply(double x) {
  return x * (6.96e10 / 1445.87003080685) / 2.99792458e10;
}
__device__ __forceinline__ double mas_rho_apply(double x) { return x * (1.6726e-16 / 1.6726e-24) / 1.0; }

// masGetNode at the position (px, py, pz) [rScale units] for node i:
// masTriLinear inside mas_radial_max, else masWind of node i (its previous
// old position, its rmag and MHD state, whatever position is asked).
// out7 = {Br, Btheta, Bphi, Vr, Vtheta, Vphi, density}.
__device__ void mas_at(const Dev &D, const MasDev &M, double px, double py, double pz, size_t i, double s,
                       double phi_offset, double radial_max, double vmin, const double *prev, size_t SN,
                       double out7[7]) {
  const double rs = D.rScale;
  // cartToSphPosAu (src/geometry.c:131-160, :35-45)
  const double ax = px * rs, ay = py * rs, az = pz * rs;
  const double pr = sqrt(ax * ax + ay * ay + az * az);
  const double pth = acos(az / pr);
  double pph = atan2(ay, ax);
  while (pph < 0.0) pph += (2.0 * kPI);
  if (pph > 2 * kPI)
    while (pph > 2 * kPI) pph -= 2 * kPI;
  if (pr <= (radial_max * kRSAU)) {
    const double rr = pr / kRSAU, th = pth;
    double ph = pph - phi_offset;
    if (ph < 0.0)
      while (ph < 0.0) ph += (2.0 * kPI);
    if (ph > (2.0 * kPI))
      while (ph > (2.0 * kPI)) ph -= (2.0 * kPI);
    double bl[7];
    if (M.fp32)
      mas_blend<float>(M, rr, th, ph, s, bl);
    else
      mas_blend<double>(M, rr, th, ph, s, bl);
    out7[2] = mas_b_apply(bl[0]);
    out7[1] = mas_b_apply(bl[1]);
    out7[0] = mas_b_apply(bl[2]);
    out7[5] = mas_v_apply(bl[3]);
    out7[4] = mas_v_apply(bl[4]);
    double vr = mas_v_apply(bl[5]);
    if (vr < (vmin / 2.99792458e10)) vr = (vmin / 2.99792458e10);
    out7[3] = vr;
    out7[6] = mas_rho_apply(bl[6]);
  } else {
    // masWind: scale the node's previous values by r_old / r
    const double ox = prev[0 * SN + i], oy = prev[1 * SN + i], oz = prev[2 * SN + i];
    const double rOldmag = sqrt(ox * ox + oy * oy + oz * oz);
    const double oneOverR = rOldmag / D.rmag[i];
    out7[0] = prev[3 * SN + i] * oneOverR * oneOverR;
    out7[1] = prev[4 * SN + i] * oneOverR;
    out7[2] = prev[5 * SN + i] * oneOverR;
    out7[3] = prev[6 * SN + i];
    out7[4] = prev[7 * SN + i] * oneOverR;
    out7[5] = prev[8 * SN + i] * oneOverR;
    out7[6] = prev[9 * SN + i] * oneOverR * oneOverR;
  }
}

// masGetNode for node i at its own position.
__device__ __forceinline__ void mas_get_node(const Dev &D, const MasDev &M, size_t i, double s, double phi_offset,
                                             double radial_max, double vmin, const double *prev, size_t SN,
                                             double out7[7]) {
  mas_at(D, M, D.x[i], D.y[i], D.z[i], i, s, phi_offset, radial_max, vmin, prev, SN, out7);
}

// sphToCartVector (src/geometry.c:89-125): (r, theta, phi) components v at
// position (x, y, z) to Cartesian.
__device__ __forceinline__ void sph_to_cart(const double *v, double x, double y, double z, double *out) {
  const double rmag = sqrt(x * x + y * y + z * z);
  if ((rmag - fabs(z)) < 1.0e-33) {
    out[0] = 0.0;
    out[1] = 0.0;
  } else {
    out[0] = x / rmag * v[0] + x * z / (rmag * sqrt(rmag * rmag - z * z)) * v[1] - y / sqrt(rmag * rmag - z * z) * v[2];
    out[1] = y / rmag * v[0] + y * z / (rmag * sqrt(rmag * rmag - z * z)) * v[1] + x / sqrt(rmag * rmag - z * z) * v[2];
  }
  out[2] = z / rmag * v[0] - sqrt(rmag * rmag - z * z) / rmag * v[1];
}

// masMoveNodes (src/masInterp.c:1344-1393): every node but the inner one
// takes the classical RK4 step of rungeKuttaFlow (:1305-1338), whose four
// velocities are vMas (:1279-1296) -- masGetNode at the stage position,
// made Cartesian there.  Stage positions r0, r0 + k1/2, r0 + k2/2, r0 + k3;
// r1 = r0 + (k1 + 2 k2 + 2 k3 + k4) / 6 in the reference's operation
// order.  The node's old position (read by masWind at every stage) is
// replaced by its position before the move only after the four stages.
// state: the MHD buffer ([0..2] old position, [3..9] B, V, density).
__global__ __launch_bounds__(256) void k_move_nodes_mas(Dev D, MasDev M, double *x, double *y, double *z,
                                                        double *rmag, double *state, double tdel, double s,
                                                        double phi_offset, double radial_max, double vmin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int st = blockIdx.y;
  if (k >= D.n_active[st]) return;
  const size_t i = (size_t)st * D.Npad + k;
  const size_t SN = (size_t)D.S * D.Npad;
  const double r0[3] = {x[i], y[i], z[i]};
  if (k > 0) {
    const double sixth = 1.0 / 6.0, two = 2.0, half = 0.5;
    const double h = tdel / D.rScale;
    double kk[4][3];
#pragma unroll
    for (int sg = 0; sg < 4; ++sg) {
      double pos[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) pos[q] = sg == 0 ? r0[q] : (sg < 3 ? r0[q] + kk[sg - 1][q] * half : r0[q] + kk[2][q]);
      double o7[7], vc[3];
      mas_at(D, M, pos[0], pos[1], pos[2], i, s, phi_offset, radial_max, vmin, state, SN, o7);
      sph_to_cart(o7 + 3, pos[0], pos[1], pos[2], vc);
#pragma unroll
      for (int q = 0; q < 3; ++q) kk[sg][q] = vc[q] * h;
    }
    const double nx = r0[0] + sixth * (kk[0][0] + two * kk[1][0] + two * kk[2][0] + kk[3][0]);
    const double ny = r0[1] + sixth * (kk[0][1] + two * kk[1][1] + two * kk[2][1] + kk[3][1]);
    const double nz = r0[2] + sixth * (kk[0][2] + two * kk[1][2] + two * kk[2][2] + kk[3][2]);
    x[i] = nx;
    y[i] = ny;
    z[i] = nz;
    rmag[i] = sqrt((nx * nx) + (ny * ny) + (nz * nz));
  }
  state[i] = r0[0];
  state[SN + i] = r0[1];
  state[2 * SN + i] = r0[2];
}

__global__ __launch_bounds__(256) void k_mas_interp(Dev D, MasDev M, double s, double phi_offset,
                                                    double radial_max, double vmin, const double *prev,
                                                    double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int st = blockIdx.y;
  if (k >= D.n_active[st]) return;
  const size_t i = (size_t)st * D.Npad + k;
  const size_t SN = (size_t)D.S * D.Npad;
  double v[7];
  mas_get_node(D, M, i, s, phi_offset, radial_max, vmin, prev, SN, v);
  for (int q = 0; q < 7; q++) out[q * SN + i] = v[q];
}

// updateMhd (src/flow.c:33-154) for every node: new B, V, density (MAS or
// the analytic wind), |B| and the derivative terms against the previous
// values, zero on the first node of each stream (src/flow.c:131-149).
struct MhdArgs {
  int mode;
  double mhdUs, mhdNsAu, mhdBsAu, omegaSun;
  double s, phi_offset, radial_max, vmin;
};

__global__ __launch_bounds__(256) void k_update_mhd(Dev D, MasDev M, MhdArgs A, double *prev,
                                                    double *out, double *bmag, double *dlnb,
                                                    double *dlnn, double *dupar) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int st = blockIdx.y;
  if (k >= D.n_active[st]) return;
  const size_t i = (size_t)st * D.Npad + k;
  const size_t SN = (size_t)D.S * D.Npad;
  double v[7];
  if (A.mode == 1) {
    mas_get_node(D, M, i, A.s, A.phi_offset, A.radial_max, A.vmin, prev, SN, v);
  } else {
    // mhdDensity, mhdV, mhdB with MHD_DEFAULT (src/flow.c:241-350)
    const double rr = D.rmag[i] * D.rScale;
    const double theta = acos(D.z[i] / D.rmag[i]);
    v[6] = A.mhdNsAu / (rr * rr);
    v[3] = A.mhdUs;
    v[4] = 0.0;
    v[5] = 0.0;
    v[0] = A.mhdBsAu / (rr * rr);
    v[1] = 0.0;
    v[2] = -1.0 * rr * v[0] * (A.omegaSun / (v[3] + 1.0e-33)) * sin(theta);
  }
  for (int q = 0; q < 7; q++) out[q * SN + i] = v[q];
  const double bm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  bmag[i] = bm;
  if (k == 0) {
    dupar[i] = 0.0;
    dlnb[i] = 0.0;
    dlnn[i] = 0.0;
  } else {
    const double obr = prev[3 * SN + i], obt = prev[4 * SN + i], obp = prev[5 * SN + i];
    const double bm_old = sqrt(obr * obr + obt * obt + obp * obp);
    double du = (v[3] - prev[6 * SN + i]) * v[0] + (v[4] - prev[7 * SN + i]) * v[1] +
                (v[5] - prev[8 * SN + i]) * v[2];
    du /= bm;
    dupar[i] = du;
    dlnb[i] = log(bm / bm_old);
    dlnn[i] = log(v[6] / prev[9 * SN + i]);
  }
  // the new values become the node's MHD state (the reference keeps them
  // in Node_t for the next updateMhd and for moveNodes); each thread only
  // touches its own node, after its reads
  for (int q = 0; q < 7; q++) prev[(3 + q) * SN + i] = v[q];
}

}  // namespace ep

using namespace ep;

static int ensure_mas_buffers();

int ensure_mhd_state() { return ensure_mas_buffers(); }

static int ensure_mas_buffers() {
  const size_t SN = (size_t)g.d.num_streams * g.Npad;
  if (!g.mas_out) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, 17 * SN * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "MAS output alloc");
    g.mas_allocs.push_back(p);
    g.mas_out = (double *)p;
    g.mas_prev = g.mas_out + 7 * SN;
    e = hipMemset(g.mas_prev, 0, 10 * SN * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "MAS prev clear");
  }
  return EP_OK;
}

static MasDev mas_dev() {
  MasDev M;
  for (int lv = 0; lv < 2; lv++) {
    for (int k = 0; k < 7; k++) M.field[lv][k] = g.mas[lv].field[k];
    for (int k = 0; k < 21; k++) {
      M.mesh[lv][k] = g.mas[lv].mesh[k];
      M.dims[lv][k] = g.mas[lv].dims[k];
    }
  }
  // canonical mesh of each (level, slot): the first one of the same axis
  // with the same points, in the order (level 0 slots, level 1 slots)
  for (int q = 0; q < 42; q++) {
    const int lv = q / 21, sl = q % 21;
    const std::vector<float> &a = g.mas[lv].hmesh[sl];
    int u = q;
    for (int p = 0; p < q && u == q; p++) {
      const int lp = p / 21, sp = p % 21;
      if (sp % 3 != sl % 3) continue;
      const std::vector<float> &b = g.mas[lp].hmesh[sp];
      if (a.size() == b.size() && memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0) u = p;
    }
    M.uq[lv][sl] = u;
  }
  M.fp32 = g.fp32;
  return M;
}

extern "C" {

int ep_mas_set_level(int32_t level, const float *const fields[7], const float *const meshes[21],
                     const int32_t dims[21]) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (level < 0 || level > 1 || !fields || !meshes || !dims)
    return fail(EP_ERR_INVALID_ARG, "bad MAS level arguments");
  MasLevel &L = g.mas[level];
  for (int k = 0; k < 7; k++) {
    for (int a = 0; a < 3; a++) {
      const int n = dims[3 * k + a];
      if (n < 2 || !meshes[3 * k + a] || !fields[k])
        return fail(EP_ERR_INVALID_ARG, "MAS field %d axis %d: need >= 2 points", k, a);
    }
  }
  // (re)allocate when the shape changes
  for (int k = 0; k < 21; k++) {
    if (L.set && L.dims[k] == dims[k]) continue;
    L.set = false;
  }
  if (!L.set) {
    for (int k = 0; k < 7; k++) {
      const size_t cnt = (size_t)dims[3 * k] * dims[3 * k + 1] * dims[3 * k + 2];
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, cnt * sizeof(float));
      if (e != hipSuccess) return hip_fail(e, "MAS field alloc");
      g.mas_allocs.push_back(p);
      L.field[k] = (float *)p;
    }
    for (int k = 0; k < 21; k++) {
      void *p = nullptr;
      hipError_t e = hipMalloc(&p, dims[k] * sizeof(float));
      if (e != hipSuccess) return hip_fail(e, "MAS mesh alloc");
      g.mas_allocs.push_back(p);
      L.mesh[k] = (float *)p;
      L.dims[k] = dims[k];
    }
  }
  for (int k = 0; k < 7; k++) {
    const size_t cnt = (size_t)dims[3 * k] * dims[3 * k + 1] * dims[3 * k + 2];
    hipError_t e = hipMemcpy(L.field[k], fields[k], cnt * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "MAS field upload");
  }
  for (int k = 0; k < 21; k++) {
    hipError_t e = hipMemcpy(L.mesh[k], meshes[k], dims[k] * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "MAS mesh upload");
    L.hmesh[k].assign(meshes[k], meshes[k] + dims[k]);
  }
  L.set = true;
  return EP_OK;
}

int ep_interp_mhd(const ep_mas_params *mp, const ep_mas_prev *prev, ep_mas_out *out,
                  double *device_ms) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  if (!mp) return fail(EP_ERR_INVALID_ARG, "null MAS params");
  // The phi wrap loops of masTriLinear add or subtract 2 pi until the angle
  // is in [0, 2 pi]; bound the offset so that they end after a few turns.
  if (!(fabs(mp->phi_offset) <= 1.0e3) || !std::isfinite(mp->s_cor) || !std::isfinite(mp->mas_radial_max))
    return fail(EP_ERR_INVALID_ARG, "phi_offset must be finite with |phi_offset| <= 1000");
  if (!g.mas[0].set || !g.mas[1].set) return fail(EP_ERR_INVALID_ARG, "set both MAS levels first");
  const size_t S = g.d.num_streams, N = g.d.num_nodes, SN = S * g.Npad;
  int rc = ensure_mas_buffers();
  if (rc) return rc;
  if (prev) {
    const double *src[10] = {prev->rold_x, prev->rold_y, prev->rold_z, prev->br, prev->btheta,
                             prev->bphi, prev->vr, prev->vtheta, prev->vphi, prev->density};
    for (int i = 0; i < 10; i++) {
      if (!src[i]) return fail(EP_ERR_INVALID_ARG, "prev array %d is null", i);
      hipError_t e = hipMemcpy2D(g.mas_prev + i * SN, g.Npad * 8, src[i], N * 8, N * 8, S,
                                 hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(e, "prev upload");
    }
  }
  MasDev M = mas_dev();
  Dev D = make_dev();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, g.st);
  hipLaunchKernelGGL(k_mas_interp, dim3((N + 255) / 256, S), dim3(256), 0, g.st, D, M, mp->s_cor,
                     mp->phi_offset, mp->mas_radial_max, mp->mas_vmin, g.mas_prev, g.mas_out);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, g.st);
  hipError_t se = hipStreamSynchronize(g.st);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (le != hipSuccess) return hip_fail(le, "k_mas_interp launch");
  if (se != hipSuccess) return hip_fail(se, "k_mas_interp");
  if (device_ms) *device_ms = ms;
  if (out) {
    double *dst[7] = {out->br, out->btheta, out->bphi, out->vr, out->vtheta, out->vphi, out->density};
    for (int i = 0; i < 7; i++) {
      if (!dst[i]) continue;
      hipError_t e = hipMemcpy2D(dst[i], N * 8, g.mas_out + i * SN, g.Npad * 8, N * 8, S,
                                 hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "MAS output download");
    }
  }
  return EP_OK;
}

int ep_set_positions(const double *x, const double *y, const double *z, const double *rmag,
                     const int32_t *n_active) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  const double *src[4] = {x, y, z, rmag};
  double *dst[4] = {g.nx, g.ny, g.nz, g.nrmag};
  const size_t S = g.d.num_streams, N = g.d.num_nodes;
  for (int i = 0; i < 4; i++) {
    if (!src[i]) return fail(EP_ERR_INVALID_ARG, "position array %d is null", i);
    hipError_t e = hipMemcpy2D(dst[i], g.Npad * 8, src[i], N * 8, N * 8, S, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "position upload");
  }
  std::vector<int32_t> na(S, (int32_t)N);
  if (n_active)
    for (size_t s = 0; s < S; s++) {
      if (n_active[s] < 2 || n_active[s] > (int32_t)N)
        return fail(EP_ERR_INVALID_ARG, "n_active[%zu] = %d outside 2..%zu", s, n_active[s], N);
      na[s] = n_active[s];
    }
  hipError_t e = hipMemcpy(g.n_active, na.data(), S * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "n_active");
  g.h_nactive = na;
  g.positions_set = true;
  return EP_OK;
}

int ep_update_mhd(const ep_mhd_params *mp, const ep_mas_prev *prev, ep_mas_out *out, double *device_ms) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.positions_set && !g.nodes_set) return fail(EP_ERR_INVALID_ARG, "set node positions first");
  if (!mp) return fail(EP_ERR_INVALID_ARG, "null MHD params");
  if (!prev && !g.mhd_state)
    return fail(EP_ERR_INVALID_ARG, "no previous MHD state on the device: pass prev");
  if (mp->mode != 0 && mp->mode != 1) return fail(EP_ERR_INVALID_ARG, "mode must be 0 or 1");
  if (mp->mode == 1) {
    if (!g.mas[0].set || !g.mas[1].set) return fail(EP_ERR_INVALID_ARG, "set both MAS levels first");
    if (!(fabs(mp->mas.phi_offset) <= 1.0e3)) return fail(EP_ERR_INVALID_ARG, "phi_offset out of range");
  }
  const size_t S = g.d.num_streams, N = g.d.num_nodes, SN = S * g.Npad;
  int rc = ensure_mas_buffers();
  if (rc) return rc;
  if (prev) {
    const double *src[10] = {prev->rold_x, prev->rold_y, prev->rold_z, prev->br, prev->btheta,
                             prev->bphi, prev->vr, prev->vtheta, prev->vphi, prev->density};
    for (int i = 0; i < 10; i++) {
      if (!src[i]) return fail(EP_ERR_INVALID_ARG, "prev array %d is null", i);
      hipError_t e = hipMemcpy2D(g.mas_prev + i * SN, g.Npad * 8, src[i], N * 8, N * 8, S,
                                 hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(e, "prev upload");
    }
  }
  MasDev M = mas_dev();
  MhdArgs A;
  A.mode = mp->mode;
  A.mhdUs = mp->mhdUs;
  A.mhdNsAu = mp->mhdNsAu;
  A.mhdBsAu = mp->mhdBsAu;
  A.omegaSun = mp->omegaSun;
  A.s = mp->mas.s_cor;
  A.phi_offset = mp->mas.phi_offset;
  A.radial_max = mp->mas.mas_radial_max;
  A.vmin = mp->mas.mas_vmin;
  Dev D = make_dev();
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, g.st);
  hipLaunchKernelGGL(k_update_mhd, dim3((N + 255) / 256, S), dim3(256), 0, g.st, D, M, A, g.mas_prev,
                     g.mas_out, g.nbmag, g.ndlnb, g.ndlnn, g.ndupar);
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, g.st);
  hipError_t se = hipStreamSynchronize(g.st);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (le != hipSuccess) return hip_fail(le, "k_update_mhd launch");
  if (se != hipSuccess) return hip_fail(se, "k_update_mhd");
  if (device_ms) *device_ms = ms;
  g.nodes_set = true;
  g.mhd_state = true;
  // the drift kernels of a captured step take the analytic field by value
  if (g.mhd_mode != mp->mode || g.mhd_bsau != mp->mhdBsAu || g.mhd_omega != mp->omegaSun) g.drop_graph();
  g.mhd_mode = mp->mode;
  g.mhd_bsau = mp->mhdBsAu;
  g.mhd_omega = mp->omegaSun;
  if (out) {
    double *dst[7] = {out->br, out->btheta, out->bphi, out->vr, out->vtheta, out->vphi, out->density};
    for (int i = 0; i < 7; i++) {
      if (!dst[i]) continue;
      hipError_t e = hipMemcpy2D(dst[i], N * 8, g.mas_out + i * SN, g.Npad * 8, N * 8, S,
                                 hipMemcpyDeviceToHost);
      if (e != hipSuccess) return hip_fail(e, "MHD output download");
    }
  }
  return EP_OK;
}

int ep_move_nodes_mas(const ep_mas_params *mp, double tdel) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.mhd_state) return fail(EP_ERR_INVALID_ARG, "ep_update_mhd first (node MHD state)");
  if (!mp) return fail(EP_ERR_INVALID_ARG, "null MAS params");
  if (!(tdel >= 0) || !std::isfinite(tdel)) return fail(EP_ERR_INVALID_ARG, "tdel must be finite and >= 0");
  if (!(fabs(mp->phi_offset) <= 1.0e3) || !std::isfinite(mp->s_cor) || !std::isfinite(mp->mas_radial_max))
    return fail(EP_ERR_INVALID_ARG, "phi_offset must be finite with |phi_offset| <= 1000");
  if (!g.mas[0].set || !g.mas[1].set) return fail(EP_ERR_INVALID_ARG, "set both MAS levels first");
  MasDev M = mas_dev();
  Dev D = make_dev();
  hipLaunchKernelGGL(k_move_nodes_mas, dim3((g.d.num_nodes + 255) / 256, g.d.num_streams), dim3(256), 0, g.st, D,
                     M, g.nx, g.ny, g.nz, g.nrmag, g.mas_prev, tdel, mp->s_cor, mp->phi_offset, mp->mas_radial_max,
                     mp->mas_vmin);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e != hipSuccess) return hip_fail(e, "move nodes (MAS)");
  return EP_OK;
}

}  // extern "C"
