// ep_api.hip -- the C ABI of include/eprem_hip.h (EP update part).
//
// Owns the device state: every buffer is allocated once in ep_init and
// reused; ep_step launches kernels on one HIP stream and synchronises twice
// per global step (after the once-per-step preparation, to size the
// streaming passes, and at the end, to read errors and diagnostics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <vector>

#include "ep_fastmath.h"
#include "ep_internal.h"
#include "ep_launch.h"
#include "ep_state.h"

using namespace ep;

namespace ep {
State g;
static std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(EP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(x, what)              \
  do {                               \
    hipError_t e_ = (x);             \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

template <typename T>
static int dalloc(T **p, size_t count, const char *what) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T) + 64);
  if (e != hipSuccess) return hip_fail(e, what);
  g.allocs.push_back(q);
  *p = (T *)q;
  return EP_OK;
}

static void free_all() {
  g.drop_graph();
  for (void *p : g.allocs) (void)hipFree(p);
  g.allocs.clear();
  for (auto &ev : g.events) (void)hipEventDestroy(ev);
  g.events.clear();
  if (g.st) (void)hipStreamDestroy(g.st);
  g.st = nullptr;
  g.mas_free();
  g.sur = SurState();
  g.init = false;
  g.nodes_set = false;
}

// Constants of the reference (src/baseTypes.h:50-115).
static const double MP = 1.6726e-24, MEV = 1.6022e-6, GEV = 1.6022e-3, C = 2.99792458e10;
static const double MZERO = (MP * C * C / GEV);
static const double MHD_DENSITY_NORM = 1.0;

// initEnergeticParticlesGrids (src/energeticParticlesInit.c:36-100) and
// the energy-only factors of sepSeedFunction
// (src/energeticParticlesBoundary.c:81-99).
static void build_grids(const ep_params &p, int E, int M) {
  const double one = 1.0, two = 2.0, half = 0.5;
  g.egrid.assign(E, 0);
  g.vgrid.assign(E, 0);
  g.vgrid_i.assign(E, 0);
  g.rigidity.assign(E, 0);
  g.pgrid.assign(E, 0);
  g.rig_lamo.assign(E, 0);
  g.plt.assign(E, 0);
  g.ext.assign(E, 0);
  g.mugrid.assign(M, 0);
  double kin = p.eMin * MEV / (MP * C * C);
  const double lnpmin = log(sqrt(kin * kin + two * kin));
  kin = p.eMax * MEV / (MP * C * C);
  const double lnpmax = log(sqrt(kin * kin + two * kin));
  g.dlnp = (lnpmax - lnpmin) / (E - 1.0);
  for (int e = 0; e < E; e++) {
    const double pp = exp(lnpmin + e * g.dlnp);
    g.pgrid[e] = pp;
    g.egrid[e] = sqrt(one + pp * pp) - one;
    g.vgrid[e] = pp / sqrt(one + pp * pp);
    g.vgrid_i[e] = sqrt(one + pp * pp) / pp;
    double rig = (p.mass / p.charge) * pp * MZERO;
    rig = pow(rig, p.rigidityPower);
    g.rigidity[e] = rig;
    g.rig_lamo[e] = rig * p.lamo;
  }
  g.dmu = two / (one * M);
  g.mpos = M;
  for (int m = 0; m < M; m++) {
    const double mup = -one + (m + one) * g.dmu;
    const double mum = -one + m * g.dmu;
    g.mugrid[m] = half * (mup + mum);
    if (g.mugrid[m] >= 0.0 && g.mpos == M) g.mpos = m;
  }
  const double efac = MEV / (MP * C * C);
  const double normE1 = 1.0 * efac;
  const double normE0 = p.boundaryFunctEcutoff * efac;
  for (int e = 0; e < E; e++) {
    g.plt[e] = pow((g.egrid[e] / normE1), -p.boundaryFunctGamma);
    g.ext[e] = exp(-(g.egrid[e] / normE0));
  }
  g.normJ0 = p.boundaryFunctAmplitude * (MP * C) / (MHD_DENSITY_NORM * MEV);
  g.normRadius = p.boundaryFunctR0 / p.rScale;
}

static int check_params(const ep_params &p) {
  if (!(p.rScale > 0) || !(p.tDel > 0) || p.numEpSteps < 1)
    return fail(EP_ERR_INVALID_ARG, "rScale, tDel must be > 0 and numEpSteps >= 1");
  if (p.mfpType != 1 && p.mfpType != 2) return fail(EP_ERR_INVALID_ARG, "mfpType must be 1 or 2");
  if (!(p.eMin > 0) || !(p.eMax >= p.eMin)) return fail(EP_ERR_INVALID_ARG, "bad eMin/eMax");
  if (p.adiabaticFocusAlg < 1 || p.adiabaticFocusAlg > 3 || p.adiabaticChangeAlg < 1 || p.adiabaticChangeAlg > 3)
    return fail(EP_ERR_INVALID_ARG, "adiabaticFocusAlg / adiabaticChangeAlg must be 1, 2 or 3");
  return EP_OK;
}

// What the operator variants need of the dimensions and storage: the
// reference's WENO3 stencils read four cells on each line (they run past
// their arrays below that), and the RK3 kernels work on FP64 storage.
static int check_algs(const ep_params &p, const ep_dims &d, int fp32) {
  if (p.adiabaticFocusAlg == 3 && d.num_mu < 4)
    return fail(EP_ERR_UNSUPPORTED, "WENO3 focusing (adiabaticFocusAlg = 3) needs at least 4 mu cells");
  if (p.adiabaticChangeAlg == 3 && d.num_energies < 4)
    return fail(EP_ERR_UNSUPPORTED, "WENO3 adiabatic change (adiabaticChangeAlg = 3) needs at least 4 energies");
  if (fp32 && (p.adiabaticFocusAlg != 1 || p.adiabaticChangeAlg != 1))
    return fail(EP_ERR_UNSUPPORTED, "the RK3 operators (alg 2, 3) need FP64 storage (precision mode 0 or 2)");
  return EP_OK;
}

// Tables of fm_exp / fm_log: 2^(j/64) and ln F (F = 1 + j/128) as double
// hi + lo pairs, and 1/F, formed in x87 extended precision.  ln F's high
// part is ln F rounded to a multiple of 2^-42 (fm_log adds it to m ln2_hi
// exactly).
static void build_fm_table(double *t) {
  for (int j = 0; j < kExpN; j++) {
    const long double v = exp2l((long double)j / kExpN);
    t[kFmExpHi + j] = (double)v;
    t[kFmExpLo + j] = (double)(v - (long double)t[kFmExpHi + j]);
  }
  const long double grid = ldexpl(1.0L, 42);
  for (int j = 0; j <= kLogN; j++) {
    const long double F = 1.0L + (long double)j / kLogN;
    const long double l = logl(F);
    t[kFmLogInv + j] = (double)(1.0L / F);
    t[kFmLogHi + j] = (double)(roundl(l * grid) / grid);
    t[kFmLogLo + j] = (double)(l - (long double)t[kFmLogHi + j]);
  }
}

static int upload_grids() {
  const int E = g.d.num_energies, M = g.d.num_mu;
  HIPCHK(hipMemcpy(g.egrid_d, g.egrid.data(), E * 8, hipMemcpyHostToDevice), "grid upload");
  HIPCHK(hipMemcpy(g.vgrid_d, g.vgrid.data(), E * 8, hipMemcpyHostToDevice), "grid upload");
  HIPCHK(hipMemcpy(g.vgrid_i_d, g.vgrid_i.data(), E * 8, hipMemcpyHostToDevice), "grid upload");
  HIPCHK(hipMemcpy(g.rig_lamo_d, g.rig_lamo.data(), E * 8, hipMemcpyHostToDevice), "grid upload");
  HIPCHK(hipMemcpy(g.plt_d, g.plt.data(), E * 8, hipMemcpyHostToDevice), "grid upload");
  HIPCHK(hipMemcpy(g.ext_d, g.ext.data(), E * 8, hipMemcpyHostToDevice), "grid upload");
  // mugrid_d holds mugrid, then 0.5 * (1 - mu^2) (the first factor of the
  // focusing velocity, src/energeticParticles.c:1620, formed in its order)
  std::vector<double> mu2(2 * M);
  for (int m = 0; m < M; m++) {
    const double mu = g.mugrid[m];
    mu2[m] = mu;
    mu2[M + m] = 0.5 * (1.0 - mu * mu);
  }
  HIPCHK(hipMemcpy(g.mugrid_d, mu2.data(), 2 * M * 8, hipMemcpyHostToDevice), "grid upload");
  std::vector<double> fmt(kFmTable);
  build_fm_table(fmt.data());
  HIPCHK(hipMemcpy(g.fm_tab_d, fmt.data(), kFmTable * 8, hipMemcpyHostToDevice), "table upload");
  return EP_OK;
}

Dev make_dev() {
  Dev D;
  memset(&D, 0, sizeof(D));
  const ep_params &p = g.p;
  D.S = g.d.num_streams;
  D.N = g.d.num_nodes;
  D.Npad = g.Npad;
  D.E = g.d.num_energies;
  D.M = g.d.num_mu;
  D.rScale = p.rScale;
  D.dt = p.tDel / (1.0 * p.numEpSteps);
  D.dt_full = D.dt * p.numEpSteps;
  D.numEpSteps = p.numEpSteps;
  D.mfpType = p.mfpType;
  D.useBC = (p.useEPBoundary > 0) && (p.useBoundaryFunction > 0);
  D.fp32 = g.fp32;
  D.fast = g.fast;
  D.stream_block = g.stream_block;
  D.stream_halo = g.stream_halo;
  D.stream_r = g.stream_r;
  D.stream_fuse = g.stream_fuse;
  D.stream_whole = g.stream_whole;
  D.focus_alg = p.adiabaticFocusAlg;
  D.change_alg = p.adiabaticChangeAlg;
  D.mfpRadialPower = p.mfpRadialPower;
  D.mfpBPower = p.mfpBPower;
  D.mhdBsAu = p.mhdBsAu;
  D.dsh_min = p.dsh_min;
  D.dampCooling = p.dampCooling;
  D.normJ0 = g.normJ0;
  D.normRadius = g.normRadius;
  D.xi = p.boundaryFunctXi;
  D.beta = p.boundaryFunctBeta;
  D.dlnp = g.dlnp;
  D.dmu = g.dmu;
  D.mpos = g.mpos;
  D.egrid = g.egrid_d;
  D.vgrid = g.vgrid_d;
  D.vgrid_i = g.vgrid_i_d;
  D.rig_lamo = g.rig_lamo_d;
  D.plt = g.plt_d;
  D.ext = g.ext_d;
  D.mugrid = g.mugrid_d;
  D.mu_c1 = g.mugrid_d + g.d.num_mu;
  D.fm_tab = g.fm_tab_d;
  D.fA = g.fA;
  D.fB = g.fB;
  D.loc = g.loc;
  D.fhead = g.fhead;
  D.x = g.nx;
  D.y = g.ny;
  D.z = g.nz;
  D.rmag = g.nrmag;
  D.bmag = g.nbmag;
  D.dlnb = g.ndlnb;
  D.dlnn = g.ndlnn;
  D.dupar = g.ndupar;
  D.n_active = g.n_active;
  D.ds = g.ds;
  D.bplus = g.bplus;
  D.bminus = g.bminus;
  D.mfp_r = g.mfp_r;
  D.radial = g.radial;
  D.list = g.list;
  D.ref = g.ref;
  D.nkept = g.nkept;
  D.thin = g.thin;
  D.dsil = g.dsil;
  D.dsmin = g.dsmin;
  D.nsteps = g.nsteps;
  D.halo = g.halo;
  D.passes = g.passes;
  D.chunks = g.chunks;
  D.dtprop = g.dtprop;
  D.del_fac = g.del_fac;
  D.seed3 = g.seed3;
  D.dsm = g.dsm;
  D.q = g.q;
  D.nsub_f = g.nsub_f;
  D.nsub_c = g.nsub_c;
  D.heavy = g.heavy;
  D.dtsub_f = g.dtsub_f;
  D.dtsub_c = g.dtsub_c;
  D.f_dlnbds = g.f_dlnbds;
  D.f_dupar = g.f_dupar;
  D.f_b = g.f_b;
  D.c_dupar = g.c_dupar;
  D.c_dlnn = g.c_dlnn;
  D.c_dlnb = g.c_dlnb;
  D.c_damp = g.c_damp;
  D.c_bc0 = g.c_bc0;
  D.c_bc1 = g.c_bc1;
  D.diag = g.diag;
  D.fphi = nullptr;
  D.logf = g.logf;
  return D;
}

// Decode an error key (layout in ep_internal.h) into the diag fields.
static int decode_err(unsigned long long key, ep_diag *dg) {
  const int group = (int)((key >> 53) & 0xF);
  const int s = (int)((key >> 39) & 0x3FFF);
  const int op = (int)((key >> 37) & 0x3);
  const int a = (int)((key >> 23) & 0x3FFF);
  const int b = (int)((key >> 15) & 0xFF);
  const int c = (int)((key >> 1) & 0x3FFF);
  const int inf = (int)(key & 1);
  int code;
  dg->err_stream = s;
  if (group == 0 && op == 0) {
    code = (a == kErrDsUnderflow) ? EP_ERR_DS_UNDERFLOW
                                  : (a == kErrStreamShort) ? EP_ERR_STREAM_SHORT : EP_ERR_UNSUPPORTED;
    dg->err_node = dg->err_energy = dg->err_mu = -1;
  } else if (group == 0) {
    code = inf ? EP_ERR_INF : EP_ERR_NAN;
    dg->err_energy = a;
    dg->err_mu = b;
    dg->err_node = c;
  } else if (group == 1) {
    code = inf ? EP_ERR_INF : EP_ERR_NAN;
    dg->err_node = a;
    dg->err_energy = b;
    dg->err_mu = c;
  } else {  // shell diffusion (2) and drift (3): [node][energy | stream][mu], the reference's order
    code = inf ? EP_ERR_INF : EP_ERR_NAN;
    dg->err_node = s;
    dg->err_energy = a >> 7;
    dg->err_stream = ((a & 0x7F) << 7) | b;
    dg->err_mu = c;
  }
  dg->err_code = code;
  return code;
}

static const char *err_text(int code) {
  switch (code) {
    case EP_ERR_NAN: return "NaN in the distribution";
    case EP_ERR_INF: return "Inf in the distribution";
    case EP_ERR_DS_UNDERFLOW: return "FindSegmentLengths(): Underflow in ds";
    case EP_ERR_STREAM_SHORT: return "fewer than 3 kept nodes on a stream";
    case EP_ERR_UNSUPPORTED: return "subcycle count exceeds the supported bound";
    default: return "error";
  }
}

static int ensure_events(int count) {
  while ((int)g.events.size() < count) {
    hipEvent_t ev;
    HIPCHK(hipEventCreate(&ev), "hipEventCreate");
    g.events.push_back(ev);
  }
  return EP_OK;
}

// Streaming schedule in use (every lane-node count has all three).
static int effective_schedule() { return g.stream_sched; }

// The EP substep loop of one global step (src/energeticParticles.c:
// 102-199): streaming passes and bookkeeping, then focusing and adiabatic
// change.  With `ev` (per-operator events, not used under graph capture)
// four events per substep bracket the operators.
static hipError_t enqueue_substeps(const Dev &D, const StepShape &sh, hipStream_t st, hipEvent_t *ev) {
  hipError_t e = hipSuccess;
#define EPQ(x)            \
  do {                    \
    e = (x);              \
    if (e != hipSuccess) return e; \
  } while (0)
  // the streaming passes and the one-pass node kernels carry the upwind
  // operators only
  // learned-surrogate form: the network's applications come first, then
  // the streaming kernels run what is left of each slab (Ds) and the node
  // stage focuses each row over the fraction of the substep it still needs
  // (Dn); the seed floor then runs as its own kernel, before the network
  const bool sur = sh.sur_levels > 0;
  const int fuse = (!sur && !D.logf && sh.do_stream && sh.do_focus && D.stream_fuse && D.focus_alg == 1) ? 1 : 0;
  const bool rk = D.focus_alg != 1 || D.change_alg != 1;
  const Dev Ds = sur ? sur_stream_dev(D) : D;
  const Dev Dn = sur ? sur_node_dev(D) : D;
  const int seed_in_pass = (sh.do_seed && !sur) ? 1 : 0;
  for (int sub = 0; sub < sh.nsub; sub++) {
    hipEvent_t *e4 = ev ? &ev[2 + 4 * sub] : nullptr;
    if (sh.do_seed && (!sh.do_stream || sur)) EPQ(launch_seed_check(D, st));
    if (sh.do_stream) {
      if (sur) EPQ(launch_sur_substep(D, sub, sh.sur_levels - 1, st));
      // focusing of full-resolution streams can ride on each slab's last pass
      if (sh.sched == 0) {
        for (int pass = 0; pass < sh.passes; pass++)
          EPQ(launch_stream_pass(Ds, pass, pass & 1, seed_in_pass && pass == 0, sub, sh.chunks, fuse, st));
      } else if (sh.sched == 1) {
        EPQ(launch_stream_slab(Ds, seed_in_pass, sub, fuse, 0, -1, st));
      } else {
        for (int s = 0; s < D.S; s++) EPQ(launch_stream_slab(Ds, seed_in_pass, sub, fuse, 0, s, st));
      }
      EPQ(launch_stream_finish(Ds, sub, st));
    }
    if (e4) EPQ(hipEventRecord(e4[0], st));
    if (rk) {
      // AdiabaticFocusing then AdiabaticChange, each with its own operator
      if (sh.do_focus)
        EPQ(D.focus_alg == 1 ? launch_focus(Dn, sub, fuse, st) : launch_focus_rk(Dn, D.focus_alg, sub, st));
      if (e4) EPQ(hipEventRecord(e4[1], st));
      if (sh.do_change)
        EPQ(D.change_alg == 1 ? launch_change(Dn, sub, st) : launch_change_rk(Dn, D.change_alg, sub, st));
    } else if (sh.node_fused == 2) {
      // the path is chosen on the device from the preparation's largest
      // change subcycle count: heavy nodes are focused first on both paths,
      // then either every node is focused and changed in one pass over f
      // (node_gate 1) or focused and changed by two kernels (node_gate 2)
      EPQ(launch_focus_heavy(Dn, sub, st));
      if (e4) EPQ(hipEventRecord(e4[1], st));
      Dev D1 = Dn, D2 = Dn;
      D1.node_gate = 1;
      D2.node_gate = 2;
      EPQ(launch_focus_change(D1, sub, fuse, 0, st));
      EPQ(launch_focus(D2, sub, fuse, st, 0, 0));
      EPQ(launch_change(D2, sub, st));
    } else if (sh.node_fused) {
      // heavy nodes are focused first; every node is then focused (unless
      // done already) and changed in one pass over f
      EPQ(launch_focus_heavy(Dn, sub, st));
      if (e4) EPQ(hipEventRecord(e4[1], st));
      EPQ(launch_focus_change(Dn, sub, fuse, 0, st));
    } else {
      if (sh.do_focus) EPQ(launch_focus(Dn, sub, fuse, st));
      if (e4) EPQ(hipEventRecord(e4[1], st));
      if (sh.do_change) EPQ(launch_change(Dn, sub, st));
    }
    // DiffuseShellData, then DriftShellData, after the stream loop
    if (sh.do_shell & 1) EPQ(launch_shell_diffuse(D, sub, st));
    if (sh.do_shell & 2) EPQ(launch_shell_drift(D, sub, st));
    if (e4) EPQ(hipEventRecord(e4[2], st));
  }
#undef EPQ
  return e;
}

}  // namespace ep

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

int ep_init(const ep_dims *dims, const ep_params *params) {
  if (!dims || !params) return fail(EP_ERR_INVALID_ARG, "null dims/params");
  if (g.init) free_all();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(EP_ERR_NO_DEVICE, "no HIP device available");
  const ep_dims d = *dims;
  if (d.num_streams < 1 || d.num_streams > 16383 || d.num_nodes < 2 || d.num_nodes > 16383 ||
      d.num_energies < 2 || d.num_energies > kMaxE || d.num_mu < 2 || d.num_mu > 32)
    return fail(EP_ERR_INVALID_ARG,
                "dims out of range (streams 1..16383, nodes 2..16383, energies 2..64, mu 2..32)");
  int rc = check_params(*params);
  if (rc) return rc;
  if ((rc = check_algs(*params, d, 0))) return rc;
  g.d = d;
  g.p = *params;
  // every f row keeps at least 64 free slots for the spawn head (fidx)
  g.Npad = (d.num_nodes + 64 + 63) / 64 * 64;
  HIPCHK(hipStreamCreate(&g.st), "hipStreamCreate");
  const size_t S = d.num_streams, E = d.num_energies, M = d.num_mu, NP = g.Npad;
  const size_t SN = S * NP, SE = S * E;
  g.init = true;  // so that free_all() releases partial allocations
  g.fp32 = 0;     // f starts as FP64 (filled below)
  g.fast = 0;
#define AL(ptr, cnt) \
  if ((rc = dalloc(&(ptr), (cnt), #ptr))) { free_all(); return rc; }
  AL(g.egrid_d, E) AL(g.vgrid_d, E) AL(g.vgrid_i_d, E) AL(g.rig_lamo_d, E) AL(g.plt_d, E)
  AL(g.ext_d, E) AL(g.mugrid_d, 2 * M) AL(g.fm_tab_d, kFmTable)
  // one extra slab each: room for the FP64 scratch of the mixed-precision form
  AL(g.fA, SE * M * NP + M * NP) AL(g.fB, SE * M * NP + M * NP) AL(g.loc, SE) AL(g.fhead, 1)
  AL(g.nx, SN) AL(g.ny, SN) AL(g.nz, SN) AL(g.nrmag, SN) AL(g.nbmag, SN) AL(g.ndlnb, SN)
  AL(g.ndlnn, SN) AL(g.ndupar, SN) AL(g.n_active, S)
  AL(g.ds, SN) AL(g.bplus, SN) AL(g.bminus, SN) AL(g.mfp_r, SN) AL(g.radial, SN)
  AL(g.list, SN) AL(g.ref, SN) AL(g.nkept, S) AL(g.thin, S) AL(g.dsil, SN) AL(g.dsmin, S)
  AL(g.nsteps, SE) AL(g.halo, SE) AL(g.passes, SE) AL(g.chunks, SE) AL(g.dtprop, SE)
  AL(g.del_fac, SE * M) AL(g.seed3, SE * 3) AL(g.dsm, SE * NP) AL(g.q, SE * NP)
  AL(g.nsub_f, SN) AL(g.nsub_c, SN) AL(g.heavy, SN) AL(g.dtsub_f, SN) AL(g.dtsub_c, SN) AL(g.f_dlnbds, SN)
  AL(g.f_dupar, SN) AL(g.f_b, SN) AL(g.c_dupar, SN) AL(g.c_dlnn, SN) AL(g.c_dlnb, SN)
  AL(g.c_damp, SN) AL(g.c_bc0, SN) AL(g.c_bc1, SN) AL(g.diag, 1)
#undef AL
  build_grids(g.p, (int)E, (int)M);
  if ((rc = upload_grids())) { free_all(); return rc; }
  // every f starts at DBL_MIN (initEnergeticParticlesGrids), in buffer A
  HIPCHK(hipMemset(g.loc, 0, SE), "hipMemset");
  g.fhead_h = (int)(NP - d.num_nodes);  // the rows start at their far end
  HIPCHK(hipMemcpy(g.fhead, &g.fhead_h, 4, hipMemcpyHostToDevice), "f head");
  HIPCHK(launch_fill(g.fA, SE * M * NP, kDblMin, g.st), "fill");
  HIPCHK(launch_fill(g.fB, SE * M * NP, kDblMin, g.st), "fill");
  std::vector<int32_t> na(S, d.num_nodes);
  HIPCHK(hipMemcpy(g.n_active, na.data(), S * 4, hipMemcpyHostToDevice), "n_active");
  g.h_nactive = na;
  HIPCHK(hipStreamSynchronize(g.st), "init sync");
  double *scr = nullptr;  // the fixed scratch's address on this device
  if ((rc = scratch_address(&scr))) { free_all(); return rc; }
  return EP_OK;
}

int ep_set_params(const ep_params *params) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!params) return fail(EP_ERR_INVALID_ARG, "null params");
  int rc = check_params(*params);
  if (rc) return rc;
  if ((rc = check_algs(*params, g.d, g.fp32))) return rc;
  g.p = *params;
  g.drop_graph();
  build_grids(g.p, g.d.num_energies, g.d.num_mu);
  return upload_grids();
}

// With FP32 storage, one node per lane and halos below this are switched
// off: steps at 8 subcycles per pass on streams of ~1.5e5 subcycles ran
// past their time limit there (one and three nodes per lane; the same
// streams at halo 12, two nodes per lane, and in FP64 at halo 8 finish),
// and the cause is not found.
constexpr int kFp32MinHalo = 12;

// A streaming tile is stream_block kept nodes (the chunk, halo included),
// carried by stream_block / lane_nodes lanes.
static bool tiling_ok(int W, int H, int R) {
  if (R < 1 || R > 3) return false;
  if (W % (64 * R) || W < 64 * R || W / R > 512) return false;
  return H >= 1 && 2 * H < W;
}

int ep_set_tuning(int32_t stream_block, int32_t stream_halo, int32_t lane_nodes, int32_t fuse_focus) {
  if (!tiling_ok(stream_block, stream_halo, lane_nodes) || (fuse_focus != 0 && fuse_focus != 1))
    return fail(EP_ERR_INVALID_ARG,
                "stream tiling: lane_nodes 1, 2 or 3, stream_block a multiple of 64*lane_nodes with at most "
                "512 lanes, 1 <= stream_halo < stream_block/2, fuse_focus 0 or 1");
  if (g.logf && (lane_nodes != 2 || fuse_focus != 0))
    return fail(EP_ERR_UNSUPPORTED, "the log-scaled storage (precision mode 3) streams two nodes per lane, "
                                    "focusing in its own kernel (lane_nodes 2, fuse_focus 0)");
  if (g.fp32 && (lane_nodes == 1 || stream_halo < kFp32MinHalo))
    return fail(EP_ERR_UNSUPPORTED, "with FP32 storage (precision modes 1 and 3) one node per lane and halos "
                                    "below %d are switched off: use lane_nodes 2 or 3 and stream_halo >= %d",
                kFp32MinHalo, kFp32MinHalo);
  g.stream_block = stream_block;
  g.stream_halo = stream_halo;
  g.stream_r = lane_nodes;
  g.stream_fuse = fuse_focus;
  g.drop_graph();
  return EP_OK;
}

// Tile-size search for the streaming kernel.  For the current nodes it
// times, per candidate (block, halo), every streaming pass of one substep
// with flip = 0 (each pass reads the current slabs and writes the spare
// buffer, so f and the step state are left as they were), keeps the
// fastest and leaves it set.  Candidates: one node per lane with 128..512
// lanes, two nodes per lane with 128..512 lanes (chunks of 256..1024
// nodes); halos keep at least half of a chunk in its interior.
static int autotune_impl(int32_t *best_block, int32_t *best_halo, int32_t *best_lane_nodes,
                         int32_t *best_fuse, double *best_ms);

int ep_autotune(int32_t *best_block, int32_t *best_halo, int32_t *best_lane_nodes, int32_t *best_fuse,
                double *best_ms) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  // the search changes the tiling while it times candidates: on an error
  // the tiling it found is restored and the task tables rebuilt for it, and
  // any captured step (taken with the tiling as a kernel argument) dropped
  const int W0 = g.stream_block, H0 = g.stream_halo, R0 = g.stream_r, F0 = g.stream_fuse, S0 = g.stream_sched;
  g.drop_graph();
  const int rc = autotune_impl(best_block, best_halo, best_lane_nodes, best_fuse, best_ms);
  if (rc != EP_OK) {
    g.stream_block = W0;
    g.stream_halo = H0;
    g.stream_r = R0;
    g.stream_fuse = F0;
    g.stream_sched = S0;
    g.drop_graph();
    Dev D = make_dev();
    if (launch_task_prep(D, g.st) == hipSuccess) (void)hipStreamSynchronize(g.st);
  }
  return rc;
}

static int autotune_impl(int32_t *best_block, int32_t *best_halo, int32_t *best_lane_nodes,
                         int32_t *best_fuse, double *best_ms) {
  int rc = ensure_events(2);
  if (rc) return rc;
  hipStream_t st = g.st;
  Dev D = make_dev();
  DevDiag hd;
  HIPCHK(launch_reset_diag(g.diag, st), "reset_diag");
  HIPCHK(launch_stream_prep(D, st), "stream_prep");
  HIPCHK(launch_node_prep(D, st), "node_prep");
  HIPCHK(hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, st), "diag copy");
  HIPCHK(hipStreamSynchronize(st), "prep sync");
  if (hd.err_key != ~0ull) {
    ep_diag dg;
    memset(&dg, 0, sizeof(dg));
    int code = decode_err(hd.err_key, &dg);
    return fail(code, "%s (stream %d)", err_text(code), dg.err_stream);
  }
  // the node stage as ep_step would run it: focusing alone, or focusing and
  // change in one pass
  const int tune_node_fused = (g.node_fuse && !g.logf && hd.max_sub_change <= 1) ? 1 : 0;
  struct Cand {
    int W, H, R, F, S;  // S: streaming schedule
    double ms;
  };
  // time the passes of one substep and the node stage after them (the
  // focusing they leave to do, with the change when the two run as one
  // kernel; all write only the spare buffer), best of two
  auto time_cand = [&](Cand &c) -> int {
    g.stream_block = c.W;
    g.stream_halo = c.H;
    g.stream_r = c.R;
    g.stream_fuse = c.F;
    D = make_dev();
    HIPCHK(launch_reset_diag(g.diag, st), "reset_diag");
    HIPCHK(launch_task_prep(D, st), "task_prep");
    HIPCHK(hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, st), "diag copy");
    HIPCHK(hipStreamSynchronize(st), "tune sync");
    c.ms = 1e300;
    // best of two, or of six when a candidate takes under 2 ms (small
    // workloads, where candidates differ by less than the run-to-run spread
    // of two)
    for (int rep = 0; rep < 2 || (rep < 6 && c.ms < 2.0); rep++) {
      if (c.S) HIPCHK(launch_reset_queue(D, st), "reset_queue");
      HIPCHK(hipEventRecord(g.events[0], st), "event");
      if (c.S)
        HIPCHK(launch_stream_slab(D, 0, 0, c.F, 1, -1, st), "stream_slab");
      else
        for (int pass = 0; pass < hd.max_passes; pass++)
          HIPCHK(launch_stream_pass(D, pass, 0, 0, 0, hd.max_chunks, c.F, st), "stream_pass");
      if (tune_node_fused)
        HIPCHK(launch_focus_change(D, 0, c.F, 1, st), "focus_change");
      else
        HIPCHK(launch_focus(D, 0, c.F, st, 1), "focus");
      HIPCHK(hipEventRecord(g.events[1], st), "event");
      HIPCHK(hipEventSynchronize(g.events[1]), "tune sync");
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, g.events[0], g.events[1]), "elapsed");
      if (ms < c.ms) c.ms = ms;
    }
    return EP_OK;
  };
  // phase 1: every tiling with the separate focusing kernel; phase 2: the
  // fused form of the best few (its extra registers can change the order)
  static const int lanes[] = {128, 192, 256, 320, 384, 448, 512};
  static const int halos[] = {4, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256};
  // phase 1 times one node per lane with a launch per pass and two or
  // three with the persistent slab launch; phase 2 the best few with
  // focusing in the last pass and each with the other schedule
  // (the log-scaled storage has the slab kernel with two nodes per lane
  // only, without fused focusing)
  std::vector<Cand> cands;
  for (int R = 1; R <= 3; R++)
    for (int Lw : lanes)
      for (int H : halos) {
        if (g.logf && R != 2) continue;
        if (g.fp32 && (R == 1 || H < kFp32MinHalo)) continue;  // switched off with FP32 storage
        Cand c{Lw * R, H, R, 0, R > 1 ? 1 : 0, 0.0};
        if (4 * H > c.W) continue;
        if ((rc = time_cand(c))) return rc;
        cands.push_back(c);
      }
  std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
  // phase 2 takes the two best of each lane-node count: timings of the
  // forms without fused focusing can sit within noise of each other (a
  // stream that fits one chunk runs every subcycle in one pass whatever
  // the halo), and fusing can reorder them
  std::vector<Cand> top;
  int per_r[4] = {0, 0, 0, 0};
  for (const Cand &c : cands)
    if (per_r[c.R] < 2) {
      ++per_r[c.R];
      top.push_back(c);
    }
  const size_t nbest = g.logf ? 0 : top.size();
  for (size_t k = 0; k < nbest; k++) {
    Cand c = top[k];
    c.F = 1;
    if ((rc = time_cand(c))) return rc;
    cands.push_back(c);
    // the other schedule, without and with fused focusing
    for (int F = 0; F <= 1; F++) {
      Cand o = top[k];
      o.S = 1 - o.S;
      o.F = F;
      if ((rc = time_cand(o))) return rc;
      cands.push_back(o);
    }
  }
  // phase 3: the four fastest so far timed again (their own best of two
  // or six more), so that candidates within the run-to-run spread of each
  // other are ranked on more than one sample each
  std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
  const size_t nfinal = std::min<size_t>(cands.size(), 4);
  for (size_t k = 0; k < nfinal; k++) {
    Cand c = cands[k];
    if ((rc = time_cand(c))) return rc;
    cands[k].ms = std::min(cands[k].ms, c.ms);
  }
  const Cand best_c = *std::min_element(cands.begin(), cands.begin() + (nfinal ? nfinal : cands.size()),
                                        [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
  const int bb = best_c.W, bh = best_c.H, br = best_c.R, bf = best_c.F;
  const double best = best_c.ms;
  g.stream_block = bb;
  g.stream_halo = bh;
  g.stream_r = br;
  g.stream_fuse = bf;
  g.stream_sched = best_c.S;
  g.drop_graph();
  // leave the task tables consistent with the chosen tiling
  D = make_dev();
  HIPCHK(launch_task_prep(D, st), "task_prep");
  HIPCHK(launch_reset_diag(g.diag, st), "reset_diag");
  HIPCHK(hipStreamSynchronize(st), "tune sync");
  if (best_block) *best_block = bb;
  if (best_halo) *best_halo = bh;
  if (best_lane_nodes) *best_lane_nodes = br;
  if (best_fuse) *best_fuse = bf;
  if (best_ms) *best_ms = best;
  return EP_OK;
}

// Storage and arithmetic of the hot path: 0 = FP64 in the reference's
// operation order (the parity path), 1 = f stored as float, streaming and
// focusing in FP32, relaxation and adiabatic change (the ln f energy flux)
// in FP64, 2 = FP64 with fused multiply-adds in the streaming kernel,
// 3 = mode 1 with the float storage holding ln f (the log-scaled variant).
// The distribution on the device is converted when the storage changes.
int ep_set_precision(int32_t mode) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (mode < 0 || mode > 3)
    return fail(EP_ERR_INVALID_ARG,
                "precision mode must be 0 (FP64), 1 (mixed), 2 (FP64, fused) or 3 (mixed, log-scaled f)");
  const int fp32 = (mode == 1 || mode == 3) ? 1 : 0;
  const int logf = (mode == 3) ? 1 : 0;
  int rc = check_algs(g.p, g.d, fp32);
  if (rc) return rc;
  if (fp32 != g.fp32 || logf != g.logf) {
    // through FP64 storage: decode with the old codec, encode with the new
    Dev D = make_dev();
    HIPCHK(launch_normalize(D, g.st), "normalize");
    if (g.fp32) HIPCHK(launch_convert(D, 0, g.logf, g.st), "convert");
    if (fp32) {
      D.fp32 = 0;
      HIPCHK(launch_convert(D, 1, logf, g.st), "convert");
    }
    HIPCHK(hipStreamSynchronize(g.st), "convert sync");
  }
  g.fp32 = fp32;
  g.logf = logf;
  g.fast = (mode != 0) ? 1 : 0;
  if (fp32 && !logf && g.stream_r == 1) {
    // one node per lane is switched off with FP32 storage: the same lane
    // count with two nodes each
    g.stream_block *= 2;
    g.stream_r = 2;
  }
  if (fp32 && g.stream_halo < kFp32MinHalo) g.stream_halo = kFp32MinHalo;
  if (logf) {
    // the log-scaled storage's streaming kernel: the slab launch with two
    // nodes per lane (same lane count), focusing in its own kernel
    if (g.stream_r != 2) {
      const int lanes = g.stream_block / g.stream_r;
      g.stream_block = 2 * lanes;
      g.stream_r = 2;
      if (2 * g.stream_halo >= g.stream_block) g.stream_halo = g.stream_block / 4;
    }
    g.stream_fuse = 0;
    if (g.stream_sched == 0) g.stream_sched = 1;
  }
  g.drop_graph();
  return EP_OK;
}

int ep_set_schedule(int32_t schedule) {
  if (schedule < 0 || schedule > 2)
    return fail(EP_ERR_INVALID_ARG, "schedule must be 0 (a launch per pass), 1 (persistent) or 2 (per stream)");
  if (g.logf && schedule == 0)
    return fail(EP_ERR_UNSUPPORTED, "the log-scaled storage (precision mode 3) streams with the slab launch");
  g.stream_sched = schedule;
  g.drop_graph();
  return EP_OK;
}

int ep_get_schedule(int32_t *schedule) {
  if (!schedule) return fail(EP_ERR_INVALID_ARG, "null schedule");
  *schedule = effective_schedule();
  return EP_OK;
}

int ep_set_graph(int32_t enable) {
  g.use_graph = enable != 0;
  g.drop_graph();
  return EP_OK;
}

int ep_set_node_fusion(int32_t fused) {
  if (fused != 0 && fused != 1) return fail(EP_ERR_INVALID_ARG, "node fusion must be 0 or 1");
  g.node_fuse = fused;
  g.drop_graph();
  return EP_OK;
}

int ep_get_grids(double *egrid, double *vgrid, double *mugrid, double *rigidity, double *dlnp,
                 double *dmu) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  const int E = g.d.num_energies, M = g.d.num_mu;
  if (egrid) memcpy(egrid, g.egrid.data(), E * 8);
  if (vgrid) memcpy(vgrid, g.vgrid.data(), E * 8);
  if (mugrid) memcpy(mugrid, g.mugrid.data(), M * 8);
  if (rigidity) memcpy(rigidity, g.rigidity.data(), E * 8);
  if (dlnp) *dlnp = g.dlnp;
  if (dmu) *dmu = g.dmu;
  return EP_OK;
}

int ep_set_nodes(const ep_node_soa *nodes, const int32_t *n_active) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!nodes) return fail(EP_ERR_INVALID_ARG, "null nodes");
  const double *src[8] = {nodes->x, nodes->y, nodes->z, nodes->rmag,
                          nodes->bmag, nodes->dlnb, nodes->dlnn, nodes->dupar};
  double *dst[8] = {g.nx, g.ny, g.nz, g.nrmag, g.nbmag, g.ndlnb, g.ndlnn, g.ndupar};
  const size_t S = g.d.num_streams, N = g.d.num_nodes;
  for (int i = 0; i < 8; i++) {
    if (!src[i]) return fail(EP_ERR_INVALID_ARG, "node array %d is null", i);
    HIPCHK(hipMemcpy2D(dst[i], g.Npad * 8, src[i], N * 8, N * 8, S, hipMemcpyHostToDevice),
           "node upload");
  }
  std::vector<int32_t> na(S, (int32_t)N);
  if (n_active) {
    for (size_t s = 0; s < S; s++) {
      if (n_active[s] < 2 || n_active[s] > (int32_t)N)
        return fail(EP_ERR_INVALID_ARG, "n_active[%zu] = %d outside 2..%zu", s, n_active[s], N);
      na[s] = n_active[s];
    }
  }
  HIPCHK(hipMemcpy(g.n_active, na.data(), S * 4, hipMemcpyHostToDevice), "n_active");
  g.h_nactive = na;
  g.nodes_set = true;
  return EP_OK;
}

// Streams s0 .. s0+count-1 of f between the host array (reference layout,
// those streams only) and the device.  Every slab is first brought back
// to buffer A, so that buffer B can stage the host data.
static int transfer_f(int32_t s0, int32_t count, double *hf, int upload) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!hf) return fail(EP_ERR_INVALID_ARG, "null f");
  if (s0 < 0 || count < 1 || s0 + count > g.d.num_streams)
    return fail(EP_ERR_INVALID_ARG, "streams %d..%d outside 0..%d", s0, s0 + count - 1, g.d.num_streams - 1);
  const size_t cnt = (size_t)count * g.d.num_nodes * g.d.num_energies * g.d.num_mu;
  Dev D = make_dev();
  HIPCHK(launch_normalize(D, g.st), "normalize");
  if (upload) {
    HIPCHK(hipStreamSynchronize(g.st), "normalize sync");
    HIPCHK(hipMemcpy(g.fB, hf, cnt * 8, hipMemcpyHostToDevice), "f upload");
    HIPCHK(launch_from_reference(D, g.fB, g.fA, g.st, s0, count), "from_reference");
    HIPCHK(hipStreamSynchronize(g.st), "upload sync");
  } else {
    HIPCHK(launch_to_reference(D, g.fA, g.fB, g.st, s0, count), "to_reference");
    HIPCHK(hipStreamSynchronize(g.st), "download sync");
    HIPCHK(hipMemcpy(hf, g.fB, cnt * 8, hipMemcpyDeviceToHost), "f download");
  }
  return EP_OK;
}

int ep_upload_f(const double *f, int32_t layout) {
  if (layout != EP_LAYOUT_REFERENCE) return fail(EP_ERR_INVALID_ARG, "unknown layout %d", layout);
  return transfer_f(0, g.init ? g.d.num_streams : 1, const_cast<double *>(f), 1);
}

int ep_download_f(double *f, int32_t layout) {
  if (layout != EP_LAYOUT_REFERENCE) return fail(EP_ERR_INVALID_ARG, "unknown layout %d", layout);
  return transfer_f(0, g.init ? g.d.num_streams : 1, f, 0);
}

int ep_upload_f_streams(int32_t stream0, int32_t count, const double *f, int32_t layout) {
  if (layout != EP_LAYOUT_REFERENCE) return fail(EP_ERR_INVALID_ARG, "unknown layout %d", layout);
  return transfer_f(stream0, count, const_cast<double *>(f), 1);
}

int ep_download_f_streams(int32_t stream0, int32_t count, double *f, int32_t layout) {
  if (layout != EP_LAYOUT_REFERENCE) return fail(EP_ERR_INVALID_ARG, "unknown layout %d", layout);
  return transfer_f(stream0, count, f, 0);
}

int ep_init_seed_population(void) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  Dev D = make_dev();
  HIPCHK(launch_init_seed(D, g.st), "init_seed");
  HIPCHK(hipStreamSynchronize(g.st), "init_seed sync");
  return EP_OK;
}

// Everything a step launches after the preparation, when the launches do
// not depend on what the preparation found (the persistent streaming
// schedule): the preparation, then the substeps with device-side gating.
static hipError_t enqueue_step(const Dev &D, const StepShape &sh, hipStream_t st, hipEvent_t *ev) {
  hipError_t e = launch_reset_diag(D.diag, st);
  if (e == hipSuccess) e = launch_stream_prep(D, st);
  if (e == hipSuccess) e = launch_node_prep(D, st);
  if (e == hipSuccess && sh.do_shell) e = launch_shell_links(D, st);
  if (e == hipSuccess && ev) e = hipEventRecord(ev[1], st);
  if (e != hipSuccess) return e;
  Dev Dg = D;
  Dg.gated = 1;
  return enqueue_substeps(Dg, sh, st, ev);
}

int ep_step(uint32_t op_flags, ep_diag *out) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!g.nodes_set) return fail(EP_ERR_INVALID_ARG, "ep_set_nodes first");
  const ep_params &p = g.p;
  const int do_stream = (p.useParallelDiffusion > 0) && (op_flags & EP_OP_STREAM);
  const int do_focus = (p.useAdiabaticFocus > 0) && (op_flags & EP_OP_FOCUS);
  const int do_change = (p.useAdiabaticChange > 0) && (op_flags & EP_OP_CHANGE);
  const int do_seed = (p.checkSeedPopulation > 0) && (op_flags & EP_OP_SEEDCHECK);
  const int nsub = p.numEpSteps;
  int rc = ensure_events(2 + 4 * nsub + 1);
  if (rc) return rc;
  Dev D = make_dev();
  hipStream_t st = g.st;
  std::vector<hipEvent_t> &ev = g.events;
  const int do_shell = ((g.kperxkpar > 0.0) && (op_flags & EP_OP_SHELL)) ? 1 : 0;
  const int do_drift = (g.shell_drift && (op_flags & EP_OP_DRIFT)) ? 1 : 0;
  if (do_shell || do_drift) {
    if (!g.mhd_state) return fail(EP_ERR_INVALID_ARG, "shell operators need the device MHD state (ep_update_mhd)");
    if (g.logf) return fail(EP_ERR_UNSUPPORTED, "shell operators do not take the log-scaled storage (precision mode 3)");
    if (do_drift && g.mhd_mode != 0)
      return fail(EP_ERR_UNSUPPORTED, "drift needs the analytic-wind field (ep_update_mhd mode 0)");
    for (int32_t n : g.h_nactive)
      if (n != g.d.num_nodes) return fail(EP_ERR_UNSUPPORTED, "shell operators need every stream full");
  }
  if ((rc = sur_check_step())) return rc;
  const bool sur = g.sur.K > 0 && do_stream;
  if (sur && (do_shell || do_drift))
    return fail(EP_ERR_UNSUPPORTED, "surrogate: shell diffusion and drift are not supported with it");
  const int sched = effective_schedule();
  if (g.logf && do_stream && (sched == 0 || g.stream_r != 2))
    return fail(EP_ERR_UNSUPPORTED, "the log-scaled storage (precision mode 3) streams with the slab launch "
                                    "(schedule 1 or 2) and two nodes per lane");
  // With the persistent streaming launch nothing the host launches depends
  // on the preparation's results: the whole global step (preparation and
  // substeps) is one sequence, captured as one graph, and the host waits
  // once, at the end.  The per-pass schedule sizes its launches from the
  // preparation and waits for it first.
  // The surrogate form sizes its network launches from the preparation's
  // largest subcycle count and takes the per-pass path.
  const bool one_sequence = sched != 0 && !sur;
  DevDiag hd;
  ep_diag dg;
  memset(&dg, 0, sizeof(dg));
  dg.err_stream = dg.err_node = dg.err_energy = dg.err_mu = -1;
  HIPCHK(hipEventRecord(ev[0], st), "event");
  int node_fused;
  StepShape shape;
  if (one_sequence) {
    node_fused = (g.node_fuse && !g.logf && do_focus && do_change && p.adiabaticFocusAlg == 1 &&
                  p.adiabaticChangeAlg == 1) ? 2 : 0;
    shape = StepShape{0, 0, nsub, do_stream, do_focus, do_change, do_seed, do_shell | (do_drift << 1), g.generation,
                      node_fused, sched};
  } else {
    HIPCHK(launch_reset_diag(g.diag, st), "reset_diag");
    HIPCHK(launch_stream_prep(D, st), "stream_prep");
    HIPCHK(launch_node_prep(D, st), "node_prep");
    HIPCHK(hipEventRecord(ev[1], st), "event");
    HIPCHK(hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, st), "diag copy");
    HIPCHK(hipStreamSynchronize(st), "prep sync");
    if (hd.err_key != ~0ull) {
      int code = decode_err(hd.err_key, &dg);
      if (out) *out = dg;
      return fail(code, "%s (stream %d)", err_text(code), dg.err_stream);
    }
    if (do_shell || do_drift) HIPCHK(launch_shell_links(D, st), "shell links");
    if (sur) HIPCHK(launch_sur_prep(D, st), "surrogate scalars");
    // the surrogate form focuses each row over its own fraction of the
    // substep, which the two-kernel node stage (k_focus) carries
    node_fused = (g.node_fuse && !g.logf && do_focus && do_change && hd.max_sub_change <= 1 &&
                  p.adiabaticFocusAlg == 1 && p.adiabaticChangeAlg == 1 && !sur) ? 1 : 0;
    shape = StepShape{hd.max_passes, hd.max_chunks, nsub, do_stream, do_focus, do_change, do_seed,
                      do_shell | (do_drift << 1), g.generation, node_fused, sched,
                      sur ? 1 + hd.max_nsteps / g.sur.K : 0};
  }
  if (g.use_graph) {
    // One HIP graph, rebuilt only when its shape (operators, schedule, pass
    // count of the per-pass schedule) or the parameters change.
    if (!g.graph_exec || !(g.graph_shape == shape)) {
      if (g.graph_exec) (void)hipGraphExecDestroy(g.graph_exec);
      g.graph_exec = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
      hipError_t ce = one_sequence ? enqueue_step(D, shape, st, nullptr) : enqueue_substeps(D, shape, st, nullptr);
      hipGraph_t graph = nullptr;
      hipError_t ee = hipStreamEndCapture(st, &graph);
      if (ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(ce, "step capture");
      }
      HIPCHK(ee, "end capture");
      hipError_t ie = hipGraphInstantiate(&g.graph_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(ie, "graph instantiate");
      g.graph_shape = shape;
    }
    HIPCHK(hipGraphLaunch(g.graph_exec, st), "graph launch");
    HIPCHK(hipEventRecord(ev[2 + 4 * (nsub - 1) + 2], st), "event");
  } else if (one_sequence) {
    HIPCHK(enqueue_step(D, shape, st, ev.data()), "step");
  } else {
    HIPCHK(enqueue_substeps(D, shape, st, ev.data()), "substeps");
  }
  HIPCHK(hipMemcpyAsync(&hd, g.diag, sizeof(hd), hipMemcpyDeviceToHost, st), "diag copy");
  HIPCHK(hipStreamSynchronize(st), "step sync");
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[2 + 4 * (nsub - 1) + 2]), "elapsed");
  dg.device_ms = ms;
  if (!(g.use_graph && one_sequence)) {
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]), "elapsed");
    dg.prep_ms = ms;
  }
  for (int sub = 0; sub < nsub && !g.use_graph; sub++) {
    hipEvent_t start = (sub == 0) ? ev[1] : ev[2 + 4 * (sub - 1) + 2];
    hipEvent_t *e4 = &ev[2 + 4 * sub];
    HIPCHK(hipEventElapsedTime(&ms, start, e4[0]), "elapsed");
    dg.stream_ms += ms;
    HIPCHK(hipEventElapsedTime(&ms, e4[0], e4[1]), "elapsed");
    dg.focus_ms += ms;
    HIPCHK(hipEventElapsedTime(&ms, e4[1], e4[2]), "elapsed");
    dg.change_ms += ms;
  }
  if (hd.prep_err) {
    // the preparation failed: the substeps did nothing (the reference
    // panics before them)
    int code = decode_err(hd.err_key, &dg);
    if (out) *out = dg;
    return fail(code, "%s (stream %d)", err_text(code), dg.err_stream);
  }
  dg.max_subcycles_change = do_change ? hd.max_sub_change : 0;
  dg.max_subcycles_focusing = do_focus ? hd.max_sub_focus : 0;
  dg.max_nsteps_stream = do_stream ? hd.max_nsteps : 0;
  double mt;
  memcpy(&mt, &hd.min_tau_bits, 8);
  dg.min_tau = mt;
  int64_t cells = 0;
  for (int32_t n : g.h_nactive) cells += (int64_t)n * g.d.num_energies * g.d.num_mu;
  dg.cell_updates = cells * nsub;
  dg.stream_sweeps = do_stream ? (int64_t)hd.stream_sweeps * nsub : 0;
  dg.focus_sweeps = do_focus ? (int64_t)hd.focus_sweeps * nsub : 0;
  dg.change_sweeps = do_change ? (int64_t)hd.change_sweeps * nsub : 0;
  dg.stream_launches = !do_stream ? 0 : sched == 0 ? hd.max_passes * nsub : sched == 1 ? nsub : nsub * g.d.num_streams;
  dg.heavy_nodes = do_focus ? (int32_t)hd.focus_heavy : 0;
  int code = EP_OK;
  if (hd.err_key != ~0ull) code = decode_err(hd.err_key, &dg);
  if (out) *out = dg;
  if (code != EP_OK)
    return fail(code, "%s (stream %d node %d energy %d mu %d)", err_text(code), dg.err_stream,
                dg.err_node, dg.err_energy, dg.err_mu);
  return EP_OK;
}

int ep_get_node_outputs(double *ds, double *bmag_plus, double *bmag_minus) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  const size_t S = g.d.num_streams, N = g.d.num_nodes;
  double *dst[3] = {ds, bmag_plus, bmag_minus};
  const double *src[3] = {g.ds, g.bplus, g.bminus};
  for (int i = 0; i < 3; i++)
    if (dst[i])
      HIPCHK(hipMemcpy2D(dst[i], N * 8, src[i], g.Npad * 8, N * 8, S, hipMemcpyDeviceToHost),
             "node download");
  return EP_OK;
}

int ep_last_error(char *buf, size_t len) {
  if (buf && len) {
    strncpy(buf, g_err.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return EP_OK;
}

int ep_finalize(void) {
  if (g.init) free_all();
  return EP_OK;
}

// Test hook: checks the DPP lane layout of one wave on the device.
int ep_selftest_dpp(const double *in64, double *shr64, double *shl64) {
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, 3 * 64 * 8), "hipMalloc");
  hipError_t e = hipMemcpy(d, in64, 64 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_dpp_probe(d, d + 64, d + 128, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(shr64, d + 64, 64 * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(shl64, d + 128, 64 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "dpp selftest");
  return EP_OK;
}

// Test hook: a plain device-to-device copy kernel of `bytes` bytes (a
// multiple of 16), run `reps` times; *ms = mean kernel time.  The reference
// point for HBM byte counters and bandwidth.
// fm_log and fm_exp of n values on the host (the same code the device
// runs), for the accuracy tests.
int ep_fastmath_host(const double *x, int64_t n, double *log_out, double *exp_out) {
  if (!x || n < 0 || !log_out || !exp_out) return fail(EP_ERR_INVALID_ARG, "bad arguments");
  std::vector<double> t(kFmTable);
  build_fm_table(t.data());
  for (int64_t i = 0; i < n; i++) {
    log_out[i] = fm_log(x[i], t.data());
    exp_out[i] = fm_exp(x[i], t.data());
  }
  return EP_OK;
}

// The same on the device (needs ep_init for the tables).
int ep_selftest_fastmath(const double *x, int64_t n, double *log_out, double *exp_out) {
  if (!g.init) return fail(EP_ERR_NOT_INIT, "ep_init first");
  if (!x || n < 1 || !log_out || !exp_out) return fail(EP_ERR_INVALID_ARG, "bad arguments");
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, (size_t)n * 3 * sizeof(double)), "hipMalloc");
  hipError_t e = hipMemcpy(d, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_fastmath_probe(g.fm_tab_d, d, n, d + n, d + 2 * n, g.st);
  if (e == hipSuccess) e = hipStreamSynchronize(g.st);
  if (e == hipSuccess) e = hipMemcpy(log_out, d + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(exp_out, d + 2 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "fastmath probe");
  return EP_OK;
}

int ep_selftest_copy(size_t bytes, int32_t reps, double *ms) {
  if (bytes == 0 || bytes % 16 || reps < 1) return fail(EP_ERR_INVALID_ARG, "bytes must be a positive multiple of 16");
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, bytes), "hipMalloc");
  hipError_t e = hipMalloc(&b, bytes);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return hip_fail(e, "hipMalloc");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  e = hipMemset(a, 0, bytes);
  if (e == hipSuccess) e = launch_copy_probe(a, b, bytes, 0);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy_probe(a, b, bytes, 0);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return hip_fail(e, "copy probe");
  if (ms) *ms = t / reps;
  return EP_OK;
}

}  // extern "C"
