// ep_state.h -- host-side record of the library's device allocations.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "ep_internal.h"

namespace ep {

// One MAS time level resident on the device: 7 FP32 fields and their
// 21 1-D meshes (src/readMAS.c:1005-1069 reads the same from HDF5).
struct MasLevel {
  float *field[7] = {nullptr};
  float *mesh[21] = {nullptr};
  int dims[21] = {0};
  std::vector<float> hmesh[21];  // host copies: which meshes are the same
  bool set = false;
};

// Counters of the learned-surrogate form over one ep_step (device side).
struct SurStats {
  unsigned long long network_slabs, fallback_slabs, plain_slabs, network_cell_subcycles;
  unsigned long long max_res_bits;  // positive doubles order like their bits
  unsigned rejected_stream_substeps, pad_;
};

// The learned-surrogate form of the EP update (ep_sur.hip): its device
// buffers, allocated by ep_set_surrogate and released with the others.
struct SurState {
  int K = 0;          // streaming subcycles per network application (0: off)
  double tol = 1e-3;  // residual above which a stream falls back to the plain path
  int G = 0;          // k_sur_apply work items per slab
  int stride = 0;     // kept nodes per sampled node of the plain check
  int NS = 0;         // most sampled nodes of a slab
  double *fC = nullptr;              // third f buffer, the layout of fA/fB
  int32_t *A = nullptr;              // [S*E] network applications this substep
  int32_t *reject = nullptr;         // [S] stream falls back this substep
  double *errsum = nullptr;          // [S*E] summed sampled errors of one application
  unsigned long long *res = nullptr; // [S*E] largest residual of the substep (bits)
  int32_t *nrem = nullptr, *hrem = nullptr, *prem = nullptr, *crem = nullptr;  // [S*E] plain remainder
  double *phi = nullptr;             // [S*E] fraction of the substep left to plain focusing
  float *scal = nullptr;             // [S*E*6*Npad] scalar inputs of every kept node
  SurStats *stats = nullptr;
  const void *packed = nullptr;      // network weights the captured graph reads
};

// What a captured substep graph depends on besides device memory.
struct StepShape {
  int passes, chunks, nsub, do_stream, do_focus, do_change, do_seed, do_shell;
  unsigned generation;
  int node_fused;  // focusing and change in one kernel (k_focus_change)
  int sched;       // streaming schedule (0: a launch per pass, 1: persistent, 2: per stream)
  int sur_levels;  // learned-surrogate form: 1 + network application levels; 0 = off
  bool operator==(const StepShape &o) const {
    return passes == o.passes && chunks == o.chunks && nsub == o.nsub && do_stream == o.do_stream &&
           do_focus == o.do_focus && do_change == o.do_change && do_seed == o.do_seed &&
           do_shell == o.do_shell && generation == o.generation && node_fused == o.node_fused &&
           sched == o.sched && sur_levels == o.sur_levels;
  }
};

struct State {
  bool init = false, nodes_set = false, positions_set = false;
  ep_dims d{};
  ep_params p{};
  int Npad = 0;
  hipStream_t st = nullptr;
  int stream_block = 384, stream_halo = 32, stream_r = 3, stream_fuse = 1, stream_whole = 1;
  // streaming schedule: 0 = one launch per pass over every slab, 1 = one
  // persistent launch per substep over whole slabs (k_stream_slab), 2 = the
  // unbatched form of 1, one launch per stream
  int stream_sched = 1;
  int node_fuse = 1;  // allow k_focus_change when every node needs <= 1 change subcycle
  // graph of the substep loop; `generation` changes whenever kernel
  // arguments captured by value (parameters, tiling) change
  bool use_graph = true;
  int fp32 = 0;  // f stored as float (mixed-precision form)
  int fast = 0;  // streaming with fused multiply-adds
  int logf = 0;  // float storage holds ln f (precision mode 3)
  unsigned generation = 0;
  hipGraphExec_t graph_exec = nullptr;
  StepShape graph_shape{};
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    generation++;
  }
  // host copies of the grids
  std::vector<double> egrid, vgrid, vgrid_i, rigidity, rig_lamo, plt, ext, mugrid, pgrid;
  double dlnp = 0, dmu = 0, normJ0 = 0, normRadius = 0;
  int mpos = 0;
  std::vector<int32_t> h_nactive;
  // device buffers
  double *egrid_d = nullptr, *vgrid_d = nullptr, *vgrid_i_d = nullptr, *rig_lamo_d = nullptr,
         *plt_d = nullptr, *ext_d = nullptr, *mugrid_d = nullptr;
  double *fm_tab_d = nullptr;  // exp/log tables of the fused precision modes
  double *fA = nullptr, *fB = nullptr;
  uint8_t *loc = nullptr;
  int32_t *fhead = nullptr;  // [1] the slot of node 0 in every f row (fidx)
  int fhead_h = 0;           // host copy
  double *nx = nullptr, *ny = nullptr, *nz = nullptr, *nrmag = nullptr, *nbmag = nullptr,
         *ndlnb = nullptr, *ndlnn = nullptr, *ndupar = nullptr;
  int32_t *n_active = nullptr;
  double *ds = nullptr, *bplus = nullptr, *bminus = nullptr, *mfp_r = nullptr, *radial = nullptr;
  int32_t *list = nullptr, *ref = nullptr, *nkept = nullptr, *thin = nullptr;
  double *dsil = nullptr, *dsmin = nullptr;
  int32_t *nsteps = nullptr, *halo = nullptr, *passes = nullptr, *chunks = nullptr;
  double *dtprop = nullptr, *del_fac = nullptr, *seed3 = nullptr, *dsm = nullptr, *q = nullptr;
  int32_t *nsub_f = nullptr, *nsub_c = nullptr, *heavy = nullptr;
  double *dtsub_f = nullptr, *dtsub_c = nullptr, *f_dlnbds = nullptr, *f_dupar = nullptr,
         *f_b = nullptr, *c_dupar = nullptr, *c_dlnn = nullptr, *c_dlnb = nullptr,
         *c_damp = nullptr, *c_bc0 = nullptr, *c_bc1 = nullptr;
  DevDiag *diag = nullptr;
  std::vector<void *> allocs;
  std::vector<hipEvent_t> events;
  // MAS interpolation
  MasLevel mas[2];
  std::vector<void *> mas_allocs;
  double *mas_out = nullptr;   // 7 x S x Npad
  double *mas_prev = nullptr;  // 10 x S x Npad
  // cross-stream operators (diffusion off while kperxkpar == 0)
  double kperxkpar = 0.0;
  int shell_drift = 0;
  int32_t *shell_nbr = nullptr, *shell_src = nullptr, *shell_dir = nullptr;
  double *shell_dlper = nullptr, *shell_dl = nullptr, *shell_curl = nullptr, *shell_factor = nullptr;
  // analytic field of the last ep_update_mhd (mode 0), for drift
  int mhd_mode = -1;
  double mhd_bsau = 0.0, mhd_omega = 0.0;
  // the MHD buffer's state half holds the latest B, V, density per node
  // once ep_update_mhd has run (ep_move_nodes reads V from it)
  bool mhd_state = false;
  SurState sur;  // learned-surrogate form (ep_sur.hip)
  // point observers, stream flux and shell-spawn staging (allocated on
  // first use)
  double *obs_buf = nullptr, *flux_buf = nullptr, *spawn_buf = nullptr;
  size_t obs_cap = 0, flux_cap = 0, spawn_cap = 0;
  void mas_free() {
    for (void *p : mas_allocs) (void)hipFree(p);
    mas_allocs.clear();
    if (obs_buf) (void)hipFree(obs_buf);
    if (flux_buf) (void)hipFree(flux_buf);
    if (spawn_buf) (void)hipFree(spawn_buf);
    obs_buf = flux_buf = spawn_buf = nullptr;
    int32_t *ip[3] = {shell_nbr, shell_src, shell_dir};
    for (int32_t *q : ip)
      if (q) (void)hipFree(q);
    double *dp[4] = {shell_dlper, shell_dl, shell_curl, shell_factor};
    for (double *q : dp)
      if (q) (void)hipFree(q);
    shell_nbr = shell_src = shell_dir = nullptr;
    shell_dlper = shell_dl = shell_curl = shell_factor = nullptr;
    kperxkpar = 0.0;
    shell_drift = 0;
    mhd_mode = -1;
    obs_cap = flux_cap = spawn_cap = 0;
    mhd_state = false;
    mas[0] = MasLevel();
    mas[1] = MasLevel();
    mas_out = nullptr;
    mas_prev = nullptr;
  }
};

extern State g;
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
// device address of ep_observe_noalloc's fixed scratch (ep_obs.hip)
int scratch_address(double **ptr);
Dev make_dev();

}  // namespace ep
