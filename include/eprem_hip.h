/*
 * eprem_hip.h -- C ABI of the MI355X energetic-particle update.
 *
 * This is the drop-in boundary for EPREM's per-stream particle update.  A C
 * host that owns the cube-shell grid, node motion, MHD background and output
 * calls these entry points in place of the reference's internal functions:
 *
 *   ep_step            replaces updateEnergeticParticles()
 *                          src/energeticParticles.h:51, called at src/eprem.c:671
 *                      (which runs, per EP substep, updateStreamValues
 *                       src/simCore.c:325, seedPopulationCheck
 *                       src/energeticParticlesBoundary.c:28, GetStreamList
 *                       src/energeticParticles.c:2147, DiffuseStreamData :2237,
 *                       AdiabaticFocusing :1533 and AdiabaticChange :805),
 *                      and the MPI stream gather/scatter src/simCore.c:177-318,
 *                      which disappears because every stream is resident.
 *   ep_set_nodes       replaces the Node_t fields the operators read
 *                          (src/cubeShellStruct.h:51-93)
 *   ep_upload_f /      replace direct access to eParts[]
 *   ep_download_f          (src/global.h:61, allocated src/global.c:112)
 *   ep_mas_set_level   replaces the MAS shared-memory windows filled by
 *                          masGetInterpData() (src/readMAS.c:645-980)
 *   ep_interp_mhd      replaces the per-node masGetNode()/masTriLinear()
 *                          loop inside updateMhd() (src/flow.c:48-77,
 *                          src/masInterp.c:30-68, :539-741, :1007-1065)
 *
 * Conventions
 *   - Every call returns EP_OK (0) or a negative EP_ERR_* code.  The
 *     reference has no return codes: it panics (src/error.c:39-49).  The
 *     host is expected to print ep_last_error() and abort, exactly like
 *     panic() does.
 *   - All physical quantities are in the reference's code units: lengths of
 *     node positions in units of rScale AU, times in TAU = AU/c, speeds in c,
 *     B in MHD_B_NORM, energies in m_p c^2 (src/baseTypes.h:50-115).
 *   - One species only (num_species = 1).  The reference's multi-species
 *     path indexes the per-energy egrid[] with idx_se(species, energy)
 *     (src/energeticParticles.c:1137-1138, src/energeticParticlesTypes.c:58-61),
 *     which reads past the array for species > 0, so there is no defined
 *     multi-species behaviour to reproduce.
 *   - Host arrays stay owned by the host.  The library owns its device
 *     buffers, allocated once in ep_init / ep_mas_set_level; ep_step does
 *     not allocate.
 *   - Single-threaded host calls, one HIP stream, no callbacks.
 */
#ifndef EPREM_HIP_H
#define EPREM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Error codes (mirror the reference's panic sites)                     */
/* ------------------------------------------------------------------ */
#define EP_OK                 0
#define EP_ERR_INVALID_ARG   -1  /* bad dims/params/pointers            */
#define EP_ERR_NOT_INIT      -2  /* call before ep_init                 */
#define EP_ERR_HIP           -3  /* HIP runtime error                   */
#define EP_ERR_NAN           -4  /* checkNaN  src/error.c:62-80          */
#define EP_ERR_INF           -5  /* checkInf  src/error.c:84-100         */
#define EP_ERR_DS_UNDERFLOW  -6  /* src/energeticParticles.c:2594-2595   */
#define EP_ERR_UNSUPPORTED   -7  /* option outside the implemented set  */
#define EP_ERR_STREAM_SHORT  -8  /* fewer than 3 kept nodes on a stream */
#define EP_ERR_NO_DEVICE     -9  /* no usable GPU                       */

/* Layout selectors for ep_upload_f / ep_download_f. */
#define EP_LAYOUT_REFERENCE   0  /* f[stream][node][energy][mu]  (idx_frcsspem
                                    of src/global.h:61 with species=1 and
                                    the buffer shell removed)          */

/* Operator flags for ep_step (AND-ed with the params' use* switches). */
#define EP_OP_STREAM      0x1u   /* DiffuseStreamData (+GetStreamList)  */
#define EP_OP_FOCUS       0x2u   /* AdiabaticFocusing                   */
#define EP_OP_CHANGE      0x4u   /* AdiabaticChange                     */
#define EP_OP_SEEDCHECK   0x8u   /* seedPopulationCheck                 */
#define EP_OP_SHELL       0x10u  /* DiffuseShellData  (when enabled by */
#define EP_OP_DRIFT       0x20u  /* DriftShellData     ep_shell_operators) */
#define EP_OP_ALL         0x3Fu

/* ------------------------------------------------------------------ */
/* Dimensions and parameters                                           */
/* ------------------------------------------------------------------ */
typedef struct ep_dims {
  int32_t num_streams;   /* NUM_STREAMS = 6 * rows * cols                */
  int32_t num_nodes;     /* TOTAL_NUM_SHELLS: max nodes per stream        */
  int32_t num_energies;  /* NUM_ESTEPS                                    */
  int32_t num_mu;        /* NUM_MUSTEPS (2..32)                           */
} ep_dims;

/* Runtime parameters, already converted to code units the way
 * setRuntimeConstants() does (src/configuration.c:324-348). */
typedef struct ep_params {
  double rScale;             /* config.rScale  [AU]                        */
  double tDel;               /* config.tDel    [TAU] (after /DAY)          */
  int32_t numEpSteps;        /* config.numEpSteps                          */
  int32_t mfpType;           /* 1: r-power, 2: B-power  (meanFreePath.c)   */
  double lamo;               /* config.lamo                                */
  double mfpRadialPower;     /* config.mfpRadialPower                      */
  double mfpBPower;          /* config.mfpBPower                           */
  double mhdBsAu;            /* config.mhdBAu / MHD_B_NORM                 */
  double rigidityPower;      /* config.rigidityPower                       */
  double eMin, eMax;         /* [MeV/nucleon]                              */
  double mass, charge;       /* species 0: config.mass[0], config.charge[0]*/
  double dsh_min;            /* config.dsh_min (dsh_hel_min not used: the
                                helio-coupled phase is out of scope)       */
  double dampCooling;        /* config.dampCooling                         */
  double boundaryFunctAmplitude, boundaryFunctXi, boundaryFunctGamma,
         boundaryFunctBeta, boundaryFunctEcutoff, boundaryFunctR0;
  int32_t useParallelDiffusion, useAdiabaticFocus, useAdiabaticChange,
          checkSeedPopulation, useEPBoundary, useBoundaryFunction;
  int32_t adiabaticFocusAlg;  /* 1 upwind Euler, 2 RK3 upwind, 3 RK3 WENO3  */
  int32_t adiabaticChangeAlg; /* (3 needs >= 4 mu cells / energies; 2 and 3
                                 need FP64 storage)                        */
} ep_params;

/* Per-node inputs, each array [num_streams][num_nodes] (node fastest).
 * x,y,z,rmag: node position (Node_t.r, .rmag); bmag: Node_t.mhdBmag;
 * dlnb, dlnn, dupar: Node_t.mhdDlnB/.mhdDlnN/.mhdDuPar (set by updateMhd). */
typedef struct ep_node_soa {
  const double *x, *y, *z, *rmag;
  const double *bmag, *dlnb, *dlnn, *dupar;
} ep_node_soa;

/* Per-global-step diagnostics (src/energeticParticles.c:149-156, :943-945,
 * :1668-1670, reduced at src/eprem.c:694-716) plus error location. */
typedef struct ep_diag {
  int32_t max_subcycles_change;    /* maxsubcycles_energychange           */
  int32_t max_subcycles_focusing;  /* maxsubcycles_focusing               */
  int32_t max_nsteps_stream;       /* largest DiffuseStreamData nsteps     */
  int32_t err_code;                /* EP_OK or the first error found      */
  double  min_tau;                 /* min_tau                             */
  int32_t err_stream, err_node, err_energy, err_mu;
  int64_t cell_updates;            /* active cells x EP substeps          */
  double  device_ms;               /* HIP-event time of the whole step    */
  double  stream_ms, focus_ms, change_ms, prep_ms; /* per-operator time   */
  /* Work done, in cell-sweeps (one cell advanced one CFL subcycle of one
   * operator), summed over the step's EP substeps. */
  int64_t stream_sweeps, focus_sweeps, change_sweeps;
  int32_t stream_launches;         /* streaming pass launches in the step  */
  int32_t heavy_nodes;             /* nodes whose focusing needs 32 or more
                                      subcycles (focused by the heavy-node
                                      walk; 0 without focusing)            */
} ep_diag;

/* ------------------------------------------------------------------ */
/* Entry points                                                        */
/* ------------------------------------------------------------------ */

/* Allocate device state for dims, build the energy/mu grids
 * (initEnergeticParticlesGrids, src/energeticParticlesInit.c:36-100) and
 * set every f to DBL_MIN like that routine does. */
int ep_init(const ep_dims *dims, const ep_params *params);

/* Replace the parameters (dims unchanged). */
int ep_set_params(const ep_params *params);

/* Copy the energy/mu grids built by ep_init back to the host; any pointer
 * may be NULL.  Arrays have num_energies / num_mu entries. */
int ep_get_grids(double *egrid, double *vgrid, double *mugrid,
                 double *rigidity, double *dlnp, double *dmu);

/* Upload node data.  n_active (may be NULL = all num_nodes) gives the
 * number of nodes of each stream (2..num_nodes); nodes past it are
 * ignored.  The reference has the same count on every stream
 * (TOTAL_NUM_SHELLS); per-stream counts serve batched workloads. */
int ep_set_nodes(const ep_node_soa *nodes, const int32_t *n_active);

/* Copy f between host and device (layout: EP_LAYOUT_*).  The host array
 * covers num_streams*num_nodes*num_energies*num_mu doubles. */
int ep_upload_f(const double *f, int32_t layout);
int ep_download_f(double *f, int32_t layout);

/* The same for streams stream0 .. stream0+count-1 only: the host array
 * holds count*num_nodes*num_energies*num_mu doubles, those streams in the
 * reference layout.  The host's access to one stream's distribution, as
 * the reference's gather update_stream_from_shells / scatter
 * update_shells_from_stream give it (src/simCore.c:177-318), without
 * moving the whole distribution over PCIe. */
int ep_upload_f_streams(int32_t stream0, int32_t count, const double *f, int32_t layout);
int ep_download_f_streams(int32_t stream0, int32_t count, double *f, int32_t layout);

/* Fill f with the seed population sepSeedFunction(E, rmag) at every node
 * (initEnergeticParticles, src/energeticParticlesInit.c:106-147). */
int ep_init_seed_population(void);

/* One global step: numEpSteps EP substeps of dt = tDel/numEpSteps, all on
 * the device (updateEnergeticParticles, src/energeticParticles.c:69-203).
 * op_flags (EP_OP_*) select operators on top of the params' switches.
 * diag may be NULL.  Returns the first device error (NaN, Inf, ds
 * underflow) with its location in diag. */
int ep_step(uint32_t op_flags, ep_diag *diag);

/* Node values the step computes for the host (updateStreamValues,
 * src/simCore.c:325-435): per node ds, mhdBmagPlus, mhdBmagMinus. */
int ep_get_node_outputs(double *ds, double *bmag_plus, double *bmag_minus);

/* ---------------- MAS trilinear interpolation ---------------------- */

/* One MAS time level.  fields[7] = {Bp, Bt, Br, Vp, Vt, Vr, D} in MAS
 * units, each (phi, theta, r) with r fastest, i.e. Vindex(r,t,p) =
 * r + nr*t + nr*nt*p (src/masInterp.c:185).  Field k has its own 1-D
 * meshes: meshes[3k+0] = r [nr_k] (in Rs), meshes[3k+1] = theta [nt_k],
 * meshes[3k+2] = phi [np_k]; dims[3k+0..2] = nr_k, nt_k, np_k.
 * level is 0 or 1 (the two time levels blended by s_cor). */
int ep_mas_set_level(int32_t level, const float *const fields[7],
                     const float *const meshes[21], const int32_t dims[21]);

/* Interpolation constants (src/masInterp.c:41, :706). */
typedef struct ep_mas_params {
  double s_cor;          /* time blend factor (readMAS.c:681-708)          */
  double phi_offset;     /* phiOffset                                      */
  double mas_radial_max; /* config.masRadialMax [Rs]                        */
  double mas_vmin;       /* config.masVmin [cm/s]                           */
} ep_mas_params;

/* Per-node previous MHD state used by the outer fallback masWind()
 * (src/masInterp.c:1007-1065); arrays [num_streams][num_nodes]. */
typedef struct ep_mas_prev {
  const double *rold_x, *rold_y, *rold_z;
  const double *br, *btheta, *bphi, *vr, *vtheta, *vphi, *density;
} ep_mas_prev;

/* Output: per node B (r, theta, phi), V (r, theta, phi) and density in
 * code units, each [num_streams][num_nodes]. */
typedef struct ep_mas_out {
  double *br, *btheta, *bphi, *vr, *vtheta, *vphi, *density;
} ep_mas_out;

/* Interpolate both MAS levels onto the node positions last given to
 * ep_set_nodes (masGetNode for every node).  prev may be NULL when every
 * node lies inside mas_radial_max.  out receives host copies (may be
 * NULL: results stay on the device).  *device_ms, if given, receives the
 * kernel time. */
int ep_interp_mhd(const ep_mas_params *mp, const ep_mas_prev *prev,
                  ep_mas_out *out, double *device_ms);

/* ---------------- MHD update on the device ------------------------- */

/* Node positions only, [num_streams][num_nodes] in rScale units, with the
 * per-stream active counts (NULL = all).  Used with ep_update_mhd, which
 * then fills the MHD node fields on the device. */
int ep_set_positions(const double *x, const double *y, const double *z, const double *rmag,
                     const int32_t *n_active);

typedef struct ep_mhd_params {
  int32_t mode;           /* 0: analytic wind (MHD_DEFAULT), 1: MAS        */
  double mhdUs, mhdNsAu, mhdBsAu, omegaSun;   /* mode 0, code units        */
  ep_mas_params mas;      /* mode 1                                        */
} ep_mhd_params;

/* updateMhd (src/flow.c:33-154) for every node at the positions on the
 * device: new B, V and density (MAS: masGetNode; mode 0: mhdDensity, mhdV,
 * mhdB, src/flow.c:241-403), |B| and the derivative terms mhdDuPar,
 * mhdDlnB, mhdDlnN against `prev` (zero on the first node of each stream,
 * src/flow.c:131-149).  The results feed ep_step directly (no host copy);
 * `out` (may be NULL) receives host copies of B, V, density. */
int ep_update_mhd(const ep_mhd_params *mp, const ep_mas_prev *prev, ep_mas_out *out,
                  double *device_ms);
/* With prev = NULL, ep_update_mhd uses the state the device already holds:
 * the B, V and density its previous call computed (carried along by
 * ep_spawn_shell) and the old positions ep_move_nodes recorded. */

/* ---------------- node motion on the device ------------------------ */

/* The new innermost shell of each stream: position [num_streams] in rScale
 * units and, optionally, its distribution [num_streams][E][M] (NULL: the
 * seed spectrum sepSeedFunction gives at its radius). */
typedef struct ep_spawn {
  const double *x, *y, *z, *rmag;
  const double *f;
} ep_spawn;

/* rotSunAndSpawnShell's ripple and spawnNewShell (src/simCore.c:442-728):
 * every node record (position, MHD state, f) moves one node outward, the
 * outermost one is dropped once a stream is full, node 0 becomes the new
 * shell, and n_active grows by one up to num_nodes.  The host computes the
 * spawn positions (the rotating template shell). */
int ep_spawn_shell(const ep_spawn *sp);

/* moveNodes with the analytic-wind flow (src/simCore.c:736-785,
 * delrFlow src/flow.c:169-178): every node but node 0 moves by
 * tdel * V / rScale, V the node's flow velocity from the last
 * ep_update_mhd made Cartesian at the node; old positions are kept for the
 * next ep_update_mhd.  tdel in code units. */
int ep_move_nodes(double tdel);

/* masMoveNodes (src/masInterp.c:1344-1393, rungeKuttaFlow :1305-1338,
 * vMas :1279-1296): every node but node 0 takes one classical RK4 step of
 * length tdel through the MAS velocity, each stage interpolated from the
 * two levels set with ep_mas_set_level at s_cor (as ep_update_mhd, mode 1)
 * at the stage position; outside mas_radial_max masWind of the node's
 * device-held state (its old position and MHD values from the last
 * ep_update_mhd).  The position before the move becomes the node's old
 * position.  Call it where the reference calls masMoveNodes: after the
 * spawn and before the next ep_update_mhd.  tdel in code units. */
int ep_move_nodes_mas(const ep_mas_params *mp, double tdel);

/* Host copies of the device node positions and active counts (any pointer
 * may be NULL). */
int ep_get_positions(double *x, double *y, double *z, double *rmag, int32_t *n_active);

/* ---------------- observers and flux -------------------------------- */

/* Point observers (config.obsR/obsTheta/obsPhi, idw_p). */
typedef struct ep_observers {
  int32_t num;                 /* 1..256                                 */
  const double *r_au;          /* radius [AU]                            */
  const double *theta, *phi;   /* colatitude, longitude [rad]            */
  double idw_p;                /* inverse-distance weight power          */
} ep_observers;

/* Distribution and omnidirectional flux at each observer for the current
 * f and node positions: per stream the projection onto the observer
 * sphere (pointObsProjectedStreamData + findIntersection,
 * src/search.c:113-268), the inverse-distance blend over streams
 * (writePointObserverDataNetCDF, src/unifiedOutput.c:846-935) and
 * flux = 2 E <f>_mu (computeFlux, src/unifiedOutput.c:153-180).
 * dist [num][E][M] and flux [num][E] (either may be NULL), code units. */
int ep_observe(const ep_observers *obs, double *dist, double *flux);

/* ep_observe without any allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  Its partial
 * results (observer coordinates, per-stream projections and weights) live
 * in a fixed 32 MiB device buffer that the library holds on each GPU it
 * runs on until the process ends, instead of the buffer ep_observe
 * allocates; a request needing more (3 num + num S (E M + 1) + num E (M +
 * 1) doubles) fails with EP_ERR_INVALID_ARG.  Results are identical to
 * ep_observe's.  Every call on a GPU shares that buffer: calls that
 * overlap on one GPU (other host threads) race. */
int ep_observe_noalloc(const ep_observers *obs, double *dist, double *flux);

/* computeFlux (src/unifiedOutput.c:153-180) on every node:
 * flux [num_streams][num_nodes][num_energies], code units. */
int ep_stream_flux(double *flux);

/* ---------------- cross-stream operators --------------------------- */
/* DiffuseShellData with the link lengths of ShellData and DriftShellData
 * with driftVelocity/curlBoverB2 (src/energeticParticles.c:208-801,
 * src/flow.c:479-615): perpendicular diffusion (kperxkpar > 0, config.
 * kperxkpar; 0 = off) and gradient/curvature drift (use_drift != 0) between
 * the streams of each shell, every EP substep after focusing and change.
 * nbr[S][4] = north, east, west, south neighbour stream of every stream
 * (the cube surface links of initNEWS, src/cubeShellInit.c:95-520).  Both
 * need every stream full and the device MHD state (ep_update_mhd); drift
 * uses the analytic-wind field (mode 0), like the reference's curlBoverB2. */
int ep_shell_operators(const int32_t *nbr, double kperxkpar, int32_t use_drift);

/* ---------------- learned surrogate: multilayer perceptron --------- */
/* No counterpart in the reference.  A fully connected network (ReLU on
 * hidden layers) whose products run on the matrix cores (bf16 operands,
 * fp32 accumulation); parameters are one flat fp32 vector, per layer
 * W [in][out] row-major then b [out].  Rows are fp32 [n][width]. */
int ep_mlp_init(const int32_t *sizes, int32_t nsizes, uint64_t seed);
int ep_mlp_num_params(int64_t *count);
int ep_mlp_get_params(float *params);
int ep_mlp_set_params(const float *params);
int ep_mlp_forward(const float *x, int32_t n, float *y);
/* mean squared error over n rows and its gradient (flat, like params) */
int ep_mlp_loss_grad(const float *x, const float *y, int32_t n, double *loss, float *grad);
/* Adam minibatch training; loss_per_epoch[epochs] may be NULL */
int ep_mlp_train(const float *x, const float *y, int32_t n, int32_t epochs, int32_t batch, float lr,
                 uint64_t seed, float *loss_per_epoch);
/* forward pass over n device-resident rows, reps times; *ms = mean time.
 * layerwise = 0: the fused one-kernel forward when the network has its
 * shape (input <= 128, hidden layers of 128, output <= 32); 1: one product
 * kernel per layer */
int ep_mlp_bench(int32_t n, int32_t reps, int32_t layerwise, double *ms);
int ep_mlp_finalize(void);

/* Training pairs for the large-step surrogate: n random kept nodes of
 * (stream, energy) slabs of full-resolution streams with at least K
 * streaming subcycles per substep, on the current nodes and f (FP64
 * storage).  x [n][3M+6]: per mu cell ln f(k-1) - ln f(k), ln f(k) minus
 * its mean over mu, ln f(k+1) - ln f(k), then T v/ds_k, T/tau_k, ln(ds_k/ds_{k-1}), ln(ds_{k+1}/ds_k),
 * asinh(a T/dmu), asinh(b T/dmu) with T = K dtProp and a, b the focusing
 * velocity's terms; y [n][M]: ln f at k after K streaming subcycles of the
 * plain path (DiffuseStreamData, src/energeticParticles.c:2352-2430) and
 * the upwind focusing over T (:1806-1896), minus the same mean; valid[n] =
 * 1 for rows that qualified; where [n][3] (may be NULL): each row's
 * (stream, energy, node).  f is unchanged. */
int ep_surrogate_samples(int32_t K, int32_t n, uint64_t seed, float *x, float *y, int32_t *valid,
                         int32_t *where);

/* Learned-surrogate form of ep_step.  No counterpart in the reference:
 * it replaces, per EP substep, K CFL-limited subcycles of streaming with
 * relaxation (DiffuseStreamData, src/energeticParticles.c:2237-2490) plus
 * the adiabatic focusing of the same time (AdiabaticFocusing, :1533-1896)
 * by one application of the network set with ep_mlp_init /
 * ep_mlp_set_params / ep_mlp_train (3M+6 inputs as in
 * ep_surrogate_samples, hidden layers of 128, M outputs), on every
 * (stream, energy) slab with at least K subcycles.  A stream whose
 * conservation residual after an application exceeds tol (or whose
 * network output is not finite) falls back to the plain path for that
 * substep.  K = 0 turns it off.  Needs FP64 storage and the upwind
 * operators; allocates a third f buffer on first use. */
int ep_set_surrogate(int32_t K, double tol);

typedef struct ep_surrogate_stats {
  int64_t network_slabs;          /* slab-substeps advanced by the network and accepted */
  int64_t fallback_slabs;         /* slab-substeps of rejected streams, redone by the plain path */
  int64_t plain_slabs;            /* slab-substeps with fewer than K subcycles (plain path) */
  int64_t network_cell_subcycles; /* cell-subcycles of streaming + focusing the accepted applications cover */
  int32_t rejected_stream_substeps; /* (stream, substep) pairs that fell back */
  int32_t pad_;
  double max_residual;            /* largest residual of any application */
} ep_surrogate_stats;

/* Counters of the last ep_step in the surrogate form. */
int ep_get_surrogate_stats(ep_surrogate_stats *out);

/* Test probes of the surrogate form on the current nodes and f: mode 0
 * applies the network once to every slab with at least K subcycles and
 * makes the result current (the other slabs keep theirs); mode 1 runs the
 * plain path for exactly K streaming subcycles into the spare buffer and
 * leaves f as it was.  Either way *max_residual receives the largest
 * conservation residual of that one application. */
int ep_surrogate_probe(int32_t mode, double *max_residual);

/* ---------------- misc --------------------------------------------- */
int ep_last_error(char *buf, size_t len);
int ep_finalize(void);

/* Tiling of the streaming kernel: kept nodes per chunk (stream_block, halo
 * included), Euler subcycles per pass (stream_halo, 1..stream_block/2-1),
 * kept nodes per lane (lane_nodes, 1, 2 or 3; stream_block must be a multiple
 * of 64*lane_nodes with at most 512 lanes; 1 is refused with FP32 storage,
 * precision modes 1 and 3) and fuse_focus (1: the
 * focusing of each node of an unthinned stream is done by its last
 * streaming pass on the values still in registers; 0: by the separate
 * focusing kernel).  Results are the same either way.  No counterpart in
 * the reference (its DiffuseStreamData loops on the CPU); set by the
 * autotuner or left at the defaults (384, 32, 3, 1).  A stream whose kept
 * nodes fit in one chunk runs all its subcycles in one pass without a
 * halo. */
int ep_set_tuning(int32_t stream_block, int32_t stream_halo, int32_t lane_nodes, int32_t fuse_focus);

/* How the streaming passes of an EP substep are launched: 1 (default) =
 * one persistent launch per substep whose blocks take whole (stream,
 * energy) slabs from a device queue, highest energies first, and run all
 * their passes and chunks; 0 = one launch per pass over every slab; 2 = the
 * unbatched form of 1, one launch per stream (the reference's loop over
 * streams, src/energeticParticles.c:102-160), kept for comparison.
 * Results are identical.  ep_get_schedule reports the schedule in effect
 * (set by the autotuner, or by ep_set_schedule).  No counterpart in the
 * reference. */
int ep_set_schedule(int32_t schedule);
int ep_get_schedule(int32_t *schedule);

/* Tile-size search for the streaming kernel on the current nodes: times
 * every streaming pass of one substep and the node stage after it (the
 * focusing the passes leave to do, with the adiabatic change when the two
 * run as one kernel) for each candidate tiling without changing f or the
 * step state, and sets the fastest.  Any output pointer may be NULL.  No counterpart in the
 * reference. */
int ep_autotune(int32_t *stream_block, int32_t *stream_halo, int32_t *lane_nodes, int32_t *fuse_focus,
                double *stream_ms);

/* ep_step runs its substep loop as one captured HIP graph (enable = 1, the
 * default; rebuilt when the pass count, grid or parameters change) or as
 * individual launches with per-operator timing in ep_diag (enable = 0).
 * No counterpart in the reference. */
int ep_set_graph(int32_t enable);

/* Node-stage kernels: fused = 1 (default) runs AdiabaticFocusing and
 * AdiabaticChange of each node in one pass over f (k_focus_change) in
 * every step whose nodes all need at most one change subcycle; fused = 0
 * always runs them as two kernels.  Results are bit-identical either way.
 * No counterpart in the reference (its two operators loop over the nodes of
 * a stream on the CPU, src/energeticParticles.c:163-177). */
int ep_set_node_fusion(int32_t fused);

/* Precision of the hot path: 0 = FP64 in the reference's operation order
 * (default; the parity path), 1 = mixed: f stored as float, streaming
 * (upwind step) and focusing in FP32 (FP32 accumulation), MAS
 * interpolation in FP32, the relaxation (pitch-angle part) and the
 * adiabatic energy change (ln f flux) in FP64, 2 = FP64 with fused
 * multiply-adds in streaming and relaxation (same results to ~1e-15
 * relative, 6 instead of 9 vector operations per cell and subcycle), 3 =
 * mode 1 with the float storage holding ln f instead of f (the log-scaled
 * variant; no shell operators).  Mode 3 streams with the slab launch, two
 * nodes per lane, and runs focusing and change as two kernels: setting it
 * moves the tiling to two nodes per lane (same lane count), switches
 * fused focusing off and the schedule to 1 if it was 0, and ep_set_tuning
 * and ep_set_schedule reject other choices while it is set.  Modes 1 and
 * 3 stream two or three nodes per lane: setting mode 1 moves a tiling of
 * one node per lane to two (same lane count).  The distribution on the
 * device is
 * converted when the storage changes; the float floor is FLT_MIN where the
 * reference floors at DBL_MIN.  No counterpart in the reference. */
int ep_set_precision(int32_t mode);

/* Device self-test of the cross-lane shift the streaming kernel uses: for
 * one 64-lane wave, shr[i] = in[i-1] (shr[0] = -1) and shl[i] = in[i+1]
 * (shl[63] = -2).  No counterpart in the reference; used by the tests. */
int ep_selftest_dpp(const double *in64, double *shr64, double *shl64);

/* Test hook: a plain copy kernel of `bytes` bytes (multiple of 16) on a
 * scratch pair of buffers, `reps` times; *ms = mean kernel time.  The
 * reference point for HBM byte counters and bandwidth. */
int ep_selftest_copy(size_t bytes, int32_t reps, double *ms);

/* The table-driven exp and log of the fused precision modes (Tang's
 * methods; the reference-order mode uses the device libm): n values of
 * log(x) and exp(x), computed on the host (ep_fastmath_host, usable
 * without a GPU) or on the device (ep_selftest_fastmath, after ep_init).
 * The two give identical bits.  No counterpart in the reference. */
int ep_fastmath_host(const double *x, int64_t n, double *log_out, double *exp_out);
int ep_selftest_fastmath(const double *x, int64_t n, double *log_out, double *exp_out);

#ifdef __cplusplus
}
#endif
#endif /* EPREM_HIP_H */
