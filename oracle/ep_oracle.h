/*
 * ep_oracle.h -- FP64 CPU oracle for the EPREM energetic-particle update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's hot path, used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker and the CPU baseline.  The
 * product (eprem_amd/) never links, loads or calls it.
 *
 * Every routine follows the reference's operation order, compiled with
 * -ffp-contract=off, so that it reproduces the reference arithmetic.  Each
 * cites the reference file:line it restates.
 *
 * Layouts: node arrays are [stream][node]; f is [stream][node][energy][mu]
 * (the reference's eParts order, src/global.h:61, with species = 1).
 */
#ifndef EP_ORACLE_H
#define EP_ORACLE_H

#include "eprem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Energy/mu grids, initEnergeticParticlesGrids
 * (src/energeticParticlesInit.c:36-100). */
void or_build_grids(const ep_params *p, int E, int M, double *egrid,
                    double *vgrid, double *vgrid_i, double *pgrid,
                    double *rigidity, double *mugrid, double *dlnp,
                    double *dmu);

/* sepSeedFunction (src/energeticParticlesBoundary.c:67-116). */
double or_seed(const ep_params *p, double energy, double r);

/* meanFreePath (src/meanFreePath.c:13-50); rig = rigidity[e]. */
double or_mfp(const ep_params *p, double rig, double rmag, double bmag);

/* Whole updateEnergeticParticles() (src/energeticParticles.c:69-203) for
 * all streams.  n_active may be NULL (all = num_nodes).  ds/bplus/bminus
 * receive updateStreamValues' per-node output (may be NULL).  nthreads:
 * OpenMP threads over streams (<=0: library default).  Returns EP_OK or
 * the first panic's EP_ERR_* code with its location in diag. */
int or_update_ep(const ep_dims *d, const ep_params *p, const int *n_active,
                 const double *x, const double *y, const double *z,
                 const double *rmag, const double *bmag, const double *dlnb,
                 const double *dlnn, const double *dupar, double *f,
                 double *ds, double *bplus, double *bminus, uint32_t op_flags,
                 int nthreads, ep_diag *diag);

/* GetStreamList + FindSegmentLengths (src/energeticParticles.c:2147-2234,
 * :2493-2601) for one stream of n nodes with positions x,y,z.  Writes
 * shell_list[<=n], shell_ref[n], ds_list[<=n], ds_i_list[<=n], *ds_min;
 * returns streamlistSize, or EP_ERR_DS_UNDERFLOW. */
int or_get_stream_list(const ep_params *p, int n, const double *x,
                       const double *y, const double *z, int *shell_list,
                       int *shell_ref, double *ds_list, double *ds_i_list,
                       double *ds_min);

/* ---------------- MAS trilinear (src/masInterp.c) ----------------- */

/* masTriLinearBinarySearch (src/masInterp.c:117-171). */
double or_mas_bsearch(const float *A, double key, int *imin, int *imax,
                      int lower, int upper);

/* masTriLinear + masGetNode + masWind for a batch of nodes
 * (src/masInterp.c:30-68, :177-200, :539-741, :1007-1065).
 * fields0/1[7], meshes0/1[21], dims0/1[21] as in ep_mas_set_level.
 * Positions x,y,z,rmag are in rScale units.  prev_* give the node's
 * previous state for masWind (may be NULL when all nodes are inside).
 * Outputs [n] in code units. */
void or_mas_interp(int n, double rScale, const double *x, const double *y,
                   const double *z, const double *rmag,
                   const float *const fields0[7], const float *const meshes0[21],
                   const int *dims0, const float *const fields1[7],
                   const float *const meshes1[21], const int *dims1,
                   double s_cor, double phi_offset, double mas_radial_max,
                   double mas_vmin, const double *rold_x,
                   const double *rold_y, const double *rold_z,
                   const double *prev_br, const double *prev_bt,
                   const double *prev_bp, const double *prev_vr,
                   const double *prev_vt, const double *prev_vp,
                   const double *prev_d, double *br, double *bt, double *bp,
                   double *vr, double *vt, double *vp, double *dens);

/* updateMhd (src/flow.c:33-154) for a batch of n nodes.  mode 0 = the
 * analytic wind (mhdDensity/mhdV/mhdB with MHD_DEFAULT, src/flow.c:241-403),
 * mode 1 = MAS (or_mas_interp, same level arguments).  old_* hold the
 * node's values before the update; is_first[i] != 0 marks the first node of
 * a stream (its derivative terms are zeroed, src/flow.c:131-136).
 * Outputs: new B (r, theta, phi), V (r, theta, phi), density, |B| and the
 * derivative terms dupar, dlnb, dlnn. */
void or_update_mhd(int n, int mode, double rScale, double mhdUs, double mhdNsAu,
                   double mhdBsAu, double omegaSun, const double *x, const double *y,
                   const double *z, const double *rmag, const int *is_first,
                   const float *const fields0[7], const float *const meshes0[21],
                   const int *dims0, const float *const fields1[7],
                   const float *const meshes1[21], const int *dims1, double s_cor,
                   double phi_offset, double mas_radial_max, double mas_vmin,
                   const double *rold_x, const double *rold_y, const double *rold_z,
                   const double *old_br, const double *old_bt, const double *old_bp,
                   const double *old_vr, const double *old_vt, const double *old_vp,
                   const double *old_d, double *br, double *bt, double *bp, double *vr,
                   double *vt, double *vp, double *dens, double *bmag, double *dupar,
                   double *dlnb, double *dlnn);

/* Point observers: per stream the projection onto the observer sphere
 * (pointObsProjectedStreamData + findIntersection, src/search.c:113-268),
 * the inverse-distance blend over streams (src/unifiedOutput.c:846-935),
 * and the omnidirectional flux 2 E <f>_mu of the blended distribution
 * (computeFlux, src/unifiedOutput.c:153-180).  Observer radii in AU,
 * angles in radians; node positions in rScale units; f in the reference
 * layout.  dist_out [nobs][E][M], flux_out [nobs][E]. */
void or_observers(const ep_dims *d, const int *n_active, double rScale, double idw_p, int nobs,
                  const double *obs_r, const double *obs_theta, const double *obs_phi,
                  const double *x, const double *y, const double *z, const double *rmag,
                  const double *f, const double *egrid, double *dist_out, double *flux_out);

/* computeFlux (src/unifiedOutput.c:153-180) for every node: flux [S][N][E]. */
void or_stream_flux(const ep_dims *d, const double *f, const double *egrid, double *flux);

/* ---------------- Parker-wind host pipeline ------------------------ */

/* Inputs of the analytic-wind pipeline (src/eprem.c:426-744 with
 * masCouple = 0), in input-file units (converted inside the way
 * setRuntimeConstants does, src/configuration.c:324-348). */
typedef struct or_sim_config {
  int rows, cols, num_nodes, num_energies, num_mu;
  double rScale, flowMag, mhdDensityAu, mhdBAu;
  double simStartTime, simStopTime, tDel;   /* [days] */
  int numEpSteps;
  double aziSunStart, omegaSun;
  double dsh_min, lamo, mfpRadialPower, mfpBPower, rigidityPower;
  int mfpType;
  double eMin, eMax, mass, charge, dampCooling;
  double boundaryFunctAmplitude, boundaryFunctXi, boundaryFunctGamma,
         boundaryFunctBeta, boundaryFunctEcutoff, boundaryFunctR0;
  int useParallelDiffusion, useAdiabaticFocus, useAdiabaticChange,
      checkSeedPopulation, useEPBoundary, useBoundaryFunction,
      boundaryFunctionInitDomain, seedFunctionTest, dumpFreq;
  double unifiedOutputTime;   /* < 0: default simStartTime + tDel/2 days */
  double epCalcStartTime;     /* < 0: default simStartTime             */
} or_sim_config;

/* Per-dump stream output (writeObserverDataNetCDF,
 * src/unifiedOutput.c:442-549), values already multiplied by the
 * netCDF scale_factor (physical units) unless raw is set.  Buffers are sized by the caller
 * for max_dumps dumps: time[max_dumps], node fields
 * [max_dumps][S][N] (R, T, P, Br, Bt, Bp, Vr, Vt, Vp, Rho),
 * dist [max_dumps][S][N][E][M].  Per-step diagnostics go to
 * step_diag[max_steps] (may be NULL). */
typedef struct or_sim_output {
  int max_dumps, num_dumps, max_steps, num_steps;
  double *time;
  double *R, *T, *P, *Br, *Bt, *Bp, *Vr, *Vt, *Vp, *Rho;
  double *dist;
  double *egrid, *vgrid, *mu;
  ep_diag *step_diag;
  int raw;  /* nonzero: store the code-unit values the file holds, before
               the scale_factor a netCDF reader applies */
} or_sim_output;

/* moveNodes for the analytic wind (src/simCore.c:736-785): n nodes move
 * by tDel * V / rScale, V = (vr, vt, vp) made Cartesian at the node; the
 * old positions go to rox/roy/roz. */
void or_move_nodes(int n, double tDel, double rScale, double *x, double *y, double *z, double *rmag,
                   double *rox, double *roy, double *roz, const double *vr, const double *vt,
                   const double *vp);

/* masMoveNodes for n nodes (RK4 through the MAS field, see the source);
 * prev[7] = the nodes' MHD state {br, bt, bp, vr, vt, vp, density}. */
void or_move_nodes_mas(int n, double tDel, double rScale, double *x, double *y, double *z,
                       double *rmag, double *rox, double *roy, double *roz,
                       const double *const prev[7], const float *const fields0[7],
                       const float *const meshes0[21], const int *dims0,
                       const float *const fields1[7], const float *const meshes1[21],
                       const int *dims1, double s, double phi_offset, double mas_radial_max,
                       double mas_vmin);

/* ShellData + DiffuseShellData (src/energeticParticles.c:247-555) for one
 * EP substep dt: f [S][N][E][M] in/out, neighbour streams nbr[S][4],
 * positions [S][N] (rScale units), B (r, theta, phi) [S][N]; node 0 is
 * skipped.  bad[4] (may be NULL) = stream, node, energy, mu of the first
 * NaN/Inf in the reference's check order. */
int or_shell_diffuse(const ep_params *p, int S, int N, int E, int M, const int32_t *nbr,
                     const double *x, const double *y, const double *z, const double *rmag,
                     const double *br, const double *bt, const double *bp, const double *rigidity,
                     const double *vgrid, double dt, double kperxkpar, double *f, int *bad);

/* curlBoverB2 (src/flow.c:479-615): curl(B/B^2) of the analytic (Parker)
 * field by centred differences at a cartesian node position (rScale
 * units) with radial flow speed vr; curl[3] = (r, theta, phi). */
void or_curl_b_over_b2(double rScale, double mhdBsAu, double omega, double x, double y, double z, double vr,
                       double *curl);

/* DriftShellData (src/energeticParticles.c:560-801) for one EP substep dt:
 * curl [S][N][3] per node, pgrid[E] momentum per energy; otherwise as
 * or_shell_diffuse. */
int or_shell_drift(const ep_params *p, int S, int N, int E, int M, const int32_t *nbr, const double *x,
                   const double *y, const double *z, const double *curl, const double *pgrid,
                   const double *vgrid, double dt, double *f, int *bad);

/* One application of the focusing operator (alg 1 upwind, 3 WENO3;
 * src/energeticParticles.c:1806-1896, :1899-2144) to f [E][M] with the
 * half-mesh velocities v_avg [E][M+1]: f1 = d(v f)/dmu [E][M]. */
void or_focus_operator(int alg, int E, int M, double dmu, const double *f, const double *v_avg, double *f1);
/* The adiabatic-change operator of p->adiabaticChangeAlg (1/2 upwind, 3
 * WENO3; :1081-1176, :1179-1427) on ln f [E][M] with v_avg [E+1][M] at
 * node radius rmag (inflow boundary = ln of the seed spectrum). */
void or_change_operator(const ep_params *p, int E, int M, double rmag, const double *f, const double *v_avg,
                        double *f1);

/* Run the whole analytic-wind simulation.  Returns EP_OK or an error. */
int or_run_parker(const or_sim_config *c, or_sim_output *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
