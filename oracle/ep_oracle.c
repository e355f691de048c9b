/*
 * ep_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement of the reference hot path, used only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product
 * never links it.  Build: oracle/Makefile (gcc -O2 -std=c99
 * -ffp-contract=off -fopenmp).  Each function cites the reference lines it
 * follows; the arithmetic keeps the reference's operation order so that it
 * reproduces the reference's rounding.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ep_oracle.h"

/* Physical constants (src/baseTypes.h:50-115). */
#define OR_PI 3.1415926535897932
#define OR_TWO_PI 6.283185307179586
#define OR_VERYSMALL 1.0e-33
#define OR_MP 1.6726e-24
#define OR_MEV 1.6022e-6
#define OR_GEV 1.6022e-3
#define OR_C 2.99792458e10
#define OR_MZERO (OR_MP * OR_C * OR_C / OR_GEV)
#define OR_AU 1.495978707e13
#define OR_RSUN 6.96e10
#define OR_RSAU (OR_RSUN / OR_AU)
#define OR_TAU (OR_AU / OR_C)
#define OR_DAY (OR_TAU / (24.0 * 60.0 * 60.0))
#define OR_MHD_DENSITY_NORM 1.0
#define OR_MHD_B_NORM (sqrt(OR_MP * OR_MHD_DENSITY_NORM) * OR_C)
#define OR_MAS_TIME_NORM 1445.87003080685
#define OR_MAS_LENGTH_NORM 6.96e10
#define OR_MAS_RHO_NORM 1.6726e-16
#define OR_MAS_V_CONVERT (OR_MAS_LENGTH_NORM / OR_MAS_TIME_NORM) / OR_C
#define OR_MAS_RHO_CONVERT (OR_MAS_RHO_NORM / OR_MP) / OR_MHD_DENSITY_NORM
#define OR_MAS_B_CONVERT \
  sqrt(4.0 * OR_PI * OR_MAS_RHO_NORM) * OR_MAS_LENGTH_NORM / OR_MAS_TIME_NORM / OR_MHD_B_NORM

/* ------------------------------------------------------------------ */
/* Grids  (src/energeticParticlesInit.c:36-100)                        */
/* ------------------------------------------------------------------ */
void or_build_grids(const ep_params *p, int E, int M, double *egrid,
                    double *vgrid, double *vgrid_i, double *pgrid,
                    double *rigidity, double *mugrid, double *dlnp_out,
                    double *dmu_out) {
  const double one = 1.0, two = 2.0, half = 0.5;
  double kin, lnpmin, lnpmax, dlnp, dmu, pp;
  int e, m;
  kin = p->eMin * OR_MEV / (OR_MP * OR_C * OR_C);
  lnpmin = log(sqrt(kin * kin + two * kin));
  kin = p->eMax * OR_MEV / (OR_MP * OR_C * OR_C);
  lnpmax = log(sqrt(kin * kin + two * kin));
  dlnp = (lnpmax - lnpmin) / (E - 1.0);
  for (e = 0; e < E; e++) {
    pp = exp(lnpmin + e * dlnp);
    if (pgrid) pgrid[e] = pp;
    egrid[e] = sqrt(one + pp * pp) - one;
    vgrid[e] = pp / sqrt(one + pp * pp);
    vgrid_i[e] = sqrt(one + pp * pp) / pp;
  }
  for (e = 0; e < E; e++) {
    pp = exp(lnpmin + e * dlnp);
    rigidity[e] = (p->mass / p->charge) * pp * OR_MZERO;
    rigidity[e] = pow(rigidity[e], p->rigidityPower);
  }
  dmu = two / (one * M);
  for (m = 0; m < M; m++) {
    double mup = -one + (m + one) * dmu;
    double mum = -one + m * dmu;
    mugrid[m] = half * (mup + mum);
  }
  *dlnp_out = dlnp;
  *dmu_out = dmu;
}

/* sepSeedFunction (src/energeticParticlesBoundary.c:67-116). */
double or_seed(const ep_params *p, double energy, double r) {
  const double two = 2.0;
  const double E1 = 1.0;
  const double efac = OR_MEV / (OR_MP * OR_C * OR_C);
  double normJ0 = p->boundaryFunctAmplitude * (OR_MP * OR_C) / (OR_MHD_DENSITY_NORM * OR_MEV);
  double normRadius = p->boundaryFunctR0 / p->rScale;
  double normE1 = E1 * efac;
  double normE0 = p->boundaryFunctEcutoff * efac;
  double radialTerm = pow((r / normRadius), -p->boundaryFunctBeta);
  double powerLawTerm = pow((energy / normE1), -p->boundaryFunctGamma);
  double expTerm = exp(-(energy / normE0));
  double normJHe = normJ0 * radialTerm * powerLawTerm * expTerm;
  double normJ = normJHe / p->boundaryFunctXi;
  double normf = normJ / (two * energy);
  if (normf < DBL_MIN) normf = DBL_MIN;
  return normf;
}

/* meanFreePath (src/meanFreePath.c:13-50). */
double or_mfp(const ep_params *p, double rig, double rmag, double bmag) {
  if (p->mfpType == 2)
    return rig * p->lamo * pow(p->mhdBsAu / bmag, p->mfpBPower);
  return rig * p->lamo * pow(rmag * p->rScale, p->mfpRadialPower);
}

/* ------------------------------------------------------------------ */
/* Shared grid context                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
  int E, M;
  double *egrid, *vgrid, *vgrid_i, *rigidity, *mugrid;
  double dlnp, dmu;
} or_grid;

static void grid_alloc(or_grid *g, const ep_params *p, int E, int M) {
  g->E = E;
  g->M = M;
  g->egrid = (double *)malloc(sizeof(double) * E);
  g->vgrid = (double *)malloc(sizeof(double) * E);
  g->vgrid_i = (double *)malloc(sizeof(double) * E);
  g->rigidity = (double *)malloc(sizeof(double) * E);
  g->mugrid = (double *)malloc(sizeof(double) * M);
  or_build_grids(p, E, M, g->egrid, g->vgrid, g->vgrid_i, NULL, g->rigidity,
                 g->mugrid, &g->dlnp, &g->dmu);
}

static void grid_free(or_grid *g) {
  free(g->egrid);
  free(g->vgrid);
  free(g->vgrid_i);
  free(g->rigidity);
  free(g->mugrid);
}

/* Error bookkeeping: the first panic in the reference's own traversal
 * order wins.  Keys order events like the reference loops do. */
typedef struct {
  int code;
  unsigned long long key;
  int stream, node, energy, mu;
} or_err;

static void err_note(or_err *e, int code, unsigned long long key, int s,
                     int n, int en, int m) {
  if (e->code == EP_OK || key < e->key) {
    e->code = code;
    e->key = key;
    e->stream = s;
    e->node = n;
    e->energy = en;
    e->mu = m;
  }
}

/* key layout: [substep 8][group 4][stream 16][op 2][a 16][b 8][c 8]
 * (a,b,c follow the reference loop nesting of the failing operator) */
static unsigned long long mkkey(int sub, int group, int s, int op, int a,
                                int b, int c) {
  unsigned long long k = 0;
  k |= (unsigned long long)(sub > 255 ? 255 : sub) << 56;
  k |= (unsigned long long)(group & 0xF) << 52;
  k |= (unsigned long long)(s & 0xFFFF) << 36;
  k |= (unsigned long long)(op & 0x3) << 34;
  k |= (unsigned long long)(a & 0xFFFF) << 18;
  k |= (unsigned long long)(b & 0xFF) << 10;
  k |= (unsigned long long)(c & 0x3FF);
  return k;
}

/* ------------------------------------------------------------------ */
/* updateStreamValues (src/simCore.c:325-435)                          */
/* ------------------------------------------------------------------ */
static void update_stream_values(const ep_params *p, int n, const double *x,
                                 const double *y, const double *z,
                                 const double *bmag, double *ds,
                                 double *bplus, double *bminus) {
  const double rs = p->rScale;
  int k;
  for (k = 0; k < n; k++) {
    if (k == 0) {
      double px = x[k + 1] * rs, py = y[k + 1] * rs, pz = z[k + 1] * rs;
      double cx = x[k] * rs, cy = y[k] * rs, cz = z[k] * rs;
      ds[k] = sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy) + (pz - cz) * (pz - cz));
      bminus[k] = bmag[k];
      bplus[k] = bmag[k + 1];
    } else if (k == n - 1) {
      double cx = x[k] * rs, cy = y[k] * rs, cz = z[k] * rs;
      double mx = x[k - 1] * rs, my = y[k - 1] * rs, mz = z[k - 1] * rs;
      ds[k] = sqrt((cx - mx) * (cx - mx) + (cy - my) * (cy - my) + (cz - mz) * (cz - mz));
      bminus[k] = bmag[k - 1];
      bplus[k] = bmag[k];
    } else {
      double px = x[k + 1] * rs, py = y[k + 1] * rs, pz = z[k + 1] * rs;
      double cx = x[k] * rs, cy = y[k] * rs, cz = z[k] * rs;
      double mx = x[k - 1] * rs, my = y[k - 1] * rs, mz = z[k - 1] * rs;
      double dsp = sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy) + (pz - cz) * (pz - cz));
      double dsm = sqrt((cx - mx) * (cx - mx) + (cy - my) * (cy - my) + (cz - mz) * (cz - mz));
      ds[k] = 0.5 * (dsp + dsm);
      bminus[k] = bmag[k - 1];
      bplus[k] = bmag[k + 1];
    }
  }
}

/* seedPopulationCheck (src/energeticParticlesBoundary.c:28-60). */
static void seed_population_check(const ep_params *p, const or_grid *g,
                                  int n, const double *rmag, double *f) {
  int k, e, m;
  for (k = 0; k < n; k++)
    for (e = 0; e < g->E; e++) {
      double sd = or_seed(p, g->egrid[e], rmag[k]);
      double *fr = f + ((size_t)k * g->E + e) * g->M;
      for (m = 0; m < g->M; m++)
        if (fr[m] < sd) fr[m] = sd;
    }
}

/* ------------------------------------------------------------------ */
/* GetStreamList + FindSegmentLengths                                  */
/*   (src/energeticParticles.c:2147-2234, :2493-2601)                  */
/* ------------------------------------------------------------------ */
static double dist3(const ep_params *p, const double *x, const double *y,
                    const double *z, int a, int b) {
  const double rs = p->rScale;
  double ax = x[a] * rs, ay = y[a] * rs, az = z[a] * rs;
  double bx = x[b] * rs, by = y[b] * rs, bz = z[b] * rs;
  return sqrt((ax - bx) * (ax - bx) + (ay - by) * (ay - by) + (az - bz) * (az - bz));
}

int or_get_stream_list(const ep_params *p, int n, const double *x,
                       const double *y, const double *z, int *list,
                       int *ref, double *dsl, double *dsil, double *ds_min) {
  int shell, size = 0, down = 0, up, sl;
  const double one = 1.0, half = 0.5;
  double dmin;
  for (shell = 0; shell < n; shell++) {
    double dsh;
    if (shell < n - 1) {
      up = shell + 1;
    } else {
      up = shell;
      down = shell - 1;
    }
    /* r1 = streamGrid[up]*rScale, r0 = streamGrid[down]*rScale */
    dsh = dist3(p, x, y, z, up, down);
    if ((dsh > p->dsh_min) || (shell == n - 1)) {
      list[size++] = shell;
      ref[shell] = shell;
      down = shell + 1;
    } else {
      int sp = shell - 1;
      ref[shell] = shell + 1;
      while ((sp >= 0) && (ref[sp] != sp)) {
        ref[sp] = ref[sp + 1];
        sp--;
      }
    }
  }
  dmin = 10000.0 * p->rScale;
  for (sl = 0; sl < size; sl++) {
    int ts = list[sl];
    if (sl == 0) {
      double d = dist3(p, x, y, z, list[sl + 1], ts);
      if (d < dmin) dmin = d;
      dsl[sl] = d;
    } else if (sl == size - 1) {
      double d = dist3(p, x, y, z, ts, list[sl - 1]);
      if (d < dmin) dmin = d;
      dsl[sl] = d;
    } else {
      double d1 = dist3(p, x, y, z, list[sl + 1], ts);
      double d9 = dist3(p, x, y, z, ts, list[sl - 1]);
      if (d1 < dmin) dmin = d1;
      if (d9 < dmin) dmin = d9;
      dsl[sl] = half * (d1 + d9);
    }
    if (dmin <= 0.0) {
      *ds_min = dmin;
      return EP_ERR_DS_UNDERFLOW;
    }
    dsil[sl] = one / dsl[sl];
  }
  *ds_min = dmin;
  return size;
}

/* ------------------------------------------------------------------ */
/* DiffuseStreamData (src/energeticParticles.c:2237-2490)              */
/* ------------------------------------------------------------------ */
static void diffuse_stream(const ep_params *p, const or_grid *g, int s,
                           int sub, int n, const double *rmag,
                           const double *bmag, const int *list,
                           const int *ref, int size, const double *dsil,
                           double dsMin, double dt, double *f, int *max_nsteps,
                           or_err *err) {
  const int E = g->E, M = g->M;
  const double one = 1.0;
  const double NUM_MUSTEPS_I = one / M;
  double *f_old = (double *)malloc(sizeof(double) * size * M);
  double *f_new = (double *)malloc(sizeof(double) * size * M);
  double *dsm = (double *)malloc(sizeof(double) * size);
  double *q = (double *)malloc(sizeof(double) * size);
  double *iso_vec = (double *)malloc(sizeof(double) * size);
  double *del_fac = (double *)malloc(sizeof(double) * M);
  double seed3[3];
  int e, m, k, step, shell;
  for (e = 0; e < E; e++) {
    double v = g->vgrid[e], vi = g->vgrid_i[e];
    double dtMin = 0.4 * dsMin * vi;
    int nsteps = (int)floor(dt / dtMin + one);
    double dtProp = dt / (one * nsteps);
    double dtProp_i = one / dtProp;
    if (nsteps > *max_nsteps) *max_nsteps = nsteps;
    for (k = 0; k < size; k++) {
      int sh = list[k];
      double tau = vi * or_mfp(p, g->rigidity[e], rmag[sh], bmag[sh]);
      if (dtProp > tau)
        dsm[k] = dsil[k] * tau * dtProp_i;
      else
        dsm[k] = dsil[k] * one;
      q[k] = exp(-dtProp / tau);
    }
    for (k = 0; k < 3; k++) seed3[k] = or_seed(p, g->egrid[e], rmag[list[k]]);
    for (m = 0; m < M; m++) del_fac[m] = v * dtProp * g->mugrid[m];
    for (m = 0; m < M; m++)
      for (k = 0; k < size; k++) f_old[size * m + k] = f[((size_t)list[k] * E + e) * M + m];
    for (step = 0; step < nsteps; step++) {
      if ((p->useEPBoundary > 0) && (p->useBoundaryFunction > 0))
        for (m = 0; m < M; m++)
          for (k = 0; k < 3; k++) f_old[size * m + k] = seed3[k];
      for (m = 0; m < M; m++) {
        double df = del_fac[m];
        double *fo = f_old + size * m, *fn = f_new + size * m;
        if (g->mugrid[m] >= 0.0) {
          fn[0] = fo[0];
          for (k = 1; k < size; k++) fn[k] = fo[k] + dsm[k] * df * (fo[k - 1] - fo[k]);
        } else {
          fn[size - 1] = fo[size - 1];
          for (k = 0; k < size - 1; k++) fn[k] = fo[k] + dsm[k] * df * (fo[k] - fo[k + 1]);
        }
      }
      for (k = 0; k < size; k++) iso_vec[k] = 0.0;
      for (m = 0; m < M; m++)
        for (k = 0; k < size; k++) iso_vec[k] = iso_vec[k] + f_new[size * m + k];
      for (m = 0; m < M; m++)
        for (k = 0; k < size; k++) {
          double iso = NUM_MUSTEPS_I * iso_vec[k];
          f_old[size * m + k] = iso + (f_new[size * m + k] - iso) * q[k];
        }
    }
    for (m = 0; m < M; m++)
      for (k = 0; k < size; k++) f[((size_t)list[k] * E + e) * M + m] = f_old[size * m + k];
    for (m = 0; m < M; m++)
      for (shell = 0; shell < n; shell++) {
        double *cell = &f[((size_t)shell * E + e) * M + m];
        if (shell != ref[shell]) *cell = f[((size_t)ref[shell] * E + e) * M + m];
        if (isnan(*cell))
          err_note(err, EP_ERR_NAN, mkkey(sub, 0, s, 1, e, m, shell), s, shell, e, m);
        else if (isinf(*cell))
          err_note(err, EP_ERR_INF, mkkey(sub, 0, s, 1, e, m, shell), s, shell, e, m);
        if (*cell < DBL_MIN) *cell = DBL_MIN;
      }
  }
  free(f_old);
  free(f_new);
  free(dsm);
  free(q);
  free(iso_vec);
  free(del_fac);
}

/* ------------------------------------------------------------------ */
/* AdiabaticFocusing + _Operator_Upwind                                */
/*   (src/energeticParticles.c:1533-1803, :1806-1896)                  */
/* ------------------------------------------------------------------ */
static void focus_operator_upwind(int E, int M, double dmu, double *f1,
                                  const double *f, const double *v_avg,
                                  double *flux) {
  const double upwind = 1.0;
  const double dmu_i = 1.0 / dmu;
  int e, m;
  /* flux / v_avg are indexed [e][mu] with stride M+1 (idx_spemp1) */
  for (e = 0; e < E; e++)
    for (m = 1; m < M; m++) {
      double va = v_avg[e * (M + 1) + m];
      double cce = copysign(upwind, va);
      flux[e * (M + 1) + m] =
          va * 0.5 * ((1.0 - cce) * f[e * M + m] + (1.0 + cce) * f[e * M + m - 1]);
    }
  for (e = 0; e < E; e++)
    for (m = 1; m < M - 1; m++)
      f1[e * M + m] = (flux[e * (M + 1) + m + 1] - flux[e * (M + 1) + m]) * dmu_i;
  for (e = 0; e < E; e++) {
    double vl = v_avg[e * (M + 1) + 1];
    double vr = v_avg[e * (M + 1) + M - 1];
    if (vl <= 0)
      f1[e * M + 0] = vl * f[e * M + 1] * dmu_i;
    else
      f1[e * M + 0] = vl * f[e * M + 0] * dmu_i;
    if (vr >= 0)
      f1[e * M + M - 1] = vr * (-f[e * M + M - 2]) * dmu_i;
    else
      f1[e * M + M - 1] = vr * (-f[e * M + M - 1]) * dmu_i;
  }
}

/* The WENO3 flux of one face from the split fluxes LN (left cells) and LP
 * (right cells) around it: lnm2, lnm1, ln0 = LN at the cells two left, one
 * left and right of the face; lpm1, lp0, lpp1 = LP at the cells left, right
 * and two right of the face (AdiabaticFocusing_Operator_WENO3 and
 * AdiabaticChange_Operator_WENO3 share it, src/energeticParticles.c:
 * 2070-2103, :1321-1354). */
static double weno3_face(double lnm2, double lnm1, double ln0, double lpm1, double lp0, double lpp1) {
  const double D_C_CPt = 1 / 2.0, D_C_MCt = 1 / 2.0, D_M_Tt = 1 / 3.0, D_CP_Tt = 2 / 3.0,
               D_P_Tt = 1 / 3.0, D_MC_Tt = 2 / 3.0;
  const double weno_eps = 10.0 * sqrt(DBL_MIN);
  double p0m = (1.0 + D_C_MCt) * lnm1 - D_C_MCt * lnm2;
  double p1m = D_C_MCt * lnm1 + D_C_CPt * ln0;
  double p0p = (1.0 + D_C_CPt) * lp0 - D_C_CPt * lpp1;
  double p1p = D_C_CPt * lp0 + D_C_MCt * lpm1;
  double B0m = 4.0 * pow((D_C_MCt * (lnm1 - lnm2)), 2);
  double B1m = 4.0 * pow((D_C_CPt * (ln0 - lnm1)), 2);
  double B0p = 4.0 * pow((D_C_CPt * (lpp1 - lp0)), 2);
  double B1p = 4.0 * pow((D_C_MCt * (lp0 - lpm1)), 2);
  double w0m = D_P_Tt / pow((weno_eps + B0m), 2);
  double w1m = D_MC_Tt / pow((weno_eps + B1m), 2);
  double w0p = D_M_Tt / pow((weno_eps + B0p), 2);
  double w1p = D_CP_Tt / pow((weno_eps + B1p), 2);
  double wm_sum = w0m + w1m, wp_sum = w0p + w1p;
  double um = (w0m / wm_sum) * p0m + (w1m / wm_sum) * p1m;
  double up = (w0p / wp_sum) * p0p + (w1p / wp_sum) * p1p;
  return up + um;
}

/* alpha of the WENO3 operators at cell i of a line of n cells whose
 * half-mesh velocities va[0..n] are extended by copies at both ends: the
 * largest |va| over faces i-2..i+3 that exist, in the reference's order
 * (src/energeticParticles.c:1961-2017 for mu, :1218-1274 for energy). */
static double weno3_alpha(const double *va, int stride, int i, int n) {
  double v = fabs(va[i * stride]);
#define OR_AL(j) \
  if (fabs(va[(j) * stride]) > v) v = fabs(va[(j) * stride]);
  if (i >= 2 && i < n - 2) {
    OR_AL(i - 1) OR_AL(i - 2) OR_AL(i + 1) OR_AL(i + 2) OR_AL(i + 3)
  } else if (i == 0) {
    OR_AL(i + 1) OR_AL(i + 2) OR_AL(i + 3)
  } else if (i == 1) {
    OR_AL(i - 1) OR_AL(i + 1) OR_AL(i + 2) OR_AL(i + 3)
  } else if (i == n - 2) {
    OR_AL(i + 1) OR_AL(i + 2) OR_AL(i - 1) OR_AL(i - 2)
  } else {
    OR_AL(i + 1) OR_AL(i - 1) OR_AL(i - 2)
  }
#undef OR_AL
  return v;
}

/* AdiabaticFocusing_Operator_WENO3 (src/energeticParticles.c:1899-2144)
 * for one energy row (v_avg / flux [M+1], f / f1 [M]); needs M >= 4.
 * ln / lp / al are scratch rows of M. */
static void focus_row_weno3(int M, double dmu, double *f1, const double *f, const double *va, double *flux,
                            double *ln, double *lp, double *al) {
  const double upwind = 1.0;
  int m;
  double cce;
  for (m = 0; m < M; m++) al[m] = weno3_alpha(va, 1, m, M);
  for (m = 0; m < M; m++) {
    lp[m] = 0.5 * f[m] * va[m + 1] - al[m];
    ln[m] = 0.5 * f[m] * va[m] + al[m];
  }
  for (m = 2; m < M - 1; m++) flux[m] = weno3_face(ln[m - 2], ln[m - 1], ln[m], lp[m - 1], lp[m], lp[m + 1]);
  m = 1;
  cce = copysign(upwind, va[m]);
  flux[m] = va[m] * 0.5 * ((1.0 - cce) * f[m] + (1.0 + cce) * f[m - 1]);
  m = M - 1;
  cce = copysign(upwind, va[m]);
  flux[m] = va[m] * 0.5 * ((1.0 - cce) * f[m] + (1.0 + cce) * f[m - 1]);
  for (m = 1; m < M - 1; m++) f1[m] = (flux[m + 1] - flux[m]) / dmu;
  if (va[1] <= 0)
    f1[0] = va[1] * f[1] / dmu;
  else
    f1[0] = va[1] * f[0] / dmu;
  if (va[M - 1] >= 0)
    f1[M - 1] = va[M - 1] * (-f[M - 2]) / dmu;
  else
    f1[M - 1] = va[M - 1] * (-f[M - 1]) / dmu;
}

/* The focusing operator of alg (1 upwind, 3 WENO3) on every energy row.
 * scratch: 3 M doubles. */
static void focus_operator(int alg, int E, int M, double dmu, double *f1, const double *f, const double *v_avg,
                           double *flux, double *scratch) {
  int e;
  if (alg != 3) {
    focus_operator_upwind(E, M, dmu, f1, f, v_avg, flux);
    return;
  }
  for (e = 0; e < E; e++)
    focus_row_weno3(M, dmu, f1 + e * M, f + e * M, v_avg + e * (M + 1), flux + e * (M + 1), scratch,
                    scratch + M, scratch + 2 * M);
}

/* The optimal third-order strong-stability-preserving Runge-Kutta step of
 * AdiabaticFocusing / AdiabaticChange for alg 2 and 3 (src/energeticParticles.c
 * :1677-1743, :952-1018): s1 = f - dt L(f); s1 = 3/4 f + 1/4 s1 - 1/4 dt L(s1);
 * f = 1/3 f + 2/3 s1 - 2/3 dt L(s1), in the reference's operation order. */
static void rk3_stage(int stage, int n, double dt, double *f, double *s1, const double *f1) {
  int i;
  if (stage == 0)
    for (i = 0; i < n; i++) s1[i] = f[i] - dt * f1[i];
  else if (stage == 1)
    for (i = 0; i < n; i++) s1[i] = (3.0 / 4.0) * f[i] + (1.0 / 4.0) * s1[i] - (1.0 / 4.0) * dt * f1[i];
  else
    for (i = 0; i < n; i++) f[i] = (1.0 / 3.0) * f[i] + (2.0 / 3.0) * s1[i] - (2.0 / 3.0) * dt * f1[i];
}

/* Test hooks: one application of the focusing operator (alg 1 upwind, 3
 * WENO3) to f [E][M] with half-mesh velocities v_avg [E][M+1], and of the
 * change operator to ln f [E][M] with v_avg [E+1][M] at radius rmag. */
void or_focus_operator(int alg, int E, int M, double dmu, const double *f, const double *v_avg, double *f1) {
  double *flux = (double *)malloc(sizeof(double) * (E * (M + 1) + 3 * M));
  focus_operator(alg, E, M, dmu, f1, f, v_avg, flux, flux + E * (M + 1));
  free(flux);
}

/* One node: returns N_subcycles (0 if vel_abs_max == 0). */
static int focus_node(const ep_params *p, const or_grid *g, double dt,
                      double ds_node, double bplus, double bminus,
                      double dlnn, double dlnb, double dupar, double *fn,
                      double *f, double *f1, double *v_avg, double *flux) {
  const int E = g->E, M = g->M;
  const double safety_factor = 0.9, half = 0.5, one = 1.0, two = 2.0, three = 3.0;
  double dt_full = dt * p->numEpSteps;
  double DuParDt = dupar / dt_full;
  double b = two * (dlnn / dt_full) - three * (dlnb / dt_full);
  double dlnBds, vel_abs_max = 0.0, dt_stable, dt_subcycle;
  int e, m, s, N;
  if ((bplus == 0.0) || (bminus == 0.0))
    dlnBds = 0.0;
  else
    dlnBds = (log(bplus) - log(bminus)) / (two * ds_node + DBL_MIN);
  for (e = 0; e < E; e++) {
    double a = -g->vgrid[e] * dlnBds - (two / g->vgrid[e]) * DuParDt;
    double vel_prev = 0.0;
    for (m = 0; m < M; m++) {
      double muval = g->mugrid[m];
      double vel = half * (one - muval * muval) * (a + muval * b);
      f[e * M + m] = fn[e * M + m];
      if (m > 0) {
        double va = half * (vel + vel_prev);
        v_avg[e * (M + 1) + m] = va;
        if (fabs(va) > vel_abs_max) vel_abs_max = fabs(va);
      }
      vel_prev = vel;
    }
    v_avg[e * (M + 1) + 0] = v_avg[e * (M + 1) + 1];
    v_avg[e * (M + 1) + M] = v_avg[e * (M + 1) + M - 1];
  }
  dt_stable = safety_factor * g->dmu / vel_abs_max;
  N = (int)ceil(dt / dt_stable);
  dt_subcycle = dt / N;
  if (p->adiabaticFocusAlg == 2 || p->adiabaticFocusAlg == 3) {
    /* s1 and the WENO3 scratch sit after the E (M + 1) flux entries */
    double *s1 = flux + E * (M + 1), *scr = s1 + E * M;
    for (s = 0; s < N; s++) {
      int st;
      for (st = 0; st < 3; st++) {
        focus_operator(p->adiabaticFocusAlg, E, M, g->dmu, f1, st == 0 ? f : s1, v_avg, flux, scr);
        rk3_stage(st, E * M, dt_subcycle, f, s1, f1);
      }
    }
  } else {
    for (s = 0; s < N; s++) {
      focus_operator_upwind(E, M, g->dmu, f1, f, v_avg, flux);
      for (e = 0; e < E * M; e++) f[e] = f[e] - dt_subcycle * f1[e];
    }
  }
  for (e = 0; e < E * M; e++) fn[e] = f[e];
  return N;
}

/* ------------------------------------------------------------------ */
/* AdiabaticChange + _Operator_Upwind                                  */
/*   (src/energeticParticles.c:805-1077, :1081-1176)                   */
/* ------------------------------------------------------------------ */
static void change_operator_upwind(const ep_params *p, const or_grid *g,
                                   double rmag, double *f1, const double *f,
                                   const double *v_avg, double *flux) {
  const int E = g->E, M = g->M;
  const double upwind = 1.0;
  const double dlnp_i = 1.0 / g->dlnp;
  double bc_val0, bc_val1;
  int e, m;
  /* v_avg / flux are [e][mu] with e in 0..E (idx_spep1m) */
  for (e = 1; e < E; e++)
    for (m = 0; m < M; m++) {
      double va = v_avg[e * M + m];
      double cce = copysign(upwind, va);
      flux[e * M + m] = va * 0.5 * ((1.0 - cce) * f[e * M + m] + (1.0 + cce) * f[(e - 1) * M + m]);
    }
  for (e = 1; e < E - 1; e++)
    for (m = 0; m < M; m++)
      f1[e * M + m] = (flux[(e + 1) * M + m] - flux[e * M + m]) * dlnp_i;
  bc_val0 = log(or_seed(p, g->egrid[0], rmag));
  bc_val1 = log(or_seed(p, g->egrid[E - 1], rmag));
  for (m = 0; m < M; m++) {
    double vl = v_avg[1 * M + m];
    double vr = v_avg[(E - 1) * M + m];
    if (vl <= 0)
      f1[0 * M + m] = vl * (f[1 * M + m] - f[0 * M + m]) * dlnp_i;
    else
      f1[0 * M + m] = vl * (f[0 * M + m] - bc_val0) * dlnp_i;
    if (vr >= 0)
      f1[(E - 1) * M + m] = vr * (f[(E - 1) * M + m] - f[(E - 2) * M + m]) * dlnp_i;
    else
      f1[(E - 1) * M + m] = vr * (bc_val1 - f[(E - 1) * M + m]) * dlnp_i;
  }
}

/* AdiabaticChange_Operator_WENO3 (src/energeticParticles.c:1179-1427) for
 * one mu column m (v_avg / flux [E+1][M], f / f1 [E][M]); needs E >= 4.
 * ln / lp / al are scratch columns of E. */
static void change_col_weno3(const ep_params *p, const or_grid *g, double rmag, int m, double *f1,
                             const double *f, const double *v_avg, double *flux, double *ln, double *lp,
                             double *al) {
  const int E = g->E, M = g->M;
  const double upwind = 1.0;
  int e;
  double cce;
  for (e = 0; e < E; e++) al[e] = weno3_alpha(v_avg + m, M, e, E);
  for (e = 0; e < E; e++) {
    lp[e] = 0.5 * f[e * M + m] * v_avg[(e + 1) * M + m] - al[e];
    ln[e] = 0.5 * f[e * M + m] * v_avg[e * M + m] + al[e];
  }
  for (e = 2; e < E - 1; e++)
    flux[e * M + m] = weno3_face(ln[e - 2], ln[e - 1], ln[e], lp[e - 1], lp[e], lp[e + 1]);
  e = 1;
  cce = copysign(upwind, v_avg[e * M + m]);
  flux[e * M + m] = v_avg[e * M + m] * 0.5 * ((1.0 - cce) * f[e * M + m] + (1.0 + cce) * f[(e - 1) * M + m]);
  e = E - 1;
  cce = copysign(upwind, v_avg[e * M + m]);
  flux[e * M + m] = v_avg[e * M + m] * 0.5 * ((1.0 - cce) * f[e * M + m] + (1.0 + cce) * f[(e - 1) * M + m]);
  for (e = 1; e < E - 1; e++) f1[e * M + m] = (flux[(e + 1) * M + m] - flux[e * M + m]) / g->dlnp;
  {
    double vl = v_avg[1 * M + m], vr = v_avg[(E - 1) * M + m];
    if (vl <= 0)
      f1[0 * M + m] = vl * (f[1 * M + m] - f[0 * M + m]) / g->dlnp;
    else
      f1[0 * M + m] = vl * (f[0 * M + m] - log(or_seed(p, g->egrid[0], rmag))) / g->dlnp;
    if (vr >= 0)
      f1[(E - 1) * M + m] = vr * (f[(E - 1) * M + m] - f[(E - 2) * M + m]) / g->dlnp;
    else
      f1[(E - 1) * M + m] = vr * (log(or_seed(p, g->egrid[E - 1], rmag)) - f[(E - 1) * M + m]) / g->dlnp;
  }
}

static void change_operator(const ep_params *p, const or_grid *g, double rmag, double *f1, const double *f,
                            const double *v_avg, double *flux, double *scratch) {
  int m;
  if (p->adiabaticChangeAlg != 3) {
    change_operator_upwind(p, g, rmag, f1, f, v_avg, flux);
    return;
  }
  for (m = 0; m < g->M; m++)
    change_col_weno3(p, g, rmag, m, f1, f, v_avg, flux, scratch, scratch + g->E, scratch + 2 * g->E);
}

void or_change_operator(const ep_params *p, int E, int M, double rmag, const double *f, const double *v_avg,
                        double *f1) {
  or_grid g;
  double *flux = (double *)malloc(sizeof(double) * ((E + 1) * M + 3 * E));
  grid_alloc(&g, p, E, M);
  change_operator(p, &g, rmag, f1, f, v_avg, flux, flux + (E + 1) * M);
  grid_free(&g);
  free(flux);
}

static int change_node(const ep_params *p, const or_grid *g, double dt,
                       double rmag, double dlnn, double dlnb, double dupar,
                       double *fn, double *f, double *f1, double *v_avg,
                       double *vel, double *flux) {
  const int E = g->E, M = g->M;
  const double safety_factor = 0.9, half = 0.5, one = 1.0;
  double dt_full = dt * p->numEpSteps;
  double DlnBDt = dlnb / dt_full;
  double DlnNDt = dlnn / dt_full;
  double cool_damp_fac = (DlnNDt < 0) ? 1.0 - p->dampCooling : 1.0;
  double DuParDt = dupar / dt_full;
  double vel_abs_max = 0.0, dt_stable, dt_subcycle;
  int e, m, s, N;
  for (e = 0; e < E; e++)
    for (m = 0; m < M; m++) {
      double muval = g->mugrid[m], a, b;
      f[e * M + m] = log(fn[e * M + m]);
      a = -muval * DuParDt;
      b = muval * muval * (DlnNDt - DlnBDt) + half * (one - muval * muval) * DlnBDt;
      vel[e * M + m] = cool_damp_fac * (a * g->vgrid_i[e] + b);
      if (fabs(vel[e * M + m]) > vel_abs_max) vel_abs_max = fabs(vel[e * M + m]);
    }
  for (e = 0; e < E - 1; e++)
    for (m = 0; m < M; m++)
      v_avg[(e + 1) * M + m] = half * (vel[(e + 1) * M + m] + vel[e * M + m]);
  for (m = 0; m < M; m++) {
    v_avg[0 * M + m] = v_avg[1 * M + m];
    v_avg[E * M + m] = v_avg[(E - 1) * M + m];
  }
  dt_stable = safety_factor * g->dlnp / vel_abs_max;
  N = (int)ceil(dt / dt_stable);
  dt_subcycle = dt / N;
  if (p->adiabaticChangeAlg == 2 || p->adiabaticChangeAlg == 3) {
    double *s1 = flux + (E + 1) * (M + 1), *scr = s1 + E * M;
    for (s = 0; s < N; s++) {
      int st;
      for (st = 0; st < 3; st++) {
        change_operator(p, g, rmag, f1, st == 0 ? f : s1, v_avg, flux, scr);
        rk3_stage(st, E * M, dt_subcycle, f, s1, f1);
      }
    }
  } else {
    for (s = 0; s < N; s++) {
      change_operator_upwind(p, g, rmag, f1, f, v_avg, flux);
      for (e = 0; e < E * M; e++) f[e] = f[e] - dt_subcycle * f1[e];
    }
  }
  for (e = 0; e < E * M; e++) fn[e] = exp(f[e]);
  return N;
}

/* ------------------------------------------------------------------ */
/* updateEnergeticParticles (src/energeticParticles.c:69-203)          */
/* ------------------------------------------------------------------ */
int or_update_ep(const ep_dims *d, const ep_params *p, const int *n_active,
                 const double *x, const double *y, const double *z,
                 const double *rmag, const double *bmag, const double *dlnb,
                 const double *dlnn, const double *dupar, double *f,
                 double *ds_out, double *bplus_out, double *bminus_out,
                 uint32_t op_flags, int nthreads, ep_diag *diag) {
  const int S = d->num_streams, N = d->num_nodes, E = d->num_energies, M = d->num_mu;
  const size_t SN = (size_t)S * N;
  const size_t cells_node = (size_t)E * M;
  or_grid g;
  or_err err;
  double *ds;
  /* the reference's WENO3 stencils read past their arrays below 4 cells */
  if ((p->adiabaticFocusAlg == 3 && M < 4) || (p->adiabaticChangeAlg == 3 && E < 4)) return EP_ERR_UNSUPPORTED;
  ds = (double *)malloc(sizeof(double) * SN);
  double *bplus = (double *)malloc(sizeof(double) * SN);
  double *bminus = (double *)malloc(sizeof(double) * SN);
  double dt = p->tDel / (1.0 * p->numEpSteps);
  int max_focus = 0, max_change = 0, max_nsteps = 0, step;
  double min_tau = DBL_MAX;
  const int do_stream = (p->useParallelDiffusion > 0) && (op_flags & EP_OP_STREAM);
  const int do_focus = (p->useAdiabaticFocus > 0) && (op_flags & EP_OP_FOCUS);
  const int do_change = (p->useAdiabaticChange > 0) && (op_flags & EP_OP_CHANGE);
  const int do_seed = (p->checkSeedPopulation > 0) && (op_flags & EP_OP_SEEDCHECK);
  long long cells = 0;
  int s;
  memset(&err, 0, sizeof(err));
  grid_alloc(&g, p, E, M);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (s = 0; s < S; s++) cells += (long long)(n_active ? n_active[s] : N) * E * M;

  for (step = 0; step < p->numEpSteps; step++) {
    /* ---- stream stage (per stream; the MPI gather/scatter of
       src/simCore.c:177-318 is a no-op on resident data) ---- */
#pragma omp parallel for schedule(dynamic, 1) reduction(max : max_nsteps)
    for (s = 0; s < S; s++) {
      const int n = n_active ? n_active[s] : N;
      const size_t o = (size_t)s * N;
      double *fs = f + o * cells_node;
      or_err lerr;
      memset(&lerr, 0, sizeof(lerr));
      update_stream_values(p, n, x + o, y + o, z + o, bmag + o, ds + o, bplus + o, bminus + o);
      if (do_seed) seed_population_check(p, &g, n, rmag + o, fs);
      if (do_stream) {
        int *list = (int *)malloc(sizeof(int) * n);
        int *ref = (int *)malloc(sizeof(int) * n);
        double *dsl = (double *)malloc(sizeof(double) * n);
        double *dsil = (double *)malloc(sizeof(double) * n);
        double dsMin;
        int size = or_get_stream_list(p, n, x + o, y + o, z + o, list, ref, dsl, dsil, &dsMin);
        int mx = 0;
        if (size == EP_ERR_DS_UNDERFLOW) {
          err_note(&lerr, EP_ERR_DS_UNDERFLOW, mkkey(step, 0, s, 0, 0, 0, 0), s, -1, -1, -1);
        } else if (size < 3 && (p->useEPBoundary > 0) && (p->useBoundaryFunction > 0)) {
          err_note(&lerr, EP_ERR_STREAM_SHORT, mkkey(step, 0, s, 0, 0, 0, 0), s, -1, -1, -1);
        } else {
          diffuse_stream(p, &g, s, step, n, rmag + o, bmag + o, list, ref, size, dsil, dsMin, dt,
                         fs, &mx, &lerr);
        }
        if (mx > max_nsteps) max_nsteps = mx;
        free(list);
        free(ref);
        free(dsl);
        free(dsil);
      }
      if (lerr.code != EP_OK) {
#pragma omp critical
        err_note(&err, lerr.code, lerr.key, lerr.stream, lerr.node, lerr.energy, lerr.mu);
      }
    }
    if (err.code != EP_OK) break;

    /* ---- node stage: focusing then change per stream (skip node 0,
       innerComputeShell, src/mpiInit.c:105-109) ---- */
#pragma omp parallel for schedule(dynamic, 1) reduction(max : max_focus, max_change) reduction(min : min_tau)
    for (s = 0; s < S; s++) {
      const int n = n_active ? n_active[s] : N;
      const size_t o = (size_t)s * N;
      double *fs = f + o * cells_node;
      double *ft = (double *)malloc(sizeof(double) * cells_node);
      double *f1 = (double *)malloc(sizeof(double) * cells_node);
      double *vel = (double *)malloc(sizeof(double) * cells_node);
      double *v_avg = (double *)malloc(sizeof(double) * (E + 1) * (M + 1));
      /* half-mesh fluxes, then the RK3 stage and the WENO3 scratch */
      double *flux = (double *)malloc(sizeof(double) * ((E + 1) * (M + 1) + E * M + 3 * (E + M + 2)));
      or_err lerr;
      int k, e, m;
      memset(&lerr, 0, sizeof(lerr));
      for (e = 0; e < E; e++)
        for (k = 1; k < n; k++) {
          double tau = or_mfp(p, g.rigidity[e], rmag[o + k], bmag[o + k]) * g.vgrid_i[e];
          if (tau < min_tau) min_tau = tau;
        }
      if (do_focus)
        for (k = 1; k < n; k++) {
          double *fn = fs + (size_t)k * cells_node;
          int nsub = focus_node(p, &g, dt, ds[o + k], bplus[o + k], bminus[o + k], dlnn[o + k],
                                dlnb[o + k], dupar[o + k], fn, ft, f1, v_avg, flux);
          if (nsub > max_focus) max_focus = nsub;
          for (e = 0; e < E; e++)
            for (m = 0; m < M; m++) {
              double *c = fn + e * M + m;
              if (isnan(*c))
                err_note(&lerr, EP_ERR_NAN, mkkey(step, 1, s, 0, k, e, m), s, k, e, m);
              else if (isinf(*c))
                err_note(&lerr, EP_ERR_INF, mkkey(step, 1, s, 0, k, e, m), s, k, e, m);
              if (*c < DBL_MIN) *c = DBL_MIN;
            }
        }
      if (do_change)
        for (k = 1; k < n; k++) {
          double *fn = fs + (size_t)k * cells_node;
          int nsub = change_node(p, &g, dt, rmag[o + k], dlnn[o + k], dlnb[o + k], dupar[o + k],
                                 fn, ft, f1, v_avg, vel, flux);
          if (nsub > max_change) max_change = nsub;
          for (e = 0; e < E; e++)
            for (m = 0; m < M; m++) {
              double *c = fn + e * M + m;
              if (isnan(*c))
                err_note(&lerr, EP_ERR_NAN, mkkey(step, 1, s, 1, k, e, m), s, k, e, m);
              else if (isinf(*c))
                err_note(&lerr, EP_ERR_INF, mkkey(step, 1, s, 1, k, e, m), s, k, e, m);
              if (*c < DBL_MIN) *c = DBL_MIN;
            }
        }
      if (lerr.code != EP_OK) {
#pragma omp critical
        err_note(&err, lerr.code, lerr.key, lerr.stream, lerr.node, lerr.energy, lerr.mu);
      }
      free(ft);
      free(f1);
      free(vel);
      free(v_avg);
      free(flux);
    }
    if (err.code != EP_OK) break;
  }
  if (ds_out) memcpy(ds_out, ds, sizeof(double) * SN);
  if (bplus_out) memcpy(bplus_out, bplus, sizeof(double) * SN);
  if (bminus_out) memcpy(bminus_out, bminus, sizeof(double) * SN);
  if (diag) {
    memset(diag, 0, sizeof(*diag));
    diag->max_subcycles_change = max_change;
    diag->max_subcycles_focusing = max_focus;
    diag->max_nsteps_stream = max_nsteps;
    diag->min_tau = min_tau;
    diag->err_code = err.code;
    diag->err_stream = err.code ? err.stream : -1;
    diag->err_node = err.code ? err.node : -1;
    diag->err_energy = err.code ? err.energy : -1;
    diag->err_mu = err.code ? err.mu : -1;
    diag->cell_updates = cells * p->numEpSteps;
  }
  free(ds);
  free(bplus);
  free(bminus);
  grid_free(&g);
  return err.code;
}

/* ------------------------------------------------------------------ */
/* MAS trilinear interpolation (src/masInterp.c)                       */
/* ------------------------------------------------------------------ */
double or_mas_bsearch(const float *A, double key, int *imin, int *imax,
                      int lower, int upper) {
  int imid;
  while (upper != (lower + 1)) {
    imid = (lower + upper) / 2;
    if (A[imid] < key)
      lower = imid;
    else
      upper = imid;
  }
  *imin = lower;
  *imax = upper;
  if (A[*imin] > key) return 0.0;
  if (A[*imax] < key) return 1.0;
  return ((key - A[lower]) / (A[upper] - A[lower]));
}

/* masInterpolate (src/masInterp.c:177-200). */
static double mas_interpolate(const float *V, int r0, int r1, int t0, int t1,
                              int p0, int p1, double rd, double td, double pd,
                              int nr, int nt) {
#define OR_VIDX(r, t, p) ((size_t)(r) + (size_t)nr * (t) + (size_t)nr * nt * (p))
  double c00 = V[OR_VIDX(r0, t0, p0)] * (1.0 - rd) + V[OR_VIDX(r1, t0, p0)] * rd;
  double c10 = V[OR_VIDX(r0, t0, p1)] * (1.0 - rd) + V[OR_VIDX(r1, t0, p1)] * rd;
  double c01 = V[OR_VIDX(r0, t1, p0)] * (1.0 - rd) + V[OR_VIDX(r1, t1, p0)] * rd;
  double c11 = V[OR_VIDX(r0, t1, p1)] * (1.0 - rd) + V[OR_VIDX(r1, t1, p1)] * rd;
  double c0 = c00 * (1.0 - pd) + c10 * pd;
  double c1 = c01 * (1.0 - pd) + c11 * pd;
#undef OR_VIDX
  return (c0 * (1.0 - td) + c1 * td);
}

/* One field on one time level at (rr [Rs], th, ph). */
static double mas_field(const float *V, const float *const meshes[21],
                        const int *dims, int k, double rr, double th,
                        double ph) {
  int r0, r1, t0, t1, p0, p1;
  const int nr = dims[3 * k], nt = dims[3 * k + 1], np = dims[3 * k + 2];
  double rd = or_mas_bsearch(meshes[3 * k + 0], rr, &r0, &r1, 0, nr - 1);
  double td = or_mas_bsearch(meshes[3 * k + 1], th, &t0, &t1, 0, nt - 1);
  double pd = or_mas_bsearch(meshes[3 * k + 2], ph, &p0, &p1, 0, np - 1);
  (void)np;
  return mas_interpolate(V, r0, r1, t0, t1, p0, p1, rd, td, pd, nr, nt);
}

/* masGetNode (src/masInterp.c:31-68) at a position (x, y, z) in rScale
   units: masTriLinear inside mas_radial_max, else masWind for the node
   whose state is given (src/masInterp.c:1007-1065, no fake co-rotation),
   which scales that node's previous values by |rOld| / node_rmag whatever
   the position asked.  prev7 = {br, bt, bp, vr, vt, vp, density} of the
   node; out7 the same order. */
static void mas_get_node(double rScale, double x, double y, double z,
                         const float *const fields0[7], const float *const meshes0[21],
                         const int *dims0, const float *const fields1[7],
                         const float *const meshes1[21], const int *dims1, double s,
                         double phi_offset, double mas_radial_max, double mas_vmin, double rox,
                         double roy, double roz, double node_rmag, const double *prev7,
                         double *out7) {
  /* cartToSphPosAu (src/geometry.c:131-160, :35-45) */
  double ax = x * rScale, ay = y * rScale, az = z * rScale;
  double pr = sqrt(ax * ax + ay * ay + az * az);
  double pth = acos(az / pr);
  double pph = atan2(ay, ax);
  while (pph < 0.0) pph += (2.0 * OR_PI);
  if (pph < 0.0)
    while (pph < 0.0) pph += 2 * OR_PI;
  if (pph > 2 * OR_PI)
    while (pph > 2 * OR_PI) pph -= 2 * OR_PI;
  if (pr <= (mas_radial_max * OR_RSAU)) {
    /* masTriLinear (src/masInterp.c:539-741) */
    double rr = pr / OR_RSAU, th = pth, ph = pph - phi_offset;
    double v;
    if (ph < 0.0)
      while (ph < 0.0) ph += (2.0 * OR_PI);
    if (ph > (2.0 * OR_PI))
      while (ph > (2.0 * OR_PI)) ph -= (2.0 * OR_PI);
    out7[2] = ((1.0 - s) * mas_field(fields0[0], meshes0, dims0, 0, rr, th, ph) +
               s * mas_field(fields1[0], meshes1, dims1, 0, rr, th, ph)) * OR_MAS_B_CONVERT;
    out7[1] = ((1.0 - s) * mas_field(fields0[1], meshes0, dims0, 1, rr, th, ph) +
               s * mas_field(fields1[1], meshes1, dims1, 1, rr, th, ph)) * OR_MAS_B_CONVERT;
    out7[0] = ((1.0 - s) * mas_field(fields0[2], meshes0, dims0, 2, rr, th, ph) +
               s * mas_field(fields1[2], meshes1, dims1, 2, rr, th, ph)) * OR_MAS_B_CONVERT;
    out7[5] = ((1.0 - s) * mas_field(fields0[3], meshes0, dims0, 3, rr, th, ph) +
               s * mas_field(fields1[3], meshes1, dims1, 3, rr, th, ph)) * OR_MAS_V_CONVERT;
    out7[4] = ((1.0 - s) * mas_field(fields0[4], meshes0, dims0, 4, rr, th, ph) +
               s * mas_field(fields1[4], meshes1, dims1, 4, rr, th, ph)) * OR_MAS_V_CONVERT;
    v = ((1.0 - s) * mas_field(fields0[5], meshes0, dims0, 5, rr, th, ph) +
         s * mas_field(fields1[5], meshes1, dims1, 5, rr, th, ph)) * OR_MAS_V_CONVERT;
    if (v < (mas_vmin / OR_C)) v = (mas_vmin / OR_C);
    out7[3] = v;
    out7[6] = ((1.0 - s) * mas_field(fields0[6], meshes0, dims0, 6, rr, th, ph) +
               s * mas_field(fields1[6], meshes1, dims1, 6, rr, th, ph)) * OR_MAS_RHO_CONVERT;
  } else {
    double rOldmag = sqrt(rox * rox + roy * roy + roz * roz);
    double oneOverR = rOldmag / node_rmag;
    out7[0] = prev7[0] * oneOverR * oneOverR;
    out7[1] = prev7[1] * oneOverR;
    out7[2] = prev7[2] * oneOverR;
    out7[3] = prev7[3];
    out7[4] = prev7[4] * oneOverR;
    out7[5] = prev7[5] * oneOverR;
    out7[6] = prev7[6] * oneOverR * oneOverR;
  }
}

void or_mas_interp(int n, double rScale, const double *x, const double *y,
                   const double *z, const double *rmag,
                   const float *const fields0[7], const float *const meshes0[21],
                   const int *dims0, const float *const fields1[7],
                   const float *const meshes1[21], const int *dims1,
                   double s, double phi_offset, double mas_radial_max,
                   double mas_vmin, const double *rold_x,
                   const double *rold_y, const double *rold_z,
                   const double *prev_br, const double *prev_bt,
                   const double *prev_bp, const double *prev_vr,
                   const double *prev_vt, const double *prev_vp,
                   const double *prev_d, double *br, double *bt, double *bp,
                   double *vr, double *vt, double *vp, double *dens) {
  int i;
  for (i = 0; i < n; i++) {
    const double prev7[7] = {prev_br[i], prev_bt[i], prev_bp[i], prev_vr[i],
                             prev_vt[i], prev_vp[i], prev_d[i]};
    double o[7];
    mas_get_node(rScale, x[i], y[i], z[i], fields0, meshes0, dims0, fields1, meshes1, dims1, s,
                 phi_offset, mas_radial_max, mas_vmin, rold_x[i], rold_y[i], rold_z[i], rmag[i],
                 prev7, o);
    br[i] = o[0];
    bt[i] = o[1];
    bp[i] = o[2];
    vr[i] = o[3];
    vt[i] = o[4];
    vp[i] = o[5];
    dens[i] = o[6];
  }
}

/* ------------------------------------------------------------------ */
/* updateMhd (src/flow.c:33-154)                                       */
/* ------------------------------------------------------------------ */
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
                   double *dlnb, double *dlnn) {
  int i;
  if (mode == 1)
    or_mas_interp(n, rScale, x, y, z, rmag, fields0, meshes0, dims0, fields1, meshes1, dims1,
                  s_cor, phi_offset, mas_radial_max, mas_vmin, rold_x, rold_y, rold_z, old_br,
                  old_bt, old_bp, old_vr, old_vt, old_vp, old_d, br, bt, bp, vr, vt, vp, dens);
  for (i = 0; i < n; i++) {
    const double bmag_old = sqrt(old_br[i] * old_br[i] + old_bt[i] * old_bt[i] +
                                 old_bp[i] * old_bp[i]);
    if (mode != 1) {
      /* mhdDensity, mhdV, mhdB with MHD_DEFAULT (src/flow.c:241-350) */
      double rr = rmag[i] * rScale;
      double theta = acos(z[i] / rmag[i]);
      dens[i] = mhdNsAu / (rr * rr);
      vr[i] = mhdUs;
      vt[i] = 0.0;
      vp[i] = 0.0;
      br[i] = mhdBsAu / (rr * rr);
      bt[i] = 0.0;
      bp[i] = -1.0 * rr * br[i] * (omegaSun / (vr[i] + OR_VERYSMALL)) * sin(theta);
    }
    bmag[i] = sqrt(br[i] * br[i] + bt[i] * bt[i] + bp[i] * bp[i]);
    if (is_first[i]) {
      dupar[i] = 0.0;
      dlnb[i] = 0.0;
      dlnn[i] = 0.0;
    } else {
      dupar[i] = (vr[i] - old_vr[i]) * br[i] + (vt[i] - old_vt[i]) * bt[i] +
                 (vp[i] - old_vp[i]) * bp[i];
      dupar[i] /= bmag[i];
      dlnb[i] = log(bmag[i] / bmag_old);
      dlnn[i] = log(dens[i] / old_d[i]);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Point observers and flux                                            */
/* ------------------------------------------------------------------ */
static double dot3(double ax, double ay, double az, double bx, double by, double bz) {
  return ax * bx + ay * by + az * bz;
}

/* findIntersection (src/search.c:234-268) */
static double find_intersection(const double *x0, const double *x1, double r) {
  double dx = x1[0] - x0[0], dy = x1[1] - x0[1], dz = x1[2] - x0[2];
  double r0square = x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2];
  double dx2 = dot3(dx, dy, dz, dx, dy, dz);
  double b = dot3(x0[0], x0[1], x0[2], dx, dy, dz) / dx2;
  double c = (r * r - r0square) / dx2;
  return -b + sqrt(b * b + c);
}

void or_observers(const ep_dims *d, const int *n_active, double rScale, double idw_p, int nobs,
                  const double *obs_r, const double *obs_theta, const double *obs_phi,
                  const double *x, const double *y, const double *z, const double *rmag,
                  const double *f, const double *egrid, double *dist_out, double *flux_out) {
  const int S = d->num_streams, N = d->num_nodes, E = d->num_energies, M = d->num_mu;
  const size_t EM = (size_t)E * M;
  double *proj = (double *)malloc(sizeof(double) * EM);
  int o, s, e, m;
  for (o = 0; o < nobs; o++) {
    double *td = dist_out + (size_t)o * EM;
    double weightSum = 0.0, rs, th, ph, cx, cy, cz;
    for (e = 0; e < (int)EM; e++) td[e] = 0.0;
    /* observer position: sphToCartPos of (r/rScale, theta, phi) after the
       phi wrap (src/unifiedOutput.c:832-851, src/geometry.c:53-60) */
    th = obs_theta[o];
    ph = obs_phi[o];
    if (ph < 0.0) ph += 2.0 * OR_PI;
    if (ph > (2.0 * OR_PI)) ph -= 2.0 * OR_PI;
    rs = obs_r[o] / rScale;
    cx = rs * sin(th) * cos(ph);
    cy = rs * sin(th) * sin(ph);
    cz = rs * cos(th);
    for (s = 0; s < S; s++) {
      const int n = n_active ? n_active[s] : N;
      const size_t on = (size_t)s * N;
      const double rSph = obs_r[o] / rScale;
      int ss0 = 0, ss1 = n - 1, in = ss0, out = ss1;
      double sfac, px, py, pz, distance, weight;
      double x0[3], x1[3];
      while ((rmag[on + ((in < ss1) ? (in + 1) : ss1)] < rSph) && (in < ss1)) in++;
      while ((rmag[on + ((out > ss0) ? (out - 1) : ss0)] >= rSph) && (out > ss0)) out--;
      x0[0] = x[on + in];
      x0[1] = y[on + in];
      x0[2] = z[on + in];
      x1[0] = x[on + out];
      x1[1] = y[on + out];
      x1[2] = z[on + out];
      sfac = (out > in) ? find_intersection(x0, x1, rSph) : 0.0;
      px = x0[0] * (1.0 - sfac) + x1[0] * sfac;
      py = x0[1] * (1.0 - sfac) + x1[1] * sfac;
      pz = x0[2] * (1.0 - sfac) + x1[2] * sfac;
      for (e = 0; e < E; e++)
        for (m = 0; m < M; m++) {
          const double fi = f[((on + in) * E + e) * M + m];
          const double fo = f[((on + out) * E + e) * M + m];
          if ((sfac >= 0.0) && (sfac <= 1.0))
            proj[e * M + m] = fi * (1.0 - sfac) + fo * sfac;
          else if (sfac < 0.0)
            proj[e * M + m] = fi;
          else
            proj[e * M + m] = fo;
        }
      distance = sqrt((cx - px) * (cx - px) + (cy - py) * (cy - py) + (cz - pz) * (cz - pz));
      weight = pow(distance, -1.0 * idw_p);
      weightSum += weight;
      for (e = 0; e < (int)EM; e++) td[e] += (weight * proj[e]);
    }
    for (e = 0; e < (int)EM; e++) td[e] /= weightSum;
    for (e = 0; e < E; e++) {
      double iso = 0.0;
      for (m = 0; m < M; m++) iso += td[e * M + m];
      iso /= M;
      flux_out[(size_t)o * E + e] = 2.0 * egrid[e] * iso;
    }
  }
  free(proj);
}

void or_stream_flux(const ep_dims *d, const double *f, const double *egrid, double *flux) {
  const size_t SN = (size_t)d->num_streams * d->num_nodes;
  const int E = d->num_energies, M = d->num_mu;
  size_t i;
  int e, m;
  for (i = 0; i < SN; i++)
    for (e = 0; e < E; e++) {
      double iso = 0.0;
      for (m = 0; m < M; m++) iso += f[(i * E + e) * M + m];
      iso /= M;
      flux[i * E + e] = 2.0 * egrid[e] * iso;
    }
}

/* ------------------------------------------------------------------ */
/* Parker-wind host pipeline (src/eprem.c:426-744, masCouple = 0)      */
/* ------------------------------------------------------------------ */
typedef struct {
  /* [stream][shell], shell 0 is the spawn template (INNER_SHELL of
     rank 0); shells 1..N are the active nodes (global node = shell-1) */
  double *x, *y, *z, *rmag, *zen, *azi;
  double *rox, *roy, *roz;
  double *dens, *densOld, *br, *bt, *bp, *bmag, *bmagOld;
  double *vr, *vt, *vp, *vrOld, *vtOld, *vpOld;
  double *vvx, *vvy, *vvz;
  double *dlnb, *dlnn, *dupar;
  double *f; /* [stream][shell][E][M] */
} or_state;

#define OR_NF 26

static void state_alloc(or_state *st, size_t nodes, size_t cells) {
  double **fld[OR_NF] = {&st->x, &st->y, &st->z, &st->rmag, &st->zen, &st->azi,
                         &st->rox, &st->roy, &st->roz, &st->dens, &st->densOld,
                         &st->br, &st->bt, &st->bp, &st->bmag, &st->bmagOld,
                         &st->vr, &st->vt, &st->vp, &st->vrOld, &st->vtOld,
                         &st->vpOld, &st->vvx, &st->vvy, &st->vvz, &st->dlnb};
  int i;
  for (i = 0; i < OR_NF; i++) *fld[i] = (double *)calloc(nodes, sizeof(double));
  st->dlnn = (double *)calloc(nodes, sizeof(double));
  st->dupar = (double *)calloc(nodes, sizeof(double));
  st->f = (double *)calloc(cells, sizeof(double));
}

static void state_free(or_state *st) {
  double *fld[OR_NF + 3] = {st->x, st->y, st->z, st->rmag, st->zen, st->azi,
                            st->rox, st->roy, st->roz, st->dens, st->densOld,
                            st->br, st->bt, st->bp, st->bmag, st->bmagOld,
                            st->vr, st->vt, st->vp, st->vrOld, st->vtOld,
                            st->vpOld, st->vvx, st->vvy, st->vvz, st->dlnb,
                            st->dlnn, st->dupar, st->f};
  int i;
  for (i = 0; i < OR_NF + 3; i++) free(fld[i]);
}

/* Copy one node's whole record (rippleShellsOut memcpy, src/simCore.c:701-728). */
static void node_copy(or_state *st, size_t dst, size_t src, size_t cells_node) {
  double *fld[OR_NF + 2] = {st->x, st->y, st->z, st->rmag, st->zen, st->azi,
                            st->rox, st->roy, st->roz, st->dens, st->densOld,
                            st->br, st->bt, st->bp, st->bmag, st->bmagOld,
                            st->vr, st->vt, st->vp, st->vrOld, st->vtOld,
                            st->vpOld, st->vvx, st->vvy, st->vvz, st->dlnb,
                            st->dlnn, st->dupar};
  int i;
  for (i = 0; i < OR_NF + 2; i++) fld[i][dst] = fld[i][src];
  memcpy(st->f + dst * cells_node, st->f + src * cells_node, sizeof(double) * cells_node);
}

/* moveNodes with delrFlow MHD_DEFAULT (src/simCore.c:736-785,
   src/flow.c:169-178) for n nodes: the flow velocity (vr, vt, vp) made
   Cartesian at the node (src/flow.c:355-403), old position saved, node
   moved by tDel * V / rScale. */
void or_move_nodes(int n, double tDel, double rScale, double *x, double *y, double *z, double *rmag,
                   double *rox, double *roy, double *roz, const double *vr, const double *vt,
                   const double *vp) {
  int k;
  for (k = 0; k < n; k++) {
    const double theta = acos(z[k] / rmag[k]);
    const double phi = atan2(y[k], x[k]);
    const double vvx = vr[k] * sin(theta) * cos(phi) + vt[k] * cos(theta) * cos(phi) - vp[k] * sin(phi);
    const double vvy = vr[k] * sin(theta) * sin(phi) + vt[k] * cos(theta) * sin(phi) + vp[k] * cos(phi);
    const double vvz = vr[k] * cos(theta) - vt[k] * sin(theta);
    double dx = tDel * vvx / rScale;
    double dy = tDel * vvy / rScale;
    double dz = tDel * vvz / rScale;
    double rx = x[k], ry = y[k], rz = z[k];
    rox[k] = rx;
    roy[k] = ry;
    roz[k] = rz;
    rx += dx;
    ry += dy;
    rz += dz;
    x[k] = rx;
    y[k] = ry;
    z[k] = rz;
    rmag[k] = sqrt((rx * rx) + (ry * ry) + (rz * rz));
  }
}

int or_run_parker(const or_sim_config *c, or_sim_output *out, int nthreads) {
  const int R = c->rows, C = c->cols, N = c->num_nodes, E = c->num_energies, M = c->num_mu;
  const int S = 6 * R * C, L = N + 1; /* shells per stream incl. template */
  const size_t cells_node = (size_t)E * M;
  or_state st;
  ep_params p;
  or_grid g;
  double tDel, t_global, t_init, simStartTimeDay, simStopTimeDay, azi_sun;
  double mhdUs, mhdNsAu, mhdBsAu, omega, unifiedOutputTime, epCalcStartTime;
  int num_loops = 0, rciter = 0, simStarted = 0, epInit = 0, rc = EP_OK;
  int face, row, col, s, k, e, m;
  double *ax, *ay, *az, *arm, *abm, *adb, *adn, *adu, *af;

  memset(&p, 0, sizeof(p));
  p.rScale = c->rScale;
  p.numEpSteps = c->numEpSteps;
  p.mfpType = c->mfpType;
  p.lamo = c->lamo;
  p.mfpRadialPower = c->mfpRadialPower;
  p.mfpBPower = c->mfpBPower;
  p.rigidityPower = c->rigidityPower;
  p.eMin = c->eMin;
  p.eMax = c->eMax;
  p.mass = c->mass;
  p.charge = c->charge;
  p.dsh_min = c->dsh_min;
  p.dampCooling = c->dampCooling;
  p.boundaryFunctAmplitude = c->boundaryFunctAmplitude;
  p.boundaryFunctXi = c->boundaryFunctXi;
  p.boundaryFunctGamma = c->boundaryFunctGamma;
  p.boundaryFunctBeta = c->boundaryFunctBeta;
  p.boundaryFunctEcutoff = c->boundaryFunctEcutoff;
  p.boundaryFunctR0 = c->boundaryFunctR0;
  p.useParallelDiffusion = c->useParallelDiffusion;
  p.useAdiabaticFocus = c->useAdiabaticFocus;
  p.useAdiabaticChange = c->useAdiabaticChange;
  p.checkSeedPopulation = c->checkSeedPopulation;
  p.useEPBoundary = c->useEPBoundary;
  p.useBoundaryFunction = c->useBoundaryFunction;
  p.adiabaticFocusAlg = 1;
  p.adiabaticChangeAlg = 1;

  /* setRuntimeConstants (src/configuration.c:324-348) */
  simStartTimeDay = c->simStartTime / OR_DAY;
  simStopTimeDay = c->simStopTime / OR_DAY;
  tDel = c->tDel / OR_DAY;
  p.tDel = tDel;
  mhdUs = c->flowMag / OR_C;
  mhdNsAu = c->mhdDensityAu / OR_MHD_DENSITY_NORM;
  mhdBsAu = c->mhdBAu / OR_MHD_B_NORM;
  p.mhdBsAu = mhdBsAu;
  omega = c->omegaSun;
  /* timeInitialization, non-MAS branch (src/simCore.c:160-165): the
     output time uses tDel after the /DAY conversion */
  unifiedOutputTime = (c->unifiedOutputTime >= 0.0) ? c->unifiedOutputTime
                                                    : c->simStartTime + 0.5 * tDel * OR_DAY;
  epCalcStartTime = (c->epCalcStartTime >= 0.0) ? c->epCalcStartTime : c->simStartTime;

  grid_alloc(&g, &p, E, M);
  state_alloc(&st, (size_t)S * L, (size_t)S * L * cells_node);
  ax = (double *)malloc(sizeof(double) * S * N);
  ay = (double *)malloc(sizeof(double) * S * N);
  az = (double *)malloc(sizeof(double) * S * N);
  arm = (double *)malloc(sizeof(double) * S * N);
  abm = (double *)malloc(sizeof(double) * S * N);
  adb = (double *)malloc(sizeof(double) * S * N);
  adn = (double *)malloc(sizeof(double) * S * N);
  adu = (double *)malloc(sizeof(double) * S * N);
  af = (double *)malloc(sizeof(double) * S * N * cells_node);

  /* simCoreInit (src/simCore.c:56-75) */
  azi_sun = c->aziSunStart;
  t_init = N * tDel;
  t_global = simStartTimeDay - t_init;

  /* gridStructInit: initCubeCoords, initSphereCoords, initCopyAll,
     initInnerShellValues (src/cubeShellInit.c:46-91, :529-822,
     :1016-1090, :1164-1236).  Stream s = idx_frc(face,row,col). */
  {
    double cw = 1.0 / C, chh = 1.0 / R;
    for (face = 0; face < 6; face++)
      for (row = 0; row < R; row++)
        for (col = 0; col < C; col++) {
          double cx = 0, cy = 0, cz = 0, rm, zen, azi;
          /* The reference walks each face accumulating the coordinate by
             repeated +/- cellWidth (cellHeight); u runs from -0.5 upward
             with the column, u2 from +0.5 downward, w from +0.5 downward
             with the row.  Reproduce that accumulation exactly. */
          double u = -0.5, u2 = 0.5, w = 0.5;
          size_t i0;
          int j;
          u += 0.5 * cw;
          u2 -= 0.5 * cw;
          w -= 0.5 * chh;
          for (j = 0; j < col; j++) {
            u += cw;
            u2 -= cw;
          }
          for (j = 0; j < row; j++) w -= chh;
          switch (face) {
            case 0: cx = u; cy = w; cz = 0.5; break;    /* TOP    */
            case 1: cx = u2; cy = w; cz = -0.5; break;  /* BOTTOM */
            case 2: cx = 0.5; cy = u; cz = w; break;    /* FRONT  */
            case 3: cx = -0.5; cy = u2; cz = w; break;  /* BACK   */
            case 4: cx = u2; cy = 0.5; cz = w; break;   /* LEFT   */
            default: cx = u; cy = -0.5; cz = w; break;  /* RIGHT  */
          }
          s = (face * R + row) * C + col;
          rm = sqrt((cx * cx) + (cy * cy) + (cz * cz));
          cx = cx / rm;
          cy = cy / rm;
          cz = cz / rm;
          rm = 1.0;
          azi = atan2(cy, cx);
          zen = acos(cz / rm);
          i0 = (size_t)s * L;
          st.rmag[i0] = 1.0;
          st.azi[i0] = azi;
          st.zen[i0] = zen;
          st.z[i0] = st.rmag[i0] * cos(zen);
          st.x[i0] = st.rmag[i0] * sin(zen) * cos(azi);
          st.y[i0] = st.rmag[i0] * sin(zen) * sin(azi);
          for (k = 1; k < L; k++) {
            size_t i = i0 + k;
            st.rmag[i] = 1.0 + 0 + k;
            st.azi[i] = azi;
            st.zen[i] = zen;
            st.z[i] = st.rmag[i] * cos(zen);
            st.x[i] = st.rmag[i] * sin(zen) * cos(azi);
            st.y[i] = st.rmag[i] * sin(zen) * sin(azi);
          }
          for (k = 0; k < L; k++) {
            size_t i = i0 + k;
            for (e = 0; e < (int)cells_node; e++) st.f[i * cells_node + e] = DBL_MIN;
          }
        }
  }

  out->num_dumps = 0;
  out->num_steps = 0;
  if (out->egrid) memcpy(out->egrid, g.egrid, sizeof(double) * E);
  if (out->vgrid) memcpy(out->vgrid, g.vgrid, sizeof(double) * E);
  if (out->mu) memcpy(out->mu, g.mugrid, sizeof(double) * M);

  do {
    rciter += 1;
    if (rciter == N + 1 && simStarted == 0) {
      simStarted = 1;
      /* initEnergeticParticles (src/energeticParticlesInit.c:106-147) */
      if (p.useBoundaryFunction > 0 && c->boundaryFunctionInitDomain > 0)
        for (s = 0; s < S; s++)
          for (k = 0; k < L; k++)
            for (e = 0; e < E; e++) {
              double sd = or_seed(&p, g.egrid[e], st.rmag[(size_t)s * L + k]);
              for (m = 0; m < M; m++) st.f[((size_t)s * L + k) * cells_node + e * M + m] = sd;
            }
      t_global = simStartTimeDay;
    }
    if (t_global * OR_DAY >= epCalcStartTime && simStarted == 1) simStarted = 2;

    /* dataDumpIO -> writeObserverDataNetCDF (src/unifiedOutput.c:442-549,
       :1479-1530) */
    if ((c->dumpFreq > 0) && (num_loops % c->dumpFreq) == 0 &&
        (t_global * OR_DAY >= unifiedOutputTime) && out->num_dumps < out->max_dumps) {
      int dd = out->num_dumps;
      out->time[dd] = out->raw ? t_global : t_global * OR_DAY;
      for (s = 0; s < S; s++)
        for (k = 0; k < N; k++) {
          size_t i = (size_t)s * L + k + 1;
          size_t o = ((size_t)dd * S + s) * N + k;
          double zen = acos(st.z[i] / st.rmag[i]);
          double azi = atan2(st.y[i], st.x[i]);
          if (azi < 0.0) azi += 2.0 * OR_PI;
          if (azi > (2.0 * OR_PI)) azi -= 2.0 * OR_PI;
          out->R[o] = out->raw ? st.rmag[i] : st.rmag[i] * c->rScale;
          out->T[o] = zen;
          out->P[o] = azi;
          out->Br[o] = out->raw ? st.br[i] : st.br[i] * (OR_MHD_B_NORM * 1.0e5);
          out->Bt[o] = out->raw ? st.bt[i] : st.bt[i] * (OR_MHD_B_NORM * 1.0e5);
          out->Bp[o] = out->raw ? st.bp[i] : st.bp[i] * (OR_MHD_B_NORM * 1.0e5);
          out->Vr[o] = out->raw ? st.vr[i] : st.vr[i] * (OR_C / 1.0e5);
          out->Vt[o] = out->raw ? st.vt[i] : st.vt[i] * (OR_C / 1.0e5);
          out->Vp[o] = out->raw ? st.vp[i] : st.vp[i] * (OR_C / 1.0e5);
          out->Rho[o] = out->raw ? st.dens[i] : st.dens[i] * OR_MHD_DENSITY_NORM;
          for (e = 0; e < (int)cells_node; e++)
            out->dist[o * cells_node + e] =
                out->raw ? st.f[i * cells_node + e] : st.f[i * cells_node + e] * (1.0 / 27.0);
        }
      out->num_dumps++;
    }

    /* rotSunAndSpawnShell (src/simCore.c:442-618): single rank, no ring */
    azi_sun += (omega * tDel);
    if (azi_sun > OR_PI) azi_sun -= OR_TWO_PI;
    for (s = 0; s < S; s++) {
      size_t i0 = (size_t)s * L;
      for (k = L - 1; k > 0; k--) node_copy(&st, i0 + k, i0 + k - 1, cells_node);
      /* spawnNewShell (src/simCore.c:629-687) */
      {
        size_t i = i0 + 1;
        double azi = st.azi[i], zen = st.zen[i], rm = st.rmag[i];
        azi += azi_sun;
        if (azi > OR_PI) azi -= OR_TWO_PI;
        st.zen[i] = zen;
        st.azi[i] = azi;
        st.x[i] = rm * sin(zen) * cos(azi);
        st.y[i] = rm * sin(zen) * sin(azi);
        st.z[i] = rm * cos(zen);
      }
    }

    /* moveNodes (src/simCore.c:736-785), delrFlow MHD_DEFAULT
       (src/flow.c:169-178): active shells 2..N move */
    for (s = 0; s < S; s++) {
      size_t i = (size_t)s * L + 2;
      or_move_nodes(L - 2, tDel, c->rScale, st.x + i, st.y + i, st.z + i, st.rmag + i, st.rox + i,
                    st.roy + i, st.roz + i, st.vr + i, st.vt + i, st.vp + i);
    }

    /* updateMhd (src/flow.c:33-154) with the Parker model, through
       or_update_mhd (moveNodes makes V Cartesian at the same position) */
    for (s = 0; s < S; s++) {
      const size_t i0 = (size_t)s * L + 1;
      const int na = L - 1;
      double *ob = (double *)malloc(sizeof(double) * 7 * na);
      int *first = (int *)calloc(na, sizeof(int));
      first[0] = 1;
      memcpy(ob + 0 * na, st.br + i0, sizeof(double) * na);
      memcpy(ob + 1 * na, st.bt + i0, sizeof(double) * na);
      memcpy(ob + 2 * na, st.bp + i0, sizeof(double) * na);
      memcpy(ob + 3 * na, st.vr + i0, sizeof(double) * na);
      memcpy(ob + 4 * na, st.vt + i0, sizeof(double) * na);
      memcpy(ob + 5 * na, st.vp + i0, sizeof(double) * na);
      memcpy(ob + 6 * na, st.dens + i0, sizeof(double) * na);
      or_update_mhd(na, 0, c->rScale, mhdUs, mhdNsAu, mhdBsAu, omega, st.x + i0, st.y + i0,
                    st.z + i0, st.rmag + i0, first, NULL, NULL, NULL, NULL, NULL, NULL, 0.0, 0.0,
                    0.0, 0.0, NULL, NULL, NULL, ob, ob + na, ob + 2 * na, ob + 3 * na,
                    ob + 4 * na, ob + 5 * na, ob + 6 * na, st.br + i0, st.bt + i0, st.bp + i0,
                    st.vr + i0, st.vt + i0, st.vp + i0, st.dens + i0, st.bmag + i0,
                    st.dupar + i0, st.dlnb + i0, st.dlnn + i0);
      free(ob);
      free(first);
    }

    /* particle computation (src/eprem.c:653-682) */
    if (t_global * OR_DAY >= epCalcStartTime) {
      ep_diag dg;
      if (epInit == 0) {
        epInit = 1;
        if (p.useBoundaryFunction > 0 && c->boundaryFunctionInitDomain > 0)
          for (s = 0; s < S; s++)
            for (k = 0; k < L; k++)
              for (e = 0; e < E; e++) {
                double sd = or_seed(&p, g.egrid[e], st.rmag[(size_t)s * L + k]);
                for (m = 0; m < M; m++) st.f[((size_t)s * L + k) * cells_node + e * M + m] = sd;
              }
      }
      /* pack the active shells as resident streams */
      for (s = 0; s < S; s++)
        for (k = 0; k < N; k++) {
          size_t i = (size_t)s * L + k + 1, o = (size_t)s * N + k;
          ax[o] = st.x[i];
          ay[o] = st.y[i];
          az[o] = st.z[i];
          arm[o] = st.rmag[i];
          abm[o] = st.bmag[i];
          adb[o] = st.dlnb[i];
          adn[o] = st.dlnn[i];
          adu[o] = st.dupar[i];
          memcpy(af + o * cells_node, st.f + i * cells_node, sizeof(double) * cells_node);
        }
      {
        ep_dims d;
        d.num_streams = S;
        d.num_nodes = N;
        d.num_energies = E;
        d.num_mu = M;
        rc = or_update_ep(&d, &p, NULL, ax, ay, az, arm, abm, adb, adn, adu, af, NULL, NULL,
                          NULL, EP_OP_ALL, nthreads, &dg);
      }
      for (s = 0; s < S; s++)
        for (k = 0; k < N; k++) {
          size_t i = (size_t)s * L + k + 1, o = (size_t)s * N + k;
          memcpy(st.f + i * cells_node, af + o * cells_node, sizeof(double) * cells_node);
        }
      if (c->seedFunctionTest > 0 && p.useBoundaryFunction > 0 && c->boundaryFunctionInitDomain > 0)
        for (s = 0; s < S; s++)
          for (k = 0; k < L; k++)
            for (e = 0; e < E; e++) {
              double sd = or_seed(&p, g.egrid[e], st.rmag[(size_t)s * L + k]);
              for (m = 0; m < M; m++) st.f[((size_t)s * L + k) * cells_node + e * M + m] = sd;
            }
      if (out->step_diag && out->num_steps < out->max_steps) out->step_diag[out->num_steps++] = dg;
      if (rc != EP_OK) break;
    }
    t_global += tDel;
    num_loops++;
  } while (t_global <= simStopTimeDay);

  free(ax);
  free(ay);
  free(az);
  free(arm);
  free(abm);
  free(adb);
  free(adn);
  free(adu);
  free(af);
  state_free(&st);
  grid_free(&g);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Shell diffusion (perpendicular diffusion between neighbouring       */
/* streams on one shell)                                               */
/* ------------------------------------------------------------------ */
#define OR_THRESH 0.025
#define OR_VERYSMALL 1.0e-33

/* ShellData (src/energeticParticles.c:247-413) and DiffuseShellData
   (:418-555) for one EP substep of length dt.  Streams s in 0..S-1 in
   (face,row,col) order, nodes k in 0..N-1 (node 0 is skipped, like global
   node 0); nbr[S][4] = north, east, west, south neighbour streams.
   Positions x,y,z,rmag [S][N] in rScale units, B (r, theta, phi) [S][N];
   f [S][N][E][M] in/out.  Returns EP_OK or EP_ERR_NAN / EP_ERR_INF with the
   first bad cell in the reference's loop order (shell, energy, stream, mu). */
int or_shell_diffuse(const ep_params *p, int S, int N, int E, int M, const int32_t *nbr,
                     const double *x, const double *y, const double *z, const double *rmag,
                     const double *br, const double *bt, const double *bp, const double *rigidity,
                     const double *vgrid, double dt, double kperxkpar, double *f, int *bad) {
  double *dlper = (double *)malloc(sizeof(double) * 4 * S);
  double *delta = (double *)malloc(sizeof(double) * S * M);
  double *mfp = (double *)malloc(sizeof(double) * S);
  int k, e, s, d, m, rc = EP_OK;
  for (k = 1; k < N && rc == EP_OK; k++) {
    /* link lengths perpendicular to B (ShellData) */
    for (s = 0; s < S; s++) {
      const size_t i = (size_t)s * N + k;
      const double theta = acos(z[i] / rmag[i]);
      const double phi = atan2(y[i], x[i]);
      const double bx = br[i] * sin(theta) * cos(phi) + bt[i] * cos(theta) * cos(phi) - bp[i] * sin(phi);
      const double by = br[i] * sin(theta) * sin(phi) + bt[i] * cos(theta) * sin(phi) + bp[i] * cos(phi);
      const double bz = br[i] * cos(theta) - bt[i] * sin(theta);
      const double bmag = sqrt(br[i] * br[i] + bt[i] * bt[i] + bp[i] * bp[i]);
      const double ebx = bx / bmag, eby = by / bmag, ebz = bz / bmag;
      const double rx = x[i] * p->rScale, ry = y[i] * p->rScale, rz = z[i] * p->rScale;
      for (d = 0; d < 4; d++) {
        const size_t j = (size_t)nbr[s * 4 + d] * N + k;
        double ex = x[j] * p->rScale - rx, ey = y[j] * p->rScale - ry, ez = z[j] * p->rScale - rz;
        const double dl = sqrt(ex * ex + ey * ey + ez * ez);
        double dot;
        ex /= dl;
        ey /= dl;
        ez /= dl;
        dot = ex * ebx + ey * eby + ez * ebz;
        dlper[s * 4 + d] = dl * sqrt(1.0 - dot * dot);
      }
    }
    for (e = 0; e < E && rc == EP_OK; e++) {
      for (s = 0; s < S; s++) {
        const size_t i = (size_t)s * N + k;
        const double bmag = sqrt(br[i] * br[i] + bt[i] * bt[i] + bp[i] * bp[i]);
        mfp[s] = or_mfp(p, rigidity[e], rmag[i], bmag);
      }
      memset(delta, 0, sizeof(double) * S * M);
      for (s = 0; s < S; s++) {
        const double dt_kper = dt * 0.33333333 * mfp[s] * vgrid[e] * kperxkpar;
        double del[4];
        const double *fs = f + (((size_t)s * N + k) * E + e) * M;
        for (d = 0; d < 4; d++) {
          del[d] = dt_kper / (dlper[s * 4 + d] * dlper[s * 4 + d] + OR_VERYSMALL);
          if (del[d] > OR_THRESH) del[d] = OR_THRESH;
        }
        for (m = 0; m < M; m++) {
          for (d = 0; d < 4; d++) delta[nbr[s * 4 + d] * M + m] += del[d] * fs[m];
          delta[s * M + m] -= (del[0] + del[1] + del[2] + del[3]) * fs[m];
        }
      }
      for (s = 0; s < S && rc == EP_OK; s++) {
        double *fs = f + (((size_t)s * N + k) * E + e) * M;
        for (m = 0; m < M; m++) {
          fs[m] += delta[s * M + m];
          if (isnan(fs[m]) || isinf(fs[m])) {
            rc = isnan(fs[m]) ? EP_ERR_NAN : EP_ERR_INF;
            if (bad) {
              bad[0] = s;
              bad[1] = k;
              bad[2] = e;
              bad[3] = m;
            }
            break;
          }
        }
      }
    }
  }
  free(dlper);
  free(delta);
  free(mfp);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Drift between neighbouring streams                                  */
/* ------------------------------------------------------------------ */
#define OR_Q 4.80320425e-10
#define OR_OM (OR_MHD_B_NORM * OR_Q * OR_AU / (OR_MP * OR_C * OR_C))

/* parkerB (src/flow.c:440-470) without the ideal shock */
static void or_parker_b(double r, double theta, double vr, double mhdBsAu, double omega, double *b) {
  b[0] = mhdBsAu / (r * r);
  b[1] = 0.0;
  b[2] = -1.0 * r * b[0] * (omega / (vr + OR_VERYSMALL)) * sin(theta);
}

/* curlBoverB2 (src/flow.c:479-615): curl(B/B^2) of the Parker field by
   centred differences, cartesian position in rScale units. */
void or_curl_b_over_b2(double rScale, double mhdBsAu, double omega, double x, double y, double z, double vr,
                       double *curl) {
  const double rw = 0.001, tw = OR_PI / 100.0, pw = OR_PI / 100.0;
  double r, th, ph, A, B;
  double rm[3], rp[3], tm[3], tp[3], pm[3], pp[3];
  double Brm[3], Brp[3], Btm[3], Btp[3], Bpm[3], Bpp[3];
  double mrm, mrp, mtm, mtp, mpm, mpp;
  r = rScale * sqrt(x * x + y * y + z * z);
  th = acos(z * rScale / r);
  ph = atan2(y, x);
  rp[0] = r + rw; rp[1] = th; rp[2] = ph;
  rm[0] = r - rw; rm[1] = th; rm[2] = ph;
  tp[0] = r; tp[1] = th + tw; tp[2] = ph;
  if (tp[1] > OR_PI) {
    tp[1] = 2.0 * OR_PI - tp[1];
    tp[2] += OR_PI;
    if (tp[2] > (2.0 * OR_PI)) tp[2] -= (2.0 * OR_PI);
  }
  tm[0] = r; tm[1] = th - tw; tm[2] = ph;
  if (tm[1] < 0.0) {
    tm[1] *= -1.0;
    tm[2] += OR_PI;
    if (tm[2] > (2.0 * OR_PI)) tm[2] -= (2.0 * OR_PI);
  }
  pp[0] = r; pp[1] = th; pp[2] = ph;
  pp[2] += pw;
  if (pp[2] > (2.0 * OR_PI)) pp[2] -= (2.0 * OR_PI);
  pm[0] = r; pm[1] = th; pm[2] = ph;
  pm[2] -= pw;
  if (pm[2] < 0.0) pm[2] += (2.0 * OR_PI);
  or_parker_b(rm[0], rm[1], vr, mhdBsAu, omega, Brm);
  or_parker_b(rp[0], rp[1], vr, mhdBsAu, omega, Brp);
  or_parker_b(tm[0], tm[1], vr, mhdBsAu, omega, Btm);
  or_parker_b(tp[0], tp[1], vr, mhdBsAu, omega, Btp);
  or_parker_b(pm[0], pm[1], vr, mhdBsAu, omega, Bpm);
  or_parker_b(pp[0], pp[1], vr, mhdBsAu, omega, Bpp);
  mrm = sqrt(Brm[0] * Brm[0] + Brm[1] * Brm[1] + Brm[2] * Brm[2]);
  mrp = sqrt(Brp[0] * Brp[0] + Brp[1] * Brp[1] + Brp[2] * Brp[2]);
  mtm = sqrt(Btm[0] * Btm[0] + Btm[1] * Btm[1] + Btm[2] * Btm[2]);
  mtp = sqrt(Btp[0] * Btp[0] + Btp[1] * Btp[1] + Btp[2] * Btp[2]);
  mpm = sqrt(Bpm[0] * Bpm[0] + Bpm[1] * Bpm[1] + Bpm[2] * Bpm[2]);
  mpp = sqrt(Bpp[0] * Bpp[0] + Bpp[1] * Bpp[1] + Bpp[2] * Bpp[2]);
  A = (0.5 / tw) * (Btp[2] * sin(tp[1]) / (mtp * mtp) - Btm[2] * sin(tm[1]) / (mtm * mtm));
  B = (0.5 / pw) * (Bpp[1] / (mpp * mpp) - Bpm[1] / (mpm * mpm));
  curl[0] = (1.0 / (r * sin(th))) * (A - B);
  A = (1.0 / sin(th)) * (0.5 / pw) * (Bpp[0] / (mpp * mpp) - Bpm[0] / (mpm * mpm));
  B = (0.5 / rw) * (rp[0] * Brp[2] / (mrp * mrp) - rm[0] * Brm[2] / (mrm * mrm));
  curl[1] = (1.0 / r) * (A - B);
  A = (0.5 / rw) * (rp[0] * Brp[1] / (mrp * mrp) - rm[0] * Brm[1] / (mrm * mrm));
  B = (0.5 / tw) * (Btp[0] / (mtp * mtp) - Btm[0] / (mtm * mtm));
  curl[2] = (1.0 / r) * (A - B);
}

/* sphToCartVector (src/geometry.c:89-125) */
static void or_sph_to_cart(const double *v, double x, double y, double z, double *out) {
  const double rmag = sqrt(x * x + y * y + z * z);
  if ((rmag - fabs(z)) < OR_VERYSMALL) {
    out[0] = 0.0;
    out[1] = 0.0;
  } else {
    out[0] = x / rmag * v[0] + x * z / (rmag * sqrt(rmag * rmag - z * z)) * v[1] - y / sqrt(rmag * rmag - z * z) * v[2];
    out[1] = y / rmag * v[0] + y * z / (rmag * sqrt(rmag * rmag - z * z)) * v[1] + x / sqrt(rmag * rmag - z * z) * v[2];
  }
  out[2] = z / rmag * v[0] - sqrt(rmag * rmag - z * z) / rmag * v[1];
}

/* masMoveNodes (src/masInterp.c:1344-1393) for n nodes: the classical RK4
   of rungeKuttaFlow (:1305-1338) with vMas (:1279-1296), the MAS velocity
   of masGetNode at each stage position made Cartesian there
   (sphToCartVector).  Outside the MAS domain masGetNode takes masWind of
   the node as it stands before the move (its position, its previous old
   position and MHD state), the same at every stage.  The node's position
   before the move becomes its old position.  prev = the node's MHD state
   {br, bt, bp, vr, vt, vp, density}, each [n]. */
void or_move_nodes_mas(int n, double tDel, double rScale, double *x, double *y, double *z,
                       double *rmag, double *rox, double *roy, double *roz,
                       const double *const prev[7], const float *const fields0[7],
                       const float *const meshes0[21], const int *dims0,
                       const float *const fields1[7], const float *const meshes1[21],
                       const int *dims1, double s, double phi_offset, double mas_radial_max,
                       double mas_vmin) {
  const double sixth = 1.0 / 6.0, two = 2.0, half = 0.5;
  const double h = tDel / rScale;
  int i, q, st;
  for (i = 0; i < n; i++) {
    const double r0[3] = {x[i], y[i], z[i]};
    const double ox = rox[i], oy = roy[i], oz = roz[i], nrm = rmag[i];
    double p7[7], kk[4][3], pos[3], o7[7], vc[3];
    for (q = 0; q < 7; q++) p7[q] = prev[q][i];
    for (st = 0; st < 4; st++) {
      /* stage positions r0, r0 + k1/2, r0 + k2/2, r0 + k3 */
      for (q = 0; q < 3; q++)
        pos[q] = st == 0 ? r0[q] : (st < 3 ? r0[q] + kk[st - 1][q] * half : r0[q] + kk[2][q]);
      mas_get_node(rScale, pos[0], pos[1], pos[2], fields0, meshes0, dims0, fields1, meshes1, dims1, s,
                   phi_offset, mas_radial_max, mas_vmin, ox, oy, oz, nrm, p7, o7);
      or_sph_to_cart(o7 + 3, pos[0], pos[1], pos[2], vc);
      for (q = 0; q < 3; q++) kk[st][q] = vc[q] * h;
    }
    rox[i] = r0[0];
    roy[i] = r0[1];
    roz[i] = r0[2];
    x[i] = r0[0] + sixth * (kk[0][0] + two * kk[1][0] + two * kk[2][0] + kk[3][0]);
    y[i] = r0[1] + sixth * (kk[0][1] + two * kk[1][1] + two * kk[2][1] + kk[3][1]);
    z[i] = r0[2] + sixth * (kk[0][2] + two * kk[1][2] + two * kk[2][2] + kk[3][2]);
    rmag[i] = sqrt((x[i] * x[i]) + (y[i] * y[i]) + (z[i] * z[i]));
  }
}

/* DriftShellData (src/energeticParticles.c:560-801) with driftVelocity
   (:208-245) for one EP substep dt; curl [S][N][3] = curlBoverB2 per node,
   pgrid[E] = momentum (exp(lnpmin + e dlnp)); other arguments as
   or_shell_diffuse. */
int or_shell_drift(const ep_params *p, int S, int N, int E, int M, const int32_t *nbr, const double *x,
                   const double *y, const double *z, const double *curl, const double *pgrid,
                   const double *vgrid, double dt, double *f, int *bad) {
  double *delta = (double *)malloc(sizeof(double) * S * M);
  double *dl = (double *)malloc(sizeof(double) * 4 * S);
  const double om1 = p->charge * OR_OM / p->mass;
  int k, e, s, d, m, rc = EP_OK;
  for (k = 1; k < N && rc == EP_OK; k++) {
    for (s = 0; s < S; s++) { /* ShellData link lengths (AU) */
      const size_t i = (size_t)s * N + k;
      for (d = 0; d < 4; d++) {
        const size_t j = (size_t)nbr[s * 4 + d] * N + k;
        const double ex = x[j] * p->rScale - x[i] * p->rScale, ey = y[j] * p->rScale - y[i] * p->rScale,
                     ez = z[j] * p->rScale - z[i] * p->rScale;
        dl[s * 4 + d] = sqrt(ex * ex + ey * ey + ez * ez);
      }
    }
    for (e = 0; e < E && rc == EP_OK; e++) {
      const double factor = (1.0 / om1) * (pgrid[e] * vgrid[e] / 3.0);
      memset(delta, 0, sizeof(double) * S * M);
      for (s = 0; s < S; s++) {
        const size_t i = (size_t)s * N + k;
        const double *fs = f + ((size_t)i * E + e) * M;
        double vds[3], vd[3];
        vds[0] = factor * curl[i * 3 + 0];
        vds[1] = factor * curl[i * 3 + 1];
        vds[2] = factor * curl[i * 3 + 2];
        or_sph_to_cart(vds, x[i], y[i], z[i], vd);
        for (d = 0; d < 4; d++) {
          const int t = nbr[s * 4 + d];
          const size_t j = (size_t)t * N + k;
          const double ex = x[j] - x[i], ey = y[j] - y[i], ez = z[j] - z[i];
          const double vdp = (vd[0] * ex + vd[1] * ey + vd[2] * ez) * p->rScale / dl[s * 4 + d];
          if (vdp > 0.0) {
            double del = vdp * dt / dl[s * 4 + d];
            if (del > OR_THRESH) del = OR_THRESH;
            for (m = 0; m < M; m++) {
              delta[t * M + m] += del * fs[m];
              delta[s * M + m] -= del * fs[m];
            }
          }
        }
      }
      for (s = 0; s < S && rc == EP_OK; s++) {
        double *fs = f + (((size_t)s * N + k) * E + e) * M;
        for (m = 0; m < M; m++) {
          fs[m] += delta[s * M + m];
          if (isnan(fs[m]) || isinf(fs[m])) {
            rc = isnan(fs[m]) ? EP_ERR_NAN : EP_ERR_INF;
            if (bad) {
              bad[0] = s;
              bad[1] = k;
              bad[2] = e;
              bad[3] = m;
            }
            break;
          }
        }
      }
    }
  }
  free(delta);
  free(dl);
  return rc;
}
