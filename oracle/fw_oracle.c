/*
 * fw_oracle.c -- CPU restatement of the FWANT elastic-wave solver and of the
 * sensitivity-kernel time integral.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product path (fwant_amd/) never links, loads or calls it.
 *
 * It is compiled twice from the same source:
 *   liborc32.so  -- REAL = float   (the reference's default SP build)
 *   liborc64.so  -- REAL = double  (the reference's -DDataTypeDouble build)
 * with -ffp-contract=off so that every expression is evaluated in the
 * left-to-right, one-rounding-per-operation order of the Fortran source.
 *
 * Index conventions follow code/srcF/mod_para.F90:79-82 converted to C:
 * arrays are column-major (i fastest) with LenFD=3 ghost layers per side;
 * Fortran index I (1-based) is C index I-1, so ni1 = 3, ni2 = 3+ni-1, ...
 *
 * Field order everywhere: Vx Vy Vz Txx Tyy Tzz Txy Txz Tyz (the order of the
 * seismogram variables in code/srcF/mod_io.F90:436-444).
 *
 * Parity status of each part (what pins it) is stated at each function.
 * The coupled step is pinned by Stokes' point-force and moment-tensor
 * solutions (tests/test_oracle_analytic.py) and, for a surface force, by
 * Lamb's problem with the buried share of the reference's smoothed source
 * (tests/test_oracle_lamb.py); mode C by tests/test_oracle_topo.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_DOUBLE
typedef double real;
#define C_(x) x
#define ORC(name) orc64_##name
#else
typedef float real;
#define C_(x) x##f
#define ORC(name) orc32_##name
#endif

#define LENFD 3
#define SEIS_ZERO C_(1.0e-20)
static const double PI_D = 3.141592653589793;

enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TXY, TXZ, TYZ, NFIELD };

/* ------------------------------------------------------------------------
 * a1: DRP/opt MacCormack one-sided operators, code/srcF/mod_macdrp.h:19,69-85
 * and the 2-4 / 2-2 low-order operators, mod_macdrp.h:230-268.
 * Pinned by tests/test_oracle_ops.py (moments on polynomials, exactness of
 * the 2-2 and 2-4 operators on linear/quadratic data).
 * ---------------------------------------------------------------------- */
static const real MCA_1 = C_(-0.30874), MCA0 = C_(-0.6326), MCA1 = C_(1.233),
                  MCA2 = C_(-0.3334), MCA3 = C_(0.04168);
static const real MAC24A0 = C_(-7.0) / C_(6.0), MAC24A1 = C_(8.0) / C_(6.0),
                  MAC24A2 = C_(-1.0) / C_(6.0);
static const real MAC22A0 = C_(-1.0), MAC22A1 = C_(1.0);

static inline real opF(const real *v, long s) {
  return (MCA_1 * v[-s] + MCA0 * v[0] + MCA1 * v[s] + MCA2 * v[2 * s] + MCA3 * v[3 * s]);
}
static inline real opB(const real *v, long s) {
  return (-MCA3 * v[-3 * s] - MCA2 * v[-2 * s] - MCA1 * v[-s] - MCA0 * v[0] - MCA_1 * v[s]);
}
static inline real opD(int fwd, const real *v, long s) { return fwd ? opF(v, s) : opB(v, s); }
static inline real op24(int fwd, const real *v, long s) {
  return fwd ? (MAC24A0 * v[0] + MAC24A1 * v[s] + MAC24A2 * v[2 * s])
             : (-MAC24A2 * v[-2 * s] - MAC24A1 * v[-s] - MAC24A0 * v[0]);
}
static inline real op22(int fwd, const real *v, long s) {
  return fwd ? (MAC22A0 * v[0] + MAC22A1 * v[s]) : (-MAC22A1 * v[-s] - MAC22A0 * v[0]);
}

/* exported single-point operator evaluations (for the operator tests) */
real ORC(op_apply)(int kind, int fwd, const real *v /* points at centre */) {
  switch (kind) {
  case 0: return opD(fwd, v, 1);
  case 1: return op24(fwd, v, 1);
  default: return op22(fwd, v, 1);
  }
}

/* ------------------------------------------------------------------------
 * Solver state
 * ---------------------------------------------------------------------- */
typedef struct {
  int on, n;          /* face active, abs_number */
  int lo[3], hi[3];   /* inclusive C index bounds of the slab */
  int ex[3];          /* extents */
  real *W[6], *h[6], *m[6], *t[6]; /* 6 auxiliary variables, 4 copies */
} pml_face;

typedef struct {
  int ni, nj, nk, nx, ny, nz;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  long sy, sz, ncell;
  real stept;
  int freenode;
  /* grid */
  real *x, *y, *z, *xsin, *xcot, *xi_x, *eta_y, *zeta_z;
  /* media */
  real *lam, *mu, *rho;
  /* wavefield and RK registers */
  real *W[NFIELD], *h[NFIELD], *m[NFIELD], *t[NFIELD];
  real *TxSrc, *TySrc, *TzSrc, *VxSrc, *VySrc, *VzSrc;
  /* PML */
  int abs_number[6];
  real abs_vel[6], cfs_bmax[6], cfs_amax[6];
  real *Ax, *Bx, *Dx, *Ay, *By, *Dy, *Az, *Bz, *Dz;
  pml_face face[6]; /* x1 x2 y1 y2 z1 z2 */
  /* sources */
  int nforce, ntw_force, frc_stf_id;
  int *force_indx;            /* 3*n, C indices */
  real *force_shift;          /* 3*n */
  real *force_t0;             /* n */
  real *frc_time, *frc_freq;  /* ntw */
  real *Fx, *Fy, *Fz;         /* ntw*n, window fastest */
  int nmoment, ntw_moment, mom_stf_id;
  int *moment_indx;
  real *moment_shift, *moment_t0, *mom_time, *mom_freq;
  real *Mxx, *Myy, *Mzz, *Mxy, *Mxz, *Myz;
  /* receivers */
  int npt, pt_tinv;
  int *pt_indx; /* 3*npt C indices */
  int nthreads;
  int cycle_origin; /* ntime at which the 8-step F/B cycle starts (a restarted run starts it anew) */
  /* mode C (terrain-following shell, r = r(i, j, k)): 0 = the reference's
   * spherical scheme; else the 3-D metric below (ncell each) */
  int curv;
  real *r3, *cp, *cq, *czz, *crr; /* radius, slopes p q, zeta_r, 1/r */
  real *tau[3];                   /* traction combinations of one sweep */
} orc_solver;

#define IDX(s, i, j, k) ((long)(i) + (s)->sy * (long)(j) + (s)->sz * (long)(k))

/* Planted mistakes, for the control tests only (0: none).  A control test
 * switches one on, shows that the check it guards fails, and switches it off:
 *   1  sign of the cot(theta) terms of hVx flipped
 *   2  1/sin(theta) of the y derivative in hVx replaced by sin(theta)
 *   3  sign of the cot(theta) term of the y-face PML grouping of hVx flipped
 *   4  sign of the cot(theta) term of E22 flipped
 *   5  source times shifted by half an RK stage (tinc + 0.25)
 *   6  smoothed body force also applied on the free-surface plane
 *   7  smoothed body force doubled
 *   8  surface renormalisation of the smoothing weights skipped
 *   9  surface traction of a surface force doubled
 *  10  mode C: sign of the slope p flipped in the free-surface solve
 *  11  mode C: the slope terms of the velocity update's zeta part left out
 *  12  mode C: sign of the slope term p*aZ of E13 flipped */
static int g_mut = 0;
void ORC(set_mutation)(int m) { g_mut = m; }

static real *zalloc(long n) {
  real *p = (real *)calloc((size_t)(n > 0 ? n : 1), sizeof(real));
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* RK4 coefficients, mod_macdrp.h:39-40 and mod_macdrp.F90:205-206 */
static real rka(int s) { static const real a[4] = {C_(0.5), C_(0.5), C_(1.0), C_(0.0)}; return a[s]; }
static real rkb(int s) {
  static real b[4];
  b[0] = C_(1.0) / C_(6.0); b[1] = C_(1.0) / C_(3.0); b[2] = C_(1.0) / C_(3.0); b[3] = C_(1.0) / C_(6.0);
  return b[s];
}

/* ------------------------------------------------------------------------
 * PML profile helpers, code/srcF/mod_abs_npml.F90:9070-9124.
 * CONSPD = CONSPB = 2, CONSPA = 1 are integer parameters (:89-91).
 * Pinned by tests/test_oracle_pml.py (end values, R(10)=1e-3, monotonicity).
 * ---------------------------------------------------------------------- */
static real pml_R(int n) {
  return (real)pow(10.0, -(log10((double)n) - 1.0) / log10(2.0) - 3.0);
}
static real pml_dmax(real L, real Vp, real Rpp) {
  /* -Vp/2/L*log(Rpp)*(CONSPD+1) */
  return -Vp / C_(2.0) / L * (real)log((double)Rpp) * (C_(2.0) + C_(1.0));
}
static real pml_d(real x, real L, real dmax) { return x < 0 ? C_(0.0) : dmax * ((x / L) * (x / L)); }
static real pml_a(real x, real L, real amax) { return x < 0 ? C_(0.0) : amax * (C_(1.0) - (x / L)); }
static real pml_b(real x, real L, real bmax) {
  return x < 0 ? C_(1.0) : C_(1.0) + (bmax - C_(1.0)) * ((x / L) * (x / L));
}

/* ------------------------------------------------------------------------
 * Source-time functions, code/srcF/mod_src.F90:438-622.  PI is double
 * precision in the reference (mod_constants.F90:12), so expressions that
 * touch it are evaluated in double and rounded back to SP where the
 * reference assigns them to a real(SP) variable.
 * Pinned by tests/test_oracle_src.py (closed forms, derivative check).
 * ---------------------------------------------------------------------- */
enum { SIG_SVF_BELL = 100, SIG_SVF_TRIANGLE = 110, SIG_SVF_GAUSS = 120, SIG_STF_RICKER = 200,
       SIG_STF_BSHIFT = 210, SIG_STF_GAUSS = 220, SIG_STF_BELL = 230, SIG_STF_STEP = 240,
       SIG_STF_DELTA = 250 };

static real fun_gauss(real t, real a, real t0) {
  if (fabs((double)t0) >= SEIS_ZERO && (t <= 0 || t >= C_(2.0) * t0)) return 0;
  return (real)(exp((double)(-((t - t0) * (t - t0)) / (a * a))) / (sqrt(PI_D) * (double)a));
}
static real fun_gauss_deriv(real t, real a, real t0) {
  if (fabs((double)t0) >= SEIS_ZERO && (t <= 0 || t >= C_(2.0) * t0)) return 0;
  return (real)(exp((double)(-((t - t0) * (t - t0)) / (a * a))) / (sqrt(PI_D) * (double)a) *
                (double)(-2 * (t - t0) / (a * a)));
}
static real fun_ricker(real t, real fc, real t0) {
  if (t <= 0) return 0;
  real f0 = (real)(sqrt(PI_D) / 2.0);
  real u = (real)((double)((t - t0) * C_(2.0)) * PI_D * (double)fc);
  return (u * u / C_(4.0) - C_(0.5)) * (real)exp((double)(-(u * u) / C_(4.0))) * f0;
}
static real fun_ricker_deriv(real t, real fc, real t0) {
  if (t <= 0) return 0;
  real f0 = (real)(sqrt(PI_D) / 2.0);
  real u = (real)((double)((t - t0) * C_(2.0)) * PI_D * (double)fc);
  return (real)((double)(u * (C_(1.5) - u * u / C_(4.0)) * (real)exp((double)(-(u * u) / C_(4.0))) * f0) *
                PI_D * (double)fc);
}
static real fun_bell(real t, real riset) {
  if (t > 0 && t < riset) return (real)((1.0 - cos(2 * PI_D * (double)t / (double)riset)) / (double)riset);
  return 0;
}
static real fun_bell_deriv(real t, real riset) {
  if (t > 0 && t < riset) return (real)(2.0 * PI_D * sin(2 * PI_D * (double)t / (double)riset) / (double)riset);
  return 0;
}
static real fun_bell_int(real t, real riset) {
  if (t <= 0) return 0;
  if (t < riset) return (real)((double)(t / riset) - sin(2 * PI_D * (double)t / (double)riset) / (2.0 * PI_D));
  return 1;
}
static real fun_triangle(real t, real riset) {
  real t0 = riset / C_(2.0);
  if (t > riset) return 0;
  if (t > t0) return C_(2.0) / t0 - t / (t0 * t0);
  if (t > 0) return t / (t0 * t0);
  return 0;
}

/* src_svf / src_stf, mod_src.F90:475-533 (stept needed by fun_delta) */
static real src_svf(real t, real t0, real f0, int id, int ntime, real stept) {
  real r;
  switch (id) {
  case SIG_SVF_GAUSS: r = fun_gauss(t, f0, t0); break;
  case SIG_STF_GAUSS: r = fun_gauss_deriv(t, f0, t0); break;
  case SIG_STF_RICKER: r = fun_ricker_deriv(t, f0, t0); break;
  case SIG_SVF_BELL: r = fun_bell(t - t0, f0); break;
  case SIG_STF_BELL: r = fun_bell_deriv(t - t0, f0); break;
  case SIG_SVF_TRIANGLE: r = fun_triangle(t - t0, f0); break;
  case SIG_STF_STEP: r = (ntime == 0) ? C_(1.0) / stept : 0; break;
  default: r = 0; break;
  }
  if (fabs((double)r) < SEIS_ZERO) r = 0;
  return r;
}
static real src_stf(real t, real t0, real f0, int id, int ntime, real stept) {
  switch (id) {
  case SIG_STF_GAUSS: return fun_gauss(t, f0, t0);
  case SIG_STF_RICKER: return fun_ricker(t, f0, t0);
  case SIG_STF_BELL: return fun_bell(t - t0, f0);
  case SIG_SVF_BELL: return fun_bell_int(t - t0, f0);
  case SIG_STF_STEP: return 1;
  case SIG_STF_DELTA: return (ntime == 0) ? C_(1.0) / stept : 0;
  default: return 0;
  }
}
real ORC(stf)(int which, int id, real t, real t0, real f0, int ntime, real stept) {
  return which ? src_svf(t, t0, f0, id, ntime, stept) : src_stf(t, t0, f0, id, ntime, stept);
}

/* cal_norm_delt, mod_src.F90:419-437: 7^3 Gaussian weights, radius LenFD/2 */
static void cal_norm_delt(real *d /*7^3, i fastest*/, real x0, real y0, real z0) {
  real r0 = (real)(LENFD / 2.0);
  double sum = 0;
  real s = 0;
  for (int k = -LENFD; k <= LENFD; k++)
    for (int j = -LENFD; j <= LENFD; j++)
      for (int i = -LENFD; i <= LENFD; i++) {
        real D1 = fun_gauss((real)i - x0, r0, 0), D2 = fun_gauss((real)j - y0, r0, 0),
             D3 = fun_gauss((real)k - z0, r0, 0);
        d[(i + 3) + 7 * (j + 3) + 49 * (k + 3)] = D1 * D2 * D3;
      }
  for (int n = 0; n < 343; n++) s += d[n];
  (void)sum;
  for (int n = 0; n < 343; n++) d[n] = d[n] / s;
}
void ORC(norm_delt)(real *d, real x0, real y0, real z0) { cal_norm_delt(d, x0, y0, z0); }

/* ------------------------------------------------------------------------
 * Construction
 * ---------------------------------------------------------------------- */
orc_solver *ORC(create)(int ni, int nj, int nk, real stept, int freenode, int nthreads) {
  orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
  s->ni = ni; s->nj = nj; s->nk = nk;
  s->nx = ni + 2 * LENFD; s->ny = nj + 2 * LENFD; s->nz = nk + 2 * LENFD;
  s->ni1 = LENFD; s->ni2 = LENFD + ni - 1;
  s->nj1 = LENFD; s->nj2 = LENFD + nj - 1;
  s->nk1 = LENFD; s->nk2 = LENFD + nk - 1;
  s->sy = s->nx; s->sz = (long)s->nx * s->ny; s->ncell = s->sz * s->nz;
  s->stept = stept; s->freenode = freenode;
  s->nthreads = nthreads > 0 ? nthreads : 1;
  s->x = zalloc(s->nx); s->xsin = zalloc(s->nx); s->xcot = zalloc(s->nx); s->xi_x = zalloc(s->nx);
  s->y = zalloc(s->ny); s->eta_y = zalloc(s->ny);
  s->z = zalloc(s->nz); s->zeta_z = zalloc(s->nz);
  s->lam = zalloc(s->ncell); s->mu = zalloc(s->ncell); s->rho = zalloc(s->ncell);
  for (int f = 0; f < NFIELD; f++) {
    s->W[f] = zalloc(s->ncell); s->h[f] = zalloc(s->ncell);
    s->m[f] = zalloc(s->ncell); s->t[f] = zalloc(s->ncell);
  }
  long n2 = (long)s->nx * s->ny;
  s->TxSrc = zalloc(n2); s->TySrc = zalloc(n2); s->TzSrc = zalloc(n2);
  s->VxSrc = zalloc(n2); s->VySrc = zalloc(n2); s->VzSrc = zalloc(n2);
  s->Ax = zalloc(s->nx); s->Bx = zalloc(s->nx); s->Dx = zalloc(s->nx);
  s->Ay = zalloc(s->ny); s->By = zalloc(s->ny); s->Dy = zalloc(s->ny);
  s->Az = zalloc(s->nz); s->Bz = zalloc(s->nz); s->Dz = zalloc(s->nz);
  for (int i = 0; i < s->nx; i++) s->Bx[i] = 1;
  for (int j = 0; j < s->ny; j++) s->By[j] = 1;
  for (int k = 0; k < s->nz; k++) s->Bz[k] = 1;
  return s;
}

void ORC(destroy)(orc_solver *s) {
  if (!s) return;
  free(s->x); free(s->xsin); free(s->xcot); free(s->xi_x); free(s->y); free(s->eta_y);
  free(s->z); free(s->zeta_z); free(s->lam); free(s->mu); free(s->rho);
  for (int f = 0; f < NFIELD; f++) { free(s->W[f]); free(s->h[f]); free(s->m[f]); free(s->t[f]); }
  free(s->TxSrc); free(s->TySrc); free(s->TzSrc); free(s->VxSrc); free(s->VySrc); free(s->VzSrc);
  free(s->Ax); free(s->Bx); free(s->Dx); free(s->Ay); free(s->By); free(s->Dy);
  free(s->Az); free(s->Bz); free(s->Dz);
  for (int n = 0; n < 6; n++)
    if (s->face[n].on)
      for (int v = 0; v < 6; v++) {
        free(s->face[n].W[v]); free(s->face[n].h[v]); free(s->face[n].m[v]); free(s->face[n].t[v]);
      }
  free(s->force_indx); free(s->force_shift); free(s->force_t0); free(s->frc_time); free(s->frc_freq);
  free(s->Fx); free(s->Fy); free(s->Fz);
  free(s->moment_indx); free(s->moment_shift); free(s->moment_t0); free(s->mom_time); free(s->mom_freq);
  free(s->Mxx); free(s->Myy); free(s->Mzz); free(s->Mxy); free(s->Mxz); free(s->Myz);
  free(s->pt_indx);
  free(s->r3); free(s->cp); free(s->cq); free(s->czz); free(s->crr);
  for (int q = 0; q < 3; q++) free(s->tau[q]);
  free(s);
}

/* grid coordinates + metrics (as read from coord_/metric_ files, mod_grid.F90:96-114) */
void ORC(set_grid)(orc_solver *s, const real *x, const real *y, const real *z, const real *xsin,
                   const real *xcot, const real *xi_x, const real *eta_y, const real *zeta_z) {
  memcpy(s->x, x, sizeof(real) * s->nx); memcpy(s->y, y, sizeof(real) * s->ny);
  memcpy(s->z, z, sizeof(real) * s->nz); memcpy(s->xsin, xsin, sizeof(real) * s->nx);
  memcpy(s->xcot, xcot, sizeof(real) * s->nx); memcpy(s->xi_x, xi_x, sizeof(real) * s->nx);
  memcpy(s->eta_y, eta_y, sizeof(real) * s->ny); memcpy(s->zeta_z, zeta_z, sizeof(real) * s->nz);
}

/* ------------------------------------------------------------------------
 * a11: metric construction, code/srcF/seis3d_metric.F90:82-112.
 * x_xi = 1/2 (B.x + F.x) with the 5-tap coefficient vectors (:38-41);
 * xi_x = 1/x_xi; sin/cot in double; ghost metric copied from the edge.
 * Computed here in double (the reference uses SP for x_xi, DP for sin/cot).
 * Pinned by tests/test_oracle_metric.py (uniform grid closed form).
 * ---------------------------------------------------------------------- */
static void metric_1d(const double *g, int n, int lo, int hi, double *out_inv) {
  static const double cF[5] = {-0.30874, -0.6326, 1.2330, -0.3334, 0.04168};
  static const double cB[5] = {-0.04168, 0.3334, -1.2330, 0.6326, 0.30874};
  for (int i = lo; i <= hi; i++) {
    double b = 0, f = 0;
    for (int q = 0; q < 5; q++) { b += g[i - 3 + q] * cB[q]; f += g[i - 1 + q] * cF[q]; }
    out_inv[i] = 1.0 / ((b + f) * 0.5);
  }
  for (int q = 1; q <= LENFD; q++) { out_inv[lo - q] = out_inv[lo]; out_inv[hi + q] = out_inv[hi]; }
  (void)n;
}
void ORC(build_metric)(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz,
                       double *xsin, double *xcot, double *xi_x, double *eta_y, double *zeta_z) {
  for (int i = 0; i < nx; i++) { xsin[i] = sin(gx[i]); xcot[i] = 1.0 / tan(gx[i]); }
  metric_1d(gx, nx, LENFD, nx - 1 - LENFD, xi_x);
  metric_1d(gy, ny, LENFD, ny - 1 - LENFD, eta_y);
  metric_1d(gz, nz, LENFD, nz - 1 - LENFD, zeta_z);
}

/* ------------------------------------------------------------------------
 * Mode C metric (no counterpart in the reference, whose grid is a spherical
 * shell with a 1-D metric).  Terrain-following shell: theta = x(i),
 * phi = y(j), r = r(i, j, k).  With xi_theta = xi_x(i), eta_phi = eta_y(j)
 * (the 1-D metric above) the physical derivatives are
 *   d/dtheta|_r = xi_x d_xi + zeta_theta d_zeta,   zeta_theta = -r_xi xi_x / r_zeta
 *   d/dphi|_r   = eta_y d_eta + zeta_phi d_zeta,   zeta_phi   = -r_eta eta_y / r_zeta
 *   d/dr        = zeta_r d_zeta,                   zeta_r     = 1 / r_zeta
 * stored as the slopes p = zeta_theta / (r zeta_r) = -xi_x r_xi / r and
 * q = zeta_phi / (r sin(theta) zeta_r) = -eta_y r_eta / (r sin(theta)), zeta_r
 * and 1/r.  Derivatives of r use the metric operator of seis3d_metric
 * (1/2 (F + B), 5 taps): along zeta in the expression of metric_1d, so that a
 * flat grid gives zeta_r = zeta_z(k) exactly; along xi and eta with B
 * written as the mirrored F taps, so that a flat grid gives p = q = 0
 * exactly.  Derivative values at the 3 outer planes are copied from the
 * nearest plane the stencil fits (the ghost rule of metric_1d); 1/r is
 * evaluated everywhere.  FP64; 1/r from the FP32-rounded radius (as the
 * spherical scheme takes 1/z of the FP32 z).
 * Pinned by tests/test_oracle_topo.py (flat grid = the spherical metric,
 * slopes of an analytic surface, the RHS on analytic fields).
 * ---------------------------------------------------------------------- */
static const double MC_F[5] = {-0.30874, -0.6326, 1.2330, -0.3334, 0.04168};
static const double MC_B[5] = {-0.04168, 0.3334, -1.2330, 0.6326, 0.30874};
static int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
/* 1/2 (F + B) along a stride, B as mirrored F taps: exactly 0 on constants */
static double dmet_mirror(const double *g, long st) {
  double f = 0, b = 0;
  for (int q = 0; q < 5; q++) { f += g[(q - 1) * st] * MC_F[q]; b -= g[(1 - q) * st] * MC_F[q]; }
  return (f + b) * 0.5;
}
void ORC(build_metric3d)(int nx, int ny, int nz, const double *gx, const double *gy, const double *r3,
                         double *p, double *q, double *zz, double *rr) {
  const long sy = nx, sz = (long)nx * ny;
  double *xi = (double *)malloc(sizeof(double) * nx), *et = (double *)malloc(sizeof(double) * ny);
  metric_1d(gx, nx, LENFD, nx - 1 - LENFD, xi);
  metric_1d(gy, ny, LENFD, ny - 1 - LENFD, et);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const long c = i + sy * j + sz * k;
        const int kc = clampi(k, LENFD, nz - 1 - LENFD), ic = clampi(i, LENFD, nx - 1 - LENFD),
                  jc = clampi(j, LENFD, ny - 1 - LENFD);
        /* zeta_r: metric_1d's expression on the column (i, j) */
        const double *g = r3 + i + sy * j;
        double b = 0, f = 0;
        for (int m = 0; m < 5; m++) { b += g[(kc - 3 + m) * sz] * MC_B[m]; f += g[(kc - 1 + m) * sz] * MC_F[m]; }
        zz[c] = 1.0 / ((b + f) * 0.5);
        const long cx = ic + sy * j + sz * k, cy = i + sy * jc + sz * k;
        p[c] = -xi[ic] * dmet_mirror(r3 + cx, 1) / r3[cx];
        q[c] = -et[jc] * dmet_mirror(r3 + cy, sy) / (r3[cy] * sin(gx[i]));
        rr[c] = 1.0 / (double)(float)r3[c];
      }
  free(xi);
  free(et);
}
/* switch the solver to mode C with the 3-D metric (ncell each: radius,
 * slopes, zeta_r, 1/r) */
void ORC(set_metric3d)(orc_solver *s, const real *r3, const real *p, const real *q, const real *zz, const real *rr) {
  real **a[5] = {&s->r3, &s->cp, &s->cq, &s->czz, &s->crr};
  const real *src[5] = {r3, p, q, zz, rr};
  for (int n = 0; n < 5; n++) {
    if (!*a[n]) *a[n] = zalloc(s->ncell);
    memcpy(*a[n], src[n], sizeof(real) * s->ncell);
  }
  for (int n = 0; n < 3; n++)
    if (!s->tau[n]) s->tau[n] = zalloc(s->ncell);
  s->curv = 1;
}

void ORC(set_media)(orc_solver *s, const real *lam, const real *mu, const real *rho) {
  memcpy(s->lam, lam, sizeof(real) * s->ncell);
  memcpy(s->mu, mu, sizeof(real) * s->ncell);
  memcpy(s->rho, rho, sizeof(real) * s->ncell);
}

/* ------------------------------------------------------------------------
 * a5: abs_init, code/srcF/mod_abs_npml.F90:135-300.
 * abs_number / velocity / bmax / amax columns are x1 x2 y1 y2 z1 z2.
 * A z2 layer clears freenode (:182-184).  Lengths along x/y are arcs at
 * r = z(nk2) (:200-208).  Dx is stored as d/b (:282-284).
 * ---------------------------------------------------------------------- */
void ORC(set_pml)(orc_solver *s, const int *abs_number, const real *vel, const real *bmax,
                  const real *amax) {
  for (int n = 0; n < 6; n++) {
    s->abs_number[n] = abs_number[n] > 0 ? abs_number[n] : 0;
    s->abs_vel[n] = vel[n]; s->cfs_bmax[n] = bmax[n]; s->cfs_amax[n] = amax[n];
  }
  if (s->abs_number[5] > 0 && s->freenode) s->freenode = 0;
  real *z = s->z;
  int N;
  if ((N = s->abs_number[0]) > 0) {
    real x0 = s->x[N + s->ni1], L0 = z[s->nk2] * (x0 - s->x[s->ni1]);
    real dmax = pml_dmax(L0, s->abs_vel[0], pml_R(N));
    for (int i = s->ni1; i <= N + s->ni1; i++) {
      real Lx = z[s->nk2] * (x0 - s->x[i]);
      s->Dx[i] = pml_d(Lx, L0, dmax); s->Ax[i] = pml_a(Lx, L0, s->cfs_amax[0]); s->Bx[i] = pml_b(Lx, L0, s->cfs_bmax[0]);
    }
  }
  if ((N = s->abs_number[1]) > 0) {
    real x0 = s->x[s->ni2 - N], L0 = z[s->nk2] * (s->x[s->ni2] - x0);
    real dmax = pml_dmax(L0, s->abs_vel[1], pml_R(N));
    for (int i = s->ni2 - N; i <= s->ni2; i++) {
      real Lx = z[s->nk2] * (s->x[i] - x0);
      s->Dx[i] = pml_d(Lx, L0, dmax); s->Ax[i] = pml_a(Lx, L0, s->cfs_amax[1]); s->Bx[i] = pml_b(Lx, L0, s->cfs_bmax[1]);
    }
  }
  if ((N = s->abs_number[2]) > 0) {
    real y0 = s->y[N + s->nj1], L0 = z[s->nk2] * (y0 - s->y[s->nj1]);
    real dmax = pml_dmax(L0, s->abs_vel[2], pml_R(N));
    for (int j = s->nj1; j <= N + s->nj1; j++) {
      real Ly = z[s->nk2] * (y0 - s->y[j]);
      s->Dy[j] = pml_d(Ly, L0, dmax); s->Ay[j] = pml_a(Ly, L0, s->cfs_amax[2]); s->By[j] = pml_b(Ly, L0, s->cfs_bmax[2]);
    }
  }
  if ((N = s->abs_number[3]) > 0) {
    real y0 = s->y[s->nj2 - N], L0 = z[s->nk2] * (s->y[s->nj2] - y0);
    real dmax = pml_dmax(L0, s->abs_vel[3], pml_R(N));
    for (int j = s->nj2 - N; j <= s->nj2; j++) {
      real Ly = z[s->nk2] * (s->y[j] - y0);
      s->Dy[j] = pml_d(Ly, L0, dmax); s->Ay[j] = pml_a(Ly, L0, s->cfs_amax[3]); s->By[j] = pml_b(Ly, L0, s->cfs_bmax[3]);
    }
  }
  if ((N = s->abs_number[4]) > 0) {
    real z0 = z[N + s->nk1], L0 = z0 - z[s->nk1];
    real dmax = pml_dmax(L0, s->abs_vel[4], pml_R(N));
    for (int k = s->nk1; k <= s->nk1 + N; k++) {
      real Lz = z0 - z[k];
      s->Dz[k] = pml_d(Lz, L0, dmax); s->Az[k] = pml_a(Lz, L0, s->cfs_amax[4]); s->Bz[k] = pml_b(Lz, L0, s->cfs_bmax[4]);
    }
  }
  if ((N = s->abs_number[5]) > 0) {
    real z0 = z[s->nk2 - N], L0 = z[s->nk2] - z0;
    real dmax = pml_dmax(L0, s->abs_vel[5], pml_R(N));
    for (int k = s->nk2 - N; k <= s->nk2; k++) {
      real Lz = z[k] - z0;
      s->Dz[k] = pml_d(Lz, L0, dmax); s->Az[k] = pml_a(Lz, L0, s->cfs_amax[5]); s->Bz[k] = pml_b(Lz, L0, s->cfs_bmax[5]);
    }
  }
  for (int i = 0; i < s->nx; i++) s->Dx[i] = s->Dx[i] / s->Bx[i];
  for (int j = 0; j < s->ny; j++) s->Dy[j] = s->Dy[j] / s->By[j];
  for (int k = 0; k < s->nz; k++) s->Dz[k] = s->Dz[k] / s->Bz[k];
  /* slabs, :300-480 */
  for (int n = 0; n < 6; n++) {
    pml_face *f = &s->face[n];
    f->on = s->abs_number[n] > 0;
    if (!f->on) continue;
    f->n = s->abs_number[n];
    f->lo[0] = s->ni1; f->hi[0] = s->ni2; f->lo[1] = s->nj1; f->hi[1] = s->nj2;
    f->lo[2] = s->nk1; f->hi[2] = s->nk2;
    int ax = n / 2;
    if (n % 2 == 0) f->hi[ax] = f->lo[ax] + f->n; else f->lo[ax] = f->hi[ax] - f->n;
    long cnt = 1;
    for (int d = 0; d < 3; d++) { f->ex[d] = f->hi[d] - f->lo[d] + 1; cnt *= f->ex[d]; }
    for (int v = 0; v < 6; v++) {
      f->W[v] = zalloc(cnt); f->h[v] = zalloc(cnt); f->m[v] = zalloc(cnt); f->t[v] = zalloc(cnt);
    }
  }
}
void ORC(get_pml_profiles)(orc_solver *s, real *A, real *B, real *D) {
  /* A,B,D each nx+ny+nz: x then y then z */
  memcpy(A, s->Ax, sizeof(real) * s->nx); memcpy(A + s->nx, s->Ay, sizeof(real) * s->ny);
  memcpy(A + s->nx + s->ny, s->Az, sizeof(real) * s->nz);
  memcpy(B, s->Bx, sizeof(real) * s->nx); memcpy(B + s->nx, s->By, sizeof(real) * s->ny);
  memcpy(B + s->nx + s->ny, s->Bz, sizeof(real) * s->nz);
  memcpy(D, s->Dx, sizeof(real) * s->nx); memcpy(D + s->nx, s->Dy, sizeof(real) * s->ny);
  memcpy(D + s->nx + s->ny, s->Dz, sizeof(real) * s->nz);
}
int ORC(get_freenode)(orc_solver *s) { return s->freenode; }

/* sources: arrays as in source_mpi*.nc (mod_src.F90:150-185), indices converted to C */
void ORC(set_force)(orc_solver *s, int n, int ntw, int stf_id, const int *indx, const real *shift,
                    const real *t0, const real *stf_time, const real *stf_freq, const real *Fx,
                    const real *Fy, const real *Fz) {
  s->nforce = n; s->ntw_force = ntw; s->frc_stf_id = stf_id;
  s->force_indx = (int *)malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
  memcpy(s->force_indx, indx, sizeof(int) * 3 * n);
  s->force_shift = zalloc(3 * n); memcpy(s->force_shift, shift, sizeof(real) * 3 * n);
  s->force_t0 = zalloc(n); memcpy(s->force_t0, t0, sizeof(real) * n);
  s->frc_time = zalloc(ntw); memcpy(s->frc_time, stf_time, sizeof(real) * ntw);
  s->frc_freq = zalloc(ntw); memcpy(s->frc_freq, stf_freq, sizeof(real) * ntw);
  s->Fx = zalloc((long)ntw * n); memcpy(s->Fx, Fx, sizeof(real) * ntw * n);
  s->Fy = zalloc((long)ntw * n); memcpy(s->Fy, Fy, sizeof(real) * ntw * n);
  s->Fz = zalloc((long)ntw * n); memcpy(s->Fz, Fz, sizeof(real) * ntw * n);
}
void ORC(set_moment)(orc_solver *s, int n, int ntw, int stf_id, const int *indx, const real *shift,
                     const real *t0, const real *stf_time, const real *stf_freq, const real *M6 /* 6*ntw*n */) {
  s->nmoment = n; s->ntw_moment = ntw; s->mom_stf_id = stf_id;
  s->moment_indx = (int *)malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
  memcpy(s->moment_indx, indx, sizeof(int) * 3 * n);
  s->moment_shift = zalloc(3 * n); memcpy(s->moment_shift, shift, sizeof(real) * 3 * n);
  s->moment_t0 = zalloc(n); memcpy(s->moment_t0, t0, sizeof(real) * n);
  s->mom_time = zalloc(ntw); memcpy(s->mom_time, stf_time, sizeof(real) * ntw);
  s->mom_freq = zalloc(ntw); memcpy(s->mom_freq, stf_freq, sizeof(real) * ntw);
  long c = (long)ntw * n;
  real **M[6] = {&s->Mxx, &s->Myy, &s->Mzz, &s->Mxy, &s->Mxz, &s->Myz};
  for (int q = 0; q < 6; q++) { *M[q] = zalloc(c); memcpy(*M[q], M6 + q * c, sizeof(real) * c); }
}
void ORC(set_receivers)(orc_solver *s, int n, int tinv, const int *indx) {
  s->npt = n; s->pt_tinv = tinv;
  s->pt_indx = (int *)malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
  memcpy(s->pt_indx, indx, sizeof(int) * 3 * n);
}

/* ------------------------------------------------------------------------
 * a3: macdrp_syn and RK stages, code/srcF/mod_macdrp.F90:268-393 and the
 * PML counterparts mod_abs_npml.F90:641-1105.
 * Pinned by tests/test_oracle_rk.py (RK4 stability polynomial on y'=lam*y).
 * ---------------------------------------------------------------------- */
static void rk_arrays(int kind, real a, real b, long n, real *W, real *m, real *t, const real *h) {
  if (kind == 0) {
    for (long c = 0; c < n; c++) { W[c] = m[c] + a * h[c]; t[c] = m[c] + b * h[c]; }
  } else if (kind == 1) {
    for (long c = 0; c < n; c++) { W[c] = m[c] + a * h[c]; t[c] = t[c] + b * h[c]; }
  } else {
    for (long c = 0; c < n; c++) W[c] = t[c] + b * h[c];
  }
}
/* exported for the RK pin: one stage of the 4-register RK on flat arrays */
void ORC(rk_stage)(int kind, real rka_, real rkb_, real stept, long n, real *W, real *m, real *t, const real *h) {
  rk_arrays(kind, rka_ * stept, rkb_ * stept, n, W, m, t, h);
}
void ORC(rk_coef)(real *a4, real *b4) { for (int q = 0; q < 4; q++) { a4[q] = rka(q); b4[q] = rkb(q); } }

static void macdrp_syn(orc_solver *s) {
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int k = s->nk1; k <= s->nk2; k++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c = IDX(s, i, j, k);
        for (int f = 0; f < NFIELD; f++) s->m[f][c] = s->W[f][c];
      }
}
static void macdrp_rk(orc_solver *s, int kind, real a, real b) {
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int k = s->nk1; k <= s->nk2; k++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c = IDX(s, i, j, k);
        for (int f = 0; f < NFIELD; f++) {
          if (kind == 0) { s->W[f][c] = s->m[f][c] + a * s->h[f][c]; s->t[f][c] = s->m[f][c] + b * s->h[f][c]; }
          else if (kind == 1) { s->W[f][c] = s->m[f][c] + a * s->h[f][c]; s->t[f][c] = s->t[f][c] + b * s->h[f][c]; }
          else s->W[f][c] = s->t[f][c] + b * s->h[f][c];
        }
      }
}
static long face_cnt(const pml_face *f) { return (long)f->ex[0] * f->ex[1] * f->ex[2]; }
static void abs_syn(orc_solver *s) {
  for (int n = 0; n < 6; n++) {
    pml_face *f = &s->face[n];
    if (!f->on) continue;
    long cnt = face_cnt(f);
    for (int v = 0; v < 6; v++) memcpy(f->m[v], f->W[v], sizeof(real) * cnt);
  }
}
static void abs_rk(orc_solver *s, int kind, real a, real b) {
  for (int n = 0; n < 6; n++) {
    pml_face *f = &s->face[n];
    if (!f->on) continue;
    long cnt = face_cnt(f);
    for (int v = 0; v < 6; v++) rk_arrays(kind, a, b, cnt, f->W[v], f->m[v], f->t[v], f->h[v]);
  }
}

/* ------------------------------------------------------------------------
 * a4: free_timg, code/srcF/mod_macdrp.F90:3414-3436.
 * ---------------------------------------------------------------------- */
/* Mode C traction imaging (no counterpart in the reference): the traction
 * combinations across the surface zeta = const, scaled so that the radial
 * stress comes first,
 *   tau_x = Txz + p Txx + q Txy,  tau_y = Tyz + p Txy + q Tyy,  tau_z = Tzz + p Txz + q Tyz
 * (sigma . N with N = (p, q, 1)), are set to the source traction at k = nk2
 * by solving for Txz, Tyz, then Tzz, and imaged oddly above it: the planes
 * nk2 + m hold tau(nk2 + m) = 2 tau_src - tau(nk2 - m) in Txz Tyz Tzz.  With
 * p = q = 0 this is free_timg. */
static void free_timg_c(orc_solver *s) {
  real **W = s->W;
  int nk2 = s->nk2;
  for (int m = 1; m <= 3; m++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c2 = i + s->sy * j, b = IDX(s, i, j, nk2 - m), a = IDX(s, i, j, nk2 + m);
        real p = s->cp[b], q = s->cq[b];
        real tx = W[TXZ][b] + p * W[TXX][b] + q * W[TXY][b];
        real ty = W[TYZ][b] + p * W[TXY][b] + q * W[TYY][b];
        real tz = W[TZZ][b] + p * W[TXZ][b] + q * W[TYZ][b];
        W[TXZ][a] = C_(2.0) * s->TxSrc[c2] - tx;
        W[TYZ][a] = C_(2.0) * s->TySrc[c2] - ty;
        W[TZZ][a] = C_(2.0) * s->TzSrc[c2] - tz;
      }
  for (int j = s->nj1; j <= s->nj2; j++)
    for (int i = s->ni1; i <= s->ni2; i++) {
      long c2 = i + s->sy * j, c = IDX(s, i, j, nk2);
      real p = s->cp[c], q = s->cq[c];
      W[TXZ][c] = s->TxSrc[c2] - p * W[TXX][c] - q * W[TXY][c];
      W[TYZ][c] = s->TySrc[c2] - p * W[TXY][c] - q * W[TYY][c];
      W[TZZ][c] = s->TzSrc[c2] - p * W[TXZ][c] - q * W[TYZ][c];
    }
}

static void free_timg(orc_solver *s) {
  if (!s->freenode) return;
  if (s->curv) { free_timg_c(s); return; }
  int nk2 = s->nk2;
  for (int k = 1; k <= 3; k++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c2 = i + s->sy * j;
        s->W[TZZ][IDX(s, i, j, nk2 + k)] = C_(2.0) * s->TzSrc[c2] - s->W[TZZ][IDX(s, i, j, nk2 - k)];
        s->W[TXZ][IDX(s, i, j, nk2 + k)] = C_(2.0) * s->TxSrc[c2] - s->W[TXZ][IDX(s, i, j, nk2 - k)];
        s->W[TYZ][IDX(s, i, j, nk2 + k)] = C_(2.0) * s->TySrc[c2] - s->W[TYZ][IDX(s, i, j, nk2 - k)];
      }
  for (int j = s->nj1; j <= s->nj2; j++)
    for (int i = s->ni1; i <= s->ni2; i++) {
      long c2 = i + s->sy * j;
      s->W[TZZ][IDX(s, i, j, nk2)] = s->TzSrc[c2];
      s->W[TXZ][IDX(s, i, j, nk2)] = s->TxSrc[c2];
      s->W[TYZ][IDX(s, i, j, nk2)] = s->TySrc[c2];
    }
}

/* ------------------------------------------------------------------------
 * a2: the operator sweep Lx?_Ly?_Lz?, code/srcF/mod_macdrp.F90:790-962
 * (the 8 variants differ only in the F/B direction per axis).
 * Pinned by tests/test_oracle_rhs.py against the continuous spherical
 * elastodynamic equations on smooth fields, plus the traction properties.
 * ---------------------------------------------------------------------- */
static void op_sweep(orc_solver *s, int xf, int yf, int zf) {
  real **W = s->W, **h = s->h;
  const long sy = s->sy, sz = s->sz;
  const int nk2 = s->nk2, freenode = s->freenode;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int k = s->nk1; k <= nk2; k++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c = IDX(s, i, j, k);
        real lam = s->lam[c], miu = s->mu[c], lam2mu = lam + C_(2.0) * miu, rrho = C_(1.0) / s->rho[c];
        real xi = s->xi_x[i], et = s->eta_y[j], ze = s->zeta_z[k];
        real DxTxx = opD(xf, W[TXX] + c, 1) * xi, DxTxy = opD(xf, W[TXY] + c, 1) * xi;
        real DxTxz = opD(xf, W[TXZ] + c, 1) * xi, DxVx = opD(xf, W[VX] + c, 1) * xi;
        real DxVy = opD(xf, W[VY] + c, 1) * xi, DxVz = opD(xf, W[VZ] + c, 1) * xi;
        real DyTyy = opD(yf, W[TYY] + c, sy) * et, DyTxy = opD(yf, W[TXY] + c, sy) * et;
        real DyTyz = opD(yf, W[TYZ] + c, sy) * et, DyVx = opD(yf, W[VX] + c, sy) * et;
        real DyVy = opD(yf, W[VY] + c, sy) * et, DyVz = opD(yf, W[VZ] + c, sy) * et;
        real DzTzz = opD(zf, W[TZZ] + c, sz) * ze, DzTxz = opD(zf, W[TXZ] + c, sz) * ze;
        real DzTyz = opD(zf, W[TYZ] + c, sz) * ze;
        real DzVx, DzVy, DzVz;
        if (freenode && k == nk2 - 1) {
          DzVx = op22(zf, W[VX] + c, sz) * ze; DzVy = op22(zf, W[VY] + c, sz) * ze; DzVz = op22(zf, W[VZ] + c, sz) * ze;
        } else if (freenode && k == nk2 - 2) {
          DzVx = op24(zf, W[VX] + c, sz) * ze; DzVy = op24(zf, W[VY] + c, sz) * ze; DzVz = op24(zf, W[VZ] + c, sz) * ze;
        } else {
          DzVx = opD(zf, W[VX] + c, sz) * ze; DzVy = opD(zf, W[VY] + c, sz) * ze; DzVz = opD(zf, W[VZ] + c, sz) * ze;
        }
        if (miu <= SEIS_ZERO) { DxTxy = 0; DxTxz = 0; DyTxy = 0; DyTyz = 0; DzTxz = 0; DzTyz = 0; }
        real zk = s->z[k], xs = s->xsin[i], xc = s->xcot[i];
        real Txx = W[TXX][c], Tyy = W[TYY][c], Tzz = W[TZZ][c], Txy = W[TXY][c], Txz = W[TXZ][c],
             Tyz = W[TYZ][c], Vx = W[VX][c], Vy = W[VY][c], Vz = W[VZ][c];
        h[VX][c] = rrho * (DxTxx / zk + DyTxy / zk / xs + DzTxz + (C_(3.0) * Txz + Txx * xc - Tyy * xc) / zk);
        if (g_mut == 1) h[VX][c] = rrho * (DxTxx / zk + DyTxy / zk / xs + DzTxz + (C_(3.0) * Txz - Txx * xc + Tyy * xc) / zk);
        if (g_mut == 2) h[VX][c] = rrho * (DxTxx / zk + DyTxy / zk * xs + DzTxz + (C_(3.0) * Txz + Txx * xc - Tyy * xc) / zk);
        h[VY][c] = rrho * (DxTxy / zk + DyTyy / zk / xs + DzTyz + (C_(2.0) * Txy * xc + C_(3.0) * Tyz) / zk);
        h[VZ][c] = rrho * (DxTxz / zk + DyTyz / zk / xs + DzTzz + (C_(2.0) * Tzz - Txx - Tyy + Txz * xc) / zk);
        real E11 = (DxVx + Vz) / zk;
        real E22 = (Vx * xc + DyVy / xs + Vz) / zk;
        if (g_mut == 4) E22 = (-Vx * xc + DyVy / xs + Vz) / zk;
        real E12 = (DxVy + DyVx / xs - Vy * xc) / zk / C_(2.0);
        real E33 = DzVz;
        real E13 = (DxVz / zk + DzVx - Vx / zk) / C_(2.0);
        real E23 = (DyVz / zk / xs + DzVy - Vy / zk) / C_(2.0);
        if (freenode && k == nk2) {
          long c2 = i + sy * j;
          E33 = -lam / lam2mu * (E11 + E22) + s->VzSrc[c2] / lam2mu;
          E13 = s->VxSrc[c2] / miu / C_(2.0);
          E23 = s->VySrc[c2] / miu / C_(2.0);
        }
        h[TXX][c] = lam2mu * E11 + lam * (E22 + E33);
        h[TYY][c] = lam2mu * E22 + lam * (E11 + E33);
        h[TZZ][c] = lam2mu * E33 + lam * (E11 + E22);
        h[TXY][c] = C_(2.0) * miu * E12;
        h[TXZ][c] = C_(2.0) * miu * E13;
        h[TYZ][c] = C_(2.0) * miu * E23;
      }
}

/* ------------------------------------------------------------------------
 * Mode C operator sweep (no counterpart in the reference).  The spherical
 * velocity-stress equations of op_sweep with the physical derivatives of a
 * terrain-following grid: d/dtheta = xi_x F_xi + zeta_theta F_zeta etc.
 * (build_metric3d).  Velocity half: the zeta parts of a row of the stress
 * divergence, zeta_r (F(sigma_ir) + p F(sigma_itheta) + q F(sigma_iphi)),
 * are evaluated as zeta_r (F(tau_i) - sigma_itheta F(p) - sigma_iphi F(q))
 * with the traction combinations tau of free_timg_c, so that the odd images
 * above the free surface enter only through tau.  Stress half: with
 * a_i = zeta_r F(V_i) (F of the VLOW orders next to the surface),
 *   E11 = (Dx Vx + Vz)/r + p aX,  E22 = (Vx cot + Dy Vy / sin + Vz)/r + q aY,
 *   E12 = ((Dx Vy + Dy Vx / sin - Vy cot)/r + p aY + q aX)/2,  E33 = aZ,
 *   E13 = (Dx Vz / r + aX - Vx / r + p aZ)/2,  E23 = (Dy Vz / (r sin) + aY - Vy / r + q aZ)/2
 * (Dx, Dy: the xi and eta parts).  At the free surface, a is solved from the
 * traction-rate condition (sigma_t . N = S, surf_solve_c); where p = q = 0
 * the reference's traction condition is used as written.  With p = q = 0
 * and r = z(k) everywhere the sweep is op_sweep up to rounding.
 * Pinned by tests/test_oracle_topo.py: the flat grid against op_sweep, the
 * RHS on analytic fields against the continuous equations on a curved
 * grid, the surface solve's traction rates, the Rayleigh speed along a
 * sloping surface.
 * ---------------------------------------------------------------------- */
/* Free-surface strain rates on a sloping surface: E holds the strain rates
 * of the xi / eta derivatives (E11 E22 E33 E12 E13 E23 with E33 = 0 and the
 * zeta parts left out); a = zeta_r F(V) is the solution of
 *   (mu n2 I + (lam + mu) u u^T) a = S - sigma(E) . N,  u = N = (p, q, 1), n2 = u.u
 * (the traction rate of Hooke's law with the zeta parts, by Sherman-Morrison),
 * and E receives the zeta parts. */
static void surf_solve_c(real lam, real miu, real p, real q, real Sx, real Sy, real Sz, real *E) {
  real lam2mu = lam + C_(2.0) * miu;
  if (g_mut == 10) p = -p;
  real tr = E[0] + E[1];
  real sxx = lam * tr + C_(2.0) * miu * E[0], syy = lam * tr + C_(2.0) * miu * E[1], szz = lam * tr;
  real sxy = C_(2.0) * miu * E[3], sxz = C_(2.0) * miu * E[4], syz = C_(2.0) * miu * E[5];
  real R0 = Sx - (p * sxx + q * sxy + sxz);
  real R1 = Sy - (p * sxy + q * syy + syz);
  real R2 = Sz - (p * sxz + q * syz + szz);
  real n2 = C_(1.0) + p * p + q * q;
  real g = (lam + miu) / (lam2mu * n2) * (p * R0 + q * R1 + R2);
  real mn = miu * n2;
  real aX = (R0 - g * p) / mn, aY = (R1 - g * q) / mn, aZ = (R2 - g) / mn;
  E[0] = E[0] + p * aX;
  E[1] = E[1] + q * aY;
  E[2] = aZ;
  E[3] = E[3] + (p * aY + q * aX) * C_(0.5);
  E[4] = E[4] + (aX + p * aZ) * C_(0.5);
  E[5] = E[5] + (aY + q * aZ) * C_(0.5);
}
static void op_sweep_c(orc_solver *s, int xf, int yf, int zf) {
  real **W = s->W, **h = s->h, **tau = s->tau;
  const long sy = s->sy, sz = s->sz;
  const int nk2 = s->nk2, freenode = s->freenode;
  /* traction combinations on every plane of the computed columns */
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int k = 0; k < s->nz; k++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c = IDX(s, i, j, k);
        if (freenode && k > nk2) {
          tau[0][c] = W[TXZ][c]; tau[1][c] = W[TYZ][c]; tau[2][c] = W[TZZ][c];
        } else {
          real p = s->cp[c], q = s->cq[c];
          if (g_mut == 11) { p = 0; q = 0; }
          tau[0][c] = W[TXZ][c] + p * W[TXX][c] + q * W[TXY][c];
          tau[1][c] = W[TYZ][c] + p * W[TXY][c] + q * W[TYY][c];
          tau[2][c] = W[TZZ][c] + p * W[TXZ][c] + q * W[TYZ][c];
        }
      }
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int k = s->nk1; k <= nk2; k++)
    for (int j = s->nj1; j <= s->nj2; j++)
      for (int i = s->ni1; i <= s->ni2; i++) {
        long c = IDX(s, i, j, k);
        real lam = s->lam[c], miu = s->mu[c], lam2mu = lam + C_(2.0) * miu, rrho = C_(1.0) / s->rho[c];
        real xi = s->xi_x[i], et = s->eta_y[j], ze = s->czz[c], rzk = s->crr[c];
        real p = s->cp[c], q = s->cq[c];
        real rxs = C_(1.0) / s->xsin[i], xc = s->xcot[i];
        real DxTxx = opD(xf, W[TXX] + c, 1) * xi, DxTxy = opD(xf, W[TXY] + c, 1) * xi;
        real DxTxz = opD(xf, W[TXZ] + c, 1) * xi, DxVx = opD(xf, W[VX] + c, 1) * xi;
        real DxVy = opD(xf, W[VY] + c, 1) * xi, DxVz = opD(xf, W[VZ] + c, 1) * xi;
        real DyTyy = opD(yf, W[TYY] + c, sy) * et, DyTxy = opD(yf, W[TXY] + c, sy) * et;
        real DyTyz = opD(yf, W[TYZ] + c, sy) * et, DyVx = opD(yf, W[VX] + c, sy) * et;
        real DyVy = opD(yf, W[VY] + c, sy) * et, DyVz = opD(yf, W[VZ] + c, sy) * et;
        real Txx = W[TXX][c], Tyy = W[TYY][c], Tzz = W[TZZ][c], Txy = W[TXY][c], Txz = W[TXZ][c],
             Tyz = W[TYZ][c], Vx = W[VX][c], Vy = W[VY][c], Vz = W[VZ][c];
        real dp = opD(zf, s->cp + c, sz), dq = opD(zf, s->cq + c, sz);
        if (g_mut == 11) { dp = 0; dq = 0; }
        real Zx = (opD(zf, tau[0] + c, sz) - Txx * dp - Txy * dq) * ze;
        real Zy = (opD(zf, tau[1] + c, sz) - Txy * dp - Tyy * dq) * ze;
        real Zz = (opD(zf, tau[2] + c, sz) - Txz * dp - Tyz * dq) * ze;
        real aX, aY, aZ;
        if (freenode && k == nk2 - 1) {
          aX = op22(zf, W[VX] + c, sz) * ze; aY = op22(zf, W[VY] + c, sz) * ze; aZ = op22(zf, W[VZ] + c, sz) * ze;
        } else if (freenode && k == nk2 - 2) {
          aX = op24(zf, W[VX] + c, sz) * ze; aY = op24(zf, W[VY] + c, sz) * ze; aZ = op24(zf, W[VZ] + c, sz) * ze;
        } else {
          aX = opD(zf, W[VX] + c, sz) * ze; aY = opD(zf, W[VY] + c, sz) * ze; aZ = opD(zf, W[VZ] + c, sz) * ze;
        }
        h[VX][c] = rrho * (DxTxx * rzk + DyTxy * rzk * rxs + Zx + (C_(3.0) * Txz + Txx * xc - Tyy * xc) * rzk);
        h[VY][c] = rrho * (DxTxy * rzk + DyTyy * rzk * rxs + Zy + (C_(2.0) * Txy * xc + C_(3.0) * Tyz) * rzk);
        h[VZ][c] = rrho * (DxTxz * rzk + DyTyz * rzk * rxs + Zz + (C_(2.0) * Tzz - Txx - Tyy + Txz * xc) * rzk);
        real E[6];
        if (freenode && k == nk2) {
          long c2 = i + sy * j;
          if (p == 0 && q == 0) {
            E[0] = (DxVx + Vz) * rzk;
            E[1] = (Vx * xc + DyVy * rxs + Vz) * rzk;
            E[3] = (DxVy + DyVx * rxs - Vy * xc) * rzk * C_(0.5);
            E[2] = -lam / lam2mu * (E[0] + E[1]) + s->VzSrc[c2] / lam2mu;
            E[4] = s->VxSrc[c2] / miu / C_(2.0);
            E[5] = s->VySrc[c2] / miu / C_(2.0);
          } else {
            E[0] = (DxVx + Vz) * rzk;
            E[1] = (Vx * xc + DyVy * rxs + Vz) * rzk;
            E[2] = 0;
            E[3] = (DxVy + DyVx * rxs - Vy * xc) * rzk * C_(0.5);
            E[4] = (DxVz * rzk - Vx * rzk) * C_(0.5);
            E[5] = (DyVz * rzk * rxs - Vy * rzk) * C_(0.5);
            surf_solve_c(lam, miu, p, q, s->VxSrc[c2], s->VySrc[c2], s->VzSrc[c2], E);
          }
        } else {
          E[0] = (DxVx + Vz) * rzk + p * aX;
          E[1] = (Vx * xc + DyVy * rxs + Vz) * rzk + q * aY;
          E[3] = ((DxVy + DyVx * rxs - Vy * xc) * rzk + p * aY + q * aX) * C_(0.5);
          E[2] = aZ;
          E[4] = (DxVz * rzk + aX - Vx * rzk + (g_mut == 12 ? -p : p) * aZ) * C_(0.5);
          E[5] = (DyVz * rzk * rxs + aY - Vy * rzk + q * aZ) * C_(0.5);
        }
        h[TXX][c] = lam2mu * E[0] + lam * (E[1] + E[2]);
        h[TYY][c] = lam2mu * E[1] + lam * (E[0] + E[2]);
        h[TZZ][c] = lam2mu * E[2] + lam * (E[0] + E[1]);
        h[TXY][c] = C_(2.0) * miu * E[3];
        h[TXZ][c] = C_(2.0) * miu * E[4];
        h[TYZ][c] = C_(2.0) * miu * E[5];
      }
}

/* exported: evaluate the RHS of one operator pattern on the current state
 * (after the free-surface imaging a stage would apply) into h (9 fields) */
void ORC(rhs)(orc_solver *s, int xf, int yf, int zf, real *hout /* 9*ncell */) {
  free_timg(s);
  if (s->curv) op_sweep_c(s, xf, yf, zf);
  else op_sweep(s, xf, yf, zf);
  for (int f = 0; f < NFIELD; f++) memcpy(hout + (long)f * s->ncell, s->h[f], sizeof(real) * s->ncell);
}

/* ------------------------------------------------------------------------
 * a6: abs_vsymm (mod_abs_npml.F90:9014-9067) and the PML corrections
 * Lx?_Ly?_Lz? (mod_abs_npml.F90:1149-2123; the 8 variants differ only in F/B).
 * Pinned by tests/test_oracle_pml.py (CFS stretching transfer function of
 * the ADE update on one cell; absorption of the whole solver).
 * ---------------------------------------------------------------------- */
static void abs_vsymm(orc_solver *s) {
  if (!s->freenode) return;
  for (int n = 0; n < 4; n++) {
    pml_face *f = &s->face[n];
    if (!f->on) continue;
    for (int j = f->lo[1]; j <= f->hi[1]; j++)
      for (int i = f->lo[0]; i <= f->hi[0]; i++)
        for (int m = 1; m <= LENFD; m++)
          for (int v = VX; v <= VZ; v++)
            s->W[v][IDX(s, i, j, s->nk2 + m)] = C_(2.0) * s->W[v][IDX(s, i, j, s->nk2)] - s->W[v][IDX(s, i, j, s->nk2 - m)];
  }
}

/* aux variable order per face:
 *   x: Vx Vy Vz Txx Txy Txz   y: Vx Vy Vz Tyy Txy Tyz   z: Vx Vy Vz Tzz Txz Tyz */
static void abs_correct(orc_solver *s, int xf, int yf, int zf) {
  real **W = s->W, **h = s->h;
  const long sy = s->sy, sz = s->sz;
  for (int n = 0; n < 6; n++) {
    pml_face *f = &s->face[n];
    if (!f->on) continue;
    int ax = n / 2;
    for (int k = f->lo[2]; k <= f->hi[2]; k++)
      for (int j = f->lo[1]; j <= f->hi[1]; j++)
        for (int i = f->lo[0]; i <= f->hi[0]; i++) {
          long c = IDX(s, i, j, k);
          long a = (i - f->lo[0]) + (long)f->ex[0] * ((j - f->lo[1]) + (long)f->ex[1] * (k - f->lo[2]));
          real rrho = C_(1.0) / s->rho[c];
          real tc13 = s->lam[c], tc66 = s->mu[c];
          real tc11 = tc13 + C_(2.0) * tc66, tc12 = tc13, tc22 = tc11, tc23 = tc13, tc33 = tc11,
               tc44 = tc66, tc55 = tc66;
          real zk = s->z[k], xs = s->xsin[i], xc = s->xcot[i];
          real b1, b1_1, ad, Dd;
          real **ps = f->W, **ph = f->h;
          if (ax == 0) {
            b1 = C_(1.0) / s->Bx[i]; b1_1 = b1 - C_(1.0); ad = s->Ax[i] + s->Dx[i]; Dd = s->Dx[i];
            real xi = s->xi_x[i];
            real DxTxx = opD(xf, W[TXX] + c, 1) * xi;
            DxTxx = (DxTxx + C_(2.0) * W[TXZ][c]) / zk;
            h[VX][c] = h[VX][c] + (b1_1 * DxTxx - b1 * ps[3][a]) * rrho;
            ph[3][a] = Dd * DxTxx - ad * ps[3][a];
            real DxTxy = opD(xf, W[TXY] + c, 1) * xi;
            DxTxy = (DxTxy + W[TYZ][c]) / zk;
            h[VY][c] = h[VY][c] + (b1_1 * DxTxy - b1 * ps[4][a]) * rrho;
            ph[4][a] = Dd * DxTxy - ad * ps[4][a];
            real DxTxz = opD(xf, W[TXZ] + c, 1) * xi;
            DxTxz = (DxTxz - W[TXX][c] + W[TZZ][c]) / zk;
            h[VZ][c] = h[VZ][c] + (b1_1 * DxTxz - b1 * ps[5][a]) * rrho;
            ph[5][a] = Dd * DxTxz - ad * ps[5][a];
            real DxVx = opD(xf, W[VX] + c, 1) * xi;
            DxVx = (DxVx + W[VZ][c]) / zk;
            h[TXX][c] = h[TXX][c] + b1_1 * tc11 * DxVx - b1 * tc11 * ps[0][a];
            h[TYY][c] = h[TYY][c] + b1_1 * tc12 * DxVx - b1 * tc12 * ps[0][a];
            h[TZZ][c] = h[TZZ][c] + b1_1 * tc13 * DxVx - b1 * tc13 * ps[0][a];
            ph[0][a] = Dd * DxVx - ad * ps[0][a];
            real DxVy = opD(xf, W[VY] + c, 1) * xi;
            DxVy = DxVy / zk;
            h[TXY][c] = h[TXY][c] + b1_1 * tc66 * DxVy - b1 * tc66 * ps[1][a];
            ph[1][a] = Dd * DxVy - ad * ps[1][a];
            real DxVz = opD(xf, W[VZ] + c, 1) * xi;
            DxVz = (DxVz - W[VX][c]) / zk;
            h[TXZ][c] = h[TXZ][c] + b1_1 * tc55 * DxVz - b1 * tc55 * ps[2][a];
            ph[2][a] = Dd * DxVz - ad * ps[2][a];
            if (s->freenode && k == s->nk2) {
              h[TXX][c] = h[TXX][c] - b1_1 * tc13 * tc13 / tc33 * DxVx + b1 * tc13 * tc13 / tc33 * ps[0][a];
              h[TYY][c] = h[TYY][c] - b1_1 * tc23 * tc13 / tc33 * DxVx + b1 * tc23 * tc13 / tc33 * ps[0][a];
            }
          } else if (ax == 1) {
            b1 = C_(1.0) / s->By[j]; b1_1 = b1 - C_(1.0); ad = s->Ay[j] + s->Dy[j]; Dd = s->Dy[j];
            real et = s->eta_y[j];
            real DyTxy = opD(yf, W[TXY] + c, sy) * et;
            DyTxy = g_mut == 3 ? (DyTxy / xs - (W[TXX][c] - W[TYY][c]) * xc + W[TXZ][c]) / zk
                                : (DyTxy / xs + (W[TXX][c] - W[TYY][c]) * xc + W[TXZ][c]) / zk;
            h[VX][c] = h[VX][c] + (b1_1 * DyTxy - b1 * ps[4][a]) * rrho;
            ph[4][a] = Dd * DyTxy - ad * ps[4][a];
            real DyTyy = opD(yf, W[TYY] + c, sy) * et;
            DyTyy = (DyTyy / xs + C_(2.0) * W[TXY][c] * xc + C_(2.0) * W[TYZ][c]) / zk;
            h[VY][c] = h[VY][c] + (b1_1 * DyTyy - b1 * ps[3][a]) * rrho;
            ph[3][a] = Dd * DyTyy - ad * ps[3][a];
            real DyTyz = opD(yf, W[TYZ] + c, sy) * et;
            DyTyz = (DyTyz / xs + W[TXZ][c] * xc + (W[TZZ][c] - W[TYY][c])) / zk;
            h[VZ][c] = h[VZ][c] + (b1_1 * DyTyz - b1 * ps[5][a]) * rrho;
            ph[5][a] = Dd * DyTyz - ad * ps[5][a];
            real DyVx = opD(yf, W[VX] + c, sy) * et;
            DyVx = (DyVx / xs - W[VY][c] * xc) / zk;
            h[TXY][c] = h[TXY][c] + b1_1 * tc66 * DyVx - b1 * tc66 * ps[0][a];
            ph[0][a] = Dd * DyVx - ad * ps[0][a];
            real DyVy = opD(yf, W[VY] + c, sy) * et;
            DyVy = (DyVy / xs + W[VX][c] * xc + W[VZ][c]) / zk;
            h[TXX][c] = h[TXX][c] + b1_1 * tc12 * DyVy - b1 * tc12 * ps[1][a];
            h[TYY][c] = h[TYY][c] + b1_1 * tc22 * DyVy - b1 * tc22 * ps[1][a];
            h[TZZ][c] = h[TZZ][c] + b1_1 * tc23 * DyVy - b1 * tc23 * ps[1][a];
            ph[1][a] = Dd * DyVy - ad * ps[1][a];
            real DyVz = opD(yf, W[VZ] + c, sy) * et;
            DyVz = (DyVz / xs - W[VY][c]) / zk;
            h[TYZ][c] = h[TYZ][c] + b1_1 * tc44 * DyVz - b1 * tc44 * ps[2][a];
            ph[2][a] = Dd * DyVz - ad * ps[2][a];
            if (s->freenode && k == s->nk2) {
              h[TXX][c] = h[TXX][c] - b1_1 * tc13 * tc23 / tc33 * DyVy + b1 * tc13 * tc23 / tc33 * ps[1][a];
              h[TYY][c] = h[TYY][c] - b1_1 * tc23 * tc23 / tc33 * DyVy + b1 * tc23 * tc23 / tc33 * ps[1][a];
            }
          } else {
            b1 = C_(1.0) / s->Bz[k]; b1_1 = b1 - C_(1.0); ad = s->Az[k] + s->Dz[k]; Dd = s->Dz[k];
            real ze = s->zeta_z[k];
            real DzTxz = opD(zf, W[TXZ] + c, sz) * ze;
            h[VX][c] = h[VX][c] + (b1_1 * DzTxz - b1 * ps[4][a]) * rrho;
            ph[4][a] = Dd * DzTxz - ad * ps[4][a];
            real DzTyz = opD(zf, W[TYZ] + c, sz) * ze;
            h[VY][c] = h[VY][c] + (b1_1 * DzTyz - b1 * ps[5][a]) * rrho;
            ph[5][a] = Dd * DzTyz - ad * ps[5][a];
            real DzTzz = opD(zf, W[TZZ] + c, sz) * ze;
            h[VZ][c] = h[VZ][c] + (b1_1 * DzTzz - b1 * ps[3][a]) * rrho;
            ph[3][a] = Dd * DzTzz - ad * ps[3][a];
            real DzVx = opD(zf, W[VX] + c, sz) * ze;
            h[TXZ][c] = h[TXZ][c] + b1_1 * tc55 * DzVx - b1 * tc55 * ps[0][a];
            ph[0][a] = Dd * DzVx - ad * ps[0][a];
            real DzVy = opD(zf, W[VY] + c, sz) * ze;
            h[TYZ][c] = h[TYZ][c] + b1_1 * tc44 * DzVy - b1 * tc44 * ps[1][a];
            ph[1][a] = Dd * DzVy - ad * ps[1][a];
            real DzVz = opD(zf, W[VZ] + c, sz) * ze;
            h[TXX][c] = h[TXX][c] + b1_1 * tc13 * DzVz - b1 * tc13 * ps[2][a];
            h[TYY][c] = h[TYY][c] + b1_1 * tc23 * DzVz - b1 * tc23 * ps[2][a];
            h[TZZ][c] = h[TZZ][c] + b1_1 * tc33 * DzVz - b1 * tc33 * ps[2][a];
            ph[2][a] = Dd * DzVz - ad * ps[2][a];
          }
        }
  }
}

/* ------------------------------------------------------------------------
 * a7: sources, code/srcF/mod_src.F90:207-414 (SrcSmooth + SrcSurface build).
 * Pinned by tests/test_oracle_src.py (Gaussian weights sum to one) and
 * tests/test_oracle_src_injection.py (one stage's terms through src_terms:
 * total injected force F*stf against exact spherical cell volumes, the cut
 * and renormalised spread below a free surface, surface traction times the
 * exact cell area = F*stf and its rate, moment sum = -M*rate), and by
 * Stokes' solutions (tests/test_oracle_analytic.py).
 * ---------------------------------------------------------------------- */
static real cell_volume(orc_solver *s, int i, int j, int k) {
  if (s->curv) {
    /* mode C: the radial extent along the column, the radius of the cell */
    long c = IDX(s, i, j, k);
    real r = s->r3[c];
    return (s->x[i + 1] - s->x[i - 1]) * (s->y[j + 1] - s->y[j - 1]) * (s->r3[c + s->sz] - s->r3[c - s->sz]) /
           C_(8.0) * (r * r) * s->xsin[i];
  }
  return (s->x[i + 1] - s->x[i - 1]) * (s->y[j + 1] - s->y[j - 1]) * (s->z[k + 1] - s->z[k - 1]) / C_(8.0) *
         (s->z[k] * s->z[k]) * s->xsin[i];
}
static void renorm_surface(orc_solver *s, real *normd, int sk) {
  if (s->freenode && sk + LENFD > s->nk2 && g_mut != 8) {
    real sum = 0;
    int kmax = s->nk2 - sk;
    for (int k = -LENFD; k <= kmax; k++)
      for (int q = 0; q < 49; q++) sum += normd[q + 49 * (k + 3)];
    for (int n = 0; n < 343; n++) normd[n] = normd[n] / sum;
  }
}
static void src_surface(orc_solver *s, int ntime, real tinc) {
  if (!s->freenode || s->nforce <= 0) return;
  real t = ((real)ntime + tinc) * s->stept;
  for (int n = 0; n < s->nforce; n++) {
    int i = s->force_indx[3 * n], j = s->force_indx[3 * n + 1], k = s->force_indx[3 * n + 2];
    for (int m = 0; m < s->ntw_force; m++) {
      real disp = src_stf(t - s->force_t0[n], s->frc_time[m], s->frc_freq[m], s->frc_stf_id, ntime, s->stept);
      real rate = src_svf(t - s->force_t0[n], s->frc_time[m], s->frc_freq[m], s->frc_stf_id, ntime, s->stept);
      long q = m + (long)s->ntw_force * n;
      real fx0 = s->Fx[q], fy0 = s->Fy[q], fz0 = s->Fz[q];
      if (i <= s->ni2 && i >= s->ni1 && j <= s->nj2 && j >= s->nj1 && k == s->nk2) {
        /* mode C: the projected area at the surface radius of the column
         * (sigma . N = F / S_projected: N's length cancels the slope's area) */
        real zs = s->curv ? s->r3[IDX(s, i, j, k)] : s->z[k];
        real S = (s->x[i + 1] - s->x[i - 1]) * (s->y[j + 1] - s->y[j - 1]) / C_(4.0) * (zs * zs) * s->xsin[i];
        long c2 = i + s->sy * j;
        if (g_mut == 9) S = S / C_(2.0);
        s->TxSrc[c2] = fx0 / S * disp; s->TySrc[c2] = fy0 / S * disp; s->TzSrc[c2] = fz0 / S * disp;
        s->VxSrc[c2] = fx0 / S * rate; s->VySrc[c2] = fy0 / S * rate; s->VzSrc[c2] = fz0 / S * rate;
      }
    }
  }
}
static void src_force(orc_solver *s, int ntime, real tinc) {
  if (s->nforce <= 0) return;
  real t = ((real)ntime + tinc) * s->stept;
  real normd[343];
  for (int n = 0; n < s->nforce; n++) {
    int si = s->force_indx[3 * n], sj = s->force_indx[3 * n + 1], sk = s->force_indx[3 * n + 2];
    cal_norm_delt(normd, s->force_shift[3 * n], s->force_shift[3 * n + 1], s->force_shift[3 * n + 2]);
    renorm_surface(s, normd, sk);
    for (int m = 0; m < s->ntw_force; m++) {
      real disp = src_stf(t - s->force_t0[n], s->frc_time[m], s->frc_freq[m], s->frc_stf_id, ntime, s->stept);
      long q = m + (long)s->ntw_force * n;
      real fx0 = s->Fx[q] * disp, fy0 = s->Fy[q] * disp, fz0 = s->Fz[q] * disp;
      for (int Lk = -LENFD; Lk <= LENFD; Lk++)
        for (int Lj = -LENFD; Lj <= LENFD; Lj++)
          for (int Li = -LENFD; Li <= LENFD; Li++) {
            int i = Li + si, j = Lj + sj, k = Lk + sk;
            real d = normd[(Li + 3) + 7 * (Lj + 3) + 49 * (Lk + 3)];
            if (i <= s->ni2 && i >= s->ni1 && j <= s->nj2 && j >= s->nj1 && k <= s->nk2 && k >= s->nk1) {
              long c = IDX(s, i, j, k);
              real V = cell_volume(s, i, j, k);
              V = C_(1.0) / (V * s->rho[c]) * d;
              real fx = fx0 * V, fy = fy0 * V, fz = fz0 * V;
              if (s->freenode && k == s->nk2 && g_mut != 6) continue;
              if (g_mut == 7) { fx *= C_(2.0); fy *= C_(2.0); fz *= C_(2.0); }
              s->h[VX][c] = s->h[VX][c] + fx; s->h[VY][c] = s->h[VY][c] + fy; s->h[VZ][c] = s->h[VZ][c] + fz;
            }
          }
    }
  }
}
static void src_stress(orc_solver *s, int ntime, real tinc) {
  if (s->nmoment <= 0) return;
  real t = ((real)ntime + tinc) * s->stept;
  real normd[343];
  for (int n = 0; n < s->nmoment; n++) {
    int si = s->moment_indx[3 * n], sj = s->moment_indx[3 * n + 1], sk = s->moment_indx[3 * n + 2];
    cal_norm_delt(normd, s->moment_shift[3 * n], s->moment_shift[3 * n + 1], s->moment_shift[3 * n + 2]);
    renorm_surface(s, normd, sk);
    for (int m = 0; m < s->ntw_moment; m++) {
      real rate = src_svf(t - s->moment_t0[n], s->mom_time[m], s->mom_freq[m], s->mom_stf_id, ntime, s->stept);
      long q = m + (long)s->ntw_moment * n;
      real Mxx = s->Mxx[q], Myy = s->Myy[q], Mzz = s->Mzz[q], Mxy = s->Mxy[q], Mxz = s->Mxz[q], Myz = s->Myz[q];
      for (int Lk = -LENFD; Lk <= LENFD; Lk++)
        for (int Lj = -LENFD; Lj <= LENFD; Lj++)
          for (int Li = -LENFD; Li <= LENFD; Li++) {
            int i = Li + si, j = Lj + sj, k = Lk + sk;
            real d = normd[(Li + 3) + 7 * (Lj + 3) + 49 * (Lk + 3)];
            if (i <= s->ni2 && i >= s->ni1 && j <= s->nj2 && j >= s->nj1 && k <= s->nk2 && k >= s->nk1) {
              long c = IDX(s, i, j, k);
              real V = cell_volume(s, i, j, k);
              V = d * rate / V;
              s->h[TXX][c] = s->h[TXX][c] - Mxx * V; s->h[TYY][c] = s->h[TYY][c] - Myy * V;
              s->h[TZZ][c] = s->h[TZZ][c] - Mzz * V; s->h[TXY][c] = s->h[TXY][c] - Mxy * V;
              s->h[TXZ][c] = s->h[TXZ][c] - Mxz * V; s->h[TYZ][c] = s->h[TYZ][c] - Myz * V;
            }
          }
    }
  }
}

/* The source terms of one stage alone (test entry): h after src_surface,
 * src_stress and src_force on a zero h (hout: 9*ncell, field-major), and
 * the surface traction / traction-rate planes TxSrc TySrc TzSrc VxSrc VySrc
 * VzSrc (surf: 6*nx*ny) that free_timg and the traction condition read. */
void ORC(src_terms)(orc_solver *s, int ntime, real tinc, real *hout, real *surf) {
  long n2 = (long)s->nx * s->ny;
  for (int f = 0; f < NFIELD; f++) memset(s->h[f], 0, sizeof(real) * s->ncell);
  real *sp[6] = {s->TxSrc, s->TySrc, s->TzSrc, s->VxSrc, s->VySrc, s->VzSrc};
  for (int q = 0; q < 6; q++) memset(sp[q], 0, sizeof(real) * n2);
  src_surface(s, ntime, tinc);
  src_stress(s, ntime, tinc);
  src_force(s, ntime, tinc);
  for (int f = 0; f < NFIELD; f++) memcpy(hout + (long)f * s->ncell, s->h[f], sizeof(real) * s->ncell);
  for (int q = 0; q < 6; q++) memcpy(surf + q * n2, sp[q], sizeof(real) * n2);
}

/* ------------------------------------------------------------------------
 * a8: one time step of the 8-block cycle, code/srcF/seis3d_wave_8op_p3.F90:92-594.
 * Block b = ntime mod 8 runs the patterns P, ~P, P, ~P with
 * P = BBB, FFB, FFF, BBF, BFB, FBB, FBF, BFF (1 = forward).
 * ---------------------------------------------------------------------- */
static const int BLOCK_PAT[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                                    {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
void ORC(stage_pattern)(int ntime, int stage, int *xyz) {
  const int *p = BLOCK_PAT[((ntime % 8) + 8) % 8];
  int flip = stage & 1;
  for (int d = 0; d < 3; d++) xyz[d] = flip ? !p[d] : p[d];
}

void ORC(step)(orc_solver *s, int ntime) {
  macdrp_syn(s);
  abs_syn(s);
  for (int st = 0; st < 4; st++) {
    real tinc = st == 0 ? C_(0.0) : rka(st - 1);
    if (g_mut == 5) tinc += C_(0.25);
    int pat[3];
    ORC(stage_pattern)(ntime - s->cycle_origin, st, pat);
    src_surface(s, ntime, tinc);
    free_timg(s);
    if (s->curv) op_sweep_c(s, pat[0], pat[1], pat[2]);
    else op_sweep(s, pat[0], pat[1], pat[2]);
    abs_vsymm(s);
    abs_correct(s, pat[0], pat[1], pat[2]);
    src_stress(s, ntime, tinc);
    src_force(s, ntime, tinc);
    int kind = st == 0 ? 0 : (st == 3 ? 2 : 1);
    real a = rka(st) * s->stept, b = rkb(st) * s->stept;
    macdrp_rk(s, kind, a, b);
    abs_rk(s, kind, a, b);
  }
}

/* run steps ntime0 .. ntime0+nsteps-1; after each step (ntime+1) sample the
 * receivers when mod(ntime+1, tinv)==0 (io_seismo_put, mod_io.F90:461-497).
 * rec: [nrec_records][npt][9]; returns number of records written. */
int ORC(run)(orc_solver *s, int ntime0, int nsteps, real *rec) {
  int nrec = 0;
  for (int n = 0; n < nsteps; n++) {
    int nt = ntime0 + n;
    ORC(step)(s, nt);
    int t1 = nt + 1;
    if (s->npt > 0 && s->pt_tinv > 0 && t1 % s->pt_tinv == 0 && rec) {
      for (int p = 0; p < s->npt; p++) {
        long c = IDX(s, s->pt_indx[3 * p], s->pt_indx[3 * p + 1], s->pt_indx[3 * p + 2]);
        for (int f = 0; f < NFIELD; f++) rec[((long)nrec * s->npt + p) * NFIELD + f] = s->W[f][c];
      }
      nrec++;
    }
  }
  return nrec;
}

/* a restarted run re-enters the cycle at its first block, seis3d_wave_8op_p3.F90:72-95 */
void ORC(set_cycle_origin)(orc_solver *s, int origin) { s->cycle_origin = origin; }
void ORC(get_field)(orc_solver *s, int f, real *out) { memcpy(out, s->W[f], sizeof(real) * s->ncell); }
void ORC(set_field)(orc_solver *s, int f, const real *in) { memcpy(s->W[f], in, sizeof(real) * s->ncell); }
long ORC(ncell)(orc_solver *s) { return s->ncell; }

/* ========================================================================
 * Sensitivity-kernel time integral, code/srcF.kernel/SI_ker_sta.F90.
 * Series layout: [n][point] (point fastest), n = 0..knt-1.
 * ====================================================================== */

/* a12: stress2strain, SI_ker_sta.F90:1185-1211.
 * Pinned by tests/test_oracle_kernel.py (Hooke round trip). */
void ORC(stress2strain)(long npt, int knt, real *Exx, real *Eyy, real *Ezz, real *Exy, real *Exz,
                        real *Eyz, const real *lam, const real *miu) {
  for (int n = 0; n < knt; n++)
    for (long p = 0; p < npt; p++) {
      long q = (long)n * npt + p;
      real l = lam[p], u = miu[p];
      real E1 = (l + u) / (u * (C_(3.0) * l + C_(2.0) * u));
      real E2 = -l / (C_(2.0) * u * (C_(3.0) * l + C_(2.0) * u));
      real E3 = C_(1.0) / u;
      real E0 = E2 * (Exx[q] + Eyy[q] + Ezz[q]);
      Exx[q] = E0 - (E2 - E1) * Exx[q];
      Eyy[q] = E0 - (E2 - E1) * Eyy[q];
      Ezz[q] = E0 - (E2 - E1) * Ezz[q];
      Exy[q] = C_(0.5) * E3 * Exy[q];
      Exz[q] = C_(0.5) * E3 * Exz[q];
      Eyz[q] = C_(0.5) * E3 * Eyz[q];
    }
}

/* a13: sgt2scatter3 for ONE receiver component, SI_ker_sta.F90:1257-1337.
 * R[6], S[6] are strain series (xx yy zz xy xz yz).  Outputs the alpha and
 * beta scattered series (reference writes them into ExxS/ExyS, EyyS/EyzS or
 * EzzS/ExzS depending on the component, :1299-1330).  Accumulation in real
 * (the reference accumulates in SP), sequential in m.
 * Pinned by tests/test_oracle_kernel.py (numpy FP64 convolution). */
void ORC(scatter1)(long npt, int knt, real dt, const real *const *R, const real *const *S, const real *lam,
                   const real *miu, real *outA, real *outB) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < npt; p++) {
    for (int n = 1; n <= knt; n++) {
      real x1 = 0, x2 = 0;
      for (int m = 1; m <= n - 1; m++) {
        long qr = (long)(m - 1) * npt + p, qs = (long)(n - m - 1) * npt + p;
        x1 = x1 + C_(1.0) * dt * (R[0][qr] + R[1][qr] + R[2][qr]) * (S[0][qs] + S[1][qs] + S[2][qs]);
        x2 = x2 + C_(1.0) * dt *
                      (R[0][qr] * S[0][qs] + R[1][qr] * S[1][qs] + R[2][qr] * S[2][qs] +
                       C_(2.0) * R[3][qr] * S[3][qs] + C_(2.0) * R[4][qr] * S[4][qs] + C_(2.0) * R[5][qr] * S[5][qs]);
      }
      long qo = (long)(n - 1) * npt + p;
      outA[qo] = x1 * C_(2.0) * (lam[p] + C_(2.0) * miu[p]);
      outB[qo] = C_(2.0) * (x2 - x1) * C_(2.0) * miu[p];
    }
  }
}

/* a14: IIR filter, SI_ker_sta.F90:1547-1557 (double precision always) */
void orc_filter(int NOD, int NEL, const double *b, const double *a, const double *x, double *y) {
  for (int n = 0; n < NEL; n++) {
    double v = b[0] * x[n];
    int mm = (NOD + 1) < (n + 1) ? (NOD + 1) : (n + 1);
    for (int m = 2; m <= mm; m++) v = v + b[m - 1] * x[n - m + 1] - a[m - 1] * y[n - m + 1];
    y[n] = v / a[0];
  }
}
/* filtfilt, :1538-1545 */
static void orc_filtfilt(int NOD, int NEL, const double *b, const double *a, double *x, double *y, double *w) {
  orc_filter(NOD, NEL, b, a, x, y);
  for (int n = 0; n < NEL; n++) w[n] = y[NEL - 1 - n];
  orc_filter(NOD, NEL, b, a, w, y);
  for (int n = 0; n < NEL / 2; n++) { double tmp = y[n]; y[n] = y[NEL - 1 - n]; y[NEL - 1 - n] = tmp; }
}
/* scatter_filter (mode 1, P1: causal, zero front pad, odd end pad) and
 * scatter_filtfilt (mode 2, P2: odd padding both ends), :1433-1475.
 * w0, w: [knt][npt]; NOD1 coefficients. */
void ORC(series_filter)(int mode, long npt, int knt, int NOD1, const double *b, const double *a, const real *w0,
                        real *w) {
  int NOD = NOD1 - 1, NFACT = 3 * NOD, NEL = knt + 2 * NFACT;
#pragma omp parallel
  {
    double *vx = (double *)calloc(NEL, sizeof(double)), *vy = (double *)calloc(NEL, sizeof(double)),
           *vw = (double *)calloc(NEL, sizeof(double));
#pragma omp for schedule(static)
    for (long p = 0; p < npt; p++) {
      for (int n = 0; n < knt; n++) vx[NFACT + n] = (double)w0[(long)n * npt + p];
      /* Fortran 1-based: vecx(n) n=1..NFACT; vecx(NFACT+1) is the first sample */
      for (int n = 1; n <= NFACT; n++) {
        if (mode == 2) vx[n - 1] = 2 * vx[NFACT] - vx[2 * NFACT + 1 - (n - 1) - 1];
        else vx[n - 1] = 0.0;
        vx[NEL - NFACT + n - 1] = 2 * vx[NEL - NFACT - 1] - vx[NEL - NFACT - n - 1];
      }
      if (mode == 2) orc_filtfilt(NOD, NEL, b, a, vx, vy, vw);
      else orc_filter(NOD, NEL, b, a, vx, vy);
      for (int n = 0; n < knt; n++) w[(long)n * npt + p] = (real)vy[NFACT + n];
    }
    free(vx); free(vy); free(vw);
  }
}

/* a15: scatter2kernel, SI_ker_sta.F90:1375-1431.  nt1, nt2 are 1-based
 * inclusive sample indices.  kdt is the series interval (the reference uses
 * its global kdt in the end-point terms of the normalisation, :1392, and the
 * argument dt elsewhere; both are the same value in every call, :796-798).
 * Pinned by tests/test_oracle_kernel.py (trapezoid on polynomials). */
void ORC(scatter2kernel)(long npt, int knt, int nt1, int nt2, real dt, real M0, const real *V, const real *U,
                         const real *Sa, const real *Sb, real *Kap, real *Kaq, real *Kbp, real *Kbq) {
  (void)knt;
  const real *Vv = V - 1, *Uu = U - 1; /* 1-based views */
  real V0 = (C_(0.5) * Vv[nt1] * Vv[nt1] + C_(0.5) * Vv[nt2] * Vv[nt2]) * dt;
  real U0 = (C_(0.5) * Uu[nt1] * Uu[nt1] + C_(0.5) * Uu[nt2] * Uu[nt2]) * dt;
  for (int n = nt1 + 1; n <= nt2 - 1; n++) { V0 = V0 + Vv[n] * Vv[n] * dt; U0 = U0 + Uu[n] * Uu[n] * dt; }
  V0 = V0 * M0; U0 = U0 * M0;
  for (long p = 0; p < npt; p++) {
#define SA(n) Sa[(long)((n)-1) * npt + p]
#define SB(n) Sb[(long)((n)-1) * npt + p]
    real kap = C_(0.5) * dt * (Vv[nt1] * SA(nt1) + Vv[nt2] * SA(nt2));
    real kaq = C_(0.5) * dt * (Uu[nt1] * SA(nt1) + Uu[nt2] * SA(nt2));
    real kbp = C_(0.5) * dt * (Vv[nt1] * SB(nt1) + Vv[nt2] * SB(nt2));
    real kbq = C_(0.5) * dt * (Uu[nt1] * SB(nt1) + Uu[nt2] * SB(nt2));
    for (int n = nt1 + 1; n <= nt2 - 1; n++) {
      kap = kap + Vv[n] * SA(n) * dt; kaq = kaq + Uu[n] * SA(n) * dt;
      kbp = kbp + Vv[n] * SB(n) * dt; kbq = kbq + Uu[n] * SB(n) * dt;
    }
#undef SA
#undef SB
    Kap[p] = kap / V0; Kaq[p] = kaq / U0; Kbp[p] = kbp / V0; Kbq[p] = kbq / U0;
  }
}

/* scatter_rotate, SI_ker_sta.F90:1354-1373 (rot is column-major 3x3) */
void ORC(scatter_rotate)(long npt, int knt, real *X, real *Y, real *Z, const real *rot) {
  for (long q = 0; q < (long)knt * npt; q++) {
    real v1 = X[q], v2 = Y[q], v3 = Z[q];
    X[q] = v1 * rot[0] + v2 * rot[1] + v3 * rot[2];
    Y[q] = v1 * rot[3] + v2 * rot[4] + v3 * rot[5];
    Z[q] = v1 * rot[6] + v2 * rot[7] + v3 * rot[8];
  }
}

/* ========================================================================
 * Reduced-precision SGT storage (no reference counterpart: the reference
 * stores FP32).  Per 16^3 block of snapshot points and per slab: amax =
 * largest |x|; BF16 codes of x/amax (scale amax) or FP16 codes of
 * x*32768/amax (scale amax/32768), round to nearest even; x' = code*scale.
 * A block with amax below 2^-100 is stored as zeros with scale 0 (the
 * factor would overflow).  Always float arithmetic, whatever REAL is.  Pinned by
 * tests/test_oracle_sgt_codec.py: conversions against numpy float16 and
 * torch bfloat16, block scales against numpy maxima, restore error bounds.
 * ====================================================================== */
#include <stdint.h>
static uint16_t c_bf16_rne(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float c_bf16_val(uint16_t c) {
  uint32_t u = (uint32_t)c << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint16_t c_fp16_rne(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  const uint32_t sign = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to infinity */
  const uint32_t e = a >> 23, m = (a & 0x7fffffu) | 0x800000u;
  if (e < 113) { /* result subnormal (or zero): k units of 2^-24 */
    if (e < 102) return (uint16_t)sign;
    const uint32_t sh = 126 - e, k = m >> sh, rem = m & ((1u << sh) - 1), half = 1u << (sh - 1);
    const uint32_t r = k + ((rem > half || (rem == half && (k & 1u))) ? 1u : 0u);
    return (uint16_t)(sign | r);
  }
  const uint32_t k = ((e - 112) << 10) | ((a & 0x7fffffu) >> 13), rem = a & 0x1fffu;
  const uint32_t r = k + ((rem > 0x1000u || (rem == 0x1000u && (k & 1u))) ? 1u : 0u);
  return (uint16_t)(sign | r);
}
static float c_fp16_val(uint16_t c) {
  const uint32_t sign = (uint32_t)(c & 0x8000u) << 16, e = (c >> 10) & 0x1fu, m = c & 0x3ffu;
  float f;
  if (e == 0) {
    f = (float)m * 5.9604644775390625e-08f; /* m * 2^-24 */
    return sign ? -f : f;
  }
  uint32_t u = sign | (e == 31 ? (0xffu << 23) | (m << 13) : ((e + 112) << 23) | (m << 13));
  memcpy(&f, &u, 4);
  return f;
}
void ORC(sgt_encode)(int fmt, int c0, int c1, int c2, long nslab, const float *x, uint16_t *codes, float *scales) {
  const int nb0 = (c0 + 15) / 16, nb1 = (c1 + 15) / 16, nb2 = (c2 + 15) / 16;
  const long npts = (long)c0 * c1 * c2, nblk = (long)nb0 * nb1 * nb2;
  for (long sl = 0; sl < nslab; sl++)
    for (int b2 = 0; b2 < nb2; b2++)
      for (int b1 = 0; b1 < nb1; b1++)
        for (int b0 = 0; b0 < nb0; b0++) {
          const long blk = b0 + (long)nb0 * (b1 + (long)nb1 * b2);
          float amax = 0.0f;
          for (int d = b2 * 16; d < c2 && d < b2 * 16 + 16; d++)
            for (int b = b1 * 16; b < c1 && b < b1 * 16 + 16; b++)
              for (int a = b0 * 16; a < c0 && a < b0 * 16 + 16; a++)
                amax = fmaxf(amax, fabsf(x[sl * npts + a + (long)c0 * (b + (long)c1 * d)]));
          /* maxima below 2^-100 (where 32768 / amax would overflow) store zeros */
          const int tiny = !(amax >= 7.888609052210118e-31f);
          const float inv = tiny ? 0.0f : (fmt == 2 ? 32768.0f / amax : 1.0f / amax);
          scales[sl * nblk + blk] = tiny ? 0.0f : (fmt == 2 ? amax * (1.0f / 32768.0f) : amax);
          for (int d = b2 * 16; d < c2 && d < b2 * 16 + 16; d++)
            for (int b = b1 * 16; b < c1 && b < b1 * 16 + 16; b++)
              for (int a = b0 * 16; a < c0 && a < b0 * 16 + 16; a++) {
                const long p = sl * npts + a + (long)c0 * (b + (long)c1 * d);
                const float q = x[p] * inv;
                codes[p] = fmt == 2 ? c_fp16_rne(q) : c_bf16_rne(q);
              }
        }
}
void ORC(sgt_decode)(int fmt, int c0, int c1, int c2, long nslab, const uint16_t *codes, const float *scales, float *out) {
  const int nb0 = (c0 + 15) / 16, nb1 = (c1 + 15) / 16, nb2 = (c2 + 15) / 16;
  const long npts = (long)c0 * c1 * c2, nblk = (long)nb0 * nb1 * nb2;
  for (long sl = 0; sl < nslab; sl++)
    for (long p = 0; p < npts; p++) {
      const int a = (int)(p % c0), b = (int)((p / c0) % c1), d = (int)(p / ((long)c0 * c1));
      const long blk = a / 16 + (long)nb0 * (b / 16 + (long)nb1 * (d / 16));
      const uint16_t c = codes[sl * npts + p];
      out[sl * npts + p] = (fmt == 2 ? c_fp16_val(c) : c_bf16_val(c)) * scales[sl * nblk + blk];
    }
}
/* single-value conversions, exported for the codec pins */
uint16_t ORC(f32_to_bf16)(float x) { return c_bf16_rne(x); }
uint16_t ORC(f32_to_fp16)(float x) { return c_fp16_rne(x); }
float ORC(fp16_to_f32)(uint16_t c) { return c_fp16_val(c); }

/* ========================================================================
 * Kernel variants (C14) and re-parameterisation (C15).
 * ====================================================================== */
/* sgt2scatter3 with lambda/mu outputs, SI_ker_sta_lamu.F90:1290-1336:
 * outA = x1*lambda, outB = 2*x2*mu.  Pinned by
 * tests/test_oracle_kernel_variants.py (numpy FP64 convolution). */
void ORC(scatter1_lamu)(long npt, int knt, real dt, const real *const *R, const real *const *S, const real *lam,
                        const real *miu, real *outA, real *outB) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < npt; p++) {
    for (int n = 1; n <= knt; n++) {
      real x1 = 0, x2 = 0;
      for (int m = 1; m <= n - 1; m++) {
        long qr = (long)(m - 1) * npt + p, qs = (long)(n - m - 1) * npt + p;
        x1 = x1 + C_(1.0) * dt * (R[0][qr] + R[1][qr] + R[2][qr]) * (S[0][qs] + S[1][qs] + S[2][qs]);
        x2 = x2 + C_(1.0) * dt *
                      (R[0][qr] * S[0][qs] + R[1][qr] * S[1][qs] + R[2][qr] * S[2][qs] +
                       C_(2.0) * R[3][qr] * S[3][qs] + C_(2.0) * R[4][qr] * S[4][qs] + C_(2.0) * R[5][qr] * S[5][qs]);
      }
      long qo = (long)(n - 1) * npt + p;
      outA[qo] = x1 * lam[p];
      outB[qo] = C_(2.0) * x2 * miu[p];
    }
  }
}
/* vel2scatter3, SI_ker_sta_rho.F90:1478-1552 (one receiver component):
 * R[3], S[3] velocity series; out = rho * dt * sum (sum_c R_c)(m) (sum_c S_c)(n-m). */
void ORC(vscatter1)(long npt, int knt, real dt, const real *const *R, const real *const *S, const real *rho,
                    real *out) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < npt; p++) {
    for (int n = 1; n <= knt; n++) {
      real x1 = 0;
      for (int m = 1; m <= n - 1; m++) {
        long qr = (long)(m - 1) * npt + p, qs = (long)(n - m - 1) * npt + p;
        x1 = x1 + C_(1.0) * dt * (R[0][qr] + R[1][qr] + R[2][qr]) * (S[0][qs] + S[1][qs] + S[2][qs]);
      }
      out[(long)(n - 1) * npt + p] = x1 * C_(1.0) * rho[p];
    }
  }
}
/* eoshift along time of one [knt][npt] series, Fortran semantics (shift > 0:
 * out(n) = in(n+shift), zero beyond the end), SI_ker_EQ.F90:574-587. */
void ORC(eoshift)(long npt, int knt, int shift, const real *in, real *out) {
  for (int n = 0; n < knt; n++)
    for (long p = 0; p < npt; p++) {
      int s = n + shift;
      out[(long)n * npt + p] = (s >= 0 && s < knt) ? in[(long)s * npt + p] : C_(0.0);
    }
}
/* re-parameterisation of (a, b) kernel pairs, FP32 order of operations of
 * SI_ker_lamu2kamu.F90:823-830 (mode 0), SI_ker_lamu2vpvs.F90:823-830
 * (mode 1), SI_ker_vpvs2kamu.F90:823-831 (mode 2).  Pinned by the
 * chain-rule invariance test in tests/test_oracle_kernel_variants.py. */
void ORC(reparam)(int mode, long npt, const real *lam, const real *mu, real *a, real *b) {
  for (long p = 0; p < npt; p++) {
    real l = lam[p], u = mu[p], x = a[p], y = b[p];
    if (mode == 0) {
      x = x / l; y = y / u / C_(2.0);
      y = (C_(2.0) * y - C_(2.0) * x / C_(3.0)) * u;
      x = x * (l + C_(2.0) * u / C_(3.0));
    } else if (mode == 1) {
      x = x / l; y = y / u / C_(2.0);
      y = C_(2.0) * (y - x) * C_(2.0) * u;
      x = x * C_(2.0) * (l + C_(2.0) * u);
    } else {
      x = x / (C_(2.0) * (l + C_(2.0) * u));
      y = y / u / C_(4.0) + x;
      y = (C_(2.0) * y - C_(2.0) * x / C_(3.0)) * u;
      x = x * (l + C_(2.0) * u / C_(3.0));
    }
    a[p] = x; b[p] = y;
  }
}

/* ========================================================================
 * Inversion (C15/C16): block assembly and LSQR, always in double.
 * ====================================================================== */
/* SI_ker2assm.F90:300-320: B(n) = sum over block n's pixels of K * volume */
void ORC(assemble)(long nblk, const long *ptr, const long *pix, const float *vol, const float *K, double *B) {
  for (long n = 0; n < nblk; n++) {
    double s = 0.0;
    for (long q = ptr[n]; q < ptr[n + 1]; q++) s += (double)K[pix[q]] * (double)vol[q];
    B[n] = s;
  }
}
/* slsqr + lgtsol, inv.LSQR/mod_LSQR.f90:142-292, on CSR rows (lldrow's
 * ra/ja/na storage with 0-based columns).  Pinned by
 * tests/test_oracle_inversion.py against numpy.linalg.lstsq. */
static double l_dot(const double *a, const double *b, long n) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static double l_nrliz(double *a, long n) {
  double v0 = sqrt(l_dot(a, a, n));
  if (fabs(v0) <= 1.0e-20) return v0;
  for (long i = 0; i < n; i++) a[i] = a[i] / v0;
  return v0;
}
long ORC(lsqr)(long m, long n, const long *ptr, const long *col, const double *val, const double *b, double atol,
               double btol, long maxit, double *x, double *err, double *res) {
  double *u = calloc(m, sizeof(double)), *v = calloc(n, sizeof(double)), *w = calloc(n, sizeof(double)),
         *q = calloc(m > n ? m : n, sizeof(double)), *sig = calloc(n, sizeof(double));
  double anorm = 0.0;
  for (long k = 0; k < ptr[m]; k++) anorm += val[k] * val[k];
  anorm = sqrt(anorm);
  for (long i = 0; i < m; i++) u[i] = b[i];
  double beta = l_nrliz(u, m);
  for (long i = 0; i < n; i++) v[i] = 0.0;
  for (long r = 0; r < m; r++)
    for (long k = ptr[r]; k < ptr[r + 1]; k++) v[col[k]] += val[k] * u[r];
  double alpha = l_nrliz(v, n);
  for (long i = 0; i < n; i++) { w[i] = v[i]; x[i] = 0.0; sig[i] = 0.0; }
  double phibar = beta, rhobar = alpha, res2 = 0.0;
  long itct = 0;
  for (;;) {
    for (long r = 0; r < m; r++) {
      double y0 = 0.0;
      for (long k = ptr[r]; k < ptr[r + 1]; k++) y0 += val[k] * v[col[k]];
      q[r] = y0;
    }
    for (long i = 0; i < m; i++) u[i] = q[i] - u[i] * alpha;
    beta = l_nrliz(u, m);
    for (long i = 0; i < n; i++) q[i] = 0.0;
    for (long r = 0; r < m; r++)
      for (long k = ptr[r]; k < ptr[r + 1]; k++) q[col[k]] += val[k] * u[r];
    for (long i = 0; i < n; i++) v[i] = q[i] - v[i] * beta;
    alpha = l_nrliz(v, n);
    double rho = sqrt(rhobar * rhobar + beta * beta);
    if (rho == 0.0) break;
    double c = rhobar / rho, s = beta / rho, theta = s * alpha;
    rhobar = -c * alpha;
    double phi = c * phibar;
    phibar = s * phibar;
    for (long i = 0; i < n; i++) {
      double t = w[i] / rho;
      x[i] = x[i] + phi * t;
      w[i] = v[i] - theta * t;
      sig[i] = sig[i] + t * t;
    }
    double test1 = phibar / anorm / sqrt(l_dot(x, x, n));
    double test2 = alpha * fabs(c) / anorm;
    itct++;
    if (test1 <= atol || test2 <= btol || itct == maxit) {
      res2 = phibar;
      break;
    }
  }
  double temp = res2 / (double)(m - n > 1 ? m - n : 1);
  if (err)
    for (long i = 0; i < n; i++) err[i] = sqrt(temp * sig[i]);
  if (res) *res = res2;
  free(u); free(v); free(w); free(q); free(sig);
  return itct;
}
