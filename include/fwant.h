/*
 * fwant.h -- C ABI of the MI355X-native FWANT solver and sensitivity-kernel
 * library (libfwant.so).  Plain pointers and sizes only; every function
 * returns FW_OK (0) or a negative status, with a message in fw_last_error().
 *
 * Conventions shared with the reference Fortran code:
 *  - 3-D arrays are column-major (i fastest) over the ghosted index space
 *    nx = ni+6, ny = nj+6, nz = nk+6 (code/srcF/mod_para.F90:79-82).
 *  - Grid indices passed through this ABI are the reference's 1-based local
 *    indices (ghost offset included), exactly as stored in source_mpi*.nc
 *    and station_mpi*.nc.
 *  - Field order: Vx Vy Vz Txx Tyy Tzz Txy Txz Tyz (seismogram variable
 *    order, code/srcF/mod_io.F90:436-444).
 *
 * One host thread per handle.  Each handle owns one HIP stream; calls that
 * launch work are stream-ordered and return before the work completes
 * unless stated; fw_sync() waits.  Host output buffers are owned by the
 * caller.
 */
#ifndef FWANT_H
#define FWANT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FW_OK 0
#define FW_ERR_ARG (-1)
#define FW_ERR_HIP (-2)
#define FW_ERR_STATE (-3)
#define FW_ERR_NOMEM (-4)

enum { FW_VX = 0, FW_VY, FW_VZ, FW_TXX, FW_TYY, FW_TZZ, FW_TXY, FW_TXZ, FW_TYZ, FW_NFIELD };

/* source-time-function ids, code/srcF/mod_src.F90:44-52 */
enum {
  FW_SVF_BELL = 100, FW_SVF_TRIANGLE = 110, FW_SVF_GAUSS = 120, FW_STF_RICKER = 200,
  FW_STF_BSHIFT = 210, FW_STF_GAUSS = 220, FW_STF_BELL = 230, FW_STF_STEP = 240, FW_STF_DELTA = 250
};

typedef struct fw_solver_s *fw_handle;

/* Solver configuration: the SeisFD3D.conf keys that shape the solve.
 *   ni nj nk stept         code/srcF/mod_para.F90:75-84
 *   freenode               top rank owns the free surface, mod_mpi.F90:206-216
 *   abs_number, abs_velocity, CFS_bmax, CFS_amax (x1 x2 y1 y2 z1 z2)
 *                          code/srcF/mod_abs_npml.F90:160-169
 * A z2 absorbing layer clears freenode, as abs_init does (mod_abs_npml.F90:182-184). */
typedef struct fw_config {
  int ni, nj, nk;
  float stept;
  int freenode;
  int abs_number[6];
  float abs_velocity[6];
  float cfs_bmax[6];
  float cfs_amax[6];
} fw_config;

/* Replaces macdrp_init + the module allocations (code/srcF/mod_macdrp.F90:124-218)
 * and the para/grid/media allocations of the driver (seis3d_wave_8op_p3.F90:40-66). */
int fw_create(const fw_config *cfg, fw_handle *out);
/* Replaces macdrp_destroy / abs_destroy / src_destroy (mod_macdrp.F90:219-227). */
int fw_destroy(fw_handle h);
/* Last error message of this handle (h may be NULL: last error of fw_create). */
const char *fw_last_error(fw_handle h);
/* Ghosted extents and the device pitch used for 3-D arrays. */
int fw_dims(fw_handle h, int *nx, int *ny, int *nz, int *freenode);

/* Replaces grid_coord_import + grid_metric_import (code/srcF/mod_grid.F90:96-114):
 * 1-D coordinate and metric vectors of lengths nx (x, xsin, xcot, xi_x),
 * ny (y, eta_y), nz (z, zeta_z).  Copied. */
int fw_set_grid(fw_handle h, const float *x, const float *y, const float *z, const float *xsin,
                const float *xcot, const float *xi_x, const float *eta_y, const float *zeta_z);
/* Replaces seis3d_metric (code/srcF/seis3d_metric.F90:82-112): builds xsin, xcot,
 * xi_x, eta_y, zeta_z on the GPU in FP64 from double coordinates and installs
 * them (stored FP32).  If the out_* pointers are non-NULL the FP32 metric is
 * copied back (lengths as in fw_set_grid). */
int fw_build_metric(fw_handle h, const double *x, const double *y, const double *z, float *out_xsin,
                    float *out_xcot, float *out_xi_x, float *out_eta_y, float *out_zeta_z);
/* Mode C: a terrain-following grid under topography.  No counterpart in the
 * reference, whose grid is a spherical shell with the 1-D metric above (its
 * coord files hold z(K) only); this is the 3-D extension of seis3d_metric
 * (code/srcF/seis3d_metric.F90:82-112) for the "curvilinear" workloads.
 * Node (i, j, k) sits at colatitude x[i], longitude y[j] and radius
 * r[i + nx*(j + ny*k)] (ghosted, column-major, metres, FP64).  Builds on the
 * GPU in FP64 the slopes p = -xi_x r_xi / r, q = -eta_y r_eta / (r sin x),
 * zeta_r = 1 / r_zeta and 1/r (stored FP32, with the one-sided zeta
 * derivatives of p and q), and switches the solver to the mode C scheme:
 * derivatives d/dtheta = xi_x d_xi + zeta_theta d_zeta etc., traction
 * imaging of sigma . (p, q, 1) at the free surface, source volumes and
 * surface areas at the nodes' radii.  Call after fw_set_grid /
 * fw_build_metric (which give xi_x, eta_y, xsin, xcot) and before fw_init.
 * fw_init rejects a grid that is not flat (r = z(k)) on every CFS-PML slab
 * cell and 3 cells beyond, and fluid cells.  A flat r (r = z(k)) runs the
 * spherical scheme up to rounding.  out_* (nx*ny*nz each, dense) may be
 * NULL; else p, q, zeta_r, 1/r are copied back. */
int fw_build_metric3d(fw_handle h, const double *x, const double *y, const double *r, float *out_p,
                      float *out_q, float *out_zeta_r, float *out_rr);
/* Replaces media_import (code/srcF/mod_media.F90:241-255): ghosted nx*ny*nz arrays. */
int fw_set_media(fw_handle h, const float *lambda, const float *mu, const float *rho);

/* Replaces src_import for force sources (code/srcF/mod_src.F90:168-185):
 *   indx[3*n] (1-based local), shift[3*n], t0[n], stf_time[ntw], stf_freq[ntw],
 *   Fx/Fy/Fz[ntw*n] (window index fastest).  n = 0: no force sources; the
 *   arrays may be NULL and ntw is ignored. */
int fw_set_force(fw_handle h, int n, int ntw, int stf_id, const int *indx, const float *shift,
                 const float *t0, const float *stf_time, const float *stf_freq, const float *Fx,
                 const float *Fy, const float *Fz);
/* Replaces src_import for moment sources (mod_src.F90:150-167):
 *   M6[6*ntw*n] = Mxx, Myy, Mzz, Mxy, Mxz, Myz blocks, each [ntw*n].  n = 0: no
 *   moment sources; the arrays may be NULL and ntw is ignored. */
int fw_set_moment(fw_handle h, int n, int ntw, int stf_id, const int *indx, const float *shift,
                  const float *t0, const float *stf_time, const float *stf_freq, const float *M6);
/* Replaces io_pt_import + io_seismo_init (code/srcF/mod_io.F90:160-176, 430-459):
 * n receivers at 1-based local indices indx[3*n], sampled every tinv steps;
 * device storage for max_records records of 9 values each is reserved. */
int fw_set_receivers(fw_handle h, int n, int tinv, const int *indx, int max_records);
/* Replaces io_snap_read/locate + the snapshot part of io_wave_export
 * (code/srcF/mod_io.F90:180-226, 505-645): a strided sub-volume subs/subc/subt
 * (1-based local start, counts, strides) of the fields in field_mask
 * (bit f = field f) is captured every tinv steps into device storage for
 * max_records records.  Returns its id in *snap_id. */
int fw_add_snapshot(fw_handle h, const int *subs, const int *subc, const int *subt, int tinv,
                    unsigned field_mask, int max_records, int *snap_id);

/* Replaces abs_init (code/srcF/mod_abs_npml.F90:135-300) and finishes set-up:
 * PML profiles, source tables, zero wavefield.  Must follow the set_* calls. */
int fw_init(fw_handle h);
/* Host copies of the PML profiles (A, B and D/B as abs_init stores them),
 * each of length nx+ny+nz (x part, then y, then z). */
int fw_get_pml_profiles(fw_handle h, float *A, float *B, float *D);

/* One time step of seis3d_wave_8op_p3.F90:97-157: the 4 RK stages with the
 * operator pattern of block (ntime mod 8) (seis3d_wave_8op_p3.F90:82-591),
 * then receiver and snapshot capture for step ntime+1.  Asynchronous. */
int fw_step(fw_handle h, int ntime);
/* One RK stage (stage 0..3) of step ntime with the operator pattern the
 * caller picks (xf, yf, zf: 1 = forward, 0 = backward), for drivers that keep
 * the reference's loop body (seis3d_wave_8op_p3.F90:97-157).  One call
 * replaces the stage's src_surface (mod_src.F90:373-414), macdrp_Lx?_Ly?_Lz?
 * (mod_macdrp.F90:790-2181, with free_timg), abs_Lx?_Ly?_Lz?
 * (mod_abs_npml.F90:1149-8955), src_stress / src_force (mod_src.F90:207-345)
 * and macdrp_RK_beg / _inn / _fin with abs_RK_* (mod_macdrp.F90:291-393,
 * mod_abs_npml.F90:738-1105): stage 0 is RK_beg at tinc 0, stages 1 and 2
 * RK_inn at tinc a1 / a2, stage 3 RK_fin at tinc a3, with the classical RK4
 * coefficients (firRKa, firRKb).  Stage 3 also captures receivers and
 * snapshots for step ntime+1 (io_seismo_put, io_wave_export).  The stages
 * of a step must come in order 0..3 (FW_ERR_STATE otherwise); fw_run over
 * the same steps with the 8-step cycle's patterns gives identical results.
 * Asynchronous. */
int fw_stage(fw_handle h, int ntime, int stage, int xf, int yf, int zf);
/* nsteps consecutive steps starting at ntime0. Asynchronous. */
int fw_run(fw_handle h, int ntime0, int nsteps);
/* Steps the reference driver runs for a given nt: 8*ceil((nt+1)/8)
 * (the exit test sits at the top of each 8-step block, seis3d_wave_8op_p3.F90:92-95). */
int fw_reference_steps(int nt);
/* Wait for all work of the handle. */
int fw_sync(fw_handle h);

/* Receiver records captured so far: out[nrec][npt][9]; *nrecords receives the
 * record count (io_seismo_put layout, mod_io.F90:461-497).  Synchronises. */
int fw_get_receivers(fw_handle h, float *out, int *nrecords);
/* Snapshot records: out[nrec][nfield_in_mask][subc0*subc1*subc2] (point index
 * i fastest; the fields of the mask in increasing field index order).  Synchronises.  fw_snapshot_device returns the device pointer of
 * the same storage (valid until fw_destroy) for zero-copy consumers. */
int fw_get_snapshot(fw_handle h, int snap_id, float *out, int *nrecords);
int fw_snapshot_device(fw_handle h, int snap_id, float **dptr, int *nrecords, long *npoints);

/* Reduced-precision SGT storage (no counterpart in the reference, which
 * writes FP32 NetCDF): like fw_add_snapshot, but each record of each field
 * is stored as 16-bit codes (FW_SGT_BF16 or FW_SGT_FP16) with one FP32 scale
 * per 16x16x16 block of snapshot points.  A value x of a block whose largest
 * magnitude is amax is coded as BF16(x/amax) or FP16(x*32768/amax), rounded
 * to nearest even, with scale amax or amax/32768; it is restored as
 * code * scale.  codes[nrec][nfield][npts] keep the point order of the FP32
 * snapshot; scales[nrec][nfield][nblocks], block index b0 + nb0*(b1 + nb1*b2)
 * with nbd = ceil(subc[d]/16).  fw_get_snapshot / fw_snapshot_device reject
 * packed snapshots and the packed calls reject FP32 ones. */
enum { FW_SGT_FP32 = 0, FW_SGT_BF16 = 1, FW_SGT_FP16 = 2 };
int fw_add_snapshot_packed(fw_handle h, const int *subs, const int *subc, const int *subt, int tinv,
                           unsigned field_mask, int max_records, int format, int *snap_id);
int fw_get_snapshot_packed(fw_handle h, int snap_id, uint16_t *codes, float *scales, int *nrecords, long *nblocks);
int fw_snapshot_packed_device(fw_handle h, int snap_id, uint16_t **codes, float **scales, int *nrecords,
                              long *npoints, long *nblocks);
/* Wavefield state (between steps): ghosted nx*ny*nz host arrays.  Replaces
 * io_rest_export / io_rest_import's array transfer (mod_io.F90:342-427).
 * Ghost values set here are seen by every later stage (no stage writes the
 * x / y ghosts), as with the reference's single array per field. */
int fw_get_wavefield(fw_handle h, int field, float *out);
int fw_set_wavefield(fw_handle h, int field, const float *in);

/* Batched virtual sources (no counterpart in the reference, which runs one
 * solver job per virtual source).  nbatch members share the grid, media,
 * metric, PML profiles, receivers and snapshot definitions; each has its own
 * wavefield, PML auxiliaries and output records.  Force source n belongs to
 * member force_member[n], moment source n to moment_member[n]; member b runs
 * nsteps[b] steps from ntime0 = 0 (uneven record lengths) and then stays
 * frozen.  Call after fw_set_force / fw_set_moment and before fw_init.  All
 * running members advance in the same stage launches (fw_set_batch_width
 * members per launch; 0 = all).  The get/set calls above address the member
 * chosen with fw_select_member (default 0).  Storage scales with nbatch. */
int fw_set_batch(fw_handle h, int nbatch, const int *force_member, const int *moment_member, const int *nsteps);
int fw_select_member(fw_handle h, int member);
int fw_set_batch_width(fw_handle h, int width);
/* Run each step as a captured HIP graph (one per F/B block position and
 * number of running members, built on first use); source times and record
 * slots come from a device-side step clock, so replays need no host
 * arguments.  Results are identical to the launch-by-launch path. */
int fw_use_graphs(fw_handle h, int on);
/* Time trial_steps steps for every stage-kernel tile shape (all running
 * members in one launch), then for the batch widths (powers of two below
 * nbatch) with the fastest shape; keep the fastest, and restore the zero
 * initial state with empty records.  Call after fw_init. */
int fw_autotune(fw_handle h, int trial_steps, int *tile, int *width, float *ms_per_step);

/* Restart.  The reference restarts from the 9 wavefields alone: its PML
 * memory variables start again at zero and the 8-step F/B cycle starts
 * again at its first block (seis3d_wave_8op_p3.F90:72-95, mod_io.F90:
 * 376-427).  fw_set_wavefield + fw_set_cycle_origin(h, ntime0) reproduce
 * that.  fw_checkpoint_* (no counterpart in the reference) save and load
 * the complete state between steps (wavefield and PML auxiliaries of the
 * selected member): a run continued from a checkpoint with the original
 * cycle origin is identical to an uninterrupted run.  Pattern of step
 * ntime: block (ntime - origin) mod 8; default origin 0. */
int fw_set_cycle_origin(fw_handle h, int origin);
int fw_checkpoint_size(fw_handle h, size_t *bytes);
int fw_checkpoint_save(fw_handle h, void *buf);
int fw_checkpoint_load(fw_handle h, const void *buf);

/* Per-kernel timing of the last fw_run (milliseconds of the main stage
 * kernels summed, measured with HIP events on the handle's stream). */
int fw_enable_timing(fw_handle h, int on);
int fw_get_timing(fw_handle h, double *stage_ms, double *total_ms, int *nstage_launches);

/* ---------------------------------------------------------------------------
 * Sensitivity-kernel time integral (code/srcF.kernel/SI_ker_sta.F90).
 * All arrays are DEVICE pointers.  A "series" array holds component c,
 * sample n (0-based), point p at base[c*cs + n*ts + p].  stream may be NULL
 * (default stream).
 * ------------------------------------------------------------------------- */
/* stress2strain, SI_ker_sta.F90:1185-1211 (in place, 6 components xx yy zz xy xz yz) */
int fwk_stress2strain(float *E, long cs, long ts, long npt, int knt, const float *lam, const float *mu,
                      void *stream);
/* sgt2scatter3 for one receiver component, SI_ker_sta.F90:1257-1337:
 * outA (alpha series) and outB (beta series), [knt][npt].  FP32 products,
 * FP64 accumulation. */
int fwk_scatter(const float *R, long rcs, long rts, const float *S, long scs, long sts, long npt, int knt,
                float dt, const float *lam, const float *mu, float *outA, float *outB, void *stream);
/* The same with the output pair selected: param 0 alpha/beta (as fwk_scatter),
 * param 1 lambda/mu: outA = lambda*x1, outB = 2*mu*x2
 * (SI_ker_sta_lamu.F90:1333-1336). */
int fwk_scatter_param(int param, const float *R, long rcs, long rts, const float *S, long scs, long sts, long npt,
                      int knt, float dt, const float *lam, const float *mu, float *outA, float *outB, void *stream);
/* vel2scatter3 for one receiver component, SI_ker_sta_rho.F90:1478-1552:
 * R and S are 3-component velocity series; out[knt][npt] = rho * dt *
 * sum_{m<n} (VxR+VyR+VzR)(m) (VxS+VyS+VzS)(n-m).  The density kernel is
 * then fwk_scatter2kernel with Sa = Sb = out (scatter2kernel_rho). */
int fwk_vscatter(const float *R, long rcs, long rts, const float *S, long scs, long sts, long npt, int knt, float dt,
                 const float *rho, float *out, void *stream);
/* Fortran eoshift along time (dim 4) of ncomp series in place, zero fill
 * (the earthquake-source SGT shift of SI_ker_EQ.F90:574-587). */
int fwk_eoshift(float *E, long cs, long ts, long npt, int knt, int ncomp, int shift, void *stream);
/* Kernel re-parameterisation in place on the phase (p) and amplitude (q)
 * kernels of the first (a) and second (b) parameter: mode 0 lambda/mu ->
 * kappa/mu (SI_ker_lamu2kamu.F90:823-830), 1 lambda/mu -> Vp/Vs
 * (SI_ker_lamu2vpvs.F90:823-830), 2 Vp/Vs -> kappa/mu
 * (SI_ker_vpvs2kamu.F90:823-831). */
int fwk_reparam(int mode, long npt, const float *lam, const float *mu, float *Kap, float *Kaq, float *Kbp,
                float *Kbq, void *stream);
/* scatter_filter (mode 1, "P1") / scatter_filtfilt (mode 2, "P2"),
 * SI_ker_sta.F90:1433-1475; b, a are HOST arrays of nod1 doubles. w0, w: [knt][npt]. */
int fwk_filter(int mode, long npt, int knt, int nod1, const double *b, const double *a, const float *w0,
               float *w, void *stream);
/* scatter2kernel, SI_ker_sta.F90:1375-1431; V, U are HOST arrays of knt floats;
 * nt1, nt2 1-based inclusive.  Outputs Kap Kaq Kbp Kbq [npt]. */
int fwk_scatter2kernel(long npt, int knt, int nt1, int nt2, float dt, float M0, const float *V,
                       const float *U, const float *Sa, const float *Sb, float *Kap, float *Kaq, float *Kbp,
                       float *Kbq, void *stream);
/* Allocation-free variant of fwk_scatter2kernel, same arguments and results
 * (bit for bit).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  V and U are staged in a fixed device
 * array of 2 x 16384 floats per GPU (held until the process ends) instead of
 * a per-call allocation; knt above 16384 is rejected with FW_ERR_ARG.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (from
 * several host threads) race.  Returns after the kernel has finished. */
int fwk_scatter2kernel_noalloc(long npt, int knt, int nt1, int nt2, float dt, float M0, const float *V,
                               const float *U, const float *Sa, const float *Sb, float *Kap, float *Kaq,
                               float *Kbp, float *Kbq, void *stream);
/* scatter_rotate, SI_ker_sta.F90:1354-1373 (rot: HOST 3x3 column-major). In place. */
int fwk_rotate(long n, float *X, float *Y, float *Z, const float *rot, void *stream);
/* Restore packed SGT records (fw_add_snapshot_packed layout) to FP32:
 * nslab = records x fields consecutive [npts] slabs of a c0 x c1 x c2
 * snapshot, codes and scales on the device, out[nslab][npts]. */
int fwk_sgt_decode(int format, const uint16_t *codes, const float *scales, int c0, int c1, int c2, long nslab,
                   float *out, void *stream);
/* Assembly of a kernel onto inversion blocks, SI_ker2assm.F90:300-320:
 * B[n] = sum_{q in [blk_ptr[n], blk_ptr[n+1])} K[pix[q]] * vol[q] (FP64
 * accumulation).  K, blk_ptr, pix (0-based point indices), vol and B are
 * DEVICE arrays (the CSR form of indx_of_pix / num_of_pix / volume). */
int fwk_assemble(long npt, const float *K, long nblk, const long *blk_ptr, const long *pix, const float *vol,
                 double *B, void *stream);
/* Last error of the fwk_* family and of the device-memory helpers below. */
const char *fwk_last_error(void);

/* ---------------------------------------------------------------------------
 * SGT block autoencoder (no counterpart in the reference): primitives of a
 * 3-D convolutional network on 16^3 blocks, trained and run on the GPU.
 * Device arrays, NCDHW layout, cubic kernels (K, stride S, padding P,
 * Do = (Di + 2P - K)/S + 1).  fwae_conv_bwd_data adds b2[ci] when b2 is not
 * NULL (the bias of a transposed convolution, whose forward pass it is).
 * fwae_conv_bwd_weight writes dw [Co][Ci][K^3] then db [Co] into dwb,
 * using part (nchunk * (Co*Ci*K^3 + Co) floats) for a two-pass reduction.
 * fwae_act_fwd applies leaky ReLU (slope; 1 = identity) and rounds the
 * stored activation to BF16; fwae_conv_fwd_act and fwae_conv_bwd_data_act
 * are fwae_conv_fwd / fwae_conv_bwd_data with that activation applied to
 * their output as it is stored.  fwae_blocks cuts snapshot records
 * [nrec][nc][c2][c1][c0] into [nrec][nblk][nc][16^3] blocks scaled to
 * [-1, 1] per block and component (scale [nrec][nblk][nc]).
 * ------------------------------------------------------------------------- */
int fwae_conv_fwd(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float *x, const float *w,
                  const float *b, float *y, void *stream);
int fwae_conv_bwd_data(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float *dy, const float *w,
                       const float *b2, float *dx, void *stream);
int fwae_conv_bwd_weight(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float *x, const float *dy,
                         int nchunk, float *part, float *dwb, void *stream);
int fwae_conv_fwd_act(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float *x, const float *w,
                      const float *b, float slope, float *y, void *stream);
int fwae_conv_bwd_data_act(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float *dy,
                           const float *w, const float *b2, float slope, float *dx, void *stream);
int fwae_act_fwd(long n, float slope, float *y, void *stream);
int fwae_act_bwd(long n, float slope, const float *y, float *dy, void *stream);
int fwae_mse(int N, long per, const float *y, const float *t, float *grad, double *loss, void *stream);
int fwae_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps, int step,
              void *stream);
int fwae_blocks(const float *rec, int nrec, int nc, int c0, int c1, int c2, float *out, float *scale, void *stream);

/* ---------------------------------------------------------------------------
 * Damped least squares by LSQR on the GPU (inv.LSQR/mod_LSQR.f90:142-292,
 * Paige & Saunders, FP64): A in CSR (the rows lldrow adds: data rows,
 * station/event terms, damping, smoothing; 0-based columns), rhs b.
 * fwl_solve runs slsqr's iteration with its stopping tests (test1 =
 * rnorm/(|A|_F |x|) <= atol, test2 = alpha|c|/|A|_F <= btol, or maxit) and
 * returns x, the standard errors of lgtsol (err may be NULL), the
 * iteration count and the final residual norm.  Host arrays.
 * ------------------------------------------------------------------------- */
typedef struct fwl_s *fwl_handle;
int fwl_create(long nrow, long ncol, long nnz, const long *row_ptr, const long *col, const double *val,
               const double *rhs, fwl_handle *out);
int fwl_solve(fwl_handle h, double atol, double btol, long maxit, double *x, double *err, long *itct, double *res);
int fwl_destroy(fwl_handle h);
const char *fwl_last_error(fwl_handle h);

/* Device-memory helpers for hosts without their own HIP allocator (the
 * reference keeps these arrays in Fortran allocatables; a Fortran or Python
 * caller of the fwk_* functions uses these to stage them in HBM). */
int fw_device_alloc(size_t bytes, void **dptr);
int fw_device_free(void *dptr);
int fw_memcpy_h2d(void *dst, const void *src, size_t bytes);
int fw_memcpy_d2h(void *dst, const void *src, size_t bytes);
int fw_memcpy_d2d(void *dst, const void *src, size_t bytes);
int fw_device_sync(void);

/* Attainable HBM bandwidth of this GPU (no counterpart in the reference; a
 * measurement aid for the roofline): `reps` timed launches of `blocks`
 * 256-thread blocks streaming two device buffers of `bytes` each with
 * 16-byte non-temporal accesses, read-only and copy (read + write 1:1);
 * GB/s of bytes moved.  Allocates and frees its buffers. */
int fw_hbm_probe(size_t bytes, int blocks, int reps, float *read_gbs, float *copy_gbs);

#ifdef __cplusplus
}
#endif
#endif
