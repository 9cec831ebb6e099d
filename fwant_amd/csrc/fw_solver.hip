// libfwant: MI355X-native velocity-stress solver of FWANT's seis3d_wave
// (DRP/opt MacCormack + RK4 + CFS-PML + free-surface imaging + sources).
//
// Device layout (per handle):
//   4 wavefield register sets A, B, C, T; each is 9 SoA fields of a pitched
//   column-major ghosted volume (i fastest, row pitch px = nx rounded up to
//   16 floats, interior column ni1 aligned to 64 bytes).
//   A holds W_n (and doubles as the RK "m" register); B and C ping-pong the
//   stage states; T is the RK accumulator.  Per step (RK forms of
//   fw_stage.h; r = 1/3):
//     stage 1: A -> B                      stage 2: B -> C, T = A + r (B - A) + b2 h
//     stage 3: C -> B                      stage 4: B -> A = T + r (B - A) + b4 h
//   which is classical RK4 as macdrp_syn/RK_beg/inn/fin compute it
//   (code/srcF/mod_macdrp.F90:268-393) with the accumulator written once.
//   PML auxiliary variables use the same A/B/C/T rotation per face slab.
//
// Kernels per RK stage:
//   k_surface   free-surface stress imaging (mod_macdrp.F90:3414-3436),
//               saving the pre-image surface stresses the RK form reads
//   k_stage     operator sweep + CFS-PML + RK update, velocity and stress
//               halves as two wave groups of one block      (fw_stage.h)
//               (mod_macdrp.F90:790-962, 291-393; mod_abs_npml.F90:1149-2123,
//               738-1105)
//   k_src       point-source terms (mod_src.F90:207-414), RK fix-up
#include "fw_common.h"
#include "fw_stage.h"
#include "../../include/fwant.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace fw;

namespace {

thread_local std::string g_create_err;

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      return fail(h, FW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                             \
  } while (0)

// ---------------------------------------------------------------------------
// Free-surface stress imaging (free_timg, mod_macdrp.F90:3414-3436) on the
// stage input, with the surface traction of surface force sources
// (src_surface, mod_src.F90:373-414).  The pre-image surface stresses are
// saved first where the RK form reads them: at stage 1 (save = 1, planes
// 0-2: the "m" values at k = nk2) and at stages 2 and 4 (save = 2, planes
// 3-5: the W_1 / W_3 values there).
// ---------------------------------------------------------------------------
struct SurfEntry {  // one surface force point (the last (n, m) writing it wins)
  int n, m;
  float rS;  // unused placeholder for alignment
  float S;   // area (x(i+1)-x(i-1))(y(j+1)-y(j-1))/4 z^2 sin
};
struct SurfArgs {
  float* win;
  float* msurf;
  const int* __restrict__ smap;  // plane map -> entry or -1 (one plane map per member)
  const SurfEntry* __restrict__ ent;
  const float* __restrict__ ft0;    // force start time [n]
  const float* __restrict__ ftime;  // window t0 [ntw]
  const float* __restrict__ ffreq;  // window f0 [ntw]
  const float* __restrict__ F;      // Fx Fy Fz blocks [3][ntw*n]
  long fs, sy, sz, ps, fblk, bstride;
  const int* __restrict__ clk;  // device step clock: clk[0] = ntime
  int ni1, ni2, nj1, nj2, nk2, ntw, stf_id, save;  // save: 0 none, 1 planes 0-2, 2 planes 3-5
  float tinc, stept;
  // mode C: slopes p, q per cell (null: the spherical grid)
  const float* __restrict__ cp;
  const float* __restrict__ cq;
};
// Mode C imaging (no counterpart in the reference): the traction
// combinations tau = sigma . N, N = (p, q, 1) (tau_z = Tzz + p Txz + q Tyz,
// tau_x = Txz + p Txx + q Txy, tau_y = Tyz + p Txy + q Tyy) are set to the
// source traction at k = nk2 by solving for Txz, Tyz, then Tzz, and the
// planes above hold tau's odd images in Tzz Txz Tyz.  With p = q = 0 this is
// free_timg.
__device__ __forceinline__ void surface_curv(const SurfArgs& g, float* Tzz, float* Txz, float* Tyz, long c2, long cs,
                                             float tx, float ty, float tz, const float* bzz, const float* bxz,
                                             const float* byz, long slot) {
  const float* Txx = g.win + slot * g.bstride + TXX * g.fs;
  const float* Tyy = g.win + slot * g.bstride + TYY * g.fs;
  const float* Txy = g.win + slot * g.bstride + TXY * g.fs;
  float ixx[4], iyy[4], ixy[4], pp[4], qq[4];
#pragma unroll
  for (int k = 0; k <= 3; ++k) {
    const long c = cs - k * g.sz;
    ixx[k] = Txx[c]; iyy[k] = Tyy[c]; ixy[k] = Txy[c];
    pp[k] = g.cp[c]; qq[k] = g.cq[c];
  }
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    const float p = pp[k], q = qq[k];
    Tzz[cs + k * g.sz] = 2.0f * tz - (bzz[k] + p * bxz[k] + q * byz[k]);
    Txz[cs + k * g.sz] = 2.0f * tx - (bxz[k] + p * ixx[k] + q * ixy[k]);
    Tyz[cs + k * g.sz] = 2.0f * ty - (byz[k] + p * ixy[k] + q * iyy[k]);
  }
  const float p = pp[0], q = qq[0];
  const float sxz = tx - p * ixx[0] - q * ixy[0];
  const float syz = ty - p * ixy[0] - q * iyy[0];
  Txz[cs] = sxz;
  Tyz[cs] = syz;
  Tzz[cs] = tz - p * sxz - q * syz;
  (void)c2;
}
__global__ void k_surface(SurfArgs g) {
  const int i = g.ni1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = g.nj1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > g.ni2 || j > g.nj2) return;
  const long slot = blockIdx.z;
  const int ntime = g.clk[0];
  const float ts = ((float)ntime + g.tinc) * g.stept;
  const long c2 = i + g.sy * j;
  const long cs = c2 + g.sz * g.nk2;
  float* Tzz = g.win + slot * g.bstride + TZZ * g.fs;
  float* Txz = g.win + slot * g.bstride + TXZ * g.fs;
  float* Tyz = g.win + slot * g.bstride + TYZ * g.fs;
  float* msurf = g.msurf + slot * 6 * g.ps + (g.save == 2 ? 3 * g.ps : 0);
  // every load before any store: the images above the surface and the
  // stresses below it are distinct cells, and issuing the loads together
  // takes one memory latency instead of one per image plane
  float bzz[4], bxz[4], byz[4];
#pragma unroll
  for (int k = 0; k <= 3; ++k) {
    bzz[k] = Tzz[cs - k * g.sz];
    bxz[k] = Txz[cs - k * g.sz];
    byz[k] = Tyz[cs - k * g.sz];
  }
  if (g.save) {
    msurf[c2] = bzz[0];
    msurf[g.ps + c2] = bxz[0];
    msurf[2 * g.ps + c2] = byz[0];
  }
  float tx = 0.f, ty = 0.f, tz = 0.f;
  const int e = g.smap[slot * g.ps + c2];
  if (e >= 0) {
    const SurfEntry se = g.ent[e];
    const float disp = stf_eval(0, g.stf_id, ts - g.ft0[se.n], g.ftime[se.m], g.ffreq[se.m], ntime, g.stept);
    const long q = se.m + (long)g.ntw * se.n;
    tx = g.F[q] / se.S * disp;
    ty = g.F[g.fblk + q] / se.S * disp;
    tz = g.F[2 * g.fblk + q] / se.S * disp;
  }
  if (g.cp) {
    surface_curv(g, Tzz, Txz, Tyz, c2, cs, tx, ty, tz, bzz, bxz, byz, slot);
    return;
  }
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    Tzz[cs + k * g.sz] = 2.0f * tz - bzz[k];
    Txz[cs + k * g.sz] = 2.0f * tx - bxz[k];
    Tyz[cs + k * g.sz] = 2.0f * ty - byz[k];
  }
  Tzz[cs] = tz;
  Txz[cs] = tx;
  Tyz[cs] = ty;
}

// ---------------------------------------------------------------------------
// Point-source terms as an RK fix-up, one thread per affected cell:
//   type 0  smoothed body force  hV += F*stf*w,  w = d/(V rho)   (src_force)
//   type 1  moment source        hT -= M*rate*w, w = d/V          (src_stress)
//   type 2  surface traction rate of a surface force at k = nk2 (the VSrc
//           terms of the traction condition, mod_macdrp.F90:942-950)
// ---------------------------------------------------------------------------
struct SrcContrib {
  int type, n, m;  // m: window for type 2 (last writer); unused otherwise
  float w;         // weight; for type 2: 1/S
};
struct SrcArgs {
  float* wout;
  float* t;
  const float* __restrict__ lam;
  const float* __restrict__ mu;
  const long* __restrict__ cell;
  const int* __restrict__ ptr;
  const SrcContrib* __restrict__ con;
  // force tables
  const float* __restrict__ ft0;
  const float* __restrict__ ftime;
  const float* __restrict__ ffreq;
  const float* __restrict__ F;  // [3][ntw*n]
  long fblk;
  int fntw, fstf;
  // moment tables
  const float* __restrict__ mt0;
  const float* __restrict__ mtime;
  const float* __restrict__ mfreq;
  const float* __restrict__ M;  // [6][ntw*n]
  long mblk;
  int mntw, mstf;
  long fs, bstride;
  const int* __restrict__ rowslot;  // member slot of each affected cell
  const int* __restrict__ clk;      // device step clock: clk[0] = ntime
  int ncell, nact, kind;  // kind: the RK form (stage index)
  float tinc, stept, a, b;
  // mode C: slopes p, q per cell (null: the spherical grid)
  const float* __restrict__ cp;
  const float* __restrict__ cq;
};
__global__ void k_src(SrcArgs g) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= g.ncell) return;
  const long slot = g.rowslot[u];
  if (slot >= g.nact) return;  // member past its record length
  const int ntime = g.clk[0];
  const float ts = ((float)ntime + g.tinc) * g.stept;
  const long c = g.cell[u];
  float dh[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) dh[f] = 0.f;
  for (int q = g.ptr[u]; q < g.ptr[u + 1]; ++q) {
    const SrcContrib s = g.con[q];
    if (s.type == 0) {
      for (int m = 0; m < g.fntw; ++m) {
        const float disp = stf_eval(0, g.fstf, ts - g.ft0[s.n], g.ftime[m], g.ffreq[m], ntime, g.stept);
        const long qq = m + (long)g.fntw * s.n;
        dh[VX] += g.F[qq] * disp * s.w;
        dh[VY] += g.F[g.fblk + qq] * disp * s.w;
        dh[VZ] += g.F[2 * g.fblk + qq] * disp * s.w;
      }
    } else if (s.type == 1) {
      for (int m = 0; m < g.mntw; ++m) {
        const float rate = stf_eval(1, g.mstf, ts - g.mt0[s.n], g.mtime[m], g.mfreq[m], ntime, g.stept);
        const float v = s.w * rate;
        const long qq = m + (long)g.mntw * s.n;
        dh[TXX] -= g.M[qq] * v;
        dh[TYY] -= g.M[g.mblk + qq] * v;
        dh[TZZ] -= g.M[2 * g.mblk + qq] * v;
        dh[TXY] -= g.M[3 * g.mblk + qq] * v;
        dh[TXZ] -= g.M[4 * g.mblk + qq] * v;
        dh[TYZ] -= g.M[5 * g.mblk + qq] * v;
      }
    } else {
      const float rate = stf_eval(1, g.fstf, ts - g.ft0[s.n], g.ftime[s.m], g.ffreq[s.m], ntime, g.stept);
      const long qq = s.m + (long)g.fntw * s.n;
      const float vx = g.F[qq] * s.w * rate, vy = g.F[g.fblk + qq] * s.w * rate,
                  vz = g.F[2 * g.fblk + qq] * s.w * rate;
      const float lam = g.lam[c], miu = g.mu[c], lam2mu = lam + 2.0f * miu;
      const float p = g.cp ? g.cp[c] : 0.f, qs = g.cq ? g.cq[c] : 0.f;
      if (p == 0.f && qs == 0.f) {
        const float dE33 = vz / lam2mu, dE13 = vx / miu / 2.0f, dE23 = vy / miu / 2.0f;
        dh[TXX] += lam * dE33;
        dh[TYY] += lam * dE33;
        dh[TZZ] += lam2mu * dE33;
        dh[TXZ] += 2.0f * miu * dE13;
        dh[TYZ] += 2.0f * miu * dE23;
      } else {
        // mode C: the traction rate's share of the surface solve (the
        // stage kernel solved with a zero source; the solve is linear)
        float aX, aY, aZ;
        surf_solve(lam, miu, p, qs, vx, vy, vz, aX, aY, aZ);
        const float E11 = p * aX, E22 = qs * aY, E33 = aZ, E12 = (p * aY + qs * aX) * 0.5f;
        const float E13 = (aX + p * aZ) * 0.5f, E23 = (aY + qs * aZ) * 0.5f;
        dh[TXX] += lam2mu * E11 + lam * (E22 + E33);
        dh[TYY] += lam2mu * E22 + lam * (E11 + E33);
        dh[TZZ] += lam2mu * E33 + lam * (E11 + E22);
        dh[TXY] += 2.0f * miu * E12;
        dh[TXZ] += 2.0f * miu * E13;
        dh[TYZ] += 2.0f * miu * E23;
      }
    }
  }
  float* wout = g.wout + slot * g.bstride;
  float* t = g.t + slot * g.bstride;
  // the RK form of the stage (fw_stage.h): forms 0 and 2 add a dh to the
  // stage output only (the accumulator takes it through W_1 / W_3), form 1
  // adds a dh and b dh, form 3 b dh.  All loads, then all stores (wout and
  // t are distinct register sets): one memory latency instead of one per field
  const bool upd_t = g.kind == 1;
  const float cw = g.kind == 3 ? g.b : g.a;
  float wv[9], tv[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) {
    wv[f] = wout[f * g.fs + c];
    tv[f] = upd_t ? t[f * g.fs + c] : 0.f;
  }
#pragma unroll
  for (int f = 0; f < 9; ++f) {
    wout[f * g.fs + c] = wv[f] + cw * dh[f];
    if (upd_t) t[f * g.fs + c] = tv[f] + g.b * dh[f];
  }
}

// receivers after step ntime = clk[0] when (ntime+1) % tinv == 0: member
// slot blockIdx.y writes record clk[1] of its rec[max_rec][npt][9]
__global__ void k_receivers(const float* __restrict__ W, long fs, long bstride, const long* __restrict__ pt, int npt,
                            float* __restrict__ rec, int max_rec, int tinv, const int* __restrict__ clk) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npt * 9 || (clk[0] + 1) % tinv != 0) return;
  const int r = clk[1];
  if (r >= max_rec) return;
  const long slot = blockIdx.y;
  const int p = q / 9, f = q % 9;
  rec[((size_t)slot * max_rec + r) * npt * 9 + q] = W[slot * bstride + f * fs + pt[p]];
}

// device step clock: after the outputs of step clk[0], advance the record
// counters of the captures that happened and the step itself
struct TickArgs {
  int* clk;
  int rec_tinv;  // 0: no receivers
  int nsnap;
  int tinv[16];
};
__global__ void k_tick(TickArgs g) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int t1 = g.clk[0] + 1;
  if (g.rec_tinv > 0 && t1 % g.rec_tinv == 0) g.clk[1] += 1;
  for (int s = 0; s < g.nsnap; ++s)
    if (t1 % g.tinv[s] == 0) g.clk[2 + s] += 1;
  g.clk[0] = t1;
}
struct ClockSet {
  int n;
  int v[18];
};
__global__ void k_setclk(int* clk, ClockSet c) {
  const int q = threadIdx.x;
  if (q < c.n) clk[q] = c.v[q];
}

// One volume between the dense host order (row pitch nx) and the pitched
// device layout (row pitch px), through a dense device staging copy: the
// host transfer stays one contiguous copy (a 2-D copy from pageable host
// memory moves one row at a time).
__global__ __launch_bounds__(256) void k_repitch(float* __restrict__ dst, const float* __restrict__ src, int nx, long px,
                                                 long n, int to_pitched) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const long r = q / nx, i = q - r * nx;
  if (to_pitched)
    dst[r * px + i] = src[q];
  else
    dst[q] = src[r * px + i];
}

// snapshot: out[nf][npts] for the fields in the mask list
struct SnapArgs {
  const float* __restrict__ W;
  float* __restrict__ out;  // member 0, record 0
  long fs, sy, sz, npts, bstride, mstride;
  const int* __restrict__ clk;
  int cidx, tinv, max_records;
  int s0, s1, s2, c0, c1, t0, t1, t2;
  int nf;
  int flist[9];
};
__global__ void k_snapshot(SnapArgs g) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.npts || (g.clk[0] + 1) % g.tinv != 0) return;
  const int rec = g.clk[g.cidx];
  if (rec >= g.max_records) return;
  const long slot = blockIdx.y;
  const float* W = g.W + slot * g.bstride;
  float* out = g.out + slot * g.mstride + (long)rec * g.nf * g.npts;
  const int a = (int)(p % g.c0);
  const long r = p / g.c0;
  const int b = (int)(r % g.c1);
  const int d = (int)(r / g.c1);
  const long c = (g.s0 + a * g.t0) + g.sy * (long)(g.s1 + b * g.t1) + g.sz * (long)(g.s2 + d * g.t2);
  for (int q = 0; q < g.nf; ++q) out[q * g.npts + p] = W[g.flist[q] * g.fs + c];
}

// snapshot capture into reduced-precision storage: one workgroup per 16^3
// block of snapshot points and field; a thread gathers a column of 16
// points, the block maximum sets the scale, codes keep the point order of
// the FP32 snapshot ([field][point], point index i fastest).
struct PackArgs {
  const float* __restrict__ W;
  uint16_t* __restrict__ codes;  // member 0, record 0: [nf][npts] per record
  float* __restrict__ scales;    // member 0, record 0: [nf][nblk] per record
  long fs, sy, sz, npts, bstride, cstride, sstride;
  const int* __restrict__ clk;
  int cidx, tinv, max_records;
  int s0, s1, s2, c0, c1, c2, t0, t1, t2, nb0, nb1, nblk;
  int nf, fmt;
  int flist[9];
};
__global__ __launch_bounds__(256) void k_snapshot_pack(PackArgs g) {
  if ((g.clk[0] + 1) % g.tinv != 0) return;
  const int rr = g.clk[g.cidx];
  if (rr >= g.max_records) return;
  const long slot = blockIdx.z;
  const int blk = blockIdx.x, q = blockIdx.y;
  const int b0 = blk % g.nb0, r = blk / g.nb0, b1 = r % g.nb1, b2 = r / g.nb1;
  const int a = b0 * SGT_B + (int)(threadIdx.x & 15), b = b1 * SGT_B + (int)(threadIdx.x >> 4);
  const bool in_ab = a < g.c0 && b < g.c1;
  const float* W = g.W + slot * g.bstride + g.flist[q] * g.fs;
  float v[SGT_B];
  float m = 0.f;
#pragma unroll
  for (int d = 0; d < SGT_B; ++d) {
    const int dd = b2 * SGT_B + d;
    v[d] = 0.f;
    if (in_ab && dd < g.c2) {
      const long c = (g.s0 + a * g.t0) + g.sy * (long)(g.s1 + b * g.t1) + g.sz * (long)(g.s2 + dd * g.t2);
      v[d] = W[c];
      m = fmaxf(m, fabsf(v[d]));
    }
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __shared__ float wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  const float amax = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
  const float inv = sgt_inv(g.fmt, amax);
  if (threadIdx.x == 0) g.scales[slot * g.sstride + ((long)rr * g.nf + q) * g.nblk + blk] = sgt_scale(g.fmt, amax);
  uint16_t* out = g.codes + slot * g.cstride + ((long)rr * g.nf + q) * g.npts;
#pragma unroll
  for (int d = 0; d < SGT_B; ++d) {
    const int dd = b2 * SGT_B + d;
    if (in_ab && dd < g.c2) out[a + (long)g.c0 * (b + (long)g.c1 * dd)] = sgt_code(g.fmt, v[d], inv);
  }
}

// FP64 metric construction (seis3d_metric.F90:82-112): sin/cot of colatitude
// and the inverse of x_xi = (B.x + F.x)/2; ghost values copied from the edge.
__global__ void k_metric(const double* __restrict__ g, int n, int lo, int hi, float* __restrict__ inv,
                         const double* __restrict__ gx_for_trig, float* __restrict__ xsin, float* __restrict__ xcot) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (xsin) {
    const double th = gx_for_trig[i];
    xsin[i] = (float)sin(th);
    xcot[i] = (float)(1.0 / tan(th));
  }
  const double cF[5] = {-0.30874, -0.6326, 1.2330, -0.3334, 0.04168};
  const double cB[5] = {-0.04168, 0.3334, -1.2330, 0.6326, 0.30874};
  const int ic = i < lo ? lo : (i > hi ? hi : i);
  double b = 0.0, f = 0.0;
  for (int q = 0; q < 5; ++q) {
    b += g[ic - 3 + q] * cB[q];
    f += g[ic - 1 + q] * cF[q];
  }
  inv[i] = (float)(1.0 / ((b + f) * 0.5));
}

// Mode C metric (no counterpart in the reference): terrain-following grid
// theta = x(i), phi = y(j), r = r(i, j, k) (dense FP64 input r).  FP64:
//   zeta_r = 1 / r_zeta   (k_metric's expression along the column),
//   p = -xi_x r_xi / r,  q = -eta_y r_eta / (r sin(theta)),
// the xi / eta derivatives as 1/2 (F + B) with B written as the mirrored F
// taps (exactly 0 where r is constant along the row); derivative values on
// the 3 outer planes of each axis copied from the nearest plane the stencil
// fits; 1/r of the FP32-rounded radius.  Outputs in the pitched field
// layout (cell i + sy j + sz k).  Like k_metric, every FP64 operation is
// rounded on its own (no contraction), so that a flat grid gives the 1-D
// metric bit for bit.
__device__ double inv_metric1d(const double* __restrict__ g, int i) {
#pragma clang fp contract(off)
  const double cF[5] = {-0.30874, -0.6326, 1.2330, -0.3334, 0.04168};
  const double cB[5] = {-0.04168, 0.3334, -1.2330, 0.6326, 0.30874};
  double b = 0.0, f = 0.0;
  for (int q = 0; q < 5; ++q) {
    b += g[i - 3 + q] * cB[q];
    f += g[i - 1 + q] * cF[q];
  }
  return 1.0 / ((b + f) * 0.5);
}
__device__ double dmet_mirror(const double* __restrict__ g, long st) {
#pragma clang fp contract(off)
  const double cF[5] = {-0.30874, -0.6326, 1.2330, -0.3334, 0.04168};
  double f = 0.0, b = 0.0;
  for (int q = 0; q < 5; ++q) {
    f += g[(q - 1) * st] * cF[q];
    b -= g[(1 - q) * st] * cF[q];
  }
  return (f + b) * 0.5;
}
__global__ __launch_bounds__(256) void k_metric3d(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ r, int nx, int ny, int nz, long sy, long sz,
                                                  float* __restrict__ p, float* __restrict__ q, float* __restrict__ zz,
                                                  float* __restrict__ rr) {
#pragma clang fp contract(off)
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= (long)nx * ny * nz) return;
  const int i = (int)(u % nx), j = (int)((u / nx) % ny), k = (int)(u / ((long)nx * ny));
  const long dsy = nx, dsz = (long)nx * ny;
  const int kc = min(max(k, 3), nz - 4), ic = min(max(i, 3), nx - 4), jc = min(max(j, 3), ny - 4);
  const double cF[5] = {-0.30874, -0.6326, 1.2330, -0.3334, 0.04168};
  const double cB[5] = {-0.04168, 0.3334, -1.2330, 0.6326, 0.30874};
  const double* col = r + i + dsy * j;
  double b = 0.0, f = 0.0;
  for (int m = 0; m < 5; ++m) {
    b += col[(kc - 3 + m) * dsz] * cB[m];
    f += col[(kc - 1 + m) * dsz] * cF[m];
  }
  const long c = i + sy * j + sz * k;
  zz[c] = (float)(1.0 / ((b + f) * 0.5));
  const long cx = ic + dsy * j + dsz * k, cy = i + dsy * jc + dsz * k;
  p[c] = (float)(-inv_metric1d(x, ic) * dmet_mirror(r + cx, 1) / r[cx]);
  q[c] = (float)(-inv_metric1d(y, jc) * dmet_mirror(r + cy, dsy) / (r[cy] * sin(x[i])));
  rr[c] = (float)(1.0 / (double)(float)r[u]);
}
// one-sided zeta derivatives F(p), F(q) (fwd) or B (bwd) of the slopes, FP32,
// on planes 3 .. nz-4 (0 elsewhere)
__global__ __launch_bounds__(256) void k_metric3d_dz(const float* __restrict__ p, const float* __restrict__ q, int nx,
                                                     int ny, int nz, long sy, long sz, float* __restrict__ dpF,
                                                     float* __restrict__ dqF, float* __restrict__ dpB,
                                                     float* __restrict__ dqB) {
  const long u = (long)blockIdx.x * 256 + threadIdx.x;
  if (u >= (long)nx * ny * nz) return;
  const int i = (int)(u % nx), j = (int)((u / nx) % ny), k = (int)(u / ((long)nx * ny));
  const long c = i + sy * j + sz * k;
  const bool in = k >= 3 && k <= nz - 4;
  dpF[c] = in ? dop<1>(p + c, sz) : 0.f;
  dqF[c] = in ? dop<1>(q + c, sz) : 0.f;
  dpB[c] = in ? dop<0>(p + c, sz) : 0.f;
  dqB[c] = in ? dop<0>(q + c, sz) : 0.f;
}

// grid-stride fill: a dispatch holds fewer than 2^32 work-items, so a batch's
// register sets (tens of billions of floats) take several passes per thread
template <typename T>
__global__ void k_fill(T* p, long n, T v) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) p[q] = v;
}
unsigned fill_blocks(long n) { return (unsigned)std::min<long>((n + 255) / 256, 1L << 20); }

// 8-block operator cycle, seis3d_wave_8op_p3.F90:82-591 (1 = forward)
const int BLOCK_PAT[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
struct PmlFace {
  bool on = false;
  int n = 0, lo[3]{}, hi[3]{}, ex[3]{};
  long cnt = 0;
  float* d = nullptr;  // 4 sets x 6 vars x cnt
};
struct Snapshot {
  int subs[3], subc[3], subt[3], tinv, nf, flist[9];
  long npts;
  int max_records, nrec = 0;   // nrec: captures so far (record index of the active members)
  std::vector<int> cnt;         // records held by each member slot
  float* d = nullptr;           // FP32 records [slot][max_records][nf][npts]
  int fmt = SGT_FP32;           // or BF16 / FP16 codes with per-block scales
  int nb[3]{}, nblk = 0;
  uint16_t* codes = nullptr;    // [slot][max_records][nf][npts]
  float* scales = nullptr;      // [slot][max_records][nf][nblk]
};
constexpr int MAX_SNAP = 16;    // snapshots per handle (device clock counters)

struct fw_solver_s {
  fw_config cfg{};
  int nx, ny, nz, px;
  long sy, sz, fs, ps;  // pitched strides; ps = plane stride of 2-D arrays
  long off;             // leading pad so that column ni1 is 64-byte aligned
  int ni1, ni2, nj1, nj2, nk1, nk2;
  int freenode;
  hipStream_t stream = nullptr;
  std::string err;
  bool ready = false;

  float* d_sets = nullptr;  // nbatch * 4 * 9 * fs (+ off)
  float* set[4]{};          // A B C T of member slot 0
  float* d_media = nullptr; // lam mu rrho
  float *lam = nullptr, *mu = nullptr, *rrho = nullptr;
  float* d_1d = nullptr;
  float *xi_x, *eta_y, *zeta_z, *rz, *rxsin, *xcot;
  float* d_msurf = nullptr;
  int* d_fluidcol = nullptr;  // per (i, j) column: 1 if any cell has mu <= SEIS_ZERO
  float* d_stage = nullptr;   // dense staging copy of one volume (host transfers)
  int tilecfg = 0;            // stage-kernel tile shape index
  int stage_next = 0, stage_ntime = 0, stage_nact = 0;  // fw_stage: next stage, its step, running members
  int4* d_items = nullptr;    // stage work items: PML-shell items, then interior items
  int n_narrow = 0, n_lean = 0, n_pml = 0;
  int xg = 1;  // stage-kernel items per XCD run
  // mode C (terrain-following grid): 8 volumes in the field layout:
  // p, q, zeta_r, 1/r, then F(p), F(q), B(p), B(q) along zeta
  bool curv = false;
  float* d_met3 = nullptr;
  float* met3[8]{};
  std::vector<float> hr3;  // radius of every node (dense), for the source weights
  // host copies
  std::vector<float> hx, hy, hz, hxsin, hxcot, hxi, het, hze;
  std::vector<float> hlam, hmu, hrho;
  bool have_grid = false, have_media = false;
  // PML
  std::vector<float> Ax, Bx, Dx, Ay, By, Dy, Az, Bz, Dz;
  float* d_pml = nullptr;  // b1, ad, dd for x, y, z
  PmlFace face[6];
  // sources
  int nforce = 0, fntw = 0, fstf = 0;
  std::vector<int> findx;
  std::vector<float> fshift, ft0, ftime, ffreq, F;
  int nmom = 0, mntw = 0, mstf = 0;
  std::vector<int> mindx;
  std::vector<float> mshift, mt0, mtime, mfreq, M;
  void* d_src = nullptr;  // packed source tables
  std::vector<void*> d_src_allocs;
  float *d_ft0 = nullptr, *d_ftime = nullptr, *d_ffreq = nullptr, *d_F = nullptr;
  float *d_mt0 = nullptr, *d_mtime = nullptr, *d_mfreq = nullptr, *d_M = nullptr;
  int* d_smap = nullptr;
  SurfEntry* d_sent = nullptr;
  long* d_scell = nullptr;
  int* d_sptr = nullptr;
  SrcContrib* d_scon = nullptr;
  int n_srccell = 0, n_surf = 0;
  // receivers
  int npt = 0, pt_tinv = 1, max_rec = 0, nrec = 0;
  std::vector<int> pt_indx;
  long* d_pt = nullptr;
  float* d_rec = nullptr;
  std::vector<Snapshot> snaps;
  // batched virtual sources: members share media, metric and PML profiles;
  // member slots are ordered by record length (longest first), so the
  // members still running at any step are the first nact slots
  int nbatch = 1;
  std::vector<int> mem_steps;         // steps each member runs, by slot
  std::vector<int> slot_of;           // member -> slot
  std::vector<int> fmember, mmember;  // member of each force / moment source
  long bstride = 0;                   // floats between members' register sets
  int width = 0;                      // members per stage launch (0: all running)
  int sel = 0;                        // member addressed by the get/set calls
  int cycle_origin = 0;               // ntime at which the 8-step F/B cycle starts
  std::vector<int> rec_cnt;           // receiver records held by each slot
  int* d_clk = nullptr;               // device step clock: ntime, record counters
  int* d_srow = nullptr;              // member slot of each source row
  // captured step graphs keyed by (ntime mod 8, running members)
  bool use_graphs = false;
  std::map<long, hipGraphExec_t> gexec;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  double stage_ms = 0, total_ms = 0;
  int nstage = 0;
};

static int fail(fw_handle h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_err = msg;
  return code;
}

extern "C" {

const char* fw_last_error(fw_handle h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int fw_reference_steps(int nt) { return nt < 0 ? 0 : 8 * ((nt + 1 + 7) / 8); }

int fw_create(const fw_config* cfg, fw_handle* out) {
  fw_handle h = nullptr;
  if (!cfg || !out) return fail(nullptr, FW_ERR_ARG, "fw_create: null argument");
  if (cfg->ni < 8 || cfg->nj < 8 || cfg->nk < 8) return fail(nullptr, FW_ERR_ARG, "fw_create: ni, nj, nk must be >= 8");
  if (!(cfg->stept > 0.f)) return fail(nullptr, FW_ERR_ARG, "fw_create: stept must be > 0");
  for (int n = 0; n < 6; ++n) {
    const int N = cfg->abs_number[n];
    const int ext = n < 2 ? cfg->ni : (n < 4 ? cfg->nj : cfg->nk);
    if (N < 0 || N + 1 > ext) return fail(nullptr, FW_ERR_ARG, "fw_create: abs_number out of range");
    if (N > 0 && cfg->cfs_bmax[n] < 1.0f)
      return fail(nullptr, FW_ERR_ARG, "fw_create: CFS_bmax should be large or equal to 1");
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, FW_ERR_HIP, "fw_create: no HIP device");
  h = new fw_solver_s();
  h->cfg = *cfg;
  h->nx = cfg->ni + 6; h->ny = cfg->nj + 6; h->nz = cfg->nk + 6;
  h->px = (h->nx + 31) / 32 * 32;
  h->sy = h->px; h->sz = (long)h->px * h->ny; h->ps = h->sz;
  h->off = 29;  // (off + ni1) % 32 == 0 with ni1 = 3: every tile's rows start on a 128-byte line
  h->fs = h->sz * h->nz + 32;
  h->ni1 = 3; h->ni2 = 3 + cfg->ni - 1; h->nj1 = 3; h->nj2 = 3 + cfg->nj - 1; h->nk1 = 3; h->nk2 = 3 + cfg->nk - 1;
  h->freenode = cfg->freenode ? 1 : 0;
  if (const char* tv = getenv("FWANT_TILE")) h->tilecfg = std::max(0, std::min(NTILECFG - 1, atoi(tv)));
  if (cfg->abs_number[5] > 0) h->freenode = 0;  // mod_abs_npml.F90:182-184
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete h; return fail(nullptr, FW_ERR_HIP, "hipStreamCreate failed"); }
  h->bstride = 4L * 9 * h->fs;  // multiple of 32 floats: every member's rows stay 128-byte aligned
  h->mem_steps.assign(1, 0x7fffffff);
  h->slot_of.assign(1, 0);
  e = hipMalloc(&h->d_clk, sizeof(int) * (2 + MAX_SNAP));
  if (e != hipSuccess) { fw_destroy(h); return fail(nullptr, FW_ERR_NOMEM, "hipMalloc clock failed"); }
  e = hipMalloc(&h->d_media, sizeof(float) * (size_t)(3 * h->fs + 64));
  if (e != hipSuccess) { fw_destroy(h); return fail(nullptr, FW_ERR_NOMEM, "hipMalloc media failed"); }
  h->lam = h->d_media + h->off; h->mu = h->lam + h->fs; h->rrho = h->mu + h->fs;
  const long n1 = h->nx + h->ny + h->nz;
  e = hipMalloc(&h->d_1d, sizeof(float) * 2 * n1 + sizeof(float) * 4 * (h->nx + h->nz));
  if (e != hipSuccess) { fw_destroy(h); return fail(nullptr, FW_ERR_NOMEM, "hipMalloc 1-d failed"); }
  h->xi_x = h->d_1d; h->eta_y = h->xi_x + h->nx; h->zeta_z = h->eta_y + h->ny;
  h->rz = h->zeta_z + h->nz; h->rxsin = h->rz + h->nz; h->xcot = h->rxsin + h->nx;
  *out = h;
  return FW_OK;
}

int fw_destroy(fw_handle h) {
  if (!h) return FW_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  hipFree(h->d_fluidcol);
  hipFree(h->d_sets); hipFree(h->d_media); hipFree(h->d_1d); hipFree(h->d_msurf); hipFree(h->d_pml);
  hipFree(h->d_met3);
  hipFree(h->d_stage);
  for (auto& f : h->face) hipFree(f.d);
  for (void* p : h->d_src_allocs) hipFree(p);
  hipFree(h->d_pt); hipFree(h->d_rec);
  for (auto& s : h->snaps) { hipFree(s.d); hipFree(s.codes); hipFree(s.scales); }
  for (auto& e : h->ev) hipEventDestroy(e);
  hipFree(h->d_items);
  hipFree(h->d_clk);
  for (auto& kv : h->gexec) hipGraphExecDestroy(kv.second);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return FW_OK;
}

int fw_dims(fw_handle h, int* nx, int* ny, int* nz, int* freenode) {
  if (!h) return FW_ERR_ARG;
  if (nx) *nx = h->nx;
  if (ny) *ny = h->ny;
  if (nz) *nz = h->nz;
  if (freenode) *freenode = h->freenode;
  return FW_OK;
}

static int upload_1d(fw_handle h) {
  std::vector<float> rz(h->nz), rxs(h->nx);
  for (int k = 0; k < h->nz; ++k) rz[k] = (float)(1.0 / (double)h->hz[k]);
  for (int i = 0; i < h->nx; ++i) rxs[i] = (float)(1.0 / (double)h->hxsin[i]);
  HIPCHK(hipMemcpyAsync(h->xi_x, h->hxi.data(), sizeof(float) * h->nx, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->eta_y, h->het.data(), sizeof(float) * h->ny, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->zeta_z, h->hze.data(), sizeof(float) * h->nz, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->rz, rz.data(), sizeof(float) * h->nz, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->rxsin, rxs.data(), sizeof(float) * h->nx, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->xcot, h->hxcot.data(), sizeof(float) * h->nx, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

int fw_set_grid(fw_handle h, const float* x, const float* y, const float* z, const float* xsin, const float* xcot,
                const float* xi_x, const float* eta_y, const float* zeta_z) {
  if (!h || !x || !y || !z || !xsin || !xcot || !xi_x || !eta_y || !zeta_z) return fail(h, FW_ERR_ARG, "fw_set_grid: null argument");
  h->hx.assign(x, x + h->nx); h->hy.assign(y, y + h->ny); h->hz.assign(z, z + h->nz);
  h->hxsin.assign(xsin, xsin + h->nx); h->hxcot.assign(xcot, xcot + h->nx);
  h->hxi.assign(xi_x, xi_x + h->nx); h->het.assign(eta_y, eta_y + h->ny); h->hze.assign(zeta_z, zeta_z + h->nz);
  for (int k = 0; k < h->nz; ++k)
    if (!(h->hz[k] != 0.f)) return fail(h, FW_ERR_ARG, "fw_set_grid: z must be nonzero");
  for (int i = 0; i < h->nx; ++i)
    if (!(h->hxsin[i] != 0.f)) return fail(h, FW_ERR_ARG, "fw_set_grid: xsin must be nonzero");
  h->have_grid = true;
  h->ready = false;
  return upload_1d(h);
}

int fw_build_metric(fw_handle h, const double* x, const double* y, const double* z, float* oxs, float* oxc,
                    float* oxi, float* oet, float* oze) {
  if (!h || !x || !y || !z) return fail(h, FW_ERR_ARG, "fw_build_metric: null argument");
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  double* d = nullptr;
  float* f = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double) * (nx + ny + nz)));
  if (hipMalloc(&f, sizeof(float) * (3 * nx + ny + nz)) != hipSuccess) {
    hipFree(d);
    return fail(h, FW_ERR_NOMEM, "fw_build_metric: hipMalloc failed");
  }
  hipMemcpyAsync(d, x, sizeof(double) * nx, hipMemcpyHostToDevice, h->stream);
  hipMemcpyAsync(d + nx, y, sizeof(double) * ny, hipMemcpyHostToDevice, h->stream);
  hipMemcpyAsync(d + nx + ny, z, sizeof(double) * nz, hipMemcpyHostToDevice, h->stream);
  float *fxs = f, *fxc = f + nx, *fxi = f + 2 * nx, *fet = f + 3 * nx, *fze = fet + ny;
  hipLaunchKernelGGL(k_metric, dim3((nx + 63) / 64), dim3(64), 0, h->stream, d, nx, 3, nx - 4, fxi, d, fxs, fxc);
  hipLaunchKernelGGL(k_metric, dim3((ny + 63) / 64), dim3(64), 0, h->stream, d + nx, ny, 3, ny - 4, fet,
                     (const double*)nullptr, (float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_metric, dim3((nz + 63) / 64), dim3(64), 0, h->stream, d + nx + ny, nz, 3, nz - 4, fze,
                     (const double*)nullptr, (float*)nullptr, (float*)nullptr);
  std::vector<float> hf(3 * nx + ny + nz);
  hipError_t e = hipMemcpyAsync(hf.data(), f, sizeof(float) * hf.size(), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(d);
  hipFree(f);
  if (e != hipSuccess) return fail(h, FW_ERR_HIP, std::string("fw_build_metric: ") + hipGetErrorString(e));
  h->hx.resize(nx); h->hy.resize(ny); h->hz.resize(nz);
  for (int i = 0; i < nx; ++i) h->hx[i] = (float)x[i];
  for (int j = 0; j < ny; ++j) h->hy[j] = (float)y[j];
  for (int k = 0; k < nz; ++k) h->hz[k] = (float)z[k];
  h->hxsin.assign(hf.begin(), hf.begin() + nx);
  h->hxcot.assign(hf.begin() + nx, hf.begin() + 2 * nx);
  h->hxi.assign(hf.begin() + 2 * nx, hf.begin() + 3 * nx);
  h->het.assign(hf.begin() + 3 * nx, hf.begin() + 3 * nx + ny);
  h->hze.assign(hf.begin() + 3 * nx + ny, hf.end());
  if (oxs) std::copy(h->hxsin.begin(), h->hxsin.end(), oxs);
  if (oxc) std::copy(h->hxcot.begin(), h->hxcot.end(), oxc);
  if (oxi) std::copy(h->hxi.begin(), h->hxi.end(), oxi);
  if (oet) std::copy(h->het.begin(), h->het.end(), oet);
  if (oze) std::copy(h->hze.begin(), h->hze.end(), oze);
  h->have_grid = true;
  h->ready = false;
  return upload_1d(h);
}

static int upload_volume(fw_handle h, float* dst, const float* src) {
  const long n = (long)h->nx * h->ny * h->nz;
  if (!h->d_stage) HIPCHK(hipMalloc(&h->d_stage, sizeof(float) * n));
  HIPCHK(hipMemcpyAsync(h->d_stage, src, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_repitch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, dst, h->d_stage, h->nx,
                     h->px, n, 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}
static int download_volume(fw_handle h, float* dst, const float* src) {
  const long n = (long)h->nx * h->ny * h->nz;
  if (!h->d_stage) HIPCHK(hipMalloc(&h->d_stage, sizeof(float) * n));
  hipLaunchKernelGGL(k_repitch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->d_stage, src, h->nx,
                     h->px, n, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, h->d_stage, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

int fw_set_media(fw_handle h, const float* lam, const float* mu, const float* rho) {
  if (!h || !lam || !mu || !rho) return fail(h, FW_ERR_ARG, "fw_set_media: null argument");
  const size_t n = (size_t)h->nx * h->ny * h->nz;
  h->hlam.assign(lam, lam + n); h->hmu.assign(mu, mu + n); h->hrho.assign(rho, rho + n);
  std::vector<float> rr(n);
  std::vector<int> fluid((size_t)h->ps, 0);
  for (int k = 0; k < h->nz; ++k)
    for (int j = 0; j < h->ny; ++j)
      for (int i = 0; i < h->nx; ++i)
        if (mu[i + (size_t)h->nx * (j + (size_t)h->ny * k)] <= SEIS_ZERO) fluid[i + h->sy * j] = 1;
  if (!h->d_fluidcol) HIPCHK(hipMalloc(&h->d_fluidcol, sizeof(int) * h->ps));
  HIPCHK(hipMemcpy(h->d_fluidcol, fluid.data(), sizeof(int) * h->ps, hipMemcpyHostToDevice));
  for (size_t q = 0; q < n; ++q) {
    if (!(rho[q] > 0.f)) return fail(h, FW_ERR_ARG, "fw_set_media: rho must be > 0 everywhere");
    rr[q] = 1.0f / rho[q];
  }
  int rc;
  if ((rc = upload_volume(h, h->lam, lam))) return rc;
  if ((rc = upload_volume(h, h->mu, mu))) return rc;
  if ((rc = upload_volume(h, h->rrho, rr.data()))) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  h->have_media = true;
  h->ready = false;
  return FW_OK;
}

int fw_build_metric3d(fw_handle h, const double* x, const double* y, const double* r, float* op, float* oq, float* ozz,
                      float* orr) {
  if (!h || !x || !y || !r) return fail(h, FW_ERR_ARG, "fw_build_metric3d: null argument");
  if (!h->have_grid) return fail(h, FW_ERR_STATE, "fw_build_metric3d: set the 1-D grid first (fw_set_grid / fw_build_metric)");
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  const size_t n = (size_t)nx * ny * nz;
  for (size_t u = 0; u < n; ++u)
    if (!(r[u] > 0.0)) return fail(h, FW_ERR_ARG, "fw_build_metric3d: radii must be > 0");
  for (int i = 0; i < nx; ++i)
    if (!(std::sin(x[i]) != 0.0)) return fail(h, FW_ERR_ARG, "fw_build_metric3d: sin(x) must be nonzero");
  if (!h->d_met3) {
    HIPCHK(hipMalloc(&h->d_met3, sizeof(float) * (8 * (size_t)h->fs + 64)));
    for (int v = 0; v < 8; ++v) h->met3[v] = h->d_met3 + h->off + (size_t)v * h->fs;
  }
  double* d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double) * (n + nx + ny)));
  hipError_t e = hipMemcpyAsync(d, r, sizeof(double) * n, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + n, x, sizeof(double) * nx, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + n + nx, y, sizeof(double) * ny, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    const unsigned nb = (unsigned)((n + 255) / 256);
    float** m = h->met3;
    hipLaunchKernelGGL(k_metric3d, dim3(nb), dim3(256), 0, h->stream, d + n, d + n + nx, d, nx, ny, nz, h->sy, h->sz,
                       m[0], m[1], m[2], m[3]);
    hipLaunchKernelGGL(k_metric3d_dz, dim3(nb), dim3(256), 0, h->stream, m[0], m[1], nx, ny, nz, h->sy, h->sz, m[4],
                       m[5], m[6], m[7]);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(h, FW_ERR_HIP, std::string("fw_build_metric3d: ") + hipGetErrorString(e));
  float* outs[4] = {op, oq, ozz, orr};
  for (int v = 0; v < 4; ++v)
    if (outs[v]) {
      int rc = download_volume(h, outs[v], h->met3[v]);
      if (rc) return rc;
    }
  h->hr3.resize(n);
  for (size_t u = 0; u < n; ++u) h->hr3[u] = (float)r[u];
  h->curv = true;
  h->ready = false;
  return FW_OK;
}

int fw_set_force(fw_handle h, int n, int ntw, int stf_id, const int* indx, const float* shift, const float* t0,
                 const float* stf_time, const float* stf_freq, const float* Fx, const float* Fy, const float* Fz) {
  if (!h || n < 0 || (n > 0 && (ntw <= 0 || !indx || !shift || !t0 || !stf_time || !stf_freq || !Fx || !Fy || !Fz)))
    return fail(h, FW_ERR_ARG, "fw_set_force: bad argument");
  if (n == 0) {  // no forces: the arrays may be NULL and ntw is ignored
    h->nforce = 0; h->fntw = 0; h->fstf = stf_id;
    h->findx.clear(); h->fshift.clear(); h->ft0.clear(); h->ftime.clear(); h->ffreq.clear(); h->F.clear();
    h->ready = false;
    return FW_OK;
  }
  h->nforce = n; h->fntw = ntw; h->fstf = stf_id;
  h->findx.assign(indx, indx + 3 * n);
  h->fshift.assign(shift, shift + 3 * n);
  h->ft0.assign(t0, t0 + n);
  h->ftime.assign(stf_time, stf_time + ntw);
  h->ffreq.assign(stf_freq, stf_freq + ntw);
  h->F.resize(3 * (size_t)ntw * n);
  std::copy(Fx, Fx + ntw * n, h->F.begin());
  std::copy(Fy, Fy + ntw * n, h->F.begin() + ntw * n);
  std::copy(Fz, Fz + ntw * n, h->F.begin() + 2 * ntw * n);
  h->ready = false;
  return FW_OK;
}

int fw_set_moment(fw_handle h, int n, int ntw, int stf_id, const int* indx, const float* shift, const float* t0,
                  const float* stf_time, const float* stf_freq, const float* M6) {
  if (!h || n < 0 || (n > 0 && (ntw <= 0 || !indx || !shift || !t0 || !stf_time || !stf_freq || !M6)))
    return fail(h, FW_ERR_ARG, "fw_set_moment: bad argument");
  if (n == 0) {  // no moment sources: the arrays may be NULL and ntw is ignored
    h->nmom = 0; h->mntw = 0; h->mstf = stf_id;
    h->mindx.clear(); h->mshift.clear(); h->mt0.clear(); h->mtime.clear(); h->mfreq.clear(); h->M.clear();
    h->ready = false;
    return FW_OK;
  }
  h->nmom = n; h->mntw = ntw; h->mstf = stf_id;
  h->mindx.assign(indx, indx + 3 * n);
  h->mshift.assign(shift, shift + 3 * n);
  h->mt0.assign(t0, t0 + n);
  h->mtime.assign(stf_time, stf_time + ntw);
  h->mfreq.assign(stf_freq, stf_freq + ntw);
  h->M.assign(M6, M6 + 6 * (size_t)ntw * n);
  h->ready = false;
  return FW_OK;
}

int fw_set_receivers(fw_handle h, int n, int tinv, const int* indx, int max_records) {
  if (!h || n < 0 || (n > 0 && (!indx || tinv <= 0 || max_records < 0))) return fail(h, FW_ERR_ARG, "fw_set_receivers: bad argument");
  for (int p = 0; p < n; ++p) {
    const int i = indx[3 * p] - 1, j = indx[3 * p + 1] - 1, k = indx[3 * p + 2] - 1;
    if (i < 0 || i >= h->nx || j < 0 || j >= h->ny || k < 0 || k >= h->nz)
      return fail(h, FW_ERR_ARG, "fw_set_receivers: receiver index outside the ghosted grid");
  }
  hipFree(h->d_pt); hipFree(h->d_rec);
  h->d_pt = nullptr; h->d_rec = nullptr;
  h->npt = n; h->pt_tinv = tinv > 0 ? tinv : 1; h->max_rec = max_records; h->nrec = 0;
  h->pt_indx.assign(indx, indx + 3 * n);
  if (n > 0) {
    std::vector<long> lin(n);
    for (int p = 0; p < n; ++p)
      lin[p] = (indx[3 * p] - 1) + h->sy * (long)(indx[3 * p + 1] - 1) + h->sz * (long)(indx[3 * p + 2] - 1);
    HIPCHK(hipMalloc(&h->d_pt, sizeof(long) * n));
    HIPCHK(hipMemcpy(h->d_pt, lin.data(), sizeof(long) * n, hipMemcpyHostToDevice));
  }
  h->ready = false;  // record storage is sized by fw_init
  return FW_OK;
}

static int add_snapshot(fw_handle h, const int* subs, const int* subc, const int* subt, int tinv, unsigned mask,
                        int max_records, int fmt, int* snap_id, const char* who) {
  if (!h || !subs || !subc || !subt || tinv <= 0 || max_records <= 0 || !(mask & 0x1ff) || !snap_id)
    return fail(h, FW_ERR_ARG, std::string(who) + ": bad argument");
  if (fmt != SGT_FP32 && fmt != SGT_BF16 && fmt != SGT_FP16)
    return fail(h, FW_ERR_ARG, std::string(who) + ": format must be FW_SGT_FP32, FW_SGT_BF16 or FW_SGT_FP16");
  Snapshot s{};
  const int ext[3] = {h->nx, h->ny, h->nz};
  for (int d = 0; d < 3; ++d) {
    s.subs[d] = subs[d] - 1; s.subc[d] = subc[d]; s.subt[d] = subt[d];
    if (subc[d] <= 0 || subt[d] <= 0 || s.subs[d] < 0 || s.subs[d] + (subc[d] - 1) * subt[d] >= ext[d])
      return fail(h, FW_ERR_ARG, std::string(who) + ": sub-volume outside the ghosted grid");
    s.nb[d] = (subc[d] + SGT_B - 1) / SGT_B;
  }
  s.tinv = tinv;
  s.nf = 0;
  for (int f = 0; f < 9; ++f)
    if (mask & (1u << f)) s.flist[s.nf++] = f;
  s.npts = (long)subc[0] * subc[1] * subc[2];
  s.nblk = s.nb[0] * s.nb[1] * s.nb[2];
  s.max_records = max_records;
  s.fmt = fmt;
  if ((int)h->snaps.size() >= MAX_SNAP) return fail(h, FW_ERR_ARG, std::string(who) + ": at most 16 snapshots");
  h->snaps.push_back(s);  // storage is sized by fw_init
  h->ready = false;
  *snap_id = (int)h->snaps.size() - 1;
  return FW_OK;
}

int fw_add_snapshot(fw_handle h, const int* subs, const int* subc, const int* subt, int tinv, unsigned mask,
                    int max_records, int* snap_id) {
  return add_snapshot(h, subs, subc, subt, tinv, mask, max_records, SGT_FP32, snap_id, "fw_add_snapshot");
}

int fw_add_snapshot_packed(fw_handle h, const int* subs, const int* subc, const int* subt, int tinv, unsigned mask,
                           int max_records, int format, int* snap_id) {
  return add_snapshot(h, subs, subc, subt, tinv, mask, max_records, format, snap_id, "fw_add_snapshot_packed");
}

// PML profiles, abs_init (mod_abs_npml.F90:135-300) with the helpers of
// :9070-9124, evaluated in float like the reference.
static float pml_R(int n) { return (float)std::pow(10.0, -(std::log10((double)n) - 1.0) / std::log10(2.0) - 3.0); }
static float pml_dmax(float L, float Vp, float Rpp) { return -Vp / 2.0f / L * (float)std::log((double)Rpp) * (2.0f + 1.0f); }
static float pml_d(float x, float L, float dmax) { return x < 0 ? 0.f : dmax * ((x / L) * (x / L)); }
static float pml_a(float x, float L, float amax) { return x < 0 ? 0.f : amax * (1.0f - (x / L)); }
static float pml_b(float x, float L, float bmax) { return x < 0 ? 1.f : 1.0f + (bmax - 1.0f) * ((x / L) * (x / L)); }

static void build_pml(fw_handle h) {
  const fw_config& c = h->cfg;
  h->Ax.assign(h->nx, 0.f); h->Bx.assign(h->nx, 1.f); h->Dx.assign(h->nx, 0.f);
  h->Ay.assign(h->ny, 0.f); h->By.assign(h->ny, 1.f); h->Dy.assign(h->ny, 0.f);
  h->Az.assign(h->nz, 0.f); h->Bz.assign(h->nz, 1.f); h->Dz.assign(h->nz, 0.f);
  const std::vector<float>&x = h->hx, &y = h->hy, &z = h->hz;
  const float ztop = z[h->nk2];
  int N;
  if ((N = c.abs_number[0]) > 0) {
    const float x0 = x[N + h->ni1], L0 = ztop * (x0 - x[h->ni1]);
    const float dm = pml_dmax(L0, c.abs_velocity[0], pml_R(N));
    for (int i = h->ni1; i <= N + h->ni1; ++i) {
      const float L = ztop * (x0 - x[i]);
      h->Dx[i] = pml_d(L, L0, dm); h->Ax[i] = pml_a(L, L0, c.cfs_amax[0]); h->Bx[i] = pml_b(L, L0, c.cfs_bmax[0]);
    }
  }
  if ((N = c.abs_number[1]) > 0) {
    const float x0 = x[h->ni2 - N], L0 = ztop * (x[h->ni2] - x0);
    const float dm = pml_dmax(L0, c.abs_velocity[1], pml_R(N));
    for (int i = h->ni2 - N; i <= h->ni2; ++i) {
      const float L = ztop * (x[i] - x0);
      h->Dx[i] = pml_d(L, L0, dm); h->Ax[i] = pml_a(L, L0, c.cfs_amax[1]); h->Bx[i] = pml_b(L, L0, c.cfs_bmax[1]);
    }
  }
  if ((N = c.abs_number[2]) > 0) {
    const float y0 = y[N + h->nj1], L0 = ztop * (y0 - y[h->nj1]);
    const float dm = pml_dmax(L0, c.abs_velocity[2], pml_R(N));
    for (int j = h->nj1; j <= N + h->nj1; ++j) {
      const float L = ztop * (y0 - y[j]);
      h->Dy[j] = pml_d(L, L0, dm); h->Ay[j] = pml_a(L, L0, c.cfs_amax[2]); h->By[j] = pml_b(L, L0, c.cfs_bmax[2]);
    }
  }
  if ((N = c.abs_number[3]) > 0) {
    const float y0 = y[h->nj2 - N], L0 = ztop * (y[h->nj2] - y0);
    const float dm = pml_dmax(L0, c.abs_velocity[3], pml_R(N));
    for (int j = h->nj2 - N; j <= h->nj2; ++j) {
      const float L = ztop * (y[j] - y0);
      h->Dy[j] = pml_d(L, L0, dm); h->Ay[j] = pml_a(L, L0, c.cfs_amax[3]); h->By[j] = pml_b(L, L0, c.cfs_bmax[3]);
    }
  }
  if ((N = c.abs_number[4]) > 0) {
    const float z0 = z[N + h->nk1], L0 = z0 - z[h->nk1];
    const float dm = pml_dmax(L0, c.abs_velocity[4], pml_R(N));
    for (int k = h->nk1; k <= h->nk1 + N; ++k) {
      const float L = z0 - z[k];
      h->Dz[k] = pml_d(L, L0, dm); h->Az[k] = pml_a(L, L0, c.cfs_amax[4]); h->Bz[k] = pml_b(L, L0, c.cfs_bmax[4]);
    }
  }
  if ((N = c.abs_number[5]) > 0) {
    const float z0 = z[h->nk2 - N], L0 = z[h->nk2] - z0;
    const float dm = pml_dmax(L0, c.abs_velocity[5], pml_R(N));
    for (int k = h->nk2 - N; k <= h->nk2; ++k) {
      const float L = z[k] - z0;
      h->Dz[k] = pml_d(L, L0, dm); h->Az[k] = pml_a(L, L0, c.cfs_amax[5]); h->Bz[k] = pml_b(L, L0, c.cfs_bmax[5]);
    }
  }
  for (int i = 0; i < h->nx; ++i) h->Dx[i] = h->Dx[i] / h->Bx[i];
  for (int j = 0; j < h->ny; ++j) h->Dy[j] = h->Dy[j] / h->By[j];
  for (int k = 0; k < h->nz; ++k) h->Dz[k] = h->Dz[k] / h->Bz[k];
}

static int build_sources(fw_handle h) {
  for (void* p : h->d_src_allocs) hipFree(p);
  h->d_src_allocs.clear();
  h->n_srccell = 0;
  h->n_surf = 0;
  auto dalloc = [&](void** p, size_t bytes, const void* src) -> int {
    if (bytes == 0) bytes = 16;
    if (hipMalloc(p, bytes) != hipSuccess) return -1;
    h->d_src_allocs.push_back(*p);
    if (src && hipMemcpy(*p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
  };
  const std::vector<float>&x = h->hx, &y = h->hy, &z = h->hz;
  // mode C: the radial extent along the column and the radius of the cell
  auto rad = [&](int i, int j, int k) { return h->curv ? h->hr3[i + (size_t)h->nx * (j + (size_t)h->ny * k)] : z[k]; };
  auto volume = [&](int i, int j, int k) {
    const float rk = rad(i, j, k);
    return (x[i + 1] - x[i - 1]) * (y[j + 1] - y[j - 1]) * (rad(i, j, k + 1) - rad(i, j, k - 1)) / 8.0f * (rk * rk) *
           h->hxsin[i];
  };
  auto norm_delt = [&](float* d, float x0, float y0, float z0, int sk) {
    const float r0 = 1.5f;
    for (int k = -3; k <= 3; ++k)
      for (int j = -3; j <= 3; ++j)
        for (int i = -3; i <= 3; ++i)
          d[(i + 3) + 7 * (j + 3) + 49 * (k + 3)] =
              fun_gauss((float)i - x0, r0, 0.f) * fun_gauss((float)j - y0, r0, 0.f) * fun_gauss((float)k - z0, r0, 0.f);
    float s = 0.f;
    for (int q = 0; q < 343; ++q) s += d[q];
    for (int q = 0; q < 343; ++q) d[q] = d[q] / s;
    if (h->freenode && sk + 3 > h->nk2) {  // mod_src.F90:230-232
      float s2 = 0.f;
      for (int k = -3; k <= h->nk2 - sk; ++k)
        for (int q = 0; q < 49; ++q) s2 += d[q + 49 * (k + 3)];
      for (int q = 0; q < 343; ++q) d[q] = d[q] / s2;
    }
  };
  // rows keyed by (member slot, cell); one surface plane map per slot
  std::map<std::pair<int, long>, std::vector<SrcContrib>> cells;
  std::vector<int> smap((size_t)h->ps * h->nbatch, -1);
  std::vector<SurfEntry> sent;
  std::map<long, int> surf_of;  // slot * ps + plane index -> entry
  auto fslot = [&](int n) { return h->nbatch > 1 ? h->slot_of[h->fmember[n]] : 0; };
  auto mslot = [&](int n) { return h->nbatch > 1 ? h->slot_of[h->mmember[n]] : 0; };
  float d[343];
  for (int n = 0; n < h->nforce; ++n) {
    const int si = h->findx[3 * n] - 1, sj = h->findx[3 * n + 1] - 1, sk = h->findx[3 * n + 2] - 1;
    // surface traction entries (src_surface), last (n, m) wins
    if (h->freenode && si >= h->ni1 && si <= h->ni2 && sj >= h->nj1 && sj <= h->nj2 && sk == h->nk2) {
      // projected area at the surface radius (mode C: sigma . N = F / S, N's
      // length cancels the slope's area)
      const float zs = rad(si, sj, sk);
      const float S = (x[si + 1] - x[si - 1]) * (y[sj + 1] - y[sj - 1]) / 4.0f * (zs * zs) * h->hxsin[si];
      const long c2 = si + h->sy * (long)sj + h->ps * (long)fslot(n);
      auto it = surf_of.find(c2);
      SurfEntry e{n, h->fntw - 1, 0.f, S};
      if (it == surf_of.end()) {
        surf_of[c2] = (int)sent.size();
        sent.push_back(e);
      } else {
        sent[it->second] = e;
      }
    }
    norm_delt(d, h->fshift[3 * n], h->fshift[3 * n + 1], h->fshift[3 * n + 2], sk);
    for (int Lk = -3; Lk <= 3; ++Lk)
      for (int Lj = -3; Lj <= 3; ++Lj)
        for (int Li = -3; Li <= 3; ++Li) {
          const int i = Li + si, j = Lj + sj, k = Lk + sk;
          if (i < h->ni1 || i > h->ni2 || j < h->nj1 || j > h->nj2 || k < h->nk1 || k > h->nk2) continue;
          if (h->freenode && k == h->nk2) continue;  // SrcSurface: surface plane carried by traction
          const long c = i + h->sy * (long)j + h->sz * (long)k;
          const float V = volume(i, j, k);
          const float w = 1.0f / (V * h->hrho[i + (size_t)h->nx * (j + (size_t)h->ny * k)]) * d[(Li + 3) + 7 * (Lj + 3) + 49 * (Lk + 3)];
          cells[{fslot(n), c}].push_back(SrcContrib{0, n, 0, w});
        }
  }
  for (int n = 0; n < h->nmom; ++n) {
    const int si = h->mindx[3 * n] - 1, sj = h->mindx[3 * n + 1] - 1, sk = h->mindx[3 * n + 2] - 1;
    norm_delt(d, h->mshift[3 * n], h->mshift[3 * n + 1], h->mshift[3 * n + 2], sk);
    for (int Lk = -3; Lk <= 3; ++Lk)
      for (int Lj = -3; Lj <= 3; ++Lj)
        for (int Li = -3; Li <= 3; ++Li) {
          const int i = Li + si, j = Lj + sj, k = Lk + sk;
          if (i < h->ni1 || i > h->ni2 || j < h->nj1 || j > h->nj2 || k < h->nk1 || k > h->nk2) continue;
          const long c = i + h->sy * (long)j + h->sz * (long)k;
          const float V = volume(i, j, k);
          cells[{mslot(n), c}].push_back(SrcContrib{1, n, 0, d[(Li + 3) + 7 * (Lj + 3) + 49 * (Lk + 3)] / V});
        }
  }
  for (auto& kv : surf_of) {
    const SurfEntry& e = sent[kv.second];
    smap[kv.first] = kv.second;
    const int slot = (int)(kv.first / h->ps);
    const long c = kv.first % h->ps + h->sz * (long)h->nk2;
    cells[{slot, c}].push_back(SrcContrib{2, e.n, e.m, 1.0f / e.S});
  }
  // device tables
  std::vector<long> cell;
  std::vector<int> ptr{0}, rowslot;
  std::vector<SrcContrib> con;
  for (auto& kv : cells) {
    rowslot.push_back(kv.first.first);
    cell.push_back(kv.first.second);
    for (auto& s : kv.second) con.push_back(s);
    ptr.push_back((int)con.size());
  }
  h->n_srccell = (int)cell.size();
  h->n_surf = (int)sent.size();
  int bad = 0;
  bad |= dalloc((void**)&h->d_scell, sizeof(long) * cell.size(), cell.data());
  bad |= dalloc((void**)&h->d_srow, sizeof(int) * rowslot.size(), rowslot.data());
  bad |= dalloc((void**)&h->d_sptr, sizeof(int) * ptr.size(), ptr.data());
  bad |= dalloc((void**)&h->d_scon, sizeof(SrcContrib) * con.size(), con.data());
  bad |= dalloc((void**)&h->d_smap, sizeof(int) * smap.size(), smap.data());
  bad |= dalloc((void**)&h->d_sent, sizeof(SurfEntry) * sent.size(), sent.data());
  bad |= dalloc((void**)&h->d_ft0, sizeof(float) * h->ft0.size(), h->ft0.data());
  bad |= dalloc((void**)&h->d_ftime, sizeof(float) * h->ftime.size(), h->ftime.data());
  bad |= dalloc((void**)&h->d_ffreq, sizeof(float) * h->ffreq.size(), h->ffreq.data());
  bad |= dalloc((void**)&h->d_F, sizeof(float) * h->F.size(), h->F.data());
  bad |= dalloc((void**)&h->d_mt0, sizeof(float) * h->mt0.size(), h->mt0.data());
  bad |= dalloc((void**)&h->d_mtime, sizeof(float) * h->mtime.size(), h->mtime.data());
  bad |= dalloc((void**)&h->d_mfreq, sizeof(float) * h->mfreq.size(), h->mfreq.data());
  bad |= dalloc((void**)&h->d_M, sizeof(float) * h->M.size(), h->M.data());
  return bad ? fail(h, FW_ERR_NOMEM, "build_sources: device allocation failed") : FW_OK;
}

static int zero_state(fw_handle h) {
  const long n = (long)h->nbatch * h->bstride + h->off;
  hipLaunchKernelGGL(k_fill<float>, dim3(fill_blocks(n)), dim3(256), 0, h->stream, h->d_sets, n, 0.f);
  const long ns = 6L * h->ps * h->nbatch;
  hipLaunchKernelGGL(k_fill<float>, dim3(fill_blocks(ns)), dim3(256), 0, h->stream, h->d_msurf, ns, 0.f);
  for (auto& f : h->face)
    if (f.on) {
      const long m = 4L * 6 * f.cnt * h->nbatch;
      hipLaunchKernelGGL(k_fill<float>, dim3(fill_blocks(m)), dim3(256), 0, h->stream, f.d, m, 0.f);
    }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  h->nrec = 0;
  for (auto& s : h->snaps) s.nrec = 0;
  h->stage_next = 0;
  return FW_OK;
}

static int build_items(fw_handle h);

// Device storage sized by the member count: register sets, surface copies,
// receiver and snapshot records; record counters reset.
static int alloc_state(fw_handle h) {
  const int B = h->nbatch;
  hipFree(h->d_sets); hipFree(h->d_msurf); hipFree(h->d_rec);
  h->d_sets = nullptr; h->d_msurf = nullptr; h->d_rec = nullptr;
  HIPCHK(hipMalloc(&h->d_sets, sizeof(float) * ((size_t)B * h->bstride + 64)));
  for (int q = 0; q < 4; ++q) h->set[q] = h->d_sets + h->off + (size_t)q * 9 * h->fs;
  HIPCHK(hipMalloc(&h->d_msurf, sizeof(float) * 6 * (size_t)h->ps * B));
  if (h->npt > 0 && h->max_rec > 0) HIPCHK(hipMalloc(&h->d_rec, sizeof(float) * 9 * (size_t)h->npt * h->max_rec * B));
  h->nrec = 0;
  h->rec_cnt.assign(B, 0);
  for (auto& sn : h->snaps) {
    hipFree(sn.d); hipFree(sn.codes); hipFree(sn.scales);
    sn.d = nullptr; sn.codes = nullptr; sn.scales = nullptr;
    const size_t nv = (size_t)sn.nf * sn.npts * sn.max_records * B;
    if (sn.fmt == SGT_FP32) {
      HIPCHK(hipMalloc(&sn.d, sizeof(float) * nv));
    } else {
      HIPCHK(hipMalloc(&sn.codes, sizeof(uint16_t) * nv));
      HIPCHK(hipMalloc(&sn.scales, sizeof(float) * (size_t)sn.nf * sn.nblk * sn.max_records * B));
    }
    sn.nrec = 0;
    sn.cnt.assign(B, 0);
  }
  for (auto& kv : h->gexec) hipGraphExecDestroy(kv.second);
  h->gexec.clear();
  return FW_OK;
}

int fw_set_batch(fw_handle h, int nbatch, const int* force_member, const int* moment_member, const int* nsteps) {
  if (!h || nbatch < 1 || nbatch > 4095 || !nsteps || (h->nforce > 0 && !force_member) ||
      (h->nmom > 0 && !moment_member))
    return fail(h, FW_ERR_ARG, "fw_set_batch: bad argument (1 <= nbatch <= 4095)");
  std::vector<int> fm(force_member, force_member + h->nforce), mm(moment_member, moment_member + h->nmom);
  for (int v : fm)
    if (v < 0 || v >= nbatch) return fail(h, FW_ERR_ARG, "fw_set_batch: force member out of range");
  for (int v : mm)
    if (v < 0 || v >= nbatch) return fail(h, FW_ERR_ARG, "fw_set_batch: moment member out of range");
  for (int b = 0; b < nbatch; ++b)
    if (nsteps[b] < 0) return fail(h, FW_ERR_ARG, "fw_set_batch: negative record length");
  std::vector<int> order(nbatch);
  for (int b = 0; b < nbatch; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nsteps[a] > nsteps[b]; });
  h->nbatch = nbatch;
  h->slot_of.assign(nbatch, 0);
  h->mem_steps.assign(nbatch, 0);
  for (int q = 0; q < nbatch; ++q) {
    h->slot_of[order[q]] = q;
    h->mem_steps[q] = nsteps[order[q]];
  }
  h->fmember = fm;
  h->mmember = mm;
  h->sel = 0;
  h->ready = false;
  return FW_OK;
}

int fw_init(fw_handle h) {
  if (!h) return FW_ERR_ARG;
  if (!h->have_grid || !h->have_media) return fail(h, FW_ERR_STATE, "fw_init: grid and media must be set first");
  if (h->nbatch > 1 && ((int)h->fmember.size() != h->nforce || (int)h->mmember.size() != h->nmom))
    return fail(h, FW_ERR_STATE, "fw_init: sources changed after fw_set_batch; call fw_set_batch again");
  {
    int rc = alloc_state(h);
    if (rc) return rc;
  }
  build_pml(h);
  // device PML profiles: per axis b1 = 1/B, ad = A + D/B, dd = D/B
  hipFree(h->d_pml);
  const long n1 = h->nx + h->ny + h->nz;
  std::vector<float> prof(3 * n1);
  auto put = [&](long base, const std::vector<float>& A, const std::vector<float>& B, const std::vector<float>& D) {
    for (size_t q = 0; q < A.size(); ++q) {
      prof[base + q] = 1.0f / B[q];
      prof[n1 + base + q] = A[q] + D[q];
      prof[2 * n1 + base + q] = D[q];
    }
  };
  put(0, h->Ax, h->Bx, h->Dx);
  put(h->nx, h->Ay, h->By, h->Dy);
  put(h->nx + h->ny, h->Az, h->Bz, h->Dz);
  HIPCHK(hipMalloc(&h->d_pml, sizeof(float) * prof.size()));
  HIPCHK(hipMemcpy(h->d_pml, prof.data(), sizeof(float) * prof.size(), hipMemcpyHostToDevice));
  for (int n = 0; n < 6; ++n) {
    PmlFace& f = h->face[n];
    hipFree(f.d);
    f = PmlFace{};
    const int N = h->cfg.abs_number[n];
    if (N <= 0) continue;
    f.on = true; f.n = N;
    f.lo[0] = h->ni1; f.hi[0] = h->ni2; f.lo[1] = h->nj1; f.hi[1] = h->nj2; f.lo[2] = h->nk1; f.hi[2] = h->nk2;
    const int ax = n / 2;
    if (n % 2 == 0) f.hi[ax] = f.lo[ax] + N; else f.lo[ax] = f.hi[ax] - N;
    f.cnt = 1;
    for (int d = 0; d < 3; ++d) { f.ex[d] = f.hi[d] - f.lo[d] + 1; f.cnt *= f.ex[d]; }
    // the stage kernel addresses a member's 24 auxiliary volumes of a face
    // with 32-bit offsets
    if (24 * f.cnt >= (1L << 32)) return fail(h, FW_ERR_ARG, "CFS-PML slab of 2^32/24 or more cells");
    HIPCHK(hipMalloc(&f.d, sizeof(float) * 4 * 6 * (size_t)f.cnt * h->nbatch));
  }
  if (h->curv) {
    // mode C: the CFS-PML corrections are the spherical scheme's, so every
    // slab cell must sit on a flat stretch of grid (slopes 0, radius z(k));
    // fluid cells are not supported on a terrain-following grid
    // (a cell whose +-3 box is flat has p = q = 0, F(p) = F(q) = 0, the 1-D
    // zeta_r and 1/z: the non-flat cells are dilated by 3 along each axis)
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const size_t n = (size_t)nx * ny * nz;
    std::vector<uint8_t> bad(n), tmp(n);
    for (int k = 0; k < nz; ++k)
      for (size_t u = (size_t)nx * ny * k; u < (size_t)nx * ny * (k + 1); ++u) bad[u] = h->hr3[u] != h->hz[k];
    const long st[3] = {1, nx, (long)nx * ny};
    const int ext[3] = {nx, ny, nz};
    for (int ax = 0; ax < 3; ++ax) {
      // each line along axis ax: a sliding count of bad cells over +-3
      const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;  // the two other axes
      for (int b = 0; b < ext[o2]; ++b)
        for (int a = 0; a < ext[o1]; ++a) {
          const size_t base = (size_t)a * st[o1] + (size_t)b * st[o2];
          const int L = ext[ax];
          int cnt = 0;
          for (int c = 0; c < 3 && c < L; ++c) cnt += bad[base + c * st[ax]];
          for (int c = 0; c < L; ++c) {
            if (c + 3 < L) cnt += bad[base + (c + 3) * st[ax]];
            if (c - 4 >= 0) cnt -= bad[base + (c - 4) * st[ax]];
            tmp[base + c * st[ax]] = cnt > 0;
          }
        }
      bad.swap(tmp);
    }
    for (int n6 = 0; n6 < 6; ++n6) {
      const PmlFace& f = h->face[n6];
      if (!f.on) continue;
      for (int k = f.lo[2]; k <= f.hi[2]; ++k)
        for (int j = f.lo[1]; j <= f.hi[1]; ++j)
          for (int i = f.lo[0]; i <= f.hi[0]; ++i)
            if (bad[i + (size_t)nx * (j + (size_t)ny * k)])
              return fail(h, FW_ERR_ARG, "fw_init: mode C needs a flat grid (r = z(k)) on every CFS-PML slab cell "
                                         "and 3 cells beyond");
    }
    for (size_t u = 0; u < n; ++u)
      if (h->hmu[u] <= SEIS_ZERO) return fail(h, FW_ERR_ARG, "fw_init: mode C does not support fluid cells (mu <= 0)");
  }
  int rc = build_items(h);
  if (rc) return rc;
  rc = build_sources(h);
  if (rc) return rc;
  rc = zero_state(h);
  if (rc) return rc;
  h->ready = true;
  return FW_OK;
}

int fw_get_pml_profiles(fw_handle h, float* A, float* B, float* D) {
  if (!h || !A || !B || !D) return fail(h, FW_ERR_ARG, "fw_get_pml_profiles: null argument");
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_get_pml_profiles: call fw_init first");
  std::copy(h->Ax.begin(), h->Ax.end(), A); std::copy(h->Ay.begin(), h->Ay.end(), A + h->nx);
  std::copy(h->Az.begin(), h->Az.end(), A + h->nx + h->ny);
  std::copy(h->Bx.begin(), h->Bx.end(), B); std::copy(h->By.begin(), h->By.end(), B + h->nx);
  std::copy(h->Bz.begin(), h->Bz.end(), B + h->nx + h->ny);
  std::copy(h->Dx.begin(), h->Dx.end(), D); std::copy(h->Dy.begin(), h->Dy.end(), D + h->nx);
  std::copy(h->Dz.begin(), h->Dz.end(), D + h->nx + h->ny);
  return FW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kernel dispatch over the tile shapes, 8 F/B patterns and 4 RK forms (the
// instantiations live in fw_stage_inst.hip).
// ---------------------------------------------------------------------------
static StageFn stage_table(int c, int p, int form, bool curv) {
  static StageFn (*const T[NTILECFG][2])(int, int) = {{stage_fn_0_0, stage_fn_0_1}, {stage_fn_1_0, stage_fn_1_1},
                                                       {stage_fn_2_0, stage_fn_2_1}, {stage_fn_3_0, stage_fn_3_1}};
  static StageFn (*const TC[NTILECFG][2])(int, int) = {{stage_fnc_0_0, stage_fnc_0_1}, {stage_fnc_1_0, stage_fnc_1_1},
                                                        {stage_fnc_2_0, stage_fnc_2_1}, {stage_fnc_3_0, stage_fnc_3_1}};
  return (curv ? TC : T)[c][(p >> 2) & 1](p, form);
}

// Work items of the stage kernels: tiles over chunks of k planes in three
// parts: narrow (TXN-column) tiles covering the remainder of the x extent
// after whole TX-column tiles when it is at most 32 columns (narrow body,
// with the CFS-PML corrections), wide tiles that touch a CFS-PML slab (PML
// body), wide interior tiles (lean body).  Each part is ordered by chunk,
// tile row, tile column.
static int build_items(fw_handle h) {
  const int TY = TILE_BYT[h->tilecfg] * TILE_RY[h->tilecfg];
  const int TYN = TY * (TX / TXN);
  const PmlFace* F = h->face;
  // A wide tile that touches an x or y slab runs the PML body over its whole
  // depth; cutting tiles at the slab boundaries instead costs an extra
  // (mostly idle) tile per row, which measures slower than running 51
  // interior columns through the PML body.
  struct Tile { int lo, n; bool pml; };
  auto tile = [](int a, int n, const PmlFace& f1, const PmlFace& f2, int ax) {
    return Tile{a, n, (f1.on && a <= f1.hi[ax]) || (f2.on && a + n - 1 >= f2.lo[ax])};
  };
  auto tiles = [&](int lo, int hi, int w, const PmlFace& f1, const PmlFace& f2, int ax) {
    std::vector<Tile> v;
    for (int a = lo; a <= hi; a += w) v.push_back(tile(a, std::min(w, hi - a + 1), f1, f2, ax));
    return v;
  };
  const int ni = h->ni2 - h->ni1 + 1, rem = ni % TX;
  const int nwide = rem > 0 && rem <= 32 && ni > TX ? ni / TX : (ni + TX - 1) / TX;
  std::vector<Tile> xt, xn;  // wide, narrow
  for (int t = 0; t < nwide; ++t) {
    const int a = h->ni1 + t * TX;
    xt.push_back(tile(a, std::min(TX, h->ni2 - a + 1), F[0], F[1], 0));
  }
  for (int a = h->ni1 + nwide * TX; a <= h->ni2; a += TXN) xn.push_back(tile(a, std::min(TXN, h->ni2 - a + 1), F[0], F[1], 0));
  const std::vector<Tile> yt = tiles(h->nj1, h->nj2, TY, F[2], F[3], 1);
  const std::vector<Tile> ytn = tiles(h->nj1, h->nj2, TYN, F[2], F[3], 1);
  const int zlo = F[4].on ? F[4].hi[2] : h->nk1 - 1;  // last plane of the z1 slab
  const int zhi = F[5].on ? F[5].lo[2] : h->nk2 + 1;  // first plane of the z2 slab
  const int nk = h->nk2 - h->nk1 + 1;
  const int ntile = (int)(xt.size() * yt.size() + xn.size() * ytn.size());
  // items of about 14 planes, 32768 / TY of them at 400x400x200 (8192 for
  // 4-row tiles): enough items to keep every XCD busy to the end of the
  // launch, few enough that the z-window restarts at item starts stay small
  // (the stage kernel at 400x400x200: 2048 items 1157 us, 4096 1139, 8192
  // 1113, 12288 1125, 24576 1157, 49152 1186)
  int target = 32768 / TY;
  if (const char* ev = getenv("FWANT_ITEM_TARGET")) target = std::max(1, atoi(ev));
  int nch = std::max(1, std::min(std::max(1, nk / 4), (target + ntile - 1) / ntile));
  const int cs = std::max(1, (nk + nch - 1) / nch);
  std::vector<int4> narrow, lean, pml;
  auto add_range = [&](std::vector<int4>& v, const Tile& x, const Tile& y, int ka, int kb) {
    for (int k0 = ka; k0 <= kb; k0 += cs)
      v.push_back(make_int4(x.lo | (x.n << 16), y.lo | (y.n << 16), k0, std::min(k0 + cs - 1, kb)));
  };
  for (const Tile& y : ytn)
    for (const Tile& x : xn) add_range(narrow, x, y, h->nk1, h->nk2);
  for (const Tile& y : yt)
    for (const Tile& x : xt) {
      if (x.pml || y.pml) {
        add_range(pml, x, y, h->nk1, h->nk2);
      } else {
        // the same chunks over the whole depth as every other tile; one that
        // holds planes of a z slab runs the PML body
        for (int k0 = h->nk1; k0 <= h->nk2; k0 += cs) {
          const int k1 = std::min(k0 + cs - 1, h->nk2);
          const int4 it = make_int4(x.lo | (x.n << 16), y.lo | (y.n << 16), k0, k1);
          (k0 <= zlo || k1 >= zhi ? pml : lean).push_back(it);
        }
      }
    }
  // items of a chunk in row-major (0) or column-major (1) tile order
  int colmajor = 0;
  if (const char* ev = getenv("FWANT_ITEM_ORDER")) colmajor = atoi(ev);
  auto order = [colmajor](const int4& a, const int4& b) {
    if (a.z != b.z) return a.z < b.z;
    const int ay = a.y & 0xffff, by = b.y & 0xffff, ax = a.x & 0xffff, bx = b.x & 0xffff;
    if (colmajor) return ax != bx ? ax < bx : ay < by;
    return ay != by ? ay < by : ax < bx;
  };
  std::stable_sort(narrow.begin(), narrow.end(), order);
  std::stable_sort(lean.begin(), lean.end(), order);
  std::stable_sort(pml.begin(), pml.end(), order);
  hipFree(h->d_items);
  h->d_items = nullptr;
  h->n_narrow = (int)narrow.size();
  h->xg = std::max(1, ((int)xt.size() + 1) / 2);  // half a tile row (3 of 6 measured 1% faster than 1, 2, 4, 6, 12)
  if (const char* ev = getenv("FWANT_XG")) h->xg = std::max(1, atoi(ev));
  h->n_lean = (int)lean.size();
  h->n_pml = (int)pml.size();
  std::vector<int4> all(narrow);
  all.insert(all.end(), pml.begin(), pml.end());
  all.insert(all.end(), lean.begin(), lean.end());
  if (!all.empty()) {
    HIPCHK(hipMalloc(&h->d_items, sizeof(int4) * all.size()));
    HIPCHK(hipMemcpy(h->d_items, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice));
  }
  return FW_OK;
}

extern "C" {

// One RK stage st (0..3) of step ntime for the nact running members, with
// the F/B pattern of the 8-step cycle, or pattern `pat` ((xf << 2) | (yf << 1)
// | zf) when it is >= 0.
static int launch_stage(fw_handle h, int ntime, int st, int nact, int pat = -1) {
  if (pat < 0) {
    const int* bp = BLOCK_PAT[(((ntime - h->cycle_origin) % 8) + 8) % 8];
    const int flip = st & 1;
    const int xf = flip ? !bp[0] : bp[0], yf = flip ? !bp[1] : bp[1], zf = flip ? !bp[2] : bp[2];
    pat = (xf << 2) | (yf << 1) | zf;
  }
  // register rotation: stage in/out sets
  static const int IN[4] = {0, 1, 2, 1}, OUT[4] = {1, 2, 1, 0};
  float* win = h->set[IN[st]];
  float* wout = h->set[OUT[st]];
  float* A = h->set[0];
  float* T = h->set[3];
  static const float RKA[4] = {0.5f, 0.5f, 1.0f, 0.0f};
  const float RKB[4] = {1.0f / 6.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 6.0f};
  const float tinc = st == 0 ? 0.f : RKA[st - 1];
  const float a = RKA[st] * h->cfg.stept, b = RKB[st] * h->cfg.stept;
  // b_{s-1} / a_{s-1}: the share of W_s - m that the accumulator takes at
  // the RK forms 1 and 3 (1/3 at both)
  const float r = st == 1 || st == 3 ? RKB[st - 1] / RKA[st - 1] : 0.f;

  if (h->freenode) {
    SurfArgs s{};
    s.win = win; s.msurf = h->d_msurf; s.smap = h->d_smap; s.ent = h->d_sent;
    s.ft0 = h->d_ft0; s.ftime = h->d_ftime; s.ffreq = h->d_ffreq; s.F = h->d_F;
    s.fs = h->fs; s.sy = h->sy; s.sz = h->sz; s.ps = h->ps; s.fblk = (long)h->fntw * h->nforce;
    s.bstride = h->bstride; s.clk = h->d_clk;
    s.ni1 = h->ni1; s.ni2 = h->ni2; s.nj1 = h->nj1; s.nj2 = h->nj2; s.nk2 = h->nk2;
    s.ntw = h->fntw; s.stf_id = h->fstf; s.save = st == 0 ? 1 : (st == 2 ? 0 : 2); s.tinc = tinc; s.stept = h->cfg.stept;
    s.cp = h->curv ? h->met3[0] : nullptr; s.cq = h->curv ? h->met3[1] : nullptr;
    dim3 blk(64, 4), grd((h->cfg.ni + 63) / 64, (h->cfg.nj + 3) / 4, nact);
    hipLaunchKernelGGL(k_surface, grd, blk, 0, h->stream, s);
  }
  StageArgs g{};
  g.win = win; g.wout = wout; g.m = A; g.t = T;
  g.lam = h->lam; g.mu = h->mu; g.rrho = h->rrho;
  g.xi_x = h->xi_x; g.eta_y = h->eta_y; g.zeta_z = h->zeta_z; g.rz = h->rz; g.rxsin = h->rxsin; g.xcot = h->xcot;
  g.msurf = h->d_msurf;
  g.fs = h->fs; g.sy = h->sy; g.sz = h->sz; g.ps = h->ps;
  g.ni1 = h->ni1; g.ni2 = h->ni2; g.nj1 = h->nj1; g.nj2 = h->nj2; g.nk1 = h->nk1; g.nk2 = h->nk2;
  g.freenode = h->freenode; g.a = a; g.b = b; g.r = r;
  g.c = st == 1 ? 1.0f - r - RKB[2] / RKA[2] : 0.f;
  g.nx = h->nx; g.ny = h->ny; g.fluidcol = h->d_fluidcol;
  g.n1 = h->nx + h->ny + h->nz;
  g.pb1 = h->d_pml; g.pad = h->d_pml + g.n1; g.pdd = h->d_pml + 2 * g.n1;
  g.setin = IN[st]; g.setout = OUT[st];
  if (h->curv) {
    g.cp = h->met3[0]; g.cq = h->met3[1]; g.czz = h->met3[2]; g.crr = h->met3[3];
    const bool zf = (pat & 1) != 0;
    g.cdp = h->met3[zf ? 4 : 6]; g.cdq = h->met3[zf ? 5 : 7];
  }
  g.pml = 0;
  for (int n = 0; n < 6; ++n) {
    const PmlFace& f = h->face[n];
    FaceArgs& fa = g.face[n];
    fa.on = f.on ? 1 : 0;
    g.pml |= fa.on;
    fa.d = f.d; fa.cnt = f.cnt;
    fa.lo0 = f.lo[0]; fa.lo1 = f.lo[1]; fa.lo2 = f.lo[2];
    fa.hi0 = f.hi[0]; fa.hi1 = f.hi[1]; fa.hi2 = f.hi[2];
    fa.ex0 = f.ex[0]; fa.ex1 = f.ex[1];
  }
  const int tc = h->tilecfg;
  const dim3 sblk(TX, 2 * TILE_BYT[tc]);
  if (h->timing) hipEventRecord(h->ev[2 * h->nstage], h->stream);
  g.items = h->d_items;
  g.nitems = h->n_narrow + h->n_pml + h->n_lean;
  g.p1 = h->n_narrow;
  g.p2 = h->n_narrow + h->n_pml;
  g.xg = h->xg;
  g.bstride = h->bstride;
  // a dispatch holds fewer than 2^32 work-items: at most maxm members per launch
  const long maxm = std::max(1L, 0xffffffffL / ((long)std::max(1, g.nitems) * sblk.x * sblk.y));
  const int W = (int)std::min<long>(h->width > 0 ? h->width : nact, maxm);
  for (int s0 = 0; s0 < nact && g.nitems > 0; s0 += W) {  // members s0 .. s0+W-1 in one launch
    g.slot0 = s0;
    g.nm = std::min(W, nact - s0);
    const dim3 grid(g.nitems * g.nm);
    hipLaunchKernelGGL(stage_table(tc, pat, st, h->curv), grid, sblk, 0, h->stream, g);
  }
  if (h->timing) {
    hipEventRecord(h->ev[2 * h->nstage + 1], h->stream);
    h->nstage++;
  }
  // point sources
  if (h->n_srccell > 0) {
    SrcArgs s{};
    s.wout = wout; s.t = T; s.lam = h->lam; s.mu = h->mu;
    s.cell = h->d_scell; s.ptr = h->d_sptr; s.con = h->d_scon;
    s.ft0 = h->d_ft0; s.ftime = h->d_ftime; s.ffreq = h->d_ffreq; s.F = h->d_F;
    s.fblk = (long)h->fntw * h->nforce; s.fntw = h->fntw; s.fstf = h->fstf;
    s.mt0 = h->d_mt0; s.mtime = h->d_mtime; s.mfreq = h->d_mfreq; s.M = h->d_M;
    s.mblk = (long)h->mntw * h->nmom; s.mntw = h->mntw; s.mstf = h->mstf;
    s.fs = h->fs; s.bstride = h->bstride; s.rowslot = h->d_srow; s.clk = h->d_clk;
    s.ncell = h->n_srccell; s.nact = nact; s.kind = st;
    s.tinc = tinc; s.stept = h->cfg.stept; s.a = a; s.b = b;
    s.cp = h->curv ? h->met3[0] : nullptr; s.cq = h->curv ? h->met3[1] : nullptr;
    hipLaunchKernelGGL(k_src, dim3((h->n_srccell + 127) / 128), dim3(128), 0, h->stream, s);
  }
  return FW_OK;
}

// Host bookkeeping of the captures after step ntime (the device kernels
// find their record slots on the device clock): storage checks, then the
// record counts of the running members.
static int account_outputs(fw_handle h, int ntime, int nact) {
  const int t1 = ntime + 1;
  if (h->npt > 0 && h->d_rec && t1 % h->pt_tinv == 0) {
    if (h->nrec >= h->max_rec) return fail(h, FW_ERR_STATE, "receiver storage full (max_records)");
    h->nrec++;
    for (int q = 0; q < nact; ++q) h->rec_cnt[q] = h->nrec;
  }
  for (auto& sn : h->snaps) {
    if (t1 % sn.tinv != 0) continue;
    if (sn.nrec >= sn.max_records) return fail(h, FW_ERR_STATE, "snapshot storage full (max_records)");
    sn.nrec++;
    for (int q = 0; q < nact; ++q) sn.cnt[q] = sn.nrec;
  }
  return FW_OK;
}

// Output kernels after step ntime (all of them when recording a graph: each
// decides on the device clock whether this step is a capture step), then
// the clock tick.
static void enqueue_outputs(fw_handle h, int ntime, int nact, bool all) {
  const int t1 = ntime + 1;
  if (h->npt > 0 && h->d_rec && (all || t1 % h->pt_tinv == 0))
    hipLaunchKernelGGL(k_receivers, dim3((h->npt * 9 + 255) / 256, nact), dim3(256), 0, h->stream, h->set[0], h->fs,
                       h->bstride, h->d_pt, h->npt, h->d_rec, h->max_rec, h->pt_tinv, h->d_clk);
  for (size_t q = 0; q < h->snaps.size(); ++q) {
    const Snapshot& sn = h->snaps[q];
    if (!all && t1 % sn.tinv != 0) continue;
    if (sn.fmt != SGT_FP32) {
      PackArgs a{};
      a.W = h->set[0]; a.codes = sn.codes; a.scales = sn.scales;
      a.fs = h->fs; a.sy = h->sy; a.sz = h->sz; a.npts = sn.npts; a.bstride = h->bstride;
      a.cstride = (long)sn.max_records * sn.nf * sn.npts; a.sstride = (long)sn.max_records * sn.nf * sn.nblk;
      a.clk = h->d_clk; a.cidx = 2 + (int)q; a.tinv = sn.tinv; a.max_records = sn.max_records;
      a.s0 = sn.subs[0]; a.s1 = sn.subs[1]; a.s2 = sn.subs[2];
      a.c0 = sn.subc[0]; a.c1 = sn.subc[1]; a.c2 = sn.subc[2];
      a.t0 = sn.subt[0]; a.t1 = sn.subt[1]; a.t2 = sn.subt[2];
      a.nb0 = sn.nb[0]; a.nb1 = sn.nb[1]; a.nblk = sn.nblk; a.nf = sn.nf; a.fmt = sn.fmt;
      for (int f = 0; f < sn.nf; ++f) a.flist[f] = sn.flist[f];
      hipLaunchKernelGGL(k_snapshot_pack, dim3((unsigned)sn.nblk, (unsigned)sn.nf, (unsigned)nact), dim3(256), 0,
                         h->stream, a);
      continue;
    }
    SnapArgs a{};
    a.W = h->set[0]; a.out = sn.d;
    a.fs = h->fs; a.sy = h->sy; a.sz = h->sz; a.npts = sn.npts; a.bstride = h->bstride;
    a.mstride = (long)sn.max_records * sn.nf * sn.npts;
    a.clk = h->d_clk; a.cidx = 2 + (int)q; a.tinv = sn.tinv; a.max_records = sn.max_records;
    a.s0 = sn.subs[0]; a.s1 = sn.subs[1]; a.s2 = sn.subs[2]; a.c0 = sn.subc[0]; a.c1 = sn.subc[1];
    a.t0 = sn.subt[0]; a.t1 = sn.subt[1]; a.t2 = sn.subt[2]; a.nf = sn.nf;
    for (int f = 0; f < sn.nf; ++f) a.flist[f] = sn.flist[f];
    hipLaunchKernelGGL(k_snapshot, dim3((unsigned)((sn.npts + 255) / 256), (unsigned)nact), dim3(256), 0, h->stream, a);
  }
  TickArgs tk{};
  tk.clk = h->d_clk;
  tk.rec_tinv = (h->npt > 0 && h->d_rec) ? h->pt_tinv : 0;
  tk.nsnap = (int)h->snaps.size();
  for (size_t q = 0; q < h->snaps.size(); ++q) tk.tinv[q] = h->snaps[q].tinv;
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, h->stream, tk);
}

// members still running at step ntime (slots are ordered longest first)
static int running(fw_handle h, int ntime) {
  int n = 0;
  while (n < h->nbatch && h->mem_steps[n] > ntime) ++n;
  return n;
}

static int enqueue_step(fw_handle h, int ntime, int nact, bool all) {
  for (int st = 0; st < 4; ++st) {
    int rc = launch_stage(h, ntime, st, nact);
    if (rc) return rc;
  }
  enqueue_outputs(h, ntime, nact, all);
  return FW_OK;
}

// The step as a graph: launches depend on ntime only through its block
// position (the F/B pattern); times and record slots come from the clock.
static int step_graph(fw_handle h, int ntime, int nact, hipGraphExec_t* out) {
  const long key = (long)((((ntime - h->cycle_origin) % 8) + 8) % 8) * 4096 + nact;
  auto it = h->gexec.find(key);
  if (it != h->gexec.end()) {
    *out = it->second;
    return FW_OK;
  }
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(h, ntime, nact, true);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(h, FW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(h, FW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  h->gexec[key] = ex;
  *out = ex;
  return FW_OK;
}

static int set_clock(fw_handle h, int ntime) {
  ClockSet c{};
  c.n = 2 + (int)h->snaps.size();
  c.v[0] = ntime;
  c.v[1] = h->nrec;
  for (size_t q = 0; q < h->snaps.size(); ++q) c.v[2 + q] = h->snaps[q].nrec;
  hipLaunchKernelGGL(k_setclk, dim3(1), dim3(64), 0, h->stream, h->d_clk, c);
  HIPCHK(hipGetLastError());
  return FW_OK;
}

int fw_step(fw_handle h, int ntime) { return fw_run(h, ntime, 1); }

// One RK stage at a time, in the order a Fortran driver calls the
// reference's routines (see fwant.h).  Stage 0 starts step ntime (device
// clock, output accounting), stage 3 ends it (receivers, snapshots, clock
// tick); the stages of a step must come in order 0, 1, 2, 3.
int fw_stage(fw_handle h, int ntime, int stage, int xf, int yf, int zf) {
  if (!h) return FW_ERR_ARG;
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_stage: call fw_init first");
  if (ntime < 0 || stage < 0 || stage > 3 || (xf | yf | zf) & ~1)
    return fail(h, FW_ERR_ARG, "fw_stage: stage must be 0..3, ntime >= 0 and the F/B flags 0 or 1");
  if (stage != h->stage_next || (stage > 0 && ntime != h->stage_ntime))
    return fail(h, FW_ERR_STATE, "fw_stage: the stages of a step must run in order 0, 1, 2, 3");
  int rc;
  if (stage == 0) {
    h->stage_nact = running(h, ntime);
    h->stage_ntime = ntime;
    if (h->stage_nact > 0) {
      if ((rc = set_clock(h, ntime))) return rc;
      if ((rc = account_outputs(h, ntime, h->stage_nact))) return rc;
    }
  }
  if (h->stage_nact > 0) {
    if ((rc = launch_stage(h, ntime, stage, h->stage_nact, (xf << 2) | (yf << 1) | zf))) return rc;
    if (stage == 3) enqueue_outputs(h, ntime, h->stage_nact, false);
    HIPCHK(hipGetLastError());
  }
  h->stage_next = (stage + 1) & 3;
  return FW_OK;
}

int fw_enable_timing(fw_handle h, int on) {
  if (!h) return FW_ERR_ARG;
  h->timing = on != 0;
  return FW_OK;
}

int fw_run(fw_handle h, int ntime0, int nsteps) {
  if (!h) return FW_ERR_ARG;
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_run: call fw_init first");
  if (nsteps < 0 || ntime0 < 0) return fail(h, FW_ERR_ARG, "fw_run: ntime0 and nsteps must be >= 0");
  if (h->stage_next != 0) return fail(h, FW_ERR_STATE, "fw_run: a step begun with fw_stage is not finished");
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (h->timing) {
    const size_t need = 2 * (size_t)nsteps * 4;
    while (h->ev.size() < need) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      h->ev.push_back(e);
    }
    h->nstage = 0;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    hipEventRecord(t0, h->stream);
  }
  int rc = set_clock(h, ntime0);
  if (rc) return rc;
  for (int n = 0; n < nsteps; ++n) {
    const int ntime = ntime0 + n;
    const int nact = running(h, ntime);
    if (nact == 0) break;  // every member has run its record length
    if ((rc = account_outputs(h, ntime, nact))) return rc;
    if (h->use_graphs && !h->timing) {
      hipGraphExec_t ex = nullptr;
      if ((rc = step_graph(h, ntime, nact, &ex))) return rc;
      HIPCHK(hipGraphLaunch(ex, h->stream));
    } else {
      if ((rc = enqueue_step(h, ntime, nact, false))) return rc;
    }
  }
  HIPCHK(hipGetLastError());
  if (h->timing) {
    hipEventRecord(t1, h->stream);
    HIPCHK(hipEventSynchronize(t1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, t0, t1);
    h->total_ms = ms;
    double s = 0;
    for (int q = 0; q < h->nstage; ++q) {
      float m = 0.f;
      hipEventElapsedTime(&m, h->ev[2 * q], h->ev[2 * q + 1]);
      s += m;
    }
    h->stage_ms = s;
    hipEventDestroy(t0);
    hipEventDestroy(t1);
  }
  return FW_OK;
}

int fw_get_timing(fw_handle h, double* stage_ms, double* total_ms, int* nst) {
  if (!h) return FW_ERR_ARG;
  if (stage_ms) *stage_ms = h->stage_ms;
  if (total_ms) *total_ms = h->total_ms;
  if (nst) *nst = h->nstage;
  return FW_OK;
}

int fw_sync(fw_handle h) {
  if (!h) return FW_ERR_ARG;
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

// slot of the member the get/set calls address
static int sel_slot(fw_handle h) { return h->nbatch > 1 ? h->slot_of[h->sel] : 0; }

int fw_get_receivers(fw_handle h, float* out, int* nrecords) {
  if (!h) return FW_ERR_ARG;
  HIPCHK(hipStreamSynchronize(h->stream));
  const int sl = sel_slot(h);
  const int n = h->rec_cnt.empty() ? 0 : h->rec_cnt[sl];
  if (nrecords) *nrecords = n;
  if (out && n > 0 && h->npt > 0)
    HIPCHK(hipMemcpy(out, h->d_rec + (size_t)sl * h->max_rec * h->npt * 9, sizeof(float) * 9 * (size_t)h->npt * n,
                     hipMemcpyDeviceToHost));
  return FW_OK;
}

static int snap_check(fw_handle h, int id, bool packed, const char* who) {
  if (!h || id < 0 || id >= (int)h->snaps.size()) return fail(h, FW_ERR_ARG, std::string(who) + ": bad id");
  if (!h->ready) return fail(h, FW_ERR_STATE, std::string(who) + ": call fw_init first");
  if (packed && h->snaps[id].fmt == SGT_FP32) return fail(h, FW_ERR_ARG, std::string(who) + ": FP32 snapshot");
  if (!packed && h->snaps[id].fmt != SGT_FP32) return fail(h, FW_ERR_ARG, std::string(who) + ": packed snapshot");
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

int fw_get_snapshot(fw_handle h, int id, float* out, int* nrecords) {
  if (int rc = snap_check(h, id, false, "fw_get_snapshot")) return rc;
  const Snapshot& s = h->snaps[id];
  const int sl = sel_slot(h), n = s.cnt[sl];
  if (nrecords) *nrecords = n;
  if (out && n > 0)
    HIPCHK(hipMemcpy(out, s.d + (size_t)sl * s.max_records * s.nf * s.npts, sizeof(float) * (size_t)s.nf * s.npts * n,
                     hipMemcpyDeviceToHost));
  return FW_OK;
}

int fw_get_snapshot_packed(fw_handle h, int id, uint16_t* codes, float* scales, int* nrecords, long* nblocks) {
  if (int rc = snap_check(h, id, true, "fw_get_snapshot_packed")) return rc;
  const Snapshot& s = h->snaps[id];
  const int sl = sel_slot(h), n = s.cnt[sl];
  if (nrecords) *nrecords = n;
  if (nblocks) *nblocks = s.nblk;
  if (codes && n > 0)
    HIPCHK(hipMemcpy(codes, s.codes + (size_t)sl * s.max_records * s.nf * s.npts,
                     sizeof(uint16_t) * (size_t)s.nf * s.npts * n, hipMemcpyDeviceToHost));
  if (scales && n > 0)
    HIPCHK(hipMemcpy(scales, s.scales + (size_t)sl * s.max_records * s.nf * s.nblk,
                     sizeof(float) * (size_t)s.nf * s.nblk * n, hipMemcpyDeviceToHost));
  return FW_OK;
}

int fw_snapshot_packed_device(fw_handle h, int id, uint16_t** codes, float** scales, int* nrecords, long* npoints,
                              long* nblocks) {
  if (int rc = snap_check(h, id, true, "fw_snapshot_packed_device")) return rc;
  const Snapshot& s = h->snaps[id];
  const int sl = sel_slot(h);
  if (codes) *codes = s.codes + (size_t)sl * s.max_records * s.nf * s.npts;
  if (scales) *scales = s.scales + (size_t)sl * s.max_records * s.nf * s.nblk;
  if (nrecords) *nrecords = s.cnt[sl];
  if (npoints) *npoints = s.npts;
  if (nblocks) *nblocks = s.nblk;
  return FW_OK;
}

int fw_snapshot_device(fw_handle h, int id, float** dptr, int* nrecords, long* npoints) {
  if (int rc = snap_check(h, id, false, "fw_snapshot_device")) return rc;
  const Snapshot& s = h->snaps[id];
  const int sl = sel_slot(h);
  if (dptr) *dptr = s.d + (size_t)sl * s.max_records * s.nf * s.npts;
  if (nrecords) *nrecords = s.cnt[sl];
  if (npoints) *npoints = s.npts;
  return FW_OK;
}

int fw_get_wavefield(fw_handle h, int f, float* out) {
  if (!h || f < 0 || f >= 9 || !out) return fail(h, FW_ERR_ARG, "fw_get_wavefield: bad argument");
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_get_wavefield: call fw_init first");
  return download_volume(h, out, h->set[0] + (size_t)sel_slot(h) * h->bstride + (size_t)f * h->fs);
}

int fw_set_wavefield(fw_handle h, int f, const float* in) {
  if (!h || f < 0 || f >= 9 || !in) return fail(h, FW_ERR_ARG, "fw_set_wavefield: bad argument");
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_set_wavefield: call fw_init first");
  const size_t off = (size_t)sel_slot(h) * h->bstride + (size_t)f * h->fs;
  int rc = upload_volume(h, h->set[0] + off, in);
  if (rc) return rc;
  // The stages read their input from sets A, B and C in turn, and no stage
  // writes the x / y ghost layers: the array goes to all three, so that its
  // ghost values are seen by every stage, as with the reference's single
  // array per field (the interiors of B and C are written before any read).
  for (int q = 1; q <= 2; ++q)
    HIPCHK(hipMemcpyAsync(h->set[q] + off, h->set[0] + off, sizeof(float) * h->fs, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

int fw_set_cycle_origin(fw_handle h, int origin) {
  if (!h) return FW_ERR_ARG;
  h->cycle_origin = origin;
  return FW_OK;
}

// Exact-restart state of the selected member: register set A (the
// wavefield between steps) and set A of every CFS-PML face's auxiliaries.
static size_t checkpoint_floats(fw_handle h) {
  size_t n = (size_t)9 * h->nx * h->ny * h->nz;
  for (const auto& f : h->face)
    if (f.on) n += 6 * (size_t)f.cnt;
  return n;
}

int fw_checkpoint_size(fw_handle h, size_t* bytes) {
  if (!h || !bytes) return FW_ERR_ARG;
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_checkpoint_size: call fw_init first");
  *bytes = sizeof(float) * checkpoint_floats(h);
  return FW_OK;
}

int fw_checkpoint_save(fw_handle h, void* buf) {
  if (!h || !buf) return fail(h, FW_ERR_ARG, "fw_checkpoint_save: null argument");
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_checkpoint_save: call fw_init first");
  float* o = (float*)buf;
  const size_t vol = (size_t)h->nx * h->ny * h->nz;
  for (int f = 0; f < 9; ++f) {
    int rc = fw_get_wavefield(h, f, o + f * vol);
    if (rc) return rc;
  }
  o += 9 * vol;
  const int sl = sel_slot(h);
  for (const auto& f : h->face) {
    if (!f.on) continue;
    HIPCHK(hipMemcpy(o, f.d + (size_t)sl * 24 * f.cnt, sizeof(float) * 6 * f.cnt, hipMemcpyDeviceToHost));
    o += 6 * f.cnt;
  }
  return FW_OK;
}

int fw_checkpoint_load(fw_handle h, const void* buf) {
  if (!h || !buf) return fail(h, FW_ERR_ARG, "fw_checkpoint_load: null argument");
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_checkpoint_load: call fw_init first");
  const float* in = (const float*)buf;
  const size_t vol = (size_t)h->nx * h->ny * h->nz;
  for (int f = 0; f < 9; ++f) {
    int rc = fw_set_wavefield(h, f, in + f * vol);
    if (rc) return rc;
  }
  in += 9 * vol;
  const int sl = sel_slot(h);
  for (const auto& f : h->face) {
    if (!f.on) continue;
    HIPCHK(hipMemcpy(f.d + (size_t)sl * 24 * f.cnt, in, sizeof(float) * 6 * f.cnt, hipMemcpyHostToDevice));
    in += 6 * f.cnt;
  }
  return FW_OK;
}

int fw_select_member(fw_handle h, int member) {
  if (!h || member < 0 || member >= h->nbatch) return fail(h, FW_ERR_ARG, "fw_select_member: member out of range");
  h->sel = member;
  return FW_OK;
}

int fw_use_graphs(fw_handle h, int on) {
  if (!h) return FW_ERR_ARG;
  h->use_graphs = on != 0;
  return FW_OK;
}

int fw_set_batch_width(fw_handle h, int width) {
  if (!h || width < 0) return fail(h, FW_ERR_ARG, "fw_set_batch_width: width must be >= 0");
  h->width = width;
  for (auto& kv : h->gexec) hipGraphExecDestroy(kv.second);
  h->gexec.clear();
  return FW_OK;
}

static void reset_records(fw_handle h) {
  h->nrec = 0;
  std::fill(h->rec_cnt.begin(), h->rec_cnt.end(), 0);
  for (auto& sn : h->snaps) {
    sn.nrec = 0;
    std::fill(sn.cnt.begin(), sn.cnt.end(), 0);
  }
}

int fw_autotune(fw_handle h, int trial_steps, int* tile, int* width, float* ms_per_step) {
  if (!h || trial_steps < 1) return fail(h, FW_ERR_ARG, "fw_autotune: trial_steps must be >= 1");
  if (!h->ready) return fail(h, FW_ERR_STATE, "fw_autotune: call fw_init first");
  const bool timing = h->timing;
  h->timing = false;
  std::vector<int> widths{0};
  for (int w = 1; w < h->nbatch; w *= 2) widths.push_back(w);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 3.4e38f;
  int bt = h->tilecfg, bw = h->width, rc = FW_OK;
  auto trial = [&](int tc, int w) -> int {
    h->tilecfg = tc;
    int r = build_items(h);
    if (!r) r = fw_set_batch_width(h, w);
    if (!r) r = zero_state(h);
    if (r) return r;
    reset_records(h);
    if ((r = fw_run(h, 0, 1))) return r;  // warm-up (graph capture, first touch)
    hipEventRecord(e0, h->stream);
    if ((r = fw_run(h, 1, trial_steps))) return r;
    hipEventRecord(e1, h->stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) { best = ms; bt = tc; bw = w; }
    return FW_OK;
  };
  // the tile shapes with all running members in one launch, then the batch
  // widths with the fastest shape (NTILECFG + widths - 1 trials)
  for (int tc = 0; tc < NTILECFG && !rc; ++tc) rc = trial(tc, 0);
  const int tbest = bt;
  for (size_t q = 1; q < widths.size() && !rc; ++q) rc = trial(tbest, widths[q]);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  h->timing = timing;
  h->tilecfg = bt;
  if (!rc) rc = build_items(h);
  if (!rc) rc = fw_set_batch_width(h, bw);
  if (!rc) rc = zero_state(h);
  reset_records(h);
  if (rc) return rc;
  if (tile) *tile = bt;
  if (width) *width = bw;
  if (ms_per_step) *ms_per_step = best / trial_steps;
  return FW_OK;
}

}  // extern "C"
