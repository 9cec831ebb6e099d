// libfwant: sensitivity-kernel time integral of FWANT's SI_ker_sta
// (code/srcF.kernel/SI_ker_sta.F90:1185-1557) on the GPU.
//
// Series layout: component c, sample n, point p at base[c*cs + n*ts + p];
// threads run over points, so every access is coalesced.  Products are FP32,
// accumulations FP64 (the compensated form of the reference's FP32 sums).
#include "fw_common.h"
#include "../../include/fwant.h"

#include <mutex>
#include <string>
#include <vector>

using fw::SGT_B;
using fw::SGT_BF16;
using fw::SGT_FP16;
using fw::sgt_value;

namespace {
thread_local std::string g_kerr;
int kfail(int code, const std::string& m) {
  g_kerr = m;
  return code;
}
}  // namespace
// shared with the other fwk_* sources (fw_inv.hip)
int fwk_fail(int code, const std::string& m) { return kfail(code, m); }
namespace {
#define KCHK(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return kfail(FW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// stress2strain, SI_ker_sta.F90:1185-1211
__global__ void k_stress2strain(float* __restrict__ E, long cs, long ts, long npt, int knt,
                                const float* __restrict__ lam, const float* __restrict__ mu) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  const float l = lam[p], u = mu[p];
  const float E1 = (l + u) / (u * (3.0f * l + 2.0f * u));
  const float E2 = -l / (2.0f * u * (3.0f * l + 2.0f * u));
  const float E3 = 1.0f / u;
  for (int n = blockIdx.y; n < knt; n += gridDim.y) {
    float* e = E + (long)n * ts + p;
    const float xx = e[0], yy = e[cs], zz = e[2 * cs];
    const float E0 = E2 * (xx + yy + zz);
    e[0] = E0 - (E2 - E1) * xx;
    e[cs] = E0 - (E2 - E1) * yy;
    e[2 * cs] = E0 - (E2 - E1) * zz;
    e[3 * cs] = 0.5f * E3 * e[3 * cs];
    e[4 * cs] = 0.5f * E3 * e[4 * cs];
    e[5 * cs] = 0.5f * E3 * e[5 * cs];
  }
}

// reduced-precision SGT restore: out[slab][p] = code * scale of p's 16^3 block
__global__ void k_sgt_decode(int fmt, const uint16_t* __restrict__ codes, const float* __restrict__ scales, int c0, int c1,
                             long npts, int nb0, int nb1, long nblk, float* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const long slab = blockIdx.y;
  const int a = (int)(p % c0);
  const long r = p / c0;
  const int b = (int)(r % c1), d = (int)(r / c1);
  const long blk = a / SGT_B + (long)nb0 * (b / SGT_B + (long)nb1 * (d / SGT_B));
  out[slab * npts + p] = sgt_value(fmt, codes[slab * npts + p], scales[slab * nblk + blk]);
}

// sgt2scatter3 for one receiver component (SI_ker_sta.F90:1290-1331).
// Each thread owns one point and NB consecutive outputs n0..n0+NB-1 (1-based);
// it walks m = 1 .. n0+NB-2 keeping a sliding register window of the source
// strain S(n-m) for its NB outputs, so R(m) is loaded once per m.
// PARAM selects the output pair: 0 alpha/beta (2(l+2u) x1, 4u (x2-x1)),
// 1 lambda/mu (l x1, 2u x2; SI_ker_sta_lamu.F90:1333-1336).
constexpr int NB = 8;
template <int PARAM>
__global__ __launch_bounds__(256) void k_scatter(const float* __restrict__ R, long rcs, long rts,
                                                 const float* __restrict__ S, long scs, long sts, long npt, int knt,
                                                 float dt, const float* __restrict__ lam, const float* __restrict__ mu,
                                                 float* __restrict__ outA, float* __restrict__ outB) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  // heavy tiles (large n0) first: blockIdx.y = 0 gets the last n block
  const int nblk = (knt + NB - 1) / NB;
  const int n0 = 1 + (nblk - 1 - (int)blockIdx.y) * NB;
  float s[6][NB], tr[NB];
  double a1[NB], a2[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    a1[q] = 0.0;
    a2[q] = 0.0;
    const int idx = n0 + q - 1;  // S index for m = 1
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c][q] = (idx >= 1 && idx <= knt) ? S[c * scs + (long)(idx - 1) * sts + p] : 0.f;
    tr[q] = s[0][q] + s[1][q] + s[2][q];
  }
  const int mmax = min(n0 + NB - 2, knt - 1);
  for (int m = 1; m <= mmax; ++m) {
    const float* r = R + (long)(m - 1) * rts + p;
    const float r0 = r[0], r1 = r[rcs], r2 = r[2 * rcs], r3 = r[3 * rcs], r4 = r[4 * rcs], r5 = r[5 * rcs];
    const float trR = r0 + r1 + r2;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const float x1 = trR * tr[q];
      const float x2 = r0 * s[0][q] + r1 * s[1][q] + r2 * s[2][q] +
                       2.0f * (r3 * s[3][q] + r4 * s[4][q] + r5 * s[5][q]);
      a1[q] += (double)x1;
      a2[q] += (double)x2;
    }
    // slide the window: s[q] <- S(n0 + q - (m+1))
#pragma unroll
    for (int q = NB - 1; q > 0; --q) {
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c][q] = s[c][q - 1];
      tr[q] = tr[q - 1];
    }
    const int idx = n0 - m - 1;
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c][0] = idx >= 1 ? S[c * scs + (long)(idx - 1) * sts + p] : 0.f;
    tr[0] = s[0][0] + s[1][0] + s[2][0];
  }
  const double l = lam[p], u = mu[p];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int n = n0 + q;
    if (n > knt) continue;
    const double x1 = a1[q] * (double)dt, x2 = a2[q] * (double)dt;
    if constexpr (PARAM == 0) {
      outA[(long)(n - 1) * npt + p] = (float)(x1 * 2.0 * (l + 2.0 * u));
      outB[(long)(n - 1) * npt + p] = (float)(2.0 * (x2 - x1) * 2.0 * u);
    } else {
      outA[(long)(n - 1) * npt + p] = (float)(x1 * l);
      outB[(long)(n - 1) * npt + p] = (float)(2.0 * x2 * u);
    }
  }
}

// The same convolution with the loads shared through LDS: a block of 256
// threads owns 64 consecutive points and 32 consecutive outputs n0..n0+31,
// as 4 groups of 8 outputs (one thread per point and group, the register
// window of k_scatter).  Per m the block loads R(m) once into LDS for all
// four groups, and one new S row: group g needs S(n0 + 8g - m - 1) at the
// end of step m, which group 0 needed 8g steps earlier, so the S rows live
// in a 32-row LDS ring indexed by row mod 32 (at most 25 rows are live).
// Global traffic per block and m: 2 x 6 x 64 floats for 2048 (n, m) pairs.
constexpr int SC_P = 64, SC_G = 4, SC_T = SC_G * NB, SC_RING = 32;
template <int PARAM>
__global__ __launch_bounds__(256) void k_scatter_t(const float* __restrict__ R, long rcs, long rts,
                                                   const float* __restrict__ S, long scs, long sts, long npt, int knt,
                                                   float dt, const float* __restrict__ lam,
                                                   const float* __restrict__ mu, float* __restrict__ outA,
                                                   float* __restrict__ outB) {
  __shared__ float sR[2][6][SC_P];
  __shared__ float sS[SC_RING][6][SC_P];
  const int tid = threadIdx.x, pl = tid % SC_P, g = tid / SC_P;
  const long p0 = (long)blockIdx.x * SC_P, p = p0 + pl;
  // heavy tiles (large n0) first: blockIdx.y = 0 gets the last n block
  const int nblk = (knt + SC_T - 1) / SC_T;
  const int n0 = 1 + (nblk - 1 - (int)blockIdx.y) * SC_T;
  const int ng = n0 + g * NB;  // this thread's first output
  const int mmax = min(n0 + SC_T - 2, knt - 1);
  // S row r (1-based; zero outside 1..knt) of point p0 + q into ring slot r mod 32
  auto s_row = [&](int r, int c, int q) -> float {
    const long pq = p0 + q;
    return (r >= 1 && r <= knt && pq < npt) ? S[c * scs + (long)(r - 1) * sts + pq] : 0.f;
  };
  auto r_row = [&](int m, int c, int q) -> float {
    const long pq = p0 + q;
    return (m >= 1 && m <= mmax && pq < npt) ? R[c * rcs + (long)(m - 1) * rts + pq] : 0.f;
  };
  // prologue: S rows n0-1 .. n0+30 (the windows at m = 1) and R(1)
  for (int j = tid; j < SC_RING * 6 * SC_P; j += 256) {
    const int q = j % SC_P, c = (j / SC_P) % 6, k = j / (6 * SC_P);
    const int r = n0 - 1 + k;
    sS[((r % SC_RING) + SC_RING) % SC_RING][c][q] = s_row(r, c, q);
  }
  for (int j = tid; j < 6 * SC_P; j += 256) sR[1][j / SC_P][j % SC_P] = r_row(1, j / SC_P, j % SC_P);
  __syncthreads();
  float s[6][NB], tr[NB];
  double a1[NB], a2[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    a1[q] = 0.0;
    a2[q] = 0.0;
    const int slot = (((ng + q - 1) % SC_RING) + SC_RING) % SC_RING;
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c][q] = sS[slot][c][pl];
    tr[q] = s[0][q] + s[1][q] + s[2][q];
  }
  __syncthreads();  // every window read before the ring is refilled
  // rows stored at the end of step m: R(m+1) and S(n0 - m - 1) (group 0's
  // new row), 3 loads per thread, issued one step ahead
  auto rows_for = [&](int m, float (&v)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int j = tid + 256 * k, q = j % SC_P, c = (j / SC_P) % 6;
      v[k] = j < 6 * SC_P ? r_row(m + 1, c, q) : s_row(n0 - m - 1, c, q);
    }
  };
  float ld[3], ldn[3];
  rows_for(1, ld);
  for (int m = 1; m <= mmax; ++m) {
    rows_for(m + 1, ldn);
    const float* rr = &sR[m & 1][0][0];
    const float r0 = rr[pl], r1 = rr[SC_P + pl], r2 = rr[2 * SC_P + pl], r3 = rr[3 * SC_P + pl],
                r4 = rr[4 * SC_P + pl], r5 = rr[5 * SC_P + pl];
    const float trR = r0 + r1 + r2;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const float x1 = trR * tr[q];
      const float x2 = r0 * s[0][q] + r1 * s[1][q] + r2 * s[2][q] +
                       2.0f * (r3 * s[3][q] + r4 * s[4][q] + r5 * s[5][q]);
      a1[q] += (double)x1;
      a2[q] += (double)x2;
    }
    const int rnew = n0 - m - 1, snew = ((rnew % SC_RING) + SC_RING) % SC_RING;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int j = tid + 256 * k, q = j % SC_P, c = (j / SC_P) % 6;
      if (j < 6 * SC_P)
        sR[(m + 1) & 1][c][q] = ld[k];
      else
        sS[snew][c][q] = ld[k];
    }
    __syncthreads();
    // slide the window: s[q] <- S(ng + q - (m+1))
#pragma unroll
    for (int q = NB - 1; q > 0; --q) {
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c][q] = s[c][q - 1];
      tr[q] = tr[q - 1];
    }
    const int slot = (((ng - m - 1) % SC_RING) + SC_RING) % SC_RING;
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c][0] = sS[slot][c][pl];
    tr[0] = s[0][0] + s[1][0] + s[2][0];
#pragma unroll
    for (int k = 0; k < 3; ++k) ld[k] = ldn[k];
  }
  if (p >= npt) return;
  const double l = lam[p], u = mu[p];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int n = ng + q;
    if (n > knt) continue;
    const double x1 = a1[q] * (double)dt, x2 = a2[q] * (double)dt;
    if constexpr (PARAM == 0) {
      outA[(long)(n - 1) * npt + p] = (float)(x1 * 2.0 * (l + 2.0 * u));
      outB[(long)(n - 1) * npt + p] = (float)(2.0 * (x2 - x1) * 2.0 * u);
    } else {
      outA[(long)(n - 1) * npt + p] = (float)(x1 * l);
      outB[(long)(n - 1) * npt + p] = (float)(2.0 * x2 * u);
    }
  }
}

// vel2scatter3 for one receiver component (SI_ker_sta_rho.F90:1478-1552):
// x1(n) = dt sum_{m<n} (VxR+VyR+VzR)(m) (VxS+VyS+VzS)(n-m), output x1*rho.
// Same sliding-window structure as k_scatter on the component sums.
__global__ __launch_bounds__(256) void k_vscatter(const float* __restrict__ R, long rcs, long rts,
                                                  const float* __restrict__ S, long scs, long sts, long npt, int knt,
                                                  float dt, const float* __restrict__ rho, float* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  const int nblk = (knt + NB - 1) / NB;
  const int n0 = 1 + (nblk - 1 - (int)blockIdx.y) * NB;
  float sv[NB];
  double a1[NB];
  auto ssum = [&](int idx) {
    if (idx < 1 || idx > knt) return 0.f;
    const float* q = S + (long)(idx - 1) * sts + p;
    return q[0] + q[scs] + q[2 * scs];
  };
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    a1[q] = 0.0;
    sv[q] = ssum(n0 + q - 1);
  }
  const int mmax = min(n0 + NB - 2, knt - 1);
  for (int m = 1; m <= mmax; ++m) {
    const float* r = R + (long)(m - 1) * rts + p;
    const float rv = r[0] + r[rcs] + r[2 * rcs];
#pragma unroll
    for (int q = 0; q < NB; ++q) a1[q] += (double)(rv * sv[q]);
#pragma unroll
    for (int q = NB - 1; q > 0; --q) sv[q] = sv[q - 1];
    sv[0] = ssum(n0 - m - 1);
  }
  const double rh = rho[p];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int n = n0 + q;
    if (n > knt) continue;
    out[(long)(n - 1) * npt + p] = (float)(a1[q] * (double)dt * rh);
  }
}

// eoshift along time of ncomp series (Fortran semantics: shift > 0 moves
// sample n+shift to n, the vacated end is zero; shift < 0 the reverse),
// as SI_ker_EQ.F90 applies it to the receiver SGT.  In place: a thread owns
// one point and walks n in the direction that reads ahead of its writes.
__global__ void k_eoshift(float* __restrict__ E, long cs, long ts, long npt, int knt, int ncomp, int shift) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  for (int c = 0; c < ncomp; ++c) {
    float* e = E + c * cs + p;
    if (shift >= 0) {
      for (int n = 0; n < knt; ++n) e[(long)n * ts] = n + shift < knt ? e[(long)(n + shift) * ts] : 0.f;
    } else {
      for (int n = knt - 1; n >= 0; --n) e[(long)n * ts] = n + shift >= 0 ? e[(long)(n + shift) * ts] : 0.f;
    }
  }
}

// kernel re-parameterisation, in place on (a, b) phase and amplitude kernels:
// mode 0 lambda/mu -> kappa/mu (SI_ker_lamu2kamu.F90:823-830), 1 lambda/mu ->
// Vp/Vs (SI_ker_lamu2vpvs.F90:823-830), 2 Vp/Vs -> kappa/mu
// (SI_ker_vpvs2kamu.F90:823-831); FP32 as the reference.
__device__ __forceinline__ void reparam1(int mode, float l, float u, float& a, float& b) {
  if (mode == 0) {
    a = a / l;
    b = b / u / 2.0f;
    b = (2.0f * b - 2.0f * a / 3.0f) * u;
    a = a * (l + 2.0f * u / 3.0f);
  } else if (mode == 1) {
    a = a / l;
    b = b / u / 2.0f;
    b = 2.0f * (b - a) * 2.0f * u;
    a = a * 2.0f * (l + 2.0f * u);
  } else {
    a = a / (2.0f * (l + 2.0f * u));
    b = b / u / 4.0f + a;
    b = (2.0f * b - 2.0f * a / 3.0f) * u;
    a = a * (l + 2.0f * u / 3.0f);
  }
}
__global__ void k_reparam(int mode, long npt, const float* __restrict__ lam, const float* __restrict__ mu,
                          float* __restrict__ Kap, float* __restrict__ Kaq, float* __restrict__ Kbp,
                          float* __restrict__ Kbq) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  const float l = lam[p], u = mu[p];
  float ap = Kap[p], bp = Kbp[p], aq = Kaq[p], bq = Kbq[p];
  reparam1(mode, l, u, ap, bp);
  reparam1(mode, l, u, aq, bq);
  Kap[p] = ap; Kbp[p] = bp; Kaq[p] = aq; Kbq[p] = bq;
}

// IIR filter of SI_ker_sta.F90:1547-1557 over the padded series of
// scatter_filter (mode 1) / scatter_filtfilt (mode 2), :1433-1475.
constexpr int MAXNOD1 = 16;
struct FiltCoef {
  double b[MAXNOD1], a[MAXNOD1];
};
__device__ __forceinline__ double padded_x(const float* __restrict__ w0, long npt, long p, int n /*1-based*/, int knt,
                                           int NFACT, int mode) {
  if (n <= NFACT) {
    if (mode == 1) return 0.0;
    return 2.0 * (double)w0[p] - (double)w0[(long)(NFACT + 1 - n) * npt + p];
  }
  if (n <= NFACT + knt) return (double)w0[(long)(n - NFACT - 1) * npt + p];
  const int q = n - NFACT - knt;  // 1..NFACT
  return 2.0 * (double)w0[(long)(knt - 1) * npt + p] - (double)w0[(long)(knt - 1 - q) * npt + p];
}
__global__ __launch_bounds__(256) void k_filter(int mode, long npt, int knt, int nod1, FiltCoef cf,
                                                const float* __restrict__ w0, float* __restrict__ w,
                                                double* __restrict__ scratch) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  const int NOD = nod1 - 1, NFACT = 3 * NOD, NEL = knt + 2 * NFACT;
  double xh[MAXNOD1], yh[MAXNOD1];  // xh[m] = x(n-m), yh[m] = y(n-m)
#pragma unroll
  for (int q = 0; q < MAXNOD1; ++q) { xh[q] = 0.0; yh[q] = 0.0; }
  for (int n = 1; n <= NEL; ++n) {
#pragma unroll
    for (int q = MAXNOD1 - 1; q > 0; --q) xh[q] = xh[q - 1];
    xh[0] = padded_x(w0, npt, p, n, knt, NFACT, mode);
    double v = cf.b[0] * xh[0];
#pragma unroll
    for (int m = 2; m <= MAXNOD1; ++m)
      if (m <= nod1 && m <= n) v = v + cf.b[m - 1] * xh[m - 1] - cf.a[m - 1] * yh[m - 2];
    v = v / cf.a[0];
#pragma unroll
    for (int q = MAXNOD1 - 1; q > 0; --q) yh[q] = yh[q - 1];
    yh[0] = v;
    if (mode == 1) {
      if (n > NFACT && n <= NFACT + knt) w[(long)(n - NFACT - 1) * npt + p] = (float)v;
    } else {
      scratch[(long)(n - 1) * npt + p] = v;
    }
  }
  if (mode == 1) return;
  // backward pass over the reversed forward output
#pragma unroll
  for (int q = 0; q < MAXNOD1; ++q) { xh[q] = 0.0; yh[q] = 0.0; }
  for (int n = 1; n <= NEL; ++n) {
#pragma unroll
    for (int q = MAXNOD1 - 1; q > 0; --q) xh[q] = xh[q - 1];
    xh[0] = scratch[(long)(NEL - n) * npt + p];
    double v = cf.b[0] * xh[0];
#pragma unroll
    for (int m = 2; m <= MAXNOD1; ++m)
      if (m <= nod1 && m <= n) v = v + cf.b[m - 1] * xh[m - 1] - cf.a[m - 1] * yh[m - 2];
    v = v / cf.a[0];
#pragma unroll
    for (int q = MAXNOD1 - 1; q > 0; --q) yh[q] = yh[q - 1];
    yh[0] = v;
    // output sample NEL+1-n of the reversed result
    const int o = NEL + 1 - n;
    if (o > NFACT && o <= NFACT + knt) w[(long)(o - NFACT - 1) * npt + p] = (float)v;
  }
}

// scatter2kernel, SI_ker_sta.F90:1375-1431 (trapezoid on [nt1, nt2])
__global__ __launch_bounds__(256) void k_kernel(long npt, int nt1, int nt2, float dt, float V0, float U0,
                                                const float* __restrict__ V, const float* __restrict__ U,
                                                const float* __restrict__ Sa, const float* __restrict__ Sb,
                                                float* __restrict__ Kap, float* __restrict__ Kaq,
                                                float* __restrict__ Kbp, float* __restrict__ Kbq) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  auto sa = [&](int n) { return (double)Sa[(long)(n - 1) * npt + p]; };
  auto sb = [&](int n) { return (double)Sb[(long)(n - 1) * npt + p]; };
  const double v1 = V[nt1 - 1], v2 = V[nt2 - 1], u1 = U[nt1 - 1], u2 = U[nt2 - 1];
  double kap = 0.5 * (v1 * sa(nt1) + v2 * sa(nt2));
  double kaq = 0.5 * (u1 * sa(nt1) + u2 * sa(nt2));
  double kbp = 0.5 * (v1 * sb(nt1) + v2 * sb(nt2));
  double kbq = 0.5 * (u1 * sb(nt1) + u2 * sb(nt2));
  for (int n = nt1 + 1; n <= nt2 - 1; ++n) {
    const double a = sa(n), b = sb(n), v = V[n - 1], u = U[n - 1];
    kap += v * a;
    kaq += u * a;
    kbp += v * b;
    kbq += u * b;
  }
  Kap[p] = (float)(kap * (double)dt / (double)V0);
  Kaq[p] = (float)(kaq * (double)dt / (double)U0);
  Kbp[p] = (float)(kbp * (double)dt / (double)V0);
  Kbq[p] = (float)(kbq * (double)dt / (double)U0);
}

struct Rot {
  float r[9];
};
__global__ void k_rotate(long n, float* __restrict__ X, float* __restrict__ Y, float* __restrict__ Z, Rot R) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float v1 = X[q], v2 = Y[q], v3 = Z[q];
  X[q] = v1 * R.r[0] + v2 * R.r[1] + v3 * R.r[2];
  Y[q] = v1 * R.r[3] + v2 * R.r[4] + v3 * R.r[5];
  Z[q] = v1 * R.r[6] + v2 * R.r[7] + v3 * R.r[8];
}

unsigned nblocks(long n, int b) { return (unsigned)((n + b - 1) / b); }
}  // namespace

extern "C" {

const char* fwk_last_error(void) { return g_kerr.c_str(); }

int fwk_sgt_decode(int format, const uint16_t* codes, const float* scales, int c0, int c1, int c2, long nslab, float* out,
                   void* stream) {
  if (!codes || !scales || !out || c0 <= 0 || c1 <= 0 || c2 <= 0 || nslab <= 0 || nslab > 65535 ||
      (format != SGT_BF16 && format != SGT_FP16))
    return kfail(FW_ERR_ARG, "fwk_sgt_decode: bad argument");
  const long npts = (long)c0 * c1 * c2;
  const int nb0 = (c0 + SGT_B - 1) / SGT_B, nb1 = (c1 + SGT_B - 1) / SGT_B, nb2 = (c2 + SGT_B - 1) / SGT_B;
  hipLaunchKernelGGL(k_sgt_decode, dim3((unsigned)((npts + 255) / 256), (unsigned)nslab), dim3(256), 0,
                     (hipStream_t)stream, format, codes, scales, c0, c1, npts, nb0, nb1, (long)nb0 * nb1 * nb2, out);
  KCHK(hipGetLastError());
  return FW_OK;
}

// device memory helpers for hosts without their own HIP allocator
int fw_device_alloc(size_t bytes, void** dptr) {
  if (!dptr) return kfail(FW_ERR_ARG, "fw_device_alloc: null argument");
  *dptr = nullptr;
  KCHK(hipMalloc(dptr, bytes ? bytes : 16));
  return FW_OK;
}
int fw_device_free(void* dptr) {
  KCHK(hipFree(dptr));
  return FW_OK;
}
int fw_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  KCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FW_OK;
}
int fw_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  KCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FW_OK;
}
int fw_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  KCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return FW_OK;
}
int fw_device_sync(void) {
  KCHK(hipDeviceSynchronize());
  return FW_OK;
}

// Attainable HBM bandwidth: grid-stride streaming over 16-byte vectors with
// non-temporal accesses, read-only (a sum kept only to hold the loads) and
// copy (read + write, 1:1)
typedef float v4f_t __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_probe_copy(const v4f_t* __restrict__ a, v4f_t* __restrict__ b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ __launch_bounds__(256) void k_probe_read(const v4f_t* __restrict__ a, float* __restrict__ out, long n) {
  float s = 0.f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const v4f_t v = __builtin_nontemporal_load(a + i);
    s += v.x + v.y + v.z + v.w;
  }
  if (s == -1.2345e37f) out[0] = s;  // never true for the zero-filled buffer
}
int fw_hbm_probe(size_t bytes, int blocks, int reps, float* read_gbs, float* copy_gbs) {
  if (bytes < (1u << 20) || blocks <= 0 || reps <= 0 || !read_gbs || !copy_gbs)
    return kfail(FW_ERR_ARG, "fw_hbm_probe: bad argument");
  const long n = (long)(bytes / 16);
  void *a = nullptr, *b = nullptr, *o = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FW_OK;
  auto fin = [&]() {
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(a); hipFree(b); hipFree(o);
    return rc;
  };
  if (hipMalloc(&a, 16 * (size_t)n) != hipSuccess || hipMalloc(&b, 16 * (size_t)n) != hipSuccess ||
      hipMalloc(&o, 64) != hipSuccess || hipMemset(a, 0, 16 * (size_t)n) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = kfail(FW_ERR_NOMEM, "fw_hbm_probe: device buffers");
    return fin();
  }
  float ms[2] = {0.f, 0.f};
  for (int which = 0; which < 2; ++which) {
    for (int r = -1; r < reps; ++r) {  // one untimed launch first
      if (r == 0) hipEventRecord(e0, nullptr);
      if (which == 0)
        hipLaunchKernelGGL(k_probe_read, dim3(blocks), dim3(256), 0, nullptr, (const v4f_t*)a, (float*)o, n);
      else
        hipLaunchKernelGGL(k_probe_copy, dim3(blocks), dim3(256), 0, nullptr, (const v4f_t*)a, (v4f_t*)b, n);
    }
    hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[which], e0, e1) != hipSuccess) {
      rc = kfail(FW_ERR_HIP, "fw_hbm_probe: timing failed");
      return fin();
    }
  }
  *read_gbs = (float)(16.0 * n * reps / (ms[0] * 1e-3) / 1e9);
  *copy_gbs = (float)(2.0 * 16.0 * n * reps / (ms[1] * 1e-3) / 1e9);
  return fin();
}

int fwk_stress2strain(float* E, long cs, long ts, long npt, int knt, const float* lam, const float* mu, void* stream) {
  if (!E || !lam || !mu || npt <= 0 || knt <= 0) return kfail(FW_ERR_ARG, "fwk_stress2strain: bad argument");
  const unsigned gy = (unsigned)std::min(knt, 64);
  hipLaunchKernelGGL(k_stress2strain, dim3(nblocks(npt, 256), gy), dim3(256), 0, (hipStream_t)stream, E, cs, ts, npt,
                     knt, lam, mu);
  KCHK(hipGetLastError());
  return FW_OK;
}

int fwk_scatter(const float* R, long rcs, long rts, const float* S, long scs, long sts, long npt, int knt, float dt,
                const float* lam, const float* mu, float* outA, float* outB, void* stream) {
  if (!R || !S || !lam || !mu || !outA || !outB || npt <= 0 || knt <= 0)
    return kfail(FW_ERR_ARG, "fwk_scatter: bad argument");
  return fwk_scatter_param(0, R, rcs, rts, S, scs, sts, npt, knt, dt, lam, mu, outA, outB, stream);
}

int fwk_scatter_param(int param, const float* R, long rcs, long rts, const float* S, long scs, long sts, long npt,
                      int knt, float dt, const float* lam, const float* mu, float* outA, float* outB, void* stream) {
  if (!R || !S || !lam || !mu || !outA || !outB || npt <= 0 || knt <= 0 || (param != 0 && param != 1))
    return kfail(FW_ERR_ARG, "fwk_scatter_param: bad argument");
  // long series: the LDS-tiled kernel (262144 points, knt = 400: 20.9 ms
  // against 23.6 ms); shorter ones: the register-window kernel (knt = 100:
  // 2.35 against 3.25 ms, knt = 200: 6.64 against 7.15 ms)
  if (knt >= 320) {
    const unsigned gy = (unsigned)((knt + SC_T - 1) / SC_T);
    if (param == 0)
      hipLaunchKernelGGL(k_scatter_t<0>, dim3(nblocks(npt, SC_P), gy), dim3(256), 0, (hipStream_t)stream, R, rcs,
                         rts, S, scs, sts, npt, knt, dt, lam, mu, outA, outB);
    else
      hipLaunchKernelGGL(k_scatter_t<1>, dim3(nblocks(npt, SC_P), gy), dim3(256), 0, (hipStream_t)stream, R, rcs,
                         rts, S, scs, sts, npt, knt, dt, lam, mu, outA, outB);
  } else {
    const unsigned gy = (unsigned)((knt + NB - 1) / NB);
    if (param == 0)
      hipLaunchKernelGGL(k_scatter<0>, dim3(nblocks(npt, 256), gy), dim3(256), 0, (hipStream_t)stream, R, rcs, rts, S,
                         scs, sts, npt, knt, dt, lam, mu, outA, outB);
    else
      hipLaunchKernelGGL(k_scatter<1>, dim3(nblocks(npt, 256), gy), dim3(256), 0, (hipStream_t)stream, R, rcs, rts, S,
                         scs, sts, npt, knt, dt, lam, mu, outA, outB);
  }
  KCHK(hipGetLastError());
  return FW_OK;
}

int fwk_vscatter(const float* R, long rcs, long rts, const float* S, long scs, long sts, long npt, int knt, float dt,
                 const float* rho, float* out, void* stream) {
  if (!R || !S || !rho || !out || npt <= 0 || knt <= 0) return kfail(FW_ERR_ARG, "fwk_vscatter: bad argument");
  const unsigned gy = (unsigned)((knt + NB - 1) / NB);
  hipLaunchKernelGGL(k_vscatter, dim3(nblocks(npt, 256), gy), dim3(256), 0, (hipStream_t)stream, R, rcs, rts, S, scs,
                     sts, npt, knt, dt, rho, out);
  KCHK(hipGetLastError());
  return FW_OK;
}

int fwk_eoshift(float* E, long cs, long ts, long npt, int knt, int ncomp, int shift, void* stream) {
  if (!E || npt <= 0 || knt <= 0 || ncomp <= 0) return kfail(FW_ERR_ARG, "fwk_eoshift: bad argument");
  hipLaunchKernelGGL(k_eoshift, dim3(nblocks(npt, 256)), dim3(256), 0, (hipStream_t)stream, E, cs, ts, npt, knt, ncomp,
                     shift);
  KCHK(hipGetLastError());
  return FW_OK;
}

int fwk_reparam(int mode, long npt, const float* lam, const float* mu, float* Kap, float* Kaq, float* Kbp, float* Kbq,
                void* stream) {
  if (mode < 0 || mode > 2 || npt <= 0 || !lam || !mu || !Kap || !Kaq || !Kbp || !Kbq)
    return kfail(FW_ERR_ARG, "fwk_reparam: bad argument");
  hipLaunchKernelGGL(k_reparam, dim3(nblocks(npt, 256)), dim3(256), 0, (hipStream_t)stream, mode, npt, lam, mu, Kap,
                     Kaq, Kbp, Kbq);
  KCHK(hipGetLastError());
  return FW_OK;
}

int fwk_filter(int mode, long npt, int knt, int nod1, const double* b, const double* a, const float* w0, float* w,
               void* stream) {
  if ((mode != 1 && mode != 2) || npt <= 0 || knt <= 0 || nod1 < 1 || nod1 > MAXNOD1 || !b || !a || !w0 || !w)
    return kfail(FW_ERR_ARG, "fwk_filter: bad argument (mode 1|2, 1 <= nod1 <= 16)");
  if (knt <= 3 * (nod1 - 1)) return kfail(FW_ERR_ARG, "fwk_filter: knt must exceed 3*(nod1-1)");
  if (a[0] == 0.0) return kfail(FW_ERR_ARG, "fwk_filter: a(1) must be nonzero");
  FiltCoef cf{};
  for (int q = 0; q < nod1; ++q) { cf.b[q] = b[q]; cf.a[q] = a[q]; }
  const int NEL = knt + 6 * (nod1 - 1);
  double* fwd = nullptr;
  hipStream_t st = (hipStream_t)stream;
  // filtfilt keeps the whole forward-pass output (NEL doubles per point)
  if (mode == 2) KCHK(hipMallocAsync((void**)&fwd, sizeof(double) * (size_t)NEL * npt, st));
  hipLaunchKernelGGL(k_filter, dim3(nblocks(npt, 256)), dim3(256), 0, st, mode, npt, knt, nod1, cf, w0, w, fwd);
  KCHK(hipGetLastError());
  if (fwd) KCHK(hipFreeAsync(fwd, st));
  return FW_OK;
}

}  // extern "C"

// normalisation of SI_ker_sta.F90:1391-1397 (host, double): M0 * int V^2 dt
// and M0 * int U^2 dt by the trapezoid rule over [nt1, nt2]
static void s2k_norms(int nt1, int nt2, float dt, float M0, const float* V, const float* U, double* V0, double* U0) {
  double v = (0.5 * (double)V[nt1 - 1] * V[nt1 - 1] + 0.5 * (double)V[nt2 - 1] * V[nt2 - 1]) * dt;
  double u = (0.5 * (double)U[nt1 - 1] * U[nt1 - 1] + 0.5 * (double)U[nt2 - 1] * U[nt2 - 1]) * dt;
  for (int n = nt1 + 1; n <= nt2 - 1; ++n) {
    v += (double)V[n - 1] * V[n - 1] * dt;
    u += (double)U[n - 1] * U[n - 1] * dt;
  }
  *V0 = v * M0;
  *U0 = u * M0;
}

// Staging of V and U for the allocation-free scatter2kernel: one copy per
// GPU, held until the process ends.  Host code never passes this name to a
// kernel or a copy (it is a host-side stand-in there): it uses the device
// address from hipGetSymbolAddress, fetched once per device.
constexpr int S2K_MAX_KNT = 16384;
static __device__ __align__(256) float scratch[2 * S2K_MAX_KNT];
namespace {
constexpr int S2K_MAX_DEV = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[S2K_MAX_DEV] = {};
}  // namespace

extern "C" {

int fwk_scatter2kernel(long npt, int knt, int nt1, int nt2, float dt, float M0, const float* V, const float* U,
                       const float* Sa, const float* Sb, float* Kap, float* Kaq, float* Kbp, float* Kbq, void* stream) {
  if (npt <= 0 || knt <= 0 || nt1 < 1 || nt2 > knt || nt1 >= nt2 || !V || !U || !Sa || !Sb || !Kap || !Kaq || !Kbp ||
      !Kbq)
    return kfail(FW_ERR_ARG, "fwk_scatter2kernel: bad argument (1 <= nt1 < nt2 <= knt)");
  double V0, U0;
  s2k_norms(nt1, nt2, dt, M0, V, U, &V0, &U0);
  hipStream_t st = (hipStream_t)stream;
  float* dVU = nullptr;
  KCHK(hipMallocAsync((void**)&dVU, sizeof(float) * 2 * knt, st));
  KCHK(hipMemcpyAsync(dVU, V, sizeof(float) * knt, hipMemcpyHostToDevice, st));
  KCHK(hipMemcpyAsync(dVU + knt, U, sizeof(float) * knt, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_kernel, dim3(nblocks(npt, 256)), dim3(256), 0, st, npt, nt1, nt2, dt, (float)V0, (float)U0, dVU,
                     dVU + knt, Sa, Sb, Kap, Kaq, Kbp, Kbq);
  KCHK(hipGetLastError());
  KCHK(hipFreeAsync(dVU, st));
  KCHK(hipStreamSynchronize(st));
  return FW_OK;
}

int fwk_scatter2kernel_noalloc(long npt, int knt, int nt1, int nt2, float dt, float M0, const float* V,
                               const float* U, const float* Sa, const float* Sb, float* Kap, float* Kaq, float* Kbp,
                               float* Kbq, void* stream) {
  if (npt <= 0 || knt <= 0 || nt1 < 1 || nt2 > knt || nt1 >= nt2 || !V || !U || !Sa || !Sb || !Kap || !Kaq || !Kbp ||
      !Kbq)
    return kfail(FW_ERR_ARG, "fwk_scatter2kernel_noalloc: bad argument (1 <= nt1 < nt2 <= knt)");
  if (knt > S2K_MAX_KNT)
    return kfail(FW_ERR_ARG, "fwk_scatter2kernel_noalloc: knt = " + std::to_string(knt) +
                                 " exceeds the staging size of 16384 samples (use fwk_scatter2kernel)");
  int dev = -1;
  KCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= S2K_MAX_DEV) return kfail(FW_ERR_ARG, "fwk_scatter2kernel_noalloc: device index above 63");
  float* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return kfail(FW_ERR_HIP, std::string("fwk_scatter2kernel_noalloc: hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
      g_scratch_ptr[dev] = (float*)p;
    }
    ptr = g_scratch_ptr[dev];
  }
  double V0, U0;
  s2k_norms(nt1, nt2, dt, M0, V, U, &V0, &U0);
  hipStream_t st = (hipStream_t)stream;
  // the copies overwrite every staged value the kernel reads (V, U: knt each)
  KCHK(hipMemcpyAsync(ptr, V, sizeof(float) * knt, hipMemcpyHostToDevice, st));
  KCHK(hipMemcpyAsync(ptr + knt, U, sizeof(float) * knt, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_kernel, dim3(nblocks(npt, 256)), dim3(256), 0, st, npt, nt1, nt2, dt, (float)V0, (float)U0, ptr,
                     ptr + knt, Sa, Sb, Kap, Kaq, Kbp, Kbq);
  KCHK(hipGetLastError());
  KCHK(hipStreamSynchronize(st));
  return FW_OK;
}

int fwk_rotate(long n, float* X, float* Y, float* Z, const float* rot, void* stream) {
  if (n <= 0 || !X || !Y || !Z || !rot) return kfail(FW_ERR_ARG, "fwk_rotate: bad argument");
  Rot R;
  for (int q = 0; q < 9; ++q) R.r[q] = rot[q];
  hipLaunchKernelGGL(k_rotate, dim3(nblocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n, X, Y, Z, R);
  KCHK(hipGetLastError());
  return FW_OK;
}

}  // extern "C"
