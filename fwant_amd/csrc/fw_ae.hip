// libfwant: kernels of the SGT block autoencoder (training and inference).
//
// The network works on 16^3 blocks of the 6 SGT components, scaled per block
// and component to [-1, 1].  Every layer is a cubic 3-D convolution (k, stride
// s, zero padding p) or its transpose; the three primitives below cover both:
//   conv forward        y[n,co,o] = b[co] + sum_{ci,t} w[co,ci,t] x[n,ci,o*s-p+t]
//   conv backward data  dx[n,ci,i] = sum_{co,t: (i+p-t) = o*s} w[co,ci,t] dy[n,co,o]
//   conv backward wts   dw[co,ci,t] = sum_{n,o} dy[n,co,o] x[n,ci,o*s-p+t], db = sum dy
// (a transposed convolution runs backward data as its forward, forward as
// its backward data).  Layout NCDHW, FP32 weights and accumulation;
// activations are rounded to BF16 at every layer output.  Weight gradients
// are reduced over fixed chunks in two passes, so training is deterministic.
#include "fw_common.h"
#include "../../include/fwant.h"

#include <string>

int fwk_fail(int code, const std::string& m);  // error of the fwk_* family (fw_kernel.hip)

namespace {

struct ConvDims {
  int N, Ci, Co, Di, Do, K, S, P;
};

__device__ __forceinline__ float bf16_round(float x) { return fw::bf16_val(fw::bf16_rne(x)); }

// ---------------------------------------------------------------------------
// The three convolution primitives as implicit GEMMs on the matrix cores
// (v_mfma_f32_16x16x4f32: FP32 operands and accumulation, so the products of
// the FP32 weights are exact as in a scalar loop; only the summation order
// differs).  One wave computes a 16 (rows m) x 16*NSUB (columns n) tile of
// D = A B over the reduction index k, 4 values of k per MFMA.  Lane l
// supplies A[m0 + l%16][k0 + l/16] and, per 16-column subtile j,
// B[k0 + l/16][n0 + 16 j + l%16]; it receives D[m0 + 4(l/16) + r][n0 + 16 j +
// l%16], r = 0..3.  B is gathered from the activations (im2col on the fly),
// with the per-k offsets in an LDS table.
//
//   forward      m = co, n = (batch, output voxel), k = (ci, tap):
//                A = w[co][ci][t], B = x[.., ci, o*S - P + t]
//   backward     m = ci, n = (batch, input voxel), k = (co, tap), one launch
//   data         slice per residue class of the input voxel mod S (blockIdx.z),
//                so that only the K/S taps per axis that reach an output are
//                visited: A = w[co][ci][t], B = dy[.., co, (i + P - t)/S]
//   backward     m = co, n = (ci, tap) plus one column of ones (the bias
//   weights      gradient), k = (batch, output voxel) in chunks (blockIdx.z)
//                whose partial tiles are summed in a fixed order afterwards
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int MT = 16;        // rows of an MFMA tile
constexpr int WPB = 4;        // waves per block
constexpr int TABMAX = 1024;  // reduction entries of the LDS table per pass
constexpr int UNR = 4;        // MFMA steps whose loads are issued together

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int pack3(int a, int b, int c) {
  return (a & 0xff) | ((b & 0xff) << 8) | ((c & 0xff) << 16);
}
__device__ __forceinline__ int unpack(int v, int q) { return (int)(signed char)((v >> (8 * q)) & 0xff); }

// Geometry of one launch slice of the forward / backward-data GEMM.
struct Slice {
  int M, Kd;          // rows, reduction length
  long Nn;            // columns
  int Ds;             // edge of the gathered volume (Di forward, Do backward)
  int m_str;          // A row stride in w (forward Ci K^3, backward K^3)
  int tz0, ty0, tx0;  // backward: first tap per axis (residue r), taps nt per axis
  int ntz, nty, ntx;
  int ez, ey, ex;     // backward: o = i' + e - t'
  int mz, my, mx;     // backward: residue class of the input voxel
  int niz, niy, nix;  // backward: input voxels of the class per axis
};
__device__ Slice make_slice(const ConvDims& d, bool bwd, int cls) {
  Slice s{};
  const int K3 = d.K * d.K * d.K;
  if (!bwd) {
    s.M = d.Co; s.Kd = d.Ci * K3; s.Nn = (long)d.N * d.Do * d.Do * d.Do; s.Ds = d.Di; s.m_str = d.Ci * K3;
    return s;
  }
  const int m3[3] = {cls / (d.S * d.S), (cls / d.S) % d.S, cls % d.S};
  int t0[3], nt[3], e[3], ni[3];
  for (int a = 0; a < 3; ++a) {
    t0[a] = (m3[a] + d.P) % d.S;
    nt[a] = t0[a] < d.K ? (d.K - t0[a] + d.S - 1) / d.S : 0;
    e[a] = (m3[a] + d.P - t0[a]) / d.S;
    ni[a] = m3[a] < d.Di ? (d.Di - m3[a] + d.S - 1) / d.S : 0;
  }
  s.M = d.Ci; s.Kd = d.Co * nt[0] * nt[1] * nt[2]; s.Ds = d.Do; s.m_str = K3;
  s.Nn = (long)d.N * ni[0] * ni[1] * ni[2];
  s.tz0 = t0[0]; s.ty0 = t0[1]; s.tx0 = t0[2]; s.ntz = nt[0]; s.nty = nt[1]; s.ntx = nt[2];
  s.ez = e[0]; s.ey = e[1]; s.ex = e[2]; s.mz = m3[0]; s.my = m3[1]; s.mx = m3[2];
  s.niz = ni[0]; s.niy = ni[1]; s.nix = ni[2];
  return s;
}

// forward (BWD false) and backward data (BWD true); bias may be null.
// KSPLIT (few columns, long reductions): the block's waves share one column
// tile and take every WPB-th group of UNR steps; their tiles are added in
// wave order through LDS
// act: the stored output is leaky ReLU (slope) of the sum, rounded to BF16
// (the network's layer output, fused into the store)
template <bool BWD, int NSUB, bool KSPLIT>
__global__ __launch_bounds__(64 * WPB) void k_conv_mfma(ConvDims d, const float* __restrict__ src,
                                                        const float* __restrict__ w, const float* __restrict__ bias,
                                                        float* __restrict__ dst, int act, float slope) {
  __shared__ int4 tab[TABMAX];  // per k: gather offset, A offset, packed tap offsets
  __shared__ float red[KSPLIT ? WPB - 1 : 1][NSUB][4][64];
  const Slice s = make_slice(d, BWD, blockIdx.z);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m0 = blockIdx.y * MT;
  const long n0 = (KSPLIT ? (long)blockIdx.x : (long)blockIdx.x * WPB + wv) * (16 * NSUB);
  const int kbeg = KSPLIT ? 4 * UNR * wv : 0, kstep = KSPLIT ? 4 * UNR * WPB : 4 * UNR;
  const long Ds3 = (long)s.Ds * s.Ds * s.Ds;
  const int K3 = d.K * d.K * d.K;
  // this lane's columns: gather base, voxel base per axis, output index
  long nbase[NSUB], dsti[NSUB];
  int bz[NSUB], by[NSUB], bx[NSUB];
  bool nok[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    const long n = n0 + 16 * j + (lane & 15);
    nok[j] = n < s.Nn;
    const long nn = nok[j] ? n : 0;
    if (!BWD) {
      const long o3 = (long)d.Do * d.Do * d.Do;
      const long nb = nn / o3, o = nn % o3;
      const int ox = (int)(o % d.Do), oy = (int)((o / d.Do) % d.Do), oz = (int)(o / ((long)d.Do * d.Do));
      bz[j] = oz * d.S - d.P; by[j] = oy * d.S - d.P; bx[j] = ox * d.S - d.P;
      nbase[j] = nb * d.Ci * Ds3 + ((long)bz[j] * s.Ds + by[j]) * s.Ds + bx[j];
      dsti[j] = nb * d.Co * o3 + o;
    } else {
      const long c3 = (long)s.niz * s.niy * s.nix;
      const long nb = nn / c3, c = nn % c3;
      const int jx = (int)(c % s.nix), jy = (int)((c / s.nix) % s.niy), jz = (int)(c / ((long)s.nix * s.niy));
      bz[j] = jz + s.ez; by[j] = jy + s.ey; bx[j] = jx + s.ex;
      nbase[j] = nb * d.Co * Ds3 + ((long)bz[j] * s.Ds + by[j]) * s.Ds + bx[j];
      const int iz = d.S * jz + s.mz, iy = d.S * jy + s.my, ix = d.S * jx + s.mx;
      dsti[j] = nb * d.Ci * (long)d.Di * d.Di * d.Di + ((long)iz * d.Di + iy) * d.Di + ix;
    }
  }
  const int ma = m0 + (lane & 15);
  const float* __restrict__ wrow = w + (ma < s.M ? (long)ma * s.m_str : 0);
  const bool mok = ma < s.M;
  const bool live = n0 < s.Nn && m0 < s.M;
  v4f acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < s.Kd; kc += TABMAX) {
    const int kn = min(TABMAX, s.Kd - kc);
    __syncthreads();
    for (int q = threadIdx.x; q < kn; q += 64 * WPB) {
      const int k = kc + q;
      int4 e;
      if (!BWD) {
        const int ci = k / K3, t = k % K3, tz = t / (d.K * d.K), ty = (t / d.K) % d.K, tx = t % d.K;
        e.x = (int)(ci * Ds3 + ((long)tz * s.Ds + ty) * s.Ds + tx);
        e.y = k;
        e.z = pack3(tz, ty, tx);
      } else {
        const int nt = s.ntz * s.nty * s.ntx, co = k / nt, u = k % nt;
        const int uz = u / (s.nty * s.ntx), uy = (u / s.ntx) % s.nty, ux = u % s.ntx;
        const int tz = s.tz0 + d.S * uz, ty = s.ty0 + d.S * uy, tx = s.tx0 + d.S * ux;
        e.x = (int)(co * Ds3 - ((long)uz * s.Ds + uy) * s.Ds - ux);
        e.y = co * d.Ci * K3 + (tz * d.K + ty) * d.K + tx;
        e.z = pack3(-uz, -uy, -ux);
      }
      e.w = 0;
      tab[q] = e;
    }
    __syncthreads();
    if (!live) continue;
    // UNR steps of 4 k per iteration: all their loads are issued before the
    // first of their MFMAs
    for (int k0 = kbeg; k0 < kn; k0 += kstep) {
      float a[UNR], b[UNR][NSUB];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int q = k0 + 4 * u + (lane >> 4);
        const bool kok = q < kn;
        const int4 e = tab[kok ? q : 0];
        a[u] = (kok && mok) ? wrow[e.y] : 0.f;
        const int dz = unpack(e.z, 0), dy = unpack(e.z, 1), dx = unpack(e.z, 2);
#pragma unroll
        for (int j = 0; j < NSUB; ++j) {
          const int iz = bz[j] + dz, iy = by[j] + dy, ix = bx[j] + dx;
          const bool ok = kok && nok[j] && (unsigned)iz < (unsigned)s.Ds && (unsigned)iy < (unsigned)s.Ds &&
                          (unsigned)ix < (unsigned)s.Ds;
          const float v = src[ok ? nbase[j] + e.x : 0];
          b[u][j] = ok ? v : 0.f;
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int j = 0; j < NSUB; ++j) acc[j] = mfma4(a[u], b[u][j], acc[j]);
    }
  }
  if (KSPLIT) {
    if (wv > 0 && live)
#pragma unroll
      for (int j = 0; j < NSUB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wv - 1][j][r][lane] = acc[j][r];
    __syncthreads();
    if (wv > 0) return;
    if (live)
      for (int q = 0; q < WPB - 1; ++q)
#pragma unroll
        for (int j = 0; j < NSUB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[j][r] += red[q][j][r][lane];
  }
  if (!live) return;
  const long mstr = BWD ? (long)d.Di * d.Di * d.Di : (long)d.Do * d.Do * d.Do;
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    if (!nok[j]) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + 4 * (lane >> 4) + r;
      if (m < s.M) {
        const float v = acc[j][r] + (bias ? bias[m] : 0.f);
        dst[dsti[j] + m * mstr] = act ? bf16_round(v > 0.f ? v : slope * v) : v;
      }
    }
  }
}

// backward weights, pass 1: the partial tile of chunk blockIdx.z of the
// reduction over (batch, output voxel); columns n < Ci K^3 are weights
// (ci, tap), column Ci K^3 is the bias (a column of ones)
template <int NSUB>
__global__ __launch_bounds__(64 * WPB) void k_conv_wgrad_mfma(ConvDims d, const float* __restrict__ x,
                                                              const float* __restrict__ dy, int nchunk,
                                                              float* __restrict__ part) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int K3 = d.K * d.K * d.K, nwt = d.Ci * K3, ncol = nwt + 1;
  const long nw = (long)d.Co * nwt + d.Co;
  const int m0 = blockIdx.y * MT;
  const int n0 = (blockIdx.x * WPB + wv) * (16 * NSUB);
  if (n0 >= ncol || m0 >= d.Co) return;
  const long o3 = (long)d.Do * d.Do * d.Do, i3 = (long)d.Di * d.Di * d.Di;
  const long tot = (long)d.N * o3;
  const int c = blockIdx.z;
  const long lo = tot * c / nchunk, hi = tot * (c + 1) / nchunk;
  long xoff[NSUB];
  int tz[NSUB], ty[NSUB], tx[NSUB];
  bool wcol[NSUB], ones[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    const int n = n0 + 16 * j + (lane & 15);
    wcol[j] = n < nwt;
    ones[j] = n == nwt;
    const int nn = wcol[j] ? n : 0;
    const int ci = nn / K3, t = nn % K3;
    tz[j] = t / (d.K * d.K); ty[j] = (t / d.K) % d.K; tx[j] = t % d.K;
    xoff[j] = ci * i3 + ((long)tz[j] * d.Di + ty[j]) * d.Di + tx[j];
  }
  const int ma = m0 + (lane & 15);
  const bool mok = ma < d.Co;
  // this lane's reduction index k = lo + k0 + l/16 as (batch, oz, oy, ox),
  // advanced by 4 per step
  long k = lo + (lane >> 4);
  long nb = k / o3;
  int oz, oy, ox;
  {
    const long o = k % o3;
    ox = (int)(o % d.Do); oy = (int)((o / d.Do) % d.Do); oz = (int)(o / ((long)d.Do * d.Do));
  }
  v4f acc[NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
  // UNR steps of 4 k per iteration, loads first (as in k_conv_mfma)
  for (long k0 = lo; k0 < hi; k0 += 4 * UNR) {
    float a[UNR], b[UNR][NSUB];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const bool kok = k < hi;
      const long o = ((long)oz * d.Do + oy) * d.Do + ox;
      a[u] = (kok && mok) ? dy[(nb * d.Co + ma) * o3 + o] : 0.f;
      const int bz = oz * d.S - d.P, by = oy * d.S - d.P, bx = ox * d.S - d.P;
      const long xb = nb * d.Ci * i3 + ((long)bz * d.Di + by) * d.Di + bx;
#pragma unroll
      for (int j = 0; j < NSUB; ++j) {
        const int iz = bz + tz[j], iy = by + ty[j], ix = bx + tx[j];
        const bool ok = kok && wcol[j] && (unsigned)iz < (unsigned)d.Di && (unsigned)iy < (unsigned)d.Di &&
                        (unsigned)ix < (unsigned)d.Di;
        const float v = x[ok ? xb + xoff[j] : 0];
        b[u][j] = ok ? v : ((kok && ones[j]) ? 1.f : 0.f);
      }
      k += 4;
      ox += 4;
      while (ox >= d.Do) { ox -= d.Do; ++oy; }
      while (oy >= d.Do) { oy -= d.Do; ++oz; }
      while (oz >= d.Do) { oz -= d.Do; ++nb; }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int j = 0; j < NSUB; ++j) acc[j] = mfma4(a[u], b[u][j], acc[j]);
  }
  float* const out = part + (long)c * nw;
#pragma unroll
  for (int j = 0; j < NSUB; ++j) {
    const int n = n0 + 16 * j + (lane & 15);
    if (n >= ncol) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + 4 * (lane >> 4) + r;
      if (m < d.Co) out[n < nwt ? (long)m * nwt + n : (long)d.Co * nwt + m] = acc[j][r];
    }
  }
}
// pass 2: out[e] = sum over the chunks of part[c][e], in a fixed order: a
// block takes 32 consecutive e; thread group g (of 8) sums the chunks
// c = g mod 8 in order, then the 8 group sums are added in order
constexpr int SUM_E = 32, SUM_G = 8;
__global__ __launch_bounds__(SUM_E * SUM_G) void k_sum_parts(long nw, int nchunk, const float* __restrict__ part,
                                                             float* __restrict__ out) {
  __shared__ float sh[SUM_G][SUM_E];
  const int el = threadIdx.x % SUM_E, g = threadIdx.x / SUM_E;
  const long e = (long)blockIdx.x * SUM_E + el;
  float s = 0.f;
  if (e < nw)
    for (int c = g; c < nchunk; c += SUM_G) s += part[(long)c * nw + e];
  sh[g][el] = s;
  __syncthreads();
  if (g == 0 && e < nw) {
    float t = sh[0][el];
    for (int q = 1; q < SUM_G; ++q) t += sh[q][el];
    out[e] = t;
  }
}

// leaky ReLU (slope a; a = 1: identity) with BF16 rounding of the stored output
__global__ void k_act_fwd(long n, float a, float* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = y[i];
  y[i] = bf16_round(v > 0.f ? v : a * v);
}
__global__ void k_act_bwd(long n, float a, const float* __restrict__ y, float* __restrict__ dy) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!(y[i] > 0.f)) dy[i] *= a;
}
// mean squared error: per-sample sums (deterministic) and the gradient 2(y-t)/n
__global__ void k_mse(long per, int N, const float* __restrict__ y, const float* __restrict__ t,
                      float* __restrict__ grad, double* __restrict__ loss) {
  const int n = blockIdx.x;
  if (n >= N) return;
  __shared__ double sh[256];
  double s = 0.0;
  const float sc = 2.0f / (float)((long)N * per);
  for (long i = threadIdx.x; i < per; i += blockDim.x) {
    const long q = (long)n * per + i;
    const float e = y[q] - t[q];
    s += (double)e * e;
    grad[q] = sc * e;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[n] = sh[0];
}
// Adam (Kingma & Ba) with bias correction
__global__ void k_adam(long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// 16^3 blocks of nc components from snapshot records [rec][nc][c2][c1][c0]
// (blocks outside the snapshot are zero-padded), scaled per block and
// component by the largest magnitude; scale 0 for an all-zero block.
__global__ void k_blocks(const float* __restrict__ rec, int nc, int c0, int c1, int c2, int nb0, int nb1, int nb2,
                         float* __restrict__ out, float* __restrict__ scale) {
  const int blk = blockIdx.x, c = blockIdx.y, r = blockIdx.z;
  const int b0 = blk % nb0, b1 = (blk / nb0) % nb1, b2 = blk / (nb0 * nb1);
  const long npts = (long)c0 * c1 * c2;
  const float* src = rec + ((long)r * nc + c) * npts;
  const long nblk = (long)nb0 * nb1 * nb2;
  float* dst = out + (((long)r * nblk + blk) * nc + c) * 4096;
  float m = 0.f;
  for (int t = threadIdx.x; t < 4096; t += blockDim.x) {
    const int x = b0 * 16 + (t & 15), y = b1 * 16 + ((t >> 4) & 15), z = b2 * 16 + (t >> 8);
    const float v = (x < c0 && y < c1 && z < c2) ? src[x + (long)c0 * (y + (long)c1 * z)] : 0.f;
    dst[t] = v;
    m = fmaxf(m, fabsf(v));
  }
  __shared__ float sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  // a component whose maximum is below SGT_TINY is a zero block (1 / s would
  // overflow for subnormal maxima and turn its zeros into NaN)
  const float s = sh[0] >= fw::SGT_TINY ? sh[0] : 0.f;
  const float inv = s > 0.f ? 1.0f / s : 0.f;
  for (int t = threadIdx.x; t < 4096; t += blockDim.x) dst[t] *= inv;
  if (threadIdx.x == 0) scale[((long)r * nblk + blk) * nc + c] = s;
}

unsigned nbl(long n) { return (unsigned)((n + 255) / 256); }

// the tap offsets of a table entry are packed as signed bytes (K up to 64),
// the S^3 residue classes of the backward-data launch are its grid's z
// extent (S up to 32); gather offsets and A offsets are 32-bit
bool dims_ok(const ConvDims& d) {
  if (d.N <= 0 || d.Ci <= 0 || d.Co <= 0 || d.Di <= 0 || d.Do <= 0 || d.K <= 0 || d.S <= 0 || d.P < 0) return false;
  if (d.K > 64 || d.S > 32) return false;
  const long i3 = (long)d.Di * d.Di * d.Di, o3 = (long)d.Do * d.Do * d.Do, K3 = (long)d.K * d.K * d.K;
  if ((long)d.Ci * i3 >= (1L << 31) || (long)d.Co * o3 >= (1L << 31) || (long)d.Co * d.Ci * K3 >= (1L << 31))
    return false;
  return (d.Di + 2 * d.P - d.K) / d.S + 1 == d.Do;
}

// columns per wave: 64 when that still gives 1024 blocks, else 16; with
// fewer than 1024 blocks of 16-column waves, the block's waves split the
// reduction of one 16-column tile instead
template <bool BWD>
void launch_conv(const ConvDims& d, int nslice, long ncol, int M, const float* src, const float* w, const float* bias,
                 float* dst, hipStream_t st, int act = 0, float slope = 1.f) {
  const unsigned gy = (unsigned)((M + MT - 1) / MT);
  if ((ncol + 64 * WPB - 1) / (64 * WPB) >= 1024)
    hipLaunchKernelGGL((k_conv_mfma<BWD, 4, false>), dim3((unsigned)((ncol + 64 * WPB - 1) / (64 * WPB)), gy, nslice),
                       dim3(64 * WPB), 0, st, d, src, w, bias, dst, act, slope);
  else if ((ncol + 16 * WPB - 1) / (16 * WPB) >= 1024)
    hipLaunchKernelGGL((k_conv_mfma<BWD, 1, false>), dim3((unsigned)((ncol + 16 * WPB - 1) / (16 * WPB)), gy, nslice),
                       dim3(64 * WPB), 0, st, d, src, w, bias, dst, act, slope);
  else
    hipLaunchKernelGGL((k_conv_mfma<BWD, 1, true>), dim3((unsigned)((ncol + 15) / 16), gy, nslice), dim3(64 * WPB), 0,
                       st, d, src, w, bias, dst, act, slope);
}

}  // namespace

extern "C" {

int fwae_conv_fwd(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* x, const float* w,
                  const float* b, float* y, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !x || !w || !y) return fwk_fail(FW_ERR_ARG, "fwae_conv_fwd: bad dimensions or argument");
  launch_conv<false>(d, 1, (long)N * Do * Do * Do, Co, x, w, b, y, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_fwd: launch failed");
}

int fwae_conv_bwd_data(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* dy, const float* w,
                       const float* b2, float* dx, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !dy || !w || !dx) return fwk_fail(FW_ERR_ARG, "fwae_conv_bwd_data: bad dimensions or argument");
  // the largest residue class holds ceil(Di/S)^3 input voxels per batch member
  const long c = (Di + S - 1) / S;
  launch_conv<true>(d, S * S * S, (long)N * c * c * c, Ci, dy, w, b2, dx, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_bwd_data: launch failed");
}

// the same with the layer's activation fused into the store (fwae_act_fwd's
// leaky ReLU and BF16 rounding)
int fwae_conv_fwd_act(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* x, const float* w,
                      const float* b, float slope, float* y, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !x || !w || !y) return fwk_fail(FW_ERR_ARG, "fwae_conv_fwd_act: bad dimensions or argument");
  launch_conv<false>(d, 1, (long)N * Do * Do * Do, Co, x, w, b, y, (hipStream_t)stream, 1, slope);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_fwd_act: launch failed");
}
int fwae_conv_bwd_data_act(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* dy, const float* w,
                           const float* b2, float slope, float* dx, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !dy || !w || !dx)
    return fwk_fail(FW_ERR_ARG, "fwae_conv_bwd_data_act: bad dimensions or argument");
  const long c = (Di + S - 1) / S;
  launch_conv<true>(d, S * S * S, (long)N * c * c * c, Ci, dy, w, b2, dx, (hipStream_t)stream, 1, slope);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_bwd_data_act: launch failed");
}

// dw [Co][Ci][K^3] followed by db [Co] in dwb; part: scratch of nchunk*(Co*Ci*K^3+Co) floats
int fwae_conv_bwd_weight(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* x, const float* dy,
                         int nchunk, float* part, float* dwb, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !x || !dy || !part || !dwb || nchunk <= 0 || nchunk > 65535)
    return fwk_fail(FW_ERR_ARG, "fwae_conv_bwd_weight: bad dimensions or argument");
  const long nw = (long)Co * Ci * K * K * K + Co;
  const long ncol = (long)Ci * K * K * K + 1;
  const dim3 g((unsigned)((ncol + 64 * WPB - 1) / (64 * WPB)), (unsigned)((Co + MT - 1) / MT), (unsigned)nchunk);
  hipLaunchKernelGGL(k_conv_wgrad_mfma<4>, g, dim3(64 * WPB), 0, (hipStream_t)stream, d, x, dy, nchunk, part);
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((nw + SUM_E - 1) / SUM_E)), dim3(SUM_E * SUM_G), 0,
                     (hipStream_t)stream, nw, nchunk, part, dwb);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_bwd_weight: launch failed");
}

int fwae_act_fwd(long n, float slope, float* y, void* stream) {
  if (n <= 0 || !y) return fwk_fail(FW_ERR_ARG, "fwae_act_fwd: bad argument");
  hipLaunchKernelGGL(k_act_fwd, dim3(nbl(n)), dim3(256), 0, (hipStream_t)stream, n, slope, y);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_act_fwd: launch failed");
}

int fwae_act_bwd(long n, float slope, const float* y, float* dy, void* stream) {
  if (n <= 0 || !y || !dy) return fwk_fail(FW_ERR_ARG, "fwae_act_bwd: bad argument");
  hipLaunchKernelGGL(k_act_bwd, dim3(nbl(n)), dim3(256), 0, (hipStream_t)stream, n, slope, y, dy);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_act_bwd: launch failed");
}

int fwae_mse(int N, long per, const float* y, const float* t, float* grad, double* loss, void* stream) {
  if (N <= 0 || per <= 0 || !y || !t || !grad || !loss) return fwk_fail(FW_ERR_ARG, "fwae_mse: bad argument");
  hipLaunchKernelGGL(k_mse, dim3(N), dim3(256), 0, (hipStream_t)stream, per, N, y, t, grad, loss);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_mse: launch failed");
}

int fwae_adam(long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int step,
              void* stream) {
  if (n <= 0 || !p || !g || !m || !v || step < 1) return fwk_fail(FW_ERR_ARG, "fwae_adam: bad argument");
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nbl(n)), dim3(256), 0, (hipStream_t)stream, n, p, g, m, v, lr, b1, b2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_adam: launch failed");
}

int fwae_blocks(const float* rec, int nrec, int nc, int c0, int c1, int c2, float* out, float* scale, void* stream) {
  if (!rec || !out || !scale || nrec <= 0 || nc <= 0 || c0 <= 0 || c1 <= 0 || c2 <= 0 || nrec > 65535)
    return fwk_fail(FW_ERR_ARG, "fwae_blocks: bad argument");
  const int nb0 = (c0 + 15) / 16, nb1 = (c1 + 15) / 16, nb2 = (c2 + 15) / 16;
  hipLaunchKernelGGL(k_blocks, dim3(nb0 * nb1 * nb2, nc, nrec), dim3(256), 0, (hipStream_t)stream, rec, nc, c0, c1, c2,
                     nb0, nb1, nb2, out, scale);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_blocks: launch failed");
}

}  // extern "C"
