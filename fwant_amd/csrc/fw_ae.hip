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

__global__ void k_conv_fwd(ConvDims d, const float* __restrict__ x, const float* __restrict__ w,
                           const float* __restrict__ b, float* __restrict__ y) {
  const long o3 = (long)d.Do * d.Do * d.Do;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)d.N * d.Co * o3) return;
  const int ox = (int)(q % d.Do), oy = (int)((q / d.Do) % d.Do), oz = (int)((q / ((long)d.Do * d.Do)) % d.Do);
  const int co = (int)((q / o3) % d.Co);
  const long n = q / (o3 * d.Co);
  const long i3 = (long)d.Di * d.Di * d.Di;
  const int K3 = d.K * d.K * d.K;
  float acc = b ? b[co] : 0.f;
  for (int ci = 0; ci < d.Ci; ++ci) {
    const float* xc = x + (n * d.Ci + ci) * i3;
    const float* wc = w + ((long)co * d.Ci + ci) * K3;
    for (int kz = 0; kz < d.K; ++kz) {
      const int iz = oz * d.S - d.P + kz;
      if (iz < 0 || iz >= d.Di) continue;
      for (int ky = 0; ky < d.K; ++ky) {
        const int iy = oy * d.S - d.P + ky;
        if (iy < 0 || iy >= d.Di) continue;
        for (int kx = 0; kx < d.K; ++kx) {
          const int ix = ox * d.S - d.P + kx;
          if (ix < 0 || ix >= d.Di) continue;
          acc += wc[(kz * d.K + ky) * d.K + kx] * xc[((long)iz * d.Di + iy) * d.Di + ix];
        }
      }
    }
  }
  y[q] = acc;
}

__global__ void k_conv_bwd_data(ConvDims d, const float* __restrict__ dy, const float* __restrict__ w,
                                const float* __restrict__ b2, float* __restrict__ dx) {
  const long i3 = (long)d.Di * d.Di * d.Di;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)d.N * d.Ci * i3) return;
  const int ix = (int)(q % d.Di), iy = (int)((q / d.Di) % d.Di), iz = (int)((q / ((long)d.Di * d.Di)) % d.Di);
  const int ci = (int)((q / i3) % d.Ci);
  const long n = q / (i3 * d.Ci);
  const long o3 = (long)d.Do * d.Do * d.Do;
  const int K3 = d.K * d.K * d.K;
  float acc = b2 ? b2[ci] : 0.f;  // bias of a transposed convolution
  for (int co = 0; co < d.Co; ++co) {
    const float* dyc = dy + (n * d.Co + co) * o3;
    const float* wc = w + ((long)co * d.Ci + ci) * K3;
    for (int kz = 0; kz < d.K; ++kz) {
      const int tz = iz + d.P - kz;
      if (tz < 0 || tz % d.S) continue;
      const int oz = tz / d.S;
      if (oz >= d.Do) continue;
      for (int ky = 0; ky < d.K; ++ky) {
        const int ty = iy + d.P - ky;
        if (ty < 0 || ty % d.S) continue;
        const int oy = ty / d.S;
        if (oy >= d.Do) continue;
        for (int kx = 0; kx < d.K; ++kx) {
          const int tx = ix + d.P - kx;
          if (tx < 0 || tx % d.S) continue;
          const int ox = tx / d.S;
          if (ox >= d.Do) continue;
          acc += wc[(kz * d.K + ky) * d.K + kx] * dyc[((long)oz * d.Do + oy) * d.Do + ox];
        }
      }
    }
  }
  dx[q] = acc;
}

// pass 1: thread (weight element or bias, chunk) sums its chunk of (n, o)
__global__ void k_conv_bwd_w_part(ConvDims d, const float* __restrict__ x, const float* __restrict__ dy, int nchunk,
                                  float* __restrict__ part) {
  const int K3 = d.K * d.K * d.K;
  const long nw = (long)d.Co * d.Ci * K3 + d.Co;  // weights, then biases
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nw * nchunk) return;
  const long e = q % nw;
  const int c = (int)(q / nw);
  const long o3 = (long)d.Do * d.Do * d.Do, i3 = (long)d.Di * d.Di * d.Di;
  const long tot = (long)d.N * o3;
  const long lo = tot * c / nchunk, hi = tot * (c + 1) / nchunk;
  float acc = 0.f;
  if (e >= (long)d.Co * d.Ci * K3) {  // bias
    const int co = (int)(e - (long)d.Co * d.Ci * K3);
    for (long r = lo; r < hi; ++r) {
      const long n = r / o3, o = r % o3;
      acc += dy[(n * d.Co + co) * o3 + o];
    }
  } else {
    const int t = (int)(e % K3), ci = (int)((e / K3) % d.Ci), co = (int)(e / ((long)K3 * d.Ci));
    const int kx = t % d.K, ky = (t / d.K) % d.K, kz = t / (d.K * d.K);
    for (long r = lo; r < hi; ++r) {
      const long n = r / o3, o = r % o3;
      const int ox = (int)(o % d.Do), oy = (int)((o / d.Do) % d.Do), oz = (int)(o / ((long)d.Do * d.Do));
      const int iz = oz * d.S - d.P + kz, iy = oy * d.S - d.P + ky, ix = ox * d.S - d.P + kx;
      if (iz < 0 || iz >= d.Di || iy < 0 || iy >= d.Di || ix < 0 || ix >= d.Di) continue;
      acc += dy[(n * d.Co + co) * o3 + o] * x[(n * d.Ci + ci) * i3 + ((long)iz * d.Di + iy) * d.Di + ix];
    }
  }
  part[q] = acc;
}
__global__ void k_sum_parts(long nw, int nchunk, const float* __restrict__ part, float* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nw) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(long)c * nw + e];
  out[e] = s;
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

bool dims_ok(const ConvDims& d) {
  if (d.N <= 0 || d.Ci <= 0 || d.Co <= 0 || d.Di <= 0 || d.Do <= 0 || d.K <= 0 || d.S <= 0 || d.P < 0) return false;
  return (d.Di + 2 * d.P - d.K) / d.S + 1 == d.Do;
}

}  // namespace

extern "C" {

int fwae_conv_fwd(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* x, const float* w,
                  const float* b, float* y, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !x || !w || !y) return fwk_fail(FW_ERR_ARG, "fwae_conv_fwd: bad dimensions or argument");
  hipLaunchKernelGGL(k_conv_fwd, dim3(nbl((long)N * Co * Do * Do * Do)), dim3(256), 0, (hipStream_t)stream, d, x, w,
                     b, y);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_fwd: launch failed");
}

int fwae_conv_bwd_data(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* dy, const float* w,
                       const float* b2, float* dx, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !dy || !w || !dx) return fwk_fail(FW_ERR_ARG, "fwae_conv_bwd_data: bad dimensions or argument");
  hipLaunchKernelGGL(k_conv_bwd_data, dim3(nbl((long)N * Ci * Di * Di * Di)), dim3(256), 0, (hipStream_t)stream, d,
                     dy, w, b2, dx);
  return hipGetLastError() == hipSuccess ? FW_OK : fwk_fail(FW_ERR_HIP, "fwae_conv_bwd_data: launch failed");
}

// dw [Co][Ci][K^3] followed by db [Co] in dwb; part: scratch of nchunk*(Co*Ci*K^3+Co) floats
int fwae_conv_bwd_weight(int N, int Ci, int Co, int Di, int Do, int K, int S, int P, const float* x, const float* dy,
                         int nchunk, float* part, float* dwb, void* stream) {
  const ConvDims d{N, Ci, Co, Di, Do, K, S, P};
  if (!dims_ok(d) || !x || !dy || !part || !dwb || nchunk <= 0)
    return fwk_fail(FW_ERR_ARG, "fwae_conv_bwd_weight: bad dimensions or argument");
  const long nw = (long)Co * Ci * K * K * K + Co;
  hipLaunchKernelGGL(k_conv_bwd_w_part, dim3(nbl(nw * nchunk)), dim3(256), 0, (hipStream_t)stream, d, x, dy, nchunk,
                     part);
  hipLaunchKernelGGL(k_sum_parts, dim3(nbl(nw)), dim3(256), 0, (hipStream_t)stream, nw, nchunk, part, dwb);
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
