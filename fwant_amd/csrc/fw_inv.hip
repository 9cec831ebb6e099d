// libfwant: the inversion steps after the kernels -- assembly of kernel
// values onto inversion blocks (code/srcF.kernel/SI_ker2assm.F90:300-320)
// and the damped least-squares solve by LSQR (inv.LSQR/mod_LSQR.f90:
// 142-292, Paige & Saunders) -- in FP64 on the GPU.
//
// LSQR keeps A in CSR (the reference's row storage ra/ja/na) and A^T in CSR
// built once on the host; every product is one wavefront per row with a
// shuffle reduction, every dot product a fixed-order two-pass reduction, so
// a solve is deterministic.  The scalar recurrences run on the host in
// double, exactly as slsqr orders them.
#include "../../include/fwant.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

int fwk_fail(int code, const std::string& m);  // error of the fwk_* family (fw_kernel.hip)

namespace {

thread_local std::string g_ierr;
int ifail(int code, const std::string& m) {
  g_ierr = m;
  return code;
}
#define ICHK(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return ifail(FW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int WAVE = 64;
constexpr int RED_BLOCKS = 256;

// B[n] = sum over the pixels of block n of K(pixel) * volume(pixel)
__global__ void k_assemble(const float* __restrict__ K, const long* __restrict__ ptr, const long* __restrict__ pix,
                           const float* __restrict__ vol, long nblk, double* __restrict__ B) {
  const long n = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (n >= nblk) return;
  double s = 0.0;
  for (long q = ptr[n] + lane; q < ptr[n + 1]; q += WAVE) s += (double)K[pix[q]] * (double)vol[q];
  for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) B[n] = s;
}

// y = A x, one wavefront per row
__global__ void k_spmv(long nrow, const long* __restrict__ ptr, const long* __restrict__ col,
                       const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  const long r = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (r >= nrow) return;
  double s = 0.0;
  for (long q = ptr[r] + lane; q < ptr[r + 1]; q += WAVE) s += val[q] * x[col[q]];
  for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) y[r] = s;
}

// partial dot products (fixed grid), then one block sums the partials in order
__global__ void k_dot_part(long n, const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) s += a[i] * b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_dot_final(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}
// a = b - a*s (lupdat '-')
__global__ void k_update(long n, double* __restrict__ a, const double* __restrict__ b, double s) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i] - a[i] * s;
}
__global__ void k_scale(long n, double* __restrict__ a, double inv) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = a[i] / inv;
}
__global__ void k_copy(long n, double* __restrict__ a, const double* __restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i];
}
__global__ void k_fill0(long n, double* __restrict__ a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}
// x += phi*w/rho; w = v - theta*w/rho; sig += (w/rho)^2  (slsqr's solution update)
__global__ void k_xw(long n, double* __restrict__ x, double* __restrict__ w, const double* __restrict__ v,
                     double* __restrict__ sig, double rho, double phi, double theta) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = w[i] / rho;
  x[i] = x[i] + phi * t;
  w[i] = v[i] - theta * t;
  sig[i] = sig[i] + t * t;
}

unsigned nb(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct fwl_s {
  long nrow = 0, ncol = 0, nnz = 0;
  long *ptr = nullptr, *col = nullptr, *tptr = nullptr, *tcol = nullptr;
  double *val = nullptr, *tval = nullptr;
  double *b = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *x = nullptr, *q = nullptr, *sig = nullptr;
  double *part = nullptr, *scal = nullptr;
  double anorm = 0.0;
  hipStream_t stream = nullptr;
};

static double ddot(fwl_handle h, long n, const double* a, const double* b) {
  hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, h->stream, n, a, b, h->part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, h->stream, h->part, RED_BLOCKS, h->scal);
  double r = 0.0;
  hipMemcpyAsync(&r, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream);
  hipStreamSynchronize(h->stream);
  return r;
}
// lnrliz: norm, then divide unless the norm is (numerically) zero
static double normalize(fwl_handle h, long n, double* a) {
  const double v0 = std::sqrt(ddot(h, n, a, a));
  if (std::fabs(v0) > 1.0e-20) hipLaunchKernelGGL(k_scale, dim3(nb(n, 256)), dim3(256), 0, h->stream, n, a, v0);
  return v0;
}
static void spmv(fwl_handle h, long nrow, const long* ptr, const long* col, const double* val, const double* x,
                 double* y) {
  hipLaunchKernelGGL(k_spmv, dim3(nb(nrow, 256 / WAVE)), dim3(256), 0, h->stream, nrow, ptr, col, val, x, y);
}

extern "C" {

int fwk_assemble(long npt, const float* K, long nblk, const long* blk_ptr, const long* pix, const float* vol,
                 double* B, void* stream) {
  if (npt <= 0 || nblk <= 0 || !K || !blk_ptr || !pix || !vol || !B)
    return fwk_fail(FW_ERR_ARG, "fwk_assemble: bad argument");
  hipLaunchKernelGGL(k_assemble, dim3(nb(nblk, 256 / WAVE)), dim3(256), 0, (hipStream_t)stream, K, blk_ptr, pix, vol,
                     nblk, B);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fwk_fail(FW_ERR_HIP, std::string("fwk_assemble: ") + hipGetErrorString(e));
  return FW_OK;
}

const char* fwl_last_error(fwl_handle h) {
  (void)h;  // one message per host thread, as for fw_create
  return g_ierr.c_str();
}

int fwl_destroy(fwl_handle h) {
  if (!h) return FW_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void* p : {(void*)h->ptr, (void*)h->col, (void*)h->tptr, (void*)h->tcol, (void*)h->val, (void*)h->tval,
                  (void*)h->b, (void*)h->u, (void*)h->v, (void*)h->w, (void*)h->x, (void*)h->q, (void*)h->sig,
                  (void*)h->part, (void*)h->scal})
    hipFree(p);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return FW_OK;
}

int fwl_create(long nrow, long ncol, long nnz, const long* row_ptr, const long* col, const double* val,
               const double* rhs, fwl_handle* out) {
  if (!out || nrow <= 0 || ncol <= 0 || nnz < 0 || !row_ptr || !col || !val || !rhs)
    return ifail(FW_ERR_ARG, "fwl_create: bad argument");
  if (row_ptr[0] != 0 || row_ptr[nrow] != nnz) return ifail(FW_ERR_ARG, "fwl_create: row_ptr must run 0 .. nnz");
  for (long q = 0; q < nnz; ++q)
    if (col[q] < 0 || col[q] >= ncol) return ifail(FW_ERR_ARG, "fwl_create: column index out of range");
  // A^T in CSR
  std::vector<long> tptr(ncol + 1, 0), tcol(nnz);
  std::vector<double> tval(nnz);
  for (long q = 0; q < nnz; ++q) tptr[col[q] + 1]++;
  for (long c = 0; c < ncol; ++c) tptr[c + 1] += tptr[c];
  std::vector<long> fill(tptr.begin(), tptr.end() - 1);
  for (long r = 0; r < nrow; ++r)
    for (long q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
      const long d = fill[col[q]]++;
      tcol[d] = r;
      tval[d] = val[q];
    }
  double an = 0.0;
  for (long q = 0; q < nnz; ++q) an += val[q] * val[q];
  fwl_handle h = new fwl_s();
  h->nrow = nrow; h->ncol = ncol; h->nnz = nnz; h->anorm = std::sqrt(an);
  auto fail = [&](const char* m) {
    fwl_destroy(h);
    return ifail(FW_ERR_NOMEM, m);
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail("fwl_create: no HIP device");
  const size_t nn = (size_t)(nnz > 0 ? nnz : 1);
  if (hipMalloc(&h->ptr, sizeof(long) * (nrow + 1)) || hipMalloc(&h->col, sizeof(long) * nn) ||
      hipMalloc(&h->val, sizeof(double) * nn) || hipMalloc(&h->tptr, sizeof(long) * (ncol + 1)) ||
      hipMalloc(&h->tcol, sizeof(long) * nn) || hipMalloc(&h->tval, sizeof(double) * nn) ||
      hipMalloc(&h->b, sizeof(double) * nrow) || hipMalloc(&h->u, sizeof(double) * nrow) ||
      hipMalloc(&h->q, sizeof(double) * std::max(nrow, ncol)) || hipMalloc(&h->v, sizeof(double) * ncol) ||
      hipMalloc(&h->w, sizeof(double) * ncol) || hipMalloc(&h->x, sizeof(double) * ncol) ||
      hipMalloc(&h->sig, sizeof(double) * ncol) || hipMalloc(&h->part, sizeof(double) * RED_BLOCKS) ||
      hipMalloc(&h->scal, sizeof(double)))
    return fail("fwl_create: hipMalloc failed");
  if (hipMemcpy(h->ptr, row_ptr, sizeof(long) * (nrow + 1), hipMemcpyHostToDevice) ||
      (nnz && hipMemcpy(h->col, col, sizeof(long) * nnz, hipMemcpyHostToDevice)) ||
      (nnz && hipMemcpy(h->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice)) ||
      hipMemcpy(h->tptr, tptr.data(), sizeof(long) * (ncol + 1), hipMemcpyHostToDevice) ||
      (nnz && hipMemcpy(h->tcol, tcol.data(), sizeof(long) * nnz, hipMemcpyHostToDevice)) ||
      (nnz && hipMemcpy(h->tval, tval.data(), sizeof(double) * nnz, hipMemcpyHostToDevice)) ||
      hipMemcpy(h->b, rhs, sizeof(double) * nrow, hipMemcpyHostToDevice))
    return fail("fwl_create: hipMemcpy failed");
  *out = h;
  return FW_OK;
}

// slsqr (mod_LSQR.f90:142-277) and lgtsol (:278-291)
int fwl_solve(fwl_handle h, double atol, double btol, long maxit, double* x_out, double* err_out, long* itct_out,
              double* res_out) {
  if (!h || !x_out || maxit <= 0) return ifail(FW_ERR_ARG, "fwl_solve: bad argument");
  const long m = h->nrow, n = h->ncol;
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_copy, dim3(nb(m, 256)), dim3(256), 0, st, m, h->u, h->b);
  double beta = normalize(h, m, h->u);
  spmv(h, n, h->tptr, h->tcol, h->tval, h->u, h->v);  // v = A^T u
  double alpha = normalize(h, n, h->v);
  hipLaunchKernelGGL(k_copy, dim3(nb(n, 256)), dim3(256), 0, st, n, h->w, h->v);
  hipLaunchKernelGGL(k_fill0, dim3(nb(n, 256)), dim3(256), 0, st, n, h->x);
  hipLaunchKernelGGL(k_fill0, dim3(nb(n, 256)), dim3(256), 0, st, n, h->sig);
  double phibar = beta, rhobar = alpha, res2 = 0.0;
  long itct = 0;
  for (;;) {
    spmv(h, m, h->ptr, h->col, h->val, h->v, h->q);  // q = A v
    hipLaunchKernelGGL(k_update, dim3(nb(m, 256)), dim3(256), 0, st, m, h->u, h->q, alpha);
    beta = normalize(h, m, h->u);
    spmv(h, n, h->tptr, h->tcol, h->tval, h->u, h->q);  // q = A^T u
    hipLaunchKernelGGL(k_update, dim3(nb(n, 256)), dim3(256), 0, st, n, h->v, h->q, beta);
    alpha = normalize(h, n, h->v);
    const double rho = std::sqrt(rhobar * rhobar + beta * beta);
    if (rho == 0.0) break;
    const double c = rhobar / rho, s = beta / rho, theta = s * alpha;
    rhobar = -c * alpha;
    const double phi = c * phibar;
    phibar = s * phibar;
    hipLaunchKernelGGL(k_xw, dim3(nb(n, 256)), dim3(256), 0, st, n, h->x, h->w, h->v, h->sig, rho, phi, theta);
    const double rnorm = phibar;
    const double test1 = rnorm / h->anorm / std::sqrt(ddot(h, n, h->x, h->x));
    const double test2 = alpha * std::fabs(c) / h->anorm;
    itct++;
    if (test1 <= atol || test2 <= btol || itct == maxit) {
      res2 = phibar;
      break;
    }
  }
  ICHK(hipStreamSynchronize(st));
  ICHK(hipMemcpy(x_out, h->x, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (err_out) {
    std::vector<double> sig(n);
    ICHK(hipMemcpy(sig.data(), h->sig, sizeof(double) * n, hipMemcpyDeviceToHost));
    const double temp = res2 / (double)std::max(m - n, 1L);
    for (long i = 0; i < n; ++i) err_out[i] = std::sqrt(temp * sig[i]);
  }
  if (itct_out) *itct_out = itct;
  if (res_out) *res_out = res2;
  ICHK(hipGetLastError());
  return FW_OK;
}

}  // extern "C"
