// libfwant: the inversion steps after the kernels -- assembly of kernel
// values onto inversion blocks (code/srcF.kernel/SI_ker2assm.F90:300-320)
// and the damped least-squares solve by LSQR (inv.LSQR/mod_LSQR.f90:
// 142-292, Paige & Saunders) -- in FP64 on the GPU.
//
// LSQR keeps A in CSR (the reference's row storage ra/ja/na) and A^T in CSR
// built once on the host; every product is one wavefront per row with a
// shuffle reduction, every dot product a fixed-order two-pass reduction, so
// a solve is deterministic.  The scalar recurrences (norms, the plane
// rotation, the stopping tests) run on the device in one thread, in slsqr's
// order with one rounding per operation, so that the host enqueues
// iterations without waiting for them: it reads the device's stop flag once
// per LSQR_POLL iterations, and every kernel of an iteration enqueued after
// the stop returns at once.
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
constexpr int LSQR_POLL = 16;

// scalar state of a solve (slsqr's variables), in device memory
struct LsqrState {
  double alpha, beta, rhobar, phibar, rho, c, phi, theta, res2;
  long itct;
  int done;
};

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
                       const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
                       const LsqrState* __restrict__ S) {
  const long r = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (r >= nrow || S->done) return;
  double s = 0.0;
  for (long q = ptr[r] + lane; q < ptr[r + 1]; q += WAVE) s += val[q] * x[col[q]];
  for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) y[r] = s;
}

// partial dot products (fixed grid), then one block sums the partials in order
__global__ void k_dot_part(long n, const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ part, const LsqrState* __restrict__ S) {
  if (S->done) return;
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
// the sum of the partials, in a fixed order, by one block
__device__ double dot_final(const double* __restrict__ part, int np) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  return sh[0];
}
// lnrliz's norm: S->beta (which 0) or S->alpha (which 1) = sqrt(a . a)
__global__ void k_norm(const double* __restrict__ part, int np, LsqrState* __restrict__ S, int which) {
  if (S->done) return;
  const double s = dot_final(part, np);
  if (threadIdx.x == 0) {
    if (which) S->alpha = sqrt(s);
    else S->beta = sqrt(s);
  }
}
// lnrliz's division, unless the norm is (numerically) zero
__global__ void k_scale(long n, double* __restrict__ a, const LsqrState* __restrict__ S, int which) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || S->done) return;
  const double v0 = which ? S->alpha : S->beta;
  if (fabs(v0) > 1.0e-20) a[i] = a[i] / v0;
}
// a = b - a*s (lupdat '-'), s = S->alpha (which 1) or S->beta (which 0)
__global__ void k_update(long n, double* __restrict__ a, const double* __restrict__ b, const LsqrState* __restrict__ S,
                         int which) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || S->done) return;
  a[i] = b[i] - a[i] * (which ? S->alpha : S->beta);
}
__global__ void k_lsqr_start(LsqrState* __restrict__ S) {
  S->phibar = S->beta;
  S->rhobar = S->alpha;
  S->res2 = 0.0;
  S->itct = 0;
}
// slsqr's plane rotation (mod_LSQR.f90:224-233); rho = 0 ends the solve
__global__ void k_lsqr_rotate(LsqrState* __restrict__ S) {
#pragma clang fp contract(off)
  if (S->done) return;
  const double rhobar = S->rhobar, beta = S->beta, alpha = S->alpha;
  const double rho = sqrt(rhobar * rhobar + beta * beta);
  if (rho == 0.0) {
    S->done = 1;
    return;
  }
  const double c = rhobar / rho, s = beta / rho;
  S->rho = rho;
  S->c = c;
  S->theta = s * alpha;
  S->rhobar = -c * alpha;
  S->phi = c * S->phibar;
  S->phibar = s * S->phibar;
}
// the stopping tests after the solution update (mod_LSQR.f90:252-268),
// from the partials of x . x
__global__ void k_lsqr_test(const double* __restrict__ part, int np, LsqrState* __restrict__ S, double anorm,
                            double atol, double btol, long maxit) {
#pragma clang fp contract(off)
  if (S->done) return;
  const double xx = dot_final(part, np);
  if (threadIdx.x != 0) return;
  const double test1 = S->phibar / anorm / sqrt(xx);
  const double test2 = S->alpha * fabs(S->c) / anorm;
  S->itct += 1;
  if (test1 <= atol || test2 <= btol || S->itct == maxit) {
    S->res2 = S->phibar;
    S->done = 1;
  }
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
                     double* __restrict__ sig, const LsqrState* __restrict__ S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || S->done) return;
  const double t = w[i] / S->rho;
  x[i] = x[i] + S->phi * t;
  w[i] = v[i] - S->theta * t;
  sig[i] = sig[i] + t * t;
}

unsigned nb(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct fwl_s {
  long nrow = 0, ncol = 0, nnz = 0;
  long *ptr = nullptr, *col = nullptr, *tptr = nullptr, *tcol = nullptr;
  double *val = nullptr, *tval = nullptr;
  double *b = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *x = nullptr, *q = nullptr, *sig = nullptr;
  double* part = nullptr;
  LsqrState* st = nullptr;
  double anorm = 0.0;
  hipStream_t stream = nullptr;
};

// lnrliz: the norm into S->beta (which 0) or S->alpha (which 1), then the
// division unless the norm is (numerically) zero
static void normalize(fwl_handle h, long n, double* a, int which) {
  hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, h->stream, n, a, a, h->part, h->st);
  hipLaunchKernelGGL(k_norm, dim3(1), dim3(256), 0, h->stream, h->part, RED_BLOCKS, h->st, which);
  hipLaunchKernelGGL(k_scale, dim3(nb(n, 256)), dim3(256), 0, h->stream, n, a, h->st, which);
}
static void spmv(fwl_handle h, long nrow, const long* ptr, const long* col, const double* val, const double* x,
                 double* y) {
  hipLaunchKernelGGL(k_spmv, dim3(nb(nrow, 256 / WAVE)), dim3(256), 0, h->stream, nrow, ptr, col, val, x, y, h->st);
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
                  (void*)h->part, (void*)h->st})
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
      hipMalloc(&h->st, sizeof(LsqrState)))
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
  LsqrState s0 = {};
  ICHK(hipMemcpyAsync(h->st, &s0, sizeof(LsqrState), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_copy, dim3(nb(m, 256)), dim3(256), 0, st, m, h->u, h->b);
  normalize(h, m, h->u, 0);                           // beta u = b
  spmv(h, n, h->tptr, h->tcol, h->tval, h->u, h->v);  // v = A^T u
  normalize(h, n, h->v, 1);                           // alpha v = A^T u
  hipLaunchKernelGGL(k_copy, dim3(nb(n, 256)), dim3(256), 0, st, n, h->w, h->v);
  hipLaunchKernelGGL(k_fill0, dim3(nb(n, 256)), dim3(256), 0, st, n, h->x);
  hipLaunchKernelGGL(k_fill0, dim3(nb(n, 256)), dim3(256), 0, st, n, h->sig);
  hipLaunchKernelGGL(k_lsqr_start, dim3(1), dim3(1), 0, st, h->st);
  // iterations are enqueued LSQR_POLL at a time; the device sets done at
  // the stopping test (at the latest at itct == maxit)
  LsqrState sh = {};
  for (long it = 0; it < maxit; it += LSQR_POLL) {
    for (long j = it; j < std::min(maxit, it + LSQR_POLL); ++j) {
      spmv(h, m, h->ptr, h->col, h->val, h->v, h->q);  // q = A v
      hipLaunchKernelGGL(k_update, dim3(nb(m, 256)), dim3(256), 0, st, m, h->u, h->q, h->st, 1);
      normalize(h, m, h->u, 0);                           // beta u = A v - alpha u
      spmv(h, n, h->tptr, h->tcol, h->tval, h->u, h->q);  // q = A^T u
      hipLaunchKernelGGL(k_update, dim3(nb(n, 256)), dim3(256), 0, st, n, h->v, h->q, h->st, 0);
      normalize(h, n, h->v, 1);                           // alpha v = A^T u - beta v
      hipLaunchKernelGGL(k_lsqr_rotate, dim3(1), dim3(1), 0, st, h->st);
      hipLaunchKernelGGL(k_xw, dim3(nb(n, 256)), dim3(256), 0, st, n, h->x, h->w, h->v, h->sig, h->st);
      hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, st, n, h->x, h->x, h->part, h->st);
      hipLaunchKernelGGL(k_lsqr_test, dim3(1), dim3(256), 0, st, h->part, RED_BLOCKS, h->st, h->anorm, atol, btol,
                         maxit);
    }
    ICHK(hipGetLastError());
    ICHK(hipMemcpyAsync(&sh, h->st, sizeof(LsqrState), hipMemcpyDeviceToHost, st));
    ICHK(hipStreamSynchronize(st));
    if (sh.done) break;
  }
  const long itct = sh.itct;
  const double res2 = sh.res2;
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
