// Resident vectors and the BLAS-1 pieces of the Lanczos iteration
// (wave_product / wave_norm / Gram_Schmidt_process of the reference,
// hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/block.f90:1092-1150,1251-1289).
// All reductions are FP64 with a fixed two-level order (per-block partials,
// then a host sum in block order), so results are run-to-run reproducible.
#include "tsc_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace tsc {

constexpr int kRedThreads = 256;

// partial[b * nv + j] = sum over the grid-stride slice of block b of Re(conj(V_j) . w)
__global__ __launch_bounds__(kRedThreads) void dot_many_kernel(const double2 *const *__restrict__ V, int nv,
                                                               const double2 *__restrict__ w, int64_t n,
                                                               double *__restrict__ partial) {
  __shared__ double sh[kRedThreads / 64][64];
  const int tid = threadIdx.x;
  for (int j0 = 0; j0 < nv; j0 += 64) {
    const int jn = min(64, nv - j0);
    double acc[8];
    for (int jb = 0; jb < jn; jb += 8) {
      const int cnt = min(8, jn - jb);
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] = 0.0;
      for (int64_t i = (int64_t)blockIdx.x * kRedThreads + tid; i < n; i += (int64_t)gridDim.x * kRedThreads) {
        const double2 x = w[i];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (q < cnt) {
            const double2 v = V[j0 + jb + q][i];
            acc[q] += v.x * x.x + v.y * x.y;
          }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (q >= cnt) break;
        double s = acc[q];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if ((tid & 63) == 0) sh[tid >> 6][q] = s;
        __syncthreads();
        if (tid == 0) {
          double t = 0.0;
          for (int wv = 0; wv < kRedThreads / 64; ++wv) t += sh[wv][q];
          partial[(int64_t)blockIdx.x * nv + j0 + jb + q] = t;
        }
        __syncthreads();
      }
    }
  }
}

// w = a*w + sum_j c_j V_j   (c on the device)
__global__ void combine_kernel(double2 *__restrict__ w, double a, const double2 *const *__restrict__ V,
                               const double *__restrict__ c, int nv, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 x = w[i];
    x.x *= a;
    x.y *= a;
    for (int j = 0; j < nv; ++j) {
      const double2 v = V[j][i];
      x.x += c[j] * v.x;
      x.y += c[j] * v.y;
    }
    w[i] = x;
  }
}

__global__ void scale_kernel(double2 *__restrict__ w, double a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 x = w[i];
    x.x *= a;
    x.y *= a;
    w[i] = x;
  }
}

// Counter-based complex Gaussian: splitmix64 of (seed, key, index) feeds a
// Box-Muller pair. Deterministic per element, independent of launch shape.
__device__ __host__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void gaussian_kernel(double2 *__restrict__ out, int64_t n, uint64_t seed, uint64_t key, double sigma) {
  const uint64_t h = splitmix64(seed ^ splitmix64(key));
  const double s = sigma * 0.70710678118654752440;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r1 = splitmix64(h + 2ull * (uint64_t)i);
    const uint64_t r2 = splitmix64(h + 2ull * (uint64_t)i + 1ull);
    const double u1 = ((double)(r1 >> 11) + 0.5) * 1.1102230246251565e-16;
    const double u2 = (double)(r2 >> 11) * 1.1102230246251565e-16;
    const double rad = sqrt(-2.0 * log(u1)) * s;
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    out[i] = make_double2(rad * cs, rad * sn);
  }
}

__global__ void axpby_kernel(double2 *__restrict__ y, double a, const double2 *__restrict__ x, double b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 xv = x[i];
    double2 yv = y[i];
    yv.x = a * xv.x + b * yv.x;
    yv.y = a * xv.y + b * yv.y;
    y[i] = yv;
  }
}

}  // namespace tsc

using namespace tsc;

// ---------------------------------------------------------------------------
// internal helpers used by the Lanczos driver
namespace tsc {

static int grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

// dots[j] = Re<V_j | w>, j < nv. V pointers are staged through ctx scratch.
int dots_re(tsc_ctx *ctx, const double2 *const *V_host, int nv, const double2 *w, int64_t n, double *dots) {
  if (nv == 0) return TSC_OK;
  TSC_REQUIRE(nv <= kMaxVecs, "too many vectors in one reduction");
  std::memcpy(ctx->h_ptrs, V_host, sizeof(double2 *) * nv);
  TSC_HIP_CHECK(hipMemcpyAsync(ctx->d_ptrs, ctx->h_ptrs, sizeof(double2 *) * nv, hipMemcpyHostToDevice, ctx->stream));
  const int nb = std::min<int>(kRedBlocks, grid_for(n, kRedThreads));
  hipLaunchKernelGGL(dot_many_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                     (const double2 *const *)ctx->d_ptrs, nv, w, n, ctx->d_red);
  TSC_HIP_CHECK(hipGetLastError());
  TSC_HIP_CHECK(hipMemcpyAsync(ctx->h_red, ctx->d_red, sizeof(double) * nb * nv, hipMemcpyDeviceToHost, ctx->stream));
  TSC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (int j = 0; j < nv; ++j) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += ctx->h_red[(size_t)b * nv + j];
    dots[j] = s;
  }
  return TSC_OK;
}

// w = a*w + sum_j c_j V_j
int combine(tsc_ctx *ctx, double2 *w, double a, const double2 *const *V_host, const double *c, int nv, int64_t n) {
  TSC_REQUIRE(nv <= kMaxVecs, "too many vectors in one combination");
  if (nv > 0) {
    // stage through pinned memory so the host arrays may go out of scope
    std::memcpy(ctx->h_ptrs, V_host, sizeof(double2 *) * nv);
    std::memcpy(ctx->h_coef, c, sizeof(double) * nv);
    TSC_HIP_CHECK(hipMemcpyAsync(ctx->d_ptrs2, ctx->h_ptrs, sizeof(double2 *) * nv, hipMemcpyHostToDevice, ctx->stream));
    TSC_HIP_CHECK(hipMemcpyAsync(ctx->d_coef, ctx->h_coef, sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream));
  }
  hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, w, a,
                     (const double2 *const *)ctx->d_ptrs2, (const double *)ctx->d_coef, nv, n);
  TSC_HIP_CHECK(hipGetLastError());
  TSC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return TSC_OK;
}

int fill_gaussian(tsc_ctx *ctx, double2 *out, int64_t n, uint64_t seed, uint64_t key, double sigma) {
  if (n <= 0) return TSC_OK;
  hipLaunchKernelGGL(gaussian_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, out, n, seed, key, sigma);
  TSC_HIP_CHECK(hipGetLastError());
  return TSC_OK;
}

int scale(tsc_ctx *ctx, double2 *w, double a, int64_t n) {
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, w, a, n);
  TSC_HIP_CHECK(hipGetLastError());
  return TSC_OK;
}

}  // namespace tsc

// ---------------------------------------------------------------------------
// C-ABI: vectors
extern "C" int tsc_vec_create(tsc_ctx *ctx, int64_t n, tsc_vec **out) {
  TSC_REQUIRE(ctx && out && n >= 0, "tsc_vec_create arguments");
  tsc_vec *v = new tsc_vec();
  v->ctx = ctx;
  v->n = n;
  if (n > 0) {
    hipError_t e = hipMalloc((void **)&v->d, sizeof(double2) * n);
    if (e != hipSuccess) {
      delete v;
      set_error(std::string("hipMalloc(vec) failed: ") + hipGetErrorString(e));
      return TSC_ENOMEM;
    }
    TSC_HIP_CHECK(hipMemsetAsync(v->d, 0, sizeof(double2) * n, ctx->stream));
  }
  *out = v;
  return TSC_OK;
}

extern "C" int tsc_vec_destroy(tsc_vec *v) {
  if (!v) return TSC_OK;
  if (v->d) {
    hipStreamSynchronize(v->ctx->stream);
    hipFree(v->d);
  }
  delete v;
  return TSC_OK;
}

extern "C" int tsc_vec_upload(tsc_vec *v, const double *host) {
  TSC_REQUIRE(v && (host || v->n == 0), "tsc_vec_upload arguments");
  if (v->n == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(v->d, host, sizeof(double2) * v->n, hipMemcpyHostToDevice, v->ctx->stream));
  TSC_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_vec_download(const tsc_vec *v, double *host) {
  TSC_REQUIRE(v && (host || v->n == 0), "tsc_vec_download arguments");
  if (v->n == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(host, v->d, sizeof(double2) * v->n, hipMemcpyDeviceToHost, v->ctx->stream));
  TSC_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  return TSC_OK;
}

extern "C" void *tsc_vec_dptr(tsc_vec *v) { return v ? (void *)v->d : nullptr; }

extern "C" int tsc_vec_dot_re(const tsc_vec *u, const tsc_vec *v, double *out) {
  TSC_REQUIRE(u && v && out && u->n == v->n && u->ctx == v->ctx, "tsc_vec_dot_re arguments");
  const double2 *V[1] = {u->d};
  return tsc::dots_re(u->ctx, V, 1, v->d, v->n, out);
}

extern "C" int tsc_vec_norm(const tsc_vec *v, double *out) {
  TSC_REQUIRE(v && out, "tsc_vec_norm arguments");
  double d = 0.0;
  const double2 *V[1] = {v->d};
  int rc = tsc::dots_re(v->ctx, V, 1, v->d, v->n, &d);
  if (rc) return rc;
  *out = std::sqrt(d);
  return TSC_OK;
}

extern "C" int tsc_vec_fill_gaussian(tsc_vec *v, uint64_t seed, uint64_t key, double sigma) {
  TSC_REQUIRE(v, "tsc_vec_fill_gaussian: null vector");
  int rc = tsc::fill_gaussian(v->ctx, v->d, v->n, seed, key, sigma);
  if (rc) return rc;
  TSC_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_vec_axpby(tsc_vec *y, double a, const tsc_vec *x, double b) {
  TSC_REQUIRE(y && x && y->n == x->n && y->ctx == x->ctx, "tsc_vec_axpby arguments");
  if (y->n == 0) return TSC_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(y->n, 256)), dim3(256), 0, y->ctx->stream, y->d, a, x->d, b, y->n);
  TSC_HIP_CHECK(hipGetLastError());
  TSC_HIP_CHECK(hipStreamSynchronize(y->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_vec_copy_range(tsc_vec *y, int64_t y_off, const tsc_vec *x, int64_t x_off, int64_t n) {
  TSC_REQUIRE(y && x && y->ctx == x->ctx && n >= 0 && y_off >= 0 && x_off >= 0 && y_off + n <= y->n &&
                  x_off + n <= x->n,
              "tsc_vec_copy_range arguments");
  if (n == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(y->d + y_off, x->d + x_off, sizeof(double2) * n, hipMemcpyDeviceToDevice,
                               y->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_vec_upload_range(tsc_vec *v, int64_t off, const double *host, int64_t n) {
  TSC_REQUIRE(v && (host || n == 0) && n >= 0 && off >= 0 && off + n <= v->n, "tsc_vec_upload_range arguments");
  if (n == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(v->d + off, host, sizeof(double2) * n, hipMemcpyHostToDevice, v->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_vec_download_range(const tsc_vec *v, int64_t off, double *host, int64_t n) {
  TSC_REQUIRE(v && (host || n == 0) && n >= 0 && off >= 0 && off + n <= v->n, "tsc_vec_download_range arguments");
  if (n == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(host, v->d + off, sizeof(double2) * n, hipMemcpyDeviceToHost, v->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_vec_copy(tsc_vec *y, const tsc_vec *x) {
  TSC_REQUIRE(y && x && y->n == x->n && y->ctx == x->ctx, "tsc_vec_copy arguments");
  if (y->n == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(y->d, x->d, sizeof(double2) * y->n, hipMemcpyDeviceToDevice, y->ctx->stream));
  TSC_HIP_CHECK(hipStreamSynchronize(y->ctx->stream));
  return TSC_OK;
}
