// Reduced density matrices, their Hermitian eigensolve, the truncation rule
// and the U^H O U renormalisation, on the grouped-contraction engine.
//
// Follows hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/den_mat.f90:4-99 (rho per
// sector), :184-418 (Get_truncation_operator) and block.f90:3465-3683
// (update_trun_dia / update_trun_ndia).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "tsc_internal.h"

using namespace tsc;

namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

// rho_s for every sector, into d_rho (sum d_s^2), through a stage-2-only plan
int build_rho(tsc_ctx *ctx, const double2 *d_psi, int64_t psi_elems, int32_t nsec, const int32_t *sys_dim,
              const int32_t *env_dim, int side, double2 *d_rho, int64_t rho_elems, std::vector<int64_t> &rho_off) {
  std::vector<tsc_contrib> c(nsec);
  std::vector<tsc_outblock> o(nsec);
  int64_t poff = 0, roff = 0;
  rho_off.assign(nsec, 0);
  for (int32_t s = 0; s < nsec; ++s) {
    const int sd = sys_dim[s], ed = env_dim[s];
    TSC_REQUIRE(sd >= 0 && ed >= 0, "negative sector dimension");
    const int d = side == 0 ? sd : ed;
    rho_off[s] = roff;
    tsc_contrib &t = c[s];
    std::memset(&t, 0, sizeof(t));
    t.a_off = poff;
    t.b_off = poff;
    t.lda = std::max(1, sd);
    t.ldb = std::max(1, sd);
    t.a_buf = TSC_BUF_IN;
    t.b_buf = TSC_BUF_IN;
    if (side == 0) {  // psi psi^H : (sd x ed)(ed x sd)
      t.a_op = TSC_OP_N;
      t.b_op = TSC_OP_C;
      t.k = ed;
    } else {  // psi^H psi : (ed x sd)(sd x ed)
      t.a_op = TSC_OP_C;
      t.b_op = TSC_OP_N;
      t.k = sd;
    }
    t.coef_re = 1.0;
    tsc_outblock &b = o[s];
    b.c_off = roff;
    b.ldc = std::max(1, d);
    b.m = d;
    b.n = d;
    b.c_buf = TSC_BUF_OUT;
    b.cbeg = s;
    b.cend = s + 1;
    b.mode = 0;
    poff += (int64_t)sd * ed;
    roff += (int64_t)d * d;
  }
  TSC_REQUIRE(poff <= psi_elems, "sector dimensions exceed psi length");
  TSC_REQUIRE(roff <= rho_elems, "rho buffer too small");
  tsc_plan *p = nullptr;
  int rc = plan_create_impl(ctx, nullptr, nullptr, 0, 0, nullptr, 0, nullptr, 0, c.data(), nsec, o.data(), nsec,
                            psi_elems, rho_elems, &p);
  if (rc) return rc;
  rc = hpsi_dev(p, d_psi, d_rho, TSC_PREC_F64);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("rho build: ") + hipGetErrorString(e));
      rc = TSC_EHIP;
    }
  }
  tsc_plan_destroy(p);
  return rc;
}

// Batched eigensolve. Sectors are sorted by dimension; the small ones are
// cut into buckets of similar size, each one batched Jacobi call
// (rocsolver_zheevj_strided_batched) on copies padded to the bucket's
// largest dimension P with -1 on the padding diagonal ([[rho, 0], [0, -I]]:
// rho's eigenpairs are untouched and its eigenvalues, >= 0, sort after the
// -1s, so they are the last d of P); each large sector is a plain zheevd.
struct EigSlot {
  int64_t rho_off;   // rho_s in the packed rho array (d x d, column-major)
  int64_t slot_off;  // padded copy in the bucket buffer (P x P)
  int64_t w_off;     // eigenvalues of sector s in the output array
  int32_t d, pad;
};

__global__ void eig_pad_kernel(const double2 *__restrict__ rho, double2 *__restrict__ buf,
                               const EigSlot *__restrict__ slots, int P) {
  const EigSlot sl = slots[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)P * P) return;
  const int i = (int)(e % P), j = (int)(e / P);
  double2 v = make_double2(i == j ? -1.0 : 0.0, 0.0);
  if (i < sl.d && j < sl.d) v = rho[sl.rho_off + i + (int64_t)j * sl.d];
  buf[sl.slot_off + e] = v;
}

__global__ void eig_unpad_kernel(double2 *__restrict__ rho, const double2 *__restrict__ buf,
                                 const double *__restrict__ wbuf, double *__restrict__ w,
                                 const EigSlot *__restrict__ slots, int P) {
  const EigSlot sl = slots[blockIdx.y];
  const int d = sl.d;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (int64_t)d * d) {
    const int i = (int)(e % d), j = (int)(e / d);
    rho[sl.rho_off + e] = buf[sl.slot_off + i + (int64_t)(P - d + j) * P];
  }
  if (e < d) w[sl.w_off + e] = wbuf[(int64_t)blockIdx.y * P + P - d + e];
}

constexpr int kJacobiMax = 64;  // largest sector solved in a padded Jacobi batch
constexpr int kEigStreams = 16;  // side streams (and submitting threads) for concurrent eigensolves

// Side streams (default priority) with one rocBLAS handle each, per device
// and host thread, kept for the life of the process.
struct EigLanes {
  int device = -1;
  hipStream_t stream[kEigStreams] = {};
  rocblas_handle handle[kEigStreams] = {};
  hipEvent_t done[kEigStreams] = {};
  hipEvent_t start = nullptr;
};

int eig_lanes(tsc_ctx *ctx, EigLanes **out) {
  static thread_local std::vector<EigLanes> lanes;
  for (auto &l : lanes)
    if (l.device == ctx->device) {
      *out = &l;
      return TSC_OK;
    }
  EigLanes l;
  l.device = ctx->device;
  for (int i = 0; i < kEigStreams; ++i) {
    TSC_HIP_CHECK(hipStreamCreate(&l.stream[i]));
    TSC_HIP_CHECK(hipEventCreateWithFlags(&l.done[i], hipEventDisableTiming));
    if (rocblas_create_handle(&l.handle[i]) != rocblas_status_success) {
      set_error("rocblas_create_handle failed");
      return TSC_EHIP;
    }
    rocblas_set_stream(l.handle[i], l.stream[i]);
  }
  TSC_HIP_CHECK(hipEventCreateWithFlags(&l.start, hipEventDisableTiming));
  lanes.push_back(l);
  *out = &lanes.back();
  return TSC_OK;
}

// ascending eigensolve of every rho_s (ZHEEV 'V','U' role, den_mat.f90:229-253):
// eigenvectors replace rho_s in place, eigenvalues go to d_w. Buckets run
// concurrently on the side streams (the solver is launch-bound at these
// sizes), largest first onto the least-loaded stream, one submitting host
// thread per stream.
int eig_sectors(tsc_ctx *ctx, int32_t nsec, const int32_t *dim, double2 *d_rho, const std::vector<int64_t> &rho_off,
                double *d_w, const std::vector<int64_t> &w_off) {
  EigLanes *L = nullptr;
  int rc = eig_lanes(ctx, &L);
  if (rc) return rc;
  std::vector<int32_t> order;
  for (int32_t s = 0; s < nsec; ++s)
    if (dim[s] > 0) order.push_back(s);
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return dim[x] < dim[y]; });
  struct Bucket {
    size_t b0, cnt;
    int P;
    int64_t buf_off, w_off, slot0;  // into the shared work buffers
  };
  std::vector<Bucket> buckets;
  std::vector<EigSlot> slots;
  int64_t buf_len = 0, wlen = 0;
  for (size_t b0 = 0; b0 < order.size();) {
    // bucket: sectors of dimension <= kJacobiMax within 25% (+8) of its
    // smallest; larger sectors one by one
    const int dmin = dim[order[b0]];
    size_t b1 = b0 + 1;
    if (dmin <= kJacobiMax)
      while (b1 < order.size() && dim[order[b1]] <= std::min(kJacobiMax, dmin + dmin / 4 + 8)) ++b1;
    const int P = dim[order[b1 - 1]];
    Bucket bk{b0, b1 - b0, P, buf_len, wlen, (int64_t)slots.size()};
    for (size_t q = 0; q < bk.cnt; ++q) {
      const int32_t s = order[b0 + q];
      slots.push_back(EigSlot{rho_off[s], buf_len + (int64_t)q * P * P, w_off[s], dim[s], 0});
    }
    buf_len += (int64_t)bk.cnt * P * P;
    wlen += (int64_t)bk.cnt * P;
    buckets.push_back(bk);
    b0 = b1;
  }
  if (buckets.empty()) return TSC_OK;
  const int64_t nslots = (int64_t)slots.size();
  DevBuf buf, wbuf, ebuf, info, dsl, resid, sweeps;
  TSC_HIP_CHECK(hipMalloc(&buf.p, sizeof(double2) * (size_t)buf_len));
  TSC_HIP_CHECK(hipMalloc(&wbuf.p, sizeof(double) * (size_t)wlen));
  TSC_HIP_CHECK(hipMalloc(&ebuf.p, sizeof(double) * (size_t)wlen));
  TSC_HIP_CHECK(hipMalloc(&info.p, sizeof(rocblas_int) * nslots));
  TSC_HIP_CHECK(hipMalloc(&resid.p, sizeof(double) * nslots));
  TSC_HIP_CHECK(hipMalloc(&sweeps.p, sizeof(rocblas_int) * nslots));
  TSC_HIP_CHECK(hipMalloc(&dsl.p, sizeof(EigSlot) * nslots));
  hipStream_t st = ctx->stream;
  TSC_HIP_CHECK(hipMemcpyAsync(dsl.p, slots.data(), sizeof(EigSlot) * nslots, hipMemcpyHostToDevice, st));
  TSC_HIP_CHECK(hipEventRecord(L->start, st));
  for (int i = 0; i < kEigStreams; ++i) TSC_HIP_CHECK(hipStreamWaitEvent(L->stream[i], L->start, 0));
  // largest buckets first, each onto the least-loaded side stream
  std::vector<size_t> border(buckets.size());
  std::iota(border.begin(), border.end(), 0);
  auto work = [&](const Bucket &bk) { return (double)bk.cnt * bk.P * bk.P * bk.P; };
  std::stable_sort(border.begin(), border.end(), [&](size_t x, size_t y) { return work(buckets[x]) > work(buckets[y]); });
  std::vector<double> load(kEigStreams, 0.0);
  std::vector<std::vector<size_t>> lane_buckets(kEigStreams);
  for (size_t bi : border) {
    const int lane = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[lane] += work(buckets[bi]) + 1e6;  // fixed cost per call
    lane_buckets[lane].push_back(bi);
  }
  // one bucket on its lane: pad, solve, unpad; returns 0, 1 (launch) or 2
  // (solver status), and never touches the shared error string
  auto run_bucket = [&](size_t bi, int lane) -> int {
    const Bucket &bk = buckets[bi];
    hipStream_t ls = L->stream[lane];
    rocblas_handle h = L->handle[lane];
    const int P = bk.P, cnt = (int)bk.cnt;
    const int64_t pp = (int64_t)P * P;
    const EigSlot *dslot = (const EigSlot *)dsl.p + bk.slot0;
    double2 *bbuf = (double2 *)buf.p + bk.buf_off;
    double *bw = (double *)wbuf.p + bk.w_off;
    rocblas_int *binfo = (rocblas_int *)info.p + bk.slot0;
    dim3 grid((unsigned)((pp + 255) / 256), (unsigned)cnt);
    // slot offsets address the shared buffer; pass its base
    hipLaunchKernelGGL(eig_pad_kernel, grid, dim3(256), 0, ls, (const double2 *)d_rho, (double2 *)buf.p, dslot, P);
    if (hipGetLastError() != hipSuccess) return 1;
    rocblas_status rs;
    if (P > kJacobiMax) {
      // a large sector alone: no padding (P == d), divide and conquer
      rs = rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, P,
                            reinterpret_cast<rocblas_double_complex *>(bbuf), P, bw, (double *)ebuf.p + bk.w_off,
                            binfo);
    } else {
      // padded batch: Jacobi (exactly decoupled, degenerate padding is harmless to it)
      rs = rocsolver_zheevj_strided_batched(h, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper,
                                            P, reinterpret_cast<rocblas_double_complex *>(bbuf), P, pp, 0.0,
                                            (double *)resid.p + bk.slot0, 100, (rocblas_int *)sweeps.p + bk.slot0,
                                            bw, P, binfo, cnt);
    }
    if (rs != rocblas_status_success) return 2;
    hipLaunchKernelGGL(eig_unpad_kernel, grid, dim3(256), 0, ls, d_rho, (const double2 *)buf.p, bw, d_w, dslot, P);
    return hipGetLastError() != hipSuccess ? 1 : 0;
  };
  // Each lane is fed by its own host thread: the divide-and-conquer solver
  // waits on the host inside a call, so a single submitting thread would
  // leave the other lanes idle meanwhile.
  std::vector<int> lane_rc(kEigStreams, 0);
  {
    std::vector<std::thread> pool;
    const int dev = ctx->device;
    auto feed = [&](int lane) {
      for (size_t bi : lane_buckets[lane])
        if ((lane_rc[lane] = run_bucket(bi, lane)) != 0) return;
    };
    for (int lane = 0; lane < kEigStreams; ++lane) {
      if (lane_buckets[lane].empty()) continue;
      try {
        pool.emplace_back([&, lane, dev]() {
          if (hipSetDevice(dev) != hipSuccess) {
            lane_rc[lane] = 1;
            return;
          }
          feed(lane);
        });
      } catch (...) {
        feed(lane);  // no thread available: this lane is fed from the calling thread
      }
    }
    for (auto &t : pool) t.join();
  }
  for (int lane = 0; lane < kEigStreams; ++lane)
    if (lane_rc[lane]) {
      set_error(lane_rc[lane] == 2 ? "rocSOLVER Hermitian eigensolver failed" : "eigensolver launch failed");
      return TSC_EHIP;
    }
  for (int i = 0; i < kEigStreams; ++i) {
    TSC_HIP_CHECK(hipEventRecord(L->done[i], L->stream[i]));
    TSC_HIP_CHECK(hipStreamWaitEvent(st, L->done[i], 0));
  }
  std::vector<rocblas_int> hinfo(nslots);
  TSC_HIP_CHECK(hipMemcpyAsync(hinfo.data(), info.p, sizeof(rocblas_int) * nslots, hipMemcpyDeviceToHost, st));
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  for (const Bucket &bk : buckets)
    for (size_t q = 0; q < bk.cnt; ++q)
      if (hinfo[bk.slot0 + q] != 0) {
        set_error("Hermitian eigensolver did not converge for sector " + std::to_string(order[bk.b0 + q]));
        return TSC_EHIP;
      }
  return TSC_OK;
}

// The reference's descending quicksort with its index array
// (QuickSort_Double, block.f90:2394-2436): midpoint pivot, Hoare-style
// two-pointer partition, recursion on both parts. Kept step for step so
// that ties at the truncation cut are ordered the same way. 1-based.
void quicksort_desc(std::vector<double> &a, std::vector<int> &idx, int lower, int upper) {
  int i = lower, j = upper;
  const int mid = (int)((lower + upper) / 2.0);
  const double piv = a[mid];
  while (i <= j) {
    while (a[i] > piv) ++i;
    while (a[j] < piv) --j;
    if (i <= j) {
      std::swap(a[i], a[j]);
      std::swap(idx[i], idx[j]);
      ++i;
      --j;
    }
  }
  if (lower < j) quicksort_desc(a, idx, lower, j);
  if (i < upper) quicksort_desc(a, idx, i, upper);
}

// Kept-state selection of Get_truncation_operator (den_mat.f90:255-364):
// eig (ascending per sector) -> kept counts per sector and scalars.
void select_states(int32_t nsec, const int32_t *dim, const int32_t *twos, const double *eig, int32_t kept_min,
                   int32_t kept_max, double trun_err, int32_t *kept, double *scalars) {
  int64_t tot = 0;
  for (int32_t s = 0; s < nsec; ++s) tot += dim[s];
  const int n = (int)tot;
  std::vector<double> ev(n + 1);
  std::vector<int> idx(n + 1), owner(n + 1);
  std::vector<int> wts(n + 1);
  int pos = 0;
  for (int32_t s = 0; s < nsec; ++s)
    for (int x = 0; x < dim[s]; ++x) {
      ++pos;
      ev[pos] = eig[pos - 1];
      idx[pos] = pos;
      owner[pos] = s;
      wts[pos] = twos[s] + 1;
    }
  // U(1)-resolved entropy on the raw spectrum (den_mat.f90:275-282)
  double entropy1 = 0.0;
  for (int i = 1; i <= n; ++i)
    if (ev[i] >= 1.0e-15) {
      const double v = ev[i] / wts[i];
      entropy1 -= v * std::log(v) * wts[i];
    }
  if (n > 0) quicksort_desc(ev, idx, 1, n);
  double tot_pm = 0.0;
  for (int i = 1; i <= n; ++i) tot_pm += ev[i];
  for (int i = 1; i <= n; ++i) ev[i] /= tot_pm;
  int num_state;
  double err = 0.0;
  if (n <= kept_min) {
    num_state = n;
    err = 0.0;
  } else {
    for (int i = kept_min + 1; i <= n; ++i) err += ev[i];
    num_state = kept_min;
    // the reference's loop (den_mat.f90:306-314) has no bound at the end of
    // the spectrum: with trun_err = 0 a rounding remainder of err could step
    // past it, so the count stops at n
    while (err > trun_err && num_state < n) {
      ++num_state;
      err -= ev[num_state];
      if (err <= trun_err || num_state >= kept_max) break;
    }
  }
  double ent = 0.0;
  for (int i = 1; i <= n; ++i)
    if (ev[i] >= 1.0e-15) ent -= ev[i] * std::log(ev[i]);
  for (int32_t s = 0; s < nsec; ++s) kept[s] = 0;
  for (int k = 1; k <= num_state; ++k) kept[owner[idx[k]]] += 1;
  scalars[0] = num_state;
  scalars[1] = err;
  scalars[2] = ent;
  scalars[3] = entropy1;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int truncate_device(tsc_ctx *ctx, int32_t nsec, const int32_t *dim, const int32_t *twos, double2 *d_rho,
                    const std::vector<int64_t> &rho_off, int32_t kept_min, int32_t kept_max, double trun_err,
                    double *eigval_out, int32_t *kept_out, double *u_out, int64_t u_cap, double *scalars_out) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int64_t> w_off(nsec);
  int64_t wt = 0;
  for (int32_t s = 0; s < nsec; ++s) {
    w_off[s] = wt;
    wt += dim[s];
  }
  DevBuf dw;
  TSC_HIP_CHECK(hipMalloc(&dw.p, sizeof(double) * std::max<int64_t>(wt, 1)));
  int rc = eig_sectors(ctx, nsec, dim, d_rho, rho_off, (double *)dw.p, w_off);
  if (rc) return rc;
  ctx->rho_ms[1] = (float)ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  std::vector<double> w(std::max<int64_t>(wt, 1));
  TSC_HIP_CHECK(hipMemcpy(w.data(), dw.p, sizeof(double) * wt, hipMemcpyDeviceToHost));
  select_states(nsec, dim, twos, w.data(), kept_min, kept_max, trun_err, kept_out, scalars_out);
  if (eigval_out) std::memcpy(eigval_out, w.data(), sizeof(double) * wt);
  int64_t u_need = 0;
  for (int32_t s = 0; s < nsec; ++s) u_need += (int64_t)dim[s] * kept_out[s];
  if (u_need > u_cap) {
    set_error("u_out holds " + std::to_string(u_cap) + " complex values; the kept eigenvectors need " +
              std::to_string(u_need) + " (sum over sectors of dim * kept; sum dim^2 always suffices)");
    return TSC_EINVAL;
  }
  ctx->rho_ms[2] = (float)ms_since(t0);
  t0 = std::chrono::steady_clock::now();
  // U_s = last kept_s eigenvector columns of sector s (den_mat.f90:392-398)
  int64_t uoff = 0;
  for (int32_t s = 0; s < nsec; ++s) {
    const int d = dim[s], k = kept_out[s];
    if (k == 0) continue;
    const double2 *src = d_rho + rho_off[s] + (int64_t)d * (d - k);
    TSC_HIP_CHECK(hipMemcpyAsync(u_out + 2 * uoff, src, sizeof(double2) * (int64_t)d * k, hipMemcpyDeviceToHost,
                                 ctx->stream));
    uoff += (int64_t)d * k;
  }
  TSC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < 2 * uoff; ++i)
    if (std::isnan(u_out[i])) {
      set_error("density-matrix eigensolver returned NaN eigenvectors");
      return TSC_EHIP;
    }
  ctx->rho_ms[3] = (float)ms_since(t0);
  return TSC_OK;
}

}  // namespace

extern "C" int tsc_density_matrix(tsc_ctx *ctx, const tsc_vec *psi, int32_t nsec, const int32_t *sys_dim,
                                  const int32_t *env_dim, int side, double *rho_out) {
  TSC_REQUIRE(ctx && psi && (nsec == 0 || (sys_dim && env_dim)) && rho_out, "tsc_density_matrix arguments");
  TSC_REQUIRE(side == 0 || side == 1, "side must be 0 (sys) or 1 (env)");
  int64_t rho_elems = 0;
  for (int32_t s = 0; s < nsec; ++s) {
    const int64_t d = side == 0 ? sys_dim[s] : env_dim[s];
    rho_elems += d * d;
  }
  DevBuf rho;
  TSC_HIP_CHECK(hipMalloc(&rho.p, sizeof(double2) * std::max<int64_t>(rho_elems, 1)));
  std::vector<int64_t> off;
  int rc = build_rho(ctx, psi->d, psi->n, nsec, sys_dim, env_dim, side, (double2 *)rho.p, rho_elems, off);
  if (rc) return rc;
  TSC_HIP_CHECK(hipMemcpy(rho_out, rho.p, sizeof(double2) * rho_elems, hipMemcpyDeviceToHost));
  return TSC_OK;
}

extern "C" int tsc_truncation_operator(tsc_ctx *ctx, int32_t nsec, const int32_t *dim, const int32_t *twos,
                                       const double *rho, int32_t kept_min, int32_t kept_max, double trun_err,
                                       double *eigval_out, int32_t *kept_out, double *u_out, int64_t u_cap,
                                       double *scalars_out) {
  TSC_REQUIRE(ctx && (nsec == 0 || (dim && twos && rho)) && kept_out && u_out && scalars_out,
              "tsc_truncation_operator arguments");
  TSC_REQUIRE(kept_min >= 0 && kept_max >= 0 && trun_err >= 0.0, "kept_min, kept_max and trun_err must be >= 0");
  int64_t rho_elems = 0;
  std::vector<int64_t> off(nsec);
  for (int32_t s = 0; s < nsec; ++s) {
    TSC_REQUIRE(dim[s] >= 0 && twos[s] >= 0, "negative dim / 2S");
    off[s] = rho_elems;
    rho_elems += (int64_t)dim[s] * dim[s];
  }
  auto t0 = std::chrono::steady_clock::now();
  DevBuf drho;
  TSC_HIP_CHECK(hipMalloc(&drho.p, sizeof(double2) * std::max<int64_t>(rho_elems, 1)));
  TSC_HIP_CHECK(hipMemcpy(drho.p, rho, sizeof(double2) * rho_elems, hipMemcpyHostToDevice));
  ctx->rho_ms[0] = (float)ms_since(t0);
  return truncate_device(ctx, nsec, dim, twos, (double2 *)drho.p, off, kept_min, kept_max, trun_err, eigval_out,
                         kept_out, u_out, u_cap, scalars_out);
}

extern "C" int tsc_rho_truncate(tsc_ctx *ctx, const tsc_vec *psi, int32_t nsec, const int32_t *sys_dim,
                                const int32_t *env_dim, const int32_t *twos, int side, int32_t kept_min,
                                int32_t kept_max, double trun_err, double *eigval_out, int32_t *kept_out,
                                double *u_out, int64_t u_cap, double *scalars_out) {
  TSC_REQUIRE(ctx && psi && (nsec == 0 || (sys_dim && env_dim && twos)) && kept_out && u_out && scalars_out,
              "tsc_rho_truncate arguments");
  TSC_REQUIRE(kept_min >= 0 && kept_max >= 0 && trun_err >= 0.0, "kept_min, kept_max and trun_err must be >= 0");
  TSC_REQUIRE(side == 0 || side == 1, "side must be 0 (sys) or 1 (env)");
  std::vector<int32_t> dim(nsec);
  int64_t rho_elems = 0;
  for (int32_t s = 0; s < nsec; ++s) {
    dim[s] = side == 0 ? sys_dim[s] : env_dim[s];
    rho_elems += (int64_t)dim[s] * dim[s];
  }
  auto t0 = std::chrono::steady_clock::now();
  DevBuf rho;
  TSC_HIP_CHECK(hipMalloc(&rho.p, sizeof(double2) * std::max<int64_t>(rho_elems, 1)));
  std::vector<int64_t> off;
  int rc = build_rho(ctx, psi->d, psi->n, nsec, sys_dim, env_dim, side, (double2 *)rho.p, rho_elems, off);
  if (rc) return rc;
  ctx->rho_ms[0] = (float)ms_since(t0);
  return truncate_device(ctx, nsec, dim.data(), twos, (double2 *)rho.p, off, kept_min, kept_max, trun_err,
                         eigval_out, kept_out, u_out, u_cap, scalars_out);
}

extern "C" int tsc_rho_last_times(const tsc_ctx *ctx, float *ms4) {
  TSC_REQUIRE(ctx && ms4, "tsc_rho_last_times arguments");
  for (int i = 0; i < 4; ++i) ms4[i] = ctx->rho_ms[i];
  return TSC_OK;
}

extern "C" int tsc_select_states(int32_t nsec, const int32_t *dim, const int32_t *twos, const double *eigval,
                                 int32_t kept_min, int32_t kept_max, double trun_err, int32_t *kept_out,
                                 double *scalars_out) {
  TSC_REQUIRE(nsec >= 0 && (nsec == 0 || (dim && twos && eigval)) && kept_out && scalars_out,
              "tsc_select_states arguments");
  TSC_REQUIRE(kept_min >= 0 && kept_max >= 0 && trun_err >= 0.0, "kept_min, kept_max and trun_err must be >= 0");
  for (int32_t s = 0; s < nsec; ++s) TSC_REQUIRE(dim[s] >= 0 && twos[s] >= 0, "negative dim / 2S");
  select_states(nsec, dim, twos, eigval, kept_min, kept_max, trun_err, kept_out, scalars_out);
  return TSC_OK;
}

extern "C" int tsc_update_trun(tsc_ctx *ctx, int32_t nblk, const int32_t *row, const int32_t *col, const int32_t *kl,
                               const int32_t *kr, const int64_t *ori_off, const int64_t *ul_off,
                               const int64_t *ur_off, const int64_t *eff_off, const double *ori, int64_t ori_elems,
                               const double *u, int64_t u_elems, double *eff, int64_t eff_elems, int prec) {
  TSC_REQUIRE(ctx && nblk >= 0 && eff, "tsc_update_trun arguments");
  TSC_REQUIRE(nblk == 0 || (row && col && kl && kr && ori_off && ul_off && ur_off && eff_off && ori && u),
              "tsc_update_trun arrays");
  std::vector<tsc_contrib> c1(nblk), c2(nblk);
  std::vector<tsc_outblock> o1(nblk), o2(nblk);
  int64_t moff = 0;
  for (int32_t i = 0; i < nblk; ++i) {
    TSC_REQUIRE(row[i] >= 0 && col[i] >= 0 && kl[i] >= 0 && kr[i] >= 0, "negative block dims");
    // stage 1: mid = O (row x col) . U_rhs (col x kr)
    tsc_contrib &a = c1[i];
    std::memset(&a, 0, sizeof(a));
    a.a_off = ori_off[i];
    a.lda = std::max(1, row[i]);
    a.a_buf = TSC_BUF_OPS;
    a.a_op = TSC_OP_N;
    a.b_off = ur_off[i];
    a.ldb = std::max(1, col[i]);
    a.b_buf = TSC_BUF_IN;
    a.b_op = TSC_OP_N;
    a.k = col[i];
    a.coef_re = 1.0;
    tsc_outblock &ob1 = o1[i];
    ob1.c_off = moff;
    ob1.ldc = std::max(1, row[i]);
    ob1.m = row[i];
    ob1.n = kr[i];
    ob1.c_buf = TSC_BUF_MID;
    ob1.cbeg = i;
    ob1.cend = i + 1;
    ob1.mode = 0;
    // stage 2: eff = U_lhs^H (kl x row) . mid (row x kr)
    tsc_contrib &b = c2[i];
    std::memset(&b, 0, sizeof(b));
    b.a_off = ul_off[i];
    b.lda = std::max(1, row[i]);
    b.a_buf = TSC_BUF_IN;
    b.a_op = TSC_OP_C;
    b.b_off = moff;
    b.ldb = std::max(1, row[i]);
    b.b_buf = TSC_BUF_MID;
    b.b_op = TSC_OP_N;
    b.k = row[i];
    b.coef_re = 1.0;
    tsc_outblock &ob2 = o2[i];
    ob2.c_off = eff_off[i];
    ob2.ldc = std::max(1, kl[i]);
    ob2.m = kl[i];
    ob2.n = kr[i];
    ob2.c_buf = TSC_BUF_OUT;
    ob2.cbeg = i;
    ob2.cend = i + 1;
    ob2.mode = 0;
    moff += (int64_t)row[i] * kr[i];
  }
  tsc_plan *p = nullptr;
  int rc = plan_create_impl(ctx, ori, nullptr, ori_elems, moff, c1.data(), nblk, o1.data(), nblk, c2.data(), nblk,
                            o2.data(), nblk, u_elems, eff_elems, &p);
  if (rc) return rc;
  rc = tsc_hpsi_host(p, u, eff, prec);
  tsc_plan_destroy(p);
  return rc;
}

// Racah formula for the 6j symbol, doubled arguments (the role of W6JS,
// block.f90:3744-3795): {j1 j2 j3; l1 l2 l3} with triads (j1 j2 j3),
// (j1 l2 l3), (l1 j2 l3), (l1 l2 j3). Factorials exact in double to 170!.
extern "C" double tsc_w6j(int32_t j1, int32_t j2, int32_t j3, int32_t l1, int32_t l2, int32_t l3) {
  static double fact[171];
  static bool init = false;
  if (!init) {
    fact[0] = 1.0;
    for (int i = 1; i <= 170; ++i) fact[i] = fact[i - 1] * (double)i;
    init = true;
  }
  auto tri_ok = [](int a, int b, int c) {
    return a >= 0 && b >= 0 && c >= 0 && c <= a + b && c >= std::abs(a - b) && ((a + b + c) % 2 == 0);
  };
  if (!tri_ok(j1, j2, j3) || !tri_ok(j1, l2, l3) || !tri_ok(l1, j2, l3) || !tri_ok(l1, l2, j3)) return 0.0;
  auto delta = [&](int a, int b, int c) {
    return std::sqrt(fact[(a + b - c) / 2] * fact[(a - b + c) / 2] * fact[(-a + b + c) / 2] / fact[(a + b + c) / 2 + 1]);
  };
  const int a1 = (j1 + j2 + j3) / 2, a2 = (j1 + l2 + l3) / 2, a3 = (l1 + j2 + l3) / 2, a4 = (l1 + l2 + j3) / 2;
  const int b1 = (j1 + j2 + l1 + l2) / 2, b2 = (j2 + j3 + l2 + l3) / 2, b3 = (j3 + j1 + l3 + l1) / 2;
  const int tmin = std::max(std::max(a1, a2), std::max(a3, a4));
  const int tmax = std::min(b1, std::min(b2, b3));
  if (tmax + 1 > 170) return std::nan("");
  double sum = 0.0;
  for (int t = tmin; t <= tmax; ++t) {
    const double den = fact[t - a1] * fact[t - a2] * fact[t - a3] * fact[t - a4] * fact[b1 - t] * fact[b2 - t] *
                       fact[b3 - t];
    const double term = fact[t + 1] / den;
    sum += (t % 2 == 0) ? term : -term;
  }
  return sum * delta(j1, j2, j3) * delta(j1, l2, l3) * delta(l1, j2, l3) * delta(l1, l2, j3);
}
