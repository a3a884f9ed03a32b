// Per-product loop path of a contraction plan: every GEMM contribution is
// its own rocBLAS ZGEMM call, issued block by block on the context stream.
// This is the structure of the reference's H.psi (model_wave and its
// kernels, DMRG_code/get_sys.f90:1783-5204 and 6915-7393, call ZGEMM once
// per operator pair and sector), kept as the baseline that the grouped
// kernels are measured against; it runs the same task arrays.
#include <rocblas/rocblas.h>

#include <algorithm>

#include "tsc_internal.h"

using namespace tsc;

namespace {

struct LoopHandle {
  int device = -1;
  rocblas_handle h = nullptr;
};

// one rocBLAS handle per device and host thread, kept for the life of the
// process (destroying it at exit would race the runtime's own teardown)
int loop_handle(tsc_ctx *ctx, rocblas_handle *out) {
  static thread_local std::vector<LoopHandle> hs;
  for (auto &x : hs)
    if (x.device == ctx->device) {
      *out = x.h;
      return rocblas_set_stream(x.h, ctx->stream) == rocblas_status_success ? TSC_OK : TSC_EHIP;
    }
  LoopHandle x;
  x.device = ctx->device;
  if (rocblas_create_handle(&x.h) != rocblas_status_success) {
    set_error("rocblas_create_handle failed");
    return TSC_EHIP;
  }
  hs.push_back(x);
  *out = x.h;
  return rocblas_set_stream(x.h, ctx->stream) == rocblas_status_success ? TSC_OK : TSC_EHIP;
}

rocblas_operation blas_op(int op) {
  return op == TSC_OP_T ? rocblas_operation_transpose
                        : (op == TSC_OP_C ? rocblas_operation_conjugate_transpose : rocblas_operation_none);
}

}  // namespace

extern "C" int tsc_hpsi_loop(tsc_plan *p, const tsc_vec *in, tsc_vec *out) {
  TSC_REQUIRE(p && in && out, "tsc_hpsi_loop arguments");
  TSC_REQUIRE(in->n >= p->in_elems && out->n >= p->out_elems, "vector shorter than the plan's input/output");
  // every contribution must be a plain GEMM with operand modes BLAS has
  for (int s = 0; s < 2; ++s)
    for (const tsc_contrib &c : p->h_c[s]) {
      TSC_REQUIRE(c.a_op != TSC_OP_I && c.b_op != TSC_OP_I,
                  "the per-product loop runs GEMM contributions only (no elementwise terms)");
      TSC_REQUIRE(c.a_op != TSC_OP_CONJ && c.b_op != TSC_OP_CONJ,
                  "the per-product loop has no conjugate-without-transpose operand");
    }
  rocblas_handle h = nullptr;
  int rc = loop_handle(p->ctx, &h);
  if (rc) {
    set_error("rocBLAS handle for the per-product loop");
    return rc;
  }
  rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
  double2 *base[4] = {p->d_ops64, const_cast<double2 *>(in->d), p->d_mid64, out->d};
  for (int s = 0; s < 2; ++s) {
    const std::vector<tsc_outblock> &ob = p->h_o[s];
    const std::vector<tsc_contrib> &cs = p->h_c[s];
    for (const tsc_outblock &o : ob) {
      if (o.m <= 0 || o.n <= 0) continue;
      double2 *C = base[o.c_buf] + o.c_off;
      bool first = (o.mode == 0);
      if (first && o.cbeg == o.cend) {
        TSC_HIP_CHECK(hipMemset2DAsync(C, sizeof(double2) * (size_t)o.ldc, 0, sizeof(double2) * (size_t)o.m,
                                       (size_t)o.n, p->ctx->stream));
        continue;
      }
      for (int i = o.cbeg; i < o.cend; ++i) {
        const tsc_contrib &c = cs[i];
        const rocblas_double_complex alpha(c.coef_re, c.coef_im);
        const rocblas_double_complex beta(first ? 0.0 : 1.0, 0.0);
        const double2 *A = base[(int)c.a_buf] + c.a_off;
        const double2 *B = base[(int)c.b_buf] + c.b_off;
        rocblas_status st = rocblas_zgemm(h, blas_op(c.a_op), blas_op(c.b_op), o.m, o.n, c.k, &alpha,
                                          (const rocblas_double_complex *)A, std::max(1, c.lda),
                                          (const rocblas_double_complex *)B, std::max(1, c.ldb), &beta,
                                          (rocblas_double_complex *)C, std::max(1, o.ldc));
        if (st != rocblas_status_success) {
          set_error("rocblas_zgemm failed in the per-product loop: " + std::string(rocblas_status_to_string(st)));
          return TSC_EHIP;
        }
        first = false;
      }
    }
  }
  return TSC_OK;
}
