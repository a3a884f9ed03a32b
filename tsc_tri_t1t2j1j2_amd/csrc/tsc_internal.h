// Internal declarations shared by the HIP translation units of libtsc_hotpath.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tsc_hotpath.h"

namespace tsc {

void set_error(const std::string &msg);

#define TSC_HIP_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::tsc::set_error(std::string(#expr " failed: ") + hipGetErrorString(_e) + \
                       " at " __FILE__ ":" + std::to_string(__LINE__));         \
      return TSC_EHIP;                                                          \
    }                                                                           \
  } while (0)

#define TSC_REQUIRE(cond, msg)                        \
  do {                                                \
    if (!(cond)) {                                    \
      ::tsc::set_error(std::string("invalid: ") + msg); \
      return TSC_EINVAL;                              \
    }                                                 \
  } while (0)

// Device-side tile descriptor: one TM x TN window of an output block.
struct Tile {
  int32_t blk;
  int16_t ti, tj;
  int32_t cbeg, cend;  // contribution range of this work unit
  int32_t slot;        // -1: write C directly; else partial-tile slot
  int32_t pad;
};

// Sum of the partial tiles [slot0, slot0 + nslot) into one output tile.
struct Reduce {
  int32_t blk;
  int16_t ti, tj;
  int32_t slot0, nslot;
};

// Base pointers of the four buffers a contribution may address.
struct Bufs {
  const void *p[4];
};

// Launch a grouped contraction (one stage) on `stream`.
// prec: TSC_PREC_*; tiles are for the 64x64 tile shape.
hipError_t launch_grouped(int prec, const Bufs &bufs, const tsc_outblock *d_blocks,
                          const tsc_contrib *d_contribs, const Tile *d_tiles, int ntiles,
                          const Reduce *d_reds, int nreds, void *part, hipStream_t stream,
                          const int *modes);
// Determine the accumulator row layout of the f64 / f32 / bf16 MFMAs.
int probe_layouts(hipStream_t stream, int modes[3]);

// Tile shape of the grouped contraction kernels.
constexpr int kTM = 64;
constexpr int kTN = 64;

constexpr int kRedBlocks = 1024;
constexpr int kMaxVecs = 256;

// Lanczos BLAS-1 helpers (tsc_vec.hip)
int dots_re(tsc_ctx *ctx, const double2 *const *V_host, int nv, const double2 *w, int64_t n, double *dots);
int combine(tsc_ctx *ctx, double2 *w, double a, const double2 *const *V_host, const double *c, int nv, int64_t n);
int scale(tsc_ctx *ctx, double2 *w, double a, int64_t n);
int fill_gaussian(tsc_ctx *ctx, double2 *out, int64_t n, uint64_t seed, uint64_t key, double sigma);

}  // namespace tsc

struct tsc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int mfma_rowmode[3] = {0, 0, 0};  // accumulator layouts (f64, f32, bf16)
  double *d_red = nullptr;   // reduction partials (kRedBlocks * kMaxVecs doubles)
  double *h_red = nullptr;   // pinned host mirror of d_red
  void **d_ptrs = nullptr;   // device copy of vector pointer lists
  void **d_ptrs2 = nullptr;
  double *d_coef = nullptr;  // combination coefficients
  void **h_ptrs = nullptr;   // pinned staging
  double *h_coef = nullptr;
};

struct tsc_vec {
  tsc_ctx *ctx = nullptr;
  int64_t n = 0;
  double2 *d = nullptr;
};

struct tsc_plan {
  tsc_ctx *ctx = nullptr;
  int64_t ops_elems = 0, mid_elems = 0, in_elems = 0, out_elems = 0;
  double2 *d_ops64 = nullptr;
  bool own_ops = true;        // false when the arena is a borrowed device buffer
  float2 *d_ops32 = nullptr;
  double2 *d_mid64 = nullptr;
  float2 *d_mid32 = nullptr;
  tsc_contrib *d_c[2] = {nullptr, nullptr};
  tsc_outblock *d_o[2] = {nullptr, nullptr};
  tsc::Tile *d_t[2] = {nullptr, nullptr};
  int nt[2] = {0, 0};
  tsc::Reduce *d_r[2] = {nullptr, nullptr};
  int nr[2] = {0, 0};
  int64_t nslot[2] = {0, 0};           // partial-tile slots per stage
  void *d_part64 = nullptr;            // partial tiles, FP64 path
  void *d_part32 = nullptr;            // partial tiles, FP32 / BF16x3 paths
  double flops = 0.0;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool timed = false;
  // graph cache: one per precision, keyed by the in/out pointers captured
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  const void *g_in[3] = {nullptr, nullptr, nullptr};
  const void *g_out[3] = {nullptr, nullptr, nullptr};
};

namespace tsc {
int plan_create_impl(tsc_ctx *ctx, const double *ops, const double2 *ops_dev, int64_t ops_elems, int64_t mid_elems,
                     const tsc_contrib *c1, int32_t nc1, const tsc_outblock *o1, int32_t no1, const tsc_contrib *c2,
                     int32_t nc2, const tsc_outblock *o2, int32_t no2, int64_t in_elems, int64_t out_elems,
                     tsc_plan **out);
int hpsi_dev(tsc_plan *p, const double2 *in, double2 *out, int prec);
double tridiag_lowest(const double *d, const double *e, int n);
void tridiag_vector(const double *d, const double *e, int n, double lam, double *z);
}  // namespace tsc
