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
// Determine the accumulator row layout of the f64 / f32 MFMAs and check the
// 32x32x16 bf16 one.
int probe_layouts(hipStream_t stream, int modes[3]);

// ---------------------------------------------------------------------------
// Packed-operand BF16x3 path.
//
// Every GEMM operand is read from a "panel": P[row][k] = op(X) restricted to
// one contribution's window, rows = output rows (A) or output columns (B),
// stored as bf16 hi/lo splits of the FP32 value. k is zero-padded to a
// multiple of 16 (one MFMA step) and cut into chunks of 8; rows are padded
// to a multiple of 64. Storage unit: an "entry" of 16 bytes = 8 bf16 = one
// MFMA fragment (8 consecutive k of one row and plane). A panel is a
// sequence of 1 KB pieces, piece (rb, c, p) = the entries of rows
// 64*rb .. 64*rb+63 for chunk c and plane p (0 re hi, 1 im hi, 2 re lo,
// 3 im lo):
//     entry(r, c, p) = ((r / 64 * nch + c) * 4 + p) * 64 + r % 64
// so a workgroup moves one step's operands with contiguous 1 KB loads and
// a half-wave reads one fragment set from 512 contiguous LDS bytes.
enum PkBuf { PK_OPS = 0, PK_IN = 1, PK_MID = 2 };
constexpr int kPkTile = 128;  // output tile of one packed-path workgroup (4 waves of 64x64)

struct PkContrib {
  int64_t a_off, b_off;  // entry index of the operand panels
  int32_t a_nch, b_nch;  // chunks per panel row (padded k / 8)
  int32_t ksteps;        // padded k / 16; 0: no GEMM (elementwise or zero coefficient)
  int32_t flags;         // bit0 conj A, bit1 conj B, bits 4-5 A PkBuf, bits 6-7 B PkBuf
  float rs_re, rs_im;    // coefficient of the previous GEMM contribution / this one
  float g_re, g_im;      // this contribution's coefficient
};

// Packed destinations of one stage-1 (MID) output block: panel as a row
// operand (r_off, rows = block rows) and as a transposed one (c_off, rows =
// block columns); -1 when no contribution reads it that way.
struct PkOut {
  int64_t r_off, c_off;
  int32_t r_nch, c_nch;
};

// One packing job: the (row, chunk) items [first, first + 256) of one panel,
// item = chunk * padded_rows + row.
struct PkPanel {
  int64_t src_off;  // element offset of the window in its complex128 buffer
  int64_t dst;      // entry index of the panel
  int32_t ld;       // leading dimension of the window
  int32_t rows;     // panel rows (before padding to 64)
  int32_t kdim;     // k extent before padding
  int32_t nch;      // chunks per row
  int32_t trans;    // 0: P[r][k] = X[r + k*ld]; 1: P[r][k] = X[k + r*ld]
  int32_t pad;
};
struct PkJob {
  int32_t panel;
  int32_t pad;
  int64_t first;
};

struct PkStage {
  std::vector<tsc_contrib> h_cp;  // host copy of d_cp (work estimates when the units are re-cut)
  PkContrib *d_pc = nullptr;   // packed-path contribution records (permuted order)
  tsc_contrib *d_cp = nullptr; // the same contributions, original records (elementwise terms)
  PkOut *d_po = nullptr;       // stage-1 blocks: packed destinations
  Tile *d_t = nullptr;         // 128x128 units
  Reduce *d_r = nullptr;
  int nt = 0, nr = 0;
  int64_t nslot = 0;
};

struct PkState {
  std::string unsupported;  // non-empty: reason the plan cannot run on the packed path
  void *d_ops = nullptr, *d_in = nullptr, *d_mid = nullptr;
  int64_t ops_len = 0, in_len = 0, mid_len = 0;  // entries
  PkPanel *d_ops_panels = nullptr, *d_in_panels = nullptr;
  PkJob *d_ops_jobs = nullptr, *d_in_jobs = nullptr;
  int n_ops_jobs = 0, n_in_jobs = 0;
  PkStage st[2];
  void *d_part = nullptr;  // float2 partial tiles
  int64_t part_slots = 0;  // capacity of d_part in 128x128 tiles
  double desired = 2048.0; // work-unit target of make_tiles for both stages
};

void pk_free(PkState *s);

// FP64 path (tsc_gd.hip): 128x128 units of 8 waves over the plan's own
// complex128 buffers.
constexpr int kGdTile = 128;
struct GdStage {
  std::vector<tsc_contrib> h_c;  // host copy of d_c (work estimates when the units are re-cut)
  tsc_contrib *d_c = nullptr;  // contributions, each block's GEMM terms first
  Tile *d_t = nullptr;
  Reduce *d_r = nullptr;
  int nt = 0, nr = 0;
};
struct GdState {
  GdStage st[2];
  double2 *d_part = nullptr;  // partial tiles of split units
  int64_t part_slots = 0;     // capacity of d_part in units
  double desired = 2048.0;    // work-unit target of make_tiles for both stages
};
void gd_free(GdState *s);

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
  // wall milliseconds of the last tsc_rho_truncate / tsc_truncation_operator:
  // rho (allocation, upload or build), eigensolves, eigenvalues to the host
  // and kept-state selection, kept eigenvectors to the host
  float rho_ms[4] = {0.f, 0.f, 0.f, 0.f};
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
  void *d_part32 = nullptr;            // partial tiles, FP32 path
  void *d_part64 = nullptr;            // partial tiles, FP64 64x64 path
  int f64_kernel = 0;                  // FP64 kernel: 0 the 128x64 LDS units (tsc_gd.hip), 1 the 64x64 grouped tiles
  double flops = 0.0;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool timed = false;
  // graph cache: one per precision, keyed by the in/out pointers captured
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  const void *g_in[3] = {nullptr, nullptr, nullptr};
  const void *g_out[3] = {nullptr, nullptr, nullptr};
  // host copies of the task arrays (the packed path is built from them on first use)
  std::vector<tsc_contrib> h_c[2];
  std::vector<tsc_outblock> h_o[2];
  tsc::PkState *pk = nullptr;
  tsc::GdState *gd = nullptr;
  // Lanczos Krylov vectors (kry_n elements each), kept between calls
  std::vector<double2 *> kry;
  int64_t kry_n = 0;
  // Lanczos H.psi replays: graphs captured on two fixed vectors (lz_io,
  // kry_n elements each) per precision; Krylov vectors are copied in and out
  double2 *lz_io[2] = {nullptr, nullptr};
  hipGraphExec_t lz_gexec[3] = {nullptr, nullptr, nullptr};
  int lz_replays = 0;  // H.psi products of the last tsc_lanczos call that ran as graph replays
  // last tsc_lanczos call: working-precision Ritz value, and of its Ritz
  // vector the FP64 <psi|H psi> and |H psi - <psi|H psi> psi| from the
  // refinement pass's first product, the refined energy, and the weight of
  // the refined vector outside psi (NaN without a refinement pass)
  double lz_refine[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
};

namespace tsc {
int plan_create_impl(tsc_ctx *ctx, const double *ops, const double2 *ops_dev, int64_t ops_elems, int64_t mid_elems,
                     const tsc_contrib *c1, int32_t nc1, const tsc_outblock *o1, int32_t no1, const tsc_contrib *c2,
                     int32_t nc2, const tsc_outblock *o2, int32_t no2, int64_t in_elems, int64_t out_elems,
                     tsc_plan **out);
int hpsi_dev(tsc_plan *p, const double2 *in, double2 *out, int prec);
// packed BF16x3 path (tsc_pk.hip)
int pk_build(tsc_plan *p);         // build on first use; packs the operator arena
int pk_repack_ops(tsc_plan *p);    // after the arena changed
int pk_enqueue(tsc_plan *p, const double2 *in, double2 *out, bool events);
int gd_build(tsc_plan *p);  // FP64 units and contribution order, on first use
int gd_enqueue(tsc_plan *p, const double2 *in, double2 *out, bool events);
// cut the work units of a built path again for another make_tiles target
int pk_retile(tsc_plan *p, double desired);
int gd_retile(tsc_plan *p, double desired);
// work units of one stage (shared by all paths)
struct Units {
  std::vector<Tile> tiles;
  std::vector<Reduce> reds;
  int64_t nslot = 0;
};
// include: optional per-block mask (blocks outside it get no units)
Units make_tiles(const tsc_contrib *c, const tsc_outblock *o, int32_t no, int tm = kTM, int tn = kTN,
                 double desired_units = 6144.0, const std::vector<char> *include = nullptr);
double tridiag_lowest(const double *d, const double *e, int n);
void tridiag_vector(const double *d, const double *e, int n, double lam, double *z);
}  // namespace tsc
