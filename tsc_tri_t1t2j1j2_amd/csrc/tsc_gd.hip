// FP64 grouped contraction with operands moved straight into LDS (gfx950).
//
// Same sums as the other paths: every contribution C += coef * op(A) op(B)
// of model_wave's kernels (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:
// 1817, 1869, 2555-2560, 3385-3408, ...), in complex FP64 with
// v_mfma_f64_16x16x4f64. A complex product takes three real MFMAs instead
// of four (Gauss's product: with op(A) = ar + i ai, op(B) = br + i bi,
// P = ar br, Q = ai bi, S = (ar + ai)(br + bi), the product is
// (P - Q) + i (S - P - Q)); the accumulators hold P, Q and S.
//
//  * A workgroup owns one GM x GN output unit (128 x 64); each wave computes
//    a (16 NG) x 32 part of it as NG x 2 blocks of 16x16.
//  * The operands are read where they live (operator arena, psi, stage-1
//    intermediates: column-major complex128, any of the ops N/T/C/CONJ).
//    A step of GK k is NP pieces of 1 KB; a piece is one MFMA operand in
//    lane order, (16 rows or columns) x (4 k), lane l holding element
//    (16 blk + l%16, 4 chunk + l/16) as (re, im). Lane l of a loading wave
//    fetches exactly that element, so the direct global->LDS load (16 bytes
//    per lane) lays the piece out in fragment order for either storage order
//    of the operand: no repacking pass.
//  * Two step buffers: the loads of step s+1 are in flight while the MFMAs
//    of step s run; one barrier per step.
//  * Rows / columns past the block and k past a contribution are clamped to
//    the last one; the A values of k past the end are zeroed in registers,
//    and 16-row / 16-column blocks without output skip their MFMAs.
//  * Each block's GEMM terms are ordered by coefficient on the host; the
//    accumulators hold (partial sum) / (current coefficient) and are rescaled
//    when it changes. Conjugation is a sign flip of the A fragments.
//  * Elementwise terms (an identity operand) are added after the products.
#include <algorithm>
#include <numeric>

#include "tsc_internal.h"

namespace tsc {

namespace {

// Shape (measured at YC6 M=6000, FP64 H.psi: this 8-wave 128x64 unit with
// 8-k steps, two workgroups and four waves per SIMD, 52.0 ms; 4 waves of
// 64x32 parts 53.1 ms; 128x128 units, one workgroup per CU, 60-64 ms; the
// 4x4x4 MFMA form and a three-buffer ring were not faster).
constexpr int NG = 2;              // 16-row groups per wave
constexpr int WI = 8 / NG;         // waves along the 128 rows
constexpr int WJ = 2;              // waves along the 64 columns (32 each)
constexpr int NW = WI * WJ;        // waves per workgroup
constexpr int GM = kGdTile;        // 128
constexpr int GN = 32 * WJ;        // 64
constexpr int GK = 8;              // k per step
constexpr int WAVES_PER_SIMD = 4;
constexpr int NC = GK / 4;         // chunks per step
constexpr int NPA = 8 * NC;        // A pieces per step: 8 row blocks x NC chunks
constexpr int NBJ = GN / 16;       // column blocks
constexpr int NP = NPA + NBJ * NC; // pieces per step
constexpr int PPW = NP / NW;       // pieces loaded by each wave
static_assert(NP % NW == 0, "pieces must divide among the waves");

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double2 Step[NP][64];  // the pieces of one step, 64 lanes each

__device__ __forceinline__ bool gd_is_gemm(const tsc_contrib &c) {
  return c.a_op != TSC_OP_I && c.b_op != TSC_OP_I && c.k > 0 && (c.coef_re != 0.0 || c.coef_im != 0.0);
}

// base pointer of buffer `buf` (constant indices into the argument struct only)
__device__ __forceinline__ const double2 *gd_buf(const Bufs &B, int buf) {
  const void *p = buf == TSC_BUF_OPS ? B.p[TSC_BUF_OPS]
                  : buf == TSC_BUF_IN ? B.p[TSC_BUF_IN]
                  : buf == TSC_BUF_MID ? B.p[TSC_BUF_MID]
                                       : B.p[TSC_BUF_OUT];
  return reinterpret_cast<const double2 *>(p);
}

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// A contribution record read as whole dwords (scalar loads): a field-wise
// read of its 8-bit fields can become a vector load, and waiting for that
// would also wait for the operand loads in flight.
__device__ __forceinline__ tsc_contrib load_contrib(const tsc_contrib *p) {
  static_assert(sizeof(tsc_contrib) == 48, "contribution record size");
  const uint32_t *w = reinterpret_cast<const uint32_t *>(p);
  uint32_t x[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) x[i] = w[i];
#pragma unroll
  for (int i = 0; i < 12; ++i) __asm__ __volatile__("" : "+s"(x[i]));
  return __builtin_bit_cast(tsc_contrib, x);
}

__global__ __launch_bounds__(64 * NW, WAVES_PER_SIMD) void gd_kernel(Bufs bufs, const tsc_outblock *__restrict__ blocks,
                                                                 const tsc_contrib *__restrict__ contribs,
                                                                 const Tile *__restrict__ tiles, int rowmode,
                                                                 double2 *__restrict__ part) {
  __shared__ Step lds0, lds1;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int wa = wave % WI, wb = wave / WI;
  const Tile t = tiles[blockIdx.x];
  if (t.blk < 0) return;  // padding unit
  const tsc_outblock ob = blocks[t.blk];
  const int m = ob.m, n = ob.n;
  const int i0 = t.ti * GM, j0 = t.tj * GN;
  const int iw = i0 + 16 * NG * wa, jw = j0 + 32 * wb;
  // 16-row groups and 16-column blocks of this wave that hold output
  const int na = max(0, min(NG, (m - iw + 15) >> 4));
  const int nb = max(0, min(2, (n - jw + 15) >> 4));
  const int lrow = lane & 15, lk = lane >> 4;

  // accumulator block (g, b), register r of lane l: row 16g + orow(r),
  // column 16b + l%16 of the wave's part; the row layout of the 16x16x4
  // accumulator is measured once per context (rowmode)
  auto orow = [&](int r) { return rowmode ? (lk + 4 * r) : (4 * lk + r); };
  d4 aP[NG][2], aQ[NG][2], aS[NG][2];
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      aP[g][b] = d4{0.0, 0.0, 0.0, 0.0};
      aQ[g][b] = d4{0.0, 0.0, 0.0, 0.0};
      aS[g][b] = d4{0.0, 0.0, 0.0, 0.0};
    }

  // piece q of a step: q < NPA: A, chunk q/8, row block q%8; else B, chunk
  // (q-NPA)/NBJ, column block (q-NPA)%NBJ. Wave w loads pieces
  // [w PPW, (w+1) PPW).
  auto issue = [&](const tsc_contrib &c, int k0, Step &L) {
#pragma unroll
    for (int i = 0; i < PPW; ++i) {
      const int q = wave * PPW + i;
      const bool isA = q < NPA;
      const int qq = isA ? q : q - NPA;
      const int ch = isA ? (qq >> 3) : (qq / NBJ);
      const int blk = isA ? (qq & 7) : (qq % NBJ);
      const int lim = isA ? m : n;
      const int rc = (isA ? i0 : j0) + 16 * blk;
      if (rc >= lim) continue;  // nothing of this block: its MFMAs are skipped
      const int kk = min(k0 + 4 * ch + lk, c.k - 1);
      const int op = isA ? c.a_op : c.b_op;
      const int ld = isA ? c.lda : c.ldb;
      const bool tr = (op == TSC_OP_T || op == TSC_OP_C);
      // A[r][k] = X[r + k ld] (N, CONJ) or X[k + r ld] (T, C);
      // B[k][j] = Y[k + j ld] (N, CONJ) or Y[j + k ld] (T, C)
      const bool rfast = isA ? !tr : tr;
      const double2 *X = gd_buf(bufs, isA ? c.a_buf : c.b_buf) + (isA ? c.a_off : c.b_off);
      const int r = min(rc + lrow, lim - 1);
      const double2 *src = rfast ? X + r + (int64_t)kk * ld : X + kk + (int64_t)r * ld;
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                       (__attribute__((address_space(3))) void *)&L[q][0], 16, 0, 0);
    }
  };

  // op(A) = ar + i sA ai, op(B) = br + i sB bi (s = -1 for a conjugated
  // operand): P += ar br, Q += (sA ai)(sB bi), S += (ar + sA ai)(br + sB bi)
  auto compute = [&](const tsc_contrib &c, int k0, const Step &L) {
    if (na == 0 || nb == 0) return;
    const int nch = min(NC, (c.k - k0 + 3) >> 2);
    const bool cja = (c.a_op == TSC_OP_C || c.a_op == TSC_OP_CONJ);
    const bool cjb = (c.b_op == TSC_OP_C || c.b_op == TSC_OP_CONJ);
    const uint64_t sgn = 0x8000000000000000ull;
    const uint64_t ma = cja ? sgn : 0, mb = cjb ? sgn : 0;
    auto flip = [](double x, uint64_t msk) {
      return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ msk);
    };
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
      if (ch >= nch) break;
      const bool kz = k0 + 4 * ch + lk >= c.k;  // k past the contribution: A is zero
      // one B fragment at a time, the A fragments read again for each (LDS
      // reads are cheap next to FP64 MFMAs; this keeps registers for the
      // three accumulator sets)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (b >= nb) break;
        const double2 bv = L[NPA + NBJ * ch + 2 * wb + b][lane];
        const double br = bv.x, bi = flip(bv.y, mb), bs = br + bi;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          if (g >= na) break;
          double2 v = L[8 * ch + NG * wa + g][lane];
          if (kz) v = make_double2(0.0, 0.0);
          const double ar = v.x, ai = flip(v.y, ma), as = ar + ai;
          aP[g][b] = mfma16(ar, br, aP[g][b]);
          aQ[g][b] = mfma16(ai, bi, aQ[g][b]);
          aS[g][b] = mfma16(as, bs, aS[g][b]);
        }
      }
    }
  };

  // multiply the accumulated sum by sr + i si: (P, Q, S) becomes
  // (Re, 0, Re + Im) of the scaled product
  auto rescale = [&](double sr, double si) {
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const d4 x = aP[g][b] - aQ[g][b], y = aS[g][b] - aP[g][b] - aQ[g][b];
        const d4 re = x * sr - y * si, im = x * si + y * sr;
        aP[g][b] = re;
        aQ[g][b] = d4{0.0, 0.0, 0.0, 0.0};
        aS[g][b] = re + im;
      }
  };
  // accumulators hold (sum) / (gr + i gi)
  double gr = 0.0, gi = 0.0;
  auto set_coef = [&](const tsc_contrib &c) {
    if (c.coef_re == gr && c.coef_im == gi) return;
    if (gr != 0.0 || gi != 0.0) {
      const double den = c.coef_re * c.coef_re + c.coef_im * c.coef_im;
      rescale((gr * c.coef_re + gi * c.coef_im) / den, (gi * c.coef_re - gr * c.coef_im) / den);
    }
    gr = c.coef_re;
    gi = c.coef_im;
  };

  // own loads landed, then every wave's
  auto step_sync = [&]() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
  };

  // The steps of a unit: k0 = 0, GK, ... of each GEMM term in its range (the
  // host orders each block's contributions GEMMs first, so these are a
  // prefix). Positions are (contribution index, k0); records are re-read
  // (scalar loads) where needed. next_step moves a position to the next
  // step and returns true, or returns false after the last one. The loop
  // body is written out once per buffer so that every LDS access names its
  // array.
  auto gemm_at = [&](int ci) { return ci < t.cend && gd_is_gemm(load_contrib(contribs + ci)); };
  auto next_step = [&](int &ci, int &kk) -> bool {
    kk += GK;
    if (kk < load_contrib(contribs + ci).k) return true;
    ++ci;
    kk = 0;
    return gemm_at(ci);
  };
  int cc = t.cbeg, k0 = 0;
  bool live = gemm_at(cc);
  if (live) issue(load_contrib(contribs + cc), 0, lds0);
  step_sync();
  auto step = [&](const Step &X, Step &Z) -> bool {
    int nc = cc, nk0 = k0;
    const bool more = next_step(nc, nk0);
    if (more) issue(load_contrib(contribs + nc), nk0, Z);
    const tsc_contrib cur = load_contrib(contribs + cc);
    if (k0 == 0) set_coef(cur);
    compute(cur, k0, X);
    step_sync();
    cc = nc;
    k0 = nk0;
    return more;
  };
  while (live) {
    if (!step(lds0, lds1)) break;
    live = step(lds1, lds0);
  }
  if (gr != 0.0 || gi != 0.0) rescale(gr, gi);
  // from here on aP holds the real part and aS the imaginary part
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const d4 re = aP[g][b] - aQ[g][b];
      aS[g][b] = aS[g][b] - aP[g][b] - aQ[g][b];
      aP[g][b] = re;
    }
  d4(&aR)[NG][2] = aP;
  d4(&aI)[NG][2] = aS;

  // elementwise terms (one operand the identity)
  for (int q = t.cbeg; q < t.cend; ++q) {
    const tsc_contrib c = contribs[q];
    if (c.a_op != TSC_OP_I && c.b_op != TSC_OP_I) continue;
    const bool useA = (c.b_op == TSC_OP_I);
    const int op = useA ? c.a_op : c.b_op;
    if (op == TSC_OP_I) continue;
    const double2 *X = gd_buf(bufs, useA ? c.a_buf : c.b_buf) + (useA ? c.a_off : c.b_off);
    const int ldx = useA ? c.lda : c.ldb;
    const bool tr = (op == TSC_OP_T || op == TSC_OP_C);
    const bool cj = (op == TSC_OP_C || op == TSC_OP_CONJ);
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = iw + 16 * g + orow(r), j = jw + 16 * b + lrow;
          if (i < m && j < n) {
            const double2 x = tr ? X[(int64_t)j + (int64_t)i * ldx] : X[(int64_t)i + (int64_t)j * ldx];
            const double xr = x.x, xi = cj ? -x.y : x.y;
            aR[g][b][r] += c.coef_re * xr - c.coef_im * xi;
            aI[g][b][r] += c.coef_re * xi + c.coef_im * xr;
          }
        }
  }

  // write the wave's part of the unit: to C, or as a partial tile
  double2 *C = const_cast<double2 *>(gd_buf(bufs, ob.c_buf));
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = 16 * NG * wa + 16 * g + orow(r), jl = 32 * wb + 16 * b + lrow;
        const int i = i0 + il, j = j0 + jl;
        if (i >= m || j >= n) continue;
        double2 v = make_double2(aR[g][b][r], aI[g][b][r]);
        if (t.slot >= 0) {
          part[(int64_t)t.slot * (GM * GN) + (int64_t)jl * GM + il] = v;
          continue;
        }
        const int64_t idx = ob.c_off + (int64_t)i + (int64_t)j * ob.ldc;
        if (ob.mode) {
          const double2 o = C[idx];
          v.x += o.x;
          v.y += o.y;
        }
        C[idx] = v;
      }
}

// Sum the partial tiles of split units in slot order and write C.
__global__ __launch_bounds__(256) void gd_reduce_kernel(Bufs bufs, const tsc_outblock *__restrict__ blocks,
                                                        const Reduce *__restrict__ reds,
                                                        const double2 *__restrict__ part) {
  const Reduce rd = reds[blockIdx.x];
  const tsc_outblock ob = blocks[rd.blk];
  const int i0 = rd.ti * GM, j0 = rd.tj * GN;
  double2 *C = const_cast<double2 *>(gd_buf(bufs, ob.c_buf));
  for (int e = threadIdx.x; e < GM * GN; e += 256) {
    const int il = e % GM, jl = e / GM;
    const int i = i0 + il, j = j0 + jl;
    if (i >= ob.m || j >= ob.n) continue;
    double sr = 0.0, si = 0.0;
    for (int s = 0; s < rd.nslot; ++s) {
      const double2 v = part[(int64_t)(rd.slot0 + s) * (GM * GN) + e];
      sr += v.x;
      si += v.y;
    }
    const int64_t idx = ob.c_off + (int64_t)i + (int64_t)j * ob.ldc;
    double2 v = make_double2(sr, si);
    if (ob.mode) {
      const double2 o = C[idx];
      v.x += o.x;
      v.y += o.y;
    }
    C[idx] = v;
  }
}

template <typename T>
int gd_upload(T **dst, const std::vector<T> &v, hipStream_t s) {
  *dst = nullptr;
  if (v.empty()) return TSC_OK;
  TSC_HIP_CHECK(hipMalloc((void **)dst, sizeof(T) * v.size()));
  TSC_HIP_CHECK(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return TSC_OK;
}

int gd_build_impl(tsc_plan *p, GdState *S) {
  hipStream_t st = p->ctx->stream;
  int64_t slots = 0;
  for (int s = 0; s < 2; ++s) {
    const std::vector<tsc_contrib> &C = p->h_c[s];
    const std::vector<tsc_outblock> &O = p->h_o[s];
    // each block's contributions: GEMMs with work first, then elementwise
    // terms, then terms with nothing to do (original order within a class)
    std::vector<tsc_contrib> cs(C.begin(), C.end());
    for (const tsc_outblock &ob : O) {
      auto cls = [&](const tsc_contrib &c) {
        if (c.a_op == TSC_OP_I || c.b_op == TSC_OP_I) return 1;
        if (c.k <= 0 || (c.coef_re == 0.0 && c.coef_im == 0.0) || ob.m <= 0 || ob.n <= 0) return 2;
        return 0;
      };
      // GEMM terms by coefficient: the kernel rescales its accumulators
      // only when the coefficient changes
      std::stable_sort(cs.begin() + ob.cbeg, cs.begin() + ob.cend, [&](const tsc_contrib &x, const tsc_contrib &y) {
        const int cx = cls(x), cy = cls(y);
        if (cx != cy) return cx < cy;
        if (cx != 0) return false;
        if (x.coef_re != y.coef_re) return x.coef_re < y.coef_re;
        return x.coef_im < y.coef_im;
      });
    }
    Units un = make_tiles(cs.data(), O.data(), (int32_t)O.size(), GM, GN, S->desired);
    GdStage &G = S->st[s];
    G.nt = (int)un.tiles.size();
    G.nr = (int)un.reds.size();
    slots = std::max(slots, un.nslot);
    int rc;
    if ((rc = gd_upload(&G.d_c, cs, st))) return rc;
    if ((rc = gd_upload(&G.d_t, un.tiles, st))) return rc;
    if ((rc = gd_upload(&G.d_r, un.reds, st))) return rc;
    G.h_c.swap(cs);
  }
  if (slots > 0) {
    hipError_t e = hipMalloc((void **)&S->d_part, sizeof(double2) * (size_t)GM * GN * (size_t)slots);
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc(FP64 partial tiles) failed: ") + hipGetErrorString(e));
      return TSC_ENOMEM;
    }
    S->part_slots = slots;
  }
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  return TSC_OK;
}

}  // namespace

void gd_free(GdState *S) {
  if (!S) return;
  for (int s = 0; s < 2; ++s) {
    hipFree(S->st[s].d_c);
    hipFree(S->st[s].d_t);
    hipFree(S->st[s].d_r);
  }
  hipFree(S->d_part);
  delete S;
}

int gd_build(tsc_plan *p) {
  if (p->gd) return TSC_OK;
  GdState *S = new GdState();
  const int rc = gd_build_impl(p, S);
  if (rc) {
    gd_free(S);
    return rc;
  }
  p->gd = S;
  return TSC_OK;
}

int gd_retile(tsc_plan *p, double desired) {
  GdState *S = p->gd;
  if (!S) {
    set_error("FP64 path re-cut before it was built");
    return TSC_ESTATE;
  }
  hipStream_t st = p->ctx->stream;
  // the current units may still be in use by queued launches
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  int64_t slots = 0;
  for (int s = 0; s < 2; ++s) {
    GdStage &G = S->st[s];
    const std::vector<tsc_outblock> &O = p->h_o[s];
    Units un = make_tiles(G.h_c.data(), O.data(), (int32_t)O.size(), GM, GN, desired);
    hipFree(G.d_t);
    hipFree(G.d_r);
    G.d_t = nullptr;
    G.d_r = nullptr;
    G.nt = (int)un.tiles.size();
    G.nr = (int)un.reds.size();
    slots = std::max(slots, un.nslot);
    int rc;
    if ((rc = gd_upload(&G.d_t, un.tiles, st))) return rc;
    if ((rc = gd_upload(&G.d_r, un.reds, st))) return rc;
  }
  if (slots > S->part_slots) {
    hipFree(S->d_part);
    S->d_part = nullptr;
    S->part_slots = 0;
    hipError_t e = hipMalloc((void **)&S->d_part, sizeof(double2) * (size_t)GM * GN * (size_t)slots);
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc(FP64 partial tiles) failed: ") + hipGetErrorString(e));
      return TSC_ENOMEM;
    }
    S->part_slots = slots;
  }
  S->desired = desired;
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  return TSC_OK;
}

int gd_enqueue(tsc_plan *p, const double2 *in, double2 *out, bool events) {
  GdState *S = p->gd;
  if (!S) {
    set_error("FP64 path used before it was built");
    return TSC_ESTATE;
  }
  hipStream_t s = p->ctx->stream;
  Bufs b;
  b.p[TSC_BUF_OPS] = p->d_ops64;
  b.p[TSC_BUF_IN] = in;
  b.p[TSC_BUF_MID] = p->d_mid64;
  b.p[TSC_BUF_OUT] = out;
  if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[0], s));
  for (int g = 0; g < 2; ++g) {
    const GdStage &G = S->st[g];
    if (G.nt > 0) {
      hipLaunchKernelGGL(gd_kernel, dim3(G.nt), dim3(64 * NW), 0, s, b, p->d_o[g], G.d_c, G.d_t, p->ctx->mfma_rowmode[0],
                         S->d_part);
      TSC_HIP_CHECK(hipGetLastError());
    }
    if (G.nr > 0) {
      hipLaunchKernelGGL(gd_reduce_kernel, dim3(G.nr), dim3(256), 0, s, b, p->d_o[g], G.d_r,
                         (const double2 *)S->d_part);
      TSC_HIP_CHECK(hipGetLastError());
    }
    if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[g + 1], s));
  }
  return TSC_OK;
}

}  // namespace tsc
