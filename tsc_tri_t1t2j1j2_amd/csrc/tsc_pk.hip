// Packed-operand BF16x3 path of the grouped contraction (gfx950).
//
// The FP64 path evaluates every contribution C += coef * op(A) op(B) of
// model_wave's kernels (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:
// 1817, 1869, 2555-2560, 3385-3408, ...) with FP64 MFMAs. This path does the
// same sums with v_mfma_f32_32x32x16_bf16 on bf16 hi/lo splits of the FP32
// operands (a*b ~ ah*bh + ah*bl + al*bh, FP32 accumulation):
//
//  * every operand is read from a pre-split panel (layout in tsc_internal.h):
//    the operator arena is packed once per plan, the psi windows once per
//    H.psi call, and stage 1 writes its intermediates directly as panels;
//    the contraction kernel therefore does no conversion work and loads each
//    MFMA fragment with one 16-byte load straight into registers;
//  * one wave owns one 64x64 output unit as 2x2 blocks of 32x32 MFMAs
//    (no LDS, no barriers); the MFMA is issued with the B panel as its first
//    operand, so a lane's accumulators run along output columns and stores
//    of psi / panels are contiguous across lanes;
//  * the contributions of a unit are ordered by coefficient; the
//    accumulators hold (partial sum) / (current coefficient) and are rescaled
//    when the coefficient changes, so coefficients never touch the operands.
#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>

#include "tsc_internal.h"

namespace tsc {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int TM = kTM, TN = kTN;

__device__ __forceinline__ void split2(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

__device__ __forceinline__ bf8 neg8(bf8 x) {
  u4v u = __builtin_bit_cast(u4v, x);
  u ^= 0x80008000u;
  return __builtin_bit_cast(bf8, u);
}

// ---------------------------------------------------------------------------
// Packing: one thread per (row, chunk) of a panel. Row-operand panels
// (trans 0) walk rows fastest so the strided reads of a column-major window
// coalesce; transposed panels read 8 consecutive elements per thread.
__global__ __launch_bounds__(256) void pack_kernel(const double2 *__restrict__ src, bf8 *__restrict__ dst,
                                                   const PkPanel *__restrict__ panels,
                                                   const PkJob *__restrict__ jobs) {
  const PkJob jb = jobs[blockIdx.x];
  const PkPanel P = panels[jb.panel];
  const int64_t t = jb.first + threadIdx.x;
  if (t >= (int64_t)P.rows * P.nch) return;
  int r, c;
  if (P.trans == 0) {
    c = (int)(t / P.rows);
    r = (int)(t % P.rows);
  } else {
    r = (int)(t / P.nch);
    c = (int)(t % P.nch);
  }
  const double2 *X = src + P.src_off;
  bf8 v0, v1, v2, v3;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * c + e;
    double2 x = make_double2(0.0, 0.0);
    if (k < P.kdim) x = P.trans ? X[(int64_t)k + (int64_t)r * P.ld] : X[(int64_t)r + (int64_t)k * P.ld];
    __bf16 a, b, cc, d;
    split2((float)x.x, a, b);
    split2((float)x.y, cc, d);
    v0[e] = a;
    v2[e] = b;
    v1[e] = cc;
    v3[e] = d;
  }
  bf8 *q = dst + P.dst + ((int64_t)r * P.nch + c) * 4;
  q[0] = v0;
  q[1] = v1;
  q[2] = v2;
  q[3] = v3;
}

// ---------------------------------------------------------------------------
// element helpers for packed intermediates
__device__ __forceinline__ float unsplit(__bf16 hi, __bf16 lo) { return (float)hi + (float)lo; }

// one element (plane position `pos` of the chunk at 16-byte index idx)
__device__ __forceinline__ void store_pk_elem(bf8 *mid, int64_t idx, int pos, float re, float im, bool acc) {
  __bf16 *q = reinterpret_cast<__bf16 *>(mid + idx);
  if (acc) {
    re += unsplit(q[pos], q[16 + pos]);
    im += unsplit(q[8 + pos], q[24 + pos]);
  }
  __bf16 a, b, c, d;
  split2(re, a, b);
  split2(im, c, d);
  q[pos] = a;
  q[8 + pos] = c;
  q[16 + pos] = b;
  q[24 + pos] = d;
}

struct PkPtrs {
  const bf8 *p[3];
};

// op(X)[r][c] of a complex128 window (elementwise contributions)
__device__ __forceinline__ double2 ld_op64(const double2 *X, int64_t off, int ld, int op, int r, int c) {
  const bool tr = (op == TSC_OP_T || op == TSC_OP_C);
  double2 v = tr ? X[off + (int64_t)c + (int64_t)r * ld] : X[off + (int64_t)r + (int64_t)c * ld];
  if (op == TSC_OP_C || op == TSC_OP_CONJ) v.y = -v.y;
  return v;
}

__device__ __forceinline__ void scale_acc(f16v (&R)[2][2], f16v (&I)[2][2], float sr, float si) {
  if (si == 0.f) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        R[a][b] *= sr;
        I[a][b] *= sr;
      }
  } else {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const f16v x = R[a][b], y = I[a][b];
        R[a][b] = x * sr - y * si;
        I[a][b] = x * si + y * sr;
      }
  }
}

#define MFMA32(x, y, acc) __builtin_amdgcn_mfma_f32_32x32x16_bf16((x), (y), (acc), 0, 0, 0)

// ---------------------------------------------------------------------------
// Contraction kernel: one wave, one 64x64 unit. Accumulator element r of
// block (a, b) held by `lane` is output (i0 + 32a + lane%32,
// j0 + 32b + 8(r/4) + 4(lane/32) + r%4).
__global__ __launch_bounds__(64, 2) void pk_gemm_kernel(PkPtrs P, bf8 *__restrict__ mid, const double2 *__restrict__ in64,
                                                     const double2 *__restrict__ ops64, double2 *__restrict__ out64,
                                                     const tsc_outblock *__restrict__ blocks,
                                                     const PkOut *__restrict__ pouts,
                                                     const tsc_contrib *__restrict__ cps,
                                                     const PkContrib *__restrict__ pcs,
                                                     const Tile *__restrict__ tiles, float2 *__restrict__ part) {
  const int lane = threadIdx.x, l32 = lane & 31, h = lane >> 5;
  const Tile t = tiles[blockIdx.x];
  const tsc_outblock ob = blocks[t.blk];
  const int i0 = t.ti * TM, j0 = t.tj * TN;
  const int m = ob.m, n = ob.n;

  f16v accR[2][2], accI[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        accR[a][b][r] = 0.f;
        accI[a][b][r] = 0.f;
      }

  double g_re = 1.0, g_im = 0.0;  // accumulators hold sum / g
  for (int ci = t.cbeg; ci < t.cend; ++ci) {
    const PkContrib c = pcs[ci];
    if (c.ksteps == 0) continue;
    if (c.coef_re != g_re || c.coef_im != g_im) {
      const double den = c.coef_re * c.coef_re + c.coef_im * c.coef_im;
      const double rr = (g_re * c.coef_re + g_im * c.coef_im) / den;
      const double ri = (g_im * c.coef_re - g_re * c.coef_im) / den;
      scale_acc(accR, accI, (float)rr, (float)ri);
      g_re = c.coef_re;
      g_im = c.coef_im;
    }
    const bf8 *pa = P.p[(c.flags >> 4) & 3] + c.a_off;
    const bf8 *pb = P.p[(c.flags >> 6) & 3] + c.b_off;
    const bf8 *ra[2], *rb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) ra[a] = pa + ((int64_t)min(i0 + 32 * a + l32, c.a_rmax) * c.a_nch + h) * 4;
#pragma unroll
    for (int b = 0; b < 2; ++b) rb[b] = pb + ((int64_t)min(j0 + 32 * b + l32, c.b_rmax) * c.b_nch + h) * 4;
    const bool cja = (c.flags & 1) != 0, cjb = (c.flags & 2) != 0;
    for (int s = 0; s < c.ksteps; ++s) {
      bf8 A[2][4], B[2][4];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) A[a][q] = ra[a][8 * s + q];
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) B[b][q] = rb[b][8 * s + q];
      if (cja) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          A[a][1] = neg8(A[a][1]);
          A[a][3] = neg8(A[a][3]);
        }
      }
      if (cjb) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          B[b][1] = neg8(B[b][1]);
          B[b][3] = neg8(B[b][3]);
        }
      }
      bf8 NB[2][2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        NB[b][0] = neg8(B[b][1]);
        NB[b][1] = neg8(B[b][3]);
      }
      // planes: 0 re hi, 1 im hi, 2 re lo, 3 im lo
      // Re += br*ar - bi*ai ; Im += bi*ar + br*ai (each product in three terms)
#define BLOCKS(stmt)                    \
  _Pragma("unroll") for (int a = 0; a < 2; ++a) \
      _Pragma("unroll") for (int b = 0; b < 2; ++b) stmt
      BLOCKS(accR[a][b] = MFMA32(B[b][0], A[a][0], accR[a][b]));
      BLOCKS(accI[a][b] = MFMA32(B[b][1], A[a][0], accI[a][b]));
      BLOCKS(accR[a][b] = MFMA32(B[b][2], A[a][0], accR[a][b]));
      BLOCKS(accI[a][b] = MFMA32(B[b][3], A[a][0], accI[a][b]));
      BLOCKS(accR[a][b] = MFMA32(B[b][0], A[a][2], accR[a][b]));
      BLOCKS(accI[a][b] = MFMA32(B[b][1], A[a][2], accI[a][b]));
      BLOCKS(accR[a][b] = MFMA32(NB[b][0], A[a][1], accR[a][b]));
      BLOCKS(accI[a][b] = MFMA32(B[b][0], A[a][1], accI[a][b]));
      BLOCKS(accR[a][b] = MFMA32(NB[b][1], A[a][1], accR[a][b]));
      BLOCKS(accI[a][b] = MFMA32(B[b][2], A[a][1], accI[a][b]));
      BLOCKS(accR[a][b] = MFMA32(NB[b][0], A[a][3], accR[a][b]));
      BLOCKS(accI[a][b] = MFMA32(B[b][0], A[a][3], accI[a][b]));
#undef BLOCKS
    }
  }
  if (g_re != 1.0 || g_im != 0.0) scale_acc(accR, accI, (float)g_re, (float)g_im);

  // elementwise contributions (one operand the identity), FP32 accumulation
  for (int ci = t.cbeg; ci < t.cend; ++ci) {
    const tsc_contrib c = cps[ci];
    if (c.a_op != TSC_OP_I && c.b_op != TSC_OP_I) continue;
    const bool useA = (c.b_op == TSC_OP_I);
    const int buf = useA ? c.a_buf : c.b_buf;
    const int op = useA ? c.a_op : c.b_op;
    if (op == TSC_OP_I) continue;
    const double2 *X = (buf == TSC_BUF_OPS) ? ops64 : in64;
    const int64_t off = useA ? c.a_off : c.b_off;
    const int ldx = useA ? c.lda : c.ldb;
    const float cre = (float)c.coef_re, cim = (float)c.coef_im;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = i0 + 32 * a + l32;
          const int j = j0 + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (i < m && j < n) {
            const double2 x = ld_op64(X, off, ldx, op, i, j);
            const float xr = (float)x.x, xi = (float)x.y;
            accR[a][b][r] += cre * xr - cim * xi;
            accI[a][b][r] += cre * xi + cim * xr;
          }
        }
  }

  // ---- write the unit
  if (t.slot >= 0) {
    float2 *Pp = part + (int64_t)t.slot * (TM * TN);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int il = 32 * a + l32, jl = 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (i0 + il < m && j0 + jl < n) Pp[jl * TM + il] = make_float2(accR[a][b][r], accI[a][b][r]);
        }
    return;
  }
  if (ob.c_buf != TSC_BUF_MID) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = i0 + 32 * a + l32, j = j0 + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (i >= m || j >= n) continue;
          const int64_t idx = ob.c_off + (int64_t)i + (int64_t)j * ob.ldc;
          double2 v = make_double2((double)accR[a][b][r], (double)accI[a][b][r]);
          if (ob.mode) {
            const double2 o = out64[idx];
            v.x += o.x;
            v.y += o.y;
          }
          out64[idx] = v;
        }
    return;
  }
  const PkOut po = pouts[t.blk];
  if (po.r_off >= 0) {
    // row panel: P[i][j]; the 4 consecutive j of (a, b, q) are half a chunk
    const int kp = 8 * po.r_nch;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = i0 + 32 * a + l32;
          const int jb = j0 + 32 * b + 8 * q + 4 * h;
          if (i >= m || jb >= kp) continue;
          bf4v p0, p1, p2, p3;
          bf4v *dst = reinterpret_cast<bf4v *>(mid + po.r_off + ((int64_t)i * po.r_nch + (jb >> 3)) * 4) + h;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float vr = (jb + e < n) ? accR[a][b][4 * q + e] : 0.f;
            float vi = (jb + e < n) ? accI[a][b][4 * q + e] : 0.f;
            if (ob.mode) {
              const bf4v o0 = dst[0], o1 = dst[2], o2 = dst[4], o3 = dst[6];
              vr += unsplit(o0[e], o2[e]);
              vi += unsplit(o1[e], o3[e]);
            }
            __bf16 x0, x1, x2, x3;
            split2(vr, x0, x2);
            split2(vi, x1, x3);
            p0[e] = x0;
            p1[e] = x1;
            p2[e] = x2;
            p3[e] = x3;
          }
          // plane stride is 16 bytes = 2 bf4v
          dst[0] = p0;
          dst[2] = p1;
          dst[4] = p2;
          dst[6] = p3;
        }
  }
  if (po.c_off >= 0) {
    // transposed panel: P[j][i]
    const int kp = 8 * po.c_nch;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = i0 + 32 * a + l32, j = j0 + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (j >= n || i >= kp) continue;
          const bool in_blk = i < m;
          store_pk_elem(mid, po.c_off + ((int64_t)j * po.c_nch + (i >> 3)) * 4, i & 7,
                        in_blk ? accR[a][b][r] : 0.f, in_blk ? accI[a][b][r] : 0.f, ob.mode != 0);
        }
  }
}

// Sum the partial tiles of split units in slot order (FP64 sums) and write
// psi or the packed intermediate, padding included.
__global__ __launch_bounds__(256) void pk_reduce_kernel(bf8 *__restrict__ mid, double2 *__restrict__ out64,
                                                        const tsc_outblock *__restrict__ blocks,
                                                        const PkOut *__restrict__ pouts,
                                                        const Reduce *__restrict__ reds,
                                                        const float2 *__restrict__ part) {
  const Reduce rd = reds[blockIdx.x];
  const tsc_outblock ob = blocks[rd.blk];
  const int i0 = rd.ti * TM, j0 = rd.tj * TN;
  const bool packed = (ob.c_buf == TSC_BUF_MID);
  PkOut po{-1, -1, 0, 0};
  if (packed) po = pouts[rd.blk];
  for (int e = threadIdx.x; e < TM * TN; e += 256) {
    const int il = e % TM, jl = e / TM;
    const int i = i0 + il, j = j0 + jl;
    const bool inside = (i < ob.m && j < ob.n);
    double sr = 0.0, si = 0.0;
    if (inside) {
      for (int s = 0; s < rd.nslot; ++s) {
        const float2 v = part[(int64_t)(rd.slot0 + s) * (TM * TN) + e];
        sr += (double)v.x;
        si += (double)v.y;
      }
    }
    if (!packed) {
      if (!inside) continue;
      const int64_t idx = ob.c_off + (int64_t)i + (int64_t)j * ob.ldc;
      double2 v = make_double2(sr, si);
      if (ob.mode) {
        const double2 o = out64[idx];
        v.x += o.x;
        v.y += o.y;
      }
      out64[idx] = v;
      continue;
    }
    if (po.r_off >= 0 && i < ob.m && j < 8 * po.r_nch)
      store_pk_elem(mid, po.r_off + ((int64_t)i * po.r_nch + (j >> 3)) * 4, j & 7, inside ? (float)sr : 0.f,
                    inside ? (float)si : 0.f, ob.mode != 0);
    if (po.c_off >= 0 && j < ob.n && i < 8 * po.c_nch)
      store_pk_elem(mid, po.c_off + ((int64_t)j * po.c_nch + (i >> 3)) * 4, i & 7, inside ? (float)sr : 0.f,
                    inside ? (float)si : 0.f, ob.mode != 0);
  }
}

// ---------------------------------------------------------------------------
// host side

template <typename T>
int upload_vec(T **dst, const std::vector<T> &v, hipStream_t s) {
  *dst = nullptr;
  if (v.empty()) return TSC_OK;
  TSC_HIP_CHECK(hipMalloc((void **)dst, sizeof(T) * v.size()));
  TSC_HIP_CHECK(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return TSC_OK;
}

int alloc_bytes(void **p, int64_t bytes, const char *what) {
  *p = nullptr;
  if (bytes <= 0) return TSC_OK;
  hipError_t e = hipMalloc(p, (size_t)bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + what + ") failed: " + hipGetErrorString(e));
    return TSC_ENOMEM;
  }
  return TSC_OK;
}

std::vector<PkJob> make_jobs(const std::vector<PkPanel> &pan) {
  std::vector<PkJob> jobs;
  for (size_t i = 0; i < pan.size(); ++i) {
    const int64_t total = (int64_t)pan[i].rows * pan[i].nch;
    for (int64_t f = 0; f < total; f += 256) jobs.push_back(PkJob{(int32_t)i, 0, f});
  }
  return jobs;
}

int launch_pack(const double2 *src, void *dst, const PkPanel *panels, const PkJob *jobs, int njobs, hipStream_t s) {
  if (njobs <= 0) return TSC_OK;
  hipLaunchKernelGGL(pack_kernel, dim3(njobs), dim3(256), 0, s, src, reinterpret_cast<bf8 *>(dst), panels, jobs);
  TSC_HIP_CHECK(hipGetLastError());
  return TSC_OK;
}

int32_t chunks_for(int k) { return 2 * ((k + 15) / 16); }

}  // namespace

void pk_free(PkState *S) {
  if (!S) return;
  hipFree(S->d_ops);
  hipFree(S->d_in);
  hipFree(S->d_mid);
  hipFree(S->d_ops_panels);
  hipFree(S->d_in_panels);
  hipFree(S->d_ops_jobs);
  hipFree(S->d_in_jobs);
  for (int s = 0; s < 2; ++s) {
    hipFree(S->st[s].d_pc);
    hipFree(S->st[s].d_cp);
    hipFree(S->st[s].d_po);
    hipFree(S->st[s].d_t);
    hipFree(S->st[s].d_r);
  }
  hipFree(S->d_part);
  delete S;
}

// Build the packed path of a plan: panels of every GEMM operand (deduplicated
// windows of the arena and of psi; intermediates as written by stage 1),
// contribution records ordered by coefficient within each output block,
// work units, device buffers; then pack the arena.
namespace {
int pk_build_impl(tsc_plan *p, PkState *S) {
  hipStream_t st = p->ctx->stream;
  auto unsupported = [&](const std::string &why) {
    S->unsupported = why;
    set_error("BF16x3 path: " + why);
    return TSC_EINVAL;
  };

  // panels of the arena (0) and of psi (1), deduplicated by window
  typedef std::tuple<int64_t, int, int, int, int> WinKey;  // off, ld, rows, kdim, trans
  std::map<WinKey, int64_t> win[2];
  std::vector<PkPanel> pan[2];
  int64_t plen[2] = {0, 0};
  auto panel = [&](int which, int64_t off, int ld, int rows, int kdim, int trans) -> int64_t {
    const WinKey key(off, ld, rows, kdim, trans);
    auto it = win[which].find(key);
    if (it != win[which].end()) return it->second;
    PkPanel P{};
    P.src_off = off;
    P.dst = plen[which];
    P.ld = ld;
    P.rows = rows;
    P.kdim = kdim;
    P.nch = chunks_for(kdim);
    P.trans = trans;
    plen[which] += (int64_t)rows * P.nch * 4;
    pan[which].push_back(P);
    win[which][key] = P.dst;
    return P.dst;
  };

  // intermediates written by stage 1, found by their offset
  const std::vector<tsc_outblock> &O1 = p->h_o[0];
  std::map<int64_t, int> mid_at;
  for (size_t b = 0; b < O1.size(); ++b) mid_at.emplace(O1[b].c_off, (int)b);
  std::vector<char> need_r(O1.size(), 0), need_c(O1.size(), 0);
  struct MidRef {
    int stage;
    int64_t pos;
    int operand;  // 0: A, 1: B
    int blk;
    int trans;
  };
  std::vector<MidRef> midrefs;

  std::vector<PkContrib> pcs[2];
  std::vector<tsc_contrib> cps[2];
  for (int s = 0; s < 2; ++s) {
    const std::vector<tsc_contrib> &C = p->h_c[s];
    const std::vector<tsc_outblock> &O = p->h_o[s];
    pcs[s].assign(C.size(), PkContrib{});
    cps[s].assign(C.begin(), C.end());
    for (size_t b = 0; b < O.size(); ++b) {
      const tsc_outblock &ob = O[b];
      std::vector<int32_t> order(ob.cend - ob.cbeg);
      std::iota(order.begin(), order.end(), ob.cbeg);
      auto cls = [&](int32_t ci) {
        const tsc_contrib &c = C[ci];
        if (c.a_op == TSC_OP_I || c.b_op == TSC_OP_I) return 1;
        return (c.coef_re == 0.0 && c.coef_im == 0.0) ? 2 : 0;
      };
      std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        const int cx = cls(x), cy = cls(y);
        if (cx != cy) return cx < cy;
        if (cx != 0) return false;
        if (C[x].coef_re != C[y].coef_re) return C[x].coef_re < C[y].coef_re;
        return C[x].coef_im < C[y].coef_im;
      });
      for (size_t q = 0; q < order.size(); ++q) {
        const int64_t pos = ob.cbeg + (int64_t)q;
        const tsc_contrib &c = C[order[q]];
        cps[s][pos] = c;
        PkContrib &pc = pcs[s][pos];
        pc.coef_re = c.coef_re;
        pc.coef_im = c.coef_im;
        const int k = c.k;
        if (c.a_op == TSC_OP_I || c.b_op == TSC_OP_I) {
          const int buf = (c.b_op == TSC_OP_I) ? c.a_buf : c.b_buf;
          const int op = (c.b_op == TSC_OP_I) ? c.a_op : c.b_op;
          if (op != TSC_OP_I && buf != TSC_BUF_OPS && buf != TSC_BUF_IN)
            return unsupported("an elementwise term reads an intermediate");
          continue;
        }
        if (cls(order[q]) == 2 || k <= 0 || ob.m <= 0 || ob.n <= 0) continue;
        pc.ksteps = (k + 15) / 16;
        // A panel: rows = output rows; B panel: rows = output columns
        const int ta = (c.a_op == TSC_OP_T || c.a_op == TSC_OP_C) ? 1 : 0;
        const int tb = (c.b_op == TSC_OP_N || c.b_op == TSC_OP_CONJ) ? 1 : 0;
        const bool cja = (c.a_op == TSC_OP_C || c.a_op == TSC_OP_CONJ);
        const bool cjb = (c.b_op == TSC_OP_C || c.b_op == TSC_OP_CONJ);
        const int bufs[2] = {c.a_buf, c.b_buf};
        const int64_t offs[2] = {c.a_off, c.b_off};
        const int lds[2] = {c.lda, c.ldb};
        const int rows[2] = {ob.m, ob.n};
        const int trs[2] = {ta, tb};
        int kinds[2] = {0, 0};
        for (int w = 0; w < 2; ++w) {
          int64_t dst = 0;
          int kind;
          if (bufs[w] == TSC_BUF_OPS || bufs[w] == TSC_BUF_IN) {
            const int which = (bufs[w] == TSC_BUF_OPS) ? 0 : 1;
            dst = panel(which, offs[w], lds[w], rows[w], k, trs[w]);
            kind = (which == 0) ? PK_OPS : PK_IN;
          } else if (bufs[w] == TSC_BUF_MID && s == 1) {
            auto it = mid_at.find(offs[w]);
            if (it == mid_at.end()) return unsupported("an operand reads inside an intermediate");
            const tsc_outblock &mb = O1[it->second];
            if (lds[w] != mb.ldc) return unsupported("an intermediate is read with another leading dimension");
            // row panel of the block: rows <= m_b, k == n_b; transposed: rows <= n_b, k == m_b
            if (trs[w] == 0 ? (rows[w] > mb.m || k != mb.n) : (rows[w] > mb.n || k != mb.m))
              return unsupported("an intermediate is read as a partial window");
            (trs[w] == 0 ? need_r : need_c)[it->second] = 1;
            midrefs.push_back(MidRef{s, pos, w, it->second, trs[w]});
            kind = PK_MID;
          } else {
            return unsupported("operand buffer");
          }
          if (w == 0) {
            pc.a_off = dst;
            pc.a_nch = chunks_for(k);
            pc.a_rmax = rows[w] - 1;
          } else {
            pc.b_off = dst;
            pc.b_nch = chunks_for(k);
            pc.b_rmax = rows[w] - 1;
          }
          kinds[w] = kind;
        }
        pc.flags = (cja ? 1 : 0) | (cjb ? 2 : 0) | (kinds[0] << 4) | (kinds[1] << 6);
      }
    }
  }
  // intermediate layout: each stage-1 block as a row panel and/or a transposed panel
  std::vector<PkOut> pouts(O1.size(), PkOut{-1, -1, 0, 0});
  for (size_t b = 0; b < O1.size(); ++b) {
    const tsc_outblock &ob = O1[b];
    if (need_r[b]) {
      pouts[b].r_nch = chunks_for(ob.n);
      pouts[b].r_off = S->mid_len;
      S->mid_len += (int64_t)ob.m * pouts[b].r_nch * 4;
    }
    if (need_c[b]) {
      pouts[b].c_nch = chunks_for(ob.m);
      pouts[b].c_off = S->mid_len;
      S->mid_len += (int64_t)ob.n * pouts[b].c_nch * 4;
    }
  }
  for (const MidRef &r : midrefs) {
    PkContrib &pc = pcs[r.stage][r.pos];
    const PkOut &po = pouts[r.blk];
    const int64_t off = r.trans == 0 ? po.r_off : po.c_off;
    if (r.operand == 0)
      pc.a_off = off;
    else
      pc.b_off = off;
  }
  S->ops_len = plen[0];
  S->in_len = plen[1];

  // device buffers
  int rc;
  if ((rc = alloc_bytes(&S->d_ops, S->ops_len * 16, "packed arena"))) return rc;
  if ((rc = alloc_bytes(&S->d_in, S->in_len * 16, "packed psi"))) return rc;
  if ((rc = alloc_bytes(&S->d_mid, S->mid_len * 16, "packed intermediates"))) return rc;
  std::vector<PkJob> jobs_ops = make_jobs(pan[0]), jobs_in = make_jobs(pan[1]);
  S->n_ops_jobs = (int)jobs_ops.size();
  S->n_in_jobs = (int)jobs_in.size();
  if ((rc = upload_vec(&S->d_ops_panels, pan[0], st))) return rc;
  if ((rc = upload_vec(&S->d_in_panels, pan[1], st))) return rc;
  if ((rc = upload_vec(&S->d_ops_jobs, jobs_ops, st))) return rc;
  if ((rc = upload_vec(&S->d_in_jobs, jobs_in, st))) return rc;
  int64_t slots = 0;
  for (int s = 0; s < 2; ++s) {
    PkStage &G = S->st[s];
    Units un = make_tiles(cps[s].data(), p->h_o[s].data(), (int32_t)p->h_o[s].size());
    G.nt = (int)un.tiles.size();
    G.nr = (int)un.reds.size();
    G.nslot = un.nslot;
    slots = std::max(slots, un.nslot);
    if ((rc = upload_vec(&G.d_pc, pcs[s], st))) return rc;
    if ((rc = upload_vec(&G.d_cp, cps[s], st))) return rc;
    if ((rc = upload_vec(&G.d_t, un.tiles, st))) return rc;
    if ((rc = upload_vec(&G.d_r, un.reds, st))) return rc;
  }
  if ((rc = upload_vec(&S->st[0].d_po, pouts, st))) return rc;
  if ((rc = alloc_bytes(&S->d_part, slots * (int64_t)TM * TN * (int64_t)sizeof(float2), "partial tiles"))) return rc;
  if ((rc = launch_pack(p->d_ops64, S->d_ops, S->d_ops_panels, S->d_ops_jobs, S->n_ops_jobs, st))) return rc;
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  return TSC_OK;
}
}  // namespace

int pk_build(tsc_plan *p) {
  if (p->pk) {
    if (!p->pk->unsupported.empty()) {
      set_error("BF16x3 path: " + p->pk->unsupported);
      return TSC_EINVAL;
    }
    return TSC_OK;
  }
  PkState *S = new PkState();
  const int rc = pk_build_impl(p, S);
  if (rc && S->unsupported.empty()) {
    // allocation or launch failure: leave nothing behind, a later call retries
    pk_free(S);
    return rc;
  }
  p->pk = S;
  return rc;
}

int pk_repack_ops(tsc_plan *p) {
  PkState *S = p->pk;
  if (!S || !S->unsupported.empty() || !S->d_ops) return TSC_OK;
  return launch_pack(p->d_ops64, S->d_ops, S->d_ops_panels, S->d_ops_jobs, S->n_ops_jobs, p->ctx->stream);
}

int pk_enqueue(tsc_plan *p, const double2 *in, double2 *out, bool events) {
  PkState *S = p->pk;
  if (!S || !S->unsupported.empty()) {
    set_error("BF16x3 path used before it was built");
    return TSC_ESTATE;
  }
  hipStream_t s = p->ctx->stream;
  if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[0], s));
  int rc = launch_pack(in, S->d_in, S->d_in_panels, S->d_in_jobs, S->n_in_jobs, s);
  if (rc) return rc;
  PkPtrs P{{reinterpret_cast<const bf8 *>(S->d_ops), reinterpret_cast<const bf8 *>(S->d_in),
            reinterpret_cast<const bf8 *>(S->d_mid)}};
  for (int g = 0; g < 2; ++g) {
    const PkStage &G = S->st[g];
    if (G.nt > 0) {
      hipLaunchKernelGGL(pk_gemm_kernel, dim3(G.nt), dim3(64), 0, s, P, reinterpret_cast<bf8 *>(S->d_mid), in,
                         p->d_ops64, out, p->d_o[g], S->st[0].d_po, G.d_cp, G.d_pc, G.d_t,
                         reinterpret_cast<float2 *>(S->d_part));
      TSC_HIP_CHECK(hipGetLastError());
    }
    if (G.nr > 0) {
      hipLaunchKernelGGL(pk_reduce_kernel, dim3(G.nr), dim3(256), 0, s, reinterpret_cast<bf8 *>(S->d_mid), out,
                         p->d_o[g], S->st[0].d_po, G.d_r, reinterpret_cast<const float2 *>(S->d_part));
      TSC_HIP_CHECK(hipGetLastError());
    }
    if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[g + 1], s));
  }
  return TSC_OK;
}

}  // namespace tsc
