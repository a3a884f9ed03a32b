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
//    the contraction kernel does no conversion work;
//  * a workgroup of 4 waves owns one 128x128 output unit, one 64x64 quarter
//    per wave as 2x2 blocks of 32x32 MFMAs; each 16-k step of A (128 rows)
//    and B (128 rows) is moved once into LDS with contiguous 1 KB loads
//    (double-buffered: the loads of step s+1 are in flight while step s is
//    computed) and shared by the two waves that need each half;
//  * the MFMA is issued with the B fragment as its first operand, so a
//    lane's accumulators run along output columns and stores of psi /
//    panels are contiguous across lanes;
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

constexpr int WT = kPkTile;  // 128

__host__ __device__ __forceinline__ int64_t entry(int r, int c, int p, int nch) {
  return (((int64_t)(r >> 6) * nch + c) * 4 + p) * 64 + (r & 63);
}

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
// Packing: one thread per (row, chunk) of a panel, rows fastest (padding rows
// up to the next multiple of 64 are written as zeros), so the 16-byte entry
// stores of consecutive threads are contiguous.
__global__ __launch_bounds__(256) void pack_kernel(const double2 *__restrict__ src, bf8 *__restrict__ dst,
                                                   const PkPanel *__restrict__ panels,
                                                   const PkJob *__restrict__ jobs) {
  const PkJob jb = jobs[blockIdx.x];
  const PkPanel P = panels[jb.panel];
  const int rows_p = (P.rows + 63) & ~63;
  const int64_t t = jb.first + threadIdx.x;
  if (t >= (int64_t)rows_p * P.nch) return;
  const int c = (int)(t / rows_p), r = (int)(t % rows_p);
  const double2 *X = src + P.src_off;
  bf8 v0, v1, v2, v3;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 8 * c + e;
    double2 x = make_double2(0.0, 0.0);
    if (r < P.rows && k < P.kdim) x = P.trans ? X[(int64_t)k + (int64_t)r * P.ld] : X[(int64_t)r + (int64_t)k * P.ld];
    __bf16 a, b, cc, d;
    split2((float)x.x, a, b);
    split2((float)x.y, cc, d);
    v0[e] = a;
    v2[e] = b;
    v1[e] = cc;
    v3[e] = d;
  }
  bf8 *q = dst + P.dst;
  q[entry(r, c, 0, P.nch)] = v0;
  q[entry(r, c, 1, P.nch)] = v1;
  q[entry(r, c, 2, P.nch)] = v2;
  q[entry(r, c, 3, P.nch)] = v3;
}

// ---------------------------------------------------------------------------
// element helpers for packed intermediates
__device__ __forceinline__ float unsplit(__bf16 hi, __bf16 lo) { return (float)hi + (float)lo; }

// one element of a panel: row r, k, at panel entry base `pan`
__device__ __forceinline__ void store_pk_elem(bf8 *pan, int nch, int r, int k, float re, float im, bool acc) {
  __bf16 *q0 = reinterpret_cast<__bf16 *>(pan + entry(r, k >> 3, 0, nch)) + (k & 7);
  // planes are 64 entries (512 bf16) apart
  if (acc) {
    re += unsplit(q0[0], q0[1024]);
    im += unsplit(q0[512], q0[1536]);
  }
  __bf16 a, b, c, d;
  split2(re, a, b);
  split2(im, c, d);
  q0[0] = a;
  q0[512] = c;
  q0[1024] = b;
  q0[1536] = d;
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
  // element by element, one temporary at a time: the accumulators fill most
  // of the register budget (whole-vector temporaries spilled to scratch);
  // with si == 0 the results equal x * sr exactly
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float x = R[a][b][r], y = I[a][b][r];
        R[a][b][r] = x * sr - y * si;
        I[a][b][r] = x * si + y * sr;
      }
}


#define MFMA32(x, y, acc) __builtin_amdgcn_mfma_f32_32x32x16_bf16((x), (y), (acc), 0, 0, 0)

// ---------------------------------------------------------------------------
// Contraction kernel: one 128x128 unit per workgroup of 4 waves. Wave w owns
// the quarter (wa, wb) = (w & 1, w >> 1); accumulator element r of block
// (a, b) held by `lane` is output (i0 + 64wa + 32a + lane%32,
// j0 + 64wb + 32b + 8(r/4) + 4(lane/32) + r%4).
// LDS holds two step buffers of 32 pieces (1 KB each): pieces 0-15 are A
// (row half rb, chunk h, plane p at 8rb + 4h + p), 16-31 the same for B.
// Waves 0/1 load the two A halves of a step, waves 2/3 the two B halves.
// Per step: read this step's fragments from LDS, start the direct
// global->LDS loads of the next step into the other buffer, run the 48
// MFMAs, wait for the own loads, workgroup barrier.
__global__ __launch_bounds__(256, 2) void pk_gemm_kernel(PkPtrs P, bf8 *__restrict__ mid,
                                                         const double2 *__restrict__ in64,
                                                         const double2 *__restrict__ ops64,
                                                         double2 *__restrict__ out64,
                                                         const tsc_outblock *__restrict__ blocks,
                                                         const PkOut *__restrict__ pouts,
                                                         const tsc_contrib *__restrict__ cps,
                                                         const PkContrib *__restrict__ pcs,
                                                         const Tile *__restrict__ tiles, float2 *__restrict__ part) {
  __shared__ bf8 lds0[32][64], lds1[32][64];
  // the wave index is wave-uniform: per-wave conditions become scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, l32 = lane & 31, h = lane >> 5;
  const int wa = wave & 1, wb = wave >> 1;
  const Tile t = tiles[blockIdx.x];
  if (t.blk < 0) return;  // padding unit
  const tsc_outblock ob = blocks[t.blk];
  const int m = ob.m, n = ob.n;
  const int i0 = t.ti * WT, j0 = t.tj * WT;
  const int iw = i0 + 64 * wa, jw = j0 + 64 * wb;
  const bool active = iw < m && jw < n;
  const bool a1 = iw + 32 < m;  // second 32-row block of this wave's quarter holds rows
  // loader role of this wave
  const bool ldA = wave < 2;
  const int lrb = wave & 1;
  const int grb = (ldA ? 2 * t.ti : 2 * t.tj) + lrb;  // 64-row block of the panel
  const bool ld_on = 64 * grb < (ldA ? m : n);
  const int q0 = (ldA ? 0 : 16) + 8 * lrb;

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

  // one step's operands go straight from the panels into LDS (no registers):
  // lane l of the wave writes LDS bytes [16 l, 16 l + 16) of each 1 KB piece
  auto issue_step = [&](const PkContrib &c, int s, bf8 (*L)[64]) {
    if (!ld_on) return;
    const int nch = ldA ? c.a_nch : c.b_nch;
    const bf8 *base = P.p[ldA ? ((c.flags >> 4) & 3) : ((c.flags >> 6) & 3)] + (ldA ? c.a_off : c.b_off) +
                      ((int64_t)grb * nch + 2 * s) * 256 + lane;
#pragma unroll
    for (int q = 0; q < 8; ++q)  // chunk 2s + q/4, plane q%4
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(base + q * 64),
                                       (__attribute__((address_space(3))) void *)&L[q0 + q][0], 16, 0, 0);
  };
  // all of this wave's loads have landed in LDS; then every wave's have
  auto step_sync = [&]() {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
  };
  auto compute = [&](const PkContrib &c, bf8 (*L)[64], bf8 (*Lnext)[64], bool more, const PkContrib &nc, int ns) {
    bf8 A[2][4], B[2][4];
    if (active) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) A[a][q] = L[8 * wa + 4 * h + q][32 * a + l32];
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) B[b][q] = L[16 + 8 * wb + 4 * h + q][32 * b + l32];
    }
    // fragments are in registers: the next step may now be loaded into the other buffer
    if (more) issue_step(nc, ns, Lnext);
    if (!active) return;
    if (c.flags & 1) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        A[a][1] = neg8(A[a][1]);
        A[a][3] = neg8(A[a][3]);
      }
    }
    if (c.flags & 2) {
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        B[b][1] = neg8(B[b][1]);
        B[b][3] = neg8(B[b][3]);
      }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (jw + 32 * b >= n) continue;  // columns past the block: nothing to compute
      const bf8 nbh = neg8(B[b][1]), nbl = neg8(B[b][3]);
      // planes: 0 re hi, 1 im hi, 2 re lo, 3 im lo
      // Re += br*ar - bi*ai ; Im += bi*ar + br*ai (each product in three terms)
      // (the second 32-row block only when it holds rows of the output block)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !a1) break;
        accR[a][b] = MFMA32(B[b][0], A[a][0], accR[a][b]);
        accI[a][b] = MFMA32(B[b][1], A[a][0], accI[a][b]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !a1) break;
        accR[a][b] = MFMA32(B[b][2], A[a][0], accR[a][b]);
        accI[a][b] = MFMA32(B[b][3], A[a][0], accI[a][b]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !a1) break;
        accR[a][b] = MFMA32(B[b][0], A[a][2], accR[a][b]);
        accI[a][b] = MFMA32(B[b][1], A[a][2], accI[a][b]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !a1) break;
        accR[a][b] = MFMA32(nbh, A[a][1], accR[a][b]);
        accI[a][b] = MFMA32(B[b][0], A[a][1], accI[a][b]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !a1) break;
        accR[a][b] = MFMA32(nbl, A[a][1], accR[a][b]);
        accI[a][b] = MFMA32(B[b][2], A[a][1], accI[a][b]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !a1) break;
        accR[a][b] = MFMA32(nbh, A[a][3], accR[a][b]);
        accI[a][b] = MFMA32(B[b][0], A[a][3], accI[a][b]);
      }
    }
  };

  // the GEMM terms of a unit are a prefix of its range (the host orders each
  // block's contributions GEMMs first), so the next one is the next record
  int cc = t.cbeg;
  PkContrib cur{};
  bool live = false;
  if (cc < t.cend) {
    cur = pcs[cc];
    live = cur.ksteps > 0;
  }
  if (live) issue_step(cur, 0, lds0);
  step_sync();
  int cs = 0, st = 0;
  bool first = true;
  float gr = 1.f, gi = 0.f;
  while (live) {
    int nc = cc, ns = cs + 1;
    PkContrib nxt = cur;
    bool more = true;
    if (ns >= cur.ksteps) {
      nc = cc + 1;
      ns = 0;
      more = nc < t.cend;
      if (more) {
        nxt = pcs[nc];
        more = nxt.ksteps > 0;
      }
    }
    if (cs == 0) {
      if (!first && (cur.rs_re != 1.f || cur.rs_im != 0.f)) scale_acc(accR, accI, cur.rs_re, cur.rs_im);
      first = false;
      gr = cur.g_re;
      gi = cur.g_im;
    }
    compute(cur, st ? lds1 : lds0, st ? lds0 : lds1, more, nxt, ns);
    step_sync();
    live = more;
    cc = nc;
    cs = ns;
    cur = nxt;
    st ^= 1;
  }
  if (!active) return;
  if (gr != 1.f || gi != 0.f) scale_acc(accR, accI, gr, gi);

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
          const int i = iw + 32 * a + l32;
          const int j = jw + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (i < m && j < n) {
            const double2 x = ld_op64(X, off, ldx, op, i, j);
            const float xr = (float)x.x, xi = (float)x.y;
            accR[a][b][r] += cre * xr - cim * xi;
            accI[a][b][r] += cre * xi + cim * xr;
          }
        }
  }

  // ---- write this wave's quarter
  if (t.slot >= 0) {
    float2 *Pp = part + (int64_t)t.slot * (WT * WT);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int il = 64 * wa + 32 * a + l32, jl = 64 * wb + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (i0 + il < m && j0 + jl < n) Pp[jl * WT + il] = make_float2(accR[a][b][r], accI[a][b][r]);
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
          const int i = iw + 32 * a + l32, j = jw + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
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
    // row panel P[i][j]: the 4 consecutive j of (a, b, q) are half of one entry
    const int kp = 8 * po.r_nch;
    bf8 *pan = mid + po.r_off;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = iw + 32 * a + l32;
          const int jb = jw + 32 * b + 8 * q + 4 * h;
          if (i >= m || jb >= kp) continue;
          bf4v *dst = reinterpret_cast<bf4v *>(pan + entry(i, jb >> 3, 0, po.r_nch)) + h;
          // planes are 64 entries = 128 bf4v apart
          bf4v p0, p1, p2, p3;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            float vr = (jb + e < n) ? accR[a][b][4 * q + e] : 0.f;
            float vi = (jb + e < n) ? accI[a][b][4 * q + e] : 0.f;
            if (ob.mode) {
              vr += unsplit(dst[0][e], dst[256][e]);
              vi += unsplit(dst[128][e], dst[384][e]);
            }
            __bf16 x0, x1, x2, x3;
            split2(vr, x0, x2);
            split2(vi, x1, x3);
            p0[e] = x0;
            p1[e] = x1;
            p2[e] = x2;
            p3[e] = x3;
          }
          dst[0] = p0;
          dst[128] = p1;
          dst[256] = p2;
          dst[384] = p3;
        }
  }
  if (po.c_off >= 0) {
    // transposed panel P[j][i]
    const int kp = 8 * po.c_nch;
    bf8 *pan = mid + po.c_off;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = iw + 32 * a + l32, j = jw + 32 * b + 8 * (r >> 2) + 4 * h + (r & 3);
          if (j >= n || i >= kp) continue;
          const bool in_blk = i < m;
          store_pk_elem(pan, po.c_nch, j, i, in_blk ? accR[a][b][r] : 0.f, in_blk ? accI[a][b][r] : 0.f,
                        ob.mode != 0);
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
  const int i0 = rd.ti * WT, j0 = rd.tj * WT;
  const bool packed = (ob.c_buf == TSC_BUF_MID);
  PkOut po{-1, -1, 0, 0};
  if (packed) po = pouts[rd.blk];
  for (int e = threadIdx.x; e < WT * WT; e += 256) {
    const int il = e % WT, jl = e / WT;
    const int i = i0 + il, j = j0 + jl;
    const bool inside = (i < ob.m && j < ob.n);
    double sr = 0.0, si = 0.0;
    if (inside) {
      for (int s = 0; s < rd.nslot; ++s) {
        const float2 v = part[(int64_t)(rd.slot0 + s) * (WT * WT) + e];
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
      store_pk_elem(mid + po.r_off, po.r_nch, i, j, inside ? (float)sr : 0.f, inside ? (float)si : 0.f,
                    ob.mode != 0);
    if (po.c_off >= 0 && j < ob.n && i < 8 * po.c_nch)
      store_pk_elem(mid + po.c_off, po.c_nch, j, i, inside ? (float)sr : 0.f, inside ? (float)si : 0.f,
                    ob.mode != 0);
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

int64_t rows64(int rows) { return (rows + 63) & ~63; }

std::vector<PkJob> make_jobs(const std::vector<PkPanel> &pan) {
  std::vector<PkJob> jobs;
  for (size_t i = 0; i < pan.size(); ++i) {
    const int64_t total = rows64(pan[i].rows) * pan[i].nch;
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
    plen[which] += rows64(rows) * P.nch * 4;
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
      // 0: GEMM with work, 1: elementwise, 2: nothing to do (zero coefficient,
      // k = 0, empty block); GEMMs come first so the kernel walks a prefix
      auto cls = [&](int32_t ci) {
        const tsc_contrib &c = C[ci];
        if (c.a_op == TSC_OP_I || c.b_op == TSC_OP_I) return 1;
        if ((c.coef_re == 0.0 && c.coef_im == 0.0) || c.k <= 0 || ob.m <= 0 || ob.n <= 0) return 2;
        return 0;
      };
      // GEMM terms ordered by coefficient (few accumulator rescales)
      std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        const int cx = cls(x), cy = cls(y);
        if (cx != cy) return cx < cy;
        if (cx != 0) return false;
        if (C[x].coef_re != C[y].coef_re) return C[x].coef_re < C[y].coef_re;
        return C[x].coef_im < C[y].coef_im;
      });
      double prev_re = 0.0, prev_im = 0.0;
      for (size_t q = 0; q < order.size(); ++q) {
        const int64_t pos = ob.cbeg + (int64_t)q;
        const tsc_contrib &c = C[order[q]];
        cps[s][pos] = c;
        PkContrib &pc = pcs[s][pos];
        pc.g_re = (float)c.coef_re;
        pc.g_im = (float)c.coef_im;
        pc.rs_re = 1.f;
        pc.rs_im = 0.f;
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
        // rescale from the previous GEMM contribution of this block (in this order)
        if (prev_re != 0.0 || prev_im != 0.0) {
          const double den = c.coef_re * c.coef_re + c.coef_im * c.coef_im;
          pc.rs_re = (float)((prev_re * c.coef_re + prev_im * c.coef_im) / den);
          pc.rs_im = (float)((prev_im * c.coef_re - prev_re * c.coef_im) / den);
        }
        prev_re = c.coef_re;
        prev_im = c.coef_im;
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
          } else {
            pc.b_off = dst;
            pc.b_nch = chunks_for(k);
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
      S->mid_len += rows64(ob.m) * pouts[b].r_nch * 4;
    }
    if (need_c[b]) {
      pouts[b].c_nch = chunks_for(ob.m);
      pouts[b].c_off = S->mid_len;
      S->mid_len += rows64(ob.n) * pouts[b].c_nch * 4;
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
  // padding rows of intermediate panels are never written; keep them zero
  if (S->mid_len > 0) TSC_HIP_CHECK(hipMemsetAsync(S->d_mid, 0, (size_t)S->mid_len * 16, st));
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
    // 128x128 units, 4 waves each, two workgroups per CU. (128x256 units of
    // 8 waves for blocks 129..256 wide read each A step once instead of twice,
    // 19% less HBM traffic, but ran slower at YC6 M=6000: one workgroup per
    // CU and a second launch with its own tail.)
    Units un = make_tiles(cps[s].data(), p->h_o[s].data(), (int32_t)p->h_o[s].size(), WT, WT, S->desired);
    G.nt = (int)un.tiles.size();
    G.nr = (int)un.reds.size();
    G.nslot = un.nslot;
    slots = std::max(slots, un.nslot);
    if ((rc = upload_vec(&G.d_pc, pcs[s], st))) return rc;
    if ((rc = upload_vec(&G.d_cp, cps[s], st))) return rc;
    if ((rc = upload_vec(&G.d_t, un.tiles, st))) return rc;
    if ((rc = upload_vec(&G.d_r, un.reds, st))) return rc;
    G.h_cp.swap(cps[s]);
  }
  if ((rc = upload_vec(&S->st[0].d_po, pouts, st))) return rc;
  if ((rc = alloc_bytes(&S->d_part, slots * (int64_t)WT * WT * (int64_t)sizeof(float2), "partial tiles"))) return rc;
  S->part_slots = slots;
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

int pk_retile(tsc_plan *p, double desired) {
  PkState *S = p->pk;
  if (!S || !S->unsupported.empty()) {
    set_error("BF16x3 path re-cut before it was built");
    return TSC_ESTATE;
  }
  hipStream_t st = p->ctx->stream;
  // the current units may still be in use by queued launches
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  int64_t slots = 0;
  int rc;
  for (int s = 0; s < 2; ++s) {
    PkStage &G = S->st[s];
    Units un = make_tiles(G.h_cp.data(), p->h_o[s].data(), (int32_t)p->h_o[s].size(), WT, WT, desired);
    hipFree(G.d_t);
    hipFree(G.d_r);
    G.d_t = nullptr;
    G.d_r = nullptr;
    G.nt = (int)un.tiles.size();
    G.nr = (int)un.reds.size();
    G.nslot = un.nslot;
    slots = std::max(slots, un.nslot);
    if ((rc = upload_vec(&G.d_t, un.tiles, st))) return rc;
    if ((rc = upload_vec(&G.d_r, un.reds, st))) return rc;
  }
  if (slots > S->part_slots) {
    hipFree(S->d_part);
    S->d_part = nullptr;
    S->part_slots = 0;
    if ((rc = alloc_bytes(&S->d_part, slots * (int64_t)WT * WT * (int64_t)sizeof(float2), "partial tiles")))
      return rc;
    S->part_slots = slots;
  }
  S->desired = desired;
  TSC_HIP_CHECK(hipStreamSynchronize(st));
  return TSC_OK;
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
      hipLaunchKernelGGL(pk_gemm_kernel, dim3(G.nt), dim3(256), 0, s, P, reinterpret_cast<bf8 *>(S->d_mid), in,
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
