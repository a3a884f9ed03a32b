// Grouped complex contraction kernels for gfx950 (CDNA4).
//
// A work unit is one 64x64 tile of one output block together with a
// contiguous range of that block's contributions:
//     C_tile (+)= sum_c coef_c * op_a(A_c) * op_b(B_c)
// Units that own their whole range write C directly; a long range may be
// cut into several units that write partial tiles, summed afterwards in a
// fixed order by reduce_kernel (deterministic, no atomics). This is the
// batched replacement for the per-sector ZGEMM pairs of model_wave's
// kernels (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90: 1817, 1869,
// 2555-2560, 3385-3408, ...).
//
// The K loop runs over the flattened sequence of (contribution, k0) steps
// of the unit; the global loads of the next step are issued before the
// MFMAs of the current one, so memory latency hides behind the matrix
// cores. Partial K steps issue only the MFMA k-groups they need.
//
// Waves are arranged 2x2 over the tile; each wave computes a 32x32 sub-tile
// as 2x2 MFMA 16x16 blocks, real and imaginary parts in separate
// accumulators (4 real MFMAs per complex k-group), here with
// v_mfma_f32_16x16x4f32, TK = 16 (complex FP32 products: the FP32 path).
// The FP64 path (tsc_gd.hip) and the BF16x3 path (tsc_pk.hip) have kernels
// of their own; this file also holds the MFMA layout probes.
#include "tsc_internal.h"

namespace tsc {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int TM = kTM, TN = kTN;

// ---------------------------------------------------------------------
// element loads: buffers IN/OUT are complex128, OPS/MID complex<CT>
template <typename CT>
__device__ __forceinline__ void ldbuf(const Bufs &B, int buf, int64_t off, CT &re, CT &im) {
  if (buf == TSC_BUF_IN || buf == TSC_BUF_OUT || sizeof(CT) == 8) {
    double2 v = reinterpret_cast<const double2 *>(B.p[buf])[off];
    re = (CT)v.x;
    im = (CT)v.y;
  } else {
    float2 v = reinterpret_cast<const float2 *>(B.p[buf])[off];
    re = (CT)v.x;
    im = (CT)v.y;
  }
}

// op(X)[r, c] for a column-major X with leading dimension ld
template <typename CT>
__device__ __forceinline__ void ld_op(const Bufs &B, int buf, int64_t off, int ld, int op, int r, int c, CT &re,
                                      CT &im) {
  const bool tr = (op == TSC_OP_T || op == TSC_OP_C);
  const int64_t idx = tr ? (off + (int64_t)c + (int64_t)r * ld) : (off + (int64_t)r + (int64_t)c * ld);
  ldbuf<CT>(B, buf, idx, re, im);
  if (op == TSC_OP_C || op == TSC_OP_CONJ) im = -im;
}

// Row of accumulator element r held by `lane` in a 16x16 MFMA output tile
// (the column is lane & 15). The two candidate layouts are told apart once
// per context by probe_layouts(); rowmode selects the one the device uses.
__device__ __forceinline__ int acc_row(int rowmode, int lane, int r) {
  return rowmode ? ((lane >> 4) + 4 * r) : (4 * (lane >> 4) + r);
}

template <typename CT>
__device__ __forceinline__ void cmul(CT ar, CT ai, CT br, CT bi, CT &cr, CT &ci) {
  cr = ar * br - ai * bi;
  ci = ar * bi + ai * br;
}

__device__ __forceinline__ bool is_elementwise(const tsc_contrib &c) {
  return c.a_op == TSC_OP_I || c.b_op == TSC_OP_I;
}

__device__ __forceinline__ int next_gemm(const tsc_contrib *__restrict__ contribs, int ci, int cend) {
  while (ci < cend && is_elementwise(contribs[ci])) ++ci;
  return ci;
}

// elementwise contributions (one operand the identity) of a unit
template <typename CT, typename ACC>
__device__ __forceinline__ void elementwise_pass(const Bufs &bufs, const tsc_contrib *__restrict__ contribs, int cbeg,
                                                 int cend, int i0, int j0, int m, int n, int wm, int wn, int lane,
                                                 int rowmode, ACC (&accR)[2][2], ACC (&accI)[2][2]) {
  for (int ci = cbeg; ci < cend; ++ci) {
    const tsc_contrib c = contribs[ci];
    if (!is_elementwise(c)) continue;
    const bool useA = (c.b_op == TSC_OP_I);
    const int buf = useA ? c.a_buf : c.b_buf;
    const int64_t off = useA ? c.a_off : c.b_off;
    const int ldx = useA ? c.lda : c.ldb;
    const int op = useA ? c.a_op : c.b_op;
    if (op == TSC_OP_I) continue;
    const CT cre = (CT)c.coef_re, cim = (CT)c.coef_im;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = i0 + wm + a * 16 + acc_row(rowmode, lane, r);
          const int j = j0 + wn + b * 16 + (lane & 15);
          if (i < m && j < n) {
            CT xr, xi, pr, pi;
            ld_op<CT>(bufs, buf, off, ldx, op, i, j, xr, xi);
            cmul<CT>(cre, cim, xr, xi, pr, pi);
            accR[a][b][r] += pr;
            accI[a][b][r] += pi;
          }
        }
  }
}

// write the unit's accumulators: to C, or as a partial tile to part[slot]
template <typename CT, typename ACC>
__device__ __forceinline__ void write_unit(const Bufs &bufs, const tsc_outblock &ob, const Tile &t, void *part,
                                           int i0, int j0, int wm, int wn, int lane, int rowmode,
                                           const ACC (&accR)[2][2], const ACC (&accI)[2][2]) {
  const int m = ob.m, n = ob.n;
  const bool out_f64 = (ob.c_buf == TSC_BUF_IN || ob.c_buf == TSC_BUF_OUT || sizeof(CT) == 8);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = wm + a * 16 + acc_row(rowmode, lane, r);
        const int jl = wn + b * 16 + (lane & 15);
        const int i = i0 + il, j = j0 + jl;
        if (i >= m || j >= n) continue;
        if (t.slot >= 0) {
          if (sizeof(CT) == 8) {
            double2 *P = reinterpret_cast<double2 *>(part) + (int64_t)t.slot * (TM * TN);
            P[jl * TM + il] = make_double2((double)accR[a][b][r], (double)accI[a][b][r]);
          } else {
            float2 *P = reinterpret_cast<float2 *>(part) + (int64_t)t.slot * (TM * TN);
            P[jl * TM + il] = make_float2((float)accR[a][b][r], (float)accI[a][b][r]);
          }
          continue;
        }
        const int64_t idx = ob.c_off + (int64_t)i + (int64_t)j * ob.ldc;
        if (out_f64) {
          double2 *C = reinterpret_cast<double2 *>(const_cast<void *>(bufs.p[ob.c_buf]));
          double2 v = make_double2((double)accR[a][b][r], (double)accI[a][b][r]);
          if (ob.mode) {
            const double2 o = C[idx];
            v.x += o.x;
            v.y += o.y;
          }
          C[idx] = v;
        } else {
          float2 *C = reinterpret_cast<float2 *>(const_cast<void *>(bufs.p[ob.c_buf]));
          float2 v = make_float2((float)accR[a][b][r], (float)accI[a][b][r]);
          if (ob.mode) {
            const float2 o = C[idx];
            v.x += o.x;
            v.y += o.y;
          }
          C[idx] = v;
        }
      }
}

// ---------------------------------------------------------------------
// FP64 / FP32 kernel
template <typename CT>
struct MfmaTraits;
template <>
struct MfmaTraits<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct MfmaTraits<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Per-thread staging registers of one (contribution, k0) step.
template <typename CT, int NA, int NB>
struct Stage {
  CT ar[NA], ai[NA], br[NB], bi[NB];
  CT cre, cim;
  int nk;  // valid k of this step (<= TK)
};

// Thread mapping of the A tile (TM x TK) and B tile (TK x TN) for a step:
// the contiguous global index runs fastest across threads.
template <int TK>
__device__ __forceinline__ void a_coord(bool a_tr, int tid, int r, int &ii, int &kk) {
  if (!a_tr) {
    ii = tid % TM;
    kk = tid / TM + r * (256 / TM);
  } else {
    kk = tid % TK;
    ii = tid / TK + r * (256 / TK);
  }
}
template <int TK>
__device__ __forceinline__ void b_coord(bool b_tr, int tid, int r, int &jj, int &kk) {
  if (b_tr) {
    jj = tid % TN;
    kk = tid / TN + r * (256 / TN);
  } else {
    kk = tid % TK;
    jj = tid / TK + r * (256 / TK);
  }
}

template <typename CT, int TK, int NA, int NB>
__device__ __forceinline__ void load_stage(const Bufs &bufs, const tsc_contrib &c, int k0, int i0, int j0, int m, int n,
                                           int tid, Stage<CT, NA, NB> &s) {
  const int K = c.k;
  const bool a_tr = (c.a_op == TSC_OP_T || c.a_op == TSC_OP_C);
  const bool b_tr = (c.b_op == TSC_OP_T || c.b_op == TSC_OP_C);
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    int ii, kk;
    a_coord<TK>(a_tr, tid, r, ii, kk);
    const int gi = i0 + ii, gk = k0 + kk;
    CT xr = 0, xi = 0;
    if (gi < m && gk < K) ld_op<CT>(bufs, c.a_buf, c.a_off, c.lda, c.a_op, gi, gk, xr, xi);
    s.ar[r] = xr;
    s.ai[r] = xi;
  }
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    int jj, kk;
    b_coord<TK>(b_tr, tid, r, jj, kk);
    const int gj = j0 + jj, gk = k0 + kk;
    CT xr = 0, xi = 0;
    if (gj < n && gk < K) ld_op<CT>(bufs, c.b_buf, c.b_off, c.ldb, c.b_op, gk, gj, xr, xi);
    s.br[r] = xr;
    s.bi[r] = xi;
  }
  s.cre = (CT)c.coef_re;
  s.cim = (CT)c.coef_im;
  s.nk = min(TK, K - k0);
}

template <typename CT, int TK>
__global__ __launch_bounds__(256) void grouped_kernel(Bufs bufs, const tsc_outblock *__restrict__ blocks,
                                                      const tsc_contrib *__restrict__ contribs,
                                                      const Tile *__restrict__ tiles, int rowmode, void *part) {
  typedef typename MfmaTraits<CT>::acc_t acc_t;
  constexpr int PAD = (sizeof(CT) == 8) ? 16 : 32;  // row padding (elements)
  constexpr int NA = (TM * TK) / 256, NB = (TN * TK) / 256;
  __shared__ CT As_r[TK][TM + PAD], As_i[TK][TM + PAD];
  __shared__ CT Bs_r[TK][TN + PAD], Bs_i[TK][TN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * 32;
  const int wn = (wave >> 1) * 32;
  const Tile t = tiles[blockIdx.x];
  if (t.blk < 0) return;  // padding unit
  const tsc_outblock ob = blocks[t.blk];
  const int i0 = t.ti * TM, j0 = t.tj * TN;
  const int m = ob.m, n = ob.n;

  acc_t accR[2][2], accI[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      accR[a][b] = acc_t{0, 0, 0, 0};
      accI[a][b] = acc_t{0, 0, 0, 0};
    }

  elementwise_pass<CT>(bufs, contribs, t.cbeg, t.cend, i0, j0, m, n, wm, wn, lane, rowmode, accR, accI);

  // flattened, software-pipelined (contribution, k0) loop
  int ci = next_gemm(contribs, t.cbeg, t.cend);
  int k0 = 0;
  Stage<CT, NA, NB> st;
  bool a_tr = false, b_tr = false;
  if (ci < t.cend) {
    const tsc_contrib c = contribs[ci];
    load_stage<CT, TK, NA, NB>(bufs, c, 0, i0, j0, m, n, tid, st);
    a_tr = (c.a_op == TSC_OP_T || c.a_op == TSC_OP_C);
    b_tr = (c.b_op == TSC_OP_T || c.b_op == TSC_OP_C);
  }
  int Kcur = (ci < t.cend) ? contribs[ci].k : 0;
  while (ci < t.cend) {
    // stage the current step (A scaled by its coefficient)
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int ii, kk;
      a_coord<TK>(a_tr, tid, r, ii, kk);
      CT pr, pi;
      cmul<CT>(st.cre, st.cim, st.ar[r], st.ai[r], pr, pi);
      As_r[kk][ii] = pr;
      As_i[kk][ii] = pi;
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int jj, kk;
      b_coord<TK>(b_tr, tid, r, jj, kk);
      Bs_r[kk][jj] = st.br[r];
      Bs_i[kk][jj] = st.bi[r];
    }
    const int nks = (st.nk + 3) >> 2;
    __syncthreads();
    // issue the next step's loads
    int nci = ci, nk0 = k0 + TK;
    if (nk0 >= Kcur) {
      nci = next_gemm(contribs, ci + 1, t.cend);
      nk0 = 0;
    }
    if (nci < t.cend) {
      const tsc_contrib c = contribs[nci];
      load_stage<CT, TK, NA, NB>(bufs, c, nk0, i0, j0, m, n, tid, st);
      a_tr = (c.a_op == TSC_OP_T || c.a_op == TSC_OP_C);
      b_tr = (c.b_op == TSC_OP_T || c.b_op == TSC_OP_C);
      if (nci != ci) Kcur = c.k;
    }
    // MFMAs of the current step
#pragma unroll
    for (int ks = 0; ks < TK / 4; ++ks) {
      if (ks < nks) {
        const int kr = ks * 4 + (lane >> 4);
        CT ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          ar[a] = As_r[kr][wm + a * 16 + (lane & 15)];
          ai[a] = As_i[kr][wm + a * 16 + (lane & 15)];
          nai[a] = -ai[a];
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          br[b] = Bs_r[kr][wn + b * 16 + (lane & 15)];
          bi[b] = Bs_i[kr][wn + b * 16 + (lane & 15)];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            accR[a][b] = MfmaTraits<CT>::mma(ar[a], br[b], accR[a][b]);
            accR[a][b] = MfmaTraits<CT>::mma(nai[a], bi[b], accR[a][b]);
            accI[a][b] = MfmaTraits<CT>::mma(ar[a], bi[b], accI[a][b]);
            accI[a][b] = MfmaTraits<CT>::mma(ai[a], br[b], accI[a][b]);
          }
      }
    }
    __syncthreads();
    ci = nci;
    k0 = nk0;
  }

  write_unit<CT>(bufs, ob, t, part, i0, j0, wm, wn, lane, rowmode, accR, accI);
}

// ---------------------------------------------------------------------
// Sum the partial tiles of split units in slot order and write C.
template <typename CT>
__global__ __launch_bounds__(256) void reduce_kernel(Bufs bufs, const tsc_outblock *__restrict__ blocks,
                                                     const Reduce *__restrict__ reds, const void *part) {
  const Reduce rd = reds[blockIdx.x];
  const tsc_outblock ob = blocks[rd.blk];
  const int i0 = rd.ti * TM, j0 = rd.tj * TN;
  const bool out_f64 = (ob.c_buf == TSC_BUF_IN || ob.c_buf == TSC_BUF_OUT || sizeof(CT) == 8);
  for (int e = threadIdx.x; e < TM * TN; e += 256) {
    const int il = e % TM, jl = e / TM;
    const int i = i0 + il, j = j0 + jl;
    if (i >= ob.m || j >= ob.n) continue;
    double sr = 0.0, si = 0.0;
    for (int s = 0; s < rd.nslot; ++s) {
      const int64_t pidx = (int64_t)(rd.slot0 + s) * (TM * TN) + e;
      if (sizeof(CT) == 8) {
        const double2 v = reinterpret_cast<const double2 *>(part)[pidx];
        sr += v.x;
        si += v.y;
      } else {
        const float2 v = reinterpret_cast<const float2 *>(part)[pidx];
        sr += (double)v.x;
        si += (double)v.y;
      }
    }
    const int64_t idx = ob.c_off + (int64_t)i + (int64_t)j * ob.ldc;
    if (out_f64) {
      double2 *C = reinterpret_cast<double2 *>(const_cast<void *>(bufs.p[ob.c_buf]));
      double2 v = make_double2(sr, si);
      if (ob.mode) {
        const double2 o = C[idx];
        v.x += o.x;
        v.y += o.y;
      }
      C[idx] = v;
    } else {
      float2 *C = reinterpret_cast<float2 *>(const_cast<void *>(bufs.p[ob.c_buf]));
      float2 v = make_float2((float)sr, (float)si);
      if (ob.mode) {
        const float2 o = C[idx];
        v.x += o.x;
        v.y += o.y;
      }
      C[idx] = v;
    }
  }
}

// ---------------------------------------------------------------------
// Layout probe: C = A B with A[i][k] = i (or 1) and B[k][j] = 1 (or j)
// for every k; out[kind][pass][lane*4 + r] = C element held in register r,
// i.e. 4*row or 4*col of that element. Uses only the row-of-A = lane & 15
// and col-of-B = lane & 15 conventions of the 16x16 MFMA family.
__global__ void probe_kernel(float *out) {
  const int lane = threadIdx.x;
  const int i = lane & 15;
  {
    d4 c1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)i, 1.0, d4{0, 0, 0, 0}, 0, 0, 0);
    d4 c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(1.0, (double)i, d4{0, 0, 0, 0}, 0, 0, 0);
    for (int r = 0; r < 4; ++r) {
      out[0 * 512 + lane * 4 + r] = (float)c1[r];
      out[0 * 512 + 256 + lane * 4 + r] = (float)c2[r];
    }
  }
  {
    f4 c1 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)i, 1.0f, f4{0, 0, 0, 0}, 0, 0, 0);
    f4 c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(1.0f, (float)i, f4{0, 0, 0, 0}, 0, 0, 0);
    for (int r = 0; r < 4; ++r) {
      out[1 * 512 + lane * 4 + r] = c1[r];
      out[1 * 512 + 256 + lane * 4 + r] = c2[r];
    }
  }
}

typedef float f16v __attribute__((ext_vector_type(16)));

// Row of accumulator register r (0..15) held by `lane` in a 32x32 MFMA
// output block (the column is lane & 31), as the packed BF16x3 kernels use it.
__host__ __device__ __forceinline__ int acc_row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// 32x32x16 bf16: out[lane*16 + r] = 16*row, out[1024 + lane*16 + r] = 16*col
__global__ void probe32_kernel(float *out) {
  const int lane = threadIdx.x;
  const int i = lane & 31;
  bf8 vi, v1;
  for (int e = 0; e < 8; ++e) {
    vi[e] = (__bf16)(float)i;
    v1[e] = (__bf16)1.0f;
  }
  f16v z;
  for (int r = 0; r < 16; ++r) z[r] = 0.f;
  f16v c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(vi, v1, z, 0, 0, 0);
  f16v c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(v1, vi, z, 0, 0, 0);
  for (int r = 0; r < 16; ++r) {
    out[lane * 16 + r] = c1[r];
    out[1024 + lane * 16 + r] = c2[r];
  }
}

// modes[0], modes[1]: row layout of the f64 / f32 16x16 accumulators;
// modes[2]: 0 if the 32x32x16 bf16 accumulator layout is the one
// acc_row32() assumes, -1 otherwise (the BF16x3 path then refuses to run).
int probe_layouts(hipStream_t stream, int modes[3]) {
  float *d = nullptr;
  TSC_HIP_CHECK(hipMalloc((void **)&d, sizeof(float) * (2 * 512 + 2048)));
  hipLaunchKernelGGL(probe_kernel, dim3(1), dim3(64), 0, stream, d);
  hipLaunchKernelGGL(probe32_kernel, dim3(1), dim3(64), 0, stream, d + 2 * 512);
  hipError_t le = hipGetLastError();
  std::vector<float> h(2 * 512 + 2048);
  hipError_t ce = hipMemcpyAsync(h.data(), d, sizeof(float) * h.size(), hipMemcpyDeviceToHost, stream);
  hipError_t se = hipStreamSynchronize(stream);
  hipFree(d);
  TSC_HIP_CHECK(le);
  TSC_HIP_CHECK(ce);
  TSC_HIP_CHECK(se);
  for (int kind = 0; kind < 2; ++kind) {
    int ok[2] = {1, 1};
    for (int lane = 0; lane < 64; ++lane)
      for (int r = 0; r < 4; ++r) {
        const float row = h[kind * 512 + lane * 4 + r] / 4.f;
        const float col = h[kind * 512 + 256 + lane * 4 + r] / 4.f;
        if (col != (float)(lane & 15)) ok[0] = ok[1] = 0;
        if (row != (float)(4 * (lane >> 4) + r)) ok[0] = 0;
        if (row != (float)((lane >> 4) + 4 * r)) ok[1] = 0;
      }
    if (ok[0])
      modes[kind] = 0;
    else if (ok[1])
      modes[kind] = 1;
    else {
      set_error("MFMA accumulator layout probe: unknown layout for kind " + std::to_string(kind));
      return TSC_EHIP;
    }
  }
  const float *h32 = h.data() + 2 * 512;
  modes[2] = 0;
  for (int lane = 0; lane < 64; ++lane)
    for (int r = 0; r < 16; ++r)
      if (h32[lane * 16 + r] != 16.f * acc_row32(lane, r) || h32[1024 + lane * 16 + r] != 16.f * (lane & 31))
        modes[2] = -1;
  return TSC_OK;
}

hipError_t launch_grouped(int prec, const Bufs &bufs, const tsc_outblock *d_blocks, const tsc_contrib *d_contribs,
                          const Tile *d_tiles, int ntiles, const Reduce *d_reds, int nreds, void *part,
                          hipStream_t stream, const int *modes) {
  if (ntiles <= 0) return hipSuccess;
  dim3 grid(ntiles), block(256);
  // FP64 runs by default on the direct-to-LDS kernels (tsc_gd.hip; this
  // 64x64 register-staged one is the tile shape the autotuner weighs against
  // them), BF16x3 on the packed-operand ones (tsc_pk.hip)
  if (prec == TSC_PREC_F64) {
    hipLaunchKernelGGL((grouped_kernel<double, 16>), grid, block, 0, stream, bufs, d_blocks, d_contribs, d_tiles,
                       modes[0], part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nreds <= 0) return e;
    hipLaunchKernelGGL(reduce_kernel<double>, dim3(nreds), dim3(256), 0, stream, bufs, d_blocks, d_reds,
                       (const void *)part);
    return hipGetLastError();
  }
  if (prec != TSC_PREC_F32) return hipErrorNotSupported;
  hipLaunchKernelGGL((grouped_kernel<float, 16>), grid, block, 0, stream, bufs, d_blocks, d_contribs, d_tiles,
                     modes[1], part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nreds <= 0) return e;
  hipLaunchKernelGGL(reduce_kernel<float>, dim3(nreds), dim3(256), 0, stream, bufs, d_blocks, d_reds,
                     (const void *)part);
  return hipGetLastError();
}

}  // namespace tsc
