// C-ABI of the hot path: contexts, grouped-contraction plans (H.psi),
// and the Lanczos driver.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>

#include "tsc_internal.h"

namespace tsc {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

__global__ void f64_to_f32_kernel(const double2 *__restrict__ src, float2 *__restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}

// ---------------------------------------------------------------------------
// Symmetric tridiagonal eigen-solve for the Lanczos matrix (DSTEV's role,
// eigen.f90:99,117,141): lowest eigenvalue by Sturm-sequence bisection,
// its eigenvector by inverse iteration with a pivoted tridiagonal solve.
static int sturm_below(const double *d, const double *e, int n, double x) {
  int cnt = 0;
  double q = d[0] - x;
  const double tiny = 1e-300;
  if (q < 0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (std::fabs(q) < tiny) q = (q < 0 ? -tiny : tiny);
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

double tridiag_lowest(const double *d, const double *e, int n) {
  if (n == 1) return d[0];
  double lo = d[0], hi = d[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  const double scale = std::max(std::fabs(lo), std::fabs(hi));
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_below(d, e, n, mid) >= 1)
      hi = mid;
    else
      lo = mid;
    if (hi - lo <= 2.0 * 2.2e-16 * scale) break;
  }
  return 0.5 * (lo + hi);
}

// eigenvector of the lowest eigenvalue lam (unit norm, first nonzero
// component made positive)
void tridiag_vector(const double *d, const double *e, int n, double lam, double *z) {
  if (n == 1) {
    z[0] = 1.0;
    return;
  }
  double scale = 0.0;
  for (int i = 0; i < n; ++i) scale = std::max(scale, std::fabs(d[i]) + (i < n - 1 ? std::fabs(e[i]) : 0.0));
  const double shift = lam - 4.0 * 2.2e-16 * std::max(scale, 1e-300);
  std::vector<double> b(n, 1.0), dl(n), dd(n), du(n), du2(n);
  for (int it = 0; it < 4; ++it) {
    // (T - shift I) x = b, Gaussian elimination with partial pivoting
    for (int i = 0; i < n; ++i) {
      dd[i] = d[i] - shift;
      if (i < n - 1) {
        dl[i] = e[i];
        du[i] = e[i];
      }
      du2[i] = 0.0;
    }
    std::vector<double> x(b);
    for (int i = 0; i < n - 1; ++i) {
      if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
        double f = (dd[i] != 0.0) ? dl[i] / dd[i] : 0.0;
        dd[i + 1] -= f * du[i];
        x[i + 1] -= f * x[i];
        dl[i] = 0.0;
      } else {
        double f = dd[i] / dl[i];
        dd[i] = dl[i];
        double t = dd[i + 1];
        dd[i + 1] = du[i] - f * t;
        if (i < n - 2) {
          du2[i] = du[i + 1];
          du[i + 1] = -f * du[i + 1];
        }
        du[i] = t;
        double tx = x[i];
        x[i] = x[i + 1];
        x[i + 1] = tx - f * x[i + 1];
      }
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = x[i];
      if (i < n - 1) s -= du[i] * x[i + 1];
      if (i < n - 2) s -= du2[i] * x[i + 2];
      double piv = dd[i];
      if (std::fabs(piv) < 1e-300) piv = 1e-300;
      x[i] = s / piv;
    }
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; ++i) b[i] = x[i] / nrm;
  }
  int first = 0;
  while (first < n - 1 && std::fabs(b[first]) < 1e-300) ++first;
  const double sg = b[first] < 0 ? -1.0 : 1.0;
  for (int i = 0; i < n; ++i) z[i] = sg * b[i];
}

}  // namespace tsc

using namespace tsc;

// ---------------------------------------------------------------------------
// Scratch memory of tsc_hpsi_noalloc: one static device buffer, which every
// device the library runs on holds until the process ends. On the host the
// array's name is only a stand-in; its address on a device is looked up once
// per device, with that device current, when the first context is created
// on it, and host code uses that address alone (kernel arguments, copies).
static __device__ __align__(256) unsigned char scratch[TSC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices] = {};

// the device of this call is current
int scratch_setup(int device) {
  TSC_REQUIRE(device >= 0 && device < kScratchDevices, "device index beyond the scratch address table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return TSC_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    return TSC_EHIP;
  }
  g_scratch_ptr[device] = ptr;
  return TSC_OK;
}

void *scratch_address(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace


extern "C" const char *tsc_last_error(void) { return g_err.c_str(); }
extern "C" int tsc_version(void) { return 1; }

extern "C" int tsc_ctx_mfma_layout(const tsc_ctx *c, int32_t *modes) {
  TSC_REQUIRE(c && modes, "tsc_ctx_mfma_layout arguments");
  for (int i = 0; i < 3; ++i) modes[i] = c->mfma_rowmode[i];
  return TSC_OK;
}

extern "C" int tsc_ctx_mem_info(const tsc_ctx *c, uint64_t *free_bytes, uint64_t *total_bytes) {
  TSC_REQUIRE(c && free_bytes && total_bytes, "tsc_ctx_mem_info arguments");
  TSC_HIP_CHECK(hipSetDevice(c->device));
  size_t f = 0, t = 0;
  TSC_HIP_CHECK(hipMemGetInfo(&f, &t));
  *free_bytes = (uint64_t)f;
  *total_bytes = (uint64_t)t;
  return TSC_OK;
}

extern "C" int tsc_ctx_create(int device, tsc_ctx **out) {
  TSC_REQUIRE(out, "tsc_ctx_create: null out");
  int ndev = 0;
  TSC_HIP_CHECK(hipGetDeviceCount(&ndev));
  TSC_REQUIRE(device >= 0 && device < ndev, "tsc_ctx_create: no such device");
  TSC_HIP_CHECK(hipSetDevice(device));
  {
    const int rc = scratch_setup(device);
    if (rc) return rc;
  }
  tsc_ctx *c = new tsc_ctx();
  c->device = device;
  TSC_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  TSC_HIP_CHECK(hipMalloc((void **)&c->d_red, sizeof(double) * kRedBlocks * kMaxVecs));
  TSC_HIP_CHECK(hipHostMalloc((void **)&c->h_red, sizeof(double) * kRedBlocks * kMaxVecs, hipHostMallocDefault));
  TSC_HIP_CHECK(hipMalloc((void **)&c->d_ptrs, sizeof(void *) * kMaxVecs));
  TSC_HIP_CHECK(hipMalloc((void **)&c->d_ptrs2, sizeof(void *) * kMaxVecs));
  TSC_HIP_CHECK(hipMalloc((void **)&c->d_coef, sizeof(double) * kMaxVecs));
  TSC_HIP_CHECK(hipHostMalloc((void **)&c->h_ptrs, sizeof(void *) * kMaxVecs, hipHostMallocDefault));
  TSC_HIP_CHECK(hipHostMalloc((void **)&c->h_coef, sizeof(double) * kMaxVecs, hipHostMallocDefault));
  {
    int rc = probe_layouts(c->stream, c->mfma_rowmode);
    if (rc) {
      tsc_ctx_destroy(c);
      return rc;
    }
  }
  *out = c;
  return TSC_OK;
}

extern "C" int tsc_ctx_destroy(tsc_ctx *c) {
  if (!c) return TSC_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipFree(c->d_red);
  hipHostFree(c->h_red);
  hipFree(c->d_ptrs);
  hipFree(c->d_ptrs2);
  hipFree(c->d_coef);
  hipHostFree(c->h_ptrs);
  hipHostFree(c->h_coef);
  hipStreamDestroy(c->stream);
  delete c;
  return TSC_OK;
}

extern "C" int tsc_ctx_sync(tsc_ctx *c) {
  TSC_REQUIRE(c, "tsc_ctx_sync: null ctx");
  TSC_HIP_CHECK(hipStreamSynchronize(c->stream));
  return TSC_OK;
}

// ---------------------------------------------------------------------------
// plan construction: validate every contribution against its buffer, build
// the tile lists (largest work first) and upload everything.
namespace {

struct StageCheck {
  int64_t size[4];
};

int check_stage(const tsc_contrib *c, int32_t nc, const tsc_outblock *o, int32_t no, const StageCheck &sz,
                int out_buf, const int *readable, double *flops) {
  for (int32_t b = 0; b < no; ++b) {
    const tsc_outblock &ob = o[b];
    TSC_REQUIRE(ob.c_buf == out_buf, "output block writes the wrong buffer");
    TSC_REQUIRE(ob.m >= 0 && ob.n >= 0 && ob.ldc >= std::max(1, ob.m), "output block shape");
    TSC_REQUIRE(ob.cbeg >= 0 && ob.cend >= ob.cbeg && ob.cend <= nc, "output block contribution range");
    TSC_REQUIRE(ob.mode == 0 || ob.mode == 1, "output block mode");
    TSC_REQUIRE((int64_t)ob.m * ob.n <= (int64_t)1 << 40, "output block too large");
    TSC_REQUIRE((ob.m + kTM - 1) / kTM < 32768 && (ob.n + kTN - 1) / kTN < 32768, "output block too many tiles");
    if (ob.m > 0 && ob.n > 0) {
      const int64_t last = ob.c_off + (int64_t)(ob.m - 1) + (int64_t)(ob.n - 1) * ob.ldc;
      TSC_REQUIRE(ob.c_off >= 0 && last < sz.size[out_buf], "output block outside its buffer");
    }
    for (int32_t ci = ob.cbeg; ci < ob.cend; ++ci) {
      const tsc_contrib &t = c[ci];
      TSC_REQUIRE(t.a_op >= 0 && t.a_op <= TSC_OP_I && t.b_op >= 0 && t.b_op <= TSC_OP_I, "bad op code");
      TSC_REQUIRE(!(t.a_op == TSC_OP_I && t.b_op == TSC_OP_I), "identity times identity");
      // stored shape of each operand: op(A) is m x k (or m x n when B = I)
      auto check_operand = [&](int buf, int64_t off, int ld, int op, int rows, int cols) -> int {
        if (op == TSC_OP_I) return TSC_OK;
        TSC_REQUIRE(buf >= 0 && buf < 4 && readable[buf], "operand reads a buffer this stage may not read");
        const bool tr = (op == TSC_OP_T || op == TSC_OP_C);
        const int sr = tr ? cols : rows, sc = tr ? rows : cols;
        if (sr == 0 || sc == 0) return TSC_OK;
        TSC_REQUIRE(ld >= std::max(1, sr), "operand leading dimension");
        const int64_t last = off + (int64_t)(sr - 1) + (int64_t)(sc - 1) * ld;
        TSC_REQUIRE(off >= 0 && last < sz.size[buf], "operand outside its buffer");
        return TSC_OK;
      };
      int rc;
      if (t.b_op == TSC_OP_I) {
        rc = check_operand(t.a_buf, t.a_off, t.lda, t.a_op, ob.m, ob.n);
        if (rc) return rc;
        *flops += 8.0 * ob.m * ob.n;
      } else if (t.a_op == TSC_OP_I) {
        rc = check_operand(t.b_buf, t.b_off, t.ldb, t.b_op, ob.m, ob.n);
        if (rc) return rc;
        *flops += 8.0 * ob.m * ob.n;
      } else {
        TSC_REQUIRE(t.k >= 0, "negative k");
        rc = check_operand(t.a_buf, t.a_off, t.lda, t.a_op, ob.m, t.k);
        if (rc) return rc;
        rc = check_operand(t.b_buf, t.b_off, t.ldb, t.b_op, t.k, ob.n);
        if (rc) return rc;
        *flops += 8.0 * ob.m * ob.n * (double)t.k;
      }
    }
  }
  return TSC_OK;
}

}  // namespace

namespace tsc {
// Work units of one stage: every 64x64 tile of every output block; tiles
// whose contribution list is long are cut into contiguous chunks of about
// `target` k-steps that write partial tiles (summed by a Reduce entry).
//
// Placement: the device hands workgroup b to XCD b % kXcds, and each XCD
// has its own L2. All units of one output block read the same operand
// windows, and output blocks whose smallest B operand window is the same
// (blocks of one environment window: they read the same operator blocks)
// do too; such blocks form one group that goes to a single XCD, units of
// the same contribution chunk next to each other. A block costing more
// than half an XCD's share is cut into near-square rectangles of tiles that
// form groups of their own. Groups (capped at a quarter of an XCD's share
// of the work) are dealt to the XCDs largest first onto the least-loaded
// one, and the per-XCD lists are interleaved (padded with empty units,
// blk = -1, that exit at once).
constexpr int kXcds = 8;  // MI355X: 256 CUs in 8 XCDs

Units make_tiles(const tsc_contrib *c, const tsc_outblock *o, int32_t no, int tm, int tn, double desired,
                 const std::vector<char> *include) {
  auto ccost = [&](int32_t ci) {
    return (c[ci].a_op == TSC_OP_I || c[ci].b_op == TSC_OP_I) ? 1.0 : (double)std::max(c[ci].k, 1);
  };
  double total = 0.0;
  int64_t ntile = 0;
  auto skip = [&](int32_t b) { return o[b].m == 0 || o[b].n == 0 || (include && !(*include)[b]); };
  for (int32_t b = 0; b < no; ++b) {
    if (skip(b)) continue;
    double w = 0.0;
    for (int32_t ci = o[b].cbeg; ci < o[b].cend; ++ci) w += ccost(ci);
    const int64_t nt = (int64_t)((o[b].m + tm - 1) / tm) * ((o[b].n + tn - 1) / tn);
    total += std::max(w, 1.0) * nt;
    ntile += nt;
  }
  // Cost (k) at which a tile's contribution list is cut into units: about
  // `desired` units for plans with fewer tiles, and in any case no unit
  // longer than kUnitCostCap, because the units of one output window that
  // are resident together drift apart over a long list and stop sharing the
  // XCD's L2. The partial tiles of cut units are capped at kMaxSlots: beyond
  // that the cut cost grows.
  constexpr double kUnitCostCap = 8192.0;
  constexpr int64_t kMaxSlots = 16384;
  auto slots_for = [&](double cut_cost) {
    int64_t slots = 0;
    for (int32_t b = 0; b < no; ++b) {
      if (skip(b) || o[b].cend - o[b].cbeg <= 1) continue;
      double w = 0.0;
      for (int32_t ci = o[b].cbeg; ci < o[b].cend; ++ci) w += ccost(ci);
      if (w <= cut_cost) continue;
      int64_t nch = 0;
      double acc = 0.0;
      for (int32_t ci = o[b].cbeg; ci < o[b].cend; ++ci) {
        acc += ccost(ci);
        if (acc >= cut_cost || ci + 1 == o[b].cend) {
          ++nch;
          acc = 0.0;
        }
      }
      slots += nch * (int64_t)((o[b].m + tm - 1) / tm) * ((o[b].n + tn - 1) / tn);
    }
    return slots;
  };
  double target = (ntile >= (int64_t)desired) ? kUnitCostCap
                                              : std::min(std::max(total / desired, 64.0), kUnitCostCap);
  for (int it = 0; it < 80 && slots_for(target) > kMaxSlots; ++it) target *= 1.25;
  Units u;
  struct Group {
    double cost = 0.0;
    std::vector<Tile> units;
  };
  std::vector<Group> groups;
  std::map<std::pair<int, int64_t>, size_t> open_group;  // smallest B window -> group being filled
  const double group_cap = total / (4.0 * kXcds);
  for (int32_t b = 0; b < no; ++b) {
    const tsc_outblock &ob = o[b];
    if (skip(b)) continue;
    double w = 0.0;
    std::pair<int, int64_t> key(1 << 30, 0);
    for (int32_t ci = ob.cbeg; ci < ob.cend; ++ci) {
      w += ccost(ci);
      if (c[ci].a_op != TSC_OP_I && c[ci].b_op != TSC_OP_I) key = std::min(key, std::make_pair((int)c[ci].b_buf, c[ci].b_off));
    }
    const int nti = (ob.m + tm - 1) / tm, ntj = (ob.n + tn - 1) / tn;
    std::vector<int32_t> cut{ob.cbeg};
    std::vector<double> ccut;
    const bool split = !(w <= target || ob.cend - ob.cbeg <= 1);
    if (!split) {
      cut.push_back(ob.cend);
      ccut.push_back(std::max(w, 1.0));
    } else {
      double acc = 0.0;
      for (int32_t ci = ob.cbeg; ci < ob.cend; ++ci) {
        acc += ccost(ci);
        if (acc >= target || ci + 1 == ob.cend) {
          cut.push_back(ci + 1);
          ccut.push_back(acc);
          acc = 0.0;
        }
      }
    }
    const int nch = (int)ccut.size();
    const int64_t slot_base = u.nslot;
    if (split) {
      for (int ti = 0; ti < nti; ++ti)
        for (int tj = 0; tj < ntj; ++tj)
          u.reds.push_back(Reduce{b, (int16_t)ti, (int16_t)tj, (int32_t)(slot_base + (int64_t)(ti * ntj + tj) * nch),
                                  nch});
      u.nslot += (int64_t)nti * ntj * nch;
    }
    auto add_units = [&](Group &g, int ti0, int ti1, int tj0, int tj1) {
      for (int ch = 0; ch < nch; ++ch)
        for (int ti = ti0; ti < ti1; ++ti)
          for (int tj = tj0; tj < tj1; ++tj) {
            const int32_t slot = split ? (int32_t)(slot_base + (int64_t)(ti * ntj + tj) * nch + ch) : -1;
            g.units.push_back(Tile{b, (int16_t)ti, (int16_t)tj, cut[ch], cut[ch + 1], slot, 0});
            g.cost += ccut[ch];
          }
    };
    double bcost = 0.0;
    for (double x : ccut) bcost += x;
    bcost *= (double)nti * ntj;
    if (bcost > 2.0 * group_cap && (int64_t)nti * ntj > 1) {
      // a block larger than half an XCD's share, which the dealing below could
      // not balance: near-square rectangles of tiles of at most a group's
      // cost, each a group of its own (one XCD reads the rectangle's A rows
      // and B columns; smaller blocks stay whole, their operands read by one
      // XCD only)
      const int64_t want = std::min<int64_t>((int64_t)nti * ntj, (int64_t)std::ceil(bcost / group_cap));
      int pr = (int)std::lround(std::sqrt((double)want * nti / ntj));
      pr = std::max(1, std::min(nti, pr));
      int pc = (int)std::min<int64_t>(ntj, (want + pr - 1) / pr);
      pc = std::max(1, pc);
      if ((int64_t)pr * pc < want) pr = (int)std::min<int64_t>(nti, (want + pc - 1) / pc);
      for (int r = 0; r < pr; ++r)
        for (int q = 0; q < pc; ++q) {
          const int ti0 = (int)((int64_t)r * nti / pr), ti1 = (int)((int64_t)(r + 1) * nti / pr);
          const int tj0 = (int)((int64_t)q * ntj / pc), tj1 = (int)((int64_t)(q + 1) * ntj / pc);
          if (ti0 == ti1 || tj0 == tj1) continue;
          groups.emplace_back();
          add_units(groups.back(), ti0, ti1, tj0, tj1);
        }
      continue;
    }
    auto it = open_group.find(key);
    if (it == open_group.end() || groups[it->second].cost >= group_cap) {
      groups.emplace_back();
      open_group[key] = groups.size() - 1;
      it = open_group.find(key);
    }
    add_units(groups[it->second], 0, nti, 0, ntj);
  }
  std::stable_sort(groups.begin(), groups.end(), [](const Group &x, const Group &y) { return x.cost > y.cost; });
  std::vector<std::vector<Tile>> lists(kXcds);
  std::vector<double> load(kXcds, 0.0);
  for (const Group &g : groups) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lists[x].insert(lists[x].end(), g.units.begin(), g.units.end());
    load[x] += g.cost;
  }
  size_t len = 0;
  for (const auto &l : lists) len = std::max(len, l.size());
  u.tiles.assign(len * kXcds, Tile{-1, 0, 0, 0, 0, -1, 0});
  for (int x = 0; x < kXcds; ++x)
    for (size_t i = 0; i < lists[x].size(); ++i) u.tiles[i * kXcds + x] = lists[x][i];
  while (!u.tiles.empty() && u.tiles.back().blk < 0) u.tiles.pop_back();
  return u;
}
}  // namespace tsc

namespace {

template <typename T>
int upload(T **dst, const T *src, size_t n, hipStream_t s) {
  if (n == 0) {
    *dst = nullptr;
    return TSC_OK;
  }
  TSC_HIP_CHECK(hipMalloc((void **)dst, sizeof(T) * n));
  TSC_HIP_CHECK(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
  return TSC_OK;
}

}  // namespace

extern "C" int tsc_plan_destroy(tsc_plan *p) {
  if (!p) return TSC_OK;
  hipSetDevice(p->ctx->device);
  hipStreamSynchronize(p->ctx->stream);
  if (p->own_ops) hipFree(p->d_ops64);
  hipFree(p->d_ops32);
  hipFree(p->d_mid64);
  hipFree(p->d_mid32);
  for (int s = 0; s < 2; ++s) {
    hipFree(p->d_c[s]);
    hipFree(p->d_o[s]);
    hipFree(p->d_t[s]);
    hipFree(p->d_r[s]);
  }
  hipFree(p->d_part32);
  hipFree(p->d_part64);
  tsc::pk_free(p->pk);
  tsc::gd_free(p->gd);
  for (auto v : p->kry) hipFree(v);
  hipFree(p->lz_io[0]);
  hipFree(p->lz_io[1]);
  for (int i = 0; i < 3; ++i) {
    if (p->ev[i]) hipEventDestroy(p->ev[i]);
    if (p->gexec[i]) hipGraphExecDestroy(p->gexec[i]);
    if (p->lz_gexec[i]) hipGraphExecDestroy(p->lz_gexec[i]);
  }
  delete p;
  return TSC_OK;
}

namespace tsc {
int plan_create_impl(tsc_ctx *ctx, const double *ops, const double2 *ops_dev, int64_t ops_elems, int64_t mid_elems,
                     const tsc_contrib *c1, int32_t nc1, const tsc_outblock *o1, int32_t no1, const tsc_contrib *c2,
                     int32_t nc2, const tsc_outblock *o2, int32_t no2, int64_t in_elems, int64_t out_elems,
                     tsc_plan **out) {
  TSC_REQUIRE(ctx && out, "tsc_plan_create: null ctx/out");
  TSC_REQUIRE(ops_elems >= 0 && mid_elems >= 0 && in_elems >= 0 && out_elems >= 0, "negative sizes");
  TSC_REQUIRE(nc1 >= 0 && no1 >= 0 && nc2 >= 0 && no2 >= 0, "negative counts");

  StageCheck sz{{ops_elems, in_elems, mid_elems, out_elems}};
  double flops = 0.0;
  {
    const int rd1[4] = {1, 1, 0, 0};
    int rc = check_stage(c1, nc1, o1, no1, sz, TSC_BUF_MID, rd1, &flops);
    if (rc) return rc;
    const int rd2[4] = {1, 1, 1, 0};
    rc = check_stage(c2, nc2, o2, no2, sz, TSC_BUF_OUT, rd2, &flops);
    if (rc) return rc;
  }
  TSC_HIP_CHECK(hipSetDevice(ctx->device));
  tsc_plan *p = new tsc_plan();
  p->ctx = ctx;
  p->ops_elems = ops_elems;
  p->mid_elems = mid_elems;
  p->in_elems = in_elems;
  p->out_elems = out_elems;
  p->flops = flops;
  int rc = TSC_OK;
  auto fail = [&](int r) {
    tsc_plan_destroy(p);
    return r;
  };
  if (ops_dev) {
    p->d_ops64 = const_cast<double2 *>(ops_dev);
    p->own_ops = false;
  } else if (ops) {
    if ((rc = upload<double2>(&p->d_ops64, reinterpret_cast<const double2 *>(ops), ops_elems, ctx->stream)))
      return fail(rc);
  } else if (ops_elems > 0) {
    hipError_t e = hipMalloc((void **)&p->d_ops64, sizeof(double2) * ops_elems);
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc(ops arena) failed: ") + hipGetErrorString(e));
      return fail(TSC_ENOMEM);
    }
    e = hipMemsetAsync(p->d_ops64, 0, sizeof(double2) * ops_elems, ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("hipMemsetAsync(ops arena) failed: ") + hipGetErrorString(e));
      return fail(TSC_EHIP);
    }
  }
  const tsc_contrib *cs[2] = {c1, c2};
  const tsc_outblock *os[2] = {o1, o2};
  const int32_t ncs[2] = {nc1, nc2}, nos[2] = {no1, no2};
  for (int s = 0; s < 2; ++s) {
    p->h_c[s].assign(cs[s], cs[s] + ncs[s]);
    p->h_o[s].assign(os[s], os[s] + nos[s]);
    if ((rc = upload<tsc_contrib>(&p->d_c[s], cs[s], ncs[s], ctx->stream))) return fail(rc);
    if ((rc = upload<tsc_outblock>(&p->d_o[s], os[s], nos[s], ctx->stream))) return fail(rc);
    Units un = make_tiles(cs[s], os[s], nos[s]);
    p->nt[s] = (int)un.tiles.size();
    p->nr[s] = (int)un.reds.size();
    p->nslot[s] = un.nslot;
    if ((rc = upload<Tile>(&p->d_t[s], un.tiles.data(), un.tiles.size(), ctx->stream))) return fail(rc);
    if ((rc = upload<Reduce>(&p->d_r[s], un.reds.data(), un.reds.size(), ctx->stream))) return fail(rc);
    TSC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  for (int i = 0; i < 3; ++i) TSC_HIP_CHECK(hipEventCreate(&p->ev[i]));
  *out = p;
  return TSC_OK;
}
}  // namespace tsc

extern "C" int tsc_plan_create(tsc_ctx *ctx, const double *ops, int64_t ops_elems, int64_t mid_elems,
                               const tsc_contrib *c1, int32_t nc1, const tsc_outblock *o1, int32_t no1,
                               const tsc_contrib *c2, int32_t nc2, const tsc_outblock *o2, int32_t no2,
                               int64_t in_elems, int64_t out_elems, tsc_plan **out) {
  return plan_create_impl(ctx, ops, nullptr, ops_elems, mid_elems, c1, nc1, o1, no1, c2, nc2, o2, no2, in_elems,
                          out_elems, out);
}

extern "C" int tsc_plan_create_dev(tsc_ctx *ctx, const tsc_vec *ops, int64_t mid_elems, const tsc_contrib *c1,
                                   int32_t nc1, const tsc_outblock *o1, int32_t no1, const tsc_contrib *c2,
                                   int32_t nc2, const tsc_outblock *o2, int32_t no2, int64_t in_elems,
                                   int64_t out_elems, tsc_plan **out) {
  TSC_REQUIRE(ctx && ops && ops->ctx == ctx, "tsc_plan_create_dev: operator arena of another context");
  return plan_create_impl(ctx, nullptr, ops->d, ops->n, mid_elems, c1, nc1, o1, no1, c2, nc2, o2, no2, in_elems,
                          out_elems, out);
}

extern "C" int tsc_plan_update_ops(tsc_plan *p, const double *ops) {
  TSC_REQUIRE(p && (ops || p->ops_elems == 0), "tsc_plan_update_ops arguments");
  if (p->ops_elems == 0) return TSC_OK;
  TSC_HIP_CHECK(hipMemcpyAsync(p->d_ops64, ops, sizeof(double2) * p->ops_elems, hipMemcpyHostToDevice, p->ctx->stream));
  if (p->d_ops32) {
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(4096), dim3(256), 0, p->ctx->stream, p->d_ops64, p->d_ops32,
                       p->ops_elems);
    TSC_HIP_CHECK(hipGetLastError());
  }
  int rc = tsc::pk_repack_ops(p);
  if (rc) return rc;
  TSC_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_plan_fill_ops_gaussian(tsc_plan *p, int64_t off, int64_t count, uint64_t seed, uint64_t key,
                                          double sigma) {
  TSC_REQUIRE(p && off >= 0 && count >= 0 && off + count <= p->ops_elems, "tsc_plan_fill_ops_gaussian range");
  TSC_REQUIRE(p->own_ops, "tsc_plan_fill_ops_gaussian: borrowed arena");
  int rc = fill_gaussian(p->ctx, p->d_ops64 + off, count, seed, key, sigma);
  if (rc) return rc;
  if (p->d_ops32) {
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(4096), dim3(256), 0, p->ctx->stream, p->d_ops64 + off,
                       p->d_ops32 + off, count);
    TSC_HIP_CHECK(hipGetLastError());
  }
  rc = tsc::pk_repack_ops(p);
  if (rc) return rc;
  TSC_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  return TSC_OK;
}

extern "C" int tsc_plan_flops(const tsc_plan *p, double *flops) {
  TSC_REQUIRE(p && flops, "tsc_plan_flops arguments");
  *flops = p->flops;
  return TSC_OK;
}

namespace {
int ensure_precision(tsc_plan *p, int prec) {
  hipStream_t s = p->ctx->stream;
  if (prec == TSC_PREC_BF16X3) {
    if (p->ctx->mfma_rowmode[2] != 0) {
      set_error("BF16x3 path: this device's 32x32x16 bf16 MFMA accumulator layout is not the one the kernel uses");
      return TSC_EHIP;
    }
    return pk_build(p);
  }
  if (prec == TSC_PREC_F64) {
    if (!p->d_mid64 && p->mid_elems > 0) TSC_HIP_CHECK(hipMalloc((void **)&p->d_mid64, sizeof(double2) * p->mid_elems));
    if (p->f64_kernel == 1) {
      const int64_t slots64 = std::max(p->nslot[0], p->nslot[1]);
      if (slots64 > 0 && !p->d_part64)
        TSC_HIP_CHECK(hipMalloc(&p->d_part64, sizeof(double2) * (size_t)kTM * kTN * (size_t)slots64));
      return TSC_OK;
    }
    return gd_build(p);
  }
  const int64_t slots = std::max(p->nslot[0], p->nslot[1]);
  if (slots > 0 && !p->d_part32)
    TSC_HIP_CHECK(hipMalloc(&p->d_part32, sizeof(float2) * (size_t)kTM * kTN * (size_t)slots));
  {
    if (!p->d_ops32 && p->ops_elems > 0) {
      TSC_HIP_CHECK(hipMalloc((void **)&p->d_ops32, sizeof(float2) * p->ops_elems));
      hipLaunchKernelGGL(f64_to_f32_kernel, dim3(4096), dim3(256), 0, s, p->d_ops64, p->d_ops32, p->ops_elems);
      TSC_HIP_CHECK(hipGetLastError());
    }
    if (!p->d_mid32 && p->mid_elems > 0) TSC_HIP_CHECK(hipMalloc((void **)&p->d_mid32, sizeof(float2) * p->mid_elems));
  }
  return TSC_OK;
}

int enqueue_hpsi(tsc_plan *p, const double2 *in, double2 *out, int prec, bool events) {
  hipStream_t s = p->ctx->stream;
  if (prec == TSC_PREC_BF16X3) return pk_enqueue(p, in, out, events);
  if (prec == TSC_PREC_F64 && p->f64_kernel == 0) return gd_enqueue(p, in, out, events);
  const bool f64 = prec == TSC_PREC_F64;
  Bufs b;
  b.p[TSC_BUF_OPS] = f64 ? (const void *)p->d_ops64 : (const void *)p->d_ops32;
  b.p[TSC_BUF_IN] = in;
  b.p[TSC_BUF_MID] = f64 ? (const void *)p->d_mid64 : (const void *)p->d_mid32;
  b.p[TSC_BUF_OUT] = out;
  if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[0], s));
  void *part = f64 ? p->d_part64 : p->d_part32;
  TSC_HIP_CHECK(launch_grouped(prec, b, p->d_o[0], p->d_c[0], p->d_t[0], p->nt[0], p->d_r[0], p->nr[0], part, s,
                               p->ctx->mfma_rowmode));
  if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[1], s));
  TSC_HIP_CHECK(launch_grouped(prec, b, p->d_o[1], p->d_c[1], p->d_t[1], p->nt[1], p->d_r[1], p->nr[1], part, s,
                               p->ctx->mfma_rowmode));
  if (events) TSC_HIP_CHECK(hipEventRecord(p->ev[2], s));
  return TSC_OK;
}
}  // namespace

namespace tsc {
int hpsi_dev(tsc_plan *p, const double2 *in, double2 *out, int prec) {
  TSC_REQUIRE(prec >= TSC_PREC_F64 && prec <= TSC_PREC_BF16X3, "unknown precision");
  int rc = ensure_precision(p, prec);
  if (rc) return rc;
  rc = enqueue_hpsi(p, in, out, prec, true);
  if (rc) return rc;
  p->timed = true;
  return TSC_OK;
}
}  // namespace tsc

extern "C" int tsc_hpsi(tsc_plan *p, const tsc_vec *in, tsc_vec *out, int prec) {
  TSC_REQUIRE(p && in && out, "tsc_hpsi arguments");
  TSC_REQUIRE(in->n >= p->in_elems && out->n >= p->out_elems, "tsc_hpsi: vector shorter than the plan needs");
  TSC_REQUIRE(in->d != out->d || p->in_elems == 0, "tsc_hpsi: in and out must differ");
  return hpsi_dev(p, in->d, out->d, prec);
}

extern "C" int tsc_hpsi_host(tsc_plan *p, const double *in, double *out, int prec) {
  TSC_REQUIRE(p && (in || p->in_elems == 0) && (out || p->out_elems == 0), "tsc_hpsi_host arguments");
  tsc_vec *vi = nullptr, *vo = nullptr;
  int rc = tsc_vec_create(p->ctx, p->in_elems, &vi);
  if (rc) return rc;
  rc = tsc_vec_create(p->ctx, p->out_elems, &vo);
  if (!rc) rc = tsc_vec_upload(vi, in);
  if (!rc) rc = tsc_hpsi(p, vi, vo, prec);
  if (!rc) rc = tsc_vec_download(vo, out);
  tsc_vec_destroy(vi);
  tsc_vec_destroy(vo);
  return rc;
}

// FP64 H.psi on the plan's 64x64 grouped tiles with the intermediates and
// the partial tiles in the static scratch buffer: [intermediates | partial
// tiles], each part starting on a 256-byte boundary. Nothing there has to
// start at zero: the stage-1 blocks write every intermediate window in full
// before stage 2 reads it, and each split unit writes every in-block element
// of its partial tile before the reduction reads it (the plan-owned buffers
// of tsc_hpsi are uninitialised device memory too). No allocation, no
// library call that allocates, no synchronisation.
extern "C" int tsc_hpsi_noalloc(tsc_plan *p, const double *in_c128, double *out_c128) {
  TSC_REQUIRE(p && p->ctx, "tsc_hpsi_noalloc: null plan");
  TSC_REQUIRE((in_c128 || p->in_elems == 0) && (out_c128 || p->out_elems == 0), "tsc_hpsi_noalloc: null vector");
  TSC_REQUIRE(in_c128 != out_c128 || p->in_elems == 0 || p->out_elems == 0, "tsc_hpsi_noalloc: in and out must differ");
  void *base = scratch_address(p->ctx->device);
  TSC_REQUIRE(base, "tsc_hpsi_noalloc: no scratch address for the plan's device");
  const int64_t mid_bytes = (p->mid_elems * (int64_t)sizeof(double2) + 255) & ~(int64_t)255;
  const int64_t slots = std::max(p->nslot[0], p->nslot[1]);
  const int64_t part_bytes = slots * (int64_t)kTM * kTN * (int64_t)sizeof(double2);
  if (mid_bytes + part_bytes > (int64_t)TSC_NOALLOC_SCRATCH_BYTES) {
    set_error("invalid: tsc_hpsi_noalloc: the plan needs " + std::to_string(mid_bytes) +
              " bytes of intermediates and " + std::to_string(part_bytes) + " bytes of partial tiles; the scratch "
              "buffer holds " + std::to_string((int64_t)TSC_NOALLOC_SCRATCH_BYTES) + " bytes");
    return TSC_EINVAL;
  }
  Bufs b;
  b.p[TSC_BUF_OPS] = p->d_ops64;
  b.p[TSC_BUF_IN] = in_c128;
  b.p[TSC_BUF_MID] = base;
  b.p[TSC_BUF_OUT] = out_c128;
  void *part = static_cast<unsigned char *>(base) + mid_bytes;
  hipStream_t s = p->ctx->stream;
  for (int g = 0; g < 2; ++g)
    TSC_HIP_CHECK(launch_grouped(TSC_PREC_F64, b, p->d_o[g], p->d_c[g], p->d_t[g], p->nt[g], p->d_r[g], p->nr[g],
                                 part, s, p->ctx->mfma_rowmode));
  return TSC_OK;
}

extern "C" int tsc_hpsi_graph(tsc_plan *p, const tsc_vec *in, tsc_vec *out, int prec) {
  TSC_REQUIRE(p && in && out, "tsc_hpsi_graph arguments");
  TSC_REQUIRE(prec >= TSC_PREC_F64 && prec <= TSC_PREC_BF16X3, "unknown precision");
  TSC_REQUIRE(in->n >= p->in_elems && out->n >= p->out_elems && in->d != out->d, "tsc_hpsi_graph vectors");
  int rc = ensure_precision(p, prec);
  if (rc) return rc;
  hipStream_t s = p->ctx->stream;
  if (!p->gexec[prec] || p->g_in[prec] != in->d || p->g_out[prec] != out->d) {
    if (p->gexec[prec]) {
      TSC_HIP_CHECK(hipGraphExecDestroy(p->gexec[prec]));
      p->gexec[prec] = nullptr;
    }
    TSC_HIP_CHECK(hipStreamSynchronize(s));
    hipGraph_t g;
    TSC_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = enqueue_hpsi(p, in->d, out->d, prec, false);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
      return TSC_EHIP;
    }
    TSC_HIP_CHECK(hipGraphInstantiate(&p->gexec[prec], g, nullptr, nullptr, 0));
    TSC_HIP_CHECK(hipGraphDestroy(g));
    p->g_in[prec] = in->d;
    p->g_out[prec] = out->d;
  }
  TSC_HIP_CHECK(hipEventRecord(p->ev[0], s));
  TSC_HIP_CHECK(hipGraphLaunch(p->gexec[prec], s));
  TSC_HIP_CHECK(hipEventRecord(p->ev[2], s));
  p->timed = false;
  return TSC_OK;
}

namespace {
// tuned choices (work-unit target, or the FP64 tile shape), shared by plans
// of the same shape: precision, 2 log2 flops, 2 log2 stage-2 output
// blocks, and the sector-size histogram of the work, as the share (in
// tenths) of the stage-2 flops in output blocks whose larger side is < 64,
// < 128, < 256, < 512 and >= 512
typedef std::array<int, 8> TuneKey;
std::map<TuneKey, double> g_tune_cache;
std::mutex g_tune_mu;

TuneKey tune_key(const tsc_plan *p, int prec) {
  double share[5] = {0, 0, 0, 0, 0}, tot = 0.0;
  const std::vector<tsc_outblock> &o = p->h_o[1];
  const std::vector<tsc_contrib> &c = p->h_c[1];
  for (const tsc_outblock &ob : o) {
    double f = 0.0;
    for (int k = ob.cbeg; k < ob.cend; ++k) f += (double)ob.m * ob.n * std::max(1, c[k].k);
    const int big = std::max(ob.m, ob.n);
    const int cls = big < 64 ? 0 : big < 128 ? 1 : big < 256 ? 2 : big < 512 ? 3 : 4;
    share[cls] += f;
    tot += f;
  }
  TuneKey k;
  k[0] = prec;
  k[1] = (int)std::lround(2.0 * std::log2(std::max(p->flops, 1.0)));
  k[2] = (int)std::lround(2.0 * std::log2(std::max<double>((double)o.size(), 1.0)));
  for (int i = 0; i < 5; ++i) k[3 + i] = tot > 0.0 ? (int)std::lround(10.0 * share[i] / tot) : 0;
  return k;
}

void drop_graph(tsc_plan *p, int prec) {
  if (p->gexec[prec]) hipGraphExecDestroy(p->gexec[prec]);
  p->gexec[prec] = nullptr;
  p->g_in[prec] = p->g_out[prec] = nullptr;
  if (p->lz_gexec[prec]) hipGraphExecDestroy(p->lz_gexec[prec]);
  p->lz_gexec[prec] = nullptr;
}
}  // namespace

extern "C" int tsc_plan_autotune(tsc_plan *p, const tsc_vec *in, tsc_vec *out, int prec, int flags,
                                 double *chosen) {
  TSC_REQUIRE(p && in && out && chosen, "tsc_plan_autotune arguments");
  TSC_REQUIRE(prec >= TSC_PREC_F64 && prec <= TSC_PREC_BF16X3, "unknown precision");
  TSC_REQUIRE(in->n >= p->in_elems && out->n >= p->out_elems && in->d != out->d, "tsc_plan_autotune vectors");
  int rc = ensure_precision(p, prec);
  if (rc) return rc;
  if (prec == TSC_PREC_F32) {  // the FP32 path keeps its fixed units
    *chosen = 0.0;
    return TSC_OK;
  }
  // a candidate: a work-unit target of the 128-row units (>= 64), or 1 =
  // FP64 on the 64x64 grouped tiles (the other tile shape of that path)
  auto retile = [&](double d) {
    if (prec == TSC_PREC_F64) {
      p->f64_kernel = d == 1.0 ? 1 : 0;
      int r = ensure_precision(p, prec);
      if (r || d == 1.0) return r;
      return gd_retile(p, d);
    }
    return pk_retile(p, d);
  };
  if (flags & 2) {
    // a given target, no timing
    const double d = *chosen;
    TSC_REQUIRE((d >= 64.0 && d <= 1e7) || (d == 1.0 && prec == TSC_PREC_F64),
                "tsc_plan_autotune: work-unit target outside [64, 1e7] (or 1: FP64 64x64 tiles)");
    rc = retile(d);
    drop_graph(p, prec);
    return rc;
  }
  const auto key = tune_key(p, prec);
  if (!(flags & 1)) {
    double v = 0.0;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) v = it->second;
    }
    if (v > 0.0) {
      rc = retile(v);
      drop_graph(p, prec);
      *chosen = v;
      return rc;
    }
  }
  // time each candidate: one H.psi to settle, then three on the plan's events
  std::vector<double> cand = {2048.0, 4096.0, 8192.0};
  if (prec == TSC_PREC_F64) cand.push_back(1.0);
  hipStream_t s = p->ctx->stream;
  double best = cand[0], best_ms = 1e300;
  for (double d : cand) {
    if ((rc = retile(d))) return rc;
    if ((rc = enqueue_hpsi(p, in->d, out->d, prec, false))) return rc;
    TSC_HIP_CHECK(hipEventRecord(p->ev[0], s));
    for (int r = 0; r < 3; ++r)
      if ((rc = enqueue_hpsi(p, in->d, out->d, prec, false))) return rc;
    TSC_HIP_CHECK(hipEventRecord(p->ev[2], s));
    TSC_HIP_CHECK(hipEventSynchronize(p->ev[2]));
    float ms = 0.f;
    TSC_HIP_CHECK(hipEventElapsedTime(&ms, p->ev[0], p->ev[2]));
    if (ms < best_ms) {
      best_ms = ms;
      best = d;
    }
  }
  if ((rc = retile(best))) return rc;
  drop_graph(p, prec);
  p->timed = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  *chosen = best;
  return TSC_OK;
}

extern "C" int tsc_plan_last_times(const tsc_plan *p, float *ms1, float *ms2) {
  TSC_REQUIRE(p && ms1 && ms2, "tsc_plan_last_times arguments");
  TSC_HIP_CHECK(hipEventSynchronize(p->ev[2]));
  if (p->timed) {
    TSC_HIP_CHECK(hipEventElapsedTime(ms1, p->ev[0], p->ev[1]));
    TSC_HIP_CHECK(hipEventElapsedTime(ms2, p->ev[1], p->ev[2]));
  } else {
    float t = 0.f;
    TSC_HIP_CHECK(hipEventElapsedTime(&t, p->ev[0], p->ev[2]));
    *ms1 = 0.f;
    *ms2 = t;
  }
  return TSC_OK;
}

// ---------------------------------------------------------------------------
// Lanczos with full re-orthogonalisation, following lanczos_memo_new
// (hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/eigen.f90:4-199) for levels = 1:
// up to max_outer+1 passes of at most max_inner Krylov vectors; inner
// products keep the real part only (wave_product, block.f90:1092-1108);
// classical Gram-Schmidt against every previous vector then normalise
// (Gram_Schmidt_process, block.f90:1251-1289); convergence when the lowest
// Ritz value moves by <= tol, or the residual norm beta <= tol.
namespace {

struct LanczosResult {
  double energy;
  int matvecs;
  int restarts;
  // alpha_1 = <v_1|H v_1> and beta_1 = |H v_1 - alpha_1 v_1| of the last
  // outer iteration (v_1: its normalised start vector), and the weight of
  // the returned Ritz vector outside v_1 (1 - z_1^2, z the normalised
  // tridiagonal eigenvector)
  double alpha1 = 0.0, beta1 = 0.0, w_rest = 0.0;
};

// Plans below this many flops per H.psi replay their products from a
// captured graph inside the Lanczos loop: their launches, not their
// kernels, set the time. Larger plans launch eagerly (the two vector
// copies a replay needs would cost more than the launches they save).
constexpr double kLanczosGraphFlops = 2.0e11;

// out = H in for Lanczos: eager, or a replay of the graph captured on the
// plan's two fixed vectors with `in` copied in and the result copied out
int lanczos_hpsi(tsc_plan *p, const double2 *in, double2 *out, int prec, bool replay) {
  if (!replay) return hpsi_dev(p, in, out, prec);
  hipStream_t s = p->ctx->stream;
  const int64_t n = p->kry_n;
  if (!p->lz_io[0]) {
    TSC_HIP_CHECK(hipMalloc((void **)&p->lz_io[0], sizeof(double2) * std::max<int64_t>(n, 1)));
    TSC_HIP_CHECK(hipMalloc((void **)&p->lz_io[1], sizeof(double2) * std::max<int64_t>(n, 1)));
  }
  if (!p->lz_gexec[prec]) {
    int rc = ensure_precision(p, prec);
    if (rc) return rc;
    TSC_HIP_CHECK(hipStreamSynchronize(s));
    hipGraph_t g;
    TSC_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = enqueue_hpsi(p, p->lz_io[0], p->lz_io[1], prec, false);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
      return TSC_EHIP;
    }
    e = hipGraphInstantiate(&p->lz_gexec[prec], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    TSC_HIP_CHECK(e);
  }
  TSC_HIP_CHECK(hipMemcpyAsync(p->lz_io[0], in, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
  TSC_HIP_CHECK(hipGraphLaunch(p->lz_gexec[prec], s));
  TSC_HIP_CHECK(hipMemcpyAsync(out, p->lz_io[1], sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
  p->timed = false;
  ++p->lz_replays;
  return TSC_OK;
}

// prec: the first H.psi of each restart (from the start or Ritz vector);
// prec_rest: the products of the later Krylov vectors
int lanczos_pass(tsc_plan *p, tsc_vec *psi, double tol, int max_inner, int max_outer, int prec, int prec_rest,
                 LanczosResult *res) {
  tsc_ctx *ctx = p->ctx;
  const int64_t n = psi->n;
  TSC_REQUIRE(max_inner >= 1 && max_inner + 1 <= kMaxVecs, "max_inner out of range");
  TSC_REQUIRE(n == p->in_elems && n == p->out_elems, "psi size does not match the plan");
  // Krylov vectors stay allocated with the plan between calls
  if (p->kry_n != n) {
    hipStreamSynchronize(ctx->stream);
    for (auto v : p->kry) hipFree(v);
    p->kry.clear();
    hipFree(p->lz_io[0]);
    hipFree(p->lz_io[1]);
    p->lz_io[0] = p->lz_io[1] = nullptr;
    for (int i = 0; i < 3; ++i) {
      if (p->lz_gexec[i]) hipGraphExecDestroy(p->lz_gexec[i]);
      p->lz_gexec[i] = nullptr;
    }
    p->kry_n = n;
  }
  const bool replay = p->flops < kLanczosGraphFlops && std::getenv("TSC_LANCZOS_EAGER") == nullptr;
  while ((int)p->kry.size() < max_inner + 1) {
    double2 *v = nullptr;
    TSC_HIP_CHECK(hipMalloc((void **)&v, sizeof(double2) * std::max<int64_t>(n, 1)));
    p->kry.push_back(v);
  }
  std::vector<double2 *> V(p->kry.begin(), p->kry.begin() + max_inner + 1);
  auto free_all = [&]() { hipStreamSynchronize(ctx->stream); };
  std::vector<double> alpha(max_inner + 1, 0.0), beta(max_inner + 1, 0.0), z(max_inner + 1), uu(max_inner + 1, 1.0);
  double lastmin = 1024.0;
  int matvecs = 0, iter_out = 0, last = 1;
  bool done = false;
  int rc = TSC_OK;
#define LZ_CALL(x)        \
  do {                    \
    rc = (x);             \
    if (rc) {             \
      free_all();         \
      return rc;          \
    }                     \
  } while (0)
  while (iter_out <= max_outer && !done) {
    ++iter_out;
    last = 1;
    std::fill(alpha.begin(), alpha.end(), 0.0);
    std::fill(beta.begin(), beta.end(), 0.0);
    double ritzmin = lastmin;
    // v1 = psi / |psi|
    TSC_HIP_CHECK(hipMemcpyAsync(V[0], psi->d, sizeof(double2) * n, hipMemcpyDeviceToDevice, ctx->stream));
    {
      double nn;
      const double2 *vv[1] = {V[0]};
      LZ_CALL(dots_re(ctx, vv, 1, V[0], n, &nn));
      LZ_CALL(scale(ctx, V[0], 1.0 / std::sqrt(nn), n));
    }
    {
      double nn;
      const double2 *vv[1] = {V[0]};
      LZ_CALL(dots_re(ctx, vv, 1, V[0], n, &nn));
      uu[0] = nn;
    }
    LZ_CALL(lanczos_hpsi(p, V[0], V[1], prec, replay));
    ++matvecs;
    {
      const double2 *vv[1] = {V[0]};
      LZ_CALL(dots_re(ctx, vv, 1, V[1], n, &alpha[1]));
    }
    ritzmin = alpha[1];
    res->alpha1 = alpha[1];
    res->beta1 = 0.0;
    bool converged_inner = false;
    while (last < max_inner) {
      // w = H v_last - alpha v_last - beta v_{last-1}   (w lives in V[last])
      {
        const double2 *vv[2] = {V[last - 1], last >= 2 ? V[last - 2] : nullptr};
        const double cc[2] = {-alpha[last], last >= 2 ? -beta[last - 1] : 0.0};
        LZ_CALL(combine(ctx, V[last], 1.0, vv, cc, last >= 2 ? 2 : 1, n));
      }
      {
        double nn;
        const double2 *vv[1] = {V[last]};
        LZ_CALL(dots_re(ctx, vv, 1, V[last], n, &nn));
        beta[last] = std::sqrt(nn);
      }
      if (last == 1) res->beta1 = beta[1];
      if (std::fabs(beta[last]) <= tol) {
        converged_inner = true;
        done = true;
        break;
      }
      LZ_CALL(scale(ctx, V[last], 1.0 / beta[last], n));
      // classical Gram-Schmidt with real overlaps, then normalise
      {
        std::vector<const double2 *> vv(V.begin(), V.begin() + last);
        std::vector<double> uv(last), cc(last);
        LZ_CALL(dots_re(ctx, vv.data(), last, V[last], n, uv.data()));
        // <u_j|u_j> of the stored Krylov vectors (prod_uu of Gram_Schmidt_process)
        for (int j = 0; j < last; ++j) cc[j] = -uv[j] / uu[j];
        LZ_CALL(combine(ctx, V[last], 1.0, vv.data(), cc.data(), last, n));
        double nn;
        const double2 *me[1] = {V[last]};
        LZ_CALL(dots_re(ctx, me, 1, V[last], n, &nn));
        LZ_CALL(scale(ctx, V[last], 1.0 / std::sqrt(nn), n));
        LZ_CALL(dots_re(ctx, me, 1, V[last], n, &nn));
        uu[last] = nn;
      }
      ++last;
      LZ_CALL(lanczos_hpsi(p, V[last - 1], V[last], prec_rest, replay));
      ++matvecs;
      {
        const double2 *vv[1] = {V[last - 1]};
        LZ_CALL(dots_re(ctx, vv, 1, V[last], n, &alpha[last]));
      }
      const double e0 = tridiag_lowest(&alpha[1], &beta[1], last);
      if (std::fabs(e0 - ritzmin) <= tol || (last == max_inner && iter_out == max_outer)) {
        lastmin = e0;
        converged_inner = true;
        done = true;
        break;
      }
      ritzmin = e0;
      lastmin = ritzmin;
    }
    // Ritz vector of the current Krylov basis -> psi (restart or final)
    const double e0 = tridiag_lowest(&alpha[1], &beta[1], last);
    tridiag_vector(&alpha[1], &beta[1], last, e0, z.data());
    {
      double zz = 0.0;
      for (int i = 0; i < last; ++i) zz += z[i] * z[i];
      res->w_rest = zz > 0.0 ? std::max(0.0, 1.0 - z[0] * z[0] / zz) : 0.0;
    }
    {
      std::vector<const double2 *> vv(V.begin(), V.begin() + last);
      LZ_CALL(combine(ctx, psi->d, 0.0, vv.data(), z.data(), last, n));
      double nn;
      const double2 *me[1] = {psi->d};
      LZ_CALL(dots_re(ctx, me, 1, psi->d, n, &nn));
      LZ_CALL(scale(ctx, psi->d, 1.0 / std::sqrt(nn), n));
    }
    lastmin = e0;
    res->energy = e0;
    (void)converged_inner;
  }
#undef LZ_CALL
  TSC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  free_all();
  res->matvecs = matvecs;
  res->restarts = iter_out - 1;
  return TSC_OK;
}

}  // namespace

extern "C" int tsc_lanczos(tsc_plan *p, tsc_vec *psi, double tol, int max_inner, int max_outer, int prec,
                           int refine_iters, double *energy, int32_t *info) {
  TSC_REQUIRE(p && psi && energy, "tsc_lanczos arguments");
  TSC_REQUIRE(max_outer >= 0 && refine_iters >= 0, "tsc_lanczos iteration counts");
  LanczosResult r{0.0, 0, 0};
  p->lz_replays = 0;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (double &x : p->lz_refine) x = nan;
  int rc = lanczos_pass(p, psi, tol, max_inner, max_outer, prec, prec, &r);
  if (rc) return rc;
  p->lz_refine[0] = r.energy;
  int mv = r.matvecs, rs = r.restarts;
  if (prec != TSC_PREC_F64 && refine_iters > 0) {
    // Refinement: a Lanczos pass from the Ritz vector psi whose first
    // product is FP64. It fixes alpha_1 = <psi|H psi> and the residual
    // r = H psi - alpha_1 psi (v_2 = r / beta_1) to FP64; the later products
    // only enter the lowest Ritz value through the weight of v_2.. in it,
    // O(beta_1^2 / gap^2), so their mixed-precision rounding moves the energy
    // by O(beta_1^2 / gap^2 * eps * |H|) and they run in the working precision.
    LanczosResult r2{0.0, 0, 0};
    rc = lanczos_pass(p, psi, tol, std::max(2, refine_iters), 0, TSC_PREC_F64, prec, &r2);
    if (rc) return rc;
    p->lz_refine[1] = r2.alpha1;
    p->lz_refine[2] = r2.beta1;
    p->lz_refine[3] = r2.energy;
    p->lz_refine[4] = r2.w_rest;
    r.energy = r2.energy;
    mv += r2.matvecs;
  }
  *energy = r.energy;
  if (info) {
    info[0] = mv;
    info[1] = rs;
  }
  return TSC_OK;
}

extern "C" int tsc_lanczos_refinement(const tsc_plan *p, double *out) {
  TSC_REQUIRE(p && out, "tsc_lanczos_refinement arguments");
  for (int i = 0; i < 5; ++i) out[i] = p->lz_refine[i];
  return TSC_OK;
}

extern "C" int tsc_lanczos_replays(const tsc_plan *p, int32_t *replays) {
  TSC_REQUIRE(p && replays, "tsc_lanczos_replays arguments");
  *replays = p->lz_replays;
  return TSC_OK;
}
