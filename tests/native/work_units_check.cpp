// Host-side check of the work-unit builder (tsc::make_tiles, library
// internal): on random output-block lists, every (block, tile row, tile
// column) gets units whose contribution ranges tile the block's range
// exactly once, and the per-XCD work of the interleaved grid (workgroup b
// runs on XCD b % 8) stays balanced, also for one block much larger than an
// XCD's share. Prints "ok" or the first failure; exit status 0 / 1.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <tuple>
#include <vector>

#include "tsc_internal.h"

using tsc::Tile;
using tsc::Units;

namespace {

double unit_cost(const std::vector<tsc_contrib> &c, const Tile &t) {
  double w = 0.0;
  for (int ci = t.cbeg; ci < t.cend; ++ci)
    w += (c[ci].a_op == TSC_OP_I || c[ci].b_op == TSC_OP_I) ? 1.0 : (double)std::max(c[ci].k, 1);
  return w;
}

// 0 on success; fills the max/mean ratio of the per-XCD work
int check(const std::vector<tsc_contrib> &c, const std::vector<tsc_outblock> &o, int tm, int tn, double desired,
          double *imbalance) {
  const Units u = tsc::make_tiles(c.data(), o.data(), (int)o.size(), tm, tn, desired);
  std::map<std::tuple<int, int, int, int>, int> units;  // (blk, ti, tj, cbeg) -> cend
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < u.tiles.size(); ++i) {
    const Tile &t = u.tiles[i];
    if (t.blk < 0) continue;
    if (!units.emplace(std::make_tuple(t.blk, (int)t.ti, (int)t.tj, t.cbeg), t.cend).second) {
      std::printf("duplicate unit: block %d tile (%d, %d) from %d\n", t.blk, t.ti, t.tj, t.cbeg);
      return 1;
    }
    load[i % 8] += unit_cost(c, t);
  }
  size_t covered = 0;
  for (int b = 0; b < (int)o.size(); ++b) {
    if (o[b].m == 0 || o[b].n == 0) continue;
    const int nti = (o[b].m + tm - 1) / tm, ntj = (o[b].n + tn - 1) / tn;
    for (int ti = 0; ti < nti; ++ti)
      for (int tj = 0; tj < ntj; ++tj)
        for (int pos = o[b].cbeg; pos < o[b].cend || (pos == o[b].cbeg && o[b].cbeg == o[b].cend);) {
          auto it = units.find(std::make_tuple(b, ti, tj, pos));
          if (it == units.end()) {
            std::printf("no unit: block %d tile (%d, %d) at contribution %d\n", b, ti, tj, pos);
            return 1;
          }
          ++covered;
          if (it->second == pos) break;  // empty range
          pos = it->second;
        }
  }
  if (covered != units.size()) {
    std::printf("%zu units, %zu expected\n", units.size(), covered);
    return 1;
  }
  double mx = 0.0, sum = 0.0;
  for (double x : load) {
    mx = std::max(mx, x);
    sum += x;
  }
  *imbalance = sum > 0.0 ? mx / (sum / 8.0) : 1.0;
  return 0;
}

tsc_contrib gemm(int k, int64_t b_off) {
  tsc_contrib t{};
  t.k = k;
  t.a_buf = TSC_BUF_OPS;
  t.b_buf = TSC_BUF_IN;
  t.a_op = TSC_OP_N;
  t.b_op = TSC_OP_N;
  t.b_off = b_off;
  t.coef_re = 1.0;
  return t;
}

}  // namespace

int main() {
  std::mt19937 rng(20261204);
  double worst = 0.0;
  for (int trial = 0; trial < 150; ++trial) {
    std::vector<tsc_contrib> c;
    std::vector<tsc_outblock> o;
    const int nb = 1 + (int)(rng() % 40);
    const int dmax = (trial % 3 == 0) ? 5000 : 900;
    for (int b = 0; b < nb; ++b) {
      tsc_outblock ob{};
      ob.m = 1 + (int)(rng() % dmax);
      ob.n = 1 + (int)(rng() % dmax);
      ob.ldc = ob.m;
      ob.c_buf = TSC_BUF_OUT;
      ob.cbeg = (int)c.size();
      const int nc = 1 + (int)(rng() % 30);
      for (int q = 0; q < nc; ++q) {
        tsc_contrib t = gemm(1 + (int)(rng() % 2000), (int64_t)(rng() % 7));
        if (rng() % 10 == 0) t.a_op = TSC_OP_I;
        c.push_back(t);
      }
      ob.cend = (int)c.size();
      o.push_back(ob);
    }
    for (int shape = 0; shape < 3; ++shape) {
      const int tm = shape == 0 ? 64 : 128, tn = shape == 1 ? 64 : (shape == 0 ? 64 : 128);
      double imb = 0.0;
      if (check(c, o, tm, tn, 2048.0 * (1 + trial % 4), &imb)) return 1;
      // balance is asked of the lists with large blocks (enough work to go round)
      if (trial % 3 == 0) worst = std::max(worst, imb);
    }
  }
  // one block far larger than an XCD's share: it must spread over all XCDs
  {
    std::vector<tsc_contrib> c{gemm(4096, 0)};
    tsc_outblock ob{};
    ob.m = ob.n = ob.ldc = 4096;
    ob.c_buf = TSC_BUF_OUT;
    ob.cbeg = 0;
    ob.cend = 1;
    std::vector<tsc_outblock> o{ob};
    double imb = 0.0;
    if (check(c, o, 128, 128, 2048.0, &imb)) return 1;
    if (imb > 1.2) {
      std::printf("one 4096 x 4096 block: per-XCD work max/mean %.3f\n", imb);
      return 1;
    }
  }
  if (worst > 1.35) {
    std::printf("per-XCD work max/mean %.3f on lists with large blocks\n", worst);
    return 1;
  }
  std::printf("ok (worst per-XCD max/mean %.3f)\n", worst);
  return 0;
}
