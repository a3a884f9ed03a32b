// Native SU(2) superblock planner: turns block descriptors and a term list
// into the task arrays of the grouped contraction engine (host code).
//
// Replaces the per-call work of model_wave's kernels in
// hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/get_sys.f90:1783-5204 (linear
// quantum-number searches, 6j look-ups, per-pair GEMM calls): for every
// term [X(L) x Y(R)]^0 it enumerates the sub-block transitions of the L side
// (operator on the block or on the site of L = block (x) site) and of the R
// side, combines them over the super basis with the 6j recoupling factor of
// the cut, and emits one contribution per target window, sharing the
// stage-1 intermediate X.psi between all targets that need it.
//
// Conventions (same as the host mirror in hamiltonian.py): quantum numbers
// (N, 2S), site basis electron (1,1) then hole (0,0), enlarged basis order of
// Get_basis (block.f90:2087-2189), super basis of Get_super_basis
// (block.f90:2194-2245), c~ from c+ by <j'||c~||j> = (-1)^(j-j'+1/2)
// conj(<j||c+||j'>), fermion sign (-1)^N of the part an odd operator passes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <array>
#include <tuple>
#include <unordered_map>

#include "tsc_internal.h"

namespace {

using tsc::set_error;

// ---------------------------------------------------------------------
// recoupling coefficients (Edmonds 7.1.5-7.1.8), doubled spins
struct SixJCache {
  std::unordered_map<uint64_t, double> m;
  double operator()(int a, int b, int c, int d, int e, int f) {
    const uint64_t k = ((uint64_t)a << 50) | ((uint64_t)b << 40) | ((uint64_t)c << 30) | ((uint64_t)d << 20) |
                       ((uint64_t)e << 10) | (uint64_t)f;
    auto it = m.find(k);
    if (it != m.end()) return it->second;
    const double v = tsc_w6j(a, b, c, d, e, f);
    m.emplace(k, v);
    return v;
  }
};

inline double par(int x2) { return ((x2 / 2) % 2) ? -1.0 : 1.0; }

struct Coef {
  SixJCache w;
  double embed_block(int ja2, int ja1, int js, int J2, int J1, int k2) {
    const double v = w(ja2, J2, js, J1, ja1, k2);
    return v == 0.0 ? 0.0 : par(ja2 + js + J1 + k2) * std::sqrt((double)(J1 + 1) * (J2 + 1)) * v;
  }
  double embed_site(int js2, int js1, int ja, int J2, int J1, int k2) {
    const double v = w(js2, J2, ja, J1, js1, k2);
    return v == 0.0 ? 0.0 : par(ja + js1 + J2 + k2) * std::sqrt((double)(J1 + 1) * (J2 + 1)) * v;
  }
  double scalar_pair(int ja2, int ja1, int js2, int js1, int J, int k2) {
    const double v = w(ja2, ja1, k2, js1, js2, J);
    return v == 0.0 ? 0.0 : par(ja1 + js2 + k2 + J) / std::sqrt((double)k2 + 1) * v;
  }
  double cut_pair(int jx2, int jx1, int jy2, int jy1, int S2, int k2) {
    const double v = w(jx2, jx1, k2, jy1, jy2, S2);
    return v == 0.0 ? 0.0 : par(jx1 + jy2 + k2 + S2) / std::sqrt((double)k2 + 1) * v;
  }
  static double ctilde_phase(int j_ket, int j_bra) { return par(j_ket - j_bra + 1); }
};

// site of the t-J model: electron, hole
const int SITE_N[2] = {1, 0};
const int SITE_J[2] = {1, 0};
// reduced site operators: c+ (e<-h), S (e<-e), n (e<-e)
struct SiteEl {
  int bra, ket;
  double v;
};
const SiteEl SITE_C = {0, 1, -1.4142135623730951};
const SiteEl SITE_S = {0, 0, 1.2247448713915890};
const SiteEl SITE_NUM = {0, 0, 1.4142135623730951};
const int SITE_K2[3] = {1, 2, 0};
const int SITE_DN[3] = {1, 0, 0};

// ---------------------------------------------------------------------
struct Sub {
  int a, s, pos, d;
};
struct EnlSector {
  int n, j, dim;
  std::vector<Sub> subs;
};
struct Basis {
  std::vector<EnlSector> sec;
  std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> where;  // (a,s) -> [(x,pos)]
};

Basis get_basis(const std::vector<tsc_sector> &blk) {
  // pre-list, QuickSort_Integer (2S desc, then N desc), unique
  std::vector<std::pair<int, int>> pre;
  for (const auto &b : blk)
    for (int s = 0; s < 2; ++s)
      for (int j = std::abs(b.twos - SITE_J[s]); j <= b.twos + SITE_J[s]; j += 2) pre.push_back({b.n + SITE_N[s], j});
  const size_t n = pre.size();
  for (size_t i = 0; i < n; ++i)
    for (size_t x = i + 1; x < n; ++x)
      if (pre[i].second < pre[x].second) std::swap(pre[i], pre[x]);
  for (size_t i = 0; i < n; ++i)
    for (size_t x = i + 1; x < n; ++x)
      if (pre[i].second == pre[x].second && pre[i].first < pre[x].first) std::swap(pre[i], pre[x]);
  Basis B;
  for (size_t i = 0; i < n; ++i) {
    if (!B.sec.empty() && B.sec.back().n == pre[i].first && B.sec.back().j == pre[i].second) continue;
    EnlSector e;
    e.n = pre[i].first;
    e.j = pre[i].second;
    e.dim = 0;
    for (int a = 0; a < (int)blk.size(); ++a)
      for (int s = 0; s < 2; ++s)
        if (blk[a].n + SITE_N[s] == e.n && std::abs(blk[a].twos - SITE_J[s]) <= e.j && e.j <= blk[a].twos + SITE_J[s]) {
          e.subs.push_back(Sub{a, s, e.dim, blk[a].dim});
          e.dim += blk[a].dim;
        }
    B.sec.push_back(e);
  }
  for (int x = 0; x < (int)B.sec.size(); ++x)
    for (const auto &sb : B.sec[x].subs) B.where[{sb.a, sb.s}].push_back({x, sb.pos});
  return B;
}

// matrix reference used by a transition: stored block at arena offset,
// leading dimension, and the op code turning it into the L-side factor
struct MatRef {
  int64_t off;
  int ld;
  int op;  // TSC_OP_N or TSC_OP_C (c~ from c+)
};

struct Tr {
  int x, p, d, x2, p2, d2;
  double f;
  int mat;  // index into mats, -1 = identity on the block part
};

struct Ctx {
  const tsc_sb_desc *D;
  Coef C;
  std::vector<MatRef> mats;
  std::map<std::tuple<int64_t, int>, int> mat_index;
  int matref(int64_t off, int ld, int op) {
    auto key = std::make_tuple(off, op);
    auto it = mat_index.find(key);
    if (it != mat_index.end()) return it->second;
    mats.push_back(MatRef{off, ld, op});
    mat_index[key] = (int)mats.size() - 1;
    return (int)mats.size() - 1;
  }
};

// operator entries (bra, ket, arena offset, ld, op, phase) of op or of its c~
struct Ent {
  int bra, ket;
  int64_t off;
  int ld, op;
  double ph;
};

std::vector<Ent> entries(const tsc_op_desc &op, const std::vector<tsc_sector> &sec, bool tilde) {
  std::vector<Ent> out;
  for (int i = 0; i < op.nblk; ++i) {
    const tsc_op_block &b = op.blk[i];
    const int ld = std::max(1, sec[b.bra].dim);
    if (!tilde)
      out.push_back(Ent{b.bra, b.ket, b.off, ld, TSC_OP_N, 1.0});
    else
      out.push_back(Ent{b.ket, b.bra, b.off, ld, TSC_OP_C, Coef::ctilde_phase(sec[b.bra].twos, sec[b.ket].twos)});
  }
  return out;
}

// transitions of a block operator embedded in block (x) site
void blk_transitions(Ctx &cx, const Basis &B, const std::vector<tsc_sector> &sec, const tsc_op_desc &op, bool tilde,
                     std::vector<Tr> &out) {
  for (const Ent &e : entries(op, sec, tilde)) {
    const int ja2 = sec[e.bra].twos, ja = sec[e.ket].twos;
    const int m = cx.matref(e.off, e.ld, e.op);
    for (int s = 0; s < 2; ++s) {
      auto it1 = B.where.find({e.ket, s});
      auto it2 = B.where.find({e.bra, s});
      if (it1 == B.where.end() || it2 == B.where.end()) continue;
      for (auto [x, p] : it1->second)
        for (auto [x2, p2] : it2->second) {
          const double f = cx.C.embed_block(ja2, ja, SITE_J[s], B.sec[x2].j, B.sec[x].j, op.k2) * e.ph;
          if (f != 0.0) out.push_back(Tr{x, p, sec[e.ket].dim, x2, p2, sec[e.bra].dim, f, m});
        }
    }
  }
}

// transitions of a site operator (0 c+, 1 S, 2 n) embedded in block (x) site
void site_transitions(Ctx &cx, const Basis &B, const std::vector<tsc_sector> &sec, int which, bool tilde,
                      std::vector<Tr> &out) {
  const SiteEl el = which == 0 ? SITE_C : (which == 1 ? SITE_S : SITE_NUM);
  const int k2 = SITE_K2[which];
  const bool odd = SITE_DN[which] % 2 != 0;
  int s2 = el.bra, s1 = el.ket;
  double v = el.v, ph = 1.0;
  if (tilde) {
    std::swap(s1, s2);
    ph = Coef::ctilde_phase(SITE_J[el.bra], SITE_J[el.ket]);
  }
  for (int a = 0; a < (int)sec.size(); ++a) {
    const double sgn = (odd && (sec[a].n % 2)) ? -1.0 : 1.0;
    auto it1 = B.where.find({a, s1});
    auto it2 = B.where.find({a, s2});
    if (it1 == B.where.end() || it2 == B.where.end()) continue;
    for (auto [x, p] : it1->second)
      for (auto [x2, p2] : it2->second) {
        const double f = cx.C.embed_site(SITE_J[s2], SITE_J[s1], sec[a].twos, B.sec[x2].j, B.sec[x].j, k2) * v * ph * sgn;
        if (f != 0.0) out.push_back(Tr{x, p, sec[a].dim, x2, p2, sec[a].dim, f, -1});
      }
  }
}

// scalar side terms: block Hamiltonian and [T(block) x U(site)]^0 pairs
void scalar_transitions(Ctx &cx, const Basis &B, const std::vector<tsc_sector> &sec, const tsc_op_desc *ham,
                        const tsc_pair_desc *pairs, int npairs, const tsc_op_desc *ops, bool reduced,
                        std::vector<Tr> &out) {
  if (ham) {
    for (int i = 0; i < ham->nblk; ++i) {
      const tsc_op_block &b = ham->blk[i];
      if (b.bra != b.ket) continue;
      const int a = b.ket;
      const int m = cx.matref(b.off, std::max(1, sec[a].dim), TSC_OP_N);
      for (int x = 0; x < (int)B.sec.size(); ++x)
        for (const Sub &sb : B.sec[x].subs)
          if (sb.a == a) {
            const double sc = reduced ? std::sqrt((double)B.sec[x].j + 1) : 1.0;
            out.push_back(Tr{x, sb.pos, sb.d, x, sb.pos, sb.d, sc / std::sqrt((double)sec[a].twos + 1), m});
          }
    }
  }
  for (int q = 0; q < npairs; ++q) {
    const tsc_pair_desc &pd = pairs[q];
    const tsc_op_desc &T = ops[pd.block_op];
    const SiteEl el = pd.site_op == 0 ? SITE_C : (pd.site_op == 1 ? SITE_S : SITE_NUM);
    const bool uodd = SITE_DN[pd.site_op] % 2 != 0;
    int us2 = el.bra, us1 = el.ket;
    double uph = 1.0;
    if (pd.site_tilde) {
      std::swap(us1, us2);
      uph = Coef::ctilde_phase(SITE_J[el.bra], SITE_J[el.ket]);
    }
    const std::vector<Ent> te = entries(T, sec, pd.block_tilde != 0);
    for (int x = 0; x < (int)B.sec.size(); ++x) {
      const int J = B.sec[x].j;
      const double sc = reduced ? std::sqrt((double)J + 1) : 1.0;
      for (const Sub &sb : B.sec[x].subs) {
        if (sb.s != us1) continue;
        const double sgn = (uodd && (sec[sb.a].n % 2)) ? -1.0 : 1.0;
        for (const Ent &e : te) {
          if (e.ket != sb.a) continue;
          const Sub *tgt = nullptr;
          for (const Sub &t2 : B.sec[x].subs)
            if (t2.a == e.bra && t2.s == us2) tgt = &t2;
          if (!tgt) continue;
          double f = cx.C.scalar_pair(sec[e.bra].twos, sec[sb.a].twos, SITE_J[us2], SITE_J[us1], J, T.k2);
          f *= pd.coef * el.v * e.ph * uph * sgn * sc;
          if (f == 0.0) continue;
          out.push_back(Tr{x, sb.pos, sb.d, x, tgt->pos, tgt->d, f, cx.matref(e.off, e.ld, e.op)});
        }
      }
    }
  }
}

void identity_transitions(const Basis &B, std::vector<Tr> &out) {
  for (int x = 0; x < (int)B.sec.size(); ++x)
    for (const Sub &sb : B.sec[x].subs) out.push_back(Tr{x, sb.pos, sb.d, x, sb.pos, sb.d, 1.0, -1});
}

}  // namespace

// ---------------------------------------------------------------------
struct tsc_tasks {
  std::vector<tsc_contrib> c1, c2;
  std::vector<tsc_outblock> o1, o2;
  int64_t mid_len = 0;
  int64_t psi_len = 0;
  int64_t arena_len = 0;           // operator arena length the descriptors address
  std::vector<int32_t> super_xy;   // (x, y, dx, dy) per psi sector
  std::vector<int32_t> lsec, rsec; // (n, 2J, dim) of the enlarged bases
  // algorithmic flops with every (L transition, R transition, source
  // window) product charged at its cheaper order and nothing shared
  double flops_pairmin = 0.0;
  // input / output lengths of the plan (psi_len for superblocks)
  int64_t in_len = 0, out_len = 0;
  // operator arena built by the host logic (a copy of the caller's arena
  // followed by merged operators); used when a plan is made without an arena
  std::vector<double2> arena;
  // renormalisation output layout: (operator, new bra sector, new ket
  // sector) per window with its offset; operator -1 = Hamiltonian, else the
  // carried operator 3 * (kept-site position) + (0 c+, 1 S, 2 n)
  std::vector<int32_t> layout;
  std::vector<int64_t> layout_off;
  std::vector<int32_t> new_sec;  // (N, 2S, kept) of the new block's sectors
};

namespace {

struct MidKey {
  int mat, i, p, q;
  bool operator<(const MidKey &o) const {
    return mat != o.mat ? mat < o.mat : (i != o.i ? i < o.i : (p != o.p ? p < o.p : q < o.q));
  }
};

struct Builder {
  tsc_tasks *T;
  std::vector<std::vector<tsc_contrib>> per_out;
  std::map<MidKey, int64_t> mid;    // (L matrix, psi sector, row pos, col pos) -> MID offset of O_L psi
  std::map<MidKey, int64_t> mid_r;  // (R matrix, psi sector, row pos, col pos) -> MID offset of psi O_R^T

  static tsc_contrib make(int64_t aoff, int lda, int abuf, int aop, int64_t boff, int ldb, int bbuf, int bop, int k,
                          double f) {
    tsc_contrib c;
    std::memset(&c, 0, sizeof(c));
    c.a_off = aoff;
    c.lda = lda;
    c.a_buf = (int8_t)abuf;
    c.a_op = (int8_t)aop;
    c.b_off = boff;
    c.ldb = ldb;
    c.b_buf = (int8_t)bbuf;
    c.b_op = (int8_t)bop;
    c.k = k;
    c.coef_re = f;
    return c;
  }
};

}  // namespace

extern "C" int tsc_superblock_tasks(const tsc_sb_desc *D, tsc_tasks **out) {
  TSC_REQUIRE(D && out, "tsc_superblock_tasks arguments");
  TSC_REQUIRE(D->nsys >= 0 && D->nenv >= 0 && D->nops >= 0 && D->nterms >= 0, "negative counts");
  std::vector<tsc_sector> A(D->sys, D->sys + D->nsys), Bv(D->env, D->env + D->nenv);
  for (int i = 0; i < D->nops; ++i) {
    const tsc_op_desc &op = D->ops[i];
    const std::vector<tsc_sector> &sec = op.side == 0 ? A : Bv;
    for (int b = 0; b < op.nblk; ++b) {
      const tsc_op_block &bl = op.blk[b];
      TSC_REQUIRE(bl.bra >= 0 && bl.bra < (int)sec.size() && bl.ket >= 0 && bl.ket < (int)sec.size(),
                  "operator block sector index");
      TSC_REQUIRE(bl.off >= 0 && bl.off + (int64_t)sec[bl.bra].dim * sec[bl.ket].dim <= D->arena_len,
                  "operator block outside the arena");
    }
  }
  Ctx cx;
  cx.D = D;
  Basis L = get_basis(A), R = get_basis(Bv);
  tsc_tasks *T = new tsc_tasks();
  T->arena_len = D->arena_len;
  // super basis
  std::vector<std::array<int, 4>> sb;
  std::map<std::pair<int, int>, int> sbi;
  std::vector<int64_t> soff;
  int64_t off = 0;
  for (int x = 0; x < (int)L.sec.size(); ++x)
    for (int y = 0; y < (int)R.sec.size(); ++y)
      if (R.sec[y].n == D->ntot - L.sec[x].n && std::abs(L.sec[x].j - R.sec[y].j) <= D->s2tot &&
          D->s2tot <= L.sec[x].j + R.sec[y].j) {
        sbi[{x, y}] = (int)sb.size();
        sb.push_back({x, y, L.sec[x].dim, R.sec[y].dim});
        soff.push_back(off);
        off += (int64_t)L.sec[x].dim * R.sec[y].dim;
      }
  T->psi_len = off;
  T->in_len = T->out_len = off;
  for (auto &e : sb) T->super_xy.insert(T->super_xy.end(), e.begin(), e.end());
  for (auto &e : L.sec) T->lsec.insert(T->lsec.end(), {e.n, e.j, e.dim});
  for (auto &e : R.sec) T->rsec.insert(T->rsec.end(), {e.n, e.j, e.dim});
  // output windows
  // window of (psi sector i, row offset p, column offset q); p, q < 2^24, i < 2^16 checked below
  std::unordered_map<uint64_t, int> win;
  auto wkey = [](int i, int p, int q) { return ((uint64_t)i << 48) | ((uint64_t)p << 24) | (uint64_t)q; };
  for (const auto &e : sb)
    if (e[2] >= (1 << 24) || e[3] >= (1 << 24) || sb.size() >= (1u << 16)) {
      delete T;
      set_error("superblock too large for the window index");
      return TSC_EINVAL;
    }
  for (int i = 0; i < (int)sb.size(); ++i) {
    const int x = sb[i][0], y = sb[i][1], dx = sb[i][2];
    for (const Sub &ls : L.sec[x].subs)
      for (const Sub &rs : R.sec[y].subs) {
        tsc_outblock ob;
        std::memset(&ob, 0, sizeof(ob));
        ob.c_off = soff[i] + ls.pos + (int64_t)rs.pos * dx;
        ob.ldc = std::max(1, dx);
        ob.m = ls.d;
        ob.n = rs.d;
        ob.c_buf = TSC_BUF_OUT;
        win[wkey(i, ls.pos, rs.pos)] = (int)T->o2.size();
        T->o2.push_back(ob);
      }
  }
  Builder bd;
  bd.T = T;
  bd.per_out.resize(T->o2.size());

  // contraction order of the two-sided products (both sides a stored
  // matrix): "left" forms O_L psi first, "right" psi O_R^T first, "auto"
  // (default) takes per source window and term whichever costs fewer
  // flops, intermediates shared between the products of the window (and
  // ones already formed by earlier terms) counted once. TSC_PLAN_ORDER
  // overrides it (tests compare the orders).
  int order_mode = 0;
  if (const char *e = std::getenv("TSC_PLAN_ORDER")) {
    if (std::strcmp(e, "left") == 0) order_mode = 1;
    else if (std::strcmp(e, "right") == 0) order_mode = 2;
  }
  struct PairRec {
    const Tr *tl, *tr;
    double f;
    int w;
  };
  auto cross = [&](const std::vector<Tr> &TL, const std::vector<Tr> &TR, double g, bool cut, int k2, bool y_odd) {
    std::vector<std::vector<const Tr *>> byx(L.sec.size()), byy(R.sec.size());
    for (const Tr &t : TL) byx[t.x].push_back(&t);
    for (const Tr &t : TR) byy[t.x].push_back(&t);
    std::vector<PairRec> both;
    for (int i = 0; i < (int)sb.size(); ++i) {
      const int x = sb[i][0], y = sb[i][1], dx = sb[i][2];
      if (byx[x].empty() || byy[y].empty()) continue;
      const int Jx = L.sec[x].j, Jy = R.sec[y].j;
      const double fsgn = (y_odd && (L.sec[x].n % 2)) ? -1.0 : 1.0;
      both.clear();
      for (const Tr *tl : byx[x])
        for (const Tr *tr : byy[y]) {
          auto it = sbi.find({tl->x2, tr->x2});
          if (it == sbi.end()) continue;
          const int i2 = it->second;
          double f = g * tl->f * tr->f;
          if (cut) f *= cx.C.cut_pair(L.sec[tl->x2].j, Jx, R.sec[tr->x2].j, Jy, D->s2tot, k2) * fsgn;
          if (f == 0.0) continue;
          const int w = win[wkey(i2, tl->p2, tr->p2)];
          const int64_t src = soff[i] + tl->p + (int64_t)tr->p * dx;
          if (tl->mat >= 0 && tr->mat >= 0) {
            both.push_back(PairRec{tl, tr, f, w});
          } else if (tl->mat >= 0) {
            const MatRef &lm = cx.mats[tl->mat];
            bd.per_out[w].push_back(Builder::make(lm.off, lm.ld, TSC_BUF_OPS, lm.op, src, std::max(1, dx), TSC_BUF_IN,
                                                  TSC_OP_N, tl->d, f));
            T->flops_pairmin += 8.0 * tl->d2 * tl->d * tr->d;
          } else if (tr->mat >= 0) {
            const MatRef &rm = cx.mats[tr->mat];
            bd.per_out[w].push_back(Builder::make(src, std::max(1, dx), TSC_BUF_IN, TSC_OP_N, rm.off, rm.ld,
                                                  TSC_BUF_OPS, rm.op == TSC_OP_N ? TSC_OP_T : TSC_OP_CONJ, tr->d, f));
            T->flops_pairmin += 8.0 * tl->d * tr->d * tr->d2;
          } else {
            bd.per_out[w].push_back(Builder::make(src, std::max(1, dx), TSC_BUF_IN, TSC_OP_N, 0, 1, 0, TSC_OP_I, 0, f));
            T->flops_pairmin += 8.0 * tl->d * tr->d;
          }
        }
      // two-sided products, one group per source window (row pos, col pos)
      std::stable_sort(both.begin(), both.end(), [](const PairRec &u, const PairRec &v) {
        return u.tl->p != v.tl->p ? u.tl->p < v.tl->p : u.tr->p < v.tr->p;
      });
      for (size_t g0 = 0; g0 < both.size();) {
        size_t g1 = g0;
        while (g1 < both.size() && both[g1].tl->p == both[g0].tl->p && both[g1].tr->p == both[g0].tr->p) ++g1;
        const int p = both[g0].tl->p, q = both[g0].tr->p;
        const double dl = both[g0].tl->d, dr = both[g0].tr->d;
        const int64_t src = soff[i] + p + (int64_t)q * dx;
        double cl = 0.0, cr = 0.0;
        std::vector<int> seen_l, seen_r;
        for (size_t k = g0; k < g1; ++k) {
          const Tr *tl = both[k].tl, *tr = both[k].tr;
          const double d2l = tl->d2, d2r = tr->d2;
          if (std::find(seen_l.begin(), seen_l.end(), tl->mat) == seen_l.end()) {
            seen_l.push_back(tl->mat);
            if (!bd.mid.count(MidKey{tl->mat, i, p, q})) cl += d2l * dl * dr;
          }
          if (std::find(seen_r.begin(), seen_r.end(), tr->mat) == seen_r.end()) {
            seen_r.push_back(tr->mat);
            if (!bd.mid_r.count(MidKey{tr->mat, i, p, q})) cr += dl * dr * d2r;
          }
          cl += d2l * dr * d2r;
          cr += d2l * dl * d2r;
          T->flops_pairmin += 8.0 * std::min(d2l * dl * dr + d2l * dr * d2r, dl * dr * d2r + d2l * dl * d2r);
        }
        const bool left = order_mode == 1 || (order_mode == 0 && cl <= cr);
        for (size_t k = g0; k < g1; ++k) {
          const Tr *tl = both[k].tl, *tr = both[k].tr;
          const MatRef &lm = cx.mats[tl->mat];
          const MatRef &rm = cx.mats[tr->mat];
          const int rop = rm.op == TSC_OP_N ? TSC_OP_T : TSC_OP_CONJ;
          if (left) {
            // MID = O_L psi_window (d2l x dr); out += f MID op(O_R)^T
            const MidKey key{tl->mat, i, p, q};
            int64_t moff;
            auto mit = bd.mid.find(key);
            if (mit != bd.mid.end()) {
              moff = mit->second;
            } else {
              moff = T->mid_len;
              T->mid_len += (int64_t)tl->d2 * tr->d;
              bd.mid.emplace(key, moff);
              tsc_outblock ob;
              std::memset(&ob, 0, sizeof(ob));
              ob.c_off = moff;
              ob.ldc = std::max(1, tl->d2);
              ob.m = tl->d2;
              ob.n = tr->d;
              ob.c_buf = TSC_BUF_MID;
              ob.cbeg = (int32_t)T->c1.size();
              ob.cend = ob.cbeg + 1;
              T->o1.push_back(ob);
              T->c1.push_back(Builder::make(lm.off, lm.ld, TSC_BUF_OPS, lm.op, src, std::max(1, dx), TSC_BUF_IN,
                                            TSC_OP_N, tl->d, 1.0));
            }
            bd.per_out[both[k].w].push_back(Builder::make(moff, std::max(1, tl->d2), TSC_BUF_MID, TSC_OP_N, rm.off,
                                                          rm.ld, TSC_BUF_OPS, rop, tr->d, both[k].f));
          } else {
            // MID = psi_window op(O_R)^T (dl x d2r); out += f O_L MID
            const MidKey key{tr->mat, i, p, q};
            int64_t moff;
            auto mit = bd.mid_r.find(key);
            if (mit != bd.mid_r.end()) {
              moff = mit->second;
            } else {
              moff = T->mid_len;
              T->mid_len += (int64_t)tl->d * tr->d2;
              bd.mid_r.emplace(key, moff);
              tsc_outblock ob;
              std::memset(&ob, 0, sizeof(ob));
              ob.c_off = moff;
              ob.ldc = std::max(1, tl->d);
              ob.m = tl->d;
              ob.n = tr->d2;
              ob.c_buf = TSC_BUF_MID;
              ob.cbeg = (int32_t)T->c1.size();
              ob.cend = ob.cbeg + 1;
              T->o1.push_back(ob);
              T->c1.push_back(Builder::make(src, std::max(1, dx), TSC_BUF_IN, TSC_OP_N, rm.off, rm.ld, TSC_BUF_OPS, rop,
                                            tr->d, 1.0));
            }
            bd.per_out[both[k].w].push_back(Builder::make(lm.off, lm.ld, TSC_BUF_OPS, lm.op, moff, std::max(1, tl->d),
                                                          TSC_BUF_MID, TSC_OP_N, tl->d, both[k].f));
          }
        }
        g0 = g1;
      }
    }
  };

  std::vector<Tr> idL, idR, sL, sR;
  identity_transitions(L, idL);
  identity_transitions(R, idR);
  scalar_transitions(cx, L, A, D->sys_ham, D->lpairs, D->nlpairs, D->ops, false, sL);
  scalar_transitions(cx, R, Bv, D->env_ham, D->rpairs, D->nrpairs, D->ops, false, sR);
  cross(sL, idR, 1.0, false, 0, false);
  cross(idL, sR, 1.0, false, 0, false);
  for (int t = 0; t < D->nterms; ++t) {
    const tsc_term_desc &td = D->terms[t];
    std::vector<Tr> TL, TR;
    int k2 = 0;
    bool yodd = false;
    if (td.l_kind == 1) {
      blk_transitions(cx, L, A, D->ops[td.l_op], td.l_tilde != 0, TL);
      k2 = D->ops[td.l_op].k2;
    } else {
      site_transitions(cx, L, A, td.l_op, td.l_tilde != 0, TL);
      k2 = SITE_K2[td.l_op];
    }
    if (td.r_kind == 1) {
      blk_transitions(cx, R, Bv, D->ops[td.r_op], td.r_tilde != 0, TR);
      yodd = D->ops[td.r_op].up_dif % 2 != 0;
    } else {
      site_transitions(cx, R, Bv, td.r_op, td.r_tilde != 0, TR);
      yodd = SITE_DN[td.r_op] % 2 != 0;
    }
    cross(TL, TR, td.coef, true, k2, yodd);
  }
  for (size_t w = 0; w < T->o2.size(); ++w) {
    T->o2[w].cbeg = (int32_t)T->c2.size();
    T->c2.insert(T->c2.end(), bd.per_out[w].begin(), bd.per_out[w].end());
    T->o2[w].cend = (int32_t)T->c2.size();
  }
  *out = T;
  return TSC_OK;
}

extern "C" int tsc_tasks_info(const tsc_tasks *T, int64_t *counts) {
  TSC_REQUIRE(T && counts, "tsc_tasks_info arguments");
  counts[0] = (int64_t)T->c1.size();
  counts[1] = (int64_t)T->o1.size();
  counts[2] = (int64_t)T->c2.size();
  counts[3] = (int64_t)T->o2.size();
  counts[4] = T->mid_len;
  counts[5] = T->psi_len;
  counts[6] = (int64_t)T->super_xy.size() / 4;
  counts[7] = (int64_t)T->lsec.size() / 3;
  counts[8] = (int64_t)T->rsec.size() / 3;
  return TSC_OK;
}

extern "C" int tsc_tasks_flops(const tsc_tasks *T, double *flops2) {
  TSC_REQUIRE(T && flops2, "tsc_tasks_flops arguments");
  double ex = 0.0;
  for (int s = 0; s < 2; ++s) {
    const std::vector<tsc_contrib> &c = s == 0 ? T->c1 : T->c2;
    for (const tsc_outblock &ob : (s == 0 ? T->o1 : T->o2))
      for (int k = ob.cbeg; k < ob.cend; ++k)
        ex += 8.0 * ob.m * ob.n * ((c[k].a_op == TSC_OP_I || c[k].b_op == TSC_OP_I) ? 1.0 : (double)c[k].k);
  }
  flops2[0] = ex;
  flops2[1] = T->flops_pairmin;
  return TSC_OK;
}

extern "C" int tsc_tasks_copy(const tsc_tasks *T, tsc_contrib *c1, tsc_outblock *o1, tsc_contrib *c2,
                              tsc_outblock *o2, int32_t *super_xy, int32_t *lsec, int32_t *rsec) {
  TSC_REQUIRE(T, "tsc_tasks_copy: null tasks");
  if (c1) std::copy(T->c1.begin(), T->c1.end(), c1);
  if (o1) std::copy(T->o1.begin(), T->o1.end(), o1);
  if (c2) std::copy(T->c2.begin(), T->c2.end(), c2);
  if (o2) std::copy(T->o2.begin(), T->o2.end(), o2);
  if (super_xy) std::copy(T->super_xy.begin(), T->super_xy.end(), super_xy);
  if (lsec) std::copy(T->lsec.begin(), T->lsec.end(), lsec);
  if (rsec) std::copy(T->rsec.begin(), T->rsec.end(), rsec);
  return TSC_OK;
}

extern "C" int tsc_tasks_plan(tsc_ctx *ctx, const tsc_tasks *T, const double *arena, int64_t arena_len,
                              tsc_plan **plan) {
  TSC_REQUIRE(ctx && T && plan, "tsc_tasks_plan arguments");
  if (!arena && !T->arena.empty()) {
    arena = reinterpret_cast<const double *>(T->arena.data());
    arena_len = (int64_t)T->arena.size();
  }
  TSC_REQUIRE(arena_len >= T->arena_len, "tsc_tasks_plan: arena shorter than the task descriptors need");
  return tsc::plan_create_impl(ctx, arena, nullptr, arena_len, T->mid_len, T->c1.data(), (int32_t)T->c1.size(),
                               T->o1.data(), (int32_t)T->o1.size(), T->c2.data(), (int32_t)T->c2.size(),
                               T->o2.data(), (int32_t)T->o2.size(), T->in_len, T->out_len, plan);
}

extern "C" int tsc_tasks_plan_dev(tsc_ctx *ctx, const tsc_tasks *T, const tsc_vec *arena, tsc_plan **plan) {
  TSC_REQUIRE(ctx && T && arena && plan && arena->ctx == ctx, "tsc_tasks_plan_dev arguments");
  TSC_REQUIRE(arena->n >= T->arena_len, "tsc_tasks_plan_dev: arena shorter than the task descriptors need");
  return tsc::plan_create_impl(ctx, nullptr, arena->d, arena->n, T->mid_len, T->c1.data(), (int32_t)T->c1.size(),
                               T->o1.data(), (int32_t)T->o1.size(), T->c2.data(), (int32_t)T->c2.size(),
                               T->o2.data(), (int32_t)T->o2.size(), T->in_len, T->out_len, plan);
}

extern "C" int tsc_tasks_destroy(tsc_tasks *T) {
  delete T;
  return TSC_OK;
}

// ---------------------------------------------------------------------
// Host logic of model_wave / hmerge and of Get_Hamiltonian_Sys/Env on block
// models: bonds of the lattice classified by the parts of the superblock
// they join, partner operators merged (sum_j w_j O_j, hmerge,
// get_sys.f90:7398-7626, block_add_block_two, block.f90:177-215), terms and
// pairs formed, and the contraction tasks built by the planner above.
namespace {

constexpr double kSq2 = 1.4142135623730951;
constexpr double kSq3 = 1.7320508075688772;

// (coupling, P operator (0 c+, 1 S, 2 n), P tilde, Q operator, Q tilde) of
// one bond: t (c+ c~ + c~ c+) / sqrt(2)-normalised hopping and
// J (S.S - n n / 4) in reduced form; tag = 0 c+c~, 1 c~c+, 2 SS, 3 nn
struct BondProd {
  double g;
  int p;
  bool ptil;
  int q;
  bool qtil;
  int tag;
};

std::vector<BondProd> bond_products(double t, double J) {
  std::vector<BondProd> out;
  if (t != 0.0) {
    out.push_back(BondProd{t * kSq2, 0, false, 0, true, 0});
    out.push_back(BondProd{t * kSq2, 0, true, 0, false, 1});
  }
  if (J != 0.0) {
    out.push_back(BondProd{-kSq3 * J, 1, false, 1, false, 2});
    out.push_back(BondProd{-0.25 * J, 2, false, 2, false, 3});
  }
  return out;
}

// operators made by the host logic (merged sums), kept alive with their blocks
struct OpStore {
  std::deque<std::vector<tsc_op_block>> blks;
  std::deque<tsc_op_desc> descs;
  const tsc_op_desc *add(int side, int up_dif, int k2, std::vector<tsc_op_block> b) {
    blks.push_back(std::move(b));
    tsc_op_desc d;
    d.side = side;
    d.up_dif = up_dif;
    d.k2 = k2;
    d.nblk = (int32_t)blks.back().size();
    d.blk = blks.back().data();
    descs.push_back(d);
    return &descs.back();
  }
};

// sum_j w_j O_j over the blocks of the operators (union of their sub-blocks,
// in order of first appearance), appended to the arena
const tsc_op_desc *merge_ops(OpStore &st, std::vector<double2> &arena, const std::vector<tsc_sector> &sec,
                             const std::vector<std::pair<const tsc_op_desc *, double>> &parts) {
  std::vector<tsc_op_block> out;
  std::map<std::pair<int, int>, int64_t> where;
  for (const auto &pw : parts)
    for (int i = 0; i < pw.first->nblk; ++i) {
      const tsc_op_block &b = pw.first->blk[i];
      if (where.count({b.bra, b.ket})) continue;
      const int64_t off = (int64_t)arena.size();
      arena.resize(arena.size() + (size_t)sec[b.bra].dim * sec[b.ket].dim, make_double2(0.0, 0.0));
      where[{b.bra, b.ket}] = off;
      out.push_back(tsc_op_block{b.bra, b.ket, off});
    }
  for (const auto &pw : parts)
    for (int i = 0; i < pw.first->nblk; ++i) {
      const tsc_op_block &b = pw.first->blk[i];
      const int64_t dst = where[{b.bra, b.ket}], n = (int64_t)sec[b.bra].dim * sec[b.ket].dim;
      for (int64_t e = 0; e < n; ++e) {
        arena[dst + e].x += pw.second * arena[b.off + e].x;
        arena[dst + e].y += pw.second * arena[b.off + e].y;
      }
    }
  const tsc_op_desc *o = parts.front().first;
  return st.add(o->side, o->up_dif, o->k2, std::move(out));
}

// operator `which` (0 c+, 1 S, 2 n) of a kept site of a model, or null
const tsc_op_desc *site_op(const tsc_model_desc *m, int site, int which) {
  for (int i = 0; i < m->nkeep; ++i)
    if (m->keep[i].site == site) return which == 0 ? m->keep[i].c : (which == 1 ? m->keep[i].S : m->keep[i].n);
  return nullptr;
}

int check_model(const tsc_model_desc *m, int64_t arena_len, const char *what) {
  TSC_REQUIRE(m && m->nsec >= 0 && (m->nsec == 0 || m->sec) && m->ham && m->nsites >= 0 && m->nkeep >= 0,
              std::string(what) + " model descriptor");
  auto check_op = [&](const tsc_op_desc *op) -> int {
    TSC_REQUIRE(op && op->nblk >= 0 && (op->nblk == 0 || op->blk), std::string(what) + " operator descriptor");
    for (int b = 0; b < op->nblk; ++b) {
      const tsc_op_block &bl = op->blk[b];
      TSC_REQUIRE(bl.bra >= 0 && bl.bra < m->nsec && bl.ket >= 0 && bl.ket < m->nsec,
                  std::string(what) + " operator block sector index");
      TSC_REQUIRE(bl.off >= 0 && bl.off + (int64_t)m->sec[bl.bra].dim * m->sec[bl.ket].dim <= arena_len,
                  std::string(what) + " operator block outside the arena");
    }
    return TSC_OK;
  };
  int rc = check_op(m->ham);
  if (rc) return rc;
  for (int i = 0; i < m->nkeep; ++i)
    for (const tsc_op_desc *op : {m->keep[i].c, m->keep[i].S, m->keep[i].n})
      if ((rc = check_op(op))) return rc;
  return TSC_OK;
}

}  // namespace

extern "C" int tsc_model_tasks(const tsc_model_desc *sys, const tsc_model_desc *env, int32_t s1, int32_t s2,
                               int32_t nbonds, const tsc_bond *bonds, int32_t ntot, int32_t s2tot,
                               const double *arena, int64_t arena_len, tsc_tasks **out) {
  TSC_REQUIRE(out && nbonds >= 0 && (nbonds == 0 || bonds) && arena_len >= 0 && (arena || arena_len == 0),
              "tsc_model_tasks arguments");
  int rc = check_model(sys, arena_len, "sys");
  if (!rc) rc = check_model(env, arena_len, "env");
  if (rc) return rc;
  std::vector<double2> ar(reinterpret_cast<const double2 *>(arena),
                          reinterpret_cast<const double2 *>(arena) + arena_len);
  std::vector<tsc_sector> A(sys->sec, sys->sec + sys->nsec), B(env->sec, env->sec + env->nsec);
  // parts: 0 A, 1 s1, 2 B, 3 s2 (the order classes are written in)
  std::map<int, int> part;
  for (int i = 0; i < sys->nsites; ++i) part[sys->sites[i]] = 0;
  for (int i = 0; i < env->nsites; ++i) part[env->sites[i]] = 2;
  part[s1] = 1;
  part[s2] = 3;
  std::map<std::pair<int, int>, std::vector<tsc_bond>> cls;
  for (int b = 0; b < nbonds; ++b) {
    tsc_bond bd = bonds[b];
    auto pi = part.find(bd.i), pj = part.find(bd.j);
    if (pi == part.end() || pj == part.end() || pi->second == pj->second) continue;
    int a = pi->second, c = pj->second;
    if (a > c) {
      std::swap(bd.i, bd.j);
      std::swap(a, c);
    }
    cls[{a, c}].push_back(bd);
  }
  OpStore st;
  std::vector<tsc_op_desc> ops;   // every operator the descriptors address
  auto op_index = [&](const tsc_op_desc *d, int side) {
    tsc_op_desc c = *d;
    c.side = side;
    ops.push_back(c);
    return (int32_t)ops.size() - 1;
  };
  auto need = [&](const tsc_model_desc *m, int site, int which) -> const tsc_op_desc * {
    const tsc_op_desc *o = site_op(m, site, which);
    if (!o) set_error("invalid: block site " + std::to_string(site) + " has a bond outside its block but no operators");
    return o;
  };
  std::vector<tsc_pair_desc> lpairs, rpairs;
  std::vector<tsc_term_desc> terms;
  // (A, s1) and (B, s2): block-site pairs of one side, merged over block sites
  for (int side = 0; side < 2; ++side) {
    const tsc_model_desc *m = side == 0 ? sys : env;
    auto it = cls.find(side == 0 ? std::make_pair(0, 1) : std::make_pair(2, 3));
    if (it == cls.end()) continue;
    for (int tag = 0; tag < 4; ++tag) {
      std::vector<std::pair<const tsc_op_desc *, double>> parts;
      BondProd spec{};
      for (const tsc_bond &bd : it->second)
        for (const BondProd &bp : bond_products(bd.t, bd.J))
          if (bp.tag == tag) {
            const tsc_op_desc *o = need(m, bd.i, bp.p);
            if (!o) return TSC_EINVAL;
            parts.push_back({o, bp.g});
            spec = bp;
          }
      if (parts.empty()) continue;
      const tsc_op_desc *M = merge_ops(st, ar, side == 0 ? A : B, parts);
      const tsc_pair_desc pd{1.0, op_index(M, side), spec.ptil ? 1 : 0, spec.q, spec.qtil ? 1 : 0};
      (side == 0 ? lpairs : rpairs).push_back(pd);
    }
  }
  // L-R classes in the order (A, B), (A, s2), (s1, B), (s1, s2)
  for (auto key : {std::make_pair(0, 2), std::make_pair(0, 3), std::make_pair(1, 2), std::make_pair(1, 3)}) {
    auto it = cls.find(key);
    if (it == cls.end()) continue;
    const std::vector<tsc_bond> &bl = it->second;
    if (key == std::make_pair(0, 2)) {
      // one term per A site, its B partners merged
      std::map<int, std::vector<tsc_bond>> by_i;
      for (const tsc_bond &bd : bl) by_i[bd.i].push_back(bd);
      for (auto &kv : by_i)
        for (int tag = 0; tag < 4; ++tag) {
          std::vector<std::pair<const tsc_op_desc *, double>> parts;
          BondProd spec{};
          for (const tsc_bond &bd : kv.second)
            for (const BondProd &bp : bond_products(bd.t, bd.J))
              if (bp.tag == tag) {
                const tsc_op_desc *o = need(env, bd.j, bp.q);
                if (!o) return TSC_EINVAL;
                parts.push_back({o, bp.g});
                spec = bp;
              }
          if (parts.empty()) continue;
          const tsc_op_desc *P = need(sys, kv.first, spec.p);
          if (!P) return TSC_EINVAL;
          const tsc_op_desc *MB = merge_ops(st, ar, B, parts);
          terms.push_back(tsc_term_desc{1.0, 1, op_index(P, 0), spec.ptil ? 1 : 0, 1, op_index(MB, 1),
                                        spec.qtil ? 1 : 0});
        }
      continue;
    }
    for (int tag = 0; tag < 4; ++tag) {
      std::vector<std::pair<const tsc_op_desc *, double>> parts;
      double gsum = 0.0;
      BondProd spec{};
      bool any = false;
      for (const tsc_bond &bd : bl)
        for (const BondProd &bp : bond_products(bd.t, bd.J))
          if (bp.tag == tag) {
            any = true;
            spec = bp;
            gsum += bp.g;
            if (key.first == 0) {   // (A, s2): merge over A sites
              const tsc_op_desc *o = need(sys, bd.i, bp.p);
              if (!o) return TSC_EINVAL;
              parts.push_back({o, bp.g});
            } else if (key.second == 2) {   // (s1, B): merge over B sites
              const tsc_op_desc *o = need(env, bd.j, bp.q);
              if (!o) return TSC_EINVAL;
              parts.push_back({o, bp.g});
            }
          }
      if (!any) continue;
      if (key.first == 0) {
        const tsc_op_desc *M = merge_ops(st, ar, A, parts);
        terms.push_back(tsc_term_desc{1.0, 1, op_index(M, 0), spec.ptil ? 1 : 0, 2, spec.q, spec.qtil ? 1 : 0});
      } else if (key.second == 2) {
        const tsc_op_desc *M = merge_ops(st, ar, B, parts);
        terms.push_back(tsc_term_desc{1.0, 2, spec.p, spec.ptil ? 1 : 0, 1, op_index(M, 1), spec.qtil ? 1 : 0});
      } else {
        terms.push_back(tsc_term_desc{gsum, 2, spec.p, spec.ptil ? 1 : 0, 2, spec.q, spec.qtil ? 1 : 0});
      }
    }
  }
  const int32_t hs = op_index(sys->ham, 0), he = op_index(env->ham, 1);
  tsc_sb_desc D;
  std::memset(&D, 0, sizeof(D));
  D.nsys = sys->nsec;
  D.nenv = env->nsec;
  D.sys = sys->sec;
  D.env = env->sec;
  D.arena_len = (int64_t)ar.size();
  D.nops = (int32_t)ops.size();
  D.ops = ops.data();
  D.sys_ham = &ops[hs];
  D.env_ham = &ops[he];
  D.nlpairs = (int32_t)lpairs.size();
  D.lpairs = lpairs.data();
  D.nrpairs = (int32_t)rpairs.size();
  D.rpairs = rpairs.data();
  D.nterms = (int32_t)terms.size();
  D.terms = terms.data();
  D.ntot = ntot;
  D.s2tot = s2tot;
  tsc_tasks *T = nullptr;
  rc = tsc_superblock_tasks(&D, &T);
  if (rc) return rc;
  T->arena = std::move(ar);
  *out = T;
  return TSC_OK;
}

extern "C" int tsc_renorm_tasks(const tsc_model_desc *blk, int32_t new_site, int32_t nbonds, const tsc_bond *bonds,
                                int32_t nkeep, const int32_t *keep_sites, int32_t norder, const int32_t *order,
                                const int32_t *kept, const double *arena, int64_t arena_len, tsc_tasks **out) {
  TSC_REQUIRE(out && nbonds >= 0 && (nbonds == 0 || bonds) && nkeep >= 0 && (nkeep == 0 || keep_sites) &&
                  norder >= 0 && (norder == 0 || (order && kept)) && arena_len >= 0 && (arena || arena_len == 0),
              "tsc_renorm_tasks arguments");
  int rc = check_model(blk, arena_len, "block");
  if (rc) return rc;
  std::vector<double2> ar(reinterpret_cast<const double2 *>(arena),
                          reinterpret_cast<const double2 *>(arena) + arena_len);
  std::vector<tsc_sector> sec(blk->sec, blk->sec + blk->nsec);
  const Basis basis = get_basis(sec);
  const int nx = (int)basis.sec.size();
  std::vector<int> kx(nx, -1), newidx(nx, -1);
  std::vector<int64_t> uo(nx, 0);
  int64_t ulen = 0;
  for (int q = 0; q < norder; ++q) {
    const int x = order[q];
    TSC_REQUIRE(x >= 0 && x < nx && kx[x] < 0, "order: enlarged sector index out of range or repeated");
    TSC_REQUIRE(kept[q] >= 1 && kept[q] <= basis.sec[x].dim, "kept states of a sector outside [1, dim]");
    kx[x] = kept[q];
    newidx[x] = q;
    uo[x] = ulen;
    ulen += (int64_t)basis.sec[x].dim * kept[q];
  }
  // merged block-site pairs of the enlarged block's Hamiltonian
  OpStore st;
  std::vector<tsc_op_desc> ops;
  std::vector<tsc_pair_desc> pairs;
  for (int tag = 0; tag < 4; ++tag) {
    std::vector<std::pair<const tsc_op_desc *, double>> parts;
    BondProd spec{};
    for (int b = 0; b < nbonds; ++b) {
      int i = bonds[b].i, j = bonds[b].j;
      if (j != new_site) std::swap(i, j);
      if (j != new_site || i == new_site) continue;
      bool inside = false;
      for (int k = 0; k < blk->nsites; ++k) inside |= blk->sites[k] == i;
      if (!inside) continue;
      for (const BondProd &bp : bond_products(bonds[b].t, bonds[b].J))
        if (bp.tag == tag) {
          const tsc_op_desc *o = site_op(blk, i, bp.p);
          TSC_REQUIRE(o, "block site " + std::to_string(i) + " bonded to the new site has no operators");
          parts.push_back({o, bp.g});
          spec = bp;
        }
    }
    if (parts.empty()) continue;
    const tsc_op_desc *M = merge_ops(st, ar, sec, parts);
    ops.push_back(*M);
    pairs.push_back(tsc_pair_desc{1.0, (int32_t)ops.size() - 1, spec.ptil ? 1 : 0, spec.q, spec.qtil ? 1 : 0});
  }
  tsc_sb_desc D;
  std::memset(&D, 0, sizeof(D));
  D.arena_len = (int64_t)ar.size();
  Ctx cx;
  cx.D = &D;
  // transitions: H, then the carried operators of every kept site
  std::vector<std::vector<Tr>> trans;
  std::vector<int32_t> opkey;
  trans.emplace_back();
  opkey.push_back(-1);
  scalar_transitions(cx, basis, sec, blk->ham, pairs.data(), (int)pairs.size(), ops.data(), true, trans.back());
  for (int k = 0; k < nkeep; ++k)
    for (int which = 0; which < 3; ++which) {
      trans.emplace_back();
      opkey.push_back(3 * k + which);
      if (keep_sites[k] == new_site) {
        site_transitions(cx, basis, sec, which, false, trans.back());
      } else {
        const tsc_op_desc *o = site_op(blk, keep_sites[k], which);
        TSC_REQUIRE(o, "kept site " + std::to_string(keep_sites[k]) + " has no operators in the block");
        blk_transitions(cx, basis, sec, *o, false, trans.back());
      }
    }
  tsc_tasks *T = new tsc_tasks();
  T->arena_len = (int64_t)ar.size();
  T->in_len = ulen;
  // output windows (operator, x2, x) in first-appearance order
  std::map<std::tuple<int, int, int>, int> wid;
  std::vector<std::vector<tsc_contrib>> per_w;
  std::map<std::tuple<int, int, int>, int64_t> mid;   // (matrix, x, p) -> MID offset of T U_x[rows]
  int64_t wout = 0;
  auto window = [&](int key, int x2, int x) {
    auto k = std::make_tuple(key, x2, x);
    auto it = wid.find(k);
    if (it != wid.end()) return it->second;
    tsc_outblock ob;
    std::memset(&ob, 0, sizeof(ob));
    ob.c_off = wout;
    ob.ldc = std::max(1, kx[x2]);
    ob.m = kx[x2];
    ob.n = kx[x];
    ob.c_buf = TSC_BUF_OUT;
    T->o2.push_back(ob);
    per_w.emplace_back();
    T->layout.insert(T->layout.end(), {key, newidx[x2], newidx[x]});
    T->layout_off.push_back(wout);
    wout += (int64_t)kx[x2] * kx[x];
    const int w = (int)T->o2.size() - 1;
    wid[k] = w;
    return w;
  };
  for (size_t o = 0; o < trans.size(); ++o)
    for (const Tr &t : trans[o]) {
      if (kx[t.x] < 0 || kx[t.x2] < 0 || t.f == 0.0) continue;
      const int w = window(opkey[o], t.x2, t.x);
      const int dx = basis.sec[t.x].dim, dx2 = basis.sec[t.x2].dim;
      if (t.mat >= 0) {
        const MatRef &m = cx.mats[t.mat];
        auto key = std::make_tuple(t.mat, t.x, t.p);
        int64_t moff;
        auto it = mid.find(key);
        if (it != mid.end()) {
          moff = it->second;
        } else {
          moff = T->mid_len;
          T->mid_len += (int64_t)t.d2 * kx[t.x];
          mid[key] = moff;
          tsc_outblock ob;
          std::memset(&ob, 0, sizeof(ob));
          ob.c_off = moff;
          ob.ldc = std::max(1, t.d2);
          ob.m = t.d2;
          ob.n = kx[t.x];
          ob.c_buf = TSC_BUF_MID;
          ob.cbeg = (int32_t)T->c1.size();
          ob.cend = ob.cbeg + 1;
          T->o1.push_back(ob);
          T->c1.push_back(Builder::make(m.off, m.ld, TSC_BUF_OPS, m.op, uo[t.x] + t.p, std::max(1, dx), TSC_BUF_IN,
                                        TSC_OP_N, t.d, 1.0));
        }
        per_w[w].push_back(Builder::make(uo[t.x2] + t.p2, std::max(1, dx2), TSC_BUF_IN, TSC_OP_C, moff,
                                         std::max(1, t.d2), TSC_BUF_MID, TSC_OP_N, t.d2, t.f));
      } else {
        per_w[w].push_back(Builder::make(uo[t.x2] + t.p2, std::max(1, dx2), TSC_BUF_IN, TSC_OP_C, uo[t.x] + t.p,
                                         std::max(1, dx), TSC_BUF_IN, TSC_OP_N, t.d, t.f));
      }
    }
  for (size_t w = 0; w < T->o2.size(); ++w) {
    T->o2[w].cbeg = (int32_t)T->c2.size();
    T->c2.insert(T->c2.end(), per_w[w].begin(), per_w[w].end());
    T->o2[w].cend = (int32_t)T->c2.size();
  }
  // diagonal Hamiltonian blocks that no term reaches: zero windows (not
  // written by the plan; the output vector starts zeroed)
  for (int q = 0; q < norder; ++q) {
    const int x = order[q];
    if (wid.count(std::make_tuple(-1, x, x))) continue;
    T->layout.insert(T->layout.end(), {-1, q, q});
    T->layout_off.push_back(wout);
    wout += (int64_t)kx[x] * kx[x];
  }
  T->out_len = std::max<int64_t>(wout, 1);
  for (int q = 0; q < norder; ++q)
    T->new_sec.insert(T->new_sec.end(), {basis.sec[order[q]].n, basis.sec[order[q]].j, kx[order[q]]});
  for (const EnlSector &e : basis.sec) T->lsec.insert(T->lsec.end(), {e.n, e.j, e.dim});
  T->arena = std::move(ar);
  *out = T;
  return TSC_OK;
}

extern "C" int tsc_tasks_layout(const tsc_tasks *T, int64_t *counts3, int32_t *layout, int64_t *offsets,
                                int32_t *new_sec, int32_t *enl_sec) {
  TSC_REQUIRE(T && counts3, "tsc_tasks_layout arguments");
  counts3[0] = (int64_t)T->layout_off.size();
  counts3[1] = (int64_t)T->new_sec.size() / 3;
  counts3[2] = (int64_t)T->lsec.size() / 3;
  if (layout) std::copy(T->layout.begin(), T->layout.end(), layout);
  if (offsets) std::copy(T->layout_off.begin(), T->layout_off.end(), offsets);
  if (new_sec) std::copy(T->new_sec.begin(), T->new_sec.end(), new_sec);
  if (enl_sec) std::copy(T->lsec.begin(), T->lsec.end(), enl_sec);
  return TSC_OK;
}

extern "C" int tsc_tasks_lengths(const tsc_tasks *T, int64_t *lens3) {
  TSC_REQUIRE(T && lens3, "tsc_tasks_lengths arguments");
  lens3[0] = T->in_len;
  lens3[1] = T->out_len;
  lens3[2] = (int64_t)T->arena.size();
  return TSC_OK;
}

extern "C" int tsc_tasks_copy_arena(const tsc_tasks *T, double *arena_out) {
  TSC_REQUIRE(T && (arena_out || T->arena.empty()), "tsc_tasks_copy_arena arguments");
  if (!T->arena.empty()) std::memcpy(arena_out, T->arena.data(), sizeof(double2) * T->arena.size());
  return TSC_OK;
}
