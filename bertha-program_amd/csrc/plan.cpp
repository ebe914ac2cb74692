// plan.cpp -- host side of the engine: basis labels, spinor E-coefficient term
// tables and the Fock-build plan (which block quartets are evaluated and how
// each RR batch is digested).  The plan reproduces the decisions of the
// reference's CLMFAST loop nest (bertha-scf.f:11630-12158); the arithmetic
// happens in kernels.hip.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <limits>
#include <map>
#include <tuple>

namespace bdf {

static int lval(int k) { return k < 0 ? -k - 1 : k; }
static int isgn(int x) { return x >= 0 ? 1 : -1; }
static double m1pow(int n) { return (std::abs(n) % 2 == 0) ? 1.0 : -1.0; }

// ---------------------------------------------------------------------------
// basis and labels
// ---------------------------------------------------------------------------
int Basis::kappa_of(int c, int k) const {  // k is 1-based KAPA index
  int l = k / 2;
  return (k % 2 == 0) ? l : -l - 1;
}

int axis_symmetry(const ClassDesc &cl) {
  int m = 0;
  for (int d = 0; d < 3; ++d)
    if (cl.xyz[0][d] == 0.0 && cl.xyz[1][d] == 0.0 && cl.xyz[2][d] == 0.0 && cl.xyz[3][d] == 0.0)
      m |= 1 << d;
  return m;
}

std::string build_basis(const bdf_system &sys, Basis &B) {
  if (sys.ncnt <= 0 || !sys.xyz || !sys.lmax || !sys.nfnc || !sys.exps)
    return "bdf_system: null or empty arrays";
  B.ncnt = sys.ncnt;
  B.xyz.assign(sys.xyz, sys.xyz + 3 * sys.ncnt);
  B.lmax.assign(sys.lmax, sys.lmax + sys.ncnt);
  B.nf.assign(sys.ncnt * (kMaxL + 1), 0);
  B.expoff.assign(sys.ncnt * (kMaxL + 1), 0);
  B.exps.clear();
  B.ndim = 0;
  for (int c = 0; c < sys.ncnt; ++c) {
    if (sys.lmax[c] < 0 || sys.lmax[c] > kMaxL) return "lmax outside 0..3";
    if (sys.lmax[c] >= sys.lstride) return "lmax >= lstride";
    for (int l = 0; l <= sys.lmax[c]; ++l) {
      int n = sys.nfnc[c * sys.lstride + l];
      if (n <= 0 || n > kMaxPrim || n > sys.mbs) return "shell size outside 1..32";
      B.nf[c * (kMaxL + 1) + l] = n;
      B.expoff[c * (kMaxL + 1) + l] = (int)B.exps.size();
      for (int i = 0; i < n; ++i) {
        double e = sys.exps[((long)c * sys.lstride + l) * sys.mbs + i];
        if (!(e > 0.0)) return "non-positive exponent";
        B.exps.push_back(e);
      }
      B.ndim += 4 * (2 * l + 1) * n;
    }
  }
  B.nskp = B.ndim / 2;
  // FOCKIND (bertha-scf.f:2605-2645): per centre, per |mj| shell, first all
  // -|mj| blocks in kappa order, then all +|mj| blocks.
  B.lrge.assign(sys.ncnt * 16 * 16, -1);
  B.labicn.assign(B.nskp, 0);
  B.labkqn.assign(B.nskp, 0);
  B.labmqn.assign(B.nskp, 0);
  int ilst = 0;
  for (int c = 0; c < sys.ncnt; ++c) {
    int nkap = 2 * B.lmax[c] + 1;
    for (int imv = 0; imv <= B.lmax[c]; ++imv) {
      int mqn = 2 * imv + 1;
      for (int half = 0; half < 2; ++half)
        for (int k = 1; k <= nkap; ++k) {
          int kq = B.kappa_of(c, k);
          int n = B.nfn(c, lval(kq));
          if (2 * std::abs(kq) - 1 >= mqn) {
            B.lrge[(c * 16 + k) * 16 + mqn + half] = ilst;
            for (int i = 0; i < n; ++i) {
              B.labicn[ilst + i] = c + 1;
              B.labkqn[ilst + i] = kq;
              B.labmqn[ilst + i] = half ? mqn : -mqn;
            }
            ilst += n;
          }
        }
    }
  }
  // normalisation of large and small radial functions (RNORM1,
  // bertha-scf.f:37770-37800): N^2 = 2 (2a)^(l+3/2) / Gamma(l+3/2) and
  // 2 (2a)^(l+1/2) / Gamma(l+5/2).
  B.normL.assign(B.exps.size(), 0.0);
  B.normS.assign(B.exps.size(), 0.0);
  for (int c = 0; c < sys.ncnt; ++c)
    for (int l = 0; l <= B.lmax[c]; ++l)
      for (int i = 0; i < B.nfn(c, l); ++i) {
        int o = B.expo(c, l) + i;
        double le = std::log(2.0 * B.exps[o]);
        B.normL[o] = std::exp(0.5 * (std::log(2.0) - std::lgamma(l + 1.5) + (l + 1.5) * le));
        B.normS[o] = std::exp(0.5 * (std::log(2.0) - std::lgamma(l + 2.5) + (l + 0.5) * le));
      }
  return "";
}

// ---------------------------------------------------------------------------
// solid harmonics and spinor coupling
// ---------------------------------------------------------------------------
// Unnormalised solid harmonics with the Schiff phase, S[l,m](x,y,z), built by
// the same recurrences the reference applies to Hermite expansions (Saunders
// 1983 Eq. 64, bertha-scf.f:37105-37566): S[0,0] = 1,
// S[l+1,+-(l+1)] = (2l+1)(x +- iy) S[l,+-l], S[|m|+1,m] = (2|m|+1) z S[|m|,m],
// (l-|m|+1) S[l+1,m] = (2l+1) z S[l,m] - (l+|m|) r^2 S[l-1,m].
namespace {
constexpr int D = kMaxPoly + 3;
struct Poly {
  std::complex<double> c[D][D][D];
  Poly() { clear(); }
  void clear() {
    for (auto &a : c) for (auto &b : a) for (auto &v : b) v = 0.0;
  }
};
Poly mul_lin(const Poly &p, std::complex<double> cx, std::complex<double> cy,
             std::complex<double> cz) {
  Poly o;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      for (int k = 0; k < D; ++k) {
        auto v = p.c[i][j][k];
        if (v == 0.0) continue;
        if (i + 1 < D) o.c[i + 1][j][k] += cx * v;
        if (j + 1 < D) o.c[i][j + 1][k] += cy * v;
        if (k + 1 < D) o.c[i][j][k + 1] += cz * v;
      }
  return o;
}
Poly mul_r2(const Poly &p) {
  Poly o;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      for (int k = 0; k < D; ++k) {
        auto v = p.c[i][j][k];
        if (v == 0.0) continue;
        if (i + 2 < D) o.c[i + 2][j][k] += v;
        if (j + 2 < D) o.c[i][j + 2][k] += v;
        if (k + 2 < D) o.c[i][j][k + 2] += v;
      }
  return o;
}
Poly axpy(double a, const Poly &x, double b, const Poly &y) {
  Poly o;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      for (int k = 0; k < D; ++k) o.c[i][j][k] = a * x.c[i][j][k] + b * y.c[i][j][k];
  return o;
}
Poly schiff(int l, int m) {
  Poly p;
  p.c[0][0][0] = 1.0;
  double sg = m >= 0 ? 1.0 : -1.0;
  for (int lq = 0; lq < std::abs(m); ++lq)
    p = mul_lin(p, 2.0 * lq + 1.0, std::complex<double>(0.0, sg * (2.0 * lq + 1.0)), 0.0);
  if (l <= std::abs(m)) return p;
  int am = std::abs(m);
  Poly prev = p;
  Poly cur = mul_lin(p, 0.0, 0.0, 2.0 * am + 1.0);
  for (int ls = am + 1; ls <= l - 1; ++ls) {
    double r1 = (2.0 * ls + 1.0) / (ls - am + 1.0);
    double r2 = -(double)(ls + am) / (ls - am + 1.0);
    Poly nxt = axpy(r1, mul_lin(cur, 0.0, 0.0, 1.0), r2, mul_r2(prev));
    prev = cur;
    cur = nxt;
  }
  return cur;
}
}  // namespace

int Etabs::poly_id(int l, int m, int nr2) {
  auto key = std::make_tuple(l, m, nr2);
  auto it = poly_index.find(key);
  if (it != poly_index.end()) return it->second;
  Poly p = schiff(l, m);
  if (nr2) p = mul_r2(p);
  int begin = (int)monos.size();
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      for (int k = 0; k < D; ++k) {
        auto v = p.c[i][j][k];
        if (std::abs(v) == 0.0) continue;
        Mono mo;
        mo.re = v.real(); mo.im = v.imag();
        mo.i = (int8_t)i; mo.j = (int8_t)j; mo.k = (int8_t)k; mo.pad = 0;
        monos.push_back(mo);
      }
  int id = begin * 65536 + ((int)monos.size() - begin);
  poly_index[key] = id;
  return id;
}

// Angular factor of conj(Y[l1,m1]) Y[l2,m2] with the Condon-Shortley phase
// (ESGTF, bertha-scf.f:36948-36964).
static double ang(int l1, int m1, int l2, int m2) {
  double ph1 = m1pow((m1 + std::abs(m1)) / 2), ph2 = m1pow((m2 + std::abs(m2)) / 2);
  double f1 = std::tgamma(l1 - std::abs(m1) + 1.0) / std::tgamma(l1 + std::abs(m1) + 1.0);
  double f2 = std::tgamma(l2 - std::abs(m2) + 1.0) / std::tgamma(l2 + std::abs(m2) + 1.0);
  return ph1 * ph2 * (0.25 / M_PI) * std::sqrt((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * f1 * f2);
}

// Clebsch-Gordan weights of the two spin components of a 2-spinor with
// kappa kq and 2*mj = mq (EQLL bertha-scf.f:34722-34734); the small
// component of kappa has the coefficients of -kappa (EQSS 36012-36024).
static void cg(int kq, int mq, bool small, double &up, double &lo) {
  int jq = 2 * std::abs(kq) - 1;
  bool neg = kq < 0;
  if (small) neg = !neg;
  if (neg) {
    up = std::sqrt((double)(jq + mq) / (2.0 * jq));
    lo = std::sqrt((double)(jq - mq) / (2.0 * jq));
  } else {
    up = -std::sqrt((double)(jq - mq + 2) / (2.0 * jq + 4));
    lo = std::sqrt((double)(jq + mq + 2) / (2.0 * jq + 4));
  }
}

// Terms of the overlap density of spinors (ka, mqa) and (kb, mqb) in
// polynomial form, for the large (small=false) or kinetically balanced small
// (small=true) components.  Small components carry the radial factors of
// sigma.p acting on r^l exp(-a r^2): -2a r^(l+1) for kappa<0 and
// (2l+1) r^(l-1) - 2a r^(l+1) for kappa>0 (EQSS cases 1-4, 35996-36860).
void Etabs::spinor_terms(int ka, int mqa, int kb, int mqb, int kind, int &begin, int &count) {
  auto key = std::make_tuple(ka, mqa, kb, mqb, kind);
  auto it = term_index.find(key);
  if (it != term_index.end()) {
    begin = it->second.first;
    count = it->second.second;
    return;
  }
  begin = (int)terms.size();
  if (kind >= kLSx) {
    ls_terms(ka, mqa, kb, mqb, kind - kLSx + 1);
    count = (int)terms.size() - begin;
    term_index[key] = {begin, count};
    return;
  }
  const bool small = kind == kSS;
  int la = lval(ka), lb = lval(kb);
  double cau, cal, cbu, cbl;
  cg(ka, mqa, small, cau, cal);
  cg(kb, mqb, small, cbu, cbl);
  for (int s = 0; s < 2; ++s) {
    int mla = s == 0 ? (mqa - 1) / 2 : (mqa + 1) / 2;
    int mlb = s == 0 ? (mqb - 1) / 2 : (mqb + 1) / 2;
    double cab = s == 0 ? cau * cbu : cal * cbl;
    if (std::fabs(cab) < 1.0e-14) continue;
    int l1 = la, l2 = lb;
    if (small) {
      l1 = ka < 0 ? la + 1 : la - 1;
      l2 = kb < 0 ? lb + 1 : lb - 1;
    }
    if (std::abs(mla) > l1 || std::abs(mlb) > l2) continue;
    double w0 = cab * ang(l1, mla, l2, mlb);
    auto add = [&](double w, int pa, int pb, int ra, int rb) {
      ETerm t;
      t.w = w0 * w;
      t.pa = (int16_t)pa;
      t.pb = (int16_t)pb;
      t.polyA = poly_id(l1, -mla, ra);
      t.polyB = poly_id(l2, mlb, rb);
      terms.push_back(t);
    };
    if (!small) {
      add(1.0, 0, 0, 0, 0);
    } else {
      double t20 = -2.0 * (2 * lb + 1), t02 = -2.0 * (2 * la + 1), t00 = (2.0 * la + 1) * (2.0 * lb + 1);
      if (ka < 0 && kb < 0) {
        add(4.0, 1, 1, 0, 0);
      } else if (ka < 0 && kb > 0) {
        add(t20, 1, 0, 0, 0);
        add(4.0, 1, 1, 0, 1);
      } else if (ka > 0 && kb < 0) {
        add(t02, 0, 1, 0, 0);
        add(4.0, 1, 1, 1, 0);
      } else {
        add(t00, 0, 0, 0, 0);
        add(t20, 1, 0, 1, 0);
        add(t02, 0, 1, 0, 1);
        add(4.0, 1, 1, 1, 1);
      }
    }
  }
  count = (int)terms.size() - begin;
  term_index[key] = {begin, count};
}

// Large component of (ka, mqa) against the kinetically balanced small
// component of (kb, mqb), coupled by the Pauli matrix sigma_iq (EQLS,
// bertha-scf.f:34937-35430): spin terms 11/22 (z) or 12/21 (x, y) with the
// reference's signs; radial factors of the small component -2b r^(l+1)
// (kappa < 0) or (2l+1) r^(l-1) - 2b r^(l+1) (kappa > 0).  The factor i
// (and i^2 for sigma_y) is applied by the table kernel.
void Etabs::ls_terms(int ka, int mqa, int kb, int mqb, int iq) {
  int la = lval(ka), lb = lval(kb);
  double cau, cal, cbu, cbl;
  cg(ka, mqa, false, cau, cal);
  cg(kb, mqb, true, cbu, cbl);
  double sig = iq == 1 ? 1.0 : -1.0;
  int tsa[2], tsb[2];
  double tw[2];
  if (iq == 3) {
    tsa[0] = 0; tsb[0] = 0; tw[0] = 1.0;
    tsa[1] = 1; tsb[1] = 1; tw[1] = sig;
  } else {
    tsa[0] = 0; tsb[0] = 1; tw[0] = sig;
    tsa[1] = 1; tsb[1] = 0; tw[1] = 1.0;
  }
  int l2 = kb < 0 ? lb + 1 : lb - 1;
  for (int s = 0; s < 2; ++s) {
    int mla = tsa[s] == 0 ? (mqa - 1) / 2 : (mqa + 1) / 2;
    int mlb = tsb[s] == 0 ? (mqb - 1) / 2 : (mqb + 1) / 2;
    double cab = (tsa[s] == 0 ? cau : cal) * (tsb[s] == 0 ? cbu : cbl);
    if (std::fabs(cab) < 1.0e-14) continue;
    if (std::abs(mla) > la || std::abs(mlb) > l2) continue;
    double w0 = tw[s] * cab * ang(la, mla, l2, mlb);
    auto add = [&](double w, int pb, int rb) {
      ETerm t;
      t.w = w0 * w;
      t.pa = 0;
      t.pb = (int16_t)pb;
      t.polyA = poly_id(la, -mla, 0);
      t.polyB = poly_id(l2, mlb, rb);
      terms.push_back(t);
    };
    if (kb < 0) {
      add(-2.0, 1, 0);
    } else {
      add(2.0 * lb + 1.0, 0, 0);
      add(-2.0, 1, 1);
    }
  }
}

// E0LLGN/E0SSGN (bertha-scf.f:33378, 33511): one table per ordered pair of
// (centre, kappa, |mj|) blocks and component type; E11 = (-|mA|,-|mB|),
// E21 = (+|mA|,-|mB|) (EQLLMK 34003-34034).
void build_etables(const Basis &B, Etabs &T, bool with_ls) {
  T.pairs.clear();
  T.pair_index.clear();
  int64_t off = 0;
  const int nkind = with_ls ? 5 : 2;
  for (int small = 0; small < nkind; ++small)
    for (int a = 0; a < B.ncnt; ++a)
      for (int b = 0; b < B.ncnt; ++b)
        for (int ka = 1; ka <= 2 * B.lmax[a] + 1; ++ka)
          for (int kb = 1; kb <= 2 * B.lmax[b] + 1; ++kb) {
            int kqa = B.kappa_of(a, ka), kqb = B.kappa_of(b, kb);
            int la = lval(kqa), lb = lval(kqb);
            for (int ma = 1; ma <= std::abs(kqa); ++ma)
              for (int mb = 1; mb <= std::abs(kqb); ++mb) {
                EPair p;
                p.ax = B.xyz[3 * a]; p.ay = B.xyz[3 * a + 1]; p.az = B.xyz[3 * a + 2];
                p.bx = B.xyz[3 * b]; p.by = B.xyz[3 * b + 1]; p.bz = B.xyz[3 * b + 2];
                p.expA = B.expo(a, la); p.expB = B.expo(b, lb);
                p.nA = B.nfn(a, la); p.nB = B.nfn(b, lb);
                p.lam = la + lb + (small == kSS ? 2 : (small >= kLSx ? 1 : 0));
                p.small = small;
                p.phase = small >= kLSx ? (small == kLSy ? 2 : 1) : 0;
                p.la = la; p.lb = lb;
                p.off = off;
                T.spinor_terms(kqa, -(2 * ma - 1), kqb, -(2 * mb - 1), small, p.term11, p.nterm11);
                T.spinor_terms(kqa, (2 * ma - 1), kqb, -(2 * mb - 1), small, p.term21, p.nterm21);
                off += (int64_t)ntuv(p.lam) * p.nA * p.nB;
                T.pair_index[Etabs::key(a, b, ka, kb, ma, mb, small)] = (int)T.pairs.size();
                T.pairs.push_back(p);
              }
          }
  T.total = off;
  // E(CD) column masks: 4 bytes per tuv per pair
  T.flag_off.assign(T.pairs.size(), 0);
  int64_t fo = 0;
  for (size_t i = 0; i < T.pairs.size(); ++i) {
    T.flag_off[i] = fo;
    fo += 4 * ntuv(T.pairs[i].lam);
  }
  T.nflag = fo;
}

// ---------------------------------------------------------------------------
// the Fock-build plan: CLMFAST's loop nest (bertha-scf.f:11630-12158)
// ---------------------------------------------------------------------------
// ISCF (bertha-scf.f:11555-11561) as bit masks over batches 1..11
static const int kISCF[6] = {
    0b00011111111, 0b00001110011, 0b00001011101, 0b00000110111, 0b00000010101, 0b00000010001};
static const int kMAPTTTT[4][4] = {{1, 0, 0, 2}, {0, 5, 6, 0}, {0, 7, 8, 0}, {3, 0, 0, 4}};

static int ncntrs(int a, int b, int c, int d) {
  int n = 1;
  if (b != a) ++n;
  if (c != a && c != b) ++n;
  if (d != a && d != b && d != c) ++n;
  return n;
}

// Items of a plan: every (class, ij), its AB combinations cut into
// consecutive runs of cost at most max(total / 8192, one combination), then
// ordered by cost, largest first.
// The kernels address the plan's tables with 32-bit offsets and counts
// (ABCombo::gmask, cd_begin/cd_end, item indices): refuse a plan that does
// not fit them rather than let an offset wrap.
template <class PlanT>
static std::string check_offsets(const PlanT &P) {
  const size_t lim = (size_t)std::numeric_limits<int>::max();
  if (P.masks.size() > lim || P.cds.size() > lim || P.abs.size() > lim || P.items.size() > lim ||
      P.classes.size() > lim)
    return "plan exceeds the 32-bit offsets of its tables (masks " +
           std::to_string(P.masks.size()) + " bytes, " + std::to_string(P.cds.size()) +
           " calls, " + std::to_string(P.items.size()) + " items)";
  return "";
}

static void make_items(const std::vector<ClassDesc> &classes, const std::vector<double> &abcost,
                       const std::vector<double> &rcost, std::vector<WorkItem> &items) {
  double total = 0.0;
  for (size_t c = 0; c < classes.size(); ++c) {
    double w = rcost[c];
    for (int a = classes[c].ab_begin; a < classes[c].ab_end; ++a) w += abcost[a];
    total += w * classes[c].nbas[0] * classes[c].nbas[1];
  }
  const double target = total / 8192.0;
  std::vector<std::pair<double, WorkItem>> runs;
  for (size_t c = 0; c < classes.size(); ++c) {
    const ClassDesc &cl = classes[c];
    int lo = cl.ab_begin;
    while (lo < cl.ab_end) {
      double w = rcost[c] + abcost[lo];
      int hi = lo + 1;
      while (hi < cl.ab_end && w + abcost[hi] <= target) w += abcost[hi++];
      for (int ij = 0; ij < cl.nbas[0] * cl.nbas[1]; ++ij)
        runs.push_back({w, WorkItem{(int)c, ij, lo, hi}});
      lo = hi;
    }
  }
  std::stable_sort(runs.begin(), runs.end(),
                   [](const std::pair<double, WorkItem> &x, const std::pair<double, WorkItem> &y) {
                     return x.first > y.first;
                   });
  items.clear();
  items.reserve(runs.size());
  for (const auto &r : runs) items.push_back(r.second);
}

// Hermite bookkeeping of the matrix-core kernel's signature blocks, in
// CARTIND order: components of index h, index of (t,u,v), parity class.
namespace {
struct HermiteOrder {
  std::vector<int> t, u, v;
  HermiteOrder() {
    for (int L = 0; L <= kMaxLfl; ++L)
      for (int a = 0; a <= L; ++a)
        for (int b = 0; b <= L - a; ++b) {
          t.push_back(a);
          u.push_back(b);
          v.push_back(L - a - b);
        }
  }
  static int index(int a, int b, int c) {
    const int L = a + b + c;
    return L * (L + 1) * (L + 2) / 6 + a * (L + 1) - a * (a - 1) / 2 + b;
  }
  int cls(int h, int m) const { return ((t[h] & 1) | ((u[h] & 1) << 1) | ((v[h] & 1) << 2)) & m; }
  // h < N(L) class by class, ascending inside a class
  std::vector<int> by_class(int m, int L, std::vector<int> &coff) const {
    std::vector<int> out;
    coff.assign(9, 0);
    for (int c = 0; c < 8; ++c) {
      coff[c] = (int)out.size();
      for (int h = 0; h < ntuv(L); ++h)
        if (cls(h, m) == c) out.push_back(h);
    }
    coff[8] = (int)out.size();
    return out;
  }
  // rank of h inside its class
  int rank(int h, int m) const {
    int r = 0;
    for (int x = 0; x < h; ++x) r += cls(x, m) == cls(h, m);
    return r;
  }
};
}  // namespace

std::string build_plan(const Basis &B, const Etabs &T, const std::vector<uint8_t> &eflags,
                       int isym, int intsym, int racah1, int ilev, Plan &P,
                       const std::vector<uint8_t> *zflags) {
  P = Plan();
  P.ndim = B.ndim;
  P.nskp = B.nskp;
  P.isym = isym;
  P.intsym = intsym;
  P.racah1 = racah1;
  P.ilev = ilev;
  // IQ block index (CLMFAST :11606-11626)
  std::map<std::tuple<int, int, int>, int> iqidx;
  int cnt = 0;
  for (int c = 0; c < B.ncnt; ++c)
    for (int k = 1; k <= 2 * B.lmax[c] + 1; ++k) {
      int kq = B.kappa_of(c, k);
      for (int mj = 1; mj <= 2 * std::abs(kq) - 1; mj += 2) iqidx[{c, kq, mj}] = ++cnt;
    }
  auto kmask_at = [&](int pair, int tuv, int comp) -> int {
    return eflags[T.flag_off[pair] + 4 * tuv + comp];
  };
  for (int ca = 0; ca < B.ncnt; ++ca) {
    int ncb = intsym ? ca + 1 : B.ncnt;
    for (int cb = 0; cb < ncb; ++cb)
      for (int cc = 0; cc < B.ncnt; ++cc)
        for (int cd = 0; cd < B.ncnt; ++cd) {
          int mcnt = ncntrs(ca, cb, cc, cd);
          if (mcnt == 1 && racah1) continue;
          int ic[4] = {ca, cb, cc, cd};
          for (int la = 0; la <= B.lmax[ca]; ++la)
          for (int lb = 0; lb <= B.lmax[cb]; ++lb)
          for (int lc = 0; lc <= B.lmax[cc]; ++lc)
          for (int ld = 0; ld <= B.lmax[cd]; ++ld) {
            if (mcnt == 1 && ((la + lb) % 2) != ((lc + ld) % 2)) continue;
            int lq[4] = {la, lb, lc, ld};
            ClassDesc cl;
            for (int k = 0; k < 4; ++k) {
              for (int x = 0; x < 3; ++x) cl.xyz[k][x] = B.xyz[3 * ic[k] + x];
              cl.exp[k] = B.expo(ic[k], lq[k]);
              cl.nbas[k] = B.nfn(ic[k], lq[k]);
            }
            int lsum = la + lb + lc + ld;
            if (mcnt == 1) cl.lamfl = lsum + 4;
            else cl.lamfl = ilev == 1 ? lsum : (ilev == 2 ? lsum + 2 : lsum + 4);
            cl.maxcd = cl.nbas[2] * cl.nbas[3];
            cl.mcnt = mcnt;
            cl.flagoff = 0;
            cl.sym = axis_symmetry(cl);
            cl.pad = 0;
            cl.ab_begin = (int)P.abs.size();
            // aggregated E(CD) flags of the (C,D,lc,ld) class (E0LLCD0/E0SSCD0)
            int ntss = ntuv(lc + ld + 2);
            std::vector<int> agg[2];
            for (int s = 0; s < 2; ++s) {
              agg[s].assign(4 * ntss, 0);
              // at stage 1 the reference never refreshes the SS lists
              // (E0SSCD0 is skipped, bertha-scf.f:11742): a fresh run sees
              // them zero, so one-centre G built from SS flags vanish
              if (s == 1 && ilev == 1) continue;
              for (int nc = (lc == 0); nc <= 1; ++nc)
                for (int nd = (ld == 0); nd <= 1; ++nd) {
                  int kc = 2 * lc + nc, kd = 2 * ld + nd;
                  for (int mc = 1; mc <= lc + nc; ++mc)
                    for (int md = 1; md <= ld + nd; ++md) {
                      int pr = T.pair_index.at(Etabs::key(cc, cd, kc, kd, mc, md, s));
                      int nt = ntuv(lc + ld + (s ? 2 : 0));
                      for (int t = 0; t < nt; ++t)
                        for (int c = 0; c < 4; ++c) agg[s][4 * t + c] += kmask_at(pr, t, c);
                    }
                }
            }
            for (int na = (la == 0); na <= 1; ++na) {
              int ka = 2 * la + na, kqa = B.kappa_of(ca, ka), jqa = 2 * std::abs(kqa) - 1;
              for (int nb = (lb == 0); nb <= 1; ++nb) {
                int kb = 2 * lb + nb, kqb = B.kappa_of(cb, kb), jqb = 2 * std::abs(kqb) - 1;
                for (int ma = 1; ma <= std::abs(kqa); ++ma)
                  for (int mb = 1; mb <= std::abs(kqb); ++mb) {
                    int mqa = 2 * ma - 1, mqb = 2 * mb - 1;
                    for (int it1 = 4; it1 >= 1; it1 -= 3) {
                      int naddab = it1 == 1 ? 0 : B.nskp;
                      ABCombo ab{};
                      int pab_pair = T.pair_index.at(Etabs::key(ca, cb, ka, kb, ma, mb, it1 == 4));
                      ab.eoff = T.pairs[pab_pair].off;
                      ab.ntab = ntuv(la + lb + (it1 == 4 ? 2 : 0));
                      ab.fa[0] = B.lrge_at(ca, ka, 2 * ma - 1) + naddab;
                      ab.fa[1] = B.lrge_at(ca, ka, 2 * ma) + naddab;
                      ab.fb[0] = B.lrge_at(cb, kb, 2 * mb - 1) + naddab;
                      ab.fb[1] = B.lrge_at(cb, kb, 2 * mb) + naddab;
                      ab.nb = cl.nbas[1];
                      ab.cd_begin = (int)P.cds.size();
                      ab.lab = la + lb + (it1 == 4 ? 2 : 0);
                      int first_it2 = 0;
                      for (int nc = (lc == 0); nc <= 1; ++nc) {
                        int kc = 2 * lc + nc, kqc = B.kappa_of(cc, kc), jqc = 2 * std::abs(kqc) - 1;
                        for (int nd = (ld == 0); nd <= 1; ++nd) {
                          int kd = 2 * ld + nd, kqd = B.kappa_of(cd, kd), jqd = 2 * std::abs(kqd) - 1;
                          if (mcnt == 1) {
                            int nui = std::max(std::abs(jqa - jqb) / 2, std::abs(jqc - jqd) / 2);
                            int nuf = std::min((jqa + jqb) / 2, (jqc + jqd) / 2);
                            if (nui > nuf) continue;
                            bool selk = false;
                            for (int nu = nui; nu <= nuf; ++nu)
                              if ((la + lb + nu) % 2 == 0 && (lc + ld + nu) % 2 == 0) selk = true;
                            if (!selk) continue;
                          }
                          for (int mc = 1; mc <= std::abs(kqc); ++mc)
                            for (int md = 1; md <= std::abs(kqd); ++md) {
                              int mqc = 2 * mc - 1, mqd = 2 * md - 1;
                              if (isym == 1 || isym == 2) {
                                bool ok = (mqa == mqb && mqc == mqd) || (mqa == mqc && mqb == mqd) ||
                                          (mqa == mqd && mqb == mqc);
                                if (!ok) continue;
                              }
                              if (mcnt == 1) {
                                bool selm = false;
                                for (int sa = -1; sa <= 1; sa += 2)
                                  for (int sb = -1; sb <= 1; sb += 2)
                                    for (int sc = -1; sc <= 1; sc += 2)
                                      for (int sd = -1; sd <= 1; sd += 2)
                                        if (sa * mqa - sb * mqb == sd * mqd - sc * mqc) selm = true;
                                if (!selm) continue;
                              }
                              int iq1 = iqidx.at({ca, kqa, mqa}), iq2 = iqidx.at({cb, kqb, mqb});
                              int iq3 = iqidx.at({cc, kqc, mqc}), iq4 = iqidx.at({cd, kqd, mqd});
                              int iql = iq1 * (iq1 - 1) / 2 + iq2, iqr = iq3 * (iq3 - 1) / 2 + iq4;
                              if (intsym && (iq1 < iq2 || iq3 < iq4 || iql < iqr)) continue;
                              int itscf;
                              if (iq1 > iq2 && iq3 > iq4 && iql > iqr) itscf = 1;
                              else if (iq1 > iq2 && iq3 > iq4 && iql == iqr) itscf = 2;
                              else if (iq1 > iq2 && iq3 == iq4 && iql > iqr) itscf = 3;
                              else if (iq1 == iq2 && iq3 > iq4 && iql > iqr) itscf = 4;
                              else if (iq1 == iq2 && iq3 == iq4 && iql > iqr) itscf = 5;
                              else if (iq1 == iq2 && iq3 == iq4 && iql == iqr) itscf = 6;
                              else {
                                if (intsym) continue;
                                itscf = 6;
                              }
                              int iflg = kISCF[itscf - 1];
                              if (iq2 == iq3 && (itscf == 1 || itscf == 3 || itscf == 4)) iflg |= 1 << 8;
                              if (itscf == 1) {
                                if (iq1 == iq3) iflg |= 1 << 9;
                                else if (iq2 == iq4) iflg |= 1 << 10;
                              }
                              int sab = isgn(kqa * kqb), scd = isgn(kqc * kqd);
                              for (int it2 = 4; it2 >= 1; it2 -= 3) {
                                int itt = kMAPTTTT[it1 - 1][it2 - 1];
                                if (ilev == 1 && itt > 1 && mcnt != 1) continue;
                                if (ilev == 2 && itt > 3 && mcnt != 1) continue;
                                int naddcd = it2 == 1 ? 0 : B.nskp;
                                CDCall call{};
                                int pcd_pair = T.pair_index.at(Etabs::key(cc, cd, kc, kd, mc, md, it2 == 4));
                                call.eoff = T.pairs[pcd_pair].off;
                                call.ntcd = ntuv(lc + ld + (it2 == 4 ? 2 : 0));
                                call.fc[0] = B.lrge_at(cc, kc, 2 * mc - 1) + naddcd;
                                call.fc[1] = B.lrge_at(cc, kc, 2 * mc) + naddcd;
                                call.fd[0] = B.lrge_at(cd, kd, 2 * md - 1) + naddcd;
                                call.fd[1] = B.lrge_at(cd, kd, 2 * md) + naddcd;
                                call.iflg = iflg;
                                call.pab[0] = (int8_t)(sab * m1pow((-mqa + mqb) / 2));
                                call.pab[1] = (int8_t)(sab * m1pow((mqa + mqb) / 2));
                                call.pcd[0] = (int8_t)(scd * m1pow((-mqc + mqd) / 2));
                                call.pcd[1] = (int8_t)(scd * m1pow((mqc + mqd) / 2));
                                if (!first_it2) first_it2 = it2;
                                P.cds.push_back(call);
                                P.quartets += (double)cl.nbas[0] * cl.nbas[1] * cl.maxcd;
                              }
                            }
                        }
                      }
                      ab.cd_end = (int)P.cds.size();
                      if (ab.cd_end == ab.cd_begin) continue;
                      // G is built once per (ij) with the flags of the first
                      // call's component type (ERIFAST IGAB, :12674-12799)
                      int gs = first_it2 == 4 ? 1 : 0;
                      ab.ntg = ntuv(lc + ld + (gs ? 2 : 0));
                      ab.gmask = (int)P.masks.size();
                      for (int t = 0; t < ab.ntg; ++t) {
                        const int *g = &agg[gs][4 * t];
                        int all = g[0] + g[1] + g[2] + g[3];
                        for (int c = 0; c < 4; ++c)
                          P.masks.push_back((uint8_t)(all != 0 && (isym <= 1 || g[c] != 0)));
                      }
                      P.max_ntg = std::max(P.max_ntg, ab.ntg);
                      P.abs.push_back(ab);
                    }
                  }
              }
            }
            cl.ab_end = (int)P.abs.size();
            if (cl.ab_end == cl.ab_begin) continue;
            // combinations of one shape (N(L_AB), N(L_CD) of G) side by side,
            // so the kernel's first contraction can share each R load among
            // several of them; only the order of accumulation changes
            std::stable_sort(P.abs.begin() + cl.ab_begin, P.abs.begin() + cl.ab_end,
                             [](const ABCombo &x, const ABCombo &y) {
                               return x.ntab != y.ntab ? x.ntab > y.ntab : x.ntg > y.ntg;
                             });
            P.classes.push_back(cl);
            P.max_lfl = std::max(P.max_lfl, cl.lamfl);
            P.max_maxcd = std::max(P.max_maxcd, cl.maxcd);
          }
        }
  }
  // Algorithmic FP64 work of one build, counted per unit of the reference's
  // loop nest: R batch per (ij, class) MAXCD*(120 + 2L + 6*C(L+3,4) + 3N(L));
  // first contraction per (ij, AB combination) 8*N(L_AB)*N(L_CD)*MAXCD;
  // second contraction per (ij, call) 32*N(L_CD)*MAXCD (8 on the real
  // linear/atomic path) and digestion 1408 (88 real) flops per RR batch
  // element.
  for (const ClassDesc &cl : P.classes) {
    double maxab = (double)cl.nbas[0] * cl.nbas[1], maxcd = cl.maxcd;
    double L = cl.lamfl;
    double c4 = (L + 3) * (L + 2) * (L + 1) * L / 24.0;
    P.flops += maxab * maxcd * (120.0 + 2.0 * L + 6.0 * c4 + 3.0 * ntuv(cl.lamfl));
    for (int a = cl.ab_begin; a < cl.ab_end; ++a) {
      const ABCombo &ab = P.abs[a];
      P.flops += maxab * 8.0 * ab.ntab * ab.ntg * maxcd;
      for (int k = ab.cd_begin; k < ab.cd_end; ++k) {
        double c2 = (isym ? 8.0 : 32.0) * P.cds[k].ntcd;
        double dig = isym ? 88.0 : 1408.0;
        P.flops += maxab * maxcd * (c2 + dig);
      }
    }
  }
  // Work items: (class, ij) split into runs of AB combinations so that no
  // item is much longer than the average share of a workgroup (each item
  // recomputes its R batch, which costs far less than one combination's
  // contractions); largest first, so the persistent kernel's queue drains
  // evenly.
  std::vector<double> abcost(P.abs.size(), 0.0), rcost(P.classes.size(), 0.0);
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    rcost[c] = (double)ntuv(cl.lamfl) * cl.maxcd;
    for (int a = cl.ab_begin; a < cl.ab_end; ++a) {
      const ABCombo &ab = P.abs[a];
      double w = (double)ab.ntab * ab.ntg * cl.maxcd;
      for (int k = ab.cd_begin; k < ab.cd_end; ++k) w += 8.0 * P.cds[k].ntcd * cl.maxcd + 64.0 * cl.maxcd;
      abcost[a] = w;
    }
  }
  make_items(P.classes, abcost, rcost, P.items);
  // one nonzero component per parity class of t in each E(AB) and E(CD)
  // table (axial systems, from the exact-zero flags): the component of G in
  // t''s class is E(AB)'s in the same class, since R(t + t') vanishes off
  // class 0
  if (zflags && isym != 0) {
    static const HermiteOrder H;
    std::map<int64_t, int> pair_of;
    for (size_t p = 0; p < T.pairs.size(); ++p) pair_of[T.pairs[p].off] = (int)p;
    auto bits_at = [&](int p, int t) {
      int b = 0;
      for (int c = 0; c < 4; ++c) b |= ((*zflags)[T.flag_off[p] + 4 * t + c] != 0) << c;
      return b;
    };
    P.one_comp = 1;
    for (const ClassDesc &cl : P.classes)
      for (int a = cl.ab_begin; a < cl.ab_end && P.one_comp; ++a) {
        ABCombo &ab = P.abs[a];
        const int pa = pair_of.at(ab.eoff);
        int cb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < ab.ntab; ++t) cb[H.cls(t, cl.sym)] |= bits_at(pa, t);
        ab.gcomp = 0;
        for (int c = 0; c < 8; ++c) {
          if (__builtin_popcount(cb[c]) > 1) P.one_comp = 0;
          ab.gcomp |= (uint32_t)(cb[c] ? __builtin_ctz(cb[c]) : 4) << (4 * c);
        }
        for (int k = ab.cd_begin; k < ab.cd_end; ++k) {
          CDCall &cd = P.cds[k];
          const int pc = pair_of.at(cd.eoff);
          int cc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          for (int t = 0; t < cd.ntcd; ++t) cc[H.cls(t, cl.sym)] |= bits_at(pc, t);
          // the kernel walks each class of t' through one 64-bit mask
          int lcd = 0, sz[8];
          while (ntuv(lcd) < cd.ntcd) ++lcd;
          hermite_class_sizes(cl.sym, lcd, sz);
          if (*std::max_element(sz, sz + 8) > 64) P.one_comp = 0;
          cd.ccomp = 0;
          for (int c = 0; c < 8; ++c) {
            if (__builtin_popcount(cc[c]) > 1) P.one_comp = 0;
            cd.ccomp |= (uint32_t)(cc[c] ? __builtin_ctz(cc[c]) : 4) << (4 * c);
          }
        }
      }
  }
  return check_offsets(P);
}

// Tiles of the matrix-core Breit kernel (real path): per class, the first
// combination alone (it reads this pair's own R-column flags, and also joins
// a compatible stale tile as an extra row), then the others sorted by their
// G components and the CD blocks they call and packed up to kTcRows to a
// tile while the tile's calls stay within kTcMaxU, so that combinations
// sharing their calls share one E(CD) load.  Work items are runs of a class's
// tiles for one primitive pair ij, costed like the VALU kernel's items.

// One signature block of the matrix-core Breit kernel for (axis mask m,
// L_AB, L_CD): the ket's Hermite indices t' in groups of up to kTcR of one
// parity class; the steps (group, ix) dealt to the kTcWaves waves of a
// workgroup by their matrix-product count (largest first to the least
// loaded wave); for each group the compact table row of t + t' for every
// bra t (class order) and group member r; and for each (ix, ket class) the
// first contraction's chunks of four bra indices, table by table.  Layout,
// in ints: ngrp, wave ranges gw[0..kTcWaves] into the step list, the chunk
// region's offset, then per group (ctp, p0, n, chunks), then the step list
// (3 ngrp codes group * 3 + ix, wave by wave), then rows [ngrp][kTcR][nab +
// 1] as 16-bit pairs, then the chunk region: 25 offsets (ix * 8 + class, and
// the end) and the chunks, tab | src << 4 | pos << 6 | count << 14 (tab: W
// table, src: E^j table, pos: first bra position, count: positions of the
// chunk, 1..4).
static void tc_signature(const HermiteOrder &H, int m, int lab, int lcd, int gaunt,
                         std::vector<int> &out, int &ngrp_out) {
  std::vector<int> cab, ccd;
  const std::vector<int> tab = H.by_class(m, lab, cab), tcd = H.by_class(m, lcd, ccd);
  const int nab = ntuv(lab), Lw = lab + lcd;
  int sz[8];
  hermite_class_sizes(m, Lw, sz);
  const int zrow = *std::max_element(sz, sz + 8);
  struct Grp { int ctp, p0, n, cost; };
  std::vector<Grp> gs;
  for (int c = 0; c < 8; ++c)
    for (int p = ccd[c]; p < ccd[c + 1]; p += kTcR) {
      Grp g{c, p, std::min(kTcR, ccd[c + 1] - p), 0};
      for (int ix = 0; ix < 3; ++ix)
        for (int ty = 0; ty < (gaunt ? 1 : 4); ++ty) {
          const int sh = ty == 0 ? 0 : (((1 << ix) ^ (1 << (ty - 1))) & m);
          const int tc = c ^ sh;
          g.cost += (cab[tc + 1] - cab[tc] + 3) / 4;
        }
      gs.push_back(g);
    }
  // steps (group, ix) dealt to the waves, largest first to the least loaded:
  // cost = the first contraction's chunk products plus the second's
  struct Step { int code, cost; };
  std::vector<Step> steps;
  for (size_t gi = 0; gi < gs.size(); ++gi)
    for (int ix = 0; ix < 3; ++ix) {
      int cost = kTcQ * kTcR / 4;
      for (int ty = 0; ty < (gaunt ? 1 : 4); ++ty) {
        const int sh = ty == 0 ? 0 : (((1 << ix) ^ (1 << (ty - 1))) & m);
        const int tc = gs[gi].ctp ^ sh;
        cost += (cab[tc + 1] - cab[tc] + 3) / 4;
      }
      steps.push_back({(int)gi * 3 + ix, cost});
    }
  std::stable_sort(steps.begin(), steps.end(),
                   [](const Step &x, const Step &y) { return x.cost > y.cost; });
  std::vector<std::vector<int>> wv(kTcWaves);
  int load[kTcWaves] = {};
  for (const Step &st : steps) {
    const int w = (int)(std::min_element(load, load + kTcWaves) - load);
    wv[w].push_back(st.code);
    load[w] += st.cost;
  }
  out.clear();
  out.push_back((int)gs.size());
  int acc = 0;
  for (int w = 0; w < kTcWaves; ++w) {
    out.push_back(acc);
    acc += (int)wv[w].size();
  }
  out.push_back(acc);
  const size_t choff_at = out.size();
  out.push_back(0);
  for (const Grp &g : gs) {
    out.push_back(g.ctp);
    out.push_back(g.p0);
    out.push_back(g.n);
    out.push_back(g.cost);
  }
  for (int w = 0; w < kTcWaves; ++w) out.insert(out.end(), wv[w].begin(), wv[w].end());
  std::vector<int> seq(gs.size());
  for (size_t i = 0; i < gs.size(); ++i) seq[i] = (int)i;
  std::vector<uint16_t> rows;
  for (int gi : seq)
    for (int r = 0; r < kTcR; ++r)
      for (int pos = 0; pos <= nab; ++pos) {
        int row = zrow;
        if (pos < nab && r < gs[gi].n) {
          const int a = tab[pos], b = tcd[gs[gi].p0 + r];
          const int h = HermiteOrder::index(H.t[a] + H.t[b], H.u[a] + H.u[b], H.v[a] + H.v[b]);
          row = H.rank(h, m);
        }
        rows.push_back((uint16_t)row);
      }
  if (rows.size() & 1) rows.push_back(0);
  for (size_t i = 0; i < rows.size(); i += 2) out.push_back((int)(rows[i] | ((uint32_t)rows[i + 1] << 16)));
  out[choff_at] = (int)out.size();
  const size_t hdr = out.size();
  out.resize(hdr + 25, 0);
  std::vector<int> ch;
  for (int ix = 0; ix < 3; ++ix)
    for (int c = 0; c < 8; ++c) {
      out[hdr + ix * 8 + c] = (int)ch.size();
      for (int ty = 0; ty < (gaunt ? 1 : 4); ++ty) {
        const int src = ty == 0 ? ix : ty - 1;
        const int tb = ty == 0 ? 0 : 1 + 3 * ix + (ty - 1);
        const int sh = ty == 0 ? 0 : (((1 << ix) ^ (1 << (ty - 1))) & m);
        const int tc = c ^ sh;
        for (int p = cab[tc]; p < cab[tc + 1]; p += 4)
          ch.push_back(tb | (src << 4) | (p << 6) | (std::min(4, cab[tc + 1] - p) << 14));
      }
    }
  out[hdr + 24] = (int)ch.size();
  out.insert(out.end(), ch.begin(), ch.end());
  ngrp_out = (int)gs.size();
}

static void build_breit_tiles(BreitPlan &P, int gaunt, const Etabs &T,
                              const std::vector<uint8_t> &zflags) {
  P.tiles.clear();
  P.tile_cd.clear();
  P.tile_pair.clear();
  P.comp.clear();
  P.titems.clear();
  P.rest_items.clear();
  P.tc_sig.clear();
  P.tc_csig.assign(P.classes.size(), 0);
  P.max_nu = P.max_npair = 0;
  P.tc_lds = 0;
  P.tc_ok = 0;
  static const HermiteOrder H;
  // components: kap from the exact-zero flags of each LS table, cout from kap
  P.comp.resize(P.abs.size());
  for (const ClassDesc &cl : P.classes) {
    const int m = cl.sym;
    int lab = 0;
    while (ntuv(lab) < P.abs[cl.ab_begin].ntab) ++lab;
    for (int a = cl.ab_begin; a < cl.ab_end; ++a) {
      BrComp &bc = P.comp[a];
      for (int j = 0; j < 3; ++j) {
        const int pr = P.ab_pair[3 * a + j];
        int bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < ntuv(lab); ++t)
          for (int c = 0; c < 4; ++c)
            if (zflags[T.flag_off[pr] + 4 * t + c]) bits[H.cls(t, m)] |= 1 << c;
        for (int c = 0; c < 8; ++c) {
          if (__builtin_popcount(bits[c]) > 1) return;  // not one component per class
          bc.kap[j][c] = bits[c] ? (uint8_t)__builtin_ctz(bits[c]) : (uint8_t)kTcNone;
        }
      }
      for (int ix = 0; ix < 3; ++ix)
        for (int c = 0; c < 8; ++c) {
          int out = bc.kap[ix][c];
          if (!gaunt)
            for (int j = 0; j < 3; ++j) {
              const int k = bc.kap[j][c ^ (((1 << ix) ^ (1 << j)) & m)];
              if (k == kTcNone) continue;
              if (out != kTcNone && out != k) return;  // G^ix with two components
              out = k;
            }
          bc.cout[ix][c] = (uint8_t)out;
        }
    }
  }
  std::vector<double> tcost;
  std::vector<std::pair<int, int>> ctile(P.classes.size());
  // -> (offset, groups, chunks)
  std::map<std::tuple<int, int, int>, std::tuple<int, int, int>> sigs;
  std::vector<char> fits(P.classes.size(), 0);
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    auto calls = [&](int a) {
      std::vector<int64_t> s;
      for (int k = P.abs[a].cd_begin; k < P.abs[a].cd_end; ++k) s.push_back(P.cds[k].eoff[0]);
      std::sort(s.begin(), s.end());
      return s;
    };
    // rows of one tile share the component of G^ix per ket class (cout,
    // where defined): combinations are sorted by cout, then by their calls,
    // and a tile closes when it is full, its calls would exceed the slots,
    // or the next combination's cout conflicts with the tile's
    std::vector<std::vector<int>> groups;
    groups.push_back({cl.ab_begin});
    auto csig = [&](int a) {
      return std::vector<uint8_t>(&P.comp[a].cout[0][0], &P.comp[a].cout[0][0] + 24);
    };
    std::vector<std::tuple<std::vector<uint8_t>, std::vector<int64_t>, int>> rest;
    for (int a = cl.ab_begin + 1; a < cl.ab_end; ++a) rest.push_back({csig(a), calls(a), a});
    std::stable_sort(rest.begin(), rest.end());
    std::vector<int> cur;
    std::vector<int64_t> uni;
    uint8_t tco[24];
    for (const auto &r : rest) {
      const std::vector<uint8_t> &co = std::get<0>(r);
      std::vector<int64_t> u2 = uni;
      for (int64_t x : std::get<1>(r))
        if (std::find(u2.begin(), u2.end(), x) == u2.end()) u2.push_back(x);
      bool conflict = false;
      if (!cur.empty())
        for (int k = 0; k < 24; ++k)
          if (co[k] != kTcNone && tco[k] != kTcNone && co[k] != tco[k]) conflict = true;
      if (!cur.empty() &&
          (cur.size() == (size_t)kTcRows || u2.size() > (size_t)kTcMaxU || conflict)) {
        groups.push_back(cur);
        cur.clear();
        uni.clear();
        u2 = std::get<1>(r);
      }
      if (cur.empty())
        for (int k = 0; k < 24; ++k) tco[k] = (uint8_t)kTcNone;
      for (int k = 0; k < 24; ++k)
        if (co[k] != kTcNone) tco[k] = co[k];
      cur.push_back(std::get<2>(r));
      uni = u2;
    }
    if (!cur.empty()) groups.push_back(cur);
    // the first combination also joins a stale tile with a free row, no
    // conflicting cout and room for its calls
    int mg = -1;
    {
      const std::vector<uint8_t> fco = csig(cl.ab_begin);
      const std::vector<int64_t> fca = calls(cl.ab_begin);
      for (size_t gi = 1; gi < groups.size() && mg < 0; ++gi) {
        if (groups[gi].size() >= (size_t)kTcRows) continue;
        bool ok = true;
        std::vector<int64_t> u = fca;
        for (int a : groups[gi]) {
          const std::vector<uint8_t> co = csig(a);
          for (int k = 0; k < 24; ++k)
            if (fco[k] != kTcNone && co[k] != kTcNone && fco[k] != co[k]) ok = false;
          for (int64_t x : calls(a))
            if (std::find(u.begin(), u.end(), x) == u.end()) u.push_back(x);
        }
        if (ok && u.size() <= (size_t)kTcMaxU) mg = (int)gi;
      }
      if (mg > 0) groups[mg].push_back(cl.ab_begin);
    }
    ctile[c].first = (int)P.tiles.size();
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      BrTile t;
      std::vector<int64_t> ids;
      for (int g = 0; g < kTcRows; ++g) t.ab[g] = g < (int)groups[gi].size() ? groups[gi][g] : -1;
      for (int g = 0; g < kTcRows; ++g)
        if (t.ab[g] >= 0)
          for (int k = P.abs[t.ab[g]].cd_begin; k < P.abs[t.ab[g]].cd_end; ++k)
            if (std::find(ids.begin(), ids.end(), P.cds[k].eoff[0]) == ids.end())
              ids.push_back(P.cds[k].eoff[0]);
      if (ids.size() > (size_t)kTcMaxU) return;  // one combination beyond the call slots
      t.nu = (int)ids.size();
      t.fresh = gi == 0;
      t.fresh_row = gi == 0 ? (mg > 0 ? (int)groups[mg].size() - 1 : -1)
                            : ((int)gi == mg ? (int)groups[gi].size() - 1 : -1);
      for (int ix = 0; ix < 3; ++ix)
        for (int cc = 0; cc < 8; ++cc) {
          t.cout[ix][cc] = (uint8_t)kTcNone;
          for (int g = 0; g < kTcRows; ++g)
            if (t.ab[g] >= 0 && P.comp[t.ab[g]].cout[ix][cc] != kTcNone)
              t.cout[ix][cc] = P.comp[t.ab[g]].cout[ix][cc];
        }
      t.pair_begin = (int)P.tile_pair.size();
      std::vector<int> cdi((size_t)kTcMaxU * kTcRows, -1);
      double w = 0.0;
      for (int g = 0; g < kTcRows; ++g) {
        if (t.ab[g] < 0) continue;
        const BrAB &ab = P.abs[t.ab[g]];
        w += 30.0 * ab.ntab * ab.ntg * cl.maxcd / 4.0;
        for (int k = ab.cd_begin; k < ab.cd_end; ++k) {
          const int u = (int)(std::find(ids.begin(), ids.end(), P.cds[k].eoff[0]) - ids.begin());
          cdi[(size_t)u * kTcRows + g] = k;
          w += 6.0 * P.cds[k].ntcd * cl.maxcd + 64.0 * cl.maxcd;
        }
      }
      for (int u = 0; u < kTcMaxU; ++u)
        for (int g = 0; g < kTcRows; ++g)
          if (cdi[(size_t)u * kTcRows + g] >= 0) P.tile_pair.push_back(u | (g << 8));
      t.npair = (int)P.tile_pair.size() - t.pair_begin;
      P.max_nu = std::max(P.max_nu, t.nu);
      P.max_npair = std::max(P.max_npair, t.npair);
      P.tiles.push_back(t);
      P.tile_cd.insert(P.tile_cd.end(), cdi.begin(), cdi.end());
      tcost.push_back(w);
    }
    ctile[c].second = (int)P.tiles.size();
  }
  // signature blocks and LDS per class; classes beyond a compute unit's LDS
  // go to the vector-ALU kernel
  std::vector<std::pair<double, WorkItem>> runs, rrest;
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    const BrAB &ab0 = P.abs[cl.ab_begin];
    int lab = 0, lcd = 0;
    while (ntuv(lab) < ab0.ntab) ++lab;
    while (ntuv(lcd) < ab0.ntg) ++lcd;
    auto key = std::make_tuple(cl.sym, lab, lcd);
    if (!sigs.count(key)) {
      std::vector<int> blk;
      int ng = 0;
      tc_signature(H, cl.sym, lab, lcd, gaunt, blk, ng);
      for (int w = 0; w < kTcWaves; ++w)
        if (blk[2 + w] - blk[1 + w] > 63) return;  // a wave's steps beyond one lane each
      // a step's code, first t' and size fit the kernel's packed step word
      if (3 * ng > 256 || ntuv(lcd) > 512) return;
      sigs[key] = {(int)P.tc_sig.size(), ng, blk[blk[2 + kTcWaves] + 24]};
      P.tc_sig.insert(P.tc_sig.end(), blk.begin(), blk.end());
    }
    P.tc_csig[c] = std::get<0>(sigs[key]);
    int sz[8], szr[8];
    hermite_class_sizes(cl.sym, lab + lcd, sz);
    hermite_class_sizes(cl.sym, cl.lamfl, szr);
    const int wrows = *std::max_element(sz, sz + 8) + 1;
    const TcLds lo = tc_lds_layout(ab0.ntab, wrows, szr[0], ntuv(cl.lamfl), gaunt ? 1 : 10,
                                   std::get<1>(sigs[key]), ab0.ntg, P.max_npair,
                                   std::get<2>(sigs[key]));
    fits[c] = lo.bytes <= kTcLdsMax;
    if (fits[c]) P.tc_lds = std::max(P.tc_lds, (size_t)lo.bytes);
  }
  // items: runs of consecutive tiles of one class, about 1/8192 of the build each
  double total = 0.0;
  for (size_t c = 0; c < P.classes.size(); ++c) {
    if (!fits[c]) continue;
    const ClassDesc &cl = P.classes[c];
    double r = (double)ntuv(cl.lamfl) * cl.maxcd;
    for (int t = ctile[c].first; t < ctile[c].second; ++t) r += tcost[t];
    total += r * cl.nbas[0] * cl.nbas[1];
  }
  const double target = total / 8192.0;
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    if (!fits[c]) continue;
    const double rc = (double)ntuv(cl.lamfl) * cl.maxcd;
    int lo = ctile[c].first;
    while (lo < ctile[c].second) {
      double w = rc + tcost[lo];
      int hi = lo + 1;
      while (hi < ctile[c].second && w + tcost[hi] <= target) w += tcost[hi++];
      for (int ij = 0; ij < cl.nbas[0] * cl.nbas[1]; ++ij)
        runs.push_back({w, WorkItem{(int)c, ij, lo, hi}});
      lo = hi;
    }
  }
  std::stable_sort(runs.begin(), runs.end(),
                   [](const std::pair<double, WorkItem> &x, const std::pair<double, WorkItem> &y) {
                     return x.first > y.first;
                   });
  P.titems.reserve(runs.size());
  for (const auto &r : runs) P.titems.push_back(r.second);
  for (const WorkItem &w : P.items)
    if (!fits[w.cls]) P.rest_items.push_back(w);
  P.tc_ok = 1;
}

// ---------------------------------------------------------------------------
// the Breit plan: BRTFAST's loop nest (bertha-scf.f:17185-17600)
// ---------------------------------------------------------------------------
std::string build_breit_plan(const Basis &B, const Etabs &T, const std::vector<uint8_t> &eflags,
                             int isym, int intsym, int racah1, int ilev, int gaunt,
                             BreitPlan &P, const std::vector<uint8_t> *zflags) {
  P = BreitPlan();
  std::map<std::tuple<int, int, int>, int> iqidx;
  int cnt = 0;
  for (int c = 0; c < B.ncnt; ++c)
    for (int k = 1; k <= 2 * B.lmax[c] + 1; ++k) {
      int kq = B.kappa_of(c, k);
      for (int mj = 1; mj <= 2 * std::abs(kq) - 1; mj += 2) iqidx[{c, kq, mj}] = ++cnt;
    }
  auto flag_at = [&](int pair, int tuv, int comp) -> int {
    return eflags[T.flag_off[pair] + 4 * tuv + comp];
  };
  auto ls_pair = [&](int a, int b, int ka, int kb, int ma, int mb, int ix) {
    return T.pair_index.at(Etabs::key(a, b, ka, kb, ma, mb, kLSx + ix));
  };
  for (int ca = 0; ca < B.ncnt; ++ca)
    for (int cb = 0; cb < B.ncnt; ++cb)
      for (int cc = 0; cc < B.ncnt; ++cc)
        for (int cd = 0; cd < B.ncnt; ++cd) {
          int mcnt = ncntrs(ca, cb, cc, cd);
          if (mcnt == 1 && racah1) continue;
          if (mcnt != 1 && ilev < 3) continue;
          int ic[4] = {ca, cb, cc, cd};
          for (int la = 0; la <= B.lmax[ca]; ++la)
          for (int lb = 0; lb <= B.lmax[cb]; ++lb)
          for (int lc = 0; lc <= B.lmax[cc]; ++lc)
          for (int ld = 0; ld <= B.lmax[cd]; ++ld) {
            if (mcnt == 1 && ((la + lb) % 2) != ((lc + ld) % 2)) continue;
            int lq[4] = {la, lb, lc, ld};
            ClassDesc cl;
            for (int k = 0; k < 4; ++k) {
              for (int x = 0; x < 3; ++x) cl.xyz[k][x] = B.xyz[3 * ic[k] + x];
              cl.exp[k] = B.expo(ic[k], lq[k]);
              cl.nbas[k] = B.nfn(ic[k], lq[k]);
            }
            int lamab = la + lb + 1, lamcd = lc + ld + 1;
            cl.lamfl = gaunt ? lamab + lamcd : lamab + lamcd + 2;
            cl.maxcd = cl.nbas[2] * cl.nbas[3];
            cl.mcnt = mcnt;
            cl.flagoff = P.nflag;
            cl.sym = axis_symmetry(cl);
            cl.pad = 0;
            cl.ab_begin = (int)P.abs.size();
            int ntcd = ntuv(lamcd);
            // EILSCD0 (:17997): E(CD) flags summed over the class, [tuv][ix][c]
            std::vector<int> agg(12 * ntcd, 0);
            for (int nc = (lc == 0); nc <= 1; ++nc)
              for (int nd = (ld == 0); nd <= 1; ++nd)
                for (int mc = 1; mc <= lc + nc; ++mc)
                  for (int md = 1; md <= ld + nd; ++md)
                    for (int ix = 0; ix < 3; ++ix) {
                      int pr = ls_pair(cc, cd, 2 * lc + nc, 2 * ld + nd, mc, md, ix);
                      for (int t = 0; t < ntcd; ++t)
                        for (int c = 0; c < 4; ++c) agg[(t * 3 + ix) * 4 + c] += flag_at(pr, t, c);
                    }
            for (int na = (la == 0); na <= 1; ++na) {
              int ka = 2 * la + na, kqa = B.kappa_of(ca, ka), jqa = 2 * std::abs(kqa) - 1;
              for (int nb = (lb == 0); nb <= 1; ++nb) {
                int kb = 2 * lb + nb, kqb = B.kappa_of(cb, kb), jqb = 2 * std::abs(kqb) - 1;
                for (int ma = 1; ma <= std::abs(kqa); ++ma)
                  for (int mb = 1; mb <= std::abs(kqb); ++mb) {
                    int mqa = 2 * ma - 1, mqb = 2 * mb - 1;
                    BrAB ab;
                    int abpr[3];
                    for (int ix = 0; ix < 3; ++ix) {
                      abpr[ix] = ls_pair(ca, cb, ka, kb, ma, mb, ix);
                      ab.eoff[ix] = T.pairs[abpr[ix]].off;
                    }
                    ab.ntab = ntuv(lamab);
                    ab.ntg = ntcd;
                    // IT1 = 2: A large, B small (:17432-17440)
                    ab.fa[0] = B.lrge_at(ca, ka, 2 * ma - 1);
                    ab.fa[1] = B.lrge_at(ca, ka, 2 * ma);
                    ab.fb[0] = B.lrge_at(cb, kb, 2 * mb - 1) + B.nskp;
                    ab.fb[1] = B.lrge_at(cb, kb, 2 * mb) + B.nskp;
                    ab.cd_begin = (int)P.cds.size();
                    ab.pad = 0;
                    int sab = isgn(kqa * kqb);
                    for (int nc = (lc == 0); nc <= 1; ++nc) {
                      int kc = 2 * lc + nc, kqc = B.kappa_of(cc, kc), jqc = 2 * std::abs(kqc) - 1;
                      for (int nd = (ld == 0); nd <= 1; ++nd) {
                        int kd = 2 * ld + nd, kqd = B.kappa_of(cd, kd), jqd = 2 * std::abs(kqd) - 1;
                        if (mcnt == 1) {
                          // Breit selection rule (:17484-17499)
                          int nui = std::max(std::abs(jqa - jqb) / 2, std::abs(jqc - jqd) / 2);
                          int nuf = std::min((jqa + jqb) / 2, (jqc + jqd) / 2);
                          if (nui > nuf) continue;
                          bool selk = false;
                          for (int nu = nui; nu <= nuf; ++nu) {
                            int pab_ = (la + lb + nu) % 2, pcd_ = (lc + ld + nu) % 2;
                            if (pab_ == 1 && pcd_ == 1 && nu != 0) selk = true;
                            if (pab_ == 0 && pcd_ == 0) selk = true;
                          }
                          if (!selk) continue;
                        }
                        for (int mc = 1; mc <= std::abs(kqc); ++mc)
                          for (int md = 1; md <= std::abs(kqd); ++md) {
                            int mqc = 2 * mc - 1, mqd = 2 * md - 1;
                            if (isym == 1 || isym == 2) {
                              bool ok = (mqa == mqb && mqc == mqd) || (mqa == mqc && mqb == mqd) ||
                                        (mqa == mqd && mqb == mqc);
                              if (!ok) continue;
                            }
                            if (mcnt == 1) {
                              bool selm = false;
                              for (int sa = -1; sa <= 1; sa += 2)
                                for (int sb = -1; sb <= 1; sb += 2)
                                  for (int sc = -1; sc <= 1; sc += 2)
                                    for (int sd = -1; sd <= 1; sd += 2)
                                      if (sa * mqa - sb * mqb == sd * mqd - sc * mqc) selm = true;
                              if (!selm) continue;
                            }
                            int iq1 = iqidx.at({ca, kqa, mqa}), iq2 = iqidx.at({cb, kqb, mqb});
                            int iq3 = iqidx.at({cc, kqc, mqc}), iq4 = iqidx.at({cd, kqd, mqd});
                            int iql = iq1 * (iq1 - 1) / 2 + iq2, iqr = iq3 * (iq3 - 1) / 2 + iq4;
                            if (intsym && iql < iqr) continue;
                            int scd = isgn(kqc * kqd);
                            BrCD call;
                            for (int ix = 0; ix < 3; ++ix)
                              call.eoff[ix] = T.pairs[ls_pair(cc, cd, kc, kd, mc, md, ix)].off;
                            call.ntcd = ntcd;
                            // IT2 = 2: C large, D small
                            call.fc[0] = B.lrge_at(cc, kc, 2 * mc - 1);
                            call.fc[1] = B.lrge_at(cc, kc, 2 * mc);
                            call.fd[0] = B.lrge_at(cd, kd, 2 * md - 1) + B.nskp;
                            call.fd[1] = B.lrge_at(cd, kd, 2 * md) + B.nskp;
                            // exchange batches 5, 6, 7 and, off the diagonal
                            // IQL = IQR, 9, 10, 11 (BRTMTG :20043-20350)
                            call.iflg = (1 << 4) | (1 << 5) | (1 << 6);
                            if (iql != iqr) call.iflg |= (1 << 8) | (1 << 9) | (1 << 10);
                            // Breit phases carry an extra minus (:17530-17533)
                            call.pab[0] = (int8_t)(-sab * m1pow((-mqa + mqb) / 2));
                            call.pab[1] = (int8_t)(-sab * m1pow((mqa + mqb) / 2));
                            call.pcd[0] = (int8_t)(-scd * m1pow((-mqc + mqd) / 2));
                            call.pcd[1] = (int8_t)(-scd * m1pow((mqc + mqd) / 2));
                            P.cds.push_back(call);
                            P.quartets += (double)cl.nbas[0] * cl.nbas[1] * cl.maxcd;
                          }
                      }
                    }
                    ab.cd_end = (int)P.cds.size();
                    if (ab.cd_end == ab.cd_begin) continue;
                    // BIIFAST builds G^ix only where some E(CD) column of the
                    // class is non-negligible (:18505-18508, :18545-18548)
                    ab.gmask = (int)P.masks.size();
                    for (int t = 0; t < ntcd; ++t)
                      for (int ix = 0; ix < 3; ++ix) {
                        const int *g = &agg[(t * 3 + ix) * 4];
                        int all = g[0] + g[1] + g[2] + g[3];
                        for (int c = 0; c < 4; ++c)
                          P.masks.push_back((uint8_t)(all != 0 && (isym == 0 || g[c] != 0)));
                      }
                    P.max_ntg = std::max(P.max_ntg, ab.ntg);
                    P.abs.push_back(ab);
                    P.ab_pair.insert(P.ab_pair.end(), abpr, abpr + 3);
                  }
              }
            }
            cl.ab_end = (int)P.abs.size();
            if (cl.ab_end == cl.ab_begin) continue;
            P.classes.push_back(cl);
            P.nflag += ntuv(cl.lamfl);
            P.max_lfl = std::max(P.max_lfl, cl.lamfl);
            P.max_maxcd = std::max(P.max_maxcd, cl.maxcd);
          }
        }
  // Algorithmic FP64 work, counted as for the Coulomb build: R batch per
  // (ij, class); first contraction per (ij, AB combination) 3 x (8 flops of
  // the Gaunt term + 3 x 14 of the gauge term) per (tuv, tuv', kl); second
  // contraction 3 x 32 (3 x 8 real) per (tuv', kl) and call; digestion 128
  // (8 real) flops per active batch and element.  The first contraction's
  // count is BIIFAST's statements per (IX, ICD, IAB, N) (bertha-scf.f
  // :18700-18810): Gaunt term G += E R for four components, 4 x 2 = 8; per
  // JX of the gauge term T1 = R 0.5 / APH (2), T2 = R PQ (1), T3 = R RTP
  // (1), T = T1 - T2 + T3 (2) and G -= E T for four components (8): 14.
  for (const ClassDesc &cl : P.classes) {
    double maxab = (double)cl.nbas[0] * cl.nbas[1], maxcd = cl.maxcd;
    double L = cl.lamfl;
    double c4 = (L + 3) * (L + 2) * (L + 1) * L / 24.0;
    P.flops += maxab * maxcd * (120.0 + 2.0 * L + 6.0 * c4 + 3.0 * ntuv(cl.lamfl));
    for (int a = cl.ab_begin; a < cl.ab_end; ++a) {
      const BrAB &ab = P.abs[a];
      P.flops += maxab * 3.0 * (gaunt ? 8.0 : 50.0) * ab.ntab * ab.ntg * maxcd;
      for (int k = ab.cd_begin; k < ab.cd_end; ++k) {
        double c2 = 3.0 * (isym ? 8.0 : 32.0) * P.cds[k].ntcd;
        double dig = (isym ? 8.0 : 128.0) * __builtin_popcount(P.cds[k].iflg);
        P.flops += maxab * maxcd * (c2 + dig);
      }
    }
  }
  // work items as for the Coulomb plan
  std::vector<double> abcost(P.abs.size(), 0.0), rcost(P.classes.size(), 0.0);
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    rcost[c] = (double)ntuv(cl.lamfl) * cl.maxcd;
    for (int a = cl.ab_begin; a < cl.ab_end; ++a) {
      const BrAB &ab = P.abs[a];
      double w = 30.0 * ab.ntab * ab.ntg * cl.maxcd;
      for (int k = ab.cd_begin; k < ab.cd_end; ++k) w += 24.0 * P.cds[k].ntcd * cl.maxcd + 64.0 * cl.maxcd;
      abcost[a] = w;
    }
  }
  make_items(P.classes, abcost, rcost, P.items);
  if (isym != 0 && zflags) build_breit_tiles(P, gaunt, T, *zflags);
  return check_offsets(P);
}

}  // namespace bdf
