// racah.cpp -- host plan of the one-centre Racah-algebra Coulomb build that
// replaces the reference's per-centre COULOMB1 calls (bertha-scf.f:21540,
// called from HFSCF at 5415-5421 when RACAH1 is on).
//
// COULOMB1 factorises every one-centre integral (AB|CD) of a centre into
// radial integrals per tensor order NU (RKCLM1 25664, over incomplete beta
// integrals BTACLM1 25380) times angular factors DKAB(NU) DKCD(NU) (ANGCLM1
// 25865), and digests the sum X = sum_NU DKAB DKCD RJ(NU) into GDIR/GXCH by the
// index permutations of CLMMT1Z 22809 (axial molecules and atoms, real part of
// D) or CLMMT1G 24443 (general molecules, complex D).  Everything that does not
// depend on the primitive pair is decided here: the LQN blocks and their
// tensor orders (TNSCLM1 25551), the KQN combinations and their angular
// factors, and for every (|m|, sign) combination the digestion terms with
// their matrix offsets, phases and upper-triangle guards.  The kernel
// (racah.hip) evaluates radial integrals per (ij, kl) and applies the terms.
//
// The reference evaluates a block's beta integrals only in the block's first
// KQN combination (K4AD = 1, 21842) and keeps them in COMMON /BETA/ by flat
// pair index.  Blocks whose first combination its triangle rule IKL <= IKR
// drops (for instance (pp|sd), la = lb = 1, lc = 0, ld = 2) therefore read
// the integrals the last writing block left at the same flat index.  The plan
// records, for every such block, which earlier block wrote each range it
// reads (RcSeg), so that the kernel reproduces the reference's values; the
// kernel can also evaluate every block's own integrals instead.
#include <cmath>
#include <cstdlib>

#include "plan.h"

namespace bdf {

namespace {

// n! for n <= 40 (3j symbols of l <= 3 shells need at most 2j + l + 1 <= 14)
struct Factorials {
  double f[41];
  Factorials() {
    f[0] = 1.0;
    for (int n = 1; n <= 40; ++n) f[n] = f[n - 1] * n;
  }
};
const Factorials kFact;

// Wigner 3j symbol (j1 m1, j2 m2, j3 m3) by Racah's single-sum formula; the
// arguments are doubled (half-integers allowed)
double three_j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2 || (j1 + j2 + j3) % 2) return 0.0;
  if ((j1 + m1) % 2 || (j2 + m2) % 2 || (j3 + m3) % 2) return 0.0;
  const double *F = kFact.f;
  auto h = [](int x) { return x / 2; };
  double tri = F[h(j1 + j2 - j3)] * F[h(j1 - j2 + j3)] * F[h(-j1 + j2 + j3)] / F[h(j1 + j2 + j3) + 1];
  double pre = std::sqrt(tri * F[h(j1 + m1)] * F[h(j1 - m1)] * F[h(j2 + m2)] * F[h(j2 - m2)] *
                         F[h(j3 + m3)] * F[h(j3 - m3)]);
  int kmin = std::max(0, std::max(h(j2 - j3 - m1), h(j1 - j3 + m2)));
  int kmax = std::min(h(j1 + j2 - j3), std::min(h(j1 - m1), h(j2 + m2)));
  double sum = 0.0;
  for (int k = kmin; k <= kmax; ++k) {
    double den = F[k] * F[h(j1 + j2 - j3) - k] * F[h(j1 - m1) - k] * F[h(j2 + m2) - k] *
                 F[h(j3 - j2 + m1) + k] * F[h(j3 - j1 - m2) + k];
    sum += (k % 2 ? -1.0 : 1.0) / den;
  }
  int ph = h(j1 - j2 - m3);
  return (ph % 2 ? -1.0 : 1.0) * pre * sum;
}

// DK (29166): coefficient of the order-nu tensor between spinors (ja, ma) and
// (jb, mb), doubled j and m: (-1)^((ma+1)/2) sqrt((ja+1)(jb+1))
// (ja nu jb; 1/2 0 -1/2) (ja nu jb; -ma/2 (ma-mb)/2 mb/2)
double dk(int ja, int ma, int jb, int mb, int nu) {
  int e = (ma + 1) / 2;
  double phs = (e % 2 == 0) ? 1.0 : -1.0;
  return phs * std::sqrt((double)((ja + 1) * (jb + 1))) * three_j(ja, 2 * nu, jb, 1, 0, -1) *
         three_j(ja, 2 * nu, jb, -ma, ma - mb, mb);
}

// BRCOEF1 (29092): coefficients of the eight (LS|LS)-type radial integrals
// in the orders nu-1 (column 0) and nu+1 (column 1) of an even-parity order
// nu (Grant and Pyper 1976, Table 2)
void breit_coefficients(const int kq[4], int nu, double co[8][2]) {
  const double kca = kq[1] - kq[0], kdb = kq[3] - kq[2], n = nu;
  if (nu >= 1) {
    const double b = n * (n + 1) / (2 * n * (2 * n - 1)), c = (n - 2) / (2 * n * (2 * n - 1));
    co[0][0] = (n + kca) * (b - c * kdb);
    co[1][0] = (n + kdb) * (b - c * kca);
    co[2][0] = (n - kca) * (b + c * kdb);
    co[3][0] = (n - kdb) * (b + c * kca);
    co[4][0] = -(n + kca) * (b + c * kdb);
    co[5][0] = -(n - kdb) * (b - c * kca);
    co[6][0] = -(n - kca) * (b - c * kdb);
    co[7][0] = -(n + kdb) * (b + c * kca);
  } else {
    for (int u = 0; u < 8; ++u) co[u][0] = 0.0;
  }
  const double b = n * (n + 1) / ((2 * n + 2) * (2 * n + 3)), c = (n + 3) / ((2 * n + 2) * (2 * n + 3));
  co[0][1] = -(n + 1 - kdb) * (b + c * kca);
  co[1][1] = -(n + 1 - kca) * (b + c * kdb);
  co[2][1] = -(n + 1 + kdb) * (b - c * kca);
  co[3][1] = -(n + 1 + kca) * (b - c * kdb);
  co[4][1] = (n + 1 + kdb) * (b + c * kca);
  co[5][1] = (n + 1 - kca) * (b - c * kdb);
  co[6][1] = (n + 1 - kdb) * (b - c * kca);
  co[7][1] = (n + 1 + kca) * (b + c * kdb);
}

// A FOCKIND block offset named as the reference names it: N/K (the sign slot
// or its Kramers partner), A..D, L/S (large or small component).
struct Label {
  int kram, letter, small;
};
constexpr Label lab(const char *s) {
  return Label{s[0] == 'K' ? 1 : 0, s[1] - 'A', s[2] == 'S' ? 1 : 0};
}

// one digestion line: M[dr, dc] += f X_xt D[sr, sc]
struct Term {
  Label dr, dc, sr, sc;
  int xt;  // 0 LLLL, 1 LLSS, 2 SSLL, 3 SSSS
};
// terms sharing an upper-triangle guard on two block offsets
struct Group {
  bool swap;    // (CD|AB) variable swap: only when IKL != IKR
  Label gr, gc; // guard: offset(gr) <= offset(gc) (< when strict)
  bool strict;
  int nt;
  Term t[4];
};
// one index permutation (8 sections: 4 GDIR, 4 GXCH)
struct Section {
  int mat;      // 0 GDIR, 1 GXCH
  int phase;    // 0: 1, 1: PAB, 2: PCD, 3: PAB PCD
  bool skip12;  // whole section skipped when KQN(A) == KQN(B)
  bool skip34;
  int ng;
  Group g[4];
};

constexpr Term T(const char *dr, const char *dc, int xt, const char *sr, const char *sc) {
  return Term{lab(dr), lab(dc), lab(sr), lab(sc), xt};
}

// CLMMT1Z 22809-23760 / CLMMT1G 24443-25377, closed-shell part.  Per-group
// skips that depend on the kappas (GOTO 512/513 in the direct sections) are
// applied in emit_entry.
const Section kSections[8] = {
    // GDIR (AB|CD)
    {0, 0, false, false, 2,
     {{false, lab("NAL"), lab("NBL"), false, 4,
       {T("NAL", "NBL", 0, "NCL", "NDL"), T("NAL", "NBL", 1, "NCS", "NDS"),
        T("NAS", "NBS", 2, "NCL", "NDL"), T("NAS", "NBS", 3, "NCS", "NDS")}},
      {true, lab("NCL"), lab("NDL"), false, 4,
       {T("NCL", "NDL", 0, "NAL", "NBL"), T("NCL", "NDL", 2, "NAS", "NBS"),
        T("NCS", "NDS", 1, "NAL", "NBL"), T("NCS", "NDS", 3, "NAS", "NBS")}}}},
    // GDIR (BA|CD) = PAB (AB|CD)
    {0, 1, false, false, 2,
     {{false, lab("KBL"), lab("KAL"), false, 4,
       {T("KBL", "KAL", 0, "NCL", "NDL"), T("KBL", "KAL", 1, "NCS", "NDS"),
        T("KBS", "KAS", 2, "NCL", "NDL"), T("KBS", "KAS", 3, "NCS", "NDS")}},
      {true, lab("NCL"), lab("NDL"), false, 4,
       {T("NCL", "NDL", 0, "KBL", "KAL"), T("NCL", "NDL", 2, "KBS", "KAS"),
        T("NCS", "NDS", 1, "KBL", "KAL"), T("NCS", "NDS", 3, "KBS", "KAS")}}}},
    // GDIR (AB|DC) = PCD (AB|CD)
    {0, 2, false, false, 2,
     {{false, lab("NAL"), lab("NBL"), false, 4,
       {T("NAL", "NBL", 0, "KDL", "KCL"), T("NAL", "NBL", 1, "KDS", "KCS"),
        T("NAS", "NBS", 2, "KDL", "KCL"), T("NAS", "NBS", 3, "KDS", "KCS")}},
      {true, lab("KDL"), lab("KCL"), false, 4,
       {T("KDL", "KCL", 0, "NAL", "NBL"), T("KDL", "KCL", 2, "NAS", "NBS"),
        T("KDS", "KCS", 1, "NAL", "NBL"), T("KDS", "KCS", 3, "NAS", "NBS")}}}},
    // GDIR (BA|DC) = PAB PCD (AB|CD)
    {0, 3, true, true, 2,
     {{false, lab("KBL"), lab("KAL"), false, 4,
       {T("KBL", "KAL", 0, "KDL", "KCL"), T("KBL", "KAL", 1, "KDS", "KCS"),
        T("KBS", "KAS", 2, "KDL", "KCL"), T("KBS", "KAS", 3, "KDS", "KCS")}},
      {true, lab("KDL"), lab("KCL"), false, 4,
       {T("KDL", "KCL", 0, "KBL", "KAL"), T("KDL", "KCL", 2, "KBS", "KAS"),
        T("KDS", "KCS", 1, "KBL", "KAL"), T("KDS", "KCS", 3, "KBS", "KAS")}}}},
    // GXCH (AD|CB)
    {1, 0, false, false, 4,
     {{false, lab("NAL"), lab("NDL"), false, 2,
       {T("NAL", "NDL", 0, "NCL", "NBL"), T("NAS", "NDS", 3, "NCS", "NBS")}},
      {false, lab("NAL"), lab("NDS"), false, 1, {T("NAL", "NDS", 1, "NCS", "NBL")}},
      {true, lab("NCL"), lab("NBL"), false, 2,
       {T("NCL", "NBL", 0, "NAL", "NDL"), T("NCS", "NBS", 3, "NAS", "NDS")}},
      {true, lab("NCL"), lab("NBS"), false, 1, {T("NCL", "NBS", 2, "NAS", "NDL")}}}},
    // GXCH (BD|CA) = PAB (AD|CB)
    {1, 1, true, false, 4,
     {{false, lab("KBL"), lab("NDL"), false, 2,
       {T("KBL", "NDL", 0, "NCL", "KAL"), T("KBS", "NDS", 3, "NCS", "KAS")}},
      {false, lab("KBL"), lab("NDS"), false, 1, {T("KBL", "NDS", 1, "NCS", "KAL")}},
      {true, lab("NCL"), lab("KAL"), false, 2,
       {T("NCL", "KAL", 0, "KBL", "NDL"), T("NCS", "KAS", 3, "KBS", "NDS")}},
      {true, lab("NCL"), lab("KAS"), true, 1, {T("NCL", "KAS", 2, "KBS", "NDL")}}}},
    // GXCH (AC|DB) = PCD (AD|CB)
    {1, 2, false, true, 4,
     {{false, lab("NAL"), lab("KCL"), false, 2,
       {T("NAL", "KCL", 0, "KDL", "NBL"), T("NAS", "KCS", 3, "KDS", "NBS")}},
      {false, lab("NAL"), lab("KCS"), false, 1, {T("NAL", "KCS", 1, "KDS", "NBL")}},
      {true, lab("KDL"), lab("NBL"), false, 2,
       {T("KDL", "NBL", 0, "NAL", "KCL"), T("KDS", "NBS", 3, "NAS", "KCS")}},
      {true, lab("KDL"), lab("NBS"), false, 1, {T("KDL", "NBS", 2, "NAS", "KCL")}}}},
    // GXCH (BC|DA) = PAB PCD (AD|CB)
    {1, 3, true, true, 4,
     {{false, lab("KBL"), lab("KCL"), false, 2,
       {T("KBL", "KCL", 0, "KDL", "KAL"), T("KBS", "KCS", 3, "KDS", "KAS")}},
      {false, lab("KBL"), lab("KCS"), false, 1, {T("KBL", "KCS", 1, "KDS", "KAL")}},
      {true, lab("KDL"), lab("KAL"), false, 2,
       {T("KDL", "KAL", 0, "KBL", "KCL"), T("KDS", "KAS", 3, "KBS", "KCS")}},
      {true, lab("KDL"), lab("KAS"), false, 1, {T("KDL", "KAS", 2, "KBS", "KCL")}}}},
};

// BRTMT1Z 26553-27460 / BRTMT1G 27462-28350, closed-shell part: BXCH from
// the four permutations of (AD|CB), with the same sign rules as COULOMB1's
// exchange sections (section_allows 4..7)
const Section kBreitSections[4] = {
    // BXCH (AD|CB)
    {1, 0, false, false, 4,
     {{false, lab("NAL"), lab("NDL"), false, 2,
       {T("NAL", "NDL", 1, "NCS", "NBS"), T("NAS", "NDS", 2, "NCL", "NBL")}},
      {false, lab("NAL"), lab("NDS"), false, 1, {T("NAL", "NDS", 0, "NCL", "NBS")}},
      {true, lab("NCL"), lab("NBL"), false, 2,
       {T("NCL", "NBL", 2, "NAS", "NDS"), T("NCS", "NBS", 1, "NAL", "NDL")}},
      {true, lab("NCL"), lab("NBS"), false, 1, {T("NCL", "NBS", 0, "NAL", "NDS")}}}},
    // BXCH (BD|CA) = PAB (AD|CB)
    {1, 1, true, false, 4,
     {{false, lab("KBL"), lab("NDL"), false, 2,
       {T("KBL", "NDL", 3, "NCS", "KAS"), T("KBS", "NDS", 0, "NCL", "KAL")}},
      {false, lab("KBL"), lab("NDS"), false, 1, {T("KBL", "NDS", 2, "NCL", "KAS")}},
      {true, lab("NCL"), lab("KAL"), false, 2,
       {T("NCL", "KAL", 0, "KBS", "NDS"), T("NCS", "KAS", 3, "KBL", "NDL")}},
      {true, lab("NCL"), lab("KAS"), false, 1, {T("NCL", "KAS", 2, "KBL", "NDS")}}}},
    // BXCH (AC|DB) = PCD (AD|CB)
    {1, 2, false, true, 4,
     {{false, lab("NAL"), lab("KCL"), false, 2,
       {T("NAL", "KCL", 0, "KDS", "NBS"), T("NAS", "KCS", 3, "KDL", "NBL")}},
      {false, lab("NAL"), lab("KCS"), false, 1, {T("NAL", "KCS", 1, "KDL", "NBS")}},
      {true, lab("KDL"), lab("NBL"), false, 2,
       {T("KDL", "NBL", 3, "NAS", "KCS"), T("KDS", "NBS", 0, "NAL", "KCL")}},
      {true, lab("KDL"), lab("NBS"), false, 1, {T("KDL", "NBS", 1, "NAL", "KCS")}}}},
    // BXCH (BC|DA) = PAB PCD (AD|CB)
    {1, 3, true, true, 4,
     {{false, lab("KBL"), lab("KCL"), false, 2,
       {T("KBL", "KCL", 2, "KDS", "KAS"), T("KBS", "KCS", 1, "KDL", "KAL")}},
      {false, lab("KBL"), lab("KCS"), false, 1, {T("KBL", "KCS", 3, "KDL", "KAS")}},
      {true, lab("KDL"), lab("KAL"), false, 2,
       {T("KDL", "KAL", 1, "KBS", "KCS"), T("KDS", "KAS", 2, "KBL", "KCL")}},
      {true, lab("KDL"), lab("KAS"), false, 1, {T("KDL", "KAS", 3, "KBL", "KCS")}}}},
};

// MQN-sign rule of each section: the sum rule both paths apply, and the
// axial path's (CLMMT1Z) extra conditions on the sign indices, as written
// there (sections 5 and 6 ask for unequal signs, section 3 pairs B with C)
bool section_allows(int sec, const int mj[4], const int mm[4], bool axial) {
  const int a = mj[0], b = mj[1], c = mj[2], d = mj[3];
  const int A = mm[0], B = mm[1], C = mm[2], D = mm[3];
  switch (sec) {
    case 0: return a - b == d - c && !(axial && (A != B || C != D));
    case 1: return b - a == d - c && !(axial && (B != A || C != D));
    case 2: return a - b == c - d && !(axial && (A != B || D != C));
    case 3: return b - a == c - d && !(axial && (B != C || D != A));
    case 4: return a - d == b - c && !(axial && (A != D || C != B));
    case 5: return -b - d == -a - c && !(axial && (B == D || C == A));
    case 6: return a + c == b + d && !(axial && (A == C || D == B));
    default: return -b + c == -a + d && !(axial && (B != C || D != A));
  }
}

struct EntryCtx {
  int base[4][2][2];  // [letter][kram][small] block offsets
  int kq[4];
  int ikl, ikr;
  double pab, pcd;
  int mj[4], mm[4];
};

void emit_entry(const EntryCtx &c, bool axial, bool breit, std::vector<RcOp> &ops) {
  const bool k12 = c.kq[0] == c.kq[1], k34 = c.kq[2] == c.kq[3];
  auto off = [&](Label l) { return c.base[l.letter][l.kram][l.small]; };
  const int nsec = breit ? 4 : 8;
  for (int s = 0; s < nsec; ++s) {
    const Section &S = breit ? kBreitSections[s] : kSections[s];
    if ((S.skip12 && k12) || (S.skip34 && k34)) continue;
    if (!section_allows(breit ? s + 4 : s, c.mj, c.mm, axial)) continue;
    const double f = S.phase == 0 ? 1.0 : S.phase == 1 ? c.pab : S.phase == 2 ? c.pcd : c.pab * c.pcd;
    for (int g = 0; g < S.ng; ++g) {
      const Group &G = S.g[g];
      if (G.swap && c.ikl == c.ikr) continue;
      // the direct (BA|CD) and (AB|DC) sections skip one half on equal kappas
      if (!breit && s == 1 && ((!G.swap && k12) || (G.swap && k34))) continue;
      if (!breit && s == 2 && ((!G.swap && k34) || (G.swap && k12))) continue;
      const int gr = off(G.gr), gc = off(G.gc);
      if (G.strict ? !(gr < gc) : !(gr <= gc)) continue;
      for (int t = 0; t < G.nt; ++t) {
        const Term &x = G.t[t];
        RcOp o{};
        o.dst_r = off(x.dr); o.dst_c = off(x.dc);
        o.src_r = off(x.sr); o.src_c = off(x.sc);
        o.sdr = (uint8_t)x.dr.letter; o.sdc = (uint8_t)x.dc.letter;
        o.ssr = (uint8_t)x.sr.letter; o.ssc = (uint8_t)x.sc.letter;
        o.mat = (uint8_t)S.mat;
        o.xt = (uint8_t)x.xt;
        o.uniform = (x.dr.letter <= 1 && x.dc.letter <= 1) ? 1 : 0;
        o.f = f;
        ops.push_back(o);
      }
    }
  }
}

}  // namespace

std::string build_racah_plan(const Basis &B, int isym, int breit, RacahPlan &P) {
  P = RacahPlan();
  const bool axial = isym != 0;
  const int nskp = B.nskp;
  for (int c = 0; c < B.ncnt; ++c) {
    const int lmx = B.lmax[c], nkap = 2 * lmx + 1;
    // IKMS (21620-21636): running index of (kappa, |m|, sign)
    int ikms[2 * kMaxL + 2][kMaxL + 2][2];
    int nkms = 0;
    for (int kn = 1; kn <= nkap; ++kn)
      for (int mj = 1; mj <= std::abs(B.kappa_of(c, kn)); ++mj)
        for (int s = 0; s < 2; ++s) ikms[kn][mj][s] = ++nkms;
    for (int la = 0; la <= lmx; ++la)
    for (int lb = 0; lb <= lmx; ++lb)
    for (int lc = 0; lc <= lmx; ++lc)
    for (int ld = 0; ld <= lmx; ++ld) {
      if (la > lb || lc > ld) continue;
      if (la * (la + 1) / 2 + lb > lc * (lc + 1) / 2 + ld) continue;
      // parity of the pairs and triangle rules (TNSCLM1 25551, TNSBRT1 28691)
      if ((la + lb) % 2 != (lc + ld) % 2) continue;
      int nui = std::max(std::abs(la - lb), std::abs(lc - ld));
      int nuf = std::min(la + lb, lc + ld);
      if (nui > nuf) continue;
      RcBlock bk{};
      bk.breit = breit;
      bk.nunum = 0;
      if (!breit) {
        // Coulomb: the orders of even parity
        for (int nu = nui; nu <= nuf; ++nu)
          if ((la + lb + nu) % 2 == 0) {
            if (bk.nunum == kRcMaxNu) return "racah plan: too many tensor orders";
            bk.nus[bk.nunum++] = nu;
          }
      } else {
        // Breit: an odd-parity order nu > 0 itself, an even one nu-1 and
        // nu+1 (nu-1 dropped at nu = 0), as TNSBRT1 lists them
        int lt = 0;
        bool any = false;
        int nus[2 * kRcMaxNu + 2] = {};
        for (int nu = nui; nu <= nuf; ++nu) {
          if ((la + lb + nu) % 2 == 1 && nu != 0) {
            any = true;
            nus[lt] = nu;
          } else if ((la + lb + nu) % 2 == 0) {
            any = true;
            nus[lt] = nu - 1;
            if (nu > 0) ++lt;
            nus[lt] = nu + 1;
          }
        }
        if (!any) continue;
        if (lt + 1 > kRcMaxNu) return "racah plan: too many tensor orders";
        bk.nunum = lt + 1;
        for (int t = 0; t < bk.nunum; ++t) bk.nus[t] = nus[t];
      }
      if (bk.nunum == 0) continue;
      const int lq[4] = {la, lb, lc, ld};
      for (int q = 0; q < 4; ++q) {
        bk.l[q] = lq[q];
        bk.nb[q] = B.nfn(c, lq[q]);
        bk.exp[q] = B.expo(c, lq[q]);
      }
      bk.nui = bk.nus[0];
      bk.nuf = bk.nus[bk.nunum - 1];
      // beta levels BTACLM1 / BTABRT1 write
      bk.nvals = (bk.nuf - bk.nui) / 2 + (breit ? 2 : 3);
      const int lsum = la + lb + lc + ld;
      for (int t = 0; t < 5; ++t) bk.cc[t] = 0.25 * std::tgamma(0.5 * (lsum + 2 * t + 1));
      bk.maxab = bk.nb[0] * bk.nb[1];
      bk.maxcd = bk.nb[2] * bk.nb[3];
      if (bk.maxab == 0 || bk.maxcd == 0) continue;
      // the first KQN combination (all four kappas negative) writes the
      // block's beta integrals unless IKL <= IKR drops it
      {
        int ka = 2 * la + 1, kb = 2 * lb + 1, kc = 2 * lc + 1, kd = 2 * ld + 1;
        bk.fresh = (ka * (ka + 1) / 2 + kb <= kc * (kc + 1) / 2 + kd) ? 1 : 0;
      }
      const int bi = (int)P.blocks.size();
      bk.seg_begin = bk.seg_end = (int)P.segs.size();
      if (!bk.fresh) {
        // last writers before this block, per level bound nv
        const int64_t size = (int64_t)bk.maxab * bk.maxcd;
        for (int nv = 1; nv <= bk.nvals; ++nv) {
          int64_t covered = 0;
          for (int w = bi - 1; w >= 0 && covered < size; --w) {
            const RcBlock &W = P.blocks[w];
            if (!W.fresh || W.nvals < nv) continue;
            const int64_t ws = (int64_t)W.maxab * W.maxcd;
            if (ws <= covered) continue;
            P.segs.push_back(RcSeg{nv, w, std::min(ws, size)});
            covered = ws;
          }
          if (covered < size) {
            P.segs.push_back(RcSeg{nv, -1, size});
            P.uncovered += size - covered;
          }
        }
        bk.seg_end = (int)P.segs.size();
      }
      P.blocks.push_back(bk);
      // KQN combinations (2001-2004)
      for (int na = 1; na >= (la == 0 ? 1 : 0); --na)
      for (int nb = 1; nb >= (lb == 0 ? 1 : 0); --nb)
      for (int nc = 1; nc >= (lc == 0 ? 1 : 0); --nc)
      for (int nd = 1; nd >= (ld == 0 ? 1 : 0); --nd) {
        const int ka = 2 * la + na, kb = 2 * lb + nb, kc = 2 * lc + nc, kd = 2 * ld + nd;
        if (ka > kb || kc > kd) continue;
        const int ikl = ka * (ka + 1) / 2 + kb, ikr = kc * (kc + 1) / 2 + kd;
        if (ikl > ikr) continue;
        const int kidx[4] = {ka, kb, kc, kd};
        int kq[4], jq[4];
        for (int q = 0; q < 4; ++q) {
          kq[q] = B.kappa_of(c, kidx[q]);
          jq[q] = 2 * std::abs(kq[q]) - 1;
        }
        // J triangle and parity select the tensor orders in use (ANGCLM1
        // 25865, ANGBRT1 28813)
        const int nuit = std::max(std::abs(jq[0] - jq[1]) / 2, std::abs(jq[2] - jq[3]) / 2);
        const int nuft = std::min((jq[0] + jq[1]) / 2, (jq[2] + jq[3]) / 2);
        if (nuit > nuft) continue;
        RcCombo cb{};
        cb.block = bi;
        cb.inu = 0;
        auto tensor_index = [&](int nu) {
          for (int t = 0; t < bk.nunum; ++t)
            if (bk.nus[t] == nu) return t;
          return -1;
        };
        // Breit: BCOF(t, u, a, b, c, d) over the combination's |m| and signs,
        // a..d = doubled |m| index + sign (0-based 2|m|-2 + sign)
        const int na4[4] = {2 * std::abs(kq[0]), 2 * std::abs(kq[1]), 2 * std::abs(kq[2]),
                            2 * std::abs(kq[3])};
        std::vector<double> bcof;
        auto bidx = [&](int t, int u, int a, int b2, int c2, int d) {
          return ((((size_t)t * 8 + u) * na4[0] + a) * na4[1] + b2) * (size_t)na4[2] * na4[3] +
                 (size_t)c2 * na4[3] + d;
        };
        // doubled signed m of sign-index position p (0: -|m|, 1: +|m|) of |m| index mi
        auto mjs = [](int mi, int sgn) { return sgn ? 2 * mi - 1 : -(2 * mi - 1); };
        int nulng = 0;
        if (!breit) {
          for (int nu = nuit; nu <= nuft; ++nu)
            if ((la + lb + nu) % 2 == 0 && (lc + ld + nu) % 2 == 0) {
              int t = tensor_index(nu);
              if (t >= 0) cb.inu |= 1 << t;
              ++nulng;
            }
        } else {
          bcof.assign((size_t)bk.nunum * 8 * na4[0] * na4[1] * na4[2] * na4[3], 0.0);
          for (int nu = nuit; nu <= nuft; ++nu) {
            const bool odd = (la + lb + nu) % 2 == 1;
            if (odd && nu == 0) continue;
            ++nulng;
            int tk[2] = {-1, -1};
            double co[8][2];
            if (odd) {
              tk[0] = tensor_index(nu);
              const double r = -(double)((kq[0] + kq[1]) * (kq[2] + kq[3])) / (double)(nu * (nu + 1));
              for (int u = 0; u < 8; ++u) co[u][0] = r;
            } else {
              breit_coefficients(kq, nu, co);
              if (nu > 0) tk[0] = tensor_index(nu - 1);
              tk[1] = tensor_index(nu + 1);
            }
            for (int h = 0; h < 2; ++h) {
              const int t = tk[h];
              if (t < 0) continue;
              cb.inu |= 1 << t;
              for (int a = 0; a < na4[0]; ++a)
                for (int b2 = 0; b2 < na4[1]; ++b2) {
                  const double dab = dk(jq[0], mjs(a / 2 + 1, a % 2), jq[1], mjs(b2 / 2 + 1, b2 % 2), nu);
                  if (dab == 0.0) continue;
                  for (int c2 = 0; c2 < na4[2]; ++c2)
                    for (int d = 0; d < na4[3]; ++d) {
                      const double dcd =
                          dk(jq[3], mjs(d / 2 + 1, d % 2), jq[2], mjs(c2 / 2 + 1, c2 % 2), nu);
                      for (int u = 0; u < 8; ++u) bcof[bidx(t, u, a, b2, c2, d)] += dab * dcd * co[u][h];
                    }
                }
            }
          }
        }
        if (nulng == 0) continue;
        for (int q = 0; q < 4; ++q) cb.tq[q] = (double)(lq[q] + kq[q] + 1);
        cb.entry_begin = (int)P.entries.size();
        EntryCtx ec{};
        for (int q = 0; q < 4; ++q) ec.kq[q] = kq[q];
        ec.ikl = ikl;
        ec.ikr = ikr;
        for (int ma = 1; ma <= std::abs(kq[0]); ++ma)
        for (int mb = 1; mb <= std::abs(kq[1]); ++mb)
        for (int mc = 1; mc <= std::abs(kq[2]); ++mc)
        for (int md = 1; md <= std::abs(kq[3]); ++md) {
          const int mq[4] = {2 * ma - 1, 2 * mb - 1, 2 * mc - 1, 2 * md - 1};
          // |m| pairing conserved on the axis (isym 1, 2); BREIT1 admits the
          // (AB)(CD) pairing for open shells only (26370)
          if (axial && !((!breit && mq[0] == mq[1] && mq[2] == mq[3]) ||
                         (mq[0] == mq[2] && mq[1] == mq[3]) || (mq[0] == mq[3] && mq[1] == mq[2])))
            continue;
          const int mabs[4] = {ma, mb, mc, md};
          for (int s = 0; s < 16; ++s) {
            // MMA..MMD = 1 (negative m) or 2 (positive m), A slowest
            int mm[4], mj[4];
            for (int q = 0; q < 4; ++q) {
              mm[q] = ((s >> (3 - q)) & 1) + 1;
              mj[q] = mm[q] == 1 ? -mq[q] : mq[q];
            }
            // minimal block permutations: COULOMB1 on equal |m| (21960),
            // BREIT1 on |m_A| < |m_B| (26405)
            if ((breit ? ma < mb : ma == mb) && ikms[ka][ma][mm[0] - 1] > ikms[kb][mb][mm[1] - 1])
              continue;
            if ((breit ? mc < md : mc == md) && ikms[kc][mc][mm[2] - 1] > ikms[kd][md][mm[3] - 1])
              continue;
            if (axial && mj[0] - mj[1] != mj[3] - mj[2]) continue;
            RcEntry en{};
            en.ang = (int)P.ang.size();
            const int per = breit ? 8 : 1;
            P.ang.resize(P.ang.size() + (size_t)per * bk.nunum, 0.0);
            bool any = false;
            for (int t = 0; t < bk.nunum; ++t) {
              if (!(cb.inu >> t & 1)) continue;
              double *a = &P.ang[en.ang + (size_t)per * t];
              if (!breit) {
                const int nu = bk.nus[t];
                // DKAB (AB pair) and DKCD (arguments reversed, 25990-26005)
                const double v = dk(jq[0], mj[0], jq[1], mj[1], nu) * dk(jq[3], mj[3], jq[2], mj[2], nu);
                if (std::fabs(v) <= 1.0e-10) continue;  // SENS of COULOMB1
                a[0] = v;
              } else {
                double sum = 0.0;
                const int ia = 2 * (ma - 1) + mm[0] - 1, ib = 2 * (mb - 1) + mm[1] - 1;
                const int ic = 2 * (mc - 1) + mm[2] - 1, id = 2 * (md - 1) + mm[3] - 1;
                for (int u = 0; u < 8; ++u) {
                  a[u] = bcof[bidx(t, u, ia, ib, ic, id)];
                  sum += std::fabs(a[u]);
                }
                if (sum <= 1.0e-10) {  // SENS of BREIT1, on the sum of the eight
                  for (int u = 0; u < 8; ++u) a[u] = 0.0;
                  continue;
                }
              }
              any = true;
            }
            if (!any) {
              P.ang.resize(en.ang);
              continue;
            }
            for (int q = 0; q < 4; ++q) {
              ec.mj[q] = mj[q];
              ec.mm[q] = mm[q];
              const int slot = 2 * mabs[q] - 2 + mm[q], kslot = 2 * mabs[q] + 1 - mm[q];
              const int n = B.lrge_at(c, kidx[q], slot), k = B.lrge_at(c, kidx[q], kslot);
              ec.base[q][0][0] = n;
              ec.base[q][0][1] = n + nskp;
              ec.base[q][1][0] = k;
              ec.base[q][1][1] = k + nskp;
            }
            auto sgn = [](int x) { return x >= 0 ? 1.0 : -1.0; };
            auto par = [](int x) { return (x / 2) % 2 == 0 ? 1.0 : -1.0; };
            // PAB, PCD; BREIT1's carry the opposite sign (26430)
            const double ph = breit ? -1.0 : 1.0;
            ec.pab = ph * sgn(kq[0] * kq[1]) * par(mj[0] - mj[1]);
            ec.pcd = ph * sgn(kq[2] * kq[3]) * par(mj[2] - mj[3]);
            en.op_begin = (int)P.ops.size();
            emit_entry(ec, axial, breit != 0, P.ops);
            en.op_end = (int)P.ops.size();
            if (en.op_end == en.op_begin) {
              P.ang.resize(en.ang);
              continue;
            }
            P.entries.push_back(en);
            P.terms += (double)bk.maxab * bk.maxcd * (en.op_end - en.op_begin);
          }
        }
        cb.entry_end = (int)P.entries.size();
        if (cb.entry_end == cb.entry_begin) continue;
        const int ci = (int)P.combos.size();
        P.combos.push_back(cb);
        for (int ij = 0; ij < bk.maxab; ++ij) P.items.push_back(RcItem{ci, ij});
      }
    }
  }
  if (P.ops.size() > (size_t)INT32_MAX || P.entries.size() > (size_t)INT32_MAX)
    return "racah plan: too many digestion terms";
  // every element a term touches lies inside the matrix
  for (const RcCombo &cb : P.combos) {
    const RcBlock &bk = P.blocks[cb.block];
    for (int e = cb.entry_begin; e < cb.entry_end; ++e)
      for (int p = P.entries[e].op_begin; p < P.entries[e].op_end; ++p) {
        const RcOp &o = P.ops[p];
        const int r[4] = {o.dst_r, o.dst_c, o.src_r, o.src_c};
        const int q[4] = {o.sdr, o.sdc, o.ssr, o.ssc};
        for (int t = 0; t < 4; ++t)
          if (r[t] < 0 || q[t] > 3 || r[t] + bk.nb[q[t]] > B.ndim)
            return "racah plan: digestion term outside the matrix";
      }
  }
  for (const RcBlock &bk : P.blocks)
    if (bk.nuf > kRcMaxNuf) return "racah plan: tensor order beyond the kernel's tables";
  return "";
}

}  // namespace bdf
