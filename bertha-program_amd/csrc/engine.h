// engine.h -- internal data layout of the MI355X two-electron engine.
//
// Host code (plan.cpp) turns a basis into flat tables; kernels.hip reads them.
// Everything here is plain data so the same structs are valid on both sides.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace bdf {

constexpr int kMaxL = 3;                  // largest shell l (f)
constexpr int kMaxPoly = kMaxL + 1;       // largest polynomial degree (SS of f)
constexpr int kMaxLab = 2 * kMaxL + 2;    // largest pair Hermite order (SS f-f)
constexpr int kMaxLfl = 4 * kMaxL + 4;    // largest R order (stage 3)
constexpr int kMaxPrim = 32;              // primitives per shell
constexpr int kWave = 64;

__host__ __device__ constexpr inline int ntuv(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }

// ---- E-coefficient generation --------------------------------------------
// A polynomial in (x,y,z) relative to one centre, stored as monomials.
struct Mono {
  double re, im;
  int8_t i, j, k, pad;
};

// One term of a spinor-pair overlap density: w * a^pa * b^pb *
// polyA(r-A) * polyB(r-B), where polyA/polyB index into the monomial table.
struct ETerm {
  double w;
  int16_t pa, pb;
  int32_t polyA, polyB;  // [begin, end) ranges packed as begin*65536 + count
};

// component kinds of the E tables
constexpr int kLL = 0, kSS = 1, kLSx = 2, kLSy = 3, kLSz = 4;

// One (A-block, B-block) pair table: E11 and E21 for one component type.
struct EPair {
  double ax, ay, az, bx, by, bz;
  int expA, expB;       // offsets into the exponent array
  int nA, nB;
  int lam;              // Hermite order of the stored expansion
  int small;            // component kind kLL, kSS, kLSx..kLSz
  int64_t off;          // double4 offset of [tuv][ij]
  int term11, nterm11;  // term ranges for E11 and E21
  int term21, nterm21;
  int la, lb;           // large-component l of A and B (normalisation)
  int phase;            // 0: none, 1: times i, 2: times -1 (i^2, sigma_y)
  int pad;
};

// ---- Fock build plan -----------------------------------------------------
// A class fixes the centres and l of all four labels; its R batch is shared
// by every kappa/|m|/component combination inside it (the reference caches
// it in RCTTFL, bertha-scf.f:12637-12645).
struct ClassDesc {
  double xyz[4][3];
  int exp[4];       // exponent offsets
  int nbas[4];
  int lamfl;        // R order (ERIFAST LAMABCDFL)
  int maxcd;
  int ab_begin, ab_end;
  int mcnt;
  int flagoff;      // Breit: offset of the class's gauge-term column flags
  // bit d (x = 1, y = 2, z = 4) set when all four centres have coordinate d
  // exactly 0: then P_d = Q_d = 0 exactly and R_tuv vanishes exactly unless
  // the d-th Hermite index is even, so contraction terms with odd index sums
  // along d are skipped without changing any sum
  int sym;
  int pad;
};

// One (kappa_A, kappa_B, |m_A|, |m_B|, IT1) combination of a class.
struct ABCombo {
  int64_t eoff;     // E(AB) table double4 offset for IT1's component type
  int ntab;         // N(L_AB)
  int ntg;          // N(L_CD) of the first CD call: G is built that far
  int gmask;        // byte offset of G mask [ntg][4]
  int fa[2], fb[2]; // Fock offsets of A, B for sign slots (0 = -|m|)
  int cd_begin, cd_end;
  int nb;           // primitives of B (ij -> i, j)
  int lab;          // L_AB: ntab = N(lab)
  // the one component of G in each parity class c of t' (the class's axis
  // mask), 4 bits at 4 c (4: none); used where the plan's one_comp is set
  uint32_t gcomp;
};

// One ERIFAST call (CD combination and IT2) with its digestion flags.
struct CDCall {
  int64_t eoff;     // E(CD) double4 offset
  int ntcd;
  int fc[2], fd[2]; // Fock offsets of C, D for sign slots
  int iflg;         // bit b set: batch b+1 active
  int8_t pab[2], pcd[2];  // [0] same-slot phase (PAB1), [1] PAB2
  // where the plan's one_comp is set: E(CD)'s one nonzero component in each
  // parity class of t' (4 bits at 4 c, 4: none)
  uint32_t ccomp;
};

// One block pair of the Schwarz tables: the self-pair integrals
// (A B | A B) of CLMDIAG (swap = 0, CD side = the same pair, kl = ij) or the
// bound integrals (A B | B A) (swap = 1, CD side = the (B,A) pair, kl = ji).
struct DiagJob {
  int64_t eab, ecd;  // double4 offsets of the AB- and CD-side E tables
  int ntab, ntcd;
  int nA, nB;
  int expA, expB;
  int row[2], col[2];  // matrix offsets of the A and B sign slots
  int swap;
  int realpart;        // keep real parts (ERI's path for ISYM 1, 2)
  int out;             // 0: GDSC, 1: bound table
  int pad;
};

struct WorkItem {
  int cls;
  int ij;
  int ab_lo, ab_hi;  // the class's AB combinations [ab_lo, ab_hi) this item covers
};

// ---- one-electron matrices (OVRLP / ONEEL) ------------------------------
// One (A, kappa_A, |m_A|; B, kappa_B, |m_B|) block pair with LRGE(A) >= LRGE(B).
struct OneJob {
  int64_t ell, ess, essba;  // E0LL(A,B), E0SS(A,B), E0SS(B,A) double4 offsets
  int nA, nB, expA, expB;
  int la, lb;
  int il[2], jl[2];         // large-block offsets of A and B sign slots
  double phs;               // PHS of the slot relations (ONEEL :7070)
  double ax, ay, az, bx, by, bz;
};

// ---- Breit plan (BRTFAST, bertha-scf.f:17060) ------------------------------
// One (kappa_A, kappa_B, |m_A|, |m_B|) combination: LS tables for sigma_x,y,z.
struct BrAB {
  int64_t eoff[3];
  int ntab, ntg;
  int gmask;        // bytes [ntg][3][4]: G^ix component built
  int fa[2], fb[2]; // Fock offsets of A (large) and B (small) sign slots
  int cd_begin, cd_end;
  int pad;
};

// One BIIFAST call with its digestion batches (BRTMTG/BRTMTZ)
struct BrCD {
  int64_t eoff[3];
  int ntcd;
  int fc[2], fd[2];
  int iflg;         // bit b set: exchange batch b+1 active
  int8_t pab[2], pcd[2];
};

// ---- matrix-core Breit kernel (real path, ISYM 1 and 2) -------------------
// On the real path every component c of an LS table E^j (j = x, y, z) is
// nonzero for bra Hermite indices t of at most one parity class (on the
// axes where all centres sit at 0), and so is every component of G^ix for a
// given parity class of t'.  The kernel's 16 x 16 x 4 FP64 matrix products
// therefore take as rows sixteen AB combinations of one class, each at the
// component that class selects (kap), and as output the one component of
// G^ix that lives in the class of t' (cout); the plan checks this structure
// from the tables' exact-zero flags and falls back to the vector-ALU kernel
// where it does not hold.
//
// A tile holds up to kTcRows combinations; its calls are listed once
// (tile_cd[(tile * kTcMaxU + u) * kTcRows + g]: the BrCD index of
// combination g's call u, -1 where g has no such call) so that one E(CD)
// load serves every combination having the call, and its (call,
// combination) pairs are numbered for the partial sums (tile_pair, u | g <<
// 8, from pair_begin).
constexpr int kTcRows = 16;
constexpr int kTcMaxU = 16;
constexpr int kTcGbLd = 17;  // row stride of a wave's G^ix buffer (doubles)
constexpr int kTcNone = 4;  // no component
constexpr int kTcWaves = 8;  // waves of a workgroup (two per SIMD)
constexpr int kTcThreads = 64 * kTcWaves;
// the 16 columns of the first contraction: kTcQ ket pairs kl of a sub-tile
// times kTcR ket Hermite indices t' of a group
constexpr int kTcQ = 2, kTcR = 8;
struct BrTile {
  int ab[kTcRows];  // BrAB index of row g, -1 = empty
  int nu;           // calls of the tile
  int fresh;        // 1: the class's first combination, which reads this pair's
                    // own R-column flags (BIIFAST's IRC); every other tile
                    // reads the class's stale flags
  // the class's first combination also sits in a stale tile as row
  // fresh_row (-1: not this tile); on the fresh tile fresh_row is that row
  // when there is one.  A pair whose own flags keep the same gauge terms as
  // the stale flags takes the combination there and skips the fresh tile;
  // any other pair takes the fresh tile and leaves that row out.
  int fresh_row;
  int pair_begin, npair;
  // the component of G^ix living in ket class c, common to the tile's rows
  // (kTcNone: G^ix vanishes there for every row)
  uint8_t cout[3][8];
};
// per BrAB: kap[j][c] (component of E^j living in bra class c) and
// cout[ix][c] (component of G^ix living in ket class c), kTcNone if none
struct BrComp {
  uint8_t kap[3][8];
  uint8_t cout[3][8];
};

// Hermite indices h < N(L) by parity class under an axis mask m (the class
// of h is (h_x & 1 | (h_y & 1) << 1 | (h_z & 1) << 2) & m): sz[c] members of
// class c, in the Hermite order of CARTIND.
inline void hermite_class_sizes(int m, int L, int sz[8]) {
  for (int c = 0; c < 8; ++c) sz[c] = 0;
  for (int n = 0; n <= L; ++n)
    for (int t = 0; t <= n; ++t)
      for (int u = 0; u <= n - t; ++u) {
        const int v = n - t - u;
        ++sz[((t & 1) | ((u & 1) << 1) | ((v & 1) << 2)) & m];
      }
}

// Dynamic LDS of one workgroup of the matrix-core Breit kernel for a class,
// byte offsets of its arrays: nab = N(L_AB) bra Hermite rows, wrows rows per
// gauge table, nr0 class-0 R rows, nt = N(L) R columns, nw tables (10
// Breit, 1 Gaunt), ngrp t' groups, ntg = N(L_CD), npair partial-sum slots.
struct TcLds {
  int et, w, wrec, rt, gbuf, rawb, dig, cdoff, kp, pidx, comp, cls0, tpl, tco, rowtab, chl, cf, gf,
      gmk, bytes;
};
// digestion constants of one (call, combination) pair of a tile
struct TcDig {
  int fa[2], fb[2], fc[2], fd[2];
  int iflg;
  int8_t pab[2], pcd[2];
};
__host__ __device__ inline TcLds tc_lds_layout(int nab, int wrows, int nr0, int nt, int nw,
                                               int ngrp, int ntg, int npair, int nchk) {
  TcLds o;
  int b = 0;
  o.et = b;     b += 3 * (nab + 1) * kTcRows * 8;  // E(AB) [3][nab + 1][16 g]
  o.w = b;      b += nw * wrows * kTcQ * 8;        // tables [nw][wrows][Q]
  o.wrec = b;   b += nw * wrows * 8;               // their recipes (one per row)
  o.rt = b;     b += nr0 * kTcQ * 8;               // class-0 R rows [nr0][Q]
  o.gbuf = b;   b += kTcWaves * 16 * kTcGbLd * 8;  // G^ix per wave [Q][R][16 g], padded rows
  o.rawb = b;   b += npair * kTcQ * 4 * 8;         // partial sums [pair][Q][4]
  o.dig = b;    b += npair * (int)sizeof(TcDig);   // digestion constants [pair]
  o.cdoff = b;  b += kTcMaxU * 3 * 8;              // E(CD) offsets [u][3]
  o.kp = b;     b += 5 * kTcQ * 8;                 // 1/(2 alpha) [Q], PQ [3][Q], prefactor [Q]
  o.pidx = b;   b += kTcMaxU * kTcRows * 2;        // pair of (u, g), -1: none
  o.comp = b;   b += kTcRows * 48;                 // BrComp of each row
  o.cls0 = b;   b += (nr0 * 2 + 7) & ~7;           // class-0 Hermite indices [nr0]
  o.tpl = b;    b += (ntg * 2 + 7) & ~7;           // ket t' in class order [ntg]
  o.tco = b;    b += 24;                           // the tile's cout [3][8]
  o.rowtab = b; b += (ngrp * kTcR * (nab + 1) * 2 + 7) & ~7;  // rows [grp][R][nab + 1]
  o.chl = b;    b += ((25 + nchk) * 4 + 7) & ~7;   // chunk offsets [ix * 8 + class], end; chunks
  o.cf = b;     b += (nt + 7) & ~7;                // this pair's R-column flags
  o.gf = b;     b += (nt + 7) & ~7;                // the class's stale flags
  o.gmk = b;    b += (ntg * 3 * 4 * 2 + 7) & ~7;   // G masks [ntg][3][4 c] (16 bits: g)
  o.bytes = b;
  return o;
}
constexpr int kTcLdsMax = 160 * 1024 - 256;

struct BreitPlan {
  std::vector<ClassDesc> classes;
  std::vector<BrAB> abs;
  std::vector<BrCD> cds;
  std::vector<uint8_t> masks;
  std::vector<WorkItem> items;
  int max_lfl = 0, max_maxcd = 0, max_ntg = 0;
  int nflag = 0;          // gauge-term flag bytes over all classes
  double quartets = 0.0;
  double flops = 0.0;
  // matrix-core kernel (real path): tiles, their calls and pairs, the
  // components of each combination, and work items (class, ij, tiles
  // [ab_lo, ab_hi)) over the classes whose LDS fits a compute unit; the
  // vector-ALU kernel takes rest_items, the items of the other classes.
  // tc_sig holds one block per (axis mask, L_AB, L_CD) signature (layout at
  // tc_signature), tc_csig each class's block offset.  tc_ok is 0 where the
  // one-component structure does not hold (the vector-ALU kernel then takes
  // the whole build).
  std::vector<int> ab_pair;  // 3 LS table pairs per BrAB
  std::vector<BrTile> tiles;
  std::vector<int> tile_cd;
  std::vector<int> tile_pair;
  std::vector<BrComp> comp;
  std::vector<WorkItem> titems, rest_items;
  std::vector<int> tc_sig;
  std::vector<int> tc_csig;
  size_t tc_lds = 0;
  int max_nu = 0, max_npair = 0, tc_ok = 0;
};

// ---- one-centre Racah plans (COULOMB1, bertha-scf.f:21540; BREIT1, 26014) --
// Tensor orders of one LQN block: for the Coulomb interaction one parity
// class of [max(|la-lb|,|lc-ld|), min(la+lb,lc+ld)], for the Breit
// interaction the orders of the other class one step beyond it.
constexpr int kRcMaxNu = kMaxL + 2;
constexpr int kRcMaxNuf = 2 * kMaxL + 1;  // largest tensor order

// One LQN block (la <= lb, lc <= ld, la(la+1)/2+lb <= lc(lc+1)/2+ld) of one
// centre, with the tensor orders TNSCLM1 allows (25551).
struct RcBlock {
  int l[4], nb[4], exp[4];  // l, primitives and exponent offsets of A, B, C, D
  int nunum, nui, nuf, nvals;
  int nus[kRcMaxNu];
  double cc[5];             // C1, C3, C5, C7, C9 = Gamma((L + 2t + 1)/2) / 4
  int maxab, maxcd;
  int seg_begin, seg_end;   // beta sources of a block that writes none (RcSeg)
  int fresh;                // 1: its first KQN combination writes its betas
  int breit;                // 1: BREIT1's tensor orders and beta integrals
};

// Where a block that writes no beta integrals reads them (the reference's
// COMMON /BETA/ keeps the last writer's values): reads at levels
// max(NX, NY) == nv and flat pair index (IJ-1)*MAXCD+M below fend come from
// block src's integrals at that flat index; src < 0: never written (zero).
struct RcSeg {
  int nv, src;
  int64_t fend;
};

// One KQN combination (K4AD) of a block that survives ANGCLM1 (25865).
struct RcCombo {
  int block;
  int inu;                  // bit t: tensor order nus[t] in use
  double tq[4];             // l + kappa + 1 of A, B, C, D (RKCLM1's T factors)
  int entry_begin, entry_end;
};

// One (|m|, sign) combination: the offset of its angular factors (0 where
// screened) and the digestion terms it feeds.  Coulomb: DKAB DKCD per tensor
// order, nunum values; Breit: the eight BCOF factors per order, 8 nunum.
struct RcEntry {
  int ang;
  int op_begin, op_end;
};

// One digestion term: M[dst] += f * X_xt(ij, kl) * D[src] (the real part of
// D on the axial path), rows and columns = base + one of (i, j, k, l).
struct RcOp {
  int dst_r, dst_c, src_r, src_c;
  uint8_t sdr, sdc, ssr, ssc;  // index of each: 0 i, 1 j, 2 k, 3 l
  uint8_t mat;                 // 0 GDIR, 1 GXCH (Breit: BXCH)
  uint8_t xt;                  // Coulomb 0 LLLL, 1 LLSS, 2 SSLL, 3 SSSS;
                               // Breit 0 LSLS, 1 LSSL, 2 SLLS, 3 SLSL
  uint8_t uniform;             // dst indexed by i and j only
  uint8_t pad;
  double f;
};

struct RcItem {
  int combo, ij;
};

struct RacahPlan {
  std::vector<RcBlock> blocks;
  std::vector<RcSeg> segs;
  std::vector<RcCombo> combos;
  std::vector<RcEntry> entries;
  std::vector<double> ang;
  std::vector<RcOp> ops;
  std::vector<RcItem> items;
  int64_t uncovered = 0;    // beta reads no earlier block wrote
  double terms = 0.0;       // digestion terms (ij, kl, op) of one build
};

// everything the host plan produces
struct Plan {
  int ndim = 0, nskp = 0;
  int isym = 0, intsym = 1, racah1 = 0, ilev = 3;
  std::vector<ClassDesc> classes;
  std::vector<ABCombo> abs;
  std::vector<CDCall> cds;
  std::vector<uint8_t> masks;
  std::vector<WorkItem> items;
  int max_lfl = 0, max_maxcd = 0, max_ntg = 0;
  double quartets = 0.0;  // primitive quartet count (ij,kl) x calls
  double flops = 0.0;     // algorithmic FP64 work of one build (see plan.cpp)
  // every parity class of every E(AB) and E(CD) table has at most one
  // nonzero component: the second contraction reads one value of each and
  // sums one product per t' class by class
  int one_comp = 0;
};

}  // namespace bdf
