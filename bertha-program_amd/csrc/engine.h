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
};

// One ERIFAST call (CD combination and IT2) with its digestion flags.
struct CDCall {
  int64_t eoff;     // E(CD) double4 offset
  int ntcd;
  int fc[2], fd[2]; // Fock offsets of C, D for sign slots
  int iflg;         // bit b set: batch b+1 active
  int8_t pab[2], pcd[2];  // [0] same-slot phase (PAB1), [1] PAB2
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
};

}  // namespace bdf
