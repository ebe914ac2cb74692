// kernels.h -- argument blocks and launchers of kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"

namespace bdf {

struct EtabArgs {
  const EPair *pairs;
  int npairs;
  const int64_t *first;  // [npairs] first flat item of each pair
  int64_t nitems;
  const ETerm *terms;
  const Mono *monos;
  const double *exps, *normL, *normS;
  double4 *E;
};

struct EflagArgs {
  const EPair *pairs;
  int npairs;
  const int64_t *flag_off;
  const double4 *E;
  uint8_t *flags;
  double4 *Ecd;   // if set: masked, signed E(CD) copy (and its FP32 copy Ecd32)
  float4 *Ecd32;
  // if set (axial engines): each Ecd entry's one nonzero component, as the
  // sum of its four (exact where at most one is nonzero), FP64 and FP32
  double *Ecd1;
  float *Ecd1_32;
  uint8_t *zflags;  // if set: per (pair, tuv, component) 1 where any entry is nonzero
};

struct FockArgs {
  const ClassDesc *classes;
  const ABCombo *abs;
  const CDCall *cds;
  const uint8_t *masks;
  const WorkItem *items;
  int nitems;
  int *counter;
  const double *exps;
  const double4 *E;
  const double2 *dent;
  double2 *gdir, *gxch;
  int ndim;
  double *slab;
  int64_t slab_stride;  // doubles per workgroup
  int ld;               // row length of the slab (>= largest maxcd)
  int isym, intsym;
  double2 *rr_out;      // test hook: write RR[16][maxcd] instead of digesting
  // SCHWARZ masks (SCHWRZ): GDSC table, sensitivity, large-block size
  const double *gdsc;
  // the largest GDSC and |D| over the four sign slots of each block pair,
  // stored at the pair's (-|m|, -|m|) element (slot_tables_kernel)
  const double *gslot, *dslot;
  double sens;
  int schwarz, nskp;
  // Schwarz split: items with qitem * max|D| < split run in FP32 on E32
  int mixed;
  double split;
  const double *qitem;
  const unsigned long long *dmax_bits;
  const float4 *E32;
  // E(CD) with the column masks and ERIFAST's Z sign applied (FP64, FP32)
  const double4 *Ecd;
  const float4 *Ecd32;
  // exchange accumulators in dynamic LDS: 2 (12 nD + 16 nC) doubles for the
  // plan's largest class
  int xacc_n;
  // the plan's one_comp: one nonzero component per parity class of t in
  // every E table, G's in each class of t' (ABCombo::gcomp) and E(CD)'s
  // (CDCall::ccomp); the second contraction then reads E(CD) from the
  // one-value tables (FP64, FP32)
  int one_comp;
  const double *Ecd1;
  const float *Ecd1_32;
};

struct DiagArgs {
  const DiagJob *jobs;
  int njobs;
  const int64_t *first;  // [njobs] first flat thread of each job
  int64_t nitems;
  const double *exps;
  const double4 *E;
  double *gdsc, *qbnd;   // ndim x ndim column-major
  int ndim;
};

struct BreitArgs {
  const ClassDesc *classes;
  int nclasses;
  const BrAB *abs;
  const BrCD *cds;
  const uint8_t *masks;
  const WorkItem *items;
  int nitems;
  int *counter;
  const double *exps;
  const double4 *E;
  const double2 *dent;
  double2 *bdir, *bxch;
  int ndim, nskp;
  double *slab;
  int64_t slab_stride;
  int ld;
  int isym, intsym, gaunt;
  uint8_t *stale;   // per class: column flags of the last R batch (BIIFAST's IRC)
  double2 *rr_out;  // test hook: RR[16][maxcd] of the first call instead of digesting
  const double4 *Ecd;  // E(CD) with column masks and ERIFAST's Z sign applied
  const double4 *Eab;  // E with EPSAB applied (launch_eps_filter), the AB side
  int group;           // AB combinations contracted together: 1, 2 or 4
  int waves;           // waves per SIMD the kernel is built for: 2, 3 or 4 (group 1), 2
  int xacc_n;          // exchange-accumulator doubles (dynamic LDS) of the largest class
};

// matrix-core Breit kernel of the real path (breit_tc.inc): items are
// (class, ij, tiles [ab_lo, ab_hi)) of BreitPlan::titems
struct BreitTcArgs {
  const ClassDesc *classes;
  const BrAB *abs;
  const BrCD *cds;
  const uint8_t *masks;
  const BrTile *tiles;
  const int *tile_cd;
  const int *tile_pair;
  const BrComp *comp;   // per BrAB
  int max_npair;        // partial-sum slots the LDS layout reserves
  const WorkItem *items;
  int nitems;
  int *counter;
  const double *exps;
  const double4 *Eab;   // E with EPSAB applied (AB side)
  const double4 *Ecd;   // E(CD) with column masks and ERIFAST's Z sign
  const double2 *dent;
  double2 *bxch;
  int ndim;
  double *slab;         // per workgroup: R rows and Boys rows of the class
  int64_t slab_stride;
  int intsym, gaunt;
  const uint8_t *stale; // per class: column flags of the last R batch
  const int *sig;       // signature blocks (BreitPlan::tc_sig)
  const int *csig;      // per class: offset of its block
};

struct OneArgs {
  const OneJob *jobs;
  int njobs;
  const int64_t *first;
  int64_t nitems;
  const double *exps;
  const double4 *E;
  const double *xyz;    // [ncnt][3]
  const double *znuc;   // [ncnt]
  const double *xi;     // [ncnt], <= 0: point nucleus
  int ncnt;
  double cv;
  double2 *ovlp, *hnuc, *hkin;
  int ndim, nskp;
};

struct CompleteArgs {
  double2 *gdir, *gxch;
  const int *icn, *kqn, *mqn;
  int ndim, nskp, intsym;
};

// the allocation-free build's scratch: 256 MiB per device, in kernels.hip
constexpr size_t kStaticScratchDoubles = (size_t(256) << 20) / sizeof(double);
constexpr int kMaxDevices = 64;
// device address of that scratch on `device`, which must be current; looked
// up once per device (hipGetSymbolAddress) and kept
// one-centre Racah Coulomb build (racah.hip)
struct RacahArgs {
  const RcBlock *blocks;
  const RcSeg *segs;
  const RcCombo *combos;
  const RcEntry *entries;
  const RcOp *ops;
  const RcItem *items;
  const double *ang;  // angular factors of the entries
  int nitems;
  const double *exps, *normL, *normS;
  const double2 *dent;
  double2 *gdir, *gxch;  // Breit: gxch is BXCH, gdir unused
  int ndim;
  int isym;
  int exact;  // 1: every block's own beta integrals
  int breit;  // 0: COULOMB1, 1: BREIT1
};
// the racah kernel, then the lower triangles of the matrices it builds set
// to the conjugate upper ones
hipError_t launch_racah(const RacahArgs &a, hipStream_t s);

hipError_t static_scratch(int device, double **ptr);

hipError_t upload_constants();
hipError_t launch_etab(const EtabArgs &a, hipStream_t s);
hipError_t launch_eflag(const EflagArgs &a, hipStream_t s);
hipError_t launch_zero(double2 *x, double2 *y, int64_t n, int *counter,
                       unsigned long long *dmax, hipStream_t s);
// out(r, c) = max over the sign slots {r, partner[r]} x {c, partner[c]} of
// G (when G is given) or of |D|, for every (r, c)
hipError_t launch_slot_tables(const double *G, const double2 *D, const int *partner, int nd,
                              double *out, hipStream_t s);
hipError_t launch_eps_filter(const double4 *src, double4 *dst, int64_t n, hipStream_t s);
hipError_t launch_to_f32(const double4 *src, float4 *dst, int64_t n, hipStream_t s);
hipError_t launch_dmax(const double2 *d, int64_t n, unsigned long long *out, hipStream_t s);
// FP64 instance on s; with a.mixed also the FP32 instance on s32 (slab
// blocks nblocks..2*nblocks-1), forked from and joined back into s
// How launch_fock runs the Coulomb kernel for a plan whose exchange
// accumulators take xacc_n doubles of LDS, on the real (isym != 0) or complex
// path: waves per SIMD (3 or 2; workgroups per CU four times that) and the
// capacity multiple of its combination groups (1 or 2; the slab's G region
// holds cap * 4 * N(kMaxLab) rows)
struct FockLaunch {
  int waves, cap;
};
FockLaunch fock_launch_config(int xacc_n, int isym);
// the instantiated shapes: {3, 1}, {2, 2}, {2, 1}
bool fock_launch_valid(FockLaunch c);
hipError_t launch_fock(const FockArgs &a, FockLaunch c, int nblocks, hipStream_t s,
                       hipStream_t s32, hipEvent_t fork, hipEvent_t join);
hipError_t launch_complete(const CompleteArgs &a, hipStream_t s);
hipError_t launch_diag(const DiagArgs &a, hipStream_t s);
// AB combinations of one class that breit_kernel contracts together
// (BreitArgs::group: 1, 2 or 4; its scratch holds 12 N(L_CD) G rows for each)
constexpr int kMaxBreitGroup = 4;
hipError_t launch_breit(const BreitArgs &a, int nblocks, hipStream_t s);
// the matrix-core kernel for the same build (real path): the stale-flag pass
// of `flags`, then breit_tc_kernel with `lds` bytes of dynamic LDS per
// workgroup and call slots for `nu` calls per tile, then the vector-ALU
// kernel on the items of `rest` (classes beyond the LDS; its own work
// counter), then the SL completion
hipError_t launch_breit_tc(const BreitArgs &flags, const BreitTcArgs &a, int nblocks, size_t lds,
                           int nu, const BreitArgs &rest, int rest_blocks, hipStream_t s);
hipError_t launch_oneel(const OneArgs &a, hipStream_t s);

}  // namespace bdf
