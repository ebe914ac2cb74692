/*
 * bertha_amd.h -- C ABI of the MI355X two-electron engine.
 *
 * The reference has no plugin or FFI surface for this path: the SCF driver
 * HFSCF calls two parameterless Fortran subroutines that talk through COMMON
 * blocks (bertha-scf.f:5401 CALL CLMFAST, :5441 CALL BRTFAST, set up once by
 * :5287 CALL EQSAVE).  Each entry point below replaces one of those calls with
 * explicit arguments; include/bertha_amd_mod.f90 is the ISO_C_BINDING module a
 * Fortran driver uses to reach them.
 *
 * Conventions
 *   - Complex matrices are double _Complex (re,im interleaved), column-major,
 *     leading dimension ld >= ndim, in the reference's FOCKIND basis order
 *     (bertha-scf.f:2582-2648): large-component block rows 1..nskp, small
 *     component rows nskp+1..ndim.
 *   - Every call returns 0 on success or a negative BDF_E* code; it never
 *     exits the process (the reference STOPs, bertha-scf.f:12348-12352).
 *     bdf_last_error() describes the last failure on the calling thread.
 *   - The engine owns its device tables; it never keeps a host pointer after
 *     a call returns.  One engine is driven by one host thread at a time.
 */
#ifndef BERTHA_AMD_H
#define BERTHA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BDF_OK 0
#define BDF_EINVAL -1   /* bad argument */
#define BDF_ENOMEM -2   /* device or host allocation failed */
#define BDF_EDEVICE -3  /* HIP runtime / kernel failure */
#define BDF_ELIMIT -4   /* outside the engine's compiled limits */
#define BDF_EUNSUPP -5  /* option not implemented by this engine */

/* molecular shape -> integral symmetry class, as CLMFAST's SHAPE test
   (bertha-scf.f:11573-11579): ATOMIC=2, DIATOM/LINEAR=1, otherwise 0 */
#define BDF_SHAPE_NOSYMM 0
#define BDF_SHAPE_LINEAR 1
#define BDF_SHAPE_ATOMIC 2

/* Hamiltonians, as CARDIN's HMLT (bertha-scf.f:218-225) */
#define BDF_HMLT_DHFR 1 /* Dirac-Coulomb */
#define BDF_HMLT_DHFB 2 /* Dirac-Coulomb-Breit */

/* precision modes */
#define BDF_PREC_FP64 0  /* every quartet in FP64 */
#define BDF_PREC_MIXED 1 /* Schwarz-split FP32/FP64 (see bdf_options) */

/* bdf_options.schwarz */
#define BDF_SCHWARZ_OFF 0        /* no screening (the reference's default) */
#define BDF_SCHWARZ_REFERENCE 1  /* SCHWARZ masks from CLMDIAG's GDSC */
#define BDF_SCHWARZ_RIGOROUS 2   /* same masks, every group bounded by Q(a,b) Q(c,d) */

/* Basis and geometry, as CARDIN fills COMMON /BSET/ (bertha-scf.f:180-182):
   centre c has angular momenta 0..lmax[c]; nfnc[c*lstride + l] exponents for
   shell l, stored at exps[(c*lstride + l)*mbs + i]. */
typedef struct {
  int ncnt;
  const double *xyz;   /* [ncnt][3] bohr */
  const int *lmax;     /* [ncnt] */
  const int *nfnc;     /* [ncnt][lstride] */
  const double *exps;  /* [ncnt][lstride][mbs] */
  int lstride;
  int mbs;
  int shape;           /* BDF_SHAPE_* */
  int hmlt;            /* BDF_HMLT_* */
} bdf_system;

/* Toggles the reference hard-codes in HFSCF (bertha-scf.f:5215-5257) plus
   the engine's own options. */
typedef struct {
  int intsym;          /* INTSYM: permutational symmetry of RR batches (1) */
  int racah1;          /* RACAH1: bdf_coulomb/bdf_breit skip one-centre
                          quartets, which bdf_coulomb1 then adds as
                          COULOMB1 does (reference default 1); 0: this engine
                          computes them by MD inside the same build */
  int precision;       /* BDF_PREC_* */
  double split_threshold; /* MIXED: a work item (class, primitive pair ij) whose
                             largest quartet contribution bound
                             Q(ij) max Q(kl) max|D| is below it runs in FP32
                             (<= 0: 0.1); Q = sqrt|(ab|ba)| */
  int device;          /* HIP device ordinal */
  int schwarz;         /* SCHWRZ (reference default 0, bertha-scf.f:5257):
                          1 = SCHWARZ masks from CLMDIAG's GDSC, as the
                          reference; 2 = the same decision tree with every
                          group bounded by Q(a,b) Q(c,d), Q = sqrt|(ab|ba)|,
                          so no dropped term exceeds sens (GDSC is no bound
                          for complex spinor quartets) */
  double sens;         /* SENS of the masks; <= 0 selects the reference's 1e-12
                          (bertha-scf.f:11567) */
  int gaunt;           /* GAUNT2: Gaunt interaction instead of Breit in
                          bdf_breit (reference default 0, bertha-scf.f:5246) */
  int workgroups;      /* persistent workgroups of the quartet kernels (0: 12
                          per compute unit for Coulomb builds whose LDS allows
                          three waves per SIMD, else 8); fewer shrink the
                          per-workgroup scratch when several engines share
                          the device */
} bdf_options;

typedef struct bdf_engine bdf_engine;

/* Replaces EQSAVE (bertha-scf.f:32984) plus the one-time label set-up FOCKIND
   and CARTIND (bertha-scf.f:108-111): builds the E0LL/E0SS coefficient tables
   and the quartet plan on the device.  ndim/nskp receive NDIM and NSKP. */
int bdf_create(const bdf_system *sys, const bdf_options *opt, bdf_engine **out,
               int *ndim, int *nskp);

/* Replaces CALL CLMFAST (bertha-scf.f:5401; routine at :11489): closed-shell
   Dirac-Coulomb direct (GDIR) and exchange (GXCH) matrices for the total
   density DENT at integral-inclusion stage ilev (1..3, COMMON /LSHF/).
   Outputs are overwritten (zeroed then filled, :11592-11599) and
   Hermitian-completed as at :12166-12205.  Host arrays, ld >= ndim.
   deno/qdir/qxch are the open-shell counterparts and must be NULL: only
   closed shells are supported (BDF_EUNSUPP otherwise). */
int bdf_coulomb(bdf_engine *e, int ilev, const double _Complex *dent,
                const double _Complex *deno, int ld, double _Complex *gdir,
                double _Complex *gxch, double _Complex *qdir,
                double _Complex *qxch);

/* Same computation on device-resident arrays (ld = ndim), enqueued on the
   engine's stream and not synchronised: for callers that keep D, GDIR and
   GXCH in HBM (the SCF loop and the serving path).  No counterpart in the
   reference, whose matrices live in host COMMON /DENS/ and /MTRX/. */
int bdf_coulomb_device(bdf_engine *e, int ilev, const void *d_dent,
                       void *d_gdir, void *d_gxch);

/* bdf_coulomb_device without allocating: no hipMalloc, hipFree or other
   allocating call.  No counterpart in the reference; it is for callers that
   cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid it).  The quartet kernel's
   per-workgroup scratch comes from a static 256 MiB array per GPU instead
   of the engine's own, which bounds its workgroups (fewer than
   bdf_coulomb_device uses for systems with long primitive-pair lists, so it
   is slower there); a plan whose single workgroup needs more than 256 MiB
   (twice its rows in a mixed-precision engine) is refused with BDF_ELIMIT.
   The stage's plan must be built beforehand (bdf_plan_info or a
   bdf_coulomb_device call of the same stage).  Every call on one GPU
   shares that GPU's array: calls that overlap on one GPU (engines on
   different host threads or streams) race.  Enqueued on the engine's
   stream, not synchronised, never captured as a graph. */
int bdf_coulomb_device_static(bdf_engine *e, int ilev, const void *d_dent,
                              void *d_gdir, void *d_gxch);

/* Replaces the per-centre one-centre Coulomb calls of HFSCF with RACAH1 on,
   DO IZ=1,NCNT: CALL COULOMB1(IZ) (bertha-scf.f:5415-5421; routine at
   :21540): adds every centre's one-centre Coulomb contributions, evaluated
   by Racah algebra and beta integrals (TNSCLM1, ANGCLM1, BTACLM1, RKCLM1,
   CLMMT1Z/CLMMT1G), to GDIR and GXCH, which hold the CLMFAST result of an
   engine created with racah1 = 1 (one-centre quartets skipped), and then
   sets every lower-triangle element to the conjugate of its upper partner,
   as COULOMB1 does on return (:22074-22081).  All component types enter at
   every stage, as in the reference.  exact_beta = 0 reproduces the
   reference exactly, including its beta integrals for LQN blocks whose
   first KQN combination the triangle rule drops (those read the integrals
   the previous block left in COMMON /BETA/); exact_beta = 1 evaluates
   every block's own beta integrals instead (the exact one-centre
   integrals).  Host arrays, ld >= ndim, gdir/gxch updated in place;
   deno/qdir/qxch must be NULL (closed shells). */
int bdf_coulomb1(bdf_engine *e, int exact_beta, const double _Complex *dent,
                 const double _Complex *deno, int ld, double _Complex *gdir,
                 double _Complex *gxch, double _Complex *qdir,
                 double _Complex *qxch);

/* Same on device-resident arrays (ld = ndim), enqueued on the engine's
   stream and not synchronised. */
int bdf_coulomb1_device(bdf_engine *e, int exact_beta, const void *d_dent,
                        void *d_gdir, void *d_gxch);

/* Replaces HFSCF's per-centre one-centre Breit calls with RACAH1 on,
   DO IZ=1,NCNT: CALL BREIT1(IZ) (bertha-scf.f:5447-5453; routine at
   :26014): adds every centre's one-centre Breit contributions (TNSBRT1,
   ANGBRT1 with BRCOEF1, BTABRT1, RKBRT1, BRTMT1Z/BRTMT1G) to BXCH, which
   holds the BRTFAST result of an engine created with racah1 = 1, and sets
   its lower triangle to the conjugate upper one as BREIT1 does on return.
   BDIR receives nothing for closed shells and is not touched.  exact_beta
   as for bdf_coulomb1.  Host arrays, ld >= ndim; deno/wdir/wxch must be
   NULL. */
int bdf_breit1(bdf_engine *e, int exact_beta, const double _Complex *dent,
               const double _Complex *deno, int ld, double _Complex *bdir,
               double _Complex *bxch, double _Complex *wdir, double _Complex *wxch);

/* Same on device-resident arrays (ld = ndim; d_bdir may be NULL), enqueued
   on the engine's stream and not synchronised. */
int bdf_breit1_device(bdf_engine *e, int exact_beta, const void *d_dent,
                      void *d_bdir, void *d_bxch);

/* Size of the one-centre Racah build of the Coulomb (breit = 0) or Breit
   (breit = 1) interaction: digestion terms (ij, kl, term) of one build,
   work items, and beta-integral reads that no earlier block of the build
   writes (the reference reads what an earlier build or a fresh process
   left there; here they are zero).  Builds the plan on first use.  No
   counterpart in the reference. */
int bdf_racah_info(bdf_engine *e, int breit, double *terms, long *items, long *uncovered);

/* Test hook mirroring ERIFAST's output (bertha-scf.f:12269-13132) for one
   block quartet: centres icnt[4] (1-based), kappa indices kidx[4] (1-based
   position in KAPA, as KA/KB/KC/KD in CLMFAST :11771-11856), |mj| indices
   midx[4] (1..|kappa|), primitive pair (ibas,jbas) (1-based), component pair
   itn = {1|4, 1|4} (LL=1, SS=4) and stage ilev.  rr receives RR(N,1..16) for
   all maxcd = nC*nD pairs, column-major [16][maxcd]. */
int bdf_eri_batch(bdf_engine *e, const int icnt[4], const int kidx[4],
                  const int midx[4], int ibas, int jbas, const int itn[2],
                  int ilev, double _Complex *rr, int maxcd);

/* Replaces CALL BRTFAST (bertha-scf.f:5441; routine at :17060): closed-shell
   Breit (or, with opt.gaunt, Gaunt) matrices for the total density DENT at
   stage ilev (multi-centre quartets only at ilev 3, :17254).  BDIR stays
   zero for closed shells; BXCH holds the LS blocks and their conjugate
   transpose SL (:17646-17664).  Needs an engine created with
   BDF_HMLT_DHFB and intsym = 1 (BDF_EUNSUPP otherwise, as for open shells:
   deno/wdir/wxch must be NULL). */
int bdf_breit(bdf_engine *e, int ilev, const double _Complex *dent,
              const double _Complex *deno, int ld, double _Complex *bdir,
              double _Complex *bxch, double _Complex *wdir, double _Complex *wxch);

/* Same on device-resident arrays (ld = ndim), enqueued on the engine's
   stream and not synchronised. */
int bdf_breit_device(bdf_engine *e, int ilev, const void *d_dent, void *d_bdir,
                     void *d_bxch);

/* Test hook mirroring BIIFAST's output (bertha-scf.f:18337-19169) for one
   (LS|LS) block quartet at a fresh class (the gauge term sees this pair's
   own R-column flags): arguments as bdf_eri_batch without itn/ilev;
   rr[16][maxcd]. */
int bdf_bii_batch(bdf_engine *e, const int icnt[4], const int kidx[4],
                  const int midx[4], int ibas, int jbas, double _Complex *rr,
                  int maxcd);

/* Replaces CALL OVRLP / CALL ONEEL (bertha-scf.f:5317-5318; routines at
   :6683 and :6996): overlap OVLP, nuclear attraction HNUC (LL, and SS
   including -2 m c^2 S_SS) and kinetic coupling HKIN (LS/SL blocks).
   znuc[ncnt] nuclear charges; xi[ncnt] Gaussian-nucleus exponents
   (xi[c] <= 0 or xi == NULL: point nucleus; the reference's Gaussian model
   uses xi = 3/(2 R_rms^2), :1199-1204); cv the speed of light
   (reference 137.035999084, :974).  Host arrays, ld >= ndim. */
int bdf_one_electron(bdf_engine *e, const double *znuc, const double *xi, double cv,
                     double _Complex *ovlp, double _Complex *hnuc, double _Complex *hkin,
                     int ld);

/* Same into device arrays (ld = ndim); synchronises the engine's stream
   before returning. */
int bdf_one_electron_device(bdf_engine *e, const double *znuc, const double *xi, double cv,
                            void *d_ovlp, void *d_hnuc, void *d_hkin);

/* Size of the stage-ilev Breit build (as bdf_plan_info). */
int bdf_breit_plan_info(bdf_engine *e, int ilev, double *flops, double *quartets,
                        long *items);

/* Copy of one E0LL (itt=1) or E0SS (itt=4) block as the reference stores it
   (bertha-scf.f:33485-33494): out[comp][tuv][ij], comp = Re E11, Im E11,
   Re E21, Im E21.  itt=2 returns the EILS block of a DHFB engine
   (bertha-scf.f:33757-33768): 12 components, sigma_x, y, z times the four
   above.  Returns the element count per component in *n. */
int bdf_etable(bdf_engine *e, int itt, int ica, int icb, int ka, int kb,
               int ma, int mb, double *out, long cap, long *n);

/* Schwarz tables, ndim x ndim real, column-major, leading dimension ld
   (either pointer may be NULL):
   gdsc -- CLMDIAG's GDSC (bertha-scf.f:38126-38445): sqrt|(a b|a b)| of the
           self-pair integrals, row block by A's component type, column block
           by the second pair's, sign slots (--|--)=(++|++), (-+|-+)=(+-|+-);
           real parts only on the ISYM 1/2 path as ERI computes them.
   qbnd -- sqrt|(a b|b a)|, the Cauchy-Schwarz factor of the Coulomb metric
           (|(ab|cd)| <= qbnd(a,b) qbnd(c,d)); the engine's FP32/FP64 split
           uses it.  No counterpart in the reference. */
int bdf_schwarz_tables(bdf_engine *e, double *gdsc, double *qbnd, int ld);

/* Device memory held by the engine (tables, plans, scratch), in bytes.
   No counterpart in the reference. */
int bdf_memory_info(bdf_engine *e, size_t *device_bytes);

/* Engine's HIP stream (hipStream_t) for callers that order their own work
   after bdf_coulomb_device. */
void *bdf_stream(bdf_engine *e);

/* Per-build counters: number of primitive quartets evaluated and the
   duration of the last build's quartet kernel in milliseconds (HIP events on
   the engine's stream).  No counterpart in the reference, whose timers are
   the SYSTEM_CLOCK labour tables (bertha-scf.f:5941-6000). */
int bdf_last_stats(bdf_engine *e, double *quartets, double *kernel_ms);

/* Size of the stage-ilev build: algorithmic FP64 flops (the reference's loop
   nest counted per R batch, contraction and digestion), primitive quartets
   and work items.  No counterpart in the reference. */
int bdf_plan_info(bdf_engine *e, int ilev, double *flops, double *quartets, long *items);

/* Launch settings of an engine (no counterpart in the reference).  graph:
   1 = each Coulomb or Breit build is captured once per (stage, device
   buffers) as a HIP graph and replayed on later calls with the same
   buffers (bdf_coulomb and bdf_breit reuse one set of buffers); the
   captured launches are the uncaptured ones, so results agree to the
   rounding of the kernels' atomic accumulation.  waves (2, 3 or 4 waves per
   SIMD), cap (group capacity 1 or 2; three or four waves run at capacity 1) and
   per_cu (persistent workgroups per compute unit) set the Coulomb kernel's
   launch shape, breit_per_cu the Breit kernel's workgroups per CU (one
   wave each; the kernel is built for one wave per SIMD per four of them,
   2 to 4 waves; default 16) and breit_group the number of AB combinations
   it contracts together (1, 2 or 4, the last two at two waves per SIMD;
   its scratch grows with it); 0 keeps the engine's choice.  breit_kernel
   picks the Breit kernel: 1 the vector-ALU kernel (every shape; the
   per-CU and group settings above apply to it), 2 the matrix-core kernel
   (axial molecules and atoms, ISYM 1 and 2; builds that it cannot take
   fail with BDF_EUNSUPP), 0 the engine's choice (the matrix-core kernel
   wherever it applies).  Takes effect from the next build; captured builds
   are dropped. */
typedef struct {
  int graph;
  int waves;
  int cap;
  int per_cu;
  int breit_per_cu;
  int breit_group;
  int breit_kernel;
} bdf_launch;

int bdf_set_launch(bdf_engine *e, const bdf_launch *l);

/* How the stage-ilev Breit build is launched: its quartet kernel (1
   vector-ALU breit_kernel, 2 matrix-core breit_tc_kernel) under the current
   launch settings, and its persistent workgroups.  No counterpart in the
   reference. */
int bdf_breit_launch_info(bdf_engine *e, int ilev, int *kernel, int *workgroups);

/* How the stage-ilev Coulomb build is launched: waves per SIMD of its
   quartet kernel (2, 3 or 4), the capacity multiple of its AB-combination
   groups (1 or 2), persistent workgroups, and the engine's scratch in
   bytes.  No counterpart in the reference. */
int bdf_launch_info(bdf_engine *e, int ilev, int *waves, int *cap, int *workgroups,
                    size_t *scratch_bytes);

void bdf_destroy(bdf_engine *e);
const char *bdf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
