/*
 * bdf_oracle.h -- CPU FP64 restatement of the reference's Dirac-Coulomb
 * two-electron path (bertha-scf.f), used ONLY as the checker for the HIP
 * engine: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * The product never links or calls anything declared here.
 *
 * Every function follows the reference routine named in its comment, in
 * the reference's loop order, with the reference's screening thresholds and
 * Boys-function branches.  Indices in comments are bertha-scf.f line numbers.
 *
 * Parity status of each part (how it is pinned) is recorded in the header
 * comment of bdf_oracle.c.
 */
#ifndef BDF_ORACLE_H
#define BDF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* shape codes, as the reference's SHAPE -> ISYM map (bertha-scf.f:11573-11579) */
#define ORC_NOSYMM 0
#define ORC_LINEAR 1
#define ORC_ATOMIC 2

typedef struct orc_sys orc_sys;

/* basis: per centre lmax, nfnc[l] and exps[l][i] (row stride mbs).
   nfnc is [ncnt][lstride], exps is [ncnt][lstride][mbs]. */
orc_sys *orc_create(int ncnt, const double *xyz, const int *lmax,
                    const int *nfnc, int lstride, const double *exps, int mbs,
                    int shape, int intsym, int racah1);
void orc_destroy(orc_sys *s);
int orc_ndim(const orc_sys *s);
int orc_nskp(const orc_sys *s);

/* FOCKIND labels for the large block (1-based values, reference layout) */
void orc_labels(const orc_sys *s, int *icn, int *kqn, int *mqn);

/* FUNFM (bertha-scf.f:32465): F(n, lam+1) with n fastest */
void orc_funfm(const double *t, int n, int lam, int itype, double *f);
/* RMAKE (bertha-scf.f:31900): rc(n, ntuv(lam)), qp(n,3) = Q-P */
void orc_rmake(const double *qp, const double *aph, int n, int lam, double *rc);

/* E0LL (itt=1), E0SS (itt=4) or EILS (itt=2, 12 components: x, y, z
   times Re/Im E11, Re/Im E21) block; out[comp][tuv][ij]; returns ntuv*maxab
   or -1 if out too small.  ka,kb are 1-based kappa indices, ma,mb 1-based |m|
   indices, ica,icb 1-based centres (as IAD0LL's arguments). */
long orc_etable(orc_sys *s, int itt, int ica, int icb, int ka, int kb,
                int ma, int mb, double *out, long cap);

/* ERIFAST batch (bertha-scf.f:12269) after the CLMFAST set-up for a single
   (ibas,jbas) and component pair itn = {1|4, 1|4}; rr[16][maxcd] complex
   stored as re/im planes rr_re[16*maxcd], rr_im[16*maxcd]. */
int orc_erifast(orc_sys *s, const int icnt[4], const int kidx[4],
                const int midx[4], int ibas, int jbas, const int itn[2],
                int ilev, double *rr_re, double *rr_im);

/* CLMFAST (bertha-scf.f:11489): closed-shell Dirac-Coulomb G matrices.
   dent, gdir, gxch are column-major ndim x ndim complex (re/im interleaved). */
void orc_clmfast(orc_sys *s, int ilev, const double *dent, double *gdir,
                 double *gxch);

/* COULOMB1 (bertha-scf.f:21540) for every centre, as HFSCF calls it after
   CLMFAST with RACAH1 on (5415-5421): adds the one-centre Racah-algebra
   Coulomb contributions to gdir/gxch (a CLMFAST result of an oracle created
   with racah1 = 1) and overwrites their lower triangles with the conjugate
   upper ones, as COULOMB1 does on return. */
void orc_coulomb1(orc_sys *s, const double *dent, double *gdir, double *gxch);

/* BREIT1 (bertha-scf.f:26014) for every centre, as HFSCF calls it after
   BRTFAST with RACAH1 on (5447-5453): adds the one-centre Racah-algebra
   Breit contributions to bxch (a BRTFAST result of an oracle created with
   racah1 = 1) and overwrites its lower triangle with the conjugate upper
   one; bdir (no closed-shell contributions) is left alone. */
void orc_breit1(orc_sys *s, const double *dent, double *bdir, double *bxch);

/* same but with OpenMP over the outer centre-pair loop when nthreads > 1
   (digestion goes to thread-private matrices that are summed at the end) */
void orc_clmfast_mt(orc_sys *s, int ilev, const double *dent, double *gdir,
                    double *gxch, int nthreads);

/* CLMFAST as a timing sample (bench.py's CPU baseline): classes as work
   items on nthreads threads, only every stride-th primitive pair ij of each
   class evaluated; returns the fraction of the build's ERIFAST calls made.
   gdir/gxch receive the sample's partial sums (not a result). */
double orc_clmfast_sample(orc_sys *s, int ilev, const double *dent, double *gdir, double *gxch,
                          int nthreads, int stride);

/* BRTFAST (bertha-scf.f:17060): closed-shell Breit matrices BDIR (zero for
   closed shells) and BXCH, LS blocks and their conjugate SL blocks */
void orc_brtfast(orc_sys *s, int ilev, const double *dent, double *bdir,
                 double *bxch);
void orc_brtfast_mt(orc_sys *s, int ilev, const double *dent, double *bdir,
                    double *bxch, int nthreads);
/* BRTFAST as a timing sample, as orc_clmfast_sample (fraction of the
   build's BIIFAST calls made). */
double orc_brtfast_sample(orc_sys *s, int ilev, const double *dent, double *bdir, double *bxch,
                          int nthreads, int stride);

/* BIIFAST batch (bertha-scf.f:18337) for one (LS|LS) block quartet at a
   fresh class; rr planes as orc_erifast */
int orc_biifast(orc_sys *s, const int icnt[4], const int kidx[4],
                const int midx[4], int ibas, int jbas, double *rr_re,
                double *rr_im);
/* GAUNT2 toggle: Gaunt interaction only (no gauge term) */
void orc_set_gaunt(orc_sys *s, int on);

/* OVRLP (bertha-scf.f:6683) and ONEEL (:6996): overlap, nuclear attraction
   (LL, and SS including -2 c^2 S_SS) and kinetic LS/SL coupling matrices,
   ndim x ndim complex column-major.  znuc[c] nuclear charges; xi[c] the
   Gaussian-nucleus exponent (<= 0: point nucleus); cv the speed of light. */
void orc_oneel(orc_sys *s, const double *znuc, const double *xi, double cv,
               double *ovlp, double *hnuc, double *hkin);

/* SCHWARZ (bertha-scf.f:37889) for one ERIFAST call on an explicit bound
   table g (ndim x ndim real, column-major) and density dent (complex
   column-major): nlab = NA1, NA2, NB1, NB2, NC1, NC2, ND1, ND2 (0-based
   block offsets), ibas/jbas 1-based, nb3 x nb4 kl pairs; msk[m] bit b keeps
   batch b+1, 0 drops the kl. */
void orc_schwarz_probe(const orc_sys *s, const double *g, const double *dent, const int nlab[8],
                       int ibas, int jbas, int nb3, int nb4, double sens, int *msk);

/* CLMDIAG (bertha-scf.f:38126): GDSC, ndim x ndim real column-major */
void orc_clmdiag(orc_sys *s, double *gdsc);

/* sqrt|(a b|b a)| for LL and SS pairs, ndim x ndim real column-major: the
   Cauchy-Schwarz factors of the engine's FP32/FP64 split (not a reference
   routine) */
void orc_pairbound(orc_sys *s, double *out);

/* SCHWRZ toggle for orc_clmfast*: on = 1, SCHWARZ masks (bertha-scf.f:37889)
   from CLMDIAG's GDSC; on = 2, the same decision tree with every group bounded
   by the Cauchy-Schwarz factors of orc_pairbound (not the reference's: no
   dropped term exceeds sens); sensitivity sens (<= 0: the reference's 1e-12) */
void orc_set_schwarz(orc_sys *s, int on, double sens);

/* largest |term| among the digestion terms the masks dropped since the last
   orc_set_schwarz, and their number (single-threaded builds only) */
void orc_schwarz_stats(const orc_sys *s, double *drop_max, double *drop_count);

#ifdef __cplusplus
}
#endif
#endif
