// api.cpp -- the C ABI (include/bertha_amd.h): engine lifetime, device
// tables, Fock builds and test hooks.  Every entry point returns an error code
// instead of stopping the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bertha_amd.h"
#include "kernels.h"
#include "plan.h"

using namespace bdf;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t _e = (x);                                                              \
    if (_e != hipSuccess)                                                             \
      return fail(BDF_EDEVICE, std::string(#x) + ": " + hipGetErrorString(_e));      \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    n = count;
    return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
  }
  hipError_t upload(const std::vector<T> &v, hipStream_t s) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

// device copy of one stage's plan
struct DevPlan {
  int built = 0;
  int xacc_n = 0, blocks = 1;  // exchange-accumulator doubles, workgroups per launch
  FockLaunch fl{2, 1};         // waves per SIMD and group capacity of the launch
  Plan P;
  DevBuf<double> qitem;  // Schwarz factor of each work item (mixed precision)
  DevBuf<ClassDesc> classes;
  DevBuf<ABCombo> abs;
  DevBuf<CDCall> cds;
  DevBuf<uint8_t> masks;
  DevBuf<WorkItem> items;
};

struct DevBreitPlan {
  int built = 0;
  BreitPlan P;
  DevBuf<ClassDesc> classes;
  DevBuf<BrAB> abs;
  DevBuf<BrCD> cds;
  DevBuf<uint8_t> masks;
  DevBuf<WorkItem> items;
  DevBuf<uint8_t> stale;
  // matrix-core kernel (real path)
  DevBuf<BrTile> tiles;
  DevBuf<int> tile_cd;
  DevBuf<WorkItem> titems;
  DevBuf<int> sig, csig;
  DevBuf<int> tile_pair;
  DevBuf<BrComp> comp;
  DevBuf<WorkItem> rest_items;
};

// device copy of the one-centre Racah plan
struct DevRacah {
  int built = 0;
  RacahPlan P;
  DevBuf<RcBlock> blocks;
  DevBuf<RcSeg> segs;
  DevBuf<RcCombo> combos;
  DevBuf<RcEntry> entries;
  DevBuf<double> ang;
  DevBuf<RcOp> ops;
  DevBuf<RcItem> items;
};

// One captured build: the launches of a Coulomb (kind 0) or Breit (kind 1)
// build of stage ilev on fixed device buffers, replayed as a HIP graph.
struct GraphEntry {
  int kind, ilev;
  const void *d, *g, *x;
  hipGraphExec_t exec;
};

struct bdf_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  Basis B;
  Etabs T;
  int isym = 0, intsym = 1, racah1 = 0;
  std::vector<uint8_t> eflags;
  std::vector<uint8_t> zflags;  // DHFB or axial: exact-nonzero flags of the E tables
  DevBuf<double> exps, normL, normS;
  DevBuf<EPair> pairs;
  DevBuf<ETerm> terms;
  DevBuf<Mono> monos;
  DevBuf<int64_t> first, flag_off;
  DevBuf<double4> E;
  DevBuf<uint8_t> flags;
  DevBuf<int> icn, kqn, mqn;
  DevPlan plans[3];
  DevBuf<int> counter;
  DevBuf<double> slab;
  int nblocks = 0;        // the caller's persistent workgroups (user_blocks)
  int cus = 0, user_blocks = 0;
  // scratch rows per workgroup and workgroups of the Coulomb (0) and Breit
  // (1) kernels, at one row length slab_ld; the Coulomb kernel of a mixed
  // engine takes a second set for its FP32 instance
  int slab_ld = 0;
  int slab_rows[2] = {0, 0}, slab_blocks[2] = {0, 0};
  DevBuf<double2> dent, gdir, gxch;
  // Schwarz tables (GDSC of CLMDIAG and the (ab|ba) bound), built on demand
  int schwarz = 0, have_tables = 0;
  double sens = 1.0e-12;
  DevBuf<double> gdsc, qbnd;
  DevBuf<int> partner;          // the other sign slot of every row (Schwarz tables)
  DevBuf<double> gslot, dslot;  // slot maxima of GDSC and |D| (SCHWARZ's direct groups)
  std::vector<double> qbnd_host;
  // FP32/FP64 split
  int mixed = 0;
  double split = 0.1;
  int breit = 0;  // DHFB: LS tables and the Breit plans
  int gaunt = 0;  // GAUNT2
  DevBreitPlan bplans[3];
  DevRacah racah[2];  // COULOMB1 (bdf_coulomb1) and BREIT1 (bdf_breit1) replacements
  DevBuf<float4> E32;
  DevBuf<double4> Ecd;  // masked, signed E(CD) of the second contraction
  DevBuf<double4> Eab;  // E with EPSAB applied: the Breit kernel's AB side
  DevBuf<float4> Ecd32;
  DevBuf<double> Ecd1;  // axial: each Ecd entry's one nonzero component
  DevBuf<float> Ecd1_32;
  DevBuf<unsigned long long> dmax;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  hipStream_t stream32 = nullptr;  // FP32 instance of the mixed build
  hipEvent_t fork = nullptr, join = nullptr;
  double last_ms = 0.0, last_quartets = 0.0;
  // launch settings (bdf_set_launch): 0 = the engine's choice
  int tune_waves = 0, tune_cap = 0, tune_per_cu = 0, tune_breit_per_cu = 0;
  int tune_breit_group = 0;
  int tune_breit_kernel = 0;
  int use_graph = 0;
  std::vector<GraphEntry> graphs;
  // this device's static scratch (bdf_coulomb_device_static)
  double *static_scratch = nullptr;
};

// where one build's per-workgroup scratch lives: base, doubles per
// workgroup, workgroups it holds
struct SlabView {
  double *p;
  int64_t stride;
  int blocks;
};

static void drop_graphs(bdf_engine *e) {
  for (GraphEntry &g : e->graphs) (void)hipGraphExecDestroy(g.exec);
  e->graphs.clear();
}

static void item_factors(const bdf_engine *e, const Plan &P, std::vector<double> &q);

// persistent workgroups of a launch: the caller's count (bdf_options.
// workgroups), else per_cu per compute unit
static int blocks_for(const bdf_engine *e, int per_cu) {
  return e->user_blocks ? e->nblocks : e->cus * per_cu;
}

// exchange-accumulator doubles of a plan's largest class (the Coulomb
// kernel's dynamic LDS): four sign slots of (re, im) per entry, or on the
// real path (ISYM 1, 2) the two diagonal slots' real parts
static int plan_xacc(const Plan &P, int isym) {
  int n = 0;
  for (const ClassDesc &cl : P.classes) {
    const int nC = cl.nbas[2], nD = cl.nbas[3];
    n = std::max(n, isym != 0 ? 6 * nD + 8 * nC : 2 * (12 * nD + 16 * nC));
  }
  return n + (n & 1);  // even: 16-byte alignment of what follows
}

// the same for a Breit plan (its exchange batches use the same layout)
static int breit_xacc(const BreitPlan &P, int isym) {
  int n = 0;
  for (const ClassDesc &cl : P.classes) {
    const int nC = cl.nbas[2], nD = cl.nbas[3];
    n = std::max(n, isym != 0 ? 6 * nD + 8 * nC : 2 * (12 * nD + 16 * nC));
  }
  return n + (n & 1);
}

// doubles per workgroup of kernel kind (0 Coulomb, 1 Breit)
static int64_t slab_stride(const bdf_engine *e, int kind) {
  return (int64_t)e->slab_rows[kind] * e->slab_ld;
}

// per-workgroup scratch of kernel kind for `blocks` workgroups: R rows, Boys
// rows and g_rows G rows, each ld doubles (ld = largest kl count rounded to
// the wavefront); the allocation serves the larger of the two kinds
static int ensure_slab(bdf_engine *e, int kind, int maxcd, int g_rows, int blocks) {
  int ld = ((std::max(maxcd, 1) + 63) / 64) * 64;
  int rows = ntuv(kMaxLfl) + kMaxLfl + 1 + g_rows;
  if (ld > e->slab_ld || rows > e->slab_rows[kind] || blocks > e->slab_blocks[kind] ||
      !e->slab.p) {
    e->slab_ld = std::max(ld, e->slab_ld);
    e->slab_rows[kind] = std::max(rows, e->slab_rows[kind]);
    e->slab_blocks[kind] = std::max(blocks, e->slab_blocks[kind]);
    // the mixed build's FP32 instance runs concurrently on a second set of blocks
    const size_t need = std::max<size_t>(
        (size_t)slab_stride(e, 0) * e->slab_blocks[0] * (e->mixed ? 2 : 1),
        (size_t)slab_stride(e, 1) * e->slab_blocks[1]);
    if (need > e->slab.n || !e->slab.p) {
      // captured builds hold the old scratch address
      drop_graphs(e);
      HIPCHK(e->slab.alloc(need));
    }
  }
  return BDF_OK;
}

// the Coulomb kernel's launch shape for a plan: the engine's choice by LDS
// (fock_launch_config), with the settings of bdf_set_launch applied
static FockLaunch fock_config(const bdf_engine *e, const DevPlan &dp) {
  FockLaunch fl = fock_launch_config(dp.xacc_n, e->isym);
  if (e->tune_waves) fl.waves = e->tune_waves;
  if (e->tune_cap) fl.cap = e->tune_cap;
  if (e->tune_waves >= 3 && !e->tune_cap) fl.cap = 1;
  return fl;
}

// workgroups and scratch of a built plan for its launch shape
static int size_launch(bdf_engine *e, DevPlan &dp) {
  dp.fl = fock_config(e, dp);
  dp.blocks = blocks_for(e, e->tune_per_cu ? e->tune_per_cu : 4 * dp.fl.waves);
  return ensure_slab(e, 0, dp.P.max_maxcd, 4 * dp.fl.cap * ntuv(kMaxLab), dp.blocks);
}

static int breit_per_cu(const bdf_engine *e) {
  return e->tune_breit_per_cu ? e->tune_breit_per_cu : 16;
}

static int breit_blocks(const bdf_engine *e) { return blocks_for(e, breit_per_cu(e)); }

// AB combinations the Breit kernel contracts together
static int breit_group(const bdf_engine *e) {
  return e->tune_breit_group ? e->tune_breit_group : 1;
}

// the matrix-core Breit kernel for this plan: on the real path, where its
// largest class's LDS fits a compute unit, unless the vector-ALU kernel is
// asked for
static bool breit_tc(const bdf_engine *e, const BreitPlan &P) {
  return e->isym != 0 && e->tune_breit_kernel != 1 && P.tc_ok && !P.titems.empty() &&
         P.tc_lds <= (size_t)kTcLdsMax;
}

// workgroups (kTcWaves waves each) per CU of the matrix-core kernel: one, its
// two waves per SIMD sharing the LDS tables
static int breit_tc_blocks(const bdf_engine *e, const BreitPlan &P) {
  (void)P;
  const int per_cu = 1;
  return e->user_blocks ? e->nblocks : e->cus * per_cu;
}

// waves per SIMD the Breit kernel is compiled for: one per four workgroups
// per CU (its workgroups are one wave), 2..4; grouped combinations need the
// registers of two
static int breit_waves(const bdf_engine *e) {
  if (breit_group(e) > 1) return 2;
  return std::max(2, std::min(4, breit_per_cu(e) / 4));
}

static int make_plan(bdf_engine *e, int ilev) {
  DevPlan &dp = e->plans[ilev - 1];
  if (dp.built) return BDF_OK;
  std::string err = build_plan(e->B, e->T, e->eflags, e->isym, e->intsym, e->racah1, ilev, dp.P,
                               e->zflags.empty() ? nullptr : &e->zflags);
  if (!err.empty()) return fail(BDF_EINVAL, err);
  HIPCHK(dp.classes.upload(dp.P.classes, e->stream));
  HIPCHK(dp.abs.upload(dp.P.abs, e->stream));
  HIPCHK(dp.cds.upload(dp.P.cds, e->stream));
  HIPCHK(dp.masks.upload(dp.P.masks, e->stream));
  HIPCHK(dp.items.upload(dp.P.items, e->stream));
  // (held until the stream synchronises below: the upload is asynchronous)
  std::vector<double> q;
  if (e->mixed) {
    item_factors(e, dp.P, q);
    HIPCHK(dp.qitem.upload(q, e->stream));
  }
  // four workgroups per CU for each wave per SIMD the kernel runs at, G rows
  // for its combination groups
  dp.xacc_n = plan_xacc(dp.P, e->isym);
  int rc = size_launch(e, dp);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipStreamSynchronize(e->stream));
  dp.built = 1;
  return BDF_OK;
}

static int make_racah(bdf_engine *e, int breit) {
  DevRacah &r = e->racah[breit];
  if (r.built) return BDF_OK;
  std::string err;
  try {
    err = build_racah_plan(e->B, e->isym, breit, r.P);
  } catch (const std::bad_alloc &) {
    err = "racah plan: out of host memory";
  }
  if (!err.empty()) return fail(BDF_ELIMIT, err);
  HIPCHK(r.blocks.upload(r.P.blocks, e->stream));
  HIPCHK(r.segs.upload(r.P.segs, e->stream));
  HIPCHK(r.combos.upload(r.P.combos, e->stream));
  HIPCHK(r.entries.upload(r.P.entries, e->stream));
  HIPCHK(r.ang.upload(r.P.ang, e->stream));
  HIPCHK(r.ops.upload(r.P.ops, e->stream));
  HIPCHK(r.items.upload(r.P.items, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  r.built = 1;
  return BDF_OK;
}

// CLMDIAG (bertha-scf.f:38126) and the (ab|ba) bound table on the device:
// one DiagJob per (A, B, kappa_A, kappa_B, |m_A|, |m_B|) and component pair.
static int make_tables(bdf_engine *e) {
  if (e->have_tables) return BDF_OK;
  const Basis &B = e->B;
  const Etabs &T = e->T;
  std::vector<DiagJob> jobs;
  std::vector<int64_t> first;
  int64_t acc = 0;
  try {
  for (int a = 0; a < B.ncnt; ++a)
    for (int b = 0; b < B.ncnt; ++b)
      for (int ka = 1; ka <= 2 * B.lmax[a] + 1; ++ka)
        for (int kb = 1; kb <= 2 * B.lmax[b] + 1; ++kb) {
          int kqa = B.kappa_of(a, ka), kqb = B.kappa_of(b, kb);
          int la = kqa < 0 ? -kqa - 1 : kqa, lb = kqb < 0 ? -kqb - 1 : kqb;
          for (int ma = 1; ma <= std::abs(kqa); ++ma)
            for (int mb = 1; mb <= std::abs(kqb); ++mb)
              for (int job = 0; job < 6; ++job) {
                // jobs 0-3: GDSC for (T1,T2) = (S,S),(S,L),(L,S),(L,L); 4-5: bound, T = S, L
                int t1 = job < 4 ? (job < 2) : (job == 4);
                int t2 = job < 4 ? (job % 2 == 0) : t1;
                DiagJob J;
                const EPair &pab = T.pairs[T.pair_index.at(Etabs::key(a, b, ka, kb, ma, mb, t1))];
                const EPair &pcd = job < 4
                    ? T.pairs[T.pair_index.at(Etabs::key(a, b, ka, kb, ma, mb, t2))]
                    : T.pairs[T.pair_index.at(Etabs::key(b, a, kb, ka, mb, ma, t2))];
                J.eab = pab.off;
                J.ecd = pcd.off;
                J.ntab = ntuv(la + lb + (t1 ? 2 : 0));
                J.ntcd = ntuv(la + lb + (t2 ? 2 : 0));
                J.nA = B.nfn(a, la);
                J.nB = B.nfn(b, lb);
                J.expA = B.expo(a, la);
                J.expB = B.expo(b, lb);
                J.row[0] = B.lrge_at(a, ka, 2 * ma - 1) + (t1 ? B.nskp : 0);
                J.row[1] = B.lrge_at(a, ka, 2 * ma) + (t1 ? B.nskp : 0);
                J.col[0] = B.lrge_at(b, kb, 2 * mb - 1) + (t2 ? B.nskp : 0);
                J.col[1] = B.lrge_at(b, kb, 2 * mb) + (t2 ? B.nskp : 0);
                J.swap = job >= 4;
                J.realpart = job < 4 && e->isym != 0;
                J.out = job >= 4;
                J.pad = 0;
                first.push_back(acc);
                acc += (int64_t)J.nA * J.nB;
                jobs.push_back(J);
              }
        }
  } catch (const std::out_of_range &) {
    return fail(BDF_EINVAL, "Schwarz tables: missing E-coefficient block");
  }
  size_t nn = (size_t)B.ndim * B.ndim;
  HIPCHK(e->gdsc.alloc(nn));
  HIPCHK(e->qbnd.alloc(nn));
  HIPCHK(hipMemsetAsync(e->gdsc.p, 0, nn * sizeof(double), e->stream));
  HIPCHK(hipMemsetAsync(e->qbnd.p, 0, nn * sizeof(double), e->stream));
  DevBuf<DiagJob> djobs;
  DevBuf<int64_t> dfirst;
  HIPCHK(djobs.upload(jobs, e->stream));
  HIPCHK(dfirst.upload(first, e->stream));
  DiagArgs da;
  da.jobs = djobs.p;
  da.njobs = (int)jobs.size();
  da.first = dfirst.p;
  da.nitems = acc;
  da.exps = e->exps.p;
  da.E = e->E.p;
  da.gdsc = e->gdsc.p;
  da.qbnd = e->qbnd.p;
  da.ndim = B.ndim;
  HIPCHK(launch_diag(da, e->stream));
  // (held until the stream synchronises below: the upload is asynchronous)
  std::vector<int> partner;
  if (e->schwarz) {
    // row r's other sign slot: the (-|m|, +|m|) rows of one (centre, kappa,
    // |m|) block, large and small
    partner.assign(B.ndim, -1);
    for (int c = 0; c < B.ncnt; ++c)
      for (int k = 1; k <= 2 * B.lmax[c] + 1; ++k) {
        const int kq = B.kappa_of(c, k), l = kq < 0 ? -kq - 1 : kq;
        for (int m = 1; m <= std::abs(kq); ++m)
          for (int t = 0; t < 2; ++t) {
            const int lo = B.lrge_at(c, k, 2 * m - 1) + t * B.nskp;
            const int hi = B.lrge_at(c, k, 2 * m) + t * B.nskp;
            for (int i = 0; i < B.nfn(c, l); ++i) {
              partner[lo + i] = hi + i;
              partner[hi + i] = lo + i;
            }
          }
      }
    for (int r = 0; r < B.ndim; ++r)
      if (partner[r] < 0 || partner[r] >= B.ndim)
        return fail(BDF_EINVAL, "Schwarz tables: row without a sign-slot partner");
    HIPCHK(e->partner.upload(partner, e->stream));
    HIPCHK(e->gslot.alloc(nn));
    HIPCHK(e->dslot.alloc(nn));
    // mode 1 bounds with GDSC as the reference; mode 2 with the Cauchy-Schwarz
    // factors Q
    HIPCHK(launch_slot_tables(e->schwarz == 2 ? e->qbnd.p : e->gdsc.p, nullptr, e->partner.p,
                              B.ndim, e->gslot.p, e->stream));
  }
  e->qbnd_host.resize(nn);
  HIPCHK(hipMemcpyAsync(e->qbnd_host.data(), e->qbnd.p, nn * sizeof(double),
                        hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  e->have_tables = 1;
  return BDF_OK;
}

// Schwarz factor of every work item: max over the item's AB combinations and
// sign slots of Q(a,b) times the largest Q(c,d) of its class, so that every
// quartet (ij|kl) of the item obeys |(ij|kl)| <= qitem.
static void item_factors(const bdf_engine *e, const Plan &P, std::vector<double> &q) {
  const int nd = e->B.ndim;
  const double *Q = e->qbnd_host.data();
  auto at = [&](int r, int c) { return Q[(int64_t)c * nd + r]; };
  std::vector<double> qcd(P.classes.size(), 0.0);
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    double m = 0.0;
    for (int abi = cl.ab_begin; abi < cl.ab_end; ++abi)
      for (int ci = P.abs[abi].cd_begin; ci < P.abs[abi].cd_end; ++ci) {
        const CDCall &cd = P.cds[ci];
        for (int k = 0; k < cl.nbas[2]; ++k)
          for (int l = 0; l < cl.nbas[3]; ++l)
            for (int s = 0; s < 4; ++s) m = std::max(m, at(cd.fc[s >> 1] + k, cd.fd[s & 1] + l));
      }
    qcd[c] = m;
  }
  // FP32 range guard: before the prefactor an item's R batch grows like
  // (2 alpha max(1, |PQ|))^L, alpha = pq/(p+q), and tight pairs also carry
  // large Hermite expansions; an item whose growth exceeds 1e20 stays in FP64
  // whatever its Schwarz factor (FP32 overflows at 3.4e38).
  std::vector<double> qmax(P.classes.size(), 0.0), dmax(P.classes.size(), 1.0);
  const std::vector<double> &X = e->B.exps;
  for (size_t c = 0; c < P.classes.size(); ++c) {
    const ClassDesc &cl = P.classes[c];
    for (int k = 0; k < cl.nbas[2]; ++k)
      for (int l = 0; l < cl.nbas[3]; ++l)
        qmax[c] = std::max(qmax[c], X[cl.exp[2] + k] + X[cl.exp[3] + l]);
    for (int u = 0; u < 4; ++u)
      for (int v = 0; v < 4; ++v) {
        double dx = cl.xyz[u][0] - cl.xyz[v][0], dy = cl.xyz[u][1] - cl.xyz[v][1];
        double dz = cl.xyz[u][2] - cl.xyz[v][2];
        dmax[c] = std::max(dmax[c], std::sqrt(dx * dx + dy * dy + dz * dz));
      }
  }
  q.resize(P.items.size());
  for (size_t i = 0; i < P.items.size(); ++i) {
    const WorkItem &it = P.items[i];
    const ClassDesc &cl = P.classes[it.cls];
    int ib = it.ij / cl.nbas[1], jb = it.ij % cl.nbas[1];
    double m = 0.0;
    for (int abi = it.ab_lo; abi < it.ab_hi; ++abi) {
      const ABCombo &ab = P.abs[abi];
      for (int s = 0; s < 4; ++s) m = std::max(m, at(ab.fa[s >> 1] + ib, ab.fb[s & 1] + jb));
    }
    q[i] = m * qcd[it.cls];
    const double p = X[cl.exp[0] + ib] + X[cl.exp[1] + jb];
    const double alpha = p * qmax[it.cls] / (p + qmax[it.cls]);
    const double growth = cl.lamfl * std::log10(std::max(1.0, 2.0 * alpha) * dmax[it.cls]);
    if (growth > 20.0) q[i] = std::numeric_limits<double>::infinity();
  }
}

// The launches of one Coulomb build on the engine's stream; with timing,
// ev0 and ev2 bracket the quartet kernel(s).
static int enqueue_fock(bdf_engine *e, DevPlan &dp, const double2 *d_dent, double2 *d_gdir,
                        double2 *d_gxch, double2 *rr_out, bool mixed, bool timing,
                        const SlabView &sv) {
  int nd = e->B.ndim;
  if (!rr_out)
    HIPCHK(launch_zero(d_gdir, d_gxch, (int64_t)nd * nd, e->counter.p, e->dmax.p, e->stream));
  else
    HIPCHK(launch_zero(rr_out, nullptr, 0, e->counter.p, e->dmax.p, e->stream));
  if (mixed) HIPCHK(launch_dmax(d_dent, (int64_t)nd * nd, e->dmax.p, e->stream));
  if (e->schwarz && !rr_out)
    HIPCHK(launch_slot_tables(nullptr, d_dent, e->partner.p, nd, e->dslot.p, e->stream));
  FockArgs a{};
  a.classes = dp.classes.p;
  a.abs = dp.abs.p;
  a.cds = dp.cds.p;
  a.masks = dp.masks.p;
  a.items = dp.items.p;
  a.nitems = (int)dp.P.items.size();
  a.counter = e->counter.p;
  a.exps = e->exps.p;
  a.E = e->E.p;
  a.dent = d_dent;
  a.gdir = d_gdir;
  a.gxch = d_gxch;
  a.ndim = nd;
  a.slab = sv.p;
  a.slab_stride = sv.stride;
  a.ld = e->slab_ld;
  a.isym = e->isym;
  a.intsym = e->intsym;
  a.rr_out = rr_out;
  a.gdsc = e->gdsc.p;
  a.gslot = e->gslot.p;
  a.dslot = e->dslot.p;
  a.sens = e->sens;
  a.schwarz = rr_out ? 0 : e->schwarz;
  a.nskp = e->B.nskp;
  a.mixed = mixed ? 1 : 0;
  a.split = e->split;
  a.qitem = mixed ? dp.qitem.p : nullptr;
  a.dmax_bits = e->dmax.p;
  a.E32 = e->E32.p;
  a.Ecd = e->Ecd.p;
  a.Ecd32 = e->Ecd32.p;
  a.xacc_n = plan_xacc(dp.P, e->isym);
  a.one_comp = dp.P.one_comp && e->Ecd1.p && (!mixed || e->Ecd1_32.p);
  a.Ecd1 = e->Ecd1.p;
  a.Ecd1_32 = e->Ecd1_32.p;
  const FockLaunch fl = dp.fl;
  const int nb = std::min(sv.blocks, std::max(1, a.nitems));
  if (timing) HIPCHK(hipEventRecord(e->ev0, e->stream));
  if (a.nitems > 0) HIPCHK(launch_fock(a, fl, nb, e->stream, e->stream32, e->fork, e->join));
  if (timing) HIPCHK(hipEventRecord(e->ev2, e->stream));
  if (!rr_out) {
    CompleteArgs c;
    c.gdir = d_gdir;
    c.gxch = d_gxch;
    c.icn = e->icn.p;
    c.kqn = e->kqn.p;
    c.mqn = e->mqn.p;
    c.ndim = nd;
    c.nskp = e->B.nskp;
    c.intsym = e->intsym;
    HIPCHK(launch_complete(c, e->stream));
  }
  return BDF_OK;
}

// Replays the captured build of (kind, ilev, buffers), capturing it first if
// it is new; enqueue(timing) issues the build's launches.  ev0 and ev2/ev1
// bracket the whole graph.
template <class F>
static int run_graph(bdf_engine *e, int kind, int ilev, const void *d, const void *g,
                     const void *x, F enqueue) {
  GraphEntry *hit = nullptr;
  for (GraphEntry &ge : e->graphs)
    if (ge.kind == kind && ge.ilev == ilev && ge.d == d && ge.g == g && ge.x == x) hit = &ge;
  if (!hit) {
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(false);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
    if (rc != BDF_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIPCHK(ce);
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(ie);
    e->graphs.push_back({kind, ilev, d, g, x, exec});
    hit = &e->graphs.back();
  }
  HIPCHK(hipEventRecord(e->ev0, e->stream));
  HIPCHK(hipGraphLaunch(hit->exec, e->stream));
  HIPCHK(hipEventRecord(e->ev2, e->stream));
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  return BDF_OK;
}

static int run_fock(bdf_engine *e, int ilev, const double2 *d_dent, double2 *d_gdir,
                    double2 *d_gxch, double2 *rr_out, DevPlan *override_plan) {
  DevPlan &dp = override_plan ? *override_plan : e->plans[ilev - 1];
  const bool mixed = e->mixed && !rr_out && !override_plan;
  e->last_quartets = dp.P.quartets;
  const SlabView sv{e->slab.p, slab_stride(e, 0), dp.blocks};
  if (e->use_graph && !rr_out && !override_plan)
    return run_graph(e, 0, ilev, d_dent, d_gdir, d_gxch, [&](bool timing) {
      return enqueue_fock(e, dp, d_dent, d_gdir, d_gxch, nullptr, mixed, timing, sv);
    });
  int rc = enqueue_fock(e, dp, d_dent, d_gdir, d_gxch, rr_out, mixed, true, sv);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  return BDF_OK;
}

static int make_breit_plan(bdf_engine *e, int ilev) {
  // with INTSYM off BRTFAST also loops over SL component pairs (IT = 3,
  // :17139-17147) from tables EQSAVE does not build for it; HFSCF always runs
  // it with INTSYM on (:5215)
  if (!e->intsym) return fail(BDF_EUNSUPP, "Breit with intsym = 0 is not supported");
  DevBreitPlan &dp = e->bplans[ilev - 1];
  if (dp.built) return BDF_OK;
  std::string err;
  try {
    err = build_breit_plan(e->B, e->T, e->eflags, e->isym, e->intsym, e->racah1, ilev, e->gaunt,
                           dp.P, e->zflags.empty() ? nullptr : &e->zflags);
  } catch (const std::out_of_range &) {
    err = "Breit plan: missing E-coefficient block";
  }
  if (!err.empty()) return fail(BDF_EINVAL, err);
  HIPCHK(dp.classes.upload(dp.P.classes, e->stream));
  HIPCHK(dp.abs.upload(dp.P.abs, e->stream));
  HIPCHK(dp.cds.upload(dp.P.cds, e->stream));
  HIPCHK(dp.masks.upload(dp.P.masks, e->stream));
  HIPCHK(dp.items.upload(dp.P.items, e->stream));
  HIPCHK(dp.stale.alloc((size_t)std::max(dp.P.nflag, 1)));
  if (dp.P.tc_ok) {
    HIPCHK(dp.tiles.upload(dp.P.tiles, e->stream));
    HIPCHK(dp.tile_cd.upload(dp.P.tile_cd, e->stream));
    HIPCHK(dp.tile_pair.upload(dp.P.tile_pair, e->stream));
    HIPCHK(dp.comp.upload(dp.P.comp, e->stream));
    HIPCHK(dp.titems.upload(dp.P.titems, e->stream));
    HIPCHK(dp.rest_items.upload(dp.P.rest_items, e->stream));
    HIPCHK(dp.sig.upload(dp.P.tc_sig, e->stream));
    HIPCHK(dp.csig.upload(dp.P.tc_csig, e->stream));
  }
  if (e->tune_breit_kernel == 2 && !breit_tc(e, dp.P))
    return fail(BDF_EUNSUPP, "matrix-core Breit kernel: not a real-path system, or its LDS "
                             "exceeds a compute unit");
  int rc = ensure_slab(e, 1, dp.P.max_maxcd, 12 * breit_group(e) * std::max(dp.P.max_ntg, 1),
                       breit_blocks(e));
  if (rc == BDF_OK && breit_tc(e, dp.P))
    rc = ensure_slab(e, 1, dp.P.max_maxcd, 0, breit_tc_blocks(e, dp.P));
  if (rc != BDF_OK) return rc;
  HIPCHK(hipStreamSynchronize(e->stream));
  dp.built = 1;
  return BDF_OK;
}

static int enqueue_breit(bdf_engine *e, DevBreitPlan &dp, const double2 *d_dent,
                         double2 *d_bdir, double2 *d_bxch, double2 *rr_out, bool timing) {
  int nd = e->B.ndim;
  if (!rr_out)
    HIPCHK(launch_zero(d_bdir, d_bxch, (int64_t)nd * nd, e->counter.p, nullptr, e->stream));
  else
    HIPCHK(launch_zero(rr_out, nullptr, 0, e->counter.p, nullptr, e->stream));
  BreitArgs a;
  a.classes = dp.classes.p;
  a.nclasses = (int)dp.P.classes.size();
  a.abs = dp.abs.p;
  a.cds = dp.cds.p;
  a.masks = dp.masks.p;
  a.items = dp.items.p;
  a.nitems = (int)dp.P.items.size();
  a.counter = e->counter.p;
  a.exps = e->exps.p;
  a.E = e->E.p;
  a.dent = d_dent;
  a.bdir = d_bdir;
  a.bxch = d_bxch;
  a.ndim = nd;
  a.nskp = e->B.nskp;
  a.slab = e->slab.p;
  a.slab_stride = slab_stride(e, 1);
  a.ld = e->slab_ld;
  a.group = rr_out ? 1 : breit_group(e);
  a.waves = rr_out ? 2 : breit_waves(e);
  a.xacc_n = breit_xacc(dp.P, e->isym);
  a.isym = e->isym;
  a.intsym = e->intsym;
  a.gaunt = e->gaunt;
  a.stale = dp.stale.p;
  a.Ecd = e->Ecd.p;
  a.Eab = e->Eab.p;
  a.rr_out = rr_out;
  if (timing) HIPCHK(hipEventRecord(e->ev0, e->stream));
  if (!rr_out && breit_tc(e, dp.P)) {
    BreitTcArgs t;
    t.classes = dp.classes.p;
    t.abs = dp.abs.p;
    t.cds = dp.cds.p;
    t.masks = dp.masks.p;
    t.tiles = dp.tiles.p;
    t.tile_cd = dp.tile_cd.p;
    t.tile_pair = dp.tile_pair.p;
    t.comp = dp.comp.p;
    t.max_npair = dp.P.max_npair;
    t.items = dp.titems.p;
    t.nitems = (int)dp.P.titems.size();
    t.counter = e->counter.p;
    t.exps = e->exps.p;
    t.Eab = e->Eab.p;
    t.Ecd = e->Ecd.p;
    t.dent = d_dent;
    t.bxch = d_bxch;
    t.ndim = nd;
    t.slab = e->slab.p;
    t.slab_stride = slab_stride(e, 1);
    t.intsym = e->intsym;
    t.gaunt = e->gaunt;
    t.stale = dp.stale.p;
    t.sig = dp.sig.p;
    t.csig = dp.csig.p;
    BreitArgs rest = a;
    rest.items = dp.rest_items.p;
    rest.nitems = (int)dp.P.rest_items.size();
    rest.counter = e->counter.p + 1;
    rest.group = 1;
    rest.waves = 4;
    HIPCHK(launch_breit_tc(a, t, breit_tc_blocks(e, dp.P), dp.P.tc_lds, dp.P.max_nu, rest,
                           blocks_for(e, 16), e->stream));
  } else {
    HIPCHK(launch_breit(a, breit_blocks(e), e->stream));
  }
  if (timing) HIPCHK(hipEventRecord(e->ev2, e->stream));
  return BDF_OK;
}

static int run_breit(bdf_engine *e, DevBreitPlan &dp, int ilev, const double2 *d_dent,
                     double2 *d_bdir, double2 *d_bxch, double2 *rr_out) {
  e->last_quartets = dp.P.quartets;
  if (e->use_graph && !rr_out)
    return run_graph(e, 1, ilev, d_dent, d_bdir, d_bxch, [&](bool timing) {
      return enqueue_breit(e, dp, d_dent, d_bdir, d_bxch, nullptr, timing);
    });
  int rc = enqueue_breit(e, dp, d_dent, d_bdir, d_bxch, rr_out, true);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  return BDF_OK;
}

extern "C" {

const char *bdf_last_error(void) { return g_err.c_str(); }

int bdf_create(const bdf_system *sys, const bdf_options *opt, bdf_engine **out, int *ndim,
               int *nskp) {
  if (!sys || !out) return fail(BDF_EINVAL, "bdf_create: null argument");
  *out = nullptr;
  if (sys->hmlt != BDF_HMLT_DHFR && sys->hmlt != BDF_HMLT_DHFB)
    return fail(BDF_EINVAL, "bdf_create: unknown hamiltonian");
  if (sys->shape < 0 || sys->shape > 2) return fail(BDF_EINVAL, "bdf_create: unknown shape");
  bdf_options o = {1, 0, BDF_PREC_FP64, 0.0, 0, 0, 0.0, 0, 0};
  if (opt) o = *opt;
  if (o.precision != BDF_PREC_FP64 && o.precision != BDF_PREC_MIXED)
    return fail(BDF_EINVAL, "bdf_create: unknown precision mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(BDF_EDEVICE, "bdf_create: no HIP device");
  if (o.device < 0 || o.device >= ndev) return fail(BDF_EINVAL, "bdf_create: bad device ordinal");
  if (o.schwarz < 0 || o.schwarz > 2) return fail(BDF_EINVAL, "bdf_create: schwarz must be 0, 1 or 2");
  bdf_engine *e = new bdf_engine();
  e->device = o.device;
  e->isym = sys->shape;
  e->intsym = o.intsym ? 1 : 0;
  e->racah1 = o.racah1 ? 1 : 0;
  e->schwarz = o.schwarz;
  e->sens = o.sens > 0.0 ? o.sens : 1.0e-12;
  e->mixed = o.precision == BDF_PREC_MIXED;
  e->breit = sys->hmlt == BDF_HMLT_DHFB;
  e->gaunt = o.gaunt ? 1 : 0;
  if (o.split_threshold > 0.0) e->split = o.split_threshold;
  std::string err = build_basis(*sys, e->B);
  if (!err.empty()) {
    delete e;
    return fail(BDF_ELIMIT, "bdf_create: " + err);
  }
  int rc = [&]() -> int {
    HIPCHK(hipSetDevice(e->device));
    // set-up of this device's static scratch (once per device)
    HIPCHK(static_scratch(e->device, &e->static_scratch));
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&e->ev0));
    HIPCHK(hipEventCreate(&e->ev1));
    HIPCHK(hipEventCreate(&e->ev2));
    if (e->mixed) {
      HIPCHK(hipStreamCreateWithFlags(&e->stream32, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&e->join, hipEventDisableTiming));
    }
    HIPCHK(upload_constants());
    build_etables(e->B, e->T, e->breit);
    HIPCHK(e->exps.upload(e->B.exps, e->stream));
    HIPCHK(e->normL.upload(e->B.normL, e->stream));
    HIPCHK(e->normS.upload(e->B.normS, e->stream));
    HIPCHK(e->pairs.upload(e->T.pairs, e->stream));
    HIPCHK(e->terms.upload(e->T.terms, e->stream));
    HIPCHK(e->monos.upload(e->T.monos, e->stream));
    std::vector<int64_t> first(e->T.pairs.size());
    int64_t acc = 0;
    for (size_t i = 0; i < e->T.pairs.size(); ++i) {
      first[i] = acc;
      acc += (int64_t)e->T.pairs[i].nA * e->T.pairs[i].nB;
    }
    HIPCHK(e->first.upload(first, e->stream));
    HIPCHK(e->flag_off.upload(e->T.flag_off, e->stream));
    HIPCHK(e->E.alloc((size_t)e->T.total));
    HIPCHK(e->flags.alloc((size_t)e->T.nflag));
    EtabArgs ea;
    ea.pairs = e->pairs.p;
    ea.npairs = (int)e->T.pairs.size();
    ea.first = e->first.p;
    ea.nitems = acc;
    ea.terms = e->terms.p;
    ea.monos = e->monos.p;
    ea.exps = e->exps.p;
    ea.normL = e->normL.p;
    ea.normS = e->normS.p;
    ea.E = e->E.p;
    HIPCHK(launch_etab(ea, e->stream));
    if (e->breit) {
      HIPCHK(e->Eab.alloc((size_t)e->T.total));
      HIPCHK(launch_eps_filter(e->E.p, e->Eab.p, e->T.total, e->stream));
    }
    EflagArgs fa;
    fa.pairs = e->pairs.p;
    fa.npairs = ea.npairs;
    fa.flag_off = e->flag_off.p;
    fa.E = e->E.p;
    fa.flags = e->flags.p;
    HIPCHK(e->Ecd.alloc((size_t)e->T.total));
    if (e->mixed) HIPCHK(e->Ecd32.alloc((size_t)e->T.total));
    fa.Ecd = e->Ecd.p;
    fa.Ecd32 = e->mixed ? e->Ecd32.p : nullptr;
    fa.Ecd1 = nullptr;
    fa.Ecd1_32 = nullptr;
    if (e->isym != 0) {
      HIPCHK(e->Ecd1.alloc((size_t)e->T.total));
      fa.Ecd1 = e->Ecd1.p;
      if (e->mixed) {
        HIPCHK(e->Ecd1_32.alloc((size_t)e->T.total));
        fa.Ecd1_32 = e->Ecd1_32.p;
      }
    }
    DevBuf<uint8_t> zf;
    fa.zflags = nullptr;
    const bool want_z = e->breit || e->isym != 0;
    if (want_z) {
      HIPCHK(zf.alloc((size_t)e->T.nflag));
      fa.zflags = zf.p;
    }
    HIPCHK(launch_eflag(fa, e->stream));
    e->eflags.resize(e->T.nflag);
    HIPCHK(hipMemcpyAsync(e->eflags.data(), e->flags.p, e->T.nflag, hipMemcpyDeviceToHost,
                          e->stream));
    if (want_z) {
      e->zflags.resize(e->T.nflag);
      HIPCHK(hipMemcpyAsync(e->zflags.data(), zf.p, e->T.nflag, hipMemcpyDeviceToHost,
                            e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
    }
    HIPCHK(e->icn.upload(e->B.labicn, e->stream));
    HIPCHK(e->kqn.upload(e->B.labkqn, e->stream));
    HIPCHK(e->mqn.upload(e->B.labmqn, e->stream));
    HIPCHK(e->counter.alloc(2));
    HIPCHK(e->dmax.alloc(1));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, e->device));
    e->cus = prop.multiProcessorCount;
    e->user_blocks = o.workgroups > 0;
    e->nblocks = o.workgroups;
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->mixed) {
      HIPCHK(e->E32.alloc((size_t)e->T.total));
      HIPCHK(launch_to_f32(e->E.p, e->E32.p, e->T.total, e->stream));
    }
    if (e->schwarz || e->mixed) return make_tables(e);
    return BDF_OK;
  }();
  if (rc != BDF_OK) {
    bdf_destroy(e);
    return rc;
  }
  *out = e;
  if (ndim) *ndim = e->B.ndim;
  if (nskp) *nskp = e->B.nskp;
  return BDF_OK;
}

int bdf_coulomb_device(bdf_engine *e, int ilev, const void *d_dent, void *d_gdir, void *d_gxch) {
  if (!e || !d_dent || !d_gdir || !d_gxch) return fail(BDF_EINVAL, "bdf_coulomb_device: null argument");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_coulomb_device: ilev outside 1..3");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_plan(e, ilev);
  if (rc != BDF_OK) return rc;
  return run_fock(e, ilev, (const double2 *)d_dent, (double2 *)d_gdir, (double2 *)d_gxch, nullptr,
                  nullptr);
}

int bdf_coulomb_device_static(bdf_engine *e, int ilev, const void *d_dent, void *d_gdir,
                              void *d_gxch) {
  if (!e || !d_dent || !d_gdir || !d_gxch)
    return fail(BDF_EINVAL, "bdf_coulomb_device_static: null argument");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_coulomb_device_static: ilev outside 1..3");
  DevPlan &dp = e->plans[ilev - 1];
  if (!dp.built)
    return fail(BDF_EINVAL, "bdf_coulomb_device_static: the stage's plan is not built "
                            "(bdf_plan_info or bdf_coulomb_device builds it)");
  if (!e->static_scratch) return fail(BDF_EDEVICE, "bdf_coulomb_device_static: no scratch");
  HIPCHK(hipSetDevice(e->device));
  // this plan's own per-workgroup rows at its kl row length; the mixed
  // build's FP32 instance takes a second set of workgroups
  const int64_t ld = ((std::max(dp.P.max_maxcd, 1) + 63) / 64) * 64;
  const int64_t rows = ntuv(kMaxLfl) + kMaxLfl + 1 + 4 * dp.fl.cap * ntuv(kMaxLab);
  const int64_t stride = rows * ld;
  const int64_t sets = e->mixed ? 2 : 1;
  const int64_t fit = (int64_t)kStaticScratchDoubles / (stride * sets);
  if (fit < 1)
    return fail(BDF_ELIMIT, "bdf_coulomb_device_static: one workgroup's scratch (" +
                                std::to_string(stride * sets * 8) +
                                " bytes) exceeds the static scratch (" +
                                std::to_string(kStaticScratchDoubles * 8) + " bytes)");
  const SlabView sv{e->static_scratch, stride, (int)std::min<int64_t>(dp.blocks, fit)};
  e->last_quartets = dp.P.quartets;
  const bool mixed = e->mixed != 0;
  int rc = enqueue_fock(e, dp, (const double2 *)d_dent, (double2 *)d_gdir, (double2 *)d_gxch,
                        nullptr, mixed, true, sv);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  return BDF_OK;
}

int bdf_coulomb(bdf_engine *e, int ilev, const double _Complex *dent, const double _Complex *deno,
                int ld, double _Complex *gdir, double _Complex *gxch, double _Complex *qdir,
                double _Complex *qxch) {
  if (!e || !dent || !gdir || !gxch) return fail(BDF_EINVAL, "bdf_coulomb: null argument");
  if (deno || qdir || qxch) return fail(BDF_EUNSUPP, "bdf_coulomb: open shells are not supported");
  int nd = e->B.ndim;
  if (ld < nd) return fail(BDF_EINVAL, "bdf_coulomb: ld < ndim");
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(e->dent.alloc((size_t)nd * nd));
  HIPCHK(e->gdir.alloc((size_t)nd * nd));
  HIPCHK(e->gxch.alloc((size_t)nd * nd));
  HIPCHK(hipMemcpy2DAsync(e->dent.p, nd * sizeof(double2), dent, ld * sizeof(double2),
                          nd * sizeof(double2), nd, hipMemcpyHostToDevice, e->stream));
  int rc = bdf_coulomb_device(e, ilev, e->dent.p, e->gdir.p, e->gxch.p);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipMemcpy2DAsync(gdir, ld * sizeof(double2), e->gdir.p, nd * sizeof(double2),
                          nd * sizeof(double2), nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpy2DAsync(gxch, ld * sizeof(double2), e->gxch.p, nd * sizeof(double2),
                          nd * sizeof(double2), nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

// one-centre Racah build of the Coulomb (breit 0) or Breit (1) interaction
static int racah_device(bdf_engine *e, int breit, int exact_beta, const void *d_dent,
                        void *d_gdir, void *d_gxch, const char *name) {
  if (!e || !d_dent || !d_gxch || (!breit && !d_gdir))
    return fail(BDF_EINVAL, std::string(name) + ": null argument");
  if (exact_beta != 0 && exact_beta != 1)
    return fail(BDF_EINVAL, std::string(name) + ": exact_beta must be 0 or 1");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_racah(e, breit);
  if (rc != BDF_OK) return rc;
  const DevRacah &r = e->racah[breit];
  RacahArgs a;
  a.blocks = r.blocks.p;
  a.segs = r.segs.p;
  a.combos = r.combos.p;
  a.entries = r.entries.p;
  a.ops = r.ops.p;
  a.items = r.items.p;
  a.ang = r.ang.p;
  a.nitems = (int)r.P.items.size();
  a.exps = e->exps.p;
  a.normL = e->normL.p;
  a.normS = e->normS.p;
  a.dent = (const double2 *)d_dent;
  a.gdir = (double2 *)d_gdir;
  a.gxch = (double2 *)d_gxch;
  a.ndim = e->B.ndim;
  a.isym = e->isym;
  a.exact = exact_beta;
  a.breit = breit;
  HIPCHK(hipEventRecord(e->ev0, e->stream));
  HIPCHK(launch_racah(a, e->stream));
  HIPCHK(hipEventRecord(e->ev1, e->stream));
  return BDF_OK;
}

int bdf_coulomb1_device(bdf_engine *e, int exact_beta, const void *d_dent, void *d_gdir,
                        void *d_gxch) {
  return racah_device(e, 0, exact_beta, d_dent, d_gdir, d_gxch, "bdf_coulomb1_device");
}

int bdf_breit1_device(bdf_engine *e, int exact_beta, const void *d_dent, void *d_bdir,
                      void *d_bxch) {
  return racah_device(e, 1, exact_beta, d_dent, d_bdir, d_bxch, "bdf_breit1_device");
}

int bdf_coulomb1(bdf_engine *e, int exact_beta, const double _Complex *dent,
                 const double _Complex *deno, int ld, double _Complex *gdir, double _Complex *gxch,
                 double _Complex *qdir, double _Complex *qxch) {
  if (!e || !dent || !gdir || !gxch) return fail(BDF_EINVAL, "bdf_coulomb1: null argument");
  if (deno || qdir || qxch) return fail(BDF_EUNSUPP, "bdf_coulomb1: open shells are not supported");
  int nd = e->B.ndim;
  if (ld < nd) return fail(BDF_EINVAL, "bdf_coulomb1: ld < ndim");
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(e->dent.alloc((size_t)nd * nd));
  HIPCHK(e->gdir.alloc((size_t)nd * nd));
  HIPCHK(e->gxch.alloc((size_t)nd * nd));
  const size_t row = nd * sizeof(double2), lrow = ld * sizeof(double2);
  HIPCHK(hipMemcpy2DAsync(e->dent.p, row, dent, lrow, row, nd, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpy2DAsync(e->gdir.p, row, gdir, lrow, row, nd, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpy2DAsync(e->gxch.p, row, gxch, lrow, row, nd, hipMemcpyHostToDevice, e->stream));
  int rc = bdf_coulomb1_device(e, exact_beta, e->dent.p, e->gdir.p, e->gxch.p);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipMemcpy2DAsync(gdir, lrow, e->gdir.p, row, row, nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpy2DAsync(gxch, lrow, e->gxch.p, row, row, nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_breit1(bdf_engine *e, int exact_beta, const double _Complex *dent,
               const double _Complex *deno, int ld, double _Complex *bdir, double _Complex *bxch,
               double _Complex *wdir, double _Complex *wxch) {
  if (!e || !dent || !bdir || !bxch) return fail(BDF_EINVAL, "bdf_breit1: null argument");
  if (deno || wdir || wxch) return fail(BDF_EUNSUPP, "bdf_breit1: open shells are not supported");
  int nd = e->B.ndim;
  if (ld < nd) return fail(BDF_EINVAL, "bdf_breit1: ld < ndim");
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(e->dent.alloc((size_t)nd * nd));
  HIPCHK(e->gxch.alloc((size_t)nd * nd));
  const size_t row = nd * sizeof(double2), lrow = ld * sizeof(double2);
  HIPCHK(hipMemcpy2DAsync(e->dent.p, row, dent, lrow, row, nd, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipMemcpy2DAsync(e->gxch.p, row, bxch, lrow, row, nd, hipMemcpyHostToDevice, e->stream));
  int rc = bdf_breit1_device(e, exact_beta, e->dent.p, nullptr, e->gxch.p);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipMemcpy2DAsync(bxch, lrow, e->gxch.p, row, row, nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_racah_info(bdf_engine *e, int breit, double *terms, long *items, long *uncovered) {
  if (!e) return fail(BDF_EINVAL, "bdf_racah_info: null engine");
  if (breit != 0 && breit != 1) return fail(BDF_EINVAL, "bdf_racah_info: breit must be 0 or 1");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_racah(e, breit);
  if (rc != BDF_OK) return rc;
  const RacahPlan &P = e->racah[breit].P;
  if (terms) *terms = P.terms;
  if (items) *items = (long)P.items.size();
  if (uncovered) *uncovered = (long)P.uncovered;
  return BDF_OK;
}

int bdf_eri_batch(bdf_engine *e, const int icnt[4], const int kidx[4], const int midx[4], int ibas,
                  int jbas, const int itn[2], int ilev, double _Complex *rr, int maxcd) {
  if (!e || !icnt || !kidx || !midx || !itn || !rr) return fail(BDF_EINVAL, "bdf_eri_batch: null argument");
  const Basis &B = e->B;
  int lq[4], kq[4], ic[4];
  for (int k = 0; k < 4; ++k) {
    ic[k] = icnt[k] - 1;
    if (ic[k] < 0 || ic[k] >= B.ncnt) return fail(BDF_EINVAL, "bdf_eri_batch: bad centre");
    if (kidx[k] < 1 || kidx[k] > 2 * B.lmax[ic[k]] + 1) return fail(BDF_EINVAL, "bdf_eri_batch: bad kappa index");
    kq[k] = B.kappa_of(ic[k], kidx[k]);
    lq[k] = kq[k] < 0 ? -kq[k] - 1 : kq[k];
    if (midx[k] < 1 || midx[k] > std::abs(kq[k])) return fail(BDF_EINVAL, "bdf_eri_batch: bad |mj| index");
  }
  if ((itn[0] != 1 && itn[0] != 4) || (itn[1] != 1 && itn[1] != 4))
    return fail(BDF_EINVAL, "bdf_eri_batch: itn must be 1 or 4");
  int nA = B.nfn(ic[0], lq[0]), nBb = B.nfn(ic[1], lq[1]);
  if (ibas < 1 || ibas > nA || jbas < 1 || jbas > nBb) return fail(BDF_EINVAL, "bdf_eri_batch: bad ibas/jbas");
  int mcd = B.nfn(ic[2], lq[2]) * B.nfn(ic[3], lq[3]);
  if (maxcd < mcd) return fail(BDF_EINVAL, "bdf_eri_batch: rr too small");
  HIPCHK(hipSetDevice(e->device));
  // a one-item plan: one class, one AB combination, one call
  DevPlan dp;
  ClassDesc cl;
  for (int k = 0; k < 4; ++k) {
    for (int x = 0; x < 3; ++x) cl.xyz[k][x] = B.xyz[3 * ic[k] + x];
    cl.exp[k] = B.expo(ic[k], lq[k]);
    cl.nbas[k] = B.nfn(ic[k], lq[k]);
  }
  int mcnt = 1;
  if (ic[1] != ic[0]) ++mcnt;
  if (ic[2] != ic[0] && ic[2] != ic[1]) ++mcnt;
  if (ic[3] != ic[0] && ic[3] != ic[1] && ic[3] != ic[2]) ++mcnt;
  int lsum = lq[0] + lq[1] + lq[2] + lq[3];
  cl.lamfl = mcnt == 1 ? lsum + 4 : (ilev == 1 ? lsum : (ilev == 2 ? lsum + 2 : lsum + 4));
  cl.maxcd = mcd;
  cl.mcnt = mcnt;
  cl.flagoff = 0;
  cl.sym = axis_symmetry(cl);
  cl.pad = 0;
  cl.ab_begin = 0;
  cl.ab_end = 1;
  ABCombo ab;
  int pab = e->T.pair_index.at(Etabs::key(ic[0], ic[1], kidx[0], kidx[1], midx[0], midx[1], itn[0] == 4));
  int pcd = e->T.pair_index.at(Etabs::key(ic[2], ic[3], kidx[2], kidx[3], midx[2], midx[3], itn[1] == 4));
  ab.eoff = e->T.pairs[pab].off;
  ab.ntab = ntuv(lq[0] + lq[1] + (itn[0] == 4 ? 2 : 0));
  int ntcd = ntuv(lq[2] + lq[3] + (itn[1] == 4 ? 2 : 0));
  ab.ntg = ntcd;
  for (int s = 0; s < 2; ++s) ab.fa[s] = ab.fb[s] = 0;
  ab.nb = cl.nbas[1];
  ab.cd_begin = 0;
  ab.cd_end = 1;
  ab.lab = lq[0] + lq[1] + (itn[0] == 4 ? 2 : 0);
  // G mask: aggregated column flags of the (C,D,lc,ld) class for this itn
  std::vector<int> agg(4 * ntcd, 0);
  int lc = lq[2], ld = lq[3];
  for (int nc = (lc == 0); nc <= 1; ++nc)
    for (int nd = (ld == 0); nd <= 1; ++nd)
      for (int mc = 1; mc <= lc + nc; ++mc)
        for (int md = 1; md <= ld + nd; ++md) {
          int pr = e->T.pair_index.at(Etabs::key(ic[2], ic[3], 2 * lc + nc, 2 * ld + nd, mc, md, itn[1] == 4));
          for (int t = 0; t < ntcd; ++t)
            for (int c = 0; c < 4; ++c) agg[4 * t + c] += e->eflags[e->T.flag_off[pr] + 4 * t + c];
        }
  ab.gmask = 0;
  for (int t = 0; t < ntcd; ++t) {
    int all = agg[4 * t] + agg[4 * t + 1] + agg[4 * t + 2] + agg[4 * t + 3];
    for (int c = 0; c < 4; ++c)
      dp.P.masks.push_back((uint8_t)(all != 0 && (e->isym <= 1 || agg[4 * t + c] != 0)));
  }
  CDCall cd;
  cd.eoff = e->T.pairs[pcd].off;
  cd.ntcd = ntcd;
  for (int s = 0; s < 2; ++s) cd.fc[s] = cd.fd[s] = 0;
  cd.iflg = 0;
  int mq[4];
  for (int k = 0; k < 4; ++k) mq[k] = 2 * midx[k] - 1;
  int sab = (kq[0] * kq[1]) >= 0 ? 1 : -1, scd = (kq[2] * kq[3]) >= 0 ? 1 : -1;
  cd.pab[0] = (int8_t)(sab * ((std::abs(mq[1] - mq[0]) / 2) % 2 ? -1 : 1));
  cd.pab[1] = (int8_t)(sab * (((mq[0] + mq[1]) / 2) % 2 ? -1 : 1));
  cd.pcd[0] = (int8_t)(scd * ((std::abs(mq[3] - mq[2]) / 2) % 2 ? -1 : 1));
  cd.pcd[1] = (int8_t)(scd * (((mq[2] + mq[3]) / 2) % 2 ? -1 : 1));
  dp.P.classes.push_back(cl);
  dp.P.abs.push_back(ab);
  dp.P.cds.push_back(cd);
  dp.P.items.push_back({0, (ibas - 1) * cl.nbas[1] + (jbas - 1), cl.ab_begin, cl.ab_end});
  dp.P.max_maxcd = mcd;
  HIPCHK(dp.classes.upload(dp.P.classes, e->stream));
  HIPCHK(dp.abs.upload(dp.P.abs, e->stream));
  HIPCHK(dp.cds.upload(dp.P.cds, e->stream));
  HIPCHK(dp.masks.upload(dp.P.masks, e->stream));
  HIPCHK(dp.items.upload(dp.P.items, e->stream));
  dp.P.ilev = ilev;
  dp.xacc_n = plan_xacc(dp.P, e->isym);
  dp.fl = fock_config(e, dp);
  dp.blocks = 1;
  int src = ensure_slab(e, 0, mcd, 4 * dp.fl.cap * ntuv(kMaxLab), 1);
  if (src != BDF_OK) return src;
  DevBuf<double2> out;
  HIPCHK(out.alloc((size_t)16 * mcd));
  int rc = run_fock(e, ilev, nullptr, nullptr, nullptr, out.p, &dp);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipMemcpyAsync(rr, out.p, sizeof(double2) * 16 * mcd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_etable(bdf_engine *e, int itt, int ica, int icb, int ka, int kb, int ma, int mb, double *out,
               long cap, long *n) {
  if (!e || !out || !n) return fail(BDF_EINVAL, "bdf_etable: null argument");
  if (itt == 2) {
    if (!e->breit) return fail(BDF_EINVAL, "bdf_etable: LS tables need a DHFB engine");
    long cnt = -1;
    HIPCHK(hipSetDevice(e->device));
    for (int ix = 0; ix < 3; ++ix) {
      auto it = e->T.pair_index.find(Etabs::key(ica - 1, icb - 1, ka, kb, ma, mb, kLSx + ix));
      if (it == e->T.pair_index.end()) return fail(BDF_EINVAL, "bdf_etable: no such block");
      const EPair &P = e->T.pairs[it->second];
      cnt = (long)ntuv(P.lam) * P.nA * P.nB;
      if (cap < 12 * cnt) return fail(BDF_EINVAL, "bdf_etable: out too small");
      std::vector<double4> h(cnt);
      HIPCHK(hipMemcpy(h.data(), e->E.p + P.off, sizeof(double4) * cnt, hipMemcpyDeviceToHost));
      for (long k = 0; k < cnt; ++k) {
        out[(4 * ix + 0) * cnt + k] = h[k].x;
        out[(4 * ix + 1) * cnt + k] = h[k].y;
        out[(4 * ix + 2) * cnt + k] = h[k].z;
        out[(4 * ix + 3) * cnt + k] = h[k].w;
      }
    }
    *n = cnt;
    return BDF_OK;
  }
  auto it = e->T.pair_index.find(Etabs::key(ica - 1, icb - 1, ka, kb, ma, mb, itt == 4));
  if (it == e->T.pair_index.end()) return fail(BDF_EINVAL, "bdf_etable: no such block");
  const EPair &P = e->T.pairs[it->second];
  long cnt = (long)ntuv(P.lam) * P.nA * P.nB;
  *n = cnt;
  if (cap < 4 * cnt) return fail(BDF_EINVAL, "bdf_etable: out too small");
  std::vector<double4> h(cnt);
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(hipMemcpy(h.data(), e->E.p + P.off, sizeof(double4) * cnt, hipMemcpyDeviceToHost));
  for (long k = 0; k < cnt; ++k) {
    out[k] = h[k].x;
    out[cnt + k] = h[k].y;
    out[2 * cnt + k] = h[k].z;
    out[3 * cnt + k] = h[k].w;
  }
  return BDF_OK;
}

int bdf_schwarz_tables(bdf_engine *e, double *gdsc, double *qbnd, int ld) {
  if (!e) return fail(BDF_EINVAL, "bdf_schwarz_tables: null engine");
  int nd = e->B.ndim;
  if (ld < nd) return fail(BDF_EINVAL, "bdf_schwarz_tables: ld < ndim");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_tables(e);
  if (rc != BDF_OK) return rc;
  if (gdsc)
    HIPCHK(hipMemcpy2D(gdsc, ld * sizeof(double), e->gdsc.p, nd * sizeof(double),
                       nd * sizeof(double), nd, hipMemcpyDeviceToHost));
  if (qbnd)
    HIPCHK(hipMemcpy2D(qbnd, ld * sizeof(double), e->qbnd.p, nd * sizeof(double),
                       nd * sizeof(double), nd, hipMemcpyDeviceToHost));
  return BDF_OK;
}

int bdf_breit_device(bdf_engine *e, int ilev, const void *d_dent, void *d_bdir, void *d_bxch) {
  if (!e || !d_dent || !d_bdir || !d_bxch) return fail(BDF_EINVAL, "bdf_breit_device: null argument");
  if (!e->breit) return fail(BDF_EINVAL, "bdf_breit_device: engine built without DHFB");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_breit_device: ilev outside 1..3");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_breit_plan(e, ilev);
  if (rc != BDF_OK) return rc;
  return run_breit(e, e->bplans[ilev - 1], ilev, (const double2 *)d_dent, (double2 *)d_bdir,
                   (double2 *)d_bxch, nullptr);
}

int bdf_breit(bdf_engine *e, int ilev, const double _Complex *dent, const double _Complex *deno,
              int ld, double _Complex *bdir, double _Complex *bxch, double _Complex *wdir,
              double _Complex *wxch) {
  if (!e || !dent || !bdir || !bxch) return fail(BDF_EINVAL, "bdf_breit: null argument");
  if (deno || wdir || wxch) return fail(BDF_EUNSUPP, "bdf_breit: open shells are not supported");
  int nd = e->B.ndim;
  if (ld < nd) return fail(BDF_EINVAL, "bdf_breit: ld < ndim");
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(e->dent.alloc((size_t)nd * nd));
  HIPCHK(e->gdir.alloc((size_t)nd * nd));
  HIPCHK(e->gxch.alloc((size_t)nd * nd));
  HIPCHK(hipMemcpy2DAsync(e->dent.p, nd * sizeof(double2), dent, ld * sizeof(double2),
                          nd * sizeof(double2), nd, hipMemcpyHostToDevice, e->stream));
  int rc = bdf_breit_device(e, ilev, e->dent.p, e->gdir.p, e->gxch.p);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipMemcpy2DAsync(bdir, ld * sizeof(double2), e->gdir.p, nd * sizeof(double2),
                          nd * sizeof(double2), nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpy2DAsync(bxch, ld * sizeof(double2), e->gxch.p, nd * sizeof(double2),
                          nd * sizeof(double2), nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_bii_batch(bdf_engine *e, const int icnt[4], const int kidx[4], const int midx[4], int ibas,
                  int jbas, double _Complex *rr, int maxcd) {
  if (!e || !icnt || !kidx || !midx || !rr) return fail(BDF_EINVAL, "bdf_bii_batch: null argument");
  if (!e->breit) return fail(BDF_EINVAL, "bdf_bii_batch: engine built without DHFB");
  const Basis &B = e->B;
  int lq[4], kq[4], ic[4];
  for (int k = 0; k < 4; ++k) {
    ic[k] = icnt[k] - 1;
    if (ic[k] < 0 || ic[k] >= B.ncnt) return fail(BDF_EINVAL, "bdf_bii_batch: bad centre");
    if (kidx[k] < 1 || kidx[k] > 2 * B.lmax[ic[k]] + 1) return fail(BDF_EINVAL, "bdf_bii_batch: bad kappa index");
    kq[k] = B.kappa_of(ic[k], kidx[k]);
    lq[k] = kq[k] < 0 ? -kq[k] - 1 : kq[k];
    if (midx[k] < 1 || midx[k] > std::abs(kq[k])) return fail(BDF_EINVAL, "bdf_bii_batch: bad |mj| index");
  }
  int nA = B.nfn(ic[0], lq[0]), nBb = B.nfn(ic[1], lq[1]);
  if (ibas < 1 || ibas > nA || jbas < 1 || jbas > nBb) return fail(BDF_EINVAL, "bdf_bii_batch: bad ibas/jbas");
  int mcd = B.nfn(ic[2], lq[2]) * B.nfn(ic[3], lq[3]);
  if (maxcd < mcd) return fail(BDF_EINVAL, "bdf_bii_batch: rr too small");
  HIPCHK(hipSetDevice(e->device));
  // a one-item plan: one class, one AB combination (the class's first, so
  // the gauge term sees this pair's own R flags), one call
  DevBreitPlan dp;
  ClassDesc cl;
  for (int k = 0; k < 4; ++k) {
    for (int x = 0; x < 3; ++x) cl.xyz[k][x] = B.xyz[3 * ic[k] + x];
    cl.exp[k] = B.expo(ic[k], lq[k]);
    cl.nbas[k] = B.nfn(ic[k], lq[k]);
  }
  int lamab = lq[0] + lq[1] + 1, lamcd = lq[2] + lq[3] + 1;
  cl.lamfl = e->gaunt ? lamab + lamcd : lamab + lamcd + 2;
  cl.maxcd = mcd;
  cl.mcnt = 0;
  cl.flagoff = 0;
  cl.sym = axis_symmetry(cl);
  cl.pad = 0;
  cl.ab_begin = 0;
  cl.ab_end = 1;
  int ntcd = ntuv(lamcd);
  try {
    BrAB ab;
    for (int ix = 0; ix < 3; ++ix)
      ab.eoff[ix] = e->T.pairs[e->T.pair_index.at(Etabs::key(ic[0], ic[1], kidx[0], kidx[1], midx[0], midx[1], kLSx + ix))].off;
    ab.ntab = ntuv(lamab);
    ab.ntg = ntcd;
    for (int s = 0; s < 2; ++s) ab.fa[s] = ab.fb[s] = 0;
    ab.cd_begin = 0;
    ab.cd_end = 1;
    ab.pad = 0;
    // G mask: E(CD) flags summed over the (C,D,lc,ld) class (EILSCD0)
    std::vector<int> agg(12 * ntcd, 0);
    int lc = lq[2], ld = lq[3];
    for (int nc = (lc == 0); nc <= 1; ++nc)
      for (int ndd = (ld == 0); ndd <= 1; ++ndd)
        for (int mc = 1; mc <= lc + nc; ++mc)
          for (int md = 1; md <= ld + ndd; ++md)
            for (int ix = 0; ix < 3; ++ix) {
              int pr = e->T.pair_index.at(Etabs::key(ic[2], ic[3], 2 * lc + nc, 2 * ld + ndd, mc, md, kLSx + ix));
              for (int t = 0; t < ntcd; ++t)
                for (int c = 0; c < 4; ++c)
                  agg[(t * 3 + ix) * 4 + c] += e->eflags[e->T.flag_off[pr] + 4 * t + c];
            }
    ab.gmask = 0;
    for (int t = 0; t < ntcd; ++t)
      for (int ix = 0; ix < 3; ++ix) {
        const int *g = &agg[(t * 3 + ix) * 4];
        int all = g[0] + g[1] + g[2] + g[3];
        for (int c = 0; c < 4; ++c)
          dp.P.masks.push_back((uint8_t)(all != 0 && (e->isym == 0 || g[c] != 0)));
      }
    BrCD cd;
    cd.ntcd = ntcd;
    for (int ix = 0; ix < 3; ++ix)
      cd.eoff[ix] = e->T.pairs[e->T.pair_index.at(Etabs::key(ic[2], ic[3], kidx[2], kidx[3], midx[2], midx[3], kLSx + ix))].off;
    for (int s = 0; s < 2; ++s) cd.fc[s] = cd.fd[s] = 0;
    cd.iflg = 0;
    int mq[4];
    for (int k = 0; k < 4; ++k) mq[k] = 2 * midx[k] - 1;
    int sab = (kq[0] * kq[1]) >= 0 ? 1 : -1, scd = (kq[2] * kq[3]) >= 0 ? 1 : -1;
    cd.pab[0] = (int8_t)(-sab * ((std::abs(mq[1] - mq[0]) / 2) % 2 ? -1 : 1));
    cd.pab[1] = (int8_t)(-sab * (((mq[0] + mq[1]) / 2) % 2 ? -1 : 1));
    cd.pcd[0] = (int8_t)(-scd * ((std::abs(mq[3] - mq[2]) / 2) % 2 ? -1 : 1));
    cd.pcd[1] = (int8_t)(-scd * (((mq[2] + mq[3]) / 2) % 2 ? -1 : 1));
    dp.P.abs.push_back(ab);
    dp.P.cds.push_back(cd);
  } catch (const std::out_of_range &) {
    return fail(BDF_EINVAL, "bdf_bii_batch: missing E-coefficient block");
  }
  dp.P.classes.push_back(cl);
  dp.P.items.push_back({0, (ibas - 1) * cl.nbas[1] + (jbas - 1), cl.ab_begin, cl.ab_end});
  dp.P.nflag = ntuv(cl.lamfl);
  dp.P.max_maxcd = mcd;
  dp.P.max_ntg = ntcd;
  HIPCHK(dp.classes.upload(dp.P.classes, e->stream));
  HIPCHK(dp.abs.upload(dp.P.abs, e->stream));
  HIPCHK(dp.cds.upload(dp.P.cds, e->stream));
  HIPCHK(dp.masks.upload(dp.P.masks, e->stream));
  HIPCHK(dp.items.upload(dp.P.items, e->stream));
  HIPCHK(dp.stale.alloc((size_t)dp.P.nflag));
  int src = ensure_slab(e, 1, mcd, 12 * ntcd, 1);
  if (src != BDF_OK) return src;
  DevBuf<double2> out;
  HIPCHK(out.alloc((size_t)16 * mcd));
  int rc = run_breit(e, dp, 3, nullptr, nullptr, nullptr, out.p);
  if (rc != BDF_OK) return rc;
  HIPCHK(hipMemcpyAsync(rr, out.p, sizeof(double2) * 16 * mcd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_one_electron_device(bdf_engine *e, const double *znuc, const double *xi, double cv,
                            void *d_ovlp, void *d_hnuc, void *d_hkin) {
  if (!e || !znuc || !d_ovlp || !d_hnuc || !d_hkin)
    return fail(BDF_EINVAL, "bdf_one_electron_device: null argument");
  if (!(cv > 0.0)) return fail(BDF_EINVAL, "bdf_one_electron_device: cv must be positive");
  const Basis &B = e->B;
  const Etabs &T = e->T;
  std::vector<OneJob> jobs;
  std::vector<int64_t> first;
  int64_t acc = 0;
  try {
    for (int a = 0; a < B.ncnt; ++a)
      for (int b = 0; b <= a; ++b)
        for (int ka = 1; ka <= 2 * B.lmax[a] + 1; ++ka)
          for (int kb = 1; kb <= 2 * B.lmax[b] + 1; ++kb) {
            int kqa = B.kappa_of(a, ka), kqb = B.kappa_of(b, kb);
            int la = kqa < 0 ? -kqa - 1 : kqa, lb = kqb < 0 ? -kqb - 1 : kqb;
            for (int ma = 1; ma <= std::abs(kqa); ++ma)
              for (int mb = 1; mb <= std::abs(kqb); ++mb) {
                OneJob J;
                J.il[0] = B.lrge_at(a, ka, 2 * ma - 1);
                J.il[1] = B.lrge_at(a, ka, 2 * ma);
                J.jl[0] = B.lrge_at(b, kb, 2 * mb - 1);
                J.jl[1] = B.lrge_at(b, kb, 2 * mb);
                if (J.il[0] < J.jl[0]) continue;  // written as the mirror of (B,A)
                J.ell = T.pairs[T.pair_index.at(Etabs::key(a, b, ka, kb, ma, mb, kLL))].off;
                J.ess = T.pairs[T.pair_index.at(Etabs::key(a, b, ka, kb, ma, mb, kSS))].off;
                J.essba = T.pairs[T.pair_index.at(Etabs::key(b, a, kb, ka, mb, ma, kSS))].off;
                J.nA = B.nfn(a, la);
                J.nB = B.nfn(b, lb);
                J.expA = B.expo(a, la);
                J.expB = B.expo(b, lb);
                J.la = la;
                J.lb = lb;
                int mq = (2 * ma - 1 - (2 * mb - 1)) / 2;
                J.phs = ((mq % 2) ? -1.0 : 1.0) * ((kqa * kqb) > 0 ? 1.0 : -1.0);
                J.ax = B.xyz[3 * a]; J.ay = B.xyz[3 * a + 1]; J.az = B.xyz[3 * a + 2];
                J.bx = B.xyz[3 * b]; J.by = B.xyz[3 * b + 1]; J.bz = B.xyz[3 * b + 2];
                first.push_back(acc);
                acc += (int64_t)J.nA * J.nB;
                jobs.push_back(J);
              }
          }
  } catch (const std::out_of_range &) {
    return fail(BDF_EINVAL, "bdf_one_electron: missing E-coefficient block");
  }
  HIPCHK(hipSetDevice(e->device));
  int nd = B.ndim;
  size_t nn = (size_t)nd * nd;
  DevBuf<OneJob> djobs;
  DevBuf<int64_t> dfirst;
  DevBuf<double> dz, dxi, dxyz;
  std::vector<double> z(znuc, znuc + B.ncnt), x(B.ncnt, 0.0);
  if (xi) x.assign(xi, xi + B.ncnt);
  HIPCHK(djobs.upload(jobs, e->stream));
  HIPCHK(dfirst.upload(first, e->stream));
  HIPCHK(dz.upload(z, e->stream));
  HIPCHK(dxi.upload(x, e->stream));
  HIPCHK(dxyz.upload(B.xyz, e->stream));
  HIPCHK(launch_zero((double2 *)d_ovlp, (double2 *)d_hnuc, (int64_t)nn, nullptr, nullptr, e->stream));
  HIPCHK(launch_zero((double2 *)d_hkin, nullptr, (int64_t)nn, nullptr, nullptr, e->stream));
  OneArgs oa;
  oa.jobs = djobs.p;
  oa.njobs = (int)jobs.size();
  oa.first = dfirst.p;
  oa.nitems = acc;
  oa.exps = e->exps.p;
  oa.E = e->E.p;
  oa.xyz = dxyz.p;
  oa.znuc = dz.p;
  oa.xi = dxi.p;
  oa.ncnt = B.ncnt;
  oa.cv = cv;
  oa.ovlp = (double2 *)d_ovlp;
  oa.hnuc = (double2 *)d_hnuc;
  oa.hkin = (double2 *)d_hkin;
  oa.ndim = nd;
  oa.nskp = B.nskp;
  HIPCHK(launch_oneel(oa, e->stream));
  // the job tables are freed when this call returns
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_one_electron(bdf_engine *e, const double *znuc, const double *xi, double cv,
                     double _Complex *ovlp, double _Complex *hnuc, double _Complex *hkin, int ld) {
  if (!e || !ovlp || !hnuc || !hkin) return fail(BDF_EINVAL, "bdf_one_electron: null argument");
  int nd = e->B.ndim;
  if (ld < nd) return fail(BDF_EINVAL, "bdf_one_electron: ld < ndim");
  HIPCHK(hipSetDevice(e->device));
  DevBuf<double2> o, h, k;
  HIPCHK(o.alloc((size_t)nd * nd));
  HIPCHK(h.alloc((size_t)nd * nd));
  HIPCHK(k.alloc((size_t)nd * nd));
  int rc = bdf_one_electron_device(e, znuc, xi, cv, o.p, h.p, k.p);
  if (rc != BDF_OK) return rc;
  // on the engine's stream, as every other host entry point copies
  const size_t row = nd * sizeof(double2), lrow = ld * sizeof(double2);
  HIPCHK(hipMemcpy2DAsync(ovlp, lrow, o.p, row, row, nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpy2DAsync(hnuc, lrow, h.p, row, row, nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipMemcpy2DAsync(hkin, lrow, k.p, row, row, nd, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return BDF_OK;
}

int bdf_breit_plan_info(bdf_engine *e, int ilev, double *flops, double *quartets, long *items) {
  if (!e) return fail(BDF_EINVAL, "bdf_breit_plan_info: null engine");
  if (!e->breit) return fail(BDF_EINVAL, "bdf_breit_plan_info: engine built without DHFB");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_breit_plan_info: ilev outside 1..3");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_breit_plan(e, ilev);
  if (rc != BDF_OK) return rc;
  const BreitPlan &P = e->bplans[ilev - 1].P;
  if (flops) *flops = P.flops;
  if (quartets) *quartets = P.quartets;
  if (items) *items = (long)P.items.size();
  return BDF_OK;
}

int bdf_memory_info(bdf_engine *e, size_t *device_bytes) {
  if (!e || !device_bytes) return fail(BDF_EINVAL, "bdf_memory_info: null argument");
  size_t b = 0;
  auto add = [&](size_t n, size_t w) { b += n * w; };
  add(e->E.n, sizeof(double4));
  add(e->E32.n, sizeof(float4));
  add(e->Ecd.n, sizeof(double4));
  add(e->Eab.n, sizeof(double4));
  add(e->Ecd32.n, sizeof(float4));
  add(e->Ecd1.n, sizeof(double));
  add(e->Ecd1_32.n, sizeof(float));
  add(e->flags.n, 1);
  add(e->slab.n, sizeof(double));
  add(e->gdsc.n + e->qbnd.n + e->gslot.n + e->dslot.n, sizeof(double));
  add(e->partner.n, sizeof(int));
  add(e->dent.n + e->gdir.n + e->gxch.n, sizeof(double2));
  add(e->pairs.n, sizeof(EPair));
  add(e->terms.n, sizeof(ETerm));
  add(e->monos.n, sizeof(Mono));
  for (const DevPlan &p : e->plans)
    b += p.classes.n * sizeof(ClassDesc) + p.abs.n * sizeof(ABCombo) + p.cds.n * sizeof(CDCall) +
         p.masks.n + p.items.n * sizeof(WorkItem) + p.qitem.n * sizeof(double);
  for (const DevBreitPlan &p : e->bplans)
    b += p.classes.n * sizeof(ClassDesc) + p.abs.n * sizeof(BrAB) + p.cds.n * sizeof(BrCD) +
         p.masks.n + p.items.n * sizeof(WorkItem) + p.stale.n;
  *device_bytes = b;
  return BDF_OK;
}

void *bdf_stream(bdf_engine *e) { return e ? (void *)e->stream : nullptr; }

int bdf_last_stats(bdf_engine *e, double *quartets, double *kernel_ms) {
  if (!e) return fail(BDF_EINVAL, "bdf_last_stats: null engine");
  HIPCHK(hipEventSynchronize(e->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev2));
  e->last_ms = ms;
  if (quartets) *quartets = e->last_quartets;
  if (kernel_ms) *kernel_ms = e->last_ms;
  return BDF_OK;
}

int bdf_plan_info(bdf_engine *e, int ilev, double *flops, double *quartets, long *items) {
  if (!e) return fail(BDF_EINVAL, "bdf_plan_info: null engine");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_plan_info: ilev outside 1..3");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_plan(e, ilev);
  if (rc != BDF_OK) return rc;
  const Plan &P = e->plans[ilev - 1].P;
  if (flops) *flops = P.flops;
  if (quartets) *quartets = P.quartets;
  if (items) *items = (long)P.items.size();
  return BDF_OK;
}

int bdf_launch_info(bdf_engine *e, int ilev, int *waves, int *cap, int *workgroups,
                    size_t *scratch_bytes) {
  if (!e) return fail(BDF_EINVAL, "bdf_launch_info: null engine");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_launch_info: ilev outside 1..3");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_plan(e, ilev);
  if (rc != BDF_OK) return rc;
  const DevPlan &dp = e->plans[ilev - 1];
  if (waves) *waves = dp.fl.waves;
  if (cap) *cap = dp.fl.cap;
  if (workgroups) *workgroups = dp.blocks;
  if (scratch_bytes) *scratch_bytes = e->slab.n * sizeof(double);
  return BDF_OK;
}

int bdf_breit_launch_info(bdf_engine *e, int ilev, int *kernel, int *workgroups) {
  if (!e) return fail(BDF_EINVAL, "bdf_breit_launch_info: null engine");
  if (!e->breit) return fail(BDF_EINVAL, "bdf_breit_launch_info: engine built without DHFB");
  if (ilev < 1 || ilev > 3) return fail(BDF_EINVAL, "bdf_breit_launch_info: ilev outside 1..3");
  HIPCHK(hipSetDevice(e->device));
  int rc = make_breit_plan(e, ilev);
  if (rc != BDF_OK) return rc;
  const BreitPlan &P = e->bplans[ilev - 1].P;
  const bool tc = breit_tc(e, P);
  if (kernel) *kernel = tc ? 2 : 1;
  if (workgroups) *workgroups = tc ? breit_tc_blocks(e, P) : breit_blocks(e);
  return BDF_OK;
}

int bdf_set_launch(bdf_engine *e, const bdf_launch *l) {
  if (!e || !l) return fail(BDF_EINVAL, "bdf_set_launch: null argument");
  if (l->waves != 0 && (l->waves < 2 || l->waves > 4))
    return fail(BDF_EINVAL, "bdf_set_launch: waves must be 0, 2, 3 or 4");
  if (l->cap < 0 || l->cap > 2) return fail(BDF_EINVAL, "bdf_set_launch: cap must be 0, 1 or 2");
  if (l->waves >= 3 && l->cap == 2)
    return fail(BDF_EINVAL, "bdf_set_launch: three or four waves run at capacity 1 only");
  if (l->per_cu < 0 || l->per_cu > 32 || l->breit_per_cu < 0 || l->breit_per_cu > 32)
    return fail(BDF_EINVAL, "bdf_set_launch: workgroups per CU outside 0..32");
  if (l->breit_group != 0 && l->breit_group != 1 && l->breit_group != 2 && l->breit_group != 4)
    return fail(BDF_EINVAL, "bdf_set_launch: breit_group must be 0, 1, 2 or 4");
  if (l->breit_kernel < 0 || l->breit_kernel > 2)
    return fail(BDF_EINVAL, "bdf_set_launch: breit_kernel must be 0, 1 or 2");
  HIPCHK(hipSetDevice(e->device));
  HIPCHK(hipStreamSynchronize(e->stream));
  drop_graphs(e);
  e->use_graph = l->graph ? 1 : 0;
  e->tune_waves = l->waves;
  e->tune_cap = l->cap;
  e->tune_per_cu = l->per_cu;
  e->tune_breit_per_cu = l->breit_per_cu;
  e->tune_breit_group = l->breit_group;
  e->tune_breit_kernel = l->breit_kernel;
  for (DevPlan &dp : e->plans)
    if (dp.built) {
      int rc = size_launch(e, dp);
      if (rc != BDF_OK) return rc;
    }
  for (DevBreitPlan &bp : e->bplans)
    if (bp.built) {
      if (e->tune_breit_kernel == 2 && !breit_tc(e, bp.P))
        return fail(BDF_EUNSUPP, "matrix-core Breit kernel: not a real-path system, or its LDS "
                                 "exceeds a compute unit");
      int rc = ensure_slab(e, 1, bp.P.max_maxcd,
                           12 * breit_group(e) * std::max(bp.P.max_ntg, 1), breit_blocks(e));
      if (rc == BDF_OK && breit_tc(e, bp.P))
        rc = ensure_slab(e, 1, bp.P.max_maxcd, 0, breit_tc_blocks(e, bp.P));
      if (rc != BDF_OK) return rc;
    }
  return BDF_OK;
}

void bdf_destroy(bdf_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->ev2) (void)hipEventDestroy(e->ev2);
  if (e->stream32) (void)hipStreamSynchronize(e->stream32);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->join) (void)hipEventDestroy(e->join);
  if (e->stream32) (void)hipStreamDestroy(e->stream32);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"
