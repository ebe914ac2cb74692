// kernels.hip -- CDNA4 (gfx950) kernels of the two-electron engine.
//
//   etab_kernel   spinor-pair Hermite E-coefficients (replaces EQSAVE's
//                 E0LLGN/E0SSGN, bertha-scf.f:33378, 33511)
//   eflag_kernel  E(CD) column masks (E0LLCD0/E0SSCD0, :17673, :17833)
//   fock_kernel   persistent work-queue kernel: Boys function and
//                 R-integrals (FUNFM/RMAKE, :32465/:31900), the two E.R.E
//                 contractions (ERIFAST, :12269) and J/K digestion
//                 (CLMMTZ/CLMMTG, :13135/:13841)
//   complete_kernel Hermitian completion (CLMFAST :12166-12205)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.h"

namespace bdf {

// Scratch of the allocation-free Coulomb build (bdf_coulomb_device_static):
// the per-workgroup R, Boys and G rows live here instead of in the engine's
// own slab.  Each GPU the library runs on holds one copy for the life of the
// process; every such build on that GPU shares it.
static __device__ __align__(256) double scratch[kStaticScratchDoubles];
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices];

hipError_t static_scratch(int device, double **ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    // the array's device address on the current device (the caller has made
    // `device` current); host code never passes the symbol itself
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<double *>(p);
  }
  *ptr = g_scratch_dev[device];
  return hipSuccess;
}

__constant__ uint32_t c_tuv[ntuv(kMaxLfl)];  // (t,u,v) packed 5:5:5 bits (32-bit words: scalar loads)
// R row of a sum of Hermite indices: c_hsum[tp * kHsRow + t] = hidx(tuv(t) + tuv(tp))
// for t, tp < N(kMaxLab); rows padded to a multiple of 4 for 16-byte scalar loads
constexpr int kHsRow = (ntuv(kMaxLab) + 3) / 4 * 4;
__constant__ int c_hsum[ntuv(kMaxLab) * kHsRow];
// unit steps in Hermite index space: c_hstep[h][d] = index(tuv(h) + e_d)
// (d = 0..2; -1 beyond order kMaxLfl) and c_hstep[h][4 + d] = index(tuv(h) -
// e_d) (-1 where that component is zero); rows of 8 words
__constant__ int c_hstep[ntuv(kMaxLfl)][8];
// Hermite indices t < N(kMaxLab) grouped by parity class under an axis mask
// m (ClassDesc::sym): class(t) = (t_x & 1 | (t_y & 1) << 1 | (t_z & 1) << 2) & m.
// For the bra order L, c_tperm[m][L][pos] lists the t < N(L) class by class,
// ascending inside a class, class c at positions [c_coff[m][L][c],
// c_coff[m][L][c + 1]); g_hsp[m][tp][k] = index(tuv(t_k) + tuv(tp)) for the
// k-th t of tp's own class.  With m = 0 these are the identity and c_hsum.
// R(tuv + tuv') vanishes exactly unless t and tp share a class, so the first
// contraction of tp runs over that class only, in ascending t as before.
__constant__ int16_t c_tperm[8][kMaxLab + 1][ntuv(kMaxLab)];
__constant__ uint8_t c_coff[8][kMaxLab + 1][9];
__device__ int g_hsp[8][ntuv(kMaxLab)][kHsRow];
__constant__ double c_rcp[2 * kMaxLfl + 2 * 60 + 4];  // 1/(2k+1)
__constant__ float c_rcpf[2 * kMaxLfl + 2 * 60 + 4];

__device__ __forceinline__ int hidx(int t, int u, int v) {
  int L = t + u + v;
  return L * (L + 1) * (L + 2) / 6 + t * (L + 1) - t * (t - 1) / 2 + u;
}
__device__ __forceinline__ void dec(int n, int &t, int &u, int &v) {
  uint32_t x = c_tuv[n];
  t = x >> 10;
  u = (x >> 5) & 31;
  v = x & 31;
}
// parity class of Hermite index n (bit d: the d-th component is odd)
__device__ __forceinline__ int hpar(int n) {
  const uint32_t x = c_tuv[n];
  return (int)(((x >> 10) & 1) | (((x >> 5) & 1) << 1) | ((x & 1) << 2));
}

// ---------------------------------------------------------------------------
// E-coefficients
// ---------------------------------------------------------------------------
// 1D Hermite expansion coefficients of x_A^i x_B^j exp(-p x_P^2):
// E[i+1][j][t] = E[i][j][t-1]/(2p) + X_PA E[i][j][t] + (t+1) E[i][j][t+1]
// (the same multiplication rule the reference applies in ESTEPL/ESTEPN).
struct Herm1 {
  double e[kMaxPoly + 1][kMaxPoly + 1][2 * kMaxPoly + 1];
};
__device__ void herm1(Herm1 &h, double pa, double pb, double inv2p) {
  for (int i = 0; i <= kMaxPoly; ++i)
    for (int j = 0; j <= kMaxPoly; ++j)
      for (int t = 0; t <= 2 * kMaxPoly; ++t) h.e[i][j][t] = 0.0;
  h.e[0][0][0] = 1.0;
  for (int i = 0; i <= kMaxPoly; ++i) {
    if (i > 0)
      for (int t = 0; t <= i; ++t) {
        double v = pa * h.e[i - 1][0][t];
        if (t > 0) v += inv2p * h.e[i - 1][0][t - 1];
        if (t + 1 <= i - 1) v += (t + 1) * h.e[i - 1][0][t + 1];
        h.e[i][0][t] = v;
      }
    for (int j = 1; j <= kMaxPoly; ++j)
      for (int t = 0; t <= i + j; ++t) {
        double v = pb * h.e[i][j - 1][t];
        if (t > 0) v += inv2p * h.e[i][j - 1][t - 1];
        if (t + 1 <= i + j - 1) v += (t + 1) * h.e[i][j - 1][t + 1];
        h.e[i][j][t] = v;
      }
  }
}

__global__ void __launch_bounds__(128) etab_kernel(EtabArgs a) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nitems) return;
  int lo = 0, hi = a.npairs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.first[mid] <= g) lo = mid; else hi = mid - 1;
  }
  const EPair P = a.pairs[lo];
  int ij = (int)(g - a.first[lo]);
  int i = ij / P.nB, j = ij % P.nB;
  double ea = a.exps[P.expA + i], eb = a.exps[P.expB + j];
  double p = ea + eb;
  double px = (ea * P.ax + eb * P.bx) / p, py = (ea * P.ay + eb * P.by) / p,
         pz = (ea * P.az + eb * P.bz) / p;
  double dx = P.ax - P.bx, dy = P.ay - P.by, dz = P.az - P.bz;
  double kab = exp(-(ea * eb * (dx * dx + dy * dy + dz * dz)) / p);
  double inv2p = 1.0 / (2.0 * p);
  Herm1 hx, hy, hz;
  herm1(hx, px - P.ax, px - P.bx, inv2p);
  herm1(hy, py - P.ay, py - P.by, inv2p);
  herm1(hz, pz - P.az, pz - P.bz, inv2p);
  constexpr int NT = ntuv(kMaxLab);
  double acc[2][2][NT];
  int nt = ntuv(P.lam);
  for (int s = 0; s < 2; ++s)
    for (int n = 0; n < nt; ++n) acc[s][0][n] = acc[s][1][n] = 0.0;
  for (int s = 0; s < 2; ++s) {
    int tb = s == 0 ? P.term11 : P.term21;
    int tn = s == 0 ? P.nterm11 : P.nterm21;
    for (int q = tb; q < tb + tn; ++q) {
      ETerm T = a.terms[q];
      double w = T.w;
      if (T.pa) w *= ea;
      if (T.pb) w *= eb;
      int ab0 = T.polyA >> 16, abn = T.polyA & 0xffff;
      int bb0 = T.polyB >> 16, bbn = T.polyB & 0xffff;
      for (int ma = ab0; ma < ab0 + abn; ++ma) {
        Mono A = a.monos[ma];
        for (int mb = bb0; mb < bb0 + bbn; ++mb) {
          Mono B = a.monos[mb];
          double cr = w * (A.re * B.re - A.im * B.im);
          double ci = w * (A.re * B.im + A.im * B.re);
          for (int t = 0; t <= A.i + B.i; ++t) {
            double ex = hx.e[A.i][B.i][t];
            if (ex == 0.0) continue;
            for (int u = 0; u <= A.j + B.j; ++u) {
              double exy = ex * hy.e[A.j][B.j][u];
              if (exy == 0.0) continue;
              for (int v = 0; v <= A.k + B.k; ++v) {
                double e3 = exy * hz.e[A.k][B.k][v];
                int n = hidx(t, u, v);
                acc[s][0][n] += cr * e3;
                acc[s][1][n] += ci * e3;
              }
            }
          }
        }
      }
    }
  }
  // large component of A except for SS; small component of B except for LL
  double nrm = (P.small == kSS ? a.normS[P.expA + i] : a.normL[P.expA + i]) *
               (P.small == kLL ? a.normL[P.expB + j] : a.normS[P.expB + j]);
  double sc = kab * nrm;
  int maxab = P.nA * P.nB;
  for (int n = 0; n < nt; ++n) {
    double r0 = sc * acc[0][0][n], i0 = sc * acc[0][1][n];
    double r1 = sc * acc[1][0][n], i1 = sc * acc[1][1][n];
    if (P.phase == 1) {  // times i (EQLS :35410-35416)
      double t0 = r0, t1 = r1;
      r0 = -i0; i0 = t0; r1 = -i1; i1 = t1;
    } else if (P.phase == 2) {  // times i twice for sigma_y (:35418-35425)
      r0 = -r0; i0 = -i0; r1 = -r1; i1 = -i1;
    }
    a.E[P.off + (int64_t)n * maxab + ij] = make_double4(r0, i0, r1, i1);
  }
}

// per (pair, tuv, component): sum over ij of |E| > 1e-10 (E0LLCD0 EPSCD);
// with zflags also whether any entry is nonzero at all
__global__ void __launch_bounds__(64) eflag_kernel(EflagArgs a) {
  int p = blockIdx.x;
  const EPair P = a.pairs[p];
  int nt = ntuv(P.lam), maxab = P.nA * P.nB;
  for (int n = threadIdx.x; n < nt; n += blockDim.x) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int ij = 0; ij < maxab; ++ij) {
      double4 e = a.E[P.off + (int64_t)n * maxab + ij];
      s0 += fabs(e.x); s1 += fabs(e.y); s2 += fabs(e.z); s3 += fabs(e.w);
    }
    uint8_t *f = a.flags + a.flag_off[p] + 4 * n;
    f[0] = s0 > 1.0e-10; f[1] = s1 > 1.0e-10; f[2] = s2 > 1.0e-10; f[3] = s3 > 1.0e-10;
    if (a.zflags) {
      uint8_t *z = a.zflags + a.flag_off[p] + 4 * n;
      z[0] = s0 > 0.0; z[1] = s1 > 0.0; z[2] = s2 > 0.0; z[3] = s3 > 0.0;
    }
  }
}

// E(CD) copy for ERIFAST's second contraction (:12809-13119): the column
// masks and the sign Z = (-1)^(t'+u'+v') applied once per table entry
// instead of once per quartet, pairs in kl = l * nC + k order; an FP32 copy
// of it for the mixed build.
__global__ void __launch_bounds__(256) fold_kernel(EflagArgs a) {
  const int p = blockIdx.x;
  const EPair P = a.pairs[p];
  const int maxab = P.nA * P.nB;
  const int64_t nw = (int64_t)ntuv(P.lam) * maxab;
  for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) {
    const int n = (int)(w / maxab);
    int t, u, v;
    dec(n, t, u, v);
    const double z = ((t + u + v) & 1) ? -1.0 : 1.0;
    const uint8_t *f = a.flags + a.flag_off[p] + 4 * n;
    const double4 e = a.E[P.off + w];
    const double4 o = make_double4(f[0] ? z * e.x : 0.0, f[1] ? z * e.y : 0.0,
                                   f[2] ? z * e.z : 0.0, f[3] ? z * e.w : 0.0);
    // pair index transposed, kl = l * nC + k: the kernels give consecutive
    // lanes consecutive k, so the digestion's G(C,D) atomics and D(C,D) loads
    // touch consecutive matrix rows
    const int ij = (int)(w - (int64_t)n * maxab), i = ij / P.nB, j = ij - i * P.nB;
    const int64_t dst = P.off + (int64_t)n * maxab + j * P.nA + i;
    a.Ecd[dst] = o;
    if (a.Ecd32) a.Ecd32[dst] = make_float4((float)o.x, (float)o.y, (float)o.z, (float)o.w);
    if (a.Ecd1) a.Ecd1[dst] = o.x + o.y + o.z + o.w;
    if (a.Ecd1_32) a.Ecd1_32[dst] = (float)(o.x + o.y + o.z + o.w);
  }
}

// ---------------------------------------------------------------------------
// Schwarz tables
// ---------------------------------------------------------------------------
// R^(0)_{tuv} at P = Q: (-2a)^k (t-1)!!(u-1)!!(v-1)!!/(2k+1) for even t,u,v
// (k = (t+u+v)/2), zero otherwise -- the Taylor coefficients of F_0(a r^2).
__device__ __forceinline__ double r_origin(int t, int u, int v, double m2a) {
  if ((t | u | v) & 1) return 0.0;
  int k = (t + u + v) >> 1;
  double x = c_rcp[k];
  for (int i = 0; i < k; ++i) x *= m2a;
  for (int i = t - 1; i > 1; i -= 2) x *= i;
  for (int i = u - 1; i > 1; i -= 2) x *= i;
  for (int i = v - 1; i > 1; i -= 2) x *= i;
  return x;
}

// One primitive pair (i,j) of one block pair: the two sign-class integrals
// s0 = sum Z E11cd G11 and s1 = sum Z E21cd G21 (conj(G21) for the bound
// table), G = E_AB . R(P=Q); the four sign slots take sqrt|.| (CLMDIAG
// :38399-38404 for GDSC).
__global__ void __launch_bounds__(128) diag_kernel(DiagArgs a) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nitems) return;
  int lo = 0, hi = a.njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.first[mid] <= g) lo = mid; else hi = mid - 1;
  }
  const DiagJob J = a.jobs[lo];
  int ij = (int)(g - a.first[lo]);
  int i = ij / J.nB, j = ij % J.nB;
  int maxab = J.nA * J.nB;
  int kl = J.swap ? j * J.nA + i : ij;
  double p = a.exps[J.expA + i] + a.exps[J.expB + j];
  double m2a = -p;  // -2 alpha, alpha = p q / (p + q) = p / 2
  double pre = 2.0 * 1.7493418327624863e1 / (sqrt(2.0 * p) * p * p);
  double sr[2] = {0.0, 0.0}, si[2] = {0.0, 0.0};
  for (int tp = 0; tp < J.ntcd; ++tp) {
    int t2, u2, v2;
    dec(tp, t2, u2, v2);
    double z = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
    double g11r = 0, g11i = 0, g21r = 0, g21i = 0;
    for (int t = 0; t < J.ntab; ++t) {
      int t1, u1, v1;
      dec(t, t1, u1, v1);
      if (((t1 + t2) | (u1 + u2) | (v1 + v2)) & 1) continue;
      double r = r_origin(t1 + t2, u1 + u2, v1 + v2, m2a);
      double4 e = a.E[J.eab + (int64_t)t * maxab + ij];
      g11r += e.x * r; g11i += e.y * r; g21r += e.z * r; g21i += e.w * r;
    }
    double4 e = a.E[J.ecd + (int64_t)tp * maxab + kl];
    sr[0] += z * (e.x * g11r - e.y * g11i);
    si[0] += z * (e.x * g11i + e.y * g11r);
    if (J.swap) {
      sr[1] += z * (e.z * g21r + e.w * g21i);
      si[1] += z * (e.w * g21r - e.z * g21i);
    } else {
      sr[1] += z * (e.z * g21r - e.w * g21i);
      si[1] += z * (e.z * g21i + e.w * g21r);
    }
  }
  double q[2];
  for (int s = 0; s < 2; ++s) {
    double m = J.realpart ? fabs(sr[s]) : sqrt(sr[s] * sr[s] + si[s] * si[s]);
    q[s] = sqrt(pre * m);
  }
  double *out = J.out ? a.qbnd : a.gdsc;
  int nd = a.ndim;
  out[(int64_t)(J.col[0] + j) * nd + J.row[0] + i] = q[0];
  out[(int64_t)(J.col[1] + j) * nd + J.row[1] + i] = q[0];
  out[(int64_t)(J.col[1] + j) * nd + J.row[0] + i] = q[1];
  out[(int64_t)(J.col[0] + j) * nd + J.row[1] + i] = q[1];
}

// ---------------------------------------------------------------------------
// Fock build
// ---------------------------------------------------------------------------
// Arithmetic of one precision: the FP64 path and the FP32 path of the
// Schwarz split share the code below.
template <typename T>
struct Arith;
template <>
struct Arith<double> {
  __device__ static double ex(double x) { return exp(x); }
  __device__ static double sq(double x) { return sqrt(x); }
  __device__ static double rcp(int k) { return c_rcp[k]; }
};
template <>
struct Arith<float> {
  __device__ static float ex(float x) { return expf(x); }
  __device__ static float sq(float x) { return sqrtf(x); }
  __device__ static float rcp(int k) { return c_rcpf[k]; }
};

// Boys function F_0..F_L of the reference's four branches (FUNFM, RMAKE's
// thresholds 1e-11 / 17 / 30, bertha-scf.f:31927-31985, 32495-32605),
// written to f[n*ld].
template <typename T>
__device__ void boys_ref(T t, int L, T *f, int ld) {
  using A = Arith<T>;
  const T kA0 = (T)4.994501191201870e-1, kB0 = (T)4.551838436668326e-1;
  const T kSqrtPiHalf = (T)(0.5 * 1.7724538509055160);
  if (t <= (T)1.0e-11) {
    for (int m = 0; m <= L; ++m) f[m * ld] = A::rcp(m);
  } else if (t <= (T)17.0) {
    T txp = A::ex(-t), tt2 = 2 * t, tl = 1, s = 1;
    for (int k = 1; k <= 60; ++k) {
      tl = tl * tt2 * A::rcp(L + k);
      s += tl;
    }
    T fm = s * txp * A::rcp(L);
    f[L * ld] = fm;
    for (int m = L - 1; m >= 0; --m) {
      fm = (tt2 * fm + txp) * A::rcp(m);
      f[m * ld] = fm;
    }
  } else if (t <= (T)30.0) {
    T txp = A::ex(-t), rt2 = 1 / (2 * t);
    T fm = kSqrtPiHalf / A::sq(t) - txp * (kA0 / (kB0 + t));
    f[0] = fm;
    for (int m = 0; m < L; ++m) {
      fm = ((2 * m + 1) * fm - txp) * rt2;
      f[(m + 1) * ld] = fm;
    }
  } else {
    T rt2 = 1 / (2 * t);
    T fm = kSqrtPiHalf / A::sq(t);
    f[0] = fm;
    for (int m = 0; m < L; ++m) {
      fm = (2 * m + 1) * fm * rt2;
      f[(m + 1) * ld] = fm;
    }
  }
}

// R^(0)_{tuv}(P-Q) for t+u+v <= L in Hermite order, in place in r[n*ld]:
// R^(n)_{000} = (-2a)^n F_n and R^(n)_{t+1,u,v} = X R^(n+1)_{tuv} +
// t R^(n+1)_{t-1,u,v} (McMurchie-Davidson), descending levels so each slot is
// overwritten only after the lower order has been read.
// The seeds (-2a)^n F_n are formed in the precision of f (FP64 for every
// item: the FP32 items of the mixed build run the recursion in FP32 from
// FP64 Boys values).
// rank of Hermite index h within its parity class under axis mask m (classes
// enumerated in CARTIND order); class sizes of h < N(L); the class-0 members
__constant__ int16_t c_crank[8][ntuv(kMaxLfl)];
__constant__ int16_t c_csz[8][kMaxLfl + 1][8];
__constant__ int16_t c_cls0[8][ntuv(kMaxLfl)];
// the members of each parity class, ascending, of h < N(kMaxLab) (the ket
// Hermite indices t' of a pair table): c_cmem[m][c][r], r < c_csz[m][L][c]
// for those below N(L); 32-bit so that uniform reads go through the scalar
// cache
__constant__ int c_cmem[8][8][ntuv(kMaxLab)];
// R on the z axis (X = Y = 0): R_tuv = (t-1)!! (u-1)!! sum_j c_hz[v][j]
// Z^(v-2j) A_((t+u)/2+v-j), A_k = (-2 alpha)^k F_k, for even t and u (else
// zero): c_hz[v][j] = v! / (j! (v-2j)! 2^j), c_dfac[t] = (t-1)!!
__constant__ double c_hz[kMaxLfl + 1][kMaxLfl / 2 + 1];
__constant__ double c_dfac[kMaxLfl + 1];

__device__ __forceinline__ int hcls(int n, int m) { return hpar(n) & m; }

// RMAKE's R for a class on the z axis (axis mask m holds x and y, so P - Q
// lies on z): the class-0 values (parity class 0 under m) from the Boys
// values F_0..F_L at stride ld, which it scales in place to A_k; the other
// classes vanish exactly and are not written.  Each value is the closed form
// above, a sum of at most L/2 + 1 terms with no chain through r.
template <typename T>
__device__ void rints_z(T *r, double *f, int ld, int L, double alpha, double Z, int m) {
  double pw = 1.0;
  for (int k = 0; k <= L; ++k) {
    f[(int64_t)k * ld] *= pw;
    pw *= -2.0 * alpha;
  }
  const int nr0 = c_csz[m][L][0];
  for (int w = 0; w < nr0; ++w) {
    const int h = c_cls0[m][w];
    const uint32_t tv = c_tuv[h];
    const int t = (int)(tv >> 10), u = (int)((tv >> 5) & 31), v = (int)(tv & 31);
    const int k0 = (t + u) / 2 + v;
    double zp = (v & 1) ? Z : 1.0, sum = 0.0;
    for (int j = v / 2; j >= 0; --j) {  // Z^(v - 2j) rises as j falls
      sum += c_hz[v][j] * zp * f[(int64_t)(k0 - j) * ld];
      zp *= Z * Z;
    }
    r[(int64_t)h * ld] = (T)(c_dfac[t] * c_dfac[u] * sum);
  }
}

template <typename T, typename FT>
__device__ void rints(T *r, const FT *f, int ld, int L, FT alpha, T X, T Y, T Z) {
  const FT m2a = -2 * alpha;
  FT pw = 1;
  for (int k = 0; k < L; ++k) pw *= m2a;
  r[0] = (T)(pw * f[L * ld]);
  for (int n = L - 1; n >= 0; --n) {
    for (int lev = L - n; lev >= 1; --lev) {
      for (int t = lev; t >= 0; --t)
        for (int u = lev - t; u >= 0; --u) {
          int v = lev - t - u;
          T val;
          if (t > 0) {
            val = X * r[hidx(t - 1, u, v) * ld];
            if (t > 1) val += (t - 1) * r[hidx(t - 2, u, v) * ld];
          } else if (u > 0) {
            val = Y * r[hidx(t, u - 1, v) * ld];
            if (u > 1) val += (u - 1) * r[hidx(t, u - 2, v) * ld];
          } else {
            val = Z * r[hidx(t, u, v - 1) * ld];
            if (v > 1) val += (v - 1) * r[hidx(t, u, v - 2) * ld];
          }
          r[hidx(t, u, v) * ld] = val;
        }
    }
    pw = 1;
    for (int k = 0; k < n; ++k) pw *= m2a;
    r[0] = (T)(pw * f[n * ld]);
  }
}

// digestion batches of CLMMTG (bertha-scf.f:13902-14392): labels 0..3 =
// A,B,C,D; each batch adds phase * RR(sA,sB,sC,sD) * D(drow,dcol) into
// G(orow,ocol); *_f flags swap the sign slot of that label; pab/pcd select
// the PAB/PCD permutation phases.
struct Batch {
  int8_t xch, orow, ocol, oflr, oflc, drow, dcol, dflr, dflc, pab, pcd;
};
__constant__ Batch c_batch[11] = {
    {0, 0, 1, 0, 0, 2, 3, 0, 0, 0, 0}, {0, 0, 1, 0, 0, 3, 2, 1, 1, 0, 1},
    {0, 2, 3, 0, 0, 0, 1, 0, 0, 0, 0}, {0, 2, 3, 0, 0, 1, 0, 1, 1, 1, 0},
    {1, 0, 3, 0, 0, 2, 1, 0, 0, 0, 0}, {1, 0, 2, 0, 1, 3, 1, 1, 0, 0, 1},
    {1, 1, 3, 1, 0, 2, 0, 0, 1, 1, 0}, {1, 1, 2, 1, 1, 3, 0, 1, 1, 1, 1},
    {1, 2, 1, 0, 0, 0, 3, 0, 0, 0, 0}, {1, 2, 0, 0, 1, 1, 3, 1, 0, 1, 0},
    {1, 3, 1, 1, 0, 0, 2, 0, 1, 0, 1}};

__device__ __forceinline__ void atomic_add2(double2 *p, double re, double im) {
  unsafeAtomicAdd(&p->x, re);
  if (im != 0.0) unsafeAtomicAdd(&p->y, im);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

constexpr Batch kBatchTab[11] = {
    {0, 0, 1, 0, 0, 2, 3, 0, 0, 0, 0}, {0, 0, 1, 0, 0, 3, 2, 1, 1, 0, 1},
    {0, 2, 3, 0, 0, 0, 1, 0, 0, 0, 0}, {0, 2, 3, 0, 0, 1, 0, 1, 1, 1, 0},
    {1, 0, 3, 0, 0, 2, 1, 0, 0, 0, 0}, {1, 0, 2, 0, 1, 3, 1, 1, 0, 0, 1},
    {1, 1, 3, 1, 0, 2, 0, 0, 1, 1, 0}, {1, 1, 2, 1, 1, 3, 0, 1, 1, 1, 1},
    {1, 2, 1, 0, 0, 0, 3, 0, 0, 0, 0}, {1, 2, 0, 0, 1, 1, 3, 1, 0, 1, 0},
    {1, 3, 1, 1, 0, 0, 2, 0, 1, 0, 1}};

// The four raw sums of the second contraction for one kl: a_s = sum Z E_CD G,
// b_s = sum Z E_CD conj(G) for the pairings s = (E11cd,G11), (E21cd,G11),
// (E11cd,G21), (E21cd,G21), with the phases PAB1 and PAB1*PCD1 of the call.
struct Raw {
  double ar[4], ai[4], br[4], bi[4];
  double pab, pabcd;
};

// RR(n), n = 1 + 8 sA + 4 sB + 2 sC + sD (s = 1 for +|m|), from the raw sums
// (ERIFAST :12822-13119); the real path (ISYM 1, 2) keeps real parts of the
// sign-diagonal combinations only.  n is a compile-time constant after
// unrolling, so only the needed products are formed.
template <bool REAL>
__device__ __forceinline__ void rrv(const Raw &w, int n, double &re, double &im) {
  const double p = w.pab, q = w.pabcd;
  re = 0.0;
  im = 0.0;
  switch (n) {
    case 0: re = w.ar[0]; im = w.ai[0]; break;
    case 12: re = p * w.br[0]; im = p * w.bi[0]; break;
    case 15: re = q * w.ar[0]; im = -q * w.ai[0]; break;
    case 3: re = q * (p * w.br[0]); im = -q * (p * w.bi[0]); break;
    case 10: re = w.ar[3]; im = w.ai[3]; break;
    case 6: re = -p * w.br[3]; im = -p * w.bi[3]; break;
    case 5: re = q * w.ar[3]; im = -q * w.ai[3]; break;
    case 9: re = q * (-p * w.br[3]); im = -q * (-p * w.bi[3]); break;
    case 2: re = w.ar[1]; im = w.ai[1]; break;
    case 14: re = p * w.br[1]; im = p * w.bi[1]; break;
    case 13: re = -q * w.ar[1]; im = q * w.ai[1]; break;
    case 1: re = -q * (p * w.br[1]); im = q * (p * w.bi[1]); break;
    case 8: re = w.ar[2]; im = w.ai[2]; break;
    case 4: re = -p * w.br[2]; im = -p * w.bi[2]; break;
    case 7: re = -q * w.ar[2]; im = q * w.ai[2]; break;
    default: re = -q * (-p * w.br[2]); im = q * (-p * w.bi[2]); break;  // 11
  }
  if (REAL) {
    const bool diag = n == 0 || n == 12 || n == 15 || n == 3 || n == 10 || n == 6 || n == 5 ||
                      n == 9;
    if (!diag) re = 0.0;
    im = 0.0;
  }
}

// The density elements one digestion batch reads, d[2*rowslot+colslot]
template <int B, bool REAL>
__device__ __forceinline__ void load_d(const int (&off)[4][2], const int (&idx)[4],
                                       const double2 *__restrict__ Dm, int nd, double2 (&d)[4]) {
  constexpr Batch bt = kBatchTab[B];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int rs = s >> 1, cs = s & 1;
    if (REAL && rs != cs) {
      d[s] = make_double2(0.0, 0.0);
      continue;
    }
    d[s] = Dm[(int64_t)(off[bt.dcol][cs] + idx[bt.dcol]) * nd + off[bt.drow][rs] + idx[bt.drow]];
  }
}

// One digestion batch for one (ij,kl): out[2*rowslot+colslot] accumulates
// the four output sub-blocks.  Everything indexed by the batch table folds to
// constants after unrolling, so the 16 terms stay in registers.
template <int B, bool REAL>
__device__ __forceinline__ void batch_terms(const Raw &raw, const double2 (&d)[4],
                                            double pab1, double pab2, double pcd1, double pcd2,
                                            double (&out)[4][2]) {
  constexpr Batch bt = kBatchTab[B];
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int sl[4] = {(n >> 3) & 1, (n >> 2) & 1, (n >> 1) & 1, n & 1};
    const int ors = sl[bt.orow] ^ bt.oflr, ocs = sl[bt.ocol] ^ bt.oflc;
    const int drs = sl[bt.drow] ^ bt.dflr, dcs = sl[bt.dcol] ^ bt.dflc;
    if (REAL && (ors != ocs || drs != dcs)) continue;
    double ph = 1.0;
    if (bt.pab) ph *= (sl[0] != sl[1]) ? pab2 : pab1;
    if (bt.pcd) ph *= (sl[2] != sl[3]) ? pcd2 : pcd1;
    const double2 dv = d[2 * drs + dcs];
    double rr, ri;
    rrv<REAL>(raw, n, rr, ri);
    if (REAL) {
      out[2 * ors + ocs][0] += ph * rr * dv.x;
    } else {
      out[2 * ors + ocs][0] += ph * (rr * dv.x - ri * dv.y);
      out[2 * ors + ocs][1] += ph * (rr * dv.y + ri * dv.x);
    }
  }
}

// LDS accumulators of the exchange batches of one call, packed for the class
// at hand: the batches whose output varies only with l (5, 7, 11) as
// [3][slots][nD], then those varying only with k (6, 8, 9, 10) as
// [4][slots][nC].  Complex path: four sign slots, (re, im) each,
// 2 (12 nD + 16 nC) doubles; real path (ISYM 1, 2; CLMMTZ keeps the real parts
// of the sign-diagonal slots only): slots 0 and 3, one double each,
// 6 nD + 8 nC doubles.
struct XView {
  double *p;
  int nC, nD;
};
constexpr int kXAccMax = 2 * (12 * kMaxPrim + 16 * kMaxPrim);  // doubles, any class

template <bool REAL>
__device__ __forceinline__ void lds_add4(const XView &x, bool kbatch, int which, int pos,
                                         const double (&o)[4][2]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (REAL && (s == 1 || s == 2)) continue;
    if (REAL) {
      const int si = s == 3;
      double *d = kbatch ? x.p + 6 * x.nD + (which * 2 + si) * x.nC + pos
                         : x.p + (which * 2 + si) * x.nD + pos;
      if (o[s][0] != 0.0) unsafeAtomicAdd(d, o[s][0]);
    } else {
      double *d = kbatch ? x.p + 2 * (12 * x.nD + (which * 4 + s) * x.nC + pos)
                         : x.p + 2 * ((which * 4 + s) * x.nD + pos);
      if (o[s][0] != 0.0) unsafeAtomicAdd(d, o[s][0]);
      if (o[s][1] != 0.0) unsafeAtomicAdd(d + 1, o[s][1]);
    }
  }
}

// digestion of one RR batch element (CLMMTG :13902-14392, CLMMTZ
// :13196-13514 when REAL): batches 1-2 into registers (G(A,B) is fixed by
// ij), 3-4 straight to global memory, exchange batches into LDS.  With PAIR
// the density loads of two batches are issued together so their latencies
// overlap; without, each batch loads just before it uses them, which keeps
// the kernel within the registers of three waves per SIMD.  Batches 3-4 read
// D(A,B)(i,j), which is the same for every kl and comes in preloaded (dab).
template <bool REAL, bool PAIR>
__device__ __forceinline__ void digest(const Raw &raw,
                                       const int (&off)[4][2], const int (&idx)[4],
                                       const double2 *__restrict__ Dm, int nd, int iflg,
                                       int intsym, double pab1, double pab2, double pcd1,
                                       double pcd2, const double2 (&dab)[2][4],
                                       double (&accAB)[4][2], const XView &xa, double2 *gdir) {
  if (!intsym) iflg &= 0x11;
  {
    // PAIR: the density loads of batches 1 and 2 issued together (latencies
    // overlap, more registers live)
    double2 dp0[4], dp1[4];
    if (PAIR) {
      if (iflg & 1) load_d<0, REAL>(off, idx, Dm, nd, dp0);
      if (iflg & 2) load_d<1, REAL>(off, idx, Dm, nd, dp1);
    }
    if (iflg & 12) {
      double o[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      if (iflg & 4) batch_terms<2, REAL>(raw, dab[0], pab1, pab2, pcd1, pcd2, o);
      if (iflg & 8) batch_terms<3, REAL>(raw, dab[1], pab1, pab2, pcd1, pcd2, o);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int rs = s >> 1, cs = s & 1;
        if (REAL && rs != cs) continue;
        atomic_add2(&gdir[(int64_t)(off[3][cs] + idx[3]) * nd + off[2][rs] + idx[2]], o[s][0],
                    o[s][1]);
      }
    }
    if (iflg & 1) {
      if (!PAIR) load_d<0, REAL>(off, idx, Dm, nd, dp0);
      batch_terms<0, REAL>(raw, dp0, pab1, pab2, pcd1, pcd2, accAB);
    }
    if (iflg & 2) {
      if (!PAIR) load_d<1, REAL>(off, idx, Dm, nd, dp1);
      batch_terms<1, REAL>(raw, dp1, pab1, pab2, pcd1, pcd2, accAB);
    }
  }
#define XL(B, D) \
  if (iflg & (1 << B)) load_d<B, REAL>(off, idx, Dm, nd, D);
#define XB(B, D, KB, W, POS)                                             \
  if (iflg & (1 << B)) {                                                 \
    double o[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};                   \
    batch_terms<B, REAL>(raw, D, pab1, pab2, pcd1, pcd2, o);             \
    lds_add4<REAL>(xa, KB, W, POS, o);                                   \
  }
  if (PAIR) {
    {
      double2 d4[4], d5[4];
      XL(4, d4) XL(5, d5)
      XB(4, d4, false, 0, idx[3])
      XB(5, d5, true, 0, idx[2])
    }
    {
      double2 d6[4], d7[4];
      XL(6, d6) XL(7, d7)
      XB(6, d6, false, 1, idx[3])
      XB(7, d7, true, 1, idx[2])
    }
    {
      double2 d8[4], d9[4];
      XL(8, d8) XL(9, d9)
      XB(8, d8, true, 2, idx[2])
      XB(9, d9, true, 3, idx[2])
    }
  } else {
    { double2 d4[4]; XL(4, d4) XB(4, d4, false, 0, idx[3]) }
    { double2 d5[4]; XL(5, d5) XB(5, d5, true, 0, idx[2]) }
    { double2 d6[4]; XL(6, d6) XB(6, d6, false, 1, idx[3]) }
    { double2 d7[4]; XL(7, d7) XB(7, d7, true, 1, idx[2]) }
    { double2 d8[4]; XL(8, d8) XB(8, d8, true, 2, idx[2]) }
    { double2 d9[4]; XL(9, d9) XB(9, d9, true, 3, idx[2]) }
  }
  { double2 d10[4]; XL(10, d10) XB(10, d10, false, 2, idx[3]) }
#undef XL
#undef XB
}

// SCHWARZ (bertha-scf.f:37889-38123) for one kl of one ERIFAST call: bit b
// keeps digestion batch b+1 (IMTX), 0 drops the kl (ISCR = 0).  Bound groups
// in the reference's order: direct (1,2) with D(C,D), (3,4) with D(A,B);
// exchange (8,9) with D(A,D), (5) with D(B,C), (7,11) with D(A,C), (6,10)
// with D(B,D); the first group above SENS ends the test.
// gslot and dslot hold, at the (-|m|, -|m|) element of every block pair, the
// largest GDSC (the Cauchy-Schwarz factor Q in mode 2) and |D| over its four
// sign slots (slot_tables_kernel), so each bound and density factor is one
// load.  gab and dab are the (A,B) factors,
// fixed for one primitive pair ij of one AB combination.
__device__ __forceinline__ int64_t slot_at(const int (&off)[4][2], const int (&idx)[4], int r,
                                           int c, int nd) {
  return (int64_t)(off[c][0] + idx[c]) * nd + off[r][0] + idx[r];
}
// With rigorous set (mode 2) gslot holds the Cauchy-Schwarz factors Q and
// every group is bounded by Q(A,B) Q(C,D), a bound of |(ab|cd)| itself.
__device__ int schwarz_mask(const double *gslot, const double *dslot, int nd,
                            const int (&off)[4][2], const int (&idx)[4], double sens, double gab,
                            double dab, bool rigorous) {
  const int64_t cd = slot_at(off, idx, 2, 3, nd);
  const double gdir = gab * gslot[cd];
  int mask = 0x7ff;
  if (gdir * dslot[cd] > sens) return mask;
  mask &= ~0x003;
  if (gdir * dab > sens) return mask;
  mask &= ~0x00c;
  const double gx1 =
      rigorous ? gdir : gslot[slot_at(off, idx, 1, 2, nd)] * gslot[slot_at(off, idx, 0, 3, nd)];
  if (gx1 * dslot[slot_at(off, idx, 0, 3, nd)] > sens) return mask;
  mask &= ~0x180;
  if (gx1 * dslot[slot_at(off, idx, 1, 2, nd)] > sens) return mask;
  mask &= ~0x010;
  const double gx2 =
      rigorous ? gdir : gslot[slot_at(off, idx, 1, 3, nd)] * gslot[slot_at(off, idx, 0, 2, nd)];
  if (gx2 * dslot[slot_at(off, idx, 0, 2, nd)] > sens) return mask;
  mask &= ~0x440;
  if (gx2 * dslot[slot_at(off, idx, 1, 3, nd)] > sens) return mask;
  return 0;
}

// E-table rows in the working precision: the FP64 table, or its FP32 copy
__device__ __forceinline__ void lde(const FockArgs &a, int64_t i, double &x, double &y, double &z,
                                    double &w) {
  double4 e = a.E[i];
  x = e.x; y = e.y; z = e.z; w = e.w;
}
__device__ __forceinline__ void lde(const FockArgs &a, int64_t i, float &x, float &y, float &z,
                                    float &w) {
  float4 e = a.E32[i];
  x = e.x; y = e.y; z = e.z; w = e.w;
}
// the one nonzero component of an E(CD) entry (one-component plans; FP64 or
// FP32 table of one value per entry, so that consecutive lanes read
// consecutive values)
__device__ __forceinline__ void ldcd1(const FockArgs &a, int64_t i, double &x) { x = a.Ecd1[i]; }
__device__ __forceinline__ void ldcd1(const FockArgs &a, int64_t i, float &x) { x = a.Ecd1_32[i]; }
// the same for the masked, signed E(CD) table of the second contraction
__device__ __forceinline__ void ldcd(const FockArgs &a, int64_t i, double &x, double &y, double &z,
                                     double &w) {
  double4 e = a.Ecd[i];
  x = e.x; y = e.y; z = e.z; w = e.w;
}
__device__ __forceinline__ void ldcd(const FockArgs &a, int64_t i, float &x, float &y, float &z,
                                     float &w) {
  float4 e = a.Ecd32[i];
  x = e.x; y = e.y; z = e.z; w = e.w;
}

// Flush the exchange accumulators of one call into the exchange matrix (one
// atomic per output element instead of one per quartet), then clear them.
template <bool REAL>
__device__ void flush_exchange(const XView &x, int flg, const int (&offs)[4][2], int ib, int jb,
                               double2 *out, int nd, int lane) {
  __syncthreads();
  const int nC = x.nC, nD = x.nD;
  constexpr int ns = REAL ? 2 : 4;  // sign slots kept
  const int nl = 3 * ns * nD, ntot = nl + 4 * ns * nC;
  for (int w = lane; w < ntot; w += 64) {
    int b, s, pos;
    if (w < nl) {
      const int li = w / (ns * nD), r = w - li * ns * nD;
      s = r / nD;
      pos = r - s * nD;
      b = li == 0 ? 4 : (li == 1 ? 6 : 10);
    } else {
      const int ki = (w - nl) / (ns * nC), r = (w - nl) - ki * ns * nC;
      s = r / nC;
      pos = r - s * nC;
      b = ki == 0 ? 5 : ki + 6;
    }
    if (REAL) s = s ? 3 : 0;
    if (!((flg >> b) & 1)) continue;
    double *v = REAL ? x.p + w : x.p + 2 * w;
    int row, col;
    const int rs = s >> 1, cs = s & 1;
    switch (b) {
      case 4: row = offs[0][rs] + ib; col = offs[3][cs] + pos; break;
      case 5: row = offs[0][rs] + ib; col = offs[2][cs] + pos; break;
      case 6: row = offs[1][rs] + jb; col = offs[3][cs] + pos; break;
      case 7: row = offs[1][rs] + jb; col = offs[2][cs] + pos; break;
      case 8: row = offs[2][rs] + pos; col = offs[1][cs] + jb; break;
      case 9: row = offs[2][rs] + pos; col = offs[0][cs] + ib; break;
      default: row = offs[3][rs] + pos; col = offs[1][cs] + jb; break;
    }
    const double im = REAL ? 0.0 : v[1];
    if (v[0] != 0.0 || im != 0.0) atomic_add2(&out[(int64_t)col * nd + row], v[0], im);
    v[0] = 0.0;
    if (!REAL) v[1] = 0.0;
  }
  __syncthreads();
}

// AB combinations whose first contraction shares the R loads
constexpr int kGroup = 4;

// Shared-memory state of one workgroup.  C sets the capacity of a group of
// AB combinations sharing the first contraction's R loads: sum of their
// N(L_AB) and of their N(L_CD) up to C * N(kMaxLab), so C = 2 lets two
// combinations of the largest orders (N = 165, f-f SS pairs) share them.
template <int C>
struct FockShared {
  union {
    uint8_t cflag[ntuv(kMaxLfl)];       // R column flags (until the prefactor pass)
    double eab[C * ntuv(kMaxLab) * 4];  // E(AB) rows of the current AB combinations
  };
  uint8_t gm[C * ntuv(kMaxLab) * 4];    // their G masks
  double2 dab[kGroup][2][4];             // their D(A,B)(i,j), D(B,A)(j,i) slots
};
// the exchange accumulators follow in dynamic LDS, sized for the plan's
// largest class: eight workgroups per CU fit up to 28 primitives per shell
// (seven at the 32-primitive limit)
static_assert(sizeof(FockShared<1>) + 16 + 2 * (12 * 28 + 16 * 28) * sizeof(double) <=
                  160 * 1024 / 8,
              "fock_kernel's LDS exceeds 1/8 of a CU's");
// the capacity-2 instance on the real path (ISYM 1, 2: 6 nD + 8 nC
// accumulator doubles) keeps eight workgroups per CU at any shell size; on
// the complex path fock_launch_config takes it only where it fits
static_assert(sizeof(FockShared<2>) + 16 + (6 * kMaxPrim + 8 * kMaxPrim) * sizeof(double) <=
                  160 * 1024 / 8,
              "fock_kernel<.,2,2>'s LDS exceeds 1/8 of a CU's on the real path");

// One work item (class, ij) in precision T: R-integrals for every kl, the two
// E.R.E contractions (ERIFAST) and the digestion into GDIR/GXCH, which always
// accumulates in FP64.
template <typename T, int W, int C>
__device__ void fock_item(const FockArgs &a, const WorkItem &it, double *slab, FockShared<C> &sh,
                          double *xacc, int lane) {
  const ClassDesc &cl = a.classes[it.cls];
  // slab rows packed at the class's own kl count, so the rows a workgroup
  // touches occupy as few cache lines as possible
  const int ld = cl.maxcd;
  // R [rows][ld] in T, Boys values F [kMaxLfl+1][ld] in FP64 (8-byte
  // aligned), G [ntg*4][ld] in T: within the slab's rows * ld doubles
  char *base = reinterpret_cast<char *>(slab);
  T *R = reinterpret_cast<T *>(base);
  const int64_t fo = ((int64_t)ntuv(kMaxLfl) * ld * (int64_t)sizeof(T) + 7) & ~(int64_t)7;
  double *F = reinterpret_cast<double *>(base + fo);
  T *G = reinterpret_cast<T *>(base + fo + (int64_t)(kMaxLfl + 1) * ld * 8);
  T *s_eab = reinterpret_cast<T *>(sh.eab);
  const double2 *Dm = a.dent;
  const int nd = a.ndim;
  const int nB = cl.nbas[1], nC = cl.nbas[2], nD = cl.nbas[3];
  const int ib = it.ij / nB, jb = it.ij % nB;
  const XView xv{xacc, cl.nbas[2], nD};
  const double ea = a.exps[cl.exp[0] + ib], eb = a.exps[cl.exp[1] + jb];
  const double p = ea + eb;
  const double Px = (cl.xyz[0][0] * ea + cl.xyz[1][0] * eb) / p;
  const double Py = (cl.xyz[0][1] * ea + cl.xyz[1][1] * eb) / p;
  const double Pz = (cl.xyz[0][2] * ea + cl.xyz[1][2] * eb) / p;
  const int L = cl.lamfl, nt = ntuv(L), maxcd = cl.maxcd;
  const int maxab = cl.nbas[0] * nB;
  const int sym = cl.sym;
  const bool zaxis = (sym & 3) == 3;  // P - Q on the z axis for every kl

  // R-integrals for every kl of this ij, then the reference's column test
  // (sum over kl of |R| > 1e-10 before the prefactor, :12660-12675)
  for (int kl0 = 0; kl0 < maxcd; kl0 += 64) {
    int kl = kl0 + lane;
    if (kl < maxcd) {
      int kb = kl % nC, lb = kl / nC;
      double ec = a.exps[cl.exp[2] + kb], ed = a.exps[cl.exp[3] + lb];
      double q = ec + ed;
      double Qx = (cl.xyz[2][0] * ec + cl.xyz[3][0] * ed) / q;
      double Qy = (cl.xyz[2][1] * ec + cl.xyz[3][1] * ed) / q;
      double Qz = (cl.xyz[2][2] * ec + cl.xyz[3][2] * ed) / q;
      double alpha = p * q / (p + q);
      double X = Px - Qx, Y = Py - Qy, Z = Pz - Qz;
      double Targ = alpha * (X * X + Y * Y + Z * Z);
      boys_ref<double>(Targ, L, F + kl, ld);
      if (zaxis)
        rints_z<T>(R + kl, F + kl, ld, L, alpha, Z, sym);
      else
        rints<T, double>(R + kl, F + kl, ld, L, alpha, (T)X, (T)Y, (T)Z);
    }
  }
  __syncthreads();
  // the column test.  On the z axis only the class-0 columns (the others
  // vanish exactly and are neither written nor read), four at a time with
  // lanes over kl (coalesced rows) and a wave sum per column; elsewhere one
  // column per lane (every column is tested, and the wave sums would cost
  // more than the strided reads)
  for (int col = lane; col < nt; col += 64) {
    if (zaxis) {
      sh.cflag[col] = 0;
      continue;
    }
    T s = 0;
    const T *row = R + (int64_t)col * ld;
    for (int kl = 0; kl < maxcd; ++kl) s += fabs(row[kl]);
    sh.cflag[col] = s > (T)1.0e-10;
  }
  const int ncol = zaxis ? c_csz[sym][L][0] : nt;
  for (int c0 = 0; zaxis && c0 < ncol; c0 += 4) {
    int cols[4];
    T s[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = min(c0 + i, ncol - 1);
      cols[i] = zaxis ? c_cls0[sym][r] : r;
      s[i] = 0;
    }
    for (int kl = lane; kl < maxcd; kl += 64)
#pragma unroll
      for (int i = 0; i < 4; ++i) s[i] += fabs(R[(int64_t)cols[i] * ld + kl]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double tot = wave_sum((double)s[i]);
      if (lane == 0 && c0 + i < ncol) sh.cflag[cols[i]] = tot > 1.0e-10;
    }
  }
  __syncthreads();
  for (int kl0 = 0; kl0 < maxcd; kl0 += 64) {
    int kl = kl0 + lane;
    if (kl < maxcd) {
      int kb = kl % nC, lb = kl / nC;
      double q = a.exps[cl.exp[2] + kb] + a.exps[cl.exp[3] + lb];
      T pre = (T)(2.0 * 1.7493418327624863e1 / (sqrt(p + q) * p * q));
      for (int r = 0; r < ncol; ++r) {
        const int col = zaxis ? c_cls0[sym][r] : r;
        T *x = R + (int64_t)col * ld + kl;
        if (sh.cflag[col])
          *x = pre * *x;
        else
          *x = (T)0;
      }
    }
  }
  __syncthreads();

  for (int a0 = it.ab_lo; a0 < it.ab_hi;) {
    // a group of up to kGroup consecutive combinations of one shape (the plan
    // orders them so) shares every R load of the first contraction
    const ABCombo ab0 = a.abs[a0];
    const int ntab = ab0.ntab, ntg = ab0.ntg;
    int na = 1;
    while (na < kGroup && a0 + na < it.ab_hi && (na + 1) * ntab <= C * ntuv(kMaxLab) &&
           (na + 1) * ntg <= C * ntuv(kMaxLab) && a.abs[a0 + na].ntab == ntab &&
           a.abs[a0 + na].ntg == ntg)
      ++na;
    for (int w = lane; w < na * ntab; w += 64) {
      const int g = w / ntab, t = c_tperm[sym][ab0.lab][w - g * ntab];
      T ex, ey, ez, ew;
      lde(a, a.abs[a0 + g].eoff + (int64_t)t * maxab + it.ij, ex, ey, ez, ew);
      // E(AB) entries of magnitude <= 1e-16 are skipped by ERIFAST (:12445)
      s_eab[4 * w + 0] = fabs(ex) > (T)1.0e-16 ? ex : (T)0;
      s_eab[4 * w + 1] = fabs(ey) > (T)1.0e-16 ? ey : (T)0;
      s_eab[4 * w + 2] = fabs(ez) > (T)1.0e-16 ? ez : (T)0;
      s_eab[4 * w + 3] = fabs(ew) > (T)1.0e-16 ? ew : (T)0;
    }
    for (int w = lane; w < na * 4 * ntg; w += 64) {
      const int g = w / (4 * ntg);
      sh.gm[w] = a.masks[a.abs[a0 + g].gmask + (w - g * 4 * ntg)];
    }
    // D(A,B)(i,j) and D(B,A)(j,i) of batches 3 and 4: fixed for each
    // combination (no density on the RR test path)
    if (!a.rr_out && lane < 8 * na) {
      const ABCombo &abg = a.abs[a0 + (lane >> 3)];
      const int b = (lane >> 2) & 1, sl = lane & 3, rs = sl >> 1, cs = sl & 1;
      const bool off_diag = a.isym != 0 && rs != cs;
      const int row = b == 0 ? abg.fa[rs] + ib : abg.fb[rs] + jb;
      const int col = b == 0 ? abg.fb[cs] + jb : abg.fa[cs] + ib;
      sh.dab[lane >> 3][b][sl] = off_diag ? make_double2(0.0, 0.0) : Dm[(int64_t)col * nd + row];
    }
    __syncthreads();
    // first contraction G_g(kl,tuv') = sum_tuv E_AB,g(tuv) R(tuv+tuv') for
    // the group's combinations g, rows [g][tuv'][4] of the slab's G region
    for (int kl0 = 0; kl0 < maxcd; kl0 += 64) {
      int kl = kl0 + lane;
      if (kl >= maxcd) continue;
      for (int tp = 0; tp < ntg; ++tp) {
        T acc[kGroup][4];
        int any = 0;
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
          acc[g][0] = acc[g][1] = acc[g][2] = acc[g][3] = 0;
          if (g < na) {
            const uint8_t *m = &sh.gm[4 * (g * ntg + tp)];
            any |= m[0] | m[1] | m[2] | m[3];
          }
        }
        if (any && a.one_comp) {
          // one-component plan: G(t') lives in one component cg (E(AB)'s in
          // t''s class, since R(t + t') vanishes off class 0), so only that
          // component is summed and stored, eight R rows in flight (then
          // four), the sum still taken in order of t
          const int pc = hpar(tp) & sym;
          const int tb = c_coff[sym][ab0.lab][pc];
          const int nc = c_coff[sym][ab0.lab][pc + 1] - tb;
          const int *hs = &g_hsp[sym][tp][0];
          const T *eb = s_eab + 4 * tb;
          int cgv[kGroup];
          T s1[kGroup];
#pragma unroll
          for (int g = 0; g < kGroup; ++g) {
            cgv[g] = g < na ? (int)((a.abs[a0 + g].gcomp >> (4 * pc)) & 15) : 4;
            s1[g] = 0;
          }
          int t = 0;
          for (; t + 8 <= nc; t += 8) {
            T r[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) r[i] = R[(int64_t)hs[t + i] * ld + kl];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
              if (g >= na) break;
              const T *e = eb + 4 * (g * ntab + t) + (cgv[g] & 3);
#pragma unroll
              for (int i = 0; i < 8; ++i) s1[g] += e[4 * i] * r[i];
            }
          }
          for (; t + 4 <= nc; t += 4) {
            const T r0 = R[(int64_t)hs[t] * ld + kl];
            const T r1 = R[(int64_t)hs[t + 1] * ld + kl];
            const T r2 = R[(int64_t)hs[t + 2] * ld + kl];
            const T r3 = R[(int64_t)hs[t + 3] * ld + kl];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
              if (g >= na) break;
              const T *e = eb + 4 * (g * ntab + t) + (cgv[g] & 3);
              s1[g] += e[0] * r0;
              s1[g] += e[4] * r1;
              s1[g] += e[8] * r2;
              s1[g] += e[12] * r3;
            }
          }
          for (; t < nc; ++t) {
            const T r = R[(int64_t)hs[t] * ld + kl];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
              if (g >= na) break;
              s1[g] += eb[4 * (g * ntab + t) + (cgv[g] & 3)] * r;
            }
          }
          // masked components are neither stored nor read (second
          // contraction)
#pragma unroll
          for (int g = 0; g < kGroup; ++g) {
            if (g >= na) break;
            const int cg = cgv[g];
            if (cg <= 3 && sh.gm[4 * (g * ntg + tp) + cg])
              G[(int64_t)(4 * (g * ntg + tp) + cg) * ld + kl] = s1[g];
          }
          continue;
        }
        if (any) {
          // R rows from the sum table, independent loads in flight: four with
          // two waves per SIMD, two with three (more would spill); only the t
          // of tp's parity class (the other R rows are exact zeros)
          const int pc = hpar(tp) & sym;
          const int tb = c_coff[sym][ab0.lab][pc];
          const int nc = c_coff[sym][ab0.lab][pc + 1] - tb;
          const int *hs = &g_hsp[sym][tp][0];
          const T *eb = s_eab + 4 * tb;
          int t = 0;
          for (; W == 2 && t + 4 <= nc; t += 4) {
            const T r0 = R[(int64_t)hs[t] * ld + kl];
            const T r1 = R[(int64_t)hs[t + 1] * ld + kl];
            const T r2 = R[(int64_t)hs[t + 2] * ld + kl];
            const T r3 = R[(int64_t)hs[t + 3] * ld + kl];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
              if (g >= na) break;
              const T *e = eb + 4 * (g * ntab + t);
              acc[g][0] += e[0] * r0; acc[g][1] += e[1] * r0;
              acc[g][2] += e[2] * r0; acc[g][3] += e[3] * r0;
              acc[g][0] += e[4] * r1; acc[g][1] += e[5] * r1;
              acc[g][2] += e[6] * r1; acc[g][3] += e[7] * r1;
              acc[g][0] += e[8] * r2; acc[g][1] += e[9] * r2;
              acc[g][2] += e[10] * r2; acc[g][3] += e[11] * r2;
              acc[g][0] += e[12] * r3; acc[g][1] += e[13] * r3;
              acc[g][2] += e[14] * r3; acc[g][3] += e[15] * r3;
            }
          }
          for (; W != 2 && t + 2 <= nc; t += 2) {
            const T r0 = R[(int64_t)hs[t] * ld + kl];
            const T r1 = R[(int64_t)hs[t + 1] * ld + kl];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
              if (g >= na) break;
              const T *e = eb + 4 * (g * ntab + t);
              acc[g][0] += e[0] * r0; acc[g][1] += e[1] * r0;
              acc[g][2] += e[2] * r0; acc[g][3] += e[3] * r0;
              acc[g][0] += e[4] * r1; acc[g][1] += e[5] * r1;
              acc[g][2] += e[6] * r1; acc[g][3] += e[7] * r1;
            }
          }
          for (; t < nc; ++t) {
            const T r = R[(int64_t)hs[t] * ld + kl];
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
              if (g >= na) break;
              const T *e = eb + 4 * (g * ntab + t);
              acc[g][0] += e[0] * r; acc[g][1] += e[1] * r;
              acc[g][2] += e[2] * r; acc[g][3] += e[3] * r;
            }
          }
        }
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
          if (g >= na) break;
          // masked components are neither stored nor read (second
          // contraction); a one-component plan stored its items above (here
          // every mask is off)
          const uint8_t *m = &sh.gm[4 * (g * ntg + tp)];
          T *gp = G + (int64_t)(4 * (g * ntg + tp)) * ld + kl;
          if (m[0]) gp[0] = acc[g][0];
          if (m[1]) gp[ld] = acc[g][1];
          if (m[2]) gp[2 * ld] = acc[g][2];
          if (m[3]) gp[3 * ld] = acc[g][3];
        }
      }
    }
    for (int g = 0; g < na; ++g) {
      const ABCombo ab = a.abs[a0 + g];
      const T *Gg = G + (int64_t)(4 * g * ntg) * ld;
      double accAB[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};  // batches 1,2 output
      const double2 (&dab)[2][4] = sh.dab[g];
      // SCHWARZ's (A,B) factors for this ij (the same for every call and kl)
      double s_gab = 0.0, s_dab = 0.0;
      if (a.schwarz) {
        const int64_t e = (int64_t)(ab.fb[0] + jb) * nd + ab.fa[0] + ib;
        s_gab = a.gslot[e];
        s_dab = a.dslot[e];
      }
      for (int ci = ab.cd_begin; ci < ab.cd_end; ++ci) {
        const CDCall cd = a.cds[ci];
        const double pab1 = cd.pab[0], pcd1 = cd.pcd[0];
        // L_CD of the call (ntcd = N(L_CD)): the class sizes of t' < ntcd
        int lcd = 0;
        while (ntuv(lcd) < cd.ntcd) ++lcd;
        // one-component plans: for parity class c of t' (lane c), bit r set
        // where the class's r-th t' has G's component stored (not masked),
        // the terms of the second contraction (a plan's classes hold at most
        // 64 t'); formed while every lane is active (the kl loop below leaves
        // the tail lanes out, and readlane reads a lane's register whatever
        // its state)
        uint32_t okl = 0, okh = 0;
        if (a.one_comp) {
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const int n = c_csz[sym][lcd][c];
            const int cg = (int)((ab.gcomp >> (4 * c)) & 15);
            bool v = false;
            if (cg <= 3 && lane < n) v = sh.gm[4 * (g * ntg + c_cmem[sym][c][lane]) + cg] != 0;
            const uint64_t m = __ballot(v);
            if (lane == c) {
              okl = (uint32_t)m;
              okh = (uint32_t)(m >> 32);
            }
          }
        }
        const double pabcd = pab1 * pcd1;
        for (int kl0 = 0; kl0 < maxcd; kl0 += 64) {
          int kl = kl0 + lane;
          if (kl >= maxcd) continue;
          // second contraction (ERIFAST :12809-13119): a_s = sum Z E_CD G,
          // b_s = sum Z E_CD conj(G) for the four (E_CD, G) pairings
          T ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
          T br[4] = {0, 0, 0, 0}, bi[4] = {0, 0, 0, 0};
          if (a.one_comp) {
            // one nonzero component of E(CD) and of G per parity class of
            // t' (ce, cg): one product per t', summed class by class into
            // P[ce][cg] = sum E_ce G_cg, from which the sums below are
            // formed; four t' at a time, their loads issued together
            T pp[4][4];
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
              for (int y = 0; y < 4; ++y) pp[x][y] = 0;
            for (int c = 0; c < 8; ++c) {
              const int ce = (int)((cd.ccomp >> (4 * c)) & 15);
              const int cg = (int)((ab.gcomp >> (4 * c)) & 15);
              // readlane returns an int: both words through uint32_t, or the
              // low word's bit 31 would be sign-extended over the high word
              uint64_t m = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(okh, c) << 32 |
                           (uint32_t)__builtin_amdgcn_readlane(okl, c);
              if (ce > 3 || cg > 3 || m == 0) continue;
              const int *mem = c_cmem[sym][c];
              T sum = 0;
              while (m) {
                T ev[4], gv[4];
                bool ok[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                  // the next term's t', or (none left) the class's first
                  // again, loaded but not summed
                  ok[i] = m != 0;
                  const int tq = mem[ok[i] ? __builtin_ctzll(m) : 0];
                  m &= m - 1;
                  ldcd1(a, cd.eoff + (int64_t)tq * maxcd + kl, ev[i]);
                  gv[i] = Gg[(int64_t)(4 * tq + cg) * ld + kl];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) sum += ok[i] ? ev[i] * gv[i] : (T)0;
              }
#pragma unroll
              for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
                  if (ce == x && cg == y) pp[x][y] += sum;
            }
            // set s = (E11 or E21) x (G11 or G21): components 0, 1 (2, 3) are
            // the real and imaginary parts of E11 (E21), G11 (G21)
#pragma unroll
            for (int st = 0; st < 4; ++st) {
              const int x = 2 * (st & 1), y = 2 * (st >> 1);
              ar[st] = pp[x][y] - pp[x + 1][y + 1];
              ai[st] = pp[x][y + 1] + pp[x + 1][y];
              br[st] = pp[x][y] + pp[x + 1][y + 1];
              bi[st] = pp[x + 1][y] - pp[x][y + 1];
            }
          } else
  #pragma unroll 4
          for (int tp = 0; tp < cd.ntcd; ++tp) {
            // masked E(CD) times Z (fold_kernel)
            T e11r, e11i, e21r, e21i;
            ldcd(a, cd.eoff + (int64_t)tp * maxcd + kl, e11r, e11i, e21r, e21i);
            const T *gp = Gg + (int64_t)(4 * tp) * ld + kl;
            // components masked off in the first contraction were stored as
            // zeros: not read again
            const uint8_t *gmk = &sh.gm[4 * (g * ntg + tp)];
            T g11r = gmk[0] ? gp[0] : (T)0, g11i = gmk[1] ? gp[ld] : (T)0;
            T g21r = gmk[2] ? gp[2 * ld] : (T)0, g21i = gmk[3] ? gp[3 * ld] : (T)0;
            // set 0: E11cd x G11, set 1: E21cd x G11, set 2: E11cd x G21,
            // set 3: E21cd x G21
            ar[0] += e11r * g11r - e11i * g11i; ai[0] += e11r * g11i + e11i * g11r;
            br[0] += e11r * g11r + e11i * g11i; bi[0] += e11i * g11r - e11r * g11i;
            ar[1] += e21r * g11r - e21i * g11i; ai[1] += e21r * g11i + e21i * g11r;
            br[1] += e21r * g11r + e21i * g11i; bi[1] += e21i * g11r - e21r * g11i;
            ar[2] += e11r * g21r - e11i * g21i; ai[2] += e11r * g21i + e11i * g21r;
            br[2] += e11r * g21r + e11i * g21i; bi[2] += e11i * g21r - e11r * g21i;
            ar[3] += e21r * g21r - e21i * g21i; ai[3] += e21r * g21i + e21i * g21r;
            br[3] += e21r * g21r + e21i * g21i; bi[3] += e21i * g21r - e21r * g21i;
          }
          // raw sums in FP64; RR(n) is formed where the digestion needs it
          Raw raw;
          for (int k = 0; k < 4; ++k) {
            raw.ar[k] = ar[k]; raw.ai[k] = ai[k]; raw.br[k] = br[k]; raw.bi[k] = bi[k];
          }
          raw.pab = pab1;
          raw.pabcd = pabcd;
          if (a.rr_out) {
  #pragma unroll
            for (int n = 0; n < 16; ++n) {
              double re, im;
              if (a.isym != 0) rrv<true>(raw, n, re, im); else rrv<false>(raw, n, re, im);
              a.rr_out[(int64_t)n * maxcd + (kl % nC) * nD + kl / nC] = make_double2(re, im);
            }
            continue;
          }
          // digestion (CLMMTG / CLMMTZ)
          const int idx[4] = {ib, jb, kl % nC, kl / nC};
          const int off[4][2] = {{ab.fa[0], ab.fa[1]}, {ab.fb[0], ab.fb[1]},
                                 {cd.fc[0], cd.fc[1]}, {cd.fd[0], cd.fd[1]}};
          int flg = cd.iflg;
          // ITOG: screening only when an SS pair is involved (ITT > 1, :12096)
          if (a.schwarz && (ab.fa[0] >= a.nskp || cd.fc[0] >= a.nskp))
            flg &= schwarz_mask(a.gslot, a.dslot, nd, off, idx, a.sens, s_gab, s_dab, a.schwarz == 2);
          if (a.isym != 0)
            digest<true, W == 2>(raw, off, idx, Dm, nd, flg, a.intsym, pab1, cd.pab[1], pcd1,
                         cd.pcd[1], dab, accAB, xv, a.gdir);
          else
            digest<false, W == 2>(raw, off, idx, Dm, nd, flg, a.intsym, pab1, cd.pab[1], pcd1,
                          cd.pcd[1], dab, accAB, xv, a.gdir);
        }
        if (!a.rr_out) {
          const int offs[4][2] = {{ab.fa[0], ab.fa[1]}, {ab.fb[0], ab.fb[1]},
                                  {cd.fc[0], cd.fc[1]}, {cd.fd[0], cd.fd[1]}};
          const int xf = a.intsym ? cd.iflg : (cd.iflg & 0x11);
          if (a.isym != 0)
            flush_exchange<true>(xv, xf, offs, ib, jb, a.gxch, nd, lane);
          else
            flush_exchange<false>(xv, xf, offs, ib, jb, a.gxch, nd, lane);
        }
      }
      if (!a.rr_out) {
        // batches 1 and 2 write G(A sA, B sB)(i,j): reduce over the wave
        for (int s = 0; s < 4; ++s) {
          double re = wave_sum(accAB[s][0]);
          double im = wave_sum(accAB[s][1]);
          if (lane == 0 && (re != 0.0 || im != 0.0)) {
            int orr = ab.fa[s >> 1] + ib, occ = ab.fb[s & 1] + jb;
            atomic_add2(&a.gdir[(int64_t)occ * nd + orr], re, im);
          }
        }
      }
      __syncthreads();
    }
    a0 += na;
  }
}

// Persistent work-queue kernel: each 64-lane workgroup takes (class, ij)
// items from a global counter until the list is exhausted.  With the Schwarz
// split on, an item whose largest possible quartet contribution
// qitem * max|D| is below the split threshold belongs to the FP32 instance
// of the kernel, every other item to the FP64 instance; the two instances
// keep their own counters and register allocations.  W waves per SIMD:
// three on the complex path where the plan's LDS lets twelve workgroups
// share a CU, else two with more loads in flight per wave and, where eight
// workgroups still fit, C = 2 (fock_launch_config).
template <typename T, int W, int C>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, W))) fock_kernel(FockArgs a) {
  __shared__ FockShared<C> sh;
  __shared__ int s_item;
  extern __shared__ double xacc[];  // a.xacc_n doubles (launch_fock)
  constexpr bool kLow = sizeof(T) == sizeof(float);
  const int lane = threadIdx.x;
  double *slab = a.slab + (int64_t)blockIdx.x * a.slab_stride;
  int *counter = a.counter + (kLow ? 1 : 0);
  for (int w = lane; w < a.xacc_n; w += 64) xacc[w] = 0.0;
  const double dmax = a.mixed ? __longlong_as_double(*a.dmax_bits) : 0.0;
  __syncthreads();
  for (;;) {
    if (lane == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    int idx = s_item;
    __syncthreads();
    if (idx >= a.nitems) break;
    const bool low = a.mixed && a.qitem[idx] * dmax < a.split;
    if (low != kLow) continue;
    fock_item<T, W, C>(a, a.items[idx], slab, sh, xacc, lane);
  }
}

// ---------------------------------------------------------------------------
// Breit (BRTFAST / BIIFAST / BRTMTG / BRTMTZ)
// ---------------------------------------------------------------------------
// R-integrals R[col][kl] (before the prefactor) of primitive pair ij against
// every kl of a class, to order L
__device__ void r_columns(const ClassDesc &cl, const double *exps, int ij, double *R, double *F,
                          int ld, int L, int lane) {
  const int nB = cl.nbas[1], nC = cl.nbas[2], nD = cl.nbas[3];
  const int ib = ij / nB, jb = ij % nB;
  const double ea = exps[cl.exp[0] + ib], eb = exps[cl.exp[1] + jb];
  const double p = ea + eb;
  const double Px = (cl.xyz[0][0] * ea + cl.xyz[1][0] * eb) / p;
  const double Py = (cl.xyz[0][1] * ea + cl.xyz[1][1] * eb) / p;
  const double Pz = (cl.xyz[0][2] * ea + cl.xyz[1][2] * eb) / p;
  for (int kl = lane; kl < cl.maxcd; kl += 64) {
    int kb = kl % nC, lb = kl / nC;
    double ec = exps[cl.exp[2] + kb], ed = exps[cl.exp[3] + lb];
    double q = ec + ed;
    double Qx = (cl.xyz[2][0] * ec + cl.xyz[3][0] * ed) / q;
    double Qy = (cl.xyz[2][1] * ec + cl.xyz[3][1] * ed) / q;
    double Qz = (cl.xyz[2][2] * ec + cl.xyz[3][2] * ed) / q;
    double alpha = p * q / (p + q);
    double X = Px - Qx, Y = Py - Qy, Z = Pz - Qz;
    boys_ref<double>(alpha * (X * X + Y * Y + Z * Z), L, F + kl, ld);
    rints<double, double>(R + kl, F + kl, ld, L, alpha, X, Y, Z);
  }
}

// column test of ERIFAST/BIIFAST: sum over kl of |R| > 1e-10 (:18465-18474)
__device__ void r_flags(const double *R, int ld, int nt, int maxcd, uint8_t *flag, int lane) {
  for (int col = lane; col < nt; col += 64) {
    double s = 0.0;
    const double *row = R + (int64_t)col * ld;
    for (int kl = 0; kl < maxcd; ++kl) s += fabs(row[kl]);
    flag[col] = s > 1.0e-10;
  }
}

// BIIFAST's gauge term reads IRC, the column flags of the most recent R
// batch.  Inside a class the first (kappa, |m|) combination generates the R
// batch of every ij in turn, so every later combination sees the flags of
// the class's last pair ij = MAXAB: computed here, one class per pass.
__global__ void __launch_bounds__(64) breit_flags_kernel(BreitArgs a) {
  const int lane = threadIdx.x;
  double *R = a.slab + (int64_t)blockIdx.x * a.slab_stride;
  for (int c = blockIdx.x; c < a.nclasses; c += gridDim.x) {
    const ClassDesc cl = a.classes[c];
    const int ld = cl.maxcd;
    double *F = R + (int64_t)ntuv(kMaxLfl) * ld;
    r_columns(cl, a.exps, cl.nbas[0] * cl.nbas[1] - 1, R, F, ld, cl.lamfl, lane);
    __syncthreads();
    r_flags(R, ld, ntuv(cl.lamfl), cl.maxcd, a.stale + cl.flagoff, lane);
    __syncthreads();
  }
}

// static LDS of a Breit workgroup contracting GB AB combinations together;
// the exchange accumulators (XView, BreitArgs::xacc_n doubles) follow in
// dynamic LDS, sized by the plan's largest class
template <int GB>
struct BreitShared {
  uint8_t cflag[ntuv(kMaxLfl)];
  uint8_t gflag[ntuv(kMaxLfl)];
  uint8_t gm[GB][ntuv(2 * kMaxL + 1) * 12];
};

__device__ __forceinline__ int any4(const double *e) {
  return e[0] != 0.0 || e[1] != 0.0 || e[2] != 0.0 || e[3] != 0.0;
}

// One work item (class, ij) of the Breit build: R to order L_AB + L_CD + 2,
// G^ix(kl, tuv') = sum_tuv [E^ix R(tuv+tuv') - sum_jx E^jx T_ix,jx] with the
// gauge term T = R(+e_ix+e_jx)/(2 alpha) - PQ_jx R(+e_ix) +
// (t+t')_jx R(+e_ix-e_jx) (BIIFAST :18492-18810), RR = sum_ix E^ix_CD G^ix
// (:18830-19150) and the exchange digestion into BXCH.
template <int kBreitGroup>
__device__ void breit_item(const BreitArgs &a, const WorkItem &it, double *slab,
                           BreitShared<kBreitGroup> &sh, double *xacc, int lane) {
  const ClassDesc &cl = a.classes[it.cls];
  const int ld = cl.maxcd;  // slab rows packed at the class's kl count
  double *R = slab;
  double *F = R + (int64_t)ntuv(kMaxLfl) * ld;
  double *G = F + (int64_t)(kMaxLfl + 1) * ld;
  const double2 *Dm = a.dent;
  const int nd = a.ndim;
  const int nB = cl.nbas[1], nC = cl.nbas[2], nD = cl.nbas[3];
  const XView xv{xacc, cl.nbas[2], nD};
  const int ib = it.ij / nB, jb = it.ij % nB;
  const double ea = a.exps[cl.exp[0] + ib], eb = a.exps[cl.exp[1] + jb];
  const double p = ea + eb;
  const double Px = (cl.xyz[0][0] * ea + cl.xyz[1][0] * eb) / p;
  const double Py = (cl.xyz[0][1] * ea + cl.xyz[1][1] * eb) / p;
  const double Pz = (cl.xyz[0][2] * ea + cl.xyz[1][2] * eb) / p;
  const int L = cl.lamfl, nt = ntuv(L), maxcd = cl.maxcd;
  const int maxab = cl.nbas[0] * nB;
  const double bfac = a.gaunt ? -1.0 : -0.5;
  const int ijs = __builtin_amdgcn_readfirstlane(it.ij);  // uniform: E(AB) by scalar loads
  const int sym = cl.sym;

  r_columns(cl, a.exps, it.ij, R, F, ld, L, lane);
  __syncthreads();
  r_flags(R, ld, nt, maxcd, sh.cflag, lane);
  __syncthreads();
  // RCTTFL = BFAC * PRE * R for every column; the flags are applied where
  // the reference tests them
  for (int kl = lane; kl < maxcd; kl += 64) {
    int kb = kl % nC, lb = kl / nC;
    double q = a.exps[cl.exp[2] + kb] + a.exps[cl.exp[3] + lb];
    double pre = bfac * (2.0 * 1.7493418327624863e1 / (sqrt(p + q) * p * q));
    for (int col = 0; col < nt; ++col) R[(int64_t)col * ld + kl] *= pre;
  }
  __syncthreads();

  // AB combinations in groups of up to kBreitGroup: every combination of a
  // class has the same N(L_AB) and N(L_CD), reads the same R values and the
  // same gauge factors T, and differs only in E(AB).  The class's first
  // combination reads the fresh R-column flags (BIIFAST's IRC of this pair),
  // every later one the class's stale flags, so groups never mix the two.
  for (int a0 = it.ab_lo; a0 < it.ab_hi;) {
    const int na = a0 == cl.ab_begin ? 1 : min(kBreitGroup, it.ab_hi - a0);
    const int ntg = a.abs[a0].ntg, ntab = a.abs[a0].ntab;
    int64_t eoff[kBreitGroup][3];
#pragma unroll
    for (int g = 0; g < kBreitGroup; ++g)
#pragma unroll
      for (int ix = 0; ix < 3; ++ix) eoff[g][ix] = g < na ? a.abs[a0 + g].eoff[ix] : 0;
    for (int g = 0; g < na; ++g) {
      const int gmo = a.abs[a0 + g].gmask;
      for (int t = lane; t < 12 * ntg; t += 64) sh.gm[g][t] = a.masks[gmo + t];
    }
    const uint8_t *gsrc = a0 == cl.ab_begin ? sh.cflag : a.stale + cl.flagoff;
    for (int t = lane; t < nt; t += 64) sh.gflag[t] = gsrc[t];
    __syncthreads();
    for (int kl = lane; kl < maxcd; kl += 64) {
      int kb = kl % nC, lb = kl / nC;
      double ec = a.exps[cl.exp[2] + kb], ed = a.exps[cl.exp[3] + lb];
      double q = ec + ed;
      double pq[3] = {(cl.xyz[2][0] * ec + cl.xyz[3][0] * ed) / q - Px,
                      (cl.xyz[2][1] * ec + cl.xyz[3][1] * ed) / q - Py,
                      (cl.xyz[2][2] * ec + cl.xyz[3][2] * ed) / q - Pz};
      double half_ia = 0.5 / (p * q / (p + q));
      for (int tp = 0; tp < ntg; ++tp) {
        // the three Pauli components ix share every R value they read: R at
        // h = tuv + tuv', h + e_i, h + e_i + e_j and h + e_i - e_j
        bool act[kBreitGroup][3], any[3] = {false, false, false};
#pragma unroll
        for (int g = 0; g < kBreitGroup; ++g)
#pragma unroll
          for (int ix = 0; ix < 3; ++ix) {
            const uint8_t *gm = &sh.gm[g][(tp * 3 + ix) * 4];
            act[g][ix] = g < na && (gm[0] | gm[1] | gm[2] | gm[3]) != 0;
            any[ix] = any[ix] || act[g][ix];
          }
        double gv[kBreitGroup][3][4];
#pragma unroll
        for (int g = 0; g < kBreitGroup; ++g)
#pragma unroll
          for (int ix = 0; ix < 3; ++ix)
#pragma unroll
            for (int c = 0; c < 4; ++c) gv[g][ix][c] = 0.0;
        // On the class's symmetric axes (ClassDesc::sym) R(h') vanishes
        // exactly unless h' is even there, so with dl the parity class of
        // h = tuv + tuv' on those axes: the direct term of ix needs dl = 0
        // and the gauge term of (ix, jx) needs dl = (e_ix + e_jx) mod 2
        // (its PQ_jx R(h + e_ix) part is zero when jx is symmetric, where
        // PQ_jx = 0).  pairs[dl] holds the (ix, jx) bits 3 ix + jx that can
        // contribute to some combination of the group; a t whose dl has none
        // is skipped.
        int pairs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int ix = 0; ix < 3; ++ix)
#pragma unroll
          for (int jx = 0; jx < 3; ++jx)
            if (any[ix] && (!a.gaunt || ix == jx))
              pairs[((1 << ix) ^ (1 << jx)) & sym] |= 1 << (3 * ix + jx);
        if (any[0] || any[1] || any[2]) {
          const int *hs = c_hsum + tp * kHsRow;
          for (int t = 0; t < ntab; ++t) {
            const int h = hs[t];
            const int dl = hpar(h) & sym;
            const int pm = pairs[dl];
            if (pm == 0) continue;
            const bool fh = dl == 0 && sh.cflag[h] != 0;
            const double r0 = R[(int64_t)h * ld + kl];
            // E(AB) of the three components at t, EPSAB applied (the same
            // address in every lane: scalar loads)
            double e3[kBreitGroup][3][4];
#pragma unroll
            for (int g = 0; g < kBreitGroup; ++g)
#pragma unroll
              for (int ix = 0; ix < 3; ++ix) {
                double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
                if (g < na) v = a.Eab[eoff[g][ix] + (int64_t)t * maxab + ijs];
                e3[g][ix][0] = v.x; e3[g][ix][1] = v.y; e3[g][ix][2] = v.z; e3[g][ix][3] = v.w;
              }
            if (fh) {
#pragma unroll
              for (int g = 0; g < kBreitGroup; ++g)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix) {
                  if (!act[g][ix]) continue;
#pragma unroll
                  for (int c = 0; c < 4; ++c) gv[g][ix][c] += e3[g][ix][c] * r0;
                }
            }
            if (a.gaunt) continue;
            // gauge term T = R(h+e_ix+e_jx)/(2 alpha) - PQ_jx R(h+e_ix) +
            // (t+t')_jx R(h+e_ix-e_jx), skipped where all three columns are
            // flagged off (BIIFAST :18492-18810); R values are loaded only
            // for the pairs in pm
            const uint32_t s3 = c_tuv[h];
            const int sj[3] = {(int)(s3 >> 10), (int)((s3 >> 5) & 31), (int)(s3 & 31)};
            int hp[3], hpp[3][3], hpm[3][3];
            double rp[3], rpp[3][3], rpm[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              hp[i] = c_hstep[h][i];
              // the middle part of (i, j) for a non-symmetric j
              const int need = pm & (7 << (3 * i)) & ~(sym * (1 << (3 * i)));
              rp[i] = need ? R[(int64_t)hp[i] * ld + kl] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) {
                if (j < i) {
                  hpp[i][j] = hpp[j][i];
                  rpp[i][j] = rpp[j][i];
                } else {
                  hpp[i][j] = c_hstep[hp[i]][j];
                  const bool need = (pm >> (3 * i + j) & 1) || (pm >> (3 * j + i) & 1);
                  rpp[i][j] = need ? R[(int64_t)hpp[i][j] * ld + kl] : 0.0;
                }
                if (i == j) {
                  hpm[i][j] = h;
                  rpm[i][j] = r0;
                } else {
                  hpm[i][j] = c_hstep[hp[i]][4 + j];  // -1 where (t+t')_j = 0
                  const bool need = (pm >> (3 * i + j) & 1) && hpm[i][j] >= 0;
                  rpm[i][j] = need ? R[(int64_t)hpm[i][j] * ld + kl] : 0.0;
                }
              }
#pragma unroll
            for (int ix = 0; ix < 3; ++ix) {
#pragma unroll
              for (int jx = 0; jx < 3; ++jx) {
                if (!(pm >> (3 * ix + jx) & 1)) continue;
                const bool ok3 = hpm[ix][jx] >= 0;
                int fl = sh.gflag[hpp[ix][jx]] + sh.gflag[hp[ix]] + (ok3 ? sh.gflag[hpm[ix][jx]] : 0);
                if (fl == 0) continue;
                double tv = rpp[ix][jx] * half_ia - rp[ix] * pq[jx];
                if (ok3) tv += rpm[ix][jx] * (double)sj[jx];
#pragma unroll
                for (int g = 0; g < kBreitGroup; ++g) {
                  const double *e = e3[g][jx];
                  if (!act[g][ix] || !any4(e)) continue;
#pragma unroll
                  for (int c = 0; c < 4; ++c) gv[g][ix][c] -= e[c] * tv;
                }
              }
            }
          }
        }
#pragma unroll
        for (int g = 0; g < kBreitGroup; ++g) {
          if (g >= na) break;
#pragma unroll
          for (int ix = 0; ix < 3; ++ix) {
            const uint8_t *gm = &sh.gm[g][(tp * 3 + ix) * 4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
              G[(int64_t)(((g * 3 + ix) * 4 + c) * ntg + tp) * ld + kl] = gm[c] ? gv[g][ix][c] : 0.0;
          }
        }
      }
    }
    for (int g = 0; g < na; ++g) {
      const BrAB ab = a.abs[a0 + g];
      const double *Gg = G + (int64_t)(g * 12 * ntg) * ld;
      for (int ci = ab.cd_begin; ci < ab.cd_end; ++ci) {
        const BrCD cd = a.cds[ci];
        const double pab1 = cd.pab[0], pcd1 = cd.pcd[0];
        const double pabcd = pab1 * pcd1;
        for (int kl = lane; kl < maxcd; kl += 64) {
          double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
          double br[4] = {0, 0, 0, 0}, bi[4] = {0, 0, 0, 0};
          for (int ix = 0; ix < 3; ++ix)
            for (int tp = 0; tp < cd.ntcd; ++tp) {
              // masked E(CD) times Z (fold_kernel; EILSCD0's flags)
              const double4 e = a.Ecd[cd.eoff[ix] + (int64_t)tp * maxcd + kl];
              const double e11r = e.x, e11i = e.y, e21r = e.z, e21i = e.w;
              const double *gp = Gg + (int64_t)(ix * 4 * ntg + tp) * ld + kl;
              const int64_t gs = (int64_t)ntg * ld;
              double g11r = gp[0], g11i = gp[gs], g21r = gp[2 * gs], g21i = gp[3 * gs];
              ar[0] += e11r * g11r - e11i * g11i; ai[0] += e11r * g11i + e11i * g11r;
              br[0] += e11r * g11r + e11i * g11i; bi[0] += e11i * g11r - e11r * g11i;
              ar[1] += e21r * g11r - e21i * g11i; ai[1] += e21r * g11i + e21i * g11r;
              br[1] += e21r * g11r + e21i * g11i; bi[1] += e21i * g11r - e21r * g11i;
              ar[2] += e11r * g21r - e11i * g21i; ai[2] += e11r * g21i + e11i * g21r;
              br[2] += e11r * g21r + e11i * g21i; bi[2] += e11i * g21r - e11r * g21i;
              ar[3] += e21r * g21r - e21i * g21i; ai[3] += e21r * g21i + e21i * g21r;
              br[3] += e21r * g21r + e21i * g21i; bi[3] += e21i * g21r - e21r * g21i;
            }
          Raw raw;
          for (int k = 0; k < 4; ++k) {
            raw.ar[k] = ar[k]; raw.ai[k] = ai[k]; raw.br[k] = br[k]; raw.bi[k] = bi[k];
          }
          raw.pab = pab1;
          raw.pabcd = pabcd;
          if (a.rr_out) {
            if (ci == 0) {
#pragma unroll
              for (int n = 0; n < 16; ++n) {
                double re, im;
                if (a.isym != 0) rrv<true>(raw, n, re, im); else rrv<false>(raw, n, re, im);
                a.rr_out[(int64_t)n * maxcd + (kl % nC) * nD + kl / nC] = make_double2(re, im);
              }
            }
            continue;
          }
          const int idx[4] = {ib, jb, kl % nC, kl / nC};
          const int off[4][2] = {{ab.fa[0], ab.fa[1]}, {ab.fb[0], ab.fb[1]},
                                 {cd.fc[0], cd.fc[1]}, {cd.fd[0], cd.fd[1]}};
          // exchange batches only: the direct batches' accumulators and D(A,B) stay unused
          double acc_unused[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
          const double2 dab_unused[2][4] = {};
          if (a.isym != 0)
            digest<true, true>(raw, off, idx, Dm, nd, cd.iflg, a.intsym, pab1, cd.pab[1], pcd1,
                               cd.pcd[1], dab_unused, acc_unused, xv, a.bdir);
          else
            digest<false, true>(raw, off, idx, Dm, nd, cd.iflg, a.intsym, pab1, cd.pab[1], pcd1,
                                cd.pcd[1], dab_unused, acc_unused, xv, a.bdir);
        }
        if (!a.rr_out) {
          const int offs[4][2] = {{ab.fa[0], ab.fa[1]}, {ab.fb[0], ab.fb[1]},
                                  {cd.fc[0], cd.fc[1]}, {cd.fd[0], cd.fd[1]}};
          const int xf = a.intsym ? cd.iflg : (cd.iflg & 0x11);
          if (a.isym != 0)
            flush_exchange<true>(xv, xf, offs, ib, jb, a.bxch, nd, lane);
          else
            flush_exchange<false>(xv, xf, offs, ib, jb, a.bxch, nd, lane);
        }
      }
    }
    __syncthreads();
    a0 += na;
  }
}

// W waves per SIMD: the register budget 512 / W (at W = 4 the compiler
// spills some registers to scratch; the occupancy still pays, the kernel
// being latency-bound on its R-slab loads)
template <int GB, int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W, W)))
breit_kernel(BreitArgs a) {
  __shared__ BreitShared<GB> sh;
  __shared__ int s_item;
  extern __shared__ double xacc[];  // a.xacc_n doubles (launch_breit)
  const int lane = threadIdx.x;
  double *slab = a.slab + (int64_t)blockIdx.x * a.slab_stride;
  for (int w = lane; w < a.xacc_n; w += 64) xacc[w] = 0.0;
  __syncthreads();
  for (;;) {
    if (lane == 0) s_item = atomicAdd(a.counter, 1);
    __syncthreads();
    int idx = __builtin_amdgcn_readfirstlane(s_item);  // the same in every lane
    __syncthreads();
    if (idx >= a.nitems) break;
    breit_item<GB>(a, a.items[idx], slab, sh, xacc, lane);
  }
}

// BRTFAST :17646-17664: the SL blocks are the conjugate transpose of LS
__global__ void breit_complete_kernel(double2 *bdir, double2 *bxch, int ndim, int nskp) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t n = (int64_t)nskp * (ndim - nskp);
  if (g >= n) return;
  int i = (int)(g % nskp), j = nskp + (int)(g / nskp);
  double2 v = bdir[(int64_t)j * ndim + i];
  bdir[(int64_t)i * ndim + j] = make_double2(v.x, -v.y);
  v = bxch[(int64_t)j * ndim + i];
  bxch[(int64_t)i * ndim + j] = make_double2(v.x, -v.y);
}

// ---------------------------------------------------------------------------
// One-electron matrices (OVRLP :6683, ONEEL :6996)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// slots 12 = -PHS conj(21), 22 = PHS conj(11)
__device__ __forceinline__ void slots4(double2 (&v)[4], double phs) {
  v[1] = make_double2(-phs * v[2].x, phs * v[2].y);
  v[3] = make_double2(phs * v[0].x, -phs * v[0].y);
}

// a 2x2 slot block at rows r[s] + i, columns c[s] + j and its mirror, in the
// reference's statement order
__device__ void put4(double2 *M, int n, const int (&r)[2], const int (&c)[2], int i, int j,
                     const double2 (&v)[4]) {
  M[(int64_t)(c[0] + j) * n + r[0] + i] = v[0];
  M[(int64_t)(c[1] + j) * n + r[0] + i] = v[1];
  M[(int64_t)(c[0] + j) * n + r[1] + i] = v[2];
  M[(int64_t)(c[1] + j) * n + r[1] + i] = v[3];
  // mirrors read the element back, so where a block meets its own mirror
  // the later statement wins exactly as in the reference
  M[(int64_t)(r[0] + i) * n + c[0] + j] = conj2(M[(int64_t)(c[0] + j) * n + r[0] + i]);
  M[(int64_t)(r[0] + i) * n + c[1] + j] = conj2(M[(int64_t)(c[1] + j) * n + r[0] + i]);
  M[(int64_t)(r[1] + i) * n + c[0] + j] = conj2(M[(int64_t)(c[0] + j) * n + r[1] + i]);
  M[(int64_t)(r[1] + i) * n + c[1] + j] = conj2(M[(int64_t)(c[1] + j) * n + r[1] + i]);
}

// -sum over nuclei of PNC E(tuv) R(tuv) for the 11 and 21 entries
__device__ void vnuc1(const OneArgs &a, int64_t eoff, int maxab, int ij, int lam, double p,
                      double px, double py, double pz, double2 &v11, double2 &v21) {
  double f[kMaxLab + 1], r[ntuv(kMaxLab)];
  int nt = ntuv(lam);
  for (int iz = 0; iz < a.ncnt; ++iz) {
    double xi = a.xi[iz], z = a.znuc[iz];
    double alpha, pnc;
    if (xi <= 0.0) {
      alpha = p;
      pnc = 2.0 * z * 3.1415926535897932 / p;
    } else {
      double esm = p + xi;
      alpha = p * xi / esm;
      pnc = 2.0 * 3.1415926535897932 * z * 1.0 * sqrt(xi / esm) / p;
    }
    double X = px - a.xyz[3 * iz], Y = py - a.xyz[3 * iz + 1], Z = pz - a.xyz[3 * iz + 2];
    boys_ref<double>(alpha * (X * X + Y * Y + Z * Z), lam, f, 1);
    rints<double, double>(r, f, 1, lam, alpha, X, Y, Z);
    for (int t = 0; t < nt; ++t) {
      double4 e = a.E[eoff + (int64_t)t * maxab + ij];
      v11.x -= pnc * e.x * r[t];
      v11.y -= pnc * e.y * r[t];
      v21.x -= pnc * e.z * r[t];
      v21.y -= pnc * e.w * r[t];
    }
  }
}

__global__ void __launch_bounds__(64) oneel_kernel(OneArgs a) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nitems) return;
  int lo = 0, hi = a.njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.first[mid] <= g) lo = mid; else hi = mid - 1;
  }
  const OneJob J = a.jobs[lo];
  const int ij = (int)(g - a.first[lo]);
  const int i = ij / J.nB, j = ij % J.nB;
  const int maxab = J.nA * J.nB;
  const bool diag = J.il[0] == J.jl[0];
  const double ea = a.exps[J.expA + i], eb = a.exps[J.expB + j];
  const double p = ea + eb;
  const double rt = sqrt(3.1415926535897932 / p), eroot = rt * rt * rt;
  const double px = (J.ax * ea + J.bx * eb) / p, py = (J.ay * ea + J.by * eb) / p,
               pz = (J.az * ea + J.bz * eb) / p;
  const int n = a.ndim, ns = a.nskp;
  const int il[2] = {J.il[0], J.il[1]}, jl[2] = {J.jl[0], J.jl[1]};
  const int is[2] = {J.il[0] + ns, J.il[1] + ns}, js[2] = {J.jl[0] + ns, J.jl[1] + ns};
  double4 ell = a.E[J.ell + ij], ess = a.E[J.ess + ij];
  double2 sll[4], sss[4], vll[4], vss[4], tls[4], tsl[4];
  sll[0] = make_double2(eroot * ell.x, eroot * ell.y);
  sll[2] = make_double2(eroot * ell.z, eroot * ell.w);
  slots4(sll, J.phs);
  sss[0] = make_double2(eroot * ess.x, eroot * ess.y);
  sss[2] = make_double2(eroot * ess.z, eroot * ess.w);
  slots4(sss, J.phs);
  vll[0] = vll[2] = make_double2(0.0, 0.0);
  vnuc1(a, J.ell, maxab, ij, J.la + J.lb, p, px, py, pz, vll[0], vll[2]);
  slots4(vll, J.phs);
  vss[0] = vss[2] = make_double2(0.0, 0.0);
  vnuc1(a, J.ess, maxab, ij, J.la + J.lb + 2, p, px, py, pz, vss[0], vss[2]);
  const double cv2 = 2.0 * a.cv * a.cv;  // 2 m c^2, m = 1
  vss[0].x -= cv2 * sss[0].x; vss[0].y -= cv2 * sss[0].y;
  vss[2].x -= cv2 * sss[2].x; vss[2].y -= cv2 * sss[2].y;
  slots4(vss, J.phs);
  // kinetic coupling (ONEEL :7820-7880): SL from the SS overlap of (A,B);
  // LS from that of (B,A), conjugated with slots 12 and 21 exchanged
  const double ejrt = a.cv * sqrt((double)(2 * J.lb + 3)) * sqrt(eb);
  tsl[0] = make_double2(ejrt * eroot * ess.x, ejrt * eroot * ess.y);
  tsl[2] = make_double2(ejrt * eroot * ess.z, ejrt * eroot * ess.w);
  slots4(tsl, J.phs);
  double4 eba = a.E[J.essba + j * J.nA + i];
  const double eirt = a.cv * sqrt((double)(2 * J.la + 3)) * sqrt(ea);
  double2 t[4];
  t[0] = make_double2(eirt * eroot * eba.x, eirt * eroot * eba.y);
  t[2] = make_double2(eirt * eroot * eba.z, eirt * eroot * eba.w);
  slots4(t, J.phs);
  tls[0] = conj2(t[0]);
  tls[1] = conj2(t[2]);
  tls[2] = conj2(t[1]);
  tls[3] = conj2(t[3]);
  if (!diag || i >= j) {
    put4(a.ovlp, n, il, jl, i, j, sll);
    put4(a.hnuc, n, il, jl, i, j, vll);
    put4(a.ovlp, n, is, js, i, j, sss);
    put4(a.hnuc, n, is, js, i, j, vss);
  }
  put4(a.hkin, n, il, js, i, j, tls);
  if (!diag) put4(a.hkin, n, is, jl, i, j, tsl);
}

// largest |D| of the density (non-negative doubles order like their bit
// patterns, so an integer max is exact)
__global__ void dmax_kernel(const double2 *d, int64_t n, unsigned long long *out) {
  double m = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, hypot(d[g].x, d[g].y));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// Hermitian completion of GDIR/GXCH (CLMFAST :12166-12205): thread per
// (i,j), i <= j < nskp, skipping pairs inside one (centre, kappa, |mj|) block.
__global__ void complete_kernel(CompleteArgs a) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int n = a.nskp;
  if (g >= (int64_t)n * n) return;
  int i = (int)(g % n), j = (int)(g / n);
  if (i > j) return;
  // INTSYM off: every quartet was digested directly and nothing is completed
  // (:12174-12177 fall through to 401); on: diagonal sub-blocks are complete
  if (!a.intsym ||
      (a.icn[i] == a.icn[j] && a.kqn[i] == a.kqn[j] && abs(a.mqn[i]) == abs(a.mqn[j])))
    return;
  int ndim = a.ndim;
  int k = i + n, l = j + n;
  auto comp = [&](double2 *m, int r, int c) {
    double2 &x = m[(int64_t)c * ndim + r];
    double2 &y = m[(int64_t)r * ndim + c];
    x.x += y.x;
    x.y -= y.y;
    y.x = x.x;
    y.y = -x.y;
  };
  comp(a.gdir, i, j);
  comp(a.gxch, i, j);
  comp(a.gdir, k, l);
  comp(a.gxch, k, l);
  comp(a.gxch, i, l);
  comp(a.gxch, k, j);
}

__global__ void zero_kernel(double2 *a, int64_t n, double2 *b, int *counter,
                            unsigned long long *dmax) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) {
    a[g] = make_double2(0.0, 0.0);
    if (b) b[g] = make_double2(0.0, 0.0);
  }
  if (g == 0 && counter) {
    counter[0] = 0;  // FP64 (and Breit) work queue
    counter[1] = 0;  // FP32 work queue
  }
  if (g == 0 && dmax) *dmax = 0ull;
}

#include "breit_tc.inc"

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t upload_constants() {
  uint32_t tuv[ntuv(kMaxLfl)];
  // Hermite order: by L, then t ascending, then u ascending (CARTIND :2678-2700)
  int n = 0;
  for (int L = 0; L <= kMaxLfl; ++L)
    for (int t = 0; t <= L; ++t)
      for (int u = 0; u <= L - t; ++u) {
        int v = L - t - u;
        tuv[n++] = (uint32_t)((t << 10) | (u << 5) | v);
      }
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tuv), tuv, sizeof(tuv));
  if (e != hipSuccess) return e;
  static int hs[ntuv(kMaxLab) * kHsRow];
  for (int tp = 0; tp < ntuv(kMaxLab); ++tp)
    for (int t = 0; t < kHsRow; ++t) {
      int v = 0;
      if (t < ntuv(kMaxLab)) {
        const uint32_t a = tuv[t], b = tuv[tp];
        const int T = (int)((a >> 10) + (b >> 10)), U = (int)(((a >> 5) & 31) + ((b >> 5) & 31));
        const int V = (int)((a & 31) + (b & 31)), L = T + U + V;
        v = L * (L + 1) * (L + 2) / 6 + T * (L + 1) - T * (T - 1) / 2 + U;
      }
      hs[tp * kHsRow + t] = v;
    }
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_hsum), hs, sizeof(hs));
  if (e != hipSuccess) return e;
  static int st[ntuv(kMaxLfl)][8];
  auto index_of = [](int T, int U, int V) {
    const int L = T + U + V;
    return L * (L + 1) * (L + 2) / 6 + T * (L + 1) - T * (T - 1) / 2 + U;
  };
  for (int h = 0; h < ntuv(kMaxLfl); ++h) {
    const int c[3] = {(int)(tuv[h] >> 10), (int)((tuv[h] >> 5) & 31), (int)(tuv[h] & 31)};
    for (int d = 0; d < 3; ++d) {
      int up[3] = {c[0], c[1], c[2]}, dn[3] = {c[0], c[1], c[2]};
      up[d] += 1;
      dn[d] -= 1;
      st[h][d] = c[0] + c[1] + c[2] + 1 <= kMaxLfl ? index_of(up[0], up[1], up[2]) : -1;
      st[h][4 + d] = dn[d] >= 0 ? index_of(dn[0], dn[1], dn[2]) : -1;
    }
    st[h][3] = st[h][7] = -1;
  }
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_hstep), st, sizeof(st));
  if (e != hipSuccess) return e;
  // parity-class orderings of the bra's Hermite indices (c_tperm, c_coff,
  // g_hsp): see their declarations
  {
    constexpr int NT = ntuv(kMaxLab);
    static int16_t perm[8][kMaxLab + 1][NT];
    static uint8_t coff[8][kMaxLab + 1][9];
    static int hsp[8][NT][kHsRow];
    auto par = [&](int n) {
      const uint32_t x = tuv[n];
      return (int)(((x >> 10) & 1) | (((x >> 5) & 1) << 1) | ((x & 1) << 2));
    };
    for (int m = 0; m < 8; ++m) {
      for (int L = 0; L <= kMaxLab; ++L) {
        const int nt = ntuv(L);
        int pos = 0;
        for (int c = 0; c < 8; ++c) {
          coff[m][L][c] = (uint8_t)pos;
          for (int t = 0; t < nt; ++t)
            if ((par(t) & m) == c) perm[m][L][pos++] = (int16_t)t;
        }
        coff[m][L][8] = (uint8_t)pos;
        for (int t = pos; t < NT; ++t) perm[m][L][t] = 0;
      }
      for (int tp = 0; tp < NT; ++tp) {
        const int c = par(tp) & m;
        int k = 0;
        for (int t = 0; t < NT; ++t)
          if ((par(t) & m) == c) hsp[m][tp][k++] = hs[tp * kHsRow + t];
        for (; k < kHsRow; ++k) hsp[m][tp][k] = 0;
      }
    }
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_tperm), perm, sizeof(perm));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_coff), coff, sizeof(coff));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(g_hsp), hsp, sizeof(hsp));
    if (e != hipSuccess) return e;
  }
  // parity-class ranks, sizes and class-0 members of all h < N(kMaxLfl)
  // (the matrix-core Breit kernel's compact tables)
  {
    static int16_t crank[8][ntuv(kMaxLfl)], cls0[8][ntuv(kMaxLfl)];
    static int16_t csz[8][kMaxLfl + 1][8];
    for (int m = 0; m < 8; ++m) {
      int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int h = 0; h < ntuv(kMaxLfl); ++h) {
        const uint32_t x = tuv[h];
        const int c = (int)(((x >> 10) & 1) | (((x >> 5) & 1) << 1) | ((x & 1) << 2)) & m;
        if (c == 0) cls0[m][cnt[0]] = (int16_t)h;
        crank[m][h] = (int16_t)cnt[c]++;
      }
      for (int h = cnt[0]; h < ntuv(kMaxLfl); ++h) cls0[m][h] = 0;
      for (int L = 0; L <= kMaxLfl; ++L) {
        int sz[8];
        hermite_class_sizes(m, L, sz);
        for (int c = 0; c < 8; ++c) csz[m][L][c] = (int16_t)sz[c];
      }
    }
    static int cmem[8][8][ntuv(kMaxLab)];
    for (int m = 0; m < 8; ++m) {
      int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int h = 0; h < ntuv(kMaxLab); ++h) {
        const uint32_t x = tuv[h];
        const int c = (int)(((x >> 10) & 1) | (((x >> 5) & 1) << 1) | ((x & 1) << 2)) & m;
        cmem[m][c][cnt[c]++] = h;
      }
    }
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_crank), crank, sizeof(crank));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_cmem), cmem, sizeof(cmem));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_cls0), cls0, sizeof(cls0));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_csz), csz, sizeof(csz));
    if (e != hipSuccess) return e;
    // closed form of R on the z axis: v! / (j! (v - 2j)! 2^j) and (t - 1)!!
    static double hz[kMaxLfl + 1][kMaxLfl / 2 + 1], dfac[kMaxLfl + 1];
    double fact[kMaxLfl + 1];
    fact[0] = 1.0;
    for (int n = 1; n <= kMaxLfl; ++n) fact[n] = fact[n - 1] * n;
    for (int v = 0; v <= kMaxLfl; ++v)
      for (int j = 0; j <= kMaxLfl / 2; ++j)
        hz[v][j] = 2 * j <= v ? fact[v] / (fact[j] * fact[v - 2 * j] * std::ldexp(1.0, j)) : 0.0;
    for (int t = 0; t <= kMaxLfl; ++t) {
      dfac[t] = 1.0;
      for (int k = t - 1; k > 0; k -= 2) dfac[t] *= k;
    }
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_hz), hz, sizeof(hz));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_dfac), dfac, sizeof(dfac));
    if (e != hipSuccess) return e;
  }
  double rcp[2 * kMaxLfl + 2 * 60 + 4];
  float rcpf[2 * kMaxLfl + 2 * 60 + 4];
  for (int k = 0; k < (int)(sizeof(rcp) / sizeof(double)); ++k) {
    rcp[k] = 1.0 / (2.0 * k + 1.0);
    rcpf[k] = (float)rcp[k];
  }
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_rcpf), rcpf, sizeof(rcpf));
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_rcp), rcp, sizeof(rcp));
}

hipError_t launch_etab(const EtabArgs &a, hipStream_t s) {
  if (a.nitems == 0) return hipSuccess;
  int bs = 128;
  int64_t nb = (a.nitems + bs - 1) / bs;
  hipLaunchKernelGGL(etab_kernel, dim3((unsigned)nb), dim3(bs), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_eflag(const EflagArgs &a, hipStream_t s) {
  if (a.npairs == 0) return hipSuccess;
  hipLaunchKernelGGL(eflag_kernel, dim3(a.npairs), dim3(64), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !a.Ecd) return e;
  hipLaunchKernelGGL(fold_kernel, dim3(a.npairs), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_zero(double2 *x, double2 *y, int64_t n, int *counter,
                       unsigned long long *dmax, hipStream_t s) {
  int bs = 256;
  int64_t nb = (n + bs - 1) / bs;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)nb), dim3(bs), 0, s, x, n, y, counter, dmax);
  return hipGetLastError();
}
__global__ void slot_tables_kernel(const double *G, const double2 *D, const int *partner, int nd,
                                   double *out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)nd * nd) return;
  const int c = (int)(g / nd), r = (int)(g - (int64_t)c * nd);
  const int rr[2] = {r, partner[r]}, cc[2] = {c, partner[c]};
  double x = 0.0;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const int64_t k = (int64_t)cc[w] * nd + rr[u];
      x = fmax(x, G ? G[k] : hypot(D[k].x, D[k].y));
    }
  out[g] = x;
}
hipError_t launch_slot_tables(const double *G, const double2 *D, const int *partner, int nd,
                              double *out, hipStream_t s) {
  const int64_t n = (int64_t)nd * nd;
  const int bs = 256;
  hipLaunchKernelGGL(slot_tables_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, G,
                     D, partner, nd, out);
  return hipGetLastError();
}
// E with ERIFAST's EPSAB cut (|e| <= 1e-16 -> 0, :12445; BIIFAST :18436-18470)
// applied once, for kernels that read E(AB) straight from memory
__global__ void eps_filter_kernel(const double4 *src, double4 *dst, int64_t n) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) {
    const double4 v = src[g];
    dst[g] = make_double4(fabs(v.x) > 1.0e-16 ? v.x : 0.0, fabs(v.y) > 1.0e-16 ? v.y : 0.0,
                          fabs(v.z) > 1.0e-16 ? v.z : 0.0, fabs(v.w) > 1.0e-16 ? v.w : 0.0);
  }
}
hipError_t launch_eps_filter(const double4 *src, double4 *dst, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  int bs = 256;
  hipLaunchKernelGGL(eps_filter_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, src,
                     dst, n);
  return hipGetLastError();
}
__global__ void to_f32_kernel(const double4 *src, float4 *dst, int64_t n) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) {
    double4 v = src[g];
    dst[g] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
  }
}
hipError_t launch_to_f32(const double4 *src, float4 *dst, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  int bs = 256;
  hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, src, dst,
                     n);
  return hipGetLastError();
}
hipError_t launch_dmax(const double2 *d, int64_t n, unsigned long long *out, hipStream_t s) {
  hipLaunchKernelGGL(dmax_kernel, dim3(256), dim3(256), 0, s, d, n, out);
  return hipGetLastError();
}
FockLaunch fock_launch_config(int xacc_n, int isym) {
  const size_t dyn = sizeof(int) + (size_t)xacc_n * sizeof(double);
  // three waves on the complex path where twelve workgroups fit; the real
  // path (ISYM 1, 2) digests half the terms, so its time goes to the first
  // contraction, which gains more from four R loads in flight than from a
  // third wave (TlF: 1.30 s with two waves, 1.33 s with three)
  if (isym == 0 && sizeof(FockShared<1>) + dyn <= 160 * 1024 / 12) return {3, 1};
  // two waves: double group capacity while eight workgroups still fit
  if (sizeof(FockShared<2>) + dyn <= 160 * 1024 / 8) return {2, 2};
  return {2, 1};
}

template <int W, int C>
static hipError_t launch_fock_w(const FockArgs &a, int nblocks, hipStream_t s, hipStream_t s32,
                                hipEvent_t fork, hipEvent_t join) {
  const size_t lds = a.xacc_n * sizeof(double);
  if (!a.mixed) {
    hipLaunchKernelGGL((fock_kernel<double, W, C>), dim3(nblocks), dim3(64), lds, s, a);
    return hipGetLastError();
  }
  // the FP32 instance runs beside the FP64 one on its own stream, with the
  // second half of the slab; s waits for both
  hipError_t e = hipEventRecord(fork, s);
  if (e == hipSuccess) e = hipStreamWaitEvent(s32, fork, 0);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((fock_kernel<double, W, C>), dim3(nblocks), dim3(64), lds, s, a);
  FockArgs b = a;
  b.slab = a.slab + (int64_t)nblocks * a.slab_stride;
  hipLaunchKernelGGL((fock_kernel<float, W, C>), dim3(nblocks), dim3(64), lds, s32, b);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipEventRecord(join, s32);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, join, 0);
  return e;
}

bool fock_launch_valid(FockLaunch c) {
  return ((c.waves == 3 || c.waves == 4) && c.cap == 1) ||
         (c.waves == 2 && (c.cap == 1 || c.cap == 2));
}

hipError_t launch_fock(const FockArgs &a, FockLaunch c, int nblocks, hipStream_t s,
                       hipStream_t s32, hipEvent_t fork, hipEvent_t join) {
  if (!fock_launch_valid(c)) return hipErrorInvalidValue;
  if (c.waves == 4) return launch_fock_w<4, 1>(a, nblocks, s, s32, fork, join);
  if (c.waves == 3) return launch_fock_w<3, 1>(a, nblocks, s, s32, fork, join);
  if (c.cap == 2) return launch_fock_w<2, 2>(a, nblocks, s, s32, fork, join);
  return launch_fock_w<2, 1>(a, nblocks, s, s32, fork, join);
}
hipError_t launch_diag(const DiagArgs &a, hipStream_t s) {
  if (a.nitems == 0) return hipSuccess;
  int bs = 128;
  int64_t nb = (a.nitems + bs - 1) / bs;
  hipLaunchKernelGGL(diag_kernel, dim3((unsigned)nb), dim3(bs), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_breit(const BreitArgs &a, int nblocks, hipStream_t s) {
  int nf = std::max(1, std::min(nblocks, a.nclasses));
  if (a.nclasses > 0)
    hipLaunchKernelGGL(breit_flags_kernel, dim3(nf), dim3(64), 0, s, a);
  if (a.nitems > 0) {
    const dim3 grid(std::max(1, std::min(nblocks, a.nitems)));
    if (a.xacc_n < 0 || a.xacc_n > kXAccMax) return hipErrorInvalidValue;
    const size_t lds = (size_t)a.xacc_n * sizeof(double);
    if (a.group == 1 && a.waves == 2)
      hipLaunchKernelGGL((breit_kernel<1, 2>), grid, dim3(64), lds, s, a);
    else if (a.group == 1 && a.waves == 3)
      hipLaunchKernelGGL((breit_kernel<1, 3>), grid, dim3(64), lds, s, a);
    else if (a.group == 1 && a.waves == 4)
      hipLaunchKernelGGL((breit_kernel<1, 4>), grid, dim3(64), lds, s, a);
    else if (a.group == 2 && a.waves == 2)
      hipLaunchKernelGGL((breit_kernel<2, 2>), grid, dim3(64), lds, s, a);
    else if (a.group == 4 && a.waves == 2)
      hipLaunchKernelGGL((breit_kernel<4, 2>), grid, dim3(64), lds, s, a);
    else
      return hipErrorInvalidValue;
  }
  if (!a.rr_out) {
    int64_t n = (int64_t)a.nskp * (a.ndim - a.nskp);
    int bs = 256;
    hipLaunchKernelGGL(breit_complete_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s,
                       a.bdir, a.bxch, a.ndim, a.nskp);
  }
  return hipGetLastError();
}
hipError_t launch_breit_tc(const BreitArgs &flags, const BreitTcArgs &a, int nblocks, size_t lds,
                           int nu, const BreitArgs &rest, int rest_blocks, hipStream_t s) {
  int nf = std::max(1, std::min(nblocks, flags.nclasses));
  if (flags.nclasses > 0)
    hipLaunchKernelGGL(breit_flags_kernel, dim3(nf), dim3(64), 0, s, flags);
  if (a.nitems > 0) {
    if (lds > (size_t)kTcLdsMax || nu > kTcMaxU) return hipErrorInvalidValue;
    const dim3 grid(std::max(1, std::min(nblocks, a.nitems)));
    if (nu <= 8)
      hipLaunchKernelGGL(breit_tc_kernel<8>, grid, dim3(kTcThreads), lds, s, a);
    else
      hipLaunchKernelGGL(breit_tc_kernel<kTcMaxU>, grid, dim3(kTcThreads), lds, s, a);
  }
  if (rest.nitems > 0) {
    if (rest.xacc_n < 0 || rest.xacc_n > kXAccMax) return hipErrorInvalidValue;
    const dim3 grid(std::max(1, std::min(rest_blocks, rest.nitems)));
    hipLaunchKernelGGL((breit_kernel<1, 4>), grid, dim3(64), (size_t)rest.xacc_n * sizeof(double),
                       s, rest);
  }
  int64_t n = (int64_t)flags.nskp * (flags.ndim - flags.nskp);
  int bs = 256;
  hipLaunchKernelGGL(breit_complete_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s,
                     flags.bdir, flags.bxch, flags.ndim, flags.nskp);
  return hipGetLastError();
}
hipError_t launch_oneel(const OneArgs &a, hipStream_t s) {
  if (a.nitems == 0) return hipSuccess;
  int bs = 64;
  hipLaunchKernelGGL(oneel_kernel, dim3((unsigned)((a.nitems + bs - 1) / bs)), dim3(bs), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_complete(const CompleteArgs &a, hipStream_t s) {
  int64_t n = (int64_t)a.nskp * a.nskp;
  int bs = 256;
  hipLaunchKernelGGL(complete_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, a);
  return hipGetLastError();
}

}  // namespace bdf
