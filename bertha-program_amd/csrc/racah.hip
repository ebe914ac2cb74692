// racah.hip -- one-centre Racah-algebra Coulomb contributions (the
// reference's COULOMB1, bertha-scf.f:21540) on gfx950.
//
// One workgroup (one wave) per work item (KQN combination, primitive pair
// ij): the 64 lanes run over the kl pairs of the block.  Each lane evaluates
// the radial integrals RJ(kl, NU) of the four component types (RKCLM1 25664)
// from incomplete beta integrals (BTACLM1 25380), then for every (|m|, sign)
// combination of the KQN combination forms X = sum_NU ang(NU) RJ(NU) and
// applies the plan's digestion terms (CLMMT1Z/CLMMT1G).  Terms whose target
// element depends on i and j only are summed over the wave first; the rest
// add their lane's value.  All additions are hardware FP64 atomics, so the
// summation order varies from run to run at the last bit.
//
// The same kernel, instantiated for the Breit interaction, replaces BREIT1
// (26014): eight angular factors per tensor order weigh the upper and lower
// radial parts of the LSLS, SLSL, LSSL and SLLS integrals (RKBRT1 28525)
// into BXCH.
//
// The work is integer-light and transcendental-light: per (ij, kl) and
// tensor order 8-18 beta integrals (short series, one log when IB = 0) and
// per digestion term one density load and one atomic add.  It is bound by
// the atomics and the scattered density loads, not by FP64 arithmetic.
#include "kernels.h"

namespace bdf {

namespace {

__device__ __forceinline__ double rc_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// x^n for n >= 0 by repeated squaring
__device__ __forceinline__ double rc_pown(double x, int n) {
  double r = 1.0;
  while (n > 0) {
    if (n & 1) r *= x;
    x *= x;
    n >>= 1;
  }
  return r;
}

// incomplete beta integral of BTACLM1 for XJ = x, IA, IB (IB >= 0):
//   IB > 1: x^(IA+1/2) / (IA+1/2) * 2F1-type series of IB - 1 terms,
//   IB = 1: x^(IA+1/2) / (IA+1/2),
//   IB = 0: log((1+sqrt x)/(1-sqrt x)) - 2 sqrt x sum_{k<IA} x^k / (2k+1)
__device__ double rc_beta(double x, int ia, int ib) {
  const double xr = sqrt(x);
  if (ib >= 1) {
    const double a = (double)ia + 0.5;
    const double lead = rc_pown(xr, 2 * ia + 1) / a;
    if (ib == 1) return lead;
    // term_t = term_{t-1} * (a + t - 1)(t - ib) / ((a + t) t) * x
    double ra = a, rb = (double)(1 - ib), rc = a + 1.0, rd = 1.0;
    double trm = ra * rb / (rc * rd) * x, sum = 1.0 + trm;
    for (int t = 2; t <= ib - 1; ++t) {
      ra += 1.0; rb += 1.0; rc += 1.0; rd += 1.0;
      trm = ra * rb / (rc * rd) * trm * x;
      sum += trm;
    }
    return lead * sum;
  }
  const double lg = log((1.0 + xr) / (1.0 - xr));
  if (ia == 0) return lg;
  double sum = 1.0, pw = x;
  for (int k = 2; k <= ia; ++k) {
    sum += pw / (double)(2 * k - 1);
    pw *= x;
  }
  return lg - 2.0 * xr * sum;
}

// beta integral of block w at level (nx, ny) for the exponent sums tij, tkl
// of one of its (ij, kl) pairs: upper (BDU, XJ = tij / (tij + tkl)) or lower
// (BDL, XJ = tkl / (tij + tkl))
// (BTABRT1's arguments are one step higher than BTACLM1's)
__device__ __forceinline__ double rc_block_beta(const RcBlock &w, double tij, double tkl, int nx,
                                                int ny, bool upper) {
  const int lab = w.l[0] + w.l[1], lcd = w.l[2] + w.l[3], h = w.breit ? -1 : -2;
  if (upper)
    return rc_beta(tij / (tij + tkl), (lab + w.nui + 2 * nx + h) / 2, (lcd - w.nuf + 2 * ny + h) / 2);
  return rc_beta(tkl / (tij + tkl), (lcd + w.nui + 2 * nx + h) / 2, (lab - w.nuf + 2 * ny + h) / 2);
}

struct RcPair {
  double tij, tkl;
};

// the beta integral block b reads at (nx, ny) for flat pair index f = ij
// maxcd + kl: its own unless it writes none (then the writer's at f, as the
// reference's COMMON /BETA/ holds it)
__device__ double rc_read_beta(const RacahArgs &a, const RcBlock &b, int64_t f, RcPair own, int nx,
                               int ny, bool upper) {
  if (b.fresh || a.exact) return rc_block_beta(b, own.tij, own.tkl, nx, ny, upper);
  const int nv = max(nx, ny);
  for (int s = b.seg_begin; s < b.seg_end; ++s) {
    const RcSeg g = a.segs[s];
    if (g.nv != nv || f >= g.fend) continue;
    if (g.src < 0) return 0.0;
    const RcBlock &w = a.blocks[g.src];
    const int ij = (int)(f / w.maxcd), kl = (int)(f % w.maxcd);
    const double *e = a.exps;
    const double tij = e[w.exp[0] + ij / w.nb[1]] + e[w.exp[1] + ij % w.nb[1]];
    const double tkl = e[w.exp[2] + kl / w.nb[3]] + e[w.exp[3] + kl % w.nb[3]];
    return rc_block_beta(w, tij, tkl, nx, ny, upper);
  }
  return 0.0;
}

// (e_a + e_b)^(-(l_a + l_b + p)/2) for p = -nuf .. nuf + 5 (IJSET1 29301)
__device__ __forceinline__ void rc_powers(double t, int lsum, int nuf, double *ep) {
  const double r = sqrt(t), ir = 1.0 / r;
  const int ip = lsum - nuf;
  ep[0] = ip > 0 ? 1.0 / rc_pown(r, ip) : rc_pown(r, -ip);
  for (int p = 1; p <= 2 * nuf + 5; ++p) ep[p] = ep[p - 1] * ir;
}

}  // namespace

// radial integrals of the four component types per tensor order: COULOMB1
// (RKCLM1 25664) LLLL, LLSS, SSLL, SSSS in rj[q][t][0]; BREIT1 (RKBRT1 28525)
// LSLS, SLSL, LSSL, SLLS with their upper and lower parts in rj[q][t][0, 1]
template <bool kBreit>
__device__ __forceinline__ void rc_radial(const RacahArgs &a, const RcBlock &b, const RcCombo &cb,
                                          int64_t f, RcPair own, const double *eij,
                                          const double *ekl, double ei, double ej, double ek,
                                          double el, const double rnl[4], const double rns[4],
                                          double rj[4][kRcMaxNu][kBreit ? 2 : 1]) {
  const double ti = cb.tq[0], tj = cb.tq[1], tk = cb.tq[2], tl = cb.tq[3];
  const double *cc = b.cc;
  const int o = b.nuf;
#pragma unroll
  for (int t = 0; t < kRcMaxNu; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < (kBreit ? 2 : 1); ++u) rj[q][t][u] = 0.0;
    if (t >= b.nunum || !(cb.inu >> t & 1)) continue;
    const int lt = t + 1, nt = b.nunum - lt, nu = b.nus[t];
    auto BL = [&](int x, int y) { return rc_read_beta(a, b, f, own, x, y, false); };
    auto BU = [&](int x, int y) { return rc_read_beta(a, b, f, own, x, y, true); };
    if (!kBreit) {
      // E(ij)^(-nu+p) E(kl)^(nu+q+1) BDL + E(ij)^(nu+p+1) E(kl)^(-nu+q) BDU
      auto B = [&](int p, int q, int x1, int y1, int x2, int y2) {
        return eij[o - nu + p] * ekl[o + nu + q + 1] * BL(x1, y1) +
               eij[o + nu + p + 1] * ekl[o - nu + q] * BU(x2, y2);
      };
      const double b10 = B(0, 0, lt, nt + 1, lt, nt + 1);
      const double b12 = B(2, 0, lt, nt + 2, lt + 1, nt + 1);
      const double b14 = B(4, 0, lt, nt + 3, lt + 2, nt + 1);
      const double b30 = B(0, 2, lt + 1, nt + 1, lt, nt + 2);
      const double b32 = B(2, 2, lt + 1, nt + 2, lt + 1, nt + 2);
      const double b34 = B(4, 2, lt + 1, nt + 3, lt + 2, nt + 2);
      const double b50 = B(0, 4, lt + 2, nt + 1, lt, nt + 3);
      const double b52 = B(2, 4, lt + 2, nt + 2, lt + 1, nt + 3);
      const double b54 = B(4, 4, lt + 2, nt + 3, lt + 2, nt + 3);
      rj[0][t][0] = cc[2] * b32 * rnl[0] * rnl[1] * rnl[2] * rnl[3];
      rj[1][t][0] = (4.0 * ek * el * cc[3] * b52 - 2.0 * tl * ek * cc[2] * b32 -
                     2.0 * tk * el * cc[2] * b32 + tk * tl * cc[1] * b12) *
                    rnl[0] * rnl[1] * rns[2] * rns[3];
      rj[2][t][0] = (4.0 * ei * ej * cc[3] * b34 - 2.0 * tj * ei * cc[2] * b32 -
                     2.0 * ti * ej * cc[2] * b32 + ti * tj * cc[1] * b30) *
                    rns[0] * rns[1] * rnl[2] * rnl[3];
      rj[3][t][0] =
          (16.0 * ei * ej * ek * el * cc[4] * b54 - 8.0 * tl * ei * ej * ek * cc[3] * b34 -
           8.0 * tk * ei * ej * el * cc[3] * b34 - 8.0 * tj * ei * ek * el * cc[3] * b52 -
           8.0 * ti * ej * ek * el * cc[3] * b52 + 4.0 * ti * tj * ek * el * cc[2] * b50 +
           4.0 * tk * tl * ei * ej * cc[2] * b14 + 4.0 * ti * tl * ej * ek * cc[2] * b32 +
           4.0 * tj * tk * ei * el * cc[2] * b32 + 4.0 * tj * tl * ei * ek * cc[2] * b32 +
           4.0 * ti * tk * ej * el * cc[2] * b32 - 2.0 * ti * tj * tl * ek * cc[1] * b30 -
           2.0 * tj * tk * tl * ei * cc[1] * b12 - 2.0 * ti * tj * tk * el * cc[1] * b30 -
           2.0 * ti * tk * tl * ej * cc[1] * b12 + ti * tj * tk * tl * cc[0] * b10) *
          rns[0] * rns[1] * rns[2] * rns[3];
    } else {
      const double b21l = eij[o - nu + 1] * ekl[o + nu + 2] * BL(lt, nt + 1);
      const double b21u = eij[o + nu + 2] * ekl[o - nu + 1] * BU(lt, nt + 1);
      const double b23l = eij[o - nu + 3] * ekl[o + nu + 2] * BL(lt, nt + 2);
      const double b41u = eij[o + nu + 4] * ekl[o - nu + 1] * BU(lt + 1, nt + 1);
      const double b41l = eij[o - nu + 1] * ekl[o + nu + 4] * BL(lt + 1, nt + 1);
      const double b23u = eij[o + nu + 2] * ekl[o - nu + 3] * BU(lt, nt + 2);
      const double b43l = eij[o - nu + 3] * ekl[o + nu + 4] * BL(lt + 1, nt + 2);
      const double b43u = eij[o + nu + 4] * ekl[o - nu + 3] * BU(lt + 1, nt + 2);
      // (x, y) of LSLS (j, l), SLSL (i, k), LSSL (j, k), SLLS (i, l); the
      // normalisations follow the large (L) / small (S) label of A..D
      const double ex[4] = {ej, ei, ej, ei}, ey[4] = {el, ek, ek, el};
      const double tx[4] = {tj, ti, tj, ti}, ty[4] = {tl, tk, tk, tl};
      const double rn[4] = {rnl[0] * rns[1] * rnl[2] * rns[3], rns[0] * rnl[1] * rns[2] * rnl[3],
                            rnl[0] * rns[1] * rns[2] * rnl[3], rns[0] * rnl[1] * rnl[2] * rns[3]};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        rj[q][t][0] = (4.0 * ex[q] * ey[q] * cc[3] * b43u - 2.0 * ty[q] * ex[q] * cc[2] * b41u -
                       2.0 * tx[q] * ey[q] * cc[2] * b23u + tx[q] * ty[q] * cc[1] * b21u) * rn[q];
        rj[q][t][1] = (4.0 * ex[q] * ey[q] * cc[3] * b43l - 2.0 * ty[q] * ex[q] * cc[2] * b23l -
                       2.0 * tx[q] * ey[q] * cc[2] * b41l + tx[q] * ty[q] * cc[1] * b21l) * rn[q];
      }
    }
  }
}

template <bool kBreit>
__global__ void __launch_bounds__(64) racah_kernel(RacahArgs a) {
  const int lane = threadIdx.x;
  const RcItem it = a.items[blockIdx.x];
  const RcCombo cb = a.combos[it.combo];
  const RcBlock &b = a.blocks[cb.block];
  const int nB = b.nb[1], nD = b.nb[3];
  const int ib = it.ij / nB, jb = it.ij % nB;
  const int oa = b.exp[0] + ib, ob = b.exp[1] + jb;
  const double ei = a.exps[oa], ej = a.exps[ob];
  const double tij = ei + ej;
  double eij[2 * kRcMaxNuf + 6];
  rc_powers(tij, b.l[0] + b.l[1], b.nuf, eij);
  const double2 *D = a.dent;
  const int nd = a.ndim;
  const bool axial = a.isym != 0;

  for (int base = 0; base < b.maxcd; base += 64) {
    const int kl = base + lane;
    const bool on = kl < b.maxcd;
    const int kb = on ? kl / nD : 0, lb = on ? kl % nD : 0;
    const int oc = b.exp[2] + kb, od = b.exp[3] + lb;
    const double ek = a.exps[oc], el = a.exps[od];
    const double tkl = ek + el;
    const double rnl[4] = {a.normL[oa], a.normL[ob], a.normL[oc], a.normL[od]};
    const double rns[4] = {a.normS[oa], a.normS[ob], a.normS[oc], a.normS[od]};
    double ekl[2 * kRcMaxNuf + 6];
    rc_powers(tkl, b.l[2] + b.l[3], b.nuf, ekl);
    const int64_t f = (int64_t)it.ij * b.maxcd + (on ? kl : 0);
    double rj[4][kRcMaxNu][kBreit ? 2 : 1];
    rc_radial<kBreit>(a, b, cb, f, RcPair{tij, tkl}, eij, ekl, ei, ej, ek, el, rnl, rns, rj);
    const int idx[4] = {ib, jb, kb, lb};
    for (int e = cb.entry_begin; e < cb.entry_end; ++e) {
      const RcEntry en = a.entries[e];
      const double *ang = a.ang + en.ang;
      double x[4] = {0.0, 0.0, 0.0, 0.0};
      if (!kBreit) {
        for (int t = 0; t < b.nunum; ++t) {
          const double g = ang[t];
#pragma unroll
          for (int q = 0; q < 4; ++q) x[q] += g * rj[q][t][0];
        }
      } else {
        // ANGFAC 1-2 LSLS, 3-4 SLSL, 5-6 LSSL, 7-8 SLLS; the terms index
        // X as LSLS, LSSL, SLLS, SLSL
        for (int t = 0; t < b.nunum; ++t) {
          const double *g = ang + 8 * t;
          x[0] += g[0] * rj[0][t][0] + g[1] * rj[0][t][kBreit ? 1 : 0];
          x[3] += g[2] * rj[1][t][0] + g[3] * rj[1][t][kBreit ? 1 : 0];
          x[1] += g[4] * rj[2][t][0] + g[5] * rj[2][t][kBreit ? 1 : 0];
          x[2] += g[6] * rj[3][t][0] + g[7] * rj[3][t][kBreit ? 1 : 0];
        }
      }
      for (int p = en.op_begin; p < en.op_end; ++p) {
        const RcOp op = a.ops[p];
        const double v = on ? op.f * x[op.xt] : 0.0;
        const double2 d = D[(int64_t)(op.src_c + idx[op.ssc]) * nd + op.src_r + idx[op.ssr]];
        double re = v * d.x, im = axial ? 0.0 : v * d.y;
        double2 *M = op.mat ? a.gxch : a.gdir;
        const int64_t dst = (int64_t)(op.dst_c + idx[op.sdc]) * nd + op.dst_r + idx[op.sdr];
        if (op.uniform) {
          re = rc_wave_sum(re);
          if (!axial) im = rc_wave_sum(im);
          if (lane != 0) continue;
        } else if (!on) {
          continue;
        }
        if (re != 0.0) unsafeAtomicAdd(&M[dst].x, re);
        if (im != 0.0) unsafeAtomicAdd(&M[dst].y, im);
      }
    }
  }
}

// COULOMB1's closing step (22074-22081): every lower-triangle element
// becomes the conjugate of its upper-triangle partner, in GDIR and GXCH
__global__ void racah_hermitian_kernel(double2 *g, double2 *x, int n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int r = (int)(t % n), c = (int)(t / n);
  if (r <= c) return;
  if (g) {
    const double2 gu = g[(int64_t)r * n + c];
    g[t] = make_double2(gu.x, -gu.y);
  }
  const double2 xu = x[(int64_t)r * n + c];
  x[t] = make_double2(xu.x, -xu.y);
}

hipError_t launch_racah(const RacahArgs &a, hipStream_t s) {
  if (a.nitems > 0) {
    if (a.breit)
      hipLaunchKernelGGL(racah_kernel<true>, dim3(a.nitems), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL(racah_kernel<false>, dim3(a.nitems), dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // BREIT1 completes BXCH only (BDIR has no closed-shell terms)
  const int64_t n = (int64_t)a.ndim * a.ndim;
  hipLaunchKernelGGL(racah_hermitian_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     a.breit ? nullptr : a.gdir, a.gxch, a.ndim);
  return hipGetLastError();
}

}  // namespace bdf
