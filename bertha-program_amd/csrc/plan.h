// plan.h -- host-side basis, E-coefficient term tables and Fock-build plan.
#pragma once
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/bertha_amd.h"
#include "engine.h"

namespace bdf {

struct Basis {
  int ncnt = 0, ndim = 0, nskp = 0;
  std::vector<double> xyz;
  std::vector<int> lmax;
  std::vector<int> nf;      // [c][kMaxL+1]
  std::vector<int> expoff;  // [c][kMaxL+1]
  std::vector<double> exps, normL, normS;
  std::vector<int> lrge;    // [c][16][16]: FOCKIND block offsets
  std::vector<int> labicn, labkqn, labmqn;
  int kappa_of(int c, int k) const;
  int nfn(int c, int l) const { return nf[c * (kMaxL + 1) + l]; }
  int expo(int c, int l) const { return expoff[c * (kMaxL + 1) + l]; }
  int lrge_at(int c, int k, int m) const { return lrge[(c * 16 + k) * 16 + m]; }
};

struct Etabs {
  std::vector<Mono> monos;
  std::vector<ETerm> terms;
  std::vector<EPair> pairs;
  std::vector<int64_t> flag_off;
  int64_t total = 0;  // double4 words
  int64_t nflag = 0;  // bytes
  std::map<std::tuple<int, int, int>, int> poly_index;
  std::map<std::tuple<int, int, int, int, int>, std::pair<int, int>> term_index;
  std::map<long, int> pair_index;
  // kind: kLL, kSS or kLSx/kLSy/kLSz (large A, small B, sigma_x/y/z)
  static long key(int a, int b, int ka, int kb, int ma, int mb, int kind) {
    return ((((((long)a * 64 + b) * 16 + ka) * 16 + kb) * 8 + ma) * 8 + mb) * 8 + kind;
  }
  int poly_id(int l, int m, int nr2);
  void spinor_terms(int ka, int mqa, int kb, int mqb, int kind, int &begin, int &count);
  void ls_terms(int ka, int mqa, int kb, int mqb, int iq);
};

std::string build_basis(const bdf_system &sys, Basis &B);
// ClassDesc::sym of a class from its centres' coordinates
int axis_symmetry(const ClassDesc &cl);
void build_etables(const Basis &B, Etabs &T, bool with_ls);
// zflags (exact-nonzero flags of the E tables, or null) let an axial plan
// take the one-component second contraction (Plan::one_comp)
std::string build_plan(const Basis &B, const Etabs &T, const std::vector<uint8_t> &eflags,
                       int isym, int intsym, int racah1, int ilev, Plan &P,
                       const std::vector<uint8_t> *zflags = nullptr);
// zflags (exact-nonzero flags of the E tables, or null) enable the tiles of
// the matrix-core kernel on the real path
std::string build_breit_plan(const Basis &B, const Etabs &T, const std::vector<uint8_t> &eflags,
                             int isym, int intsym, int racah1, int ilev, int gaunt,
                             BreitPlan &P, const std::vector<uint8_t> *zflags = nullptr);

// the one-centre Racah plan of every centre (racah.cpp) of the Coulomb
// (breit = 0, COULOMB1) or Breit (breit = 1, BREIT1) interaction; isym: 0
// general molecule (complex density), 1, 2 axial molecule or atom
std::string build_racah_plan(const Basis &B, int isym, int breit, RacahPlan &P);

}  // namespace bdf
