// Two-site effective Hamiltonian of CheMPS2 restated on the CPU (oracle; test infrastructure only).
#pragma once
#include "operators.hpp"

namespace orc {

// y = Heff x for the two-site object S at sites (i, i+1) (Heff.cpp:43-248), with
// x and y in the symmetric convention; the left operators live at ops[i-1]
// (built moving right), the right ones at ops[i+1] (built moving left).
struct Heff {
  const Problem* P = nullptr;
  const Bookkeeper* bk = nullptr;
  Sobject* S = nullptr;
  Operators* ops = nullptr;
  int index = 0, L = 0;
  OpSet* Lo = nullptr;
  OpSet* Ro = nullptr;
  bool atLeft = false, atRight = false;
  double flops = 0.0;  // 2mnk over every GEMM issued, in the reference's association order

  void init(const Problem* P, const Bookkeeper* bk, Sobject* S, Operators* ops);
  void apply(const double* x, double* y);  // makeHeff
  // Output blocks k0 <= kappa < k1 only (a bounded sample of one application).
  void apply_range(const double* x, double* y, int k0, int k1);
  void apply_block(int ik);
  void diag(double* d);                    // fillHeffDiag, Heff.cpp:250-315

  // --- internals -------------------------------------------------------
  const double* x = nullptr;
  double* y = nullptr;
  std::vector<double> w1, w2;
  int dl(int ik) const { return bk->dim(index, S->NL[ik], S->SL[ik], S->IL[ik]); }
  int dr(int ik) const { return bk->dim(index + 2, S->NR[ik], S->SR[ik], S->IR[ik]); }
  // y[ik] += alpha * op(Lb) * x[kin] * op(Rb); op 'I' identity, 'N' as stored, 'T' transposed.
  void term(int ik, int kin, double alpha, char tl, const double* Lb, char tr, const double* Rb, bool x_times_r_first = false);
  // Pair operators at the left/right boundary, indexed by orbital pairs (a <= b).
  TOp* pl(std::vector<std::vector<std::unique_ptr<TOp>>>& v, int a, int b) { return v[b - a][index - 1 - b].get(); }   // left block pair
  TOp* cl(std::vector<std::vector<std::unique_ptr<TOp>>>& v, int a, int b) { return v[b - a][a - index].get(); }       // left boundary, pair right of the sites
  TOp* pr(std::vector<std::vector<std::unique_ptr<TOp>>>& v, int a, int b) { return v[b - a][a - index - 2].get(); }   // right block pair
  TOp* cr(std::vector<std::vector<std::unique_ptr<TOp>>>& v, int a, int b) { return v[b - a][index + 1 - b].get(); }   // right boundary, pair left of the sites

  void d1(int ik);   // 1A 1B 1C 1D
  void d2a(int ik);  // 2a1/2a2 spin 0/1, 2a3 spin 0/1
  void d2loc(int ik);  // 2b 2c 2e 2f (A, C, D local) and 2d
  void d3(int ik);
  void d4(int ik);
  void d4BC(int ik);  // 4B 4C: A/B/C/D of a pair (i or i+1, l>i+1) at the left boundary
  void d4E(int ik);   // 4E (site i) and 4H (site i+1)
  void d4I(int ik);
  void d4F(int ik);
  void d4G(int ik);
  void d4J(int ik);
  void d4KL(int ik);
  void d5(int ik);    // 5A-5F: L_a (left) x sum_b V L_b (right), pair of local sites
};

}  // namespace orc
