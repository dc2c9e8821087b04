// Two-particle reduced density matrix restated on the CPU (oracle; test infrastructure
// only): TwoDM (TwoDM.cpp:40-630) and the MPS normalisations it needs.
#pragma once
#include "dmrg.hpp"

namespace orc {

// Spin-summed 2-RDMs A and B of TwoDM (TwoDM.cpp:40-56), L^4 each, element (i, j, k, l) at
// i + L (j + L (k + L l)), DMRG orbital order.
struct TwoDM {
  const Problem* P = nullptr;
  int L = 0;
  std::vector<double> A, B;
  void init(const Problem* p);
  void set(std::vector<double>& R, int i, int j, int k, int l, double v);
  void fill_site(TensorT& T, const OpSet* left, const OpSet* right);
  void correct_multiplicity();
  double trace() const;
  double energy() const;
};

// DMRG::left_normalize / right_normalize (DMRGtechnics.cpp:276-320).
void left_normalize(TensorT& T, TensorT* next);
void right_normalize(TensorT* prev, TensorT& T);

}  // namespace orc
