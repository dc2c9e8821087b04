// Davidson eigensolver and the two-site DMRG sweep of CheMPS2 restated on the CPU
// (oracle; test infrastructure only).
#pragma once
#include "heff.hpp"

namespace orc {

// Reverse-communication Davidson for the lowest eigenpair of a symmetric matrix
// (Davidson.cpp:31-536, eigenvalue mode 'E'). Instructions: 'A' fill guess
// (ptr[0]) and diagonal (ptr[1]); 'B' ptr[1] = H ptr[0]; 'C' converged, solution
// in ptr[0] and eigenvalue in ptr[1][0].
struct Davidson {
  int n = 0, max_vec = 0, keep = 0;
  double rtol = 0.0, cutoff = 0.0;
  char state = 'I';
  int num_vec = 0, n_mult = 0;
  std::vector<std::vector<double>> vecs, hvecs;
  std::vector<double> M, Meig, Mvec;  // max_vec x max_vec, leading dimension max_vec
  std::vector<double> diag, t, u, work;
  Rng rng{12345};
  Davidson(int n, int max_vec, int keep, double rtol, double cutoff);
  char fetch(double** ptr);
 private:
  void add_new_vec();                 // Davidson.cpp:214-240
  double diagonalize_and_residual();  // Davidson.cpp:242-320
  void new_vec();                     // Davidson.cpp:322-350
  void deflate();                     // Davidson.cpp:352-412
  void mxm_after_deflation();         // Davidson.cpp:510-536
};

// Heff::SolveDAVIDSON_main (Heff.cpp:330-386): Davidson(V, 32, 3, rtol, 1e-12) warm
// started from the joined two-site object in the symmetric convention.
double solve_davidson(Heff& H, Sobject& S, double rtol, int* n_mult = nullptr);

struct TwoDM;

struct Instruction { int D; double econv; int max_sweeps; double noise; double rtol; };

// DMRG driver (DMRG.cpp:44-446, DMRGoperators.cpp:233-907, DMRGtechnics.cpp:276-298).
struct DMRG {
  Problem P;
  Bookkeeper bk;
  Operators ops;
  std::vector<TensorT> mps;
  double total_min = 1e8, max_disc_last = 0.0, last_min = 1e8;
  double matvec_flops = 0.0;
  long n_matvec = 0;
  Rng rng{1};
  void init(int D0, uint64_t seed);  // bookkeeper at D0, random MPS in [-1,1), left-normalise, PreSolve
  void presolve();                   // DMRG.cpp:255-265
  double solve(const std::vector<Instruction>& scheme);  // DMRG.cpp:268-354
  double sweep_left(bool change, const Instruction& ins);   // DMRG.cpp:356-386
  double sweep_right(bool change, const Instruction& ins);  // DMRG.cpp:387-417
  double solve_site(int index, double rtol, double noise, int D, bool moving_right, bool change);  // DMRG.cpp:419-446
  // the 2-RDM of the current state, after a right sweep (DMRGtechnics.cpp:40-110)
  void calc_2rdm(TwoDM& out);
};

// TensorT::QR restated (TensorT.cpp:188-280): per right sector stack the left blocks
// and replace them by the orthonormal Q factor; R is discarded by left_normalize.
void left_normalize(TensorT& T);

// Reorder orbitals for D2h so that irreps appear as Ag B1u B3u B2g B2u B3g B1g Au
// (Problem.cpp:57-90). Returns f2 (DMRG orbital -> Hamiltonian orbital); identity
// for other groups.
std::vector<int> reorder_d2h(int L, int group, const int* irreps);

}  // namespace orc
