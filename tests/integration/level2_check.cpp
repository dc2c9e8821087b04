// The level-2 wiring of INTEGRATION.md compiled against the reference's own headers and
// library (oracle/_ref/libchemps2_ref.so, built from the reference's sources by
// oracle/ref.mk) and against libchemps2_amd.so: a CheMPS2::Problem and
// CheMPS2::ConvergenceScheme of the reference drive chm_dmrg_init / chm_dmrg_solve exactly
// as the replaced DMRG constructor and DMRG::Solve would (DMRG.cpp:44-116, 268-352), and
// the reference's own DMRG solves the same problem on the host for comparison.
// Test infrastructure: it shows the drop-in from C++, nothing ships it.
//
//   level2_check <FCIDUMP> <group> <N> <2S> <irrep> <reorder D2h 0|1> <D> <econv> <max sweeps>
//
// Prints both energies; exits non-zero when they differ by more than 1e-8.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "ConvergenceScheme.h"
#include "DMRG.h"
#include "Hamiltonian.h"
#include "Initialize.h"
#include "Problem.h"
#include "chemps2_amd.h"

namespace {

void chm_ok(chm_ctx* c, int rc) {  // the reference's abort-on-error semantics
  if (rc != CHM_OK) {
    std::cerr << "chemps2_amd: " << chm_last_error(c) << std::endl;
    std::exit(2);
  }
}

// DMRG::DMRG(Prob, OptScheme) + DMRG::Solve() of INTEGRATION.md, level 2
double solve_on_gpu(CheMPS2::Problem* Prob, CheMPS2::ConvergenceScheme* OptScheme) {
  Prob->construct_mxelem();
  const int L = Prob->gL();
  chm_ctx* gpu = nullptr;
  if (chm_ctx_create(0, &gpu) != CHM_OK) {
    std::cerr << "chemps2_amd: no GPU context" << std::endl;
    std::exit(2);
  }
  std::vector<int> irr(L);
  std::vector<double> mx((size_t)L * L * L * L);
  for (int a = 0; a < L; a++) {
    irr[a] = Prob->gIrrep(a);
    for (int b = 0; b < L; b++)
      for (int c = 0; c < L; c++)
        for (int d = 0; d < L; d++) mx[a + L * (b + L * (c + (size_t)L * d))] = Prob->gMxElement(a, b, c, d);
  }
  chm_ok(gpu, chm_set_problem(gpu, L, Prob->gSy(), irr.data(), Prob->gN(), Prob->gTwoS(), Prob->gIrrep(), mx.data(),
                              Prob->gEconst()));
  chm_ok(gpu, chm_dmrg_init(gpu, OptScheme->get_D(0), /*seed=*/1));
  const int n = OptScheme->get_number();
  std::vector<int> D(n), sweeps(n);
  std::vector<double> econv(n), noise(n), rtol(n);
  for (int i = 0; i < n; i++) {
    D[i] = OptScheme->get_D(i);
    econv[i] = OptScheme->get_energy_conv(i);
    sweeps[i] = OptScheme->get_max_sweeps(i);
    noise[i] = OptScheme->get_noise_prefactor(i);
    rtol[i] = OptScheme->get_dvdson_rtol(i);
  }
  double e = 0.0;
  chm_ok(gpu, chm_dmrg_solve(gpu, n, D.data(), econv.data(), sweeps.data(), noise.data(), rtol.data(), &e));
  chm_ctx_destroy(gpu);
  return e;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 10) {
    std::cerr << "usage: level2_check FCIDUMP group N 2S irrep reorder D econv max_sweeps" << std::endl;
    return 2;
  }
  CheMPS2::Initialize::Init();
  const int group = std::atoi(argv[2]), N = std::atoi(argv[3]), TwoS = std::atoi(argv[4]), Irrep = std::atoi(argv[5]);
  const bool reorder = std::atoi(argv[6]) != 0;
  const int D = std::atoi(argv[7]), sweeps = std::atoi(argv[9]);
  const double econv = std::atof(argv[8]);
  CheMPS2::Hamiltonian* Ham = new CheMPS2::Hamiltonian(std::string(argv[1]), group);
  CheMPS2::Problem* Prob = new CheMPS2::Problem(Ham, TwoS, N, Irrep);
  if (reorder) Prob->SetupReorderD2h();
  CheMPS2::ConvergenceScheme* Opt = new CheMPS2::ConvergenceScheme(1);
  Opt->set_instruction(0, D, econv, sweeps, 0.0, 1e-5);
  const double e_gpu = solve_on_gpu(Prob, Opt);
  CheMPS2::DMRG* dmrg = new CheMPS2::DMRG(Prob, Opt);
  const double e_ref = dmrg->Solve();
  delete dmrg;
  delete Opt;
  delete Prob;
  delete Ham;
  std::printf("level2_check: chemps2_amd %.12f, reference %.12f, difference %.2e\n", e_gpu, e_ref, std::fabs(e_gpu - e_ref));
  return std::fabs(e_gpu - e_ref) < 1e-8 ? 0 : 1;
}
