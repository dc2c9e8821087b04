// Driver for the reference CheMPS2 library built from its own sources by oracle/ref.mk
// (test infrastructure: it writes golden vectors and times the reference's CPU path; the
// product never runs it). It uses the reference's public API, and two private DMRG
// members in the steps mode.
//
// Problem files (written by oracle/refdrv.py), little endian:
//   int32 magic 0x50484d43 ("CMHP"), L, group, N, TwoS, Irrep, has_mx
//   int32 irreps[L]
//   float64 econst, T[L*L] (row major), V[L^4] (physicist <ij|kl>, row major i,j,k,l)
//   float64 mx[L^4] if has_mx: Problem::setMxElement(a, b, c, d) = mx[a + L(b + L(c + L d))]
//   after the DMRG object exists, then DMRG::PreSolve (as tests/test9.cpp.in:84-115 and
//   tests/test12.cpp.in:62-81 do)
//
// Modes
//   dims <problem> <D>
//       SyBookkeeper(Prob, D) (SyBookkeeper.cpp:29-269): one line per boundary b and
//       sector with a non-zero FCI dimension: "b N 2S I current fci".
//   steps <problem> <D> <first> <last> <rtol> <budget_s> [chkpt]
//       The first steps of the reference's first left sweep (DMRG::sweepleft with
//       change = false, DMRG.cpp:356-386): DMRG(Prob, scheme) (PreSolve), then for
//       index = first down to last, until the steps have taken budget_s seconds:
//       solve_site + updateMovingLeftSafe. Prints
//       "init seconds <t>" and per step "step <index> energy <E> seconds <t>"
//       (Davidson products follow in the reference's own "nIt(DAVIDSON)" lines). With
//       chkpt the MPS comes from CheMPS2_MPS0.h5 in the working directory.
//   dmrg <problem> <nstates> <eshift> <ninstr> {D econv max_sweeps noise rtol}*ninstr
//        [rdm <file>] [chkpt]
//       DMRG(Prob, scheme).Solve() (DMRG.cpp:268-417); for nstates > 1 the excited-state
//       driver of tests/test5.cpp.in (activateExcitations, newExcitation(eshift)).
//       Prints "state <k> energy <E> seconds <t>" per state. With rdm, also
//       calc2DMandCorrelations (DMRGtechnics.cpp:40) of the last state and writes the
//       2-RDM A and B arrays in DMRG orbital order (getTwoDMA_DMRG / getTwoDMB_DMRG,
//       index i + L(j + L(k + L l))) to <file>, and prints the 2-RDM energy. chkpt makes
//       the DMRG load / store CheMPS2_MPS0.h5 in the working directory (DMRG.cpp:100-120).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/time.h>
#include <vector>

#include "ConvergenceScheme.h"
// The steps mode drives DMRG::solve_site and updateMovingLeftSafe (DMRG.cpp:419-446,
// DMRGoperators.cpp), which are private: the header is read with them public. The
// library itself is built from the unmodified sources; access specifiers do not change
// the class layout.
#define private public
#include "DMRG.h"
#undef private
#include "Hamiltonian.h"
#include "Initialize.h"
#include "Problem.h"
#include "SyBookkeeper.h"
#include "TwoDM.h"

namespace {

struct ProbFile {
  int L = 0, group = 0, N = 0, TwoS = 0, Irrep = 0, has_mx = 0;
  std::vector<int> irreps;
  double econst = 0.0;
  std::vector<double> T, V, mx;
};

void die(const char* msg) {
  std::fprintf(stderr, "ref_driver: %s\n", msg);
  std::exit(2);
}

void rd(FILE* f, void* p, size_t n) {
  if (std::fread(p, 1, n, f) != n) die("short problem file");
}

ProbFile read_problem(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) die("cannot open problem file");
  int32_t head[7];
  rd(f, head, sizeof(head));
  if (head[0] != 0x50484d43) die("bad magic");
  ProbFile p;
  p.L = head[1]; p.group = head[2]; p.N = head[3]; p.TwoS = head[4]; p.Irrep = head[5]; p.has_mx = head[6];
  const size_t L = p.L, L4 = L * L * L * L;
  p.irreps.resize(L);
  rd(f, p.irreps.data(), L * 4);
  rd(f, &p.econst, 8);
  p.T.resize(L * L);
  rd(f, p.T.data(), L * L * 8);
  p.V.resize(L4);
  rd(f, p.V.data(), L4 * 8);
  if (p.has_mx) {
    p.mx.resize(L4);
    rd(f, p.mx.data(), L4 * 8);
  }
  std::fclose(f);
  return p;
}

CheMPS2::Hamiltonian* make_ham(const ProbFile& p) {
  std::vector<int> irr(p.irreps);
  auto* H = new CheMPS2::Hamiltonian(p.L, p.group, irr.data());
  H->setEconst(p.econst);
  const int L = p.L;
  for (int i = 0; i < L; i++)
    for (int j = i; j < L; j++)
      if (p.T[i * L + j] != 0.0) H->setTmat(i, j, p.T[i * L + j]);
  for (int i = 0; i < L; i++)
    for (int j = 0; j < L; j++)
      for (int k = 0; k < L; k++)
        for (int l = 0; l < L; l++) {
          const double v = p.V[((size_t(i) * L + j) * L + k) * L + l];
          if (v != 0.0) H->setVmat(i, j, k, l, v);
        }
  return H;
}

double now() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

int mode_dims(const ProbFile& p, int D) {
  CheMPS2::Hamiltonian* H = make_ham(p);
  CheMPS2::Problem prob(H, p.TwoS, p.N, p.Irrep);
  if (!prob.checkConsistency()) die("inconsistent problem");
  CheMPS2::SyBookkeeper bk(&prob, D);
  for (int b = 0; b <= p.L; b++)
    for (int N = bk.gNmin(b); N <= bk.gNmax(b); N++)
      for (int S = bk.gTwoSmin(b, N); S <= bk.gTwoSmax(b, N); S += 2)
        for (int I = 0; I < bk.getNumberOfIrreps(); I++) {
          const int fci = bk.gFCIdim(b, N, S, I);
          if (fci > 0) std::printf("%d %d %d %d %d %d\n", b, N, S, I, bk.gCurrentDim(b, N, S, I), fci);
        }
  delete H;
  return 0;
}

int mode_dmrg(const ProbFile& p, int argc, char** argv) {
  if (argc < 3) die("dmrg: <nstates> <eshift> <ninstr> ...");
  const int nstates = std::atoi(argv[0]);
  const double eshift = std::atof(argv[1]);
  const int ninstr = std::atoi(argv[2]);
  if (argc < 3 + 5 * ninstr) die("dmrg: too few instruction fields");
  CheMPS2::ConvergenceScheme scheme(ninstr);
  for (int i = 0; i < ninstr; i++) {
    char** a = argv + 3 + 5 * i;
    scheme.set_instruction(i, std::atoi(a[0]), std::atof(a[1]), std::atoi(a[2]), std::atof(a[3]), std::atof(a[4]));
  }
  const char* rdm_file = nullptr;
  bool chkpt = false;
  for (int i = 3 + 5 * ninstr; i < argc; i++) {
    if (!std::strcmp(argv[i], "rdm") && i + 1 < argc) rdm_file = argv[++i];
    else if (!std::strcmp(argv[i], "chkpt")) chkpt = true;
    else die("dmrg: unknown option");
  }
  CheMPS2::Hamiltonian* H = make_ham(p);
  CheMPS2::Problem prob(H, p.TwoS, p.N, p.Irrep);
  if (!prob.checkConsistency()) die("inconsistent problem");
  double t0 = now();
  auto* dmrg = new CheMPS2::DMRG(&prob, &scheme, chkpt);
  const int L = p.L;
  if (p.has_mx) {
    for (int a = 0; a < L; a++)
      for (int b = 0; b < L; b++)
        for (int c = 0; c < L; c++)
          for (int d = 0; d < L; d++)
            prob.setMxElement(a, b, c, d, p.mx[a + size_t(L) * (b + size_t(L) * (c + size_t(L) * d))]);
    dmrg->PreSolve();
  }
  std::printf("init seconds %.6f\n", now() - t0);
  if (nstates > 1) dmrg->activateExcitations(nstates - 1);
  for (int s = 0; s < nstates; s++) {
    if (s > 0) dmrg->newExcitation(eshift);
    const double t1 = now();
    const double E = dmrg->Solve();
    std::printf("state %d energy %.15f seconds %.6f\n", s, E, now() - t1);
    std::fflush(stdout);
  }
  if (rdm_file) {
    dmrg->calc2DMandCorrelations();
    CheMPS2::TwoDM* tdm = dmrg->get2DM();
    std::printf("rdm energy %.15f trace %.15f\n", tdm->energy(), tdm->trace());
    FILE* f = std::fopen(rdm_file, "wb");
    if (!f) die("cannot write rdm file");
    std::vector<double> buf(size_t(L) * L * L * L);
    for (int pass = 0; pass < 2; pass++) {
      for (int l = 0; l < L; l++)
        for (int k = 0; k < L; k++)
          for (int j = 0; j < L; j++)
            for (int i = 0; i < L; i++)
              buf[i + size_t(L) * (j + size_t(L) * (k + size_t(L) * l))] =
                  pass == 0 ? tdm->getTwoDMA_DMRG(i, j, k, l) : tdm->getTwoDMB_DMRG(i, j, k, l);
      std::fwrite(buf.data(), 8, buf.size(), f);
    }
    std::fclose(f);
  }
  if (chkpt) {
    // keep the checkpoint for the caller; only the operator files go
    dmrg->deleteStoredOperators();
  } else {
    if (CheMPS2::DMRG_storeMpsOnDisk) dmrg->deleteStoredMPS();
    if (CheMPS2::DMRG_storeRenormOptrOnDisk) dmrg->deleteStoredOperators();
  }
  delete dmrg;
  delete H;
  return 0;
}

int mode_steps(const ProbFile& p, int argc, char** argv) {
  if (argc < 5) die("steps: <D> <first> <last> <rtol> <budget_s> [chkpt]");
  const int D = std::atoi(argv[0]), first = std::atoi(argv[1]), last = std::atoi(argv[2]);
  const double rtol = std::atof(argv[3]), budget = std::atof(argv[4]);
  const bool chkpt = argc > 5 && !std::strcmp(argv[5], "chkpt");
  CheMPS2::ConvergenceScheme scheme(1);
  scheme.set_instruction(0, D, 1e-8, 1, 0.0, rtol);
  CheMPS2::Hamiltonian* H = make_ham(p);
  CheMPS2::Problem prob(H, p.TwoS, p.N, p.Irrep);
  if (!prob.checkConsistency()) die("inconsistent problem");
  double t0 = now();
  auto* dmrg = new CheMPS2::DMRG(&prob, &scheme, chkpt);
  std::printf("init seconds %.6f\n", now() - t0);
  std::fflush(stdout);
  double spent = 0.0;
  for (int index = first; index >= last && index > 0 && spent < budget; index--) {
    const double t1 = now();
    const double E = dmrg->solve_site(index, rtol, 0.0, D, true, false, false);
    dmrg->updateMovingLeftSafe(index);
    const double dt = now() - t1;
    spent += dt;
    std::printf("step %d energy %.15f seconds %.6f\n", index, E, dt);
    std::fflush(stdout);
  }
  dmrg->deleteStoredOperators();
  delete dmrg;
  delete H;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) die("usage: ref_driver dims|dmrg <problem> ...");
  CheMPS2::Initialize::Init();
  const ProbFile p = read_problem(argv[2]);
  if (!std::strcmp(argv[1], "dims")) {
    if (argc < 4) die("dims: <D>");
    return mode_dims(p, std::atoi(argv[3]));
  }
  if (!std::strcmp(argv[1], "dmrg")) return mode_dmrg(p, argc - 3, argv + 3);
  if (!std::strcmp(argv[1], "steps")) return mode_steps(p, argc - 3, argv + 3);
  die("unknown mode");
  return 2;
}
