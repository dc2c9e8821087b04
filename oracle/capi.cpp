// C entry points of the CPU oracle, for ctypes (test infrastructure only: tests, the
// smoke check and the CPU-baseline leg of the benchmark load this library; the
// product library never does).
#include <chrono>

#include <omp.h>

#include "twodm.hpp"

using namespace orc;

namespace {

struct Handle {
  DMRG dm;
  std::unique_ptr<Sobject> S;
  std::unique_ptr<Heff> H;
  int site = -1;
};

enum Kind { K_L = 0, K_F0, K_F1, K_S0, K_S1, K_A, K_B, K_C, K_D, K_Q, K_X };

TOp* find_op(OpSet& o, int kind, int k, int m) {
  auto pick2 = [&](std::vector<std::vector<std::unique_ptr<TOp>>>& v) -> TOp* {
    if (k < 0 || k >= (int)v.size() || m < 0 || m >= (int)v[k].size()) return nullptr;
    return v[k][m].get();
  };
  switch (kind) {
    case K_L: return (k >= 0 && k < (int)o.L.size()) ? o.L[k].get() : nullptr;
    case K_F0: return pick2(o.F0);
    case K_F1: return pick2(o.F1);
    case K_S0: return pick2(o.S0);
    case K_S1: return pick2(o.S1);
    case K_A: return pick2(o.A);
    case K_B: return pick2(o.B);
    case K_C: return pick2(o.C);
    case K_D: return pick2(o.D);
    case K_Q: return (k >= 0 && k < (int)o.Q.size()) ? o.Q[k].get() : nullptr;
    case K_X: return o.X.get();
  }
  return nullptr;
}

}  // namespace

extern "C" {

long orc_site(void* p, int index);

// T: L x L hopping, V: L^4 physicist <ij|kl> at V[i + L(j + L(k + L l))], both in the
// Hamiltonian's orbital order; reorder != 0 applies the D2h ordering (group 7 only).
void* orc_create(int L, int group, const int* irreps, int N, int TwoS, int Irrep, const double* T, const double* V,
                 double econst, int reorder) {
  auto* h = new Handle();
  std::vector<int> f2(L);
  for (int i = 0; i < L; i++) f2[i] = i;
  if (reorder) f2 = reorder_d2h(L, group, irreps);
  std::vector<int> dirr(L);
  for (int i = 0; i < L; i++) dirr[i] = irreps[f2[i]];
  Problem& P = h->dm.P;
  P.init(L, group, dirr.data());
  P.N = N; P.TwoS = TwoS; P.Irrep = Irrep; P.econst = econst;
  for (int a = 0; a < L; a++)
    for (int b = 0; b < L; b++) P.T[a + L * b] = T[f2[a] + L * f2[b]];
  for (int a = 0; a < L; a++)
    for (int b = 0; b < L; b++)
      for (int c = 0; c < L; c++)
        for (int d = 0; d < L; d++)
          P.V[a + L * (b + L * (c + (size_t)L * d))] = V[f2[a] + L * (f2[b] + L * (f2[c] + (size_t)L * f2[d]))];
  return h;
}

void orc_destroy(void* p) { delete (Handle*)p; }

void orc_init(void* p, int D0, unsigned long long seed) { ((Handle*)p)->dm.init(D0, seed); }

double orc_solve(void* p, int n, const int* D, const double* econv, const int* max_sweeps, const double* noise,
                 const double* rtol) {
  std::vector<Instruction> s;
  for (int i = 0; i < n; i++) s.push_back({D[i], econv[i], max_sweeps[i], noise[i], rtol[i]});
  return ((Handle*)p)->dm.solve(s);
}

long orc_matvec_count(void* p) { return ((Handle*)p)->dm.n_matvec; }
double orc_matvec_flops(void* p) { return ((Handle*)p)->dm.matvec_flops; }

// Left-moving sweep steps index = L-2 .. stop+1 (solve, split, update) so that the
// operators right of site `stop` are built moving left.
double orc_sweep_left_to(void* p, int stop, int D, double rtol, int change) {
  DMRG& dm = ((Handle*)p)->dm;
  double E = 0.0;
  for (int index = dm.P.L - 2; index > stop; index--) {
    E = dm.solve_site(index, rtol, 0.0, D, false, change != 0);
    dm.ops.update_moving_left(index, dm.mps);
  }
  return E;
}

// Sobject::Split of the given two-site data (program convention) into the MPS of
// index, index+1; returns the discarded weight.
double orc_split_test(void* p, int index, int D, int mr, int change, const double* S_prog) {
  DMRG& dm = ((Handle*)p)->dm;
  Sobject S;
  S.init(index, &dm.bk);
  std::copy(S_prog, S_prog + S.size(), S.data.begin());
  return S.split(&dm.mps[index], &dm.mps[index + 1], D, mr != 0, change != 0);
}

// One left-moving sweep step at `index` (solve, split, update moving left).
double orc_step_left(void* p, int index, int D, double rtol, int change) {
  DMRG& dm = ((Handle*)p)->dm;
  const double E = dm.solve_site(index, rtol, 0.0, D, false, change != 0);
  dm.ops.update_moving_left(index, dm.mps);
  return E;
}

// splitmix64-based fill in [-scale, scale), element i of a tensor filled with `seed`
// (the benchmark's synthetic operator values).
static double hash_value(uint64_t seed, uint64_t i, double scale) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + i;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return scale * (2.0 * ((z >> 11) * (1.0 / 9007199254740992.0)) - 1.0);
}

// Benchmark workload: bookkeeper at D, the operator slots around `site` allocated with
// the reference's shapes and filled, tensor by tensor in allocation order, with
// seeds seed_l*1000003+17+j (slot site-1) and seed_r*1000003+17+j (slot site+1).
long orc_synthetic(void* p, int D, int site, unsigned long long seed_l, unsigned long long seed_r) {
  Handle* h = (Handle*)p;
  DMRG& dm = h->dm;
  dm.P.build_mx();
  dm.bk.init(&dm.P, D);
  dm.ops.init(&dm.P, &dm.bk);
  auto fill_set = [&](int cnt, bool mr, unsigned long long seed) {
    dm.ops.allocate(cnt, mr);
    OpSet& o = dm.ops.ops[cnt];
    uint64_t s = seed * 1000003ULL + 17;
    auto fill = [&](TOp* t) {
      if (!t) return;
      const double sc = t->data.empty() ? 0.0 : 1.0 / std::sqrt((double)t->data.size());
      for (size_t i = 0; i < t->data.size(); i++) t->data[i] = hash_value(s, i, sc);
      s++;
    };
    for (auto& t : o.L) fill(t.get());
    for (size_t k = 0; k < o.F0.size(); k++)
      for (size_t m = 0; m < o.F0[k].size(); m++) {
        fill(o.F0[k][m].get()); fill(o.F1[k][m].get()); fill(o.S0[k][m].get()); fill(o.S1[k][m].get());
      }
    for (size_t k = 0; k < o.A.size(); k++)
      for (size_t m = 0; m < o.A[k].size(); m++) {
        fill(o.A[k][m].get()); fill(o.B[k][m].get()); fill(o.C[k][m].get()); fill(o.D[k][m].get());
      }
    for (auto& t : o.Q) fill(t.get());
    fill(o.X.get());
  };
  if (site > 0) fill_set(site - 1, true, seed_l);
  if (site < dm.P.L - 2) fill_set(site + 1, false, seed_r);
  h->S = std::make_unique<Sobject>();
  h->S->init(site, &dm.bk);
  h->H = std::make_unique<Heff>();
  h->H->init(&dm.P, &dm.bk, h->S.get(), &dm.ops);
  h->site = site;
  return (long)h->S->size();
}

// Threads the OpenMP parts of the oracle use (its Heff application).
int orc_threads() { return omp_get_max_threads(); }

// Only the bookkeeper at D (sector dimensions), for layout comparisons.
void orc_bookkeeper(void* p, int D) {
  DMRG& dm = ((Handle*)p)->dm;
  dm.P.build_mx();
  dm.bk.init(&dm.P, D);
}

// Heff x restricted to output blocks [k0, k1); returns the GEMM flops of the sample
// and writes the wall time to *seconds.
double orc_site_apply_range(void* p, const double* x, double* y, int k0, int k1, double* seconds) {
  Handle* h = (Handle*)p;
  h->H->flops = 0.0;
  auto t0 = std::chrono::steady_clock::now();
  h->H->apply_range(x, y, k0, k1);
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return h->H->flops;
}

// The MPS given site by site (TensorT storage for the bookkeeper at D, as
// chm_mps_download writes it): the bookkeeper at D, every site tensor copied, the
// operators right of `site` built moving left (slots L-2 .. site+1) and left of it moving
// right (slots 0 .. site-1) with the oracle's own updates, then the two-site object and
// Heff at `site`. Returns V, or -1 if a tensor's size does not match.
long orc_load_mps(void* p, int D, int site, const double* const* data, const long* sizes) {
  Handle* h = (Handle*)p;
  DMRG& dm = h->dm;
  dm.P.build_mx();
  dm.bk.init(&dm.P, D);
  dm.mps.assign(dm.P.L, TensorT());
  for (int s = 0; s < dm.P.L; s++) {
    dm.mps[s].init(s, &dm.bk);
    if ((long)dm.mps[s].data.size() != sizes[s]) return -1;
    std::copy(data[s], data[s] + sizes[s], dm.mps[s].data.begin());
  }
  dm.ops.init(&dm.P, &dm.bk);
  for (int c = 0; c < site; c++) dm.ops.update_moving_right(c, dm.mps);
  for (int c = dm.P.L - 2; c > site; c--) dm.ops.update_moving_left(c, dm.mps);
  return orc_site(p, site);
}

// Builds the two-site object at `index` (joined from the current MPS) and its Heff.
long orc_site(void* p, int index) {
  Handle* h = (Handle*)p;
  h->S = std::make_unique<Sobject>();
  h->S->init(index, &h->dm.bk);
  h->S->join(&h->dm.mps[index], &h->dm.mps[index + 1]);
  h->S->prog2symm();
  h->H = std::make_unique<Heff>();
  h->H->init(&h->dm.P, &h->dm.bk, h->S.get(), &h->dm.ops);
  h->site = index;
  return (long)h->S->size();
}

void orc_site_x(void* p, double* x) {
  Handle* h = (Handle*)p;
  std::copy(h->S->data.begin(), h->S->data.end(), x);
}

double orc_site_apply(void* p, const double* x, double* y) {
  Handle* h = (Handle*)p;
  h->H->flops = 0.0;
  h->H->apply(x, y);
  return h->H->flops;
}

// Wall seconds for `reps` applications (single thread), for the CPU baseline.
double orc_site_apply_timed(void* p, const double* x, double* y, int reps) {
  Handle* h = (Handle*)p;
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; r++) h->H->apply(x, y);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

void orc_site_diag(void* p, double* d) { ((Handle*)p)->H->diag(d); }

// Davidson from the site's joined guess (symmetric convention); x receives the
// eigenvector in the symmetric convention.
double orc_site_davidson(void* p, double rtol, double* x, int* n_mult) {
  Handle* h = (Handle*)p;
  Sobject S2 = *h->S;
  S2.symm2prog();
  Heff H;
  H.init(&h->dm.P, &h->dm.bk, &S2, &h->dm.ops);
  const double e = solve_davidson(H, S2, rtol, n_mult);
  S2.prog2symm();
  std::copy(S2.data.begin(), S2.data.end(), x);
  return e;
}

int orc_site_nkappa(void* p) { return ((Handle*)p)->S->nk(); }

// labels: 9 ints per block (NL, 2SL, IL, N1, N2, 2J, NR, 2SR, IR); k2i: nk+1 offsets.
void orc_site_kappa(void* p, int* labels, int* k2i) {
  Sobject& S = *((Handle*)p)->S;
  for (int k = 0; k < S.nk(); k++) {
    int* o = labels + 9 * k;
    o[0] = S.NL[k]; o[1] = S.SL[k]; o[2] = S.IL[k]; o[3] = S.N1[k]; o[4] = S.N2[k];
    o[5] = S.TJ[k]; o[6] = S.NR[k]; o[7] = S.SR[k]; o[8] = S.IR[k];
  }
  for (int k = 0; k <= S.nk(); k++) k2i[k] = S.k2i[k];
}

int orc_L(void* p) { return ((Handle*)p)->dm.P.L; }
int orc_nirrep(void* p) { return ((Handle*)p)->dm.P.nirrep; }
int orc_irrep(void* p, int site) { return ((Handle*)p)->dm.P.orb_irrep[site]; }

// The L^4 table h_{ab;cd} in DMRG order (Problem.cpp:363-384).
void orc_mx(void* p, double* out) {
  const auto& mx = ((Handle*)p)->dm.P.mx;
  std::copy(mx.begin(), mx.end(), out);
}

// Sectors with non-zero current dimension at boundary b: 4 ints each (N, 2S, I, dim).
int orc_sectors(void* p, int b, int* out) {
  const Bookkeeper& bk = ((Handle*)p)->dm.bk;
  int n = 0;
  for (int N = bk.Nmin[b]; N <= bk.Nmax[b]; N++)
    for (int S = bk.Smin[b][N - bk.Nmin[b]]; S <= bk.Smax[b][N - bk.Nmin[b]]; S += 2)
      for (int I = 0; I < bk.nirrep; I++) {
        const int d = bk.dim(b, N, S, I);
        if (d <= 0) continue;
        if (out) { out[4 * n] = N; out[4 * n + 1] = S; out[4 * n + 2] = I; out[4 * n + 3] = d; }
        n++;
      }
  return n;
}

// Operator tensor at storage slot cnt (boundary cnt+1). meta: index, 2j, n_elec,
// n_irrep, moving_right, prime_last, jw_phase, direction of the set (1 right, 2 left).
// Returns the number of doubles, -1 if absent.
long orc_op(void* p, int cnt, int kind, int k, int m, int* meta, const double** data) {
  Handle* h = (Handle*)p;
  if (cnt < 0 || cnt >= (int)h->dm.ops.ops.size()) return -1;
  OpSet& o = h->dm.ops.ops[cnt];
  if (o.dir == 0) return -1;
  TOp* t = find_op(o, kind, k, m);
  if (!t) return -1;
  if (meta) {
    meta[0] = t->index; meta[1] = t->two_j; meta[2] = t->n_elec; meta[3] = t->n_irrep;
    meta[4] = t->moving_right; meta[5] = t->prime_last; meta[6] = t->jw_phase; meta[7] = o.dir;
  }
  if (data) *data = t->data.data();
  return (long)t->data.size();
}

// MPS site tensor data (TensorT storage); returns the number of doubles.
long orc_mps(void* p, int site, const double** data) {
  TensorT& T = ((Handle*)p)->dm.mps[site];
  if (data) *data = T.data.data();
  return (long)T.data.size();
}

int orc_op_qsite(void* p, int cnt, int k) { return ((Handle*)p)->dm.ops.ops[cnt].Qsite[k]; }

double orc_wigner6j(int a, int b, int c, int d, int e, int f) { return wigner6j(a, b, c, d, e, f); }
double orc_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  return wigner9j(a, b, c, d, e, f, g, h, i);
}

// Dense lowest eigenpair helpers exposed for the Davidson unit test: runs the oracle
// Davidson on a dense symmetric matrix A (n x n, column-major).
double orc_davidson_dense(int n, const double* A, const double* x0, double rtol, double* x, int* n_mult) {
  Davidson dv(n, 32, 3, rtol, 1e-12);
  double* ptr[2];
  char ins = dv.fetch(ptr);
  std::copy(x0, x0 + n, ptr[0]);
  for (int i = 0; i < n; i++) ptr[1][i] = A[i + (size_t)n * i];
  ins = dv.fetch(ptr);
  while (ins == 'B') {
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int j = 0; j < n; j++) s += A[i + (size_t)n * j] * ptr[0][j];
      ptr[1][i] = s;
    }
    ins = dv.fetch(ptr);
  }
  std::copy(ptr[0], ptr[0] + n, x);
  if (n_mult) *n_mult = dv.n_mult;
  return ptr[1][0];
}

// The 2-RDM of the current state (DMRG::calc2DMandCorrelations, 2-RDM part), after a
// solve ending with a right sweep: A and B (L^4 each, DMRG orbital order); returns
// Econst + 0.5 trace(A h) (TwoDM::energy).
double orc_2rdm(void* p, double* A, double* B) {
  DMRG& dm = ((Handle*)p)->dm;
  TwoDM t;
  dm.calc_2rdm(t);
  std::copy(t.A.begin(), t.A.end(), A);
  std::copy(t.B.begin(), t.B.end(), B);
  return t.energy();
}

}  // extern "C"
