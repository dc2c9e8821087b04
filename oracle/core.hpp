// CPU FP64 restatement of CheMPS2's two-site DMRG data structures.
//
// TEST INFRASTRUCTURE ONLY.  This directory is the oracle: tests, the smoke
// check and the CPU-baseline leg of the benchmark may call it, the product
// library (chemps2_amd) never does.  Every routine cites the reference file
// and line whose behaviour it restates; the code itself is written from the
// mathematics, not copied.
#pragma once
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

// (-1)^(two_p/2) for an even two_p (CheMPS2/include/chemps2/Special.h:36).
inline int phase(int two_p) { return (((two_p / 2) % 2) != 0) ? -1 : 1; }
// Abelian point groups in the XOR labelling (CheMPS2/include/chemps2/Irreps.h:123).
inline int irr(int a, int b) { return a ^ b; }
int num_irreps(int group);  // Irreps.cpp:100-108

// Wigner 6j symbol {a b c; d e f}, arguments are twice the angular momenta
// (Racah formula, the quantity CheMPS2/Wigner.cpp:294-339 evaluates).
double wigner6j(int a, int b, int c, int d, int e, int f);
// Wigner 9j symbol {a b c; d e f; g h i} (doubled arguments) as the sum over x of
// (-1)^{2x} (2x+1) {a b c; f i x}{d e f; b x h}{g h i; x a d} (CheMPS2/Wigner.cpp:341-369).
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);

// Column-major dense GEMM: C = alpha*op(A)*op(B) + beta*C (Fortran dgemm semantics).
void dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
           const double* B, int ldb, double beta, double* C, int ldc);

// ---------------------------------------------------------------- problem ---
// Hamiltonian + Problem restated (Hamiltonian.cpp:40-140, Problem.cpp:351-384).
struct Problem {
  int L = 0, group = 0, nirrep = 1;
  int N = 0, TwoS = 0, Irrep = 0;
  std::vector<int> orb_irrep;   // DMRG ordering
  std::vector<double> T;        // L*L  (hopping, DMRG ordering)
  std::vector<double> V;        // L^4  physicist <ij|kl>, index i+L(j+L(k+Ll))
  std::vector<double> mx;       // L^4  h_{ab;cd} = V + (T.delta + delta.T)/(N-1)
  double econst = 0.0;
  void init(int L_, int group_, const int* irreps);
  void setT(int i, int j, double v) { T[i + L * j] = v; T[j + L * i] = v; }
  void setV(int i, int j, int k, int l, double v);  // all 8 permutations (FourIndex.cpp:182-220)
  void build_mx();  // Problem.cpp:363-384
  double g(int a, int b, int c, int d) const { return mx[a + L * (b + L * (c + (size_t)L * d))]; }
};

// ------------------------------------------------------------- bookkeeper ---
// Virtual dimensions per boundary and (N, 2S, I) sector (SyBookkeeper.cpp:29-325).
struct Bookkeeper {
  const Problem* P = nullptr;
  int L = 0, nirrep = 1;
  std::vector<int> Nmin, Nmax;
  std::vector<std::vector<int>> Smin, Smax;  // [b][N-Nmin]
  std::vector<std::vector<int>> fci, cur;    // [b][flat]
  std::vector<std::vector<int>> base;        // [b][N-Nmin] start of the (S,I) slab
  void init(const Problem* P, int D);
  int idx(int b, int N, int S, int I) const;  // -1 if outside
  int fci_dim(int b, int N, int S, int I) const { int i = idx(b, N, S, I); return i < 0 ? 0 : fci[b][i]; }
  int dim(int b, int N, int S, int I) const { int i = idx(b, N, S, I); return i < 0 ? 0 : cur[b][i]; }
  void set_dim(int b, int N, int S, int I, int v) { int i = idx(b, N, S, I); if (i >= 0 && fci[b][i] > 0) cur[b][i] = v; }
  void scale(int D, int start, int stop);      // ScaleCURdim, SyBookkeeper.cpp:250-269
  int max_dim(int b) const;
  int tot_dim(int b) const;
  int irrep(int site) const { return P->orb_irrep[site]; }
};

inline uint64_t key6(int a, int b, int c, int d, int e, int f) {
  return ((uint64_t)(a & 0x3ff)) | ((uint64_t)(b & 0x3ff) << 10) | ((uint64_t)(c & 0xf) << 20) |
         ((uint64_t)(d & 0x3ff) << 24) | ((uint64_t)(e & 0x3ff) << 34) | ((uint64_t)(f & 0xf) << 44);
}

// ------------------------------------------------------------ MPS tensor ---
// Site tensor T[i]: blocks (NL,2SL,IL) x (NR,2SR,IR), column-major dL x dR
// (TensorT.cpp:29-104).
struct TensorT {
  int index = 0;
  const Bookkeeper* bk = nullptr;
  std::vector<int> NL, SL, IL, NR, SR, IR, k2i;
  std::vector<double> data;
  std::unordered_map<uint64_t, int> lut;
  void init(int site, const Bookkeeper* bk);
  int kappa(int nl, int sl, int il, int nr, int sr, int ir) const {
    auto it = lut.find(key6(nl, sl, il, nr, sr, ir));
    return it == lut.end() ? -1 : it->second;
  }
  double* blk(int nl, int sl, int il, int nr, int sr, int ir) {
    int k = kappa(nl, sl, il, nr, sr, ir);
    return k < 0 ? nullptr : data.data() + k2i[k];
  }
  int nk() const { return (int)NL.size(); }
};

// ------------------------------------------------------ operator tensor ---
// Block-sparse renormalised operator (TensorOperator.cpp:29-102): blocks
// (n_up, 2s_up, i_up) x (n_up+n_elec, 2s_down, i_up^n_irrep), dim_up x dim_down.
struct TOp {
  int index = 0, two_j = 0, n_elec = 0, n_irrep = 0;
  bool moving_right = true, prime_last = true, jw_phase = false;
  const Bookkeeper* bk = nullptr;
  std::vector<int> Nup, Sup, Iup, Sdn, k2i;
  std::vector<double> data;
  std::unordered_map<uint64_t, int> lut;
  void init(int index, int two_j, int n_elec, int n_irrep, bool mr, bool pl, bool jw, const Bookkeeper* bk);
  int kappa(int n1, int s1, int i1, int n2, int s2, int i2) const;
  double* blk(int n1, int s1, int i1, int n2, int s2, int i2) {
    int k = kappa(n1, s1, i1, n2, s2, i2);
    return k < 0 ? nullptr : data.data() + k2i[k];
  }
  int nk() const { return (int)Nup.size(); }
  void clear() { std::fill(data.begin(), data.end(), 0.0); }
  // Generic renormalisation step, TensorOperator.cpp:163-405.
  void update(TOp* prev, TensorT* Tup, TensorT* Tdown);
  void axpy(double a, const TOp* o) { for (size_t i = 0; i < data.size(); i++) data[i] += a * o->data[i]; }
  void axpy_transpose_CD(double a, TOp* o);  // TensorOperator.cpp:416-455
};

// ------------------------------------------------------- two-site object ---
// (Sobject.cpp:36-147): blocks (NL,2SL,IL,N1,N2,2J,NR,2SR,IR), dL x dR.
struct Sobject {
  int index = 0;
  Bookkeeper* bk = nullptr;
  int I1 = 0, I2 = 0;
  std::vector<int> NL, SL, IL, N1, N2, TJ, NR, SR, IR, k2i, reorder;
  std::vector<double> data;
  std::unordered_map<uint64_t, int> lut;
  void init(int index, Bookkeeper* bk);
  int kappa(int nl, int sl, int il, int n1, int n2, int tj, int nr, int sr, int ir) const;
  int nk() const { return (int)NL.size(); }
  size_t size() const { return k2i.empty() ? 0 : (size_t)k2i.back(); }
  void join(TensorT* Tl, TensorT* Tr);                                        // Sobject.cpp:212-258
  double split(TensorT* Tl, TensorT* Tr, int D, bool moving_right, bool change);  // Sobject.cpp:260-622
  void prog2symm();  // Sobject.cpp:624-636
  void symm2prog();  // Sobject.cpp:638-650
};

// ---------------------------------------------------- dense linear algebra ---
// Symmetric eigen-decomposition (cyclic Jacobi), ascending eigenvalues,
// eigenvectors in the columns of V (n x n, leading dimension ldv).
void sym_eig(int n, const double* A, int lda, double* w, double* V, int ldv);
// Thin SVD A (m x n) = U diag(s) VT, s descending; U m x r, VT r x n, r=min(m,n).
void svd_thin(int m, int n, const double* A, double* U, double* s, double* VT);
// Householder QR of A (m x n, column-major, in place): A <- Q (first min(m,n)
// columns, rest zero), R (n x n upper) returned; the LAPACK dgeqrf+dorgqr contract.
void qr_inplace(int m, int n, double* A, double* R);

// ---------------------------------------------------------------- RNG -------
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL) {}
  uint64_t next() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

}  // namespace orc
