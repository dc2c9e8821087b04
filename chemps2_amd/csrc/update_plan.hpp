// Work-list generator for the renormalised-operator updates of one boundary
// (DMRGoperators.cpp:243-907 with TensorL/F0/F1/S0/S1/Q/X and TensorOperator::update).
#pragma once
#include <functional>

#include "chm_internal.hpp"

namespace chm {

// out (m x n) += alpha * op(A) (m x k1) * op(B) (k1 x n)                     kind 0
// out (m x n) += alpha * op(A) (m x k1) * op(B) (k1 x k2) * op(C) (k2 x n)  kind 1, as (A B) C
// out (m x n) += alpha * op(A) (m x n)                                       kind 2
struct UTerm {
  Ref out;
  int m, n, k1, k2;
  double alpha;
  Ref A, B, C;
  int ta, tb, tc, lda, ldb, ldc;
  int kind;
};

// Terms of one task; their combination Refs are offsets relative to combo_base.
struct USeg {
  std::vector<UTerm> terms;
  int64_t combo_base = 0;
};

struct UpdatePlanner {
  const Sectors* sec = nullptr;
  const int* irr = nullptr;
  const double* mx = nullptr;
  int L = 0;
  int nthreads = 0;               // 0: host_threads()
  int level = 1;                  // 1: L, F, S, Q, X; 2: A, B, C, D (read level-1 results)
  std::vector<UTerm> terms[3];    // by level, while one task runs (and after flatten())
  std::vector<Combo> combos;      // x-independent operator sums, evaluated first
  int64_t combo_size = 0;
  double flops = 0.0;
  // moving_right / moving_left queue one task per output tensor (or per complementary
  // operator with its sums) and run them on the host threads; the results are
  // segments in queue order.
  struct Task {
    int level;
    std::function<void(UpdatePlanner&)> fn;
  };
  std::vector<Task> tasks;
  std::vector<USeg> segs[3];
  int64_t nterms(int lev) const {
    int64_t n = 0;
    for (const USeg& s : segs[lev]) n += (int64_t)s.terms.size();
    return n;
  }
  void run_tasks();
  void flatten();  // segments -> terms[1], terms[2] with arena offsets

  int irrep(int site) const { return irr[site]; }
  double gmx(int a, int b, int c, int d) const { return mx[a + (size_t)L * (b + (size_t)L * (c + (size_t)L * d))]; }

  // emitters with the argument conventions of the reference's dgemm_ calls
  void dgemm(char ta, char tb, int m, int n, int k, double alpha, Ref A, int lda, Ref B, int ldb, Ref out, int ldo);
  void dgemm(char ta, char tb, int m, int n, int k, double alpha, Ref A, int lda, Ref B, int ldb, double beta, Ref out,
             int ldo) {
    (void)beta;  // every new tensor starts at zero: all contributions accumulate
    dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, out, ldo);
  }
  void sand_r(Ref out, int dRu, int dRd, int dLu, int dLd, double alpha, Ref Tup, Ref mid, bool midT, Ref Tdn);
  void sand_l(Ref out, int dLu, int dLd, int dRu, int dRd, double alpha, Ref Tup, Ref mid, bool midT, Ref Tdn);
  void axpy(OpT& self, double a, OpT* o);
  void axpy_transpose_CD(OpT& self, double a, OpT* o);
  void update(OpT& self, OpT* prev, const TLay* Tup, const TLay* Tdn);  // TensorOperator.cpp:163-405

  std::vector<std::pair<Ref, double>> cur_src;
  int64_t cur_size = 0;
  void combo_begin(int64_t size) { cur_src.clear(); cur_size = size; }
  void combo_add(Ref r, double c);
  Ref combo_end();

  void L_create(OpT& self, const TLay* T);
  void F0_makenew(OpT& self, const TLay* T);
  void F0_makenew(OpT& self, OpT* Lop, const TLay* T);
  void F1_makenew(OpT& self, const TLay* T);
  void F1_makenew(OpT& self, OpT* Lop, const TLay* T);
  void S0_makenew(OpT& self, const TLay* T);
  void S0_makenew(OpT& self, OpT* Lop, const TLay* T);
  void S1_makenew(OpT& self, OpT* Lop, const TLay* T);
  void Q_add_simple(OpT& self, int site, const TLay* T);
  void Q_add_L(OpT& self, int site, std::vector<OpT*>& Lprev, const TLay* T);
  void Q_add_AB(OpT& self, OpT* A, OpT* B, const TLay* T);
  void Q_add_CD(OpT& self, OpT* C, OpT* D, const TLay* T);
  void X_makenew(OpT& self, const TLay* T, int ik);
  void X_update(OpT& self, const TLay* T);
  void X_update(OpT& self, const TLay* T, std::vector<OpT*>& Lprev, OpT* Xprev, OpT* Qprev, OpT* A, OpT* C, OpT* D);

  // Terms of the whole slot (DMRGoperators.cpp:243-574 / 576-907): o is the freshly laid
  // out set of slot `index`, prev the neighbouring slot (null at the chain ends).
  void moving_right(int index, OpSet& o, OpSet* prev, const TLay* T, const std::vector<int>& Qsite);
  void moving_left(int index, OpSet& o, OpSet* prev, const TLay* T, const std::vector<int>& Qsite);
};

}  // namespace chm
