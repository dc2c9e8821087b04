// Renormalised-operator construction of CheMPS2 restated on the CPU (oracle; test infrastructure only).
#pragma once
#include "core.hpp"

namespace orc {

// All operators living on one boundary (storage index cnt = boundary - 1),
// laid out as in DMRG::allocateTensors (DMRGoperators.cpp:909-1145).
struct OpSet {
  int dir = 0;  // 0 none, 1 built moving right, 2 built moving left
  std::vector<std::unique_ptr<TOp>> L;                                  // [k]
  std::vector<std::vector<std::unique_ptr<TOp>>> F0, F1, S0, S1;       // [k][m]
  std::vector<std::vector<std::unique_ptr<TOp>>> A, B, C, D;           // [k][m]
  std::vector<std::unique_ptr<TOp>> Q;                                  // [k]
  std::vector<int> Qsite;                                               // orbital of Q[k]
  std::unique_ptr<TOp> X;
};

struct Operators {
  const Problem* P = nullptr;
  const Bookkeeper* bk = nullptr;
  int L = 0;
  std::vector<OpSet> ops;  // [cnt] for cnt = 0..L-2
  void init(const Problem* P, const Bookkeeper* bk);
  void allocate(int cnt, bool moving_right);                    // DMRGoperators.cpp:909-1145
  void update_moving_right(int index, std::vector<TensorT>& mps);  // DMRGoperators.cpp:243-574
  void update_moving_left(int index, std::vector<TensorT>& mps);   // DMRGoperators.cpp:576-907
};

// Individual constructors (exposed for unit tests).
void L_create(TOp& self, TensorT* T);                                  // TensorL.cpp:41-208
void F0_makenew(TOp& self, TensorT* T);                                // TensorF0.cpp:53-133
void F0_makenew(TOp& self, TOp* Lop, TensorT* T);                      // TensorF0.cpp:135-285
void F1_makenew(TOp& self, TensorT* T);                                // TensorF1.cpp:54-123
void F1_makenew(TOp& self, TOp* Lop, TensorT* T);                      // TensorF1.cpp:125-269
void S0_makenew(TOp& self, TensorT* T);                                // TensorS0.cpp:53-103
void S0_makenew(TOp& self, TOp* Lop, TensorT* T);                      // TensorS0.cpp:105-246
void S1_makenew(TOp& self, TOp* Lop, TensorT* T);                      // TensorS1.cpp:48-195
void Q_add_simple(TOp& self, const Problem* P, int site, TensorT* T);  // TensorQ.cpp:45-100
void Q_add_L(TOp& self, const Problem* P, int site, std::vector<std::unique_ptr<TOp>>& Lprev, TensorT* T);  // TensorQ.cpp:102-390
void Q_add_AB(TOp& self, TOp* A, TOp* B, TensorT* T);                  // TensorQ.cpp:392-575
void Q_add_CD(TOp& self, TOp* C, TOp* D, TensorT* T);                  // TensorQ.cpp:577-785
void X_update(TOp& self, const Problem* P, TensorT* T);                // TensorX.cpp:44-56
void X_update(TOp& self, const Problem* P, TensorT* T, std::vector<std::unique_ptr<TOp>>& Lprev, TOp* Xprev,
              TOp* Qprev, TOp* A, TOp* C, TOp* D);                     // TensorX.cpp:58-125

}  // namespace orc
