// Learned preconditioner: the per-element MLP and its FP32 kernels (learned.hip).
#pragma once
#include <cstdint>

#include "dbuf.hpp"

namespace chm {

constexpr int MLP_IN = 8;   // features per element (5 used, zero padded)
constexpr int MLP_H = 128;  // width of the three hidden layers

// C (M x N, ldc) = op(A) op(B) (+ C), row-major FP32; split > 1 splits K over
// workgroups through `slices` (split M ldc floats, ldc == N).
void sgemm(const float* A, int lda, bool ta, const float* B, int ldb, bool tb, float* C, int ldc, int M, int N, int K,
           bool accumulate, float* slices, int split, hipStream_t s);

struct Mlp {
  DBuf<float> params, grad, am, av, slices;
  DBuf<float> Z1, Z2, Z3, H1, H2, H3, dA, dB, out, dO;
  DBuf<double> lpart;
  int64_t cap = 0;
  int step = 0;
  bool ready = false;  // parameters set
  static int64_t nparams();
  static void offsets(int64_t o[8]);  // W1 b1 W2 b2 W3 b3 W4 b4 (W row-major, inputs x outputs)
  void ensure(int64_t rows);
  void forward(const float* X, int64_t n, hipStream_t s);  // out[0..n)
  // mean over the batch of w (p (1 + out) - e)^2; gradient into grad
  double loss_grad(const float* X, const float* p, const float* e, const float* w, int64_t n, hipStream_t s);
  void adam(float lr, hipStream_t s);
  void reset_optimizer(hipStream_t s);
};

void launch_gather_rows(const float* src, const int64_t* idx, int64_t n, float* dst, hipStream_t s);
void launch_gather(const float* src, const int64_t* idx, int64_t n, float* dst, hipStream_t s);
// X[i] = (asinh(d_i - theta), r_i inv_rms, u_i sqrt_n, eldims[2i], eldims[2i + 1], 0, 0, 0)
void launch_features(const double* diag, const double* r, const double* u, const float* eldims, double theta,
                     double inv_rms, double sqrt_n, int64_t n, float* X, hipStream_t s);

}  // namespace chm
