// Learned correction of the Davidson diagonal preconditioner: a per-element MLP
// (8 inputs, three hidden layers of 128 with GELU, one output m) in FP32, trained on the
// device by the kernels below. The Davidson step t = -(r + alpha u) / (d - theta)
// (Davidson.cpp:322-350) becomes t = -(1 + m) (r + alpha u) / (d - theta); alpha keeps t
// orthogonal to u for any m. Activations are row-major (sample-major) matrices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "learned.hpp"

namespace chm {

namespace {

constexpr int GT = 64, GK = 16;  // output tile 64 x 64, K chunks of 16, 256 threads, 4 x 4 each

// C (M x N, ldc) = op(A) op(B) (+ C when accumulate), row-major:
// op(A)(i, k) = ta ? A[k lda + i] : A[i lda + k];  op(B)(k, j) = tb ? B[j ldb + k] : B[k ldb + j].
// The K range [k0, k1) of each workgroup is blockIdx.z's slice of K (split-K); with
// split > 1 the slices are written to C + z * M * ldc and summed by k_sum_slices.
__global__ __launch_bounds__(256) void k_sgemm(const float* __restrict__ A, int lda, int ta, const float* __restrict__ B,
                                               int ldb, int tb, float* __restrict__ C, int ldc, int M, int N, int K,
                                               int kslice, int accumulate) {
  __shared__ float As[GK][GT + 1];
  __shared__ float Bs[GK][GT + 1];
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  const int kb = blockIdx.z * kslice, ke = min(K, kb + kslice);
  float* Cz = C + (size_t)blockIdx.z * M * ldc;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += GK) {
    // 64 x 16 of op(A) and 16 x 64 of op(B), four elements per thread each
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      int r, kk;
      if (ta) { r = e & 63; kk = e >> 6; } else { kk = e & 15; r = e >> 4; }
      const int gi = m0 + r, gk = k0 + kk;
      float v = 0.f;
      if (gi < M && gk < ke) v = ta ? A[(size_t)gk * lda + gi] : A[(size_t)gi * lda + gk];
      As[kk][r] = v;
      int c, kk2;
      if (tb) { kk2 = e & 15; c = e >> 4; } else { c = e & 63; kk2 = e >> 6; }
      const int gj = n0 + c, gk2 = k0 + kk2;
      float w = 0.f;
      if (gj < N && gk2 < ke) w = tb ? B[(size_t)gj * ldb + gk2] : B[(size_t)gk2 * ldb + gj];
      Bs[kk2][c] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int gi = m0 + ty + 16 * i, gj = n0 + tx + 16 * j;
      if (gi < M && gj < N) {
        float* p = Cz + (size_t)gi * ldc + gj;
        *p = accumulate ? *p + acc[i][j] : acc[i][j];
      }
    }
}

// out[e] (+)= sum_z slices[z * n + e], in z order
__global__ __launch_bounds__(256) void k_sum_slices(const float* __restrict__ slices, int nz, int64_t n,
                                                    float* __restrict__ out, int accumulate) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    float s = 0.f;
    for (int z = 0; z < nz; z++) s += slices[(size_t)z * n + e];
    out[e] = accumulate ? out[e] + s : s;
  }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Z += b (per column); H = gelu(Z) (act) or Z
__global__ __launch_bounds__(256) void k_bias_act(float* __restrict__ Z, float* __restrict__ H, const float* __restrict__ b,
                                                  int64_t rows, int cols, int act) {
  const int64_t n = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const float z = Z[e] + b[e % cols];
    Z[e] = z;
    if (H) H[e] = act ? gelu(z) : z;
  }
}

// dZ = dH * gelu'(Z)
__global__ __launch_bounds__(256) void k_gelu_back(const float* __restrict__ dH, const float* __restrict__ Z,
                                                   float* __restrict__ dZ, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) dZ[e] = dH[e] * gelu_d(Z[e]);
}

// db[j] = sum_i dZ[i cols + j]: one workgroup per column, fixed summation order
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, int64_t rows, int cols, float* __restrict__ db) {
  __shared__ float s[256];
  const int j = blockIdx.x;
  float a = 0.f;
  for (int64_t i = threadIdx.x; i < rows; i += 256) a += dZ[i * cols + j];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = s[0];
}

// Loss sum_i w_i (p_i (1 + o_i) - e_i)^2 / n and dO_i; per-workgroup partial losses in
// double, summed by the host in order.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ o, const float* __restrict__ p,
                                              const float* __restrict__ e, const float* __restrict__ w, int64_t n,
                                              float* __restrict__ dO, double* __restrict__ part) {
  __shared__ double s[256];
  double a = 0.0;
  const double inv_n = 1.0 / (double)n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float d = p[i] * (1.f + o[i]) - e[i];
    a += (double)w[i] * (double)d * (double)d;
    if (dO) dO[i] = (float)(2.0 * inv_n) * w[i] * p[i] * d;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0] * inv_n;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ x, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    x[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// rows[q] = src[idx[q]] (MLP_IN floats per row) and the scalars of the sample
__global__ __launch_bounds__(256) void k_gather_rows(const float* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
                                                     float* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n * MLP_IN; e += (int64_t)gridDim.x * 256)
    dst[e] = src[idx[e / MLP_IN] * MLP_IN + e % MLP_IN];
}
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
                                                float* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) dst[e] = src[idx[e]];
}

// Features of element i: asinh(d - theta), r / rms(r), u sqrt(V), log(dL) / 4, log(dR) / 4.
__global__ __launch_bounds__(256) void k_features(const double* __restrict__ diag, const double* __restrict__ r,
                                                  const double* __restrict__ u, const float* __restrict__ eldims,
                                                  double theta, double inv_rms, double sqrt_n, int64_t n,
                                                  float* __restrict__ X) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    float* x = X + i * MLP_IN;
    x[0] = (float)asinh(diag[i] - theta);
    x[1] = (float)fmin(64.0, fmax(-64.0, r[i] * inv_rms));
    x[2] = (float)(u[i] * sqrt_n);
    x[3] = eldims[2 * i];
    x[4] = eldims[2 * i + 1];
    x[5] = x[6] = x[7] = 0.f;
  }
}

unsigned grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

}  // namespace

void sgemm(const float* A, int lda, bool ta, const float* B, int ldb, bool tb, float* C, int ldc, int M, int N, int K,
           bool accumulate, float* slices, int split, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  split = std::max(1, std::min(split, (K + GK - 1) / GK));
  const int kslice = ((K + split - 1) / split + GK - 1) / GK * GK;
  split = (K + kslice - 1) / kslice;
  const dim3 g((unsigned)((M + GT - 1) / GT), (unsigned)((N + GT - 1) / GT), (unsigned)split);
  if (split == 1) {
    hipLaunchKernelGGL(k_sgemm, g, dim3(256), 0, s, A, lda, (int)ta, B, ldb, (int)tb, C, ldc, M, N, K, kslice,
                       (int)accumulate);
  } else {
    // slices are M x ldc each (ldc == N for every caller of the split form)
    hipLaunchKernelGGL(k_sgemm, g, dim3(256), 0, s, A, lda, (int)ta, B, ldb, (int)tb, slices, ldc, M, N, K, kslice, 0);
    hipLaunchKernelGGL(k_sum_slices, dim3(grid((int64_t)M * ldc)), dim3(256), 0, s, slices, split, (int64_t)M * ldc, C,
                       (int)accumulate);
  }
}

int64_t Mlp::nparams() { return (int64_t)MLP_IN * MLP_H + MLP_H + 2 * ((int64_t)MLP_H * MLP_H + MLP_H) + MLP_H + 1; }

void Mlp::offsets(int64_t o[8]) {
  // W1 b1 W2 b2 W3 b3 W4 b4
  o[0] = 0;
  o[1] = o[0] + (int64_t)MLP_IN * MLP_H;
  o[2] = o[1] + MLP_H;
  o[3] = o[2] + (int64_t)MLP_H * MLP_H;
  o[4] = o[3] + MLP_H;
  o[5] = o[4] + (int64_t)MLP_H * MLP_H;
  o[6] = o[5] + MLP_H;
  o[7] = o[6] + MLP_H;
}

void Mlp::ensure(int64_t rows) {
  if (!params.p) {
    params.alloc(nparams());
    grad.alloc(nparams());
    am.alloc(nparams());
    av.alloc(nparams());
    slices.alloc((size_t)16 * MLP_H * MLP_H);
  }
  if (rows > cap) {
    cap = rows;
    for (auto* b : {&Z1, &Z2, &Z3, &H1, &H2, &H3, &dA, &dB}) b->alloc((size_t)rows * MLP_H);
    out.alloc(rows);
    dO.alloc(rows);
    lpart.alloc(8192);
  }
}

void Mlp::forward(const float* X, int64_t n, hipStream_t s) {
  ensure(n);
  int64_t o[8];
  offsets(o);
  const float* P = params.p;
  const int rows = (int)n;
  sgemm(X, MLP_IN, false, P + o[0], MLP_H, false, Z1.p, MLP_H, rows, MLP_H, MLP_IN, false, nullptr, 1, s);
  hipLaunchKernelGGL(k_bias_act, dim3(grid(n * MLP_H)), dim3(256), 0, s, Z1.p, H1.p, P + o[1], n, MLP_H, 1);
  sgemm(H1.p, MLP_H, false, P + o[2], MLP_H, false, Z2.p, MLP_H, rows, MLP_H, MLP_H, false, nullptr, 1, s);
  hipLaunchKernelGGL(k_bias_act, dim3(grid(n * MLP_H)), dim3(256), 0, s, Z2.p, H2.p, P + o[3], n, MLP_H, 1);
  sgemm(H2.p, MLP_H, false, P + o[4], MLP_H, false, Z3.p, MLP_H, rows, MLP_H, MLP_H, false, nullptr, 1, s);
  hipLaunchKernelGGL(k_bias_act, dim3(grid(n * MLP_H)), dim3(256), 0, s, Z3.p, H3.p, P + o[5], n, MLP_H, 1);
  sgemm(H3.p, MLP_H, false, P + o[6], 1, false, out.p, 1, rows, 1, MLP_H, false, nullptr, 1, s);
  hipLaunchKernelGGL(k_bias_act, dim3(grid(n)), dim3(256), 0, s, out.p, (float*)nullptr, P + o[7], n, 1, 0);
}

double Mlp::loss_grad(const float* X, const float* p, const float* e, const float* w, int64_t n, hipStream_t s) {
  forward(X, n, s);
  int64_t o[8];
  offsets(o);
  const float* P = params.p;
  float* G = grad.p;
  const int rows = (int)n;
  const int nb = (int)grid(n);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, out.p, p, e, w, n, dO.p, lpart.p);
  const int split = std::max(1, std::min(16, rows / 1024));
  // layer 4: dW4 = H3^T dO, db4 = sum dO, dH3 = dO W4^T
  sgemm(H3.p, MLP_H, true, dO.p, 1, false, G + o[6], 1, MLP_H, 1, rows, false, slices.p, split, s);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, s, dO.p, n, 1, G + o[7]);
  sgemm(dO.p, 1, false, P + o[6], 1, true, dA.p, MLP_H, rows, MLP_H, 1, false, nullptr, 1, s);
  // layer 3
  hipLaunchKernelGGL(k_gelu_back, dim3(grid(n * MLP_H)), dim3(256), 0, s, dA.p, Z3.p, dB.p, n * MLP_H);
  sgemm(H2.p, MLP_H, true, dB.p, MLP_H, false, G + o[4], MLP_H, MLP_H, MLP_H, rows, false, slices.p, split, s);
  hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, s, dB.p, n, MLP_H, G + o[5]);
  sgemm(dB.p, MLP_H, false, P + o[4], MLP_H, true, dA.p, MLP_H, rows, MLP_H, MLP_H, false, nullptr, 1, s);
  // layer 2
  hipLaunchKernelGGL(k_gelu_back, dim3(grid(n * MLP_H)), dim3(256), 0, s, dA.p, Z2.p, dB.p, n * MLP_H);
  sgemm(H1.p, MLP_H, true, dB.p, MLP_H, false, G + o[2], MLP_H, MLP_H, MLP_H, rows, false, slices.p, split, s);
  hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, s, dB.p, n, MLP_H, G + o[3]);
  sgemm(dB.p, MLP_H, false, P + o[2], MLP_H, true, dA.p, MLP_H, rows, MLP_H, MLP_H, false, nullptr, 1, s);
  // layer 1 (the input gradient is not needed)
  hipLaunchKernelGGL(k_gelu_back, dim3(grid(n * MLP_H)), dim3(256), 0, s, dA.p, Z1.p, dB.p, n * MLP_H);
  sgemm(X, MLP_IN, true, dB.p, MLP_H, false, G + o[0], MLP_H, MLP_IN, MLP_H, rows, false, slices.p, split, s);
  hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, s, dB.p, n, MLP_H, G + o[1]);
  std::vector<double> h(nb);
  hip_check(hipMemcpyAsync(h.data(), lpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, s), "loss");
  hip_check(hipStreamSynchronize(s), "loss");
  double L = 0.0;
  for (double v : h) L += v;
  return L;
}

void Mlp::adam(float lr, hipStream_t s) {
  step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid(nparams())), dim3(256), 0, s, params.p, grad.p, am.p, av.p, nparams(), lr, b1, b2,
                     1e-8f, c1, c2);
}

void Mlp::reset_optimizer(hipStream_t s) {
  step = 0;
  hip_check(hipMemsetAsync(am.p, 0, nparams() * sizeof(float), s), "adam state");
  hip_check(hipMemsetAsync(av.p, 0, nparams() * sizeof(float), s), "adam state");
}

void launch_gather_rows(const float* src, const int64_t* idx, int64_t n, float* dst, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_rows, dim3(grid(n * MLP_IN)), dim3(256), 0, s, src, idx, n, dst);
}
void launch_gather(const float* src, const int64_t* idx, int64_t n, float* dst, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather, dim3(grid(n)), dim3(256), 0, s, src, idx, n, dst);
}
void launch_features(const double* diag, const double* r, const double* u, const float* eldims, double theta,
                     double inv_rms, double sqrt_n, int64_t n, float* X, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_features, dim3(grid(n)), dim3(256), 0, s, diag, r, u, eldims, theta, inv_rms, sqrt_n, n, X);
}

}  // namespace chm
