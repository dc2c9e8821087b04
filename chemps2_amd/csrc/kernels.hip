// gfx950 kernels of the chemps2_amd library (FP64).
//
// k_tiles: output-stationary grouped GEMM. One workgroup owns one 64 x 64 tile of an
// output block and accumulates every product of its group into registers before a
// single store, so no atomics are needed and the summation order is fixed.
// Operands: op(A) is m x k, op(B) is k x n, both column-major with leading dimensions
// and optional transposition; "axpy" products add alpha * X[tile].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "chm_internal.hpp"

namespace chm {

namespace {

constexpr int TM = 64, TN = 64, TK = 16, NT = 256;

__device__ inline const double* rel(const double* p, int is_rel, const double* base) {
  return is_rel ? base + (intptr_t)p : p;
}

// word k of a record held one 32-bit word per lane (uniform result)
__device__ __forceinline__ int ts_rl(uint32_t v, int k) { return __builtin_amdgcn_readlane((int)v, k); }
__device__ __forceinline__ uint64_t ts_rl64(uint32_t v, int k) {
  return (uint64_t)(uint32_t)ts_rl(v, k) | ((uint64_t)(uint32_t)ts_rl(v, k + 1) << 32);
}

// the output of tile t: relative to y (flags bit 0), to the partial sums (TILE_C_PART) or
// absolute
__device__ inline double* tile_out(const DevTile& t, double* ybase, double* pbase) {
  return (t.flags & 1) ? ybase + (intptr_t)t.C : ((t.flags & TILE_C_PART) ? pbase + (intptr_t)t.C : t.C);
}

__global__ __launch_bounds__(NT) void k_tiles(const DevTile* __restrict__ tiles, const DevGemmOp* __restrict__ ops,
                                              const double* __restrict__ xbase, double* __restrict__ ybase,
                                              double* __restrict__ pbase) {
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  const DevTile t = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  const int mrem = min(TM, t.m - t.m0), nrem = min(TN, t.n - t.n0);

  for (int o = 0; o < t.nop; o++) {
    const DevGemmOp op = ops[t.op0 + o];
    const double* A = rel(op.A, op.flags & 1, xbase);
    if (op.kind == 1) {
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int r = tx + 16 * i, c = ty + 16 * j;
          if (r < mrem && c < nrem)
            acc[i][j] += op.alpha * (op.ta ? A[(size_t)(t.n0 + c) + (size_t)op.lda * (t.m0 + r)]
                                           : A[(size_t)(t.m0 + r) + (size_t)op.lda * (t.n0 + c)]);
        }
      continue;
    }
    const double* B = rel(op.B, (op.flags >> 1) & 1, xbase);
    for (int k0 = 0; k0 < op.k; k0 += TK) {
      const int krem = min(TK, op.k - k0);
      // stage alpha * op(A)[m0:m0+64, k0:k0+16] as As[kk][r]
      if (op.ta == 0) {
        const int r = tid & 63;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int kk = (tid >> 6) + 4 * q;
          double v = 0.0;
          if (r < mrem && kk < krem) v = A[(size_t)(t.m0 + r) + (size_t)op.lda * (k0 + kk)];
          As[kk][r] = op.alpha * v;
        }
      } else {
        const int kk = tid & 15;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int r = (tid >> 4) + 16 * q;
          double v = 0.0;
          if (r < mrem && kk < krem) v = A[(size_t)(k0 + kk) + (size_t)op.lda * (t.m0 + r)];
          As[kk][r] = op.alpha * v;
        }
      }
      // stage op(B)[k0:k0+16, n0:n0+64] as Bs[kk][c]
      if (op.tb == 0) {
        const int kk = tid & 15;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int c = (tid >> 4) + 16 * q;
          double v = 0.0;
          if (c < nrem && kk < krem) v = B[(size_t)(k0 + kk) + (size_t)op.ldb * (t.n0 + c)];
          Bs[kk][c] = v;
        }
      } else {
        const int c = tid & 63;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int kk = (tid >> 6) + 4 * q;
          double v = 0.0;
          if (c < nrem && kk < krem) v = B[(size_t)(t.n0 + c) + (size_t)op.ldb * (k0 + kk)];
          Bs[kk][c] = v;
        }
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < TK; kk++) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = As[kk][tx + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
      }
      __syncthreads();
    }
  }
  double* C = tile_out(t, ybase, pbase);
  const bool add = (t.flags & 2) != 0;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int r = tx + 16 * i, c = ty + 16 * j;
      if (r < mrem && c < nrem) {
        double* p = C + (size_t)(t.m0 + r) + (size_t)t.ldc * (t.n0 + c);
        *p = add ? *p + acc[i][j] : acc[i][j];
      }
    }
}

// The same tiles on the matrix cores: each of the four wavefronts owns a 32 x 32 quarter
// of the 64 x 64 tile as 2 x 2 v_mfma_f64_16x16x4f64 accumulators (lane l holds rows
// 4 r + l / 16 of column l % 16 of each 16 x 16 block, register r). The K chunks of all
// products of the tile form one sequence: the next chunk's operands are loaded into
// registers while the current one computes from LDS, and the two LDS buffers alternate,
// so one barrier per chunk suffices. axpy products (kind 1) are added first.
typedef double chm_t4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const double gdbl;  // global (not flat) loads
typedef __attribute__((address_space(1))) const unsigned char gbyte;

__global__ __launch_bounds__(NT) void k_tiles_mfma(const DevTile* __restrict__ tiles, const DevGemmOp* __restrict__ ops,
                                                   const double* __restrict__ xbase, double* __restrict__ ybase,
                                                   double* __restrict__ pbase) {
  __shared__ double As[2][TK][TM + 1];
  __shared__ double Bs[2][TK][TN + 1];
  const DevTile t = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
  const int mrem = min(TM, t.m - t.m0), nrem = min(TN, t.n - t.n0);
  chm_t4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = chm_t4{0.0, 0.0, 0.0, 0.0};
  // axpy products: C += alpha * op(A) on this lane's accumulator elements
  for (int o = 0; o < t.nop; o++) {
    const DevGemmOp op = ops[t.op0 + o];
    if (op.kind != 1) continue;
    const double* A = rel(op.A, op.flags & 1, xbase);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int rr = wm + 16 * i + 4 * r + lg, cc = wn + 16 * j + lr;
          if (rr < mrem && cc < nrem)
            acc[i][j][r] += op.alpha * (op.ta ? A[(size_t)(t.n0 + cc) + (size_t)op.lda * (t.m0 + rr)]
                                              : A[(size_t)(t.m0 + rr) + (size_t)op.lda * (t.n0 + cc)]);
        }
  }
  // the GEMM chunks: (product o, k0) in order
  int o = 0, k0 = 0;
  auto skip = [&]() {  // advance o to the next GEMM product with work left
    while (o < t.nop && (ops[t.op0 + o].kind == 1 || k0 >= ops[t.op0 + o].k)) { o++; k0 = 0; }
  };
  skip();
  double ra[4], rb[4];
  auto fetch = [&]() {  // this thread's share of chunk (o, k0), alpha folded into A
    const DevGemmOp op = ops[t.op0 + o];
    const double* A = rel(op.A, op.flags & 1, xbase);
    const double* B = rel(op.B, (op.flags >> 1) & 1, xbase);
    const int krem = min(TK, op.k - k0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk;
      if (op.ta == 0) { r = tid & 63; kk = (tid >> 6) + 4 * q; } else { kk = tid & 15; r = (tid >> 4) + 16 * q; }
      double v = 0.0;
      if (r < mrem && kk < krem)
        v = op.ta ? A[(size_t)(k0 + kk) + (size_t)op.lda * (t.m0 + r)] : A[(size_t)(t.m0 + r) + (size_t)op.lda * (k0 + kk)];
      ra[q] = op.alpha * v;
      int c, kb;
      if (op.tb == 0) { kb = tid & 15; c = (tid >> 4) + 16 * q; } else { c = tid & 63; kb = (tid >> 6) + 4 * q; }
      double w = 0.0;
      if (c < nrem && kb < krem)
        w = op.tb ? B[(size_t)(t.n0 + c) + (size_t)op.ldb * (k0 + kb)] : B[(size_t)(k0 + kb) + (size_t)op.ldb * (t.n0 + c)];
      rb[q] = w;
    }
  };
  auto put = [&](int buf, int ta, int tb) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (ta == 0) As[buf][(tid >> 6) + 4 * q][tid & 63] = ra[q];
      else As[buf][tid & 15][(tid >> 4) + 16 * q] = ra[q];
      if (tb == 0) Bs[buf][tid & 15][(tid >> 4) + 16 * q] = rb[q];
      else Bs[buf][(tid >> 6) + 4 * q][tid & 63] = rb[q];
    }
  };
  int buf = 0;
  if (o < t.nop) fetch();
  while (o < t.nop) {
    const DevGemmOp op = ops[t.op0 + o];
    put(buf, op.ta, op.tb);
    __syncthreads();
    k0 += TK;
    skip();
    if (o < t.nop) fetch();  // in flight while this chunk computes
#pragma unroll
    for (int ks = 0; ks < TK / 4; ks++) {
      const int kk = 4 * ks + lg;
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) a[i] = As[buf][kk][wm + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < 2; j++) b[j] = Bs[buf][kk][wn + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    buf ^= 1;
  }
  double* C = tile_out(t, ybase, pbase);
  const bool add = (t.flags & 2) != 0;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int rr = wm + 16 * i + 4 * r + lg, cc = wn + 16 * j + lr;
        if (rr < mrem && cc < nrem) {
          double* p = C + (size_t)(t.m0 + rr) + (size_t)t.ldc * (t.n0 + cc);
          *p = add ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
}

// The same tiles without the padding (k_tiles_fit): only the 16 x 16 sub-tiles that hold
// elements of the block are computed, ms x ns of them (ms, ns <= 4), dealt out to the
// four wavefronts in turn (wavefront w takes sub-tiles w, w + 4, ...; C = ceil(ms ns / 4)
// per wavefront, a template argument so that every accumulator has a fixed register). The
// staging is k_tiles_mfma's: two LDS buffers, the next chunk's operands loaded into
// registers while the current one computes, one barrier per chunk. Tiles whose flags
// carry TILE_NO_AXPY skip the search for axpy products; the others find them 64 records
// at a time with a ballot.
template <int C>
__device__ __forceinline__ void tiles_fit_body(const DevTile& t, const DevGemmOp* __restrict__ ops,
                                               const double* __restrict__ xbase, double* __restrict__ ybase,
                                               double* __restrict__ pbase,
                                               double (*As)[TK][TM + 1], double (*Bs)[TK][TN + 1], int ms, int nsub) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int mrem = min(TM, t.m - t.m0), nrem = min(TN, t.n - t.n0);
  // this wavefront's sub-tiles (q < C). A slot past nsub (only the last can be) repeats
  // a sub-tile of the block and is not stored, so that the MFMA sequence has no branch
  // (a conditional MFMA makes the compiler move the accumulators around it).
  int si[C], sj[C];
  bool on[C];
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int sq = wv + 4 * q;
    on[q] = sq < nsub;
    const int se = on[q] ? sq : 0;
    si[q] = se % ms;
    sj[q] = se / ms;
  }
  int offa[C], offb[C];
#pragma unroll
  for (int q = 0; q < C; q++) {
    offa[q] = 16 * si[q] + lr;
    offb[q] = 16 * sj[q] + lr;
  }
  chm_t4 acc[C];
#pragma unroll
  for (int q = 0; q < C; q++) acc[q] = chm_t4{0.0, 0.0, 0.0, 0.0};
  // axpy products: C += alpha * op(A) on this lane's accumulator elements
  if (!(t.flags & TILE_NO_AXPY)) {
    for (int base = 0; base < t.nop; base += 64) {
      const int oi = base + lane;
      const int kind = oi < t.nop ? ops[t.op0 + oi].kind : 0;
      uint64_t msk = __ballot(kind == 1);
      while (msk) {
        const int qb = __ffsll((unsigned long long)msk) - 1;
        msk &= msk - 1;
        const DevGemmOp op = ops[t.op0 + base + qb];
        const double* A = rel(op.A, op.flags & 1, xbase);
#pragma unroll
        for (int q = 0; q < C; q++) {
          if (!on[q]) continue;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int rr = 16 * si[q] + 4 * r + lg, cc = 16 * sj[q] + lr;
            if (rr < mrem && cc < nrem)
              acc[q][r] += op.alpha * (op.ta ? A[(size_t)(t.n0 + cc) + (size_t)op.lda * (t.m0 + rr)]
                                             : A[(size_t)(t.m0 + rr) + (size_t)op.lda * (t.n0 + cc)]);
          }
        }
      }
    }
  }
  int o = 0, k0 = 0;
  auto skip = [&]() {  // advance o to the next GEMM product with work left
    while (o < t.nop && (ops[t.op0 + o].kind == 1 || k0 >= ops[t.op0 + o].k)) { o++; k0 = 0; }
  };
  skip();
  double ra[4], rb[4];
  // this thread's elements of chunk (op, k0): (r, kk) of op(A), (kb, c) of op(B)
  auto a_at = [&](const DevGemmOp& op, int q, int& r, int& kk) {
    if (op.ta == 0) { r = tid & 63; kk = (tid >> 6) + 4 * q; } else { kk = tid & 15; r = (tid >> 4) + 16 * q; }
  };
  auto b_at = [&](const DevGemmOp& op, int q, int& kb, int& c) {
    if (op.tb == 0) { kb = tid & 15; c = (tid >> 4) + 16 * q; } else { c = tid & 63; kb = (tid >> 6) + 4 * q; }
  };
  // raw global (not flat) loads only, a lane outside the block reading the block's first
  // element: put applies the masks and alpha, so nothing here waits for the loads and the
  // LDS waits of the compute do not wait for them either
  auto fetch = [&]() {
    const DevGemmOp op = ops[t.op0 + o];
    gdbl* A = (gdbl*)rel(op.A, op.flags & 1, xbase);
    gdbl* B = (gdbl*)rel(op.B, (op.flags >> 1) & 1, xbase);
    const int krem = min(TK, op.k - k0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk, kb, c;
      a_at(op, q, r, kk);
      b_at(op, q, kb, c);
      const size_t ia = op.ta ? (size_t)(k0 + kk) + (size_t)op.lda * (t.m0 + r) : (size_t)(t.m0 + r) + (size_t)op.lda * (k0 + kk);
      const size_t ib = op.tb ? (size_t)(t.n0 + c) + (size_t)op.ldb * (k0 + kb) : (size_t)(k0 + kb) + (size_t)op.ldb * (t.n0 + c);
      ra[q] = A[(r < mrem && kk < krem) ? ia : 0];
      rb[q] = B[(c < nrem && kb < krem) ? ib : 0];
    }
  };
  auto put = [&](int buf, const DevGemmOp& op, int kc) {  // the chunk (op, kc) fetched last
    const int krem = min(TK, op.k - kc);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk, kb, c;
      a_at(op, q, r, kk);
      b_at(op, q, kb, c);
      As[buf][kk][r] = (r < mrem && kk < krem) ? op.alpha * ra[q] : 0.0;
      Bs[buf][kb][c] = (c < nrem && kb < krem) ? rb[q] : 0.0;
    }
  };
  int buf = 0;
  if (o < t.nop) fetch();
  while (o < t.nop) {
    const DevGemmOp op = ops[t.op0 + o];
    put(buf, op, k0);
    __syncthreads();
    k0 += TK;
    skip();
    if (o < t.nop) fetch();  // in flight while this chunk computes
#pragma unroll
    for (int ks = 0; ks < TK / 4; ks++) {  // k past the product's end is zero in LDS
      const int kk = 4 * ks + lg;
#pragma unroll
      for (int q = 0; q < C; q++)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[buf][kk][offa[q]], Bs[buf][kk][offb[q]], acc[q], 0, 0, 0);
    }
    buf ^= 1;
  }
  double* Cp = tile_out(t, ybase, pbase);
  const bool add = (t.flags & 2) != 0;
#pragma unroll
  for (int q = 0; q < C; q++) {
    if (!on[q]) continue;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int rr = 16 * si[q] + 4 * r + lg, cc = 16 * sj[q] + lr;
      if (rr < mrem && cc < nrem) {
        double* p = Cp + (size_t)(t.m0 + rr) + (size_t)t.ldc * (t.n0 + cc);
        *p = add ? *p + acc[q][r] : acc[q][r];
      }
    }
  }
}

// tiles_fit_body for tiles without axpy products (TILE_NO_AXPY): the record of the next
// product comes in a product ahead as one vector load of a 32-bit word per lane (read with
// v_readlane when due), so a product switch does not wait for a scalar load before the
// next chunk's operands can be requested.
struct FitOp {
  const double* A;
  const double* B;
  double alpha;
  int lda, ldb, k, ta, tb;
};
template <int C>
__device__ __forceinline__ void tiles_fit_body_pf(const DevTile& t, const DevGemmOp* __restrict__ ops,
                                                  const double* __restrict__ xbase, double* __restrict__ ybase,
                                                  double* __restrict__ pbase, double (*As)[TK][TM + 1],
                                                  double (*Bs)[TK][TN + 1], int ms, int nsub) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int mrem = min(TM, t.m - t.m0), nrem = min(TN, t.n - t.n0);
  int si[C], sj[C], offa[C], offb[C];
  bool on[C];
#pragma unroll
  for (int q = 0; q < C; q++) {
    const int sq = wv + 4 * q;
    on[q] = sq < nsub;
    const int se = on[q] ? sq : 0;
    si[q] = se % ms;
    sj[q] = se / ms;
    offa[q] = 16 * si[q] + lr;
    offb[q] = 16 * sj[q] + lr;
  }
  chm_t4 acc[C];
#pragma unroll
  for (int q = 0; q < C; q++) acc[q] = chm_t4{0.0, 0.0, 0.0, 0.0};
  const uint32_t* __restrict__ ow = (const uint32_t*)(ops + t.op0);
  auto op_words = [&](int i) -> uint32_t {
    return (i < t.nop && lane < 14) ? ((__attribute__((address_space(1))) const uint32_t*)ow)[(size_t)i * 14 + lane] : 0u;
  };
  auto decode = [&](uint32_t v) {
    FitOp f;
    const int fl = ts_rl(v, 12);
    const double* A = (const double*)ts_rl64(v, 0);
    const double* B = (const double*)ts_rl64(v, 2);
    f.A = (fl & 1) ? xbase + (intptr_t)A : A;
    f.B = (fl & 2) ? xbase + (intptr_t)B : B;
    f.alpha = __longlong_as_double((long long)ts_rl64(v, 4));
    f.lda = ts_rl(v, 6); f.ldb = ts_rl(v, 7); f.k = ts_rl(v, 8); f.ta = ts_rl(v, 9); f.tb = ts_rl(v, 10);
    return f;
  };
  auto a_at = [&](const FitOp& op, int q, int& r, int& kk) {
    if (op.ta == 0) { r = tid & 63; kk = (tid >> 6) + 4 * q; } else { kk = tid & 15; r = (tid >> 4) + 16 * q; }
  };
  auto b_at = [&](const FitOp& op, int q, int& kb, int& c) {
    if (op.tb == 0) { kb = tid & 15; c = (tid >> 4) + 16 * q; } else { c = tid & 63; kb = (tid >> 6) + 4 * q; }
  };
  double ra[4], rb[4];
  auto fetch = [&](const FitOp& op, int kc) {  // raw global loads of chunk (op, kc)
    gdbl* A = (gdbl*)op.A;
    gdbl* B = (gdbl*)op.B;
    const int krem = min(TK, op.k - kc);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk, kb, c;
      a_at(op, q, r, kk);
      b_at(op, q, kb, c);
      const size_t ia = op.ta ? (size_t)(kc + kk) + (size_t)op.lda * (t.m0 + r) : (size_t)(t.m0 + r) + (size_t)op.lda * (kc + kk);
      const size_t ib = op.tb ? (size_t)(t.n0 + c) + (size_t)op.ldb * (kc + kb) : (size_t)(kc + kb) + (size_t)op.ldb * (t.n0 + c);
      ra[q] = A[(r < mrem && kk < krem) ? ia : 0];
      rb[q] = B[(c < nrem && kb < krem) ? ib : 0];
    }
  };
  auto put = [&](int buf, const FitOp& op, int kc) {
    const int krem = min(TK, op.k - kc);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk, kb, c;
      a_at(op, q, r, kk);
      b_at(op, q, kb, c);
      As[buf][kk][r] = (r < mrem && kk < krem) ? op.alpha * ra[q] : 0.0;
      Bs[buf][kb][c] = (c < nrem && kb < krem) ? rb[q] : 0.0;
    }
  };
  // the chunk sequence: product o (its fields in cur), chunk k0; vN: words of product o + 1
  int o = 0, k0 = 0;
  FitOp cur = decode(op_words(0));
  uint32_t vN = op_words(1);
  auto next_op = [&]() {  // to the next product with k > 0 (or o = nop)
    for (;;) {
      o++;
      if (o >= t.nop) return;
      cur = decode(vN);
      vN = op_words(o + 1);
      if (cur.k > 0) return;
    }
  };
  if (cur.k <= 0) next_op();
  int buf = 0;
  FitOp pc = cur;  // the product of the chunk in ra / rb
  int pk = 0;
  if (o < t.nop) fetch(cur, k0);
  while (o < t.nop) {
    put(buf, pc, pk);
    __syncthreads();
    k0 += TK;
    if (k0 >= cur.k) {
      k0 = 0;
      next_op();
    }
    if (o < t.nop) {
      fetch(cur, k0);  // in flight while this chunk computes
      pc = cur;
      pk = k0;
    }
#pragma unroll
    for (int ks = 0; ks < TK / 4; ks++) {  // k past the product's end is zero in LDS
      const int kk = 4 * ks + lg;
#pragma unroll
      for (int q = 0; q < C; q++)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[buf][kk][offa[q]], Bs[buf][kk][offb[q]], acc[q], 0, 0, 0);
    }
    buf ^= 1;
  }
  double* Cp = tile_out(t, ybase, pbase);
  const bool add = (t.flags & 2) != 0;
#pragma unroll
  for (int q = 0; q < C; q++) {
    if (!on[q]) continue;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int rr = 16 * si[q] + 4 * r + lg, cc = 16 * sj[q] + lr;
      if (rr < mrem && cc < nrem) {
        double* p = Cp + (size_t)(t.m0 + rr) + (size_t)t.ldc * (t.n0 + cc);
        *p = add ? *p + acc[q][r] : acc[q][r];
      }
    }
  }
}

__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_tiles_fit(const DevTile* __restrict__ tiles, const DevGemmOp* __restrict__ ops,
                                                  const double* __restrict__ xbase, double* __restrict__ ybase,
                                                  double* __restrict__ pbase) {
  __shared__ double As[2][TK][TM + 1];
  __shared__ double Bs[2][TK][TN + 1];
  const DevTile t = tiles[blockIdx.x];
  const int mrem = min(TM, t.m - t.m0), nrem = min(TN, t.n - t.n0);
  const int ms = (mrem + 15) >> 4, ns = (nrem + 15) >> 4, nsub = ms * ns;
  if (t.flags & TILE_NO_AXPY) {
    switch ((nsub + 3) >> 2) {
      case 1: tiles_fit_body_pf<1>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
      case 2: tiles_fit_body_pf<2>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
      case 3: tiles_fit_body_pf<3>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
      default: tiles_fit_body_pf<4>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
    }
    return;
  }
  switch ((nsub + 3) >> 2) {
    case 1: tiles_fit_body<1>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
    case 2: tiles_fit_body<2>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
    case 3: tiles_fit_body<3>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
    default: tiles_fit_body<4>(t, ops, xbase, ybase, pbase, As, Bs, ms, nsub); break;
  }
}

// Phase-1 tiles as a stream (k_tiles_stream1): tiles of one GEMM product each (the
// intermediates W = op(L) x or x op(R) of the chained terms), each record holding the tile
// and its product. A workgroup walks tiles blockIdx.x, blockIdx.x + gridDim.x, ... as one
// sequence of 16-deep k chunks: the next chunk (of this tile or the next) is loaded while
// the current one computes, so a tile's start no longer waits for its record and
// operands. The next tile's record comes in a tile ahead as one vector load of a 32-bit
// word per lane and is read with v_readlane when due. Operands are read through global
// (not flat) pointers, whose loads the LDS wait of the compute does not wait for. All
// tiles of a launch have C sub-tiles per wavefront (the host sorts them by C); sub-tiles
// as in k_tiles_fit.
static_assert(sizeof(DevTile1) == 80 && offsetof(DevTile1, ldc) == 32 && offsetof(DevTile1, flags) == 72,
              "DevTile1 words");
struct Ts1 {
  double* C;
  const double* A;
  const double* B;
  double alpha;
  int ldc, m0, n0, mrem, nrem, lda, ldb, k, ta, tb, flags;
};
__device__ __forceinline__ Ts1 ts_rec(uint32_t v) {
  Ts1 t;
  t.C = (double*)ts_rl64(v, 0);
  t.A = (const double*)ts_rl64(v, 2);
  t.B = (const double*)ts_rl64(v, 4);
  t.alpha = __longlong_as_double((long long)ts_rl64(v, 6));
  t.ldc = ts_rl(v, 8);
  t.m0 = ts_rl(v, 9);
  t.n0 = ts_rl(v, 10);
  t.mrem = ts_rl(v, 11);
  t.nrem = ts_rl(v, 12);
  t.lda = ts_rl(v, 13);
  t.ldb = ts_rl(v, 14);
  t.k = ts_rl(v, 15);
  t.ta = ts_rl(v, 16);
  t.tb = ts_rl(v, 17);
  t.flags = ts_rl(v, 18);
  return t;
}

template <int C>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_tiles_stream1(
    const DevTile1* __restrict__ tiles, int ntiles, const double* __restrict__ xbase) {
  __shared__ double As[2][TK][TM + 1];
  __shared__ double Bs[2][TK][TN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int G = gridDim.x;
  int ti = blockIdx.x;  // tile of the chunk being fetched
  if (ti >= ntiles) return;
  const uint32_t* __restrict__ tw = (const uint32_t*)tiles;
  auto rec_words = [&](int i) -> uint32_t {
    return (i < ntiles && lane < 20) ? ((__attribute__((address_space(1))) const uint32_t*)tw)[(size_t)i * 20 + lane] : 0u;
  };

  // the fetch side: tile fT, chunk k0, and the words of the next tile's record
  Ts1 fT = ts_rec(rec_words(ti));
  uint32_t vN = rec_words(ti + G);
  int k0 = 0;

  // the compute side: the tile of the chunk in LDS, its sub-tiles and accumulators
  Ts1 cT = fT;
  int offa[C], offb[C], si[C], sj[C];
  bool on[C];
  auto map_tile = [&](const Ts1& t) {
    const int ms = (t.mrem + 15) >> 4, ns = (t.nrem + 15) >> 4, nsub = ms * ns;
#pragma unroll
    for (int q = 0; q < C; q++) {
      const int sq = wv + 4 * q;
      on[q] = sq < nsub;
      const int se = on[q] ? sq : 0;
      si[q] = se % ms;
      sj[q] = se / ms;
      offa[q] = 16 * si[q] + lr;
      offb[q] = 16 * sj[q] + lr;
    }
  };
  map_tile(cT);
  chm_t4 acc[C];
#pragma unroll
  for (int q = 0; q < C; q++) acc[q] = chm_t4{0.0, 0.0, 0.0, 0.0};

  double ra[4], rb[4];
  // this thread's elements of chunk (fT, k0): (r, kk) of op(A), (kb, c) of op(B)
  auto a_at = [&](int q, int& r, int& kk) {
    if (fT.ta == 0) { r = tid & 63; kk = (tid >> 6) + 4 * q; } else { kk = tid & 15; r = (tid >> 4) + 16 * q; }
  };
  auto b_at = [&](int q, int& kb, int& c) {
    if (fT.tb == 0) { kb = tid & 15; c = (tid >> 4) + 16 * q; } else { c = tid & 63; kb = (tid >> 6) + 4 * q; }
  };
  // raw loads only (a lane outside the block reads the block's first element): the
  // masks and alpha are applied by put, so nothing here waits for the loads
  auto fetch = [&]() {
    gdbl* A = (gdbl*)((fT.flags & 1) ? xbase + (intptr_t)fT.A : fT.A);
    gdbl* B = (gdbl*)((fT.flags & 2) ? xbase + (intptr_t)fT.B : fT.B);
    const int krem = min(TK, fT.k - k0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk, kb, c;
      a_at(q, r, kk);
      b_at(q, kb, c);
      const size_t ia = fT.ta ? (size_t)(k0 + kk) + (size_t)fT.lda * (fT.m0 + r) : (size_t)(fT.m0 + r) + (size_t)fT.lda * (k0 + kk);
      const size_t ib = fT.tb ? (size_t)(fT.n0 + c) + (size_t)fT.ldb * (k0 + kb) : (size_t)(k0 + kb) + (size_t)fT.ldb * (fT.n0 + c);
      ra[q] = A[(r < fT.mrem && kk < krem) ? ia : 0];
      rb[q] = B[(c < fT.nrem && kb < krem) ? ib : 0];
    }
  };
  // the chunk fetched last into LDS buffer buf (fT and k0 still describe it)
  auto put = [&](int buf) {
    const int krem = min(TK, fT.k - k0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, kk, kb, c;
      a_at(q, r, kk);
      b_at(q, kb, c);
      As[buf][kk][r] = (r < fT.mrem && kk < krem) ? fT.alpha * ra[q] : 0.0;
      Bs[buf][kb][c] = (c < fT.nrem && kb < krem) ? rb[q] : 0.0;
    }
  };
  auto store = [&](const Ts1& t) {
#pragma unroll
    for (int q = 0; q < C; q++) {
      if (!on[q]) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int rr = 16 * si[q] + 4 * r + lg, cc = 16 * sj[q] + lr;
        if (rr < t.mrem && cc < t.nrem) t.C[(size_t)(t.m0 + rr) + (size_t)t.ldc * (t.n0 + cc)] = acc[q][r];
      }
    }
  };

  int buf = 0;
  fetch();
  for (;;) {
    put(buf);
    __syncthreads();
    // the next chunk: the rest of this product, else the first chunk of the next tile
    bool more = true, last = false;
    k0 += TK;
    if (k0 >= fT.k) {
      last = true;
      ti += G;
      if (ti >= ntiles) {
        more = false;
      } else {
        k0 = 0;
        fT = ts_rec(vN);
        vN = rec_words(ti + G);
      }
    }
    if (more) fetch();  // in flight while this chunk computes
#pragma unroll
    for (int ks = 0; ks < TK / 4; ks++) {  // k past the product's end is zero in LDS
      const int kk = 4 * ks + lg;
#pragma unroll
      for (int q = 0; q < C; q++)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[buf][kk][offa[q]], Bs[buf][kk][offb[q]], acc[q], 0, 0, 0);
    }
    if (last) {  // the chunk just computed ended tile cT
      store(cT);
#pragma unroll
      for (int q = 0; q < C; q++) acc[q] = chm_t4{0.0, 0.0, 0.0, 0.0};
      if (!more) break;
      cT = fT;
      map_tile(cT);
    }
    buf ^= 1;
  }
}

// Wavefront tiles of at most 32 x 32 (k_tiles_wave1, k_tiles_wave2): one wavefront per
// tile, no LDS and no barrier. Each lane loads its MFMA operands for a 16-deep k chunk
// straight from global memory (a[ks][i] = alpha op(A)[16 i + lr, k0 + 4 ks + lg],
// b[ks][j] = op(B)[k0 + 4 ks + lg, 16 j + lr]) as raw loads of clamped addresses; the next
// chunk is in flight while the current one computes, and the masks and alpha are applied
// when a chunk is taken for the MFMAs. Records come in ahead as vector loads of one word
// per lane and are read with v_readlane when due.
struct TwOp {  // one product on one tile, pointers resolved
  const double* A;
  const double* B;
  double alpha;
  int lda, ldb, k, ta, tb;
  int m0, n0, mrem, nrem;
};
// The helpers below work on an MS x NS grid of 16 x 16 sub-tiles per wavefront (2 x 2:
// tiles of at most 32 x 32; 4 x 1 and 1 x 4: the thin tiles of at most 64 x 16 and 16 x 64).
// raw loads of chunk (f, k0); 32-bit byte offsets (an operator block or x holds fewer
// than 2^28 elements), so that the loads take the uniform base from scalar registers
template <int MS, int NS>
__device__ __forceinline__ void tw_fetch(const TwOp& f, int k0, int lr, int lg, double (&pa)[4][MS], double (&pb)[4][NS]) {
  gbyte* A = (gbyte*)f.A;
  gbyte* B = (gbyte*)f.B;
  const int krem = f.k - k0;
  const uint32_t sa_k = f.ta ? 1u : (uint32_t)f.lda, sa_r = f.ta ? (uint32_t)f.lda : 1u;
  const uint32_t sb_k = f.tb ? (uint32_t)f.ldb : 1u, sb_c = f.tb ? 1u : (uint32_t)f.ldb;
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const int kk = 4 * ks + lg;
    const uint32_t kg = (uint32_t)(k0 + kk);
    const bool kok = kk < krem;
#pragma unroll
    for (int i = 0; i < MS; i++) {
      const int r = 16 * i + lr;
      const uint32_t ia = kg * sa_k + (uint32_t)(f.m0 + r) * sa_r;
      pa[ks][i] = *(gdbl*)(A + ((kok && r < f.mrem) ? ia * 8u : 0u));
    }
#pragma unroll
    for (int j = 0; j < NS; j++) {
      const int c = 16 * j + lr;
      const uint32_t ib = kg * sb_k + (uint32_t)(f.n0 + c) * sb_c;
      pb[ks][j] = *(gdbl*)(B + ((kok && c < f.nrem) ? ib * 8u : 0u));
    }
  }
}
template <int MS, int NS>
__device__ __forceinline__ void tw_take(const TwOp& f, int k0, int lr, int lg, double (&pa)[4][MS], double (&pb)[4][NS]) {
  const int krem = f.k - k0;
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const bool kok = 4 * ks + lg < krem;
#pragma unroll
    for (int i = 0; i < MS; i++) pa[ks][i] = (kok && 16 * i + lr < f.mrem) ? f.alpha * pa[ks][i] : 0.0;
#pragma unroll
    for (int j = 0; j < NS; j++) pb[ks][j] = (kok && 16 * j + lr < f.nrem) ? pb[ks][j] : 0.0;
  }
}
template <int MS, int NS>
__device__ __forceinline__ void tw_mma(chm_t4 (&acc)[MS][NS], const double (&pa)[4][MS], const double (&pb)[4][NS]) {
#pragma unroll
  for (int ks = 0; ks < 4; ks++)
#pragma unroll
    for (int i = 0; i < MS; i++)
#pragma unroll
      for (int j = 0; j < NS; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[ks][i], pb[ks][j], acc[i][j], 0, 0, 0);
}
template <int MS, int NS>
__device__ __forceinline__ void tw_zero(chm_t4 (&acc)[MS][NS]) {
#pragma unroll
  for (int i = 0; i < MS; i++)
#pragma unroll
    for (int j = 0; j < NS; j++) acc[i][j] = chm_t4{0.0, 0.0, 0.0, 0.0};
}
// the tile's elements of this lane: out[(m0 + row) + ldc (n0 + col)] (+)= acc
template <int MS, int NS>
__device__ __forceinline__ void tw_store(const chm_t4 (&acc)[MS][NS], double* out, int ldc, int m0, int n0, int mrem,
                                         int nrem, bool add, int lr, int lg) {
#pragma unroll
  for (int i = 0; i < MS; i++)
#pragma unroll
    for (int j = 0; j < NS; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int rr = 16 * i + 4 * r + lg, cc = 16 * j + lr;
        if (rr < mrem && cc < nrem) {
          double* p = out + (size_t)(m0 + rr) + (size_t)ldc * (n0 + cc);
          *p = add ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
}

// Phase-1 tiles (DevTile1) of MS x NS sub-tiles: wavefront w of the grid walks tiles w,
// w + W, ... (W wavefronts in the grid) as one chunk sequence; the next tile's record
// comes in a tile ahead.
template <int MS, int NS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_tiles_wave1(
    const DevTile1* __restrict__ tiles, int ntiles, const double* __restrict__ xbase) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int W = gridDim.x * (NT / 64);
  int ti = blockIdx.x * (NT / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ti >= ntiles) return;
  const uint32_t* __restrict__ tw = (const uint32_t*)tiles;
  auto rec_words = [&](int i) -> uint32_t {
    return (i < ntiles && lane < 20) ? ((__attribute__((address_space(1))) const uint32_t*)tw)[(size_t)i * 20 + lane] : 0u;
  };
  auto as_op = [&](const Ts1& t) {
    TwOp f;
    f.A = (t.flags & 1) ? xbase + (intptr_t)t.A : t.A;
    f.B = (t.flags & 2) ? xbase + (intptr_t)t.B : t.B;
    f.alpha = t.alpha;
    f.lda = t.lda; f.ldb = t.ldb; f.k = t.k; f.ta = t.ta; f.tb = t.tb;
    f.m0 = t.m0; f.n0 = t.n0; f.mrem = t.mrem; f.nrem = t.nrem;
    return f;
  };
  Ts1 fT = ts_rec(rec_words(ti));  // the tile of the chunk being fetched
  TwOp f = as_op(fT);
  uint32_t vN = rec_words(ti + W);
  int k0 = 0;
  Ts1 cT = fT;  // the tile of the chunk being computed
  chm_t4 acc[MS][NS];
  tw_zero(acc);
  // moves the fetch side to the next chunk; last: the chunk before ended its tile
  auto advance = [&](bool& last) -> bool {
    k0 += TK;
    last = false;
    if (k0 < fT.k) return true;
    last = true;
    ti += W;
    if (ti >= ntiles) return false;
    k0 = 0;
    fT = ts_rec(vN);
    f = as_op(fT);
    vN = rec_words(ti + W);
    return true;
  };
  auto finish = [&](bool last, bool more) -> bool {  // after the MFMAs of a chunk; false: done
    if (!last) return true;
    tw_store(acc, cT.C, cT.ldc, cT.m0, cT.n0, cT.mrem, cT.nrem, false, lr, lg);
    tw_zero(acc);
    if (!more) return false;
    cT = fT;
    return true;
  };
  double a0[4][MS], b0[4][NS], a1[4][MS], b1[4][NS];
  tw_fetch(f, k0, lr, lg, a0, b0);
  for (;;) {
    bool last;
    tw_take(f, k0, lr, lg, a0, b0);
    bool more = advance(last);
    if (more) tw_fetch(f, k0, lr, lg, a1, b1);  // in flight while chunk 0 computes
    tw_mma(acc, a0, b0);
    if (!finish(last, more)) break;
    tw_take(f, k0, lr, lg, a1, b1);
    more = advance(last);
    if (more) tw_fetch(f, k0, lr, lg, a0, b0);
    tw_mma(acc, a1, b1);
    if (!finish(last, more)) break;
  }
}

// Grouped tiles (DevTile + DevGemmOp lists, phase 2) of at most 32 x 32 whose products are
// all GEMMs with k > 0: one wavefront per tile walks the chunks of its products; the next
// product's record comes in a product ahead.
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_tiles_wave2(
    const DevTile* __restrict__ tiles, int ntiles, const DevGemmOp* __restrict__ ops, const double* __restrict__ xbase,
    double* __restrict__ ybase, double* __restrict__ pbase) {
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const int ti = blockIdx.x * (NT / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ti >= ntiles) return;
  const DevTile t = tiles[ti];
  if (t.nop <= 0) {
    chm_t4 z[2][2];
    tw_zero(z);
    tw_store(z, tile_out(t, ybase, pbase), t.ldc, t.m0, t.n0, min(32, t.m - t.m0), min(32, t.n - t.n0),
             (t.flags & 2) != 0, lr, lg);
    return;
  }
  const uint32_t* __restrict__ ow = (const uint32_t*)(ops + t.op0);
  auto op_words = [&](int i) -> uint32_t {
    return (i < t.nop && lane < 14) ? ((__attribute__((address_space(1))) const uint32_t*)ow)[(size_t)i * 14 + lane] : 0u;
  };
  const int mrem = min(32, t.m - t.m0), nrem = min(32, t.n - t.n0);  // the wavefront's 32 x 32
  auto as_op = [&](uint32_t v) {
    TwOp f;
    const int fl = ts_rl(v, 12);
    const double* A = (const double*)ts_rl64(v, 0);
    const double* B = (const double*)ts_rl64(v, 2);
    f.A = (fl & 1) ? xbase + (intptr_t)A : A;
    f.B = (fl & 2) ? xbase + (intptr_t)B : B;
    f.alpha = __longlong_as_double((long long)ts_rl64(v, 4));
    f.lda = ts_rl(v, 6); f.ldb = ts_rl(v, 7); f.k = ts_rl(v, 8); f.ta = ts_rl(v, 9); f.tb = ts_rl(v, 10);
    f.m0 = t.m0; f.n0 = t.n0; f.mrem = mrem; f.nrem = nrem;
    return f;
  };
  int o = 0, k0 = 0;
  TwOp f = as_op(op_words(0));
  uint32_t vN = op_words(1);
  chm_t4 acc[2][2];
  tw_zero(acc);
  auto advance = [&]() -> bool {
    k0 += TK;
    if (k0 < f.k) return true;
    o++;
    if (o >= t.nop) return false;
    k0 = 0;
    f = as_op(vN);
    vN = op_words(o + 1);
    return true;
  };
  double a0[4][2], b0[4][2], a1[4][2], b1[4][2];
  tw_fetch(f, k0, lr, lg, a0, b0);
  for (;;) {
    tw_take(f, k0, lr, lg, a0, b0);
    bool more = advance();
    if (more) tw_fetch(f, k0, lr, lg, a1, b1);  // in flight while chunk 0 computes
    tw_mma(acc, a0, b0);
    if (!more) break;
    tw_take(f, k0, lr, lg, a1, b1);
    more = advance();
    if (more) tw_fetch(f, k0, lr, lg, a0, b0);
    tw_mma(acc, a1, b1);
    if (!more) break;
  }
  tw_store(acc, tile_out(t, ybase, pbase), t.ldc, t.m0, t.n0, mrem, nrem, (t.flags & 2) != 0, lr, lg);
}

// Assembles the explicit sub-block of every (kappa, kappa') pair:
// M[(l + dL r), (l' + dLi r')] = sum_t alpha_t op(A_t)[l, l'] op(B_t)[r', r].
// SPLIT stores each value v as two floats, hi = fl32(v) in the first nnz floats of the
// buffer and lo = fl32(v - hi) in the next nnz: hi alone is v to 2^-24, hi + lo (exact
// in double) is v to 2^-48, in the same bytes as the double matrix.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_assemble(const DevPair* __restrict__ pairs, const DevKTerm* __restrict__ terms,
                                                  double* __restrict__ M, int64_t nnz) {
  const DevPair p = pairs[blockIdx.x];
  const int ncol = p.dLi * p.dRi;
  const int n = p.R * ncol;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int row = e % p.R, col = e / p.R;
    const int l = row % p.dL, r = row / p.dL;
    const int li = col % p.dLi, ri = col / p.dLi;
    double v = 0.0;
    for (int q = 0; q < p.nt; q++) {
      const DevKTerm t = terms[p.t0 + q];
      const double a = t.A ? (t.ta ? t.A[li + (size_t)t.lda * l] : t.A[l + (size_t)t.lda * li]) : (l == li ? 1.0 : 0.0);
      const double b = t.B ? (t.tb ? t.B[r + (size_t)t.ldb * ri] : t.B[ri + (size_t)t.ldb * r]) : (r == ri ? 1.0 : 0.0);
      v += t.alpha * a * b;
    }
    if constexpr (SPLIT) {
      float* hi = reinterpret_cast<float*>(M);
      const float h = (float)v;
      hi[p.moff + e] = h;
      hi[nnz + p.moff + e] = (float)(v - (double)h);
    } else {
      M[p.moff + e] = v;
    }
  }
}

// Explicit block product, pass 1: one workgroup per work item (at most BX_COLS
// columns). The item's x values are gathered into LDS first; then thread t takes row
// t % R and every ncl-th column (ncl = 256 / R column lanes) with eight matrix loads in
// flight, streamed non-temporally since every element is used once per product. The
// column lanes are summed through LDS in a fixed order. MODE 0 reads the double
// matrix; MODE 1 the hi floats of the split storage (half the bytes); MODE 2 hi + lo.
// Products and sums are in double in every mode.
constexpr int BX_COLS = 4096;
template <int MODE>
__device__ __forceinline__ double mload(const double* M, int64_t nnz, int64_t e) {
  if constexpr (MODE == 0) {
    return __builtin_nontemporal_load(M + e);
  } else if constexpr (MODE == 1) {
    return (double)__builtin_nontemporal_load(reinterpret_cast<const float*>(M) + e);
  } else {
    const float* hi = reinterpret_cast<const float*>(M);
    return (double)__builtin_nontemporal_load(hi + e) + (double)__builtin_nontemporal_load(hi + nnz + e);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_bspmv(const DevBItem* __restrict__ items, const double* __restrict__ M,
                                               int64_t nnz, const int* __restrict__ col, const double* __restrict__ x,
                                               double* __restrict__ y, double* __restrict__ P) {
  __shared__ double xs[BX_COLS];
  __shared__ double part[256];
  const DevBItem it = items[blockIdx.x];
  const int R = it.R, ncl = 256 / R;
  const int t = threadIdx.x;
  const int nc = it.c1 - it.c0;
  const int* ci = col + it.coff + it.c0;
  for (int c = t; c < nc; c += 256) xs[c] = x[ci[c]];
  __syncthreads();
  const int r = t % R, cl = t / R;
  double a[8];
#pragma unroll
  for (int u = 0; u < 8; u++) a[u] = 0.0;
  if (cl < ncl) {
    const int64_t m = it.moff + r + (int64_t)it.ld * it.c0;
    const int64_t ld = it.ld;
    int c = cl;
    for (; c + 7 * ncl < nc; c += 8 * ncl) {
      double mv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) mv[u] = mload<MODE>(M, nnz, m + ld * (c + u * ncl));
#pragma unroll
      for (int u = 0; u < 8; u++) a[u] = fma(mv[u], xs[c + u * ncl], a[u]);
    }
    for (; c < nc; c += ncl) a[0] = fma(mload<MODE>(M, nnz, m + ld * c), xs[c], a[0]);
  }
  part[t] = (cl < ncl) ? ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7])) : 0.0;
  __syncthreads();
  if (t < R) {
    double s = 0.0;
    for (int q = 0; q < ncl; q++) s += part[t + R * q];
    if (it.direct) y[it.out + t] = s;
    else P[it.out + t] = s;
  }
}

// Explicit block product, pass 2: sum the partial results of split blocks in order.
__global__ __launch_bounds__(256) void k_bspmv_reduce(const DevBRed* __restrict__ reds, const double* __restrict__ P,
                                                      double* __restrict__ y) {
  const DevBRed d = reds[blockIdx.x];
  for (int r = threadIdx.x; r < d.R; r += 256) {
    double s = 0.0;
    for (int q = 0; q < d.nparts; q++) s += P[d.p0 + (int64_t)q * d.R + r];
    y[d.yoff + r] = s;
  }
}

// out = sum_q c_q src_q for every combination; one 1024-element chunk per workgroup.
__global__ __launch_bounds__(256) void k_combos(const DevCombo* __restrict__ combos, const DevSrc* __restrict__ src,
                                                const int64_t* __restrict__ chunk_start, int ncombo) {
  const int64_t chunk = blockIdx.x;
  // binary search for the combination owning this chunk
  int lo = 0, hi = ncombo - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chunk_start[mid] <= chunk) lo = mid; else hi = mid - 1;
  }
  const DevCombo c = combos[lo];
  const int64_t base = (chunk - chunk_start[lo]) * 1024;
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int64_t idx = base + e;
    if (idx >= c.size) break;
    double v = 0.0;
    for (int q = 0; q < c.nsrc; q++) v += src[c.src0 + q].c * src[c.src0 + q].p[idx];
    c.out[idx] = v;
  }
}

// diag(Heff) block by block: one workgroup per block, threads over its elements.
__global__ __launch_bounds__(256) void k_diag(const DevDiagBlock* __restrict__ blocks, const DevDiag* __restrict__ terms) {
  const DevDiagBlock b = blocks[blockIdx.x];
  const int n = b.dL * b.dR;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int l = e % b.dL, r = e / b.dL;
    double v = 0.0;
    for (int q = 0; q < b.nt; q++) {
      const DevDiag t = terms[b.t0 + q];
      if (t.kind == 0) v += t.coef;
      else if (t.kind == 1) v += t.coef * t.L[(size_t)l * (b.dL + 1)];
      else if (t.kind == 2) v += t.coef * t.R[(size_t)r * (b.dR + 1)];
      else v += t.coef * t.L[(size_t)l * (b.dL + 1)] * t.R[(size_t)r * (b.dR + 1)];
    }
    b.out[e] = v;
  }
}

// Chained products, output-stationary (see DevCBlock). Every product is computed in
// 2 x 2 register tiles (4 LDS reads per 4 FMAs); tile u = ti + RT tj covers rows
// 2ti, 2ti+1 and columns 2tj, 2tj+1, clamped at odd edges (duplicates are computed and
// not stored). Each thread owns the output tiles u = tid + 256 q of its block. Per term
// the operator blocks are copied into LDS (contiguous, coalesced); two-operand terms
// form their intermediate in LDS in the cheaper association order chosen on the host.
// DB (double-buffered): the copies alternate between two buffers, so one barrier
// separates copy and use; without it a barrier also precedes every copy. Dynamic LDS:
// 6 (DB) or 4 arrays of P doubles (x, intermediate, A and B copies), P = the largest
// block product of the launch. A work item covers a range of input groups of one
// output block; it adds its sum to y (direct) or stores it as a partial for
// k_chain_reduce.
struct Tile22 {
  int i0, i1, j0, j1;
};

__device__ inline Tile22 tile22(int u, int rows, int cols) {
  const int rt = (rows + 1) >> 1;
  Tile22 t;
  t.i0 = 2 * (u % rt);
  t.j0 = 2 * (u / rt);
  t.i1 = min(t.i0 + 1, rows - 1);
  t.j1 = min(t.j0 + 1, cols - 1);
  return t;
}

// s[r + 2c] = sum_k P(i_r, k) Q(k, j_c), P(i, k) = P[i * pi + k * pk], Q(k, j) = Q[k * qk + j * qj]
__device__ inline void mm22(const double* __restrict__ P, int pi, int pk, const double* __restrict__ Q, int qk, int qj,
                            int K, const Tile22& t, double s[4]) {
  double s00 = 0.0, s10 = 0.0, s01 = 0.0, s11 = 0.0;
  const double* p0 = P + t.i0 * pi;
  const double* p1 = P + t.i1 * pi;
  const double* q0 = Q + t.j0 * qj;
  const double* q1 = Q + t.j1 * qj;
  for (int k = 0; k < K; k++) {
    const double a0 = p0[k * pk], a1 = p1[k * pk], b0 = q0[k * qk], b1 = q1[k * qk];
    s00 += a0 * b0;
    s10 += a1 * b0;
    s01 += a0 * b1;
    s11 += a1 * b1;
  }
  s[0] = s00; s[1] = s10; s[2] = s01; s[3] = s11;
}

__device__ inline void store22(double* __restrict__ D, int ld, int rows, int cols, const Tile22& t, const double s[4]) {
  D[t.i0 + ld * t.j0] = s[0];
  if (t.i0 + 1 < rows) D[t.i0 + 1 + ld * t.j0] = s[1];
  if (t.j0 + 1 < cols) {
    D[t.i0 + ld * (t.j0 + 1)] = s[2];
    if (t.i0 + 1 < rows) D[t.i0 + 1 + ld * (t.j0 + 1)] = s[3];
  }
}

template <int PERT, bool DB>
__global__ __launch_bounds__(256) void k_chain(const DevCBlock* __restrict__ blocks, const DevCGroup* __restrict__ groups,
                                               const DevCTerm* __restrict__ terms, const double* __restrict__ x,
                                               double* __restrict__ y, double* __restrict__ part, int P) {
  extern __shared__ double lds[];
  double* xs = lds;
  double* ts = lds + P;
  const DevCBlock b = blocks[blockIdx.x];
  const int tid = threadIdx.x;
  const int dL = b.dL, dR = b.dR;
  const int ntile = ((dL + 1) >> 1) * ((dR + 1) >> 1);
  double acc[PERT][4];
#pragma unroll
  for (int q = 0; q < PERT; q++) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.0;
  int buf = 0;
  for (int g = b.g0; g < b.g0 + b.ng; g++) {
    const DevCGroup G = groups[g];
    const int dLi = G.dLi, dRi = G.dRi;
    __syncthreads();  // every reader of the previous group's x is done
    for (int e = tid; e < dLi * dRi; e += 256) xs[e] = x[G.xoff + e];
    for (int t = G.t0; t < G.t0 + G.nt; t++) {
      const DevCTerm T = terms[t];
      double* As = lds + (2 + 2 * buf) * P;
      double* Bs = lds + (3 + 2 * buf) * P;
      if (DB) buf ^= 1;
      else __syncthreads();  // every reader of the previous term's copies is done
      if (T.A) for (int e = tid; e < dL * dLi; e += 256) As[e] = T.A[e];
      if (T.B) for (int e = tid; e < dRi * dR; e += 256) Bs[e] = T.B[e];
      __syncthreads();
      const double al = T.alpha;
      // op(A)(i, k) = As[i * ai + k * ak], op(B)(k, j) = Bs[k * bk + j * bj]
      const int ai = T.ta ? T.lda : 1, ak = T.ta ? 1 : T.lda;
      const int bk = T.tb ? T.ldb : 1, bj = T.tb ? 1 : T.ldb;
      double s[4];
      if (T.mode == 3) {  // ts = op(A) x  (dL x dRi), then y += al ts op(B)
        const int n1 = ((dL + 1) >> 1) * ((dRi + 1) >> 1);
        for (int u = tid; u < n1; u += 256) {
          const Tile22 tl = tile22(u, dL, dRi);
          mm22(As, ai, ak, xs, 1, dLi, dLi, tl, s);
          store22(ts, dL, dL, dRi, tl, s);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PERT; q++) {
          const int u = tid + 256 * q;
          if (u < ntile) {
            const Tile22 tl = tile22(u, dL, dR);
            mm22(ts, 1, dL, Bs, bk, bj, dRi, tl, s);
            for (int r = 0; r < 4; r++) acc[q][r] += al * s[r];
          }
        }
      } else if (T.mode == 4) {  // ts = x op(B)  (dLi x dR), then y += al op(A) ts
        const int n1 = ((dLi + 1) >> 1) * ((dR + 1) >> 1);
        for (int u = tid; u < n1; u += 256) {
          const Tile22 tl = tile22(u, dLi, dR);
          mm22(xs, 1, dLi, Bs, bk, bj, dRi, tl, s);
          store22(ts, dLi, dLi, dR, tl, s);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PERT; q++) {
          const int u = tid + 256 * q;
          if (u < ntile) {
            const Tile22 tl = tile22(u, dL, dR);
            mm22(As, ai, ak, ts, 1, dLi, dLi, tl, s);
            for (int r = 0; r < 4; r++) acc[q][r] += al * s[r];
          }
        }
      } else if (T.mode == 1) {  // y += al op(A) x
#pragma unroll
        for (int q = 0; q < PERT; q++) {
          const int u = tid + 256 * q;
          if (u < ntile) {
            const Tile22 tl = tile22(u, dL, dR);
            mm22(As, ai, ak, xs, 1, dLi, dLi, tl, s);
            for (int r = 0; r < 4; r++) acc[q][r] += al * s[r];
          }
        }
      } else if (T.mode == 2) {  // y += al x op(B)
#pragma unroll
        for (int q = 0; q < PERT; q++) {
          const int u = tid + 256 * q;
          if (u < ntile) {
            const Tile22 tl = tile22(u, dL, dR);
            mm22(xs, 1, dLi, Bs, bk, bj, dRi, tl, s);
            for (int r = 0; r < 4; r++) acc[q][r] += al * s[r];
          }
        }
      } else {  // y += al x
#pragma unroll
        for (int q = 0; q < PERT; q++) {
          const int u = tid + 256 * q;
          if (u < ntile) {
            const Tile22 tl = tile22(u, dL, dR);
            acc[q][0] += al * xs[tl.i0 + dL * tl.j0];
            acc[q][1] += al * xs[tl.i1 + dL * tl.j0];
            acc[q][2] += al * xs[tl.i0 + dL * tl.j1];
            acc[q][3] += al * xs[tl.i1 + dL * tl.j1];
          }
        }
      }
    }
  }
  double* Y = b.direct ? y + b.out : part + b.out;
#pragma unroll
  for (int q = 0; q < PERT; q++) {
    const int u = tid + 256 * q;
    if (u < ntile) {
      const Tile22 tl = tile22(u, dL, dR);
      const int e00 = tl.i0 + dL * tl.j0;
      const bool r1 = tl.i0 + 1 < dL, c1 = tl.j0 + 1 < dR;
      if (b.direct) {
        Y[e00] += acc[q][0];
        if (r1) Y[e00 + 1] += acc[q][1];
        if (c1) Y[e00 + dL] += acc[q][2];
        if (r1 && c1) Y[e00 + dL + 1] += acc[q][3];
      } else {
        Y[e00] = acc[q][0];
        if (r1) Y[e00 + 1] = acc[q][1];
        if (c1) Y[e00 + dL] = acc[q][2];
        if (r1 && c1) Y[e00 + dL + 1] = acc[q][3];
      }
    }
  }
}
// Chained products on the matrix cores (v_mfma_f64_16x16x4f64), one wavefront per work
// item and no LDS: the wave keeps its output block Y (dL x dR <= 16 NT) as NT x NT
// 16 x 16 accumulator tiles. MFMA operand layout: lane (lr = lane % 16, lg = lane / 16)
// supplies A(lr, k) and B(k, lr) for the k-index lg of each 4-wide step, and receives
// D(4 r + lg, lr) in register r = 0..3. A product's D tile is therefore directly the B
// operand of the next one (register r holds k-index 4 r + lg of step r), and the D tile of a
// transposed product is the A operand: two-operand terms chain through registers.
//   mode 3: T^T = X^T op(A)^T (dRi x dL tiles), then Y += al T op(B)
//   mode 4: T = X op(B) (dLi x dR tiles), then Y += al op(A) T
//   mode 0 (Y += al X) runs as mode 1 with an identity operand
// Operands are read straight from global memory (operator blocks and x stay in the
// caches across the terms of a work item); indices past a dimension read element 0
// and contribute zero.
typedef double chm_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double ld0(const double* __restrict__ p, bool ok, int idx) {
  const double v = p[ok ? idx : 0];
  return ok ? v : 0.0;
}

#define CHM_MFMA(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)

template <int NT>
__global__ __launch_bounds__(256) void k_chain_mfma(const DevCBlock* __restrict__ blocks, int nblocks,
                                                    const DevCGroup* __restrict__ groups,
                                                    const DevCTerm* __restrict__ terms, const double* __restrict__ x,
                                                    double* __restrict__ y, double* __restrict__ part) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nblocks) return;  // the whole wave leaves; the kernel has no barriers
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  const DevCBlock b = blocks[w];
  const int dL = b.dL, dR = b.dR;
  const int nL = (dL + 15) >> 4, nR = (dR + 15) >> 4;
  chm_d4 acc[NT][NT];
#pragma unroll
  for (int i = 0; i < NT; i++)
#pragma unroll
    for (int j = 0; j < NT; j++) acc[i][j] = chm_d4{0.0, 0.0, 0.0, 0.0};
  for (int g = b.g0; g < b.g0 + b.ng; g++) {
    const DevCGroup G = groups[g];
    const int dLi = G.dLi, dRi = G.dRi;
    const double* __restrict__ X = x + G.xoff;  // dLi x dRi, column-major
    for (int t = G.t0; t < G.t0 + G.nt; t++) {
      const DevCTerm T = terms[t];
      const double al = T.alpha;
      // op(A)(i, k) = A[i * ai + k * ak] (dL x dLi), op(B)(k, j) = B[k * bk + j * bj] (dRi x dR)
      const int ai = T.ta ? T.lda : 1, ak = T.ta ? 1 : T.lda;
      const int bk = T.tb ? T.ldb : 1, bj = T.tb ? 1 : T.ldb;
      // absent operands (identity) point at x, so that no address formed is invalid
      const double* __restrict__ A = T.A ? T.A : X;
      const double* __restrict__ B = T.B ? T.B : X;
      const int mode = T.mode;
      if (mode == 3) {
        const int nRi = (dRi + 15) >> 4;
        chm_d4 t1[NT][NT];
#pragma unroll
        for (int i = 0; i < NT; i++)
#pragma unroll
          for (int j = 0; j < NT; j++) t1[i][j] = chm_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int k = lg; k < ((dLi + 3) & ~3); k += 4) {
          double av[NT], bv[NT];
#pragma unroll
          for (int i = 0; i < NT; i++) av[i] = ld0(X, k < dLi && 16 * i + lr < dRi, k + dLi * (16 * i + lr));
#pragma unroll
          for (int j = 0; j < NT; j++) bv[j] = ld0(A, k < dLi && 16 * j + lr < dL, (16 * j + lr) * ai + k * ak);
#pragma unroll
          for (int i = 0; i < NT; i++)
            if (i < nRi)
#pragma unroll
              for (int j = 0; j < NT; j++)
                if (j < nL) t1[i][j] = CHM_MFMA(av[i], bv[j], t1[i][j]);
        }
#pragma unroll
        for (int i = 0; i < NT; i++) {
          if (i >= nRi) break;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int k = 16 * i + 4 * r + lg;
            double bv[NT];
#pragma unroll
            for (int j = 0; j < NT; j++) bv[j] = ld0(B, k < dRi && 16 * j + lr < dR, k * bk + (16 * j + lr) * bj);
#pragma unroll
            for (int yi = 0; yi < NT; yi++)
              if (yi < nL) {
                const double a = al * t1[i][yi][r];
#pragma unroll
                for (int yj = 0; yj < NT; yj++)
                  if (yj < nR) acc[yi][yj] = CHM_MFMA(a, bv[yj], acc[yi][yj]);
              }
          }
        }
      } else if (mode == 4) {
        const int nLi = (dLi + 15) >> 4;
        chm_d4 t1[NT][NT];
#pragma unroll
        for (int i = 0; i < NT; i++)
#pragma unroll
          for (int j = 0; j < NT; j++) t1[i][j] = chm_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int k = lg; k < ((dRi + 3) & ~3); k += 4) {
          double av[NT], bv[NT];
#pragma unroll
          for (int i = 0; i < NT; i++) av[i] = ld0(X, 16 * i + lr < dLi && k < dRi, 16 * i + lr + dLi * k);
#pragma unroll
          for (int j = 0; j < NT; j++) bv[j] = ld0(B, k < dRi && 16 * j + lr < dR, k * bk + (16 * j + lr) * bj);
#pragma unroll
          for (int i = 0; i < NT; i++)
            if (i < nLi)
#pragma unroll
              for (int j = 0; j < NT; j++)
                if (j < nR) t1[i][j] = CHM_MFMA(av[i], bv[j], t1[i][j]);
        }
#pragma unroll
        for (int i = 0; i < NT; i++) {
          if (i >= nLi) break;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int k = 16 * i + 4 * r + lg;
            double av[NT];
#pragma unroll
            for (int yi = 0; yi < NT; yi++) av[yi] = al * ld0(A, 16 * yi + lr < dL && k < dLi, (16 * yi + lr) * ai + k * ak);
#pragma unroll
            for (int yi = 0; yi < NT; yi++)
              if (yi < nL)
#pragma unroll
                for (int yj = 0; yj < NT; yj++)
                  if (yj < nR) acc[yi][yj] = CHM_MFMA(av[yi], t1[i][yj][r], acc[yi][yj]);
          }
        }
      } else if (mode <= 1) {  // Y += al op(A) X, X is dLi x dR; mode 0: op(A) = identity
#pragma unroll 1
        for (int k = lg; k < ((dLi + 3) & ~3); k += 4) {
          double av[NT], bv[NT];
#pragma unroll
          for (int i = 0; i < NT; i++)
            av[i] = mode == 0 ? (16 * i + lr == k ? al : 0.0)
                              : al * ld0(A, 16 * i + lr < dL && k < dLi, (16 * i + lr) * ai + k * ak);
#pragma unroll
          for (int j = 0; j < NT; j++) bv[j] = ld0(X, k < dLi && 16 * j + lr < dR, k + dLi * (16 * j + lr));
#pragma unroll
          for (int i = 0; i < NT; i++)
            if (i < nL)
#pragma unroll
              for (int j = 0; j < NT; j++)
                if (j < nR) acc[i][j] = CHM_MFMA(av[i], bv[j], acc[i][j]);
        }
      } else if (mode == 2) {  // Y += al X op(B), X is dL x dRi
#pragma unroll 1
        for (int k = lg; k < ((dRi + 3) & ~3); k += 4) {
          double av[NT], bv[NT];
#pragma unroll
          for (int i = 0; i < NT; i++) av[i] = al * ld0(X, 16 * i + lr < dL && k < dRi, 16 * i + lr + dL * k);
#pragma unroll
          for (int j = 0; j < NT; j++) bv[j] = ld0(B, k < dRi && 16 * j + lr < dR, k * bk + (16 * j + lr) * bj);
#pragma unroll
          for (int i = 0; i < NT; i++)
            if (i < nL)
#pragma unroll
              for (int j = 0; j < NT; j++)
                if (j < nR) acc[i][j] = CHM_MFMA(av[i], bv[j], acc[i][j]);
        }
      }
    }
  }
  double* Y = b.direct ? y + b.out : part + b.out;
#pragma unroll
  for (int i = 0; i < NT; i++)
#pragma unroll
    for (int j = 0; j < NT; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int l = 16 * i + 4 * r + lg, c = 16 * j + lr;
        if (l < dL && c < dR) {
          if (b.direct) Y[l + dL * c] += acc[i][j][r];
          else Y[l + dL * c] = acc[i][j][r];
        }
      }
}
#undef CHM_MFMA

// Chained products of blocks up to FCHAIN_MAXD on the matrix cores, one workgroup of four
// wavefronts per work item (an output block and a range of its input groups, the
// records of k_chain), output-stationary. Every term of an item has one of the forms
//   mode 0: Y += a X,  1: Y += op(A) X,  2: Y += X op(B),  3: Y += (op(A) X) op(B);
// the other association order, op(A) (X op(B)), is mode 3 of the transposed block: an
// item with tr set computes Y^T += op(B)^T X^T op(A)^T, its terms' operators exchanged
// and their transposition flags flipped on the host. Wavefront w owns the 16-row strip w
// of Y and holds it as 16 x 16 accumulator tiles (a fifth strip is shared by all four,
// each with part of every term's work, and its partial sums are added at the end of
// the item). A two-operand term forms
// its strip of the intermediate transposed, W^T = X^T op(A)^T, whose result tiles are
// directly the A operands of W op(B) (the register chaining of k_chain_mfma), so the
// intermediate never leaves registers. op(A), op(B) (the factor a folded into one of
// them) and, per input group, X are staged in LDS zero-padded to whole tiles, with fixed
// leading dimensions (compile-time offsets; op(A) is read along rows, 16 mod 32, the
// others along columns, 4 mod 8, against bank conflicts); the next term's operators
// are loaded into registers while the current term computes, and stored after the
// barrier that ends it.
//   f64: v_mfma_f64_16x16x4f64, result register r of lane (lr, lg) holds row 4 r + lg;
//   f32 (the mixed form's low-precision products): v_mfma_f32_16x16x4f32, row 4 lg + r.
// In a chained product lane group lg supplies the k index of its register, so the k
// order of a step follows the result layout (any bijection of k onto the lanes serves).
typedef float chm_f4 __attribute__((ext_vector_type(4)));

template <bool F32>
struct FcMath;
template <>
struct FcMath<false> {
  typedef double R;
  typedef chm_d4 V4;
  static __device__ __forceinline__ V4 mma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int r, int lg) { return 4 * r + lg; }
};
template <>
struct FcMath<true> {
  typedef float R;
  typedef chm_f4 V4;
  static __device__ __forceinline__ V4 mma(R a, R b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int r, int lg) { return 4 * lg + r; }
};

template <int NT>
struct FcShape {
  static constexpr int D = 16 * NT;
  static constexpr int LDA = (NT & 1) ? D : D + 16;
  static constexpr int LDB = D + 4;
  static constexpr int MS = (NT + 3) / 4;  // strips per wavefront
  static constexpr int NP = NT * NT;       // staged elements per thread
  static constexpr int ELEMS = LDA * D + 2 * LDB * D;
};

// Loads the nf x ns matrix p (element (f, s) at p[f + ld s]) padded to whole 16 x 16
// tiles: thread (f0, s0) = (tid % 16, tid / 16) takes (f0 + 16 fi, s0 + 16 si). The loads
// are global (not flat: those would also hold the LDS counter the compute waits on), and a
// padding element reads the block's first element and is replaced by zero.
typedef __attribute__((address_space(1))) const double gdouble;
template <int NT, typename P>
__device__ __forceinline__ void fc_load(P (&v)[NT * NT], const double* __restrict__ p, int64_t ld, int nf, int ns, int tid) {
  const int f0 = tid & 15, s0 = tid >> 4;
  const int nf16 = (nf + 15) >> 4, ns16 = (ns + 15) >> 4;
  gdouble* g = (gdouble*)p;
#pragma unroll
  for (int si = 0; si < NT; si++) {
    if (si >= ns16) break;
#pragma unroll
    for (int fi = 0; fi < NT; fi++) {
      if (fi >= nf16) break;
      const int f = f0 + 16 * fi, s = s0 + 16 * si;
      const bool ok = f < nf && s < ns;
      const double e = g[ok ? f + ld * s : 0];
      v[si * NT + fi] = ok ? (P)e : (P)0;
    }
  }
}

// Stores what fc_load loaded into an LDS matrix with leading dimension LD: element
// (f, s) at (f, s), or at (s, f) when sw, scaled by sc
template <int NT, int LD, typename R, typename P>
__device__ __forceinline__ void fc_store(R* __restrict__ dst, const P (&v)[NT * NT], bool sw, int nf, int ns, int tid,
                                         double sc) {
  const int f0 = tid & 15, s0 = tid >> 4;
  const int nf16 = (nf + 15) >> 4, ns16 = (ns + 15) >> 4;
#pragma unroll
  for (int si = 0; si < NT; si++) {
    if (si >= ns16) break;
#pragma unroll
    for (int fi = 0; fi < NT; fi++) {
      if (fi >= nf16) break;
      const int f = f0 + 16 * fi, s = s0 + 16 * si;
      dst[sw ? s + LD * f : f + LD * s] = (R)(sc * v[si * NT + fi]);
    }
  }
}

template <int NT, int NJ, bool F32>
__global__ __launch_bounds__(256) void k_fchain(const DevCBlock* __restrict__ blocks, const DevCGroup* __restrict__ groups,
                                               const DevCTerm* __restrict__ terms, const double* __restrict__ x,
                                               double* __restrict__ y, double* __restrict__ part) {
  typedef FcMath<F32> M;
  typedef typename M::R R;
  typedef typename M::V4 V4;
  typedef FcShape<NT> SH;
  constexpr int LDA = SH::LDA, LDB = SH::LDB, MS = SH::MS, NP = SH::NP;
  extern __shared__ __align__(16) unsigned char fc_lds[];
  R* __restrict__ As = (R*)fc_lds;
  R* __restrict__ Bs = As + LDA * SH::D;
  R* __restrict__ Xs = Bs + LDB * SH::D;
  const DevCBlock b = blocks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool tr = b.tr != 0;
  const int dLp = tr ? b.dR : b.dL, dRp = tr ? b.dL : b.dR;  // the block as computed
  const int nL = (dLp + 15) >> 4;  // strips; the block has NJ column tiles
  V4 acc[MS][NJ];
#pragma unroll
  for (int m = 0; m < MS; m++)
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[m][j] = V4{0, 0, 0, 0};

  R pa[NP], pb[NP];  // the next term's operators (FP32 variant: rounded as loaded)
  // operators of term T (input group G) into registers: op(A) is dLp x dLi', stored with
  // element (i, k) at A[i + lda k], or at A[i lda + k] when ta; op(B) is dRi' x dRp
  auto load_ops = [&](const DevCTerm& T, const DevCGroup& G) {
    const int dLip = tr ? G.dRi : G.dLi, dRip = tr ? G.dLi : G.dRi;
    if (T.A) fc_load<NT, R>(pa, T.A, T.lda, T.ta ? dLip : dLp, T.ta ? dLp : dLip, tid);
    if (T.B) fc_load<NT, R>(pb, T.B, T.ldb, T.tb ? dRp : dRip, T.tb ? dRip : dRp, tid);
  };
  auto store_ops = [&](const DevCTerm& T, const DevCGroup& G) {
    const int dLip = tr ? G.dRi : G.dLi, dRip = tr ? G.dLi : G.dRi;
    if (T.A) fc_store<NT, LDA>(As, pa, T.ta != 0, T.ta ? dLip : dLp, T.ta ? dLp : dLip, tid, T.alpha);
    if (T.B) fc_store<NT, LDB>(Bs, pb, T.tb != 0, T.tb ? dRp : dRip, T.tb ? dRip : dRp, tid, T.A ? 1.0 : T.alpha);
  };
  // x of group G (dLi x dRi, column-major) as X' = X, or X^T when tr
  auto stage_x = [&](const DevCGroup& G) {
    R px[NP];
    fc_load<NT, R>(px, x + G.xoff, G.dLi, G.dLi, G.dRi, tid);
    fc_store<NT, LDB>(Xs, px, tr, G.dLi, G.dRi, tid, 1.0);
  };

  const int gend = b.g0 + b.ng;
  int g = b.g0;
  DevCGroup G = groups[g];
  int t = G.t0;
  DevCTerm T = terms[t];
  stage_x(G);
  load_ops(T, G);
  store_ops(T, G);
  __syncthreads();
  for (;;) {
    int tn = t + 1, gn = g;
    DevCGroup Gn = G;
    if (tn == G.t0 + G.nt) {
      gn = g + 1;
      if (gn < gend) {
        Gn = groups[gn];
        tn = Gn.t0;
      }
    }
    const bool more = gn < gend;
    DevCTerm Tn = T;
    if (more) {
      Tn = terms[tn];
      load_ops(Tn, Gn);  // in flight while this term computes
    }
    const int nLi = tr ? (G.dRi + 15) >> 4 : (G.dLi + 15) >> 4;
    const int nRi = tr ? (G.dLi + 15) >> 4 : (G.dRi + 15) >> 4;
    const int mode = T.mode;
    // Every wavefront runs the same instruction stream: one without a strip of its own in
    // a pass repeats the last strip (its sums are not stored). The tile loops have fixed trip
    // counts, so that the accumulators stay in place (tiles past the block's own see only
    // padding or other data and are discarded); the k loops run over the block's size.
#pragma unroll
    for (int m = 0; m < MS; m++) {
      if (4 * m >= nL) break;  // no strip of this pass in the block (uniform)
      // pass 0: wavefront w owns strip w; pass 1 (blocks of 5 strips): the four
      // wavefronts share strip 4, each taking every fourth k-block or intermediate tile
      // into a partial sum that the end of the item adds up
      const bool shared = m > 0;
      const int s = shared ? 4 : min(w, nL - 1);
      const int q0 = shared ? w : 0, dq = shared ? 4 : 1;
      const int ra = 16 * s + lr;  // this lane's row of op(A), and of X in mode 2
      if (mode == 3) {
#pragma unroll 1
        for (int jj = q0; jj < nRi; jj += dq) {
          V4 wt = V4{0, 0, 0, 0};  // tile jj of W^T = X^T op(A)^T
#pragma unroll 1
          for (int kb = 0; kb < nLi; kb++) {
            const R* __restrict__ ap = As + ra + LDA * (16 * kb + lg);
            const R* __restrict__ xp = Xs + (16 * kb + lg) + LDB * (16 * jj + lr);
#pragma unroll
            for (int u = 0; u < 4; u++) wt = M::mma(xp[4 * u], ap[LDA * 4 * u], wt);
          }
          // Y += W op(B) over the 16 k of tile jj, W's tile as the A operands
          const R* __restrict__ bp = Bs + (16 * jj + M::row(0, lg)) + LDB * lr;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const R a = wt[r];
#pragma unroll
            for (int j = 0; j < NJ; j++) acc[m][j] = M::mma(a, bp[M::row(r, 0) + LDB * 16 * j], acc[m][j]);
          }
        }
      } else if (mode == 1) {
#pragma unroll 1
        for (int kb = q0; kb < nLi; kb += dq) {
          const R* __restrict__ ap = As + ra + LDA * (16 * kb + lg);
          const R* __restrict__ xp = Xs + (16 * kb + lg) + LDB * lr;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const R a = ap[LDA * 4 * u];
#pragma unroll
            for (int j = 0; j < NJ; j++) acc[m][j] = M::mma(a, xp[4 * u + LDB * 16 * j], acc[m][j]);
          }
        }
      } else if (mode == 2) {
#pragma unroll 1
        for (int kb = q0; kb < nRi; kb += dq) {
          const R* __restrict__ xp = Xs + ra + LDB * (16 * kb + lg);
          const R* __restrict__ bp = Bs + (16 * kb + lg) + LDB * lr;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const R a = xp[LDB * 4 * u];
#pragma unroll
            for (int j = 0; j < NJ; j++) acc[m][j] = M::mma(a, bp[4 * u + LDB * 16 * j], acc[m][j]);
          }
        }
      } else if (!shared || w == 0) {
        const R al = (R)T.alpha;
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) acc[m][j][r] += al * Xs[(16 * s + M::row(r, lg)) + LDB * (16 * j + lr)];
      }
    }
    __syncthreads();  // every wavefront is done with this term's LDS
    if (!more) break;
    if (gn != g) stage_x(Gn);
    store_ops(Tn, Gn);
    __syncthreads();
    t = tn;
    g = gn;
    G = Gn;
    T = Tn;
  }
  if (MS > 1 && nL > 4) {  // the shared strip: partial sums through LDS (the term loop ended on a barrier)
    R* __restrict__ red = As;
#pragma unroll
    for (int j = 0; j < NJ; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) red[((w * NJ + j) * 4 + r) * 64 + lane] = acc[MS - 1][j][r];
    __syncthreads();
    if (w == 0)
#pragma unroll
      for (int j = 0; j < NJ; j++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          R v = red[(j * 4 + r) * 64 + lane];
#pragma unroll
          for (int q = 1; q < 4; q++) v += red[((q * NJ + j) * 4 + r) * 64 + lane];
          acc[MS - 1][j][r] = v;
        }
  }
  double* __restrict__ Y = b.direct ? y + b.out : part + b.out;
#pragma unroll
  for (int m = 0; m < MS; m++) {
    const int s = m > 0 ? 4 : w;  // the shared strip is stored by wavefront 0
    if (s >= nL || (m > 0 && w != 0)) break;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int l = 16 * s + M::row(r, lg), c = 16 * j + lr;  // element (l, c) of the block as computed
        if (l < dLp && c < dRp) {
          const int64_t o = tr ? c + (int64_t)b.dL * l : l + (int64_t)b.dL * c;
          if (b.direct) Y[o] += (double)acc[m][j][r];
          else Y[o] = (double)acc[m][j][r];
        }
      }
    }
  }
}


constexpr int CHAIN_PER_S = (18 * 18 + 255) / 256;                                  // 2 x 2 tiles of a 36 x 36 block
constexpr int CHAIN_PER_L = ((CHAIN_MAXD / 2) * (CHAIN_MAXD / 2) + 255) / 256;     // of a 64 x 64 block

// Small-block update products (see DevOItem). Dynamic LDS: 2 arrays of P doubles.
template <int PERT>
__global__ __launch_bounds__(256) void k_ops(const DevOItem* __restrict__ items, const DevGemmOp* __restrict__ ops, int P) {
  extern __shared__ double lds[];
  double* As = lds;
  double* Bs = lds + P;
  const DevOItem it = items[blockIdx.x];
  const int tid = threadIdx.x;
  const int m = it.m, n = it.n;
  const int ntile = ((m + 1) >> 1) * ((n + 1) >> 1);
  double acc[PERT][4];
#pragma unroll
  for (int q = 0; q < PERT; q++) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.0;
  for (int o = it.op0; o < it.op0 + it.nop; o++) {
    const DevGemmOp op = ops[o];
    const double al = op.alpha;
    const int ai = op.ta ? op.lda : 1, ak = op.ta ? 1 : op.lda;
    if (op.kind == 1) {  // acc += al op(A), op(A) m x n read in place
#pragma unroll
      for (int q = 0; q < PERT; q++) {
        const int u = tid + 256 * q;
        if (u < ntile) {
          const Tile22 tl = tile22(u, m, n);
          acc[q][0] += al * op.A[tl.i0 * ai + tl.j0 * ak];
          acc[q][1] += al * op.A[tl.i1 * ai + tl.j0 * ak];
          acc[q][2] += al * op.A[tl.i0 * ai + tl.j1 * ak];
          acc[q][3] += al * op.A[tl.i1 * ai + tl.j1 * ak];
        }
      }
      continue;
    }
    const int k = op.k;
    const int bk = op.tb ? op.ldb : 1, bj = op.tb ? 1 : op.ldb;
    __syncthreads();  // every reader of the previous copies is done
    for (int e = tid; e < m * k; e += 256) As[e] = op.A[e];
    for (int e = tid; e < k * n; e += 256) Bs[e] = op.B[e];
    __syncthreads();
    double s[4];
#pragma unroll
    for (int q = 0; q < PERT; q++) {
      const int u = tid + 256 * q;
      if (u < ntile) {
        const Tile22 tl = tile22(u, m, n);
        mm22(As, ai, ak, Bs, bk, bj, k, tl, s);
        for (int r = 0; r < 4; r++) acc[q][r] += al * s[r];
      }
    }
  }
  double s4[4];
#pragma unroll
  for (int q = 0; q < PERT; q++) {
    const int u = tid + 256 * q;
    if (u < ntile) {
      const Tile22 tl = tile22(u, m, n);
      for (int r = 0; r < 4; r++) s4[r] = acc[q][r];
      store22(it.out, m, m, n, tl, s4);
    }
  }
}
constexpr int OPS_PER = ((OPS_MAXD / 2) * (OPS_MAXD / 2) + 255) / 256;

__global__ __launch_bounds__(256) void k_part_reduce(const DevPRed* __restrict__ reds, const double* __restrict__ part) {
  const DevPRed r = reds[blockIdx.x];
  for (int e = threadIdx.x; e < r.R; e += 256) {
    double s = 0.0;
    for (int q = 0; q < r.nparts; q++) s += part[r.p0 + (int64_t)q * r.R + e];
    r.out[e] = s;
  }
}

// y[yoff + e] += sum_q part[p0 + q n + e] over the work items of a split output block
__global__ __launch_bounds__(256) void k_chain_reduce(const DevBRed* __restrict__ reds, const double* __restrict__ part,
                                                      double* __restrict__ y) {
  const DevBRed r = reds[blockIdx.x];
  for (int e = threadIdx.x; e < r.R; e += 256) {
    double s = 0.0;
    for (int q = 0; q < r.nparts; q++) s += part[r.p0 + (int64_t)q * r.R + e];
    y[r.yoff + e] += s;
  }
}

// --------------------------------------------------------- vector kernels --
__global__ __launch_bounds__(256) void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                     int64_t n, double* __restrict__ partial) {
  __shared__ double s[256];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) v += a[i] * b[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Sum nparts partial sums per output (one workgroup per output), fixed order.
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ partial, int nparts, double* __restrict__ out) {
  __shared__ double s[256];
  const double* p = partial + (size_t)blockIdx.x * nparts;
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) v += p[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// partial dots of many vectors against one: grid (nparts, nv)
__global__ __launch_bounds__(256) void k_multi_dot_partial(const double* const* __restrict__ vecs,
                                                           const double* __restrict__ b, int64_t n,
                                                           double* __restrict__ partial) {
  __shared__ double s[256];
  const double* a = vecs[blockIdx.y];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) v += a[i] * b[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s[0];
}

// y += (sign * *coef) x, with the coefficient read from device memory
__global__ __launch_bounds__(256) void k_axpy_dev(const double* __restrict__ coef, double sign,
                                                  const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const double a = sign * coef[0];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] += a * x[i];
}

__global__ __launch_bounds__(256) void k_axpy(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] += a * x[i];
}

// d[i] += sum over the nv vectors of v[q * n + i]^2, vectors in order (the diagonal of the
// lower states' projector terms, HeffDiagonal.cpp:621-642).
__global__ __launch_bounds__(256) void k_add_squares(const double* __restrict__ v, int nv, double* __restrict__ d,
                                                     int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double a = d[i];
    for (int q = 0; q < nv; q++) {
      const double w = v[(int64_t)q * n + i];
      a += w * w;
    }
    d[i] = a;
  }
}

__global__ __launch_bounds__(256) void k_scale(double a, double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) x[i] *= a;
}

// out = sum_v coefs[v] * vecs[v]
__global__ __launch_bounds__(256) void k_lincomb_vecs(const double* const* __restrict__ vecs,
                                                      const double* __restrict__ coefs, int nv,
                                                      double* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v = 0.0;
    for (int q = 0; q < nv; q++) v += coefs[q] * vecs[q][i];
    out[i] = v;
  }
}

// Davidson correction (Davidson.cpp:322-350). stage 0: work = u / K; stage 1:
// t = -(t + alpha u) / K, with K = diag - shift clamped to +-cutoff.
__global__ __launch_bounds__(256) void k_precond(const double* __restrict__ diag, const double* __restrict__ u,
                                                 double shift, double cutoff, double* __restrict__ work,
                                                 double* __restrict__ t, int64_t n, int stage, double alpha,
                                                 const float* __restrict__ m, const double* __restrict__ dshift,
                                                 const double* __restrict__ dalpha) {
  if (dshift) shift = *dshift;
  if (dalpha) alpha = *dalpha;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double d = diag[i] - shift;
    const double k = (fabs(d) > cutoff) ? d : cutoff;
    const double kk = m ? k / fmin(10.0, fmax(0.1, 1.0 + (double)m[i])) : k;  // learned factor kept in [0.1, 10]
    if (stage == 0) work[i] = u[i] / kk;
    else t[i] = -(t[i] + alpha * u[i]) / kk;
  }
}

// Scale block kappa by scale[kappa] (program <-> symmetric convention, Sobject.cpp:624-650).
__global__ __launch_bounds__(256) void k_symm_scale(double* __restrict__ x, const int64_t* __restrict__ k2i,
                                                    const double* __restrict__ scale, int nk) {
  const int k = blockIdx.x;
  if (k >= nk) return;
  const double s = scale[k];
  for (int64_t i = k2i[k] + threadIdx.x; i < k2i[k + 1]; i += 256) x[i] *= s;
}

// Deterministic pseudo-random fill in [-scale, scale) (splitmix64 of the index).
__global__ __launch_bounds__(256) void k_fill_hash(double* __restrict__ p, int64_t n, uint64_t seed, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)i;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    p[i] = scale * (2.0 * ((z >> 11) * (1.0 / 9007199254740992.0)) - 1.0);
  }
}

// Scaled block copies with up to two sources (one workgroup per item).
__global__ __launch_bounds__(256) void k_copies(const DevCopy* __restrict__ items) {
  const DevCopy it = items[blockIdx.x];
  const int n = it.rows * it.cols;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int r = e % it.rows, c = e / it.rows;
    double v = 0.0;
    if (r < it.keep_rows && c < it.keep_cols) {
      const size_t si = (size_t)r + (size_t)it.lds * c;
      v = it.c0 * it.src0[si];
      if (it.src1) v += it.c1 * it.src1[si];
      if (it.rowscale) v *= it.rowscale[r];
      if (it.colscale) v *= it.colscale[c];
    }
    if (it.trans) it.dst[(size_t)c + (size_t)it.ldd * r] = v;
    else it.dst[(size_t)r + (size_t)it.ldd * c] = v;
  }
}

// One-sided (Hestenes) Jacobi SVD of the small centre-sector matrices of Split, one
// workgroup of SVD_THREADS per matrix. Each round rotates a set of disjoint column pairs
// (round-robin tournament order), every pair on a group of SVD_TPP lanes, all n / 2 pairs
// of a round at once for n up to 2 SVD_GROUPS (one barrier per round); sweeps end when no pair
// has |a_p . a_q| > sqrt(m) eps |a_p| |a_q|, the rounding floor of an m-term dot product
// (the stopping rule of one-sided Jacobi in Drmac & Veselic, SIAM J. Matrix Anal. Appl.
// 29 (2008) 1322, as LAPACK's dgesvj uses it). A job that reaches SVD_MAX_SWEEPS without
// meeting it sets *info = 1. Columns whose norm is at most eps |A|_F are numerically zero:
// they are not rotated (their directions are rounding noise and would keep the sweeps
// going), their singular values are set to zero, and their vectors on the m-long side are
// completed to an orthonormal set like those of exactly zero singular values.
constexpr int SVD_THREADS = 1024;
constexpr int SVD_TPP = 8;
constexpr int SVD_GROUPS = SVD_THREADS / SVD_TPP;
constexpr double SVD_EPS = 1.1102230246251565e-16;  // 2^-53
constexpr int SVD_MAX_SWEEPS = 80;

__device__ inline double group_sum(double v) {
  for (int o = SVD_TPP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, SVD_TPP);
  return v;
}

__global__ __launch_bounds__(SVD_THREADS) void k_svd_jacobi(const DevSvd* __restrict__ items) {
  const DevSvd it = items[blockIdx.x];
  const int m = it.m, n = it.n;
  const int tid = threadIdx.x, g = tid / SVD_TPP, l = tid % SVD_TPP;
  double* A = it.A;
  double* Q = it.Q;
  __shared__ int rotated;
  __shared__ double sig[SVD_JACOBI_MAXN];
  __shared__ int src[SVD_JACOBI_MAXN];
  __shared__ double fro[SVD_THREADS];
  double f2 = 0.0;
  for (int e = tid; e < m * n; e += SVD_THREADS) f2 += A[e] * A[e];
  fro[tid] = f2;
  for (int e = tid; e < n * n; e += SVD_THREADS) Q[e] = (e % n == e / n) ? 1.0 : 0.0;
  __syncthreads();
  for (int h = SVD_THREADS / 2; h > 0; h >>= 1) {
    if (tid < h) fro[tid] += fro[tid + h];
    __syncthreads();
  }
  const double floor2 = SVD_EPS * SVD_EPS * fro[0];
  const int n2 = n + (n & 1);
  const double tol = sqrt((double)m) * SVD_EPS;
  bool converged = n <= 1;
  const int max_sweeps = it.max_sweeps > 0 ? min(it.max_sweeps, SVD_MAX_SWEEPS) : SVD_MAX_SWEEPS;
  for (int sweep = 0; sweep < max_sweeps && n > 1; sweep++) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < n2 - 1; r++) {
      for (int k0 = 0; k0 < n2 / 2; k0 += SVD_GROUPS) {
        const int k = k0 + g;
        int p = -1, q = -1;
        if (k < n2 / 2) {
          p = (k == 0) ? 0 : ((k - 1 + r) % (n2 - 1)) + 1;
          q = ((n2 - 2 - k + r) % (n2 - 1)) + 1;
          if (p > q) { const int t = p; p = q; q = t; }
          if (q >= n) p = q = -1;
        }
        double al = 0.0, be = 0.0, ga = 0.0;
        if (p >= 0)
          for (int i = l; i < m; i += SVD_TPP) {
            const double x = A[i + (size_t)m * p], y = A[i + (size_t)m * q];
            al += x * x; be += y * y; ga += x * y;
          }
        al = group_sum(al); be = group_sum(be); ga = group_sum(ga);
        if (p >= 0 && ga != 0.0 && al > floor2 && be > floor2 && al > 0.0 && be > 0.0 && fabs(ga) > tol * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
          for (int i = l; i < m; i += SVD_TPP) {
            const double x = A[i + (size_t)m * p], y = A[i + (size_t)m * q];
            A[i + (size_t)m * p] = c * x - s * y;
            A[i + (size_t)m * q] = s * x + c * y;
          }
          for (int i = l; i < n; i += SVD_TPP) {
            const double x = Q[i + (size_t)n * p], y = Q[i + (size_t)n * q];
            Q[i + (size_t)n * p] = c * x - s * y;
            Q[i + (size_t)n * q] = s * x + c * y;
          }
          if (l == 0) rotated = 1;
        }
        __syncthreads();
      }
    }
    __syncthreads();
    const int any = rotated;
    __syncthreads();
    if (!any) {
      converged = true;
      break;
    }
  }
  if (tid == 0 && it.info) *it.info = converged ? 0 : 1;
  // singular values and their descending order
  for (int j = g; j < n; j += SVD_GROUPS) {
    double s2 = 0.0;
    for (int i = l; i < m; i += SVD_TPP) s2 += A[i + (size_t)m * j] * A[i + (size_t)m * j];
    s2 = group_sum(s2);
    if (l == 0) sig[j] = s2 > floor2 ? sqrt(s2) : 0.0;
  }
  __syncthreads();
  for (int j = tid; j < n; j += SVD_THREADS) {
    int rank = 0;
    for (int k = 0; k < n; k++) rank += (sig[k] > sig[j] || (sig[k] == sig[j] && k < j)) ? 1 : 0;
    src[rank] = j;
  }
  __syncthreads();
  // M = sum_j sig_j u_j v_j^T: u_j = a_j / sig_j and v_j = q_j (tall), or u_j = q_j and
  // v_j = a_j / sig_j (wide, A = M^T)
  const int ldU = it.wide ? n : m;
  for (int t = g; t < n; t += SVD_GROUPS) {
    const int j = src[t];
    const double sj = sig[j];
    const double inv = sj > 1e-250 ? 1.0 / sj : 0.0;
    if (l == 0) it.S[t] = sj;
    if (!it.wide) {
      for (int i = l; i < m; i += SVD_TPP) it.U[i + (size_t)ldU * t] = A[i + (size_t)m * j] * inv;
      for (int i = l; i < n; i += SVD_TPP) it.VT[t + (size_t)n * i] = Q[i + (size_t)n * j];
    } else {
      for (int i = l; i < n; i += SVD_TPP) it.U[i + (size_t)ldU * t] = Q[i + (size_t)n * j];
      for (int i = l; i < m; i += SVD_TPP) it.VT[t + (size_t)n * i] = A[i + (size_t)m * j] * inv;
    }
  }
  __syncthreads();
  // vectors of (numerically) zero singular values: complete the m-long side to an
  // orthonormal set by Gram-Schmidt on unit vectors (one wave, in order; every lane
  // reads and writes only the elements i = lane (mod 64), lanes meet in shuffles)
  if (tid < 64) {
    for (int t = 0; t < n; t++) {
      if (sig[src[t]] > 1e-250) continue;
      double* base = it.wide ? it.VT : it.U;
      const size_t rs = it.wide ? (size_t)n : 1;         // stride between elements
      const size_t cs = it.wide ? 1 : (size_t)ldU;       // stride between vectors
      for (int e = 0; e < m; e++) {
        // v = e_e - sum_{t' < t} w_t' w_t'[e], accepted when |v| > 0.5
        for (int i = tid; i < m; i += 64) base[i * rs + t * cs] = (i == e) ? 1.0 : 0.0;
        for (int u = 0; u < t; u++) {
          double d = 0.0;
          for (int i = tid; i < m; i += 64) d += base[i * rs + u * cs] * base[i * rs + t * cs];
          for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
          for (int i = tid; i < m; i += 64) base[i * rs + t * cs] -= d * base[i * rs + u * cs];
        }
        double nn = 0.0;
        for (int i = tid; i < m; i += 64) nn += base[i * rs + t * cs] * base[i * rs + t * cs];
        for (int o = 32; o > 0; o >>= 1) nn += __shfl_xor(nn, o, 64);
        if (nn > 0.25) {
          const double f = 1.0 / sqrt(nn);
          for (int i = tid; i < m; i += 64) base[i * rs + t * cs] *= f;
          break;
        }
      }
    }
  }
}

// ---------------------------------------------------------- Davidson on the device ---
// The scalars and small matrices of one Davidson solve (Davidson.cpp:105-191, 214-350) in
// a DevDav record on the device; the kernels read the basis size from it, so the same
// launch sequence (and one captured graph) serves every iteration.
__global__ __launch_bounds__(256) void k_dav_store(const DevDav* __restrict__ st, double* const* __restrict__ vec,
                                                   double* const* __restrict__ hv, const double* __restrict__ x,
                                                   const double* __restrict__ y, int64_t n) {
  const int c = st->nv;
  double* __restrict__ v = vec[c];
  double* __restrict__ h = hv[c];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    v[i] = x[i];
    h[i] = y[i];
  }
}

// partial[c * gridDim.x + b] = chunk b of a[c] . b for c < nv + add
__global__ __launch_bounds__(256) void k_dav_dots_partial(const DevDav* __restrict__ st, int add,
                                                          const double* const* __restrict__ a,
                                                          const double* __restrict__ b, int64_t n,
                                                          double* __restrict__ partial) {
  __shared__ double red[256];
  const int c = blockIdx.y;
  if (c >= st->nv + add) return;  // uniform over the block: no barrier is skipped by part of it
  const double* __restrict__ p = a[c];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s = fma(p[i], b[i], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)c * gridDim.x + blockIdx.x] = red[0];
}

// mode 0: M[c, nv] = M[nv, c] = sum of the partials (the new column of V^T H V);
// mode 1: proj[c] = sum of the partials (projections of the new vector)
__global__ __launch_bounds__(256) void k_dav_dots_reduce(DevDav* __restrict__ st, int add,
                                                         const double* __restrict__ partial, int parts, int mode) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  const int nv = st->nv;
  if (c >= nv + add) return;
  double s = 0.0;
  for (int q = threadIdx.x; q < parts; q += 256) s += partial[(int64_t)c * parts + q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mode == 0) {
      st->M[c + DAV_MAXV * nv] = red[0];
      st->M[nv + DAV_MAXV * c] = red[0];
    } else {
      st->proj[c] = red[0];
    }
  }
}

// t -= sum_{c < nv} proj[c] vec[c]
__global__ __launch_bounds__(256) void k_dav_project_out(const DevDav* __restrict__ st,
                                                         const double* const* __restrict__ vec, double* __restrict__ t,
                                                         int64_t n) {
  __shared__ double pr[DAV_MAXV];
  __shared__ const double* vp[DAV_MAXV];
  const int nv = st->nv;
  if (threadIdx.x < DAV_MAXV) {
    pr[threadIdx.x] = threadIdx.x < nv ? st->proj[threadIdx.x] : 0.0;
    vp[threadIdx.x] = threadIdx.x < nv ? vec[threadIdx.x] : nullptr;
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v = t[i];
    for (int c = 0; c < nv; c++) v -= pr[c] * vp[c][i];
    t[i] = v;
  }
}

// out = t / |t| with |t|^2 = st->nrm2; a zero or invalid norm marks the solve as failed
__global__ __launch_bounds__(256) void k_dav_normalize(DevDav* __restrict__ st, const double* __restrict__ t,
                                                       double* __restrict__ out, int64_t n) {
  const double q = st->nrm2;
  const bool ok = q > 0.0 && q < 1e300;
  if (!ok && blockIdx.x == 0 && threadIdx.x == 0) st->bad = 1;
  const double f = ok ? 1.0 / sqrt(q) : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) out[i] = f * t[i];
}

// Eigen-decomposition of the (nv + 1) x (nv + 1) leading block of M by cyclic Jacobi in
// one workgroup, parallel ordering (disjoint pairs of each round-robin step rotated at
// once: rows first, then columns), until no off-diagonal element is above 1e-17 of the
// geometric mean of its diagonal pair (or 60 sweeps). Eigenvalues ascending with their
// vectors (columns of evec, leading dimension DAV_MAXV); nv becomes nv + 1.
__global__ __launch_bounds__(256) void k_dav_eig(DevDav* __restrict__ st) {
  __shared__ double A[DAV_MAXV][DAV_MAXV + 1];
  __shared__ double Q[DAV_MAXV][DAV_MAXV + 1];
  __shared__ double cs[DAV_MAXV / 2], sn[DAV_MAXV / 2];
  __shared__ int pp[DAV_MAXV / 2], pq[DAV_MAXV / 2];
  __shared__ int rot;
  __shared__ int order[DAV_MAXV];
  const int tid = threadIdx.x;
  const int n = st->nv + 1;
  const int m = n + (n & 1);  // even number of slots for the round-robin (a dummy index past n)
  for (int e = tid; e < DAV_MAXV * DAV_MAXV; e += 256) {
    const int r = e % DAV_MAXV, c = e / DAV_MAXV;
    A[r][c] = (r < n && c < n) ? st->M[r + DAV_MAXV * c] : 0.0;
    Q[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 60; sweep++) {
    if (tid == 0) rot = 0;
    __syncthreads();
    for (int step = 0; step < m - 1; step++) {
      // round-robin pairing: slot 0 fixed, the others rotate
      if (tid < m / 2) {
        auto slot = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + step) % (m - 1); };
        int p = slot(tid), q = slot(m - 1 - tid);
        if (p > q) { const int w = p; p = q; q = w; }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = A[p][q];
          const double app = A[p][p], aqq = A[q][q];
          if (fabs(apq) > 1e-17 * sqrt(fabs(app * aqq)) && apq != 0.0) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            rot = 1;
          }
        }
        pp[tid] = p; pq[tid] = q; cs[tid] = c; sn[tid] = s;
      }
      __syncthreads();
      // rows p, q of A  <-  J^T A
      for (int e = tid; e < (m / 2) * n; e += 256) {
        const int k = e % n, w = e / n;
        const int p = pp[w], q = pq[w];
        if (q >= n) continue;
        const double c = cs[w], s = sn[w];
        const double ap = A[p][k], aq = A[q][k];
        A[p][k] = c * ap - s * aq;
        A[q][k] = s * ap + c * aq;
      }
      __syncthreads();
      // columns p, q of A and of Q  <-  A J, Q J
      for (int e = tid; e < (m / 2) * n; e += 256) {
        const int k = e % n, w = e / n;
        const int p = pp[w], q = pq[w];
        if (q >= n) continue;
        const double c = cs[w], s = sn[w];
        const double ap = A[k][p], aq = A[k][q];
        A[k][p] = c * ap - s * aq;
        A[k][q] = s * ap + c * aq;
        const double qp = Q[k][p], qq = Q[k][q];
        Q[k][p] = c * qp - s * qq;
        Q[k][q] = s * qp + c * qq;
      }
      __syncthreads();
    }
    if (rot == 0) break;
    __syncthreads();
  }
  if (tid == 0) {  // ascending order
    for (int i = 0; i < n; i++) order[i] = i;
    for (int i = 1; i < n; i++) {
      const int v = order[i];
      int j = i - 1;
      while (j >= 0 && A[order[j]][order[j]] > A[v][v]) { order[j + 1] = order[j]; j--; }
      order[j + 1] = v;
    }
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += 256) {
    const int r = e % n, c = e / n;
    st->evec[r + DAV_MAXV * c] = Q[r][order[c]];
  }
  if (tid < n) st->eig[tid] = A[order[tid]][order[tid]];
  __syncthreads();
  if (tid == 0) {
    st->theta = A[order[0]][order[0]];
    st->nv = n;
  }
}

// u = sum_c evec[c, 0] vec[c], r = sum_c evec[c, 0] hv[c] - theta u, partial sums of r^2
__global__ __launch_bounds__(256) void k_dav_ritz(const DevDav* __restrict__ st, const double* const* __restrict__ vec,
                                                  const double* const* __restrict__ hv, double* __restrict__ u,
                                                  double* __restrict__ r, int64_t n, double* __restrict__ partial) {
  __shared__ double cf[DAV_MAXV];
  __shared__ const double* vp[DAV_MAXV];
  __shared__ const double* hp[DAV_MAXV];
  __shared__ double red[256];
  const int nv = st->nv;
  const double th = st->theta;
  if (threadIdx.x < DAV_MAXV) {
    const bool in = threadIdx.x < nv;
    cf[threadIdx.x] = in ? st->evec[threadIdx.x] : 0.0;
    vp[threadIdx.x] = in ? vec[threadIdx.x] : nullptr;
    hp[threadIdx.x] = in ? hv[threadIdx.x] : nullptr;
  }
  __syncthreads();
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double a = 0.0, b = 0.0;
    for (int c = 0; c < nv; c++) {
      a = fma(cf[c], vp[c][i], a);
      b = fma(cf[c], hp[c][i], b);
    }
    const double ri = b - th * a;
    u[i] = a;
    r[i] = ri;
    s = fma(ri, ri, s);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// st->rn2 = sum of the partials; st->alpha from the two preconditioner dots
__global__ __launch_bounds__(256) void k_dav_scalar(DevDav* __restrict__ st, const double* __restrict__ partial, int parts,
                                                    int what) {
  __shared__ double red[256];
  if (what == 1) {  // Olsen coefficient alpha = -(w . r) / (w . u) (Davidson.cpp:331-340)
    if (threadIdx.x == 0) st->alpha = -st->wr[0] / st->wr[1];
    return;
  }
  double s = 0.0;
  for (int q = threadIdx.x; q < parts; q += 256) s += partial[q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) st->rn2 = red[0];
}

// device wall clock around the Heff product: which 0 starts, 1 adds the elapsed ticks
__global__ void k_dav_clock(DevDav* __restrict__ st, int which) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long t = wall_clock64();
  if (which == 0) st->t0 = t;
  else st->heff_ticks += t - st->t0;
}

// Dot products of array pairs, one workgroup per pair: four strided partial sums per
// thread, then a fixed-order tree over the workgroup.
__global__ __launch_bounds__(256) void k_dots(const DevDot* __restrict__ items, double* __restrict__ out) {
  __shared__ double red[256];
  const DevDot d = items[blockIdx.x];
  const int t = threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t e = t;
  for (; e + 768 < d.n; e += 1024) {
    s0 = fma(d.a[e], d.b[e], s0);
    s1 = fma(d.a[e + 256], d.b[e + 256], s1);
    s2 = fma(d.a[e + 512], d.b[e + 512], s2);
    s3 = fma(d.a[e + 768], d.b[e + 768], s3);
  }
  for (; e < d.n; e += 256) s0 = fma(d.a[e], d.b[e], s0);
  red[t] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = red[0];
}

// x += level * (u - 1/2) with u uniform in [0, 1) from a splitmix64 hash of the index.
__global__ __launch_bounds__(256) void k_add_noise(double* __restrict__ x, int64_t n, uint64_t seed, double level) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)i;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    x[i] += level * ((z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  }
}

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

void launch_tiles(const DevTile* tiles, const DevGemmOp* ops, int ntiles, const double* xbase, double* ybase,
                  hipStream_t s, int mode, double* pbase) {
  if (ntiles <= 0) return;
  if (mode == 2) hipLaunchKernelGGL(k_tiles_fit, dim3(ntiles), dim3(NT), 0, s, tiles, ops, xbase, ybase, pbase);
  else if (mode == 1) hipLaunchKernelGGL(k_tiles_mfma, dim3(ntiles), dim3(NT), 0, s, tiles, ops, xbase, ybase, pbase);
  else hipLaunchKernelGGL(k_tiles, dim3(ntiles), dim3(NT), 0, s, tiles, ops, xbase, ybase, pbase);
}

void launch_tiles_stream1(const DevTile1* tiles, int ntiles, int c, const double* xbase, hipStream_t s) {
  if (ntiles <= 0) return;
  if (c == 0) {  // tiles of at most 32 x 32: one wavefront each
    const int g = std::min((ntiles + 3) / 4, TILE_WAVE_GROUPS);
    hipLaunchKernelGGL((k_tiles_wave1<2, 2>), dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase);
    return;
  }
  if (c == 5 || c == 6) {  // thin tiles of at most 64 x 16 or 16 x 64: one wavefront each
    const int g = std::min((ntiles + 3) / 4, TILE_WAVE_GROUPS);
    if (c == 5) hipLaunchKernelGGL((k_tiles_wave1<4, 1>), dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase);
    else hipLaunchKernelGGL((k_tiles_wave1<1, 4>), dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase);
    return;
  }
  const int g = std::min(ntiles, TILE_STREAM_GROUPS);
  switch (c) {
    case 1: hipLaunchKernelGGL(k_tiles_stream1<1>, dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase); break;
    case 2: hipLaunchKernelGGL(k_tiles_stream1<2>, dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase); break;
    case 3: hipLaunchKernelGGL(k_tiles_stream1<3>, dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase); break;
    default: hipLaunchKernelGGL(k_tiles_stream1<4>, dim3(g), dim3(NT), 0, s, tiles, ntiles, xbase); break;
  }
}

void launch_tiles_wave2(const DevTile* tiles, const DevGemmOp* ops, int ntiles, const double* xbase, double* ybase,
                        hipStream_t s, double* pbase) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_tiles_wave2, dim3((ntiles + 3) / 4), dim3(NT), 0, s, tiles, ntiles, ops, xbase, ybase, pbase);
}

void launch_assemble(const DevPair* pairs, const DevKTerm* terms, int64_t npairs, double* M, int64_t nnz, bool split,
                     hipStream_t s) {
  for (int64_t p0 = 0; p0 < npairs; p0 += (1LL << 30)) {
    const int64_t n = std::min<int64_t>(npairs - p0, 1LL << 30);
    if (split) hipLaunchKernelGGL(k_assemble<true>, dim3((unsigned)n), dim3(256), 0, s, pairs + p0, terms, M, nnz);
    else hipLaunchKernelGGL(k_assemble<false>, dim3((unsigned)n), dim3(256), 0, s, pairs + p0, terms, M, nnz);
  }
}

void launch_bspmv(const DevBItem* items, int nitems, const DevBRed* reds, int nreds, double* P, const double* M,
                  int64_t nnz, int mode, const int* col, const double* x, double* y, hipStream_t s) {
  if (nitems > 0) {
    if (mode == 1) hipLaunchKernelGGL(k_bspmv<1>, dim3(nitems), dim3(256), 0, s, items, M, nnz, col, x, y, P);
    else if (mode == 2) hipLaunchKernelGGL(k_bspmv<2>, dim3(nitems), dim3(256), 0, s, items, M, nnz, col, x, y, P);
    else hipLaunchKernelGGL(k_bspmv<0>, dim3(nitems), dim3(256), 0, s, items, M, nnz, col, x, y, P);
  }
  if (nreds > 0) hipLaunchKernelGGL(k_bspmv_reduce, dim3(nreds), dim3(256), 0, s, reds, P, y);
}

void launch_combos(const DevCombo* d_combos, const DevSrc* d_src, int ncombo, const int64_t* d_chunk_start,
                   int64_t nchunks, hipStream_t s) {
  if (ncombo <= 0 || nchunks <= 0) return;
  hipLaunchKernelGGL(k_combos, dim3((unsigned)nchunks), dim3(256), 0, s, d_combos, d_src, d_chunk_start, ncombo);
}

void launch_diag(const DevDiagBlock* blocks, const DevDiag* terms, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_diag, dim3(nblocks), dim3(256), 0, s, blocks, terms);
}

void launch_chain(const DevCBlock* blocks, int nblocks, const DevCGroup* groups, const DevCTerm* terms, const double* x,
                  double* y, double* part, int P, bool large, hipStream_t s) {
  if (nblocks <= 0) return;
  if (large)
    hipLaunchKernelGGL((k_chain<CHAIN_PER_L, false>), dim3(nblocks), dim3(256), (size_t)4 * P * sizeof(double), s, blocks,
                       groups, terms, x, y, part, P);
  else
    hipLaunchKernelGGL((k_chain<CHAIN_PER_S, true>), dim3(nblocks), dim3(256), (size_t)6 * P * sizeof(double), s, blocks,
                       groups, terms, x, y, part, P);
}

void launch_chain_mfma(const DevCBlock* blocks, const int* offs, const DevCGroup* groups, const DevCTerm* terms,
                       const double* x, double* y, double* part, hipStream_t s) {
  for (int nt = 1; nt <= CHAIN_MFMA_NT; nt++) {
    const int n = offs[nt] - offs[nt - 1];
    if (n <= 0) continue;
    const DevCBlock* bl = blocks + offs[nt - 1];
    const dim3 grid((unsigned)((n + 3) / 4));
    if (nt == 1) hipLaunchKernelGGL(k_chain_mfma<1>, grid, dim3(256), 0, s, bl, n, groups, terms, x, y, part);
    else if (nt == 2) hipLaunchKernelGGL(k_chain_mfma<2>, grid, dim3(256), 0, s, bl, n, groups, terms, x, y, part);
    else hipLaunchKernelGGL(k_chain_mfma<3>, grid, dim3(256), 0, s, bl, n, groups, terms, x, y, part);
  }
}

void launch_ops(const DevOItem* items, int nitems, const DevGemmOp* ops, int P, hipStream_t s) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_ops<OPS_PER>, dim3(nitems), dim3(256), (size_t)2 * P * sizeof(double), s, items, ops, P);
}

void launch_part_reduce(const DevPRed* reds, int nreds, const double* part, hipStream_t s) {
  if (nreds > 0) hipLaunchKernelGGL(k_part_reduce, dim3(nreds), dim3(256), 0, s, reds, part);
}

void launch_chain_reduce(const DevBRed* reds, int nreds, const double* part, double* y, hipStream_t s) {
  if (nreds > 0) hipLaunchKernelGGL(k_chain_reduce, dim3(nreds), dim3(256), 0, s, reds, part, y);
}

bool chain_kernel_setup() {
  return hipFuncSetAttribute((const void*)k_chain<CHAIN_PER_L, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             4 * CHAIN_MAXD * CHAIN_MAXD * (int)sizeof(double)) == hipSuccess;
}

typedef void (*FcKernel)(const DevCBlock*, const DevCGroup*, const DevCTerm*, const double*, double*, double*);

template <bool F32>
struct FcTable {
  static constexpr FcKernel k[FCHAIN_NB] = {
      k_fchain<2, 1, F32>, k_fchain<2, 2, F32>, k_fchain<3, 1, F32>, k_fchain<3, 2, F32>, k_fchain<3, 3, F32>,
      k_fchain<4, 1, F32>, k_fchain<4, 2, F32>, k_fchain<4, 3, F32>, k_fchain<4, 4, F32>, k_fchain<5, 1, F32>,
      k_fchain<5, 2, F32>, k_fchain<5, 3, F32>, k_fchain<5, 4, F32>, k_fchain<5, 5, F32>};
};
template <bool F32>
constexpr FcKernel FcTable<F32>::k[FCHAIN_NB];

static int fc_nt(int bucket) {
  int nt = FCHAIN_NT0;
  while (fchain_bucket(nt + 1, 1) <= bucket) nt++;
  return nt;
}

static size_t fc_lds_bytes(int nt, bool f32) {
  const int e = nt == 2 ? FcShape<2>::ELEMS : nt == 3 ? FcShape<3>::ELEMS : nt == 4 ? FcShape<4>::ELEMS : FcShape<5>::ELEMS;
  return (size_t)e * (f32 ? sizeof(float) : sizeof(double));
}

bool fchain_kernel_setup() {
  static_assert(16 * (FCHAIN_NT0 + 3) == FCHAIN_MAXD && fchain_bucket(5, 5) == FCHAIN_NB - 1, "fused chain buckets");
  bool ok = true;
  for (int bk = 0; bk < FCHAIN_NB; bk++)
    for (int f = 0; f < 2; f++) {
      const FcKernel k = f ? FcTable<true>::k[bk] : FcTable<false>::k[bk];
      ok = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fc_lds_bytes(fc_nt(bk), f)) ==
               hipSuccess && ok;
    }
  return ok;
}

void launch_fchain(const DevCBlock* blocks, int nblocks, const DevCGroup* groups, const DevCTerm* terms, const double* x,
                   double* y, double* part, int bucket, bool fp32, hipStream_t s) {
  if (nblocks <= 0 || bucket < 0 || bucket >= FCHAIN_NB) return;
  const FcKernel k = fp32 ? FcTable<true>::k[bucket] : FcTable<false>::k[bucket];
  hipLaunchKernelGGL(k, dim3(nblocks), dim3(256), fc_lds_bytes(fc_nt(bucket), fp32), s, blocks, groups, terms, x, y, part);
}

int dot_parts(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  return (int)g;
}

void launch_dots(const DevDot* items, int n, double* out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_dots, dim3(n), dim3(256), 0, s, items, out);
}

void launch_dot(const double* a, const double* b, int64_t n, double* partial, double* out, hipStream_t s) {
  const int parts = dot_parts(n);
  hipLaunchKernelGGL(k_dot_partial, dim3(parts), dim3(256), 0, s, a, b, n, partial);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, partial, parts, out);
}

void launch_multi_dot(const double* const* d_vecs, int nv, const double* b, int64_t n, double* partial, double* out,
                      hipStream_t s) {
  if (nv <= 0) return;
  const int parts = dot_parts(n);
  hipLaunchKernelGGL(k_multi_dot_partial, dim3(parts, nv), dim3(256), 0, s, d_vecs, b, n, partial);
  hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(256), 0, s, partial, parts, out);
}

void launch_axpy_dev(const double* coef, double sign, const double* x, double* y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_dev, dim3(grid_for(n)), dim3(256), 0, s, coef, sign, x, y, n);
}

void launch_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, s, a, x, y, n);
}

void launch_add_squares(const double* v, int nv, double* d, int64_t n, hipStream_t s) {
  if (nv > 0 && n > 0) hipLaunchKernelGGL(k_add_squares, dim3(grid_for(n)), dim3(256), 0, s, v, nv, d, n);
}

void launch_scale(double a, double* x, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, s, a, x, n);
}

void launch_lincomb_vecs(const double* const* d_vecs, const double* d_coefs, int nv, double* out, int64_t n,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_lincomb_vecs, dim3(grid_for(n)), dim3(256), 0, s, d_vecs, d_coefs, nv, out, n);
}

void launch_precond(const double* diag, const double* u, double shift, double cutoff, double* work, double* t,
                    int64_t n, int stage, double alpha, hipStream_t s, const float* m) {
  hipLaunchKernelGGL(k_precond, dim3(grid_for(n)), dim3(256), 0, s, diag, u, shift, cutoff, work, t, n, stage, alpha, m,
                     (const double*)nullptr, (const double*)nullptr);
}

void launch_precond_dev(const double* diag, const double* u, const double* dshift, double cutoff, double* work, double* t,
                        int64_t n, int stage, const double* dalpha, hipStream_t s, const float* m) {
  hipLaunchKernelGGL(k_precond, dim3(grid_for(n)), dim3(256), 0, s, diag, u, 0.0, cutoff, work, t, n, stage, 0.0, m, dshift,
                     dalpha);
}

void launch_dav_store(const DevDav* st, double* const* vec, double* const* hv, const double* x, const double* y, int64_t n,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_dav_store, dim3(grid_for(n)), dim3(256), 0, s, st, vec, hv, x, y, n);
}

void launch_dav_dots(DevDav* st, int add, const double* const* a, const double* b, int64_t n, double* partial, int parts,
                     int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_dav_dots_partial, dim3(parts, DAV_MAXV), dim3(256), 0, s, st, add, a, b, n, partial);
  hipLaunchKernelGGL(k_dav_dots_reduce, dim3(DAV_MAXV), dim3(256), 0, s, st, add, partial, parts, mode);
}

void launch_dav_project_out(const DevDav* st, const double* const* vec, double* t, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_dav_project_out, dim3(grid_for(n)), dim3(256), 0, s, st, vec, t, n);
}

void launch_dav_normalize(DevDav* st, const double* t, double* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_dav_normalize, dim3(grid_for(n)), dim3(256), 0, s, st, t, out, n);
}

void launch_dav_eig(DevDav* st, hipStream_t s) { hipLaunchKernelGGL(k_dav_eig, dim3(1), dim3(256), 0, s, st); }

void launch_dav_ritz(DevDav* st, const double* const* vec, const double* const* hv, double* u, double* r, int64_t n,
                     double* partial, int parts, hipStream_t s) {
  hipLaunchKernelGGL(k_dav_ritz, dim3(parts), dim3(256), 0, s, st, vec, hv, u, r, n, partial);
  hipLaunchKernelGGL(k_dav_scalar, dim3(1), dim3(256), 0, s, st, partial, parts, 0);
}

void launch_dav_alpha(DevDav* st, hipStream_t s) {
  hipLaunchKernelGGL(k_dav_scalar, dim3(1), dim3(64), 0, s, st, nullptr, 0, 1);
}

void launch_dav_clock(DevDav* st, int which, hipStream_t s) {
  hipLaunchKernelGGL(k_dav_clock, dim3(1), dim3(64), 0, s, st, which);
}


void launch_fill_hash(double* p, int64_t n, uint64_t seed, double scale, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_hash, dim3(grid_for(n)), dim3(256), 0, s, p, n, seed, scale);
}

void launch_copies(const DevCopy* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_copies, dim3(n), dim3(256), 0, s, items);
}

void launch_svd_jacobi(const DevSvd* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_svd_jacobi, dim3(n), dim3(SVD_THREADS), 0, s, items);
}

void launch_add_noise(double* x, int64_t n, uint64_t seed, double level, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_add_noise, dim3(grid_for(n)), dim3(256), 0, s, x, n, seed, level);
}

void launch_symm_scale(double* x, const int64_t* d_k2i, const double* d_scale, int nk, hipStream_t s) {
  if (nk <= 0) return;
  hipLaunchKernelGGL(k_symm_scale, dim3(nk), dim3(256), 0, s, x, d_k2i, d_scale, nk);
}

}  // namespace chm
