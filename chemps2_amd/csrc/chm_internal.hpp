// Host-side data structures of the chemps2_amd library: sector bookkeeping, the block
// layouts of the reference's operator tensors and two-site object, and the device work
// lists that one two-site effective-Hamiltonian application executes.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace chm {

// Error codes returned through the C ABI.
enum {
  OK = 0,
  ERR_ARG = -1,
  ERR_STATE = -2,
  ERR_MISSING = -3,
  ERR_SIZE = -4,
  ERR_HIP = -5,
  ERR_NUMERIC = -6,
};

struct Error {
  int code;
  std::string msg;
};

// (-1)^(two_p/2) for even two_p (CheMPS2 Special.h:36).
inline int phase(int two_p) { return (((two_p / 2) % 2) != 0) ? -1 : 1; }
double wigner6j(int a, int b, int c, int d, int e, int f);
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);

inline uint64_t key3(int n, int s, int i) { return ((uint64_t)(n & 0xffff)) | ((uint64_t)(s & 0xffff) << 16) | ((uint64_t)(i & 0xff) << 32); }

// Dense (N, 2S, I) -> position lookup of one boundary's sector list; immutable once
// built, shared by the layouts made from it.
struct SecIndex {
  int nmax = -1, smax = -1;
  std::vector<int32_t> pos;
  int find(int N, int S, int I) const {
    if (N < 0 || S < 0 || N > nmax || S > smax || I < 0 || I > 7) return -1;
    return pos[((size_t)N * (smax + 1) + S) * 8 + I];
  }
};

// Current virtual dimensions d(b, N, 2S, I) of every boundary b = 0..L
// (SyBookkeeper::gCurrentDim, SyBookkeeper.cpp:165).
struct Sectors {
  struct Sec { int N, S, I, d; };
  std::vector<std::vector<Sec>> list;                   // [b], sorted by (N, 2S, I), d > 0
  std::vector<std::shared_ptr<const SecIndex>> idx;     // [b]
  uint64_t gen = 0;                                      // incremented by every change
  void resize(int L) {
    list.assign(L + 1, {});
    idx.assign(L + 1, std::make_shared<const SecIndex>());
  }
  void set(int b, std::vector<Sec> v);  // replaces boundary b (entries with d <= 0 dropped)
  int dim(int b, int N, int S, int I) const {
    if (b < 0 || b >= (int)idx.size()) return 0;
    const int p = idx[b]->find(N, S, I);
    return p < 0 ? 0 : list[b][p].d;
  }
};

// Open-addressing hash map from 64-bit keys (never ~0) to int, built once from a list.
struct FlatMap {
  std::vector<uint64_t> keys;
  std::vector<int> vals;
  uint64_t mask = 0;
  static uint64_t mixh(uint64_t k) {
    k ^= k >> 33;
    k *= 0xFF51AFD7ED558CCDULL;
    k ^= k >> 33;
    return k;
  }
  void build(const std::vector<std::pair<uint64_t, int>>& kv) {
    size_t cap = 16;
    while (cap < 2 * kv.size()) cap <<= 1;
    keys.assign(cap, ~0ULL);
    vals.assign(cap, -1);
    mask = cap - 1;
    for (const auto& e : kv) {
      size_t i = mixh(e.first) & mask;
      while (keys[i] != ~0ULL && keys[i] != e.first) i = (i + 1) & mask;
      keys[i] = e.first;
      vals[i] = e.second;
    }
  }
  int get(uint64_t k) const {
    if (keys.empty()) return -1;
    size_t i = mixh(k) & mask;
    while (keys[i] != ~0ULL) {
      if (keys[i] == k) return vals[i];
      i = (i + 1) & mask;
    }
    return -1;
  }
  void clear() { keys.clear(); vals.clear(); mask = 0; }
};

// Where a dense block lives: buffer id (an operator tensor, the combination arena,
// the Heff input vector x) and the element offset inside it.
struct Ref {
  int32_t buf = -1;
  int64_t off = 0;
  bool ok() const { return buf >= 0; }
};

enum : int32_t { BUF_X = 0, BUF_COMBO = 1, BUF_TENSOR0 = 2 };

// Block layout of one operator tensor (TensorOperator.cpp:29-102): blocks
// (n, 2s, i) x (n + n_elec, 2s', i ^ n_irrep) with |s - s'| <= two_j.
struct OpT {
  int boundary = 0, two_j = 0, n_elec = 0, n_irrep = 0;
  int index = 0;  // == boundary (the reference's TensorOperator::gIndex)
  bool moving_right = true, prime_last = true, jw_phase = false;
  int32_t buf = -1;  // buffer id of its device copy
  int64_t size = 0;
  std::shared_ptr<const SecIndex> bi;  // sector index of the boundary it was laid out on
  std::vector<int64_t> offs;           // [up position * (two_j + 1) + (s' - s + two_j) / 2] -> offset or -1
  std::vector<int> Nup, Sup, Iup, Sdn;  // blocks in storage order
  std::vector<int64_t> k2i;
  void build(const Sectors& sec, int b, int two_j, int n_elec, int n_irrep);
  Ref blk(int n1, int s1, int i1, int n2, int s2, int i2) const;
  Ref blkref(int ik) const { Ref r; r.buf = buf; r.off = k2i[ik]; return r; }
  int nk() const { return (int)Nup.size(); }
};

// Block layout of an MPS site tensor (TensorT.cpp:38-104): blocks (NL, 2SL, IL) x
// (NR, 2SR, IR) with NR - NL in {0, 1, 2}, column-major dL x dR.
struct TLay {
  int index = 0;
  int32_t buf = -1;
  int64_t size = 0;
  std::vector<int> NL, SL, IL, NR, SR, IR;
  std::vector<int64_t> k2i;
  FlatMap lut;
  void build(const Sectors& sec, int site, int Is);
  int kappa(int nl, int sl, int il, int nr, int sr, int ir) const;
  Ref blk(int nl, int sl, int il, int nr, int sr, int ir) const;
  int nk() const { return (int)NL.size(); }
};

// All operators of one boundary, indexed like DMRG::allocateTensors
// (DMRGoperators.cpp:909-1145): slot cnt holds boundary cnt + 1.
struct OpSet {
  int dir = 0;       // 1 built moving right, 2 built moving left
  uint64_t gen = 0;  // Sectors::gen when it was laid out
  std::vector<OpT*> L, Q;
  std::vector<std::vector<OpT*>> F0, F1, S0, S1, A, B, C, D;
  OpT* X = nullptr;
  std::vector<std::unique_ptr<OpT>> owned;
};

// Two-site object layout (Sobject.cpp:36-147): blocks (NL, 2SL, IL, N1, N2, 2J, NR,
// 2SR, IR), column-major dL x dR at k2i[kappa].
struct SLayout {
  int index = 0;
  std::vector<int> NL, SL, IL, N1, N2, TJ, NR, SR, IR;
  std::vector<int> dL, dR;  // block dimensions
  std::vector<int64_t> k2i;
  FlatMap lut;
  void build(const Sectors& sec, int index, int I1, int I2);
  int kappa(int nl, int sl, int il, int n1, int n2, int tj, int nr, int sr, int ir) const;
  int nk() const { return (int)NL.size(); }
  int64_t size() const { return k2i.empty() ? 0 : k2i.back(); }
};

// ----------------------------------------------------------- work lists ----
// y_out += alpha * op(Lb) * x_in * op(Rb); an absent Lb or Rb is the identity. With
// both present the reference's association order is kept: (op(L) x) op(R), or
// op(L) (x op(R)) when xr_first is set.
// Ref with a 32-bit offset, the form stored in the (tens of millions of) Heff terms.
struct TRef {
  int32_t buf = -1;
  int32_t off = 0;
  TRef() = default;
  TRef(const Ref& r) : buf(r.buf), off((int32_t)r.off) {}
  operator Ref() const {
    Ref r;
    r.buf = buf;
    r.off = off;
    return r;
  }
  bool ok() const { return buf >= 0; }
};
struct Term {  // 40 bytes
  int out, in;
  double alpha;
  TRef Lb, Rb;
  char tl, tr;
  bool xr_first;
};

// out = sum_q coef_q * src_q over equally shaped blocks (x-independent, built once per
// site before the Davidson iterations).
struct Combo {
  int64_t out_off, size;
  std::vector<std::pair<Ref, double>> src;
};

// Diagonal contribution to block kappa of diag(Heff): c + a * diag(L) (x) 1 + b * 1 (x)
// diag(R) + w * diag(L) (x) diag(R).
struct DiagTerm {
  int k;
  int kind;  // 0 constant, 1 left, 2 right, 3 outer
  double coef;
  Ref Lb, Rb;
};

// ---------------------------------------------------- device-side records --
// Pointers flagged relative hold an element offset into the Heff input vector x
// (operands) or output vector y (tile outputs), resolved by the kernel.
struct DevGemmOp {  // one product accumulated into a tile
  const double* A;
  const double* B;
  double alpha;
  int lda, ldb, k;
  int ta, tb;   // 0 'N', 1 'T'
  int kind;     // 0 gemm, 1 axpy: C += alpha * A (A read with lda)
  int flags;    // bit 0: A relative to x, bit 1: B relative to x
  int pad;
};
struct DevTile {  // one 64 x 64 output tile and the products that accumulate into it
  double* C;
  int ldc, m0, n0, m, n;  // m, n: full extents of C
  int op0, nop;
  int flags;  // bit 0: C relative to y; bit 1: add to C instead of overwriting it; TILE_NO_AXPY
};
struct DevTile1 {  // a tile of one GEMM product (k > 0, no axpy) with the product itself
  double* C;
  const double* A;
  const double* B;
  double alpha;
  int ldc, m0, n0, mrem, nrem;  // mrem, nrem: rows and columns of the tile that hold elements
  int lda, ldb, k, ta, tb;
  int flags;  // bit 0: A relative to x, bit 1: B relative to x
  int pad;
};
constexpr int TILE_NO_AXPY = 4;  // no product of the tile is an axpy (kind 1)
constexpr int TILE_C_PART = 8;   // C relative to the partial-sum buffer
// Small-block products of the operator updates: one workgroup per work item computes
// out (m x n, contiguous) = sum of its products (DevGemmOp, absolute pointers, kind 0:
// alpha op(A) op(B); kind 1: alpha op(A)) in 2 x 2 register tiles with the operands
// copied into LDS; an output block with many products is split over several items
// whose partials DevPRed sums (overwriting the output).
constexpr int OPS_MAXD = 64;    // m, n of the small path
constexpr int OPS_P = 2048;     // operand elements (m k, k n) of the small path
constexpr int OPS_ITEM = 64;    // products per work item
struct DevOItem {
  double* out;
  int m, n;
  int op0, nop;
};
struct DevPRed {  // out[e] = sum_q part[p0 + q R + e], q < nparts
  double* out;
  int64_t p0;
  int R, nparts;
};
// Output-stationary chained products for output blocks of at most CHAIN_MAXD rows and
// columns (and inputs of the same bound): one workgroup per block keeps y_kappa in
// registers and runs every term of the block; terms sharing an input block kappa'
// form a group that stages x_kappa' in LDS once.
constexpr int CHAIN_MAXD = 64;  // larger blocks run as GEMM tiles
constexpr int CHAIN_SMALLP = 1296;  // block products up to 36^2: 6 LDS arrays (62 KB), double-buffered
constexpr int CHAIN_GROUP_TERMS = 32;  // terms per input group record
constexpr int CHAIN_ITEM_TERMS = 64;   // terms per work item (load balance across workgroups)
struct DevCBlock {  // one work item: input groups [g0, g0 + ng) of one output block
  int64_t out;        // y offset (direct) or partial offset
  int dL, dR;
  int g0, ng;
  int direct;
  int tr;             // k_fchain: the item computes the transposed block (terms already exchanged)
};
struct DevCGroup {
  int64_t xoff;
  int dLi, dRi;
  int t0, nt;
};
struct DevCTerm {  // mode 0: y += a x; 1: a op(A) x; 2: a x op(B); 3: a (op(A) x) op(B); 4: a op(A) (x op(B))
  const double* A;
  const double* B;
  double alpha;
  int lda, ldb;
  int ta, tb;
  int mode;
  int pad;
};
// Explicit block-sparse Heff: output block kappa holds a dense R x C column-major
// matrix (R = |kappa|) whose column c multiplies x[col[c]].
struct DevRowBlock {
  int64_t moff;  // start of the R x C matrix in the value array
  int64_t coff;  // start of its column indices
  int64_t yoff;  // k2i[kappa]
  int R, C;
};
// Work item of the explicit product: rows [r0, r0 + R) of block kappa against
// columns [c0, c1) of its matrix; the partial result goes to y (single item per
// block) or to P[poff + row] for the reduction pass.
struct DevBItem {
  int64_t moff;  // element (r0, 0) of the block matrix (leading dimension ld)
  int64_t coff;  // column indices of the block
  int64_t out;   // y offset (direct) or partial offset
  int ld, R, c0, c1;
  int direct;
  int pad;
};
struct DevBRed {  // y[yoff + r] = sum_q P[p0 + q * R + r], q < nparts
  int64_t yoff, p0;
  int R, nparts;
};
// One (kappa, kappa') pair: its R x |kappa'| sub-block is sum_t alpha_t op(A_t) (x) op(B_t).
struct DevPair {
  int64_t moff;  // first element of the sub-block (leading dimension R)
  int R, dL, dR, dLi, dRi;
  int t0, nt;
  int pad;
};
struct DevKTerm {
  const double* A;  // op(A) is dL x dLi; null: identity
  const double* B;  // op(B) is dRi x dR; null: identity
  double alpha;
  int lda, ldb, ta, tb;
};
struct DevCombo {
  double* out;
  int64_t size;
  int src0, nsrc;
};
struct DevSrc {
  const double* p;
  double c;
};
struct DevDiag {
  int kind;
  int pad;
  double coef;
  const double* L;
  const double* R;
};
struct DevDiagBlock {
  double* out;
  int dL, dR;
  int t0, nt;
};

// dst (rows x cols, leading dimension ldd) = sum_q coef_q * src_q (rows x cols,
// leading dimension lds) over at most two sources; used to assemble the centre-sector
// matrices of Sobject::Split and to scatter the SVD factors back into MPS blocks.
struct DevCopy {
  double* dst;
  const double* src0;
  const double* src1;
  const double* rowscale;  // optional per-row factor on the sources (null: 1)
  const double* colscale;  // optional per-column factor (null: 1)
  double c0, c1;
  int rows, cols, ldd, lds;
  int keep_rows, keep_cols;  // elements outside [0, keep) are written as zero
  int trans;                 // write the transpose: dst[c + ldd * r]
  int pad;
};

// Thin SVD M = U diag(S) VT of one centre sector of Split by one-sided Jacobi. The
// working matrix A (m x n, m >= n, column-major) is M, or M^T when M is wide; it is
// overwritten. Q (n x n) is workspace. U is dLt x n and VT is n x dRt (leading
// dimensions dLt and n), S descending.
struct DevSvd {
  double* A;
  double* Q;
  double* U;
  double* VT;
  double* S;
  int* info;  // set to 1 when the sweeps end before every pair is orthogonal
  int m, n;
  int wide;
  int max_sweeps;  // at most SVD_MAX_SWEEPS
};
constexpr int SVD_JACOBI_MAXN = 256;  // largest n of the batched Jacobi kernel

// Launchers (kernels.hip).
// One dot product of two device arrays (the 2-RDM's P tensor with an operator).
struct DevDot {
  const double* a;
  const double* b;
  int64_t n;
};

void launch_copies(const DevCopy* items, int n, hipStream_t s);
// out[q] = a_q . b_q for every item q (one workgroup per item, fixed summation order)
void launch_dots(const DevDot* items, int n, double* out, hipStream_t s);
void launch_svd_jacobi(const DevSvd* items, int n, hipStream_t s);
void launch_add_noise(double* x, int64_t n, uint64_t seed, double level, hipStream_t s);
// mode 2: k_tiles_fit (matrix cores, only the 16 x 16 sub-tiles and k-steps that hold
// elements), 1: k_tiles_mfma (matrix cores, whole 64 x 64 tiles), 0: k_tiles (vector FMA)
void launch_tiles(const DevTile* tiles, const DevGemmOp* ops, int ntiles, const double* xbase, double* ybase,
                  hipStream_t s, int mode = 2, double* pbase = nullptr);
// Tiles of one GEMM product each, overwriting their outputs, all of one bucket
// (tile_bucket): c = 0, at most 32 x 32, and c = 5 / 6, at most 64 x 16 / 16 x 64, one
// wavefront per tile (k_tiles_wave1); c = 1..4, c sub-tiles per wavefront of a workgroup
// per tile (k_tiles_stream1). The workgroups (TILE_STREAM_GROUPS, or TILE_WAVE_GROUPS of
// four wavefronts for the wavefront buckets: what the device holds at once) walk the list.
// Grouped tiles of at most 32 x 32 whose products are all GEMMs with k > 0: one
// wavefront per tile (k_tiles_wave2).
void launch_tiles_wave2(const DevTile* tiles, const DevGemmOp* ops, int ntiles, const double* xbase, double* ybase,
                        hipStream_t s, double* pbase = nullptr);
constexpr int TILE_STREAM_GROUPS = 1024;
constexpr int TILE_WAVE_GROUPS = 768;
constexpr int TILE_BUCKETS = 7;
inline int tile_bucket(int mrem, int nrem) {
  if (mrem <= 32 && nrem <= 32) return 0;
  if (nrem <= 16) return 5;
  if (mrem <= 16) return 6;
  return ((mrem + 15) / 16 * ((nrem + 15) / 16) + 3) / 4;
}
void launch_tiles_stream1(const DevTile1* tiles, int ntiles, int c, const double* xbase, hipStream_t s);
void launch_combos(const DevCombo* d_combos, const DevSrc* d_src, int ncombo, const int64_t* d_chunk_start,
                   int64_t nchunks, hipStream_t s);
void launch_diag(const DevDiagBlock* blocks, const DevDiag* terms, int nblocks, hipStream_t s);
// small: block products <= CHAIN_SMALLP (62 KB of LDS); large: up to CHAIN_MAXD^2 with
// 4 single-buffered LDS arrays (up to 128 KB, allowed by chain_kernel_setup)
void launch_chain(const DevCBlock* blocks, int nblocks, const DevCGroup* groups, const DevCTerm* terms, const double* x,
                  double* y, double* part, int P, bool large, hipStream_t s);
void launch_chain_reduce(const DevBRed* reds, int nreds, const double* part, double* y, hipStream_t s);
// Chained products on the matrix cores: work items sorted by tile count nt = 1..3
// (every dimension of the item's blocks at most 16 nt); items [offs[nt-1], offs[nt]).
// offs is a host array. Larger blocks run in k_chain.
constexpr int CHAIN_MFMA_NT = 3;
// k_fchain: chained products of blocks up to FCHAIN_MAXD on the matrix cores, one
// workgroup of four wavefronts per work item, each wavefront owning 16-row strips of the
// output block (as computed: transposed for items with tr set). Items are bucketed by
// nt = max(FCHAIN_NT0, ceil(largest dimension / 16)) (a bucket's LDS holds op(A), op(B)
// and x padded to 16 nt) and nj = ceil(columns of the block as computed / 16) <= nt.
constexpr int FCHAIN_NT0 = 2;
constexpr int FCHAIN_MAXD = 80;
constexpr int fchain_bucket(int nt, int nj) { return (nt - 2) * (nt + 1) / 2 + nj - 1; }
constexpr int FCHAIN_NB = 14;  // (nt, nj) = (2, 1) .. (5, 5)
// grants k_fchain the LDS of its buckets (false: the device refuses; the blocks then run
// in the other chained kernels and as GEMM tiles)
bool fchain_kernel_setup();
// bucket: fchain_bucket(nt, nj); fp32: the FP32 variant (v_mfma_f32_16x16x4f32, FP32
// operands and sums, added to y in FP64), for the low-precision phase of the mixed form
void launch_fchain(const DevCBlock* blocks, int nblocks, const DevCGroup* groups, const DevCTerm* terms, const double* x,
                   double* y, double* part, int bucket, bool fp32, hipStream_t s);
void launch_chain_mfma(const DevCBlock* blocks, const int* offs, const DevCGroup* groups, const DevCTerm* terms,
                       const double* x, double* y, double* part, hipStream_t s);
void launch_ops(const DevOItem* items, int nitems, const DevGemmOp* ops, int P, hipStream_t s);
void launch_part_reduce(const DevPRed* reds, int nreds, const double* part, hipStream_t s);
bool chain_kernel_setup();  // false: the large variant cannot get its LDS on this device
// split: the matrix is stored as hi/lo float pairs (see k_assemble); mode 0 double
// matrix, 1 hi floats only, 2 hi + lo
void launch_assemble(const DevPair* pairs, const DevKTerm* terms, int64_t npairs, double* M, int64_t nnz, bool split,
                     hipStream_t s);
void launch_bspmv(const DevBItem* items, int nitems, const DevBRed* reds, int nreds, double* P, const double* M,
                  int64_t nnz, int mode, const int* col, const double* x, double* y, hipStream_t s);
int dot_parts(int64_t n);
void launch_dot(const double* a, const double* b, int64_t n, double* partial, double* out, hipStream_t s);
void launch_multi_dot(const double* const* d_vecs, int nv, const double* b, int64_t n, double* partial, double* out,
                      hipStream_t s);
void launch_axpy_dev(const double* coef, double sign, const double* x, double* y, int64_t n, hipStream_t s);
void launch_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s);
void launch_add_squares(const double* v, int nv, double* d, int64_t n, hipStream_t s);
void launch_scale(double a, double* x, int64_t n, hipStream_t s);
void launch_lincomb_vecs(const double* const* d_vecs, const double* d_coefs, int nv, double* out, int64_t n,
                         hipStream_t s);
// Davidson.cpp:322-350: stage 0 work = g u, stage 1 t = -g (t + alpha u), with
// g = 1 / max(|d - shift|, cutoff) (signed) times 1 + m[i], clamped to [0.1, 10], when m
// is given.
// Scalars and small matrices of one Davidson solve on the device (Davidson.cpp:105-191,
// 214-350): basis size, Ritz value, residual norm, Olsen coefficient, V^T H V, its
// eigenvectors, projections of a new vector, and a wall-clock sum around the products.
constexpr int DAV_MAXV = 32;
struct DevDav {
  int nv;    // basis vectors held
  int bad;   // a new vector had a zero or invalid norm
  double theta, rn2, alpha, nrm2;
  double wr[2];
  unsigned long long t0, heff_ticks;
  double M[DAV_MAXV * DAV_MAXV];     // column-major, leading dimension DAV_MAXV
  double evec[DAV_MAXV * DAV_MAXV];  // eigenvectors of the leading nv x nv block, ascending
  double eig[DAV_MAXV];
  double proj[DAV_MAXV];
};
void launch_dav_store(const DevDav* st, double* const* vec, double* const* hv, const double* x, const double* y, int64_t n,
                      hipStream_t s);
// dots a[c] . b for c < nv + add; mode 0 writes the new column of M, mode 1 the projections
void launch_dav_dots(DevDav* st, int add, const double* const* a, const double* b, int64_t n, double* partial, int parts,
                     int mode, hipStream_t s);
void launch_dav_project_out(const DevDav* st, const double* const* vec, double* t, int64_t n, hipStream_t s);
void launch_dav_normalize(DevDav* st, const double* t, double* out, int64_t n, hipStream_t s);
void launch_dav_eig(DevDav* st, hipStream_t s);
void launch_dav_ritz(DevDav* st, const double* const* vec, const double* const* hv, double* u, double* r, int64_t n,
                     double* partial, int parts, hipStream_t s);
void launch_dav_alpha(DevDav* st, hipStream_t s);
void launch_dav_clock(DevDav* st, int which, hipStream_t s);
// the Davidson correction with the shift and the coefficient read from device memory
void launch_precond_dev(const double* diag, const double* u, const double* dshift, double cutoff, double* work, double* t,
                        int64_t n, int stage, const double* dalpha, hipStream_t s, const float* m);
void launch_precond(const double* diag, const double* u, double shift, double cutoff, double* work, double* t,
                    int64_t n, int stage, double alpha, hipStream_t s, const float* m = nullptr);
void launch_symm_scale(double* x, const int64_t* d_k2i, const double* d_scale, int nk, hipStream_t s);

}  // namespace chm
