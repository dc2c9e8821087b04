// Work-list generator for one two-site effective Hamiltonian (Heff.cpp:43-315).
#pragma once
#include "chm_internal.hpp"

namespace chm {

struct HeffPlanner {
  static constexpr int64_t ZERO = -7;  // Ref.off of an identically zero combination

  // Read-only views with the names the diagram code uses.
  struct BKView {
    const Sectors* s = nullptr;
    const int* irr = nullptr;
    int nirrep = 1;
    int dim(int b, int N, int S, int I) const { return s->dim(b, N, S, I); }
    int irrep(int site) const { return irr[site]; }
  };
  struct PView {
    const double* mx = nullptr;
    int L = 0;
    double g(int a, int b, int c, int d) const { return mx[a + (size_t)L * (b + (size_t)L * (c + (size_t)L * d))]; }
  };

  const Sectors* sec = nullptr;
  const SLayout* S = nullptr;
  BKView bkv;
  PView pv;
  const BKView* bk = &bkv;
  const PView* P = &pv;
  int index = 0, L = 0;
  bool atLeft = false, atRight = false;
  OpSet* Lo = nullptr;
  OpSet* Ro = nullptr;

  // Outputs. Chunk c holds the terms of output blocks [c * CHUNK, (c + 1) * CHUNK),
  // sorted by (out, in).
  static constexpr int CHUNK = 8;
  std::vector<std::vector<Term>> chunks;
  std::vector<std::vector<int64_t>> remap;  // chunk-local combination index -> arena offset
  int64_t nterms = 0;
  // A term's operator block with its combination reference mapped to the arena.
  Ref ref(const Term& t, Ref r) const {
    if (r.buf == BUF_COMBO) r.off = remap[t.out / CHUNK][r.off];
    return r;
  }
  std::vector<Term> terms;  // scratch of one chunk while it is generated
  std::vector<std::vector<Term>> pool;  // buffers to reuse for the chunks (optional)
  struct Part {  // one chunk's output before the merge
    std::vector<Term> terms;
    std::vector<Combo> combos;
    std::vector<uint64_t> hash;
    std::vector<DiagTerm> diag;
    double flops = 0.0;
  };
  std::vector<Part> leftover;  // released with the planner (duplicate combinations)
  std::vector<Combo> combos;
  int64_t combo_size = 0;
  std::vector<DiagTerm> diag;
  double flops = 0.0;  // F_H: 2mnk over every dgemm_ the reference issues at this site

  int nthreads = 0;  // 0: host_threads()
  double phase_s[3] = {0, 0, 0};  // seconds: generation, combination merge, rest of build()

  void init(const Sectors* sec, const std::vector<int>* irreps, int nirrep, const double* mx, int L, const SLayout* S,
            OpSet* Lo, OpSet* Ro);
  // Terms grouped by output block and sorted by (out, in), combinations, diagonal
  // terms; output blocks are processed in parallel chunks and merged in order, so the
  // result does not depend on the thread count. Throws Error.
  void build();

  // --- internals ----------------------------------------------------------
  Ref w2;
  // While a chunk is generated, combination Refs hold the chunk-local combination
  // index; the merge replaces it by the arena offset.
  std::unordered_map<uint64_t, int> combo_cache;  // content hash -> local combination
  std::vector<uint64_t> combo_hash;
  int dl(int ik) const { return S->dL[ik]; }
  int dr(int ik) const { return S->dR[ik]; }
  void diag_block(int ik);
  void term(int ik, int kin, double alpha, char tl, Ref Lb, char tr, Ref Rb, bool xr_first = false);
  Ref make_combo(std::vector<std::pair<Ref, double>>&& src, int64_t size);
  Ref lin_comb(const std::vector<OpT*>& ops, const std::vector<double>& co, int n1, int s1, int i1, int n2, int s2, int i2,
               int sz);
  Ref q_plus_l(OpT* Q, std::vector<OpT*>& Ls, int site, bool left, int n1, int s1, int i1, int n2, int s2, int i2);
  static OpT* need(OpT* t) {
    if (!t) throw Error{ERR_MISSING, "operator tensor not uploaded"};
    return t;
  }
  static OpT* at(std::vector<std::vector<OpT*>>& v, int k, int m) {
    if (k < 0 || k >= (int)v.size() || m < 0 || m >= (int)v[k].size()) throw Error{ERR_MISSING, "operator tensor not uploaded"};
    return need(v[k][m]);
  }
  // Pair operators indexed by orbital pairs (a <= b): left block pair, left boundary
  // pair right of the sites, right block pair, right boundary pair left of the sites
  // (storage of DMRGoperators.cpp:909-1145).
  OpT* pl(std::vector<std::vector<OpT*>>& v, int a, int b) { return at(v, b - a, index - 1 - b); }
  OpT* cl(std::vector<std::vector<OpT*>>& v, int a, int b) { return at(v, b - a, a - index); }
  OpT* pr(std::vector<std::vector<OpT*>>& v, int a, int b) { return at(v, b - a, a - index - 2); }
  OpT* cr(std::vector<std::vector<OpT*>>& v, int a, int b) { return at(v, b - a, index + 1 - b); }

  void d1(int ik);
  void d2a(int ik);
  void d2loc(int ik);
  void d3(int ik);
  void d4(int ik);
  void d4BC(int ik);
  void d4E(int ik);
  void d4I(int ik);
  void d4F(int ik);
  void d4G(int ik);
  void d4J(int ik);
  void d4KL(int ik);
  void d5(int ik);
};

}  // namespace chm
