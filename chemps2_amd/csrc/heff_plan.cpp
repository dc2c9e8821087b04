// Work-list generator for the two-site effective Hamiltonian of CheMPS2 (Heff.cpp:43-315,
// HeffDiagrams1-5.cpp, HeffDiagonal.cpp). Every diagram of the reference becomes one or
// more chain terms y_k += alpha op(L) x_k' op(R) executed on the GPU; the x-independent
// operator sums the reference rebuilds inside every matrix-vector product (Q + sum V L,
// sum_l V L_l, sum_b V L_b) become combinations computed once per site.
#include <algorithm>
#include <array>
#include <cassert>
#include <cstring>

#include <chrono>

#include "heff_plan.hpp"
#include "parallel.hpp"

namespace chm {

static const double SQ2 = std::sqrt(2.0);
static const double SQH = std::sqrt(0.5);

void HeffPlanner::init(const Sectors* sec_, const std::vector<int>* irreps, int nirrep, const double* mx, int L_,
                       const SLayout* S_, OpSet* Lo_, OpSet* Ro_) {
  sec = sec_;
  S = S_;
  L = L_;
  index = S->index;
  bkv = {sec_, irreps->data(), nirrep};
  pv = {mx, L_};
  atLeft = (index == 0);
  atRight = (index == L - 2);
  Lo = Lo_;
  Ro = Ro_;
}

void HeffPlanner::term(int ik, int kin, double alpha, char tl, Ref Lb, char tr, Ref Rb, bool xr_first) {
  const double dL = dl(ik), dR = dr(ik), dLi = dl(kin), dRi = dr(kin);
  if (tl != 'I' && !Lb.ok() && Lb.off != ZERO) throw Error{ERR_MISSING, "missing left operator block"};
  if (tr != 'I' && !Rb.ok() && Rb.off != ZERO) throw Error{ERR_MISSING, "missing right operator block"};
  // The reference's GEMM count (2mnk per dgemm_ call, its association order).
  if (tl != 'I' && tr != 'I') flops += xr_first ? 2.0 * dLi * dR * dRi + 2.0 * dL * dR * dLi : 2.0 * dL * dRi * dLi + 2.0 * dL * dR * dRi;
  else if (tl != 'I') flops += 2.0 * dL * dR * dLi;
  else if (tr != 'I') flops += 2.0 * dL * dR * dRi;
  if ((tl != 'I' && !Lb.ok()) || (tr != 'I' && !Rb.ok())) return;  // identically zero combination
  if (alpha == 0.0) return;
  if (Lb.off > INT32_MAX || Rb.off > INT32_MAX) throw Error{ERR_SIZE, "operator block offset beyond 32 bits"};
  Term t;
  t.out = ik; t.in = kin; t.alpha = alpha; t.Lb = Lb; t.Rb = Rb; t.tl = tl; t.tr = tr; t.xr_first = xr_first;
  terms.push_back(t);
}

static inline uint64_t hmix(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
  return h * 0xFF51AFD7ED558CCDULL;
}

static uint64_t combo_key(const std::vector<std::pair<Ref, double>>& src, int64_t size) {
  uint64_t h = hmix(0x1234567ULL, (uint64_t)size);
  for (auto& s : src) {
    uint64_t c;
    std::memcpy(&c, &s.second, sizeof(c));
    h = hmix(hmix(hmix(h, (uint64_t)(uint32_t)s.first.buf), (uint64_t)s.first.off), c);
  }
  return h;
}

static bool same_combo(const Combo& c, const std::vector<std::pair<Ref, double>>& src, int64_t size) {
  if (c.size != size || c.src.size() != src.size()) return false;
  for (size_t q = 0; q < src.size(); q++)
    if (c.src[q].first.buf != src[q].first.buf || c.src[q].first.off != src[q].first.off || c.src[q].second != src[q].second)
      return false;
  return true;
}

Ref HeffPlanner::make_combo(std::vector<std::pair<Ref, double>>&& src, int64_t size) {
  if (src.empty()) { Ref z; z.off = ZERO; return z; }
  const uint64_t h = combo_key(src, size);
  auto it = combo_cache.find(h);
  if (it != combo_cache.end() && same_combo(combos[it->second], src, size)) return Ref{BUF_COMBO, it->second};
  Combo c;
  c.out_off = -1;  // set by the merge
  c.size = size;
  c.src = std::move(src);
  combos.push_back(std::move(c));
  combo_hash.push_back(h);
  if (it == combo_cache.end()) combo_cache.emplace(h, (int)combos.size() - 1);
  return Ref{BUF_COMBO, (int64_t)combos.size() - 1};
}

Ref HeffPlanner::lin_comb(const std::vector<OpT*>& ops, const std::vector<double>& co, int n1, int s1, int i1, int n2,
                          int s2, int i2, int sz) {
  std::vector<std::pair<Ref, double>> src;
  for (size_t q = 0; q < ops.size(); q++) {
    if (co[q] == 0.0) continue;
    Ref b = ops[q]->blk(n1, s1, i1, n2, s2, i2);
    if (!b.ok()) throw Error{ERR_MISSING, "missing block in operator combination"};
    src.push_back({b, co[q]});
  }
  return make_combo(std::move(src), sz);
}

// Q block plus sum_l V L_l over the orbitals l on the same side as the boundary whose
// irrep equals that of `site` (HeffDiagrams3.cpp:60-75 and the sibling diagrams).
Ref HeffPlanner::q_plus_l(OpT* Q, std::vector<OpT*>& Ls, int site, bool left, int n1, int s1, int i1, int n2, int s2,
                          int i2) {
  std::vector<std::pair<Ref, double>> src;
  const int b = left ? index : index + 2;
  const int sz = bk->dim(b, n1, s1, i1) * bk->dim(b, n2, s2, i2);
  Ref q = Q->blk(n1, s1, i1, n2, s2, i2);
  if (!q.ok()) throw Error{ERR_MISSING, "missing Q block"};
  src.push_back({q, 1.0});
  if (left) {
    for (int l = 0; l < index; l++) {
      if (bk->irrep(l) != bk->irrep(site)) continue;
      const double a = P->g(l, site, site, site);
      if (a == 0.0) continue;
      src.push_back({Ls[index - 1 - l]->blk(n1, s1, i1, n2, s2, i2), a});
    }
  } else {
    for (int l = index + 2; l < L; l++) {
      if (bk->irrep(l) != bk->irrep(site)) continue;
      const double a = P->g(site, site, site, l);
      if (a == 0.0) continue;
      src.push_back({Ls[l - index - 2]->blk(n1, s1, i1, n2, s2, i2), a});
    }
  }
  for (auto& s : src)
    if (!s.first.ok()) throw Error{ERR_MISSING, "missing L block"};
  return make_combo(std::move(src), sz);
}

static double wall() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

void HeffPlanner::build() {  // Heff.cpp:43-248 (the loop over output blocks)
  const double w0 = wall();
  const int nk = S->nk();
  const int CH = CHUNK;
  const int64_t nch = (nk + CH - 1) / CH;
  std::vector<Part> parts(nch);
  const int nt = (int)std::min<int64_t>(nthreads > 0 ? nthreads : host_threads(), std::max<int64_t>(nch, 1));
  std::vector<std::unique_ptr<HeffPlanner>> workers(nt);
  for (auto& w : workers) {
    w = std::make_unique<HeffPlanner>();
    w->sec = sec; w->S = S; w->bkv = bkv; w->pv = pv; w->index = index; w->L = L;
    w->atLeft = atLeft; w->atRight = atRight; w->Lo = Lo; w->Ro = Ro;
  }
  parallel_for(nch, nt, [&](int64_t c, int tid) {
    HeffPlanner& w = *workers[tid];
    if (c < (int64_t)pool.size()) w.terms = std::move(pool[c]);
    w.terms.clear(); w.combos.clear(); w.combo_hash.clear(); w.combo_cache.clear(); w.diag.clear();
    w.flops = 0.0;
    const int k1 = (int)std::min<int64_t>(nk, (c + 1) * CH);
    for (int ik = (int)(c * CH); ik < k1; ik++) {
      const size_t t0 = w.terms.size();
      w.d1(ik);
      w.d2loc(ik);
      if (!atLeft && !atRight) w.d2a(ik);
      w.d3(ik);
      w.d4(ik);
      if (!atLeft && !atRight) w.d5(ik);
      w.diag_block(ik);
      std::stable_sort(w.terms.begin() + t0, w.terms.end(), [](const Term& a, const Term& b) { return a.in < b.in; });
    }
    Part& p = parts[c];
    p.terms = std::move(w.terms);
    p.combos = std::move(w.combos);
    p.hash = std::move(w.combo_hash);
    p.diag = std::move(w.diag);
    p.flops = w.flops;
    w.terms = {}; w.combos = {}; w.combo_hash = {}; w.diag = {};
  });
  // merge: global deduplication of the combinations in NSH shards (by hash bits), each
  // shard walking the chunks in order; arena offsets follow shard order, then first
  // appearance, so the result does not depend on the thread count. The terms keep
  // chunk-local combination indices; ref() maps them through remap.
  const double w1 = wall();
  terms.clear(); combos.clear(); combo_hash.clear(); combo_cache.clear(); diag.clear();
  combo_size = 0;
  flops = 0.0;
  constexpr int NSH = 64;
  auto shard_of = [](uint64_t h) { return (int)(h >> 58); };
  std::vector<std::vector<int>> by_shard(nch);        // local indices sorted by shard
  std::vector<std::array<int, NSH + 1>> sh_start(nch);
  parallel_for(nch, nt, [&](int64_t c, int) {
    const Part& p = parts[c];
    auto& st = sh_start[c];
    st.fill(0);
    for (uint64_t h : p.hash) st[shard_of(h) + 1]++;
    for (int q = 0; q < NSH; q++) st[q + 1] += st[q];
    std::array<int, NSH> fill;
    for (int q = 0; q < NSH; q++) fill[q] = st[q];
    by_shard[c].resize(p.hash.size());
    for (size_t q = 0; q < p.hash.size(); q++) by_shard[c][fill[shard_of(p.hash[q])]++] = (int)q;
  });
  struct Shard {
    std::vector<std::pair<int64_t, int>> reps;  // (chunk, local index) of every distinct combination
    std::vector<int64_t> size;                   // arena offset within the shard after the pass
    int64_t total = 0;
  };
  std::vector<Shard> shards(NSH);
  std::vector<std::vector<int64_t>> local(nch);  // local index -> (shard << 40) | index within shard
  for (int64_t c = 0; c < nch; c++) local[c].resize(parts[c].hash.size());
  parallel_for(NSH, nt, [&](int64_t s, int) {
    Shard& sh = shards[s];
    // open-addressing table hash -> first representative with that hash
    int64_t cand = 0;
    for (int64_t c = 0; c < nch; c++) cand += sh_start[c][s + 1] - sh_start[c][s];
    size_t cap = 16;
    while (cap < (size_t)cand * 2) cap <<= 1;
    std::vector<uint64_t> keys(cap);
    std::vector<int> vals(cap, -1);
    const size_t mask = cap - 1;
    for (int64_t c = 0; c < nch; c++) {
      const Part& p = parts[c];
      for (int e = sh_start[c][s]; e < sh_start[c][s + 1]; e++) {
        const int q = by_shard[c][e];
        const Combo& cb = p.combos[q];
        const uint64_t h = p.hash[q];
        size_t slot = (size_t)(h * 0x9E3779B97F4A7C15ULL >> 20) & mask;
        while (vals[slot] >= 0 && keys[slot] != h) slot = (slot + 1) & mask;
        if (vals[slot] >= 0) {
          const auto& r = sh.reps[vals[slot]];
          if (same_combo(parts[r.first].combos[r.second], cb.src, cb.size)) {
            local[c][q] = ((int64_t)s << 40) | vals[slot];
            continue;
          }
        }
        const int id = (int)sh.reps.size();
        sh.reps.push_back({c, q});
        if (vals[slot] < 0) { keys[slot] = h; vals[slot] = id; }
        local[c][q] = ((int64_t)s << 40) | id;
      }
    }
    sh.size.resize(sh.reps.size());
    for (size_t i = 0; i < sh.reps.size(); i++) {
      sh.size[i] = sh.total;
      sh.total += (parts[sh.reps[i].first].combos[sh.reps[i].second].size + 7) / 8 * 8;  // 64-byte aligned
    }
  });
  std::vector<int64_t> sbase(NSH + 1, 0);
  for (int s = 0; s < NSH; s++) sbase[s + 1] = sbase[s] + shards[s].total;
  combo_size = sbase[NSH];
  std::vector<int64_t> cbase(NSH + 1, 0);
  for (int s = 0; s < NSH; s++) cbase[s + 1] = cbase[s] + (int64_t)shards[s].reps.size();
  combos.resize(cbase[NSH]);
  parallel_for(NSH, nt, [&](int64_t s, int) {
    for (size_t i = 0; i < shards[s].reps.size(); i++) {
      Combo& cb = parts[shards[s].reps[i].first].combos[shards[s].reps[i].second];
      cb.out_off = sbase[s] + shards[s].size[i];
      combos[cbase[s] + i] = std::move(cb);
    }
  });
  remap.assign(nch, {});
  parallel_for(nch, nt, [&](int64_t c, int) {
    remap[c].resize(local[c].size());
    for (size_t q = 0; q < local[c].size(); q++) {
      const int s = (int)(local[c][q] >> 40), i = (int)(local[c][q] & ((1LL << 40) - 1));
      remap[c][q] = sbase[s] + shards[s].size[i];
    }
  });
  const double w2 = wall();
  nterms = 0;
  size_t ndiag = 0;
  chunks.assign(nch, {});
  for (int64_t c = 0; c < nch; c++) {
    flops += parts[c].flops;
    nterms += (int64_t)parts[c].terms.size();
    ndiag += parts[c].diag.size();
    chunks[c] = std::move(parts[c].terms);
  }
  diag.reserve(ndiag);
  for (auto& p : parts) diag.insert(diag.end(), p.diag.begin(), p.diag.end());
  leftover = std::move(parts);
  phase_s[0] = w1 - w0;
  phase_s[1] = w2 - w1;
  phase_s[2] = wall() - w2;
}

// ---------------------------------------------------------------- group 1 --
void HeffPlanner::d1(int ik) {  // HeffDiagrams1.cpp:26-63
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  if (S->N1[ik] == 2) term(ik, ik, P->g(index, index, index, index), 'I', Ref(), 'I', Ref());          // 1C
  if (S->N2[ik] == 2) term(ik, ik, P->g(index + 1, index + 1, index + 1, index + 1), 'I', Ref(), 'I', Ref());  // 1D
  if (!atLeft) term(ik, ik, 1.0, 'N', Lo->X->blk(nl, sl, il, nl, sl, il), 'I', Ref());   // 1A
  if (!atRight) term(ik, ik, 1.0, 'I', Ref(), 'T', Ro->X->blk(nr, sr, ir, nr, sr, ir));  // 1B
}

// ------------------------------------------------ group 2: local operators --
void HeffPlanner::d2loc(int ik) {  // HeffDiagrams2.cpp:850-1511
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  // 2d (on-site two-site integrals), HeffDiagrams2.cpp:996-1058
  {
    const int i = index, j = index + 1;
    if (n1 == 2 && n2 == 0) { int k = S->kappa(nl, sl, il, 0, 2, 0, nr, sr, ir); if (k >= 0) term(ik, k, P->g(i, i, j, j), 'I', Ref(), 'I', Ref()); }
    if (n1 == 0 && n2 == 2) { int k = S->kappa(nl, sl, il, 2, 0, 0, nr, sr, ir); if (k >= 0) term(ik, k, P->g(i, i, j, j), 'I', Ref(), 'I', Ref()); }
    if (n1 == 2 && n2 == 2) term(ik, ik, 4 * P->g(i, j, i, j) - 2 * P->g(i, j, j, i), 'I', Ref(), 'I', Ref());
    if (n1 == 1 && n2 == 1) term(ik, ik, P->g(i, j, i, j) + (tj == 0 ? 1 : -1) * P->g(i, j, j, i), 'I', Ref(), 'I', Ref());
    if ((n1 == 2 && n2 == 1) || (n1 == 1 && n2 == 2)) term(ik, ik, 2 * P->g(i, j, i, j) - P->g(i, j, j, i), 'I', Ref(), 'I', Ref());
  }
  if (!atLeft) {
    // 2b1/2b2 and 2c1/2c2: local pair annihilators at sites i, i+1 (A at the left boundary)
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2;
      OpT* A = Lo->A[0][which];
      if (nloc == 0 && bk->dim(index, nl - 2, sl, il) > 0) {
        int k = which == 0 ? S->kappa(nl - 2, sl, il, 2, n2, tj, nr, sr, ir) : S->kappa(nl - 2, sl, il, n1, 2, tj, nr, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'T', A->blk(nl - 2, sl, il, nl, sl, il), 'I', Ref());
      }
      if (nloc == 2 && bk->dim(index, nl + 2, sl, il) > 0) {
        int k = which == 0 ? S->kappa(nl + 2, sl, il, 0, n2, tj, nr, sr, ir) : S->kappa(nl + 2, sl, il, n1, 0, tj, nr, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'N', A->blk(nl, sl, il, nl + 2, sl, il), 'I', Ref());
      }
    }
    // 2b3/2c3 spin 0: C_{ii} and C_{i+1,i+1} at the left boundary
    if (n1 != 0) term(ik, ik, (n1 == 2 ? 1.0 : 0.5) * SQ2, 'T', Lo->C[0][0]->blk(nl, sl, il, nl, sl, il), 'I', Ref());
    if (n2 != 0) term(ik, ik, (n2 == 2 ? 1.0 : 0.5) * SQ2, 'T', Lo->C[0][1]->blk(nl, sl, il, nl, sl, il), 'I', Ref());
    // 2b3/2c3 spin 1
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2, nother = which == 0 ? n2 : n1;
      if (nloc != 1) continue;
      OpT* D = Lo->D[0][which];
      const int so = (nother == 1) ? 1 : 0;
      for (int sld = sl - 2; sld <= sl + 2; sld += 2) {
        if (bk->dim(index, nl, sld, il) <= 0) continue;
        Ref Db = D->blk(nl, sld, il, nl, sl, il);
        const int jstart = (sr != sld || so == 0) ? 1 + so : 0;
        for (int jd = jstart; jd <= 1 + so; jd += 2) {
          if (std::abs(sld - sr) > jd) continue;
          int k = S->kappa(nl, sld, il, n1, n2, jd, nr, sr, ir);
          if (k < 0) continue;
          const int fase = which == 0 ? phase(sld + sr + tj + so + jd - 1) : phase(sld + sr + 2 * tj + so - 1);
          const double a = fase * std::sqrt(3.0 * (tj + 1) * (jd + 1) * (sl + 1)) * wigner6j(jd, tj, 2, 1, 1, so) * wigner6j(jd, tj, 2, sl, sld, sr);
          term(ik, k, a, 'T', Db, 'I', Ref());
        }
      }
    }
  }
  if (!atRight) {
    // 2e1/2e2 (site i) and 2f1/2f2 (site i+1): A at the right boundary
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2;
      OpT* A = Ro->A[0][which == 0 ? 1 : 0];
      if (nloc == 2) {
        int k = which == 0 ? S->kappa(nl, sl, il, 0, n2, tj, nr - 2, sr, ir) : S->kappa(nl, sl, il, n1, 0, tj, nr - 2, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'I', Ref(), 'N', A->blk(nr - 2, sr, ir, nr, sr, ir));
      }
      if (nloc == 0) {
        int k = which == 0 ? S->kappa(nl, sl, il, 2, n2, tj, nr + 2, sr, ir) : S->kappa(nl, sl, il, n1, 2, tj, nr + 2, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'I', Ref(), 'T', A->blk(nr, sr, ir, nr + 2, sr, ir));
      }
    }
    // 2e3/2f3 spin 0
    if (n1 != 0) term(ik, ik, (n1 == 2 ? 1.0 : 0.5) * SQ2, 'I', Ref(), 'N', Ro->C[0][1]->blk(nr, sr, ir, nr, sr, ir));
    if (n2 != 0) term(ik, ik, (n2 == 2 ? 1.0 : 0.5) * SQ2, 'I', Ref(), 'N', Ro->C[0][0]->blk(nr, sr, ir, nr, sr, ir));
    // 2e3/2f3 spin 1
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2, nother = which == 0 ? n2 : n1;
      if (nloc != 1) continue;
      OpT* D = Ro->D[0][which == 0 ? 1 : 0];
      const int so = (nother == 1) ? 1 : 0;
      for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
        if (bk->dim(index + 2, nr, srd, ir) <= 0) continue;
        Ref Db = D->blk(nr, srd, ir, nr, sr, ir);
        const int jstart = (srd != sl || so == 0) ? 1 + so : 0;
        for (int jd = jstart; jd <= 1 + so; jd += 2) {
          if (std::abs(sl - srd) > jd) continue;
          int k = S->kappa(nl, sl, il, n1, n2, jd, nr, srd, ir);
          if (k < 0) continue;
          const int fase = which == 0 ? phase(srd + sl + 2 * tj + so + 1) : phase(srd + sl + tj + so + jd + 1);
          const double a = fase * std::sqrt(3.0 * (tj + 1) * (jd + 1) * (srd + 1)) * wigner6j(jd, tj, 2, 1, 1, so) * wigner6j(jd, tj, 2, sr, srd, sl);
          term(ik, k, a, 'I', Ref(), 'N', Db);
        }
      }
    }
  }
}

// ------------------------------------------------- group 2: pair sums 2a ----
void HeffPlanner::d2a(int ik) {  // HeffDiagrams2.cpp:28-849
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const bool leftSum = index < L * 0.5;
  // pairs (a <= b) with strict flag; returns the two pair operators (left op, right op)
  // 2a1/2a2 spin 0: S0 x A (S0 on the short side)
  for (int up = 0; up < 2; up++) {  // up == 0: 2a1 (N-2), up == 1: 2a2 (N+2)
    const int dn = up == 0 ? -2 : 2;
    if (leftSum) {
      for (int a = 0; a < index; a++)
        for (int b = a; b < index; b++) {
          OpT* Sl = pl(Lo->S0, a, b); OpT* Ar = cr(Ro->A, a, b);
          const int ild = il ^ Sl->n_irrep, ird = ir ^ Ar->n_irrep;
          int k = S->kappa(nl + dn, sl, ild, n1, n2, tj, nr + dn, sr, ird);
          if (k < 0) continue;
          if (up == 0) term(ik, k, 1.0, 'T', Sl->blk(nl - 2, sl, ild, nl, sl, il), 'N', Ar->blk(nr - 2, sr, ird, nr, sr, ir));
          else term(ik, k, 1.0, 'N', Sl->blk(nl, sl, il, nl + 2, sl, ild), 'T', Ar->blk(nr, sr, ir, nr + 2, sr, ird));
        }
    } else {
      for (int g = index + 2; g < L; g++)
        for (int d = g; d < L; d++) {
          OpT* Al = cl(Lo->A, g, d); OpT* Sr = pr(Ro->S0, g, d);
          const int ild = il ^ Al->n_irrep, ird = ir ^ Sr->n_irrep;
          int k = S->kappa(nl + dn, sl, ild, n1, n2, tj, nr + dn, sr, ird);
          if (k < 0) continue;
          if (up == 0) term(ik, k, 1.0, 'T', Al->blk(nl - 2, sl, ild, nl, sl, il), 'N', Sr->blk(nr - 2, sr, ird, nr, sr, ir));
          else term(ik, k, 1.0, 'N', Al->blk(nl, sl, il, nl + 2, sl, ild), 'T', Sr->blk(nr, sr, ir, nr + 2, sr, ird));
        }
    }
  }
  // 2a1/2a2 spin 1: S1 x B
  for (int up = 0; up < 2; up++) {
    const int dn = up == 0 ? -2 : 2;
    for (int sld = sl - 2; sld <= sl + 2; sld += 2)
      for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
        if (sld < 0 || srd < 0 || std::abs(sld - srd) > tj) continue;
        const double f = up == 0 ? phase(srd + sl + tj + 2) * std::sqrt((sr + 1) * (sl + 1.0)) * wigner6j(sld, srd, tj, sr, sl, 2)
                                 : phase(sld + sr + tj + 2) * std::sqrt((srd + 1) * (sld + 1.0)) * wigner6j(sld, srd, tj, sr, sl, 2);
        if (leftSum) {
          for (int a = 0; a < index; a++)
            for (int b = a + 1; b < index; b++) {
              OpT* Sl = pl(Lo->S1, a, b); OpT* Br = cr(Ro->B, a, b);
              const int ild = il ^ Sl->n_irrep, ird = ir ^ Br->n_irrep;
              int k = S->kappa(nl + dn, sld, ild, n1, n2, tj, nr + dn, srd, ird);
              if (k < 0) continue;
              if (up == 0) term(ik, k, f, 'T', Sl->blk(nl - 2, sld, ild, nl, sl, il), 'N', Br->blk(nr - 2, srd, ird, nr, sr, ir));
              else term(ik, k, f, 'N', Sl->blk(nl, sl, il, nl + 2, sld, ild), 'T', Br->blk(nr, sr, ir, nr + 2, srd, ird));
            }
        } else {
          for (int g = index + 2; g < L; g++)
            for (int d = g + 1; d < L; d++) {
              OpT* Bl = cl(Lo->B, g, d); OpT* Sr = pr(Ro->S1, g, d);
              const int ild = il ^ Bl->n_irrep, ird = ir ^ Sr->n_irrep;
              int k = S->kappa(nl + dn, sld, ild, n1, n2, tj, nr + dn, srd, ird);
              if (k < 0) continue;
              if (up == 0) term(ik, k, f, 'T', Bl->blk(nl - 2, sld, ild, nl, sl, il), 'N', Sr->blk(nr - 2, srd, ird, nr, sr, ir));
              else term(ik, k, f, 'N', Bl->blk(nl, sl, il, nl + 2, sld, ild), 'T', Sr->blk(nr, sr, ir, nr + 2, srd, ird));
            }
        }
      }
  }
  // 2a3 spin 0: F0 x C
  if (leftSum) {
    for (int g = 0; g < index; g++)
      for (int a = g + 1; a < index; a++) {
        OpT* Fl = pl(Lo->F0, g, a); OpT* Cr = cr(Ro->C, g, a);
        const int ild = il ^ Fl->n_irrep, ird = ir ^ Cr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'N', Fl->blk(nl, sl, il, nl, sl, ild), 'T', Cr->blk(nr, sr, ir, nr, sr, ird));
      }
    for (int a = 0; a < index; a++)
      for (int g = a; g < index; g++) {
        OpT* Fl = pl(Lo->F0, a, g); OpT* Cr = cr(Ro->C, a, g);
        const int ild = il ^ Fl->n_irrep, ird = ir ^ Cr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'T', Fl->blk(nl, sl, ild, nl, sl, il), 'N', Cr->blk(nr, sr, ird, nr, sr, ir));
      }
  } else {
    for (int d = index + 2; d < L; d++)
      for (int b = d + 1; b < L; b++) {
        OpT* Cl = cl(Lo->C, d, b); OpT* Fr = pr(Ro->F0, d, b);
        const int ild = il ^ Cl->n_irrep, ird = ir ^ Fr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'N', Cl->blk(nl, sl, il, nl, sl, ild), 'T', Fr->blk(nr, sr, ir, nr, sr, ird));
      }
    for (int b = index + 2; b < L; b++)
      for (int d = b; d < L; d++) {
        OpT* Cl = cl(Lo->C, b, d); OpT* Fr = pr(Ro->F0, b, d);
        const int ild = il ^ Cl->n_irrep, ird = ir ^ Fr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'T', Cl->blk(nl, sl, ild, nl, sl, il), 'N', Fr->blk(nr, sr, ird, nr, sr, ir));
      }
  }
  // 2a3 spin 1: F1 x D
  for (int sld = sl - 2; sld <= sl + 2; sld += 2)
    for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
      if (sld < 0 || srd < 0 || std::abs(sld - srd) > tj) continue;
      const double f1 = phase(sld + srd + tj + 2) * std::sqrt((sr + 1) * (sld + 1.0)) * wigner6j(sld, srd, tj, sr, sl, 2);
      const double f2 = phase(sl + sr + tj + 2) * std::sqrt((srd + 1) * (sl + 1.0)) * wigner6j(sld, srd, tj, sr, sl, 2);
      if (leftSum) {
        for (int g = 0; g < index; g++)
          for (int a = g + 1; a < index; a++) {
            OpT* Fl = pl(Lo->F1, g, a); OpT* Dr = cr(Ro->D, g, a);
            const int ild = il ^ Fl->n_irrep, ird = ir ^ Dr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f1, 'N', Fl->blk(nl, sl, il, nl, sld, ild), 'T', Dr->blk(nr, sr, ir, nr, srd, ird));
          }
        for (int a = 0; a < index; a++)
          for (int g = a; g < index; g++) {
            OpT* Fl = pl(Lo->F1, a, g); OpT* Dr = cr(Ro->D, a, g);
            const int ild = il ^ Fl->n_irrep, ird = ir ^ Dr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f2, 'T', Fl->blk(nl, sld, ild, nl, sl, il), 'N', Dr->blk(nr, srd, ird, nr, sr, ir));
          }
      } else {
        for (int d = index + 2; d < L; d++)
          for (int b = d + 1; b < L; b++) {
            OpT* Dl = cl(Lo->D, d, b); OpT* Fr = pr(Ro->F1, d, b);
            const int ild = il ^ Dl->n_irrep, ird = ir ^ Fr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f1, 'N', Dl->blk(nl, sl, il, nl, sld, ild), 'T', Fr->blk(nr, sr, ir, nr, srd, ird));
          }
        for (int b = index + 2; b < L; b++)
          for (int d = b; d < L; d++) {
            OpT* Dl = cl(Lo->D, b, d); OpT* Fr = pr(Ro->F1, b, d);
            const int ild = il ^ Dl->n_irrep, ird = ir ^ Fr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f2, 'T', Dl->blk(nl, sld, ild, nl, sl, il), 'N', Fr->blk(nr, srd, ird, nr, sr, ir));
          }
      }
    }
}


void HeffPlanner::d3(int ik) {  // HeffDiagrams3.cpp:28-848
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  // 3E / 3H: purely local hopping (HeffDiagrams3.cpp:399-474)
  if (bk->irrep(i) == bk->irrep(j)) {
    auto ax = [&](int a1, int a2, int aj, double a) { int k = S->kappa(nl, sl, il, a1, a2, aj, nr, sr, ir); if (k >= 0) term(ik, k, a, 'I', Ref(), 'I', Ref()); };
    const double giiij = P->g(i, i, i, j), gijjj = P->g(i, j, j, j);
    if (n1 == 2 && n2 == 0) ax(1, 1, 0, SQ2 * giiij);
    if (n1 == 2 && n2 == 1) ax(1, 2, 1, -(giiij + gijjj));
    if (n1 == 1 && n2 == 1 && tj == 0) { ax(2, 0, 0, SQ2 * giiij); ax(0, 2, 0, SQ2 * gijjj); }
    if (n1 == 1 && n2 == 2) ax(2, 1, 1, -(giiij + gijjj));
    if (n1 == 0 && n2 == 2) ax(1, 1, 0, SQ2 * gijjj);
  }
  if (!atLeft) {
    // 3A/3D (site i, which==0) and 3B/3I (site i+1, which==1): Q at the left boundary
    for (int which = 0; which < 2; which++) {
      OpT* Q = Lo->Q[which];
      const int site = index + which;
      const int nloc = which == 0 ? n1 : n2, so = which == 0 ? (n2 == 1) : (n1 == 1);
      const int ild = il ^ bk->irrep(site);
      auto kap = [&](int nld, int sld, int nloc_new, int jd) {
        return which == 0 ? S->kappa(nld, sld, ild, nloc_new, n2, jd, nr, sr, ir) : S->kappa(nld, sld, ild, n1, nloc_new, jd, nr, sr, ir);
      };
      for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
        if (nloc == 2 && bk->dim(index, nl + 1, sld, ild) > 0) {  // 3A1A+3D1 / 3B1A+3I2
          const int jstart = (sr != sld || so == 0) ? 1 + so : 0;
          for (int jd = jstart; jd <= 1 + so; jd += 2) {
            if (std::abs(sld - sr) > jd) continue;
            int k = kap(nl + 1, sld, 1, jd);
            if (k < 0) continue;
            const int fase = which == 0 ? phase(sl + sr + 2 + so) : phase(sl + sr + 3 - jd);
            const double f = std::sqrt((jd + 1) * (sld + 1.0)) * fase * wigner6j(jd, so, 1, sl, sld, sr);
            w2 = q_plus_l(Q, Lo->L, site, true, nl, sl, il, nl + 1, sld, ild);
            term(ik, k, f, 'N', w2, 'I', Ref());
          }
        }
        if (nloc == 1 && std::abs(sld - sr) <= so && sld >= 0) {  // 3A1B / 3B1B
          int k = kap(nl + 1, sld, 0, so);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sl + sr + 1 + so) : phase(sl + sr + 2 - tj);
            const double f = std::sqrt((sld + 1) * (tj + 1.0)) * fase * wigner6j(so, tj, 1, sl, sld, sr);
            term(ik, k, f, 'N', Q->blk(nl, sl, il, nl + 1, sld, ild), 'I', Ref());
          }
        }
        if (nloc == 0 && bk->dim(index, nl - 1, sld, ild) > 0) {  // 3A2A / 3B2A
          const int jstart = (sr != sld || so == 0) ? 1 + so : 0;
          for (int jd = jstart; jd <= 1 + so; jd += 2) {
            if (std::abs(sld - sr) > jd) continue;
            int k = kap(nl - 1, sld, 1, jd);
            if (k < 0) continue;
            const int fase = which == 0 ? phase(sld + sr + 1 + so) : phase(sld + sr + 2 - jd);
            const double f = fase * std::sqrt((sl + 1) * (jd + 1.0)) * wigner6j(jd, so, 1, sl, sld, sr);
            term(ik, k, f, 'T', Q->blk(nl - 1, sld, ild, nl, sl, il), 'I', Ref());
          }
        }
        if (nloc == 1 && std::abs(sld - sr) <= so && sld >= 0) {  // 3A2B+3D2 / 3B2B+3I1
          int k = kap(nl - 1, sld, 2, so);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sld + sr + 2 + so) : phase(sld + sr + 3 - tj);
            const double f = fase * std::sqrt((sl + 1) * (tj + 1.0)) * wigner6j(so, tj, 1, sl, sld, sr);
            w2 = q_plus_l(Q, Lo->L, site, true, nl - 1, sld, ild, nl, sl, il);
            term(ik, k, f, 'T', w2, 'I', Ref());
          }
        }
      }
    }
  }
  if (!atRight) {
    // 3K/3F (site i, which==0) and 3L/3G (site i+1, which==1): Q at the right boundary
    for (int which = 0; which < 2; which++) {
      OpT* Q = Ro->Q[which == 0 ? 1 : 0];
      const int site = index + which;
      const int nloc = which == 0 ? n1 : n2, so = which == 0 ? (n2 == 1) : (n1 == 1);
      const int ird = ir ^ bk->irrep(site);
      auto kap = [&](int nloc_new, int jd, int nrd, int srd) {
        return which == 0 ? S->kappa(nl, sl, il, nloc_new, n2, jd, nrd, srd, ird) : S->kappa(nl, sl, il, n1, nloc_new, jd, nrd, srd, ird);
      };
      for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
        if (nloc == 1 && std::abs(sl - srd) <= so && srd >= 0) {  // 3K1A / 3L1A
          int k = kap(0, so, nr - 1, srd);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sl + sr + tj + 2 * so) : phase(sl + sr + so + 1);
            const double f = std::sqrt((tj + 1) * (sr + 1.0)) * fase * wigner6j(so, tj, 1, sr, srd, sl);
            term(ik, k, f, 'I', Ref(), 'N', Q->blk(nr - 1, srd, ird, nr, sr, ir));
          }
        }
        if (nloc == 2 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {  // 3K1B+3F1 / 3L1B+3G1
          const int jstart = (srd != sl || so == 0) ? 1 + so : 0;
          for (int jd = jstart; jd <= 1 + so; jd += 2) {
            if (std::abs(sl - srd) > jd) continue;
            int k = kap(1, jd, nr - 1, srd);
            if (k < 0) continue;
            const int fase = which == 0 ? phase(sl + sr + jd + 1 + 2 * so) : phase(sl + sr + so + 2);
            const double f = std::sqrt((jd + 1) * (sr + 1.0)) * fase * wigner6j(jd, so, 1, sr, srd, sl);
            w2 = q_plus_l(Q, Ro->L, site, false, nr - 1, srd, ird, nr, sr, ir);
            term(ik, k, f, 'I', Ref(), 'N', w2);
          }
        }
        if (nloc == 0 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {  // 3K2A / 3L2A
          const int jstart = (srd != sl || so == 0) ? 1 + so : 0;
          for (int jd = jstart; jd <= 1 + so; jd += 2) {
            if (std::abs(sl - srd) > jd) continue;
            int k = kap(1, jd, nr + 1, srd);
            if (k < 0) continue;
            const int fase = which == 0 ? phase(sl + srd + jd + 2 * so) : phase(sl + srd + so + 1);
            const double f = std::sqrt((jd + 1) * (srd + 1.0)) * fase * wigner6j(jd, so, 1, sr, srd, sl);
            term(ik, k, f, 'I', Ref(), 'T', Q->blk(nr, sr, ir, nr + 1, srd, ird));
          }
        }
        if (nloc == 1 && std::abs(sl - srd) <= so && srd >= 0) {  // 3K2B+3F2 / 3L2B+3G2
          int k = kap(2, so, nr + 1, srd);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sl + srd + tj + 1 + 2 * so) : phase(sl + srd + so + 2);
            const double f = std::sqrt((tj + 1) * (srd + 1.0)) * fase * wigner6j(so, tj, 1, sr, srd, sl);
            w2 = q_plus_l(Q, Ro->L, site, false, nr, sr, ir, nr + 1, srd, ird);
            term(ik, k, f, 'I', Ref(), 'T', w2);
          }
        }
      }
    }
  }
  if (!atLeft && !atRight) {
    const int ex = (n1 == 1 ? 2 : 0) + (n2 == 1 ? 2 : 0);
    for (int sld = sl - 1; sld <= sl + 1; sld += 2)
      for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
        if (std::abs(sld - srd) > tj || sld < 0 || srd < 0) continue;
        const double fup = phase(sld + sr + tj + 1 + ex) * std::sqrt((sld + 1) * (srd + 1.0)) * wigner6j(sl, sr, tj, srd, sld, 1);
        const double fdn = phase(sl + srd + tj + 1 + ex) * std::sqrt((sl + 1) * (sr + 1.0)) * wigner6j(sl, sr, tj, srd, sld, 1);
        // 3C: Q_l at the left boundary, L_l at the right boundary, l right of the sites (HeffDiagrams3.cpp:294-398)
        for (int l = index + 2; l < L; l++) {
          const int ild = il ^ bk->irrep(l), ird = ir ^ bk->irrep(l);
          OpT* Q = Lo->Q[l - index];
          OpT* Lr = Ro->L[l - index - 2];
          int k = S->kappa(nl + 1, sld, ild, n1, n2, tj, nr + 1, srd, ird);
          if (k >= 0) term(ik, k, fup, 'N', Q->blk(nl, sl, il, nl + 1, sld, ild), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird));
          k = S->kappa(nl - 1, sld, ild, n1, n2, tj, nr - 1, srd, ird);
          if (k >= 0) term(ik, k, fdn, 'T', Q->blk(nl - 1, sld, ild, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir));
        }
        // 3J: L_l at the left boundary, Q_l at the right boundary, l left of the sites (HeffDiagrams3.cpp:742-848)
        for (int l = 0; l < index; l++) {
          const int ild = il ^ bk->irrep(l), ird = ir ^ bk->irrep(l);
          OpT* Ll = Lo->L[index - 1 - l];
          OpT* Q = Ro->Q[index + 1 - l];
          int k = S->kappa(nl + 1, sld, ild, n1, n2, tj, nr + 1, srd, ird);
          if (k >= 0) term(ik, k, fup, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', Q->blk(nr, sr, ir, nr + 1, srd, ird));
          k = S->kappa(nl - 1, sld, ild, n1, n2, tj, nr - 1, srd, ird);
          if (k >= 0) term(ik, k, fdn, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', Q->blk(nr - 1, srd, ird, nr, sr, ir));
        }
      }
  }
}


// ---------------------------------------------------------------- group 4 --

void HeffPlanner::d4(int ik) {  // HeffDiagrams4.cpp:28-6518
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  if (!atLeft) {
    // ---- 4A: A/B/C/D of the pair (i, i+1) at the left boundary (HeffDiagrams4.cpp:28-669)
    {
      OpT* A = Lo->A[1][0];
      const int ild = il ^ A->n_irrep;
      auto dn = [&](int a1, int a2, int aj, double f) {  // kappa' at NL-2
        int k = S->kappa(nl - 2, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', A->blk(nl - 2, sl, ild, nl, sl, il), 'I', Ref());
      };
      auto up = [&](int a1, int a2, int aj, double f) {  // kappa' at NL+2
        int k = S->kappa(nl + 2, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', A->blk(nl, sl, il, nl + 2, sl, ild), 'I', Ref());
      };
      if (n1 == 0 && n2 == 0) dn(1, 1, 0, 1.0);
      if (n1 == 1 && n2 == 0) dn(2, 1, 1, -SQH);
      if (n1 == 0 && n2 == 1) dn(1, 2, 1, -SQH);
      if (n1 == 1 && n2 == 1 && tj == 0) { dn(2, 2, 0, -1.0); up(0, 0, 0, 1.0); }
      if (n1 == 2 && n2 == 1) up(1, 0, 1, -SQH);
      if (n1 == 1 && n2 == 2) up(0, 1, 1, -SQH);
      if (n1 == 2 && n2 == 2) up(1, 1, 0, -1.0);
    }
    {
      OpT* B = Lo->B[1][0];
      const int ild = il ^ B->n_irrep;
      auto dn = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl - 2, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', B->blk(nl - 2, sld, ild, nl, sl, il), 'I', Ref());
      };
      auto up = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl + 2, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', B->blk(nl, sl, il, nl + 2, sld, ild), 'I', Ref());
      };
      for (int sld = sl - 2; sld <= sl + 2; sld += 2) {
        if (sld < 0) continue;
        const bool near = std::abs(sld - sr) <= 1;
        if (n1 == 0 && n2 == 0) dn(sld, 1, 1, 2, 1.0);
        if (n1 == 1 && n2 == 0 && near) dn(sld, 2, 1, 1, phase(sld + sr + 1) * std::sqrt(3.0 * (sl + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 0 && n2 == 1 && near) dn(sld, 1, 2, 1, phase(sld + sr + 3) * std::sqrt(3.0 * (sl + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 2 && n2 == 1 && near) up(sld, 1, 0, 1, phase(sl + sr + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 1 && n2 == 2 && near) up(sld, 0, 1, 1, phase(sl + sr + 3) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 2 && n2 == 2) up(sld, 1, 1, 2, phase(sld - sl) * std::sqrt((sld + 1.0) / (sl + 1.0)));
      }
      if (n1 == 1 && n2 == 1 && tj == 2) {
        const int sld = sr;
        dn(sld, 2, 2, 0, phase(sld - sl) * std::sqrt((sl + 1.0) / (sld + 1.0)));
        up(sld, 0, 0, 0, 1.0);
      }
    }
    {
      OpT* C = Lo->C[1][0];
      const int ild = il ^ C->n_irrep;
      auto tr = [&](int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', C->blk(nl, sl, ild, nl, sl, il), 'I', Ref());
      };
      auto nt = [&](int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', C->blk(nl, sl, il, nl, sl, ild), 'I', Ref());
      };
      if (n1 == 1 && n2 == 0) tr(0, 1, 1, SQH);
      if (n1 == 1 && n2 == 1 && tj == 0) { tr(0, 2, 0, 1.0); nt(2, 0, 0, 1.0); }
      if (n1 == 2 && n2 == 0) tr(1, 1, 0, 1.0);
      if (n1 == 2 && n2 == 1) tr(1, 2, 1, -SQH);
      if (n1 == 0 && n2 == 1) nt(1, 0, 1, SQH);
      if (n1 == 0 && n2 == 2) nt(1, 1, 0, 1.0);
      if (n1 == 1 && n2 == 2) nt(2, 1, 1, -SQH);
    }
    {
      OpT* D = Lo->D[1][0];
      const int ild = il ^ D->n_irrep;
      auto tr = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', D->blk(nl, sld, ild, nl, sl, il), 'I', Ref());
      };
      auto nt = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', D->blk(nl, sl, il, nl, sld, ild), 'I', Ref());
      };
      for (int sld = sl - 2; sld <= sl + 2; sld += 2) {
        if (sld < 0) continue;
        const bool near = std::abs(sld - sr) <= 1;
        if (n1 == 1 && n2 == 0 && near) tr(sld, 0, 1, 1, phase(sld + sr + 1) * std::sqrt(3.0 * (sl + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 2 && n2 == 0) tr(sld, 1, 1, 2, -1.0);
        if (n1 == 2 && n2 == 1 && near) tr(sld, 1, 2, 1, phase(sld + sr + 1) * std::sqrt(3.0 * (sl + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 0 && n2 == 1 && near) nt(sld, 1, 0, 1, phase(sl + sr + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
        if (n1 == 0 && n2 == 2) nt(sld, 1, 1, 2, phase(sl - sld) * std::sqrt((sld + 1.0) / (sl + 1.0)));
        if (n1 == 1 && n2 == 2 && near) nt(sld, 2, 1, 1, phase(sl + sr + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 1, 2, sl, sld, sr));
      }
      if (n1 == 1 && n2 == 1 && tj == 2) {
        const int sld = sr;
        tr(sld, 0, 2, 0, phase(sld - sl) * std::sqrt((sl + 1.0) / (sld + 1.0)));
        nt(sld, 2, 0, 0, -1.0);
      }
    }
    // ---- 4D: sum_l V L_l at the left boundary, l left of the sites, irrep(l) == irrep(i+1) (HeffDiagrams4.cpp:2334-2595)
    {
      const int ild = il ^ bk->irrep(j);
      std::vector<int> ls;
      for (int l = 0; l < index; l++) if (bk->irrep(l) == bk->irrep(j)) ls.push_back(l);
      if (!ls.empty()) {
        auto comb = [&](int nlu, int slu, int ilu, int nld, int sld, int ildn, const std::vector<double>& co) {
          std::vector<OpT*> ops;
          for (int l : ls) ops.push_back(Lo->L[index - 1 - l]);
          w2 = lin_comb(ops, co, nlu, slu, ilu, nld, sld, ildn, bk->dim(index, nlu, slu, ilu) * bk->dim(index, nld, sld, ildn));
        };
        std::vector<double> co(ls.size());
        if (n1 == 0 && n2 > 0) {  // 4D1
          const int s2d = (n2 == 1) ? 0 : 1;
          for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
            if (std::abs(sld - sr) > s2d || bk->dim(index, nl - 1, sld, ild) <= 0) continue;
            for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(ls[q], j, i, i);
            const double f = (n2 == 1) ? phase(sr + 1 - sl) * std::sqrt((sl + 1.0) / (sr + 1.0)) : -1.0;
            int k = S->kappa(nl - 1, sld, ild, 2, n2 - 1, s2d, nr, sr, ir);
            if (k < 0) continue;
            comb(nl - 1, sld, ild, nl, sl, il, co);
            term(ik, k, f, 'T', w2, 'I', Ref());
          }
        }
        if (n1 == 2 && n2 < 2) {  // 4D2
          const int s2d = (n2 == 0) ? 1 : 0;
          for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
            if (std::abs(sld - sr) > s2d || bk->dim(index, nl + 1, sld, ild) <= 0) continue;
            for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(ls[q], j, i, i);
            const double f = (n2 == 0) ? phase(sr + 1 - sld) * std::sqrt((sld + 1.0) / (sr + 1.0)) : -1.0;
            int k = S->kappa(nl + 1, sld, ild, 0, n2 + 1, s2d, nr, sr, ir);
            if (k < 0) continue;
            comb(nl, sl, il, nl + 1, sld, ild, co);
            term(ik, k, f, 'N', w2, 'I', Ref());
          }
        }
        if (n1 > 0 && n2 < 2) {  // 4D3
          for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
            if (bk->dim(index, nl - 1, sld, ild) <= 0) continue;
            const int sum = (n1 == 1 ? 1 : 0) + (n2 == 0 ? 1 : 0);
            const int jstart = (sld != sr || sum < 2) ? sum : 0;
            for (int jd = jstart; jd <= sum; jd += 2) {
              if (std::abs(sld - sr) > jd) continue;
              double af = 0.0;
              if (n1 == 1 && n2 == 0) af = phase(sld + sr + 2) * std::sqrt((sl + 1.0) * (jd + 1)) * wigner6j(jd, 1, 1, sl, sld, sr);
              if (n1 == 1 && n2 == 1) af = phase(sld + sr + 3 + tj) * std::sqrt((sl + 1.0) * (tj + 1)) * wigner6j(sl, sr, tj, 1, 1, sld);
              if (n1 == 2 && n2 == 0) af = -1.0;
              if (n1 == 2 && n2 == 1) af = phase(sl + 1 - sr) * std::sqrt((sl + 1.0) / (sr + 1.0));
              for (size_t q = 0; q < ls.size(); q++) {
                const int l = ls[q];
                if (n1 == 1 && n2 == 0) co[q] = af * (P->g(l, i, i, j) + (jd == 0 ? 1 : -1) * P->g(l, i, j, i));
                if (n1 == 1 && n2 == 1) co[q] = af * P->g(l, i, j, i) + (tj == 0 ? SQ2 * P->g(l, i, i, j) : 0.0);
                if (n1 == 2) co[q] = af * (P->g(l, i, i, j) - 2 * P->g(l, i, j, i));
              }
              int k = S->kappa(nl - 1, sld, ild, n1, n2 + 1, jd, nr, sr, ir);
              if (k < 0) continue;
              comb(nl - 1, sld, ild, nl, sl, il, co);
              term(ik, k, 1.0, 'T', w2, 'I', Ref());
            }
          }
        }
        if (n1 > 0 && n2 > 0) {  // 4D4
          for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
            if (bk->dim(index, nl + 1, sld, ild) <= 0) continue;
            const int sum = (n1 == 1 ? 1 : 0) + (n2 == 2 ? 1 : 0);
            const int jstart = (sld != sr || sum < 2) ? sum : 0;
            for (int jd = jstart; jd <= sum; jd += 2) {
              if (std::abs(sld - sr) > jd) continue;
              double af = 0.0;
              if (n1 == 1 && n2 == 1) af = phase(sl + sr + 2) * std::sqrt((sld + 1.0) * (tj + 1)) * wigner6j(tj, 1, 1, sld, sl, sr);
              if (n1 == 1 && n2 == 2) af = phase(sl + sr + 3 + jd) * std::sqrt((sld + 1.0) * (jd + 1)) * wigner6j(sld, sr, jd, 1, 1, sl);
              if (n1 == 2 && n2 == 1) af = -1.0;
              if (n1 == 2 && n2 == 2) af = phase(sld + 1 - sr) * std::sqrt((sld + 1.0) / (sr + 1.0));
              for (size_t q = 0; q < ls.size(); q++) {
                const int l = ls[q];
                if (n1 == 1 && n2 == 1) co[q] = af * (P->g(l, i, i, j) + (tj == 0 ? 1 : -1) * P->g(l, i, j, i));
                if (n1 == 1 && n2 == 2) co[q] = af * P->g(l, i, j, i) + (jd == 0 ? SQ2 * P->g(l, i, i, j) : 0.0);
                if (n1 == 2) co[q] = af * (P->g(l, i, i, j) - 2 * P->g(l, i, j, i));
              }
              int k = S->kappa(nl + 1, sld, ild, n1, n2 - 1, jd, nr, sr, ir);
              if (k < 0) continue;
              comb(nl, sl, il, nl + 1, sld, ild, co);
              term(ik, k, 1.0, 'N', w2, 'I', Ref());
            }
          }
        }
      }
    }
    d4I(ik);
  }
  if (!atRight) {
    d4J(ik);
    d4F(ik);
    d4G(ik);
  }
  if (!atLeft && !atRight) {
    d4BC(ik);
    d4E(ik);
    d4KL(ik);
  }
}

// 4B/4C: A/B/C/D of the pair (site, l) at the left boundary (site = i for 4B, i+1 for 4C)
// times L_l at the right boundary, l right of the sites (HeffDiagrams4.cpp:670-2333).
void HeffPlanner::d4BC(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  for (int which = 0; which < 2; which++) {  // 0: 4B (site i), 1: 4C (site i+1)
    const int nloc = which == 0 ? n1 : n2;
    const int so = which == 0 ? (n2 == 1) : (n1 == 1);
    auto kap = [&](int nld, int sld, int ild, int nloc_new, int jd, int nrd, int srd, int ird) {
      return which == 0 ? S->kappa(nld, sld, ild, nloc_new, n2, jd, nrd, srd, ird) : S->kappa(nld, sld, ild, n1, nloc_new, jd, nrd, srd, ird);
    };
    for (int l = index + 2; l < L; l++) {
      const int kk = which == 0 ? l - index : l - index - 1, mm = which;
      OpT* A = Lo->A[kk][mm]; OpT* B = Lo->B[kk][mm]; OpT* C = Lo->C[kk][mm]; OpT* D = Lo->D[kk][mm];
      OpT* Lr = Ro->L[l - index - 2];
      const int ird = ir ^ bk->irrep(l);
      for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
        if (srd < 0) continue;
        // ---- spin 0, A (N-2 / N+2) and C (N)
        {
          const int ilA = il ^ A->n_irrep, ilC = il ^ C->n_irrep;
          const int jstart = (sl != srd || so == 0) ? so + 1 : 0;
          for (int jd = jstart; jd <= so + 1; jd += 2) {
            if (std::abs(sl - srd) > jd) continue;
            if (nloc == 0) {  // 4B1A / 4C1A (A, down) and 4B4A / 4C4A (C, up)
              double f = which == 0 ? phase(sr + sl + 1 + jd + 2 * so) : phase(sr + sl + 2 + so);
              f *= std::sqrt(0.5 * (sr + 1) * (jd + 1)) * wigner6j(jd, so, 1, sr, srd, sl);
              int k = kap(nl - 2, sl, ilA, 1, jd, nr - 1, srd, ird);
              if (k >= 0) term(ik, k, f, 'T', A->blk(nl - 2, sl, ilA, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
              double fc = which == 0 ? phase(srd + sl + jd + 2 * so) : phase(srd + sl + 1 + so);
              fc *= std::sqrt(0.5 * (srd + 1) * (jd + 1)) * wigner6j(jd, so, 1, sr, srd, sl);
              k = kap(nl, sl, ilC, 1, jd, nr + 1, srd, ird);
              if (k >= 0) term(ik, k, fc, 'N', C->blk(nl, sl, il, nl, sl, ilC), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
            }
            if (nloc == 2) {  // 4B2B / 4C2B (A, up) and 4B3B / 4C3B (C, down)
              double f = which == 0 ? phase(srd + sl + 2 + jd + 2 * so) : phase(srd + sl + 3 + so);
              f *= std::sqrt(0.5 * (srd + 1) * (jd + 1)) * wigner6j(jd, so, 1, sr, srd, sl);
              int k = kap(nl + 2, sl, ilA, 1, jd, nr + 1, srd, ird);
              if (k >= 0) term(ik, k, f, 'N', A->blk(nl, sl, il, nl + 2, sl, ilA), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
              double fc = which == 0 ? phase(sr + sl + 1 + jd + 2 * so) : phase(sr + sl + 2 + so);
              fc *= std::sqrt(0.5 * (sr + 1) * (jd + 1)) * wigner6j(jd, so, 1, sr, srd, sl);
              k = kap(nl, sl, ilC, 1, jd, nr - 1, srd, ird);
              if (k >= 0) term(ik, k, fc, 'T', C->blk(nl, sl, ilC, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
            }
          }
          if (nloc == 1 && std::abs(sl - srd) <= so) {
            // 4B1B / 4C1B (A, down, local 1 -> 2)
            double f = which == 0 ? phase(sr + sl + 2 + tj + 2 * so) : phase(sr + sl + 3 + so);
            f *= std::sqrt(0.5 * (sr + 1) * (tj + 1)) * wigner6j(tj, so, 1, srd, sr, sl);
            int k = kap(nl - 2, sl, ilA, 2, so, nr - 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'T', A->blk(nl - 2, sl, ilA, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
            // 4B2A / 4C2A (A, up, local 1 -> 0)
            f = which == 0 ? phase(srd + sl + 1 + tj + 2 * so) : phase(srd + sl + 2 + so);
            f *= std::sqrt(0.5 * (srd + 1) * (tj + 1)) * wigner6j(tj, so, 1, srd, sr, sl);
            k = kap(nl + 2, sl, ilA, 0, so, nr + 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'N', A->blk(nl, sl, il, nl + 2, sl, ilA), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
            // 4B3A / 4C3A (C, local 1 -> 0, right N-1)
            f = which == 0 ? phase(sr + sl + tj + 2 * so) : phase(sr + sl + 1 + so);
            f *= std::sqrt(0.5 * (sr + 1) * (tj + 1)) * wigner6j(tj, so, 1, srd, sr, sl);
            k = kap(nl, sl, ilC, 0, so, nr - 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'T', C->blk(nl, sl, ilC, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
            // 4B4B / 4C4B (C, local 1 -> 2, right N+1)
            f = which == 0 ? phase(srd + sl + 1 + tj + 2 * so) : phase(srd + sl + 2 + so);
            f *= std::sqrt(0.5 * (srd + 1) * (tj + 1)) * wigner6j(tj, so, 1, srd, sr, sl);
            k = kap(nl, sl, ilC, 2, so, nr + 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'N', C->blk(nl, sl, il, nl, sl, ilC), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
          }
        }
        // ---- spin 1, B (N-2 / N+2) and D (N)
        const int ilB = B ? (il ^ B->n_irrep) : 0, ilD = il ^ D->n_irrep;
        for (int sld = sl - 2; sld <= sl + 2; sld += 2) {
          if (sld < 0) continue;
          const int jstart = (sld != srd || so == 0) ? so + 1 : 0;
          for (int jd = jstart; jd <= so + 1; jd += 2) {
            if (std::abs(sld - srd) > jd) continue;
            if (nloc == 0) {
              // 4B1A / 4C1A spin 1 (B, down)
              double f = which == 0 ? (so == 0 ? 1 : -1) : phase(1 + so - jd);
              f *= std::sqrt(3.0 * (sr + 1) * (sl + 1) * (jd + 1)) * wigner9j(2, sld, sl, 1, srd, sr, 1, jd, so);
              int k = kap(nl - 2, sld, ilB, 1, jd, nr - 1, srd, ird);
              if (k >= 0) term(ik, k, f, 'T', B->blk(nl - 2, sld, ilB, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
              // 4B4A / 4C4A spin 1 (D, up)
              double fd = which == 0 ? phase(srd - sr + sld - sl + 3 + 2 * so) : phase(srd - sr + sld - sl + so - jd);
              fd *= std::sqrt(3.0 * (srd + 1) * (jd + 1) * (sld + 1)) * wigner9j(2, sld, sl, 1, srd, sr, 1, jd, so);
              k = kap(nl, sld, ilD, 1, jd, nr + 1, srd, ird);
              if (k >= 0) term(ik, k, fd, 'N', D->blk(nl, sl, il, nl, sld, ilD), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
            }
            if (nloc == 2) {
              // 4B2B / 4C2B spin 1 (B, up)
              double f = which == 0 ? phase(sld + 3 - sl + srd - sr + 2 * so) : phase(sld - sl + srd - sr + so - jd);
              f *= std::sqrt(3.0 * (srd + 1) * (jd + 1) * (sld + 1)) * wigner9j(2, sld, sl, 1, srd, sr, 1, jd, so);
              int k = kap(nl + 2, sld, ilB, 1, jd, nr + 1, srd, ird);
              if (k >= 0) term(ik, k, f, 'N', B->blk(nl, sl, il, nl + 2, sld, ilB), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
              // 4B3B / 4C3B spin 1 (D, down)
              double fd = which == 0 ? (so == 0 ? -1 : 1) : phase(3 + so - jd);
              fd *= std::sqrt(3.0 * (sr + 1) * (jd + 1) * (sl + 1)) * wigner9j(2, sld, sl, 1, srd, sr, 1, jd, so);
              k = kap(nl, sld, ilD, 1, jd, nr - 1, srd, ird);
              if (k >= 0) term(ik, k, fd, 'T', D->blk(nl, sld, ilD, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
            }
          }
          if (nloc == 1 && std::abs(sld - srd) <= so) {
            // 4B1B / 4C1B spin 1 (B, down, local 1 -> 2)
            double f = which == 0 ? phase(sr - srd + sl + 3 - sld + 2 * so) : phase(sr - srd + sl - sld + so - tj);
            f *= std::sqrt(3.0 * (sr + 1) * (sl + 1) * (tj + 1)) * wigner9j(2, sl, sld, 1, sr, srd, 1, tj, so);
            int k = kap(nl - 2, sld, ilB, 2, so, nr - 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'T', B->blk(nl - 2, sld, ilB, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
            // 4B2A / 4C2A spin 1 (B, up, local 1 -> 0)
            f = which == 0 ? (so == 0 ? 1 : -1) : phase(1 + so - tj);
            f *= std::sqrt(3.0 * (srd + 1) * (sld + 1) * (tj + 1)) * wigner9j(2, sl, sld, 1, sr, srd, 1, tj, so);
            k = kap(nl + 2, sld, ilB, 0, so, nr + 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'N', B->blk(nl, sl, il, nl + 2, sld, ilB), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
            // 4B3A / 4C3A spin 1 (D, local 1 -> 0, right N-1)
            f = which == 0 ? phase(sl - sld + sr - srd + 3 + 2 * so) : phase(sl - sld + sr - srd + so - tj);
            f *= std::sqrt(3.0 * (sr + 1) * (tj + 1) * (sl + 1)) * wigner9j(2, sl, sld, 1, sr, srd, 1, tj, so);
            k = kap(nl, sld, ilD, 0, so, nr - 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'T', D->blk(nl, sld, ilD, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir), true);
            // 4B4B / 4C4B spin 1 (D, local 1 -> 2, right N+1)
            f = which == 0 ? (so == 0 ? -1 : 1) : phase(3 + so - tj);
            f *= std::sqrt(3.0 * (srd + 1) * (tj + 1) * (sld + 1)) * wigner9j(2, sl, sld, 1, sr, srd, 1, tj, so);
            k = kap(nl, sld, ilD, 2, so, nr + 1, srd, ird);
            if (k >= 0) term(ik, k, f, 'N', D->blk(nl, sl, il, nl, sld, ilD), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird), true);
          }
        }
      }
    }
  }
}


// 4E (site i) and 4H (site i+1): L_a at the left boundary times sum_b V L_b at the right
// boundary, a left and b right of the sites (HeffDiagrams4.cpp:2596-3028, 3563-4014).
void HeffPlanner::d4E(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  for (int which = 0; which < 2; which++) {
    const int site = index + which;
    const int nloc = which == 0 ? n1 : n2, so = which == 0 ? (n2 == 1) : (n1 == 1);
    auto kap = [&](int nld, int sld, int ild, int nloc_new, int jd, int nrd, int srd, int ird) {
      return which == 0 ? S->kappa(nld, sld, ild, nloc_new, n2, jd, nrd, srd, ird) : S->kappa(nld, sld, ild, n1, nloc_new, jd, nrd, srd, ird);
    };
    for (int irrep = 0; irrep < bk->nirrep; irrep++) {
      std::vector<int> la, lb;
      for (int a = 0; a < index; a++) if (bk->irrep(a) == irrep) la.push_back(a);
      for (int b = index + 2; b < L; b++) if (bk->irrep(b) == irrep) lb.push_back(b);
      if (la.empty() || lb.empty()) continue;
      const int ild = il ^ irrep, ird = ir ^ irrep;
      std::vector<OpT*> Rops;
      for (int b : lb) Rops.push_back(Ro->L[b - index - 2]);
      std::vector<double> co(lb.size());
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
          if (sld < 0 || srd < 0) continue;
          // local 0 -> 2 (4E1 / 4H2): left N-1, right N+1
          if (nloc == 0 && std::abs(sld - srd) <= tj && bk->dim(index, nl - 1, sld, ild) > 0 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {
            const double f = phase(sl + sr - so) * std::sqrt((sl + 1) * (srd + 1.0)) *
                             (which == 0 ? wigner6j(sl, sr, so, srd, sld, 1) : wigner6j(sld, srd, so, sr, sl, 1));
            int k = kap(nl - 1, sld, ild, 2, so, nr + 1, srd, ird);
            if (k >= 0)
              for (int a : la) {
                for (size_t q = 0; q < lb.size(); q++) co[q] = P->g(a, lb[q], site, site);
                w2 = lin_comb(Rops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
                term(ik, k, f, 'T', Lo->L[index - 1 - a]->blk(nl - 1, sld, ild, nl, sl, il), 'T', w2, true);
              }
          }
          // local 2 -> 0 (4E2 / 4H1): left N+1, right N-1
          if (nloc == 2 && std::abs(sld - srd) <= tj && bk->dim(index, nl + 1, sld, ild) > 0 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {
            const double f = phase(sld + srd - so) * std::sqrt((sld + 1) * (sr + 1.0)) *
                             (which == 0 ? wigner6j(sld, srd, so, sr, sl, 1) : wigner6j(sl, sr, so, srd, sld, 1));
            int k = kap(nl + 1, sld, ild, 0, so, nr - 1, srd, ird);
            if (k >= 0)
              for (int a : la) {
                for (size_t q = 0; q < lb.size(); q++) co[q] = P->g(a, lb[q], site, site);
                w2 = lin_comb(Rops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
                term(ik, k, f, 'N', Lo->L[index - 1 - a]->blk(nl, sl, il, nl + 1, sld, ild), 'N', w2, true);
              }
          }
          // local 2 -> 2 with exchange (4E3B / 4H3B): left N-1, right N-1
          if (nloc == 2 && std::abs(sld - srd) <= tj && bk->dim(index, nl - 1, sld, ild) > 0 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {
            const double f = phase(sl + srd - so + 3) * std::sqrt((sl + 1) * (sr + 1.0)) * wigner6j(sld, srd, so, sr, sl, 1);
            int k = kap(nl - 1, sld, ild, 2, so, nr - 1, srd, ird);
            if (k >= 0)
              for (int a : la) {
                for (size_t q = 0; q < lb.size(); q++) co[q] = P->g(a, site, site, lb[q]) - 2 * P->g(a, site, lb[q], site);
                w2 = lin_comb(Rops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
                term(ik, k, f, 'T', Lo->L[index - 1 - a]->blk(nl - 1, sld, ild, nl, sl, il), 'N', w2, true);
              }
          }
          // local 2 -> 2 with exchange (4E4B / 4H4B): left N+1, right N+1
          if (nloc == 2 && std::abs(sld - srd) <= tj && bk->dim(index, nl + 1, sld, ild) > 0 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {
            const double f = phase(sld + sr - so + 3) * std::sqrt((sld + 1) * (srd + 1.0)) * wigner6j(sl, sr, so, srd, sld, 1);
            int k = kap(nl + 1, sld, ild, 2, so, nr + 1, srd, ird);
            if (k >= 0)
              for (int a : la) {
                for (size_t q = 0; q < lb.size(); q++) co[q] = P->g(a, site, site, lb[q]) - 2 * P->g(a, site, lb[q], site);
                w2 = lin_comb(Rops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
                term(ik, k, f, 'N', Lo->L[index - 1 - a]->blk(nl, sl, il, nl + 1, sld, ild), 'T', w2, true);
              }
          }
          if (nloc == 1) {
            const int jstart = (sld != srd || so == 0) ? so + 1 : 0;
            for (int jd = jstart; jd <= so + 1; jd += 2) {
              if (std::abs(sld - srd) > jd) continue;
              // 4E3A / 4H3A: left N-1, right N-1
              if (bk->dim(index, nl - 1, sld, ild) > 0 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {
                const double f1 = phase(sl + sr + (which == 0 ? tj : jd) + sld + srd + 1 - so) * std::sqrt((tj + 1) * (jd + 1) * (sl + 1) * (sr + 1.0)) *
                                  wigner6j(sl, srd, so, jd, 1, sld) * wigner6j(tj, 1, so, srd, sl, sr);
                const double f2 = (tj == jd) ? phase(sl + srd + tj + 3 + 2 * so) * std::sqrt((sl + 1) * (sr + 1.0)) * wigner6j(sld, srd, tj, sr, sl, 1) : 0.0;
                int k = kap(nl - 1, sld, ild, 1, jd, nr - 1, srd, ird);
                if (k >= 0)
                  for (int a : la) {
                    for (size_t q = 0; q < lb.size(); q++) co[q] = f1 * P->g(a, site, site, lb[q]) + f2 * P->g(a, site, lb[q], site);
                    w2 = lin_comb(Rops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
                    term(ik, k, 1.0, 'T', Lo->L[index - 1 - a]->blk(nl - 1, sld, ild, nl, sl, il), 'N', w2, true);
                  }
              }
              // 4E4A / 4H4A: left N+1, right N+1
              if (bk->dim(index, nl + 1, sld, ild) > 0 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {
                const double f1 = phase(sl + sr + (which == 0 ? jd : tj) + sld + srd + 1 - so) * std::sqrt((tj + 1) * (jd + 1) * (sld + 1) * (srd + 1.0)) *
                                  wigner6j(sld, sr, so, tj, 1, sl) * wigner6j(jd, 1, so, sr, sld, srd);
                const double f2 = (tj == jd) ? phase(sld + sr + tj + 3 + 2 * so) * std::sqrt((sld + 1) * (srd + 1.0)) * wigner6j(sl, sr, tj, srd, sld, 1) : 0.0;
                int k = kap(nl + 1, sld, ild, 1, jd, nr + 1, srd, ird);
                if (k >= 0)
                  for (int a : la) {
                    for (size_t q = 0; q < lb.size(); q++) co[q] = f1 * P->g(a, site, site, lb[q]) + f2 * P->g(a, site, lb[q], site);
                    w2 = lin_comb(Rops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
                    term(ik, k, 1.0, 'N', Lo->L[index - 1 - a]->blk(nl, sl, il, nl + 1, sld, ild), 'T', w2, true);
                  }
              }
            }
          }
        }
    }
  }
}

// 4I: sum_l V L_l at the left boundary, l left of the sites with irrep(l) == irrep(i)
// (HeffDiagrams4.cpp:4015-4276).
void HeffPlanner::d4I(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  std::vector<OpT*> ops; std::vector<int> ls;
  for (int l = 0; l < index; l++) if (bk->irrep(l) == bk->irrep(i)) { ls.push_back(l); ops.push_back(Lo->L[index - 1 - l]); }
  if (ls.empty()) return;
  const int ild = il ^ bk->irrep(i);
  std::vector<double> co(ls.size());
  for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
    if (n1 > 0 && n2 == 0) {  // 4I1
      const int jd = n1 == 2 ? 1 : 0;
      if (std::abs(sld - sr) <= jd && bk->dim(index, nl - 1, sld, ild) > 0) {
        for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(ls[q], i, j, j);
        const double f = n1 == 1 ? phase(sr + 1 - sl) * std::sqrt((sl + 1.0) / (sr + 1.0)) : -1.0;
        int k = S->kappa(nl - 1, sld, ild, n1 - 1, 2, jd, nr, sr, ir);
        if (k >= 0) {
          w2 = lin_comb(ops, co, nl - 1, sld, ild, nl, sl, il, bk->dim(index, nl - 1, sld, ild) * bk->dim(index, nl, sl, il));
          term(ik, k, f, 'T', w2, 'I', Ref());
        }
      }
    }
    if (n1 < 2 && n2 == 2) {  // 4I2
      const int jd = n1 == 0 ? 1 : 0;
      if (std::abs(sld - sr) <= jd && bk->dim(index, nl + 1, sld, ild) > 0) {
        for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(ls[q], i, j, j);
        const double f = n1 == 0 ? phase(sr + 1 - sld) * std::sqrt((sld + 1.0) / (sr + 1.0)) : -1.0;
        int k = S->kappa(nl + 1, sld, ild, n1 + 1, 0, jd, nr, sr, ir);
        if (k >= 0) {
          w2 = lin_comb(ops, co, nl, sl, il, nl + 1, sld, ild, bk->dim(index, nl, sl, il) * bk->dim(index, nl + 1, sld, ild));
          term(ik, k, f, 'N', w2, 'I', Ref());
        }
      }
    }
    if (n1 < 2 && n2 > 0 && bk->dim(index, nl - 1, sld, ild) > 0) {  // 4I3
      const int sum = (n1 == 0 ? 1 : 0) + (n2 == 1 ? 1 : 0);
      const int jstart = (sr != sld || sum < 2) ? sum : 0;
      for (int jd = jstart; jd <= sum; jd += 2) {
        if (std::abs(sld - sr) > jd) continue;
        double pf = 0.0;
        if (n1 == 0 && n2 == 1) pf = phase(sld + sr + 2) * std::sqrt((sl + 1.0) * (jd + 1)) * wigner6j(jd, 1, 1, sl, sld, sr);
        if (n1 == 1 && n2 == 1) pf = phase(sld + sr + 3) * std::sqrt((tj + 1) * (sl + 1.0)) * wigner6j(tj, 1, 1, sld, sl, sr);
        if (n1 == 0 && n2 == 2) pf = 1.0;
        if (n1 == 1 && n2 == 2) pf = phase(sr + 1 - sl) * std::sqrt((sl + 1.0) / (sr + 1.0));
        for (size_t q = 0; q < ls.size(); q++) {
          const int l = ls[q];
          if (n1 == 0 && n2 == 1) co[q] = pf * (P->g(l, j, i, j) + (jd == 0 ? 1 : -1) * P->g(l, j, j, i));
          if (n1 == 1 && n2 == 1) co[q] = pf * P->g(l, j, i, j) + (tj == 0 ? SQ2 * P->g(l, j, j, i) : 0.0);
          if (n2 == 2) co[q] = pf * (2 * P->g(l, j, i, j) - P->g(l, j, j, i));
        }
        int k = S->kappa(nl - 1, sld, ild, n1 + 1, n2, jd, nr, sr, ir);
        if (k < 0) continue;
        w2 = lin_comb(ops, co, nl - 1, sld, ild, nl, sl, il, bk->dim(index, nl - 1, sld, ild) * bk->dim(index, nl, sl, il));
        term(ik, k, 1.0, 'T', w2, 'I', Ref());
      }
    }
    if (n1 > 0 && n2 > 0 && bk->dim(index, nl + 1, sld, ild) > 0) {  // 4I4
      const int sum = (n1 == 2 ? 1 : 0) + (n2 == 1 ? 1 : 0);
      const int jstart = (sld != sr || sum < 2) ? sum : 0;
      for (int jd = jstart; jd <= sum; jd += 2) {
        if (std::abs(sld - sr) > jd) continue;
        double pf = 0.0;
        if (n1 == 1 && n2 == 1) pf = phase(sl + sr + 2) * std::sqrt((sld + 1.0) * (tj + 1)) * wigner6j(tj, 1, 1, sld, sl, sr);
        if (n1 == 2 && n2 == 1) pf = phase(sl + sr + 3) * std::sqrt((jd + 1) * (sld + 1.0)) * wigner6j(jd, 1, 1, sl, sld, sr);
        if (n1 == 1 && n2 == 2) pf = 1.0;
        if (n1 == 2 && n2 == 2) pf = phase(sr + 1 - sld) * std::sqrt((sld + 1.0) / (sr + 1.0));
        for (size_t q = 0; q < ls.size(); q++) {
          const int l = ls[q];
          if (n1 == 1 && n2 == 1) co[q] = pf * (P->g(l, j, i, j) + (tj == 0 ? 1 : -1) * P->g(l, j, j, i));
          if (n1 == 2 && n2 == 1) co[q] = pf * P->g(l, j, i, j) + (jd == 0 ? SQ2 * P->g(l, j, j, i) : 0.0);
          if (n2 == 2) co[q] = pf * (2 * P->g(l, j, i, j) - P->g(l, j, j, i));
        }
        int k = S->kappa(nl + 1, sld, ild, n1 - 1, n2, jd, nr, sr, ir);
        if (k < 0) continue;
        w2 = lin_comb(ops, co, nl, sl, il, nl + 1, sld, ild, bk->dim(index, nl, sl, il) * bk->dim(index, nl + 1, sld, ild));
        term(ik, k, 1.0, 'N', w2, 'I', Ref());
      }
    }
  }
}

// 4F (irrep(l) == irrep(i+1)) and 4G (irrep(l) == irrep(i)): sum_l V L_l at the right
// boundary, l right of the sites (HeffDiagrams4.cpp:3029-3562).
void HeffPlanner::d4F(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  std::vector<OpT*> ops; std::vector<int> ls;
  for (int l = index + 2; l < L; l++) if (bk->irrep(l) == bk->irrep(j)) { ls.push_back(l); ops.push_back(Ro->L[l - index - 2]); }
  if (ls.empty()) return;
  const int ird = ir ^ bk->irrep(j);
  std::vector<double> co(ls.size());
  for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
    if (n1 == 2 && n2 < 2) {  // 4F1
      const int s2d = n2 == 1 ? 0 : 1;
      if (std::abs(sl - srd) <= s2d && bk->dim(index + 2, nr - 1, srd, ird) > 0) {
        for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(i, i, j, ls[q]);
        const double f = n2 == 1 ? -std::sqrt((sr + 1.0) / (srd + 1.0)) : phase(sr + 1 - srd);
        int k = S->kappa(nl, sl, il, 0, n2 + 1, s2d, nr - 1, srd, ird);
        if (k >= 0) {
          w2 = lin_comb(ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
          term(ik, k, f, 'I', Ref(), 'N', w2);
        }
      }
    }
    if (n1 == 0 && n2 > 0) {  // 4F2
      const int s2d = n2 == 1 ? 0 : 1;
      if (std::abs(sl - srd) <= s2d && bk->dim(index + 2, nr + 1, srd, ird) > 0) {
        for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(i, i, j, ls[q]);
        const double f = n2 == 2 ? -std::sqrt((srd + 1.0) / (sr + 1.0)) : phase(srd + 1 - sr);
        int k = S->kappa(nl, sl, il, 2, n2 - 1, s2d, nr + 1, srd, ird);
        if (k >= 0) {
          w2 = lin_comb(ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
          term(ik, k, f, 'I', Ref(), 'T', w2);
        }
      }
    }
    if (n1 > 0 && n2 > 0 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {  // 4F3
      const int sum = (n1 == 1 ? 1 : 0) + (n2 == 2 ? 1 : 0);
      const int jstart = (sl != srd || sum < 2) ? sum : 0;
      for (int jd = jstart; jd <= sum; jd += 2) {
        if (std::abs(sl - srd) > jd) continue;
        double f = 0.0, f2 = 0.0;
        if (n1 == 1 && n2 == 1) f = phase(sl + sr + 2) * std::sqrt((sr + 1.0) * (tj + 1)) * wigner6j(tj, 1, 1, srd, sr, sl);
        if (n1 == 1 && n2 == 2) {
          f = phase(sl + sr + 3) * std::sqrt((sr + 1.0) * (jd + 1)) * wigner6j(jd, 1, 1, sr, srd, sl);
          f2 = jd == 0 ? std::sqrt(2.0 * (sr + 1.0) / (srd + 1.0)) : 0.0;
        }
        if (n1 == 2 && n2 == 1) f = std::sqrt((sr + 1.0) / (srd + 1.0));
        if (n1 == 2 && n2 == 2) f = phase(sr + 1 - srd);
        for (size_t q = 0; q < ls.size(); q++) {
          const int l = ls[q];
          if (n1 == 1 && n2 == 1) co[q] = f * (P->g(i, j, i, l) + (tj == 0 ? 1 : -1) * P->g(i, j, l, i));
          if (n1 == 1 && n2 == 2) co[q] = f * P->g(i, j, i, l) + (jd == 0 ? f2 * P->g(i, j, l, i) : 0.0);
          if (n1 == 2) co[q] = f * (2 * P->g(i, j, i, l) - P->g(i, j, l, i));
        }
        int k = S->kappa(nl, sl, il, n1, n2 - 1, jd, nr - 1, srd, ird);
        if (k < 0) continue;
        w2 = lin_comb(ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
        term(ik, k, 1.0, 'I', Ref(), 'N', w2);
      }
    }
    if (n1 > 0 && n2 < 2 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {  // 4F4
      const int sum = (n1 == 1 ? 1 : 0) + (n2 == 0 ? 1 : 0);
      const int jstart = (sl != srd || sum < 2) ? sum : 0;
      for (int jd = jstart; jd <= sum; jd += 2) {
        if (std::abs(sl - srd) > jd) continue;
        double f = 0.0, f2 = 0.0;
        if (n1 == 1 && n2 == 0) f = phase(sl + srd + 2) * std::sqrt((srd + 1.0) * (jd + 1)) * wigner6j(jd, 1, 1, sr, srd, sl);
        if (n1 == 1 && n2 == 1) {
          f = phase(sl + srd + 3) * std::sqrt((srd + 1.0) * (tj + 1)) * wigner6j(tj, 1, 1, srd, sr, sl);
          f2 = tj == 0 ? std::sqrt(2.0 * (srd + 1.0) / (sr + 1.0)) : 0.0;
        }
        if (n1 == 2 && n2 == 0) f = std::sqrt((srd + 1.0) / (sr + 1.0));
        if (n1 == 2 && n2 == 1) f = phase(srd + 1 - sr);
        for (size_t q = 0; q < ls.size(); q++) {
          const int l = ls[q];
          if (n1 == 1 && n2 == 0) co[q] = f * (P->g(i, j, i, l) + (jd == 0 ? 1 : -1) * P->g(i, j, l, i));
          if (n1 == 1 && n2 == 1) co[q] = f * P->g(i, j, i, l) + (tj == 0 ? f2 * P->g(i, j, l, i) : 0.0);
          if (n1 == 2) co[q] = f * (2 * P->g(i, j, i, l) - P->g(i, j, l, i));
        }
        int k = S->kappa(nl, sl, il, n1, n2 + 1, jd, nr + 1, srd, ird);
        if (k < 0) continue;
        w2 = lin_comb(ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
        term(ik, k, 1.0, 'I', Ref(), 'T', w2);
      }
    }
  }
}

void HeffPlanner::d4G(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  std::vector<OpT*> ops; std::vector<int> ls;
  for (int l = index + 2; l < L; l++) if (bk->irrep(l) == bk->irrep(i)) { ls.push_back(l); ops.push_back(Ro->L[l - index - 2]); }
  if (ls.empty()) return;
  const int ird = ir ^ bk->irrep(i);
  std::vector<double> co(ls.size());
  for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
    if (n1 < 2 && n2 == 2) {  // 4G1
      const int s1d = n1 == 1 ? 0 : 1;
      if (std::abs(sl - srd) <= s1d && bk->dim(index + 2, nr - 1, srd, ird) > 0) {
        for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(i, ls[q], j, j);
        const double f = n1 == 1 ? -std::sqrt((sr + 1.0) / (srd + 1.0)) : phase(sr + 1 - srd);
        int k = S->kappa(nl, sl, il, n1 + 1, 0, s1d, nr - 1, srd, ird);
        if (k >= 0) {
          w2 = lin_comb(ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
          term(ik, k, f, 'I', Ref(), 'N', w2);
        }
      }
    }
    if (n1 > 0 && n2 == 0) {  // 4G2
      const int s1d = n1 == 1 ? 0 : 1;
      if (std::abs(sl - srd) <= s1d && bk->dim(index + 2, nr + 1, srd, ird) > 0) {
        for (size_t q = 0; q < ls.size(); q++) co[q] = P->g(i, ls[q], j, j);
        const double f = n1 == 2 ? -std::sqrt((srd + 1.0) / (sr + 1.0)) : phase(srd + 1 - sr);
        int k = S->kappa(nl, sl, il, n1 - 1, 2, s1d, nr + 1, srd, ird);
        if (k >= 0) {
          w2 = lin_comb(ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
          term(ik, k, f, 'I', Ref(), 'T', w2);
        }
      }
    }
    if (n1 < 2 && n2 > 0 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {  // 4G3
      const int sum = (n1 == 1 ? 0 : 1) + (n2 == 1 ? 1 : 0);
      const int jstart = (sl != srd || sum < 2) ? sum : 0;
      for (int jd = jstart; jd <= sum; jd += 2) {
        if (std::abs(sl - srd) > jd) continue;
        double f = 0.0, f2 = 0.0;
        if (n1 == 0 && n2 == 1) f = phase(sl + srd + 2) * std::sqrt((jd + 1) * (srd + 1.0)) * wigner6j(jd, 1, 1, sr, srd, sl);
        if (n1 == 1 && n2 == 1) {
          f = phase(sl + srd + tj + 3) * std::sqrt((tj + 1) * (srd + 1.0)) * wigner6j(tj, 1, 1, srd, sr, sl);
          f2 = tj == 0 ? std::sqrt(2.0 * (srd + 1.0) / (sr + 1.0)) : 0.0;
        }
        if (n1 == 0 && n2 == 2) f = -std::sqrt((srd + 1.0) / (sr + 1.0));
        if (n1 == 1 && n2 == 2) f = phase(sr + 1 - srd);
        for (size_t q = 0; q < ls.size(); q++) {
          const int l = ls[q];
          if (n1 == 0 && n2 == 1) co[q] = f * (P->g(i, j, j, l) + (jd == 0 ? 1 : -1) * P->g(i, j, l, j));
          if (n1 == 1 && n2 == 1) co[q] = f * P->g(i, j, l, j) + (tj == 0 ? f2 * P->g(i, j, j, l) : 0.0);
          if (n2 == 2) co[q] = f * (P->g(i, j, j, l) - 2 * P->g(i, j, l, j));
        }
        int k = S->kappa(nl, sl, il, n1 + 1, n2, jd, nr + 1, srd, ird);
        if (k < 0) continue;
        w2 = lin_comb(ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
        term(ik, k, 1.0, 'I', Ref(), 'T', w2);
      }
    }
    if (n1 > 0 && n2 > 0 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {  // 4G4
      const int sum = (n1 == 1 ? 0 : 1) + (n2 == 1 ? 1 : 0);
      const int jstart = (sl != srd || sum < 2) ? sum : 0;
      for (int jd = jstart; jd <= sum; jd += 2) {
        if (std::abs(sl - srd) > jd) continue;
        double f = 0.0, f2 = 0.0;
        if (n1 == 1 && n2 == 1) f = phase(sl + sr + 2) * std::sqrt((tj + 1) * (sr + 1.0)) * wigner6j(tj, 1, 1, srd, sr, sl);
        if (n1 == 2 && n2 == 1) {
          f = phase(sl + sr + jd + 3) * std::sqrt((jd + 1) * (sr + 1.0)) * wigner6j(jd, 1, 1, sr, srd, sl);
          f2 = jd == 0 ? std::sqrt(2.0 * (sr + 1.0) / (srd + 1.0)) : 0.0;
        }
        if (n1 == 1 && n2 == 2) f = -std::sqrt((sr + 1.0) / (srd + 1.0));
        if (n1 == 2 && n2 == 2) f = phase(srd + 1 - sr);
        for (size_t q = 0; q < ls.size(); q++) {
          const int l = ls[q];
          if (n1 == 1 && n2 == 1) co[q] = f * (P->g(i, j, j, l) + (tj == 0 ? 1 : -1) * P->g(i, j, l, j));
          if (n1 == 2 && n2 == 1) co[q] = f * P->g(i, j, l, j) + (jd == 0 ? f2 * P->g(i, j, j, l) : 0.0);
          if (n2 == 2) co[q] = f * (P->g(i, j, j, l) - 2 * P->g(i, j, l, j));
        }
        int k = S->kappa(nl, sl, il, n1 - 1, n2, jd, nr - 1, srd, ird);
        if (k < 0) continue;
        w2 = lin_comb(ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
        term(ik, k, 1.0, 'I', Ref(), 'N', w2);
      }
    }
  }
}


// 4J: A/B/C/D of the pair (i, i+1) at the right boundary (HeffDiagrams4.cpp:4277-4846).
void HeffPlanner::d4J(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  {
    OpT* A = Ro->A[1][0];
    const int ird = ir ^ A->n_irrep;
    auto up = [&](int a1, int a2, int aj, double f) {  // right sector N+2
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr + 2, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'T', A->blk(nr, sr, ir, nr + 2, sr, ird));
    };
    auto dn = [&](int a1, int a2, int aj, double f) {  // right sector N-2
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr - 2, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'N', A->blk(nr - 2, sr, ird, nr, sr, ir));
    };
    if (n1 == 0 && n2 == 0) up(1, 1, 0, 1.0);
    if (n1 == 1 && n2 == 0) up(2, 1, 1, -SQH);
    if (n1 == 0 && n2 == 1) up(1, 2, 1, -SQH);
    if (n1 == 1 && n2 == 1 && tj == 0) { up(2, 2, 0, -1.0); dn(0, 0, 0, 1.0); }
    if (n1 == 2 && n2 == 1) dn(1, 0, 1, -SQH);
    if (n1 == 1 && n2 == 2) dn(0, 1, 1, -SQH);
    if (n1 == 2 && n2 == 2) dn(1, 1, 0, -1.0);
  }
  {
    OpT* B = Ro->B[1][0];
    const int ird = ir ^ B->n_irrep;
    auto up = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr + 2, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'T', B->blk(nr, sr, ir, nr + 2, srd, ird));
    };
    auto dn = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr - 2, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'N', B->blk(nr - 2, srd, ird, nr, sr, ir));
    };
    for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
      if (srd < 0) continue;
      const bool near = std::abs(sl - srd) <= 1;
      if (n1 == 0 && n2 == 0) up(srd, 1, 1, 2, std::sqrt((srd + 1.0) / (sr + 1.0)));
      if (n1 == 1 && n2 == 0 && near) up(srd, 2, 1, 1, phase(sl + srd + 3) * std::sqrt(3.0 * (srd + 1)) * wigner6j(1, 1, 2, sr, srd, sl));
      if (n1 == 0 && n2 == 1 && near) up(srd, 1, 2, 1, phase(sl + srd + 1) * std::sqrt(3.0 * (srd + 1)) * wigner6j(1, 1, 2, sr, srd, sl));
      if (n1 == 2 && n2 == 1 && near) dn(srd, 1, 0, 1, phase(sl + sr + 3) * std::sqrt(3.0 * (sr + 1)) * wigner6j(1, 1, 2, srd, sr, sl));
      if (n1 == 1 && n2 == 2 && near) dn(srd, 0, 1, 1, phase(sl + sr + 1) * std::sqrt(3.0 * (sr + 1)) * wigner6j(1, 1, 2, srd, sr, sl));
      if (n1 == 2 && n2 == 2) dn(srd, 1, 1, 2, phase(sr - srd));
    }
    if (n1 == 1 && n2 == 1 && tj == 2) {
      const int srd = sl;
      up(srd, 2, 2, 0, phase(sr - srd));
      dn(srd, 0, 0, 0, std::sqrt((sr + 1.0) / (srd + 1.0)));
    }
  }
  {
    OpT* C = Ro->C[1][0];
    const int ird = ir ^ C->n_irrep;
    auto nt = [&](int a1, int a2, int aj, double f) {  // C block stored (down -> up)
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'N', C->blk(nr, sr, ird, nr, sr, ir));
    };
    auto tr = [&](int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'T', C->blk(nr, sr, ir, nr, sr, ird));
    };
    if (n1 == 1 && n2 == 0) nt(0, 1, 1, SQH);
    if (n1 == 2 && n2 == 0) nt(1, 1, 0, 1.0);
    if (n1 == 1 && n2 == 1 && tj == 0) { nt(0, 2, 0, 1.0); tr(2, 0, 0, 1.0); }
    if (n1 == 2 && n2 == 1) nt(1, 2, 1, -SQH);
    if (n1 == 0 && n2 == 1) tr(1, 0, 1, SQH);
    if (n1 == 0 && n2 == 2) tr(1, 1, 0, 1.0);
    if (n1 == 1 && n2 == 2) tr(2, 1, 1, -SQH);
  }
  {
    OpT* D = Ro->D[1][0];
    const int ird = ir ^ D->n_irrep;
    auto nt = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'N', D->blk(nr, srd, ird, nr, sr, ir));
    };
    auto tr = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', Ref(), 'T', D->blk(nr, sr, ir, nr, srd, ird));
    };
    for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
      if (srd < 0) continue;
      const bool near = std::abs(sl - srd) <= 1;
      if (n1 == 1 && n2 == 0 && near) nt(srd, 0, 1, 1, phase(sl + srd + 3) * std::sqrt(3.0 * (srd + 1)) * wigner6j(1, 1, 2, sr, srd, sl));
      if (n1 == 2 && n2 == 0) nt(srd, 1, 1, 2, -std::sqrt((srd + 1.0) / (sr + 1.0)));
      if (n1 == 1 && n2 == 1 && tj == 2 && sl == srd) nt(srd, 0, 2, 0, phase(sr - srd));
      if (n1 == 2 && n2 == 1 && near) nt(srd, 1, 2, 1, phase(sl + srd + 3) * std::sqrt(3.0 * (srd + 1)) * wigner6j(1, 1, 2, sr, srd, sl));
      if (n1 == 0 && n2 == 1 && near) tr(srd, 1, 0, 1, phase(sl + sr + 3) * std::sqrt(3.0 * (sr + 1)) * wigner6j(1, 1, 2, sr, srd, sl));
      if (n1 == 1 && n2 == 1 && tj == 2 && sl == srd) tr(srd, 2, 0, 0, -std::sqrt((sr + 1.0) / (srd + 1.0)));
      if (n1 == 0 && n2 == 2) tr(srd, 1, 1, 2, phase(sr - srd));
      if (n1 == 1 && n2 == 2 && near) tr(srd, 2, 1, 1, phase(sl + sr + 3) * std::sqrt(3.0 * (sr + 1)) * wigner6j(1, 1, 2, sr, srd, sl));
    }
  }
}

// 4K (site i+1) and 4L (site i): L_l at the left boundary (l left of the sites) times
// A/B/C/D of the pair (l, site) at the right boundary (HeffDiagrams4.cpp:4847-6518).
void HeffPlanner::d4KL(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  for (int which = 0; which < 2; which++) {  // 0: 4K (local site i+1), 1: 4L (local site i)
    const int nloc = which == 0 ? n2 : n1;
    const int so = which == 0 ? (n1 == 1) : (n2 == 1);
    auto kap = [&](int nld, int sld, int ild, int nloc_new, int jd, int nrd, int srd, int ird) {
      return which == 0 ? S->kappa(nld, sld, ild, n1, nloc_new, jd, nrd, srd, ird) : S->kappa(nld, sld, ild, nloc_new, n2, jd, nrd, srd, ird);
    };
    for (int l = 0; l < index; l++) {
      const int kk = which == 0 ? index + 1 - l : index - l, mm = which == 0 ? 0 : 1;
      OpT* A = Ro->A[kk][mm]; OpT* B = Ro->B[kk][mm]; OpT* C = Ro->C[kk][mm]; OpT* D = Ro->D[kk][mm];
      OpT* Ll = Lo->L[index - 1 - l];
      const int ild = il ^ bk->irrep(l);
      const int irA = ir ^ A->n_irrep, irB = B ? (ir ^ B->n_irrep) : 0, irC = ir ^ C->n_irrep, irD = ir ^ D->n_irrep;
      for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
        if (sld < 0) continue;
        // ---- spin 0: A (right N-2 / N+2) and C (right N)
        const int jstart0 = (sld != sr || so == 0) ? so + 1 : 0;
        for (int jd = jstart0; jd <= so + 1; jd += 2) {
          if (std::abs(sld - sr) > jd) continue;
          if (nloc == 2) {
            // 4K1B / 4L1B: A down (NR-2), left N-1
            double f = which == 0 ? phase(sld + sr + jd + 2 + 2 * so) : phase(sld + sr + 3 + so);
            f *= std::sqrt(0.5 * (sl + 1) * (jd + 1)) * wigner6j(jd, so, 1, sl, sld, sr);
            int k = kap(nl - 1, sld, ild, 1, jd, nr - 2, sr, irA);
            if (k >= 0) term(ik, k, f, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', A->blk(nr - 2, sr, irA, nr, sr, ir), true);
            // 4K3B / 4L3B: C, left N+1
            double fc = which == 0 ? phase(sl + sr + jd + 1 + 2 * so) : phase(sl + sr + 2 + so);
            fc *= std::sqrt(0.5 * (sld + 1) * (jd + 1)) * wigner6j(jd, so, 1, sl, sld, sr);
            k = kap(nl + 1, sld, ild, 1, jd, nr, sr, irC);
            if (k >= 0) term(ik, k, fc, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', C->blk(nr, sr, ir, nr, sr, irC), true);
          }
          if (nloc == 0) {
            // 4K2A / 4L2A: A up (NR+2), left N+1
            double f = which == 0 ? phase(sl + sr + jd + 1 + 2 * so) : phase(sl + sr + 2 + so);
            f *= std::sqrt(0.5 * (sld + 1) * (jd + 1)) * wigner6j(jd, so, 1, sl, sld, sr);
            int k = kap(nl + 1, sld, ild, 1, jd, nr + 2, sr, irA);
            if (k >= 0) term(ik, k, f, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', A->blk(nr, sr, ir, nr + 2, sr, irA), true);
            // 4K4A / 4L4A: C, left N-1
            double fc = which == 0 ? phase(sld + sr + jd + 2 * so) : phase(sld + sr + 1 + so);
            fc *= std::sqrt(0.5 * (sl + 1) * (jd + 1)) * wigner6j(jd, so, 1, sl, sld, sr);
            k = kap(nl - 1, sld, ild, 1, jd, nr, sr, irC);
            if (k >= 0) term(ik, k, fc, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', C->blk(nr, sr, irC, nr, sr, ir), true);
          }
        }
        if (nloc == 1 && std::abs(sld - sr) <= so) {
          // 4K1A / 4L1A: A down, local 1 -> 0, left N-1
          double f = which == 0 ? phase(sld + sr + tj + 1 + 2 * so) : phase(sld + sr + 2 + so);
          f *= std::sqrt(0.5 * (sl + 1) * (tj + 1)) * wigner6j(so, tj, 1, sl, sld, sr);
          int k = kap(nl - 1, sld, ild, 0, so, nr - 2, sr, irA);
          if (k >= 0) term(ik, k, f, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', A->blk(nr - 2, sr, irA, nr, sr, ir), true);
          // 4K2B / 4L2B: A up, local 1 -> 2, left N+1
          f = which == 0 ? phase(sl + sr + tj + 2 + 2 * so) : phase(sl + sr + 3 + so);
          f *= std::sqrt(0.5 * (sld + 1) * (tj + 1)) * wigner6j(tj, so, 1, sld, sl, sr);
          k = kap(nl + 1, sld, ild, 2, so, nr + 2, sr, irA);
          if (k >= 0) term(ik, k, f, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', A->blk(nr, sr, ir, nr + 2, sr, irA), true);
          // 4K3A / 4L3A: C, local 1 -> 0, left N+1
          f = which == 0 ? phase(sl + sr + tj + 2 * so) : phase(sl + sr + 1 + so);
          f *= std::sqrt(0.5 * (sld + 1) * (tj + 1)) * wigner6j(tj, so, 1, sld, sl, sr);
          k = kap(nl + 1, sld, ild, 0, so, nr, sr, irC);
          if (k >= 0) term(ik, k, f, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', C->blk(nr, sr, ir, nr, sr, irC), true);
          // 4K4B / 4L4B: C, local 1 -> 2, left N-1
          f = which == 0 ? phase(sld + sr + 1 + tj + 2 * so) : phase(sld + sr + 2 + so);
          f *= std::sqrt(0.5 * (sl + 1) * (tj + 1)) * wigner6j(tj, so, 1, sld, sl, sr);
          k = kap(nl - 1, sld, ild, 2, so, nr, sr, irC);
          if (k >= 0) term(ik, k, f, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', C->blk(nr, sr, irC, nr, sr, ir), true);
        }
        // ---- spin 1: B (right N-2 / N+2) and D (right N)
        for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
          if (srd < 0) continue;
          const int jstart = (sld != srd || so == 0) ? so + 1 : 0;
          for (int jd = jstart; jd <= so + 1; jd += 2) {
            if (std::abs(sld - srd) > jd) continue;
            if (nloc == 2) {
              // 4K1B / 4L1B spin 1: B down, left N-1
              double f = which == 0 ? phase(sr - srd + sld - sl + 3 + 2 * so) : phase(sr - srd + sld - sl + so - jd);
              f *= std::sqrt(3.0 * (sr + 1) * (sl + 1) * (jd + 1)) * wigner9j(2, srd, sr, 1, sld, sl, 1, jd, so);
              int k = kap(nl - 1, sld, ild, 1, jd, nr - 2, srd, irB);
              if (k >= 0) term(ik, k, f, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', B->blk(nr - 2, srd, irB, nr, sr, ir), true);
              // 4K3B / 4L3B spin 1: D, left N+1
              double fd = which == 0 ? phase(sr - srd + 2 * so) : phase(sr - srd + 1 + so - jd);
              fd *= std::sqrt(3.0 * (sr + 1) * (sld + 1) * (jd + 1)) * wigner9j(2, srd, sr, 1, sld, sl, 1, jd, so);
              k = kap(nl + 1, sld, ild, 1, jd, nr, srd, irD);
              if (k >= 0) term(ik, k, fd, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', D->blk(nr, sr, ir, nr, srd, irD), true);
            }
            if (nloc == 0) {
              // 4K2A / 4L2A spin 1: B up, left N+1
              double f = which == 0 ? (so == 1 ? -1 : 1) : phase(1 + so - jd);
              f *= std::sqrt(3.0 * (srd + 1) * (sld + 1) * (jd + 1)) * wigner9j(2, srd, sr, 1, sld, sl, 1, jd, so);
              int k = kap(nl + 1, sld, ild, 1, jd, nr + 2, srd, irB);
              if (k >= 0) term(ik, k, f, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', B->blk(nr, sr, ir, nr + 2, srd, irB), true);
              // 4K4A / 4L4A spin 1: D, left N-1
              double fd = which == 0 ? phase(sld + 1 - sl + 2 * so) : phase(sld + 2 - sl + so - jd);
              fd *= std::sqrt(3.0 * (srd + 1) * (sl + 1) * (jd + 1)) * wigner9j(2, srd, sr, 1, sld, sl, 1, jd, so);
              k = kap(nl - 1, sld, ild, 1, jd, nr, srd, irD);
              if (k >= 0) term(ik, k, fd, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', D->blk(nr, srd, irD, nr, sr, ir), true);
            }
          }
          if (nloc == 1 && std::abs(sld - srd) <= so) {
            // 4K1A / 4L1A spin 1: B down, local 1 -> 0, left N-1
            double f = which == 0 ? (so == 1 ? -1 : 1) : phase(1 + so - tj);
            f *= std::sqrt(3.0 * (sr + 1) * (sl + 1) * (tj + 1)) * wigner9j(2, sr, srd, 1, sl, sld, 1, tj, so);
            int k = kap(nl - 1, sld, ild, 0, so, nr - 2, srd, irB);
            if (k >= 0) term(ik, k, f, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', B->blk(nr - 2, srd, irB, nr, sr, ir), true);
            // 4K2B / 4L2B spin 1: B up, local 1 -> 2, left N+1
            f = which == 0 ? phase(srd - sr + sl - sld + 3 + 2 * so) : phase(srd - sr + sl - sld + so - tj);
            f *= std::sqrt(3.0 * (srd + 1) * (sld + 1) * (tj + 1)) * wigner9j(2, sr, srd, 1, sl, sld, 1, tj, so);
            k = kap(nl + 1, sld, ild, 2, so, nr + 2, srd, irB);
            if (k >= 0) term(ik, k, f, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', B->blk(nr, sr, ir, nr + 2, srd, irB), true);
            // 4K3A / 4L3A spin 1: D, local 1 -> 0, left N+1
            f = which == 0 ? phase(sl - sld + 1 + 2 * so) : phase(sl - sld + 2 + so - tj);
            f *= std::sqrt(3.0 * (sr + 1) * (sld + 1) * (tj + 1)) * wigner9j(2, sr, srd, 1, sl, sld, 1, tj, so);
            k = kap(nl + 1, sld, ild, 0, so, nr, srd, irD);
            if (k >= 0) term(ik, k, f, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', D->blk(nr, sr, ir, nr, srd, irD), true);
            // 4K4B / 4L4B spin 1: D, local 1 -> 2, left N-1
            f = which == 0 ? phase(sr - srd + 2 * so) : phase(sr - srd + 1 + so - tj);
            f *= std::sqrt(3.0 * (srd + 1) * (sl + 1) * (tj + 1)) * wigner9j(2, sr, srd, 1, sl, sld, 1, tj, so);
            k = kap(nl - 1, sld, ild, 2, so, nr, srd, irD);
            if (k >= 0) term(ik, k, f, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', D->blk(nr, srd, irD, nr, sr, ir), true);
          }
        }
      }
    }
  }
}


// ---------------------------------------------------------------- group 5 --
// 5A-5F: L_a at the left boundary (a left of the sites) times sum_b V L_b at the
// right boundary (b right of the sites), with two local operators on sites i and
// i+1; irrep(a) x irrep(b) = irrep(i) x irrep(i+1) (HeffDiagrams5.cpp:28-1959).
void HeffPlanner::d5(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  const int mid = bk->irrep(i) ^ bk->irrep(j);
  // One contribution: left operator block moves NL by dnl (-1: transposed L_a block
  // (NL-1 -> NL), +1: L_a block (NL -> NL+1)); right combination moves NR by dnr
  // (-1: block (NR-1 -> NR) as stored, +1: block (NR -> NR+1) transposed).
  auto add = [&](int dnl, int sld, int dnr, int srd, int a1, int a2, int aj, auto&& coef) {
    if (sld < 0 || srd < 0) return;
    for (int irrep = 0; irrep < bk->nirrep; irrep++) {
      const int irm = irrep ^ mid;
      std::vector<int> la, lb;
      for (int a = 0; a < index; a++) if (bk->irrep(a) == irrep) la.push_back(a);
      for (int b = index + 2; b < L; b++) if (bk->irrep(b) == irm) lb.push_back(b);
      if (la.empty() || lb.empty()) continue;
      const int ild = il ^ irrep, ird = ir ^ irm;
      if (bk->dim(index, nl + dnl, sld, ild) <= 0 || bk->dim(index + 2, nr + dnr, srd, ird) <= 0) continue;
      const int k = S->kappa(nl + dnl, sld, ild, a1, a2, aj, nr + dnr, srd, ird);
      if (k < 0) continue;
      std::vector<OpT*> Rops;
      for (int b : lb) Rops.push_back(Ro->L[b - index - 2]);
      std::vector<double> co(lb.size());
      const int rsz = bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + dnr, srd, ird);
      for (int a : la) {
        for (size_t q = 0; q < lb.size(); q++) co[q] = coef(a, lb[q]);
        if (dnr < 0) w2 = lin_comb(Rops, co, nr - 1, srd, ird, nr, sr, ir, rsz);
        else w2 = lin_comb(Rops, co, nr, sr, ir, nr + 1, srd, ird, rsz);
        OpT* La = Lo->L[index - 1 - a];
        Ref lbk = dnl < 0 ? La->blk(nl - 1, sld, ild, nl, sl, il) : La->blk(nl, sl, il, nl + 1, sld, ild);
        term(ik, k, 1.0, dnl < 0 ? 'T' : 'N', lbk, dnr < 0 ? 'N' : 'T', w2, true);
      }
    }
  };
  const double sgnJ = (tj == 0) ? 1.0 : -1.0;
  // ---- 5A: left N-1, right N+1
  if (n1 == 0 && n2 == 0)  // 5A1
    for (int sld = sl - 1; sld <= sl + 1; sld += 2)
      for (int srd = sr - 1; srd <= sr + 1; srd += 2)
        for (int jd = 0; jd <= 2; jd += 2) {
          if (std::abs(sld - srd) > jd || sld < 0 || srd < 0) continue;
          const double f = phase(sld + srd + 2) * std::sqrt((jd + 1) * (srd + 1.0)) * wigner6j(1, 1, jd, sld, srd, sr);
          const double sj = (jd == 0) ? 1.0 : -1.0;
          add(-1, sld, +1, srd, 1, 1, jd, [&](int a, int b) { return f * (P->g(a, b, i, j) + sj * P->g(a, b, j, i)); });
        }
  if ((n1 == 1 && n2 == 0) || (n1 == 0 && n2 == 1))  // 5A2, 5A3
    for (int sld = sl - 1; sld <= sl + 1; sld += 2)
      for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
        if (!((sl == srd) || (sld == sr)) || std::abs(sld - srd) > 1 || sld < 0 || srd < 0) continue;
        const double fa = (sl == srd) ? phase(srd + 1 - sr) : 0.0;
        const double fb = (sld == sr) ? phase(sld + 1 - sl) * std::sqrt((srd + 1) * (sl + 1.0) / ((sr + 1) * (sld + 1.0))) : 0.0;
        const double f1 = (n1 == 1) ? fb : fa, f2 = (n1 == 1) ? fa : fb;
        add(-1, sld, +1, srd, n1 + 1, n2 + 1, 1, [&](int a, int b) { return f1 * P->g(a, b, i, j) + f2 * P->g(a, b, j, i); });
      }
  if (n1 == 1 && n2 == 1)  // 5A4
    for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
      if (sld < 0 || std::abs(sr - sld) > 1) continue;
      const int srd = sld;
      const double f = (((sl + 1) % 2) ? -1.0 : 1.0) * std::sqrt((tj + 1) * (sl + 1.0)) * wigner6j(1, 1, tj, sl, sr, srd);
      add(-1, sld, +1, srd, 2, 2, 0, [&](int a, int b) { return f * (P->g(a, b, i, j) + sgnJ * P->g(a, b, j, i)); });
    }
  // ---- 5B: left N+1, right N-1
  if (n1 == 1 && n2 == 1)  // 5B1
    for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
      if (sld < 0 || std::abs(sr - sld) > 1) continue;
      const int srd = sld;
      const double f = phase(sl + sr + 2) * std::sqrt((tj + 1) * (sr + 1.0)) * wigner6j(1, 1, tj, sl, sr, srd);
      add(+1, sld, -1, srd, 0, 0, 0, [&](int a, int b) { return f * (P->g(a, b, i, j) + sgnJ * P->g(a, b, j, i)); });
    }
  if ((n1 == 2 && n2 == 1) || (n1 == 1 && n2 == 2))  // 5B2, 5B3
    for (int sld = sl - 1; sld <= sl + 1; sld += 2)
      for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
        if (!((sl == srd) || (sld == sr)) || std::abs(sld - srd) > 1 || sld < 0 || srd < 0) continue;
        const double fa = (sld == sr) ? phase(sr + 1 - srd) : 0.0;
        const double fb = (sl == srd) ? phase(sl + 1 - sld) * std::sqrt((sr + 1) * (sld + 1.0) / ((srd + 1) * (sl + 1.0))) : 0.0;
        const double f1 = (n1 == 2) ? fb : fa, f2 = (n1 == 2) ? fa : fb;
        add(+1, sld, -1, srd, n1 - 1, n2 - 1, 1, [&](int a, int b) { return f1 * P->g(a, b, i, j) + f2 * P->g(a, b, j, i); });
      }
  if (n1 == 2 && n2 == 2)  // 5B4
    for (int sld = sl - 1; sld <= sl + 1; sld += 2)
      for (int srd = sr - 1; srd <= sr + 1; srd += 2)
        for (int jd = 0; jd <= 2; jd += 2) {
          if (std::abs(sld - srd) > jd || sld < 0 || srd < 0) continue;
          const double f = (((sld + 1) % 2) ? -1.0 : 1.0) * std::sqrt((jd + 1) * (sld + 1.0)) * wigner6j(1, 1, jd, sld, srd, sr);
          const double sj = (jd == 0) ? 1.0 : -1.0;
          add(+1, sld, -1, srd, 1, 1, jd, [&](int a, int b) { return f * (P->g(a, b, i, j) + sj * P->g(a, b, j, i)); });
        }
  // ---- 5C (annihilate on i, create on i+1; left N-1, right N-1) and 5E (mirror, i <-> i+1)
  for (int which = 0; which < 2; which++) {
    const int p = which == 0 ? i : j, q = which == 0 ? j : i;
    const int np = which == 0 ? n1 : n2, nq = which == 0 ? n2 : n1;  // occupation of p and q
    auto g1 = [&](int a, int b) { return P->g(a, p, q, b); };
    auto g2 = [&](int a, int b) { return P->g(a, p, b, q); };
    auto kadd = [&](int sld, int srd, int ap, int aq, int aj, double f1, double f2) {
      const int a1 = which == 0 ? ap : aq, a2 = which == 0 ? aq : ap;
      add(-1, sld, -1, srd, a1, a2, aj, [&](int a, int b) { return f1 * g1(a, b) + f2 * g2(a, b); });
    };
    if (np == 1 && nq == 0)  // 5C1 / 5E1
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
          if (std::abs(sld - srd) > 1 || sld < 0 || srd < 0) continue;
          const double f2 = phase(sl + srd) * std::sqrt((sl + 1) * (sr + 1.0)) * wigner6j(sld, srd, 1, sr, sl, 1);
          const double f1 = (sl == srd) ? std::sqrt((sr + 1.0) / (srd + 1.0)) : 0.0;
          kadd(sld, srd, 0, 1, 1, f1, f2);
        }
    if (np == 2 && nq == 0)  // 5C2 / 5E2
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2)
          for (int jd = 0; jd <= 2; jd += 2) {
            if (std::abs(sld - srd) > jd || sld < 0 || srd < 0) continue;
            const double f1 = phase(sr + sld + 3 + (which == 0 ? jd : 0)) * std::sqrt((sr + 1) * (jd + 1.0)) * wigner6j(1, 1, jd, sld, srd, sr);
            const double f2 = (jd == 0) ? std::sqrt(2.0 * (sr + 1.0) / (srd + 1.0)) : 0.0;
            kadd(sld, srd, 1, 1, jd, f1, f2);
          }
    if (np == 1 && nq == 1)  // 5C3 / 5E3
      for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
        if (sld < 0 || std::abs(sr - sld) > 1) continue;
        const int srd = sld;
        const double f1 = phase(sr + srd + 3 + (which == 0 ? tj : 0)) * std::sqrt((sl + 1.0) * (tj + 1.0) * (sr + 1.0) / (srd + 1.0)) *
                          wigner6j(1, 1, tj, sl, sr, srd);
        const double f2 = (tj == 0) ? std::sqrt(2.0 * (sr + 1.0) / (srd + 1.0)) : 0.0;
        kadd(sld, srd, 0, 2, 0, f1, f2);
      }
    if (np == 2 && nq == 1)  // 5C4 / 5E4
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
          if (std::abs(sld - srd) > 1 || sld < 0 || srd < 0) continue;
          const double f1 = (sld == sr) ? phase(sl - srd) * std::sqrt((sl + 1.0) / (sld + 1.0)) : 0.0;
          const double f2 = phase(sl + srd + 2) * std::sqrt((sl + 1) * (sr + 1.0)) * wigner6j(sld, srd, 1, sr, sl, 1);
          kadd(sld, srd, 1, 2, 1, f1, f2);
        }
  }
  // ---- 5D (create on i, annihilate on i+1; left N+1, right N+1) and 5F (mirror)
  for (int which = 0; which < 2; which++) {
    const int p = which == 0 ? i : j, q = which == 0 ? j : i;
    const int np = which == 0 ? n1 : n2, nq = which == 0 ? n2 : n1;
    auto g1 = [&](int a, int b) { return P->g(a, p, q, b); };
    auto g2 = [&](int a, int b) { return P->g(a, p, b, q); };
    auto kadd = [&](int sld, int srd, int ap, int aq, int aj, double f1, double f2) {
      const int a1 = which == 0 ? ap : aq, a2 = which == 0 ? aq : ap;
      add(+1, sld, +1, srd, a1, a2, aj, [&](int a, int b) { return f1 * g1(a, b) + f2 * g2(a, b); });
    };
    if (np == 0 && nq == 1)  // 5D1 / 5F1
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
          if (std::abs(sld - srd) > 1 || sld < 0 || srd < 0) continue;
          const double f2 = phase(sld + sr) * std::sqrt((sld + 1) * (srd + 1.0)) * wigner6j(sl, sr, 1, srd, sld, 1);
          const double f1 = (sld == sr) ? std::sqrt((srd + 1.0) / (sr + 1.0)) : 0.0;
          kadd(sld, srd, 1, 0, 1, f1, f2);
        }
    if (np == 1 && nq == 1)  // 5D2 / 5F2
      for (int sld = sl - 1; sld <= sl + 1; sld += 2) {
        if (sld < 0 || std::abs(sr - sld) > 1) continue;
        const int srd = sld;
        const double f1 = phase(srd + sl + 3 + (which == 0 ? tj : 0)) * std::sqrt((srd + 1) * (tj + 1.0)) * wigner6j(1, 1, tj, sl, sr, srd);
        const double f2 = (tj == 0) ? std::sqrt(2.0 * (srd + 1.0) / (sr + 1.0)) : 0.0;
        kadd(sld, srd, 2, 0, 0, f1, f2);
      }
    if (np == 0 && nq == 2)  // 5D3 / 5F3
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2)
          for (int jd = 0; jd <= 2; jd += 2) {
            if (std::abs(sld - srd) > jd || sld < 0 || srd < 0) continue;
            const double f1 = phase(sr + srd + 3 + (which == 0 ? jd : 0)) * std::sqrt((sld + 1.0) * (jd + 1.0) * (srd + 1.0) / (sr + 1.0)) *
                              wigner6j(1, 1, jd, sld, srd, sr);
            const double f2 = (jd == 0) ? std::sqrt(2.0 * (srd + 1.0) / (sr + 1.0)) : 0.0;
            kadd(sld, srd, 1, 1, jd, f1, f2);
          }
    if (np == 1 && nq == 2)  // 5D4 / 5F4
      for (int sld = sl - 1; sld <= sl + 1; sld += 2)
        for (int srd = sr - 1; srd <= sr + 1; srd += 2) {
          if (std::abs(sld - srd) > 1 || sld < 0 || srd < 0) continue;
          const double f1 = (sl == srd) ? phase(sld - sr) * std::sqrt((sld + 1.0) / (sl + 1.0)) : 0.0;
          const double f2 = phase(sld + sr + 2) * std::sqrt((sld + 1) * (srd + 1.0)) * wigner6j(sl, sr, 1, srd, sld, 1);
          kadd(sld, srd, 2, 1, 1, f1, f2);
        }
  }
}


// ----------------------------------------------------------- diagonal ------
// Diagonal of Heff (Heff.cpp:250-315, HeffDiagonal.cpp:27-620): the sector-preserving
// terms, each through the diagonals of its operator blocks.
void HeffPlanner::diag_block(int ik) {  // HeffDiagonal.cpp:27-620 for block ik
  const int i = index, j = index + 1;
  const double sqh = std::sqrt(0.5);
  const bool leftSum = index < L * 0.5;
  auto need = [](Ref r) { if (!r.ok()) throw Error{ERR_MISSING, "missing block for the diagonal"}; return r; };
  {
    const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
    const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
    auto add_const = [&](double c) { diag.push_back({ik, 0, c, Ref(), Ref()}); };
    auto add_left = [&](double c, Ref b) { diag.push_back({ik, 1, c, need(b), Ref()}); };
    auto add_right = [&](double c, Ref b) { diag.push_back({ik, 2, c, Ref(), need(b)}); };
    auto add_outer = [&](double c, Ref lb, Ref rb) { diag.push_back({ik, 3, c, need(lb), need(rb)}); };
    double c0 = 0.0;  // 1C, 1D, 2d3 (HeffDiagonal.cpp:56-124)
    if (n1 == 2) c0 += P->g(i, i, i, i);
    if (n2 == 2) c0 += P->g(j, j, j, j);
    if (n1 == 2 && n2 == 2) c0 += 4 * P->g(i, j, i, j) - 2 * P->g(i, j, j, i);
    if (n1 == 1 && n2 == 1) c0 += P->g(i, j, i, j) + (tj == 2 ? -1 : 1) * P->g(i, j, j, i);
    if ((n1 == 2 && n2 == 1) || (n1 == 1 && n2 == 2)) c0 += 2 * P->g(i, j, i, j) - P->g(i, j, j, i);
    add_const(c0);
    if (!atLeft) {
      add_left(1.0, Lo->X->blk(nl, sl, il, nl, sl, il));  // 1A
      for (int w = 0; w < 2; w++) {  // 2b3 / 2c3 (HeffDiagonal.cpp:126-328)
        const int nw = w == 0 ? n1 : n2, no = w == 0 ? n2 : n1;
        if (nw != 0) add_left(nw * sqh, Lo->C[0][w]->blk(nl, sl, il, nl, sl, il));
        if (nw == 1) {
          const int so = (no == 1) ? 1 : 0;
          const double a = phase(sl + sr + 2 * tj + so - 1) * (tj + 1) * std::sqrt(3.0 * (sl + 1)) * wigner6j(tj, tj, 2, 1, 1, so) *
                           wigner6j(tj, tj, 2, sl, sl, sr);
          add_left(a, Lo->D[0][w]->blk(nl, sl, il, nl, sl, il));
        }
      }
    }
    if (!atRight) {
      add_right(1.0, Ro->X->blk(nr, sr, ir, nr, sr, ir));  // 1B
      for (int w = 0; w < 2; w++) {  // 2e3 / 2f3: pair [0][1] is site i, [0][0] site i+1
        const int nw = w == 0 ? n1 : n2, no = w == 0 ? n2 : n1;
        const int slot = w == 0 ? 1 : 0;
        if (nw != 0) add_right(nw * sqh, Ro->C[0][slot]->blk(nr, sr, ir, nr, sr, ir));
        if (nw == 1) {
          const int so = (no == 1) ? 1 : 0;
          const double a = phase(sr + sl + 2 * tj + so + 1) * (tj + 1) * std::sqrt(3.0 * (sr + 1)) * wigner6j(tj, tj, 2, 1, 1, so) *
                           wigner6j(tj, tj, 2, sr, sr, sl);
          add_right(a, Ro->D[0][slot]->blk(nr, sr, ir, nr, sr, ir));
        }
      }
    }
    if (!atLeft && !atRight) {  // 2a3 spin 0 and 1 (HeffDiagonal.cpp:380-620)
      const double a1 = phase(sl + sr + tj + 2) * std::sqrt((sr + 1) * (sl + 1.0)) * wigner6j(sl, sr, tj, sr, sl, 2);
      if (leftSum) {
        for (int a = 0; a < index; a++)
          for (int b = a; b < index; b++) {
            if (bk->irrep(a) != bk->irrep(b)) continue;
            const double mult = (a == b) ? 1.0 : 2.0;
            add_outer(mult, pl(Lo->F0, a, b)->blk(nl, sl, il, nl, sl, il), cr(Ro->C, a, b)->blk(nr, sr, ir, nr, sr, ir));
            if (a1 != 0.0)
              add_outer(mult * a1, pl(Lo->F1, a, b)->blk(nl, sl, il, nl, sl, il), cr(Ro->D, a, b)->blk(nr, sr, ir, nr, sr, ir));
          }
      } else {
        for (int a = index + 2; a < L; a++)
          for (int b = a; b < L; b++) {
            if (bk->irrep(a) != bk->irrep(b)) continue;
            const double mult = (a == b) ? 1.0 : 2.0;
            add_outer(mult, cl(Lo->C, a, b)->blk(nl, sl, il, nl, sl, il), pr(Ro->F0, a, b)->blk(nr, sr, ir, nr, sr, ir));
            if (a1 != 0.0)
              add_outer(mult * a1, cl(Lo->D, a, b)->blk(nl, sl, il, nl, sl, il), pr(Ro->F1, a, b)->blk(nr, sr, ir, nr, sr, ir));
          }
      }
    }
  }
}

}  // namespace chm
