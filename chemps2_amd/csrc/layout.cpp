// Sector bookkeeping and block layouts (host side of the chemps2_amd library).
#include <algorithm>
#include <cstdlib>

#include "chm_internal.hpp"

namespace chm {

// ------------------------------------------------------------- Wigner -----
struct FactTable {  // filled at static initialisation, before any concurrent reader
  double v[200];
  FactTable() {
    v[0] = 1.0;
    for (int i = 1; i < 200; i++) v[i] = v[i - 1] * i;
  }
};
static const FactTable g_fact_table;
static const double* const g_fact = g_fact_table.v;
static void fact_init() {}
static bool tri_bad(int a, int b, int c) { return ((a + b + c) % 2) != 0 || c < std::abs(a - b) || c > a + b; }
static double tri_coef(int a, int b, int c) {  // Delta(abc) with doubled arguments
  return std::sqrt(g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] * g_fact[(-a + b + c) / 2] / g_fact[(a + b + c) / 2 + 1]);
}

// Racah's single-sum formula for {a b c; d e f} (doubled arguments), the quantity
// Wigner::wigner6j evaluates (Wigner.cpp:294-339).
static double wigner6j_eval(int a, int b, int c, int d, int e, int f);

// Direct-mapped per-thread cache in front of the evaluation: the work-list generators
// ask for the same few thousand symbols tens of millions of times.
double wigner6j(int a, int b, int c, int d, int e, int f) {
  if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0 || f < 0) return 0.0;
  if ((a | b | c | d | e | f) >= 512) return wigner6j_eval(a, b, c, d, e, f);
  struct Cache {
    uint64_t key[1 << 14];
    double val[1 << 14];
    Cache() { for (auto& k : key) k = ~0ULL; }
  };
  thread_local Cache cache;
  const uint64_t key = (uint64_t)a | ((uint64_t)b << 9) | ((uint64_t)c << 18) | ((uint64_t)d << 27) | ((uint64_t)e << 36) |
                       ((uint64_t)f << 45);
  uint64_t h = key * 0x9E3779B97F4A7C15ULL;
  const size_t slot = (size_t)(h >> 50);
  if (cache.key[slot] == key) return cache.val[slot];
  const double v = wigner6j_eval(a, b, c, d, e, f);
  cache.key[slot] = key;
  cache.val[slot] = v;
  return v;
}

static double wigner6j_eval(int a, int b, int c, int d, int e, int f) {
  fact_init();
  if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0 || f < 0) return 0.0;
  if (tri_bad(a, b, c) || tri_bad(d, e, c) || tri_bad(a, e, f) || tri_bad(d, b, f)) return 0.0;
  const int t1 = (a + b + c) / 2, t2 = (d + e + c) / 2, t3 = (a + e + f) / 2, t4 = (d + b + f) / 2;
  const int u1 = (a + b + d + e) / 2, u2 = (a + c + d + f) / 2, u3 = (b + c + e + f) / 2;
  const int lo = std::max(std::max(t1, t2), std::max(t3, t4)), hi = std::min(u1, std::min(u2, u3));
  double sum = 0.0;
  for (int z = lo; z <= hi; z++) {
    const double term = g_fact[z + 1] / (g_fact[z - t1] * g_fact[z - t2] * g_fact[z - t3] * g_fact[z - t4] *
                                         g_fact[u1 - z] * g_fact[u2 - z] * g_fact[u3 - z]);
    sum += (z % 2) ? -term : term;
  }
  return sum * tri_coef(a, b, c) * tri_coef(d, e, c) * tri_coef(a, e, f) * tri_coef(d, b, f);
}

// {a b c; d e f; g h i} = sum_x (-1)^{2x} (2x+1) {a b c; f i x}{d e f; b x h}{g h i; x a d}.
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  if (tri_bad(a, b, c) || tri_bad(d, e, f) || tri_bad(g, h, i) || tri_bad(a, d, g) || tri_bad(b, e, h) || tri_bad(c, f, i))
    return 0.0;
  const int lo = std::max(std::abs(a - i), std::max(std::abs(d - h), std::abs(b - f)));
  const int hi = std::min(a + i, std::min(d + h, b + f));
  double v = 0.0;
  for (int x = lo; x <= hi; x += 2)
    v += ((x % 2) ? -1.0 : 1.0) * (x + 1) * wigner6j(a, b, c, f, i, x) * wigner6j(d, e, f, b, x, h) * wigner6j(g, h, i, x, a, d);
  return v;
}

// ------------------------------------------------------------ Sectors ------
void Sectors::set(int b, std::vector<Sec> v) {
  std::stable_sort(v.begin(), v.end(), [](const Sec& x, const Sec& y) {
    return x.N != y.N ? x.N < y.N : (x.S != y.S ? x.S < y.S : x.I < y.I);
  });
  std::vector<Sec> keep;
  auto ix = std::make_shared<SecIndex>();
  for (const Sec& q : v)
    if (q.d > 0) {
      keep.push_back(q);
      ix->nmax = std::max(ix->nmax, q.N);
      ix->smax = std::max(ix->smax, q.S);
    }
  if (!keep.empty()) ix->pos.assign((size_t)(ix->nmax + 1) * (ix->smax + 1) * 8, -1);
  for (size_t p = 0; p < keep.size(); p++) ix->pos[((size_t)keep[p].N * (ix->smax + 1) + keep[p].S) * 8 + keep[p].I] = (int32_t)p;
  list[b] = std::move(keep);
  idx[b] = std::move(ix);
  gen++;
}

// ---------------------------------------------------------------- OpT ------
void OpT::build(const Sectors& sec, int b, int tj, int ne, int nirr) {  // TensorOperator.cpp:29-102
  boundary = b; index = b; two_j = tj; n_elec = ne; n_irrep = nirr;
  bi = sec.idx[b];
  offs.assign(sec.list[b].size() * (tj + 1), -1);
  Nup.clear(); Sup.clear(); Iup.clear(); Sdn.clear();
  k2i.assign(1, 0);
  int64_t pos = 0;
  for (size_t p = 0; p < sec.list[b].size(); p++) {
    const auto& up = sec.list[b][p];
    const int nd = up.N + ne, id = up.I ^ nirr;
    for (int sd = up.S - tj; sd <= up.S + tj; sd += 2) {
      if (sd < 0) continue;
      const int dd = sec.dim(b, nd, sd, id);
      if (dd <= 0) continue;
      offs[p * (tj + 1) + (sd - up.S + tj) / 2] = pos;
      Nup.push_back(up.N); Sup.push_back(up.S); Iup.push_back(up.I); Sdn.push_back(sd);
      pos += (int64_t)up.d * dd;
      k2i.push_back(pos);
    }
  }
  size = pos;
}

// ---------------------------------------------------------------- TLay -----
static uint64_t key6(int nl, int sl, int il, int nr, int sr, int ir) {
  return ((uint64_t)(nl & 0xff)) | ((uint64_t)(sl & 0xff) << 8) | ((uint64_t)(il & 0x7) << 16) | ((uint64_t)(nr & 0xff) << 19) |
         ((uint64_t)(sr & 0xff) << 27) | ((uint64_t)(ir & 0x7) << 35);
}

void TLay::build(const Sectors& sec, int site, int Is) {  // TensorT.cpp:38-104
  index = site;
  NL.clear(); SL.clear(); IL.clear(); NR.clear(); SR.clear(); IR.clear(); lut.clear();
  std::vector<std::pair<uint64_t, int>> kv;
  k2i.assign(1, 0);
  for (const auto& l : sec.list[site]) {
    for (int nr = l.N; nr <= l.N + 2; nr++) {
      const int tj = (nr == l.N + 1) ? 1 : 0;
      for (int sr = l.S - tj; sr <= l.S + tj; sr += 2) {
        if (sr < 0) continue;
        const int ir = (nr == l.N + 1) ? (l.I ^ Is) : l.I;
        const int dr = sec.dim(site + 1, nr, sr, ir);
        if (dr <= 0) continue;
        kv.push_back({key6(l.N, l.S, l.I, nr, sr, ir), (int)NL.size()});
        NL.push_back(l.N); SL.push_back(l.S); IL.push_back(l.I);
        NR.push_back(nr); SR.push_back(sr); IR.push_back(ir);
        k2i.push_back(k2i.back() + (int64_t)l.d * dr);
      }
    }
  }
  size = k2i.back();
  lut.build(kv);
}

int TLay::kappa(int nl, int sl, int il, int nr, int sr, int ir) const {
  if (nl < 0 || sl < 0 || nr < 0 || sr < 0) return -1;
  return lut.get(key6(nl, sl, il, nr, sr, ir));
}

Ref TLay::blk(int nl, int sl, int il, int nr, int sr, int ir) const {
  Ref r;
  const int k = kappa(nl, sl, il, nr, sr, ir);
  if (k < 0) return r;
  r.buf = buf;
  r.off = k2i[k];
  return r;
}

Ref OpT::blk(int n1, int s1, int i1, int n2, int s2, int i2) const {
  Ref r;
  if ((i1 ^ n_irrep) != i2 || n2 != n1 + n_elec || std::abs(s1 - s2) > two_j || ((s2 - s1 + two_j) & 1)) return r;
  const int p = bi ? bi->find(n1, s1, i1) : -1;
  if (p < 0) return r;
  const int64_t o = offs[(size_t)p * (two_j + 1) + (s2 - s1 + two_j) / 2];
  if (o < 0) return r;
  r.buf = buf;
  r.off = o;
  return r;
}

// ------------------------------------------------------------- SLayout -----
static uint64_t key9(int nl, int sl, int il, int n1, int n2, int tj, int nr, int sr, int ir) {
  return ((uint64_t)(nl & 0xff)) | ((uint64_t)(sl & 0xff) << 8) | ((uint64_t)(il & 0x7) << 16) | ((uint64_t)(n1 & 3) << 19) |
         ((uint64_t)(n2 & 3) << 21) | ((uint64_t)(tj & 3) << 23) | ((uint64_t)(nr & 0xff) << 25) |
         ((uint64_t)(sr & 0xff) << 33) | ((uint64_t)(ir & 0x7) << 41);
}

void SLayout::build(const Sectors& sec, int idx, int I1, int I2) {  // Sobject.cpp:36-125
  index = idx;
  NL.clear(); SL.clear(); IL.clear(); N1.clear(); N2.clear(); TJ.clear(); NR.clear(); SR.clear(); IR.clear(); lut.clear();
  dL.clear(); dR.clear();
  std::vector<std::pair<uint64_t, int>> kv;
  k2i.assign(1, 0);
  for (const auto& l : sec.list[idx]) {
    for (int n1 = 0; n1 <= 2; n1++)
      for (int n2 = 0; n2 <= 2; n2++) {
        const int nr = l.N + n1 + n2;
        const int im = (n1 == 1) ? (l.I ^ I1) : l.I;
        const int ir = (n2 == 1) ? (im ^ I2) : im;
        const int tjmin = (n1 + n2) % 2, tjmax = (n1 == 1 && n2 == 1) ? 2 : tjmin;
        for (int tj = tjmin; tj <= tjmax; tj += 2)
          for (int sr = l.S - tj; sr <= l.S + tj; sr += 2) {
            if (sr < 0) continue;
            const int dr = sec.dim(idx + 2, nr, sr, ir);
            if (dr <= 0) continue;
            kv.push_back({key9(l.N, l.S, l.I, n1, n2, tj, nr, sr, ir), (int)NL.size()});
            NL.push_back(l.N); SL.push_back(l.S); IL.push_back(l.I); N1.push_back(n1); N2.push_back(n2);
            TJ.push_back(tj); NR.push_back(nr); SR.push_back(sr); IR.push_back(ir);
            dL.push_back(l.d); dR.push_back(dr);
            k2i.push_back(k2i.back() + (int64_t)l.d * dr);
          }
      }
  }
  lut.build(kv);
}

int SLayout::kappa(int nl, int sl, int il, int n1, int n2, int tj, int nr, int sr, int ir) const {
  if (n1 < 0 || n1 > 2 || n2 < 0 || n2 > 2 || tj < 0 || tj > 2 || sl < 0 || sr < 0 || nl < 0 || nr < 0) return -1;
  return lut.get(key9(nl, sl, il, n1, n2, tj, nr, sr, ir));
}

}  // namespace chm
