// CPU FP64 restatement of CheMPS2's data structures (oracle; test infrastructure only).
#include "core.hpp"

#include <algorithm>
#include <numeric>

namespace orc {

int num_irreps(int group) {  // Irreps.cpp:100-108
  if (group == 0) return 1;
  if (group <= 3) return 2;
  if (group <= 6) return 4;
  return 8;
}

// ------------------------------------------------------------------ 6j -----
// Factorial table, filled once before any thread can read it (static initialisation).
struct FactTable {
  long double v[256];
  FactTable() {
    v[0] = 1.0L;
    for (int i = 1; i < 256; i++) v[i] = v[i - 1] * (long double)i;
  }
};
static const FactTable g_fact_table;
static const long double* const g_fact = g_fact_table.v;
static void fact_init() {}
static bool tri_fail(int a, int b, int c) {
  if ((a + b + c) % 2 != 0) return true;
  if (c > a + b || c < std::abs(a - b)) return true;
  return false;
}
static long double delta(int a, int b, int c) {  // triangle coefficient, arguments doubled
  return std::sqrt(g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] * g_fact[(-a + b + c) / 2] /
                   g_fact[(a + b + c) / 2 + 1]);
}
double wigner6j(int a, int b, int c, int d, int e, int f) {
  fact_init();
  if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0 || f < 0) return 0.0;
  if (tri_fail(a, b, c) || tri_fail(d, e, c) || tri_fail(a, e, f) || tri_fail(d, b, f)) return 0.0;
  const int a1 = (a + b + c) / 2, a2 = (d + e + c) / 2, a3 = (a + e + f) / 2, a4 = (d + b + f) / 2;
  const int b1 = (a + b + d + e) / 2, b2 = (a + c + d + f) / 2, b3 = (b + c + e + f) / 2;
  const int tmin = std::max(std::max(a1, a2), std::max(a3, a4));
  const int tmax = std::min(b1, std::min(b2, b3));
  if (tmax < tmin) return 0.0;
  long double sum = 0.0L;
  for (int t = tmin; t <= tmax; t++) {
    long double term = g_fact[t + 1] / (g_fact[t - a1] * g_fact[t - a2] * g_fact[t - a3] * g_fact[t - a4] *
                                        g_fact[b1 - t] * g_fact[b2 - t] * g_fact[b3 - t]);
    sum += (t % 2 == 0) ? term : -term;
  }
  return (double)(sum * delta(a, b, c) * delta(d, e, c) * delta(a, e, f) * delta(d, b, f));
}

double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
  if (tri_fail(a, b, c) || tri_fail(d, e, f) || tri_fail(g, h, i) || tri_fail(a, d, g) || tri_fail(b, e, h) ||
      tri_fail(c, f, i))
    return 0.0;
  const int lo = std::max(std::abs(a - i), std::max(std::abs(d - h), std::abs(b - f)));
  const int hi = std::min(a + i, std::min(d + h, b + f));
  double v = 0.0;
  for (int x = lo; x <= hi; x += 2)
    v += ((x % 2) ? -1.0 : 1.0) * (x + 1) * wigner6j(a, b, c, f, i, x) * wigner6j(d, e, f, b, x, h) * wigner6j(g, h, i, x, a, d);
  return v;
}

// ---------------------------------------------------------------- dgemm ----
// With ORACLE_BLAS the products go to the host BLAS (the reference's own choice,
// Lapack.h:23-39) once they are large enough to gain from it; the loops below are the
// fallback and the path of small products.
#ifdef ORACLE_BLAS
extern "C" void dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* alpha,
                       const double* A, const int* lda, const double* B, const int* ldb, const double* beta, double* C,
                       const int* ldc);
#endif

void dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
           int ldb, double beta, double* C, int ldc) {
  if (m <= 0 || n <= 0) return;
#ifdef ORACLE_BLAS
  if (k > 0 && (double)m * n * k >= 4096.0) {
    dgemm_(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
    return;
  }
#endif
  for (int j = 0; j < n; j++) {
    double* c = C + (size_t)ldc * j;
    if (beta == 0.0) {
      for (int i = 0; i < m; i++) c[i] = 0.0;
    } else if (beta != 1.0) {
      for (int i = 0; i < m; i++) c[i] *= beta;
    }
  }
  if (k <= 0 || alpha == 0.0) return;
  const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  if (!tA) {
    for (int j = 0; j < n; j++) {
      double* c = C + (size_t)ldc * j;
      for (int l = 0; l < k; l++) {
        const double b = alpha * (tB ? B[j + (size_t)ldb * l] : B[l + (size_t)ldb * j]);
        if (b == 0.0) continue;
        const double* a = A + (size_t)lda * l;
        for (int i = 0; i < m; i++) c[i] += a[i] * b;
      }
    }
  } else {
    std::vector<double> bcol(k);
    for (int j = 0; j < n; j++) {
      for (int l = 0; l < k; l++) bcol[l] = tB ? B[j + (size_t)ldb * l] : B[l + (size_t)ldb * j];
      double* c = C + (size_t)ldc * j;
      for (int i = 0; i < m; i++) {
        const double* a = A + (size_t)lda * i;
        double s = 0.0;
        for (int l = 0; l < k; l++) s += a[l] * bcol[l];
        c[i] += alpha * s;
      }
    }
  }
}

// -------------------------------------------------------------- problem ----
void Problem::init(int L_, int group_, const int* irreps) {
  L = L_;
  group = group_;
  nirrep = num_irreps(group);
  orb_irrep.assign(irreps, irreps + L);
  T.assign((size_t)L * L, 0.0);
  V.assign((size_t)L * L * L * L, 0.0);
}
void Problem::setV(int i, int j, int k, int l, double v) {
  auto at = [&](int a, int b, int c, int d) -> double& { return V[a + L * (b + L * (c + (size_t)L * d))]; };
  at(i, j, k, l) = v; at(j, i, l, k) = v; at(k, j, i, l) = v; at(j, k, l, i) = v;
  at(i, l, k, j) = v; at(l, i, j, k) = v; at(k, l, i, j) = v; at(l, k, j, i) = v;
}
void Problem::build_mx() {  // Problem.cpp:363-384
  mx.assign((size_t)L * L * L * L, 0.0);
  const double pref = 1.0 / (N - 1);
  for (int a = 0; a < L; a++)
    for (int b = 0; b < L; b++)
      for (int c = 0; c < L; c++)
        for (int d = 0; d < L; d++) {
          double v = V[a + L * (b + L * (c + (size_t)L * d))];
          if (a == c) v += pref * T[b + L * d];
          if (b == d) v += pref * T[a + L * c];
          mx[a + L * (b + L * (c + (size_t)L * d))] = v;
        }
}

// ----------------------------------------------------------- bookkeeper ----
int Bookkeeper::idx(int b, int N, int S, int I) const {
  if (b < 0 || b > L) return -1;
  if (N < Nmin[b] || N > Nmax[b]) return -1;
  const int n = N - Nmin[b];
  if (S < 0 || (S % 2) != (Smin[b][n] % 2)) return -1;
  if (S < Smin[b][n] || S > Smax[b][n]) return -1;
  if (I < 0 || I >= nirrep) return -1;
  return base[b][n] + ((S - Smin[b][n]) / 2) * nirrep + I;
}

void Bookkeeper::init(const Problem* P_, int D) {  // SyBookkeeper.cpp:29-49, 91-130
  P = P_;
  L = P->L;
  nirrep = P->nirrep;
  const int N = P->N, TS = P->TwoS;
  Nmin.resize(L + 1); Nmax.resize(L + 1); Smin.resize(L + 1); Smax.resize(L + 1);
  base.resize(L + 1); fci.resize(L + 1); cur.resize(L + 1);
  for (int b = 0; b <= L; b++) {
    Nmin[b] = std::max(std::max(0, N + 2 * (b - L)), b - L + (N + TS) / 2);
    Nmax[b] = std::min(std::min(2 * b, N), b + (N - TS) / 2);
    int tot = 0;
    for (int n = Nmin[b]; n <= Nmax[b]; n++) {
      const int tmp = L - b - std::abs(N - n - L + b);
      const int smin = std::max(n % 2, TS - tmp);
      const int smax = std::min(b - std::abs(b - n), TS + tmp);
      Smin[b].push_back(smin);
      Smax[b].push_back(smax);
      base[b].push_back(tot);
      if (smax >= smin) tot += ((smax - smin) / 2 + 1) * nirrep;
    }
    fci[b].assign(tot, 0);
    cur[b].assign(tot, 0);
  }
  // FCI dimensions: left to right, then capped from right to left (SyBookkeeper.cpp:172-233).
  const int cutoff = 262144;  // Options.h:76 SYBK_dimensionCutoff
  fci[0][idx(0, 0, 0, 0)] = 1;
  for (int b = 1; b <= L; b++) {
    const int Ib = irrep(b - 1);
    for (int n = Nmin[b]; n <= Nmax[b]; n++)
      for (int s = Smin[b][n - Nmin[b]]; s <= Smax[b][n - Nmin[b]]; s += 2)
        for (int I = 0; I < nirrep; I++) {
          auto f = [&](int nn, int ss, int ii) { int i = idx(b - 1, nn, ss, ii); return i < 0 ? 0 : fci[b - 1][i]; };
          long v = (long)f(n, s, I) + f(n - 2, s, I) + f(n - 1, s + 1, I ^ Ib) + f(n - 1, s - 1, I ^ Ib);
          fci[b][idx(b, n, s, I)] = (int)std::min<long>(cutoff, v);
        }
  }
  {
    const int rhs = fci[L][idx(L, N, TS, P->Irrep)];
    for (int I = 0; I < nirrep; I++) fci[L][idx(L, N, TS, I)] = 0;
    fci[L][idx(L, N, TS, P->Irrep)] = std::min(1, rhs);
  }
  for (int b = L - 1; b >= 0; b--) {
    const int Ib = irrep(b);
    for (int n = Nmin[b]; n <= Nmax[b]; n++)
      for (int s = Smin[b][n - Nmin[b]]; s <= Smax[b][n - Nmin[b]]; s += 2)
        for (int I = 0; I < nirrep; I++) {
          auto f = [&](int nn, int ss, int ii) { int i = idx(b + 1, nn, ss, ii); return i < 0 ? 0 : fci[b + 1][i]; };
          long v = (long)f(n, s, I) + f(n + 2, s, I) + f(n + 1, s + 1, I ^ Ib) + f(n + 1, s - 1, I ^ Ib);
          int& d = fci[b][idx(b, n, s, I)];
          d = std::min<long>(d, std::min<long>(cutoff, v));
        }
  }
  cur = fci;
  scale(D, 1, L - 1);
  assert(dim(L, N, TS, P->Irrep) == 1);
}

void Bookkeeper::scale(int D, int start, int stop) {  // SyBookkeeper.cpp:250-269
  for (int b = start; b <= stop; b++) {
    const int tot = tot_dim(b);
    if (tot > D) {
      const double factor = (1.0 * D) / tot;
      for (size_t i = 0; i < cur[b].size(); i++)
        if (fci[b][i] > 0) cur[b][i] = (int)(std::ceil(factor * cur[b][i]) + 0.1);
    }
  }
}
int Bookkeeper::max_dim(int b) const { int m = 0; for (int v : cur[b]) m = std::max(m, v); return m; }
int Bookkeeper::tot_dim(int b) const { int t = 0; for (int v : cur[b]) t += v; return t; }

// -------------------------------------------------------------- TensorT ----
void TensorT::init(int site, const Bookkeeper* bk_) {  // TensorT.cpp:38-104
  index = site;
  bk = bk_;
  NL.clear(); SL.clear(); IL.clear(); NR.clear(); SR.clear(); IR.clear(); lut.clear();
  k2i.assign(1, 0);
  const int Is = bk->irrep(index);
  for (int nl = bk->Nmin[index]; nl <= bk->Nmax[index]; nl++)
    for (int sl = bk->Smin[index][nl - bk->Nmin[index]]; sl <= bk->Smax[index][nl - bk->Nmin[index]]; sl += 2)
      for (int il = 0; il < bk->nirrep; il++) {
        const int dl = bk->dim(index, nl, sl, il);
        if (dl <= 0) continue;
        for (int nr = nl; nr <= nl + 2; nr++) {
          const int tj = (nr == nl + 1) ? 1 : 0;
          for (int sr = sl - tj; sr <= sl + tj; sr += 2) {
            if (sr < 0) continue;
            const int ir = (nr == nl + 1) ? (il ^ Is) : il;
            const int dr = bk->dim(index + 1, nr, sr, ir);
            if (dr <= 0) continue;
            lut[key6(nl, sl, il, nr, sr, ir)] = (int)NL.size();
            NL.push_back(nl); SL.push_back(sl); IL.push_back(il);
            NR.push_back(nr); SR.push_back(sr); IR.push_back(ir);
            k2i.push_back(k2i.back() + dl * dr);
          }
        }
      }
  data.assign(k2i.back(), 0.0);
}

// ----------------------------------------------------------------- TOp -----
void TOp::init(int index_, int two_j_, int n_elec_, int n_irrep_, bool mr, bool pl, bool jw, const Bookkeeper* bk_) {
  // TensorOperator.cpp:29-102
  index = index_; two_j = two_j_; n_elec = n_elec_; n_irrep = n_irrep_;
  moving_right = mr; prime_last = pl; jw_phase = jw; bk = bk_;
  Nup.clear(); Sup.clear(); Iup.clear(); Sdn.clear(); lut.clear();
  k2i.assign(1, 0);
  for (int nu = bk->Nmin[index]; nu <= bk->Nmax[index]; nu++)
    for (int su = bk->Smin[index][nu - bk->Nmin[index]]; su <= bk->Smax[index][nu - bk->Nmin[index]]; su += 2)
      for (int iu = 0; iu < bk->nirrep; iu++) {
        const int du = bk->dim(index, nu, su, iu);
        if (du <= 0) continue;
        const int id = n_irrep ^ iu, nd = nu + n_elec;
        for (int sd = su - two_j; sd <= su + two_j; sd += 2) {
          if (sd < 0) continue;
          const int dd = bk->dim(index, nd, sd, id);
          if (dd <= 0) continue;
          lut[key6(nu, su, iu, 0, sd, 0)] = (int)Nup.size();
          Nup.push_back(nu); Sup.push_back(su); Iup.push_back(iu); Sdn.push_back(sd);
          k2i.push_back(k2i.back() + du * dd);
        }
      }
  data.assign(k2i.back(), 0.0);
}

int TOp::kappa(int n1, int s1, int i1, int n2, int s2, int i2) const {  // TensorOperator.cpp:119-137
  if ((i1 ^ n_irrep) != i2) return -1;
  if (n2 != n1 + n_elec) return -1;
  if (std::abs(s1 - s2) > two_j) return -1;
  auto it = lut.find(key6(n1, s1, i1, 0, s2, 0));
  return it == lut.end() ? -1 : it->second;
}

void TOp::update(TOp* prev, TensorT* Tup, TensorT* Tdn) {  // TensorOperator.cpp:163-405
  clear();
  const int jw3 = jw_phase ? 3 : 1;
  std::vector<double> work;
  for (int ik = 0; ik < nk(); ik++) {
    const int nu = Nup[ik], nd = nu + n_elec, su = Sup[ik], sd = Sdn[ik], iu = Iup[ik], id = iu ^ n_irrep;
    double* out = data.data() + k2i[ik];
    if (moving_right) {
      const int dRu = bk->dim(index, nu, su, iu), dRd = bk->dim(index, nd, sd, id);
      const int Is = bk->irrep(index - 1);
      for (int g = 0; g < 6; g++) {
        int sLu = su, sLd = sd, nLu = nu, nLd = nd, iLu = iu, iLd = id;
        if (g == 1) { nLu -= 2; nLd -= 2; }
        if (g >= 2) {
          nLu -= 1; nLd -= 1; iLu ^= Is; iLd ^= Is;
          sLu += (g == 2 || g == 3) ? -1 : 1;
          sLd += (g == 2 || g == 4) ? -1 : 1;
        }
        if (std::abs(sLu - sLd) > two_j) continue;
        const int dLu = bk->dim(index - 1, nLu, sLu, iLu), dLd = bk->dim(index - 1, nLd, sLd, iLd);
        if (dLu <= 0 || dLd <= 0) continue;
        double* bu = Tup->blk(nLu, sLu, iLu, nu, su, iu);
        double* bd = Tdn->blk(nLd, sLd, iLd, nd, sd, id);
        double* bl = prev->blk(nLu, sLu, iLu, nLd, sLd, iLd);
        double alpha = 1.0;
        if (g >= 2) {
          if (two_j == 0) alpha = jw_phase ? -1.0 : 1.0;
          else if (prime_last)
            alpha = phase(su + sLd + two_j + jw3) * std::sqrt((sLd + 1.0) * (su + 1.0)) * wigner6j(sLu, sLd, two_j, sd, su, 1);
          else
            alpha = phase(sd + sLu + two_j + jw3) * std::sqrt((sLu + 1.0) * (sd + 1.0)) * wigner6j(sLd, sLu, two_j, su, sd, 1);
        }
        work.resize((size_t)dRu * dLd);
        dgemm('T', 'N', dRu, dLd, dLu, alpha, bu, dLu, bl, dLu, 0.0, work.data(), dRu);
        dgemm('N', 'N', dRu, dRd, dLd, 1.0, work.data(), dRu, bd, dLd, 1.0, out, dRu);
      }
    } else {
      const int dLu = bk->dim(index, nu, su, iu), dLd = bk->dim(index, nd, sd, id);
      const int Is = bk->irrep(index);
      for (int g = 0; g < 6; g++) {
        int sRu = su, sRd = sd, nRu = nu, nRd = nd, iRu = iu, iRd = id;
        if (g == 1) { nRu += 2; nRd += 2; }
        if (g >= 2) {
          nRu += 1; nRd += 1; iRu ^= Is; iRd ^= Is;
          sRu += (g == 2 || g == 3) ? -1 : 1;
          sRd += (g == 2 || g == 4) ? -1 : 1;
        }
        if (std::abs(sRu - sRd) > two_j) continue;
        const int dRu = bk->dim(index + 1, nRu, sRu, iRu), dRd = bk->dim(index + 1, nRd, sRd, iRd);
        if (dRu <= 0 || dRd <= 0) continue;
        double* bu = Tup->blk(nu, su, iu, nRu, sRu, iRu);
        double* bd = Tdn->blk(nd, sd, id, nRd, sRd, iRd);
        double* br = prev->blk(nRu, sRu, iRu, nRd, sRd, iRd);
        double alpha = 1.0;
        if (g >= 2) {
          if (two_j == 0) alpha = (jw_phase ? -1.0 : 1.0) * ((sRu + 1.0) / (su + 1.0));
          else if (prime_last)
            alpha = phase(sRu + sd + two_j + jw3) * (sRd + 1) * std::sqrt((sRu + 1.0) / (sd + 1.0)) * wigner6j(sRu, sRd, two_j, sd, su, 1);
          else
            alpha = phase(sRd + su + two_j + jw3) * (sRu + 1) * std::sqrt((sRd + 1.0) / (su + 1.0)) * wigner6j(sRd, sRu, two_j, su, sd, 1);
        }
        work.resize((size_t)dLu * dRd);
        dgemm('N', 'N', dLu, dRd, dRu, alpha, bu, dLu, br, dRu, 0.0, work.data(), dLu);
        dgemm('N', 'T', dLu, dLd, dRd, 1.0, work.data(), dLu, bd, dLd, 1.0, out, dLu);
      }
    }
  }
}

void TOp::axpy_transpose_CD(double a, TOp* o) {  // TensorOperator.cpp:416-455
  for (int ik = 0; ik < nk(); ik++) {
    const int n = Nup[ik], su = Sup[ik], sd = Sdn[ik], iu = Iup[ik], id = iu ^ n_irrep;
    const int du = bk->dim(index, n, su, iu), dd = bk->dim(index, n, sd, id);
    double pref = a;
    if (su != sd) pref *= phase(su - sd) * std::sqrt(moving_right ? (su + 1.0) / (sd + 1.0) : (sd + 1.0) / (su + 1.0));
    const double* blk = o->blk(n, sd, id, n, su, iu);
    double* out = data.data() + k2i[ik];
    for (int r = 0; r < du; r++)
      for (int c = 0; c < dd; c++) out[r + du * c] += pref * blk[c + dd * r];
  }
}

// -------------------------------------------------------------- Sobject ----
static uint64_t key9(int nl, int sl, int il, int n1, int n2, int tj, int nr, int sr, int ir) {
  return key6(nl, sl, il, nr, sr, ir) | ((uint64_t)n1 << 48) | ((uint64_t)n2 << 52) | ((uint64_t)tj << 56);
}
void Sobject::init(int index_, Bookkeeper* bk_) {  // Sobject.cpp:36-147
  index = index_; bk = bk_;
  I1 = bk->irrep(index); I2 = bk->irrep(index + 1);
  k2i.assign(1, 0);
  for (int nl = bk->Nmin[index]; nl <= bk->Nmax[index]; nl++)
    for (int sl = bk->Smin[index][nl - bk->Nmin[index]]; sl <= bk->Smax[index][nl - bk->Nmin[index]]; sl += 2)
      for (int il = 0; il < bk->nirrep; il++) {
        const int dl = bk->dim(index, nl, sl, il);
        if (dl <= 0) continue;
        for (int n1 = 0; n1 <= 2; n1++)
          for (int n2 = 0; n2 <= 2; n2++) {
            const int nr = nl + n1 + n2;
            const int im = (n1 == 1) ? (il ^ I1) : il;
            const int ir = (n2 == 1) ? (im ^ I2) : im;
            const int tjmin = (n1 + n2) % 2, tjmax = (n1 == 1 && n2 == 1) ? 2 : tjmin;
            for (int tj = tjmin; tj <= tjmax; tj += 2)
              for (int sr = sl - tj; sr <= sl + tj; sr += 2) {
                if (sr < 0) continue;
                const int dr = bk->dim(index + 2, nr, sr, ir);
                if (dr <= 0) continue;
                lut[key9(nl, sl, il, n1, n2, tj, nr, sr, ir)] = (int)NL.size();
                NL.push_back(nl); SL.push_back(sl); IL.push_back(il); N1.push_back(n1); N2.push_back(n2);
                TJ.push_back(tj); NR.push_back(nr); SR.push_back(sr); IR.push_back(ir);
                k2i.push_back(k2i.back() + dl * dr);
              }
          }
      }
  data.assign(k2i.back(), 0.0);
  reorder.resize(nk());
  std::iota(reorder.begin(), reorder.end(), 0);
  std::stable_sort(reorder.begin(), reorder.end(), [&](int a, int b) {
    return (k2i[a + 1] - k2i[a]) > (k2i[b + 1] - k2i[b]);
  });
}
int Sobject::kappa(int nl, int sl, int il, int n1, int n2, int tj, int nr, int sr, int ir) const {
  if (n1 < 0 || n1 > 2 || n2 < 0 || n2 > 2 || tj < 0 || tj > 2 || sl < 0 || sr < 0) return -1;
  auto it = lut.find(key9(nl, sl, il, n1, n2, tj, nr, sr, ir));
  return it == lut.end() ? -1 : it->second;
}

void Sobject::join(TensorT* Tl, TensorT* Tr) {  // Sobject.cpp:212-258
  for (int ik = 0; ik < nk(); ik++) {
    const int nl = NL[ik], sl = SL[ik], il = IL[ik], nr = NR[ik], sr = SR[ik], ir = IR[ik];
    const int tj = TJ[ik], n1 = N1[ik], n2 = N2[ik];
    const int s1 = (n1 == 1) ? 1 : 0, s2 = (n2 == 1) ? 1 : 0;
    const int fase = phase(sl + sr + s1 + s2);
    const int dl = bk->dim(index, nl, sl, il), dr = bk->dim(index + 2, nr, sr, ir);
    double* out = data.data() + k2i[ik];
    std::fill(out, out + (size_t)dl * dr, 0.0);
    const int nm = nl + n1, im = (s1 == 1) ? (il ^ I1) : il;
    const int lo = std::max(std::abs(sl - s1), std::abs(sr - s2)), hi = std::min(sl + s1, sr + s2);
    for (int sm = lo; sm <= hi; sm += 2) {
      const int dm = bk->dim(index + 1, nm, sm, im);
      if (dm <= 0) continue;
      const double* bl = Tl->blk(nl, sl, il, nm, sm, im);
      const double* br = Tr->blk(nm, sm, im, nr, sr, ir);
      const double pref = fase * std::sqrt(1.0 * (tj + 1) * (sm + 1)) * wigner6j(sl, sr, tj, s2, s1, sm);
      dgemm('N', 'N', dl, dr, dm, pref, bl, dl, br, dm, 1.0, out, dl);
    }
  }
}

void Sobject::prog2symm() {  // Sobject.cpp:624-636
  for (int ik = 0; ik < nk(); ik++) {
    const double a = std::sqrt(SR[ik] + 1.0);
    for (int c = k2i[ik]; c < k2i[ik + 1]; c++) data[c] *= a;
  }
}
void Sobject::symm2prog() {  // Sobject.cpp:638-650
  for (int ik = 0; ik < nk(); ik++) {
    const double a = 1.0 / std::sqrt(SR[ik] + 1.0);
    for (int c = k2i[ik]; c < k2i[ik + 1]; c++) data[c] *= a;
  }
}

double Sobject::split(TensorT* Tl, TensorT* Tr, int D, bool moving_right, bool change) {  // Sobject.cpp:260-622
  struct Center { int nm, sm, im, dLt = 0, dRt = 0, cd = 0; std::vector<double> U, s, VT; };
  std::vector<Center> cs;
  for (int nm = bk->Nmin[index + 1]; nm <= bk->Nmax[index + 1]; nm++)
    for (int sm = bk->Smin[index + 1][nm - bk->Nmin[index + 1]]; sm <= bk->Smax[index + 1][nm - bk->Nmin[index + 1]]; sm += 2)
      for (int im = 0; im < bk->nirrep; im++)
        if (bk->fci_dim(index + 1, nm, sm, im) > 0) { Center c; c.nm = nm; c.sm = sm; c.im = im; cs.push_back(c); }

  // Left / right multiplet lists of a center (same enumeration order as the reference).
  auto lefts = [&](const Center& c, std::vector<int>* out) {  // (nl, sl, il, dl)
    for (int nl = c.nm - 2; nl <= c.nm; nl++) {
      const int s1 = (nl + 1 == c.nm) ? 1 : 0;
      for (int sl = c.sm - s1; sl <= c.sm + s1; sl += 2) {
        if (sl < 0) continue;
        const int il = (s1 == 1) ? (I1 ^ c.im) : c.im;
        const int dl = bk->dim(index, nl, sl, il);
        if (dl > 0) { out->push_back(nl); out->push_back(sl); out->push_back(il); out->push_back(dl); }
      }
    }
  };
  auto rights = [&](const Center& c, std::vector<int>* out) {
    for (int nr = c.nm; nr <= c.nm + 2; nr++) {
      const int s2 = (nr == c.nm + 1) ? 1 : 0;
      for (int sr = c.sm - s2; sr <= c.sm + s2; sr += 2) {
        if (sr < 0) continue;
        const int ir = (s2 == 1) ? (I2 ^ c.im) : c.im;
        const int dr = bk->dim(index + 2, nr, sr, ir);
        if (dr > 0) { out->push_back(nr); out->push_back(sr); out->push_back(ir); out->push_back(dr); }
      }
    }
  };

  for (auto& c : cs) {
    std::vector<int> lf, rt;
    lefts(c, &lf); rights(c, &rt);
    for (size_t a = 0; a < lf.size(); a += 4) c.dLt += lf[a + 3];
    for (size_t a = 0; a < rt.size(); a += 4) c.dRt += rt[a + 3];
    c.cd = std::min(c.dLt, c.dRt);
    if (c.cd <= 0) continue;
    std::vector<double> M((size_t)c.dLt * c.dRt, 0.0);
    int lo = 0;
    for (size_t a = 0; a < lf.size(); a += 4) {
      const int nl = lf[a], sl = lf[a + 1], il = lf[a + 2], dl = lf[a + 3];
      const int s1 = (nl + 1 == c.nm) ? 1 : 0;
      int ro = 0;
      for (size_t b = 0; b < rt.size(); b += 4) {
        const int nr = rt[b], sr = rt[b + 1], ir = rt[b + 2], dr = rt[b + 3];
        const int s2 = (nr == c.nm + 1) ? 1 : 0;
        const int fase = phase(sl + sr + s1 + s2);
        const int tjmin = std::max(std::abs(sr - sl), std::abs(s2 - s1)), tjmax = std::min(s1 + s2, sl + sr);
        for (int tj = tjmin; tj <= tjmax; tj += 2) {
          const double pref = fase * std::sqrt(1.0 * (tj + 1) * (sr + 1)) * wigner6j(sl, sr, tj, s2, s1, c.sm);
          const int k = kappa(nl, sl, il, c.nm - nl, nr - c.nm, tj, nr, sr, ir);
          assert(k >= 0);
          const double* blk = data.data() + k2i[k];
          for (int l = 0; l < dl; l++)
            for (int r = 0; r < dr; r++) M[lo + l + (size_t)c.dLt * (ro + r)] += pref * blk[l + dl * r];
        }
        ro += dr;
      }
      lo += dl;
    }
    c.U.resize((size_t)c.dLt * c.cd); c.s.resize(c.cd); c.VT.resize((size_t)c.cd * c.dRt);
    svd_thin(c.dLt, c.dRt, M.data(), c.U.data(), c.s.data(), c.VT.data());
  }

  double disc = 0.0;
  bool update = false;
  std::vector<int> nd(cs.size());
  if (change) {
    int tot = 0;
    for (size_t i = 0; i < cs.size(); i++) { nd[i] = cs[i].cd; tot += nd[i]; }
    if (tot > D) {
      std::vector<double> vals;
      for (auto& c : cs) for (int i = 0; i < c.cd; i++) vals.push_back(c.s[i]);
      std::sort(vals.begin(), vals.end(), std::greater<double>());
      const double lb = vals[D];
      for (size_t i = 0; i < cs.size(); i++)
        for (int k = 0; k < nd[i]; k++)
          if (cs[i].s[k] <= lb) { nd[i] = k; break; }
      double tsum = 0.0, dsum = 0.0;
      for (auto& c : cs)
        for (int k = 0; k < c.cd; k++) {
          const double t = (c.sm + 1) * c.s[k] * c.s[k];
          tsum += t;
          if (c.s[k] <= lb) dsum += t;
        }
      disc = dsum / tsum;
    }
    for (size_t i = 0; i < cs.size(); i++)
      if (nd[i] != bk->dim(index + 1, cs[i].nm, cs[i].sm, cs[i].im)) update = true;
  }
  if (update) {
    for (size_t i = 0; i < cs.size(); i++) bk->set_dim(index + 1, cs[i].nm, cs[i].sm, cs[i].im, nd[i]);
    Tl->init(Tl->index, bk);
    Tr->init(Tr->index, bk);
  }
  for (auto& c : cs) {
    const int dm = bk->dim(index + 1, c.nm, c.sm, c.im);
    if (dm <= 0) continue;
    const int lim = std::min(dm, c.cd);
    std::vector<int> lf, rt;
    lefts(c, &lf); rights(c, &rt);
    int lo = 0;
    for (size_t a = 0; a < lf.size(); a += 4) {
      const int nl = lf[a], sl = lf[a + 1], il = lf[a + 2], dl = lf[a + 3];
      double* B = Tl->blk(nl, sl, il, c.nm, c.sm, c.im);
      for (int r = 0; r < dm; r++)
        for (int l = 0; l < dl; l++)
          B[l + dl * r] = (r < lim) ? (moving_right ? 1.0 : c.s[r]) * c.U[lo + l + (size_t)c.dLt * r] : 0.0;
      lo += dl;
    }
    int ro = 0;
    for (size_t b = 0; b < rt.size(); b += 4) {
      const int nr = rt[b], sr = rt[b + 1], ir = rt[b + 2], dr = rt[b + 3];
      double* B = Tr->blk(c.nm, c.sm, c.im, nr, sr, ir);
      const double fb = std::sqrt((c.sm + 1.0) / (sr + 1));
      for (int r = 0; r < dr; r++)
        for (int l = 0; l < dm; l++)
          B[l + dm * r] = (l < lim) ? fb * (moving_right ? c.s[l] : 1.0) * c.VT[l + (size_t)c.cd * (ro + r)] : 0.0;
      ro += dr;
    }
  }
  return disc;
}

// ------------------------------------------------------ dense routines -----
void sym_eig(int n, const double* A, int lda, double* w, double* V, int ldv) {
  std::vector<double> a((size_t)n * n);
  for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) a[i + (size_t)n * j] = 0.5 * (A[i + (size_t)lda * j] + A[j + (size_t)lda * i]);
  std::vector<double> v((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) { double x = a[i + (size_t)n * j] * a[i + (size_t)n * j]; tot += x; if (i != j) off += x; }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[p + (size_t)n * q];
        if (apq == 0.0) continue;
        const double app = a[p + (size_t)n * p], aqq = a[q + (size_t)n * q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p,q
          const double akp = a[k + (size_t)n * p], akq = a[k + (size_t)n * q];
          a[k + (size_t)n * p] = c * akp - s * akq;
          a[k + (size_t)n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // rows p,q
          const double apk = a[p + (size_t)n * k], aqk = a[q + (size_t)n * k];
          a[p + (size_t)n * k] = c * apk - s * aqk;
          a[q + (size_t)n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[k + (size_t)n * p], vkq = v[k + (size_t)n * q];
          v[k + (size_t)n * p] = c * vkp - s * vkq;
          v[k + (size_t)n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x + (size_t)n * x] < a[y + (size_t)n * y]; });
  for (int j = 0; j < n; j++) {
    w[j] = a[ord[j] + (size_t)n * ord[j]];
    for (int i = 0; i < n; i++) V[i + (size_t)ldv * j] = v[i + (size_t)n * ord[j]];
  }
}

// One-sided Jacobi (Hestenes) on the columns of a tall matrix.
static void svd_tall(int m, int n, const double* A, double* U, double* s, double* VT) {
  std::vector<double> a(A, A + (size_t)m * n);
  std::vector<double> v((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 80; sweep++) {
    bool rotated = false;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double alpha = 0, beta = 0, gamma = 0;
        const double* ap = a.data() + (size_t)m * p;
        const double* aq = a.data() + (size_t)m * q;
        for (int i = 0; i < m; i++) { alpha += ap[i] * ap[i]; beta += aq[i] * aq[i]; gamma += ap[i] * aq[i]; }
        if (gamma == 0.0 || std::fabs(gamma) <= 1e-15 * std::sqrt(alpha * beta)) continue;
        rotated = true;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        double* pp = a.data() + (size_t)m * p;
        double* qq = a.data() + (size_t)m * q;
        for (int i = 0; i < m; i++) { const double x = pp[i], y = qq[i]; pp[i] = c * x - sn * y; qq[i] = sn * x + c * y; }
        double* vp = v.data() + (size_t)n * p;
        double* vq = v.data() + (size_t)n * q;
        for (int i = 0; i < n; i++) { const double x = vp[i], y = vq[i]; vp[i] = c * x - sn * y; vq[i] = sn * x + c * y; }
      }
    if (!rotated) break;
  }
  std::vector<double> nrm(n);
  for (int j = 0; j < n; j++) { double t = 0; for (int i = 0; i < m; i++) t += a[i + (size_t)m * j] * a[i + (size_t)m * j]; nrm[j] = std::sqrt(t); }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
  const double smax = n > 0 ? nrm[ord[0]] : 0.0;
  for (int j = 0; j < n; j++) {
    const int o = ord[j];
    s[j] = nrm[o];
    for (int i = 0; i < n; i++) VT[j + (size_t)n * i] = v[i + (size_t)n * o];
    if (nrm[o] > 1e-300 && nrm[o] > 1e-15 * smax) {
      for (int i = 0; i < m; i++) U[i + (size_t)m * j] = a[i + (size_t)m * o] / nrm[o];
    } else {
      // Null direction: complete U with a unit vector orthogonal to the previous columns.
      for (int trial = 0; trial < m; trial++) {
        for (int i = 0; i < m; i++) U[i + (size_t)m * j] = (i == (j + trial) % m) ? 1.0 : 0.0;
        for (int pass = 0; pass < 2; pass++)
          for (int k = 0; k < j; k++) {
            double d = 0; for (int i = 0; i < m; i++) d += U[i + (size_t)m * k] * U[i + (size_t)m * j];
            for (int i = 0; i < m; i++) U[i + (size_t)m * j] -= d * U[i + (size_t)m * k];
          }
        double t = 0; for (int i = 0; i < m; i++) t += U[i + (size_t)m * j] * U[i + (size_t)m * j];
        if (t > 1e-6) { t = 1.0 / std::sqrt(t); for (int i = 0; i < m; i++) U[i + (size_t)m * j] *= t; break; }
      }
    }
  }
}

void svd_thin(int m, int n, const double* A, double* U, double* s, double* VT) {
  if (m >= n) { svd_tall(m, n, A, U, s, VT); return; }
  // A^T = V S U^T : run on the transpose (n x m), then swap roles.
  std::vector<double> At((size_t)n * m);
  for (int j = 0; j < n; j++) for (int i = 0; i < m; i++) At[j + (size_t)n * i] = A[i + (size_t)m * j];
  std::vector<double> V((size_t)n * m), UT((size_t)m * m);
  svd_tall(n, m, At.data(), V.data(), s, UT.data());
  // U (m x m) = UT^T ; VT (m x n) = V^T
  for (int j = 0; j < m; j++) for (int i = 0; i < m; i++) U[i + (size_t)m * j] = UT[j + (size_t)m * i];
  for (int j = 0; j < n; j++) for (int i = 0; i < m; i++) VT[i + (size_t)m * j] = V[j + (size_t)n * i];
}

void qr_inplace(int m, int n, double* A, double* R) {
  const int r = std::min(m, n);
  std::vector<double> tau(r), work(n);
  for (int k = 0; k < r; k++) {
    double nrm = 0;
    for (int i = k; i < m; i++) nrm += A[i + (size_t)m * k] * A[i + (size_t)m * k];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) { tau[k] = 0.0; continue; }
    const double akk = A[k + (size_t)m * k];
    const double beta = (akk >= 0) ? -nrm : nrm;
    const double v0 = akk - beta;
    for (int i = k + 1; i < m; i++) A[i + (size_t)m * k] /= v0;
    tau[k] = (beta - akk) / beta;
    A[k + (size_t)m * k] = beta;
    for (int j = k + 1; j < n; j++) {
      double d = A[k + (size_t)m * j];
      for (int i = k + 1; i < m; i++) d += A[i + (size_t)m * k] * A[i + (size_t)m * j];
      d *= tau[k];
      A[k + (size_t)m * j] -= d;
      for (int i = k + 1; i < m; i++) A[i + (size_t)m * j] -= d * A[i + (size_t)m * k];
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) R[i + (size_t)n * j] = (i <= j && i < r) ? A[i + (size_t)m * j] : 0.0;
  // Form Q explicitly (first r columns), remaining columns zero.
  std::vector<double> Q((size_t)m * n, 0.0);
  for (int i = 0; i < r; i++) Q[i + (size_t)m * i] = 1.0;
  for (int k = r - 1; k >= 0; k--) {
    for (int j = k; j < r; j++) {
      double d = Q[k + (size_t)m * j];
      for (int i = k + 1; i < m; i++) d += A[i + (size_t)m * k] * Q[i + (size_t)m * j];
      d *= tau[k];
      Q[k + (size_t)m * j] -= d;
      for (int i = k + 1; i < m; i++) Q[i + (size_t)m * j] -= d * A[i + (size_t)m * k];
    }
  }
  std::copy(Q.begin(), Q.end(), A);
}

}  // namespace orc
