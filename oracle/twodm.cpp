// Two-particle reduced density matrix of a converged MPS, restated on the CPU
// (oracle; test infrastructure only): DMRG::calc_rdms_and_correlations
// (DMRGtechnics.cpp:40-130, the 2-RDM part), DMRG::left_normalize / right_normalize
// (DMRGtechnics.cpp:276-330), TensorT::QR / LQ / LeftMultiply / RightMultiply
// (TensorT.cpp:188-420) and TwoDM (TwoDM.cpp:40-630).
#include "twodm.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

namespace {

// Frobenius product <U, op(A) X op(B)> of column-major blocks: U is m x n, X is p x q,
// op(A) = A^T with A stored p x m (at = true) or A itself (m x p), absent = identity;
// op(B) = B (q x n) or B^T with B stored n x q (bt = true), absent = identity.
double sand(const double* U, int m, int n, const double* X, int p, int q, const double* A, bool at, const double* B,
            bool bt) {
  std::vector<double> W;
  const double* Y = X;  // m x q after the left factor
  if (A) {
    W.assign((size_t)m * q, 0.0);
    dgemm(at ? 'T' : 'N', 'N', m, q, p, 1.0, A, at ? p : m, X, p, 0.0, W.data(), m);
    Y = W.data();
  }
  std::vector<double> Z;
  const double* F = Y;  // m x n after the right factor
  if (B) {
    Z.assign((size_t)m * n, 0.0);
    dgemm('N', bt ? 'T' : 'N', m, n, q, 1.0, Y, m, B, bt ? n : q, 0.0, Z.data(), m);
    F = Z.data();
  }
  double s = 0.0;
  for (size_t e = 0; e < (size_t)m * n; e++) s += U[e] * F[e];
  return s;
}

}  // namespace

void TwoDM::init(const Problem* p) {
  P = p;
  L = p->L;
  const size_t n = (size_t)L * L * L * L;
  A.assign(n, 0.0);
  B.assign(n, 0.0);
}

// TwoDM::set_2rdm_A_DMRG / set_2rdm_B_DMRG (TwoDM.cpp:76-96): the value and its three
// index permutations (ijkl, jilk, klij, lkji).
void TwoDM::set(std::vector<double>& R, int i, int j, int k, int l, double v) {
  auto at = [&](int a, int b, int c, int d) -> double& { return R[a + (size_t)L * (b + (size_t)L * (c + (size_t)L * d))]; };
  at(i, j, k, l) = v;
  at(j, i, l, k) = v;
  at(k, l, i, j) = v;
  at(l, k, j, i) = v;
}

// TwoDM::FillSite (TwoDM.cpp:445-628) with the diagram functions doD1 ... doD24
// (TwoDM.cpp:642-1592). Lleft: L operators of boundary `site` built moving right
// (orbital g at index site-1-g); right: the operators of boundary site+1 built moving
// left (orbital j at index j-site-1, pair (j, k) at [k-j][j-site-1]).
void TwoDM::fill_site(TensorT& T, const OpSet* left, const OpSet* right) {
  const Bookkeeper* bk = T.bk;
  const int i = T.index, Ii = bk->irrep(i);
  const int b = i, c = i + 1;  // left and right boundary of the site tensor
  const int ni = bk->nirrep;
  auto dimL = [&](int n, int s, int r) { return bk->dim(b, n, s, r); };
  auto dimR = [&](int n, int s, int r) { return bk->dim(c, n, s, r); };
  auto Nlo = [&](int bb) { return bk->Nmin[bb]; };
  auto Nhi = [&](int bb) { return bk->Nmax[bb]; };
  auto Slo = [&](int bb, int n) { return bk->Smin[bb][n - bk->Nmin[bb]]; };
  auto Shi = [&](int bb, int n) { return bk->Smax[bb][n - bk->Nmin[bb]]; };
  // every left sector (NL, 2SL, IL) of the site tensor
  auto for_left = [&](auto&& fn) {
    for (int n = Nlo(b); n <= Nhi(b); n++)
      for (int s = Slo(b, n); s <= Shi(b, n); s += 2)
        for (int r = 0; r < ni; r++) fn(n, s, r);
  };

  // D1 (TwoDM.cpp:642-668): doubly occupied site
  {
    double d1 = 0.0;
    for_left([&](int n, int s, int r) {
      const int dl = dimL(n, s, r), dr = dimR(n + 2, s, r);
      if (dl <= 0 || dr <= 0) return;
      const double* t = T.blk(n, s, r, n + 2, s, r);
      double v = 0.0;
      for (size_t e = 0; e < (size_t)dl * dr; e++) v += t[e] * t[e];
      d1 += (s + 1) * v;
    });
    set(A, i, i, i, i, 2 * d1);
    set(B, i, i, i, i, -2 * d1);
  }

  if (right) {
    // D2 (TwoDM.cpp:670-713) with L_j on the right
    for (int j = i + 1; j < L; j++) {
      if (bk->irrep(j) != Ii) continue;
      TOp* Lr = right->L[j - i - 1].get();
      double d2 = 0.0;
      for_left([&](int n, int s, int r) {
        for (int sr = s - 1; sr <= s + 1; sr += 2) {
          const int iru = r ^ Ii;
          const int dl = dimL(n, s, r), drd = dimR(n + 2, s, r), dru = dimR(n + 1, sr, iru);
          if (dl <= 0 || drd <= 0 || dru <= 0) continue;
          const double f = phase(s + 1 - sr) * 0.5 * std::sqrt((s + 1) * (sr + 1.0));
          d2 += f * sand(T.blk(n, s, r, n + 1, sr, iru), dl, dru, T.blk(n, s, r, n + 2, s, r), dl, drd, nullptr, false,
                         Lr->blk(n + 1, sr, iru, n + 2, s, r), true);
        }
      });
      set(A, i, j, i, i, 2 * d2);
      set(B, i, j, i, i, -2 * d2);
    }
    // D3 - D6 (TwoDM.cpp:715-879) with S0, F0, F1 of the pair (j, k) on the right
    for (int j = i + 1; j < L; j++)
      for (int k = j; k < L; k++) {
        if (bk->irrep(j) != bk->irrep(k)) continue;
        TOp* S0 = right->S0[k - j][j - i - 1].get();
        TOp* F0 = right->F0[k - j][j - i - 1].get();
        TOp* F1 = right->F1[k - j][j - i - 1].get();
        double d3 = 0.0, d4 = 0.0, d5 = 0.0, d6 = 0.0;
        for_left([&](int n, int s, int r) {
          const int dl = dimL(n, s, r);
          if (dl <= 0) return;
          {  // D3: empty -> doubly occupied
            const int dd = dimR(n, s, r), du = dimR(n + 2, s, r);
            if (dd > 0 && du > 0)
              d3 += std::sqrt(0.5) * (s + 1) *
                    sand(T.blk(n, s, r, n + 2, s, r), dl, du, T.blk(n, s, r, n, s, r), dl, dd, nullptr, false,
                         S0->blk(n, s, r, n + 2, s, r), false);
          }
          {  // D4: doubly occupied
            const int dr = dimR(n + 2, s, r);
            if (dr > 0) {
              const double* t = T.blk(n, s, r, n + 2, s, r);
              d4 += std::sqrt(0.5) * (s + 1) * sand(t, dl, dr, t, dl, dr, nullptr, false, F0->blk(n + 2, s, r, n + 2, s, r), false);
            }
          }
          const int irr = r ^ Ii;
          for (int sr = s - 1; sr <= s + 1; sr += 2) {  // D5: singly occupied
            const int dr = dimR(n + 1, sr, irr);
            if (dr <= 0) continue;
            const double* t = T.blk(n, s, r, n + 1, sr, irr);
            d5 += 0.5 * std::sqrt(0.5) * (sr + 1) *
                  sand(t, dl, dr, t, dl, dr, nullptr, false, F0->blk(n + 1, sr, irr, n + 1, sr, irr), false);
          }
          for (int su = s - 1; su <= s + 1; su += 2)  // D6
            for (int sd = s - 1; sd <= s + 1; sd += 2) {
              const int du = dimR(n + 1, su, irr), dd = dimR(n + 1, sd, irr);
              if (du <= 0 || dd <= 0) continue;
              const double f = std::sqrt((su + 1) / 3.0) * (sd + 1) * phase(s + sd - 1) * wigner6j(1, 1, 2, su, sd, s);
              d6 += f * sand(T.blk(n, s, r, n + 1, su, irr), dl, du, T.blk(n, s, r, n + 1, sd, irr), dl, dd, nullptr,
                             false, F1->blk(n + 1, sd, irr, n + 1, su, irr), false);
            }
        });
        set(A, i, i, j, k, 2 * d3);
        set(B, i, i, j, k, -2 * d3);
        set(A, i, j, k, i, -2 * d4 - 2 * d5 - 3 * d6);
        set(B, i, j, k, i, -2 * d4 - 2 * d5 + d6);
        set(A, i, j, i, k, 4 * d4 + 4 * d5);
        set(B, i, j, i, k, 2 * d6);
      }
  }

  if (left) {
    // D7 (TwoDM.cpp:881-925) with L_g on the left
    for (int g = 0; g < i; g++) {
      if (bk->irrep(g) != Ii) continue;
      TOp* Ll = left->L[i - 1 - g].get();
      double d7 = 0.0;
      for_left([&](int nu, int su, int iu) {
        const int dlu = dimL(nu, su, iu);
        if (dlu <= 0) return;
        for (int sd = su - 1; sd <= su + 1; sd += 2) {
          const int ir = iu ^ Ii;
          const int dld = dimL(nu - 1, sd, ir), dr = dimR(nu + 1, sd, ir);
          if (dld <= 0 || dr <= 0) continue;
          const double f = 0.5 * std::sqrt((sd + 1) * (su + 1.0)) * phase(su - sd + 3);
          d7 += f * sand(T.blk(nu, su, iu, nu + 1, sd, ir), dlu, dr, T.blk(nu - 1, sd, ir, nu + 1, sd, ir), dld, dr,
                         Ll->blk(nu - 1, sd, ir, nu, su, iu), true, nullptr, false);
        }
      });
      set(A, g, i, i, i, 2 * d7);
      set(B, g, i, i, i, -2 * d7);
    }
  }

  if (!left || !right) return;
  // D8 - D12 (TwoDM.cpp:927-1106): L_g on the left, L_j on the right
  for (int g = 0; g < i; g++)
    for (int j = i + 1; j < L; j++) {
      const int Ig = bk->irrep(g);
      if (Ig != bk->irrep(j)) continue;
      TOp* Ll = left->L[i - 1 - g].get();
      TOp* Lr = right->L[j - i - 1].get();
      double d8 = 0.0, d9 = 0.0, d10 = 0.0, d11 = 0.0, d12 = 0.0;
      for_left([&](int n, int su, int iu) {
        const int dlu = dimL(n, su, iu);
        if (dlu <= 0) return;
        const int idn = iu ^ Ig;
        {  // D8
          const int dru = dimR(n + 2, su, iu);
          if (dru > 0)
            for (int sd = su - 1; sd <= su + 1; sd += 2) {
              const int dld = dimL(n - 1, sd, idn), drd = dimR(n + 1, sd, idn);
              if (dld <= 0 || drd <= 0) continue;
              d8 += -0.5 * (su + 1) *
                    sand(T.blk(n, su, iu, n + 2, su, iu), dlu, dru, T.blk(n - 1, sd, idn, n + 1, sd, idn), dld, drd,
                         Ll->blk(n - 1, sd, idn, n, su, iu), true, Lr->blk(n + 1, sd, idn, n + 2, su, iu), false);
            }
        }
        {  // D9, D10, D11
          const int iru = iu ^ Ii, ird = idn ^ Ii;
          for (int sld = su - 1; sld <= su + 1; sld += 2)
            for (int sru = su - 1; sru <= su + 1; sru += 2)
              for (int srd = sru - 1; srd <= sru + 1; srd += 2) {
                if (sld < 0 || sru < 0 || srd < 0 || std::abs(sld - srd) > 1) continue;
                const int dld = dimL(n - 1, sld, idn), dru = dimR(n + 1, sru, iru), drd = dimR(n, srd, ird);
                if (dld <= 0 || dru <= 0 || drd <= 0) continue;
                const double v = sand(T.blk(n, su, iu, n + 1, sru, iru), dlu, dru, T.blk(n - 1, sld, idn, n, srd, ird), dld,
                                      drd, Ll->blk(n - 1, sld, idn, n, su, iu), true, Lr->blk(n, srd, ird, n + 1, sru, iru),
                                      false);
                const double f1 = phase(su + srd + 2) * (sru + 1) * std::sqrt((srd + 1) * (su + 1.0)) *
                                  wigner6j(sru, 1, su, sld, 1, srd);
                const double f2 = 2 * (sru + 1) * std::sqrt((srd + 1) * (su + 1.0)) * wigner6j(sru, sld, 2, 1, 1, su) *
                                  wigner6j(sru, sld, 2, 1, 1, srd);
                const double f3 = (srd == su) ? (sru + 1) : 0;
                d9 += f1 * v;
                d10 += f2 * v;
                d11 += f3 * v;
              }
        }
        {  // D12
          const int dru = dimR(n, su, iu);
          if (dru > 0)
            for (int sd = su - 1; sd <= su + 1; sd += 2) {
              const int dld = dimL(n - 1, sd, idn), drd = dimR(n + 1, sd, idn);
              if (dld <= 0 || drd <= 0) continue;
              const double f = phase(sd + 1 - su) * 0.5 * std::sqrt((su + 1) * (sd + 1.0));
              d12 += f * sand(T.blk(n, su, iu, n, su, iu), dlu, dru, T.blk(n - 1, sd, idn, n + 1, sd, idn), dld, drd,
                              Ll->blk(n - 1, sd, idn, n, su, iu), true, Lr->blk(n, su, iu, n + 1, sd, idn), true);
            }
        }
      });
      set(A, g, i, j, i, -4 * d8 - d9);
      set(A, g, i, i, j, 2 * d8 + d11);
      set(B, g, i, j, i, d9 - 2 * d10);
      set(B, g, i, i, j, 2 * d8 + 2 * d10 - d11);
      set(A, g, j, i, i, 2 * d12);
      set(B, g, j, i, i, -2 * d12);
    }

  // D13 - D24 (TwoDM.cpp:1108-1592): L_g on the left, a pair operator (j, k) on the right
  for (int g = 0; g < i; g++)
    for (int j = i + 1; j < L; j++)
      for (int k = j; k < L; k++) {
        const int Ig = bk->irrep(g);
        if ((Ig ^ Ii) != (bk->irrep(j) ^ bk->irrep(k))) continue;
        TOp* Ll = left->L[i - 1 - g].get();
        TOp* S0 = right->S0[k - j][j - i - 1].get();
        TOp* S1 = (k > j) ? right->S1[k - j][j - i - 1].get() : nullptr;
        TOp* F0 = right->F0[k - j][j - i - 1].get();
        TOp* F1 = right->F1[k - j][j - i - 1].get();
        double d[25] = {0.0};
        for_left([&](int n, int su, int iu) {
          const int dlu = dimL(n, su, iu);
          if (dlu <= 0) return;
          const int idn = iu ^ Ig;   // left sector irrep of the lower tensor
          const int ird = idn ^ Ii;  // its right sector irrep when site i is singly occupied
          const int iru = iu ^ Ii;
          for (int sd = su - 1; sd <= su + 1; sd += 2) {
            const int dld = dimL(n - 1, sd, idn);
            if (dld <= 0) continue;
            const double* Lb = Ll->blk(n - 1, sd, idn, n, su, iu);
            {  // D13 (S0), D15 (S1): up doubly occupied, down singly occupied
              const int dru = dimR(n + 2, su, iu);
              if (dru > 0) {
                const double* Tu = T.blk(n, su, iu, n + 2, su, iu);
                const int drd = dimR(n, su, ird);
                if (drd > 0)
                  d[13] += -0.5 * std::sqrt(0.5) * (su + 1) *
                           sand(Tu, dlu, dru, T.blk(n - 1, sd, idn, n, su, ird), dld, drd, Lb, true,
                                S0->blk(n, su, ird, n + 2, su, iu), false);
                if (S1)
                  for (int srd = sd - 1; srd <= sd + 1; srd += 2) {
                    const int dr2 = dimR(n, srd, ird);
                    if (dr2 <= 0) continue;
                    const double f = phase(sd + su + 1) * (su + 1) * std::sqrt((srd + 1) / 3.0) * wigner6j(1, 1, 2, su, srd, sd);
                    d[15] += f * sand(Tu, dlu, dru, T.blk(n - 1, sd, idn, n, srd, ird), dld, dr2, Lb, true,
                                      S1->blk(n, srd, ird, n + 2, su, iu), false);
                  }
              }
            }
            {  // D14 (S0), D16 (S1): up singly occupied, down empty
              const int drd = dimR(n - 1, sd, idn);
              if (drd > 0) {
                const double* Td = T.blk(n - 1, sd, idn, n - 1, sd, idn);
                const int dru = dimR(n + 1, sd, iru);
                if (dru > 0) {
                  const double f = phase(sd + 1 - su) * 0.5 * std::sqrt(0.5 * (su + 1) * (sd + 1));
                  d[14] += f * sand(T.blk(n, su, iu, n + 1, sd, iru), dlu, dru, Td, dld, drd, Lb, true,
                                    S0->blk(n - 1, sd, idn, n + 1, sd, iru), false);
                }
                if (S1)
                  for (int sru = su - 1; sru <= su + 1; sru += 2) {
                    const int dr2 = dimR(n + 1, sru, iru);
                    if (dr2 <= 0) continue;
                    const double f = phase(sru + sd + 2) * (sru + 1) * std::sqrt((su + 1) / 3.0) * wigner6j(1, 1, 2, sru, sd, su);
                    d[16] += f * sand(T.blk(n, su, iu, n + 1, sru, iru), dlu, dr2, Td, dld, drd, Lb, true,
                                      S1->blk(n - 1, sd, idn, n + 1, sru, iru), false);
                  }
              }
            }
            {  // D17/D21 (F0), D19/D23 (F1): up empty, down singly occupied
              const int dru = dimR(n, su, iu);
              if (dru > 0) {
                const double* Tu = T.blk(n, su, iu, n, su, iu);
                const int drd = dimR(n, su, ird);
                if (drd > 0) {
                  const double* Td = T.blk(n - 1, sd, idn, n, su, ird);
                  const double f = std::sqrt(0.5) * 0.5 * (su + 1);
                  d[17] += f * sand(Tu, dlu, dru, Td, dld, drd, Lb, true, F0->blk(n, su, ird, n, su, iu), false);
                  d[21] += f * sand(Tu, dlu, dru, Td, dld, drd, Lb, true, F0->blk(n, su, iu, n, su, ird), true);
                }
                for (int srd = sd - 1; srd <= sd + 1; srd += 2) {
                  const int dr2 = dimR(n, srd, ird);
                  if (dr2 <= 0) continue;
                  const double* Td = T.blk(n - 1, sd, idn, n, srd, ird);
                  const double w = wigner6j(1, 1, 2, su, srd, sd);
                  const double f19 = phase(sd + srd - 1) * (srd + 1) * std::sqrt((su + 1) / 3.0) * w;
                  const double f23 = phase(sd + su - 1) * (su + 1) * std::sqrt((srd + 1) / 3.0) * w;
                  d[19] += f19 * sand(Tu, dlu, dru, Td, dld, dr2, Lb, true, F1->blk(n, srd, ird, n, su, iu), false);
                  d[23] += f23 * sand(Tu, dlu, dru, Td, dld, dr2, Lb, true, F1->blk(n, su, iu, n, srd, ird), true);
                }
              }
            }
            {  // D18/D22 (F0), D20/D24 (F1): up singly occupied, down doubly occupied
              const int drd = dimR(n + 1, sd, idn);
              if (drd > 0) {
                const double* Td = T.blk(n - 1, sd, idn, n + 1, sd, idn);
                const int dru = dimR(n + 1, sd, iru);
                if (dru > 0) {
                  const double* Tu = T.blk(n, su, iu, n + 1, sd, iru);
                  const double f = phase(sd + 1 - su) * 0.5 * std::sqrt(0.5 * (su + 1) * (sd + 1));
                  d[18] += f * sand(Tu, dlu, dru, Td, dld, drd, Lb, true, F0->blk(n + 1, sd, idn, n + 1, sd, iru), false);
                  d[22] += f * sand(Tu, dlu, dru, Td, dld, drd, Lb, true, F0->blk(n + 1, sd, iru, n + 1, sd, idn), true);
                }
                for (int sru = su - 1; sru <= su + 1; sru += 2) {
                  const int dr2 = dimR(n + 1, sru, iru);
                  if (dr2 <= 0) continue;
                  const double* Tu = T.blk(n, su, iu, n + 1, sru, iru);
                  const double w = wigner6j(1, 1, 2, sru, sd, su);
                  const double f20 = phase(2 * su) * std::sqrt((su + 1) * (sru + 1) * (sd + 1) / 3.0) * w;
                  const double f24 = phase(2 * su + sru - sd) * (sru + 1) * std::sqrt((su + 1) / 3.0) * w;
                  d[20] += f20 * sand(Tu, dlu, dr2, Td, dld, drd, Lb, true, F1->blk(n + 1, sd, idn, n + 1, sru, iru), false);
                  d[24] += f24 * sand(Tu, dlu, dr2, Td, dld, drd, Lb, true, F1->blk(n + 1, sru, iru, n + 1, sd, idn), true);
                }
              }
            }
          }
        });
        set(A, g, i, j, k, 2 * d[13] + 2 * d[14] + 3 * d[15] + 3 * d[16]);
        set(A, g, i, k, j, 2 * d[13] + 2 * d[14] - 3 * d[15] - 3 * d[16]);
        set(B, g, i, j, k, -2 * d[13] - 2 * d[14] + d[15] + d[16]);
        set(B, g, i, k, j, -2 * d[13] - 2 * d[14] - d[15] - d[16]);
        set(A, g, j, k, i, -2 * d[17] - 2 * d[18] - 3 * d[19] - 3 * d[20]);
        set(A, g, j, i, k, 4 * d[17] + 4 * d[18]);
        set(B, g, j, k, i, -2 * d[17] - 2 * d[18] + d[19] + d[20]);
        set(B, g, j, i, k, 2 * d[19] + 2 * d[20]);
        set(A, g, k, j, i, -2 * d[21] - 2 * d[22] - 3 * d[23] - 3 * d[24]);
        set(A, g, k, i, j, 4 * d[21] + 4 * d[22]);
        set(B, g, k, j, i, -2 * d[21] - 2 * d[22] + d[23] + d[24]);
        set(B, g, k, i, j, 2 * d[23] + 2 * d[24]);
      }
}

// TwoDM::correct_higher_multiplicities (TwoDM.cpp:630-640).
void TwoDM::correct_multiplicity() {
  if (P->TwoS == 0) return;
  const double a = 1.0 / (P->TwoS + 1.0);
  for (double& v : A) v *= a;
  for (double& v : B) v *= a;
}

double TwoDM::trace() const {  // TwoDM.cpp:205-215
  double v = 0.0;
  for (int a = 0; a < L; a++)
    for (int b = 0; b < L; b++) v += A[a + (size_t)L * (b + (size_t)L * (a + (size_t)L * b))];
  return v;
}

double TwoDM::energy() const {  // TwoDM.cpp:217-232 (h of Problem::gMxElement, DMRG orbital order)
  double v = 0.0;
  for (size_t e = 0; e < A.size(); e++) v += A[e] * P->mx[e];
  return 0.5 * v + P->econst;
}

// TensorT::QR + LeftMultiply (TensorT.cpp:188-280, 404-420) as DMRG::left_normalize
// (DMRGtechnics.cpp:276-297): per right sector the stacked left blocks become Q and R
// multiplies the next site's blocks from the left.
void left_normalize(TensorT& T, TensorT* next) {
  const Bookkeeper* bk = T.bk;
  const int b = T.index + 1;
  for (int nr = bk->Nmin[b]; nr <= bk->Nmax[b]; nr++)
    for (int sr = bk->Smin[b][nr - bk->Nmin[b]]; sr <= bk->Smax[b][nr - bk->Nmin[b]]; sr += 2)
      for (int ir = 0; ir < bk->nirrep; ir++) {
        const int dR = bk->dim(b, nr, sr, ir);
        if (dR <= 0) continue;
        std::vector<int> ks;
        int dLt = 0;
        for (int k = 0; k < T.nk(); k++)
          if (T.NR[k] == nr && T.SR[k] == sr && T.IR[k] == ir) {
            const int dL = bk->dim(T.index, T.NL[k], T.SL[k], T.IL[k]);
            if (dL > 0) { ks.push_back(k); dLt += dL; }
          }
        if (dLt == 0) continue;
        std::vector<double> mem((size_t)dLt * dR), R((size_t)dR * dR, 0.0);
        int off = 0;
        for (int k : ks) {
          const int dL = bk->dim(T.index, T.NL[k], T.SL[k], T.IL[k]);
          for (int r = 0; r < dR; r++)
            for (int l = 0; l < dL; l++) mem[off + l + (size_t)dLt * r] = T.data[T.k2i[k] + l + (size_t)dL * r];
          off += dL;
        }
        qr_inplace(dLt, dR, mem.data(), R.data());
        if (dLt < dR)  // rows of R beyond the rank are zero (TensorT.cpp:233-237)
          for (int r = dLt; r < dR; r++)
            for (int c = 0; c < dR; c++) R[r + (size_t)dR * c] = 0.0;
        off = 0;
        for (int k : ks) {
          const int dL = bk->dim(T.index, T.NL[k], T.SL[k], T.IL[k]);
          for (int r = 0; r < dR; r++)
            for (int l = 0; l < dL; l++) T.data[T.k2i[k] + l + (size_t)dL * r] = mem[off + l + (size_t)dLt * r];
          off += dL;
        }
        if (!next) continue;
        for (int k = 0; k < next->nk(); k++) {
          if (next->NL[k] != nr || next->SL[k] != sr || next->IL[k] != ir) continue;
          const int dX = bk->dim(b + 1, next->NR[k], next->SR[k], next->IR[k]);
          if (dX <= 0) continue;
          double* blk = next->data.data() + next->k2i[k];
          std::vector<double> tmp((size_t)dR * dX, 0.0);
          dgemm('N', 'N', dR, dX, dR, 1.0, R.data(), dR, blk, dR, 0.0, tmp.data(), dR);
          std::copy(tmp.begin(), tmp.end(), blk);
        }
      }
}

// TensorT::LQ + RightMultiply (TensorT.cpp:282-386, 421-440) as DMRG::right_normalize
// (DMRGtechnics.cpp:299-320): per left sector the right blocks, scaled by
// sqrt((2S_R + 1) / (2S_L + 1)), are stacked side by side and factored as L Q; the blocks
// become sqrt((2S_L + 1) / (2S_R + 1)) Q and L multiplies the previous site from the right.
// L Q is formed from the QR factorisation of the transposed stack.
void right_normalize(TensorT* prev, TensorT& T) {
  const Bookkeeper* bk = T.bk;
  const int b = T.index;
  for (int nl = bk->Nmin[b]; nl <= bk->Nmax[b]; nl++)
    for (int sl = bk->Smin[b][nl - bk->Nmin[b]]; sl <= bk->Smax[b][nl - bk->Nmin[b]]; sl += 2)
      for (int il = 0; il < bk->nirrep; il++) {
        const int dL = bk->dim(b, nl, sl, il);
        if (dL <= 0) continue;
        std::vector<int> ks;
        int dRt = 0;
        for (int k = 0; k < T.nk(); k++)
          if (T.NL[k] == nl && T.SL[k] == sl && T.IL[k] == il) {
            const int dR = bk->dim(b + 1, T.NR[k], T.SR[k], T.IR[k]);
            if (dR > 0) { ks.push_back(k); dRt += dR; }
          }
        if (dRt == 0) continue;
        // M^T (dRt x dL), M = [f_k T_k ...]
        std::vector<double> mt((size_t)dRt * dL), R((size_t)dL * dL, 0.0);
        int off = 0;
        for (int k : ks) {
          const int dR = bk->dim(b + 1, T.NR[k], T.SR[k], T.IR[k]);
          const double f = std::sqrt((T.SR[k] + 1.0) / (sl + 1.0));
          for (int r = 0; r < dR; r++)
            for (int l = 0; l < dL; l++) mt[(off + r) + (size_t)dRt * l] = f * T.data[T.k2i[k] + l + (size_t)dL * r];
          off += dR;
        }
        qr_inplace(dRt, dL, mt.data(), R.data());  // M^T = Q' R, so M = R^T Q'^T
        if (dRt < dL)
          for (int r = dRt; r < dL; r++)
            for (int c = 0; c < dL; c++) R[r + (size_t)dL * c] = 0.0;
        off = 0;
        for (int k : ks) {
          const int dR = bk->dim(b + 1, T.NR[k], T.SR[k], T.IR[k]);
          const double f = std::sqrt((sl + 1.0) / (T.SR[k] + 1.0));
          for (int r = 0; r < dR; r++)
            for (int l = 0; l < dL; l++) T.data[T.k2i[k] + l + (size_t)dL * r] = f * mt[(off + r) + (size_t)dRt * l];
          off += dR;
        }
        if (!prev) continue;
        // prev blocks (x -> this sector) times L = R^T
        for (int k = 0; k < prev->nk(); k++) {
          if (prev->NR[k] != nl || prev->SR[k] != sl || prev->IR[k] != il) continue;
          const int dX = bk->dim(b - 1, prev->NL[k], prev->SL[k], prev->IL[k]);
          if (dX <= 0) continue;
          double* blk = prev->data.data() + prev->k2i[k];
          std::vector<double> tmp((size_t)dX * dL, 0.0);
          dgemm('N', 'T', dX, dL, dL, 1.0, blk, dX, R.data(), dL, 0.0, tmp.data(), dX);
          std::copy(tmp.begin(), tmp.end(), blk);
        }
      }
}

// DMRG::calc_rdms_and_correlations, the 2-RDM part (DMRGtechnics.cpp:40-110): the MPS
// after a right sweep is brought into left-canonical form, then a left sweep fills the
// 2-RDM site by site with right-normalisation and left-moving operator updates.
void DMRG::calc_2rdm(TwoDM& out) {
  const int L = P.L;
  left_normalize(mps[L - 2], &mps[L - 1]);
  left_normalize(mps[L - 1], nullptr);
  ops.update_moving_right(L - 2, mps);
  out.init(&P);
  for (int site = L - 1; site >= 0; site--) {
    const OpSet* left = site > 0 ? &ops.ops[site - 1] : nullptr;
    const OpSet* right = site < L - 1 ? &ops.ops[site] : nullptr;
    if (left) assert(left->dir == 1);
    if (right) assert(right->dir == 2);
    out.fill_site(mps[site], left, right);
    if (site > 0) {
      right_normalize(&mps[site - 1], mps[site]);
      ops.update_moving_left(site - 1, mps);
    }
  }
  out.correct_multiplicity();
}

}  // namespace orc
