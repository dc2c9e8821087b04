// Two-site effective Hamiltonian of CheMPS2 restated on the CPU: driver and diagram groups 1-2
// (oracle; test infrastructure only).
#include "heff.hpp"

#include <algorithm>

namespace orc {

static const double SQ2 = std::sqrt(2.0);

void Heff::init(const Problem* P_, const Bookkeeper* bk_, Sobject* S_, Operators* ops_) {
  P = P_; bk = bk_; S = S_; ops = ops_;
  index = S->index; L = P->L;
  atLeft = (index == 0);
  atRight = (index == L - 2);
  Lo = atLeft ? nullptr : &ops->ops[index - 1];
  Ro = atRight ? nullptr : &ops->ops[index + 1];
  if (Lo) assert(Lo->dir == 1);
  if (Ro) assert(Ro->dir == 2);
}

void Heff::term(int ik, int kin, double alpha, char tl, const double* Lb, char tr, const double* Rb, bool xr_first) {
  const int dL = dl(ik), dR = dr(ik), dLi = dl(kin), dRi = dr(kin);
  const double* X = x + S->k2i[kin];
  double* Y = y + S->k2i[ik];
  if (tl == 'I' && tr == 'I') {
    for (int c = 0; c < dL * dR; c++) Y[c] += alpha * X[c];
    return;
  }
  if (tl == 'I') {
    dgemm('N', tr, dL, dR, dRi, alpha, X, dL, Rb, tr == 'N' ? dRi : dR, 1.0, Y, dL);
    flops += 2.0 * dL * dR * dRi;
    return;
  }
  if (tr == 'I') {
    dgemm(tl, 'N', dL, dR, dLi, alpha, Lb, tl == 'N' ? dL : dLi, X, dLi, 1.0, Y, dL);
    flops += 2.0 * dL * dR * dLi;
    return;
  }
  if (!xr_first) {
    w1.resize((size_t)dL * dRi);
    dgemm(tl, 'N', dL, dRi, dLi, alpha, Lb, tl == 'N' ? dL : dLi, X, dLi, 0.0, w1.data(), dL);
    dgemm('N', tr, dL, dR, dRi, 1.0, w1.data(), dL, Rb, tr == 'N' ? dRi : dR, 1.0, Y, dL);
    flops += 2.0 * dL * dRi * dLi + 2.0 * dL * dR * dRi;
  } else {
    w1.resize((size_t)dLi * dR);
    dgemm('N', tr, dLi, dR, dRi, alpha, X, dLi, Rb, tr == 'N' ? dRi : dR, 0.0, w1.data(), dLi);
    dgemm(tl, 'N', dL, dR, dLi, 1.0, Lb, tl == 'N' ? dL : dLi, w1.data(), dLi, 1.0, Y, dL);
    flops += 2.0 * dLi * dR * dRi + 2.0 * dL * dR * dLi;
  }
}

void Heff::apply(const double* x_, double* y_) { apply_range(x_, y_, 0, S->nk()); }

void Heff::apply_range(const double* x_, double* y_, int k0, int k1) {  // Heff.cpp:43-248
  x = x_;
  y = y_;
  // Output blocks are independent (OpenMP over kappa, as Heff.cpp:54-64); each thread
  // works on its own copy of the scratch state and adds its flop count at the end.
  double total = 0.0;
#pragma omp parallel reduction(+ : total)
  {
    Heff local = *this;
    local.flops = 0.0;
#pragma omp for schedule(dynamic)
    for (int ik = k0; ik < k1; ik++) local.apply_block(ik);
    total += local.flops;
  }
  flops += total;
}

void Heff::apply_block(int ik) {
  {
    std::fill(y + S->k2i[ik], y + S->k2i[ik + 1], 0.0);
    d1(ik);
    d2loc(ik);
    if (!atLeft && !atRight) d2a(ik);
    d3(ik);
    d4(ik);
    if (!atLeft && !atRight) d5(ik);
  }
}

// ---------------------------------------------------------------- group 1 --
void Heff::d1(int ik) {  // HeffDiagrams1.cpp:26-63
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  if (S->N1[ik] == 2) term(ik, ik, P->g(index, index, index, index), 'I', nullptr, 'I', nullptr);          // 1C
  if (S->N2[ik] == 2) term(ik, ik, P->g(index + 1, index + 1, index + 1, index + 1), 'I', nullptr, 'I', nullptr);  // 1D
  if (!atLeft) term(ik, ik, 1.0, 'N', Lo->X->blk(nl, sl, il, nl, sl, il), 'I', nullptr);   // 1A
  if (!atRight) term(ik, ik, 1.0, 'I', nullptr, 'T', Ro->X->blk(nr, sr, ir, nr, sr, ir));  // 1B
}

// ------------------------------------------------ group 2: local operators --
void Heff::d2loc(int ik) {  // HeffDiagrams2.cpp:850-1511
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  // 2d (on-site two-site integrals), HeffDiagrams2.cpp:996-1058
  {
    const int i = index, j = index + 1;
    if (n1 == 2 && n2 == 0) { int k = S->kappa(nl, sl, il, 0, 2, 0, nr, sr, ir); if (k >= 0) term(ik, k, P->g(i, i, j, j), 'I', nullptr, 'I', nullptr); }
    if (n1 == 0 && n2 == 2) { int k = S->kappa(nl, sl, il, 2, 0, 0, nr, sr, ir); if (k >= 0) term(ik, k, P->g(i, i, j, j), 'I', nullptr, 'I', nullptr); }
    if (n1 == 2 && n2 == 2) term(ik, ik, 4 * P->g(i, j, i, j) - 2 * P->g(i, j, j, i), 'I', nullptr, 'I', nullptr);
    if (n1 == 1 && n2 == 1) term(ik, ik, P->g(i, j, i, j) + (tj == 0 ? 1 : -1) * P->g(i, j, j, i), 'I', nullptr, 'I', nullptr);
    if ((n1 == 2 && n2 == 1) || (n1 == 1 && n2 == 2)) term(ik, ik, 2 * P->g(i, j, i, j) - P->g(i, j, j, i), 'I', nullptr, 'I', nullptr);
  }
  if (!atLeft) {
    // 2b1/2b2 and 2c1/2c2: local pair annihilators at sites i, i+1 (A at the left boundary)
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2;
      TOp* A = Lo->A[0][which].get();
      if (nloc == 0 && bk->dim(index, nl - 2, sl, il) > 0) {
        int k = which == 0 ? S->kappa(nl - 2, sl, il, 2, n2, tj, nr, sr, ir) : S->kappa(nl - 2, sl, il, n1, 2, tj, nr, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'T', A->blk(nl - 2, sl, il, nl, sl, il), 'I', nullptr);
      }
      if (nloc == 2 && bk->dim(index, nl + 2, sl, il) > 0) {
        int k = which == 0 ? S->kappa(nl + 2, sl, il, 0, n2, tj, nr, sr, ir) : S->kappa(nl + 2, sl, il, n1, 0, tj, nr, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'N', A->blk(nl, sl, il, nl + 2, sl, il), 'I', nullptr);
      }
    }
    // 2b3/2c3 spin 0: C_{ii} and C_{i+1,i+1} at the left boundary
    if (n1 != 0) term(ik, ik, (n1 == 2 ? 1.0 : 0.5) * SQ2, 'T', Lo->C[0][0]->blk(nl, sl, il, nl, sl, il), 'I', nullptr);
    if (n2 != 0) term(ik, ik, (n2 == 2 ? 1.0 : 0.5) * SQ2, 'T', Lo->C[0][1]->blk(nl, sl, il, nl, sl, il), 'I', nullptr);
    // 2b3/2c3 spin 1
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2, nother = which == 0 ? n2 : n1;
      if (nloc != 1) continue;
      TOp* D = Lo->D[0][which].get();
      const int so = (nother == 1) ? 1 : 0;
      for (int sld = sl - 2; sld <= sl + 2; sld += 2) {
        if (bk->dim(index, nl, sld, il) <= 0) continue;
        const double* Db = D->blk(nl, sld, il, nl, sl, il);
        const int jstart = (sr != sld || so == 0) ? 1 + so : 0;
        for (int jd = jstart; jd <= 1 + so; jd += 2) {
          if (std::abs(sld - sr) > jd) continue;
          int k = S->kappa(nl, sld, il, n1, n2, jd, nr, sr, ir);
          if (k < 0) continue;
          const int fase = which == 0 ? phase(sld + sr + tj + so + jd - 1) : phase(sld + sr + 2 * tj + so - 1);
          const double a = fase * std::sqrt(3.0 * (tj + 1) * (jd + 1) * (sl + 1)) * wigner6j(jd, tj, 2, 1, 1, so) * wigner6j(jd, tj, 2, sl, sld, sr);
          term(ik, k, a, 'T', Db, 'I', nullptr);
        }
      }
    }
  }
  if (!atRight) {
    // 2e1/2e2 (site i) and 2f1/2f2 (site i+1): A at the right boundary
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2;
      TOp* A = Ro->A[0][which == 0 ? 1 : 0].get();
      if (nloc == 2) {
        int k = which == 0 ? S->kappa(nl, sl, il, 0, n2, tj, nr - 2, sr, ir) : S->kappa(nl, sl, il, n1, 0, tj, nr - 2, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'I', nullptr, 'N', A->blk(nr - 2, sr, ir, nr, sr, ir));
      }
      if (nloc == 0) {
        int k = which == 0 ? S->kappa(nl, sl, il, 2, n2, tj, nr + 2, sr, ir) : S->kappa(nl, sl, il, n1, 2, tj, nr + 2, sr, ir);
        if (k >= 0) term(ik, k, SQ2, 'I', nullptr, 'T', A->blk(nr, sr, ir, nr + 2, sr, ir));
      }
    }
    // 2e3/2f3 spin 0
    if (n1 != 0) term(ik, ik, (n1 == 2 ? 1.0 : 0.5) * SQ2, 'I', nullptr, 'N', Ro->C[0][1]->blk(nr, sr, ir, nr, sr, ir));
    if (n2 != 0) term(ik, ik, (n2 == 2 ? 1.0 : 0.5) * SQ2, 'I', nullptr, 'N', Ro->C[0][0]->blk(nr, sr, ir, nr, sr, ir));
    // 2e3/2f3 spin 1
    for (int which = 0; which < 2; which++) {
      const int nloc = which == 0 ? n1 : n2, nother = which == 0 ? n2 : n1;
      if (nloc != 1) continue;
      TOp* D = Ro->D[0][which == 0 ? 1 : 0].get();
      const int so = (nother == 1) ? 1 : 0;
      for (int srd = sr - 2; srd <= sr + 2; srd += 2) {
        if (bk->dim(index + 2, nr, srd, ir) <= 0) continue;
        const double* Db = D->blk(nr, srd, ir, nr, sr, ir);
        const int jstart = (srd != sl || so == 0) ? 1 + so : 0;
        for (int jd = jstart; jd <= 1 + so; jd += 2) {
          if (std::abs(sl - srd) > jd) continue;
          int k = S->kappa(nl, sl, il, n1, n2, jd, nr, srd, ir);
          if (k < 0) continue;
          const int fase = which == 0 ? phase(srd + sl + 2 * tj + so + 1) : phase(srd + sl + tj + so + jd + 1);
          const double a = fase * std::sqrt(3.0 * (tj + 1) * (jd + 1) * (srd + 1)) * wigner6j(jd, tj, 2, 1, 1, so) * wigner6j(jd, tj, 2, sr, srd, sl);
          term(ik, k, a, 'I', nullptr, 'N', Db);
        }
      }
    }
  }
}

// ------------------------------------------------- group 2: pair sums 2a ----
void Heff::d2a(int ik) {  // HeffDiagrams2.cpp:28-849
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
          TOp* Sl = pl(Lo->S0, a, b); TOp* Ar = cr(Ro->A, a, b);
          const int ild = il ^ Sl->n_irrep, ird = ir ^ Ar->n_irrep;
          int k = S->kappa(nl + dn, sl, ild, n1, n2, tj, nr + dn, sr, ird);
          if (k < 0) continue;
          if (up == 0) term(ik, k, 1.0, 'T', Sl->blk(nl - 2, sl, ild, nl, sl, il), 'N', Ar->blk(nr - 2, sr, ird, nr, sr, ir));
          else term(ik, k, 1.0, 'N', Sl->blk(nl, sl, il, nl + 2, sl, ild), 'T', Ar->blk(nr, sr, ir, nr + 2, sr, ird));
        }
    } else {
      for (int g = index + 2; g < L; g++)
        for (int d = g; d < L; d++) {
          TOp* Al = cl(Lo->A, g, d); TOp* Sr = pr(Ro->S0, g, d);
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
              TOp* Sl = pl(Lo->S1, a, b); TOp* Br = cr(Ro->B, a, b);
              const int ild = il ^ Sl->n_irrep, ird = ir ^ Br->n_irrep;
              int k = S->kappa(nl + dn, sld, ild, n1, n2, tj, nr + dn, srd, ird);
              if (k < 0) continue;
              if (up == 0) term(ik, k, f, 'T', Sl->blk(nl - 2, sld, ild, nl, sl, il), 'N', Br->blk(nr - 2, srd, ird, nr, sr, ir));
              else term(ik, k, f, 'N', Sl->blk(nl, sl, il, nl + 2, sld, ild), 'T', Br->blk(nr, sr, ir, nr + 2, srd, ird));
            }
        } else {
          for (int g = index + 2; g < L; g++)
            for (int d = g + 1; d < L; d++) {
              TOp* Bl = cl(Lo->B, g, d); TOp* Sr = pr(Ro->S1, g, d);
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
        TOp* Fl = pl(Lo->F0, g, a); TOp* Cr = cr(Ro->C, g, a);
        const int ild = il ^ Fl->n_irrep, ird = ir ^ Cr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'N', Fl->blk(nl, sl, il, nl, sl, ild), 'T', Cr->blk(nr, sr, ir, nr, sr, ird));
      }
    for (int a = 0; a < index; a++)
      for (int g = a; g < index; g++) {
        TOp* Fl = pl(Lo->F0, a, g); TOp* Cr = cr(Ro->C, a, g);
        const int ild = il ^ Fl->n_irrep, ird = ir ^ Cr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'T', Fl->blk(nl, sl, ild, nl, sl, il), 'N', Cr->blk(nr, sr, ird, nr, sr, ir));
      }
  } else {
    for (int d = index + 2; d < L; d++)
      for (int b = d + 1; b < L; b++) {
        TOp* Cl = cl(Lo->C, d, b); TOp* Fr = pr(Ro->F0, d, b);
        const int ild = il ^ Cl->n_irrep, ird = ir ^ Fr->n_irrep;
        int k = S->kappa(nl, sl, ild, n1, n2, tj, nr, sr, ird);
        if (k >= 0) term(ik, k, 1.0, 'N', Cl->blk(nl, sl, il, nl, sl, ild), 'T', Fr->blk(nr, sr, ir, nr, sr, ird));
      }
    for (int b = index + 2; b < L; b++)
      for (int d = b; d < L; d++) {
        TOp* Cl = cl(Lo->C, b, d); TOp* Fr = pr(Ro->F0, b, d);
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
            TOp* Fl = pl(Lo->F1, g, a); TOp* Dr = cr(Ro->D, g, a);
            const int ild = il ^ Fl->n_irrep, ird = ir ^ Dr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f1, 'N', Fl->blk(nl, sl, il, nl, sld, ild), 'T', Dr->blk(nr, sr, ir, nr, srd, ird));
          }
        for (int a = 0; a < index; a++)
          for (int g = a; g < index; g++) {
            TOp* Fl = pl(Lo->F1, a, g); TOp* Dr = cr(Ro->D, a, g);
            const int ild = il ^ Fl->n_irrep, ird = ir ^ Dr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f2, 'T', Fl->blk(nl, sld, ild, nl, sl, il), 'N', Dr->blk(nr, srd, ird, nr, sr, ir));
          }
      } else {
        for (int d = index + 2; d < L; d++)
          for (int b = d + 1; b < L; b++) {
            TOp* Dl = cl(Lo->D, d, b); TOp* Fr = pr(Ro->F1, d, b);
            const int ild = il ^ Dl->n_irrep, ird = ir ^ Fr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f1, 'N', Dl->blk(nl, sl, il, nl, sld, ild), 'T', Fr->blk(nr, sr, ir, nr, srd, ird));
          }
        for (int b = index + 2; b < L; b++)
          for (int d = b; d < L; d++) {
            TOp* Dl = cl(Lo->D, b, d); TOp* Fr = pr(Ro->F1, b, d);
            const int ild = il ^ Dl->n_irrep, ird = ir ^ Fr->n_irrep;
            int k = S->kappa(nl, sld, ild, n1, n2, tj, nr, srd, ird);
            if (k >= 0) term(ik, k, f2, 'T', Dl->blk(nl, sld, ild, nl, sl, il), 'N', Fr->blk(nr, srd, ird, nr, sr, ir));
          }
      }
    }
}


// ---------------------------------------------------------------- group 3 --
// Q block plus sum_l V * L_l over the orbitals l on the same side with the same irrep
// as `site` (the x-independent combination of HeffDiagrams3.cpp:60-75 and siblings).
static void q_plus_l(std::vector<double>& out, TOp* Q, std::vector<std::unique_ptr<TOp>>& Ls, const Bookkeeper* bk,
                     const Problem* P, int site, bool left, int index, int n1, int s1, int i1, int n2, int s2, int i2) {
  const double* q = Q->blk(n1, s1, i1, n2, s2, i2);
  const int sz = bk->dim(left ? index : index + 2, n1, s1, i1) * bk->dim(left ? index : index + 2, n2, s2, i2);
  out.assign(q, q + sz);
  const int L = P->L;
  if (left) {
    for (int l = 0; l < index; l++) {
      if (bk->irrep(l) != bk->irrep(site)) continue;
      const double a = P->g(l, site, site, site);
      const double* b = Ls[index - 1 - l]->blk(n1, s1, i1, n2, s2, i2);
      for (int c = 0; c < sz; c++) out[c] += a * b[c];
    }
  } else {
    for (int l = index + 2; l < L; l++) {
      if (bk->irrep(l) != bk->irrep(site)) continue;
      const double a = P->g(site, site, site, l);
      const double* b = Ls[l - index - 2]->blk(n1, s1, i1, n2, s2, i2);
      for (int c = 0; c < sz; c++) out[c] += a * b[c];
    }
  }
}

void Heff::d3(int ik) {  // HeffDiagrams3.cpp:28-848
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  // 3E / 3H: purely local hopping (HeffDiagrams3.cpp:399-474)
  if (bk->irrep(i) == bk->irrep(j)) {
    auto ax = [&](int a1, int a2, int aj, double a) { int k = S->kappa(nl, sl, il, a1, a2, aj, nr, sr, ir); if (k >= 0) term(ik, k, a, 'I', nullptr, 'I', nullptr); };
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
      TOp* Q = Lo->Q[which].get();
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
            q_plus_l(w2, Q, Lo->L, bk, P, site, true, index, nl, sl, il, nl + 1, sld, ild);
            term(ik, k, f, 'N', w2.data(), 'I', nullptr);
          }
        }
        if (nloc == 1 && std::abs(sld - sr) <= so && sld >= 0) {  // 3A1B / 3B1B
          int k = kap(nl + 1, sld, 0, so);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sl + sr + 1 + so) : phase(sl + sr + 2 - tj);
            const double f = std::sqrt((sld + 1) * (tj + 1.0)) * fase * wigner6j(so, tj, 1, sl, sld, sr);
            term(ik, k, f, 'N', Q->blk(nl, sl, il, nl + 1, sld, ild), 'I', nullptr);
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
            term(ik, k, f, 'T', Q->blk(nl - 1, sld, ild, nl, sl, il), 'I', nullptr);
          }
        }
        if (nloc == 1 && std::abs(sld - sr) <= so && sld >= 0) {  // 3A2B+3D2 / 3B2B+3I1
          int k = kap(nl - 1, sld, 2, so);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sld + sr + 2 + so) : phase(sld + sr + 3 - tj);
            const double f = fase * std::sqrt((sl + 1) * (tj + 1.0)) * wigner6j(so, tj, 1, sl, sld, sr);
            q_plus_l(w2, Q, Lo->L, bk, P, site, true, index, nl - 1, sld, ild, nl, sl, il);
            term(ik, k, f, 'T', w2.data(), 'I', nullptr);
          }
        }
      }
    }
  }
  if (!atRight) {
    // 3K/3F (site i, which==0) and 3L/3G (site i+1, which==1): Q at the right boundary
    for (int which = 0; which < 2; which++) {
      TOp* Q = Ro->Q[which == 0 ? 1 : 0].get();
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
            term(ik, k, f, 'I', nullptr, 'N', Q->blk(nr - 1, srd, ird, nr, sr, ir));
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
            q_plus_l(w2, Q, Ro->L, bk, P, site, false, index, nr - 1, srd, ird, nr, sr, ir);
            term(ik, k, f, 'I', nullptr, 'N', w2.data());
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
            term(ik, k, f, 'I', nullptr, 'T', Q->blk(nr, sr, ir, nr + 1, srd, ird));
          }
        }
        if (nloc == 1 && std::abs(sl - srd) <= so && srd >= 0) {  // 3K2B+3F2 / 3L2B+3G2
          int k = kap(2, so, nr + 1, srd);
          if (k >= 0) {
            const int fase = which == 0 ? phase(sl + srd + tj + 1 + 2 * so) : phase(sl + srd + so + 2);
            const double f = std::sqrt((tj + 1) * (srd + 1.0)) * fase * wigner6j(so, tj, 1, sr, srd, sl);
            q_plus_l(w2, Q, Ro->L, bk, P, site, false, index, nr, sr, ir, nr + 1, srd, ird);
            term(ik, k, f, 'I', nullptr, 'T', w2.data());
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
          TOp* Q = Lo->Q[l - index].get();
          TOp* Lr = Ro->L[l - index - 2].get();
          int k = S->kappa(nl + 1, sld, ild, n1, n2, tj, nr + 1, srd, ird);
          if (k >= 0) term(ik, k, fup, 'N', Q->blk(nl, sl, il, nl + 1, sld, ild), 'T', Lr->blk(nr, sr, ir, nr + 1, srd, ird));
          k = S->kappa(nl - 1, sld, ild, n1, n2, tj, nr - 1, srd, ird);
          if (k >= 0) term(ik, k, fdn, 'T', Q->blk(nl - 1, sld, ild, nl, sl, il), 'N', Lr->blk(nr - 1, srd, ird, nr, sr, ir));
        }
        // 3J: L_l at the left boundary, Q_l at the right boundary, l left of the sites (HeffDiagrams3.cpp:742-848)
        for (int l = 0; l < index; l++) {
          const int ild = il ^ bk->irrep(l), ird = ir ^ bk->irrep(l);
          TOp* Ll = Lo->L[index - 1 - l].get();
          TOp* Q = Ro->Q[index + 1 - l].get();
          int k = S->kappa(nl + 1, sld, ild, n1, n2, tj, nr + 1, srd, ird);
          if (k >= 0) term(ik, k, fup, 'N', Ll->blk(nl, sl, il, nl + 1, sld, ild), 'T', Q->blk(nr, sr, ir, nr + 1, srd, ird));
          k = S->kappa(nl - 1, sld, ild, n1, n2, tj, nr - 1, srd, ird);
          if (k >= 0) term(ik, k, fdn, 'T', Ll->blk(nl - 1, sld, ild, nl, sl, il), 'N', Q->blk(nr - 1, srd, ird, nr, sr, ir));
        }
      }
  }
}


// ---------------------------------------------------------------- group 4 --
static const double SQH = std::sqrt(0.5);

void Heff::d4(int ik) {  // HeffDiagrams4.cpp:28-6518
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  if (!atLeft) {
    // ---- 4A: A/B/C/D of the pair (i, i+1) at the left boundary (HeffDiagrams4.cpp:28-669)
    {
      TOp* A = Lo->A[1][0].get();
      const int ild = il ^ A->n_irrep;
      auto dn = [&](int a1, int a2, int aj, double f) {  // kappa' at NL-2
        int k = S->kappa(nl - 2, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', A->blk(nl - 2, sl, ild, nl, sl, il), 'I', nullptr);
      };
      auto up = [&](int a1, int a2, int aj, double f) {  // kappa' at NL+2
        int k = S->kappa(nl + 2, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', A->blk(nl, sl, il, nl + 2, sl, ild), 'I', nullptr);
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
      TOp* B = Lo->B[1][0].get();
      const int ild = il ^ B->n_irrep;
      auto dn = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl - 2, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', B->blk(nl - 2, sld, ild, nl, sl, il), 'I', nullptr);
      };
      auto up = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl + 2, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', B->blk(nl, sl, il, nl + 2, sld, ild), 'I', nullptr);
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
      TOp* C = Lo->C[1][0].get();
      const int ild = il ^ C->n_irrep;
      auto tr = [&](int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', C->blk(nl, sl, ild, nl, sl, il), 'I', nullptr);
      };
      auto nt = [&](int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sl, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', C->blk(nl, sl, il, nl, sl, ild), 'I', nullptr);
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
      TOp* D = Lo->D[1][0].get();
      const int ild = il ^ D->n_irrep;
      auto tr = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'T', D->blk(nl, sld, ild, nl, sl, il), 'I', nullptr);
      };
      auto nt = [&](int sld, int a1, int a2, int aj, double f) {
        int k = S->kappa(nl, sld, ild, a1, a2, aj, nr, sr, ir);
        if (k >= 0) term(ik, k, f, 'N', D->blk(nl, sl, il, nl, sld, ild), 'I', nullptr);
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
          const int sz = bk->dim(index, nlu, slu, ilu) * bk->dim(index, nld, sld, ildn);
          w2.assign(sz, 0.0);
          for (size_t q = 0; q < ls.size(); q++) {
            const double* b = Lo->L[index - 1 - ls[q]]->blk(nlu, slu, ilu, nld, sld, ildn);
            for (int c = 0; c < sz; c++) w2[c] += co[q] * b[c];
          }
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
            term(ik, k, f, 'T', w2.data(), 'I', nullptr);
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
            term(ik, k, f, 'N', w2.data(), 'I', nullptr);
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
              term(ik, k, 1.0, 'T', w2.data(), 'I', nullptr);
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
              term(ik, k, 1.0, 'N', w2.data(), 'I', nullptr);
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
void Heff::d4BC(int ik) {
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
      TOp* A = Lo->A[kk][mm].get(); TOp* B = Lo->B[kk][mm].get(); TOp* C = Lo->C[kk][mm].get(); TOp* D = Lo->D[kk][mm].get();
      TOp* Lr = Ro->L[l - index - 2].get();
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


// sum_q co[q] * blk(ops[q]) over equally shaped blocks -> out
static void lin_comb(std::vector<double>& out, const std::vector<TOp*>& ops, const std::vector<double>& co, int n1, int s1,
                     int i1, int n2, int s2, int i2, int sz) {
  out.assign(sz, 0.0);
  for (size_t q = 0; q < ops.size(); q++) {
    if (co[q] == 0.0) continue;
    const double* b = ops[q]->blk(n1, s1, i1, n2, s2, i2);
    for (int c = 0; c < sz; c++) out[c] += co[q] * b[c];
  }
}

// 4E (site i) and 4H (site i+1): L_a at the left boundary times sum_b V L_b at the right
// boundary, a left and b right of the sites (HeffDiagrams4.cpp:2596-3028, 3563-4014).
void Heff::d4E(int ik) {
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
      std::vector<TOp*> Rops;
      for (int b : lb) Rops.push_back(Ro->L[b - index - 2].get());
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
                lin_comb(w2, Rops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
                term(ik, k, f, 'T', Lo->L[index - 1 - a]->blk(nl - 1, sld, ild, nl, sl, il), 'T', w2.data(), true);
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
                lin_comb(w2, Rops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
                term(ik, k, f, 'N', Lo->L[index - 1 - a]->blk(nl, sl, il, nl + 1, sld, ild), 'N', w2.data(), true);
              }
          }
          // local 2 -> 2 with exchange (4E3B / 4H3B): left N-1, right N-1
          if (nloc == 2 && std::abs(sld - srd) <= tj && bk->dim(index, nl - 1, sld, ild) > 0 && bk->dim(index + 2, nr - 1, srd, ird) > 0) {
            const double f = phase(sl + srd - so + 3) * std::sqrt((sl + 1) * (sr + 1.0)) * wigner6j(sld, srd, so, sr, sl, 1);
            int k = kap(nl - 1, sld, ild, 2, so, nr - 1, srd, ird);
            if (k >= 0)
              for (int a : la) {
                for (size_t q = 0; q < lb.size(); q++) co[q] = P->g(a, site, site, lb[q]) - 2 * P->g(a, site, lb[q], site);
                lin_comb(w2, Rops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
                term(ik, k, f, 'T', Lo->L[index - 1 - a]->blk(nl - 1, sld, ild, nl, sl, il), 'N', w2.data(), true);
              }
          }
          // local 2 -> 2 with exchange (4E4B / 4H4B): left N+1, right N+1
          if (nloc == 2 && std::abs(sld - srd) <= tj && bk->dim(index, nl + 1, sld, ild) > 0 && bk->dim(index + 2, nr + 1, srd, ird) > 0) {
            const double f = phase(sld + sr - so + 3) * std::sqrt((sld + 1) * (srd + 1.0)) * wigner6j(sl, sr, so, srd, sld, 1);
            int k = kap(nl + 1, sld, ild, 2, so, nr + 1, srd, ird);
            if (k >= 0)
              for (int a : la) {
                for (size_t q = 0; q < lb.size(); q++) co[q] = P->g(a, site, site, lb[q]) - 2 * P->g(a, site, lb[q], site);
                lin_comb(w2, Rops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
                term(ik, k, f, 'N', Lo->L[index - 1 - a]->blk(nl, sl, il, nl + 1, sld, ild), 'T', w2.data(), true);
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
                    lin_comb(w2, Rops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
                    term(ik, k, 1.0, 'T', Lo->L[index - 1 - a]->blk(nl - 1, sld, ild, nl, sl, il), 'N', w2.data(), true);
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
                    lin_comb(w2, Rops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
                    term(ik, k, 1.0, 'N', Lo->L[index - 1 - a]->blk(nl, sl, il, nl + 1, sld, ild), 'T', w2.data(), true);
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
void Heff::d4I(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  std::vector<TOp*> ops; std::vector<int> ls;
  for (int l = 0; l < index; l++) if (bk->irrep(l) == bk->irrep(i)) { ls.push_back(l); ops.push_back(Lo->L[index - 1 - l].get()); }
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
          lin_comb(w2, ops, co, nl - 1, sld, ild, nl, sl, il, bk->dim(index, nl - 1, sld, ild) * bk->dim(index, nl, sl, il));
          term(ik, k, f, 'T', w2.data(), 'I', nullptr);
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
          lin_comb(w2, ops, co, nl, sl, il, nl + 1, sld, ild, bk->dim(index, nl, sl, il) * bk->dim(index, nl + 1, sld, ild));
          term(ik, k, f, 'N', w2.data(), 'I', nullptr);
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
        lin_comb(w2, ops, co, nl - 1, sld, ild, nl, sl, il, bk->dim(index, nl - 1, sld, ild) * bk->dim(index, nl, sl, il));
        term(ik, k, 1.0, 'T', w2.data(), 'I', nullptr);
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
        lin_comb(w2, ops, co, nl, sl, il, nl + 1, sld, ild, bk->dim(index, nl, sl, il) * bk->dim(index, nl + 1, sld, ild));
        term(ik, k, 1.0, 'N', w2.data(), 'I', nullptr);
      }
    }
  }
}

// 4F (irrep(l) == irrep(i+1)) and 4G (irrep(l) == irrep(i)): sum_l V L_l at the right
// boundary, l right of the sites (HeffDiagrams4.cpp:3029-3562).
void Heff::d4F(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  std::vector<TOp*> ops; std::vector<int> ls;
  for (int l = index + 2; l < L; l++) if (bk->irrep(l) == bk->irrep(j)) { ls.push_back(l); ops.push_back(Ro->L[l - index - 2].get()); }
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
          lin_comb(w2, ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
          term(ik, k, f, 'I', nullptr, 'N', w2.data());
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
          lin_comb(w2, ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
          term(ik, k, f, 'I', nullptr, 'T', w2.data());
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
        lin_comb(w2, ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
        term(ik, k, 1.0, 'I', nullptr, 'N', w2.data());
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
        lin_comb(w2, ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
        term(ik, k, 1.0, 'I', nullptr, 'T', w2.data());
      }
    }
  }
}

void Heff::d4G(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  const int i = index, j = index + 1;
  std::vector<TOp*> ops; std::vector<int> ls;
  for (int l = index + 2; l < L; l++) if (bk->irrep(l) == bk->irrep(i)) { ls.push_back(l); ops.push_back(Ro->L[l - index - 2].get()); }
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
          lin_comb(w2, ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
          term(ik, k, f, 'I', nullptr, 'N', w2.data());
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
          lin_comb(w2, ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
          term(ik, k, f, 'I', nullptr, 'T', w2.data());
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
        lin_comb(w2, ops, co, nr, sr, ir, nr + 1, srd, ird, bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + 1, srd, ird));
        term(ik, k, 1.0, 'I', nullptr, 'T', w2.data());
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
        lin_comb(w2, ops, co, nr - 1, srd, ird, nr, sr, ir, bk->dim(index + 2, nr - 1, srd, ird) * bk->dim(index + 2, nr, sr, ir));
        term(ik, k, 1.0, 'I', nullptr, 'N', w2.data());
      }
    }
  }
}


// 4J: A/B/C/D of the pair (i, i+1) at the right boundary (HeffDiagrams4.cpp:4277-4846).
void Heff::d4J(int ik) {
  const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
  const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
  {
    TOp* A = Ro->A[1][0].get();
    const int ird = ir ^ A->n_irrep;
    auto up = [&](int a1, int a2, int aj, double f) {  // right sector N+2
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr + 2, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'T', A->blk(nr, sr, ir, nr + 2, sr, ird));
    };
    auto dn = [&](int a1, int a2, int aj, double f) {  // right sector N-2
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr - 2, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'N', A->blk(nr - 2, sr, ird, nr, sr, ir));
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
    TOp* B = Ro->B[1][0].get();
    const int ird = ir ^ B->n_irrep;
    auto up = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr + 2, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'T', B->blk(nr, sr, ir, nr + 2, srd, ird));
    };
    auto dn = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr - 2, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'N', B->blk(nr - 2, srd, ird, nr, sr, ir));
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
    TOp* C = Ro->C[1][0].get();
    const int ird = ir ^ C->n_irrep;
    auto nt = [&](int a1, int a2, int aj, double f) {  // C block stored (down -> up)
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'N', C->blk(nr, sr, ird, nr, sr, ir));
    };
    auto tr = [&](int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, sr, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'T', C->blk(nr, sr, ir, nr, sr, ird));
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
    TOp* D = Ro->D[1][0].get();
    const int ird = ir ^ D->n_irrep;
    auto nt = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'N', D->blk(nr, srd, ird, nr, sr, ir));
    };
    auto tr = [&](int srd, int a1, int a2, int aj, double f) {
      int k = S->kappa(nl, sl, il, a1, a2, aj, nr, srd, ird);
      if (k >= 0) term(ik, k, f, 'I', nullptr, 'T', D->blk(nr, sr, ir, nr, srd, ird));
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
void Heff::d4KL(int ik) {
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
      TOp* A = Ro->A[kk][mm].get(); TOp* B = Ro->B[kk][mm].get(); TOp* C = Ro->C[kk][mm].get(); TOp* D = Ro->D[kk][mm].get();
      TOp* Ll = Lo->L[index - 1 - l].get();
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
void Heff::d5(int ik) {
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
      std::vector<TOp*> Rops;
      for (int b : lb) Rops.push_back(Ro->L[b - index - 2].get());
      std::vector<double> co(lb.size());
      const int rsz = bk->dim(index + 2, nr, sr, ir) * bk->dim(index + 2, nr + dnr, srd, ird);
      for (int a : la) {
        for (size_t q = 0; q < lb.size(); q++) co[q] = coef(a, lb[q]);
        if (dnr < 0) lin_comb(w2, Rops, co, nr - 1, srd, ird, nr, sr, ir, rsz);
        else lin_comb(w2, Rops, co, nr, sr, ir, nr + 1, srd, ird, rsz);
        TOp* La = Lo->L[index - 1 - a].get();
        const double* lbk = dnl < 0 ? La->blk(nl - 1, sld, ild, nl, sl, il) : La->blk(nl, sl, il, nl + 1, sld, ild);
        term(ik, k, 1.0, dnl < 0 ? 'T' : 'N', lbk, dnr < 0 ? 'N' : 'T', w2.data(), true);
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
// Diagonal of Heff used as the Davidson preconditioner (Heff.cpp:250-315,
// HeffDiagonal.cpp:27-620): only the sector-preserving terms contribute, each
// through the diagonals of its operator blocks.
void Heff::diag(double* d) {
  const int i = index, j = index + 1;
  const double sqh = std::sqrt(0.5);
  const bool leftSum = index < L * 0.5;
  for (int ik = 0; ik < S->nk(); ik++) {
    const int nl = S->NL[ik], sl = S->SL[ik], il = S->IL[ik], nr = S->NR[ik], sr = S->SR[ik], ir = S->IR[ik];
    const int n1 = S->N1[ik], n2 = S->N2[ik], tj = S->TJ[ik];
    const int dL = dl(ik), dR = dr(ik);
    double* D = d + S->k2i[ik];
    std::fill(D, D + (size_t)dL * dR, 0.0);
    auto add_const = [&](double c) { for (int q = 0; q < dL * dR; q++) D[q] += c; };
    auto add_left = [&](double c, const double* blk) {  // c * diag(left block) along rows
      for (int r = 0; r < dR; r++) for (int l = 0; l < dL; l++) D[l + dL * r] += c * blk[(dL + 1) * l];
    };
    auto add_right = [&](double c, const double* blk) {  // c * diag(right block) along columns
      for (int r = 0; r < dR; r++) for (int l = 0; l < dL; l++) D[l + dL * r] += c * blk[(dR + 1) * r];
    };
    auto add_outer = [&](double c, const double* lb, const double* rb) {
      for (int l = 0; l < dL; l++) for (int r = 0; r < dR; r++) D[l + dL * r] += c * lb[(dL + 1) * l] * rb[(dR + 1) * r];
    };
    // 1C, 1D, 2d3 (HeffDiagonal.cpp:56-124)
    if (n1 == 2) add_const(P->g(i, i, i, i));
    if (n2 == 2) add_const(P->g(j, j, j, j));
    if (n1 == 2 && n2 == 2) add_const(4 * P->g(i, j, i, j) - 2 * P->g(i, j, j, i));
    if (n1 == 1 && n2 == 1) add_const(P->g(i, j, i, j) + (tj == 2 ? -1 : 1) * P->g(i, j, j, i));
    if ((n1 == 2 && n2 == 1) || (n1 == 1 && n2 == 2)) add_const(2 * P->g(i, j, i, j) - P->g(i, j, j, i));
    if (!atLeft) {
      add_left(1.0, Lo->X->blk(nl, sl, il, nl, sl, il));  // 1A
      // 2b3 / 2c3: local C (spin 0) and D (spin 1) of site i and i+1 (HeffDiagonal.cpp:126-328)
      for (int w = 0; w < 2; w++) {
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
      // 2e3 / 2f3: at the right boundary pair [0][1] is site i and [0][0] is site i+1
      for (int w = 0; w < 2; w++) {
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
    if (!atLeft && !atRight) {
      // 2a3 spin 0 and spin 1 (HeffDiagonal.cpp:380-620): pairs on the shorter side; every
      // pair a<b is visited twice (once per ordering) and a==b once.
      const double a1 = phase(sl + sr + tj + 2) * std::sqrt((sr + 1) * (sl + 1.0)) * wigner6j(sl, sr, tj, sr, sl, 2);
      if (leftSum) {
        for (int a = 0; a < index; a++)
          for (int b = a; b < index; b++) {
            if (bk->irrep(a) != bk->irrep(b)) continue;
            const double mult = (a == b) ? 1.0 : 2.0;
            add_outer(mult, pl(Lo->F0, a, b)->blk(nl, sl, il, nl, sl, il), cr(Ro->C, a, b)->blk(nr, sr, ir, nr, sr, ir));
            add_outer(mult * a1, pl(Lo->F1, a, b)->blk(nl, sl, il, nl, sl, il), cr(Ro->D, a, b)->blk(nr, sr, ir, nr, sr, ir));
          }
      } else {
        for (int a = index + 2; a < L; a++)
          for (int b = a; b < L; b++) {
            if (bk->irrep(a) != bk->irrep(b)) continue;
            const double mult = (a == b) ? 1.0 : 2.0;
            add_outer(mult, cl(Lo->C, a, b)->blk(nl, sl, il, nl, sl, il), pr(Ro->F0, a, b)->blk(nr, sr, ir, nr, sr, ir));
            add_outer(mult * a1, cl(Lo->D, a, b)->blk(nl, sl, il, nl, sl, il), pr(Ro->F1, a, b)->blk(nr, sr, ir, nr, sr, ir));
          }
      }
    }
  }
}

}  // namespace orc
