// Renormalised-operator construction of CheMPS2 restated on the CPU (oracle; test infrastructure only).
#include "operators.hpp"

#include <algorithm>

namespace orc {

static const double SQ2 = std::sqrt(2.0), SQH = std::sqrt(0.5);

// out(dRu x dRd) += alpha * Tup^T * op(mid) * Tdn ; Tup is dLu x dRu, Tdn is dLd x dRd,
// mid is dLu x dLd (or dLd x dLu when transposed); mid == nullptr means identity.
static void sand_r(double* out, int dRu, int dRd, int dLu, int dLd, double alpha, const double* Tup,
                   const double* mid, bool midT, const double* Tdn) {
  if (!mid) { dgemm('T', 'N', dRu, dRd, dLu, alpha, Tup, dLu, Tdn, dLd, 1.0, out, dRu); return; }
  std::vector<double> w((size_t)dRu * dLd);
  dgemm('T', midT ? 'T' : 'N', dRu, dLd, dLu, alpha, Tup, dLu, mid, midT ? dLd : dLu, 0.0, w.data(), dRu);
  dgemm('N', 'N', dRu, dRd, dLd, 1.0, w.data(), dRu, Tdn, dLd, 1.0, out, dRu);
}
// out(dLu x dLd) += alpha * Tup * op(mid) * Tdn^T ; Tup is dLu x dRu, Tdn is dLd x dRd,
// mid is dRu x dRd (or dRd x dRu when transposed).
static void sand_l(double* out, int dLu, int dLd, int dRu, int dRd, double alpha, const double* Tup,
                   const double* mid, bool midT, const double* Tdn) {
  if (!mid) { dgemm('N', 'T', dLu, dLd, dRu, alpha, Tup, dLu, Tdn, dLd, 1.0, out, dLu); return; }
  std::vector<double> w((size_t)dLu * dRd);
  dgemm('N', midT ? 'T' : 'N', dLu, dRd, dRu, alpha, Tup, dLu, mid, midT ? dRd : dRu, 0.0, w.data(), dLu);
  dgemm('N', 'T', dLu, dLd, dRd, 1.0, w.data(), dLu, Tdn, dLd, 1.0, out, dLu);
}

#define DIM(b, n, s, i) self.bk->dim((b), (n), (s), (i))

// ---------------------------------------------------------------- TensorL --
void L_create(TOp& self, TensorT* T) {  // TensorL.cpp:41-135 (previous == NULL branches)
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int nu = self.Nup[ik], su = self.Sup[ik], iu = self.Iup[ik], sd = self.Sdn[ik];
    const int nd = nu + 1, id = iu ^ self.n_irrep;
    double* out = self.data.data() + self.k2i[ik];
    if (self.moving_right) {
      const int dRu = DIM(b, nu, su, iu), dRd = DIM(b, nd, sd, id);
      for (int g = 0; g < 2; g++) {
        const int nl = g == 0 ? nu : nu - 1, sl = g == 0 ? su : sd, il = g == 0 ? iu : id;
        const int dl = DIM(b - 1, nl, sl, il);
        if (dl <= 0) continue;
        const double a = g == 0 ? 1.0 : phase(sd - su + 1) * std::sqrt((su + 1.0) / (sd + 1.0));
        dgemm('T', 'N', dRu, dRd, dl, a, T->blk(nl, sl, il, nu, su, iu), dl, T->blk(nl, sl, il, nd, sd, id), dl, 1.0, out, dRu);
      }
    } else {
      const int dLu = DIM(b, nu, su, iu), dLd = DIM(b, nd, sd, id);
      for (int g = 0; g < 2; g++) {
        const int nr = g == 0 ? nd : nu + 2, sr = g == 0 ? sd : su, ir = g == 0 ? id : iu;
        const int dr = DIM(b + 1, nr, sr, ir);
        if (dr <= 0) continue;
        const double a = g == 0 ? 1.0 : phase(su - sd + 1) * std::sqrt((su + 1.0) / (sd + 1.0));
        dgemm('N', 'T', dLu, dLd, dr, a, T->blk(nu, su, iu, nr, sr, ir), dLu, T->blk(nd, sd, id, nr, sr, ir), dLd, 1.0, out, dLu);
      }
    }
  }
}

// --------------------------------------------------------------- TensorF0 --
void F0_makenew(TOp& self, TensorT* T) {  // TensorF0.cpp:53-133
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
    double* out = self.data.data() + self.k2i[ik];
    if (self.moving_right) {
      const int dR = DIM(b, n, s, i), Is = self.bk->irrep(b - 1);
      for (int g = 0; g < 3; g++) {
        const int nl = g == 0 ? n - 2 : n - 1, sl = g == 0 ? s : (g == 1 ? s - 1 : s + 1), il = g == 0 ? i : (i ^ Is);
        const int dl = DIM(b - 1, nl, sl, il);
        if (dl <= 0) continue;
        const double* B = T->blk(nl, sl, il, n, s, i);
        dgemm('T', 'N', dR, dR, dl, g == 0 ? SQ2 : 0.5 * SQ2, B, dl, B, dl, 1.0, out, dR);
      }
    } else {
      const int dL = DIM(b, n, s, i), Is = self.bk->irrep(b);
      for (int g = 0; g < 3; g++) {
        const int nr = g == 0 ? n + 2 : n + 1, sr = g == 0 ? s : (g == 1 ? s - 1 : s + 1), ir = g == 0 ? i : (i ^ Is);
        const int dr = DIM(b + 1, nr, sr, ir);
        if (dr <= 0) continue;
        double a = SQ2;
        if (g >= 1) a *= 0.5 * (sr + 1.0) / (s + 1.0);
        const double* B = T->blk(n, s, i, nr, sr, ir);
        dgemm('N', 'T', dL, dL, dr, a, B, dL, B, dL, 1.0, out, dL);
      }
    }
  }
}

void F0_makenew(TOp& self, TOp* Lop, TensorT* T) {  // TensorF0.cpp:135-285
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, s, i), dD = DIM(b, n, s, idn);
    if (self.moving_right) {
      const int Is = self.bk->irrep(b - 1);
      for (int g = 0; g < 4; g++) {
        int nlu, slu, ilu, sld, ild;
        if (g <= 1) { nlu = n - 1; slu = g == 0 ? s - 1 : s + 1; ilu = i ^ Is; sld = s; ild = idn; }
        else { nlu = n - 2; slu = s; ilu = i; sld = g == 2 ? s - 1 : s + 1; ild = ilu ^ Lop->n_irrep; }
        const int dlu = DIM(b - 1, nlu, slu, ilu), dld = DIM(b - 1, nlu + 1, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double a = g <= 1 ? SQH : phase(s + 1 - sld) * std::sqrt(0.5 * (sld + 1.0) / (s + 1.0));
        sand_r(out, dU, dD, dlu, dld, a, T->blk(nlu, slu, ilu, n, s, i), Lop->blk(nlu, slu, ilu, nlu + 1, sld, ild), false,
               T->blk(nlu + 1, sld, ild, n, s, idn));
      }
    } else {
      const int Is = self.bk->irrep(b);
      for (int g = 0; g < 4; g++) {
        int nru, sru, iru, srd, ird;
        if (g <= 1) { nru = n; sru = s; iru = i; srd = g == 0 ? s - 1 : s + 1; ird = iru ^ Lop->n_irrep; }
        else { nru = n + 1; sru = g == 2 ? s - 1 : s + 1; iru = i ^ Is; srd = s; ird = idn; }
        const int dru = DIM(b + 1, nru, sru, iru), drd = DIM(b + 1, nru + 1, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double a = g <= 1 ? SQH * (srd + 1.0) / (s + 1.0) : phase(s - sru + 1) * std::sqrt(0.5 * (sru + 1.0) / (s + 1.0));
        sand_l(out, dU, dD, dru, drd, a, T->blk(n, s, i, nru, sru, iru), Lop->blk(nru, sru, iru, nru + 1, srd, ird), false,
               T->blk(n, s, idn, nru + 1, srd, ird));
      }
    }
  }
}

// --------------------------------------------------------------- TensorF1 --
void F1_makenew(TOp& self, TensorT* T) {  // TensorF1.cpp:54-123
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik];
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n, sd, i);
    if (self.moving_right) {
      const int il = i ^ self.bk->irrep(b - 1);
      for (int g = 0; g < 2; g++) {
        const int sl = su + (g == 0 ? -1 : 1);
        if (sl < 0 || std::abs(sd - sl) >= 2) continue;
        const int dl = DIM(b - 1, n - 1, sl, il);
        if (dl <= 0) continue;
        const double a = phase(sl + sd + 3) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 1, 2, su, sd, sl);
        dgemm('T', 'N', dU, dD, dl, a, T->blk(n - 1, sl, il, n, su, i), dl, T->blk(n - 1, sl, il, n, sd, i), dl, 1.0, out, dU);
      }
    } else {
      const int ir = i ^ self.bk->irrep(b);
      for (int g = 0; g < 2; g++) {
        const int sr = su + (g == 0 ? -1 : 1);
        if (sr < 0 || std::abs(sd - sr) >= 2) continue;
        const int dr = DIM(b + 1, n + 1, sr, ir);
        if (dr <= 0) continue;
        const double a = phase(sd + sr + 1) * std::sqrt(3.0 / (su + 1.0)) * (sr + 1) * wigner6j(1, 1, 2, su, sd, sr);
        dgemm('N', 'T', dU, dD, dr, a, T->blk(n, su, i, n + 1, sr, ir), dU, T->blk(n, sd, i, n + 1, sr, ir), dD, 1.0, out, dU);
      }
    }
  }
}

void F1_makenew(TOp& self, TOp* Lop, TensorT* T) {  // TensorF1.cpp:125-269
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n, sd, idn);
    if (self.moving_right) {
      const int Is = self.bk->irrep(b - 1);
      for (int g = 0; g < 4; g++) {
        int nlu, slu, ilu, sld, ild;
        if (g <= 1) { nlu = n - 1; slu = g == 0 ? su - 1 : su + 1; ilu = i ^ Is; sld = sd; ild = idn; }
        else { nlu = n - 2; slu = su; ilu = i; sld = g == 2 ? sd - 1 : sd + 1; ild = ilu ^ Lop->n_irrep; }
        const int dlu = DIM(b - 1, nlu, slu, ilu), dld = DIM(b - 1, nlu + 1, sld, ild);
        if (dlu <= 0 || dld <= 0 || std::abs(slu - sld) >= 2) continue;
        const double a = g <= 1 ? phase(slu + sd + 3) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 1, 2, su, sd, slu)
                                : phase(su + sd + 2) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 1, 2, su, sd, sld);
        sand_r(out, dU, dD, dlu, dld, a, T->blk(nlu, slu, ilu, n, su, i), Lop->blk(nlu, slu, ilu, nlu + 1, sld, ild), false,
               T->blk(nlu + 1, sld, ild, n, sd, idn));
      }
    } else {
      const int Is = self.bk->irrep(b);
      for (int g = 0; g < 4; g++) {
        int nru, sru, iru, srd, ird;
        if (g <= 1) { nru = n; sru = su; iru = i; srd = g == 0 ? sd - 1 : sd + 1; ird = iru ^ Lop->n_irrep; }
        else { nru = n + 1; sru = g == 2 ? su - 1 : su + 1; iru = i ^ Is; srd = sd; ird = idn; }
        const int dru = DIM(b + 1, nru, sru, iru), drd = DIM(b + 1, nru + 1, srd, ird);
        if (dru <= 0 || drd <= 0 || std::abs(sru - srd) >= 2) continue;
        double a;
        if (g <= 1) a = phase(sd + srd + 1) * std::sqrt(3.0 / (su + 1.0)) * (srd + 1) * wigner6j(1, 1, 2, su, sd, srd);
        else a = ((su % 2) != 0 ? -1 : 1) * std::sqrt(3.0 * (sru + 1.0) * (sd + 1.0) / (su + 1.0)) * wigner6j(1, 1, 2, su, sd, sru);
        sand_l(out, dU, dD, dru, drd, a, T->blk(n, su, i, nru, sru, iru), Lop->blk(nru, sru, iru, nru + 1, srd, ird), false,
               T->blk(n, sd, idn, nru + 1, srd, ird));
      }
    }
  }
}

// --------------------------------------------------------------- TensorS0 --
void S0_makenew(TOp& self, TensorT* T) {  // TensorS0.cpp:53-103
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, s, i), dD = DIM(b, n + 2, s, i);
    if (self.moving_right) {
      const int dl = DIM(b - 1, n, s, i);
      if (dl <= 0) continue;
      dgemm('T', 'N', dU, dD, dl, SQ2, T->blk(n, s, i, n, s, i), dl, T->blk(n, s, i, n + 2, s, i), dl, 1.0, out, dU);
    } else {
      const int dr = DIM(b + 1, n + 2, s, i);
      if (dr <= 0) continue;
      dgemm('N', 'T', dU, dD, dr, SQ2, T->blk(n, s, i, n + 2, s, i), dU, T->blk(n + 2, s, i, n + 2, s, i), dD, 1.0, out, dU);
    }
  }
}

void S0_makenew(TOp& self, TOp* Lop, TensorT* T) {  // TensorS0.cpp:105-246
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, s, i), dD = DIM(b, n + 2, s, idn);
    if (self.moving_right) {
      const int Is = self.bk->irrep(b - 1);
      for (int g = 0; g < 4; g++) {
        int nlu, slu, ilu, sld, ild;
        if (g <= 1) { nlu = n; slu = s; ilu = i; sld = g == 0 ? s - 1 : s + 1; ild = ilu ^ Lop->n_irrep; }
        else { nlu = n - 1; slu = g == 2 ? s - 1 : s + 1; ilu = i ^ Is; sld = s; ild = idn; }
        const int dlu = DIM(b - 1, nlu, slu, ilu), dld = DIM(b - 1, nlu + 1, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double a = g <= 1 ? phase(s - sld + 1) * std::sqrt(0.5 * (sld + 1.0) / (s + 1.0)) : -SQH;
        sand_r(out, dU, dD, dlu, dld, a, T->blk(nlu, slu, ilu, n, s, i), Lop->blk(nlu, slu, ilu, nlu + 1, sld, ild), false,
               T->blk(nlu + 1, sld, ild, n + 2, s, idn));
      }
    } else {
      const int Is = self.bk->irrep(b);
      for (int g = 0; g < 4; g++) {
        int nru, sru, iru, srd, ird;
        if (g <= 1) { nru = n + 1; sru = g == 0 ? s - 1 : s + 1; iru = i ^ Is; srd = s; ird = idn; }
        else { nru = n + 2; sru = s; iru = i; srd = g == 2 ? s - 1 : s + 1; ird = i ^ Lop->n_irrep; }
        const int dru = DIM(b + 1, nru, sru, iru), drd = DIM(b + 1, nru + 1, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double a = g <= 1 ? phase(s - sru + 1) * std::sqrt(0.5 * (sru + 1.0) / (s + 1.0)) : -SQH * (srd + 1.0) / (s + 1.0);
        sand_l(out, dU, dD, dru, drd, a, T->blk(n, s, i, nru, sru, iru), Lop->blk(nru, sru, iru, nru + 1, srd, ird), false,
               T->blk(n + 2, s, idn, nru + 1, srd, ird));
      }
    }
  }
}

// --------------------------------------------------------------- TensorS1 --
void S1_makenew(TOp& self, TOp* Lop, TensorT* T) {  // TensorS1.cpp:48-195
  self.clear();
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 2, sd, idn);
    if (self.moving_right) {
      const int Is = self.bk->irrep(b - 1);
      for (int g = 0; g < 4; g++) {
        int nlu, slu, ilu, sld, ild;
        if (g <= 1) { nlu = n; slu = su; ilu = i; sld = g == 0 ? sd - 1 : sd + 1; ild = ilu ^ Lop->n_irrep; }
        else { nlu = n - 1; slu = g == 2 ? su - 1 : su + 1; ilu = i ^ Is; sld = sd; ild = idn; }
        const int dlu = DIM(b - 1, nlu, slu, ilu), dld = DIM(b - 1, nlu + 1, sld, ild);
        if (dlu <= 0 || dld <= 0 || std::abs(slu - sld) >= 2) continue;
        const double a = g <= 1 ? phase(su + sd + 2) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 1, 2, su, sd, sld)
                                : phase(slu + sd + 1) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 1, 2, su, sd, slu);
        sand_r(out, dU, dD, dlu, dld, a, T->blk(nlu, slu, ilu, n, su, i), Lop->blk(nlu, slu, ilu, nlu + 1, sld, ild), false,
               T->blk(nlu + 1, sld, ild, n + 2, sd, idn));
      }
    } else {
      const int Is = self.bk->irrep(b);
      for (int g = 0; g < 4; g++) {
        int nru, sru, iru, srd, ird;
        if (g <= 1) { nru = n + 1; sru = g == 0 ? su - 1 : su + 1; iru = i ^ Is; srd = sd; ird = idn; }
        else { nru = n + 2; sru = su; iru = i; srd = g == 2 ? sd - 1 : sd + 1; ird = i ^ Lop->n_irrep; }
        const int dru = DIM(b + 1, nru, sru, iru), drd = DIM(b + 1, nru + 1, srd, ird);
        if (dru <= 0 || drd <= 0 || std::abs(srd - sru) >= 2) continue;
        const double a = g <= 1 ? phase(su + sd + 2) * std::sqrt(3.0 * (sru + 1)) * wigner6j(1, 1, 2, su, sd, sru)
                                : phase(su + srd + 1) * std::sqrt(3.0 / (sd + 1.0)) * (srd + 1) * wigner6j(1, 1, 2, su, sd, srd);
        sand_l(out, dU, dD, dru, drd, a, T->blk(n, su, i, nru, sru, iru), Lop->blk(nru, sru, iru, nru + 1, srd, ird), false,
               T->blk(n + 2, sd, idn, nru + 1, srd, ird));
      }
    }
  }
}

// ---------------------------------------------------------------- TensorQ --
void Q_add_simple(TOp& self, const Problem* P, int site, TensorT* T) {  // TensorQ.cpp:45-100
  if (self.bk->irrep(T->index) != self.n_irrep) return;
  const int b = self.index;
  const double mx = self.moving_right ? P->g(b - 1, b - 1, b - 1, site) : P->g(site, b, b, b);
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    if (self.moving_right) {
      const int dl = DIM(b - 1, n - 1, sd, idn);
      if (dl <= 0) continue;
      const double a = phase(sd + 1 - su) * mx * std::sqrt((su + 1.0) / (sd + 1.0));
      dgemm('T', 'N', dU, dD, dl, a, T->blk(n - 1, sd, idn, n, su, i), dl, T->blk(n - 1, sd, idn, n + 1, sd, idn), dl, 1.0, out, dU);
    } else {
      const int dr = DIM(b + 1, n + 2, su, i);
      if (dr <= 0) continue;
      const double a = phase(su + 1 - sd) * mx * std::sqrt((su + 1.0) / (sd + 1.0));
      dgemm('N', 'T', dU, dD, dr, a, T->blk(n, su, i, n + 2, su, i), dU, T->blk(n + 1, sd, idn, n + 2, su, i), dD, 1.0, out, dU);
    }
  }
}

void Q_add_L(TOp& self, const Problem* P, int site, std::vector<std::unique_ptr<TOp>>& Lprev, TensorT* T) {  // TensorQ.cpp:102-390
  const int b = self.index, L = P->L;
  // orbital -> previous L tensor (moving right: sites 0..b-2; moving left: sites b+1..L-1)
  auto Lof = [&](int loca) -> TOp* { return self.moving_right ? Lprev[b - 2 - loca].get() : Lprev[loca - b - 1].get(); };
  const int lo = self.moving_right ? 0 : b + 1, hi = self.moving_right ? b - 1 : L;
  bool any = false;
  for (int loca = lo; loca < hi; loca++) if (Lof(loca)->n_irrep == self.n_irrep) any = true;
  if (!any) return;
  std::vector<double> mem;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    if (self.moving_right) {
      {  // case 1
        const int dlu = DIM(b - 1, n, su, i), dld = DIM(b - 1, n - 1, sd, idn);
        if (dlu > 0 && dld > 0) {
          mem.assign((size_t)dlu * dld, 0.0);  // stored dld x dlu
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            const double* BL = Lof(loca)->blk(n - 1, sd, idn, n, su, i);
            const double a = P->g(loca, site, b - 1, b - 1);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          const double a = phase(sd + 1 - su) * std::sqrt((su + 1.0) / (sd + 1.0));
          sand_r(out, dU, dD, dlu, dld, a, T->blk(n, su, i, n, su, i), mem.data(), true, T->blk(n - 1, sd, idn, n + 1, sd, idn));
        }
      }
      {  // case 2
        const int dlu = DIM(b - 1, n - 2, su, i), dld = DIM(b - 1, n - 1, sd, idn);
        if (dlu > 0 && dld > 0) {
          mem.assign((size_t)dlu * dld, 0.0);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            const double* BL = Lof(loca)->blk(n - 2, su, i, n - 1, sd, idn);
            const double a = 2 * P->g(loca, b - 1, site, b - 1) - P->g(loca, b - 1, b - 1, site);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n - 2, su, i, n, su, i), mem.data(), false, T->blk(n - 1, sd, idn, n + 1, sd, idn));
        }
      }
      for (int slu = su - 1; slu <= su + 1; slu += 2)  // case 3
        for (int sld = sd - 1; sld <= sd + 1; sld += 2) {
          if (sld < 0 || slu < 0 || std::abs(sld - slu) >= 2) continue;
          const int ilu = i ^ self.bk->irrep(b - 1), ild = idn ^ self.bk->irrep(b - 1);
          const int dlu = DIM(b - 1, n - 1, slu, ilu), dld = DIM(b - 1, n, sld, ild);
          if (dlu <= 0 || dld <= 0) continue;
          const double f = phase(su + sld) * std::sqrt((sld + 1) * (su + 1.0)) * wigner6j(su, sd, 1, sld, slu, 1);
          mem.assign((size_t)dlu * dld, 0.0);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            const double* BL = Lof(loca)->blk(n - 1, slu, ilu, n, sld, ild);
            double a = f * P->g(loca, b - 1, site, b - 1);
            if (sld == su) a += P->g(loca, b - 1, b - 1, site);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n - 1, slu, ilu, n, su, i), mem.data(), false, T->blk(n, sld, ild, n + 1, sd, idn));
        }
    } else {
      {  // case 1
        const int dru = DIM(b + 1, n + 2, su, i), drd = DIM(b + 1, n + 1, sd, idn);
        if (dru > 0 && drd > 0) {
          mem.assign((size_t)dru * drd, 0.0);  // stored drd x dru
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            const double* BL = Lof(loca)->blk(n + 1, sd, idn, n + 2, su, i);
            const double a = P->g(site, loca, b, b);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          const double a = phase(su + 1 - sd) * std::sqrt((su + 1.0) / (sd + 1.0));
          sand_l(out, dU, dD, dru, drd, a, T->blk(n, su, i, n + 2, su, i), mem.data(), true, T->blk(n + 1, sd, idn, n + 1, sd, idn));
        }
      }
      {  // case 2
        const int dru = DIM(b + 1, n + 2, su, i), drd = DIM(b + 1, n + 3, sd, idn);
        if (dru > 0 && drd > 0) {
          mem.assign((size_t)dru * drd, 0.0);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            const double* BL = Lof(loca)->blk(n + 2, su, i, n + 3, sd, idn);
            const double a = 2 * P->g(site, b, loca, b) - P->g(site, b, b, loca);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 2, su, i), mem.data(), false, T->blk(n + 1, sd, idn, n + 3, sd, idn));
        }
      }
      for (int sru = su - 1; sru <= su + 1; sru += 2)  // case 3
        for (int srd = sd - 1; srd <= sd + 1; srd += 2) {
          if (srd < 0 || sru < 0 || std::abs(srd - sru) >= 2) continue;
          const int iru = i ^ self.bk->irrep(b), ird = idn ^ self.bk->irrep(b);
          const int dru = DIM(b + 1, n + 1, sru, iru), drd = DIM(b + 1, n + 2, srd, ird);
          if (dru <= 0 || drd <= 0) continue;
          const double f1 = phase(sd + sru) * std::sqrt((sru + 1.0) / (sd + 1.0)) * (srd + 1) * wigner6j(su, sd, 1, srd, sru, 1);
          const double f2 = (srd + 1.0) / (sd + 1.0);
          mem.assign((size_t)dru * drd, 0.0);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            const double* BL = Lof(loca)->blk(n + 1, sru, iru, n + 2, srd, ird);
            double a = f1 * P->g(site, b, loca, b);
            if (sru == sd) a += f2 * P->g(site, b, b, loca);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 1, sru, iru), mem.data(), false, T->blk(n + 1, sd, idn, n + 2, srd, ird));
        }
    }
  }
}

void Q_add_AB(TOp& self, TOp* A, TOp* B, TensorT* T) {  // TensorQ.cpp:392-575
  const int b = self.index;
  std::vector<double> mem;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    auto combine = [&](double fA, const double* BA, double fB, const double* BB, size_t sz, double* alpha) -> const double* {
      if (BA) {
        mem.resize(sz);
        for (size_t c = 0; c < sz; c++) mem[c] = fA * BA[c] + fB * BB[c];
        *alpha = 1.0;
        return mem.data();
      }
      *alpha = fB;
      return BB;
    };
    if (self.moving_right) {
      const int ilu = i ^ self.bk->irrep(b - 1);
      for (int slu = su - 1; slu <= su + 1; slu += 2) {  // case 1
        const int dlu = DIM(b - 1, n - 1, slu, ilu), dld = DIM(b - 1, n + 1, sd, idn);
        if (dlu <= 0 || dld <= 0) continue;
        const double fB = phase(slu + sd + 2) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 2, 1, sd, su, slu);
        const double fA = phase(sd + 1 - su) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        double alpha;
        const double* m = combine(fA, slu == sd ? A->blk(n - 1, slu, ilu, n + 1, sd, idn) : nullptr, fB,
                                  B->blk(n - 1, slu, ilu, n + 1, sd, idn), (size_t)dlu * dld, &alpha);
        sand_r(out, dU, dD, dlu, dld, alpha, T->blk(n - 1, slu, ilu, n, su, i), m, false, T->blk(n + 1, sd, idn, n + 1, sd, idn));
      }
      const int ild = idn ^ self.bk->irrep(b - 1);
      for (int sld = sd - 1; sld <= sd + 1; sld += 2) {  // case 2
        const int dlu = DIM(b - 1, n - 2, su, i), dld = DIM(b - 1, n, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double fB = phase(su + sd + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 2, 1, su, sd, sld);
        double alpha;
        const double* m = combine(-SQH, sld == su ? A->blk(n - 2, su, i, n, sld, ild) : nullptr, fB,
                                  B->blk(n - 2, su, i, n, sld, ild), (size_t)dlu * dld, &alpha);
        sand_r(out, dU, dD, dlu, dld, alpha, T->blk(n - 2, su, i, n, su, i), m, false, T->blk(n, sld, ild, n + 1, sd, idn));
      }
    } else {
      const int ird = idn ^ self.bk->irrep(b);
      for (int srd = sd - 1; srd <= sd + 1; srd += 2) {  // case 1
        const int dru = DIM(b + 1, n, su, i), drd = DIM(b + 1, n + 2, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double fB = phase(srd + su + 2) * std::sqrt(3.0 / (sd + 1.0)) * (srd + 1) * wigner6j(1, 1, 2, su, srd, sd);
        const double fA = phase(su + 1 - sd) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        double alpha;
        const double* m = combine(fA, srd == su ? A->blk(n, su, i, n + 2, srd, ird) : nullptr, fB,
                                  B->blk(n, su, i, n + 2, srd, ird), (size_t)dru * drd, &alpha);
        sand_l(out, dU, dD, dru, drd, alpha, T->blk(n, su, i, n, su, i), m, false, T->blk(n + 1, sd, idn, n + 2, srd, ird));
      }
      const int iru = i ^ self.bk->irrep(b);
      for (int sru = su - 1; sru <= su + 1; sru += 2) {  // case 2
        const int dru = DIM(b + 1, n + 1, sru, iru), drd = DIM(b + 1, n + 3, sd, idn);
        if (dru <= 0 || drd <= 0) continue;
        const double fB = phase(su + sd + 1) * std::sqrt(3.0 * (sru + 1)) * wigner6j(1, 1, 2, sru, sd, su);
        double alpha;
        const double* m = combine(-SQH, sru == sd ? A->blk(n + 1, sru, iru, n + 3, sd, idn) : nullptr, fB,
                                  B->blk(n + 1, sru, iru, n + 3, sd, idn), (size_t)dru * drd, &alpha);
        sand_l(out, dU, dD, dru, drd, alpha, T->blk(n, su, i, n + 1, sru, iru), m, false, T->blk(n + 1, sd, idn, n + 3, sd, idn));
      }
    }
  }
}

void Q_add_CD(TOp& self, TOp* C, TOp* D, TensorT* T) {  // TensorQ.cpp:577-785
  const int b = self.index;
  std::vector<double> mem;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    double* out = self.data.data() + self.k2i[ik];
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    auto build = [&](double fD, const double* BD, double fC, const double* BC, size_t sz) {
      mem.resize(sz);
      for (size_t c = 0; c < sz; c++) mem[c] = fD * BD[c];
      if (BC) for (size_t c = 0; c < sz; c++) mem[c] += fC * BC[c];
    };
    if (self.moving_right) {
      const int ild = idn ^ self.bk->irrep(b - 1);
      for (int sld = sd - 1; sld <= sd + 1; sld += 2) {  // case 1
        const int dlu = DIM(b - 1, n, su, i), dld = DIM(b - 1, n, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double fD = phase(su + sd + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 2, 1, su, sd, sld);
        build(fD, D->blk(n, su, i, n, sld, ild), SQH, sld == su ? C->blk(n, su, i, n, sld, ild) : nullptr, (size_t)dlu * dld);
        sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n, su, i, n, su, i), mem.data(), false, T->blk(n, sld, ild, n + 1, sd, idn));
      }
      const int ilu = i ^ self.bk->irrep(b - 1);
      for (int slu = su - 1; slu <= su + 1; slu += 2) {  // case 2
        const int dlu = DIM(b - 1, n - 1, slu, ilu), dld = DIM(b - 1, n - 1, sd, idn);
        if (dlu <= 0 || dld <= 0) continue;
        const double fD = phase(slu + sd) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 2, 1, sd, su, slu);
        const double fC = phase(sd + 1 - su) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        build(fD, D->blk(n - 1, slu, ilu, n - 1, sd, idn), fC, slu == sd ? C->blk(n - 1, slu, ilu, n - 1, sd, idn) : nullptr, (size_t)dlu * dld);
        sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n - 1, slu, ilu, n, su, i), mem.data(), false, T->blk(n - 1, sd, idn, n + 1, sd, idn));
      }
    } else {
      const int iru = i ^ self.bk->irrep(b);
      for (int sru = su - 1; sru <= su + 1; sru += 2) {  // case 1
        const int dru = DIM(b + 1, n + 1, sru, iru), drd = DIM(b + 1, n + 1, sd, idn);
        if (dru <= 0 || drd <= 0) continue;
        const double fD = phase(su + sru + 3) * std::sqrt(3.0 / (sd + 1.0)) * (sru + 1) * wigner6j(1, 1, 2, sru, sd, su);
        build(fD, D->blk(n + 1, sru, iru, n + 1, sd, idn), SQH, sru == sd ? C->blk(n + 1, sru, iru, n + 1, sd, idn) : nullptr, (size_t)dru * drd);
        sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 1, sru, iru), mem.data(), false, T->blk(n + 1, sd, idn, n + 1, sd, idn));
      }
      const int ird = idn ^ self.bk->irrep(b);
      for (int srd = sd - 1; srd <= sd + 1; srd += 2) {  // case 2
        const int dru = DIM(b + 1, n + 2, su, i), drd = DIM(b + 1, n + 2, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double fD = (((srd + 1) % 2) != 0 ? -1 : 1) * std::sqrt(3.0 * (srd + 1.0) * (su + 1.0) / (sd + 1.0)) * wigner6j(1, 1, 2, su, srd, sd);
        const double fC = phase(su + 1 - sd) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        build(fD, D->blk(n + 2, su, i, n + 2, srd, ird), fC, srd == su ? C->blk(n + 2, su, i, n + 2, srd, ird) : nullptr, (size_t)dru * drd);
        sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 2, su, i), mem.data(), false, T->blk(n + 1, sd, idn, n + 2, srd, ird));
      }
    }
  }
}

// ---------------------------------------------------------------- TensorX --
static void X_makenew(TOp& self, const Problem* P, TensorT* T, int ik) {  // TensorX.cpp:127-167
  const int b = self.index;
  const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
  double* out = self.data.data() + self.k2i[ik];
  const int sz = self.k2i[ik + 1] - self.k2i[ik];
  if (self.moving_right) {
    const int dR = DIM(b, n, s, i), dL = DIM(b - 1, n - 2, s, i);
    const double a = P->g(b - 1, b - 1, b - 1, b - 1);
    if (dL > 0 && std::fabs(a) > 0.0) {
      const double* B = T->blk(n - 2, s, i, n, s, i);
      dgemm('T', 'N', dR, dR, dL, a, B, dL, B, dL, 0.0, out, dR);
    } else std::fill(out, out + sz, 0.0);
  } else {
    const int dL = DIM(b, n, s, i), dR = DIM(b + 1, n + 2, s, i);
    const double a = P->g(b, b, b, b);
    if (dR > 0 && std::fabs(a) > 0.0) {
      const double* B = T->blk(n, s, i, n + 2, s, i);
      dgemm('N', 'T', dL, dL, dR, a, B, dL, B, dL, 0.0, out, dL);
    } else std::fill(out, out + sz, 0.0);
  }
}

void X_update(TOp& self, const Problem* P, TensorT* T) {  // TensorX.cpp:44-56
  for (int ik = 0; ik < self.nk(); ik++) X_makenew(self, P, T, ik);
}

void X_update(TOp& self, const Problem* P, TensorT* T, std::vector<std::unique_ptr<TOp>>& Lprev, TOp* Xprev,
              TOp* Qprev, TOp* A, TOp* C, TOp* D) {  // TensorX.cpp:58-125 and 169-645
  const int b = self.index, L = P->L;
  // The generic update of Xprev, evaluated block by block.
  TOp tmp;
  tmp.init(self.index, 0, 0, 0, self.moving_right, true, false, self.bk);
  const bool other = self.moving_right ? (b > 1) : (b < L - 1);
  if (other) tmp.update(Xprev, T, T);
  std::vector<double> acc, mem;
  for (int ik = 0; ik < self.nk(); ik++) {
    X_makenew(self, P, T, ik);
    if (!other) continue;
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
    double* out = self.data.data() + self.k2i[ik];
    const int sz = self.k2i[ik + 1] - self.k2i[ik];
    for (int c = 0; c < sz; c++) out[c] += tmp.data[tmp.k2i[ik] + c];
    if (self.moving_right) {
      const int dR = DIM(b, n, s, i), Is = self.bk->irrep(b - 1);
      // Q and L term (TensorX.cpp:169-263): mem + mem^T
      acc.assign((size_t)dR * dR, 0.0);
      for (int g = 0; g < 4; g++) {
        int nlu, slu, ilu, nld, sld, ild;
        if (g <= 1) { nlu = n - 1; slu = g == 0 ? s - 1 : s + 1; ilu = i ^ Is; nld = n; sld = s; ild = i; }
        else { nlu = n - 2; slu = s; ilu = i; nld = n - 1; sld = g == 2 ? s - 1 : s + 1; ild = i ^ Is; }
        const int dlu = DIM(b - 1, nlu, slu, ilu), dld = DIM(b - 1, nld, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double* BQ = Qprev->blk(nlu, slu, ilu, nld, sld, ild);
        double f = 1.0;
        const double* ptr = BQ;
        if (g >= 2) {
          f = phase(s + 1 - sld) * std::sqrt((sld + 1.0) / (s + 1.0));
          mem.assign(BQ, BQ + (size_t)dlu * dld);
          for (int loca = 0; loca < b - 1; loca++) {
            if (self.bk->irrep(b - 1) != self.bk->irrep(loca)) continue;
            const double a = P->g(loca, b - 1, b - 1, b - 1);
            const double* BL = Lprev[b - 2 - loca]->blk(nlu, slu, ilu, nld, sld, ild);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          ptr = mem.data();
        }
        sand_r(acc.data(), dR, dR, dlu, dld, f, T->blk(nlu, slu, ilu, n, s, i), ptr, false, T->blk(nld, sld, ild, n, s, i));
      }
      for (int r = 0; r < dR; r++) for (int c = 0; c < dR; c++) out[r + dR * c] += acc[r + dR * c] + acc[c + dR * r];
      {  // A term (TensorX.cpp:361-395)
        const int dlu = DIM(b - 1, n - 2, s, i), dld = DIM(b - 1, n, s, i);
        if (dlu > 0 && dld > 0) {
          acc.assign((size_t)dR * dR, 0.0);
          sand_r(acc.data(), dR, dR, dlu, dld, SQ2, T->blk(n - 2, s, i, n, s, i), A->blk(n - 2, s, i, n, s, i), false, T->blk(n, s, i, n, s, i));
          for (int r = 0; r < dR; r++) for (int c = 0; c < dR; c++) out[r + dR * c] += acc[r + dR * c] + acc[c + dR * r];
        }
      }
      for (int g = 0; g < 3; g++) {  // C term (TensorX.cpp:433-467)
        const int nl = g < 2 ? n - 1 : n - 2, sl = g == 0 ? s - 1 : (g == 1 ? s + 1 : s), il = g < 2 ? (i ^ Is) : i;
        const int dl = DIM(b - 1, nl, sl, il);
        if (dl <= 0) continue;
        const double* BT = T->blk(nl, sl, il, n, s, i);
        sand_r(out, dR, dR, dl, dl, g < 2 ? SQH : SQ2, BT, C->blk(nl, sl, il, nl, sl, il), true, BT);
      }
      {  // D term (TensorX.cpp:539-590)
        const int il = i ^ Is, nl = n - 1;
        for (int g = 0; g < 4; g++) {
          const int slu = (g == 0 || g == 2) ? s - 1 : s + 1, sld = (g <= 1) ? s - 1 : s + 1;
          const int dlu = DIM(b - 1, nl, slu, il), dld = DIM(b - 1, nl, sld, il);
          if (dlu <= 0 || dld <= 0) continue;
          const double f = phase(sld + s + 1) * std::sqrt(3.0 * (slu + 1)) * wigner6j(1, 1, 2, slu, sld, s);
          sand_r(out, dR, dR, dlu, dld, f, T->blk(nl, slu, il, n, s, i), D->blk(nl, sld, il, nl, slu, il), true, T->blk(nl, sld, il, n, s, i));
        }
      }
    } else {
      const int dL = DIM(b, n, s, i), Is = self.bk->irrep(b);
      acc.assign((size_t)dL * dL, 0.0);
      for (int g = 0; g < 4; g++) {  // Q and L term (TensorX.cpp:265-359)
        int nru, sru, iru, nrd, srd, ird;
        if (g <= 1) { nru = n; sru = s; iru = i; nrd = n + 1; srd = g == 0 ? s - 1 : s + 1; ird = i ^ Is; }
        else { nru = n + 1; sru = g == 2 ? s - 1 : s + 1; iru = i ^ Is; nrd = n + 2; srd = s; ird = i; }
        const int dru = DIM(b + 1, nru, sru, iru), drd = DIM(b + 1, nrd, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double* BQ = Qprev->blk(nru, sru, iru, nrd, srd, ird);
        double f;
        const double* ptr = BQ;
        if (g <= 1) f = (srd + 1.0) / (s + 1.0);
        else {
          f = phase(s + 1 - sru) * std::sqrt((sru + 1.0) / (s + 1.0));
          mem.assign(BQ, BQ + (size_t)dru * drd);
          for (int loca = b + 1; loca < L; loca++) {
            if (self.bk->irrep(b) != self.bk->irrep(loca)) continue;
            const double a = P->g(b, b, b, loca);
            const double* BL = Lprev[loca - b - 1]->blk(nru, sru, iru, nrd, srd, ird);
            for (size_t c = 0; c < mem.size(); c++) mem[c] += a * BL[c];
          }
          ptr = mem.data();
        }
        sand_l(acc.data(), dL, dL, dru, drd, f, T->blk(n, s, i, nru, sru, iru), ptr, false, T->blk(n, s, i, nrd, srd, ird));
      }
      for (int r = 0; r < dL; r++) for (int c = 0; c < dL; c++) out[r + dL * c] += acc[r + dL * c] + acc[c + dL * r];
      {  // A term (TensorX.cpp:397-431)
        const int dru = DIM(b + 1, n, s, i), drd = DIM(b + 1, n + 2, s, i);
        if (dru > 0 && drd > 0) {
          acc.assign((size_t)dL * dL, 0.0);
          sand_l(acc.data(), dL, dL, dru, drd, SQ2, T->blk(n, s, i, n, s, i), A->blk(n, s, i, n + 2, s, i), false, T->blk(n, s, i, n + 2, s, i));
          for (int r = 0; r < dL; r++) for (int c = 0; c < dL; c++) out[r + dL * c] += acc[r + dL * c] + acc[c + dL * r];
        }
      }
      for (int g = 0; g < 3; g++) {  // C term (TensorX.cpp:469-537)
        const int nr = g < 2 ? n + 1 : n + 2, sr = g == 0 ? s - 1 : (g == 1 ? s + 1 : s), ir = g < 2 ? (i ^ Is) : i;
        const int dr = DIM(b + 1, nr, sr, ir);
        if (dr <= 0) continue;
        const double* BT = T->blk(n, s, i, nr, sr, ir);
        sand_l(out, dL, dL, dr, dr, g < 2 ? SQH * (sr + 1.0) / (s + 1.0) : SQ2, BT, C->blk(nr, sr, ir, nr, sr, ir), true, BT);
      }
      {  // D term (TensorX.cpp:592-645)
        const int nr = n + 1, ir = i ^ Is;
        for (int g = 0; g < 4; g++) {
          const int sru = (g == 0 || g == 2) ? s - 1 : s + 1, srd = (g <= 1) ? s - 1 : s + 1;
          const int dru = DIM(b + 1, nr, sru, ir), drd = DIM(b + 1, nr, srd, ir);
          if (dru <= 0 || drd <= 0) continue;
          const double f = phase(s + srd + 3) * std::sqrt(3.0 * (sru + 1)) * ((srd + 1.0) / (s + 1.0)) * wigner6j(1, 1, 2, sru, srd, s);
          sand_l(out, dL, dL, dru, drd, f, T->blk(n, s, i, nr, sru, ir), D->blk(nr, srd, ir, nr, sru, ir), true, T->blk(n, s, i, nr, srd, ir));
        }
      }
    }
  }
}

// ------------------------------------------------------- orchestration -----
void Operators::init(const Problem* P_, const Bookkeeper* bk_) {
  P = P_; bk = bk_; L = P->L;
  ops.clear();
  ops.resize(L - 1);
}

static std::unique_ptr<TOp> mk(int b, int tj, int ne, int ir, bool mr, bool pl, bool jw, const Bookkeeper* bk) {
  auto t = std::make_unique<TOp>();
  t->init(b, tj, ne, ir, mr, pl, jw, bk);
  return t;
}

void Operators::allocate(int cnt, bool mr) {  // DMRGoperators.cpp:909-1145
  OpSet& o = ops[cnt];
  o = OpSet();
  o.dir = mr ? 1 : 2;
  const int b = cnt + 1;
  auto ir = [&](int s) { return bk->irrep(s); };
  const int nL = mr ? cnt + 1 : L - 1 - cnt;  // L, F, S extent (sites on the block side)
  const int nC = mr ? L - 1 - cnt : cnt + 1;  // A, B, C, D, Q extent (sites on the far side)
  for (int k = 0; k < nL; k++) o.L.push_back(mk(b, 1, 1, ir(mr ? cnt - k : cnt + 1 + k), mr, true, true, bk));
  o.F0.resize(nL); o.F1.resize(nL); o.S0.resize(nL); o.S1.resize(nL);
  for (int k = 0; k < nL; k++)
    for (int m = 0; m < nL - k; m++) {
      const int Ip = mr ? (ir(cnt - k - m) ^ ir(cnt - m)) : (ir(cnt + 1 + m) ^ ir(cnt + 1 + k + m));
      o.F0[k].push_back(mk(b, 0, 0, Ip, mr, true, false, bk));
      o.F1[k].push_back(mk(b, 2, 0, Ip, mr, mr, false, bk));
      o.S0[k].push_back(mk(b, 0, 2, Ip, mr, true, false, bk));
      o.S1[k].push_back(k > 0 ? mk(b, 2, 2, Ip, mr, true, false, bk) : nullptr);
    }
  o.A.resize(nC); o.B.resize(nC); o.C.resize(nC); o.D.resize(nC);
  for (int k = 0; k < nC; k++)
    for (int m = 0; m < nC - k; m++) {
      const int Id = mr ? (ir(cnt + 1 + k + m) ^ ir(cnt + 1 + m)) : (ir(cnt - k - m) ^ ir(cnt - m));
      o.A[k].push_back(mk(b, 0, 2, Id, mr, true, false, bk));
      o.B[k].push_back(k > 0 ? mk(b, 2, 2, Id, mr, true, false, bk) : nullptr);
      o.C[k].push_back(mk(b, 0, 0, Id, mr, true, false, bk));
      o.D[k].push_back(mk(b, 2, 0, Id, mr, mr, false, bk));
    }
  for (int k = 0; k < nC; k++) {
    const int site = mr ? cnt + 1 + k : cnt - k;
    o.Q.push_back(mk(b, 1, 1, ir(site), mr, true, true, bk));
    o.Qsite.push_back(site);
  }
  o.X = mk(b, 0, 0, 0, mr, true, false, bk);
}

void Operators::update_moving_right(int index, std::vector<TensorT>& mps) {  // DMRGoperators.cpp:243-574
  allocate(index, true);
  OpSet& o = ops[index];
  TensorT* T = &mps[index];
  OpSet* prev = index > 0 ? &ops[index - 1] : nullptr;
  if (index > 0) assert(prev->dir == 1);
  for (int k = 0; k <= index; k++) {
    if (k == 0) L_create(*o.L[0], T);
    else o.L[k]->update(prev->L[k - 1].get(), T, T);
  }
  for (int k = 0; k <= index; k++)
    for (int m = 0; m <= index - k; m++) {
      if (m == 0) {
        if (k == 0) { F0_makenew(*o.F0[0][0], T); F1_makenew(*o.F1[0][0], T); S0_makenew(*o.S0[0][0], T); }
        else {
          TOp* Lp = prev->L[k - 1].get();
          F0_makenew(*o.F0[k][0], Lp, T); F1_makenew(*o.F1[k][0], Lp, T);
          S0_makenew(*o.S0[k][0], Lp, T); S1_makenew(*o.S1[k][0], Lp, T);
        }
      } else {
        o.F0[k][m]->update(prev->F0[k][m - 1].get(), T, T);
        o.F1[k][m]->update(prev->F1[k][m - 1].get(), T, T);
        o.S0[k][m]->update(prev->S0[k][m - 1].get(), T, T);
        if (k > 0) o.S1[k][m]->update(prev->S1[k][m - 1].get(), T, T);
      }
    }
  const int k2 = L - 1 - index;
  for (int k = 0; k < k2; k++)
    for (int m = 0; m < k2 - k; m++) {
      const int s1 = index + 1 + m, s2 = index + 1 + k + m;
      const int ip = bk->irrep(s1) ^ bk->irrep(s2);
      if (index == 0) {
        o.A[k][m]->clear(); if (k > 0) o.B[k][m]->clear();
        o.C[k][m]->clear(); o.D[k][m]->clear();
      } else {
        o.A[k][m]->update(prev->A[k][m + 1].get(), T, T);
        if (k > 0) o.B[k][m]->update(prev->B[k][m + 1].get(), T, T);
        o.C[k][m]->update(prev->C[k][m + 1].get(), T, T);
        o.D[k][m]->update(prev->D[k][m + 1].get(), T, T);
      }
      for (int num = 0; num <= index; num++) {
        if (ip != o.S0[num][0]->n_irrep) continue;
        double a = P->g(index - num, index, s1, s2);
        if (k == 0 && num == 0) a *= 0.5;
        if (k > 0 && num > 0) a += P->g(index - num, index, s2, s1);
        o.A[k][m]->axpy(a, o.S0[num][0].get());
        if (num > 0 && k > 0) o.B[k][m]->axpy(P->g(index - num, index, s1, s2) - P->g(index - num, index, s2, s1), o.S1[num][0].get());
        o.C[k][m]->axpy(2 * P->g(index - num, s1, index, s2) - P->g(index - num, s1, s2, index), o.F0[num][0].get());
        o.D[k][m]->axpy(-P->g(index - num, s1, s2, index), o.F1[num][0].get());
        if (num > 0) {
          o.C[k][m]->axpy_transpose_CD(2 * P->g(index - num, s2, index, s1) - P->g(index - num, s2, s1, index), o.F0[num][0].get());
          o.D[k][m]->axpy_transpose_CD(-P->g(index - num, s2, s1, index), o.F1[num][0].get());
        }
      }
    }
  for (int k = 0; k < k2; k++) {
    TOp& q = *o.Q[k];
    const int site = o.Qsite[k];
    if (index == 0) { q.clear(); Q_add_simple(q, P, site, T); }
    else {
      q.update(prev->Q[k + 1].get(), T, T);
      Q_add_simple(q, P, site, T);
      Q_add_L(q, P, site, prev->L, T);
      Q_add_AB(q, prev->A[k + 1][0].get(), prev->B[k + 1][0].get(), T);
      Q_add_CD(q, prev->C[k + 1][0].get(), prev->D[k + 1][0].get(), T);
    }
  }
  if (index == 0) X_update(*o.X, P, T);
  else X_update(*o.X, P, T, prev->L, prev->X.get(), prev->Q[0].get(), prev->A[0][0].get(), prev->C[0][0].get(), prev->D[0][0].get());
}

void Operators::update_moving_left(int index, std::vector<TensorT>& mps) {  // DMRGoperators.cpp:576-907
  allocate(index, false);
  OpSet& o = ops[index];
  TensorT* T = &mps[index + 1];
  OpSet* prev = index < L - 2 ? &ops[index + 1] : nullptr;
  if (prev) assert(prev->dir == 2);
  const int k1 = L - 1 - index;
  for (int k = 0; k < k1; k++) {
    if (k == 0) L_create(*o.L[0], T);
    else o.L[k]->update(prev->L[k - 1].get(), T, T);
  }
  for (int k = 0; k < k1; k++)
    for (int m = 0; m < k1 - k; m++) {
      if (m == 0) {
        if (k == 0) { F0_makenew(*o.F0[0][0], T); F1_makenew(*o.F1[0][0], T); S0_makenew(*o.S0[0][0], T); }
        else {
          TOp* Lp = prev->L[k - 1].get();
          F0_makenew(*o.F0[k][0], Lp, T); F1_makenew(*o.F1[k][0], Lp, T);
          S0_makenew(*o.S0[k][0], Lp, T); S1_makenew(*o.S1[k][0], Lp, T);
        }
      } else {
        o.F0[k][m]->update(prev->F0[k][m - 1].get(), T, T);
        o.F1[k][m]->update(prev->F1[k][m - 1].get(), T, T);
        o.S0[k][m]->update(prev->S0[k][m - 1].get(), T, T);
        if (k > 0) o.S1[k][m]->update(prev->S1[k][m - 1].get(), T, T);
      }
    }
  const int k2 = index + 1;
  for (int k = 0; k < k2; k++)
    for (int m = 0; m < k2 - k; m++) {
      const int s1 = index - m - k, s2 = index - m;
      const int ip = bk->irrep(s1) ^ bk->irrep(s2);
      if (index == L - 2) {
        o.A[k][m]->clear(); if (k > 0) o.B[k][m]->clear();
        o.C[k][m]->clear(); o.D[k][m]->clear();
      } else {
        o.A[k][m]->update(prev->A[k][m + 1].get(), T, T);
        if (k > 0) o.B[k][m]->update(prev->B[k][m + 1].get(), T, T);
        o.C[k][m]->update(prev->C[k][m + 1].get(), T, T);
        o.D[k][m]->update(prev->D[k][m + 1].get(), T, T);
      }
      for (int num = 0; num < L - index - 1; num++) {
        if (ip != o.S0[num][0]->n_irrep) continue;
        const int r0 = index + 1, rn = index + 1 + num;
        double a = P->g(s1, s2, r0, rn);
        if (k == 0 && num == 0) a *= 0.5;
        if (k > 0 && num > 0) a += P->g(s1, s2, rn, r0);
        o.A[k][m]->axpy(a, o.S0[num][0].get());
        if (num > 0 && k > 0) o.B[k][m]->axpy(P->g(s1, s2, r0, rn) - P->g(s1, s2, rn, r0), o.S1[num][0].get());
        o.C[k][m]->axpy(2 * P->g(s1, r0, s2, rn) - P->g(s1, r0, rn, s2), o.F0[num][0].get());
        o.D[k][m]->axpy(-P->g(s1, r0, rn, s2), o.F1[num][0].get());
        if (num > 0) {
          o.C[k][m]->axpy_transpose_CD(2 * P->g(s1, rn, s2, r0) - P->g(s1, rn, r0, s2), o.F0[num][0].get());
          o.D[k][m]->axpy_transpose_CD(-P->g(s1, rn, r0, s2), o.F1[num][0].get());
        }
      }
    }
  for (int k = 0; k < k2; k++) {
    TOp& q = *o.Q[k];
    const int site = o.Qsite[k];
    if (index == L - 2) { q.clear(); Q_add_simple(q, P, site, T); }
    else {
      q.update(prev->Q[k + 1].get(), T, T);
      Q_add_simple(q, P, site, T);
      Q_add_L(q, P, site, prev->L, T);
      Q_add_AB(q, prev->A[k + 1][0].get(), prev->B[k + 1][0].get(), T);
      Q_add_CD(q, prev->C[k + 1][0].get(), prev->D[k + 1][0].get(), T);
    }
  }
  if (index == L - 2) X_update(*o.X, P, T);
  else X_update(*o.X, P, T, prev->L, prev->X.get(), prev->Q[0].get(), prev->A[0][0].get(), prev->C[0][0].get(), prev->D[0][0].get());
}

}  // namespace orc
