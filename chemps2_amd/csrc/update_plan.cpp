// Work-list generator for the renormalised-operator updates (DMRGoperators.cpp:243-907,
// TensorL.cpp, TensorF0.cpp, TensorF1.cpp, TensorS0.cpp, TensorS1.cpp, TensorQ.cpp,
// TensorX.cpp, TensorOperator.cpp:163-455). Every dgemm_ of the reference becomes a
// term of the device work list; the operator sums the reference builds in scratch
// memory become combinations evaluated before the terms.
#include <algorithm>
#include <cassert>

#include "parallel.hpp"
#include "update_plan.hpp"

namespace chm {

static const double SQ2 = std::sqrt(2.0), SQH = std::sqrt(0.5);
static constexpr int64_t ZERO = -7;  // Ref.off of an identically zero combination

#define DIM(b, n, s, i) sec->dim((b), (n), (s), (i))

static Ref need(Ref r, const char* what) {
  if (!r.ok()) throw Error{ERR_MISSING, std::string("operator update: missing block (") + what + ")"};
  return r;
}

void UpdatePlanner::dgemm(char ta, char tb, int m, int n, int k, double alpha, Ref A, int lda, Ref B, int ldb, Ref out,
                          int ldo) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  flops += 2.0 * m * n * k;
  if (ldo != m) throw Error{ERR_STATE, "operator update: unexpected output leading dimension"};
  UTerm t{};
  t.out = need(out, "out"); t.m = m; t.n = n; t.k1 = k; t.k2 = 0; t.alpha = alpha;
  t.A = need(A, "A"); t.ta = (ta == 'T'); t.lda = lda;
  t.B = need(B, "B"); t.tb = (tb == 'T'); t.ldb = ldb;
  t.kind = 0;
  terms[level].push_back(t);
}

// out(dRu x dRd) += alpha * Tup^T op(mid) Tdn; an absent mid is the identity.
void UpdatePlanner::sand_r(Ref out, int dRu, int dRd, int dLu, int dLd, double alpha, Ref Tup, Ref mid, bool midT,
                           Ref Tdn) {
  if (!mid.ok()) {
    if (mid.off == ZERO) { flops += 2.0 * dRu * dLd * dLu + 2.0 * dRu * dRd * dLd; return; }
    dgemm('T', 'N', dRu, dRd, dLu, alpha, Tup, dLu, Tdn, dLd, out, dRu);
    return;
  }
  flops += 2.0 * dRu * dLd * dLu + 2.0 * dRu * dRd * dLd;
  UTerm t{};
  t.out = need(out, "out"); t.m = dRu; t.n = dRd; t.k1 = dLu; t.k2 = dLd; t.alpha = alpha;
  t.A = need(Tup, "Tup"); t.ta = 1; t.lda = dLu;
  t.B = mid; t.tb = midT; t.ldb = midT ? dLd : dLu;
  t.C = need(Tdn, "Tdn"); t.tc = 0; t.ldc = dLd;
  t.kind = 1;
  terms[level].push_back(t);
}

// out(dLu x dLd) += alpha * Tup op(mid) Tdn^T; an absent mid is the identity.
void UpdatePlanner::sand_l(Ref out, int dLu, int dLd, int dRu, int dRd, double alpha, Ref Tup, Ref mid, bool midT,
                           Ref Tdn) {
  if (!mid.ok()) {
    if (mid.off == ZERO) { flops += 2.0 * dLu * dRd * dRu + 2.0 * dLu * dLd * dRd; return; }
    dgemm('N', 'T', dLu, dLd, dRu, alpha, Tup, dLu, Tdn, dLd, out, dLu);
    return;
  }
  flops += 2.0 * dLu * dRd * dRu + 2.0 * dLu * dLd * dRd;
  UTerm t{};
  t.out = need(out, "out"); t.m = dLu; t.n = dLd; t.k1 = dRu; t.k2 = dRd; t.alpha = alpha;
  t.A = need(Tup, "Tup"); t.ta = 0; t.lda = dLu;
  t.B = mid; t.tb = midT; t.ldb = midT ? dRd : dRu;
  t.C = need(Tdn, "Tdn"); t.tc = 1; t.ldc = dLd;
  t.kind = 1;
  terms[level].push_back(t);
}

void UpdatePlanner::combo_add(Ref r, double c) {
  if (c == 0.0) return;
  cur_src.push_back({need(r, "combination source"), c});
}

Ref UpdatePlanner::combo_end() {
  if (cur_src.empty()) { Ref z; z.off = ZERO; return z; }
  Combo c;
  c.out_off = combo_size;
  c.size = cur_size;
  c.src = std::move(cur_src);
  cur_src.clear();
  combo_size += (cur_size + 7) / 8 * 8;
  combos.push_back(std::move(c));
  return Ref{BUF_COMBO, combos.back().out_off};
}

// self += a * o, block by block (TensorOperator::daxpy).
void UpdatePlanner::axpy(OpT& self, double a, OpT* o) {
  if (a == 0.0) return;
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], iu = self.Iup[ik], sd = self.Sdn[ik];
    const int nd = n + self.n_elec, id = iu ^ self.n_irrep;
    const int du = DIM(b, n, su, iu), dd = DIM(b, nd, sd, id);
    UTerm t{};
    t.out = self.blkref(ik); t.m = du; t.n = dd; t.alpha = a;
    t.A = need(o->blk(n, su, iu, nd, sd, id), "axpy source"); t.ta = 0; t.lda = du;
    t.kind = 2;
    terms[level].push_back(t);
  }
}

// self += a * o^T with the spin factor of TensorOperator.cpp:416-455.
void UpdatePlanner::axpy_transpose_CD(OpT& self, double a, OpT* o) {
  if (a == 0.0) return;
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], sd = self.Sdn[ik], iu = self.Iup[ik], id = iu ^ self.n_irrep;
    const int du = DIM(b, n, su, iu), dd = DIM(b, n, sd, id);
    double pref = a;
    if (su != sd) pref *= phase(su - sd) * std::sqrt(self.moving_right ? (su + 1.0) / (sd + 1.0) : (sd + 1.0) / (su + 1.0));
    UTerm t{};
    t.out = self.blkref(ik); t.m = du; t.n = dd; t.alpha = pref;
    t.A = need(o->blk(n, sd, id, n, su, iu), "transposed source"); t.ta = 1; t.lda = dd;
    t.kind = 2;
    terms[level].push_back(t);
  }
}

// Generic renormalisation (TensorOperator.cpp:163-405): six local cases per block.
void UpdatePlanner::update(OpT& self, OpT* prev, const TLay* Tup, const TLay* Tdn) {
  const int index = self.index, two_j = self.two_j;
  const int jw3 = self.jw_phase ? 3 : 1;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int nu = self.Nup[ik], nd = nu + self.n_elec, su = self.Sup[ik], sd = self.Sdn[ik], iu = self.Iup[ik];
    const int id = iu ^ self.n_irrep;
    const Ref out = self.blkref(ik);
    if (self.moving_right) {
      const int dRu = DIM(index, nu, su, iu), dRd = DIM(index, nd, sd, id);
      const int Is = irrep(index - 1);
      for (int g = 0; g < 6; g++) {
        int sLu = su, sLd = sd, nLu = nu, nLd = nd, iLu = iu, iLd = id;
        if (g == 1) { nLu -= 2; nLd -= 2; }
        if (g >= 2) {
          nLu -= 1; nLd -= 1; iLu ^= Is; iLd ^= Is;
          sLu += (g == 2 || g == 3) ? -1 : 1;
          sLd += (g == 2 || g == 4) ? -1 : 1;
        }
        if (std::abs(sLu - sLd) > two_j) continue;
        const int dLu = DIM(index - 1, nLu, sLu, iLu), dLd = DIM(index - 1, nLd, sLd, iLd);
        if (dLu <= 0 || dLd <= 0) continue;
        double alpha = 1.0;
        if (g >= 2) {
          if (two_j == 0) alpha = self.jw_phase ? -1.0 : 1.0;
          else if (self.prime_last)
            alpha = phase(su + sLd + two_j + jw3) * std::sqrt((sLd + 1.0) * (su + 1.0)) * wigner6j(sLu, sLd, two_j, sd, su, 1);
          else
            alpha = phase(sd + sLu + two_j + jw3) * std::sqrt((sLu + 1.0) * (sd + 1.0)) * wigner6j(sLd, sLu, two_j, su, sd, 1);
        }
        sand_r(out, dRu, dRd, dLu, dLd, alpha, Tup->blk(nLu, sLu, iLu, nu, su, iu), need(prev->blk(nLu, sLu, iLu, nLd, sLd, iLd), "prev"),
               false, Tdn->blk(nLd, sLd, iLd, nd, sd, id));
      }
    } else {
      const int dLu = DIM(index, nu, su, iu), dLd = DIM(index, nd, sd, id);
      const int Is = irrep(index);
      for (int g = 0; g < 6; g++) {
        int sRu = su, sRd = sd, nRu = nu, nRd = nd, iRu = iu, iRd = id;
        if (g == 1) { nRu += 2; nRd += 2; }
        if (g >= 2) {
          nRu += 1; nRd += 1; iRu ^= Is; iRd ^= Is;
          sRu += (g == 2 || g == 3) ? -1 : 1;
          sRd += (g == 2 || g == 4) ? -1 : 1;
        }
        if (std::abs(sRu - sRd) > two_j) continue;
        const int dRu = DIM(index + 1, nRu, sRu, iRu), dRd = DIM(index + 1, nRd, sRd, iRd);
        if (dRu <= 0 || dRd <= 0) continue;
        double alpha = 1.0;
        if (g >= 2) {
          if (two_j == 0) alpha = (self.jw_phase ? -1.0 : 1.0) * ((sRu + 1.0) / (su + 1.0));
          else if (self.prime_last)
            alpha = phase(sRu + sd + two_j + jw3) * (sRd + 1) * std::sqrt((sRu + 1.0) / (sd + 1.0)) * wigner6j(sRu, sRd, two_j, sd, su, 1);
          else
            alpha = phase(sRd + su + two_j + jw3) * (sRu + 1) * std::sqrt((sRd + 1.0) / (su + 1.0)) * wigner6j(sRd, sRu, two_j, su, sd, 1);
        }
        sand_l(out, dLu, dLd, dRu, dRd, alpha, Tup->blk(nu, su, iu, nRu, sRu, iRu), need(prev->blk(nRu, sRu, iRu, nRd, sRd, iRd), "prev"),
               false, Tdn->blk(nd, sd, id, nRd, sRd, iRd));
      }
    }
  }
}



// ---------------------------------------------------------------- TensorL --
void UpdatePlanner::L_create(OpT& self, const TLay* T) {  // TensorL.cpp:41-135 (previous == NULL branches)
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int nu = self.Nup[ik], su = self.Sup[ik], iu = self.Iup[ik], sd = self.Sdn[ik];
    const int nd = nu + 1, id = iu ^ self.n_irrep;
    const Ref out = self.blkref(ik);
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
void UpdatePlanner::F0_makenew(OpT& self, const TLay* T) {  // TensorF0.cpp:53-133
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
    const Ref out = self.blkref(ik);
    if (self.moving_right) {
      const int dR = DIM(b, n, s, i), Is = irrep(b - 1);
      for (int g = 0; g < 3; g++) {
        const int nl = g == 0 ? n - 2 : n - 1, sl = g == 0 ? s : (g == 1 ? s - 1 : s + 1), il = g == 0 ? i : (i ^ Is);
        const int dl = DIM(b - 1, nl, sl, il);
        if (dl <= 0) continue;
        Ref B = T->blk(nl, sl, il, n, s, i);
        dgemm('T', 'N', dR, dR, dl, g == 0 ? SQ2 : 0.5 * SQ2, B, dl, B, dl, 1.0, out, dR);
      }
    } else {
      const int dL = DIM(b, n, s, i), Is = irrep(b);
      for (int g = 0; g < 3; g++) {
        const int nr = g == 0 ? n + 2 : n + 1, sr = g == 0 ? s : (g == 1 ? s - 1 : s + 1), ir = g == 0 ? i : (i ^ Is);
        const int dr = DIM(b + 1, nr, sr, ir);
        if (dr <= 0) continue;
        double a = SQ2;
        if (g >= 1) a *= 0.5 * (sr + 1.0) / (s + 1.0);
        Ref B = T->blk(n, s, i, nr, sr, ir);
        dgemm('N', 'T', dL, dL, dr, a, B, dL, B, dL, 1.0, out, dL);
      }
    }
  }
}

void UpdatePlanner::F0_makenew(OpT& self, OpT* Lop, const TLay* T) {  // TensorF0.cpp:135-285
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, s, i), dD = DIM(b, n, s, idn);
    if (self.moving_right) {
      const int Is = irrep(b - 1);
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
      const int Is = irrep(b);
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
void UpdatePlanner::F1_makenew(OpT& self, const TLay* T) {  // TensorF1.cpp:54-123
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik];
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, su, i), dD = DIM(b, n, sd, i);
    if (self.moving_right) {
      const int il = i ^ irrep(b - 1);
      for (int g = 0; g < 2; g++) {
        const int sl = su + (g == 0 ? -1 : 1);
        if (sl < 0 || std::abs(sd - sl) >= 2) continue;
        const int dl = DIM(b - 1, n - 1, sl, il);
        if (dl <= 0) continue;
        const double a = phase(sl + sd + 3) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 1, 2, su, sd, sl);
        dgemm('T', 'N', dU, dD, dl, a, T->blk(n - 1, sl, il, n, su, i), dl, T->blk(n - 1, sl, il, n, sd, i), dl, 1.0, out, dU);
      }
    } else {
      const int ir = i ^ irrep(b);
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

void UpdatePlanner::F1_makenew(OpT& self, OpT* Lop, const TLay* T) {  // TensorF1.cpp:125-269
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, su, i), dD = DIM(b, n, sd, idn);
    if (self.moving_right) {
      const int Is = irrep(b - 1);
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
      const int Is = irrep(b);
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
void UpdatePlanner::S0_makenew(OpT& self, const TLay* T) {  // TensorS0.cpp:53-103
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
    const Ref out = self.blkref(ik);
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

void UpdatePlanner::S0_makenew(OpT& self, OpT* Lop, const TLay* T) {  // TensorS0.cpp:105-246
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, s, i), dD = DIM(b, n + 2, s, idn);
    if (self.moving_right) {
      const int Is = irrep(b - 1);
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
      const int Is = irrep(b);
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
void UpdatePlanner::S1_makenew(OpT& self, OpT* Lop, const TLay* T) {  // TensorS1.cpp:48-195
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 2, sd, idn);
    if (self.moving_right) {
      const int Is = irrep(b - 1);
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
      const int Is = irrep(b);
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
void UpdatePlanner::Q_add_simple(OpT& self, int site, const TLay* T) {  // TensorQ.cpp:45-100
  if (irrep(T->index) != self.n_irrep) return;
  const int b = self.index;
  const double mx = self.moving_right ? gmx(b - 1, b - 1, b - 1, site) : gmx(site, b, b, b);
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
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


void UpdatePlanner::Q_add_L(OpT& self, int site, std::vector<OpT*>& Lprev, const TLay* T) {  // TensorQ.cpp:102-390
  const int b = self.index;
  // orbital -> previous L tensor (moving right: sites 0..b-2; moving left: sites b+1..L-1)
  auto Lof = [&](int loca) -> OpT* { return self.moving_right ? Lprev[b - 2 - loca] : Lprev[loca - b - 1]; };
  const int lo = self.moving_right ? 0 : b + 1, hi = self.moving_right ? b - 1 : L;
  bool any = false;
  for (int loca = lo; loca < hi; loca++) if (Lof(loca)->n_irrep == self.n_irrep) any = true;
  if (!any) return;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    if (self.moving_right) {
      {  // case 1
        const int dlu = DIM(b - 1, n, su, i), dld = DIM(b - 1, n - 1, sd, idn);
        if (dlu > 0 && dld > 0) {
          combo_begin((int64_t)dlu * dld);  // stored dld x dlu
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            Ref BL = Lof(loca)->blk(n - 1, sd, idn, n, su, i);
            const double a = gmx(loca, site, b - 1, b - 1);
            combo_add(BL, a);
          }
          const double a = phase(sd + 1 - su) * std::sqrt((su + 1.0) / (sd + 1.0));
          sand_r(out, dU, dD, dlu, dld, a, T->blk(n, su, i, n, su, i), combo_end(), true, T->blk(n - 1, sd, idn, n + 1, sd, idn));
        }
      }
      {  // case 2
        const int dlu = DIM(b - 1, n - 2, su, i), dld = DIM(b - 1, n - 1, sd, idn);
        if (dlu > 0 && dld > 0) {
          combo_begin((int64_t)dlu * dld);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            Ref BL = Lof(loca)->blk(n - 2, su, i, n - 1, sd, idn);
            const double a = 2 * gmx(loca, b - 1, site, b - 1) - gmx(loca, b - 1, b - 1, site);
            combo_add(BL, a);
          }
          sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n - 2, su, i, n, su, i), combo_end(), false, T->blk(n - 1, sd, idn, n + 1, sd, idn));
        }
      }
      for (int slu = su - 1; slu <= su + 1; slu += 2)  // case 3
        for (int sld = sd - 1; sld <= sd + 1; sld += 2) {
          if (sld < 0 || slu < 0 || std::abs(sld - slu) >= 2) continue;
          const int ilu = i ^ irrep(b - 1), ild = idn ^ irrep(b - 1);
          const int dlu = DIM(b - 1, n - 1, slu, ilu), dld = DIM(b - 1, n, sld, ild);
          if (dlu <= 0 || dld <= 0) continue;
          const double f = phase(su + sld) * std::sqrt((sld + 1) * (su + 1.0)) * wigner6j(su, sd, 1, sld, slu, 1);
          combo_begin((int64_t)dlu * dld);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            Ref BL = Lof(loca)->blk(n - 1, slu, ilu, n, sld, ild);
            double a = f * gmx(loca, b - 1, site, b - 1);
            if (sld == su) a += gmx(loca, b - 1, b - 1, site);
            combo_add(BL, a);
          }
          sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n - 1, slu, ilu, n, su, i), combo_end(), false, T->blk(n, sld, ild, n + 1, sd, idn));
        }
    } else {
      {  // case 1
        const int dru = DIM(b + 1, n + 2, su, i), drd = DIM(b + 1, n + 1, sd, idn);
        if (dru > 0 && drd > 0) {
          combo_begin((int64_t)dru * drd);  // stored drd x dru
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            Ref BL = Lof(loca)->blk(n + 1, sd, idn, n + 2, su, i);
            const double a = gmx(site, loca, b, b);
            combo_add(BL, a);
          }
          const double a = phase(su + 1 - sd) * std::sqrt((su + 1.0) / (sd + 1.0));
          sand_l(out, dU, dD, dru, drd, a, T->blk(n, su, i, n + 2, su, i), combo_end(), true, T->blk(n + 1, sd, idn, n + 1, sd, idn));
        }
      }
      {  // case 2
        const int dru = DIM(b + 1, n + 2, su, i), drd = DIM(b + 1, n + 3, sd, idn);
        if (dru > 0 && drd > 0) {
          combo_begin((int64_t)dru * drd);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            Ref BL = Lof(loca)->blk(n + 2, su, i, n + 3, sd, idn);
            const double a = 2 * gmx(site, b, loca, b) - gmx(site, b, b, loca);
            combo_add(BL, a);
          }
          sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 2, su, i), combo_end(), false, T->blk(n + 1, sd, idn, n + 3, sd, idn));
        }
      }
      for (int sru = su - 1; sru <= su + 1; sru += 2)  // case 3
        for (int srd = sd - 1; srd <= sd + 1; srd += 2) {
          if (srd < 0 || sru < 0 || std::abs(srd - sru) >= 2) continue;
          const int iru = i ^ irrep(b), ird = idn ^ irrep(b);
          const int dru = DIM(b + 1, n + 1, sru, iru), drd = DIM(b + 1, n + 2, srd, ird);
          if (dru <= 0 || drd <= 0) continue;
          const double f1 = phase(sd + sru) * std::sqrt((sru + 1.0) / (sd + 1.0)) * (srd + 1) * wigner6j(su, sd, 1, srd, sru, 1);
          const double f2 = (srd + 1.0) / (sd + 1.0);
          combo_begin((int64_t)dru * drd);
          for (int loca = lo; loca < hi; loca++) {
            if (Lof(loca)->n_irrep != self.n_irrep) continue;
            Ref BL = Lof(loca)->blk(n + 1, sru, iru, n + 2, srd, ird);
            double a = f1 * gmx(site, b, loca, b);
            if (sru == sd) a += f2 * gmx(site, b, b, loca);
            combo_add(BL, a);
          }
          sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 1, sru, iru), combo_end(), false, T->blk(n + 1, sd, idn, n + 2, srd, ird));
        }
    }
  }
}


void UpdatePlanner::Q_add_AB(OpT& self, OpT* A, OpT* B, const TLay* T) {  // TensorQ.cpp:392-575
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    auto combine = [&](double fA, Ref BA, double fB, Ref BB, size_t sz, double* alpha) -> Ref {
      if (BA.ok()) {
        combo_begin((int64_t)sz);
        combo_add(BA, fA);
        combo_add(BB, fB);
        *alpha = 1.0;
        return combo_end();
      }
      *alpha = fB;
      return BB;
    };
    if (self.moving_right) {
      const int ilu = i ^ irrep(b - 1);
      for (int slu = su - 1; slu <= su + 1; slu += 2) {  // case 1
        const int dlu = DIM(b - 1, n - 1, slu, ilu), dld = DIM(b - 1, n + 1, sd, idn);
        if (dlu <= 0 || dld <= 0) continue;
        const double fB = phase(slu + sd + 2) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 2, 1, sd, su, slu);
        const double fA = phase(sd + 1 - su) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        double alpha;
        Ref m = combine(fA, slu == sd ? A->blk(n - 1, slu, ilu, n + 1, sd, idn) : Ref(), fB,
                                  B->blk(n - 1, slu, ilu, n + 1, sd, idn), (size_t)dlu * dld, &alpha);
        sand_r(out, dU, dD, dlu, dld, alpha, T->blk(n - 1, slu, ilu, n, su, i), m, false, T->blk(n + 1, sd, idn, n + 1, sd, idn));
      }
      const int ild = idn ^ irrep(b - 1);
      for (int sld = sd - 1; sld <= sd + 1; sld += 2) {  // case 2
        const int dlu = DIM(b - 1, n - 2, su, i), dld = DIM(b - 1, n, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double fB = phase(su + sd + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 2, 1, su, sd, sld);
        double alpha;
        Ref m = combine(-SQH, sld == su ? A->blk(n - 2, su, i, n, sld, ild) : Ref(), fB,
                                  B->blk(n - 2, su, i, n, sld, ild), (size_t)dlu * dld, &alpha);
        sand_r(out, dU, dD, dlu, dld, alpha, T->blk(n - 2, su, i, n, su, i), m, false, T->blk(n, sld, ild, n + 1, sd, idn));
      }
    } else {
      const int ird = idn ^ irrep(b);
      for (int srd = sd - 1; srd <= sd + 1; srd += 2) {  // case 1
        const int dru = DIM(b + 1, n, su, i), drd = DIM(b + 1, n + 2, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double fB = phase(srd + su + 2) * std::sqrt(3.0 / (sd + 1.0)) * (srd + 1) * wigner6j(1, 1, 2, su, srd, sd);
        const double fA = phase(su + 1 - sd) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        double alpha;
        Ref m = combine(fA, srd == su ? A->blk(n, su, i, n + 2, srd, ird) : Ref(), fB,
                                  B->blk(n, su, i, n + 2, srd, ird), (size_t)dru * drd, &alpha);
        sand_l(out, dU, dD, dru, drd, alpha, T->blk(n, su, i, n, su, i), m, false, T->blk(n + 1, sd, idn, n + 2, srd, ird));
      }
      const int iru = i ^ irrep(b);
      for (int sru = su - 1; sru <= su + 1; sru += 2) {  // case 2
        const int dru = DIM(b + 1, n + 1, sru, iru), drd = DIM(b + 1, n + 3, sd, idn);
        if (dru <= 0 || drd <= 0) continue;
        const double fB = phase(su + sd + 1) * std::sqrt(3.0 * (sru + 1)) * wigner6j(1, 1, 2, sru, sd, su);
        double alpha;
        Ref m = combine(-SQH, sru == sd ? A->blk(n + 1, sru, iru, n + 3, sd, idn) : Ref(), fB,
                                  B->blk(n + 1, sru, iru, n + 3, sd, idn), (size_t)dru * drd, &alpha);
        sand_l(out, dU, dD, dru, drd, alpha, T->blk(n, su, i, n + 1, sru, iru), m, false, T->blk(n + 1, sd, idn, n + 3, sd, idn));
      }
    }
  }
}


void UpdatePlanner::Q_add_CD(OpT& self, OpT* C, OpT* D, const TLay* T) {  // TensorQ.cpp:577-785
  const int b = self.index;
  for (int ik = 0; ik < self.nk(); ik++) {
    const int n = self.Nup[ik], su = self.Sup[ik], i = self.Iup[ik], sd = self.Sdn[ik], idn = self.n_irrep ^ i;
    const Ref out = self.blkref(ik);
    const int dU = DIM(b, n, su, i), dD = DIM(b, n + 1, sd, idn);
    Ref mem;
    auto build = [&](double fD, Ref BD, double fC, Ref BC, size_t sz) {
      combo_begin((int64_t)sz);
      combo_add(BD, fD);
      if (BC.ok()) combo_add(BC, fC);
      mem = combo_end();
    };
    if (self.moving_right) {
      const int ild = idn ^ irrep(b - 1);
      for (int sld = sd - 1; sld <= sd + 1; sld += 2) {  // case 1
        const int dlu = DIM(b - 1, n, su, i), dld = DIM(b - 1, n, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        const double fD = phase(su + sd + 1) * std::sqrt(3.0 * (sld + 1)) * wigner6j(1, 2, 1, su, sd, sld);
        build(fD, D->blk(n, su, i, n, sld, ild), SQH, sld == su ? C->blk(n, su, i, n, sld, ild) : Ref(), (size_t)dlu * dld);
        sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n, su, i, n, su, i), mem, false, T->blk(n, sld, ild, n + 1, sd, idn));
      }
      const int ilu = i ^ irrep(b - 1);
      for (int slu = su - 1; slu <= su + 1; slu += 2) {  // case 2
        const int dlu = DIM(b - 1, n - 1, slu, ilu), dld = DIM(b - 1, n - 1, sd, idn);
        if (dlu <= 0 || dld <= 0) continue;
        const double fD = phase(slu + sd) * std::sqrt(3.0 * (su + 1)) * wigner6j(1, 2, 1, sd, su, slu);
        const double fC = phase(sd + 1 - su) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        build(fD, D->blk(n - 1, slu, ilu, n - 1, sd, idn), fC, slu == sd ? C->blk(n - 1, slu, ilu, n - 1, sd, idn) : Ref(), (size_t)dlu * dld);
        sand_r(out, dU, dD, dlu, dld, 1.0, T->blk(n - 1, slu, ilu, n, su, i), mem, false, T->blk(n - 1, sd, idn, n + 1, sd, idn));
      }
    } else {
      const int iru = i ^ irrep(b);
      for (int sru = su - 1; sru <= su + 1; sru += 2) {  // case 1
        const int dru = DIM(b + 1, n + 1, sru, iru), drd = DIM(b + 1, n + 1, sd, idn);
        if (dru <= 0 || drd <= 0) continue;
        const double fD = phase(su + sru + 3) * std::sqrt(3.0 / (sd + 1.0)) * (sru + 1) * wigner6j(1, 1, 2, sru, sd, su);
        build(fD, D->blk(n + 1, sru, iru, n + 1, sd, idn), SQH, sru == sd ? C->blk(n + 1, sru, iru, n + 1, sd, idn) : Ref(), (size_t)dru * drd);
        sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 1, sru, iru), mem, false, T->blk(n + 1, sd, idn, n + 1, sd, idn));
      }
      const int ird = idn ^ irrep(b);
      for (int srd = sd - 1; srd <= sd + 1; srd += 2) {  // case 2
        const int dru = DIM(b + 1, n + 2, su, i), drd = DIM(b + 1, n + 2, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        const double fD = (((srd + 1) % 2) != 0 ? -1 : 1) * std::sqrt(3.0 * (srd + 1.0) * (su + 1.0) / (sd + 1.0)) * wigner6j(1, 1, 2, su, srd, sd);
        const double fC = phase(su + 1 - sd) * std::sqrt(0.5 * (su + 1.0) / (sd + 1.0));
        build(fD, D->blk(n + 2, su, i, n + 2, srd, ird), fC, srd == su ? C->blk(n + 2, su, i, n + 2, srd, ird) : Ref(), (size_t)dru * drd);
        sand_l(out, dU, dD, dru, drd, 1.0, T->blk(n, su, i, n + 2, su, i), mem, false, T->blk(n + 1, sd, idn, n + 2, srd, ird));
      }
    }
  }
}


// ---------------------------------------------------------------- TensorX --
void UpdatePlanner::X_makenew(OpT& self, const TLay* T, int ik) {  // TensorX.cpp:127-167
  const int b = self.index;
  const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
  const Ref out = self.blkref(ik);
  if (self.moving_right) {
    const int dR = DIM(b, n, s, i), dL = DIM(b - 1, n - 2, s, i);
    const double a = gmx(b - 1, b - 1, b - 1, b - 1);
    if (dL > 0 && std::fabs(a) > 0.0) {
      Ref B = T->blk(n - 2, s, i, n, s, i);
      dgemm('T', 'N', dR, dR, dL, a, B, dL, B, dL, out, dR);
    }
  } else {
    const int dL = DIM(b, n, s, i), dR = DIM(b + 1, n + 2, s, i);
    const double a = gmx(b, b, b, b);
    if (dR > 0 && std::fabs(a) > 0.0) {
      Ref B = T->blk(n, s, i, n + 2, s, i);
      dgemm('N', 'T', dL, dL, dR, a, B, dL, B, dL, out, dL);
    }
  }
}

void UpdatePlanner::X_update(OpT& self, const TLay* T) {  // TensorX.cpp:44-56
  for (int ik = 0; ik < self.nk(); ik++) X_makenew(self, T, ik);
}

// TensorX.cpp:58-125 and 169-645. Terms of the form acc + acc^T are emitted as the
// product and its transpose (swapped outer operands, transposed middle).
void UpdatePlanner::X_update(OpT& self, const TLay* T, std::vector<OpT*>& Lprev, OpT* Xprev, OpT* Qprev, OpT* A,
                             OpT* C, OpT* D) {
  const int b = self.index;
  const bool other = self.moving_right ? (b > 1) : (b < L - 1);
  if (other) update(self, Xprev, T, T);  // the generic update of the previous X
  for (int ik = 0; ik < self.nk(); ik++) {
    X_makenew(self, T, ik);
    if (!other) continue;
    const int n = self.Nup[ik], s = self.Sup[ik], i = self.Iup[ik];
    const Ref out = self.blkref(ik);
    if (self.moving_right) {
      const int dR = DIM(b, n, s, i), Is = irrep(b - 1);
      for (int g = 0; g < 4; g++) {  // Q and L term (TensorX.cpp:169-263): mem + mem^T
        int nlu, slu, ilu, nld, sld, ild;
        if (g <= 1) { nlu = n - 1; slu = g == 0 ? s - 1 : s + 1; ilu = i ^ Is; nld = n; sld = s; ild = i; }
        else { nlu = n - 2; slu = s; ilu = i; nld = n - 1; sld = g == 2 ? s - 1 : s + 1; ild = i ^ Is; }
        const int dlu = DIM(b - 1, nlu, slu, ilu), dld = DIM(b - 1, nld, sld, ild);
        if (dlu <= 0 || dld <= 0) continue;
        Ref BQ = Qprev->blk(nlu, slu, ilu, nld, sld, ild);
        double f = 1.0;
        Ref ptr = BQ;
        if (g >= 2) {
          f = phase(s + 1 - sld) * std::sqrt((sld + 1.0) / (s + 1.0));
          combo_begin((int64_t)dlu * dld);
          combo_add(BQ, 1.0);
          for (int loca = 0; loca < b - 1; loca++) {
            if (irrep(b - 1) != irrep(loca)) continue;
            combo_add(Lprev[b - 2 - loca]->blk(nlu, slu, ilu, nld, sld, ild), gmx(loca, b - 1, b - 1, b - 1));
          }
          ptr = combo_end();
        }
        Ref T1 = T->blk(nlu, slu, ilu, n, s, i), T2 = T->blk(nld, sld, ild, n, s, i);
        sand_r(out, dR, dR, dlu, dld, f, T1, ptr, false, T2);
        sand_r(out, dR, dR, dld, dlu, f, T2, ptr, true, T1);
      }
      {  // A term (TensorX.cpp:361-395)
        const int dlu = DIM(b - 1, n - 2, s, i), dld = DIM(b - 1, n, s, i);
        if (dlu > 0 && dld > 0) {
          Ref T1 = T->blk(n - 2, s, i, n, s, i), T2 = T->blk(n, s, i, n, s, i), M = A->blk(n - 2, s, i, n, s, i);
          sand_r(out, dR, dR, dlu, dld, SQ2, T1, M, false, T2);
          sand_r(out, dR, dR, dld, dlu, SQ2, T2, M, true, T1);
        }
      }
      for (int g = 0; g < 3; g++) {  // C term (TensorX.cpp:433-467)
        const int nl = g < 2 ? n - 1 : n - 2, sl = g == 0 ? s - 1 : (g == 1 ? s + 1 : s), il = g < 2 ? (i ^ Is) : i;
        const int dl = DIM(b - 1, nl, sl, il);
        if (dl <= 0) continue;
        Ref BT = T->blk(nl, sl, il, n, s, i);
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
      const int dL = DIM(b, n, s, i), Is = irrep(b);
      for (int g = 0; g < 4; g++) {  // Q and L term (TensorX.cpp:265-359)
        int nru, sru, iru, nrd, srd, ird;
        if (g <= 1) { nru = n; sru = s; iru = i; nrd = n + 1; srd = g == 0 ? s - 1 : s + 1; ird = i ^ Is; }
        else { nru = n + 1; sru = g == 2 ? s - 1 : s + 1; iru = i ^ Is; nrd = n + 2; srd = s; ird = i; }
        const int dru = DIM(b + 1, nru, sru, iru), drd = DIM(b + 1, nrd, srd, ird);
        if (dru <= 0 || drd <= 0) continue;
        Ref BQ = Qprev->blk(nru, sru, iru, nrd, srd, ird);
        double f;
        Ref ptr = BQ;
        if (g <= 1) f = (srd + 1.0) / (s + 1.0);
        else {
          f = phase(s + 1 - sru) * std::sqrt((sru + 1.0) / (s + 1.0));
          combo_begin((int64_t)dru * drd);
          combo_add(BQ, 1.0);
          for (int loca = b + 1; loca < L; loca++) {
            if (irrep(b) != irrep(loca)) continue;
            combo_add(Lprev[loca - b - 1]->blk(nru, sru, iru, nrd, srd, ird), gmx(b, b, b, loca));
          }
          ptr = combo_end();
        }
        Ref T1 = T->blk(n, s, i, nru, sru, iru), T2 = T->blk(n, s, i, nrd, srd, ird);
        sand_l(out, dL, dL, dru, drd, f, T1, ptr, false, T2);
        sand_l(out, dL, dL, drd, dru, f, T2, ptr, true, T1);
      }
      {  // A term (TensorX.cpp:397-431)
        const int dru = DIM(b + 1, n, s, i), drd = DIM(b + 1, n + 2, s, i);
        if (dru > 0 && drd > 0) {
          Ref T1 = T->blk(n, s, i, n, s, i), T2 = T->blk(n, s, i, n + 2, s, i), M = A->blk(n, s, i, n + 2, s, i);
          sand_l(out, dL, dL, dru, drd, SQ2, T1, M, false, T2);
          sand_l(out, dL, dL, drd, dru, SQ2, T2, M, true, T1);
        }
      }
      for (int g = 0; g < 3; g++) {  // C term (TensorX.cpp:469-537)
        const int nr = g < 2 ? n + 1 : n + 2, sr = g == 0 ? s - 1 : (g == 1 ? s + 1 : s), ir = g < 2 ? (i ^ Is) : i;
        const int dr = DIM(b + 1, nr, sr, ir);
        if (dr <= 0) continue;
        Ref BT = T->blk(n, s, i, nr, sr, ir);
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
void UpdatePlanner::run_tasks() {
  struct Res {
    std::vector<UTerm> t[3];
    std::vector<Combo> combos;
    int64_t csize = 0;
    double flops = 0.0;
  };
  const int64_t n = (int64_t)tasks.size();
  std::vector<Res> res(n);
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads > 0 ? nthreads : host_threads(), n));
  std::vector<std::unique_ptr<UpdatePlanner>> workers(nt);
  for (auto& w : workers) {
    w = std::make_unique<UpdatePlanner>();
    w->sec = sec; w->irr = irr; w->mx = mx; w->L = L;
  }
  parallel_for(n, nt, [&](int64_t i, int tid) {
    UpdatePlanner& w = *workers[tid];
    w.terms[1].clear(); w.terms[2].clear(); w.combos.clear();
    w.combo_size = 0;
    w.flops = 0.0;
    w.level = tasks[i].level;
    tasks[i].fn(w);
    Res& r = res[i];
    r.t[1] = std::move(w.terms[1]); r.t[2] = std::move(w.terms[2]);
    r.combos = std::move(w.combos);
    r.csize = w.combo_size;
    r.flops = w.flops;
    w.terms[1] = {}; w.terms[2] = {}; w.combos = {};
  });
  for (int64_t i = 0; i < n; i++) {
    Res& r = res[i];
    const int64_t base = combo_size;
    for (Combo& c : r.combos) {
      c.out_off += base;
      combos.push_back(std::move(c));
    }
    combo_size += r.csize;
    flops += r.flops;
    for (int lev = 1; lev <= 2; lev++)
      if (!r.t[lev].empty()) segs[lev].push_back({std::move(r.t[lev]), base});
  }
  tasks.clear();
}

void UpdatePlanner::flatten() {
  for (int lev = 1; lev <= 2; lev++) {
    terms[lev].clear();
    for (const USeg& s : segs[lev])
      for (UTerm t : s.terms) {
        for (Ref* r : {&t.A, &t.B, &t.C})
          if (r->buf == BUF_COMBO) r->off += s.combo_base;
        terms[lev].push_back(t);
      }
  }
}

void UpdatePlanner::moving_right(int index, OpSet& o, OpSet* prev, const TLay* T, const std::vector<int>& Qsite) {  // DMRGoperators.cpp:243-574
  OpSet* O = &o;
  auto add = [&](int lev, std::function<void(UpdatePlanner&)> f) { tasks.push_back({lev, std::move(f)}); };
  for (int k = 0; k <= index; k++) {
    if (k == 0) add(1, [=](UpdatePlanner& w) { w.L_create(*O->L[0], T); });
    else add(1, [=](UpdatePlanner& w) { w.update(*O->L[k], prev->L[k - 1], T, T); });
  }
  for (int k = 0; k <= index; k++)
    for (int m = 0; m <= index - k; m++) {
      if (m == 0) {
        if (k == 0) {
          add(1, [=](UpdatePlanner& w) { w.F0_makenew(*O->F0[0][0], T); });
          add(1, [=](UpdatePlanner& w) { w.F1_makenew(*O->F1[0][0], T); });
          add(1, [=](UpdatePlanner& w) { w.S0_makenew(*O->S0[0][0], T); });
        } else {
          OpT* Lp = prev->L[k - 1];
          add(1, [=](UpdatePlanner& w) { w.F0_makenew(*O->F0[k][0], Lp, T); });
          add(1, [=](UpdatePlanner& w) { w.F1_makenew(*O->F1[k][0], Lp, T); });
          add(1, [=](UpdatePlanner& w) { w.S0_makenew(*O->S0[k][0], Lp, T); });
          add(1, [=](UpdatePlanner& w) { w.S1_makenew(*O->S1[k][0], Lp, T); });
        }
      } else {
        add(1, [=](UpdatePlanner& w) { w.update(*O->F0[k][m], prev->F0[k][m - 1], T, T); });
        add(1, [=](UpdatePlanner& w) { w.update(*O->F1[k][m], prev->F1[k][m - 1], T, T); });
        add(1, [=](UpdatePlanner& w) { w.update(*O->S0[k][m], prev->S0[k][m - 1], T, T); });
        if (k > 0) add(1, [=](UpdatePlanner& w) { w.update(*O->S1[k][m], prev->S1[k][m - 1], T, T); });
      }
    }
  const int k2 = L - 1 - index;
  for (int k = 0; k < k2; k++)
    for (int m = 0; m < k2 - k; m++)
      add(2, [=](UpdatePlanner& w) {
        const int s1 = index + 1 + m, s2 = index + 1 + k + m;
        const int ip = w.irrep(s1) ^ w.irrep(s2);
        if (index != 0) {  // at index 0 the new tensors start at zero
          w.update(*O->A[k][m], prev->A[k][m + 1], T, T);
          if (k > 0) w.update(*O->B[k][m], prev->B[k][m + 1], T, T);
          w.update(*O->C[k][m], prev->C[k][m + 1], T, T);
          w.update(*O->D[k][m], prev->D[k][m + 1], T, T);
        }
        for (int num = 0; num <= index; num++) {
          if (ip != O->S0[num][0]->n_irrep) continue;
          double a = w.gmx(index - num, index, s1, s2);
          if (k == 0 && num == 0) a *= 0.5;
          if (k > 0 && num > 0) a += w.gmx(index - num, index, s2, s1);
          w.axpy(*O->A[k][m], a, O->S0[num][0]);
          if (num > 0 && k > 0) w.axpy(*O->B[k][m], w.gmx(index - num, index, s1, s2) - w.gmx(index - num, index, s2, s1), O->S1[num][0]);
          w.axpy(*O->C[k][m], 2 * w.gmx(index - num, s1, index, s2) - w.gmx(index - num, s1, s2, index), O->F0[num][0]);
          w.axpy(*O->D[k][m], -w.gmx(index - num, s1, s2, index), O->F1[num][0]);
          if (num > 0) {
            w.axpy_transpose_CD(*O->C[k][m], 2 * w.gmx(index - num, s2, index, s1) - w.gmx(index - num, s2, s1, index), O->F0[num][0]);
            w.axpy_transpose_CD(*O->D[k][m], -w.gmx(index - num, s2, s1, index), O->F1[num][0]);
          }
        }
      });
  for (int k = 0; k < k2; k++) {
    const int site = Qsite[k];
    add(1, [=](UpdatePlanner& w) {
      OpT& q = *O->Q[k];
      if (index == 0) { w.Q_add_simple(q, site, T); return; }
      w.update(q, prev->Q[k + 1], T, T);
      w.Q_add_simple(q, site, T);
      w.Q_add_L(q, site, prev->L, T);
      w.Q_add_AB(q, prev->A[k + 1][0], prev->B[k + 1][0], T);
      w.Q_add_CD(q, prev->C[k + 1][0], prev->D[k + 1][0], T);
    });
  }
  if (index == 0) add(1, [=](UpdatePlanner& w) { w.X_update(*O->X, T); });
  else add(1, [=](UpdatePlanner& w) { w.X_update(*O->X, T, prev->L, prev->X, prev->Q[0], prev->A[0][0], prev->C[0][0], prev->D[0][0]); });
  run_tasks();
}

void UpdatePlanner::moving_left(int index, OpSet& o, OpSet* prev, const TLay* T, const std::vector<int>& Qsite) {  // DMRGoperators.cpp:576-907
  OpSet* O = &o;
  auto add = [&](int lev, std::function<void(UpdatePlanner&)> f) { tasks.push_back({lev, std::move(f)}); };
  const int k1 = L - 1 - index;
  for (int k = 0; k < k1; k++) {
    if (k == 0) add(1, [=](UpdatePlanner& w) { w.L_create(*O->L[0], T); });
    else add(1, [=](UpdatePlanner& w) { w.update(*O->L[k], prev->L[k - 1], T, T); });
  }
  for (int k = 0; k < k1; k++)
    for (int m = 0; m < k1 - k; m++) {
      if (m == 0) {
        if (k == 0) {
          add(1, [=](UpdatePlanner& w) { w.F0_makenew(*O->F0[0][0], T); });
          add(1, [=](UpdatePlanner& w) { w.F1_makenew(*O->F1[0][0], T); });
          add(1, [=](UpdatePlanner& w) { w.S0_makenew(*O->S0[0][0], T); });
        } else {
          OpT* Lp = prev->L[k - 1];
          add(1, [=](UpdatePlanner& w) { w.F0_makenew(*O->F0[k][0], Lp, T); });
          add(1, [=](UpdatePlanner& w) { w.F1_makenew(*O->F1[k][0], Lp, T); });
          add(1, [=](UpdatePlanner& w) { w.S0_makenew(*O->S0[k][0], Lp, T); });
          add(1, [=](UpdatePlanner& w) { w.S1_makenew(*O->S1[k][0], Lp, T); });
        }
      } else {
        add(1, [=](UpdatePlanner& w) { w.update(*O->F0[k][m], prev->F0[k][m - 1], T, T); });
        add(1, [=](UpdatePlanner& w) { w.update(*O->F1[k][m], prev->F1[k][m - 1], T, T); });
        add(1, [=](UpdatePlanner& w) { w.update(*O->S0[k][m], prev->S0[k][m - 1], T, T); });
        if (k > 0) add(1, [=](UpdatePlanner& w) { w.update(*O->S1[k][m], prev->S1[k][m - 1], T, T); });
      }
    }
  const int k2 = index + 1;
  for (int k = 0; k < k2; k++)
    for (int m = 0; m < k2 - k; m++)
      add(2, [=](UpdatePlanner& w) {
        const int s1 = index - m - k, s2 = index - m;
        const int ip = w.irrep(s1) ^ w.irrep(s2);
        if (index != L - 2) {  // at index L-2 the new tensors start at zero
          w.update(*O->A[k][m], prev->A[k][m + 1], T, T);
          if (k > 0) w.update(*O->B[k][m], prev->B[k][m + 1], T, T);
          w.update(*O->C[k][m], prev->C[k][m + 1], T, T);
          w.update(*O->D[k][m], prev->D[k][m + 1], T, T);
        }
        for (int num = 0; num < L - index - 1; num++) {
          if (ip != O->S0[num][0]->n_irrep) continue;
          const int r0 = index + 1, rn = index + 1 + num;
          double a = w.gmx(s1, s2, r0, rn);
          if (k == 0 && num == 0) a *= 0.5;
          if (k > 0 && num > 0) a += w.gmx(s1, s2, rn, r0);
          w.axpy(*O->A[k][m], a, O->S0[num][0]);
          if (num > 0 && k > 0) w.axpy(*O->B[k][m], w.gmx(s1, s2, r0, rn) - w.gmx(s1, s2, rn, r0), O->S1[num][0]);
          w.axpy(*O->C[k][m], 2 * w.gmx(s1, r0, s2, rn) - w.gmx(s1, r0, rn, s2), O->F0[num][0]);
          w.axpy(*O->D[k][m], -w.gmx(s1, r0, rn, s2), O->F1[num][0]);
          if (num > 0) {
            w.axpy_transpose_CD(*O->C[k][m], 2 * w.gmx(s1, rn, s2, r0) - w.gmx(s1, rn, r0, s2), O->F0[num][0]);
            w.axpy_transpose_CD(*O->D[k][m], -w.gmx(s1, rn, r0, s2), O->F1[num][0]);
          }
        }
      });
  for (int k = 0; k < k2; k++) {
    const int site = Qsite[k];
    add(1, [=](UpdatePlanner& w) {
      OpT& q = *O->Q[k];
      if (index == L - 2) { w.Q_add_simple(q, site, T); return; }
      w.update(q, prev->Q[k + 1], T, T);
      w.Q_add_simple(q, site, T);
      w.Q_add_L(q, site, prev->L, T);
      w.Q_add_AB(q, prev->A[k + 1][0], prev->B[k + 1][0], T);
      w.Q_add_CD(q, prev->C[k + 1][0], prev->D[k + 1][0], T);
    });
  }
  if (index == L - 2) add(1, [=](UpdatePlanner& w) { w.X_update(*O->X, T); });
  else add(1, [=](UpdatePlanner& w) { w.X_update(*O->X, T, prev->L, prev->X, prev->Q[0], prev->A[0][0], prev->C[0][0], prev->D[0][0]); });
  run_tasks();
}

}  // namespace chm
