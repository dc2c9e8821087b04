// Two-particle reduced density matrix of the current MPS on the device
// (DMRG::calc_rdms_and_correlations, 2-RDM part, DMRGtechnics.cpp:40-130; TwoDM,
// TwoDM.cpp:40-1592).
//
// Every diagram value of TwoDM::FillSite is a sum over blocks of
//   f(block) < T_up, op(L_g) T_down op(R) >_F
// with T the site tensor, L_g an operator of the left boundary (or absent) and R an
// operator of the right boundary (or absent). Moving the contraction onto one boundary,
//   < T_up, L_g^T T_down R > = < T_down^T L_g T_up , R >,
// turns each family of diagrams into one "site-contracted" tensor P laid out exactly
// like the operators it meets (same boundary, spin, charge and irrep: same blocks at the
// same offsets), built by the operator-update executor (the products T^T L T of
// TensorOperator::update), after which every diagram value is one dot product of two
// equal-length device arrays: P with each operator of the other boundary. Families whose
// values only ever appear summed in the 2-RDM (d4 + d5, d13 + d14, d15 + d16, d17 + d18,
// d19 + d20, d21 + d22, d23 + d24) share one P.
#include <array>
#include <chrono>
#include <cmath>

#include "context.hpp"

namespace chm {

namespace {

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct PTen {
  const OpT* lay = nullptr;  // block layout (built on the current sectors)
  int64_t base = 0;          // offset in the P buffer
  Ref blk(int32_t buf, int n1, int s1, int i1, int n2, int s2, int i2) const {
    Ref r = lay->blk(n1, s1, i1, n2, s2, i2);
    if (!r.ok()) return r;
    r.buf = buf;
    r.off += base;
    return r;
  }
};

struct DotRec {
  Ref a, b;
  int64_t n;
  int d;        // diagram value it adds to
  double coef;  // weight of the dot product in that value
};

}  // namespace

// TwoDM::FillSite (TwoDM.cpp:445-628) at site i; A and B are the host 2-RDMs (L^4).
void Ctx::rdm_fill_site(int i, std::vector<double>& A, std::vector<double>& B) {
  const TLay& T = mps[i];
  const int b = i, c = i + 1, Ii = irreps[i];
  const OpSet* left = nullptr;
  const OpSet* right = nullptr;
  if (i > 0) {
    left = &sets[i - 1];
    if (left->dir != 1 || left->L.size() != (size_t)i) throw Error{ERR_STATE, "2-RDM: left operators are not built moving right"};
  }
  if (i < L - 1) {
    right = &sets[i];
    if (right->dir != 2 || right->L.size() != (size_t)(L - 1 - i))
      throw Error{ERR_STATE, "2-RDM: right operators are not built moving left"};
  }
  auto dL = [&](int n, int s, int r) { return sec.dim(b, n, s, r); };
  auto dR = [&](int n, int s, int r) { return sec.dim(c, n, s, r); };
  auto Tb = [&](int nl, int sl, int il, int nr, int sr, int ir) {
    Ref r = T.blk(nl, sl, il, nr, sr, ir);
    if (!r.ok()) throw Error{ERR_STATE, "2-RDM: missing MPS block"};
    return r;
  };
  auto set4 = [&](std::vector<double>& R, int p, int q, int r, int s, double v) {  // TwoDM.cpp:76-96
    auto at = [&](int a0, int a1, int a2, int a3) -> double& {
      return R[a0 + (size_t)L * (a1 + (size_t)L * (a2 + (size_t)L * a3))];
    };
    at(p, q, r, s) = v;
    at(q, p, s, r) = v;
    at(r, s, p, q) = v;
    at(s, r, q, p) = v;
  };

  // layouts of the P tensors, one per (boundary, 2j, n, irrep)
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<OpT>> lays;
  auto layout = [&](int bb, int tj, int ne, int nirr) -> const OpT* {
    auto& p = lays[std::make_tuple(bb, tj, ne, nirr)];
    if (!p) {
      p.reset(new OpT());
      p->build(sec, bb, tj, ne, nirr);
      p->buf = 0;  // any id: PTen::blk replaces it
    }
    return p.get();
  };
  std::vector<PTen> pts;
  int64_t ptot = 0;
  auto new_p = [&](int bb, int tj, int ne, int nirr) {
    PTen p;
    p.lay = layout(bb, tj, ne, nirr);
    p.base = ptot;
    ptot += p.lay->size;
    pts.push_back(p);
    return (int)pts.size() - 1;
  };
  auto same_layout = [&](const PTen& p, const OpT* op) {
    if (!op || op->size != p.lay->size || op->buf < 0) throw Error{ERR_STATE, "2-RDM: operator layout differs from its partner"};
  };

  // ---- which P tensors the site needs
  int p2 = -1, p3 = -1, p45 = -1, p6 = -1, p7 = -1;
  if (right) {
    p2 = new_p(c, 1, 1, Ii);
    p3 = new_p(c, 0, 2, 0);
    p45 = new_p(c, 0, 0, 0);
    p6 = new_p(c, 2, 0, 0);
  }
  if (left) p7 = new_p(b, 1, 1, Ii);
  struct GP { int p8 = -1, p9, p10, p11, p12, s0 = -1, s1 = -1, f0a = -1, f1a = -1, f0b = -1, f1b = -1; };
  std::vector<GP> gp(left && right ? i : 0);
  if (left && right)
    for (int g = 0; g < i; g++) {
      const int Ig = irreps[g];
      bool need_l = false, need_pair = false, need_s1 = false;
      for (int j = i + 1; j < L; j++) {
        if (irreps[j] == Ig) need_l = true;
        for (int k = j; k < L; k++)
          if ((irreps[j] ^ irreps[k]) == (Ig ^ Ii)) {
            need_pair = true;
            if (k > j) need_s1 = true;
          }
      }
      GP& q = gp[g];
      if (need_l) {
        q.p8 = new_p(c, 1, 1, Ig); q.p9 = new_p(c, 1, 1, Ig); q.p10 = new_p(c, 1, 1, Ig);
        q.p11 = new_p(c, 1, 1, Ig); q.p12 = new_p(c, 1, 1, Ig);
      }
      if (need_pair) {
        q.s0 = new_p(c, 0, 2, Ig ^ Ii);
        if (need_s1) q.s1 = new_p(c, 2, 2, Ig ^ Ii);
        q.f0a = new_p(c, 0, 0, Ig ^ Ii); q.f1a = new_p(c, 2, 0, Ig ^ Ii);
        q.f0b = new_p(c, 0, 0, Ig ^ Ii); q.f1b = new_p(c, 2, 0, Ig ^ Ii);
      }
    }
  own_buf(rdm_pbuf, std::max<int64_t>(ptot, 1));
  const int32_t pb = rdm_pbuf;

  // ---- terms of the P tensors
  std::vector<UTerm> terms;
  auto P = [&](int p, int n1, int s1, int i1, int n2, int s2, int i2) {
    Ref r = pts[p].blk(pb, n1, s1, i1, n2, s2, i2);
    if (!r.ok()) throw Error{ERR_STATE, "2-RDM: block outside its P layout"};
    return r;
  };
  // out (m x n) += a op(A) op(B), op(A) m x k
  auto emit2 = [&](Ref out, int m, int n, int k, double a, Ref Am, int ta, int lda, Ref Bm, int tb, int ldb) {
    if (a == 0.0) return;
    UTerm t{};
    t.out = out; t.m = m; t.n = n; t.k1 = k; t.alpha = a; t.kind = 0;
    t.A = Am; t.ta = ta; t.lda = lda; t.B = Bm; t.tb = tb; t.ldb = ldb;
    terms.push_back(t);
  };
  // out (m x n) += a op(A) op(B) op(C), op(A) m x k1, op(B) k1 x k2
  auto emit3 = [&](Ref out, int m, int n, double a, Ref Am, int ta, int lda, int k1, Ref Bm, int tb, int ldb, int k2,
                   Ref Cm, int tc, int ldc) {
    if (a == 0.0) return;
    UTerm t{};
    t.out = out; t.m = m; t.n = n; t.k1 = k1; t.k2 = k2; t.alpha = a; t.kind = 1;
    t.A = Am; t.ta = ta; t.lda = lda; t.B = Bm; t.tb = tb; t.ldb = ldb; t.C = Cm; t.tc = tc; t.ldc = ldc;
    terms.push_back(t);
  };
  const auto& lsecs = sec.list[b];
  std::vector<DotRec> dots;
  std::vector<double> dval;
  auto new_d = [&]() { dval.push_back(0.0); return (int)dval.size() - 1; };

  // D1 (TwoDM.cpp:642-668): dot products of the doubly occupied blocks with themselves
  const int d1 = new_d();
  for (const auto& ls : lsecs) {
    const int n = ls.N, s = ls.S, r = ls.I, dl = ls.d;
    const int dr = dR(n + 2, s, r);
    if (dr <= 0) continue;
    const Ref t = Tb(n, s, r, n + 2, s, r);
    dots.push_back({t, t, (int64_t)dl * dr, d1, (double)(s + 1)});
  }
  for (const auto& ls : lsecs) {
    const int n = ls.N, s = ls.S, r = ls.I, dl = ls.d;
    if (right) {
      const int irr = r ^ Ii;
      for (int sr = s - 1; sr <= s + 1; sr += 2) {  // D2: P2 = sum f T_up^T T_down
        const int dru = dR(n + 1, sr, irr), drd = dR(n + 2, s, r);
        if (sr < 0 || dru <= 0 || drd <= 0) continue;
        emit2(P(p2, n + 1, sr, irr, n + 2, s, r), dru, drd, dl, phase(s + 1 - sr) * 0.5 * std::sqrt((s + 1) * (sr + 1.0)),
              Tb(n, s, r, n + 1, sr, irr), 1, dl, Tb(n, s, r, n + 2, s, r), 0, dl);
      }
      {  // D3
        const int dd = dR(n, s, r), du = dR(n + 2, s, r);
        if (dd > 0 && du > 0)
          emit2(P(p3, n, s, r, n + 2, s, r), dd, du, dl, std::sqrt(0.5) * (s + 1), Tb(n, s, r, n, s, r), 1, dl,
                Tb(n, s, r, n + 2, s, r), 0, dl);
      }
      {  // D4
        const int dr = dR(n + 2, s, r);
        if (dr > 0) {
          const Ref t = Tb(n, s, r, n + 2, s, r);
          emit2(P(p45, n + 2, s, r, n + 2, s, r), dr, dr, dl, std::sqrt(0.5) * (s + 1), t, 1, dl, t, 0, dl);
        }
      }
      for (int sr = s - 1; sr <= s + 1; sr += 2) {  // D5
        const int dr = dR(n + 1, sr, irr);
        if (sr < 0 || dr <= 0) continue;
        const Ref t = Tb(n, s, r, n + 1, sr, irr);
        emit2(P(p45, n + 1, sr, irr, n + 1, sr, irr), dr, dr, dl, 0.5 * std::sqrt(0.5) * (sr + 1), t, 1, dl, t, 0, dl);
      }
      for (int su = s - 1; su <= s + 1; su += 2)  // D6
        for (int sd = s - 1; sd <= s + 1; sd += 2) {
          if (su < 0 || sd < 0) continue;
          const int du = dR(n + 1, su, irr), dd = dR(n + 1, sd, irr);
          if (du <= 0 || dd <= 0) continue;
          const double f = std::sqrt((su + 1) / 3.0) * (sd + 1) * phase(s + sd - 1) * wigner6j(1, 1, 2, su, sd, s);
          emit2(P(p6, n + 1, sd, irr, n + 1, su, irr), dd, du, dl, f, Tb(n, s, r, n + 1, sd, irr), 1, dl,
                Tb(n, s, r, n + 1, su, irr), 0, dl);
        }
    }
    if (left) {  // D7: P7 (left boundary) = sum f T_down T_up^T
      const int nu = n, su = s, iu = r, dlu = dl, ir = iu ^ Ii;
      for (int sd = su - 1; sd <= su + 1; sd += 2) {
        if (sd < 0) continue;
        const int dld = dL(nu - 1, sd, ir), dr = dR(nu + 1, sd, ir);
        if (dld <= 0 || dr <= 0) continue;
        const double f = 0.5 * std::sqrt((sd + 1) * (su + 1.0)) * phase(su - sd + 3);
        emit2(P(p7, nu - 1, sd, ir, nu, su, iu), dld, dlu, dr, f, Tb(nu - 1, sd, ir, nu + 1, sd, ir), 0, dld,
              Tb(nu, su, iu, nu + 1, sd, ir), 1, dlu);
      }
    }
  }

  // D8 - D24 (TwoDM.cpp:927-1592): P_g = sum f T_down^T L_g T_up (or its transpose)
  for (int g = 0; g < (int)gp.size(); g++) {
    const GP& q = gp[g];
    if (q.p8 < 0 && q.s0 < 0) continue;
    const int Ig = irreps[g];
    const OpT* Lg = left->L[i - 1 - g];
    if (!Lg || Lg->buf < 0) throw Error{ERR_MISSING, "2-RDM: left L operator missing"};
    for (const auto& ls : lsecs) {
      const int n = ls.N, su = ls.S, iu = ls.I, dlu = ls.d;
      const int idn = iu ^ Ig, ird = idn ^ Ii, iru = iu ^ Ii;
      if (q.p8 >= 0) {
        for (int sd = su - 1; sd <= su + 1; sd += 2) {
          if (sd < 0) continue;
          const int dld = dL(n - 1, sd, idn);
          if (dld <= 0) continue;
          const Ref Lb = Lg->blk(n - 1, sd, idn, n, su, iu);
          if (!Lb.ok()) continue;
          const int drd = dR(n + 1, sd, idn);
          if (drd <= 0) continue;
          const Ref Td = Tb(n - 1, sd, idn, n + 1, sd, idn);
          const int dru8 = dR(n + 2, su, iu);  // D8
          if (dru8 > 0)
            emit3(P(q.p8, n + 1, sd, idn, n + 2, su, iu), drd, dru8, -0.5 * (su + 1), Td, 1, dld, dld, Lb, 0, dld, dlu,
                  Tb(n, su, iu, n + 2, su, iu), 0, dlu);
          const int dru12 = dR(n, su, iu);  // D12
          if (dru12 > 0)
            emit3(P(q.p12, n, su, iu, n + 1, sd, idn), dru12, drd, phase(sd + 1 - su) * 0.5 * std::sqrt((su + 1) * (sd + 1.0)),
                  Tb(n, su, iu, n, su, iu), 1, dlu, dlu, Lb, 1, dld, dld, Td, 0, dld);
        }
        for (int sld = su - 1; sld <= su + 1; sld += 2)  // D9, D10, D11
          for (int sru = su - 1; sru <= su + 1; sru += 2)
            for (int srd = sru - 1; srd <= sru + 1; srd += 2) {
              if (sld < 0 || sru < 0 || srd < 0 || std::abs(sld - srd) > 1) continue;
              const int dld = dL(n - 1, sld, idn), dru = dR(n + 1, sru, iru), drd = dR(n, srd, ird);
              if (dld <= 0 || dru <= 0 || drd <= 0) continue;
              const Ref Lb = Lg->blk(n - 1, sld, idn, n, su, iu);
              if (!Lb.ok()) continue;
              const Ref Td = Tb(n - 1, sld, idn, n, srd, ird), Tu = Tb(n, su, iu, n + 1, sru, iru);
              const double f1 = phase(su + srd + 2) * (sru + 1) * std::sqrt((srd + 1) * (su + 1.0)) *
                                wigner6j(sru, 1, su, sld, 1, srd);
              const double f2 = 2 * (sru + 1) * std::sqrt((srd + 1) * (su + 1.0)) * wigner6j(sru, sld, 2, 1, 1, su) *
                                wigner6j(sru, sld, 2, 1, 1, srd);
              const double f3 = (srd == su) ? (sru + 1) : 0;
              const int pq[3] = {q.p9, q.p10, q.p11};
              const double fs[3] = {f1, f2, f3};
              for (int u = 0; u < 3; u++)
                emit3(P(pq[u], n, srd, ird, n + 1, sru, iru), drd, dru, fs[u], Td, 1, dld, dld, Lb, 0, dld, dlu, Tu, 0, dlu);
            }
      }
      if (q.s0 < 0) continue;
      for (int sd = su - 1; sd <= su + 1; sd += 2) {
        if (sd < 0) continue;
        const int dld = dL(n - 1, sd, idn);
        if (dld <= 0) continue;
        const Ref Lb = Lg->blk(n - 1, sd, idn, n, su, iu);
        if (!Lb.ok()) continue;
        {  // D13 (S0), D15 (S1)
          const int dru = dR(n + 2, su, iu);
          if (dru > 0) {
            const Ref Tu = Tb(n, su, iu, n + 2, su, iu);
            const int drd = dR(n, su, ird);
            if (drd > 0)
              emit3(P(q.s0, n, su, ird, n + 2, su, iu), drd, dru, -0.5 * std::sqrt(0.5) * (su + 1),
                    Tb(n - 1, sd, idn, n, su, ird), 1, dld, dld, Lb, 0, dld, dlu, Tu, 0, dlu);
            if (q.s1 >= 0)
              for (int srd = sd - 1; srd <= sd + 1; srd += 2) {
                const int d2 = srd >= 0 ? dR(n, srd, ird) : 0;
                if (d2 <= 0) continue;
                const double f = phase(sd + su + 1) * (su + 1) * std::sqrt((srd + 1) / 3.0) * wigner6j(1, 1, 2, su, srd, sd);
                emit3(P(q.s1, n, srd, ird, n + 2, su, iu), d2, dru, f, Tb(n - 1, sd, idn, n, srd, ird), 1, dld, dld, Lb, 0,
                      dld, dlu, Tu, 0, dlu);
              }
          }
        }
        {  // D14 (S0), D16 (S1)
          const int drd = dR(n - 1, sd, idn);
          if (drd > 0) {
            const Ref Td = Tb(n - 1, sd, idn, n - 1, sd, idn);
            const int dru = dR(n + 1, sd, iru);
            if (dru > 0)
              emit3(P(q.s0, n - 1, sd, idn, n + 1, sd, iru), drd, dru, phase(sd + 1 - su) * 0.5 * std::sqrt(0.5 * (su + 1) * (sd + 1)),
                    Td, 1, dld, dld, Lb, 0, dld, dlu, Tb(n, su, iu, n + 1, sd, iru), 0, dlu);
            if (q.s1 >= 0)
              for (int sru = su - 1; sru <= su + 1; sru += 2) {
                const int d2 = sru >= 0 ? dR(n + 1, sru, iru) : 0;
                if (d2 <= 0) continue;
                const double f = phase(sru + sd + 2) * (sru + 1) * std::sqrt((su + 1) / 3.0) * wigner6j(1, 1, 2, sru, sd, su);
                emit3(P(q.s1, n - 1, sd, idn, n + 1, sru, iru), drd, d2, f, Td, 1, dld, dld, Lb, 0, dld, dlu,
                      Tb(n, su, iu, n + 1, sru, iru), 0, dlu);
              }
          }
        }
        {  // D17/D21 (F0), D19/D23 (F1): up block empty on site i
          const int dru = dR(n, su, iu);
          if (dru > 0) {
            const Ref Tu = Tb(n, su, iu, n, su, iu);
            const int drd = dR(n, su, ird);
            if (drd > 0) {
              const Ref Td = Tb(n - 1, sd, idn, n, su, ird);
              const double f = std::sqrt(0.5) * 0.5 * (su + 1);
              emit3(P(q.f0a, n, su, ird, n, su, iu), drd, dru, f, Td, 1, dld, dld, Lb, 0, dld, dlu, Tu, 0, dlu);
              emit3(P(q.f0b, n, su, iu, n, su, ird), dru, drd, f, Tu, 1, dlu, dlu, Lb, 1, dld, dld, Td, 0, dld);
            }
            for (int srd = sd - 1; srd <= sd + 1; srd += 2) {
              const int d2 = srd >= 0 ? dR(n, srd, ird) : 0;
              if (d2 <= 0) continue;
              const Ref Td = Tb(n - 1, sd, idn, n, srd, ird);
              const double w = wigner6j(1, 1, 2, su, srd, sd);
              const double f19 = phase(sd + srd - 1) * (srd + 1) * std::sqrt((su + 1) / 3.0) * w;
              const double f23 = phase(sd + su - 1) * (su + 1) * std::sqrt((srd + 1) / 3.0) * w;
              emit3(P(q.f1a, n, srd, ird, n, su, iu), d2, dru, f19, Td, 1, dld, dld, Lb, 0, dld, dlu, Tu, 0, dlu);
              emit3(P(q.f1b, n, su, iu, n, srd, ird), dru, d2, f23, Tu, 1, dlu, dlu, Lb, 1, dld, dld, Td, 0, dld);
            }
          }
        }
        {  // D18/D22 (F0), D20/D24 (F1): down block doubly occupied on site i
          const int drd = dR(n + 1, sd, idn);
          if (drd > 0) {
            const Ref Td = Tb(n - 1, sd, idn, n + 1, sd, idn);
            const int dru = dR(n + 1, sd, iru);
            if (dru > 0) {
              const Ref Tu = Tb(n, su, iu, n + 1, sd, iru);
              const double f = phase(sd + 1 - su) * 0.5 * std::sqrt(0.5 * (su + 1) * (sd + 1));
              emit3(P(q.f0a, n + 1, sd, idn, n + 1, sd, iru), drd, dru, f, Td, 1, dld, dld, Lb, 0, dld, dlu, Tu, 0, dlu);
              emit3(P(q.f0b, n + 1, sd, iru, n + 1, sd, idn), dru, drd, f, Tu, 1, dlu, dlu, Lb, 1, dld, dld, Td, 0, dld);
            }
            for (int sru = su - 1; sru <= su + 1; sru += 2) {
              const int d2 = sru >= 0 ? dR(n + 1, sru, iru) : 0;
              if (d2 <= 0) continue;
              const Ref Tu = Tb(n, su, iu, n + 1, sru, iru);
              const double w = wigner6j(1, 1, 2, sru, sd, su);
              const double f20 = phase(2 * su) * std::sqrt((su + 1) * (sru + 1) * (sd + 1) / 3.0) * w;
              const double f24 = phase(2 * su + sru - sd) * (sru + 1) * std::sqrt((su + 1) / 3.0) * w;
              emit3(P(q.f1a, n + 1, sd, idn, n + 1, sru, iru), drd, d2, f20, Td, 1, dld, dld, Lb, 0, dld, dlu, Tu, 0, dlu);
              emit3(P(q.f1b, n + 1, sru, iru, n + 1, sd, idn), d2, drd, f24, Tu, 1, dlu, dlu, Lb, 1, dld, dld, Td, 0, dld);
            }
          }
        }
      }
    }
  }

  // ---- build every P (blocks without terms are zero)
  if (!pts.empty()) {
    std::vector<std::pair<Ref, std::pair<int, int>>> outs;
    for (const PTen& p : pts) {
      const OpT& lay = *p.lay;
      for (int k = 0; k < lay.nk(); k++) {
        const int du = sec.dim(lay.boundary, lay.Nup[k], lay.Sup[k], lay.Iup[k]);
        const int dd = sec.dim(lay.boundary, lay.Nup[k] + lay.n_elec, lay.Sdn[k], lay.Iup[k] ^ lay.n_irrep);
        outs.push_back({Ref{pb, p.base + lay.k2i[k]}, {du, dd}});
      }
    }
    run_terms(terms, outs, nullptr);
  }

  // ---- dot products with the operators of the other boundary
  auto dotp = [&](int p, const OpT* op, int d, double coef) {
    if (!op) throw Error{ERR_MISSING, "2-RDM: operator missing"};
    if (pts[p].lay->size == 0 && op->size == 0) return;  // no blocks at this boundary: the value is zero
    same_layout(pts[p], op);
    dots.push_back({Ref{pb, pts[p].base}, Ref{op->buf, 0}, pts[p].lay->size, d, coef});
  };
  std::map<int, int> d2v, d7v;
  std::map<std::pair<int, int>, std::array<int, 3>> d36v;
  std::map<std::pair<int, int>, std::array<int, 5>> d812v;
  std::map<std::tuple<int, int, int>, std::array<int, 6>> d1324v;
  if (right) {
    for (int j = i + 1; j < L; j++) {
      if (irreps[j] != Ii) continue;
      const int d = new_d();
      dotp(p2, right->L[j - i - 1], d, 1.0);
      d2v[j] = d;
    }
    for (int j = i + 1; j < L; j++)
      for (int k = j; k < L; k++) {
        if (irreps[j] != irreps[k]) continue;
        std::array<int, 3> v{new_d(), new_d(), new_d()};
        dotp(p3, right->S0[k - j][j - i - 1], v[0], 1.0);
        dotp(p45, right->F0[k - j][j - i - 1], v[1], 1.0);
        dotp(p6, right->F1[k - j][j - i - 1], v[2], 1.0);
        d36v[{j, k}] = v;
      }
  }
  if (left)
    for (int g = 0; g < i; g++) {
      if (irreps[g] != Ii) continue;
      const int d = new_d();
      dotp(p7, left->L[i - 1 - g], d, 1.0);
      d7v[g] = d;
    }
  for (int g = 0; g < (int)gp.size(); g++) {
    const GP& q = gp[g];
    const int Ig = irreps[g];
    for (int j = i + 1; j < L; j++) {
      if (q.p8 >= 0 && irreps[j] == Ig) {
        std::array<int, 5> v;
        const int pq[5] = {q.p8, q.p9, q.p10, q.p11, q.p12};
        for (int u = 0; u < 5; u++) {
          v[u] = new_d();
          dotp(pq[u], right->L[j - i - 1], v[u], 1.0);
        }
        d812v[{g, j}] = v;
      }
      if (q.s0 < 0) continue;
      for (int k = j; k < L; k++) {
        if ((irreps[j] ^ irreps[k]) != (Ig ^ Ii)) continue;
        std::array<int, 6> v;
        for (int u = 0; u < 6; u++) v[u] = new_d();
        dotp(q.s0, right->S0[k - j][j - i - 1], v[0], 1.0);
        if (k > j) dotp(q.s1, right->S1[k - j][j - i - 1], v[1], 1.0);
        dotp(q.f0a, right->F0[k - j][j - i - 1], v[2], 1.0);
        dotp(q.f1a, right->F1[k - j][j - i - 1], v[3], 1.0);
        dotp(q.f0b, right->F0[k - j][j - i - 1], v[4], 1.0);
        dotp(q.f1b, right->F1[k - j][j - i - 1], v[5], 1.0);
        d1324v[{g, j, k}] = v;
      }
    }
  }
  if (!dots.empty()) {
    std::vector<DevDot> h(dots.size());
    for (size_t e = 0; e < dots.size(); e++) h[e] = DevDot{resolve(dots[e].a), resolve(dots[e].b), dots[e].n};
    rdm_items.upload(h, stream);
    rdm_res.alloc(dots.size());
    launch_dots(rdm_items.p, (int)dots.size(), rdm_res.p, stream);
    hip_check(hipGetLastError(), "2-RDM dot products");
    std::vector<double> res(dots.size());
    copy_sync(stream, res.data(), rdm_res.p, res.size() * sizeof(double), hipMemcpyDeviceToHost, "2-RDM dot products");
    for (size_t e = 0; e < dots.size(); e++) dval[dots[e].d] += dots[e].coef * res[e];
  }

  // ---- the 2-RDM entries of this site (TwoDM.cpp:455-627)
  set4(A, i, i, i, i, 2 * dval[d1]);
  set4(B, i, i, i, i, -2 * dval[d1]);
  for (const auto& e : d2v) {
    set4(A, i, e.first, i, i, 2 * dval[e.second]);
    set4(B, i, e.first, i, i, -2 * dval[e.second]);
  }
  for (const auto& e : d36v) {
    const int j = e.first.first, k = e.first.second;
    const double d3 = dval[e.second[0]], d45 = dval[e.second[1]], d6 = dval[e.second[2]];
    set4(A, i, i, j, k, 2 * d3);
    set4(B, i, i, j, k, -2 * d3);
    set4(A, i, j, k, i, -2 * d45 - 3 * d6);
    set4(B, i, j, k, i, -2 * d45 + d6);
    set4(A, i, j, i, k, 4 * d45);
    set4(B, i, j, i, k, 2 * d6);
  }
  for (const auto& e : d7v) {
    set4(A, e.first, i, i, i, 2 * dval[e.second]);
    set4(B, e.first, i, i, i, -2 * dval[e.second]);
  }
  for (const auto& e : d812v) {
    const int g = e.first.first, j = e.first.second;
    const double d8 = dval[e.second[0]], d9 = dval[e.second[1]], d10 = dval[e.second[2]], d11 = dval[e.second[3]],
                 d12 = dval[e.second[4]];
    set4(A, g, i, j, i, -4 * d8 - d9);
    set4(A, g, i, i, j, 2 * d8 + d11);
    set4(B, g, i, j, i, d9 - 2 * d10);
    set4(B, g, i, i, j, 2 * d8 + 2 * d10 - d11);
    set4(A, g, j, i, i, 2 * d12);
    set4(B, g, j, i, i, -2 * d12);
  }
  for (const auto& e : d1324v) {
    const int g = std::get<0>(e.first), j = std::get<1>(e.first), k = std::get<2>(e.first);
    const double s0 = dval[e.second[0]], s1 = dval[e.second[1]], fa = dval[e.second[2]], ga = dval[e.second[3]],
                 fb = dval[e.second[4]], gb = dval[e.second[5]];
    set4(A, g, i, j, k, 2 * s0 + 3 * s1);
    set4(A, g, i, k, j, 2 * s0 - 3 * s1);
    set4(B, g, i, j, k, -2 * s0 + s1);
    set4(B, g, i, k, j, -2 * s0 - s1);
    set4(A, g, j, k, i, -2 * fa - 3 * ga);
    set4(A, g, j, i, k, 4 * fa);
    set4(B, g, j, k, i, -2 * fa + ga);
    set4(B, g, j, i, k, 2 * ga);
    set4(A, g, k, j, i, -2 * fb - 3 * gb);
    set4(A, g, k, i, j, 4 * fb);
    set4(B, g, k, j, i, -2 * fb + gb);
    set4(B, g, k, i, j, 2 * gb);
  }
}

// DMRG::calc_rdms_and_correlations, 2-RDM part (DMRGtechnics.cpp:40-110), from the state a
// right sweep leaves (sites 0..L-3 left-normalised). left_normalize(L-2, L-1) and
// right_normalize(i-1, i) are Split of the joined pair without truncation (the same state
// in another gauge of the bond, which the 2-RDM does not see); left_normalize(L-1, NULL)
// divides the last site by its norm (the R = |T| that TensorT::QR discards).
double Ctx::two_rdm(double* Aout, double* Bout) {
  if (!have_problem || L < 2) throw Error{ERR_STATE, "2-RDM: set the problem first"};
  if ((int)fci.size() != L + 1) throw Error{ERR_STATE, "2-RDM: no state (call chm_dmrg_init and solve first)"};
  for (int s = 0; s < L; s++)
    if (s >= (int)mps.size() || mps[s].buf < 0 || !bufs[mps[s].buf]) throw Error{ERR_MISSING, "2-RDM: MPS tensor of site " + std::to_string(s) + " not set"};
  for (int c = 0; c < L - 2; c++)
    if (sets[c].dir != 1) throw Error{ERR_STATE, "2-RDM: the operators are not those of a finished right sweep"};
  drop_plan();
  site = -1;
  const double t0 = now_s();
  join(L - 2);
  split(L - 2, 0, true, false);
  {
    const int64_t n = mps[L - 1].size;
    partial.alloc(std::max<int64_t>(dot_parts(n), 1));
    scal.alloc(8);
    launch_dot(mps_buf[L - 1], mps_buf[L - 1], n, partial.p, scal.p, stream);
    double nrm2 = 0.0;
    copy_sync(stream, &nrm2, scal.p, sizeof(double), hipMemcpyDeviceToHost, "2-RDM norm");
    if (!(nrm2 > 0.0) || !std::isfinite(nrm2)) throw Error{ERR_NUMERIC, "2-RDM: the state has no norm"};
    launch_scale(1.0 / std::sqrt(nrm2), mps_buf[L - 1], n, stream);
  }
  update_operators(L - 2, true);
  const size_t n4 = (size_t)L * L * L * L;
  std::vector<double> A(n4, 0.0), B(n4, 0.0);
  for (int s = L - 1; s >= 0; s--) {
    rdm_fill_site(s, A, B);
    if (s > 0) {
      join(s - 1);
      split(s - 1, 0, false, false);
      update_operators(s - 1, false);
    }
  }
  if (TwoS != 0) {  // TwoDM::correct_higher_multiplicities (TwoDM.cpp:630-640)
    const double a = 1.0 / (TwoS + 1.0);
    for (double& v : A) v *= a;
    for (double& v : B) v *= a;
  }
  rdm_seconds = now_s() - t0;
  double e = 0.0;  // TwoDM::energy (TwoDM.cpp:217-232)
  for (size_t q = 0; q < n4; q++) e += A[q] * mx[q];
  if (Aout) std::copy(A.begin(), A.end(), Aout);
  if (Bout) std::copy(B.begin(), B.end(), Bout);
  return 0.5 * e + econst;
}

}  // namespace chm
