// Excited states: lower states kept on the device, their overlap operators with the
// current MPS, and the projector vectors VeffTilde that Heff adds per site
// (DMRG::activateExcitations / newExcitation, DMRG.cpp:464-505; TensorO.cpp:39-193;
// DMRGoperators.cpp:558-567 and 891-900; DMRG::calcVeffTilde, DMRGtechnics.cpp:540-620).
#include <cmath>

#include "context.hpp"

namespace chm {

double* Ctx::own_buf(int32_t& id, int64_t n) {
  n = std::max<int64_t>(n, 1);
  if (id >= 0 && bufs[id] && buf_len[id] >= n) return bufs[id];
  if (id >= 0 && bufs[id]) {
    hip_check(hipStreamSynchronize(stream), "buffer regrow");
    (void)hipFree(bufs[id]);
    bufs[id] = nullptr;
    buf_len[id] = 0;
  }
  double* p = nullptr;
  hip_check(hipMalloc((void**)&p, (size_t)n * sizeof(double)), "hipMalloc");
  if (id < 0) {
    id = (int32_t)bufs.size();
    bufs.push_back(p);
    buf_len.push_back(n);
  } else {
    bufs[id] = p;
    buf_len[id] = n;
  }
  return p;
}

// Forgets every lower state and frees its device memory (a new DMRG object).
void Ctx::drop_lowers() {
  hip_check(hipStreamSynchronize(stream), "drop lower states");
  auto drop = [&](int32_t id) {
    if (id >= 0 && id < (int32_t)bufs.size() && bufs[id]) {
      (void)hipFree(bufs[id]);
      bufs[id] = nullptr;
      buf_len[id] = 0;
    }
  };
  for (auto& lw : lowers) {
    for (auto& t : lw.mps) drop(t.buf);
    for (auto& o : lw.ovl) drop(o.buf);
  }
  lowers.clear();
}

// Whether an overlap operator was built in direction dir on the current dimensions of
// its boundary (they change when a Split of a later step truncates that boundary).
bool Ctx::ovl_current(const OvlT& o, int dir) const {
  if (o.dir != dir || o.boundary < 0 || o.buf < 0) return false;
  const auto& cur = sec.list[o.boundary];
  if (cur.size() != o.pos2blk.size()) return false;
  for (size_t p = 0; p < cur.size(); p++) {
    const int k = o.pos2blk[p];
    if (k >= 0 && (o.N[k] != cur[p].N || o.S[k] != cur[p].S || o.I[k] != cur[p].I || o.du[k] != cur[p].d)) return false;
  }
  return true;
}

// The current state becomes lower state number lowers.size() with energy shift eshift;
// a new state starts from SyBookkeeper(Prob, D) dimensions and a random MPS, and the
// operators (with the overlaps to every lower state) are built moving right.
void Ctx::new_excitation(double eshift, int D, uint64_t seed) {
  if ((int)fci.size() != L + 1 || mps.empty()) throw Error{ERR_STATE, "no converged state to keep (call chm_dmrg_init first)"};
  if (!(eshift > 0.0)) throw Error{ERR_ARG, "the energy shift must be positive"};
  for (int s = 0; s < L; s++)
    if (mps[s].buf < 0 || !bufs[mps[s].buf]) throw Error{ERR_MISSING, "MPS tensor of site " + std::to_string(s) + " not set"};
  drop_plan();
  site = -1;
  s_site = -1;
  LowerState ls;
  ls.eshift = eshift;
  ls.sec = sec;
  ls.mps = mps;  // layouts and buffer ids; the buffers now belong to the lower state
  ls.ovl.assign(L - 1, OvlT());
  lowers.push_back(std::move(ls));
  for (int s = 0; s < L; s++) {
    mps[s].buf = -1;
    mps_buf[s] = nullptr;
  }
  // the boundary operators of the old state go (DMRG::deleteAllBoundaryOperators)
  for (int c = 0; c < L - 1; c++) clear_set(c);
  for (auto& lw : lowers)
    for (auto& o : lw.ovl) o.dir = 0;
  auto_sectors(*this, D);
  mps_random(seed);
  presolve();
}

// Overlap operators of every lower state at slot `index` (boundary index + 1), after the
// other operators of that slot: created from the MPS tensors at the chain ends
// (TensorO::create_right / create_left), otherwise updated from the neighbouring slot
// (TensorOperator::update_moving_right / left with the spin-0 identity: prefactor 1, or
// (2S_R + 1) / (2S + 1) for the singly occupied local states moving left).
void Ctx::update_overlaps(int index, bool mr) {
  if (lowers.empty()) return;
  const int b = index + 1;
  const int tsite = mr ? index : index + 1;
  const int Is = irreps[tsite];
  const bool create = mr ? index == 0 : index == L - 2;
  std::vector<UTerm> terms;
  std::vector<std::pair<Ref, std::pair<int, int>>> outs;
  for (auto& lw : lowers) {
    // layout on the current sectors of boundary b; the device buffer of the slot is kept
    // (and grown when needed)
    OvlT& o = lw.ovl[index];
    int32_t keep = o.buf;
    const auto& cur = sec.list[b];
    o = OvlT{};
    o.boundary = b;
    o.dir = mr ? 1 : 2;
    o.pos2blk.assign(cur.size(), -1);
    int64_t off = 0;
    o.k2i.push_back(0);
    for (size_t p = 0; p < cur.size(); p++) {
      const int dd = lw.sec.dim(b, cur[p].N, cur[p].S, cur[p].I);
      if (dd <= 0) continue;
      o.pos2blk[p] = (int)o.N.size();
      o.N.push_back(cur[p].N); o.S.push_back(cur[p].S); o.I.push_back(cur[p].I);
      o.du.push_back(cur[p].d); o.dd.push_back(dd);
      off += (int64_t)cur[p].d * dd;
      o.k2i.push_back(off);
    }
    o.size = off;
    own_buf(keep, o.size);
    o.buf = keep;
  }
  for (size_t q = 0; q < lowers.size(); q++) {
    LowerState& lw = lowers[q];
    OvlT& o = lw.ovl[index];
    const TLay& Tup = mps[tsite];
    const TLay& Tdn = lw.mps[tsite];
    const OvlT* prev = nullptr;
    if (!create) {
      prev = &lw.ovl[mr ? index - 1 : index + 1];
      if (!ovl_current(*prev, mr ? 1 : 2)) throw Error{ERR_STATE, "overlap operators of the neighbouring slot are missing"};
    }
    for (int k = 0; k < (int)o.N.size(); k++) {
      const Ref out{o.buf, o.k2i[k]};
      const int n = o.N[k], s = o.S[k], ir = o.I[k], du = o.du[k], dd = o.dd[k];
      outs.push_back({out, {du, dd}});
      // local cases g: 0 empty, 1 doubly occupied, 2 and 3 singly occupied (spin -1, +1)
      for (int g = 0; g < 4; g++) {
        int n2 = n, s2 = s, i2 = ir;
        const int dn = mr ? -1 : 1;
        if (g == 1) n2 = n + 2 * dn;
        if (g >= 2) {
          n2 = n + dn;
          s2 = s + (g == 2 ? -1 : 1);
          i2 = ir ^ Is;
        }
        if (s2 < 0) continue;
        const int bo = mr ? b - 1 : b + 1;  // the other boundary of the site tensor
        const int ou = sec.dim(bo, n2, s2, i2), od = lw.sec.dim(bo, n2, s2, i2);
        if (ou <= 0 || od <= 0) continue;
        UTerm t{};
        t.out = out;
        t.m = du;
        t.n = dd;
        if (mr) {
          // O(R) += T_up(L->R)^T O_prev(L) T_down(L->R)
          const Ref A = Tup.blk(n2, s2, i2, n, s, ir), C = Tdn.blk(n2, s2, i2, n, s, ir);
          if (!A.ok() || !C.ok()) throw Error{ERR_MISSING, "overlap: missing MPS block"};
          t.alpha = 1.0;
          if (create) {
            if (ou != od) continue;
            t.kind = 0; t.k1 = ou;
            t.A = A; t.ta = 1; t.lda = ou;
            t.B = C; t.tb = 0; t.ldb = od;
          } else {
            const Ref B = prev->blk(sec, n2, s2, i2);
            if (!B.ok()) continue;
            t.kind = 1; t.k1 = ou; t.k2 = od;
            t.A = A; t.ta = 1; t.lda = ou;
            t.B = B; t.tb = 0; t.ldb = ou;
            t.C = C; t.tc = 0; t.ldc = od;
          }
        } else {
          // O(L) += alpha T_up(L->R) O_prev(R) T_down(L->R)^T
          const Ref A = Tup.blk(n, s, ir, n2, s2, i2), C = Tdn.blk(n, s, ir, n2, s2, i2);
          if (!A.ok() || !C.ok()) throw Error{ERR_MISSING, "overlap: missing MPS block"};
          t.alpha = (g >= 2) ? (s2 + 1.0) / (s + 1.0) : 1.0;
          if (create) {
            if (ou != od) continue;
            t.kind = 0; t.k1 = ou;
            t.A = A; t.ta = 0; t.lda = du;
            t.B = C; t.tb = 1; t.ldb = dd;
          } else {
            const Ref B = prev->blk(sec, n2, s2, i2);
            if (!B.ok()) continue;
            t.kind = 1; t.k1 = ou; t.k2 = od;
            t.A = A; t.ta = 0; t.lda = du;
            t.B = B; t.tb = 0; t.ldb = ou;
            t.C = C; t.tc = 1; t.ldc = dd;
          }
        }
        terms.push_back(t);
      }
    }
  }
  run_terms(terms, outs, nullptr);
}

// VeffTilde of every lower state at the prepared site (DMRG::calcVeffTilde): the lower
// state's two-site object Sup (joined from its MPS) contracted with the left and right
// overlaps, sqrt(Eshift) / (2S + 1) * sqrt(2S_R + 1) * O_L Sup O_R^T per block of the
// current two-site object (symmetric convention), zero where Sup has no such block.
// The vectors go to the lower-state projector of Heff.
void Ctx::veff_tilde(int index) {
  if (lowers.empty()) return;
  if (site != index) throw Error{ERR_STATE, "veff_tilde: the site is not prepared"};
  const int nq = (int)lowers.size();
  own_buf(low_buf, (int64_t)nq * V);
  std::vector<UTerm> terms;
  std::vector<std::pair<Ref, std::pair<int, int>>> outs;
  std::vector<SLayout> sups(nq);
  int64_t sup_tot = 0;
  std::vector<int64_t> sup_off(nq);
  for (int q = 0; q < nq; q++) {
    sups[q].build(lowers[q].sec, index, irreps[index], irreps[index + 1]);
    sup_off[q] = sup_tot;
    sup_tot += sups[q].size();
  }
  own_buf(sup_buf, sup_tot);
  for (int q = 0; q < nq; q++) {
    // the lower state's two-site object, joined into its part of the Sup buffer
    SLayout shifted = sups[q];
    for (auto& v : shifted.k2i) v += sup_off[q];
    join_terms(lowers[q].sec, lowers[q].mps[index], lowers[q].mps[index + 1], index, shifted, sup_buf);
  }
  for (int q = 0; q < nq; q++) {
    const LowerState& lw = lowers[q];
    const OvlT* OL = index > 0 ? &lw.ovl[index - 1] : nullptr;
    const OvlT* OR = index < L - 2 ? &lw.ovl[index + 1] : nullptr;
    if (OL && !ovl_current(*OL, 1)) throw Error{ERR_STATE, "left overlap operators are missing"};
    if (OR && !ovl_current(*OR, 2)) throw Error{ERR_STATE, "right overlap operators are missing"};
    const double pre = std::sqrt(lw.eshift) / (TwoS + 1.0);
    const SLayout& Su = sups[q];
    for (int ik = 0; ik < S.nk(); ik++) {
      const int nl = S.NL[ik], sl = S.SL[ik], il = S.IL[ik], nr = S.NR[ik], sr = S.SR[ik], ir = S.IR[ik];
      const Ref out{low_buf, (int64_t)q * V + S.k2i[ik]};
      const int dLd = S.dL[ik], dRd = S.dR[ik];
      outs.push_back({out, {dLd, dRd}});
      const int ku = Su.kappa(nl, sl, il, S.N1[ik], S.N2[ik], S.TJ[ik], nr, sr, ir);
      if (ku < 0) continue;
      const int dLu = Su.dL[ku], dRu = Su.dR[ku];
      const Ref B{sup_buf, sup_off[q] + Su.k2i[ku]};
      UTerm t{};
      t.out = out;
      t.m = dLd;
      t.n = dRd;
      t.alpha = pre * std::sqrt(sr + 1.0);
      const Ref A = OL ? OL->blk(sec, nl, sl, il) : Ref{};
      const Ref C = OR ? OR->blk(sec, nr, sr, ir) : Ref{};
      if ((OL && !A.ok()) || (OR && !C.ok())) continue;
      if (OL && OR) {
        t.kind = 1; t.k1 = dLu; t.k2 = dRu;
        t.A = A; t.ta = 0; t.lda = dLd;
        t.B = B; t.tb = 0; t.ldb = dLu;
        t.C = C; t.tc = 1; t.ldc = dRd;
      } else if (OL) {  // right end: dRu = dRd = 1
        t.kind = 0; t.k1 = dLu;
        t.A = A; t.ta = 0; t.lda = dLd;
        t.B = B; t.tb = 0; t.ldb = dLu;
      } else if (OR) {  // left end: dLu = dLd = 1
        t.kind = 0; t.k1 = dRu;
        t.A = B; t.ta = 0; t.lda = dLu;
        t.B = C; t.tb = 1; t.ldb = dRd;
      } else {
        t.kind = 2;
        t.A = B; t.ta = 0; t.lda = dLu;
      }
      terms.push_back(t);
    }
  }
  run_terms(terms, outs, nullptr);
  // hand the vectors to the Heff projector (HeffDiagrams1.cpp:65-86, HeffDiagonal.cpp:621-642)
  lower.alloc((size_t)nq * V);
  hip_check(hipMemcpyAsync(lower.p, bufs[low_buf], (size_t)nq * V * sizeof(double), hipMemcpyDeviceToDevice, stream),
            "lower-state vectors");
  std::vector<const double*> ptrs(nq);
  for (int q = 0; q < nq; q++) ptrs[q] = lower.p + (int64_t)q * V;
  lower_ptrs.upload(ptrs, stream);
  lower_part.alloc((size_t)dot_parts(V) * nq);
  lower_dot.alloc((size_t)nq);
  n_lower = nq;
}

}  // namespace chm
