// Context, device work-list construction, the Davidson driver and the C ABI.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>

#include "../../include/chemps2_amd.h"
#include "context.hpp"
#include "parallel.hpp"

namespace chm {

void launch_fill_hash(double* p, int64_t n, uint64_t seed, double scale, hipStream_t s);
void destroy_rocblas(void* h);

Ctx::~Ctx() {
  drop_plan();
  reap_join();
  destroy_rocblas(rocblas);
  for (auto h : svd_handles) destroy_rocblas(h);
  for (auto e : svd_ev_out)
    if (e) (void)hipEventDestroy(e);
  if (svd_ev_in) (void)hipEventDestroy(svd_ev_in);
  for (auto st : svd_streams)
    if (st) (void)hipStreamDestroy(st);
  if (s_buf >= 0) bufs[s_buf] = nullptr;  // owned by Sdev
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : mv_ev)
    if (e) (void)hipEventDestroy(e);
  if (dav_host) (void)hipHostFree(dav_host);
  if (stream) (void)hipStreamDestroy(stream);
}

void Ctx::set_problem(int L_, int group_, const int* irr, int N_, int TwoS_, int Irrep_, const double* mx_,
                      double econst_) {
  static const int nirr_of[8] = {1, 2, 2, 2, 4, 4, 4, 8};
  if (L_ < 2 || group_ < 0 || group_ > 7 || N_ < 0 || TwoS_ < 0 || N_ > 2 * L_ || (N_ % 2) != (TwoS_ % 2))
    throw Error{ERR_ARG, "inconsistent problem (L, group, N, 2S)"};
  L = L_; group = group_; nirrep = nirr_of[group_]; N = N_; TwoS = TwoS_; Irrep = Irrep_;
  if (Irrep < 0 || Irrep >= nirrep) throw Error{ERR_ARG, "target irrep out of range"};
  irreps.assign(irr, irr + L);
  for (int v : irreps)
    if (v < 0 || v >= nirrep) throw Error{ERR_ARG, "orbital irrep out of range"};
  mx.assign(mx_, mx_ + (size_t)L * L * L * L);
  econst = econst_;
  sec.resize(L);
  fci.clear();
  if (s_buf >= 0) bufs[s_buf] = nullptr;  // owned by Sdev
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  bufs.assign(BUF_TENSOR0, nullptr);
  buf_len.assign(BUF_TENSOR0, 0);
  Sdev.release();
  s_buf = -1;
  s_site = -1;
  lowers.clear();
  low_buf = sup_buf = -1;
  sets.clear();
  sets.resize(L - 1);
  mps.assign(L, TLay());
  mps_buf.assign(L, nullptr);
  site = -1;
  drop_plan();
  prealloc.clear();
  reap_join();
  plan.reset();
  have_problem = true;
}

void Ctx::set_sectors(int b, int n, const int* Ns, const int* Ss, const int* Is, const int* ds) {
  if (!have_problem) throw Error{ERR_STATE, "set the problem first"};
  if (b < 0 || b > L || n < 0) throw Error{ERR_ARG, "boundary out of range"};
  std::vector<Sectors::Sec> v;
  for (int q = 0; q < n; q++) {
    if (ds[q] <= 0) continue;
    if (Ns[q] < 0 || Ss[q] < 0 || Is[q] < 0 || Is[q] >= nirrep) throw Error{ERR_ARG, "bad sector label"};
    v.push_back({Ns[q], Ss[q], Is[q], ds[q]});
  }
  std::sort(v.begin(), v.end(), [](const Sectors::Sec& a, const Sectors::Sec& b2) {
    return std::make_tuple(a.N, a.S, a.I) < std::make_tuple(b2.N, b2.S, b2.I);
  });
  sec.set(b, std::move(v));
  site = -1;
}

OpT* Ctx::add_tensor(int cnt, bool mr, int kind, int k, int m, int two_j, int n_elec, int n_irrep, int64_t n) {
  if (!have_problem) throw Error{ERR_STATE, "set the problem first"};
  if (cnt < 0 || cnt > L - 2) throw Error{ERR_ARG, "slot out of range"};
  if (kind < 0 || kind >= KIND_COUNT || k < 0 || m < 0) throw Error{ERR_ARG, "bad tensor index"};
  OpSet& s = sets[cnt];
  const int dir = mr ? 1 : 2;
  if (s.dir != dir) clear_set(cnt);
  s.dir = dir;
  auto owned = std::make_unique<OpT>();
  owned->build(sec, cnt + 1, two_j, n_elec, n_irrep);
  if (n >= 0 && owned->size != n)
    throw Error{ERR_SIZE, "tensor size " + std::to_string(n) + " != layout size " + std::to_string(owned->size)};
  OpT* t = owned.get();
  auto slot1 = [&](std::vector<OpT*>& v) -> OpT*& {
    if ((int)v.size() <= k) v.resize(k + 1, nullptr);
    return v[k];
  };
  auto slot2 = [&](std::vector<std::vector<OpT*>>& v) -> OpT*& {
    if ((int)v.size() <= k) v.resize(k + 1);
    if ((int)v[k].size() <= m) v[k].resize(m + 1, nullptr);
    return v[k][m];
  };
  OpT** where = nullptr;
  switch (kind) {
    case KIND_L: where = &slot1(s.L); break;
    case KIND_Q: where = &slot1(s.Q); break;
    case KIND_X: where = &s.X; break;
    case KIND_F0: where = &slot2(s.F0); break;
    case KIND_F1: where = &slot2(s.F1); break;
    case KIND_S0: where = &slot2(s.S0); break;
    case KIND_S1: where = &slot2(s.S1); break;
    case KIND_A: where = &slot2(s.A); break;
    case KIND_B: where = &slot2(s.B); break;
    case KIND_C: where = &slot2(s.C); break;
    case KIND_D: where = &slot2(s.D); break;
  }
  if (*where) {  // replace: free the old device copy
    if ((*where)->buf >= 0 && bufs[(*where)->buf]) {
      (void)hipFree(bufs[(*where)->buf]);
      bufs[(*where)->buf] = nullptr;
      buf_len[(*where)->buf] = 0;
    }
  }
  double* p = nullptr;
  if (t->size > 0) hip_check(hipMalloc((void**)&p, t->size * sizeof(double)), "hipMalloc tensor");
  t->buf = (int32_t)bufs.size();
  bufs.push_back(p);
  buf_len.push_back(t->size);
  *where = t;
  s.owned.push_back(std::move(owned));
  site = -1;
  return t;
}

void Ctx::clear_set(int cnt) {
  OpSet& s = sets[cnt];
  for (auto& o : s.owned)
    if (o->buf >= 0 && bufs[o->buf]) {
      (void)hipFree(bufs[o->buf]);
      bufs[o->buf] = nullptr;
      buf_len[o->buf] = 0;
    }
  s = OpSet();
  site = -1;
}

const double* Ctx::resolve_with(const Ref& r, const double* combo_base) const {
  if (r.buf == BUF_COMBO) return combo_base + r.off;
  if (r.buf >= BUF_TENSOR0 && r.buf < (int)bufs.size() && bufs[r.buf]) return bufs[r.buf] + r.off;
  throw Error{ERR_MISSING, "unresolvable block reference"};
}

const double* Ctx::resolve(const Ref& r) const {
  if (r.buf == BUF_COMBO) return combo_arena.p + r.off;
  if (r.buf >= BUF_TENSOR0 && r.buf < (int)bufs.size() && bufs[r.buf]) return bufs[r.buf] + r.off;
  throw Error{ERR_MISSING, "unresolvable block reference"};
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void Ctx::start_plan(int st) {
  drop_plan();
  if (st < 0 || st > L - 2) return;
  OpSet* Lo = st > 0 ? &sets[st - 1] : nullptr;
  OpSet* Ro = st < L - 2 ? &sets[st + 1] : nullptr;
  job = std::make_unique<PlanJob>();
  PlanJob* j = job.get();
  j->site = st;
  j->gen = sec.gen;
  j->S.build(sec, st, irreps[st], irreps[st + 1]);
  j->plan = std::make_unique<HeffPlanner>();
  j->plan->init(&sec, &irreps, nirrep, mx.data(), L, &j->S, Lo, Ro);
  j->plan->pool = std::move(term_pool);
  j->th = std::thread([j]() {
    try {
      j->plan->build();
    } catch (...) {
      j->err = std::current_exception();
    }
  });
}

void Ctx::drop_plan() {
  if (!job) return;
  if (job->th.joinable()) job->th.join();
  job.reset();
}

// The layout part of update_operators: the new slot's tensors with the flags of
// DMRG::allocateTensors, before their values are computed.
void Ctx::alloc_slot(int index, bool mr) {
  const int prepared = site;  // a prepared Heff reads slots site - 1 and site + 1 only
  clear_set(index);
  for (const OpShape& sh : opset_shapes(*this, index, mr)) {
    OpT* t = add_tensor(index, mr, sh.kind, sh.k, sh.m, sh.two_j, sh.n_elec, sh.n_irrep, -1);
    t->moving_right = mr;
    t->jw_phase = (sh.kind == KIND_L || sh.kind == KIND_Q);
    t->prime_last = (sh.kind == KIND_F1 || sh.kind == KIND_D) ? mr : true;
  }
  sets[index].gen = sec.gen;
  if ((int)prealloc.size() < L - 1) prealloc.assign(L - 1, 0);
  prealloc[index] = 1;
  if (prepared >= 0 && index != prepared - 1 && index != prepared + 1) site = prepared;
}

void Ctx::apply_tuning() {
  site_dmax = 0;
  for (int k = 0; k < S.nk(); k++) site_dmax = std::max<int64_t>(site_dmax, std::max(S.dL[k], S.dR[k]));
  explicit_max = def_explicit_max;
  chain_item_terms = def_item_terms;
  chain_group_terms = def_group_terms;
  use_chain_mfma = def_chain_mfma;
  fchain_min = def_fchain_min;
  tune_row = -1;
  for (size_t r = 0; r < tune.size(); r++) {
    const TuneRow& t = tune[r];
    if (site_dmax < t.dmax_lo || site_dmax > t.dmax_hi || V < t.v_lo || V > t.v_hi) continue;
    if (t.explicit_max >= 0 && !env_pinned[0]) explicit_max = t.explicit_max;
    if (t.item_terms > 0 && !env_pinned[1]) chain_item_terms = (int)t.item_terms;
    if (t.group_terms > 0 && !env_pinned[2]) chain_group_terms = (int)t.group_terms;
    if (t.chain_mfma >= 0 && !env_pinned[3]) use_chain_mfma = t.chain_mfma != 0;
    if (t.fchain_min >= 0 && !env_pinned[4]) fchain_min = (int)t.fchain_min;
    tune_row = (int)r;
    break;
  }
}

void Ctx::prepare(int st) {
  const double t_start = now_s();
  if (!have_problem) throw Error{ERR_STATE, "set the problem first"};
  if (st < 0 || st > L - 2) throw Error{ERR_ARG, "site out of range"};
  site = -1;
  n_lower = 0;
  OpSet* Lo = nullptr;
  OpSet* Ro = nullptr;
  if (st > 0) { validate_set(st - 1, true); Lo = &sets[st - 1]; }
  if (st < L - 2) { validate_set(st + 1, false); Ro = &sets[st + 1]; }
  bool planned = false;
  if (job && job->site == st) {
    job->th.join();
    if (!job->err && job->gen == sec.gen && job->plan->Lo == Lo && job->plan->Ro == Ro) {
      S = std::move(job->S);
      plan = std::move(job->plan);
      plan->S = &S;
      planned = true;
    }
    job.reset();
  } else {
    drop_plan();
  }
  if (!planned) {
    S.build(sec, st, irreps[st], irreps[st + 1]);
    plan = std::make_unique<HeffPlanner>();
    plan->init(&sec, &irreps, nirrep, mx.data(), L, &S, Lo, Ro);
    plan->pool = std::move(term_pool);
    plan->build();
  }
  V = S.size();
  if (V <= 0) throw Error{ERR_STATE, "empty two-site object"};
  apply_tuning();
  F_H = plan->flops;
  const double t_built = now_s();

  // ---- x-independent combinations, evaluated now
  combo_arena.alloc(std::max<int64_t>(plan->combo_size, 1));
  {
    const auto& cv = plan->combos;
    const int64_t ncb = (int64_t)cv.size();
    std::vector<DevCombo> dc(ncb);
    std::vector<int64_t> chunk_start(ncb), s_start(ncb + 1, 0);
    int64_t nch = 0;
    for (int64_t q = 0; q < ncb; q++) {
      chunk_start[q] = nch;
      nch += (cv[q].size + 1023) / 1024;
      s_start[q + 1] = s_start[q] + (int64_t)cv[q].src.size();
    }
    if (s_start[ncb] > INT32_MAX) throw Error{ERR_SIZE, "too many combination sources"};
    std::vector<DevSrc> ds(s_start[ncb]);
    parallel_for((ncb + 4095) / 4096, host_threads(), [&](int64_t c, int) {
      for (int64_t q = c * 4096; q < std::min<int64_t>(ncb, (c + 1) * 4096); q++) {
        DevCombo& d = dc[q];
        d.out = combo_arena.p + cv[q].out_off;
        d.size = cv[q].size;
        d.src0 = (int)s_start[q];
        d.nsrc = (int)cv[q].src.size();
        int64_t e = s_start[q];
        for (auto& s : cv[q].src) ds[e++] = {resolve(s.first), s.second};
      }
    });
    if (!dc.empty()) {
      cmb_dc.upload(dc, stream);
      cmb_ds.upload(ds, stream);
      cmb_cs.upload(chunk_start, stream);
      launch_combos(cmb_dc.p, cmb_ds.p, (int)dc.size(), cmb_cs.p, nch, stream);
      hip_check(hipGetLastError(), "combination kernel launch");
      hip_check(hipStreamSynchronize(stream), "combination kernel");
    }
  }

  const double t_plan = now_s();
  // ---- split the terms: (kappa, kappa') pairs small enough are assembled into an
  // explicit block-sparse matrix once per site; the rest run as chained GEMM tiles.
  // Passes over the planner's chunks (output blocks [c * CH, (c + 1) * CH)) run on the
  // host threads; every output position comes from prefix sums over the chunks, so the
  // lists do not depend on the thread count.
  auto dLof = [&](int k) { return S.dL[k]; };
  auto dRof = [&](int k) { return S.dR[k]; };
  auto bsize = [&](int k) { return (int64_t)(S.k2i[k + 1] - S.k2i[k]); };
  const int nthr = host_threads();
  const std::vector<std::vector<Term>>& chunks = plan->chunks;
  const int64_t nch = (int64_t)chunks.size();
  const int CH = HeffPlanner::CHUNK;
  const int nk = S.nk();
  using Grp = PGrp;
  constexpr int NTHR = 64;
  if ((int64_t)p_grp.size() < nch) p_grp.resize(nch);
  std::vector<std::vector<Grp>>& grp = p_grp;
  std::vector<std::array<double, NTHR>> hist(nch);
  std::vector<std::vector<char>> seen_t(nthr, std::vector<char>(bufs.size(), 0));
  parallel_for(nch, nthr, [&](int64_t c, int tid) {
    auto& g = grp[c];
    g.clear();
    const auto& tv = chunks[c];
    std::vector<char>& seen = seen_t[tid];
    for (int64_t q = 0; q < (int64_t)tv.size(); q++) {
      const Term& t = tv[q];
      if (t.Lb.buf >= BUF_TENSOR0) seen[t.Lb.buf] = 1;
      if (t.Rb.buf >= BUF_TENSOR0) seen[t.Rb.buf] = 1;
      if (g.empty() || g.back().out != t.out || g.back().in != t.in) g.push_back({t.out, t.in, q, q, bsize(t.out) * bsize(t.in)});
      g.back().q1 = q + 1;
    }
    auto& h = hist[c];
    h.fill(0.0);
    for (const Grp& x : g)
      for (int k = 0; k < NTHR; k++) {
        if ((explicit_max >> k) < x.elems) break;
        h[k] += (double)x.elems;
      }
  });
  int64_t thr = 0;
  {
    size_t freeb = 0, totb = 0;
    hip_check(hipMemGetInfo(&freeb, &totb), "hipMemGetInfo");
    const double avail = (double)freeb + 8.0 * (double)Mvals.n;  // the matrix of the last site is reused
    const double budget = explicit_budget_gb > 0 ? explicit_budget_gb * 1e9 : 0.4 * avail;
    for (int k = 0; k < NTHR && (explicit_max >> k) > 0; k++) {
      double nnz = 0.0;
      for (int64_t c = 0; c < nch; c++) nnz += hist[c][k];
      if (nnz * 8.0 <= budget) { thr = explicit_max >> k; break; }
    }
  }
  explicit_threshold = thr;
  // counts per chunk, then offsets
  struct Cnt { int64_t pairs = 0, kt = 0, cols = 0, m = 0, chain = 0; };
  std::vector<Cnt> cnt(nch + 1);
  parallel_for(nch, nthr, [&](int64_t c, int) {
    Cnt n;
    for (const Grp& x : grp[c]) {
      if (x.elems > thr) { n.chain += x.q1 - x.q0; continue; }
      n.pairs++;
      n.kt += x.q1 - x.q0;
      n.cols += bsize(x.in);
      n.m += (int64_t)bsize(x.out) * bsize(x.in);
    }
    cnt[c + 1] = n;
  });
  for (int64_t c = 0; c < nch; c++) {
    cnt[c + 1].pairs += cnt[c].pairs; cnt[c + 1].kt += cnt[c].kt; cnt[c + 1].cols += cnt[c].cols;
    cnt[c + 1].m += cnt[c].m; cnt[c + 1].chain += cnt[c].chain;
  }
  const Cnt tot = cnt[nch];
  if (tot.kt > INT32_MAX) throw Error{ERR_SIZE, "too many explicit terms"};
  stage.reset();
  stage.reserve(256 * 4 + tot.pairs * sizeof(DevPair) + tot.kt * sizeof(DevKTerm) + std::max<int64_t>(tot.cols, 1) * sizeof(int));
  DevPair* h_dp = stage.take<DevPair>(tot.pairs);
  DevKTerm* h_kt = stage.take<DevKTerm>(tot.kt);
  int* h_col = stage.take<int>(std::max<int64_t>(tot.cols, 1));
  h_col[0] = 0;
  std::vector<DevRowBlock> rb(nk);
  std::vector<const Term*> chain(tot.chain);
  parallel_for(nch, nthr, [&](int64_t c, int) {
    int64_t ip = cnt[c].pairs, ik = cnt[c].kt, ic = cnt[c].cols, im = cnt[c].m, ich = cnt[c].chain;
    const auto& tv = chunks[c];
    const auto& g = grp[c];
    size_t h = 0;
    for (int k = (int)(c * CH); k < std::min<int>(nk, (int)((c + 1) * CH)); k++) {
      const int R = (int)bsize(k);
      rb[k].moff = im;
      rb[k].coff = ic;
      rb[k].yoff = S.k2i[k];
      rb[k].R = R;
      int cpos = 0;
      for (; h < g.size() && g[h].out == k; h++) {
        const Grp& x = g[h];
        if (x.elems > thr) {
          for (int64_t q = x.q0; q < x.q1; q++) chain[ich++] = &tv[q];
          continue;
        }
        DevPair pr{};
        pr.moff = im + (int64_t)R * cpos;
        pr.R = R; pr.dL = dLof(k); pr.dR = dRof(k); pr.dLi = dLof(x.in); pr.dRi = dRof(x.in);
        pr.t0 = (int)ik;
        for (int64_t q = x.q0; q < x.q1; q++) {
          const Term& t = tv[q];
          DevKTerm e{};
          e.alpha = t.alpha;
          if (t.tl != 'I') { e.A = resolve(plan->ref(t, t.Lb)); e.lda = (t.tl == 'N') ? pr.dL : pr.dLi; e.ta = (t.tl == 'T'); }
          if (t.tr != 'I') { e.B = resolve(plan->ref(t, t.Rb)); e.ldb = (t.tr == 'N') ? pr.dRi : pr.dR; e.tb = (t.tr == 'T'); }
          h_kt[ik++] = e;
        }
        pr.nt = (int)(ik - pr.t0);
        h_dp[ip++] = pr;
        const int64_t base = S.k2i[x.in];
        for (int q = 0; q < (int)bsize(x.in); q++) h_col[ic++] = (int)(base + q);
        cpos += (int)bsize(x.in);
      }
      rb[k].C = cpos;
      im += (int64_t)R * cpos;
    }
  });
  const double t_lists = now_s();
  {
    Mvals.alloc(std::max<int64_t>(tot.m, 1));
    M_nnz = tot.m;
    M_split = heff_form == 1;
    Mcol.alloc(std::max<int64_t>(tot.cols, 1));
    hip_check(hipMemcpyAsync(Mcol.p, h_col, std::max<int64_t>(tot.cols, 1) * sizeof(int), hipMemcpyHostToDevice, stream), "columns");
    // work items: rows in chunks of at most 256, columns in chunks of about
    // item_elems elements, so every workgroup streams a similar share of the matrix
    std::vector<DevBItem> items;
    std::vector<DevBRed> reds;
    int64_t poff = 0;
    for (int k = 0; k < nk; k++) {
      const DevRowBlock& B = rb[k];
      for (int r0 = 0; r0 < B.R; r0 += 256) {
        const int R = std::min(256, B.R - r0);
        const int64_t cols_per = std::min<int64_t>(4096, std::max<int64_t>(1, item_elems / R));  // LDS staging limit
        const int nparts = B.C == 0 ? 1 : (int)((B.C + cols_per - 1) / cols_per);
        for (int q = 0; q < nparts; q++) {
          DevBItem it{};
          it.moff = B.moff + r0;
          it.coff = B.coff;
          it.ld = B.R;
          it.R = R;
          it.c0 = (int)std::min<int64_t>(B.C, q * cols_per);
          it.c1 = (int)std::min<int64_t>(B.C, (q + 1) * cols_per);
          it.direct = nparts == 1;
          it.out = it.direct ? B.yoff + r0 : poff + (int64_t)q * R;
          items.push_back(it);
        }
        if (nparts > 1) {
          reds.push_back({B.yoff + r0, poff, R, nparts});
          poff += (int64_t)nparts * R;
        }
      }
    }
    bitems.upload(items, stream);
    breds.upload(reds, stream);
    nbitems = (int)items.size();
    nbreds = (int)reds.size();
    bpart.alloc(std::max<int64_t>(poff, 1));
    bpart_need = poff;
    if (tot.pairs > 0) {
      d_pairs.alloc(tot.pairs);
      d_kterms.alloc(tot.kt);
      hip_check(hipMemcpyAsync(d_pairs.p, h_dp, tot.pairs * sizeof(DevPair), hipMemcpyHostToDevice, stream), "pairs");
      hip_check(hipMemcpyAsync(d_kterms.p, h_kt, tot.kt * sizeof(DevKTerm), hipMemcpyHostToDevice, stream), "terms");
      launch_assemble(d_pairs.p, d_kterms.p, tot.pairs, Mvals.p, tot.m, M_split, stream);
      hip_check(hipGetLastError(), "assembly kernel launch");
    }
    hip_check(hipStreamSynchronize(stream), "assembly");
    stats[11] = tot.m;
    stats[12] = tot.cols;
    stats[13] = thr;
    stats[14] = (int64_t)nbitems;
  }
  const double t_assemble = now_s();

  // ---- chained terms: output blocks whose own and input dimensions are at most
  // CHAIN_MAXD run in the output-stationary kernel (two launches: block products up to
  // CHAIN_SMALLP, and larger ones when the device grants the LDS); the rest as GEMM tiles
  // the output-stationary chained kernels (k_chain, k_chain_mfma, k_fchain): in the
  // mixed form by default (its FP32 products are k_fchain's), else only when asked for;
  // the GEMM tiles are faster on the FP64 form
  const bool chain_on = use_chain_kernel < 0 ? heff_form == 1 : use_chain_kernel != 0;
  if (!chain_on) {  // every chained term stays for the GEMM tiles; the chained lists are empty
    ncblocks = 0;
    ncblocks_l = 0;
    for (int bk = 0; bk < FCHAIN_NB; bk++) ncblocks_f[bk] = 0;
    for (int n = 0; n <= CHAIN_MFMA_NT; n++) cblocks_m_offs[n] = 0;
    ncreds = 0;
    cpart_need = 0;
    stats[15] = 0;
  } else {
    // output blocks in chunks on the host threads, each with local records; the chunks
    // are then concatenated in order (group, term and partial offsets shifted), so the
    // lists do not depend on the thread count
    struct Part {
      std::vector<DevCBlock> cb, cbl, cbm[CHAIN_MFMA_NT], cbf[FCHAIN_NB];  // cbf: k_fchain, by tile shape
      std::vector<DevCGroup> cg;
      std::vector<DevCTerm> ct;
      std::vector<DevBRed> cr;
      std::vector<const Term*> rest;
      int64_t cpart_n = 0;
      int Pmax = 1, PmaxL = 1;
    };
    std::vector<size_t> bstart;  // first chain index of every output block
    for (size_t q = 0; q < chain.size(); q++)
      if (q == 0 || chain[q]->out != chain[q - 1]->out) bstart.push_back(q);
    bstart.push_back(chain.size());
    const int64_t nblk = (int64_t)bstart.size() - 1;
    const int64_t nparts = std::max<int64_t>(1, std::min<int64_t>(256, (nblk + 15) / 16));
    std::vector<Part> parts(nparts);
    const int maxd = chain_large_ok ? CHAIN_MAXD : 36;
    // blocks up to fchain_maxd (at most FCHAIN_MAXD) can take the fused MFMA kernel: those above maxd always,
    // the others when their largest dimension is above fchain_min
    const int fmaxd = std::min(fchain_maxd, FCHAIN_MAXD);
    const int lim = use_fchain ? std::max(maxd, fmaxd) : maxd;
    parallel_for(nparts, nthr, [&](int64_t pi, int) {
      Part& P = parts[pi];
      std::vector<DevCBlock>&cb = P.cb, &cbl = P.cbl;
      std::vector<DevCBlock>* cbm = P.cbm;
      std::vector<DevCBlock>* cbf = P.cbf;
      std::vector<DevCGroup>& cg = P.cg;
      std::vector<DevCTerm>& ct = P.ct;
      std::vector<DevBRed>& cr = P.cr;
      std::vector<const Term*>& rest = P.rest;
      int64_t& cpart_n = P.cpart_n;
      int &Pmax = P.Pmax, &PmaxL = P.PmaxL;
      const size_t qb = bstart[nblk * pi / nparts], qe = bstart[nblk * (pi + 1) / nparts];
      for (size_t q0 = qb; q0 < qe;) {
        const int k = chain[q0]->out;
        size_t q1 = q0;
        bool small = chain_on && dLof(k) <= lim && dRof(k) <= lim;
        int dmax = std::max(dLof(k), dRof(k));
        while (q1 < qe && chain[q1]->out == k) {
          const int in = chain[q1]->in;
          if (dLof(in) > lim || dRof(in) > lim) small = false;
          dmax = std::max({dmax, dLof(in), dRof(in)});
          q1++;
        }
        if (!small) {
          for (size_t q = q0; q < q1; q++) rest.push_back(chain[q]);
          q0 = q1;
          continue;
        }
        DevCBlock B{};
        B.dL = dLof(k); B.dR = dRof(k);
        const bool fused = use_fchain && dmax <= fmaxd && (dmax > maxd || dmax > fchain_min);
        int Pk = B.dL * B.dR;
        std::vector<DevCBlock> items;
        // k_fchain computes the terms of the association op(A) (x op(B)) as the transposed
        // block: those form items of their own (class 1), with operators exchanged
        for (int cls = 0; cls < (fused ? 2 : 1); cls++) {
          const int g_first = (int)cg.size();
          for (size_t a = q0; a < q1;) {
            const int in = chain[a]->in;
            size_t e = a;
            while (e < q1 && chain[e]->in == in) e++;
            DevCGroup G{};
            G.xoff = S.k2i[in]; G.dLi = dLof(in); G.dRi = dRof(in); G.t0 = (int)ct.size();
            Pk = std::max({Pk, G.dLi * G.dRi, B.dL * G.dRi, G.dLi * B.dR, B.dL * G.dLi, G.dRi * B.dR});
            for (size_t q = a; q < e; q++) {
              const Term& t = *chain[q];
              const int dL = B.dL, dR = B.dR, dLi = G.dLi, dRi = G.dRi;
              DevCTerm T{};
              T.alpha = t.alpha;
              if (t.tl != 'I') { T.A = resolve(plan->ref(t, t.Lb)); T.lda = (t.tl == 'N') ? dL : dLi; T.ta = (t.tl == 'T'); }
              if (t.tr != 'I') { T.B = resolve(plan->ref(t, t.Rb)); T.ldb = (t.tr == 'N') ? dRi : dR; T.tb = (t.tr == 'T'); }
              if (t.tl == 'I' && t.tr == 'I') T.mode = 0;
              else if (t.tr == 'I') T.mode = 1;
              else if (t.tl == 'I') T.mode = 2;
              else {
                const double c3 = (double)dL * dRi * dLi + (double)dL * dR * dRi, c4 = (double)dLi * dR * dRi + (double)dL * dR * dLi;
                T.mode = c3 <= c4 ? 3 : 4;
              }
              if (fused && (T.mode == 4) != (cls == 1)) continue;
              if (cls == 1) {  // y^T += a op(B)^T x^T op(A)^T
                std::swap(T.A, T.B);
                std::swap(T.lda, T.ldb);
                const int ta = T.ta;
                T.ta = !T.tb;
                T.tb = !ta;
                T.mode = 3;
              }
              ct.push_back(T);
            }
            // groups of at most CHAIN_GROUP_TERMS terms (x is staged once per group)
            const int t_end = (int)ct.size();
            for (int t0 = G.t0; t0 < t_end; t0 += chain_group_terms) {
              DevCGroup Gs = G;
              Gs.t0 = t0;
              Gs.nt = std::min(chain_group_terms, t_end - t0);
              cg.push_back(Gs);
            }
            a = e;
          }
          // work items of about CHAIN_ITEM_TERMS terms; a block with several items sums
          // their partials in k_chain_reduce
          for (int g = g_first; g < (int)cg.size();) {
            DevCBlock it = B;
            it.g0 = g;
            it.tr = cls;
            int nt = 0;
            while (g < (int)cg.size() && (nt == 0 || nt + cg[g].nt <= chain_item_terms)) nt += cg[g++].nt;
            it.ng = g - it.g0;
            items.push_back(it);
          }
        }
        const int64_t nel = (int64_t)B.dL * B.dR;
        if (items.size() == 1) {
          items[0].direct = 1;
          items[0].out = S.k2i[k];
        } else if (items.size() > 1) {
          cr.push_back({S.k2i[k], cpart_n, (int)nel, (int)items.size()});
          for (auto& it : items) { it.direct = 0; it.out = cpart_n; cpart_n += nel; }
        }
        const int nt = (dmax + 15) / 16;
        if (fused) {
          const int ntf = std::max(nt, FCHAIN_NT0);
          for (auto& it : items) cbf[fchain_bucket(ntf, ((it.tr ? B.dL : B.dR) + 15) / 16)].push_back(it);
        } else if (use_chain_mfma && nt <= CHAIN_MFMA_NT) {
          for (auto& it : items) cbm[nt - 1].push_back(it);
        } else {
          for (auto& it : items) {
            if (Pk <= CHAIN_SMALLP) cb.push_back(it);
            else cbl.push_back(it);
          }
          if (Pk <= CHAIN_SMALLP) Pmax = std::max(Pmax, Pk);
          else PmaxL = std::max(PmaxL, Pk);
        }
        q0 = q1;
      }
    });
    // offsets of every part in the concatenated lists (cbm: all parts of tile count 1,
    // then of 2, ...), then a parallel copy into page-locked staging with the indices
    // shifted, and asynchronous uploads (complete at the synchronisation ending prepare)
    enum { L_CG, L_CT, L_CR, L_CB, L_CBL, L_REST, L_F0, L_M0 = L_F0 + FCHAIN_NB };
    const int NL = L_M0 + CHAIN_MFMA_NT;
    std::vector<int64_t> off((size_t)NL * (nparts + 1), 0), poff(nparts + 1, 0);
    auto O = [&](int l, int64_t pi) -> int64_t& { return off[(size_t)l * (nparts + 1) + pi]; };
    int Pmax = 1, PmaxL = 1;
    for (int64_t pi = 0; pi < nparts; pi++) {
      const Part& P = parts[pi];
      size_t sz[L_M0] = {P.cg.size(), P.ct.size(), P.cr.size(), P.cb.size(), P.cbl.size(), P.rest.size()};
      for (int bk = 0; bk < FCHAIN_NB; bk++) sz[L_F0 + bk] = P.cbf[bk].size();
      for (int l = 0; l < L_M0; l++) O(l, pi + 1) = O(l, pi) + (int64_t)sz[l];
      for (int n = 0; n < CHAIN_MFMA_NT; n++) O(L_M0 + n, pi + 1) = O(L_M0 + n, pi) + (int64_t)P.cbm[n].size();
      poff[pi + 1] = poff[pi] + P.cpart_n;
      Pmax = std::max(Pmax, P.Pmax);
      PmaxL = std::max(PmaxL, P.PmaxL);
    }
    const int64_t n_cg = O(L_CG, nparts), n_ct = O(L_CT, nparts), n_cr = O(L_CR, nparts), n_cb = O(L_CB, nparts),
                  n_cbl = O(L_CBL, nparts), n_rest = O(L_REST, nparts), cpart_n = poff[nparts];
    int64_t n_f[FCHAIN_NB], n_fall = 0;
    for (int bk = 0; bk < FCHAIN_NB; bk++) { n_f[bk] = O(L_F0 + bk, nparts); n_fall += n_f[bk]; }
    int64_t n_m[CHAIN_MFMA_NT], n_mall = 0;
    for (int n = 0; n < CHAIN_MFMA_NT; n++) { n_m[n] = O(L_M0 + n, nparts); n_mall += n_m[n]; }
    if (n_ct > INT32_MAX || n_cg > INT32_MAX) throw Error{ERR_SIZE, "too many chained terms"};
    cstage.reset();
    cstage.reserve(8 * 256 + n_cg * sizeof(DevCGroup) + n_ct * sizeof(DevCTerm) + n_cr * sizeof(DevBRed) +
                   (n_cb + n_cbl + n_mall + n_fall) * sizeof(DevCBlock));
    DevCGroup* hcg = cstage.take<DevCGroup>(n_cg);
    DevCTerm* hct = cstage.take<DevCTerm>(n_ct);
    DevBRed* hcr = cstage.take<DevBRed>(n_cr);
    DevCBlock* hcb = cstage.take<DevCBlock>(n_cb);
    DevCBlock* hcbl = cstage.take<DevCBlock>(n_cbl);
    DevCBlock* hcbm = cstage.take<DevCBlock>(n_mall);
    DevCBlock* hcbf[FCHAIN_NB];
    for (int bk = 0; bk < FCHAIN_NB; bk++) hcbf[bk] = cstage.take<DevCBlock>(n_f[bk]);
    std::vector<const Term*> rest(n_rest);
    parallel_for(nparts, nthr, [&](int64_t pi, int) {
      const Part& P = parts[pi];
      const int g_off = (int)O(L_CG, pi), t_off = (int)O(L_CT, pi);
      const int64_t p_off = poff[pi];
      DevCGroup* g = hcg + O(L_CG, pi);
      for (DevCGroup x : P.cg) { x.t0 += t_off; *g++ = x; }
      if (!P.ct.empty()) std::memcpy(hct + O(L_CT, pi), P.ct.data(), P.ct.size() * sizeof(DevCTerm));
      DevBRed* r = hcr + O(L_CR, pi);
      for (DevBRed x : P.cr) { x.p0 += p_off; *r++ = x; }
      auto shift = [&](DevCBlock* dst, const std::vector<DevCBlock>& src) {
        for (DevCBlock b : src) {
          b.g0 += g_off;
          if (!b.direct) b.out += p_off;
          *dst++ = b;
        }
      };
      shift(hcb + O(L_CB, pi), P.cb);
      shift(hcbl + O(L_CBL, pi), P.cbl);
      for (int bk = 0; bk < FCHAIN_NB; bk++) shift(hcbf[bk] + O(L_F0 + bk, pi), P.cbf[bk]);
      int64_t mbase = 0;
      for (int n = 0; n < CHAIN_MFMA_NT; n++) {
        shift(hcbm + mbase + O(L_M0 + n, pi), P.cbm[n]);
        mbase += n_m[n];
      }
      std::copy(P.rest.begin(), P.rest.end(), rest.begin() + O(L_REST, pi));
    });
    auto up = [&](auto& dbuf, const auto* h, int64_t n) {
      dbuf.alloc(std::max<int64_t>(n, 1));
      if (n > 0) hip_check(hipMemcpyAsync(dbuf.p, h, n * sizeof(*h), hipMemcpyHostToDevice, stream), "chain lists");
    };
    up(cblocks, hcb, n_cb);
    up(cblocks_m, hcbm, n_mall);
    cblocks_m_offs[0] = 0;
    for (int n = 0; n < CHAIN_MFMA_NT; n++) cblocks_m_offs[n + 1] = cblocks_m_offs[n] + (int)n_m[n];
    up(cgroups, hcg, n_cg);
    up(cterms, hct, n_ct);
    ncblocks = (int)n_cb;
    chain_P = std::min((Pmax + 31) / 32 * 32, CHAIN_SMALLP);
    up(cblocks_l, hcbl, n_cbl);
    ncblocks_l = (int)n_cbl;
    for (int bk = 0; bk < FCHAIN_NB; bk++) {
      up(cblocks_f[bk], hcbf[bk], n_f[bk]);
      ncblocks_f[bk] = (int)n_f[bk];
    }
    up(creds, hcr, n_cr);
    ncreds = (int)n_cr;
    cpart.alloc(std::max<int64_t>(cpart_n, 1));
    cpart_need = cpart_n;
    chain_P_l = std::min((PmaxL + 31) / 32 * 32, CHAIN_MAXD * CHAIN_MAXD);
    stats[15] = n_ct;
    chain.swap(rest);
  }
  const double t_route = now_s();
  // ---- chained GEMM tiles for the remaining terms (in (out, in) order)
  std::vector<DevGemmOp> o1;  // phase-1 products (host only: copied into the tile records)
  DevTile1* t1 = nullptr;      // the device records, built in page-locked staging
  DevGemmOp* o2 = nullptr;
  DevTile* t2 = nullptr;
  struct WRec { int64_t off; int rows, cols; };
  std::vector<WRec> wrec(chain.size(), {-1, 0, 0});
  auto rel_x = [&](int k) { return (const double*)(intptr_t)S.k2i[k]; };
  // The intermediates W of one batch of output blocks share the scratch buffer, so that
  // it stays within a cap (one block larger than that is a batch of its own); apply()
  // runs phase 1 and phase 2 batch after batch. The cap is tile_w_cap doubles if set,
  // else a quarter of the free memory (the current buffer counted as free), at least
  // 4 GiB.
  int64_t TILE_W_CAP = tile_w_cap;
  if (TILE_W_CAP <= 0) {
    size_t freeb = 0, totb = 0;
    hip_check(hipMemGetInfo(&freeb, &totb), "hipMemGetInfo");
    TILE_W_CAP = std::max<int64_t>((int64_t)1 << 29, (int64_t)((freeb + 8.0 * W.n) / 4 / 8));
  }
  auto wneed = [&](const Term& t) -> int64_t {
    if (t.tl == 'I' || t.tr == 'I') return 0;
    const int dL = dLof(t.out), dR = dRof(t.out), dLi = dLof(t.in), dRi = dRof(t.in);
    return t.xr_first ? ((int64_t)dLi * dR + 7) / 8 * 8 : ((int64_t)dL * dRi + 7) / 8 * 8;
  };
  std::vector<size_t> bstart;  // first chain index of every batch
  int64_t Wmax = 0;
  {
    int64_t cur = 0;
    for (size_t q0 = 0; q0 < chain.size();) {
      size_t q1 = q0;
      int64_t need = 0;
      while (q1 < chain.size() && chain[q1]->out == chain[q0]->out) need += wneed(*chain[q1++]);
      if (bstart.empty() || (cur > 0 && cur + need > TILE_W_CAP)) { bstart.push_back(q0); cur = 0; }
      cur += need;
      Wmax = std::max(Wmax, cur);
      q0 = q1;
    }
  }
  // Index pass: every two-operand term's phase-1 op and its W offset, and the phase-1
  // tiles per batch and per bucket (tile_bucket), whose ranges tb1 holds as
  // [B b + c, B b + c + 1) with B = TILE_BUCKETS. It runs on the host threads over chunks
  // of terms that stay within one batch (counts, a prefix over the chunks, then the
  // offsets); the records are filled the same way after the ops.
  const size_t nq = chain.size();
  std::vector<int> opi1(nq, -1);
  tb1.assign(1, 0);
  tb2.assign(1, 0);
  // tiles of 64 x 64 by bucket, or (wave_tiles) of 32 x 32, all on wavefronts
  const int ts1 = wave_tiles ? 32 : 64;
  auto each_tile = [&](const WRec& w, auto&& f) {
    for (int m0 = 0; m0 < w.rows; m0 += ts1)
      for (int n0 = 0; n0 < w.cols; n0 += ts1)
        f(m0, n0, wave_tiles ? 0 : tile_bucket(std::min(64, w.rows - m0), std::min(64, w.cols - n0)));
  };
  auto w_of = [&](const Term& t) -> WRec {
    return t.xr_first ? WRec{0, dLof(t.in), dRof(t.out)} : WRec{0, dLof(t.out), dRof(t.in)};
  };
  struct QChunk {
    size_t q0, q1;
    int batch;
    int64_t ops = 0, w = 0;
    std::array<int64_t, TILE_BUCKETS> tiles{};
  };
  std::vector<QChunk> qch;
  const int nbatch = std::max<int>(1, (int)bstart.size());
  for (int bt = 0; bt < nbatch; bt++) {
    const size_t b0 = bstart.empty() ? 0 : bstart[bt], b1 = (bt + 1 < (int)bstart.size()) ? bstart[bt + 1] : nq;
    for (size_t q = b0; q < b1; q += 8192) qch.push_back(QChunk{q, std::min(b1, q + 8192), bt});
  }
  parallel_for((int64_t)qch.size(), nthr, [&](int64_t ci, int) {
    QChunk& C = qch[ci];
    for (size_t q = C.q0; q < C.q1; q++) {
      const Term& t = *chain[q];
      if (t.tl == 'I' || t.tr == 'I') continue;
      const WRec w = w_of(t);
      C.ops++;
      C.w += ((int64_t)w.rows * w.cols + 7) / 8 * 8;
      each_tile(w, [&](int, int, int c) { C.tiles[c]++; });
    }
  });
  // prefix over the chunks: op index and W offset (per batch) at every chunk's start,
  // tiles per batch and bucket
  std::vector<int64_t> cops(qch.size()), cw(qch.size());
  std::vector<std::array<int64_t, TILE_BUCKETS>> bcount(nbatch, std::array<int64_t, TILE_BUCKETS>{});
  std::vector<std::array<int64_t, TILE_BUCKETS>> ctile(qch.size());
  int no1 = 0;
  {
    int64_t ops = 0, wcur = 0;
    int bt = -1;
    for (size_t ci = 0; ci < qch.size(); ci++) {
      if (qch[ci].batch != bt) { bt = qch[ci].batch; wcur = 0; }
      cops[ci] = ops;
      cw[ci] = wcur;
      ops += qch[ci].ops;
      wcur += qch[ci].w;
      for (int c = 0; c < TILE_BUCKETS; c++) {
        ctile[ci][c] = bcount[bt][c];  // position within (batch, bucket), shifted below
        bcount[bt][c] += qch[ci].tiles[c];
      }
    }
    no1 = (int)ops;
  }
  int64_t nt1 = 0;
  for (auto& bc : bcount)
    for (int c = 0; c < TILE_BUCKETS; c++) {
      nt1 += bc[c];
      tb1.push_back((int)std::min<int64_t>(nt1, INT32_MAX));
    }
  for (size_t ci = 0; ci < qch.size(); ci++)
    for (int c = 0; c < TILE_BUCKETS; c++) ctile[ci][c] += tb1[TILE_BUCKETS * qch[ci].batch + c];
  parallel_for((int64_t)qch.size(), nthr, [&](int64_t ci, int) {
    const QChunk& C = qch[ci];
    int64_t op = cops[ci], wo = cw[ci];
    for (size_t q = C.q0; q < C.q1; q++) {
      const Term& t = *chain[q];
      if (t.tl == 'I' || t.tr == 'I') continue;
      WRec w = w_of(t);
      w.off = wo;
      wrec[q] = w;
      wo += ((int64_t)w.rows * w.cols + 7) / 8 * 8;
      opi1[q] = (int)op++;
    }
  });
  // Phase 2: an output block whose products are long (in 16-deep k chunks per tile) is
  // split into segments of consecutive products, so that the launch has many more
  // workgroups than the device holds at once; each segment writes its partial block into
  // the partial-sum buffer and k_chain_reduce adds them to y.
  auto chunks_of = [&](const Term& t) -> int64_t {
    if (t.tl == 'I' && t.tr == 'I') return 0;
    const int k = (t.tr == 'I' || (t.tl != 'I' && t.xr_first)) ? dLof(t.in) : dRof(t.in);
    return (k + 15) / 16;
  };
  // output blocks (the chained terms are in (out, in) order) and their work, on the host
  // threads over chunks of terms
  std::vector<size_t> bq;      // first term of every output block
  std::vector<int64_t> bwork;  // 16-deep k chunks per tile of every output block
  std::vector<int> cq(nq);     // chunks of every term
  {
    const int64_t nc = ((int64_t)nq + 16383) / 16384;
    std::vector<std::vector<size_t>> starts(nc);
    parallel_for(nc, nthr, [&](int64_t c, int) {
      const size_t q0 = (size_t)c * 16384, q1 = std::min(nq, q0 + 16384);
      for (size_t q = q0; q < q1; q++) {
        cq[q] = (int)chunks_of(*chain[q]);
        if (q == 0 || chain[q]->out != chain[q - 1]->out) starts[c].push_back(q);
      }
    });
    for (auto& v : starts) bq.insert(bq.end(), v.begin(), v.end());
    bq.push_back(nq);
    bwork.resize(bq.size() - 1);
    parallel_for((int64_t)bwork.size(), nthr, [&](int64_t bi, int) {
      int64_t w = 0;
      for (size_t q = bq[bi]; q < bq[bi + 1]; q++) w += cq[q];
      bwork[bi] = w;
    });
  }
  const int64_t nblk = (int64_t)bwork.size();
  std::vector<int> bseg(nblk, 1);
  {
    int64_t wtot = 0;
    for (int64_t bi = 0; bi < nblk; bi++) {
      const int k = chain[bq[bi]]->out;
      wtot += bwork[bi] * ((dLof(k) + 63) / 64) * ((dRof(k) + 63) / 64);
    }
    const int64_t wseg = std::max<int64_t>(std::max(1, tile_seg_min), wtot / std::max(1, tile_segments));
    for (int64_t bi = 0; bi < nblk; bi++) bseg[bi] = (int)std::max<int64_t>(1, std::min<int64_t>(bwork[bi] / wseg, 64));
  }
  // Segment boundaries (about equal chunk counts) and whether a segment holds an axpy
  // product, per block on the host threads; then the tiles per batch in two ranges,
  // tb2[2 b] (tiles of at most 32 x 32 without axpy: one wavefront each, k_tiles_wave2)
  // and tb2[2 b + 1] (the rest, k_tiles_fit).
  std::vector<int64_t> segq0(nblk + 1, 0);  // first segment of every block in segq / segax
  for (int64_t bi = 0; bi < nblk; bi++) segq0[bi + 1] = segq0[bi] + bseg[bi];
  std::vector<size_t> segq(segq0[nblk] + 1);  // first term of every segment
  std::vector<char> segax(segq0[nblk]);
  parallel_for(nblk, nthr, [&](int64_t bi, int) {
    const size_t q0 = bq[bi], q1 = bq[bi + 1];
    const int ns = bseg[bi];
    size_t qa = q0;
    int64_t done = 0;
    for (int sg = 0; sg < ns; sg++) {
      size_t qb = qa;
      const int64_t goal = bwork[bi] * (sg + 1) / ns;
      if (sg == ns - 1) qb = q1;
      else
        while (qb < q1 && done < goal) done += cq[qb++];
      char ax = 0;
      for (size_t q = qa; q < qb; q++) ax |= (chain[q]->tl == 'I' && chain[q]->tr == 'I') ? 1 : 0;
      segq[segq0[bi] + sg] = qa;
      segax[segq0[bi] + sg] = ax;
      qa = qb;
    }
  });
  segq[segq0[nblk]] = nq;
  // the tiles of an output block (segment): 64 x 64, those of at most 32 x 32 without axpy
  // on wavefronts; or (wave_tiles) 32 x 32 on wavefronts unless the segment has an axpy
  auto each_tile2 = [&](int dL, int dR, bool ax, auto&& f) {
    const bool q32 = wave_tiles && !ax;
    const int st = q32 ? 32 : 64;
    for (int m0 = 0; m0 < dL; m0 += st)
      for (int n0 = 0; n0 < dR; n0 += st)
        f(m0, n0, q32 || (!ax && std::min(64, dL - m0) <= 32 && std::min(64, dR - n0) <= 32));
  };
  std::vector<int64_t> tile2w(nblk), tile2f(nblk), bpart(nblk, -1);
  int64_t nt2 = 0, tpart_n = 0;
  std::vector<DevBRed> tred;
  {
    // per batch: wavefront tiles, then the rest
    size_t nextb = 1;
    int64_t bb = 0;  // first block of the batch
    auto close_batch = [&](int64_t be) {
      int64_t nw = 0, nf = 0;
      for (int64_t bi = bb; bi < be; bi++) {
        const int k = chain[bq[bi]]->out;
        const int dL = dLof(k), dR = dRof(k);
        int64_t cw = 0, cf = 0;
        for (int64_t sgi = segq0[bi]; sgi < segq0[bi + 1]; sgi++)
          each_tile2(dL, dR, segax[sgi] != 0, [&](int, int, bool wv) { (wv ? cw : cf)++; });
        tile2w[bi] = nt2 + nw;
        tile2f[bi] = nf;  // shifted below, once the batch's wavefront tiles are counted
        nw += cw;
        nf += cf;
      }
      for (int64_t bi = bb; bi < be; bi++) tile2f[bi] += nt2 + nw;
      tb2.push_back((int)(nt2 + nw));
      nt2 += nw + nf;
      tb2.push_back((int)nt2);
      bb = be;
    };
    for (int64_t bi = 0; bi < nblk; bi++) {
      if (nextb < bstart.size() && bq[bi] == bstart[nextb]) {
        close_batch(bi);
        nextb++;
      }
      const int k = chain[bq[bi]]->out;
      const int64_t nel = (int64_t)dLof(k) * dRof(k);
      if (bseg[bi] > 1) {
        bpart[bi] = tpart_n;
        tred.push_back({S.k2i[k], cpart_need + tpart_n, (int)nel, bseg[bi]});
        tpart_n += nel * bseg[bi];
      }
    }
    close_batch(nblk);
  }
  if (nt1 > INT32_MAX || nt2 > INT32_MAX || nq > (size_t)INT32_MAX) throw Error{ERR_SIZE, "too many Heff tiles"};
  tpart_off = cpart_need;
  cpart_need += tpart_n;
  cpart.alloc(std::max<int64_t>(cpart_need, 1));
  treds.upload(tred, stream);
  ntreds = (int)tred.size();
  W.alloc(std::max<int64_t>(Wmax, 1));
  o1.resize(no1);
  tstage.reset();
  tstage.reserve(4 * 256 + nt1 * sizeof(DevTile1) + nq * sizeof(DevGemmOp) + nt2 * sizeof(DevTile));
  t1 = tstage.take<DevTile1>(nt1);
  o2 = tstage.take<DevGemmOp>(nq);
  t2 = tstage.take<DevTile>(nt2);
  // phase 1: one group per two-operand term
  parallel_for(((int64_t)nq + 4095) / 4096, nthr, [&](int64_t c, int) {
    const size_t qe = std::min(nq, (size_t)(c + 1) * 4096);
    for (size_t q = (size_t)c * 4096; q < qe; q++) {
      if (opi1[q] < 0) continue;
      const Term& t = *chain[q];
      const int dL = dLof(t.out), dR = dRof(t.out), dLi = dLof(t.in), dRi = dRof(t.in);
      DevGemmOp op{};
      if (!t.xr_first) {
        op.A = resolve(plan->ref(t, t.Lb)); op.lda = (t.tl == 'N') ? dL : dLi; op.ta = (t.tl == 'T');
        op.B = rel_x(t.in); op.ldb = dLi; op.tb = 0; op.k = dLi; op.flags = 2;
      } else {
        op.A = rel_x(t.in); op.lda = dLi; op.ta = 0;
        op.B = resolve(plan->ref(t, t.Rb)); op.ldb = (t.tr == 'N') ? dRi : dR; op.tb = (t.tr == 'T'); op.k = dRi; op.flags = 1;
      }
      op.alpha = t.alpha;
      op.kind = 0;
      if (op.k <= 0) throw Error{ERR_STATE, "Heff plan: a phase-1 product with no inner dimension"};
      o1[opi1[q]] = op;
    }
  });
  parallel_for((int64_t)qch.size(), nthr, [&](int64_t ci, int) {
    std::array<int64_t, TILE_BUCKETS> pos = ctile[ci];
    for (size_t q = qch[ci].q0; q < qch[ci].q1; q++) {
      if (opi1[q] < 0) continue;
      const WRec& w = wrec[q];
      const DevGemmOp& op = o1[opi1[q]];
      each_tile(w, [&](int m0, int n0, int c) {
        DevTile1 tl{};
        tl.C = W.p + w.off;
        tl.A = op.A; tl.B = op.B; tl.alpha = op.alpha;
        tl.ldc = w.rows; tl.m0 = m0; tl.n0 = n0; tl.mrem = std::min(ts1, w.rows - m0); tl.nrem = std::min(ts1, w.cols - n0);
        tl.lda = op.lda; tl.ldb = op.ldb; tl.k = op.k; tl.ta = op.ta; tl.tb = op.tb; tl.flags = op.flags;
        t1[pos[c]++] = tl;
      });
    }
  });
  // phase 2: one group per output block (or per segment of it), added to the explicit
  // product
  parallel_for(nblk, nthr, [&](int64_t bi, int) {
    const size_t q0 = bq[bi], q1 = bq[bi + 1];
    const int k = chain[q0]->out;
    const int dL = dLof(k), dR = dRof(k);
    for (size_t q = q0; q < q1; q++) {
      const Term& t = *chain[q];
      const int dLi = dLof(t.in), dRi = dRof(t.in);
      DevGemmOp op{};
      op.alpha = t.alpha;
      if (t.tl == 'I' && t.tr == 'I') {
        op.kind = 1; op.A = rel_x(t.in); op.lda = dL; op.flags = 1;
      } else if (t.tr == 'I') {
        op.A = resolve(plan->ref(t, t.Lb)); op.lda = (t.tl == 'N') ? dL : dLi; op.ta = (t.tl == 'T');
        op.B = rel_x(t.in); op.ldb = dLi; op.k = dLi; op.flags = 2;
      } else if (t.tl == 'I') {
        op.A = rel_x(t.in); op.lda = dLi; op.ta = 0;
        op.B = resolve(plan->ref(t, t.Rb)); op.ldb = (t.tr == 'N') ? dRi : dR; op.tb = (t.tr == 'T'); op.k = dRi; op.flags = 1;
      } else if (!t.xr_first) {
        const WRec& w = wrec[q];
        op.alpha = 1.0;
        op.A = W.p + w.off; op.lda = w.rows; op.ta = 0;
        op.B = resolve(plan->ref(t, t.Rb)); op.ldb = (t.tr == 'N') ? dRi : dR; op.tb = (t.tr == 'T'); op.k = dRi;
      } else {
        const WRec& w = wrec[q];
        op.alpha = 1.0;
        op.A = resolve(plan->ref(t, t.Lb)); op.lda = (t.tl == 'N') ? dL : dLi; op.ta = (t.tl == 'T');
        op.B = W.p + w.off; op.ldb = w.rows; op.tb = 0; op.k = dLi;
      }
      o2[q] = op;
    }
    // segments as the index pass cut them
    const int ns = bseg[bi];
    int64_t tw = tile2w[bi], tf = tile2f[bi];
    for (int sg = 0; sg < ns; sg++) {
      const int64_t sgi = segq0[bi] + sg;
      const size_t qa = segq[sgi], qb = (sg == ns - 1) ? q1 : segq[sgi + 1];
      const bool axpy = segax[sgi] != 0;
      each_tile2(dL, dR, axpy, [&](int m0, int n0, bool wv) {
          DevTile tl{};
          if (ns == 1) {
            tl.C = (double*)(intptr_t)S.k2i[k];
            tl.flags = 3;
          } else {
            tl.C = (double*)(intptr_t)(tpart_off + bpart[bi] + (int64_t)sg * dL * dR);
            tl.flags = TILE_C_PART;
          }
          if (!axpy) tl.flags |= TILE_NO_AXPY;
          tl.ldc = dL; tl.m0 = m0; tl.n0 = n0; tl.m = dL; tl.n = dR; tl.op0 = (int)qa; tl.nop = (int)(qb - qa);
          t2[wv ? tw++ : tf++] = tl;
        });
    }
  });
  {  // asynchronous uploads from page-locked memory (complete at the synchronisation ending prepare)
    auto up = [&](auto& dbuf, const auto* h, int64_t n) {
      dbuf.alloc(std::max<int64_t>(n, 1));
      if (n > 0) hip_check(hipMemcpyAsync(dbuf.p, h, n * sizeof(*h), hipMemcpyHostToDevice, stream), "tile lists");
    };
    up(ops2, o2, (int64_t)nq);
    up(tiles1, t1, nt1);
    up(tiles2, t2, nt2);
  }
  ntiles1 = (int)nt1;
  ntiles2 = (int)nt2;
  stats[6] = tot.pairs;
  stats[7] = (int64_t)chain.size() + stats[15];

  const double t_tiles = now_s();
  // ---- diagonal
  {
    std::vector<DevDiagBlock> blocks(S.nk());
    std::vector<DevDiag> terms;
    std::vector<std::vector<const DiagTerm*>> per(S.nk());
    for (const DiagTerm& d : plan->diag) per[d.k].push_back(&d);
    for (int k = 0; k < S.nk(); k++) {
      blocks[k].out = (double*)(intptr_t)S.k2i[k];
      blocks[k].dL = dLof(k);
      blocks[k].dR = dRof(k);
      blocks[k].t0 = (int)terms.size();
      for (const DiagTerm* d : per[k]) {
        DevDiag t{};
        t.kind = d->kind;
        t.coef = d->coef;
        t.L = d->Lb.ok() ? resolve(d->Lb) : nullptr;
        t.R = d->Rb.ok() ? resolve(d->Rb) : nullptr;
        terms.push_back(t);
      }
      blocks[k].nt = (int)terms.size() - blocks[k].t0;
    }
    hblocks = blocks;
    dblocks.alloc(blocks.size());
    dterms.upload(terms, stream);
    stats[10] = (int64_t)terms.size();
  }

  const double t_diag = now_s();
  // ---- program <-> symmetric convention factors (Sobject.cpp:624-650)
  {
    std::vector<double> fs(S.nk()), fp(S.nk());
    for (int k = 0; k < S.nk(); k++) { fs[k] = std::sqrt(S.SR[k] + 1.0); fp[k] = 1.0 / fs[k]; }
    d_sym.upload(fs, stream);
    d_prog.upload(fp, stream);
    d_k2i.upload(S.k2i, stream);
  }
  dx.alloc(V);
  dy.alloc(V);
  partial.alloc((size_t)dot_parts(V) * 64);
  scal.alloc(256);
  hip_check(hipStreamSynchronize(stream), "prepare");
  const double t_sync = now_s();

  // ---- algorithmic bytes: the vectors and every operator tensor the step reads
  {
    std::vector<char> seen(bufs.size(), 0);
    for (const auto& st2 : seen_t)
      for (size_t b = 0; b < seen.size(); b++) seen[b] |= st2[b];
    auto mark = [&](const Ref& r) { if (r.buf >= BUF_TENSOR0) seen[r.buf] = 1; };
    for (const Combo& c : plan->combos) for (auto& s : c.src) mark(s.first);
    double bytes = 16.0 * V;
    for (size_t b = 0; b < bufs.size(); b++) if (seen[b]) bytes += 8.0 * buf_len[b];
    B_H = bytes;
  }
  stats[0] = plan->nterms;
  {
    const double t_end = now_s();
    const double ph[6] = {t_built - t_start, t_plan - t_built, t_lists - t_plan, t_assemble - t_lists, t_end - t_assemble,
                          t_end - t_start};
    for (int q = 0; q < 6; q++) prep_s[q] = ph[q];
    if (verbose)
    {
      int maxd = 0;
      for (int k = 0; k < S.nk(); k++) maxd = std::max({maxd, S.dL[k], S.dR[k]});
      std::fprintf(stderr, "prepare site %d: plan %.3f s (generate %.3f, merge %.3f, finish %.3f), combinations %.3f s, lists %.3f s, assembly %.3f s, rest %.3f s (chain %.3f, tiles %.3f, diagonal %.3f, sync %.3f, bytes %.3f), total %.3f s; "
                   "max block dim %d, terms explicit %lld / chain kernel %lld / tiles %lld\n",
                   st, ph[0], plan->phase_s[0], plan->phase_s[1], plan->phase_s[2], ph[1], ph[2], ph[3], ph[4],
                   t_route - t_assemble, t_tiles - t_route, t_diag - t_tiles, t_sync - t_diag, t_end - t_sync, ph[5], maxd,
                   (long long)(plan->nterms - stats[7]), (long long)stats[15], (long long)(stats[7] - stats[15]));
    }
  }
  stats[1] = (int64_t)plan->combos.size();
  stats[2] = plan->combo_size;
  term_pool = std::move(plan->chunks);  // their capacity serves the next site
  reap(std::move(plan));
  stats[3] = Wmax;
  stats[4] = ntiles1;
  stats[5] = ntiles2;
  stats[8] = S.nk();
  stats[9] = V;
  site = st;
}

void Ctx::apply(const double* xp, double* yp) { apply_with(xp, yp, bpart.p, cpart.p, lower_part.p, lower_dot.p); }

// The product with the given partial-sum scratch: bp (explicit rows split over work
// items), cp (chained blocks split over work items), lp / ld (lower-state dots).
void Ctx::apply_with(const double* xp, double* yp, double* bp, double* cp, double* lp, double* ld) {
  if (site < 0) throw Error{ERR_STATE, "no prepared site"};
  if (timing) hip_check(hipEventRecord(ev[0], stream), "event");
  launch_bspmv(bitems.p, nbitems, breds.p, nbreds, bp, Mvals.p, M_nnz, M_split ? (mv_low ? 1 : 2) : 0, Mcol.p, xp,
               yp, stream);
  if (timing) hip_check(hipEventRecord(ev[1], stream), "event");
  for (size_t b = 0; 2 * b + 2 < tb2.size(); b++) {  // GEMM tiles, batch by batch (W is shared)
    for (int c = 0; c < TILE_BUCKETS; c++) {
      const int i = TILE_BUCKETS * (int)b + c;
      launch_tiles_stream1(tiles1.p + tb1[i], tb1[i + 1] - tb1[i], c, xp, stream);
    }
    const int w0 = tb2[2 * b], w1 = tb2[2 * b + 1], f1 = tb2[2 * b + 2];
    launch_tiles_wave2(tiles2.p + w0, ops2.p, w1 - w0, xp, yp, stream, cp);
    launch_tiles(tiles2.p + w1, ops2.p, f1 - w1, xp, yp, stream, tiles_kernel, cp);
  }
  launch_chain_reduce(treds.p, ntreds, cp, yp, stream);  // segments of split output blocks
  if (timing) hip_check(hipEventRecord(ev[2], stream), "event");
  launch_chain_mfma(cblocks_m.p, cblocks_m_offs, cgroups.p, cterms.p, xp, yp, cp, stream);
  launch_chain(cblocks.p, ncblocks, cgroups.p, cterms.p, xp, yp, cp, chain_P, false, stream);
  launch_chain(cblocks_l.p, ncblocks_l, cgroups.p, cterms.p, xp, yp, cp, chain_P_l, true, stream);
  for (int bk = 0; bk < FCHAIN_NB; bk++)
    launch_fchain(cblocks_f[bk].p, ncblocks_f[bk], cgroups.p, cterms.p, xp, yp, cp, bk, mv_low && fchain_fp32, stream);
  launch_chain_reduce(creds.p, ncreds, cp, yp, stream);
  if (n_lower > 0) {  // lower states, in state order (HeffDiagrams1.cpp:65-86)
    launch_multi_dot(lower_ptrs.p, n_lower, xp, V, lp, ld, stream);
    for (int q = 0; q < n_lower; q++) launch_axpy_dev(ld + q, 1.0, lower.p + (int64_t)q * V, yp, V, stream);
  }
  if (timing) hip_check(hipEventRecord(ev[3], stream), "event");
  hip_check(hipGetLastError(), "Heff kernels");
  if (timing) {
    hip_check(hipEventSynchronize(ev[3]), "event sync");
    static const int slot[3] = {1, 0, 2};  // explicit product, GEMM tiles, chained kernels
    for (int q = 0; q < 3; q++) {
      float a = 0.f;
      hip_check(hipEventElapsedTime(&a, ev[q], ev[q + 1]), "elapsed");
      t_ms[slot[q]] += a;
    }
    t_count++;
  }
}

void Ctx::diag(double* d) {
  if (site < 0) throw Error{ERR_STATE, "no prepared site"};
  std::vector<DevDiagBlock> blocks = hblocks;  // outputs relative to d
  for (int k = 0; k < S.nk(); k++) blocks[k].out = d + S.k2i[k];
  dblocks.upload(blocks, stream);
  launch_diag(dblocks.p, dterms.p, S.nk(), stream);
  launch_add_squares(lower.p, n_lower, d, V, stream);  // HeffDiagonal.cpp:621-642
  hip_check(hipGetLastError(), "diag kernel");
  hip_check(hipStreamSynchronize(stream), "diag");
}

void Ctx::mv_flush() {
  if (mv_used == 0) return;
  hip_check(hipEventSynchronize(mv_ev[mv_used - 1]), "event sync");
  for (int q = 0; q < mv_used; q += 2) {
    float a = 0.f;
    hip_check(hipEventElapsedTime(&a, mv_ev[q], mv_ev[q + 1]), "elapsed");
    mv_ms += a;
  }
  mv_used = 0;
}

void Ctx::symm(double* x, bool to_symm) {
  launch_symm_scale(x, d_k2i.p, to_symm ? d_sym.p : d_prog.p, S.nk(), stream);
}

// ------------------------------------------------------ operator updates -----
void Ctx::mps_alloc(int st) {
  if (st < 0 || st >= L) throw Error{ERR_ARG, "site out of range"};
  TLay& T = mps[st];
  T.build(sec, st, irreps[st]);
  if (T.buf >= 0 && bufs[T.buf]) { (void)hipFree(bufs[T.buf]); bufs[T.buf] = nullptr; buf_len[T.buf] = 0; }
  double* p = nullptr;
  if (T.size > 0) hip_check(hipMalloc((void**)&p, T.size * sizeof(double)), "hipMalloc MPS");
  if (T.buf < 0) { T.buf = (int32_t)bufs.size(); bufs.push_back(p); buf_len.push_back(T.size); }
  else { bufs[T.buf] = p; buf_len[T.buf] = T.size; }
  mps_buf[st] = p;
}

OpT* Ctx::find_tensor(int cnt, int kind, int k, int m) {
  if (cnt < 0 || cnt > L - 2) return nullptr;
  OpSet& s = sets[cnt];
  auto g1 = [&](std::vector<OpT*>& v) { return k < (int)v.size() ? v[k] : nullptr; };
  auto g2 = [&](std::vector<std::vector<OpT*>>& v) { return (k < (int)v.size() && m < (int)v[k].size()) ? v[k][m] : nullptr; };
  switch (kind) {
    case KIND_L: return g1(s.L);
    case KIND_Q: return g1(s.Q);
    case KIND_X: return s.X;
    case KIND_F0: return g2(s.F0);
    case KIND_F1: return g2(s.F1);
    case KIND_S0: return g2(s.S0);
    case KIND_S1: return g2(s.S1);
    case KIND_A: return g2(s.A);
    case KIND_B: return g2(s.B);
    case KIND_C: return g2(s.C);
    case KIND_D: return g2(s.D);
  }
  return nullptr;
}

void Ctx::run_terms(const std::vector<UTerm>& terms, const std::vector<std::pair<Ref, std::pair<int, int>>>& outs,
                    const double* combo_base) {
  std::vector<USeg> segs(1);
  segs[0].terms = terms;
  run_terms(segs, outs, combo_base);
}

// Executes update terms: every output block is overwritten by the sum of its terms.
// Three-operand terms first form W = alpha op(A) op(B); then every output block sums its
// two-operand products, W op(C) and the transposed axpys. Products whose blocks are
// small run in k_ops (2 x 2 register tiles, output blocks with many products split
// over several work items and summed by k_part_reduce), the rest in k_tiles. The lists
// are built on the host threads; positions come from prefix sums, so they do not
// depend on the thread count.
void Ctx::run_terms(const std::vector<USeg>& segs, const std::vector<std::pair<Ref, std::pair<int, int>>>& outs,
                    const double* combo_base) {
  const double t_begin = now_s();
  const int nthr = host_threads();
  const int64_t nseg = (int64_t)segs.size();
  const int64_t nq = (int64_t)outs.size();
  // output index of a (buffer, offset) reference: the outputs come grouped by buffer in
  // increasing offset, so a buffer's range of outputs and a binary search find it (the
  // last of equal offsets: an empty block shares its offset with the block after it)
  std::vector<int> qfirst(bufs.size() + 1, -1), qcount(bufs.size() + 1, 0);
  for (int64_t q = 0; q < nq; q++) {
    const Ref& r = outs[q].first;
    if (r.buf < 0 || r.buf >= (int)qfirst.size()) throw Error{ERR_STATE, "operator update output outside the buffers"};
    if (qfirst[r.buf] < 0) qfirst[r.buf] = (int)q;
    else if (qfirst[r.buf] + qcount[r.buf] != q || outs[q - 1].first.off > r.off)
      throw Error{ERR_STATE, "operator update outputs not grouped by buffer in offset order"};
    qcount[r.buf]++;
  }
  struct Where {
    const std::vector<std::pair<Ref, std::pair<int, int>>>& outs;
    const std::vector<int>& first;
    const std::vector<int>& count;
    int get(const Ref& r) const {
      if (r.buf < 0 || r.buf >= (int)first.size() || first[r.buf] < 0) return -1;
      int lo = first[r.buf], hi = lo + count[r.buf] - 1;
      if (outs[lo].first.off > r.off) return -1;
      while (lo < hi) {  // the last output with offset <= r.off
        const int mid = (lo + hi + 1) >> 1;
        if (outs[mid].first.off <= r.off) lo = mid; else hi = mid - 1;
      }
      return outs[lo].first.off == r.off ? lo : -1;
    }
  } where{outs, qfirst, qcount};
  const double t_map = now_s();
  std::vector<int64_t> s0(nseg + 1, 0);
  for (int64_t s = 0; s < nseg; s++) s0[s + 1] = s0[s] + (int64_t)segs[s].terms.size();
  const int64_t n = s0[nseg];
  if (n > INT32_MAX) throw Error{ERR_SIZE, "too many update terms"};
  u_qof.resize(n);
  u_seg.resize(n);
  u_woff.resize(n);
  u_order.resize(n);
  int* qof = u_qof.data();
  int* seg_of = u_seg.data();
  int64_t* woff = u_woff.data();
  int64_t* order = u_order.data();
  const int omax = ops_maxd;
  auto small_gemm = [omax](int m, int nn, int k, int ta, int lda, int tb, int ldb) {
    return m <= omax && nn <= omax && (int64_t)m * k <= OPS_P && (int64_t)k * nn <= OPS_P &&
           lda == (ta ? k : m) && ldb == (tb ? nn : k);
  };
  const bool ops_on = use_ops_kernel;
  // output block of every term; phase-1 counts per segment (k_ops items or GEMM tiles)
  struct P1 { int64_t ops = 0, tiles = 0, items = 0, w = 0; };
  std::vector<P1> p1(nseg + 1);
  parallel_for(nseg, nthr, [&](int64_t s, int) {
    P1 c;
    int last = -1;  // consecutive terms mostly write the same block
    for (size_t j = 0; j < segs[s].terms.size(); j++) {
      const UTerm& u = segs[s].terms[j];
      const int qq = (last >= 0 && outs[last].first.buf == u.out.buf && outs[last].first.off == u.out.off)
                         ? last
                         : where.get(u.out);
      last = qq;
      if (qq < 0) throw Error{ERR_STATE, "operator update term writes outside the new tensors"};
      qof[s0[s] + j] = qq;
      seg_of[s0[s] + j] = (int)s;
      if (u.kind != 1) continue;
      c.ops++;
      if (ops_on && small_gemm(u.m, u.k2, u.k1, u.ta, u.lda, u.tb, u.ldb)) c.items++;
      else c.tiles += (int64_t)((u.m + 63) / 64) * ((u.k2 + 63) / 64);
      c.w += ((int64_t)u.m * u.k2 + 7) / 8 * 8;
    }
    p1[s + 1] = c;
  });
  for (int64_t s = 0; s < nseg; s++) {
    p1[s + 1].ops += p1[s].ops; p1[s + 1].tiles += p1[s].tiles; p1[s + 1].items += p1[s].items; p1[s + 1].w += p1[s].w;
  }
  const int64_t n1 = p1[nseg].ops, nt1 = p1[nseg].tiles, ni1 = p1[nseg].items;
  u_W.alloc(std::max<int64_t>(p1[nseg].w, 1));
  double* Wp = u_W.p;
  // terms grouped by output block, stable in (segment, term) order: a counting sort
  // with one histogram per range of terms
  std::vector<int64_t> qstart(nq + 1, 0);
  {
    const int R = (int)std::max<int64_t>(1, std::min<int64_t>(nthr, (n + 65535) / 65536));
    if (u_hist.size() < (size_t)R * nq) u_hist.resize((size_t)R * nq);
    int64_t* hist = u_hist.data();
    auto range = [&](int r) { return std::make_pair(n * r / R, n * (r + 1) / R); };
    parallel_for(R, nthr, [&](int64_t r, int) {
      std::fill(hist + r * nq, hist + (r + 1) * nq, (int64_t)0);
      const auto rg = range((int)r);
      for (int64_t g = rg.first; g < rg.second; g++) hist[r * nq + qof[g]]++;
    });
    // exclusive prefix over (q, r) in q-major order: chunk sums, their prefix, then fill
    const int64_t NC = std::max<int64_t>(1, std::min<int64_t>(64, nq / 4096));
    std::vector<int64_t> csum(NC + 1, 0);
    parallel_for(NC, nthr, [&](int64_t c, int) {
      int64_t t = 0;
      for (int64_t q = nq * c / NC; q < nq * (c + 1) / NC; q++)
        for (int r = 0; r < R; r++) t += hist[r * nq + q];
      csum[c + 1] = t;
    });
    for (int64_t c = 0; c < NC; c++) csum[c + 1] += csum[c];
    parallel_for(NC, nthr, [&](int64_t c, int) {
      int64_t pos = csum[c];
      for (int64_t q = nq * c / NC; q < nq * (c + 1) / NC; q++) {
        qstart[q] = pos;
        for (int r = 0; r < R; r++) {
          const int64_t cnt = hist[r * nq + q];
          hist[r * nq + q] = pos;
          pos += cnt;
        }
      }
    });
    qstart[nq] = csum[NC];
    parallel_for(R, nthr, [&](int64_t r, int) {
      const auto rg = range((int)r);
      for (int64_t g = rg.first; g < rg.second; g++) order[hist[r * nq + qof[g]]++] = g;
    });
  }
  const double t_sort = now_s();
  // phase 2 per output block: small path or tiles; counts, then positions
  std::vector<char> qsmall(nq, 0);
  std::vector<int64_t> t2start(nq + 1, 0), i2start(nq + 1, 0), p2start(nq + 1, 0);
  parallel_for((nq + 1023) / 1024, nthr, [&](int64_t c, int) {
    for (int64_t q = c * 1024; q < std::min<int64_t>(nq, (c + 1) * 1024); q++) {
      const int m = outs[q].second.first, nn = outs[q].second.second;
      if (m <= 0 || nn <= 0) continue;
      bool sm = ops_on && m <= omax && nn <= omax;
      for (int64_t pos = qstart[q]; sm && pos < qstart[q + 1]; pos++) {
        const int64_t g = order[pos];
        const UTerm& u = segs[seg_of[g]].terms[g - s0[seg_of[g]]];
        if (u.kind == 1) sm = small_gemm(m, nn, u.k2, 0, m, u.tc, u.ldc);
        else if (u.kind == 0) sm = small_gemm(m, nn, u.k1, u.ta, u.lda, u.tb, u.ldb);
      }
      qsmall[q] = sm;
      const int64_t nops = qstart[q + 1] - qstart[q];
      if (sm) {
        const int64_t ni = std::max<int64_t>(1, (nops + OPS_ITEM - 1) / OPS_ITEM);
        i2start[q + 1] = ni;
        p2start[q + 1] = ni > 1 ? ni * (int64_t)m * nn : 0;
      } else {
        t2start[q + 1] = (int64_t)((m + 63) / 64) * ((nn + 63) / 64);
      }
    }
  });
  int64_t nred = 0;
  for (int64_t q = 0; q < nq; q++) {
    if (qsmall[q] && i2start[q + 1] > 1) nred++;
    t2start[q + 1] += t2start[q]; i2start[q + 1] += i2start[q]; p2start[q + 1] += p2start[q];
  }
  const int64_t nt2 = t2start[nq], ni2 = i2start[nq], npart = p2start[nq];
  if (nt1 > INT32_MAX || nt2 > INT32_MAX || ni1 + ni2 > INT32_MAX) throw Error{ERR_SIZE, "too many update tiles"};
  const double t_counts = now_s();
  ustage.reset();
  ustage.reserve(256 * 8 + (n1 + n) * sizeof(DevGemmOp) + (nt1 + nt2) * sizeof(DevTile) + (ni1 + ni2) * sizeof(DevOItem) +
                 nred * sizeof(DevPRed));
  DevGemmOp* o1 = ustage.take<DevGemmOp>(n1);
  DevGemmOp* o2 = ustage.take<DevGemmOp>(n);
  DevTile* t1 = ustage.take<DevTile>(nt1);
  DevTile* t2 = ustage.take<DevTile>(nt2);
  DevOItem* it1 = ustage.take<DevOItem>(ni1);
  DevOItem* it2 = ustage.take<DevOItem>(ni2);
  DevPRed* rd = ustage.take<DevPRed>(nred);
  u_part.alloc(std::max<int64_t>(npart, 1));
  std::atomic<int> Pneed{1};
  auto need = [&](int v) {
    int cur = Pneed.load();
    while (v > cur && !Pneed.compare_exchange_weak(cur, v)) {}
  };
  // phase 1: W = alpha op(A) op(B) for three-operand terms
  parallel_for(nseg, nthr, [&](int64_t s, int) {
    int64_t io = p1[s].ops, it = p1[s].tiles, ii = p1[s].items, iw = p1[s].w;
    const double* cb = combo_base ? combo_base + segs[s].combo_base : nullptr;
    int pl = 1;
    for (size_t j = 0; j < segs[s].terms.size(); j++) {
      const UTerm& u = segs[s].terms[j];
      if (u.kind != 1) continue;
      DevGemmOp op{};
      op.A = resolve_with(u.A, cb); op.ta = u.ta; op.lda = u.lda;
      op.B = resolve_with(u.B, cb); op.tb = u.tb; op.ldb = u.ldb;
      op.k = u.k1; op.alpha = u.alpha; op.kind = 0;
      woff[s0[s] + j] = iw;
      const int opi = (int)io;
      o1[io++] = op;
      if (ops_on && small_gemm(u.m, u.k2, u.k1, u.ta, u.lda, u.tb, u.ldb)) {
        it1[ii++] = DevOItem{Wp + iw, u.m, u.k2, opi, 1};
        pl = std::max({pl, u.m * u.k1, u.k1 * u.k2});
      } else {
        for (int m0 = 0; m0 < u.m; m0 += 64)
          for (int n0 = 0; n0 < u.k2; n0 += 64) {
            DevTile tl{};
            tl.C = Wp + iw;
            tl.ldc = u.m; tl.m0 = m0; tl.n0 = n0; tl.m = u.m; tl.n = u.k2; tl.op0 = opi; tl.nop = 1; tl.flags = TILE_NO_AXPY;
            t1[it++] = tl;
          }
      }
      iw += ((int64_t)u.m * u.k2 + 7) / 8 * 8;
    }
    need(pl);
  });
  // phase 2: one op per term at its grouped position
  parallel_for((n + 65535) / 65536, nthr, [&](int64_t c, int) {
    for (int64_t pos = c * 65536; pos < std::min<int64_t>(n, (c + 1) * 65536); pos++) {
      const int64_t g = order[pos];
      const int s = seg_of[g];
      const UTerm& u = segs[s].terms[g - s0[s]];
      const double* cb = combo_base ? combo_base + segs[s].combo_base : nullptr;
      DevGemmOp op{};
      if (u.kind == 1) {
        op.A = Wp + woff[g]; op.ta = 0; op.lda = u.m;
        op.B = resolve_with(u.C, cb); op.tb = u.tc; op.ldb = u.ldc;
        op.k = u.k2; op.alpha = 1.0; op.kind = 0;
      } else if (u.kind == 0) {
        op.A = resolve_with(u.A, cb); op.ta = u.ta; op.lda = u.lda;
        op.B = resolve_with(u.B, cb); op.tb = u.tb; op.ldb = u.ldb;
        op.k = u.k1; op.alpha = u.alpha; op.kind = 0;
      } else {
        op.A = resolve_with(u.A, cb); op.ta = u.ta; op.lda = u.lda;
        op.alpha = u.alpha; op.kind = 1;
      }
      o2[pos] = op;
    }
  });
  // phase 2 records per output block
  std::vector<int64_t> rstart(nq + 1, 0);
  for (int64_t q = 0; q < nq; q++) rstart[q + 1] = rstart[q] + ((qsmall[q] && i2start[q + 1] - i2start[q] > 1) ? 1 : 0);
  parallel_for((nq + 1023) / 1024, nthr, [&](int64_t c, int) {
    int pl = 1;
    for (int64_t q = c * 1024; q < std::min<int64_t>(nq, (c + 1) * 1024); q++) {
      const int m = outs[q].second.first, nn = outs[q].second.second;
      if (m <= 0 || nn <= 0) continue;
      double* C = const_cast<double*>(resolve_with(outs[q].first, combo_base));
      if (qsmall[q]) {
        const int64_t ni = i2start[q + 1] - i2start[q];
        const int64_t nops = qstart[q + 1] - qstart[q];
        for (int64_t i = 0; i < ni; i++) {
          DevOItem itm{};
          itm.m = m; itm.n = nn;
          itm.op0 = (int)(qstart[q] + i * OPS_ITEM);
          itm.nop = (int)std::min<int64_t>(OPS_ITEM, nops - i * OPS_ITEM);
          itm.out = ni > 1 ? u_part.p + p2start[q] + i * (int64_t)m * nn : C;
          it2[i2start[q] + i] = itm;
        }
        if (ni > 1) rd[rstart[q]] = DevPRed{C, p2start[q], m * nn, (int)ni};
        for (int64_t pos = qstart[q]; pos < qstart[q + 1]; pos++)
          if (o2[pos].kind == 0) pl = std::max({pl, m * o2[pos].k, o2[pos].k * nn});
      } else {
        int64_t it = t2start[q];
        bool axpy = false;
        for (int64_t pos = qstart[q]; pos < qstart[q + 1]; pos++) axpy |= o2[pos].kind == 1;
        for (int m0 = 0; m0 < m; m0 += 64)
          for (int n0 = 0; n0 < nn; n0 += 64) {
            DevTile tl{};
            tl.C = C; tl.ldc = m; tl.m0 = m0; tl.n0 = n0; tl.m = m; tl.n = nn;
            tl.op0 = (int)qstart[q]; tl.nop = (int)(qstart[q + 1] - qstart[q]);
            tl.flags = axpy ? 0 : TILE_NO_AXPY;
            t2[it++] = tl;
          }
      }
    }
    need(pl);
  });
  const double t_lists = now_s();
  u_d1.alloc(std::max<int64_t>(n1, 1));
  u_d2.alloc(std::max<int64_t>(n, 1));
  u_dt1.alloc(std::max<int64_t>(nt1, 1));
  u_dt2.alloc(std::max<int64_t>(nt2, 1));
  u_it1.alloc(std::max<int64_t>(ni1, 1));
  u_it2.alloc(std::max<int64_t>(ni2, 1));
  u_rd.alloc(std::max<int64_t>(nred, 1));
  if (n1) hip_check(hipMemcpyAsync(u_d1.p, o1, n1 * sizeof(DevGemmOp), hipMemcpyHostToDevice, stream), "ops");
  if (n) hip_check(hipMemcpyAsync(u_d2.p, o2, n * sizeof(DevGemmOp), hipMemcpyHostToDevice, stream), "ops");
  if (nt1) hip_check(hipMemcpyAsync(u_dt1.p, t1, nt1 * sizeof(DevTile), hipMemcpyHostToDevice, stream), "tiles");
  if (nt2) hip_check(hipMemcpyAsync(u_dt2.p, t2, nt2 * sizeof(DevTile), hipMemcpyHostToDevice, stream), "tiles");
  if (ni1) hip_check(hipMemcpyAsync(u_it1.p, it1, ni1 * sizeof(DevOItem), hipMemcpyHostToDevice, stream), "items");
  if (ni2) hip_check(hipMemcpyAsync(u_it2.p, it2, ni2 * sizeof(DevOItem), hipMemcpyHostToDevice, stream), "items");
  if (nred) hip_check(hipMemcpyAsync(u_rd.p, rd, nred * sizeof(DevPRed), hipMemcpyHostToDevice, stream), "reductions");
  const double t_up = now_s();
  const int P = std::min(OPS_P, (Pneed.load() + 31) / 32 * 32);
  launch_tiles(u_dt1.p, u_d1.p, (int)nt1, nullptr, nullptr, stream, tiles_kernel);
  launch_ops(u_it1.p, (int)ni1, u_d1.p, P, stream);
  launch_tiles(u_dt2.p, u_d2.p, (int)nt2, nullptr, nullptr, stream, tiles_kernel);
  launch_ops(u_it2.p, (int)ni2, u_d2.p, P, stream);
  launch_part_reduce(u_rd.p, (int)nred, u_part.p, stream);
  hip_check(hipGetLastError(), "update kernels");
  hip_check(hipStreamSynchronize(stream), "update");
  if (verbose)
    std::fprintf(stderr, "    terms %lld: lists %.3f s (map %.3f, sort %.3f, counts %.3f, fill %.3f), upload %.3f s, kernels %.3f s (%lld + %lld tiles, %lld + %lld small items)\n",
                 (long long)n, t_lists - t_begin, t_map - t_begin, t_sort - t_map, t_counts - t_sort, t_lists - t_counts,
                 t_up - t_lists, now_s() - t_up, (long long)nt1, (long long)nt2, (long long)ni1, (long long)ni2);
}

// DMRG::updateMovingRight/Left (DMRGoperators.cpp:243-907) for slot `index`.
void Ctx::update_operators(int index, bool mr) {
  if (!have_problem) throw Error{ERR_STATE, "set the problem first"};
  if (index < 0 || index > L - 2) throw Error{ERR_ARG, "slot out of range"};
  const int tsite = mr ? index : index + 1;
  if (mps[tsite].buf < 0 || !bufs[mps[tsite].buf]) throw Error{ERR_MISSING, "MPS tensor of site " + std::to_string(tsite) + " not set"};
  {
    TLay check;
    check.build(sec, tsite, irreps[tsite]);
    if (check.size != mps[tsite].size) throw Error{ERR_STATE, "MPS tensor layout does not match the current sectors"};
  }
  OpSet* prev = nullptr;
  if (mr && index > 0) { validate_set(index - 1, true); prev = &sets[index - 1]; }
  if (!mr && index < L - 2) { validate_set(index + 1, false); prev = &sets[index + 1]; }
  // the new slot laid out with the flags of DMRG::allocateTensors (already done by
  // alloc_slot when the next site was planned ahead with these dimensions)
  const bool laid_out = index < (int)prealloc.size() && prealloc[index] && sets[index].dir == (mr ? 1 : 2) &&
                        sets[index].gen == sec.gen;
  if (!laid_out) alloc_slot(index, mr);
  prealloc[index] = 0;
  OpSet& o = sets[index];
  std::vector<int> Qsite;
  const int nC = mr ? L - 1 - index : index + 1;
  for (int k = 0; k < nC; k++) Qsite.push_back(mr ? index + 1 + k : index - k);
  const double t0 = now_s();
  auto upp = std::make_unique<UpdatePlanner>();
  UpdatePlanner& up = *upp;
  up.sec = &sec;
  up.irr = irreps.data();
  up.mx = mx.data();
  up.L = L;
  if (mr) up.moving_right(index, o, prev, &mps[tsite], Qsite);
  else up.moving_left(index, o, prev, &mps[tsite], Qsite);
  const double t1 = now_s();
  last_update_flops = up.flops;
  last_update_terms = up.nterms(1) + up.nterms(2);
  // combinations first (they read the previous slot only)
  DBuf<double>& arena = u_arena;  // kept between updates (grown as needed)
  hip_check(hipStreamSynchronize(stream), "update arena");
  arena.alloc(std::max<int64_t>(up.combo_size, 1));
  if (!up.combos.empty()) {
    // records by prefix sums, filled on the host threads
    const int64_t ncb = (int64_t)up.combos.size();
    std::vector<DevCombo> dc(ncb);
    std::vector<int64_t> chunk_start(ncb), src0(ncb + 1, 0);
    int64_t nch = 0;
    for (int64_t i = 0; i < ncb; i++) {
      const Combo& c = up.combos[i];
      src0[i + 1] = src0[i] + (int64_t)c.src.size();
      chunk_start[i] = nch;
      nch += (c.size + 1023) / 1024;
    }
    if (src0[ncb] > INT32_MAX) throw Error{ERR_SIZE, "too many combination sources"};
    std::vector<DevSrc> ds(src0[ncb]);
    parallel_for((ncb + 4095) / 4096, host_threads(), [&](int64_t blk, int) {
      for (int64_t i = blk * 4096; i < std::min<int64_t>(ncb, (blk + 1) * 4096); i++) {
        const Combo& c = up.combos[i];
        DevCombo d;
        d.out = arena.p + c.out_off; d.size = c.size; d.src0 = (int)src0[i]; d.nsrc = (int)c.src.size();
        for (size_t q = 0; q < c.src.size(); q++) ds[src0[i] + q] = {resolve(c.src[q].first), c.src[q].second};
        dc[i] = d;
      }
    });
    cmb_dc.upload(dc, stream);
    cmb_ds.upload(ds, stream);
    cmb_cs.upload(chunk_start, stream);
    launch_combos(cmb_dc.p, cmb_ds.p, (int)dc.size(), cmb_cs.p, nch, stream);
    hip_check(hipGetLastError(), "combination kernel launch");
    hip_check(hipStreamSynchronize(stream), "combinations");
  }
  // outputs of each level
  std::vector<std::pair<Ref, std::pair<int, int>>> out1, out2;
  auto add_outs = [&](OpT* t, std::vector<std::pair<Ref, std::pair<int, int>>>& v) {
    if (!t) return;
    for (int ik = 0; ik < t->nk(); ik++) {
      const int du = sec.dim(t->boundary, t->Nup[ik], t->Sup[ik], t->Iup[ik]);
      const int dd = sec.dim(t->boundary, t->Nup[ik] + t->n_elec, t->Sdn[ik], t->Iup[ik] ^ t->n_irrep);
      v.push_back({t->blkref(ik), {du, dd}});
    }
  };
  for (OpT* t : o.L) add_outs(t, out1);
  for (auto& row : o.F0) for (OpT* t : row) add_outs(t, out1);
  for (auto& row : o.F1) for (OpT* t : row) add_outs(t, out1);
  for (auto& row : o.S0) for (OpT* t : row) add_outs(t, out1);
  for (auto& row : o.S1) for (OpT* t : row) add_outs(t, out1);
  for (OpT* t : o.Q) add_outs(t, out1);
  add_outs(o.X, out1);
  for (auto& row : o.A) for (OpT* t : row) add_outs(t, out2);
  for (auto& row : o.B) for (OpT* t : row) add_outs(t, out2);
  for (auto& row : o.C) for (OpT* t : row) add_outs(t, out2);
  for (auto& row : o.D) for (OpT* t : row) add_outs(t, out2);
  run_terms(up.segs[1], out1, arena.p);
  run_terms(up.segs[2], out2, arena.p);
  update_overlaps(index, mr);  // DMRGoperators.cpp:555-567, 888-900
  const size_t ncombos = up.combos.size();
  reap(std::move(upp));
  site = -1;
  if (verbose)
    std::fprintf(stderr, "update slot %d (%s): plan %.3f s (%lld terms, %zu combinations), device %.3f s\n", index,
                 mr ? "right" : "left", t1 - t0, (long long)last_update_terms, ncombos, now_s() - t1);
}

// ------------------------------------------------------------ Davidson -----
static void jacobi_eig(int n, const std::vector<double>& A, int lda, std::vector<double>& w, std::vector<double>& Vv) {
  std::vector<double> a((size_t)n * n), v((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) a[i + (size_t)n * j] = 0.5 * (A[i + (size_t)lda * j] + A[j + (size_t)lda * i]);
  for (int i = 0; i < n; i++) v[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const double x = a[i + (size_t)n * j] * a[i + (size_t)n * j];
        tot += x;
        if (i != j) off += x;
      }
    if (off == 0.0 || off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[p + (size_t)n * q];
        if (apq == 0.0) continue;
        // after a few sweeps an element below the rounding of both diagonal entries is
        // set to zero (the cyclic Jacobi threshold rule), so the sweeps terminate
        const double app = std::fabs(a[p + (size_t)n * p]), aqq = std::fabs(a[q + (size_t)n * q]);
        if (sweep > 3 && app + 100.0 * std::fabs(apq) == app && aqq + 100.0 * std::fabs(apq) == aqq) {
          a[p + (size_t)n * q] = 0.0;
          a[q + (size_t)n * p] = 0.0;
          continue;
        }
        const double th = (a[q + (size_t)n * q] - a[p + (size_t)n * p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          const double kp = a[k + (size_t)n * p], kq = a[k + (size_t)n * q];
          a[k + (size_t)n * p] = c * kp - s * kq;
          a[k + (size_t)n * q] = s * kp + c * kq;
        }
        for (int k = 0; k < n; k++) {
          const double pk = a[p + (size_t)n * k], qk = a[q + (size_t)n * k];
          a[p + (size_t)n * k] = c * pk - s * qk;
          a[q + (size_t)n * k] = s * pk + c * qk;
        }
        for (int k = 0; k < n; k++) {
          const double kp = v[k + (size_t)n * p], kq = v[k + (size_t)n * q];
          v[k + (size_t)n * p] = c * kp - s * kq;
          v[k + (size_t)n * q] = s * kp + c * kq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x + (size_t)n * x] < a[y + (size_t)n * y]; });
  w.assign(n, 0.0);
  Vv.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++) {
    w[j] = a[ord[j] + (size_t)n * ord[j]];
    for (int i = 0; i < n; i++) Vv[i + (size_t)n * j] = v[i + (size_t)n * ord[j]];
  }
}

// Per-element inputs of the learned preconditioner: log(dL) / 4 and log(dR) / 4 of the
// element's block at the prepared site.
void Ctx::build_eldims() {
  if (eldims_site == site && eldims_gen == sec.gen && eldims.p) return;
  std::vector<float> h((size_t)2 * V);
  for (int k = 0; k < S.nk(); k++) {
    const float a = (float)(std::log((double)S.dL[k]) / 4.0), b = (float)(std::log((double)S.dR[k]) / 4.0);
    for (int64_t i = S.k2i[k]; i < S.k2i[k + 1]; i++) { h[2 * i] = a; h[2 * i + 1] = b; }
  }
  eldims.upload(h, stream);
  eldims_site = site;
  eldims_gen = sec.gen;
}

// Davidson.cpp:31-536 in eigenvalue mode with the reference's constants
// (DAVIDSON_NUM_VEC = 32, DAVIDSON_NUM_VEC_KEEP = 3, DAVIDSON_PRECOND_CUTOFF = 1e-12;
// Options.h:70-74). Vectors live on the device; the projected problem on the host.
double Ctx::davidson_host(double* x, double rtol, int* n_matvec) {
  if (site < 0) throw Error{ERR_STATE, "no prepared site"};
  const int MAXV = 32, KEEP = 3;
  const double CUT = 1e-12;
  const int64_t n = V;
  DBuf<double> store;
  store.alloc((size_t)(2 * MAXV + 4 + KEEP) * n);
  std::vector<double*> vec(MAXV), hv(MAXV);
  for (int i = 0; i < MAXV; i++) { vec[i] = store.p + (size_t)i * n; hv[i] = store.p + (size_t)(MAXV + i) * n; }
  double* t = store.p + (size_t)(2 * MAXV) * n;
  double* u = t + n;
  double* work = u + n;
  double* dg = work + n;
  double* E = dg + n;
  DBuf<const double*> d_ptrs;
  DBuf<double> d_coef;
  d_ptrs.alloc(2 * MAXV + 8);
  d_coef.alloc(2 * MAXV + 8);
  struct HarvestRec {
    std::vector<int64_t> idx;
    std::vector<float> feats, p;
    std::vector<double> u;
  };
  std::vector<HarvestRec> recs;
  auto host_scalars = [&](int cnt) {
    std::vector<double> h(cnt);
    hip_check(hipMemcpyAsync(h.data(), scal.p, cnt * sizeof(double), hipMemcpyDeviceToHost, stream), "d2h");
    hip_check(hipStreamSynchronize(stream), "sync");
    return h;
  };
  auto dot = [&](const double* a, const double* b) {
    launch_dot(a, b, n, partial.p, scal.p, stream);
    return host_scalars(1)[0];
  };
  auto lincomb = [&](const std::vector<const double*>& vs, const std::vector<double>& cs, double* out) {
    copy_sync(stream, d_ptrs.p, vs.data(), vs.size() * sizeof(double*), hipMemcpyHostToDevice, "ptrs");
    copy_sync(stream, d_coef.p, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice, "coefs");
    launch_lincomb_vecs(d_ptrs.p, d_coef.p, (int)vs.size(), out, n, stream);
  };
  // mixed form: products with the hi floats of the explicit matrix first, hi + lo once
  // the residual has fallen below mixed_switch (restarting from the Ritz vector)
  struct LowPhase {
    bool& f;
    ~LowPhase() { f = false; }
  } low_phase{mv_low};
  mv_low = M_split;
  auto matvec = [&](int i) {
    (mv_low ? n_low_mults : n_full_mults)++;
    if (mv_timing) {
      if (mv_ev.empty()) {
        mv_ev.assign(512, nullptr);
        for (auto& e : mv_ev) hip_check(hipEventCreate(&e), "hipEventCreate");
      }
      if (mv_used + 2 > (int)mv_ev.size()) mv_flush();
      hip_check(hipEventRecord(mv_ev[mv_used], stream), "event");
      apply(vec[i], hv[i]);
      hip_check(hipEventRecord(mv_ev[mv_used + 1], stream), "event");
      mv_used += 2;
      mv_flops += F_H;
    } else {
      apply(vec[i], hv[i]);
    }
  };

  // 'A': guess in the symmetric convention and the diagonal
  hip_check(hipMemcpyAsync(t, x, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "guess");
  symm(t, true);
  diag(dg);
  std::vector<double> M((size_t)MAXV * MAXV, 0.0), eig, evec;
  int num_vec = 0, nmult = 0;
  auto add_new_vec = [&]() {  // Davidson.cpp:214-240 (modified Gram-Schmidt, scalars stay on the device)
    for (int c = 0; c < num_vec; c++) {
      launch_dot(t, vec[c], n, partial.p, scal.p, stream);
      launch_axpy_dev(scal.p, -1.0, vec[c], t, n, stream);
    }
    const double nrm = std::sqrt(dot(t, t));
    if (!(nrm > 0.0) || !std::isfinite(nrm)) throw Error{ERR_NUMERIC, "Davidson: new vector has zero or invalid norm"};
    launch_scale(1.0 / nrm, t, n, stream);
    std::swap(vec[num_vec], t);
  };
  // 'U': SafetyCheckGuess (Davidson.cpp:202-212)
  if (dot(t, t) == 0.0) {
    std::vector<double> h(n);
    uint64_t s = 0x9E3779B97F4A7C15ULL;
    for (int64_t i = 0; i < n; i++) {
      s ^= s << 13; s ^= s >> 7; s ^= s << 17;
      h[i] = (s >> 11) * (1.0 / 9007199254740992.0);
    }
    copy_sync(stream, t, h.data(), n * sizeof(double), hipMemcpyHostToDevice, "random guess");
  }
  add_new_vec();
  matvec(num_vec);
  nmult++;
  // learned preconditioner guard: the learned correction is used while every step it
  // makes lowers the residual norm; the first step that does not switches this solve
  // back to the plain diagonal correction of Davidson.cpp:322-350
  bool learned_ok = precond_mode == 1 && mlp.ready, last_learned = false;
  double prev_rnorm = HUGE_VAL, prev_low = -1.0;
  for (int iter = 0; iter < 100000; iter++) {
    // Davidson.cpp:242-320
    {
      std::vector<const double*> hp(hv.begin(), hv.begin() + num_vec + 1);
      copy_sync(stream, d_ptrs.p, hp.data(), hp.size() * sizeof(double*), hipMemcpyHostToDevice, "ptrs");
      launch_multi_dot(d_ptrs.p, num_vec + 1, vec[num_vec], n, partial.p, scal.p, stream);
      std::vector<double> h = host_scalars(num_vec + 1);
      for (int c = 0; c <= num_vec; c++) { M[c + MAXV * num_vec] = h[c]; M[num_vec + MAXV * c] = h[c]; }
    }
    num_vec++;
    jacobi_eig(num_vec, M, MAXV, eig, evec);
    {
      std::vector<const double*> vs;
      std::vector<double> cs;
      for (int c = 0; c < num_vec; c++) { vs.push_back(vec[c]); cs.push_back(evec[c]); }
      lincomb(vs, cs, u);
      for (int c = 0; c < num_vec; c++) { vs[c] = hv[c]; }
      vs.push_back(u);
      cs.push_back(-eig[0]);
      lincomb(vs, cs, t);
    }
    const double rnorm = std::sqrt(dot(t, t));
    if (!std::isfinite(rnorm)) throw Error{ERR_NUMERIC, "Davidson: residual is not finite"};
    if (last_learned && !(rnorm < prev_rnorm)) {
      learned_ok = false;
      n_learned_fallbacks++;
    }
    prev_rnorm = rnorm;
    // the low-precision products have a residual floor near 1e-6 |theta| (FP32 chained
    // products): the refinement also starts once the residual stops falling
    const bool low_stalled = mv_low && prev_low > 0.0 && rnorm > 0.7 * prev_low;
    prev_low = mv_low ? rnorm : -1.0;
    if (mv_low && (rnorm <= std::max(rtol, mixed_switch) || low_stalled)) {
      mv_low = false;
      prev_rnorm = HUGE_VAL;
      hip_check(hipMemcpyAsync(t, u, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "refinement start");
      num_vec = 0;
      add_new_vec();
      matvec(num_vec);
      nmult++;
      continue;
    }
    if (rnorm <= rtol) break;
    // Davidson.cpp:322-350, with the learned factor in precond_mode 1
    const bool learned = learned_ok;
    last_learned = learned;
    const bool harvest = harvest_per_iter > 0 && !learned;
    const float* mvals = nullptr;
    if (learned || harvest) {
      build_eldims();
      featX.alloc((size_t)n * MLP_IN);
      launch_features(dg, t, u, eldims.p, eig[0], std::sqrt((double)n) / rnorm, std::sqrt((double)n), n, featX.p, stream);
      if (learned) {
        mlp.forward(featX.p, n, stream);
        mvals = mlp.out.p;
      }
    }
    HarvestRec* rec = nullptr;
    if (harvest) {
      const int64_t S = std::min<int64_t>(harvest_per_iter, n), stride = n / S;
      const int64_t off = (int64_t)((hv_counter++ * 2654435761ULL) % (uint64_t)stride);
      recs.emplace_back();
      rec = &recs.back();
      rec->idx.resize(S);
      for (int64_t q = 0; q < S; q++) rec->idx[q] = off + q * stride;
      hv_idx.alloc(S);
      hv_rows.alloc((size_t)S * MLP_IN);
      copy_sync(stream, hv_idx.p, rec->idx.data(), S * sizeof(int64_t), hipMemcpyHostToDevice, "harvest indices");
      launch_gather_rows(featX.p, hv_idx.p, S, hv_rows.p, stream);
      rec->feats.resize((size_t)S * MLP_IN);
      copy_sync(stream, rec->feats.data(), hv_rows.p, rec->feats.size() * sizeof(float), hipMemcpyDeviceToHost, "harvest rows");
      std::vector<double> full(n);
      copy_sync(stream, full.data(), u, n * sizeof(double), hipMemcpyDeviceToHost, "harvest u");
      rec->u.resize(S);
      for (int64_t q = 0; q < S; q++) rec->u[q] = full[rec->idx[q]];
    }
    launch_precond(dg, u, eig[0], CUT, work, t, n, 0, 0.0, stream, mvals);
    launch_dot(work, t, n, partial.p, scal.p, stream);
    launch_dot(work, u, n, partial.p, scal.p + 1, stream);
    std::vector<double> wr = host_scalars(2);
    const double alpha = -wr[0] / wr[1];
    launch_precond(dg, u, eig[0], CUT, work, t, n, 1, alpha, stream, mvals);
    if (rec) {  // the plain preconditioned residual of the samples
      std::vector<double> full(n);
      copy_sync(stream, full.data(), t, n * sizeof(double), hipMemcpyDeviceToHost, "harvest t");
      rec->p.resize(rec->idx.size());
      for (size_t q = 0; q < rec->idx.size(); q++) rec->p[q] = (float)full[rec->idx[q]];
    }
    if (num_vec == MAXV) {  // Davidson.cpp:352-412, 510-536
      hip_check(hipMemcpyAsync(E, u, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "deflate");
      for (int c = 1; c < KEEP; c++) {
        std::vector<const double*> vs(vec.begin(), vec.end());
        std::vector<double> cs(MAXV);
        for (int v = 0; v < MAXV; v++) cs[v] = evec[v + (size_t)num_vec * c];
        lincomb(vs, cs, E + (size_t)c * n);
      }
      std::vector<double> Sm(KEEP * KEEP);
      for (int a = 0; a < KEEP; a++)
        for (int b = a; b < KEEP; b++) Sm[a + KEEP * b] = Sm[b + KEEP * a] = dot(E + (size_t)a * n, E + (size_t)b * n);
      std::vector<double> se, sv;
      jacobi_eig(KEEP, Sm, KEEP, se, sv);
      for (int c = 0; c < KEEP; c++) {
        const double f = std::pow(se[c], -0.25);
        for (int r = 0; r < KEEP; r++) sv[r + KEEP * c] *= f;
      }
      for (int v = 0; v < KEEP; v++) {
        std::vector<const double*> vs;
        std::vector<double> cs;
        for (int v2 = 0; v2 < KEEP; v2++) {
          double low = 0.0;
          for (int c = 0; c < KEEP; c++) low += sv[v2 + KEEP * c] * sv[v + KEEP * c];
          vs.push_back(E + (size_t)v2 * n);
          cs.push_back(low);
        }
        lincomb(vs, cs, vec[v]);
      }
      num_vec = 0;
      prev_rnorm = HUGE_VAL;
      for (int v = 0; v < KEEP; v++) { matvec(v); nmult++; num_vec++; }
      for (int a = 0; a < KEEP; a++)
        for (int b = a; b < KEEP; b++) M[a + MAXV * b] = M[b + MAXV * a] = dot(vec[a], hv[b]);
    }
    add_new_vec();
    matvec(num_vec);
    nmult++;
  }
  if (mv_timing) mv_flush();
  if (!recs.empty()) {  // ideal corrections e = x* - u_iteration of the harvested samples
    std::vector<double> xs(n);
    copy_sync(stream, xs.data(), u, n * sizeof(double), hipMemcpyDeviceToHost, "harvest solution");
    for (const HarvestRec& r : recs) {
      if (r.p.size() != r.idx.size()) continue;
      double ov = 0.0;
      for (size_t q = 0; q < r.idx.size(); q++) ov += xs[r.idx[q]] * r.u[q];
      const double sg = ov < 0.0 ? -1.0 : 1.0;
      std::vector<double> e(r.idx.size());
      double ss = 0.0, pe = 0.0, pp = 0.0;
      for (size_t q = 0; q < r.idx.size(); q++) {
        e[q] = sg * xs[r.idx[q]] - r.u[q];
        pe += (double)r.p[q] * e[q];
        pp += (double)r.p[q] * r.p[q];
      }
      // the correction's scale is free (the new vector is normalised): e is rescaled by
      // the best scale of the plain p, and each application weighs 1 / sum e^2
      const double sc = pp > 0.0 ? pe / pp : 0.0;
      if (!(sc > 0.0)) continue;
      for (size_t q = 0; q < r.idx.size(); q++) {
        e[q] /= sc;
        ss += e[q] * e[q];
      }
      if (!(ss > 1e-28)) continue;
      const float w = (float)(1.0 / ss);
      hv_feats.insert(hv_feats.end(), r.feats.begin(), r.feats.end());
      for (size_t q = 0; q < r.idx.size(); q++) {
        hv_p.push_back(r.p[q]);
        hv_e.push_back((float)e[q]);
        hv_w.push_back(w);
      }
    }
  }
  // 'C': copy back and convert to the program convention
  hip_check(hipMemcpyAsync(x, u, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "solution");
  symm(x, false);
  hip_check(hipStreamSynchronize(stream), "davidson");
  if (n_matvec) *n_matvec = nmult;
  return eig[0];
}

// Davidson eigensolver with its scalars on the device (Heff::SolveDAVIDSON_main,
// Heff.cpp:330-386, with Davidson(V, 32, 3, rtol, 1e-12), Davidson.cpp:31-536). One
// iteration is a fixed launch sequence: the Olsen correction (Davidson.cpp:322-350) with
// the shift and its coefficient read on the device, orthogonalisation against the basis
// (classical Gram-Schmidt applied twice) and normalisation, the Heff product, the new
// column of V^T H V, the Rayleigh-Ritz step of the small matrix in one workgroup (cyclic
// Jacobi, FP64), the Ritz vector and residual, and a copy of the scalars to page-locked
// host memory. The sequence reads the basis size from the device record, so it is
// captured once per solve as a HIP graph and replayed every iteration: one launch and
// one synchronisation per iteration. Deflation at 32 vectors (Davidson.cpp:352-412,
// 510-536), the mixed form's switch to refined products and convergence are decided on
// the host from the copied scalars.
namespace {
// bytes of DevDav copied to the host after each iteration (the scalars before M)
constexpr size_t kDavHead = offsetof(DevDav, M);
}  // namespace

double Ctx::davidson(double* x, double rtol, int* n_matvec) {
  if (site < 0) throw Error{ERR_STATE, "no prepared site"};
  if (harvest_per_iter > 0 || timing || !dav_device) return davidson_host(x, rtol, n_matvec);
  const int MAXV = DAV_MAXV, KEEP = 3;
  const double CUT = 1e-12;
  const int64_t n = V;
  DBuf<double> store;
  store.alloc((size_t)(2 * MAXV + 6 + KEEP) * n);
  std::vector<double*> vec(MAXV), hv(MAXV);
  for (int i = 0; i < MAXV; i++) {
    vec[i] = store.p + (size_t)i * n;
    hv[i] = store.p + (size_t)(MAXV + i) * n;
  }
  double* t = store.p + (size_t)(2 * MAXV) * n;  // residual, then correction
  double* u = t + n;                                // Ritz vector
  double* work = u + n;
  double* dg = work + n;                            // diagonal of Heff
  double* xc = dg + n;                              // new basis vector
  double* yc = xc + n;                              // its product
  double* E = yc + n;                               // deflation
  dav_vec.upload(vec, stream);
  dav_hv.upload(hv, stream);
  const int parts = dot_parts(n);
  partial.alloc((size_t)std::max(parts, 1));
  dav_part.alloc((size_t)MAXV * parts);
  dav_state.alloc(1);
  if (!dav_host) hip_check(hipHostMalloc((void**)&dav_host, sizeof(DevDav), hipHostMallocDefault), "hipHostMalloc");
  DevDav* st = dav_state.p;
  std::memset(dav_host, 0, sizeof(DevDav));
  hip_check(hipMemcpyAsync(st, dav_host, sizeof(DevDav), hipMemcpyHostToDevice, stream), "Davidson state");
  hip_check(hipStreamSynchronize(stream), "Davidson state");
  double* const* dvec = dav_vec.p;
  double* const* dhv = dav_hv.p;

  // 'A': guess in the symmetric convention and the diagonal
  hip_check(hipMemcpyAsync(t, x, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "guess");
  symm(t, true);
  diag(dg);
  {  // 'U': SafetyCheckGuess (Davidson.cpp:202-212)
    launch_dot(t, t, n, partial.p, scal.p, stream);
    double q = 0.0;
    copy_sync(stream, &q, scal.p, sizeof(double), hipMemcpyDeviceToHost, "guess norm");
    if (q == 0.0) {
      std::vector<double> h(n);
      uint64_t s = 0x9E3779B97F4A7C15ULL;
      for (int64_t i = 0; i < n; i++) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        h[i] = (s >> 11) * (1.0 / 9007199254740992.0);
      }
      copy_sync(stream, t, h.data(), n * sizeof(double), hipMemcpyHostToDevice, "random guess");
    }
  }
  struct LowPhase {
    bool& f;
    ~LowPhase() { f = false; }
  } low_phase{mv_low};
  mv_low = M_split;
  int nmult = 0, napply = 0;
  // correction from the residual in t: t = -(r + alpha u) / K with K = (diag - theta),
  // or K / (1 + m) with the learned factor m, and alpha making t orthogonal to u in the
  // Olsen sense (Davidson.cpp:322-350)
  auto enqueue_P = [&](const float* mvals) {
    launch_precond_dev(dg, u, &st->theta, CUT, work, t, n, 0, nullptr, stream, mvals);
    launch_dot(work, t, n, partial.p, &st->wr[0], stream);
    launch_dot(work, u, n, partial.p, &st->wr[1], stream);
    launch_dav_alpha(st, stream);
    launch_precond_dev(dg, u, &st->theta, CUT, work, t, n, 1, &st->alpha, stream, mvals);
  };
  // new vector from t, its product, and the Ritz step
  auto enqueue_Q = [&]() {
    for (int pass = 0; pass < 2; pass++) {
      launch_dav_dots(st, 0, dvec, t, n, dav_part.p, parts, 1, stream);
      launch_dav_project_out(st, dvec, t, n, stream);
    }
    launch_dot(t, t, n, partial.p, &st->nrm2, stream);
    launch_dav_normalize(st, t, xc, n, stream);
    launch_dav_clock(st, 0, stream);
    apply(xc, yc);
    launch_dav_clock(st, 1, stream);
    launch_dav_store(st, dvec, dhv, xc, yc, n, stream);
    launch_dav_dots(st, 1, dhv, xc, n, dav_part.p, parts, 0, stream);
    launch_dav_eig(st, stream);
    launch_dav_ritz(st, dvec, dhv, u, t, n, partial.p, parts, stream);
    hip_check(hipMemcpyAsync(dav_host, st, kDavHead, hipMemcpyDeviceToHost, stream), "Davidson scalars");
  };
  // iteration graphs, captured on first use: [0] plain correction + new vector, [1] new
  // vector only (after a correction made outside the graph)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  auto drop_graphs = [&]() {
    for (int g = 0; g < 2; g++) {
      if (gexec[g]) (void)hipGraphExecDestroy(gexec[g]);
      if (graph[g]) (void)hipGraphDestroy(graph[g]);
      gexec[g] = nullptr;
      graph[g] = nullptr;
    }
  };
  struct GraphGuard {
    std::function<void()> f;
    ~GraphGuard() { f(); }
  } graph_guard{drop_graphs};
  auto run = [&](int g) {  // g = 0: P + Q, 1: Q
    if (!dav_graph) {
      if (g == 0) enqueue_P(nullptr);
      enqueue_Q();
      return;
    }
    if (!gexec[g]) {
      hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "graph capture");
      try {
        if (g == 0) enqueue_P(nullptr);
        enqueue_Q();
      } catch (...) {
        hipGraph_t gg = nullptr;
        (void)hipStreamEndCapture(stream, &gg);
        if (gg) (void)hipGraphDestroy(gg);
        throw;
      }
      hip_check(hipStreamEndCapture(stream, &graph[g]), "graph capture end");
      hip_check(hipGraphInstantiate(&gexec[g], graph[g], nullptr, nullptr, 0), "graph instantiate");
      dav_graphs++;
    }
    hip_check(hipGraphLaunch(gexec[g], stream), "graph launch");
    dav_graph_launches++;
  };
  auto count_apply = [&]() {
    (mv_low ? n_low_mults : n_full_mults)++;
    napply++;
    nmult++;
  };
  auto sync_read = [&]() {
    hip_check(hipStreamSynchronize(stream), "Davidson iteration");
    if (dav_host->bad) throw Error{ERR_NUMERIC, "Davidson: new vector has zero or invalid norm"};
  };
  auto dot = [&](const double* a, const double* b) {
    launch_dot(a, b, n, partial.p, scal.p, stream);
    double q = 0.0;
    copy_sync(stream, &q, scal.p, sizeof(double), hipMemcpyDeviceToHost, "dot");
    return q;
  };

  enqueue_Q();
  count_apply();
  int host_nv = 1;
  sync_read();
  // learned preconditioner guard: the learned correction is used while every step it
  // makes lowers the residual norm; the first step that does not switches this solve
  // back to the plain diagonal correction of Davidson.cpp:322-350
  bool learned_ok = precond_mode == 1 && mlp.ready, last_learned = false;
  double prev_rnorm = HUGE_VAL, prev_low = -1.0;
  for (int iter = 0; iter < 100000; iter++) {
    const double rnorm = std::sqrt(dav_host->rn2);
    const double theta = dav_host->theta;
    if (!std::isfinite(rnorm)) throw Error{ERR_NUMERIC, "Davidson: residual is not finite"};
    if (last_learned && !(rnorm < prev_rnorm)) {
      learned_ok = false;
      n_learned_fallbacks++;
    }
    prev_rnorm = rnorm;
    // the low-precision products have a residual floor near 1e-6 |theta| (FP32 chained
    // products): the refinement also starts once the residual stops falling
    const bool low_stalled = mv_low && prev_low > 0.0 && rnorm > 0.7 * prev_low;
    prev_low = mv_low ? rnorm : -1.0;
    if (mv_low && (rnorm <= std::max(rtol, mixed_switch) || low_stalled)) {  // refined products from the Ritz vector on
      mv_low = false;
      prev_rnorm = HUGE_VAL;
      last_learned = false;
      drop_graphs();
      hip_check(hipMemcpyAsync(t, u, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "refinement start");
      DevDav h;
      copy_sync(stream, &h, st, kDavHead, hipMemcpyDeviceToHost, "Davidson state");
      h.nv = 0;
      copy_sync(stream, st, &h, kDavHead, hipMemcpyHostToDevice, "basis reset");
      enqueue_Q();
      count_apply();
      host_nv = 1;
      sync_read();
      continue;
    }
    if (rnorm <= rtol) break;
    const bool learned = learned_ok;
    last_learned = learned;
    const float* mvals = nullptr;
    if (learned) {
      build_eldims();
      featX.alloc((size_t)n * MLP_IN);
      launch_features(dg, t, u, eldims.p, theta, std::sqrt((double)n) / rnorm, std::sqrt((double)n), n, featX.p, stream);
      mlp.forward(featX.p, n, stream);
      mvals = mlp.out.p;
    }
    if (host_nv == MAXV) {  // Davidson.cpp:352-412, 510-536
      enqueue_P(mvals);
      DevDav h;
      copy_sync(stream, &h, st, sizeof(DevDav), hipMemcpyDeviceToHost, "Davidson state");
      hip_check(hipMemcpyAsync(E, u, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "deflate");
      auto lincomb = [&](const std::vector<const double*>& vs, const std::vector<double>& cs, double* out) {
        dav_ptrs.upload(vs, stream);
        dav_coef.upload(cs, stream);
        launch_lincomb_vecs(dav_ptrs.p, dav_coef.p, (int)vs.size(), out, n, stream);
      };
      for (int c = 1; c < KEEP; c++) {
        std::vector<const double*> vs(vec.begin(), vec.end());
        std::vector<double> cs(MAXV);
        for (int v = 0; v < MAXV; v++) cs[v] = h.evec[v + (size_t)MAXV * c];
        lincomb(vs, cs, E + (size_t)c * n);
      }
      std::vector<double> Sm(KEEP * KEEP);
      for (int a = 0; a < KEEP; a++)
        for (int b = a; b < KEEP; b++) Sm[a + KEEP * b] = Sm[b + KEEP * a] = dot(E + (size_t)a * n, E + (size_t)b * n);
      std::vector<double> se, sv;
      jacobi_eig(KEEP, Sm, KEEP, se, sv);
      for (int c = 0; c < KEEP; c++) {
        const double f = std::pow(se[c], -0.25);
        for (int r = 0; r < KEEP; r++) sv[r + KEEP * c] *= f;
      }
      for (int v = 0; v < KEEP; v++) {
        std::vector<const double*> vs;
        std::vector<double> cs;
        for (int v2 = 0; v2 < KEEP; v2++) {
          double low = 0.0;
          for (int c = 0; c < KEEP; c++) low += sv[v2 + KEEP * c] * sv[v + KEEP * c];
          vs.push_back(E + (size_t)v2 * n);
          cs.push_back(low);
        }
        lincomb(vs, cs, vec[v]);
      }
      for (int v = 0; v < KEEP; v++) {
        launch_dav_clock(st, 0, stream);
        apply(vec[v], hv[v]);
        launch_dav_clock(st, 1, stream);
        count_apply();
      }
      copy_sync(stream, &h, st, sizeof(DevDav), hipMemcpyDeviceToHost, "Davidson state");
      for (int a = 0; a < KEEP; a++)
        for (int b = a; b < KEEP; b++) h.M[a + MAXV * b] = h.M[b + MAXV * a] = dot(vec[a], hv[b]);
      h.nv = KEEP;
      copy_sync(stream, st, &h, sizeof(DevDav), hipMemcpyHostToDevice, "Davidson state");
      dav_deflations++;
      prev_rnorm = HUGE_VAL;
      run(1);
      host_nv = KEEP;
    } else if (learned) {
      enqueue_P(mvals);
      run(1);
    } else {
      run(0);
    }
    count_apply();
    host_nv++;
    sync_read();
  }
  const double theta = dav_host->theta;
  if (mv_timing) {
    mv_flops += F_H * napply;
    mv_ms += (double)dav_host->heff_ticks / wall_khz;
  }
  // 'C': copy back and convert to the program convention
  hip_check(hipMemcpyAsync(x, u, n * sizeof(double), hipMemcpyDeviceToDevice, stream), "solution");
  symm(x, false);
  hip_check(hipStreamSynchronize(stream), "davidson");
  if (n_matvec) *n_matvec = nmult;
  return theta;
}

}  // namespace chm

// ================================================================= C ABI =====
using chm::Ctx;
using chm::Error;

template <class F>
static int guard(chm_ctx* ctx, F&& f) {
  if (!ctx) return CHM_ERR_ARG;
  try {
    f(ctx->c);
    ctx->c.err.clear();
    return CHM_OK;
  } catch (const Error& e) {
    ctx->c.err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return CHM_ERR_STATE;
  }
}

extern "C" {

int chm_ctx_create(int device, chm_ctx** out) {
  if (!out) return CHM_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return CHM_ERR_HIP;
  auto* ctx = new (std::nothrow) chm_ctx();
  if (!ctx) return CHM_ERR_STATE;
  ctx->c.device = device;
  if (const char* e = std::getenv("CHM_EXPLICIT_MAX")) ctx->c.explicit_max = std::atoll(e);
  if (const char* e = std::getenv("CHM_EXPLICIT_BUDGET_GB")) ctx->c.explicit_budget_gb = std::atof(e);
  if (const char* e = std::getenv("CHM_ITEM_ELEMS")) ctx->c.item_elems = std::max(256LL, std::atoll(e));
  if (const char* e = std::getenv("CHM_VERBOSE")) ctx->c.verbose = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_SVD_JACOBI_MAXN")) ctx->c.svd_jacobi_maxn = std::atoi(e);
  if (const char* e = std::getenv("CHM_CHAIN_KERNEL")) ctx->c.use_chain_kernel = std::atoi(e) != 0 ? 1 : 0;
  if (const char* e = std::getenv("CHM_CHAIN_MFMA")) ctx->c.use_chain_mfma = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_TILES_MFMA")) ctx->c.tiles_kernel = std::max(0, std::min(2, std::atoi(e)));
  if (const char* e = std::getenv("CHM_TILE_SEGMENTS")) ctx->c.tile_segments = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("CHM_TILE_SEG_MIN")) ctx->c.tile_seg_min = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("CHM_WAVE_TILES")) ctx->c.wave_tiles = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_CHAIN_ITEM_TERMS")) ctx->c.chain_item_terms = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("CHM_CHAIN_GROUP_TERMS")) ctx->c.chain_group_terms = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("CHM_SVD_GESVDJ")) ctx->c.svd_large_gesvdj = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_SVD_MAX_SWEEPS")) ctx->c.svd_max_sweeps = std::max(0, std::atoi(e));
  if (const char* e = std::getenv("CHM_TILE_W_CAP")) ctx->c.tile_w_cap = std::max(1LL, std::atoll(e));
  if (const char* e = std::getenv("CHM_OPS_KERNEL")) ctx->c.use_ops_kernel = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_OPS_MAXD")) ctx->c.ops_maxd = std::min(chm::OPS_MAXD, std::max(1, std::atoi(e)));
  ctx->c.env_pinned[0] = std::getenv("CHM_EXPLICIT_MAX") != nullptr;
  ctx->c.env_pinned[1] = std::getenv("CHM_CHAIN_ITEM_TERMS") != nullptr;
  ctx->c.env_pinned[2] = std::getenv("CHM_CHAIN_GROUP_TERMS") != nullptr;
  ctx->c.env_pinned[3] = std::getenv("CHM_CHAIN_MFMA") != nullptr;
  ctx->c.env_pinned[4] = std::getenv("CHM_FCHAIN_MIN") != nullptr;
  ctx->c.def_explicit_max = ctx->c.explicit_max;
  ctx->c.def_item_terms = ctx->c.chain_item_terms;
  ctx->c.def_group_terms = ctx->c.chain_group_terms;
  ctx->c.def_chain_mfma = ctx->c.use_chain_mfma;
  if (const char* e = std::getenv("CHM_FCHAIN")) ctx->c.use_fchain = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_FCHAIN_MIN")) ctx->c.fchain_min = std::atoi(e);
  if (const char* e = std::getenv("CHM_FCHAIN_FP32")) ctx->c.fchain_fp32 = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_FCHAIN_MAXD")) ctx->c.fchain_maxd = std::atoi(e);
  ctx->c.def_fchain_min = ctx->c.fchain_min;
  if (const char* e = std::getenv("CHM_DAVIDSON_DEVICE")) ctx->c.dav_device = std::atoi(e) != 0;
  if (const char* e = std::getenv("CHM_DAVIDSON_GRAPH")) ctx->c.dav_graph = std::atoi(e) != 0;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return CHM_ERR_HIP;
  }
  ctx->c.chain_large_ok = chm::chain_kernel_setup();
  ctx->c.fchain_ok = chm::fchain_kernel_setup();
  (void)hipGetLastError();
  if (!ctx->c.fchain_ok) ctx->c.use_fchain = false;
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0) ctx->c.wall_khz = khz;
  }
  (void)hipGetLastError();  // a refused attribute leaves no error behind
  for (auto& e : ctx->c.ev)
    if (hipEventCreate(&e) != hipSuccess) { delete ctx; return CHM_ERR_HIP; }
  *out = ctx;
  return CHM_OK;
}

int chm_ctx_destroy(chm_ctx* ctx) {
  if (!ctx) return CHM_ERR_ARG;
  (void)hipSetDevice(ctx->c.device);
  delete ctx;
  return CHM_OK;
}

const char* chm_last_error(const chm_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int chm_set_problem(chm_ctx* ctx, int L, int group, const int* orb_irreps, int N, int two_s, int irrep,
                    const double* mx, double econst) {
  return guard(ctx, [&](Ctx& c) {
    if (!orb_irreps || !mx) throw Error{CHM_ERR_ARG, "null pointer"};
    (void)hipSetDevice(c.device);
    c.set_problem(L, group, orb_irreps, N, two_s, irrep, mx, econst);
  });
}

int chm_set_sectors(chm_ctx* ctx, int boundary, int n, const int* N, const int* two_s, const int* irrep,
                    const int* dim) {
  return guard(ctx, [&](Ctx& c) {
    if (n > 0 && (!N || !two_s || !irrep || !dim)) throw Error{CHM_ERR_ARG, "null pointer"};
    c.set_sectors(boundary, n, N, two_s, irrep, dim);
  });
}

int chm_set_sectors_auto(chm_ctx* ctx, int D) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem) throw Error{CHM_ERR_STATE, "set the problem first"};
    if (D < 1) throw Error{CHM_ERR_ARG, "D must be positive"};
    chm::auto_sectors(c, D);
    c.site = -1;
  });
}

int chm_get_sectors(const chm_ctx* ctx, int boundary, int* n, int* N, int* two_s, int* irrep, int* dim) {
  if (!ctx || !n) return CHM_ERR_ARG;
  const Ctx& c = ctx->c;
  if (!c.have_problem || boundary < 0 || boundary > c.L) return CHM_ERR_ARG;
  const auto& v = c.sec.list[boundary];
  if (N && two_s && irrep && dim) {
    if (*n < (int)v.size()) return CHM_ERR_SIZE;
    for (size_t q = 0; q < v.size(); q++) { N[q] = v[q].N; two_s[q] = v[q].S; irrep[q] = v[q].I; dim[q] = v[q].d; }
  }
  *n = (int)v.size();
  return CHM_OK;
}

int chm_tensor_upload(chm_ctx* ctx, int cnt, int moving_right, int kind, int k, int m, int two_j, int n_elec,
                      int n_irrep, const double* data, int64_t n) {
  return guard(ctx, [&](Ctx& c) {
    if (n > 0 && !data) throw Error{CHM_ERR_ARG, "null pointer"};
    (void)hipSetDevice(c.device);
    chm::OpT* t = c.add_tensor(cnt, moving_right != 0, kind, k, m, two_j, n_elec, n_irrep, n);
    if (t->size > 0)
      chm::copy_sync(c.stream, c.bufs[t->buf], data, t->size * sizeof(double), hipMemcpyHostToDevice, "tensor upload");
  });
}

int chm_tensor_fill_synthetic(chm_ctx* ctx, int cnt, int moving_right, uint64_t seed) {
  return guard(ctx, [&](Ctx& c) {
    (void)hipSetDevice(c.device);
    if (cnt < 0 || cnt > c.L - 2) throw Error{CHM_ERR_ARG, "slot out of range"};
    c.clear_set(cnt);
    uint64_t s = seed * 1000003ULL + 17;
    for (const chm::OpShape& sh : chm::opset_shapes(c, cnt, moving_right != 0)) {
      chm::OpT* t = c.add_tensor(cnt, moving_right != 0, sh.kind, sh.k, sh.m, sh.two_j, sh.n_elec, sh.n_irrep, -1);
      const uint64_t ts = s++;  // one seed per tensor in allocation order, empty ones included
      if (t->size > 0) chm::launch_fill_hash(c.bufs[t->buf], t->size, ts, 1.0 / std::sqrt((double)t->size), c.stream);
    }
    chm::hip_check(hipGetLastError(), "fill kernel");
    chm::hip_check(hipStreamSynchronize(c.stream), "fill");
  });
}

int chm_tensor_clear(chm_ctx* ctx, int cnt) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem || cnt < 0 || cnt > c.L - 2) throw Error{CHM_ERR_ARG, "slot out of range"};
    c.clear_set(cnt);
  });
}

int chm_mps_upload(chm_ctx* ctx, int site, const double* data, int64_t n) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem) throw Error{CHM_ERR_STATE, "set the problem first"};
    if (n > 0 && !data) throw Error{CHM_ERR_ARG, "null pointer"};
    (void)hipSetDevice(c.device);
    c.mps_alloc(site);
    if (c.mps[site].size != n)
      throw Error{CHM_ERR_SIZE, "MPS size " + std::to_string(n) + " != layout size " + std::to_string(c.mps[site].size)};
    if (n > 0) chm::copy_sync(c.stream, c.mps_buf[site], data, n * sizeof(double), hipMemcpyHostToDevice, "MPS upload");
  });
}

int chm_mps_layout_size(const chm_ctx* ctx, int site, int64_t* n) {
  if (!ctx || !n) return CHM_ERR_ARG;
  const Ctx& c = ctx->c;
  if (!c.have_problem || site < 0 || site >= c.L) return CHM_ERR_ARG;
  chm::TLay T;
  T.build(c.sec, site, c.irreps[site]);
  *n = T.size;
  return CHM_OK;
}

int chm_mps_download(chm_ctx* ctx, int site, double* data, int64_t n) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem || site < 0 || site >= c.L || c.mps[site].buf < 0) throw Error{CHM_ERR_STATE, "no MPS tensor at this site"};
    if (n != c.mps[site].size) throw Error{CHM_ERR_SIZE, "size mismatch"};
    if (n > 0) chm::copy_sync(c.stream, data, c.mps_buf[site], n * sizeof(double), hipMemcpyDeviceToHost, "MPS download");
  });
}

int chm_tensor_download(chm_ctx* ctx, int cnt, int kind, int k, int m, double* data, int64_t n) {
  return guard(ctx, [&](Ctx& c) {
    chm::OpT* t = c.find_tensor(cnt, kind, k, m);
    if (!t) throw Error{CHM_ERR_MISSING, "no such tensor"};
    if (n != t->size) throw Error{CHM_ERR_SIZE, "size " + std::to_string(n) + " != " + std::to_string(t->size)};
    if (n > 0) chm::copy_sync(c.stream, data, c.bufs[t->buf], n * sizeof(double), hipMemcpyDeviceToHost, "tensor download");
  });
}

int chm_update_operators(chm_ctx* ctx, int index, int moving_right) {
  return guard(ctx, [&](Ctx& c) {
    (void)hipSetDevice(c.device);
    c.update_operators(index, moving_right != 0);
  });
}

int chm_dmrg_init(chm_ctx* ctx, int D, uint64_t seed) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem) throw Error{CHM_ERR_STATE, "set the problem first"};
    if (D < 1) throw Error{CHM_ERR_ARG, "D must be positive"};
    (void)hipSetDevice(c.device);
    const double t0 = chm::now_s();
    c.drop_lowers();  // a new DMRG object has no excitations
    c.site_log.clear();
    chm::auto_sectors(c, D);
    c.mps_random(seed);
    const double t1 = chm::now_s();
    c.presolve();
    if (c.verbose) std::fprintf(stderr, "dmrg init: sectors and MPS %.3f s, operators %.3f s\n", t1 - t0, chm::now_s() - t1);
    c.n_matvec_total = 0;
    for (double& t : c.tm) t = 0.0;
    c.mv_flops = c.mv_ms = 0.0;
  });
}

int chm_dmrg_new_excitation(chm_ctx* ctx, double eshift, int D, uint64_t seed) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem) throw Error{CHM_ERR_STATE, "set the problem first"};
    if (D < 1) throw Error{CHM_ERR_ARG, "D must be positive"};
    (void)hipSetDevice(c.device);
    c.new_excitation(eshift, D, seed);
    c.n_matvec_total = 0;
    for (double& t : c.tm) t = 0.0;
    c.mv_flops = c.mv_ms = 0.0;
  });
}

int chm_dmrg_2rdm(chm_ctx* ctx, double* A, double* B, double* energy) {
  return guard(ctx, [&](Ctx& c) {
    (void)hipSetDevice(c.device);
    const double e = c.two_rdm(A, B);
    if (energy) *energy = e;
  });
}

int chm_davidson_stats(const chm_ctx* ctx, int64_t* out) {
  if (!ctx || !out) return CHM_ERR_ARG;
  out[0] = ctx->c.dav_deflations;
  out[1] = ctx->c.dav_graphs;
  out[2] = ctx->c.dav_graph_launches;
  return CHM_OK;
}

int chm_dmrg_nstates(const chm_ctx* ctx, int* n) {
  if (!ctx || !n) return CHM_ERR_ARG;
  *n = (int)ctx->c.lowers.size() + 1;
  return CHM_OK;
}

int chm_get_problem(const chm_ctx* ctx, int* L, int* N, int* two_s, int* irrep, int* nirrep) {
  if (!ctx || !L || !N || !two_s || !irrep || !nirrep) return CHM_ERR_ARG;
  if (!ctx->c.have_problem) return CHM_ERR_STATE;
  *L = ctx->c.L; *N = ctx->c.N; *two_s = ctx->c.TwoS; *irrep = ctx->c.Irrep; *nirrep = ctx->c.nirrep;
  return CHM_OK;
}

int chm_dmrg_presolve(chm_ctx* ctx) {
  return guard(ctx, [&](Ctx& c) {
    if (!c.have_problem) throw Error{CHM_ERR_STATE, "set the problem first"};
    if ((int)c.fci.size() != c.L + 1) throw Error{CHM_ERR_STATE, "call chm_set_sectors_auto first (FCI dimensions)"};
    for (int s = 0; s < c.L; s++)
      if (c.mps[s].buf < 0 || !c.bufs[c.mps[s].buf]) throw Error{CHM_ERR_MISSING, "MPS tensor of site " + std::to_string(s) + " not set"};
    (void)hipSetDevice(c.device);
    c.presolve();
  });
}

int chm_dmrg_solve(chm_ctx* ctx, int n, const int* D, const double* econv, const int* max_sweeps, const double* noise,
                   const double* rtol, double* energy) {
  return guard(ctx, [&](Ctx& c) {
    if (n < 1 || !D || !econv || !max_sweeps || !noise || !rtol || !energy) throw Error{CHM_ERR_ARG, "bad instructions"};
    (void)hipSetDevice(c.device);
    *energy = c.dmrg_solve(n, D, econv, max_sweeps, noise, rtol);
  });
}

int chm_dmrg_sweep(chm_ctx* ctx, int moving_right, int D, double rtol, double noise, int change, double* energy) {
  return guard(ctx, [&](Ctx& c) {
    if (!energy || D < 1 || !(rtol > 0.0)) throw Error{CHM_ERR_ARG, "bad sweep arguments"};
    if (c.mps.empty() || c.fci.empty()) throw Error{CHM_ERR_STATE, "call chm_dmrg_init first"};
    (void)hipSetDevice(c.device);
    *energy = c.sweep(moving_right != 0, D, rtol, noise, change != 0);
  });
}

int chm_dmrg_sweep_range(chm_ctx* ctx, int moving_right, int first, int last, int D, double rtol, double noise,
                         int change, double* energy) {
  return guard(ctx, [&](Ctx& c) {
    if (!energy || D < 1 || !(rtol > 0.0) || first < 0 || last < 0) throw Error{CHM_ERR_ARG, "bad sweep arguments"};
    if (c.mps.empty() || c.fci.empty()) throw Error{CHM_ERR_STATE, "call chm_dmrg_init first"};
    (void)hipSetDevice(c.device);
    *energy = c.sweep(moving_right != 0, D, rtol, noise, change != 0, first, last);
  });
}

int chm_dmrg_heff_stats(const chm_ctx* ctx, double* flops, double* kernel_ms) {
  if (!ctx || !flops || !kernel_ms) return CHM_ERR_ARG;
  *flops = ctx->c.mv_flops;
  *kernel_ms = ctx->c.mv_ms;
  return CHM_OK;
}

int chm_set_heff_knobs(chm_ctx* ctx, int64_t explicit_max, int item_terms, int group_terms, int chain_mfma,
                       int fchain_min) {
  return guard(ctx, [&](Ctx& c) {
    if (explicit_max >= 0) c.def_explicit_max = explicit_max;
    if (item_terms > 0) c.def_item_terms = item_terms;
    if (group_terms > 0) c.def_group_terms = group_terms;
    if (chain_mfma >= 0) c.def_chain_mfma = chain_mfma != 0;
    if (fchain_min >= 0) c.def_fchain_min = fchain_min;
    c.site = -1;
  });
}

int chm_set_tuning(chm_ctx* ctx, int n, const int64_t* rows) {
  return guard(ctx, [&](Ctx& c) {
    if (n < 0 || (n > 0 && !rows)) throw Error{CHM_ERR_ARG, "bad tuning table"};
    std::vector<Ctx::TuneRow> t;
    for (int r = 0; r < n; r++) {
      const int64_t* v = rows + 9 * r;
      if (v[0] > v[1] || v[2] > v[3]) throw Error{CHM_ERR_ARG, "tuning row with an empty range"};
      t.push_back({v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]});
    }
    c.tune = std::move(t);
    c.site = -1;
  });
}

int chm_tuning_used(const chm_ctx* ctx, int64_t* out) {
  if (!ctx || !out) return CHM_ERR_ARG;
  const Ctx& c = ctx->c;
  out[0] = c.tune_row; out[1] = c.site_dmax; out[2] = c.explicit_max; out[3] = c.chain_item_terms;
  out[4] = c.chain_group_terms; out[5] = c.use_chain_mfma ? 1 : 0; out[6] = c.fchain_min;
  return CHM_OK;
}

int chm_dmrg_site_log(chm_ctx* ctx, int* n, int* sites, int* products, double* values) {
  if (!ctx || !n) return CHM_ERR_ARG;
  auto& log = ctx->c.site_log;
  *n = (int)log.size();
  if (!sites) return CHM_OK;
  if (!products || !values) return CHM_ERR_ARG;
  for (size_t q = 0; q < log.size(); q++) {
    sites[q] = log[q].site;
    products[q] = log[q].products;
    double* v = values + 5 * q;
    v[0] = log[q].energy; v[1] = log[q].seconds; v[2] = log[q].F_H; v[3] = log[q].V; v[4] = log[q].heff_ms;
  }
  log.clear();
  return CHM_OK;
}

int chm_dmrg_stats(const chm_ctx* ctx, double* seconds, int64_t* n_matvec) {
  if (!ctx || !seconds || !n_matvec) return CHM_ERR_ARG;
  for (int q = 0; q < 6; q++) seconds[q] = ctx->c.tm[q];
  *n_matvec = ctx->c.n_matvec_total;
  return CHM_OK;
}

int chm_join(chm_ctx* ctx, int site, int64_t* V) {
  return guard(ctx, [&](Ctx& c) {
    (void)hipSetDevice(c.device);
    c.join(site);
    if (V) *V = (int64_t)c.buf_len[c.s_buf];
  });
}

int chm_s_get(chm_ctx* ctx, double* x, int64_t n, int on_device) {
  return guard(ctx, [&](Ctx& c) {
    if (c.s_site < 0 || n != c.buf_len[c.s_buf]) throw Error{CHM_ERR_STATE, "no two-site object of this size"};
    chm::copy_sync(c.stream, x, c.Sdev.p, n * sizeof(double), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, "S get");
  });
}

int chm_s_set(chm_ctx* ctx, const double* x, int64_t n, int on_device) {
  return guard(ctx, [&](Ctx& c) {
    if (c.s_site < 0 || n != c.buf_len[c.s_buf]) throw Error{CHM_ERR_STATE, "no two-site object of this size"};
    chm::copy_sync(c.stream, c.Sdev.p, x, n * sizeof(double), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, "S set");
  });
}

int chm_split(chm_ctx* ctx, int site, int D, int moving_right, int change, double* disc_weight) {
  return guard(ctx, [&](Ctx& c) {
    (void)hipSetDevice(c.device);
    const double d = c.split(site, D, moving_right != 0, change != 0);
    if (disc_weight) *disc_weight = d;
  });
}

int chm_heff_prepare(chm_ctx* ctx, int site) {
  return guard(ctx, [&](Ctx& c) {
    (void)hipSetDevice(c.device);
    c.prepare(site);
  });
}

int chm_heff_veclen(const chm_ctx* ctx, int64_t* V) {
  if (!ctx || !V) return CHM_ERR_ARG;
  if (ctx->c.site < 0) return CHM_ERR_STATE;
  *V = ctx->c.V;
  return CHM_OK;
}

static void apply_abi(Ctx& c, const double* x, double* y, int on_device) {
  if (!x || !y) throw Error{CHM_ERR_ARG, "null pointer"};
  if (c.site < 0) throw Error{CHM_ERR_STATE, "no prepared site"};
  (void)hipSetDevice(c.device);
  if (on_device) {
    c.apply(x, y);
    chm::hip_check(hipStreamSynchronize(c.stream), "apply");
  } else {
    chm::hip_check(hipMemcpyAsync(c.dx.p, x, c.V * sizeof(double), hipMemcpyHostToDevice, c.stream), "x upload");
    c.apply(c.dx.p, c.dy.p);
    chm::hip_check(hipMemcpyAsync(y, c.dy.p, c.V * sizeof(double), hipMemcpyDeviceToHost, c.stream), "y download");
    chm::hip_check(hipStreamSynchronize(c.stream), "apply");
  }
}

int chm_heff_apply(chm_ctx* ctx, const double* x, double* y, int on_device) {
  return guard(ctx, [&](Ctx& c) {
    c.mv_low = false;
    apply_abi(c, x, y, on_device);
  });
}

int chm_heff_apply_f32(chm_ctx* ctx, const double* x, double* y, int on_device) {
  return guard(ctx, [&](Ctx& c) {
    if (c.site >= 0 && !c.M_split) throw Error{CHM_ERR_STATE, "the site was prepared in the FP64 form"};
    c.mv_low = true;
    try {
      apply_abi(c, x, y, on_device);
    } catch (...) {
      c.mv_low = false;
      throw;
    }
    c.mv_low = false;
  });
}

int chm_heff_set_lower(chm_ctx* ctx, int n_lower, const double* veff, int on_device) {
  return guard(ctx, [&](Ctx& c) {
    if (c.site < 0) throw Error{CHM_ERR_STATE, "no prepared site"};
    if (n_lower < 0 || (n_lower > 0 && !veff)) throw Error{CHM_ERR_ARG, "bad lower-state vectors"};
    (void)hipSetDevice(c.device);
    c.n_lower = 0;
    if (n_lower == 0) return;
    const int64_t n = (int64_t)n_lower * c.V;
    c.lower.alloc((size_t)n);
    chm::copy_sync(c.stream, c.lower.p, veff, (size_t)n * sizeof(double),
                   on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, "lower-state upload");
    std::vector<const double*> ptrs(n_lower);
    for (int q = 0; q < n_lower; q++) ptrs[q] = c.lower.p + (int64_t)q * c.V;
    c.lower_ptrs.upload(ptrs, c.stream);
    c.lower_part.alloc((size_t)chm::dot_parts(c.V) * n_lower);
    c.lower_dot.alloc((size_t)n_lower);
    c.n_lower = n_lower;
  });
}

int chm_set_heff_form(chm_ctx* ctx, int form) {
  return guard(ctx, [&](Ctx& c) {
    if (form != CHM_FORM_FP64 && form != CHM_FORM_MIXED) throw Error{CHM_ERR_ARG, "unknown Heff form"};
    c.heff_form = form;
  });
}

int chm_precond_mode(chm_ctx* ctx, int mode) {
  return guard(ctx, [&](Ctx& c) {
    if (mode != CHM_PRECOND_DIAGONAL && mode != CHM_PRECOND_LEARNED) throw Error{CHM_ERR_ARG, "unknown preconditioner"};
    if (mode == CHM_PRECOND_LEARNED && !c.mlp.ready) throw Error{CHM_ERR_STATE, "set the network parameters first"};
    c.precond_mode = mode;
  });
}

int chm_mlp_nparams(int64_t* n) {
  if (!n) return CHM_ERR_ARG;
  *n = chm::Mlp::nparams();
  return CHM_OK;
}

int chm_mlp_set(chm_ctx* ctx, const float* params, int64_t n) {
  return guard(ctx, [&](Ctx& c) {
    if (!params || n != chm::Mlp::nparams()) throw Error{CHM_ERR_SIZE, "parameter count"};
    (void)hipSetDevice(c.device);
    c.mlp.ensure(1);
    chm::copy_sync(c.stream, c.mlp.params.p, params, n * sizeof(float), hipMemcpyHostToDevice, "parameters");
    c.mlp.reset_optimizer(c.stream);
    c.mlp.ready = true;
  });
}

int chm_mlp_get(chm_ctx* ctx, float* params, int64_t n) {
  return guard(ctx, [&](Ctx& c) {
    if (!params || n != chm::Mlp::nparams()) throw Error{CHM_ERR_SIZE, "parameter count"};
    if (!c.mlp.ready) throw Error{CHM_ERR_STATE, "no parameters"};
    chm::copy_sync(c.stream, params, c.mlp.params.p, n * sizeof(float), hipMemcpyDeviceToHost, "parameters");
  });
}

namespace {
// Uploads n samples (X n x CHM_MLP_IN, p, e, w) into one device buffer.
struct Samples {
  chm::DBuf<float> buf;
  float *X, *p, *e, *w;
  void load(Ctx& c, const float* hX, const float* hp, const float* he, const float* hw, int64_t n) {
    buf.alloc((size_t)n * (chm::MLP_IN + 3));
    X = buf.p;
    p = X + (size_t)n * chm::MLP_IN;
    e = p + n;
    w = e + n;
    chm::copy_sync(c.stream, X, hX, (size_t)n * chm::MLP_IN * sizeof(float), hipMemcpyHostToDevice, "samples");
    chm::copy_sync(c.stream, p, hp, n * sizeof(float), hipMemcpyHostToDevice, "samples");
    chm::copy_sync(c.stream, e, he, n * sizeof(float), hipMemcpyHostToDevice, "samples");
    chm::copy_sync(c.stream, w, hw, n * sizeof(float), hipMemcpyHostToDevice, "samples");
  }
};
}  // namespace

int chm_mlp_forward(chm_ctx* ctx, const float* X, int64_t n, float* out) {
  return guard(ctx, [&](Ctx& c) {
    if (!X || !out || n < 1) throw Error{CHM_ERR_ARG, "bad arguments"};
    if (!c.mlp.ready) throw Error{CHM_ERR_STATE, "no parameters"};
    (void)hipSetDevice(c.device);
    chm::DBuf<float> dX;
    dX.alloc((size_t)n * chm::MLP_IN);
    chm::copy_sync(c.stream, dX.p, X, (size_t)n * chm::MLP_IN * sizeof(float), hipMemcpyHostToDevice, "inputs");
    c.mlp.forward(dX.p, n, c.stream);
    chm::copy_sync(c.stream, out, c.mlp.out.p, n * sizeof(float), hipMemcpyDeviceToHost, "outputs");
  });
}

int chm_mlp_loss_grad(chm_ctx* ctx, const float* X, const float* p, const float* e, const float* w, int64_t n,
                      float* grad, double* loss) {
  return guard(ctx, [&](Ctx& c) {
    if (!X || !p || !e || !w || !grad || !loss || n < 1) throw Error{CHM_ERR_ARG, "bad arguments"};
    if (!c.mlp.ready) throw Error{CHM_ERR_STATE, "no parameters"};
    (void)hipSetDevice(c.device);
    Samples d;
    d.load(c, X, p, e, w, n);
    *loss = c.mlp.loss_grad(d.X, d.p, d.e, d.w, n, c.stream);
    chm::copy_sync(c.stream, grad, c.mlp.grad.p, chm::Mlp::nparams() * sizeof(float), hipMemcpyDeviceToHost, "gradient");
  });
}

int chm_mlp_train(chm_ctx* ctx, const float* X, const float* p, const float* e, const float* w, int64_t n, int batch,
                  int epochs, double lr, uint64_t seed, double* loss_first, double* loss_last) {
  return guard(ctx, [&](Ctx& c) {
    if (!X || !p || !e || !w || n < 1 || batch < 1 || epochs < 1 || !(lr > 0.0)) throw Error{CHM_ERR_ARG, "bad arguments"};
    if (!c.mlp.ready) throw Error{CHM_ERR_STATE, "no parameters"};
    (void)hipSetDevice(c.device);
    Samples d;
    d.load(c, X, p, e, w, n);
    const int64_t B = std::min<int64_t>(batch, n);
    chm::DBuf<float> bX, bs;
    chm::DBuf<int64_t> bidx;
    bX.alloc((size_t)B * chm::MLP_IN);
    bs.alloc((size_t)3 * B);
    bidx.alloc(B);
    std::vector<int64_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    for (int ep = 0; ep < epochs; ep++) {
      std::shuffle(perm.begin(), perm.end(), rng);
      double tot = 0.0;
      int64_t cnt = 0;
      for (int64_t b0 = 0; b0 + B <= n; b0 += B) {
        chm::copy_sync(c.stream, bidx.p, perm.data() + b0, B * sizeof(int64_t), hipMemcpyHostToDevice, "batch");
        chm::launch_gather_rows(d.X, bidx.p, B, bX.p, c.stream);
        chm::launch_gather(d.p, bidx.p, B, bs.p, c.stream);
        chm::launch_gather(d.e, bidx.p, B, bs.p + B, c.stream);
        chm::launch_gather(d.w, bidx.p, B, bs.p + 2 * B, c.stream);
        tot += c.mlp.loss_grad(bX.p, bs.p, bs.p + B, bs.p + 2 * B, B, c.stream);
        cnt++;
        c.mlp.adam((float)lr, c.stream);
      }
      const double mean = cnt ? tot / cnt : 0.0;
      if (ep == 0 && loss_first) *loss_first = mean;
      if (loss_last) *loss_last = mean;
    }
    chm::hip_check(hipStreamSynchronize(c.stream), "training");
  });
}

int chm_harvest(chm_ctx* ctx, int per_iteration) {
  return guard(ctx, [&](Ctx& c) {
    if (per_iteration < 0) throw Error{CHM_ERR_ARG, "negative sample count"};
    c.harvest_per_iter = per_iteration;
  });
}

int chm_harvest_take(chm_ctx* ctx, int64_t* n, float* X, float* p, float* e, float* w) {
  return guard(ctx, [&](Ctx& c) {
    if (!n) throw Error{CHM_ERR_ARG, "null count"};
    const int64_t have = (int64_t)c.hv_p.size();
    if (!X || !p || !e || !w) { *n = have; return; }
    if (*n < have) throw Error{CHM_ERR_SIZE, "arrays too small"};
    std::copy(c.hv_feats.begin(), c.hv_feats.end(), X);
    std::copy(c.hv_p.begin(), c.hv_p.end(), p);
    std::copy(c.hv_e.begin(), c.hv_e.end(), e);
    std::copy(c.hv_w.begin(), c.hv_w.end(), w);
    *n = have;
    c.hv_feats.clear(); c.hv_p.clear(); c.hv_e.clear(); c.hv_w.clear();
  });
}

int chm_precond_stats(const chm_ctx* ctx, int64_t* fallbacks) {
  if (!ctx || !fallbacks) return CHM_ERR_ARG;
  *fallbacks = ctx->c.n_learned_fallbacks;
  return CHM_OK;
}

int chm_mixed_stats(const chm_ctx* ctx, int64_t* low_products, int64_t* full_products) {
  if (!ctx || !low_products || !full_products) return CHM_ERR_ARG;
  *low_products = ctx->c.n_low_mults;
  *full_products = ctx->c.n_full_mults;
  return CHM_OK;
}

int chm_heff_diag(chm_ctx* ctx, double* d, int on_device) {
  return guard(ctx, [&](Ctx& c) {
    if (!d) throw Error{CHM_ERR_ARG, "null pointer"};
    if (c.site < 0) throw Error{CHM_ERR_STATE, "no prepared site"};
    (void)hipSetDevice(c.device);
    if (on_device) {
      c.diag(d);
      chm::hip_check(hipStreamSynchronize(c.stream), "diag");
    } else {
      c.diag(c.dy.p);
      chm::copy_sync(c.stream, d, c.dy.p, c.V * sizeof(double), hipMemcpyDeviceToHost, "diag download");
    }
  });
}

int chm_davidson(chm_ctx* ctx, double* x, int on_device, double rtol, double* eigenvalue, int* n_matvec) {
  return guard(ctx, [&](Ctx& c) {
    if (!x || !eigenvalue) throw Error{CHM_ERR_ARG, "null pointer"};
    if (c.site < 0) throw Error{CHM_ERR_STATE, "no prepared site"};
    if (!(rtol > 0.0)) throw Error{CHM_ERR_ARG, "rtol must be positive"};
    (void)hipSetDevice(c.device);
    if (on_device) {
      *eigenvalue = c.davidson(x, rtol, n_matvec);
      chm::hip_check(hipStreamSynchronize(c.stream), "davidson");
    } else {
      chm::DBuf<double> buf;
      buf.alloc(c.V);
      chm::copy_sync(c.stream, buf.p, x, c.V * sizeof(double), hipMemcpyHostToDevice, "x upload");
      *eigenvalue = c.davidson(buf.p, rtol, n_matvec);
      chm::copy_sync(c.stream, x, buf.p, c.V * sizeof(double), hipMemcpyDeviceToHost, "x download");
    }
  });
}

int chm_flops(const chm_ctx* ctx, double* F_H, double* B_H) {
  if (!ctx || !F_H || !B_H) return CHM_ERR_ARG;
  if (ctx->c.site < 0) return CHM_ERR_STATE;
  *F_H = ctx->c.F_H;
  *B_H = ctx->c.B_H;
  return CHM_OK;
}

int chm_plan_stats(const chm_ctx* ctx, int64_t* stats) {
  if (!ctx || !stats) return CHM_ERR_ARG;
  if (ctx->c.site < 0) return CHM_ERR_STATE;
  std::memcpy(stats, ctx->c.stats, sizeof(ctx->c.stats));
  return CHM_OK;
}

int chm_stream(const chm_ctx* ctx, void** stream) {
  if (!ctx || !stream) return CHM_ERR_ARG;
  *stream = (void*)ctx->c.stream;
  return CHM_OK;
}

int chm_timing_enable(chm_ctx* ctx, int enable) {
  if (!ctx) return CHM_ERR_ARG;
  ctx->c.timing = enable != 0;
  for (double& t : ctx->c.t_ms) t = 0.0;
  ctx->c.t_count = 0;
  return CHM_OK;
}

int chm_timing_read(const chm_ctx* ctx, double* ms, int64_t* count) {
  if (!ctx || !ms || !count) return CHM_ERR_ARG;
  for (int q = 0; q < 3; q++) ms[q] = ctx->c.t_ms[q];
  *count = ctx->c.t_count;
  return CHM_OK;
}

}  // extern "C"
