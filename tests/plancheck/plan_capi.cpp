// Host-only access to the chemps2_amd work-list generator, for CPU tests: the tests
// execute the generated terms with numpy and compare with the oracle's Heff.
#include <array>
#include <cstring>

#include "../../chemps2_amd/csrc/context.hpp"
#include "../../chemps2_amd/csrc/heff_plan.hpp"
#include "../../chemps2_amd/csrc/update_plan.hpp"

using namespace chm;

namespace {
struct PC {
  HostState c;
  std::vector<std::unique_ptr<OpT>> owned;
  std::vector<OpT*> tensors;  // by buf id - BUF_TENSOR0
  SLayout S;
  std::unique_ptr<HeffPlanner> plan;
  std::string err;
  std::vector<TLay> mps;
  int next_buf = 1000000;  // buffer ids of MPS tensors and new operator tensors
  std::unique_ptr<UpdatePlanner> up;
  std::vector<std::array<int64_t, 5>> newt;  // kind, k, m, buf, size
};
}  // namespace

extern "C" {

void* pc_create(int L, int group, const int* irreps, int N, int TwoS, int Irrep, const double* mx) {
  auto* p = new PC();
  static const int nirr_of[8] = {1, 2, 2, 2, 4, 4, 4, 8};
  p->c.L = L; p->c.group = group; p->c.nirrep = nirr_of[group]; p->c.N = N; p->c.TwoS = TwoS; p->c.Irrep = Irrep;
  p->c.irreps.assign(irreps, irreps + L);
  p->c.mx.assign(mx, mx + (size_t)L * L * L * L);
  p->c.sec.resize(L);
  p->c.sets.resize(L - 1);
  p->c.have_problem = true;
  return p;
}

void pc_destroy(void* h) { delete (PC*)h; }

void pc_sectors(void* h, int b, int n, const int* N, const int* S, const int* I, const int* d) {
  PC* p = (PC*)h;
  std::vector<Sectors::Sec> v;
  for (int q = 0; q < n; q++) v.push_back({N[q], S[q], I[q], d[q]});
  p->c.sec.set(b, std::move(v));
}

void pc_sectors_auto(void* h, int D) { auto_sectors(((PC*)h)->c, D); }

int pc_nsectors(void* h, int b, int* out) {
  PC* p = (PC*)h;
  const auto& v = p->c.sec.list[b];
  if (out)
    for (size_t q = 0; q < v.size(); q++) { out[4 * q] = v[q].N; out[4 * q + 1] = v[q].S; out[4 * q + 2] = v[q].I; out[4 * q + 3] = v[q].d; }
  return (int)v.size();
}

// FCI dimensions of boundary b that auto_sectors derived (N, 2S, I, dim per sector).
int pc_fci(void* h, int b, int* out) {
  PC* p = (PC*)h;
  const auto& v = p->c.fci[b];
  if (out)
    for (size_t q = 0; q < v.size(); q++) { out[4 * q] = v[q].N; out[4 * q + 1] = v[q].S; out[4 * q + 2] = v[q].I; out[4 * q + 3] = v[q].d; }
  return (int)v.size();
}

// Registers a tensor; returns its buffer id and writes its size.
int pc_tensor(void* h, int cnt, int mr, int kind, int k, int m, int two_j, int n_elec, int n_irrep, long* size) {
  PC* p = (PC*)h;
  OpSet& s = p->c.sets[cnt];
  s.dir = mr ? 1 : 2;
  auto t = std::make_unique<OpT>();
  t->build(p->c.sec, cnt + 1, two_j, n_elec, n_irrep);
  t->buf = BUF_TENSOR0 + (int)p->tensors.size();
  *size = t->size;
  auto s1 = [&](std::vector<OpT*>& v) -> OpT*& { if ((int)v.size() <= k) v.resize(k + 1, nullptr); return v[k]; };
  auto s2 = [&](std::vector<std::vector<OpT*>>& v) -> OpT*& {
    if ((int)v.size() <= k) v.resize(k + 1);
    if ((int)v[k].size() <= m) v[k].resize(m + 1, nullptr);
    return v[k][m];
  };
  OpT* raw = t.get();
  switch (kind) {
    case KIND_L: s1(s.L) = raw; break;
    case KIND_Q: s1(s.Q) = raw; break;
    case KIND_X: s.X = raw; break;
    case KIND_F0: s2(s.F0) = raw; break;
    case KIND_F1: s2(s.F1) = raw; break;
    case KIND_S0: s2(s.S0) = raw; break;
    case KIND_S1: s2(s.S1) = raw; break;
    case KIND_A: s2(s.A) = raw; break;
    case KIND_B: s2(s.B) = raw; break;
    case KIND_C: s2(s.C) = raw; break;
    case KIND_D: s2(s.D) = raw; break;
  }
  p->tensors.push_back(raw);
  p->owned.push_back(std::move(t));
  return raw->buf;
}

// Registers every tensor of slot cnt with the library's own shapes; returns the count.
int pc_synthetic_set(void* h, int cnt, int mr) {
  PC* p = (PC*)h;
  int n = 0;
  for (const OpShape& sh : opset_shapes(p->c, cnt, mr != 0)) {
    long sz;
    pc_tensor(h, cnt, mr, sh.kind, sh.k, sh.m, sh.two_j, sh.n_elec, sh.n_irrep, &sz);
    n++;
  }
  return n;
}

// MPS layout of `site`; returns its buffer id and writes the size.
int pc_mps(void* h, int site, long* size) {
  PC* p = (PC*)h;
  if ((int)p->mps.size() < p->c.L) p->mps.resize(p->c.L);
  TLay& T = p->mps[site];
  T.build(p->c.sec, site, p->c.irreps[site]);
  T.buf = p->next_buf++;
  *size = T.size;
  return T.buf;
}

// Work list of updateMovingRight/Left(index); returns the number of new tensors (-1 on error).
long pc_update(void* h, int index, int mr) {
  PC* p = (PC*)h;
  try {
    HostState& c = p->c;
    OpSet* prev = nullptr;
    if (mr && index > 0) { c.validate_set(index - 1, true); prev = &c.sets[index - 1]; }
    if (!mr && index < c.L - 2) { c.validate_set(index + 1, false); prev = &c.sets[index + 1]; }
    c.sets[index] = OpSet();
    OpSet& o = c.sets[index];
    o.dir = mr ? 1 : 2;
    p->newt.clear();
    for (const OpShape& sh : opset_shapes(c, index, mr != 0)) {
      auto t = std::make_unique<OpT>();
      t->build(c.sec, index + 1, sh.two_j, sh.n_elec, sh.n_irrep);
      t->moving_right = mr != 0;
      t->jw_phase = (sh.kind == KIND_L || sh.kind == KIND_Q);
      t->prime_last = (sh.kind == KIND_F1 || sh.kind == KIND_D) ? (mr != 0) : true;
      t->buf = p->next_buf++;
      OpT* raw = t.get();
      auto s1 = [&](std::vector<OpT*>& v) -> OpT*& { if ((int)v.size() <= sh.k) v.resize(sh.k + 1, nullptr); return v[sh.k]; };
      auto s2 = [&](std::vector<std::vector<OpT*>>& v) -> OpT*& {
        if ((int)v.size() <= sh.k) v.resize(sh.k + 1);
        if ((int)v[sh.k].size() <= sh.m) v[sh.k].resize(sh.m + 1, nullptr);
        return v[sh.k][sh.m];
      };
      switch (sh.kind) {
        case KIND_L: s1(o.L) = raw; break;
        case KIND_Q: s1(o.Q) = raw; break;
        case KIND_X: o.X = raw; break;
        case KIND_F0: s2(o.F0) = raw; break;
        case KIND_F1: s2(o.F1) = raw; break;
        case KIND_S0: s2(o.S0) = raw; break;
        case KIND_S1: s2(o.S1) = raw; break;
        case KIND_A: s2(o.A) = raw; break;
        case KIND_B: s2(o.B) = raw; break;
        case KIND_C: s2(o.C) = raw; break;
        case KIND_D: s2(o.D) = raw; break;
      }
      p->newt.push_back({sh.kind, sh.k, sh.m, raw->buf, raw->size});
      o.owned.push_back(std::move(t));
    }
    std::vector<int> Qsite;
    const int nC = mr ? c.L - 1 - index : index + 1;
    for (int k = 0; k < nC; k++) Qsite.push_back(mr ? index + 1 + k : index - k);
    p->up = std::make_unique<UpdatePlanner>();
    UpdatePlanner& u = *p->up;
    u.sec = &c.sec; u.irr = c.irreps.data(); u.mx = c.mx.data(); u.L = c.L;
    const TLay* T = &p->mps[mr ? index : index + 1];
    if (mr) u.moving_right(index, o, prev, T, Qsite);
    else u.moving_left(index, o, prev, T, Qsite);
    u.flatten();
    return (long)p->newt.size();
  } catch (const Error& e) {
    p->err = e.msg;
    return -1;
  }
}

void pc_update_newt(void* h, long* out) {
  PC* p = (PC*)h;
  for (size_t q = 0; q < p->newt.size(); q++)
    for (int j = 0; j < 5; j++) out[5 * q + j] = p->newt[q][j];
}

long pc_update_nterms(void* h, int level) { return (long)((PC*)h)->up->terms[level].size(); }

// ints[16]: out.buf, m, n, k1, k2, A.buf, B.buf, C.buf, ta, tb, tc, lda, ldb, ldc, kind, 0;
// offs[4]: out, A, B, C; alpha.
void pc_update_terms(void* h, int level, int* ints, long* offs, double* alpha) {
  PC* p = (PC*)h;
  const auto& v = p->up->terms[level];
  for (size_t q = 0; q < v.size(); q++) {
    const UTerm& t = v[q];
    int* o = ints + 16 * q;
    o[0] = t.out.buf; o[1] = t.m; o[2] = t.n; o[3] = t.k1; o[4] = t.k2; o[5] = t.A.buf; o[6] = t.B.buf; o[7] = t.C.buf;
    o[8] = t.ta; o[9] = t.tb; o[10] = t.tc; o[11] = t.lda; o[12] = t.ldb; o[13] = t.ldc; o[14] = t.kind; o[15] = 0;
    long* f = offs + 4 * q;
    f[0] = t.out.off; f[1] = t.A.off; f[2] = t.B.off; f[3] = t.C.off;
    alpha[q] = t.alpha;
  }
}

long pc_update_ncombos(void* h, long* nsrc_total, long* arena) {
  PC* p = (PC*)h;
  long n = 0;
  for (auto& c : p->up->combos) n += (long)c.src.size();
  *nsrc_total = n;
  *arena = p->up->combo_size;
  return (long)p->up->combos.size();
}

void pc_update_combos(void* h, long* meta, int* sbuf, long* soff, double* scoef) {
  PC* p = (PC*)h;
  long s = 0;
  for (size_t q = 0; q < p->up->combos.size(); q++) {
    const Combo& c = p->up->combos[q];
    meta[3 * q] = c.out_off; meta[3 * q + 1] = c.size; meta[3 * q + 2] = (long)c.src.size();
    for (auto& e : c.src) { sbuf[s] = e.first.buf; soff[s] = e.first.off; scoef[s] = e.second; s++; }
  }
}

double pc_update_flops(void* h) { return ((PC*)h)->up->flops; }

long pc_plan(void* h, int site) {
  PC* p = (PC*)h;
  try {
    if (site > 0) p->c.validate_set(site - 1, true);
    if (site < p->c.L - 2) p->c.validate_set(site + 1, false);
    p->S.build(p->c.sec, site, p->c.irreps[site], p->c.irreps[site + 1]);
    p->plan = std::make_unique<HeffPlanner>();
    p->plan->init(&p->c.sec, &p->c.irreps, p->c.nirrep, p->c.mx.data(), p->c.L, &p->S, site > 0 ? &p->c.sets[site - 1] : nullptr,
                  site < p->c.L - 2 ? &p->c.sets[site + 1] : nullptr);
    p->plan->build();
    return (long)p->plan->nterms;
  } catch (const Error& e) {
    p->err = e.msg;
    return -1;
  }
}

const char* pc_error(void* h) { return ((PC*)h)->err.c_str(); }
double pc_flops(void* h) { return ((PC*)h)->plan->flops; }
int pc_nk(void* h) { return ((PC*)h)->S.nk(); }
void pc_k2i(void* h, long* k2i, int* dims) {
  PC* p = (PC*)h;
  for (int k = 0; k <= p->S.nk(); k++) k2i[k] = p->S.k2i[k];
  for (int k = 0; k < p->S.nk(); k++) {
    dims[2 * k] = p->c.sec.dim(p->S.index, p->S.NL[k], p->S.SL[k], p->S.IL[k]);
    dims[2 * k + 1] = p->c.sec.dim(p->S.index + 2, p->S.NR[k], p->S.SR[k], p->S.IR[k]);
  }
}

// terms: ints[8] = out, in, tl, tr, xr_first, Lbuf, Rbuf, 0; offs[2]; alpha
void pc_terms(void* h, int* ints, long* offs, double* alpha) {
  PC* p = (PC*)h;
  size_t q = 0;
  for (const auto& ch : p->plan->chunks)
  for (const Term& t : ch) {
    int* o = ints + 8 * q;
    o[0] = t.out; o[1] = t.in; o[2] = t.tl; o[3] = t.tr; o[4] = t.xr_first; o[5] = t.Lb.buf; o[6] = t.Rb.buf; o[7] = 0;
    offs[2 * q] = p->plan->ref(t, t.Lb).off;
    offs[2 * q + 1] = p->plan->ref(t, t.Rb).off;
    alpha[q] = t.alpha;
    q++;
  }
}

long pc_ncombos(void* h, long* nsrc_total, long* arena) {
  PC* p = (PC*)h;
  long n = 0;
  for (auto& c : p->plan->combos) n += (long)c.src.size();
  *nsrc_total = n;
  *arena = p->plan->combo_size;
  return (long)p->plan->combos.size();
}

// combos: meta[3] = out_off, size, nsrc; src: buf, off, coef
void pc_combos(void* h, long* meta, int* sbuf, long* soff, double* scoef) {
  PC* p = (PC*)h;
  long s = 0;
  for (size_t q = 0; q < p->plan->combos.size(); q++) {
    const Combo& c = p->plan->combos[q];
    meta[3 * q] = c.out_off; meta[3 * q + 1] = c.size; meta[3 * q + 2] = (long)c.src.size();
    for (auto& e : c.src) { sbuf[s] = e.first.buf; soff[s] = e.first.off; scoef[s] = e.second; s++; }
  }
}

long pc_ndiag(void* h) { return (long)((PC*)h)->plan->diag.size(); }
// diag: ints[4] = k, kind, Lbuf, Rbuf; offs[2]; coef
void pc_diag(void* h, int* ints, long* offs, double* coef) {
  PC* p = (PC*)h;
  for (size_t q = 0; q < p->plan->diag.size(); q++) {
    const DiagTerm& d = p->plan->diag[q];
    ints[4 * q] = d.k; ints[4 * q + 1] = d.kind; ints[4 * q + 2] = d.Lb.buf; ints[4 * q + 3] = d.Rb.buf;
    offs[2 * q] = d.Lb.off; offs[2 * q + 1] = d.Rb.off;
    coef[q] = d.coef;
  }
}

}  // extern "C"
