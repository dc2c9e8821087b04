// Two-site sweep on the device: Sobject::Join and Split, the initial MPS, and the
// DMRG driver (DMRG.cpp:44-446, Sobject.cpp:212-622, DMRGtechnics.cpp:276-298).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <numeric>

#include "context.hpp"

namespace chm {

namespace {
double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// xorshift64 generator of the initial MPS values
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL) {}
  double uniform() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (s >> 11) * (1.0 / 9007199254740992.0);
  }
};

// Householder QR of the m x n column-major A (in place): A <- first min(m, n) columns of
// Q (the rest zero), the contract of dgeqrf + dorgqr used by TensorT::QR.
void householder_q(int m, int n, double* A) {
  const int r = std::min(m, n);
  std::vector<double> tau(r, 0.0);
  for (int k = 0; k < r; k++) {
    double nrm = 0;
    for (int i = k; i < m; i++) nrm += A[i + (size_t)m * k] * A[i + (size_t)m * k];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
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
  std::vector<double> Q((size_t)m * n, 0.0);
  for (int i = 0; i < r; i++) Q[i + (size_t)m * i] = 1.0;
  for (int k = r - 1; k >= 0; k--) {
    if (tau[k] == 0.0) continue;
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
}  // namespace

// Random site tensors in [-1, 1) (TensorT::random, TensorT.cpp:167-173), each left
// normalised per right sector (TensorT::QR, TensorT.cpp:188-280), uploaded.
void Ctx::mps_random(uint64_t seed) {
  Rng rng(seed);
  for (int st = 0; st < L; st++) {
    mps_alloc(st);
    TLay& T = mps[st];
    std::vector<double> h(T.size);
    for (auto& v : h) v = 2.0 * rng.uniform() - 1.0;
    for (const auto& r : sec.list[st + 1]) {
      std::vector<int> ks;
      int dLt = 0;
      for (int k = 0; k < T.nk(); k++)
        if (T.NR[k] == r.N && T.SR[k] == r.S && T.IR[k] == r.I) { ks.push_back(k); dLt += (int)((T.k2i[k + 1] - T.k2i[k]) / r.d); }
      if (dLt == 0) continue;
      std::vector<double> mem((size_t)dLt * r.d);
      int off = 0;
      for (int k : ks) {
        const int dl = (int)((T.k2i[k + 1] - T.k2i[k]) / r.d);
        for (int c = 0; c < r.d; c++)
          for (int l = 0; l < dl; l++) mem[off + l + (size_t)dLt * c] = h[T.k2i[k] + l + (size_t)dl * c];
        off += dl;
      }
      householder_q(dLt, r.d, mem.data());
      off = 0;
      for (int k : ks) {
        const int dl = (int)((T.k2i[k + 1] - T.k2i[k]) / r.d);
        for (int c = 0; c < r.d; c++)
          for (int l = 0; l < dl; l++) h[T.k2i[k] + l + (size_t)dl * c] = mem[off + l + (size_t)dLt * c];
        off += dl;
      }
    }
    if (T.size > 0) copy_sync(stream, mps_buf[st], h.data(), T.size * sizeof(double), hipMemcpyHostToDevice, "MPS init");
  }
}

// Sobject::Join (Sobject.cpp:212-258): S_kappa = sum_SM factor * T_i(left -> M) T_i+1(M -> right).
void Ctx::join(int st) {
  if (!have_problem || st < 0 || st > L - 2) throw Error{ERR_ARG, "site out of range"};
  for (int q : {st, st + 1})
    if (mps[q].buf < 0) throw Error{ERR_MISSING, "MPS tensor of site " + std::to_string(q) + " not set"};
  SLayout Sj;
  Sj.build(sec, st, irreps[st], irreps[st + 1]);
  Sdev.alloc(std::max<int64_t>(Sj.size(), 1));
  if (s_buf < 0) { s_buf = (int32_t)bufs.size(); bufs.push_back(Sdev.p); buf_len.push_back(Sj.size()); }
  else { bufs[s_buf] = Sdev.p; buf_len[s_buf] = Sj.size(); }
  join_terms(sec, mps[st], mps[st + 1], st, Sj, s_buf);
  s_site = st;
}

void Ctx::join_terms(const Sectors& sc, const TLay& Tl, const TLay& Tr, int st, const SLayout& Sj, int32_t out_buf) {
  const int I1 = irreps[st];
  std::vector<UTerm> terms;
  std::vector<std::pair<Ref, std::pair<int, int>>> outs;
  for (int ik = 0; ik < Sj.nk(); ik++) {
    const int nl = Sj.NL[ik], sl = Sj.SL[ik], il = Sj.IL[ik], nr = Sj.NR[ik], sr = Sj.SR[ik], ir = Sj.IR[ik];
    const int tj = Sj.TJ[ik], n1 = Sj.N1[ik];
    const int s1 = (n1 == 1) ? 1 : 0, s2 = (Sj.N2[ik] == 1) ? 1 : 0;
    const int dl = sc.dim(st, nl, sl, il), dr = sc.dim(st + 2, nr, sr, ir);
    Ref out{out_buf, Sj.k2i[ik]};
    outs.push_back({out, {dl, dr}});
    const int fase = phase(sl + sr + s1 + s2);
    const int nm = nl + n1, im = (s1 == 1) ? (il ^ I1) : il;
    const int lo = std::max(std::abs(sl - s1), std::abs(sr - s2)), hi = std::min(sl + s1, sr + s2);
    for (int sm = lo; sm <= hi; sm += 2) {
      const int dm = sc.dim(st + 1, nm, sm, im);
      if (dm <= 0) continue;
      UTerm t{};
      t.out = out; t.m = dl; t.n = dr; t.k1 = dm; t.kind = 0;
      t.alpha = fase * std::sqrt(1.0 * (tj + 1) * (sm + 1)) * wigner6j(sl, sr, tj, s2, s1, sm);
      t.A = Tl.blk(nl, sl, il, nm, sm, im); t.ta = 0; t.lda = dl;
      t.B = Tr.blk(nm, sm, im, nr, sr, ir); t.tb = 0; t.ldb = dm;
      if (!t.A.ok() || !t.B.ok()) throw Error{ERR_MISSING, "join: missing MPS block"};
      terms.push_back(t);
    }
  }
  run_terms(terms, outs, nullptr);
}

// Sobject::Split (Sobject.cpp:260-622): per centre sector an SVD of the recoupled
// two-site matrix, truncation to the D largest singular values over all sectors,
// and the new MPS tensors (U, and sqrt((2SM+1)/(2SR+1)) V^T, with the singular values
// on the side the sweep moves to).
double Ctx::split(int st, int D, bool mr, bool change) {
  const double t0 = now();
  if (s_site != st) throw Error{ERR_STATE, "split: the two-site object is not at this site"};
  if ((int)fci.size() != L + 1) throw Error{ERR_STATE, "split needs the FCI dimensions (chm_set_sectors_auto)"};
  if (!rocblas) {
    rocblas_handle hnd;
    if (rocblas_create_handle(&hnd) != rocblas_status_success) throw Error{ERR_HIP, "rocblas_create_handle"};
    rocblas = hnd;
  }
  rocblas_handle hnd = (rocblas_handle)rocblas;
  rocblas_set_stream(hnd, stream);
  SLayout Sj;
  Sj.build(sec, st, irreps[st], irreps[st + 1]);
  const int I1 = irreps[st], I2 = irreps[st + 1];
  struct Mult { int n, s, i, d; };
  struct Center {
    int nm, sm, im, dLt = 0, dRt = 0, cd = 0;
    bool jac = false;  // batched Jacobi kernel (small sectors) or rocSOLVER
    int64_t moff = 0, uoff = 0, voff = 0, soff = 0, qoff = 0;
    std::vector<Mult> lf, rt;
  };
  std::vector<Center> cs;
  for (const auto& f : fci[st + 1]) {
    Center c;
    c.nm = f.N; c.sm = f.S; c.im = f.I;
    for (int nl = c.nm - 2; nl <= c.nm; nl++) {
      const int s1 = (nl + 1 == c.nm) ? 1 : 0;
      for (int sl = c.sm - s1; sl <= c.sm + s1; sl += 2) {
        if (sl < 0) continue;
        const int il = (s1 == 1) ? (I1 ^ c.im) : c.im;
        const int dl = sec.dim(st, nl, sl, il);
        if (dl > 0) { c.lf.push_back({nl, sl, il, dl}); c.dLt += dl; }
      }
    }
    for (int nr = c.nm; nr <= c.nm + 2; nr++) {
      const int s2 = (nr == c.nm + 1) ? 1 : 0;
      for (int sr = c.sm - s2; sr <= c.sm + s2; sr += 2) {
        if (sr < 0) continue;
        const int ir = (s2 == 1) ? (I2 ^ c.im) : c.im;
        const int dr = sec.dim(st + 2, nr, sr, ir);
        if (dr > 0) { c.rt.push_back({nr, sr, ir, dr}); c.dRt += dr; }
      }
    }
    c.cd = std::min(c.dLt, c.dRt);
    cs.push_back(c);
  }
  int64_t msz = 0, usz = 0, vsz = 0, ssz = 0, qsz = 0;
  for (auto& c : cs) {
    if (c.cd <= 0) continue;
    c.jac = c.cd <= std::min(svd_jacobi_maxn, SVD_JACOBI_MAXN) && std::max(c.dLt, c.dRt) <= 4096;
    c.moff = msz; msz += (int64_t)c.dLt * c.dRt;
    c.uoff = usz; usz += (int64_t)c.dLt * c.cd;
    c.voff = vsz; vsz += (int64_t)c.cd * c.dRt;
    c.soff = ssz; ssz += c.cd;
    if (c.jac) { c.qoff = qsz; qsz += (int64_t)c.cd * c.cd; }
  }
  DBuf<double> M, U, VT, Sg, E, Qw;
  Qw.alloc(std::max<int64_t>(qsz, 1));
  M.alloc(std::max<int64_t>(msz, 1));
  U.alloc(std::max<int64_t>(usz, 1));
  VT.alloc(std::max<int64_t>(vsz, 1));
  Sg.alloc(std::max<int64_t>(ssz, 1));
  E.alloc(std::max<int64_t>(ssz, 1));
  DBuf<rocblas_int> info;
  info.alloc(std::max<size_t>(cs.size(), 1));
  hip_check(hipMemsetAsync(info.p, 0, std::max<size_t>(cs.size(), 1) * sizeof(rocblas_int), stream), "clear info");
  // Assembly and SVD of the selected sectors (sel[q]): the centre matrices from the S
  // blocks with their recoupling factors, then the batched Jacobi kernel for the small
  // sectors and rocSOLVER for the others.
  auto assemble_and_svd = [&](const std::vector<char>& sel) {
    std::vector<DevCopy> items;
    for (size_t q = 0; q < cs.size(); q++) {
      const auto& c = cs[q];
      if (c.cd <= 0 || !sel[q]) continue;
      int lo = 0;
      for (const Mult& a : c.lf) {
        const int s1 = (a.n + 1 == c.nm) ? 1 : 0;
        int ro = 0;
        for (const Mult& b : c.rt) {
          const int s2 = (b.n == c.nm + 1) ? 1 : 0;
          const int fase = phase(a.s + b.s + s1 + s2);
          const int tjmin = std::max(std::abs(b.s - a.s), std::abs(s2 - s1)), tjmax = std::min(s1 + s2, a.s + b.s);
          DevCopy it{};
          if (c.jac && c.dLt < c.dRt) {  // the Jacobi kernel works on M^T
            it.dst = M.p + c.moff + ro + (int64_t)c.dRt * lo;
            it.ldd = c.dRt; it.trans = 1;
          } else {
            it.dst = M.p + c.moff + lo + (int64_t)c.dLt * ro;
            it.ldd = c.dLt;
          }
          it.rows = a.d; it.cols = b.d; it.lds = a.d;
          int nsrc = 0;
          for (int tj = tjmin; tj <= tjmax; tj += 2) {
            const int k = Sj.kappa(a.n, a.s, a.i, c.nm - a.n, b.n - c.nm, tj, b.n, b.s, b.i);
            if (k < 0) throw Error{ERR_STATE, "split: missing two-site block"};
            const double pref = fase * std::sqrt(1.0 * (tj + 1) * (b.s + 1)) * wigner6j(a.s, b.s, tj, s2, s1, c.sm);
            if (nsrc == 0) { it.src0 = Sdev.p + Sj.k2i[k]; it.c0 = pref; }
            else { it.src1 = Sdev.p + Sj.k2i[k]; it.c1 = pref; }
            nsrc++;
          }
          if (nsrc > 2) throw Error{ERR_STATE, "split: more than two spin couplings"};
          it.keep_rows = nsrc ? a.d : 0;
          it.keep_cols = nsrc ? b.d : 0;
          if (!it.src0) it.src0 = Sdev.p;  // not read: keep_rows = 0
          items.push_back(it);
          ro += b.d;
        }
        lo += a.d;
      }
    }
    DBuf<DevCopy> d_items;
    d_items.upload(items, stream);
    launch_copies(d_items.p, (int)items.size(), stream);
    hip_check(hipGetLastError(), "split assembly");
    std::vector<DevSvd> jobs;
    for (size_t q = 0; q < cs.size(); q++) {
      const auto& c = cs[q];
      if (c.cd > 0 && c.jac && sel[q]) {
        DevSvd j{};
        j.A = M.p + c.moff; j.Q = Qw.p + c.qoff; j.U = U.p + c.uoff; j.VT = VT.p + c.voff; j.S = Sg.p + c.soff;
        j.info = info.p + q;
        j.m = std::max(c.dLt, c.dRt); j.n = c.cd; j.wide = c.dLt < c.dRt;
        j.max_sweeps = svd_max_sweeps;
        jobs.push_back(j);
      }
    }
    DBuf<DevSvd> d_jobs;
    d_jobs.upload(jobs, stream);
    launch_svd_jacobi(d_jobs.p, (int)jobs.size(), stream);
    hip_check(hipGetLastError(), "split SVD kernel");
    DBuf<double> jres;
    DBuf<rocblas_int> jsw;
    if (svd_large_gesvdj) {
      jres.alloc(std::max<size_t>(cs.size(), 1));
      jsw.alloc(std::max<size_t>(cs.size(), 1));
    }
    // The large sectors are independent and one rocSOLVER SVD fills little of the device:
    // they run on SVD_STREAMS side streams (default priority, one rocBLAS handle each),
    // which wait for the assembled matrices on the context's stream and which that stream
    // waits for in turn.
    int nlarge = 0;
    for (size_t q = 0; q < cs.size(); q++) nlarge += (cs[q].cd > 0 && !cs[q].jac && sel[q]) ? 1 : 0;
    const int nside = nlarge > 1 ? std::min(nlarge, SVD_STREAMS) : 0;
    if (nside > 0) {
      side_streams(nside);
      hip_check(hipEventRecord(svd_ev_in, stream), "svd event");
      for (int q = 0; q < nside; q++) hip_check(hipStreamWaitEvent(svd_streams[q], svd_ev_in, 0), "svd wait");
    }
    int li = 0;
    for (size_t q = 0; q < cs.size(); q++) {
      const auto& c = cs[q];
      if (c.cd <= 0 || c.jac || !sel[q]) continue;
      rocblas_handle h = nside > 0 ? (rocblas_handle)svd_handles[li % nside] : hnd;
      li++;
      rocblas_status st2;
      if (svd_large_gesvdj)  // Jacobi eigenvectors of the Gram matrix, then QR (rocSOLVER's gesvdj)
        st2 = rocsolver_dgesvdj(h, rocblas_svect_singular, rocblas_svect_singular, c.dLt, c.dRt, M.p + c.moff, c.dLt,
                                0.0, jres.p + q, 100, jsw.p + q, Sg.p + c.soff, U.p + c.uoff, c.dLt, VT.p + c.voff,
                                c.cd, info.p + q);
      else
        st2 = rocsolver_dgesvd(h, rocblas_svect_singular, rocblas_svect_singular, c.dLt, c.dRt, M.p + c.moff, c.dLt,
                               Sg.p + c.soff, U.p + c.uoff, c.dLt, VT.p + c.voff, c.cd, E.p + c.soff,
                               rocblas_outofplace, info.p + q);
      if (st2 != rocblas_status_success) throw Error{ERR_HIP, "rocSOLVER SVD failed"};
    }
    for (int q = 0; q < nside; q++) {
      hip_check(hipEventRecord(svd_ev_out[q], svd_streams[q]), "svd event");
      hip_check(hipStreamWaitEvent(stream, svd_ev_out[q], 0), "svd wait");
    }
    hip_check(hipStreamSynchronize(stream), "split SVD");
  };
  assemble_and_svd(std::vector<char>(cs.size(), 1));
  {
    // a Jacobi sector whose sweeps ended before every pair was orthogonal (the kernel's
    // info flag) is assembled again and decomposed by rocSOLVER's dgesvd, as the
    // reference decomposes every sector (Sobject.cpp:444-518); a rocSOLVER failure is an
    // error
    std::vector<rocblas_int> inf(std::max<size_t>(cs.size(), 1));
    copy_sync(stream, inf.data(), info.p, inf.size() * sizeof(rocblas_int), hipMemcpyDeviceToHost, "svd info");
    std::vector<char> redo(cs.size(), 0);
    bool any = false;
    for (size_t q = 0; q < cs.size(); q++) {
      if (inf[q] == 0) continue;
      if (!cs[q].jac) throw Error{ERR_NUMERIC, "SVD did not converge"};
      cs[q].jac = false;
      redo[q] = 1;
      any = true;
      jacobi_retries++;
    }
    if (any) {
      hip_check(hipMemsetAsync(info.p, 0, std::max<size_t>(cs.size(), 1) * sizeof(rocblas_int), stream), "clear info");
      const bool gj = svd_large_gesvdj;
      svd_large_gesvdj = false;
      try {
        assemble_and_svd(redo);
      } catch (...) {
        svd_large_gesvdj = gj;
        throw;
      }
      svd_large_gesvdj = gj;
      copy_sync(stream, inf.data(), info.p, inf.size() * sizeof(rocblas_int), hipMemcpyDeviceToHost, "svd info");
      for (size_t q = 0; q < cs.size(); q++)
        if (redo[q] && inf[q] != 0) throw Error{ERR_NUMERIC, "SVD did not converge"};
    }
  }
  const double t1 = now();
  std::vector<double> sv(std::max<int64_t>(ssz, 1));
  copy_sync(stream, sv.data(), Sg.p, sv.size() * sizeof(double), hipMemcpyDeviceToHost, "singular values");
  // truncation (Sobject.cpp:444-518)
  double disc = 0.0;
  bool update = false;
  std::vector<int> nd(cs.size());
  if (change) {
    int tot = 0;
    for (size_t i = 0; i < cs.size(); i++) { nd[i] = cs[i].cd; tot += nd[i]; }
    if (tot > D) {
      std::vector<double> vals;
      for (auto& c : cs) for (int i = 0; i < c.cd; i++) vals.push_back(sv[c.soff + i]);
      std::sort(vals.begin(), vals.end(), std::greater<double>());
      const double lb = vals[D];
      for (size_t i = 0; i < cs.size(); i++)
        for (int k = 0; k < nd[i]; k++)
          if (sv[cs[i].soff + k] <= lb) { nd[i] = k; break; }
      double tsum = 0.0, dsum = 0.0;
      for (auto& c : cs)
        for (int k = 0; k < c.cd; k++) {
          const double t = (c.sm + 1) * sv[c.soff + k] * sv[c.soff + k];
          tsum += t;
          if (sv[c.soff + k] <= lb) dsum += t;
        }
      disc = dsum / tsum;
    }
    for (size_t i = 0; i < cs.size(); i++)
      if (nd[i] != sec.dim(st + 1, cs[i].nm, cs[i].sm, cs[i].im)) update = true;
  }
  if (update) {
    std::vector<Sectors::Sec> v;
    for (size_t i = 0; i < cs.size(); i++) v.push_back({cs[i].nm, cs[i].sm, cs[i].im, nd[i]});
    sec.set(st + 1, std::move(v));
  }
  mps_alloc(st);
  mps_alloc(st + 1);
  if (mps[st].size > 0) hip_check(hipMemsetAsync(mps_buf[st], 0, mps[st].size * sizeof(double), stream), "clear");
  if (mps[st + 1].size > 0) hip_check(hipMemsetAsync(mps_buf[st + 1], 0, mps[st + 1].size * sizeof(double), stream), "clear");
  std::vector<DevCopy> items;
  for (auto& c : cs) {
    const int dm = sec.dim(st + 1, c.nm, c.sm, c.im);
    if (dm <= 0) continue;
    const int lim = std::min(dm, c.cd);
    int lo = 0;
    for (const Mult& a : c.lf) {
      Ref r = mps[st].blk(a.n, a.s, a.i, c.nm, c.sm, c.im);
      if (!r.ok()) throw Error{ERR_STATE, "split: missing left block"};
      DevCopy it{};
      it.dst = mps_buf[st] + r.off; it.ldd = a.d; it.rows = a.d; it.cols = dm;
      it.src0 = U.p + c.uoff + lo; it.lds = c.dLt; it.c0 = 1.0;
      it.colscale = mr ? nullptr : Sg.p + c.soff;
      it.keep_rows = a.d; it.keep_cols = lim;
      items.push_back(it);
      lo += a.d;
    }
    int ro = 0;
    for (const Mult& b : c.rt) {
      Ref r = mps[st + 1].blk(c.nm, c.sm, c.im, b.n, b.s, b.i);
      if (!r.ok()) throw Error{ERR_STATE, "split: missing right block"};
      DevCopy it{};
      it.dst = mps_buf[st + 1] + r.off; it.ldd = dm; it.rows = dm; it.cols = b.d;
      it.src0 = VT.p + c.voff + (int64_t)c.cd * ro; it.lds = c.cd; it.c0 = std::sqrt((c.sm + 1.0) / (b.s + 1));
      it.rowscale = mr ? Sg.p + c.soff : nullptr;
      it.keep_rows = lim; it.keep_cols = b.d;
      items.push_back(it);
      ro += b.d;
    }
  }
  {
    DBuf<DevCopy> d_items;
    d_items.upload(items, stream);
    launch_copies(d_items.p, (int)items.size(), stream);
    hip_check(hipGetLastError(), "split scatter");
    hip_check(hipStreamSynchronize(stream), "split");
  }
  if (verbose) {
    int nj = 0, nr = 0, maxcd = 0;
    for (auto& c : cs)
      if (c.cd > 0) { (c.jac ? nj : nr)++; maxcd = std::max(maxcd, c.cd); }
    std::fprintf(stderr, "  split %d: %zu centre sectors (%d Jacobi, %d rocSOLVER, largest %d): SVD %.3f s, rest %.3f s; "
                 "Jacobi sectors redone so far %lld\n", st, cs.size(), nj, nr, maxcd, t1 - t0, now() - t1,
                 (long long)jacobi_retries);
  }
  s_site = -1;
  site = -1;
  return disc;
}

void Ctx::presolve() {  // DMRG.cpp:255-265
  for (int c = 0; c < L - 2; c++) update_operators(c, true);
  total_min = 1e8;
  max_disc_last = 0.0;
}

double Ctx::solve_site(int index, double rtol, double noise, int D, bool mr, bool change, int plan_next) {  // DMRG.cpp:419-446
  double t0 = now();
  join(index);
  double t1 = now();
  tm[0] += t1 - t0;
  prepare(index);
  veff_tilde(index);  // lower states' projector vectors (DMRG.cpp:434, calcVeffTilde)
  if (plan_next >= 0 && !change) {
    // dimensions stay fixed in this sweep: lay out this step's new operator slot and
    // generate the next site's work lists while the GPU solves this one
    alloc_slot(index, mr);
    start_plan(plan_next);
  }
  double t2 = now();
  tm[1] += t2 - t1;
  int nm = 0;
  const double mv0 = mv_ms;
  const double E = davidson(Sdev.p, rtol, &nm) + econst;
  n_matvec_total += nm;
  double t3 = now();
  tm[2] += t3 - t2;
  site_log.push_back({index, nm, E, t3 - t2, F_H, (double)V, mv_ms - mv0});
  {
    // under memory pressure the explicit matrix and the GEMM-tile scratch of this site go
    // back to the pool before Split and the operator update allocate (the next prepare
    // builds its own)
    size_t fb = 0, tb = 0;
    if (hipMemGetInfo(&fb, &tb) == hipSuccess && fb < tb / 4) release_site_scratch();
  }
  if (verbose)
    std::fprintf(stderr, "  site %d: join %.3f s, prepare %.3f s, davidson %.3f s (%d products, %.3f s in Heff kernels)\n", index,
                 t1 - t0, t2 - t1, t3 - t2, nm, (mv_ms - mv0) * 1e-3);
  s_site = index;  // prepare/davidson keep the two-site object
  if (noise > 0.0) {
    launch_add_noise(Sdev.p, S.size(), noise_seed++, noise, stream);
    hip_check(hipStreamSynchronize(stream), "noise");
  }
  const double disc = split(index, D, mr, change);
  if (plan_next >= 0 && change) {
    // the new dimensions are known once Split is done: plan the next site while the
    // operator update of this one runs
    alloc_slot(index, mr);
    start_plan(plan_next);
  }
  tm[3] += now() - t3;
  if (disc > max_disc_last) max_disc_last = disc;
  return E;
}

// DMRG::sweepleft / sweepright (DMRG.cpp:354-417): energy of the last site.
double Ctx::sweep(bool mr, int D, double rtol, double noise_prefactor, bool change, int first, int last) {
  const double t_start = now();
  const int begin = mr ? 0 : L - 2, end = mr ? L - 3 : 1;
  if (first < 0) first = begin;
  if (last < 0) last = end;
  if (mr ? !(0 <= first && first <= last && last <= L - 3) : !(L - 2 >= first && first >= last && last >= 1))
    throw Error{ERR_ARG, "sweep range out of order or out of the chain"};
  // a range that starts mid-sweep continues the sweep begun by an earlier range
  if (first == begin) {
    sweep_noise = std::fabs(noise_prefactor) * max_disc_last;
    max_disc_last = 0.0;
    last_min = 1e8;
  }
  mv_timing = true;
  const double nl = sweep_noise;
  double E = 0.0;
  auto step = [&](int index) {
    const int next = mr ? (index + 1 <= L - 3 ? index + 1 : -1) : (index - 1 >= 1 ? index - 1 : -1);
    E = solve_site(index, rtol, nl, D, mr, change, next);
    total_min = std::min(total_min, E);
    last_min = std::min(last_min, E);
    const double t = now();
    update_operators(index, mr);
    tm[4] += now() - t;
    if (verbose) std::fprintf(stderr, "  site %d: E = %.12f (%.2f s)\n", index, E, now() - t_start);
  };
  try {
    if (mr)
      for (int index = first; index <= last; index++) step(index);
    else
      for (int index = first; index >= last; index--) step(index);
  } catch (...) {
    mv_timing = false;
    mv_used = 0;
    drop_plan();
    throw;
  }
  drop_plan();
  mv_timing = false;
  tm[5] += now() - t_start;
  return E;
}

double Ctx::dmrg_solve(int n, const int* D, const double* econv, const int* max_sweeps, const double* noise,
                       const double* rtol) {  // DMRG.cpp:268-352
  bool change = total_min < 1e8;
  double E = 0.0;
  for (int ins = 0; ins < n; ins++) {
    int it = 0;
    double Eprev = E + 10 * econv[ins];
    while (std::fabs(E - Eprev) > econv[ins] && it < max_sweeps[ins]) {
      Eprev = E;
      E = sweep(false, D[ins], rtol[ins], noise[ins], change);
      change = true;
      E = sweep(true, D[ins], rtol[ins], noise[ins], change);
      it++;
    }
  }
  return total_min;
}

void destroy_rocblas(void* h) {
  if (h) rocblas_destroy_handle((rocblas_handle)h);
}

// n side streams for the large Split SVDs, each with its rocBLAS handle, created once
void Ctx::side_streams(int n) {
  if (!svd_ev_in) hip_check(hipEventCreateWithFlags(&svd_ev_in, hipEventDisableTiming), "svd event");
  while ((int)svd_streams.size() < n) {
    hipStream_t st = nullptr;
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "svd stream");
    svd_streams.push_back(st);
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw Error{ERR_HIP, "rocblas_create_handle"};
    rocblas_set_stream(h, st);
    svd_handles.push_back(h);
    hipEvent_t e = nullptr;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "svd event");
    svd_ev_out.push_back(e);
  }
}

}  // namespace chm
