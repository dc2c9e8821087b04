// Davidson eigensolver and two-site DMRG sweep of CheMPS2 restated on the CPU
// (oracle; test infrastructure only).
#include "dmrg.hpp"

#include <algorithm>

namespace orc {

static double dot(int n, const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static double nrm2(int n, const double* a) { return std::sqrt(dot(n, a, a)); }
static void axpy(int n, double a, const double* x, double* y) {
  for (int i = 0; i < n; i++) y[i] += a * x[i];
}

// ------------------------------------------------------------ Davidson -----
Davidson::Davidson(int n_, int max_vec_, int keep_, double rtol_, double cutoff_)
    : n(n_), max_vec(max_vec_), keep(keep_), rtol(rtol_), cutoff(cutoff_) {
  M.assign((size_t)max_vec * max_vec, 0.0);
  Mvec.assign((size_t)max_vec * max_vec, 0.0);
  Meig.assign(max_vec, 0.0);
  diag.assign(n, 0.0); t.assign(n, 0.0); u.assign(n, 0.0); work.assign(n, 0.0);
  vecs.resize(max_vec); hvecs.resize(max_vec);
}

char Davidson::fetch(double** ptr) {  // Davidson.cpp:105-191
  if (state == 'I') {
    ptr[0] = t.data();
    ptr[1] = diag.data();
    state = 'U';
    return 'A';
  }
  if (state == 'U') {
    if (nrm2(n, t.data()) == 0.0)  // SafetyCheckGuess, Davidson.cpp:202-212
      for (int i = 0; i < n; i++) t[i] = rng.uniform();
    add_new_vec();
    ptr[0] = vecs[num_vec].data();
    ptr[1] = hvecs[num_vec].data();
    n_mult++;
    state = 'N';
    return 'B';
  }
  if (state == 'N') {
    const double rnorm = diagonalize_and_residual();
    if (rnorm > rtol) {
      new_vec();
      if (num_vec == max_vec) {
        deflate();
        ptr[0] = vecs[num_vec].data();
        ptr[1] = hvecs[num_vec].data();
        n_mult++;
        num_vec++;
        state = 'F';
        return 'B';
      }
      add_new_vec();
      ptr[0] = vecs[num_vec].data();
      ptr[1] = hvecs[num_vec].data();
      n_mult++;
      state = 'N';
      return 'B';
    }
    state = 'C';
    ptr[0] = u.data();
    ptr[1] = work.data();
    work[0] = Meig[0];
    return 'C';
  }
  if (state == 'F') {
    if (num_vec == keep) {
      mxm_after_deflation();
      add_new_vec();
      ptr[0] = vecs[num_vec].data();
      ptr[1] = hvecs[num_vec].data();
      n_mult++;
      state = 'N';
      return 'B';
    }
    ptr[0] = vecs[num_vec].data();
    ptr[1] = hvecs[num_vec].data();
    n_mult++;
    num_vec++;
    state = 'F';
    return 'B';
  }
  return 'D';
}

void Davidson::add_new_vec() {
  for (int c = 0; c < num_vec; c++) axpy(n, -dot(n, t.data(), vecs[c].data()), vecs[c].data(), t.data());
  const double a = 1.0 / nrm2(n, t.data());
  for (int i = 0; i < n; i++) t[i] *= a;
  if (vecs[num_vec].empty()) {
    vecs[num_vec].assign(n, 0.0);
    hvecs[num_vec].assign(n, 0.0);
  }
  std::swap(vecs[num_vec], t);
}

double Davidson::diagonalize_and_residual() {
  const int mv = max_vec;
  for (int c = 0; c < num_vec; c++) {
    M[c + mv * num_vec] = dot(n, vecs[num_vec].data(), hvecs[c].data());
    M[num_vec + mv * c] = M[c + mv * num_vec];
  }
  M[num_vec + mv * num_vec] = dot(n, vecs[num_vec].data(), hvecs[num_vec].data());
  num_vec++;
  sym_eig(num_vec, M.data(), mv, Meig.data(), Mvec.data(), mv);  // ascending, like dsyev
  std::fill(t.begin(), t.end(), 0.0);
  std::fill(u.begin(), u.end(), 0.0);
  for (int c = 0; c < num_vec; c++) {
    axpy(n, Mvec[c], hvecs[c].data(), t.data());
    axpy(n, Mvec[c], vecs[c].data(), u.data());
  }
  axpy(n, -Meig[0], u.data(), t.data());
  return nrm2(n, t.data());
}

void Davidson::new_vec() {
  const double shift = Meig[0];
  for (int i = 0; i < n; i++) {
    const double d = diag[i] - shift;
    work[i] = (std::fabs(d) > cutoff) ? u[i] / d : u[i] / cutoff;
  }
  const double a = -dot(n, work.data(), t.data()) / dot(n, work.data(), u.data());
  axpy(n, a, u.data(), t.data());
  for (int i = 0; i < n; i++) {
    const double d = diag[i] - shift;
    t[i] = (std::fabs(d) > cutoff) ? -t[i] / d : -t[i] / cutoff;
  }
}

void Davidson::deflate() {
  const int mv = max_vec;
  if (keep <= 1) {
    const double a = 1.0 / nrm2(n, u.data());
    for (int i = 0; i < n; i++) vecs[0][i] = u[i] * a;
  } else {
    std::vector<double> E((size_t)n * keep, 0.0);
    std::copy(u.begin(), u.end(), E.begin());
    for (int c = 1; c < keep; c++)
      for (int v = 0; v < mv; v++) axpy(n, Mvec[v + mv * c], vecs[v].data(), E.data() + (size_t)n * c);
    std::vector<double> S((size_t)keep * keep), se(keep), sv((size_t)keep * keep), low((size_t)keep * keep, 0.0);
    for (int a = 0; a < keep; a++)
      for (int b = 0; b < keep; b++) S[a + keep * b] = dot(n, E.data() + (size_t)n * a, E.data() + (size_t)n * b);
    sym_eig(keep, S.data(), keep, se.data(), sv.data(), keep);
    for (int c = 0; c < keep; c++) {
      const double f = std::pow(se[c], -0.25);
      for (int r = 0; r < keep; r++) sv[r + keep * c] *= f;
    }
    for (int a = 0; a < keep; a++)  // Lowdin = sv sv^T = S^{-1/2}
      for (int b = 0; b < keep; b++) {
        double s = 0;
        for (int c = 0; c < keep; c++) s += sv[a + keep * c] * sv[b + keep * c];
        low[a + keep * b] = s;
      }
    for (int v = 0; v < keep; v++) {
      std::fill(vecs[v].begin(), vecs[v].end(), 0.0);
      for (int v2 = 0; v2 < keep; v2++) axpy(n, low[v2 + keep * v], E.data() + (size_t)n * v2, vecs[v].data());
    }
  }
  num_vec = 0;
}

void Davidson::mxm_after_deflation() {
  const int mv = max_vec;
  for (int a = 0; a < keep; a++)
    for (int b = a; b < keep; b++) {
      M[a + mv * b] = dot(n, vecs[a].data(), hvecs[b].data());
      M[b + mv * a] = M[a + mv * b];
    }
}

double solve_davidson(Heff& H, Sobject& S, double rtol, int* n_mult) {
  const int V = (int)S.size();
  Davidson dv(V, 32, 3, rtol, 1e-12);  // DAVIDSON_NUM_VEC, _KEEP, _PRECOND_CUTOFF (Options.h:70-74)
  double* p[2];
  char ins = dv.fetch(p);
  assert(ins == 'A');
  S.prog2symm();
  std::copy(S.data.begin(), S.data.end(), p[0]);
  H.diag(p[1]);
  ins = dv.fetch(p);
  while (ins == 'B') {
    H.apply(p[0], p[1]);
    ins = dv.fetch(p);
  }
  assert(ins == 'C');
  std::copy(p[0], p[0] + V, S.data.begin());
  S.symm2prog();
  if (n_mult) *n_mult = dv.n_mult;
  return p[1][0];
}

// ----------------------------------------------------------------- DMRG -----
void left_normalize(TensorT& T) {  // TensorT.cpp:188-280 via DMRGtechnics.cpp:276-298
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
        std::vector<double> mem((size_t)dLt * dR), R((size_t)dR * dR);
        int off = 0;
        for (int k : ks) {
          const int dL = bk->dim(T.index, T.NL[k], T.SL[k], T.IL[k]);
          for (int r = 0; r < dR; r++)
            for (int l = 0; l < dL; l++) mem[off + l + (size_t)dLt * r] = T.data[T.k2i[k] + l + (size_t)dL * r];
          off += dL;
        }
        qr_inplace(dLt, dR, mem.data(), R.data());
        off = 0;
        for (int k : ks) {
          const int dL = bk->dim(T.index, T.NL[k], T.SL[k], T.IL[k]);
          for (int r = 0; r < dR; r++)
            for (int l = 0; l < dL; l++) T.data[T.k2i[k] + l + (size_t)dL * r] = mem[off + l + (size_t)dLt * r];
          off += dL;
        }
      }
}

std::vector<int> reorder_d2h(int L, int group, const int* irreps) {  // Problem.cpp:57-90
  std::vector<int> f2(L);
  for (int i = 0; i < L; i++) f2[i] = i;
  if (group != 7) return f2;
  static const int order[8] = {0, 5, 7, 2, 6, 3, 1, 4};
  int d = 0;
  for (int o = 0; o < 8; o++)
    for (int h = 0; h < L; h++)
      if (irreps[h] == order[o]) f2[d++] = h;
  assert(d == L);
  return f2;
}

void DMRG::init(int D0, uint64_t seed) {  // DMRG.cpp:44-116, 118-253
  P.build_mx();
  bk.init(&P, D0);
  rng = Rng(seed);
  mps.assign(P.L, TensorT());
  for (int s = 0; s < P.L; s++) {
    mps[s].init(s, &bk);
    for (auto& v : mps[s].data) v = 2.0 * rng.uniform() - 1.0;  // TensorT::random, TensorT.cpp:167-173
    left_normalize(mps[s]);
  }
  ops.init(&P, &bk);
  presolve();
}

void DMRG::presolve() {
  for (int c = 0; c < P.L - 2; c++) ops.update_moving_right(c, mps);
  total_min = 1e8;
  max_disc_last = 0.0;
}

double DMRG::solve_site(int index, double rtol, double noise, int D, bool moving_right, bool change) {
  Sobject S;
  S.init(index, &bk);
  S.join(&mps[index], &mps[index + 1]);
  Heff H;
  H.init(&P, &bk, &S, &ops);
  int nm = 0;
  double E = solve_davidson(H, S, rtol, &nm) + P.econst;
  n_matvec += nm;
  matvec_flops += H.flops;
  if (noise > 0.0)  // Sobject::addNoise, Sobject.cpp:652-659
    for (auto& v : S.data) v += (rng.uniform() - 0.5) * noise;
  const double disc = S.split(&mps[index], &mps[index + 1], D, moving_right, change);
  if (disc > max_disc_last) max_disc_last = disc;
  return E;
}

double DMRG::sweep_left(bool change, const Instruction& ins) {
  double E = 0.0;
  const double noise = std::fabs(ins.noise) * max_disc_last;
  max_disc_last = 0.0;
  last_min = 1e8;
  for (int index = P.L - 2; index > 0; index--) {
    E = solve_site(index, ins.rtol, noise, ins.D, false, change);
    if (E < total_min) total_min = E;
    if (E < last_min) last_min = E;
    ops.update_moving_left(index, mps);
  }
  return E;
}

double DMRG::sweep_right(bool change, const Instruction& ins) {
  double E = 0.0;
  const double noise = std::fabs(ins.noise) * max_disc_last;
  max_disc_last = 0.0;
  last_min = 1e8;
  for (int index = 0; index < P.L - 2; index++) {
    E = solve_site(index, ins.rtol, noise, ins.D, true, change);
    if (E < total_min) total_min = E;
    if (E < last_min) last_min = E;
    ops.update_moving_right(index, mps);
  }
  return E;
}

double DMRG::solve(const std::vector<Instruction>& scheme) {
  bool change = total_min < 1e8;
  double E = 0.0;
  for (const Instruction& ins : scheme) {
    int it = 0;
    double Eprev = E + 10 * ins.econv;
    while (std::fabs(E - Eprev) > ins.econv && it < ins.max_sweeps) {
      Eprev = E;
      E = sweep_left(change, ins);
      change = true;
      E = sweep_right(change, ins);
      it++;
    }
  }
  return total_min;
}

}  // namespace orc
