// Virtual-dimension bookkeeping and operator-set shapes, for callers that let the
// library derive sector dimensions and tensor shapes itself.
#include <algorithm>
#include <cstdlib>

#include "context.hpp"

namespace chm {

// FCI dimensions of every boundary capped at 262144 and scaled to a total of about D
// per boundary (SyBookkeeper.cpp:29-49, 91-130, 172-269).
void auto_sectors(HostState& c, int D) {
  const int L = c.L, N = c.N, TS = c.TwoS, nirr = c.nirrep;
  std::vector<int> Nmin(L + 1), Nmax(L + 1);
  std::vector<std::vector<int>> Smin(L + 1), Smax(L + 1);
  for (int b = 0; b <= L; b++) {
    Nmin[b] = std::max(std::max(0, N + 2 * (b - L)), b - L + (N + TS) / 2);
    Nmax[b] = std::min(std::min(2 * b, N), b + (N - TS) / 2);
    for (int n = Nmin[b]; n <= Nmax[b]; n++) {
      const int tmp = L - b - std::abs(N - n - L + b);
      Smin[b].push_back(std::max(n % 2, TS - tmp));
      Smax[b].push_back(std::min(b - std::abs(b - n), TS + tmp));
    }
  }
  std::vector<std::map<std::tuple<int, int, int>, long>> fci(L + 1);
  auto get = [&](int b, int n, int s, int i) -> long {
    auto it = fci[b].find({n, s, i});
    return it == fci[b].end() ? 0 : it->second;
  };
  const long cutoff = 262144;
  fci[0][{0, 0, 0}] = 1;
  for (int b = 1; b <= L; b++) {
    const int Ib = c.irreps[b - 1];
    for (int n = Nmin[b]; n <= Nmax[b]; n++)
      for (int s = Smin[b][n - Nmin[b]]; s <= Smax[b][n - Nmin[b]]; s += 2)
        for (int I = 0; I < nirr; I++) {
          const long v = get(b - 1, n, s, I) + get(b - 1, n - 2, s, I) + get(b - 1, n - 1, s + 1, I ^ Ib) +
                         get(b - 1, n - 1, s - 1, I ^ Ib);
          fci[b][{n, s, I}] = std::min(cutoff, v);
        }
  }
  {
    const long rhs = get(L, N, TS, c.Irrep);
    for (int I = 0; I < nirr; I++) fci[L][{N, TS, I}] = 0;
    fci[L][{N, TS, c.Irrep}] = std::min(1L, rhs);
  }
  for (int b = L - 1; b >= 0; b--) {
    const int Ib = c.irreps[b];
    for (int n = Nmin[b]; n <= Nmax[b]; n++)
      for (int s = Smin[b][n - Nmin[b]]; s <= Smax[b][n - Nmin[b]]; s += 2)
        for (int I = 0; I < nirr; I++) {
          const long v = get(b + 1, n, s, I) + get(b + 1, n + 2, s, I) + get(b + 1, n + 1, s + 1, I ^ Ib) +
                         get(b + 1, n + 1, s - 1, I ^ Ib);
          long& d = fci[b][{n, s, I}];
          d = std::min(d, std::min(cutoff, v));
        }
  }
  c.sec.resize(L);
  c.fci.assign(L + 1, {});
  for (int b = 0; b <= L; b++)
    for (auto& kv : fci[b])
      if (kv.second > 0) c.fci[b].push_back({std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), (int)kv.second});
  for (int b = 0; b <= L; b++) {
    long tot = 0;
    for (auto& kv : fci[b]) tot += kv.second;
    const double factor = (b >= 1 && b <= L - 1 && tot > D) ? (1.0 * D) / tot : 1.0;
    std::vector<Sectors::Sec> v;
    for (auto& kv : fci[b]) {  // std::map order == (N, 2S, I) ascending
      if (kv.second <= 0) continue;
      const int d = (factor < 1.0) ? (int)(std::ceil(factor * kv.second) + 0.1) : (int)kv.second;
      v.push_back({std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), d});
    }
    c.sec.set(b, std::move(v));
  }
}

// Shapes (kind, k, m, 2j, n_elec, irrep) of every operator of slot cnt, as
// DMRG::allocateTensors creates them (DMRGoperators.cpp:909-1145).
std::vector<OpShape> opset_shapes(const HostState& c, int cnt, bool mr) {
  std::vector<OpShape> out;
  const int L = c.L;
  auto ir = [&](int s) { return c.irreps[s]; };
  const int nL = mr ? cnt + 1 : L - 1 - cnt;
  const int nC = mr ? L - 1 - cnt : cnt + 1;
  for (int k = 0; k < nL; k++) out.push_back({KIND_L, k, 0, 1, 1, ir(mr ? cnt - k : cnt + 1 + k)});
  for (int k = 0; k < nL; k++)
    for (int m = 0; m < nL - k; m++) {
      const int Ip = mr ? (ir(cnt - k - m) ^ ir(cnt - m)) : (ir(cnt + 1 + m) ^ ir(cnt + 1 + k + m));
      out.push_back({KIND_F0, k, m, 0, 0, Ip});
      out.push_back({KIND_F1, k, m, 2, 0, Ip});
      out.push_back({KIND_S0, k, m, 0, 2, Ip});
      if (k > 0) out.push_back({KIND_S1, k, m, 2, 2, Ip});
    }
  for (int k = 0; k < nC; k++)
    for (int m = 0; m < nC - k; m++) {
      const int Id = mr ? (ir(cnt + 1 + k + m) ^ ir(cnt + 1 + m)) : (ir(cnt - k - m) ^ ir(cnt - m));
      out.push_back({KIND_A, k, m, 0, 2, Id});
      if (k > 0) out.push_back({KIND_B, k, m, 2, 2, Id});
      out.push_back({KIND_C, k, m, 0, 0, Id});
      out.push_back({KIND_D, k, m, 2, 0, Id});
    }
  for (int k = 0; k < nC; k++) out.push_back({KIND_Q, k, 0, 1, 1, ir(mr ? cnt + 1 + k : cnt - k)});
  out.push_back({KIND_X, 0, 0, 0, 0, 0});
  return out;
}

void HostState::validate_set(int cnt, bool mr) const {
  const OpSet& s = sets[cnt];
  if (s.dir != (mr ? 1 : 2))
    throw Error{ERR_MISSING, "operator slot " + std::to_string(cnt) + (mr ? " not built moving right" : " not built moving left")};
  for (const OpShape& sh : opset_shapes(*this, cnt, mr)) {
    OpT* t = nullptr;
    auto g1 = [&](const std::vector<OpT*>& v) { return sh.k < (int)v.size() ? v[sh.k] : nullptr; };
    auto g2 = [&](const std::vector<std::vector<OpT*>>& v) {
      return (sh.k < (int)v.size() && sh.m < (int)v[sh.k].size()) ? v[sh.k][sh.m] : nullptr;
    };
    switch (sh.kind) {
      case KIND_L: t = g1(s.L); break;
      case KIND_Q: t = g1(s.Q); break;
      case KIND_X: t = s.X; break;
      case KIND_F0: t = g2(s.F0); break;
      case KIND_F1: t = g2(s.F1); break;
      case KIND_S0: t = g2(s.S0); break;
      case KIND_S1: t = g2(s.S1); break;
      case KIND_A: t = g2(s.A); break;
      case KIND_B: t = g2(s.B); break;
      case KIND_C: t = g2(s.C); break;
      case KIND_D: t = g2(s.D); break;
    }
    if (!t) throw Error{ERR_MISSING, "slot " + std::to_string(cnt) + ": tensor kind " + std::to_string(sh.kind) + " [" +
                                         std::to_string(sh.k) + "][" + std::to_string(sh.m) + "] not uploaded"};
    if (t->two_j != sh.two_j || t->n_elec != sh.n_elec || t->n_irrep != sh.n_irrep)
      throw Error{ERR_ARG, "slot " + std::to_string(cnt) + ": tensor kind " + std::to_string(sh.kind) + " has the wrong symmetry"};
  }
}

}  // namespace chm
