// nhf_pairing.cpp -- host side of the BCS pairing step (the P x P pairing
// matrix and the time-reversal overlaps come from device kernels).
//
// Follows BCS::BCSiter / solveBCS (reference src/util/bcs.cpp:65-321):
// Kramers partners are matched greedily by |<psi_k|T psi_j>|^2, each pair
// gets one gap Delta_p; the gap equation is iterated to a fixed point with
// the algebraic update of the chemical potential
//   lambda <- 0.9 (A - sum(1 - eps/E)) / sum(1/E) + 0.1 lambda
// and Delta <- 0.1 Delta + 0.9 sum_q G'_pq w_q Delta_q / E_q, G' = G/2.
#include <algorithm>
#include <cmath>
#include <tuple>

#include "nhf_internal.h"
#include "nhf_pairing.h"

namespace nhf {

std::vector<int> match_time_reversal(const std::vector<cd>& O, int M, double threshold) {
  std::vector<std::tuple<double, int, int>> list;
  for (int k = 0; k < M; ++k)
    for (int j = 0; j < M; ++j) {
      double v = std::norm(O[(size_t)j * M + k]);
      if (v > threshold) list.emplace_back(v, k, j);
    }
  std::stable_sort(list.begin(), list.end(), [](const auto& x, const auto& y) { return std::get<0>(x) > std::get<0>(y); });
  std::vector<int> partner(M, -1);
  std::vector<char> used(M, 0);
  for (const auto& e : list) {
    int k = std::get<1>(e), j = std::get<2>(e);
    if (used[k] || used[j] || k == j) continue;
    partner[k] = j;
    partner[j] = k;
    used[k] = used[j] = 1;
  }
  return partner;
}

void pair_representatives(const std::vector<int>& partner, std::vector<int>& uniq, std::vector<int>& s2p,
                          std::vector<char>& vis) {
  const int M = (int)partner.size();
  uniq.clear();
  s2p.assign(M, 0);
  vis.assign(M, 0);
  for (int i = 0; i < M; ++i) {
    if (vis[i]) continue;
    int pt = partner[i];
    if (pt != i && pt != -1) {
      int pi = (int)uniq.size();
      uniq.push_back(i);
      s2p[i] = pi;
      s2p[pt] = pi;
      vis[i] = vis[pt] = 1;
    }
  }
}

BcsState solve_bcs(const std::vector<double>& eps, const std::vector<double>& Gsym, int A, const nhf_pairing& prm,
                   const std::vector<double>& oldDelta, double oldLambda) {
  const double EPS_SMALL = 1e-12, tol = 1e-10, prec = 1e-10;
  const int P = (int)eps.size();
  BcsState R;
  if (P == 0) return R;
  std::vector<double> G(Gsym.size());
  for (size_t i = 0; i < G.size(); ++i) G[i] = 0.5 * Gsym[i];
  std::vector<double> Delta(P, 1.0);
  if ((int)oldDelta.size() == P && *std::max_element(oldDelta.begin(), oldDelta.end()) > 1e-3) Delta = oldDelta;
  const double ws = prm.window / 10.0;
  double lambda = oldLambda;
  if (!std::isfinite(lambda)) lambda = eps[std::clamp((A / 2) - 1, 0, P - 1)];
  std::vector<double> wv(P), E(P), DN(P);
  auto window = [&](double lam) {
    for (int p = 0; p < P; ++p) {
      wv[p] = 1.0 / (1.0 + std::exp((eps[p] - lam - prm.window) / ws));
      if (prm.window_both) wv[p] *= 1.0 / (1.0 + std::exp(-(eps[p] - lam + prm.window) / ws));
    }
  };
  for (int iter = 0; iter < 2000; ++iter) {
    const double prevLambda = lambda;
    window(lambda);
    double s1 = 0.0, sInv = 0.0;
    for (int p = 0; p < P; ++p) {
      const double xi = eps[p] - lambda;
      E[p] = std::sqrt(xi * xi + Delta[p] * Delta[p] * wv[p]);
      s1 += 1.0 - eps[p] / E[p];
      sInv += 1.0 / E[p];
    }
    const double raw = (sInv < EPS_SMALL) ? prevLambda : ((double)A - s1) / sInv;
    lambda = 0.9 * raw + 0.1 * lambda;
    for (int p = 0; p < P; ++p) {
      double s = 0.0;
      for (int q = 0; q < P; ++q)
        if (wv[q] > EPS_SMALL) s += G[(size_t)q * P + p] * wv[q] * Delta[q] / E[q];
      DN[p] = s;
    }
    double dn = 0.0;
    for (int p = 0; p < P; ++p) {
      const double mixed = 0.1 * Delta[p] + 0.9 * DN[p];
      dn += (mixed - Delta[p]) * (mixed - Delta[p]);
      Delta[p] = mixed;
    }
    if (std::abs(raw - prevLambda) < prec && std::sqrt(dn) < tol) break;
  }
  window(lambda);
  R.u2.resize(P);
  R.v2.resize(P);
  for (int i = 0; i < P; ++i) {
    const double xi = eps[i] - lambda;
    E[i] = std::sqrt(xi * xi + Delta[i] * Delta[i] * wv[i]);
    const double v2 = std::clamp(0.5 * (1 - xi / E[i]), 0.0, 1.0);
    R.v2[i] = v2;
    R.u2[i] = 1.0 - v2;
  }
  double Ep = 0.0;
  for (int p = 0; p < P; ++p)
    for (int q = 0; q < P; ++q)
      if (wv[q] > EPS_SMALL) Ep += -0.5 * wv[q] * wv[p] * Delta[p] * Delta[q] * G[(size_t)q * P + p] / E[p] / E[q];
  R.Delta = Delta;
  R.lambda = lambda;
  R.Epair = Ep;
  return R;
}

}  // namespace nhf
