// nhf_pairing.h -- host BCS step of libnhf_hip.so
#pragma once
#include <vector>

#include "nhf_internal.h"

namespace nhf {

struct BcsState {
  std::vector<double> u2, v2, Delta, qp;  // per state (qp = eps v^2)
  double lambda = 0.0, Epair = 0.0;
};

// greedy Kramers matching from O_kj = <psi_k | T psi_j> (M x M column-major)
std::vector<int> match_time_reversal(const std::vector<cd>& O, int M, double threshold = 1e-6);
void pair_representatives(const std::vector<int>& partner, std::vector<int>& uniq, std::vector<int>& s2p,
                          std::vector<char>& vis);
// gap equation on P pairs with the symmetrised pairing matrix Gsym (P x P)
BcsState solve_bcs(const std::vector<double>& eps, const std::vector<double>& Gsym, int A, const nhf_pairing& prm,
                   const std::vector<double>& oldDelta, double oldLambda);

}  // namespace nhf
