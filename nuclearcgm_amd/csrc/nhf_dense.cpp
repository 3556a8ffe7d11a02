// nhf_dense.cpp -- small dense FP64 host algebra of the GCG step: the
// Hermitian Rayleigh-Ritz eigenproblem and the Cholesky-QR coefficients.
//
// eigh_lower: the lower triangle of a k x k Hermitian matrix (the semantics
// of Eigen::SelfAdjointEigenSolver used by the reference, solver.cpp:768) is
// reduced to a real symmetric tridiagonal matrix by complex Householder
// reflections H = I - tau v v^H (the unblocked LAPACK zhetd2 scheme), the
// tridiagonal eigenproblem is solved by the implicit QL method with Wilkinson
// shifts, and the eigenvectors are transformed back.  Eigenvalues ascending.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "nhf_internal.h"

namespace nhf {

bool eigh_lower(int k, const cd* Ain, double* w, cd* V) {
  if (k <= 0) return false;
  std::vector<cd> A((size_t)k * k);
  auto a = [&](int r, int c) -> cd& { return A[(size_t)c * k + r]; };
  for (int c = 0; c < k; ++c)
    for (int r = 0; r < k; ++r) {
      if (r > c)
        a(r, c) = Ain[(size_t)c * k + r];
      else if (r < c)
        a(r, c) = std::conj(Ain[(size_t)r * k + c]);
      else
        a(r, c) = cd(Ain[(size_t)c * k + r].real(), 0.0);
    }
  std::vector<double> d(k, 0.0), e(k, 0.0);
  std::vector<cd> tau(k, 0.0);
  std::vector<std::vector<cd>> vs(k);
  std::vector<cd> x(k);
  for (int i = 0; i < k - 1; ++i) {
    const int m = k - i - 1;  // length of the reflected vector A(i+1:k, i)
    cd alpha = a(i + 1, i);
    double xn2 = 0.0;
    for (int r = i + 2; r < k; ++r) xn2 += std::norm(a(r, i));
    cd ti = 0.0;
    std::vector<cd> v(m, 0.0);
    v[0] = 1.0;
    if (!(xn2 == 0.0 && alpha.imag() == 0.0)) {
      double beta = -std::copysign(std::sqrt(std::norm(alpha) + xn2), alpha.real());
      ti = cd((beta - alpha.real()) / beta, -alpha.imag() / beta);
      cd sc = 1.0 / (alpha - beta);
      for (int r = i + 2; r < k; ++r) v[r - i - 1] = a(r, i) * sc;
      alpha = beta;
    }
    e[i] = alpha.real();
    if (ti != cd(0.0, 0.0)) {
      // x = tau A22 v ; x += (-1/2 tau x^H v) v ; A22 -= v x^H + x v^H
      for (int r = 0; r < m; ++r) {
        cd s = 0.0;
        for (int c = 0; c < m; ++c) s += a(i + 1 + r, i + 1 + c) * v[c];
        x[r] = ti * s;
      }
      cd xv = 0.0;
      for (int r = 0; r < m; ++r) xv += std::conj(x[r]) * v[r];
      cd al2 = -0.5 * ti * xv;
      for (int r = 0; r < m; ++r) x[r] += al2 * v[r];
      for (int c = 0; c < m; ++c) {
        cd xc = std::conj(x[c]), vc = std::conj(v[c]);
        for (int r = 0; r < m; ++r) a(i + 1 + r, i + 1 + c) -= v[r] * xc + x[r] * vc;
      }
    }
    a(i + 1, i + 1) = cd(a(i + 1, i + 1).real(), 0.0);
    d[i] = a(i, i).real();
    tau[i] = ti;
    vs[i] = std::move(v);
  }
  d[k - 1] = a(k - 1, k - 1).real();
  e[k - 1] = 0.0;
  // implicit QL on (d, e), eigenvectors accumulated in Z (column-major)
  std::vector<double> Z((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) Z[(size_t)i * k + i] = 1.0;
  for (int l = 0; l < k; ++l) {
    int iter = 0, mm;
    do {
      for (mm = l; mm < k - 1; ++mm) {
        double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
        if (std::fabs(e[mm]) <= DBL_EPSILON * dd) break;
      }
      if (mm != l) {
        if (iter++ == 60) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + std::copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = mm - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[mm] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          double* zi = &Z[(size_t)i * k];
          double* zi1 = &Z[(size_t)(i + 1) * k];
          for (int q = 0; q < k; ++q) {
            f = zi1[q];
            zi1[q] = s * zi[q] + c * f;
            zi[q] = c * zi[q] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
    } while (mm != l);
  }
  // back-transformation V = H(0) H(1) ... H(k-2) Z, ascending order
  std::vector<int> ord(k);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x1, int x2) { return d[x1] < d[x2]; });
  for (int c = 0; c < k; ++c) {
    w[c] = d[ord[c]];
    cd* col = V + (size_t)c * k;
    const double* z = &Z[(size_t)ord[c] * k];
    for (int r = 0; r < k; ++r) col[r] = z[r];
    for (int i = k - 2; i >= 0; --i) {
      if (tau[i] == cd(0.0, 0.0)) continue;
      const std::vector<cd>& v = vs[i];
      const int m = k - i - 1;
      cd s = 0.0;
      for (int r = 0; r < m; ++r) s += std::conj(v[r]) * col[i + 1 + r];
      s *= tau[i];
      for (int r = 0; r < m; ++r) col[i + 1 + r] -= v[r] * s;
    }
  }
  for (int c = 0; c < k; ++c)
    if (!std::isfinite(w[c])) return false;
  return true;
}

// Cholesky of the kept columns of G (lower triangle used) in column order:
// column j is dropped when the squared norm of its component orthogonal to
// the kept earlier columns is <= drop_rel * G_jj (the column-by-column rule
// of the modified Gram-Schmidt of solver.cpp:389-416 and :756-764, on a Gram
// matrix).  Cinv (k x ke, column-major) = rows `keep` of R^-1.
int cholqr_coeffs(int k, const cd* G, double drop_rel, std::vector<int>& keep, std::vector<cd>& Cinv,
                  double* min_rel) {
  double mr = 1.0;
  auto g = [&](int r, int c) -> cd { return r >= c ? G[(size_t)c * k + r] : std::conj(G[(size_t)r * k + c]); };
  keep.clear();
  std::vector<cd> R;  // column-major ke x ke upper triangular, grown column by column
  std::vector<std::vector<cd>> cols;
  for (int j = 0; j < k; ++j) {
    const int ke = (int)keep.size();
    std::vector<cd> rj(ke + 1, 0.0);
    double s = 0.0;
    for (int i = 0; i < ke; ++i) {
      cd v = g(keep[i], j);
      for (int l = 0; l < i; ++l) v -= std::conj(cols[i][l]) * rj[l];
      v /= cols[i][i];
      rj[i] = v;
      s += std::norm(v);
    }
    double gjj = g(j, j).real();
    double dd = gjj - s;
    if (!(dd > drop_rel * gjj) || !(gjj > 0.0)) continue;
    mr = std::min(mr, dd / gjj);
    rj[ke] = std::sqrt(dd);
    keep.push_back(j);
    cols.push_back(rj);
  }
  const int ke = (int)keep.size();
  if (min_rel) *min_rel = mr;
  // R^-1 by back substitution, column by column
  std::vector<cd> Ri((size_t)ke * ke, 0.0);
  for (int c = 0; c < ke; ++c) {
    Ri[(size_t)c * ke + c] = 1.0 / cols[c][c];
    for (int r = c - 1; r >= 0; --r) {
      cd s = 0.0;
      for (int l = r + 1; l <= c; ++l) s += cols[l][r] * Ri[(size_t)c * ke + l];
      Ri[(size_t)c * ke + r] = -s / cols[r][r];
    }
  }
  Cinv.assign((size_t)k * ke, 0.0);
  for (int c = 0; c < ke; ++c)
    for (int r = 0; r <= c; ++r) Cinv[(size_t)c * k + keep[r]] = Ri[(size_t)c * ke + r];
  return ke;
}

}  // namespace nhf

// ---------------------------------------------------------------- C ABI
extern "C" int nhf_eigh_lower(int k, const double* A, double* w, double* V) {
  if (k < 1 || !A || !w || !V) return NHF_EINVAL;
  return nhf::eigh_lower(k, reinterpret_cast<const nhf::cd*>(A), w, reinterpret_cast<nhf::cd*>(V)) ? NHF_OK
                                                                                                   : NHF_ESOLVER;
}

extern "C" int nhf_cholqr_coeffs(int k, const double* G, double drop_rel, int* keep, double* Cinv) {
  if (k < 1 || !G || !keep || !Cinv) return NHF_EINVAL;
  std::vector<int> kp;
  std::vector<nhf::cd> Ci;
  int ke = nhf::cholqr_coeffs(k, reinterpret_cast<const nhf::cd*>(G), drop_rel, kp, Ci);
  for (int i = 0; i < ke; ++i) keep[i] = kp[i];
  std::copy(Ci.begin(), Ci.end(), reinterpret_cast<nhf::cd*>(Cinv));
  return ke;
}
