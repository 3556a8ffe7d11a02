// nhf_guess.cpp -- start-up states and fields on the host: the
// harmonic-oscillator spinor guess and the deformed Woods-Saxon fields.
//
// Spin-orbit-coupled |n_r l j m_j> states ordered by major shell N, then
// n_r, then j = l + 1/2 before j = l - 1/2, |m_j| ascending with +m_j before
// -m_j (the enumeration of reference guess.cpp:126-163), radial functions
// R_{n_r l}(r) = sqrt(2 n_r! / (b^3 Gamma(n_r + l + 3/2))) (r/b)^l
// L_{n_r}^{l+1/2}((r/b)^2) exp(-(r/b)^2 / 2), Clebsch-Gordan coupling with
// the Condon-Shortley spherical harmonics, Euclidean-normalised columns.
#include <cmath>
#include <complex>
#include <tuple>
#include <vector>

#include "../../include/nhf.h"

namespace {
using cd = std::complex<double>;
const double PI = 3.14159265358979323846;

double gen_laguerre(int n, double alpha, double x) {
  if (n == 0) return 1.0;
  double lm2 = 1.0, lm1 = 1.0 + alpha - x;
  for (int k = 2; k <= n; ++k) {
    double l = ((2.0 * k - 1.0 + alpha - x) * lm1 - (k - 1.0 + alpha) * lm2) / k;
    lm2 = lm1;
    lm1 = l;
  }
  return lm1;
}
// associated Legendre P_l^m(x), m >= 0, without the Condon-Shortley phase
double assoc_legendre(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    double df = 1.0;
    for (int i = 1; i <= 2 * m - 1; i += 2) df *= i;
    pmm = df * std::pow(1.0 - x * x, 0.5 * m);
  }
  if (l == m) return pmm;
  double pm1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pm1;
  double pl = 0.0;
  for (int ll = m + 2; ll <= l; ++ll) {
    pl = ((2 * ll - 1) * x * pm1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pm1;
    pm1 = pl;
  }
  return pl;
}
cd sph_y(int l, int m, double th, double ph) {
  const int am = std::abs(m);
  double f = 1.0;
  for (int i = l - am + 1; i <= l + am; ++i) f *= i;  // (l+|m|)! / (l-|m|)!
  const double nrm = std::sqrt((2.0 * l + 1.0) / (4.0 * PI) / f);
  const double sgn = (am % 2) ? -1.0 : 1.0;
  cd y = sgn * nrm * assoc_legendre(l, am, std::cos(th)) * std::polar(1.0, am * ph);
  if (m < 0) return sgn * std::conj(y);
  return y;
}
}  // namespace

extern "C" int nhf_ho_guess(int n, double box_size, int nev, double b, double* out) {
  if (n < 2 || nev < 1 || !(b > 0) || !out) return NHF_EINVAL;
  const double a = 0.5 * box_size, h = 2.0 * a / (n - 1);
  const size_t N3 = (size_t)n * n * n, rows = 2 * N3;
  std::vector<std::tuple<int, int, int, int>> orb;  // n_r, l, 2j, 2m_j
  for (int N = 0; (int)orb.size() < nev; ++N)
    for (int nr = 0; nr <= N / 2 && (int)orb.size() < nev; ++nr) {
      const int l = N - 2 * nr;
      for (int j2 : {2 * l + 1, 2 * l - 1}) {
        if (j2 < 0) continue;
        for (int m2 = 1; m2 <= j2 && (int)orb.size() < nev; m2 += 2) {
          orb.emplace_back(nr, l, j2, m2);
          if ((int)orb.size() < nev) orb.emplace_back(nr, l, j2, -m2);
        }
      }
    }
  cd* G = reinterpret_cast<cd*>(out);
  for (int s = 0; s < nev; ++s) {
    int nr, l, j2, m2;
    std::tie(nr, l, j2, m2) = orb[s];
    const double mj = 0.5 * m2, den = 2.0 * l + 1.0;
    double cu, cdn;
    if (j2 == 2 * l + 1) {
      cu = std::sqrt((l + mj + 0.5) / den);
      cdn = std::sqrt((l - mj + 0.5) / den);
    } else {
      cu = std::sqrt((l - mj + 0.5) / den);
      cdn = -std::sqrt((l + mj - 0.5) / den);
    }
    const int mlu = (m2 - 1) / 2, mld = (m2 + 1) / 2;
    const bool hu = std::abs(mlu) <= l, hd = std::abs(mld) <= l;
    const double rn = std::sqrt(2.0 * std::tgamma(nr + 1.0) / (b * b * b * std::tgamma(nr + l + 1.5)));
    cd* col = G + (size_t)s * rows;
    double nsq = 0.0;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double x = -a + i * h, y = -a + j * h, z = -a + k * h;
          const double r = std::sqrt(x * x + y * y + z * z);
          const double th = r > 1e-12 ? std::acos(z / r) : 0.0, ph = std::atan2(y, x);
          const double rho = r / b;
          const double R = rn * std::pow(rho, l) * gen_laguerre(nr, l + 0.5, rho * rho) * std::exp(-0.5 * rho * rho);
          const cd up = hu ? cu * R * sph_y(l, mlu, th, ph) : cd(0.0, 0.0);
          const cd dn = hd ? cdn * R * sph_y(l, mld, th, ph) : cd(0.0, 0.0);
          const size_t p = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
          col[2 * p] = up;
          col[2 * p + 1] = dn;
          nsq += std::norm(up) + std::norm(dn);
        }
    const double inv = 1.0 / std::sqrt(nsq);
    for (size_t r = 0; r < rows; ++r) col[r] *= inv;
  }
  return NHF_OK;
}

// Deformed Woods-Saxon start-up fields.  Central: -V0 (1 + kappa (A-2Z)/A)
// / (1 + exp((r - R(theta) - beta3 R0 Y30(theta)) / a)) with
// R(theta) = R0 (1 + beta2 Y20(theta)), R0 = r0 A^(1/3).  Spin-orbit:
// W = ls g, g = V0_so / a * e^t / (1 + e^t)^2 * (r_hat - grad R) with
// t = (r - R(theta)) / a (no beta3 shift), ls = -35 (hbar / 2m)^2, and the
// reference's grad R = R0 beta2 (6 sqrt(5/16pi) x z^2 / r^3, 0, 0).
// e^t / (1 + e^t)^2 is evaluated as e^-|t| / (1 + e^-|t|)^2 (the function
// is even in t), which cannot overflow.
extern "C" int nhf_ws_fields(int n, double box_size, int A, int Z, double beta2, double beta3, int spin_orbit,
                             const nhf_ws_params* p, double* f, double* gradf, double* U, double* W) {
  static const nhf_ws_params file_values = {55.5, 1.20, 0.67, 0.86, 55.5};
  if (n < 2 || !(box_size > 0) || A < 1 || Z < 0 || Z > A || !f || !gradf || !U || !W) return NHF_EINVAL;
  const nhf_ws_params& P = p ? *p : file_values;
  if (!(P.diff > 0) || !(P.r0 > 0)) return NHF_EINVAL;
  const double hbar = 197.327053, mnuc = 938.91869;
  const double a = 0.5 * box_size, h = 2.0 * a / (n - 1);
  const size_t N3 = (size_t)n * n * n;
  const double R0 = P.r0 * std::pow((double)A, 1.0 / 3.0);
  const double depth = P.V0 * (1.0 + P.kappa * (A - 2 * Z) / (double)A);
  const double c20 = std::sqrt(5.0 / (16.0 * PI)), c30 = std::sqrt(7.0 / (4.0 * PI));
  const double hb2m = hbar * hbar / (2.0 * mnuc);
  const double ls = -35.0 * (0.5 * hbar / mnuc) * (0.5 * hbar / mnuc);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t q = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
        const double x = -a + i * h, y = -a + j * h, z = -a + k * h;
        const double r = std::sqrt(x * x + y * y + z * z);
        f[q] = hb2m;
        gradf[q] = gradf[N3 + q] = gradf[2 * N3 + q] = 0.0;
        W[q] = W[N3 + q] = W[2 * N3 + q] = 0.0;
        if (r == 0.0) {  // the centre (odd meshes only): the spherical limit
          U[q] = -depth / (1.0 + std::exp(-R0 / P.diff));
          continue;
        }
        const double ct = z / r;
        const double Rt = R0 * (1.0 + beta2 * c20 * (3.0 * ct * ct - 1.0));
        const double y30 = c30 * 0.5 * ct * (5.0 * ct * ct - 3.0);
        U[q] = -depth / (1.0 + std::exp((r - Rt - beta3 * R0 * y30) / P.diff));
        if (!spin_orbit) continue;
        const double e = std::exp(-std::fabs((r - Rt) / P.diff));
        const double g = ls * P.V0_so / P.diff * e / ((1.0 + e) * (1.0 + e));
        const double dRx = R0 * beta2 * 6.0 * c20 * x * z * z / (r * r * r);
        W[q] = g * (x / r - dRx);
        W[N3 + q] = g * (y / r);
        W[2 * N3 + q] = g * (z / r);
      }
  return NHF_OK;
}
