// Host-side tables the grid kernels consume, computed once per grid in FP64.
#include "grid_tables.h"

#include <cmath>
#include <cstdlib>
#include <utility>

namespace dg {

#include "lebedev_table.inc"

void radial_grid(int n, std::vector<double>& r, std::vector<double>& w) {
  // Node p of N: x = cos(p pi/(N+1)), r = (1+x)/(1-x); the Chebyshev weight
  // (pi/(N+1)) sin^2 is divided by sqrt(1-x^2) and multiplied by dr/dx = 2/(1-x)^2.
  // Written with the same library calls, in the same order, as
  // src/atomicgrid.cpp:46-63 so that the nodes agree to the last bit.
  r.assign(n, 0.0);
  w.assign(n, 0.0);
  const double f = kPi / (double)(n + 1);
  for (int p = 1; p <= n; ++p) {
    double wp = f * std::pow(std::sin(f * (double)p), 2.0);
    const double x = std::cos(f * (double)p);
    r[p - 1] = (1.0 + x) / (1.0 - x);
    wp = wp / std::sqrt(1.0 - std::pow(x, 2.0)) * 2.0 / std::pow(1.0 - x, 2.0);
    w[p - 1] = wp;
  }
}

const double* lebedev_rule(int npoints) {
  int off = 0;
  for (int i = 0; i < kLebedevNumOrders; ++i) {
    if (kLebedevCounts[i] == npoints) return &kLebedevTable[off][0];
    off += kLebedevCounts[i];
  }
  return nullptr;
}

static double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

double sh_prefactor(int l, int m) {
  const int am = std::abs(m);
  const double base = 1.0 / std::sqrt(4 * kPi);
  return base * (m == 0 ? 1.0 : std::sqrt(2.0)) *
         std::sqrt((double)(2 * l + 1) * factorial(l - am) / factorial(l + am));
}

// Associated Legendre P_l^|m|(x) with the Condon-Shortley phase, by the upward
// recurrence in l from P_m^m = (-1)^m (2m-1)!! (1-x^2)^{m/2}.
static double legendre(int l, int m, double x) {
  double pmm = 1.0, fact = 1.0;
  const double s = std::sqrt(1.0 - x * x);
  for (int k = 0; k < m; ++k) {
    pmm *= -fact * s;
    fact += 2.0;
  }
  if (l == m) return pmm;
  double p1 = x * (double)(2 * m + 1) * pmm;  // P_{m+1}^m
  if (l == m + 1) return p1;
  double p0 = pmm;
  for (int j = m + 2; j <= l; ++j) {
    const double pj = ((double)(2 * j - 1) * x * p1 - (double)(j + m - 1) * p0) / (double)(j - m);
    p0 = p1;
    p1 = pj;
  }
  return p1;
}

void ylm_table(int nang, const double* leb, int lmax, std::vector<double>& y) {
  const int nlm = (lmax + 1) * (lmax + 1);
  y.assign((size_t)nang * nlm, 0.0);
  for (int j = 0; j < nang; ++j) {
    const double x = leb[4 * j], yy = leb[4 * j + 1], z = leb[4 * j + 2];
    const double r = std::sqrt(x * x + yy * yy + z * z);
    const double phi = std::atan2(yy, x);
    const double ct = std::cos(std::acos(z / r));
    for (int l = 0; l <= lmax; ++l)
      for (int m = -l; m <= l; ++m) {
        const double ang = m == 0 ? 1.0 : (m > 0 ? std::cos(m * phi) : std::sin(-m * phi));
        y[(size_t)j * nlm + l * l + l + m] = sh_prefactor(l, m) * legendre(l, std::abs(m), ct) * ang;
      }
  }
}

void poisson_matrix(int n, const double* r, int l, std::vector<double>& a) {
  // Radial Poisson equation U'' - l(l+1) U / r^2 = -4 pi r rho_lm on the
  // uniform grid z = acos((r-1)/(r+1))/pi, h = 1/(N+1), node i <-> r_{i-1};
  // (dz/dr)^2 = 1/(pi^2 r (1+r)^2), d2z/dr2 = (1+3r)/(2 pi r^{3/2} (1+r)^2).
  // Interior rows: 7-point 6th-order stencils; rows 1, 2, N-1, N: one-sided
  // 5/6-point stencils; rows 0 and N+1: Dirichlet (src/atomicgrid.cpp:316-389).
  const int N = n, M = n + 2;
  const double h = 1.0 / (double)(N + 1);
  a.assign((size_t)M * M, 0.0);
  auto at = [&](int i, int j) -> double& { return a[(size_t)i * M + j]; };
  at(0, 0) = 1.0;
  at(N + 1, N + 1) = 1.0;
  for (int i = 1; i <= N; ++i) {
    const double rr = r[i - 1];
    double c1 = 1.0 / (kPi * kPi * rr * (1.0 + rr) * (1.0 + rr));
    double c2 = (1.0 + 3.0 * rr) /
                (2.0 * kPi * std::pow(rr / ((1.0 + rr) * (1.0 + rr)), 1.5) * std::pow(1.0 + rr, 5.0));
    // row kind, tested in the reference's order: 1, 2, N-1, N, interior
    const int kind = (i == 1) ? 1 : (i == 2) ? 2 : (i == N - 1) ? 3 : (i == N) ? 4 : 0;
    if (kind == 1 || kind == 4) {
      c1 /= 12.0 * h * h;
      c2 /= 12.0 * h;
    } else if (kind == 2 || kind == 3) {
      c1 /= 12.0 * h * h;
      c2 /= 60.0 * h;
    } else {
      c1 /= 180.0 * h * h;
      c2 /= 60.0 * h;
    }
    if (kind == 1) {
      const double s1[5] = {11, -20, 6, 4, -1}, s2[5] = {-3, -10, 18, -6, 1};
      for (int k = 0; k < 5; ++k) at(i, k) = s1[k] * c1 + s2[k] * c2;
    } else if (kind == 2) {
      const double s1[6] = {-1, 16, -30, 16, -1, 0}, s2[6] = {3, -30, -20, 60, -15, 2};
      for (int k = 0; k < 6; ++k) at(i, k) = s1[k] * c1 + s2[k] * c2;
    } else if (kind == 3) {
      const double s1[6] = {0, -1, 16, -30, 16, -1}, s2[6] = {-2, 15, -60, 20, 30, -3};
      for (int k = 0; k < 6; ++k) at(i, N - 4 + k) = s1[k] * c1 + s2[k] * c2;
    } else if (kind == 4) {
      const double s1[5] = {-1, 4, 6, -20, 11}, s2[5] = {-1, 6, -18, 10, 3};
      for (int k = 0; k < 5; ++k) at(i, N - 3 + k) = s1[k] * c1 + s2[k] * c2;
    } else {
      const double s1[7] = {2, -27, 270, -490, 270, -27, 2}, s2[7] = {-1, 9, -45, 0, 45, -9, 1};
      for (int k = 0; k < 7; ++k) at(i, i - 3 + k) = s1[k] * c1 + s2[k] * c2;
    }
    at(i, i) -= (double)l * (double)(l + 1) / (rr * rr);
  }
}

bool invert_matrix(int n, std::vector<double>& a) {
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A(i, k)) > std::fabs(A(p, k))) p = i;
    if (A(p, k) == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
      std::swap(perm[k], perm[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      A(i, k) /= A(k, k);
      const double f = A(i, k);
      if (f != 0.0)
        for (int j = k + 1; j < n; ++j) A(i, j) -= f * A(k, j);
    }
  }
  // solve LU x = e_perm for every unit vector
  std::vector<double> inv((size_t)n * n, 0.0), col(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (perm[i] == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i) {
      double s = col[i];
      for (int j = 0; j < i; ++j) s -= A(i, j) * col[j];
      col[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = col[i];
      for (int j = i + 1; j < n; ++j) s -= A(i, j) * col[j];
      col[i] = s / A(i, i);
    }
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = col[i];
  }
  a.swap(inv);
  return true;
}

static double double_factorial(int n) {
  double f = 1.0;
  for (int i = n; i > 1; i -= 2) f *= i;
  return f;
}

double gto_norm(double alpha, int l, int m, int n) {
  const int L = l + m + n;
  const double num = std::pow(2.0, 2.0 * L + 1.5) * std::pow(alpha, L + 1.5);
  const double den = double_factorial(2 * l - 1) * double_factorial(2 * m - 1) *
                     double_factorial(2 * n - 1) * std::pow(kPiTruncated, 1.5);
  return std::sqrt(num / den);
}

SplineSystem spline_system(const std::vector<double>& x) {
  // Rows of the slope system (D_i = spline derivative at knot i): row 0 and
  // row n-1 carry the not-a-knot conditions, rows 1..n-2 the continuity of the
  // second derivative (src/cspline.cpp:86-109).  The sweep of :112-121 only
  // needs a_i, the pivots den_i and the reduced super-diagonal cmod_i.
  const size_t n = x.size();
  SplineSystem s;
  s.a.assign(n, 0.0);
  s.den.assign(n, 0.0);
  s.cmod.assign(n, 0.0);
  std::vector<double> b(n, 0.0), c(n, 0.0);
  double h0 = x[1] - x[0], h1 = x[2] - x[1];
  b[0] = h1 * (h0 + h1);
  c[0] = (h0 + h1) * (h0 + h1);
  for (size_t i = 1; i + 1 < n; ++i) {
    h0 = x[i] - x[i - 1];
    h1 = x[i + 1] - x[i];
    s.a[i] = h1;
    b[i] = 2 * (h0 + h1);
    c[i] = h0;
  }
  s.a[n - 1] = (h0 + h1) * (h0 + h1);
  b[n - 1] = h0 * (h0 + h1);
  s.den[0] = b[0];
  s.cmod[0] = c[0] / b[0];
  for (size_t i = 1; i + 1 < n; ++i) {
    s.den[i] = b[i] - s.a[i] * s.cmod[i - 1];
    s.cmod[i] = c[i] / s.den[i];
  }
  s.den[n - 1] = b[n - 1] - s.a[n - 1] * s.cmod[n - 2];
  return s;
}

std::vector<double> spline_slope_matrix(const std::vector<double>& x) {
  const size_t n = x.size();
  const SplineSystem sys = spline_system(x);
  std::vector<double> S(n * n, 0.0), Y(n), yp(n), D(n);
  for (size_t c = 0; c < n; ++c) {
    // right-hand side of the slope system for y = e_c
    auto y = [&](size_t i) { return i == c ? 1.0 : 0.0; };
    double h0 = x[1] - x[0], h1 = x[2] - x[1];
    double r0 = (y(1) - y(0)) / h0, r1 = (y(2) - y(1)) / h1;
    Y[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0;
    for (size_t i = 1; i + 1 < n; ++i) {
      h0 = x[i] - x[i - 1];
      h1 = x[i + 1] - x[i];
      r0 = (y(i) - y(i - 1)) / h0;
      r1 = (y(i + 1) - y(i)) / h1;
      Y[i] = 3 * (r0 * h1 + r1 * h0);
    }
    Y[n - 1] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0);
    yp[0] = Y[0] / sys.den[0];
    for (size_t i = 1; i < n; ++i) yp[i] = (Y[i] - sys.a[i] * yp[i - 1]) / sys.den[i];
    D[n - 1] = yp[n - 1];
    for (size_t i = n - 1; i > 0; --i) D[i - 1] = yp[i - 1] - sys.cmod[i - 1] * D[i];
    for (size_t r = 0; r < n; ++r) S[r * n + c] = D[r];
  }
  return S;
}

}  // namespace dg
