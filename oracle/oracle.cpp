// FP64 CPU restatement of the reference's grid-integration path (the oracle).
//
// TEST INFRASTRUCTURE ONLY: tests, the smoke check and bench.py's CPU-baseline
// leg use it as the checker; the product library never links or calls it.
//
// Each routine restates the reference routine cited next to it (file:line of
// the reference source tree), keeping loop structure, evaluation order and
// OpenMP placement.  Deliberate differences, all at rounding level:
//  * Eigen's blocked PartialPivLU and its GEMV/dot kernels are replaced by
//    plain unblocked loops (same pivot rule);
//  * Eigen's SelfAdjointEigenSolver is replaced by cyclic Jacobi (the density
//    matrix is invariant to the eigenvector basis of each eigenvalue);
//  * the Numerical-Recipes gammln inside the Boys function is std::lgamma;
//  * OpenMP reductions are done serially in index order (deterministic).
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;       // src/atomicgrid.h:33, functionals.h:49
const double kPiTrunc = 3.14159265359;           // src/cgf.cpp:103, integrals.cpp:342,394

struct Vec3 {
  double x, y, z;
};
inline Vec3 sub(const Vec3& a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double sqnorm(const Vec3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double norm(const Vec3& a) { return std::sqrt(sqnorm(a)); }

double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}
double double_factorial(int n) {  // n!! for n >= -1
  double f = 1.0;
  for (int i = n; i > 1; i -= 2) f *= i;
  return f;
}

// ---------------------------------------------------------------------------
// Basis: GTO / CGF (src/cgf.cpp)
// ---------------------------------------------------------------------------
struct GTO {
  double c, alpha, norm;
  int l, m, n;
  Vec3 pos;
};
struct CGF {
  std::vector<GTO> gtos;
  Vec3 pos;
};

// src/cgf.cpp:102-114 (truncated pi)
double gto_norm(double alpha, int l, int m, int n) {
  double nom = std::pow(2.0, 2.0 * (l + m + n) + 3.0 / 2.0) *
               std::pow(alpha, (l + m + n) + 3.0 / 2.0);
  double denom = (l < 1 ? 1 : double_factorial(2 * l - 1)) *
                 (m < 1 ? 1 : double_factorial(2 * m - 1)) *
                 (n < 1 ? 1 : double_factorial(2 * n - 1)) * std::pow(kPiTrunc, 3.0 / 2.0);
  return std::sqrt(nom / denom);
}

// src/cgf.cpp:49-57
double gto_amp(const GTO& g, const Vec3& r) {
  double r2 = sqnorm(sub(r, g.pos));
  return g.norm * std::pow(r.x - g.pos.x, g.l) * std::pow(r.y - g.pos.y, g.m) *
         std::pow(r.z - g.pos.z, g.n) * std::exp(-g.alpha * r2);
}

// src/cgf.cpp:146-154
double cgf_amp(const CGF& f, const Vec3& r) {
  double sum = 0.0;
  for (const auto& g : f.gtos) sum += g.c * gto_amp(g, r);
  return sum;
}

// Shell expansion of src/molecule.cpp:176-218 with the component order of
// src/cgf.h:208-221.
const int kCart[3][6][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};

std::vector<CGF> build_cgfs(const double* xyz, int nshell, const int* shell_atom, const int* shell_L,
                            const int* shell_nprim, const int* shell_offset, const double* alpha,
                            const double* coef) {
  std::vector<CGF> out;
  for (int s = 0; s < nshell; ++s) {
    int L = shell_L[s];
    if (L < 0 || L > 2) throw std::runtime_error("Undefined orbital type");
    int a = shell_atom[s];
    Vec3 p{xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    int ncomp = (L + 1) * (L + 2) / 2;
    for (int c = 0; c < ncomp; ++c) {
      CGF f;
      f.pos = p;
      for (int k = 0; k < shell_nprim[s]; ++k) {
        GTO g;
        g.c = coef[shell_offset[s] + k];
        g.alpha = alpha[shell_offset[s] + k];
        g.l = kCart[L][c][0];
        g.m = kCart[L][c][1];
        g.n = kCart[L][c][2];
        g.norm = gto_norm(g.alpha, g.l, g.m, g.n);
        g.pos = p;
        f.gtos.push_back(g);
      }
      out.push_back(f);
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Spherical harmonics (src/spherical_harmonics.cpp:24-117)
// ---------------------------------------------------------------------------
double legendre_p(int n, int m, double x) {  // :81-117
  std::vector<double> v(n + 1, 0.0);
  if (m <= n) {
    v[m] = 1.0;
    double fact = 1.0;
    for (int k = 0; k < m; k++) {
      v[m] *= -fact * std::sqrt(1.0 - x * x);
      fact += 2.0;
    }
  }
  if (m + 1 <= n) v[m + 1] = x * (double)(2 * m + 1) * v[m];
  for (int j = m + 2; j <= n; j++) {
    v[j] = ((double)(2 * j - 1) * x * v[j - 1] + (double)(-j - m + 1) * v[j - 2]) /
           (double)(j - m);
  }
  return v[n];
}
double sh_prefactor(int l, int m) {  // :28-33
  static const double pre = 1.0 / std::sqrt(4 * kPi);
  int am = std::abs(m);
  return pre * (m == 0 ? 1 : std::sqrt(2.0)) *
         std::sqrt((double)(2 * l + 1) * factorial(l - am) / factorial(l + am));
}
double azimuthal(int m, double phi) {  // :39-47
  if (m == 0) return 1.0;
  if (m > 0) return std::cos((double)m * phi);
  return std::sin(-(double)m * phi);
}
double spherical_harmonic(int l, int m, double pole, double azimuth) {  // :24-26, 35-37
  return legendre_p(l, std::abs(m), std::cos(pole)) * azimuthal(m, azimuth);
}

// ---------------------------------------------------------------------------
// Cubic spline (src/cspline.cpp:36-172)
// ---------------------------------------------------------------------------
struct Spline {
  std::vector<double> x, y;
  std::vector<double> c;  // (n-1) x 4
};
void spline_generate(Spline& s) {  // :66-142
  const std::vector<double>& x = s.x;
  const std::vector<double>& y = s.y;
  size_t n = x.size();
  if (n != y.size()) throw std::runtime_error("Cspline x and y data ranges do not match");
  if (n < 3) throw std::runtime_error("Cspline data range has to contain 3 of more items");
  for (size_t i = 1; i < n; ++i)
    if (x[i] <= x[i - 1]) throw std::runtime_error("Spline x-data should be continuously increasing");
  std::vector<double> A(n, 0.0), B(n, 0.0), C(n, 0.0), D(n, 0.0), Y(n, 0.0);
  double h0 = x[1] - x[0], h1 = x[2] - x[1];
  double r0 = (y[1] - y[0]) / h0, r1 = (y[2] - y[1]) / h1;
  B[0] = h1 * (h0 + h1);
  C[0] = (h0 + h1) * (h0 + h1);
  Y[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0;
  for (size_t i = 1; i < n - 1; ++i) {
    h0 = x[i] - x[i - 1];
    h1 = x[i + 1] - x[i];
    r0 = (y[i] - y[i - 1]) / h0;
    r1 = (y[i + 1] - y[i]) / h1;
    A[i] = h1;
    B[i] = 2 * (h0 + h1);
    C[i] = h0;
    Y[i] = 3 * (r0 * h1 + r1 * h0);
  }
  A[n - 1] = (h0 + h1) * (h0 + h1);
  B[n - 1] = h0 * (h0 + h1);
  Y[n - 1] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0);
  C[0] = C[0] / B[0];
  for (size_t i = 1; i < n - 1; ++i) C[i] = C[i] / (B[i] - A[i] * C[i - 1]);
  Y[0] = Y[0] / B[0];
  for (size_t i = 1; i < n; ++i) Y[i] = (Y[i] - A[i] * Y[i - 1]) / (B[i] - A[i] * C[i - 1]);
  D[n - 1] = Y[n - 1];
  for (size_t i = n - 1; i > 0; i--) D[i - 1] = Y[i - 1] - C[i - 1] * D[i];
  s.c.assign(4 * (n - 1), 0.0);
  for (size_t i = 0; i < n - 1; ++i) {
    double dx = 1.0 / (x[i + 1] - x[i]);
    double dy = (y[i + 1] - y[i]) * dx;
    s.c[4 * i + 0] = y[i];
    s.c[4 * i + 1] = D[i];
    s.c[4 * i + 2] = dx * (3 * dy - 2 * D[i] - D[i + 1]);
    s.c[4 * i + 3] = dx * dx * (-2 * dy + D[i] + D[i + 1]);
  }
}
double spline_eval(const double* x, const double* y, const double* c, size_t n, double xx) {  // :151-172
  if (xx < x[0]) return y[0];
  if (xx >= x[n - 1]) return y[n - 1];
  for (size_t i = 1; i < n; ++i) {
    if (xx <= x[i]) {
      double xpar = (xx - x[i - 1]);
      const double* s = c + 4 * (i - 1);
      return s[0] + s[1] * xpar + s[2] * xpar * xpar + s[3] * xpar * xpar * xpar;
    }
  }
  return 0.0;
}

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting (stands in for Eigen::PartialPivLU,
// src/atomicgrid.cpp:424-425).  A column-major n x n, overwritten.
// ---------------------------------------------------------------------------
void lu_factor(int n, std::vector<double>& A, std::vector<int>& perm) {
  perm.resize(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[k + (size_t)k * n]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(A[i + (size_t)k * n]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k + (size_t)j * n], A[p + (size_t)j * n]);
      std::swap(perm[k], perm[p]);
    }
    double piv = A[k + (size_t)k * n];
    if (piv != 0.0)
      for (int i = k + 1; i < n; ++i) A[i + (size_t)k * n] /= piv;
    for (int j = k + 1; j < n; ++j) {
      double akj = A[k + (size_t)j * n];
      if (akj == 0.0) continue;
      for (int i = k + 1; i < n; ++i) A[i + (size_t)j * n] -= A[i + (size_t)k * n] * akj;
    }
  }
}
void lu_solve(int n, const std::vector<double>& LU, const std::vector<int>& perm, const double* b,
              double* x) {
  std::vector<double> y(n);
  for (int i = 0; i < n; ++i) y[i] = b[perm[i]];
  for (int i = 0; i < n; ++i) {
    double s = y[i];
    for (int j = 0; j < i; ++j) s -= LU[i + (size_t)j * n] * y[j];
    y[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int j = i + 1; j < n; ++j) s -= LU[i + (size_t)j * n] * y[j];
    y[i] = s / LU[i + (size_t)i * n];
  }
  for (int i = 0; i < n; ++i) x[i] = y[i];
}

// src/atomicgrid.cpp:502-510 (m = 1)
double d2zdr2(double r, double m) {
  double nom = m * m * (m + 3.0 * r);
  double denom = 2.0 * kPi * std::pow((m * r) / ((m + r) * (m + r)), 1.5) * std::pow(m + r, 5.0);
  return nom / denom;
}
double dzdrsq(double r, double m) { return m / (kPi * kPi * r * (m + r) * (m + r)); }

// FD matrix of src/atomicgrid.cpp:316-389, column-major (N+2)^2
void build_fd_matrix(int N, const double* r_n, std::vector<double>& A) {
  const int n = N + 2;
  const double h = 1.0 / (double)(N + 1);
  A.assign((size_t)n * n, 0.0);
  auto at = [&](int i, int j) -> double& { return A[i + (size_t)j * n]; };
  for (int i = 0; i < n; i++) {
    double c1 = 0.0, c2 = 0.0;
    if (i > 0 && i < N + 1) {
      c1 = dzdrsq(r_n[i - 1], 1.0);
      c2 = d2zdr2(r_n[i - 1], 1.0);
    }
    if (i == 0) {
      at(0, 0) = 1.0;
      continue;
    }
    if (i == 1) {
      c1 /= 12.0 * h * h;
      c2 /= 12.0 * h;
      at(i, 0) = 11.0 * c1 - 3.0 * c2;
      at(i, 1) = -20.0 * c1 - 10.0 * c2;
      at(i, 2) = 6.0 * c1 + 18.0 * c2;
      at(i, 3) = 4.0 * c1 - 6.0 * c2;
      at(i, 4) = -1.0 * c1 + 1.0 * c2;
      continue;
    }
    if (i == 2) {
      c1 /= 12.0 * h * h;
      c2 /= 60.0 * h;
      at(i, 0) = -1.0 * c1 + 3.0 * c2;
      at(i, 1) = 16.0 * c1 - 30.0 * c2;
      at(i, 2) = -30.0 * c1 - 20.0 * c2;
      at(i, 3) = 16.0 * c1 + 60.0 * c2;
      at(i, 4) = -1.0 * c1 - 15.0 * c2;
      at(i, 5) = 0.0 * c1 + 2.0 * c2;
      continue;
    }
    if (i == N - 1) {
      c1 /= 12.0 * h * h;
      c2 /= 60.0 * h;
      at(i, N - 4) = 0.0 * c1 - 2.0 * c2;
      at(i, N - 3) = -1.0 * c1 + 15.0 * c2;
      at(i, N - 2) = 16.0 * c1 - 60.0 * c2;
      at(i, N - 1) = -30.0 * c1 + 20.0 * c2;
      at(i, N) = 16.0 * c1 + 30.0 * c2;
      at(i, N + 1) = -1.0 * c1 - 3.0 * c2;
      continue;
    }
    if (i == N) {
      c1 /= 12.0 * h * h;
      c2 /= 12.0 * h;
      at(i, N - 3) = -1.0 * c1 - 1.0 * c2;
      at(i, N - 2) = 4.0 * c1 + 6.0 * c2;
      at(i, N - 1) = 6.0 * c1 - 18.0 * c2;
      at(i, N) = -20.0 * c1 + 10.0 * c2;
      at(i, N + 1) = 11.0 * c1 + 3.0 * c2;
      continue;
    }
    if (i == N + 1) {
      at(i, i) = 1.0;
      continue;
    }
    c1 /= 180.0 * h * h;
    c2 /= 60.0 * h;
    at(i, i - 3) = 2.0 * c1 - 1.0 * c2;
    at(i, i - 2) = -27.0 * c1 + 9.0 * c2;
    at(i, i - 1) = 270.0 * c1 - 45.0 * c2;
    at(i, i) = -490.0 * c1;
    at(i, i + 1) = 270.0 * c1 + 45.0 * c2;
    at(i, i + 2) = -27.0 * c1 - 9.0 * c2;
    at(i, i + 3) = 2.0 * c1 + 1.0 * c2;
  }
}

// ---------------------------------------------------------------------------
// LDA (src/functionals.cpp)
// ---------------------------------------------------------------------------
double vwn_xx(double x, double b, double c) { return x * x + b * x + c; }  // :116-118
double vwn_eps(double x, double a, double x0, double b, double c) {        // :128-134
  double q = std::sqrt(4.0 * c - b * b);
  double eps = a * (std::log(x * x / vwn_xx(x, b, c)) -
                    b * (x0 / vwn_xx(x0, b, c)) * std::log(std::pow(x - x0, 2.0) / vwn_xx(x, b, c)) +
                    (2.0 * b / q) * (1.0 - (x0 * (2.0 * x0 + b) / vwn_xx(x0, b, c))) *
                        std::atan(q / (2.0 * x + b)));
  return eps;
}
double vwn_deps(double x, double a, double x0, double b, double c) {  // :144-150
  double q = std::sqrt(4.0 * c - b * b);
  double deps = a * (2.0 / x - (2.0 * x + b) / vwn_xx(x, b, c) -
                     4.0 * b / (std::pow(2.0 * x + b, 2.0) + q * q) -
                     (b * x0 / vwn_xx(x0, b, c)) *
                         (2.0 / (x - x0) - (2.0 * x + b) / vwn_xx(x, b, c) -
                          4.0 * (2.0 * x0 + b) / (std::pow(2.0 * x + b, 2.0) + q * q)));
  return deps;
}
double vwn_epsp(double x) { return vwn_eps(x, 0.0310907, -0.10498, 3.72744, 12.9352); }    // :120-122
double vwn_epsf(double x) { return vwn_eps(x, 0.01554535, -0.32500, 7.06042, 13.0045); }   // :124-126
double vwn_depsp(double x) { return vwn_deps(x, 0.0310907, -0.10498, 3.72744, 12.9352); }  // :136-138
double vwn_depsf(double x) { return vwn_deps(x, 0.01554535, -0.32500, 7.06042, 13.0045); } // :140-142
double vwn_g(double z) {  // :152-154
  return 1.125 * (std::pow(1.0 + z, 4.0 / 3.0) + std::pow(1.0 - z, 4.0 / 3.0) - 2.0);
}
double vwn_dg(double z) { return 1.5 * (std::pow(1.0 + z, 1.0 / 3.0) - std::pow(1.0 - z, 1.0 / 3.0)); }

void xalpha(long n, const double* da, const double* db, double* ex, double* vxa, double* vxb) {  // :24-63
  static const double tol = 1e-10;
  static const double xa = 2.0 / 3.0;
  static const double fac = -2.25 * xa * std::pow(3.0 / 4.0 / kPi, 1.0 / 3.0);
  for (long i = 0; i < n; i++) ex[i] = vxa[i] = vxb[i] = 0.0;
#pragma omp parallel for
  for (long i = 0; i < n; i++) {
    if (da[i] < tol) continue;
    double rho3 = std::pow(da[i], 1.0 / 3.0);
    ex[i] += fac * da[i] * rho3;
    vxa[i] += 4.0 / 3.0 * fac * rho3;
  }
#pragma omp parallel for
  for (long i = 0; i < n; i++) {
    if (db[i] < tol) continue;
    double rho3 = std::pow(db[i], 1.0 / 3.0);
    ex[i] += fac * db[i] * rho3;
    vxb[i] += 4.0 / 3.0 * fac * rho3;
  }
}

void vwn(long n, const double* da, const double* db, double* ec, double* vca, double* vcb) {  // :65-114
  static const double tol = 1e-10;
#pragma omp parallel for
  for (long i = 0; i < n; i++) {
    double dens = da[i] + db[i];
    if (dens < tol) {
      ec[i] = vca[i] = vcb[i] = 0.0;
      continue;
    }
    double zeta = (da[i] - db[i]) / dens;
    double x = std::pow(3.0 / 4.0 / kPi / dens, 1.0 / 6.0);
    double epsp = vwn_epsp(x);
    double depsp = vwn_depsp(x);
    double g = vwn_g(zeta);
    if (g < tol) {
      ec[i] = epsp * dens;
      vca[i] = epsp - (x / 6.0) * depsp;
      vcb[i] = epsp - (x / 6.0) * depsp;
      continue;
    }
    double epsf = vwn_epsf(x);
    double eps = epsp + g * (epsf - epsp);
    ec[i] = eps * dens;
    double depsf = vwn_depsf(x);
    double dg = vwn_dg(zeta);
    double deps_dx = depsp + g * (depsf - depsp);
    double deps_dg = (epsf - epsp) * dg;
    vca[i] = eps - (x / 6.0) * deps_dx + deps_dg * (1.0 - zeta);
    vcb[i] = eps - (x / 6.0) * deps_dx - deps_dg * (1.0 + zeta);
  }
}

// ---------------------------------------------------------------------------
// Becke partition (src/moleculargrid.cpp:275-329)
// ---------------------------------------------------------------------------
double fk(unsigned k, double mu) {
  for (unsigned i = 0; i < k; i++) mu = 3.0 / 2.0 * mu - 0.5 * std::pow(mu, 3.0);
  return mu;
}
double cutoff(double mu) { return 0.5 * (1.0 - fk(3, mu)); }

}  // namespace

// ---------------------------------------------------------------------------
// The grid object
// ---------------------------------------------------------------------------
struct or_grid {
  int natoms = 0, nbf = 0, nrad = 0, nang = 0, lmax = 0, nlm = 0;
  // cells that carry points: all natoms, or the first ncell of them when a
  // timing sample builds only part of the grid (every atom still enters the
  // Becke partition and the cross-atom potential)
  int ncell = 0;
  std::vector<Vec3> atoms;
  std::vector<int> Z;
  std::vector<CGF> cgfs;
  std::vector<double> leb;  // nang x 4
  std::vector<double> r_n;  // nrad
  // per point (atom-major, then radial p = 1..N, then angular)
  std::vector<Vec3> r, r_at;
  std::vector<double> w, wb, rho;
  std::vector<double> amp;  // npts x nbf
  std::vector<double> cell_density;
  std::vector<char> cell_cached;
  // Hartree state
  std::vector<double> rho_lm, U_lm, vcell, V;  // natoms x nrad x nlm ; npts
  std::vector<std::vector<Spline>> splines;    // natoms x nlm
  // XC state
  std::vector<double> exc_density, wva;
  // wall seconds of the last or_grid_hartree: projection, Poisson + own cell +
  // splines, cross-atom potential, J (for CPU timing samples)
  double t_hartree[4] = {0.0, 0.0, 0.0, 0.0};
  long npts() const { return (long)ncell * nrad * nang; }
  long cell() const { return (long)nrad * nang; }
};

namespace {

// AtomicGrid::create_atomic_grid, src/atomicgrid.cpp:32-88
void create_atomic_grid(or_grid* g, int a) {
  const Vec3 p1 = g->atoms[a];
  const double f = kPi / (double)(g->nrad + 1);
  const long base = (long)a * g->cell();
  for (int p = 1; p <= g->nrad; p++) {
    double w = f * std::pow(std::sin(f * (double)p), 2.0);
    double x = std::cos(f * (double)p);
    const double r = (1.0 + x) / (1.0 - x);
    if (a == 0) g->r_n[p - 1] = r;
    w = w / std::sqrt(1.0 - std::pow(x, 2.0)) * 2.0 / std::pow(1.0 - x, 2.0);
    for (int j = 0; j < g->nang; j++) {
      const double* L = &g->leb[4 * (size_t)j];
      Vec3 p2{L[0] * r, L[1] * r, L[2] * r};
      long idx = base + (long)(p - 1) * g->nang + j;
      g->r[idx] = Vec3{p1.x + p2.x, p1.y + p2.y, p1.z + p2.z};
      g->r_at[idx] = p2;
      g->w[idx] = w * L[3];
      for (int b = 0; b < g->nbf; b++) g->amp[(size_t)idx * g->nbf + b] = cgf_amp(g->cgfs[b], g->r[idx]);
    }
  }
  for (long i = 0; i < g->cell(); i++) g->w[base + i] *= sqnorm(g->r_at[base + i]) * 4.0 * kPi;
}

// MolecularGrid::get_becke_weight_pn, src/moleculargrid.cpp:275-300
double becke_pn(const or_grid* g, int i, const Vec3& p0) {
  double wprod = 1.0;
  const Vec3 p1 = g->atoms[i];
  for (int j = 0; j < g->natoms; j++) {
    if (i == j) continue;
    const Vec3 p2 = g->atoms[j];
    double mu = (norm(sub(p0, p1)) - norm(sub(p0, p2))) / norm(sub(p2, p1));
    wprod *= cutoff(mu);
  }
  return wprod;
}

// AtomicGrid::calculate_density, src/atomicgrid.cpp:520-530 (serial sum)
double cell_density(or_grid* g, int a) {
  if (!g->cell_cached[a]) {
    double d = 0.0;
    const long base = (long)a * g->cell();
    for (long i = 0; i < g->cell(); i++) d += g->w[base + i] * g->rho[base + i];
    g->cell_density[a] = d;
    g->cell_cached[a] = 1;
  }
  return g->cell_density[a];
}

inline int lm_index(int l, int m) { return l * l + l + m; }  // src/atomicgrid.h:272-274

// AtomicGrid::calculate_rho_lm, src/atomicgrid.cpp:241-296
void calculate_rho_lm(or_grid* g, int a) {
  const long base = (long)a * g->cell();
  double* rho_lm = &g->rho_lm[(size_t)a * g->nrad * g->nlm];
  for (int i = 0; i < g->nrad * g->nlm; i++) rho_lm[i] = 0.0;
  if (a >= g->ncell) return;  // timing sample: a cell without points
#pragma omp parallel for
  for (int i = 0; i < g->nrad; i++) {
    int cnt = 0;
    for (int l = 0; l <= g->lmax; l++) {
      for (int m = -l; m <= l; m++) {
        const double pre = sh_prefactor(l, m);
        double& acc = rho_lm[(size_t)i * g->nlm + cnt];
        for (int j = 0; j < g->nang; j++) {
          long idx = base + (long)i * g->nang + j;
          const Vec3& pos = g->r_at[idx];
          const double w = g->leb[4 * (size_t)j + 3];
          const double wb = g->wb[idx];
          const double r = norm(pos);
          const double azimuth = std::atan2(pos.y, pos.x);
          const double pole = std::acos(pos.z / r);
          const double y_lm = pre * spherical_harmonic(l, m, pole, azimuth);
          acc += g->rho[idx] * y_lm * w * wb;
        }
        acc *= 4.0 * M_PI;
        cnt++;
      }
    }
  }
}

// AtomicGrid::calculate_U_lm, src/atomicgrid.cpp:301-466
void calculate_U_lm(or_grid* g, int a) {
  const int N = g->nrad;
  static const double sqrt4pi = std::sqrt(4.0 * M_PI);
  const double q_n = a < g->ncell ? cell_density(g, a) : 0.0;
  std::vector<double> A;
  build_fd_matrix(N, g->r_n.data(), A);
  const double* rho_lm = &g->rho_lm[(size_t)a * N * g->nlm];
  double* U = &g->U_lm[(size_t)a * N * g->nlm];
  const int n = N + 2;
#pragma omp parallel for
  for (int l = 0; l <= g->lmax; l++) {
    for (int m = -l; m <= l; m++) {
      int lm = lm_index(l, m);
      std::vector<double> gv(n, 0.0);
      std::vector<double> M(A);
      for (int i = 0; i < n; i++) {
        if (i == 0) {
          gv[i] = (lm == 0) ? sqrt4pi * q_n : 0.0;
          continue;
        }
        if (i == N + 1) {
          gv[i] = 0.0;
          continue;
        }
        M[i + (size_t)i * n] -= (double)l * (double)(l + 1) / (g->r_n[i - 1] * g->r_n[i - 1]);
        gv[i] = -4.0 * M_PI * g->r_n[i - 1] * rho_lm[(size_t)(i - 1) * g->nlm + lm];
      }
      std::vector<int> perm;
      lu_factor(n, M, perm);
      std::vector<double> b(n);
      lu_solve(n, M, perm, gv.data(), b.data());
      for (int i = 1; i < N + 1; i++) U[(size_t)(i - 1) * g->nlm + lm] = b[i];
    }
  }
  const long base = (long)a * g->cell();
  const int nshell_pts = a < g->ncell ? N : 0;  // timing sample: a cell without points
  for (long i = 0; i < (long)nshell_pts * g->nang; i++) g->vcell[base + i] = 0.0;
#pragma omp parallel for
  for (int i = 0; i < nshell_pts; i++) {
    for (int l = 0; l <= g->lmax; l++) {
      for (int m = -l; m <= l; m++) {
        int lm = lm_index(l, m);
        const double pre = sh_prefactor(l, m);
        for (int j = 0; j < g->nang; j++) {
          long idx = base + (long)i * g->nang + j;
          const Vec3 pos = g->r_at[idx];
          const double r = norm(pos);
          const double azimuth = std::atan2(pos.y, pos.x);
          const double pole = std::acos(pos.z / r);
          const double y_lm = pre * spherical_harmonic(l, m, pole, azimuth);
          g->vcell[idx] += 1.0 / r * y_lm * U[(size_t)i * g->nlm + lm];
        }
      }
    }
  }
  // AtomicGrid::interpolate_sh_coeff, src/atomicgrid.cpp:535-556
  g->splines[a].assign(g->nlm, Spline());
  for (int lm = 0; lm < g->nlm; lm++) {
    Spline& s = g->splines[a][lm];
    for (int i = N - 1; i >= 0; i--) s.x.push_back(g->r_n[i]);
    for (int j = N - 1; j >= 0; j--) s.y.push_back(U[(size_t)j * g->nlm + lm]);
    spline_generate(s);
  }
}

}  // namespace

extern "C" {

or_grid* or_grid_create_cells(int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
                              const int* shell_L, const int* shell_nprim, const int* shell_offset,
                              const double* alpha, const double* coef, int nrad, int nang, const double* leb,
                              int lmax, int ncell) {
  or_grid* g = new or_grid;
  g->natoms = natoms;
  g->ncell = ncell < 1 || ncell > natoms ? natoms : ncell;
  g->nrad = nrad;
  g->nang = nang;
  g->lmax = lmax;
  g->nlm = (lmax + 1) * (lmax + 1);
  for (int a = 0; a < natoms; a++) {
    g->atoms.push_back(Vec3{xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]});
    g->Z.push_back(Z[a]);
  }
  g->cgfs = build_cgfs(xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef);
  g->nbf = (int)g->cgfs.size();
  g->leb.assign(leb, leb + 4 * (size_t)nang);
  g->r_n.assign(nrad, 0.0);
  const long np = g->npts();
  g->r.resize(np);
  g->r_at.resize(np);
  g->w.assign(np, 0.0);
  g->wb.assign(np, 0.0);
  g->rho.assign(np, 0.0);
  g->vcell.assign(np, 0.0);
  g->V.assign(np, 0.0);
  g->amp.assign((size_t)np * g->nbf, 0.0);
  g->cell_density.assign(natoms, 0.0);
  g->cell_cached.assign(natoms, 0);
  g->rho_lm.assign((size_t)natoms * nrad * g->nlm, 0.0);
  g->U_lm.assign((size_t)natoms * nrad * g->nlm, 0.0);
  g->splines.resize(natoms);
  // src/moleculargrid.cpp:215-226 (grids serially, as the reference)
  for (int a = 0; a < g->ncell; a++) create_atomic_grid(g, a);
  // src/moleculargrid.cpp:229-254 (serial triple loop)
  for (int i = 0; i < g->ncell; i++) {
    const long base = (long)i * g->cell();
    for (long j = 0; j < g->cell(); j++) {
      double denom = 0.0, nom = 1.0;
      for (int k = 0; k < natoms; k++) {
        const double term = becke_pn(g, k, g->r[base + j]);
        denom += term;
        if (i == k) nom = term;
      }
      const double vw = nom / denom;
      // AtomicGrid::set_becke_weights, src/atomicgrid.cpp:196-206
      g->w[base + j] *= vw;
      g->wb[base + j] = vw;
    }
  }
  return g;
}

or_grid* or_grid_create(int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
                        const int* shell_L, const int* shell_nprim, const int* shell_offset,
                        const double* alpha, const double* coef, int nrad, int nang, const double* leb,
                        int lmax) {
  return or_grid_create_cells(natoms, Z, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef,
                              nrad, nang, leb, lmax, natoms);
}

void or_grid_destroy(or_grid* g) { delete g; }
int or_grid_nbf(const or_grid* g) { return g->nbf; }
long or_grid_npts(const or_grid* g) { return g->npts(); }

void or_grid_points(const or_grid* g, double* xyz, double* w, double* wb) {
  for (long i = 0; i < g->npts(); i++) {
    if (xyz) {
      xyz[3 * i] = g->r[i].x;
      xyz[3 * i + 1] = g->r[i].y;
      xyz[3 * i + 2] = g->r[i].z;
    }
    if (w) w[i] = g->w[i];
    if (wb) wb[i] = g->wb[i];
  }
}
void or_grid_amplitudes(const or_grid* g, double* phi) {
  std::memcpy(phi, g->amp.data(), sizeof(double) * g->amp.size());
}
void or_grid_radial(const or_grid* g, double* r) {
  for (int i = 0; i < g->nrad; i++) r[i] = g->r_n[i];
}

// MolecularGrid::set_density -> AtomicGrid::set_density -> GridPoint::set_density
// (src/moleculargrid.cpp:48-53, src/atomicgrid.cpp:101-108, src/gridpoint.cpp:82-84)
void or_grid_set_density(or_grid* g, const double* P) {
  const int nbf = g->nbf;
#pragma omp parallel for
  for (int a = 0; a < g->ncell; a++) {
    std::vector<double> t(nbf);
    const long base = (long)a * g->cell();
    for (long i = 0; i < g->cell(); i++) {
      const double* phi = &g->amp[(size_t)(base + i) * nbf];
      for (int r = 0; r < nbf; r++) t[r] = 0.0;
      for (int c = 0; c < nbf; c++)
        for (int r = 0; r < nbf; r++) t[r] += P[r + (size_t)c * nbf] * phi[c];
      double d = 0.0;
      for (int r = 0; r < nbf; r++) d += phi[r] * t[r];
      g->rho[base + i] = 2.0 * d;
    }
    g->cell_cached[a] = 0;
  }
}

// MolecularGrid::correct_densities, src/moleculargrid.cpp:132-146
void or_grid_correct_densities(or_grid* g) {
  double sum = 0.0, charge = 0.0;
  for (int a = 0; a < g->ncell; a++) {
    sum += cell_density(g, a);
    charge += g->Z[a];
  }
  double factor = charge / sum;
  for (int a = 0; a < g->ncell; a++) {
    const long base = (long)a * g->cell();
    for (long i = 0; i < g->cell(); i++) g->rho[base + i] *= factor;
    g->cell_cached[a] = 0;
  }
}

// MolecularGrid::calculate_density, src/moleculargrid.cpp:157-166
double or_grid_calculate_density(or_grid* g) {
  double sum = 0.0;
  for (int a = 0; a < g->ncell; a++) sum += cell_density(g, a);
  return sum;
}
void or_grid_densities(const or_grid* g, double* rho) {
  std::memcpy(rho, g->rho.data(), sizeof(double) * g->rho.size());
}
void or_grid_cell_densities(or_grid* g, double* q) {
  for (int a = 0; a < g->ncell; a++) q[a] = cell_density(g, a);
}

// MolecularGrid::calculate_hartree_potential, src/moleculargrid.cpp:336-389
void or_grid_hartree(or_grid* g, double* J) {
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  const auto t0 = clk::now();
  for (int a = 0; a < g->natoms; a++) calculate_rho_lm(g, a);
  const auto t1 = clk::now();
  for (int a = 0; a < g->natoms; a++) calculate_U_lm(g, a);
  const auto t2 = clk::now();
  const int N = g->nrad;
  for (int i = 0; i < g->ncell; i++) {
    const long base = (long)i * g->cell();
#pragma omp parallel for
    for (long j = 0; j < g->cell(); j++) {
      double V = 0.0;
      for (int k = 0; k < g->natoms; k++) {
        if (i == k) {
          V += g->vcell[base + j];
        } else {
          int lm = 0;
          for (int l = 0; l <= g->lmax; l++) {
            for (int m = -l; m <= l; m++) {
              const double pre = sh_prefactor(l, m);
              const Vec3 pos = sub(g->r[base + j], g->atoms[k]);
              const double r = norm(pos);
              const double azimuth = std::atan2(pos.y, pos.x);
              const double pole = std::acos(pos.z / r);
              const double y_lm = pre * spherical_harmonic(l, m, pole, azimuth);
              const Spline& s = g->splines[k][lm];
              V += 1.0 / r * y_lm * spline_eval(s.x.data(), s.y.data(), s.c.data(), N, r);
              lm++;
            }
          }
        }
      }
      g->V[base + j] = V;
    }
  }
  const auto t3 = clk::now();
  // AtomicGrid::calculate_J_matrix, src/atomicgrid.cpp:471-488, summed at :382-388
  const int nbf = g->nbf;
  for (size_t i = 0; i < (size_t)nbf * nbf; i++) J[i] = 0.0;
  std::vector<double> Jj((size_t)nbf * nbf);
  for (int a = 0; a < g->ncell; a++) {
    const long base = (long)a * g->cell();
    const long np = g->cell();
#pragma omp parallel for
    for (int i = 0; i < nbf; i++) {
      std::vector<double> row_i(np);
      for (long p = 0; p < np; p++)
        row_i[p] = (g->w[base + p] * g->V[base + p]) * g->amp[(size_t)(base + p) * nbf + i];
      for (int j = i; j < nbf; j++) {
        double d = 0.0;
        for (long p = 0; p < np; p++) d += row_i[p] * g->amp[(size_t)(base + p) * nbf + j];
        Jj[i + (size_t)j * nbf] = Jj[j + (size_t)i * nbf] = d * 0.5;
      }
    }
    for (size_t i = 0; i < (size_t)nbf * nbf; i++) J[i] += Jj[i];
  }
  const auto t4 = clk::now();
  g->t_hartree[0] = secs(t0, t1);
  g->t_hartree[1] = secs(t1, t2);
  g->t_hartree[2] = secs(t2, t3);
  g->t_hartree[3] = secs(t3, t4);
}

void or_grid_hartree_times(const or_grid* g, double* t) {
  for (int i = 0; i < 4; i++) t[i] = g->t_hartree[i];
}

void or_grid_hartree_parts(const or_grid* g, double* rho_lm, double* U_lm, double* vcell, double* v,
                           double* spl) {
  if (rho_lm) std::memcpy(rho_lm, g->rho_lm.data(), sizeof(double) * g->rho_lm.size());
  if (U_lm) std::memcpy(U_lm, g->U_lm.data(), sizeof(double) * g->U_lm.size());
  if (vcell) std::memcpy(vcell, g->vcell.data(), sizeof(double) * g->vcell.size());
  if (v) std::memcpy(v, g->V.data(), sizeof(double) * g->V.size());
  if (spl) {
    const size_t per = 4 * (size_t)(g->nrad - 1);
    for (int a = 0; a < g->natoms; a++)
      for (int lm = 0; lm < g->nlm; lm++)
        if (!g->splines[a].empty())
          std::memcpy(spl + ((size_t)a * g->nlm + lm) * per, g->splines[a][lm].c.data(),
                      sizeof(double) * per);
  }
}

// DFT::calculate_exchange_correlation_matrix, src/dft.cpp:394-433
void or_grid_xc(or_grid* g, double* XC, double* exc) {
  const long n = g->npts();
  const int nbf = g->nbf;
  std::vector<double> da(n), db(n), ex(n), vxa(n), vxb(n), ec(n), vca(n), vcb(n);
  for (long i = 0; i < n; i++) da[i] = db[i] = g->rho[i] * 0.5;
  xalpha(n, da.data(), db.data(), ex.data(), vxa.data(), vxb.data());
  vwn(n, da.data(), db.data(), ec.data(), vca.data(), vcb.data());
  g->exc_density.assign(n, 0.0);
  g->wva.assign(n, 0.0);
  double e = 0.0;
  for (long i = 0; i < n; i++) {
    g->exc_density[i] = ex[i] + ec[i];
    e += g->w[i] * g->exc_density[i];
  }
  *exc = e;
  for (long i = 0; i < n; i++) g->wva[i] = g->w[i] * ((vxa[i] + vxb[i] + vca[i] + vcb[i]) * 0.5);
  std::vector<double> wva_i(n);
  for (int i = 0; i < nbf; i++) {  // serial, as the reference
    for (long p = 0; p < n; p++) wva_i[p] = g->wva[p] * g->amp[(size_t)p * nbf + i];
    for (int j = 0; j < nbf; j++) {
      double d = 0.0;
      for (long p = 0; p < n; p++) d += wva_i[p] * g->amp[(size_t)p * nbf + j];
      XC[i + (size_t)j * nbf] = d;
    }
  }
}

void or_grid_xc_parts(const or_grid* g, double* exc_density, double* wva) {
  if (exc_density) std::memcpy(exc_density, g->exc_density.data(), sizeof(double) * g->exc_density.size());
  if (wva) std::memcpy(wva, g->wva.data(), sizeof(double) * g->wva.size());
}

// --- primitives -------------------------------------------------------------
double or_gto_norm(double alpha, int l, int m, int n) { return gto_norm(alpha, l, m, n); }
double or_sh_prefactor(int l, int m) { return sh_prefactor(l, m); }
double or_spherical_harmonic(int l, int m, double pole, double azimuth) {
  return spherical_harmonic(l, m, pole, azimuth);
}
double or_ylm_xyz(int l, int m, double x, double y, double z) {
  const double r = std::sqrt(x * x + y * y + z * z);
  const double azimuth = std::atan2(y, x);
  const double pole = std::acos(z / r);
  return sh_prefactor(l, m) * spherical_harmonic(l, m, pole, azimuth);
}
double or_becke_cutoff(double mu) { return cutoff(mu); }
void or_xalpha(long n, const double* a, const double* b, double* ex, double* vxa, double* vxb) {
  xalpha(n, a, b, ex, vxa, vxb);
}
void or_vwn(long n, const double* a, const double* b, double* ec, double* vca, double* vcb) {
  vwn(n, a, b, ec, vca, vcb);
}
int or_spline_build(int n, const double* x, const double* y, double* coef) {
  try {
    Spline s;
    s.x.assign(x, x + n);
    s.y.assign(y, y + n);
    spline_generate(s);
    std::memcpy(coef, s.c.data(), sizeof(double) * s.c.size());
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}
double or_spline_eval(int n, const double* x, const double* y, const double* coef, double xx) {
  return spline_eval(x, y, coef, (size_t)n, xx);
}
void or_poisson_matrix(int N, int l, const double* r, double* M) {
  std::vector<double> A;
  build_fd_matrix(N, r, A);
  const int n = N + 2;
  for (int i = 1; i < N + 1; i++) A[i + (size_t)i * n] -= (double)l * (double)(l + 1) / (r[i - 1] * r[i - 1]);
  std::memcpy(M, A.data(), sizeof(double) * A.size());
}
void or_lu_solve(int n, const double* A, const double* b, double* x) {
  std::vector<double> M(A, A + (size_t)n * n);
  std::vector<int> perm;
  lu_factor(n, M, perm);
  lu_solve(n, M, perm, b, x);
}

// RectangularGrid::build_grid positions (src/rectangulargrid.cpp:34-57): dp^3
// points, x fastest, spanning [-size/2, size/2].
void or_rectangular_points(double size, unsigned dp, double* xyz) {
  const double gdsize = size / (double)(dp - 1);
  long q = 0;
  for (unsigned i = 0; i < dp; i++)
    for (unsigned j = 0; j < dp; j++)
      for (unsigned k = 0; k < dp; k++) {
        xyz[q++] = (double)k * gdsize - size / 2.0;
        xyz[q++] = (double)j * gdsize - size / 2.0;
        xyz[q++] = (double)i * gdsize - size / 2.0;
      }
}

// RectangularGrid::set_density (src/rectangulargrid.cpp:68-73) at the given
// points: GridPoint::set_basis_func_amp / set_basis_func_grad
// (src/gridpoint.cpp:45-72), set_density (82-84) and set_gradient (94-108),
// in the reference's product-rule form with both products D*x and D*phi.
// The basis gradient is the derivative of the amplitude formula (the
// reference's GTO::get_grad, src/cgf.cpp:67-94, is not followed: it omits N,
// applies c twice and drops the factor l).  P column-major; rho [npts],
// grad [npts][3], dphi [npts][3][nbf]; any output may be NULL.
void or_points_density(int natoms, const double* xyz, int nshell, const int* shell_atom, const int* shell_L,
                       const int* shell_nprim, const int* shell_offset, const double* alpha, const double* coef,
                       long npts, const double* pts, const double* P, double* rho, double* grad, double* dphi) {
  (void)natoms;
  const std::vector<CGF> cgfs =
      build_cgfs(xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef);
  const int n = (int)cgfs.size();
#pragma omp parallel for
  for (long p = 0; p < npts; ++p) {
    const Vec3 r{pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    std::vector<double> amp(n), gx(n), gy(n), gz(n);
    for (int i = 0; i < n; ++i) {
      amp[i] = cgf_amp(cgfs[i], r);
      double g[3] = {0.0, 0.0, 0.0};
      for (const GTO& t : cgfs[i].gtos) {
        const double x = r.x - t.pos.x, y = r.y - t.pos.y, z = r.z - t.pos.z;
        const double e = std::exp(-t.alpha * (x * x + y * y + z * z));
        const double fx = std::pow(x, t.l), fy = std::pow(y, t.m), fz = std::pow(z, t.n);
        const double dfx = (t.l > 0 ? t.l * std::pow(x, t.l - 1) : 0.0) - 2.0 * t.alpha * x * fx;
        const double dfy = (t.m > 0 ? t.m * std::pow(y, t.m - 1) : 0.0) - 2.0 * t.alpha * y * fy;
        const double dfz = (t.n > 0 ? t.n * std::pow(z, t.n - 1) : 0.0) - 2.0 * t.alpha * z * fz;
        const double cn = t.c * t.norm * e;
        g[0] += cn * dfx * fy * fz;
        g[1] += cn * fx * dfy * fz;
        g[2] += cn * fx * fy * dfz;
      }
      gx[i] = g[0];
      gy[i] = g[1];
      gz[i] = g[2];
    }
    // D * v for a vector v (column-major D)
    auto dmul = [&](const std::vector<double>& v) {
      std::vector<double> out(n, 0.0);
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) out[i] += P[i + (long)j * n] * v[j];
      return out;
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
      double t = 0.0;
      for (int i = 0; i < n; ++i) t += a[i] * b[i];
      return t;
    };
    const std::vector<double> Dphi = dmul(amp);
    if (rho) rho[p] = 2.0 * dot(amp, Dphi);
    if (grad) {
      grad[3 * p] = 2.0 * dot(amp, dmul(gx)) + 2.0 * dot(gx, Dphi);
      grad[3 * p + 1] = 2.0 * dot(amp, dmul(gy)) + 2.0 * dot(gy, Dphi);
      grad[3 * p + 2] = 2.0 * dot(amp, dmul(gz)) + 2.0 * dot(gz, Dphi);
    }
    if (dphi)
      for (int i = 0; i < n; ++i) {
        dphi[(p * 3 + 0) * n + i] = gx[i];
        dphi[(p * 3 + 1) * n + i] = gy[i];
        dphi[(p * 3 + 2) * n + i] = gz[i];
      }
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Integrals (src/integrals.cpp) and the Boys function (src/gamma.cpp)
// ---------------------------------------------------------------------------
namespace {

// GammaInc::gser / gcf / gammp / gamm_inc / Fgamma, src/gamma.cpp:41-161
// (gammln -> std::lgamma; the a >= 100 quadrature branch is unreachable for
// s, p, d shells and is not restated)
double gser(double a, double x) {
  const double EPS = 2.220446049250313e-16;
  double gln = std::lgamma(a);
  double ap = a, del, sum;
  del = sum = 1.0 / a;
  for (;;) {
    ++ap;
    del *= x / ap;
    sum += del;
    if (std::fabs(del) < std::fabs(sum) * EPS) return sum * std::exp(-x + a * std::log(x) - gln);
  }
}
double gcf(double a, double x) {
  const double EPS = 2.220446049250313e-16;
  const double FPMIN = 2.2250738585072014e-308 / EPS;
  double gln = std::lgamma(a);
  double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
  for (int i = 1;; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < FPMIN) d = FPMIN;
    c = b + an / c;
    if (std::fabs(c) < FPMIN) c = FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) <= EPS) break;
  }
  return std::exp(-x + a * std::log(x) - gln) * h;
}
double gammp(double a, double x) {
  if (x < 0.0 || a <= 0.0) return 0.0;
  if (x == 0.0) return 0.0;
  if (x < a + 1.0) return gser(a, x);
  return 1.0 - gcf(a, x);
}
double Fgamma(double m, double x) {
  double tiny = 0.00000001;
  x = std::max(std::fabs(x), tiny);
  double val = std::exp(std::lgamma(m + 0.5)) * gammp(m + 0.5, x);
  return 0.5 * std::pow(x, -m - 0.5) * val;
}

double binomial(int a, int b) {  // src/integrals.cpp:318-324
  if ((a < 0) | (b < 0) | (a - b < 0)) return 1.0;
  return factorial(a) / (factorial(b) * factorial(a - b));
}
double binomial_prefactor(int s, int ia, int ib, double xpa, double xpb) {  // :301-316
  double sum = 0.0;
  for (int t = 0; t < s + 1; t++) {
    if ((s - ia <= t) && (t <= ib)) {
      sum += binomial(ia, s - t) * binomial(ib, t) * std::pow(xpa, ia - s + t) * std::pow(xpb, ib - t);
    }
  }
  return sum;
}
double overlap_1D(int l1, int l2, double x1, double x2, double gamma) {  // :278-288
  double sum = 0.0;
  for (int i = 0; i < (1 + std::floor(0.5 * (l1 + l2))); i++) {
    sum += binomial_prefactor(2 * i, l1, l2, x1, x2) * (i == 0 ? 1 : double_factorial(2 * i - 1)) /
           std::pow(2 * gamma, i);
  }
  return sum;
}
Vec3 product_center(double a1, const Vec3& a, double a2, const Vec3& b) {  // :290-294
  double g = a1 + a2;
  return Vec3{(a1 * a.x + a2 * b.x) / g, (a1 * a.y + a2 * b.y) / g, (a1 * a.z + a2 * b.z) / g};
}
double overlap_prim(double alpha1, int l1, int m1, int n1, const Vec3& a, double alpha2, int l2, int m2,
                    int n2, const Vec3& b) {  // :252-276
  double rab2 = sqnorm(sub(a, b));
  double gamma = alpha1 + alpha2;
  Vec3 p = product_center(alpha1, a, alpha2, b);
  double pre = std::pow(kPi / gamma, 1.5) * std::exp(-alpha1 * alpha2 * rab2 / gamma);
  double wx = overlap_1D(l1, l2, p.x - a.x, p.x - b.x, gamma);
  double wy = overlap_1D(m1, m2, p.y - a.y, p.y - b.y, gamma);
  double wz = overlap_1D(n1, n2, p.z - a.z, p.z - b.z, gamma);
  return pre * wx * wy * wz;
}
double overlap_gto(const GTO& g1, const GTO& g2) {
  return overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l, g2.m, g2.n, g2.pos);
}
double kinetic_gto(const GTO& g1, const GTO& g2) {  // :88-135
  double term0 = g2.alpha * (2.0 * (g2.l + g2.m + g2.n) + 3.0) * overlap_gto(g1, g2);
  double term1 = -2.0 * std::pow(g2.alpha, 2.0) *
                 (overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l + 2, g2.m, g2.n, g2.pos) +
                  overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l, g2.m + 2, g2.n, g2.pos) +
                  overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l, g2.m, g2.n + 2, g2.pos));
  double term2 =
      -0.5 * (g2.l * (g2.l - 1) *
                  overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l - 2, g2.m, g2.n, g2.pos) +
              g2.m * (g2.m - 1) *
                  overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l, g2.m - 2, g2.n, g2.pos) +
              g2.n * (g2.n - 1) *
                  overlap_prim(g1.alpha, g1.l, g1.m, g1.n, g1.pos, g2.alpha, g2.l, g2.m, g2.n - 2, g2.pos));
  return term0 + term1 + term2;
}
double A_term(int i, int r, int u, int l1, int l2, double pax, double pbx, double cpx, double gamma) {  // :355-366
  return std::pow(-1, i) * binomial_prefactor(i, l1, l2, pax, pbx) * std::pow(-1, u) * factorial(i) *
         std::pow(cpx, i - 2 * r - 2 * u) * std::pow(0.25 / gamma, r + u) / factorial(r) / factorial(u) /
         factorial(i - 2 * r - 2 * u);
}
std::vector<double> A_array(int l1, int l2, double pa, double pb, double cp, double g) {  // :339-353
  int imax = l1 + l2 + 1;
  std::vector<double> arrA(imax, 0);
  for (int i = 0; i < imax; i++)
    for (int r = 0; r <= i / 2; r++)
      for (int u = 0; u <= (i - 2 * r) / 2; u++) {
        int iI = i - 2 * r - u;
        arrA[iI] += A_term(i, r, u, l1, l2, pa, pb, cp, g);
      }
  return arrA;
}
double nuclear_gto(const GTO& g1, const GTO& g2, const Vec3& c) {  // :336-387 (truncated pi)
  const Vec3& a = g1.pos;
  const Vec3& b = g2.pos;
  double gamma = g1.alpha + g2.alpha;
  Vec3 p = product_center(g1.alpha, a, g2.alpha, b);
  double rab2 = sqnorm(sub(a, b));
  double rcp2 = sqnorm(sub(c, p));
  std::vector<double> ax = A_array(g1.l, g2.l, p.x - a.x, p.x - b.x, p.x - c.x, gamma);
  std::vector<double> ay = A_array(g1.m, g2.m, p.y - a.y, p.y - b.y, p.y - c.y, gamma);
  std::vector<double> az = A_array(g1.n, g2.n, p.z - a.z, p.z - b.z, p.z - c.z, gamma);
  double sum = 0.0;
  for (int i = 0; i <= g1.l + g2.l; i++)
    for (int j = 0; j <= g1.m + g2.m; j++)
      for (int k = 0; k <= g1.n + g2.n; k++) sum += ax[i] * ay[j] * az[k] * Fgamma(i + j + k, rcp2 * gamma);
  return -2.0 * kPiTrunc / gamma * std::exp(-g1.alpha * g2.alpha * rab2 / gamma) * sum;
}
double fact_ratio2(int a, int b) { return factorial(a) / factorial(b) / factorial(a - 2 * b); }  // :466-468
double B0(int i, int r, double g) { return fact_ratio2(i, r) * std::pow(4 * g, r - i); }
double fB(int i, int l1, int l2, double p, double a, double b, int r, double g) {
  return binomial_prefactor(i, l1, l2, p - a, p - b) * B0(i, r, g);
}
double B_term(int i1, int i2, int r1, int r2, int u, int l1, int l2, int l3, int l4, double px, double ax,
              double bx, double qx, double cx, double dx, double gamma1, double gamma2, double delta) {  // :437-449
  return fB(i1, l1, l2, px, ax, bx, r1, gamma1) * std::pow(-1, i2) * fB(i2, l3, l4, qx, cx, dx, r2, gamma2) *
         std::pow(-1, u) * fact_ratio2(i1 + i2 - 2 * (r1 + r2), u) *
         std::pow(qx - px, i1 + i2 - 2 * (r1 + r2) - 2 * u) / std::pow(delta, i1 + i2 - 2 * (r1 + r2) - u);
}
std::vector<double> B_array(int l1, int l2, int l3, int l4, double p, double a, double b, double q, double c,
                            double d, double g1, double g2, double delta) {  // :415-435
  int imax = l1 + l2 + l3 + l4 + 1;
  std::vector<double> arrB(imax, 0);
  for (int i1 = 0; i1 < l1 + l2 + 1; i1++)
    for (int i2 = 0; i2 < l3 + l4 + 1; i2++)
      for (int r1 = 0; r1 < i1 / 2 + 1; r1++)
        for (int r2 = 0; r2 < i2 / 2 + 1; r2++)
          for (int u = 0; u < (i1 + i2) / 2 - r1 - r2 + 1; u++) {
            int i = i1 + i2 - 2 * (r1 + r2) - u;
            arrB[i] += B_term(i1, i2, r1, r2, u, l1, l2, l3, l4, p, a, b, q, c, d, g1, g2, delta);
          }
  return arrB;
}
double repulsion_gto(const GTO& A, const GTO& B, const GTO& C, const GTO& D) {  // :389-413 (truncated pi)
  double rab2 = sqnorm(sub(A.pos, B.pos));
  double rcd2 = sqnorm(sub(C.pos, D.pos));
  Vec3 p = product_center(A.alpha, A.pos, B.alpha, B.pos);
  Vec3 q = product_center(C.alpha, C.pos, D.alpha, D.pos);
  double rpq2 = sqnorm(sub(p, q));
  double gamma1 = A.alpha + B.alpha, gamma2 = C.alpha + D.alpha;
  double delta = 0.25 * (1.0 / gamma1 + 1.0 / gamma2);
  std::vector<double> bx = B_array(A.l, B.l, C.l, D.l, p.x, A.pos.x, B.pos.x, q.x, C.pos.x, D.pos.x, gamma1, gamma2, delta);
  std::vector<double> by = B_array(A.m, B.m, C.m, D.m, p.y, A.pos.y, B.pos.y, q.y, C.pos.y, D.pos.y, gamma1, gamma2, delta);
  std::vector<double> bz = B_array(A.n, B.n, C.n, D.n, p.z, A.pos.z, B.pos.z, q.z, C.pos.z, D.pos.z, gamma1, gamma2, delta);
  double sum = 0.0;
  for (int i = 0; i <= A.l + B.l + C.l + D.l; i++)
    for (int j = 0; j <= A.m + B.m + C.m + D.m; j++)
      for (int k = 0; k <= A.n + B.n + C.n + D.n; k++)
        sum += bx[i] * by[j] * bz[k] * Fgamma(i + j + k, 0.25 * rpq2 / delta);
  return 2.0 * std::pow(kPiTrunc, 2.5) / (gamma1 * gamma2 * std::sqrt(gamma1 + gamma2)) *
         std::exp(-A.alpha * B.alpha * rab2 / gamma1) * std::exp(-C.alpha * D.alpha * rcd2 / gamma2) * sum *
         A.norm * B.norm * C.norm * D.norm;
}
double repulsion_cgf(const CGF& a, const CGF& b, const CGF& c, const CGF& d) {  // :195-217
  double sum = 0;
  for (const auto& ga : a.gtos)
    for (const auto& gb : b.gtos)
      for (const auto& gc : c.gtos)
        for (const auto& gd : d.gtos) sum += ga.c * gb.c * gc.c * gd.c * repulsion_gto(ga, gb, gc, gd);
  return sum;
}

// cyclic Jacobi eigen-solver (stands in for Eigen::SelfAdjointEigenSolver)
void jacobi_eigh(int n, const double* Ain, double* evals, double* V) {
  std::vector<double> A(Ain, Ain + (size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) V[i + (size_t)j * n] = (i == j) ? 1.0 : 0.0;
  auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * n]; };
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a(p, q);
        if (apq == 0.0) continue;
        double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
          V[k + (size_t)p * n] = c * vkp - s * vkq;
          V[k + (size_t)q * n] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  std::vector<double> Vs((size_t)n * n);
  for (int j = 0; j < n; j++) {
    evals[j] = a(idx[j], idx[j]);
    for (int i = 0; i < n; i++) Vs[i + (size_t)j * n] = V[i + (size_t)idx[j] * n];
  }
  std::memcpy(V, Vs.data(), sizeof(double) * Vs.size());
}

void matmul(int n, const double* A, const double* B, double* C) {  // C = A B, column-major
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int k = 0; k < n; k++) s += A[i + (size_t)k * n] * B[k + (size_t)j * n];
      C[i + (size_t)j * n] = s;
    }
}

}  // namespace

extern "C" {

double or_boys(double m, double x) { return Fgamma(m, x); }

// DFT::construct_matrices, src/dft.cpp:184-201
void or_one_electron(int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
                     const int* shell_L, const int* shell_nprim, const int* shell_offset,
                     const double* alpha, const double* coef, double* S, double* T, double* V) {
  std::vector<CGF> cg = build_cgfs(xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef);
  const int n = (int)cg.size();
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; i++) {
    for (int j = i; j < n; j++) {
      double s = 0.0, t = 0.0;
      for (const auto& g1 : cg[i].gtos)
        for (const auto& g2 : cg[j].gtos) {
          s += g1.norm * g2.norm * g1.c * g2.c * overlap_gto(g1, g2);
          t += g1.norm * g2.norm * g1.c * g2.c * kinetic_gto(g1, g2);
        }
      double v_sum = 0.0;
      for (int k = 0; k < natoms; k++) {
        Vec3 c{xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]};
        double v = 0.0;
        for (const auto& g1 : cg[i].gtos)
          for (const auto& g2 : cg[j].gtos) v += g1.norm * g2.norm * g1.c * g2.c * nuclear_gto(g1, g2, c);
        v_sum += v * (double)Z[k];
      }
      S[i + (size_t)j * n] = S[j + (size_t)i * n] = s;
      T[i + (size_t)j * n] = T[j + (size_t)i * n] = t;
      V[i + (size_t)j * n] = V[j + (size_t)i * n] = v_sum;
    }
  }
}

// DFT::calculate_hartree_potential_te_int, src/dft.cpp:452-477 (all ERIs direct)
void or_eri_J(int natoms, const double* xyz, int nshell, const int* shell_atom, const int* shell_L,
              const int* shell_nprim, const int* shell_offset, const double* alpha, const double* coef,
              const double* P, double* J) {
  (void)natoms;
  std::vector<CGF> cg = build_cgfs(xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef);
  const int n = (int)cg.size();
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      double s = 0.0;
      for (int k = 0; k < n; k++)
        for (int l = 0; l <= k; l++) {
          double eri = repulsion_cgf(cg[i], cg[j], cg[k], cg[l]);
          s += (k == l ? P[k + (size_t)l * n] : P[k + (size_t)l * n] + P[l + (size_t)k * n]) * eri;
        }
      J[i + (size_t)j * n] = J[j + (size_t)i * n] = s;
    }
}

void or_eigh(int n, const double* A, double* evals, double* evecs) { jacobi_eigh(n, A, evals, evecs); }

}  // extern "C"

namespace {

// The SCF state of DFT (src/dft.h:41-216): one-electron matrices, the
// canonical orthogonalisation, the grid matrices and the energy terms.
struct ScfCore {
  or_grid* g;
  int n = 0, nelec = 0;
  size_t nn = 0;
  std::vector<double> S, T, Vn, H, X, Xt, J, XC, P;
  double enuc = 0.0, exc = 0.0, et = 0.0, e_one = 0.0, e_j = 0.0, c = 1.0;
  double mixing = 0.5;
  bool is_first = true;

  ScfCore(or_grid* grid, int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
          const int* shell_L, const int* shell_nprim, const int* shell_offset, const double* alpha,
          const double* coef)
      : g(grid), n(grid->nbf), nn((size_t)grid->nbf * grid->nbf) {
    S.assign(nn, 0.0);
    T.assign(nn, 0.0);
    Vn.assign(nn, 0.0);
    H.assign(nn, 0.0);
    X.assign(nn, 0.0);
    Xt.assign(nn, 0.0);
    J.assign(nn, 0.0);
    XC.assign(nn, 0.0);
    P.assign(nn, 0.0);
    // DFT::construct_matrices, src/dft.cpp:184-201
    or_one_electron(natoms, Z, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef,
                    S.data(), T.data(), Vn.data());
    for (size_t i = 0; i < nn; i++) H[i] = T[i] + Vn[i];
    // DFT::calculate_nuclear_repulsion, src/dft.cpp:235-246
    for (int i = 0; i < natoms; i++)
      for (int j = i + 1; j < natoms; j++) {
        Vec3 a{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, b{xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2]};
        enuc += Z[i] * Z[j] / norm(sub(a, b));
      }
    for (int i = 0; i < natoms; i++) nelec += Z[i];
    // canonical orthogonalisation, src/dft.cpp:300-316
    std::vector<double> ev(n), U(nn);
    jacobi_eigh(n, S.data(), ev.data(), U.data());
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) X[i + (size_t)j * n] = U[i + (size_t)j * n] * (1.0 / std::sqrt(ev[j]));
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) Xt[i + (size_t)j * n] = X[j + (size_t)i * n];
  }

  // F = H + 2J + XC, src/dft.cpp:334
  void fock(std::vector<double>& F) const {
    for (size_t i = 0; i < nn; i++) F[i] = H[i] + 2.0 * J[i] + XC[i];
  }

  // DFT::calculate_density_matrix, src/dft.cpp:330-366 (P mixed with the old P
  // after the first pass), without its grid calls
  void density_matrix() {
    std::vector<double> F(nn), tmp(nn), Fp(nn), Cc(nn), C(nn), ev(n);
    fock(F);
    matmul(n, Xt.data(), F.data(), tmp.data());
    matmul(n, tmp.data(), X.data(), Fp.data());
    jacobi_eigh(n, Fp.data(), ev.data(), Cc.data());
    matmul(n, X.data(), Cc.data(), C.data());
    std::vector<double> Pnew(nn, 0.0);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < nelec / 2; k++) Pnew[i + (size_t)j * n] += C[i + (size_t)k * n] * C[j + (size_t)k * n];
    if (is_first) {
      P = Pnew;
      is_first = false;
    } else {
      for (size_t i = 0; i < nn; i++) P[i] = (1.0 - mixing) * Pnew[i] + mixing * P[i];
    }
  }

  // set_density + correct_densities (src/dft.cpp:362-365); hold_c > 0 scales
  // by that factor instead of the renormalisation's own
  void grid_density(double hold_c) {
    or_grid_set_density(g, P.data());
    if (hold_c > 0.0) {
      for (double& v : g->rho) v *= hold_c;
      std::fill(g->cell_cached.begin(), g->cell_cached.end(), 0);
      c = hold_c;
    } else {
      const double before = or_grid_calculate_density(g);
      or_grid_correct_densities(g);
      double charge = 0.0;
      for (int a = 0; a < g->natoms; a++) charge += g->Z[a];
      c = charge / before;
    }
  }

  // DFT::calculate_energy, src/dft.cpp:441-447
  void energy() {
    double t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < n; k++) {
        t1 += P[i + (size_t)k * n] * H[k + (size_t)i * n];
        t2 += P[i + (size_t)k * n] * J[k + (size_t)i * n];
      }
    e_one = 2.0 * t1;
    e_j = 2.0 * t2;
    et = e_one + e_j + enuc + exc;
  }

  void terms(double* t) const {
    t[0] = et;
    t[1] = e_one;
    t[2] = e_j;
    t[3] = exc;
    t[4] = enuc;
  }
};

}  // namespace

extern "C" {

// DFT::scf with J from the Becke-grid Poisson path (eri = false) or from the
// two-electron integrals (eri = true: hartree_evaluation = two_electron_integrals,
// src/dft.cpp:378-385, 452-477); rho is set and corrected every iteration in
// both (src/dft.cpp:362-365) since XC needs it.
static int scf_loop(or_grid* g, int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
                    const int* shell_L, const int* shell_nprim, const int* shell_offset, const double* alpha,
                    const double* coef, double threshold, int min_iter, int max_iter, double mixing,
                    double* energies, double* terms, double* P_out, bool eri) {
  ScfCore s(g, natoms, Z, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef);
  s.mixing = mixing;
  auto coulomb = [&] {
    if (eri)
      or_eri_J(natoms, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef, s.P.data(),
               s.J.data());
    else
      or_grid_hartree(g, s.J.data());
  };
  // DFT::DFT -> construct_matrices: the core guess, J of it, the energy (E_xc = 0, Q3)
  s.density_matrix();
  s.grid_density(0.0);
  coulomb();
  s.energy();
  // DFT::scf, src/dft.cpp:84-142
  double old_energy = s.et, difference = 1.0;
  int iteration = 0;
  while (difference > threshold || iteration < min_iter) {
    iteration++;
    s.density_matrix();
    s.grid_density(0.0);
    coulomb();
    or_grid_xc(g, s.XC.data(), &s.exc);
    s.energy();
    if (energies) energies[iteration - 1] = s.et;
    difference = std::fabs(s.et - old_energy);
    old_energy = s.et;
    if (iteration >= max_iter) break;
  }
  if (terms) s.terms(terms);
  if (P_out) std::memcpy(P_out, s.P.data(), sizeof(double) * s.nn);
  return iteration;
}

int or_scf(or_grid* g, int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
           const int* shell_L, const int* shell_nprim, const int* shell_offset, const double* alpha,
           const double* coef, double threshold, int min_iter, int max_iter, double mixing, double* energies,
           double* terms, double* P_out) {
  return scf_loop(g, natoms, Z, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef,
                  threshold, min_iter, max_iter, mixing, energies, terms, P_out, false);
}

int or_scf_eri(or_grid* g, int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
               const int* shell_L, const int* shell_nprim, const int* shell_offset, const double* alpha,
               const double* coef, double threshold, int min_iter, int max_iter, double mixing,
               double* energies, double* terms, double* P_out) {
  return scf_loop(g, natoms, Z, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef,
                  threshold, min_iter, max_iter, mixing, energies, terms, P_out, true);
}

int or_scf_evaluate(or_grid* g, int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
                    const int* shell_L, const int* shell_nprim, const int* shell_offset, const double* alpha,
                    const double* coef, const double* P, double hold_c, double* terms, double* F, double* H,
                    double* J, double* XC) {
  ScfCore s(g, natoms, Z, xyz, nshell, shell_atom, shell_L, shell_nprim, shell_offset, alpha, coef);
  std::memcpy(s.P.data(), P, sizeof(double) * s.nn);
  s.grid_density(hold_c);
  or_grid_hartree(g, s.J.data());
  or_grid_xc(g, s.XC.data(), &s.exc);
  s.energy();
  if (terms) {
    s.terms(terms);
    terms[5] = s.c;
  }
  if (F) {
    std::vector<double> f(s.nn);
    s.fock(f);
    std::memcpy(F, f.data(), sizeof(double) * s.nn);
  }
  if (H) std::memcpy(H, s.H.data(), sizeof(double) * s.nn);
  if (J) std::memcpy(J, s.J.data(), sizeof(double) * s.nn);
  if (XC) std::memcpy(XC, s.XC.data(), sizeof(double) * s.nn);
  return 0;
}

}  // extern "C"
