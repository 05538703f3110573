// Host-side tables the grid kernels consume, computed once per grid in FP64.
#ifndef DFTGRID_GRID_TABLES_H
#define DFTGRID_GRID_TABLES_H

#include <string>
#include <vector>

namespace dg {

// Sets the message dg_last_error() returns on this thread (library-internal).
void set_last_error(const std::string& msg);

constexpr double kPi = 3.14159265358979323846;
// The reference normalises primitives and evaluates nuclear attraction with
// this truncated value of pi (src/cgf.cpp:103, src/integrals.cpp:342).
constexpr double kPiTruncated = 3.14159265359;

// Gauss-Chebyshev (second kind) radial nodes mapped by r = (1+x)/(1-x), r_m = 1,
// and their weights before the Lebedev weight and the 4 pi r^2 Jacobian
// (src/atomicgrid.cpp:46-63).  r is decreasing.
void radial_grid(int n, std::vector<double>& r, std::vector<double>& w);

// Compiled-in Lebedev rule with npoints points (x, y, z, w; sum w = 1), or nullptr.
const double* lebedev_rule(int npoints);

// Real orthonormal spherical-harmonic normalisation of (l, m), Condon-Shortley
// phase carried by the Legendre function (src/spherical_harmonics.cpp:28-33).
double sh_prefactor(int l, int m);

// Y_lm at the nang Lebedev directions, lm = l^2 + l + m, laid out [j][lm].
void ylm_table(int nang, const double* leb, int lmax, std::vector<double>& y);

// FD radial Poisson matrix for angular momentum l on the nodes r (size n),
// (n+2)^2 row-major (src/atomicgrid.cpp:316-389, 419).
void poisson_matrix(int n, const double* r, int l, std::vector<double>& m);

// In-place inverse of a row-major n x n matrix by LU with partial pivoting.
// Returns false if a pivot is exactly zero.
bool invert_matrix(int n, std::vector<double>& a);

// Normalisation of a primitive Cartesian Gaussian x^l y^m z^n exp(-alpha r^2)
// (src/cgf.cpp:102-114, truncated pi).
double gto_norm(double alpha, int l, int m, int n);

// Coefficients of the not-a-knot spline's tridiagonal system for ascending
// knots x (src/cspline.cpp:76-121), reduced so that a right-hand side can be
// swept with  y'_i = (Y_i - a_i y'_{i-1}) / den_i  and  D_i = y'_i - cmod_i D_{i+1}.
struct SplineSystem {
  std::vector<double> a, den, cmod;
};
SplineSystem spline_system(const std::vector<double>& x);

// Matrix S (n x n, row-major) with D = S y: the not-a-knot slopes at the knots
// x of the spline through the values y (src/cspline.cpp:66-127 applied to the
// unit vectors; the slope system is linear in y).
std::vector<double> spline_slope_matrix(const std::vector<double>& x);

}  // namespace dg

#endif
