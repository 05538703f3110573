// Host SCF driver around the GPU grid library: one-electron integrals, the
// symmetric eigen-solver and the SCF loop of the reference's DFT class
// (src/dft.cpp:84-142, 160-227, 235-246, 300-366, 441-447), in FP64 on the host.
// The grid work of every iteration runs on the GPU through dg_build_device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dftcxx/molecular_grid.hpp"
#include "../../include/dftgrid.h"
#include "grid_tables.h"
#include "scf_device.h"

namespace dftcxx {
namespace {

// ---------------------------------------------------------------------------
// Boys function F_n(T) = int_0^1 t^{2n} exp(-T t^2) dt.  The argument is
// clamped to >= 1e-8 as the reference does (src/gamma.cpp:41-46).  Series
// e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)) for T < 40 with downward
// recursion, F_0 = sqrt(pi/T)/2 and upward recursion above.
// ---------------------------------------------------------------------------
void boys(int nmax, double T, double* F) {
  T = std::max(std::fabs(T), 1e-8);
  const double et = std::exp(-T);
  if (T < 40.0) {
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[nmax] = et * sum;
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
  } else {
    F[0] = 0.5 * std::sqrt(dg::kPi / T);
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
  }
}

// ---------------------------------------------------------------------------
// McMurchie-Davidson Hermite coefficients E^{ij}_t for one Cartesian direction
// ---------------------------------------------------------------------------
constexpr int kMaxE = 5;  // i, j up to 2 (+2 for the kinetic energy)
struct Herm {
  double e[kMaxE][kMaxE][2 * kMaxE];  // [i][j][t]
};
void hermite(int imax, int jmax, double a, double b, double xa, double xb, Herm& H) {
  const double p = a + b, mu = a * b / p, xab = xa - xb;
  const double xp = (a * xa + b * xb) / p, xpa = xp - xa, xpb = xp - xb, o2p = 0.5 / p;
  std::memset(&H, 0, sizeof(Herm));
  H.e[0][0][0] = std::exp(-mu * xab * xab);
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0) {
          const auto& s = H.e[i - 1][j];
          v = xpa * s[t] + (t + 1 < 2 * kMaxE ? (t + 1) * s[t + 1] : 0.0) + (t > 0 ? o2p * s[t - 1] : 0.0);
        } else {
          const auto& s = H.e[i][j - 1];
          v = xpb * s[t] + (t + 1 < 2 * kMaxE ? (t + 1) * s[t + 1] : 0.0) + (t > 0 ? o2p * s[t - 1] : 0.0);
        }
        H.e[i][j][t] = v;
      }
    }
}

const int kCart[3][6][3] = {{{0, 0, 0}},
                            {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
                            {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};

// Threads for the host algebra of large bases: OMP_NUM_THREADS when set, else
// the hardware's count, at most 16.
int host_threads() {
  static const int nt = [] {
    int t = (int)std::thread::hardware_concurrency();
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
      const int v = std::atoi(e);
      if (v > 0) t = v;
    }
    return std::max(1, std::min(t, 16));
  }();
  return nt;
}

// f(j0, j1) over the column ranges of [0, m), split over the host threads when
// the work (multiply-adds) is large.  A column is done by one thread in the
// serial order, so the result does not depend on the split; if a thread cannot
// be started its range runs here.
template <class F>
void parallel_cols(int m, double work, F f) {
  const int nt = work >= 2e6 ? std::min(host_threads(), m) : 1;
  if (nt <= 1) {
    f(0, m);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) {
    const int j0 = (int)((long)m * t / nt), j1 = (int)((long)m * (t + 1) / nt);
    try {
      th.emplace_back(f, j0, j1);
    } catch (...) {
      f(j0, j1);
    }
  }
  f(0, (int)((long)m / nt));
  for (auto& x : th) x.join();
}

// S, T, V (nuclear attraction of all nuclei) over contracted Cartesian
// Gaussians with primitive norms of src/cgf.cpp:102-114.  The reference's
// nuclear-attraction prefactor uses pi = 3.14159265359 (src/integrals.cpp:342,
// 386), reproduced by the factor kPiTruncated / kPi.
void one_electron(const MoleculeData& m, std::vector<double>& S, std::vector<double>& T,
                  std::vector<double>& V) {
  const int n = m.nbf(), ns = (int)m.shell_L.size();
  S.assign((size_t)n * n, 0.0);
  T.assign((size_t)n * n, 0.0);
  V.assign((size_t)n * n, 0.0);
  std::vector<int> off(ns);
  for (int s = 0, o = 0; s < ns; ++s) {
    off[s] = o;
    o += (m.shell_L[s] + 1) * (m.shell_L[s] + 2) / 2;
  }
  const double vfac = dg::kPiTruncated / dg::kPi;
  // shell rows (sa, all sb >= sa) split over the host threads for large bases; the
  // order pairs heavy early rows with light late ones so that ranges carry similar
  // work; every pair's block is computed as in the serial loop and written once
  std::vector<int> order(ns);
  for (int t = 0, lo = 0, hi = ns - 1; t < ns; ++t) order[t] = (t & 1) ? hi-- : lo++;
  parallel_cols(ns, 50.0 * ns * ns * m.natoms(), [&](int r0, int r1) {
    for (int rr = r0; rr < r1; ++rr) {
      const int sa = order[rr];
      for (int sb = sa; sb < ns; ++sb) {
        const int La = m.shell_L[sa], Lb = m.shell_L[sb];
        const int na = (La + 1) * (La + 2) / 2, nb = (Lb + 1) * (Lb + 2) / 2;
        const double* A = &m.xyz[3 * m.shell_atom[sa]];
        const double* B = &m.xyz[3 * m.shell_atom[sb]];
        std::vector<double> s_blk(na * nb, 0.0), t_blk(na * nb, 0.0), v_blk(na * nb, 0.0);
        for (int ka = 0; ka < m.shell_nprim[sa]; ++ka)
          for (int kb = 0; kb < m.shell_nprim[sb]; ++kb) {
            const double a = m.alpha[m.shell_offset[sa] + ka], b = m.alpha[m.shell_offset[sb] + kb];
            const double ca = m.coef[m.shell_offset[sa] + ka], cb = m.coef[m.shell_offset[sb] + kb];
            const double p = a + b;
            Herm H[3];
            for (int d = 0; d < 3; ++d) hermite(La, Lb + 2, a, b, A[d], B[d], H[d]);
            const double sq = std::sqrt(dg::kPi / p);
            double P[3];
            for (int d = 0; d < 3; ++d) P[d] = (a * A[d] + b * B[d]) / p;
            // R_tuv for every nucleus, t+u+v <= La+Lb
            const int Lt = La + Lb;
            std::vector<double> Rsum((Lt + 1) * (Lt + 1) * (Lt + 1), 0.0);
            for (int c = 0; c < m.natoms(); ++c) {
              const double* C = &m.xyz[3 * c];
              const double X = P[0] - C[0], Y = P[1] - C[1], Z = P[2] - C[2];
              double F[16];
              boys(Lt, p * (X * X + Y * Y + Z * Z), F);
              // R^n_{tuv}, built by the MD recursion from R^n_{000} = (-2p)^n F_n
              const int D = Lt + 1;
              std::vector<double> R((size_t)(Lt + 1) * D * D * D, 0.0);
              auto r = [&](int nn, int t, int u, int v) -> double& {
                return R[(((size_t)nn * D + t) * D + u) * D + v];
              };
              double f = 1.0;
              for (int nn = 0; nn <= Lt; ++nn) {
                r(nn, 0, 0, 0) = f * F[nn];
                f *= -2.0 * p;
              }
              for (int nn = Lt - 1; nn >= 0; --nn)
                for (int t = 0; t <= Lt - nn; ++t)
                  for (int u = 0; u <= Lt - nn - t; ++u)
                    for (int v = 0; v <= Lt - nn - t - u; ++v) {
                      if (t + u + v == 0) continue;
                      double val;
                      if (t > 0)
                        val = (t > 1 ? (t - 1) * r(nn + 1, t - 2, u, v) : 0.0) + X * r(nn + 1, t - 1, u, v);
                      else if (u > 0)
                        val = (u > 1 ? (u - 1) * r(nn + 1, t, u - 2, v) : 0.0) + Y * r(nn + 1, t, u - 1, v);
                      else
                        val = (v > 1 ? (v - 1) * r(nn + 1, t, u, v - 2) : 0.0) + Z * r(nn + 1, t, u, v - 1);
                      r(nn, t, u, v) = val;
                    }
              for (int t = 0; t <= Lt; ++t)
                for (int u = 0; u <= Lt - t; ++u)
                  for (int v = 0; v <= Lt - t - u; ++v)
                    Rsum[(t * (Lt + 1) + u) * (Lt + 1) + v] += -(double)m.Z[c] * r(0, t, u, v);
            }
            for (int ia = 0; ia < na; ++ia)
              for (int ib = 0; ib < nb; ++ib) {
                const int* la = kCart[La][ia];
                const int* lb = kCart[Lb][ib];
                const double Na = dg::gto_norm(a, la[0], la[1], la[2]);
                const double Nb = dg::gto_norm(b, lb[0], lb[1], lb[2]);
                const double pre = ca * cb * Na * Nb;
                double s1[3], t1[3];
                for (int d = 0; d < 3; ++d) {
                  const int i = la[d], j = lb[d];
                  s1[d] = H[d].e[i][j][0] * sq;
                  const double sp2 = H[d].e[i][j + 2][0] * sq;
                  const double sm2 = j >= 2 ? H[d].e[i][j - 2][0] * sq : 0.0;
                  t1[d] = -2.0 * b * b * sp2 + b * (2 * j + 1) * s1[d] - 0.5 * j * (j - 1) * sm2;
                }
                const double ov = s1[0] * s1[1] * s1[2];
                const double kin = t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2];
                double vn = 0.0;
                for (int t = 0; t <= la[0] + lb[0]; ++t)
                  for (int u = 0; u <= la[1] + lb[1]; ++u)
                    for (int v = 0; v <= la[2] + lb[2]; ++v)
                      vn += H[0].e[la[0]][lb[0]][t] * H[1].e[la[1]][lb[1]][u] * H[2].e[la[2]][lb[2]][v] *
                            Rsum[(t * (Lt + 1) + u) * (Lt + 1) + v];
                vn *= 2.0 * dg::kPi / p * vfac;
                s_blk[ia * nb + ib] += pre * ov;
                t_blk[ia * nb + ib] += pre * kin;
                v_blk[ia * nb + ib] += pre * vn;
              }
          }
        for (int ia = 0; ia < na; ++ia)
          for (int ib = 0; ib < nb; ++ib) {
            const size_t i = off[sa] + ia, j = off[sb] + ib;
            S[i + j * n] = S[j + i * n] = s_blk[ia * nb + ib];
            T[i + j * n] = T[j + i * n] = t_blk[ia * nb + ib];
            V[i + j * n] = V[j + i * n] = v_blk[ia * nb + ib];
          }
      }
    }
  });
}

// ---------------------------------------------------------------------------
// Symmetric eigen-solver: Householder reduction to tridiagonal form, then the
// implicit QL iteration with Wilkinson-type shifts (EISPACK tred2/tql2
// algorithms).  A column-major n x n; eigenvalues ascending, vectors in Z's columns.
// Every O(n^3) loop runs down contiguous columns (A v as column axpys, since A
// is symmetric; Z H as u = Z v, then rank-1 column updates).
// ---------------------------------------------------------------------------
void eigh(int n, const std::vector<double>& Ain, std::vector<double>& w, std::vector<double>& Z) {
  std::vector<double> A(Ain);
  auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * n]; };
  Z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] = 1.0;
  auto z = [&](int i, int j) -> double& { return Z[i + (size_t)j * n]; };
  std::vector<double> v(n), pv(n), qv(n), uv(n), d(n), e(n, 0.0);
  for (int k = 0; k + 2 < n; ++k) {
    double norm2 = 0.0;
    for (int i = k + 1; i < n; ++i) norm2 += a(i, k) * a(i, k);
    const double xnorm = std::sqrt(norm2);
    if (xnorm == 0.0) continue;
    const double alpha = a(k + 1, k) > 0 ? -xnorm : xnorm;
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
    v[k + 1] -= alpha;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += v[i] * v[i];
    if (vn2 == 0.0) continue;
    const double inv = 1.0 / std::sqrt(vn2);
    for (int i = k + 1; i < n; ++i) v[i] *= inv;
    // A <- H A H on the trailing block, H = I - 2 v v^T
    for (int i = k + 1; i < n; ++i) pv[i] = 0.0;
    for (int j = k + 1; j < n; ++j) {
      const double vj = v[j];
      const double* col = &a(0, j);
      for (int i = k + 1; i < n; ++i) pv[i] += col[i] * vj;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; ++i) K += v[i] * pv[i];
    for (int i = k + 1; i < n; ++i) qv[i] = pv[i] - K * v[i];
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i) a(i, j) -= 2.0 * (v[i] * qv[j] + qv[i] * v[j]);
    a(k + 1, k) = a(k, k + 1) = alpha;
    for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
    // Z <- Z H = Z - 2 (Z v) v^T
    for (int i = 0; i < n; ++i) uv[i] = 0.0;
    for (int j = k + 1; j < n; ++j) {
      const double vj = v[j];
      const double* col = &z(0, j);
      for (int i = 0; i < n; ++i) uv[i] += col[i] * vj;
    }
    for (int j = k + 1; j < n; ++j) {
      const double vj = v[j];
      double* col = &z(0, j);
      for (int i = 0; i < n; ++i) col[i] -= 2.0 * uv[i] * vj;
    }
  }
  for (int i = 0; i < n; ++i) d[i] = a(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
  // implicit QL; e[i] couples d[i] and d[i+1]
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 200; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          const double fz = z(k, i + 1);
          z(k, i + 1) = s * z(k, i) + c * fz;
          z(k, i) = c * z(k, i) - s * fz;
        }
      }
      if (underflow) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(n);
  std::vector<double> Zs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    std::memcpy(&Zs[(size_t)j * n], &Z[(size_t)idx[j] * n], sizeof(double) * n);
  }
  Z.swap(Zs);
}

// Lowest nv eigenpairs of a symmetric n x n matrix (column-major): Householder
// tridiagonalisation T = Q^T A Q with the reflections kept (Q = H_0 ... H_{n-3}),
// the eigenvalues of T by the implicit QL iteration without vectors (the same
// iteration as eigh, so the same values), the eigenvectors of T for the nv
// lowest by inverse iteration on T - lambda I (LU with partial pivoting, three
// solves, Gram-Schmidt against the earlier vectors of a cluster of eigenvalues
// closer than 1e-3 |T|_1), and x = Q y.  Z holds the nv vectors as columns.
void eigh_lowest(int n, int nv, const std::vector<double>& Ain, std::vector<double>& w, std::vector<double>& Z) {
  std::vector<double> A(Ain);
  auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * n]; };
  const int nh = std::max(0, n - 2);
  std::vector<double> V((size_t)n * std::max(nh, 1), 0.0);  // reflection k in column k (zero: none)
  std::vector<double> pv(n), qv(n), d(n), e(n, 0.0);
  for (int k = 0; k + 2 < n; ++k) {
    double* v = &V[(size_t)k * n];
    double norm2 = 0.0;
    for (int i = k + 1; i < n; ++i) norm2 += a(i, k) * a(i, k);
    const double xnorm = std::sqrt(norm2);
    if (xnorm == 0.0) continue;
    const double alpha = a(k + 1, k) > 0 ? -xnorm : xnorm;
    for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
    v[k + 1] -= alpha;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += v[i] * v[i];
    if (vn2 == 0.0) {
      for (int i = k + 1; i < n; ++i) v[i] = 0.0;
      continue;
    }
    const double inv = 1.0 / std::sqrt(vn2);
    for (int i = k + 1; i < n; ++i) v[i] *= inv;
    for (int i = k + 1; i < n; ++i) pv[i] = 0.0;
    for (int j = k + 1; j < n; ++j) {
      const double vj = v[j];
      const double* col = &a(0, j);
      for (int i = k + 1; i < n; ++i) pv[i] += col[i] * vj;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; ++i) K += v[i] * pv[i];
    for (int i = k + 1; i < n; ++i) qv[i] = pv[i] - K * v[i];
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i) a(i, j) -= 2.0 * (v[i] * qv[j] + qv[i] * v[j]);
    a(k + 1, k) = a(k, k + 1) = alpha;
    for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
  }
  for (int i = 0; i < n; ++i) d[i] = a(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
  const std::vector<double> td(d), te(e);  // T for the inverse iteration
  // eigenvalues: implicit QL without vectors
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 200; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-16 * dd) break;
      }
      if (m == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
      }
      if (underflow) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::sort(d.begin(), d.end());
  w.assign(d.begin(), d.begin() + nv);
  // inverse iteration on T
  double tnorm = 0.0;  // |T|_1
  for (int i = 0; i < n; ++i)
    tnorm = std::max(tnorm, std::fabs(td[i]) + (i > 0 ? std::fabs(te[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(te[i]) : 0.0));
  if (tnorm == 0.0) tnorm = 1.0;
  const double eps = 2.220446049250313e-16, ortol = 1e-3 * tnorm, pertol = 10.0 * eps * tnorm;
  Z.assign((size_t)n * nv, 0.0);
  std::vector<double> dl(n), du(n), du2(n), dg(n), y(n);
  std::vector<int> piv(n);
  int c0 = 0;  // first vector of the current cluster
  double xprev = 0.0;
  uint64_t seed = 0x9e3779b97f4a7c15ull;
  for (int j = 0; j < nv; ++j) {
    double x = w[j];
    if (j > 0) {
      if (x - w[j - 1] > ortol) c0 = j;
      if (x - xprev < pertol) x = xprev + pertol;  // separate the shifts within a cluster
    }
    xprev = x;
    // LU of T - x I with partial pivoting (rows i, i+1 swapped when |sub| > |diag|)
    for (int i = 0; i < n; ++i) dg[i] = td[i] - x;
    for (int i = 0; i + 1 < n; ++i) du[i] = te[i], dl[i] = te[i];
    for (int i = 0; i + 2 < n; ++i) du2[i] = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
      if (std::fabs(dg[i]) >= std::fabs(dl[i])) {
        piv[i] = 0;
        if (dg[i] == 0.0) dg[i] = pertol;
        const double f = dl[i] / dg[i];
        dl[i] = f;
        dg[i + 1] -= f * du[i];
      } else {
        piv[i] = 1;
        const double f = dg[i] / dl[i];
        dg[i] = dl[i];
        dl[i] = f;
        const double t = du[i];
        du[i] = dg[i + 1];
        dg[i + 1] = t - f * dg[i + 1];
        if (i + 2 < n) {
          du2[i] = du[i + 1];
          du[i + 1] = -f * du[i + 1];
        }
      }
    }
    if (dg[n - 1] == 0.0) dg[n - 1] = pertol;
    // deterministic start vector
    for (int i = 0; i < n; ++i) {
      seed = seed * 6364136223846793005ull + 1442695040888963407ull;
      y[i] = (double)(seed >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    for (int it = 0; it < 3; ++it) {
      // solve L U y = P b: forward with the row swaps, then back substitution
      for (int i = 0; i + 1 < n; ++i) {
        if (piv[i]) {
          const double t = y[i];
          y[i] = y[i + 1];
          y[i + 1] = t - dl[i] * y[i];
        } else {
          y[i + 1] -= dl[i] * y[i];
        }
      }
      y[n - 1] /= dg[n - 1];
      if (n > 1) y[n - 2] = (y[n - 2] - du[n - 2] * y[n - 1]) / dg[n - 2];
      for (int i = n - 3; i >= 0; --i) y[i] = (y[i] - du[i] * y[i + 1] - du2[i] * y[i + 2]) / dg[i];
      // orthogonalise against the cluster's earlier vectors (still in T's basis: column j of Z)
      for (int q = c0; q < j; ++q) {
        const double* zq = &Z[(size_t)q * n];
        double dot = 0.0;
        for (int i = 0; i < n; ++i) dot += zq[i] * y[i];
        for (int i = 0; i < n; ++i) y[i] -= dot * zq[i];
      }
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; ++i) y[i] /= nrm;
    }
    std::memcpy(&Z[(size_t)j * n], y.data(), sizeof(double) * n);
  }
  // back-transformation x = H_0 (H_1 (... H_{n-3} y)), vectors independent of each other
  const double* Vd = V.data();
  double* Zd = Z.data();
  parallel_cols(nv, 2.0 * n * n * nv, [=](int j0, int j1) {
    for (int j = j0; j < j1; ++j) {
      double* z = Zd + (size_t)j * n;
      for (int k = nh - 1; k >= 0; --k) {
        const double* v = Vd + (size_t)k * n;
        double dot = 0.0;
        for (int i = k + 1; i < n; ++i) dot += v[i] * z[i];
        const double f = 2.0 * dot;
        for (int i = k + 1; i < n; ++i) z[i] -= f * v[i];
      }
    }
  });
}

// C = A B, column-major, A n x n, B n x m, each column of C as axpys over
// contiguous columns of A (every C_ij is summed over k in ascending order)
void gemm(int n, const double* A, const double* B, double* C, int m = -1) {
  if (m < 0) m = n;
  parallel_cols(m, (double)n * n * m, [=](int j0, int j1) {
    for (int j = j0; j < j1; ++j) {
      double* c = C + (size_t)j * n;
      for (int i = 0; i < n; ++i) c[i] = 0.0;
      for (int k = 0; k < n; ++k) {
        const double b = B[k + (size_t)j * n];
        const double* a = A + (size_t)k * n;
        for (int i = 0; i < n; ++i) c[i] += a[i] * b;
      }
    }
  });
}

// Host algebra of the reference's DFT class (src/dft.h:41-216): integrals,
// orthogonalisation, the density matrix with mixing and the energy.  The grid
// results (J, XC, scalars) are supplied by a single-molecule handle (DFT) or a
// batch (batch_scf).
class HostSCF {
 public:
  explicit HostSCF(const MoleculeData& mol, bool mixed = false) : mol_(mol), mixed_(mixed) {
    n_ = mol.nbf();
    nelec_ = mol.nelec();
  }

  // src/dft.cpp:160-227 without the grid part
  void construct_matrices() {
    std::vector<double> Tm, Vm;
    one_electron(mol_, S_, Tm, Vm);
    H_.resize(S_.size());
    for (size_t i = 0; i < S_.size(); ++i) H_[i] = Tm[i] + Vm[i];
    enuc_ = 0.0;
    for (int i = 0; i < mol_.natoms(); ++i)
      for (int j = i + 1; j < mol_.natoms(); ++j) {
        const double dx = mol_.xyz[3 * i] - mol_.xyz[3 * j], dy = mol_.xyz[3 * i + 1] - mol_.xyz[3 * j + 1],
                     dz = mol_.xyz[3 * i + 2] - mol_.xyz[3 * j + 2];
        enuc_ += mol_.Z[i] * mol_.Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    // canonical orthogonalisation X = U s^{-1/2}, src/dft.cpp:300-316
    std::vector<double> s, U;
    eigh(n_, S_, s, U);
    X_.assign((size_t)n_ * n_, 0.0);
    for (int j = 0; j < n_; ++j) {
      if (!(s[j] > 0.0)) throw std::runtime_error("overlap matrix is not positive definite");
      const double f = 1.0 / std::sqrt(s[j]);
      for (int i = 0; i < n_; ++i) X_[i + (size_t)j * n_] = U[i + (size_t)j * n_] * f;
    }
    XT_.resize(X_.size());
    for (int j = 0; j < n_; ++j)
      for (int i = 0; i < n_; ++i) XT_[j + (size_t)i * n_] = X_[i + (size_t)j * n_];
    J_.assign((size_t)n_ * n_, 0.0);
    XC_.assign((size_t)n_ * n_, 0.0);
    P_.assign((size_t)n_ * n_, 0.0);
    exc_ = 0.0;  // the reference reads it before the first XC evaluation
  }

  // src/dft.cpp:330-366
  void density_matrix() {
    const size_t nn = (size_t)n_ * n_;
    std::vector<double> F(nn), tmp(nn), Fp(nn), Cc, C(nn), w;
    for (size_t i = 0; i < nn; ++i) F[i] = H_[i] + 2.0 * J_[i] + XC_[i];
    gemm(n_, XT_.data(), F.data(), tmp.data());
    gemm(n_, tmp.data(), X_.data(), Fp.data());
    // only the occupied orbitals enter P
    const int nocc = nelec_ / 2;
    eigh_lowest(n_, nocc, Fp, w, Cc);
    gemm(n_, X_.data(), Cc.data(), C.data(), nocc);  // C_occ = X Cc
    // P = C_occ C_occ^T, each element summed over the occupied orbitals in order
    std::vector<double> Pn(nn, 0.0);
    const int n = n_;
    const double* Cd = C.data();
    double* Pd = Pn.data();
    parallel_cols(n, (double)n * n * nocc, [=](int j0, int j1) {
      for (int j = j0; j < j1; ++j) {
        double* pcol = Pd + (size_t)j * n;
        for (int k = 0; k < nocc; ++k) {
          const double* ck = Cd + (size_t)k * n;
          const double cj = ck[j];
          for (int i = 0; i < n; ++i) pcol[i] += ck[i] * cj;
        }
      }
    });
    if (first_) {
      P_ = Pn;
      first_ = false;
    } else {
      for (size_t i = 0; i < nn; ++i) P_[i] = (1.0 - mixing) * Pn[i] + mixing * P_[i];
    }
  }

  // results of one grid build: J, XC (taken only with_xc) and {E_xc, N before, N after, E_J}
  void take_grid(const double* J, const double* XC, const double* sc, bool with_xc) {
    const size_t nn = (size_t)n_ * n_;
    std::memcpy(J_.data(), J, nn * 8);
    nelec_grid_ = sc[2];
    ej_point_ = sc[3];
    if (with_xc) {
      std::memcpy(XC_.data(), XC, nn * 8);
      exc_ = sc[0];
    }
  }

  // src/dft.cpp:441-447
  void energy() {
    double t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < n_; ++i)
      for (int k = 0; k < n_; ++k) {
        t1 += P_[i + (size_t)k * n_] * H_[k + (size_t)i * n_];
        t2 += P_[i + (size_t)k * n_] * J_[k + (size_t)i * n_];
      }
    e_one_ = 2.0 * t1;
    // In reduced precision J carries FP32-product errors that would enter 2 tr(PJ)
    // at first order; the same quantity evaluated pointwise in FP64 does not.
    // A J from the two-electron integrals is FP64 in either mode.
    e_j_ = (mixed_ && !eri_mode_) ? ej_point_ : 2.0 * t2;
    et_ = e_one_ + e_j_ + enuc_ + exc_;
  }

  void fill(dg_scf_result* r, int iterations, double setup, double sec_iter, bool converged,
            double delta) const {
    r->iterations = iterations;
    r->converged = converged ? 1 : 0;
    r->device_step = 0;
    r->host_fallbacks = 0;
    r->delta = delta;
    r->energy = et_;
    r->e_one = e_one_;
    r->e_j = e_j_;
    r->e_xc = exc_;
    r->e_nuc = enuc_;
    r->nelec = nelec_grid_;
    r->seconds_setup = setup;
    r->seconds_iter = sec_iter;
  }

  // for a step formed elsewhere (the batched device step): later steps mix
  void mark_started() { first_ = false; }
  bool first() const { return first_; }
  int nocc() const { return nelec_ / 2; }
  const std::vector<double>& H() const { return H_; }
  const std::vector<double>& X() const { return X_; }
  std::vector<double>& J() { return J_; }
  std::vector<double>& XC() { return XC_; }

  double mixing = 0.5;
  double et_ = 0.0, e_one_ = 0.0, e_j_ = 0.0, exc_ = 0.0, enuc_ = 0.0, nelec_grid_ = 0.0;
  std::vector<double> P_;
  int n_ = 0;

 protected:
  MoleculeData mol_;
  bool mixed_ = false;
  bool eri_mode_ = false;  // J from the analytic two-electron integrals
  double ej_point_ = 0.0;
  int nelec_ = 0;
  bool first_ = true;
  std::vector<double> S_, H_, X_, XT_, J_, XC_;
};

// The reference's DFT class with the grid work of every iteration on the GPU
// (dg_build_device on one molecule's handle).  With on_device, the
// density-matrix step and the energy run on the GPU as well (DeviceStep:
// F, X^T F X, the occupied projector by SP2, X P' X^T, mixing, energy), so
// an iteration moves only the energy terms to the host; a projection that
// does not converge (no HOMO-LUMO gap) is redone by the host eigen-solver.
class DFT : public HostSCF {
 public:
  DFT(const MoleculeData& mol, const dg_grid_params& grid, bool on_device, bool eri)
      : HostSCF(mol, grid.precision == DG_PREC_MIXED), grid_(mol), on_device_(on_device) {
    eri_mode_ = eri;
    grid_.set_grid_points(grid.n_radial, grid.n_angular, grid.lmax, grid.precision, grid.screen);
    grid_.create_grid();
  }
  ~DFT() {
    ds_.reset();
    if (eri_) dg_eri_destroy(eri_);
    if (dP_) (void)hipFree(dP_);
    if (hbuf_) (void)hipHostFree(hbuf_);
  }

  void construct_matrices() {
    HostSCF::construct_matrices();
    if (eri_mode_) {  // DFT::construct_matrices computes every ERI once (src/dft.cpp:204-211)
      dg_molecule m = mol_.c_molecule();
      dg_basis b = mol_.c_basis();
      dg_throw(dg_eri_create(&m, &b, &eri_));
    }
    if (on_device_) {
      void* sp = nullptr;
      dg_throw(dg_get_stream(grid_.handle(), &sp));
      ds_.reset(new DeviceStep(n_, nelec_ / 2, (hipStream_t)sp));
      ds_->upload_core(H_.data(), X_.data());
      device_iteration(false);
      return;
    }
    if (hipMalloc((void**)&dP_, sizeof(double) * (3 * (size_t)n_ * n_ + 4)) != hipSuccess)
      throw std::runtime_error("hipMalloc failed");
    // pinned staging for P, J, XC and the scalars (page-locked copies each iteration)
    if (hipHostMalloc((void**)&hbuf_, sizeof(double) * (3 * (size_t)n_ * n_ + 4), hipHostMallocDefault) !=
        hipSuccess)
      throw std::runtime_error("hipHostMalloc failed");
    density_matrix();
    grid_build(false);
    energy();
  }

  // src/dft.cpp:84-142
  // converged: the loop ended on |dE| <= threshold, not at max_iter (src/dft.cpp:126-131)
  int scf(double threshold, int min_iter, int max_iter, double* energies, double* sec_iter, bool* converged,
          double* delta) {
    double old_energy = et_, difference = 1.0;
    int iteration = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (difference > threshold || iteration < min_iter) {
      iteration++;
      if (on_device_) {
        device_iteration(true);
      } else {
        density_matrix();
        grid_build(true);
        energy();
      }
      if (energies) energies[iteration - 1] = et_;
      difference = std::fabs(et_ - old_energy);
      old_energy = et_;
      if (iteration >= max_iter) break;
    }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (sec_iter) *sec_iter = iteration ? dt / iteration : 0.0;
    *converged = difference <= threshold && iteration >= min_iter;
    *delta = difference;
    if (on_device_) ds_->download(P_.data(), ds_->dP());
    return iteration;
  }

  int host_fallbacks = 0;  // density steps the host eigen-solver redid

 private:
  // One iteration with everything on the device: density step, grid build, energy.
  // The set-up build (no XC) forms J alone; every later build forms
  // G = 2J + XC in one contraction (dg_build_fock_device), the only part of
  // them the next density step and the energy (E_J pointwise) use.
  void device_iteration(bool with_xc) {
    if (!ds_->density(with_xc, first_, mixing, fused_)) {
      // no gap (degenerate frontier orbitals): this step by the host eigen-solver
      // from the same F, whose choice within the degenerate space is the one the
      // host path makes
      ++host_fallbacks;
      ds_->download(P_.data(), ds_->dP());
      if (fused_) {  // dXC holds G = 2J + XC
        std::fill(J_.begin(), J_.end(), 0.0);
        ds_->download(XC_.data(), ds_->dXC());
      } else {
        ds_->download(J_.data(), ds_->dJ());
        if (with_xc) ds_->download(XC_.data(), ds_->dXC());
      }
      HostSCF::density_matrix();
      ds_->upload(ds_->dP(), P_.data());
    }
    device_build(with_xc);
    ds_->finish(&e_one_, &e_j_, &exc_, &et_, &nelec_grid_);
    first_ = false;
    fused_ = with_xc && !eri_mode_;
  }

  void device_build(bool with_xc) {
    if (eri_mode_) {  // J from the integrals, the grid for rho and XC
      void* sp = nullptr;
      dg_throw(dg_get_stream(grid_.handle(), &sp));
      dg_throw(dg_eri_coulomb_device(eri_, ds_->dP(), ds_->dJ(), sp));
      dg_throw(dg_build_xc_device(grid_.handle(), ds_->dP(), ds_->dXC(), ds_->dscal()));
      if (!with_xc) ds_->clear_xc();
      ds_->energy(enuc_, with_xc, false);
      return;
    }
    if (with_xc)
      dg_throw(dg_build_fock_device(grid_.handle(), ds_->dP(), ds_->dXC(), ds_->dscal()));
    else
      dg_throw(dg_build_device(grid_.handle(), ds_->dP(), ds_->dJ(), nullptr, ds_->dscal()));
    ds_->energy(enuc_, with_xc, with_xc || mixed_);
  }

  // set_density + correct_densities + Hartree J (+ XC) on the GPU
  void grid_build(bool with_xc) {
    const size_t nn = (size_t)n_ * n_;
    double* dJ = dP_ + nn;
    double* dXC = dJ + nn;
    double* dS = dXC + nn;
    void* sp = nullptr;
    dg_throw(dg_get_stream(grid_.handle(), &sp));
    hipStream_t st = (hipStream_t)sp;
    double* hP = hbuf_;
    double* hJ = hP + nn;  // then XC and the scalars, contiguous like dJ, dXC, dS
    std::memcpy(hP, P_.data(), nn * 8);
    if (hipMemcpyAsync(dP_, hP, nn * 8, hipMemcpyHostToDevice, st) != hipSuccess)
      throw std::runtime_error("copy of P failed");
    if (eri_mode_) {
      dg_throw(dg_eri_coulomb_device(eri_, dP_, dJ, st));
      dg_throw(dg_build_xc_device(grid_.handle(), dP_, dXC, dS));
    } else {
      dg_throw(dg_build_device(grid_.handle(), dP_, dJ, dXC, dS));
    }
    if (hipMemcpyAsync(hJ, dJ, (2 * nn + 4) * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      throw std::runtime_error("copy of grid results failed");
    take_grid(hJ, hJ + nn, hJ + 2 * nn, with_xc);
  }

  MolecularGrid grid_;
  bool on_device_ = false;
  bool fused_ = false;  // the last device build formed G = 2J + XC
  dg_eri* eri_ = nullptr;
  std::unique_ptr<DeviceStep> ds_;
  double* dP_ = nullptr;
  double* hbuf_ = nullptr;
};

// Batched SCF: every molecule runs the reference's loop (src/dft.cpp:84-142)
// with its own convergence test; the grid work of one iteration of all
// molecules is one dg_batch_build.  A converged molecule keeps its density
// matrix and results (its later builds are discarded), so its energy sequence
// is the one a single-molecule run produces.
int batch_scf(std::vector<HostSCF>& mols, dg_batch* batch, const dg_scf_params& prm, int* iterations,
              double* sec_iter, std::vector<char>& converged, std::vector<double>& diff) {
  const int nm = (int)mols.size();
  std::vector<int> nbf(nm);
  std::vector<long> off(nm + 1);
  dg_throw(dg_batch_layout(batch, nbf.data(), off.data(), nullptr));
  std::vector<double> P(off[nm]), J(off[nm]), XC(off[nm]), sc(4 * (size_t)nm);
  auto build = [&](bool with_xc, const std::vector<char>& active) {
    for (int m = 0; m < nm; ++m) std::memcpy(&P[off[m]], mols[m].P_.data(), (size_t)nbf[m] * nbf[m] * 8);
    dg_throw(dg_batch_build(batch, P.data(), J.data(), XC.data(), sc.data()));
    for (int m = 0; m < nm; ++m)
      if (active[m]) {
        mols[m].take_grid(&J[off[m]], &XC[off[m]], &sc[4 * (size_t)m], with_xc);
        mols[m].energy();
      }
  };
  // the molecules' density-matrix steps are independent: split over the host
  // threads (each molecule's result is the serial one); an error is rethrown here
  double work = 0.0;
  for (int m = 0; m < nm; ++m) work += (double)nbf[m] * nbf[m] * nbf[m];
  auto densities = [&](const std::vector<char>& act) {
    std::exception_ptr err = nullptr;
    std::mutex mu;
    parallel_cols(nm, work, [&](int m0, int m1) {
      try {
        for (int m = m0; m < m1; ++m)
          if (act[m]) mols[m].density_matrix();
      } catch (...) {
        std::lock_guard<std::mutex> lock(mu);
        if (!err) err = std::current_exception();
      }
    });
    if (err) std::rethrow_exception(err);
  };
  std::vector<char> active(nm, 1);
  densities(active);
  build(false, active);
  std::vector<double> old(nm);
  diff.assign(nm, 1.0);
  converged.assign(nm, 0);
  for (int m = 0; m < nm; ++m) old[m] = mols[m].et_;
  int iteration = 0, done = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (done < nm) {
    iteration++;
    densities(active);
    build(true, active);
    for (int m = 0; m < nm; ++m) {
      if (!active[m]) continue;
      diff[m] = std::fabs(mols[m].et_ - old[m]);
      old[m] = mols[m].et_;
      iterations[m] = iteration;
      converged[m] = diff[m] <= prm.threshold && iteration >= prm.min_iter;
      if (converged[m] || iteration >= prm.max_iter) {
        active[m] = 0;
        ++done;
      }
    }
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (sec_iter) *sec_iter = iteration ? dt / iteration : 0.0;
  return iteration;
}

// The batched SCF with every molecule's density-matrix step and energy on the
// GPU (BatchStep, one work-group per molecule): per iteration the step, one
// batch build (a graph launch) and the energies, with only the energies and the
// per-molecule projection flags read back; a molecule whose projection has no
// gap has that step formed on the host.  Same loop and convergence rule as
// batch_scf.
int batch_scf_device(std::vector<HostSCF>& mols, dg_batch* batch, const dg_scf_params& prm, int* iterations,
                     double* sec_iter, std::vector<char>& converged, std::vector<double>& diff, int* fallbacks) {
  const int nm = (int)mols.size();
  std::vector<int> nbf(nm), nocc(nm);
  std::vector<long> off(nm + 1);
  dg_throw(dg_batch_layout(batch, nbf.data(), off.data(), nullptr));
  for (int m = 0; m < nm; ++m) nocc[m] = mols[m].nocc();
  dg_handle* h = nullptr;
  dg_throw(dg_batch_handle(batch, &h));
  void* sp = nullptr;
  dg_throw(dg_get_stream(h, &sp));
  BatchStep bs(nm, nbf.data(), off.data(), nocc.data(), (hipStream_t)sp);
  {
    std::vector<double> Hp(off[nm]), Xp(off[nm]), en(nm);
    for (int m = 0; m < nm; ++m) {
      std::memcpy(&Hp[off[m]], mols[m].H().data(), sizeof(double) * nbf[m] * nbf[m]);
      std::memcpy(&Xp[off[m]], mols[m].X().data(), sizeof(double) * nbf[m] * nbf[m]);
      en[m] = mols[m].enuc_;
    }
    bs.upload_core(Hp.data(), Xp.data(), en.data());
  }
  std::vector<char> active(nm, 1);
  bs.set_active(active.data());
  *fallbacks = 0;
  auto step = [&](bool with_xc, bool first) {
    if (bs.density(with_xc, first, mols[0].mixing) > 0) {
      const int* f = bs.failed();
      for (int m = 0; m < nm; ++m) {
        if (!active[m] || !f[m]) continue;
        ++*fallbacks;
        const size_t nn = (size_t)nbf[m] * nbf[m];
        bs.download(mols[m].P_.data(), bs.dP() + off[m], nn);
        bs.download(mols[m].J().data(), bs.dJ() + off[m], nn);
        if (with_xc) bs.download(mols[m].XC().data(), bs.dXC() + off[m], nn);
        mols[m].density_matrix();
        bs.upload(bs.dP() + off[m], mols[m].P_.data(), nn);
      }
    }
    for (HostSCF& hm : mols) hm.mark_started();
    dg_throw(dg_batch_build_device(batch, bs.dP(), bs.dJ(), bs.dXC(), bs.dscal()));
    if (!with_xc) bs.clear_xc();  // the set-up step has no XC yet (src/dft.cpp:218-226)
    bs.energy(with_xc);
    bs.finish();
    const double* e = bs.energies();
    for (int m = 0; m < nm; ++m) {
      if (!active[m]) continue;
      mols[m].e_one_ = e[5 * m];
      mols[m].e_j_ = e[5 * m + 1];
      mols[m].exc_ = e[5 * m + 2];
      mols[m].et_ = e[5 * m + 3];
      mols[m].nelec_grid_ = e[5 * m + 4];
    }
  };
  step(false, true);
  std::vector<double> old(nm);
  diff.assign(nm, 1.0);
  converged.assign(nm, 0);
  for (int m = 0; m < nm; ++m) old[m] = mols[m].et_;
  int iteration = 0, done = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (done < nm) {
    iteration++;
    step(true, false);
    bool changed = false;
    for (int m = 0; m < nm; ++m) {
      if (!active[m]) continue;
      diff[m] = std::fabs(mols[m].et_ - old[m]);
      old[m] = mols[m].et_;
      iterations[m] = iteration;
      converged[m] = diff[m] <= prm.threshold && iteration >= prm.min_iter;
      if (converged[m] || iteration >= prm.max_iter) {
        active[m] = 0;
        changed = true;
        ++done;
      }
    }
    if (changed) bs.set_active(active.data());
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (sec_iter) *sec_iter = iteration ? dt / iteration : 0.0;
  return iteration;
}

}  // namespace
}  // namespace dftcxx

namespace {
dftcxx::MoleculeData to_data(const dg_molecule* mol, const dg_basis* basis) {
  dftcxx::MoleculeData m;
  m.Z.assign(mol->Z, mol->Z + mol->natoms);
  m.xyz.assign(mol->xyz, mol->xyz + 3 * mol->natoms);
  int nprim = 0;
  for (int s = 0; s < basis->nshell; ++s) nprim = std::max(nprim, basis->offset[s] + basis->nprim[s]);
  m.shell_atom.assign(basis->atom, basis->atom + basis->nshell);
  m.shell_L.assign(basis->L, basis->L + basis->nshell);
  m.shell_nprim.assign(basis->nprim, basis->nprim + basis->nshell);
  m.shell_offset.assign(basis->offset, basis->offset + basis->nshell);
  m.alpha.assign(basis->alpha, basis->alpha + nprim);
  m.coef.assign(basis->coef, basis->coef + nprim);
  for (int s = 0; s < basis->nshell; ++s) {
    if (m.shell_L[s] < 0 || m.shell_L[s] > 2) throw std::runtime_error("shell L must be 0, 1 or 2");
    if (m.shell_atom[s] < 0 || m.shell_atom[s] >= mol->natoms) throw std::runtime_error("shell atom out of range");
  }
  return m;
}
// The reference hard-codes these (src/dft.cpp:85, 95, 331); as parameters they are checked.
bool check_scf_params(const dg_scf_params* p, const char* who) {
  const char* bad = nullptr;
  if (p->max_iter < 1) bad = "max_iter must be >= 1";
  else if (p->min_iter < 0) bad = "min_iter must be >= 0";
  else if (!std::isfinite(p->threshold) || p->threshold < 0.0) bad = "threshold must be finite and >= 0";
  else if (!(p->mixing >= 0.0 && p->mixing < 1.0)) bad = "mixing must lie in [0, 1)";
  else if (p->density_step < DG_STEP_AUTO || p->density_step > DG_STEP_DEVICE)
    bad = "density_step must be DG_STEP_AUTO, DG_STEP_HOST or DG_STEP_DEVICE";
  else if (p->hartree != DG_HARTREE_BECKE && p->hartree != DG_HARTREE_ERI)
    bad = "hartree must be DG_HARTREE_BECKE or DG_HARTREE_ERI";
  if (bad) dg::set_last_error(std::string(who) + ": " + bad);
  return bad == nullptr;
}
}  // namespace

extern "C" int dg_one_electron(const dg_molecule* mol, const dg_basis* basis, double* S, double* T, double* V) {
  if (!mol || !basis || mol->natoms < 1 || basis->nshell < 1) {
    dg::set_last_error("dg_one_electron: empty molecule or basis");
    return DG_ERR_INVALID;
  }
  try {
    dftcxx::MoleculeData m = to_data(mol, basis);
    std::vector<double> s, t, v;
    dftcxx::one_electron(m, s, t, v);
    if (S) std::memcpy(S, s.data(), s.size() * 8);
    if (T) std::memcpy(T, t.data(), t.size() * 8);
    if (V) std::memcpy(V, v.data(), v.size() * 8);
    return DG_OK;
  } catch (const std::exception& e) {
    dg::set_last_error(std::string("dg_one_electron: ") + e.what());
    return DG_ERR_INVALID;
  }
}

extern "C" int dg_eigh(int n, const double* A, double* w, double* Z) {
  if (n < 1 || !A || !w || !Z) {
    dg::set_last_error("dg_eigh: bad argument");
    return DG_ERR_INVALID;
  }
  std::vector<double> a(A, A + (size_t)n * n), ww, zz;
  dftcxx::eigh(n, a, ww, zz);
  std::memcpy(w, ww.data(), (size_t)n * 8);
  std::memcpy(Z, zz.data(), (size_t)n * n * 8);
  return DG_OK;
}

extern "C" int dg_scf_run(const dg_molecule* mol, const dg_basis* basis, const dg_grid_params* grid,
                          const dg_scf_params* scf, dg_scf_result* result, double* energies, double* P_out) {
  if (!mol || !basis || !grid || !scf || !result) {
    dg::set_last_error("dg_scf_run: null argument");
    return DG_ERR_INVALID;
  }
  if (!check_scf_params(scf, "dg_scf_run")) return DG_ERR_INVALID;
  try {
    dftcxx::MoleculeData m = to_data(mol, basis);
    if (m.nelec() % 2 != 0) throw std::runtime_error("closed-shell SCF needs an even electron count");
    const auto t0 = std::chrono::steady_clock::now();
    const bool dev = scf->density_step == DG_STEP_DEVICE ||
                     (scf->density_step == DG_STEP_AUTO && m.nbf() >= DG_STEP_AUTO_NBF);
    dftcxx::DFT dft(m, *grid, dev, scf->hartree == DG_HARTREE_ERI);
    dft.mixing = scf->mixing;
    dft.construct_matrices();
    const double setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double sec_iter = 0.0, delta = 0.0;
    bool converged = false;
    const int it = dft.scf(scf->threshold, scf->min_iter, scf->max_iter, energies, &sec_iter, &converged, &delta);
    dft.fill(result, it, setup, sec_iter, converged, delta);
    result->device_step = dev ? 1 : 0;
    result->host_fallbacks = dft.host_fallbacks;
    if (P_out) std::memcpy(P_out, dft.P_.data(), sizeof(double) * dft.P_.size());
    return DG_OK;
  } catch (const std::exception& e) {
    dg::set_last_error(std::string("dg_scf_run: ") + e.what());
    return DG_ERR_STATE;
  }
}

extern "C" int dg_batch_scf_run(int nmol, const dg_molecule* mols, const dg_basis* bases,
                                const dg_grid_params* grid, const dg_scf_params* scf, dg_scf_result* results,
                                int* iterations) {
  if (nmol < 1 || !mols || !bases || !grid || !scf || !results) {
    dg::set_last_error("dg_batch_scf_run: null or empty argument");
    return DG_ERR_INVALID;
  }
  if (!check_scf_params(scf, "dg_batch_scf_run")) return DG_ERR_INVALID;
  if (scf->hartree != DG_HARTREE_BECKE) {
    dg::set_last_error("dg_batch_scf_run: batches use the Becke-grid Hartree path");
    return DG_ERR_UNSUPPORTED;
  }
  dg_batch* batch = nullptr;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<dftcxx::HostSCF> dfts;
    dfts.reserve(nmol);
    for (int m = 0; m < nmol; ++m) {
      dftcxx::MoleculeData md = to_data(&mols[m], &bases[m]);
      if (md.nelec() % 2 != 0) throw std::runtime_error("closed-shell SCF needs an even electron count");
      dfts.emplace_back(md);
      dfts.back().mixing = scf->mixing;
      dfts.back().construct_matrices();
    }
    if (dg_batch_create(nmol, mols, bases, grid, &batch) != DG_OK) throw std::runtime_error(dg_last_error());
    const double setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<int> its(nmol, 0);
    double sec_iter = 0.0;
    std::vector<char> conv;
    std::vector<double> delta;
    const bool dev = scf->density_step != DG_STEP_HOST;
    int fallbacks = 0;
    const int total = dev ? dftcxx::batch_scf_device(dfts, batch, *scf, its.data(), &sec_iter, conv, delta, &fallbacks)
                          : dftcxx::batch_scf(dfts, batch, *scf, its.data(), &sec_iter, conv, delta);
    for (int m = 0; m < nmol; ++m) {
      dfts[m].fill(&results[m], its[m], setup, sec_iter, conv[m] != 0, delta[m]);
      results[m].device_step = dev ? 1 : 0;
      results[m].host_fallbacks = m == 0 ? fallbacks : 0;
    }
    if (iterations) *iterations = total;
    dg_batch_destroy(batch);
    return DG_OK;
  } catch (const std::exception& e) {
    if (batch) dg_batch_destroy(batch);
    dg::set_last_error(std::string("dg_batch_scf_run: ") + e.what());
    return DG_ERR_STATE;
  }
}
