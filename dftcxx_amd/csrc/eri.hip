// Analytic two-electron integrals and the Coulomb matrix from them, the
// reference's alternative Hartree path (hartree_evaluation =
// two_electron_integrals): DFT::calculate_two_electron_integrals
// (src/dft.cpp:254-288) computes every unique (ij|kl) once, and
// DFT::calculate_hartree_potential_te_int (src/dft.cpp:452-477) forms
// J_ij = sum_kl P_kl (ij|kl) every iteration.
//
// Here the integrals come from the McMurchie-Davidson scheme (Hermite
// expansion of each product Gaussian, Hermite Coulomb integrals R_tuv by the
// recursion on the Boys function) rather than the reference's
// Taketa-Huzinaga-O-ohata sums (src/integrals.cpp:389-468); both are exact, so
// they agree to rounding.  The prefactor 2 pi^(5/2) and the primitive norms
// use the reference's pi = 3.14159265359 (src/integrals.cpp:394, src/cgf.cpp:103),
// and the Boys argument is clamped to >= 1e-8 as src/gamma.cpp:41-46 does.
//
// Layout: function pairs IJ = i(i+1)/2 + j (i >= j) and integrals at
// IJ(IJ+1)/2 + KL (IJ >= KL), the reference's teindex (src/integrals.cpp:470-486).
// One GPU thread per unique integral; per primitive pair of every function pair
// a record {p, P, K, E_x[5], E_y[5], E_z[5]} computed once on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dftgrid.h"
#include "grid_tables.h"

namespace {

constexpr int kRec = 20;         // doubles per primitive-pair record
constexpr int kMaxLsum = 8;      // (dd|dd): four Cartesian d functions
constexpr double kPiTrunc = 3.14159265359;

int fail(int code, const std::string& msg) {
  dg::set_last_error(msg);
  return code;
}

#define ERI_HIP(expr)                                                                     \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) return fail(DG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

const int kCart[3][6][3] = {{{0, 0, 0}},
                            {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
                            {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}}};

// Hermite expansion coefficients E^{ij}_t (t = 0..i+j) of x_A^i x_B^j exp(-a x_A^2 - b x_B^2)
// about the product centre, E^{00}_0 = exp(-ab/(a+b) X_AB^2); i, j <= 2.
void hermite_1d(int i, int j, double a, double b, double xa, double xb, double* E /* [5] */) {
  double e[3][3][5];
  std::memset(e, 0, sizeof(e));
  const double p = a + b, xp = (a * xa + b * xb) / p;
  const double xpa = xp - xa, xpb = xp - xb, o2p = 0.5 / p;
  e[0][0][0] = std::exp(-a * b / p * (xa - xb) * (xa - xb));
  for (int ii = 0; ii <= i; ++ii)
    for (int jj = 0; jj <= j; ++jj) {
      if (ii == 0 && jj == 0) continue;
      const double(*s)[5] = ii > 0 ? e[ii - 1] : e[ii];
      const double* src = ii > 0 ? s[jj] : s[jj - 1];
      const double x = ii > 0 ? xpa : xpb;
      for (int t = 0; t <= ii + jj; ++t)
        e[ii][jj][t] = x * src[t] + (t + 1 <= 4 ? (t + 1) * src[t + 1] : 0.0) + (t > 0 ? o2p * src[t - 1] : 0.0);
    }
  for (int t = 0; t < 5; ++t) E[t] = t <= i + j ? e[i][j][t] : 0.0;
}

// Boys function F_n(T), n = 0..nmax, T clamped to >= 1e-8 (src/gamma.cpp:41-46):
// series e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)) and downward recursion
// below T = 40, F_0 = sqrt(pi/T)/2 and upward recursion above.
__device__ void boys_dev(int nmax, double T, double* F) {
  T = fmax(fabs(T), 1e-8);
  const double et = exp(-T);
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
    F[0] = 0.5 * sqrt(3.14159265358979323846 / T);
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
  }
}

// index of (t, u, v), t + u + v = s, in the Hermite tetrahedron
__device__ __forceinline__ int tuv(int t, int u, int v) {
  const int s = t + u + v, m = u + v;
  return s * (s + 1) * (s + 2) / 6 + m * (m + 1) / 2 + v;
}

// One unique contracted integral (IJ|KL) per thread, IJ >= KL, written at
// IJ(IJ+1)/2 + KL.  pair_off: primitive-pair records of pair IJ are
// [pair_off[IJ], pair_off[IJ+1]); pair_l: per pair the Hermite orders
// (lx_i + lx_j, ly_i + ly_j, lz_i + lz_j).  LMAX bounds the total order.
template <int LMAX>
__global__ __launch_bounds__(256) void k_eri(long nints, const int* __restrict__ pair_off,
                                             const int* __restrict__ pair_l, const double* __restrict__ rec,
                                             double* ints) {
  constexpr int NT = (LMAX + 1) * (LMAX + 2) * (LMAX + 3) / 6;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nints) return;
  long IJ = (long)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while (IJ * (IJ + 1) / 2 > q) --IJ;
  while ((IJ + 1) * (IJ + 2) / 2 <= q) ++IJ;
  const long KL = q - IJ * (IJ + 1) / 2;
  const int ax = pair_l[3 * IJ], ay = pair_l[3 * IJ + 1], az = pair_l[3 * IJ + 2];
  const int cx = pair_l[3 * KL], cy = pair_l[3 * KL + 1], cz = pair_l[3 * KL + 2];
  const int L = ax + ay + az + cx + cy + cz;
  double Ra[NT], Rb[NT], F[LMAX + 1];
  double total = 0.0;
  for (int ab = pair_off[IJ]; ab < pair_off[IJ + 1]; ++ab) {
    const double* A = rec + (long)ab * kRec;
    const double p = A[0];
    for (int cd = pair_off[KL]; cd < pair_off[KL + 1]; ++cd) {
      const double* C = rec + (long)cd * kRec;
      const double qe = C[0];
      const double X = A[1] - C[1], Y = A[2] - C[2], Z = A[3] - C[3];
      const double al = p * qe / (p + qe);
      boys_dev(L, al * (X * X + Y * Y + Z * Z), F);
      // R^n_tuv from n = L down to 0 (two levels in Ra / Rb)
      double* cur = Ra;
      double* nxt = Rb;
      double m2a = 1.0;
      for (int k = 0; k < L; ++k) m2a *= -2.0 * al;
      cur[0] = m2a * F[L];
      for (int n = L - 1; n >= 0; --n) {
        m2a /= -2.0 * al;
        const int lev = L - n;
        for (int s = 0; s <= lev; ++s)
          for (int t = s; t >= 0; --t)
            for (int v = 0; v <= s - t; ++v) {
              const int u = s - t - v;
              double r;
              if (s == 0)
                r = m2a * F[n];
              else if (t > 0)
                r = (t > 1 ? (t - 1) * cur[tuv(t - 2, u, v)] : 0.0) + X * cur[tuv(t - 1, u, v)];
              else if (u > 0)
                r = (u > 1 ? (u - 1) * cur[tuv(t, u - 2, v)] : 0.0) + Y * cur[tuv(t, u - 1, v)];
              else
                r = (v > 1 ? (v - 1) * cur[tuv(t, u, v - 2)] : 0.0) + Z * cur[tuv(t, u, v - 1)];
              nxt[tuv(t, u, v)] = r;
            }
        double* tmp = cur;
        cur = nxt;
        nxt = tmp;
      }
      double val = 0.0;
      for (int t = 0; t <= ax; ++t)
        for (int u = 0; u <= ay; ++u)
          for (int v = 0; v <= az; ++v) {
            const double eab = A[5 + t] * A[10 + u] * A[15 + v];
            double inner = 0.0;
            for (int tt = 0; tt <= cx; ++tt)
              for (int uu = 0; uu <= cy; ++uu)
                for (int vv = 0; vv <= cz; ++vv) {
                  const double ecd = C[5 + tt] * C[10 + uu] * C[15 + vv];
                  const double r = cur[tuv(t + tt, u + uu, v + vv)];
                  inner += ((tt + uu + vv) & 1) ? -ecd * r : ecd * r;
                }
            val += eab * inner;
          }
      const double pref = 2.0 * pow(kPiTrunc, 2.5) / (p * qe * sqrt(p + qe));
      total += pref * A[4] * C[4] * val;
    }
  }
  ints[q] = total;
}

// v_KL = P_kl (k == l) or P_kl + P_lk: J_IJ = sum_KL (IJ|KL) v_KL.
__global__ void k_eri_vec(const double* __restrict__ P, int n, const int* __restrict__ pij, int npair, double* v) {
  const int IJ = blockIdx.x * blockDim.x + threadIdx.x;
  if (IJ >= npair) return;
  const int i = pij[2 * IJ], j = pij[2 * IJ + 1];
  v[IJ] = (i == j) ? P[i + (long)i * n] : P[i + (long)j * n] + P[j + (long)i * n];
}

// Jp[IJ] = sum_{KL <= IJ} (IJ|KL) v_KL: one work-group per pair, the row of the
// packed lower triangle read contiguously, fixed-order reduction.
__global__ __launch_bounds__(256) void k_eri_rows(const double* __restrict__ ints, const double* __restrict__ v,
                                                  double* Jp) {
  __shared__ double red[256];
  const long IJ = blockIdx.x;
  const double* row = ints + IJ * (IJ + 1) / 2;
  double s = 0.0;
  for (long KL = threadIdx.x; KL <= IJ; KL += 256) s += row[KL] * v[KL];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) Jp[IJ] = red[0];
}

// part[chunk][IJ] = sum_{KL > IJ, KL in chunk} (KL|IJ) v_KL: lane IJ of a 64-pair
// block, KL over a chunk of kColChunk pairs; the 64 lanes read 64 consecutive
// integrals of row KL.
constexpr int kColChunk = 1024;
__global__ __launch_bounds__(64) void k_eri_cols(const double* __restrict__ ints, const double* __restrict__ v,
                                                 int npair, double* part) {
  const long IJ = (long)blockIdx.x * 64 + threadIdx.x;
  const long k0 = (long)blockIdx.y * kColChunk, k1 = std::min<long>(k0 + kColChunk, npair);
  double s = 0.0;
  if (IJ < npair)
    for (long KL = std::max(k0, IJ + 1); KL < k1; ++KL) s += ints[KL * (KL + 1) / 2 + IJ] * v[KL];
  if (IJ < npair) part[(long)blockIdx.y * npair + IJ] = s;
}

// J_ij = J_ji = Jp[IJ] + sum over chunks of part (in chunk order).
__global__ void k_eri_unpack(const double* __restrict__ Jp, const double* __restrict__ part, int nchunk,
                             const int* __restrict__ pij, int npair, int n, double* J) {
  const int IJ = blockIdx.x * blockDim.x + threadIdx.x;
  if (IJ >= npair) return;
  double s = Jp[IJ];
  for (int c = 0; c < nchunk; ++c) s += part[(long)c * npair + IJ];
  const int i = pij[2 * IJ], j = pij[2 * IJ + 1];
  J[i + (long)j * n] = s;
  J[j + (long)i * n] = s;
}

unsigned nb(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

struct dg_eri {
  int nbf = 0, npair = 0, nchunk = 0;
  long nints = 0;
  hipStream_t stream = nullptr;
  double *ints = nullptr, *rec = nullptr, *v = nullptr, *Jp = nullptr, *part = nullptr, *P = nullptr, *J = nullptr;
  int *pair_off = nullptr, *pair_l = nullptr, *pij = nullptr;
  ~dg_eri() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (double* b : {ints, rec, v, Jp, part, P, J})
      if (b) (void)hipFree(b);
    for (int* b : {pair_off, pair_l, pij})
      if (b) (void)hipFree(b);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" {

int dg_eri_create(const dg_molecule* mol, const dg_basis* basis, dg_eri** out) {
  if (!mol || !basis || !out) return fail(DG_ERR_INVALID, "dg_eri_create: null argument");
  *out = nullptr;
  if (mol->natoms < 1 || !mol->xyz) return fail(DG_ERR_INVALID, "dg_eri_create: empty molecule");
  if (basis->nshell < 1 || !basis->atom || !basis->L || !basis->nprim || !basis->offset || !basis->alpha ||
      !basis->coef)
    return fail(DG_ERR_INVALID, "dg_eri_create: empty basis");
  // functions: centre, Cartesian powers, primitives (alpha, c N)
  struct Fn {
    double A[3];
    int l[3];
    std::vector<double> a, cn;
  };
  std::vector<Fn> fn;
  int lmax = 0;
  for (int s = 0; s < basis->nshell; ++s) {
    const int L = basis->L[s], at = basis->atom[s];
    if (L < 0 || L > 2) return fail(DG_ERR_INVALID, "dg_eri_create: shell L must be 0, 1 or 2");
    if (at < 0 || at >= mol->natoms) return fail(DG_ERR_INVALID, "dg_eri_create: shell atom out of range");
    if (basis->nprim[s] < 1 || basis->offset[s] < 0) return fail(DG_ERR_INVALID, "dg_eri_create: bad primitive range");
    lmax = std::max(lmax, L);
    for (int c = 0; c < (L + 1) * (L + 2) / 2; ++c) {
      Fn f;
      for (int d = 0; d < 3; ++d) {
        f.A[d] = mol->xyz[3 * at + d];
        f.l[d] = kCart[L][c][d];
      }
      for (int k = 0; k < basis->nprim[s]; ++k) {
        const double a = basis->alpha[basis->offset[s] + k];
        if (!(a > 0.0)) return fail(DG_ERR_INVALID, "dg_eri_create: primitive exponent must be positive");
        f.a.push_back(a);
        f.cn.push_back(basis->coef[basis->offset[s] + k] * dg::gto_norm(a, f.l[0], f.l[1], f.l[2]));
      }
      fn.push_back(f);
    }
  }
  const int n = (int)fn.size();
  std::unique_ptr<dg_eri> e(new dg_eri);
  e->nbf = n;
  e->npair = n * (n + 1) / 2;
  e->nints = (long)e->npair * (e->npair + 1) / 2;
  if (e->nints > (1L << 36)) return fail(DG_ERR_UNSUPPORTED, "dg_eri_create: more than 2^36 integrals");
  // primitive-pair records, pair IJ = i(i+1)/2 + j
  std::vector<double> rec;
  std::vector<int> off(e->npair + 1, 0), pl(3 * (size_t)e->npair), pij(2 * (size_t)e->npair);
  for (int i = 0, IJ = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j, ++IJ) {
      pij[2 * IJ] = i;
      pij[2 * IJ + 1] = j;
      for (int d = 0; d < 3; ++d) pl[3 * IJ + d] = fn[i].l[d] + fn[j].l[d];
      off[IJ] = (int)(rec.size() / kRec);
      for (size_t ka = 0; ka < fn[i].a.size(); ++ka)
        for (size_t kb = 0; kb < fn[j].a.size(); ++kb) {
          const double a = fn[i].a[ka], b = fn[j].a[kb], p = a + b;
          double r[kRec];
          r[0] = p;
          for (int d = 0; d < 3; ++d) r[1 + d] = (a * fn[i].A[d] + b * fn[j].A[d]) / p;
          r[4] = fn[i].cn[ka] * fn[j].cn[kb];
          for (int d = 0; d < 3; ++d) hermite_1d(fn[i].l[d], fn[j].l[d], a, b, fn[i].A[d], fn[j].A[d], r + 5 + 5 * d);
          rec.insert(rec.end(), r, r + kRec);
        }
    }
  off[e->npair] = (int)(rec.size() / kRec);
  e->nchunk = (e->npair + kColChunk - 1) / kColChunk;
  ERI_HIP(hipStreamCreate(&e->stream));
  ERI_HIP(hipMalloc((void**)&e->ints, sizeof(double) * std::max<long>(e->nints, 1)));
  ERI_HIP(hipMalloc((void**)&e->rec, sizeof(double) * std::max<size_t>(rec.size(), 1)));
  ERI_HIP(hipMalloc((void**)&e->pair_off, sizeof(int) * off.size()));
  ERI_HIP(hipMalloc((void**)&e->pair_l, sizeof(int) * pl.size()));
  ERI_HIP(hipMalloc((void**)&e->pij, sizeof(int) * pij.size()));
  ERI_HIP(hipMalloc((void**)&e->v, sizeof(double) * e->npair));
  ERI_HIP(hipMalloc((void**)&e->Jp, sizeof(double) * e->npair));
  ERI_HIP(hipMalloc((void**)&e->part, sizeof(double) * (size_t)e->nchunk * e->npair));
  ERI_HIP(hipMalloc((void**)&e->P, sizeof(double) * (size_t)n * n));
  ERI_HIP(hipMalloc((void**)&e->J, sizeof(double) * (size_t)n * n));
  ERI_HIP(hipMemcpyAsync(e->rec, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice, e->stream));
  ERI_HIP(hipMemcpyAsync(e->pair_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, e->stream));
  ERI_HIP(hipMemcpyAsync(e->pair_l, pl.data(), sizeof(int) * pl.size(), hipMemcpyHostToDevice, e->stream));
  ERI_HIP(hipMemcpyAsync(e->pij, pij.data(), sizeof(int) * pij.size(), hipMemcpyHostToDevice, e->stream));
  const unsigned blocks = nb(e->nints, 256);
  if (lmax <= 1)
    hipLaunchKernelGGL(k_eri<4>, dim3(blocks), dim3(256), 0, e->stream, e->nints, e->pair_off, e->pair_l, e->rec,
                       e->ints);
  else
    hipLaunchKernelGGL(k_eri<kMaxLsum>, dim3(blocks), dim3(256), 0, e->stream, e->nints, e->pair_off, e->pair_l,
                       e->rec, e->ints);
  ERI_HIP(hipGetLastError());
  ERI_HIP(hipStreamSynchronize(e->stream));
  *out = e.release();
  return DG_OK;
}

int dg_eri_destroy(dg_eri* e) {
  delete e;
  return DG_OK;
}

int dg_eri_count(const dg_eri* e, long* nints, int* nbf) {
  if (!e) return fail(DG_ERR_INVALID, "dg_eri_count: null object");
  if (nints) *nints = e->nints;
  if (nbf) *nbf = e->nbf;
  return DG_OK;
}

int dg_eri_get(dg_eri* e, double* ints) {
  if (!e || !ints) return fail(DG_ERR_INVALID, "dg_eri_get: null argument");
  ERI_HIP(hipMemcpyAsync(ints, e->ints, sizeof(double) * e->nints, hipMemcpyDeviceToHost, e->stream));
  ERI_HIP(hipStreamSynchronize(e->stream));
  return DG_OK;
}

int dg_eri_coulomb_device(dg_eri* e, const double* dP, double* dJ, void* stream) {
  if (!e || !dP || !dJ) return fail(DG_ERR_INVALID, "dg_eri_coulomb_device: null argument");
  hipStream_t st = stream ? (hipStream_t)stream : e->stream;
  const int np = e->npair;
  hipLaunchKernelGGL(k_eri_vec, dim3(nb(np, 256)), dim3(256), 0, st, dP, e->nbf, e->pij, np, e->v);
  hipLaunchKernelGGL(k_eri_rows, dim3(np), dim3(256), 0, st, e->ints, e->v, e->Jp);
  hipLaunchKernelGGL(k_eri_cols, dim3(nb(np, 64), e->nchunk), dim3(64), 0, st, e->ints, e->v, np, e->part);
  hipLaunchKernelGGL(k_eri_unpack, dim3(nb(np, 256)), dim3(256), 0, st, e->Jp, e->part, e->nchunk, e->pij, np,
                     e->nbf, dJ);
  ERI_HIP(hipGetLastError());
  return DG_OK;
}

int dg_eri_coulomb(dg_eri* e, const double* P, double* J) {
  if (!e || !P || !J) return fail(DG_ERR_INVALID, "dg_eri_coulomb: null argument");
  const size_t bytes = sizeof(double) * (size_t)e->nbf * e->nbf;
  ERI_HIP(hipMemcpyAsync(e->P, P, bytes, hipMemcpyHostToDevice, e->stream));
  const int rc = dg_eri_coulomb_device(e, e->P, e->J, nullptr);
  if (rc != DG_OK) return rc;
  ERI_HIP(hipMemcpyAsync(J, e->J, bytes, hipMemcpyDeviceToHost, e->stream));
  ERI_HIP(hipStreamSynchronize(e->stream));
  return DG_OK;
}

}  // extern "C"
