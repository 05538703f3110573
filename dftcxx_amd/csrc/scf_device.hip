// Device-resident density-matrix step of the SCF (src/dft.cpp:330-366, 441-447):
// F = H + 2J + XC, F' = X^T F X, the projector onto the nocc lowest
// eigenvectors of F', P = X P' X^T, the mixing with the previous P, and the
// energy 2 tr(PH) + E_J + E_nuc + E_xc -- all on the device, so that an SCF
// iteration moves no matrix between host and device.
//
// The reference diagonalises F' (Eigen::SelfAdjointEigenSolver) and forms
// P = C_occ C_occ^T.  Only that projector enters the SCF, so here it is
// computed by second-order spectral projection (SP2; A. M. N. Niklasson,
// Phys. Rev. B 66, 155115 (2002)): starting from the Gershgorin-bounded linear
// map P0 = (h_max I - F')/(h_max - h_min), each step replaces P by P^2 or by
// 2P - P^2, whichever trace lies closer to nocc; the eigenvalues of P go to 1
// below the HOMO-LUMO gap and to 0 above it.  Every step is one symmetric
// FP64 matrix product on the matrix cores.  Without a gap (degenerate
// frontier orbitals) the iteration stalls; the step then reports failure and
// the caller falls back to the host eigen-solver.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "scf_device.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma_f64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Control block of one projection: done, steps taken, failed; the idempotency
// error tr(P) - tr(P^2) of the last two steps.
struct Sp2Ctl {
  int done, iter, fail, pad;
  double err, err_prev, hmin, hmax;
};

// C = alpha op(A) op(B) + beta D (D may be NULL), column-major, M x N with inner
// dimension K.  A work-group computes one 16 x 16 tile of C on the FP64 16x16x4
// MFMA, its four waves a quarter of the K range each: lane (g, c) supplies
// op(A)(i0 + c, k + g) and op(B)(k + g, j0 + c), and result register v holds
// C(i0 + 4v + g, j0 + c).  A wave issues the loads of up to kGemmBatch 16-wide
// K blocks before its first MFMA (at N <= 512 all of its range), so that it
// waits for memory once, not once per block; the four partial tiles are added
// in wave order through LDS.  TA reads A transposed (A(k, i)), TB reads B
// transposed (B(j, k)); for a symmetric B, TB = true gives the same product
// with lane-contiguous loads.  The launch returns at once when ctl->done.
constexpr int kGemmBatch = 8;
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, double alpha, const double* __restrict__ A,
                                              int lda, const double* __restrict__ B, int ldb, double beta,
                                              const double* __restrict__ D, double* __restrict__ C, int ldc,
                                              const Sp2Ctl* __restrict__ ctl) {
  if (ctl && ctl->done) return;
  __shared__ double part[3][256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  const int ia = i0 + c, jb = j0 + c;
  const bool arow = ia < M, bcol = jb < N;
  const int nkb = (K + 15) / 16;
  const int kb0 = wave * nkb / 4, kb1 = (wave + 1) * nkb / 4;
  dbl4 d = {0.0, 0.0, 0.0, 0.0};
  for (int kb = kb0; kb < kb1; kb += kGemmBatch) {
    double a[kGemmBatch][4], b[kGemmBatch][4];
#pragma unroll
    for (int q = 0; q < kGemmBatch; ++q)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = 16 * (kb + q) + 4 * u + g;
        const bool in = kb + q < kb1 && k < K;
        a[q][u] = (in && arow) ? (TA ? A[k + (long)ia * lda] : A[ia + (long)k * lda]) : 0.0;
        b[q][u] = (in && bcol) ? (TB ? B[jb + (long)k * ldb] : B[k + (long)jb * ldb]) : 0.0;
      }
#pragma unroll
    for (int q = 0; q < kGemmBatch; ++q)
      if (kb + q < kb1) {
#pragma unroll
        for (int u = 0; u < 4; ++u) d = mfma_f64(a[q][u], b[q][u], d);
      }
  }
  if (wave > 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) part[wave - 1][v * 64 + lane] = d[v];
  }
  __syncthreads();
  if (wave != 0 || !bcol) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = i0 + 4 * v + g;
    if (row >= M) continue;
    const double x = ((d[v] + part[0][v * 64 + lane]) + part[1][v * 64 + lane]) + part[2][v * 64 + lane];
    const long e = row + (long)jb * ldc;
    C[e] = alpha * x + (D ? beta * D[e] : 0.0);
  }
}

// F = H + 2J + XC (src/dft.cpp:334), or F = H + G with G = 2J + XC from one
// contraction (J NULL); XC may be NULL (the set-up step, before any XC).
__global__ void k_fock(const double* __restrict__ H, const double* __restrict__ J, const double* __restrict__ XC,
                       long nn, double* F) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nn) F[e] = H[e] + (J ? 2.0 * J[e] : 0.0) + (XC ? XC[e] : 0.0);
}

// Gershgorin discs of the symmetric F': row i (its column i, contiguous) gives
// lo[i] = F'_ii - sum_{j != i} |F'_ij| and hi[i] = F'_ii + the same; one wave per row.
__global__ __launch_bounds__(256) void k_gershgorin(const double* __restrict__ F, int n, double* lo, double* hi) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* col = F + (long)i * n;
  double r = 0.0;
  for (int j = lane; j < n; j += 64)
    if (j != i) r += fabs(col[j]);
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
  if (lane == 0) {
    lo[i] = col[i] - r;
    hi[i] = col[i] + r;
  }
}

// h_min = min lo, h_max = max hi (exact, so the order does not matter); resets
// the control block.  One work-group.
__global__ __launch_bounds__(1024) void k_sp2_init(const double* __restrict__ lo, const double* __restrict__ hi,
                                                   int n, Sp2Ctl* ctl) {
  __shared__ double lo_s[1024], hi_s[1024];
  double l = 1e300, h = -1e300;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    l = fmin(l, lo[i]);
    h = fmax(h, hi[i]);
  }
  lo_s[threadIdx.x] = l;
  hi_s[threadIdx.x] = h;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      lo_s[threadIdx.x] = fmin(lo_s[threadIdx.x], lo_s[threadIdx.x + s]);
      hi_s[threadIdx.x] = fmax(hi_s[threadIdx.x], hi_s[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctl->done = ctl->iter = ctl->fail = 0;
    ctl->err = ctl->err_prev = 1e300;
    ctl->hmin = lo_s[0];
    ctl->hmax = hi_s[0];
  }
}

// P0 = (h_max I - F') / (h_max - h_min), the linear map of the spectrum of F' into [0, 1].
__global__ void k_sp2_p0(const double* __restrict__ F, int n, double* P, const Sp2Ctl* __restrict__ ctl) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const double hmin = ctl->hmin, hmax = ctl->hmax;
  const double inv = 1.0 / fmax(hmax - hmin, 1e-300);
  const long i = e % n, j = e / n;
  P[e] = ((i == j ? hmax : 0.0) - F[e]) * inv;
}

// One SP2 step after P2 = P^2, over the elements of P: every work-group forms
// tr(P) and tr(P^2) in the same fixed order, so all take the same decision: the
// projection stops once the idempotency error tr(P) - tr(P^2) falls below tol
// (the eigenvalues of P are then within about tol of 0 or 1; two McWeeny steps
// follow); it fails at step maxit - 1.  Else Pout = P^2 or 2P - P^2, whichever
// trace lies closer to nocc (Pout is not P, whose diagonal other work-groups
// may still be reading).  Work-group 0 records the outcome in ctl (iter = the
// index of the step whose input holds the current P).  (A test on the error's
// decrease from one step to the next would not do: a step squares the
// eigenvalue errors on one side of the gap but doubles them on the other.)
__global__ __launch_bounds__(256) void k_sp2_step(const double* __restrict__ P, const double* __restrict__ P2,
                                                  double* __restrict__ Pout, int n, int nocc, double tol, int step,
                                                  int maxit, Sp2Ctl* ctl) {
  if (ctl->done) return;
  __shared__ double t1s[256], t2s[256];
  double t1 = 0.0, t2 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    t1 += P[i + (long)i * n];
    t2 += P2[i + (long)i * n];
  }
  t1s[threadIdx.x] = t1;
  t2s[threadIdx.x] = t2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      t1s[threadIdx.x] += t1s[threadIdx.x + s];
      t2s[threadIdx.x] += t2s[threadIdx.x + s];
    }
    __syncthreads();
  }
  const double tr1 = t1s[0], tr2 = t2s[0];
  const double err = fabs(tr1 - tr2);
  const bool stop = err < tol, fail = !stop && step + 1 >= maxit;
  // trace after P -> P^2 is tr2, after P -> 2P - P^2 it is 2 tr1 - tr2
  const bool sq = fabs(tr2 - nocc) < fabs(2.0 * tr1 - tr2 - nocc);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->err_prev = ctl->err;
    ctl->err = err;
    if (stop || fail) {
      ctl->done = 1;
      ctl->fail = fail ? 1 : 0;
    } else {
      ctl->iter = step + 1;
      ctl->pad = sq ? 1 : 0;
    }
  }
  if (stop || fail) return;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  Pout[e] = sq ? P2[e] : 2.0 * P[e] - P2[e];
}

// dst = src elementwise: the result Q = 3 P^2 - 2 P^2 P of a McWeeny step
// (which squares the eigenvalue errors on both sides of the gap, once the
// projection has brought them within 1e-6 of 0 or 1) copied back over P.
__global__ void k_copy(double* dst, const double* __restrict__ src, long nn) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nn) dst[e] = src[e];
}

// P = Pn on the first step, else (1 - m) Pn + m P (src/dft.cpp:331, 358).
__global__ void k_mix(double* P, const double* __restrict__ Pn, long nn, double m, int first) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  P[e] = first ? Pn[e] : (1.0 - m) * Pn[e] + m * P[e];
}

// Partial sums of tr(PH) and tr(PJ) = sum_e P_e H_e, sum_e P_e J_e (P, H, J
// symmetric): work-group b sums its contiguous range of the n^2 elements in a
// fixed order into part[2b], part[2b + 1].
constexpr int kEnergyParts = 64;
__global__ __launch_bounds__(256) void k_energy_part(const double* __restrict__ P, const double* __restrict__ H,
                                                     const double* __restrict__ J, long nn, double* part) {
  __shared__ double s1[256], s2[256];
  const long e0 = nn * blockIdx.x / gridDim.x, e1 = nn * (blockIdx.x + 1) / gridDim.x;
  double t1 = 0.0, t2 = 0.0;
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    t1 += P[e] * H[e];
    t2 += P[e] * J[e];
  }
  s1[threadIdx.x] = t1;
  s2[threadIdx.x] = t2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// E = 2 tr(PH) + E_J + E_nuc + E_xc (src/dft.cpp:441-447) with E_J = 2 tr(PJ), or
// the pointwise value scal[3] (use_point_ej); E_xc = scal[0] when use_exc.
// out = {E_one, E_J, E_xc, total, N after correction}.  The partial sums of
// k_energy_part are added in order by one thread.
__global__ void k_energy(const double* __restrict__ part, int nparts, const double* __restrict__ scal, double enuc,
                         int use_exc, int use_point_ej, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double t1 = 0.0, t2 = 0.0;
  for (int b = 0; b < nparts; ++b) {
    t1 += part[2 * b];
    t2 += part[2 * b + 1];
  }
  const double e1 = 2.0 * t1;
  const double ej = use_point_ej ? scal[3] : 2.0 * t2;
  const double exc = use_exc ? scal[0] : 0.0;
  out[0] = e1;
  out[1] = ej;
  out[2] = exc;
  out[3] = e1 + ej + enuc + exc;
  out[4] = scal[2];
}

// ---------------------------------------------------------------------------
// Batched step: one work-group per molecule of a batch (n <= 32 functions),
// every matrix in LDS.  The same algebra as the single-molecule step: F, X^T F X,
// SP2 to an idempotency error of 1e-6 and two McWeeny steps, X P' X^T, mixing.
// Matrices are packed column-major at off[m] (the batch layout).
// ---------------------------------------------------------------------------
constexpr int kBatchN = 32;
constexpr int kBatchMaxSteps = 100;

// C = A op(B) for n x n LDS matrices (column-major, leading dimension n);
// TB reads B transposed.  All 256 threads; ends with a barrier.
template <bool TA, bool TB>
__device__ void lds_mm(int n, const double* A, const double* B, double* C) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e % n, j = e / n;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += (TA ? A[k + i * n] : A[i + k * n]) * (TB ? B[j + k * n] : B[k + j * n]);
    C[e] = s;
  }
  __syncthreads();
}

// sum over the work-group of v (fixed order); every thread gets the result
__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// One density-matrix step of every active molecule; fail[m] = 1, with P_m left
// as it was, when its projection did not reach the tolerance (no gap: the
// caller forms that molecule's step on the host).
__global__ __launch_bounds__(256) void k_batch_density(
    const int* __restrict__ nbf, const long* __restrict__ off, const int* __restrict__ nocc,
    const char* __restrict__ active, const double* __restrict__ H, const double* __restrict__ X,
    const double* __restrict__ J, const double* __restrict__ XC, int with_xc, int first, double mixing, double* P,
    int* fail) {
  __shared__ double A[kBatchN * kBatchN], B[kBatchN * kBatchN], Cm[kBatchN * kBatchN], Xs[kBatchN * kBatchN];
  __shared__ double red[256];
  const int m = blockIdx.x;
  if (!active[m]) return;
  const int n = nbf[m], nn = n * n;
  const long o = off[m];
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    A[e] = H[o + e] + 2.0 * J[o + e] + (with_xc ? XC[o + e] : 0.0);
    Xs[e] = X[o + e];
  }
  __syncthreads();
  lds_mm<true, false>(n, Xs, A, B);   // X^T F
  lds_mm<false, false>(n, B, Xs, A);  // F' = X^T F X
  // Gershgorin bounds of F'
  double lo = 1e300, hi = -1e300;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double r = 0.0;
    for (int j = 0; j < n; ++j)
      if (j != i) r += fabs(A[j + i * n]);
    lo = fmin(lo, A[i + i * n] - r);
    hi = fmax(hi, A[i + i * n] + r);
  }
  red[threadIdx.x] = lo;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const double hmin = red[0];
  __syncthreads();
  red[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const double hmax = red[0];
  __syncthreads();
  const double inv = 1.0 / fmax(hmax - hmin, 1e-300);
  for (int e = threadIdx.x; e < nn; e += blockDim.x) B[e] = ((e % n == e / n ? hmax : 0.0) - A[e]) * inv;
  __syncthreads();
  // SP2 on B (P'), squares in A
  const int no = nocc[m];
  bool ok = false;
  for (int it = 0; it < kBatchMaxSteps; ++it) {
    lds_mm<false, true>(n, B, B, A);
    double t1 = 0.0, t2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      t1 += B[i + i * n];
      t2 += A[i + i * n];
    }
    const double tr1 = block_sum(t1, red), tr2 = block_sum(t2, red);
    if (fabs(tr1 - tr2) < 1e-6) {
      ok = true;
      break;
    }
    const bool sq = fabs(tr2 - no) < fabs(2.0 * tr1 - tr2 - no);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) B[e] = sq ? A[e] : 2.0 * B[e] - A[e];
    __syncthreads();
  }
  // two McWeeny steps P <- 3P^2 - 2P^3
  for (int k = 0; k < 2; ++k) {
    lds_mm<false, true>(n, B, B, A);
    lds_mm<false, true>(n, A, B, Cm);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) B[e] = 3.0 * A[e] - 2.0 * Cm[e];
    __syncthreads();
  }
  lds_mm<false, false>(n, Xs, B, A);  // X P'
  lds_mm<false, true>(n, A, Xs, Cm);  // Pn = X P' X^T
  if (ok)
    for (int e = threadIdx.x; e < nn; e += blockDim.x)
      P[o + e] = first ? Cm[e] : (1.0 - mixing) * Cm[e] + mixing * P[o + e];
  if (threadIdx.x == 0) fail[m] = ok ? 0 : 1;
}

// Energy of every active molecule: out[5m..5m+4] = {E_one, E_J = 2 tr(PJ), E_xc
// (scal[4m], when with_xc), total, N after correction (scal[4m+2])}.
__global__ __launch_bounds__(256) void k_batch_energy(const int* __restrict__ nbf, const long* __restrict__ off,
                                                      const char* __restrict__ active, const double* __restrict__ P,
                                                      const double* __restrict__ H, const double* __restrict__ J,
                                                      const double* __restrict__ scal,
                                                      const double* __restrict__ enuc, int with_xc, double* out) {
  __shared__ double red[256];
  const int m = blockIdx.x;
  if (!active[m]) return;
  const int nn = nbf[m] * nbf[m];
  const long o = off[m];
  double t1 = 0.0, t2 = 0.0;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    t1 += P[o + e] * H[o + e];
    t2 += P[o + e] * J[o + e];
  }
  const double s1 = block_sum(t1, red), s2 = block_sum(t2, red);
  if (threadIdx.x == 0) {
    const double exc = with_xc ? scal[4 * m] : 0.0;
    out[5 * m] = 2.0 * s1;
    out[5 * m + 1] = 2.0 * s2;
    out[5 * m + 2] = exc;
    out[5 * m + 3] = 2.0 * s1 + 2.0 * s2 + enuc[m] + exc;
    out[5 * m + 4] = scal[4 * m + 2];
  }
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

unsigned nb(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

namespace dftcxx {

DeviceStep::DeviceStep(int n, int nocc, hipStream_t stream) : n_(n), nocc_(nocc), st_(stream) {
  const size_t nn = (size_t)n * n;
  // H, X, F (also P'), F', Pn, P2 (also X P'), P, J, XC, scalars(4), energy(5 of 8),
  // energy partial sums (2 kEnergyParts), Gershgorin bounds (2 n)
  const size_t total = 9 * nn + 12 + 2 * kEnergyParts + 2 * (size_t)n;
  check(hipMalloc((void**)&buf_, sizeof(double) * total), "hipMalloc of the SCF buffers");
  check(hipMalloc((void**)&ctl_, sizeof(Sp2Ctl)), "hipMalloc of the SCF control block");
  check(hipHostMalloc((void**)&host_, sizeof(double) * 8 + sizeof(Sp2Ctl), hipHostMallocDefault),
        "hipHostMalloc of the SCF staging");
  H_ = buf_;
  X_ = H_ + nn;
  F_ = X_ + nn;
  Fp_ = F_ + nn;
  Pn_ = Fp_ + nn;
  P2_ = Pn_ + nn;
  P_ = P2_ + nn;
  J_ = P_ + nn;
  XC_ = J_ + nn;
  scal_ = XC_ + nn;
  eout_ = scal_ + 4;
  epart_ = eout_ + 8;
  glo_ = epart_ + 2 * kEnergyParts;
  ghi_ = glo_ + n;
  check(hipMemsetAsync(buf_, 0, sizeof(double) * total, st_), "hipMemsetAsync");
}

DeviceStep::~DeviceStep() {
  if (st_) (void)hipStreamSynchronize(st_);
  if (buf_) (void)hipFree(buf_);
  if (ctl_) (void)hipFree(ctl_);
  if (host_) (void)hipHostFree(host_);
}

void DeviceStep::upload_core(const double* H, const double* X) {
  const size_t bytes = sizeof(double) * (size_t)n_ * n_;
  check(hipMemcpyAsync(H_, H, bytes, hipMemcpyHostToDevice, st_), "copy of H");
  check(hipMemcpyAsync(X_, X, bytes, hipMemcpyHostToDevice, st_), "copy of X");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

template <bool TA, bool TB>
static void gemm(hipStream_t st, int n, const double* A, const double* B, double* C, const Sp2Ctl* ctl) {
  const dim3 grid(nb(n, 16), nb(n, 16));
  hipLaunchKernelGGL((k_gemm<TA, TB>), grid, dim3(256), 0, st, n, n, n, 1.0, A, n, B, n, 0.0, nullptr, C, n, ctl);
}

bool DeviceStep::density(bool with_xc, bool first, double mixing, bool fused) {
  const long nn = (long)n_ * n_;
  Sp2Ctl* ctl = static_cast<Sp2Ctl*>(ctl_);
  if (fused)  // XC_ holds G = 2J + XC of the last build
    hipLaunchKernelGGL(k_fock, dim3(nb(nn, 256)), dim3(256), 0, st_, H_, nullptr, XC_, nn, F_);
  else
    hipLaunchKernelGGL(k_fock, dim3(nb(nn, 256)), dim3(256), 0, st_, H_, J_, with_xc ? XC_ : nullptr, nn, F_);
  // F' = X^T F X: Pn <- X^T F (A transposed), F' <- Pn X
  gemm<true, false>(st_, n_, X_, F_, Pn_, nullptr);
  gemm<false, false>(st_, n_, Pn_, X_, Fp_, nullptr);
  // SP2 on F' into F (as P'), P2 as the square
  hipLaunchKernelGGL(k_gershgorin, dim3(nb(n_, 4)), dim3(256), 0, st_, Fp_, n_, glo_, ghi_);
  hipLaunchKernelGGL(k_sp2_init, dim3(1), dim3(1024), 0, st_, glo_, ghi_, n_, ctl);
  // the iterates alternate between F and F' (free once P0 is formed): step k
  // reads pb[k % 2] and writes pb[(k + 1) % 2]
  hipLaunchKernelGGL(k_sp2_p0, dim3(nb(nn, 256)), dim3(256), 0, st_, Fp_, n_, F_, ctl);
  double* pb[2] = {F_, Fp_};
  const double tol = 1e-6;
  // steps in groups, the control block read back between groups (a converged
  // projection makes the rest of a group return at once)
  Sp2Ctl* hc = reinterpret_cast<Sp2Ctl*>(host_ + 8);
  for (int it = 0; it < kMaxSteps;) {
    const int group = it == 0 ? 24 : 8;
    for (int k = it; k < it + group; ++k) {
      gemm<false, true>(st_, n_, pb[k & 1], pb[k & 1], P2_, ctl);  // P' symmetric: B read transposed
      hipLaunchKernelGGL(k_sp2_step, dim3(nb(nn, 256)), dim3(256), 0, st_, pb[k & 1], P2_, pb[(k + 1) & 1], n_,
                         nocc_, tol, k, kMaxSteps, ctl);
    }
    it += group;
    check(hipMemcpyAsync(hc, ctl_, sizeof(Sp2Ctl), hipMemcpyDeviceToHost, st_), "copy of the SP2 state");
    check(hipStreamSynchronize(st_), "hipStreamSynchronize");
    if (hc->done) break;
  }
  // no gap: P is left as it was, the caller forms this step on the host
  if (!hc->done || hc->fail) return false;
  // the converged P' is the input of step hc->iter
  if (hc->iter & 1) hipLaunchKernelGGL(k_copy, dim3(nb(nn, 256)), dim3(256), 0, st_, F_, Fp_, nn);
  // two McWeeny steps: eigenvalue errors 1e-6 -> 1e-12 -> rounding
  for (int k = 0; k < 2; ++k) {
    gemm<false, true>(st_, n_, F_, F_, P2_, nullptr);
    const dim3 grid(nb(n_, 16), nb(n_, 16));
    hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, st_, n_, n_, n_, -2.0, P2_, n_, F_, n_, 3.0, P2_,
                       Fp_, n_, (const Sp2Ctl*)nullptr);
    hipLaunchKernelGGL(k_copy, dim3(nb(nn, 256)), dim3(256), 0, st_, F_, Fp_, nn);
  }
  // Pn = X P' X^T: P2 <- X P', Pn <- P2 X^T
  gemm<false, false>(st_, n_, X_, F_, P2_, nullptr);
  gemm<false, true>(st_, n_, P2_, X_, Pn_, nullptr);
  hipLaunchKernelGGL(k_mix, dim3(nb(nn, 256)), dim3(256), 0, st_, P_, Pn_, nn, mixing, first ? 1 : 0);
  check(hipGetLastError(), "launch of the density-matrix step");
  return true;
}

void DeviceStep::energy(double enuc, bool with_xc, bool point_ej) {
  hipLaunchKernelGGL(k_energy_part, dim3(kEnergyParts), dim3(256), 0, st_, P_, H_, J_, (long)n_ * n_, epart_);
  hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st_, epart_, kEnergyParts, scal_, enuc, with_xc ? 1 : 0,
                     point_ej ? 1 : 0, eout_);
  check(hipGetLastError(), "launch of the energy");
  check(hipMemcpyAsync(host_, eout_, 5 * sizeof(double), hipMemcpyDeviceToHost, st_), "copy of the energy");
}

void DeviceStep::finish(double* e_one, double* e_j, double* e_xc, double* total, double* nelec) {
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
  *e_one = host_[0];
  *e_j = host_[1];
  *e_xc = host_[2];
  *total = host_[3];
  *nelec = host_[4];
}

void DeviceStep::clear_xc() {
  check(hipMemsetAsync(XC_, 0, sizeof(double) * (size_t)n_ * n_, st_), "hipMemsetAsync");
}

void DeviceStep::download(double* dst, const double* src) {
  check(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n_ * n_, hipMemcpyDeviceToHost, st_), "copy to host");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

void DeviceStep::upload(double* dst, const double* src) {
  check(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)n_ * n_, hipMemcpyHostToDevice, st_), "copy to device");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

BatchStep::BatchStep(int nmol, const int* nbf, const long* off, const int* nocc, hipStream_t stream)
    : nmol_(nmol), packed_(off[nmol]), st_(stream) {
  for (int m = 0; m < nmol; ++m)
    if (nbf[m] > kBatchN) throw std::runtime_error("batched device step: more than 32 functions in a molecule");
  const size_t np = (size_t)packed_;
  // H, X, P, J, XC, scalars (4/mol), energies (5/mol), enuc
  check(hipMalloc((void**)&buf_, sizeof(double) * (5 * np + 10 * (size_t)nmol)), "hipMalloc of the batch SCF");
  check(hipMalloc((void**)&ibuf_, sizeof(long) * (nmol + 1) + sizeof(int) * 3 * (size_t)nmol + nmol),
        "hipMalloc of the batch SCF");
  check(hipHostMalloc((void**)&host_, sizeof(double) * 5 * (size_t)nmol + sizeof(int) * (size_t)nmol,
                      hipHostMallocDefault),
        "hipHostMalloc of the batch SCF");
  H_ = buf_;
  X_ = H_ + np;
  P_ = X_ + np;
  J_ = P_ + np;
  XC_ = J_ + np;
  scal_ = XC_ + np;
  eout_ = scal_ + 4 * (size_t)nmol;
  enuc_ = eout_ + 5 * (size_t)nmol;
  off_ = reinterpret_cast<long*>(ibuf_);
  nbf_ = reinterpret_cast<int*>(off_ + nmol + 1);
  nocc_ = nbf_ + nmol;
  fail_ = nocc_ + nmol;
  active_ = reinterpret_cast<char*>(fail_ + nmol);
  check(hipMemsetAsync(buf_, 0, sizeof(double) * (5 * np + 10 * (size_t)nmol), st_), "hipMemsetAsync");
  check(hipMemcpyAsync(off_, off, sizeof(long) * (nmol + 1), hipMemcpyHostToDevice, st_), "copy");
  check(hipMemcpyAsync(nbf_, nbf, sizeof(int) * nmol, hipMemcpyHostToDevice, st_), "copy");
  check(hipMemcpyAsync(nocc_, nocc, sizeof(int) * nmol, hipMemcpyHostToDevice, st_), "copy");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

BatchStep::~BatchStep() {
  if (st_) (void)hipStreamSynchronize(st_);
  if (buf_) (void)hipFree(buf_);
  if (ibuf_) (void)hipFree(ibuf_);
  if (host_) (void)hipHostFree(host_);
}

void BatchStep::upload_core(const double* H, const double* X, const double* enuc) {
  check(hipMemcpyAsync(H_, H, sizeof(double) * packed_, hipMemcpyHostToDevice, st_), "copy of H");
  check(hipMemcpyAsync(X_, X, sizeof(double) * packed_, hipMemcpyHostToDevice, st_), "copy of X");
  check(hipMemcpyAsync(enuc_, enuc, sizeof(double) * nmol_, hipMemcpyHostToDevice, st_), "copy of E_nuc");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

void BatchStep::set_active(const char* active) {
  check(hipMemcpyAsync(active_, active, nmol_, hipMemcpyHostToDevice, st_), "copy of the active flags");
}

int BatchStep::density(bool with_xc, bool first, double mixing) {
  hipLaunchKernelGGL(k_batch_density, dim3(nmol_), dim3(256), 0, st_, nbf_, off_, nocc_, active_, H_, X_, J_, XC_,
                     with_xc ? 1 : 0, first ? 1 : 0, mixing, P_, fail_);
  check(hipGetLastError(), "launch of the batched density step");
  int* hf = reinterpret_cast<int*>(host_ + 5 * (size_t)nmol_);
  check(hipMemcpyAsync(hf, fail_, sizeof(int) * nmol_, hipMemcpyDeviceToHost, st_), "copy of the projection flags");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
  int nf = 0;
  for (int m = 0; m < nmol_; ++m) nf += hf[m] != 0;
  return nf;
}

void BatchStep::clear_xc() {
  check(hipMemsetAsync(XC_, 0, sizeof(double) * (size_t)packed_, st_), "hipMemsetAsync");
}

void BatchStep::energy(bool with_xc) {
  hipLaunchKernelGGL(k_batch_energy, dim3(nmol_), dim3(256), 0, st_, nbf_, off_, active_, P_, H_, J_, scal_, enuc_,
                     with_xc ? 1 : 0, eout_);
  check(hipGetLastError(), "launch of the batched energy");
  check(hipMemcpyAsync(host_, eout_, sizeof(double) * 5 * nmol_, hipMemcpyDeviceToHost, st_), "copy of energies");
}

void BatchStep::finish() { check(hipStreamSynchronize(st_), "hipStreamSynchronize"); }

void BatchStep::download(double* dst, const double* src, size_t n) {
  check(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, st_), "copy to host");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

void BatchStep::upload(double* dst, const double* src, size_t n) {
  check(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, st_), "copy to device");
  check(hipStreamSynchronize(st_), "hipStreamSynchronize");
}

}  // namespace dftcxx
