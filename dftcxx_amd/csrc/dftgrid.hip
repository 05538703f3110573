// MI355X (gfx950) grid-integration library: device kernels and the C ABI.
//
// Device data layout (HBM), per handle:
//   points      SoA px/py/pz/w/wb, npts_pad = npts rounded up to 256, padding
//               points carry zero weight and zero amplitudes;
//   amplitudes  phi[npts_pad][ldphi], ldphi = nbf rounded up to 32, zero padded;
//   Hartree     rho_lm, U [natoms][nrad][nlm]; spline table [natoms][nrad][nlm]{y,D};
//               partial potentials vpart[natoms][npts_pad];
//   matrices    nbf x nbf column-major.
// Kernels (one SCF build, in launch order):
//   pfrag -> density (FP64 MFMA) -> cell_sum -> scale -> cell_sum -> lda ->
//   rho_lm -> poisson -> spline -> vcell -> hartree_cross -> vreduce ->
//   contract (FP64 MFMA) -> contract_reduce -> scalars.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dftgrid.h"
#include "grid_tables.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace

void dg::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return fail(DG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));         \
  } while (0)

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// s_waitcnt immediate that waits until at most n of this wave's vector-memory
// operations are outstanding (gfx9 encoding: vmcnt in bits 3:0 and 15:14,
// expcnt 6:4 and lgkmcnt 11:8 left at their maxima)
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// Wave-level sum in a fixed butterfly order (all 64 lanes end with the total).
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

constexpr int kMaxL = 12;  // largest lmax compiled
constexpr double kPiD = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Setup kernels (once per geometry)
// ---------------------------------------------------------------------------

// Position R_a + r_i Omega_j of a grid point (explicit fused multiply-adds: the
// grid kernel and the cross-potential kernel, which re-forms the coordinates
// from the indices, see the same values).
__device__ __forceinline__ void grid_xyz(const double* __restrict__ atoms, const double* __restrict__ rad,
                                         const double* __restrict__ leb, int a, int i, int j, double& x, double& y,
                                         double& z) {
  const double r = rad[i];
  x = __fma_rn(leb[4 * j], r, atoms[4 * a]);
  y = __fma_rn(leb[4 * j + 1], r, atoms[4 * a + 1]);
  z = __fma_rn(leb[4 * j + 2], r, atoms[4 * a + 2]);
}

// Grid points: atom a, radial node i, Lebedev point j of p = (a*nrad + i)*nang + j.
// Position R_a + r_i*Omega_j, weight w_i * w_j * |r_i Omega_j|^2 * 4 pi
// (src/atomicgrid.cpp:50-87).
__global__ void k_grid(const double* __restrict__ atoms, const double* __restrict__ rad,
                       const double* __restrict__ wrad, const double* __restrict__ leb, long npts,
                       long npts_pad, int nrad, int nang, double* px, double* py, double* pz,
                       double* w) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts_pad) return;
  if (p >= npts) {
    px[p] = py[p] = pz[p] = 0.0;
    w[p] = 0.0;
    return;
  }
  const int j = (int)(p % nang);
  const long ai = p / nang;
  const int i = (int)(ai % nrad);
  const int a = (int)(ai / nrad);
  const double r = rad[i];
  const double ox = leb[4 * j] * r, oy = leb[4 * j + 1] * r, oz = leb[4 * j + 2] * r;
  grid_xyz(atoms, rad, leb, a, i, j, px[p], py[p], pz[p]);
  double wp = wrad[i] * leb[4 * j + 3];
  wp *= (ox * ox + oy * oy + oz * oz) * 4.0 * kPiD;
  w[p] = wp;
}

// Becke fuzzy-cell weight of every point (src/moleculargrid.cpp:229-329):
// w_a = P_a / sum_n P_n, P_n = prod_{m != n} s(mu_nm), s = (1 - f(f(f(mu))))/2,
// f(x) = 1.5x - 0.5x^3, mu_nm = (|r-R_n| - |r-R_m|) / |R_n-R_m|.
// One wave per block; the point's atom distances live in LDS [natoms][64].
__global__ __launch_bounds__(64) void k_becke(const double* __restrict__ atoms,
                                              const double* __restrict__ invR, int natoms, long npts,
                                              long cell, const double* __restrict__ px,
                                              const double* __restrict__ py,
                                              const double* __restrict__ pz, double* w, double* wb) {
  extern __shared__ double dist[];
  const int lane = threadIdx.x;
  const long p = (long)blockIdx.x * 64 + lane;
  const bool valid = p < npts;
  const double x = valid ? px[p] : 0.0, y = valid ? py[p] : 0.0, z = valid ? pz[p] : 0.0;
  for (int m = 0; m < natoms; ++m) {
    const double dx = x - atoms[4 * m], dy = y - atoms[4 * m + 1], dz = z - atoms[4 * m + 2];
    dist[m * 64 + lane] = sqrt(dx * dx + dy * dy + dz * dz);
  }
  if (!valid) {
    if (p < (long)gridDim.x * 64) wb[p] = 0.0;
    return;
  }
  const int own = (int)(p / cell);
  double denom = 0.0, nom = 1.0;
  for (int n = 0; n < natoms; ++n) {
    const double dn = dist[n * 64 + lane];
    double prod = 1.0;
    for (int m = 0; m < natoms; ++m) {
      if (m == n) continue;
      double mu = (dn - dist[m * 64 + lane]) * invR[n * natoms + m];
      for (int it = 0; it < 3; ++it) mu = 1.5 * mu - 0.5 * (mu * mu * mu);
      prod *= 0.5 * (1.0 - mu);
    }
    denom += prod;
    if (n == own) nom = prod;
  }
  const double v = nom / denom;
  wb[p] = v;
  w[p] *= v;
}

// Basis amplitudes phi[p][cgf], one thread per (point, shell).  A shell's
// primitives share one exponential per point; the CGF value is
// sum_k c_k N_k x^l y^m z^n exp(-alpha_k r^2) (src/cgf.cpp:49-57, 146-154).
// cn0 = c N for (L,0,0)-type components, cn1 = c N for mixed d components.
__device__ __forceinline__ void phi_shell(const double* __restrict__ atoms, const int* __restrict__ sh_atom,
                                          const int* __restrict__ sh_L, const int* __restrict__ sh_nprim,
                                          const int* __restrict__ sh_off, const int* __restrict__ sh_cgf,
                                          const double* __restrict__ alpha, const double* __restrict__ cn0,
                                          const double* __restrict__ cn1, const double* __restrict__ px,
                                          const double* __restrict__ py, const double* __restrict__ pz,
                                          int ldphi, double* phi, long p, int s) {
  const int a = sh_atom[s];
  const double dx = px[p] - atoms[4 * a], dy = py[p] - atoms[4 * a + 1], dz = pz[p] - atoms[4 * a + 2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  double r0 = 0.0, r1 = 0.0;
  const int off = sh_off[s], np = sh_nprim[s];
  for (int k = 0; k < np; ++k) {
    const double e = exp(-alpha[off + k] * r2);
    r0 += cn0[off + k] * e;
    r1 += cn1[off + k] * e;
  }
  double* out = phi + p * ldphi + sh_cgf[s];
  switch (sh_L[s]) {
    case 0:
      out[0] = r0;
      break;
    case 1:
      out[0] = r0 * dx;
      out[1] = r0 * dy;
      out[2] = r0 * dz;
      break;
    default:
      out[0] = r0 * dx * dx;
      out[1] = r1 * dx * dy;
      out[2] = r1 * dx * dz;
      out[3] = r0 * dy * dy;
      out[4] = r1 * dy * dz;
      out[5] = r0 * dz * dz;
      break;
  }
}

__global__ void k_phi(const double* __restrict__ atoms, int nshell, const int* __restrict__ sh_atom,
                      const int* __restrict__ sh_L, const int* __restrict__ sh_nprim,
                      const int* __restrict__ sh_off, const int* __restrict__ sh_cgf,
                      const double* __restrict__ alpha, const double* __restrict__ cn0,
                      const double* __restrict__ cn1, long npts, const double* __restrict__ px,
                      const double* __restrict__ py, const double* __restrict__ pz, int ldphi,
                      double* phi) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts * nshell) return;
  phi_shell(atoms, sh_atom, sh_L, sh_nprim, sh_off, sh_cgf, alpha, cn0, cn1, px, py, pz, ldphi, phi, g / nshell,
            (int)(g % nshell));
}

// Screening mask: bit T of mask[b] is set when some |phi| of basis tile T
// (16 functions) exceeds screen on the 64-point block b.  One wave per block.
__global__ __launch_bounds__(64) void k_block_mask(const double* __restrict__ phi, int ldphi, int ntb,
                                                   double screen, unsigned* mask) {
  const long p = (long)blockIdx.x * 64 + threadIdx.x;
  const double* row = phi + p * ldphi;
  unsigned bits = 0u;
  for (int T = 0; T < ntb; ++T) {
    double m = 0.0;
    for (int i = 0; i < 16; ++i) m = fmax(m, fabs(row[16 * T + i]));
    if (m > screen) bits |= 1u << T;
  }
  for (int o = 32; o > 0; o >>= 1) bits |= (unsigned)__shfl_xor((int)bits, o);
  if (threadIdx.x == 0) mask[blockIdx.x] = bits;
}

// ---------------------------------------------------------------------------
// Density: rho_p = 2 phi_p^T P phi_p  (src/gridpoint.cpp:82-84)
// ---------------------------------------------------------------------------

// P in the A-operand order of the 16x16x4 FP64 matrix-core op, halved by
// symmetry: rho/2 = sum_b phi_b^T T_b with T_b = sum_{b' <= b} P'_{bb'} phi_b',
// P'_{bb'} = 2 P_{bb'} for b' < b and P_{bb} on the diagonal tile (16-wide tiles).
// Tile b holds 4(b+1) reduction steps s; entry [s][lane] is
// P'(16b + pi(lane%16), kk(s, lane/16)), kk(s,g) = 16(s/4) + 4g + s%4,
// pi(r) = 4(r%4) + r/4, stored from offset 128 b (b+1).  The same permutation
// kk of the reduction index is used for phi in k_density, and pi puts
// T_b[p][16b + 4g + v] in result register v of lane (c = p%16, g): the lane
// that holds phi[p][16b + 4g + v] as its B operand of step 4b + v, so the
// final dot product needs no further loads.
__device__ __forceinline__ int pfrag_row(int r) { return 4 * (r & 3) + (r >> 2); }
__device__ __forceinline__ double pfrag_value(const double* __restrict__ P, int nbf, long e) {
  int b = 0;
  while (128L * (b + 1) * (b + 2) <= e) ++b;
  const int rem = (int)(e - 128L * b * (b + 1));
  const int s = rem / 64, l = rem % 64;
  const int j = 16 * b + pfrag_row(l & 15);
  const int k = 16 * (s / 4) + 4 * (l >> 4) + (s & 3);
  const double f = (s / 4 < b) ? 2.0 : 1.0;
  return (j < nbf && k < nbf) ? f * P[j + (long)k * nbf] : 0.0;
}

__global__ void k_pfrag(const double* __restrict__ P, int nbf, int nt, double* frag) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 128L * nt * (nt + 1)) return;
  frag[e] = pfrag_value(P, nbf, e);
}

// One wave handles 16 points; lane (c = lane%16, g = lane/16) keeps
// phi[p0+c][kk(s,g)] for every reduction step s in registers (B operand).
// For output tile b the wave forms T_b[p][16b+4v+g] (result register v) with
// 4(b+1) MFMAs in two accumulation chains (steps j even / odd), then dots it
// with phi[p][16b+4v+g], which is its own B-operand register f[4b+v] (the tile
// loop is unrolled, so every register index is a constant).  The P fragments
// of one tile are staged in LDS and shared by the work-group's W waves
// (16 W points), copied from global memory straight into LDS (global_load_lds,
// 1 KB per wave instruction).  NT = ceil(nbf/16).
// NBUF = 1: one tile buffer (<= 50 KB at NT <= 25), so that two work-groups of
// W = 4 share a CU, one wave of each per SIMD: while one work-group copies a
// tile, waits at its barriers or loads its amplitude rows, the other issues
// MFMAs.  NBUF = 2: the next live tile is copied into the other buffer while the
// current tile's MFMAs run.
// With screening, mask (one word per 64-point block, bit T for basis tile T)
// drops the tiles whose amplitudes are negligible on the whole block: a wave
// skips the MFMAs of the tiles off its block's word, and the work-group walks
// (and fetches) the tiles on any of its blocks' words, so every branch on the
// walk is uniform over the work-group.
template <int NT, int W, int NBUF>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(NBUF == 1 ? 2 : 1))) void k_density(
    const double* __restrict__ phi, int ldphi, const double* __restrict__ frag, double* rho,
    const unsigned* __restrict__ mask) {
  static_assert(W % 4 == 0 && (NBUF == 1 || NBUF == 2), "whole 64-point blocks; one or two tile buffers");
  __shared__ double pl[NBUF][4 * NT * 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long p0 = (long)blockIdx.x * 16 * W + wave * 16;
  const double* row = phi + (p0 + c) * ldphi;
  constexpr unsigned all = NT >= 32 ? 0xffffffffu : ((1u << NT) - 1u);
  const long blk0 = (long)blockIdx.x * W / 4;  // first 64-point block of the work-group
  // wave-uniform (the compiler cannot see that from threadIdx)
  const unsigned mine = __builtin_amdgcn_readfirstlane((mask ? mask[blk0 + wave / 4] : 0xffffffffu) & all);
  unsigned mk = 0;
  for (int b = 0; b < W / 4; ++b) mk |= (mask ? mask[blk0 + b] : 0xffffffffu) & all;
  double f[4 * NT];
#pragma unroll
  for (int s = 0; s < 4 * NT; ++s) f[s] = ((mine >> (s >> 2)) & 1u) ? row[16 * (s >> 2) + 4 * g + (s & 3)] : 0.0;
  // tile b's fragments, 256 (b + 1) doubles from offset 128 b (b + 1): chunks of
  // 128 doubles, chunk q copied by wave q % W, 16 bytes per lane
  auto fetch = [&](int b, double* dst) {
    const double* src = frag + 128L * b * (b + 1);
    for (int q = wave; q < 2 * (b + 1); q += W)
      __builtin_amdgcn_global_load_lds((const void*)(src + q * 128 + lane * 2),
                                       (__attribute__((address_space(3))) void*)(dst + q * 128), 16, 0, 0);
  };
  auto next_live = [&](int b) {  // first tile > b with its mask bit set, NT if none
    const unsigned rest = (b + 1 >= 32) ? 0u : (mk >> (b + 1));
    return rest ? b + 1 + __builtin_ctz(rest) : NT;
  };
  double acc = 0.0;
  int buf = 0;
  if (NBUF == 2) {
    const int b0 = next_live(-1);
    if (b0 < NT) fetch(b0, pl[0]);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // tiles in compile-time order (each skipped when off the mask), so that the
  // amplitudes of the final dot are the B-operand registers f[4b + v]
#pragma unroll
  for (int b = 0; b < NT; ++b) {
    if (!((mk >> b) & 1u)) continue;
    if (NBUF == 2) {
      const int nb = next_live(b);
      if (nb < NT) fetch(nb, pl[buf ^ 1]);
    } else {
      fetch(b, pl[0]);
      __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
      __syncthreads();
    }
    const double* cur = pl[buf];
    // two accumulation chains (steps j even / odd) over the reduction tiles k <= b
    dbl4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k <= b; ++k) {
      if ((mine >> k) & 1u) {
        double av[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) av[j] = cur[(4 * k + j) * 64 + lane];
        d0 = mfma16x16x4(av[0], f[4 * k], d0);
        d1 = mfma16x16x4(av[1], f[4 * k + 1], d1);
        d0 = mfma16x16x4(av[2], f[4 * k + 2], d0);
        d1 = mfma16x16x4(av[3], f[4 * k + 3], d1);
      }
    }
    // register v of lane (c, g) holds T_b[p0+c][16b + 4g + v]; f[4b + v] = phi[p0+c][16b + 4g + v]
    if ((mine >> b) & 1u)
      acc += (d0[0] + d1[0]) * f[4 * b] + (d0[1] + d1[1]) * f[4 * b + 1] + (d0[2] + d1[2]) * f[4 * b + 2] +
             (d0[3] + d1[3]) * f[4 * b + 3];
    if (NBUF == 2) {
      __builtin_amdgcn_s_waitcnt(0);  // this wave's copies into the other buffer have landed
      __syncthreads();
      buf ^= 1;
    } else {
      __syncthreads();  // every wave is done with the buffer before the next copy
    }
  }
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (g == 0) rho[p0 + c] = 2.0 * acc;
}

// k_density without screening, its P fragments streamed through a ring.  The
// fragments are stored tile after tile, reduction tile after reduction tile
// (unit (b, k), k <= b: 4 steps x 64 lanes = 2 KB), which is the order the
// tiles consume them; the stream is cut into chunks of kDensityChunk units
// (16 KB, four 1 KB copy instructions per wave; the last chunk ends at the end
// of the stream and overlaps the one before) and each chunk is copied two
// chunks ahead into one of three LDS slots (48 KB: two work-groups per CU).  One wait and barrier per chunk, which then also
// tells that every wave is done with the slot the next copy overwrites.  W = 4
// waves share the chunks.
constexpr int kDensityChunk = 8;

// Chunks of C units of a fragment stream of U >= C units (256 doubles each):
// chunk j starts at unit C j, except the last, which ends at unit U (so every
// chunk is whole and no copy reads past the stream).
template <int C, int U>
__device__ __forceinline__ constexpr int stream_start(int j) {
  return j < (U + C - 1) / C - 1 ? j * C : U - C;
}

// Copy of chunk j into LDS slot dst: 1 KB pieces, wave w takes pieces w, w + 4,
// ..., C / 2 instructions per wave.  (Every wave issues the same count, which the
// wait below relies on; a conditional copy here would also cost registers.)
template <int C, int U>
__device__ __forceinline__ void stream_copy(const double* src, double* dst, int j, int wave, int lane) {
  static_assert(U >= C && C % 2 == 0, "whole chunks");
  const double* s0 = src + (long)stream_start<C, U>(j) * 256;
#pragma unroll
  for (int q = 0; q < C * 2 / 4; ++q) {
    const int piece = wave + 4 * q;
    __builtin_amdgcn_global_load_lds((const void*)(s0 + piece * 128 + lane * 2),
                                     (__attribute__((address_space(3))) void*)(dst + piece * 128), 16, 0, 0);
  }
}

// The wait at the start of chunk j, chunk j + 1 in flight (its C / 2 copies).
template <int C, int U>
__device__ __forceinline__ void stream_wait(int j) {
  constexpr int NCH = (U + C - 1) / C;
  if (j + 1 < NCH)
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(C * 2 / 4));
  else
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
}

// A work-group barrier without the release fence of __syncthreads (which would
// also wait for every copy in flight): this wave's LDS reads are waited for
// (lgkmcnt 0, vmcnt and expcnt left at their maxima), since the compiler may
// issue a read before the barrier and consume it after, while the slot it reads
// is overwritten by another wave's copy issued after the barrier; the empty asm
// statements keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void bare_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((15) | (3 << 14) | (7 << 4));  // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_density_ring(
    const double* __restrict__ phi, int ldphi, const double* __restrict__ frag, double* rho) {
  constexpr int C = kDensityChunk, U = NT * (NT + 1) / 2, NCH = (U + C - 1) / C;
  __shared__ double ring[3][C * 256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long p0 = (long)blockIdx.x * 64 + wave * 16;
  const double* row = phi + (p0 + c) * ldphi;
  double f[4 * NT];
#pragma unroll
  for (int s = 0; s < 4 * NT; ++s) f[s] = row[16 * (s >> 2) + 4 * g + (s & 3)];
  auto copy = [&](int j) { stream_copy<C, U>(frag, ring[j % 3], j, wave, lane); };
  copy(0);
  if (NCH > 1) copy(1);
  double acc = 0.0;
  dbl4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0;
  int u = 0;
#pragma unroll
  for (int b = 0; b < NT; ++b) {
#pragma unroll
    for (int k = 0; k <= b; ++k, ++u) {
      if (u % C == 0) {
        const int j = u / C;
        stream_wait<C, U>(j);  // this wave's copies of chunk j have landed
        bare_barrier();        // and every wave's; all are done with chunk j - 1
        if (j + 2 < NCH) copy(j + 2);  // into the slot of chunk j - 1
      }
      if (k == 0) d0 = d1 = dbl4{0.0, 0.0, 0.0, 0.0};
      const double* cur = ring[(u / C) % 3] + (u - stream_start<C, U>(u / C)) * 256;
      double av[4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) av[jj] = cur[jj * 64 + lane];
      d0 = mfma16x16x4(av[0], f[4 * k], d0);
      d1 = mfma16x16x4(av[1], f[4 * k + 1], d1);
      d0 = mfma16x16x4(av[2], f[4 * k + 2], d0);
      d1 = mfma16x16x4(av[3], f[4 * k + 3], d1);
    }
    // register v of lane (c, g) holds T_b[p0+c][16b + 4g + v]; f[4b + v] = phi[p0+c][16b + 4g + v]
    acc += (d0[0] + d1[0]) * f[4 * b] + (d0[1] + d1[1]) * f[4 * b + 1] + (d0[2] + d1[2]) * f[4 * b + 2] +
           (d0[3] + d1[3]) * f[4 * b + 3];
  }
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (g == 0) rho[p0 + c] = 2.0 * acc;
}

// k_density for small bases (NT <= 7): the whole fragment set (128 NT(NT+1)
// doubles, formed from P as k_pfrag does) is staged in LDS once per work-group, which then walks the 64-point
// blocks b = blockIdx.x, blockIdx.x + G, ... with no further barriers; each
// wave loads the next block's amplitude row while it multiplies the current.
template <int NT>
__global__ __launch_bounds__(256) void k_density_pl(const double* __restrict__ phi, int ldphi,
                                                    const double* __restrict__ P, int nbf, long nblk, double* rho,
                                                    const unsigned* __restrict__ mask) {
  __shared__ double pl[128 * NT * (NT + 1)];
  for (int e = threadIdx.x; e < 128 * NT * (NT + 1); e += 256) pl[e] = pfrag_value(P, nbf, e);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  double f[4 * NT];
  long blk = blockIdx.x;
  if (blk < nblk) {
    const double* row = phi + (blk * 64 + wave * 16 + c) * ldphi;
#pragma unroll
    for (int s = 0; s < 4 * NT; ++s) f[s] = row[16 * (s >> 2) + 4 * g + (s & 3)];
  }
  for (; blk < nblk; blk += gridDim.x) {
    const long p0 = blk * 64 + wave * 16;
    const unsigned mk = mask ? mask[blk] : 0xffffffffu;
    double cf[4 * NT];
#pragma unroll
    for (int s = 0; s < 4 * NT; ++s) cf[s] = ((mk >> (s >> 2)) & 1u) ? f[s] : 0.0;
    const long nb = blk + gridDim.x;
    if (nb < nblk) {
      const double* row = phi + (nb * 64 + wave * 16 + c) * ldphi;
#pragma unroll
      for (int s = 0; s < 4 * NT; ++s) f[s] = row[16 * (s >> 2) + 4 * g + (s & 3)];
    }
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      if (!((mk >> b) & 1u)) continue;
      const double* src = pl + 128 * b * (b + 1);
      dbl4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4 * (b + 1); ++s)
        if ((mk >> (s >> 2)) & 1u) d = mfma16x16x4(src[s * 64 + lane], cf[s], d);
      // register v of lane (c, g) holds T_b[p0+c][16b + 4g + v]
      acc += d[0] * cf[4 * b] + d[1] * cf[4 * b + 1] + d[2] * cf[4 * b + 2] + d[3] * cf[4 * b + 3];
    }
    acc += __shfl_xor(acc, 16);
    acc += __shfl_xor(acc, 32);
    if (g == 0) rho[p0 + c] = 2.0 * acc;
  }
}

// out[a] = sum over the points of cell a of x*y, fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_cell_sum(const double* __restrict__ x,
                                                   const double* __restrict__ y, long cell,
                                                   double* out) {
  __shared__ double red[1024];
  const long base = (long)blockIdx.x * cell;
  double s = 0.0;
  long i = threadIdx.x;
  for (; i + 7 * 1024 < cell; i += 8 * 1024) {  // eight products per memory round, added in order
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = x[base + i + u * 1024] * y[base + i + u * 1024];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += t[u];
  }
  for (; i < cell; i += 1024) s += x[base + i] * y[base + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// The same sums split S ways per cell (block b: cell b / S, part b % S), for
// molecules with few atoms, whose cells would otherwise occupy few CUs.
__global__ __launch_bounds__(256) void k_cell_sum_split(const double* __restrict__ x,
                                                        const double* __restrict__ y, long cell, int S,
                                                        double* part) {
  __shared__ double red[256];
  const int a = blockIdx.x / S, sp = blockIdx.x % S;
  const long base = (long)a * cell, lo = cell * sp / S, hi = cell * (sp + 1) / S;
  double s = 0.0;
  long i = lo + threadIdx.x;
  for (; i + 7 * 256 < hi; i += 8 * 256) {  // eight products per memory round, added in order
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = x[base + i + u * 256] * y[base + i + u * 256];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += t[u];
  }
  for (; i < hi; i += 256) s += x[base + i] * y[base + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// k_scale from the split populations: q_a = sum_s part[a S + s], in order;
// the total is summed by the first wave of every block in one fixed order
// (lane strides, then a butterfly), so every block gets the same factor.
__global__ __launch_bounds__(256) void k_scale_parts(double* rho, long n, const double* __restrict__ part, int S,
                                                     int natoms, double zsum, double* q, double* qc,
                                                     double* scalars) {
  __shared__ double qsum;
  if (threadIdx.x < 64) {
    double t = 0.0;
    for (int e = threadIdx.x; e < natoms * S; e += 64) t += part[e];
    t = wave_sum(t);
    if (threadIdx.x == 0) qsum = t;
  }
  __syncthreads();
  const double qs = qsum;
  const double f = zsum / qs;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) rho[p] *= f;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0 && scalars) scalars[1] = qs;
    for (int a = threadIdx.x; a < natoms; a += blockDim.x) {
      double qa = 0.0;
      for (int sp = 0; sp < S; ++sp) qa += part[a * S + sp];
      q[a] = qa;
      qc[a] = f * qa;
    }
  }
}

// rho *= sum(Z) / sum_a q_a  (src/moleculargrid.cpp:132-146); scalars[1] = sum q
// before; block 0 writes the populations of the corrected density, qc_a = f q_a.
__global__ void k_scale(double* rho, long n, const double* __restrict__ q, int natoms, double zsum,
                        double* qc, double* scalars) {
  double qs = 0.0;
  for (int a = 0; a < natoms; ++a) qs += q[a];
  const double f = zsum / qs;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) rho[p] *= f;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0 && scalars) scalars[1] = qs;
    for (int a = threadIdx.x; a < natoms; a += blockDim.x) qc[a] = f * q[a];
  }
}

// ---------------------------------------------------------------------------
// LDA: Slater exchange (alpha = 2/3) + VWN correlation (VWN5 parameters,
// paramagnetic branch: rho_a = rho_b = rho/2 gives zeta = 0), per point
// (src/functionals.cpp:24-158, src/dft.cpp:407-424).
// uxc = w (v_xa + v_xb + v_ca + v_cb)/2, exc_part[block] = sum w (e_x + e_c).
// ---------------------------------------------------------------------------
struct VwnConst {
  double A, x0, b, c, Q, X0, k1, k2;
};

__device__ __forceinline__ void vwn_para(double x, const VwnConst& k, double& eps, double& deps) {
  // eps(x) = A [ ln(x^2/X) - k2 ln((x-x0)^2/X) + k1 atan(Q/(2x+b)) ],  X = x^2+bx+c,
  // k1 = (2b/Q)(1 - x0(2x0+b)/X0), k2 = b x0/X0; deps = d eps/dx.
  const double X = x * x + k.b * x + k.c;
  const double tx = 2.0 * x + k.b;
  const double at = atan(k.Q / tx);
  eps = k.A * (log(x * x / X) - k.k2 * log((x - k.x0) * (x - k.x0) / X) + k.k1 * at);
  const double qq = tx * tx + k.Q * k.Q;
  deps = k.A * (2.0 / x - tx / X - 4.0 * k.b / qq -
                k.k2 * (2.0 / (x - k.x0) - tx / X - 4.0 * (2.0 * k.x0 + k.b) / qq));
}

// SCALE (the device build, fused with the renormalisation): rho is scaled
// here, by f = sum Z / sum q computed in every block exactly as k_scale_parts
// (SCALE = 1, populations split S ways in part) or k_scale (SCALE = 2, S = 1,
// per-atom q in part) compute it, and written back; block 0 also writes q, qc
// and scalars[1] as those kernels do.
template <int SCALE>
__global__ __launch_bounds__(256) void k_lda(double* rho, const double* __restrict__ w, long n, VwnConst vk,
                                             double xfac, double* uxc, double* exd, double* exc_part,
                                             const double* __restrict__ part, int S, int natoms, double zsum,
                                             double* q, double* qc, double* scalars) {
  __shared__ double red[256];
  __shared__ double qsum;
  double f = 1.0;
  if constexpr (SCALE == 1) {
    if (threadIdx.x < 64) {
      double t = 0.0;
      for (int e = threadIdx.x; e < natoms * S; e += 64) t += part[e];
      t = wave_sum(t);
      if (threadIdx.x == 0) qsum = t;
    }
    __syncthreads();
    f = zsum / qsum;
    if (blockIdx.x == 0) {
      if (threadIdx.x == 0) scalars[1] = qsum;
      for (int a = threadIdx.x; a < natoms; a += blockDim.x) {
        double qa = 0.0;
        for (int sp = 0; sp < S; ++sp) qa += part[a * S + sp];
        q[a] = qa;
        qc[a] = f * qa;
      }
    }
  } else if constexpr (SCALE == 2) {
    double qs = 0.0;
    for (int a = 0; a < natoms; ++a) qs += part[a];
    f = zsum / qs;
    if (blockIdx.x == 0) {
      if (threadIdx.x == 0) scalars[1] = qs;
      for (int a = threadIdx.x; a < natoms; a += blockDim.x) qc[a] = f * part[a];
    }
  }
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  double e = 0.0;
  if (SCALE > 0 && p < n) rho[p] *= f;
  if (p < n) {
    const double r = rho[p];
    const double ra = r * 0.5;
    double ex = 0.0, vx = 0.0, ec = 0.0, vc = 0.0;
    if (ra >= 1e-10) {
      const double r3 = cbrt(ra);
      ex = 2.0 * (xfac * ra * r3);
      vx = 4.0 / 3.0 * xfac * r3;
    }
    if (r >= 1e-10) {
      const double x = sqrt(cbrt(3.0 / 4.0 / kPiD / r));  // (3/(4 pi rho))^{1/6}
      double eps, deps;
      vwn_para(x, vk, eps, deps);
      ec = eps * r;
      vc = eps - (x / 6.0) * deps;
    }
    const double wp = w[p];
    uxc[p] = wp * ((vx + vx + vc + vc) * 0.5);
    exd[p] = ex + ec;
    e = wp * (ex + ec);
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) exc_part[blockIdx.x] = red[0];
}

// scalars[0] = sum of exc_part (fixed tree order); scalars[2] = sum q;
// scalars[3] = E_J = 2 tr(P J) evaluated pointwise: J = (1/2) Phi^T diag(w V) Phi
// and rho = c 2 phi^T P phi give 2 tr(P J) = sum w V rho / (2c), c = sum Z / q_raw.
// The whole work-group (256 threads) evaluates them; out2 (if set) receives a copy
// of all four.
struct ScalarArgs {
  const double* exc_part;
  long nparts;
  const double* ej_part;
  long nej;
  const double* q;
  int natoms;
  double zsum;
  double* scalars;
  double* out2;
};

__device__ void scalars_block(const ScalarArgs& a) {
  __shared__ double red[256];
  __shared__ double red2[256];
  const double* __restrict__ exc_part = a.exc_part;
  const double* __restrict__ ej_part = a.ej_part;
  const long nparts = a.nparts, nej = a.nej;
  double* scalars = a.scalars;
  double s = 0.0, s2 = 0.0;
  for (long i = threadIdx.x; i < nparts; i += 256) s += exc_part[i];
  if (ej_part)
    for (long i = threadIdx.x; i < nej; i += 256) s2 += ej_part[i];
  red[threadIdx.x] = s;
  red2[threadIdx.x] = s2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      red[threadIdx.x] += red[threadIdx.x + st];
      red2[threadIdx.x] += red2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    scalars[0] = red[0];
    if (a.q) {
      double qs = 0.0;
      for (int at = 0; at < a.natoms; ++at) qs += a.q[at];
      scalars[2] = qs;
    }
    scalars[3] = ej_part ? red2[0] * scalars[1] / (2.0 * a.zsum) : 0.0;
    if (a.out2)
      for (int i = 0; i < 4; ++i) a.out2[i] = scalars[i];
  }
}

__global__ __launch_bounds__(256) void k_scalars(ScalarArgs a) { scalars_block(a); }

// ---------------------------------------------------------------------------
// Hartree potential (src/atomicgrid.cpp:241-466, src/moleculargrid.cpp:336-389)
// ---------------------------------------------------------------------------

// The three small per-atom products of the Hartree stage run as FP64 MFMA
// tiles (lane (c, g) feeds A[c][g] and B[g][c]; result register v holds row
// 4v + g, column c), summing the inner index in steps of 4 with zero padding.
// Each wave loads the operands of kSteps steps (kSteps2 in the Poisson kernel,
// kStepsV in the own-cell one: 84 inner indices, all 81 harmonics of lmax 8)
// before it issues their MFMAs, so that one memory latency covers them.
constexpr int kSteps = 4;
constexpr int kSteps2 = 8;
constexpr int kStepsV = 21;

// rho_lm[a][i][lm] = 4 pi sum_j F[i][j] Y[j][lm] with F[i][j] = rho_p w_j wb_p,
// p = (a nrad + i) nang + j (src/atomicgrid.cpp:241-296).  Block (a, 16 shells
// i0.., group of NT lm tiles): the 8 waves split the angular points (wave w
// takes j = 4 (w + 8 t) + 0..3), each A fragment feeds the NT tiles, and the
// wave partials are added in wave order through LDS.
constexpr int kRhoWaves = 8;
template <int NT>
__global__ __launch_bounds__(64 * kRhoWaves) void k_rho_lm(const double* __restrict__ rho,
                                                           const double* __restrict__ wb,
                                                           const double* __restrict__ leb,
                                                           const double* __restrict__ ylm, int nrad, int nang,
                                                           int nlm, double* rho_lm) {
  __shared__ double part[kRhoWaves][NT * 256];
  const int ti = (nrad + 15) / 16;
  const int a = blockIdx.x / ti, i0 = 16 * (blockIdx.x % ti), l0 = 16 * NT * blockIdx.y;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int i = i0 + c;
  dbl4 d[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) d[t] = dbl4{0.0, 0.0, 0.0, 0.0};
  const long prow = ((long)a * nrad + i) * nang;
  for (int jb = 4 * wave; jb < nang; jb += 4 * kRhoWaves * kSteps) {
    double av[kSteps], bv[kSteps][NT];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
      const int j = jb + 4 * kRhoWaves * u + g;
      av[u] = (j < nang && i < nrad) ? rho[prow + j] * leb[4 * j + 3] * wb[prow + j] : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int lm = l0 + 16 * t + c;
        bv[u][t] = (j < nang && lm < nlm) ? ylm[(long)j * nlm + lm] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kSteps; ++u)
#pragma unroll
      for (int t = 0; t < NT; ++t) d[t] = mfma16x16x4(av[u], bv[u][t], d[t]);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) part[wave][t * 256 + v * 64 + lane] = d[t][v];
  __syncthreads();
  for (int e = threadIdx.x; e < NT * 256; e += 64 * kRhoWaves) {
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < kRhoWaves; ++w) sum += part[w][e];
    const int t = e >> 8, v = (e >> 6) & 3, ln = e & 63;
    const int row = i0 + 4 * v + (ln >> 4), lm = l0 + 16 * t + (ln & 15);
    if (row < nrad && lm < nlm) rho_lm[((long)a * nrad + row) * nlm + lm] = sum * (4.0 * kPiD);
  }
}

// c_lm of the scaled associated-Legendre recurrence: the standard recurrence
// P_l = ((2l-1) ct P_{l-1} - (l+m-1) P_{l-2}) / (l-m) over Q_l = P_l / c_lm, with
// c_mm = 1, c_{m+1,m} = 2m+1, c_lm = c_{l-1,m} (2l-1)/(l-m), needs one multiply
// fewer per step (Q_{m+1} = ct Q_m, Q_l = ct Q_{l-1} - beta_l Q_{l-2}).
__host__ __device__ inline double legendre_scale(int l, int m) {
  double c = l >= m + 1 ? 2.0 * m + 1.0 : 1.0;
  for (int k = m + 2; k <= l; ++k) c *= (double)(2 * k - 1) / (double)(k - m);
  return c;
}

// Radial Poisson solve and spline slopes (src/atomicgrid.cpp:395-432,
// src/cspline.cpp:66-127, src/atomicgrid.cpp:535-556): per l, with the columns
// (a, m) of all atoms, U_i = sum_k KU_l[k][i] g_k and D_t = sum_k KD_l[k][t] g_k,
// where g_0 = sqrt(4 pi) q_a [lm = 0], g_k = -4 pi r_{k-1} rho_lm[a][k-1][lm],
// g_{N+1} = 0, KU_l = rows 1..N of the inverse FD matrix and KD_l = S KU_l (the
// not-a-knot slope map S over the ascending knots).  Block x = (l, row tile,
// column tile, output U or D); tiles past the l's columns exit.
// tab[a][t][lm] = N_lm c_lm {U at knot t, D_t}; knot t is radial node N-1-t, and
// N_lm c_lm (the harmonic's normalisation and the recurrence scale, shc) is folded
// in for the cross kernel.
// (Four waves per work-group, one tile each: tile 4 blockIdx.x + wave of the
// ntiles; small work-groups would cap the waves resident per CU.)
__global__ __launch_bounds__(256) void k_poisson(const double* __restrict__ rho_lm, const double* __restrict__ q,
                                                 const double* __restrict__ rad, const double* __restrict__ KU,
                                                 const double* __restrict__ KD, const double* __restrict__ shc,
                                                 int natoms, int nrad, int lmax, int tcmax, long ntiles, double* U,
                                                 double* tab) {
  const int nlm = (lmax + 1) * (lmax + 1), n2 = nrad + 2, ti = (nrad + 15) / 16;
  const long tile = 4L * blockIdx.x + (threadIdx.x >> 6);
  if (tile >= ntiles) return;
  int b = (int)tile;
  const int which = b & 1;
  b >>= 1;
  const int tc = b % tcmax;
  b /= tcmax;
  const int it = b % ti, l = b / ti;
  const int w = 2 * l + 1, ncol = natoms * w;
  if (16 * tc >= ncol) return;
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int i = 16 * it + c, col = 16 * tc + c;
  const int ca = col / w, clm = l * l + col % w;
  const double* M = (which ? KD : KU) + (long)l * n2 * nrad;
  dbl4 d = {0.0, 0.0, 0.0, 0.0};
  for (int kb = 0; kb < n2; kb += 4 * kSteps2) {
    double av[kSteps2], bv[kSteps2];
#pragma unroll
    for (int u = 0; u < kSteps2; ++u) {
      const int k = kb + 4 * u + g;
      av[u] = bv[u] = 0.0;
      if (k < n2) {
        if (i < nrad) av[u] = M[(long)k * nrad + i];
        if (col < ncol) {
          if (k == 0)
            bv[u] = (clm == 0) ? sqrt(4.0 * kPiD) * q[ca] : 0.0;
          else if (k <= nrad)
            bv[u] = -4.0 * kPiD * rad[k - 1] * rho_lm[((long)ca * nrad + (k - 1)) * nlm + clm];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kSteps2; ++u) d = mfma16x16x4(av[u], bv[u], d);
  }
  if (col >= ncol) return;
  const int m = clm - l * l - l;
  const double nl = shc[3 * (l * (l + 1) / 2 + (m < 0 ? -m : m))];
  double* T = tab + (long)ca * nrad * nlm * 2 + clm * 2;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 16 * it + 4 * v + g;
    if (row >= nrad) continue;
    if (which == 0) {
      U[((long)ca * nrad + row) * nlm + clm] = d[v];
      T[(long)(nrad - 1 - row) * nlm * 2] = nl * d[v];
    } else {
      T[(long)row * nlm * 2 + 1] = nl * d[v];
    }
  }
}

// Own-cell potential V_cell(p) = sum_lm U[a][i][lm] Y_lm(Omega_j) / |r_p - R_a|
// (src/atomicgrid.cpp:434-462): tile (shells i0.., angular points j0..) of atom a.
// |r_p - R_a| is the radial node r_i (the point is R_a + r_i Omega_j with a unit
// Omega_j), which spares the coordinate reads.
// (Four waves per work-group, one tile each: tile 4 blockIdx.x + wave.)
__global__ __launch_bounds__(256) void k_vcell(const double* __restrict__ U, const double* __restrict__ ylm,
                                               const double* __restrict__ rad, int nrad, int nang, int nlm,
                                               long ntiles, double* vcell) {
  const int ti = (nrad + 15) / 16, tj = (nang + 15) / 16;
  const long tile = 4L * blockIdx.x + (threadIdx.x >> 6);
  if (tile >= ntiles) return;
  const int a = (int)(tile / (ti * tj)), rem = (int)(tile % (ti * tj));
  const int i0 = 16 * (rem / tj), j0 = 16 * (rem % tj);
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int i = i0 + c, j = j0 + c;
  dbl4 d = {0.0, 0.0, 0.0, 0.0};
  for (int kb = 0; kb < nlm; kb += 4 * kStepsV) {
    double av[kStepsV], bv[kStepsV];
#pragma unroll
    for (int u = 0; u < kStepsV; ++u) {
      const int k = kb + 4 * u + g;
      av[u] = (k < nlm && i < nrad) ? U[((long)a * nrad + i) * nlm + k] : 0.0;
      bv[u] = (k < nlm && j < nang) ? ylm[(long)j * nlm + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kStepsV; ++u) d = mfma16x16x4(av[u], bv[u], d);
  }
  if (j >= nang) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = i0 + 4 * v + g;
    if (row >= nrad) continue;
    vcell[((long)a * nrad + row) * nang + j] = d[v] / rad[row];
  }
}

struct CrossArgs {
  const double* rad;    // radial nodes (nrad)
  const double* leb;    // Lebedev directions and weights (4 nang)
  const double* atoms;
  const double* knots;  // ascending, nrad
  const double* invh;   // 1/(x_t - x_{t-1})
  const double* tab;    // [natoms][nrad][nlm]{y, D}
  const int* items;     // {k, p_begin, p_end, perm offset} point ranges outside cell k, k ascending
  const unsigned short* perm;  // per item, its points by distance to atom k, interval groups padded to even
  const unsigned short* pkey;  // interval key (spline_key) of every perm slot
  const int* plen;             // per item, the padded length of its perm range
  double* vpart;        // partial potentials, atom k's at vpart + vbase[k] + p
  const long* vbase;
  int nitems, nrad, nlm, l0, l1, accumulate, nang;
  double inv_nang, inv_nrad;  // 1/nang, 1/nrad: grid indices of a point
};

// Distance r of a point from an atom centre and its distance rxy from the
// atom's z axis, given the offset (dx, dy, dz).  One definition for the item
// sort and the potential kernel, so that both see the same r.
__device__ __forceinline__ void point_polar(double dx, double dy, double dz, double& rxy, double& r) {
  const double rxy2 = dx * dx + dy * dy;
  r = sqrt(rxy2 + dz * dz);
  rxy = sqrt(rxy2);
}

// Spline interval of r over the ascending knots kn[0..n): the reference's first
// knot s >= 1 with r <= kn[s] (the interval is [s-1, s]), or 0 / n for r below
// kn[0] / at or beyond kn[n-1], where the end values are used (src/cspline.cpp:156-168).
__device__ __forceinline__ int spline_key(double r, const double* kn, int n) {
  if (r < kn[0]) return 0;
  if (r >= kn[n - 1]) return n;
  int lo = 1, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (r <= kn[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Inclusive scan (maximum or sum) of a[0..len) by the whole work-group, ping-ponging
// between a and b; returns the buffer that holds the result.
template <bool MAX>
__device__ int* block_scan(int* a, int* b, int len) {
  for (int d = 1; d < len; d <<= 1) {
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
      const int v = a[i];
      b[i] = (i >= d) ? (MAX ? max(v, a[i - d]) : v + a[i - d]) : v;
    }
    __syncthreads();
    int* t = a;
    a = b;
    b = t;
  }
  return a;
}

// Points of every cross-potential item in order of their distance to the
// item's atom, grouped by spline interval, each group padded to an even count:
// perm[off + i] is the offset (from p_begin) of the i-th point, 0xFFFF a pad
// slot (always at an odd position), pkey[off + i] the slot's interval key
// (spline_key) and plen[item] the padded length.  Two consecutive slots from an
// even position then always share their interval.  One work-group per item (at
// most kMaxItem points), bitonic sort in LDS.
constexpr int kMaxItem = 4096;
constexpr unsigned short kPadSlot = 0xFFFF;
__global__ __launch_bounds__(1024) void k_sort_items(const int* __restrict__ items, const double* __restrict__ atoms,
                                                     const double* __restrict__ px, const double* __restrict__ py,
                                                     const double* __restrict__ pz, const double* __restrict__ knots,
                                                     int nrad, unsigned short* perm, unsigned short* pkey,
                                                     int* plen) {
  __shared__ double key[kMaxItem];
  __shared__ unsigned short idx[kMaxItem];
  __shared__ int grp[kMaxItem];
  const int* it = items + 4 * blockIdx.x;
  const int k = it[0];
  const long pb = it[1];
  const int len = it[2] - it[1];
  int n2 = 1;
  while (n2 < len) n2 <<= 1;
  const double ax = atoms[4 * k], ay = atoms[4 * k + 1], az = atoms[4 * k + 2];
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    double r = 1e300;
    if (i < len) {
      double rxy;
      point_polar(px[pb + i] - ax, py[pb + i] - ay, pz[pb + i] - az, rxy, r);
    }
    key[i] = r;
    idx[i] = (unsigned short)i;
  }
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const double ki = key[i], kj = key[j];
          if ((ki > kj) == up && ki != kj) {
            key[i] = kj;
            key[j] = ki;
            const unsigned short t = idx[i];
            idx[i] = idx[j];
            idx[j] = t;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < len; i += blockDim.x) grp[i] = spline_key(key[i], knots, nrad);
  __syncthreads();
  // the keys are done with: their space holds two scan buffers
  int* sa = reinterpret_cast<int*>(key);
  int* sb = sa + kMaxItem;
  for (int i = threadIdx.x; i < len; i += blockDim.x) sa[i] = (i == 0 || grp[i] != grp[i - 1]) ? i : 0;
  __syncthreads();
  const int* gs = block_scan<true>(sa, sb, len);  // first sorted position of each point's group
  int flag[kMaxItem / 1024];
#pragma unroll
  for (int r = 0; r < kMaxItem / 1024; ++r) {
    const int i = threadIdx.x + 1024 * r;
    flag[r] = 0;
    if (i < len) {
      const bool last = i == len - 1 || grp[i + 1] != grp[i];
      flag[r] = (last && ((i - gs[i] + 1) & 1)) ? 1 : 0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kMaxItem / 1024; ++r) {
    const int i = threadIdx.x + 1024 * r;
    if (i < len) sa[i] = flag[r];
  }
  __syncthreads();
  const int* pads = block_scan<false>(sa, sb, len);  // pad slots up to and including i's group end
  unsigned short* out = perm + it[3];
  unsigned short* okey = pkey + it[3];
#pragma unroll
  for (int r = 0; r < kMaxItem / 1024; ++r) {
    const int i = threadIdx.x + 1024 * r;
    if (i < len) {
      const int pos = i + pads[i] - flag[r];
      out[pos] = idx[i];
      okey[pos] = (unsigned short)grp[i];
      if (flag[r]) {
        out[pos + 1] = kPadSlot;
        okey[pos + 1] = (unsigned short)grp[i];
      }
    }
  }
  if (threadIdx.x == 0) plen[blockIdx.x] = len > 0 ? len + pads[len - 1] : 0;
}

// Potential of atom k's spherical-harmonic expansion at every point outside
// cell k: V_k(p) = sum_lm Y_lm(r_p - R_k) S_{k,lm}(|r_p - R_k|) / |r_p - R_k|
// (src/moleculargrid.cpp:349-376).  Persistent: work-group g walks the items
// [g n/G, (g+1) n/G) of the (atom, point range) list, reloading atom k's spline
// table for l in [l0, l1) into LDS when k changes.  Each thread takes Q
// consecutive slots of the item's padded order (perm: by distance to atom k,
// interval groups padded to even), runs the associated-Legendre / Chebyshev
// recurrences for each and forms the four sums sum Y y_{s-1}, sum Y D_{s-1},
// sum Y y_s, sum Y D_s per point, from which the cubic on interval s is
// evaluated once (linear in the per-lm coefficients).  The Q = 2 slots of a
// thread share their interval (a pad slot repeats its partner's point and is
// not written), taken from the sort (pkey), so each table entry is read from
// LDS once for both.  A slot's coordinates are formed from its grid indices
// (grid_xyz), not read: the sorted order scatters reads over the item's range.
// The item's potentials are collected in LDS (vbuf, by offset in the item) and
// written to vpart in point order once the item is done.
template <int LMAX, int Q>
__global__ __launch_bounds__(1024) void k_hartree_cross(CrossArgs A) {
  static_assert(Q == 2, "slots are paired by the item sort");
  extern __shared__ double lds[];
  const int n = A.nrad;
  const int lmb = A.l0 * A.l0, nlmc = A.l1 * A.l1 - lmb;
  const unsigned lbits = __builtin_amdgcn_readfirstlane(((1u << A.l1) - 1u) & ~((1u << A.l0) - 1u));
  double* kn = lds;
  double* ih = lds + n;
  double* tab = lds + 2 * n;
  double* vbuf = tab + (long)n * nlmc * 2;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    kn[i] = A.knots[i];
    ih[i] = A.invh[i];
  }
  const int it0 = (int)((long)blockIdx.x * A.nitems / gridDim.x);
  const int it1 = (int)((long)(blockIdx.x + 1) * A.nitems / gridDim.x);
  int cur = -1;
  for (int it = it0; it < it1; ++it) {
    const int k = A.items[4 * it];
    if (k != cur) {
      __syncthreads();
      const double* src = A.tab + (long)k * n * A.nlm * 2;
      for (int i = threadIdx.x; i < n * nlmc * 2; i += blockDim.x) {
        const int t = i / (nlmc * 2), rem = i % (nlmc * 2);
        tab[i] = src[(long)t * A.nlm * 2 + lmb * 2 + rem];
      }
      __syncthreads();
      cur = k;
    }
    const int pb = A.items[4 * it + 1], npt = A.items[4 * it + 2] - pb;
    const int len = A.plen[it];
    const unsigned short* pm = A.perm + A.items[4 * it + 3];
    const unsigned short* pk = A.pkey + A.items[4 * it + 3];
    const double Ax = A.atoms[4 * k], Ay = A.atoms[4 * k + 1], Az = A.atoms[4 * k + 2];
    for (int q0 = threadIdx.x * Q; q0 < len; q0 += blockDim.x * Q) {
      int eq[Q];
      double rinv[Q], ct[Q], st[Q], cp[Q], sp[Q], tq[Q];
      bool valid[Q];
      unsigned short e0 = pm[q0];
      if (e0 == kPadSlot) e0 = 0;  // not produced by the sort; kept in range regardless
      const int key = pk[q0];
      const bool clamp = key == 0 || key == n;
      const int s1 = clamp ? (key == 0 ? 0 : n - 1) : key;
      const int s0 = clamp ? s1 : key - 1;

      // knot rows s0 and s1 of the table, indexed by lm from 0 (entries below lmb are not read)
      const double* T0 = tab + (s0 * nlmc - lmb) * 2;
      const double* T1 = tab + (s1 * nlmc - lmb) * 2;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        // a pad slot or a missing tail slot repeats the thread's first point
        const unsigned short e = (q0 + q < len) ? pm[q0 + q] : kPadSlot;
        valid[q] = e != kPadSlot;
        eq[q] = valid[q] ? e : e0;
        // p = (a nrad + i) nang + j
        const int p = pb + eq[q];
        const int ai = (int)(((double)p + 0.5) * A.inv_nang);
        const int j = p - ai * A.nang;
        const int a = (int)(((double)ai + 0.5) * A.inv_nrad);
        const int i = ai - a * n;
        double x, y, z;
        grid_xyz(A.atoms, A.rad, A.leb, a, i, j, x, y, z);
        const double dx = x - Ax, dy = y - Ay, dz = z - Az;
        double rxy, r;
        point_polar(dx, dy, dz, rxy, r);
        rinv[q] = 1.0 / r;
        ct[q] = dz * rinv[q];
        st[q] = rxy * rinv[q];
        cp[q] = 1.0;
        sp[q] = 0.0;
        if (rxy > 0.0) {
          const double ir = 1.0 / rxy;
          cp[q] = dx * ir;
          sp[q] = dy * ir;
        }
        tq[q] = clamp ? 0.0 : r - kn[s0];
      }
      double sy0[Q], sd0[Q], sy1[Q], sd1[Q], pmm[Q], cm[Q], sm[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        sy0[q] = sd0[q] = sy1[q] = sd1[q] = 0.0;
        pmm[q] = 1.0;
        cm[q] = 1.0;
        sm[q] = 0.0;
      }
#pragma unroll
      for (int m = 0; m <= LMAX; ++m) {
        if (m > 0) {
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            pmm[q] *= -(double)(2 * m - 1) * st[q];
            const double c2 = cm[q] * cp[q] - sm[q] * sp[q];
            sm[q] = sm[q] * cp[q] + cm[q] * sp[q];
            cm[q] = c2;
          }
        }
        double p1[Q], p2[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) p1[q] = p2[q] = 0.0;
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
          // scaled recurrence Q_l = P_l / c_lm (legendre_scale; c_lm is folded into
          // the table): Q_l = ct Q_{l-1} - beta Q_{l-2}, beta = (l+m-1)/(l-m) c_{l-2,m} / c_lm,
          // a constant after unrolling (a global read of it waited on its full latency)
          const double beta = (l >= m + 2) ? (double)(l + m - 1) / (double)(l - m) * legendre_scale(l - 2, m) /
                                                 legendre_scale(l, m)
                                           : 0.0;
          double yp[Q], yn[Q];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            double pl;
            if (l == m)
              pl = pmm[q];
            else if (l == m + 1)
              pl = ct[q] * pmm[q];
            else
              pl = ct[q] * p1[q] - beta * p2[q];
            p2[q] = p1[q];
            p1[q] = pl;
            const double base = pl;  // N_lm c_lm is folded into the table
            yp[q] = (m == 0) ? base : base * cm[q];
            yn[q] = base * sm[q];
          }
          // (a run-time l range, one bit per l of a scalar, tested where it is used:
          // the opaque copy keeps the tests from being hoisted and carried as lane
          // masks, and the branches keep the table reads from being hoisted en masse,
          // which would spill at 1024 threads; the rows are addressed from lm = 0, so
          // every read's offset is a compile-time immediate)
          unsigned lb = lbits;
          asm volatile("" : "+s"(lb));
          if ((__builtin_amdgcn_readfirstlane(lb) >> l) & 1u) {
            const int ip = (l * l + l + m) * 2;
            const int in = (l * l + l - m) * 2;
            // 16-byte {y, D} pairs (the table and its rows start on 16-byte boundaries)
            const double2 t0p = *reinterpret_cast<const double2*>(T0 + ip);
            const double2 t1p = *reinterpret_cast<const double2*>(T1 + ip);
            const double a0 = t0p.x, a1 = t0p.y, b0 = t1p.x, b1 = t1p.y;
            double c0 = 0.0, c1 = 0.0, e0 = 0.0, e1 = 0.0;
            if (m > 0) {
              const double2 t0n = *reinterpret_cast<const double2*>(T0 + in);
              const double2 t1n = *reinterpret_cast<const double2*>(T1 + in);
              c0 = t0n.x;
              c1 = t0n.y;
              e0 = t1n.x;
              e1 = t1n.y;
            }
            // all of the step's table reads issued before its first use
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
              sy0[q] += yp[q] * a0;
              sd0[q] += yp[q] * a1;
              sy1[q] += yp[q] * b0;
              sd1[q] += yp[q] * b1;
            }
            if (m > 0) {
#pragma unroll
              for (int q = 0; q < Q; ++q) {
                sy0[q] += yn[q] * c0;
                sd0[q] += yn[q] * c1;
                sy1[q] += yn[q] * e0;
                sd1[q] += yn[q] * e1;
              }
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (!valid[q]) continue;
        double val;
        if (clamp) {
          val = sy0[q];
        } else {
          const double hi_ = ih[s1];
          const double t = tq[q];
          const double dy = (sy1[q] - sy0[q]) * hi_;
          const double c = hi_ * (3.0 * dy - 2.0 * sd0[q] - sd1[q]);
          const double d = hi_ * hi_ * (-2.0 * dy + sd0[q] + sd1[q]);
          val = sy0[q] + sd0[q] * t + c * t * t + d * t * t * t;
        }
        vbuf[eq[q]] = val * rinv[q];
      }
    }
    // every point of the item has its slot: write the potentials in point order
    __syncthreads();
    double* out = A.vpart + A.vbase[k] + pb;
    for (int e = threadIdx.x; e < npt; e += blockDim.x) out[e] = A.accumulate ? out[e] + vbuf[e] : vbuf[e];
    __syncthreads();
  }
}

// V[p] = sum_k (k == own ? vcell[p] : vpart_k[p]) over the atoms k of the
// point's molecule, k ascending (src/moleculargrid.cpp:347-378); uJ[p] = w[p] V[p] / 2
// (src/atomicgrid.cpp:475-485); ej_part[block] = sum w V rho over the block (fixed
// tree order).  atom_mol/mol_atom0 (NULL for one molecule) give the atom range.
__global__ __launch_bounds__(256) void k_vreduce(const double* __restrict__ vpart,
                                                 const long* __restrict__ vbase,
                                                 const double* __restrict__ vcell,
                                                 const double* __restrict__ w,
                                                 const double* __restrict__ rho, long npts,
                                                 long npts_pad, long cell, int natoms,
                                                 const int* __restrict__ atom_mol,
                                                 const int* __restrict__ mol_atom0, double* V,
                                                 double* uJ, double* ej_part) {
  __shared__ double red[256];
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (p < npts) {
    const int own = (int)(p / cell);
    int k0 = 0, k1 = natoms;
    if (atom_mol) {
      const int m = atom_mol[own];
      k0 = mol_atom0[m];
      k1 = mol_atom0[m + 1];
    }
    // sixteen partials loaded together per round (one memory latency), added in k order
    double v = 0.0;
    for (int kb = k0; kb < k1; kb += 16) {
      double x[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int k = kb + u;
        x[u] = (k < k1) ? ((k == own) ? vcell[p] : vpart[vbase[k] + p]) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (kb + u < k1) v += x[u];
    }
    V[p] = v;
    uJ[p] = (w[p] * v) * 0.5;
    e = w[p] * v * rho[p];
  } else if (p < npts_pad) {
    V[p] = 0.0;
    uJ[p] = 0.0;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) ej_part[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Contractions  C_o = Phi^T diag(u_o) Phi  (src/dft.cpp:424-432, src/atomicgrid.cpp:471-488)
// ---------------------------------------------------------------------------
// Contraction for larger bases with 4 x NJ tile blocks: block {I0, nI, J0, nJ}
// covers the tiles (I0 + r, J0 + t), r < nI, t < nJ <= NJ, I <= J, of the NOUT
// outputs, so a wave loads 4 + NJ amplitude tiles per step (4 points) for
// 4 NJ NOUT MFMAs: NJ = 2 with two outputs (J and XC), NJ = 4 with one.  The
// block's 64-point blocks (its list: all, or with screening those where its
// tiles are significant) are split evenly over nsplit work-groups (split-K);
// the 4 waves interleave the 16 points of each block.  The 4 waves are added in
// LDS in wave order and written as 4 reduction groups (I0 + r, J0, nJ) in the
// k_contract_reduce layout.
// Work-group order: the hardware deals work-groups round-robin to the kXcds
// XCDs, each with its own L2.  Work-group L (1-D grid, padded to a multiple of
// kXcds) takes virtual id v = (L % kXcds) per + L / kXcds and order[v] = {tile
// block, split}: consecutive v are the tile blocks of one split, so an XCD runs
// whole splits, and within an XCD's range the blocks with the most MFMAs come
// first, so that the short diagonal blocks fill the last round.
constexpr int kXcds = 8;

// The FP64 2-D contraction, its operands staged through LDS: each wave copies
// the amplitude rows and weights of step st + D - 1 into slot (st + D - 1) % D
// of its own ring (global_load_lds, 16 bytes per lane, so the copies hold no
// registers) while the MFMAs of step st read slot st % D; a step's copy is
// NI vector-memory instructions, so waiting for vmcnt <= NI (D - 1) leaves the
// later steps in flight.  Slot layout (doubles): the 4 points' rows of tiles
// I0..I0+3 (point-major, 64 each), of tiles J0..J0+NJ-1 (16 NJ each), then
// u0 and u1 of the 4 points.  The epilogue's LDS reduction reuses the ring.
template <int NOUT, int NJ, int D, bool DENSE>
__global__ __launch_bounds__(256, 3) void k_contract2d_lds(const double* __restrict__ phi, int ldphi,
                                                          const double* __restrict__ u0,
                                                          const double* __restrict__ u1,
                                                          const int* __restrict__ cblocks,
                                                          const int* __restrict__ blist,
                                                          const int* __restrict__ loff,
                                                          const int2* __restrict__ order, int ngroups, int nsplit,
                                                          double* part) {
  constexpr int SA = 256, SB = 64 * NJ, SW = 8, SLOT = SA + SB + SW;  // doubles per slot
  constexpr int NI = 2 + NJ / 2 + 1;                                    // copy instructions per step
  static_assert(4 * SLOT * D >= 4 * NJ * 256, "the epilogue reduction fits in the ring");
  __shared__ __attribute__((aligned(16))) double ring[4 * SLOT * D];
  double* red = ring;
  const int ncb = ngroups / 4, per = gridDim.x / kXcds;
  const int v = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (v >= ncb * nsplit) return;
  const int cbi = order[v].x, split = order[v].y;
  const int I0 = cblocks[4 * cbi], nI = cblocks[4 * cbi + 1], J0 = cblocks[4 * cbi + 2], nJ = cblocks[4 * cbi + 3];
  // the wave index as a scalar, so that the ring addresses stay in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long l0 = loff[cbi], cnt = loff[cbi + 1] - l0;
  const int s0 = (int)(cnt * split / nsplit);
  const int* bl = blist + l0 + s0;
  const int nst = 4 * (int)(cnt * (split + 1) / nsplit - s0);
  double* my = ring + wave * SLOT * D;
  // step t covers this wave's points p .. p + 3, p = 4 wave + rows(t) with
  // rows(t) = 64 s0 + 16 t dense, 64 blist[t / 4] + 16 (t % 4) screened
  auto rows = [&](int t) { return DENSE ? 16L * t : 64L * bl[t >> 2] + 16 * (t & 3); };
  // copy lane roles: A pair k = 64 i + lane -> point k / 32, columns 2 (k % 32), +1;
  // weights: lanes 0, 1 u0 of points 0-1, 2-3, lanes 2, 3 u1 (unused with one output).
  // Each lane's sources at rows(t) = 0, so that a copy adds one row offset.
  const int ag = lane >> 5, acol = 16 * I0 + 2 * (lane & 31);
  const int bg = NJ == 2 ? lane >> 4 : lane >> 5, bcol = 16 * J0 + 2 * (NJ == 2 ? (lane & 15) : (lane & 31));
  const long p0 = (DENSE ? 64L * s0 : 0L) + 4 * wave;
  const double* pa = phi + (p0 + ag) * ldphi + acol;
  const double* pb = phi + (p0 + bg) * ldphi + bcol;
  const double* pw = (lane < 2 ? u0 : u1) + p0 + 2 * (lane & 1);
  auto copy = [&](int t, int slot) {
    double* dst = my + slot * SLOT;
    typedef __attribute__((address_space(3))) void* lptr;
    const long r = rows(t), o = r * ldphi;
#pragma unroll
    for (int i = 0; i < 2; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(pa + o + 2L * i * ldphi), (lptr)(dst + i * 128), 16, 0, 0);
#pragma unroll
    for (int i = 0; i < NJ / 2; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(pb + o + (NJ == 2 ? 0L : 2L * i * ldphi)),
                                       (lptr)(dst + SA + i * 128), 16, 0, 0);
    if (lane < 4) __builtin_amdgcn_global_load_lds((const void*)(pw + r), (lptr)(dst + SA + SB), 16, 0, 0);
  };
  dbl4 acc[NOUT][4][NJ];
#pragma unroll
  for (int o = 0; o < NOUT; ++o)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int t = 0; t < NJ; ++t) acc[o][r][t] = dbl4{0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < D - 1; ++k)
    if (k < nst) copy(k, k);
  for (int st = 0; st < nst; ++st) {
    if (st + D - 1 < nst) {
      copy(st + D - 1, (st + D - 1) % D);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(NI * (D - 1)));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    const double* cur = my + (st % D) * SLOT;
    double ca[4], cb[NJ];
#pragma unroll
    for (int r = 0; r < 4; ++r) ca[r] = cur[g * 64 + 16 * r + c];
#pragma unroll
    for (int t = 0; t < NJ; ++t) cb[t] = cur[SA + g * 16 * NJ + 16 * t + c];
    const double cw0 = cur[SA + SB + g], cw1 = NOUT == 2 ? cur[SA + SB + 4 + g] : 0.0;
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
      const double b0 = cw0 * cb[t], b1 = cw1 * cb[t];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < nI && t < nJ && I0 + r <= J0 + t) {
          acc[0][r][t] = mfma16x16x4(ca[r], b0, acc[0][r][t]);
          if (NOUT == 2) acc[NOUT - 1][r][t] = mfma16x16x4(ca[r], b1, acc[NOUT - 1][r][t]);
        }
    }
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    for (int wv = 0; wv < 4; ++wv) {
      __syncthreads();
      if (wave == wv) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int t = 0; t < NJ; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              // register q of lane (c, g) = row 4q + g
              double* x = red + (r * NJ + t) * 256 + q * 64 + lane;
              *x = (wv == 0) ? acc[o][r][t][q] : *x + acc[o][r][t][q];
            }
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 4 * NJ * 256; e += 256) {
      const int r = e / (NJ * 256), t = (e >> 8) % NJ, rem = e & 255;
      part[(((long)split * ngroups + 4 * cbi + r) * NOUT + o) * 1024 + t * 256 + rem] = red[e];
    }
  }
}

// Reduced-precision 2-D contraction: FP32 amplitudes on the FP32 16x16x4 MFMA
// (register v of lane (c, g) holds row 4g + v), staged through the same kind of
// per-wave LDS ring (slot: the 4 points' FP32 rows of tiles I0..I0+3, of tiles
// J0..J0+NJ-1, then u0 and u1 of the 4 points in FP64; three copy instructions per
// step).  Every 16 steps (64 points) a wave adds its FP32 sums into FP64 sums it
// keeps in registers, so the waves never wait for each other before the epilogue;
// the FP64 sums (64 doubles per lane) leave room for two waves per SIMD, and the
// ring, with two work-groups per CU, is D steps deep.  The weights are rounded to
// FP32 and multiply the B operand in FP32.
template <int NOUT, int NJ, int D, bool DENSE>
__global__ __launch_bounds__(256, 2) void k_contract2d_lds_f32(const float* __restrict__ phi, int ldphi,
                                                              const double* __restrict__ u0,
                                                              const double* __restrict__ u1,
                                                              const int* __restrict__ cblocks,
                                                              const int* __restrict__ blist,
                                                              const int* __restrict__ loff,
                                                              const int2* __restrict__ order, int ngroups,
                                                              int nsplit, double* part) {
  constexpr int SA = 256, SB = 64 * NJ;            // floats per slot: A and B rows
  constexpr int SLOT = (SA + SB) * 4 + 8 * 8;      // bytes per slot (+ 8 FP64 weights)
  constexpr int NI = 3;                            // copy instructions per step
  static_assert(4 * SLOT * D >= 4 * NJ * 256 * 8, "the epilogue reduction fits in the ring");
  __shared__ __attribute__((aligned(16))) char ring[4 * SLOT * D];
  double* red = reinterpret_cast<double*>(ring);
  const int ncb = ngroups / 4, per = gridDim.x / kXcds;
  const int v = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (v >= ncb * nsplit) return;
  const int cbi = order[v].x, split = order[v].y;
  const int I0 = cblocks[4 * cbi], nI = cblocks[4 * cbi + 1], J0 = cblocks[4 * cbi + 2], nJ = cblocks[4 * cbi + 3];
  // the wave index as a scalar, so that the ring addresses stay in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long l0 = loff[cbi], cnt = loff[cbi + 1] - l0;
  const int s0 = (int)(cnt * split / nsplit);
  const int* bl = blist + l0 + s0;
  const int nst = 4 * (int)(cnt * (split + 1) / nsplit - s0);
  char* my = ring + wave * SLOT * D;
  // step t covers this wave's points p .. p + 3, p = 4 wave + rows(t) with
  // rows(t) = 64 s0 + 16 t dense, 64 blist[t / 4] + 16 (t % 4) screened
  auto rows = [&](int t) { return DENSE ? 16L * t : 64L * bl[t >> 2] + 16 * (t & 3); };
  // copy lane roles: A: point lane / 16, 4 columns from 16 I0 + 4 (lane % 16);
  // B (the first 16 NJ lanes): point lane / (4 NJ), 4 columns from 16 J0 + 4 (lane % 4 NJ);
  // weights: lanes 0, 1 u0 of points 0-1, 2-3, lanes 2, 3 u1 (unused with one output).
  // Each lane's sources at rows(t) = 0, so that a copy adds one row offset.
  const int ag = lane >> 4, acol = 16 * I0 + 4 * (lane & 15);
  const int bg = lane / (4 * NJ), bcol = 16 * J0 + 4 * (lane % (4 * NJ));
  const long p0 = (DENSE ? 64L * s0 : 0L) + 4 * wave;
  const float* pa = phi + (p0 + ag) * ldphi + acol;
  const float* pb = phi + (p0 + bg) * ldphi + bcol;
  const double* pw = (lane < 2 ? u0 : u1) + p0 + 2 * (lane & 1);
  auto copy = [&](int t, int slot) {
    char* dst = my + slot * SLOT;
    typedef __attribute__((address_space(3))) void* lptr;
    const long r = rows(t), o = r * ldphi;
    __builtin_amdgcn_global_load_lds((const void*)(pa + o), (lptr)dst, 16, 0, 0);
    if (lane < 16 * NJ) __builtin_amdgcn_global_load_lds((const void*)(pb + o), (lptr)(dst + SA * 4), 16, 0, 0);
    if (lane < 4) __builtin_amdgcn_global_load_lds((const void*)(pw + r), (lptr)(dst + (SA + SB) * 4), 16, 0, 0);
  };
  flt4 acc[NOUT][4][NJ];
  double sum[NOUT][4][NJ][4];
#pragma unroll
  for (int o = 0; o < NOUT; ++o)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int t = 0; t < NJ; ++t) {
        acc[o][r][t] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q) sum[o][r][t][q] = 0.0;
      }
#pragma unroll
  for (int k = 0; k < D - 1; ++k)
    if (k < nst) copy(k, k);
  for (int st = 0; st < nst; ++st) {
    if (st + D - 1 < nst) {
      copy(st + D - 1, (st + D - 1) % D);
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(NI * (D - 1)));
    } else {
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    }
    const char* cur = my + (st % D) * SLOT;
    const float* cf = reinterpret_cast<const float*>(cur);
    const double* cwp = reinterpret_cast<const double*>(cur + (SA + SB) * 4);
    float ca[4], cb[NJ];
#pragma unroll
    for (int r = 0; r < 4; ++r) ca[r] = cf[g * 64 + 16 * r + c];
#pragma unroll
    for (int t = 0; t < NJ; ++t) cb[t] = cf[SA + g * 16 * NJ + 16 * t + c];
    const float cw0 = (float)cwp[g], cw1 = NOUT == 2 ? (float)cwp[4 + g] : 0.f;
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
      const float b0 = cw0 * cb[t], b1 = cw1 * cb[t];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < nI && t < nJ && I0 + r <= J0 + t) {
          acc[0][r][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[r], b0, acc[0][r][t], 0, 0, 0);
          if (NOUT == 2)
            acc[NOUT - 1][r][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[r], b1, acc[NOUT - 1][r][t], 0, 0, 0);
        }
    }
    if ((st & 15) == 15 || st + 1 == nst) {
#pragma unroll
      for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int t = 0; t < NJ; ++t) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sum[o][r][t][q] += (double)acc[o][r][t][q];
            acc[o][r][t] = flt4{0.f, 0.f, 0.f, 0.f};
          }
    }
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    for (int wv = 0; wv < 4; ++wv) {
      __syncthreads();
      if (wave == wv) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int t = 0; t < NJ; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              // register q of lane (c, g) = row 4g + q
              double* x = red + (r * NJ + t) * 256 + g * 64 + 16 * q + c;
              *x = (wv == 0) ? sum[o][r][t][q] : *x + sum[o][r][t][q];
            }
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 4 * NJ * 256; e += 256) {
      const int r = e / (NJ * 256), t = (e >> 8) % NJ, rem = e & 255;
      part[(((long)split * ngroups + 4 * cbi + r) * NOUT + o) * 1024 + t * 256 + rem] = red[e];
    }
  }
}

// Contraction for small bases (NT = ceil(nbf/16) <= 5 tiles, dense): a wave
// accumulates all NT(NT+1)/2 upper tiles of one output, so each amplitude is
// read from HBM once per build and serves as both operands.  Work-group s
// takes the 64-point blocks [nblk s/G, nblk (s+1)/G); its 4 waves are split
// over the outputs (4/NOUT per output) and interleave the 16 four-point steps
// of each block (next step's operands prefetched).  The waves of an output
// are summed in LDS in wave order and written in the k_contract_reduce layout
// (group of row I and columns J0 = I + 4k.., tile t = (J - I) % 4, register v
// holding row 4v + g).  EDGE > 0: the last tile holds only EDGE (<= 4)
// functions; its NT tile pairs (I, NT-1) are summed on the VALU instead of by
// MFMAs (15/16 of which would multiply padding): lane (c, g) accumulates
// phi[p][16I + c] u_p phi[p][16(NT-1) + j] with the second factor taken from
// lane (j, g), and the four point lanes g are added before the LDS stage.
template <int NT, int NOUT, int EDGE>
__global__ __launch_bounds__(256, 2) void k_contract_small(const double* __restrict__ phi, int ldphi,
                                                        const double* __restrict__ u0,
                                                        const double* __restrict__ u1, long nblk,
                                                        int ngroups, double* part) {
  constexpr int NP = NT * (NT + 1) / 2;
  constexpr int NW = 4 / NOUT;  // waves per output
  __shared__ double red[NP * 256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int o = wave / NW, sw = wave % NW;
  const double* u = (o == 0) ? u0 : u1;
  const long b0 = nblk * blockIdx.x / gridDim.x, b1 = nblk * (blockIdx.x + 1) / gridDim.x;
  constexpr int SPB = 16 / NW;  // this wave's steps per block
  const int nst = SPB * (int)(b1 - b0);
  auto point = [&](int t) { return 64L * (b0 + t / SPB) + 4 * (sw + NW * (t % SPB)) + g; };
  dbl4 acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = dbl4{0.0, 0.0, 0.0, 0.0};
  double e[NT][EDGE > 0 ? EDGE : 1];
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int j = 0; j < (EDGE > 0 ? EDGE : 1); ++j) e[I][j] = 0.0;
  double f[NT], w = 0.0;
  if (nst > 0) {
    const long p = point(0);
#pragma unroll
    for (int T = 0; T < NT; ++T) f[T] = phi[p * ldphi + 16 * T + c];
    w = u[p];
  }
  for (int st = 0; st < nst; ++st) {
    double cf[NT];
#pragma unroll
    for (int T = 0; T < NT; ++T) cf[T] = f[T];
    const double cw = w;
    if (st + 1 < nst) {
      const long p = point(st + 1);
#pragma unroll
      for (int T = 0; T < NT; ++T) f[T] = phi[p * ldphi + 16 * T + c];
      w = u[p];
    }
    double bw[NT];
#pragma unroll
    for (int T = 0; T < NT; ++T) bw[T] = cw * cf[T];
    int q = 0;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = I; J < NT; ++J, ++q)
        if (EDGE == 0 || J < NT - 1) acc[q] = mfma16x16x4(cf[I], bw[J], acc[q]);
    if constexpr (EDGE > 0) {
#pragma unroll
      for (int j = 0; j < EDGE; ++j) {
        const double wj = __shfl(bw[NT - 1], (lane & 48) + j);
#pragma unroll
        for (int I = 0; I < NT; ++I) e[I][j] += cf[I] * wj;
      }
    }
  }
  if constexpr (EDGE > 0) {
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int j = 0; j < EDGE; ++j) {
        e[I][j] += __shfl_xor(e[I][j], 16);
        e[I][j] += __shfl_xor(e[I][j], 32);
      }
  }
  for (int oo = 0; oo < NOUT; ++oo) {
    for (int s2 = 0; s2 < NW; ++s2) {
      __syncthreads();
      if (o == oo && sw == s2) {
#pragma unroll
        for (int q = 0; q < NP; ++q)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double* r = red + q * 256 + v * 64 + lane;
            *r = (s2 == 0) ? acc[q][v] : *r + acc[q][v];
          }
        if constexpr (EDGE > 0) {
          // element (row c, column j) of tile (I, NT-1): register c/4 of lane 16 (c%4) + j
          if (g == 0) {
            int qi = 0;
#pragma unroll
            for (int I = 0; I < NT; ++I) {
              qi += (I == 0) ? NT - 1 : NT - I;  // index of the pair (I, NT-1)
#pragma unroll
              for (int j = 0; j < EDGE; ++j) red[qi * 256 + (c >> 2) * 64 + (c & 3) * 16 + j] += e[I][j];
            }
          }
        }
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NP * 256; e += 256) {
      const int q = e >> 8, rem = e & 255;
      int I = 0, qq = q;
      while (qq >= NT - I) {
        qq -= NT - I;
        ++I;
      }
      const int J = I + qq;
      int grp = 0;
      for (int I2 = 0; I2 < I; ++I2) grp += (NT - I2 + 3) / 4;
      grp += (J - I) / 4;
      part[(((long)blockIdx.x * ngroups + grp) * NOUT + oo) * 1024 + ((J - I) & 3) * 256 + rem] = red[e];
    }
  }
}

// k_contract_small in reduced precision: FP32 amplitudes and u Phi on the FP32
// 16x16x4 MFMA (register v holds row 4g + v), whose sums are added every 16
// steps (64 points per wave), wave by wave in a fixed order, into the
// work-group's FP64 accumulators in LDS (one NP-tile set per output); a sparse
// last tile (EDGE) is summed in FP64 on the VALU as in k_contract_small.
template <int NT, int NOUT, int EDGE>
__global__ __launch_bounds__(256, 2) void k_contract_small_f32(const float* __restrict__ phi, int ldphi,
                                                            const double* __restrict__ u0,
                                                            const double* __restrict__ u1, long nblk,
                                                            int ngroups, double* part) {
  constexpr int NP = NT * (NT + 1) / 2;
  constexpr int NW = 4 / NOUT;  // waves per output
  __shared__ double red[NOUT * NP * 256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int o = wave / NW, sw = wave % NW;
  const double* u = (o == 0) ? u0 : u1;
  const long b0 = nblk * blockIdx.x / gridDim.x, b1 = nblk * (blockIdx.x + 1) / gridDim.x;
  constexpr int SPB = 16 / NW;  // this wave's steps per block
  const int nst = SPB * (int)(b1 - b0);
  auto point = [&](int t) { return 64L * (b0 + t / SPB) + 4 * (sw + NW * (t % SPB)) + g; };
  for (int i = threadIdx.x; i < NOUT * NP * 256; i += 256) red[i] = 0.0;
  __syncthreads();
  flt4 acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
  double e[NT][EDGE > 0 ? EDGE : 1];
#pragma unroll
  for (int I = 0; I < NT; ++I)
#pragma unroll
    for (int j = 0; j < (EDGE > 0 ? EDGE : 1); ++j) e[I][j] = 0.0;
  double* my = red + o * NP * 256;
  auto flush = [&]() {  // same steps in every wave
    for (int s2 = 0; s2 < NW; ++s2) {
      __syncthreads();
      if (sw == s2) {
#pragma unroll
        for (int q = 0; q < NP; ++q)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            my[q * 256 + g * 64 + v * 16 + c] += (double)acc[q][v];
            acc[q][v] = 0.0f;
          }
      }
    }
  };
  float f[NT];
  double w = 0.0;
  if (nst > 0) {
    const long p = point(0);
#pragma unroll
    for (int T = 0; T < NT; ++T) f[T] = phi[p * ldphi + 16 * T + c];
    w = u[p];
  }
  for (int st = 0; st < nst; ++st) {
    float cf[NT];
#pragma unroll
    for (int T = 0; T < NT; ++T) cf[T] = f[T];
    const double cw = w;
    if (st + 1 < nst) {
      const long p = point(st + 1);
#pragma unroll
      for (int T = 0; T < NT; ++T) f[T] = phi[p * ldphi + 16 * T + c];
      w = u[p];
    }
    float bw[NT];
#pragma unroll
    for (int T = 0; T < NT; ++T) bw[T] = (float)cw * cf[T];
    int q = 0;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int J = I; J < NT; ++J, ++q)
        if (EDGE == 0 || J < NT - 1) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(cf[I], bw[J], acc[q], 0, 0, 0);
    if constexpr (EDGE > 0) {
#pragma unroll
      for (int j = 0; j < EDGE; ++j) {
        const double wj = __shfl((double)bw[NT - 1], (lane & 48) + j);
#pragma unroll
        for (int I = 0; I < NT; ++I) e[I][j] += (double)cf[I] * wj;
      }
    }
    if ((st & 15) == 15) flush();
  }
  if (nst & 15) flush();
  if constexpr (EDGE > 0) {
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
      for (int j = 0; j < EDGE; ++j) {
        e[I][j] += __shfl_xor(e[I][j], 16);
        e[I][j] += __shfl_xor(e[I][j], 32);
      }
    for (int s2 = 0; s2 < NW; ++s2) {
      __syncthreads();
      if (sw == s2 && g == 0) {
        int qi = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
          qi += (I == 0) ? NT - 1 : NT - I;  // index of the pair (I, NT-1)
#pragma unroll
          for (int j = 0; j < EDGE; ++j) my[qi * 256 + (c >> 2) * 64 + (c & 3) * 16 + j] += e[I][j];
        }
      }
    }
  }
  __syncthreads();
  for (int oo = 0; oo < NOUT; ++oo)
    for (int ei = threadIdx.x; ei < NP * 256; ei += 256) {
      const int q = ei >> 8, rem = ei & 255;
      int I = 0, qq = q;
      while (qq >= NT - I) {
        qq -= NT - I;
        ++I;
      }
      const int J = I + qq;
      int grp = 0;
      for (int I2 = 0; I2 < I; ++I2) grp += (NT - I2 + 3) / 4;
      grp += (J - I) / 4;
      part[(((long)blockIdx.x * ngroups + grp) * NOUT + oo) * 1024 + ((J - I) & 3) * 256 + rem] =
          red[oo * NP * 256 + ei];
    }
}

// ---------------------------------------------------------------------------
// Adjoint of E_xc (no counterpart in the reference)
// ---------------------------------------------------------------------------
// Full P in fragment order (no symmetric halving): frag[b][s][lane] =
// P(16b + lane%16, kk(s, lane/16)) for all 4*nt reduction steps.
__global__ void k_pfrag_full(const double* __restrict__ P, int nbf, int nt, double* frag) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = 256L * nt;
  if (e >= per * nt) return;
  const int b = (int)(e / per);
  const int rem = (int)(e % per);
  const int s = rem / 64, l = rem % 64;
  const int j = 16 * b + (l & 15);
  const int k = 16 * (s / 4) + 4 * (l >> 4) + (s & 3);
  frag[e] = (j < nbf && k < nbf) ? P[j + (long)k * nbf] : 0.0;
}

// T = Phi P (T[p][j] = sum_k phi_pk P_kj) with the FP64 MFMA, same operand
// arrangement as k_density; T is written [npts_pad][ldphi].  W waves (16
// points each) share each column tile's fragments in LDS, double-buffered: the
// next tile's are copied by global_load_lds while the current tile's MFMAs run
// in four accumulation chains; one wait and barrier per tile.
template <int NT, int W>
__global__ __launch_bounds__(64 * W) void k_tmat(const double* __restrict__ phi, int ldphi,
                                                 const double* __restrict__ frag, double* tmat) {
  __shared__ double pl[2][4 * NT * 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long p0 = (long)blockIdx.x * 16 * W + wave * 16;
  const double* row = phi + (p0 + c) * ldphi;
  double f[4 * NT];
#pragma unroll
  for (int s = 0; s < 4 * NT; ++s) f[s] = row[16 * (s >> 2) + 4 * g + (s & 3)];
  // column tile b's fragments: 256 NT doubles from 256 NT b, chunks of 128
  // doubles, chunk q copied by wave q % W, 16 bytes per lane
  auto fetch = [&](int b, double* dst) {
    const double* src = frag + 256L * NT * b;
    for (int q = wave; q < 2 * NT; q += W)
      __builtin_amdgcn_global_load_lds((const void*)(src + q * 128 + lane * 2),
                                       (__attribute__((address_space(3))) void*)(dst + q * 128), 16, 0, 0);
  };
  fetch(0, pl[0]);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int b = 0; b < NT; ++b) {
    if (b + 1 < NT) fetch(b + 1, pl[(b + 1) & 1]);
    const double* cur = pl[b & 1];
    dbl4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0, d2 = d0, d3 = d0;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      d0 = mfma16x16x4(cur[(4 * k) * 64 + lane], f[4 * k], d0);
      d1 = mfma16x16x4(cur[(4 * k + 1) * 64 + lane], f[4 * k + 1], d1);
      d2 = mfma16x16x4(cur[(4 * k + 2) * 64 + lane], f[4 * k + 2], d2);
      d3 = mfma16x16x4(cur[(4 * k + 3) * 64 + lane], f[4 * k + 3], d3);
    }
    double* out = tmat + (p0 + c) * ldphi + 16 * b + g;
    out[0] = (d0[0] + d1[0]) + (d2[0] + d3[0]);
    out[4] = (d0[1] + d1[1]) + (d2[1] + d3[1]);
    out[8] = (d0[2] + d1[2]) + (d2[2] + d3[2]);
    out[12] = (d0[3] + d1[3]) + (d2[3] + d3[3]);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies into the other buffer have landed
    __syncthreads();
  }
}

// k_tmat with the full-P fragments streamed through a ring: unit (b, k) = the
// 4 steps of reduction tile k for column tile b (2 KB), stored b-major, which
// is the order consumed; chunks of kDensityChunk units copied two ahead into
// three LDS slots (48 KB: two work-groups of W = 4 waves per CU), one wait and
// barrier per chunk.  Same chains and sums as k_tmat, so the same T.
// With tmask (bit b of tmask[block]: tile b matters at the block's 64 points,
// k_tile_mask), the block forms and writes only the column tiles in its mask
// and sums only over the reduction tiles in it (the others' amplitudes are
// below e^-100 there); the ring's copies, waits and barriers run unchanged.
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_tmat_ring(
    const double* __restrict__ phi, int ldphi, const double* __restrict__ frag, double* tmat,
    const unsigned* __restrict__ tmask) {
  constexpr int C = kDensityChunk, U = NT * NT, NCH = (U + C - 1) / C;
  const unsigned tm = tmask ? tmask[blockIdx.x] : ~0u;
  __shared__ double ring[3][C * 256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long p0 = (long)blockIdx.x * 64 + wave * 16;
  const double* row = phi + (p0 + c) * ldphi;
  double f[4 * NT];
#pragma unroll
  for (int s = 0; s < 4 * NT; ++s) f[s] = row[16 * (s >> 2) + 4 * g + (s & 3)];
  auto copy = [&](int j) { stream_copy<C, U>(frag, ring[j % 3], j, wave, lane); };
  copy(0);
  if (NCH > 1) copy(1);
  // the ring's wait, barrier and next copy at the first unit of every chunk
  auto ring_step = [&](int u) {
    if (u % C == 0) {
      const int j = u / C;
      stream_wait<C, U>(j);
      bare_barrier();
      if (j + 2 < NCH) copy(j + 2);
    }
  };
#pragma unroll
  for (int b = 0; b < NT; ++b) {
    if ((tm >> b) & 1u) {
      dbl4 d0 = {0.0, 0.0, 0.0, 0.0}, d1 = d0, d2 = d0, d3 = d0;
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const int u = b * NT + k;
        ring_step(u);
        if ((tm >> k) & 1u) {
          const double* cur = ring[(u / C) % 3] + (u - stream_start<C, U>(u / C)) * 256;
          d0 = mfma16x16x4(cur[lane], f[4 * k], d0);
          d1 = mfma16x16x4(cur[64 + lane], f[4 * k + 1], d1);
          d2 = mfma16x16x4(cur[128 + lane], f[4 * k + 2], d2);
          d3 = mfma16x16x4(cur[192 + lane], f[4 * k + 3], d3);
        }
      }
      double* out = tmat + (p0 + c) * ldphi + 16 * b + g;
      out[0] = (d0[0] + d1[0]) + (d2[0] + d3[0]);
      out[4] = (d0[1] + d1[1]) + (d2[1] + d3[1]);
      out[8] = (d0[2] + d1[2]) + (d2[2] + d3[2]);
      out[12] = (d0[3] + d1[3]) + (d2[3] + d3[3]);
    } else {
#pragma unroll
      for (int k = 0; k < NT; ++k) ring_step(b * NT + k);
    }
  }
}

// Value and gradient of component `comp` of a shell with radial sums
// r0 = sum cN0 e, d0 = sum cN0 (-2 alpha) e (and r1, d1 for the mixed d
// components): phi = R x^l y^m z^n, d/dx phi = l x^(l-1) y^m z^n R + x phi' with R' = d.
__device__ __forceinline__ void shell_grad(int L, int comp, double r0, double d0, double r1, double d1,
                                           double x, double y, double z, double* gr) {
  int l = 0, m = 0, n = 0;
  bool mixedd = false;
  if (L == 1) {
    l = comp == 0;
    m = comp == 1;
    n = comp == 2;
  } else if (L == 2) {
    const int t[6][3] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
    l = t[comp][0];
    m = t[comp][1];
    n = t[comp][2];
    mixedd = (comp == 1 || comp == 2 || comp == 4);
  }
  const double R = mixedd ? r1 : r0, Rp = mixedd ? d1 : d0;
  auto pw = [](double v, int k) { return k == 0 ? 1.0 : (k == 1 ? v : v * v); };
  const double px = pw(x, l), py = pw(y, m), pz = pw(z, n);
  const double poly = px * py * pz;
  const double dpx = l == 0 ? 0.0 : (l == 1 ? 1.0 : 2.0 * x);
  const double dpy = m == 0 ? 0.0 : (m == 1 ? 1.0 : 2.0 * y);
  const double dpz = n == 0 ? 0.0 : (n == 1 ? 1.0 : 2.0 * z);
  gr[0] = dpx * py * pz * R + x * poly * Rp;
  gr[1] = px * dpy * pz * R + y * poly * Rp;
  gr[2] = px * py * dpz * R + z * poly * Rp;
}


// A shell is taken at a point while alpha_min r^2 <= kShellReach (some
// primitive's exp(-alpha r^2) at least e^-100).
constexpr double kShellReach = 100.0;

// Function tiles the force terms read, per 64-point block: bit t of mask[blk]
// is set when a shell with a function in tile t (16 functions) passes the
// shell test at one of the block's points, with a 1 % margin on kShellReach so
// that every shell k_force_basis takes is covered.  Four threads per point,
// each taking every fourth shell; the bits are OR-ed over the block.
__global__ __launch_bounds__(256) void k_tile_mask(const double* __restrict__ atoms, int nshell,
                                                   const int* __restrict__ sh_atom, const int* __restrict__ sh_L,
                                                   const int* __restrict__ sh_nprim, const int* __restrict__ sh_off,
                                                   const int* __restrict__ sh_cgf, const double* __restrict__ alpha,
                                                   long npts, const double* __restrict__ px,
                                                   const double* __restrict__ py, const double* __restrict__ pz,
                                                   unsigned* mask) {
  __shared__ unsigned red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long p = (long)blockIdx.x * 64 + lane;
  unsigned m = 0;
  if (p < npts) {
    const double x = px[p], y = py[p], z = pz[p];
    for (int s = wave; s < nshell; s += 4) {
      const int a = sh_atom[s];
      const double dx = x - atoms[4 * a], dy = y - atoms[4 * a + 1], dz = z - atoms[4 * a + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      double amin = alpha[sh_off[s]];
      for (int k = 1; k < sh_nprim[s]; ++k) amin = fmin(amin, alpha[sh_off[s] + k]);
      if (amin * r2 <= 1.01 * kShellReach) {
        const int L = sh_L[s], i0 = sh_cgf[s], i1 = i0 + (L + 1) * (L + 2) / 2 - 1;
        for (int t = i0 >> 4; t <= (i1 >> 4); ++t) m |= 1u << t;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) m |= (unsigned)__shfl_xor((int)m, o);
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) mask[blockIdx.x] = red[0] | red[1] | red[2] | red[3];
}

// Basis-centre and grid-motion terms of dE_xc/dR, one wave per 64 points.
// For point p and atom A: g_A(p) = ub_p sum_{i on A} T_pi grad phi_i(r_p).
// part[(A*3+d)*nblk + blk] = -4 sum_{p in blk} g_A(p)  (functions move with A)
// mot[d*npts + p]          = +4 sum_A g_A(p)         (the point moves with its atom)
// Shells are visited atom by atom through ash_start/ash_list (shells of atom
// A are ash_list[ash_start[A] .. ash_start[A+1])).
__global__ __launch_bounds__(64) void k_force_basis(
    const double* __restrict__ atoms, int natoms, const int* __restrict__ ash_start,
    const int* __restrict__ ash_list, const int* __restrict__ sh_L, const int* __restrict__ sh_nprim,
    const int* __restrict__ sh_off, const int* __restrict__ sh_cgf, const double* __restrict__ alpha,
    const double* __restrict__ cn0, const double* __restrict__ cn1, long npts,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    int ldphi, const double* __restrict__ tmat, const double* __restrict__ ub, long nblk, double* part,
    double* mot) {
  const long p = (long)blockIdx.x * 64 + threadIdx.x;
  const bool valid = p < npts;
  const double x = valid ? px[p] : 0.0, y = valid ? py[p] : 0.0, z = valid ? pz[p] : 0.0;
  const double u = valid ? ub[p] : 0.0;
  const double* trow = tmat + (valid ? p : 0) * ldphi;
  double G[3] = {0.0, 0.0, 0.0};
  for (int A = 0; A < natoms; ++A) {
    double g[3] = {0.0, 0.0, 0.0};
    if (valid) {
      const double dx = x - atoms[4 * A], dy = y - atoms[4 * A + 1], dz = z - atoms[4 * A + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      for (int t = ash_start[A]; t < ash_start[A + 1]; ++t) {
        const int s = ash_list[t];
        // every primitive's exp(-alpha r2) below e^-kShellReach (1e-43): the shell's
        // terms are below the rounding of the sums they would enter, and it is skipped
        double amin = alpha[sh_off[s]];
        for (int k = 1; k < sh_nprim[s]; ++k) amin = fmin(amin, alpha[sh_off[s] + k]);
        if (amin * r2 > kShellReach) continue;
        double r0 = 0.0, d0 = 0.0, r1 = 0.0, d1 = 0.0;
        for (int k = 0; k < sh_nprim[s]; ++k) {
          const int q = sh_off[s] + k;
          const double e = exp(-alpha[q] * r2);
          const double m2a = -2.0 * alpha[q];
          r0 += cn0[q] * e;
          r1 += cn1[q] * e;
          d0 += cn0[q] * m2a * e;
          d1 += cn1[q] * m2a * e;
        }
        const int L = sh_L[s], nc = (L + 1) * (L + 2) / 2;
        for (int c = 0; c < nc; ++c) {
          double gr[3];
          shell_grad(L, c, r0, d0, r1, d1, dx, dy, dz, gr);
          const double tv = u * trow[sh_cgf[s] + c];
          g[0] += tv * gr[0];
          g[1] += tv * gr[1];
          g[2] += tv * gr[2];
        }
      }
    }
    for (int d = 0; d < 3; ++d) {
      G[d] += g[d];
      const double t = wave_sum(g[d]);
      if (threadIdx.x == 0) part[((long)A * 3 + d) * nblk + blockIdx.x] = -4.0 * t;
    }
  }
  if (valid)
    for (int d = 0; d < 3; ++d) mot[d * npts + p] = 4.0 * G[d];
}

// ---------------------------------------------------------------------------
// Basis-function and density gradients (GridPoint::set_basis_func_grad and
// set_gradient, src/gridpoint.cpp:62-108).  The gradients are those of the
// amplitude formula; the reference's GTO::get_grad (src/cgf.cpp:67-94) leaves
// out the normalisation, applies the contraction coefficient twice and drops
// the factor l of x^(l-1), and is not followed.
// ---------------------------------------------------------------------------
// Radial sums of shell s at squared distance r2: r0 = sum cN0 e, d0 = sum cN0 (-2 alpha) e,
// and r1, d1 with the mixed-d normalisation.
__device__ __forceinline__ void shell_radial(int s, double r2, const int* __restrict__ sh_nprim,
                                            const int* __restrict__ sh_off, const double* __restrict__ alpha,
                                            const double* __restrict__ cn0, const double* __restrict__ cn1,
                                            double& r0, double& d0, double& r1, double& d1) {
  r0 = d0 = r1 = d1 = 0.0;
  for (int k = 0; k < sh_nprim[s]; ++k) {
    const int q = sh_off[s] + k;
    const double e = exp(-alpha[q] * r2);
    const double m2a = -2.0 * alpha[q];
    r0 += cn0[q] * e;
    r1 += cn1[q] * e;
    d0 += cn0[q] * m2a * e;
    d1 += cn1[q] * m2a * e;
  }
}

// Gradients of every basis function at every point, out[p][d][i] (d = x, y, z;
// per point the nbf x 3 column-major matrix of src/gridpoint.cpp:62-72), one
// thread per (point, shell); points p0 .. p0 + npts of px/py/pz.
__global__ void k_phi_grad(const double* __restrict__ atoms, int nshell, const int* __restrict__ sh_atom,
                           const int* __restrict__ sh_L, const int* __restrict__ sh_nprim,
                           const int* __restrict__ sh_off, const int* __restrict__ sh_cgf,
                           const double* __restrict__ alpha, const double* __restrict__ cn0,
                           const double* __restrict__ cn1, long p0, long npts, const double* __restrict__ px,
                           const double* __restrict__ py, const double* __restrict__ pz, int nbf, double* out) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts * nshell) return;
  const long p = g / nshell;
  const int s = (int)(g % nshell);
  const int a = sh_atom[s];
  const double dx = px[p0 + p] - atoms[4 * a], dy = py[p0 + p] - atoms[4 * a + 1], dz = pz[p0 + p] - atoms[4 * a + 2];
  double r0, d0, r1, d1;
  shell_radial(s, dx * dx + dy * dy + dz * dz, sh_nprim, sh_off, alpha, cn0, cn1, r0, d0, r1, d1);
  const int L = sh_L[s], nc = (L + 1) * (L + 2) / 2;
  double* o = out + p * 3 * nbf + sh_cgf[s];
  for (int c = 0; c < nc; ++c) {
    double gr[3];
    shell_grad(L, c, r0, d0, r1, d1, dx, dy, dz, gr);
    o[c] = gr[0];
    o[nbf + c] = gr[1];
    o[2 * nbf + c] = gr[2];
  }
}

// Weights of the two-electron Fock part G = 2J + XC = Phi^T diag(2 uJ + uxc) Phi.
__global__ void k_fock_weights(const double* __restrict__ uJ, const double* __restrict__ uxc, long n, double* uG) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) uG[i] = 2.0 * uJ[i] + uxc[i];
}

// x[i] = v for i < n.
__global__ void k_fill(double* x, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Points [n][3] -> x, y, z arrays.
__global__ void k_split_xyz(const double* __restrict__ xyz, long n, double* px, double* py, double* pz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  px[i] = xyz[3 * i];
  py[i] = xyz[3 * i + 1];
  pz[i] = xyz[3 * i + 2];
}

// Density and its gradient at points p0 .. p0 + npts from T = Phi P (rows
// [p][ldphi] of tmat, row 0 = point p0): rho = 2 sum_j T_j phi_j and
// grad rho = 2 phi^T P grad phi + 2 grad phi^T P phi = 4 sum_j T_j grad phi_j
// (P symmetric), both times fac; fac = zsum / scal[1] (the renormalisation of
// src/moleculargrid.cpp:132-146) when scal is given, else 1.  One thread per
// point, shells in order; rho (may be NULL) [npts], grad [npts][3].
__global__ __launch_bounds__(256) void k_density_grad(
    const double* __restrict__ atoms, int nshell, const int* __restrict__ sh_atom, const int* __restrict__ sh_L,
    const int* __restrict__ sh_nprim, const int* __restrict__ sh_off, const int* __restrict__ sh_cgf,
    const double* __restrict__ alpha, const double* __restrict__ cn0, const double* __restrict__ cn1, long p0,
    long npts, const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    int ldphi, const double* __restrict__ tmat, const double* __restrict__ scal, double zsum, double* rho,
    double* grad) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const double x = px[p0 + p], y = py[p0 + p], z = pz[p0 + p];
  const double* trow = tmat + p * ldphi;
  double v = 0.0, g[3] = {0.0, 0.0, 0.0};
  for (int s = 0; s < nshell; ++s) {
    const int a = sh_atom[s];
    const double dx = x - atoms[4 * a], dy = y - atoms[4 * a + 1], dz = z - atoms[4 * a + 2];
    double r0, d0, r1, d1;
    shell_radial(s, dx * dx + dy * dy + dz * dz, sh_nprim, sh_off, alpha, cn0, cn1, r0, d0, r1, d1);
    const int L = sh_L[s], nc = (L + 1) * (L + 2) / 2;
    for (int c = 0; c < nc; ++c) {
      const double t = trow[sh_cgf[s] + c];
      double gr[3];
      shell_grad(L, c, r0, d0, r1, d1, dx, dy, dz, gr);
      double phi;
      if (L == 0)
        phi = r0;
      else if (L == 1)
        phi = r0 * (c == 0 ? dx : (c == 1 ? dy : dz));
      else {
        const double cx[6] = {dx * dx, dx * dy, dx * dz, dy * dy, dy * dz, dz * dz};
        phi = ((c == 1 || c == 2 || c == 4) ? r1 : r0) * cx[c];
      }
      v += t * phi;
      g[0] += t * gr[0];
      g[1] += t * gr[1];
      g[2] += t * gr[2];
    }
  }
  const double fac = scal ? zsum / scal[1] : 1.0;
  if (rho) rho[p] = 2.0 * v * fac;
  for (int d = 0; d < 3; ++d) grad[3 * p + d] = 4.0 * g[d] * fac;
}

// Becke-weight term: sum_p w0_p a_p d wb_p / d R_B (w0 = weight without the
// Becke factor), one wave per 64 points with each point's active cell functions
// (atom and P_n) in LDS.  The B != owner terms are added to part[B][d][blk]
// (after k_force_basis wrote it), the owner's share to mot[d][p].
// d wb/d R_B = (dP_a/dR_B)/Z - (wb/Z) sum_n dP_n/dR_B for B != a, and
// d wb/d R_a = -sum_{B != a} d wb/d R_B (translating everything leaves wb fixed).
// dP_n/dR_B = P_n sum_m (s'/s)(mu_nm) d mu_nm/d R_B with
// d mu_nm/d R_n = -(u_n + mu_nm e_nm)/R_nm, d mu_nm/d R_m = (u_m + mu_nm e_nm)/R_nm,
// u_k = (r - R_k)/|r - R_k|, e_nm = (R_n - R_m)/R_nm.
__device__ __forceinline__ void cutoff_and_ratio(double mu, double& s, double& ratio) {
  // s = (1 - f(f(f(mu))))/2, s' = -f'(f2) f'(f1) f'(mu)/2, f'(x) = 1.5 - 1.5 x^2
  const double f1 = 1.5 * mu - 0.5 * mu * mu * mu;
  const double f2 = 1.5 * f1 - 0.5 * f1 * f1 * f1;
  const double f3 = 1.5 * f2 - 0.5 * f2 * f2 * f2;
  s = 0.5 * (1.0 - f3);
  const double sp = -0.5 * (1.5 - 1.5 * f2 * f2) * (1.5 - 1.5 * f1 * f1) * (1.5 - 1.5 * mu * mu);
  ratio = (s > 0.0) ? sp / s : 0.0;
}

__device__ __forceinline__ double becke_s(double mu) {
  for (int it = 0; it < 3; ++it) mu = 1.5 * mu - 0.5 * (mu * mu * mu);
  return 0.5 * (1.0 - mu);
}

// Terms of an atom n whose cell function P_n is at most kBeckeSkip times
// Z = sum_n P_n at the point are skipped.  Such a term is s'(mu) times the
// other factors of P_n (at most 1); with |s'| <= 10.5 s^{7/8} (exact
// arithmetic, the bound is approached as mu -> 1) it is at most
// 10.5 (kBeckeSkip Z)^{7/8} ~ 3e-17 Z^{7/8}.
constexpr double kBeckeSkip = 1e-20;

// At most kBeckeActive cell functions per point may lie above the skip level
// (about 15 on a linear alkane); a point with more sets *overflow and the
// caller reports the gradient as unsupported.
constexpr int kBeckeActive = 24;

// Active entries per pass of the Becke-term loops (held in registers).
constexpr int kBeckeBlock = 8;

__device__ __forceinline__ double atom_dist(const double* __restrict__ atoms, int m, double x, double y, double z) {
  const double dx = x - atoms[4 * m], dy = y - atoms[4 * m + 1], dz = z - atoms[4 * m + 2];
  return fmax(sqrt(dx * dx + dy * dy + dz * dz), 1e-300);
}

// Cell function P_n at the point, its factors taken with the nearest atom
// first and stopped once the product is at most cut (then it is below the
// level that matters and is returned as is).
__device__ __forceinline__ double cell_function(const double* __restrict__ atoms,
                                               const double* __restrict__ invR, int natoms, int n, int near,
                                               double dn, double dnear, double x, double y, double z,
                                               double cut) {
  double prod = (n == near) ? 1.0 : becke_s((dn - dnear) * invR[n * natoms + near]);
  for (int m = 0; m < natoms && prod > cut; ++m) {
    if (m == n || m == near) continue;
    prod *= becke_s((dn - atom_dist(atoms, m, x, y, z)) * invR[n * natoms + m]);
  }
  return prod;
}

__global__ __launch_bounds__(64) void k_force_becke(const double* __restrict__ atoms,
                                                    const double* __restrict__ invR, int natoms,
                                                    long npts, long cell, int nrad, int nang,
                                                    const double* __restrict__ rad,
                                                    const double* __restrict__ wrad,
                                                    const double* __restrict__ leb,
                                                    const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz,
                                                    const double* __restrict__ ab, long nblk,
                                                    double* part, double* mot, int* overflow) {
  // the point's cell functions above the skip level: atom and P_n (|r - R_n| is
  // recomputed where needed: 10 bytes per entry keep ten work-groups per CU)
  __shared__ unsigned short act_n[kBeckeActive * 64];
  __shared__ double act_P[kBeckeActive * 64];
  const int lane = threadIdx.x;
  const long p = (long)blockIdx.x * 64 + lane;
  double w0a = 0.0;  // weight without the Becke factor, times a_p
  if (p < npts) {
    const int j = (int)(p % nang);
    const int i = (int)((p / nang) % nrad);
    const double ox = leb[4 * j] * rad[i], oy = leb[4 * j + 1] * rad[i], oz = leb[4 * j + 2] * rad[i];
    w0a = wrad[i] * leb[4 * j + 3] * (ox * ox + oy * oy + oz * oz) * 4.0 * kPiD * ab[p];
  }
  // points with |w0 a_p| <= 1e-30 (far out, where the density has vanished and
  // many cells overlap) add nothing that survives rounding and are skipped
  const bool valid = p < npts && fabs(w0a) > 1e-30;
  const double x = valid ? px[p] : 0.0, y = valid ? py[p] : 0.0, z = valid ? pz[p] : 0.0;
  int near = 0;
  double dnear = atom_dist(atoms, 0, x, y, z);
  for (int m = 1; m < natoms; ++m) {
    const double d = atom_dist(atoms, m, x, y, z);
    if (d < dnear) {
      dnear = d;
      near = m;
    }
  }
  // Z: the nearest atom's P is a lower bound, and a product that drops below
  // 1e-4 kBeckeSkip of it (below the skip level) is kept as truncated, which
  // changes Z by less than 1e-24 Z per atom.  The same pass lists the
  // candidates P > kBeckeSkip pnear, a superset of the active P > kBeckeSkip Z
  // (Z >= pnear) whose P are exact (a product above the skip level never
  // reaches the truncation); the list is then cut to the active ones.
  double pnear = 1.0, Z = 1.0, Pown = 0.0;
  const int own = valid ? (int)(p / cell) : 0;
  int ncand = 0;
  bool cand_full = false;
  if (valid) {
    pnear = cell_function(atoms, invR, natoms, near, near, dnear, dnear, x, y, z, 0.0);
    const double cut0 = 1e-4 * kBeckeSkip * pnear, pre = kBeckeSkip * pnear;
    Z = pnear;
    // the cell functions of kBeckeBlock atoms at a time, each factor sequence as in
    // cell_function (the nearest atom first, then m ascending, stopped once the
    // product is at most cut0), the distance to each atom m shared by the block
    for (int n0 = 0; n0 < natoms; n0 += kBeckeBlock) {
      double dn[kBeckeBlock], prod[kBeckeBlock];
      bool alive = false;
#pragma unroll
      for (int k = 0; k < kBeckeBlock; ++k) {
        const int n = n0 + k;
        const bool in = n < natoms && n != near;
        const int ns = in ? n : near;
        dn[k] = atom_dist(atoms, ns, x, y, z);
        prod[k] = in ? becke_s((dn[k] - dnear) * invR[ns * natoms + near]) : 0.0;
        alive |= prod[k] > cut0;
      }
      for (int m = 0; m < natoms && __ballot(alive) != 0; ++m) {
        const double dm = atom_dist(atoms, m, x, y, z);
        alive = false;
#pragma unroll
        for (int k = 0; k < kBeckeBlock; ++k) {
          const int n = min(n0 + k, natoms - 1);  // past the last atom: prod[k] = 0, no factor taken
          if (m != n && m != near && prod[k] > cut0) prod[k] *= becke_s((dn[k] - dm) * invR[n * natoms + m]);
          alive |= prod[k] > cut0;
        }
      }
#pragma unroll
      for (int k = 0; k < kBeckeBlock; ++k) {
        const int n = n0 + k;
        if (n >= natoms) break;
        double P = pnear;
        if (n != near) {
          P = prod[k];
          Z += P;
        }
        if (n == own) Pown = P;
        if (P > pre) {
          if (ncand < kBeckeActive) {
            act_n[ncand * 64 + lane] = (unsigned short)n;
            act_P[ncand * 64 + lane] = P;
            ++ncand;
          } else {
            cand_full = true;
          }
        }
      }
    }
  }
  const double skip = kBeckeSkip * Z;
  int nact = 0;
  if (!cand_full) {
    for (int t = 0; t < ncand; ++t) {
      const double P = act_P[t * 64 + lane];
      if (P > skip) {
        act_n[nact * 64 + lane] = act_n[t * 64 + lane];
        act_P[nact * 64 + lane] = P;
        ++nact;
      }
    }
  } else {
    // more candidates than slots: list the active cells from scratch
    for (int n = 0; n < natoms; ++n) {
      const double dn = (n == near) ? dnear : atom_dist(atoms, n, x, y, z);
      const double P = (n == near) ? pnear : cell_function(atoms, invR, natoms, n, near, dn, dnear, x, y, z, skip);
      if (P > skip) {
        if (nact < kBeckeActive) {
          act_n[nact * 64 + lane] = (unsigned short)n;
          act_P[nact * 64 + lane] = P;
          ++nact;
        } else {
          *overflow = 1;
        }
      }
    }
  }
  const double pref = (valid && Z > 0.0) ? w0a / Z : 0.0;
  const double wbv = (valid && Z > 0.0) ? Pown / Z : 0.0;
  const bool on = valid && pref != 0.0;
  double own_acc[3] = {0.0, 0.0, 0.0};
  // Every pair (active n, atom m != n) enters twice with the same
  // rho = ratio(mu_nm), f = P_n rho / R_nm and e_nm: into dP_n/dR_m as
  // f (u_m + mu e_nm), and into dP_n/dR_n as -f (u_n + mu e_nm).  Passes over
  // kBeckeBlock active entries at a time (registers), atoms m in order: output m
  // is complete at the end of its step; output n sums F = sum_m f and
  // G = sum_m f mu e_nm over the pass and is added after it, as
  // dP_n/dR_n = -(u_n F + G).  Per atom B the sums go to part[B] in a fixed order:
  // the m terms of pass 0, the n terms of pass 0, then those of pass 1, ...
  int nmax = on ? nact : 0;
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
  for (int t0 = 0; t0 < nmax; t0 += kBeckeBlock) {
    int nk[kBeckeBlock];
    double Pk[kBeckeBlock], dk[kBeckeBlock], wk[kBeckeBlock][3], F[kBeckeBlock], G[kBeckeBlock][3];
#pragma unroll
    for (int k = 0; k < kBeckeBlock; ++k) {
      const bool in = on && t0 + k < nact;
      nk[k] = in ? act_n[(t0 + k) * 64 + lane] : -1;
      Pk[k] = in ? act_P[(t0 + k) * 64 + lane] : 0.0;
      const int n = in ? nk[k] : 0;
      // w = R_n - r
      wk[k][0] = atoms[4 * n] - x;
      wk[k][1] = atoms[4 * n + 1] - y;
      wk[k][2] = atoms[4 * n + 2] - z;
      dk[k] = atom_dist(atoms, n, x, y, z);
      F[k] = G[k][0] = G[k][1] = G[k][2] = 0.0;
    }
    for (int m = 0; m < natoms; ++m) {
      const double dm = atom_dist(atoms, m, x, y, z);
      const double vm[3] = {x - atoms[4 * m], y - atoms[4 * m + 1], z - atoms[4 * m + 2]};  // r - R_m
      const double um[3] = {vm[0] / dm, vm[1] / dm, vm[2] / dm};
      double S[3] = {0.0, 0.0, 0.0}, Dp[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < kBeckeBlock; ++k) {
        const int n = nk[k];
        if (n < 0 || n == m) continue;
        const double iR = invR[n * natoms + m];
        const double mu = (dk[k] - dm) * iR;
        double sv, rt;
        cutoff_and_ratio(mu, sv, rt);
        const double f = Pk[k] * rt * iR, fm = f * mu;
        F[k] += f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double me = fm * ((wk[k][d] + vm[d]) * iR);  // f mu e_nm
          const double tv = f * um[d] + me;
          S[d] += tv;
          if (n == own) Dp[d] = tv;
          G[k][d] += me;
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double gm = (on && m != own) ? pref * (Dp[d] - wbv * S[d]) : 0.0;
        own_acc[d] -= gm;
        const double t = wave_sum(gm);
        if (lane == 0) part[((long)m * 3 + d) * nblk + blockIdx.x] += t;
      }
    }
    // the pass's entries n: d wb/dR_n gains pref wbv (u_n F + G), u_n = -w / d_n
#pragma unroll
    for (int k = 0; k < kBeckeBlock; ++k) {
      const bool use = nk[k] >= 0 && nk[k] != own;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        G[k][d] = use ? pref * wbv * (G[k][d] - wk[k][d] / dk[k] * F[k]) : 0.0;
        own_acc[d] -= G[k][d];
      }
    }
    for (int B = 0; B < natoms; ++B) {
      double v[3] = {0.0, 0.0, 0.0};
      bool has = false;
#pragma unroll
      for (int k = 0; k < kBeckeBlock; ++k)
        if (nk[k] == B) {
          has = true;
#pragma unroll
          for (int d = 0; d < 3; ++d) v[d] = G[k][d];
        }
      if (__ballot(has) == 0) continue;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double t = wave_sum(v[d]);
        if (lane == 0) part[((long)B * 3 + d) * nblk + blockIdx.x] += t;
      }
    }
  }
  // the owner atom's share (translation invariance), per point
  if (valid)
    for (int d = 0; d < 3; ++d) mot[d * npts + p] += own_acc[d];
}

// part[blk] = sum over the block's 256 points of x*y (fixed tree order).
__global__ __launch_bounds__(256) void k_dot_parts(const double* __restrict__ x,
                                                   const double* __restrict__ y, long n, double* part) {
  __shared__ double red[256];
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  red[threadIdx.x] = p < n ? x[p] * y[p] : 0.0;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[r] = sum_c in[r * len + c], one block per row, fixed order.
__global__ __launch_bounds__(1024) void k_row_sum(const double* __restrict__ in, long len, double* out) {
  __shared__ double red[1024];
  const double* src = in + (long)blockIdx.x * len;
  double s = 0.0;
  for (long i = threadIdx.x; i < len; i += 1024) s += src[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Per-point factors of the adjoint, k = lam/Z (0 with c held), c = Z / scalars[1]:
// ub = c w (v - k), ab = e - k rho, udp = 2 c w (v - k) (the dE/dP weights);
// uxc holds w v and exd the energy density e.
__global__ void k_adjoint_weights(const double* __restrict__ uxc, const double* __restrict__ w,
                                  const double* __restrict__ rho, const double* __restrict__ exd,
                                  const double* __restrict__ scal, const double* __restrict__ lam,
                                  double zsum, int hold_c, long n, double* ub, double* ab, double* udp) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double c = zsum / scal[1];
  const double k = hold_c ? 0.0 : lam[0] / zsum;
  const double wv = uxc[p] - k * w[p];
  ub[p] = c * wv;
  ab[p] = exd[p] - k * rho[p];
  udp[p] = 2.0 * c * wv;
}

// FP32 copy of the amplitudes for the reduced-precision contractions.
__global__ void k_phi_f32(const double* __restrict__ phi, long n, float* phf) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) phf[e] = (float)phi[e];
}

__device__ __forceinline__ flt4 mfma16x16x4f32(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// out_o = sum over splits of the partial tiles.  A block of 256 threads takes
// 16 tile elements; thread (element, r) sums the splits r, r+16, ... and the 16
// partial sums are added in order, so the result does not depend on timing.
// Tiles on the diagonal contribute their upper triangle only and every element
// is mirrored, so the output is exactly symmetric.
// With sa.scalars set, the last work-group evaluates the build's scalars
// instead (k_scalars, which depends on nothing this kernel writes).
__global__ __launch_bounds__(256) void k_contract_reduce(const double* __restrict__ part, int nsplit,
                                                         int ngroups, int nout,
                                                         const int* __restrict__ groups, int nbf,
                                                         double* out0, double* out1, ScalarArgs sa) {
  if (sa.scalars && blockIdx.x == gridDim.x - 1) {
    scalars_block(sa);
    return;
  }
  __shared__ double red[16][17];
  const int el = threadIdx.x & 15, r = threadIdx.x >> 4;
  const long e = (long)blockIdx.x * 16 + el;
  bool live = e < (long)ngroups * nout * 1024;
  double s = 0.0;
  const int idx = (int)(e & 1023);
  const int o = live ? (int)((e >> 10) % nout) : 0;
  const int grp = live ? (int)((e >> 10) / nout) : 0;
  if (live && (idx >> 8) >= groups[3 * grp + 2]) live = false;  // tile beyond the group: not read
  if (live) {
    const long stride = (long)ngroups * nout * 1024;
    const double* src = part + ((long)grp * nout + o) * 1024 + idx;
    // eight splits' partials loaded per round (one memory latency), added in split order
    int sp = r;
    for (; sp + 16 * 7 < nsplit; sp += 16 * 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = src[(long)(sp + 16 * u) * stride];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; sp < nsplit; sp += 16) s += src[sp * stride];
  }
  red[el][r] = s;
  __syncthreads();
  if (r != 0 || !live) return;
  double tot = 0.0;
  for (int q = 0; q < 16; ++q) tot += red[el][q];
  const int t = idx >> 8, v = (idx >> 6) & 3, ln = idx & 63;
  // tile (I, J0 + t); result register v of lane (c, g) holds row 4v + g
  if (t >= groups[3 * grp + 2]) return;
  const int ti = groups[3 * grp], tj = groups[3 * grp + 1] + t, il = 4 * v + (ln >> 4);
  const int jl = ln & 15;
  const int i = 16 * ti + il, j = 16 * tj + jl;
  if (i >= nbf || j >= nbf || ti > tj || (ti == tj && il > jl)) return;
  double* out = (o == 0) ? out0 : out1;
  if (!out) return;
  out[i + (long)j * nbf] = tot;
  out[j + (long)i * nbf] = tot;
}

// One wave: d[v*64 + lane] = FP32 register v of lane after the FP32 16x16x4 MFMA.
__global__ void k_mfma_raw_f32(const float* a, const float* b, float* d) {
  const int lane = threadIdx.x;
  flt4 r = mfma16x16x4f32(a[lane], b[lane], flt4{0.f, 0.f, 0.f, 0.f});
  for (int v = 0; v < 4; ++v) d[v * 64 + lane] = r[v];
}

// One wave: d[v*64 + lane] = register v of lane after D = mfma(a[lane], b[lane], 0).
__global__ void k_mfma_raw(const double* a, const double* b, double* d) {
  const int lane = threadIdx.x;
  dbl4 r = mfma16x16x4(a[lane], b[lane], dbl4{0.0, 0.0, 0.0, 0.0});
  for (int v = 0; v < 4; ++v) d[v * 64 + lane] = r[v];
}

// Probe of the 16x16x4 FP64 MFMA lane layout assumed above:
// A[i][k] on lane i + 16k, B[k][j] on lane j + 16k, D[4v+g][c] in register v of lane c + 16g.
__global__ void k_mfma_probe(double* out) {
  const int lane = threadIdx.x;
  const int c = lane & 15, g = lane >> 4;
  const double a = 1.0 + c + 100.0 * g;          // A[c][g]
  const double b = 1.0 + 0.5 * c - 3.0 * g;       // B[g][c]
  dbl4 d = mfma16x16x4(a, b, dbl4{0.0, 0.0, 0.0, 0.0});
  for (int v = 0; v < 4; ++v) out[v * 64 + lane] = d[v];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(DG_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                  " bytes failed: " + hipGetErrorString(e));
  }
  return DG_OK;
}

long round_up(long x, long m) { return (x + m - 1) / m * m; }

}  // namespace

struct ProfEntry {
  std::string name;
  hipEvent_t e0, e1;
};
struct ProfStat {
  double ms = 0.0;
  int n = 0;
};

// Tile blocks of the 2-D contraction for one block width NJ: {I0, nI, J0, nJ} per
// block, its four reduction groups (I0 + r, J0, nJ) in the k_contract_reduce
// layout, the 64-point block list of each block (all blocks, or with screening
// those where its tiles are significant), and per virtual work-group id the
// {tile block, split} it runs.
struct TileBlocks {
  int *cb = nullptr, *groups = nullptr, *list = nullptr, *loff = nullptr;
  int2* order = nullptr;
  int ncb = 0, nsplit = 0;
};

struct dg_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  int natoms = 0, nbf = 0, nshell = 0, nrad = 0, nang = 0, lmax = 0, nlm = 0, precision = 0;
  long npts = 0, npts_pad = 0, cell = 0;
  int ldphi = 0, ntb = 0, ngroups = 0, nsplit_small = 0;
  int cus = 256;  // compute units of the device
  // tile blocks of the 2-D contraction: [0] 4 x 2 (two outputs), [1] 4 x 4 (one output)
  TileBlocks tb[2];
  long nlda_blocks = 0;
  double zsum = 0.0, xfac = 0.0;
  VwnConst vwn{};
  std::vector<std::pair<int, int>> lchunks;  // l ranges of the cross kernel
  size_t cross_lds = 0;
  // device buffers
  double *atoms = nullptr, *invR = nullptr, *rad = nullptr, *wrad = nullptr, *leb = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr, *w = nullptr, *wb = nullptr;
  double* phi = nullptr;
  int *sh_atom = nullptr, *sh_L = nullptr, *sh_nprim = nullptr, *sh_off = nullptr, *sh_cgf = nullptr;
  double *pr_alpha = nullptr, *pr_cn0 = nullptr, *pr_cn1 = nullptr;
  double *P = nullptr, *pfrag = nullptr, *rho = nullptr, *uxc = nullptr, *exd = nullptr;
  double *uJ = nullptr, *V = nullptr, *vcell = nullptr, *vpart = nullptr;
  double* uG = nullptr;  // 2 uJ + uxc (dg_build_fock_device, allocated on first use)
  double *q = nullptr, *qc = nullptr, *exc_part = nullptr, *scalars = nullptr;
  double* qpart = nullptr;  // split cell populations (cell_split > 1)
  int cell_split = 1;
  double *ylm = nullptr, *minvT = nullptr, *shc = nullptr;
  double *knots = nullptr, *invh = nullptr, *KD = nullptr;
  double *rho_lm = nullptr, *U = nullptr, *tab = nullptr;
  int* blocks = nullptr;
  int* items = nullptr;
  unsigned short* perm = nullptr;  // per item, point order by distance to the item's atom (padded)
  unsigned short* pkey = nullptr;  // interval key of every perm slot
  int* plen = nullptr;             // per item, padded length of its order
  long npairs = 0;                 // capacity of perm
  int nitems = 0, cross_wgs = 0;
  // screening masks
  double screen = 0.0;
  unsigned* bmask = nullptr;
  // reduced-precision mode
  float* phf = nullptr;  // FP32 amplitudes
  double* ej_part = nullptr;
  double *cpart = nullptr, *J = nullptr, *XC = nullptr;
  long* vbase = nullptr;  // [natoms] offset of atom k's partial potentials in vpart (minus first point)
  // batch mode: several molecules in one handle (nmol = 0 for one molecule)
  int nmol = 0;
  int *atom_mol = nullptr, *mol_atom0 = nullptr, *mol_shell0 = nullptr, *mol_nbf = nullptr;
  long* mol_poff = nullptr;  // offsets of the packed nbf_m x nbf_m matrices
  double *mol_z = nullptr, *mol_fac = nullptr, *mol_qraw = nullptr, *atom_exc = nullptr, *atom_ej = nullptr;
  int* bchunks = nullptr;   // contraction chunks {m, p_begin, p_end}
  int* mol_chunk0 = nullptr;
  int nbchunks = 0, max_mol_shells = 0, max_mol_atoms = 0, bchunk_len = 2048, item_len = 2048;
  long packed = 0;  // sum of nbf_m^2
  std::vector<int> h_mol_atom0, h_mol_nbf;
  std::vector<long> h_mol_poff;
  size_t items_cap = 0, bchunks_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
  // adjoint of E_xc (allocated on its first use)
  double *adj_ub = nullptr, *adj_ab = nullptr, *adj_udp = nullptr, *adj_T = nullptr, *adj_frag = nullptr;
  double *adj_part = nullptr, *adj_mot = nullptr, *adj_lam = nullptr, *adj_out = nullptr, *adj_dP = nullptr;
  int *ash_start = nullptr, *ash_list = nullptr;
  unsigned* adj_tmask = nullptr;  // per 64-point block, the function tiles the force terms read
  int *adj_list = nullptr, *adj_loff = nullptr;  // dE/dP: point blocks per one-output tile block
  // state
  bool has_density = false, has_hartree = false, has_xc = false, corrected = false;
  const double* last_P = nullptr;  // device density matrix of the last density evaluation
  // profiling
  bool prof = false;
  std::vector<ProfEntry> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, ProfStat> stats;

  ~dg_handle() {
    double* bufs[] = {atoms, invR,  rad,   wrad,   leb,   px,     py,    pz,   w,     wb,
                      phi,   pr_alpha, pr_cn0, pr_cn1, P, pfrag, rho,  uxc,   exd,   uJ,
                      V,     vcell, vpart, q,      qc,    exc_part, scalars, ylm, minvT, shc,
                      knots, invh,  KD,    rho_lm, U,     tab,  cpart, J,   XC,
                      adj_ub, adj_ab, adj_udp, adj_T, adj_frag, adj_part, adj_mot, adj_lam, adj_out, adj_dP,
                      mol_z, mol_fac, mol_qraw, atom_exc, atom_ej, qpart, uG};
    for (double* b : bufs)
      if (b) (void)hipFree(b);
    int* ibufs[] = {sh_atom, sh_L, sh_nprim, sh_off, sh_cgf, blocks, items, plen, ash_start, ash_list, adj_list, adj_loff,
                    atom_mol, mol_atom0, mol_shell0, mol_nbf, bchunks, mol_chunk0,
                    tb[0].cb, tb[0].groups, tb[0].list, tb[0].loff, tb[1].cb, tb[1].groups, tb[1].list,
                    tb[1].loff};
    if (bmask) (void)hipFree(bmask);
    if (adj_tmask) (void)hipFree(adj_tmask);
    for (TileBlocks& t : tb)
      if (t.order) (void)hipFree(t.order);
    if (vbase) (void)hipFree(vbase);
    if (perm) (void)hipFree(perm);
    if (pkey) (void)hipFree(pkey);
    if (mol_poff) (void)hipFree(mol_poff);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (phf) (void)hipFree(phf);
    if (ej_part) (void)hipFree(ej_part);
    for (int* b : ibufs)
      if (b) (void)hipFree(b);
    for (auto& e : pending) {
      (void)hipEventDestroy(e.e0);
      (void)hipEventDestroy(e.e1);
    }
    for (auto& e : event_pool) (void)hipEventDestroy(e);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

hipEvent_t take_event(dg_handle* h) {
  hipEvent_t e;
  if (!h->event_pool.empty()) {
    e = h->event_pool.back();
    h->event_pool.pop_back();
  } else {
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
  }
  return e;
}

// Runs launch() on the handle's stream, bracketed by events when profiling.
template <typename F>
int run(dg_handle* h, const char* name, F launch) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->prof) {
    e0 = take_event(h);
    e1 = take_event(h);
    if (e0) (void)hipEventRecord(e0, h->stream);
  }
  launch();
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(DG_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(err));
  if (h->prof && e0 && e1) {
    (void)hipEventRecord(e1, h->stream);
    h->pending.push_back({name, e0, e1});
  }
  return DG_OK;
}

#define DG_TRY(expr)              \
  do {                            \
    int _rc = (expr);             \
    if (_rc != DG_OK) return _rc; \
  } while (0)

unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Per 4 x 2 tile block {I0, nI, J0, nJ}, the 64-point blocks on which one of its
// row tiles and one of its column tiles are significant (all blocks without
// screening), and the list offsets.
void tile_block_lists(const std::vector<unsigned>& mk, const std::vector<int>& cbs, std::vector<int>& list,
                      std::vector<int>& off) {
  const int n = (int)cbs.size() / 4;
  off.assign(n + 1, 0);
  list.clear();
  for (int q = 0; q < n; ++q) {
    const unsigned ri = ((1u << cbs[4 * q + 1]) - 1u) << cbs[4 * q];
    const unsigned cj = ((1u << cbs[4 * q + 3]) - 1u) << cbs[4 * q + 2];
    for (size_t b = 0; b < mk.size(); ++b)
      if ((mk[b] & ri) && (mk[b] & cj)) list.push_back((int)b);
    off[q + 1] = (int)list.size();
  }
}

// Orders the points of every cross-potential item by distance to its atom.
int sort_items(dg_handle* h) {
  if (h->nitems == 0) return DG_OK;
  return run(h, "sort_items", [&] {
    hipLaunchKernelGGL(k_sort_items, dim3(h->nitems), dim3(1024), 0, h->stream, h->items, h->atoms, h->px, h->py,
                       h->pz, h->knots, h->nrad, h->perm, h->pkey, h->plen);
  });
}

template <int NT>
void launch_density_nt(dg_handle* h) {
  // Without screening, up to 400 functions: the fragments streamed through a
  // ring, two work-groups per CU (the amplitude rows, 8 NT registers, leave room
  // for two waves per SIMD).  With screening: small bases (NT <= 12) eight waves
  // share double-buffered tiles; up to 400 functions one tile buffer and two
  // work-groups per CU; beyond, double-buffered with one wave per SIMD.
  // npts_pad is a multiple of 256.
  constexpr int W = NT <= 12 ? 8 : 4;
  constexpr int NBUF = (NT > 12 && NT <= 25) ? 1 : 2;
  if constexpr (NT >= 8 && NT <= 25) {
    if (!h->bmask) {
      hipLaunchKernelGGL(k_density_ring<NT>, dim3(nblocks(h->npts_pad, 64)), dim3(256), 0, h->stream, h->phi,
                         h->ldphi, h->pfrag, h->rho);
      return;
    }
  }
  hipLaunchKernelGGL((k_density<NT, W, NBUF>), dim3(nblocks(h->npts_pad, 16 * W)), dim3(64 * W), 0, h->stream,
                     h->phi, h->ldphi, h->pfrag, h->rho, h->bmask);
}

template <int NT>
void launch_density_pl_nt(dg_handle* h, const double* dP) {
  const long nblk = h->npts_pad / 64;
  const unsigned G = (unsigned)std::min<long>(nblk, 1024);
  hipLaunchKernelGGL(k_density_pl<NT>, dim3(G), dim3(256), 0, h->stream, h->phi, h->ldphi, dP, h->nbf, nblk, h->rho,
                     h->bmask);
}

int launch_density(dg_handle* h, const double* dP) {
  if (h->ntb <= 7)  // fragments within 64 KB of LDS
    return run(h, "density", [&] {
      switch (h->ntb) {
        case 1: launch_density_pl_nt<1>(h, dP); break;
        case 2: launch_density_pl_nt<2>(h, dP); break;
        case 3: launch_density_pl_nt<3>(h, dP); break;
        case 4: launch_density_pl_nt<4>(h, dP); break;
        case 5: launch_density_pl_nt<5>(h, dP); break;
        case 6: launch_density_pl_nt<6>(h, dP); break;
        case 7: launch_density_pl_nt<7>(h, dP); break;
        default: break;
      }
    });
  return run(h, "density", [&] {
    switch (h->ntb) {
#define DG_NT(N) \
  case N:        \
    launch_density_nt<N>(h); \
    break;
      DG_NT(1) DG_NT(2) DG_NT(3) DG_NT(4) DG_NT(5) DG_NT(6) DG_NT(7) DG_NT(8) DG_NT(9) DG_NT(10) DG_NT(11)
      DG_NT(12) DG_NT(13) DG_NT(14) DG_NT(15) DG_NT(16) DG_NT(17) DG_NT(18) DG_NT(19) DG_NT(20) DG_NT(21)
      DG_NT(22) DG_NT(23) DG_NT(24) DG_NT(25) DG_NT(26) DG_NT(27) DG_NT(28) DG_NT(29) DG_NT(30) DG_NT(31)
      DG_NT(32)
#undef DG_NT
      default:
        break;
    }
  });
}

// tmask: the tiles that matter per 64-point block (k_tile_mask), or null for
// all; only the ring kernel uses it, the others form every tile.
template <int NT>
void launch_tmat_nt(hipStream_t st, long npts_pad, const double* phi, int ldphi, const double* frag, double* tmat,
                    const unsigned* tmask) {
  // 8 to 25 tiles: the fragments streamed through a ring, two work-groups per CU
  if constexpr (NT >= 8 && NT <= 25) {
    hipLaunchKernelGGL(k_tmat_ring<NT>, dim3(nblocks(npts_pad, 64)), dim3(256), 0, st, phi, ldphi, frag, tmat,
                       tmask);
    return;
  }
  // eight waves (two per SIMD) when the rows come in 128s and the amplitude
  // registers (8 NT) leave room for both
  if constexpr (NT <= 25) {
    if (npts_pad % 128 == 0) {
      hipLaunchKernelGGL((k_tmat<NT, 8>), dim3(npts_pad / 128), dim3(512), 0, st, phi, ldphi, frag, tmat);
      return;
    }
  }
  hipLaunchKernelGGL((k_tmat<NT, 4>), dim3(nblocks(npts_pad, 64)), dim3(256), 0, st, phi, ldphi, frag, tmat);
}

// T = Phi P on npts_pad (a multiple of 64) rows, from full P fragments of nt tiles;
// with tmask, tiles outside a block's mask may be left unformed.
void launch_tmat_rows(int nt, hipStream_t st, long npts_pad, const double* phi, int ldphi, const double* frag,
                      double* tmat, const unsigned* tmask = nullptr) {
  switch (nt) {
#define DG_NT(N)                                                         \
  case N:                                                                \
    launch_tmat_nt<N>(st, npts_pad, phi, ldphi, frag, tmat, tmask); \
    break;
    DG_NT(1) DG_NT(2) DG_NT(3) DG_NT(4) DG_NT(5) DG_NT(6) DG_NT(7) DG_NT(8) DG_NT(9) DG_NT(10) DG_NT(11)
    DG_NT(12) DG_NT(13) DG_NT(14) DG_NT(15) DG_NT(16) DG_NT(17) DG_NT(18) DG_NT(19) DG_NT(20) DG_NT(21)
    DG_NT(22) DG_NT(23) DG_NT(24) DG_NT(25) DG_NT(26) DG_NT(27) DG_NT(28) DG_NT(29) DG_NT(30) DG_NT(31)
    DG_NT(32)
#undef DG_NT
    default:
      break;
  }
}

// T = Phi P for the force terms (P = the last density matrix, full fragments).
// The function tiles within reach of each 64-point block (k_tile_mask), on
// first use; none beyond 32 tiles (the masks are 32-bit).
int ensure_tile_mask(dg_handle* h) {
  if (h->adj_tmask || h->ntb > 32) return DG_OK;
  const long nblk = h->npts_pad / 64;
  DG_TRY(dalloc(&h->adj_tmask, nblk));
  return run(h, "adj_tile_mask", [&] {
    hipLaunchKernelGGL(k_tile_mask, dim3(nblk), dim3(256), 0, h->stream, h->atoms, h->nshell, h->sh_atom, h->sh_L,
                       h->sh_nprim, h->sh_off, h->sh_cgf, h->pr_alpha, h->npts, h->px, h->py, h->pz, h->adj_tmask);
  });
}

// T = Phi P for the last density matrix.  masked (the force terms): only the
// tiles k_force_basis reads are formed (k_tile_mask).
int launch_tmat(dg_handle* h, bool masked = false) {
  if (masked) DG_TRY(ensure_tile_mask(h));
  DG_TRY(run(h, "adj_pfrag", [&] {
    const long nf = 256L * h->ntb * h->ntb;
    hipLaunchKernelGGL(k_pfrag_full, dim3(nblocks(nf, 256)), dim3(256), 0, h->stream, h->last_P, h->nbf, h->ntb,
                       h->adj_frag);
  }));
  const unsigned* tmask = masked ? h->adj_tmask : nullptr;
  return run(h, "adj_tmat", [&] {
    launch_tmat_rows(h->ntb, h->stream, h->npts_pad, h->phi, h->ldphi, h->adj_frag, h->adj_T, tmask);
  });
}

int stage_density(dg_handle* h, const double* dP) {
  const long nfrag = 128L * h->ntb * (h->ntb + 1);
  if (h->ntb > 7)  // the persistent kernel forms its fragments itself
    DG_TRY(run(h, "pfrag", [&] {
      hipLaunchKernelGGL(k_pfrag, dim3(nblocks(nfrag, 256)), dim3(256), 0, h->stream, dP, h->nbf, h->ntb,
                         h->pfrag);
    }));
  DG_TRY(launch_density(h, dP));
  h->last_P = dP;
  h->corrected = false;
  h->has_density = true;
  h->has_hartree = h->has_xc = false;
  return DG_OK;
}

int stage_correct(dg_handle* h) {
  const int S = h->cell_split;
  if (S > 1) {
    DG_TRY(run(h, "cell_sum", [&] {
      hipLaunchKernelGGL(k_cell_sum_split, dim3(h->natoms * S), dim3(256), 0, h->stream, h->rho, h->w, h->cell, S,
                         h->qpart);
    }));
    DG_TRY(run(h, "scale", [&] {
      hipLaunchKernelGGL(k_scale_parts, dim3(nblocks(h->npts_pad, 256)), dim3(256), 0, h->stream, h->rho,
                         h->npts_pad, h->qpart, S, h->natoms, h->zsum, h->q, h->qc, h->scalars);
    }));
  } else {
    DG_TRY(run(h, "cell_sum", [&] {
      hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->rho, h->w, h->cell, h->q);
    }));
    DG_TRY(run(h, "scale", [&] {
      hipLaunchKernelGGL(k_scale, dim3(nblocks(h->npts_pad, 256)), dim3(256), 0, h->stream, h->rho,
                         h->npts_pad, h->q, h->natoms, h->zsum, h->qc, h->scalars);
    }));
  }
  h->corrected = true;
  return DG_OK;
}

// points per thread of the cross-potential kernel (a work-group pass covers 1024 kCrossQ slots)
constexpr int kCrossQ = 2;

template <int L>
void launch_cross_l(dg_handle* h, CrossArgs& a, dim3 grid) {
  hipLaunchKernelGGL((k_hartree_cross<L, kCrossQ>), grid, dim3(1024), h->cross_lds, h->stream, a);
}

template <int L>
hipError_t allow_cross_lds(size_t bytes) {
  return hipFuncSetAttribute((const void*)k_hartree_cross<L, kCrossQ>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// NT lm tiles per block: all of them up to lmax 8 (81 functions), else two
// groups; one per block when the (atom, shell tile) blocks would not fill the CUs.
void launch_rho_lm(dg_handle* h, int ti) {
  const int tiles = (h->nlm + 15) / 16;
  const int nt = (long)h->natoms * ti < h->cus ? 1 : (tiles <= 6 ? tiles : (tiles + 1) / 2);
  const dim3 grid(h->natoms * ti, (tiles + nt - 1) / nt), block(64 * kRhoWaves);
#define DG_RHO(N) \
  case N: \
    hipLaunchKernelGGL(k_rho_lm<N>, grid, block, 0, h->stream, h->rho, h->wb, h->leb, h->ylm, h->nrad, h->nang, \
                       h->nlm, h->rho_lm); \
    break;
  switch (nt) {
    DG_RHO(1) DG_RHO(2) DG_RHO(3) DG_RHO(4) DG_RHO(5) DG_RHO(6)
  }
#undef DG_RHO
}

hipError_t allow_cross_lds_l(int lmax, size_t bytes) {
  switch (lmax) {
#define DG_L(N) \
  case N:       \
    return allow_cross_lds<N>(bytes);
    DG_L(0) DG_L(1) DG_L(2) DG_L(3) DG_L(4) DG_L(5) DG_L(6) DG_L(7) DG_L(8) DG_L(9) DG_L(10) DG_L(11) DG_L(12)
#undef DG_L
    default:
      return hipErrorInvalidValue;
  }
}

void launch_cross(dg_handle* h, CrossArgs& a, dim3 grid) {
  switch (h->lmax) {
#define DG_L(N)                    \
  case N:                          \
    launch_cross_l<N>(h, a, grid); \
    break;
    DG_L(0) DG_L(1) DG_L(2) DG_L(3) DG_L(4) DG_L(5) DG_L(6) DG_L(7) DG_L(8) DG_L(9) DG_L(10) DG_L(11) DG_L(12)
#undef DG_L
    default:
      break;
  }
}

int stage_hartree(dg_handle* h) {
  const int ti = (h->nrad + 15) / 16;
  DG_TRY(run(h, "rho_lm", [&] { launch_rho_lm(h, ti); }));
  DG_TRY(run(h, "poisson", [&] {
    const int tcmax = (h->natoms * (2 * h->lmax + 1) + 15) / 16;
    const long ntiles = (long)(h->lmax + 1) * ti * tcmax * 2;
    hipLaunchKernelGGL(k_poisson, dim3(nblocks(ntiles, 4)), dim3(256), 0, h->stream, h->rho_lm, h->qc, h->rad,
                       h->minvT, h->KD, h->shc, h->natoms, h->nrad, h->lmax, tcmax, ntiles, h->U, h->tab);
  }));
  DG_TRY(run(h, "vcell", [&] {
    const long ntiles = (long)h->natoms * ti * ((h->nang + 15) / 16);
    hipLaunchKernelGGL(k_vcell, dim3(nblocks(ntiles, 4)), dim3(256), 0, h->stream, h->U, h->ylm, h->rad, h->nrad,
                       h->nang, h->nlm, ntiles, h->vcell);
  }));
  if (h->natoms > 1) {
    CrossArgs a;
    a.rad = h->rad;
    a.leb = h->leb;
    a.nang = h->nang;
    a.inv_nang = 1.0 / h->nang;
    a.inv_nrad = 1.0 / h->nrad;
    a.atoms = h->atoms;
    a.knots = h->knots;
    a.invh = h->invh;
    a.tab = h->tab;
    a.items = h->items;
    a.perm = h->perm;
    a.plen = h->plen;
    a.pkey = h->pkey;
    a.vpart = h->vpart;
    a.vbase = h->vbase;
    a.nitems = h->nitems;
    a.nrad = h->nrad;
    a.nlm = h->nlm;
    dim3 grid(h->cross_wgs);
    for (size_t ci = 0; ci < h->lchunks.size(); ++ci) {
      a.l0 = h->lchunks[ci].first;
      a.l1 = h->lchunks[ci].second;
      a.accumulate = ci > 0;
      DG_TRY(run(h, "hartree_cross", [&] { launch_cross(h, a, grid); }));
    }
  }
  DG_TRY(run(h, "vreduce", [&] {
    hipLaunchKernelGGL(k_vreduce, dim3(nblocks(h->npts_pad, 256)), dim3(256), 0, h->stream, h->vpart,
                       h->vbase, h->vcell, h->w, h->rho, h->npts, h->npts_pad, h->cell, h->natoms,
                       h->atom_mol, h->mol_atom0, h->V, h->uJ, h->ej_part);
  }));
  return DG_OK;
}

int stage_lda(dg_handle* h) {
  return run(h, "lda", [&] {
    hipLaunchKernelGGL(k_lda<0>, dim3((unsigned)h->nlda_blocks), dim3(256), 0, h->stream, h->rho, h->w,
                       h->npts_pad, h->vwn, h->xfac, h->uxc, h->exd, h->exc_part, nullptr, 1, h->natoms, 0.0,
                       nullptr, nullptr, nullptr);
  });
}

// The renormalisation and the functional in two launches: the cell sums, then
// k_lda scaling rho as it goes (the device build; every scalar and every value of
// rho, q and qc is the one stage_correct + stage_lda give).
int stage_correct_lda(dg_handle* h) {
  const int S = h->cell_split;
  if (S > 1) {
    DG_TRY(run(h, "cell_sum", [&] {
      hipLaunchKernelGGL(k_cell_sum_split, dim3(h->natoms * S), dim3(256), 0, h->stream, h->rho, h->w, h->cell, S,
                         h->qpart);
    }));
  } else {
    DG_TRY(run(h, "cell_sum", [&] {
      hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->rho, h->w, h->cell, h->q);
    }));
  }
  h->corrected = true;
  return run(h, "lda", [&] {
    if (S > 1)
      hipLaunchKernelGGL(k_lda<1>, dim3((unsigned)h->nlda_blocks), dim3(256), 0, h->stream, h->rho, h->w,
                         h->npts_pad, h->vwn, h->xfac, h->uxc, h->exd, h->exc_part, h->qpart, S, h->natoms,
                         h->zsum, h->q, h->qc, h->scalars);
    else
      hipLaunchKernelGGL(k_lda<2>, dim3((unsigned)h->nlda_blocks), dim3(256), 0, h->stream, h->rho, h->w,
                         h->npts_pad, h->vwn, h->xfac, h->uxc, h->exd, h->exc_part, h->q, 1, h->natoms, h->zsum,
                         nullptr, h->qc, h->scalars);
  });
}

template <int N>
void launch_contract_small_nt(dg_handle* h, const double* u0, const double* u1, int nout) {
  const long nblk = h->npts_pad / 64;
  auto go = [&](auto kern2, auto kern1) {
    if (nout == 2)
      hipLaunchKernelGGL(kern2, dim3(h->nsplit_small), dim3(256), 0, h->stream, h->phi, h->ldphi, u0, u1, nblk,
                         h->ngroups, h->cpart);
    else
      hipLaunchKernelGGL(kern1, dim3(h->nsplit_small), dim3(256), 0, h->stream, h->phi, h->ldphi, u0, u0, nblk,
                         h->ngroups, h->cpart);
  };
  // the VALU edge path when the last of N >= 2 tiles holds at most 4 functions
  switch (N >= 2 ? h->nbf - 16 * (N - 1) : 0) {
    case 1: go(k_contract_small<N, 2, 1>, k_contract_small<N, 1, 1>); break;
    case 2: go(k_contract_small<N, 2, 2>, k_contract_small<N, 1, 2>); break;
    case 3: go(k_contract_small<N, 2, 3>, k_contract_small<N, 1, 3>); break;
    case 4: go(k_contract_small<N, 2, 4>, k_contract_small<N, 1, 4>); break;
    default: go(k_contract_small<N, 2, 0>, k_contract_small<N, 1, 0>); break;
  }
}

template <int N>
void launch_contract_small_f32_nt(dg_handle* h, const double* u0, const double* u1, int nout) {
  const long nblk = h->npts_pad / 64;
  auto go = [&](auto kern2, auto kern1) {
    if (nout == 2)
      hipLaunchKernelGGL(kern2, dim3(h->nsplit_small), dim3(256), 0, h->stream, h->phf, h->ldphi, u0, u1, nblk,
                         h->ngroups, h->cpart);
    else
      hipLaunchKernelGGL(kern1, dim3(h->nsplit_small), dim3(256), 0, h->stream, h->phf, h->ldphi, u0, u0, nblk,
                         h->ngroups, h->cpart);
  };
  switch (N >= 2 ? h->nbf - 16 * (N - 1) : 0) {
    case 1: go(k_contract_small_f32<N, 2, 1>, k_contract_small_f32<N, 1, 1>); break;
    case 2: go(k_contract_small_f32<N, 2, 2>, k_contract_small_f32<N, 1, 2>); break;
    case 3: go(k_contract_small_f32<N, 2, 3>, k_contract_small_f32<N, 1, 3>); break;
    case 4: go(k_contract_small_f32<N, 2, 4>, k_contract_small_f32<N, 1, 4>); break;
    default: go(k_contract_small_f32<N, 2, 0>, k_contract_small_f32<N, 1, 0>); break;
  }
}

void launch_contract_small(dg_handle* h, const double* u0, const double* u1, int nout) {
  if (h->precision == DG_PREC_MIXED) {
    switch (h->ntb) {
      case 1: launch_contract_small_f32_nt<1>(h, u0, u1, nout); break;
      case 2: launch_contract_small_f32_nt<2>(h, u0, u1, nout); break;
      case 3: launch_contract_small_f32_nt<3>(h, u0, u1, nout); break;
      case 4: launch_contract_small_f32_nt<4>(h, u0, u1, nout); break;
      case 5: launch_contract_small_f32_nt<5>(h, u0, u1, nout); break;
      default: break;
    }
    return;
  }
  switch (h->ntb) {
    case 1: launch_contract_small_nt<1>(h, u0, u1, nout); break;
    case 2: launch_contract_small_nt<2>(h, u0, u1, nout); break;
    case 3: launch_contract_small_nt<3>(h, u0, u1, nout); break;
    case 4: launch_contract_small_nt<4>(h, u0, u1, nout); break;
    case 5: launch_contract_small_nt<5>(h, u0, u1, nout); break;
    default: break;
  }
}

// two outputs on 4 x 2 tile blocks, one output on 4 x 4 (the same 16 MFMAs per step);
// list/loff (optional): point-block lists per tile block to use instead of the
// handle's (the screened kernel then runs whatever the handle's screening)
template <typename T, int NOUT>
void launch_contract2d(dg_handle* h, const T* phi, const double* u0, const double* u1, const int* list = nullptr,
                       const int* loff = nullptr) {
  constexpr int NJ = NOUT == 2 ? 2 : 4;
  const TileBlocks& t = h->tb[NOUT == 2 ? 0 : 1];
  const int* bl = list ? list : t.list;
  const int* bo = list ? loff : t.loff;
  const bool dense = h->screen == 0.0 && !list;
  const dim3 grid(round_up((long)t.ncb * t.nsplit, kXcds));
  if constexpr (sizeof(T) == 8) {
    constexpr int D = NJ == 2 ? 4 : 3;  // ring slots: 3 work-groups per CU within 160 KB of LDS
    if (dense)
      hipLaunchKernelGGL((k_contract2d_lds<NOUT, NJ, D, true>), grid, dim3(256), 0, h->stream, phi, h->ldphi, u0, u1,
                         t.cb, bl, bo, t.order, 4 * t.ncb, t.nsplit, h->cpart);
    else
      hipLaunchKernelGGL((k_contract2d_lds<NOUT, NJ, D, false>), grid, dim3(256), 0, h->stream, phi, h->ldphi, u0,
                         u1, t.cb, bl, bo, t.order, 4 * t.ncb, t.nsplit, h->cpart);
  } else {
    constexpr int D = NJ == 2 ? 12 : 9;  // ring slots: 2 work-groups per CU within 160 KB of LDS
    if (dense)
      hipLaunchKernelGGL((k_contract2d_lds_f32<NOUT, NJ, D, true>), grid, dim3(256), 0, h->stream, phi, h->ldphi,
                         u0, u1, t.cb, bl, bo, t.order, 4 * t.ncb, t.nsplit, h->cpart);
    else
      hipLaunchKernelGGL((k_contract2d_lds_f32<NOUT, NJ, D, false>), grid, dim3(256), 0, h->stream, phi, h->ldphi,
                         u0, u1, t.cb, bl, bo, t.order, 4 * t.ncb, t.nsplit, h->cpart);
  }
}

// sa (optional): the build's scalars, evaluated by one more work-group of the reduction
// list/loff (optional, one output, 2-D path): point-block lists per tile block
int stage_contract(dg_handle* h, const double* u0, const double* u1, double* out0, double* out1,
                   const ScalarArgs* sa = nullptr, const int* list = nullptr, const int* loff = nullptr) {
  const int nout = u1 ? 2 : 1;
  const bool mx = h->precision == DG_PREC_MIXED;
  // up to 80 functions the dense all-tiles kernel (FP64, or FP32 in mixed mode)
  // reads each amplitude once and is faster than skipping tiles, so it also
  // serves screened handles
  const bool small = h->ntb <= 5;
  DG_TRY(run(h, "contract", [&] {
    if (small) {
      launch_contract_small(h, u0, u1, nout);
    } else if (mx) {
      if (nout == 2)
        launch_contract2d<float, 2>(h, h->phf, u0, u1);
      else
        launch_contract2d<float, 1>(h, h->phf, u0, u0);
    } else {
      if (nout == 2)
        launch_contract2d<double, 2>(h, h->phi, u0, u1);
      else
        launch_contract2d<double, 1>(h, h->phi, u0, u0, list, loff);
    }
  }));
  const bool two_d = !small;
  const TileBlocks& tb = h->tb[nout == 2 ? 0 : 1];
  const int ng = two_d ? 4 * tb.ncb : h->ngroups;
  ScalarArgs none{};
  const unsigned nb = nblocks((long)ng * nout * 1024, 16) + (sa ? 1 : 0);
  return run(h, "contract_reduce", [&] {
    hipLaunchKernelGGL(k_contract_reduce, dim3(nb), dim3(256), 0, h->stream, h->cpart,
                       small ? h->nsplit_small : tb.nsplit, ng, nout, two_d ? tb.groups : h->blocks, h->nbf,
                       out0, out1, sa ? *sa : none);
  });
}

ScalarArgs scalar_args(dg_handle* h, bool with_ej, double* out2) {
  ScalarArgs a;
  a.exc_part = h->exc_part;
  a.nparts = h->nlda_blocks;
  a.ej_part = with_ej ? h->ej_part : nullptr;
  a.nej = h->npts_pad / 256;
  a.q = h->qc;
  a.natoms = h->natoms;
  a.zsum = h->zsum;
  a.scalars = h->scalars;
  a.out2 = out2;
  return a;
}

int stage_scalars(dg_handle* h, bool with_ej) {
  return run(h, "scalars", [&] {
    hipLaunchKernelGGL(k_scalars, dim3(1), dim3(256), 0, h->stream, scalar_args(h, with_ej, nullptr));
  });
}

int sync(dg_handle* h) {
  HIP_TRY(hipStreamSynchronize(h->stream));
  return DG_OK;
}

int copy_out(dg_handle* h, double* dst, const double* src, size_t n) {
  HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}

// T = Phi P buffers (the adjoint's force terms and the density gradients), on first use.
int ensure_tmat_bufs(dg_handle* h) {
  if (!h->adj_T) DG_TRY(dalloc(&h->adj_T, (size_t)h->npts_pad * h->ldphi));
  if (!h->adj_frag) DG_TRY(dalloc(&h->adj_frag, (size_t)256 * h->ntb * h->ntb));
  return DG_OK;
}

// ---------------------------------------------------------------------------
// Density and density gradient at arbitrary points (dg_density_at_points and
// its variant on the static scratch array)
// ---------------------------------------------------------------------------
constexpr int kPtsMaxShell = 512, kPtsMaxPrim = 4096, kPtsMaxAtoms = 1024;
constexpr long kPointChunk = 1L << 16;  // points per device pass at most

// Host tables of the basis (fixed capacity, so that a caller-provided object
// needs no heap): CGF offset per shell, exponents and c N products, atoms.
struct PointTables {
  int nshell = 0, na = 0, nbf = 0, nprim = 0, nt = 0, ld = 0;
  int ishell[5 * kPtsMaxShell];  // atom, L, nprim, offset, cgf offset per shell (five blocks)
  double alpha[kPtsMaxPrim], cn0[kPtsMaxPrim], cn1[kPtsMaxPrim];
  double atoms[4 * kPtsMaxAtoms];
};

// Device buffers of one evaluation; chunk = points per pass.
struct PointWork {
  double *atoms = nullptr, *alpha = nullptr, *cn0 = nullptr, *cn1 = nullptr, *P = nullptr, *frag = nullptr;
  double *xyz = nullptr, *pxyz = nullptr, *phi = nullptr, *T = nullptr, *rho = nullptr, *grad = nullptr;
  int* ishell = nullptr;
  long chunk = 0;
};

int point_check(const dg_molecule* mol, const dg_basis* basis, long npts, const double* xyz, const double* P,
                const char* who) {
  if (!mol || !basis || !P || (npts > 0 && !xyz)) return fail(DG_ERR_INVALID, std::string(who) + ": null argument");
  if (npts < 0) return fail(DG_ERR_INVALID, std::string(who) + ": negative point count");
  if (mol->natoms < 1 || !mol->xyz) return fail(DG_ERR_INVALID, std::string(who) + ": empty molecule");
  if (basis->nshell < 1 || !basis->atom || !basis->L || !basis->nprim || !basis->offset || !basis->alpha ||
      !basis->coef)
    return fail(DG_ERR_INVALID, std::string(who) + ": empty basis");
  return DG_OK;
}

int point_tables(const dg_molecule* mol, const dg_basis* basis, PointTables& t, const char* who) {
  // (messages are built only on failure: the success path allocates nothing)
  auto w = [who](const char* msg) { return std::string(who) + msg; };
  t.nshell = basis->nshell;
  t.na = mol->natoms;
  if (t.nshell > kPtsMaxShell) return fail(DG_ERR_UNSUPPORTED, w(": more than 512 shells"));
  if (t.na > kPtsMaxAtoms) return fail(DG_ERR_UNSUPPORTED, w(": more than 1024 atoms"));
  int nbf = 0, nprim = 0;
  for (int s = 0; s < t.nshell; ++s) {
    const int L = basis->L[s];
    if (L < 0 || L > 2) return fail(DG_ERR_INVALID, w(": shell L must be 0, 1 or 2"));
    if (basis->atom[s] < 0 || basis->atom[s] >= t.na) return fail(DG_ERR_INVALID, w(": shell atom out of range"));
    if (basis->nprim[s] < 1 || basis->offset[s] < 0) return fail(DG_ERR_INVALID, w(": bad primitive range"));
    if (basis->offset[s] + basis->nprim[s] > kPtsMaxPrim)
      return fail(DG_ERR_UNSUPPORTED, w(": more than 4096 primitives"));
    t.ishell[s] = basis->atom[s];
    t.ishell[kPtsMaxShell + s] = L;
    t.ishell[2 * kPtsMaxShell + s] = basis->nprim[s];
    t.ishell[3 * kPtsMaxShell + s] = basis->offset[s];
    t.ishell[4 * kPtsMaxShell + s] = nbf;
    nbf += (L + 1) * (L + 2) / 2;
    nprim = std::max(nprim, basis->offset[s] + basis->nprim[s]);
  }
  t.nbf = nbf;
  t.nprim = nprim;
  t.nt = (nbf + 15) / 16;
  t.ld = (int)round_up(nbf, 32);
  if (t.nt > 32) return fail(DG_ERR_UNSUPPORTED, w(": more than 512 basis functions"));
  for (int i = 0; i < nprim; ++i) {
    t.alpha[i] = 1.0;
    t.cn0[i] = t.cn1[i] = 0.0;
  }
  for (int s = 0; s < t.nshell; ++s)
    for (int k = 0; k < basis->nprim[s]; ++k) {
      const int i = basis->offset[s] + k, L = basis->L[s];
      if (!(basis->alpha[i] > 0.0)) return fail(DG_ERR_INVALID, w(": primitive exponent must be positive"));
      t.alpha[i] = basis->alpha[i];
      t.cn0[i] = basis->coef[i] * dg::gto_norm(t.alpha[i], L, 0, 0);
      t.cn1[i] = basis->coef[i] * dg::gto_norm(t.alpha[i], L == 2 ? 1 : L, L == 2 ? 1 : 0, 0);
    }
  for (int a = 0; a < t.na; ++a) {
    for (int d = 0; d < 3; ++d) t.atoms[4 * a + d] = mol->xyz[3 * a + d];
    t.atoms[4 * a + 3] = mol->Z ? mol->Z[a] : 0;
  }
  return DG_OK;
}

// The evaluation on the buffers of w, in passes of w.chunk points on stream st:
// amplitudes (k_phi), T = Phi P (k_tmat), rho and grad rho (k_density_grad).
int point_run(const PointTables& t, const dg_basis* basis, const double* P, long npts, const double* xyz, double* rho,
              double* grad, const PointWork& w, hipStream_t st) {
  (void)basis;
  const int ns = t.nshell, K = kPtsMaxShell;
  auto up = [&](void* d, const void* src, size_t bytes) {
    return hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, st);
  };
  HIP_TRY(up(w.atoms, t.atoms, 4 * (size_t)t.na * 8));
  HIP_TRY(up(w.alpha, t.alpha, (size_t)t.nprim * 8));
  HIP_TRY(up(w.cn0, t.cn0, (size_t)t.nprim * 8));
  HIP_TRY(up(w.cn1, t.cn1, (size_t)t.nprim * 8));
  for (int b = 0; b < 5; ++b) HIP_TRY(up(w.ishell + (size_t)b * ns, t.ishell + (size_t)b * K, (size_t)ns * 4));
  HIP_TRY(up(w.P, P, (size_t)t.nbf * t.nbf * 8));
  const int *sa = w.ishell, *sL = sa + ns, *snp = sL + ns, *soff = snp + ns, *scgf = soff + ns;
  const long C = w.chunk;
  double *px = w.pxyz, *py = px + C, *pz = py + C;
  // padding rows and columns of Phi stay zero (a fill kernel: the buffer may hold an earlier call's values)
  hipLaunchKernelGGL(k_fill, dim3(nblocks(C * t.ld, 256)), dim3(256), 0, st, w.phi, C * t.ld, 0.0);
  hipLaunchKernelGGL(k_pfrag_full, dim3(nblocks(256L * t.nt * t.nt, 256)), dim3(256), 0, st, w.P, t.nbf, t.nt,
                     w.frag);
  for (long p0 = 0; p0 < npts; p0 += C) {
    const long n = std::min(C, npts - p0);
    HIP_TRY(up(w.xyz, xyz + 3 * p0, (size_t)n * 3 * 8));
    hipLaunchKernelGGL(k_split_xyz, dim3(nblocks(n, 256)), dim3(256), 0, st, w.xyz, n, px, py, pz);
    hipLaunchKernelGGL(k_phi, dim3(nblocks(n * ns, 256)), dim3(256), 0, st, w.atoms, ns, sa, sL, snp, soff, scgf,
                       w.alpha, w.cn0, w.cn1, n, px, py, pz, t.ld, w.phi);
    launch_tmat_rows(t.nt, st, round_up(n, 64), w.phi, t.ld, w.frag, w.T);
    hipLaunchKernelGGL(k_density_grad, dim3(nblocks(n, 256)), dim3(256), 0, st, w.atoms, ns, sa, sL, snp, soff, scgf,
                       w.alpha, w.cn0, w.cn1, 0L, n, px, py, pz, t.ld, w.T, nullptr, 0.0, w.rho, w.grad);
    HIP_TRY(hipGetLastError());
    if (rho) HIP_TRY(hipMemcpyAsync(rho + p0, w.rho, n * 8, hipMemcpyDeviceToHost, st));
    if (grad) HIP_TRY(hipMemcpyAsync(grad + 3 * p0, w.grad, 3 * n * 8, hipMemcpyDeviceToHost, st));
    // the next pass overwrites the position and result buffers
    HIP_TRY(hipStreamSynchronize(st));
  }
  return DG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

// Batch mode (several molecules in one handle; defined in batch.inl).
// Atoms, shells and primitives of the molecules are concatenated molecule by
// molecule; mol_atom0 / mol_shell0 hold each molecule's first atom / shell.
struct BatchSpec {
  int nmol = 0;
  std::vector<int> mol_atom0, mol_shell0;
};
std::vector<int> cross_items(const dg_handle* h, int len, long* perm_len = nullptr);
int batch_alloc(dg_handle* h, const BatchSpec* bs, const std::vector<int>& amol);
int batch_setup(dg_handle* h);

}  // namespace

extern "C" {

const char* dg_last_error(void) { return g_last_error.c_str(); }
int dg_version(void) { return 100; }
int dg_lebedev_available(int npoints) { return dg::lebedev_rule(npoints) != nullptr; }

// Shared by dg_create (bs = NULL) and dg_batch_create (bs lists the molecules
// whose atoms and shells are concatenated in mol and basis).
static int create_impl(const dg_molecule* mol, const dg_basis* basis, const dg_grid_params* prm,
                       const BatchSpec* bs, dg_handle** out) {
  g_last_error.clear();
  if (!mol || !basis || !prm || !out) return fail(DG_ERR_INVALID, "dg_create: null argument");
  *out = nullptr;
  if (mol->natoms < 1 || !mol->Z || !mol->xyz) return fail(DG_ERR_INVALID, "dg_create: empty molecule");
  if (basis->nshell < 1 || !basis->atom || !basis->L || !basis->nprim || !basis->offset || !basis->alpha ||
      !basis->coef)
    return fail(DG_ERR_INVALID, "dg_create: empty basis");
  if (prm->n_radial < 5) return fail(DG_ERR_INVALID, "dg_create: n_radial must be >= 5");
  if (prm->lmax < 0 || prm->lmax > kMaxL)
    return fail(DG_ERR_INVALID, "dg_create: lmax must be in 0.." + std::to_string(kMaxL));
  const double* lebt = dg::lebedev_rule(prm->n_angular);
  if (!lebt) return fail(DG_ERR_INVALID, "dg_create: no Lebedev rule with " + std::to_string(prm->n_angular) + " points");
  if (prm->precision != DG_PREC_FP64 && prm->precision != DG_PREC_MIXED)
    return fail(DG_ERR_INVALID, "dg_create: unknown precision mode");
  for (int a = 0; a < mol->natoms; ++a)
    if (mol->Z[a] < 1 || mol->Z[a] > 18) return fail(DG_ERR_INVALID, "dg_create: Z outside 1..18");
  std::vector<int> amol(mol->natoms, 0);
  if (bs)
    for (int m = 0; m < bs->nmol; ++m)
      for (int a = bs->mol_atom0[m]; a < bs->mol_atom0[m + 1]; ++a) amol[a] = m;
  for (int i = 0; i < mol->natoms; ++i)
    for (int j = i + 1; j < mol->natoms && amol[j] == amol[i]; ++j) {
      const double dx = mol->xyz[3 * i] - mol->xyz[3 * j], dy = mol->xyz[3 * i + 1] - mol->xyz[3 * j + 1],
                   dz = mol->xyz[3 * i + 2] - mol->xyz[3 * j + 2];
      if (dx * dx + dy * dy + dz * dz == 0.0) return fail(DG_ERR_INVALID, "dg_create: two atoms coincide");
    }
  if (!bs && mol->natoms > 1024) return fail(DG_ERR_UNSUPPORTED, "dg_create: more than 1024 atoms");
  if ((long)mol->natoms * prm->n_radial * prm->n_angular > (1L << 31) - 4096)
    return fail(DG_ERR_UNSUPPORTED, "dg_create: more than 2^31 grid points");
  if (bs && prm->precision != DG_PREC_FP64)
    return fail(DG_ERR_UNSUPPORTED, "dg_batch_create: batches run in FP64");
  if (!(prm->screen >= 0.0) || !std::isfinite(prm->screen))
    return fail(DG_ERR_INVALID, "dg_create: screen must be a finite value >= 0");
  if (bs && prm->screen != 0.0) return fail(DG_ERR_UNSUPPORTED, "dg_batch_create: batches do not screen");

  auto h = new dg_handle();
  std::unique_ptr<dg_handle> guard(h);
  h->natoms = mol->natoms;
  h->nrad = prm->n_radial;
  h->nang = prm->n_angular;
  h->lmax = prm->lmax;
  h->nlm = (prm->lmax + 1) * (prm->lmax + 1);
  h->precision = prm->precision;
  h->screen = prm->screen;
  h->cell = (long)h->nrad * h->nang;
  h->npts = h->cell * h->natoms;
  h->npts_pad = round_up(h->npts, 256);

  // shells -> CGF offsets, primitive tables
  h->nshell = basis->nshell;
  std::vector<int> cgf_off(h->nshell);
  int nbf = 0, nprim_tot = 0, cur_mol = 0, max_nbf = 0;
  for (int s = 0; s < h->nshell; ++s) {
    const int L = basis->L[s];
    if (L < 0 || L > 2) return fail(DG_ERR_INVALID, "dg_create: shell L must be 0, 1 or 2");
    if (basis->atom[s] < 0 || basis->atom[s] >= mol->natoms) return fail(DG_ERR_INVALID, "dg_create: shell atom out of range");
    if (basis->nprim[s] < 1 || basis->offset[s] < 0) return fail(DG_ERR_INVALID, "dg_create: bad primitive range");
    if (bs) {  // CGF offsets restart with every molecule
      while (cur_mol < bs->nmol && s >= bs->mol_shell0[cur_mol + 1]) {
        h->h_mol_nbf.push_back(nbf);
        nbf = 0;
        ++cur_mol;
      }
      if (amol[basis->atom[s]] != cur_mol) return fail(DG_ERR_INVALID, "dg_batch_create: shell atom outside its molecule");
    }
    cgf_off[s] = nbf;
    nbf += (L + 1) * (L + 2) / 2;
    max_nbf = std::max(max_nbf, nbf);
    nprim_tot = std::max(nprim_tot, basis->offset[s] + basis->nprim[s]);
  }
  if (bs) {
    h->h_mol_nbf.push_back(nbf);
    while ((int)h->h_mol_nbf.size() < bs->nmol) h->h_mol_nbf.push_back(0);
    for (int m = 0; m < bs->nmol; ++m)
      if (h->h_mol_nbf[m] < 1) return fail(DG_ERR_INVALID, "dg_batch_create: a molecule has no basis functions");
    if (max_nbf > 32) return fail(DG_ERR_UNSUPPORTED, "dg_batch_create: more than 32 basis functions in a molecule");
    nbf = 32;  // batch Phi rows are 32 wide; each molecule uses its first nbf_m columns
  }
  h->nbf = nbf;
  h->ldphi = (int)round_up(nbf, 32);
  h->ntb = (nbf + 15) / 16;
  if (h->ntb > 32) return fail(DG_ERR_UNSUPPORTED, "dg_create: more than 512 basis functions");
  std::vector<double> alpha(nprim_tot), cn0(nprim_tot), cn1(nprim_tot);
  for (int s = 0; s < h->nshell; ++s) {
    const int L = basis->L[s];
    for (int k = 0; k < basis->nprim[s]; ++k) {
      const int i = basis->offset[s] + k;
      const double a = basis->alpha[i], c = basis->coef[i];
      if (!(a > 0.0)) return fail(DG_ERR_INVALID, "dg_create: primitive exponent must be positive");
      alpha[i] = a;
      cn0[i] = c * dg::gto_norm(a, L, 0, 0);
      cn1[i] = c * dg::gto_norm(a, L == 2 ? 1 : L, L == 2 ? 1 : 0, 0);
    }
  }
  std::vector<double> atoms(4 * h->natoms), invR((size_t)h->natoms * h->natoms, 0.0);
  for (int a = 0; a < h->natoms; ++a) {
    for (int d = 0; d < 3; ++d) atoms[4 * a + d] = mol->xyz[3 * a + d];
    atoms[4 * a + 3] = mol->Z[a];
    h->zsum += mol->Z[a];
  }
  for (int n = 0; n < h->natoms; ++n)
    for (int m = 0; m < h->natoms; ++m)
      if (n != m) {
        const double dx = atoms[4 * m] - atoms[4 * n], dy = atoms[4 * m + 1] - atoms[4 * n + 1],
                     dz = atoms[4 * m + 2] - atoms[4 * n + 2];
        invR[(size_t)n * h->natoms + m] = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
      }

  // radial grid, angular tables
  std::vector<double> rad, wrad;
  dg::radial_grid(h->nrad, rad, wrad);
  std::vector<double> ylm;
  dg::ylm_table(h->nang, lebt, h->lmax, ylm);
  // Poisson inverses, rows 1..N, stored [l][k][i]
  const int n2 = h->nrad + 2;
  std::vector<double> minvT((size_t)(h->lmax + 1) * n2 * h->nrad);
  for (int l = 0; l <= h->lmax; ++l) {
    std::vector<double> M;
    dg::poisson_matrix(h->nrad, rad.data(), l, M);
    if (!dg::invert_matrix(n2, M)) return fail(DG_ERR_INVALID, "dg_create: singular Poisson matrix");
    for (int k = 0; k < n2; ++k)
      for (int i = 0; i < h->nrad; ++i) minvT[((size_t)l * n2 + k) * h->nrad + i] = M[(size_t)(i + 1) * n2 + k];
  }
  // spline over the ascending knots: slopes D = S y, then KD_l = S KU_l (knots reversed)
  std::vector<double> knots(h->nrad), invh(h->nrad, 0.0);
  for (int t = 0; t < h->nrad; ++t) knots[t] = rad[h->nrad - 1 - t];
  for (int t = 1; t < h->nrad; ++t) invh[t] = 1.0 / (knots[t] - knots[t - 1]);
  const std::vector<double> Sspl = dg::spline_slope_matrix(knots);  // [t][t'] row-major
  std::vector<double> KD(minvT.size(), 0.0);
  for (int l = 0; l <= h->lmax; ++l)
    for (int k = 0; k < n2; ++k) {
      const double* ku = &minvT[((size_t)l * n2 + k) * h->nrad];  // ku[i], i = radial node
      double* kd = &KD[((size_t)l * n2 + k) * h->nrad];
      for (int t = 0; t < h->nrad; ++t) {
        double acc = 0.0;
        for (int tp = 0; tp < h->nrad; ++tp) acc += Sspl[(size_t)t * h->nrad + tp] * ku[h->nrad - 1 - tp];
        kd[t] = acc;
      }
    }
  // spherical-harmonic recurrence constants per (l, m >= 0): {N_lm c_lm (the table
  // scale for the cross kernel's scaled recurrence), (2l-1)/(l-m), (l+m-1)/(l-m)}
  std::vector<double> shc(3 * (kMaxL + 1) * (kMaxL + 2) / 2, 0.0);
  for (int l = 0; l <= kMaxL; ++l)
    for (int m = 0; m <= l; ++m) {
      const int ci = 3 * (l * (l + 1) / 2 + m);
      shc[ci] = dg::sh_prefactor(l, m) * legendre_scale(l, m);
      if (l >= m + 2) {
        shc[ci + 1] = (double)(2 * l - 1) / (double)(l - m);
        shc[ci + 2] = (double)(l + m - 1) / (double)(l - m);
      }
    }
  // LDA constants
  h->xfac = -2.25 * (2.0 / 3.0) * std::pow(3.0 / 4.0 / dg::kPi, 1.0 / 3.0);
  {
    const double A = 0.0310907, x0 = -0.10498, b = 3.72744, c = 12.9352;
    const double Q = std::sqrt(4.0 * c - b * b), X0 = x0 * x0 + b * x0 + c;
    h->vwn = VwnConst{A, x0, b, c, Q, X0, (2.0 * b / Q) * (1.0 - x0 * (2.0 * x0 + b) / X0), b * x0 / X0};
  }
  // contraction: groups of up to 4 upper tiles of one tile row, split-K over points
  std::vector<int> blocks;
  for (int I = 0; I < h->ntb; ++I)
    for (int J0 = I; J0 < h->ntb; J0 += 4) {
      blocks.push_back(I);
      blocks.push_back(J0);
      blocks.push_back(std::min(4, h->ntb - J0));
    }
  h->ngroups = (int)blocks.size() / 3;
  std::vector<int> cbs[2], cbg[2];
  for (int q = 0; q < 2; ++q) {
    const int nj = q ? 4 : 2;
    for (int I0 = 0; I0 < h->ntb; I0 += 4)
      for (int J0 = I0; J0 < h->ntb; J0 += nj) {
        const int nI = std::min(4, h->ntb - I0), nJ = std::min(nj, h->ntb - J0);
        cbs[q].insert(cbs[q].end(), {I0, nI, J0, nJ});
        for (int r = 0; r < 4; ++r) cbg[q].insert(cbg[q].end(), {I0 + r, J0, r < nI ? nJ : 0});
      }
    h->tb[q].ncb = (int)cbs[q].size() / 4;
  }
  h->nlda_blocks = h->npts_pad / 256;
  // cross kernel: l ranges whose spline table fits in LDS beside an item's potentials
  {
    const size_t limit = 150 * 1024;
    auto bytes = [&](int l0, int l1) {
      return (size_t)(2 * h->nrad + (size_t)h->nrad * (l1 * l1 - l0 * l0) * 2 + kMaxItem) * sizeof(double);
    };
    int l0 = 0;
    size_t maxb = 0;
    while (l0 <= h->lmax) {
      int l1 = l0 + 1;
      if (bytes(l0, l1) > limit)
        return fail(DG_ERR_UNSUPPORTED, "dg_create: radial grid too large for the Hartree kernel");
      while (l1 <= h->lmax && bytes(l0, l1 + 1) <= limit) ++l1;
      h->lchunks.push_back({l0, l1});
      maxb = std::max(maxb, bytes(l0, l1));
      l0 = l1;
    }
    h->cross_lds = maxb;
  }
  // cross-potential work list: point ranges of at most 4096 (two work-group passes)
  // outside each cell; fewer, longer items spend less of the kernel between items
  std::vector<long> vbase(h->natoms);
  long vtotal = 0;
  if (bs) {
    h->nmol = bs->nmol;
    h->h_mol_atom0 = bs->mol_atom0;
    for (int k = 0; k < h->natoms; ++k) {
      const int m = amol[k];
      const long p0 = (long)bs->mol_atom0[m] * h->cell, p1 = (long)bs->mol_atom0[m + 1] * h->cell;
      vbase[k] = vtotal - p0;
      vtotal += p1 - p0;
    }
  } else {
    for (int k = 0; k < h->natoms; ++k) vbase[k] = (long)k * h->npts_pad;
    vtotal = (long)h->natoms * h->npts_pad;
  }
  std::vector<int> items = cross_items(h, 2048 * kCrossQ, &h->npairs);
  h->nitems = (int)items.size() / 4;
  h->items_cap = std::max<size_t>(items.size(), 4);
  if (bs) {  // the batch autotuner may split the items down to 1024 points
    long cap = 0;
    h->items_cap = std::max(h->items_cap, cross_items(h, 1024, &cap).size());
    h->npairs = std::max(h->npairs, cap);
  }
  // host-side checks and tables are done: from here on the device is used
  HIP_TRY(hipGetDevice(&h->device));
  HIP_TRY(hipStreamCreate(&h->stream));
  {
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    h->cus = cus;
    // small contraction: exactly one round of resident work-groups (two per CU);
    // a partial second round measured slower, and so did fewer splits
    h->nsplit_small = (int)std::max(1L, std::min(2L * cus, h->npts_pad / 256));
    const size_t lds = h->cross_lds;
    const int per_cu = std::max(1, (int)((160 * 1024) / std::max<size_t>(lds, 1)));
    h->cross_wgs = std::max(1, std::min(h->nitems, cus * std::min(per_cu, 2)));
    // 2-D contraction: splits per tile block such that the work-groups make about
    // three rounds of the resident slots (3 FP64 / 2 FP32 work-groups per CU); with
    // the longest blocks first, the short ones even out the last round (C30H62: one
    // round 14.3 ms, two 12.8, three 12.35, four 12.4)
    const long slots = (long)cus * (h->precision == DG_PREC_MIXED ? 2 : 3);
    for (int q = 0; q < 2; ++q) {
      TileBlocks& t = h->tb[q];
      const std::vector<int>& c = cbs[q];
      t.nsplit = (int)std::max(1L, std::min(std::min(512L, 3 * slots / std::max(t.ncb, 1)), h->npts_pad / 64));
      // virtual work-group order of the 2-D contraction: split-major, each XCD's range
      // re-ordered by decreasing tile pairs of the block (stable)
      const long nv = (long)t.ncb * t.nsplit, per = round_up(nv, kXcds) / kXcds;
      std::vector<int> work(std::max(t.ncb, 1), 0);
      for (int b = 0; b < t.ncb; ++b)
        for (int r = 0; r < c[4 * b + 1]; ++r)
          for (int u = 0; u < c[4 * b + 3]; ++u) work[b] += (c[4 * b] + r <= c[4 * b + 2] + u) ? 1 : 0;
      std::vector<int2> order(std::max<long>(nv, 1));
      for (long v = 0; v < nv; ++v)
        order[v] = make_int2((int)(v % std::max(t.ncb, 1)), (int)(v / std::max(t.ncb, 1)));
      for (long x0 = 0; x0 < nv; x0 += per)
        std::stable_sort(order.begin() + x0, order.begin() + std::min(nv, x0 + per),
                         [&](const int2& a, const int2& b) { return work[a.x] > work[b.x]; });
      DG_TRY(dalloc(&t.order, order.size()));
      HIP_TRY(hipMemcpy(t.order, order.data(), order.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
  }

  // allocations
  const long np = h->npts_pad;
  DG_TRY(dalloc(&h->atoms, atoms.size()));
  DG_TRY(dalloc(&h->invR, invR.size()));
  DG_TRY(dalloc(&h->rad, h->nrad));
  DG_TRY(dalloc(&h->wrad, h->nrad));
  DG_TRY(dalloc(&h->leb, 4 * (size_t)h->nang));
  for (double** b : {&h->px, &h->py, &h->pz, &h->w, &h->wb, &h->rho, &h->uxc, &h->exd, &h->uJ, &h->V, &h->vcell})
    DG_TRY(dalloc(b, np));
  // one row of slack: the 2-D contraction reads whole 4 x NJ tile blocks, whose tiles
  // past the last function tile reach up to 48 values beyond a row
  DG_TRY(dalloc(&h->phi, (size_t)(np + 1) * h->ldphi));
  DG_TRY(dalloc(&h->sh_atom, h->nshell));
  DG_TRY(dalloc(&h->sh_L, h->nshell));
  DG_TRY(dalloc(&h->sh_nprim, h->nshell));
  DG_TRY(dalloc(&h->sh_off, h->nshell));
  DG_TRY(dalloc(&h->sh_cgf, h->nshell));
  DG_TRY(dalloc(&h->pr_alpha, nprim_tot));
  DG_TRY(dalloc(&h->pr_cn0, nprim_tot));
  DG_TRY(dalloc(&h->pr_cn1, nprim_tot));
  if (bs) {
    h->h_mol_poff.assign(bs->nmol + 1, 0);
    for (int m = 0; m < bs->nmol; ++m) h->h_mol_poff[m + 1] = h->h_mol_poff[m] + (long)h->h_mol_nbf[m] * h->h_mol_nbf[m];
    h->packed = h->h_mol_poff[bs->nmol];
  } else {
    h->packed = (long)h->nbf * h->nbf;
  }
  DG_TRY(dalloc(&h->P, (size_t)h->packed));
  DG_TRY(dalloc(&h->pfrag, (size_t)128 * h->ntb * (h->ntb + 1) * std::max(1, h->nmol)));
  DG_TRY(dalloc(&h->vpart, (size_t)vtotal));
  DG_TRY(dalloc(&h->vbase, (size_t)h->natoms));
  DG_TRY(dalloc(&h->q, h->natoms));
  h->cell_split = (int)std::max(1, std::min(32, 256 / h->natoms));
  DG_TRY(dalloc(&h->qpart, (size_t)h->natoms * h->cell_split));
  DG_TRY(dalloc(&h->qc, h->natoms));
  DG_TRY(dalloc(&h->exc_part, h->nlda_blocks));
  DG_TRY(dalloc(&h->scalars, 4 * (size_t)std::max(1, h->nmol)));
  DG_TRY(dalloc(&h->ylm, ylm.size()));
  DG_TRY(dalloc(&h->minvT, minvT.size()));
  DG_TRY(dalloc(&h->shc, shc.size()));
  DG_TRY(dalloc(&h->knots, h->nrad));
  DG_TRY(dalloc(&h->invh, h->nrad));
  DG_TRY(dalloc(&h->KD, KD.size()));
  DG_TRY(dalloc(&h->rho_lm, (size_t)h->natoms * h->nrad * h->nlm));
  DG_TRY(dalloc(&h->U, (size_t)h->natoms * h->nrad * h->nlm));
  DG_TRY(dalloc(&h->tab, (size_t)h->natoms * h->nrad * h->nlm * 2));
  DG_TRY(dalloc(&h->blocks, blocks.size()));
  DG_TRY(dalloc(&h->items, h->items_cap));
  DG_TRY(dalloc(&h->plen, h->items_cap / 4));
  HIP_TRY(hipMalloc((void**)&h->perm, std::max<long>(h->npairs, 1) * sizeof(unsigned short)));
  HIP_TRY(hipMalloc((void**)&h->pkey, std::max<long>(h->npairs, 1) * sizeof(unsigned short)));
  // partials: two outputs of the 4 x 2 blocks, one of the 4 x 4 blocks
  DG_TRY(dalloc(&h->cpart, std::max({(size_t)h->nsplit_small * h->ngroups * 2 * 1024,
                                      (size_t)h->tb[0].nsplit * h->tb[0].ncb * 4 * 2 * 1024,
                                      (size_t)h->tb[1].nsplit * h->tb[1].ncb * 4 * 1024})));
  for (int q = 0; q < 2; ++q) {
    DG_TRY(dalloc(&h->tb[q].cb, cbs[q].size()));
    DG_TRY(dalloc(&h->tb[q].groups, cbg[q].size()));
  }
  DG_TRY(dalloc(&h->ej_part, h->npts_pad / 256));
  if (h->precision == DG_PREC_MIXED) {
    HIP_TRY(hipMalloc((void**)&h->phf, (size_t)(np + 1) * h->ldphi * sizeof(float)));
  }
  DG_TRY(dalloc(&h->J, (size_t)h->packed));
  DG_TRY(dalloc(&h->XC, (size_t)h->packed));
  if (bs) DG_TRY(batch_alloc(h, bs, amol));

  std::vector<int> shv(h->nshell);
  auto up = [&](void* d, const void* s, size_t bytes) {
    return hipMemcpyAsync(d, s, bytes, hipMemcpyHostToDevice, h->stream);
  };
  HIP_TRY(up(h->atoms, atoms.data(), atoms.size() * 8));
  HIP_TRY(up(h->invR, invR.data(), invR.size() * 8));
  HIP_TRY(up(h->rad, rad.data(), rad.size() * 8));
  HIP_TRY(up(h->wrad, wrad.data(), wrad.size() * 8));
  HIP_TRY(up(h->leb, lebt, 4 * (size_t)h->nang * 8));
  HIP_TRY(up(h->sh_atom, basis->atom, h->nshell * 4));
  HIP_TRY(up(h->sh_L, basis->L, h->nshell * 4));
  HIP_TRY(up(h->sh_nprim, basis->nprim, h->nshell * 4));
  HIP_TRY(up(h->sh_off, basis->offset, h->nshell * 4));
  HIP_TRY(up(h->sh_cgf, cgf_off.data(), h->nshell * 4));
  HIP_TRY(up(h->pr_alpha, alpha.data(), nprim_tot * 8));
  HIP_TRY(up(h->pr_cn0, cn0.data(), nprim_tot * 8));
  HIP_TRY(up(h->pr_cn1, cn1.data(), nprim_tot * 8));
  HIP_TRY(up(h->ylm, ylm.data(), ylm.size() * 8));
  HIP_TRY(up(h->minvT, minvT.data(), minvT.size() * 8));
  HIP_TRY(up(h->shc, shc.data(), shc.size() * 8));
  HIP_TRY(up(h->knots, knots.data(), knots.size() * 8));
  HIP_TRY(up(h->invh, invh.data(), invh.size() * 8));
  HIP_TRY(up(h->KD, KD.data(), KD.size() * 8));
  HIP_TRY(up(h->blocks, blocks.data(), blocks.size() * 4));
  for (int q = 0; q < 2; ++q) {
    HIP_TRY(up(h->tb[q].cb, cbs[q].data(), cbs[q].size() * 4));
    HIP_TRY(up(h->tb[q].groups, cbg[q].data(), cbg[q].size() * 4));
  }
  if (!items.empty()) HIP_TRY(up(h->items, items.data(), items.size() * 4));
  HIP_TRY(hipMemsetAsync(h->phi, 0, (size_t)(np + 1) * h->ldphi * 8, h->stream));
  HIP_TRY(up(h->vbase, vbase.data(), vbase.size() * 8));
  HIP_TRY(hipMemsetAsync(h->vpart, 0, (size_t)vtotal * 8, h->stream));
  HIP_TRY(hipMemsetAsync(h->rho, 0, (size_t)np * 8, h->stream));
  HIP_TRY(hipMemsetAsync(h->vcell, 0, (size_t)np * 8, h->stream));
  HIP_TRY(hipMemsetAsync(h->scalars, 0, 4 * 8 * (size_t)std::max(1, h->nmol), h->stream));

  HIP_TRY(allow_cross_lds_l(h->lmax, h->cross_lds));
  if (!bs) {
    if ((size_t)h->natoms * 64 * 8 > 160 * 1024)
      return fail(DG_ERR_UNSUPPORTED, "dg_create: too many atoms for the Becke kernel's LDS");
    HIP_TRY(hipFuncSetAttribute((const void*)k_becke, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)std::max<size_t>((size_t)h->natoms * 64 * 8, 1024)));
  }
  DG_TRY(run(h, "grid", [&] {
    hipLaunchKernelGGL(k_grid, dim3(nblocks(np, 256)), dim3(256), 0, h->stream, h->atoms, h->rad, h->wrad,
                       h->leb, h->npts, np, h->nrad, h->nang, h->px, h->py, h->pz, h->w);
  }));
  if (bs) {
    DG_TRY(batch_setup(h));
    DG_TRY(sync(h));
    guard.release();
    *out = h;
    return DG_OK;
  }
  DG_TRY(run(h, "becke", [&] {
    hipLaunchKernelGGL(k_becke, dim3(nblocks(np, 64)), dim3(64), (size_t)h->natoms * 64 * 8, h->stream,
                       h->atoms, h->invR, h->natoms, h->npts, h->cell, h->px, h->py, h->pz, h->w, h->wb);
  }));
  DG_TRY(run(h, "phi", [&] {
    hipLaunchKernelGGL(k_phi, dim3(nblocks(h->npts * h->nshell, 256)), dim3(256), 0, h->stream, h->atoms,
                       h->nshell, h->sh_atom, h->sh_L, h->sh_nprim, h->sh_off, h->sh_cgf, h->pr_alpha,
                       h->pr_cn0, h->pr_cn1, h->npts, h->px, h->py, h->pz, h->ldphi, h->phi);
  }));
  if (h->precision == DG_PREC_MIXED) {
    DG_TRY(run(h, "phi_f32", [&] {
      hipLaunchKernelGGL(k_phi_f32, dim3(nblocks(np * h->ldphi, 256)), dim3(256), 0, h->stream, h->phi,
                         np * h->ldphi, h->phf);
    }));
  }
  DG_TRY(sort_items(h));
  {
    const long nblk = np / 64;
    std::vector<unsigned> mk(nblk, 0xffffffffu);
    if (h->screen > 0.0) {
      DG_TRY(dalloc(&h->bmask, nblk));
      DG_TRY(run(h, "block_mask", [&] {
        hipLaunchKernelGGL(k_block_mask, dim3(nblk), dim3(64), 0, h->stream, h->phi, h->ldphi, h->ntb, h->screen,
                           h->bmask);
      }));
      HIP_TRY(hipMemcpyAsync(mk.data(), h->bmask, nblk * 4, hipMemcpyDeviceToHost, h->stream));
      DG_TRY(sync(h));
    }
    for (int q = 0; q < 2; ++q) {
      std::vector<int> l2, o2;
      tile_block_lists(mk, cbs[q], l2, o2);
      DG_TRY(dalloc(&h->tb[q].list, std::max<size_t>(l2.size(), 1)));
      DG_TRY(dalloc(&h->tb[q].loff, o2.size()));
      HIP_TRY(hipMemcpy(h->tb[q].list, l2.data(), l2.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(h->tb[q].loff, o2.data(), o2.size() * 4, hipMemcpyHostToDevice));
    }
  }
  DG_TRY(sync(h));
  guard.release();
  *out = h;
  return DG_OK;
}

int dg_create(const dg_molecule* mol, const dg_basis* basis, const dg_grid_params* prm, dg_handle** out) {
  return create_impl(mol, basis, prm, nullptr, out);
}

int dg_destroy(dg_handle* h) {
  if (!h) return DG_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
  return DG_OK;
}

int dg_info(const dg_handle* h, int* nbf, long* npts, int* natoms) {
  if (!h) return fail(DG_ERR_INVALID, "dg_info: null handle");
  if (nbf) *nbf = h->nbf;
  if (npts) *npts = h->npts;
  if (natoms) *natoms = h->natoms;
  return DG_OK;
}

int dg_set_stream(dg_handle* h, void* stream) {
  if (!h) return fail(DG_ERR_INVALID, "dg_set_stream: null handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (stream) {
    if (h->own_stream) HIP_TRY(hipStreamDestroy(h->stream));
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
  } else if (!h->own_stream) {
    HIP_TRY(hipStreamCreate(&h->stream));
    h->own_stream = true;
  }
  return DG_OK;
}

int dg_get_stream(const dg_handle* h, void** stream) {
  if (!h || !stream) return fail(DG_ERR_INVALID, "dg_get_stream: null argument");
  *stream = (void*)h->stream;
  return DG_OK;
}

int dg_set_density(dg_handle* h, const double* P, int nbf) {
  if (!h || !P) return fail(DG_ERR_INVALID, "dg_set_density: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_set_density: not for a batch handle (use dg_batch_build)");
  if (nbf != h->nbf) return fail(DG_ERR_INVALID, "dg_set_density: nbf mismatch");
  for (long i = 0; i < (long)nbf * nbf; ++i)
    if (!std::isfinite(P[i])) return fail(DG_ERR_INVALID, "dg_set_density: P contains a non-finite value");
  HIP_TRY(hipMemcpyAsync(h->P, P, (size_t)nbf * nbf * 8, hipMemcpyHostToDevice, h->stream));
  DG_TRY(stage_density(h, h->P));
  return sync(h);
}

int dg_correct_densities(dg_handle* h, double* nelec_before) {
  if (!h) return fail(DG_ERR_INVALID, "dg_correct_densities: null handle");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_correct_densities: not for a batch handle (use dg_batch_build)");
  if (!h->has_density) return fail(DG_ERR_STATE, "dg_correct_densities: no density set");
  DG_TRY(stage_correct(h));
  h->has_hartree = h->has_xc = false;
  if (nelec_before) {
    double s[4];
    HIP_TRY(hipMemcpyAsync(s, h->scalars, 4 * 8, hipMemcpyDeviceToHost, h->stream));
    DG_TRY(sync(h));
    *nelec_before = s[1];
  }
  return sync(h);
}

int dg_calculate_density(dg_handle* h, double* nelec) {
  if (!h || !nelec) return fail(DG_ERR_INVALID, "dg_calculate_density: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_calculate_density: not for a batch handle (use dg_batch_build)");
  if (!h->has_density) return fail(DG_ERR_STATE, "dg_calculate_density: no density set");
  DG_TRY(run(h, "cell_sum", [&] {
    hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->rho, h->w, h->cell, h->q);
  }));
  std::vector<double> q(h->natoms);
  DG_TRY(copy_out(h, q.data(), h->q, h->natoms));
  double s = 0.0;
  for (double v : q) s += v;
  *nelec = s;
  return DG_OK;
}

int dg_hartree(dg_handle* h, double* J) {
  if (!h) return fail(DG_ERR_INVALID, "dg_hartree: null handle");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_hartree: not for a batch handle (use dg_batch_build)");
  if (!h->has_density) return fail(DG_ERR_STATE, "dg_hartree: no density set");
  // cell populations of the current density (the Poisson boundary values)
  DG_TRY(run(h, "cell_sum", [&] {
    hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->rho, h->w, h->cell, h->qc);
  }));
  DG_TRY(stage_hartree(h));
  DG_TRY(stage_contract(h, h->uJ, nullptr, h->J, nullptr));
  h->has_hartree = true;
  if (J) return copy_out(h, J, h->J, (size_t)h->nbf * h->nbf);
  return sync(h);
}

int dg_xc(dg_handle* h, double* XC, double* exc) {
  if (!h) return fail(DG_ERR_INVALID, "dg_xc: null handle");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_xc: not for a batch handle (use dg_batch_build)");
  if (!h->has_density) return fail(DG_ERR_STATE, "dg_xc: no density set");
  DG_TRY(stage_lda(h));
  DG_TRY(stage_contract(h, h->uxc, nullptr, h->XC, nullptr));
  DG_TRY(stage_scalars(h, false));
  h->has_xc = true;
  if (XC) DG_TRY(copy_out(h, XC, h->XC, (size_t)h->nbf * h->nbf));
  if (exc) {
    double s[4];
    DG_TRY(copy_out(h, s, h->scalars, 4));
    *exc = s[0];
  }
  return sync(h);
}

int dg_build_device(dg_handle* h, const double* dP, double* dJ, double* dXC, double* dscalars) {
  if (!h || !dP) return fail(DG_ERR_INVALID, "dg_build_device: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_build_device: not for a batch handle (use dg_batch_build)");
  DG_TRY(stage_density(h, dP));
  DG_TRY(stage_correct_lda(h));
  DG_TRY(stage_hartree(h));
  const ScalarArgs sa = scalar_args(h, true, dscalars);
  DG_TRY(stage_contract(h, h->uxc, h->uJ, dXC ? dXC : h->XC, dJ ? dJ : h->J, &sa));
  h->has_hartree = h->has_xc = true;
  return DG_OK;
}

int dg_build_xc_device(dg_handle* h, const double* dP, double* dXC, double* dscalars) {
  if (!h || !dP || !dXC) return fail(DG_ERR_INVALID, "dg_build_xc_device: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_build_xc_device: not for a batch handle");
  DG_TRY(stage_density(h, dP));
  DG_TRY(stage_correct_lda(h));
  const ScalarArgs sa = scalar_args(h, false, dscalars);
  DG_TRY(stage_contract(h, h->uxc, nullptr, dXC, nullptr, &sa));
  h->has_xc = true;
  return DG_OK;
}

int dg_build_fock_device(dg_handle* h, const double* dP, double* dG, double* dscalars) {
  if (!h || !dP || !dG) return fail(DG_ERR_INVALID, "dg_build_fock_device: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_build_fock_device: not for a batch handle");
  if (!h->uG) DG_TRY(dalloc(&h->uG, (size_t)h->npts_pad));
  DG_TRY(stage_density(h, dP));
  DG_TRY(stage_correct_lda(h));
  DG_TRY(stage_hartree(h));
  DG_TRY(run(h, "fock_weights", [&] {
    hipLaunchKernelGGL(k_fock_weights, dim3(nblocks(h->npts_pad, 256)), dim3(256), 0, h->stream, h->uJ, h->uxc,
                       h->npts_pad, h->uG);
  }));
  const ScalarArgs sa = scalar_args(h, true, dscalars);
  DG_TRY(stage_contract(h, h->uG, nullptr, dG, nullptr, &sa));
  h->has_hartree = h->has_xc = true;
  return DG_OK;
}

int dg_get_weights(dg_handle* h, double* w) {
  if (!h || !w) return fail(DG_ERR_INVALID, "dg_get_weights: null argument");
  return copy_out(h, w, h->w, h->npts);
}
int dg_get_becke_weights(dg_handle* h, double* wb) {
  if (!h || !wb) return fail(DG_ERR_INVALID, "dg_get_becke_weights: null argument");
  return copy_out(h, wb, h->wb, h->npts);
}
int dg_get_positions(dg_handle* h, double* xyz) {
  if (!h || !xyz) return fail(DG_ERR_INVALID, "dg_get_positions: null argument");
  std::vector<double> x(h->npts), y(h->npts), z(h->npts);
  DG_TRY(copy_out(h, x.data(), h->px, h->npts));
  DG_TRY(copy_out(h, y.data(), h->py, h->npts));
  DG_TRY(copy_out(h, z.data(), h->pz, h->npts));
  for (long i = 0; i < h->npts; ++i) {
    xyz[3 * i] = x[i];
    xyz[3 * i + 1] = y[i];
    xyz[3 * i + 2] = z[i];
  }
  return DG_OK;
}
int dg_get_densities(dg_handle* h, double* rho) {
  if (!h || !rho) return fail(DG_ERR_INVALID, "dg_get_densities: null argument");
  if (!h->has_density) return fail(DG_ERR_STATE, "dg_get_densities: no density set");
  return copy_out(h, rho, h->rho, h->npts);
}
int dg_get_amplitudes(dg_handle* h, double* phi) {
  if (!h || !phi) return fail(DG_ERR_INVALID, "dg_get_amplitudes: null argument");
  HIP_TRY(hipMemcpy2DAsync(phi, (size_t)h->nbf * 8, h->phi, (size_t)h->ldphi * 8, (size_t)h->nbf * 8,
                           h->npts, hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}
int dg_get_hartree_potential(dg_handle* h, double* V) {
  if (!h || !V) return fail(DG_ERR_INVALID, "dg_get_hartree_potential: null argument");
  if (!h->has_hartree) return fail(DG_ERR_STATE, "dg_get_hartree_potential: no Hartree potential computed");
  return copy_out(h, V, h->V, h->npts);
}
int dg_get_hartree_parts(dg_handle* h, double* rho_lm, double* U, double* vcell) {
  if (!h) return fail(DG_ERR_INVALID, "dg_get_hartree_parts: null handle");
  if (!h->has_hartree) return fail(DG_ERR_STATE, "dg_get_hartree_parts: no Hartree potential computed");
  const size_t n = (size_t)h->natoms * h->nrad * h->nlm;
  if (rho_lm) DG_TRY(copy_out(h, rho_lm, h->rho_lm, n));
  if (U) DG_TRY(copy_out(h, U, h->U, n));
  if (vcell) DG_TRY(copy_out(h, vcell, h->vcell, h->npts));
  return DG_OK;
}
int dg_get_xc_parts(dg_handle* h, double* exd, double* wv) {
  if (!h) return fail(DG_ERR_INVALID, "dg_get_xc_parts: null handle");
  if (!h->has_xc) return fail(DG_ERR_STATE, "dg_get_xc_parts: no XC computed");
  if (exd) DG_TRY(copy_out(h, exd, h->exd, h->npts));
  if (wv) DG_TRY(copy_out(h, wv, h->uxc, h->npts));
  return DG_OK;
}

int dg_xc_adjoint(dg_handle* h, int hold_c, double* dEdP, double* grad) {
  if (!h) return fail(DG_ERR_INVALID, "dg_xc_adjoint: null handle");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_xc_adjoint: not for a batch handle (use dg_batch_build)");
  if (!h->has_xc || !h->corrected || !h->last_P)
    return fail(DG_ERR_STATE, "dg_xc_adjoint: needs a corrected density and dg_xc (or a build) first");
  const long np = h->npts_pad, nblk = np / 64;
  const int na = h->natoms;
  if (grad && na > 65535)
    return fail(DG_ERR_UNSUPPORTED, "dg_xc_adjoint: forces for at most 65535 atoms");
  if (!h->adj_ub) {
    for (double** b : {&h->adj_ub, &h->adj_ab, &h->adj_udp}) DG_TRY(dalloc(b, np));
    DG_TRY(dalloc(&h->adj_lam, np / 256 + 1));
    DG_TRY(dalloc(&h->adj_dP, (size_t)h->nbf * h->nbf));
    DG_TRY(dalloc(&h->adj_out, (size_t)6 * na + 1));
  }
  if (grad) DG_TRY(ensure_tmat_bufs(h));
  if (grad && !h->adj_part) {
    DG_TRY(dalloc(&h->adj_part, (size_t)3 * na * nblk));
    DG_TRY(dalloc(&h->adj_mot, (size_t)3 * h->npts));
    // shells grouped by centre atom
    std::vector<int> at(h->nshell), start(na + 1, 0), list;
    HIP_TRY(hipMemcpyAsync(at.data(), h->sh_atom, h->nshell * 4, hipMemcpyDeviceToHost, h->stream));
    DG_TRY(sync(h));
    for (int a = 0; a < na; ++a) {
      start[a] = (int)list.size();
      for (int s = 0; s < h->nshell; ++s)
        if (at[s] == a) list.push_back(s);
    }
    start[na] = (int)list.size();
    DG_TRY(dalloc(&h->ash_start, na + 1));
    DG_TRY(dalloc(&h->ash_list, std::max<size_t>(list.size(), 1)));
    HIP_TRY(hipMemcpyAsync(h->ash_start, start.data(), (na + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->ash_list, list.data(), list.size() * 4, hipMemcpyHostToDevice, h->stream));
    DG_TRY(sync(h));
  }
  // Lambda = sum w v rho (uxc = w v), then the per-point factors
  const long nparts = np / 256;
  DG_TRY(run(h, "adj_lambda", [&] {
    hipLaunchKernelGGL(k_dot_parts, dim3(nparts), dim3(256), 0, h->stream, h->uxc, h->rho, h->npts, h->adj_lam + 1);
    hipLaunchKernelGGL(k_row_sum, dim3(1), dim3(1024), 0, h->stream, h->adj_lam + 1, nparts, h->adj_lam);
  }));
  DG_TRY(run(h, "adj_weights", [&] {
    hipLaunchKernelGGL(k_adjoint_weights, dim3(nblocks(np, 256)), dim3(256), 0, h->stream, h->uxc, h->w, h->rho,
                       h->exd, h->scalars, h->adj_lam, h->zsum, hold_c ? 1 : 0, np, h->adj_ub, h->adj_ab,
                       h->adj_udp);
  }));
  if (dEdP) {
    // dE/dP = Phi^T diag(udp) Phi with the FP64 contraction in either precision mode.
    // Unscreened handles on the 2-D path take, per tile block, only the point
    // blocks where both its row and its column tiles are within reach
    // (k_tile_mask; the terms left out are below e^-100 of an amplitude).
    const bool lists = h->screen == 0.0 && h->ntb > 5 && h->ntb <= 32;
    if (lists && !h->adj_list) {
      DG_TRY(ensure_tile_mask(h));
      const long nblk_m = h->npts_pad / 64;
      std::vector<unsigned> mk(nblk_m);
      std::vector<int> cb(4 * (size_t)h->tb[1].ncb), l2, o2;
      HIP_TRY(hipMemcpyAsync(mk.data(), h->adj_tmask, nblk_m * 4, hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipMemcpyAsync(cb.data(), h->tb[1].cb, cb.size() * 4, hipMemcpyDeviceToHost, h->stream));
      DG_TRY(sync(h));
      tile_block_lists(mk, cb, l2, o2);
      DG_TRY(dalloc(&h->adj_list, std::max<size_t>(l2.size(), 1)));
      DG_TRY(dalloc(&h->adj_loff, o2.size()));
      HIP_TRY(hipMemcpyAsync(h->adj_list, l2.data(), l2.size() * 4, hipMemcpyHostToDevice, h->stream));
      HIP_TRY(hipMemcpyAsync(h->adj_loff, o2.data(), o2.size() * 4, hipMemcpyHostToDevice, h->stream));
      DG_TRY(sync(h));
    }
    const int saved = h->precision;
    h->precision = DG_PREC_FP64;
    const int rc = stage_contract(h, h->adj_udp, nullptr, h->adj_dP, nullptr, nullptr, lists ? h->adj_list : nullptr,
                                  lists ? h->adj_loff : nullptr);
    h->precision = saved;
    DG_TRY(rc);
    DG_TRY(copy_out(h, dEdP, h->adj_dP, (size_t)h->nbf * h->nbf));
  }
  if (grad) {
    DG_TRY(launch_tmat(h, true));
    DG_TRY(run(h, "adj_force_basis", [&] {
      hipLaunchKernelGGL(k_force_basis, dim3(nblk), dim3(64), 0, h->stream, h->atoms, na, h->ash_start, h->ash_list,
                         h->sh_L, h->sh_nprim, h->sh_off, h->sh_cgf, h->pr_alpha, h->pr_cn0, h->pr_cn1, h->npts,
                         h->px, h->py, h->pz, h->ldphi, h->adj_T, h->adj_ub, nblk, h->adj_part, h->adj_mot);
    }));
    HIP_TRY(hipMemsetAsync(h->adj_out + 6 * na, 0, sizeof(double), h->stream));
    DG_TRY(run(h, "adj_force_becke", [&] {
      hipLaunchKernelGGL(k_force_becke, dim3(nblk), dim3(64), 0, h->stream, h->atoms, h->invR, na, h->npts,
                         h->cell, h->nrad, h->nang, h->rad, h->wrad, h->leb, h->px, h->py, h->pz, h->adj_ab, nblk,
                         h->adj_part, h->adj_mot, (int*)(h->adj_out + 6 * na));
    }));
    DG_TRY(run(h, "adj_force_sum", [&] {
      hipLaunchKernelGGL(k_row_sum, dim3(3 * na), dim3(1024), 0, h->stream, h->adj_part, nblk, h->adj_out);
      hipLaunchKernelGGL(k_row_sum, dim3(3 * na), dim3(1024), 0, h->stream, h->adj_mot, h->cell,
                         h->adj_out + 3 * na);
    }));
    std::vector<double> o((size_t)6 * na + 1);
    DG_TRY(copy_out(h, o.data(), h->adj_out, o.size()));
    int over = 0;
    std::memcpy(&over, &o[6 * na], sizeof(int));
    if (over)
      return fail(DG_ERR_UNSUPPORTED, "dg_xc_adjoint: a grid point has more than " + std::to_string(kBeckeActive) +
                                          " significant Becke cell functions");
    for (int a = 0; a < na; ++a)
      for (int d = 0; d < 3; ++d) grad[3 * a + d] = o[3 * a + d] + o[3 * na + d * na + a];
  }
  return sync(h);
}

int dg_get_amplitude_gradients(dg_handle* h, double* dphi) {
  if (!h || !dphi) return fail(DG_ERR_INVALID, "dg_get_amplitude_gradients: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_get_amplitude_gradients: not for a batch handle");
  // in chunks of at most 2^27 doubles of device memory
  const long chunk = std::max(1L, std::min(h->npts, (1L << 27) / (3L * h->nbf)));
  double* buf = nullptr;
  DG_TRY(dalloc(&buf, (size_t)chunk * 3 * h->nbf));
  int rc = DG_OK;
  for (long p0 = 0; p0 < h->npts && rc == DG_OK; p0 += chunk) {
    const long n = std::min(chunk, h->npts - p0);
    rc = run(h, "phi_grad", [&] {
      hipLaunchKernelGGL(k_phi_grad, dim3(nblocks(n * h->nshell, 256)), dim3(256), 0, h->stream, h->atoms,
                         h->nshell, h->sh_atom, h->sh_L, h->sh_nprim, h->sh_off, h->sh_cgf, h->pr_alpha, h->pr_cn0,
                         h->pr_cn1, p0, n, h->px, h->py, h->pz, h->nbf, buf);
    });
    if (rc == DG_OK) rc = copy_out(h, dphi + p0 * 3 * h->nbf, buf, (size_t)n * 3 * h->nbf);
  }
  (void)hipFree(buf);
  return rc;
}

int dg_get_density_gradients(dg_handle* h, double* grad) {
  if (!h || !grad) return fail(DG_ERR_INVALID, "dg_get_density_gradients: null argument");
  if (h->nmol) return fail(DG_ERR_STATE, "dg_get_density_gradients: not for a batch handle");
  if (!h->has_density || !h->last_P)
    return fail(DG_ERR_STATE, "dg_get_density_gradients: no density evaluated (dg_set_density or a build first)");
  DG_TRY(ensure_tmat_bufs(h));
  DG_TRY(launch_tmat(h));
  double* buf = nullptr;
  DG_TRY(dalloc(&buf, (size_t)3 * h->npts));
  int rc = run(h, "density_grad", [&] {
    hipLaunchKernelGGL(k_density_grad, dim3(nblocks(h->npts, 256)), dim3(256), 0, h->stream, h->atoms, h->nshell,
                       h->sh_atom, h->sh_L, h->sh_nprim, h->sh_off, h->sh_cgf, h->pr_alpha, h->pr_cn0, h->pr_cn1, 0L,
                       h->npts, h->px, h->py, h->pz, h->ldphi, h->adj_T, h->corrected ? h->scalars : nullptr,
                       h->zsum, nullptr, buf);
  });
  if (rc == DG_OK) rc = copy_out(h, grad, buf, (size_t)3 * h->npts);
  (void)hipFree(buf);
  return rc;
}

int dg_density_at_points(const dg_molecule* mol, const dg_basis* basis, long npts, const double* xyz,
                         const double* P, double* rho, double* grad) {
  g_last_error.clear();
  const char* who = "dg_density_at_points";
  DG_TRY(point_check(mol, basis, npts, xyz, P, who));
  if (npts == 0 || (!rho && !grad)) return DG_OK;
  std::unique_ptr<PointTables> t(new (std::nothrow) PointTables);
  if (!t) return fail(DG_ERR_ALLOC, "dg_density_at_points: out of host memory");
  DG_TRY(point_tables(mol, basis, *t, who));
  PointWork w;
  w.chunk = std::min(round_up(npts, 64), kPointChunk);
  struct Bufs {  // device buffers and the stream of this call, released on every return
    std::vector<void*> p;
    hipStream_t st = nullptr;
    ~Bufs() {
      if (st) (void)hipStreamSynchronize(st);
      for (void* b : p) (void)hipFree(b);
      if (st) (void)hipStreamDestroy(st);
    }
  } B;
  auto get = [&](auto** ptr, size_t n) {
    const int rc = dalloc(ptr, n);
    if (rc == DG_OK) B.p.push_back((void*)*ptr);
    return rc;
  };
  const long C = w.chunk;
  DG_TRY(get(&w.atoms, 4 * (size_t)t->na));
  DG_TRY(get(&w.alpha, t->nprim));
  DG_TRY(get(&w.cn0, t->nprim));
  DG_TRY(get(&w.cn1, t->nprim));
  DG_TRY(get(&w.ishell, 5 * (size_t)t->nshell));
  DG_TRY(get(&w.P, (size_t)t->nbf * t->nbf));
  DG_TRY(get(&w.frag, (size_t)256 * t->nt * t->nt));
  DG_TRY(get(&w.xyz, 3 * (size_t)C));
  DG_TRY(get(&w.pxyz, 3 * (size_t)C));
  DG_TRY(get(&w.phi, (size_t)C * t->ld));
  DG_TRY(get(&w.T, (size_t)C * t->ld));
  DG_TRY(get(&w.rho, C));
  DG_TRY(get(&w.grad, 3 * (size_t)C));
  HIP_TRY(hipStreamCreate(&B.st));
  return point_run(*t, basis, P, npts, xyz, rho, grad, w, B.st);
}

int dg_profile_enable(dg_handle* h, int enable) {
  if (!h) return fail(DG_ERR_INVALID, "dg_profile_enable: null handle");
  h->prof = enable != 0;
  return DG_OK;
}

static int collect(dg_handle* h) {
  if (h->pending.empty()) return DG_OK;
  DG_TRY(sync(h));
  for (auto& e : h->pending) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e.e0, e.e1));
    ProfStat& s = h->stats[e.name];
    s.ms += ms;
    s.n += 1;
    h->event_pool.push_back(e.e0);
    h->event_pool.push_back(e.e1);
  }
  h->pending.clear();
  return DG_OK;
}

int dg_profile_read(dg_handle* h, int max, char* names, double* total_ms, int* launches, int* n) {
  if (!h) return fail(DG_ERR_INVALID, "dg_profile_read: null handle");
  DG_TRY(collect(h));
  int i = 0;
  for (auto& kv : h->stats) {
    if (i < max) {
      if (names) {
        std::strncpy(names + 48 * i, kv.first.c_str(), 47);
        names[48 * i + 47] = 0;
      }
      if (total_ms) total_ms[i] = kv.second.ms;
      if (launches) launches[i] = kv.second.n;
    }
    ++i;
  }
  if (n) *n = i;
  return DG_OK;
}

int dg_profile_reset(dg_handle* h) {
  if (!h) return fail(DG_ERR_INVALID, "dg_profile_reset: null handle");
  DG_TRY(collect(h));
  h->stats.clear();
  return DG_OK;
}

int dg_mfma_probe(const double* a, const double* b, double* d) {
  if (!a || !b || !d) return fail(DG_ERR_INVALID, "dg_mfma_probe: null argument");
  double* buf = nullptr;
  DG_TRY(dalloc(&buf, 64 + 64 + 256));
  hipError_t e = hipMemcpy(buf, a, 64 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(buf + 64, b, 64 * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_raw, dim3(1), dim3(64), 0, 0, buf, buf + 64, buf + 128);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(d, buf + 128, 256 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(DG_ERR_HIP, std::string("dg_mfma_probe: ") + hipGetErrorString(e));
  return DG_OK;
}

int dg_mfma_probe_f32(const float* a, const float* b, float* d) {
  if (!a || !b || !d) return fail(DG_ERR_INVALID, "dg_mfma_probe_f32: null argument");
  float* buf = nullptr;
  hipError_t e = hipMalloc((void**)&buf, (64 + 64 + 256) * sizeof(float));
  if (e != hipSuccess) return fail(DG_ERR_ALLOC, "dg_mfma_probe_f32: hipMalloc failed");
  e = hipMemcpy(buf, a, 64 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(buf + 64, b, 64 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_raw_f32, dim3(1), dim3(64), 0, 0, buf, buf + 64, buf + 128);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(d, buf + 128, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(DG_ERR_HIP, std::string("dg_mfma_probe_f32: ") + hipGetErrorString(e));
  return DG_OK;
}

int dg_selftest_mfma(int* ok) {
  if (!ok) return fail(DG_ERR_INVALID, "dg_selftest_mfma: null argument");
  *ok = 0;
  double* d = nullptr;
  DG_TRY(dalloc(&d, 256));
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, d);
  std::vector<double> out(256);
  hipError_t e = hipMemcpy(out.data(), d, 256 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(DG_ERR_HIP, std::string("dg_selftest_mfma: ") + hipGetErrorString(e));
  int good = 1;
  for (int lane = 0; lane < 64; ++lane)
    for (int v = 0; v < 4; ++v) {
      const int i = 4 * v + (lane >> 4), j = lane & 15;
      double ref = 0.0;
      for (int k = 0; k < 4; ++k) ref += (1.0 + i + 100.0 * k) * (1.0 + 0.5 * j - 3.0 * k);
      if (std::fabs(out[v * 64 + lane] - ref) > 1e-9 * std::fabs(ref) + 1e-12) good = 0;
    }
  *ok = good;
  return DG_OK;
}

}  // extern "C"

#include "batch.inl"
#include "noalloc.inl"
