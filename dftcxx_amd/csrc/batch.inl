// Batch mode: many independent molecules in one handle, one build per SCF
// iteration for all of them (no counterpart in the reference, which runs one
// molecule per process, src/dftcxx.cpp:53-54).  Included at the end of
// dftgrid.hip (one translation unit).
//
// Layout: the molecules' atoms are concatenated, so global atom a owns the
// points [a cell, (a+1) cell) exactly as in a single-molecule handle; the
// Becke partition, the basis, the density matrix, the renormalisation, the
// cross-atom potentials and the contractions are restricted to each atom's
// own molecule.  Phi rows are 32 wide (molecules of at most 32 CGFs); P, J and
// XC are packed: molecule m's nbf_m x nbf_m column-major matrix starts at
// mol_poff[m].  All stages are launched on the handle's stream and captured
// once into a hipGraph, which is replayed every iteration.

// Becke weights with the partition over the atoms of the point's molecule
// (src/moleculargrid.cpp:229-329); the same arithmetic, in the same order, as k_becke.
__global__ __launch_bounds__(64) void k_becke_batch(const double* __restrict__ atoms,
                                                    const double* __restrict__ invR, int natoms,
                                                    long npts, long cell, const int* __restrict__ atom_mol,
                                                    const int* __restrict__ mol_atom0,
                                                    const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz, double* w, double* wb) {
  extern __shared__ double dist[];  // max atoms per molecule x 64
  const int lane = threadIdx.x;
  const long p = (long)blockIdx.x * 64 + lane;
  if (p >= npts) {
    if (p < (long)gridDim.x * 64) wb[p] = 0.0;
    return;
  }
  const double x = px[p], y = py[p], z = pz[p];
  const int own = (int)(p / cell);
  const int m = atom_mol[own], a0 = mol_atom0[m], a1 = mol_atom0[m + 1];
  for (int n = a0; n < a1; ++n) {
    const double dx = x - atoms[4 * n], dy = y - atoms[4 * n + 1], dz = z - atoms[4 * n + 2];
    dist[(n - a0) * 64 + lane] = sqrt(dx * dx + dy * dy + dz * dz);
  }
  double denom = 0.0, nom = 1.0;
  for (int n = a0; n < a1; ++n) {
    const double dn = dist[(n - a0) * 64 + lane];
    double prod = 1.0;
    for (int k = a0; k < a1; ++k) {
      if (k == n) continue;
      double mu = (dn - dist[(k - a0) * 64 + lane]) * invR[(long)n * natoms + k];
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

// Amplitudes of the point's own molecule: thread per (point, local shell).
__global__ void k_phi_batch(const double* __restrict__ atoms, const int* __restrict__ atom_mol,
                            const int* __restrict__ mol_shell0, int max_shells, long cell,
                            const int* __restrict__ sh_atom, const int* __restrict__ sh_L,
                            const int* __restrict__ sh_nprim, const int* __restrict__ sh_off,
                            const int* __restrict__ sh_cgf, const double* __restrict__ alpha,
                            const double* __restrict__ cn0, const double* __restrict__ cn1, long npts,
                            const double* __restrict__ px, const double* __restrict__ py,
                            const double* __restrict__ pz, int ldphi, double* phi) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts * max_shells) return;
  const long p = g / max_shells;
  const int m = atom_mol[p / cell];
  const int s = mol_shell0[m] + (int)(g % max_shells);
  if (s >= mol_shell0[m + 1]) return;
  phi_shell(atoms, sh_atom, sh_L, sh_nprim, sh_off, sh_cgf, alpha, cn0, cn1, px, py, pz, ldphi, phi, p, s);
}

// P fragments of every molecule (two 16-wide tiles, symmetric halving as
// k_pfrag): frag[m][768] from the packed P.
constexpr int kBatchFrag = 128 * 2 * 3;
__global__ void k_pfrag_batch(const double* __restrict__ P, const long* __restrict__ poff,
                              const int* __restrict__ mol_nbf, double* frag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (e >= kBatchFrag) return;
  frag[(long)m * kBatchFrag + e] = pfrag_value(P + poff[m], mol_nbf[m], e);
}

// rho_p = 2 phi_p^T P_m phi_p, the k_density scheme with two tiles.  A wave's
// 16 points may span two molecules (molecule ranges are whole atomic cells):
// the wave then runs the MFMAs once per molecule and each point keeps the
// products of its own molecule.
__global__ __launch_bounds__(256) void k_density_batch(const double* __restrict__ phi,
                                                       const double* __restrict__ frag,
                                                       const int* __restrict__ atom_mol,
                                                       const int* __restrict__ mol_nbf, long cell,
                                                       long npts, double* rho) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const long p0 = (long)blockIdx.x * 64 + wave * 16;
  const long pc = p0 + c;
  const double* row = phi + pc * 32;
  const int mine = pc < npts ? atom_mol[pc / cell] : -1;
  const int mfirst = p0 < npts ? atom_mol[p0 / cell] : -1;
  const long plast = (p0 + 15 < npts) ? p0 + 15 : npts - 1;
  const int mlast = p0 < npts ? atom_mol[plast / cell] : -1;
  // the second 16-function tile only for molecules of more than 16 functions
  // (its amplitudes are zero otherwise, and so are its products)
  bool two = false;
  for (int m = mfirst; m >= 0 && m <= mlast; ++m) two = two || mol_nbf[m] > 16;
  double f[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) f[s] = (s < 4 || two) ? row[16 * (s >> 2) + 4 * g + (s & 3)] : 0.0;
  double acc = 0.0;
  for (int m = mfirst; m >= 0 && m <= mlast; ++m) {
    const double* fr = frag + (long)m * kBatchFrag;
    const int nb = mol_nbf[m] > 16 ? 2 : 1;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (b >= nb) continue;
      dbl4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4 * (b + 1); ++s) d = mfma16x16x4(fr[128 * b * (b + 1) + s * 64 + lane], f[s], d);
      // register v of lane (c, g) holds T_b[pc][16b + 4g + v]
      const double t = d[0] * f[4 * b] + d[1] * f[4 * b + 1] + d[2] * f[4 * b + 2] + d[3] * f[4 * b + 3];
      if (m == mine) acc += t;
    }
  }
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (g == 0) rho[pc] = 2.0 * acc;
}

// Per molecule: q_raw = sum of its atoms' populations (atom order), the
// renormalisation factor f = Z_m / q_raw (src/moleculargrid.cpp:132-146) and
// the corrected populations qc_a = f q_a.
__global__ void k_mol_factors(const double* __restrict__ q, const int* __restrict__ mol_atom0,
                              const double* __restrict__ mol_z, int nmol, double* fac, double* qraw,
                              double* qc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  double qs = 0.0;
  for (int a = mol_atom0[m]; a < mol_atom0[m + 1]; ++a) qs += q[a];
  const double f = mol_z[m] / qs;
  fac[m] = f;
  qraw[m] = qs;
  for (int a = mol_atom0[m]; a < mol_atom0[m + 1]; ++a) qc[a] = f * q[a];
}

__global__ void k_scale_batch(double* rho, long npts, long cell, const int* __restrict__ atom_mol,
                              const double* __restrict__ fac) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npts) rho[p] *= fac[atom_mol[p / cell]];
}

// C_o[m] = Phi_m^T diag(u_o) Phi_m.  A work-group takes one chunk {m, p_begin,
// p_end} of molecule m's points; its 4 waves interleave 4-point steps.  Lane
// (c, g) feeds A = phi[p_g][16I + c] and B = u phi[p_g][16J + c] for the tiles
// t = (I, J) = (0,0), (0,1), (1,1) (the last two only when nbf_m > 16) and
// receives C[16I + 4v + g][16J + c] in result register v.  The second tile of
// amplitudes is read only when nbf_m > 16, and the operands of the next two
// steps are in flight while a step's MFMAs run.  The 4 waves are summed in LDS
// in a fixed order: part[chunk][o][t][v][lane].
template <int NOUT>
__global__ __launch_bounds__(256) void k_contract_batch(const double* __restrict__ phi,
                                                        const double* __restrict__ u0,
                                                        const double* __restrict__ u1,
                                                        const int* __restrict__ chunks,
                                                        const int* __restrict__ mol_nbf, double* part) {
  __shared__ double red[4][NOUT * 768];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int m = chunks[3 * blockIdx.x];
  const long pb = chunks[3 * blockIdx.x + 1], pe = chunks[3 * blockIdx.x + 2];
  const bool two = mol_nbf[m] > 16;
  dbl4 acc[NOUT][3];
#pragma unroll
  for (int o = 0; o < NOUT; ++o)
#pragma unroll
    for (int t = 0; t < 3; ++t) acc[o][t] = dbl4{0.0, 0.0, 0.0, 0.0};
  constexpr int PF = 3;  // the current step and the next two
  double ra0[PF], ra1[PF], rw[PF][NOUT];
  auto load = [&](int k, long p0) {
    const long p = p0 + g;
    ra0[k] = ra1[k] = 0.0;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) rw[k][o] = 0.0;
    if (p < pe) {
      ra0[k] = phi[p * 32 + c];
      if (two) ra1[k] = phi[p * 32 + 16 + c];
      rw[k][0] = u0[p];
      if (NOUT > 1) rw[k][NOUT - 1] = u1[p];
    }
  };
#pragma unroll
  for (int k = 0; k < PF; ++k) load(k, pb + 4 * wave + 16 * k);
  for (long p0 = pb + 4 * wave; p0 < pe; p0 += 16) {
    const double a0 = ra0[0], a1 = ra1[0];
    double w[NOUT];
#pragma unroll
    for (int o = 0; o < NOUT; ++o) w[o] = rw[0][o];
#pragma unroll
    for (int k = 0; k + 1 < PF; ++k) {
      ra0[k] = ra0[k + 1];
      ra1[k] = ra1[k + 1];
#pragma unroll
      for (int o = 0; o < NOUT; ++o) rw[k][o] = rw[k + 1][o];
    }
    load(PF - 1, p0 + 16 * PF);
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
      acc[o][0] = mfma16x16x4(a0, w[o] * a0, acc[o][0]);
      if (two) {
        acc[o][1] = mfma16x16x4(a0, w[o] * a1, acc[o][1]);
        acc[o][2] = mfma16x16x4(a1, w[o] * a1, acc[o][2]);
      }
    }
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o)
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave][(o * 3 + t) * 256 + v * 64 + lane] = acc[o][t][v];
  __syncthreads();
  double* out = part + (long)blockIdx.x * NOUT * 768;
  for (int e = threadIdx.x; e < NOUT * 768; e += 256) out[e] = red[0][e] + red[1][e] + red[2][e] + red[3][e];
}

// out_o[m] (packed, column-major, exactly symmetric) = sum over molecule m's
// chunks, in chunk order, of the partial tiles.  One block per molecule.
__global__ __launch_bounds__(256) void k_contract_batch_reduce(const double* __restrict__ part, int nout,
                                                               const int* __restrict__ mol_chunk0,
                                                               const int* __restrict__ mol_nbf,
                                                               const long* __restrict__ poff, double* out0,
                                                               double* out1) {
  const int m = blockIdx.x;
  const int n = mol_nbf[m], c0 = mol_chunk0[m], c1 = mol_chunk0[m + 1];
  for (int e = threadIdx.x; e < nout * n * n; e += 256) {
    const int o = e / (n * n), rem = e % (n * n);
    const int i = rem % n, j = rem / n;
    const int ii = min(i, j), jj = max(i, j);
    const int t = (jj < 16) ? 0 : (ii < 16 ? 1 : 2);
    const int r = ii & 15;
    const int idx = (o * 3 + t) * 256 + (r >> 2) * 64 + (jj & 15) + 16 * (r & 3);
    // eight chunks' partials loaded per round (one memory latency), added in chunk order
    double s = 0.0;
    int ch = c0;
    for (; ch + 8 <= c1; ch += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = part[(long)(ch + u) * nout * 768 + idx];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; ch < c1; ++ch) s += part[(long)ch * nout * 768 + idx];
    (o == 0 ? out0 : out1)[poff[m] + i + (long)j * n] = s;
  }
}

// Per molecule {E_xc, N before correction, N after, E_J = sum w V rho / (2c)}
// from the per-atom sums (atom order).  atom_ej holds sum uJ rho = sum w V rho / 2.
__global__ void k_scalars_batch(const double* __restrict__ atom_exc, const double* __restrict__ atom_ej,
                                const double* __restrict__ qc, const double* __restrict__ qraw,
                                const int* __restrict__ mol_atom0, const double* __restrict__ mol_z, int nmol,
                                double* scal) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  double e = 0.0, ej = 0.0, nc = 0.0;
  for (int a = mol_atom0[m]; a < mol_atom0[m + 1]; ++a) {
    e += atom_exc[a];
    ej += atom_ej[a];
    nc += qc[a];
  }
  scal[4 * m] = e;
  scal[4 * m + 1] = qraw[m];
  scal[4 * m + 2] = nc;
  scal[4 * m + 3] = 2.0 * ej * qraw[m] / (2.0 * mol_z[m]);
}

struct dg_batch {
  dg_handle* h = nullptr;
  ~dg_batch() {
    if (h) {
      if (h->stream) (void)hipStreamSynchronize(h->stream);
      delete h;
    }
  }
};

namespace {

int batch_mol_of_atom(const dg_handle* h, int k) {
  int m = 0;
  while (m + 1 < h->nmol && h->h_mol_atom0[m + 1] <= k) ++m;
  return m;
}

// Work list of the cross-atom potentials: for every atom k, the points of its
// molecule outside its own cell, in ranges of at most len (<= kMaxItem) points,
// k ascending: {k, p_begin, p_end, offset of the item's point order in perm}.
// An item's order holds its points and at most one pad slot per spline
// interval group (nrad + 1 groups); *perm_len receives the total.  A range is
// cut into equal items of at most cap points, cap chosen so that an item's
// padded order never exceeds len slots: the cross kernel walks an item in
// passes of 2048 slots, and an order just over a multiple of that would take
// a nearly empty extra pass.
std::vector<int> cross_items(const dg_handle* h, int len, long* perm_len) {
  std::vector<int> items;
  long off = 0;
  const long cap = std::max<long>(len / 2, len - (h->nrad + 1));
  auto add = [&](int k, long b, long e) {
    const long m = (e - b + cap - 1) / cap;
    for (long r = 0; r < m; ++r) {
      const long p = b + (e - b) * r / m, q = b + (e - b) * (r + 1) / m;
      items.push_back(k);
      items.push_back((int)p);
      items.push_back((int)q);
      items.push_back((int)off);
      off += (q - p) + std::min<long>(q - p, h->nrad + 1);
    }
  };
  for (int k = 0; k < h->natoms; ++k) {
    long p0 = 0, p1 = h->npts;
    if (h->nmol) {
      const int m = batch_mol_of_atom(h, k);
      p0 = (long)h->h_mol_atom0[m] * h->cell;
      p1 = (long)h->h_mol_atom0[m + 1] * h->cell;
    }
    if (p1 - p0 <= h->cell) continue;
    add(k, p0, (long)k * h->cell);
    add(k, (long)(k + 1) * h->cell, p1);
  }
  if (perm_len) *perm_len = off;
  return items;
}

int set_items(dg_handle* h, int len) {
  long plen = 0;
  std::vector<int> items = cross_items(h, len, &plen);
  if (items.size() > h->items_cap || plen > h->npairs)
    return fail(DG_ERR_INVALID, "cross-potential item list exceeds its buffer");
  if (!items.empty())
    HIP_TRY(hipMemcpy(h->items, items.data(), items.size() * 4, hipMemcpyHostToDevice));
  h->nitems = (int)items.size() / 4;
  DG_TRY(sort_items(h));
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const int per_cu = std::max(1, (int)((160 * 1024) / std::max<size_t>(h->cross_lds, 1)));
  h->cross_wgs = std::max(1, std::min(h->nitems, cus * std::min(per_cu, 2)));
  h->item_len = len;
  return DG_OK;
}

std::vector<int> batch_chunks(const dg_handle* h, int len, std::vector<int>* first) {
  std::vector<int> ch;
  first->assign(h->nmol + 1, 0);
  for (int m = 0; m < h->nmol; ++m) {
    (*first)[m] = (int)ch.size() / 3;
    const long p0 = (long)h->h_mol_atom0[m] * h->cell, p1 = (long)h->h_mol_atom0[m + 1] * h->cell;
    for (long p = p0; p < p1; p += len) {
      ch.push_back(m);
      ch.push_back((int)p);
      ch.push_back((int)std::min(p1, p + len));
    }
  }
  (*first)[h->nmol] = (int)ch.size() / 3;
  return ch;
}

int set_bchunks(dg_handle* h, int len) {
  std::vector<int> first;
  std::vector<int> ch = batch_chunks(h, len, &first);
  if (ch.size() > h->bchunks_cap) return fail(DG_ERR_INVALID, "contraction chunk list exceeds its buffer");
  HIP_TRY(hipMemcpy(h->bchunks, ch.data(), ch.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->mol_chunk0, first.data(), first.size() * 4, hipMemcpyHostToDevice));
  h->nbchunks = (int)ch.size() / 3;
  h->bchunk_len = len;
  return DG_OK;
}

constexpr int kChunkCandidates[] = {256, 512, 1024, 2048, 4096};
constexpr int kItemCandidates[] = {1024, 2048, 4096};

int batch_alloc(dg_handle* h, const BatchSpec* bs, const std::vector<int>&) {
  const int nm = bs->nmol;
  for (int m = 0; m < nm; ++m) {
    h->max_mol_atoms = std::max(h->max_mol_atoms, bs->mol_atom0[m + 1] - bs->mol_atom0[m]);
    h->max_mol_shells = std::max(h->max_mol_shells, bs->mol_shell0[m + 1] - bs->mol_shell0[m]);
  }
  DG_TRY(dalloc(&h->atom_mol, h->natoms));
  DG_TRY(dalloc(&h->mol_atom0, nm + 1));
  DG_TRY(dalloc(&h->mol_shell0, nm + 1));
  DG_TRY(dalloc(&h->mol_nbf, nm));
  DG_TRY(dalloc(&h->mol_poff, nm + 1));
  for (double** b : {&h->mol_z, &h->mol_fac, &h->mol_qraw}) DG_TRY(dalloc(b, nm));
  DG_TRY(dalloc(&h->atom_exc, h->natoms));
  DG_TRY(dalloc(&h->atom_ej, h->natoms));
  std::vector<int> first;
  const size_t cap = batch_chunks(h, kChunkCandidates[0], &first).size();
  h->bchunks_cap = cap;
  DG_TRY(dalloc(&h->bchunks, cap));
  DG_TRY(dalloc(&h->mol_chunk0, nm + 1));
  if (h->cpart) {
    HIP_TRY(hipFree(h->cpart));
    h->cpart = nullptr;
  }
  DG_TRY(dalloc(&h->cpart, cap / 3 * 2 * 768));
  return DG_OK;
}

int batch_setup(dg_handle* h) {
  const int nm = h->nmol;
  std::vector<int> amol(h->natoms);
  for (int k = 0; k < h->natoms; ++k) amol[k] = batch_mol_of_atom(h, k);
  std::vector<double> atoms(4 * (size_t)h->natoms), mz(nm, 0.0);
  HIP_TRY(hipMemcpy(atoms.data(), h->atoms, atoms.size() * 8, hipMemcpyDeviceToHost));
  for (int k = 0; k < h->natoms; ++k) mz[amol[k]] += atoms[4 * k + 3];
  std::vector<int> shell0(nm + 1, 0), sat(h->nshell);
  HIP_TRY(hipMemcpy(sat.data(), h->sh_atom, h->nshell * 4, hipMemcpyDeviceToHost));
  for (int s = 0, m = 0; m <= nm; ++m) {
    while (s < h->nshell && amol[sat[s]] < m) ++s;
    shell0[m] = s;
  }
  HIP_TRY(hipMemcpy(h->atom_mol, amol.data(), amol.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->mol_atom0, h->h_mol_atom0.data(), (nm + 1) * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->mol_shell0, shell0.data(), (nm + 1) * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->mol_nbf, h->h_mol_nbf.data(), nm * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->mol_poff, h->h_mol_poff.data(), (nm + 1) * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->mol_z, mz.data(), nm * 8, hipMemcpyHostToDevice));
  DG_TRY(set_bchunks(h, h->bchunk_len));
  DG_TRY(set_items(h, h->item_len));
  const size_t lds = (size_t)h->max_mol_atoms * 64 * 8;
  if (lds > 160 * 1024) return fail(DG_ERR_UNSUPPORTED, "dg_batch_create: too many atoms in a molecule");
  HIP_TRY(hipFuncSetAttribute((const void*)k_becke_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)std::max<size_t>(lds, 1024)));
  const long np = h->npts_pad;
  DG_TRY(run(h, "becke", [&] {
    hipLaunchKernelGGL(k_becke_batch, dim3(nblocks(np, 64)), dim3(64), std::max<size_t>(lds, 1024), h->stream,
                       h->atoms, h->invR, h->natoms, h->npts, h->cell, h->atom_mol, h->mol_atom0, h->px, h->py,
                       h->pz, h->w, h->wb);
  }));
  DG_TRY(run(h, "phi", [&] {
    hipLaunchKernelGGL(k_phi_batch, dim3(nblocks(h->npts * h->max_mol_shells, 256)), dim3(256), 0, h->stream,
                       h->atoms, h->atom_mol, h->mol_shell0, h->max_mol_shells, h->cell, h->sh_atom, h->sh_L,
                       h->sh_nprim, h->sh_off, h->sh_cgf, h->pr_alpha, h->pr_cn0, h->pr_cn1, h->npts, h->px,
                       h->py, h->pz, h->ldphi, h->phi);
  }));
  return DG_OK;
}

int batch_contract(dg_handle* h, double* dJ, double* dXC) {
  DG_TRY(run(h, "contract", [&] {
    hipLaunchKernelGGL(k_contract_batch<2>, dim3(h->nbchunks), dim3(256), 0, h->stream, h->phi, h->uxc, h->uJ,
                       h->bchunks, h->mol_nbf, h->cpart);
  }));
  return run(h, "contract_reduce", [&] {
    hipLaunchKernelGGL(k_contract_batch_reduce, dim3(h->nmol), dim3(256), 0, h->stream, h->cpart, 2,
                       h->mol_chunk0, h->mol_nbf, h->mol_poff, dXC, dJ);
  });
}

// One build of every molecule: the single-molecule stage sequence with the
// per-molecule density, renormalisation, contraction and scalars.
int batch_build(dg_handle* h, const double* dP, double* dJ, double* dXC, double* dscal) {
  const long np = h->npts_pad;
  const int nm = h->nmol;
  DG_TRY(run(h, "pfrag", [&] {
    hipLaunchKernelGGL(k_pfrag_batch, dim3(nblocks(kBatchFrag, 256), nm), dim3(256), 0, h->stream, dP,
                       h->mol_poff, h->mol_nbf, h->pfrag);
  }));
  DG_TRY(run(h, "density", [&] {
    hipLaunchKernelGGL(k_density_batch, dim3(nblocks(np, 64)), dim3(256), 0, h->stream, h->phi, h->pfrag,
                       h->atom_mol, h->mol_nbf, h->cell, h->npts, h->rho);
  }));
  DG_TRY(run(h, "cell_sum", [&] {
    hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->rho, h->w, h->cell, h->q);
  }));
  DG_TRY(run(h, "scale", [&] {
    hipLaunchKernelGGL(k_mol_factors, dim3(nblocks(nm, 64)), dim3(64), 0, h->stream, h->q, h->mol_atom0,
                       h->mol_z, nm, h->mol_fac, h->mol_qraw, h->qc);
    hipLaunchKernelGGL(k_scale_batch, dim3(nblocks(h->npts, 256)), dim3(256), 0, h->stream, h->rho, h->npts,
                       h->cell, h->atom_mol, h->mol_fac);
  }));
  DG_TRY(stage_lda(h));
  DG_TRY(stage_hartree(h));
  DG_TRY(run(h, "atom_sums", [&] {
    hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->w, h->exd, h->cell, h->atom_exc);
    hipLaunchKernelGGL(k_cell_sum, dim3(h->natoms), dim3(1024), 0, h->stream, h->uJ, h->rho, h->cell, h->atom_ej);
  }));
  DG_TRY(batch_contract(h, dJ, dXC));
  return run(h, "scalars", [&] {
    hipLaunchKernelGGL(k_scalars_batch, dim3(nblocks(nm, 64)), dim3(64), 0, h->stream, h->atom_exc, h->atom_ej,
                       h->qc, h->mol_qraw, h->mol_atom0, h->mol_z, nm, dscal);
  });
}

void drop_graph(dg_handle* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  for (auto& k : h->graph_key) k = nullptr;
}

template <typename F>
int time_ms(dg_handle* h, int reps, F body, double* ms) {
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = body();  // warm-up
  if (rc == DG_OK) {
    (void)hipEventRecord(e0, h->stream);
    for (int r = 0; r < reps && rc == DG_OK; ++r) rc = body();
    (void)hipEventRecord(e1, h->stream);
    (void)hipEventSynchronize(e1);
    float f = 0.0f;
    (void)hipEventElapsedTime(&f, e0, e1);
    *ms = f / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

extern "C" {

int dg_batch_create(int nmol, const dg_molecule* mols, const dg_basis* bases, const dg_grid_params* prm,
                    dg_batch** out) {
  if (!mols || !bases || !prm || !out || nmol < 1) return fail(DG_ERR_INVALID, "dg_batch_create: null or empty argument");
  *out = nullptr;
  BatchSpec bs;
  bs.nmol = nmol;
  std::vector<int> Z, atom, L, nprim, offset;
  std::vector<double> xyz, alpha, coef;
  bs.mol_atom0.push_back(0);
  bs.mol_shell0.push_back(0);
  for (int m = 0; m < nmol; ++m) {
    const dg_molecule& mo = mols[m];
    const dg_basis& ba = bases[m];
    if (mo.natoms < 1 || !mo.Z || !mo.xyz) return fail(DG_ERR_INVALID, "dg_batch_create: empty molecule");
    if (ba.nshell < 1 || !ba.atom || !ba.L || !ba.nprim || !ba.offset || !ba.alpha || !ba.coef)
      return fail(DG_ERR_INVALID, "dg_batch_create: empty basis");
    const int a0 = (int)Z.size(), q0 = (int)alpha.size();
    for (int a = 0; a < mo.natoms; ++a) {
      Z.push_back(mo.Z[a]);
      for (int d = 0; d < 3; ++d) xyz.push_back(mo.xyz[3 * a + d]);
    }
    int nq = 0;
    for (int s = 0; s < ba.nshell; ++s) {
      if (ba.atom[s] < 0 || ba.atom[s] >= mo.natoms)
        return fail(DG_ERR_INVALID, "dg_batch_create: shell atom out of range");
      if (ba.nprim[s] < 1 || ba.offset[s] < 0) return fail(DG_ERR_INVALID, "dg_batch_create: bad primitive range");
      atom.push_back(a0 + ba.atom[s]);
      L.push_back(ba.L[s]);
      nprim.push_back(ba.nprim[s]);
      offset.push_back(q0 + ba.offset[s]);
      nq = std::max(nq, ba.offset[s] + ba.nprim[s]);
    }
    for (int i = 0; i < nq; ++i) {
      alpha.push_back(ba.alpha[i]);
      coef.push_back(ba.coef[i]);
    }
    bs.mol_atom0.push_back((int)Z.size());
    bs.mol_shell0.push_back((int)atom.size());
  }
  dg_molecule sm{(int)Z.size(), Z.data(), xyz.data()};
  dg_basis sb{(int)atom.size(), atom.data(), L.data(), nprim.data(), offset.data(), alpha.data(), coef.data()};
  dg_handle* h = nullptr;
  DG_TRY(create_impl(&sm, &sb, prm, &bs, &h));
  auto b = new dg_batch();
  b->h = h;
  *out = b;
  return DG_OK;
}

int dg_batch_destroy(dg_batch* b) {
  delete b;
  return DG_OK;
}

int dg_batch_info(const dg_batch* b, int* nmol, long* npts, int* natoms, long* packed) {
  if (!b) return fail(DG_ERR_INVALID, "dg_batch_info: null batch");
  if (nmol) *nmol = b->h->nmol;
  if (npts) *npts = b->h->npts;
  if (natoms) *natoms = b->h->natoms;
  if (packed) *packed = b->h->packed;
  return DG_OK;
}

int dg_batch_layout(const dg_batch* b, int* nbf, long* offset, long* point0) {
  if (!b) return fail(DG_ERR_INVALID, "dg_batch_layout: null batch");
  const dg_handle* h = b->h;
  for (int m = 0; m < h->nmol; ++m) {
    if (nbf) nbf[m] = h->h_mol_nbf[m];
    if (point0) point0[m] = (long)h->h_mol_atom0[m] * h->cell;
  }
  for (int m = 0; m <= h->nmol; ++m) {
    if (offset) offset[m] = h->h_mol_poff[m];
  }
  if (point0) point0[h->nmol] = h->npts;
  return DG_OK;
}

int dg_batch_handle(dg_batch* b, dg_handle** h) {
  if (!b || !h) return fail(DG_ERR_INVALID, "dg_batch_handle: null argument");
  *h = b->h;
  return DG_OK;
}

int dg_batch_set_stream(dg_batch* b, void* stream) {
  if (!b) return fail(DG_ERR_INVALID, "dg_batch_set_stream: null batch");
  drop_graph(b->h);
  return dg_set_stream(b->h, stream);
}

int dg_batch_build_device(dg_batch* b, const double* dP, double* dJ, double* dXC, double* dscalars) {
  if (!b || !dP || !dJ || !dXC || !dscalars) return fail(DG_ERR_INVALID, "dg_batch_build_device: null argument");
  dg_handle* h = b->h;
  if (h->prof) {  // profiled builds run eagerly so that every kernel is timed
    DG_TRY(batch_build(h, dP, dJ, dXC, dscalars));
  } else {
    const void* key[4] = {dP, dJ, dXC, dscalars};
    if (!h->gexec || !std::equal(key, key + 4, h->graph_key)) {
      drop_graph(h);
      HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
      const int rc = batch_build(h, dP, dJ, dXC, dscalars);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(h->stream, &g);
      if (rc != DG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(DG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      h->graph = g;
      HIP_TRY(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
      std::copy(key, key + 4, h->graph_key);
    }
    HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
  }
  h->has_density = h->has_hartree = h->has_xc = true;
  h->corrected = true;
  return DG_OK;
}

int dg_batch_build(dg_batch* b, const double* P, double* J, double* XC, double* scalars) {
  if (!b || !P) return fail(DG_ERR_INVALID, "dg_batch_build: null argument");
  dg_handle* h = b->h;
  for (long i = 0; i < h->packed; ++i)
    if (!std::isfinite(P[i])) return fail(DG_ERR_INVALID, "dg_batch_build: P contains a non-finite value");
  HIP_TRY(hipMemcpyAsync(h->P, P, (size_t)h->packed * 8, hipMemcpyHostToDevice, h->stream));
  DG_TRY(dg_batch_build_device(b, h->P, h->J, h->XC, h->scalars));
  if (J) HIP_TRY(hipMemcpyAsync(J, h->J, (size_t)h->packed * 8, hipMemcpyDeviceToHost, h->stream));
  if (XC) HIP_TRY(hipMemcpyAsync(XC, h->XC, (size_t)h->packed * 8, hipMemcpyDeviceToHost, h->stream));
  if (scalars) HIP_TRY(hipMemcpyAsync(scalars, h->scalars, (size_t)h->nmol * 4 * 8, hipMemcpyDeviceToHost, h->stream));
  return sync(h);
}

int dg_batch_autotune(dg_batch* b, int* chunk_len, int* item_len, double* best_ms) {
  if (!b) return fail(DG_ERR_INVALID, "dg_batch_autotune: null batch");
  dg_handle* h = b->h;
  if (!h->has_xc) return fail(DG_ERR_STATE, "dg_batch_autotune: run dg_batch_build once first");
  const bool prof = h->prof;
  h->prof = false;
  drop_graph(h);
  double best_c = 1e30, best_i = 1e30;
  int bc = h->bchunk_len, bi = h->item_len;
  for (int len : kChunkCandidates) {
    DG_TRY(set_bchunks(h, len));
    double ms = 0.0;
    DG_TRY(time_ms(h, 5, [&] { return batch_contract(h, h->J, h->XC); }, &ms));
    if (ms < best_c) {
      best_c = ms;
      bc = len;
    }
  }
  DG_TRY(set_bchunks(h, bc));
  for (int len : kItemCandidates) {
    DG_TRY(set_items(h, len));
    double ms = 0.0;
    DG_TRY(time_ms(h, 3, [&] { return stage_hartree(h); }, &ms));
    if (ms < best_i) {
      best_i = ms;
      bi = len;
    }
  }
  DG_TRY(set_items(h, bi));
  h->prof = prof;
  if (chunk_len) *chunk_len = bc;
  if (item_len) *item_len = bi;
  if (best_ms) {
    best_ms[0] = best_c;
    best_ms[1] = best_i;
  }
  return sync(h);
}

}  // extern "C"
