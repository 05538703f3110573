// Allocation-free variant of dg_density_at_points (included at the end of
// dftgrid.hip, after the point-evaluation helpers it shares).
namespace {

// Device scratch of dg_density_at_points_noalloc: 32 MiB on every device the
// library runs on, held until the process ends.  In host code the array's name
// is only a stand-in; its device address comes from hipGetSymbolAddress, once
// per device (scratch_address).
constexpr long kScratchDoubles = 4L << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;           // guards the tables and the addresses below
PointTables g_noalloc_tables;      // static storage: no heap in the call
double* g_scratch_ptr[kMaxDevices];  // device address of scratch on each device

// Device address of the scratch array on the current device (g_noalloc_mu held).
int scratch_address(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DG_ERR_UNSUPPORTED, "scratch: device index beyond 63");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return DG_OK;
}

}  // namespace

extern "C" {

int dg_density_at_points_noalloc(const dg_molecule* mol, const dg_basis* basis, long npts, const double* xyz,
                                 const double* P, double* rho, double* grad) {
  g_last_error.clear();
  const char* who = "dg_density_at_points_noalloc";
  DG_TRY(point_check(mol, basis, npts, xyz, P, who));
  if (npts == 0 || (!rho && !grad)) return DG_OK;
  // one call at a time in this process: the host tables and each device's
  // scratch array are shared by all calls
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  PointTables& t = g_noalloc_tables;
  DG_TRY(point_tables(mol, basis, t, who));
  double* base = nullptr;
  DG_TRY(scratch_address(&base));
  // carve the scratch array: fixed tables, then the per-pass buffers (256-byte aligned)
  long off = 0;
  auto take = [&](long n) {
    double* q = base + off;
    off += round_up(n, 32);
    return q;
  };
  PointWork w;
  w.atoms = take(4L * t.na);
  w.alpha = take(t.nprim);
  w.cn0 = take(t.nprim);
  w.cn1 = take(t.nprim);
  w.ishell = reinterpret_cast<int*>(take((5L * t.nshell + 1) / 2));
  w.P = take((long)t.nbf * t.nbf);
  w.frag = take(256L * t.nt * t.nt);
  const long per_point = 2L * t.ld + 10;  // phi, T, xyz (3), x y z (3), rho, grad (3)
  long C = std::min(kPointChunk, (kScratchDoubles - off - 8 * 32) / per_point);
  C = std::min(C, round_up(npts, 64)) / 64 * 64;
  if (C < 64)
    return fail(DG_ERR_UNSUPPORTED, std::string(who) + ": the basis needs more than the " +
                                        std::to_string(kScratchDoubles * 8 / (1 << 20)) + " MiB scratch array");
  w.chunk = C;
  w.xyz = take(3 * C);
  w.pxyz = take(3 * C);
  w.phi = take(C * t.ld);
  w.T = take(C * t.ld);
  w.rho = take(C);
  w.grad = take(3 * C);
  if (off > kScratchDoubles) return fail(DG_ERR_UNSUPPORTED, std::string(who) + ": scratch layout overflow");
  const int rc = point_run(t, basis, P, npts, xyz, rho, grad, w, hipStreamPerThread);
  (void)hipStreamSynchronize(hipStreamPerThread);
  return rc;
}

}  // extern "C"
