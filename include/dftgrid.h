/*
 * dftgrid -- C ABI of the MI355X grid-integration library.
 *
 * One handle owns the device state of one molecular grid: points, weights,
 * basis amplitudes, density, Hartree potential and the matrices built from
 * them.  It is the drop-in for the reference's MolecularGrid (src/moleculargrid.h)
 * together with the exchange-correlation assembly of DFT
 * (src/dft.cpp:394-433).  The reference has no C or FFI interface of its own;
 * each entry point below names the reference member function whose result it
 * reproduces.
 *
 * Conventions
 *  - Every function returns DG_OK (0) or an error code; dg_last_error() gives
 *    a message for the calling thread.  No C++ exception crosses the ABI.
 *  - Matrices are nbf x nbf, column-major (as the reference's Eigen
 *    matrices), double precision.  P is the reference's half density matrix
 *    (sum over occupied orbitals of C C^T, src/dft.cpp:345-352).
 *  - Coordinates are in bohr.
 *  - Grid points are ordered atom by atom (input order), then radial node
 *    p = 1..N (r decreasing), then Lebedev point (src/atomicgrid.cpp:50-82).
 *  - The basis is a list of shells in the reference's CGF order
 *    (src/molecule.cpp:135-238); shell L = 0, 1, 2 expands to 1, 3, 6 CGFs in
 *    the order s; px, py, pz; dxx, dxy, dxz, dyy, dyz, dzz (src/cgf.h:208-221).
 *  - A handle is not thread-safe.  All its work runs on one HIP stream (its own
 *    by default, or one set with dg_set_stream).  "_device" entry points take
 *    device pointers and enqueue work without synchronising; the host-pointer
 *    entry points synchronise before they return.
 */
#ifndef DFTGRID_H
#define DFTGRID_H

#ifdef __cplusplus
extern "C" {
#endif

#define DG_OK 0
#define DG_ERR_INVALID 1  /* bad argument or size */
#define DG_ERR_HIP 2      /* a HIP runtime call failed */
#define DG_ERR_ALLOC 3    /* device or host allocation failed */
#define DG_ERR_STATE 4    /* call out of order (e.g. hartree before set_density) */
#define DG_ERR_UNSUPPORTED 5

#define DG_PREC_FP64 0  /* every stage in FP64 */
#define DG_PREC_MIXED 1 /* FP64 density/potentials/energies, FP32-MFMA Fock contractions */

typedef struct dg_handle dg_handle;

typedef struct {
  int natoms;
  const int* Z;       /* [natoms] nuclear charges (1..18) */
  const double* xyz;  /* [natoms][3] positions in bohr */
} dg_molecule;

typedef struct {
  int nshell;
  const int* atom;      /* [nshell] centre atom index */
  const int* L;         /* [nshell] 0, 1 or 2 */
  const int* nprim;     /* [nshell] primitives per shell */
  const int* offset;    /* [nshell] first primitive in alpha/coef */
  const double* alpha;  /* primitive exponents */
  const double* coef;   /* contraction coefficients (unnormalised primitives) */
} dg_basis;

typedef struct {
  int n_radial;   /* Gauss-Chebyshev radial nodes per atom */
  int n_angular;  /* Lebedev points per shell (a point count, e.g. 110, 302) */
  int lmax;       /* spherical-harmonic cut-off of the Hartree expansion */
  int precision;  /* DG_PREC_FP64 or DG_PREC_MIXED */
  /* Amplitude screening (no counterpart in the reference, which sums every
   * basis function at every point): 0 keeps every product; screen > 0 skips,
   * in the density and in the Fock contractions, each 64-point block's 16-wide
   * groups of basis functions whose |phi| stays at or below screen on the
   * whole block.  Single-molecule handles only. */
  double screen;
} dg_grid_params;

/* Message of the last failing call on this thread ("" if none). */
const char* dg_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int dg_version(void);
/* 1 if a Lebedev rule with that many points is compiled in, else 0. */
int dg_lebedev_available(int npoints);

/* Build the molecular grid on the current HIP device: points, Becke weights
 * and basis amplitudes.  Replaces MolecularGrid(mol) + set_grid_parameters()
 * + create_grid() (src/moleculargrid.h:84,104,118; src/moleculargrid.cpp:175-261). */
int dg_create(const dg_molecule* mol, const dg_basis* basis, const dg_grid_params* params,
              dg_handle** out);
int dg_destroy(dg_handle* h);
int dg_info(const dg_handle* h, int* nbf, long* npts, int* natoms);
/* Use this HIP stream (a hipStream_t) for all later work; NULL = own stream. */
int dg_set_stream(dg_handle* h, void* stream);
int dg_get_stream(const dg_handle* h, void** stream);

/* --- host-pointer API, one call per reference member function ---------- */
/* MolecularGrid::set_density(P) (src/moleculargrid.cpp:48-53): rho = 2 phi^T P phi. */
int dg_set_density(dg_handle* h, const double* P, int nbf);
/* MolecularGrid::correct_densities() (src/moleculargrid.cpp:132-146): scale rho to
 * sum(Z) electrons.  nelec_before (may be NULL) receives the integral before. */
int dg_correct_densities(dg_handle* h, double* nelec_before);
/* MolecularGrid::calculate_density() (src/moleculargrid.cpp:157-166). */
int dg_calculate_density(dg_handle* h, double* nelec);
/* MolecularGrid::calculate_hartree_potential() (src/moleculargrid.cpp:336-389): J. */
int dg_hartree(dg_handle* h, double* J);
/* DFT::calculate_exchange_correlation_matrix() (src/dft.cpp:394-433): XC and E_xc. */
int dg_xc(dg_handle* h, double* XC, double* exc);
/* Getters (src/moleculargrid.cpp:61-127): weights include the Becke factor. */
int dg_get_weights(dg_handle* h, double* w);            /* [npts] */
int dg_get_becke_weights(dg_handle* h, double* wb);     /* [npts] */
int dg_get_positions(dg_handle* h, double* xyz);        /* [npts][3] */
int dg_get_densities(dg_handle* h, double* rho);        /* [npts] */
int dg_get_amplitudes(dg_handle* h, double* phi);       /* [npts][nbf] */
int dg_get_hartree_potential(dg_handle* h, double* V);  /* [npts], after dg_hartree */
/* Hartree intermediates after dg_hartree: rho_lm, U_lm [natoms][nrad][nlm]
 * (src/atomicgrid.cpp:241-296, 301-432) and the own-cell potential [npts]
 * (src/atomicgrid.cpp:434-462).  Any pointer may be NULL. */
int dg_get_hartree_parts(dg_handle* h, double* rho_lm, double* U_lm, double* vcell);
/* Pointwise XC after dg_xc: energy density e_x+e_c and w*(v_xa+v_xb+v_ca+v_cb)/2 [npts]. */
int dg_get_xc_parts(dg_handle* h, double* exc_density, double* wv);

/* --- basis-function and density gradients ---------------------------------- */
/* Gradients of every basis function at every grid point, [npts][3][nbf]: per
 * point the nbf x 3 column-major matrix GridPoint::basis_func_grad holds
 * (GridPoint::set_basis_func_grad, src/gridpoint.cpp:62-72).  They are the
 * derivatives of the amplitude formula; the reference's GTO::get_grad
 * (src/cgf.cpp:67-94) leaves out the normalisation, applies the contraction
 * coefficient twice and drops the factor l of x^(l-1), and is not followed. */
int dg_get_amplitude_gradients(dg_handle* h, double* dphi);
/* Density gradient [npts][3] of the last evaluated density,
 * grad rho = 2 phi^T P grad phi + 2 grad phi^T P phi (GridPoint::set_gradient,
 * src/gridpoint.cpp:94-108), times the renormalisation factor once the density
 * is corrected (so that it is the gradient of dg_get_densities' rho).  Needs the
 * density matrix still in the device buffer it was read from. */
int dg_get_density_gradients(dg_handle* h, double* grad);
/* Density and density gradient at arbitrary points, the work of the reference's
 * RectangularGrid (src/rectangulargrid.cpp:34-95: build_grid, set_density and
 * GridPoint::set_density/set_gradient): xyz [npts][3] in bohr, P the
 * column-major nbf x nbf density matrix (host pointers); rho [npts] and
 * grad [npts][3] (either may be NULL).  Runs on the current HIP device and
 * synchronises. */
int dg_density_at_points(const dg_molecule* mol, const dg_basis* basis, long npts, const double* xyz,
                         const double* P, double* rho, double* grad);
/* The same evaluation without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that
 * runs beside HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its work buffers live in a 32 MiB array in device memory that
 * the library holds on every device it runs on until the process ends; the
 * call allocates nothing on the host heap either on its success path.  Same
 * arguments and results as dg_density_at_points; bases of at most 512
 * functions, 512 shells, 4096 primitives and 1024 atoms (DG_ERR_UNSUPPORTED
 * beyond), points in passes that fit the array.  The array keeps the last
 * call's contents; calls are serialised within the process (calls that
 * overlapped on one device would race on its array).  Runs on the calling
 * thread's default stream (hipStreamPerThread) and synchronises. */
int dg_density_at_points_noalloc(const dg_molecule* mol, const dg_basis* basis, long npts, const double* xyz,
                                 const double* P, double* rho, double* grad);

/* --- adjoint of E_xc ------------------------------------------------------ */
/* No counterpart in the reference: its GTO gradient (src/cgf.cpp:67-94,
 * 164-172) is never called on this path.  Differentiates the E_xc of the
 * last dg_xc (or build), E_xc = sum_p w_p e(c rho_raw_p) with the
 * renormalisation c = sum Z / sum_p w_p rho_raw_p of src/moleculargrid.cpp:132-146.
 *   hold_c = 1: c is a constant, and dEdP = 2 c XC;
 *   hold_c = 0: c varies with P and R (the derivative of the E_xc the SCF reports).
 * dEdP [nbf*nbf] column-major; grad [natoms*3] = dE_xc/dR_A (atom-major xyz;
 * the force is -grad), including the basis-centre, grid-motion and Becke-weight
 * terms.  Shells out of reach of a point (exp(-alpha r^2) below e^-100 for
 * their most diffuse primitive) are left out of both, which is below the
 * rounding of the sums.  Needs a corrected density (dg_correct_densities or a build) and the
 * density matrix of that evaluation still in the device buffer it was read from.
 * Either output may be NULL.  Returns DG_ERR_STATE before dg_xc. */
int dg_xc_adjoint(dg_handle* h, int hold_c, double* dEdP, double* grad);

/* --- device-pointer API --------------------------------------------------- */
/* One full grid build for one SCF iteration, enqueued on the handle's stream:
 * set_density(P) + correct_densities() + calculate_hartree_potential() +
 * the XC assembly.  dP, dJ, dXC are device arrays nbf*nbf (column-major);
 * dscalars (device, 4 doubles) receives {E_xc, N_elec before correction,
 * N_elec after correction, E_J = 2 tr(P J) evaluated pointwise as
 * sum_p w_p V_p rho_p / (2c)}.  Any output pointer may be NULL. */
int dg_build_device(dg_handle* h, const double* dP, double* dJ, double* dXC, double* dscalars);

/* The same build for an SCF step that needs only the two-electron part of the
 * Fock matrix F = H + 2J + XC (src/dft.cpp:334): dG (device, nbf*nbf) receives
 * G = 2J + XC from one contraction, Phi^T diag(w V + w v_xc) Phi, instead of
 * two; dscalars as dg_build_device (E_J evaluated pointwise).  No counterpart
 * in the reference, which forms J and XC separately. */
int dg_build_fock_device(dg_handle* h, const double* dP, double* dG, double* dscalars);

/* The build without the Hartree stages, for the SCF whose J comes from the
 * two-electron integrals (dg_eri_*): rho, its renormalisation, the functional
 * and XC (device, nbf*nbf); dscalars as dg_build_device with E_J = 0. */
int dg_build_xc_device(dg_handle* h, const double* dP, double* dXC, double* dscalars);

/* --- analytic two-electron integrals ------------------------------------------ */
/* The reference's other Hartree path (hartree_evaluation = two_electron_integrals,
 * src/dft.cpp:378-385): every unique (ij|kl) over the basis, computed once on
 * the current HIP device and kept there, packed at
 * IJ(IJ+1)/2 + KL with IJ = i(i+1)/2 + j, i >= j, IJ >= KL
 * (DFT::calculate_two_electron_integrals and Integrator::teindex,
 * src/dft.cpp:254-288, src/integrals.cpp:470-486). */
typedef struct dg_eri dg_eri;
int dg_eri_create(const dg_molecule* mol, const dg_basis* basis, dg_eri** out);
int dg_eri_destroy(dg_eri* e);
int dg_eri_count(const dg_eri* e, long* nints, int* nbf);
/* the packed integrals, nints doubles (host) */
int dg_eri_get(dg_eri* e, double* ints);
/* DFT::calculate_hartree_potential_te_int (src/dft.cpp:452-477):
 * J_ij = sum_kl P_kl (ij|kl), device pointers (nbf*nbf, column-major),
 * enqueued on stream (a hipStream_t; NULL = the object's own). */
int dg_eri_coulomb_device(dg_eri* e, const double* dP, double* dJ, void* stream);
/* the same from and to host buffers (synchronous) */
int dg_eri_coulomb(dg_eri* e, const double* P, double* J);

/* --- profiling ------------------------------------------------------------- */
/* When enabled, every kernel of dg_build_device is bracketed by HIP events on
 * the handle's stream and its time accumulated per kernel name. */
int dg_profile_enable(dg_handle* h, int enable);
/* Copies up to max entries: name (NUL-terminated, <= 47 chars) into names[i*48],
 * total milliseconds and launch count.  *n receives the number of kernels. */
int dg_profile_read(dg_handle* h, int max, char* names, double* total_ms, int* launches, int* n);
int dg_profile_reset(dg_handle* h);

/* Checks the lane layout of the FP64 16x16x4 matrix-core instruction the
 * contraction kernels rely on; *ok = 1 if it matches. */
int dg_selftest_mfma(int* ok);
/* Runs the FP64 16x16x4 MFMA once on one wave: lane i supplies a[i] and b[i];
 * d[v*64 + i] receives result register v of lane i (diagnostics). */
int dg_mfma_probe(const double* a, const double* b, double* d);
/* The same for the FP32 16x16x4 MFMA of the reduced-precision contractions:
 * lane i supplies a[i] and b[i]; d[v*64 + i] receives result register v. */
int dg_mfma_probe_f32(const float* a, const float* b, float* d);

/* --- SCF driver (host C++, src/dft.cpp:84-142, 160-227, 300-366, 441-447) --- */
#define DG_STEP_AUTO 0   /* device step for bases of at least DG_STEP_AUTO_NBF functions */
#define DG_STEP_HOST 1   /* density-matrix step on the host (eigen-solver) */
#define DG_STEP_DEVICE 2 /* density-matrix step and energy on the GPU */
#define DG_STEP_AUTO_NBF 128

typedef struct {
  double threshold;  /* stop when |dE| <= threshold (after min_iter) */
  int min_iter;      /* reference: 3 */
  int max_iter;      /* reference: 100 */
  double mixing;     /* reference: 0.5 (weight of the old P) */
  /* Where the density-matrix step runs (no counterpart in the reference, which
   * diagonalises on the host): DG_STEP_HOST diagonalises X^T F X on the host
   * and copies P to and J, XC from the device each iteration; DG_STEP_DEVICE
   * keeps F, P and the energy on the GPU and forms P as the projector onto
   * the occupied space by second-order spectral projection (the same P up to
   * rounding; steps without a HOMO-LUMO gap fall back to the host). */
  int density_step;
  /* DG_HARTREE_BECKE (the Becke-grid Poisson path) or DG_HARTREE_ERI (J from
   * the analytic two-electron integrals, hartree_evaluation =
   * two_electron_integrals; the grid then serves the XC part only). */
  int hartree;
} dg_scf_params;

typedef struct {
  int iterations;
  double energy;        /* total */
  double e_one;         /* 2 tr(P H) */
  double e_j;           /* 2 tr(P J) */
  double e_xc;
  double e_nuc;
  double nelec;         /* grid electron count after the last iteration */
  double seconds_setup; /* integrals + grid construction */
  double seconds_iter;  /* mean wall seconds per SCF iteration */
  int converged;        /* 1 if |dE| <= threshold was reached after min_iter, 0 if the
                         * loop stopped at max_iter instead (the reference prints
                         * "Stopping because maximum number of iterations has been
                         * reached", src/dft.cpp:126-131) */
  double delta;         /* |dE| of the last iteration */
  int device_step;      /* 1 if the density-matrix steps ran on the GPU */
  int host_fallbacks;   /* device steps redone on the host (no gap) */
} dg_scf_result;

/* Runs the reference SCF with the grid work on the GPU.  energies (may be NULL)
 * receives the total energy of each iteration (max_iter doubles); P_out (may
 * be NULL) the final density matrix.  DG_ERR_INVALID for max_iter < 1,
 * min_iter < 0, a negative or non-finite threshold, or mixing outside [0, 1). */
int dg_scf_run(const dg_molecule* mol, const dg_basis* basis, const dg_grid_params* grid,
               const dg_scf_params* scf, dg_scf_result* result, double* energies, double* P_out);

/* --- batches of molecules ----------------------------------------------------- */
/* No counterpart in the reference, which runs one molecule per process
 * (src/dftcxx.cpp:53-54).  A batch holds many independent molecules (each at
 * most 32 CGFs) on one grid specification (FP64 only) and builds J, XC and the
 * scalars of all of them in one hipGraph launch per SCF iteration.  P, J and XC
 * are packed: molecule m's nbf_m x nbf_m column-major matrix starts at
 * offset[m] (dg_batch_layout).  Each molecule's results equal those of a
 * single-molecule handle up to summation order. */
typedef struct dg_batch dg_batch;
int dg_batch_create(int nmol, const dg_molecule* mols, const dg_basis* bases, const dg_grid_params* prm,
                    dg_batch** out);
int dg_batch_destroy(dg_batch* b);
/* molecules, total grid points, total atoms, packed matrix length (any may be NULL) */
int dg_batch_info(const dg_batch* b, int* nmol, long* npts, int* natoms, long* packed);
/* per molecule: nbf [nmol], packed offsets [nmol+1], first grid point [nmol+1] (any may be NULL) */
int dg_batch_layout(const dg_batch* b, int* nbf, long* offset, long* point0);
/* The underlying grid handle, for the getters (weights, densities, ...) and
 * profiling; its single-molecule build calls return DG_ERR_STATE. */
int dg_batch_handle(dg_batch* b, dg_handle** h);
int dg_batch_set_stream(dg_batch* b, void* stream);
/* One build of every molecule on device buffers (packed dP, dJ, dXC; dscalars
 * 4 doubles per molecule: E_xc, N before correction, N after, E_J), enqueued
 * as one graph launch on the batch's stream; the graph is captured on the first
 * call with a given set of pointers and replayed afterwards. */
int dg_batch_build_device(dg_batch* b, const double* dP, double* dJ, double* dXC, double* dscalars);
/* The same from and to host buffers (synchronous); any output may be NULL. */
int dg_batch_build(dg_batch* b, const double* P, double* J, double* XC, double* scalars);
/* Times the candidate contraction chunk lengths and cross-potential item
 * lengths on the batch (after one build) and keeps the fastest of each;
 * best_ms (2 doubles, may be NULL) receives their times. */
int dg_batch_autotune(dg_batch* b, int* chunk_len, int* item_len, double* best_ms);

/* Runs the reference SCF (dg_scf_run) for every molecule of a batch, all of
 * them advancing together with one dg_batch_build per iteration; each molecule
 * stops at its own convergence (its later builds are discarded).  results[m]
 * per molecule (seconds_iter is the batch's wall time per iteration);
 * iterations (may be NULL) receives the number of batch iterations. */
int dg_batch_scf_run(int nmol, const dg_molecule* mols, const dg_basis* bases, const dg_grid_params* grid,
                     const dg_scf_params* scf, dg_scf_result* results, int* iterations);

/* --- reference input files (host only, no GPU needed) ---------------------- */
#define DG_HARTREE_BECKE 0 /* Becke-grid Poisson path (the reference's default) */
#define DG_HARTREE_ERI 1   /* "hartree_evaluation = two_electron_integrals" */

/* A parsed input file: settings, atoms and the expanded basis set. */
typedef struct dg_input dg_input;
/* Reads a reference input file (molecules/<name>.in) and the basis set
 * <basis_dir>/<basis>.dat it names.  Replaces Settings(filename) and
 * Molecule(filename, settings) (src/settings.cpp:30-187, src/molecule.cpp:49-238):
 * "key = value" lines up to the line "system:", then the atom count and
 * "Element x y z" lines (angstrom converted with 1.889725989 when
 * units = angstrom, else bohr); grid = coarse | medium (default) | fine |
 * ultrafine sets radial points, Lebedev rule and lmax, which radial_points,
 * lebedev_order (the reference's index 0..10) and lmax override.  The basis
 * file is read up to the molecule's highest element; the shells come out in
 * the reference's CGF order.  Errors: DG_ERR_INVALID with the reference's
 * message ("Could not find name", "Unknown element: X", "Cannot open ...!"),
 * and also for an unknown grid name and for an element without basis
 * functions (where the reference runs on with an unset grid or no functions). */
int dg_input_read(const char* path, const char* basis_dir, dg_input** out);
int dg_input_destroy(dg_input* in);
/* Views into the parsed input, valid until dg_input_destroy. */
int dg_input_molecule(const dg_input* in, dg_molecule* mol);
int dg_input_basis(const dg_input* in, dg_basis* basis);
/* Grid parameters (precision FP64, no screening), the Lebedev index and the
 * Hartree method (DG_HARTREE_BECKE or DG_HARTREE_ERI); any output may be NULL. */
int dg_input_grid(const dg_input* in, dg_grid_params* grid, int* lebedev_index, int* hartree);
/* Settings::get_value (src/settings.cpp:82-89): the trimmed value of a key, or
 * DG_ERR_INVALID "Could not find <key>". */
int dg_input_value(const dg_input* in, const char* key, const char** value);

/* --- host utilities (no GPU needed) ----------------------------------------- */
/* Overlap, kinetic and nuclear-attraction matrices (nbf x nbf, column-major)
 * of the basis, as DFT::construct_matrices builds them (src/dft.cpp:184-201,
 * src/integrals.cpp:43-180).  Any output may be NULL. */
int dg_one_electron(const dg_molecule* mol, const dg_basis* basis, double* S, double* T, double* V);
/* Eigenvalues (ascending) and column-major eigenvectors of a symmetric
 * column-major n x n matrix (stands in for Eigen::SelfAdjointEigenSolver). */
int dg_eigh(int n, const double* A, double* w, double* Z);

#ifdef __cplusplus
}
#endif
#endif /* DFTGRID_H */
