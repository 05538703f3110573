/*
 * FP64 CPU restatement of the reference's grid-integration path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Tests, the smoke check and the CPU-baseline leg
 * of bench.py load this library to check and time against; the product
 * library (libdftgrid) never links, loads or calls it.
 *
 * Every function restates a reference routine, cited as file:line relative to
 * the reference source tree, keeping its loop structure, evaluation order and
 * OpenMP placement so that it also serves as the reference-faithful CPU
 * baseline.  Arrays are row-major unless stated; N x N matrices are symmetric
 * or stored column-major as the reference's Eigen matrices are.
 */
#ifndef DFTCXX_ORACLE_H
#define DFTCXX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_grid or_grid;

/* --- grid object: MolecularGrid + AtomicGrid + GridPoint ------------------ */
/* Build the molecular grid (src/moleculargrid.cpp:193-261): per atom a
 * Gauss-Chebyshev x Lebedev grid with amplitudes, then Becke weights.
 * leb is nang x 4 (x, y, z, w), weights summing to 1. */
or_grid* or_grid_create(int natoms, const int* Z, const double* xyz,
                        int nshell, const int* shell_atom, const int* shell_L,
                        const int* shell_nprim, const int* shell_offset,
                        const double* alpha, const double* coef,
                        int nrad, int nang, const double* leb, int lmax);
/* The same grid with points in only the first ncell cells (every atom still
 * enters the Becke partition, the Poisson solves and the cross-atom
 * potential); the per-iteration calls then do the work of those cells, and
 * the renormalisation uses their atoms' charge.  For CPU timing samples. */
or_grid* or_grid_create_cells(int natoms, const int* Z, const double* xyz,
                              int nshell, const int* shell_atom, const int* shell_L,
                              const int* shell_nprim, const int* shell_offset,
                              const double* alpha, const double* coef,
                              int nrad, int nang, const double* leb, int lmax, int ncell);
void or_grid_destroy(or_grid* g);
int or_grid_nbf(const or_grid* g);
long or_grid_npts(const or_grid* g);
/* positions (npts x 3), total weights and Becke weights (npts) */
void or_grid_points(const or_grid* g, double* xyz, double* w, double* wb);
/* amplitudes, npts x nbf (point-major) */
void or_grid_amplitudes(const or_grid* g, double* phi);
/* radial nodes r_n (nrad), decreasing */
void or_grid_radial(const or_grid* g, double* r);
void or_grid_set_density(or_grid* g, const double* P);
void or_grid_correct_densities(or_grid* g);
double or_grid_calculate_density(or_grid* g);
void or_grid_densities(const or_grid* g, double* rho);
/* per-atom cell populations q_n = sum w rho */
void or_grid_cell_densities(or_grid* g, double* q);
/* Hartree (src/moleculargrid.cpp:336-389), J is nbf x nbf */
void or_grid_hartree(or_grid* g, double* J);
/* after or_grid_hartree: rho_lm and U_lm (natoms x nrad x nlm), V_cell and V (npts),
 * spline coefficients (natoms x nlm x (nrad-1) x 4); any pointer may be NULL */
void or_grid_hartree_parts(const or_grid* g, double* rho_lm, double* U_lm,
                           double* vcell, double* v, double* spl);
/* wall seconds of the last or_grid_hartree's stages: projection rho_lm,
 * Poisson solves + own-cell potential + splines, cross-atom potential, J */
void or_grid_hartree_times(const or_grid* g, double* t);
/* exchange-correlation matrix and energy (src/dft.cpp:394-433) */
void or_grid_xc(or_grid* g, double* XC, double* exc);
/* pointwise LDA outputs of the last or_grid_xc: ex+ec (energy density) and
 * wva = w (vxa+vxb+vca+vcb)/2 */
void or_grid_xc_parts(const or_grid* g, double* exc_density, double* wva);

/* --- primitives ----------------------------------------------------------- */
double or_gto_norm(double alpha, int l, int m, int n);
double or_sh_prefactor(int l, int m);
double or_spherical_harmonic(int l, int m, double pole, double azimuth);
/* pre * SH of the direction of (x, y, z), angles as the reference derives them */
double or_ylm_xyz(int l, int m, double x, double y, double z);
double or_becke_cutoff(double mu);
void or_xalpha(long n, const double* rhoa, const double* rhob, double* ex, double* vxa, double* vxb);
void or_vwn(long n, const double* rhoa, const double* rhob, double* ec, double* vca, double* vcb);
/* not-a-knot spline: coef is (n-1) x 4 */
int or_spline_build(int n, const double* x, const double* y, double* coef);
double or_spline_eval(int n, const double* x, const double* y, const double* coef, double xx);
/* radial Poisson: FD matrix M_l ((N+2)^2, column-major) and the solved U for rhs rho_lm */
void or_poisson_matrix(int N, int l, const double* r, double* M);
void or_lu_solve(int n, const double* A, const double* b, double* x);

/* --- one-electron integrals, ERIs, eigensolver, SCF (src/integrals.cpp, src/dft.cpp) */
void or_one_electron(int natoms, const int* Z, const double* xyz,
                     int nshell, const int* shell_atom, const int* shell_L,
                     const int* shell_nprim, const int* shell_offset,
                     const double* alpha, const double* coef,
                     double* S, double* T, double* V);
double or_boys(double m, double x);
/* J_ij = sum_kl P_kl (ij|kl) from analytic ERIs (src/dft.cpp:452-477) */
void or_eri_J(int natoms, const double* xyz,
              int nshell, const int* shell_atom, const int* shell_L,
              const int* shell_nprim, const int* shell_offset,
              const double* alpha, const double* coef,
              const double* P, double* J);
/* symmetric eigensolver, ascending eigenvalues, column-major eigenvectors */
void or_eigh(int n, const double* A, double* evals, double* evecs);
/* SCF (src/dft.cpp:84-142, 160-227, 330-366, 441-447).  Returns iterations.
 * energies[k] is the total energy after iteration k (k < max_iter);
 * terms = {E_total, E_one, E_J, E_xc, E_nuc}; P_out nbf x nbf. */
int or_scf(or_grid* g, int natoms, const int* Z, const double* xyz,
           int nshell, const int* shell_atom, const int* shell_L,
           const int* shell_nprim, const int* shell_offset,
           const double* alpha, const double* coef,
           double threshold, int min_iter, int max_iter, double mixing,
           double* energies, double* terms, double* P_out);
/* The same SCF with J from the two-electron integrals (or_eri_J each iteration),
 * the reference's hartree_evaluation = two_electron_integrals (src/dft.cpp:378-385). */
int or_scf_eri(or_grid* g, int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
               const int* shell_L, const int* shell_nprim, const int* shell_offset, const double* alpha,
               const double* coef, double threshold, int min_iter, int max_iter, double mixing,
               double* energies, double* terms, double* P_out);
/* One evaluation of the SCF's grid work and energy at a given P, by the same
 * code or_scf runs each iteration: set_density, correct_densities (or, with
 * hold_c > 0, rho scaled by hold_c instead), Hartree J, XC, calculate_energy.
 * terms = {E_total, E_one, E_J, E_xc, E_nuc, c}; F = H + 2J + XC
 * (src/dft.cpp:334).  Any output may be NULL. */
int or_scf_evaluate(or_grid* g, int natoms, const int* Z, const double* xyz,
                    int nshell, const int* shell_atom, const int* shell_L,
                    const int* shell_nprim, const int* shell_offset,
                    const double* alpha, const double* coef,
                    const double* P, double hold_c, double* terms,
                    double* F, double* H, double* J, double* XC);

/* --- rectangular grid and density gradients (src/rectangulargrid.cpp:34-95,
 *     src/gridpoint.cpp:45-108) ------------------------------------------ */
void or_rectangular_points(double size, unsigned dp, double* xyz);
void or_points_density(int natoms, const double* xyz, int nshell, const int* shell_atom, const int* shell_L,
                       const int* shell_nprim, const int* shell_offset, const double* alpha, const double* coef,
                       long npts, const double* pts, const double* P, double* rho, double* grad, double* dphi);

#ifdef __cplusplus
}
#endif
#endif
