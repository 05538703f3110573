// Drives the reference-shaped C++ facade (include/dftcxx/molecular_grid.hpp)
// through the call sequence DFT makes on MolecularGrid (src/dft.cpp:57-61,
// 362-365, 395-397, 483), for a test to compare with the ctypes path.
// Built into tests/facade/libfacade_probe.so by the Makefile beside it.
#include <cstring>
#include <exception>
#include <string>

#include "dftcxx/molecular_grid.hpp"

namespace {
std::string g_err;
}

extern "C" const char* facade_probe_error() { return g_err.c_str(); }

// Molecule and basis as flat arrays; lebedev_order is the reference's enum
// index.  Outputs: J and XC (nbf x nbf column-major), scalars {E_xc, N before
// correction, N from calculate_density after it}.  Returns 0 or -1 (message in
// facade_probe_error).
extern "C" int facade_probe(int natoms, const int* Z, const double* xyz, int nshell, const int* shell_atom,
                            const int* shell_L, const int* shell_nprim, const int* shell_offset,
                            const double* alpha, const double* coef, int nprim, int radial, int lebedev_order,
                            int lmax, const double* P, double* J, double* XC, double* scalars) {
  try {
    dftcxx::MoleculeData mol;
    mol.Z.assign(Z, Z + natoms);
    mol.xyz.assign(xyz, xyz + 3 * natoms);
    mol.shell_atom.assign(shell_atom, shell_atom + nshell);
    mol.shell_L.assign(shell_L, shell_L + nshell);
    mol.shell_nprim.assign(shell_nprim, shell_nprim + nshell);
    mol.shell_offset.assign(shell_offset, shell_offset + nshell);
    mol.alpha.assign(alpha, alpha + nprim);
    mol.coef.assign(coef, coef + nprim);
    dftcxx::MolecularGrid grid(mol);
    grid.set_grid_parameters(radial, lebedev_order, lmax);
    grid.create_grid();
    const int n = grid.nbf();
    grid.set_density(std::vector<double>(P, P + (size_t)n * n));
    double before = 0.0;
    dftcxx::dg_throw(dg_correct_densities(grid.handle(), &before));
    const std::vector<double> j = grid.calculate_hartree_potential();
    double exc = 0.0;
    const std::vector<double> xc = grid.calculate_exchange_correlation(&exc);
    std::memcpy(J, j.data(), j.size() * sizeof(double));
    std::memcpy(XC, xc.data(), xc.size() * sizeof(double));
    scalars[0] = exc;
    scalars[1] = before;
    scalars[2] = grid.calculate_density();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
