// C++ facade with the reference's MolecularGrid call surface
// (src/moleculargrid.h:48-190) over the dftgrid C ABI.  Header-only; link
// against libdftgrid.so.  Errors are thrown as std::runtime_error, as the
// reference does.  Matrices are plain column-major std::vector<double>
// (the reference uses Eigen::MatrixXd, also column-major).
#ifndef DFTCXX_MOLECULAR_GRID_HPP
#define DFTCXX_MOLECULAR_GRID_HPP

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dftgrid.h"

namespace dftcxx {

// Flattened molecule + basis in the reference's CGF order.
struct MoleculeData {
  std::vector<int> Z;
  std::vector<double> xyz;  // bohr, natoms x 3
  std::vector<int> shell_atom, shell_L, shell_nprim, shell_offset;
  std::vector<double> alpha, coef;
  int natoms() const { return (int)Z.size(); }
  int nbf() const {
    int n = 0;
    for (int L : shell_L) n += (L + 1) * (L + 2) / 2;
    return n;
  }
  int nelec() const {
    int n = 0;
    for (int z : Z) n += z;
    return n;
  }
  dg_molecule c_molecule() const { return dg_molecule{natoms(), Z.data(), xyz.data()}; }
  dg_basis c_basis() const {
    return dg_basis{(int)shell_atom.size(), shell_atom.data(), shell_L.data(), shell_nprim.data(),
                    shell_offset.data(), alpha.data(), coef.data()};
  }
};

inline void dg_throw(int rc) {
  if (rc != DG_OK) throw std::runtime_error(std::string("dftgrid: ") + dg_last_error());
}

// Settings + Molecule of a reference input file (src/settings.cpp:30-187,
// src/molecule.cpp:49-238): the molecule, its basis from basis_dir/<basis>.dat,
// and the grid parameters the file sets.
struct InputData {
  MoleculeData mol;
  dg_grid_params grid{15, 110, 8, DG_PREC_FP64, 0.0};
  int lebedev_order = 7;
  int hartree = DG_HARTREE_BECKE;
  std::string name, basis;
};

inline InputData read_input(const std::string& path, const std::string& basis_dir) {
  dg_input* in = nullptr;
  dg_throw(dg_input_read(path.c_str(), basis_dir.c_str(), &in));
  InputData d;
  dg_molecule m;
  dg_basis b;
  const char* v = nullptr;
  int rc = dg_input_molecule(in, &m);
  if (rc == DG_OK) rc = dg_input_basis(in, &b);
  if (rc == DG_OK) rc = dg_input_grid(in, &d.grid, &d.lebedev_order, &d.hartree);
  if (rc == DG_OK) {
    d.mol.Z.assign(m.Z, m.Z + m.natoms);
    d.mol.xyz.assign(m.xyz, m.xyz + 3 * m.natoms);
    int nprim = 0;
    for (int s = 0; s < b.nshell; ++s) nprim = std::max(nprim, b.offset[s] + b.nprim[s]);
    d.mol.shell_atom.assign(b.atom, b.atom + b.nshell);
    d.mol.shell_L.assign(b.L, b.L + b.nshell);
    d.mol.shell_nprim.assign(b.nprim, b.nprim + b.nshell);
    d.mol.shell_offset.assign(b.offset, b.offset + b.nshell);
    d.mol.alpha.assign(b.alpha, b.alpha + nprim);
    d.mol.coef.assign(b.coef, b.coef + nprim);
    rc = dg_input_value(in, "name", &v);
    if (rc == DG_OK) d.name = v;
    if (rc == DG_OK) rc = dg_input_value(in, "basis", &v);
    if (rc == DG_OK) d.basis = v;
  }
  dg_input_destroy(in);
  dg_throw(rc);
  return d;
}

// Reference lebedev_order enum index (src/quadrature.h:30-44) -> point count.
inline int lebedev_points_from_index(unsigned idx) {
  static const int n[11] = {6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194};
  if (idx >= 11) throw std::runtime_error("lebedev_order index out of range");
  return n[idx];
}

class MolecularGrid {
 public:
  explicit MolecularGrid(const MoleculeData& mol) : mol_(mol) {}
  ~MolecularGrid() {
    if (h_) dg_destroy(h_);
  }
  MolecularGrid(const MolecularGrid&) = delete;
  MolecularGrid& operator=(const MolecularGrid&) = delete;

  // src/moleculargrid.h:104 -- lebedev_order is the reference's enum index
  void set_grid_parameters(unsigned radial_points, unsigned lebedev_order, unsigned lmax) {
    set_grid_points(radial_points, lebedev_points_from_index(lebedev_order), lmax);
  }
  // the same with the angular rule given by its point count (e.g. 302)
  void set_grid_points(unsigned radial_points, unsigned angular_points, unsigned lmax,
                       int precision = DG_PREC_FP64, double screen = 0.0) {
    prm_ = dg_grid_params{(int)radial_points, (int)angular_points, (int)lmax, precision, screen};
  }
  // src/moleculargrid.h:118
  void create_grid() {
    if (h_) throw std::runtime_error("grid already created");
    dg_molecule m = mol_.c_molecule();
    dg_basis b = mol_.c_basis();
    dg_throw(dg_create(&m, &b, &prm_, &h_));
    long np = 0;
    dg_throw(dg_info(h_, &nbf_, &np, nullptr));
    npts_ = np;
  }
  // src/moleculargrid.h:131
  void set_density(const std::vector<double>& P) { dg_throw(dg_set_density(h_, P.data(), nbf_)); }
  // src/moleculargrid.h:167
  void correct_densities() { dg_throw(dg_correct_densities(h_, nullptr)); }
  // src/moleculargrid.h:95
  double calculate_density() {
    double n = 0.0;
    dg_throw(dg_calculate_density(h_, &n));
    return n;
  }
  // src/moleculargrid.h:138
  std::vector<double> calculate_hartree_potential() {
    std::vector<double> J((size_t)nbf_ * nbf_);
    dg_throw(dg_hartree(h_, J.data()));
    return J;
  }
  // DFT::calculate_exchange_correlation_matrix, src/dft.cpp:394-433
  std::vector<double> calculate_exchange_correlation(double* exc) {
    std::vector<double> XC((size_t)nbf_ * nbf_);
    dg_throw(dg_xc(h_, XC.data(), exc));
    return XC;
  }
  // Derivatives of the last E_xc (no counterpart in the reference): dE/dP
  // (column-major nbf x nbf) and dE/dR (atom-major xyz); hold_c keeps the
  // renormalisation factor fixed.  Either pointer may be null.
  void xc_adjoint(bool hold_c, std::vector<double>* dEdP, std::vector<double>* grad) {
    if (dEdP) dEdP->assign((size_t)nbf_ * nbf_, 0.0);
    if (grad) grad->assign((size_t)3 * mol_.natoms(), 0.0);
    dg_throw(dg_xc_adjoint(h_, hold_c ? 1 : 0, dEdP ? dEdP->data() : nullptr, grad ? grad->data() : nullptr));
  }
  // src/moleculargrid.h:146-162
  std::vector<double> get_weights() { return vec(dg_get_weights, npts_); }
  std::vector<double> get_densities() { return vec(dg_get_densities, npts_); }
  std::vector<double> get_amplitudes() { return vec(dg_get_amplitudes, npts_ * nbf_); }
  // per point the nbf x 3 column-major basis_func_grad (src/gridpoint.cpp:62-72)
  std::vector<double> get_amplitude_gradients() { return vec(dg_get_amplitude_gradients, npts_ * 3 * nbf_); }
  // per point grad rho (src/gridpoint.cpp:94-108), npts x 3
  std::vector<double> get_density_gradients() { return vec(dg_get_density_gradients, npts_ * 3); }

  dg_handle* handle() { return h_; }
  int nbf() const { return nbf_; }
  long npts() const { return npts_; }

 private:
  std::vector<double> vec(int (*fn)(dg_handle*, double*), long n) {
    std::vector<double> v(n);
    dg_throw(fn(h_, v.data()));
    return v;
  }
  MoleculeData mol_;
  dg_grid_params prm_{15, 110, 8, DG_PREC_FP64, 0.0};  // reference "medium" grid
  dg_handle* h_ = nullptr;
  int nbf_ = 0;
  long npts_ = 0;
};

// The reference's RectangularGrid (src/rectangulargrid.h:30-80,
// src/rectangulargrid.cpp:34-95): a cube of dp^3 points spanning
// [-size/2, size/2] in each direction (x fastest, then y, then z), the density
// and its gradient at every point, and the gradient file.
class RectangularGrid {
 public:
  explicit RectangularGrid(const MoleculeData& mol) : mol_(mol) {}
  void build_grid(double size, unsigned dp) {
    if (dp < 2) throw std::runtime_error("RectangularGrid: at least two points per direction");
    const double gd = size / (double)(dp - 1);
    pos_.clear();
    for (unsigned i = 0; i < dp; ++i)
      for (unsigned j = 0; j < dp; ++j)
        for (unsigned k = 0; k < dp; ++k) {
          pos_.push_back((double)k * gd - size / 2.0);
          pos_.push_back((double)j * gd - size / 2.0);
          pos_.push_back((double)i * gd - size / 2.0);
        }
  }
  void set_density(const std::vector<double>& P) {
    const long n = (long)(pos_.size() / 3);
    rho_.assign(n, 0.0);
    grad_.assign(3 * n, 0.0);
    dg_molecule m = mol_.c_molecule();
    dg_basis b = mol_.c_basis();
    dg_throw(dg_density_at_points(&m, &b, n, pos_.data(), P.data(), rho_.data(), grad_.data()));
  }
  // one line "x y z gx gy gz" per point, %12.8f (src/rectangulargrid.cpp:82-95)
  void write_gradient(const std::string& filename) const {
    FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) throw std::runtime_error("RectangularGrid: cannot write " + filename);
    for (size_t i = 0; i < pos_.size() / 3; ++i)
      std::fprintf(f, "%12.8f  %12.8f  %12.8f  %12.8f  %12.8f  %12.8f\n", pos_[3 * i], pos_[3 * i + 1],
                   pos_[3 * i + 2], grad_[3 * i], grad_[3 * i + 1], grad_[3 * i + 2]);
    std::fclose(f);
  }
  const std::vector<double>& positions() const { return pos_; }
  const std::vector<double>& densities() const { return rho_; }
  const std::vector<double>& gradients() const { return grad_; }

 private:
  MoleculeData mol_;
  std::vector<double> pos_, rho_, grad_;
};

}  // namespace dftcxx

#endif
