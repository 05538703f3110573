// C++ facade with the reference's MolecularGrid call surface
// (src/moleculargrid.h:48-190) over the dftgrid C ABI.  Header-only; link
// against libdftgrid.so.  Errors are thrown as std::runtime_error, as the
// reference does.  Matrices are plain column-major std::vector<double>
// (the reference uses Eigen::MatrixXd, also column-major).
#ifndef DFTCXX_MOLECULAR_GRID_HPP
#define DFTCXX_MOLECULAR_GRID_HPP

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

}  // namespace dftcxx

#endif
