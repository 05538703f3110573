// Device-resident density-matrix step and energy of the SCF (scf_device.hip).
#ifndef DFTGRID_SCF_DEVICE_H
#define DFTGRID_SCF_DEVICE_H

#include <hip/hip_runtime.h>

namespace dftcxx {

// Device state of one molecule's SCF: H, X, the Fock and density matrices and
// the grid results J, XC and scalars, all n x n column-major on the device.
// A build writes J, XC and the scalars (dJ(), dXC(), dscal()); density() forms
// P from them, energy() evaluates the energy of P, finish() waits and returns
// the energy terms and whether the projection converged.
class DeviceStep {
 public:
  static constexpr int kMaxSteps = 100;  // SP2 steps launched per density step (converged ones return at once)

  DeviceStep(int n, int nocc, hipStream_t stream);
  ~DeviceStep();
  DeviceStep(const DeviceStep&) = delete;
  DeviceStep& operator=(const DeviceStep&) = delete;

  void upload_core(const double* H, const double* X);
  // F = H + 2J (+ XC when with_xc), or F = H + G when fused (dXC() holding
  // G = 2J + XC); P <- projector of X^T F X, mixed into P unless first.
  // Returns false, with P untouched, when the projection does not converge
  // (no HOMO-LUMO gap); the caller then forms this step on the host.
  bool density(bool with_xc, bool first, double mixing, bool fused);
  // energy of the current P with the last build's J and scalars (asynchronous)
  void energy(double enuc, bool with_xc, bool point_ej);
  // waits for the stream; the energy terms of energy()
  void finish(double* e_one, double* e_j, double* e_xc, double* total, double* nelec);
  // XC = 0 (a set-up build that formed XC the reference does not use yet)
  void clear_xc();
  void download(double* dst, const double* src);
  void upload(double* dst, const double* src);

  double* dP() { return P_; }
  double* dJ() { return J_; }
  double* dXC() { return XC_; }  // XC, or G = 2J + XC of a fused build
  double* dscal() { return scal_; }

 private:
  int n_ = 0, nocc_ = 0;
  hipStream_t st_ = nullptr;
  double* buf_ = nullptr;
  void* ctl_ = nullptr;
  double* host_ = nullptr;  // pinned: energy terms [0..4], SP2 control block from [8]
  double *H_ = nullptr, *X_ = nullptr, *F_ = nullptr, *Fp_ = nullptr, *Pn_ = nullptr, *P2_ = nullptr;
  double *P_ = nullptr, *J_ = nullptr, *XC_ = nullptr, *scal_ = nullptr, *eout_ = nullptr;
  double *epart_ = nullptr, *glo_ = nullptr, *ghi_ = nullptr;  // energy partial sums, Gershgorin bounds
};

// The same step for every molecule of a batch (each at most 32 functions), one
// work-group per molecule; matrices packed at off[m] as the batch builds them.
// A build writes dJ(), dXC(), dscal(); energy() then gives, per molecule,
// {E_one, E_J, E_xc, total, N} (energies()) and whether its last projection
// converged (failed()).  Inactive molecules (set_active) keep P and results.
class BatchStep {
 public:
  BatchStep(int nmol, const int* nbf, const long* off, const int* nocc, hipStream_t stream);
  ~BatchStep();
  BatchStep(const BatchStep&) = delete;
  BatchStep& operator=(const BatchStep&) = delete;

  void upload_core(const double* H, const double* X, const double* enuc);  // packed H, X; E_nuc per molecule
  void set_active(const char* active);                                     // nmol flags
  // returns the number of molecules whose projection did not converge (P_m
  // untouched, failed()[m] = 1); waits for the stream
  int density(bool with_xc, bool first, double mixing);
  void clear_xc();  // XC = 0 (a set-up build formed an XC the reference does not use yet)
  void energy(bool with_xc);
  void finish();  // waits; energies() are then current
  const double* energies() const { return host_; }
  const int* failed() const { return reinterpret_cast<const int*>(host_ + 5 * (size_t)nmol_); }
  void download(double* dst, const double* src, size_t n);
  void upload(double* dst, const double* src, size_t n);

  double* dP() { return P_; }
  double* dJ() { return J_; }
  double* dXC() { return XC_; }
  double* dscal() { return scal_; }

 private:
  int nmol_ = 0;
  long packed_ = 0;
  hipStream_t st_ = nullptr;
  double* buf_ = nullptr;
  void* ibuf_ = nullptr;
  double* host_ = nullptr;
  double *H_ = nullptr, *X_ = nullptr, *P_ = nullptr, *J_ = nullptr, *XC_ = nullptr;
  double *scal_ = nullptr, *eout_ = nullptr, *enuc_ = nullptr;
  long* off_ = nullptr;
  int *nbf_ = nullptr, *nocc_ = nullptr, *fail_ = nullptr;
  char* active_ = nullptr;
};

}  // namespace dftcxx

#endif
