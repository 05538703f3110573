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
  // Asynchronous; the projection's outcome comes with finish().
  void density(bool with_xc, bool first, double mixing, bool fused);
  // energy of the current P with the last build's J and scalars (asynchronous)
  void energy(double enuc, bool with_xc, bool point_ej);
  // waits for the stream; energy terms of energy() and the state of the last density()
  void finish(double* e_one, double* e_j, double* e_xc, double* total, double* nelec, bool* sp2_ok,
              int* sp2_steps);
  void download(double* dst, const double* src);
  void upload(double* dst, const double* src);

  double* dP() { return P_; }
  double* dJ() { return J_; }
  double* dXC() { return XC_; }  // XC, or G = 2J + XC of a fused build
  double* dscal() { return scal_; }
  double* dPold() { return Pold_; }  // P before the last density()

 private:
  int n_ = 0, nocc_ = 0;
  hipStream_t st_ = nullptr;
  double* buf_ = nullptr;
  void* ctl_ = nullptr;
  double* host_ = nullptr;  // pinned: energy terms [0..4], SP2 control block from [8]
  double *H_ = nullptr, *X_ = nullptr, *F_ = nullptr, *Fp_ = nullptr, *Pn_ = nullptr, *P2_ = nullptr;
  double *P_ = nullptr, *J_ = nullptr, *XC_ = nullptr, *Pold_ = nullptr, *scal_ = nullptr, *eout_ = nullptr;
};

}  // namespace dftcxx

#endif
