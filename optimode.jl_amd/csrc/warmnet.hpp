// The learned warm start of solve_k: a small fully convolutional network that maps a
// downsampled eps^-1 map plus omega to per-band effective-index guesses and low-resolution
// start vectors.  It has no counterpart in the reference (OptiMode starts every solve from
// random vectors and k_guess, src/MaxwellEigenmodes.jl:63); it is trained on the GPU by the
// kernels in warmnet.hip on modes the solver itself converged.
//
// Tensors are FP32, [batch][channel][y][x] with y, x in 0..hw-1 (the convolutions may run in
// bf16 on the matrix cores, set_bf16).  Network (3x3 convolutions, zero padding, stride 1):
//   c1: 7 -> 16 at hw, ReLU                         (input: 6 eps^-1 components + omega)
//   pool: 2x2 average                               -> hw/2
//   c2: 16 -> 32 at hw/2, ReLU
//   pool                                            -> hw/4
//   c3: 32 -> 32 at hw/4, ReLU
//   head: mean over pixels of c3 (32) -> dense -> nev            (effective-index guesses)
//   c4: [up2(c3) | c2] 64 -> 32 at hw/2, ReLU       (nearest-neighbour upsampling)
//   c5: [up2(c4) | c1] 48 -> 32 at hw, ReLU
//   c6: 32 -> 4 nev at hw (linear)                  (start vectors: re/im of 2 components)
// Loss: mean over field entries of (y - T)^2 + wk * mean over bands of (k - Tk)^2.
#pragma once

#include <cstdint>
#include <vector>

#include "solver.hpp"

namespace omx {

// Input and output convention of the warm-start network (the shipped network was trained
// with it): input channels (eps^-1 block mean - WS_DIAG0) / WS_DIAG_SCALE for xx, yy, zz,
// block mean / WS_OFF_SCALE for xy, xz, yz, (omega - WS_OMEGA0) / WS_OMEGA_SCALE; effective
// index of band j = WS_N_SHIFT + WS_N_SCALE * k_j.
constexpr double WS_DIAG0 = 0.55, WS_DIAG_SCALE = 0.25, WS_OFF_SCALE = 0.05;
constexpr double WS_OMEGA0 = 0.65, WS_OMEGA_SCALE = 0.05;
constexpr double WS_N_SHIFT = 1.8, WS_N_SCALE = 0.2;

class WarmNet {
 public:
  static constexpr int NIN = 7;
  WarmNet(int device, int hw, int nev, uint64_t seed);
  ~WarmNet();
  WarmNet(const WarmNet&) = delete;
  WarmNet& operator=(const WarmNet&) = delete;
  int hw() const { return hw_; }
  int nev() const { return nev_; }
  long nparams() const { return nparams_; }
  void get_params(float* out);
  void set_params(const float* in);
  // x: B x 7 x hw x hw -> fields B x 4nev x hw x hw (may be NULL), k B x nev (may be NULL)
  void forward(int B, const float* x, float* fields, float* kout);
  // loss and its gradient with respect to the flat parameter vector (grad may be NULL)
  double loss_grad(int B, const float* x, const float* tf, const float* tk, float wk, float* grad);
  // Adam over n samples held on the device, minibatches of `batch` in a seeded shuffle;
  // hist[e] = mean minibatch loss of epoch e
  // (resume: continue the previous call's Adam moments and step count instead of resetting them)
  void train(int n, const float* x, const float* tf, const float* tk, float wk, int epochs, int batch, float lr,
             uint64_t seed, double* hist, bool resume = false);
  int device() const { return device_; }
  // convolutions (forward and data gradient) in bf16 on the matrix cores with FP32
  // accumulation (true) or FP32 on the vector units (false); the weight gradient, the
  // parameters and Adam stay FP32 either way
  void set_bf16(bool on) { bf16_ = on; }
  bool bf16() const { return bf16_; }
  // The guess for one 2D cell on the device: features from eps6 (FP64 per pixel xx yy zz xy
  // xz yz, pixel ix + nx iy), one forward, the band outputs to k_host (nev floats), and the
  // start vectors' hw x hw lowest plane waves (inverse 2D DFT of the field maps, unit norm per
  // band; entry c + 2(gx + hw gy) of band j at j 2 hw^2), held in the network's buffer whose
  // device address is returned (valid until the next call).
  const c64* guess_device(const double* eps6, int nx, int ny, double omega, float* k_host);

 private:
  struct Conv {
    int ci, co, res;
    long w, b;  // offsets into the parameter vector
  };
  int device_, hw_, nev_;
  bool bf16_ = false;
  long nparams_ = 0;
  Conv c_[6];
  long dw_ = 0, db_ = 0;  // dense head
  hipStream_t s_ = nullptr;
  DevMem P_, G_, M1_, M2_;  // parameters, gradient, Adam moments
  int B_ = 0;               // batch capacity of the activation buffers
  DevMem x_, a1_, p1_, a2_, p2_, a3_, g_, k_, cat4_, a4_, cat5_, a5_, y_;
  DevMem dy_, da5_, dcat5_, da4_, dcat4_, da3_, dp2_, da2_, dp1_, da1_, dk_, dg_, part_, lossd_;
  DevMem dft_, low_;  // inverse-DFT intermediate and result of guess_device
  long adam_t_ = 0;

  void ensure_batch(int B);
  void fwd(int B);
  // gradient of the loss for targets already on the device; accumulates the loss sum into
  // lossd_[slot] (device double)
  void bwd(int B, const float* tf_dev, const float* tk_dev, float wk, int slot);
  void conv(const Conv& c, const float* in, float* out, int B, bool relu);
  void conv_bwd(const Conv& c, const float* in, const float* dout, float* din, int B);
};

// solve_k from the network's guess, all of it on the device but nev floats: the shared first
// eigensolve starts at k0_scale times the predicted band-1 k from the predicted block (its
// hw x hw lowest plane waves scattered into the LOBPCG block); stopping rules and results
// are those of Solver::solve_k.
void solve_k_learned(Solver& S, WarmNet& net, double omega, int nev, double k0_scale, double k_tol, double eig_tol,
                     int max_eigsolves, int maxiter, double* kmags, double* evecs, omx_info* info);

}  // namespace omx
