// Host-side engine behind the omx C ABI: device buffers, operator plans, the batched
// LOBPCG driver and the Newton solve_k.
#pragma once
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "omx.h"
#include "omx_common.hpp"
#include "omx_errors.hpp"
#include "operator.hpp"
#include "block.hpp"
#include "dense.hpp"

namespace omx {

// Owning device allocation.
class DevMem {
 public:
  DevMem() = default;
  explicit DevMem(size_t bytes) { alloc(bytes); }
  ~DevMem() { release(); }
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  DevMem(DevMem&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevMem& operator=(DevMem&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes) OMX_HIP_CHECK(hipMalloc(&p_, bytes));
    n_ = bytes;
  }
  void ensure(size_t bytes) {
    if (bytes > n_) alloc(bytes);
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename U> U* as() const { return static_cast<U*>(p_); }
  size_t size() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

template <typename T> struct PrecisionData {
  OpPlan<T> plan{};
  DevMem eps6, eps6c, twx, twy, twz;
  int reg_ok = 0;   // the grid admits the register-FFT operator
  int slab_ok = 0;  // 3D cells: x-columns per block of the yz-slab path (0: five passes only)
};

struct SolveStats {
  int newton = 0;
  int it32 = 0;
  int it64 = 0;
  long matvec = 0;
  long precon = 0;
  double max_residual = 0;
  double device_ms = 0;
};

// Live timing of one kernel family on the solver's stream (HIP events), with the
// algorithmic byte count of each launch.
struct KernelStat {
  double ms = 0;
  long launches = 0;
  double bytes = 0;
};

// Page-locked host memory (hipHostMalloc) for small device-to-host snapshots.
class PinnedMem {
 public:
  PinnedMem() = default;
  ~PinnedMem() {
    if (p_) (void)hipHostFree(p_);
  }
  PinnedMem(const PinnedMem&) = delete;
  PinnedMem& operator=(const PinnedMem&) = delete;
  void alloc(size_t bytes) {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    if (bytes) OMX_HIP_CHECK(hipHostMalloc(&p_, bytes, hipHostMallocDefault));
  }
  template <typename U> U* as() const { return static_cast<U*>(p_); }

 private:
  void* p_ = nullptr;
};

class Profiler {
 public:
  bool enabled = false;
  KernelStat op32, op64;
  void record(hipStream_t s, KernelStat* st, double bytes, hipEvent_t* out_pair);
  void flush();
  void reset() {
    flush();
    op32 = KernelStat{};
    op64 = KernelStat{};
  }
  ~Profiler();

 private:
  struct Pending {
    hipEvent_t a, b;
    KernelStat* st;
    double bytes;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  hipEvent_t take();

 public:
  hipEvent_t begin(hipStream_t s);
  void end(hipStream_t s, hipEvent_t a, KernelStat* st, double bytes);
  // Two recorded events whose interval is attributed later, once its byte count is known.
  void commit(hipEvent_t a, hipEvent_t b, KernelStat* st, double bytes) { pending_.push_back(Pending{a, b, st, bytes}); }
  // Return two recorded events unused (an interval that did no work).
  void recycle(hipEvent_t a, hipEvent_t b) {
    pool_.push_back(a);
    pool_.push_back(b);
  }
};

struct Options {
  double fp32_tol_rel = 2e-3;   // solve_k: FP32 LOBPCG locking threshold of the Newton phase, relative
                                // to omega^2 (the FP64 refinement takes the vectors from there)
  double fp32_tol_omega2 = 1e-5;  // solve_omega2: FP32 locking threshold before its FP64 LOBPCG
  double fp32_switch = 1e-6;    // |Δω²|/ω² below which Newton moves to the FP64 phase
  double seed_tol = 0;          // locking threshold / omega^2 of the shared first eigensolve
                                // (0: 1e-1 on 2D cells in the two-stream FP32 phase of a single
                                // solve, 5e-2 in batches and the other paths, 3e-3 on 3D cells,
                                // as measured)
  int maxiter_fp64 = 200;       // LOBPCG iteration cap in the FP64 phase
  int guard = 2;                // block columns of band i: min(nev, i + 1 + guard); < 0: nev
  uint64_t seed = 0x5eed;
  int fp64_only = 0;            // skip the FP32 phase
  int reg_ops = 1;              // register-FFT operator where the grid admits it
  int reg_rows32 = 0;           // register path, FP32: rows per block of passes 1/3 (16 or 8; 0: by
                                // launch size, 8 for M/P launches of few vectors, else 16)
  int reg_rows64 = 0;           // register path, FP64: the same with 8 or 4
  int reg_nt2 = 0;              // register path: threads per block of pass 2 (256 or 128; 0: by launch size)
  int slab3d = 1;               // 3D cells: the yz-slab operator where the cell admits it
  int plain = 0;                // the reference's algorithm: FP64 only, nev-column block per band,
                                // random start at k_guess per band, every eigensolve to eig_tol
  int warm_start = 0;           // solve_k without kguess/Hguess starts every band from the previous
                                // solve's k and eigenvectors (kept on the device)
  int band_guess = 0;           // solve_k with kguess (one k per band) and Hguess: every band starts at
                                // its own guess with the Hguess block, no shared first eigensolve
                                // (the learned warm start)
  int guess_lowres = 0;         // > 0: solve_k's Hguess vectors hold only their guess_lowres^2
                                // lowest plane waves (2 hw^2 complex each), scattered on the device
  int refine = 1;               // FP64 phase by Rayleigh-Ritz + Jacobi-Davidson refinement of each
                                // band's vector (0: FP64 LOBPCG on the whole block)
  int fp32_newton = 1;          // FP32 Newton steps before the FP64 phase (0: the FP64 phase starts
                                // from the shared first FP32 eigensolve)
  double jd_rtol = 5e-3;        // relative tolerance of the refinement's correction solves
  int jd_maxit = 60;            // iteration cap of one correction solve
  int jd_fp32 = 1;              // correction solves with FP32 vectors and operator (FP64 recurrences)
  int jd_deflate = 1;           // FP32 correction solves on the complement of the band's whole block
                                // (not only of its vector): the neighbouring bands leave the spectrum
  double jd_first_tol = 3e-5;   // residual (relative to omega^2) at which the refinement's first
                                // Newton step stops correcting and moves k (0: at the eigensolve
                                // tolerance, as the later steps): Newton needs rho, whose error is
                                // the residual squared over the band gap
  double jac_stop32 = 1e-4;     // Jacobi stop (max |sin| of a sweep) of the FP32 batches' Rayleigh-Ritz
                                // (off-diagonal left ~1e-8 |A|, below FP32 data; FP64 batches: 1e-9)
  int gram_blocks = 0;          // Gram products: about this many row-chunk blocks over the active
                                // instances (0: 1024 on 2D cells, 512 on 3D cells, as measured)
  int combine_blocks = 0;       // block updates: about this many row blocks (0: 2048 on 2D cells,
                                // 1024 on 3D cells)
  // FP32 operator plan of 2D cells on the LDS Stockham path (picked by omx_autotune):
  int fft_variant32 = 0;        // 0: radix-8 passes first (8, 8, 4 for 256); 1: radix-4 passes
                                // where the power of two is even (4, 4, 4, 4 for 256)
  int tile_rows32 = 0;          // rows per block of passes 1 and 3 (0: the largest that fits)
  int tile_cols32 = 0;          // columns per block of pass 2 (0: the largest that fits)
  int streams2 = 1;             // FP32 Newton phase of solve_k as two half-batches of bands on two
                                // streams in lockstep (one half's latency-bound dense steps overlap
                                // the other half's memory-bound kernels)
  int graphs = 0;               // MINRES iterations replayed from a captured HIP graph (groups of 12;
                                // same results, no measured gain: the gaps are GPU-side)
};

// Device-resident EME mode set: E and H of nev modes, planar per mode and component
// ((v * 3 + c) * N + p, p = ix + Nx iy), complex FP64.
constexpr int MODE_SLOTS = 4;
struct ModeSlot {
  int nev = 0;
  DevMem E, H;
  // for the pull-back (omx_mode_pullback): the eigenvectors (nev columns of ld), their k,
  // the norms the fields were divided by, the inner-product convention, and the field
  // cotangents accumulated by omx_overlap_pullback (planar as E and H; bar_live = 0: zero)
  DevMem ev, kd, Ebar, Hbar;
  std::vector<c64> nrm;
  int conj = 0;
  int normalized = 1;
  int bar_live = 0;
};

// Buffers of the FP64 eigenpair refinement, kept on the handle between solves (grown on
// demand, so that a repeated solve allocates and frees nothing there).
struct RefineWorkspace {
  int ninst = 0, nblk_g = 0, nblk_r = 0, dps = 0;
  long ld = 0;
  DevMem n1d, zero, cold, lamd, Tm, T2, ATm, XB, RBm, part, Gd, Bd, Cd, normp, scl, dots;
  DevMem XBf, RBf, Tf;  // FP32 copies for the correction solves (option jd_fp32)
  DevMem Xf, nqd;       // FP32 copy of the block and its column counts (option jd_deflate)
};

// One problem of a batched solve on a handle's grid (Solver::solve_k_multi).
struct MultiProblem {
  const double* epsinv;  // (3,3,Nx,Ny) column-major
  double omega;
  int nev;
  double k_tol, eig_tol;
  int max_eigsolves, maxiter;
  double* kmags;         // out: nev
  double* evecs;         // out: NULL or nev vectors of 2N complex
  omx_info info;         // out
  int status;            // out
};

class Solver {
 public:
  Solver(int device, int nx, int ny, double Dx, double Dy, const double* epsinv);
  // 3D cell periodic along z (Grid{3}): nz planes, period Dz; epsinv is (3,3,Nx,Ny,Nz)
  Solver(int device, int nx, int ny, int nz, double Dx, double Dy, double Dz, const double* epsinv);
  ~Solver();

  void set_epsinv(const double* epsinv);
  void apply(OpKind op, int prec, double k, const double* Hin, double* Hout, int m, double* dots);
  void solve_omega2(double k, int nev, double tol, int maxiter, double* H_io, double* w2_out, omx_info* info);
  void solve_k(double omega, int nev, double k_tol, double eig_tol, int max_eigsolves, int maxiter,
               const double* kguess, const double* Hguess, int nguess, double* kmags, double* evecs, omx_info* info);
  void get_evecs(double* evecs);
  // solve_k for several problems on this handle's grid (own eps^-1, omega, nev and tolerances
  // each) as one batch: every band of every problem is an instance of the same LOBPCG,
  // Newton and refinement launches (2D cells)
  void solve_k_multi(std::vector<MultiProblem>& probs);
  void hmh(double k, const double* A, const double* H, int m, double* out);
  void group_index(double omega, int nev, const double* kmags, const double* evecs, const double* deps_domega,
                   double* ng_out);
  // Lbar: NULL, or (3D cells) the period cotangent of solve_k_periodic
  void adjoint(double omega, int nev, const double* kmags, const double* evecs, const double* kbar,
               const double* evbar, const uint8_t* evbar_mask, double* omegabar, double* epsinvbar,
               double* Lbar = nullptr);
  void ng_gvd(double omega, int nev, const double* kmags, const double* evecs, const double* deps_domega,
              const double* d2eps_domega2, double* ng_out, double* gvd_out, double* E_out);
  void build_modes(int slot, int nev, const double* kmags, const double* evecs, int conjugate, int normalize,
                   double* norms_out);
  void overlap(int slot_a, int slot_b, int conjugate, double* out);
  void get_mode_fields(int slot, double* E_out, double* H_out);
  // pull-backs of overlap_matrix and build_mode (eme.hip)
  void overlap_pullback(int slot_a, int slot_b, int conjugate, const double* obar, int which);
  void mode_pullback(int slot, double* kbar_out, double* evbar_out, double* epsinvbar_out);
  void set_option(const std::string& key, double value);
  // timed trials of the operator path, FFT radix variant, Stockham tile shape and the Gram /
  // block-update partitions for nev-mode solves; the fastest become the handle's options
  void autotune(int nev, int* choice_out, double* us_out);
  // halves: the single solve's FP32 phase on two streams (its Newton steps make up for a
  // looser seed; the batched launch groups wait for their slowest problem and do not)
  double seed_tol_rel(bool halves = false) const {
    return opt.seed_tol > 0 ? opt.seed_tol : (nz > 1 ? 3e-3 : (halves ? 1e-1 : 5e-2));
  }
  // per mode j: dw2dk[j] = 2 HMkH(ev_j) and dw2dL[j] = 2 <ev_j| dM/dDz |ev_j> (3D cells)
  void period_derivatives(int nev, const double* kmags, const double* evecs, double* dw2dk, double* dw2dL);
  void require_2d(const char* what) const;

  std::mutex mu;
  std::string last_error;

  // shared state, public for the templated engine
  int device;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // second stream of the two-stream FP32 Newton phase
  int nx, ny;
  double Dx, Dy;
  int nz = 1;       // planes along z (1: a 2D cell)
  double Dz = 0.0;  // period along z (3D cells)
  long N, K, ld;
  PrecisionData<float> p32;
  PrecisionData<double> p64;
  Options opt;
  // set only during solve_k_learned: the start vectors' lowest plane waves already on the
  // device (used in place of Hguess when guess_lowres > 0)
  const c64* guess_dev = nullptr;
  std::vector<double> epsinv_host;  // (3,3,Nx,Ny) column-major copy
  DevMem last_evecs;                // FP64, nev columns of ld complex
  int last_nev = 0;
  std::vector<double> last_kmags;   // k of the eigenvectors in last_evecs
  int last_adjoint_iters = 0;
  ModeSlot modes[MODE_SLOTS];
  bool trace = false;               // OMX_TRACE=1: per-Newton-step iteration log on stderr
  DevMem scratch_a, scratch_b, scratch_c;
  DevMem krylov_ws;                  // MINRES vectors and scratch (grown on demand)
  DevMem vec_part, vec_coef;          // partial sums and coefficients of the vector helpers
  RefineWorkspace refine_ws;          // buffers of the FP64 refinement phase
  DevMem field_scratch, argmax_v, argmax_i;  // canonicalize_phase
  // solve_k_multi: the problems' eps^-1 stacked per precision and layout, the instance ->
  // field map, and the band eigenvectors of all instances
  DevMem multi_e32r, multi_e32c, multi_e64r, multi_e64c, multi_slot, multi_evecs;
  Profiler prof;

  // LOBPCG workspaces kept between solves (a solve allocates nothing once its shape has
  // been seen): type-erased batches keyed by (precision bytes, instances, block, maxiter).
  struct Workspace {
    int tbytes, ninst, M, maxiter, tag;  // tag 1: a batch on stream2
    std::shared_ptr<void> batch;
    size_t bytes;  // device memory the batch holds
  };
  size_t workspace_bytes() const {
    size_t b = 0;
    for (const auto& w : workspaces) b += w.bytes;
    return b;
  }
  std::vector<Workspace> workspaces;

  template <typename T> PrecisionData<T>& pd();
  void canonicalize_phase(const double* d_k, c64* vecs, long ldv, long istride, int nvec, int mper);
};

// A symmetric per-pixel 3x3 field (column-major (3,3,Nx,Ny), FP64 host) uploaded in the two
// device layouts the operator reads (row-major and column-major pixel order, 6 entries each).
template <typename T> void upload_tensor6(const Solver& S, const double* E, DevMem& rowm, DevMem& colm);
// throws unless every pixel of E is finite and symmetric
void check_symmetric(long N, const double* E, const char* what);
// per pixel, out = sym(L M R) for column-major 3x3 fields
void pixel_product3(long N, const double* L, const double* M, const double* R, double* out);

// Allocation-free M(k) H (noalloc.hip): every device buffer in a fixed per-device array whose
// address noalloc_init looks up once per device (called when a handle is created there).
void noalloc_init(int device);
size_t noalloc_bytes(long N, int m, int prec);
void apply_helmholtz_noalloc(Solver& s, double k, const double* Hin, double* Hout, int m, int prec);

// Many solve_k problems of uneven grid, omega and mode count (batch.cpp).
void solve_k_batch(int device, int nprob, omx_problem* probs, int max_group, int fp64_only);

// Kottke-smoothed eps^-1 of convex polygons (smooth.hip); no handle needed.
void smooth_epsinv(int device, int nx, int ny, double Dx, double Dy, int nshape, const int* nverts,
                   const double* verts, const double* eps_shapes, const double* eps_background, int nsub,
                   double* epsinv_out);

// per-pixel 3x3 inverse and its pull-back Abar = -Y^T Ybar Y^T (tensor.hip); no handle
void sliceinv_3x3(int device, long npix, const double* A, double* Y_out);
void sliceinv_3x3_pullback(int device, long npix, const double* Y, const double* Ybar, double* Abar_out);

}  // namespace omx
