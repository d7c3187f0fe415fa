/*
 * omx -- MI355X (gfx950) plane-wave Maxwell eigensolver, C ABI.
 *
 * Drop-in backend for OptiMode.jl's MaxwellEigenmodes `solve_ω²` / `solve_k` and the
 * `solve_k` adjoint rrule.  Every entry point takes plain host pointers in the
 * reference's column-major layouts and returns an `int` status (OMX_OK = 0); the text of
 * the last error on a handle is available from omx_last_error().
 *
 * Layouts (0-based flat indices, complex values interleaved (re, im) as Julia's
 * ComplexF64):
 *   epsinv  (3,3,Nx,Ny) Float64      : a + 3*b + 9*(ix + Nx*iy)
 *   evec    (2,Nx,Ny)   ComplexF64   : c + 2*(ix + Nx*iy), i.e. 2N complex = 4N doubles
 *   blocks of m vectors              : vector j starts at j * 4N doubles
 *
 * Threading: calls on different handles may run concurrently (one HIP stream per
 * handle); calls on one handle are serialised by a per-handle mutex.
 */
#ifndef OMX_H
#define OMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMX_ABI_VERSION 1

/* status codes */
#define OMX_OK 0
#define OMX_ERR_NO_DEVICE 1      /* no usable gfx950 device */
#define OMX_ERR_BAD_SHAPE 2      /* grid / block sizes not supported or inconsistent */
#define OMX_ERR_NOT_CONVERGED 3  /* eigensolve or Newton iteration hit its iteration cap */
#define OMX_ERR_CHOLESKY 4       /* orthogonalisation failed even after shifting */
#define OMX_ERR_INVALID_ARG 5    /* null handle/pointer, non-finite or out-of-range value */
#define OMX_ERR_HIP 6            /* HIP runtime failure */
#define OMX_ERR_INTERNAL 7
#define OMX_ERR_UNSUPPORTED 8    /* valid input outside what this backend implements */

/* arithmetic of an operator application */
#define OMX_PREC_FP32 0
#define OMX_PREC_FP64 1

typedef struct omx_solver omx_solver; /* one grid + one eps^-1 field + one HIP stream */

/* per-solve statistics (LOBPCG counters follow DFTK's n_matvec / resid_history,
   lib/MaxwellEigenmodes/src/solvers/dftk.jl:288-295) */
typedef struct omx_info {
  int32_t converged;         /* 1 when every requested mode met k_tol / tol */
  int32_t n_newton;          /* Newton steps (eigensolves), summed over bands */
  int32_t lobpcg_iters_fp32; /* LOBPCG iterations in FP32, summed over bands and steps */
  int32_t lobpcg_iters_fp64; /* FP64 phase: Rayleigh-Ritz steps of the eigenpair refinement
                                (FP64 LOBPCG iterations with option refine = 0) */
  int64_t n_matvec;          /* operator applications (vectors) */
  int64_t n_precon;          /* preconditioner applications (vectors) */
  double max_residual;       /* final FP64 residual norm ||M x - w2 x||, max over modes */
  double device_ms;          /* GPU time of the solve (HIP events) */
} omx_info;

/* One problem of omx_solve_k_batch: inputs, then outputs written by the call. */
typedef struct omx_problem {
  int32_t nx, ny;            /* grid */
  double Dx, Dy;             /* cell (um) */
  const double* epsinv;      /* (3,3,Nx,Ny) column-major, symmetric per pixel */
  double omega;
  int32_t nev;
  int32_t max_eigsolves, maxiter;
  double k_tol, eig_tol;
  double* kmags_out;         /* nev */
  double* evecs_out;         /* NULL, or nev vectors of 2N complex */
  omx_info info;             /* out */
  int32_t status;            /* out: OMX_OK or the problem's error */
} omx_problem;

/* library */
const char* omx_version(void);
int omx_device_count(int* count);
const char* omx_status_string(int status);

/* Handle: grid (Dx x Dy um, Nx x Ny pixels; Grid, lib/DielectricSmoothing/src/grid.jl:46-71)
   and eps^-1 (symmetric 3x3 per pixel), uploaded once.  Replaces the HelmholtzMap /
   ModeSolver construction (lib/MaxwellEigenmodes/src/maxwell.jl:922-972). */
int omx_create(omx_solver** out, int device, int nx, int ny, double Dx, double Dy, const double* epsinv);
/* Handle on a 3D cell periodic along z (Grid{3}: Dx x Dy x Dz um, Nx x Ny x Nz pixels; Dz is
   the period Λ of lib/MaxwellEigenmodes/src/grads/period.jl:1-27).  eps^-1 is (3,3,Nx,Ny,Nz)
   column-major, eigenvectors have 2 Nx Ny Nz complex entries, flat index
   c + 2(ix + Nx(iy + Ny iz)).  The operator entry points, omx_solve_omega2, omx_solve_k (the
   solve_k of solve_k_periodic, period.jl:144-148), omx_hmh and omx_group_index accept
   such handles; the adjoint, ng_gvd and the EME mode sets return OMX_ERR_UNSUPPORTED. */
int omx_create_3d(omx_solver** out, int device, int nx, int ny, int nz, double Dx, double Dy, double Dz,
                  const double* epsinv);
int omx_destroy(omx_solver* h);
/* update_ε⁻¹ (maxwell.jl:1114-1136) */
int omx_set_epsinv(omx_solver* h, const double* epsinv);
const char* omx_last_error(const omx_solver* h);
/* tuning knobs: "fp32_tol" (solve_k's FP32 locking threshold / omega^2, 2e-3), "fp32_tol_omega2"
   (solve_omega2's, 1e-5), "fp32_switch", "jd_rtol", "jd_first_tol" (residual / omega^2 at which
   the FP64 refinement's first Newton step moves k, 3e-5), "maxiter_fp64", "guard", "seed"; start modes
   "warm_start" and "band_guess" (omx_solve_k); "guess_lowres" = hw > 0 (even, <= Nx, Ny):
   omx_solve_k's Hguess vectors hold only their hw x hw lowest plane waves, 2 hw^2 complex
   each at c + 2(gx + hw gy) in FFT order, and are scattered into the start block on the GPU */
int omx_set_option(omx_solver* h, const char* key, double value);

/* Autotuner (the reference's counterpart is FFTW's PATIENT planning when a ModeSolver is built,
   lib/MaxwellEigenmodes/src/maxwell.jl:940-943): timed trials on this handle's grid of the
   kernels of an nev-mode omx_solve_k, with the shapes its FP32 phase launches them with --
   the operator path (register FFT, with its rows per block of passes 1/3 and threads of pass
   2, or LDS Stockham), the Stockham FFT radix variant and tile shape (rows / columns per
   block), on 3D cells the yz-slab or five-pass form, and the Gram and block-update row
   partitions.  A candidate plan is used only if its output equals the default plan's to 1e-5.
   The fastest of each becomes the handle's option; the same choice can be applied to another
   handle of the grid with omx_set_option ("reg_ops", "fft_variant32", "tile_rows32",
   "tile_cols32", "gram_blocks", "combine_blocks", "slab3d", "reg_rows32", "reg_nt2").
   choice_out: NULL or 9 ints in that order; us_out: NULL or 3 doubles, microseconds per launch
   of the chosen operator application, Gram and block update. */
int omx_autotune(omx_solver* h, int nev, int* choice_out, double* us_out);

/* HelmholtzMap * H for m vectors at k_vec = (0,0,k): _M! (maxwell.jl:471-481) */
int omx_apply_helmholtz(omx_solver* h, double k, const double* H_in, double* H_out, int m, int precision);
/* omx_apply_helmholtz that allocates no device memory.  No counterpart in the reference: it is
   for callers that cannot allocate (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Same arguments and results as omx_apply_helmholtz; every
   device buffer of the call (k, the FP64 staging of H_in / H_out, the FP32 copies and the
   operator's 3-component intermediate) lives in a fixed 64 MiB array per device that the
   library holds until the process ends.  A call needs 256 + 112 N m bytes (FP64) or
   256 + 88 N m bytes (FP32), N = Nx Ny (Nx Ny Nz on 3D handles): at 256x256, m <= 9 in FP64 and
   m <= 11 in FP32; larger calls return OMX_ERR_BAD_SHAPE.  Calls on one device share the array
   (it keeps the last call's contents) and are serialised by a per-device lock, since
   overlapping calls would race on it. */
int omx_apply_helmholtz_noalloc(omx_solver* h, double k, const double* H_in, double* H_out, int m, int precision);
/* HelmholtzPreconditioner * H: _P! (maxwell.jl:459-469) */
int omx_apply_preconditioner(omx_solver* h, double k, const double* H_in, double* H_out, int m,
                             int precision);
/* HMₖH (maxwell.jl:267-275) per vector: out[j] = -Re<H_j, kx_ct(ifft(eps^-1 fft(zx_tc H_j)))> */
int omx_hmkh(omx_solver* h, double k, const double* H, int m, double* out, int precision);

/* Period sensitivity on a 3D handle, FP64, per mode j at k = kmags[j]:
   dw2dk_out[j] = 2 HMₖH(ev_j) (∂ω²/∂k) and dw2dL_out[j] = ∂ω²_∂Λ(ev_j) = 2 HMₖH_weighted(ev_j,
   g_z/Λ) (lib/MaxwellEigenmodes/src/grads/period.jl:69-74, :88-90).  With them the pull-back
   of solve_k_periodic for cotangents on kmags is ω̄ = Σ 2ω k̄_j/∂ω²∂k_j and
   Λ̄ = Σ k̄_j (-∂ω²∂Λ_j/∂ω²∂k_j) (period.jl:224-232). */
int omx_period_derivatives(omx_solver* h, int nev, const double* kmags, const double* evecs, double* dw2dk_out,
                           double* dw2dL_out);

/* HMH (src/maxwell.jl:288-291) per vector, FP64: out[j] = Re<H_j, kx_ct(ifft(A fft(kx_tc H_j)))>
   with normalised ifft, for a symmetric per-pixel field A (3,3,Nx,Ny) column-major, or
   A = NULL for the handle's eps^-1 (then out[j] = <H_j, M(k) H_j>, = ω² for a unit eigenvector). */
int omx_hmh(omx_solver* h, double k, const double* A, const double* H, int m, double* out);

/* group_index(k, ev, ω, ε⁻¹, ∂ε_∂ω, grid) (lib/ModeAnalysis/src/group_index.jl:21-29), FP64:
   ng_out[j] = (ω + HMH(ev_j, ε⁻¹ ∂ε_∂ω ε⁻¹)/2) / HMₖH(ev_j) at k = kmags[j], for nev modes
   (evecs: nev vectors of 2N complex, as omx_solve_k returns them).  deps_domega: (3,3,Nx,Ny)
   column-major, symmetric per pixel, or NULL for a non-dispersive medium. */
int omx_group_index(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
                    const double* deps_domega, double* ng_out);

/* ng_gvd(ω, k, ev, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid) and ng_gvd_E (lib/ModeAnalysis/src/
   group_index.jl:126-162, :55-95), FP64, for nev modes at once: the group index ng_out[j]
   and the group-velocity dispersion gvd_out[j] = d n_g / d ω (one linear solve with
   M - ω² per mode).  deps_domega, d2eps_domega2: (3,3,Nx,Ny) column-major, symmetric per
   pixel (not NULL).  E_out: NULL, or nev fields E = ε⁻¹ fft(kx_tc(ev_j)) of 3N complex each
   ((3,Nx,Ny) column-major, unnormalised), as ng_gvd_E returns them. */
int omx_ng_gvd(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
               const double* deps_domega, const double* d2eps_domega2, double* ng_out, double* gvd_out,
               double* E_out);

/* Batched solve_k (no single reference counterpart: the 64-instance mixed-grid batch of
   independent solve_k calls, src/solve.jl:182-196, that sweeps and EME run one by one,
   lib/ModeSweeps/src/batch.jl:256-265).  Problems may differ in grid, eps^-1, ω, nev and
   tolerances.  They are bucketed by grid; a bucket runs on one handle and one HIP stream in
   launch groups of as many problems as fit half the free device memory (at most max_group
   problems when max_group > 0): every band of every problem of a group is an instance of the
   same LOBPCG, Newton, refinement and phase launches.  Per-problem results and status go into
   probs (info.lobpcg_iters_fp32 and converged are the problem's own; the FP64 step count,
   matvecs and device_ms are its launch group's); the call returns OMX_OK when every problem
   succeeded, else the first failing status.  fp64_only != 0 runs every solve in FP64. */
int omx_solve_k_batch(int device, int nprob, omx_problem* probs, int max_group, int fp64_only);

/* EME per-slice mode set (lib/EigenmodeExpansion/src/modes.jl:105-122, mode_fields and
   build_mode): E = i ε⁻¹ fft(kx_tc(ev_j)) and H = fft(tc(ev_j)) of nev <= 16 modes at
   k = kmags[j], kept on the device in mode set `slot` (0..3) of the handle.  evecs: nev host
   vectors (2N complex each), or NULL for the eigenvectors of the handle's last omx_solve_k.
   normalize != 0 divides both fields of mode j by nrm_j = sqrt(inner_product(m_j, m_j;
   conjugate)) (principal square root); norms_out: NULL or nev complex nrm_j. */
int omx_build_modes(omx_solver* h, int slot, int nev, const double* kmags, const double* evecs, int conjugate,
                    int normalize, double* norms_out);

/* overlap_matrix(modes_a, modes_b; conjugate) (modes.jl:135-170, MEOW inner product):
   out[i + na j] = 1/2 Σ (E_a,i,x H_b,j,y − E_a,i,y H_b,j,x) δA over the pixels, E_a
   conjugated when conjugate != 0; na x nb complex, column-major. */
int omx_overlap_matrix(omx_solver* h, int slot_a, int slot_b, int conjugate, double* out);

/* Pull-back of overlap_matrix (modes.jl:156-170, the reverse-mode rule the reference takes
   from Zygote for it): given the cotangent obar (na x nb complex, column-major) of
   O = overlap_matrix(set slot_a, set slot_b; conjugate), add the field cotangents to the sets'
   device accumulators -- of set a's E when which & 1, of set b's H when which & 2.  Cotangent
   convention: a real scalar L of a complex array Z changes by Re sum conj(Zbar) dZ.  The
   accumulators start at zero after omx_build_modes and after omx_mode_pullback. */
int omx_overlap_pullback(omx_solver* h, int slot_a, int slot_b, int conjugate, const double* obar, int which);

/* Pull-back of build_mode (modes.jl:116-122: mode_fields :105-113 and the power
   normalisation) for the nev modes of a set, from its accumulated field cotangents (which
   are then cleared): kbar_out[j] (through the plane-wave basis' dependence on k_j),
   evbar_out (nev vectors of 2N complex: the eigenvector cotangents omx_solve_k_adjoint
   takes) and epsinvbar_out ((3,3,Nx,Ny) column-major: the direct dependence of
   E = i eps^-1 fft(kx_tc(ev)) on eps^-1, summed over the modes, each entry its own
   derivative as the contraction eps^-1_dot's rule gives it).  The handle's eps^-1 must be
   the one the set was built with. */
int omx_mode_pullback(omx_solver* h, int slot, double* kbar_out, double* evbar_out, double* epsinvbar_out);

/* The fields of mode set `slot`: E_out, H_out (either may be NULL) receive nev fields of
   3N complex each, (3,Nx,Ny) column-major, as Mode.E / Mode.H hold them. */
int omx_get_mode_fields(omx_solver* h, int slot, double* E_out, double* H_out);

/* solve_ω²(ms, solver; nev, tol, maxiter) (src/solve.jl:38-63, solvers/dftk.jl:32-37):
   lowest nev eigenpairs of M(k).  H_io: in = start block (2N x nev complex) or NULL for a
   seeded random start; out = eigenvectors (written when non-NULL). */
int omx_solve_omega2(omx_solver* h, double k, int nev, double tol, int maxiter, double* H_io,
                     double* omega2_out, omx_info* info);

/* solve_k(ω, ε⁻¹, grid, solver; nev, k_tol, eig_tol, max_eigsolves, maxiter, kguess, Hguess)
   (src/solve.jl:182-196, :104-115; solvers/gpu.jl:213-242): per band, Newton on
   ω²(k) - ω² = 0 with slope 2 HMₖH; eigenvectors unit-norm and phase-canonicalised
   (fields.jl:258-260).  kguess may be NULL (k_guess, MaxwellEigenmodes.jl:63); Hguess may
   be NULL or hold nguess start vectors.  evecs_out may be NULL: the eigenvectors then stay
   on the device (fetch them with omx_get_evecs).  With option "warm_start" set and both
   kguess and Hguess NULL, every band starts at the previous solve's k of that band with
   the previous eigenvectors (on the device) as its block -- the EME slice chain
   (lib/EigenmodeExpansion/src/eme.jl:58-76).  With option "band_guess" set, kguess holds nev
   values and Hguess at least one vector: band j starts at kguess[j] with the Hguess block
   and the shared first eigensolve is skipped (the learned warm start, omx_warmnet_*). */
int omx_solve_k(omx_solver* h, double omega, int nev, double k_tol, double eig_tol, int max_eigsolves,
                int maxiter, const double* kguess, const double* Hguess, int nguess, double* kmags_out,
                double* evecs_out, omx_info* info);
/* eigenvectors of the last omx_solve_k on this handle (2N x nev complex) */
int omx_get_evecs(omx_solver* h, double* evecs_out);

/* Pull-back of rrule(solve_k) (src/solve.jl:213-289, solvers/gpu.jl:292-365): given the
   primal (kmags, evecs) and cotangents kbar (nev, NULL = zero) and evbar (2N x nev complex,
   NULL = zero; evbar_mask[j] = 0 marks band j's cotangent as zero), returns
   omegabar and epsinvbar: (3,3,Nx,Ny) with the 2D mirrored off-diagonal convention
   (grads/solve.jl:490-515), or on a 3D handle (3,3,Nx,Ny,Nz) with each off-diagonal sum
   above the diagonal and zeros below (grads/solve.jl:464-487). */
int omx_solve_k_adjoint(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
                        const double* kbar, const double* evbar, const uint8_t* evbar_mask,
                        double* omegabar_out, double* epsinvbar_out);

/* Pull-back of rrule(solve_k_periodic, ω, ε⁻¹, Λ, grid, solver) (grads/period.jl:163-249,
   _solve_k_period_grads) on a 3D handle whose cell period Dz is Λ: the arguments and the
   ω̄, ε̄⁻¹ outputs of omx_solve_k_adjoint on that handle, plus the period cotangent
   Lbar_out = Σ_j Λ̄_H,j + k̄tot_j (-∂ω²∂Λ_j / ∂ω²∂k_j), where k̄tot_j = k̄_j + k̄ₕ,j, and
   k̄ₕ and Λ̄_H = -Σ_G w̄_z g_z/Λ come from the eigenvector channel (eig_adjt) through the
   plane-wave basis (period.jl:97-119).  OMX_ERR_UNSUPPORTED on a 2D handle. */
int omx_solve_k_periodic_adjoint(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
                                 const double* kbar, const double* evbar, const uint8_t* evbar_mask,
                                 double* omegabar_out, double* epsinvbar_out, double* Lbar_out);

/* Kottke sub-pixel smoothing on the GPU (lib/DielectricSmoothing/src/kottke.jl:44-95,
   smooth.jl:115-170, smooth_eps_single and smooth_eps): eps^-1 (3,3,Nx,Ny) column-major
   of nshapes convex polygons over a background, on the grid of cell Dx x Dy with pixel centres at -D/2 + i D/N.  Shape s has
   nverts[s] counter-clockwise vertices, (x, y) pairs concatenated in verts, and the
   symmetric tensor eps_shapes[9 s .. 9 s + 8] (column-major 3x3); earlier shapes win where
   shapes overlap.  Pixels are supersampled nsub x nsub to find the fill fraction and the
   interface normal (the reference uses GeometryPrimitives' surfpt_nearby / volfrac).  No
   handle: errors are reported through omx_last_error(NULL).  Limits: nshapes <= 15,
   1 <= nsub <= 64. */
int omx_smooth_epsinv(int device, int nx, int ny, double Dx, double Dy, int nshapes, const int* nverts,
                      const double* verts, const double* eps_shapes, const double* eps_background, int nsub,
                      double* epsinv_out);

/* sliceinv_3x3 (lib/MaxwellEigenmodes/src/linalg.jl:126-149): the per-pixel inverse Y = A^-1
   of npix 3x3 tensors, column-major (3,3,Nx,Ny[,Nz]) (pixel p at 9p + a + 3b), FP64 on the
   GPU; and its pull-back (rrule, src/grads/linalg.jl:40-47) Abar = -Y^T Ybar Y^T, which
   turns an eps^-1 cotangent (omx_solve_k_adjoint) into an eps cotangent.  No handle:
   errors are reported through omx_last_error(NULL). */
int omx_sliceinv_3x3(int device, int64_t npix, const double* A, double* Y_out);
int omx_sliceinv_3x3_pullback(int device, int64_t npix, const double* Y, const double* Ybar, double* Abar_out);

/* Learned warm start (no counterpart in the reference, which starts every solve from random
   vectors and k_guess, src/MaxwellEigenmodes.jl:63): a small convolutional network on the GPU,
   FP32, mapping a hw x hw map of 7 channels (the 6 unique eps^-1 components and omega) to
   nev effective-index guesses and 4 nev channels of low-resolution start vectors (re/im of the
   two plane-wave components per band).  Tensors are [batch][channel][y][x] row-major floats.
   Layers (3x3 convolutions): 7->16 @hw, pool, 16->32 @hw/2, pool, 32->32 @hw/4 (-> pixel mean
   -> dense -> nev), [up(.) | 32] 64->32 @hw/2, [up(.) | 16] 48->32 @hw, 32->4nev @hw.
   Loss: mean (y - t_fields)^2 + wk mean (k - t_k)^2.  hw: multiple of 4 in 4..256; nev 1..16.
   Errors are reported through omx_last_error(NULL); calls on one network are serialised. */
typedef struct omx_warmnet omx_warmnet;
int omx_warmnet_create(int device, int hw, int nev, uint64_t seed, omx_warmnet** out);
int omx_warmnet_destroy(omx_warmnet* h);
int omx_warmnet_num_params(omx_warmnet* h, int64_t* n);
int omx_warmnet_get_params(omx_warmnet* h, float* out);
int omx_warmnet_set_params(omx_warmnet* h, const float* in);
/* fields_out (batch x 4nev x hw x hw) and k_out (batch x nev) may each be NULL */
int omx_warmnet_forward(omx_warmnet* h, int batch, const float* x, float* fields_out, float* k_out);
/* loss of one batch and its gradient with respect to the flat parameters (grad_out may be NULL) */
int omx_warmnet_loss_grad(omx_warmnet* h, int batch, const float* x, const float* t_fields, const float* t_k,
                          float wk, double* loss_out, float* grad_out);
/* Adam (β 0.9, 0.999) over n samples held on the device: `epochs` seeded shuffles, minibatches
   of `batch` (the remainder of each shuffle is skipped); loss_hist_out[e] (may be NULL) is the
   mean minibatch loss of epoch e. */
int omx_warmnet_train(omx_warmnet* h, int n, const float* x, const float* t_fields, const float* t_k, float wk,
                      int epochs, int batch, float lr, uint64_t seed, double* loss_hist_out);
/* Convolutions of the network (forward and data gradient) in bf16 on the matrix cores with
   FP32 accumulation (bf16 != 0) or FP32 on the vector units (0, the default); parameters,
   weight gradient and Adam stay FP32. */
int omx_warmnet_set_precision(omx_warmnet* h, int bf16);
/* omx_warmnet_train continuing the previous call's Adam moments and step count (a training
   run in pieces whose targets change between them, e.g. re-aligned to the current output). */
int omx_warmnet_train_resume(omx_warmnet* h, int n, const float* x, const float* t_fields, const float* t_k, float wk,
                             int epochs, int batch, float lr, uint64_t seed, double* loss_hist_out);

/* The guess omx_solve_k_learned starts from, for inspection: k_out (nev floats, the
   normalised effective indices) and, unless NULL, lowres_out (nev vectors of 2 hw^2 complex,
   entry c + 2(gx + hw gy), unit norm). */
int omx_warmnet_guess(omx_warmnet* net, omx_solver* h, double omega, float* k_out, double* lowres_out);

/* omx_solve_k from the network's guess with everything on the device but the nev predicted
   indices: features from the handle's eps^-1, one network forward, the inverse 2D DFT of the
   predicted maps into the start block's hw x hw lowest plane waves; the shared first
   eigensolve starts at k0_scale times the predicted band-1 k (1.0 is the tested choice).
   Arguments and outputs otherwise as omx_solve_k; 2D handles whose Nx, Ny are multiples of
   the network's hw, nev <= the network's band count, both on one device. */
int omx_solve_k_learned(omx_solver* h, omx_warmnet* net, double omega, int nev, double k0_scale, double k_tol,
                        double eig_tol, int max_eigsolves, int maxiter, double* kmags_out, double* evecs_out,
                        omx_info* info);

/* Live kernel statistics (no counterpart in the reference: instrumentation for
   measurement).  With omx_set_option(h, "profile", 1) every operator application inside
   a solve is bracketed by HIP events on the handle's stream; `family` is "op_fp32" or
   "op_fp64".  Returns total milliseconds, launches and algorithmic bytes
   (32 N m + 24 N per FP32 application of m vectors, twice that in FP64). */
int omx_get_kernel_stats(omx_solver* h, const char* family, double* total_ms, int64_t* launches, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* OMX_H */
