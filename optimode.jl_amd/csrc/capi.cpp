// extern "C" entry points of libomx (declared in include/omx.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "solver.hpp"
#include "warmnet.hpp"

struct omx_warmnet {
  std::unique_ptr<omx::WarmNet> n;
  std::mutex mu;
};

struct omx_solver {
  std::unique_ptr<omx::Solver> s;
  std::string create_error;
};

namespace {

thread_local std::string g_error;  // errors raised before a handle exists

template <typename F> int guarded(omx_solver* h, F&& f) {
  if (!h || !h->s) return OMX_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(h->s->mu);
  try {
    h->s->last_error.clear();
    f(*h->s);
    return OMX_OK;
  } catch (const omx::Error& e) {
    h->s->last_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    h->s->last_error = e.what();
    return OMX_ERR_INTERNAL;
  } catch (...) {
    h->s->last_error = "unknown error";
    return OMX_ERR_INTERNAL;
  }
}

// handle-free errors (and those of warm-start networks) go to the thread's last error
template <typename F> int guarded_free(F&& f) {
  try {
    f();
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  } catch (...) {
    g_error = "unknown error";
    return OMX_ERR_INTERNAL;
  }
}

template <typename F> int guarded_net(omx_warmnet* h, F&& f) {
  if (!h || !h->n) return OMX_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  return guarded_free([&] { f(*h->n); });
}

}  // namespace

extern "C" {

const char* omx_version(void) { return "omx 0.1 (gfx950)"; }

int omx_device_count(int* count) {
  if (!count) return OMX_ERR_INVALID_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return OMX_OK;
}

const char* omx_status_string(int status) {
  switch (status) {
    case OMX_OK: return "ok";
    case OMX_ERR_NO_DEVICE: return "no usable gfx950 device";
    case OMX_ERR_BAD_SHAPE: return "unsupported or inconsistent shape";
    case OMX_ERR_NOT_CONVERGED: return "iteration did not converge";
    case OMX_ERR_CHOLESKY: return "orthogonalisation failed after shifting";
    case OMX_ERR_INVALID_ARG: return "invalid argument";
    case OMX_ERR_HIP: return "HIP runtime error";
    case OMX_ERR_INTERNAL: return "internal error";
    case OMX_ERR_UNSUPPORTED: return "unsupported input";
    default: return "unknown status";
  }
}

int omx_create(omx_solver** out, int device, int nx, int ny, double Dx, double Dy, const double* epsinv) {
  if (!out) return OMX_ERR_INVALID_ARG;
  *out = nullptr;
  try {
    auto h = std::make_unique<omx_solver>();
    h->s = std::make_unique<omx::Solver>(device, nx, ny, Dx, Dy, epsinv);
    *out = h.release();
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  }
}

int omx_create_3d(omx_solver** out, int device, int nx, int ny, int nz, double Dx, double Dy, double Dz,
                  const double* epsinv) {
  if (!out) return OMX_ERR_INVALID_ARG;
  *out = nullptr;
  try {
    auto h = std::make_unique<omx_solver>();
    h->s = std::make_unique<omx::Solver>(device, nx, ny, nz, Dx, Dy, Dz, epsinv);
    *out = h.release();
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  }
}

int omx_period_derivatives(omx_solver* h, int nev, const double* kmags, const double* evecs, double* dw2dk_out,
                           double* dw2dL_out) {
  return guarded(h, [&](omx::Solver& s) { s.period_derivatives(nev, kmags, evecs, dw2dk_out, dw2dL_out); });
}

int omx_destroy(omx_solver* h) {
  if (!h) return OMX_ERR_INVALID_ARG;
  delete h;
  return OMX_OK;
}

int omx_smooth_epsinv(int device, int nx, int ny, double Dx, double Dy, int nshapes, const int* nverts,
                      const double* verts, const double* eps_shapes, const double* eps_background, int nsub,
                      double* epsinv_out) {
  try {
    omx::smooth_epsinv(device, nx, ny, Dx, Dy, nshapes, nverts, verts, eps_shapes, eps_background, nsub, epsinv_out);
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  }
}

int omx_sliceinv_3x3(int device, int64_t npix, const double* A, double* Y_out) {
  try {
    omx::sliceinv_3x3(device, long(npix), A, Y_out);
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  }
}

int omx_sliceinv_3x3_pullback(int device, int64_t npix, const double* Y, const double* Ybar, double* Abar_out) {
  try {
    omx::sliceinv_3x3_pullback(device, long(npix), Y, Ybar, Abar_out);
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  }
}

int omx_solve_k_batch(int device, int nprob, omx_problem* probs, int max_group, int fp64_only) {
  try {
    omx::solve_k_batch(device, nprob, probs, max_group, fp64_only);
    return OMX_OK;
  } catch (const omx::Error& e) {
    g_error = e.what();
    return e.code();
  } catch (const std::exception& e) {
    g_error = e.what();
    return OMX_ERR_INTERNAL;
  }
}

const char* omx_last_error(const omx_solver* h) {
  if (!h || !h->s) return g_error.c_str();
  return h->s->last_error.c_str();
}

int omx_set_epsinv(omx_solver* h, const double* epsinv) {
  return guarded(h, [&](omx::Solver& s) { s.set_epsinv(epsinv); });
}

int omx_set_option(omx_solver* h, const char* key, double value) {
  if (!key) return OMX_ERR_INVALID_ARG;
  return guarded(h, [&](omx::Solver& s) { s.set_option(key, value); });
}

int omx_apply_helmholtz(omx_solver* h, double k, const double* H_in, double* H_out, int m, int precision) {
  return guarded(h, [&](omx::Solver& s) { s.apply(omx::OP_M, precision, k, H_in, H_out, m, nullptr); });
}

int omx_apply_helmholtz_noalloc(omx_solver* h, double k, const double* H_in, double* H_out, int m, int precision) {
  return guarded(h, [&](omx::Solver& s) { omx::apply_helmholtz_noalloc(s, k, H_in, H_out, m, precision); });
}

int omx_autotune(omx_solver* h, int nev, int* choice_out, double* us_out) {
  return guarded(h, [&](omx::Solver& s) { s.autotune(nev, choice_out, us_out); });
}

int omx_apply_preconditioner(omx_solver* h, double k, const double* H_in, double* H_out, int m, int precision) {
  return guarded(h, [&](omx::Solver& s) { s.apply(omx::OP_P, precision, k, H_in, H_out, m, nullptr); });
}

int omx_hmkh(omx_solver* h, double k, const double* H, int m, double* out, int precision) {
  return guarded(h, [&](omx::Solver& s) { s.apply(omx::OP_MKH, precision, k, H, nullptr, m, out); });
}

int omx_hmh(omx_solver* h, double k, const double* A, const double* H, int m, double* out) {
  return guarded(h, [&](omx::Solver& s) { s.hmh(k, A, H, m, out); });
}

int omx_group_index(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
                    const double* deps_domega, double* ng_out) {
  return guarded(h, [&](omx::Solver& s) { s.group_index(omega, nev, kmags, evecs, deps_domega, ng_out); });
}

int omx_ng_gvd(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
               const double* deps_domega, const double* d2eps_domega2, double* ng_out, double* gvd_out,
               double* E_out) {
  return guarded(h, [&](omx::Solver& s) {
    s.ng_gvd(omega, nev, kmags, evecs, deps_domega, d2eps_domega2, ng_out, gvd_out, E_out);
  });
}

int omx_build_modes(omx_solver* h, int slot, int nev, const double* kmags, const double* evecs, int conjugate,
                    int normalize, double* norms_out) {
  return guarded(h, [&](omx::Solver& s) { s.build_modes(slot, nev, kmags, evecs, conjugate, normalize, norms_out); });
}

int omx_overlap_matrix(omx_solver* h, int slot_a, int slot_b, int conjugate, double* out) {
  return guarded(h, [&](omx::Solver& s) { s.overlap(slot_a, slot_b, conjugate, out); });
}

int omx_overlap_pullback(omx_solver* h, int slot_a, int slot_b, int conjugate, const double* obar, int which) {
  return guarded(h, [&](omx::Solver& s) { s.overlap_pullback(slot_a, slot_b, conjugate, obar, which); });
}

int omx_mode_pullback(omx_solver* h, int slot, double* kbar_out, double* evbar_out, double* epsinvbar_out) {
  return guarded(h, [&](omx::Solver& s) { s.mode_pullback(slot, kbar_out, evbar_out, epsinvbar_out); });
}

int omx_get_mode_fields(omx_solver* h, int slot, double* E_out, double* H_out) {
  return guarded(h, [&](omx::Solver& s) { s.get_mode_fields(slot, E_out, H_out); });
}

int omx_solve_omega2(omx_solver* h, double k, int nev, double tol, int maxiter, double* H_io, double* omega2_out,
                     omx_info* info) {
  return guarded(h, [&](omx::Solver& s) { s.solve_omega2(k, nev, tol, maxiter, H_io, omega2_out, info); });
}

int omx_solve_k(omx_solver* h, double omega, int nev, double k_tol, double eig_tol, int max_eigsolves, int maxiter,
                const double* kguess, const double* Hguess, int nguess, double* kmags_out, double* evecs_out,
                omx_info* info) {
  return guarded(h, [&](omx::Solver& s) {
    s.solve_k(omega, nev, k_tol, eig_tol, max_eigsolves, maxiter, kguess, Hguess, nguess, kmags_out, evecs_out, info);
  });
}

int omx_get_evecs(omx_solver* h, double* evecs_out) {
  return guarded(h, [&](omx::Solver& s) { s.get_evecs(evecs_out); });
}

int omx_solve_k_adjoint(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
                        const double* kbar, const double* evbar, const uint8_t* evbar_mask, double* omegabar_out,
                        double* epsinvbar_out) {
  return guarded(h, [&](omx::Solver& s) {
    s.adjoint(omega, nev, kmags, evecs, kbar, evbar, evbar_mask, omegabar_out, epsinvbar_out);
  });
}

int omx_solve_k_periodic_adjoint(omx_solver* h, double omega, int nev, const double* kmags, const double* evecs,
                                 const double* kbar, const double* evbar, const uint8_t* evbar_mask,
                                 double* omegabar_out, double* epsinvbar_out, double* Lbar_out) {
  return guarded(h, [&](omx::Solver& s) {
    if (!Lbar_out) throw omx::Error(OMX_ERR_INVALID_ARG, "NULL Lbar_out");
    s.adjoint(omega, nev, kmags, evecs, kbar, evbar, evbar_mask, omegabar_out, epsinvbar_out, Lbar_out);
  });
}

int omx_warmnet_create(int device, int hw, int nev, uint64_t seed, omx_warmnet** out) {
  if (!out) return OMX_ERR_INVALID_ARG;
  *out = nullptr;
  return guarded_free([&] {
    auto h = std::make_unique<omx_warmnet>();
    h->n = std::make_unique<omx::WarmNet>(device, hw, nev, seed);
    *out = h.release();
  });
}

int omx_warmnet_destroy(omx_warmnet* h) {
  delete h;
  return OMX_OK;
}

int omx_warmnet_num_params(omx_warmnet* h, int64_t* n) {
  if (!n) return OMX_ERR_INVALID_ARG;
  return guarded_net(h, [&](omx::WarmNet& w) { *n = w.nparams(); });
}

int omx_warmnet_get_params(omx_warmnet* h, float* out) {
  if (!out) return OMX_ERR_INVALID_ARG;
  return guarded_net(h, [&](omx::WarmNet& w) { w.get_params(out); });
}

int omx_warmnet_set_params(omx_warmnet* h, const float* in) {
  if (!in) return OMX_ERR_INVALID_ARG;
  return guarded_net(h, [&](omx::WarmNet& w) { w.set_params(in); });
}

int omx_warmnet_forward(omx_warmnet* h, int batch, const float* x, float* fields_out, float* k_out) {
  return guarded_net(h, [&](omx::WarmNet& w) { w.forward(batch, x, fields_out, k_out); });
}

int omx_warmnet_loss_grad(omx_warmnet* h, int batch, const float* x, const float* t_fields, const float* t_k,
                          float wk, double* loss_out, float* grad_out) {
  if (!loss_out) return OMX_ERR_INVALID_ARG;
  return guarded_net(h, [&](omx::WarmNet& w) { *loss_out = w.loss_grad(batch, x, t_fields, t_k, wk, grad_out); });
}

int omx_warmnet_train(omx_warmnet* h, int n, const float* x, const float* t_fields, const float* t_k, float wk,
                      int epochs, int batch, float lr, uint64_t seed, double* loss_hist_out) {
  return guarded_net(h, [&](omx::WarmNet& w) { w.train(n, x, t_fields, t_k, wk, epochs, batch, lr, seed, loss_hist_out); });
}

int omx_warmnet_set_precision(omx_warmnet* h, int bf16) {
  return guarded_net(h, [&](omx::WarmNet& w) { w.set_bf16(bf16 != 0); });
}

int omx_warmnet_train_resume(omx_warmnet* h, int n, const float* x, const float* t_fields, const float* t_k, float wk,
                             int epochs, int batch, float lr, uint64_t seed, double* loss_hist_out) {
  return guarded_net(h, [&](omx::WarmNet& w) {
    w.train(n, x, t_fields, t_k, wk, epochs, batch, lr, seed, loss_hist_out, true);
  });
}

int omx_warmnet_guess(omx_warmnet* net, omx_solver* h, double omega, float* k_out, double* lowres_out) {
  if (!net || !net->n || !k_out) return OMX_ERR_INVALID_ARG;
  return guarded(h, [&](omx::Solver& s) {
    std::lock_guard<std::mutex> lk(net->mu);
    s.require_2d("the learned start");
    if (net->n->device() != s.device) throw omx::Error(OMX_ERR_INVALID_ARG, "the network and the solver are on different devices");
    OMX_HIP_CHECK(hipSetDevice(s.device));
    const omx::c64* low = net->n->guess_device(s.p64.plan.eps6, s.nx, s.ny, omega, k_out);
    if (lowres_out) {
      const size_t n = size_t(net->n->nev()) * 2 * net->n->hw() * net->n->hw();
      OMX_HIP_CHECK(hipMemcpy(lowres_out, low, sizeof(omx::c64) * n, hipMemcpyDeviceToHost));
    }
  });
}

int omx_solve_k_learned(omx_solver* h, omx_warmnet* net, double omega, int nev, double k0_scale, double k_tol,
                        double eig_tol, int max_eigsolves, int maxiter, double* kmags_out, double* evecs_out,
                        omx_info* info) {
  if (!net || !net->n) return OMX_ERR_INVALID_ARG;
  return guarded(h, [&](omx::Solver& s) {
    std::lock_guard<std::mutex> lk(net->mu);
    omx::solve_k_learned(s, *net->n, omega, nev, k0_scale, k_tol, eig_tol, max_eigsolves, maxiter, kmags_out,
                         evecs_out, info);
  });
}

int omx_get_kernel_stats(omx_solver* h, const char* family, double* total_ms, int64_t* launches, double* bytes) {
  if (!family || !total_ms || !launches || !bytes) return OMX_ERR_INVALID_ARG;
  return guarded(h, [&](omx::Solver& s) {
    s.prof.flush();
    const omx::KernelStat* st = nullptr;
    if (std::strcmp(family, "op_fp32") == 0)
      st = &s.prof.op32;
    else if (std::strcmp(family, "op_fp64") == 0)
      st = &s.prof.op64;
    else
      throw omx::Error(OMX_ERR_INVALID_ARG, std::string("unknown kernel family '") + family + "'");
    *total_ms = st->ms;
    *launches = st->launches;
    *bytes = st->bytes;
  });
}

}  // extern "C"
