// Batched vector helpers and preconditioned MINRES on the complement of a vector, shared by
// the adjoint (eig_adjt), GVD (FP64) and the eigenpair refinement's correction solves (FP32).
#pragma once
#include <initializer_list>
#include <vector>

#include "solver.hpp"

namespace omx {

// sum of nb partial values
double sum_part(const double* p, int nb);

// Vector helpers over nb instances of K complex entries (FP64 or FP32; dots and
// coefficients in FP64) at stride ld, against the unit vectors EV (vector i at EV + i ld).
// Coefficients of lincomb live on the device (coef: 4 complex per instance), so that no
// step waits for the host.
template <typename T> struct VecOpsT {
  using C = cplx<T>;
  hipStream_t s;
  int nb;
  long ld, K;
  int nblk;
  const C* EV;
  DevMem& part;  // the handle's buffers, grown on demand (no allocation per solve)
  DevMem& coef;
  VecOpsT(Solver& S, hipStream_t s_, int nb_, long ld_, long K_, const C* EV_);
  // part[i nblk + b] = partial sums of Re <A_i, B_i>
  void vdot(const C* A, const C* B, const int* act);
  // out_i = sum_j coef[slot][4 i + j] in_j_i (out may alias an input); coefficient slot
  // s (of four) at coef + 4 nb s
  void lincomb(std::initializer_list<const C*> ins, C* outp, const int* act, int slot = 0);
  // the same, also leaving the partials of Re <d_i, out_i> in part (as vdot)
  void lincomb_dot(std::initializer_list<const C*> ins, C* outp, const int* act, const C* d, int slot = 0);
  // x_i <- x_i - ev_i <ev_i, x_i> (coefficient slot 3); with dot_with, also the partials of
  // Re <dot_with_i, x_i> of the projected x in part
  void project(C* x, const int* act, const C* dot_with = nullptr);
  // the same, the partials of Re <dot_with_i, x_i> into out (layout of part)
  void project_dot_to(C* x, const int* act, const C* dot_with, double* out);
  // Block deflation: from here on the projections remove the components of x_i along the
  // nq[i] orthonormal columns of instance i's block Qb (column c at Qb + i qis + c ld, at
  // most QMAX columns; nq on the device) instead of along ev_i.
  static constexpr int QMAX = 16;
  void set_block(const C* Qb, long qis, const int* nq, int nq_max);
  const C* Q = nullptr;
  long qistride = 0;
  const int* qn = nullptr;
  int qmax = 0;   // the largest nq[i] (host): columns the block kernels are compiled for
  int qblk = 0;   // row blocks per instance of the block dots
};
using VecOps = VecOpsT<double>;

// Solves (M(k_i) - shift_i) x_i = r_i on the complement of ev_i for the instances with
// want[i] (MINRES, Paige & Saunders 1975, preconditioned by P projected onto the
// complement; M - shift is Hermitian, singular along ev when shift is its eigenvalue, and
// indefinite above the lowest band), with vectors and operator applications in T and the
// recurrences in FP64.  r is projected on entry and overwritten; xout receives the
// solution orthogonalised against ev twice.  Stops when the preconditioned residual
// estimate falls below rtol times its start, or after maxit iterations (an error only when
// throw_on_cap).  Returns the iteration count.
template <typename T>
int minres_complement_t(Solver& S, VecOpsT<T>& V, const std::vector<double>& shifts, const double* dk, cplx<T>* r,
                        cplx<T>* xout, const std::vector<int>& want, double rtol, int maxit = 1000,
                        bool throw_on_cap = true);

inline int minres_complement(Solver& S, VecOps& V, const std::vector<double>& shifts, const double* dk, c64* r,
                             c64* xout, const std::vector<int>& want, double rtol, int maxit = 1000,
                             bool throw_on_cap = true) {
  return minres_complement_t<double>(S, V, shifts, dk, r, xout, want, rtol, maxit, throw_on_cap);
}

}  // namespace omx
