// Adjoint (pull-back) of solve_k: placeholder until the GPU adjoint lands.
#include "solver.hpp"

namespace omx {

void Solver::adjoint(double, int, const double*, const double*, const double*, const double*, const uint8_t*, double*,
                     double*) {
  throw Error(OMX_ERR_UNSUPPORTED, "solve_k adjoint is not implemented yet");
}

}  // namespace omx
