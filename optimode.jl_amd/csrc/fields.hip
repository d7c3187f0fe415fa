// Phase canonicalisation of eigenvectors (lib/MaxwellEigenmodes/src/fields.jl:258-260):
// evec *= exp(-i arg E_max), E_max the first entry of maximal |E|^2 in column-major
// (3, Nx, Ny) order of E = eps^-1 . fft(kx_tc(evec)).
#include "solver.hpp"

namespace omx {

namespace {

constexpr int FB = 256;

// per block: (max |E|^2, first index attaining it) over a slice of the 3N entries of vector v
__global__ __launch_bounds__(FB) void k_argmax_partial(const c64* D, long n3, double* pval, long* pidx) {
  const int v = blockIdx.y;
  const c64* E = D + long(v) * n3;
  double best = -1.0;
  long bi = 0;
  for (long e = long(blockIdx.x) * FB + threadIdx.x; e < n3; e += long(gridDim.x) * FB) {
    const double a = E[e].x * E[e].x + E[e].y * E[e].y;
    if (a > best) {
      best = a;
      bi = e;
    }
  }
  __shared__ double sv[FB];
  __shared__ long si[FB];
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = FB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = sv[threadIdx.x + s];
      const long oi = si[threadIdx.x + s];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pval[long(v) * gridDim.x + blockIdx.x] = sv[0];
    pidx[long(v) * gridDim.x + blockIdx.x] = si[0];
  }
}

// blocks (x, v): each finishes the argmax of vector v (the same fixed-order scan in every
// block), then multiplies its slice of the vector by exp(-i arg E_max)
__global__ __launch_bounds__(FB) void k_apply_phase(const c64* D, long n3, const double* pval, const long* pidx,
                                                    int nb, c64* vecs, long ldv, long K) {
  const int v = blockIdx.y;
  __shared__ c64 ph;
  if (threadIdx.x == 0) {
    double best = -1.0;
    long bi = 0;
    for (int b = 0; b < nb; ++b) {
      const double o = pval[long(v) * nb + b];
      const long oi = pidx[long(v) * nb + b];
      if (o > best || (o == best && oi < bi)) {
        best = o;
        bi = oi;
      }
    }
    const c64 e = D[long(v) * n3 + bi];
    const double a = sqrt(e.x * e.x + e.y * e.y);
    ph = a > 0 ? c64{e.x / a, -e.y / a} : c64{1.0, 0.0};
  }
  __syncthreads();
  const c64 p = ph;
  c64* x = vecs + long(v) * ldv;
  for (long r = long(blockIdx.x) * FB + threadIdx.x; r < K; r += long(gridDim.x) * FB) {
    const c64 a = x[r];
    x[r] = c64{a.x * p.x - a.y * p.y, a.x * p.y + a.y * p.x};
  }
}

}  // namespace

// vecs: nvec FP64 vectors (column j at vecs + j*ldv); d_k: per-vector k (device), laid out
// as instances of mper vectors each.
void Solver::canonicalize_phase(const double* d_k, c64* vecs, long ldv, long istride, int nvec, int mper) {
  const long n3 = 3 * N;
  field_scratch.ensure(sizeof(c64) * size_t(nvec) * n3);
  VecSel sel{nvec / mper, mper, nullptr, nullptr};
  BView<double> X{vecs, ldv, istride};
  launch_fields<double>(FIELD_E, p64.plan, X, field_scratch.as<c64>(), sel, d_k, 1.0, stream);
  const int nb = 64;
  argmax_v.ensure(sizeof(double) * nvec * nb);
  argmax_i.ensure(sizeof(long) * nvec * nb);
  hipLaunchKernelGGL(k_argmax_partial, dim3(nb, nvec), dim3(FB), 0, stream, field_scratch.as<c64>(), n3,
                     argmax_v.as<double>(), argmax_i.as<long>());
  hipLaunchKernelGGL(k_apply_phase, dim3(64, nvec), dim3(FB), 0, stream, field_scratch.as<c64>(), n3,
                     argmax_v.as<double>(), argmax_i.as<long>(), nb, vecs, ldv, K);
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace omx
