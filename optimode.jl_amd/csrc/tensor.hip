// Per-pixel 3x3 tensor fields: the inverse sliceinv_3x3 (lib/MaxwellEigenmodes/src/
// linalg.jl:126-149) and its pull-back Abar = -Y^T Ybar Y^T (src/grads/linalg.jl:40-47),
// FP64, one thread per pixel.  Fields are column-major (3,3,...) host arrays: pixel p
// holds entries (a, b) at 9p + a + 3b.
#include "solver.hpp"

namespace omx {

namespace {

constexpr int TB = 256;

// inverse by the adjugate: Y = adj(A) / det(A)
__global__ __launch_bounds__(TB) void k_sliceinv(const double* A, double* Y, long npix) {
  const long p = long(blockIdx.x) * TB + threadIdx.x;
  if (p >= npix) return;
  double a[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = A[9 * p + i];
  // a[r + 3c] = A(r, c); C(r, c) the cofactors
  const double C00 = a[4] * a[8] - a[7] * a[5];
  const double C01 = a[7] * a[2] - a[1] * a[8];
  const double C02 = a[1] * a[5] - a[4] * a[2];
  const double C10 = a[6] * a[5] - a[3] * a[8];
  const double C11 = a[0] * a[8] - a[6] * a[2];
  const double C12 = a[3] * a[2] - a[0] * a[5];
  const double C20 = a[3] * a[7] - a[6] * a[4];
  const double C21 = a[6] * a[1] - a[0] * a[7];
  const double C22 = a[0] * a[4] - a[3] * a[1];
  const double det = a[0] * C00 + a[1] * C10 + a[2] * C20;  // along the first column
  const double id = 1.0 / det;
  // Y(r, c) = C(c, r) / det, stored at r + 3c
  double* y = Y + 9 * p;
  y[0] = C00 * id;
  y[1] = C01 * id;
  y[2] = C02 * id;
  y[3] = C10 * id;
  y[4] = C11 * id;
  y[5] = C12 * id;
  y[6] = C20 * id;
  y[7] = C21 * id;
  y[8] = C22 * id;
}

// Abar = -Y^T Ybar Y^T per pixel
__global__ __launch_bounds__(TB) void k_sliceinv_pullback(const double* Y, const double* Yb, double* Ab, long npix) {
  const long p = long(blockIdx.x) * TB + threadIdx.x;
  if (p >= npix) return;
  double y[9], b[9], t[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    y[i] = Y[9 * p + i];
    b[i] = Yb[9 * p + i];
  }
  // t = Y^T Ybar: t(r, c) = sum_k Y(k, r) Ybar(k, c)
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) t[r + 3 * c] = y[3 * r] * b[3 * c] + y[1 + 3 * r] * b[1 + 3 * c] + y[2 + 3 * r] * b[2 + 3 * c];
  // Abar(r, c) = -sum_k t(r, k) Y(c, k)
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Ab[9 * p + r + 3 * c] = -(t[r] * y[c] + t[r + 3] * y[c + 3] + t[r + 6] * y[c + 6]);
}

void check_device(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
    throw Error(OMX_ERR_NO_DEVICE, "no such HIP device");
  OMX_HIP_CHECK(hipSetDevice(device));
}

}  // namespace

void sliceinv_3x3(int device, long npix, const double* A, double* Y_out) {
  if (npix < 0 || (npix > 0 && (!A || !Y_out))) throw Error(OMX_ERR_INVALID_ARG, "bad sliceinv_3x3 arguments");
  if (npix == 0) return;
  check_device(device);
  const size_t bytes = sizeof(double) * 9 * size_t(npix);
  DevMem dA(bytes), dY(bytes);
  OMX_HIP_CHECK(hipMemcpy(dA.as<void>(), A, bytes, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sliceinv, dim3(unsigned((npix + TB - 1) / TB)), dim3(TB), 0, 0, dA.as<double>(), dY.as<double>(),
                     npix);
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipMemcpy(Y_out, dY.as<void>(), bytes, hipMemcpyDeviceToHost));
}

void sliceinv_3x3_pullback(int device, long npix, const double* Y, const double* Ybar, double* Abar_out) {
  if (npix < 0 || (npix > 0 && (!Y || !Ybar || !Abar_out)))
    throw Error(OMX_ERR_INVALID_ARG, "bad sliceinv_3x3_pullback arguments");
  if (npix == 0) return;
  check_device(device);
  const size_t bytes = sizeof(double) * 9 * size_t(npix);
  DevMem dY(bytes), dB(bytes), dA(bytes);
  OMX_HIP_CHECK(hipMemcpy(dY.as<void>(), Y, bytes, hipMemcpyHostToDevice));
  OMX_HIP_CHECK(hipMemcpy(dB.as<void>(), Ybar, bytes, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sliceinv_pullback, dim3(unsigned((npix + TB - 1) / TB)), dim3(TB), 0, 0, dY.as<double>(),
                     dB.as<double>(), dA.as<double>(), npix);
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipMemcpy(Abar_out, dA.as<void>(), bytes, hipMemcpyDeviceToHost));
}

}  // namespace omx
