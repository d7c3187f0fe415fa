// Common device/host helpers for the omx HIP library (gfx950 / MI355X).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace omx {

// Complex number of real type T, 8 B (float) or 16 B (double), interleaved (re, im) --
// the same memory layout as Julia's ComplexF32/ComplexF64.
template <typename T> struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};
using c32 = cplx<float>;
using c64 = cplx<double>;

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T a, T b) { return cplx<T>{a, b}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return {s * a.x, s * a.y}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, T s) { return {s * a.x, s * a.y}; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.x, -a.y}; }
// conj(a) * b
template <typename T> __host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename T> __host__ __device__ __forceinline__ T abs2(cplx<T> a) { return a.x * a.x + a.y * a.y; }

// Per-grid constants shared by every kernel of one operator family.
struct GridDesc {
  int nx, ny;        // pixels
  double dx, dy;     // cell size (um)
};

// fftfreq index -> signed integer frequency (lib/DielectricSmoothing/src/grid.jl:241-243)
__host__ __device__ __forceinline__ int fft_index(int i, int n) {
  return (n % 2 == 0) ? (i < n / 2 ? i : i - n) : (i <= (n - 1) / 2 ? i : i - n);
}

}  // namespace omx

#define OMX_HIP_CHECK(expr)                                                       \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) throw omx::HipError(_e, #expr, __FILE__, __LINE__);     \
  } while (0)
