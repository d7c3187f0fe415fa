// EME per-slice mode sets on the GPU: mode fields, power normalisation and MEOW overlaps
// (lib/EigenmodeExpansion/src/modes.jl:95-170).
//
// A mode set (slot) holds E = i eps^-1 fft(kx_tc(h)) and H = fft(tc(h)) of nev modes,
// planar per mode and component so that the overlap kernel streams the x and y planes
// only.  build_modes divides both fields of mode j by sqrt(<m_j, m_j>) (build_mode,
// modes.jl:116-122); overlap forms O[i, j] = 1/2 sum_p (Ea_x,i Hb_y,j - Ea_y,i Hb_x,j) dA
// (inner_product / overlap_matrix, modes.jl:135-170), Ea conjugated on request.
#include <cmath>
#include <complex>
#include <vector>

#include "solver.hpp"

namespace omx {

namespace {

constexpr int EB = 256;
constexpr int OV_P = 32;     // pixels per staged chunk
constexpr int OV_MAX = 16;   // modes per set
constexpr int OV_BLOCKS = 1024;

// interleaved [v][p][c] fields of the field helpers -> planar slot fields; E gains the
// factor i of mode_fields
__global__ __launch_bounds__(EB) void k_modes_pack(const c64* Et, const c64* Ht, long N, c64* E, c64* H) {
  const int v = blockIdx.y;
  const c64* et = Et + long(v) * 3 * N;
  const c64* ht = Ht + long(v) * 3 * N;
  c64* e = E + long(v) * 3 * N;
  c64* h = H + long(v) * 3 * N;
  for (long p = long(blockIdx.x) * EB + threadIdx.x; p < N; p += long(gridDim.x) * EB)
    for (int c = 0; c < 3; ++c) {
      const c64 a = et[3 * p + c];
      e[c * N + p] = c64{-a.y, a.x};
      h[c * N + p] = ht[3 * p + c];
    }
}

// planar slot fields -> interleaved (3, Nx, Ny) column-major per mode
__global__ __launch_bounds__(EB) void k_modes_unpack(const c64* F, long N, c64* out) {
  const int v = blockIdx.y;
  const c64* f = F + long(v) * 3 * N;
  c64* o = out + long(v) * 3 * N;
  for (long p = long(blockIdx.x) * EB + threadIdx.x; p < N; p += long(gridDim.x) * EB)
    for (int c = 0; c < 3; ++c) o[3 * p + c] = f[c * N + p];
}

// both fields of mode v times s[v]
__global__ __launch_bounds__(EB) void k_modes_scale(c64* E, c64* H, long N, const c64* s) {
  const int v = blockIdx.y;
  const c64 f = s[v];
  c64* e = E + long(v) * 3 * N;
  c64* h = H + long(v) * 3 * N;
  for (long q = long(blockIdx.x) * EB + threadIdx.x; q < 3 * N; q += long(gridDim.x) * EB) {
    const c64 a = e[q], b = h[q];
    e[q] = c64{a.x * f.x - a.y * f.y, a.x * f.y + a.y * f.x};
    h[q] = c64{b.x * f.x - b.y * f.y, b.x * f.y + b.y * f.x};
  }
}

// Partial overlap sums: thread t of a block owns the pair (i, j) = (t % 16, t / 16) and
// accumulates Ea_x,i Hb_y,j - Ea_y,i Hb_x,j over the chunks of OV_P pixels the block
// visits; the chunk's x and y planes of both sets are staged in LDS (rows padded by one
// value against bank conflicts).  part[blockIdx.x * 256 + t].
__global__ __launch_bounds__(EB) void k_overlap(const c64* Ea, const c64* Hb, int na, int nb, long N, int conj_a,
                                                c64* part) {
  __shared__ c64 ex[OV_MAX][OV_P + 1], ey[OV_MAX][OV_P + 1], hx[OV_MAX][OV_P + 1], hy[OV_MAX][OV_P + 1];
  const int t = threadIdx.x, i = t % OV_MAX, j = t / OV_MAX;
  const bool mine = i < na && j < nb;
  const double sg = conj_a ? -1.0 : 1.0;
  double sr = 0, si = 0;
  for (long p0 = long(blockIdx.x) * OV_P; p0 < N; p0 += long(gridDim.x) * OV_P) {
    for (int e = t; e < OV_MAX * OV_P; e += EB) {
      const int v = e / OV_P, q = e % OV_P;
      const long p = p0 + q;
      const bool ina = v < na && p < N, inb = v < nb && p < N;
      const c64 z{0, 0};
      ex[v][q] = ina ? Ea[long(v * 3 + 0) * N + p] : z;
      ey[v][q] = ina ? Ea[long(v * 3 + 1) * N + p] : z;
      hx[v][q] = inb ? Hb[long(v * 3 + 0) * N + p] : z;
      hy[v][q] = inb ? Hb[long(v * 3 + 1) * N + p] : z;
    }
    __syncthreads();
    if (mine) {
#pragma unroll 8
      for (int q = 0; q < OV_P; ++q) {
        const c64 a = ex[i][q], b = hy[j][q], c = ey[i][q], d = hx[j][q];
        // (a b - c d) with a, c conjugated when sg = -1
        sr += a.x * b.x - sg * a.y * b.y - (c.x * d.x - sg * c.y * d.y);
        si += a.x * b.y + sg * a.y * b.x - (c.x * d.y + sg * c.y * d.x);
      }
    }
    __syncthreads();
  }
  part[long(blockIdx.x) * EB + t] = c64{sr, si};
}

// out[i + na j] = scale * sum over blocks of part[b * 256 + (i + 16 j)]
__global__ __launch_bounds__(EB) void k_overlap_reduce(const c64* part, int nblk, int na, int nb, double scale,
                                                       c64* out) {
  const int t = threadIdx.x, i = t % OV_MAX, j = t / OV_MAX;
  if (i >= na || j >= nb) return;
  double sr = 0, si = 0;
  for (int b = 0; b < nblk; ++b) {
    const c64 v = part[long(b) * EB + t];
    sr += v.x;
    si += v.y;
  }
  out[i + na * j] = c64{scale * sr, scale * si};
}

int grid_for(long n) { return int(std::min<long>(1024, (n + EB - 1) / EB)); }

}  // namespace

void Solver::build_modes(int slot, int nev, const double* kmags, const double* evecs, int conjugate, int normalize,
                         double* norms_out) {
  require_2d("EME mode sets");
  if (slot < 0 || slot >= MODE_SLOTS) throw Error(OMX_ERR_INVALID_ARG, "mode slot must be in 0..3");
  if (nev < 1 || nev > OV_MAX) throw Error(OMX_ERR_UNSUPPORTED, "nev must be in 1..16");
  if (!kmags) throw Error(OMX_ERR_INVALID_ARG, "kmags is NULL");
  for (int j = 0; j < nev; ++j)
    if (!(kmags[j] > 0) || !std::isfinite(kmags[j])) throw Error(OMX_ERR_INVALID_ARG, "kmags must be positive");
  if (!evecs && nev > last_nev)
    throw Error(OMX_ERR_INVALID_ARG, "evecs is NULL and the last solve holds fewer eigenvectors");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const long n3 = 3 * N;
  ModeSlot& ms = modes[slot];
  ms.E.ensure(sizeof(c64) * nev * n3);
  ms.H.ensure(sizeof(c64) * nev * n3);
  ms.nev = 0;
  DevMem dk(sizeof(double) * nev), Et(sizeof(c64) * nev * n3), Ht(sizeof(c64) * nev * n3), up;
  OMX_HIP_CHECK(hipMemcpyAsync(dk.as<void>(), kmags, sizeof(double) * nev, hipMemcpyHostToDevice, s));
  const c64* ev = last_evecs.as<c64>();
  if (evecs) {
    up.alloc(sizeof(c64) * nev * ld);
    for (int j = 0; j < nev; ++j)
      OMX_HIP_CHECK(hipMemcpyAsync(up.as<c64>() + j * ld, evecs + 2 * size_t(j) * K, sizeof(c64) * K,
                                   hipMemcpyHostToDevice, s));
    ev = up.as<c64>();
  }
  const VecSel sel{nev, 1, nullptr, nullptr};
  BView<double> EVv{const_cast<c64*>(ev), ld, ld};
  launch_fields<double>(FIELD_E, p64.plan, EVv, Et.as<c64>(), sel, dk.as<double>(), 1.0, s);
  launch_fields<double>(FIELD_H, p64.plan, EVv, Ht.as<c64>(), sel, dk.as<double>(), 1.0, s);
  hipLaunchKernelGGL(k_modes_pack, dim3(grid_for(N), nev), dim3(EB), 0, s, Et.as<c64>(), Ht.as<c64>(), N,
                     ms.E.as<c64>(), ms.H.as<c64>());
  OMX_HIP_CHECK(hipGetLastError());
  ms.nev = nev;
  std::vector<c64> self(size_t(nev) * nev);
  overlap(slot, slot, conjugate, reinterpret_cast<double*>(self.data()));
  std::vector<c64> scale(nev);
  for (int j = 0; j < nev; ++j) {
    const std::complex<double> nrm = std::sqrt(std::complex<double>(self[j + size_t(nev) * j].x,
                                                                    self[j + size_t(nev) * j].y));
    if (norms_out) {
      norms_out[2 * j] = nrm.real();
      norms_out[2 * j + 1] = nrm.imag();
    }
    if (!(std::abs(nrm) > 0)) throw Error(OMX_ERR_INVALID_ARG, "a mode has zero self overlap");
    const std::complex<double> inv = 1.0 / nrm;
    scale[j] = normalize ? c64{inv.real(), inv.imag()} : c64{1.0, 0.0};
  }
  if (normalize) {
    DevMem sd(sizeof(c64) * nev);
    OMX_HIP_CHECK(hipMemcpyAsync(sd.as<void>(), scale.data(), sizeof(c64) * nev, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_modes_scale, dim3(grid_for(n3), nev), dim3(EB), 0, s, ms.E.as<c64>(), ms.H.as<c64>(), N,
                       sd.as<c64>());
    OMX_HIP_CHECK(hipGetLastError());
    OMX_HIP_CHECK(hipStreamSynchronize(s));
  }
}

void Solver::overlap(int slot_a, int slot_b, int conjugate, double* out) {
  if (slot_a < 0 || slot_a >= MODE_SLOTS || slot_b < 0 || slot_b >= MODE_SLOTS)
    throw Error(OMX_ERR_INVALID_ARG, "mode slot must be in 0..3");
  if (!out) throw Error(OMX_ERR_INVALID_ARG, "out is NULL");
  const ModeSlot &A = modes[slot_a], &B = modes[slot_b];
  if (A.nev == 0 || B.nev == 0) throw Error(OMX_ERR_INVALID_ARG, "mode slot is empty");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const int nblk = int(std::min<long>(OV_BLOCKS, (N + OV_P - 1) / OV_P));
  DevMem part(sizeof(c64) * size_t(nblk) * EB), od(sizeof(c64) * A.nev * B.nev);
  hipLaunchKernelGGL(k_overlap, dim3(nblk), dim3(EB), 0, s, A.E.as<c64>(), B.H.as<c64>(), A.nev, B.nev, N,
                     conjugate ? 1 : 0, part.as<c64>());
  const double dA = (Dx / nx) * (Dy / ny);
  hipLaunchKernelGGL(k_overlap_reduce, dim3(1), dim3(EB), 0, s, part.as<c64>(), nblk, A.nev, B.nev, 0.5 * dA,
                     od.as<c64>());
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipMemcpyAsync(out, od.as<void>(), sizeof(c64) * A.nev * B.nev, hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
}

void Solver::get_mode_fields(int slot, double* E_out, double* H_out) {
  if (slot < 0 || slot >= MODE_SLOTS) throw Error(OMX_ERR_INVALID_ARG, "mode slot must be in 0..3");
  const ModeSlot& ms = modes[slot];
  if (ms.nev == 0) throw Error(OMX_ERR_INVALID_ARG, "mode slot is empty");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const long n3 = 3 * N;
  DevMem tmp(sizeof(c64) * ms.nev * n3);
  for (int f = 0; f < 2; ++f) {
    double* dst = f == 0 ? E_out : H_out;
    if (!dst) continue;
    hipLaunchKernelGGL(k_modes_unpack, dim3(grid_for(N), ms.nev), dim3(EB), 0, s,
                       (f == 0 ? ms.E : ms.H).as<c64>(), N, tmp.as<c64>());
    OMX_HIP_CHECK(hipGetLastError());
    OMX_HIP_CHECK(hipMemcpyAsync(dst, tmp.as<void>(), sizeof(c64) * ms.nev * n3, hipMemcpyDeviceToHost, s));
    OMX_HIP_CHECK(hipStreamSynchronize(s));
  }
}

}  // namespace omx
