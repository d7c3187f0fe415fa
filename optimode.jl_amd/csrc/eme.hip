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
#include <cstring>
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
  // kept for the pull-back: the eigenvectors and their k
  ms.ev.ensure(sizeof(c64) * nev * ld);
  ms.kd.ensure(sizeof(double) * nev);
  OMX_HIP_CHECK(hipMemcpyAsync(ms.ev.as<void>(), ev, sizeof(c64) * nev * ld, hipMemcpyDeviceToDevice, s));
  OMX_HIP_CHECK(hipMemcpyAsync(ms.kd.as<void>(), dk.as<void>(), sizeof(double) * nev, hipMemcpyDeviceToDevice, s));
  ms.conj = conjugate ? 1 : 0;
  ms.normalized = normalize ? 1 : 0;
  ms.bar_live = 0;
  ms.nrm.assign(nev, c64{1.0, 0.0});
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
    if (normalize) ms.nrm[j] = c64{nrm.real(), nrm.imag()};
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


// ---------------------------------------------------------------------------------------
// Pull-backs (cotangent convention of the reference's Zygote rules: a real scalar L of a
// complex array Z changes by Re sum conj(Zbar) dZ)
// ---------------------------------------------------------------------------------------
namespace {

__device__ __forceinline__ c64 cm(c64 a, c64 b) { return c64{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ c64 cj(c64 a) { return c64{a.x, -a.y}; }
__device__ __forceinline__ c64 ca(c64 a, c64 b) { return c64{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c64 cs(double s, c64 a) { return c64{s * a.x, s * a.y}; }

// Overlap pull-back (inner_product, modes.jl:135-148): with O[i, j] = h sum_p (ex_i hy_j -
// ey_i hx_j), ex_i conjugated when conj, the cotangent Obar[i + na j] adds
//   Ebar_a,i,x += h sum_j Obar_ij conj(hy_j)   (conj: conj(Obar_ij) hy_j), Ebar_a,i,y likewise with -hx_j
//   Hbar_b,j,y += h sum_i Obar_ij conj(ex_i)   (conj: Obar_ij ex_i),       Hbar_b,j,x likewise with -ey_i
// one thread per pixel; which & 1: the E side, which & 2: the H side.
__global__ __launch_bounds__(EB) void k_overlap_pb(const c64* Ea, const c64* Hb, const c64* obar, int na, int nb,
                                                   long N, int conj, double h, int which, c64* Eab, c64* Hbb) {
  __shared__ c64 o[OV_MAX * OV_MAX];
  for (int e = threadIdx.x; e < na * nb; e += EB) o[e] = obar[e];
  __syncthreads();
  for (long p = long(blockIdx.x) * EB + threadIdx.x; p < N; p += long(gridDim.x) * EB) {
    if (which & 1) {
      c64 hx[OV_MAX], hy[OV_MAX];
      for (int j = 0; j < nb; ++j) {
        hx[j] = Hb[long(j * 3 + 0) * N + p];
        hy[j] = Hb[long(j * 3 + 1) * N + p];
      }
      for (int i = 0; i < na; ++i) {
        c64 gx{0, 0}, gy{0, 0};
        for (int j = 0; j < nb; ++j) {
          const c64 w = o[i + na * j];
          if (conj) {
            gx = ca(gx, cm(cj(w), hy[j]));
            gy = ca(gy, cm(cj(w), hx[j]));
          } else {
            gx = ca(gx, cm(w, cj(hy[j])));
            gy = ca(gy, cm(w, cj(hx[j])));
          }
        }
        c64* ex = Eab + long(i * 3 + 0) * N + p;
        c64* ey = Eab + long(i * 3 + 1) * N + p;
        *ex = ca(*ex, cs(h, gx));
        *ey = ca(*ey, cs(-h, gy));
      }
    }
    if (which & 2) {
      c64 ex[OV_MAX], ey[OV_MAX];
      for (int i = 0; i < na; ++i) {
        ex[i] = Ea[long(i * 3 + 0) * N + p];
        ey[i] = Ea[long(i * 3 + 1) * N + p];
      }
      for (int j = 0; j < nb; ++j) {
        c64 gy{0, 0}, gx{0, 0};
        for (int i = 0; i < na; ++i) {
          const c64 w = o[i + na * j];
          if (conj) {
            gy = ca(gy, cm(w, ex[i]));
            gx = ca(gx, cm(w, ey[i]));
          } else {
            gy = ca(gy, cm(w, cj(ex[i])));
            gx = ca(gx, cm(w, cj(ey[i])));
          }
        }
        c64* hy = Hbb + long(j * 3 + 1) * N + p;
        c64* hx = Hbb + long(j * 3 + 0) * N + p;
        *hy = ca(*hy, cs(h, gy));
        *hx = ca(*hx, cs(-h, gx));
      }
    }
  }
}

// per mode v and block: partial sums of conj(Ebar) . E + conj(Hbar) . H over the 3N values
__global__ __launch_bounds__(EB) void k_mode_dots(const c64* E, const c64* H, const c64* Eb, const c64* Hb, long n3,
                                                  c64* part) {
  const int v = blockIdx.y;
  const c64 *e = E + long(v) * n3, *hh = H + long(v) * n3, *eb = Eb + long(v) * n3, *hb = Hb + long(v) * n3;
  double sr = 0, si = 0;
  for (long q = long(blockIdx.x) * EB + threadIdx.x; q < n3; q += long(gridDim.x) * EB) {
    const c64 a = cm(cj(eb[q]), e[q]), b = cm(cj(hb[q]), hh[q]);
    sr += a.x + b.x;
    si += a.y + b.y;
  }
  __shared__ double rr[EB], ri[EB];
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  for (int st = EB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      rr[threadIdx.x] += rr[threadIdx.x + st];
      ri[threadIdx.x] += ri[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[long(v) * gridDim.x + blockIdx.x] = c64{rr[0], ri[0]};
}

// Normalisation pull-back (build_mode, modes.jl:116-122) and the start of the field pull-back,
// one thread per pixel over all modes.  With E = E_r / nrm, H = H_r / nrm, s = nrm^2 and
// c_v = 1/2 s^-1 (conj(Ebar) . E + conj(Hbar) . H):
//   Ebar_r = Ebar / conj(nrm) + (c terms of ds), Hbar_r likewise (cf[v] = c_v, ninv[v] = 1/conj(nrm))
// then gE = -i eps^-1 Ebar_r and gH = Hbar_r go out interleaved [v][p][c] for the inverse FFT,
// and epsinvbar[a + 3b + 9p] = sum_v Re(conj(Ebar_r,a) i D_b) (D = fft2(kx_tc(h)), [v][p][c]).
__global__ __launch_bounds__(EB) void k_mode_raw(const c64* E, const c64* H, const c64* Eb, const c64* Hb,
                                                 const c64* nrm, const c64* cf, const c64* ninv, int nev, long N,
                                                 int conj, double h, const double* eps6, const c64* D, c64* gE,
                                                 c64* gH, double* eibar) {
  for (long p = long(blockIdx.x) * EB + threadIdx.x; p < N; p += long(gridDim.x) * EB) {
    const double* ep = eps6 + 6 * p;
    const double exx = ep[0], eyy = ep[1], ezz = ep[2], exy = ep[3], exz = ep[4], eyz = ep[5];
    double acc[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) acc[a] = 0.0;
    for (int v = 0; v < nev; ++v) {
      c64 er[3], hr[3], ebr[3], hbr[3];
      const c64 n0 = nrm[v], c = cf[v], ni = ninv[v];
      for (int a = 0; a < 3; ++a) {
        er[a] = cm(n0, E[long(v * 3 + a) * N + p]);
        hr[a] = cm(n0, H[long(v * 3 + a) * N + p]);
        ebr[a] = cm(ni, Eb[long(v * 3 + a) * N + p]);
        hbr[a] = cm(ni, Hb[long(v * 3 + a) * N + p]);
      }
      if (conj) {
        ebr[0] = ca(ebr[0], cs(-h, cm(c, hr[1])));
        ebr[1] = ca(ebr[1], cs(h, cm(c, hr[0])));
        hbr[1] = ca(hbr[1], cs(-h, cm(cj(c), er[0])));
        hbr[0] = ca(hbr[0], cs(h, cm(cj(c), er[1])));
      } else {
        ebr[0] = ca(ebr[0], cs(-h, cm(cj(c), cj(hr[1]))));
        ebr[1] = ca(ebr[1], cs(h, cm(cj(c), cj(hr[0]))));
        hbr[1] = ca(hbr[1], cs(-h, cm(cj(c), cj(er[0]))));
        hbr[0] = ca(hbr[0], cs(h, cm(cj(c), cj(er[1]))));
      }
      // eps^-1 . Ebar_r (symmetric), times -i
      const c64 t0 = ca(ca(cs(exx, ebr[0]), cs(exy, ebr[1])), cs(exz, ebr[2]));
      const c64 t1 = ca(ca(cs(exy, ebr[0]), cs(eyy, ebr[1])), cs(eyz, ebr[2]));
      const c64 t2 = ca(ca(cs(exz, ebr[0]), cs(eyz, ebr[1])), cs(ezz, ebr[2]));
      c64* ge = gE + (long(v) * N + p) * 3;
      ge[0] = c64{t0.y, -t0.x};
      ge[1] = c64{t1.y, -t1.x};
      ge[2] = c64{t2.y, -t2.x};
      c64* gh = gH + (long(v) * N + p) * 3;
      gh[0] = hbr[0];
      gh[1] = hbr[1];
      gh[2] = hbr[2];
      const c64* d = D + (long(v) * N + p) * 3;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          // Re(conj(ebr_a) i d_b) = -Im(conj(ebr_a) d_b)
          const c64 z = cm(cj(ebr[a]), d[b]);
          acc[a + 3 * b] += -z.y;
        }
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) eibar[9 * p + a] = acc[a];
  }
}

// Plane-wave end of the field pull-back, mode v, plane wave (ix, iy) at k_v (mode_fields,
// modes.jl:105-113): evbar = kx_tc^H gE + tc^H gH (gE, gH after the inverse FFT, [v][p][c]),
// and partial sums of the k derivative Re(conj(gE) . dkx_tc/dk(h) + conj(gH) . dtc/dk(h)) with
// d(m |k+G|)/dk = n x z, d|k+G|/dk = k/|k+G| and n independent of k.
__global__ __launch_bounds__(EB) void k_mode_pw(const c64* gE, const c64* gH, const c64* ev, long ld, const double* kd,
                                                int nx, int ny, double invdx, double invdy, c64* evbar, double* kpart) {
  const int v = blockIdx.y;
  const long N = long(nx) * ny;
  const double k = kd[v];
  double acc = 0.0;
  for (long p = long(blockIdx.x) * EB + threadIdx.x; p < N; p += long(gridDim.x) * EB) {
    const int ix = int(p % nx), iy = int(p / nx);
    const Basis<double> bs = basis_at<double>(ix, iy, nx, ny, invdx, invdy, k);
    const double m[3] = {bs.m0, bs.m1, bs.m2}, n[3] = {bs.n0, bs.n1, bs.n2};
    const c64* ge = gE + (long(v) * N + p) * 3;
    const c64* gh = gH + (long(v) * N + p) * 3;
    c64 en{0, 0}, em{0, 0}, hm{0, 0}, hn{0, 0};
    for (int a = 0; a < 3; ++a) {
      en = ca(en, cs(n[a], ge[a]));
      em = ca(em, cs(m[a], ge[a]));
      hm = ca(hm, cs(m[a], gh[a]));
      hn = ca(hn, cs(n[a], gh[a]));
    }
    c64* o = evbar + long(v) * 2 * N + 2 * p;
    o[0] = ca(cs(-bs.mag, en), hm);
    o[1] = ca(cs(bs.mag, em), hn);
    const c64 h1 = ev[long(v) * ld + 2 * p], h2 = ev[long(v) * ld + 2 * p + 1];
    const double nxz[3] = {n[1], -n[0], 0.0};
    const double kr = k / bs.mag;
    double s = 0.0;
    for (int a = 0; a < 3; ++a) {
      const c64 dkx = ca(cs(nxz[a], h2), cs(-n[a] * kr, h1));
      const c64 dtc = cs((nxz[a] - m[a] * kr) / bs.mag, h1);
      s += (ge[a].x * dkx.x + ge[a].y * dkx.y) + (gh[a].x * dtc.x + gh[a].y * dtc.y);
    }
    acc += s;
  }
  __shared__ double red[EB];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = EB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) kpart[long(v) * gridDim.x + blockIdx.x] = red[0];
}

}  // namespace

void Solver::overlap_pullback(int slot_a, int slot_b, int conjugate, const double* obar, int which) {
  require_2d("EME mode sets");
  if (slot_a < 0 || slot_a >= MODE_SLOTS || slot_b < 0 || slot_b >= MODE_SLOTS)
    throw Error(OMX_ERR_INVALID_ARG, "mode slot must be in 0..3");
  if (!obar) throw Error(OMX_ERR_INVALID_ARG, "obar is NULL");
  if (slot_a == slot_b && (which & 3) == 3)
    throw Error(OMX_ERR_INVALID_ARG, "a set's overlap with itself: pull back one side at a time");
  ModeSlot &A = modes[slot_a], &B = modes[slot_b];
  if (A.nev == 0 || B.nev == 0) throw Error(OMX_ERR_INVALID_ARG, "mode slot is empty");
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const long n3 = 3 * N;
  for (ModeSlot* m : {&A, &B}) {
    if (m->bar_live) continue;
    m->Ebar.ensure(sizeof(c64) * m->nev * n3);
    m->Hbar.ensure(sizeof(c64) * m->nev * n3);
    OMX_HIP_CHECK(hipMemsetAsync(m->Ebar.as<void>(), 0, sizeof(c64) * m->nev * n3, s));
    OMX_HIP_CHECK(hipMemsetAsync(m->Hbar.as<void>(), 0, sizeof(c64) * m->nev * n3, s));
    m->bar_live = 1;
  }
  DevMem od(sizeof(c64) * A.nev * B.nev);
  OMX_HIP_CHECK(hipMemcpyAsync(od.as<void>(), obar, sizeof(c64) * A.nev * B.nev, hipMemcpyHostToDevice, s));
  const double h = 0.5 * (Dx / nx) * (Dy / ny);
  hipLaunchKernelGGL(k_overlap_pb, dim3(grid_for(N)), dim3(EB), 0, s, A.E.as<c64>(), B.H.as<c64>(), od.as<c64>(),
                     A.nev, B.nev, N, conjugate ? 1 : 0, h, which & 3, A.Ebar.as<c64>(), B.Hbar.as<c64>());
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipStreamSynchronize(s));
}

void Solver::mode_pullback(int slot, double* kbar_out, double* evbar_out, double* epsinvbar_out) {
  require_2d("EME mode sets");
  if (slot < 0 || slot >= MODE_SLOTS) throw Error(OMX_ERR_INVALID_ARG, "mode slot must be in 0..3");
  ModeSlot& ms = modes[slot];
  if (ms.nev == 0) throw Error(OMX_ERR_INVALID_ARG, "mode slot is empty");
  if (!kbar_out || !evbar_out || !epsinvbar_out) throw Error(OMX_ERR_INVALID_ARG, "NULL output");
  const int nev = ms.nev;
  OMX_HIP_CHECK(hipSetDevice(device));
  hipStream_t s = stream;
  const long n3 = 3 * N;
  if (!ms.bar_live) {  // no cotangent reached the set: every output is zero
    std::memset(kbar_out, 0, sizeof(double) * nev);
    std::memset(evbar_out, 0, sizeof(c64) * nev * K);
    std::memset(epsinvbar_out, 0, sizeof(double) * 9 * N);
    return;
  }
  // c_v = 1/2 s^-1 (conj(Ebar) . E + conj(Hbar) . H)
  const int nb = grid_for(n3);
  DevMem part(sizeof(c64) * nev * nb);
  hipLaunchKernelGGL(k_mode_dots, dim3(nb, nev), dim3(EB), 0, s, ms.E.as<c64>(), ms.H.as<c64>(), ms.Ebar.as<c64>(),
                     ms.Hbar.as<c64>(), n3, part.as<c64>());
  std::vector<c64> hp(size_t(nev) * nb);
  OMX_HIP_CHECK(hipMemcpyAsync(hp.data(), part.as<void>(), sizeof(c64) * hp.size(), hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<c64> cf(3 * size_t(nev));  // nrm, c, 1/conj(nrm)
  for (int v = 0; v < nev; ++v) {
    std::complex<double> d(0, 0);
    for (int b = 0; b < nb; ++b) d += std::complex<double>(hp[size_t(v) * nb + b].x, hp[size_t(v) * nb + b].y);
    const std::complex<double> n0(ms.nrm[v].x, ms.nrm[v].y);
    const std::complex<double> c = ms.normalized ? 0.5 * d / (n0 * n0) : std::complex<double>(0, 0);
    const std::complex<double> ni = 1.0 / std::conj(n0);
    cf[v] = c64{n0.real(), n0.imag()};
    cf[nev + v] = c64{c.real(), c.imag()};
    cf[2 * nev + v] = c64{ni.real(), ni.imag()};
  }
  DevMem cd(sizeof(c64) * 3 * nev), D(sizeof(c64) * nev * n3), G(sizeof(c64) * 2 * nev * n3),
      eib(sizeof(double) * 9 * N), evb(sizeof(c64) * nev * 2 * N), kp;
  OMX_HIP_CHECK(hipMemcpyAsync(cd.as<void>(), cf.data(), sizeof(c64) * cf.size(), hipMemcpyHostToDevice, s));
  const VecSel sel{nev, 1, nullptr, nullptr};
  BView<double> EVv{ms.ev.as<c64>(), ld, ld};
  launch_fields<double>(FIELD_D, p64.plan, EVv, D.as<c64>(), sel, ms.kd.as<double>(), 1.0, s);
  const double h = 0.5 * (Dx / nx) * (Dy / ny);
  c64* gE = G.as<c64>();
  c64* gH = gE + nev * n3;
  hipLaunchKernelGGL(k_mode_raw, dim3(grid_for(N)), dim3(EB), 0, s, ms.E.as<c64>(), ms.H.as<c64>(), ms.Ebar.as<c64>(),
                     ms.Hbar.as<c64>(), cd.as<c64>(), cd.as<c64>() + nev, cd.as<c64>() + 2 * nev, nev, N, ms.conj, h,
                     p64.plan.eps6, D.as<c64>(), gE, gH, eib.as<double>());
  // both sets of cotangent fields through one inverse FFT launch (2 nev vectors)
  const VecSel sel2{2 * nev, 1, nullptr, nullptr};
  launch_fields<double>(FIELD_BFFT, p64.plan, EVv, gE, sel2, ms.kd.as<double>(), 1.0, s);
  const int nbp = grid_for(N);
  kp.alloc(sizeof(double) * nev * nbp);
  hipLaunchKernelGGL(k_mode_pw, dim3(nbp, nev), dim3(EB), 0, s, gE, gH, ms.ev.as<c64>(), ld, ms.kd.as<double>(), nx, ny,
                     1.0 / Dx, 1.0 / Dy, evb.as<c64>(), kp.as<double>());
  OMX_HIP_CHECK(hipGetLastError());
  std::vector<double> hk(size_t(nev) * nbp);
  OMX_HIP_CHECK(hipMemcpyAsync(hk.data(), kp.as<void>(), sizeof(double) * hk.size(), hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipMemcpyAsync(evbar_out, evb.as<void>(), sizeof(c64) * nev * 2 * N, hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipMemcpyAsync(epsinvbar_out, eib.as<void>(), sizeof(double) * 9 * N, hipMemcpyDeviceToHost, s));
  OMX_HIP_CHECK(hipStreamSynchronize(s));
  for (int v = 0; v < nev; ++v) {
    double t = 0;
    for (int b = 0; b < nbp; ++b) t += hk[size_t(v) * nbp + b];
    kbar_out[v] = t;
  }
  ms.bar_live = 0;  // consumed
}

}  // namespace omx
