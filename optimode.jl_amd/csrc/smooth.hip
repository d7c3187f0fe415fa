// Kottke sub-pixel smoothing of convex-polygon cross-sections into eps^-1, on the GPU.
//
// Follows lib/DielectricSmoothing/src/kottke.jl:44-95 (tau transform, avg_param) and
// smooth.jl:115-170 (uniform / single-interface / multi-material pixels).  The reference
// locates the interface with GeometryPrimitives (surfpt_nearby, volfrac), which is not
// available; here each pixel is supersampled on an nsub x nsub lattice:
//   - every sample takes the first shape that contains it (the background otherwise);
//   - one material: that material's tensor;
//   - two or more: the two most frequent materials m1, m2 (ties to the lower index), fill
//     fraction f = count(m1) / nsub^2, interface normal from the first moment of the m1
//     indicator (pointing from m1 into m2), and the Kottke average
//       eps = S tau^-1( f tau(S^T eps1 S) + (1 - f) tau(S^T eps2 S) ) S^T,
//     S = [n, z x n, z]; when the moment vanishes or more than two materials are present,
//     the harmonic mean inv(f inv(eps1) + (1 - f) inv(eps2)).
// The result is symmetrised, inverted and symmetrised again.  One thread per pixel.
#include "solver.hpp"
#include "omx_errors.hpp"

#include <vector>

// sample positions and the containment test round exactly as the host's numpy code does
#pragma clang fp contract(off)

namespace omx {

namespace {

constexpr int SB = 256;
constexpr int MAXMAT = 16;  // background + up to 15 shapes

struct M3 {
  double a[3][3];  // a[row][col]
};

__device__ __forceinline__ M3 mmul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j];
  return r;
}
__device__ __forceinline__ M3 mtrans(const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[j][i];
  return r;
}
__device__ __forceinline__ M3 minv(const M3& m) {
  const double(&a)[3][3] = m.a;
  const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  const double id = 1.0 / det;
  M3 r;
  r.a[0][0] = c00 * id;
  r.a[1][0] = c01 * id;
  r.a[2][0] = c02 * id;
  r.a[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  r.a[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  r.a[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  r.a[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  r.a[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  r.a[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
  return r;
}
// tau transform in interface coordinates (axis 0 normal), kottke.jl tau_trans
__device__ __forceinline__ M3 tau(const M3& e) {
  const double i = 1.0 / e.a[0][0];
  M3 t;
  t.a[0][0] = -i;
  t.a[1][0] = e.a[1][0] * i;
  t.a[2][0] = e.a[2][0] * i;
  t.a[0][1] = e.a[0][1] * i;
  t.a[1][1] = e.a[1][1] - e.a[1][0] * e.a[0][1] * i;
  t.a[2][1] = e.a[2][1] - e.a[2][0] * e.a[0][1] * i;
  t.a[0][2] = e.a[0][2] * i;
  t.a[1][2] = e.a[1][2] - e.a[1][0] * e.a[0][2] * i;
  t.a[2][2] = e.a[2][2] - e.a[2][0] * e.a[0][2] * i;
  return t;
}
// inverse tau transform, kottke.jl tau_inv_trans
__device__ __forceinline__ M3 tau_inv(const M3& t) {
  const double i = 1.0 / t.a[0][0];
  M3 e;
  e.a[0][0] = -i;
  e.a[1][0] = -t.a[1][0] * i;
  e.a[2][0] = -t.a[2][0] * i;
  e.a[0][1] = -t.a[0][1] * i;
  e.a[1][1] = t.a[1][1] - t.a[1][0] * t.a[0][1] * i;
  e.a[2][1] = t.a[2][1] - t.a[2][0] * t.a[0][1] * i;
  e.a[0][2] = -t.a[0][2] * i;
  e.a[1][2] = t.a[1][2] - t.a[1][0] * t.a[0][2] * i;
  e.a[2][2] = t.a[2][2] - t.a[2][0] * t.a[0][2] * i;
  return e;
}

struct Geometry {
  int nshape;
  const int* v0;      // first vertex of shape s
  const int* nv;      // vertex count of shape s
  const double* vxy;  // (x, y) pairs, counter-clockwise per shape
  const double* mats; // 9 per material, column-major 3x3; material 0 = background, s + 1 = shape s
};

__device__ __forceinline__ int material_at(const Geometry& G, double x, double y) {
  for (int s = 0; s < G.nshape; ++s) {
    const double* v = G.vxy + 2 * G.v0[s];
    const int n = G.nv[s];
    bool in = true;
    for (int k = 0; k < n && in; ++k) {
      const int k1 = (k + 1 == n) ? 0 : k + 1;
      const double x0 = v[2 * k], y0 = v[2 * k + 1], x1 = v[2 * k1], y1 = v[2 * k1 + 1];
      in = (x1 - x0) * (y - y0) - (y1 - y0) * (x - x0) >= 0.0;
    }
    if (in) return s + 1;
  }
  return 0;
}

__device__ __forceinline__ M3 material(const Geometry& G, int m) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = G.mats[9 * m + i + 3 * j];
  return r;
}

__global__ __launch_bounds__(SB) void k_smooth(int nx, int ny, double Dx, double Dy, int nsub, Geometry G,
                                               double* out) {
  const int p = blockIdx.x * SB + threadIdx.x;
  if (p >= nx * ny) return;
  const int ix = p % nx, iy = p / nx;
  const double dx = Dx / nx, dy = Dy / ny;
  const double xc = -0.5 * Dx + ix * dx, yc = -0.5 * Dy + iy * dy;
  const int nm = G.nshape + 1;
  int counts[MAXMAT];
  for (int m = 0; m < MAXMAT; ++m) counts[m] = 0;
  for (int a = 0; a < nsub; ++a) {
    const double x = xc + ((a + 0.5) / nsub - 0.5) * dx;
    for (int b = 0; b < nsub; ++b) {
      const double y = yc + ((b + 0.5) / nsub - 0.5) * dy;
      ++counts[material_at(G, x, y)];
    }
  }
  int m1 = 0, npres = 0;
  for (int m = 0; m < nm; ++m) {
    npres += counts[m] > 0 ? 1 : 0;
    if (counts[m] > counts[m1]) m1 = m;
  }
  M3 eps;
  if (npres == 1) {
    eps = material(G, m1);
  } else {
    int m2 = (m1 == 0) ? 1 : 0;
    for (int m = 0; m < nm; ++m)
      if (m != m1 && counts[m] > counts[m2]) m2 = m;
    const int ns = nsub * nsub;
    const double f = double(counts[m1]) / ns;
    // first moment of the m1 indicator about its mean, in exact integers: sub-sample
    // offsets are (2a + 1 - nsub) d / (2 nsub), which sum to zero, so the mean term drops
    // out and the common factor 1 / (2 nsub) cancels in the normalisation
    int ix1 = 0, iy1 = 0;
    for (int a = 0; a < nsub; ++a) {
      const double x = xc + ((a + 0.5) / nsub - 0.5) * dx;
      for (int b = 0; b < nsub; ++b) {
        const double y = yc + ((b + 0.5) / nsub - 0.5) * dy;
        if (material_at(G, x, y) == m1) {
          ix1 += 2 * a + 1 - nsub;
          iy1 += 2 * b + 1 - nsub;
        }
      }
    }
    const double gx = double(ix1) * dx, gy = double(iy1) * dy;
    const double nrm = sqrt(gx * gx + gy * gy);
    const M3 e1 = material(G, m1), e2 = material(G, m2);
    if (nrm == 0.0 || npres > 2) {
      const M3 i1 = minv(e1), i2 = minv(e2);
      M3 h;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) h.a[i][j] = f * i1.a[i][j] + (1.0 - f) * i2.a[i][j];
      eps = minv(h);
    } else {
      const double n0 = -gx / nrm, n1 = -gy / nrm;
      M3 S;  // columns n, z x n, z
      S.a[0][0] = n0;
      S.a[1][0] = n1;
      S.a[2][0] = 0.0;
      S.a[0][1] = -n1;
      S.a[1][1] = n0;
      S.a[2][1] = 0.0;
      S.a[0][2] = 0.0;
      S.a[1][2] = 0.0;
      S.a[2][2] = 1.0;
      const M3 St = mtrans(S);
      const M3 t1 = tau(mmul(St, mmul(e1, S))), t2 = tau(mmul(St, mmul(e2, S)));
      M3 ta;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ta.a[i][j] = f * t1.a[i][j] + (1.0 - f) * t2.a[i][j];
      eps = mmul(S, mmul(tau_inv(ta), St));
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j) eps.a[i][j] = eps.a[j][i] = 0.5 * (eps.a[i][j] + eps.a[j][i]);
  M3 ei = minv(eps);
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j) ei.a[i][j] = ei.a[j][i] = 0.5 * (ei.a[i][j] + ei.a[j][i]);
  double* o = out + 9 * long(p);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[i + 3 * j] = ei.a[i][j];
}

}  // namespace

void smooth_epsinv(int device, int nx, int ny, double Dx, double Dy, int nshape, const int* nverts,
                   const double* verts, const double* eps_shapes, const double* eps_background, int nsub,
                   double* epsinv_out) {
  if (nx < 2 || ny < 2 || !(Dx > 0) || !(Dy > 0)) throw Error(OMX_ERR_BAD_SHAPE, "bad grid");
  if (nshape < 0 || nshape >= MAXMAT) throw Error(OMX_ERR_UNSUPPORTED, "at most 15 shapes");
  if (nsub < 1 || nsub > 64) throw Error(OMX_ERR_INVALID_ARG, "nsub must be in 1..64");
  if (!eps_background || !epsinv_out || (nshape > 0 && (!nverts || !verts || !eps_shapes)))
    throw Error(OMX_ERR_INVALID_ARG, "NULL buffer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    throw Error(OMX_ERR_NO_DEVICE, "no such HIP device");
  std::vector<int> v0(nshape + 1, 0), nv(nshape + 1, 0);
  int tot = 0;
  for (int s = 0; s < nshape; ++s) {
    if (nverts[s] < 3) throw Error(OMX_ERR_INVALID_ARG, "a shape needs at least 3 vertices");
    v0[s] = tot;
    nv[s] = nverts[s];
    tot += nverts[s];
  }
  std::vector<double> mats(size_t(9) * (nshape + 1));
  for (int k = 0; k < 9; ++k) mats[k] = eps_background[k];
  for (int s = 0; s < nshape; ++s)
    for (int k = 0; k < 9; ++k) mats[9 * (s + 1) + k] = eps_shapes[9 * s + k];
  for (double v : mats)
    if (!std::isfinite(v)) throw Error(OMX_ERR_INVALID_ARG, "material tensors must be finite");
  OMX_HIP_CHECK(hipSetDevice(device));
  DevMem dv0(sizeof(int) * (nshape + 1)), dnv(sizeof(int) * (nshape + 1)), dxy(sizeof(double) * 2 * (tot + 1)),
      dm(sizeof(double) * mats.size()), dout(sizeof(double) * 9 * size_t(nx) * ny);
  OMX_HIP_CHECK(hipMemcpy(dv0.as<void>(), v0.data(), sizeof(int) * v0.size(), hipMemcpyHostToDevice));
  OMX_HIP_CHECK(hipMemcpy(dnv.as<void>(), nv.data(), sizeof(int) * nv.size(), hipMemcpyHostToDevice));
  if (tot > 0) OMX_HIP_CHECK(hipMemcpy(dxy.as<void>(), verts, sizeof(double) * 2 * tot, hipMemcpyHostToDevice));
  OMX_HIP_CHECK(hipMemcpy(dm.as<void>(), mats.data(), sizeof(double) * mats.size(), hipMemcpyHostToDevice));
  Geometry G{nshape, dv0.as<int>(), dnv.as<int>(), dxy.as<double>(), dm.as<double>()};
  const long np = long(nx) * ny;
  hipLaunchKernelGGL(k_smooth, dim3(unsigned((np + SB - 1) / SB)), dim3(SB), 0, 0, nx, ny, Dx, Dy, nsub, G,
                     dout.as<double>());
  OMX_HIP_CHECK(hipGetLastError());
  OMX_HIP_CHECK(hipMemcpy(epsinv_out, dout.as<void>(), sizeof(double) * 9 * np, hipMemcpyDeviceToHost));
}

}  // namespace omx
