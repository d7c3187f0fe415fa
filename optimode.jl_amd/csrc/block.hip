// Tall-skinny block kernels (see block.hpp).
#include "block.hpp"
#include "omx_errors.hpp"

namespace omx {

namespace {

constexpr int BLK = 256;

__device__ __forceinline__ int sref_off(const ShapeRef& s, int i) { return s.off ? s.off[i] : s.fixed_off; }
__device__ __forceinline__ int sref_cnt(const ShapeRef& s, int i) { return s.cnt ? s.cnt[i] : s.fixed_cnt; }

// Resolve the column pointers of list L for instance i into ptrs[0..count); returns count.
template <typename T>
__device__ int resolve_cols(const ColList<T>& L, int i, const cplx<T>** ptrs, int maxc) {
  int n = 0;
  for (int p = 0; p < L.n; ++p) {
    const int c = sref_cnt(L.p[p].s, i);
    const int o = sref_off(L.p[p].s, i);
    const cplx<T>* b = L.p[p].base + i * L.p[p].istride;
    for (int q = 0; q < c && n < maxc; ++q) ptrs[n++] = b + long(o + q) * L.p[p].ld;
  }
  return n;
}

// ---------------------------------------------------------------------------------------
// Gram: partial[a][b] = sum_rows conj(A_a) B_b.  Thread (ty, tx) of a 16x16 layout owns a
// TILE x TILE patch of the PM x PM output; rows are staged through LDS SR at a time.
// ---------------------------------------------------------------------------------------
template <typename T, int PM>
__global__ __launch_bounds__(BLK) void k_gram(ColList<T> A, ColList<T> B, long K, long rows_per_chunk,
                                              c64* partial) {
  constexpr int TILE = PM / 16;
  constexpr int SR = 32;
  __shared__ const cplx<T>* pa_ptr[PM];
  __shared__ const cplx<T>* pb_ptr[PM];
  __shared__ int npa_s, npb_s;
  __shared__ cplx<T> As[SR][PM + 1];
  __shared__ cplx<T> Bs[SR][PM + 1];
  const int inst = blockIdx.y;
  const int chunk = blockIdx.x;
  if (threadIdx.x == 0) {
    npa_s = resolve_cols(A, inst, pa_ptr, PM);
    npb_s = resolve_cols(B, inst, pb_ptr, PM);
  }
  __syncthreads();
  const int npa = npa_s, npb = npb_s;
  if (npa == 0 || npb == 0) return;
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  const bool busy = ty * TILE < npa && tx * TILE < npb;
  cplx<T> acc[TILE][TILE];
#pragma unroll
  for (int i = 0; i < TILE; ++i)
#pragma unroll
    for (int j = 0; j < TILE; ++j) acc[i][j] = cplx<T>{T(0), T(0)};
  const long r0 = long(chunk) * rows_per_chunk;
  const long r1 = (r0 + rows_per_chunk < K) ? r0 + rows_per_chunk : K;
  for (long rs = r0; rs < r1; rs += SR) {
    for (int idx = threadIdx.x; idx < SR * PM; idx += BLK) {
      const int a = idx / SR, r = idx - (idx / SR) * SR;
      const long row = rs + r;
      const bool ok = row < r1;
      As[r][a] = (a < npa && ok) ? pa_ptr[a][row] : cplx<T>{T(0), T(0)};
      Bs[r][a] = (a < npb && ok) ? pb_ptr[a][row] : cplx<T>{T(0), T(0)};
    }
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < (busy ? SR : 0); ++r) {
      cplx<T> av[TILE], bv[TILE];
#pragma unroll
      for (int i = 0; i < TILE; ++i) av[i] = As[r][ty * TILE + i];
#pragma unroll
      for (int j = 0; j < TILE; ++j) bv[j] = Bs[r][tx * TILE + j];
#pragma unroll
      for (int i = 0; i < TILE; ++i)
#pragma unroll
        for (int j = 0; j < TILE; ++j) {
          // conj(a) * b
          acc[i][j].x += av[i].x * bv[j].x + av[i].y * bv[j].y;
          acc[i][j].y += av[i].x * bv[j].y - av[i].y * bv[j].x;
        }
    }
    __syncthreads();
  }
  c64* out = partial + (long(inst) * gridDim.x + chunk) * PM * PM;
#pragma unroll
  for (int i = 0; i < TILE; ++i)
#pragma unroll
    for (int j = 0; j < TILE; ++j) {
      const int a = ty * TILE + i, b = tx * TILE + j;
      out[a * PM + b] = c64{double(acc[i][j].x), double(acc[i][j].y)};
    }
}

__global__ __launch_bounds__(BLK) void k_reduce(const c64* partial, int nchunk, int esz, c64* out) {
  const int inst = blockIdx.y;
  for (int e = blockIdx.x * BLK + threadIdx.x; e < esz; e += gridDim.x * BLK) {
    double sx = 0.0, sy = 0.0;
    const c64* p = partial + long(inst) * nchunk * esz + e;
    for (int c = 0; c < nchunk; ++c) {
      sx += p[long(c) * esz].x;
      sy += p[long(c) * esz].y;
    }
    out[long(inst) * esz + e] = c64{sx, sy};
  }
}

// ---------------------------------------------------------------------------------------
// Combine: row-parallel Y C products (see CombineMode).
// ---------------------------------------------------------------------------------------
constexpr int CMB_MAXP = 48;

template <typename T, int QM, int MODE>
__global__ __launch_bounds__(BLK) void k_combine(ColList<T> Y, ColList<T> Y2, CombineCoef cf, OutView<T> O1,
                                                 OutView<T> O2, OutView<T> O3, long K, double* norm_partial) {
  __shared__ const cplx<T>* py[CMB_MAXP];
  __shared__ const cplx<T>* py2[CMB_MAXP];
  __shared__ cplx<T> Cs[CMB_MAXP][QM];
  __shared__ T lam_s[QM];
  __shared__ int np_s, nq_s;
  __shared__ double red[BLK];
  const int inst = blockIdx.y;
  if (threadIdx.x == 0) {
    np_s = resolve_cols(Y, inst, py, CMB_MAXP);
    if (MODE == CMB_TWO || MODE == CMB_RESID) resolve_cols(Y2, inst, py2, CMB_MAXP);
    nq_s = cf.qcnt ? cf.qcnt[inst] : cf.qfixed;
  }
  __syncthreads();
  const int np = np_s, nq = nq_s;
  if (nq <= 0) return;
  for (int idx = threadIdx.x; idx < CMB_MAXP * QM; idx += BLK) {
    const int a = idx / QM, q = idx - (idx / QM) * QM;
    cplx<T> v{T(0), T(0)};
    if (a < np && q < nq) {
      const double* c = cf.C + 2 * (inst * cf.cstride + long(a) * cf.ldc + q);
      v = cplx<T>{T(c[0]), T(c[1])};
    }
    Cs[a][q] = v;
  }
  if (MODE == CMB_RESID && threadIdx.x < QM) lam_s[threadIdx.x] = threadIdx.x < nq ? T(cf.lam[inst * cf.lstride + threadIdx.x]) : T(0);
  __syncthreads();
  const int o1 = O1.off ? O1.off[inst] : O1.fixed_off;
  cplx<T>* b1 = O1.base + inst * O1.istride + long(o1) * O1.ld;
  cplx<T>* b2 = nullptr;
  cplx<T>* b3 = nullptr;
  if (MODE == CMB_TWO || MODE == CMB_RESID) {
    const int o2 = O2.off ? O2.off[inst] : O2.fixed_off;
    b2 = O2.base + inst * O2.istride + long(o2) * O2.ld;
  }
  if (MODE == CMB_RESID) {
    const int o3 = O3.off ? O3.off[inst] : O3.fixed_off;
    b3 = O3.base + inst * O3.istride + long(o3) * O3.ld;
  }
  double nrm[QM];
#pragma unroll
  for (int q = 0; q < QM; ++q) nrm[q] = 0.0;
  for (long row = long(blockIdx.x) * BLK + threadIdx.x; row < K; row += long(gridDim.x) * BLK) {
    cplx<T> a1[QM], a2[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      a1[q] = cplx<T>{T(0), T(0)};
      a2[q] = cplx<T>{T(0), T(0)};
    }
    for (int a = 0; a < np; ++a) {
      const cplx<T> y = py[a][row];
      cplx<T> y2{T(0), T(0)};
      if (MODE == CMB_TWO || MODE == CMB_RESID) y2 = py2[a][row];
#pragma unroll
      for (int q = 0; q < QM; ++q) {
        const cplx<T> c = Cs[a][q];
        a1[q] = a1[q] + y * c;
        if (MODE == CMB_TWO || MODE == CMB_RESID) a2[q] = a2[q] + y2 * c;
      }
    }
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      if (q < nq) {
        const long idx = long(q) * O1.ld + row;
        if (MODE == CMB_SUBTRACT) {
          b1[idx] = b1[idx] - a1[q];
        } else {
          b1[idx] = a1[q];
        }
        if (MODE == CMB_TWO || MODE == CMB_RESID) b2[long(q) * O2.ld + row] = a2[q];
        if (MODE == CMB_RESID) {
          const cplx<T> r = a2[q] - a1[q] * lam_s[q];
          b3[long(q) * O3.ld + row] = r;
          nrm[q] += double(r.x) * double(r.x) + double(r.y) * double(r.y);
        }
      }
    }
  }
  if (MODE == CMB_RESID) {
    for (int q = 0; q < QM; ++q) {
      red[threadIdx.x] = nrm[q];
      __syncthreads();
      for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) norm_partial[(long(inst) * gridDim.x + blockIdx.x) * QM + q] = red[0];
      __syncthreads();
    }
  }
}

// splitmix64-based counter RNG -> standard normal via Box-Muller
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_randn(cplx<T>* base, long ld, long istride, int ncol, long K, uint64_t seed) {
  const int inst = blockIdx.y;
  const long total = long(ncol) * K;
  for (long e = long(blockIdx.x) * BLK + threadIdx.x; e < total; e += long(gridDim.x) * BLK) {
    const int c = int(e / K);
    const long row = e - long(c) * K;
    const uint64_t key = mix64(seed ^ mix64(uint64_t(inst) * 0x100000001b3ull + uint64_t(c))) + uint64_t(row) * 2;
    const uint64_t h1 = mix64(key), h2 = mix64(key + 1);
    const double u1 = (double(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = double(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double rad = sqrt(-2.0 * log(u1)) * 0.70710678118654752;
    double s, co;
    sincospi(2.0 * u2, &s, &co);
    base[inst * istride + long(c) * ld + row] = cplx<T>{T(rad * co), T(rad * s)};
  }
}

template <typename TS, typename TD>
__global__ __launch_bounds__(BLK) void k_convert(const cplx<TS>* src, long lds, long istrs, cplx<TD>* dst, long ldd,
                                                 long istrd, int ncol, long K) {
  const int inst = blockIdx.y;
  const long total = long(ncol) * K;
  for (long e = long(blockIdx.x) * BLK + threadIdx.x; e < total; e += long(gridDim.x) * BLK) {
    const int c = int(e / K);
    const long row = e - long(c) * K;
    const cplx<TS> v = src[inst * istrs + long(c) * lds + row];
    dst[inst * istrd + long(c) * ldd + row] = cplx<TD>{TD(v.x), TD(v.y)};
  }
}

// Copy columns [off_i, off_i + cnt_i) of instance i from src to dst (same positions).
template <typename T>
__global__ __launch_bounds__(BLK) void k_copy_cols(const cplx<T>* src, cplx<T>* dst, long ld, long istride,
                                                   const int* off, const int* cnt, long K) {
  const int inst = blockIdx.y;
  const int c0 = off ? off[inst] : 0;
  const int nc = cnt[inst];
  const long total = long(nc) * K;
  for (long e = long(blockIdx.x) * BLK + threadIdx.x; e < total; e += long(gridDim.x) * BLK) {
    const int c = int(e / K);
    const long row = e - long(c) * K;
    const long idx = inst * istride + long(c0 + c) * ld + row;
    dst[idx] = src[idx];
  }
}

// ---------------------------------------------------------------------------------------
// Gram on the matrix cores (FP32): G = A^H B with A, B complex.  Over the interleaved
// (re, im) rows of one complex row r, v_mfma_f32_32x32x2f32 with k = (re, im) gives
//   Re G += [ar ai] . [br bi]^T      (B operand = (br, bi))
//   Im G += [ar ai] . [bi -br]^T     (B operand = (bi, -br))
// so each complex row costs two MFMAs per 32x32 output tile.  Lane layout (checked on
// gfx950): A operand lane l -> (row l%32, k l/32); B operand lane l -> (k l/32, col l%32);
// accumulator register q of lane l -> (row 4(l/32) + q%4 + 8(q/4), col l%32).
// The 4 waves of a block split the (TA x TB) output tiles and, when fewer than 4 tiles,
// the staged rows; partial tiles are summed through LDS and written once per block.
// ---------------------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));

template <int TA, int TB>
__global__ __launch_bounds__(BLK) void k_gram_mfma(ColList<float> A, ColList<float> B, long K, long rows_per_chunk,
                                                   c64* partial, int pm) {
  constexpr int NT = TA * TB;          // output tiles
  constexpr int G = 4 / NT;            // row groups per tile (1, 2 or 4)
  constexpr int SR = 64;               // complex rows staged per step
  constexpr int WA = TA * 32, WB = TB * 32;
  __shared__ const cplx<float>* pa_ptr[WA];
  __shared__ const cplx<float>* pb_ptr[WB];
  __shared__ int npa_s, npb_s;
  __shared__ cplx<float> As[SR][WA + 1];
  __shared__ cplx<float> Bs[SR][WB + 1];
  __shared__ float red[4][32][64];
  const int inst = blockIdx.y, chunk = blockIdx.x;
  if (threadIdx.x == 0) {
    npa_s = resolve_cols(A, inst, pa_ptr, WA);
    npb_s = resolve_cols(B, inst, pb_ptr, WB);
  }
  __syncthreads();
  const int npa = npa_s, npb = npb_s;
  if (npa == 0 || npb == 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = wave % NT, grp = wave / NT;
  const int ta = tile / TB, tb = tile % TB;
  const int hi = lane >> 5, col = lane & 31;
  floatx16 accr = {0}, acci = {0};
  const long r0 = long(chunk) * rows_per_chunk;
  const long r1 = (r0 + rows_per_chunk < K) ? r0 + rows_per_chunk : K;
  for (long rs = r0; rs < r1; rs += SR) {
    for (int idx = threadIdx.x; idx < SR * WA; idx += BLK) {
      const int a = idx / SR, r = idx - (idx / SR) * SR;
      const long row = rs + r;
      As[r][a] = (a < npa && row < r1) ? pa_ptr[a][row] : cplx<float>{0.f, 0.f};
    }
    for (int idx = threadIdx.x; idx < SR * WB; idx += BLK) {
      const int b = idx / SR, r = idx - (idx / SR) * SR;
      const long row = rs + r;
      Bs[r][b] = (b < npb && row < r1) ? pb_ptr[b][row] : cplx<float>{0.f, 0.f};
    }
    __syncthreads();
#pragma unroll 4
    for (int r = grp; r < SR; r += G) {
      const cplx<float> av = As[r][ta * 32 + col];
      const cplx<float> bv = Bs[r][tb * 32 + col];
      const float a = hi ? av.y : av.x;
      const float b = hi ? bv.y : bv.x;
      const float b2 = hi ? -bv.x : bv.y;
      accr = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, accr, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b2, acci, 0, 0, 0);
    }
    __syncthreads();
  }
  // sum the G row groups of each tile through LDS, then write FP64 partials
  for (int q = 0; q < 16; ++q) {
    red[wave][q][lane] = accr[q];
    red[wave][16 + q][lane] = acci[q];
  }
  __syncthreads();
  if (grp == 0) {
    c64* out = partial + (long(inst) * gridDim.x + chunk) * pm * pm;
    for (int q = 0; q < 16; ++q) {
      float sr = 0.f, si = 0.f;
      for (int g = 0; g < G; ++g) {
        sr += red[tile + g * NT][q][lane];
        si += red[tile + g * NT][16 + q][lane];
      }
      const int a = ta * 32 + 4 * hi + (q & 3) + 8 * (q >> 2);
      const int b = tb * 32 + col;
      if (a < pm && b < pm) out[a * pm + b] = c64{double(sr), double(si)};
    }
  }
}

}  // namespace

template <typename T>
void launch_gram(const ColList<T>& A, const ColList<T>& B, int ninst, long K, int nchunk, int pm, c64* partial,
                 hipStream_t stream, int ca, int cb) {
  const long rpc = (K + nchunk - 1) / nchunk;
  dim3 grid(nchunk, ninst);
  if (sizeof(T) == 4 && ca > 0 && cb > 0 && ca <= 64 && cb <= 64) {
    // matrix-core path: ca, cb bound the column counts of A and B
    const ColList<float>& Af = reinterpret_cast<const ColList<float>&>(A);
    const ColList<float>& Bf = reinterpret_cast<const ColList<float>&>(B);
    const int ta = ca > 32 ? 2 : 1, tb = cb > 32 ? 2 : 1;
    if (ta == 1 && tb == 1)
      hipLaunchKernelGGL((k_gram_mfma<1, 1>), grid, dim3(BLK), 0, stream, Af, Bf, K, rpc, partial, pm);
    else if (ta == 2 && tb == 1)
      hipLaunchKernelGGL((k_gram_mfma<2, 1>), grid, dim3(BLK), 0, stream, Af, Bf, K, rpc, partial, pm);
    else if (ta == 1 && tb == 2)
      hipLaunchKernelGGL((k_gram_mfma<1, 2>), grid, dim3(BLK), 0, stream, Af, Bf, K, rpc, partial, pm);
    else
      hipLaunchKernelGGL((k_gram_mfma<2, 2>), grid, dim3(BLK), 0, stream, Af, Bf, K, rpc, partial, pm);
    OMX_HIP_CHECK(hipGetLastError());
    return;
  }
  switch (pm) {
    case 16: hipLaunchKernelGGL((k_gram<T, 16>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial); break;
    case 32: hipLaunchKernelGGL((k_gram<T, 32>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial); break;
    case 48: hipLaunchKernelGGL((k_gram<T, 48>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial); break;
    default: throw Error(OMX_ERR_INTERNAL, "gram: unsupported tile size");
  }
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_reduce(const c64* partial, int ninst, int nchunk, int esz, c64* out, hipStream_t stream) {
  dim3 grid((esz + BLK - 1) / BLK, ninst);
  hipLaunchKernelGGL(k_reduce, grid, dim3(BLK), 0, stream, partial, nchunk, esz, out);
  OMX_HIP_CHECK(hipGetLastError());
}

template <typename T, int QM>
static void combine_q(CombineMode mode, const ColList<T>& Y, const ColList<T>& Y2, const CombineCoef& cf,
                      OutView<T> O1, OutView<T> O2, OutView<T> O3, int ninst, long K, int nblk, double* np,
                      hipStream_t s) {
  dim3 grid(nblk, ninst);
  switch (mode) {
    case CMB_PLAIN: hipLaunchKernelGGL((k_combine<T, QM, CMB_PLAIN>), grid, dim3(BLK), 0, s, Y, Y2, cf, O1, O2, O3, K, np); break;
    case CMB_TWO: hipLaunchKernelGGL((k_combine<T, QM, CMB_TWO>), grid, dim3(BLK), 0, s, Y, Y2, cf, O1, O2, O3, K, np); break;
    case CMB_RESID: hipLaunchKernelGGL((k_combine<T, QM, CMB_RESID>), grid, dim3(BLK), 0, s, Y, Y2, cf, O1, O2, O3, K, np); break;
    case CMB_SUBTRACT: hipLaunchKernelGGL((k_combine<T, QM, CMB_SUBTRACT>), grid, dim3(BLK), 0, s, Y, Y2, cf, O1, O2, O3, K, np); break;
  }
}

template <typename T>
void launch_combine(CombineMode mode, const ColList<T>& Y, const ColList<T>& Y2, const CombineCoef& cf, int qmax,
                    OutView<T> O1, OutView<T> O2, OutView<T> O3, int ninst, long K, int nblk, double* norm_partial,
                    hipStream_t stream) {
  switch (qmax) {
    case 4: combine_q<T, 4>(mode, Y, Y2, cf, O1, O2, O3, ninst, K, nblk, norm_partial, stream); break;
    case 8: combine_q<T, 8>(mode, Y, Y2, cf, O1, O2, O3, ninst, K, nblk, norm_partial, stream); break;
    case 16: combine_q<T, 16>(mode, Y, Y2, cf, O1, O2, O3, ninst, K, nblk, norm_partial, stream); break;
    default: throw Error(OMX_ERR_INTERNAL, "combine: unsupported column bound");
  }
  OMX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_randn(cplx<T>* base, long ld, long istride, int ninst, int ncol, long K, uint64_t seed,
                  hipStream_t stream) {
  dim3 grid(256, ninst);
  hipLaunchKernelGGL((k_randn<T>), grid, dim3(BLK), 0, stream, base, ld, istride, ncol, K, seed);
  OMX_HIP_CHECK(hipGetLastError());
}

template <typename TS, typename TD>
void launch_convert(const cplx<TS>* src, long lds, long istrs, cplx<TD>* dst, long ldd, long istrd, int ninst,
                    int ncol, long K, hipStream_t stream) {
  dim3 grid(256, ninst);
  hipLaunchKernelGGL((k_convert<TS, TD>), grid, dim3(BLK), 0, stream, src, lds, istrs, dst, ldd, istrd, ncol, K);
  OMX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_copy_cols(const cplx<T>* src, cplx<T>* dst, long ld, long istride, int ninst, const int* off,
                      const int* cnt, long K, hipStream_t stream) {
  dim3 grid(128, ninst);
  hipLaunchKernelGGL((k_copy_cols<T>), grid, dim3(BLK), 0, stream, src, dst, ld, istride, off, cnt, K);
  OMX_HIP_CHECK(hipGetLastError());
}

template void launch_copy_cols<float>(const cplx<float>*, cplx<float>*, long, long, int, const int*, const int*, long, hipStream_t);
template void launch_copy_cols<double>(const cplx<double>*, cplx<double>*, long, long, int, const int*, const int*, long, hipStream_t);
template void launch_gram<float>(const ColList<float>&, const ColList<float>&, int, long, int, int, c64*, hipStream_t, int, int);
template void launch_gram<double>(const ColList<double>&, const ColList<double>&, int, long, int, int, c64*, hipStream_t, int, int);
template void launch_combine<float>(CombineMode, const ColList<float>&, const ColList<float>&, const CombineCoef&, int,
                                    OutView<float>, OutView<float>, OutView<float>, int, long, int, double*, hipStream_t);
template void launch_combine<double>(CombineMode, const ColList<double>&, const ColList<double>&, const CombineCoef&, int,
                                     OutView<double>, OutView<double>, OutView<double>, int, long, int, double*,
                                     hipStream_t);
template void launch_randn<float>(cplx<float>*, long, long, int, int, long, uint64_t, hipStream_t);
template void launch_randn<double>(cplx<double>*, long, long, int, int, long, uint64_t, hipStream_t);
template void launch_convert<float, double>(const cplx<float>*, long, long, cplx<double>*, long, long, int, int, long, hipStream_t);
template void launch_convert<double, float>(const cplx<double>*, long, long, cplx<float>*, long, long, int, int, long, hipStream_t);
template void launch_convert<double, double>(const cplx<double>*, long, long, cplx<double>*, long, long, int, int, long, hipStream_t);
template void launch_convert<float, float>(const cplx<float>*, long, long, cplx<float>*, long, long, int, int, long, hipStream_t);

}  // namespace omx
