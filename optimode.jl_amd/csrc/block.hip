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
// Gram on the matrix cores: partial[a][b] = sum over a row chunk of conj(A_a) B_b, complex,
// FP32 or FP64, with v_mfma_{f32,f64}_16x16x4.  Lane layout (checked on gfx950): A operand
// lane l -> (i = l%16, k = l/16), B operand lane l -> (k = l/16, j = l%16) for both types;
// accumulator register r of lane l -> (row 4(l/16) + r, col l%16) for f32 but
// (row l/16 + 4r, col l%16) for f64.
//
// The four k slots of one MFMA are four rows of the chunk: lane (i, g) holds S consecutive
// complex rows [rb, rb + S) of column i of its A tile and of column i of its B tile, with
// rb = stage base + g S, loaded straight from HBM with 16-byte loads (no LDS staging).
// Step s of the stage multiplies row rb + s:
//   Re G += ar br + ai bi,   Im G += ar bi - ai br       (four real MFMAs per step).
// Every 16 x 16 output tile of an instance that holds at least one column pair is computed
// (runtime tile count <= TA x TB); the next stage's rows are loaded while the current one
// is multiplied.  The 4 waves of a block take interleaved stages of the block's row chunk
// and their tiles are summed through LDS before one FP64 write per chunk.
// ---------------------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
template <typename T> struct Acc4;
template <> struct Acc4<float> { typedef floatx4 type; };
template <> struct Acc4<double> { typedef doublex4 type; };

__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma16(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// S consecutive complex rows of one column: 16-byte global loads when the column exists
// and all S rows lie below `lim` (every stage of a chunk but a final partial one), else
// element-wise with zeros past `lim` or for a missing column.
template <typename T> using gcptr = const __attribute__((address_space(1))) T*;
typedef float gf4 __attribute__((ext_vector_type(4)));
typedef double gd2 __attribute__((ext_vector_type(2)));

template <int S>
__device__ __forceinline__ void load_rows(const cplx<float>* col, long row, long lim, cplx<float> (&v)[S]) {
  if (col != nullptr && row + S <= lim) {
    gcptr<gf4> q = (gcptr<gf4>)(col + row);
#pragma unroll
    for (int j = 0; j < S / 2; ++j) {
      const gf4 w = q[j];
      v[2 * j] = cplx<float>{w.x, w.y};
      v[2 * j + 1] = cplx<float>{w.z, w.w};
    }
  } else {
#pragma unroll
    for (int j = 0; j < S; ++j) v[j] = (col != nullptr && row + j < lim) ? col[row + j] : cplx<float>{0.f, 0.f};
  }
}
template <int S>
__device__ __forceinline__ void load_rows(const cplx<double>* col, long row, long lim, cplx<double> (&v)[S]) {
  if (col != nullptr && row + S <= lim) {
    gcptr<gd2> q = (gcptr<gd2>)(col + row);
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const gd2 w = q[j];
      v[j] = cplx<double>{w.x, w.y};
    }
  } else {
#pragma unroll
    for (int j = 0; j < S; ++j) v[j] = (col != nullptr && row + j < lim) ? col[row + j] : cplx<double>{0.0, 0.0};
  }
}

// The body of k_gram16 for exactly NA x NB active output tiles (compile-time, so the MFMA
// sequence has no predicates); S rows per lane and stage.  Returns the accumulators.
template <typename T, int NA, int NB, int S, bool BPRE, bool HERM, typename Acc>
__device__ __forceinline__ void gram_tiles(const cplx<T>* const (&ap)[3], const cplx<T>* const (&bp)[3], bool bown,
                                           long rs, long r1, long stride, int g, Acc (&accr)[3][3],
                                           Acc (&acci)[3][3]) {
  cplx<T> a[NA][S], b[NB][S];
  auto load = [&](long r, cplx<T>(&xa)[NA][S], cplx<T>(&xb)[NB][S]) {
#pragma unroll
    for (int t = 0; t < NA; ++t) load_rows<S>(ap[t], r + g * S, r1, xa[t]);
    if constexpr (!BPRE) {
#pragma unroll
      for (int t = 0; t < NB; ++t) load_rows<S>(bp[t], r + g * S, r1, xb[t]);
    }
  };
  if (rs < r1) load(rs, a, b);
  while (rs < r1) {
    const long rn = rs + stride;
    cplx<T> na[NA][S], nb[NB][S];
    if (rn < r1) load(rn, na, nb);
    if constexpr (BPRE) {
#pragma unroll
      for (int s = 0; s < S; ++s) b[0][s] = bown ? a[0][s] : cplx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int ta = 0; ta < NA; ++ta)
#pragma unroll
        for (int tb = 0; tb < NB; ++tb) {
          if (HERM && ta > tb) continue;
          const cplx<T> x = a[ta][s], y = b[tb][s];
          accr[ta][tb] = mfma16(x.x, y.x, accr[ta][tb]);
          acci[ta][tb] = mfma16(x.x, y.y, acci[ta][tb]);
          accr[ta][tb] = mfma16(x.y, y.y, accr[ta][tb]);
          acci[ta][tb] = mfma16(x.y, -y.x, acci[ta][tb]);
        }
#pragma unroll
    for (int t = 0; t < NA; ++t)
#pragma unroll
      for (int s = 0; s < S; ++s) a[t][s] = na[t][s];
    if constexpr (!BPRE) {
#pragma unroll
      for (int t = 0; t < NB; ++t)
#pragma unroll
        for (int s = 0; s < S; ++s) b[t][s] = nb[t][s];
    }
    rs = rn;
  }
}

// BPRE: B is the leading npb (<= 16) columns of A, so the B tile is taken from A's first
// tile instead of being loaded again.  HERM: A^H B is Hermitian (B = M A for Hermitian M):
// only the tiles on and above the tile diagonal are computed and written.
template <typename T, int TA, int TB, bool BPRE, bool HERM>
__global__ __launch_bounds__(BLK) void k_gram16(ColList<T> A, ColList<T> B, long K, long rows_per_chunk, c64* partial,
                                                int pm) {
  static_assert(!BPRE || TB == 1, "a B prefix of A fits one tile");
  static_assert(!HERM || TA == TB, "a Hermitian Gram is square");
  using acc_t = typename Acc4<T>::type;
  // rows per lane and stage: 64 bytes of each tile column (FP32) or 32 (FP64), halved for
  // wide tile sets, to bound the registers of the prefetched stage
  constexpr int S = (sizeof(T) == 4 ? (TA + TB > 4 ? 4 : 8) : (TA + TB >= 4 ? 2 : 4));
  constexpr int WR = 4 * S;       // rows per wave and stage
  constexpr int STRIDE = 4 * WR;  // rows per block and stage
  __shared__ const cplx<T>* pa_ptr[16 * TA];
  __shared__ const cplx<T>* pb_ptr[16 * TB];
  __shared__ int npa_s, npb_s;
  __shared__ T red[4][8][64];
  const int inst = blockIdx.y, chunk = blockIdx.x;
  if (threadIdx.x == 0) {
    npa_s = resolve_cols(A, inst, pa_ptr, 16 * TA);
    npb_s = resolve_cols(B, inst, pb_ptr, 16 * TB);
  }
  __syncthreads();
  const int npa = npa_s, npb = npb_s;
  if (npa == 0 || npb == 0) return;
  const int nta = (npa + 15) >> 4, ntb = (npb + 15) >> 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, g = lane >> 4;
  const cplx<T>* ap[3] = {nullptr, nullptr, nullptr};
  const cplx<T>* bp[3] = {nullptr, nullptr, nullptr};
#pragma unroll
  for (int t = 0; t < TA; ++t) ap[t] = (t * 16 + i16 < npa) ? pa_ptr[t * 16 + i16] : nullptr;
#pragma unroll
  for (int t = 0; t < TB; ++t) bp[t] = (t * 16 + i16 < npb) ? pb_ptr[t * 16 + i16] : nullptr;
  acc_t accr[3][3], acci[3][3];
#pragma unroll
  for (int ta = 0; ta < 3; ++ta)
#pragma unroll
    for (int tb = 0; tb < 3; ++tb) {
      accr[ta][tb] = acc_t{0, 0, 0, 0};
      acci[ta][tb] = acc_t{0, 0, 0, 0};
    }
  const long r0 = long(chunk) * rows_per_chunk;
  const long r1 = (r0 + rows_per_chunk < K) ? r0 + rows_per_chunk : K;
  const long rs = r0 + long(wave) * WR;
  const bool bown = i16 < npb;
  // one fully unrolled body per active tile shape (block-uniform)
  switch (nta * 4 + ntb) {
#define OMX_GRAM_CASE(na, nb)                                                                                   \
  case na * 4 + nb:                                                                                             \
    if constexpr (na <= TA && nb <= TB && (!BPRE || nb == 1) && (!HERM || na == nb))                            \
      gram_tiles<T, na, nb, S, BPRE, HERM>(ap, bp, bown, rs, r1, STRIDE, g, accr, acci);                        \
    break;
    OMX_GRAM_CASE(1, 1)
    OMX_GRAM_CASE(1, 2)
    OMX_GRAM_CASE(1, 3)
    OMX_GRAM_CASE(2, 1)
    OMX_GRAM_CASE(2, 2)
    OMX_GRAM_CASE(2, 3)
    OMX_GRAM_CASE(3, 1)
    OMX_GRAM_CASE(3, 2)
    OMX_GRAM_CASE(3, 3)
#undef OMX_GRAM_CASE
    default: break;
  }
  // sum the four waves tile by tile; wave w finishes accumulator register w
  c64* out = partial + (long(inst) * gridDim.x + chunk) * pm * pm;
#pragma unroll
  for (int ta = 0; ta < TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) {
      if (ta >= nta || tb >= ntb || (HERM && ta > tb)) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        red[wave][r][lane] = accr[ta][tb][r];
        red[wave][4 + r][lane] = acci[ta][tb][r];
      }
      __syncthreads();
      const int r = wave;
      const double sr = (double(red[0][r][lane]) + double(red[1][r][lane])) +
                        (double(red[2][r][lane]) + double(red[3][r][lane]));
      const double si = (double(red[0][4 + r][lane]) + double(red[1][4 + r][lane])) +
                        (double(red[2][4 + r][lane]) + double(red[3][4 + r][lane]));
      const int row = ta * 16 + (sizeof(T) == 4 ? 4 * g + r : g + 4 * r);
      out[row * pm + tb * 16 + i16] = c64{sr, si};
      __syncthreads();
    }
}

// out_i[e] = sum over chunks of partial_i[chunk][e]: 16 elements x 16 chunk groups per
// block (16 consecutive elements are 256 contiguous bytes of one chunk), the groups
// combined in a fixed order (deterministic).  Instances with cnt[i] == 0 are skipped.
__global__ __launch_bounds__(BLK) void k_reduce(const c64* partial, int nchunk, int esz, c64* out, const int* cnt) {
  __shared__ double sx[16][17], sy[16][17];
  const int inst = blockIdx.y;
  if (cnt != nullptr && cnt[inst] == 0) return;
  const int el = threadIdx.x & 15, grp = threadIdx.x >> 4;
  const int e = blockIdx.x * 16 + el;
  double ax = 0.0, ay = 0.0;
  if (e < esz) {
    const c64* p = partial + long(inst) * nchunk * esz + e;
    // unrolled: the loads of eight chunks are in flight together (the sums keep their order)
#pragma unroll 8
    for (int c = grp; c < nchunk; c += 16) {
      const c64 v = p[long(c) * esz];
      ax += v.x;
      ay += v.y;
    }
  }
  sx[grp][el] = ax;
  sy[grp][el] = ay;
  __syncthreads();
  if (grp == 0 && e < esz) {
    double tx = 0.0, ty = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      tx += sx[q][el];
      ty += sy[q][el];
    }
    out[long(inst) * esz + e] = c64{tx, ty};
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
    // four input columns per step: their loads are issued together before the products
    int a = 0;
    for (; a + 4 <= np; a += 4) {
      cplx<T> y[4], y2[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        y[u] = py[a + u][row];
        y2[u] = cplx<T>{T(0), T(0)};
        if (MODE == CMB_TWO || MODE == CMB_RESID) y2[u] = py2[a + u][row];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int q = 0; q < QM; ++q) {
          const cplx<T> c = Cs[a + u][q];
          a1[q] = a1[q] + y[u] * c;
          if (MODE == CMB_TWO || MODE == CMB_RESID) a2[q] = a2[q] + y2[u] * c;
        }
    }
    for (; a < np; ++a) {
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
    // per column: a wavefront reduction (fixed butterfly order), then the BLK/64 wavefront
    // sums in order through LDS
    constexpr int NW = BLK / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < QM; ++q) {
      double v = nrm[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) red[wave * QM + q] = v;
    }
    __syncthreads();
    if (threadIdx.x < QM) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) t += red[w * QM + threadIdx.x];
      norm_partial[(long(inst) * gridDim.x + blockIdx.x) * QM + threadIdx.x] = t;
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
__global__ __launch_bounds__(BLK) void k_copy_cols(const cplx<T>* src, cplx<T>* dst, const cplx<T>* src2,
                                                   cplx<T>* dst2, long ld, long istride, const int* off,
                                                   const int* cnt, long K) {
  // blockIdx.z selects the pair: (src, dst) or (src2, dst2)
  if (blockIdx.z == 1) {
    src = src2;
    dst = dst2;
  }
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

}  // namespace

template <typename T, int TA>
static void gram_tb(int tb, bool bpre, bool herm, dim3 grid, hipStream_t stream, const ColList<T>& A,
                    const ColList<T>& B, long K, long rpc, c64* partial, int pm) {
  if (bpre) {
    hipLaunchKernelGGL((k_gram16<T, TA, 1, true, false>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial, pm);
    return;
  }
  if (herm) {
    hipLaunchKernelGGL((k_gram16<T, TA, TA, false, true>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial, pm);
    return;
  }
  switch (tb) {
    case 1:
      hipLaunchKernelGGL((k_gram16<T, TA, 1, false, false>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial, pm);
      break;
    case 2:
      hipLaunchKernelGGL((k_gram16<T, TA, 2, false, false>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial, pm);
      break;
    default:
      hipLaunchKernelGGL((k_gram16<T, TA, 3, false, false>), grid, dim3(BLK), 0, stream, A, B, K, rpc, partial, pm);
      break;
  }
}

template <typename T>
void launch_gram(const ColList<T>& A, const ColList<T>& B, int ninst, long K, int nchunk, int pm, c64* partial,
                 hipStream_t stream, int ca, int cb, bool b_prefix_of_a, bool hermitian) {
  if (ca <= 0 || cb <= 0 || ca > 48 || cb > 48 || ca > pm || cb > pm || pm % 16 != 0 || (b_prefix_of_a && cb > 16) ||
      (hermitian && (b_prefix_of_a || (ca + 15) / 16 != (cb + 15) / 16)))
    throw Error(OMX_ERR_INTERNAL, "gram: column bounds out of range");
  // chunks are whole block stages (128 rows), so only the last one can end mid-stage
  long rpc = (K + nchunk - 1) / nchunk;
  rpc = (rpc + 127) / 128 * 128;
  dim3 grid(nchunk, ninst);
  const int ta = (ca + 15) / 16, tb = (cb + 15) / 16;
  switch (ta) {
    case 1: gram_tb<T, 1>(tb, b_prefix_of_a, hermitian, grid, stream, A, B, K, rpc, partial, pm); break;
    case 2: gram_tb<T, 2>(tb, b_prefix_of_a, hermitian, grid, stream, A, B, K, rpc, partial, pm); break;
    default: gram_tb<T, 3>(tb, b_prefix_of_a, hermitian, grid, stream, A, B, K, rpc, partial, pm); break;
  }
  OMX_HIP_CHECK(hipGetLastError());
}

void launch_reduce(const c64* partial, int ninst, int nchunk, int esz, c64* out, hipStream_t stream, const int* cnt) {
  dim3 grid((esz + 15) / 16, ninst);
  hipLaunchKernelGGL(k_reduce, grid, dim3(BLK), 0, stream, partial, nchunk, esz, out, cnt);
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
                      const int* cnt, long K, hipStream_t stream, const cplx<T>* src2, cplx<T>* dst2) {
  dim3 grid(128, ninst, src2 ? 2 : 1);
  hipLaunchKernelGGL((k_copy_cols<T>), grid, dim3(BLK), 0, stream, src, dst, src2, dst2, ld, istride, off, cnt, K);
  OMX_HIP_CHECK(hipGetLastError());
}

template void launch_copy_cols<float>(const cplx<float>*, cplx<float>*, long, long, int, const int*, const int*, long,
                                      hipStream_t, const cplx<float>*, cplx<float>*);
template void launch_copy_cols<double>(const cplx<double>*, cplx<double>*, long, long, int, const int*, const int*, long,
                                       hipStream_t, const cplx<double>*, cplx<double>*);
template void launch_gram<float>(const ColList<float>&, const ColList<float>&, int, long, int, int, c64*, hipStream_t, int, int,
                                 bool, bool);
template void launch_gram<double>(const ColList<double>&, const ColList<double>&, int, long, int, int, c64*, hipStream_t,
                                  int, int, bool, bool);
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
