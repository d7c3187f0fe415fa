// Allocation-free operator application: omx_apply_helmholtz_noalloc.
//
// Same arguments and results as omx_apply_helmholtz (M(k) H for m vectors, maxwell.jl:471-481),
// for callers that must not allocate device memory.  Every device buffer of the call lives
// in `scratch`, a fixed file-scope device array of NOALLOC_SCRATCH_BYTES:
//
//   [0, 256)            k (FP64), the operator's per-instance k
//   A: 16 m K bytes     FP64 staging of H_in; in FP32, also the FP64 result before download
//   B: 16 m K bytes     FP64: the output block; FP32: the FP32 input (8 m K) and output (8 m K)
//   D: 3 N m cplx<T>    the operator's 3-component intermediate
//
// (K = 2N complex entries per vector, each region rounded up to 256 bytes).  In total
// 256 + 112 N m bytes in FP64 and 256 + 88 N m bytes in FP32.  Nothing here needs to
// start at zero: every region is written before it is read.
//
// `scratch` names a host-side stand-in in host code, not a device address: its device
// address is looked up once per device with hipGetSymbolAddress (noalloc_init, called
// when a handle is created on the device) and only that address is ever used.  All calls
// on one device share that device's array, so they are serialised by a per-device lock
// (overlapping calls would otherwise race on it); the array keeps the last call's contents.
#include <mutex>

#include "block.hpp"
#include "solver.hpp"

namespace omx {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_init_mu;
void* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
std::mutex g_call_mu[NOALLOC_MAX_DEVICES];

constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// Device address of this device's `scratch` (the caller has made `device` current).
void noalloc_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw Error(OMX_ERR_NO_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(OMX_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_dev[device] = p;
}

size_t noalloc_bytes(long N, int m, int prec) {
  const size_t K = size_t(2) * N;
  const size_t a = round256(sizeof(c64) * m * K);
  const size_t b = round256(sizeof(c64) * m * K);
  const size_t d = round256((prec == OMX_PREC_FP32 ? sizeof(c32) : sizeof(c64)) * 3 * size_t(N) * m);
  return 256 + a + b + d;
}

void apply_helmholtz_noalloc(Solver& s, double k, const double* Hin, double* Hout, int m, int prec) {
  if (m <= 0) throw Error(OMX_ERR_INVALID_ARG, "m must be positive");
  if (!Hin || !Hout) throw Error(OMX_ERR_INVALID_ARG, "NULL buffer");
  if (!std::isfinite(k)) throw Error(OMX_ERR_INVALID_ARG, "k must be finite");
  if (prec != OMX_PREC_FP32 && prec != OMX_PREC_FP64)
    throw Error(OMX_ERR_INVALID_ARG, "precision must be OMX_PREC_FP32 or OMX_PREC_FP64");
  const size_t need = noalloc_bytes(s.N, m, prec);
  if (need > NOALLOC_SCRATCH_BYTES)
    throw Error(OMX_ERR_BAD_SHAPE, "omx_apply_helmholtz_noalloc: " + std::to_string(m) + " vectors of " +
                                       std::to_string(s.N) + " pixels need " + std::to_string(need) +
                                       " bytes of scratch, more than its fixed " +
                                       std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes");
  OMX_HIP_CHECK(hipSetDevice(s.device));
  unsigned char* base = static_cast<unsigned char*>(g_scratch_dev[s.device]);
  if (!base) throw Error(OMX_ERR_INTERNAL, "the allocation-free scratch of this device was not set up");
  std::lock_guard<std::mutex> lk(g_call_mu[s.device]);

  const size_t K = size_t(s.K);
  const size_t nbytes64 = sizeof(c64) * m * K;
  double* kbuf = reinterpret_cast<double*>(base);
  c64* A = reinterpret_cast<c64*>(base + 256);
  unsigned char* Bp = base + 256 + round256(nbytes64);
  unsigned char* Dp = Bp + round256(nbytes64);
  OMX_HIP_CHECK(hipMemcpyAsync(kbuf, &k, sizeof(double), hipMemcpyHostToDevice, s.stream));
  OMX_HIP_CHECK(hipMemcpyAsync(A, Hin, nbytes64, hipMemcpyHostToDevice, s.stream));
  VecSel sel{1, m, nullptr, nullptr};
  if (prec == OMX_PREC_FP64) {
    c64* Y = reinterpret_cast<c64*>(Bp);
    BView<double> X{A, long(K), 0}, Yv{Y, long(K), 0};
    launch_operator<double>(OP_M, s.p64.plan, X, Yv, reinterpret_cast<c64*>(Dp), sel, kbuf, nullptr, 0, s.stream);
    OMX_HIP_CHECK(hipMemcpyAsync(Hout, Y, nbytes64, hipMemcpyDeviceToHost, s.stream));
  } else {
    c32* x32 = reinterpret_cast<c32*>(Bp);
    c32* y32 = x32 + m * K;
    launch_convert<double, float>(A, long(K), 0, x32, long(K), 0, 1, m, long(K), s.stream);
    BView<float> X{x32, long(K), 0}, Yv{y32, long(K), 0};
    launch_operator<float>(OP_M, s.p32.plan, X, Yv, reinterpret_cast<c32*>(Dp), sel, kbuf, nullptr, 0, s.stream);
    launch_convert<float, double>(y32, long(K), 0, A, long(K), 0, 1, m, long(K), s.stream);
    OMX_HIP_CHECK(hipMemcpyAsync(Hout, A, nbytes64, hipMemcpyDeviceToHost, s.stream));
  }
  OMX_HIP_CHECK(hipStreamSynchronize(s.stream));
}

}  // namespace omx
