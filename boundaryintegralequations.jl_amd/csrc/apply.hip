// Dense multi-right-hand-side operator apply Y = A X for complex128 column-major A
// (the assembled F or G, M x K, leading dimension lda) and X (K x nrhs), on FP64 MFMA.
//
// This is the step after assembly in the reference's workflows (F*p, G*vs:
// examples/3d_pulsating_sphere.jl:42,62; BLAS ZGEMM behind `*`), batched over 32
// right-hand sides as the reference's FMM operators batch them
// (src/3d/fast_multipole_method.jl:249-287).
//
// v_mfma_f64_16x16x4_f64 operand layout (one double per lane for A and B, four per lane
// for the 16 x 16 accumulator), as measured on gfx950 with lane-tagged operands:
//   A[i][k] in lane i + 16 k,  B[k][j] in lane j + 16 k,
//   D[i][j] in lane j + 16 (i % 4), element i / 4.
// A complex product uses four real MFMAs (Re += Ar Xr + (-Ai) Xi, Im += Ar Xi + Ai Xr).
// Each wave owns 32 rows x 32 columns (2 x 2 tiles); a workgroup of 4 waves owns 128
// rows; K is split over blockIdx.y (as many splits as fill one round of resident
// workgroups) and the partial sums are added into a zeroed Y with FP64 atomics.  X is
// first repacked k-major (K x ncp, ncp = nrhs rounded up to 32) so the B operand is a
// coalesced 16-byte load per lane.  Each wave keeps kRing k steps of operands in flight
// (bounds-checked buffer loads, so the compiler counts them instead of draining them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "bie_common.hpp"
#include "device_common.hpp"

namespace bie {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kRing = 4;  // k steps (of 4) of operands in flight per wave

__device__ __forceinline__ v4d mfma4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ void k_pack_x(const double2* __restrict__ X, int64_t ldx, int64_t K, int nrhs,
                         int ncp, double2* __restrict__ Xp) {
  const int64_t total = K * (int64_t)ncp;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t k = t / ncp;
    const int c = (int)(t - k * ncp);
    Xp[t] = (c < nrhs) ? X[k + (int64_t)c * ldx] : make_double2(0.0, 0.0);
  }
}

__global__ void k_zero_y(double2* __restrict__ Y, int64_t ldy, int64_t M, int nrhs) {
  const int64_t total = M * (int64_t)nrhs;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t c = t / M;
    const int64_t r = t - c * M;
    Y[r + c * ldy] = make_double2(0.0, 0.0);
  }
}

__global__ void __launch_bounds__(256)
k_zgemm_mfma(const double2* __restrict__ A, int64_t lda, const double2* __restrict__ Xp,
             int ncp, double2* __restrict__ Y, int64_t ldy, int64_t M, int64_t K, int nrhs,
             int64_t klen) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int li = lane & 15;   // row within a 16-row A tile / column within a B tile
  const int lk = lane >> 4;   // k within the 4-deep step
  const int64_t r0 = (int64_t)blockIdx.x * 128 + wave * 32;
  const int c0 = blockIdx.z * 32;
  const int64_t kb = (int64_t)blockIdx.y * klen;
  const int64_t ke = min(K, kb + klen);
  v4d re[2][2], im[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      re[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
      im[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    }
  // Operand loads are bounds-checked buffer loads over this split's k range [kb, ke): k >= ke
  // reads zeros with no lane mask or branch, so the compiler can count the loads in flight
  // instead of waiting for all of them.  Rows past M read in-range values (columns kb..ke-1
  // of A) or zeros, and their results are never stored (each output row depends on its own
  // A row only).  The host keeps (klen + 64) 16 max(lda, ncp) below 2^31.
  const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double2*>(A + kb * lda), (short)0, (int)((ke - kb) * lda * 16), 0x00020000);
  const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double2*>(Xp + kb * ncp), (short)0, (int)((ke - kb) * ncp * 16), 0x00020000);
  const int64_t row0 = r0 + li;
  auto ld = [](__amdgpu_buffer_rsrc_t rs, int64_t idx) {
    return __builtin_bit_cast(double2,
                              __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(idx * 16), 0, 0));
  };
  auto loadA = [&](int64_t k, double2& a0, double2& a1) {
    a0 = ld(rsA, (k - kb) * lda + row0);
    a1 = ld(rsA, (k - kb) * lda + row0 + 16);
  };
  auto loadB = [&](int64_t k, double2& b0, double2& b1) {
    b0 = ld(rsB, (k - kb) * ncp + c0 + li);
    b1 = ld(rsB, (k - kb) * ncp + c0 + 16 + li);
  };
  // ring of kRing k steps of A and X operands in flight: HBM latency is several k steps of
  // MFMA work, so one step of prefetch left the matrix pipe waiting
  double2 ra[kRing][2], rb[kRing][2];
#pragma unroll
  for (int p = 0; p < kRing; ++p) {
    loadA(kb + 4 * p + lk, ra[p][0], ra[p][1]);
    loadB(kb + 4 * p + lk, rb[p][0], rb[p][1]);
  }
  for (int64_t k0 = kb; k0 < ke; k0 += 4 * kRing) {
#pragma unroll
    for (int p = 0; p < kRing; ++p) {
      const double ar[2] = {ra[p][0].x, ra[p][1].x}, ai[2] = {ra[p][0].y, ra[p][1].y};
      const double br[2] = {rb[p][0].x, rb[p][1].x}, bi[2] = {rb[p][0].y, rb[p][1].y};
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          re[a][b] = mfma4(ar[a], br[b], re[a][b]);
          im[a][b] = mfma4(ar[a], bi[b], im[a][b]);
        }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          re[a][b] = mfma4(-ai[a], bi[b], re[a][b]);
          im[a][b] = mfma4(ai[a], br[b], im[a][b]);
        }
      const int64_t kn = k0 + 4 * (p + kRing) + lk;  // refill this ring entry
      loadA(kn, ra[p][0], ra[p][1]);
      loadB(kn, rb[p][0], rb[p][1]);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = c0 + b * 16 + li;
      if (col >= nrhs) continue;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int64_t row = r0 + a * 16 + 4 * t + lk;
        if (row < M) {
          double* y = reinterpret_cast<double*>(Y + row + (int64_t)col * ldy);
          unsafeAtomicAdd(y, re[a][b][t]);
          unsafeAtomicAdd(y + 1, im[a][b][t]);
        }
      }
    }
}

// Workgroups of k_zgemm_mfma resident at once on a device: CUs x occupancy (cached).
int resident_workgroups(int dev) {
  static std::mutex mu;
  static int cached[16] = {};
  std::lock_guard<std::mutex> lk(mu);
  if (cached[dev] == 0) {
    int cus = 0, per_cu = 0;
    BIE_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BIE_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_zgemm_mfma), 256, 0));
    cached[dev] = std::max(1, cus * std::max(1, per_cu));
  }
  return cached[dev];
}

// y_q += A_q x_q for the systems q = idx[blockIdx.z] of a strided batch of column-major
// complex matrices (A_q at A + q strideA, leading dimension lda), x_q at X + q ldx, y_q at
// Y + q ldy: the matrix-vector product of Krylov solves over a frequency sweep's operators.
// One row per thread, so a wave reads 1 KB of a column at a time; kCols columns are loaded
// before use to keep enough bytes in flight; x is wave-uniform and comes through scalar
// loads; K is split over blockIdx.y with FP64 atomics into y.  Memory-bound: 16 bytes of
// A per complex multiply-add.
constexpr int kGemvCols = 8;

__global__ void __launch_bounds__(256)
k_zgemv_batched(const double2* __restrict__ A, int64_t strideA, int64_t lda,
                const double2* __restrict__ X, int64_t ldx, double2* __restrict__ Y,
                int64_t ldy, const int* __restrict__ idx, int64_t M, int64_t K, int64_t klen) {
  const int q = sload(idx + blockIdx.z);
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t kb = (int64_t)blockIdx.y * klen;
  const int64_t ke = min(K, kb + klen);
  if (r >= M || kb >= ke) return;
  const double2* __restrict__ a = A + (size_t)q * strideA + r;
  const double* xq = reinterpret_cast<const double*>(X + (size_t)q * ldx);
  auto xload = [&](int64_t jj) { return make_double2(sload(xq + 2 * jj), sload(xq + 2 * jj + 1)); };
  double yr = 0.0, yi = 0.0;
  int64_t j = kb;
  for (; j + kGemvCols <= ke; j += kGemvCols) {
    double2 av[kGemvCols];
#pragma unroll
    for (int t = 0; t < kGemvCols; ++t) av[t] = a[(j + t) * lda];
#pragma unroll
    for (int t = 0; t < kGemvCols; ++t) {
      const double2 x = xload(j + t);
      yr = fma(av[t].x, x.x, fma(-av[t].y, x.y, yr));
      yi = fma(av[t].x, x.y, fma(av[t].y, x.x, yi));
    }
  }
  for (; j < ke; ++j) {
    const double2 av = a[j * lda];
    const double2 x = xload(j);
    yr = fma(av.x, x.x, fma(-av.y, x.y, yr));
    yi = fma(av.x, x.y, fma(av.y, x.x, yi));
  }
  double* y = reinterpret_cast<double*>(Y + (size_t)q * ldy + r);
  unsafeAtomicAdd(y, yr);
  unsafeAtomicAdd(y + 1, yi);
}

// Per-device packed-X workspace shared by all calls.  Its mutex is held while a call
// enqueues its kernels; the call's end is recorded in done[dev], and a call on another
// stream waits for it before overwriting the workspace (a growth waits on the host before
// the old buffer is freed).
struct Workspace {
  std::mutex mu;
  double2* p[16] = {};
  size_t bytes[16] = {};
  hipEvent_t done[16] = {};
  hipStream_t done_stream[16] = {};
  bool done_valid[16] = {};
};
Workspace g_ws;

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  BIE_HIP_CHECK(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}

}  // namespace

namespace {

// Checks the shapes of Y = A X; false when there is nothing to compute.
bool zgemm_validate(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                    const double* dX, int64_t ldx, double* dY, int64_t ldy) {
  BIE_REQUIRE(M >= 0 && nrhs >= 0 && K >= 0, BIE_EINVAL, "negative size");
  BIE_REQUIRE(lda >= (M > 0 ? M : 1) && ldx >= (K > 0 ? K : 1) && ldy >= (M > 0 ? M : 1),
              BIE_EINVAL, "leading dimension too small");
  BIE_REQUIRE(nrhs <= (1 << 20), BIE_EINVAL, "too many right-hand sides");
  if (M == 0 || nrhs == 0) return false;
  BIE_REQUIRE(dA && dX && dY, BIE_EINVAL, "NULL pointer");
  const int64_t ncp = (nrhs + 31) / 32 * 32;
  const int64_t wide = std::max<int64_t>(lda, ncp) * 16;
  BIE_REQUIRE(wide * (64 + 4 * kRing) < (1LL << 31), BIE_EINVAL,
              "apply: leading dimension too large (lda <= 1.6e6)");
  return true;
}

// The kernels of Y = A X with the packed copy of X in Xp (K x ncp complex entries), all on
// stream s: Y zeroed by a fill kernel, X packed, the MFMA product with K splits.
void zgemm_launch(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                  const double* dX, int64_t ldx, double* dY, int64_t ldy, double2* Xp, int dev,
                  hipStream_t s) {
  const int ncp = (int)((nrhs + 31) / 32 * 32);
  const int64_t wide = std::max<int64_t>(lda, ncp) * 16;
  {
    int64_t blocks = (M * nrhs + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero_y, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<double2*>(dY), ldy, M, (int)nrhs);
  }
  if (K == 0) {
    BIE_HIP_CHECK(hipGetLastError());
    return;
  }
  {
    int64_t blocks = (K * ncp + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pack_x, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(dX), ldx, K, (int)nrhs, ncp, Xp);
  }
  const int64_t rowblocks = (M + 127) / 128;
  const int colblocks = ncp / 32;
  // K splits: one round of resident workgroups, each with as long a K range as fits (more
  // splits than fill one round add a mostly empty round, or more partial-sum atomics)
  int64_t splits = resident_workgroups(dev) / (rowblocks * colblocks);
  const int64_t kmax_splits = (K + 255) / 256;
  if (splits > kmax_splits) splits = kmax_splits;
  if (splits < 1) splits = 1;
  if (splits > 65535) splits = 65535;
  const int64_t klen_max = ((1LL << 31) - 1) / wide - 64;
  int64_t klen = (K + splits - 1) / splits;
  klen = (klen + 4 * kRing - 1) / (4 * kRing) * (4 * kRing);  // whole rings but in the last split
  if (klen > klen_max) klen = klen_max / (4 * kRing) * (4 * kRing);
  splits = (K + klen - 1) / klen;
  BIE_REQUIRE(splits <= 65535, BIE_EINVAL, "apply: too many K splits");
  BIE_REQUIRE(rowblocks < 2147483647LL && colblocks <= 65535, BIE_EINVAL, "apply too large");
  hipLaunchKernelGGL(k_zgemm_mfma, dim3((unsigned)rowblocks, (unsigned)splits, colblocks),
                     dim3(256), 0, s, reinterpret_cast<const double2*>(dA), lda, Xp, ncp,
                     reinterpret_cast<double2*>(dY), ldy, M, K, (int)nrhs, klen);
  BIE_HIP_CHECK(hipGetLastError());
}

int current_device() {
  int dev = 0;
  BIE_HIP_CHECK(hipGetDevice(&dev));
  BIE_REQUIRE(dev >= 0 && dev < 16, BIE_EINVAL, "device ordinal out of range");
  return dev;
}

}  // namespace

// Packed-X workspace of the allocation-free apply (zgemm_device_noalloc): a file-scope
// device array, one copy on each GPU the library runs on, held until the process ends.
// Its host-side name is not a device address: host code takes the address each device
// reports for it (scratch_address) and passes that to kernels.
static __device__ __align__(256) double2 scratch[kApplyScratchComplex];

namespace {

// Device address of `scratch` on device dev (the current device), looked up once per
// device.
double2* scratch_address(int dev) {
  static std::mutex mu;
  static double2* addr[16] = {};
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    BIE_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BIE_REQUIRE(p != nullptr, BIE_EHIP, "no device address for the apply scratch array");
    addr[dev] = static_cast<double2*>(p);
  }
  return addr[dev];
}

}  // namespace

void zgemm_device_noalloc(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                          const double* dX, int64_t ldx, double* dY, int64_t ldy,
                          hipStream_t s) {
  if (!zgemm_validate(M, nrhs, K, dA, lda, dX, ldx, dY, ldy)) return;
  const int64_t ncp = (nrhs + 31) / 32 * 32;
  BIE_REQUIRE(K <= kApplyScratchComplex / ncp, BIE_EINVAL,
              "apply without allocation: K * ceil(nrhs/32)*32 = " + std::to_string(K * ncp) +
                  " complex entries exceed the scratch array's " +
                  std::to_string(kApplyScratchComplex) + " (16 MiB)");
  const int dev = current_device();
  zgemm_launch(M, nrhs, K, dA, lda, dX, ldx, dY, ldy, scratch_address(dev), dev, s);
}

void zgemm_device(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                  const double* dX, int64_t ldx, double* dY, int64_t ldy, hipStream_t s) {
  if (!zgemm_validate(M, nrhs, K, dA, lda, dX, ldx, dY, ldy)) return;
  const int ncp = (int)((nrhs + 31) / 32 * 32);
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(g_ws.mu);  // held until this call's kernels are enqueued
  const bool cap = stream_capturing(s);
  if (!g_ws.done[dev]) BIE_HIP_CHECK(hipEventCreateWithFlags(&g_ws.done[dev], hipEventDisableTiming));
  double2* Xp = nullptr;
  if (K > 0) {
    const size_t need = (size_t)K * ncp * sizeof(double2);
    if (g_ws.bytes[dev] < need) {
      BIE_REQUIRE(!cap, BIE_EINVAL,
                  "apply: the workspace must grow, which a stream under graph capture cannot "
                  "do (call once outside the capture first, or use bie_zgemm_device_noalloc)");
      if (g_ws.p[dev]) {
        if (g_ws.done_valid[dev]) BIE_HIP_CHECK(hipEventSynchronize(g_ws.done[dev]));
        (void)hipFree(g_ws.p[dev]);
      }
      g_ws.p[dev] = nullptr;
      g_ws.bytes[dev] = 0;
      void* p = nullptr;
      if (hipMalloc(&p, need) != hipSuccess) {
        (void)hipGetLastError();
        throw Error{BIE_ENOMEM, "apply workspace allocation failed"};
      }
      g_ws.p[dev] = static_cast<double2*>(p);
      g_ws.bytes[dev] = need;
    } else if (g_ws.done_valid[dev] && g_ws.done_stream[dev] != s && !cap) {
      BIE_HIP_CHECK(hipStreamWaitEvent(s, g_ws.done[dev], 0));
    }
    Xp = g_ws.p[dev];
  }
  zgemm_launch(M, nrhs, K, dA, lda, dX, ldx, dY, ldy, Xp, dev, s);
  if (!cap && K > 0) {
    BIE_HIP_CHECK(hipEventRecord(g_ws.done[dev], s));
    g_ws.done_stream[dev] = s;
    g_ws.done_valid[dev] = true;
  }
}

void zgemv_batched_device(int64_t M, int64_t K, const double* dA, int64_t strideA, int64_t lda,
                          const double* dX, int64_t ldx, double* dY, int64_t ldy,
                          const int32_t* d_idx, int32_t n_idx, hipStream_t s) {
  BIE_REQUIRE(M >= 0 && K >= 0 && n_idx >= 0, BIE_EINVAL, "negative size");
  BIE_REQUIRE(n_idx <= 65535, BIE_EINVAL, "more than 65535 systems in one call");
  BIE_REQUIRE(lda >= (M > 0 ? M : 1) && ldx >= (K > 0 ? K : 1) && ldy >= (M > 0 ? M : 1),
              BIE_EINVAL, "leading dimension too small");
  BIE_REQUIRE(strideA >= lda * K, BIE_EINVAL, "matrix stride smaller than one matrix");
  if (M == 0 || K == 0 || n_idx == 0) return;
  BIE_REQUIRE(dA && dX && dY && d_idx, BIE_EINVAL, "NULL pointer");
  int dev = 0;
  BIE_HIP_CHECK(hipGetDevice(&dev));
  int cus = 0;
  BIE_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t rowblocks = (M + 255) / 256;
  // enough workgroups for the memory system (8 per CU), but columns of at least 64
  int64_t splits = (8LL * cus + rowblocks * n_idx - 1) / (rowblocks * n_idx);
  splits = std::max<int64_t>(1, std::min<int64_t>(splits, (K + 63) / 64));
  int64_t klen = (K + splits - 1) / splits;
  splits = (K + klen - 1) / klen;
  BIE_REQUIRE(rowblocks < 2147483647LL && splits <= 65535, BIE_EINVAL, "gemv too large");
  hipLaunchKernelGGL(k_zgemv_batched, dim3((unsigned)rowblocks, (unsigned)splits, (unsigned)n_idx),
                     dim3(256), 0, s, reinterpret_cast<const double2*>(dA), strideA, lda,
                     reinterpret_cast<const double2*>(dX), ldx, reinterpret_cast<double2*>(dY),
                     ldy, reinterpret_cast<const int*>(d_idx), M, K, klen);
  BIE_HIP_CHECK(hipGetLastError());
}

namespace {

struct TaylorCoefs {
  double c[2 * kTaylorMaxOrders];  // interleaved complex coefficients (1 KiB of kernel arguments)
};

// One complex entry per thread and pass: the n_terms loads are independent, so a wave keeps
// several in flight; every byte is touched once, so loads and stores bypass the caches
// (non-temporal).  Y may be A_0 itself (each entry is read before it is written, by the
// same thread).
__global__ void __launch_bounds__(256)
k_taylor_expand(int nt, int64_t n, const double2* A, int64_t strideA, const TaylorCoefs cf,
                double2* Y) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += step) {
    double yr = 0.0, yi = 0.0;
#pragma unroll 4
    for (int t = 0; t < nt; ++t) {
      const double* pa = reinterpret_cast<const double*>(A + (size_t)t * strideA + e);
      const double vx = __builtin_nontemporal_load(pa), vy = __builtin_nontemporal_load(pa + 1);
      const double cr = cf.c[2 * t], ci = cf.c[2 * t + 1];
      yr = fma(cr, vx, fma(-ci, vy, yr));
      yi = fma(cr, vy, fma(ci, vx, yi));
    }
    double* py = reinterpret_cast<double*>(Y + e);
    __builtin_nontemporal_store(yr, py);
    __builtin_nontemporal_store(yi, py + 1);
  }
}

}  // namespace

void taylor_expand_device(int n_terms, int64_t n, const double* dA, int64_t strideA,
                          const double* coef, double* dY, hipStream_t s) {
  BIE_REQUIRE(n_terms >= 1 && n_terms <= kTaylorMaxOrders, BIE_EINVAL,
              "n_terms must be in 1..64");
  BIE_REQUIRE(n >= 0, BIE_EINVAL, "negative size");
  BIE_REQUIRE(n_terms == 1 || strideA >= n, BIE_EINVAL, "stride smaller than one array");
  BIE_REQUIRE(coef != nullptr, BIE_EINVAL, "coefficients are NULL");
  if (n == 0) return;
  BIE_REQUIRE(dA && dY, BIE_EINVAL, "NULL pointer");
  {  // dY may be A_0 itself, but must not overlap any other term
    const uintptr_t y0 = reinterpret_cast<uintptr_t>(dY), y1 = y0 + 16 * (uintptr_t)n;
    for (int t = 0; t < n_terms; ++t) {
      const uintptr_t a0 = reinterpret_cast<uintptr_t>(dA) + 16 * (uintptr_t)t * strideA;
      const uintptr_t a1 = a0 + 16 * (uintptr_t)n;
      const bool same = t == 0 && a0 == y0;
      BIE_REQUIRE(same || y1 <= a0 || a1 <= y0, BIE_EINVAL,
                  "output overlaps a term other than the first");
    }
  }
  TaylorCoefs cf{};
  for (int t = 0; t < 2 * n_terms; ++t) {
    BIE_REQUIRE(std::isfinite(coef[t]), BIE_EINVAL, "coefficient is not finite");
    cf.c[t] = coef[t];
  }
  int dev = 0, cus = 0;
  BIE_HIP_CHECK(hipGetDevice(&dev));
  BIE_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 64LL * cus));
  hipLaunchKernelGGL(k_taylor_expand, dim3((unsigned)blocks), dim3(256), 0, s, n_terms, n,
                     reinterpret_cast<const double2*>(dA), strideA, cf,
                     reinterpret_cast<double2*>(dY));
  BIE_HIP_CHECK(hipGetLastError());
}

}  // namespace bie
