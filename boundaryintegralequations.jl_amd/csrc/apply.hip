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
// rows; K is split over blockIdx.y and the partial sums are added into a zeroed Y with
// FP64 atomics.  X is first repacked k-major (K x ncp, ncp = nrhs rounded up to 32) so
// the B operand is a coalesced 16-byte load per lane.
#include <hip/hip_runtime.h>

#include <mutex>

#include "bie_common.hpp"

namespace bie {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

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
  const int64_t row0 = r0 + li, row1 = r0 + 16 + li;
  const bool ok0 = row0 < M, ok1 = row1 < M;
  auto loadA = [&](int64_t k, double2& a0, double2& a1) {
    const bool kk = k < ke;
    a0 = (kk && ok0) ? A[row0 + k * lda] : make_double2(0.0, 0.0);
    a1 = (kk && ok1) ? A[row1 + k * lda] : make_double2(0.0, 0.0);
  };
  auto loadB = [&](int64_t k, double2& b0, double2& b1) {
    const bool kk = k < ke;
    b0 = kk ? Xp[k * ncp + c0 + li] : make_double2(0.0, 0.0);
    b1 = kk ? Xp[k * ncp + c0 + 16 + li] : make_double2(0.0, 0.0);
  };
  double2 a0, a1, b0, b1;
  loadA(kb + lk, a0, a1);
  loadB(kb + lk, b0, b1);
  for (int64_t k0 = kb; k0 < ke; k0 += 4) {
    double2 na0, na1, nb0, nb1;
    loadA(k0 + 4 + lk, na0, na1);
    loadB(k0 + 4 + lk, nb0, nb1);
    const double ar[2] = {a0.x, a1.x}, ai[2] = {a0.y, a1.y};
    const double br[2] = {b0.x, b1.x}, bi[2] = {b0.y, b1.y};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        re[a][b] = mfma4(ar[a], br[b], re[a][b]);
        re[a][b] = mfma4(-ai[a], bi[b], re[a][b]);
        im[a][b] = mfma4(ar[a], bi[b], im[a][b]);
        im[a][b] = mfma4(ai[a], br[b], im[a][b]);
      }
    a0 = na0;
    a1 = na1;
    b0 = nb0;
    b1 = nb1;
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

struct Workspace {
  std::mutex mu;
  double2* p[16] = {};
  size_t bytes[16] = {};
};
Workspace g_ws;

}  // namespace

void zgemm_device(int64_t M, int64_t nrhs, int64_t K, const double* dA, int64_t lda,
                  const double* dX, int64_t ldx, double* dY, int64_t ldy, hipStream_t s) {
  BIE_REQUIRE(M >= 0 && nrhs >= 0 && K >= 0, BIE_EINVAL, "negative size");
  BIE_REQUIRE(lda >= (M > 0 ? M : 1) && ldx >= (K > 0 ? K : 1) && ldy >= (M > 0 ? M : 1),
              BIE_EINVAL, "leading dimension too small");
  BIE_REQUIRE(nrhs <= (1 << 20), BIE_EINVAL, "too many right-hand sides");
  if (M == 0 || nrhs == 0) return;
  BIE_REQUIRE(dA && dX && dY, BIE_EINVAL, "NULL pointer");
  {
    int64_t blocks = (M * nrhs + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero_y, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<double2*>(dY), ldy, M, (int)nrhs);
  }
  if (K == 0) return;
  const int ncp = (int)((nrhs + 31) / 32 * 32);
  int dev = 0;
  BIE_HIP_CHECK(hipGetDevice(&dev));
  BIE_REQUIRE(dev >= 0 && dev < 16, BIE_EINVAL, "device ordinal out of range");
  double2* Xp = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ws.mu);
    const size_t need = (size_t)K * ncp * sizeof(double2);
    if (g_ws.bytes[dev] < need) {
      if (g_ws.p[dev]) {
        BIE_HIP_CHECK(hipDeviceSynchronize());
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
    }
    Xp = g_ws.p[dev];
  }
  {
    int64_t blocks = (K * ncp + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pack_x, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(dX), ldx, K, (int)nrhs, ncp, Xp);
  }
  const int64_t rowblocks = (M + 127) / 128;
  const int colblocks = ncp / 32;
  int64_t splits = (1024 + rowblocks * colblocks - 1) / (rowblocks * colblocks);
  const int64_t kmax_splits = (K + 255) / 256;
  if (splits > kmax_splits) splits = kmax_splits;
  if (splits < 1) splits = 1;
  if (splits > 65535) splits = 65535;
  int64_t klen = (K + splits - 1) / splits;
  klen = (klen + 3) / 4 * 4;
  splits = (K + klen - 1) / klen;
  BIE_REQUIRE(rowblocks < 2147483647LL && colblocks <= 65535, BIE_EINVAL, "apply too large");
  hipLaunchKernelGGL(k_zgemm_mfma, dim3((unsigned)rowblocks, (unsigned)splits, colblocks),
                     dim3(256), 0, s, reinterpret_cast<const double2*>(dA), lda, Xp, ncp,
                     reinterpret_cast<double2*>(dY), ldy, M, K, (int)nrhs, klen);
  BIE_HIP_CHECK(hipGetLastError());
}

}  // namespace bie
