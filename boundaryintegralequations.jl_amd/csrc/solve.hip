// Solves on the device for the systems the assembly produces, so F never leaves HBM.
//
// The reference forms the collocation system A = F + diag(1 - C) and solves A p = b with
// Julia's `\` (LAPACK getrf/getrs on the host, examples/3d_rigid_sphere.jl:47-52,
// test/3d/assembly_test.jl:37) or with IterativeSolvers.gmres; the field pressure is
// p_inc - F_field p (examples/3d_rosebem.jl:81).  Here:
//   * LU: rocSOLVER zgetrf / zgetrs on the assembled matrix in place (A := F + diag(1 - C)
//     by a small kernel first);
//   * y = op(A) x: rocBLAS zgemv;
//   * restarted GMRES over many systems (F_q + diag(1 - C)) x_q = b_q at once (a frequency
//     sweep's operators): the Arnoldi process on the device (operator application by the
//     batched GEMV kernel, classical Gram-Schmidt applied twice by the kernels below), the
//     small Hessenberg least-squares problems (Givens rotations) on the host, one small
//     device-to-host transfer per Arnoldi step for all systems.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <string>
#include <vector>

#include "bie_common.hpp"

namespace bie {
namespace {

// A_ii += 1 - C_i (A column-major, leading dimension lda).
__global__ void k_add_free_term(double2* __restrict__ A, int64_t lda, const double2* __restrict__ C,
                                int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 a = A[i + i * lda];
  const double2 c = C[i];
  A[i + i * lda] = make_double2(a.x + 1.0 - c.x, a.y - c.y);
}

// One rocBLAS handle per device (rocSOLVER runs on it), guarded by a mutex because the
// stream is a property of the handle.
struct Handles {
  std::mutex mu;
  rocblas_handle h[16] = {};
};
Handles g_handles;

void check_blas(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    throw Error{BIE_EHIP, std::string(what) + " failed: " + rocblas_status_to_string(st)};
}

rocblas_handle handle_for(int dev) {
  if (!g_handles.h[dev]) check_blas(rocblas_create_handle(&g_handles.h[dev]), "rocblas_create_handle");
  return g_handles.h[dev];
}

int current_device() {
  int dev = 0;
  BIE_HIP_CHECK(hipGetDevice(&dev));
  BIE_REQUIRE(dev >= 0 && dev < 16, BIE_EINVAL, "device ordinal out of range");
  return dev;
}

rocblas_operation op_of(int trans) {
  BIE_REQUIRE(trans >= 0 && trans <= 2, BIE_EINVAL, "trans must be 0 (N), 1 (T) or 2 (C)");
  return trans == 0 ? rocblas_operation_none
                    : (trans == 1 ? rocblas_operation_transpose
                                  : rocblas_operation_conjugate_transpose);
}

// ---- GMRES kernels ---------------------------------------------------------------------
// V: per system (m + 1) Krylov vectors of n entries, system q at V + q * sv; W, X, R, B:
// system q at + q * ldw (etc.).  sys[] lists the systems a launch works on (blockIdx.y).

constexpr int kRed = 256;

__device__ __forceinline__ double2 block_sum2(double2 v, double2* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRed / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[threadIdx.x].x += sh[threadIdx.x + s].x;
      sh[threadIdx.x].y += sh[threadIdx.x + s].y;
    }
    __syncthreads();
  }
  return sh[0];
}

// h[q][i] = V_q[i]^H w_q for i <= j  (blockIdx.x = i, blockIdx.y = listed system)
__global__ void __launch_bounds__(kRed)
k_gs_dot(const double2* __restrict__ V, int64_t sv, int64_t n, const double2* __restrict__ W,
         int64_t ldw, const int* __restrict__ sys, double2* __restrict__ h, int hs) {
  __shared__ double2 sh[kRed];
  const int q = sys[blockIdx.y];
  const int i = blockIdx.x;
  const double2* v = V + q * sv + (int64_t)i * n;
  const double2* w = W + q * ldw;
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t t = threadIdx.x; t < n; t += kRed) {
    const double2 a = v[t], b = w[t];
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));   // conj(a) b
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
  }
  const double2 s = block_sum2(acc, sh);
  if (threadIdx.x == 0) h[(int64_t)q * hs + i] = s;
}

// w_q -= sum_{i <= j} h[q][i] V_q[i]   (or, with add, x_q += sum_i y[q][i] V_q[i])
__global__ void __launch_bounds__(256)
k_gs_update(const double2* __restrict__ V, int64_t sv, int64_t n, int jn,
            const double2* __restrict__ h, int hs, double2* __restrict__ W, int64_t ldw,
            const int* __restrict__ sys, double sign) {
  const int q = sys[blockIdx.y];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double2* v = V + q * sv + t;
  const double2* hq = h + (int64_t)q * hs;
  double2 acc = make_double2(0.0, 0.0);
  for (int i = 0; i < jn; ++i) {
    const double2 a = v[(int64_t)i * n], c = hq[i];
    acc.x = fma(a.x, c.x, fma(-a.y, c.y, acc.x));
    acc.y = fma(a.x, c.y, fma(a.y, c.x, acc.y));
  }
  double2* w = W + q * ldw + t;
  *w = make_double2(w->x + sign * acc.x, w->y + sign * acc.y);
}

// out[q] = ||w_q||_2 (real, in the .x of a complex slot; .y = 0)
__global__ void __launch_bounds__(kRed)
k_norm(const double2* __restrict__ W, int64_t ldw, int64_t n, const int* __restrict__ sys,
       double2* __restrict__ out, int os, int oi) {
  __shared__ double2 sh[kRed];
  const int q = sys[blockIdx.x];
  const double2* w = W + q * ldw;
  double acc = 0.0;
  for (int64_t t = threadIdx.x; t < n; t += kRed) acc = fma(w[t].x, w[t].x, fma(w[t].y, w[t].y, acc));
  const double2 s = block_sum2(make_double2(acc, 0.0), sh);
  if (threadIdx.x == 0) out[(int64_t)q * os + oi] = make_double2(sqrt(s.x), 0.0);
}

// dst_q = src_q * scale[q]  (scale real)
__global__ void k_scale(const double2* __restrict__ src, int64_t lds_, double2* __restrict__ dst,
                        int64_t ldd, int64_t n, const int* __restrict__ sys,
                        const double* __restrict__ scale) {
  const int q = sys[blockIdx.y];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double s = scale[q];
  const double2 a = src[q * lds_ + t];
  dst[q * ldd + t] = make_double2(a.x * s, a.y * s);
}

// y_q = (1 - C) .* x_q   (the diagonal part of the operator; the F part is added after)
__global__ void k_free_term_apply(const double2* __restrict__ X, int64_t ldx,
                                  const double2* __restrict__ C, double2* __restrict__ Y,
                                  int64_t ldy, int64_t n, const int* __restrict__ sys) {
  const int q = sys[blockIdx.y];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double2 x = X[q * ldx + t], c = C[t];
  const double dr = 1.0 - c.x, di = -c.y;
  Y[q * ldy + t] = make_double2(dr * x.x - di * x.y, dr * x.y + di * x.x);
}

// r_q = b_q - r_q
__global__ void k_residual(const double2* __restrict__ B, int64_t ldb, double2* __restrict__ R,
                           int64_t ldr, int64_t n, const int* __restrict__ sys) {
  const int q = sys[blockIdx.y];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double2 b = B[q * ldb + t], r = R[q * ldr + t];
  R[q * ldr + t] = make_double2(b.x - r.x, b.y - r.y);
}

template <class T>
struct DevArray {
  T* p = nullptr;
  explicit DevArray(size_t count) {
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      throw Error{BIE_ENOMEM, "solver workspace allocation failed"};
    }
  }
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
};

}  // namespace

void lu_factor_device(int64_t n, double* dA, int64_t lda, const double* dC, int32_t* d_ipiv,
                      hipStream_t s) {
  BIE_REQUIRE(n >= 0 && lda >= std::max<int64_t>(n, 1), BIE_EINVAL, "bad size or leading dimension");
  BIE_REQUIRE(n < (int64_t)1 << 31 && lda < (int64_t)1 << 31, BIE_EINVAL, "matrix too large");
  if (n == 0) return;
  BIE_REQUIRE(dA && d_ipiv, BIE_EINVAL, "NULL pointer");
  const int dev = current_device();
  if (dC) {
    hipLaunchKernelGGL(k_add_free_term, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<double2*>(dA), lda, reinterpret_cast<const double2*>(dC), n);
    BIE_HIP_CHECK(hipGetLastError());
  }
  DevArray<rocblas_int> info(1);
  rocblas_int h_info = 0;
  {
    std::lock_guard<std::mutex> lk(g_handles.mu);
    rocblas_handle h = handle_for(dev);
    check_blas(rocblas_set_stream(h, s), "rocblas_set_stream");
    check_blas(rocsolver_zgetrf(h, (rocblas_int)n, (rocblas_int)n,
                                reinterpret_cast<rocblas_double_complex*>(dA), (rocblas_int)lda,
                                reinterpret_cast<rocblas_int*>(d_ipiv), info.p),
               "rocsolver_zgetrf");
    BIE_HIP_CHECK(hipMemcpyAsync(&h_info, info.p, sizeof(h_info), hipMemcpyDeviceToHost, s));
    BIE_HIP_CHECK(hipStreamSynchronize(s));
  }
  BIE_REQUIRE(h_info == 0, BIE_EINVAL,
              "singular matrix: zero pivot in column " + std::to_string(h_info));
}

void lu_solve_device(int64_t n, const double* dA, int64_t lda, const int32_t* d_ipiv, int trans,
                     double* dB, int64_t ldb, int64_t nrhs, hipStream_t s) {
  BIE_REQUIRE(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) &&
                  ldb >= std::max<int64_t>(n, 1),
              BIE_EINVAL, "bad size or leading dimension");
  BIE_REQUIRE(n < (int64_t)1 << 31 && nrhs < (int64_t)1 << 31 && lda < (int64_t)1 << 31 &&
                  ldb < (int64_t)1 << 31,
              BIE_EINVAL, "matrix too large");
  const rocblas_operation op = op_of(trans);
  if (n == 0 || nrhs == 0) return;
  BIE_REQUIRE(dA && d_ipiv && dB, BIE_EINVAL, "NULL pointer");
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(g_handles.mu);
  rocblas_handle h = handle_for(dev);
  check_blas(rocblas_set_stream(h, s), "rocblas_set_stream");
  check_blas(rocsolver_zgetrs(h, op, (rocblas_int)n, (rocblas_int)nrhs,
                              reinterpret_cast<rocblas_double_complex*>(const_cast<double*>(dA)),
                              (rocblas_int)lda, reinterpret_cast<const rocblas_int*>(d_ipiv),
                              reinterpret_cast<rocblas_double_complex*>(dB), (rocblas_int)ldb),
             "rocsolver_zgetrs");
}

void zgemv_device(int trans, int64_t m, int64_t n, const double* dA, int64_t lda,
                  const double* dx, double* dy, hipStream_t s) {
  BIE_REQUIRE(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1), BIE_EINVAL,
              "bad size or leading dimension");
  BIE_REQUIRE(m < (int64_t)1 << 31 && n < (int64_t)1 << 31 && lda < (int64_t)1 << 31,
              BIE_EINVAL, "matrix too large");
  const rocblas_operation op = op_of(trans);
  const int64_t ny = trans == 0 ? m : n, nx = trans == 0 ? n : m;
  if (ny == 0) return;
  BIE_REQUIRE(dA && dx && dy, BIE_EINVAL, "NULL pointer");
  if (nx == 0) {  // empty product: y = 0 (a fill kernel, so the call can be captured)
    launch_fill_zero(dy, 2 * ny, s);
    return;
  }
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(g_handles.mu);
  rocblas_handle h = handle_for(dev);
  check_blas(rocblas_set_stream(h, s), "rocblas_set_stream");
  check_blas(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "rocblas_set_pointer_mode");
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
  check_blas(rocblas_zgemv(h, op, (rocblas_int)m, (rocblas_int)n, &one,
                           reinterpret_cast<const rocblas_double_complex*>(dA), (rocblas_int)lda,
                           reinterpret_cast<const rocblas_double_complex*>(dx), 1, &zero,
                           reinterpret_cast<rocblas_double_complex*>(dy), 1),
             "rocblas_zgemv");
}

// Restarted GMRES over nsys systems in lockstep (the algorithm of one system: Arnoldi with
// classical Gram-Schmidt applied twice, Givens rotations on the small Hessenberg matrix,
// restart after `restart` steps, stop at ||b_q - A_q x_q|| <= tol ||b_q||).
void gmres_batched_device(int64_t n, int nsys, const double* dF, int64_t strideF, int64_t ldf,
                          const double* dC, const double* dB, int64_t ldb, double* dX,
                          int64_t ldx, double tol, int restart, int maxiter, int32_t* iters,
                          double* resid, hipStream_t s) {
  BIE_REQUIRE(n >= 0 && nsys >= 0 && nsys <= 65535, BIE_EINVAL, "bad system count or size");
  BIE_REQUIRE(ldf >= std::max<int64_t>(n, 1) && strideF >= ldf * n && ldb >= n && ldx >= n,
              BIE_EINVAL, "bad leading dimension or stride");
  BIE_REQUIRE(restart >= 1 && maxiter >= 0 && tol >= 0.0, BIE_EINVAL,
              "restart must be >= 1, maxiter and tol >= 0");
  BIE_REQUIRE(iters && resid, BIE_EINVAL, "NULL result arrays");
  if (nsys == 0) return;
  BIE_REQUIRE(dF && dC && dB && dX, BIE_EINVAL, "NULL pointer");
  for (int q = 0; q < nsys; ++q) {
    iters[q] = 0;
    resid[q] = 0.0;
  }
  if (n == 0) return;
  const int m = (int)std::min<int64_t>(restart, n);
  const int hs = m + 1;                        // h row length per system
  const int64_t sv = (int64_t)(m + 1) * n;     // V stride per system
  DevArray<double2> V((size_t)nsys * sv), W((size_t)nsys * n), H((size_t)nsys * hs * 2),
      Y((size_t)nsys * hs);
  DevArray<int> sys(nsys);
  DevArray<double> scale(nsys);
  const double2* C = reinterpret_cast<const double2*>(dC);
  const double2* B = reinterpret_cast<const double2*>(dB);
  double2* X = reinterpret_cast<double2*>(dX);
  const dim3 blk(256);
  const unsigned nb = (unsigned)((n + 255) / 256);
  std::vector<int> list;
  auto upload_list = [&](const std::vector<int>& l) {
    BIE_HIP_CHECK(hipMemcpyAsync(sys.p, l.data(), l.size() * sizeof(int), hipMemcpyHostToDevice, s));
  };
  // Y_q (rows of Wd, leading dimension ldw) = A_q X_q for the listed systems
  auto apply = [&](const double2* Xs, int64_t ldxs, double2* Wd, int64_t ldw, int count) {
    hipLaunchKernelGGL(k_free_term_apply, dim3(nb, count), blk, 0, s, Xs, ldxs, C, Wd, ldw, n, sys.p);
    zgemv_batched_device(n, n, dF, strideF, ldf, reinterpret_cast<const double*>(Xs), ldxs,
                         reinterpret_cast<double*>(Wd), ldw, sys.p, count, s);
  };
  std::vector<double2> hbuf((size_t)nsys * hs);
  std::vector<double> bnorm(nsys), beta(nsys);
  // ||b_q||, and the first residual r = b (x0 = 0)
  list.resize(nsys);
  for (int q = 0; q < nsys; ++q) list[q] = q;
  upload_list(list);
  for (int q = 0; q < nsys; ++q) launch_fill_zero(reinterpret_cast<double*>(X + q * ldx), 2 * n, s);
  // Krylov vectors never written in a cycle are multiplied by zero coefficients: keep them
  // finite
  BIE_HIP_CHECK(hipMemsetAsync(V.p, 0, sizeof(double2) * (size_t)nsys * sv, s));
  hipLaunchKernelGGL(k_norm, dim3(nsys), dim3(kRed), 0, s, B, ldb, n, sys.p, H.p, hs, 0);
  BIE_HIP_CHECK(hipMemcpyAsync(hbuf.data(), H.p, hbuf.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
  BIE_HIP_CHECK(hipStreamSynchronize(s));
  for (int q = 0; q < nsys; ++q) beta[q] = bnorm[q] = hbuf[(size_t)q * hs].x;
  // R lives in W between restarts: W = b
  for (int q = 0; q < nsys; ++q)
    BIE_HIP_CHECK(hipMemcpyAsync(W.p + q * n, B + q * ldb, 16 * (size_t)n, hipMemcpyDeviceToDevice, s));
  std::vector<char> done(nsys);
  for (int q = 0; q < nsys; ++q) done[q] = !(beta[q] > tol * bnorm[q]);
  int total = 0;
  typedef std::complex<double> cd;
  while (total < maxiter && std::count(done.begin(), done.end(), 0) > 0) {
    const int mm = std::min(m, maxiter - total);
    // V_q[0] = r_q / beta_q for the systems not done
    list.clear();
    std::vector<double> sc(nsys, 0.0);
    for (int q = 0; q < nsys; ++q)
      if (!done[q]) {
        list.push_back(q);
        sc[q] = beta[q] > 0 ? 1.0 / beta[q] : 0.0;
      }
    std::vector<char> active(nsys, 0);
    for (int q : list) active[q] = 1;
    upload_list(list);
    BIE_HIP_CHECK(hipMemcpyAsync(scale.p, sc.data(), nsys * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scale, dim3(nb, (unsigned)list.size()), blk, 0, s, W.p, n, V.p, sv, n,
                       sys.p, scale.p);
    std::vector<std::vector<cd>> Hm(nsys, std::vector<cd>((size_t)(mm + 1) * mm, cd(0)));
    std::vector<std::vector<double>> cs(nsys, std::vector<double>(mm, 0.0));
    std::vector<std::vector<cd>> sn(nsys, std::vector<cd>(mm, cd(0)));
    std::vector<std::vector<cd>> g(nsys, std::vector<cd>(mm + 1, cd(0)));
    std::vector<int> kused(nsys, 0);
    for (int q = 0; q < nsys; ++q) g[q][0] = done[q] ? cd(0) : cd(beta[q]);
    int steps = 0;
    for (int j = 0; j < mm && !list.empty(); ++j) {
      const unsigned cnt = (unsigned)list.size();
      upload_list(list);
      apply(V.p + (int64_t)j * n, sv, W.p, n, (int)cnt);                // w = A v_j
      // CGS2: h = V^H w; w -= V h; twice (the second pass's coefficients added to the first)
      hipLaunchKernelGGL(k_gs_dot, dim3(j + 1, cnt), dim3(kRed), 0, s, V.p, sv, n, W.p, (int64_t)n,
                         sys.p, H.p, hs);
      hipLaunchKernelGGL(k_gs_update, dim3(nb, cnt), blk, 0, s, V.p, sv, n, j + 1, H.p, hs, W.p,
                         (int64_t)n, sys.p, -1.0);
      hipLaunchKernelGGL(k_gs_dot, dim3(j + 1, cnt), dim3(kRed), 0, s, V.p, sv, n, W.p, (int64_t)n,
                         sys.p, H.p + (size_t)nsys * hs, hs);
      hipLaunchKernelGGL(k_gs_update, dim3(nb, cnt), blk, 0, s, V.p, sv, n, j + 1,
                         H.p + (size_t)nsys * hs, hs, W.p, (int64_t)n, sys.p, -1.0);
      hipLaunchKernelGGL(k_norm, dim3(cnt), dim3(kRed), 0, s, W.p, (int64_t)n, n, sys.p, H.p, hs,
                         j + 1);
      BIE_HIP_CHECK(hipGetLastError());
      std::vector<double2> h2((size_t)nsys * hs);
      BIE_HIP_CHECK(hipMemcpyAsync(hbuf.data(), H.p, hbuf.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
      BIE_HIP_CHECK(hipMemcpyAsync(h2.data(), H.p + (size_t)nsys * hs, h2.size() * sizeof(double2),
                                   hipMemcpyDeviceToHost, s));
      BIE_HIP_CHECK(hipStreamSynchronize(s));
      std::vector<double> inv(nsys, 0.0);
      std::vector<int> next;
      for (int q : list) {
        const double2* hq = &hbuf[(size_t)q * hs];
        const double2* hq2 = &h2[(size_t)q * hs];
        std::vector<cd>& Hq = Hm[q];
        auto Hat = [&](int r, int c) -> cd& { return Hq[(size_t)r * mm + c]; };
        for (int i = 0; i <= j; ++i) Hat(i, j) = cd(hq[i].x + hq2[i].x, hq[i].y + hq2[i].y);
        const double hn = hq[j + 1].x;
        Hat(j + 1, j) = cd(hn);
        inv[q] = hn > 0 ? 1.0 / hn : 0.0;
        for (int i = 0; i < j; ++i) {  // previous rotations on column j
          const cd t = cs[q][i] * Hat(i, j) + sn[q][i] * Hat(i + 1, j);
          Hat(i + 1, j) = -std::conj(sn[q][i]) * Hat(i, j) + cs[q][i] * Hat(i + 1, j);
          Hat(i, j) = t;
        }
        const cd a = Hat(j, j), bb = Hat(j + 1, j);
        const double den = std::sqrt(std::norm(a) + std::norm(bb));
        if (den == 0.0) {
          cs[q][j] = 1.0;
          sn[q][j] = 0.0;
        } else {
          cs[q][j] = std::abs(a) / den;
          const cd phase = std::abs(a) > 0 ? a / std::abs(a) : cd(1.0);
          sn[q][j] = phase * std::conj(bb) / den;
        }
        Hat(j, j) = cs[q][j] * a + sn[q][j] * bb;
        Hat(j + 1, j) = 0.0;
        g[q][j + 1] = -std::conj(sn[q][j]) * g[q][j];
        g[q][j] = cs[q][j] * g[q][j];
        iters[q] += 1;
        kused[q] = j + 1;
        if (!(std::abs(g[q][j + 1]) <= tol * bnorm[q]) && hn != 0.0) next.push_back(q);
      }
      // v_{j+1} = w / ||w|| for the systems still iterating
      BIE_HIP_CHECK(hipMemcpyAsync(scale.p, inv.data(), nsys * sizeof(double), hipMemcpyHostToDevice, s));
      if (!next.empty()) {
        upload_list(next);
        hipLaunchKernelGGL(k_scale, dim3(nb, (unsigned)next.size()), blk, 0, s, W.p, (int64_t)n,
                           V.p + (int64_t)(j + 1) * n, sv, n, sys.p, scale.p);
      }
      list.swap(next);
      steps = j + 1;
    }
    total += steps;
    // y_q from the triangular system, x_q += V_q y_q for every system that iterated
    std::vector<double2> yh((size_t)nsys * hs, make_double2(0.0, 0.0));
    std::vector<int> upd;
    for (int q = 0; q < nsys; ++q) {
      const int kq = kused[q];
      if (!active[q] || kq == 0) continue;
      upd.push_back(q);
      std::vector<cd> y(kq);
      for (int i = kq - 1; i >= 0; --i) {
        cd acc = g[q][i];
        for (int c = i + 1; c < kq; ++c) acc -= Hm[q][(size_t)i * mm + c] * y[c];
        y[i] = acc / Hm[q][(size_t)i * mm + i];
      }
      for (int i = 0; i < kq; ++i) yh[(size_t)q * hs + i] = make_double2(y[i].real(), y[i].imag());
    }
    if (!upd.empty()) {
      int kmax = 0;
      for (int q : upd) kmax = std::max(kmax, kused[q]);
      BIE_HIP_CHECK(hipMemcpyAsync(Y.p, yh.data(), yh.size() * sizeof(double2), hipMemcpyHostToDevice, s));
      upload_list(upd);
      hipLaunchKernelGGL(k_gs_update, dim3(nb, (unsigned)upd.size()), blk, 0, s, V.p, sv, n, kmax,
                         Y.p, hs, X, ldx, sys.p, 1.0);
      // r = b - A x and its norm for those systems
      apply(X, ldx, W.p, n, (int)upd.size());
      hipLaunchKernelGGL(k_residual, dim3(nb, (unsigned)upd.size()), blk, 0, s, B, ldb, W.p,
                         (int64_t)n, n, sys.p);
      hipLaunchKernelGGL(k_norm, dim3((unsigned)upd.size()), dim3(kRed), 0, s, W.p, (int64_t)n, n,
                         sys.p, H.p, hs, 0);
      BIE_HIP_CHECK(hipGetLastError());
      BIE_HIP_CHECK(hipMemcpyAsync(hbuf.data(), H.p, hbuf.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
      BIE_HIP_CHECK(hipStreamSynchronize(s));
      for (int q : upd) {
        beta[q] = hbuf[(size_t)q * hs].x;
        done[q] = !(beta[q] > tol * bnorm[q]);
      }
    }
    if (steps == 0) break;
  }
  for (int q = 0; q < nsys; ++q) resid[q] = bnorm[q] > 0 ? beta[q] / bnorm[q] : 0.0;
}

}  // namespace bie
