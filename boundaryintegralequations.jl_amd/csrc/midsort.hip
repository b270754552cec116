// Radix sort of 64-bit keys on the device (hipCUB), used once per mesh to order the
// mid-edge pairs of collocation by (element, target).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "bie_common.hpp"

namespace bie {

void sort_keys(unsigned long long* keys, int64_t n, int bits, hipStream_t s) {
  if (n <= 1) return;
  BIE_REQUIRE(n < (int64_t)1 << 31, BIE_EINVAL, "too many keys to sort");
  unsigned long long* out = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  BIE_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, keys, (int)n, 0,
                                                  bits, s));
  if (hipMalloc(&out, sizeof(unsigned long long) * n) != hipSuccess) {
    (void)hipGetLastError();
    throw Error{BIE_ENOMEM, "mid-edge sort buffer allocation failed"};
  }
  if (hipMalloc(&tmp, tmp_bytes) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(out);
    throw Error{BIE_ENOMEM, "mid-edge sort workspace allocation failed"};
  }
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, out, (int)n, 0, bits, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(keys, out, sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp);
  (void)hipFree(out);
  BIE_HIP_CHECK(e);
}

}  // namespace bie
