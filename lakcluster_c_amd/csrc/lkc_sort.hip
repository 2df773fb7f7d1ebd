// lkc_sort.hip -- ascending sort of the score columns on the device (bcc_sumscores_xij,
// bcc_sumscores.c:225-322).  A stable radix sort of (score, position) pairs yields the
// ascending value sequence, which is unique; where the values in the prefix that the
// elimination drops are strictly increasing, it also equals the reference quicksort's
// permutation there (dQuickSort_xij, Quicksort.c:83-141).  The host checks that and
// falls back to the reference's quicksort otherwise (ties or NaN).
#include "lkc_kernels.h"
#include <rocprim/device/device_radix_sort.hpp>

namespace lkc {

__global__ void k_iota(int* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

size_t sort_pairs_temp_bytes(int n)
{
  size_t b = 0;
  if (n <= 0) return 256;
  if (rocprim::radix_sort_pairs(nullptr, b, (const double*)nullptr, (double*)nullptr, (const int*)nullptr,
                                (int*)nullptr, (unsigned int)n) != hipSuccess) return 0;
  return b + 256;
}

hipError_t sort_pairs(const double* keys, double* keys_out, int* idx_in, int* idx_out, int n,
                      void* temp, size_t temp_bytes, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, st, idx_in, n);
  size_t b = temp_bytes;
  return rocprim::radix_sort_pairs(temp, b, keys, keys_out, (const int*)idx_in, idx_out, (unsigned int)n, 0u, 64u, st);
}

}  // namespace lkc
