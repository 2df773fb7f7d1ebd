// lkc_noalloc.hip -- get_An_ZtSn_ww with its xcalc (An_ZtSn_ww.c:5-155, xcalc.c:41-246) on
// caller-owned device buffers, with every intermediate in one file-scope device array, so
// that the call allocates no device memory.  For callers that cannot allocate: code under
// HIP graph capture, serving loops that keep their buffers, harnesses that forbid it.
//
// The intermediates (the quantised integers, their bit planes, the active-n row, the
// covariate list, the min/max partials and scalars, the covariate-major result before it
// is transposed into the caller's row-major output) are carved out of `scratch`, whose
// size bounds the problem: calls needing more are rejected.  Each device holds its own
// copy of the array for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns for the current device (the array's name in host
// code is not a device address).  Every call on a device shares that device's copy, so
// calls that overlap on one device race; calls on one stream are ordered.
#include "lkc_kernels.h"
#include "lkc_stages.h"

#include <mutex>

namespace lkc {

static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

namespace {

std::mutex g_mu;
void* g_ptr[64] = {};     // the current device's copy of `scratch`, per device ordinal

inline size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// row0 holds the bits of nlist (MSB first within each byte) over prb bytes
__global__ void k_zero_words(uint32_t* p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0u;
}
__global__ void k_set_bits(uint32_t* row0, const int* nlist, int nn)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
    const int n = nlist[i];
    atomicOr(row0 + (n >> 5), 1u << (((n >> 3) & 3) * 8 + (7 - (n & 7))));
  }
}
__global__ void k_iota_small(int* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
// out[a*ns + s] = cm[s*nr + a]
__global__ void k_to_row_major(const double* cm, int nr, int ns, double* out)
{
  const long long tot = (long long)nr * ns;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(e / ns), s = (int)(e % ns);
    out[e] = cm[(long long)s * nr + a];
  }
}

inline unsigned grid_for(long long n) { long long g = (n + 255) / 256; return (unsigned)(g < 1 ? 1 : (g > 65535 ? 65535 : g)); }

}  // namespace

size_t an_ztsn_noalloc_bytes(int nr_total, int n_total, int ns, int B)
{
  const size_t prb = bsize(n_total);
  return up256((size_t)ns * n_total * sizeof(uint32_t))      // cz
       + up256((size_t)ns * B * prb)                         // planes
       + up256(prb)                                          // row0
       + up256((size_t)ns * sizeof(int))                     // slist
       + up256(16 * sizeof(double))                          // scal
       + up256(1024 * sizeof(double))                        // xcalc min/max partials
       + up256((size_t)ns * nr_total * sizeof(double));      // covariate-major result
}

hipError_t noalloc_scratch(void** ptr)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}

hipError_t an_ztsn_noalloc(const uint8_t* A, int nr_total, const int* alist, int na, const uint8_t* amask,
                           const double* L, const int* nlist, int nn, int n_total, int ns, int B, long long D_MLT,
                           double* out, hipStream_t st)
{
  if (ns == 0) return hipSuccess;   // an empty result
  void* base = nullptr;
  hipError_t e = noalloc_scratch(&base);
  if (e != hipSuccess) return e;
  const size_t prb = bsize(n_total);
  uint8_t* p = (uint8_t*)base;
  uint32_t* cz = (uint32_t*)p;      p += up256((size_t)ns * n_total * sizeof(uint32_t));
  uint8_t* planes = p;              p += up256((size_t)ns * B * prb);
  uint32_t* row0 = (uint32_t*)p;    p += up256(prb);
  int* slist = (int*)p;             p += up256((size_t)ns * sizeof(int));
  double* scal = (double*)p;        p += up256(16 * sizeof(double));
  double* work = (double*)p;        p += up256(1024 * sizeof(double));
  double* cm = (double*)p;
  // every piece of scratch this call reads is written by this call first (the previous
  // call's contents are never read): row0 and the result are cleared, cz is read only at
  // the listed n, the planes are written whole
  const long long nwd = (long long)(prb / 4);
  hipLaunchKernelGGL(k_zero_words, dim3(grid_for(nwd)), dim3(256), 0, st, row0, nwd);
  if (nn > 0) hipLaunchKernelGGL(k_set_bits, dim3(grid_for(nn)), dim3(256), 0, st, row0, nlist, nn);
  hipLaunchKernelGGL(k_iota_small, dim3((ns + 255) / 256), dim3(256), 0, st, slist, ns);
  fill(cm, (long long)ns * nr_total, 0.0, st);
  xcalc(L, ns, 1, nullptr, nullptr, nn, ns, nlist, B, D_MLT, cz, n_total, scal, work, st);
  cz_planes(cz, n_total, ns, n_total, row0, B, planes, prb, st);
  an_ztsn(Bits{ A, nr_total, n_total, prb }, alist, na, (const uint32_t*)amask, planes, prb, row0,
          ns, slist, B, scal, cm, nr_total, st);
  const long long tot = (long long)nr_total * ns;
  hipLaunchKernelGGL(k_to_row_major, dim3(grid_for(tot)), dim3(256), 0, st, cm, nr_total, ns, out);
  return hipGetLastError();
}

}  // namespace lkc
