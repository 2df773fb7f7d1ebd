// lkc_attyn_mfma.hip -- the QC stage-b terms of the low-rank update on the int8 matrix cores:
// get_AtTYn_vv (AtTYn_vv.c:79-120) for the three (A, T, Y, mask) triples of
// bcc_lrup_QC_YnWt_stage_b1 fused with the four get_AtTYn____WtSZn_vv sums of
// stage b2 (AtTYn____WtSZn_vv.c:60-80; bcc_lrup.c:422-590).
//
// For every (k, y, t) the reference needs the integer
//     hhAY = |mask| - 2 popcount(~(A_k ^ T_t ^ Y_y) & mask) = sum_{n in mask} a t y   (+-1 values)
// Here a t = 2 pos - 1 on the mask (pos = A xnor T), so
//     hhAY = 2 sum_n pos(n) y(n) - sum_{n in mask} y(n)
// with the first sum one int8 MFMA contraction (pos as 0/1, y as +-1).  The double
// epilogue follows the reference's expressions and sums over y in ascending order.
#include "lkc_kernels.h"
#include <stdio.h>

namespace lkc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
__device__ __forceinline__ int popc_u(uint32_t x) { return (int)__builtin_popcount(x); }
__device__ __forceinline__ long long wsum_l(long long v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// cy[q*ny + y] = sum_{n in mask_q} Y_y(n) (+-1);  block ny: hhay[q*kMaxT + t] = sum_{n in mask_q} T_t(n)
__global__ void k_attyn_pre(AttynArgs a, long long* cy, long long* hhay)
{
  const int q = blockIdx.y, y = blockIdx.x;
  const AttynTerm& tm = a.term[q];
  const int nwd = (int)(tm.ldn / 32);
  const uint32_t* m = (const uint32_t*)tm.mask;
  if (y < a.ny) {
    const uint32_t* Yr = (const uint32_t*)(tm.Y + (size_t)a.ylist[y] * tm.rbY);
    long long c = 0;
    for (int w = threadIdx.x; w < nwd; w += 64) c += popc_u(Yr[w] & m[w]);
    c = wsum_l(c);
    if (threadIdx.x == 0) cy[(size_t)q * a.ny + y] = 2 * c - tm.mcount;
  } else {
    for (int t = 0; t < a.nt; t++) {
      const uint32_t* Tr = (const uint32_t*)(tm.T + (size_t)a.tlist[t] * tm.rbT);
      long long c = 0;
      for (int w = threadIdx.x; w < nwd; w += 64) c += popc_u(Tr[w] & m[w]);
      c = wsum_l(c);
      if (threadIdx.x == 0) hhay[q * kMaxT + t] = -(long long)tm.mcount + 2 * c;
    }
  }
}

// byte x -> 8 bytes for the 8 elements it packs (element j = bit 7-j): 0/1, or -1/+1
__device__ __forceinline__ uint64_t expand01(uint32_t x)
{
  uint64_t r = 0;
  for (int j = 0; j < 8; j++) r |= (uint64_t)((x >> (7 - j)) & 1u) << (8 * j);
  return r;
}
__device__ __forceinline__ uint64_t expandpm(uint32_t x)
{
  uint64_t r = 0;
  for (int j = 0; j < 8; j++) r |= (uint64_t)(((x >> (7 - j)) & 1u) ? 0x01u : 0xFFu) << (8 * j);
  return r;
}
__device__ __forceinline__ v4i frag(const uint64_t* lut, uint64_t w, int s)
{
  const uint64_t e0 = lut[(w >> (16 * s)) & 0xFF], e1 = lut[(w >> (16 * s + 8)) & 0xFF];
  v4i f;
  f[0] = (int)(uint32_t)e0; f[1] = (int)(uint32_t)(e0 >> 32);
  f[2] = (int)(uint32_t)e1; f[3] = (int)(uint32_t)(e1 >> 32);
  return f;
}

template <int NTG, int NQ>
__global__ __launch_bounds__(256) void k_attyn_mfma(AttynArgs a, const long long* cy, const long long* hhay)
{
  __shared__ uint64_t lut01[256], lutpm[256];
  __shared__ int Gs[4][NQ][32][33];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  lut01[tid] = expand01(tid);
  lutpm[tid] = expandpm(tid);
  const int t0 = blockIdx.y * NTG;
  const int nts = min(NTG, a.nt - t0);
  const int kbase = blockIdx.x * 128 + wave * 32;
  const int krow = min(kbase + r, a.nk - 1);
  const int kr_orig = a.klist[krow];
  // per-row constants of the epilogue (lane < 32 owns row kbase + lane)
  const int myk = min(kbase + lane, a.nk - 1), kk = a.klist[myk];
  const double* alpha = a.ajdk + (size_t)1 * a.pcols;
  const double* Dt = a.ajdk + (size_t)5 * a.pcols;
  const double aa = alpha[kk / kPopBits];
  long long hAy[NQ][NTG];
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int t = 0; t < NTG; t++)
      hAy[q][t] = (long long)a.term[q].hAy[(long long)a.tlist[t0 + min(t, nts - 1)] * a.term[q].ldh + kk];
  double o[4][NTG];
#pragma unroll
  for (int p = 0; p < 4; p++)
#pragma unroll
    for (int t = 0; t < NTG; t++) o[p][t] = 0.0;
  __syncthreads();

  for (int y0 = 0; y0 < a.ny; y0 += 32) {
    const int yrow = min(y0 + r, a.ny - 1);
    const int yr_orig = a.ylist[yrow];
    v16i acc[NQ][NTG];
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
      for (int t = 0; t < NTG; t++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[q][t][v] = 0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const AttynTerm& tm = a.term[q];
      const uint8_t* Arow = tm.A + (size_t)kr_orig * tm.rbA;
      const uint8_t* Yrow = tm.Y + (size_t)yr_orig * tm.rbY;
      const uint8_t* Trow[NTG];
#pragma unroll
      for (int t = 0; t < NTG; t++) Trow[t] = tm.T + (size_t)a.tlist[t0 + min(t, nts - 1)] * tm.rbT;
      for (long long n0 = 0; n0 < tm.ldn; n0 += 128) {
        const long long bo = n0 / 8 + 8 * g;
        const uint64_t m8 = *(const uint64_t*)(tm.mask + bo);
        if (!__any(m8 != 0)) continue;
        const uint64_t a8 = *(const uint64_t*)(Arow + bo);
        const uint64_t y8 = *(const uint64_t*)(Yrow + bo);
        uint64_t pos8[NTG];
#pragma unroll
        for (int t = 0; t < NTG; t++) pos8[t] = ~(a8 ^ *(const uint64_t*)(Trow[t] + bo)) & m8;
#pragma unroll 2
        for (int s = 0; s < 4; s++) {
          const v4i bf = frag(lutpm, y8, s);
#pragma unroll
          for (int t = 0; t < NTG; t++)
            acc[q][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(frag(lut01, pos8[t], s), bf, acc[q][t], 0, 0, 0);
        }
      }
    }
    // epilogue: lane < 32 walks its row's y in ascending order
    const int yn = min(32, a.ny - y0);
#pragma unroll
    for (int t = 0; t < NTG; t++) {
      if (t >= nts) break;
#pragma unroll
      for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int v = 0; v < 16; v++) Gs[wave][q][8 * (v / 4) + 4 * g + (v % 4)][r] = acc[q][t][v];
      __syncthreads();
      if (lane < 32) {
        const int tt = a.tlist[t0 + t];
        double L[NQ];
        double p0 = o[0][t], p1 = o[1][t], p2 = o[2][t], p3 = o[3][t];
        for (int j = 0; j < yn; j++) {
          const int yy = a.ylist[y0 + j];
          const double ay = alpha[yy / kPopBits], Dy = Dt[yy / kPopBits];
#pragma unroll
          for (int q = 0; q < NQ; q++) {
            const long long hhAY = 2LL * Gs[wave][q][lane][j] - cy[(size_t)q * a.ny + y0 + j];
            const long long hhaY = (long long)a.term[q].haY[(long long)tt * a.term[q].ldh + yy];
            const long long hhay_ = hhay[q * kMaxT + t0 + t];
            L[q] = (hhAY - hAy[q][t] * ay - aa * hhaY + aa * hhay_ * ay);
          }
          // attjn = L[0], ktrhn = L[1], ztsvn = L[2]
          if (NQ > 2) p0 += L[0] * Dy * L[NQ - 1];
          p1 += L[0] * Dy * L[0];
          p2 += L[0] * Dy * L[1];
          p3 += L[1] * Dy * L[1];
        }
        o[0][t] = p0; o[1][t] = p1; o[2][t] = p2; o[3][t] = p3;
      }
      __syncthreads();
    }
  }
  if (lane < 32 && kbase + lane < a.nk) {
#pragma unroll
    for (int t = 0; t < NTG; t++) {
      if (t >= nts) break;
      const long long oo = (long long)a.tlist[t0 + t] * a.ldo + a.klist[kbase + lane];
      if (NQ > 2) a.vtszn[oo] = o[0][t];
      a.jttan[oo] = o[1][t];
      a.htrkn[oo] = o[2][t];
      a.hhtrkn[oo] = o[3][t];
    }
  }
}

size_t attyn_mfma_work_bytes(int ny, int nt)
{
  (void)nt;
  return sizeof(long long) * (3 * (size_t)(ny + 1) + 3 * kMaxT + 16) + 256;
}

void attyn_mfma(const AttynArgs& a, void* work, size_t work_bytes, hipStream_t st)
{
  if (a.nk <= 0 || a.nt <= 0) return;
  if (work_bytes < attyn_mfma_work_bytes(a.ny, a.nt)) { fprintf(stderr, "lkc: attyn work buffer too small\n"); return; }
  long long* cy = (long long*)work;
  long long* hhay = cy + 3 * (size_t)(a.ny + 1);
  hipLaunchKernelGGL(k_attyn_pre, dim3(a.ny + 1, a.nterm), dim3(64), 0, st, a, cy, hhay);
  const int ntg = a.nt >= 3 ? 3 : a.nt;
  dim3 grid(cdiv(a.nk, 128), cdiv(a.nt, ntg));
#define AT_LAUNCH(NTG, NQ) hipLaunchKernelGGL((k_attyn_mfma<NTG, NQ>), grid, dim3(256), 0, st, a, cy, hhay)
  if (a.nterm == 3) {
    if (ntg == 3) AT_LAUNCH(3, 3); else if (ntg == 2) AT_LAUNCH(2, 3); else AT_LAUNCH(1, 3);
  } else {
    if (ntg == 3) AT_LAUNCH(3, 2); else if (ntg == 2) AT_LAUNCH(2, 2); else AT_LAUNCH(1, 2);
  }
#undef AT_LAUNCH
}

}  // namespace lkc
