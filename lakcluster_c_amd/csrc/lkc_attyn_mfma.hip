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

// 16 packed columns -> 16 bytes 0/1 in column order.  wr is the bit-reversed 32-bit
// word of the packed row (bfrev: column 8b + j, MSB-first in byte b, lands on bit
// 8(3-b) + j); half h selects columns 16h .. 16h+15.  Each nibble becomes four bytes
// with one multiply: (n * 0x204081) & 0x01010101 puts bit j of n into byte j.
__device__ __forceinline__ v4i spread16_cols(uint32_t wr, int h)
{
  v4i f;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int sh = 24 - 16 * h - 8 * (d >> 1) + 4 * (d & 1);
    f[d] = (int)((((wr >> sh) & 15u) * 0x204081u) & 0x01010101u);
  }
  return f;
}
// bytes 0/1 -> -1/+1 (0xFF / 0x01): ~(b * 0xFE) per byte, no carries between bytes
__device__ __forceinline__ uint32_t pm_bytes(uint32_t b) { return ~((b << 8) - (b << 1)); }

#define AT_YROW 144   // LDS row of one y: 128 bytes of +-1 values + 16 bytes of padding

// Block: 4 waves x 32 rows k, looping over the y in tiles of 32.  Per term and
// 128-column chunk of n, the tile's y rows are spread once into +-1 bytes in LDS
// (double buffered; the next chunk's bits are loaded while the current one is used);
// each wave spreads its pos = (A xnor T) & mask bits in registers.  Lane (r, g) of
// step s feeds columns n0 + 64g + 16s .. +15.
template <int NTG, int NQ>
__global__ __launch_bounds__(256, 2) void k_attyn_mfma(AttynArgs a, const long long* cy, const long long* hhay)
{
  __shared__ __attribute__((aligned(16))) int8_t Ys[2][32][AT_YROW];
  __shared__ int Gs[4][NQ][32][33];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  const int t0 = blockIdx.y * NTG;
  const int nts = min(NTG, a.nt - t0);
  const int kbase = blockIdx.x * 128 + wave * 32;
  const int krow = min(kbase + r, a.nk - 1);
  const int kr_orig = a.klist[krow];
  // y loader role: row ly of the tile, columns 16 lp .. 16 lp + 15 of the chunk
  const int ly = tid >> 3, lp = tid & 7;
  double o[4][NTG];
#pragma unroll
  for (int p = 0; p < 4; p++)
#pragma unroll
    for (int t = 0; t < NTG; t++) o[p][t] = 0.0;

  for (int y0 = 0; y0 < a.ny; y0 += 32) {
    const int yl_orig = a.ylist[min(y0 + ly, a.ny - 1)];
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
      const uint8_t* Arow = tm.A + (size_t)kr_orig * tm.rbA + 8 * g;
      const uint8_t* Trow[NTG];
#pragma unroll
      for (int t = 0; t < NTG; t++) Trow[t] = tm.T + (size_t)a.tlist[t0 + min(t, nts - 1)] * tm.rbT + 8 * g;
      const uint8_t* mrow = tm.mask + 8 * g;
      const uint8_t* Ysrc = tm.Y + (size_t)yl_orig * tm.rbY + 2 * lp;
      const int nchunk = (int)(tm.ldn / 128);
      uint64_t an = 0, tn[NTG], mk = 0;
      uint32_t yb = 0;
#pragma unroll
      for (int t = 0; t < NTG; t++) tn[t] = 0;
#define AT_LOAD(c) do { const long long bo_ = 16LL * (c); an = *(const uint64_t*)(Arow + bo_); \
        for (int t_ = 0; t_ < NTG; t_++) tn[t_] = *(const uint64_t*)(Trow[t_] + bo_); \
        mk = *(const uint64_t*)(mrow + bo_); yb = *(const uint16_t*)(Ysrc + bo_); } while (0)
#define AT_YSTORE(buf) do { const uint32_t wr_ = __builtin_bitreverse32(yb); \
        const v4i f_ = spread16_cols(wr_, 0); uint4 u_; u_.x = pm_bytes((uint32_t)f_[0]); u_.y = pm_bytes((uint32_t)f_[1]); \
        u_.z = pm_bytes((uint32_t)f_[2]); u_.w = pm_bytes((uint32_t)f_[3]); *(uint4*)&Ys[buf][ly][16 * lp] = u_; } while (0)
      __syncthreads();   // the previous readers of Ys are done
      if (nchunk > 0) { AT_LOAD(0); AT_YSTORE(0); }
      __syncthreads();
      for (int c = 0; c < nchunk; c++) {
        const int buf = c & 1;
        uint32_t prv[NTG][2];
#pragma unroll
        for (int t = 0; t < NTG; t++) {
          const uint64_t pos8 = ~(an ^ tn[t]) & mk;
          prv[t][0] = __builtin_bitreverse32((uint32_t)pos8);
          prv[t][1] = __builtin_bitreverse32((uint32_t)(pos8 >> 32));
        }
        if (c + 1 < nchunk) AT_LOAD(c + 1);
#pragma unroll 2
        for (int s = 0; s < 4; s++) {
          const v4i bf = *(const v4i*)&Ys[buf][r][64 * g + 16 * s];
#pragma unroll
          for (int t = 0; t < NTG; t++)
            acc[q][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(spread16_cols(prv[t][s >> 1], s & 1), bf, acc[q][t], 0, 0, 0);
        }
        if (c + 1 < nchunk) AT_YSTORE(buf ^ 1);
        __syncthreads();
      }
#undef AT_LOAD
#undef AT_YSTORE
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
        // per-row constants (recomputed here to keep them out of the main loop's registers)
        const double* alpha = a.ajdk + (size_t)1 * a.pcols;
        const double* Dt = a.ajdk + (size_t)5 * a.pcols;
        const int kk = a.klist[min(kbase + lane, a.nk - 1)];
        const double aa = alpha[kk / kPopBits];
        const int tt = a.tlist[t0 + t];
        long long hAy[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) hAy[q] = (long long)a.term[q].hAy[(long long)tt * a.term[q].ldh + kk];
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
            L[q] = (hhAY - hAy[q] * ay - aa * hhaY + aa * hhay_ * ay);
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
