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

#define AT_PROW 144   // LDS row of one k: 128 position bytes + 16 bytes of padding (no bank conflicts)

// Block: 32 rows k shared by its 4 waves; wave w owns the y sub-tile y0 + 32 w of a
// 128-y tile.  Per term and 128-column chunk of n, the block spreads the position
// bits pos_t = (A_k xnor T_t) & mask of its 32 k once into 0/1 bytes in LDS (double
// buffered, the next chunk's bits loaded meanwhile), and each wave spreads its own
// y bits (0/1) in registers: one y fragment feeds the NTG MFMAs of a step.
//     C = sum_n pos y01,   hhAY = sum_{n in mask} a t y (+-1) = 4 C - 2 |pos| - sum_{n in mask} y
// Lane (r, g) of step s feeds columns n0 + 64g + 16s .. +15.  The ordered y sums run
// sub-tile by sub-tile in y order, one thread per (k, t).
template <int NTG, int NQ>
__global__ __launch_bounds__(256, 2) void k_attyn_mfma(AttynArgs a, const long long* cy, const long long* hhay)
{
  __shared__ __attribute__((aligned(16))) int8_t Ps[2][NTG][32][AT_PROW];
  __shared__ int Dm[NQ][NTG][32][33];
  __shared__ int cpos[NQ][NTG][32];
  // per-y values of the sub-tile the chains walk (staged by all threads)
  __shared__ double s_ay[32], s_Dy[32], s_haY[NQ][NTG][32];
  __shared__ long long s_cy[NQ][32];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  const int t0 = blockIdx.y * NTG;
  const int nts = min(NTG, a.nt - t0);
  const int kb = blockIdx.x * 32;
  // staging role: k row sk of the tile, 16-column piece sp of the chunk
  const int sk = tid >> 3, sp = tid & 7;
  const int sk_orig = a.klist[min(kb + sk, a.nk - 1)];
  // chain role (epilogue): thread ce < 32 * nts walks (k = ce % 32, t = ce / 32)
  const int ck = tid & 31, ct = tid >> 5;
  const bool chain = ct < nts;
  double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0;
  for (int e = tid; e < NQ * NTG * 32; e += 256) (&cpos[0][0][0])[e] = 0;

  for (int y0 = 0; y0 < a.ny; y0 += 128) {
    const int yrow = a.ylist[min(y0 + 32 * wave + r, a.ny - 1)];
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
      const uint8_t* Arow = tm.A + (size_t)sk_orig * tm.rbA + 2 * sp;
      const uint8_t* Trow[NTG];
#pragma unroll
      for (int t = 0; t < NTG; t++) Trow[t] = tm.T + (size_t)a.tlist[t0 + min(t, nts - 1)] * tm.rbT + 2 * sp;
      const uint8_t* mrow = tm.mask + 2 * sp;
      const uint8_t* Yrow = tm.Y + (size_t)yrow * tm.rbY + 8 * g;
      const int nchunk = (int)(tm.ldn / 128);
      uint32_t sa = 0, sm = 0, stv[NTG];
      uint64_t yb = 0;
      int cnt[NTG];
#pragma unroll
      for (int t = 0; t < NTG; t++) { stv[t] = 0; cnt[t] = 0; }
#define AT_LOAD(c) do { const long long bo_ = 16LL * (c); sa = *(const uint16_t*)(Arow + bo_); \
        sm = *(const uint16_t*)(mrow + bo_); \
        _Pragma("unroll") for (int t_ = 0; t_ < NTG; t_++) stv[t_] = *(const uint16_t*)(Trow[t_] + bo_); \
        yb = *(const uint64_t*)(Yrow + bo_); } while (0)
#define AT_STAGE(buf) do { _Pragma("unroll") for (int t_ = 0; t_ < NTG; t_++) { \
          const uint32_t p_ = ~(sa ^ stv[t_]) & sm & 0xFFFFu; cnt[t_] += __builtin_popcount(p_); \
          const v4i f_ = spread16_cols(__builtin_bitreverse32(p_), 0); \
          *(v4i*)&Ps[buf][t_][sk][16 * sp] = f_; } } while (0)
      __syncthreads();   // the previous readers of Ps are done
      uint64_t ycur = 0;
      if (nchunk > 0) { AT_LOAD(0); AT_STAGE(0); ycur = yb; }
      __syncthreads();
      for (int c = 0; c < nchunk; c++) {
        const int buf = c & 1;
        const uint32_t yr0 = __builtin_bitreverse32((uint32_t)ycur), yr1 = __builtin_bitreverse32((uint32_t)(ycur >> 32));
        if (c + 1 < nchunk) AT_LOAD(c + 1);
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const v4i bf = spread16_cols(s < 2 ? yr0 : yr1, s & 1);
#pragma unroll
          for (int t = 0; t < NTG; t++)
            acc[q][t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(*(const v4i*)&Ps[buf][t][r][64 * g + 16 * s], bf,
                                                               acc[q][t], 0, 0, 0);
        }
        if (c + 1 < nchunk) { AT_STAGE(buf ^ 1); ycur = yb; }
        __syncthreads();
      }
#undef AT_LOAD
#undef AT_STAGE
      // |pos| per (k, t) of this term: the 8 pieces of a row add up (integers)
      if (y0 == 0) {
#pragma unroll
        for (int t = 0; t < NTG; t++) atomicAdd(&cpos[q][t][sk], cnt[t]);
      }
    }
    // ordered sums over the tile's y: sub-tile by sub-tile, in y order
    for (int w = 0; w < 4; w++) {
      const int ys = y0 + 32 * w;
      if (ys >= a.ny) break;
      __syncthreads();
      if (wave == w) {
#pragma unroll
        for (int q = 0; q < NQ; q++)
#pragma unroll
          for (int t = 0; t < NTG; t++)
#pragma unroll
            for (int v = 0; v < 16; v++) Dm[q][t][8 * (v / 4) + 4 * g + (v % 4)][r] = acc[q][t][v];
      }
      // the sub-tile's per-y inputs of the chain, loaded once by all threads
      {
        const double* alpha = a.ajdk + (size_t)1 * a.pcols;
        const double* Dt = a.ajdk + (size_t)5 * a.pcols;
        for (int e = tid; e < 32 * (1 + NQ + NQ * NTG); e += 256) {
          const int j = e & 31, f = e >> 5, yi = min(ys + j, a.ny - 1);
          const int yy = a.ylist[yi];
          if (f == 0) { s_ay[j] = alpha[yy / kPopBits]; s_Dy[j] = Dt[yy / kPopBits]; }
          else if (f <= NQ) s_cy[f - 1][j] = cy[(size_t)(f - 1) * a.ny + yi];
          else {
            const int qt = f - 1 - NQ, q = qt / NTG, t = qt % NTG;
            const AttynTerm& tm = a.term[q];
            const int tt = a.tlist[t0 + min(t, nts - 1)];
            s_haY[q][t][j] = (double)(long long)tm.haY[(long long)tt * tm.ldh + yy];
          }
        }
      }
      __syncthreads();
      if (chain) {
        const double* alpha = a.ajdk + (size_t)1 * a.pcols;
        const int kk = a.klist[min(kb + ck, a.nk - 1)];
        const double aa = alpha[kk / kPopBits];
        const int tt = a.tlist[t0 + ct];
        long long hAy[NQ], base[NQ], hhay_[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          hAy[q] = (long long)a.term[q].hAy[(long long)tt * a.term[q].ldh + kk];
          base[q] = 2LL * cpos[q][ct][ck];
          hhay_[q] = hhay[q * kMaxT + t0 + ct];
        }
        const int yn = min(32, a.ny - ys);
        for (int j = 0; j < yn; j++) {
          const double ay = s_ay[j], Dy = s_Dy[j];
          double L[NQ];
#pragma unroll
          for (int q = 0; q < NQ; q++) {
            const long long hhAY = 4LL * Dm[q][ct][ck][j] - base[q] - s_cy[q][j];
            // (aa * hhaY with the reference's integer hhaY converted to double)
            L[q] = (hhAY - hAy[q] * ay - aa * s_haY[q][ct][j] + aa * hhay_[q] * ay);
          }
          // attjn = L[0], ktrhn = L[1], ztsvn = L[2]
          if (NQ > 2) o0 += L[0] * Dy * L[NQ - 1];
          o1 += L[0] * Dy * L[0];
          o2 += L[0] * Dy * L[1];
          o3 += L[1] * Dy * L[1];
        }
      }
    }
  }
  if (chain && kb + ck < a.nk) {
    const long long oo = (long long)a.tlist[t0 + ct] * a.ldo + a.klist[kb + ck];
    if (NQ > 2) a.vtszn[oo] = o0;
    a.jttan[oo] = o1;
    a.htrkn[oo] = o2;
    a.hhtrkn[oo] = o3;
  }
}

size_t attyn_mfma_work_bytes(int ny, int nt)
{
  (void)nt;
  return sizeof(long long) * (3 * (size_t)(ny + 1) + 3 * kMaxT + 16) + 256;
}

hipError_t attyn_mfma(const AttynArgs& a, void* work, size_t work_bytes, hipStream_t st)
{
  if (a.nk <= 0 || a.nt <= 0) return hipSuccess;
  if (work_bytes < attyn_mfma_work_bytes(a.ny, a.nt)) return hipErrorInvalidValue;
  long long* cy = (long long*)work;
  long long* hhay = cy + 3 * (size_t)(a.ny + 1);
  hipLaunchKernelGGL(k_attyn_pre, dim3(a.ny + 1, a.nterm), dim3(64), 0, st, a, cy, hhay);
  const int ntg = a.nt >= 3 ? 3 : a.nt;
  dim3 grid(cdiv(a.nk, 32), cdiv(a.nt, ntg));
#define AT_LAUNCH(NTG, NQ) hipLaunchKernelGGL((k_attyn_mfma<NTG, NQ>), grid, dim3(256), 0, st, a, cy, hhay)
  if (a.nterm == 3) {
    if (ntg == 3) AT_LAUNCH(3, 3); else if (ntg == 2) AT_LAUNCH(2, 3); else AT_LAUNCH(1, 3);
  } else {
    if (ntg == 3) AT_LAUNCH(3, 2); else if (ntg == 2) AT_LAUNCH(2, 2); else AT_LAUNCH(1, 2);
  }
#undef AT_LAUNCH
  return hipGetLastError();
}

}  // namespace lkc
