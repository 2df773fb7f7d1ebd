// lkc_ww_mfma.hip -- At_T_YnWt_S_Zn_ww (At_T_YnWt_S_Zn_ww.c:44-110) on the int8 matrix cores.
//
// The reference forms, for every (k, t, w), the exact integer
//     G = sum_{n in mask} 2^b pmpm0(P_k, T_t, plane_b(w))  =  sum_{n in mask} s_kt(n) cz(w, n)
// with s = +1 where P_k(n) == T_t(n) and -1 otherwise, and cz the B-bit quantised values.
// Here s = 2 pos - 1 on the mask (pos = P xnor T), and cz is split into signed bytes
// d_i = byte_i(cz) - 128, so that
//     sum pos*cz = sum_i 256^i (sum_n pos d_i) + 128 (sum_i 256^i) |pos|
//     G          = 2 sum pos*cz - sum_{n in mask} cz(w, n)
// where every inner sum is one int8 x int8 -> int32 MFMA contraction over n.  The
// integer is then turned into the reference's double term and summed over w in
// the reference's order (ascending w), per (k, t).
#include "lkc_kernels.h"

#include <vector>
#include <stdio.h>
#include <algorithm>

namespace lkc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
__device__ __forceinline__ int popc32(uint32_t x) { return (int)__builtin_popcount(x); }
__device__ __forceinline__ int bitv(const uint8_t* row, int i) { return (row[i >> 3] >> (7 - (i & 7))) & 1; }
__device__ __forceinline__ long long wsum64(long long v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// signed digits of cz: Dg[(i*nwp + j)*ldn + n] = byte_i(cz[j*ldc + n]) - 128 for n in the mask, else 0
__global__ void k_ww_digits(const uint32_t* cz, long long ldc, int nw, int nwp, int n_total, const uint8_t* mask,
                            int ndig, long long ldn, int8_t* Dg)
{
  long long per = ldn / 4;
  long long tot = (long long)nwp * per;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int j = (int)(e / per);
    long long n = (e % per) * 4;
    uint32_t d[4] = { 0, 0, 0, 0 };
    if (j < nw) {
      for (int q = 0; q < 4; q++) {
        long long m = n + q;
        if (m < n_total && bitv(mask, (int)m)) {
          uint32_t v = cz[(long long)j * ldc + m];
          for (int i = 0; i < ndig; i++) d[i] |= (((v >> (8 * i)) & 0xFFu) ^ 0x80u) << (8 * q);
        }
      }
    }
    for (int i = 0; i < ndig; i++) *(uint32_t*)(Dg + ((long long)i * nwp + j) * ldn + n) = d[i];
  }
}

// The per-w and per-k inputs of the contraction, in one launch: blocks 0 .. nw - 1 the
// per-w sums, block nw the covariate sums, then four rows k per block (a wave each).
//   per w: Gm[w] = sum_{n in mask} cz(w,n);  H[t][w] = (double) sum_{n in mask} T_t(n)+-1 cz(w,n)
//   block nw: bp[t] = (double)(sum_{n in mask} T_t(n)+-1) * min * mlt
//   per k: cntp[t][k] = |mask & (P_k xnor T_t)|, llk[t][k] = pmpm0(P_k, T_t, mask, mask)
__global__ __launch_bounds__(256) void k_ww_pre(WwArgs a, double* H, double* bp, long long* Gm, int* cntp, int* llk)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwd = (int)(a.prb / 4);       // 32-column words of the n extent
  if ((int)blockIdx.x > a.nw) {
    const int k = ((int)blockIdx.x - a.nw - 1) * 4 + wave;
    if (k >= a.nk) return;
    const uint32_t* Pr = (const uint32_t*)(a.P.p + (size_t)a.klist[k] * a.P.rb);
    for (int t = 0; t < a.nt; t++) {
      const uint32_t* Tr = (const uint32_t*)(a.Tc.p + (size_t)a.tlist[t] * a.Tc.rb);
      long long p = 0, q = 0;
      for (int w = lane; w < nwd; w += 64) {
        uint32_t e = Pr[w] ^ Tr[w], m = a.pm[w];
        p += popc32(~e & m); q += popc32(e & m);
      }
      p = wsum64(p); q = wsum64(q);
      if (lane == 0) { cntp[(size_t)t * a.nk + k] = (int)p; llk[(size_t)t * a.nk + k] = (int)(p - q); }
    }
    return;
  }
  __shared__ long long red[4][kMaxT + 1];
  const int j = blockIdx.x;
  const uint32_t* pm = a.pm;
  long long g = 0, h[kMaxT];
#pragma unroll
  for (int t = 0; t < kMaxT; t++) h[t] = 0;
  const uint32_t* Tw[kMaxT];
#pragma unroll
  for (int t = 0; t < kMaxT; t++) Tw[t] = (const uint32_t*)(a.Tc.p + (size_t)a.tlist[min(t, a.nt - 1)] * a.Tc.rb);
  // thread per word: the mask word and the covariate words are loaded once, the 32
  // values of the word in eight 16-byte loads (words of an all-zero mask are skipped)
  for (int w = threadIdx.x; w < nwd; w += 256) {
    const uint32_t m = pm[w];
    if (!m) continue;
    uint32_t tw[kMaxT];
#pragma unroll
    for (int t = 0; t < kMaxT; t++) tw[t] = Tw[t][w];
#pragma unroll 8
    for (int q = 0; q < 32; q++) {
      // bit of column 32 w + q: byte q / 8 of the word, MSB first
      const int sh = 8 * (q >> 3) + 7 - (q & 7);
      if (!((m >> sh) & 1u)) continue;
      const long long v = (j < a.nw) ? (long long)a.cz[(long long)j * a.ldc + 32 * w + q] : 1;
      g += v;
#pragma unroll
      for (int t = 0; t < kMaxT; t++)
        if (t < a.nt) h[t] += ((tw[t] >> sh) & 1u) ? v : -v;
    }
  }
  // integer sums: the order of the partial sums does not matter
  g = wsum64(g);
#pragma unroll
  for (int t = 0; t < kMaxT; t++) if (t < a.nt) h[t] = wsum64(h[t]);
  if (lane == 0) {
    red[wave][kMaxT] = g;
#pragma unroll
    for (int t = 0; t < kMaxT; t++) red[wave][t] = h[t];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  g = red[0][kMaxT] + red[1][kMaxT] + red[2][kMaxT] + red[3][kMaxT];
  for (int t = 0; t < a.nt; t++) h[t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
  if (j < a.nw) {
    Gm[j] = g;
    for (int t = 0; t < a.nt; t++) H[(size_t)t * a.nw + j] = (double)h[t];
  } else {
    for (int t = 0; t < a.nt; t++) { double d = h[t] * 1.0; bp[t] = d * a.scal[2] * a.scal[4]; }
  }
}

// 16 packed columns -> 16 bytes 0/1 in column order.  wr is the bit-reversed 32-bit
// word of the packed row (bfrev: column 8b + j of the word, MSB-first in byte b, lands
// on bit 8(3-b) + j); half h selects columns 16h .. 16h+15.  Each nibble becomes four
// bytes with one multiply: (n * 0x204081) & 0x01010101 puts bit j of n into byte j.
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

#define WW_BROW 144   // LDS row of one w: 128 digit bytes + 16 bytes of padding (no bank conflicts)

// the compiler keeps LDS accesses of one wave in program order across this point; the
// LDS operations of a wave are performed in order, so lanes of the same wave then see
// each other's writes
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block: 4 waves x 32 rows k; the 32 w of a step are shared, their digits staged in LDS
// per 128-column chunk of n (double buffered, the next chunk loaded while the current
// one is used).  Lane (r, g) of a step s feeds columns n0 + 64g + 16s .. +15: the
// packed bits of P, T and the mask come as one 8-byte word per lane and chunk.
//
// After the last chunk of a w step, each lane holds the integer sums of 16 rows k at one
// w.  Every lane turns its 16 into the reference's double terms (the per-w inputs -- Gm,
// H, the S signs -- and the per-k ones -- alpha, the counts, the Q signs -- come from LDS,
// staged once per block or per w step), writes them to LDS, and lane k of each wave then
// only adds its row's terms in ascending w, as the reference's loop over w does.
//
// MODE 0 is that fused form.  For grids too small to fill the device (few rows k), the
// w steps, and parts of the n chunks, spread over blocks (blockIdx.z = w step * nsplit +
// part) and the terms go through a buffer Pbuf[(t * nwp + w) * nk + k]: MODE 1 (one part)
// forms the terms in the block and stores them; MODE 2 (several parts) adds its integer
// sums there with 64-bit atomics (integer sums in any order are exact) and k_ww_terms
// forms the terms in place.  k_ww_sum then adds each row's terms in ascending w.
template <int NTG, int ND, int MODE>
__global__ __launch_bounds__(256, NTG == 1 ? 2 : 1) void k_ww_mfma(WwArgs a, const int8_t* Dg, long long ldn, int nwp,
                                                 const double* X, const double* bp, const long long* Gm,
                                                 const int* cntp, const int* llk, int nsplit,
                                                 unsigned long long* Pbuf)
{
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][ND][32][WW_BROW];
  __shared__ double Tm[4][32][33];                    // terms [wave][k][w]
  __shared__ double KcQ[128];                         // alpha of q's column block
  __shared__ int KcB[128];                            // k's column block
  __shared__ int WvB[4];                              // the wave's rows' one column block, or -1
  __shared__ int Kcnt[NTG][128], Kllk[NTG][128];      // cntp, llk of the block's rows k
  __shared__ double Cs[2 + NTG];                      // min, mlt, bp[t]
  __shared__ long long WGm[32];                       // the w step's Gm, X, S signs, w indices
  __shared__ double WX[4][NTG][32];
  __shared__ int WvS[NTG][32];
  __shared__ int Wws[32];
  __shared__ uint32_t Zb[128];                        // bit j: Q(q_k, w0 + j) == +1
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  const int t0 = blockIdx.y * NTG;
  const int nts = min(NTG, a.nt - t0);
  const int kbase = blockIdx.x * 128 + wave * 32;
  const int krow = min(kbase + r, a.nk - 1);
  const uint8_t* Prow = a.P.p + (size_t)a.klist[krow] * a.P.rb + 8 * g;
  const uint8_t* pm = (const uint8_t*)a.pm + 8 * g;
  const uint8_t* Trow[NTG];
#pragma unroll
  for (int t = 0; t < NTG; t++) Trow[t] = a.Tc.p + (size_t)a.tlist[t0 + min(t, nts - 1)] * a.Tc.rb + 8 * g;
  constexpr bool SPLIT = MODE != 0;
  // per-block constants (visible after the first barrier of the w loop)
  if (MODE == 2) {
  } else if (tid < 128) {
    const int myk = min(blockIdx.x * 128 + tid, a.nk - 1);
    const double* alpha = a.ajdk + (size_t)1 * a.pcols;
    KcB[tid] = a.klist[myk] / kPopBits;
    KcQ[tid] = alpha[a.qlist[myk] / kPopBits];
#pragma unroll
    for (int t = 0; t < NTG; t++) {
      const int tt = t0 + min(t, nts - 1);
      Kcnt[t][tid] = cntp[(size_t)tt * a.nk + myk];
      Kllk[t][tid] = llk[(size_t)tt * a.nk + myk];
    }
  } else if (tid == 128) {
    Cs[0] = a.scal[2]; Cs[1] = a.scal[4];
#pragma unroll
    for (int t = 0; t < NTG; t++) Cs[2 + t] = bp[t0 + min(t, nts - 1)];
  }
  __syncthreads();
  // a wave whose 32 rows share one column block reads the per-(block, covariate, w) part of
  // its terms once per w (staged below); otherwise each row reads its own
  if (MODE != 2 && tid < 4) {
    const int b0 = KcB[32 * tid];
    bool one = true;
    for (int j = 1; j < 32; j++) one = one && KcB[32 * tid + j] == b0;
    WvB[tid] = one ? b0 : -1;
  }
  // ordered sums over w of row kbase + lane (lane < 32), per covariate
  double o[NTG];
#pragma unroll
  for (int t = 0; t < NTG; t++) o[t] = 0.0;
  const int nchunk_all = (int)(ldn / 128);
  int c_lo = 0, c_hi = nchunk_all, w_first = 0, w_end = a.nw;
  if (SPLIT) {
    const int ws = blockIdx.z / nsplit, sp = blockIdx.z % nsplit;
    c_lo = (int)((long long)sp * nchunk_all / nsplit);
    c_hi = (int)((long long)(sp + 1) * nchunk_all / nsplit);
    w_first = 32 * ws; w_end = min(a.nw, w_first + 32);
  }
  // digit loader: piece e = tid + 256 j -> digit plane j, w row tid / 8, 16 bytes (tid % 8)
  const int lw = tid >> 3, lp = tid & 7;
  long long S = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) S += 1LL << (8 * i);

  for (int w0 = w_first; w0 < w_end; w0 += 32) {
    const int wn = min(32, a.nw - w0);
    v16i acc[NTG][ND];
#pragma unroll
    for (int t = 0; t < NTG; t++)
#pragma unroll
      for (int i = 0; i < ND; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[t][i][v] = 0;
    const int8_t* Bsrc = Dg + (long long)(w0 + lw) * ldn + 16 * lp;
    const long long bplane = (long long)nwp * ldn;
    uint4 b0 = make_uint4(0, 0, 0, 0), b1 = b0, b2 = b0, b3 = b0;
    uint64_t pn = 0, tn0 = 0, tn1 = 0, tn2 = 0, mk = 0;
#define WW_BLOAD(c) do { const int8_t* q_ = Bsrc + 128LL * (c); b0 = *(const uint4*)q_; \
      if (ND > 1) b1 = *(const uint4*)(q_ + bplane); if (ND > 2) b2 = *(const uint4*)(q_ + 2 * bplane); \
      if (ND > 3) b3 = *(const uint4*)(q_ + 3 * bplane); } while (0)
#define WW_BSTORE(buf) do { *(uint4*)&Bs[buf][0][lw][16 * lp] = b0; if (ND > 1) *(uint4*)&Bs[buf][ND > 1 ? 1 : 0][lw][16 * lp] = b1; \
      if (ND > 2) *(uint4*)&Bs[buf][ND > 2 ? 2 : 0][lw][16 * lp] = b2; if (ND > 3) *(uint4*)&Bs[buf][ND > 3 ? 3 : 0][lw][16 * lp] = b3; } while (0)
#define WW_ALOAD(c) do { const long long bo_ = 16LL * (c); pn = *(const uint64_t*)(Prow + bo_); \
      tn0 = *(const uint64_t*)(Trow[0] + bo_); if (NTG > 1) tn1 = *(const uint64_t*)(Trow[NTG > 1 ? 1 : 0] + bo_); \
      if (NTG > 2) tn2 = *(const uint64_t*)(Trow[NTG > 2 ? 2 : 0] + bo_); mk = *(const uint64_t*)(pm + bo_); } while (0)
    __syncthreads();   // the previous w step's readers are done with Bs and the staged inputs
    // the step's per-w inputs
    if (MODE != 2 && tid < 32) {
      const int w = w0 + min(tid, wn - 1);
      const int ww = a.wlist[w];
      Wws[tid] = ww;
      WGm[tid] = Gm[w];
#pragma unroll
      for (int t = 0; t < NTG; t++) {
        const int tt = t0 + min(t, nts - 1);
        WvS[t][tid] = bitv(a.Sc.p + (size_t)a.slist[tt] * a.Sc.rb, ww) ? 1 : -1;
      }
    }
    if (MODE != 2 && tid < 128) {
      const int wv = tid >> 5, j = tid & 31, b = WvB[wv];
      if (b >= 0) {
        const int w = w0 + min(j, wn - 1);
#pragma unroll
        for (int t = 0; t < NTG; t++) WX[wv][t][j] = X[((size_t)b * a.nt + t0 + min(t, nts - 1)) * a.nw + w];
      }
    }
    if (c_hi > c_lo) { WW_BLOAD(c_lo); WW_ALOAD(c_lo); WW_BSTORE(0); }
    __syncthreads();
    for (int c = c_lo; c < c_hi; c++) {
      const int buf = (c - c_lo) & 1;
      uint64_t pos8[NTG];
      pos8[0] = ~(pn ^ tn0) & mk;
      if (NTG > 1) pos8[NTG > 1 ? 1 : 0] = ~(pn ^ tn1) & mk;
      if (NTG > 2) pos8[NTG > 2 ? 2 : 0] = ~(pn ^ tn2) & mk;
      uint32_t prv[NTG][2];
#pragma unroll
      for (int t = 0; t < NTG; t++) {
        prv[t][0] = __builtin_bitreverse32((uint32_t)pos8[t]);
        prv[t][1] = __builtin_bitreverse32((uint32_t)(pos8[t] >> 32));
      }
      // the next chunk (the last one again at the end: no branch in the loop body, so the
      // scheduler can interleave the spreads and LDS reads with the MFMAs of the whole chunk)
      const int cn = min(c + 1, c_hi - 1);
      WW_BLOAD(cn); WW_ALOAD(cn);
      __builtin_amdgcn_iglp_opt(0);
#pragma unroll
      for (int s = 0; s < 4; s++) {
        v4i bf[ND];
#pragma unroll
        for (int i = 0; i < ND; i++) bf[i] = *(const v4i*)&Bs[buf][i][r][64 * g + 16 * s];
#pragma unroll
        for (int t = 0; t < NTG; t++) {
          const v4i af = spread16_cols(prv[t][s >> 1], s & 1);
#pragma unroll
          for (int i = 0; i < ND; i++) acc[t][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[i], acc[t][i], 0, 0, 0);
        }
      }
      WW_BSTORE(buf ^ 1);
      __syncthreads();
    }
#undef WW_BLOAD
#undef WW_BSTORE
#undef WW_ALOAD
    if (MODE == 2) {
      // this block's part of the integer sums of rows kbase .. kbase + 31 at w = w0 + r
      if (w0 + r < a.nw) {
#pragma unroll
        for (int t = 0; t < NTG; t++) {
          if (t >= nts) break;
#pragma unroll
          for (int v = 0; v < 16; v++) {
            const int k = kbase + 8 * (v / 4) + 4 * g + (v % 4);
            long long spc = 0;
#pragma unroll
            for (int i = ND - 1; i >= 0; i--) spc = spc * 256 + (long long)acc[t][i][v];
            if (k < a.nk) atomicAdd(Pbuf + ((size_t)(t0 + t) * nwp + w0 + r) * a.nk + k, (unsigned long long)spc);
          }
        }
      }
      continue;
    }
    // the Q signs of the block's 128 rows k at the step's w: two threads per k, 16 w each
    {
      const int kq = tid >> 1, half = tid & 1;
      const int myk = min(blockIdx.x * 128 + kq, a.nk - 1);
      const uint8_t* Qrow = a.Q.p + (size_t)a.qlist[myk] * a.Q.rb;
      uint32_t bits = 0;
#pragma unroll 4
      for (int jj = 0; jj < 16; jj++) {
        const int j = 16 * half + jj;
        if (j < wn && bitv(Qrow, Wws[j])) bits |= 1u << j;
      }
      bits |= __shfl_xor(bits, 1, 64);
      if (!half) Zb[kq] = bits;
    }
    __syncthreads();
    // terms: lane (g, r) holds rows k = 8 (v / 4) + 4 g + v % 4 of its wave at w = w0 + r
#pragma unroll
    for (int t = 0; t < NTG; t++) {
      if (t >= nts) break;
      const double mn = Cs[0], mlt = Cs[1];
      const int tt = t0 + min(t, nts - 1);
      if (r < wn) {
        const long long gm = WGm[r];
        const double vS = (double)WvS[t][r];
        const int bw = WvB[wave];
        const double xw = bw >= 0 ? WX[wave][t][r] : 0.0;
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const int kw = 8 * (v / 4) + 4 * g + (v % 4);        // row within the wave
          const int kl = 32 * wave + kw;                        // row within the block
          long long spc = 0;
#pragma unroll
          for (int i = ND - 1; i >= 0; i--) spc = spc * 256 + (long long)acc[t][i][v];
          const long long G = 2 * (spc + 128 * S * (long long)Kcnt[t][kl]) - gm;
          const double alpha_q = KcQ[kl];
          const double bkk = (double)Kllk[t][kl] * mn * mlt;
          const double gd = (double)G;
          // X = (bp alpha_k + H alpha_k) / mlt of the row's column block, covariate and w
          const double xk = bw >= 0 ? xw : X[((size_t)KcB[kl] * a.nt + tt) * a.nw + w0 + r];
          const double tmp = (double)(bkk + gd) / mlt - xk;
          const int vZ = ((Zb[kl] >> r) & 1u) ? 1 : -1;
          const double term = tmp * vS * (vZ - alpha_q);
          if (MODE == 1) {
            const int k = kbase + kw;
            if (k < a.nk) ((double*)Pbuf)[((size_t)(t0 + t) * nwp + w0 + r) * a.nk + k] = term;
          } else {
            Tm[wave][kw][r] = term;
          }
        }
      }
      if (MODE == 1) continue;
      wave_lds_sync();
      if (lane < 32) {
        double acc_o = o[t];
        for (int j = 0; j < wn; j++) acc_o += Tm[wave][lane][j];
        o[t] = acc_o;
      }
      wave_lds_sync();
    }
  }
  if (!SPLIT && lane < 32 && kbase + lane < a.nk)
#pragma unroll
    for (int t = 0; t < NTG; t++) if (t < nts) a.out[(long long)a.tlist[t0 + t] * a.ldo + a.klist[kbase + lane]] = o[t];
}

// The terms of the split form with several parts, in place: one thread per (t, w, k), the
// same double expression as the fused kernel's.
__global__ void k_ww_terms(WwArgs a, int nwp, unsigned long long* Pbuf, const double* X,
                           const long long* Gm, const int* cntp, const int* llk, long long S)
{
  const long long tot = (long long)a.nt * a.nw * a.nk;
  const double* alpha = a.ajdk + (size_t)1 * a.pcols;
  const double mn = a.scal[2], mlt = a.scal[4];
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(e % a.nk);
    const long long tw = e / a.nk;
    const int w = (int)(tw % a.nw), tt = (int)(tw / a.nw);
    const int kk = a.klist[k], qq = a.qlist[k], ww = a.wlist[w];
    const double alpha_q = alpha[qq / kPopBits];
    const double bkk = (double)llk[(size_t)tt * a.nk + k] * mn * mlt;
    const long long off = 128 * S * (long long)cntp[(size_t)tt * a.nk + k];
    unsigned long long* P = Pbuf + ((size_t)tt * nwp + w) * a.nk + k;
    const long long G = 2 * ((long long)*P + off) - Gm[w];
    const double gd = (double)G;
    const double tmp = (double)(bkk + gd) / mlt - X[((size_t)(kk / kPopBits) * a.nt + tt) * a.nw + w];
    const int vS = bitv(a.Sc.p + (size_t)a.slist[tt] * a.Sc.rb, ww) ? 1 : -1;
    const int vZ = bitv(a.Q.p + (size_t)qq * a.Q.rb, ww) ? 1 : -1;
    *(double*)P = tmp * vS * (vZ - alpha_q);
  }
}
// the ordered sums over w of the terms, one thread per (k, t): loads along k coalesce and
// do not wait on the sum
__global__ void k_ww_sum(WwArgs a, int nwp, const unsigned long long* Pbuf)
{
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)a.nk * a.nt) return;
  const int k = (int)(e % a.nk), tt = (int)(e / a.nk);
  const double* T = (const double*)Pbuf + (size_t)tt * nwp * a.nk + k;
  double acc_o = 0.0;
  for (int w = 0; w < a.nw; w++) acc_o += T[(size_t)w * a.nk];
  a.out[(long long)a.tlist[tt] * a.ldo + a.klist[k]] = acc_o;
}
__global__ void k_ww_zero(unsigned long long* p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0ull;
}

// The split form is taken when the fused grid (128 rows k per block, covariates in groups of
// 3) leaves CUs idle and the w steps can fill them: then one block per (rows, covariates,
// w step), and the n chunks of a step in parts (at least 4 chunks each) until about one
// block per CU.  (One block of this kernel fills a CU: its registers allow one wave per SIMD.)
// A fused grid of one to a few rounds of blocks with a ragged last round (257 to 511
// blocks: the last round leaves up to half the CUs idle for a whole block's w loop) also
// takes the one-part split when its estimated time is lower: rounds x w steps x chunks of
// per-chunk block time (about ndig / 3 us on MI355X, measured at 258-500 blocks) against
// the split's rounds x chunks plus writing and reading the terms (8 bytes per k, t, w at
// about 2.3 TB/s, plus the sum kernel's launch).
static const int kWwCUs = 256;
static int ww_split_parts(int nk, int nt, int nw, size_t prb, int ndig)
{
  if (nk <= 0 || nt <= 0 || nw <= 0) return 0;
  const long long fused = (long long)cdiv(nk, 128) * cdiv(nt, 3);
  const int nchunk = (int)(prb * 8 / 128);
  if (nchunk < 1 || nw < 1) return 0;
  // the terms buffer (8 bytes per k, t and w) stays under 1 GiB
  const double pbytes = (double)nt * (double)(cdiv(nw, 32) * 32) * (double)nk * 8.0;
  if (pbytes > (double)(1ll << 30)) return 0;
  const long long nws = cdiv(nw, 32);
  if (fused >= kWwCUs) {
    const double c_us = ndig / 3.0;
    const double t_fused = (double)cdiv(fused, kWwCUs) * nws * nchunk * c_us;
    const double t_split = (double)cdiv(fused * nws, kWwCUs) * nchunk * c_us + 2.0 * pbytes / 2.3e6 + 20.0;
    return t_split < 0.9 * t_fused ? 1 : 0;
  }
  const long long per = fused * nws;
  if (per >= 2 * fused && per >= kWwCUs) return 1;
  long long parts = (kWwCUs + per - 1) / per;
  parts = std::min<long long>(parts, std::max(1, nchunk / 4));
  if (parts <= 1 && per < 2 * fused) return 0;       // nothing to spread
  return (int)std::max<long long>(parts, 1);
}

// X[(b*nt + t)*nw + w] = (bp[t] alpha_b + H[t][w] alpha_b) / mlt for every column block b:
// the part of the reference's term that depends on k only through its column block
// (At_T_YnWt_S_Zn_ww.c:95-104), formed once per (block, covariate, w) with the same
// operations in the same order
__global__ void k_ww_xtab(const double* H, const double* bp, const double* ajdk, int pcols, int nt, int nw,
                          const double* scal, double* X)
{
  const long long tot = (long long)pcols * nt * nw;
  const double* alpha = ajdk + (size_t)1 * pcols;
  const double mlt = scal[4];
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int w = (int)(e % nw);
    const long long bt = e / nw;
    const int t = (int)(bt % nt), b = (int)(bt / nt);
    const double alpha_k = alpha[b];
    const double at_base = bp[t] * alpha_k;
    const double at_tmp = H[(size_t)t * nw + w] * alpha_k;
    X[e] = (double)(at_base + at_tmp) / mlt;
  }
}

// the split forms the tuner may try besides the cost model's: terms buffers up to 256 MiB
static const double kTunePbufMax = 256.0 * 1024 * 1024;
static double ww_pbytes(int nk, int nt, int nw) { return (double)nt * (double)(cdiv(nw, 32) * 32) * (double)nk * 8.0; }

size_t ww_mfma_work_bytes(int nk, int nt, int nw, size_t prb, int B, int pcols)
{
  int ndig = (B + 7) / 8;
  size_t nwp = (size_t)((nw + 31) / 32) * 32, ldn = prb * 8;
  size_t b = (size_t)ndig * nwp * ldn;
  b = (b + 255) & ~(size_t)255;
  b += sizeof(double) * ((size_t)nt * nw + nt + 8);
  b += sizeof(long long) * ((size_t)nw + 8);
  b += 2 * sizeof(int) * ((size_t)nt * nk + 8);
  b += sizeof(double) * ((size_t)pcols * nt * nw + 8) + 256;
  const bool terms = ww_split_parts(nk, nt, nw, prb, ndig) > 0 || (tune_mode() != 0 && ww_pbytes(nk, nt, nw) <= kTunePbufMax);
  if (terms) b += 256 + sizeof(unsigned long long) * (size_t)nt * nwp * nk;
  return b + 1024;
}

namespace {
// one form of the contraction after the digit and per-w / per-k kernels: nsplit 0 the fused
// kernel, 1 one block per (rows, w step) with k_ww_sum, >1 also n parts with k_ww_terms;
// kWwOneT + nsplit the same with one covariate per block (fewer registers: two blocks per
// CU, which hides the double epilogue of contractions over few n)
constexpr int kWwOneT = 100;
struct WwLaunch {
  WwArgs a; const int8_t* Dg; long long ldn; int nwp, ndig; const double* X; const double* bp; const long long* Gm;
  const int* cntp; const int* llk; unsigned long long* Pbuf; long long S; hipStream_t st;
};
void ww_launch(void* ctx, int choice)
{
  const WwLaunch& L = *(const WwLaunch*)ctx;
  const WwArgs& a = L.a;
  hipStream_t st = L.st;
  const bool one_t = choice >= kWwOneT;
  const int nsplit = one_t ? choice - kWwOneT : choice;
  const long long np = (long long)a.nt * L.nwp * a.nk;
  if (nsplit > 1)
    hipLaunchKernelGGL(k_ww_zero, dim3(cdiv(np, 256) < 65536 ? cdiv(np, 256) : 65536), dim3(256), 0, st, L.Pbuf, np);
  const int ntg = one_t ? 1 : (a.nt >= 3 ? 3 : a.nt);
  dim3 grid(cdiv(a.nk, 128), cdiv(a.nt, ntg), nsplit > 0 ? cdiv(a.nw, 32) * nsplit : 1);
#define WW_LAUNCH(NTG, ND) do { if (nsplit > 1) \
      hipLaunchKernelGGL((k_ww_mfma<NTG, ND, 2>), grid, dim3(256), 0, st, a, L.Dg, L.ldn, L.nwp, L.X, L.bp, L.Gm, L.cntp, L.llk, nsplit, L.Pbuf); \
    else if (nsplit == 1) \
      hipLaunchKernelGGL((k_ww_mfma<NTG, ND, 1>), grid, dim3(256), 0, st, a, L.Dg, L.ldn, L.nwp, L.X, L.bp, L.Gm, L.cntp, L.llk, 1, L.Pbuf); \
    else hipLaunchKernelGGL((k_ww_mfma<NTG, ND, 0>), grid, dim3(256), 0, st, a, L.Dg, L.ldn, L.nwp, L.X, L.bp, L.Gm, L.cntp, L.llk, 1, L.Pbuf); } while (0)
#define WW_ND(NTG) switch (L.ndig) { case 1: WW_LAUNCH(NTG, 1); break; case 2: WW_LAUNCH(NTG, 2); break; \
                                     case 3: WW_LAUNCH(NTG, 3); break; default: WW_LAUNCH(NTG, 4); break; }
  if (ntg == 3) { WW_ND(3) } else if (ntg == 2) { WW_ND(2) } else { WW_ND(1) }
#undef WW_ND
#undef WW_LAUNCH
  if (nsplit > 1) {
    const long long nte = (long long)a.nt * a.nw * a.nk;
    hipLaunchKernelGGL(k_ww_terms, dim3(cdiv(nte, 256) < 65536 ? cdiv(nte, 256) : 65536), dim3(256), 0, st,
                       a, L.nwp, L.Pbuf, L.X, L.Gm, L.cntp, L.llk, L.S);
  }
  if (nsplit > 0) {
    const long long nc = (long long)a.nk * a.nt;
    hipLaunchKernelGGL(k_ww_sum, dim3(cdiv(nc, 64)), dim3(64), 0, st, a, L.nwp, L.Pbuf);
  }
}
}  // namespace

hipError_t ww_mfma(const WwArgs& a, void* work, size_t work_bytes, hipStream_t st, const WwHook* hook)
{
  if (a.nk <= 0 || a.nt <= 0) return hipSuccess;
  int ndig = (a.B + 7) / 8;
  int nwp = ((a.nw + 31) / 32) * 32;
  long long ldn = (long long)a.prb * 8;
  uint8_t* p = (uint8_t*)work;
  int8_t* Dg = (int8_t*)p;
  size_t off = ((size_t)ndig * nwp * ldn + 255) & ~(size_t)255;
  double* H = (double*)(p + off); off += sizeof(double) * ((size_t)a.nt * a.nw + 8);
  double* bp = (double*)(p + off); off += sizeof(double) * ((size_t)a.nt + 8);
  long long* Gm = (long long*)(p + off); off += sizeof(long long) * ((size_t)a.nw + 8);
  int* cntp = (int*)(p + off); off += sizeof(int) * ((size_t)a.nt * a.nk + 8);
  int* llk = (int*)(p + off); off += sizeof(int) * ((size_t)a.nt * a.nk + 8);
  off = (off + 255) & ~(size_t)255;
  double* X = (double*)(p + off); off += sizeof(double) * ((size_t)a.pcols * a.nt * a.nw + 8);
  unsigned long long* Pbuf = (unsigned long long*)(((uintptr_t)(p + off) + 255) & ~(uintptr_t)255);
  const long long np = (long long)a.nt * nwp * a.nk;
  // the caller sizes work with ww_mfma_work_bytes; the split forms need the terms buffer
  if (work_bytes < off) return hipErrorInvalidValue;
  const bool terms_fit = (size_t)((uint8_t*)(Pbuf + np) - p) <= work_bytes;
  int nsplit = ww_split_parts(a.nk, a.nt, a.nw, a.prb, ndig);
  if (nsplit > 0 && !terms_fit) return hipErrorInvalidValue;
  if (nwp > 0) {
    long long tot = (long long)nwp * (ldn / 4);
    hipLaunchKernelGGL(k_ww_digits, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                       a.cz, a.ldc, a.nw, nwp, (int)(a.prb * 8) < a.ldc ? (int)(a.prb * 8) : (int)a.ldc,
                       (const uint8_t*)a.pm, ndig, ldn, Dg);
  }
  hipLaunchKernelGGL(k_ww_pre, dim3(a.nw + 1 + cdiv(a.nk, 4)), dim3(256), 0, st, a, H, bp, Gm, cntp, llk);
  {
    const long long nx = (long long)a.pcols * a.nt * a.nw;
    hipLaunchKernelGGL(k_ww_xtab, dim3(cdiv(nx, 256) < 65536 ? cdiv(nx, 256) : 65536), dim3(256), 0, st,
                       H, bp, a.ajdk, a.pcols, a.nt, a.nw, a.scal, X);
  }
  long long S = 0;
  for (int i = 0; i < ndig; i++) S += 1LL << (8 * i);
  WwLaunch L{ a, Dg, ldn, nwp, ndig, X, bp, Gm, cntp, llk, Pbuf, S, st };
  // the tuned form of this shape class, measured here on first use (every form writes the
  // same outputs, so the measured launches also leave the result)
  const int nchunk = (int)(ldn / 128);
  const int max_parts = std::min(16, nchunk);
  const int tmode = terms_fit ? tune_mode() : 0;
  if (tmode != 0 && a.nw > 0 && nchunk > 0) {
    const TuneKey key{ kTuneWw, tune_bucket(a.nk), tune_bucket(a.nw), tune_bucket(nchunk), a.nt, ndig };
    int ch = -1;
    if (tune_get(key, &ch)) {
      const int c = ch >= kWwOneT ? ch - kWwOneT : ch;
      if ((ch < kWwOneT || a.nt > 1) && (c == 0 || c == 1 || (c > 1 && c <= max_parts))) nsplit = ch;
    } else if (tune_measuring(tmode)) {
      int cand[12], nc = 0, mi = -1;
      cand[nc++] = 0; cand[nc++] = 1;
      for (int q = 2; q <= max_parts; q *= 2) cand[nc++] = q;
      if (a.nt > 1) { cand[nc++] = kWwOneT; cand[nc++] = kWwOneT + 1; }
      for (int i = 0; i < nc; i++) if (cand[i] == nsplit) mi = i;
      if (mi < 0) { mi = nc; cand[nc++] = nsplit; }
      float best_ms = 0, model_ms = 0;
      const int b = tune_measure(st, cand, nc, mi, ww_launch, &L, &best_ms, &model_ms, a.out,
                                 (size_t)std::max(a.out_elems, 0LL) * sizeof(double));
      tune_put(key, cand[b], best_ms, model_ms);
      nsplit = cand[b];
    }
  }
  if (hook && hook->before) hook->before(hook->ctx, st);
  ww_launch(&L, nsplit);
  if (hook && hook->after) hook->after(hook->ctx, st);
  return hipGetLastError();
}

// ------------------------------------------------------------------
// An_ZtSn_ww (An_ZtSn_ww.c:44-70) on the matrix cores.  The reference's plane sums
// give the exact integer x_s = sum_{n in m} a(n) cz_s(n) (+-1 a; every partial sum
// stays below 2^53), so any exact evaluation reproduces it:
//     x_s = 2 (sum_i 256^i P_i + 128 S |a & m|) - sum_{n in m} cz_s(n)
// with P_i = sum_n (a & m)(n) d_i(n) over the signed digits d_i = byte_i(cz) - 128.
// One block per 32 rows; its 4 waves split n and add their integer sums in LDS.
// ------------------------------------------------------------------
// Gpart[s*ZT_PARTS + b] = sum over the n of part b (n in mask, n < n_total) of cz[s*ldc + n]
#define ZT_PARTS 32
__global__ __launch_bounds__(256) void k_cz_sums(const uint32_t* cz, long long ldc, int n_total, const uint32_t* mask,
                                                 long long* Gpart)
{
  const int s = blockIdx.y, b = blockIdx.x;
  const int per = (n_total + ZT_PARTS - 1) / ZT_PARTS;
  const int n0 = b * per, n1 = min(n_total, n0 + per);
  long long v = 0;
  for (int n = n0 + (int)threadIdx.x; n < n1; n += 256)
    if ((mask[n >> 5] >> (((n >> 3) & 3) * 8 + 7 - (n & 7))) & 1u) v += cz[(long long)s * ldc + n];
  v = wsum64(v);
  __shared__ long long part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) Gpart[s * ZT_PARTS + b] = part[0] + part[1] + part[2] + part[3];
}

// The rows' packed bits stream from HBM once per launch; the kernel keeps several chunks
// of them in flight per wave (loads issued one iteration ahead) and, when the rows alone
// give too few blocks to fill the device, also splits n over blockIdx.y.  Each block adds
// its waves' integer sums in LDS; with one part it forms the outputs itself, else it adds
// them to Pg / Cg / Mg with 64-bit integer atomics (exact in any order) and k_ztsn_fin
// forms the outputs.
template <int ND, bool PARTS>
__device__ __forceinline__ void ztsn_block(const uint8_t* A, size_t rbA, const int* alist, int na,
                                           const uint8_t* mask, const int8_t* Dg, long long ldn, int ns,
                                           const long long* Gsum, const int* slist, const double* scal,
                                           double* out, long long ldo, unsigned long long* Pg, int bx, int part,
                                           int nparts)
{
  __shared__ int Ps[32][33];
  __shared__ int Cs[32], Ms[32];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  const int i0 = bx * 32;
  const uint8_t* Arow = A + (size_t)alist[min(i0 + r, na - 1)] * rbA + 8 * g;
  const uint8_t* mrow = mask + 8 * g;
  const int ncol = ND * ns;
  const int8_t* Bcol = Dg + (long long)min(r, ncol - 1) * ldn + 64 * g;
  for (int e = tid; e < 32 * 33; e += 256) (&Ps[0][0])[e] = 0;
  if (tid < 32) { Cs[tid] = 0; Ms[tid] = 0; }
  v16i acc;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0;
  int cnt = 0, mcnt = 0;
  const int nchunk = (int)(ldn / 128);
  const int c_lo = (int)((long long)part * nchunk / nparts);
  const int c_hi = (int)((long long)(part + 1) * nchunk / nparts);
  // chunk c of this wave: its A and mask words and its 4 digit fragments
  uint64_t a_n = 0, m_n = 0;
  v4i b_n[4];
  auto load = [&](int c) {
    m_n = *(const uint64_t*)(mrow + 16LL * c);
    a_n = *(const uint64_t*)(Arow + 16LL * c);
#pragma unroll
    for (int s = 0; s < 4; s++) b_n[s] = *(const v4i*)(Bcol + 128LL * c + 16 * s);
  };
  int c = c_lo + wave;
  if (c < c_hi) load(c);
  for (; c < c_hi; c += 4) {
    const uint64_t m8 = m_n, a8 = a_n & m_n;
    v4i bf[4];
#pragma unroll
    for (int s = 0; s < 4; s++) bf[s] = b_n[s];
    if (c + 4 < c_hi) load(c + 4);
    cnt += __popcll(a8); mcnt += __popcll(m8);
    const uint32_t wr0 = __builtin_bitreverse32((uint32_t)a8), wr1 = __builtin_bitreverse32((uint32_t)(a8 >> 32));
#pragma unroll
    for (int s = 0; s < 4; s++)
      acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(spread16_cols(s < 2 ? wr0 : wr1, s & 1), bf[s], acc, 0, 0, 0);
  }
  __syncthreads();
  // integer sums over the waves (order free): D row 8(v/4) + 4g + v%4, column r
#pragma unroll
  for (int v = 0; v < 16; v++) atomicAdd(&Ps[8 * (v / 4) + 4 * g + (v % 4)][r], acc[v]);
  atomicAdd(&Cs[r], cnt);
  atomicAdd(&Ms[r], mcnt);
  __syncthreads();
  if (PARTS) {
    // rows i0 .. i0 + 31: 32 digit columns, then the two counts
    for (int e = tid; e < 32 * 34; e += 256) {
      const int rr = e / 34, f = e % 34;
      if (i0 + rr >= na) continue;
      const long long v = f < 32 ? (long long)Ps[rr][f] : (f == 32 ? (long long)Cs[rr] : (long long)Ms[rr]);
      if (v) atomicAdd(Pg + (size_t)(i0 + rr) * 34 + f, (unsigned long long)v);
    }
    return;
  }
  const double mn = scal[2], mlt = scal[4];
  long long S = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) S += 1LL << (8 * i);
  for (int e = tid; e < 32 * ns; e += 256) {
    const int rr = e & 31, s = e >> 5;
    if (i0 + rr >= na) continue;
    long long P = 0;
#pragma unroll
    for (int i = ND - 1; i >= 0; i--) P = P * 256 + (long long)Ps[rr][i * ns + s];
    long long G = 0;
    for (int b = 0; b < ZT_PARTS; b++) G += Gsum[s * ZT_PARTS + b];
    const long long x = 2 * (P + 128 * S * (long long)Cs[rr]) - G;
    const double ll = (double)(2LL * Cs[rr] - Ms[rr]);
    const double base = ll * mn * mlt;
    const double tmp = (double)x;
    out[(long long)slist[s] * ldo + alist[i0 + rr]] = (double)(base + tmp) / mlt;
  }
}
template <int ND, bool PARTS>
__global__ __launch_bounds__(256) void k_ztsn_mfma(const uint8_t* A, size_t rbA, const int* alist, int na,
                                                   const uint8_t* mask, const int8_t* Dg, long long ldn, int ns,
                                                   const long long* Gsum, const int* slist, const double* scal,
                                                   double* out, long long ldo, unsigned long long* Pg)
{
  ztsn_block<ND, PARTS>(A, rbA, alist, na, mask, Dg, ldn, ns, Gsum, slist, scal, out, ldo, Pg, blockIdx.x, blockIdx.y,
                        gridDim.y);
}
// several contractions in one launch (blockIdx.z picks the job), each split over its own
// number of n parts into its own Pg; blocks past a job's rows or parts leave at once
template <int ND>
__global__ __launch_bounds__(256) void k_ztsn_mfma_jobs(ZtsnJobs J)
{
  const ZtsnJob& j = J.job[blockIdx.z];
  if ((int)blockIdx.x * 32 >= j.na || (int)blockIdx.y >= j.parts) return;
  ztsn_block<ND, true>(j.A, j.rbA, j.alist, j.na, j.mask, j.Dg, j.ldn, j.ns, j.Gsum, j.slist, j.scal, j.out, j.ldo,
                       j.Pg, blockIdx.x, blockIdx.y, j.parts);
}

// the outputs of the split form from the summed integers, one thread per (row, covariate):
// the same expressions as the one-part kernel's
template <int ND>
__device__ __forceinline__ void ztsn_fin_at(const unsigned long long* Pg, int na, int ns, const long long* Gsum,
                                            const int* alist, const int* slist, const double* scal, double* out,
                                            long long ldo, int e)
{
  const int rr = e % na, s = e / na;
  const long long* row = (const long long*)Pg + (size_t)rr * 34;
  const double mn = scal[2], mlt = scal[4];
  long long S = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) S += 1LL << (8 * i);
  long long P = 0;
#pragma unroll
  for (int i = ND - 1; i >= 0; i--) P = P * 256 + row[i * ns + s];
  long long G = 0;
  for (int b = 0; b < ZT_PARTS; b++) G += Gsum[s * ZT_PARTS + b];
  const long long cnt = row[32], mcnt = row[33];
  const long long x = 2 * (P + 128 * S * cnt) - G;
  const double ll = (double)(2LL * cnt - mcnt);
  const double base = ll * mn * mlt;
  const double tmp = (double)x;
  out[(long long)slist[s] * ldo + alist[rr]] = (double)(base + tmp) / mlt;
}
template <int ND>
__global__ void k_ztsn_fin(const unsigned long long* Pg, int na, int ns, const long long* Gsum, const int* alist,
                           const int* slist, const double* scal, double* out, long long ldo)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < na * ns) ztsn_fin_at<ND>(Pg, na, ns, Gsum, alist, slist, scal, out, ldo, e);
}
template <int ND>
__global__ void k_ztsn_fin_jobs(ZtsnJobs J)
{
  const ZtsnJob& j = J.job[blockIdx.y];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < j.na * j.ns) ztsn_fin_at<ND>(j.Pg, j.na, j.ns, j.Gsum, j.alist, j.slist, j.scal, j.out, j.ldo, e);
}
__global__ void k_ztsn_zero_jobs(ZtsnJobs J)
{
  const ZtsnJob& j = J.job[blockIdx.y];
  const long long n = 34LL * j.na;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) j.Pg[i] = 0ull;
}

__global__ void k_ztsn_zero(unsigned long long* p, long long n)
{
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0ull;
}

// n parts: enough blocks for about four per CU, each wave keeping at least 8 chunks
static int ztsn_parts(int na, size_t prb)
{
  if (na <= 0) return 1;
  const long long rows_blocks = cdiv(na, 32);
  const int nchunk = (int)(prb * 8 / 128);
  long long parts = (1024 + rows_blocks - 1) / rows_blocks;
  parts = std::min<long long>(parts, std::max(1, nchunk / 32));
  return (int)std::max<long long>(parts, 1);
}

size_t ztsn_mfma_work_bytes(int na, int ns, size_t prb, int B)
{
  const int ndig = (B + 7) / 8;
  size_t b = (((size_t)ndig * ns * prb * 8 + 255) & ~(size_t)255) + sizeof(long long) * ((size_t)ns * ZT_PARTS + 8) + 256;
  if (ztsn_parts(na, prb) > 1) b += sizeof(unsigned long long) * 34 * (size_t)na + 256;
  return b;
}

// The digit planes of cz and their column sums (the first half of ztsn_mfma): they depend on
// cz, its mask and the row width only, so several row sets may share them.
hipError_t ztsn_mfma_prep(size_t rb, const uint32_t* mask, const uint32_t* cz, long long ldc, int n_total, int ns, int B,
                          void* work, size_t work_bytes, hipStream_t st)
{
  if (ns <= 0) return hipSuccess;
  const int ndig = (B + 7) / 8;
  const long long ldn = (long long)rb * 8;
  if (ndig * ns > 32 || work_bytes < ztsn_mfma_work_bytes(0, ns, rb, B)) return hipErrorInvalidValue;
  int8_t* Dg = (int8_t*)work;
  long long* Gsum = (long long*)((uint8_t*)work + (((size_t)ndig * ns * ldn + 255) & ~(size_t)255));
  const int nt_ = n_total < ldn ? n_total : (int)ldn;
  const long long tot = (long long)ns * (ldn / 4);
  hipLaunchKernelGGL(k_ww_digits, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     cz, ldc, ns, ns, nt_, (const uint8_t*)mask, ndig, ldn, Dg);
  hipLaunchKernelGGL(k_cz_sums, dim3(ZT_PARTS, ns), dim3(256), 0, st, cz, ldc, nt_, mask, Gsum);
  return hipGetLastError();
}

// The contraction of the listed rows of A with digit planes ztsn_mfma_prep left in work.
hipError_t ztsn_mfma_run(const Bits& A, const int* alist, int na, const uint32_t* mask, int ns, const int* slist, int B,
                         const double* scal, double* out, long long ldo, void* work, size_t work_bytes, hipStream_t st)
{
  if (na <= 0 || ns <= 0) return hipSuccess;
  const int ndig = (B + 7) / 8;
  const long long ldn = (long long)A.rb * 8;
  if (ndig * ns > 32 || work_bytes < ztsn_mfma_work_bytes(na, ns, A.rb, B)) return hipErrorInvalidValue;
  int8_t* Dg = (int8_t*)work;
  long long* Gsum = (long long*)((uint8_t*)work + (((size_t)ndig * ns * ldn + 255) & ~(size_t)255));
  unsigned long long* Pg = (unsigned long long*)((((uintptr_t)(Gsum + (size_t)ns * ZT_PARTS + 8)) + 255) & ~(uintptr_t)255);
  const int parts = ztsn_parts(na, A.rb);
  if (parts > 1) {
    const long long np = 34LL * na;
    hipLaunchKernelGGL(k_ztsn_zero, dim3(cdiv(np, 256) < 65536 ? cdiv(np, 256) : 65536), dim3(256), 0, st, Pg, np);
  }
  const dim3 grid(cdiv(na, 32), parts);
#define ZT_LAUNCH(N) do { if (parts > 1) { \
      hipLaunchKernelGGL((k_ztsn_mfma<N, true>), grid, dim3(256), 0, st, A.p, A.rb, alist, na, (const uint8_t*)mask, Dg, \
                         ldn, ns, Gsum, slist, scal, out, ldo, Pg); \
      hipLaunchKernelGGL((k_ztsn_fin<N>), dim3(cdiv((long long)na * ns, 256)), dim3(256), 0, st, Pg, na, ns, Gsum, alist, \
                         slist, scal, out, ldo); \
    } else hipLaunchKernelGGL((k_ztsn_mfma<N, false>), grid, dim3(256), 0, st, A.p, A.rb, alist, na, (const uint8_t*)mask, \
                              Dg, ldn, ns, Gsum, slist, scal, out, ldo, Pg); } while (0)
  switch (ndig) { case 1: ZT_LAUNCH(1); break; case 2: ZT_LAUNCH(2); break; case 3: ZT_LAUNCH(3); break; default: ZT_LAUNCH(4); break; }
#undef ZT_LAUNCH
  return hipGetLastError();
}

int ztsn_job_parts(int na, size_t rb) { return ztsn_parts(na, rb); }

// Many contractions with prepared digit planes (ztsn_mfma_prep), each with its own Pg of
// 34 * na words: zeroed, accumulated and finished in three launches per group of jobs.  The
// integer sums are exact in any order, so each output is the one ztsn_mfma_run writes.
hipError_t ztsn_mfma_jobs(const std::vector<ZtsnJob>& v, int B, hipStream_t st)
{
  const int ndig = (B + 7) / 8;
  for (size_t j0 = 0; j0 < v.size(); j0 += ZtsnJobs::kMax) {
    ZtsnJobs J;
    J.n = (int)std::min<size_t>(ZtsnJobs::kMax, v.size() - j0);
    int rows = 0, parts = 0; long long outs = 0, words = 0;
    for (int j = 0; j < J.n; j++) {
      J.job[j] = v[j0 + j];
      if (ndig * J.job[j].ns > 32) return hipErrorInvalidValue;
      rows = std::max(rows, J.job[j].na); parts = std::max(parts, J.job[j].parts);
      outs = std::max(outs, (long long)J.job[j].na * J.job[j].ns); words = std::max(words, 34LL * J.job[j].na);
    }
    if (rows <= 0) continue;
    hipLaunchKernelGGL(k_ztsn_zero_jobs, dim3(std::min<unsigned>(cdiv(words, 256), 1024), J.n), dim3(256), 0, st, J);
    const dim3 grid(cdiv(rows, 32), parts, J.n);
    const dim3 gfin(cdiv(outs, 256), J.n);
    switch (ndig) {
      case 1: hipLaunchKernelGGL((k_ztsn_mfma_jobs<1>), grid, dim3(256), 0, st, J);
              hipLaunchKernelGGL((k_ztsn_fin_jobs<1>), gfin, dim3(256), 0, st, J); break;
      case 2: hipLaunchKernelGGL((k_ztsn_mfma_jobs<2>), grid, dim3(256), 0, st, J);
              hipLaunchKernelGGL((k_ztsn_fin_jobs<2>), gfin, dim3(256), 0, st, J); break;
      case 3: hipLaunchKernelGGL((k_ztsn_mfma_jobs<3>), grid, dim3(256), 0, st, J);
              hipLaunchKernelGGL((k_ztsn_fin_jobs<3>), gfin, dim3(256), 0, st, J); break;
      default: hipLaunchKernelGGL((k_ztsn_mfma_jobs<4>), grid, dim3(256), 0, st, J);
               hipLaunchKernelGGL((k_ztsn_fin_jobs<4>), gfin, dim3(256), 0, st, J); break;
    }
  }
  return hipGetLastError();
}

hipError_t ztsn_mfma(const Bits& A, const int* alist, int na, const uint32_t* mask, const uint32_t* cz, long long ldc,
               int n_total, int ns, const int* slist, int B, const double* scal, double* out, long long ldo,
               void* work, size_t work_bytes, hipStream_t st)
{
  if (na <= 0 || ns <= 0) return hipSuccess;
  if ((B + 7) / 8 * ns > 32 || work_bytes < ztsn_mfma_work_bytes(na, ns, A.rb, B)) return hipErrorInvalidValue;
  hipError_t e = ztsn_mfma_prep(A.rb, mask, cz, ldc, n_total, ns, B, work, work_bytes, st);
  if (e != hipSuccess) return e;
  return ztsn_mfma_run(A, alist, na, mask, ns, slist, B, scal, out, ldo, work, work_bytes, st);
}

}  // namespace lkc
