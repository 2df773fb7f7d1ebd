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
#include <stdio.h>

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

// per w: Gm[w] = sum_{n in mask} cz(w,n);  H[t][w] = (double) sum_{n in mask} T_t(n)+-1 cz(w,n)
// block nw: bp[t] = (double)(sum_{n in mask} T_t(n)+-1) * min * mlt
__global__ __launch_bounds__(256) void k_ww_hg(WwArgs a, double* H, double* bp, long long* Gm)
{
  __shared__ long long red[4][kMaxT + 1];
  const int j = blockIdx.x, wave = threadIdx.x >> 6;
  const uint32_t* pm = a.pm;
  const int nwd = (int)(a.prb / 4);       // 32-column words of the n extent
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
  if ((threadIdx.x & 63) == 0) {
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

// per k: cntp[t][k] = |mask & (P_k xnor T_t)|, llk[t][k] = pmpm0(P_k, T_t, mask, mask)
__global__ void k_ww_cntp(WwArgs a, int* cntp, int* llk)
{
  int k = blockIdx.x;
  const uint32_t* Pr = (const uint32_t*)(a.P.p + (size_t)a.klist[k] * a.P.rb);
  int nwd = (int)(a.prb / 4);
  for (int t = 0; t < a.nt; t++) {
    const uint32_t* Tr = (const uint32_t*)(a.Tc.p + (size_t)a.tlist[t] * a.Tc.rb);
    long long p = 0, q = 0;
    for (int w = threadIdx.x; w < nwd; w += 64) {
      uint32_t e = Pr[w] ^ Tr[w], m = a.pm[w];
      p += popc32(~e & m); q += popc32(e & m);
    }
    p = wsum64(p); q = wsum64(q);
    if (threadIdx.x == 0) { cntp[(size_t)t * a.nk + k] = (int)p; llk[(size_t)t * a.nk + k] = (int)(p - q); }
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

// Block: 4 waves x 32 rows k; the 32 w of a step are shared, their digits staged in LDS
// per 128-column chunk of n (double buffered, the next chunk loaded while the current
// one is used).  Lane (r, g) of a step s feeds columns n0 + 64g + 16s .. +15: the
// packed bits of P, T and the mask come as one 8-byte word per lane and chunk.
template <int NTG, int ND>
__global__ __launch_bounds__(256) void k_ww_mfma(WwArgs a, const int8_t* Dg, long long ldn, int nwp,
                                                 const double* H, const double* bp, const long long* Gm,
                                                 const int* cntp, const int* llk)
{
  __shared__ __attribute__((aligned(16))) int8_t Bs[2][ND][32][WW_BROW];
  __shared__ long long Gs[4][32][33];
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
  // ordered sums over w of row kbase + lane (lane < 32), per covariate
  double o[NTG];
#pragma unroll
  for (int t = 0; t < NTG; t++) o[t] = 0.0;
  const int nchunk = (int)(ldn / 128);
  // digit loader: piece e = tid + 256 j -> digit plane j, w row tid / 8, 16 bytes (tid % 8)
  const int lw = tid >> 3, lp = tid & 7;

  for (int w0 = 0; w0 < a.nw; w0 += 32) {
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
    __syncthreads();   // the previous w step's readers are done with Bs
    if (nchunk > 0) { WW_BLOAD(0); WW_ALOAD(0); WW_BSTORE(0); }
    __syncthreads();
    for (int c = 0; c < nchunk; c++) {
      const int buf = c & 1;
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
      const int cn = min(c + 1, nchunk - 1);
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
    // epilogue: each lane < 32 walks its row's 32 w in ascending order
    const int wn = min(32, a.nw - w0);
#pragma unroll
    for (int t = 0; t < NTG; t++) {
      if (t >= nts) break;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        long long spc = 0;
#pragma unroll
        for (int i = ND - 1; i >= 0; i--) spc = spc * 256 + (long long)acc[t][i][v];
        Gs[wave][8 * (v / 4) + 4 * g + (v % 4)][r] = spc;
      }
      __syncthreads();
      if (lane < 32) {
        // per-row constants (recomputed here to keep them out of the main loop's registers)
        const double mn = a.scal[2], mlt = a.scal[4];
        const double* alpha = a.ajdk + (size_t)1 * a.pcols;
        const int myk = min(kbase + lane, a.nk - 1);
        const int kk = a.klist[myk], qq = a.qlist[myk];
        const double alpha_k = alpha[kk / kPopBits], alpha_q = alpha[qq / kPopBits];
        const uint8_t* Qrow = a.Q.p + (size_t)qq * a.Q.rb;
        long long S = 0;
#pragma unroll
        for (int i = 0; i < ND; i++) S += 1LL << (8 * i);
        const int tt = t0 + t;
        const double bkk = (double)llk[(size_t)tt * a.nk + myk] * mn * mlt;
        const long long off = 128 * S * (long long)cntp[(size_t)tt * a.nk + myk];
        const double at_base = bp[tt] * alpha_k;
        const uint8_t* Srow = a.Sc.p + (size_t)a.slist[tt] * a.Sc.rb;
        const double* Ht = H + (size_t)tt * a.nw;
        double acc_o = o[t];
        for (int j = 0; j < wn; j++) {
          const int w = w0 + j, ww = a.wlist[w];
          const long long G = 2 * (Gs[wave][lane][j] + off) - Gm[w];
          const double at_tmp = Ht[w] * alpha_k;
          const double gd = (double)G;
          const double tmp = (double)(bkk + gd) / mlt - (double)(at_base + at_tmp) / mlt;
          const int vS = bitv(Srow, ww) ? 1 : -1, vZ = bitv(Qrow, ww) ? 1 : -1;
          acc_o += tmp * vS * (vZ - alpha_q);
        }
        o[t] = acc_o;
      }
      __syncthreads();
    }
  }
  if (lane < 32 && kbase + lane < a.nk)
#pragma unroll
    for (int t = 0; t < NTG; t++) if (t < nts) a.out[(long long)a.tlist[t0 + t] * a.ldo + a.klist[kbase + lane]] = o[t];
}

size_t ww_mfma_work_bytes(int nk, int nt, int nw, size_t prb, int B)
{
  int ndig = (B + 7) / 8;
  size_t nwp = (size_t)((nw + 31) / 32) * 32, ldn = prb * 8;
  size_t b = (size_t)ndig * nwp * ldn;
  b = (b + 255) & ~(size_t)255;
  b += sizeof(double) * ((size_t)nt * nw + nt + 8);
  b += sizeof(long long) * ((size_t)nw + 8);
  b += 2 * sizeof(int) * ((size_t)nt * nk + 8);
  return b + 1024;
}

hipError_t ww_mfma(const WwArgs& a, void* work, size_t work_bytes, hipStream_t st, const WwHook* hook)
{
  if (a.nk <= 0 || a.nt <= 0) return hipSuccess;
  int ndig = (a.B + 7) / 8;
  int nwp = ((a.nw + 31) / 32) * 32;
  long long ldn = (long long)a.prb * 8;
  if (work_bytes < ww_mfma_work_bytes(a.nk, a.nt, a.nw, a.prb, a.B)) return hipErrorInvalidValue;   // the caller sizes work with ww_mfma_work_bytes
  uint8_t* p = (uint8_t*)work;
  int8_t* Dg = (int8_t*)p;
  size_t off = ((size_t)ndig * nwp * ldn + 255) & ~(size_t)255;
  double* H = (double*)(p + off); off += sizeof(double) * ((size_t)a.nt * a.nw + 8);
  double* bp = (double*)(p + off); off += sizeof(double) * ((size_t)a.nt + 8);
  long long* Gm = (long long*)(p + off); off += sizeof(long long) * ((size_t)a.nw + 8);
  int* cntp = (int*)(p + off); off += sizeof(int) * ((size_t)a.nt * a.nk + 8);
  int* llk = (int*)(p + off);
  if (nwp > 0) {
    long long tot = (long long)nwp * (ldn / 4);
    hipLaunchKernelGGL(k_ww_digits, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                       a.cz, a.ldc, a.nw, nwp, (int)(a.prb * 8) < a.ldc ? (int)(a.prb * 8) : (int)a.ldc,
                       (const uint8_t*)a.pm, ndig, ldn, Dg);
  }
  hipLaunchKernelGGL(k_ww_hg, dim3(a.nw + 1), dim3(256), 0, st, a, H, bp, Gm);
  hipLaunchKernelGGL(k_ww_cntp, dim3(a.nk), dim3(64), 0, st, a, cntp, llk);
  dim3 grid(cdiv(a.nk, 128), cdiv(a.nt, a.nt >= 3 ? 3 : a.nt));
  if (hook && hook->before) hook->before(hook->ctx, st);
#define WW_LAUNCH(NTG, ND) hipLaunchKernelGGL((k_ww_mfma<NTG, ND>), grid, dim3(256), 0, st, a, Dg, ldn, nwp, H, bp, Gm, cntp, llk)
#define WW_ND(NTG) switch (ndig) { case 1: WW_LAUNCH(NTG, 1); break; case 2: WW_LAUNCH(NTG, 2); break; \
                                   case 3: WW_LAUNCH(NTG, 3); break; default: WW_LAUNCH(NTG, 4); break; }
  if (a.nt >= 3) { WW_ND(3) } else if (a.nt == 2) { WW_ND(2) } else { WW_ND(1) }
#undef WW_ND
#undef WW_LAUNCH
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

template <int ND>
__global__ __launch_bounds__(256) void k_ztsn_mfma(const uint8_t* A, size_t rbA, const int* alist, int na,
                                                   const uint8_t* mask, const int8_t* Dg, long long ldn, int ns,
                                                   const long long* Gsum, const int* slist, const double* scal,
                                                   double* out, long long ldo)
{
  __shared__ int Ps[32][33];
  __shared__ int Cs[32], Ms[32];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  const int i0 = blockIdx.x * 32;
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
  for (int c = wave; c < nchunk; c += 4) {
    const uint64_t m8 = *(const uint64_t*)(mrow + 16LL * c);
    const uint64_t a8 = *(const uint64_t*)(Arow + 16LL * c) & m8;
    v4i bf[4];
#pragma unroll
    for (int s = 0; s < 4; s++) bf[s] = *(const v4i*)(Bcol + 128LL * c + 16 * s);
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

size_t ztsn_mfma_work_bytes(int ns, size_t prb, int B)
{
  const int ndig = (B + 7) / 8;
  return (((size_t)ndig * ns * prb * 8 + 255) & ~(size_t)255) + sizeof(long long) * ((size_t)ns * ZT_PARTS + 8) + 256;
}

hipError_t ztsn_mfma(const Bits& A, const int* alist, int na, const uint32_t* mask, const uint32_t* cz, long long ldc,
               int n_total, int ns, const int* slist, int B, const double* scal, double* out, long long ldo,
               void* work, size_t work_bytes, hipStream_t st)
{
  if (na <= 0 || ns <= 0) return hipSuccess;
  const int ndig = (B + 7) / 8;
  const long long ldn = (long long)A.rb * 8;
  if (ndig * ns > 32 || work_bytes < ztsn_mfma_work_bytes(ns, A.rb, B)) return hipErrorInvalidValue;
  int8_t* Dg = (int8_t*)work;
  long long* Gsum = (long long*)((uint8_t*)work + (((size_t)ndig * ns * ldn + 255) & ~(size_t)255));
  const int nt_ = n_total < ldn ? n_total : (int)ldn;
  const long long tot = (long long)ns * (ldn / 4);
  hipLaunchKernelGGL(k_ww_digits, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     cz, ldc, ns, ns, nt_, (const uint8_t*)mask, ndig, ldn, Dg);
  hipLaunchKernelGGL(k_cz_sums, dim3(ZT_PARTS, ns), dim3(256), 0, st, cz, ldc, nt_, mask, Gsum);
#define ZT_LAUNCH(N) hipLaunchKernelGGL((k_ztsn_mfma<N>), dim3(cdiv(na, 32)), dim3(256), 0, st, A.p, A.rb, alist, na, \
                                        (const uint8_t*)mask, Dg, ldn, ns, Gsum, slist, scal, out, ldo)
  switch (ndig) { case 1: ZT_LAUNCH(1); break; case 2: ZT_LAUNCH(2); break; case 3: ZT_LAUNCH(3); break; default: ZT_LAUNCH(4); break; }
#undef ZT_LAUNCH
  return hipGetLastError();
}

}  // namespace lkc
