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
__global__ void k_ww_hg(WwArgs a, double* H, double* bp, long long* Gm)
{
  int j = blockIdx.x;
  const uint8_t* pm = (const uint8_t*)a.pm;
  int n_total = (int)(a.prb * 8);
  long long g = 0, h[kMaxT];
  for (int t = 0; t < kMaxT; t++) h[t] = 0;
  for (int n = threadIdx.x; n < n_total; n += 64) {
    if (!bitv(pm, n)) continue;
    long long v = (j < a.nw) ? (long long)a.cz[(long long)j * a.ldc + n] : 1;
    g += v;
    for (int t = 0; t < a.nt; t++) h[t] += bitv(a.Tc.p + (size_t)a.tlist[t] * a.Tc.rb, n) ? v : -v;
  }
  g = wsum64(g);
  for (int t = 0; t < a.nt; t++) h[t] = wsum64(h[t]);
  if (threadIdx.x != 0) return;
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

// byte x -> 8 bytes, byte j = bit (7-j) of x (the bit of element 8m+j of a packed row)
__device__ __forceinline__ uint64_t expand_byte(uint32_t x)
{
  uint64_t r = 0;
  for (int j = 0; j < 8; j++) r |= (uint64_t)((x >> (7 - j)) & 1u) << (8 * j);
  return r;
}

template <int NTG, int ND>
__global__ __launch_bounds__(256) void k_ww_mfma(WwArgs a, const int8_t* Dg, long long ldn, int nwp,
                                                 const double* H, const double* bp, const long long* Gm,
                                                 const int* cntp, const int* llk)
{
  __shared__ uint64_t lut[256];
  __shared__ long long Gs[4][32][33];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, r = lane & 31;
  for (int i = tid; i < 256; i += 256) lut[i] = expand_byte(i);
  const int t0 = blockIdx.y * NTG;
  const int nts = min(NTG, a.nt - t0);
  const int kbase = blockIdx.x * 128 + wave * 32;
  const int krow = min(kbase + r, a.nk - 1);
  const uint8_t* Prow = a.P.p + (size_t)a.klist[krow] * a.P.rb;
  const uint8_t* pm = (const uint8_t*)a.pm;
  const uint8_t* Trow[NTG];
#pragma unroll
  for (int t = 0; t < NTG; t++) Trow[t] = a.Tc.p + (size_t)a.tlist[t0 + min(t, nts - 1)] * a.Tc.rb;
  // per-row constants of the epilogue (lane < 32 owns row kbase + lane)
  const double mn = a.scal[2], mlt = a.scal[4];
  const double* alpha = a.ajdk + (size_t)1 * a.pcols;
  const int myk = min(kbase + lane, a.nk - 1);
  const int kk = a.klist[myk], qq = a.qlist[myk];
  const double alpha_k = alpha[kk / kPopBits], alpha_q = alpha[qq / kPopBits];
  const uint8_t* Qrow = a.Q.p + (size_t)qq * a.Q.rb;
  long long S = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) S += 1LL << (8 * i);
  double o[NTG], bkk[NTG];
  long long off[NTG];
#pragma unroll
  for (int t = 0; t < NTG; t++) {
    o[t] = 0.0;
    const int tt = t0 + min(t, nts - 1);
    bkk[t] = (double)llk[(size_t)tt * a.nk + myk] * mn * mlt;
    off[t] = 128 * S * (long long)cntp[(size_t)tt * a.nk + myk];
  }
  __syncthreads();

  for (int w0 = 0; w0 < a.nw; w0 += 32) {
    v16i acc[NTG][ND];
#pragma unroll
    for (int t = 0; t < NTG; t++)
#pragma unroll
      for (int i = 0; i < ND; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) acc[t][i][v] = 0;
    const int8_t* Bcol = Dg + (long long)(w0 + r) * ldn + 64 * g;
    for (long long n0 = 0; n0 < ldn; n0 += 128) {
      const long long bo = n0 / 8 + 8 * g;
      const uint64_t m8 = *(const uint64_t*)(pm + bo);
      if (!__any(m8 != 0)) continue;
      const uint64_t p8 = *(const uint64_t*)(Prow + bo);
      uint64_t pos8[NTG];
#pragma unroll
      for (int t = 0; t < NTG; t++) pos8[t] = ~(p8 ^ *(const uint64_t*)(Trow[t] + bo)) & m8;
#pragma unroll 2
      for (int s = 0; s < 4; s++) {
        v4i bf[ND];
#pragma unroll
        for (int i = 0; i < ND; i++) bf[i] = *(const v4i*)(Bcol + (long long)i * nwp * ldn + n0 + 16 * s);
#pragma unroll
        for (int t = 0; t < NTG; t++) {
          v4i af;
          const uint64_t e0 = lut[(pos8[t] >> (16 * s)) & 0xFF], e1 = lut[(pos8[t] >> (16 * s + 8)) & 0xFF];
          af[0] = (int)(uint32_t)e0; af[1] = (int)(uint32_t)(e0 >> 32);
          af[2] = (int)(uint32_t)e1; af[3] = (int)(uint32_t)(e1 >> 32);
#pragma unroll
          for (int i = 0; i < ND; i++) acc[t][i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[i], acc[t][i], 0, 0, 0);
        }
      }
    }
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
        const int tt = t0 + t;
        const double at_base = bp[tt] * alpha_k;
        const uint8_t* Srow = a.Sc.p + (size_t)a.slist[tt] * a.Sc.rb;
        const double* Ht = H + (size_t)tt * a.nw;
        double acc_o = o[t];
        for (int j = 0; j < wn; j++) {
          const int w = w0 + j, ww = a.wlist[w];
          const long long G = 2 * (Gs[wave][lane][j] + off[t]) - Gm[w];
          const double at_tmp = Ht[w] * alpha_k;
          const double gd = (double)G;
          const double tmp = (double)(bkk[t] + gd) / mlt - (double)(at_base + at_tmp) / mlt;
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

void ww_mfma(const WwArgs& a, void* work, size_t work_bytes, hipStream_t st, const WwHook* hook)
{
  if (a.nk <= 0 || a.nt <= 0) return;
  int ndig = (a.B + 7) / 8;
  int nwp = ((a.nw + 31) / 32) * 32;
  long long ldn = (long long)a.prb * 8;
  if (work_bytes < ww_mfma_work_bytes(a.nk, a.nt, a.nw, a.prb, a.B)) { fprintf(stderr, "lkc: ww work buffer too small\n"); return; }
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
  hipLaunchKernelGGL(k_ww_hg, dim3(a.nw + 1), dim3(64), 0, st, a, H, bp, Gm);
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
}

}  // namespace lkc
