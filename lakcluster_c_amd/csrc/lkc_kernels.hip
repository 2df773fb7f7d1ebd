// lkc_kernels.hip -- HIP kernels (gfx950) for the lakcluster bcc elimination path.
//
// Integer work (popcounts, +-1 products, quantised-plane sums) may run in any
// order; every double reduction keeps the order of the reference restatement
// cited at its kernel, and the file is compiled with -ffp-contract=off so no
// multiply/add pair is fused.
#include "lkc_kernels.h"
#include <stdio.h>

namespace lkc {

size_t bsize(int n)
{
  long long e = n + ((16 - (n % 16)) % 16);
  long long r = ((e + kPopBits - 1) / kPopBits) * kPopBits;
  return (size_t)(r / 8);
}
int psize(int n) { return (int)(bsize(n) / kPopBytes); }

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// signed popcount: HIP's __popc returns unsigned int, so a difference of two of them
// would wrap before widening
__device__ __forceinline__ int popc(uint32_t x) { return (int)__builtin_popcount(x); }
__device__ __forceinline__ int bit_at(const uint8_t* row, int i) { return (row[i >> 3] >> (7 - (i & 7))) & 1; }
__device__ __forceinline__ int pm_at(const uint8_t* row, int i) { return bit_at(row, i) ? 1 : -1; }

__device__ __forceinline__ int wave_sum_i(int v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------------
// mask_lf: block counts of a mask and the ordered weighted sums
// (dtmp_base An_ajdk_v.c:312, dtmp_a0d1/a2d1 AnZt_vv.c:62-63, lf_A_a*d* bcc_flattenloop.c:1520-1534)
// ------------------------------------------------------------------
__global__ void k_mask_lf(const uint32_t* mask, int pcols, const double* ajdk, double* out16)
{
  extern __shared__ int cnt[];
  for (int b = threadIdx.x; b < pcols; b += blockDim.x) {
    int c = 0;
    for (int w = 0; w < kPopWords; w++) c += popc(mask[(size_t)b * kPopWords + w]);
    cnt[b] = c;
  }
  __syncthreads();
  int jk = threadIdx.x;
  if (jk < kAjdk) {
    if (!ajdk) return;
    double s = 0;
    for (int b = 0; b < pcols; b++) s += (double)ajdk[(size_t)jk * pcols + b] * (double)cnt[b];
    out16[jk] = s;
  } else if (jk == kAjdk) {
    double s = 0;
    for (int b = 0; b < pcols; b++) s += (double)1 * (double)cnt[b];
    out16[kAjdk] = s;
  }
}
void mask_lf(const uint32_t* mask, int pcols, const double* ajdk, double* out16, hipStream_t st)
{
  hipLaunchKernelGGL(k_mask_lf, dim3(1), dim3(64), sizeof(int) * (pcols + 1), st, mask, pcols, ajdk, out16);
}
// several masks in one launch, one block each
__global__ void k_mask_lf_jobs(MaskLfJobs J)
{
  const MaskLfJob& j = J.job[blockIdx.x];
  extern __shared__ int cnt[];
  for (int b = threadIdx.x; b < j.pcols; b += blockDim.x) {
    int c = 0;
    for (int w = 0; w < kPopWords; w++) c += popc(j.mask[(size_t)b * kPopWords + w]);
    cnt[b] = c;
  }
  __syncthreads();
  int jk = threadIdx.x;
  if (jk < kAjdk) {
    if (!j.ajdk) return;
    double s = 0;
    for (int b = 0; b < j.pcols; b++) s += (double)j.ajdk[(size_t)jk * j.pcols + b] * (double)cnt[b];
    j.out16[jk] = s;
  } else if (jk == kAjdk) {
    double s = 0;
    for (int b = 0; b < j.pcols; b++) s += (double)1 * (double)cnt[b];
    j.out16[kAjdk] = s;
  }
}
void mask_lf_jobs(const MaskLfJobs& J, hipStream_t st)
{
  if (J.n <= 0) return;
  int pmax = 0;
  for (int i = 0; i < J.n; i++) pmax = J.job[i].pcols > pmax ? J.job[i].pcols : pmax;
  hipLaunchKernelGGL(k_mask_lf_jobs, dim3(J.n), dim3(64), sizeof(int) * (pmax + 1), st, J);
}

// ------------------------------------------------------------------
// An_ajdk_v (An_ajdk_v.c:272-363): one wave per row
// ------------------------------------------------------------------
// One wave per row; thread b counts block b (15 x 16-byte words, independent loads),
// then one thread per table adds the weighted counts over the blocks in order.
__global__ __launch_bounds__(64) void k_an_ajdk(const uint8_t* X, size_t rb, const int* rows, int nrows,
                                                const uint32_t* mask, int pcols, const double* ajdk,
                                                const double* base, double* out, int ld)
{
  extern __shared__ int cnt[];
  int r = blockIdx.x;
  if (r >= nrows) return;
  int row = rows[r];
  const uint4* xr = (const uint4*)(X + (size_t)row * rb);
  const uint4* mr = (const uint4*)mask;
  for (int b = threadIdx.x; b < pcols; b += 64) {
    int c = 0;
#pragma unroll
    for (int q = 0; q < kPopWords / 4; q++) {
      const uint4 x = xr[(size_t)b * (kPopWords / 4) + q], m = mr[(size_t)b * (kPopWords / 4) + q];
      c += popc(x.x & m.x) + popc(x.y & m.y) + popc(x.z & m.z) + popc(x.w & m.w);
    }
    cnt[b] = c;
  }
  __syncthreads();
  int jk = threadIdx.x;
  if (jk < kAjdk) {
    const double* w = ajdk + (size_t)jk * pcols;
    double d = 0;
    for (int b = 0; b < pcols; b++) d += (double)w[b] * (double)(unsigned)cnt[b];
    double v = -base[jk];
    v += (2 * d);
    out[(size_t)jk * ld + row] = v;
  }
}
void an_ajdk(const Bits& X, const int* rows, int nrows, const uint32_t* mask, int pcols,
             const double* ajdk, const double* base16, double* out, int ld, hipStream_t st)
{
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_an_ajdk, dim3(nrows), dim3(64), sizeof(int) * (pcols + 1), st,
                     X.p, X.rb, rows, nrows, mask, pcols, ajdk, base16, out, ld);
}

// ------------------------------------------------------------------
// AnZt_vv (AnZt_vv.c:62-95): 64x64 pair tile per block, 4x4 pairs per thread,
// one 1920-bit block of both operands staged (masked) in LDS at a time.
// ------------------------------------------------------------------
#define GT 64
#define GCW 60   // words staged per round
#define GLD 61   // padded LDS row stride in words
// Segments: the full layout walks the pcols 1920-bit blocks (60 words each); a
// column-compacted layout (only some columns kept, block by block) passes for each
// non-empty block its word offset, word count (<= 60) and block index.
__global__ __launch_bounds__(256) void k_gram(const uint8_t* A, size_t rbA, const int* alist, int na,
                                              const uint8_t* Z, size_t rbZ, const int* zlist, int nz,
                                              const uint32_t* mask, int pcols, const int* seg, int nseg,
                                              const double* ajdk, const double* base,
                                              const double* xa, const double* xz, int centred,
                                              double* out, long long ldo, int a_orig, int z_orig, int op_sub)
{
  __shared__ uint32_t As[GT * GLD];
  __shared__ uint32_t Zs[GT * GLD];
  int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int a0 = blockIdx.x * GT, z0 = blockIdx.y * GT;
  const double* D = centred ? ajdk + (size_t)5 * pcols : nullptr;
  double dsum[4][4];
  long long isum[4][4];
  for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) { dsum[i][j] = 0; isum[i][j] = 0; }
  const int ns = seg ? nseg : pcols;
  // stage runs of consecutive segments (at most GCW words) at a time
  int q = 0;
  while (q < ns) {
    const int woff0 = seg ? seg[3 * q] : q * kPopWords;
    int q1 = q, wtot = 0;
    while (q1 < ns) {
      const int l = seg ? seg[3 * q1 + 1] : kPopWords;
      if (q1 > q && wtot + l > GCW) break;
      wtot += l; q1++;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < GT * wtot; e += 256) {
      int r = e / wtot, w = e % wtot;
      uint32_t m = mask[woff0 + w];
      uint32_t va = 0, vz = 0;
      if (a0 + r < na) va = ((const uint32_t*)(A + (size_t)alist[a0 + r] * rbA))[woff0 + w] & m;
      if (z0 + r < nz) vz = ((const uint32_t*)(Z + (size_t)zlist[z0 + r] * rbZ))[woff0 + w] & m;
      As[r * GLD + w] = va; Zs[r * GLD + w] = vz;
    }
    __syncthreads();
    for (int qq = q; qq < q1; qq++) {
      const int wrel = (seg ? seg[3 * qq] : qq * kPopWords) - woff0;
      const int wlen = seg ? seg[3 * qq + 1] : kPopWords;
      const int b = seg ? seg[3 * qq + 2] : qq;
      int acc[4][4];
      for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) acc[i][j] = 0;
      for (int w = wrel; w < wrel + wlen; w++) {
        uint32_t av[4], zv[4];
        for (int i = 0; i < 4; i++) av[i] = As[(tx + 16 * i) * GLD + w];
        for (int j = 0; j < 4; j++) zv[j] = Zs[(ty * 4 + j) * GLD + w];
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) acc[i][j] += popc(av[i] ^ zv[j]);
      }
      if (centred) {
        double wb = D[b];
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) dsum[i][j] += (double)wb * (double)(unsigned)acc[i][j];
      } else {
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) isum[i][j] += acc[i][j];
      }
    }
    q = q1;
  }
  double a0d1 = centred ? base[5] : base[kAjdk];
  double a2d1 = centred ? base[7] : 0.0;
  for (int j = 0; j < 4; j++) {
    int zc = z0 + ty * 4 + j; if (zc >= nz) continue;
    for (int i = 0; i < 4; i++) {
      int ac = a0 + tx + 16 * i; if (ac >= na) continue;
      double dtmp = centred ? dsum[i][j] : (double)isum[i][j];
      double o = a0d1;
      o -= (2 * dtmp);
      if (centred) o -= xa[alist[ac]] + xz[zlist[zc]] - a2d1;
      long long ai = a_orig ? alist[ac] : ac, zi = z_orig ? zlist[zc] : zc;
      double* dst = out + zi * ldo + ai;
      if (op_sub) *dst -= o; else *dst = o;
    }
  }
}
void gram(const Bits& A, const int* alist, int na, const Bits& Z, const int* zlist, int nz,
          const uint32_t* mask, int pcols, const double* ajdk, const double* base16,
          const double* xa, const double* xz, int centred,
          double* out, long long ldo, int a_orig, int z_orig, int op_sub, hipStream_t st,
          const int* seg, int nseg)
{
  if (na <= 0 || nz <= 0) return;
  hipLaunchKernelGGL(k_gram, dim3(cdiv(na, GT), cdiv(nz, GT)), dim3(256), 0, st,
                     A.p, A.rb, alist, na, Z.p, Z.rb, zlist, nz, mask, pcols, seg, nseg, ajdk, base16, xa, xz, centred,
                     out, ldo, a_orig, z_orig, op_sub);
}

// column compaction: out word w of row r gathers bits src[32w + q] (q = 0..31, MSB-first
// within each byte; -1 = padding, 0); rows rows[i] (or i) of X -> row i of out.
__global__ void k_compact(const uint8_t* X, size_t rb, const int* rows, int nrows, const int* src, int W,
                          uint32_t* out, int ldw)
{
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nrows * W) return;
  int i = (int)(e / W), w = (int)(e % W);
  const uint8_t* xr = X + (size_t)(rows ? rows[i] : i) * rb;
  uint32_t v = 0;
  for (int q = 0; q < 32; q++) {
    int c = src[32 * w + q];
    if (c >= 0 && bit_at(xr, c)) v |= 1u << ((q >> 3) * 8 + (7 - (q & 7)));
  }
  out[(size_t)i * ldw + w] = v;
}
void compact_cols(const Bits& X, const int* rows, int nrows, const int* src, int W, uint32_t* out, int ldw, hipStream_t st)
{
  long long n = (long long)nrows * W;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_compact, dim3(cdiv(n, 256)), dim3(256), 0, st, X.p, X.rb, rows, nrows, src, W, out, ldw);
}

// ------------------------------------------------------------------
// xcalc (xcalc.c:5-246): min/max, then (u64)(mlt*(x-min)) clamped to [1, D_MLT-1]
// ------------------------------------------------------------------
__device__ __forceinline__ double xval(const double* src, long long sn, long long sw, const int* rl, const int* cl, int i, int j)
{
  long long r = rl ? rl[i] : i, c = cl ? cl[j] : j;
  return src[r * sn + c * sw];
}
__global__ void k_minmax(const double* src, long long sn, long long sw, const int* rl, const int* cl, int nn, int nw, double* part)
{
  __shared__ double smin[256], smax[256]; __shared__ int sset[256];
  long long tot = (long long)nn * nw;
  double mn = 0, mx = 0; int set = 0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int j = (int)(e / nn), i = (int)(e % nn);
    double v = xval(src, sn, sw, rl, cl, i, j);
    if (!set) { mn = v; mx = v; set = 1; } else { mn = v < mn ? v : mn; mx = v > mx ? v : mx; }
  }
  smin[threadIdx.x] = mn; smax[threadIdx.x] = mx; sset[threadIdx.x] = set;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      int o = threadIdx.x + s;
      if (sset[o]) {
        if (!sset[threadIdx.x]) { smin[threadIdx.x] = smin[o]; smax[threadIdx.x] = smax[o]; sset[threadIdx.x] = 1; }
        else { smin[threadIdx.x] = smin[o] < smin[threadIdx.x] ? smin[o] : smin[threadIdx.x];
               smax[threadIdx.x] = smax[o] > smax[threadIdx.x] ? smax[o] : smax[threadIdx.x]; }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[3 * blockIdx.x] = smin[0]; part[3 * blockIdx.x + 1] = smax[0]; part[3 * blockIdx.x + 2] = sset[0]; }
}
// The partials (at most 256, one per k_minmax block, each with a flag for "saw a value")
// fold in a tree over the block's threads: a min / max of the values is the same in any
// order (as within the k_minmax blocks), then one lane derives the quantiser's range.
__global__ void k_minmax_final(const double* part, int np, long long D_MLT, double* scal)
{
  __shared__ double smin[256], smax[256]; __shared__ int sset[256];
  const int t = threadIdx.x;
  smin[t] = t < np ? part[3 * t] : 0.0; smax[t] = t < np ? part[3 * t + 1] : 0.0;
  sset[t] = t < np ? (part[3 * t + 2] != 0) : 0;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      const int o = t + s;
      if (sset[o]) {
        if (!sset[t]) { smin[t] = smin[o]; smax[t] = smax[o]; sset[t] = 1; }
        else { smin[t] = smin[o] < smin[t] ? smin[o] : smin[t]; smax[t] = smax[o] > smax[t] ? smax[o] : smax[t]; }
      }
    }
    __syncthreads();
  }
  if (t != 0) return;
  double mn = sset[0] ? smin[0] : 0.0, mx = sset[0] ? smax[0] : 0.0;
  if (mn >= mx) mx = mn + 1;
  double lo = mn - 0.0625 * (mx - mn);
  double hi = mx + 0.0625 * (mx - mn);
  double mlt = (double)(D_MLT) / (hi - lo);
  scal[0] = mn; scal[1] = mx; scal[2] = lo; scal[3] = hi;
  scal[4] = (!isfinite(mlt) || mlt <= 0) ? 1.0 : mlt;
  scal[5] = mlt;   // the value used for quantisation (before the guard)
}
__global__ void k_quantize(const double* src, long long sn, long long sw, const int* rl, const int* cl, int nn, int nw,
                           const int* nout, long long D_MLT, const double* scal, uint32_t* cz, long long ldc)
{
  long long tot = (long long)nn * nw;
  double lo = scal[2], mlt = scal[5];
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int j = (int)(e / nn), i = (int)(e % nn);
    double v = xval(src, sn, sw, rl, cl, i, j);
    unsigned long long c = (unsigned long long)(mlt * (double)(v - lo));
    if (c < 1ull) c = 1ull;
    if (c > (unsigned long long)(D_MLT - 1)) c = (unsigned long long)(D_MLT - 1);
    cz[(long long)j * ldc + nout[i]] = (uint32_t)c;
  }
}
void xcalc(const double* src, long long sn, long long sw, const int* rlist, const int* clist,
           int nn, int nw, const int* nout, int B, long long D_MLT,
           uint32_t* cz, long long ldc, double* scal, double* work, hipStream_t st)
{
  int np = 256;
  long long tot = (long long)nn * nw;
  if (tot < 256LL * 64) np = (int)cdiv(tot > 0 ? tot : 1, 256);
  hipLaunchKernelGGL(k_minmax, dim3(np), dim3(256), 0, st, src, sn, sw, rlist, clist, nn, nw, work);
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, st, work, np, D_MLT, scal);
  if (tot > 0)
    hipLaunchKernelGGL(k_quantize, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                       src, sn, sw, rlist, clist, nn, nw, nout, D_MLT, scal, cz, ldc);
}

// planes[(j*B+nb)*rb] bit n = bit nb of cz[j*ldc+n]; one thread per 32 n
__global__ void k_cz_planes(const uint32_t* cz, long long ldc, int nw, int n_total, const uint32_t* nmask, int B, uint8_t* planes, size_t rb)
{
  int nwords = (int)(rb / 4);
  long long tot = (long long)nw * nwords;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int j = (int)(e / nwords), wd = (int)(e % nwords);
    uint32_t pw[32];
    for (int nb = 0; nb < 32; nb++) pw[nb] = 0;
    for (int q = 0; q < 32; q++) {
      int n = wd * 32 + q;
      if (n >= n_total) break;
      if (!((nmask[wd] >> ((q >> 3) * 8 + (7 - (q & 7)))) & 1u)) continue;
      uint32_t v = cz[(long long)j * ldc + n];
      uint32_t bitpos = (uint32_t)((q >> 3) * 8 + (7 - (q & 7)));
      for (int nb = 0; nb < B; nb++) pw[nb] |= ((v >> nb) & 1u) << bitpos;
    }
    uint32_t m = nmask[wd];
    for (int nb = 0; nb < B; nb++) ((uint32_t*)(planes + ((size_t)j * B + nb) * rb))[wd] = pw[nb] & m;
  }
}
void cz_planes(const uint32_t* cz, long long ldc, int nw, int n_total, const uint32_t* nmask, int B, uint8_t* planes, size_t rb, hipStream_t st)
{
  long long tot = (long long)nw * (long long)(rb / 4);
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_cz_planes, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     cz, ldc, nw, n_total, nmask, B, planes, rb);
}

// ------------------------------------------------------------------
// At_T_YnWt_S_Zn_ww (At_T_YnWt_S_Zn_ww.c:44-110), bit-plane VALU form.
//   H[t][w] and base'[t] (row0 terms), bk[t][k] (P_k terms), then per (k,w) the
//   integer G and its epilogue term, then the ordered sum over w per (k,t).
// ------------------------------------------------------------------
__device__ __forceinline__ int pmpm0_word(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t m)
{
  uint32_t e = x0 ^ x1, b = x2 & m;
  return popc(~e & b) - popc(e & b);
}
// one wave per (t, w) with w == nw meaning the base' term
__global__ void k_ww_H(WwArgs a, double* H, double* bp)
{
  int t = blockIdx.y, j = blockIdx.x;
  const uint32_t* Tr = (const uint32_t*)(a.Tc.p + (size_t)a.tlist[t] * a.Tc.rb);
  int nwd = (int)(a.prb / 4);
  double mn = a.scal[2], mlt = a.scal[4];
  if (j == a.nw) {
    long long c = 0;
    for (int w = threadIdx.x; w < nwd; w += 64) c += pmpm0_word(a.row0[w], Tr[w], a.row0[w], a.pm[w]);
    c = wave_sum_ll(c);
    if (threadIdx.x == 0) { double d = c * 1.0; bp[t] = d * mn * mlt; }
    return;
  }
  double h = 0; long long n2 = 1;
  for (int nb = 0; nb < a.B; nb++) {
    const uint32_t* pl = (const uint32_t*)(a.planes + ((size_t)j * a.B + nb) * a.prb);
    long long c = 0;
    for (int w = threadIdx.x; w < nwd; w += 64) c += pmpm0_word(a.row0[w], Tr[w], pl[w], a.pm[w]);
    c = wave_sum_ll(c);
    double d = c * 1.0;
    h += n2 * d; n2 *= 2;
  }
  if (threadIdx.x == 0) H[(size_t)t * a.nw + j] = h;
}
// one wave per (t, k)
__global__ void k_ww_bk(WwArgs a, double* bk)
{
  int t = blockIdx.y, k = blockIdx.x;
  const uint32_t* Tr = (const uint32_t*)(a.Tc.p + (size_t)a.tlist[t] * a.Tc.rb);
  const uint32_t* Pr = (const uint32_t*)(a.P.p + (size_t)a.klist[k] * a.P.rb);
  int nwd = (int)(a.prb / 4);
  long long c = 0;
  for (int w = threadIdx.x; w < nwd; w += 64) c += pmpm0_word(Pr[w], Tr[w], a.row0[w], a.pm[w]);
  c = wave_sum_ll(c);
  if (threadIdx.x == 0) bk[(size_t)t * a.nk + k] = c * a.scal[2] * a.scal[4];
}
// one thread per (k, w) for a chunk of w; term[(w-w0)*ntk + t*nk + k]
__global__ __launch_bounds__(256) void k_ww_term_valu(WwArgs a, int t, int w0, int wn, const double* H, const double* bp,
                                                      const double* bk, double* term)
{
  int k = blockIdx.x * 256 + threadIdx.x;
  int jw = blockIdx.y;
  if (k >= a.nk || jw >= wn) return;
  int j = w0 + jw;
  const uint32_t* Tr = (const uint32_t*)(a.Tc.p + (size_t)a.tlist[t] * a.Tc.rb);
  const uint32_t* Pr = (const uint32_t*)(a.P.p + (size_t)a.klist[k] * a.P.rb);
  int nwd = (int)(a.prb / 4);
  long long G = 0;
  for (int w = 0; w < nwd; w++) {
    uint32_t m = a.pm[w];
    uint32_t e = Pr[w] ^ Tr[w];
    uint32_t up = ~e & m, un = e & m;
    for (int nb = 0; nb < a.B; nb++) {
      uint32_t pl = ((const uint32_t*)(a.planes + ((size_t)j * a.B + nb) * a.prb))[w];
      G += (long long)(popc(up & pl) - popc(un & pl)) << nb;
    }
  }
  double mlt = a.scal[4];
  const double* alpha = a.ajdk + (size_t)1 * a.pcols;
  int kk = a.klist[k], qq = a.qlist[k], ww = a.wlist[j];
  double alpha_k = alpha[kk / kPopBits], alpha_q = alpha[qq / kPopBits];
  int vS = pm_at(a.Sc.p + (size_t)a.slist[t] * a.Sc.rb, ww);
  int vZ = pm_at(a.Q.p + (size_t)qq * a.Q.rb, ww);
  double bkk = bk[(size_t)t * a.nk + k];
  double at_base = bp[t] * alpha_k;
  double at_tmp = H[(size_t)t * a.nw + j] * alpha_k;
  double g = (double)G;
  double tmp = (double)(bkk + g) / mlt - (double)(at_base + at_tmp) / mlt;
  term[(size_t)jw * a.nk + k] = tmp * vS * (vZ - alpha_q);
}
__global__ void k_fill0(double* p, long long n)
{
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
// ordered sum over the chunk's w for each k: acc[t*nk+k] += term[...]
__global__ void k_ww_accum(const double* term, int nk, int wn, double* acc)
{
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nk) return;
  double o = acc[k];
  for (int jw = 0; jw < wn; jw++) o += term[(size_t)jw * nk + k];
  acc[k] = o;
}
__global__ void k_ww_store(const double* acc, const int* klist, const int* tlist, int nk, int nt, double* out, long long ldo)
{
  int k = blockIdx.x * 256 + threadIdx.x, t = blockIdx.y;
  if (k >= nk || t >= nt) return;
  out[(long long)tlist[t] * ldo + klist[k]] = acc[(size_t)t * nk + k];
}
size_t ww_work_bytes(int nk, int nt, int nw)
{
  return sizeof(double) * ((size_t)nt * nw + nt + 2 * (size_t)nt * nk + (size_t)64 * (nk > 0 ? nk : 1));
}
hipError_t ww_valu(const WwArgs& a, double* work, size_t work_bytes, hipStream_t st)
{
  if (a.nk <= 0 || a.nt <= 0) return hipSuccess;
  // work layout: H[nt*nw] bp[nt] bk[nt*nk] acc[nt*nk] term[chunk*nk]
  double* H = work;
  double* bp = H + (size_t)a.nt * a.nw;
  double* bk = bp + a.nt;
  double* acc = bk + (size_t)a.nt * a.nk;
  double* term = acc + (size_t)a.nt * a.nk;
  size_t used = (size_t)(term - work) * sizeof(double);
  long long chunk = (long long)((work_bytes - used) / sizeof(double)) / (a.nk > 0 ? a.nk : 1);
  if (chunk > 65535) chunk = 65535;
  if (chunk < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ww_H, dim3(a.nw + 1, a.nt), dim3(64), 0, st, a, H, bp);
  hipLaunchKernelGGL(k_ww_bk, dim3(a.nk, a.nt), dim3(64), 0, st, a, bk);
  hipLaunchKernelGGL(k_fill0, dim3(cdiv((long long)a.nt * a.nk, 256)), dim3(256), 0, st, acc, (long long)a.nt * a.nk);
  for (int t = 0; t < a.nt; t++) {
    for (int w0 = 0; w0 < a.nw; w0 += (int)chunk) {
      int wn = (int)((a.nw - w0) < chunk ? (a.nw - w0) : chunk);
      hipLaunchKernelGGL(k_ww_term_valu, dim3(cdiv(a.nk, 256), wn), dim3(256), 0, st, a, t, w0, wn, H, bp, bk, term);
      hipLaunchKernelGGL(k_ww_accum, dim3(cdiv(a.nk, 256)), dim3(256), 0, st, term, a.nk, wn, acc + (size_t)t * a.nk);
    }
  }
  hipLaunchKernelGGL(k_ww_store, dim3(cdiv(a.nk, 256), a.nt), dim3(256), 0, st, acc, a.klist, a.tlist, a.nk, a.nt, a.out, a.ldo);
  return hipGetLastError();
}

// ------------------------------------------------------------------
// An_ZtSn_ww (An_ZtSn_ww.c:44-70): one wave per row
// ------------------------------------------------------------------
__global__ void k_an_ztsn(const uint8_t* A, size_t rbA, const int* alist, int na, const uint32_t* amask,
                          const uint8_t* planes, size_t prb, const uint32_t* row0, int ns, const int* slist, int B,
                          const double* scal, double* out, long long ldo)
{
  int i = blockIdx.x; if (i >= na) return;
  const uint32_t* Ar = (const uint32_t*)(A + (size_t)alist[i] * rbA);
  int nwd = (int)(prb / 4);
  long long c = 0;
  for (int w = threadIdx.x; w < nwd; w += 64) { uint32_t b = row0[w] & amask[w]; c += popc(Ar[w] & b) - popc(~Ar[w] & b); }
  c = wave_sum_ll(c);
  double mn = scal[2], mlt = scal[4];
  double ll = (double)c;
  double base = ll * mn * mlt;
  for (int s = 0; s < ns; s++) {
    double tmp = 0; long long n2 = 1;
    for (int nb = 0; nb < B; nb++) {
      const uint32_t* pl = (const uint32_t*)(planes + ((size_t)s * B + nb) * prb);
      long long d = 0;
      for (int w = threadIdx.x; w < nwd; w += 64) { uint32_t b = pl[w] & amask[w]; d += popc(Ar[w] & b) - popc(~Ar[w] & b); }
      d = wave_sum_ll(d);
      double l2 = (double)d;
      tmp += n2 * l2; n2 *= 2;
    }
    if (threadIdx.x == 0) out[(long long)slist[s] * ldo + alist[i]] = (double)(base + tmp) / mlt;
  }
}
void an_ztsn(const Bits& A, const int* alist, int na, const uint32_t* amask,
             const uint8_t* planes, size_t prb, const uint32_t* row0, int ns, const int* slist, int B,
             const double* scal, double* out, long long ldo, hipStream_t st)
{
  if (na <= 0) return;
  hipLaunchKernelGGL(k_an_ztsn, dim3(na), dim3(64), 0, st, A.p, A.rb, alist, na, amask, planes, prb, row0, ns, slist, B, scal, out, ldo);
}

// ------------------------------------------------------------------
// AnZt_S_WnYt_vv (AnZt_S_WnYt_vv.c:49-67) with X==Y
// ------------------------------------------------------------------
// Signs of the covariate columns at the listed z: sg[t*nzp + i] = S_t(zlist[i]) (+-1).
__global__ void k_qr_signs(const uint8_t* Sc, size_t rbS, const int* tlist, int nt, const int* zlist, int nz,
                           int nzp, signed char* sg)
{
  long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)nt * nzp) return;
  int t = (int)(e / nzp), i = (int)(e % nzp);
  sg[e] = i < nz ? (signed char)pm_at(Sc + (size_t)tlist[t] * rbS, zlist[i]) : 0;
}

// One block per 16 rows a.  Per chunk of 128 z all 256 threads load X and form the
// reference's products p = (X * S_t) * X (X == Y), one per (t, z, a), into LDS; the
// chain threads (t, a) of the first wave then only add p in ascending z, so the ordered
// chain costs one read and one add per z.  The next chunk's X and signs are loaded into
// registers meanwhile.  x_rows selects the layout X[a*ldx + z] (thread (z, 8 rows a):
// runs along z) or X[z*ldx + a] (thread (a, 8 consecutive z): runs along a).
#define QR_ZC 128
#define QR_AC 16
#define QR_AP 17        // padded row of the product tile
template <int XROWS>
__global__ __launch_bounds__(256) void k_qr(const double* X, long long ldx, const int* alist, int na, const int* zlist,
                                            int nz, const signed char* sg, int nzp, const int* tlist, int nt,
                                            double* out, long long ldo)
{
  extern __shared__ double pr[];          // [t][z][a], nt x QR_ZC x QR_AP
  const int tid = threadIdx.x, i0 = blockIdx.x * QR_AC;
  // load role: XROWS: z = tid % 128, rows a = 8 (tid / 128) + k;  else: a = tid % 16, z = 8 (tid / 16) + k
  const int lz = XROWS ? (tid & 127) : 8 * (tid >> 4), la = XROWS ? 8 * (tid >> 7) : (tid & 15);
  long long arow[8];
#pragma unroll
  for (int k = 0; k < 8; k++) arow[k] = (long long)alist[min(i0 + (XROWS ? la + k : la), na - 1)];
  // chain role: thread (t, a) = (tid / 16, tid % 16)
  const int ct = tid / QR_AC, cc = tid & (QR_AC - 1);
  const bool chain = ct < nt && i0 + cc < na;
  const int nch = (nz + QR_ZC - 1) / QR_ZC;
  double xr[8];
  uint64_t sv[kMaxT];                     // the signs of the thread's z (8 consecutive z: one per byte)
  // every address is valid (clamped); entries past nz or na are never added
  auto load = [&](int c) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int zp = c * QR_ZC + lz + (XROWS ? 0 : k);
      const long long z = zlist[min(zp, nz - 1)];
      xr[k] = XROWS ? X[arow[k] * ldx + z] : X[z * ldx + arow[0]];
    }
#pragma unroll
    for (int t = 0; t < kMaxT; t++) {
      const signed char* sr = sg + (long long)min(t, nt - 1) * nzp + c * QR_ZC + lz;
      sv[t] = XROWS ? (uint64_t)(uint8_t)sr[0] : *(const uint64_t*)sr;
    }
  };
  auto products = [&]() {
#pragma unroll
    for (int t = 0; t < kMaxT; t++) {
      if (t >= nt) break;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int z = XROWS ? lz : lz + k, a = XROWS ? la + k : la;
        const int vS = (int)(signed char)(uint8_t)(sv[t] >> (XROWS ? 0 : 8 * k));
        pr[((size_t)t * QR_ZC + z) * QR_AP + a] = xr[k] * vS * xr[k];
      }
    }
  };
  double acc = 0.0;
  if (nch > 0) load(0);
  // the barriers order LDS only: the next chunk's loads stay in flight while the chains add
  for (int c = 0; c < nch; c++) {
    block_lds_sync();              // the previous chunk's chains are done with pr
    products();
    if (c + 1 < nch) load(c + 1);
    block_lds_sync();
    if (chain) {
      const double* pp = pr + (size_t)ct * QR_ZC * QR_AP + cc;
      const int zc = min(QR_ZC, nz - c * QR_ZC);
      if (zc == QR_ZC) {
        // a whole chunk: the next 16 values are read while the current 16 are added, so
        // the adds do not wait for LDS (same order, one add per z)
        double u[16], v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) u[k] = pp[k * QR_AP];
#pragma unroll
        for (int g = 0; g < QR_ZC / 16; g += 2) {
#pragma unroll
          for (int k = 0; k < 16; k++) v[k] = pp[(16 * (g + 1) + k) * QR_AP];
#pragma unroll
          for (int k = 0; k < 16; k++) acc += u[k];
          if (g + 2 < QR_ZC / 16) {
#pragma unroll
            for (int k = 0; k < 16; k++) u[k] = pp[(16 * (g + 2) + k) * QR_AP];
          }
#pragma unroll
          for (int k = 0; k < 16; k++) acc += v[k];
        }
      } else {
#pragma unroll 16
        for (int zi = 0; zi < zc; zi++) acc += pp[zi * QR_AP];
      }
    }
  }
  if (chain) out[(long long)tlist[ct] * ldo + alist[i0 + cc]] = acc;
}

static int qr_nzp(int nz) { return ((nz + QR_ZC - 1) / QR_ZC) * QR_ZC + QR_ZC; }
void qr(const double* X, long long ldx, const int* alist, int na, const int* zlist, int nz,
        const Bits& Sc, const int* tlist, int nt, double* out, long long ldo, signed char* sgbuf, hipStream_t st,
        int x_rows)
{
  if (na <= 0 || nt <= 0) return;
  const int nzp = qr_nzp(nz);
  hipLaunchKernelGGL(k_qr_signs, dim3(cdiv((long long)nt * nzp, 256)), dim3(256), 0, st, Sc.p, Sc.rb, tlist, nt, zlist, nz, nzp, sgbuf);
  const size_t lds = sizeof(double) * (size_t)nt * QR_ZC * QR_AP;
  if (x_rows)
    hipLaunchKernelGGL((k_qr<1>), dim3(cdiv(na, QR_AC)), dim3(256), lds, st, X, ldx, alist, na, zlist, nz, sgbuf, nzp, tlist, nt, out, ldo);
  else
    hipLaunchKernelGGL((k_qr<0>), dim3(cdiv(na, QR_AC)), dim3(256), lds, st, X, ldx, alist, na, zlist, nz, sgbuf, nzp, tlist, nt, out, ldo);
}
int qr_signs(const Bits& Sc, const int* tlist, int nt, const int* zlist, int nz, signed char* sgbuf, hipStream_t st)
{
  const int nzp = qr_nzp(nz);
  if (nt > 0)
    hipLaunchKernelGGL(k_qr_signs, dim3(cdiv((long long)nt * nzp, 256)), dim3(256), 0, st, Sc.p, Sc.rb, tlist, nt, zlist, nz, nzp, sgbuf);
  return nzp;
}
size_t qr_sign_bytes(int nz, int nt) { return (size_t)nt * qr_nzp(nz) + 256; }

// ------------------------------------------------------------------
// AtTYn_vv (AtTYn_vv.c:79-120)
// ------------------------------------------------------------------
__global__ void k_attyn(const uint8_t* At, size_t rbA, const int* alist, int nk, const uint32_t* amask, int acpop,
                        const uint8_t* Tc, size_t rbT, const int* tlist, int nt, const uint32_t* tmask, int tcpop,
                        const uint8_t* Yt, size_t rbY, const int* ylist, int ny,
                        const double* lf_AtTn, const double* lf_YtTn, long long ldT, const double* ajdk, int pcols, double* out)
{
  int k = blockIdx.x * 64 + (threadIdx.x & 63);
  int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  int t = blockIdx.z;
  if (k >= nk || y >= ny) return;
  const uint32_t* Ar = (const uint32_t*)(At + (size_t)alist[k] * rbA);
  const uint32_t* Tr = (const uint32_t*)(Tc + (size_t)tlist[t] * rbT);
  const uint32_t* Yr = (const uint32_t*)(Yt + (size_t)ylist[y] * rbY);
  int nwd = (int)(rbA / 4);
  long long cnt = 0, tc = 0;
  for (int w = 0; w < nwd; w++) {
    cnt += popc(~(Ar[w] ^ Tr[w] ^ Yr[w]) & amask[w]);
    tc += popc(Tr[w] & tmask[w]);
  }
  long long hhay = -(long long)tcpop + 2 * tc;
  long long hhAy = (long long)lf_AtTn[(long long)tlist[t] * ldT + alist[k]];
  long long hhaY = (long long)lf_YtTn[(long long)tlist[t] * ldT + ylist[y]];
  long long hhAY = +(long long)acpop;
  hhAY -= (long long)(2 * cnt);
  const double* al = ajdk + (size_t)pcols;
  double aa = al[alist[k] / kPopBits], ay = al[ylist[y] / kPopBits];
  out[((size_t)y * nt + t) * nk + k] = (hhAY - hhAy * ay - aa * hhaY + aa * hhay * ay);
}
void attyn(const Bits& At, const int* alist, int nk, const uint32_t* amask, int acpop,
           const Bits& Tc, const int* tlist, int nt, const uint32_t* tmask, int tcpop,
           const Bits& Yt, const int* ylist, int ny, const double* lf_AtTn, const double* lf_YtTn, long long ldT,
           const double* ajdk, int pcols, double* out, hipStream_t st)
{
  if (nk <= 0 || ny <= 0 || nt <= 0) return;
  hipLaunchKernelGGL(k_attyn, dim3(cdiv(nk, 64), cdiv(ny, 4), nt), dim3(256), 0, st,
                     At.p, At.rb, alist, nk, amask, acpop, Tc.p, Tc.rb, tlist, nt, tmask, tcpop,
                     Yt.p, Yt.rb, ylist, ny, lf_AtTn, lf_YtTn, ldT, ajdk, pcols, out);
}
__global__ void k_attyn_wtszn(const double* L1, const double* L2, const int* klist, int nk, const int* ylist, int ny,
                              const int* tlist, int nt, const double* ajdk, int pcols, double* out, long long ldo)
{
  int k = blockIdx.x * 256 + threadIdx.x, t = blockIdx.y;
  if (k >= nk) return;
  const double* D = ajdk + (size_t)5 * pcols;
  double o = 0;
  for (int y = 0; y < ny; y++) {
    size_t q = ((size_t)y * nt + t) * nk + k;
    o += L1[q] * D[ylist[y] / kPopBits] * L2[q];
  }
  out[(long long)tlist[t] * ldo + klist[k]] = o;
}
void attyn_wtszn(const double* L1, const double* L2, const int* klist, int nk, const int* ylist, int ny,
                 const int* tlist, int nt, const double* ajdk, int pcols, double* out, long long ldo, hipStream_t st)
{
  if (nk <= 0 || nt <= 0) return;
  hipLaunchKernelGGL(k_attyn_wtszn, dim3(cdiv(nk, 256), nt), dim3(256), 0, st, L1, L2, klist, nk, ylist, ny, tlist, nt, ajdk, pcols, out, ldo);
}

// ------------------------------------------------------------------
// row popcounts under a mask: one wave per row
// ------------------------------------------------------------------
__global__ void k_popc_rows(const uint8_t* X, size_t rb, const int* rows, int nrows, const uint32_t* mask, int nwords, int* out)
{
  int i = blockIdx.x; if (i >= nrows) return;
  const uint32_t* xr = (const uint32_t*)(X + (size_t)(rows ? rows[i] : i) * rb);
  int c = 0;
  for (int w = threadIdx.x; w < nwords; w += 64) c += popc(xr[w] & mask[w]);
  c = wave_sum_i(c);
  if (threadIdx.x == 0) out[i] = c;
}
void popc_rows(const Bits& X, const int* rows, int nrows, const uint32_t* mask, int nwords, int* out, hipStream_t st)
{
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_popc_rows, dim3(nrows), dim3(64), 0, st, X.p, X.rb, rows, nrows, mask, nwords, out);
}

// Bit-matrix transpose (the on-device counterpart of wrap_transpose, binary_read.c:174-269):
// Y row j, bit i = X row i, bit j.  One wave per 64x64 tile: lane i holds 64 bits of
// row r0 + i; a ballot per column gives that column's 64 row bits (lane i -> bit i),
// which bit-reverse + byte-swap turn into the packed MSB-first order.
__global__ __launch_bounds__(256) void k_transpose_bits(const uint8_t* X, size_t rbX, int nr, int nc,
                                                        uint8_t* Y, size_t rbY)
{
  const int lane = threadIdx.x & 63;
  const long long r0 = 64LL * blockIdx.y, c0 = 64LL * (blockIdx.x * 4 + (threadIdx.x >> 6));
  if (c0 >= (long long)rbX * 8) return;
  const long long r = r0 + lane;
  const uint64_t x = r < nr ? *(const uint64_t*)(X + (size_t)r * rbX + c0 / 8) : 0ull;
  uint64_t mine = 0;
#pragma unroll 8
  for (int j = 0; j < 64; j++) {
    const uint64_t m = __ballot((x >> (8 * (j >> 3) + 7 - (j & 7))) & 1ull);
    if (lane == j) mine = m;
  }
  const long long cj = c0 + lane;
  if (cj < nc) *(uint64_t*)(Y + (size_t)cj * rbY + r0 / 8) = __builtin_bswap64(__builtin_bitreverse64(mine));
}
void transpose_bits(const Bits& X, uint8_t* Y, size_t rbY, hipStream_t st)
{
  if (X.nr <= 0 || X.nc <= 0) return;
  const long long tiles_c = ((long long)X.rb * 8 + 63) / 64;
  dim3 grid((unsigned)((tiles_c + 3) / 4), (unsigned)((X.nr + 63) / 64));
  hipLaunchKernelGGL(k_transpose_bits, grid, dim3(256), 0, st, X.p, X.rb, X.nr, X.nc, Y, rbY);
}

// Row gathers: dst row i = src row rows[i] (rb bytes, a multiple of 16; rows[i] < 0
// gives a zero row), and dst row t bit i = src row t bit idx[i] (bits past n stay zero).
// With the bit transpose the first also compacts columns: the columns of X are the rows
// of its transposed copy.
__global__ void k_gather_rows(const uint8_t* src, size_t rb, const int* rows, int n, uint8_t* dst)
{
  const long long per = (long long)(rb / 16), tot = (long long)n * per;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const long long i = e / per, q = e % per;
    const int ri = rows[i];
    ((uint4*)(dst + (size_t)i * rb))[q] = ri >= 0 ? ((const uint4*)(src + (size_t)ri * rb))[q] : make_uint4(0, 0, 0, 0);
  }
}
void gather_rows(const Bits& src, const int* rows, int n, uint8_t* dst, hipStream_t st)
{
  const long long tot = (long long)n * (long long)(src.rb / 16);
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     src.p, src.rb, rows, n, dst);
}
__global__ void k_gather_bits(const uint8_t* src, size_t rbs, int nr, const int* idx, int n, uint8_t* dst, size_t rbd)
{
  const long long per = (long long)(rbd / 4), tot = (long long)nr * per;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const long long t = e / per, w = e % per;
    const uint8_t* sr = src + (size_t)t * rbs;
    uint32_t v = 0;
    for (int q = 0; q < 32; q++) {
      const long long i = 32 * w + q;
      if (i < n && bit_at(sr, idx[i])) v |= 1u << ((q >> 3) * 8 + (7 - (q & 7)));
    }
    ((uint32_t*)(dst + (size_t)t * rbd))[w] = v;
  }
}
void gather_bits(const Bits& src, const int* idx, int n, uint8_t* dst, size_t rbd, hipStream_t st)
{
  const long long tot = (long long)src.nr * (long long)(rbd / 4);
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_gather_bits, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     src.p, src.rb, src.nr, idx, n, dst, rbd);
}

// Plain +-1 products of many rows with a few covariate rows (AnZt_vv with Z = T,
// AnZt_vv.c:62-95): out[tlist[t]*ldo + alist[i]] = base[15] - 2 pop((A_i ^ T_t) & mask).
// One wave per row, 4 rows per block; the covariate rows stay in L1/L2.
__global__ __launch_bounds__(256) void k_pm_few(const uint8_t* A, size_t rbA, const int* alist, int na,
                                                const uint8_t* T, size_t rbT, const int* tlist, int nt,
                                                const uint32_t* mask, int nwords, const double* base,
                                                double* out, long long ldo)
{
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= na) return;
  const uint32_t* ar = (const uint32_t*)(A + (size_t)alist[i] * rbA);
  const uint32_t* tr[kMaxT];
#pragma unroll
  for (int t = 0; t < kMaxT; t++) tr[t] = (const uint32_t*)(T + (size_t)tlist[min(t, nt - 1)] * rbT);
  int c[kMaxT];
#pragma unroll
  for (int t = 0; t < kMaxT; t++) c[t] = 0;
  for (int w = lane; w < nwords; w += 64) {
    const uint32_t a = ar[w], m = mask[w];
#pragma unroll
    for (int t = 0; t < kMaxT; t++) if (t < nt) c[t] += popc((a ^ tr[t][w]) & m);
  }
#pragma unroll
  for (int t = 0; t < kMaxT; t++) if (t < nt) c[t] = wave_sum_i(c[t]);
  if (lane == 0) {
    const double a0 = base[kAjdk];
#pragma unroll
    for (int t = 0; t < kMaxT; t++)
      if (t < nt) {
        double o = a0;
        o -= (2 * (double)c[t]);
        out[(long long)tlist[t] * ldo + alist[i]] = o;
      }
  }
}
// bcc_scramble's new bits: bit q of bits[u*wpr + q/32] becomes bit (rows[u], cols[q]) of
// An and bit (cols[q], rows[u]) of At.  Rows and columns are unique, so every bit has one
// writer; neighbours in a word are written with atomic and / or.
__device__ __forceinline__ void set_bit_word(uint32_t* w, int i, bool v)
{
  const uint32_t m = 1u << (8 * ((i & 31) >> 3) + 7 - (i & 7));   // MSB first in each byte
  if (v) atomicOr(w + (i >> 5), m); else atomicAnd(w + (i >> 5), ~m);
}
__global__ void k_scramble_bits(uint8_t* An, size_t rbA, uint8_t* At, size_t rbT, const int* rows, int nr,
                                const int* cols, int nc, const uint32_t* bits)
{
  const int wpr = (nc + 31) / 32;
  const long long tot = (long long)nr * nc;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int u = (int)(e / nc), q = (int)(e % nc);
    const bool v = (bits[(size_t)u * wpr + q / 32] >> (q % 32)) & 1u;
    const int a = rows[u], k = cols[q];
    set_bit_word((uint32_t*)(An + (size_t)a * rbA), k, v);
    set_bit_word((uint32_t*)(At + (size_t)k * rbT), a, v);
  }
}
void scramble_bits(uint8_t* An, size_t rbA, uint8_t* At, size_t rbT, const int* rows, int nr, const int* cols, int nc,
                   const uint32_t* bits, hipStream_t st)
{
  const long long tot = (long long)nr * nc;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_scramble_bits, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     An, rbA, At, rbT, rows, nr, cols, nc, bits);
}

void pm_few(const Bits& A, const int* alist, int na, const Bits& T, const int* tlist, int nt,
            const uint32_t* mask, int pcols, const double* base16, double* out, long long ldo, hipStream_t st,
            int nwords)
{
  if (na <= 0 || nt <= 0) return;
  hipLaunchKernelGGL(k_pm_few, dim3(cdiv(na, 4)), dim3(256), 0, st, A.p, A.rb, alist, na, T.p, T.rb, tlist, nt,
                     mask, nwords >= 0 ? nwords : pcols * kPopWords, base16, out, ldo);
}

// ------------------------------------------------------------------
// Synthetic patient pools drawn on the device.  A counter-based generator: the uniform of
// (seed, row, column) is the splitmix64 finaliser (Steele, Lea, Flood 2014) of a Weyl-style
// combination of the three, so any entry can be recomputed on its own and a pool does not
// depend on launch shape.  Row ~0u gives the column frequencies.
// ------------------------------------------------------------------
__device__ __host__ inline double pool_uniform(unsigned long long seed, unsigned int row, unsigned int col)
{
  unsigned long long x = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)row * 0xD1B54A32D192ED03ull +
                         ((unsigned long long)col + 1ull) * 0xABC98388FB8FAC03ull;
  x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27; x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}
__global__ void k_pool_p(double* p, int ncols, double pmin, double pmax, unsigned long long seed)
{
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < ncols; j += gridDim.x * blockDim.x)
    p[j] = pmin + (pmax - pmin) * pool_uniform(seed, 0xFFFFFFFFu, (unsigned)j);
}
// one 32-bit word (32 columns, MSB-first bytes) per thread; padding bits stay 0
__global__ void k_pool_draw(uint32_t* rows, long long words_per_row, int nrows, int ncols, const double* p,
                            unsigned long long seed)
{
  const long long tot = words_per_row * nrows;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / words_per_row);
    const int w = (int)(e % words_per_row);
    uint32_t v = 0;
    for (int q = 0; q < 32; q++) {
      const int j = w * 32 + q;
      if (j >= ncols) break;
      if (pool_uniform(seed, (unsigned)r, (unsigned)j) < p[j]) v |= 1u << (((q >> 3) << 3) + (7 - (q & 7)));
    }
    rows[e] = v;
  }
}
// overwrite an m x mc block: bit (prow[i], pcol[c]) = bits[i*mc + c]
__global__ void k_pool_plant(uint32_t* rows, long long words_per_row, const int* prow, const int* pcol,
                             const uint8_t* bits, int m, int mc)
{
  const long long tot = (long long)m * mc;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / mc), c = (int)(e % mc);
    const int j = pcol[c];
    uint32_t* wd = rows + (long long)prow[i] * words_per_row + (j >> 5);
    const uint32_t b = 1u << (((j >> 3) & 3) * 8 + (7 - (j & 7)));
    if (bits[e]) atomicOr(wd, b); else atomicAnd(wd, ~b);
  }
}
void pool_draw(uint8_t* rows, size_t rb, int nrows, int ncols, double pmin, double pmax, unsigned long long seed,
               double* pbuf, hipStream_t st)
{
  hipLaunchKernelGGL(k_pool_p, dim3(cdiv(ncols, 256) < 4096 ? cdiv(ncols, 256) : 4096), dim3(256), 0, st,
                     pbuf, ncols, pmin, pmax, seed);
  const long long tot = (long long)(rb / 4) * nrows;
  hipLaunchKernelGGL(k_pool_draw, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     (uint32_t*)rows, (long long)(rb / 4), nrows, ncols, pbuf, seed);
}
void pool_plant(uint8_t* rows, size_t rb, const int* prow, const int* pcol, const uint8_t* bits, int m, int mc,
                hipStream_t st)
{
  const long long tot = (long long)m * mc;
  if (tot <= 0) return;
  hipLaunchKernelGGL(k_pool_plant, dim3(cdiv(tot, 256) < 65536 ? cdiv(tot, 256) : 65536), dim3(256), 0, st,
                     (uint32_t*)rows, (long long)(rb / 4), prow, pcol, bits, m, mc);
}

}  // namespace lkc
