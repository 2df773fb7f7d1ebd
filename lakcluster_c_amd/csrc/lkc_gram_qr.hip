// lkc_gram_qr.hip -- the QR stage of the low-rank update fused with the QR row sums.
//
// bcc_lrup_QR_YnWt_stage_0..3 (bcc_lrup.c:863-892) subtract from lf_AnZt (and lf_AnAt)
// the centred Gram of the retained rows over the dropped columns
// (get_AnZt_vv__run, AnZt_vv.c:62-95); the next bcc_lf_AnZt_S_WnYt
// (AnZt_S_WnYt_vv.c:49-67) sums x*S*x of every updated row in ascending z.
// One block owns 64 rows a and walks the z tiles in order: it forms the 64x64 Gram
// tile from column-compacted rows staged in LDS, updates X in place, and adds the
// updated values into its rows' ordered sums, so X is read and written once.
#include "lkc_kernels.h"

namespace lkc {

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
__device__ __forceinline__ int popc_s(uint32_t x) { return (int)__builtin_popcount(x); }

#define GQ_LD (kGramQrMaxW + 1)

__global__ __launch_bounds__(256) void k_gram_qr(GramQrArgs g)
{
  __shared__ uint32_t As[64 * GQ_LD];
  __shared__ uint32_t Zs[64 * GQ_LD];
  __shared__ double Xt[64][65];
  __shared__ signed char sgs[kMaxT][64];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4, wave = tid >> 6, lane = tid & 63;
  const int a0 = blockIdx.x * 64, W = g.W, LD = GQ_LD;
  for (int e = tid; e < 64 * W; e += 256) {
    const int r = e / W, w = e % W;
    As[r * LD + w] = (a0 + r < g.na) ? ((const uint32_t*)(g.A.p + (size_t)g.alist[a0 + r] * g.A.rb))[w] & g.mask[w] : 0u;
  }
  double xav[4];
  int aorig[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int ac = a0 + tx + 16 * i;
    aorig[i] = ac < g.na ? g.alist[ac] : -1;
    xav[i] = ac < g.na ? g.xa[aorig[i]] : 0.0;
  }
  const double a0d1 = g.base16[5], a2d1 = g.base16[7];
  const double* D = g.ajdk + (size_t)5 * g.pcols;
  const int ta = wave, tb = wave + 4;
  double q0 = 0.0, q1 = 0.0;
  for (int z0 = 0; z0 < g.nz; z0 += 64) {
    __syncthreads();
    for (int e = tid; e < 64 * W; e += 256) {
      const int r = e / W, w = e % W;
      Zs[r * LD + w] = (z0 + r < g.nz) ? ((const uint32_t*)(g.Z.p + (size_t)g.zlist[z0 + r] * g.Z.rb))[w] & g.mask[w] : 0u;
    }
    for (int e = tid; e < kMaxT * 64; e += 256) {
      const int t = e / 64, zi = e % 64;
      sgs[t][zi] = (t < g.nt) ? g.sg[(long long)t * g.nzp + z0 + zi] : 0;
    }
    __syncthreads();
    double dsum[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) dsum[i][j] = 0;
    for (int q = 0; q < g.nseg; q++) {
      const int wrel = g.seg[3 * q], wlen = g.seg[3 * q + 1];
      const double wb = D[g.seg[3 * q + 2]];
      int acc[4][4];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
      for (int w = wrel; w < wrel + wlen; w++) {
        uint32_t av[4], zv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) av[i] = As[(tx + 16 * i) * LD + w];
#pragma unroll
        for (int j = 0; j < 4; j++) zv[j] = Zs[(ty + 16 * j) * LD + w];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i][j] += popc_s(av[i] ^ zv[j]);
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) dsum[i][j] += (double)wb * (double)(unsigned)acc[i][j];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int zc = z0 + ty + 16 * j;
      if (zc >= g.nz) continue;
      const int zo = g.zlist[zc];
      const double xzv = g.xz[zo];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (aorig[i] < 0) continue;
        double o = a0d1;
        o -= (2 * dsum[i][j]);
        o -= xav[i] + xzv - a2d1;
        double* dst = g.X + (long long)zo * g.ldx + aorig[i];
        const double nv = *dst - o;
        *dst = nv;
        Xt[ty + 16 * j][tx + 16 * i] = nv;
      }
    }
    __syncthreads();
    const int zn = min(64, g.nz - z0);
    if (ta < g.nt) {
      double s = q0;
      for (int zi = 0; zi < zn; zi++) { const double x = Xt[zi][lane]; const int vS = sgs[ta][zi]; s += x * vS * x; }
      q0 = s;
    }
    if (tb < g.nt) {
      double s = q1;
      for (int zi = 0; zi < zn; zi++) { const double x = Xt[zi][lane]; const int vS = sgs[tb][zi]; s += x * vS * x; }
      q1 = s;
    }
  }
  if (a0 + lane < g.na) {
    const long long a = g.alist[a0 + lane];
    if (ta < g.nt) g.qr[(long long)g.tlist[ta] * g.ldq + a] = q0;
    if (tb < g.nt) g.qr[(long long)g.tlist[tb] * g.ldq + a] = q1;
  }
}

void gram_qr(const GramQrArgs& g, hipStream_t st)
{
  if (g.na <= 0) return;
  hipLaunchKernelGGL(k_gram_qr, dim3(cdiv(g.na, 64)), dim3(256), 0, st, g);
}

}  // namespace lkc
