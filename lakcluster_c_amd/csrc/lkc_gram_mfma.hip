// lkc_gram_mfma.hip -- centred AnZt_vv Grams (AnZt_vv.c:62-95) on the int8 matrix cores.
//
// The reference forms, per pair (a, z), dtmp = sum_b D_b pop((a ^ z) & m)_b over the
// 1920-bit blocks b in order.  Per block
//     pop((a ^ z) & m)_b = |a & m|_b + |z & m|_b - 2 sum_{n in b} (a & m)(n) z(n)
// and the last sum is an int8 MFMA contraction of 0/1 values.  The kernel reads the
// packed bit rows (1 bit per column, 8x less traffic than int8 rows) and spreads
// every 16-bit half word into 16 bytes in registers; A and B operands use the same
// bit-to-slot map, so the contraction is exact.  The double sum over the blocks keeps
// the reference's order; the epilogue is the reference's expression.
#include "lkc_kernels.h"

namespace lkc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// cnt[q*ld + i] = pop(X row rows[i] & mask) over the words of segment q
// (seg == NULL: segment q = block q, words 60q .. 60q+59; mask == NULL: no mask)
__global__ __launch_bounds__(64) void k_seg_pop(const uint8_t* X, size_t rb, const int* rows, int nrows,
                                                const uint32_t* mask, const int* seg, int nseg, int* cnt, long long ld)
{
  const int i = blockIdx.x;
  if (i >= nrows) return;
  const uint32_t* xr = (const uint32_t*)(X + (size_t)(rows ? rows[i] : i) * rb);
  // thread q counts segment q (words contiguous, independent loads; 16-byte loads for the
  // full 60-word blocks, single words for the compacted segments of any length)
  for (int q = threadIdx.x; q < nseg; q += 64) {
    const int w0 = seg ? seg[3 * q] : kPopWords * q, nw = seg ? seg[3 * q + 1] : kPopWords;
    int c = 0;
    if (!seg) {
      for (int w = 0; w < nw; w += 4) {
        const uint4 x = *(const uint4*)(xr + w0 + w);
        uint4 m = make_uint4(~0u, ~0u, ~0u, ~0u);
        if (mask) m = *(const uint4*)(mask + w0 + w);
        c += __builtin_popcount(x.x & m.x) + __builtin_popcount(x.y & m.y) + __builtin_popcount(x.z & m.z) +
             __builtin_popcount(x.w & m.w);
      }
    } else {
      for (int w = 0; w < nw; w++) c += __builtin_popcount(xr[w0 + w] & (mask ? mask[w0 + w] : ~0u));
    }
    cnt[(long long)q * ld + i] = c;
  }
}
void seg_counts(const Bits& X, const int* rows, int nrows, const uint32_t* mask, const int* seg, int nseg,
                int* cnt, long long ld, hipStream_t st)
{
  if (nrows <= 0 || nseg <= 0) return;
  hipLaunchKernelGGL(k_seg_pop, dim3(nrows), dim3(64), 0, st, X.p, X.rb, rows, nrows, mask, seg, nseg, cnt, ld);
}

// 16 bits -> 16 bytes 0/1: byte 4d + j of the fragment = bit 4d + j of h
__device__ __forceinline__ v4i spread16(uint32_t h)
{
  v4i f;
#pragma unroll
  for (int d = 0; d < 4; d++) f[d] = (int)((((h >> (4 * d)) & 15u) * 0x204081u) & 0x01010101u);
  return f;
}

#define GM_QG 48   // segments whose counts are staged in LDS at a time

// Block tile: 64 WI z (M, the MFMA rows) x 128 a (N, the MFMA columns); 4 waves of
// (32 WI) x 64.
// Columns a run along the lanes, so the update of out[z*ldo + a] is coalesced.
// Segment lengths are multiples of 4 words (128 columns, four MFMA steps): lane (r, h)
// loads words 2h and 2h+1 of its rows and feeds one half word to each step, the lanes
// h = 0 and 1 covering the four words.  Two chunks are in flight while one is used.
template <int OPSUB, int WI>
__global__ __launch_bounds__(256, 3 - WI) void k_gram_mfma(GramMArgs g)
{
  constexpr int BZ = 64 * WI;   // block rows z (the waves wm = 0, 1 take 32 WI each)
  __shared__ int czs[GM_QG][BZ], cas[GM_QG][128];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, hg = lane >> 5, r = lane & 31;
  const int wm = wave & 1, wn = wave >> 1;
  // XCD-aware tile order: consecutive block ids go to different XCDs; each XCD takes a band
  const int ntx = (g.na + 127) / 128, nty = (g.nz + BZ - 1) / BZ;
  const int nb = ntx * nty, bid = blockIdx.x;
  const int per_xcd = (nb + 7) / 8;
  const int lin = (bid % 8) * per_xcd + bid / 8;
  if (lin >= nb) return;
  const int tyi = lin / ntx, txi = lin % ntx;
  const int zt0 = tyi * BZ, at0 = txi * 128;
  const int zb = zt0 + 32 * WI * wm, ab = at0 + 64 * wn;
  // lane (r, h) reads words 2h, 2h+1 of every 4-word chunk of its rows
  const uint2* Zr[WI];
  const uint2* Ar[2];
#pragma unroll
  for (int i = 0; i < WI; i++) Zr[i] = (const uint2*)(g.Z + (size_t)g.zlist[min(zb + 32 * i + r, g.nz - 1)] * g.rbZ) + hg;
#pragma unroll
  for (int j = 0; j < 2; j++) Ar[j] = (const uint2*)(g.A + (size_t)g.alist[min(ab + 32 * j + r, g.na - 1)] * g.rbA) + hg;
  const uint2* mk = g.mask ? (const uint2*)g.mask + hg : nullptr;
  const double* D = g.ajdk + (size_t)5 * g.pcols;
  double dsum[WI][2][16];
#pragma unroll
  for (int i = 0; i < WI; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) dsum[i][j][v] = 0.0;
  // total chunks (4 words each); two chunks are in flight ahead of the one in use
  const int wtot = g.seg ? (g.nseg > 0 ? g.seg[3 * (g.nseg - 1)] + g.seg[3 * (g.nseg - 1) + 1] : 0) : kPopWords * g.nseg;
  const int nch = wtot / 4;
  uint2 z1[WI], a1[2], m1, z2[WI], a2[2], m2;
  auto fetch = [&](int c, uint2* zz, uint2* aa, uint2& mm) {
    const int cc = 2 * min(c, nch - 1);
#pragma unroll
    for (int i = 0; i < WI; i++) zz[i] = Zr[i][cc];
#pragma unroll
    for (int j = 0; j < 2; j++) aa[j] = Ar[j][cc];
    // (a literal, not a captured variable: a pointer select would send the load to scratch)
    mm = mk ? mk[cc] : make_uint2(~0u, ~0u);
  };
  if (nch > 0) { fetch(0, z1, a1, m1); fetch(1, z2, a2, m2); }
  int ch = 0;
  for (int q = 0; q < g.nseg; q++) {
    const int qq = q % GM_QG;
    if (qq == 0) {
      __syncthreads();
      for (int e = tid; e < GM_QG * 128; e += 256) {
        const int s = e / 128, p = e % 128, qs = q + s;
        int va = 0;
        if (qs < g.nseg && at0 + p < g.na) va = g.ca[(long long)qs * g.ldca + at0 + p];
        cas[s][p] = va;
      }
      for (int e = tid; e < GM_QG * BZ; e += 256) {
        const int s = e / BZ, p = e % BZ, qs = q + s;
        int vz = 0;
        if (qs < g.nseg && zt0 + p < g.nz) vz = g.cz[(long long)qs * g.ldcz + zt0 + p];
        czs[s][p] = vz;
      }
      __syncthreads();
    }
    const int qch = g.seg ? g.seg[3 * q + 1] / 4 : kPopWords / 4;
    // The accumulators start at |z&m|_q + |a&m|_q and the z operand enters as -2 z, so
    // the contraction leaves x_q = pop((a ^ z) & m)_q = |z&m| + |a&m| - 2 sum a z itself.
    v16i acc[WI][2];
    {
      int caj[2];
#pragma unroll
      for (int j = 0; j < 2; j++) caj[j] = cas[qq][64 * wn + 32 * j + r];
#pragma unroll
      for (int i = 0; i < WI; i++)
#pragma unroll
        for (int v4 = 0; v4 < 4; v4++) {
          const int4 cz4 = *(const int4*)&czs[qq][32 * WI * wm + 32 * i + 8 * v4 + 4 * hg];
          const int czv[4] = { cz4.x, cz4.y, cz4.z, cz4.w };
#pragma unroll
          for (int u = 0; u < 4; u++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j][4 * v4 + u] = czv[u] + caj[j];
        }
    }
    for (int c = 0; c < qch; c++, ch++) {
      uint32_t zw[WI][2];
#pragma unroll
      for (int i = 0; i < WI; i++) { zw[i][0] = z1[i].x; zw[i][1] = z1[i].y; z1[i] = z2[i]; }
      const uint2 ac0 = a1[0], ac1 = a1[1], mc = m1;
      a1[0] = a2[0]; a1[1] = a2[1]; m1 = m2;
      if (ch + 2 < nch) fetch(ch + 2, z2, a2, m2);
      const uint32_t aw[2][2] = { { ac0.x & mc.x, ac0.y & mc.y }, { ac1.x & mc.x, ac1.y & mc.y } };
      // step p takes half (p & 1) of word (p >> 1) of the lane's pair
#pragma unroll
      for (int p = 0; p < 4; p++) {
        v4i af[WI], bf[2];
#pragma unroll
        for (int i = 0; i < WI; i++) {
          af[i] = spread16((zw[i][p >> 1] >> (16 * (p & 1))) & 0xFFFFu);
#pragma unroll
          for (int d = 0; d < 4; d++) af[i][d] = (int)((uint32_t)af[i][d] * 0xFEu);   // bytes 0 / -2
        }
#pragma unroll
        for (int j = 0; j < 2; j++) bf[j] = spread16((aw[j][p >> 1] >> (16 * (p & 1))) & 0xFFFFu);
#pragma unroll
        for (int i = 0; i < WI; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
    }
    // segment epilogue: dtmp += D_b * pop((a ^ z) & m)_b
    const double wb = D[g.seg ? g.seg[3 * q + 2] : q];
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int v = 0; v < 16; v++) dsum[i][j][v] += (double)wb * (double)(unsigned)acc[i][j][v];
  }
  // Epilogue.  Every load below reads a valid (clamped) address without a branch, so
  // they all go out together; only the stores are predicated.
  const double a0d1 = g.base16[5], a2d1 = g.base16[7];
  int zrow[WI][16];
  double xzv[WI][16];
#pragma unroll
  for (int i = 0; i < WI; i++)
#pragma unroll
    for (int v = 0; v < 16; v++) {
      zrow[i][v] = g.zlist[min(zb + 32 * i + 8 * (v / 4) + 4 * hg + (v % 4), g.nz - 1)];
      xzv[i][v] = g.xz[zrow[i][v]];
    }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int ac = ab + 32 * j + r, acl = min(ac, g.na - 1);
    const int aor = g.alist[acl];
    const double xav = g.xa[aor];
    const long long aidx = g.a_orig ? aor : acl;
    double val[WI][16];
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int zc = min(zb + 32 * i + 8 * (v / 4) + 4 * hg + (v % 4), g.nz - 1);
        double o = a0d1;
        o -= (2 * dsum[i][j][v]);
        o -= xav + xzv[i][v] - a2d1;
        const long long zidx = g.z_orig ? zrow[i][v] : zc;
        val[i][v] = o;
        if (OPSUB) val[i][v] = g.out[zidx * g.ldo + aidx] - o;
      }
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int zc = zb + 32 * i + 8 * (v / 4) + 4 * hg + (v % 4);
        const long long zidx = g.z_orig ? zrow[i][v] : min(zc, g.nz - 1);
        if (ac < g.na && zc < g.nz) g.out[zidx * g.ldo + aidx] = val[i][v];
      }
  }
}

// The same Gram over a compacted layout (segments of any whole number of words): one MFMA
// step per 32-column word, the lanes of half h taking half word h, so a segment costs
// only the words it holds.  Rows are read as aligned 4-word groups, two groups ahead;
// a segment may end at any word of a group, where its double term is added (the segment
// order, hence the double order, is the reference's).  Block tile 64 z x 128 a, waves of
// 32 x 64, two blocks per CU.
template <int OPSUB>
__global__ __launch_bounds__(256, 2) void k_gram_seg(GramMArgs g)
{
  __shared__ int czs[GM_QG][64], cas[GM_QG][128];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, hg = lane >> 5, r = lane & 31;
  const int wm = wave & 1, wn = wave >> 1;
  const int ntx = (g.na + 127) / 128, nty = (g.nz + 63) / 64;
  const int nb = ntx * nty, bid = blockIdx.x;
  const int per_xcd = (nb + 7) / 8;
  const int lin = (bid % 8) * per_xcd + bid / 8;
  if (lin >= nb || g.nseg <= 0) return;
  const int tyi = lin / ntx, txi = lin % ntx;
  const int zt0 = tyi * 64, at0 = txi * 128;
  const int zb = zt0 + 32 * wm, ab = at0 + 64 * wn;
  const uint4* Zr = (const uint4*)(g.Z + (size_t)g.zlist[min(zb + r, g.nz - 1)] * g.rbZ);
  const uint4* Ar[2];
#pragma unroll
  for (int j = 0; j < 2; j++) Ar[j] = (const uint4*)(g.A + (size_t)g.alist[min(ab + 32 * j + r, g.na - 1)] * g.rbA);
  const uint4* mk = (const uint4*)g.mask;
  const double* D = g.ajdk + (size_t)5 * g.pcols;
  double dsum[2][16];
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int v = 0; v < 16; v++) dsum[j][v] = 0.0;
  const int wtot = g.seg[3 * (g.nseg - 1)] + g.seg[3 * (g.nseg - 1) + 1];
  const int ngr = (wtot + 3) / 4;
  // group buffers: current (1) and next (2); the row stride holds whole groups
  uint4 z1, z2, a1[2], a2[2], m1, m2;
  auto fetch = [&](int c, uint4& zz, uint4* aa, uint4& mm) {
    const int cc = min(c, ngr - 1);
    zz = Zr[cc];
#pragma unroll
    for (int j = 0; j < 2; j++) aa[j] = Ar[j][cc];
    mm = mk ? mk[cc] : make_uint4(~0u, ~0u, ~0u, ~0u);
  };
  fetch(0, z1, a1, m1);
  fetch(1, z2, a2, m2);
  auto stage = [&](int q) {
    __syncthreads();
    for (int e = tid; e < GM_QG * 128; e += 256) {
      const int s_ = e / 128, p = e % 128, qs = q + s_;
      int va = 0;
      if (qs < g.nseg && at0 + p < g.na) va = g.ca[(long long)qs * g.ldca + at0 + p];
      cas[s_][p] = va;
    }
    for (int e = tid; e < GM_QG * 64; e += 256) {
      const int s_ = e / 64, p = e % 64, qs = q + s_;
      int vz = 0;
      if (qs < g.nseg && zt0 + p < g.nz) vz = g.cz[(long long)qs * g.ldcz + zt0 + p];
      czs[s_][p] = vz;
    }
    __syncthreads();
  };
  v16i acc[2];
  // accumulators start at |z&m|_q + |a&m|_q; the z operand enters as -2 z
  auto init = [&](int q) {
    const int qq = q % GM_QG;
    int caj[2];
#pragma unroll
    for (int j = 0; j < 2; j++) caj[j] = cas[qq][64 * wn + 32 * j + r];
#pragma unroll
    for (int v4 = 0; v4 < 4; v4++) {
      const int4 cz4 = *(const int4*)&czs[qq][32 * wm + 8 * v4 + 4 * hg];
      const int czv[4] = { cz4.x, cz4.y, cz4.z, cz4.w };
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[j][4 * v4 + u] = czv[u] + caj[j];
    }
  };
  int q = 0, qend = g.seg[1];
  stage(0);
  init(0);
  for (int c = 0; c < ngr; c++) {
    const uint4 zc = z1, ac0 = a1[0], ac1 = a1[1], mc = m1;
    z1 = z2; a1[0] = a2[0]; a1[1] = a2[1]; m1 = m2;
    if (c + 2 < ngr) fetch(c + 2, z2, a2, m2);
    const uint32_t zw[4] = { zc.x, zc.y, zc.z, zc.w };
    const uint32_t aw[2][4] = { { ac0.x & mc.x, ac0.y & mc.y, ac0.z & mc.z, ac0.w & mc.w },
                                { ac1.x & mc.x, ac1.y & mc.y, ac1.z & mc.z, ac1.w & mc.w } };
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int w = 4 * c + k;
      if (w >= wtot) break;
      v4i af = spread16((zw[k] >> (16 * hg)) & 0xFFFFu);
#pragma unroll
      for (int d = 0; d < 4; d++) af[d] = (int)((uint32_t)af[d] * 0xFEu);   // bytes 0 / -2
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, spread16((aw[j][k] >> (16 * hg)) & 0xFFFFu), acc[j], 0, 0, 0);
      if (w + 1 == qend) {
        // segment q ends: dtmp += D_b * pop((a ^ z) & m)_b
        const double wb = D[g.seg[3 * q + 2]];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int v = 0; v < 16; v++) dsum[j][v] += (double)wb * (double)(unsigned)acc[j][v];
        q++;
        if (q < g.nseg) {
          if (q % GM_QG == 0) stage(q);
          qend = g.seg[3 * q] + g.seg[3 * q + 1];
          init(q);
        }
      }
    }
  }
  // epilogue (clamped loads, predicated stores), as in k_gram_mfma
  const double a0d1 = g.base16[5], a2d1 = g.base16[7];
  int zrow[16];
  double xzv[16];
#pragma unroll
  for (int v = 0; v < 16; v++) {
    zrow[v] = g.zlist[min(zb + 8 * (v / 4) + 4 * hg + (v % 4), g.nz - 1)];
    xzv[v] = g.xz[zrow[v]];
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int ac = ab + 32 * j + r, acl = min(ac, g.na - 1);
    const int aor = g.alist[acl];
    const double xav = g.xa[aor];
    const long long aidx = g.a_orig ? aor : acl;
    double val[16];
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int zc = min(zb + 8 * (v / 4) + 4 * hg + (v % 4), g.nz - 1);
      double o = a0d1;
      o -= (2 * dsum[j][v]);
      o -= xav + xzv[v] - a2d1;
      const long long zidx = g.z_orig ? zrow[v] : zc;
      val[v] = o;
      if (OPSUB) val[v] = g.out[zidx * g.ldo + aidx] - o;
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int zc = zb + 8 * (v / 4) + 4 * hg + (v % 4);
      const long long zidx = g.z_orig ? zrow[v] : min(zc, g.nz - 1);
      if (ac < g.na && zc < g.nz) g.out[zidx * g.ldo + aidx] = val[v];
    }
  }
}

// Split over the blocks, for Grams with few output tiles (the kept-column Grams of the
// removed rows): block (tile, q) contracts block q's 60 words alone and stores the
// integers x_q = |a&m|_q + |z&m|_q - 2 sum_q a z; a second kernel adds D_q x_q over q in
// order per pair and applies the epilogue, so the double arithmetic is unchanged.
// X[(q*nz + zpos)*na + apos].
__global__ __launch_bounds__(256) void k_gram_split(GramMArgs g, uint32_t* X)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, hg = lane >> 5, r = lane & 31;
  const int wm = wave & 1, wn = wave >> 1;
  const int ntx = (g.na + 127) / 128;
  const int tyi = blockIdx.x / ntx, txi = blockIdx.x % ntx, q = blockIdx.y;
  const int zb = tyi * 128 + 64 * wm, ab = txi * 128 + 64 * wn;
  const uint2* Zr[2];
  const uint2* Ar[2];
#pragma unroll
  for (int i = 0; i < 2; i++) Zr[i] = (const uint2*)(g.Z + (size_t)g.zlist[min(zb + 32 * i + r, g.nz - 1)] * g.rbZ) + hg;
#pragma unroll
  for (int j = 0; j < 2; j++) Ar[j] = (const uint2*)(g.A + (size_t)g.alist[min(ab + 32 * j + r, g.na - 1)] * g.rbA) + hg;
  const uint2* mk = (const uint2*)g.mask + hg;
  v16i acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[i][j][v] = 0;
  const int c0 = q * (kPopWords / 4);
  uint2 zn[2], an[2], mn;
#pragma unroll
  for (int i = 0; i < 2; i++) zn[i] = Zr[i][2 * c0];
#pragma unroll
  for (int j = 0; j < 2; j++) an[j] = Ar[j][2 * c0];
  mn = mk[2 * c0];
  for (int c = 0; c < kPopWords / 4; c++) {
    uint32_t zw[2][2], aw[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++) { zw[i][0] = zn[i].x; zw[i][1] = zn[i].y; }
#pragma unroll
    for (int j = 0; j < 2; j++) { aw[j][0] = an[j].x & mn.x; aw[j][1] = an[j].y & mn.y; }
    if (c + 1 < kPopWords / 4) {
      const int cc = 2 * (c0 + c + 1);
#pragma unroll
      for (int i = 0; i < 2; i++) zn[i] = Zr[i][cc];
#pragma unroll
      for (int j = 0; j < 2; j++) an[j] = Ar[j][cc];
      mn = mk[cc];
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
      v4i af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; i++) af[i] = spread16((zw[i][p >> 1] >> (16 * (p & 1))) & 0xFFFFu);
#pragma unroll
      for (int j = 0; j < 2; j++) bf[j] = spread16((aw[j][p >> 1] >> (16 * (p & 1))) & 0xFFFFu);
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  }
  const int* cq_a = g.ca + (long long)q * g.ldca;
  const int* cq_z = g.cz + (long long)q * g.ldcz;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int ac = ab + 32 * j + r;
    const int caj = cq_a[min(ac, g.na - 1)];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int zc = zb + 32 * i + 8 * (v / 4) + 4 * hg + (v % 4);
        const int x = cq_z[min(zc, g.nz - 1)] + caj - 2 * acc[i][j][v];
        if (ac < g.na && zc < g.nz) X[((long long)q * g.nz + zc) * g.na + ac] = (uint32_t)x;
      }
  }
}

__global__ void k_gram_combine(GramMArgs g, const uint32_t* X)
{
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.na * g.nz) return;
  const int zc = (int)(e / g.na), ac = (int)(e % g.na);
  const double* D = g.ajdk + (size_t)5 * g.pcols;
  const long long plane = (long long)g.na * g.nz;
  double dsum = 0.0;
  for (int q = 0; q < g.nseg; q++) dsum += (double)D[q] * (double)X[q * plane + e];
  const int aor = g.alist[ac], zor = g.zlist[zc];
  double o = g.base16[5];
  o -= (2 * dsum);
  o -= g.xa[aor] + g.xz[zor] - g.base16[7];
  const long long aidx = g.a_orig ? aor : ac, zidx = g.z_orig ? zor : zc;
  double* dst = g.out + zidx * g.ldo + aidx;
  if (g.op_sub) *dst -= o; else *dst = o;
}

size_t gram_split_bytes_any(const GramMArgs& g)
{
  if (g.seg || g.na <= 0 || g.nz <= 0 || g.nseg < 2) return 0;
  const double b = 4.0 * g.nseg * (double)g.na * g.nz;
  return b <= 4.0e9 ? (size_t)b + 256 : 0;
}

size_t gram_split_bytes(const GramMArgs& g)
{
  // the split pays an int32 store and reload per segment and pair: only for Grams too
  // small to fill the CUs with whole tiles
  const long long tiles = (long long)cdiv(g.na, 128) * cdiv(g.nz, 128);
  if (tiles >= 128) return 0;
  return gram_split_bytes_any(g);
}

void gram_split(const GramMArgs& g, uint32_t* X, hipStream_t st)
{
  const unsigned tiles = cdiv(g.na, 128) * cdiv(g.nz, 128);
  hipLaunchKernelGGL(k_gram_split, dim3(tiles, g.nseg), dim3(256), 0, st, g, X);
  const long long n = (long long)g.na * g.nz;
  hipLaunchKernelGGL(k_gram_combine, dim3(cdiv(n, 256)), dim3(256), 0, st, g, (const uint32_t*)X);
}

void gram_mfma(const GramMArgs& g, hipStream_t st)
{
  if (g.na <= 0 || g.nz <= 0) return;
  if (g.seg) {
    // compacted layout: word steps, 64 x 128 tiles (more waves resident for the
    // per-segment double work that bounds these Grams)
    if (g.nseg <= 0) return;
    const long long nb = (long long)cdiv(g.na, 128) * cdiv(g.nz, 64);
    const long long grid = ((nb + 7) / 8) * 8;
    if (g.op_sub) hipLaunchKernelGGL((k_gram_seg<1>), dim3((unsigned)grid), dim3(256), 0, st, g);
    else hipLaunchKernelGGL((k_gram_seg<0>), dim3((unsigned)grid), dim3(256), 0, st, g);
    return;
  }
  const long long nb = (long long)cdiv(g.na, 128) * cdiv(g.nz, 128);
  const long long grid = ((nb + 7) / 8) * 8;
  if (g.op_sub) hipLaunchKernelGGL((k_gram_mfma<1, 2>), dim3((unsigned)grid), dim3(256), 0, st, g);
  else hipLaunchKernelGGL((k_gram_mfma<0, 2>), dim3((unsigned)grid), dim3(256), 0, st, g);
}

}  // namespace lkc
