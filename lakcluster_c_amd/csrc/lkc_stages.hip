// lkc_stages.hip -- element-wise and ordered-sum kernels of the bcc stages that sit
// between the bit-matrix contractions: single/double-study terms, flattenloop,
// score normalisation, the QC stage-c update.
//
// Arrays indexed by (row, covariate) are covariate-major, X[t*ld + row].
// Every expression keeps the operand order of the reference line it cites
// (compiled with -ffp-contract=off), so results are bit-identical to it.
#include "lkc_stages.h"

namespace lkc {

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
__device__ __forceinline__ int bit_of(const uint8_t* row, int i) { return (row[i >> 3] >> (7 - (i & 7))) & 1; }
__device__ __forceinline__ int pm_of(const uint8_t* row, int i) { return bit_of(row, i) ? 1 : -1; }
__device__ __forceinline__ int bit_w(const uint32_t* m, int i) { return bit_of((const uint8_t*)m, i); }

// ------------------------------------------------------------------
// ordered sums Σ_i s_i·v[list[i]] (get_singlestudy_vv, bcc_flattenloop.c:359-731):
// one block per job; chunks of 2048 values are staged in LDS (the next chunk's
// gathers in flight meanwhile) and lane 0 adds them in order.
// ------------------------------------------------------------------
#define OS_CH 2048
__global__ __launch_bounds__(256) void k_ordered_sums(OrderedJobs J)
{
  __shared__ double buf[OS_CH];
  int q = blockIdx.x;
  if (q >= J.n) return;
  const OrderedJob& jb = J.job[q];
  constexpr int PER = OS_CH / 256;
  double reg[PER];
  auto load = [&](int i0) {
#pragma unroll
    for (int u = 0; u < PER; u++) {
      // clamped index, no branch: all the gathers of a thread are in flight together
      const int i = i0 + threadIdx.x + 256 * u, ic = min(i, jb.n - 1);
      const int r = jb.list[ic];
      double v = jb.vals[r];
      if (jb.sign) v = pm_of(jb.sign, r) * v;
      reg[u] = i < jb.n ? v : 0.0;
    }
  };
  double o = 0.0;
  if (jb.n > 0) load(0);
  for (int i0 = 0; i0 < jb.n; i0 += OS_CH) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; u++) buf[threadIdx.x + 256 * u] = reg[u];
    __syncthreads();
    if (i0 + OS_CH < jb.n) load(i0 + OS_CH);
    if (threadIdx.x == 0) {
      const int m = jb.n - i0 < OS_CH ? jb.n - i0 : OS_CH;
#pragma unroll 16
      for (int e = 0; e < m; e++) o += buf[e];
    }
  }
  if (threadIdx.x == 0) *jb.out = o;
}
void ordered_sums(const OrderedJobs& J, hipStream_t st)
{
  if (J.n <= 0) return;
  hipLaunchKernelGGL(k_ordered_sums, dim3(J.n), dim3(256), 0, st, J);
}

// out[list[i]] = 0 + (double)(-(long long)base + 2*cnt[i])  (bcc_flattenloop.c:690-701)
__global__ void k_et(const int* cnt, const int* list, int n, long long base, double* out, int assign_only)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  long long v = -base + 2 * (long long)cnt[i];
  int r = list ? list[i] : i;
  out[r] = assign_only ? (double)v : 0.0 + (double)v;
}
void et_from_counts(const int* cnt, const int* list, int n, long long base, double* out, int assign_only, hipStream_t st)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_et, dim3(cdiv(n, 256)), dim3(256), 0, st, cnt, list, n, base, out, assign_only);
}

// T_AnAt_AnAt (bcc_flattenloop.c:702-731): the A row's centred Gram with itself, squared, signed by T
__global__ void k_t_anat(const int* ar, int nar, const int* tl, int nt, const uint8_t* Tt, size_t rbT,
                         const double* x6, const double* mlf, double* out, long long ld)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nar) return;
  int a = ar[i];
  double x11 = x6[a];
  double Ad = mlf[5];
  Ad -= 2 * 0.0;
  Ad -= x11 + x11 - mlf[7];
  for (int t = 0; t < nt; t++) {
    int tt = tl[t];
    int vT = pm_of(Tt + (size_t)tt * rbT, a);
    out[(long long)tt * ld + a] = vT * Ad * Ad;
  }
}
void t_anat(const int* ar, int nar, const int* tl, int nt, const Bits& Tt, const double* x6, const double* mlf,
            double* out, long long ld, hipStream_t st)
{
  if (nar <= 0) return;
  hipLaunchKernelGGL(k_t_anat, dim3(cdiv(nar, 256)), dim3(256), 0, st, ar, nar, tl, nt, Tt.p, Tt.rb, x6, mlf, out, ld);
}

// L[t*nk + k] = w[kl[k]/1920] * src[tl[t]*ld + kl[k]]  (An_a2d2_ZtSn_excerpt, An_ZtSn_ww.c:267-290)
__global__ void k_scale_rows(const double* src, long long ld, const int* kl, int nk, const int* tl, int nt,
                             const double* w, double* L)
{
  long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)nk * nt) return;
  int t = (int)(e / nk), k = (int)(e % nk);
  int kk = kl[k];
  L[e] = w[kk / kPopBits] * src[(long long)tl[t] * ld + kk];
}
void scale_rows(const double* src, long long ld, const int* kl, int nk, const int* tl, int nt, const double* w,
                double* L, hipStream_t st)
{
  long long n = (long long)nk * nt;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_rows, dim3(cdiv(n, 256)), dim3(256), 0, st, src, ld, kl, nk, tl, nt, w, L);
}

// ------------------------------------------------------------------
// get_flattenloop, QR half (bcc_flattenloop.c:1599-1716)
// ------------------------------------------------------------------
__global__ void k_flat_qr(FlatQR f)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= f.nar) return;
  int a = f.ar[i];
  const double* x11p = f.ajdk_rows + (size_t)11 * f.NA;
  const double* x13p = f.ajdk_rows + (size_t)13 * f.NA;
  double x11 = x11p[a], x13 = x13p[a];
  double a0d2 = f.mlf[10], a2d2 = f.mlf[12], a4d2 = f.mlf[14];
  for (int t = 0; t < f.nt; t++) {
    int tt = f.tl[t];
    long long o = (long long)tt * f.NA + a;
    double eT = f.sc[4 * f.T + tt];
    double dtmp;
    if (f.has_Z) {
      double eS = f.sc[5 * f.T + tt];
      double q = f.lf_AnZtSZnAt[o];
      dtmp = + 1*a0d2*eS - 2*f.sc[2 * f.T + tt] + 1*a2d2*eS - 2*x11*eS + 4*f.An_a2d2_ZtSn[o]
             - 2*x13*eS + 1*a2d2*eS - 2*f.sc[3 * f.T + tt] + 1*a4d2*eS;
      q += -dtmp;
      f.QR_ZSZ[o] = q;
    }
    double r = f.lf_AnAtTAnAt[o];
    dtmp = + 1*a0d2*eT - 2*f.sc[0 * f.T + tt] + 1*a2d2*eT - 2*x11*eT + 4*f.An_a2d2_AtTn[o]
           - 2*x13*eT + 1*a2d2*eT - 2*f.sc[1 * f.T + tt] + 1*a4d2*eT;
    r += -dtmp;
    dtmp = f.T_AnAt_AnAt[o];
    r += -dtmp;
    int vT = pm_of(f.Tt + (size_t)tt * f.rbT, a);
    dtmp = vT*( + 1*a0d2 - 4*(x11) + 6*a2d2 - 4*(x13) + 1*a4d2 );
    r += +dtmp;
    f.QR_TAT[o] = r;
  }
}
void flat_qr(const FlatQR& f, hipStream_t st)
{
  if (f.nar <= 0) return;
  hipLaunchKernelGGL(k_flat_qr, dim3(cdiv(f.nar, 256)), dim3(256), 0, st, f);
}

// get_flattenloop, QC half (bcc_flattenloop.c:1720-1889)
__global__ void k_flat_qc(FlatQC f)
{
  int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= f.nk) return;
  int kk = f.kl[j];
  int b = kk / kPopBits;
  const double* J = f.ajdk; int P = f.pcols;
  double D01 = J[b + 5 * P], D11 = J[b + 6 * P], D21 = J[b + 7 * P], a10 = J[b + 1 * P], a20 = J[b + 2 * P];
  double eA = f.et_An[kk], eAa = f.sc[6 * f.T], AtA = f.At_An_a1d1[kk];
  double a0d1 = f.mlf[5], a2d1 = f.mlf[7];
  long long rpop = f.rpop;
  for (int t = 0; t < f.nt; t++) {
    int tt = f.tl[t];
    long long o = (long long)tt * f.M + kk;
    double eT = f.sc[4 * f.T + tt];
    double dtmp;
    double D = J[kk / kPopBits + 5 * P];
    if (f.has_Z) {
      double q = D * f.lf_AtTAnZtSZn[o];
      double eS = f.sc[5 * f.T + tt];
      dtmp = D01 * ( +(1*eT) - (2*a10 * f.lf_AtTn[o]) + (a20 * eT) ) * D01 * ( +(1*eS) - (2*a10 * f.lf_ZtSn[o]) + (a20 * eS) );
      q += -dtmp;
      f.QC_ZSZ[o] = q;
    }
    double r = D * f.lf_AtTAnAtTAn[o];
    dtmp = D01 * ( +(1*eT) - (2*a10 * f.lf_AtTn[o]) + (a20 * eT) ) * D01 * ( +(1*eT) - (2*a10 * f.lf_AtTn[o]) + (a20 * eT) );
    r += -dtmp;
    dtmp = ( +(1*D01*(rpop)*a0d1) - (2*D11*eA*a0d1) + (1*D21*(rpop)*a0d1)
             - (2*D01*eAa) + (4*D11*AtA) - (2*D21*eAa)
             + (1*D01*(rpop)*a2d1) - (2*D11*eA*a2d1) + (1*D21*(rpop)*a2d1) );
    r += -dtmp;
    dtmp = ( +(1*J[b + 10 * P]*(rpop)) - (4*J[b + 11 * P]*eA) + (6*J[b + 12 * P]*(rpop))
             - (4*J[b + 13 * P]*eA) + (1*J[b + 14 * P]*(rpop)) );
    r += +dtmp;
    f.QC_TAT[o] = r;
  }
}
void flat_qc(const FlatQC& f, hipStream_t st)
{
  if (f.nk <= 0) return;
  hipLaunchKernelGGL(k_flat_qc, dim3(cdiv(f.nk, 256)), dim3(256), 0, st, f);
}

// ------------------------------------------------------------------
// bcc_sumscores_nrm / _ifT / _cmb (bcc_sumscores.c:324-779; dcc_sumscores.c:248-699), one
// bin, GLOBAL_Ireq <= 0, over the listed rows (or columns) r and the active covariates:
//   nrm: dst[t*ld + r] = 0.0 + src[t*ld + r] / den (dra_plusdivequals into a zeroed array)
//   ifT: column 0 corrected by the others (dra_mds_pow / dra_mds_nrm, ra_stats.c:77-104)
//   cmb: x[0*ld + r] += -1.0 * z[0*ld + r]
// then xij's gather of column 0 into list order.
// ------------------------------------------------------------------
__global__ void k_nrm(ScoreStage s, const double* src, double den, double* dst)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= s.n) return;
  const int r = s.list[i];
  for (int t = 0; t < s.nt; t++) {
    const long long o = (long long)s.tl[t] * s.ld + r;
    dst[o] = 0.0 + src[o] / den;
  }
}
__global__ void k_ifT(ScoreStage s, double* v_, double den, int pow_form)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= s.n) return;
  const int r = s.list[i];
  double* v = v_ + r;                      // v[t * ld]
  const long long ld = s.ld;
  if (pow_form) {
    if (v[0] >= 0) {
      double v0 = v[0] * v[0];
      for (int c = 0; c < s.nt; c++) if (s.tl[c] >= 1) v0 -= v[s.tl[c] * ld] * v[s.tl[c] * ld] / den;
      v[0] = v0 >= 0 ? sqrt(v0) : -sqrt(-v0);
    }
  } else {
    double v0 = v[0];
    for (int c = 0; c < s.nt; c++) if (s.tl[c] >= 1) v0 -= v[s.tl[c] * ld] / den;
    v[0] = v0;
  }
}
__global__ void k_cmb(ScoreStage s, double* x, const double* z)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= s.n) return;
  const int r = s.list[i];
  x[r] += -1.0 * z[r];
}
__global__ void k_gather0(ScoreStage s, const double* x, double* out)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= s.n) return;
  out[i] = x[s.list[i]];
}
void score_nrm(const ScoreStage& s, const double* src, double den, double* dst, hipStream_t st)
{
  if (s.n > 0) hipLaunchKernelGGL(k_nrm, dim3(cdiv(s.n, 256)), dim3(256), 0, st, s, src, den, dst);
}
void score_ifT(const ScoreStage& s, double* v, double den, int pow_form, hipStream_t st)
{
  if (s.n > 0) hipLaunchKernelGGL(k_ifT, dim3(cdiv(s.n, 256)), dim3(256), 0, st, s, v, den, pow_form);
}
void score_cmb(const ScoreStage& s, double* x, const double* z, hipStream_t st)
{
  if (s.n > 0) hipLaunchKernelGGL(k_cmb, dim3(cdiv(s.n, 256)), dim3(256), 0, st, s, x, z);
}
void score_gather0(const ScoreStage& s, const double* x, double* out, hipStream_t st)
{
  if (s.n > 0) hipLaunchKernelGGL(k_gather0, dim3(cdiv(s.n, 256)), dim3(256), 0, st, s, x, out);
}

// ------------------------------------------------------------------
// bcc_lrup_QC_YnWt_stage_c (bcc_lrup.c:592-655): L2_clean, then subtract the low-rank terms
// ------------------------------------------------------------------
__global__ void k_stage_c(StageC c)
{
  long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)c.M * c.T) return;
  int t = (int)(e / c.M), k = (int)(e % c.M);
  int keep = bit_w(c.Ac_rtn, k) && bit_w(c.Ac_b, k) && bit_w(c.Tc_j, t);
  if (!keep) { if (c.has_Z) c.ZSZ[e] = 0; c.TAT[e] = 0; return; }
  if (c.has_Z) {
    double q = c.ZSZ[e];
    q += -1.0 * c.vtszn[e];
    q += -1.0 * c.t_knzt[e];
    q += -1.0 * c.t_hnvt[e];
    c.ZSZ[e] = q;
  }
  double r = c.TAT[e];
  r += -1.0 * c.jttan[e];
  r += -1.0 * c.htrkn[e];
  r += -1.0 * c.htrkn[e];
  r += -1.0 * c.hhtrkn[e];
  r += -1.0 * c.t_ankt[e];
  r += -1.0 * c.t_ankt[e];
  r += -1.0 * c.t_knkt[e];
  c.TAT[e] = r;
}
void stage_c(const StageC& c, hipStream_t st)
{
  long long n = (long long)c.M * c.T;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_stage_c, dim3(cdiv(n, 256)), dim3(256), 0, st, c);
}

// grid-stride: a dispatch counts its work-items in 32 bits, and the square score
// arrays reach 2^32 entries (65536 x 65536)
__global__ void k_fill(double* p, long long n, double v)
{
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}
void fill(double* p, long long n, double v, hipStream_t st)
{
  if (n <= 0) return;
  const unsigned nb = cdiv(n, 256);
  hipLaunchKernelGGL(k_fill, dim3(nb < 65536 ? nb : 65536), dim3(256), 0, st, p, n, v);
}

// ------------------------------------------------------------------
// dcc QC terms, one thread per (k, t): get_D_AtTn_ZtSn_vv and, with Z = A, the
// get_D_AtTn_AtTn_vv form that removes the a = a' terms (D_AtTn_ZtSn.c:5-310):
//   QC_TAnZtS = (AtTn - etTn a) D (ZtSn - a etSn)
//   QC_TAnAtT = (AtTn - etTn a) D (AtTn - a etTn) - D (|rows| - 2 etAn a + |rows| a a)
// with a, D the column's block values of A_ajdk (AJDK_1_0, AJDK_0_1).
// ------------------------------------------------------------------
__global__ void k_dcc_qc(DccQC a)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)a.nk * a.nt) return;
  const int k = a.kl[i % a.nk], tt = a.tl[i / a.nk];
  const long long o = (long long)tt * a.ld + k;
  const double al = a.ajdk[(size_t)1 * a.pcols + k / kPopBits], D = a.ajdk[(size_t)5 * a.pcols + k / kPopBits];
  const double At = a.AtTn[o], eT = a.etT[tt];
  if (a.has_Z) a.QC_ZSZ[o] = (At - eT * al) * D * (a.ZtSn[o] - al * a.etS[tt]);
  double v = (At - eT * al) * D * (At - al * eT);
  const double r = (double)a.rpop;
  v -= D * (r - 2 * a.etA[k] * al + r * al * al);
  a.QC_TAT[o] = v;
}
void dcc_qc(const DccQC& a, hipStream_t st)
{
  const long long n = (long long)a.nk * a.nt;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dcc_qc, dim3(cdiv(n, 256)), dim3(256), 0, st, a);
}

// dcc QR terms, one thread per (a, t) (stage 2 and the half-loop copy):
//   QR_TAnZtS = 0 + (An_a0d1_ZtSn - a1d1_ZtSn - An_a1d1 etSn + A_a2d1 etSn)
//   QR_TAnAtT = 0 + (An_a0d1_AtTn - a1d1_AtTn - An_a1d1 etTn + A_a2d1 etTn)
//               - T(a,t) (A_a0d1 - 2 An_a1d1 + A_a2d1)
// A_a0d1 = mlf[5], A_a2d1 = mlf[7] (the weighted counts of the active columns).
__global__ void k_dcc_qr(DccQR a)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)a.na * a.nt) return;
  const int r = a.ar[i % a.na], tt = a.tl[i / a.na];
  const long long o = (long long)tt * a.ld + r;
  const double x6 = a.x6[r], a0d1 = a.mlf[5], a2d1 = a.mlf[7];
  if (a.has_Z) {
    double z = 0.0;
    z += +(a.xZ[o] - a.a1S[tt] - x6 * a.etS[tt] + a2d1 * a.etS[tt]);
    a.QR_ZSZ[o] = z;
  }
  double v = 0.0;
  v += +(a.xA[o] - a.a1T[tt] - x6 * a.etT[tt] + a2d1 * a.etT[tt]);
  const double l = pm_of(a.Tt + (size_t)tt * a.rbT, r) * (a0d1 - 2 * x6 + a2d1);
  v += -l;
  a.QR_TAT[o] = v;
}
void dcc_qr(const DccQR& a, hipStream_t st)
{
  const long long n = (long long)a.na * a.nt;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dcc_qr, dim3(cdiv(n, 256)), dim3(256), 0, st, a);
}

}  // namespace lkc
