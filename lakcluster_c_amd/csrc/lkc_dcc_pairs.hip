// lkc_dcc_pairs.hip -- the dcc half-loop terms of one bin pair and the score stages
// that sum them over the bins (several bins, Y/W genetic controls).
//
// Arrays indexed by (row or column, covariate) are covariate-major, X[t*ld + r].
// Every expression keeps the operand order of the reference line it cites
// (compiled with -ffp-contract=off), so results are bit-identical to it.
#include "lkc_dcc_bins.h"

#include <algorithm>

namespace lkc {

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
static_assert(sizeof(PairQCJobs) <= 4096 && sizeof(PairQRJobs) <= 4096 && sizeof(NrmJobs) <= 4096 &&
              sizeof(FillJobs) <= 4096, "job groups must fit the kernel arguments");
__device__ __forceinline__ int pm_at(const uint8_t* row, int i) { return ((row[i >> 3] >> (7 - (i & 7))) & 1) ? 1 : -1; }

// get_D_AtTn_ZtSn_vv / get_D_AtTn_AtTn_vv (D_AtTn_ZtSn.c:5-310) of one bin pair, one thread
// per (column k, covariate t); the half-loop (dcc_lf.c:461-486) stores it at the covariate's
// loaded index.
__device__ __forceinline__ void pair_qc_at(const PairQC& a, long long i)
{
  const int k = a.kl[i % a.nk], tt = a.tl[i / a.nk];
  const long long o = (long long)tt * a.ld + k;
  const double al = a.ajdk[(size_t)1 * a.pcols + k / kPopBits], D = a.ajdk[(size_t)5 * a.pcols + k / kPopBits];
  double v = (a.L1[o] - a.etT[tt] * al) * D * (a.L2[o] - al * a.etS[tt]);
  if (a.etA) {
    const double r = (double)a.rpop;
    v -= D * (r - 2 * a.etA[k] * al + r * al * al);
  }
  a.out[(long long)a.tb[tt] * a.ld + k] = v;
}
__global__ void k_pair_qc(PairQC a)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < (long long)a.nk * a.nt) pair_qc_at(a, i);
}
void pair_qc(const PairQC& a, hipStream_t st)
{
  const long long n = (long long)a.nk * a.nt;
  if (n > 0) hipLaunchKernelGGL(k_pair_qc, dim3(cdiv(n, 256)), dim3(256), 0, st, a);
}
// several pairs' QC terms in one launch: blockIdx.y picks the pair
__global__ void k_pair_qc_jobs(PairQCJobs J)
{
  const PairQC& a = J.job[blockIdx.y];
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < (long long)a.nk * a.nt) pair_qc_at(a, i);
}

// get_dcc_TAnZtS_ww_stage_2 (TAnZtS.c:380-495) and the half-loop copy (dcc_lf.c:425-458) of one
// bin pair, one thread per (row r of the first bin, covariate t):
//   0 + (X_a0d1(r,t) - a1d1(t) - x6(r) et(t) + a2d1 et(t))  [ - T(r,t) (a0d1 - 2 x6(r) + a2d1) ]
// the bracket for the same bin's A-with-A (and Y-with-Y) term only.
__device__ __forceinline__ void pair_qr_at(const PairQR& a, long long i)
{
  const int r = a.ar[i % a.na], tt = a.tl[i / a.na];
  const long long o = (long long)tt * a.ld + r;
  const double x6 = a.x6[r], a0d1 = a.mlf[5], a2d1 = a.mlf[7];
  double v = 0.0;
  v += +(a.x[o] - a.a1[tt] - x6 * a.et[tt] + a2d1 * a.et[tt]);
  if (a.Tt) {
    const double l = pm_at(a.Tt + (size_t)tt * a.rbT, r) * (a0d1 - 2 * x6 + a2d1);
    v += -l;
  }
  a.out[(long long)a.tb[tt] * a.ld + r] = v;
}
__global__ void k_pair_qr(PairQR a)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < (long long)a.na * a.nt) pair_qr_at(a, i);
}
void pair_qr(const PairQR& a, hipStream_t st)
{
  const long long n = (long long)a.na * a.nt;
  if (n > 0) hipLaunchKernelGGL(k_pair_qr, dim3(cdiv(n, 256)), dim3(256), 0, st, a);
}
__global__ void k_pair_qr_jobs(PairQRJobs J)
{
  const PairQR& a = J.job[blockIdx.y];
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < (long long)a.na * a.nt) pair_qr_at(a, i);
}

// Job lists launched in groups that fit one kernel's arguments; each group's grid is as
// wide as its largest job.
template <class Job, class Jobs, class Size>
static void launch_jobs(void (*kern)(Jobs), const std::vector<Job>& v, Size size, hipStream_t st)
{
  for (size_t j0 = 0; j0 < v.size(); j0 += Jobs::kMax) {
    Jobs J;
    J.n = (int)std::min<size_t>(Jobs::kMax, v.size() - j0);
    long long widest = 0;
    for (int j = 0; j < J.n; j++) { J.job[j] = v[j0 + j]; widest = std::max<long long>(widest, size(J.job[j])); }
    if (widest > 0) hipLaunchKernelGGL(kern, dim3(cdiv(widest, 256), J.n), dim3(256), 0, st, J);
  }
}
void pair_qc_jobs(const std::vector<PairQC>& v, hipStream_t st)
{
  launch_jobs(k_pair_qc_jobs, v, [](const PairQC& a) { return (long long)a.nk * a.nt; }, st);
}
void pair_qr_jobs(const std::vector<PairQR>& v, hipStream_t st)
{
  launch_jobs(k_pair_qr_jobs, v, [](const PairQR& a) { return (long long)a.na * a.nt; }, st);
}

// dcc_sumscores_nrm's divisions (dra_plusdivequals_s___m_m) of several arrays in one launch:
// dst[t*ld + r] = 0 + src[t*ld + r] / den over the listed rows and covariates, as k_nrm
__global__ void k_nrm_jobs(NrmJobs J)
{
  const NrmJob& a = J.job[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.s.n) return;
  const int r = a.s.list[i];
  for (int t = 0; t < a.s.nt; t++) {
    const long long o = (long long)a.s.tl[t] * a.s.ld + r;
    a.dst[o] = 0.0 + a.src[o] / a.den;
  }
}
void nrm_jobs(const std::vector<NrmJob>& v, hipStream_t st)
{
  launch_jobs(k_nrm_jobs, v, [](const NrmJob& a) { return (long long)a.s.n; }, st);
}

// zero fills of several arrays in one launch
__global__ void k_fill_jobs(FillJobs J)
{
  const FillJob& a = J.job[blockIdx.y];
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < a.n; i += (long long)gridDim.x * 256) a.p[i] = 0.0;
}
void fill_jobs(const std::vector<FillJob>& v, hipStream_t st)
{
  for (size_t j0 = 0; j0 < v.size(); j0 += FillJobs::kMax) {
    FillJobs J;
    J.n = (int)std::min<size_t>(FillJobs::kMax, v.size() - j0);
    long long widest = 0;
    for (int j = 0; j < J.n; j++) { J.job[j] = v[j0 + j]; widest = std::max<long long>(widest, J.job[j].n); }
    const long long nb = std::min<long long>(cdiv(widest, 256), 1024);
    if (widest > 0) hipLaunchKernelGGL(k_fill_jobs, dim3((unsigned)nb, J.n), dim3(256), 0, st, J);
  }
}

// dra_sumx (ra_stats.c:106-109) as dcc_sumscores_nrm applies it (dcc_sumscores.c:604-695):
// x[t*ld + r] += x[t*ld + r + q*stride] for q = 1 .. nsum-1, in that order
__global__ void k_bins_sumx(const int* list, int n, const int* tl, int nt, long long ld, double* x, long long stride, int nsum)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)n * nt) return;
  double* p = x + (long long)tl[i / n] * ld + list[i % n];
  double v = p[0];
  for (int q = 1; q < nsum; q++) v += p[(long long)q * stride];
  p[0] = v;
}
void bins_sumx(const int* list, int n, const int* tl, int nt, long long ld, double* x, long long stride, int nsum, hipStream_t st)
{
  const long long m = (long long)n * nt;
  if (m > 0 && nsum > 1) hipLaunchKernelGGL(k_bins_sumx, dim3(cdiv(m, 256)), dim3(256), 0, st, list, n, tl, nt, ld, x, stride, nsum);
}

// dcc_sumscores_cmb's row scores (dcc_sumscores.c:265-272), first covariate column of every
// bin block: x += +1 yw, then x += -1 az, then x += -1 yy (absent terms skipped)
__global__ void k_bins_cmb_rows(const int* list, int n, int nblk, long long blk, double* x, const double* yw,
                                const double* az, const double* yy)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long long)n * nblk) return;
  const long long o = (i / n) * blk + list[i % n];
  double v = x[o];
  if (yw) v += +1.0 * yw[o];
  if (az) v += -1.0 * az[o];
  if (yy) v += -1.0 * yy[o];
  x[o] = v;
}
void bins_cmb_rows(const int* list, int n, int nblk, long long blk, double* x, const double* yw, const double* az,
                   const double* yy, hipStream_t st)
{
  const long long m = (long long)n * nblk;
  if (m > 0) hipLaunchKernelGGL(k_bins_cmb_rows, dim3(cdiv(m, 256)), dim3(256), 0, st, list, n, nblk, blk, x, yw, az, yy);
}

}  // namespace lkc
