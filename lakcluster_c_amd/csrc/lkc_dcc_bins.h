// lkc_dcc_bins.h -- launchers of the multi-bin dcc stage kernels (lkc_dcc_pairs.hip) and the
// host side of the multi-bin dcc eliminations on the device (lkc_dcc_bins.cpp):
// dcc_sumscores_test on its synthetic problem and dexcluster_driver on files.
#pragma once
#include "lkc_kernels.h"
#include "lkc_stages.h"
#include <string>
#include <vector>

namespace lkc {

// one QC half-loop term of a bin pair: out[tb[t]*ld + k] for the listed columns k and
// active covariates t; L1/L2 [t*ld + k]; etA (with rpop) adds the same-bin correction
struct PairQC {
  const int* kl; int nk; const int* tl; int nt; long long ld;
  const double* L1; const double* L2; const double* etT; const double* etS;
  const double* etA; int rpop;
  const double* ajdk; int pcols;
  int tb[kMaxT];
  double* out;
};
void pair_qc(const PairQC& a, hipStream_t st);

// one QR half-loop term of a bin pair: out[tb[t]*ld + r] over the first bin's listed rows;
// x the An_ZtSn_ww product [t*ld + r], a1/et of the second bin [t], x6 = lf_*_ajdk[AJDK_1_1],
// mlf the weighted counts of the active columns ([5] a0d1, [7] a2d1); Tt (rows of rbT
// bytes) adds the same-bin diagonal term
struct PairQR {
  const int* ar; int na; const int* tl; int nt; long long ld;
  const double* x; const double* a1; const double* et; const double* x6; const double* mlf;
  const uint8_t* Tt; size_t rbT;
  int tb[kMaxT];
  double* out;
};
void pair_qr(const PairQR& a, hipStream_t st);

// the same terms for many bin pairs, a few launches each (kernel-argument sized groups)
struct PairQCJobs { static constexpr int kMax = 24; int n; PairQC job[kMax]; };
struct PairQRJobs { static constexpr int kMax = 24; int n; PairQR job[kMax]; };
void pair_qc_jobs(const std::vector<PairQC>& v, hipStream_t st);
void pair_qr_jobs(const std::vector<PairQR>& v, hipStream_t st);
// dcc_sumscores_nrm's divisions, and zero fills, of many arrays per launch
struct NrmJob { ScoreStage s; const double* src; double den; double* dst; };
struct NrmJobs { static constexpr int kMax = 48; int n; NrmJob job[kMax]; };
void nrm_jobs(const std::vector<NrmJob>& v, hipStream_t st);
struct FillJob { double* p; long long n; };
struct FillJobs { static constexpr int kMax = 192; int n; FillJob job[kMax]; };
void fill_jobs(const std::vector<FillJob>& v, hipStream_t st);

void bins_sumx(const int* list, int n, const int* tl, int nt, long long ld, double* x, long long stride, int nsum,
               hipStream_t st);
void bins_cmb_rows(const int* list, int n, int nblk, long long blk, double* x, const double* yw, const double* az,
                   const double* yy, hipStream_t st);

// dcc_sumscores_test (dcc_sumscores.c:700-747) on dcc_init's synthetic problem; outputs
struct DccTestOut { std::vector<double> trace; std::vector<int> xdrop_a, xdrop_b; int n_iter = 0; long long ops = 0; };
struct Cfg;
void dcc_sumscores_test(const Cfg& c, DccTestOut& out, const std::string& out_dir);
// dexcluster_driver (dexcluster_driver.c:5-57) on the .in's per-bin files with Y/W controls,
// run to the end; the three output files into out_dir when it is not empty
void dcc_driver_bins(const Cfg& c, DccTestOut& out, const std::string& out_dir);

}  // namespace lkc
