// lkc_stages.h -- launchers for the element-wise / ordered-sum stage kernels.
#pragma once
#include "lkc_kernels.h"

namespace lkc {

struct OrderedJob { const int* list; int n; const uint8_t* sign; const double* vals; double* out; };
struct OrderedJobs { OrderedJob job[48]; int n; };   // up to 4 sums per covariate (8) + 1
void ordered_sums(const OrderedJobs& J, hipStream_t st);

void et_from_counts(const int* cnt, const int* list, int n, long long base, double* out, int assign_only, hipStream_t st);
void t_anat(const int* ar, int nar, const int* tl, int nt, const Bits& Tt, const double* x6, const double* mlf,
            double* out, long long ld, hipStream_t st);
void scale_rows(const double* src, long long ld, const int* kl, int nk, const int* tl, int nt, const double* w,
                double* L, hipStream_t st);

// sc layout: [0,T) a1d2_AtTn, [T,2T) a3d2_AtTn, [2T,3T) a1d2_ZtSn, [3T,4T) a3d2_ZtSn,
//            [4T,5T) et_Tn, [5T,6T) et_Sn, [6T] et_An_a1d1
struct FlatQR {
  const int* ar; int nar; const int* tl; int nt; int T; int NA; int has_Z;
  const double* ajdk_rows;      // lf_An_ajdk [jk*NA + a]
  const double* mlf;            // mask_lf of the active columns (16 values)
  const double* sc;
  const double* lf_AnZtSZnAt; const double* lf_AnAtTAnAt;
  const double* An_a2d2_ZtSn; const double* An_a2d2_AtTn; const double* T_AnAt_AnAt;
  const uint8_t* Tt; size_t rbT;
  double* QR_ZSZ; double* QR_TAT;
};
void flat_qr(const FlatQR& f, hipStream_t st);

struct FlatQC {
  const int* kl; int nk; const int* tl; int nt; int T; int M; int has_Z;
  const double* ajdk; int pcols; const double* mlf; const double* sc;
  const double* lf_AtTAnZtSZn; const double* lf_AtTAnAtTAn; const double* lf_AtTn; const double* lf_ZtSn;
  const double* et_An; const double* At_An_a1d1; long long rpop;
  double* QC_ZSZ; double* QC_TAT;
};
void flat_qc(const FlatQC& f, hipStream_t st);

// the score stages over the listed rows (or columns) and active covariates; arrays [t*ld + r]
struct ScoreStage { const int* list; int n; const int* tl; int nt; long long ld; };
void score_nrm(const ScoreStage& s, const double* src, double den, double* dst, hipStream_t st);
void score_ifT(const ScoreStage& s, double* v, double den, int pow_form, hipStream_t st);
void score_cmb(const ScoreStage& s, double* x, const double* z, hipStream_t st);
void score_gather0(const ScoreStage& s, const double* x, double* out, hipStream_t st);

struct StageC {
  int M; int T; int has_Z; const uint32_t* Ac_rtn; const uint32_t* Ac_b; const uint32_t* Tc_j;
  const double* vtszn; const double* t_knzt; const double* t_hnvt;
  const double* jttan; const double* htrkn; const double* hhtrkn; const double* t_ankt; const double* t_knkt;
  double* ZSZ; double* TAT;
};
void stage_c(const StageC& c, hipStream_t st);

void fill(double* p, long long n, double v, hipStream_t st);

// dcc half-loop subscores of one bin (get_D_AtTn_ZtSn_vv / get_D_AtTn_AtTn_vv,
// D_AtTn_ZtSn.c:5-310; get_dcc_TAnZtS_ww_stage_2 and get_dcc_halfloop, TAnZtS.c:380-495,
// dcc_lf.c:379-488).  (row, covariate) arrays are [t*rows + row].
struct DccQC {
  const int* kl; int nk; const int* tl; int nt; long long ld;
  const double* AtTn; const double* ZtSn; const double* etT; const double* etS; const double* etA; int rpop;
  const double* ajdk; int pcols; int has_Z; double* QC_ZSZ; double* QC_TAT;
};
void dcc_qc(const DccQC& a, hipStream_t st);
struct DccQR {
  const int* ar; int na; const int* tl; int nt; long long ld;
  const double* xZ; const double* xA; const double* a1T; const double* a1S; const double* etT; const double* etS;
  const double* x6; const double* mlf; const uint8_t* Tt; size_t rbT; int has_Z; double* QR_ZSZ; double* QR_TAT;
};
void dcc_qr(const DccQR& a, hipStream_t st);

}  // namespace lkc
