// lkc_engine.h -- host-side state of one bcc elimination on the device.
#pragma once
#include "lkc_kernels.h"
#include "../../include/lkc.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace lkc {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Cfg {
  int B = 32; long long D_MLT = 4294967296LL; double gamma = 0.125; int gamma_type = 0; double kappa2 = 0;
  int sparse = 1; int nbins = 0; int Ireq = 0; int A_ncols = 0, Y_ncols = 0, T_ncols = 0; int scramble_num = 0;
  std::string QR, QC, A_n, A_t, A_rind, A_cind, A_p, Z_n, Z_t, Z_rind, T_n, T_t, T_cind, S_n, S_t;
  std::string xpre, dir_name, dir_base, out_name;
  // dexcluster_driver with several bins: bin n's files (item n of each comma-separated list,
  // updateglobals.c:76-117, GLOBAL_DIR_XPRE applied; empty when not named) and the Y column
  // mask; the fields above hold bin 0's
  std::vector<std::string> A_n_, A_t_, A_rind_, Z_n_, Z_t_, Z_rind_, Y_n_, Y_t_, W_n_, W_t_, T_n_, T_t_, S_n_, S_t_;
  std::string Y_cind, Y_p;
  // bcc_scramble: per scramble an out_xdrop file and a seed (GLOBAL_scramble_out_xdrop_,
  // GLOBAL_scramble_rseed_; the seed of scramble n defaults to n + 1)
  std::vector<std::string> scramble_files; std::vector<unsigned long> scramble_seeds;
  std::string scramble_files_raw, scramble_seeds_raw;
  // GLOBAL_TEST_TYPE and the lakcluster_scorebox grid (updateglobals.c:18, 136-140)
  std::string test_type, sb_file; int sb_row_max = 0, sb_row_num = 0, sb_col_max = 0, sb_col_num = 0;
  // the self-tests' synthetic problem (GLOBAL_TEST_*, defaults of lakcluster_ver18.c:70-79):
  // read by dcc_sumscores_test, ignored by the file-driven eliminations
  std::string test_typ2; double test_mrand = 0.0;
  int test_A_rows = 12, test_A_cols = 13, test_Z_rows = 15, test_Y_cols = 11, test_T_cols = 3, test_niter = 2;
};
// lakcluster_scorebox output (struct S_handle): row_num x col_num entries at [nr + nc*row_num]
struct Scorebox {
  int row_max = 0, row_num = 0, col_max = 0, col_num = 0;
  std::vector<int> rpop, cpop, irem; std::vector<double> qr, qc;
};
void scorebox_xdrop_array(int nxols, int x_max, int x_num, int* xval, int* xdrop);
void write_scorebox(const Scorebox& S, const std::string& dir);
void parse_cfg(const char* text, Cfg& c);
// .b16 input (binary_read.c:5-55): the file's (fnr x fnc) matrix as fnc rows of fnr bits
struct HostBits { std::vector<uint8_t> b; int nr = 0, nc = 0; size_t rb = 0; };
HostBits read_b16(const std::string& path);
void b16_size(const std::string& path, int* fnr, int* fnc);   // the header's sizes only
std::vector<uint8_t> read_mask(const std::string& path, int n);
std::string out_dir_of(const Cfg& c);   // GLOBAL_DIR_NAME, else GLOBAL_DIR_BASE/dir_<GLOBAL_out_name>
void get_xdrop(double lrij, double lcij, double gamma, int gamma_type, int* rdrop, int* cdrop);
int get_xdrop_length(double lrij, double lcij, double gamma, int gamma_type);
unsigned sort_scores(double* d, int* perm, int n);
unsigned quicksort_xij(double* d, int* perm, int l, int r);

// Device allocations of one owner.  With a stream set, they are stream-ordered
// (hipMallocAsync / hipFreeAsync from the device's memory pool): growing or freeing a
// buffer then never synchronises the device, which matters when jobs run side by side.
struct Dev {
  std::vector<void*> ptrs;
  hipStream_t st = nullptr;
  ~Dev();
  void* raw(size_t bytes);
  void release(void* p);
  void release_all();
  template <class T> T* alloc(size_t n) { return (T*)raw(n * sizeof(T)); }
};
struct Grow { void* p = nullptr; size_t cap = 0; void* need(Dev& d, size_t bytes); };

struct DMat { uint8_t* p = nullptr; int nr = 0, nc = 0; size_t rb = 0; Bits b() const { return Bits{ p, nr, nc, rb }; } };
// Device-resident pool of patients that jobs draw their rows from (batched nulls,
// serving): Pn holds P rows of M bits, Tp the T covariate columns as rows of P bits.
struct Pool { Dev dev; int P = 0, M = 0, T = 0; DMat Pn, Tp; };
struct JobIn { const int* rowsA = nullptr; int nA = 0; const int* rowsZ = nullptr; int nZ = 0;
               const uint8_t* colmask = nullptr; Cfg cfg; double arrival_s = 0; };
struct JobOut { int status = 0; std::string err; int n_iter = 0; long long ops = 0;
                double arrival_s = 0, start_s = 0, end_s = 0;
                std::vector<double> trace; std::vector<int> xdrop_a, xdrop_b; };
double job_bytes(int nA, int nZ, int M, int B);
// Runs the jobs with at most max_conc at a time, admitted in order within budget bytes.
void run_batch(const Pool& P, const JobIn* jobs, int n, int max_conc, double budget, JobOut* out);
struct IList { std::vector<int> h; int* d = nullptr; int n() const { return (int)h.size(); } };
// one patient group of an At_T_YnWt_S_Zn term: SNP-major bits, active-row mask and list, covariates
struct Side { const DMat* X; const uint32_t* mask; const IList* rows; const DMat* Cov; };
// the same group with its patients packed to the front: SNP-major rows and covariate rows of
// rb bytes whose first nbits positions are the group's rows in list order (mask: those in use)
struct Packed { const uint8_t* X = nullptr; const uint8_t* Cov = nullptr; size_t rb = 0; const uint32_t* mask = nullptr; int nbits = 0; };
// one At_T_YnWt_S_Zn term: L(r1, r2) = src[idx1(r1)*st1 + idx2(r2)*st2] over the sides' rows, the
// n side (s2 when trm, else s1) contracted, the result into out[t*M + k] for k in kls
struct WwTerm { int trm; Side s1, s2; const IList* kls; const double* src; long long st1, st2; int full; double* out; Packed pk; };

// HIP-event timing of named kernel groups on the run's stream
struct Prof {
  struct Pending { int id; hipEvent_t a, b; };
  struct Stat { std::string name; long long launches = 0; double ms = 0, work = 0; };
  bool on = false;
  std::vector<Stat> stats;
  std::vector<Pending> pend;
  std::vector<hipEvent_t> pool;
  ~Prof();
  int id(const char* name);
  hipEvent_t ev();
  void begin(const char* name, double work, hipStream_t st, Pending& p);
  void end(Pending& p, hipStream_t st);
  void collect();                     // after a stream sync
};

struct Run {
  Cfg c;
  Prof prof;
  hipStream_t st = nullptr;
  Dev dev;
  int NA = 0, NZ = 0, M = 0, T = 0, pcols = 0, has_Z = 0, Arb = 0, Zrb = 0;
  DMat An, At, Zn, Zt, Tt, St;
  std::vector<uint8_t> hAr_b, hAr_j, hAr_rtn, hAr_rmv, hZr_b, hZr_j, hAc_b, hAc_j, hAc_rtn, hAc_rmv, hTc_j;
  uint32_t *Ar_b = nullptr, *Ar_j = nullptr, *Ar_rtn = nullptr, *Ar_rmv = nullptr, *Zr_b = nullptr, *Zr_j = nullptr,
           *Ac_b = nullptr, *Ac_j = nullptr, *Ac_rtn = nullptr, *Ac_rmv = nullptr, *Tc_j = nullptr;
  IList ar, kl, tl, zr, rtn, rmv, crt, crm;
  int* d_zero = nullptr;
  std::vector<int> Ar_bidx, Ac_bidx;
  std::vector<double> hA_p, hajdk;
  double* ajdk = nullptr;
  int trm_AnZt = 0, trm_knzt = 0, trm_hnvt = 0;
  // control block
  uint8_t* hctl = nullptr; uint8_t* dctl = nullptr; size_t ctl_bytes = 0, ctl_head = 0;
  size_t ctl_mask_off[11] = {}, ctl_list_off[8] = {}, ctl_cmp_off[9] = {};
  // column compaction of the dropped columns (crm), block by block
  int cmp_W = 0, cmp_nseg = 0, cmp_Wmax = 0;
  int *d_seg = nullptr, *d_src = nullptr; uint32_t* d_cmask = nullptr;
  uint32_t *An_c = nullptr, *Zn_c = nullptr;
  // the same compaction of the kept columns (crt), used once fewer than half are kept
  int kc_on = 0, kc_W = 0, kc_nseg = 0;
  int *d_seg_k = nullptr, *d_src_k = nullptr;
  Grow g_ank, g_znk;
  // compaction of the removed rows (rmv) for the patient dimension of At / Tt
  int rmv_W = 0, rmv_Wmax = 0;
  int* d_src_rmv = nullptr; uint32_t* d_cmask_rmv = nullptr;
  uint32_t *At_rmvc = nullptr, *Tt_rmvc = nullptr;
  // and of the retained rows (rtn); d_iota[i] = i (positions in a packed group)
  int rtn_W = 0;
  int* d_src_rtn = nullptr; uint32_t* d_cmask_rtn = nullptr;
  uint32_t *At_rtnc = nullptr, *Tt_rtnc = nullptr;
  int* d_iota = nullptr;
  // ones per (1920-block segment, list position) of the matrix-core Grams' rows
  int* gcnt[6] = {};
  Grow g_attyn, g_qrs, g_qrs2, g_ztsn, g_split, g_ctmp;
  // device arrays; (row, covariate) arrays are [t*rows + row]
  double *mlf = nullptr;
  double *lf_An_ajdk = nullptr, *lf_Zn_ajdk = nullptr;              // [jk*rows + r]
  double *jn_ajdk = nullptr, *vn_ajdk = nullptr, *kn_ajdk = nullptr, *hn_ajdk = nullptr, *an_ajdk_ = nullptr, *zn_ajdk = nullptr;
  double *lf_AtTn = nullptr, *lf_ZtSn = nullptr;
  double *lf_AnZt = nullptr, *lf_AnAt = nullptr;                      // [a*NZ + z], [a2*NA + a1] (symmetric)
  double *lf_AtTAnZtSZn = nullptr, *lf_AtTAnAtTAn = nullptr;
  double *lf_AnZtSZnAt = nullptr, *lf_AnAtTAnAt = nullptr;
  double *At_An_a1d1 = nullptr, *et_An = nullptr, *An_a2d2_ZtSn = nullptr, *An_a2d2_AtTn = nullptr;
  double *sc = nullptr, *T_AnAt_AnAt = nullptr;
  double *QR_ZSZ = nullptr, *QR_TAT = nullptr, *QC_ZSZ = nullptr, *QC_TAT = nullptr;
  double *QR_ZSZ_nrm = nullptr, *QR_TAT_nrm = nullptr, *QC_ZSZ_nrm = nullptr, *QC_TAT_nrm = nullptr;   // [t*rows + r]
  bool xij_done = false, pending_dup = false;   // stage order within a pass
  long ctl_epoch = 0, vn_epoch = -1;            // uploads of the masks; the one vn_ajdk was formed under
  bool a0_valid = false; std::vector<uint8_t> a0_Ar, a0_Ac, a0_Zr;   // the masks of stage a0's An_ajdk
  bool b0_valid = false; std::vector<uint8_t> b0_Ar, b0_Ac, b0_Zr;   // and of stage b0's products
  double *mt[14] = {};
  int* cnt_buf = nullptr;
  double *xwork = nullptr, *scal = nullptr, *dsv = nullptr, *hsv = nullptr;
  double* skey = nullptr; int *sidx_in = nullptr, *sidx = nullptr, *hidx = nullptr;   // device score sort
  void* sort_tmp = nullptr; size_t sort_tmp_bytes = 0;
  long n_sort_fallback = 0;                 // passes that needed the host quicksort
  Grow g_cz, g_planes, g_work, g_L, g_knzt, g_hnvt, g_ankt, g_knkt, g_ztsvn, g_attjn, g_ktrhn;
  Grow g_czs[4], g_planess[4];            // the quantised values of the update's four ww terms
  double* scal_s = nullptr;               // and their [min, mlt] scalars, 16 doubles each
  // host sort state and outputs
  std::vector<double> QC_sv, QR_sv; std::vector<int> QC_perm, QR_perm;
  std::vector<double> trace; std::vector<int> xdrop_a, xdrop_b;
  int n_iter = 0; bool inited = false;
  int ww_kernel = 0;                        // contractions: 0 int8 matrix cores, 1 VALU popcount
  int dcc = 0;                              // the dcc half-loop elimination (dexcluster_driver)
  long long ops = 0;
  double load_s = 0, init_s = 0, loop_s = 0;
  std::string err;

  ~Run();
  bool synthetic = false;                   // bcc_init's problem: masks and A_p drawn, not read
  std::vector<uint8_t> hAr_j0, hZr_j0, hAc_j0;   // the working masks a run starts from
  void load(const char* in_text);
  void load_synthetic(double t0);
  void load_job(const struct Pool& P, const int* rowsA, int nA, const int* rowsZ, int nZ, const uint8_t* colmask,
                const Cfg& params);
  void masks_from_files();
  void finish_masks();
  void setup_device(double t0);
  void scramble(const std::vector<int>& zcnt, int Zrb);
  void push_ctl();
  void st_ajdk(); void st_ZtSn(); void st_YnWt(); void st_At_T(); void st_QR();
  void st_single(); void st_double(); void st_flatten(); void st_dcc();
  void ztsn_weighted(const double* w, double* outZ, double* outA);
  void ww_excerpt(int trm, const Side& s1, const Side& s2, const IList& kls, const double* src, long long st1, long long st2,
                  int full, double* out, const Packed* pk = nullptr);
  void ww_quantise(int slot, const WwTerm& w);
  void ww_contract(int slot, const WwTerm& w);
  void pack_rows(const int* src, int W, uint32_t* Atc, uint32_t* Ttc);
  void mask_lfs();
  void init_lf();
  void score_lists();
  void st_nrm(); void st_ifT(); void st_srt(); void st_cmb(); void st_xij(); void st_mxA(bool dup); void st_mxdup();
  void sumscores(bool defer = false);
  void finish_trace();
  bool defer_means = false, trace_pending = false;
  size_t trace_at = 0;
  int trace_rpop = 0, trace_cpop = 0;
  void lrup();
  void lrup_QR_0(); void lrup_QR_1(); void lrup_QR_3();
  void lrup_QC_a0(); void lrup_QC_a1(); void lrup_QC_a2(); void lrup_QC_a3();
  void lrup_QC_b0(); void lrup_QC_b1(); void lrup_QC_b2(); void lrup_QC_c();
  const WwTerm* lrup_terms();
  WwTerm lrup_w[4] = {};
  void sb_drop(const std::vector<int>& rord, const std::vector<int>& cord, int rdrop, int cdrop);
  void sb_record(Scorebox& S, int tab);
  void scorebox(Scorebox& S);
  void scorebox_dcc(Scorebox& S, const std::vector<int>& rord, const std::vector<int>& cord,
                    const std::vector<int>& rdrop, const std::vector<int>& cdrop, bool rows_outer);
  void gram_any(const GramMArgs& g);
  bool active() const;
  void iterate(int* done);
  void run(int max_iter);
  void reset();
  void write_outputs(const std::string& dir) const;
  std::string default_out_dir() const;
  long debug_get(const std::string& name, std::vector<double>& out);
};

}  // namespace lkc
