// lkc_capi.cpp -- the extern "C" boundary declared in include/lkc.h.
#include "lkc_engine.h"
#include "lkc_dcc_bins.h"
#include <cmath>
#include "lkc_stages.h"

#include <string.h>
#include <string>
#include <vector>

struct lkc_bcc { lkc::Run run; lkc::Scorebox sb; bool has_sb = false; };
struct lkc_pool { lkc::Pool pool; };

static thread_local std::string g_last_error;

template <class F>
static int guard(lkc_bcc* D, F&& f)
{
  try { f(); return LKC_OK; }
  catch (const lkc::Error& e) { g_last_error = e.what(); if (D) D->run.err = e.what(); return e.code; }
  catch (const std::exception& e) { g_last_error = e.what(); if (D) D->run.err = e.what(); return LKC_ERR_HIP; }
}

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  throw lkc::Error(LKC_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ")"); } while (0)

extern "C" {

const char* lkc_last_error(void) { return g_last_error.c_str(); }
const char* lkc_bcc_error(const lkc_bcc* D) { return D ? D->run.err.c_str() : "null handle"; }
size_t lkc_bsize(int ncols) { return lkc::bsize(ncols); }

int lkc_bcc_load(const char* in_text, lkc_bcc** out)
{
  if (!out) { g_last_error = "null output pointer"; return LKC_ERR_STATE; }
  *out = nullptr;
  lkc_bcc* D = new lkc_bcc();
  int rc = guard(nullptr, [&] { D->run.load(in_text); });
  if (rc != LKC_OK) { delete D; return rc; }
  *out = D;
  return LKC_OK;
}

void lkc_bcc_free(lkc_bcc* D) { delete D; }

#define STAGE(fn, call) int fn(lkc_bcc* D) { if (!D) return LKC_ERR_STATE; return guard(D, [&] { call; }); }
STAGE(lkc_bcc_init_lf, D->run.init_lf())
STAGE(lkc_bcc_An_ajdk, D->run.st_ajdk())
STAGE(lkc_bcc_lf_ZtSn, D->run.st_ZtSn())
STAGE(lkc_bcc_lf_YnWt, D->run.st_YnWt())
STAGE(lkc_bcc_lf_At_T_YnWt_S_Zn, D->run.st_At_T())
STAGE(lkc_bcc_lf_AnZt_S_WnYt, D->run.st_QR())
STAGE(lkc_bcc_singlestudy_ww, D->run.st_single())
STAGE(lkc_bcc_doublestudy_ww, D->run.st_double())
// the initial pass is complete once bcc_flattenloop has run (lakcluster_driver.c:104)
STAGE(lkc_bcc_flattenloop, D->run.st_flatten(); D->run.inited = true)
#undef STAGE

// the stages of one pass of the elimination loop (lakcluster_driver.c:107-246); they need
// the initial pass
#define LOOP_STAGE(fn, call) int fn(lkc_bcc* D) { if (!D) return LKC_ERR_STATE; return guard(D, [&] { \
  if (!D->run.inited) throw lkc::Error(LKC_ERR_STATE, #fn " needs the initial pass (through bcc_flattenloop)"); \
  call; }); }
LOOP_STAGE(lkc_bcc_sumscores, D->run.sumscores())
LOOP_STAGE(lkc_bcc_lrup, D->run.lrup(); D->run.st_mxdup())
LOOP_STAGE(lkc_bcc_sumscores_nrm, D->run.st_nrm())
LOOP_STAGE(lkc_bcc_sumscores_ifT, D->run.st_ifT())
LOOP_STAGE(lkc_bcc_sumscores_srt, D->run.st_srt())
LOOP_STAGE(lkc_bcc_sumscores_cmb, D->run.st_cmb())
LOOP_STAGE(lkc_bcc_sumscores_xij, D->run.st_xij())
LOOP_STAGE(lkc_bcc_sumscores_mxA, D->run.st_mxA(false))
LOOP_STAGE(lkc_bcc_lrup_QR_YnWt_stage_0, D->run.lrup_QR_0())
LOOP_STAGE(lkc_bcc_lrup_QR_YnWt_stage_1, D->run.lrup_QR_1())
LOOP_STAGE(lkc_bcc_lrup_QR_YnWt_stage_3, D->run.lrup_QR_3())
LOOP_STAGE(lkc_bcc_lrup_QC_ZtSWn_stage_0, (void)0)
LOOP_STAGE(lkc_bcc_lrup_QC_ZtSWn_stage_1, (void)0)
LOOP_STAGE(lkc_bcc_lrup_QC_ZtSWn_stage_2, (void)0)
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_a0, D->run.lrup_QC_a0())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_a1, D->run.lrup_QC_a1())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_a2, D->run.lrup_QC_a2())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_a3, D->run.lrup_QC_a3())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_b0, D->run.lrup_QC_b0())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_b1, D->run.lrup_QC_b1())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_b2, D->run.lrup_QC_b2())
LOOP_STAGE(lkc_bcc_lrup_QC_YnWt_stage_c, D->run.lrup_QC_c())
LOOP_STAGE(lkc_bcc_lrup_mxdup, D->run.st_mxdup())
LOOP_STAGE(lkc_bcc_M_mxset, (void)0)
LOOP_STAGE(lkc_bcc_M_ZtSWn, (void)0)
LOOP_STAGE(lkc_bcc_M_YnWt, (void)0)
LOOP_STAGE(lkc_bcc_lf_AtTYn____WtSZn, (void)0)
#undef LOOP_STAGE

int lkc_bcc_sumscores_dmp(const lkc_bcc* D, const char* out_dir) { return lkc_bcc_write_outputs(D, out_dir); }

int lkc_bcc_active(const lkc_bcc* D) { return D && D->run.active() ? 1 : 0; }

unsigned int lkc_dquicksort_xij(double* d, int* idx, int l, int r)
{
  if (!d || !idx || l > r) return 0;
  return lkc::quicksort_xij(d, idx, l, r);
}

int lkc_get_xdrop(double lrij, double lcij, double gamma, int gamma_type, int* rdrop, int* cdrop)
{
  if (!std::isfinite(lrij) || !std::isfinite(lcij)) { g_last_error = "get_xdrop: non-finite row or column count"; return LKC_ERR_STATE; }
  lkc::get_xdrop(lrij, lcij, gamma, gamma_type, rdrop, cdrop);
  return LKC_OK;
}

int lkc_get_xdrop_length(double lrij, double lcij, double gamma, int gamma_type)
{
  if (!std::isfinite(lrij) || !std::isfinite(lcij)) { g_last_error = "get_xdrop_length: non-finite row or column count"; return LKC_ERR_STATE; }
  return lkc::get_xdrop_length(lrij, lcij, gamma, gamma_type);
}

int lkc_get_xdrop_array(double lrij, double lcij, double gamma, int gamma_type, int cap, int* length, int* rdrop,
                        int* cdrop, int* rkeep, int* ckeep)
{
  const int n = lkc_get_xdrop_length(lrij, lcij, gamma, gamma_type);
  if (n < 0) return n;
  if (length) *length = n;
  if (!rdrop && !cdrop && !rkeep && !ckeep) return LKC_OK;
  if (cap < n) { g_last_error = "get_xdrop_array: the arrays hold fewer entries than the schedule"; return LKC_ERR_STATE; }
  double r = lrij, c = lcij;
  for (int i = 0; i < n; i++) {
    int rd = 0, cd = 0;
    lkc::get_xdrop(r, c, gamma, gamma_type, &rd, &cd);
    r -= rd; c -= cd;
    if (rdrop) rdrop[i] = rd;
    if (cdrop) cdrop[i] = cd;
    if (rkeep) rkeep[i] = (int)std::round(r);
    if (ckeep) ckeep[i] = (int)std::round(c);
  }
  return LKC_OK;
}

int lkc_bcc_iterate(lkc_bcc* D, int* done)
{
  if (!D) return LKC_ERR_STATE;
  return guard(D, [&] { D->run.iterate(done); });
}

int lkc_bcc_run(lkc_bcc* D, int max_iter)
{
  if (!D) return LKC_ERR_STATE;
  return guard(D, [&] { D->run.run(max_iter); });
}

int lkc_bcc_write_outputs(const lkc_bcc* D, const char* out_dir)
{
  if (!D) return LKC_ERR_STATE;
  return guard(const_cast<lkc_bcc*>(D), [&] {
    D->run.write_outputs(out_dir && out_dir[0] ? std::string(out_dir) : D->run.default_out_dir());
  });
}

int lkc_lakcluster_driver(const char* in_text, const char* out_dir)
{
  lkc_bcc* D = nullptr;
  int rc = lkc_bcc_load(in_text, &D);
  if (rc) return rc;
  D->run.dcc = 0;
  rc = lkc_bcc_run(D, -1);
  if (!rc) rc = lkc_bcc_write_outputs(D, out_dir);
  lkc_bcc_free(D);
  return rc;
}

int lkc_bcc_scorebox(lkc_bcc* D)
{
  if (!D) return LKC_ERR_STATE;
  return guard(D, [&] { D->has_sb = false; D->run.scorebox(D->sb); D->has_sb = true; });
}

int lkc_bcc_get_scorebox(const lkc_bcc* D, int* dims4, int* rpop, int* cpop, int* irem, double* qr, double* qc)
{
  if (!D || !dims4 || !D->has_sb) return LKC_ERR_STATE;
  const lkc::Scorebox& S = D->sb;
  dims4[0] = S.row_max; dims4[1] = S.row_num; dims4[2] = S.col_max; dims4[3] = S.col_num;
  const size_t n = (size_t)S.row_num * S.col_num;
  if (rpop) memcpy(rpop, S.rpop.data(), sizeof(int) * n);
  if (cpop) memcpy(cpop, S.cpop.data(), sizeof(int) * n);
  if (irem) memcpy(irem, S.irem.data(), sizeof(int) * n);
  if (qr) memcpy(qr, S.qr.data(), sizeof(double) * n);
  if (qc) memcpy(qc, S.qc.data(), sizeof(double) * n);
  return LKC_OK;
}

int lkc_bcc_write_scorebox(const lkc_bcc* D, const char* out_dir)
{
  if (!D || !D->has_sb) return LKC_ERR_STATE;
  return guard(const_cast<lkc_bcc*>(D), [&] {
    lkc::write_scorebox(D->sb, out_dir && out_dir[0] ? std::string(out_dir) : D->run.default_out_dir());
  });
}

int lkc_bcc_is_scorebox(const lkc_bcc* D)
{
  return D && (D->run.c.test_type == "lakcluster_scorebox" || D->run.c.test_type == "dexcluster_scorebox") ? 1 : 0;
}

static int scorebox_run(const char* in_text, const char* out_dir, int dcc)
{
  lkc_bcc* D = nullptr;
  int rc = lkc_bcc_load(in_text, &D);
  if (rc) return rc;
  D->run.dcc = dcc;
  rc = lkc_bcc_scorebox(D);
  if (!rc) rc = lkc_bcc_write_scorebox(D, out_dir);
  lkc_bcc_free(D);
  return rc;
}

int lkc_lakcluster_scorebox(const char* in_text, const char* out_dir) { return scorebox_run(in_text, out_dir, 0); }
int lkc_dexcluster_scorebox(const char* in_text, const char* out_dir) { return scorebox_run(in_text, out_dir, 1); }

int lkc_scorebox_xdrop_array(int nxols, int x_max, int x_num, int* xval, int* xdrop)
{
  if (x_num < 1 || !xval || !xdrop) { g_last_error = "x_num must be >= 1 with both outputs given"; return LKC_ERR_STATE; }
  lkc::scorebox_xdrop_array(nxols, x_max, x_num, xval, xdrop);
  return LKC_OK;
}

int lkc_dcc_driver_bins(const char* in_text, const char* out_dir)
{
  return guard(nullptr, [&] {
    lkc::Cfg c;
    lkc::parse_cfg(in_text, c);
    lkc::DccTestOut out;
    lkc::dcc_driver_bins(c, out, out_dir && out_dir[0] ? std::string(out_dir) : lkc::out_dir_of(c));
  });
}

int lkc_dexcluster_driver(const char* in_text, const char* out_dir)
{
  // several bins or Y/W genetic controls: the multi-bin elimination; else the one-bin engine
  bool bins = false;
  int rc = guard(nullptr, [&] {
    lkc::Cfg c;
    lkc::parse_cfg(in_text, c);
    bins = c.nbins > 1 || c.Y_ncols > 0;
    for (const std::string& f : c.Y_t_) bins = bins || !f.empty();
  });
  if (rc) return rc;
  if (bins) return lkc_dcc_driver_bins(in_text, out_dir);
  lkc_bcc* D = nullptr;
  rc = lkc_bcc_load(in_text, &D);
  if (rc) return rc;
  D->run.dcc = 1;
  rc = lkc_bcc_run(D, -1);
  if (!rc) rc = lkc_bcc_write_outputs(D, out_dir);
  lkc_bcc_free(D);
  return rc;
}

int lkc_bcc_is_dcc(const lkc_bcc* D) { return D && D->run.dcc ? 1 : 0; }

int lkc_dcc_sumscores_test(const char* in_text, const char* out_dir)
{
  return guard(nullptr, [&] {
    lkc::Cfg c;
    lkc::parse_cfg(in_text, c);
    lkc::DccTestOut out;
    lkc::dcc_sumscores_test(c, out, out_dir && out_dir[0] ? std::string(out_dir) : lkc::out_dir_of(c));
  });
}

int lkc_bcc_sumscores_test(const char* in_text, const char* out_dir)
{
  lkc_bcc* D = nullptr;
  int rc = lkc_bcc_load(in_text, &D);
  if (rc) return rc;
  if (!D->run.synthetic) {
    g_last_error = "bcc_sumscores_test needs GLOBAL_TEST_TYPE= bcc_sumscores";
    lkc_bcc_free(D);
    return LKC_ERR_CONFIG;
  }
  rc = lkc_bcc_run(D, D->run.c.test_niter > 0 ? D->run.c.test_niter : -1);
  if (!rc) rc = lkc_bcc_write_outputs(D, out_dir);
  lkc_bcc_free(D);
  return rc;
}

int lkc_bcc_n_iter(const lkc_bcc* D) { return D ? D->run.n_iter : -1; }
int lkc_bcc_n_trace(const lkc_bcc* D) { return D ? (int)(D->run.trace.size() / 6) : -1; }
int lkc_bcc_get_trace(const lkc_bcc* D, double* out)
{
  if (!D || !out) return LKC_ERR_STATE;
  if (!D->run.trace.empty()) memcpy(out, D->run.trace.data(), sizeof(double) * D->run.trace.size());
  return LKC_OK;
}
int lkc_bcc_n_xdrop(const lkc_bcc* D) { return D ? (int)(D->run.xdrop_a.size() / 2) : -1; }
int lkc_bcc_get_xdrop(const lkc_bcc* D, int* a, int* b)
{
  if (!D || !a || !b) return LKC_ERR_STATE;
  if (!D->run.xdrop_a.empty()) {
    memcpy(a, D->run.xdrop_a.data(), sizeof(int) * D->run.xdrop_a.size());
    memcpy(b, D->run.xdrop_b.data(), sizeof(int) * D->run.xdrop_b.size());
  }
  return LKC_OK;
}
long long lkc_bcc_ops(const lkc_bcc* D) { return D ? D->run.ops : -1; }
long lkc_bcc_sort_fallbacks(const lkc_bcc* D) { return D ? D->run.n_sort_fallback : -1; }
int lkc_bcc_times(const lkc_bcc* D, double* t3)
{
  if (!D || !t3) return LKC_ERR_STATE;
  t3[0] = D->run.load_s; t3[1] = D->run.init_s; t3[2] = D->run.loop_s;
  return LKC_OK;
}
int lkc_bcc_shape(const lkc_bcc* D, int* o)
{
  if (!D || !o) return LKC_ERR_STATE;
  o[0] = D->run.NA; o[1] = D->run.NZ; o[2] = D->run.M; o[3] = D->run.T;
  return LKC_OK;
}

int lkc_bcc_reset(lkc_bcc* D)
{
  if (!D) return LKC_ERR_STATE;
  return guard(D, [&] { D->run.reset(); });
}

int lkc_bcc_set_profiling(lkc_bcc* D, int on)
{
  if (!D) return LKC_ERR_STATE;
  D->run.prof.on = on != 0;
  return LKC_OK;
}

int lkc_bcc_set_kernels(lkc_bcc* D, int which)
{
  if (!D || which < 0 || which > 1) return LKC_ERR_STATE;
  D->run.ww_kernel = which;
  return LKC_OK;
}

int lkc_tune_set(int mode)
{
  if (mode < 0 || mode > 3) { g_last_error = "tuning mode must be 0, 1, 2 or 3"; return LKC_ERR_STATE; }
  lkc::tune_set_mode(mode);
  return LKC_OK;
}

int lkc_tune_get(int* mode, int* classes, long long* measured, double* saved_ms)
{
  long long m = 0;
  double s = 0;
  lkc::tune_stats(&m, &s);
  if (mode) *mode = lkc::tune_mode();
  if (classes) *classes = lkc::tune_entries();
  if (measured) *measured = m;
  if (saved_ms) *saved_ms = s;
  return LKC_OK;
}

int lkc_tune_verify(long long* compared, long long* mismatched)
{
  long long c = 0, m = 0;
  lkc::tune_verify_stats(&c, &m);
  if (compared) *compared = c;
  if (mismatched) *mismatched = m;
  return LKC_OK;
}

int lkc_tune_save(const char* path)
{
  if (!path) { g_last_error = "null path"; return LKC_ERR_STATE; }
  const int n = lkc::tune_save(path);
  if (n < 0) { g_last_error = std::string("cannot write ") + path; return LKC_ERR_IO; }
  return n;
}

int lkc_tune_load(const char* path)
{
  if (!path) { g_last_error = "null path"; return LKC_ERR_STATE; }
  const int n = lkc::tune_load(path);
  if (n < 0) { g_last_error = std::string("cannot read ") + path; return LKC_ERR_IO; }
  return n;
}

int lkc_tune_clear(void)
{
  lkc::tune_clear();
  return LKC_OK;
}

int lkc_bcc_kernel_stats(const lkc_bcc* D, const char* name, long long* launches, double* ms, double* work)
{
  if (!D || !name) return LKC_ERR_STATE;
  for (const auto& k : D->run.prof.stats) if (k.name == name) {
    if (launches) *launches = k.launches;
    if (ms) *ms = k.ms;
    if (work) *work = k.work;
    return LKC_OK;
  }
  if (launches) *launches = 0;
  if (ms) *ms = 0;
  if (work) *work = 0;
  return LKC_OK;
}

long lkc_bcc_debug_get(lkc_bcc* D, const char* name, double* out, long n)
{
  if (!D || !name) return -1;
  long len = -1;
  std::vector<double> v;
  if (guard(D, [&] { len = D->run.debug_get(name, v); }) != LKC_OK) return -1;
  if (len > 0 && out) memcpy(out, v.data(), sizeof(double) * (size_t)std::min(len, n));
  return len;
}

}  // extern "C"

// ---- single kernels on host buffers ---------------------------------------
namespace {
struct Scoped {
  lkc::Dev dev; hipStream_t st = nullptr;
  Scoped() { HIPCHK(hipSetDevice(0)); HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); }
  ~Scoped() { if (st) (void)hipStreamDestroy(st); }
  template <class T> T* put(const T* h, size_t n)
  {
    T* d = dev.alloc<T>(n ? n : 1);
    if (n) HIPCHK(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return d;
  }
};
// repack caller rows (stride rb) into rows of bsize(ncols) bytes
static lkc::DMat put_bits(Scoped& s, const uint8_t* X, size_t rb, int nrows, int ncols, std::vector<uint8_t>& tmp)
{
  size_t b = lkc::bsize(ncols);
  if (rb < b) throw lkc::Error(LKC_ERR_STATE, "row stride smaller than bsize(ncols)");
  tmp.assign((size_t)std::max(nrows, 1) * b, 0);
  for (int r = 0; r < nrows; r++) memcpy(tmp.data() + (size_t)r * b, X + (size_t)r * rb, b);
  lkc::DMat m; m.nr = nrows; m.nc = ncols; m.rb = b;
  m.p = s.put(tmp.data(), tmp.size());
  return m;
}
static void check_rows(const int* rows, int n, int total)
{
  for (int i = 0; i < n; i++) if (rows[i] < 0 || rows[i] >= total) throw lkc::Error(LKC_ERR_STATE, "row index out of range");
}
}  // namespace

extern "C" {

int lkc_An_ajdk_v(const uint8_t* X, size_t rb, int nrows_total, const int* rows, int nrows,
                  const uint8_t* colmask, int ncols, const double* ajdk, double* out)
{
  return guard(nullptr, [&] {
    if (nrows < 0 || ncols <= 0 || !X || !colmask || !ajdk || !out || (nrows && !rows)) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    check_rows(rows, nrows, nrows_total);
    Scoped s;
    int pc = lkc::psize(ncols);
    std::vector<uint8_t> tmp;
    lkc::DMat dX = put_bits(s, X, rb, nrows_total, ncols, tmp);
    uint32_t* dm = (uint32_t*)s.put(colmask, lkc::bsize(ncols));
    int* dr = s.put(rows, (size_t)nrows);
    double* da = s.put(ajdk, (size_t)pc * lkc::kAjdk);
    double* base = s.dev.alloc<double>(16);
    double* dout = s.dev.alloc<double>((size_t)lkc::kAjdk * std::max(nrows_total, 1));
    lkc::mask_lf(dm, pc, da, base, s.st);
    lkc::an_ajdk(dX.b(), dr, nrows, dm, pc, da, base, dout, nrows_total, s.st);
    std::vector<double> h((size_t)lkc::kAjdk * std::max(nrows_total, 1));
    HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipStreamSynchronize(s.st));
    for (int i = 0; i < nrows; i++)
      for (int jk = 0; jk < lkc::kAjdk; jk++) out[(size_t)i * lkc::kAjdk + jk] = h[(size_t)jk * nrows_total + rows[i]];
  });
}

int lkc_AnZt_vv(const uint8_t* A, size_t rbA, int na_total, const int* arows, int na,
                const uint8_t* Z, size_t rbZ, int nz_total, const int* zrows, int nz,
                const uint8_t* colmask, int ncols, const double* ajdk,
                const double* xa, const double* xz, double* out)
{
  return guard(nullptr, [&] {
    if (na < 0 || nz < 0 || ncols <= 0 || !A || !Z || !colmask || !out) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    if (ajdk && (!xa || !xz)) throw lkc::Error(LKC_ERR_STATE, "xa/xz required with ajdk");
    check_rows(arows, na, na_total); check_rows(zrows, nz, nz_total);
    Scoped s;
    int pc = lkc::psize(ncols);
    std::vector<uint8_t> ta, tz;
    lkc::DMat dA = put_bits(s, A, rbA, na_total, ncols, ta);
    lkc::DMat dZ = put_bits(s, Z, rbZ, nz_total, ncols, tz);
    uint32_t* dm = (uint32_t*)s.put(colmask, lkc::bsize(ncols));
    int* dar = s.put(arows, (size_t)na);
    int* dzr = s.put(zrows, (size_t)nz);
    double *da = nullptr, *dxa = nullptr, *dxz = nullptr;
    if (ajdk) {
      da = s.put(ajdk, (size_t)pc * lkc::kAjdk);
      std::vector<double> fa(std::max(na_total, 1), 0.0), fz(std::max(nz_total, 1), 0.0);
      for (int i = 0; i < na; i++) fa[arows[i]] = xa[i];
      for (int j = 0; j < nz; j++) fz[zrows[j]] = xz[j];
      dxa = s.put(fa.data(), fa.size());
      dxz = s.put(fz.data(), fz.size());
    }
    double* base = s.dev.alloc<double>(16);
    double* dout = s.dev.alloc<double>((size_t)std::max(na, 1) * std::max(nz, 1));
    lkc::mask_lf(dm, pc, da, base, s.st);
    lkc::gram(dA.b(), dar, na, dZ.b(), dzr, nz, dm, pc, da, base, dxa, dxz, ajdk ? 1 : 0, dout, na, 0, 0, 0, s.st);
    std::vector<double> h((size_t)std::max(na, 1) * std::max(nz, 1));
    HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipStreamSynchronize(s.st));
    for (int i = 0; i < na; i++) for (int j = 0; j < nz; j++) out[(size_t)i * nz + j] = h[(size_t)j * na + i];
  });
}

int lkc_xcalc(const double* L, const int* nrows_list, int nn, int n_total, int nw, int B,
              long long D_MLT, unsigned long long* cz, double* min_mlt)
{
  return guard(nullptr, [&] {
    if (nn < 0 || nw < 0 || B < 1 || B > 32 || !min_mlt || n_total < nn || (nn * (long long)nw && (!L || !cz || !nrows_list)))
      throw lkc::Error(LKC_ERR_STATE, "bad argument");
    for (int i = 0; i < nn; i++)
      if (nrows_list[i] < 0 || nrows_list[i] >= n_total) throw lkc::Error(LKC_ERR_STATE, "nrows_list entry outside [0, n_total)");
    Scoped s;
    double* dL = s.put(L, (size_t)nn * nw);
    int* dn = nn ? s.put(nrows_list, (size_t)nn) : nullptr;
    const size_t ncz = (size_t)std::max(n_total, 1) * std::max(nw, 1);
    uint32_t* dcz = s.dev.alloc<uint32_t>(ncz);
    HIPCHK(hipMemsetAsync(dcz, 0, sizeof(uint32_t) * ncz, s.st));
    double* scal = s.dev.alloc<double>(16);
    double* work = s.dev.alloc<double>(1024);
    lkc::xcalc(dL, nw, 1, nullptr, nullptr, nn, nw, dn, B, D_MLT, dcz, std::max(n_total, 1), scal, work, s.st);
    std::vector<uint32_t> h(ncz);
    double hs[16];
    HIPCHK(hipMemcpyAsync(h.data(), dcz, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipMemcpyAsync(hs, scal, sizeof(hs), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipStreamSynchronize(s.st));
    for (int j = 0; j < nw; j++) for (int n = 0; n < n_total; n++) cz[(size_t)j * n_total + n] = h[(size_t)j * std::max(n_total, 1) + n];
    min_mlt[0] = hs[2]; min_mlt[1] = hs[4];
  });
}

static lkc::Cfg job_cfg(const lkc_job_params* p)
{
  lkc::Cfg c;
  if (p) {
    c.gamma = p->gamma; c.gamma_type = p->gamma_type; c.B = p->B_MLT; c.kappa2 = p->kappa_squared; c.sparse = p->sparse;
    c.D_MLT = (long long)pow(2.0, (double)c.B);
  }
  return c;
}

int lkc_pool_create(const uint8_t* rows, size_t rb, int nrows, int ncols, const uint8_t* cov, size_t rbc, int ncov,
                    lkc_pool** out)
{
  if (!out) { g_last_error = "null output pointer"; return LKC_ERR_STATE; }
  *out = nullptr;
  lkc_pool* P = new lkc_pool();
  int rc = guard(nullptr, [&] {
    if (nrows <= 0 || ncols <= 0 || ncov < 1 || ncov > lkc::kMaxT || !rows || !cov) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    const size_t b = lkc::bsize(ncols), bc = lkc::bsize(nrows);
    if (rb < b || rbc < bc) throw lkc::Error(LKC_ERR_STATE, "row stride smaller than bsize");
    HIPCHK(hipSetDevice(0));
    lkc::Pool& p = P->pool;
    p.P = nrows; p.M = ncols; p.T = ncov;
    std::vector<uint8_t> h((size_t)nrows * b);
    for (int r = 0; r < nrows; r++) memcpy(h.data() + (size_t)r * b, rows + (size_t)r * rb, b);
    p.Pn.nr = nrows; p.Pn.nc = ncols; p.Pn.rb = b; p.Pn.p = p.dev.alloc<uint8_t>(h.size());
    HIPCHK(hipMemcpy(p.Pn.p, h.data(), h.size(), hipMemcpyHostToDevice));
    std::vector<uint8_t> hc((size_t)ncov * bc);
    for (int t = 0; t < ncov; t++) memcpy(hc.data() + (size_t)t * bc, cov + (size_t)t * rbc, bc);
    p.Tp.nr = ncov; p.Tp.nc = nrows; p.Tp.rb = bc; p.Tp.p = p.dev.alloc<uint8_t>(hc.size());
    HIPCHK(hipMemcpy(p.Tp.p, hc.data(), hc.size(), hipMemcpyHostToDevice));
  });
  if (rc != LKC_OK) { delete P; return rc; }
  *out = P;
  return LKC_OK;
}

int lkc_pool_create_random(int nrows, int ncols, double pmin, double pmax, unsigned long long seed, int nplant,
                           const int* plant_m, const int* plant_mc, const int* plant_rows, const int* plant_cols,
                           const uint8_t* plant_bits, const uint8_t* cov, size_t rbc, int ncov, lkc_pool** out)
{
  if (!out) { g_last_error = "null output pointer"; return LKC_ERR_STATE; }
  *out = nullptr;
  lkc_pool* P = new lkc_pool();
  int rc = guard(nullptr, [&] {
    if (nrows <= 0 || ncols <= 0 || ncov < 1 || ncov > lkc::kMaxT || !cov || nplant < 0 || !(pmin <= pmax) ||
        (nplant && (!plant_m || !plant_mc || !plant_rows || !plant_cols || !plant_bits)))
      throw lkc::Error(LKC_ERR_STATE, "bad argument");
    const size_t b = lkc::bsize(ncols), bc = lkc::bsize(nrows);
    if (rbc < bc) throw lkc::Error(LKC_ERR_STATE, "row stride smaller than bsize");
    size_t nr = 0, nc = 0;
    for (int k = 0; k < nplant; k++) {
      if (plant_m[k] < 0 || plant_mc[k] < 0) throw lkc::Error(LKC_ERR_STATE, "bad block size");
      nr += plant_m[k]; nc += plant_mc[k];
    }
    for (size_t i = 0; i < nr; i++) if (plant_rows[i] < 0 || plant_rows[i] >= nrows) throw lkc::Error(LKC_ERR_STATE, "block row out of range");
    for (size_t i = 0; i < nc; i++) if (plant_cols[i] < 0 || plant_cols[i] >= ncols) throw lkc::Error(LKC_ERR_STATE, "block column out of range");
    HIPCHK(hipSetDevice(0));
    lkc::Pool& p = P->pool;
    p.P = nrows; p.M = ncols; p.T = ncov;
    p.Pn.nr = nrows; p.Pn.nc = ncols; p.Pn.rb = b; p.Pn.p = p.dev.alloc<uint8_t>((size_t)nrows * b);
    Scoped s;
    double* pb = s.dev.alloc<double>((size_t)ncols);
    lkc::pool_draw(p.Pn.p, b, nrows, ncols, pmin, pmax, seed, pb, s.st);
    size_t orr = 0, oc = 0, ob = 0;
    for (int k = 0; k < nplant; k++) {
      int* dr = s.put(plant_rows + orr, (size_t)plant_m[k]);
      int* dc = s.put(plant_cols + oc, (size_t)plant_mc[k]);
      uint8_t* db = s.put(plant_bits + ob, (size_t)plant_m[k] * plant_mc[k]);
      lkc::pool_plant(p.Pn.p, b, dr, dc, db, plant_m[k], plant_mc[k], s.st);
      orr += plant_m[k]; oc += plant_mc[k]; ob += (size_t)plant_m[k] * plant_mc[k];
    }
    HIPCHK(hipStreamSynchronize(s.st));
    HIPCHK(hipGetLastError());
    std::vector<uint8_t> hc((size_t)ncov * bc);
    for (int t = 0; t < ncov; t++) memcpy(hc.data() + (size_t)t * bc, cov + (size_t)t * rbc, bc);
    p.Tp.nr = ncov; p.Tp.nc = nrows; p.Tp.rb = bc; p.Tp.p = p.dev.alloc<uint8_t>(hc.size());
    HIPCHK(hipMemcpy(p.Tp.p, hc.data(), hc.size(), hipMemcpyHostToDevice));
  });
  if (rc != LKC_OK) { delete P; return rc; }
  *out = P;
  return LKC_OK;
}

int lkc_pool_read(const lkc_pool* P, int r0, int n, uint8_t* out)
{
  return guard(nullptr, [&] {
    if (!P || !out || r0 < 0 || n < 0 || r0 + (long long)n > P->pool.P) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    const lkc::DMat& m = P->pool.Pn;
    if (n) HIPCHK(hipMemcpy(out, m.p + (size_t)r0 * m.rb, (size_t)n * m.rb, hipMemcpyDeviceToHost));
  });
}

void lkc_pool_free(lkc_pool* P) { delete P; }

int lkc_bcc_load_job(const lkc_pool* P, const int* rowsA, int nA, const int* rowsZ, int nZ, const uint8_t* colmask,
                     const lkc_job_params* prm, lkc_bcc** out)
{
  if (!out || !P) { g_last_error = "null pointer"; return LKC_ERR_STATE; }
  *out = nullptr;
  lkc_bcc* D = new lkc_bcc();
  int rc = guard(nullptr, [&] { D->run.load_job(P->pool, rowsA, nA, rowsZ, nZ, colmask, job_cfg(prm)); });
  if (rc != LKC_OK) { delete D; return rc; }
  *out = D;
  return LKC_OK;
}

double lkc_job_bytes(int nA, int nZ, int ncols, int B_MLT) { return lkc::job_bytes(nA, nZ, ncols, B_MLT); }

int lkc_batch_run(const lkc_pool* P, const lkc_job* jobs, int njobs, int max_concurrent, double hbm_budget_bytes,
                  lkc_job_result* results)
{
  return guard(nullptr, [&] {
    if (!P || njobs < 0 || (njobs && (!jobs || !results))) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    if (hbm_budget_bytes <= 0) {
      // free device memory plus what the stream-ordered pool holds idle from earlier runs
      // (its release threshold keeps freed blocks for reuse, so hipMemGetInfo counts them
      // as used although every job can take them)
      size_t fr = 0, tot = 0;
      HIPCHK(hipSetDevice(0));
      HIPCHK(hipMemGetInfo(&fr, &tot));
      double idle = 0.0;
      hipMemPool_t mp;
      if (hipDeviceGetDefaultMemPool(&mp, 0) == hipSuccess) {
        uint64_t reserved = 0, used = 0;
        if (hipMemPoolGetAttribute(mp, hipMemPoolAttrReservedMemCurrent, &reserved) == hipSuccess &&
            hipMemPoolGetAttribute(mp, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess && reserved > used)
          idle = (double)(reserved - used);
      }
      hbm_budget_bytes = 0.8 * ((double)fr + idle);
    }
    std::vector<lkc::JobIn> in(njobs);
    for (int j = 0; j < njobs; j++) {
      in[j].rowsA = jobs[j].rowsA; in[j].nA = jobs[j].nA; in[j].rowsZ = jobs[j].rowsZ; in[j].nZ = jobs[j].nZ;
      in[j].colmask = jobs[j].colmask; in[j].cfg = job_cfg(&jobs[j].params); in[j].arrival_s = jobs[j].arrival_s;
    }
    std::vector<lkc::JobOut> o(njobs);
    lkc::run_batch(P->pool, in.data(), njobs, max_concurrent, hbm_budget_bytes, o.data());
    for (int j = 0; j < njobs; j++) {
      lkc_job_result& r = results[j];
      memset(&r, 0, sizeof r);
      r.status = o[j].status; r.n_iter = o[j].n_iter; r.ops = o[j].ops;
      r.arrival_s = o[j].arrival_s; r.start_s = o[j].start_s; r.end_s = o[j].end_s;
      r.n_trace = (long)(o[j].trace.size() / 6);
      r.trace = (double*)malloc(sizeof(double) * std::max<size_t>(o[j].trace.size(), 1));
      if (!o[j].trace.empty()) memcpy(r.trace, o[j].trace.data(), sizeof(double) * o[j].trace.size());
      r.n_xdrop = (long)(o[j].xdrop_a.size() / 2);
      r.xdrop_a = (int*)malloc(sizeof(int) * std::max<size_t>(o[j].xdrop_a.size(), 1));
      r.xdrop_b = (int*)malloc(sizeof(int) * std::max<size_t>(o[j].xdrop_b.size(), 1));
      if (!o[j].xdrop_a.empty()) memcpy(r.xdrop_a, o[j].xdrop_a.data(), sizeof(int) * o[j].xdrop_a.size());
      if (!o[j].xdrop_b.empty()) memcpy(r.xdrop_b, o[j].xdrop_b.data(), sizeof(int) * o[j].xdrop_b.size());
      if (o[j].status != LKC_OK) g_last_error = o[j].err;
    }
  });
}

void lkc_batch_results_free(lkc_job_result* results, int njobs)
{
  if (!results) return;
  for (int j = 0; j < njobs; j++) {
    free(results[j].trace); free(results[j].xdrop_a); free(results[j].xdrop_b);
    results[j].trace = nullptr; results[j].xdrop_a = results[j].xdrop_b = nullptr;
  }
}

int lkc_transpose_bits(const uint8_t* X, size_t rbX, int nrows, int ncols, uint8_t* Y, size_t rbY)
{
  return guard(nullptr, [&] {
    if (nrows < 0 || ncols < 0 || (nrows && ncols && (!X || !Y))) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    if (nrows == 0 || ncols == 0) return;
    if (rbY < lkc::bsize(nrows)) throw lkc::Error(LKC_ERR_STATE, "row stride smaller than bsize(nrows)");
    Scoped s;
    std::vector<uint8_t> tmp;
    lkc::DMat dX = put_bits(s, X, rbX, nrows, ncols, tmp);
    const size_t rbd = lkc::bsize(nrows);
    uint8_t* dY = s.dev.alloc<uint8_t>((size_t)ncols * rbd);
    HIPCHK(hipMemsetAsync(dY, 0, (size_t)ncols * rbd, s.st));
    lkc::transpose_bits(dX.b(), dY, rbd, s.st);
    std::vector<uint8_t> h((size_t)ncols * rbd);
    HIPCHK(hipMemcpyAsync(h.data(), dY, h.size(), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipStreamSynchronize(s.st));
    for (int j = 0; j < ncols; j++) {
      memcpy(Y + (size_t)j * rbY, h.data() + (size_t)j * rbd, rbd);
      if (rbY > rbd) memset(Y + (size_t)j * rbY + rbd, 0, rbY - rbd);
    }
  });
}

int lkc_An_ZtSn_ww(const uint8_t* A, size_t rbA, int nr_total, const int* alist, int na, const uint8_t* amask,
                   const double* L, const int* nlist, int nn, int n_total, int ns, int B, long long D_MLT, double* out)
{
  return guard(nullptr, [&] {
    if (na < 0 || nn < 0 || ns < 0 || n_total <= 0 || B < 1 || B > 32 || !A || !amask || !out) throw lkc::Error(LKC_ERR_STATE, "bad argument");
    check_rows(alist, na, nr_total); check_rows(nlist, nn, n_total);
    Scoped s;
    std::vector<uint8_t> ta;
    lkc::DMat dA = put_bits(s, A, rbA, nr_total, n_total, ta);
    size_t prb = dA.rb;
    std::vector<uint8_t> hm(prb, 0);
    memcpy(hm.data(), amask, lkc::bsize(n_total));
    uint32_t* dm = (uint32_t*)s.put(hm.data(), prb);
    int* dal = s.put(alist, (size_t)na);
    int* dnl = s.put(nlist, (size_t)nn);
    double* dL = s.put(L, (size_t)nn * ns);
    std::vector<int> sl(std::max(ns, 1));
    for (int i = 0; i < ns; i++) sl[i] = i;
    int* dsl = s.put(sl.data(), sl.size());
    // the active n rows: the bits of nlist
    std::vector<uint8_t> row0(prb, 0);
    for (int i = 0; i < nn; i++) row0[nlist[i] >> 3] |= (uint8_t)(1u << (7 - (nlist[i] & 7)));
    uint32_t* drow0 = (uint32_t*)s.put(row0.data(), prb);
    uint32_t* cz = s.dev.alloc<uint32_t>((size_t)std::max(ns, 1) * n_total);
    uint8_t* planes = s.dev.alloc<uint8_t>((size_t)std::max(ns, 1) * B * prb);
    double* scal = s.dev.alloc<double>(16);
    double* work = s.dev.alloc<double>(1024);
    double* dout = s.dev.alloc<double>((size_t)std::max(ns, 1) * std::max(nr_total, 1));
    lkc::fill(dout, (long long)std::max(ns, 1) * std::max(nr_total, 1), 0.0, s.st);
    lkc::xcalc(dL, ns, 1, nullptr, nullptr, nn, ns, dnl, B, D_MLT, cz, n_total, scal, work, s.st);
    lkc::cz_planes(cz, n_total, ns, n_total, drow0, B, planes, prb, s.st);
    lkc::an_ztsn(dA.b(), dal, na, dm, planes, prb, drow0, ns, dsl, B, scal, dout, nr_total, s.st);
    std::vector<double> h((size_t)std::max(ns, 1) * std::max(nr_total, 1));
    HIPCHK(hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipStreamSynchronize(s.st));
    for (int a = 0; a < nr_total; a++) for (int j = 0; j < ns; j++) out[(size_t)a * ns + j] = h[(size_t)j * nr_total + a];
  });
}

int lkc_An_ZtSn_ww_noalloc(const uint8_t* A, int nr_total, const int* alist, int na, const uint8_t* amask,
                           const double* L, const int* nlist, int nn, int n_total, int ns, int B, long long D_MLT,
                           double* out, void* stream)
{
  return guard(nullptr, [&] {
    if (na < 0 || nn < 0 || ns < 0 || nr_total < 0 || n_total <= 0 || nn > n_total || B < 1 || B > 32 || D_MLT < 2 ||
        !A || !amask || (ns && nr_total && !out) || (na && !alist) || (nn && (!nlist || !L)))
      throw lkc::Error(LKC_ERR_STATE, "bad argument");
    const size_t need = lkc::an_ztsn_noalloc_bytes(nr_total, n_total, ns, B);
    if (need > lkc::kNoallocScratchBytes)
      throw lkc::Error(LKC_ERR_STATE, "lkc_An_ZtSn_ww_noalloc: the problem needs " + std::to_string(need) +
                                      " bytes of scratch, more than its fixed " +
                                      std::to_string(lkc::kNoallocScratchBytes) + " (use lkc_An_ZtSn_ww)");
    HIPCHK(lkc::an_ztsn_noalloc(A, nr_total, alist, na, amask, L, nlist, nn, n_total, ns, B, D_MLT, out,
                                (hipStream_t)stream));
  });
}

}  // extern "C"
