// lkc_engine.cpp -- host side of the bcc elimination: configuration, loading,
// device-resident state, stage sequencing, sort and drop selection, outputs.
//
// Stage order and semantics follow lakcluster_driver (lakcluster_driver.c:14-251)
// for one bin, no Y/W controls and the YnWt strategies.  All score arrays live
// on the device; per step only the two score columns that are sorted travel to
// the host, and the new masks and index lists travel back.
#include "lkc_engine.h"
#include "lkc_testgen.h"

#include <algorithm>
#include <mutex>
#include <thread>
#include "lkc_stages.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <string>
#include <vector>

namespace lkc {

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  throw Error(LKC_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ")"); } while (0)

static inline int bget(const uint8_t* b, long i) { return (b[i >> 3] >> (7 - (i & 7))) & 1; }
static inline void bset(uint8_t* b, long i) { b[i >> 3] |= (uint8_t)(1u << (7 - (i & 7))); }
static inline void bclr(uint8_t* b, long i) { b[i >> 3] &= (uint8_t)~(1u << (7 - (i & 7))); }
static int popmask(const std::vector<uint8_t>& m, int n)
{
  int s = 0;
  for (int b = 0; b < n / 8; b++) s += __builtin_popcount(m[b]);
  for (int i = n / 8 * 8; i < n; i++) s += bget(m.data(), i);
  return s;
}
static double now_s() { struct timeval tv; gettimeofday(&tv, NULL); return tv.tv_sec + 1e-6 * tv.tv_usec; }

// ------------------------------------------------------------------
// configuration (updateglobals, updateglobals.c:5-199)
// ------------------------------------------------------------------
static std::string trim(const std::string& s)
{
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// Keys that updateglobals reads (updateglobals.c:5-186) but that the bcc / dcc elimination,
// the scorebox, bcc_scramble and dcc_sumscores_test never consult: thread and print
// settings, the row counts the loader takes from the files, the dcg / PCA /
// PLINK-conversion settings.  Any other key is an error, as the reference's exit() on an
// unknown name (updateglobals.c:185).
static bool reference_key_without_effect(const std::string& n)
{
  static const char* keys[] = {
    "GLOBAL_verbose", "GLOBAL_thread_count", "GLOBAL_omp_type", "GLOBAL_LBITS", "GLOBAL_A_n_rows_", "GLOBAL_Z_n_rows_",
    "GLOBAL_J_n_rind_", "GLOBAL_K_n_cind", "GLOBAL_pca_out_xdrop", "GLOBAL_pca_V_",
    "GLOBAL_pca_V_pre_", "GLOBAL_pca_infix", "GLOBAL_pca_iteration_num", "GLOBAL_pca_iteration_max",
    "GLOBAL_pca_iteration_min", "GLOBAL_pca_rank", "GLOBAL_pca_tolerance", "GLOBAL_fname_bed_0in",
    "GLOBAL_fname_b16_out", "GLOBAL_fname_bim_0in", "GLOBAL_fname_bim_out", "GLOBAL_fname_fam_0in",
    "GLOBAL_fname_fam_out", "GLOBAL_fname_flip_flag", "GLOBAL_snp_mss_threshold", "GLOBAL_snp_maf_threshold",
    "GLOBAL_snp_I_opt_threshold", "GLOBAL_pat_mss_threshold", "GLOBAL_n_fam_char_max", "GLOBAL_n_bim_char_max",
    "GLOBAL_n_file", "GLOBAL_fname_b16_0in_", "GLOBAL_fname_bim_0in_", "GLOBAL_fname_fam_0in_",
    "GLOBAL_fname_b16_0in", "GLOBAL_n_bytes_per_read" };
  for (const char* k : keys) if (n == k) return true;
  return false;
}

void parse_cfg(const char* text, Cfg& c)
{
  c = Cfg();
  std::string s(text ? text : "");
  size_t p = 0;
  while (p < s.size()) {
    size_t eq = s.find('=', p);
    if (eq == std::string::npos) break;
    std::string name = trim(s.substr(p, eq - p));
    size_t semi = s.find(';', eq + 1);
    std::string val = trim(s.substr(eq + 1, (semi == std::string::npos ? s.size() : semi) - eq - 1));
    p = semi == std::string::npos ? s.size() : semi + 1;
    if (name == "END") break;
    std::string first = trim(val.substr(0, val.find(',')));   // list-valued keys: the first bin
    auto path = [&](const std::string& v) { return c.xpre.empty() ? v : c.xpre + "/" + v; };
    auto list = [&](std::vector<std::string>& out) {
      out.clear();
      size_t q = 0;
      while (q <= val.size()) {
        size_t k = val.find(',', q); if (k == std::string::npos) k = val.size();
        const std::string item = trim(val.substr(q, k - q));
        out.push_back(item.empty() ? item : path(item));
        q = k + 1;
      }
      return out.empty() ? std::string() : out[0];
    };
    if (name == "GLOBAL_B_MLT") { c.B = atoi(val.c_str()); c.D_MLT = (long long)pow(2.0, (double)c.B); }
    else if (name == "GLOBAL_D_MLT") { c.D_MLT = atoll(val.c_str()); c.B = (int)llround(log2((double)c.D_MLT)); }
    else if (name == "GLOBAL_gamma") c.gamma = atof(val.c_str());
    else if (name == "GLOBAL_gamma_type") c.gamma_type = atoi(val.c_str());
    else if (name == "GLOBAL_kappa_squared") c.kappa2 = atof(val.c_str());
    else if (name == "GLOBAL_TEST_sparse") c.sparse = atoi(val.c_str());
    else if (name == "GLOBAL_NBINS") c.nbins = atoi(val.c_str());
    else if (name == "GLOBAL_Ireq") c.Ireq = atoi(val.c_str());
    else if (name == "GLOBAL_QR_strategy") c.QR = val;
    else if (name == "GLOBAL_QC_strategy") c.QC = val;
    else if (name == "GLOBAL_A_n_cols") c.A_ncols = atoi(val.c_str());
    else if (name == "GLOBAL_Y_n_cols") c.Y_ncols = atoi(val.c_str());
    else if (name == "GLOBAL_T_n_cols") c.T_ncols = atoi(val.c_str());
    else if (name == "GLOBAL_DIR_XPRE") c.xpre = val;
    else if (name == "GLOBAL_DIR_NAME") c.dir_name = val;
    else if (name == "GLOBAL_DIR_BASE") c.dir_base = val;
    else if (name == "GLOBAL_out_name") c.out_name = val;
    else if (name == "GLOBAL_A_n_name_") c.A_n = list(c.A_n_);
    else if (name == "GLOBAL_A_t_name_") c.A_t = list(c.A_t_);
    else if (name == "GLOBAL_A_n_rind_") c.A_rind = list(c.A_rind_);
    else if (name == "GLOBAL_A_n_cind") c.A_cind = path(val);
    else if (name == "GLOBAL_A_p_name") c.A_p = path(val);
    else if (name == "GLOBAL_Z_n_name_") c.Z_n = list(c.Z_n_);
    else if (name == "GLOBAL_Z_t_name_") c.Z_t = list(c.Z_t_);
    else if (name == "GLOBAL_Z_n_rind_") c.Z_rind = list(c.Z_rind_);
    else if (name == "GLOBAL_T_n_name_") c.T_n = list(c.T_n_);
    else if (name == "GLOBAL_T_t_name_") c.T_t = list(c.T_t_);
    else if (name == "GLOBAL_T_n_cind") c.T_cind = path(val);
    else if (name == "GLOBAL_S_n_name_") c.S_n = list(c.S_n_);
    else if (name == "GLOBAL_S_t_name_") c.S_t = list(c.S_t_);
    else if (name == "GLOBAL_Y_n_name_") list(c.Y_n_);
    else if (name == "GLOBAL_Y_t_name_") list(c.Y_t_);
    else if (name == "GLOBAL_W_n_name_") list(c.W_n_);
    else if (name == "GLOBAL_W_t_name_") list(c.W_t_);
    else if (name == "GLOBAL_Y_n_cind") c.Y_cind = path(val);
    else if (name == "GLOBAL_Y_p_name") c.Y_p = path(val);
    else if (name == "GLOBAL_scramble_num") c.scramble_num = atoi(val.c_str());
    else if (name == "GLOBAL_scramble_out_xdrop_") c.scramble_files_raw = val;
    else if (name == "GLOBAL_scramble_rseed_") c.scramble_seeds_raw = val;
    else if (name == "GLOBAL_TEST_TYPE") c.test_type = first;
    else if (name == "GLOBAL_TEST_TYP2") c.test_typ2 = first;
    else if (name == "GLOBAL_TEST_mrand") c.test_mrand = atof(val.c_str());
    else if (name == "GLOBAL_TEST_A_n_rows") c.test_A_rows = atoi(val.c_str());
    else if (name == "GLOBAL_TEST_A_n_cols") c.test_A_cols = atoi(val.c_str());
    else if (name == "GLOBAL_TEST_Z_n_rows") c.test_Z_rows = atoi(val.c_str());
    else if (name == "GLOBAL_TEST_Y_n_cols") c.test_Y_cols = atoi(val.c_str());
    else if (name == "GLOBAL_TEST_T_n_cols") c.test_T_cols = atoi(val.c_str());
    else if (name == "GLOBAL_TEST_niter") c.test_niter = atoi(val.c_str());
    else if (name == "GLOBAL_scorebox_out_xdrop") c.sb_file = first;
    else if (name == "GLOBAL_scorebox_row_max") c.sb_row_max = atoi(val.c_str());
    else if (name == "GLOBAL_scorebox_row_num") c.sb_row_num = atoi(val.c_str());
    else if (name == "GLOBAL_scorebox_col_max") c.sb_col_max = atoi(val.c_str());
    else if (name == "GLOBAL_scorebox_col_num") c.sb_col_num = atoi(val.c_str());
    else if (reference_key_without_effect(name)) { /* a key of updateglobals that this path does not read */ }
    else throw Error(LKC_ERR_CONFIG, "unknown key " + name);
  }
  // the scramble lists are sized by GLOBAL_scramble_num (updateglobals.c:122-135)
  if (c.scramble_num > 0) {
    auto split = [](const std::string& v) {
      std::vector<std::string> out; size_t q = 0;
      while (q <= v.size()) {
        size_t k = v.find(',', q); if (k == std::string::npos) k = v.size();
        out.push_back(trim(v.substr(q, k - q))); q = k + 1;
      }
      return out;
    };
    c.scramble_files.assign(c.scramble_num, std::string());
    c.scramble_seeds.resize(c.scramble_num);
    for (int i = 0; i < c.scramble_num; i++) c.scramble_seeds[i] = (unsigned long)(i + 1);
    if (!c.scramble_files_raw.empty()) {
      std::vector<std::string> f = split(c.scramble_files_raw);
      for (int i = 0; i < c.scramble_num && i < (int)f.size(); i++) c.scramble_files[i] = f[i];
    }
    if (!c.scramble_seeds_raw.empty()) {
      std::vector<std::string> f = split(c.scramble_seeds_raw);
      for (int i = 0; i < c.scramble_num && i < (int)f.size(); i++) if (!f[i].empty()) c.scramble_seeds[i] = (unsigned long)atol(f[i].c_str());
    }
  }
}

// ------------------------------------------------------------------
// .b16 input (binary_read.c:5-55; M_load, M_init.c:229-242): a file holding an
// (fnr x fnc) matrix column by column yields fnc rows of fnr bits.
// ------------------------------------------------------------------
HostBits read_b16(const std::string& path)
{
  FILE* fp = fopen(path.c_str(), "rb");
  if (!fp) throw Error(LKC_ERR_IO, "cannot open " + path);
  int hdr[3];
  if (fread(hdr, sizeof(int), 3, fp) != 3) { fclose(fp); throw Error(LKC_ERR_IO, "short header in " + path); }
  int bitj = hdr[0] > 0 ? hdr[0] : 16, fnr = hdr[1], fnc = hdr[2];
  if (fnr < 0 || fnc < 0) { fclose(fp); throw Error(LKC_ERR_IO, "bad header in " + path); }
  size_t wp = (size_t)(fnr + (bitj - (fnr % bitj)) % bitj) / 8;
  HostBits h; h.nr = fnc; h.nc = fnr; h.rb = bsize(fnr);
  h.b.assign((size_t)std::max(fnc, 1) * h.rb, 0);
  size_t take = std::min(wp, h.rb);
  std::vector<uint8_t> col(wp);
  for (int r = 0; r < fnc; r++) {
    if (wp && fread(col.data(), 1, wp, fp) != wp) { fclose(fp); throw Error(LKC_ERR_IO, "short data in " + path); }
    memcpy(h.b.data() + (size_t)r * h.rb, col.data(), take);
  }
  fclose(fp);
  return h;
}

void b16_size(const std::string& path, int* fnr, int* fnc)
{
  FILE* fp = fopen(path.c_str(), "rb");
  if (!fp) throw Error(LKC_ERR_IO, "cannot open " + path);
  int hdr[3];
  const bool ok = fread(hdr, sizeof(int), 3, fp) == 3;
  fclose(fp);
  if (!ok || hdr[1] < 0 || hdr[2] < 0) throw Error(LKC_ERR_IO, "bad header in " + path);
  *fnr = hdr[1]; *fnc = hdr[2];
}

std::vector<uint8_t> read_mask(const std::string& path, int n)
{
  HostBits m = read_b16(path);
  std::vector<uint8_t> out(bsize(n) + 8, 0);
  if (m.nr > 0) memcpy(out.data(), m.b.data(), std::min(bsize(n), m.rb));
  for (long i = n; i < (long)bsize(n) * 8; i++) bclr(out.data(), i);
  return out;
}

// ------------------------------------------------------------------
// get_xdrop (xdrop.c:5-64) and dQuickSort_xij (Quicksort.c:83-141)
// ------------------------------------------------------------------
void get_xdrop(double lrij, double lcij, double gamma, int gamma_type, int* rdrop_p, int* cdrop_p)
{
  if (lrij > lcij) { get_xdrop(lcij, lrij, gamma, gamma_type, cdrop_p, rdrop_p); return; }
  const double dbl1 = 1.000000 - 0.000001;
  double gtr = 0, atr = 0, gtc = 0, atc = 0; int rdrop = 0, cdrop = 0;
  if (lrij <= 2) { rdrop = (int)lrij; cdrop = (int)lcij; }
  else {
    double one_r = std::max(1.0, lrij);
    if (gamma > 0) { gtr = std::max(gamma, dbl1 / one_r); atr = std::min(1 - gamma, (lrij - 1) / one_r); }
    else { gtr = dbl1 / one_r; atr = (lrij - 1) / one_r; }
    if (gamma_type == 0) atc = exp(log(atr) * log(lcij) / std::max(1.0, log(one_r)));
    else if (gamma_type == 1) atc = sqrt(atr * lrij) * sqrt(lcij) / std::max(1.0, sqrt(one_r)) / ((double)std::max(1.0, lcij));
    else if (gamma_type == 2) atc = atr;
    else if (gamma_type == 3) atc = 0;
    gtc = dbl1 - atc; rdrop = (int)ceil(gtr * lrij); cdrop = (int)ceil(gtc * lcij);
    if (gamma_type == 3) cdrop = 0;
  }
  rdrop = (int)std::min((double)rdrop, lrij); cdrop = (int)std::min((double)cdrop, lcij);
  if (rdrop_p) *rdrop_p = rdrop;
  if (cdrop_p) *cdrop_p = cdrop;
}

// get_xdrop_length (xdrop.c:66-83): passes until the rows or the columns run out
int get_xdrop_length(double lrij, double lcij, double gamma, int gamma_type)
{
  int length = 0;
  for (bool more = true; more;) {
    length++;
    int rdrop = 0, cdrop = 0;
    get_xdrop(lrij, lcij, gamma, gamma_type, &rdrop, &cdrop);
    lrij -= rdrop; lcij -= cdrop;
    more = lrij > 0 && lcij > 0;
  }
  return length;
}

// Hoare partition around the first element; equal keys stay left of the pivot.
static int partition_first(double* d, int* perm, int l, int r)
{
  double pivot = d[l];
  int i = l, j = r + 1;
  for (;;) {
    do { i++; } while (i <= r && d[i] <= pivot);
    do { j--; } while (d[j] > pivot);
    if (i >= j) break;
    std::swap(d[i], d[j]); std::swap(perm[i], perm[j]);
  }
  std::swap(d[l], d[j]); std::swap(perm[l], perm[j]);
  return j;
}
// dQuickSort_xij (Quicksort.c:124-141): returns the deepest level reached (the recursion
// stops below GLOBAL_recursion_limit = 32768, leaving such a part as partitioned so far)
static unsigned quicksort_first(double* d, int* perm, int l, int r, unsigned depth)
{
  unsigned n1 = depth, n2 = depth;
  if (l < r) {
    const int j = partition_first(d, perm, l, r);
    if (depth < 32768u) { n1 = quicksort_first(d, perm, l, j - 1, depth + 1); n2 = quicksort_first(d, perm, j + 1, r, depth + 1); }
  }
  return std::max(n1, n2);
}
// The same recursion with its two halves sorted on two host threads near the top: the
// halves are disjoint and each keeps its own depth count, so the result is the sequential
// one's exactly (a list of tied scores longer than 32769 entries is common in the dcc
// elimination, whose column scores are functions of a few per-column counts).
static unsigned quicksort_par(double* d, int* perm, int l, int r, unsigned depth, int spawn)
{
  if (spawn <= 0 || r - l < 16384) return quicksort_first(d, perm, l, r, depth);
  unsigned n1 = depth, n2 = depth;
  const int j = partition_first(d, perm, l, r);
  if (depth < 32768u) {
    std::thread left([&n1, d, perm, l, j, depth, spawn] { n1 = quicksort_par(d, perm, l, j - 1, depth + 1, spawn - 1); });
    n2 = quicksort_par(d, perm, j + 1, r, depth + 1, spawn - 1);
    left.join();
  }
  return std::max(n1, n2);
}
unsigned sort_scores(double* d, int* perm, int n) { return n > 0 ? quicksort_par(d, perm, 0, n - 1, 0, 2) : 0; }
unsigned quicksort_xij(double* d, int* perm, int l, int r) { return quicksort_par(d, perm, l, r, 0, 2); }

// ------------------------------------------------------------------
// device memory
// ------------------------------------------------------------------
Dev::~Dev() { release_all(); }
void Dev::release_all()
{
  for (void* p : ptrs) (void)(st ? hipFreeAsync(p, st) : hipFree(p));
  ptrs.clear();
}
void* Dev::raw(size_t bytes)
{
  void* p = nullptr;
  bytes = std::max<size_t>(bytes, 256);
  if (st) HIPCHK(hipMallocAsync(&p, bytes, st));
  else HIPCHK(hipMalloc(&p, bytes));
  ptrs.push_back(p);
  return p;
}
void Dev::release(void* p)
{
  if (st) HIPCHK(hipFreeAsync(p, st));
  else { HIPCHK(hipDeviceSynchronize()); HIPCHK(hipFree(p)); }
  ptrs.erase(std::find(ptrs.begin(), ptrs.end(), p));
}
void* Grow::need(Dev& d, size_t bytes)
{
  if (bytes <= cap && p) return p;
  if (p) d.release(p);
  cap = std::max(bytes, cap + cap / 2);
  p = d.raw(cap);
  return p;
}

// sizes of the pinned host buffers handed out (freed ones are kept with their size)
namespace {
std::mutex g_pin_sz_mu;
std::vector<std::pair<void*, size_t>> g_pin_sz;
}
static void pinned_size_note(void* p, size_t bytes)
{
  std::lock_guard<std::mutex> lk(g_pin_sz_mu);
  g_pin_sz.push_back({ p, bytes });
}
static size_t pinned_size_of(void* p)
{
  std::lock_guard<std::mutex> lk(g_pin_sz_mu);
  for (const auto& e : g_pin_sz) if (e.first == p) return e.second;
  return 0;
}

// Streams and pinned host buffers outlive the runs that used them: a finished run hands
// them back and the next one takes them, so jobs that start and end while others run
// create and free neither (freeing pinned memory waits for the whole device).
namespace {
std::mutex g_res_mu;
std::vector<hipStream_t> g_free_streams;
std::vector<std::pair<size_t, void*>> g_free_pinned;
}
static hipStream_t stream_take()
{
  {
    std::lock_guard<std::mutex> lk(g_res_mu);
    if (!g_free_streams.empty()) { hipStream_t s = g_free_streams.back(); g_free_streams.pop_back(); return s; }
  }
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}
static void stream_give(hipStream_t s)
{
  if (!s) return;
  std::lock_guard<std::mutex> lk(g_res_mu);
  g_free_streams.push_back(s);
}
// the smallest free buffer of at least bytes (and at most four times that), else a new one
static void* pinned_take(size_t bytes)
{
  {
    std::lock_guard<std::mutex> lk(g_res_mu);
    size_t best = (size_t)-1;
    for (size_t i = 0; i < g_free_pinned.size(); i++) {
      const size_t b = g_free_pinned[i].first;
      if (b >= bytes && b / 4 <= bytes && (best == (size_t)-1 || b < g_free_pinned[best].first)) best = i;
    }
    if (best != (size_t)-1) {
      void* p = g_free_pinned[best].second;
      g_free_pinned.erase(g_free_pinned.begin() + best);
      return p;
    }
  }
  void* p = nullptr;
  HIPCHK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
  pinned_size_note(p, bytes);
  return p;
}
static void pinned_give(void* p)
{
  if (!p) return;
  const size_t bytes = pinned_size_of(p);
  std::lock_guard<std::mutex> lk(g_res_mu);
  g_free_pinned.push_back({ bytes, p });
}

// the run's stream; device memory of the run comes from the pool, kept for reuse
static void make_stream(hipStream_t* st, Dev& dev)
{
  HIPCHK(hipSetDevice(0));
  *st = stream_take();
  static const bool pool_set = [] {   // once per process, thread-safe
    hipMemPool_t mp;
    if (hipDeviceGetDefaultMemPool(&mp, 0) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep);
    }
    return true;
  }();
  (void)pool_set;
  dev.st = *st;
}

static void upload_bits(Dev& dev, DMat& m, const HostBits& h)
{
  m.nr = h.nr; m.nc = h.nc; m.rb = h.rb;
  size_t bytes = (size_t)std::max(h.nr, 1) * h.rb;
  m.p = dev.alloc<uint8_t>(bytes);
  HIPCHK(hipMemcpyAsync(m.p, h.b.data(), bytes, hipMemcpyHostToDevice, dev.st));
  HIPCHK(hipStreamSynchronize(dev.st));
}

// ------------------------------------------------------------------
// kernel-group timing
// ------------------------------------------------------------------
Prof::~Prof()
{
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  for (Pending& p : pend) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
}
int Prof::id(const char* name)
{
  for (size_t i = 0; i < stats.size(); i++) if (stats[i].name == name) return (int)i;
  stats.push_back(Stat()); stats.back().name = name;
  return (int)stats.size() - 1;
}
hipEvent_t Prof::ev()
{
  if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
  hipEvent_t e; HIPCHK(hipEventCreate(&e)); return e;
}
void Prof::begin(const char* name, double work, hipStream_t st, Pending& p)
{
  p.id = -1;
  if (!on) return;
  p.id = id(name); p.a = ev(); p.b = ev();
  stats[p.id].launches++; stats[p.id].work += work;
  HIPCHK(hipEventRecord(p.a, st));
}
void Prof::end(Pending& p, hipStream_t st)
{
  if (p.id < 0) return;
  HIPCHK(hipEventRecord(p.b, st));
  pend.push_back(p);
}
void Prof::collect()
{
  for (Pending& p : pend) {
    float ms = 0; HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
    stats[p.id].ms += ms;
    pool.push_back(p.a); pool.push_back(p.b);
  }
  pend.clear();
}

// ------------------------------------------------------------------
// Run: construction (bcc_load, bcc.c:1482-1540)
// ------------------------------------------------------------------
Run::~Run()
{
  dev.release_all();
  if (st) (void)hipStreamSynchronize(st);
  pinned_give(hctl);
  pinned_give(hsv);
  pinned_give(hidx);
  stream_give(st);
}

void Run::load(const char* in_text)
{
  double t0 = now_s();
  parse_cfg(in_text, c);
  if (c.test_type == "bcc_sumscores") { load_synthetic(t0); return; }
  if (c.nbins != 1) throw Error(LKC_ERR_CONFIG, "only GLOBAL_NBINS= 1 is supported");
  bool y_files = false;
  for (const std::string& f : c.Y_t_) y_files = y_files || !f.empty();
  if (c.Y_ncols != 0 || y_files) throw Error(LKC_ERR_CONFIG, "Y/W genetic controls are not supported here (the file-driven dexcluster_driver takes them)");
  // GLOBAL_TEST_TYPE selects the elimination (lakcluster_ver18.c:337-345); the dcc one has
  // no low-rank update, so no QR/QC strategy
  dcc = (c.test_type == "dexcluster_driver" || c.test_type == "dexcluster_scorebox") ? 1 : 0;
  if (!dcc && (c.QR.find("YnWt") == std::string::npos || c.QC.find("YnWt") == std::string::npos ||
               c.QR.find("condense") != std::string::npos || c.QC.find("store one") != std::string::npos ||
               c.QR.find("ZtSWn") != std::string::npos || c.QC.find("ZtSWn") != std::string::npos))
    throw Error(LKC_ERR_CONFIG, "only the YnWt strategies are supported");
  // GLOBAL_Ireq > 0 sorts each row's scores over the bins and normalises per bin
  // (bcc_sumscores_srt / _nrm, bcc_sumscores.c:358-481, 583-628; dcc_sumscores.c:283-407,
  // 490-541): with one bin that sorts one value and gives the same normalisers, so the run is
  // the GLOBAL_Ireq <= 0 one bit for bit; with several bins it is not supported
  if (c.Ireq > 0 && c.nbins > 1) throw Error(LKC_ERR_CONFIG, "GLOBAL_Ireq > 0 (per-bin sorting) is not supported");
  for (int i = 0; i < c.scramble_num; i++)
    if (c.scramble_files[i].empty()) throw Error(LKC_ERR_CONFIG, "GLOBAL_scramble_out_xdrop_ names fewer files than GLOBAL_scramble_num");
  if (c.B < 1 || c.B > 32) throw Error(LKC_ERR_CONFIG, "GLOBAL_B_MLT must be in [1,32]");
  if (c.A_t.empty() && c.A_n.empty()) throw Error(LKC_ERR_CONFIG, "an A_n or A_t file is required");
  make_stream(&st, dev);

  // Either copy of a matrix may be omitted: the other one is transposed on the device
  // (the file pair of the reference, wrap_transpose in binary_read.c:174-269).
  HostBits hAn, hAt;
  if (!c.A_t.empty()) hAn = read_b16(c.A_t);
  if (!c.A_n.empty()) hAt = read_b16(c.A_n);
  if (c.A_t.empty()) { hAn.nr = hAt.nc; hAn.nc = hAt.nr; hAn.rb = bsize(hAn.nc); }
  if (c.A_n.empty()) { hAt.nr = hAn.nc; hAt.nc = hAn.nr; hAt.rb = bsize(hAt.nc); }
  NA = hAn.nr; M = hAn.nc;
  if (c.A_ncols && c.A_ncols != M) throw Error(LKC_ERR_SHAPE, "GLOBAL_A_n_cols does not match A");
  if (hAt.nr != M || hAt.nc != NA) throw Error(LKC_ERR_SHAPE, "A_n/A_t shapes disagree");
  if (NA <= 0 || M <= 0) throw Error(LKC_ERR_SHAPE, "A is empty");
  has_Z = !c.Z_t.empty() || !c.Z_n.empty();
  HostBits hZn, hZt, hTt, hSt;
  if (has_Z) {
    if (!c.Z_t.empty()) hZn = read_b16(c.Z_t);
    if (!c.Z_n.empty()) hZt = read_b16(c.Z_n);
    if (c.Z_t.empty()) { hZn.nr = hZt.nc; hZn.nc = hZt.nr; hZn.rb = bsize(hZn.nc); }
    if (c.Z_n.empty()) { hZt.nr = hZn.nc; hZt.nc = hZn.nr; hZt.rb = bsize(hZt.nc); }
    NZ = hZn.nr;
    if (hZn.nc != M || hZt.nr != M || hZt.nc != NZ) throw Error(LKC_ERR_SHAPE, "Z shapes disagree");
  } else {
    NZ = 0; hZn.nr = 0; hZn.nc = M; hZn.rb = bsize(M); hZn.b.assign(hZn.rb, 0);
    hZt.nr = M; hZt.nc = 0; hZt.rb = bsize(0); hZt.b.assign((size_t)M * hZt.rb, 0);
  }
  if (!c.T_n.empty()) {
    hTt = read_b16(c.T_n);
    T = hTt.nr;
    if (hTt.nc != NA) throw Error(LKC_ERR_SHAPE, "T shape disagrees");
    if (has_Z) {
      if (c.S_n.empty()) throw Error(LKC_ERR_CONFIG, "S_n file required with T and Z");
      hSt = read_b16(c.S_n);
      if (hSt.nr != T || hSt.nc != NZ) throw Error(LKC_ERR_SHAPE, "S shape disagrees");
    } else { hSt.nr = T; hSt.nc = 0; hSt.rb = bsize(0); hSt.b.assign((size_t)std::max(T, 1) * hSt.rb, 0); }
  } else {
    // no covariate file: one all-ones intercept column (bcc.c:361-373)
    T = 1;
    hTt.nr = 1; hTt.nc = NA; hTt.rb = bsize(NA); hTt.b.assign(hTt.rb, 0);
    for (int a = 0; a < NA; a++) bset(hTt.b.data(), a);
    hSt.nr = 1; hSt.nc = NZ; hSt.rb = bsize(NZ); hSt.b.assign(hSt.rb, 0);
    for (int z = 0; z < NZ; z++) bset(hSt.b.data(), z);
  }
  if (T < 1 || T > kMaxT) throw Error(LKC_ERR_CONFIG, "between 1 and 8 covariate columns are supported");
  pcols = psize(M);

  masks_from_files();
  finish_masks();

  // device-resident bit matrices
  auto place = [&](DMat& m, const HostBits& h, bool from_file) {
    if (from_file) { upload_bits(dev, m, h); return; }
    m.nr = h.nr; m.nc = h.nc; m.rb = h.rb;
    size_t bytes = (size_t)std::max(h.nr, 1) * h.rb;
    m.p = dev.alloc<uint8_t>(bytes);
    HIPCHK(hipMemsetAsync(m.p, 0, bytes, st));
  };
  place(An, hAn, !c.A_t.empty()); place(At, hAt, !c.A_n.empty());
  if (c.A_t.empty()) transpose_bits(At.b(), An.p, An.rb, st);
  if (c.A_n.empty()) transpose_bits(An.b(), At.p, At.rb, st);
  const bool zn_file = !has_Z || !c.Z_t.empty(), zt_file = !has_Z || !c.Z_n.empty();
  place(Zn, hZn, zn_file); place(Zt, hZt, zt_file);
  if (!zn_file) transpose_bits(Zt.b(), Zn.p, Zn.rb, st);
  if (!zt_file) transpose_bits(Zn.b(), Zt.p, Zt.rb, st);
  upload_bits(dev, Tt, hTt); upload_bits(dev, St, hSt);
  { HostBits().b.swap(hAn.b); }

  setup_device(t0);
}

// bcc_init (bcc.c:1542-1600) for one bin without Y/W controls, GLOBAL_TEST_TYP2= error, as
// bcc_sumscores_test (bcc_sumscores.c:781-880) calls it: bcc_ajdk_init's column masks
// (bcc.c:117-174; the covariate mask with its first and last column forced on),
// bcc_single_init_M_An's row masks and matrices (bcc.c:415-520; the draws of
// wrap_M_setup_test_excerpt_1 / _0, An_ajdk_v.c:32-98: column by column over every row,
// bit = draw > 1/2; the covariates' column 0 set in every row), then bcc_init_A_p's block
// frequencies (bcc.c:1315-1330).  The draws replay glibc rand() from its initial state.  A
// mask bit is on in _b when its draw exceeds mrand^2 and in _j when it exceeds mrand: rows
// and columns may start loaded but inactive.  Every covariate column must be drawn active.
void Run::load_synthetic(double t0)
{
  if (c.nbins != 1) throw Error(LKC_ERR_CONFIG, "bcc_sumscores: only GLOBAL_NBINS= 1 is supported");
  if (c.test_Y_cols != 0) throw Error(LKC_ERR_CONFIG, "bcc_sumscores: Y/W genetic controls are not supported (GLOBAL_TEST_Y_n_cols must be 0)");
  if (c.test_typ2 != "error") throw Error(LKC_ERR_CONFIG, "bcc_sumscores: GLOBAL_TEST_TYP2 must be error (random data)");
  if (c.QR.find("YnWt") == std::string::npos || c.QC.find("YnWt") == std::string::npos ||
      c.QR.find("condense") != std::string::npos || c.QC.find("store one") != std::string::npos ||
      c.QR.find("ZtSWn") != std::string::npos || c.QC.find("ZtSWn") != std::string::npos)
    throw Error(LKC_ERR_CONFIG, "only the YnWt strategies are supported");
  if (c.Ireq > 0 && c.nbins > 1) throw Error(LKC_ERR_CONFIG, "GLOBAL_Ireq > 0 (per-bin sorting) is not supported");
  if (c.scramble_num > 0) throw Error(LKC_ERR_CONFIG, "bcc_scramble is not part of bcc_sumscores_test");
  if (c.B < 1 || c.B > 32) throw Error(LKC_ERR_CONFIG, "GLOBAL_B_MLT must be in [1,32]");
  if (c.test_A_cols < 1 || c.test_T_cols < 1 || c.test_T_cols > kMaxT) throw Error(LKC_ERR_CONFIG, "bcc_sumscores: bad problem size");
  synthetic = true; dcc = 0;
  GlibcRand g;
  const double mrnd = c.test_mrand;
  M = c.test_A_cols; T = c.test_T_cols;
  NA = std::max(0, c.test_A_rows + (0 % 3) - 1); NZ = std::max(0, c.test_Z_rows + (0 % 5) - 2);
  has_Z = NZ > 0;
  auto mk = [](int n) { return std::vector<uint8_t>(bsize(n) + 8, 0); };
  auto mask_draw = [&](int n, std::vector<uint8_t>& mb, std::vector<uint8_t>& mj) {
    mb = mk(n); mj = mk(n);
    for (int i = 0; i < n; i++) { const double d = g.r01(); if (d > mrnd * mrnd) bset(mb.data(), i); if (d > mrnd) bset(mj.data(), i); }
  };
  auto bits_draw = [&](int nrows, int ncols, HostBits& Xn, HostBits& Xt) {
    Xn.nr = nrows; Xn.nc = ncols; Xn.rb = bsize(ncols); Xn.b.assign((size_t)std::max(nrows, 1) * Xn.rb, 0);
    Xt.nr = ncols; Xt.nc = nrows; Xt.rb = bsize(nrows); Xt.b.assign((size_t)std::max(ncols, 1) * Xt.rb, 0);
    for (int k = 0; k < ncols; k++) for (int r = 0; r < nrows; r++)
      if (g.r01() > 0.5) { bset(Xn.b.data() + (size_t)r * Xn.rb, k); bset(Xt.b.data() + (size_t)k * Xt.rb, r); }
  };
  std::vector<uint8_t> tcb, tcj;
  mask_draw(M, hAc_b, hAc_j);
  mask_draw(T, tcb, tcj);
  bset(tcb.data(), 0); bset(tcj.data(), 0); bset(tcb.data(), T - 1); bset(tcj.data(), T - 1);
  HostBits hAn, hAt, hZn, hZt, hTn, hTt, hSn, hSt;
  mask_draw(NA, hAr_b, hAr_j);
  bits_draw(NA, M, hAn, hAt);
  mask_draw(NZ, hZr_b, hZr_j);
  bits_draw(NZ, M, hZn, hZt);
  bits_draw(NA, T, hTn, hTt);   // Y and W have no columns: no draws
  for (int r = 0; r < NA; r++) bset(hTt.b.data(), r);
  bits_draw(NZ, T, hSn, hSt);
  for (int r = 0; r < NZ; r++) bset(hSt.b.data(), r);
  pcols = psize(M);
  hA_p.assign(pcols, 0.5);
  for (int b = 0; b < pcols; b++) hA_p[b] = crop_draw(g);
  if (c.sparse == 0) std::fill(hA_p.begin(), hA_p.end(), 0.5);
  if (popmask(tcj, T) != T) throw Error(LKC_ERR_CONFIG, "bcc_sumscores: the drawn covariate mask leaves a column out (not supported)");
  hTc_j = tcj;
  hAr_rtn = mk(NA); hAr_rmv = mk(NA); hAc_rtn = mk(M); hAc_rmv = mk(M);
  make_stream(&st, dev);
  finish_masks();
  upload_bits(dev, An, hAn); upload_bits(dev, At, hAt);
  upload_bits(dev, Zn, hZn); upload_bits(dev, Zt, hZt);
  upload_bits(dev, Tt, hTt); upload_bits(dev, St, hSt);
  setup_device(t0);
}

// A job drawn from a device-resident pool (permutation nulls, serving): the listed
// rows become A and Z, compacted in the order given; all of them are active; colmask
// (bsize(M) bytes, NULL = all) selects the columns, as the mr_/mc_ masks of a .in do.
// The matrices never leave the device: rows are gathered, the transposed copies built
// with the bit transpose, the covariate columns gathered bit by bit.
void Run::load_job(const Pool& P, const int* rowsA, int nA, const int* rowsZ, int nZ, const uint8_t* colmask,
                   const Cfg& params)
{
  double t0 = now_s();
  c = params; c.nbins = 1; c.QR = c.QC = "YnWt";
  if (c.B < 1 || c.B > 32) throw Error(LKC_ERR_CONFIG, "B_MLT must be in [1,32]");
  if (nA <= 0 || nZ < 0 || P.M <= 0) throw Error(LKC_ERR_SHAPE, "a job needs at least one A row");
  for (int i = 0; i < nA; i++) if (rowsA[i] < 0 || rowsA[i] >= P.P) throw Error(LKC_ERR_SHAPE, "A row outside the pool");
  for (int i = 0; i < nZ; i++) if (rowsZ[i] < 0 || rowsZ[i] >= P.P) throw Error(LKC_ERR_SHAPE, "Z row outside the pool");
  make_stream(&st, dev);
  NA = nA; NZ = nZ; M = P.M; T = P.T; has_Z = nZ > 0; pcols = psize(M);
  if (T < 1 || T > kMaxT) throw Error(LKC_ERR_CONFIG, "between 1 and 8 covariate columns are supported");
  auto mk = [](int n) { return std::vector<uint8_t>(bsize(n) + 8, 0); };
  hAr_b = mk(NA); hAr_j = mk(NA); hAr_rtn = mk(NA); hAr_rmv = mk(NA);
  hZr_b = mk(NZ); hZr_j = mk(NZ);
  hAc_b = mk(M); hAc_j = mk(M); hAc_rtn = mk(M); hAc_rmv = mk(M);
  hTc_j = mk(T);
  for (int i = 0; i < NA; i++) bset(hAr_b.data(), i);
  for (int i = 0; i < NZ; i++) bset(hZr_b.data(), i);
  if (colmask) { memcpy(hAc_b.data(), colmask, bsize(M)); for (long i = M; i < (long)bsize(M) * 8; i++) bclr(hAc_b.data(), i); }
  else for (int i = 0; i < M; i++) bset(hAc_b.data(), i);
  for (int i = 0; i < T; i++) bset(hTc_j.data(), i);
  finish_masks();

  int* drows = dev.alloc<int>((size_t)nA + nZ + 1);
  HIPCHK(hipMemcpyAsync(drows, rowsA, sizeof(int) * nA, hipMemcpyHostToDevice, st));
  if (nZ) HIPCHK(hipMemcpyAsync(drows + nA, rowsZ, sizeof(int) * nZ, hipMemcpyHostToDevice, st));
  auto alloc0 = [&](DMat& m, int nr, int nc) {
    m.nr = nr; m.nc = nc; m.rb = bsize(nc);
    size_t bytes = (size_t)std::max(nr, 1) * m.rb;
    m.p = dev.alloc<uint8_t>(bytes);
    HIPCHK(hipMemsetAsync(m.p, 0, bytes, st));
  };
  alloc0(An, NA, M); gather_rows(P.Pn.b(), drows, NA, An.p, st);
  alloc0(At, M, NA); transpose_bits(An.b(), At.p, At.rb, st);
  alloc0(Zn, NZ, M); alloc0(Zt, M, NZ);
  if (has_Z) { gather_rows(P.Pn.b(), drows + nA, NZ, Zn.p, st); transpose_bits(Zn.b(), Zt.p, Zt.rb, st); }
  alloc0(Tt, T, NA); gather_bits(P.Tp.b(), drows, NA, Tt.p, Tt.rb, st);
  alloc0(St, T, NZ); if (has_Z) gather_bits(P.Tp.b(), drows + nA, NZ, St.p, St.rb, st);
  setup_device(t0);
}

// row / column / covariate masks from the .in files (mr_*, mc_*, T column mask)
void Run::masks_from_files()
{
  auto mk = [](int n) { return std::vector<uint8_t>(bsize(n) + 8, 0); };
  hAr_b = mk(NA); hAr_j = mk(NA); hAr_rtn = mk(NA); hAr_rmv = mk(NA);
  hZr_b = mk(NZ); hZr_j = mk(NZ);
  hAc_b = mk(M); hAc_j = mk(M); hAc_rtn = mk(M); hAc_rmv = mk(M);
  hTc_j = mk(T);
  if (!c.A_rind.empty()) hAr_b = read_mask(c.A_rind, NA); else for (int i = 0; i < NA; i++) bset(hAr_b.data(), i);
  if (!c.Z_rind.empty() && has_Z) hZr_b = read_mask(c.Z_rind, NZ); else for (int i = 0; i < NZ; i++) bset(hZr_b.data(), i);
  if (!c.A_cind.empty()) hAc_b = read_mask(c.A_cind, M); else for (int i = 0; i < M; i++) bset(hAc_b.data(), i);
  if (!c.T_cind.empty()) hTc_j = read_mask(c.T_cind, T);
  else { int nt = c.T_ncols > 0 ? std::min(c.T_ncols, T) : T; for (int i = 0; i < nt; i++) bset(hTc_j.data(), i); }
}

// working masks, active counts and storage index maps (bcc_M_mxset, bcc.c:1414-1438)
void Run::finish_masks()
{
  for (int i = 1; i < T; i++) if (bget(hTc_j.data(), i) && !bget(hTc_j.data(), i - 1)) throw Error(LKC_ERR_CONFIG, "T column mask must be a prefix");
  if (!synthetic) { hAr_j = hAr_b; hZr_j = hZr_b; hAc_j = hAc_b; }   // a loaded problem starts with all of it active
  hAr_j0 = hAr_j; hZr_j0 = hZr_j; hAc_j0 = hAc_j;
  Arb = popmask(hAr_b, NA); Zrb = popmask(hZr_b, NZ);
  if (Arb == 0) throw Error(LKC_ERR_SHAPE, "A has no active rows");
  Ar_bidx.assign(NA + 1, 0); Ac_bidx.assign(M + 1, 0);
  for (int i = 0, b = 0; i < NA; i++) { Ar_bidx[i] = b; if (bget(hAr_b.data(), i)) b++; }
  for (int i = 0, b = 0; i < M; i++) { Ac_bidx[i] = b; if (bget(hAc_b.data(), i)) b++; }
}

// everything after the bit matrices are resident: control block, A_p / ajdk tables,
// persistent score arrays (bcc_init, bcc.c:1005-1150, 1197-1313)
void Run::setup_device(double t0)
{
  // control block: masks and index lists, uploaded together once per step
  size_t off = 0;
  auto reserve = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
  size_t o_m[11]; int mn[11] = { NA, NA, NA, NA, NZ, NZ, M, M, M, M, T };
  for (int i = 0; i < 11; i++) o_m[i] = reserve(bsize(mn[i]) + 8);
  size_t o_zero = reserve(sizeof(int) * 4);
  ctl_head = off;                             // masks and the constant zero: fixed offsets
  IList* L[8] = { &ar, &kl, &tl, &zr, &rtn, &rmv, &crt, &crm };
  int lcap[8] = { NA, M, T, NZ, NA, NA, M, M };
  size_t o_l[8];
  for (int i = 0; i < 8; i++) o_l[i] = reserve(sizeof(int) * (size_t)(lcap[i] + 1));
  cmp_Wmax = 4 * pcols + (M + 31) / 32 + 4;   // whole-word segments, rows padded to 4-word groups
  ctl_cmp_off[0] = reserve(sizeof(int) * 3 * (size_t)(pcols + 1));
  ctl_cmp_off[1] = reserve(sizeof(int) * 32 * (size_t)cmp_Wmax);
  ctl_cmp_off[2] = reserve(sizeof(uint32_t) * (size_t)cmp_Wmax);
  rmv_Wmax = (((NA + 31) / 32) + 3) / 4 * 4;
  ctl_cmp_off[3] = reserve(sizeof(int) * 32 * (size_t)rmv_Wmax);
  ctl_cmp_off[4] = reserve(sizeof(uint32_t) * (size_t)rmv_Wmax);
  ctl_cmp_off[5] = reserve(sizeof(int) * 32 * (size_t)rmv_Wmax);
  ctl_cmp_off[6] = reserve(sizeof(uint32_t) * (size_t)rmv_Wmax);
  ctl_cmp_off[7] = reserve(sizeof(int) * 3 * (size_t)(pcols + 1));
  ctl_cmp_off[8] = reserve(sizeof(int) * 32 * (size_t)cmp_Wmax);
  ctl_bytes = off;
  hctl = (uint8_t*)pinned_take(ctl_bytes);
  memset(hctl, 0, ctl_bytes);
  dctl = dev.alloc<uint8_t>(ctl_bytes);
  uint32_t** dm[11] = { &Ar_b, &Ar_j, &Ar_rtn, &Ar_rmv, &Zr_b, &Zr_j, &Ac_b, &Ac_j, &Ac_rtn, &Ac_rmv, &Tc_j };
  for (int i = 0; i < 11; i++) { *dm[i] = (uint32_t*)(dctl + o_m[i]); ctl_mask_off[i] = o_m[i]; }
  for (int i = 0; i < 8; i++) { L[i]->d = (int*)(dctl + o_l[i]); ctl_list_off[i] = o_l[i]; }
  d_zero = (int*)(dctl + o_zero);
  d_seg = (int*)(dctl + ctl_cmp_off[0]); d_src = (int*)(dctl + ctl_cmp_off[1]); d_cmask = (uint32_t*)(dctl + ctl_cmp_off[2]);
  d_src_rmv = (int*)(dctl + ctl_cmp_off[3]); d_cmask_rmv = (uint32_t*)(dctl + ctl_cmp_off[4]);
  d_src_rtn = (int*)(dctl + ctl_cmp_off[5]); d_cmask_rtn = (uint32_t*)(dctl + ctl_cmp_off[6]);
  d_seg_k = (int*)(dctl + ctl_cmp_off[7]); d_src_k = (int*)(dctl + ctl_cmp_off[8]);
  for (int i = 0; i < 6; i++) gcnt[i] = dev.alloc<int>((size_t)std::max(NA, NZ) * (pcols + 1));
  At_rmvc = dev.alloc<uint32_t>((size_t)M * rmv_Wmax);
  Tt_rmvc = dev.alloc<uint32_t>((size_t)T * rmv_Wmax);
  At_rtnc = dev.alloc<uint32_t>((size_t)M * rmv_Wmax);
  Tt_rtnc = dev.alloc<uint32_t>((size_t)T * rmv_Wmax);
  {
    std::vector<int> io((size_t)std::max(NA, NZ) + 1);
    for (size_t i = 0; i < io.size(); i++) io[i] = (int)i;
    d_iota = dev.alloc<int>(io.size());
    HIPCHK(hipMemcpyAsync(d_iota, io.data(), sizeof(int) * io.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  An_c = dev.alloc<uint32_t>((size_t)NA * cmp_Wmax);
  Zn_c = dev.alloc<uint32_t>((size_t)std::max(NZ, 1) * cmp_Wmax);

  // A_p per 1920-column block (bcc_load_A_p, bcc.c:1197-1313); bcc_init's problem draws its own
  if (!synthetic) hA_p.assign(pcols, 0.5);
  push_ctl();
  if (!synthetic) {
    int* cnt = dev.alloc<int>((size_t)M * 2);
    popc_rows(At.b(), nullptr, M, Ar_b, (int)(At.rb / 4), cnt, st);
    if (has_Z) popc_rows(Zt.b(), nullptr, M, Zr_b, (int)(Zt.rb / 4), cnt + M, st);
    std::vector<int> hc((size_t)M * 2, 0);
    HIPCHK(hipMemcpyAsync(hc.data(), cnt, sizeof(int) * (size_t)M * (has_Z ? 2 : 1), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (c.scramble_num > 0) {
      // bcc_scramble runs after bcc_load, so A_p above is the loaded data's; the control
      // frequency of column k is its count over the loaded Z rows (lf_Zt_rsum, An_v.c:5-57)
      std::vector<int> zc(M, 0);
      if (has_Z) for (int k = 0; k < M; k++) zc[k] = hc[(size_t)M + k];
      scramble(zc, popmask(hZr_b, NZ));
    }
    std::vector<double> AZ(M, 0.0);
    for (int k = 0; k < M; k++) {
      double a = 0, z = 0;
      if (bget(hAc_b.data(), k)) { a = (double)hc[k]; if (has_Z) z = (double)hc[(size_t)M + k]; }
      AZ[k] = 0; AZ[k] += a; AZ[k] += z;
    }
    if (!c.A_p.empty()) {
      FILE* fp = fopen(c.A_p.c_str(), "rb"); int hdr[3];
      bool ok = fp && fread(hdr, sizeof(int), 3, fp) == 3 && hdr[1] == pcols &&
                fread(hA_p.data(), sizeof(double), pcols, fp) == (size_t)pcols;
      if (fp) fclose(fp);
      if (!ok) throw Error(LKC_ERR_IO, "cannot read A_p");
    } else if (c.sparse) {
      for (int b = 0; b < pcols; b++) {
        int s0 = std::min(M - 1, std::max(0, b * kPopBits)), s1 = std::min(M - 1, std::max(0, (b + 1) * kPopBits - 1));
        double tsum = 0, tnum = 0;
        for (int k = s0; k <= s1; k++) if (bget(hAc_b.data(), k)) { tsum += AZ[k]; tnum += 1; }
        if (tnum == 0) hA_p[b] = 0.5;
        else { double v = (tsum) / (tnum * (Arb + Zrb)); hA_p[b] = v < 0.01 ? 0.01 : (v > 0.99 ? 0.99 : v); }
      }
    }
  }
  // calc_A_ajdk (An_ajdk_v.c:5-30)
  hajdk.assign((size_t)pcols * kAjdk, 0.0);
  for (int j = 0; j < 5; j++) for (int k = 0; k < 3; k++) {
    int jk = j + 5 * k;
    for (int nc = 0; nc < pcols; nc++)
      hajdk[nc + (size_t)jk * pcols] = pow(hA_p[nc] - (1 - hA_p[nc]), j) * pow(1.0 / std::max(0.01, (4.0 * hA_p[nc] * (1 - hA_p[nc]))), k);
  }
  ajdk = dev.alloc<double>(hajdk.size());
  HIPCHK(hipMemcpyAsync(ajdk, hajdk.data(), sizeof(double) * hajdk.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));

  // xcalc orientation flags (bcc_double_init_lf, bcc.c:1005-1150)
  // bcc_double_init_lf (bcc.c:1005-1040): the first step's row drop from the active counts,
  // the sides by their loaded counts
  int rdrop0 = 0; get_xdrop(popmask(hAr_j, NA), popmask(hAc_j, M), c.gamma, c.gamma_type, &rdrop0, NULL);
  int rdrop2 = std::max(2, rdrop0);
  trm_AnZt = (Zrb <= Arb) ? 0 : 1;
  trm_knzt = (Zrb <= std::min(rdrop2, Arb)) ? 0 : 1;
  trm_hnvt = trm_knzt;

  // persistent device arrays
  size_t nA = NA, nZ = std::max(NZ, 1), nM = M, nT = T;
  mlf = dev.alloc<double>(16 * 8);
  mask_lfs();
  lf_An_ajdk = dev.alloc<double>(nA * kAjdk); lf_Zn_ajdk = dev.alloc<double>(nZ * kAjdk);
  jn_ajdk = dev.alloc<double>(nA * kAjdk); vn_ajdk = dev.alloc<double>(nZ * kAjdk);
  kn_ajdk = dev.alloc<double>(nA * kAjdk); hn_ajdk = dev.alloc<double>(nA * kAjdk);
  an_ajdk_ = dev.alloc<double>(nA * kAjdk); zn_ajdk = dev.alloc<double>(nZ * kAjdk);
  lf_AtTn = dev.alloc<double>(nM * nT); lf_ZtSn = dev.alloc<double>(nM * nT);
  lf_AnZt = dev.alloc<double>(nA * nZ); lf_AnAt = dev.alloc<double>(nA * nA);
  lf_AtTAnZtSZn = dev.alloc<double>(nM * nT); lf_AtTAnAtTAn = dev.alloc<double>(nM * nT);
  lf_AnZtSZnAt = dev.alloc<double>(nA * nT); lf_AnAtTAnAt = dev.alloc<double>(nA * nT);
  At_An_a1d1 = dev.alloc<double>(nM); et_An = dev.alloc<double>(nM);
  An_a2d2_ZtSn = dev.alloc<double>(nA * nT); An_a2d2_AtTn = dev.alloc<double>(nA * nT);
  sc = dev.alloc<double>(6 * nT + 2);
  T_AnAt_AnAt = dev.alloc<double>(nT * nA);
  QR_ZSZ = dev.alloc<double>(nA * nT); QR_TAT = dev.alloc<double>(nA * nT);
  QC_ZSZ = dev.alloc<double>(nM * nT); QC_TAT = dev.alloc<double>(nM * nT);
  QR_ZSZ_nrm = dev.alloc<double>(nA * nT); QR_TAT_nrm = dev.alloc<double>(nA * nT);
  QC_ZSZ_nrm = dev.alloc<double>(nM * nT); QC_TAT_nrm = dev.alloc<double>(nM * nT);
  for (int i = 0; i < 14; i++) mt[i] = dev.alloc<double>(nM * nT);
  {
    // zero the persistent arrays once: entries outside the active sets then read as 0
    double* zs[] = { lf_An_ajdk, lf_Zn_ajdk, lf_AtTn, lf_ZtSn, lf_AnZt, lf_AnAt, lf_AtTAnZtSZn, lf_AtTAnAtTAn,
                     lf_AnZtSZnAt, lf_AnAtTAnAt, At_An_a1d1, et_An, An_a2d2_ZtSn, An_a2d2_AtTn, T_AnAt_AnAt,
                     QR_ZSZ, QR_TAT, QC_ZSZ, QC_TAT };
    size_t zn[] = { nA * kAjdk, nZ * kAjdk, nM * nT, nM * nT, nA * nZ, nA * nA, nM * nT, nM * nT,
                    nA * nT, nA * nT, nM, nM, nA * nT, nA * nT, nT * nA, nA * nT, nA * nT, nM * nT, nM * nT };
    for (size_t i = 0; i < sizeof zs / sizeof zs[0]; i++) fill(zs[i], (long long)zn[i], 0.0, st);
  }
  cnt_buf = dev.alloc<int>(std::max(nM, nA) + nT + 8);
  xwork = dev.alloc<double>(1024);
  scal = dev.alloc<double>(16);
  scal_s = dev.alloc<double>(4 * 16);
  dsv = dev.alloc<double>(nM + nA);
  hsv = (double*)pinned_take(sizeof(double) * (nM + nA));
  skey = dev.alloc<double>(nM + nA);
  sidx_in = dev.alloc<int>(nM + nA); sidx = dev.alloc<int>(nM + nA);
  hidx = (int*)pinned_take(sizeof(int) * (nM + nA));
  sort_tmp_bytes = std::max(sort_pairs_temp_bytes(M), sort_pairs_temp_bytes(NA));
  if (!sort_tmp_bytes) throw Error(LKC_ERR_HIP, "radix sort set-up failed");
  sort_tmp = dev.raw(sort_tmp_bytes);
  trace.reserve((size_t)(M + NA + 2) * 6);
  HIPCHK(hipStreamSynchronize(st));
  load_s = now_s() - t0;
}

// ------------------------------------------------------------------
// bcc_scramble (R_init.c:165-193 and 256-285, with R_handle_make R_init.c:5-29,
// xcc_out_xdrop_load / _sort bcc_scorebox.c:163-196, bcc_out_xdrop_prune
// R_init.c:195-238).  The rows and the columns named in an out_xdrop file (each in
// ascending order, only those of the loaded masks) have the bits of every (row, column)
// pair redrawn, row by row and column by column, as Bernoulli(p_k), p_k = the column's
// count over the loaded Z rows / their number, from the reference's Lehmer generator
// (R01GET, boxmuller.c:57-66: seed = seed * 35 mod 2^31 - 1, r = seed / (2^31 - 1) in
// long double).  Both copies of A get the new bits.
// ------------------------------------------------------------------
static const unsigned long kRngMul = 35UL, kRngMod = 2147483647UL;
static inline double r01_next(unsigned long* seed)
{
  *seed = (*seed * kRngMul) % kRngMod;
  return (double)((long double)*seed / (long double)kRngMod);
}
// seed after n steps: seed * 35^n mod (2^31 - 1)
static unsigned long rng_jump(unsigned long seed, unsigned long long n)
{
  unsigned long long r = seed % kRngMod, b = kRngMul;
  while (n) { if (n & 1) r = r * b % kRngMod; b = b * b % kRngMod; n >>= 1; }
  return (unsigned long)r;
}

// xcc_out_xdrop_load + xcc_out_xdrop_sort (bcc_scorebox.c:163-196): one pair per line;
// the rows are the first entries >= 0 and the columns the second, each in file order
static void read_out_xdrop(const std::string& fn, const char* what, std::vector<int>& rows, std::vector<int>& cols)
{
  FILE* fp = fopen(fn.c_str(), "r");
  if (!fp) throw Error(LKC_ERR_IO, std::string("cannot open ") + what + " file " + fn);
  int nl = 0;
  for (int ch; (ch = fgetc(fp)) != EOF;) nl += ch == '\n';
  fseek(fp, 0, SEEK_SET);
  rows.clear(); cols.clear();
  for (int l = 0; l < nl; l++) {
    int a = -1, b = -1;
    if (fscanf(fp, "%d %d", &a, &b) != 2) break;
    if (a >= 0) rows.push_back(a);
    if (b >= 0) cols.push_back(b);
  }
  fclose(fp);
}

void Run::scramble(const std::vector<int>& zcnt, int Zrb_)
{
  for (int s = 0; s < c.scramble_num; s++) {
    std::vector<int> rows, cols;
    read_out_xdrop(c.scramble_files[s], "scramble", rows, cols);
    std::sort(rows.begin(), rows.end());
    std::sort(cols.begin(), cols.end());
    // bcc_out_xdrop_prune: keep the indices inside the loaded masks
    std::vector<int> pr, pc;
    for (int a : rows) if (a >= 0 && a < NA && bget(hAr_b.data(), a)) pr.push_back(a);
    for (int k : cols) if (k >= 0 && k < M && bget(hAc_b.data(), k)) pc.push_back(k);
    if (pr.empty() || pc.empty()) continue;
    // A repeated index is redrawn in place, so its last occurrence decides: the unique
    // rows and columns with the draw index of their last occurrence.
    std::vector<int> ur, uc; std::vector<long long> ri, cj;
    for (size_t i = 0; i < pr.size(); i++) {
      if (!ur.empty() && ur.back() == pr[i]) ri.back() = (long long)i;
      else { ur.push_back(pr[i]); ri.push_back((long long)i); }
    }
    for (size_t j = 0; j < pc.size(); j++) {
      if (!uc.empty() && uc.back() == pc[j]) cj.back() = (long long)j;
      else { uc.push_back(pc[j]); cj.push_back((long long)j); }
    }
    const long long ncol = (long long)pc.size();
    std::vector<double> pk(uc.size());
    for (size_t j = 0; j < uc.size(); j++) pk[j] = (double)zcnt[uc[j]] / (double)Zrb_;
    const int wpr = (int)((uc.size() + 31) / 32);
    std::vector<uint32_t> bits((size_t)ur.size() * wpr, 0u);
    const unsigned long seed0 = c.scramble_seeds[s];
    // rows in parallel: row i of the draw order starts i * ncol draws after the seed
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int t) {
      for (size_t u = t; u < ur.size(); u += nth) {
        unsigned long sd = rng_jump(seed0, (unsigned long long)ri[u] * (unsigned long long)ncol);
        size_t q = 0;
        for (long long j = 0; j < ncol && q < uc.size(); j++) {
          const double r = r01_next(&sd);
          if (j == cj[q]) {
            if (r < pk[q]) bits[u * wpr + q / 32] |= 1u << (q % 32);
            q++;
          }
        }
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; t++) th.emplace_back(work, t);
    work(0);
    for (auto& t : th) t.join();
    int* dr = dev.alloc<int>(ur.size() + uc.size());
    uint32_t* db = dev.alloc<uint32_t>(bits.size());
    HIPCHK(hipMemcpyAsync(dr, ur.data(), sizeof(int) * ur.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dr + ur.size(), uc.data(), sizeof(int) * uc.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(db, bits.data(), sizeof(uint32_t) * bits.size(), hipMemcpyHostToDevice, st));
    scramble_bits(An.p, An.rb, At.p, At.rb, dr, (int)ur.size(), dr + ur.size(), (int)uc.size(), db, st);
    HIPCHK(hipStreamSynchronize(st));
    dev.release(dr); dev.release(db);
  }
}

// masks and lists: host -> pinned block -> device (one copy)
void Run::push_ctl()
{
  const std::vector<uint8_t>* hm[11] = { &hAr_b, &hAr_j, &hAr_rtn, &hAr_rmv, &hZr_b, &hZr_j, &hAc_b, &hAc_j, &hAc_rtn, &hAc_rmv, &hTc_j };
  int mn[11] = { NA, NA, NA, NA, NZ, NZ, M, M, M, M, T };
  for (int i = 0; i < 11; i++) memcpy(hctl + ctl_mask_off[i], hm[i]->data(), bsize(mn[i]));
  // set positions of a mask in ascending order: zero bytes skipped, set bits taken
  // MSB-first with a count of leading zeros
  // (written into a list sized for every position, then cut: full bytes take eight
  // consecutive positions at once, empty 8-byte groups are skipped)
  auto fill_list = [&](IList& l, const std::vector<uint8_t>& m, int n) {
    l.h.resize((size_t)std::max(n, 0));
    int* o = l.h.data();
    int c = 0;
    const int nb = (n + 7) / 8;
    int b = 0;
    while (b < nb) {
      if (b + 8 <= nb) {
        uint64_t w;
        memcpy(&w, m.data() + b, 8);
        if (!w) { b += 8; continue; }
      }
      const unsigned v0 = m[b];
      if (v0 == 0xFFu && 8 * b + 8 <= n) {
        for (int k = 0; k < 8; k++) o[c++] = 8 * b + k;
      } else {
        unsigned v = v0;
        while (v) {
          const int k = __builtin_clz(v) - 24, i = 8 * b + k;
          if (i >= n) break;
          o[c++] = i;
          v &= ~(0x80u >> k);
        }
      }
      b++;
    }
    l.h.resize(c);
  };
  // the retained sets equal the active ones after the drop (bcc_lrup_mxdup): copied
  auto same = [](const std::vector<uint8_t>& x, const std::vector<uint8_t>& y, int n) {
    return memcmp(x.data(), y.data(), (size_t)(n + 7) / 8) == 0;
  };
  fill_list(ar, hAr_j, NA); fill_list(kl, hAc_j, M); fill_list(tl, hTc_j, T); fill_list(zr, hZr_j, NZ);
  if (same(hAr_rtn, hAr_j, NA)) rtn.h = ar.h; else fill_list(rtn, hAr_rtn, NA);
  fill_list(rmv, hAr_rmv, NA);
  if (same(hAc_rtn, hAc_j, M)) crt.h = kl.h; else fill_list(crt, hAc_rtn, M);
  fill_list(crm, hAc_rmv, M);
  // compacted column layout: per non-empty 1920-block a segment of whole words (offset
  // and length in 32-column words, block), the source column of each position and the
  // mask of the valid positions; the row length is padded to whole 4-word groups
  auto layout = [&](const std::vector<int>& cols, int* seg, int* src, uint32_t* cm, int& Wout, int& nsout) {
    int W = 0, ns = 0;
    size_t i = 0;
    while (i < cols.size()) {
      int b = cols[i] / kPopBits;
      size_t j = i;
      while (j < cols.size() && cols[j] / kPopBits == b) j++;
      int cnt = (int)(j - i), nw = (cnt + 31) / 32;
      seg[3 * ns] = W; seg[3 * ns + 1] = nw; seg[3 * ns + 2] = b; ns++;
      for (int q = 0; q < 32 * nw; q++) src[32 * W + q] = q < cnt ? cols[i + q] : -1;
      if (cm)
        for (int w = 0; w < nw; w++) {
          uint32_t v = 0;
          for (int q = 0; q < 32; q++) if (32 * w + q < cnt) v |= 1u << ((q >> 3) * 8 + (7 - (q & 7)));
          cm[W + w] = v;
        }
      W += nw; i = j;
    }
    const int Wp = (W + 3) / 4 * 4;
    for (int q = 32 * W; q < 32 * Wp; q++) src[q] = -1;
    if (cm) for (int w = W; w < Wp; w++) cm[w] = 0;
    Wout = Wp; nsout = ns;
  };
  {
    layout(crm.h, (int*)(hctl + ctl_cmp_off[0]), (int*)(hctl + ctl_cmp_off[1]), (uint32_t*)(hctl + ctl_cmp_off[2]),
           cmp_W, cmp_nseg);
    // kept columns: compacted once fewer than half of the columns remain (the full layout
    // would walk every column of every row under the mask)
    kc_on = ww_kernel == 0 && 2LL * crt.n() < (long long)M && crt.n() > 0;
    if (kc_on) layout(crt.h, (int*)(hctl + ctl_cmp_off[7]), (int*)(hctl + ctl_cmp_off[8]), nullptr, kc_W, kc_nseg);
    else { kc_W = 0; kc_nseg = 0; }
    // removed and retained rows, each packed in whole 128-bit groups
    auto pack = [&](const IList& l, int* srr, uint32_t* cmr) {
      int nr = l.n(), W = (((nr + 31) / 32) + 3) / 4 * 4;
      for (int q = 0; q < 32 * W; q++) srr[q] = q < nr ? l.h[q] : -1;
      for (int w = 0; w < W; w++) {
        uint32_t v = 0;
        for (int q = 0; q < 32; q++) if (32 * w + q < nr) v |= 1u << ((q >> 3) * 8 + (7 - (q & 7)));
        cmr[w] = v;
      }
      return W;
    };
    rmv_W = pack(rmv, (int*)(hctl + ctl_cmp_off[3]), (uint32_t*)(hctl + ctl_cmp_off[4]));
    rtn_W = pack(rtn, (int*)(hctl + ctl_cmp_off[5]), (uint32_t*)(hctl + ctl_cmp_off[6]));
  }
  // Only what this pass uses travels: after the fixed head, the lists and tables are
  // packed in order at their used sizes (each at most its reserved size, so a table moved
  // down never overwrites one not yet moved), their device pointers set to match, and the
  // packed block goes up in one copy.
  size_t off = ctl_head;
  auto next = [&](size_t bytes) { const size_t o = off; off += std::max<size_t>(256, (bytes + 255) & ~(size_t)255); return o; };
  IList* L[8] = { &ar, &kl, &tl, &zr, &rtn, &rmv, &crt, &crm };
  for (int i = 0; i < 8; i++) {
    const size_t n = L[i]->h.size(), o = next(sizeof(int) * (n + 1));
    if (n) memcpy(hctl + o, L[i]->h.data(), sizeof(int) * n);
    L[i]->d = (int*)(dctl + o);
  }
  auto move = [&](int k, size_t bytes) {
    const size_t o = next(bytes);
    if (bytes && o != ctl_cmp_off[k]) memmove(hctl + o, hctl + ctl_cmp_off[k], bytes);
    return dctl + o;
  };
  d_seg = (int*)move(0, sizeof(int) * 3 * (size_t)cmp_nseg);
  d_src = (int*)move(1, sizeof(int) * 32 * (size_t)cmp_W);
  d_cmask = (uint32_t*)move(2, sizeof(uint32_t) * (size_t)cmp_W);
  d_src_rmv = (int*)move(3, sizeof(int) * 32 * (size_t)rmv_W);
  d_cmask_rmv = (uint32_t*)move(4, sizeof(uint32_t) * (size_t)rmv_W);
  d_src_rtn = (int*)move(5, sizeof(int) * 32 * (size_t)rtn_W);
  d_cmask_rtn = (uint32_t*)move(6, sizeof(uint32_t) * (size_t)rtn_W);
  d_seg_k = (int*)move(7, sizeof(int) * 3 * (size_t)kc_nseg);
  d_src_k = (int*)move(8, sizeof(int) * 32 * (size_t)kc_W);
  if (off > ctl_bytes) throw Error(LKC_ERR_STATE, "control block overflow");
  HIPCHK(hipMemcpyAsync(dctl, hctl, off, hipMemcpyHostToDevice, st));
  ctl_epoch++;
  mask_lfs();
}

// the mask_lf slots of every mask a pass reads (each a function of its mask only), in
// one launch after each upload: 0 active columns, 1 / 2 removed / retained columns (with
// the ajdk weights), 3 active A rows, 4 and 7 active Z rows, 5 / 6 retained / removed A rows
void Run::mask_lfs()
{
  if (!mlf || !ajdk) return;     // before the tables exist (the first upload at load)
  MaskLfJobs J{};
  auto job = [&](const uint32_t* m, int pc, const double* w, int slot) { J.job[J.n++] = MaskLfJob{ m, pc, w, mlf + 16 * slot }; };
  job(Ac_j, pcols, ajdk, 0); job(Ac_rmv, pcols, ajdk, 1); job(Ac_rtn, pcols, ajdk, 2);
  job(Ar_j, psize(NA), nullptr, 3); job(Ar_rtn, psize(NA), nullptr, 5); job(Ar_rmv, psize(NA), nullptr, 6);
  if (has_Z) { job(Zr_j, psize(NZ), nullptr, 4); job(Zr_j, psize(NZ), nullptr, 7); }
  mask_lf_jobs(J, st);
}

// ------------------------------------------------------------------
// stages
// ------------------------------------------------------------------
// bcc_An_ajdk (bcc_lf.c:7-51)
void Run::st_ajdk()
{
  // After the update the active rows and columns are the retained ones, whose An_ajdk
  // stage a0 has just formed with the same kernel on the same inputs (its an_ajdk /
  // zn_ajdk): those are copied instead of recomputed.
  const bool reuse = a0_valid && hAr_j == a0_Ar && hAc_j == a0_Ac && hZr_j == a0_Zr;
  a0_valid = false;
  if (reuse) HIPCHK(hipMemcpyAsync(lf_An_ajdk, an_ajdk_, sizeof(double) * kAjdk * (size_t)NA, hipMemcpyDeviceToDevice, st));
  else an_ajdk(An.b(), ar.d, ar.n(), Ac_j, pcols, ajdk, mlf, lf_An_ajdk, NA, st);
  ops += 15LL * ar.n() * (long long)An.rb * 8;
  if (has_Z) {
    if (reuse) HIPCHK(hipMemcpyAsync(lf_Zn_ajdk, zn_ajdk, sizeof(double) * kAjdk * (size_t)NZ, hipMemcpyDeviceToDevice, st));
    else an_ajdk(Zn.b(), zr.d, zr.n(), Ac_j, pcols, ajdk, mlf, lf_Zn_ajdk, NZ, st);
    ops += 15LL * zr.n() * (long long)Zn.rb * 8;
  }
}

// bcc_lf_ZtSn (bcc_lf.c:53-101): plain +-1 products At_k . T_t over the active rows
void Run::st_ZtSn()
{
  // After the update the active rows and columns are the retained ones, whose products
  // stage b0 has just formed (attn, ztsn: the same integers and mask weights): copied.
  const bool reuse = b0_valid && hAr_j == b0_Ar && hAc_j == b0_Ac && hZr_j == b0_Zr;
  b0_valid = false;
  const size_t bytes = sizeof(double) * (size_t)M * T;
  if (reuse) HIPCHK(hipMemcpyAsync(lf_AtTn, mt[4], bytes, hipMemcpyDeviceToDevice, st));
  else pm_few(At.b(), kl.d, kl.n(), Tt.b(), tl.d, tl.n(), Ar_j, psize(NA), mlf + 3 * 16, lf_AtTn, M, st);
  ops += (long long)kl.n() * tl.n() * (long long)bsize(NA) * 8;
  if (has_Z) {
    if (reuse) HIPCHK(hipMemcpyAsync(lf_ZtSn, mt[8], bytes, hipMemcpyDeviceToDevice, st));
    else pm_few(Zt.b(), kl.d, kl.n(), St.b(), tl.d, tl.n(), Zr_j, psize(NZ), mlf + 4 * 16, lf_ZtSn, M, st);
    ops += (long long)kl.n() * tl.n() * (long long)bsize(NZ) * 8;
  }
}

// bcc_lf_YnWt (bcc_lf.c:223-302): centred Grams; lf_AnZt[a*NZ + z], lf_AnAt[a2*NA + a1].
// lf_AnZt is stored with z contiguous (the controls are never dropped, so its rows stay
// dense while A rows leave): the Grams run with the roles of A and Z swapped, which the
// symmetric formula allows (pop((a ^ z) & m) and xa + xz are symmetric, bit for bit).
void Run::st_YnWt()
{
  if (ww_kernel == 0) {
    int *cA = gcnt[0], *cZ = gcnt[1];
    seg_counts(An.b(), ar.d, ar.n(), Ac_j, nullptr, pcols, cA, NA, st);
    if (has_Z) {
      seg_counts(Zn.b(), zr.d, zr.n(), Ac_j, nullptr, pcols, cZ, NZ, st);
      GramMArgs g{ Zn.p, Zn.rb, zr.d, zr.n(), An.p, An.rb, ar.d, ar.n(), Ac_j, nullptr, pcols, cZ, NZ, cA, NA,
                   ajdk, pcols, mlf, lf_Zn_ajdk + 6 * (size_t)NZ, lf_An_ajdk + 6 * (size_t)NA, lf_AnZt, NZ, 1, 1, 0 };
      gram_mfma(g, st);
      ops += (long long)ar.n() * zr.n() * (long long)bsize(M) * 8;
    }
    GramMArgs g{ An.p, An.rb, ar.d, ar.n(), An.p, An.rb, ar.d, ar.n(), Ac_j, nullptr, pcols, cA, NA, cA, NA,
                 ajdk, pcols, mlf, lf_An_ajdk + 6 * (size_t)NA, lf_An_ajdk + 6 * (size_t)NA, lf_AnAt, NA, 1, 1, 0 };
    gram_mfma(g, st);
    ops += (long long)ar.n() * ar.n() * (long long)bsize(M) * 8;
    return;
  }
  if (has_Z) {
    gram(Zn.b(), zr.d, zr.n(), An.b(), ar.d, ar.n(), Ac_j, pcols, ajdk, mlf,
         lf_Zn_ajdk + 6 * (size_t)NZ, lf_An_ajdk + 6 * (size_t)NA, 1, lf_AnZt, NZ, 1, 1, 0, st);
    ops += (long long)ar.n() * zr.n() * (long long)bsize(M) * 8;
  }
  gram(An.b(), ar.d, ar.n(), An.b(), ar.d, ar.n(), Ac_j, pcols, ajdk, mlf,
       lf_An_ajdk + 6 * (size_t)NA, lf_An_ajdk + 6 * (size_t)NA, 1, lf_AnAt, NA, 1, 1, 0, st);
  ops += (long long)ar.n() * ar.n() * (long long)bsize(M) * 8;
}

// bcc_M_YnWt_excerpt + bcc_lf_At_T_YnWt_S_Zn_excerpt (bcc_lf.c:304-316, 615-634).
// L(r1, r2) = src[idx1(r1)*st1 + idx2(r2)*st2] with idx = the side's original index when full.
// The quantisation (xcalc into the slot's integers and scalars) and the contraction are
// separate steps: the low-rank update quantises its four terms (stage a2) before it
// contracts them (stage a3).
void Run::ww_quantise(int slot, const WwTerm& w)
{
  const Side& pn = w.trm ? w.s2 : w.s1;
  const Side& pw = w.trm ? w.s1 : w.s2;
  int nn = pn.rows->n(), nw = pw.rows->n();
  ops += (long long)nn * nw * (c.B + 1) * 8;
  if (nn == 0 || nw == 0) return;
  long long sn, sw; const int *rl = nullptr, *cl = nullptr;
  if (!w.trm) { sn = w.st1; sw = w.st2; if (w.full) { rl = w.s1.rows->d; cl = w.s2.rows->d; } }
  else { sn = w.st2; sw = w.st1; if (w.full) { rl = w.s2.rows->d; cl = w.s1.rows->d; } }
  int n_total = pn.X->nc;
  size_t prb = pn.X->rb;
  long long ldc = std::max(n_total, 1);
  uint32_t* cz = (uint32_t*)g_czs[slot].need(dev, sizeof(uint32_t) * (size_t)std::max(nw, 1) * ldc);
  // The matrix-core contraction runs over the patients of the n side only: packed to the
  // front when a packed copy is given (the quantised values then go to list positions),
  // else the rows' whole 128-bit groups.  The integer sums over n do not depend on where
  // each patient sits, so both forms give the same result.
  const bool packed = ww_kernel == 0 && w.pk.nbits > 0;
  const int* nout = packed ? d_iota : pn.rows->d;
  xcalc(w.src, sn, sw, rl, cl, nn, nw, nout, c.B, c.D_MLT, cz, ldc, scal_s + 16 * slot, xwork, st);
  if (ww_kernel == 1) {
    uint8_t* planes = (uint8_t*)g_planess[slot].need(dev, (size_t)std::max(nw, 1) * c.B * prb);
    cz_planes(cz, ldc, nw, n_total, pn.mask, c.B, planes, prb, st);
  }
}

void Run::ww_contract(int slot, const WwTerm& w)
{
  const Side& pn = w.trm ? w.s2 : w.s1;
  const Side& pw = w.trm ? w.s1 : w.s2;
  int nn = pn.rows->n(), nw = pw.rows->n();
  int n_total = pn.X->nc;
  size_t prb = pn.X->rb;
  ops += (long long)tl.n() * w.kls->n() * nw * (c.B + 1) * (long long)prb * 8;
  if (nn == 0 || nw == 0) {
    // no rows on one side: every sum of the reference's kernel is over an empty set, so
    // each output is +0.0 (the low-rank update after a drop of only rows or only columns)
    fill(w.out, (long long)M * T, 0.0, st);
    return;
  }
  long long ldc = std::max(n_total, 1);
  const bool packed = ww_kernel == 0 && w.pk.nbits > 0;
  WwArgs a;
  a.P = pn.X->b(); a.klist = w.kls->d; a.qlist = w.kls->d; a.nk = w.kls->n(); a.pm = pn.mask;
  a.Tc = pn.Cov->b(); a.tlist = tl.d; a.slist = tl.d; a.nt = tl.n();
  a.planes = ww_kernel == 1 ? (const uint8_t*)g_planess[slot].p : nullptr; a.prb = prb; a.row0 = pn.mask;
  a.nw = nw; a.wlist = pw.rows->d; a.B = c.B;
  a.cz = (const uint32_t*)g_czs[slot].p; a.ldc = ldc;
  a.Sc = pw.Cov->b(); a.Q = pw.X->b(); a.ajdk = ajdk; a.pcols = pcols; a.scal = scal_s + 16 * slot;
  a.out = w.out; a.ldo = M; a.out_elems = (long long)M * T;
  if (packed) {
    a.P = Bits{ w.pk.X, pn.X->nr, w.pk.nbits, w.pk.rb }; a.Tc = Bits{ w.pk.Cov, pn.Cov->nr, w.pk.nbits, w.pk.rb };
    a.pm = w.pk.mask; a.row0 = w.pk.mask; a.prb = (size_t)w.pk.nbits / 8;
  } else if (ww_kernel == 0) {
    a.prb = std::min(prb, (size_t)((n_total + 127) / 128) * 16);
  }
  // timed: the contraction kernel itself (the quantised-value preparation is not included)
  struct Ctx { Prof* prof; Prof::Pending pp; double work; };
  Ctx cx{ &prof, Prof::Pending(), 2.0 * tl.n() * w.kls->n() * nw * (double)nn * ((c.B + 7) / 8) };
  WwHook hook{ [](void* p, hipStream_t s) { Ctx* x = (Ctx*)p; x->prof->begin("ww", x->work, s, x->pp); },
               [](void* p, hipStream_t s) { Ctx* x = (Ctx*)p; x->prof->end(x->pp, s); }, &cx };
  if (ww_kernel == 1) {
    size_t wb = ww_work_bytes(a.nk, a.nt, nw);
    hook.before(&cx, st);
    HIPCHK(ww_valu(a, (double*)g_work.need(dev, wb), wb, st));
    hook.after(&cx, st);
  } else {
    size_t wb = ww_mfma_work_bytes(a.nk, a.nt, nw, a.prb, c.B, a.pcols);
    HIPCHK(ww_mfma(a, g_work.need(dev, wb), wb, st, &hook));
  }
}

void Run::ww_excerpt(int trm, const Side& s1, const Side& s2, const IList& kls, const double* src, long long st1,
                     long long st2, int full, double* out, const Packed* pk)
{
  const WwTerm w{ trm, s1, s2, &kls, src, st1, st2, full, out, pk ? *pk : Packed{} };
  ww_quantise(0, w);
  ww_contract(0, w);
}

// Packed copies of one patient group (row list src, -1 = padding, 32 W entries):
// SNP-major rows Atc (M x 4W bytes) through a row gather of An and a bit transpose,
// covariate rows Ttc (T x 4W bytes) by column compaction of Tt.
void Run::pack_rows(const int* src, int W, uint32_t* Atc, uint32_t* Ttc)
{
  if (W <= 0) return;
  uint8_t* tmp = (uint8_t*)g_ctmp.need(dev, (size_t)32 * W * An.rb);
  gather_rows(An.b(), src, 32 * W, tmp, st);
  transpose_bits(Bits{ tmp, 32 * W, M, An.rb }, (uint8_t*)Atc, (size_t)W * 4, st);
  compact_cols(Tt.b(), nullptr, T, src, W, Ttc, W, st);
}

// bcc_M_YnWt + bcc_lf_At_T_YnWt_S_Zn (bcc_lf.c:318-356, 636-709), initial pass
void Run::st_At_T()
{
  Side sA{ &At, Ar_j, &ar, &Tt }, sZ{ &Zt, Zr_j, &zr, &St };
  if (has_Z) ww_excerpt(trm_AnZt, sA, sZ, kl, lf_AnZt, NZ, 1, 1, lf_AtTAnZtSZn);
  ww_excerpt(0, sA, sA, kl, lf_AnAt, 1, NA, 1, lf_AtTAnAtTAn);
}

// bcc_lf_AnZt_S_WnYt (bcc_lf.c:360-409)
// timed as "qr": work = the bytes of the loop-subscores the row sums read (8 per pair)
void Run::st_QR()
{
  Prof::Pending pq;
  if (has_Z) {
    signed char* sg = (signed char*)g_qrs.need(dev, qr_sign_bytes(zr.n(), tl.n()));
    prof.begin("qr", 8.0 * ar.n() * (double)zr.n(), st, pq);
    qr(lf_AnZt, NZ, ar.d, ar.n(), zr.d, zr.n(), St.b(), tl.d, tl.n(), lf_AnZtSZnAt, NA, sg, st, 1);
    prof.end(pq, st);
  }
  signed char* sg2 = (signed char*)g_qrs2.need(dev, qr_sign_bytes(ar.n(), tl.n()));
  prof.begin("qr", 8.0 * ar.n() * (double)ar.n(), st, pq);
  qr(lf_AnAt, NA, ar.d, ar.n(), ar.d, ar.n(), Tt.b(), tl.d, tl.n(), lf_AnAtTAnAt, NA, sg2, st);
  prof.end(pq, st);
}

// bcc_singlestudy_ww (bcc_flattenloop.c:753-831) and get_singlestudy_vv (bcc_flattenloop.c:359-731)
void Run::st_single()
{
  const double* x6 = lf_An_ajdk + 6 * (size_t)NA;
  // At_Yn_a1d1: xcalc of the rows' a1d1 values, then An_ZtSn_ww over the columns
  size_t prb = At.rb;
  uint32_t* cz = (uint32_t*)g_cz.need(dev, sizeof(uint32_t) * (size_t)NA);
  uint8_t* planes = (uint8_t*)g_planes.need(dev, (size_t)c.B * prb);
  xcalc(x6, 1, 0, ar.d, nullptr, ar.n(), 1, ar.d, c.B, c.D_MLT, cz, NA, scal, xwork, st);
  ops += (long long)ar.n() * (c.B + 1) * 8;
  if (ww_kernel == 0) {
    size_t wb = ztsn_mfma_work_bytes(kl.n(), 1, prb, c.B);
    HIPCHK(ztsn_mfma(At.b(), kl.d, kl.n(), Ar_j, cz, NA, NA, 1, d_zero, c.B, scal, At_An_a1d1, 0, g_ztsn.need(dev, wb), wb, st));
  } else {
    cz_planes(cz, NA, 1, NA, Ar_j, c.B, planes, prb, st);
    an_ztsn(At.b(), kl.d, kl.n(), Ar_j, planes, prb, Ar_j, 1, d_zero, c.B, scal, At_An_a1d1, 0, st);
  }
  ops += (long long)kl.n() * (c.B + 1) * (long long)prb * 8;
  // ordered sums over the active rows
  fill(sc, 6 * (long long)T + 2, 0.0, st);
  OrderedJobs J; J.n = 0;
  auto add = [&](const IList& l, const DMat* sgn, int t, const double* vals, double* out) {
    OrderedJob& q = J.job[J.n++];
    q.list = l.d; q.n = l.n(); q.sign = sgn ? sgn->p + (size_t)t * sgn->rb : nullptr; q.vals = vals; q.out = out;
  };
  for (int t : tl.h) {
    add(ar, &Tt, t, lf_An_ajdk + 11 * (size_t)NA, sc + 0 * T + t);
    add(ar, &Tt, t, lf_An_ajdk + 13 * (size_t)NA, sc + 1 * T + t);
    if (has_Z) {
      add(zr, &St, t, lf_Zn_ajdk + 11 * (size_t)NZ, sc + 2 * T + t);
      add(zr, &St, t, lf_Zn_ajdk + 13 * (size_t)NZ, sc + 3 * T + t);
    }
  }
  add(ar, nullptr, 0, x6, sc + 6 * T);
  ordered_sums(J, st);
  // et_An[k] = -|rows| + 2 pop(At_k & rows); et_Tn, et_Sn likewise
  popc_rows(At.b(), kl.d, kl.n(), Ar_j, (int)(At.rb / 4), cnt_buf, st);
  et_from_counts(cnt_buf, kl.d, kl.n(), ar.n(), et_An, 0, st);
  int* ct = cnt_buf + std::max(M, NA);
  popc_rows(Tt.b(), tl.d, tl.n(), Ar_j, (int)(Tt.rb / 4), ct, st);
  et_from_counts(ct, tl.d, tl.n(), ar.n(), sc + 4 * T, 1, st);
  if (has_Z) {
    popc_rows(St.b(), tl.d, tl.n(), Zr_j, (int)(St.rb / 4), ct, st);
    et_from_counts(ct, tl.d, tl.n(), zr.n(), sc + 5 * T, 1, st);
  }
  t_anat(ar.d, ar.n(), tl.d, tl.n(), Tt.b(), x6, mlf, T_AnAt_AnAt, NA, st);
}

// bcc_doublestudy_ww (bcc_flattenloop.c:977-1029) with An_a2d2_ZtSn_excerpt (An_ZtSn_ww.c:267-290)
void Run::st_double() { ztsn_weighted(ajdk + 12 * (size_t)pcols, An_a2d2_ZtSn, An_a2d2_AtTn); }

// An_ZtSn_ww of the block-weighted ZtSn and AtTn (An_a2d2_ZtSn_excerpt / An_a0d1_ZtSn_excerpt,
// An_ZtSn_ww.c:243-290, then xcalc): out[t*NA + a] over the active rows and columns
void Run::ztsn_weighted(const double* a2d2, double* outZ, double* outA)
{
  int nk = kl.n(), nt = tl.n();
  double* L = (double*)g_L.need(dev, sizeof(double) * ((size_t)std::max(nk, 1) * std::max(nt, 1)));
  size_t prb = An.rb;
  uint32_t* cz = (uint32_t*)g_cz.need(dev, sizeof(uint32_t) * (size_t)std::max(nt, 1) * M);
  uint8_t* planes = (uint8_t*)g_planes.need(dev, (size_t)std::max(nt, 1) * c.B * prb);
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 0 && !has_Z) continue;
    const double* src = pass == 0 ? lf_ZtSn : lf_AtTn;
    double* out = pass == 0 ? outZ : outA;
    scale_rows(src, M, kl.d, nk, tl.d, nt, a2d2, L, st);
    xcalc(L, 1, nk, nullptr, nullptr, nk, nt, kl.d, c.B, c.D_MLT, cz, M, scal, xwork, st);
    if (ww_kernel == 0) {
      size_t wb = ztsn_mfma_work_bytes(ar.n(), nt, prb, c.B);
      Prof::Pending pz;
      // timed with its HBM bytes: the active rows (whole packed rows) and the quantised values
      prof.begin("ztsn", (double)ar.n() * prb + 4.0 * nt * (double)M, st, pz);
      HIPCHK(ztsn_mfma(An.b(), ar.d, ar.n(), Ac_j, cz, M, M, nt, tl.d, c.B, scal, out, NA, g_ztsn.need(dev, wb), wb, st));
      prof.end(pz, st);
    } else {
      cz_planes(cz, M, nt, M, Ac_j, c.B, planes, prb, st);
      an_ztsn(An.b(), ar.d, ar.n(), Ac_j, planes, prb, Ac_j, nt, tl.d, c.B, scal, out, NA, st);
    }
    ops += (long long)ar.n() * nt * (c.B + 1) * (long long)prb * 8;
  }
}

// get_flattenloop (bcc_flattenloop.c:1431-1889), one bin, no Y
void Run::st_flatten()
{
  FlatQR f;
  f.ar = ar.d; f.nar = ar.n(); f.tl = tl.d; f.nt = tl.n(); f.T = T; f.NA = NA; f.has_Z = has_Z;
  f.ajdk_rows = lf_An_ajdk; f.mlf = mlf; f.sc = sc;
  f.lf_AnZtSZnAt = lf_AnZtSZnAt; f.lf_AnAtTAnAt = lf_AnAtTAnAt;
  f.An_a2d2_ZtSn = An_a2d2_ZtSn; f.An_a2d2_AtTn = An_a2d2_AtTn; f.T_AnAt_AnAt = T_AnAt_AnAt;
  f.Tt = Tt.p; f.rbT = Tt.rb; f.QR_ZSZ = QR_ZSZ; f.QR_TAT = QR_TAT;
  flat_qr(f, st);
  FlatQC g;
  g.kl = kl.d; g.nk = kl.n(); g.tl = tl.d; g.nt = tl.n(); g.T = T; g.M = M; g.has_Z = has_Z;
  g.ajdk = ajdk; g.pcols = pcols; g.mlf = mlf; g.sc = sc;
  g.lf_AtTAnZtSZn = lf_AtTAnZtSZn; g.lf_AtTAnAtTAn = lf_AtTAnAtTAn; g.lf_AtTn = lf_AtTn; g.lf_ZtSn = lf_ZtSn;
  g.et_An = et_An; g.At_An_a1d1 = At_An_a1d1; g.rpop = ar.n();
  g.QC_ZSZ = QC_ZSZ; g.QC_TAT = QC_TAT;
  flat_qc(g, st);
}

// ------------------------------------------------------------------
// dexcluster_driver's half-loop subscores (dexcluster_driver.c:22-25, 44-47), one bin, no
// Y/W: dcc_lf_D_AtTn_ZtSn_vv (dcc_lf.c:102-163), dcc_lf_TAnZtS_ww stages 0-2 (dcc_lf.c:267-288,
// TAnZtS.c:5-540) and dcc_wrap_dcc_halfloop (dcc_lf.c:379-530).  The scores land in the bcc
// score arrays: QR_TAT / QR_ZSZ hold QR_TAnAtT / QR_TAnZtS, QC_TAT / QC_ZSZ hold
// QC_TAnAtT / QC_TAnZtS.  Runs after st_ajdk and st_ZtSn of the same masks.
// ------------------------------------------------------------------
void Run::st_dcc()
{
  const int nk = kl.n(), nt = tl.n();
  // etTn, etSn (-|rows| + 2 pop of the covariate rows over the active patients) into
  // sc + 4T, sc + 5T; etAn of the active SNP rows into et_An
  int* ct = cnt_buf + std::max(M, NA);
  popc_rows(Tt.b(), tl.d, nt, Ar_j, (int)(Tt.rb / 4), ct, st);
  et_from_counts(ct, tl.d, nt, ar.n(), sc + 4 * T, 1, st);
  if (has_Z) {
    popc_rows(St.b(), tl.d, nt, Zr_j, (int)(St.rb / 4), ct, st);
    et_from_counts(ct, tl.d, nt, zr.n(), sc + 5 * T, 1, st);
  }
  popc_rows(At.b(), kl.d, nk, Ar_j, (int)(At.rb / 4), cnt_buf, st);
  et_from_counts(cnt_buf, kl.d, nk, ar.n(), et_An, 1, st);
  // stage 0: a1d1_AtTn, a1d1_ZtSn (sc, sc + 2T) as ordered sums over the active rows
  OrderedJobs J; J.n = 0;
  for (int t : tl.h) {
    OrderedJob& q = J.job[J.n++];
    q.list = ar.d; q.n = ar.n(); q.sign = Tt.p + (size_t)t * Tt.rb; q.vals = lf_An_ajdk + 6 * (size_t)NA; q.out = sc + t;
    if (has_Z) {
      OrderedJob& z = J.job[J.n++];
      z.list = zr.d; z.n = zr.n(); z.sign = St.p + (size_t)t * St.rb; z.vals = lf_Zn_ajdk + 6 * (size_t)NZ; z.out = sc + 2 * T + t;
    }
  }
  ordered_sums(J, st);
  DccQC q{ kl.d, nk, tl.d, nt, M, lf_AtTn, lf_ZtSn, sc + 4 * T, sc + 5 * T, et_An, ar.n(), ajdk, pcols, has_Z, QC_ZSZ, QC_TAT };
  dcc_qc(q, st);
  // stage 1: the a0d1-weighted products (AJDK_0_1) through xcalc and An_ZtSn_ww
  ztsn_weighted(ajdk + 5 * (size_t)pcols, An_a2d2_ZtSn, An_a2d2_AtTn);
  // stage 2 and the half-loop copy
  DccQR r{ ar.d, ar.n(), tl.d, nt, NA, An_a2d2_ZtSn, An_a2d2_AtTn, sc, sc + 2 * T, sc + 4 * T, sc + 5 * T,
           lf_An_ajdk + 6 * (size_t)NA, mlf, Tt.p, Tt.rb, has_Z, QR_ZSZ, QR_TAT };
  dcc_qr(r, st);
}

void Run::init_lf()
{
  if (inited) throw Error(LKC_ERR_STATE, "already initialised");
  double t0 = now_s();
  if (dcc) { st_ajdk(); st_ZtSn(); st_dcc(); }
  else {
    st_ajdk(); st_ZtSn(); st_YnWt(); st_At_T(); st_QR();
    st_single(); st_double(); st_flatten();
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  prof.collect();
  inited = true;
  init_s = now_s() - t0;
}

// bcc_sumscores_nrm (bcc_sumscores.c:549-779; dcc_sumscores.c:478-699), one bin,
// GLOBAL_Ireq <= 0: the four subscores over the active rows and columns divided by their
// counts of summands, into zeroed arrays
void Run::st_nrm()
{
  unsigned int Ar = (unsigned int)ar.n(), Zr = (unsigned int)zr.n();
  long long cpop = kl.n();
  auto den = [](long long lld) { double d = (double)lld; if (fabs(d) < 0.000000000001) d = 1; return d; };
  fill(QR_ZSZ_nrm, (long long)NA * T, 0.0, st); fill(QR_TAT_nrm, (long long)NA * T, 0.0, st);
  fill(QC_ZSZ_nrm, (long long)M * T, 0.0, st); fill(QC_TAT_nrm, (long long)M * T, 0.0, st);
  const ScoreStage qr{ ar.d, ar.n(), tl.d, tl.n(), NA }, qc{ kl.d, kl.n(), tl.d, tl.n(), M };
  if (has_Z)
    score_nrm(qr, QR_ZSZ, den(dcc ? (long long)(Zr - 0) * (long long)(cpop) : (long long)(Zr - 0) * (long long)(cpop) * (long long)(cpop - 1)),
              QR_ZSZ_nrm, st);
  score_nrm(qr, QR_TAT, den(dcc ? (long long)(Ar - 1) * (long long)(cpop) : (long long)(Ar - 1) * (long long)(cpop) * (long long)(cpop - 1)),
            QR_TAT_nrm, st);
  if (has_Z)
    score_nrm(qc, QC_ZSZ, den(dcc ? (long long)(Ar) * (long long)(Zr) : (long long)(Ar) * (long long)(cpop - 1) * (long long)(Zr)),
              QC_ZSZ_nrm, st);
  score_nrm(qc, QC_TAT, den(dcc ? (long long)(Ar) * (long long)(Ar - 1) : (long long)(Ar) * (long long)(cpop - 1) * (long long)(Ar - 1)),
            QC_TAT_nrm, st);
}
// bcc_sumscores_ifT (bcc_sumscores.c:483-548): n_mds = min(6, T - 1); the scale is
// GLOBAL_kappa_squared whatever its sign (the reference's `if (...);{ ... }`), so nothing
// changes when it is 0
void Run::st_ifT()
{
  int n_mds = std::max(0, tl.n() - 1); if (n_mds > 6) n_mds = 6;
  const double dmds = n_mds > 0 ? n_mds * c.kappa2 : 0.0;
  if (!(dmds > 0)) return;
  const ScoreStage qr{ ar.d, ar.n(), tl.d, tl.n(), NA }, qc{ kl.d, kl.n(), tl.d, tl.n(), M };
  if (has_Z) score_ifT(qr, QR_ZSZ_nrm, dmds, 1, st);
  score_ifT(qr, QR_TAT_nrm, dmds, 1, st);
  if (has_Z) score_ifT(qc, QC_ZSZ_nrm, dmds, 0, st);
  score_ifT(qc, QC_TAT_nrm, dmds, 0, st);
}
// bcc_sumscores_srt (bcc_sumscores.c:358-481): with GLOBAL_Ireq <= 0 the scores summed over
// the bins are used as they are (one bin here): nothing to do
void Run::st_srt() {}
// bcc_sumscores_cmb (bcc_sumscores.c:324-356): QC_AtTAnAtTAn -= QC_AtTAnZtSZn and
// QR_AnAtTAnAt -= QR_AnZtSZnAt, first covariate column
void Run::st_cmb()
{
  const ScoreStage qr{ ar.d, ar.n(), tl.d, tl.n(), NA }, qc{ kl.d, kl.n(), tl.d, tl.n(), M };
  if (has_Z && Zrb > 0 && Arb > 0) score_cmb(qc, QC_TAT_nrm, QC_ZSZ_nrm, st);
  if (has_Z) score_cmb(qr, QR_TAT_nrm, QR_ZSZ_nrm, st);
}
// the combined scores of the active columns into dsv[0, nk) and of the active rows into
// dsv[M, M + na), in list (index) order
void Run::score_lists()
{
  st_nrm(); st_ifT(); st_srt(); st_cmb();
  score_gather0(ScoreStage{ kl.d, kl.n(), tl.d, tl.n(), M }, QC_TAT_nrm, dsv, st);
  score_gather0(ScoreStage{ ar.d, ar.n(), tl.d, tl.n(), NA }, QR_TAT_nrm, dsv + M, st);
}

// bcc_sumscores_xij (bcc_sumscores.c:225-322): the active column scores then row scores in
// ascending order (dQuickSort_xij, Quicksort.c:83-141)
void Run::st_xij()
{
  score_gather0(ScoreStage{ kl.d, kl.n(), tl.d, tl.n(), M }, QC_TAT_nrm, dsv, st);
  score_gather0(ScoreStage{ ar.d, ar.n(), tl.d, tl.n(), NA }, QR_TAT_nrm, dsv + M, st);
  int nk = kl.n(), na = ar.n();
  int rdrop = 0, cdrop = 0;
  get_xdrop(na, nk, c.gamma, c.gamma_type, &rdrop, &cdrop);
  int nc_drop = std::min(cdrop, nk), nr_drop = std::min(rdrop, na);
  // The device sorts both
  // columns (stable radix sort of (score, position)); the sorted values are unique, and
  // where the dropped prefix is strictly increasing its positions are the reference
  // quicksort's as well.  Otherwise (ties or NaN in the prefix) the host runs the
  // reference's quicksort on the raw scores.
  HIPCHK(sort_pairs(dsv, skey, sidx_in, sidx, nk, sort_tmp, sort_tmp_bytes, st));
  HIPCHK(sort_pairs(dsv + M, skey + M, sidx_in + M, sidx + M, na, sort_tmp, sort_tmp_bytes, st));
  if (nk) HIPCHK(hipMemcpyAsync(hsv, skey, sizeof(double) * nk, hipMemcpyDeviceToHost, st));
  if (na) HIPCHK(hipMemcpyAsync(hsv + M, skey + M, sizeof(double) * na, hipMemcpyDeviceToHost, st));
  if (nc_drop) HIPCHK(hipMemcpyAsync(hidx, sidx, sizeof(int) * nc_drop, hipMemcpyDeviceToHost, st));
  if (nr_drop) HIPCHK(hipMemcpyAsync(hidx + M, sidx + M, sizeof(int) * nr_drop, hipMemcpyDeviceToHost, st));
  finish_trace();
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  prof.collect();
  QC_sv.assign(hsv, hsv + nk); QC_perm.resize(nk);
  QR_sv.assign(hsv + M, hsv + M + na); QR_perm.resize(na);
  // The device order is the reference quicksort's where the sorted values of the dropped
  // prefix are strictly increasing and below the rest, no value is NaN (a NaN moves the
  // quicksort's pivots), and the quicksort sorts the whole list.  It stops recursing at
  // depth 32768 (GLOBAL_recursion_limit, Quicksort.c:33-40), which a list of at most 32769
  // entries cannot reach; in a longer one a run of equal scores peels one entry per level,
  // so there any tie sends the list to the host quicksort too.  (A longer list of distinct
  // values reaches that depth only when tens of thousands of its entries arrive in sorted
  // order, which scores in column or row order do not.)
  auto strict = [](const double* v, int n, int k) {
    for (int i = 0; i < n; i++) if (v[i] != v[i]) return false;
    const int upto = n > 32769 ? n - 1 : k;
    for (int i = 0; i < upto && i + 1 < n; i++) if (!(v[i] < v[i + 1])) return false;
    return true;
  };
  auto exact_order = [&](int off, int n, std::vector<double>& sv, std::vector<int>& perm) {
    std::vector<double> raw(n);
    if (n) HIPCHK(hipMemcpyAsync(raw.data(), dsv + off, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int j = 0; j < n; j++) perm[j] = j;
    sort_scores(raw.data(), perm.data(), n);
    sv.assign(raw.begin(), raw.end());
  };
  if (strict(QC_sv.data(), nk, nc_drop)) { for (int i = 0; i < nc_drop; i++) QC_perm[i] = hidx[i]; }
  else { exact_order(0, nk, QC_sv, QC_perm); n_sort_fallback++; }
  if (strict(QR_sv.data(), na, nr_drop)) { for (int i = 0; i < nr_drop; i++) QR_perm[i] = hidx[M + i]; }
  else { exact_order(M, na, QR_sv, QR_perm); n_sort_fallback++; }
  xij_done = true;
}

// bcc_sumscores_mxA (bcc_sumscores.c:159-224) with _mxB / _mxC (37-125): get_xdrop's counts, the
// first cdrop columns and rdrop rows of the sorted order leave (_rmv / _rtn masks), the
// trace line and the xdrop lines; bcc_lrup_mxset's masks reach the device.  With dup, the
// retained sets also become the active ones at once (bcc_lrup_mxdup, which the update
// itself does not read: it uses the _rtn / _rmv sets).
void Run::st_mxA(bool dup)
{
  if (!xij_done) throw Error(LKC_ERR_STATE, "bcc_sumscores_mxA needs bcc_sumscores_xij of the same pass");
  xij_done = false;
  int nk = kl.n(), na = ar.n();
  int rpop = na, cpop_i = nk, rdrop = 0, cdrop = 0;
  get_xdrop(rpop, cpop_i, c.gamma, c.gamma_type, &rdrop, &cdrop);
  int nc_drop = std::min(cdrop, cpop_i), nr_drop = std::min(rdrop, rpop);
  int Irem = na > 0 ? 1 : 0;
  hAc_rtn = hAc_j; std::fill(hAc_rmv.begin(), hAc_rmv.end(), 0);
  hAr_rtn = hAr_j; std::fill(hAr_rmv.begin(), hAr_rmv.end(), 0);
  for (int i = 0; i < nc_drop; i++) { int k = kl.h[QC_perm[i]]; bset(hAc_rmv.data(), k); bclr(hAc_rtn.data(), k); }
  for (int i = 0; i < nr_drop; i++) { int a = ar.h[QR_perm[i]]; bset(hAr_rmv.data(), a); bclr(hAr_rtn.data(), a); }
  double tr[6] = { (double)n_iter, (double)rpop, (double)cpop_i, 0.0, 0.0, (double)Irem };
  trace.insert(trace.end(), tr, tr + 6);
  // the trace line's mean scores (ordered sums over the sorted values); inside the
  // engine's own loop they are formed while the device runs the update
  trace_pending = true; trace_at = trace.size() - 6; trace_rpop = rpop; trace_cpop = cpop_i;
  if (!defer_means) finish_trace();
  n_iter++;                        // D->out_iteration (bcc_sumscores_mxC)
  for (int i = 0; i < nr_drop; i++) {
    int a = ar.h[QR_perm[i]];
    xdrop_a.push_back(a); xdrop_a.push_back(-1); xdrop_b.push_back(Ar_bidx[a]); xdrop_b.push_back(-1);
  }
  for (int i = 0; i < nc_drop; i++) {
    int k = kl.h[QC_perm[i]];
    xdrop_a.push_back(-1); xdrop_a.push_back(k); xdrop_b.push_back(-1); xdrop_b.push_back(Ac_bidx[k]);
  }
  if (dup) { hAr_j = hAr_rtn; hAc_j = hAc_rtn; }
  push_ctl();
  pending_dup = !dup;
}

// bcc_lrup_mxdup (bcc_lrup.c:1128-1139) + bcc_M_mxset (bcc.c:1414-1438): the retained rows
// and columns become the active sets, and the masks and index lists reach the device
void Run::st_mxdup()
{
  if (!pending_dup) return;
  hAr_j = hAr_rtn; hAc_j = hAc_rtn;
  push_ctl();
  pending_dup = false;
}

// one pass's bcc_sumscores_nrm/ifT/srt/cmb/xij/mxA (bcc_sumscores.c:37-779), one bin,
// Ireq <= 0, with the active sets moved to the retained ones at once.  With defer, the
// trace line's means wait for finish_trace (the engine's loops call it once the pass's
// update is enqueued).
void Run::sumscores(bool defer)
{
  st_nrm(); st_ifT(); st_srt(); st_cmb(); st_xij();
  defer_means = defer;
  st_mxA(true);
  defer_means = false;
}

// the mean row and column scores of the last trace line: the sorted values summed in
// order (bcc_sumscores_mxB, bcc_sumscores.c:37-125), a non-finite mean written as 1
void Run::finish_trace()
{
  if (!trace_pending) return;
  trace_pending = false;
  auto mean = [](const std::vector<double>& v, unsigned long long n) {
    double m = 0; for (unsigned long long i = 0; i < n; i++) m += v[i]; m /= (double)n; return m;
  };
  double mr = mean(QR_sv, (unsigned long long)trace_rpop); if (!std::isfinite(mr)) mr = 1.0;
  double mc = mean(QC_sv, (unsigned long long)trace_cpop); if (!std::isfinite(mc)) mc = 1.0;
  trace[trace_at + 3] = mr;
  trace[trace_at + 4] = mc;
}

// a Gram over the full column layout: split over the 1920-blocks when it has few tiles
// (the cost model), or the form the autotuner measured faster for its shape class.  A
// Gram that subtracts in place (op_sub) runs once, so only the others are measured.
void Run::gram_any(const GramMArgs& g)
{
  size_t b = gram_split_bytes(g);
  const size_t bany = gram_split_bytes_any(g);
  const int tm = g.op_sub ? 0 : tune_mode();
  if (tm != 0 && bany > 0 && bany <= ((size_t)256 << 20)) {
    const TuneKey key{ kTuneGram, tune_bucket(g.na), tune_bucket(g.nz), tune_bucket(g.nseg), 0, 0 };
    int ch = -1;
    if (tune_get(key, &ch)) b = ch == 1 ? bany : 0;
    else if (tune_measuring(tm)) {
      struct Ctx { GramMArgs g; uint32_t* X; hipStream_t st; };
      Ctx cx{ g, (uint32_t*)g_split.need(dev, bany), st };
      const int cand[2] = { 0, 1 };
      float best_ms = 0, model_ms = 0;
      const int bi = tune_measure(st, cand, 2, b ? 1 : 0, [](void* p, int v) {
        const Ctx* c = (const Ctx*)p;
        if (v) gram_split(c->g, c->X, c->st); else gram_mfma(c->g, c->st);
      }, &cx, &best_ms, &model_ms, g.out,
         (!g.a_orig && !g.z_orig) ? sizeof(double) * (size_t)g.nz * (size_t)g.ldo : 0);
      tune_put(key, bi, best_ms, model_ms);
      return;   // every measured launch wrote the same Gram
    }
  }
  if (b) gram_split(g, (uint32_t*)g_split.need(dev, b), st);
  else gram_mfma(g, st);
}

// the low-rank update's shared quantities (mask_lf slots, strides, the compacted dropped
// columns, the list sizes of the pass)
#define LRUP_LOCALS \
  [[maybe_unused]] const int sA = psize(NA), sZ = psize(NZ); \
  [[maybe_unused]] double* mlf_rmv = mlf + 1 * 16;     /* column masks */ \
  [[maybe_unused]] double* mlf_rtn = mlf + 2 * 16; \
  [[maybe_unused]] double* mlf_nrtn = mlf + 5 * 16;    /* row masks */ \
  [[maybe_unused]] double* mlf_nrmv = mlf + 6 * 16; \
  [[maybe_unused]] double* mlf_nz = mlf + 7 * 16; \
  [[maybe_unused]] const long long rbm = (long long)An.rb * 8, rbn = (long long)At.rb * 8; \
  [[maybe_unused]] const int W = cmp_W; \
  [[maybe_unused]] Bits Anc{ (const uint8_t*)An_c, NA, W * 32, (size_t)W * 4 }, Znc{ (const uint8_t*)Zn_c, NZ, W * 32, (size_t)W * 4 }; \
  [[maybe_unused]] const bool cols_dropped = cmp_nseg > 0; \
  [[maybe_unused]] int *c_mr = gcnt[0], *c_mm = gcnt[1], *c_mz = gcnt[2]; \
  [[maybe_unused]] const int nrm = rmv.n(), nrt = rtn.n(), nz = zr.n(); \
  [[maybe_unused]] double *t_knzt = mt[0], *t_hnvt = mt[1], *t_ankt = mt[2], *t_knkt = mt[3]; \
  [[maybe_unused]] double *attn = mt[4], *jttn = mt[5], *ktrn = mt[6], *htrn = mt[7], *ztsn = mt[8], *vtsn = mt[9]; \
  [[maybe_unused]] double *vtszn = mt[10], *jttan = mt[11], *htrkn = mt[12], *hhtrkn = mt[13]; \
  [[maybe_unused]] double *knzt = (double*)g_knzt.p, *hnvt = (double*)g_hnvt.p, *ankt = (double*)g_ankt.p, *knkt = (double*)g_knkt.p; \
  [[maybe_unused]] Side s_rmv{ &At, Ar_rmv, &rmv, &Tt }, s_rtn{ &At, Ar_rtn, &rtn, &Tt }, s_Z{ &Zt, Zr_j, &zr, &St }; \
  [[maybe_unused]] const Packed pk_rmv{ (const uint8_t*)At_rmvc, (const uint8_t*)Tt_rmvc, (size_t)rmv_W * 4, d_cmask_rmv, 32 * rmv_W }; \
  [[maybe_unused]] const Packed pk_rtn{ (const uint8_t*)At_rtnc, (const uint8_t*)Tt_rtnc, (size_t)rtn_W * 4, d_cmask_rtn, 32 * rtn_W }

// bcc_lrup_QR_YnWt_stage_0 (bcc_lrup.c:659-701): the mask weights of the dropped and the
// retained columns, the An_ajdk of the retained rows and the controls over the dropped columns
void Run::lrup_QR_0()
{
  LRUP_LOCALS;
  an_ajdk(An.b(), rtn.d, rtn.n(), Ac_rmv, pcols, ajdk, mlf_rmv, jn_ajdk, NA, st);
  ops += 15LL * rtn.n() * rbm;
  if (has_Z) {
    an_ajdk(Zn.b(), zr.d, zr.n(), Ac_rmv, pcols, ajdk, mlf_rmv, vn_ajdk, NZ, st);
    vn_epoch = ctl_epoch;
    ops += 15LL * zr.n() * (long long)Zn.rb * 8;
  }
}

// bcc_lrup_QR_YnWt_stage_1 + _stage_3 (bcc_lrup.c:702-771, 863-895): the Grams of the retained
// rows with the controls and with themselves over the dropped columns, subtracted from
// lf_AnZt and lf_AnAt as they are formed (the Grams are not kept: stage 3 has nothing left)
void Run::lrup_QR_1()
{
  LRUP_LOCALS;
  // compacted rows over the dropped columns: gather those columns as rows of the
  // transposed copies (padding positions give zero rows), transpose back
  {
    const size_t need = (size_t)32 * std::max(W, 1) * std::max(At.rb, Zt.rb);
    uint8_t* tmp = (uint8_t*)g_ctmp.need(dev, need);
    gather_rows(At.b(), d_src, 32 * W, tmp, st);
    transpose_bits(Bits{ tmp, 32 * W, NA, At.rb }, (uint8_t*)An_c, (size_t)W * 4, st);
    if (has_Z) {
      gather_rows(Zt.b(), d_src, 32 * W, tmp, st);
      transpose_bits(Bits{ tmp, 32 * W, NZ, Zt.rb }, (uint8_t*)Zn_c, (size_t)W * 4, st);
    }
  }
  // with no column dropped (a drop of rows only) the QR updates subtract +0.0: skipped
  if (!cols_dropped) {
  } else if (ww_kernel == 0) {
    seg_counts(Anc, rtn.d, rtn.n(), nullptr, d_seg, cmp_nseg, c_mr, NA, st);
    seg_counts(Anc, rmv.d, rmv.n(), nullptr, d_seg, cmp_nseg, c_mm, NA, st);
    if (has_Z) {
      seg_counts(Znc, zr.d, zr.n(), nullptr, d_seg, cmp_nseg, c_mz, NZ, st);
      GramMArgs g{ Znc.p, Znc.rb, zr.d, zr.n(), Anc.p, Anc.rb, rtn.d, rtn.n(), nullptr, d_seg, cmp_nseg, c_mz, NZ, c_mr, NA,
                   ajdk, pcols, mlf_rmv, vn_ajdk + 6 * (size_t)NZ, jn_ajdk + 6 * (size_t)NA, lf_AnZt, NZ, 1, 1, 1 };
      gram_mfma(g, st);
    }
    GramMArgs g{ Anc.p, Anc.rb, rtn.d, rtn.n(), Anc.p, Anc.rb, rtn.d, rtn.n(), nullptr, d_seg, cmp_nseg, c_mr, NA, c_mr, NA,
                 ajdk, pcols, mlf_rmv, jn_ajdk + 6 * (size_t)NA, jn_ajdk + 6 * (size_t)NA, lf_AnAt, NA, 1, 1, 1 };
    gram_mfma(g, st);
  } else {
    if (has_Z)
      gram(Znc, zr.d, zr.n(), Anc, rtn.d, rtn.n(), d_cmask, pcols, ajdk, mlf_rmv,
           vn_ajdk + 6 * (size_t)NZ, jn_ajdk + 6 * (size_t)NA, 1, lf_AnZt, NZ, 1, 1, 1, st, d_seg, cmp_nseg);
    gram(Anc, rtn.d, rtn.n(), Anc, rtn.d, rtn.n(), d_cmask, pcols, ajdk, mlf_rmv,
         jn_ajdk + 6 * (size_t)NA, jn_ajdk + 6 * (size_t)NA, 1, lf_AnAt, NA, 1, 1, 1, st, d_seg, cmp_nseg);
  }
  if (has_Z) ops += (long long)rtn.n() * zr.n() * rbm;
  ops += (long long)rtn.n() * rtn.n() * rbm;
}

// bcc_lrup_QR_YnWt_stage_3 (bcc_lrup.c:863-895): done inside stage 1
void Run::lrup_QR_3() {}

// bcc_lrup_QC_YnWt_stage_a0 (bcc_lrup.c:5-66): An_ajdk of the removed, the retained rows and
// the controls over the retained and the dropped columns
void Run::lrup_QC_a0()
{
  LRUP_LOCALS;
  an_ajdk(An.b(), rmv.d, rmv.n(), Ac_rtn, pcols, ajdk, mlf_rtn, kn_ajdk, NA, st);
  an_ajdk(An.b(), rmv.d, rmv.n(), Ac_rmv, pcols, ajdk, mlf_rmv, hn_ajdk, NA, st);
  an_ajdk(An.b(), rtn.d, rtn.n(), Ac_rtn, pcols, ajdk, mlf_rtn, an_ajdk_, NA, st);
  ops += 15LL * (2LL * rmv.n() + rtn.n()) * rbm;
  if (has_Z) {
    an_ajdk(Zn.b(), zr.d, zr.n(), Ac_rtn, pcols, ajdk, mlf_rtn, zn_ajdk, NZ, st);
    // the controls' An_ajdk over the dropped columns: stage 0 of the QR update formed it
    // since the masks last changed (the same call)
    if (vn_epoch != ctl_epoch) an_ajdk(Zn.b(), zr.d, zr.n(), Ac_rmv, pcols, ajdk, mlf_rmv, vn_ajdk, NZ, st);
    ops += 30LL * zr.n() * (long long)Zn.rb * 8;
  }
  // what the next An_ajdk stage may copy
  a0_valid = true; a0_Ar = hAr_rtn; a0_Ac = hAc_rtn; a0_Zr = hZr_j;
}

// bcc_lrup_QC_YnWt_stage_a1 (bcc_lrup.c:67-196): the compact Grams knzt, hnvt, ankt, knkt
void Run::lrup_QC_a1()
{
  LRUP_LOCALS;
  ankt = (double*)g_ankt.need(dev, sizeof(double) * ((size_t)nrt * nrm + 1));
  knkt = (double*)g_knkt.need(dev, sizeof(double) * ((size_t)nrm * nrm + 1));
  if (has_Z) {
    knzt = (double*)g_knzt.need(dev, sizeof(double) * ((size_t)nrm * nz + 1));
    hnvt = (double*)g_hnvt.need(dev, sizeof(double) * ((size_t)nrm * nz + 1));
  }
  if (ww_kernel == 0 && kc_on) {
    // over the compacted kept columns: gather those columns as rows of the SNP-major
    // copies, transpose back; the segments are the kept columns' 1920-blocks in order
    const int Wk = kc_W;
    uint32_t* Ank = (uint32_t*)g_ank.need(dev, sizeof(uint32_t) * (size_t)NA * Wk);
    uint32_t* Znk = has_Z ? (uint32_t*)g_znk.need(dev, sizeof(uint32_t) * (size_t)NZ * Wk) : nullptr;
    {
      uint8_t* tmp = (uint8_t*)g_ctmp.need(dev, (size_t)32 * Wk * std::max(At.rb, Zt.rb));
      gather_rows(At.b(), d_src_k, 32 * Wk, tmp, st);
      transpose_bits(Bits{ tmp, 32 * Wk, NA, At.rb }, (uint8_t*)Ank, (size_t)Wk * 4, st);
      if (has_Z) {
        gather_rows(Zt.b(), d_src_k, 32 * Wk, tmp, st);
        transpose_bits(Bits{ tmp, 32 * Wk, NZ, Zt.rb }, (uint8_t*)Znk, (size_t)Wk * 4, st);
      }
    }
    const Bits Ak{ (const uint8_t*)Ank, NA, Wk * 32, (size_t)Wk * 4 }, Zk{ (const uint8_t*)Znk, NZ, Wk * 32, (size_t)Wk * 4 };
    int *c_kr = gcnt[3], *c_km = gcnt[4], *c_kz = gcnt[5];
    seg_counts(Ak, rtn.d, nrt, nullptr, d_seg_k, kc_nseg, c_kr, NA, st);
    seg_counts(Ak, rmv.d, nrm, nullptr, d_seg_k, kc_nseg, c_km, NA, st);
    if (has_Z) {
      seg_counts(Zk, zr.d, nz, nullptr, d_seg_k, kc_nseg, c_kz, NZ, st);
      GramMArgs g1{ Ak.p, Ak.rb, rmv.d, nrm, Zk.p, Zk.rb, zr.d, nz, nullptr, d_seg_k, kc_nseg, c_km, NA, c_kz, NZ,
                    ajdk, pcols, mlf_rtn, kn_ajdk + 6 * (size_t)NA, zn_ajdk + 6 * (size_t)NZ, knzt, nrm, 0, 0, 0 };
      gram_mfma(g1, st);
      GramMArgs g2{ Anc.p, Anc.rb, rmv.d, nrm, Znc.p, Znc.rb, zr.d, nz, nullptr, d_seg, cmp_nseg, c_mm, NA, c_mz, NZ,
                    ajdk, pcols, mlf_rmv, hn_ajdk + 6 * (size_t)NA, vn_ajdk + 6 * (size_t)NZ, hnvt, nrm, 0, 0, 0 };
      if (cols_dropped) gram_mfma(g2, st);
      ops += 2LL * nrm * nz * rbm;
    }
    GramMArgs g3{ Ak.p, Ak.rb, rtn.d, nrt, Ak.p, Ak.rb, rmv.d, nrm, nullptr, d_seg_k, kc_nseg, c_kr, NA, c_km, NA,
                  ajdk, pcols, mlf_rtn, an_ajdk_ + 6 * (size_t)NA, kn_ajdk + 6 * (size_t)NA, ankt, nrt, 0, 0, 0 };
    gram_mfma(g3, st);
    GramMArgs g4{ Ak.p, Ak.rb, rmv.d, nrm, Ak.p, Ak.rb, rmv.d, nrm, nullptr, d_seg_k, kc_nseg, c_km, NA, c_km, NA,
                  ajdk, pcols, mlf_rtn, kn_ajdk + 6 * (size_t)NA, kn_ajdk + 6 * (size_t)NA, knkt, nrm, 0, 0, 0 };
    gram_mfma(g4, st);
  } else if (ww_kernel == 0) {
    int *c_kr = gcnt[3], *c_km = gcnt[4], *c_kz = gcnt[5];
    seg_counts(An.b(), rtn.d, nrt, Ac_rtn, nullptr, pcols, c_kr, NA, st);
    seg_counts(An.b(), rmv.d, nrm, Ac_rtn, nullptr, pcols, c_km, NA, st);
    if (has_Z) {
      seg_counts(Zn.b(), zr.d, nz, Ac_rtn, nullptr, pcols, c_kz, NZ, st);
      GramMArgs g1{ An.p, An.rb, rmv.d, nrm, Zn.p, Zn.rb, zr.d, nz, Ac_rtn, nullptr, pcols, c_km, NA, c_kz, NZ,
                    ajdk, pcols, mlf_rtn, kn_ajdk + 6 * (size_t)NA, zn_ajdk + 6 * (size_t)NZ, knzt, nrm, 0, 0, 0 };
      gram_any(g1);
      GramMArgs g2{ Anc.p, Anc.rb, rmv.d, nrm, Znc.p, Znc.rb, zr.d, nz, nullptr, d_seg, cmp_nseg, c_mm, NA, c_mz, NZ,
                    ajdk, pcols, mlf_rmv, hn_ajdk + 6 * (size_t)NA, vn_ajdk + 6 * (size_t)NZ, hnvt, nrm, 0, 0, 0 };
      if (cols_dropped) gram_mfma(g2, st);
      ops += 2LL * nrm * nz * rbm;
    }
    GramMArgs g3{ An.p, An.rb, rtn.d, nrt, An.p, An.rb, rmv.d, nrm, Ac_rtn, nullptr, pcols, c_kr, NA, c_km, NA,
                  ajdk, pcols, mlf_rtn, an_ajdk_ + 6 * (size_t)NA, kn_ajdk + 6 * (size_t)NA, ankt, nrt, 0, 0, 0 };
    gram_any(g3);
    GramMArgs g4{ An.p, An.rb, rmv.d, nrm, An.p, An.rb, rmv.d, nrm, Ac_rtn, nullptr, pcols, c_km, NA, c_km, NA,
                  ajdk, pcols, mlf_rtn, kn_ajdk + 6 * (size_t)NA, kn_ajdk + 6 * (size_t)NA, knkt, nrm, 0, 0, 0 };
    gram_any(g4);
  } else {
    if (has_Z) {
      gram(An.b(), rmv.d, nrm, Zn.b(), zr.d, nz, Ac_rtn, pcols, ajdk, mlf_rtn,
           kn_ajdk + 6 * (size_t)NA, zn_ajdk + 6 * (size_t)NZ, 1, knzt, nrm, 0, 0, 0, st);
      if (cols_dropped)
        gram(Anc, rmv.d, nrm, Znc, zr.d, nz, d_cmask, pcols, ajdk, mlf_rmv,
             hn_ajdk + 6 * (size_t)NA, vn_ajdk + 6 * (size_t)NZ, 1, hnvt, nrm, 0, 0, 0, st, d_seg, cmp_nseg);
      ops += 2LL * nrm * nz * rbm;
    }
    gram(An.b(), rtn.d, nrt, An.b(), rmv.d, nrm, Ac_rtn, pcols, ajdk, mlf_rtn,
         an_ajdk_ + 6 * (size_t)NA, kn_ajdk + 6 * (size_t)NA, 1, ankt, nrt, 0, 0, 0, st);
    gram(An.b(), rmv.d, nrm, An.b(), rmv.d, nrm, Ac_rtn, pcols, ajdk, mlf_rtn,
         kn_ajdk + 6 * (size_t)NA, kn_ajdk + 6 * (size_t)NA, 1, knkt, nrm, 0, 0, 0, st);
  }
  ops += ((long long)nrt * nrm + (long long)nrm * nrm) * rbm;
  // the Gram of the removed rows and the controls over no dropped column: every entry
  // is the reference's empty-mask value, +0.0
  if (has_Z && !cols_dropped && nrm > 0 && nz > 0) fill(hnvt, (long long)nrm * nz, 0.0, st);
}

// the update's four At_T_YnWt_S_Zn terms (bcc_lrup.c:238-304): knzt and hnvt (with Z),
// ankt and knkt, each over the retained columns
const WwTerm* Run::lrup_terms()
{
  LRUP_LOCALS;
  lrup_w[0] = WwTerm{ trm_knzt, s_rmv, s_Z, &crt, knzt, 1, nrm, 0, t_knzt, trm_knzt ? Packed{} : pk_rmv };
  lrup_w[1] = WwTerm{ trm_hnvt, s_rmv, s_Z, &crt, hnvt, 1, nrm, 0, t_hnvt, trm_hnvt ? Packed{} : pk_rmv };
  lrup_w[2] = WwTerm{ 0, s_rtn, s_rmv, &crt, ankt, 1, nrt, 0, t_ankt, pk_rtn };
  lrup_w[3] = WwTerm{ 0, s_rmv, s_rmv, &crt, knkt, 1, nrm, 0, t_knkt, pk_rmv };
  return lrup_w;
}

// bcc_lrup_QC_YnWt_stage_a2 (bcc_lrup.c:197-237): the four Grams quantised (xcalc), each into
// its own slot of integers and scalars
void Run::lrup_QC_a2()
{
  LRUP_LOCALS;
  // the removed and the retained A rows packed to the front of the patient dimension
  if (ww_kernel == 0) {
    pack_rows(d_src_rmv, rmv_W, At_rmvc, Tt_rmvc);
    pack_rows(d_src_rtn, rtn_W, At_rtnc, Tt_rtnc);
  }
  const WwTerm* w = lrup_terms();
  for (int i = 0; i < 4; i++) if (has_Z || i >= 2) ww_quantise(i, w[i]);
}

// bcc_lrup_QC_YnWt_stage_a3 (bcc_lrup.c:238-304): the four At_T_YnWt_S_Zn_ww contractions
void Run::lrup_QC_a3()
{
  LRUP_LOCALS;
  const WwTerm* w = lrup_terms();
  for (int i = 0; i < 4; i++) if (has_Z || i >= 2) ww_contract(i, w[i]);
}

// bcc_lrup_QC_YnWt_stage_b0 (bcc_lrup.c:305-362): plain +-1 products of the retained and the
// dropped columns with the covariates, [t*M + k]
void Run::lrup_QC_b0()
{
  LRUP_LOCALS;
  if (ww_kernel == 0) {
    // over the packed retained / removed patients (the counts are the same integers)
    const Bits Ptn{ (const uint8_t*)At_rtnc, M, 32 * rtn_W, (size_t)rtn_W * 4 }, Ttn{ (const uint8_t*)Tt_rtnc, T, 32 * rtn_W, (size_t)rtn_W * 4 };
    const Bits Pmv{ (const uint8_t*)At_rmvc, M, 32 * rmv_W, (size_t)rmv_W * 4 }, Tmv{ (const uint8_t*)Tt_rmvc, T, 32 * rmv_W, (size_t)rmv_W * 4 };
    pm_few(Ptn, crt.d, crt.n(), Ttn, tl.d, tl.n(), d_cmask_rtn, sA, mlf_nrtn, attn, M, st, rtn_W);
    pm_few(Ptn, crm.d, crm.n(), Ttn, tl.d, tl.n(), d_cmask_rtn, sA, mlf_nrtn, jttn, M, st, rtn_W);
    pm_few(Pmv, crt.d, crt.n(), Tmv, tl.d, tl.n(), d_cmask_rmv, sA, mlf_nrmv, ktrn, M, st, rmv_W);
    pm_few(Pmv, crm.d, crm.n(), Tmv, tl.d, tl.n(), d_cmask_rmv, sA, mlf_nrmv, htrn, M, st, rmv_W);
  } else {
    pm_few(At.b(), crt.d, crt.n(), Tt.b(), tl.d, tl.n(), Ar_rtn, sA, mlf_nrtn, attn, M, st);
    pm_few(At.b(), crm.d, crm.n(), Tt.b(), tl.d, tl.n(), Ar_rtn, sA, mlf_nrtn, jttn, M, st);
    pm_few(At.b(), crt.d, crt.n(), Tt.b(), tl.d, tl.n(), Ar_rmv, sA, mlf_nrmv, ktrn, M, st);
    pm_few(At.b(), crm.d, crm.n(), Tt.b(), tl.d, tl.n(), Ar_rmv, sA, mlf_nrmv, htrn, M, st);
  }
  ops += 2LL * (crt.n() + crm.n()) * tl.n() * rbn;
  if (has_Z) {
    pm_few(Zt.b(), crt.d, crt.n(), St.b(), tl.d, tl.n(), Zr_j, sZ, mlf_nz, ztsn, M, st);
    pm_few(Zt.b(), crm.d, crm.n(), St.b(), tl.d, tl.n(), Zr_j, sZ, mlf_nz, vtsn, M, st);
    ops += (long long)(crt.n() + crm.n()) * tl.n() * (long long)Zt.rb * 8;
  }
  // what the next ZtSn stage may copy
  b0_valid = true; b0_Ar = hAr_rtn; b0_Ac = hAc_rtn; b0_Zr = hZr_j;
}

// bcc_lrup_QC_YnWt_stage_b1 (bcc_lrup.c:363-408): the AtTYn_vv triples; on the matrix cores
// fused with stage b2's ordered sums (the triples are then not kept)
void Run::lrup_QC_b1()
{
  LRUP_LOCALS;
  if (crm.n() == 0) {
    // no dropped column: the stage-b2 sums over the dropped columns are empty, +0.0
    ops += (long long)tl.n() * crt.n() * crm.n() * (rbn * 2 + (has_Z ? (long long)Zt.rb * 8 : 0));
    double* zb[4] = { vtszn, jttan, htrkn, hhtrkn };
    for (double* z : zb) fill(z, (long long)M * T, 0.0, st);
  } else if (ww_kernel == 0) {
    // QC b1 + b2 fused on the matrix cores; the removed rows are packed first
    // the A terms run over the packed removed / retained rows (packed above), the Z term
    // over the controls' whole 128-bit groups
    const int Wr = rmv_W, Wt = rtn_W;
    AttynArgs x;
    const uint8_t* Arc = (const uint8_t*)At_rmvc;
    const uint8_t* Trc = (const uint8_t*)Tt_rmvc;
    const uint8_t* Atc = (const uint8_t*)At_rtnc;
    const uint8_t* Ttc = (const uint8_t*)Tt_rtnc;
    x.term[0] = AttynTerm{ Atc, (size_t)Wt * 4, Ttc, (size_t)Wt * 4, Atc, (size_t)Wt * 4, (const uint8_t*)d_cmask_rtn,
                           (long long)Wt * 32, nrt, attn, jttn, M };
    x.term[1] = AttynTerm{ Arc, (size_t)Wr * 4, Trc, (size_t)Wr * 4, Arc, (size_t)Wr * 4, (const uint8_t*)d_cmask_rmv,
                           (long long)Wr * 32, nrm, ktrn, htrn, M };
    const long long zbits = std::min((long long)Zt.rb * 8, (long long)(NZ + 127) / 128 * 128);
    x.term[2] = AttynTerm{ Zt.p, Zt.rb, St.p, St.rb, Zt.p, Zt.rb, (const uint8_t*)Zr_j, zbits, nz,
                           ztsn, vtsn, M };
    x.nterm = has_Z ? 3 : 2;
    x.klist = crt.d; x.nk = crt.n(); x.ylist = crm.d; x.ny = crm.n(); x.tlist = tl.d; x.nt = tl.n();
    x.ajdk = ajdk; x.pcols = pcols;
    x.vtszn = vtszn; x.jttan = jttan; x.htrkn = htrkn; x.hhtrkn = hhtrkn; x.ldo = M;
    size_t wb = attyn_mfma_work_bytes(x.ny, x.nt);
    Prof::Pending pa;
    prof.begin("attyn", 2.0 * crt.n() * crm.n() * tl.n() * ((double)nrt + nrm + (has_Z ? nz : 0)), st, pa);
    HIPCHK(attyn_mfma(x, g_attyn.need(dev, wb), wb, st));
    prof.end(pa, st);
    ops += (long long)tl.n() * crt.n() * crm.n() * (rbn * 2 + (has_Z ? (long long)Zt.rb * 8 : 0));
  } else {
  // QC b1: compact [(y*nt + t)*nk + k]
  size_t n3 = (size_t)crt.n() * crm.n() * tl.n() + 1;
  double* ztsvn = has_Z ? (double*)g_ztsvn.need(dev, sizeof(double) * n3) : nullptr;
  double* attjn = (double*)g_attjn.need(dev, sizeof(double) * n3);
  double* ktrhn = (double*)g_ktrhn.need(dev, sizeof(double) * n3);
  if (has_Z) attyn(Zt.b(), crt.d, crt.n(), Zr_j, nz, St.b(), tl.d, tl.n(), Zr_j, nz, Zt.b(), crm.d, crm.n(),
                   ztsn, vtsn, M, ajdk, pcols, ztsvn, st);
  attyn(At.b(), crt.d, crt.n(), Ar_rtn, nrt, Tt.b(), tl.d, tl.n(), Ar_rtn, nrt, At.b(), crm.d, crm.n(),
        attn, jttn, M, ajdk, pcols, attjn, st);
  attyn(At.b(), crt.d, crt.n(), Ar_rmv, nrm, Tt.b(), tl.d, tl.n(), Ar_rmv, nrm, At.b(), crm.d, crm.n(),
        ktrn, htrn, M, ajdk, pcols, ktrhn, st);
  ops += (long long)tl.n() * crt.n() * crm.n() * (rbn * 2 + (has_Z ? (long long)Zt.rb * 8 : 0));
  }
}

// bcc_lrup_QC_YnWt_stage_b2 (bcc_lrup.c:409-482): the ordered WtSZn sums of the triples
void Run::lrup_QC_b2()
{
  LRUP_LOCALS;
  // with the matrix-core kernels, and with no dropped column, stage b1 has formed the sums
  if (ww_kernel == 0 || crm.n() == 0) return;
  const double* ztsvn = (const double*)g_ztsvn.p;
  const double* attjn = (const double*)g_attjn.p;
  const double* ktrhn = (const double*)g_ktrhn.p;
  if (has_Z) attyn_wtszn(attjn, ztsvn, crt.d, crt.n(), crm.d, crm.n(), tl.d, tl.n(), ajdk, pcols, vtszn, M, st);
  attyn_wtszn(attjn, attjn, crt.d, crt.n(), crm.d, crm.n(), tl.d, tl.n(), ajdk, pcols, jttan, M, st);
  attyn_wtszn(attjn, ktrhn, crt.d, crt.n(), crm.d, crm.n(), tl.d, tl.n(), ajdk, pcols, htrkn, M, st);
  attyn_wtszn(ktrhn, ktrhn, crt.d, crt.n(), crm.d, crm.n(), tl.d, tl.n(), ajdk, pcols, hhtrkn, M, st);
}

// bcc_lrup_QC_YnWt_stage_c (bcc_lrup.c:592-655): L2_clean of the column scores, then the
// low-rank terms subtracted
void Run::lrup_QC_c()
{
  LRUP_LOCALS;
  StageC s;
  s.M = M; s.T = T; s.has_Z = has_Z; s.Ac_rtn = Ac_rtn; s.Ac_b = Ac_b; s.Tc_j = Tc_j;
  s.vtszn = vtszn; s.t_knzt = t_knzt; s.t_hnvt = t_hnvt; s.jttan = jttan; s.htrkn = htrkn; s.hhtrkn = hhtrkn;
  s.t_ankt = t_ankt; s.t_knkt = t_knkt; s.ZSZ = lf_AtTAnZtSZn; s.TAT = lf_AtTAnAtTAn;
  stage_c(s, st);
}

// the whole low-rank update after a drop (bcc_lrup.c:5-1138), in the driver's order
// (lakcluster_driver.c:185-203); stage c, which the driver calls after the re-quantisation
// of the Grams (bcc_M_YnWt), reads nothing that re-quantisation changes
void Run::lrup()
{
  lrup_QR_0(); lrup_QR_1(); lrup_QR_3();
  lrup_QC_a0(); lrup_QC_a1(); lrup_QC_a2(); lrup_QC_a3();
  lrup_QC_b0(); lrup_QC_b1(); lrup_QC_b2();
  lrup_QC_c();
}

bool Run::active() const { return !ar.h.empty() || !kl.h.empty(); }

// ------------------------------------------------------------------
// lakcluster_scorebox (lakcluster_scorebox.c:5-271; S_handle_make / S_init_bcc
// S_init.c:5-89; bcc_scorebox_mxA / _svalue / _srt and xcc_scorebox_xdrop_array
// bcc_scorebox.c:5-133, 282-308).  The elimination order of an out_xdrop file is
// replayed over a grid of (rows kept, columns kept).  The first grid point drops its
// rows and columns and recomputes every loop-subscore; the outer loop (rows, or columns
// when row_max > col_max) then follows by the low-rank update, and from each outer
// point the inner loop does the same over the other index, from a copy of the outer
// state.  Here the copy is a snapshot of the device arrays and host masks that the
// inner loop starts from and that is restored before the next outer step.
// ------------------------------------------------------------------
void scorebox_xdrop_array(int nxols, int x_max, int x_num, int* xval, int* xdrop)
{
  const double d = ((double)(x_max - 0)) / ((double)std::max(1, x_num - 1));
  for (int nx = 0; nx < x_num; nx++) xval[nx] = (int)round(x_max - nx * d);
  xdrop[0] = std::max(0, nxols - (x_max + 1));
  for (int nx = 1; nx < x_num; nx++) xdrop[nx] = xval[nx - 1] - xval[nx];
}

// bcc_scorebox_srt + bcc_scorebox_mxA: the current order is the last rpop (cpop) entries of
// the file's rows (columns); its first rdrop (cdrop) entries leave.  Then bcc_lrup_mxdup.
void Run::sb_drop(const std::vector<int>& rord, const std::vector<int>& cord, int rdrop, int cdrop)
{
  const int rpop = ar.n(), cpop = kl.n(), nro = (int)rord.size(), nco = (int)cord.size();
  const int nrd = std::min(rdrop, rpop), ncd = std::min(cdrop, cpop);
  const char* msg = "the scorebox out_xdrop order does not cover the active rows and columns";
  if ((nro < rpop && nrd > 0) || (nco < cpop && ncd > 0)) throw Error(LKC_ERR_CONFIG, msg);
  hAc_rtn = hAc_j; std::fill(hAc_rmv.begin(), hAc_rmv.end(), 0);
  for (int i = 0; i < ncd; i++) {
    const int k = cord[nco - cpop + i];
    if (k < 0 || k >= M || !bget(hAc_j.data(), k)) throw Error(LKC_ERR_CONFIG, msg);
    bset(hAc_rmv.data(), k); bclr(hAc_rtn.data(), k);
  }
  hAr_rtn = hAr_j; std::fill(hAr_rmv.begin(), hAr_rmv.end(), 0);
  for (int i = 0; i < nrd; i++) {
    const int a = rord[nro - rpop + i];
    if (a < 0 || a >= NA || !bget(hAr_j.data(), a)) throw Error(LKC_ERR_CONFIG, msg);
    bset(hAr_rmv.data(), a); bclr(hAr_rtn.data(), a);
  }
  hAr_j = hAr_rtn; hAc_j = hAc_rtn;
  push_ctl();
}

// lakcluster_scorebox_excerpt_0 with bcc_scorebox_svalue: the active counts, Irem and the
// means of the first covariate column's row and column scores, in index order
void Run::sb_record(Scorebox& S, int tab)
{
  score_lists();
  const int nk = kl.n(), na = ar.n();
  if (nk) HIPCHK(hipMemcpyAsync(hsv, dsv, sizeof(double) * nk, hipMemcpyDeviceToHost, st));
  if (na) HIPCHK(hipMemcpyAsync(hsv + M, dsv + M, sizeof(double) * na, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  prof.collect();
  double m = 0;
  for (int i = 0; i < na; i++) m += hsv[M + i];
  m /= (double)na; if (!std::isfinite(m)) m = 1.0;
  S.qr[tab] = m;
  m = 0;
  for (int j = 0; j < nk; j++) m += hsv[j];
  m /= (double)nk; if (!std::isfinite(m)) m = 1.0;
  S.qc[tab] = m;
  S.rpop[tab] = na; S.cpop[tab] = nk; S.irem[tab] = na > 0 ? 1 : 0;
}

void Run::scorebox(Scorebox& S)
{
  if (inited || n_iter) throw Error(LKC_ERR_STATE, "the scorebox starts from a freshly loaded handle");
  S.row_max = c.sb_row_max; S.row_num = c.sb_row_num; S.col_max = c.sb_col_max; S.col_num = c.sb_col_num;
  const int row_num = S.row_num, col_num = S.col_num;
  if (row_num < 1 || col_num < 1) throw Error(LKC_ERR_CONFIG, "GLOBAL_scorebox_row_num and _col_num must be >= 1");
  std::vector<int> rord, cord;
  read_out_xdrop(c.sb_file, "scorebox out_xdrop", rord, cord);
  std::vector<int> rkeep(row_num), rdrop(row_num), ckeep(col_num), cdrop(col_num);
  scorebox_xdrop_array(kl.n(), S.col_max, col_num, ckeep.data(), cdrop.data());
  scorebox_xdrop_array(ar.n(), S.row_max, row_num, rkeep.data(), rdrop.data());
  const size_t ng = (size_t)row_num * col_num;
  S.rpop.assign(ng, 0); S.cpop.assign(ng, 0); S.irem.assign(ng, 0); S.qr.assign(ng, 0.0); S.qc.assign(ng, 0.0);
  const bool rows_outer = !(S.row_max > S.col_max);     // lakcluster_scorebox_cr, else _rc
  const int n_out = rows_outer ? row_num : col_num, n_in = rows_outer ? col_num : row_num;
  if (dcc) { scorebox_dcc(S, rord, cord, rdrop, cdrop, rows_outer); return; }
  // everything an update reads or accumulates into
  size_t nA = NA, nZ = std::max(NZ, 1), nM = M, nT = T;
  std::vector<std::pair<double*, size_t>> keep = {
    { mlf, 16 * 8 }, { lf_An_ajdk, nA * kAjdk }, { lf_Zn_ajdk, nZ * kAjdk }, { jn_ajdk, nA * kAjdk },
    { vn_ajdk, nZ * kAjdk }, { kn_ajdk, nA * kAjdk }, { hn_ajdk, nA * kAjdk }, { an_ajdk_, nA * kAjdk },
    { zn_ajdk, nZ * kAjdk }, { lf_AtTn, nM * nT }, { lf_ZtSn, nM * nT }, { lf_AnZt, nA * nZ }, { lf_AnAt, nA * nA },
    { lf_AtTAnZtSZn, nM * nT }, { lf_AtTAnAtTAn, nM * nT }, { lf_AnZtSZnAt, nA * nT }, { lf_AnAtTAnAt, nA * nT },
    { At_An_a1d1, nM }, { et_An, nM }, { An_a2d2_ZtSn, nA * nT }, { An_a2d2_AtTn, nA * nT }, { sc, 6 * nT + 2 },
    { T_AnAt_AnAt, nT * nA }, { QR_ZSZ, nA * nT }, { QR_TAT, nA * nT }, { QC_ZSZ, nM * nT }, { QC_TAT, nM * nT } };
  for (int i = 0; i < 14; i++) keep.push_back({ mt[i], nM * nT });
  std::vector<double*> snap;
  std::vector<std::vector<uint8_t>> hsnap;
  auto save = [&]() {
    if (snap.empty()) for (auto& e : keep) snap.push_back(dev.alloc<double>(e.second));
    for (size_t i = 0; i < keep.size(); i++)
      HIPCHK(hipMemcpyAsync(snap[i], keep[i].first, sizeof(double) * keep[i].second, hipMemcpyDeviceToDevice, st));
    hsnap = { hAr_j, hAr_rtn, hAr_rmv, hAc_j, hAc_rtn, hAc_rmv };
  };
  auto restore = [&]() {
    for (size_t i = 0; i < keep.size(); i++)
      HIPCHK(hipMemcpyAsync(keep[i].first, snap[i], sizeof(double) * keep[i].second, hipMemcpyDeviceToDevice, st));
    hAr_j = hsnap[0]; hAr_rtn = hsnap[1]; hAr_rmv = hsnap[2]; hAc_j = hsnap[3]; hAc_rtn = hsnap[4]; hAc_rmv = hsnap[5];
    push_ctl();
  };
  auto update = [&]() {
    lrup();
    st_QR(); st_ajdk(); st_ZtSn(); st_single(); st_double(); st_flatten();
  };
  double t0 = now_s();
  for (int o = 0; o < n_out; o++) {
    if (o == 0) {
      sb_drop(rord, cord, rdrop[0], cdrop[0]);
      st_ajdk(); st_ZtSn(); st_YnWt(); st_At_T(); st_QR();
      st_single(); st_double(); st_flatten();
      init_s = now_s() - t0;
    } else {
      sb_drop(rord, cord, rows_outer ? rdrop[o] : 0, rows_outer ? 0 : cdrop[o]);
      update();
    }
    sb_record(S, rows_outer ? o : o * row_num);
    if (n_in > 1) save();
    for (int i = 1; i < n_in; i++) {
      sb_drop(rord, cord, rows_outer ? 0 : rdrop[i], rows_outer ? cdrop[i] : 0);
      update();
      sb_record(S, rows_outer ? o + i * row_num : i + o * row_num);
    }
    if (n_in > 1 && o + 1 < n_out) restore();
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  prof.collect();
  for (double* p : snap) dev.release(p);
  inited = true;
  loop_s = now_s() - t0 - init_s;
}

// dexcluster_scorebox_rc / _cr (dexcluster_scorebox.c:67-165).  The dcc subscores are
// recomputed from the masks at every grid point, so a problem state is its masks: D_ori
// along the outer loop, D_sub along the inner one.  D_sub starts from the loaded masks and is
// never copied from D_ori, so it keeps its own drops from one inner loop to the next and never
// takes the outer loop's.  The first point runs dexcluster_scorebox_excerpt_1, which computes
// before its drop takes effect (no dcc_lrup_mxdup): that drop is lost.
void Run::scorebox_dcc(Scorebox& S, const std::vector<int>& rord, const std::vector<int>& cord,
                       const std::vector<int>& rdrop, const std::vector<int>& cdrop, bool rows_outer)
{
  const int row_num = S.row_num, col_num = S.col_num;
  const int n_out = rows_outer ? row_num : col_num, n_in = rows_outer ? col_num : row_num;
  struct Masks { std::vector<uint8_t> ar, ac; };
  auto get = [&]() { return Masks{ hAr_j, hAc_j }; };
  auto set = [&](const Masks& m) { hAr_j = m.ar; hAc_j = m.ac; push_ctl(); };
  auto compute = [&]() { st_ajdk(); st_ZtSn(); st_dcc(); };
  const Masks loaded = get();
  Masks ori = loaded, sub = loaded;
  double t0 = now_s();
  for (int o = 0; o < n_out; o++) {
    set(ori);
    if (o == 0) {
      sb_drop(rord, cord, rdrop[0], cdrop[0]);   // checks the order; excerpt_1 drops nothing
      set(ori);
    } else {
      sb_drop(rord, cord, rows_outer ? rdrop[o] : 0, rows_outer ? 0 : cdrop[o]);
    }
    compute();
    if (o == 0) init_s = now_s() - t0;
    sb_record(S, rows_outer ? o : o * row_num);
    ori = get();
    set(sub);
    for (int i = 1; i < n_in; i++) {
      sb_drop(rord, cord, rows_outer ? 0 : rdrop[i], rows_outer ? cdrop[i] : 0);
      compute();
      sb_record(S, rows_outer ? o + i * row_num : i + o * row_num);
    }
    sub = get();
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  prof.collect();
  inited = true;
  loop_s = now_s() - t0 - init_s;
}

// S_handle_dmp (S_init.c:79-89) with mda_write_d2_i4 / _r8 (mda_io.c:5-31): an int32 rank
// 2, the dimensions row_num and col_num, then the entries [nr + nc*row_num]
void write_scorebox(const Scorebox& S, const std::string& dir)
{
  mkdir(dir.c_str(), 0755);
  char pre[160];
  snprintf(pre, sizeof pre, "scorebox_r%dn%d_c%dn%d", S.row_max, S.row_num, S.col_max, S.col_num);
  auto put = [&](const char* nm, const void* data, size_t el) {
    std::string p = dir + "/" + pre + "_" + nm + ".mda";
    FILE* fp = fopen(p.c_str(), "wb");
    if (!fp) throw Error(LKC_ERR_IO, "cannot write " + p);
    int hdr[3] = { 2, S.row_num, S.col_num };
    const size_t n = (size_t)S.row_num * S.col_num;
    bool ok = fwrite(hdr, sizeof(int), 3, fp) == 3 && fwrite(data, el, n, fp) == n;
    ok = (fclose(fp) == 0) && ok;
    if (!ok) throw Error(LKC_ERR_IO, "cannot write " + p);
  };
  put("A_rpop_j_total", S.rpop.data(), sizeof(int));
  put("A_cpop_j", S.cpop.data(), sizeof(int));
  put("Irem", S.irem.data(), sizeof(int));
  put("QR_avg", S.qr.data(), sizeof(double));
  put("QC_avg", S.qc.data(), sizeof(double));
}

void Run::iterate(int* done)
{
  if (!inited) throw Error(LKC_ERR_STATE, "call init first");
  if (!active()) { if (done) *done = 1; return; }
  double t0 = now_s();
  sumscores(true);
  if (dcc) { st_ajdk(); st_ZtSn(); st_dcc(); }     // dcc: recomputed, no low-rank update
  else {
    lrup();
    st_QR(); st_ajdk(); st_ZtSn(); st_single(); st_double(); st_flatten();
  }
  finish_trace();
  HIPCHK(hipGetLastError());
  if (done) *done = active() ? 0 : 1;
  if (!active()) HIPCHK(hipStreamSynchronize(st));
  loop_s += now_s() - t0;
}

void Run::run(int max_iter)
{
  if (!inited) init_lf();
  int done = active() ? 0 : 1;
  double t0 = now_s();
  loop_s = 0;
  while (!done && (max_iter < 0 || n_iter < max_iter)) {
    sumscores(true);
    if (dcc) { st_ajdk(); st_ZtSn(); st_dcc(); }
    else {
      lrup();
      st_QR(); st_ajdk(); st_ZtSn(); st_single(); st_double(); st_flatten();
    }
    finish_trace();
    done = active() ? 0 : 1;
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  prof.collect();
  loop_s = now_s() - t0;
}

// back to the state right after loading: all rows and columns active, no outputs
void Run::reset()
{
  HIPCHK(hipStreamSynchronize(st));
  prof.collect();
  hAr_j = hAr_j0; hZr_j = hZr_j0; hAc_j = hAc_j0;
  std::fill(hAr_rtn.begin(), hAr_rtn.end(), 0); std::fill(hAr_rmv.begin(), hAr_rmv.end(), 0);
  std::fill(hAc_rtn.begin(), hAc_rtn.end(), 0); std::fill(hAc_rmv.begin(), hAc_rmv.end(), 0);
  push_ctl();
  HIPCHK(hipStreamSynchronize(st));
  trace.clear(); xdrop_a.clear(); xdrop_b.clear(); trace_pending = false;
  n_iter = 0; inited = false; ops = 0; init_s = 0; loop_s = 0; n_sort_fallback = 0;
  xij_done = false; pending_dup = false; a0_valid = false; b0_valid = false; vn_epoch = -1;
  for (auto& k : prof.stats) { k.launches = 0; k.ms = 0; k.work = 0; }
}

void Run::write_outputs(const std::string& dir) const
{
  mkdir(dir.c_str(), 0755);
  std::string p = dir + "/out_trace.txt";
  FILE* fp = fopen(p.c_str(), "w");
  if (!fp) throw Error(LKC_ERR_IO, "cannot write " + p);
  for (size_t l = 0; l + 5 < trace.size(); l += 6) {
    const double* t = &trace[l];
    fprintf(fp, "%d %d %d %0.16f %0.16f %d\n", (int)t[0], (int)t[1], (int)t[2], t[3], t[4], (int)t[5]);
  }
  fclose(fp);
  std::string pa = dir + "/out_xdrop_a.txt", pb = dir + "/out_xdrop_b.txt";
  FILE* fa = fopen(pa.c_str(), "w");
  if (!fa) throw Error(LKC_ERR_IO, "cannot write " + pa);
  FILE* fb = fopen(pb.c_str(), "w");
  if (!fb) { fclose(fa); throw Error(LKC_ERR_IO, "cannot write " + pb); }
  for (size_t l = 0; l + 1 < xdrop_a.size(); l += 2) {
    fprintf(fa, "%d %d\n", xdrop_a[l], xdrop_a[l + 1]);
    fprintf(fb, "%d %d\n", xdrop_b[l], xdrop_b[l + 1]);
  }
  fclose(fa); fclose(fb);
}

// copy of a named state array in the layout of the CPU restatement ([row*T + t], [a*NZ + z], [r*15 + jk])
long Run::debug_get(const std::string& name, std::vector<double>& out)
{
  HIPCHK(hipStreamSynchronize(st));
  auto fetch = [&](const double* p, size_t n) {
    std::vector<double> h(n);
    if (n) HIPCHK(hipMemcpyAsync(h.data(), p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
  };
  auto transpose = [&](const double* p, size_t rows, size_t cols) {   // device [c*rows + r] -> [r*cols + c]
    std::vector<double> h = fetch(p, rows * cols);
    out.assign(rows * cols, 0.0);
    for (size_t r = 0; r < rows; r++) for (size_t c = 0; c < cols; c++) out[r * cols + c] = h[c * rows + r];
  };
  size_t nA = NA, nZ = NZ, nM = M, nT = T;
  struct { const char* nm; const double* p; size_t rows, cols; } mt_tab[] = {
    { "lf_An_ajdk", lf_An_ajdk, nA, (size_t)kAjdk }, { "lf_Zn_ajdk", lf_Zn_ajdk, nZ, (size_t)kAjdk },
    { "lf_AtTn", lf_AtTn, nM, nT }, { "lf_ZtSn", lf_ZtSn, nM, nT },
    { "lf_AnZt", lf_AnZt, nA, nZ }, { "lf_AnAt", lf_AnAt, nA, nA },
    { "lf_AtTAnZtSZn", lf_AtTAnZtSZn, nM, nT }, { "lf_AtTAnAtTAn", lf_AtTAnAtTAn, nM, nT },
    { "lf_AnZtSZnAt", lf_AnZtSZnAt, nA, nT }, { "lf_AnAtTAnAt", lf_AnAtTAnAt, nA, nT },
    { "At_An_a1d1", At_An_a1d1, nM, 1 }, { "An_a2d2_ZtSn", An_a2d2_ZtSn, nA, nT }, { "An_a2d2_AtTn", An_a2d2_AtTn, nA, nT },
    { "et_An", et_An, nM, 1 }, { "T_AnAt_AnAt", T_AnAt_AnAt, nT * nA, 1 },
    { "QR_ZSZ", QR_ZSZ, nA, nT }, { "QR_TAT", QR_TAT, nA, nT }, { "QC_ZSZ", QC_ZSZ, nM, nT }, { "QC_TAT", QC_TAT, nM, nT },
  };
  if (name == "lf_AnZt") { out = fetch(lf_AnZt, nA * nZ); return (long)out.size(); }   // [a*NZ + z] on the device
  for (auto& e : mt_tab) if (name == e.nm) { transpose(e.p, e.rows, e.cols); return (long)out.size(); }
  const char* scn[] = { "a1d2_AtTn", "a3d2_AtTn", "a1d2_ZtSn", "a3d2_ZtSn", "et_Tn", "et_Sn" };
  std::vector<double> h = fetch(sc, 6 * nT + 1);
  for (int i = 0; i < 6; i++) if (name == scn[i]) { out.assign(h.begin() + i * nT, h.begin() + (i + 1) * nT); return (long)nT; }
  if (name == "et_An_a1d1") { out.assign(1, h[6 * nT]); return 1; }
  return -1;
}

std::string out_dir_of(const Cfg& c)
{
  if (!c.dir_name.empty()) return c.dir_name;
  if (!c.out_name.empty()) return (c.dir_base.empty() ? std::string(".") : c.dir_base) + "/dir_" + c.out_name;
  return ".";
}
std::string Run::default_out_dir() const { return out_dir_of(c); }

}  // namespace lkc
