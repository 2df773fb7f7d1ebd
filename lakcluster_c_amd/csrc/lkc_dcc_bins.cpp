// lkc_dcc_bins.cpp -- the dcc elimination with several bins and Y/W genetic controls on the
// device: dcc_sumscores_test (dcc_sumscores.c:700-747) on the synthetic problem of dcc_init
// (dcc.c:1217-1275), and dexcluster_driver (dexcluster_driver.c:5-57) on the problem that
// dcc_load (dcc.c:1157-1215) reads from the .in's per-bin files.  Both run the dcc half-loop subscores of every bin pair (dcc_An_ajdk, dcc_lf_ZtSn,
// dcc_lf_D_AtTn_ZtSn_vv, dcc_lf_TAnZtS_ww, dcc_wrap_dcc_halfloop; dcc_lf.c, TAnZtS.c,
// D_AtTn_ZtSn.c), the score stages that sum the bins (dcc_sumscores.c:5-699) and the three
// output files.
//
// Every double reduction keeps the reference's order; the integer contractions run on
// the kernels of the one-bin path (An_ajdk_v, plain +-1 products, xcalc, the int8
// An_ZtSn_ww), the per-pair terms and the bin sums on lkc_dcc_pairs.hip.  The sort is the
// reference's quicksort on the host, as in the reference.
#include "lkc_dcc_bins.h"
#include "lkc_engine.h"
#include "lkc_stages.h"
#include "lkc_testgen.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <vector>

namespace lkc {

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
  throw Error(LKC_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ")"); } while (0)

namespace {

inline int bg(const uint8_t* b, long i) { return (b[i >> 3] >> (7 - (i & 7))) & 1; }
inline void bs(uint8_t* b, long i) { b[i >> 3] |= (uint8_t)(1u << (7 - (i & 7))); }
inline void bc(uint8_t* b, long i) { b[i >> 3] &= (uint8_t)~(1u << (7 - (i & 7))); }
int popn(const std::vector<uint8_t>& m, int n) { int s = 0; for (int i = 0; i < n; i++) s += bg(m.data(), i); return s; }

struct HBits { std::vector<uint8_t> b; int nr = 0, nc = 0; size_t rb = 0;
  void init(int r, int c) { nr = r; nc = c; rb = bsize(c); b.assign((size_t)std::max(r, 1) * rb, 0); } };

enum { YW = 0, AZ = 1, YY = 2, AA = 3 };   // TYnWtS, TAnZtS, TYnYtT, TAnAtT

struct Bin {
  int nA = 0, nZ = 0;
  std::vector<uint8_t> hAr_b, hAr_j, hAr_rtn, hZr_b, hZr_j;
  HBits hAn, hAt, hZn, hZt, hYn, hYt, hWn, hWt, hTt, hSt;
  DMat An, At, Zn, Zt, Yn, Yt, Wn, Wt, Tt, St;
  uint32_t *Ar_j = nullptr, *Zr_j = nullptr;
  IList ar, zr;
  double *mlfAr = nullptr, *mlfZr = nullptr;                 // mask_lf of the active rows (count at [15])
  double *lfA = nullptr, *lfY = nullptr, *lfZ = nullptr, *lfW = nullptr;   // [jk*rows + r]
  double *AtTn = nullptr, *YtTn = nullptr, *ZtSn = nullptr, *WtSn = nullptr;   // [t*cols + k]
  double* sc = nullptr;   // a1d1_AtTn, a1d1_ZtSn, a1d1_YtTn, a1d1_WtSn, et_Tn, et_Sn: [T] each
  double *etA = nullptr, *etY = nullptr;   // -|rows| + 2 pop per column over the active rows
  uint32_t* cz[4] = {}; double* scal[4] = {};                // quantised a0d1 products, by term
  double* QRn[4] = {};                                       // [nb2*T*nA + t*nA + r]
  std::vector<int> g_a, g_b;                                 // global row indices (dcc_init_QX)
};
struct Pair { double *x[4] = {}, *qr[4] = {}, *qc[4] = {}; unsigned long long* pg[4] = {}; };

struct DccDev {
  Cfg c;
  hipStream_t st = nullptr;
  Dev dev;
  int nbins = 0, M = 0, NY = 0, T = 0, B = 32, Apc = 0, Ypc = 0;
  std::vector<uint8_t> hAc_b, hAc_j, hAc_rtn, hYc_b, hYc_j, hTc_b, hTc_j;
  std::vector<double> hA_p, hY_p, hA_ajdk, hY_ajdk;
  int tb[kMaxT] = {};
  std::vector<Bin> E; std::vector<Pair> F;
  uint32_t *Ac_j = nullptr, *Yc_j = nullptr;
  IList kl, yl, tl;
  double *A_ajdk = nullptr, *Y_ajdk = nullptr, *mlfA = nullptr, *mlfY = nullptr;
  double* QCn[4] = {};                                       // [nbx*T*X + t*X + k]
  double *L = nullptr, *xwork = nullptr, *dsv = nullptr;
  int* cnt = nullptr;
  std::vector<FillJob> pass_fills;                           // zeroed at the start of every pass
  std::vector<void*> prep;                                   // digit planes per (nb2, term)
  size_t prep_bytes = 0;
  int A_nrows_total = 0, A_rpop_b_total = 0, Z_rpop_b_total = 0, A_rpop_j_total = 0, Z_rpop_j_total = 0, A_cpop_j = 0, Irem = 0;
  std::vector<double> QC_sv, QR_sv; std::vector<int> QC_perm, QR_perm, QR_bin, QR_row;
  DccTestOut* out = nullptr;
  long long ops = 0;

  ~DccDev() { dev.release_all(); if (st) (void)hipStreamDestroy(st); }
  int X(int q) const { return (q == YW || q == YY) ? NY : M; }
  bool on(int q, int nb1, int nb2) const   // the `bother` flags of a term (dcc_lf.c, TAnZtS.c)
  {
    const Bin &e1 = E[nb1], &e2 = E[nb2];
    if (e1.nA <= 0) return false;
    switch (q) { case YW: return NY > 0 && e2.nZ > 0; case AZ: return M > 0 && e2.nZ > 0;
                 case YY: return NY > 0 && e2.nA > 0; default: return M > 0 && e2.nA > 0; }
  }

  bool files = false;   // the problem came from the .in's files: A_p / Y_p from the data

  void generate();
  void load();
  void finish_host();
  void setup();
  void data_freqs();
  void upload_list(IList& l, const std::vector<uint8_t>& m, int n);
  void mxset();
  void subscores();
  void nrm(); void ifT(); void cmb(); void xij(); void mxA(); void mxdup();
  void write(const std::string& dir) const;
};

// dcc_init: dcc_ajdk_init's column masks (dcc.c:117-174), per bin dcc_single_init_M_An's row
// masks and matrices (dcc.c:543-619: wrap_M_setup_test_excerpt_0, "error", An_ajdk_v.c:33-52,
// column-major draws, bit = draw > 1/2; the covariates' column 0 set in every row), then
// dcc_init_A_p's block frequencies (dcc.c:993-1011)
void DccDev::generate()
{
  GlibcRand g;
  const double mrnd = c.test_mrand;
  auto mask_draw = [&](int n, std::vector<uint8_t>& mb, std::vector<uint8_t>& mj) {
    mb.assign(bsize(n) + 8, 0); mj.assign(bsize(n) + 8, 0);
    for (int i = 0; i < n; i++) { double d = g.r01(); if (d > mrnd * mrnd) bs(mb.data(), i); if (d > mrnd) bs(mj.data(), i); }
  };
  auto bits_draw = [&](int nrows, int ncols, HBits& Xn, HBits& Xt) {
    Xn.init(nrows, ncols); Xt.init(ncols, nrows);
    const double half = 0.5;
    for (int nc = 0; nc < ncols; nc++) for (int nr = 0; nr < nrows; nr++)
      if (g.r01() > half) { bs(Xn.b.data() + (size_t)nr * Xn.rb, nc); bs(Xt.b.data() + (size_t)nc * Xt.rb, nr); }
  };
  mask_draw(M, hAc_b, hAc_j);
  mask_draw(NY, hYc_b, hYc_j);
  mask_draw(T, hTc_b, hTc_j);
  bs(hTc_b.data(), 0); bs(hTc_j.data(), 0); bs(hTc_b.data(), T - 1); bs(hTc_j.data(), T - 1);
  E.assign(nbins, Bin());
  for (int nb = 0; nb < nbins; nb++) {
    Bin& e = E[nb];
    HBits Tn, Sn;
    e.nA = std::max(0, c.test_A_rows + (nb % 3) - 1);
    mask_draw(e.nA, e.hAr_b, e.hAr_j);
    bits_draw(e.nA, M, e.hAn, e.hAt);
    e.nZ = std::max(0, c.test_Z_rows + (nb % 5) - 2);
    mask_draw(e.nZ, e.hZr_b, e.hZr_j);
    bits_draw(e.nZ, M, e.hZn, e.hZt);
    bits_draw(e.nA, NY, e.hYn, e.hYt);
    bits_draw(e.nZ, NY, e.hWn, e.hWt);
    bits_draw(e.nA, T, Tn, e.hTt);
    for (int r = 0; r < e.nA; r++) bs(e.hTt.b.data(), r);
    bits_draw(e.nZ, T, Sn, e.hSt);
    for (int r = 0; r < e.nZ; r++) bs(e.hSt.b.data(), r);
  }
  Apc = psize(M); Ypc = psize(NY);
  hA_p.resize(Apc); hY_p.resize(Ypc);
  for (int i = 0; i < Apc; i++) hA_p[i] = crop_draw(g);
  if (c.sparse == 0) std::fill(hA_p.begin(), hA_p.end(), 0.5);
  for (int i = 0; i < Ypc; i++) hY_p[i] = crop_draw(g);
  if (c.sparse == 0) std::fill(hY_p.begin(), hY_p.end(), 0.5);
  finish_host();
}

// calc_A_ajdk (An_ajdk_v.c:5-30), libm pow on the host
static void ajdk_of(const std::vector<double>& p, std::vector<double>& out)
{
  const int pc = (int)p.size(); out.assign((size_t)kAjdk * std::max(pc, 1), 0.0);
  for (int j = 0; j < 5; j++) for (int k = 0; k < 3; k++) for (int n = 0; n < pc; n++)
    out[(size_t)(j + 5 * k) * pc + n] = pow(p[n] - (1 - p[n]), j) * pow(1.0 / std::max(0.01, (4.0 * p[n] * (1 - p[n]))), k);
}

// dcc_load (dcc.c:1157-1215): dcc_ajdk_load's column masks (dcc.c:54-115), per bin
// dcc_single_load_M_An (dcc.c:251-419): A from A_t (patients) and A_n (SNPs), Z, Y, W, T
// and S likewise, each pair optional as there (T and S default to the intercept column),
// the row masks from the A_n_rind_ / Z_n_rind_ lists.  Every bin names the same kinds of
// file.  The covariates in use are the T_n_cind mask, else the first GLOBAL_T_n_cols of
// them (all when 0), as in the one-bin loader.  A_p and Y_p follow from the data once it
// is on the device (data_freqs).
void DccDev::load()
{
  files = true;
  nbins = c.nbins;
  auto at = [](const std::vector<std::string>& v, int nb) { return nb < (int)v.size() ? v[nb] : std::string(); };
  for (int nb = 0; nb < nbins; nb++) {
    if (at(c.A_t_, nb).empty() || at(c.A_n_, nb).empty())
      throw Error(LKC_ERR_CONFIG, "bin " + std::to_string(nb) + ": GLOBAL_A_n_name_ and GLOBAL_A_t_name_ must name its files");
    if (at(c.Y_t_, nb).empty() != at(c.Y_t_, 0).empty() || at(c.T_t_, nb).empty() != at(c.T_t_, 0).empty() ||
        at(c.Z_t_, nb).empty() != at(c.Z_n_, nb).empty())
      throw Error(LKC_ERR_CONFIG, "bin " + std::to_string(nb) + " names a different set of files");
  }
  int fnr = 0, fnc = 0;
  b16_size(c.A_t_[0], &fnr, &fnc); M = fnr;
  NY = 0; if (!at(c.Y_t_, 0).empty()) { b16_size(c.Y_t_[0], &fnr, &fnc); NY = fnr; }
  T = 1; if (!at(c.T_t_, 0).empty()) { b16_size(c.T_t_[0], &fnr, &fnc); T = fnr; }
  if ((c.A_ncols && c.A_ncols != M) || (c.Y_ncols && c.Y_ncols != NY))
    throw Error(LKC_ERR_SHAPE, "GLOBAL_A_n_cols / GLOBAL_Y_n_cols do not match the files");
  if (T < 1 || T > kMaxT) throw Error(LKC_ERR_CONFIG, "between 1 and 8 covariate columns are supported");
  auto all = [](int n) { std::vector<uint8_t> m(bsize(n) + 8, 0); for (int i = 0; i < n; i++) bs(m.data(), i); return m; };
  auto mask = [&](const std::string& f, int n) {
    if (f.empty()) return all(n);
    int r = 0, k = 0; b16_size(f, &r, &k);
    if (r != n || k != 1) throw Error(LKC_ERR_SHAPE, "mask " + f + " has " + std::to_string(r) + " entries, expected " + std::to_string(n));
    return read_mask(f, n);
  };
  hAc_b = mask(c.A_cind, M); hYc_b = mask(c.Y_cind, NY);
  if (!c.T_cind.empty()) hTc_b = mask(c.T_cind, T);
  else { hTc_b.assign(bsize(T) + 8, 0); const int nt = c.T_ncols > 0 ? std::min(c.T_ncols, T) : T; for (int t = 0; t < nt; t++) bs(hTc_b.data(), t); }
  hAc_j = hAc_b; hYc_j = hYc_b; hTc_j = hTc_b;
  auto take = [](HBits& d, HostBits&& h) { d.nr = h.nr; d.nc = h.nc; d.rb = h.rb; d.b.swap(h.b); };
  // X (rows x cols) from its patient-major file pt and SNP-major file pn
  auto pair = [&](const std::string& pt, const std::string& pn, int rows, int cols, HBits& Xn, HBits& Xt) {
    take(Xn, read_b16(pt)); take(Xt, read_b16(pn));
    if (Xn.nr != rows || Xn.nc != cols || Xt.nr != cols || Xt.nc != rows)
      throw Error(LKC_ERR_SHAPE, pt + " / " + pn + " do not hold a " + std::to_string(rows) + " x " + std::to_string(cols) + " matrix");
  };
  auto intercept = [](int rows, HBits& Xt) { Xt.init(1, rows); for (int r = 0; r < rows; r++) bs(Xt.b.data(), r); };
  E.assign(nbins, Bin());
  for (int nb = 0; nb < nbins; nb++) {
    Bin& e = E[nb];
    HBits Tn, Sn;
    b16_size(c.A_t_[nb], &fnr, &fnc); e.nA = fnc;
    pair(c.A_t_[nb], c.A_n_[nb], e.nA, M, e.hAn, e.hAt);
    e.nZ = 0;
    if (!at(c.Z_t_, nb).empty()) { b16_size(c.Z_t_[nb], &fnr, &fnc); e.nZ = fnc; }
    if (e.nZ > 0) pair(c.Z_t_[nb], c.Z_n_[nb], e.nZ, M, e.hZn, e.hZt);
    else { e.hZn.init(0, M); e.hZt.init(M, 0); }
    if (NY > 0) pair(c.Y_t_[nb], at(c.Y_n_, nb), e.nA, NY, e.hYn, e.hYt);
    else { e.hYn.init(e.nA, 0); e.hYt.init(0, e.nA); }
    if (NY > 0 && e.nZ > 0) {
      if (at(c.W_t_, nb).empty() || at(c.W_n_, nb).empty()) throw Error(LKC_ERR_CONFIG, "bin " + std::to_string(nb) + ": Y without W");
      pair(c.W_t_[nb], c.W_n_[nb], e.nZ, NY, e.hWn, e.hWt);
    } else { e.hWn.init(e.nZ, NY); e.hWt.init(NY, e.nZ); }
    if (!at(c.T_t_, nb).empty()) pair(c.T_t_[nb], at(c.T_n_, nb), e.nA, T, Tn, e.hTt);
    else intercept(e.nA, e.hTt);
    if (e.nZ > 0 && !at(c.S_t_, nb).empty()) pair(c.S_t_[nb], at(c.S_n_, nb), e.nZ, T, Sn, e.hSt);
    else if (T == 1 || e.nZ == 0) intercept(e.nZ, e.hSt);
    else throw Error(LKC_ERR_CONFIG, "bin " + std::to_string(nb) + ": T without S");
    e.hAr_b = mask(at(c.A_rind_, nb), e.nA); e.hAr_j = e.hAr_b;
    e.hZr_b = mask(e.nZ > 0 ? at(c.Z_rind_, nb) : std::string(), e.nZ); e.hZr_j = e.hZr_b;
  }
  Apc = psize(M); Ypc = psize(NY);
  hA_p.assign(Apc, 0.5); hY_p.assign(Ypc, 0.5);
  finish_host();
}

// dcc_X_nrows_total and dcc_init_QX (dcc.c:1115-1131, 1047-1090): row indices over the
// bins, the totals, the covariates' compacted indices
void DccDev::finish_host()
{
  ajdk_of(hA_p, hA_ajdk); ajdk_of(hY_p, hY_ajdk);
  int mx_a = 0, mx_b = 0;
  for (Bin& e : E) {
    e.g_a.assign(e.nA, 0); e.g_b.assign(e.nA, 0);
    for (int r = 0; r < e.nA; r++) { e.g_a[r] = mx_a++; if (bg(e.hAr_b.data(), r)) e.g_b[r] = mx_b++; }
    A_nrows_total += e.nA; A_rpop_b_total += popn(e.hAr_b, e.nA); Z_rpop_b_total += popn(e.hZr_b, e.nZ);
  }
  for (int t = 0, b = 0; t < T; t++) { tb[t] = b; if (bg(hTc_b.data(), t)) b++; }
}

void DccDev::upload_list(IList& l, const std::vector<uint8_t>& m, int n)
{
  l.h.clear();
  for (int i = 0; i < n; i++) if (bg(m.data(), i)) l.h.push_back(i);
  if (!l.h.empty()) HIPCHK(hipMemcpyAsync(l.d, l.h.data(), sizeof(int) * l.h.size(), hipMemcpyHostToDevice, st));
}

void DccDev::setup()
{
  HIPCHK(hipSetDevice(0));
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  dev.st = st;
  auto up = [&](DMat& d, const HBits& h) {
    d.nr = h.nr; d.nc = h.nc; d.rb = h.rb; d.p = dev.alloc<uint8_t>(h.b.size());
    HIPCHK(hipMemcpyAsync(d.p, h.b.data(), h.b.size(), hipMemcpyHostToDevice, st));
  };
  auto dbl = [&](size_t n) { return dev.alloc<double>(std::max<size_t>(n, 1)); };
  auto upd = [&](const std::vector<double>& h) {
    double* p = dbl(h.size());
    if (!h.empty()) HIPCHK(hipMemcpyAsync(p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
    return p;
  };
  auto mask = [&](int n) { return dev.alloc<uint32_t>((bsize(n) + 8) / 4); };
  A_ajdk = upd(hA_ajdk); Y_ajdk = upd(hY_ajdk);
  Ac_j = mask(M); Yc_j = mask(NY);
  kl.d = dev.alloc<int>(M + 1); yl.d = dev.alloc<int>(NY + 1); tl.d = dev.alloc<int>(T + 1);
  HIPCHK(hipMemcpyAsync(Yc_j, hYc_j.data(), bsize(NY), hipMemcpyHostToDevice, st));
  upload_list(yl, hYc_j, NY); upload_list(tl, hTc_j, T);
  mlfA = dbl(16); mlfY = dbl(16);
  mask_lf(Yc_j, Ypc, Y_ajdk, mlfY, st);
  for (Bin& e : E) {
    up(e.An, e.hAn); up(e.At, e.hAt); up(e.Zn, e.hZn); up(e.Zt, e.hZt); up(e.Yn, e.hYn); up(e.Yt, e.hYt);
    up(e.Wn, e.hWn); up(e.Wt, e.hWt); up(e.Tt, e.hTt); up(e.St, e.hSt);
    e.Ar_j = mask(e.nA); e.Zr_j = mask(e.nZ);
    e.ar.d = dev.alloc<int>(e.nA + 1); e.zr.d = dev.alloc<int>(e.nZ + 1);
    e.mlfAr = dbl(16); e.mlfZr = dbl(16);
    e.lfA = dbl((size_t)kAjdk * e.nA); e.lfY = dbl((size_t)kAjdk * e.nA);
    e.lfZ = dbl((size_t)kAjdk * e.nZ); e.lfW = dbl((size_t)kAjdk * e.nZ);
    e.AtTn = dbl((size_t)T * M); e.ZtSn = dbl((size_t)T * M); e.YtTn = dbl((size_t)T * NY); e.WtSn = dbl((size_t)T * NY);
    e.sc = dbl((size_t)6 * T); e.etA = dbl(M); e.etY = dbl(NY);
    for (int q = 0; q < 4; q++) {
      e.cz[q] = dev.alloc<uint32_t>(std::max<size_t>((size_t)T * X(q), 1)); e.scal[q] = dbl(8);
      e.QRn[q] = dbl((size_t)nbins * T * e.nA);
      fill(e.QRn[q], (long long)nbins * T * e.nA, 0.0, st);
    }
    // the Z rows never change: their mask, list and count
    HIPCHK(hipMemcpyAsync(e.Zr_j, e.hZr_j.data(), bsize(e.nZ), hipMemcpyHostToDevice, st));
    upload_list(e.zr, e.hZr_j, e.nZ);
    mask_lf(e.Zr_j, psize(e.nZ), nullptr, e.mlfZr, st);
  }
  F.assign((size_t)nbins * nbins, Pair());
  for (int nb1 = 0; nb1 < nbins; nb1++) for (int nb2 = 0; nb2 < nbins; nb2++) {
    Pair& f = F[nb1 + nb2 * nbins];
    for (int q = 0; q < 4; q++) {
      f.x[q] = dbl((size_t)T * E[nb1].nA); f.qr[q] = dbl((size_t)T * E[nb1].nA); f.qc[q] = dbl((size_t)T * X(q));
      f.pg[q] = dev.alloc<unsigned long long>(34 * (size_t)std::max(E[nb1].nA, 1));
    }
  }
  for (int q = 0; q < 4; q++) { QCn[q] = dbl((size_t)nbins * nbins * T * X(q)); fill(QCn[q], (long long)nbins * nbins * T * X(q), 0.0, st); }
  // what every pass zeroes: the pair terms, the bins' scalar sums, and the normalised block of
  // each pair whose term is on (dcc_sumscores_nrm zeroes only those, right before it writes
  // them; a block of a term that is off keeps what the in-place bin sums left there)
  pass_fills.clear();
  for (int nb1 = 0; nb1 < nbins; nb1++) for (int nb2 = 0; nb2 < nbins; nb2++) {
    Pair& f = F[nb1 + nb2 * nbins]; Bin& e1 = E[nb1];
    for (int q = 0; q < 4; q++) {
      pass_fills.push_back(FillJob{ f.qr[q], (long long)T * e1.nA });
      pass_fills.push_back(FillJob{ f.qc[q], (long long)T * X(q) });
      if (!on(q, nb1, nb2)) continue;
      pass_fills.push_back(FillJob{ e1.QRn[q] + (size_t)nb2 * T * e1.nA, (long long)T * e1.nA });
      pass_fills.push_back(FillJob{ QCn[q] + (size_t)(nb1 + nb2 * nbins) * T * X(q), (long long)T * X(q) });
    }
  }
  for (Bin& e : E) pass_fills.push_back(FillJob{ e.sc, 6LL * T });
  // the digit planes of each bin's quantised products, shared by the first bins' row sets
  int max_nA = 0; for (const Bin& e : E) max_nA = std::max(max_nA, e.nA);
  // (room for the row partials of any row count up to the largest bin's: how many parts a
  // contraction splits into grows as its row count falls)
  prep_bytes = std::max(ztsn_mfma_work_bytes(0, T, bsize(M), B), ztsn_mfma_work_bytes(0, T, bsize(NY), B)) +
               sizeof(unsigned long long) * 34 * (size_t)std::max(max_nA, 1) + 512;
  prep.assign((size_t)nbins * 4, nullptr);
  for (auto& p : prep) p = dev.alloc<uint8_t>(prep_bytes);
  L = dbl((size_t)T * std::max(M, NY)); xwork = dbl(768);
  dsv = dbl((size_t)M + A_nrows_total);
  cnt = dev.alloc<int>(std::max(std::max(M, NY), T) + 1);
  if (files) data_freqs();
  HIPCHK(hipStreamSynchronize(st));
}

// dcc_load_A_p (dcc.c:875-991): per column the counts of the loaded (mr_b) rows summed over
// the bins, A with Z and Y with W; per 1920-column block the mean count over the block's
// loaded columns divided by the loaded rows, cropped to [0.01, 0.99] (0.5 everywhere when
// GLOBAL_TEST_sparse is 0); then the ajdk tables and the Y columns' weights again
void DccDev::data_freqs()
{
  std::vector<double> AZ(M + 1, 0.0), YW(NY + 1, 0.0);
  std::vector<int> h(std::max(M, NY) + 1, 0);
  auto count = [&](const DMat& Xt, int rows, const uint32_t* mask, const std::vector<uint8_t>& cm, std::vector<double>& acc) {
    if (rows <= 0 || Xt.nr <= 0) return;
    popc_rows(Xt.b(), nullptr, Xt.nr, mask, (int)(Xt.rb / 4), cnt, st);
    HIPCHK(hipMemcpyAsync(h.data(), cnt, sizeof(int) * Xt.nr, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < Xt.nr; k++) if (bg(cm.data(), k)) acc[k] += (double)h[k];
  };
  for (Bin& e : E) {
    // the row masks start as loaded (mr_j = mr_b), so the working masks hold mr_b here
    HIPCHK(hipMemcpyAsync(e.Ar_j, e.hAr_b.data(), bsize(e.nA), hipMemcpyHostToDevice, st));
    count(e.At, e.nA, e.Ar_j, hAc_b, AZ);
    count(e.Zt, e.nZ, e.Zr_j, hAc_b, AZ);
    count(e.Yt, e.nA, e.Ar_j, hYc_b, YW);
    count(e.Wt, e.nZ, e.Zr_j, hYc_b, YW);
  }
  auto freq = [&](const std::vector<double>& rs, const std::vector<uint8_t>& cm, int n, std::vector<double>& p) {
    for (int b = 0; b < (int)p.size(); b++) {
      p[b] = 0.5;
      if (!c.sparse) continue;
      const int s0 = std::min(n - 1, std::max(0, b * kPopBits)), s1 = std::min(n - 1, std::max(0, (b + 1) * kPopBits - 1));
      double tsum = 0, tnum = 0;
      for (int k = s0; k <= s1; k++) if (bg(cm.data(), k)) { tsum += rs[k]; tnum += 1; }
      if (tnum == 0) p[b] = 0.5;
      else { const double v = (tsum) / (tnum * (A_rpop_b_total + Z_rpop_b_total)); p[b] = v < 0.01 ? 0.01 : (v > 0.99 ? 0.99 : v); }
    }
  };
  freq(AZ, hAc_b, M, hA_p); freq(YW, hYc_b, NY, hY_p);
  // GLOBAL_A_p_name / GLOBAL_Y_p_name (mda_read_d2_r8): the frequencies from a file instead
  const std::string* pf[2] = { &c.A_p, &c.Y_p };
  std::vector<double>* pv[2] = { &hA_p, &hY_p };
  for (int v = 0; v < 2; v++) {
    if (pf[v]->empty()) continue;
    FILE* fp = fopen(pf[v]->c_str(), "rb"); int hdr[3];
    const bool ok = fp && fread(hdr, sizeof(int), 3, fp) == 3 && hdr[1] == (int)pv[v]->size() &&
                    fread(pv[v]->data(), sizeof(double), pv[v]->size(), fp) == pv[v]->size();
    if (fp) fclose(fp);
    if (!ok) throw Error(LKC_ERR_IO, "cannot read " + *pf[v]);
  }
  ajdk_of(hA_p, hA_ajdk); ajdk_of(hY_p, hY_ajdk);
  HIPCHK(hipMemcpyAsync(A_ajdk, hA_ajdk.data(), sizeof(double) * hA_ajdk.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(Y_ajdk, hY_ajdk.data(), sizeof(double) * hY_ajdk.size(), hipMemcpyHostToDevice, st));
  mask_lf(Yc_j, Ypc, Y_ajdk, mlfY, st);
  HIPCHK(hipStreamSynchronize(st));
}

// dcc_M_mxset (dcc.c:1092-1113): the active masks, lists and counts reach the device
void DccDev::mxset()
{
  HIPCHK(hipMemcpyAsync(Ac_j, hAc_j.data(), bsize(M), hipMemcpyHostToDevice, st));
  upload_list(kl, hAc_j, M);
  A_cpop_j = kl.n();
  A_rpop_j_total = 0; Z_rpop_j_total = 0;
  for (Bin& e : E) {
    HIPCHK(hipMemcpyAsync(e.Ar_j, e.hAr_j.data(), bsize(e.nA), hipMemcpyHostToDevice, st));
    upload_list(e.ar, e.hAr_j, e.nA);
    A_rpop_j_total += e.ar.n(); Z_rpop_j_total += e.zr.n();
    mask_lf(e.Ar_j, psize(e.nA), nullptr, e.mlfAr, st);
  }
  mask_lf(Ac_j, Apc, A_ajdk, mlfA, st);
  HIPCHK(hipStreamSynchronize(st));   // the host lists may change before the next upload
}

// the half-loop subscores of every bin pair on the current masks
void DccDev::subscores()
{
  const int nt = tl.n(), nk = kl.n(), ny = yl.n();
  fill_jobs(pass_fills, st);
  for (Bin& e : E) {
    // dcc_An_ajdk (dcc_lf.c:7-51)
    if (e.nA > 0 && M > 0) an_ajdk(e.An.b(), e.ar.d, e.ar.n(), Ac_j, Apc, A_ajdk, mlfA, e.lfA, e.nA, st);
    if (e.nA > 0 && NY > 0) an_ajdk(e.Yn.b(), e.ar.d, e.ar.n(), Yc_j, Ypc, Y_ajdk, mlfY, e.lfY, e.nA, st);
    if (e.nZ > 0 && M > 0) an_ajdk(e.Zn.b(), e.zr.d, e.zr.n(), Ac_j, Apc, A_ajdk, mlfA, e.lfZ, e.nZ, st);
    if (e.nZ > 0 && NY > 0) an_ajdk(e.Wn.b(), e.zr.d, e.zr.n(), Yc_j, Ypc, Y_ajdk, mlfY, e.lfW, e.nZ, st);
    ops += 15LL * e.ar.n() * (long long)(e.An.rb + e.Yn.rb) * 8 + 15LL * e.zr.n() * (long long)(e.Zn.rb + e.Wn.rb) * 8;
    // dcc_lf_ZtSn (dcc_lf.c:53-100): plain +-1 products over the bin's active rows
    if (e.nA > 0) {
      pm_few(e.At.b(), kl.d, nk, e.Tt.b(), tl.d, nt, e.Ar_j, psize(e.nA), e.mlfAr, e.AtTn, M, st);
      pm_few(e.Yt.b(), yl.d, ny, e.Tt.b(), tl.d, nt, e.Ar_j, psize(e.nA), e.mlfAr, e.YtTn, NY, st);
    }
    if (e.nZ > 0) {
      pm_few(e.Zt.b(), kl.d, nk, e.St.b(), tl.d, nt, e.Zr_j, psize(e.nZ), e.mlfZr, e.ZtSn, M, st);
      pm_few(e.Wt.b(), yl.d, ny, e.St.b(), tl.d, nt, e.Zr_j, psize(e.nZ), e.mlfZr, e.WtSn, NY, st);
    }
    ops += (long long)(nk + ny) * nt * ((long long)bsize(e.nA) + (long long)bsize(e.nZ)) * 8;
    // TAnZtS stage 0 (TAnZtS.c:5-209): et_Tn, et_Sn, the columns' et over the active rows,
    // and the ordered a1d1 sums
    popc_rows(e.Tt.b(), tl.d, nt, e.Ar_j, (int)(e.Tt.rb / 4), cnt, st);
    et_from_counts(cnt, tl.d, nt, e.ar.n(), e.sc + 4 * T, 1, st);
    popc_rows(e.St.b(), tl.d, nt, e.Zr_j, (int)(e.St.rb / 4), cnt, st);
    et_from_counts(cnt, tl.d, nt, e.zr.n(), e.sc + 5 * T, 1, st);
    popc_rows(e.At.b(), kl.d, nk, e.Ar_j, (int)(e.At.rb / 4), cnt, st);
    et_from_counts(cnt, kl.d, nk, e.ar.n(), e.etA, 1, st);
    popc_rows(e.Yt.b(), yl.d, ny, e.Ar_j, (int)(e.Yt.rb / 4), cnt, st);
    et_from_counts(cnt, yl.d, ny, e.ar.n(), e.etY, 1, st);
    OrderedJobs J; J.n = 0;
    for (int t : tl.h) {
      const uint8_t* Tr = e.Tt.p + (size_t)t * e.Tt.rb; const uint8_t* Sr = e.St.p + (size_t)t * e.St.rb;
      J.job[J.n++] = OrderedJob{ e.ar.d, e.ar.n(), Tr, e.lfA + 6 * (size_t)e.nA, e.sc + t };
      J.job[J.n++] = OrderedJob{ e.zr.d, e.zr.n(), Sr, e.lfZ + 6 * (size_t)e.nZ, e.sc + T + t };
      J.job[J.n++] = OrderedJob{ e.ar.d, e.ar.n(), Tr, e.lfY + 6 * (size_t)e.nA, e.sc + 2 * T + t };
      J.job[J.n++] = OrderedJob{ e.zr.d, e.zr.n(), Sr, e.lfW + 6 * (size_t)e.nZ, e.sc + 3 * T + t };
    }
    ordered_sums(J, st);
    // TAnZtS stage 1, first half: the AJDK_0_1-weighted products quantised (An_a0d1_ZtSn_excerpt,
    // An_ZtSn_ww.c:243-265, then xcalc)
    struct Qz { const double* src; const IList* l; int ld; const double* ajdk; int pc; int q; };
    const Qz qs[4] = { { e.WtSn, &yl, NY, Y_ajdk, Ypc, YW }, { e.ZtSn, &kl, M, A_ajdk, Apc, AZ },
                       { e.YtTn, &yl, NY, Y_ajdk, Ypc, YY }, { e.AtTn, &kl, M, A_ajdk, Apc, AA } };
    for (const Qz& z : qs) {
      const int n = z.l->n();
      if (n <= 0 || nt <= 0) continue;
      if ((z.q == YW || z.q == AZ) ? e.nZ <= 0 : e.nA <= 0) continue;
      scale_rows(z.src, z.ld, z.l->d, n, tl.d, nt, z.ajdk + 5 * (size_t)z.pc, L, st);
      xcalc(L, 1, n, nullptr, nullptr, n, nt, z.l->d, B, c.D_MLT, e.cz[z.q], z.ld, e.scal[z.q], xwork, st);
      ops += (long long)n * nt * (B + 1) * 8;
    }
  }
  // the digit planes of each second bin's quantised products, once per pass
  for (int nb2 = 0; nb2 < nbins; nb2++) for (int q = 0; q < 4; q++) {
    const Bin& e2 = E[nb2];
    const IList& l = (q == YW || q == YY) ? yl : kl;
    const bool used = (q == YW || q == AZ) ? e2.nZ > 0 : e2.nA > 0;
    if (!used || nt <= 0 || l.n() <= 0 || X(q) <= 0) continue;
    const size_t rb = bsize(X(q));
    HIPCHK(ztsn_mfma_prep(rb, (q == YW || q == YY) ? Yc_j : Ac_j, e2.cz[q], X(q), X(q), nt, B, prep[(size_t)nb2 * 4 + q],
                          prep_bytes, st));
  }
  std::vector<PairQC> qcj; std::vector<PairQR> qrj; std::vector<ZtsnJob> zj;
  for (int nb1 = 0; nb1 < nbins; nb1++) for (int nb2 = 0; nb2 < nbins; nb2++) {
    Bin &e1 = E[nb1], &e2 = E[nb2];
    Pair& f = F[nb1 + nb2 * nbins];
    const int na = e1.ar.n();
    const bool same = nb1 == nb2;
    // dcc_lf_D_AtTn_ZtSn_vv (dcc_lf.c:102-163) into the half-loop's QC arrays
    auto qc = [&](int q, const IList& l, long long ld, const double* L1, const double* L2, const double* etS,
                  const double* etA, const double* ajdk, int pc) {
      if (!on(q, nb1, nb2)) return;
      PairQC a{ l.d, l.n(), tl.d, nt, ld, L1, L2, e1.sc + 4 * T, etS, etA, na, ajdk, pc, {}, f.qc[q] };
      std::copy(tb, tb + kMaxT, a.tb);
      qcj.push_back(a);
    };
    qc(YW, yl, NY, e1.YtTn, e2.WtSn, e2.sc + 5 * T, nullptr, Y_ajdk, Ypc);
    qc(YY, yl, NY, e1.YtTn, e2.YtTn, e2.sc + 4 * T, same ? e1.etY : nullptr, Y_ajdk, Ypc);
    qc(AZ, kl, M, e1.AtTn, e2.ZtSn, e2.sc + 5 * T, nullptr, A_ajdk, Apc);
    qc(AA, kl, M, e1.AtTn, e2.AtTn, e2.sc + 4 * T, same ? e1.etA : nullptr, A_ajdk, Apc);
    // TAnZtS stage 1, second half (An_ZtSn_ww of the first bin's rows with the second bin's
    // quantised products), stage 2 and the half-loop's QR arrays
    auto qr = [&](int q, const DMat& Xn, const uint32_t* cmask, const IList& l, int ld, const double* a1,
                  const double* et, const double* x6, const double* mlf, const DMat* Tt) {
      if (!on(q, nb1, nb2) || na <= 0 || nt <= 0 || l.n() <= 0) return;
      (void)ld;
      const void* w = prep[(size_t)nb2 * 4 + q];
      zj.push_back(ZtsnJob{ Xn.p, Xn.rb, e1.ar.d, na, (const uint8_t*)cmask, ztsn_prep_digits(w), (long long)Xn.rb * 8,
                            nt, ztsn_prep_sums(w, nt, Xn.rb, B), tl.d, e2.scal[q], f.x[q], e1.nA, f.pg[q],
                            ztsn_job_parts(na, Xn.rb) });
      ops += (long long)na * nt * (B + 1) * (long long)Xn.rb * 8;
      PairQR a{ e1.ar.d, na, tl.d, nt, e1.nA, f.x[q], a1, et, x6, mlf, Tt ? Tt->p : nullptr, Tt ? Tt->rb : 0, {}, f.qr[q] };
      std::copy(tb, tb + kMaxT, a.tb);
      qrj.push_back(a);
    };
    qr(YW, e1.Yn, Yc_j, yl, NY, e2.sc + 3 * T, e2.sc + 5 * T, e1.lfY + 6 * (size_t)e1.nA, mlfY, nullptr);
    qr(AZ, e1.An, Ac_j, kl, M, e2.sc + 1 * T, e2.sc + 5 * T, e1.lfA + 6 * (size_t)e1.nA, mlfA, nullptr);
    qr(YY, e1.Yn, Yc_j, yl, NY, e2.sc + 2 * T, e2.sc + 4 * T, e1.lfY + 6 * (size_t)e1.nA, mlfY, same ? &e1.Tt : nullptr);
    qr(AA, e1.An, Ac_j, kl, M, e2.sc + 0 * T, e2.sc + 4 * T, e1.lfA + 6 * (size_t)e1.nA, mlfA, same ? &e1.Tt : nullptr);
  }
  HIPCHK(ztsn_mfma_jobs(zj, B, st));
  pair_qc_jobs(qcj, st);
  pair_qr_jobs(qrj, st);
}

// dcc_sumscores_nrm with GLOBAL_Ireq <= 0 (dcc_sumscores.c:569-699): each bin pair's terms over
// its counts of summands (the half-loop arrays read at the covariate's original index, as the
// reference reads them), then the pairs summed into the first block in pair order
void DccDev::nrm()
{
  const long long At = A_rpop_j_total, Zt = Z_rpop_j_total, Yc = yl.n(), Ac = kl.n();
  auto den = [](long long lld) { double d = (double)lld; if (fabs(d) < 0.000000000001) d = 1; return d; };
  const long long lldr[4] = { Zt * Yc, Zt * Ac, (At - 1) * Yc, (At - 1) * Ac };
  const long long lldc[4] = { At * Zt, At * Zt, At * (At - 1), At * (At - 1) };
  // (the blocks of the terms that are on were zeroed at the start of the pass)
  std::vector<NrmJob> jobs;
  for (int nb1 = 0; nb1 < nbins; nb1++) for (int nb2 = 0; nb2 < nbins; nb2++) {
    const int nbx = nb1 + nb2 * nbins; Bin& e = E[nb1]; Pair& f = F[nbx];
    for (int q = 0; q < 4; q++) {
      if (!on(q, nb1, nb2)) continue;
      double* dr = e.QRn[q] + (size_t)nb2 * T * e.nA;
      jobs.push_back(NrmJob{ ScoreStage{ e.ar.d, e.ar.n(), tl.d, tl.n(), e.nA }, f.qr[q], den(lldr[q]), dr });
      const IList& l = (q == YW || q == YY) ? yl : kl;
      double* dc = QCn[q] + (size_t)nbx * T * X(q);
      jobs.push_back(NrmJob{ ScoreStage{ l.d, l.n(), tl.d, tl.n(), X(q) }, f.qc[q], den(lldc[q]), dc });
    }
  }
  nrm_jobs(jobs, st);
  for (int nb1 = 0; nb1 < nbins; nb1++) {
    Bin& e = E[nb1];
    if (e.nA <= 0) continue;
    const bool z = e.nZ > 0, y = NY > 0;
    const bool use[4] = { z && y, z && M > 0, y, M > 0 };
    for (int q = 0; q < 4; q++) if (use[q])
      bins_sumx(e.ar.d, e.ar.n(), tl.d, tl.n(), e.nA, e.QRn[q], (long long)T * e.nA, nbins, st);
  }
  const bool ar = A_rpop_b_total > 0, zr = Z_rpop_b_total > 0;
  const bool usec[4] = { ar && NY > 0 && zr, ar && M > 0 && zr, ar && NY > 0 && zr, ar && M > 0 };   // by term
  for (int q = 0; q < 4; q++) if (usec[q]) {
    const IList& l = (q == YW || q == YY) ? yl : kl;
    bins_sumx(l.d, l.n(), tl.d, tl.n(), X(q), QCn[q], (long long)T * X(q), nbins * nbins, st);
  }
}
// dcc_sumscores_ifT (dcc_sumscores.c:418-476): n_mds = min(6, active covariates - 1), scale
// GLOBAL_kappa_squared whatever its sign (the stray ';' of `if (GLOBAL_kappa_squared>0);`),
// so nothing changes when it is 0; with GLOBAL_Ireq <= 0 only block nb2 = 0 of each bin
void DccDev::ifT()
{
  int n_mds = std::max(0, tl.n() - 1); if (n_mds > 6) n_mds = 6;
  const double dmds = n_mds > 0 ? n_mds * c.kappa2 : 0.0;
  if (!(dmds > 0)) return;
  const bool arb = A_rpop_b_total > 0, zrb = Z_rpop_b_total > 0;
  for (int nb1 = 0; nb1 < nbins; nb1++) {
    Bin& e = E[nb1];
    const int nbx = nb1;
    if (e.nA > 0) {
      const bool use[4] = { e.nZ > 0 && NY > 0, e.nZ > 0 && M > 0, NY > 0, M > 0 };
      for (int q : { YW, YY, AZ, AA }) if (use[q]) score_ifT(ScoreStage{ e.ar.d, e.ar.n(), tl.d, tl.n(), e.nA }, e.QRn[q], dmds, 1, st);
    }
    const bool usec[4] = { arb && NY > 0 && zrb, arb && M > 0 && zrb, arb && NY > 0, arb && M > 0 };
    for (int q : { YW, YY, AZ, AA }) if (usec[q]) {
      const IList& l = (q == YW || q == YY) ? yl : kl;
      score_ifT(ScoreStage{ l.d, l.n(), tl.d, tl.n(), X(q) }, QCn[q] + (size_t)nbx * T * X(q), dmds, 0, st);
    }
  }
}
// dcc_sumscores_cmb (dcc_sumscores.c:243-278), first covariate column
void DccDev::cmb()
{
  if (M > 0 && A_rpop_b_total > 0 && Z_rpop_b_total > 0)
    for (int nbx = 0; nbx < nbins * nbins; nbx++)
      score_cmb(ScoreStage{ kl.d, kl.n(), tl.d, tl.n(), M }, QCn[AA] + (size_t)nbx * T * M, QCn[AZ] + (size_t)nbx * T * M, st);
  for (Bin& e : E) {
    if (e.nA <= 0 || M <= 0) continue;
    bins_cmb_rows(e.ar.d, e.ar.n(), nbins, (long long)T * e.nA, e.QRn[AA], e.nZ > 0 && NY > 0 ? e.QRn[YW] : nullptr,
                  e.nZ > 0 ? e.QRn[AZ] : nullptr, NY > 0 ? e.QRn[YY] : nullptr, st);
  }
}
// dcc_sumscores_xij (dcc_sumscores.c:158-241), GLOBAL_Ireq = 0: the first covariate column of
// the summed A-with-A scores, columns then the rows of every bin in bin order, sorted by the
// reference's quicksort
void DccDev::xij()
{
  Irem = 0; for (Bin& e : E) Irem += e.ar.n() > 0 ? 1 : 0;
  const int nk = kl.n();
  score_gather0(ScoreStage{ kl.d, nk, tl.d, tl.n(), M }, QCn[AA], dsv, st);
  QR_bin.clear(); QR_row.clear();
  int off = 0;
  for (int nb = 0; nb < nbins; nb++) {
    Bin& e = E[nb];
    score_gather0(ScoreStage{ e.ar.d, e.ar.n(), tl.d, tl.n(), e.nA }, e.QRn[AA], dsv + M + off, st);
    for (int r : e.ar.h) { QR_bin.push_back(nb); QR_row.push_back(r); }
    off += e.ar.n();
  }
  std::vector<double> h((size_t)M + off);
  HIPCHK(hipMemcpyAsync(h.data(), dsv, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  QC_sv.assign(h.begin(), h.begin() + nk); QC_perm.resize(nk);
  for (int i = 0; i < nk; i++) QC_perm[i] = i;
  if (M > 0 && A_rpop_b_total > 0) sort_scores(QC_sv.data(), QC_perm.data(), nk);
  QR_sv.assign(h.begin() + M, h.begin() + M + off); QR_perm.resize(off);
  for (int i = 0; i < off; i++) QR_perm[i] = i;
  if (A_rpop_b_total > 0 && M > 0) sort_scores(QR_sv.data(), QR_perm.data(), off);
  A_rpop_j_total = off;
}
// dcc_sumscores_mxA / mxB / mxC (dcc_sumscores.c:5-156)
void DccDev::mxA()
{
  int rdrop = 0, cdrop = 0;
  const int nk = kl.n(), na = A_rpop_j_total;
  get_xdrop(na, nk, c.gamma, c.gamma_type, &rdrop, &cdrop);
  const int nc = std::min(cdrop, nk), nr = std::min(rdrop, na);
  hAc_rtn = hAc_j;
  for (int i = 0; i < nc; i++) bc(hAc_rtn.data(), kl.h[QC_perm[i]]);
  for (Bin& e : E) e.hAr_rtn = e.hAr_j;
  for (int i = 0; i < nr; i++) { const int p = QR_perm[i]; bc(E[QR_bin[p]].hAr_rtn.data(), QR_row[p]); }
  auto mean = [](const std::vector<double>& v, unsigned long long n) {
    double m = 0; for (unsigned long long i = 0; i < n; i++) m += v[i]; m /= (double)n; return m;
  };
  double mr = mean(QR_sv, (unsigned long long)na); if (!std::isfinite(mr)) mr = 1.0;
  double mc = mean(QC_sv, (unsigned long long)nk); if (!std::isfinite(mc)) mc = 1.0;
  const double tr[6] = { (double)out->n_iter, (double)na, (double)nk, mr, mc, (double)Irem };
  out->trace.insert(out->trace.end(), tr, tr + 6);
  out->n_iter++;
  std::vector<int> cb(M, 0);
  for (int k = 0, b = 0; k < M; k++) { cb[k] = b; if (bg(hAc_b.data(), k)) b++; }
  for (int i = 0; i < nr; i++) {
    const int p = QR_perm[i]; const Bin& e = E[QR_bin[p]];
    out->xdrop_a.push_back(e.g_a[QR_row[p]]); out->xdrop_a.push_back(-1);
    out->xdrop_b.push_back(e.g_b[QR_row[p]]); out->xdrop_b.push_back(-1);
  }
  for (int i = 0; i < nc; i++) {
    const int k = kl.h[QC_perm[i]];
    out->xdrop_a.push_back(-1); out->xdrop_a.push_back(k); out->xdrop_b.push_back(-1); out->xdrop_b.push_back(cb[k]);
  }
}
// dcc_lrup_mxdup (dcc_lf.c:531-542)
void DccDev::mxdup()
{
  hAc_j = hAc_rtn;
  for (Bin& e : E) e.hAr_j = e.hAr_rtn;
}

// dcc_sumscores_dmp (dcc_sumscores.c:90-118)
void DccDev::write(const std::string& dir) const
{
  mkdir(dir.c_str(), 0755);
  auto open = [&](const char* f) {
    FILE* fp = fopen((dir + "/" + f).c_str(), "w");
    if (!fp) throw Error(LKC_ERR_IO, "cannot write " + dir + "/" + f);
    return fp;
  };
  FILE* ft = open("out_trace.txt");
  for (size_t i = 0; i + 5 < out->trace.size(); i += 6) {
    const double* t = &out->trace[i];
    fprintf(ft, "%d %d %d %0.16f %0.16f %d\n", (int)t[0], (int)t[1], (int)t[2], t[3], t[4], (int)t[5]);
  }
  fclose(ft);
  FILE* fa = open("out_xdrop_a.txt");
  FILE* fb = open("out_xdrop_b.txt");
  for (size_t i = 0; i + 1 < out->xdrop_a.size(); i += 2) {
    fprintf(fa, "%d %d\n", out->xdrop_a[i], out->xdrop_a[i + 1]);
    fprintf(fb, "%d %d\n", out->xdrop_b[i], out->xdrop_b[i + 1]);
  }
  fclose(fa); fclose(fb);
}

// the dexcluster_driver loop (dexcluster_driver.c:22-53) from the first subscores to the end
void run_loop(DccDev& D, int max_iter)
{
  D.mxset();
  D.subscores();
  int nl = 0;
  while ((max_iter <= 0 || nl < max_iter) && (D.A_rpop_j_total > 0 || D.A_cpop_j > 0)) {
    D.nrm(); D.ifT(); /* dcc_sumscores_srt: nothing with GLOBAL_Ireq <= 0 */ D.cmb();
    D.xij(); D.mxA();
    D.mxdup(); D.mxset();
    D.subscores();
    nl++;
  }
  HIPCHK(hipStreamSynchronize(D.st));
  HIPCHK(hipGetLastError());
}

}  // namespace

void dcc_driver_bins(const Cfg& c, DccTestOut& out, const std::string& out_dir)
{
  if (c.nbins < 1 || c.nbins > 64) throw Error(LKC_ERR_CONFIG, "GLOBAL_NBINS must be in [1,64]");
  // (GLOBAL_Ireq > 0 with one bin is the GLOBAL_Ireq <= 0 run: one value sorted per row)
  if (c.Ireq > 0 && c.nbins > 1) throw Error(LKC_ERR_CONFIG, "GLOBAL_Ireq > 0 (per-bin sorting) is not supported");
  if (c.scramble_num > 0) throw Error(LKC_ERR_CONFIG, "bcc_scramble is not part of dexcluster_driver");
  if (c.B < 1 || c.B > 32) throw Error(LKC_ERR_CONFIG, "GLOBAL_B_MLT must be in [1,32]");
  DccDev D;
  D.c = c; D.out = &out; D.B = c.B;
  out = DccTestOut();
  D.load();
  D.setup();
  run_loop(D, 0);
  out.ops = D.ops;
  if (!out_dir.empty()) D.write(out_dir);
}

void dcc_sumscores_test(const Cfg& c, DccTestOut& out, const std::string& out_dir)
{
  if (c.nbins < 1) throw Error(LKC_ERR_CONFIG, "dcc_sumscores needs GLOBAL_NBINS >= 1");
  if (c.test_typ2 != "error") throw Error(LKC_ERR_CONFIG, "dcc_sumscores: GLOBAL_TEST_TYP2 must be error (random data)");
  if (c.Ireq > 0 && c.nbins > 1) throw Error(LKC_ERR_CONFIG, "GLOBAL_Ireq > 0 (per-bin sorting) is not supported");
  if (c.test_T_cols < 1 || c.test_T_cols > kMaxT) throw Error(LKC_ERR_CONFIG, "dcc_sumscores: GLOBAL_TEST_T_n_cols must be 1..8");
  if (c.test_A_cols < 1 || c.test_Y_cols < 0 || c.test_A_rows < 0 || c.test_Z_rows < 0 || c.B < 1 || c.B > 32)
    throw Error(LKC_ERR_CONFIG, "dcc_sumscores: bad problem size");
  DccDev D;
  D.c = c; D.out = &out;
  D.nbins = c.nbins; D.M = c.test_A_cols; D.NY = c.test_Y_cols; D.T = c.test_T_cols; D.B = c.B;
  out = DccTestOut();
  D.generate();
  D.setup();
  run_loop(D, c.test_niter);
  out.ops = D.ops;
  if (!out_dir.empty()) D.write(out_dir);
}

}  // namespace lkc
