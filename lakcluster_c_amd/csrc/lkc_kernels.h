// lkc_kernels.h -- launchers for the HIP kernels of the bcc elimination path (gfx950).
//
// Bit matrices are row-major packed bit rows, most significant bit first in
// each byte, bit 1 <=> +1; every row is bsize(ncols) bytes (a multiple of
// the 240-byte / 1920-bit weighting block).  Double arrays indexed by a
// (row, covariate) pair are stored covariate-major: X[t*ld + row].
// All launchers are asynchronous on `st`.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace lkc {

// A block barrier that orders LDS only: the waves' LDS writes before it are visible to
// LDS reads after it, while global loads issued before it stay in flight (__syncthreads
// would also wait for them).  For kernels that prefetch from memory across barriers.
__device__ __forceinline__ void block_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kPopBits = 1920;   // weighting block (lakcluster_header.h:268)
constexpr int kPopBytes = 240;
constexpr int kPopWords = 60;    // 32-bit words per block
constexpr int kAjdk = 15;
constexpr int kMaxT = 8;         // covariate columns supported by the fused kernels

size_t bsize(int n);             // bytes of one packed row of n bits
int psize(int n);                // number of 1920-bit blocks

struct Bits { const uint8_t* p; int nr; int nc; size_t rb; };

// Σ_b w_jk[b]·|mask ∩ block b| for the 15 tables (out[0..14], only when ajdk
// is non-null) and the plain count Σ_b 1.0·|mask ∩ block b| (out[15]); ordered over b.
void mask_lf(const uint32_t* mask, int pcols, const double* ajdk, double* out16, hipStream_t st);
struct MaskLfJob { const uint32_t* mask; int pcols; const double* ajdk; double* out16; };
struct MaskLfJobs { MaskLfJob job[8]; int n; };
void mask_lf_jobs(const MaskLfJobs& J, hipStream_t st);

// An_ajdk_v: out[jk*ld + rows[i]] = -base[jk] + 2 Σ_b ajdk[jk][b]·pop(X_row ∩ mask ∩ b)
void an_ajdk(const Bits& X, const int* rows, int nrows, const uint32_t* mask, int pcols,
             const double* ajdk, const double* base16, double* out, int ld, hipStream_t st);

// AnZt_vv.  Output element for (alist[i], zlist[j]) goes to
//   out[zidx*ldo + aidx]  with aidx = a_orig ? alist[i] : i, zidx = z_orig ? zlist[j] : j.
// centred: D-weighted xor counts + ajdk correction with xa[alist[i]], xz[zlist[j]];
// otherwise the plain +-1 product.  op_sub: out -= value instead of out = value.
void gram(const Bits& A, const int* alist, int na, const Bits& Z, const int* zlist, int nz,
          const uint32_t* mask, int pcols, const double* ajdk, const double* base16,
          const double* xa, const double* xz, int centred,
          double* out, long long ldo, int a_orig, int z_orig, int op_sub, hipStream_t st,
          const int* seg = nullptr, int nseg = 0);
// Column compaction for grams over a sparse column set: seg (3 ints per non-empty
// 1920-block: word offset, word count, block index) describes the compacted rows.
// out[i*ldw + w] = bits of X row (rows ? rows[i] : i) at columns src[32w .. 32w+31] (-1: none), w < W.
void compact_cols(const Bits& X, const int* rows, int nrows, const int* src, int W, uint32_t* out, int ldw, hipStream_t st);

// xcalc.  Values v(i,j) = src[ridx(i)*sn + cidx(j)*sw] for i<nn, j<nw, with
// ridx(i) = rlist ? rlist[i] : i (same for cidx).  scal receives
// [min_data, max_data, min, max, mlt (guarded), mlt (raw)].  cz[j*ldc + nout[i]]
// receives the B-bit integers.  work: 768 doubles.
void xcalc(const double* src, long long sn, long long sw, const int* rlist, const int* clist,
           int nn, int nw, const int* nout, int B, long long D_MLT,
           uint32_t* cz, long long ldc, double* scal, double* work, hipStream_t st);
// bit planes of cz restricted to nmask: planes[(j*B + nb)*rb + ...] over n (n_total bits per row)
void cz_planes(const uint32_t* cz, long long ldc, int nw, int n_total, const uint32_t* nmask,
               int B, uint8_t* planes, size_t rb, hipStream_t st);

// At_T_YnWt_S_Zn_ww.
//   P rows klist (n-dimension bits), mask pm (n), covariate rows Tc[tlist] over n,
//   planes (nw x B) over n with row0 (active n mask), Sc covariates over w, Q rows
//   qlist over w.  Result: out[tlist[t]*ldo + klist[k]].
struct WwArgs {
  Bits P; const int* klist; const int* qlist; int nk; const uint32_t* pm;
  Bits Tc; const int* tlist; const int* slist; int nt;
  const uint8_t* planes; size_t prb; const uint32_t* row0; int nw; const int* wlist; int B;
  const uint32_t* cz; long long ldc;      // cz[j*ldc + n]
  Bits Sc; Bits Q; const double* ajdk; int pcols; const double* scal;
  double* out; long long ldo;
  long long out_elems = 0;                // doubles of out a call may write (0: unknown)
};
// bit-plane VALU formulation (popcount of plane bit-rows)
size_t ww_work_bytes(int nk, int nt, int nw);
hipError_t ww_valu(const WwArgs& a, double* work, size_t work_bytes, hipStream_t st);

// int8 matrix-core formulation (cz split into signed bytes); planes are not used
size_t ww_mfma_work_bytes(int nk, int nt, int nw, size_t prb, int B, int pcols);
// hook: called on the stream right before and after the matrix-core kernel (timing)
struct WwHook { void (*before)(void*, hipStream_t); void (*after)(void*, hipStream_t); void* ctx; };
hipError_t ww_mfma(const WwArgs& a, void* work, size_t work_bytes, hipStream_t st, const WwHook* hook = nullptr);

// An_ZtSn_ww: out[slist[s]*ldo + alist[i]] = (base + Σ 2^nb pm0(A_a, plane(s,nb)))/mlt
void an_ztsn(const Bits& A, const int* alist, int na, const uint32_t* amask,
             const uint8_t* planes, size_t prb, const uint32_t* row0, int ns, const int* slist, int B,
             const double* scal, double* out, long long ldo, hipStream_t st);

// The same An_ZtSn_ww with its xcalc on caller device buffers, every intermediate in a
// file-scope device array of kNoallocScratchBytes (lkc_noalloc.hip): no device allocation.
// A: nr_total rows of bsize(n_total) bytes; out[a*ns + s] (zero for rows not in alist).
constexpr size_t kNoallocScratchBytes = (size_t)32 << 20;
size_t an_ztsn_noalloc_bytes(int nr_total, int n_total, int ns, int B);
hipError_t noalloc_scratch(void** ptr);   // the current device's copy of the array
hipError_t an_ztsn_noalloc(const uint8_t* A, int nr_total, const int* alist, int na, const uint8_t* amask,
                           const double* L, const int* nlist, int nn, int n_total, int ns, int B, long long D_MLT,
                           double* out, hipStream_t st);

// An_ZtSn_ww on the int8 matrix cores from the quantised values themselves (no planes):
// cz[s*ldc + n] for n < n_total, mask = the active n (the planes' row0 and amask, equal
// at every call site).  Same output as an_ztsn.
size_t ztsn_mfma_work_bytes(int na, int ns, size_t prb, int B);
hipError_t ztsn_mfma(const Bits& A, const int* alist, int na, const uint32_t* mask, const uint32_t* cz, long long ldc,
               int n_total, int ns, const int* slist, int B, const double* scal, double* out, long long ldo,
               void* work, size_t work_bytes, hipStream_t st);
// ztsn_mfma in two halves: the digit planes of cz (shared by every row set contracted with
// the same cz, mask and row width) and the contraction of one row set with them
hipError_t ztsn_mfma_prep(size_t rb, const uint32_t* mask, const uint32_t* cz, long long ldc, int n_total, int ns, int B,
                          void* work, size_t work_bytes, hipStream_t st);
hipError_t ztsn_mfma_run(const Bits& A, const int* alist, int na, const uint32_t* mask, int ns, const int* slist, int B,
                         const double* scal, double* out, long long ldo, void* work, size_t work_bytes, hipStream_t st);
// one contraction of a group: digit planes and column sums from ztsn_mfma_prep (Dg at the
// start of its work, Gsum after them), its own Pg of 34 * na words, and its n parts
struct ZtsnJob {
  const uint8_t* A; size_t rbA; const int* alist; int na; const uint8_t* mask; const int8_t* Dg; long long ldn;
  int ns; const long long* Gsum; const int* slist; const double* scal; double* out; long long ldo;
  unsigned long long* Pg; int parts;
};
struct ZtsnJobs { static constexpr int kMax = 24; int n; ZtsnJob job[kMax]; };
int ztsn_job_parts(int na, size_t rb);
// where ztsn_mfma_prep leaves the digit planes and column sums in work
inline const int8_t* ztsn_prep_digits(const void* work) { return (const int8_t*)work; }
inline const long long* ztsn_prep_sums(const void* work, int ns, size_t rb, int B)
{
  return (const long long*)((const uint8_t*)work + (((size_t)((B + 7) / 8) * ns * rb * 8 + 255) & ~(size_t)255));
}
hipError_t ztsn_mfma_jobs(const std::vector<ZtsnJob>& v, int B, hipStream_t st);

// AnZt_S_WnYt_vv with X==Y: out[tlist[t]*ldo + a] = Σ_z (X(a,z)*vS(z))*X(a,z) in list
// order of z, with X(a,z) = X[z*ldx + a] (x_rows 0) or X[a*ldx + z] (x_rows 1).
// sgbuf: qr_sign_bytes(nz, nt) bytes of device scratch
void qr(const double* X, long long ldx, const int* alist, int na, const int* zlist, int nz,
        const Bits& Sc, const int* tlist, int nt, double* out, long long ldo, signed char* sgbuf, hipStream_t st,
        int x_rows = 0);
size_t qr_sign_bytes(int nz, int nt);

// AtTYn_vv: out[(y*nt + t)*nk + k] (compact indices); lf_AtTn/lf_YtTn read at [t*ldT + row]
void attyn(const Bits& At, const int* alist, int nk, const uint32_t* amask, int acpop,
           const Bits& Tc, const int* tlist, int nt, const uint32_t* tmask, int tcpop,
           const Bits& Yt, const int* ylist, int ny, const double* lf_AtTn, const double* lf_YtTn, long long ldT,
           const double* ajdk, int pcols, double* out, hipStream_t st);
// AtTYn____WtSZn_vv: out[tlist[t]*ldo + klist[k]] = Σ_y (L1*D_y)*L2 (L1/L2 in attyn layout)
void attyn_wtszn(const double* L1, const double* L2, const int* klist, int nk, const int* ylist, int ny,
                 const int* tlist, int nt, const double* ajdk, int pcols, double* out, long long ldo, hipStream_t st);

// AtTYn_vv for the three (A, T, Y, mask) triples of the QC stage b1 fused with the
// four AtTYn____WtSZn_vv sums of stage b2, on the int8 matrix cores.
//   term 0 (attjn): A = At, T = Tt, Y = At, mask = retained rows
//   term 1 (ktrhn): the same over the removed rows
//   term 2 (ztsvn): A = Zt, T = St, Y = Zt, mask = active Z rows (absent without Z)
// A and Y rows are indexed by the original column index; mask has ldn bits (a
// multiple of 128) and mcount set bits; hAy/haY are lf_AtTn-like [t*ldh + row].
// Outputs [t*ldo + k]: vtszn (term0 x term2), jttan (0 x 0), htrkn (0 x 1), hhtrkn (1 x 1).
struct AttynTerm {
  const uint8_t* A; size_t rbA; const uint8_t* T; size_t rbT; const uint8_t* Y; size_t rbY;
  const uint8_t* mask; long long ldn; int mcount;
  const double* hAy; const double* haY; long long ldh;
};
struct AttynArgs {
  AttynTerm term[3]; int nterm;
  const int* klist; int nk; const int* ylist; int ny; const int* tlist; int nt;
  const double* ajdk; int pcols;
  double* vtszn; double* jttan; double* htrkn; double* hhtrkn; long long ldo;
};
size_t attyn_mfma_work_bytes(int ny, int nt);
hipError_t attyn_mfma(const AttynArgs& a, void* work, size_t work_bytes, hipStream_t st);

// Centred Grams on the int8 matrix cores, over packed bit rows.  A column layout is
// either the full one (seg == NULL: segment q = 1920-block q, words 60q .. 60q+59,
// with the column mask applied) or a compacted one (segments of 3 ints: word offset,
// word count, block; segments of whole words, contiguous from word 0, the row stride a
// whole number of 4-word groups; rows hold only the kept columns, the rest of every
// word zero; mask NULL or of the same layout).  seg_counts: cnt[q*ld + i] = ones of
// (row rows[i] & mask) in segment q, indexed by list position.  gram_mfma: output
// and semantics as gram() with centred = 1; ca/cz are seg_counts of alist/zlist.
struct GramMArgs {
  const uint8_t* A; size_t rbA; const int* alist; int na;
  const uint8_t* Z; size_t rbZ; const int* zlist; int nz;
  const uint32_t* mask; const int* seg; int nseg;
  const int* ca; long long ldca; const int* cz; long long ldcz;
  const double* ajdk; int pcols; const double* base16; const double* xa; const double* xz;
  double* out; long long ldo; int a_orig; int z_orig; int op_sub;
};
void seg_counts(const Bits& X, const int* rows, int nrows, const uint32_t* mask, const int* seg, int nseg,
                int* cnt, long long ld, hipStream_t st);
void gram_mfma(const GramMArgs& g, hipStream_t st);
// The same Gram split over the 1920-blocks for few output tiles (full layout only):
// gram_split_bytes(g) = scratch bytes it needs, 0 when the split does not apply.
size_t gram_split_bytes(const GramMArgs& g);
size_t gram_split_bytes_any(const GramMArgs& g);   // the same without the tile-count rule
void gram_split(const GramMArgs& g, uint32_t* X, hipStream_t st);
// sg[t*nzp + i] = S_t(zlist[i]) (+-1) for the QR sums; returns nzp.  sgbuf: qr_sign_bytes(nz, nt)
int qr_signs(const Bits& Sc, const int* tlist, int nt, const int* zlist, int nz, signed char* sgbuf, hipStream_t st);

// bcc_scramble's redrawn bits: bit q of bits[u*ceil(nc/32) + q/32] -> An(rows[u], cols[q]) and
// At(cols[q], rows[u]); rows and cols unique
void scramble_bits(uint8_t* An, size_t rbA, uint8_t* At, size_t rbT, const int* rows, int nr, const int* cols, int nc,
                   const uint32_t* bits, hipStream_t st);

// plain +-1 products with a few (<= kMaxT) covariate rows, as gram() with centred = 0,
// a_orig = z_orig = 1: out[tlist[t]*ldo + alist[i]] = base16[15] - 2 pop((A_i ^ T_t) & mask)
// over pcols 1920-blocks, or over nwords 32-bit words when nwords >= 0 (packed rows)
void pm_few(const Bits& A, const int* alist, int na, const Bits& T, const int* tlist, int nt,
            const uint32_t* mask, int pcols, const double* base16, double* out, long long ldo, hipStream_t st,
            int nwords = -1);

// dst row i = src row rows[i] (rows[i] < 0: zeros); dst row t, bit i = src row t, bit idx[i] (rows of rbd bytes)
void gather_rows(const Bits& src, const int* rows, int n, uint8_t* dst, hipStream_t st);
void gather_bits(const Bits& src, const int* idx, int n, uint8_t* dst, size_t rbd, hipStream_t st);

// Y (X.nc rows of rbY = bsize(X.nr) bytes, zero padded by the caller) = transpose of X
void transpose_bits(const Bits& X, uint8_t* Y, size_t rbY, hipStream_t st);
// synthetic pools: rows (nrows x rb bytes, rb % 4 == 0) with entry (r, j) = 1 with probability
// p_j = pmin + (pmax - pmin) U(seed, ~0, j), drawn as U(seed, r, j) < p_j (pbuf: ncols doubles);
// then blocks overwritten with given bits
void pool_draw(uint8_t* rows, size_t rb, int nrows, int ncols, double pmin, double pmax, unsigned long long seed,
               double* pbuf, hipStream_t st);
void pool_plant(uint8_t* rows, size_t rb, const int* prow, const int* pcol, const uint8_t* bits, int m, int mc,
                hipStream_t st);

// stable ascending radix sort of (keys[i], i): keys_out, idx_out; temp of sort_pairs_temp_bytes(n)
size_t sort_pairs_temp_bytes(int n);
hipError_t sort_pairs(const double* keys, double* keys_out, int* idx_in, int* idx_out, int n,
                      void* temp, size_t temp_bytes, hipStream_t st);

// out[i] = pop(X[rows[i]] & mask) over nwords 32-bit words (rows==NULL: row i)
void popc_rows(const Bits& X, const int* rows, int nrows, const uint32_t* mask, int nwords, int* out, hipStream_t st);

// launch-shape autotuner (lkc_tune.cpp): per shape class, the fastest of a kernel's
// bit-identical forms, measured on first use.  mode 0 off, 1 measure new classes, 2 frozen,
// 3 measure and check that every form wrote the same bits as the first.
struct TuneKey {
  int op, a, b, c, d, e;
  bool operator<(const TuneKey& o) const
  {
    const int x[6] = { op, a, b, c, d, e }, y[6] = { o.op, o.a, o.b, o.c, o.d, o.e };
    for (int i = 0; i < 6; i++) if (x[i] != y[i]) return x[i] < y[i];
    return false;
  }
};
enum { kTuneWw = 1, kTuneGram = 2 };
int tune_mode();
void tune_set_mode(int m);
inline bool tune_measuring(int mode) { return mode == 1 || mode == 3; }
int tune_bucket(long long x);            // round(2 log2 x): half-octave size classes
bool tune_get(const TuneKey& k, int* choice);
void tune_put(const TuneKey& k, int choice, float best_ms, float model_ms);
// runs launch(ctx, cand[i]) for every candidate twice on st between HIP events, waits, and
// returns the index of the fastest; model_idx is the cost model's candidate.  In mode 3 the
// out_bytes at out (what every form writes) are compared after each launch with the first
// form's, the mismatching 8-byte words counted (tune_verify_stats)
int tune_measure(hipStream_t st, const int* cand, int n, int model_idx, void (*launch)(void*, int), void* ctx,
                 float* best_ms, float* model_ms, const void* out = nullptr, size_t out_bytes = 0);
void tune_verify_stats(long long* compared, long long* mismatched);
int tune_entries();
void tune_stats(long long* measured, double* saved_ms);
void tune_clear();
int tune_save(const char* path);
int tune_load(const char* path);

}  // namespace lkc
