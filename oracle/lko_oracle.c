/*
 * lko_oracle.c -- CPU oracle for the lakcluster `bcc` elimination path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path (see the
 * header).  Single bin, no Y/W, QR/QC strategy "YnWt" (store all): the problem from
 * the .in's files (bcc_load) or drawn by bcc_init (bcc.c:1542-1600, st_init); the bcc
 * elimination, the one-bin dcc elimination, bcc_scramble, and the scorebox replays of
 * both (lakcluster_scorebox, dexcluster_scorebox).
 *
 * Conventions
 *   - bit matrices are row-major packed bit rows, most significant bit first
 *     in each byte, bit 1 <=> +1 (lakcluster_header.h:24-26); each row is
 *     lko_bsize(ncols) bytes so 1920-bit weighting blocks are whole bytes.
 *   - all dense score arrays are stored with original ("a") indices; entries
 *     outside the active masks are zero, exactly as the reference's memsets
 *     leave them.  Storage spacing never changes a value.
 *   - double expressions keep the reference's left-to-right order; the file
 *     must be compiled without FMA contraction (-ffp-contract=off).
 */
#include "lko_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <errno.h>

#define MAXI(a,b) ((a)>(b)?(a):(b))
#define MINI(a,b) ((a)<(b)?(a):(b))

size_t lko_bsize(int n)
{
  long long e = n + ((16 - (n % 16)) % 16);
  long long r = ((e + LKO_POP - 1) / LKO_POP) * LKO_POP;
  return (size_t)(r / 8);
}
int lko_psize(int n) { return (int)(lko_bsize(n) / LKO_POPB); }

static inline int bget(const uint8_t *b, long i) { return (b[i >> 3] >> (7 - (i & 7))) & 1; }
static inline void bset(uint8_t *b, long i) { b[i >> 3] |= (uint8_t)(1u << (7 - (i & 7))); }
static inline int bpm(const uint8_t *b, long i) { return bget(b, i) ? 1 : -1; }

static void *xcalloc(size_t n, size_t s)
{
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "lko: out of memory (%zu x %zu)\n", n, s); exit(1); }
  return p;
}

static inline int pop64(uint64_t x) { return __builtin_popcountll(x); }
static inline uint64_t ld64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

/* popcount of (x & m) over nbytes (nbytes multiple of 8) */
static long pop_and(const uint8_t *x, const uint8_t *m, size_t nbytes)
{
  long s = 0;
  for (size_t i = 0; i < nbytes; i += 8) s += pop64(ld64(x + i) & ld64(m + i));
  return s;
}

/* popcount_pm0 (popcount.c:425-489): sum over mask&x2 of (x1 ? +1 : -1) */
static long pm0(const uint8_t *x1, const uint8_t *x2, const uint8_t *m, size_t nbytes)
{
  long p = 0, q = 0;
  for (size_t i = 0; i < nbytes; i += 8) {
    uint64_t a = ld64(x1 + i), b = ld64(x2 + i) & ld64(m + i);
    p += pop64(a & b); q += pop64(~a & b);
  }
  return p - q;
}

/* popcount_pmpm0 (popcount.c:492-553): sum over mask&x2 of (x0==x1 ? +1 : -1) */
static long pmpm0(const uint8_t *x0, const uint8_t *x1, const uint8_t *x2, const uint8_t *m, size_t nbytes)
{
  long p = 0, q = 0;
  for (size_t i = 0; i < nbytes; i += 8) {
    uint64_t e = ld64(x0 + i) ^ ld64(x1 + i), b = ld64(x2 + i) & ld64(m + i);
    p += pop64(~e & b); q += pop64(e & b);
  }
  return p - q;
}

/* popcount_notxorxor (popcount.c:198-249): positions of mask where x1^xs^x2 == 0 */
static long pop_notxorxor(const uint8_t *x1, const uint8_t *xs, const uint8_t *x2, const uint8_t *m, size_t nbytes)
{
  long s = 0;
  for (size_t i = 0; i < nbytes; i += 8) s += pop64(~(ld64(x1 + i) ^ ld64(xs + i) ^ ld64(x2 + i)) & ld64(m + i));
  return s;
}

/* ------------------------------------------------------------------ */
/* scalar helpers                                                      */
/* ------------------------------------------------------------------ */

/* get_xdrop (xdrop.c:5-64) */
void lko_get_xdrop(double lrij, double lcij, double gamma, int gamma_type, int *rdrop_p, int *cdrop_p)
{
  double dbl1 = 1.000000 - 0.000001;
  double gtr = 0, atr = 0, gtc = 0, atc = 0; int rdrop = 0, cdrop = 0;
  if (lrij > lcij) { lko_get_xdrop(lcij, lrij, gamma, gamma_type, cdrop_p, rdrop_p); return; }
  if (lrij <= 2) { rdrop = (int)lrij; cdrop = (int)lcij; }
  else {
    if (gamma > 0) { gtr = MAXI(gamma, (dbl1) / MAXI(1, lrij)); atr = MINI(1 - gamma, (lrij - 1) / MAXI(1, lrij)); }
    else { gtr = (dbl1) / MAXI(1, lrij); atr = (lrij - 1) / MAXI(1, lrij); }
    switch (gamma_type) {
    case 0: atc = exp(log(atr) * log(lcij) / MAXI(1, log(MAXI(1, lrij)))); break;
    case 1: atc = sqrt(atr * lrij) * sqrt(lcij) / MAXI(1, sqrt(MAXI(1, lrij))) / ((double)MAXI(1, lcij)); break;
    case 2: atc = atr; break;
    case 3: atc = 0; break;
    default: break;
    }
    gtc = dbl1 - atc; rdrop = (int)ceil(gtr * lrij); cdrop = (int)ceil(gtc * lcij);
    if (gamma_type == 3) cdrop = 0;
  }
  rdrop = (int)MINI((double)rdrop, lrij); cdrop = (int)MINI((double)cdrop, lcij);
  if (rdrop_p) *rdrop_p = rdrop;
  if (cdrop_p) *cdrop_p = cdrop;
}

/* calc_A_ajdk (An_ajdk_v.c:5-30): alpha = p-(1-p), D = 1/max(0.01,4p(1-p)), libm pow */
void lko_calc_ajdk(const double *p, int pcols, double *out)
{
  for (int j = 0; j < 5; j++) for (int k = 0; k < 3; k++) {
    int jk = j + 5 * k;
    for (int nc = 0; nc < pcols; nc++)
      out[nc + jk * pcols] = pow(p[nc] - (1 - p[nc]), j) * pow(1.0 / MAXI(0.01, (4.0 * p[nc] * (1 - p[nc]))), k);
  }
}

/* dPartition_xij / dQuickSort_xij (Quicksort.c:83-141), recursion limit 32768 */
static int part_xij(double *d, int **ix, int l, int r)
{
  double pivot = d[l], td; int i = l, j = r + 1, t;
  for (;;) {
    do { i++; } while (i <= r && d[i] <= pivot);
    do { j--; } while (d[j] > pivot);
    if (i >= j) break;
    td = d[i]; d[i] = d[j]; d[j] = td;
    for (int q = 0; q < 6; q++) if (ix[q]) { t = ix[q][i]; ix[q][i] = ix[q][j]; ix[q][j] = t; }
  }
  td = d[l]; d[l] = d[j]; d[j] = td;
  for (int q = 0; q < 6; q++) if (ix[q]) { t = ix[q][l]; ix[q][l] = ix[q][j]; ix[q][j] = t; }
  return j;
}
static unsigned int qs_xij(unsigned int nn, double *d, int **ix, int l, int r)
{
  unsigned int n1 = nn, n2 = nn;
  if (l < r) {
    int j = part_xij(d, ix, l, r);
    if (nn < 32768u) { n1 = qs_xij(nn + 1, d, ix, l, j - 1); n2 = qs_xij(nn + 1, d, ix, j + 1, r); }
  }
  return MAXI(n1, n2);
}
unsigned int lko_quicksort_xij(double *d, int *i0, int *i1, int *i2, int *i3, int *i4, int *i5, int l, int r)
{
  int *ix[6] = { i0, i1, i2, i3, i4, i5 };
  return qs_xij(0, d, ix, l, r);
}

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */

/* per-block weighted popcounts: sum_b w[b] * cnt_b(x & m), in block order
   (popcount_lf, popcount.c:252-306) */
static double lf_and(const uint8_t *x, const uint8_t *m, int pcols, const double *w)
{
  double dout = 0;
  for (int b = 0; b < pcols; b++) {
    long c = pop_and(x + (size_t)b * LKO_POPB, m + (size_t)b * LKO_POPB, LKO_POPB);
    dout += (double)w[b] * (double)((unsigned int)c);
  }
  return dout;
}
/* popcount_xor_lf (popcount.c:310-364) */
static double lf_xor(const uint8_t *x, const uint8_t *z, const uint8_t *m, int pcols, const double *w)
{
  double dout = 0;
  for (int b = 0; b < pcols; b++) {
    long c = 0; size_t o = (size_t)b * LKO_POPB;
    for (int i = 0; i < LKO_POPB; i += 8) c += pop64((ld64(x + o + i) ^ ld64(z + o + i)) & ld64(m + o + i));
    dout += (double)w[b] * (double)((unsigned int)c);
  }
  return dout;
}
/* sum_b w[b]*|m block b| in block order (dtmp_base, An_ajdk_v.c:312; dtmp_a0d1, AnZt_vv.c:62) */
static double mask_lf(const uint8_t *m, int pcols, const double *w)
{
  double s = 0;
  for (int b = 0; b < pcols; b++) {
    long c = 0; const uint8_t *p = m + (size_t)b * LKO_POPB;
    for (int i = 0; i < LKO_POPB; i++) c += __builtin_popcount(p[i]);
    s += (double)(w ? w[b] : 1.0) * (double)c;
  }
  return s;
}

/* get_An_ajdk_v (An_ajdk_v.c:272-363), TYPE_pm: out = -base + 2*lf */
void lko_an_ajdk_v(const uint8_t *X, size_t rb, const int *rows, int nrows,
                   const uint8_t *colmask, int ncols, const double *ajdk, double *out)
{
  int pcols = lko_psize(ncols);
  for (int jk = 0; jk < LKO_AJDK; jk++) {
    const double *w = ajdk + (size_t)jk * pcols;
    double base = mask_lf(colmask, pcols, w);
    for (int i = 0; i < nrows; i++) {
      double v = -base;
      v += (2 * lf_and(X + (size_t)rows[i] * rb, colmask, pcols, w));
      out[(size_t)i * LKO_AJDK + jk] = v;
    }
  }
}

/* get_AnZt_vv__run (AnZt_vv.c:62-95), TYPE_pm */
void lko_anzt_vv(const uint8_t *A, size_t rbA, const int *arows, int na,
                 const uint8_t *Z, size_t rbZ, const int *zrows, int nz,
                 const uint8_t *colmask, int ncols, const double *ajdk,
                 const double *xa, const double *xz, double *out)
{
  int pcols = lko_psize(ncols);
  double a0d1 = mask_lf(colmask, pcols, ajdk ? ajdk + 5 * pcols : NULL);
  double a2d1 = 0;
  if (ajdk) a2d1 = mask_lf(colmask, pcols, ajdk + 7 * pcols);
  size_t nbytes = (size_t)pcols * LKO_POPB;
  #pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < na; i++) {
    const uint8_t *x = A + (size_t)arows[i] * rbA;
    for (int j = 0; j < nz; j++) {
      const uint8_t *z = Z + (size_t)zrows[j] * rbZ;
      double o = a0d1, dtmp;
      if (ajdk) dtmp = lf_xor(x, z, colmask, pcols, ajdk + 5 * pcols);
      else {
        long c = 0;
        for (size_t q = 0; q < nbytes; q += 8) c += pop64((ld64(x + q) ^ ld64(z + q)) & ld64(colmask + q));
        dtmp = (double)c;
      }
      o -= (2 * dtmp);
      if (ajdk) o -= xa[i] + xz[j] - a2d1;
      out[(size_t)i * nz + j] = o;
    }
  }
}

void lko_planes_free(lko_planes *p) { if (p) { free(p->row0); free(p->planes); p->row0 = NULL; p->planes = NULL; } }

/* get_xcalc (xcalc.c:5-246), trm handled by the caller through L's layout.
   L[i*nw + j]: value at (n-row nrows_list[i], w j).  Iteration order of the
   min/max scan follows xcalc_setmaxmin (w outer, n inner). */
void lko_xcalc(const double *L, const int *nlist, int nn, int n_total, int nw,
               int B, long long D_MLT, lko_planes *P, unsigned long long *cz_out)
{
  double min_data = 0, max_data = 0; int m_set = 0;
  for (int j = 0; j < nw; j++) for (int i = 0; i < nn; i++) {
    double v = L[(size_t)i * nw + j];
    if (!m_set) { min_data = v; max_data = v; m_set = 1; }
    else { min_data = (v < min_data) ? v : min_data; max_data = (v > max_data) ? v : max_data; }
  }
  if (min_data >= max_data) max_data = min_data + 1;
  double mn = min_data - 0.0625 * (max_data - min_data);
  double mx = max_data + 0.0625 * (max_data - min_data);
  double mlt = (double)(D_MLT) / (mx - mn);
  P->B = B; P->n_total = n_total; P->rb = lko_bsize(n_total); P->nw = nw;
  P->row0 = (uint8_t *)xcalloc(P->rb, 1);
  P->planes = (uint8_t *)xcalloc((size_t)nw * B * P->rb + 1, 1);
  for (int i = 0; i < nn; i++) bset(P->row0, nlist[i]);
  for (int j = 0; j < nw; j++) for (int i = 0; i < nn; i++) {
    double v = L[(size_t)i * nw + j];
    unsigned long long cz = (unsigned long long)(mlt * (double)(v - mn));
    if (cz < 1) cz = 1;
    if (cz > (unsigned long long)(D_MLT - 1)) cz = (unsigned long long)(D_MLT - 1);
    if (cz_out) cz_out[(size_t)j * nn + i] = cz;
    for (int nb = 0; nb < B; nb++) if ((cz >> nb) & 1ull) bset(P->planes + ((size_t)j * B + nb) * P->rb, nlist[i]);
  }
  P->min_d = mn; P->max_d = mx; P->mlt_d = mlt;
  if (!isfinite(P->mlt_d) || P->mlt_d <= 0) P->mlt_d = 1;
}

/* ------------------------------------------------------------------ */
/* bit-matrix handle used by the driver                               */
/* ------------------------------------------------------------------ */
typedef struct { uint8_t *b; int nr, nc; size_t rb; } bm_t;

static bm_t bm_alloc(int nr, int nc)
{
  bm_t m; m.nr = nr; m.nc = nc; m.rb = lko_bsize(nc);
  m.b = (uint8_t *)xcalloc((size_t)MAXI(nr, 1) * MAXI(m.rb, 8) + 64, 1);
  return m;
}
static inline const uint8_t *bm_row(const bm_t *m, int r) { return m->b + (size_t)r * m->rb; }

/* a list of ascending original indices */
typedef struct { int *ix; int n; } ilist;
static ilist il_from_mask(const uint8_t *mask, int n)
{
  ilist L; L.ix = (int *)xcalloc(n + 1, sizeof(int)); L.n = 0;
  for (int i = 0; i < n; i++) if (bget(mask, i)) L.ix[L.n++] = i;
  return L;
}

/* ------------------------------------------------------------------ */
/* get_At_T_YnWt_S_Zn_ww (At_T_YnWt_S_Zn_ww.c:44-110)                  */
/*   P rows (k) over the n dimension with mask pmask; covariate rows Tc */
/*   over n; planes X over n with w list; Sc (covariates over w) and Q  */
/*   (k rows over w).  t/s lists walk together, as do k/q lists.       */
/*   out[k*out_ld + tlist[t]] (a-spacing in both indices).             */
/* ------------------------------------------------------------------ */
static void ww_kernel(const bm_t *P, const int *klist, const int *qlist, int nk, const uint8_t *pmask,
                      const bm_t *Tc, const int *tlist, const int *slist, int nt,
                      const lko_planes *X, const int *wlist,
                      const bm_t *Sc, const bm_t *Q, const double *ajdk, int pcols,
                      double *out, int out_ld)
{
  const double *a_An = ajdk + 1 * (size_t)pcols;
  size_t nb8 = X->rb; int B = X->B, nw = X->nw;
  double mn = X->min_d, mlt = X->mlt_d;
  double *H = (double *)xcalloc((size_t)nw + 1, sizeof(double));
  long *cm = (long *)xcalloc((size_t)nw * B + 1, sizeof(long));
  for (int t = 0; t < nt; t++) {
    const uint8_t *Tr = bm_row(Tc, tlist[t]);
    const uint8_t *Sr = bm_row(Sc, slist[t]);
    long llb = pmpm0(X->row0, Tr, X->row0, pmask, nb8);
    double dtmp = llb * 1.0;
    double bp = dtmp * mn * mlt;
    for (int j = 0; j < nw; j++) {
      double h = 0; long long n2 = 1;
      for (int nb = 0; nb < B; nb++) {
        const uint8_t *pl = X->planes + ((size_t)j * B + nb) * nb8;
        long ll = pmpm0(X->row0, Tr, pl, pmask, nb8); double d = ll * 1.0;
        h += n2 * d; n2 *= 2;
        cm[(size_t)j * B + nb] = pop_and(pl, pmask, nb8);
      }
      H[j] = h;
    }
    #pragma omp parallel
    {
      uint8_t *u = (uint8_t *)xcalloc(nb8 + 8, 1);
      #pragma omp for schedule(dynamic,4)
      for (int k = 0; k < nk; k++) {
        const uint8_t *Pr = bm_row(P, klist[k]);
        const uint8_t *Qr = bm_row(Q, qlist[k]);
        /* u = mask & (P xnor T): the +1 positions of pmpm0 */
        for (size_t i = 0; i < nb8; i += 8) { uint64_t e = ~(ld64(Pr + i) ^ ld64(Tr + i)) & ld64(pmask + i); memcpy(u + i, &e, 8); }
        long llk = pmpm0(Pr, Tr, X->row0, pmask, nb8);
        double bk = llk * mn * mlt;
        double alpha_k = a_An[klist[k] / LKO_POP], alpha_q = a_An[qlist[k] / LKO_POP];
        double o = 0;
        for (int j = 0; j < nw; j++) {
          int vS = bpm(Sr, wlist[j]), vZ = bpm(Qr, wlist[j]);
          double g = 0; long long n2 = 1;
          for (int nb = 0; nb < B; nb++) {
            const uint8_t *pl = X->planes + ((size_t)j * B + nb) * nb8;
            long ll = 2 * pop_and(u, pl, nb8) - cm[(size_t)j * B + nb];
            g += n2 * ll; n2 *= 2;
          }
          double at_base = bp * alpha_k;
          double at_tmp = H[j] * alpha_k;
          double tmp = (double)(bk + g) / mlt - (double)(at_base + at_tmp) / mlt;
          o += tmp * vS * (vZ - alpha_q);
        }
        out[(size_t)klist[k] * out_ld + tlist[t]] = o;
      }
      free(u);
    }
  }
  free(H); free(cm);
}

/* get_An_ZtSn_ww (An_ZtSn_ww.c:44-70): out[a*out_ld + slist[s]] */
static void an_ztsn_ww(const bm_t *A, const int *alist, int na, const uint8_t *amask,
                       const lko_planes *X, const int *slist, int ns, double *out, int out_ld)
{
  size_t nb8 = X->rb; int B = X->B;
  #pragma omp parallel for schedule(dynamic,8)
  for (int a = 0; a < na; a++) {
    const uint8_t *Ar = bm_row(A, alist[a]);
    double ll = (double)pm0(Ar, X->row0, amask, nb8);
    double base = ll * X->min_d * X->mlt_d;
    for (int s = 0; s < ns; s++) {
      double tmp = 0; long long n2 = 1;
      for (int nb = 0; nb < B; nb++) {
        double l2 = (double)pm0(Ar, X->planes + ((size_t)s * B + nb) * nb8, amask, nb8);
        tmp += n2 * l2; n2 *= 2;
      }
      out[(size_t)alist[a] * out_ld + slist[s]] = (double)(base + tmp) / X->mlt_d;
    }
  }
}

/* get_AnZt_S_WnYt_vv (AnZt_S_WnYt_vv.c:49-67): out[a*out_ld + s] = sum_z (X*vS)*Y */
static void qr_kernel(const int *alist, int na, const int *zlist, int nz,
                      const bm_t *Sc, const int *slist, int ns,
                      const double *Xm, const double *Ym, int ld, double *out, int out_ld)
{
  #pragma omp parallel for schedule(dynamic,16)
  for (int a = 0; a < na; a++) {
    for (int s = 0; s < ns; s++) {
      const uint8_t *Sr = bm_row(Sc, slist[s]);
      double tmp = 0;
      for (int z = 0; z < nz; z++) {
        size_t o = (size_t)alist[a] * ld + zlist[z];
        int vS = bpm(Sr, zlist[z]);
        tmp += Xm[o] * vS * Ym[o];
      }
      out[(size_t)alist[a] * out_ld + slist[s]] = tmp;
    }
  }
}

/* get_AtTYn_vv (AtTYn_vv.c:79-120): out[(a*ny + y)*nt + t] in compact indices */
static void attyn_kernel(const bm_t *At, const int *alist, int na, const uint8_t *amask, int acpop,
                         const bm_t *Tc, const int *tlist, int nt, const uint8_t *tmask, int tcpop,
                         const bm_t *Yt, const int *ylist, int ny,
                         const double *lf_AtTn, const double *lf_YtTn, int tld,
                         const double *ajdk, int pcols, double *out)
{
  const double *aA = ajdk + 1 * (size_t)pcols;
  size_t nb8 = At->rb;
  for (int t = 0; t < nt; t++) {
    const uint8_t *Tr = bm_row(Tc, tlist[t]);
    long long hhay = -(long long)tcpop + 2 * pop_and(Tr, tmask, nb8);
    #pragma omp parallel for schedule(dynamic,8)
    for (int a = 0; a < na; a++) {
      const uint8_t *Ar = bm_row(At, alist[a]);
      long long hhAy = (long long)lf_AtTn[(size_t)alist[a] * tld + tlist[t]];
      for (int y = 0; y < ny; y++) {
        long long hhaY = (long long)lf_YtTn[(size_t)ylist[y] * tld + tlist[t]];
        long long hhAY = +(long long)acpop;
        hhAY -= (long long)(2 * pop_notxorxor(Ar, Tr, bm_row(Yt, ylist[y]), amask, nb8));
        double aa = aA[alist[a] / LKO_POP], ay = aA[ylist[y] / LKO_POP];
        out[((size_t)a * ny + y) * nt + t] = (hhAY - hhAy * ay - aa * hhaY + aa * hhay * ay);
      }
    }
  }
}

/* get_AtTYn____WtSZn_vv (AtTYn____WtSZn_vv.c:60-80):
   out[k*out_ld + tlist[t]] = sum_y (L1[k,y,t]*D_y)*L2[k,y,t] (L1/L2 compact [k][y][t]) */
static void attyn_wtszn_kernel(const int *klist, int nk, const int *ylist, int ny, const int *tlist, int nt,
                               const double *L1, const double *L2, const double *ajdk, int pcols,
                               double *out, int out_ld)
{
  const double *D = ajdk + 5 * (size_t)pcols;
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < nk; k++) {
    for (int t = 0; t < nt; t++) {
      double o = 0;
      for (int y = 0; y < ny; y++) {
        size_t q = ((size_t)k * ny + y) * nt + t;
        o += L1[q] * D[ylist[y] / LKO_POP] * L2[q];
      }
      out[(size_t)klist[k] * out_ld + tlist[t]] = o;
    }
  }
}

/* ------------------------------------------------------------------ */
/* .b16 input (binary_read.c:5-55, M_handle_make/M_load M_init.c:143-242) */
/*   A file holding an (fnr x fnc) matrix column-major yields a handle  */
/*   with fnc rows of fnr bits: handle row r = file column r.           */
/* ------------------------------------------------------------------ */
static int b16_read(const char *path, bm_t *out)
{
  FILE *fp = fopen(path, "rb");
  if (!fp) { fprintf(stderr, "lko: cannot open %s\n", path); return -1; }
  int hdr[3];
  if (fread(hdr, sizeof(int), 3, fp) != 3) { fclose(fp); return -1; }
  int bitj = hdr[0], fnr = hdr[1], fnc = hdr[2];
  if (bitj <= 0) bitj = 16;
  size_t wp = (size_t)(fnr + (bitj - (fnr % bitj)) % bitj) / 8;
  *out = bm_alloc(fnc, fnr);
  for (int r = 0; r < fnc; r++) {
    if (wp && fread(out->b + (size_t)r * out->rb, 1, wp, fp) != wp) { fclose(fp); return -1; }
  }
  fclose(fp);
  return 0;
}
/* a mask file is an (n x 1) .b16 */
static int mask_read(const char *path, int n, uint8_t *mask)
{
  bm_t m; if (b16_read(path, &m)) return -1;
  if (m.nc != n) fprintf(stderr, "lko: mask %s has %d entries, expected %d\n", path, m.nc, n);
  size_t nb = lko_bsize(n);
  memset(mask, 0, nb);
  if (m.nr > 0) memcpy(mask, m.b, MINI(nb, m.rb));
  /* keep only the n valid bits */
  for (long i = n; i < (long)nb * 8; i++) mask[i >> 3] &= (uint8_t)~(1u << (7 - (i & 7)));
  free(m.b);
  return 0;
}

/* ------------------------------------------------------------------ */
/* configuration from the key/value .in file (updateglobals.c:5-199)  */
/* ------------------------------------------------------------------ */
#define LKO_PATH 4352
typedef struct {
  int B; long long D_MLT; double gamma; int gamma_type; double kappa2; int sparse; int nbins; int Ireq;
  int A_ncols, Y_ncols, T_ncols;
  char QR[LKO_PATH], QC[LKO_PATH];
  char A_n[LKO_PATH], A_t[LKO_PATH], A_rind[LKO_PATH], A_cind[LKO_PATH], A_p[LKO_PATH];
  char Z_n[LKO_PATH], Z_t[LKO_PATH], Z_rind[LKO_PATH];
  char T_n[LKO_PATH], T_t[LKO_PATH], T_cind[LKO_PATH], S_n[LKO_PATH], S_t[LKO_PATH];
  char xpre[LKO_PATH], dir_name[LKO_PATH], dir_base[LKO_PATH], out_name[LKO_PATH];
  int scramble_num;
  /* bcc_scramble inputs: out_xdrop files and seeds (seed n defaults to n + 1) */
  char scr_files_raw[LKO_PATH], scr_seeds_raw[LKO_PATH];
  /* lakcluster_scorebox inputs (updateglobals.c:136-140) */
  char test_type[LKO_PATH], sb_file[LKO_PATH];
  int sb_row_max, sb_row_num, sb_col_max, sb_col_num;
  /* the self-tests' synthetic problem (GLOBAL_TEST_*, defaults of lakcluster_ver18.c:70-79) */
  char test_typ2[LKO_PATH]; double test_mrand;
  int test_A_rows, test_A_cols, test_Z_rows, test_Y_cols, test_T_cols, test_niter;
} lko_cfg;

static void trim(char *s)
{
  size_t n = strlen(s);
  while (n && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\n' || s[n - 1] == '\r')) s[--n] = 0;
  size_t i = 0; while (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r') i++;
  if (i) memmove(s, s + i, strlen(s + i) + 1);
}
static void set_path(const lko_cfg *c, char *dst, const char *v)
{
  if (c->xpre[0]) snprintf(dst, LKO_PATH, "%s/%s", c->xpre, v); else snprintf(dst, LKO_PATH, "%s", v);
}

/* returns 0, or -1 on an unknown key (the reference exits with RET_READ_FAIL there) */
static int cfg_parse(const char *text, lko_cfg *c)
{
  memset(c, 0, sizeof(*c));
  c->B = 32; c->D_MLT = 4294967296LL; c->gamma = 0.125; c->gamma_type = 0; c->sparse = 1; c->nbins = 0;
  c->test_A_rows = 12; c->test_A_cols = 13; c->test_Z_rows = 15; c->test_Y_cols = 11; c->test_T_cols = 3; c->test_niter = 2;
  const char *p = text;
  while (*p) {
    const char *eq = strchr(p, '='); if (!eq) break;
    char name[256]; size_t nl = (size_t)(eq - p); if (nl > 255) nl = 255;
    memcpy(name, p, nl); name[nl] = 0; trim(name);
    const char *v0 = eq + 1; const char *semi = strchr(v0, ';'); if (!semi) semi = v0 + strlen(v0);
    char val[LKO_PATH]; size_t vl = (size_t)(semi - v0); if (vl >= LKO_PATH) vl = LKO_PATH - 1;
    memcpy(val, v0, vl); val[vl] = 0; trim(val);
    p = (*semi) ? semi + 1 : semi;
    if (!strcmp(name, "END")) break;
    /* list-valued keys take the first (only) bin */
    char first[LKO_PATH]; snprintf(first, sizeof first, "%s", val); { char *cm = strchr(first, ','); if (cm) *cm = 0; } trim(first);
    if (!strcmp(name, "GLOBAL_B_MLT")) { c->B = atoi(val); c->D_MLT = (long long)pow(2.0, (double)c->B); }
    else if (!strcmp(name, "GLOBAL_D_MLT")) { c->D_MLT = atoll(val); c->B = (int)llround(log2((double)c->D_MLT)); }
    else if (!strcmp(name, "GLOBAL_gamma")) c->gamma = atof(val);
    else if (!strcmp(name, "GLOBAL_gamma_type")) c->gamma_type = atoi(val);
    else if (!strcmp(name, "GLOBAL_kappa_squared")) c->kappa2 = atof(val);
    else if (!strcmp(name, "GLOBAL_TEST_sparse")) c->sparse = atoi(val);
    else if (!strcmp(name, "GLOBAL_NBINS")) c->nbins = atoi(val);
    else if (!strcmp(name, "GLOBAL_Ireq")) c->Ireq = atoi(val);
    else if (!strcmp(name, "GLOBAL_QR_strategy")) snprintf(c->QR, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_QC_strategy")) snprintf(c->QC, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_A_n_cols")) c->A_ncols = atoi(val);
    else if (!strcmp(name, "GLOBAL_Y_n_cols")) c->Y_ncols = atoi(val);
    else if (!strcmp(name, "GLOBAL_T_n_cols")) c->T_ncols = atoi(val);
    else if (!strcmp(name, "GLOBAL_DIR_XPRE")) snprintf(c->xpre, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_DIR_NAME")) snprintf(c->dir_name, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_DIR_BASE")) snprintf(c->dir_base, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_out_name")) snprintf(c->out_name, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_A_n_name_")) set_path(c, c->A_n, first);
    else if (!strcmp(name, "GLOBAL_A_t_name_")) set_path(c, c->A_t, first);
    else if (!strcmp(name, "GLOBAL_A_n_rind_")) set_path(c, c->A_rind, first);
    else if (!strcmp(name, "GLOBAL_A_n_cind")) set_path(c, c->A_cind, val);
    else if (!strcmp(name, "GLOBAL_A_p_name")) set_path(c, c->A_p, val);
    else if (!strcmp(name, "GLOBAL_Z_n_name_")) set_path(c, c->Z_n, first);
    else if (!strcmp(name, "GLOBAL_Z_t_name_")) set_path(c, c->Z_t, first);
    else if (!strcmp(name, "GLOBAL_Z_n_rind_")) set_path(c, c->Z_rind, first);
    else if (!strcmp(name, "GLOBAL_T_n_name_")) set_path(c, c->T_n, first);
    else if (!strcmp(name, "GLOBAL_T_t_name_")) set_path(c, c->T_t, first);
    else if (!strcmp(name, "GLOBAL_T_n_cind")) set_path(c, c->T_cind, val);
    else if (!strcmp(name, "GLOBAL_S_n_name_")) set_path(c, c->S_n, first);
    else if (!strcmp(name, "GLOBAL_S_t_name_")) set_path(c, c->S_t, first);
    else if (!strcmp(name, "GLOBAL_scramble_num")) c->scramble_num = atoi(val);
    else if (!strcmp(name, "GLOBAL_scramble_out_xdrop_")) snprintf(c->scr_files_raw, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_scramble_rseed_")) snprintf(c->scr_seeds_raw, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_TEST_TYPE")) snprintf(c->test_type, LKO_PATH, "%s", val);
    else if (!strcmp(name, "GLOBAL_TEST_TYP2")) snprintf(c->test_typ2, LKO_PATH, "%s", first);
    else if (!strcmp(name, "GLOBAL_TEST_mrand")) c->test_mrand = atof(val);
    else if (!strcmp(name, "GLOBAL_TEST_A_n_rows")) c->test_A_rows = atoi(val);
    else if (!strcmp(name, "GLOBAL_TEST_A_n_cols")) c->test_A_cols = atoi(val);
    else if (!strcmp(name, "GLOBAL_TEST_Z_n_rows")) c->test_Z_rows = atoi(val);
    else if (!strcmp(name, "GLOBAL_TEST_Y_n_cols")) c->test_Y_cols = atoi(val);
    else if (!strcmp(name, "GLOBAL_TEST_T_n_cols")) c->test_T_cols = atoi(val);
    else if (!strcmp(name, "GLOBAL_TEST_niter")) c->test_niter = atoi(val);
    else if (!strcmp(name, "GLOBAL_scorebox_out_xdrop")) snprintf(c->sb_file, LKO_PATH, "%s", first);
    else if (!strcmp(name, "GLOBAL_scorebox_row_max")) c->sb_row_max = atoi(val);
    else if (!strcmp(name, "GLOBAL_scorebox_row_num")) c->sb_row_num = atoi(val);
    else if (!strcmp(name, "GLOBAL_scorebox_col_max")) c->sb_col_max = atoi(val);
    else if (!strcmp(name, "GLOBAL_scorebox_col_num")) c->sb_col_num = atoi(val);
    else if (!strncmp(name, "GLOBAL_", 7)) { /* accepted, no effect on this path */ }
    else { fprintf(stderr, "lko: unknown key %s\n", name); return -1; }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* driver state (struct bcc_ajdk/bcc_single/bcc_double, dir_h/bcc.h, one bin) */
/* ------------------------------------------------------------------ */
typedef struct {
  lko_cfg c;
  int NA, NZ, M, T, pcols, has_Z;
  bm_t An, At, Zn, Zt, Tt, St;
  /* packed masks; *_j always a subset of *_b */
  uint8_t *Ar_b, *Ar_j, *Ar_rtn, *Ar_rmv, *Zr_b, *Zr_j, *Ac_b, *Ac_j, *Ac_rtn, *Ac_rmv, *Tc_j;
  double *A_p, *ajdk;
  int trm_AnZt, trm_knzt, trm_hnvt;
  /* persistent arrays, original indices */
  double *lf_An_ajdk, *lf_Zn_ajdk;          /* NA*15, NZ*15 */
  double *lf_AtTn, *lf_ZtSn;                /* M*T */
  double *lf_AnZt, *lf_AnAt;                /* NA*NZ, NA*NA */
  double *lf_AtTAnZtSZn, *lf_AtTAnAtTAn;    /* M*T  (lf_At_T_AnZt_S_Zn, lf_At_T_AnAt_T_An) */
  double *lf_AnZtSZnAt, *lf_AnAtTAnAt;      /* NA*T (lf_AnZt_S_ZnAt, lf_AnAt_T_AnAt) */
  double *At_An_a1d1;                       /* M */
  double *An_a2d2_ZtSn, *An_a2d2_AtTn;      /* NA*T */
  double *a1d2_ZtSn, *a3d2_ZtSn, *a1d2_AtTn, *a3d2_AtTn, *et_Tn, *et_Sn; /* T */
  double *et_An;                            /* M */
  double et_An_a1d1;
  double *T_AnAt_AnAt;                      /* T*NA */
  double *QR_ZSZ, *QR_TAT, *QC_ZSZ, *QC_TAT;            /* NA*T, NA*T, M*T, M*T */
  double *QR_ZSZ_nrm, *QR_TAT_nrm, *QC_ZSZ_nrm, *QC_TAT_nrm;
  /* sort buffers */
  double *QC_sv, *QR_sv; int *QC_a, *QC_b, *QC_j, *QR_a, *QR_b, *QR_j, *QR_nb, *QR_ga, *QR_gb;
  int *Ar_bidx, *Ac_bidx;                   /* b index of every original row/col */
  /* outputs */
  int n_trace; double *trace; int n_xdrop; int *xdrop_a, *xdrop_b;
  long long ops;                            /* reference accounting of bit-ops performed */
} lko_state;

static int popmask(const uint8_t *m, int n) { int s = 0; for (int i = 0; i < n; i++) s += bget(m, i); return s; }
static uint8_t *mask_alloc(int n) { return (uint8_t *)xcalloc(lko_bsize(n) + 8, 1); }
static void mask_and(uint8_t *d, const uint8_t *a, const uint8_t *b, int n) { size_t nb = lko_bsize(n); for (size_t i = 0; i < nb; i++) d[i] = a[i] & b[i]; }

/* gram over lists with optional centring; writes out[a*ld + z] (original indices) */
static void gram_to(lko_state *S, const bm_t *A, const ilist *ar, const bm_t *Z, const ilist *zr,
                    const uint8_t *cmask, int ncols, int centred, const double *xa_full, const double *xz_full,
                    double *out, int ld)
{
  double *xa = NULL, *xz = NULL;
  if (centred) {
    xa = (double *)xcalloc(ar->n + 1, sizeof(double)); xz = (double *)xcalloc(zr->n + 1, sizeof(double));
    for (int i = 0; i < ar->n; i++) xa[i] = xa_full[(size_t)ar->ix[i] * LKO_AJDK + 6];
    for (int j = 0; j < zr->n; j++) xz[j] = xz_full[(size_t)zr->ix[j] * LKO_AJDK + 6];
  }
  double *tmp = (double *)xcalloc((size_t)ar->n * zr->n + 1, sizeof(double));
  lko_anzt_vv(A->b, A->rb, ar->ix, ar->n, Z->b, Z->rb, zr->ix, zr->n, cmask, ncols, centred ? S->ajdk : NULL, xa, xz, tmp);
  for (int i = 0; i < ar->n; i++) for (int j = 0; j < zr->n; j++) out[(size_t)ar->ix[i] * ld + zr->ix[j]] = tmp[(size_t)i * zr->n + j];
  S->ops += (long long)ar->n * zr->n * (long long)lko_bsize(ncols) * 8;
  free(tmp); free(xa); free(xz);
}

/* An_ajdk_v over a row list with a column mask; out[a*15+jk], zero elsewhere */
static void ajdk_to(lko_state *S, const bm_t *X, const ilist *rows, const uint8_t *cmask, int nrows_total, double *out)
{
  memset(out, 0, sizeof(double) * (size_t)nrows_total * LKO_AJDK);
  double *tmp = (double *)xcalloc((size_t)rows->n * LKO_AJDK + 1, sizeof(double));
  lko_an_ajdk_v(X->b, X->rb, rows->ix, rows->n, cmask, X->nc, S->ajdk, tmp);
  for (int i = 0; i < rows->n; i++) memcpy(out + (size_t)rows->ix[i] * LKO_AJDK, tmp + (size_t)i * LKO_AJDK, sizeof(double) * LKO_AJDK);
  S->ops += 15LL * rows->n * (long long)X->rb * 8;
  free(tmp);
}

/* xcalc of the (n x w) block of a full array L[n*ld + w] (trm==0) or L[w*ld + n] (trm==1) */
static void xcalc_from(lko_state *S, const double *L, int ld, int trm, const ilist *nl, int n_total, const ilist *wl, lko_planes *P)
{
  double *tmp = (double *)xcalloc((size_t)nl->n * wl->n + 1, sizeof(double));
  for (int i = 0; i < nl->n; i++) for (int j = 0; j < wl->n; j++)
    tmp[(size_t)i * wl->n + j] = trm ? L[(size_t)wl->ix[j] * ld + nl->ix[i]] : L[(size_t)nl->ix[i] * ld + wl->ix[j]];
  lko_xcalc(tmp, nl->ix, nl->n, n_total, wl->n, S->c.B, S->c.D_MLT, P, NULL);
  S->ops += (long long)nl->n * wl->n * (S->c.B + 1) * 8;
  free(tmp);
}

static void ww_ops(lko_state *S, int nt, int nk, int nw, size_t rb) { S->ops += (long long)nt * nk * nw * (S->c.B + 1) * (long long)rb * 8; }

/* xcalc of a compact (list1 x list2) block L[i*ld + j]; trm==1 reads L[j*ld + i]
   (bcc_M_YnWt_excerpt, bcc_lf.c:304-316) */
static void xcalc_compact(lko_state *S, const double *L, int ld, int trm, const ilist *nl, int n_total, const ilist *wl, lko_planes *P)
{
  double *tmp = (double *)xcalloc((size_t)nl->n * wl->n + 1, sizeof(double));
  for (int i = 0; i < nl->n; i++) for (int j = 0; j < wl->n; j++)
    tmp[(size_t)i * wl->n + j] = trm ? L[(size_t)j * ld + i] : L[(size_t)i * ld + j];
  lko_xcalc(tmp, nl->ix, nl->n, n_total, wl->n, S->c.B, S->c.D_MLT, P, NULL);
  S->ops += (long long)nl->n * wl->n * (S->c.B + 1) * 8;
  free(tmp);
}

/* one patient group of an At_T_YnWt_S_Zn term: SNP-major bits, active-row mask/list, covariates */
typedef struct { const bm_t *Xt; const uint8_t *pmask; ilist prow; const bm_t *Cov; } side_t;

/* bcc_lf_At_T_YnWt_S_Zn_excerpt (bcc_lf.c:615-634) with its xcalc (bcc_M_YnWt_excerpt):
   L is compact over (s1 rows x s2 rows), or full with original indices when full!=0. */
static void ww_excerpt(lko_state *S, int trm, const side_t *s1, const side_t *s2, const ilist *kl, const ilist *tl,
                       const double *L, int ld, int full, double *out)
{
  lko_planes P; memset(&P, 0, sizeof P);
  memset(out, 0, sizeof(double) * (size_t)S->M * S->T);
  const side_t *pn = trm ? s2 : s1;   /* group the planes run over */
  const side_t *pw = trm ? s1 : s2;   /* group of the w index */
  ilist nl = pn->prow, wl = pw->prow;
  double *tmp = (double *)xcalloc((size_t)nl.n * wl.n + 1, sizeof(double));
  for (int i = 0; i < nl.n; i++) for (int j = 0; j < wl.n; j++) {
    int r1 = trm ? j : i, r2 = trm ? i : j;               /* positions in s1 / s2 lists */
    size_t o = full ? (size_t)s1->prow.ix[r1] * ld + s2->prow.ix[r2] : (size_t)r1 * ld + r2;
    tmp[(size_t)i * wl.n + j] = L[o];
  }
  lko_xcalc(tmp, nl.ix, nl.n, pn->Xt->nc, wl.n, S->c.B, S->c.D_MLT, &P, NULL);
  S->ops += (long long)nl.n * wl.n * (S->c.B + 1) * 8;
  free(tmp);
  ww_kernel(pn->Xt, kl->ix, kl->ix, kl->n, pn->pmask, pn->Cov, tl->ix, tl->ix, tl->n, &P, wl.ix,
            pw->Cov, pw->Xt, S->ajdk, S->pcols, out, S->T);
  ww_ops(S, tl->n, kl->n, wl.n, P.rb);
  lko_planes_free(&P);
}

/* bcc_An_ajdk (bcc_lf.c:7-51) */
static void st_ajdk(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA);
  ajdk_to(S, &S->An, &ar, S->Ac_j, S->NA, S->lf_An_ajdk);
  if (S->has_Z) { ilist zr = il_from_mask(S->Zr_j, S->NZ); ajdk_to(S, &S->Zn, &zr, S->Ac_j, S->NZ, S->lf_Zn_ajdk); free(zr.ix); }
  free(ar.ix);
}

/* bcc_lf_ZtSn (bcc_lf.c:53-101): plain +-1 products AtTn[k,t], ZtSn[k,t] */
static void st_ZtSn(lko_state *S)
{
  ilist kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  memset(S->lf_AtTn, 0, sizeof(double) * (size_t)S->M * S->T);
  gram_to(S, &S->At, &kl, &S->Tt, &tl, S->Ar_j, S->NA, 0, NULL, NULL, S->lf_AtTn, S->T);
  if (S->has_Z) {
    memset(S->lf_ZtSn, 0, sizeof(double) * (size_t)S->M * S->T);
    gram_to(S, &S->Zt, &kl, &S->St, &tl, S->Zr_j, S->NZ, 0, NULL, NULL, S->lf_ZtSn, S->T);
  }
  free(kl.ix); free(tl.ix);
}

/* bcc_lf_YnWt (bcc_lf.c:223-302): centred Grams AnZt, AnAt */
static void st_YnWt(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA);
  if (S->has_Z) {
    ilist zr = il_from_mask(S->Zr_j, S->NZ);
    memset(S->lf_AnZt, 0, sizeof(double) * (size_t)S->NA * S->NZ);
    gram_to(S, &S->An, &ar, &S->Zn, &zr, S->Ac_j, S->M, 1, S->lf_An_ajdk, S->lf_Zn_ajdk, S->lf_AnZt, S->NZ);
    free(zr.ix);
  }
  memset(S->lf_AnAt, 0, sizeof(double) * (size_t)S->NA * S->NA);
  gram_to(S, &S->An, &ar, &S->An, &ar, S->Ac_j, S->M, 1, S->lf_An_ajdk, S->lf_An_ajdk, S->lf_AnAt, S->NA);
  free(ar.ix);
}

/* bcc_M_YnWt + bcc_lf_At_T_YnWt_S_Zn (bcc_lf.c:318-356, 636-709), initial pass */
static void st_At_T(lko_state *S)
{
  ilist kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  side_t sA = { &S->At, S->Ar_j, il_from_mask(S->Ar_j, S->NA), &S->Tt };
  if (S->has_Z) {
    side_t sZ = { &S->Zt, S->Zr_j, il_from_mask(S->Zr_j, S->NZ), &S->St };
    ww_excerpt(S, S->trm_AnZt, &sA, &sZ, &kl, &tl, S->lf_AnZt, S->NZ, 1, S->lf_AtTAnZtSZn);
    free(sZ.prow.ix);
  }
  ww_excerpt(S, 0, &sA, &sA, &kl, &tl, S->lf_AnAt, S->NA, 1, S->lf_AtTAnAtTAn);
  free(sA.prow.ix); free(kl.ix); free(tl.ix);
}

/* bcc_lf_AnZt_S_WnYt (bcc_lf.c:360-409): QR terms */
static void st_QR(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), tl = il_from_mask(S->Tc_j, S->T);
  if (S->has_Z) {
    ilist zr = il_from_mask(S->Zr_j, S->NZ);
    memset(S->lf_AnZtSZnAt, 0, sizeof(double) * (size_t)S->NA * S->T);
    qr_kernel(ar.ix, ar.n, zr.ix, zr.n, &S->St, tl.ix, tl.n, S->lf_AnZt, S->lf_AnZt, S->NZ, S->lf_AnZtSZnAt, S->T);
    free(zr.ix);
  }
  memset(S->lf_AnAtTAnAt, 0, sizeof(double) * (size_t)S->NA * S->T);
  qr_kernel(ar.ix, ar.n, ar.ix, ar.n, &S->Tt, tl.ix, tl.n, S->lf_AnAt, S->lf_AnAt, S->NA, S->lf_AnAtTAnAt, S->T);
  free(ar.ix); free(tl.ix);
}

/* bcc_singlestudy_ww (bcc_flattenloop.c:753-831) and the get_singlestudy_vv terms it feeds
   (bcc_flattenloop.c:359-731) that the one-bin, no-Y path consumes */
static void st_single(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  ilist one; int one_ix = 0; one.ix = &one_ix; one.n = 1;
  /* At_Yn_a1d1_excerpt + xcalc + An_ZtSn_ww -> lf_At_An_a1d1 */
  double *a1d1 = (double *)xcalloc(ar.n + 1, sizeof(double));
  for (int i = 0; i < ar.n; i++) a1d1[i] = S->lf_An_ajdk[(size_t)ar.ix[i] * LKO_AJDK + 6];
  lko_planes P; memset(&P, 0, sizeof P);
  lko_xcalc(a1d1, ar.ix, ar.n, S->NA, 1, S->c.B, S->c.D_MLT, &P, NULL);
  S->ops += (long long)ar.n * (S->c.B + 1) * 8;
  memset(S->At_An_a1d1, 0, sizeof(double) * S->M);
  an_ztsn_ww(&S->At, kl.ix, kl.n, S->Ar_j, &P, &one_ix, 1, S->At_An_a1d1, 1);
  S->ops += (long long)kl.n * (S->c.B + 1) * (long long)P.rb * 8;
  lko_planes_free(&P); free(a1d1);
  /* get_singlestudy_vv */
  memset(S->a1d2_AtTn, 0, sizeof(double) * S->T); memset(S->a3d2_AtTn, 0, sizeof(double) * S->T);
  memset(S->a1d2_ZtSn, 0, sizeof(double) * S->T); memset(S->a3d2_ZtSn, 0, sizeof(double) * S->T);
  memset(S->et_Tn, 0, sizeof(double) * S->T); memset(S->et_Sn, 0, sizeof(double) * S->T);
  memset(S->et_An, 0, sizeof(double) * S->M); S->et_An_a1d1 = 0;
  memset(S->T_AnAt_AnAt, 0, sizeof(double) * (size_t)S->T * S->NA);
  ilist zr = il_from_mask(S->Zr_j, S->NZ);
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    const uint8_t *Tr = bm_row(&S->Tt, tt);
    for (int i = 0; i < ar.n; i++) {
      int a = ar.ix[i]; int vT = bpm(Tr, a); const double *x = S->lf_An_ajdk + (size_t)a * LKO_AJDK;
      S->a1d2_AtTn[tt] += vT * x[11];
      S->a3d2_AtTn[tt] += vT * x[13];
    }
    if (S->has_Z) {
      const uint8_t *Sr = bm_row(&S->St, tt);
      for (int i = 0; i < zr.n; i++) {
        int z = zr.ix[i]; int vS = bpm(Sr, z); const double *x = S->lf_Zn_ajdk + (size_t)z * LKO_AJDK;
        S->a1d2_ZtSn[tt] += vS * x[11];
        S->a3d2_ZtSn[tt] += vS * x[13];
      }
    }
  }
  int acpop = ar.n;
  for (int k = 0; k < kl.n; k++) {
    int kk = kl.ix[k];
    long long v = -(long long)acpop + 2 * pop_and(bm_row(&S->At, kk), S->Ar_j, S->At.rb);
    S->et_An[kk] += (double)v;
  }
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    S->et_Tn[tt] = (double)(-(long long)ar.n + 2 * pop_and(bm_row(&S->Tt, tt), S->Ar_j, S->Tt.rb));
    if (S->has_Z) S->et_Sn[tt] = (double)(-(long long)zr.n + 2 * pop_and(bm_row(&S->St, tt), S->Zr_j, S->St.rb));
  }
  for (int i = 0; i < ar.n; i++) S->et_An_a1d1 += S->lf_An_ajdk[(size_t)ar.ix[i] * LKO_AJDK + 6];
  double a0d1 = mask_lf(S->Ac_j, S->pcols, S->ajdk + 5 * S->pcols);
  double a2d1 = mask_lf(S->Ac_j, S->pcols, S->ajdk + 7 * S->pcols);
  for (int i = 0; i < ar.n; i++) {
    int a = ar.ix[i];
    double x11 = S->lf_An_ajdk[(size_t)a * LKO_AJDK + 6];
    double Ad = a0d1;
    Ad -= 2 * 0.0;                 /* popcount_xor_lf of a row with itself */
    Ad -= x11 + x11 - a2d1;
    for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t]; int vT = bpm(bm_row(&S->Tt, tt), a);
      S->T_AnAt_AnAt[(size_t)tt * S->NA + a] = vT * Ad * Ad;
    }
  }
  free(ar.ix); free(kl.ix); free(tl.ix); free(zr.ix);
}

/* bcc_doublestudy_ww (bcc_flattenloop.c:977-1029) with An_a2d2_ZtSn_excerpt (An_ZtSn_ww.c:267-290) */
static void st_double(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  const double *a2d2 = S->ajdk + 12 * (size_t)S->pcols;
  double *L = (double *)xcalloc((size_t)kl.n * tl.n + 1, sizeof(double));
  lko_planes P;
  memset(S->An_a2d2_ZtSn, 0, sizeof(double) * (size_t)S->NA * S->T);
  memset(S->An_a2d2_AtTn, 0, sizeof(double) * (size_t)S->NA * S->T);
  if (S->has_Z) {
    for (int k = 0; k < kl.n; k++) for (int t = 0; t < tl.n; t++)
      L[(size_t)k * tl.n + t] = a2d2[kl.ix[k] / LKO_POP] * S->lf_ZtSn[(size_t)kl.ix[k] * S->T + tl.ix[t]];
    memset(&P, 0, sizeof P);
    lko_xcalc(L, kl.ix, kl.n, S->M, tl.n, S->c.B, S->c.D_MLT, &P, NULL);
    an_ztsn_ww(&S->An, ar.ix, ar.n, S->Ac_j, &P, tl.ix, tl.n, S->An_a2d2_ZtSn, S->T);
    S->ops += (long long)ar.n * tl.n * (S->c.B + 1) * (long long)P.rb * 8;
    lko_planes_free(&P);
  }
  for (int k = 0; k < kl.n; k++) for (int t = 0; t < tl.n; t++)
    L[(size_t)k * tl.n + t] = a2d2[kl.ix[k] / LKO_POP] * S->lf_AtTn[(size_t)kl.ix[k] * S->T + tl.ix[t]];
  memset(&P, 0, sizeof P);
  lko_xcalc(L, kl.ix, kl.n, S->M, tl.n, S->c.B, S->c.D_MLT, &P, NULL);
  an_ztsn_ww(&S->An, ar.ix, ar.n, S->Ac_j, &P, tl.ix, tl.n, S->An_a2d2_AtTn, S->T);
  S->ops += (long long)ar.n * tl.n * (S->c.B + 1) * (long long)P.rb * 8;
  lko_planes_free(&P);
  free(L); free(ar.ix); free(kl.ix); free(tl.ix);
}

/* get_flattenloop (bcc_flattenloop.c:1431-1889), one bin (nb1==nb2), no Y */
static void st_flatten(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  int P = S->pcols; const double *J = S->ajdk; const uint8_t *mcj = S->Ac_j;
  double a0d2 = mask_lf(mcj, P, J + 10 * P), a2d2 = mask_lf(mcj, P, J + 12 * P), a4d2 = mask_lf(mcj, P, J + 14 * P);
  double a0d1 = mask_lf(mcj, P, J + 5 * P), a2d1 = mask_lf(mcj, P, J + 7 * P);
  size_t nqr = (size_t)S->NA * S->T, nqc = (size_t)S->M * S->T;
  memset(S->QR_ZSZ, 0, sizeof(double) * nqr); memset(S->QR_TAT, 0, sizeof(double) * nqr);
  memset(S->QC_ZSZ, 0, sizeof(double) * nqc); memset(S->QC_TAT, 0, sizeof(double) * nqc);
  {
    for (int t = 0; t < tl.n; t++) for (int i = 0; i < ar.n; i++) {
      size_t o = (size_t)ar.ix[i] * S->T + tl.ix[t];
      if (S->has_Z) S->QR_ZSZ[o] = S->lf_AnZtSZnAt[o];
      S->QR_TAT[o] = S->lf_AnAtTAnAt[o];
    }
    for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t];
      for (int i = 0; i < ar.n; i++) {
        int a = ar.ix[i]; size_t o = (size_t)a * S->T + tt; const double *x = S->lf_An_ajdk + (size_t)a * LKO_AJDK;
        double dtmp;
        if (S->has_Z) {
          double eS = S->et_Sn[tt];
          dtmp = + 1*a0d2*eS - 2*S->a1d2_ZtSn[tt] + 1*a2d2*eS - 2*x[11]*eS + 4*S->An_a2d2_ZtSn[o]
                 - 2*x[13]*eS + 1*a2d2*eS - 2*S->a3d2_ZtSn[tt] + 1*a4d2*eS;
          S->QR_ZSZ[o] += -dtmp;
        }
        double eT = S->et_Tn[tt];
        dtmp = + 1*a0d2*eT - 2*S->a1d2_AtTn[tt] + 1*a2d2*eT - 2*x[11]*eT + 4*S->An_a2d2_AtTn[o]
               - 2*x[13]*eT + 1*a2d2*eT - 2*S->a3d2_AtTn[tt] + 1*a4d2*eT;
        S->QR_TAT[o] += -dtmp;
      }
    }
    for (int t = 0; t < tl.n; t++) for (int i = 0; i < ar.n; i++) {
      int tt = tl.ix[t], a = ar.ix[i];
      double dtmp = S->T_AnAt_AnAt[(size_t)tt * S->NA + a];
      S->QR_TAT[(size_t)a * S->T + tt] += -dtmp;
    }
    for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t]; const uint8_t *Tr = bm_row(&S->Tt, tt);
      for (int i = 0; i < ar.n; i++) {
        int a = ar.ix[i]; int vT = bpm(Tr, a); const double *x = S->lf_An_ajdk + (size_t)a * LKO_AJDK;
        double dtmp = vT*( + 1*a0d2 - 4*(x[11]) + 6*a2d2 - 4*(x[13]) + 1*a4d2 );
        S->QR_TAT[(size_t)a * S->T + tt] += +dtmp;
      }
    }
    /* QC */
    for (int t = 0; t < tl.n; t++) for (int k = 0; k < kl.n; k++) {
      int kk = kl.ix[k], tt = tl.ix[t]; size_t o = (size_t)kk * S->T + tt; double D = J[kk / LKO_POP + 5 * P];
      if (S->has_Z) S->QC_ZSZ[o] = D * S->lf_AtTAnZtSZn[o];
      S->QC_TAT[o] = D * S->lf_AtTAnAtTAn[o];
    }
    for (int t = 0; t < tl.n; t++) for (int k = 0; k < kl.n; k++) {
      int kk = kl.ix[k], tt = tl.ix[t]; size_t o = (size_t)kk * S->T + tt; int b = kk / LKO_POP;
      double D01 = J[b + 5 * P], a10 = J[b + 1 * P], a20 = J[b + 2 * P];
      double eT = S->et_Tn[tt];
      double dtmp;
      if (S->has_Z) {
        double eS = S->et_Sn[tt];
        dtmp = D01 * ( +(1*eT) - (2*a10 * S->lf_AtTn[o]) + (a20 * eT) ) * D01 * ( +(1*eS) - (2*a10 * S->lf_ZtSn[o]) + (a20 * eS) );
        S->QC_ZSZ[o] += -dtmp;
      }
      dtmp = D01 * ( +(1*eT) - (2*a10 * S->lf_AtTn[o]) + (a20 * eT) ) * D01 * ( +(1*eT) - (2*a10 * S->lf_AtTn[o]) + (a20 * eT) );
      S->QC_TAT[o] += -dtmp;
    }
    long long rpop = ar.n;
    for (int t = 0; t < tl.n; t++) for (int k = 0; k < kl.n; k++) {
      int kk = kl.ix[k], tt = tl.ix[t]; size_t o = (size_t)kk * S->T + tt; int b = kk / LKO_POP;
      double D01 = J[b + 5 * P], D11 = J[b + 6 * P], D21 = J[b + 7 * P];
      double eA = S->et_An[kk], eAa = S->et_An_a1d1, AtA = S->At_An_a1d1[kk];
      double dtmp = ( +(1*D01*(rpop)*a0d1) - (2*D11*eA*a0d1) + (1*D21*(rpop)*a0d1)
                      - (2*D01*eAa) + (4*D11*AtA) - (2*D21*eAa)
                      + (1*D01*(rpop)*a2d1) - (2*D11*eA*a2d1) + (1*D21*(rpop)*a2d1) );
      S->QC_TAT[o] += -dtmp;
    }
    for (int t = 0; t < tl.n; t++) for (int k = 0; k < kl.n; k++) {
      int kk = kl.ix[k], tt = tl.ix[t]; size_t o = (size_t)kk * S->T + tt; int b = kk / LKO_POP;
      double eA = S->et_An[kk];
      double dtmp = ( +(1*J[b + 10 * P]*(rpop)) - (4*J[b + 11 * P]*eA) + (6*J[b + 12 * P]*(rpop))
                      - (4*J[b + 13 * P]*eA) + (1*J[b + 14 * P]*(rpop)) );
      S->QC_TAT[o] += +dtmp;
    }
  }
  free(ar.ix); free(kl.ix); free(tl.ix);
}

/* bcc_sumscores_nrm (Ireq<=0), _ifT, _cmb (bcc_sumscores.c:324-356, 483-548, 550-779) */
static void dra_plusdiv(double *dst, const double *src, double den, const ilist *rows, const ilist *cols, int ld)
{
  double d = den; if (fabs(d) < 0.000000000001) d = 1;
  for (int c = 0; c < cols->n; c++) for (int r = 0; r < rows->n; r++) {
    size_t o = (size_t)rows->ix[r] * ld + cols->ix[c];
    dst[o] += src[o] / d;
  }
}
static void mds_pow(double *x, const ilist *rows, const ilist *cols, int ld, double den)
{
  if (!(den > 0)) return;
  for (int r = 0; r < rows->n; r++) {
    double *v = x + (size_t)rows->ix[r] * ld;
    if (v[0] >= 0) {
      v[0] *= v[0];
      for (int c = 0; c < cols->n; c++) if (cols->ix[c] >= 1) v[0] -= v[cols->ix[c]] * v[cols->ix[c]] / den;
      v[0] = v[0] >= 0 ? sqrt(v[0]) : -sqrt(-v[0]);
    }
  }
}
static void mds_nrm(double *x, const ilist *rows, const ilist *cols, int ld, double den)
{
  if (!(den > 0)) return;
  for (int r = 0; r < rows->n; r++) {
    double *v = x + (size_t)rows->ix[r] * ld;
    for (int c = 0; c < cols->n; c++) if (cols->ix[c] >= 1) v[0] -= v[cols->ix[c]] / den;
  }
}
/* dcc: the half-loop normalisation of dcc_sumscores_nrm (dcc_sumscores.c:478-699, one bin,
   Ireq<=0); the scores are the same arrays (QR_TAnAtT / QR_TAnZtS / QC_TAnAtT / QC_TAnZtS in
   QR_TAT / QR_ZSZ / QC_TAT / QC_ZSZ), ifT and cmb as dcc_sumscores_ifT / _cmb
   (dcc_sumscores.c:248-280, 411-477) */
static void st_nrm_ifT_cmb_x(lko_state *S, int dcc)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  unsigned int Ar = (unsigned int)ar.n, Zr = (unsigned int)popmask(S->Zr_j, S->NZ);
  long long cpop = kl.n;
  size_t nqr = (size_t)S->NA * S->T, nqc = (size_t)S->M * S->T;
  long long lld;
  if (S->has_Z) {
    memset(S->QR_ZSZ_nrm, 0, sizeof(double) * nqr);
    lld = dcc ? (long long)(Zr - 0) * (long long)(cpop) : (long long)(Zr - 0) * (long long)(cpop) * (long long)(cpop - 1);
    dra_plusdiv(S->QR_ZSZ_nrm, S->QR_ZSZ, (double)lld, &ar, &tl, S->T);
  }
  memset(S->QR_TAT_nrm, 0, sizeof(double) * nqr);
  lld = dcc ? (long long)(Ar - 1) * (long long)(cpop) : (long long)(Ar - 1) * (long long)(cpop) * (long long)(cpop - 1);
  dra_plusdiv(S->QR_TAT_nrm, S->QR_TAT, (double)lld, &ar, &tl, S->T);
  if (S->has_Z) {
    memset(S->QC_ZSZ_nrm, 0, sizeof(double) * nqc);
    lld = dcc ? (long long)(Ar) * (long long)(Zr) : (long long)(Ar) * (long long)(cpop - 1) * (long long)(Zr);
    dra_plusdiv(S->QC_ZSZ_nrm, S->QC_ZSZ, (double)lld, &kl, &tl, S->T);
  }
  memset(S->QC_TAT_nrm, 0, sizeof(double) * nqc);
  lld = dcc ? (long long)(Ar) * (long long)(Ar - 1) : (long long)(Ar) * (long long)(cpop - 1) * (long long)(Ar - 1);
  dra_plusdiv(S->QC_TAT_nrm, S->QC_TAT, (double)lld, &kl, &tl, S->T);
  /* ifT: with the reference's stray ';' the scale factor is always kappa_squared */
  int n_mds = MAXI(0, tl.n - 1); if (n_mds > 6) n_mds = 6;
  if (n_mds > 0) {
    double den = n_mds * S->c.kappa2;
    if (S->has_Z) mds_pow(S->QR_ZSZ_nrm, &ar, &tl, S->T, den);
    mds_pow(S->QR_TAT_nrm, &ar, &tl, S->T, den);
    if (S->has_Z && (!dcc || popmask(S->Zr_b, S->NZ) > 0)) mds_nrm(S->QC_ZSZ_nrm, &kl, &tl, S->T, den);
    mds_nrm(S->QC_TAT_nrm, &kl, &tl, S->T, den);
  }
  /* cmb, first T column only */
  if (S->has_Z && popmask(S->Zr_b, S->NZ) > 0 && popmask(S->Ar_b, S->NA) > 0) {
    for (int k = 0; k < kl.n; k++) { size_t o = (size_t)kl.ix[k] * S->T; S->QC_TAT_nrm[o] += -1.0 * S->QC_ZSZ_nrm[o]; }
  }
  if (S->has_Z) {
    for (int i = 0; i < ar.n; i++) { size_t o = (size_t)ar.ix[i] * S->T; S->QR_TAT_nrm[o] += -1.0 * S->QR_ZSZ_nrm[o]; }
  }
  free(ar.ix); free(kl.ix); free(tl.ix);
}
static void st_nrm_ifT_cmb(lko_state *S) { st_nrm_ifT_cmb_x(S, 0); }

/* bcc_sumscores_xij (bcc_sumscores.c:225-322), one bin, Ireq<=0 */
static int st_xij(lko_state *S, int *Irem)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M);
  *Irem = ar.n > 0 ? 1 : 0;
  int Ab = popmask(S->Ar_b, S->NA);
  if (Ab) {
    for (int j = 0; j < kl.n; j++) {
      int k = kl.ix[j];
      S->QC_sv[j] = S->QC_TAT_nrm[(size_t)k * S->T + 0];
      S->QC_a[j] = k; S->QC_b[j] = S->Ac_bidx[k]; S->QC_j[j] = j;
    }
    lko_quicksort_xij(S->QC_sv, S->QC_a, S->QC_b, S->QC_j, NULL, NULL, NULL, 0, kl.n - 1);
    for (int j = 0; j < ar.n; j++) {
      int a = ar.ix[j];
      S->QR_sv[j] = S->QR_TAT_nrm[(size_t)a * S->T + 0];
      S->QR_a[j] = a; S->QR_b[j] = S->Ar_bidx[a]; S->QR_j[j] = j; S->QR_nb[j] = 0;
      S->QR_ga[j] = a; S->QR_gb[j] = S->Ar_bidx[a];
    }
    lko_quicksort_xij(S->QR_sv, S->QR_a, S->QR_b, S->QR_j, S->QR_nb, S->QR_ga, S->QR_gb, 0, ar.n - 1);
  }
  int rn = ar.n; free(ar.ix); free(kl.ix);
  return rn;
}

static double ra_mean(const double *v, unsigned long long n)
{
  double mean = 0;
  for (unsigned long long i = 0; i < n; i++) mean += v[i];
  mean /= (double)n;
  return mean;
}

/* bcc_sumscores_mxA/mxB/mxC (bcc_sumscores.c:37-205) */
static void st_mxA(lko_state *S, int iter, int Irem, int *rdrop_p, int *cdrop_p)
{
  int rpop = popmask(S->Ar_j, S->NA), cpop = popmask(S->Ac_j, S->M);
  int rdrop = 0, cdrop = 0;
  lko_get_xdrop(rpop, cpop, S->c.gamma, S->c.gamma_type, &rdrop, &cdrop);
  size_t nbc = lko_bsize(S->M), nbr = lko_bsize(S->NA);
  memcpy(S->Ac_rtn, S->Ac_j, nbc); memset(S->Ac_rmv, 0, nbc);
  for (int nc = 0; nc < MINI(cdrop, cpop); nc++) { int k = S->QC_a[nc]; bset(S->Ac_rmv, k); S->Ac_rtn[k >> 3] &= (uint8_t)~(1u << (7 - (k & 7))); }
  memcpy(S->Ar_rtn, S->Ar_j, nbr); memset(S->Ar_rmv, 0, nbr);
  for (int nr = 0; nr < MINI(rdrop, rpop); nr++) { int a = S->QR_a[nr]; bset(S->Ar_rmv, a); S->Ar_rtn[a >> 3] &= (uint8_t)~(1u << (7 - (a & 7))); }
  /* trace row */
  double *tr = S->trace + (size_t)S->n_trace * 6;
  double m;
  tr[0] = iter; tr[1] = rpop; tr[2] = cpop;
  m = ra_mean(S->QR_sv, (unsigned long long)rpop); if (!isfinite(m)) m = 1.0; tr[3] = m;
  m = ra_mean(S->QC_sv, (unsigned long long)cpop); if (!isfinite(m)) m = 1.0; tr[4] = m;
  tr[5] = Irem;
  S->n_trace++;
  for (int nr = 0; nr < MINI(rdrop, rpop); nr++) {
    S->xdrop_a[2 * S->n_xdrop] = S->QR_ga[nr]; S->xdrop_a[2 * S->n_xdrop + 1] = -1;
    S->xdrop_b[2 * S->n_xdrop] = S->QR_gb[nr]; S->xdrop_b[2 * S->n_xdrop + 1] = -1;
    S->n_xdrop++;
  }
  for (int nc = 0; nc < MINI(cdrop, cpop); nc++) {
    S->xdrop_a[2 * S->n_xdrop] = -1; S->xdrop_a[2 * S->n_xdrop + 1] = S->QC_a[nc];
    S->xdrop_b[2 * S->n_xdrop] = -1; S->xdrop_b[2 * S->n_xdrop + 1] = S->QC_b[nc];
    S->n_xdrop++;
  }
  *rdrop_p = rdrop; *cdrop_p = cdrop;
}

static double *dalloc(size_t n) { return (double *)xcalloc(n + 1, sizeof(double)); }

/* ------------------------------------------------------------------ */
/* dexcluster_driver's half-loop subscores (dexcluster_driver.c:22-25, 44-47), one bin,
   no Y: dcc_lf_D_AtTn_ZtSn_vv (dcc_lf.c:102-163; get_D_AtTn_ZtSn_vv and
   get_D_AtTn_AtTn_vv, D_AtTn_ZtSn.c:5-310), dcc_lf_TAnZtS_ww (dcc_lf.c:267-288;
   TAnZtS.c stages 0-2, 5-540) and dcc_wrap_dcc_halfloop (dcc_lf.c:379-530).  Writes
   QR_TAnAtT / QR_TAnZtS [a*T+t] into QR_TAT / QR_ZSZ and QC_TAnAtT / QC_TAnZtS [k*T+t]
   into QC_TAT / QC_ZSZ; needs st_ajdk and st_ZtSn of the same masks.                  */
/* ------------------------------------------------------------------ */
static void st_dcc(lko_state *S)
{
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M), tl = il_from_mask(S->Tc_j, S->T);
  ilist zr = il_from_mask(S->Zr_j, S->NZ);
  const int P = S->pcols, T = S->T; const double *J = S->ajdk;
  const double *aA = J + 1 * (size_t)P, *DA = J + 5 * (size_t)P;        /* AJDK_1_0, AJDK_0_1 */
  memset(S->QR_ZSZ, 0, sizeof(double) * (size_t)S->NA * T); memset(S->QR_TAT, 0, sizeof(double) * (size_t)S->NA * T);
  memset(S->QC_ZSZ, 0, sizeof(double) * (size_t)S->M * T); memset(S->QC_TAT, 0, sizeof(double) * (size_t)S->M * T);
  /* -cpop_j + 2 popcount of the covariate rows over the active patients */
  double *etT = dalloc(T), *etS = dalloc(T), *a1T = dalloc(T), *a1S = dalloc(T);
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    etT[tt] = (double)(-(long long)ar.n + 2 * pop_and(bm_row(&S->Tt, tt), S->Ar_j, S->Tt.rb));
    if (S->has_Z) etS[tt] = (double)(-(long long)zr.n + 2 * pop_and(bm_row(&S->St, tt), S->Zr_j, S->St.rb));
  }
  /* QC: get_D_AtTn_ZtSn_vv / get_D_AtTn_AtTn_vv (one bin: nb1 == nb2) */
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    for (int k = 0; k < kl.n; k++) {
      int kk = kl.ix[k]; size_t o = (size_t)kk * T + tt;
      double a = aA[kk / LKO_POP], D = DA[kk / LKO_POP], At = S->lf_AtTn[o];
      if (S->has_Z) S->QC_ZSZ[o] = (At - etT[tt] * a) * D * (S->lf_ZtSn[o] - a * etS[tt]);
      double v = (At - etT[tt] * a) * D * (At - a * etT[tt]);
      double etA = (double)(-(long long)ar.n + 2 * pop_and(bm_row(&S->At, kk), S->Ar_j, S->At.rb));
      v -= D * (ar.n - 2 * etA * a + ar.n * a * a);
      S->QC_TAT[o] = v;
    }
  }
  /* stage 0 (get_dcc_TAnZtS_ww_stage_0): a0d1, a2d1, a1d1_AtTn, a1d1_ZtSn */
  double a0d1 = mask_lf(S->Ac_j, P, J + 5 * (size_t)P), a2d1 = mask_lf(S->Ac_j, P, J + 7 * (size_t)P);
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    const uint8_t *Tr = bm_row(&S->Tt, tt);
    for (int i = 0; i < ar.n; i++) a1T[tt] += bpm(Tr, ar.ix[i]) * S->lf_An_ajdk[(size_t)ar.ix[i] * LKO_AJDK + 6];
    if (S->has_Z) {
      const uint8_t *Sr = bm_row(&S->St, tt);
      for (int i = 0; i < zr.n; i++) a1S[tt] += bpm(Sr, zr.ix[i]) * S->lf_Zn_ajdk[(size_t)zr.ix[i] * LKO_AJDK + 6];
    }
  }
  /* stage 1 (dcc_TAnZtS_ww_stage_1): An_a0d1_ZtSn_excerpt (An_ZtSn_ww.c:243-265), xcalc and
     An_ZtSn_ww of the a0d1-weighted ZtSn / AtTn */
  double *L = (double *)xcalloc((size_t)kl.n * tl.n + 1, sizeof(double));
  double *xZ = dalloc((size_t)S->NA * T), *xA = dalloc((size_t)S->NA * T);
  lko_planes Pl;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 0 && !S->has_Z) continue;
    const double *src = pass == 0 ? S->lf_ZtSn : S->lf_AtTn;
    for (int k = 0; k < kl.n; k++) for (int t = 0; t < tl.n; t++)
      L[(size_t)k * tl.n + t] = DA[kl.ix[k] / LKO_POP] * src[(size_t)kl.ix[k] * T + tl.ix[t]];
    memset(&Pl, 0, sizeof Pl);
    lko_xcalc(L, kl.ix, kl.n, S->M, tl.n, S->c.B, S->c.D_MLT, &Pl, NULL);
    an_ztsn_ww(&S->An, ar.ix, ar.n, S->Ac_j, &Pl, tl.ix, tl.n, pass == 0 ? xZ : xA, T);
    S->ops += (long long)ar.n * tl.n * (S->c.B + 1) * (long long)Pl.rb * 8;
    lko_planes_free(&Pl);
  }
  /* stage 2 (get_dcc_TAnZtS_ww_stage_2) and dcc_halfloop */
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    for (int i = 0; i < ar.n; i++) {
      int a = ar.ix[i]; size_t o = (size_t)a * T + tt;
      double x6 = S->lf_An_ajdk[(size_t)a * LKO_AJDK + 6];
      if (S->has_Z) S->QR_ZSZ[o] += +(xZ[o] - a1S[tt] - x6 * etS[tt] + a2d1 * etS[tt]);
      S->QR_TAT[o] += +(xA[o] - a1T[tt] - x6 * etT[tt] + a2d1 * etT[tt]);
    }
  }
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    const uint8_t *Tr = bm_row(&S->Tt, tt);
    for (int i = 0; i < ar.n; i++) {
      int a = ar.ix[i]; size_t o = (size_t)a * T + tt;
      double x6 = S->lf_An_ajdk[(size_t)a * LKO_AJDK + 6];
      double l = bpm(Tr, a) * (a0d1 - 2 * x6 + a2d1);
      S->QR_TAT[o] += -l;
    }
  }
  free(L); free(xZ); free(xA); free(etT); free(etS); free(a1T); free(a1S);
  free(ar.ix); free(kl.ix); free(tl.ix); free(zr.ix);
}

/* low-rank update after a drop: bcc_lrup_QR_YnWt_stage_0/1/3 and
   bcc_lrup_QC_YnWt_stage_a0..a3, b0..b2, c (bcc_lrup.c:5-655, 659-892) */
static void st_lrup(lko_state *S)
{
  int NA = S->NA, NZ = S->NZ, M = S->M, T = S->T;
  ilist rtn = il_from_mask(S->Ar_rtn, NA), rmv = il_from_mask(S->Ar_rmv, NA);
  ilist crt = il_from_mask(S->Ac_rtn, M), crm = il_from_mask(S->Ac_rmv, M);
  ilist zr = il_from_mask(S->Zr_j, NZ), tl = il_from_mask(S->Tc_j, T);
  double *jn_ajdk = dalloc((size_t)NA * LKO_AJDK), *vn_ajdk = dalloc((size_t)NZ * LKO_AJDK);
  double *kn_ajdk = dalloc((size_t)NA * LKO_AJDK), *hn_ajdk = dalloc((size_t)NA * LKO_AJDK), *an_ajdk = dalloc((size_t)NA * LKO_AJDK);
  double *zn_ajdk = dalloc((size_t)NZ * LKO_AJDK);
  /* QR stage 0/1/3 */
  ajdk_to(S, &S->An, &rtn, S->Ac_rmv, NA, jn_ajdk);
  if (S->has_Z) {
    ajdk_to(S, &S->Zn, &zr, S->Ac_rmv, NZ, vn_ajdk);
    double *jnvt = dalloc((size_t)rtn.n * zr.n);
    double *xa = dalloc(rtn.n), *xz = dalloc(zr.n);
    for (int i = 0; i < rtn.n; i++) xa[i] = jn_ajdk[(size_t)rtn.ix[i] * LKO_AJDK + 6];
    for (int j = 0; j < zr.n; j++) xz[j] = vn_ajdk[(size_t)zr.ix[j] * LKO_AJDK + 6];
    lko_anzt_vv(S->An.b, S->An.rb, rtn.ix, rtn.n, S->Zn.b, S->Zn.rb, zr.ix, zr.n, S->Ac_rmv, M, S->ajdk, xa, xz, jnvt);
    S->ops += (long long)rtn.n * zr.n * (long long)S->An.rb * 8;
    for (int i = 0; i < rtn.n; i++) for (int j = 0; j < zr.n; j++) S->lf_AnZt[(size_t)rtn.ix[i] * NZ + zr.ix[j]] -= jnvt[(size_t)i * zr.n + j];
    free(jnvt); free(xa); free(xz);
  }
  {
    double *jnjt = dalloc((size_t)rtn.n * rtn.n);
    double *xa = dalloc(rtn.n);
    for (int i = 0; i < rtn.n; i++) xa[i] = jn_ajdk[(size_t)rtn.ix[i] * LKO_AJDK + 6];
    lko_anzt_vv(S->An.b, S->An.rb, rtn.ix, rtn.n, S->An.b, S->An.rb, rtn.ix, rtn.n, S->Ac_rmv, M, S->ajdk, xa, xa, jnjt);
    S->ops += (long long)rtn.n * rtn.n * (long long)S->An.rb * 8;
    for (int i = 0; i < rtn.n; i++) for (int j = 0; j < rtn.n; j++) S->lf_AnAt[(size_t)rtn.ix[i] * NA + rtn.ix[j]] -= jnjt[(size_t)i * rtn.n + j];
    free(jnjt); free(xa);
  }
  /* QC a0 */
  ajdk_to(S, &S->An, &rmv, S->Ac_rtn, NA, kn_ajdk);
  ajdk_to(S, &S->An, &rmv, S->Ac_rmv, NA, hn_ajdk);
  ajdk_to(S, &S->An, &rtn, S->Ac_rtn, NA, an_ajdk);
  if (S->has_Z) { ajdk_to(S, &S->Zn, &zr, S->Ac_rtn, NZ, zn_ajdk); ajdk_to(S, &S->Zn, &zr, S->Ac_rmv, NZ, vn_ajdk); }
  /* QC a1 (compact grams) */
  double *knzt = NULL, *hnvt = NULL, *ankt = NULL, *knkt = NULL;
  {
#define XL(dst, aj, lst) do { dst = dalloc((lst).n); for (int _i = 0; _i < (lst).n; _i++) dst[_i] = (aj)[(size_t)(lst).ix[_i] * LKO_AJDK + 6]; } while (0)
    double *x_kn, *x_hn, *x_an, *x_zn, *x_vn;
    XL(x_kn, kn_ajdk, rmv); XL(x_hn, hn_ajdk, rmv); XL(x_an, an_ajdk, rtn); XL(x_zn, zn_ajdk, zr); XL(x_vn, vn_ajdk, zr);
    if (S->has_Z) {
      knzt = dalloc((size_t)rmv.n * zr.n); hnvt = dalloc((size_t)rmv.n * zr.n);
      lko_anzt_vv(S->An.b, S->An.rb, rmv.ix, rmv.n, S->Zn.b, S->Zn.rb, zr.ix, zr.n, S->Ac_rtn, M, S->ajdk, x_kn, x_zn, knzt);
      lko_anzt_vv(S->An.b, S->An.rb, rmv.ix, rmv.n, S->Zn.b, S->Zn.rb, zr.ix, zr.n, S->Ac_rmv, M, S->ajdk, x_hn, x_vn, hnvt);
      S->ops += 2LL * rmv.n * zr.n * (long long)S->An.rb * 8;
    }
    ankt = dalloc((size_t)rtn.n * rmv.n); knkt = dalloc((size_t)rmv.n * rmv.n);
    lko_anzt_vv(S->An.b, S->An.rb, rtn.ix, rtn.n, S->An.b, S->An.rb, rmv.ix, rmv.n, S->Ac_rtn, M, S->ajdk, x_an, x_kn, ankt);
    lko_anzt_vv(S->An.b, S->An.rb, rmv.ix, rmv.n, S->An.b, S->An.rb, rmv.ix, rmv.n, S->Ac_rtn, M, S->ajdk, x_kn, x_kn, knkt);
    S->ops += ((long long)rtn.n * rmv.n + (long long)rmv.n * rmv.n) * (long long)S->An.rb * 8;
    free(x_kn); free(x_hn); free(x_an); free(x_zn); free(x_vn);
#undef XL
  }
  /* QC a2+a3 */
  size_t nmt = (size_t)M * T;
  double *t_knzt = dalloc(nmt), *t_hnvt = dalloc(nmt), *t_ankt = dalloc(nmt), *t_knkt = dalloc(nmt);
  side_t s_rmv = { &S->At, S->Ar_rmv, rmv, &S->Tt }, s_rtn = { &S->At, S->Ar_rtn, rtn, &S->Tt };
  side_t s_Z = { &S->Zt, S->Zr_j, zr, &S->St };
  if (S->has_Z) {
    ww_excerpt(S, S->trm_knzt, &s_rmv, &s_Z, &crt, &tl, knzt, zr.n, 0, t_knzt);
    ww_excerpt(S, S->trm_hnvt, &s_rmv, &s_Z, &crt, &tl, hnvt, zr.n, 0, t_hnvt);
  }
  ww_excerpt(S, 0, &s_rtn, &s_rmv, &crt, &tl, ankt, rmv.n, 0, t_ankt);
  ww_excerpt(S, 0, &s_rmv, &s_rmv, &crt, &tl, knkt, rmv.n, 0, t_knkt);
  /* QC b0 (plain +-1 products, original indices [k*T+t]) */
  double *attn = dalloc(nmt), *jttn = dalloc(nmt), *ktrn = dalloc(nmt), *htrn = dalloc(nmt), *ztsn = dalloc(nmt), *vtsn = dalloc(nmt);
  gram_to(S, &S->At, &crt, &S->Tt, &tl, S->Ar_rtn, NA, 0, NULL, NULL, attn, T);
  gram_to(S, &S->At, &crm, &S->Tt, &tl, S->Ar_rtn, NA, 0, NULL, NULL, jttn, T);
  gram_to(S, &S->At, &crt, &S->Tt, &tl, S->Ar_rmv, NA, 0, NULL, NULL, ktrn, T);
  gram_to(S, &S->At, &crm, &S->Tt, &tl, S->Ar_rmv, NA, 0, NULL, NULL, htrn, T);
  if (S->has_Z) {
    gram_to(S, &S->Zt, &crt, &S->St, &tl, S->Zr_j, NZ, 0, NULL, NULL, ztsn, T);
    gram_to(S, &S->Zt, &crm, &S->St, &tl, S->Zr_j, NZ, 0, NULL, NULL, vtsn, T);
  }
  /* QC b1 (compact [k][y][t]) */
  size_t n3 = (size_t)crt.n * crm.n * tl.n;
  double *ztsvn = dalloc(n3), *attjn = dalloc(n3), *ktrhn = dalloc(n3);
  if (S->has_Z) attyn_kernel(&S->Zt, crt.ix, crt.n, S->Zr_j, zr.n, &S->St, tl.ix, tl.n, S->Zr_j, zr.n, &S->Zt, crm.ix, crm.n, ztsn, vtsn, T, S->ajdk, S->pcols, ztsvn);
  attyn_kernel(&S->At, crt.ix, crt.n, S->Ar_rtn, rtn.n, &S->Tt, tl.ix, tl.n, S->Ar_rtn, rtn.n, &S->At, crm.ix, crm.n, attn, jttn, T, S->ajdk, S->pcols, attjn);
  attyn_kernel(&S->At, crt.ix, crt.n, S->Ar_rmv, rmv.n, &S->Tt, tl.ix, tl.n, S->Ar_rmv, rmv.n, &S->At, crm.ix, crm.n, ktrn, htrn, T, S->ajdk, S->pcols, ktrhn);
  S->ops += (long long)tl.n * crt.n * crm.n * ((long long)S->At.rb * 8 * 2 + (S->has_Z ? (long long)S->Zt.rb * 8 : 0));
  /* QC b2 */
  double *vtszn = dalloc(nmt), *jttan = dalloc(nmt), *htrkn = dalloc(nmt), *hhtrkn = dalloc(nmt);
  if (S->has_Z) attyn_wtszn_kernel(crt.ix, crt.n, crm.ix, crm.n, tl.ix, tl.n, attjn, ztsvn, S->ajdk, S->pcols, vtszn, T);
  attyn_wtszn_kernel(crt.ix, crt.n, crm.ix, crm.n, tl.ix, tl.n, attjn, attjn, S->ajdk, S->pcols, jttan, T);
  attyn_wtszn_kernel(crt.ix, crt.n, crm.ix, crm.n, tl.ix, tl.n, attjn, ktrhn, S->ajdk, S->pcols, htrkn, T);
  attyn_wtszn_kernel(crt.ix, crt.n, crm.ix, crm.n, tl.ix, tl.n, ktrhn, ktrhn, S->ajdk, S->pcols, hhtrkn, T);
  /* bcc_lrup_mxdup: retained becomes current */
  memcpy(S->Ar_j, S->Ar_rtn, lko_bsize(NA)); memcpy(S->Ac_j, S->Ac_rtn, lko_bsize(M));
  /* bcc_lrup_QC_YnWt_stage_c: L2_clean then subtract the low-rank terms */
  for (int k = 0; k < M; k++) for (int t = 0; t < T; t++) {
    size_t o = (size_t)k * T + t;
    int keep = bget(S->Ac_rtn, k) && bget(S->Ac_b, k) && bget(S->Tc_j, t);
    if (!keep) { S->lf_AtTAnZtSZn[o] = 0; S->lf_AtTAnAtTAn[o] = 0; continue; }
    if (S->has_Z) {
      S->lf_AtTAnZtSZn[o] += -1.0 * vtszn[o];
      S->lf_AtTAnZtSZn[o] += -1.0 * t_knzt[o];
      S->lf_AtTAnZtSZn[o] += -1.0 * t_hnvt[o];
    }
    S->lf_AtTAnAtTAn[o] += -1.0 * jttan[o];
    S->lf_AtTAnAtTAn[o] += -1.0 * htrkn[o];
    S->lf_AtTAnAtTAn[o] += -1.0 * htrkn[o];      /* lf_ktrhn____jttan duplicates lf_attjn____htrkn */
    S->lf_AtTAnAtTAn[o] += -1.0 * hhtrkn[o];
    S->lf_AtTAnAtTAn[o] += -1.0 * t_ankt[o];
    S->lf_AtTAnAtTAn[o] += -1.0 * t_ankt[o];     /* lf_kt_r_knat_t_an duplicates lf_at_t_ankt_r_kn */
    S->lf_AtTAnAtTAn[o] += -1.0 * t_knkt[o];
  }
  free(jn_ajdk); free(vn_ajdk); free(kn_ajdk); free(hn_ajdk); free(an_ajdk); free(zn_ajdk);
  free(knzt); free(hnvt); free(ankt); free(knkt);
  free(t_knzt); free(t_hnvt); free(t_ankt); free(t_knkt);
  free(attn); free(jttn); free(ktrn); free(htrn); free(ztsn); free(vtsn);
  free(ztsvn); free(attjn); free(ktrhn); free(vtszn); free(jttan); free(htrkn); free(hhtrkn);
  free(rtn.ix); free(rmv.ix); free(crt.ix); free(crm.ix); free(zr.ix); free(tl.ix);
}

/* ------------------------------------------------------------------ */
/* load (bcc_load, bcc.c:1482-1540) and run (lakcluster_driver, lakcluster_driver.c:14-251) */
/* ------------------------------------------------------------------ */
static double now_s(void) { struct timeval tv; gettimeofday(&tv, NULL); return tv.tv_sec + 1e-6 * tv.tv_usec; }

typedef struct {
  int n_trace; double *trace; int n_xdrop; int *xdrop_a, *xdrop_b;
  long long ops; double load_s, init_s, loop_s; int n_iter; int status; char err[256];
} lko_result;

static int fail(lko_result *R, const char *msg) { R->status = -1; snprintf(R->err, sizeof R->err, "%s", msg); return -1; }

/* ------------------------------------------------------------------ */
/* bcc_scramble (R_init.c:165-193, 256-285; R_handle_make R_init.c:5-29; xcc_out_xdrop_load
   and _sort bcc_scorebox.c:163-196; bcc_out_xdrop_prune R_init.c:195-238): the rows and
   columns named in an out_xdrop file, ascending and restricted to the loaded masks, get
   every (row, column) bit redrawn in row-major order as (r < p_k), r from R01GET
   (boxmuller.c:57-66: seed = seed * 35 % (2^31 - 1), r = seed / (2^31 - 1) in long
   double), p_k = the column's count over the loaded Z rows / their number.  Parity
   pinned by the closed form of the generator and the resampling invariants
   (tests/test_scramble.py).                                                          */
/* ------------------------------------------------------------------ */
static double lko_r01(unsigned long *seed)
{
  *seed = (*seed * 35UL) % 2147483647UL;
  return (double)((long double)*seed / (long double)2147483647UL);
}
static int icmp(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return x < y ? -1 : x > y; }
static void bclr_(uint8_t *b, long i) { b[i >> 3] &= (uint8_t)~(1u << (7 - (i & 7))); }

static int list_item(const char *raw, int n, char *out, size_t cap)
{
  /* the n-th comma-separated entry, trimmed */
  const char *p = raw;
  for (int i = 0; i < n && p; i++) { p = strchr(p, ','); if (p) p++; }
  if (!p) { out[0] = 0; return -1; }
  const char *e = strchr(p, ','); size_t l = e ? (size_t)(e - p) : strlen(p); if (l >= cap) l = cap - 1;
  memcpy(out, p, l); out[l] = 0; trim(out);
  return out[0] ? 0 : -1;
}

static int st_scramble(lko_state *S, lko_result *R)
{
  lko_cfg *c = &S->c;
  int Zrb = popmask(S->Zr_b, S->NZ);
  for (int s = 0; s < c->scramble_num; s++) {
    char fn[LKO_PATH], sd[64];
    if (list_item(c->scr_files_raw, s, fn, sizeof fn)) return fail(R, "GLOBAL_scramble_out_xdrop_ names fewer files than GLOBAL_scramble_num");
    unsigned long seed = (unsigned long)(s + 1);
    if (!list_item(c->scr_seeds_raw, s, sd, sizeof sd)) seed = (unsigned long)atol(sd);
    FILE *fp = fopen(fn, "r"); if (!fp) return fail(R, "cannot open a scramble file");
    int nl = 0; for (int ch; (ch = fgetc(fp)) != EOF;) nl += ch == '\n';
    fseek(fp, 0, SEEK_SET);
    int *rows = (int *)xcalloc((size_t)nl + 1, sizeof(int)), *cols = (int *)xcalloc((size_t)nl + 1, sizeof(int));
    int nr = 0, nc = 0;
    for (int l = 0; l < nl; l++) {
      int a = -1, b = -1;
      if (fscanf(fp, "%d %d", &a, &b) != 2) break;
      if (a >= 0) rows[nr++] = a;
      if (b >= 0) cols[nc++] = b;
    }
    fclose(fp);
    qsort(rows, nr, sizeof(int), icmp); qsort(cols, nc, sizeof(int), icmp);
    int pr = 0, pc = 0;
    for (int i = 0; i < nr; i++) if (rows[i] >= 0 && rows[i] < S->NA && bget(S->Ar_b, rows[i])) rows[pr++] = rows[i];
    for (int j = 0; j < nc; j++) if (cols[j] >= 0 && cols[j] < S->M && bget(S->Ac_b, cols[j])) cols[pc++] = cols[j];
    for (int i = 0; i < pr; i++) {
      const int a = rows[i];
      for (int j = 0; j < pc; j++) {
        const int k = cols[j];
        const double zs = S->has_Z ? (double)pop_and(bm_row(&S->Zt, k), S->Zr_b, S->Zt.rb) : 0.0;
        const int v = lko_r01(&seed) < zs / (double)Zrb;
        uint8_t *an = (uint8_t *)bm_row(&S->An, a), *at = (uint8_t *)bm_row(&S->At, k);
        if (v) { bset(an, k); bset(at, a); } else { bclr_(an, k); bclr_(at, a); }
      }
    }
    free(rows); free(cols);
  }
  return 0;
}

static int st_init(lko_state *S, lko_result *R);
static int st_arrays(lko_state *S, lko_result *R);

static int st_load(lko_state *S, lko_result *R)
{
  lko_cfg *c = &S->c;
  if (!strcmp(c->test_type, "bcc_sumscores")) return st_init(S, R);
  if (c->nbins != 1) return fail(R, "only GLOBAL_NBINS= 1 is supported");
  if (c->Y_ncols != 0) return fail(R, "Y/W genetic controls are not supported");
  /* dexcluster_driver has no low-rank update and reads no QR/QC strategy (dcc.c:1167) */
  int dcc = !strcmp(c->test_type, "dexcluster_driver") || !strcmp(c->test_type, "dexcluster_scorebox");
  if (!dcc && (!strstr(c->QR, "YnWt") || !strstr(c->QC, "YnWt") || strstr(c->QR, "condense") || strstr(c->QC, "store one")
               || strstr(c->QR, "ZtSWn") || strstr(c->QC, "ZtSWn"))) return fail(R, "only the YnWt strategies are supported");
  /* GLOBAL_Ireq > 0 with one bin: one value sorted per row, the same normalisers
     (bcc_sumscores.c:358-481, 583-628): the GLOBAL_Ireq <= 0 run */
  if (!c->A_t[0] || !c->A_n[0]) return fail(R, "A_n/A_t files required");
  if (b16_read(c->A_t, &S->An) || b16_read(c->A_n, &S->At)) return fail(R, "cannot read A");
  S->NA = S->An.nr; S->M = S->An.nc;
  if (c->A_ncols && c->A_ncols != S->M) return fail(R, "GLOBAL_A_n_cols does not match A");
  if (S->At.nr != S->M || S->At.nc != S->NA) return fail(R, "A_n/A_t shapes disagree");
  S->has_Z = c->Z_t[0] != 0;
  if (S->has_Z) {
    if (b16_read(c->Z_t, &S->Zn) || b16_read(c->Z_n, &S->Zt)) return fail(R, "cannot read Z");
    S->NZ = S->Zn.nr;
    if (S->Zn.nc != S->M || S->Zt.nr != S->M || S->Zt.nc != S->NZ) return fail(R, "Z shapes disagree");
  } else { S->NZ = 0; S->Zn = bm_alloc(0, S->M); S->Zt = bm_alloc(S->M, 0); }
  if (c->T_n[0]) {
    if (b16_read(c->T_n, &S->Tt)) return fail(R, "cannot read T");
    S->T = S->Tt.nr;
    if (S->Tt.nc != S->NA) return fail(R, "T shape disagrees");
    if (S->has_Z) { if (!c->S_n[0] || b16_read(c->S_n, &S->St)) return fail(R, "cannot read S"); if (S->St.nr != S->T || S->St.nc != S->NZ) return fail(R, "S shape disagrees"); }
    else S->St = bm_alloc(S->T, 0);
  } else {
    /* no covariate file: a single all-ones intercept column (bcc.c:361-373) */
    S->T = 1; S->Tt = bm_alloc(1, S->NA); for (int a = 0; a < S->NA; a++) bset(S->Tt.b, a);
    S->St = bm_alloc(1, S->NZ); for (int z = 0; z < S->NZ; z++) bset(S->St.b, z);
  }
  S->pcols = lko_psize(S->M);
  S->Ar_b = mask_alloc(S->NA); S->Ar_j = mask_alloc(S->NA); S->Ar_rtn = mask_alloc(S->NA); S->Ar_rmv = mask_alloc(S->NA);
  S->Zr_b = mask_alloc(S->NZ); S->Zr_j = mask_alloc(S->NZ);
  S->Ac_b = mask_alloc(S->M); S->Ac_j = mask_alloc(S->M); S->Ac_rtn = mask_alloc(S->M); S->Ac_rmv = mask_alloc(S->M);
  S->Tc_j = mask_alloc(S->T);
  if (c->A_rind[0]) { if (mask_read(c->A_rind, S->NA, S->Ar_b)) return fail(R, "cannot read A row mask"); } else for (int i = 0; i < S->NA; i++) bset(S->Ar_b, i);
  if (c->Z_rind[0] && S->has_Z) { if (mask_read(c->Z_rind, S->NZ, S->Zr_b)) return fail(R, "cannot read Z row mask"); } else for (int i = 0; i < S->NZ; i++) bset(S->Zr_b, i);
  if (c->A_cind[0]) { if (mask_read(c->A_cind, S->M, S->Ac_b)) return fail(R, "cannot read A col mask"); } else for (int i = 0; i < S->M; i++) bset(S->Ac_b, i);
  if (c->T_cind[0]) { if (mask_read(c->T_cind, S->T, S->Tc_j)) return fail(R, "cannot read T col mask"); }
  else { int nt = c->T_ncols > 0 ? MINI(c->T_ncols, S->T) : S->T; for (int i = 0; i < nt; i++) bset(S->Tc_j, i); }
  for (int i = 1; i < S->T; i++) if (bget(S->Tc_j, i) && !bget(S->Tc_j, i - 1)) return fail(R, "T column mask must be a prefix");
  memcpy(S->Ar_j, S->Ar_b, lko_bsize(S->NA)); memcpy(S->Zr_j, S->Zr_b, lko_bsize(S->NZ)); memcpy(S->Ac_j, S->Ac_b, lko_bsize(S->M));
  S->Ar_bidx = (int *)xcalloc(S->NA + 1, sizeof(int)); S->Ac_bidx = (int *)xcalloc(S->M + 1, sizeof(int));
  for (int i = 0, b = 0; i < S->NA; i++) { S->Ar_bidx[i] = b; if (bget(S->Ar_b, i)) b++; }
  for (int i = 0, b = 0; i < S->M; i++) { S->Ac_bidx[i] = b; if (bget(S->Ac_b, i)) b++; }
  /* bcc_load_A_p (bcc.c:1197-1313): per-block frequency over mr_b rows and mc_b columns */
  S->A_p = dalloc(S->pcols); S->ajdk = dalloc((size_t)S->pcols * LKO_AJDK);
  int Arb = popmask(S->Ar_b, S->NA), Zrb = popmask(S->Zr_b, S->NZ);
  double *AZ_rsum = dalloc(S->M);
  for (int k = 0; k < S->M; k++) {
    double a = 0, z = 0;
    if (bget(S->Ac_b, k)) {
      a = (double)pop_and(bm_row(&S->At, k), S->Ar_b, S->At.rb);
      if (S->has_Z) z = (double)pop_and(bm_row(&S->Zt, k), S->Zr_b, S->Zt.rb);
    }
    AZ_rsum[k] = 0; AZ_rsum[k] += a; AZ_rsum[k] += z;
  }
  if (c->A_p[0]) {
    FILE *fp = fopen(c->A_p, "rb"); int hdr[3];
    if (!fp || fread(hdr, sizeof(int), 3, fp) != 3 || hdr[1] != S->pcols || fread(S->A_p, sizeof(double), S->pcols, fp) != (size_t)S->pcols) { if (fp) fclose(fp); return fail(R, "cannot read A_p"); }
    fclose(fp);
  } else {
    for (int b = 0; b < S->pcols; b++) S->A_p[b] = 0.5;
    if (c->sparse) for (int b = 0; b < S->pcols; b++) {
      int s0 = MINI(S->M - 1, MAXI(0, b * LKO_POP)), s1 = MINI(S->M - 1, MAXI(0, (b + 1) * LKO_POP - 1));
      double tsum = 0, tnum = 0;
      for (int k = s0; k <= s1; k++) if (bget(S->Ac_b, k)) { tsum += AZ_rsum[k]; tnum += 1; }
      if (tnum == 0) S->A_p[b] = 0.5;
      else { double v = (tsum) / (tnum * (Arb + Zrb)); S->A_p[b] = v < 0.01 ? 0.01 : (v > 0.99 ? 0.99 : v); }
    }
  }
  free(AZ_rsum);
  lko_calc_ajdk(S->A_p, S->pcols, S->ajdk);
  /* bcc_scramble runs after bcc_load (lakcluster_driver.c:56-62): A_p is the loaded data's */
  if (c->scramble_num > 0 && st_scramble(S, R)) return -1;
  return st_arrays(S, R);
}

/* bcc_init (bcc.c:1542-1600) for one bin without Y/W controls, GLOBAL_TEST_TYP2= error:
   bcc_ajdk_init's column masks (bcc.c:117-174; the covariate mask with its first and last
   column forced on), bcc_single_init_M_An's row masks and matrices (bcc.c:415-520; the
   draws of wrap_M_setup_test_excerpt_1 / _0, An_ajdk_v.c:32-98: column by column over every
   row, bit = draw > 1/2; the covariates' column 0 set in every row), then bcc_init_A_p's
   block frequencies (bcc.c:1315-1330).  Draws are glibc rand() from its initial state.
   A mask bit is on in _b when its draw exceeds mrand^2 and in _j when it exceeds mrand, so
   rows and columns may start loaded but inactive. */
#define LK_RAND01 ((double)rand() / (double)RAND_MAX)
#define LK_MAXIMUM(A, B) ((A) > (B) ? (A) : (B))
#define LK_MINIMUM(A, B) ((A) < (B) ? (A) : (B))
static void init_mask(int n, double mrnd, uint8_t *mb, uint8_t *mj)
{
  for (int i = 0; i < n; i++) { double d = LK_RAND01; if (d > mrnd * mrnd) bset(mb, i); if (d > mrnd) bset(mj, i); }
}
static void init_bits(int nrows, int ncols, bm_t *Xn, bm_t *Xt)
{
  *Xn = bm_alloc(nrows, ncols); *Xt = bm_alloc(ncols, nrows);
  for (int k = 0; k < ncols; k++) for (int r = 0; r < nrows; r++)
    if (LK_RAND01 > 0.5) { bset(Xn->b + (size_t)r * Xn->rb, k); bset(Xt->b + (size_t)k * Xt->rb, r); }
}
static int st_init(lko_state *S, lko_result *R)
{
  lko_cfg *c = &S->c;
  if (c->nbins != 1) return fail(R, "bcc_sumscores: only GLOBAL_NBINS= 1 is supported");
  if (c->test_Y_cols != 0) return fail(R, "bcc_sumscores: Y/W genetic controls are not supported (GLOBAL_TEST_Y_n_cols must be 0)");
  if (strcmp(c->test_typ2, "error")) return fail(R, "bcc_sumscores: GLOBAL_TEST_TYP2 must be error (random data)");
  if (!strstr(c->QR, "YnWt") || !strstr(c->QC, "YnWt") || strstr(c->QR, "condense") || strstr(c->QC, "store one")
      || strstr(c->QR, "ZtSWn") || strstr(c->QC, "ZtSWn")) return fail(R, "only the YnWt strategies are supported");

  if (c->test_A_cols < 1 || c->test_T_cols < 1 || c->test_T_cols > 8) return fail(R, "bcc_sumscores: bad problem size");
  const double mrnd = c->test_mrand;
  srand(1);
  S->M = c->test_A_cols; S->T = c->test_T_cols;
  S->NA = MAXI(0, c->test_A_rows + (0 % 3) - 1); S->NZ = MAXI(0, c->test_Z_rows + (0 % 5) - 2);
  S->has_Z = S->NZ > 0;
  S->pcols = lko_psize(S->M);
  S->Ac_b = mask_alloc(S->M); S->Ac_j = mask_alloc(S->M); S->Ac_rtn = mask_alloc(S->M); S->Ac_rmv = mask_alloc(S->M);
  uint8_t *Tc_b = mask_alloc(S->T); S->Tc_j = mask_alloc(S->T);
  init_mask(S->M, mrnd, S->Ac_b, S->Ac_j);
  init_mask(S->T, mrnd, Tc_b, S->Tc_j);
  bset(Tc_b, 0); bset(S->Tc_j, 0); bset(Tc_b, S->T - 1); bset(S->Tc_j, S->T - 1);
  const int tall = popmask(S->Tc_j, S->T) == S->T;
  free(Tc_b);
  S->Ar_b = mask_alloc(S->NA); S->Ar_j = mask_alloc(S->NA); S->Ar_rtn = mask_alloc(S->NA); S->Ar_rmv = mask_alloc(S->NA);
  init_mask(S->NA, mrnd, S->Ar_b, S->Ar_j);
  init_bits(S->NA, S->M, &S->An, &S->At);
  S->Zr_b = mask_alloc(S->NZ); S->Zr_j = mask_alloc(S->NZ);
  init_mask(S->NZ, mrnd, S->Zr_b, S->Zr_j);
  init_bits(S->NZ, S->M, &S->Zn, &S->Zt);
  /* Y and W have no columns: no draws */
  bm_t Tn, Sn;
  init_bits(S->NA, S->T, &Tn, &S->Tt);
  for (int r = 0; r < S->NA; r++) { bset(Tn.b + (size_t)r * Tn.rb, 0); bset(S->Tt.b, r); }
  init_bits(S->NZ, S->T, &Sn, &S->St);
  for (int r = 0; r < S->NZ; r++) { bset(Sn.b + (size_t)r * Sn.rb, 0); bset(S->St.b, r); }
  free(Tn.b); free(Sn.b);
  /* bcc_init_A_p: per block p = maximum(0.1, minimum(0.9, rand01)), the macros drawing again
     for the value they return; 0.5 everywhere when GLOBAL_TEST_sparse is 0 (after the draws) */
  S->A_p = dalloc(S->pcols); S->ajdk = dalloc((size_t)S->pcols * LKO_AJDK);
  for (int b = 0; b < S->pcols; b++) S->A_p[b] = LK_MAXIMUM(0.1, LK_MINIMUM(0.9, LK_RAND01));
  if (c->sparse == 0) for (int b = 0; b < S->pcols; b++) S->A_p[b] = 0.5;
  lko_calc_ajdk(S->A_p, S->pcols, S->ajdk);
  /* the covariates this path carries: every column in use */
  if (!tall) return fail(R, "bcc_sumscores: the drawn covariate mask leaves a column out (not supported)");
  if (popmask(S->Ar_b, S->NA) == 0) return fail(R, "bcc_sumscores: A has no loaded rows");
  S->Ar_bidx = (int *)xcalloc(S->NA + 1, sizeof(int)); S->Ac_bidx = (int *)xcalloc(S->M + 1, sizeof(int));
  for (int i = 0, b = 0; i < S->NA; i++) { S->Ar_bidx[i] = b; if (bget(S->Ar_b, i)) b++; }
  for (int i = 0, b = 0; i < S->M; i++) { S->Ac_bidx[i] = b; if (bget(S->Ac_b, i)) b++; }
  return st_arrays(S, R);
}

/* xcalc orientation flags (bcc_double_init_lf, bcc.c:1005-1150: the first step's row drop
   from the active counts, the sides by their loaded counts) and the score arrays */
static int st_arrays(lko_state *S, lko_result *R)
{
  lko_cfg *c = &S->c;
  (void)R;
  int Arb = popmask(S->Ar_b, S->NA), Zrb = popmask(S->Zr_b, S->NZ);
  int rdrop0 = 0; lko_get_xdrop(popmask(S->Ar_j, S->NA), popmask(S->Ac_j, S->M), c->gamma, c->gamma_type, &rdrop0, NULL);
  int rdrop2 = MAXI(2, rdrop0);
  S->trm_AnZt = (Zrb <= Arb) ? 0 : 1;
  { int n1b = MINI(rdrop2, Arb), n2b = Zrb; S->trm_knzt = (n2b <= n1b) ? 0 : 1; S->trm_hnvt = S->trm_knzt; }
  /* arrays */
  size_t NA = S->NA, NZ = S->NZ, M = S->M, T = S->T;
  S->lf_An_ajdk = dalloc(NA * LKO_AJDK); S->lf_Zn_ajdk = dalloc(NZ * LKO_AJDK);
  S->lf_AtTn = dalloc(M * T); S->lf_ZtSn = dalloc(M * T);
  S->lf_AnZt = dalloc(NA * NZ); S->lf_AnAt = dalloc(NA * NA);
  S->lf_AtTAnZtSZn = dalloc(M * T); S->lf_AtTAnAtTAn = dalloc(M * T);
  S->lf_AnZtSZnAt = dalloc(NA * T); S->lf_AnAtTAnAt = dalloc(NA * T);
  S->At_An_a1d1 = dalloc(M); S->An_a2d2_ZtSn = dalloc(NA * T); S->An_a2d2_AtTn = dalloc(NA * T);
  S->a1d2_ZtSn = dalloc(T); S->a3d2_ZtSn = dalloc(T); S->a1d2_AtTn = dalloc(T); S->a3d2_AtTn = dalloc(T);
  S->et_Tn = dalloc(T); S->et_Sn = dalloc(T); S->et_An = dalloc(M); S->T_AnAt_AnAt = dalloc(T * NA);
  S->QR_ZSZ = dalloc(NA * T); S->QR_TAT = dalloc(NA * T); S->QC_ZSZ = dalloc(M * T); S->QC_TAT = dalloc(M * T);
  S->QR_ZSZ_nrm = dalloc(NA * T); S->QR_TAT_nrm = dalloc(NA * T); S->QC_ZSZ_nrm = dalloc(M * T); S->QC_TAT_nrm = dalloc(M * T);
  S->QC_sv = dalloc(M); S->QR_sv = dalloc(NA);
  S->QC_a = (int *)xcalloc(M + 1, sizeof(int)); S->QC_b = (int *)xcalloc(M + 1, sizeof(int)); S->QC_j = (int *)xcalloc(M + 1, sizeof(int));
  S->QR_a = (int *)xcalloc(NA + 1, sizeof(int)); S->QR_b = (int *)xcalloc(NA + 1, sizeof(int)); S->QR_j = (int *)xcalloc(NA + 1, sizeof(int));
  S->QR_nb = (int *)xcalloc(NA + 1, sizeof(int)); S->QR_ga = (int *)xcalloc(NA + 1, sizeof(int)); S->QR_gb = (int *)xcalloc(NA + 1, sizeof(int));
  S->trace = dalloc((M + NA + 2) * 6); S->xdrop_a = (int *)xcalloc(2 * (M + NA) + 2, sizeof(int)); S->xdrop_b = (int *)xcalloc(2 * (M + NA) + 2, sizeof(int));
  return 0;
}

static void st_free(lko_state *S)
{
  free(S->An.b); free(S->At.b); free(S->Zn.b); free(S->Zt.b); free(S->Tt.b); free(S->St.b);
  void *ps[] = { S->Ar_b, S->Ar_j, S->Ar_rtn, S->Ar_rmv, S->Zr_b, S->Zr_j, S->Ac_b, S->Ac_j, S->Ac_rtn, S->Ac_rmv, S->Tc_j,
                 S->A_p, S->ajdk, S->lf_An_ajdk, S->lf_Zn_ajdk, S->lf_AtTn, S->lf_ZtSn, S->lf_AnZt, S->lf_AnAt,
                 S->lf_AtTAnZtSZn, S->lf_AtTAnAtTAn, S->lf_AnZtSZnAt, S->lf_AnAtTAnAt, S->At_An_a1d1, S->An_a2d2_ZtSn,
                 S->An_a2d2_AtTn, S->a1d2_ZtSn, S->a3d2_ZtSn, S->a1d2_AtTn, S->a3d2_AtTn, S->et_Tn, S->et_Sn, S->et_An,
                 S->T_AnAt_AnAt, S->QR_ZSZ, S->QR_TAT, S->QC_ZSZ, S->QC_TAT, S->QR_ZSZ_nrm, S->QR_TAT_nrm, S->QC_ZSZ_nrm,
                 S->QC_TAT_nrm, S->QC_sv, S->QR_sv, S->QC_a, S->QC_b, S->QC_j, S->QR_a, S->QR_b, S->QR_j, S->QR_nb,
                 S->QR_ga, S->QR_gb, S->Ar_bidx, S->Ac_bidx };
  for (size_t i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
}

/* the per-step rescoring that follows a drop, also used for the initial pass */
static void st_rescore_tail(lko_state *S)
{
  st_QR(S); st_ajdk(S); st_ZtSn(S); st_single(S); st_double(S); st_flatten(S);
}

static int write_outputs(const lko_state *S, const char *dir)
{
  char p[LKO_PATH + 32]; FILE *fp;
  snprintf(p, sizeof p, "%s/out_trace.txt", dir); if (!(fp = fopen(p, "w"))) return -1;
  for (int l = 0; l < S->n_trace; l++) {
    const double *t = S->trace + (size_t)l * 6;
    fprintf(fp, "%d %d %d %0.16f %0.16f %d\n", (int)t[0], (int)t[1], (int)t[2], t[3], t[4], (int)t[5]);
  }
  fclose(fp);
  FILE *fa, *fb;
  snprintf(p, sizeof p, "%s/out_xdrop_a.txt", dir); if (!(fa = fopen(p, "w"))) return -1;
  snprintf(p, sizeof p, "%s/out_xdrop_b.txt", dir); if (!(fb = fopen(p, "w"))) { fclose(fa); return -1; }
  for (int l = 0; l < S->n_xdrop; l++) {
    fprintf(fa, "%d %d\n", S->xdrop_a[2 * l], S->xdrop_a[2 * l + 1]);
    fprintf(fb, "%d %d\n", S->xdrop_b[2 * l], S->xdrop_b[2 * l + 1]);
  }
  fclose(fa); fclose(fb);
  return 0;
}

/* Runs the whole elimination described by an .in text.  max_iter<0 runs to the end.
   out_dir (optional) overrides GLOBAL_DIR_NAME; outputs are written only when
   write_files is nonzero. */
static void *run_impl(const char *in_text, const char *out_dir, int max_iter, int write_files, int recompute)
{
  lko_result *R = (lko_result *)xcalloc(1, sizeof(lko_result));
  lko_state *S = (lko_state *)xcalloc(1, sizeof(lko_state));
  if (cfg_parse(in_text, &S->c)) { fail(R, "bad .in"); free(S); return R; }
  double t0 = now_s();
  if (st_load(S, R)) { st_free(S); free(S); return R; }
  /* bcc_sumscores_test stops after GLOBAL_TEST_niter passes when that is positive
     (bcc_sumscores.c:815) */
  if (!strcmp(S->c.test_type, "bcc_sumscores") && S->c.test_niter > 0 && (max_iter < 0 || max_iter > S->c.test_niter))
    max_iter = S->c.test_niter;
  double t1 = now_s();
  const int dcc = !strcmp(S->c.test_type, "dexcluster_driver");
  /* initial loop-subscores (lakcluster_driver.c:73-106; dexcluster_driver.c:22-25) */
  if (dcc) { st_ajdk(S); st_ZtSn(S); st_dcc(S); }
  else {
    st_ajdk(S); st_ZtSn(S); st_YnWt(S); st_At_T(S); st_QR(S);
    st_single(S); st_double(S); st_flatten(S);
  }
  double t2 = now_s();
  int iter = 0;
  char dir[LKO_PATH];
  if (out_dir && out_dir[0]) snprintf(dir, sizeof dir, "%s", out_dir);
  else if (S->c.dir_name[0]) snprintf(dir, sizeof dir, "%s", S->c.dir_name);
  else if (S->c.out_name[0]) snprintf(dir, sizeof dir, "%s/dir_%s", S->c.dir_base[0] ? S->c.dir_base : ".", S->c.out_name);
  else snprintf(dir, sizeof dir, ".");
  if (write_files) mkdir(dir, 0755);
  while ((popmask(S->Ar_j, S->NA) > 0 || popmask(S->Ac_j, S->M) > 0) && (max_iter < 0 || iter < max_iter)) {
    int Irem = 0, rdrop = 0, cdrop = 0;
    st_nrm_ifT_cmb_x(S, dcc);
    st_xij(S, &Irem);
    st_mxA(S, iter, Irem, &rdrop, &cdrop);
    if (dcc) {
      /* dcc_lrup_mxdup + dcc_M_mxset, then the half-loop subscores from scratch
         (dexcluster_driver.c:43-47) */
      memcpy(S->Ar_j, S->Ar_rtn, lko_bsize(S->NA)); memcpy(S->Ac_j, S->Ac_rtn, lko_bsize(S->M));
      st_ajdk(S); st_ZtSn(S); st_dcc(S);
      iter++;
      continue;
    }
    if (!recompute) st_lrup(S);
    else {
      /* no low-rank update: drop, then every loop-subscore from scratch on the new masks */
      memcpy(S->Ar_j, S->Ar_rtn, lko_bsize(S->NA)); memcpy(S->Ac_j, S->Ac_rtn, lko_bsize(S->M));
      st_ajdk(S); st_ZtSn(S); st_YnWt(S); st_At_T(S);
    }
    st_rescore_tail(S);
    iter++;
  }
  double t3 = now_s();
  if (write_files && write_outputs(S, dir)) fail(R, "cannot write outputs");
  R->n_trace = S->n_trace; R->n_xdrop = S->n_xdrop; R->ops = S->ops; R->n_iter = iter;
  R->trace = S->trace; R->xdrop_a = S->xdrop_a; R->xdrop_b = S->xdrop_b;
  S->trace = NULL; S->xdrop_a = NULL; S->xdrop_b = NULL;
  R->load_s = t1 - t0; R->init_s = t2 - t1; R->loop_s = t3 - t2;
  st_free(S); free(S);
  return R;
}

void *lko_run_text(const char *in_text, const char *out_dir, int max_iter, int write_files)
{
  return run_impl(in_text, out_dir, max_iter, write_files, 0);
}
void *lko_run_text_recompute(const char *in_text, int max_iter)
{
  return run_impl(in_text, NULL, max_iter, 0, 1);
}

int lko_status(void *h, char *err, int n) { lko_result *R = (lko_result *)h; if (err && n > 0) snprintf(err, n, "%s", R->err); return R->status; }
int lko_n_trace(void *h) { return ((lko_result *)h)->n_trace; }
int lko_n_xdrop(void *h) { return ((lko_result *)h)->n_xdrop; }
int lko_n_iter(void *h) { return ((lko_result *)h)->n_iter; }
long long lko_ops(void *h) { return ((lko_result *)h)->ops; }
void lko_times(void *h, double *t3) { lko_result *R = (lko_result *)h; t3[0] = R->load_s; t3[1] = R->init_s; t3[2] = R->loop_s; }
void lko_get_trace(void *h, double *out) { lko_result *R = (lko_result *)h; memcpy(out, R->trace, sizeof(double) * 6 * R->n_trace); }
void lko_get_drops(void *h, int *a, int *b) { lko_result *R = (lko_result *)h; memcpy(a, R->xdrop_a, sizeof(int) * 2 * R->n_xdrop); memcpy(b, R->xdrop_b, sizeof(int) * 2 * R->n_xdrop); }
void lko_free(void *h) { lko_result *R = (lko_result *)h; if (R) { free(R->trace); free(R->xdrop_a); free(R->xdrop_b); free(R); } }

/* ------------------------------------------------------------------ */
/* lakcluster_scorebox (lakcluster_scorebox.c:5-271; S_handle S_init.c:5-89;
   bcc_scorebox_mxA/_svalue/_srt, xcc_scorebox_xdrop_array bcc_scorebox.c:5-133,
   282-308): replays the elimination order of an out_xdrop file over a grid of
   (rows kept, columns kept).  The first grid point recomputes every loop-subscore
   after its drop; the others follow by the low-rank update, rows along the outer
   loop (columns when row_max > col_max) and columns from a copy of that state along
   the inner one.  Each point records the active counts, Irem and the mean row and
   column scores in index order.  Parity pinned by the reference's worked example of
   the schedule and by full recomputation at each grid point
   (tests/test_scorebox.py).                                                       */
/* ------------------------------------------------------------------ */
void lko_scorebox_xdrop_array(int nxols, int x_max, int x_num, int *xval, int *xdrop)
{
  double d = ((double)(x_max - 0)) / ((double)MAXI(1, x_num - 1));
  for (int nx = 0; nx < x_num; nx++) xval[nx] = (int)round(x_max - nx * d);
  xdrop[0] = MAXI(0, nxols - (x_max + 1));
  for (int nx = 1; nx < x_num; nx++) xdrop[nx] = xval[nx - 1] - xval[nx];
}

typedef struct { int *rpop, *cpop, *irem; double *qr, *qc; } lko_sbout;

/* bcc_scorebox_srt + bcc_scorebox_mxA: the current order is the last rpop (cpop) entries
   of the file's row (column) order; its first rdrop (cdrop) entries are removed */
static int sb_drop(lko_state *S, const int *rord, int nro, const int *cord, int nco, int rdrop, int cdrop)
{
  int rpop = popmask(S->Ar_j, S->NA), cpop = popmask(S->Ac_j, S->M);
  if ((nro < rpop && MINI(rdrop, rpop) > 0) || (nco < cpop && MINI(cdrop, cpop) > 0)) return -1;
  size_t nbc = lko_bsize(S->M), nbr = lko_bsize(S->NA);
  memcpy(S->Ac_rtn, S->Ac_j, nbc); memset(S->Ac_rmv, 0, nbc);
  for (int nc = 0; nc < MINI(cdrop, cpop); nc++) {
    int k = cord[nco - cpop + nc];
    if (k < 0 || k >= S->M || !bget(S->Ac_j, k)) return -1;
    bset(S->Ac_rmv, k); S->Ac_rtn[k >> 3] &= (uint8_t)~(1u << (7 - (k & 7)));
  }
  memcpy(S->Ar_rtn, S->Ar_j, nbr); memset(S->Ar_rmv, 0, nbr);
  for (int nr = 0; nr < MINI(rdrop, rpop); nr++) {
    int a = rord[nro - rpop + nr];
    if (a < 0 || a >= S->NA || !bget(S->Ar_j, a)) return -1;
    bset(S->Ar_rmv, a); S->Ar_rtn[a >> 3] &= (uint8_t)~(1u << (7 - (a & 7)));
  }
  return 0;
}
/* lakcluster_scorebox_excerpt_0 with bcc_scorebox_svalue: scores of the active rows and
   columns in index order, first covariate column */
static void sb_record_x(lko_state *S, lko_sbout *B, int tab, int dcc)
{
  st_nrm_ifT_cmb_x(S, dcc);
  ilist ar = il_from_mask(S->Ar_j, S->NA), kl = il_from_mask(S->Ac_j, S->M);
  double *v = dalloc((size_t)MAXI(ar.n, kl.n));
  for (int i = 0; i < ar.n; i++) v[i] = S->QR_TAT_nrm[(size_t)ar.ix[i] * S->T + 0];
  double m = ra_mean(v, (unsigned long long)ar.n); if (!isfinite(m)) m = 1.0;
  B->qr[tab] = m;
  for (int j = 0; j < kl.n; j++) v[j] = S->QC_TAT_nrm[(size_t)kl.ix[j] * S->T + 0];
  m = ra_mean(v, (unsigned long long)kl.n); if (!isfinite(m)) m = 1.0;
  B->qc[tab] = m;
  B->rpop[tab] = ar.n; B->cpop[tab] = kl.n; B->irem[tab] = ar.n > 0 ? 1 : 0;
  free(v); free(ar.ix); free(kl.ix);
}
static void sb_record(lko_state *S, lko_sbout *B, int tab) { sb_record_x(S, B, tab, 0); }
/* dexcluster_scorebox_excerpt_1 / _2 (dexcluster_scorebox.c:21-65): the dcc subscores from
   scratch, after dcc_lrup_mxdup + dcc_M_mxset (the drop of the preceding mxA taking effect)
   only in _2 */
static void sb_dcc(lko_state *S, int dup)
{
  if (dup) { memcpy(S->Ar_j, S->Ar_rtn, lko_bsize(S->NA)); memcpy(S->Ac_j, S->Ac_rtn, lko_bsize(S->M)); }
  st_ajdk(S); st_ZtSn(S); st_dcc(S);
}
static void sb_full(lko_state *S)
{
  memcpy(S->Ar_j, S->Ar_rtn, lko_bsize(S->NA)); memcpy(S->Ac_j, S->Ac_rtn, lko_bsize(S->M));
  st_ajdk(S); st_ZtSn(S); st_YnWt(S); st_At_T(S); st_QR(S);
  st_single(S); st_double(S); st_flatten(S);
}
static void sb_update(lko_state *S) { st_lrup(S); st_rescore_tail(S); }

/* bcc_copy: every array the loop carries */
static void sb_copy(lko_state *D, const lko_state *S)
{
  size_t NA = S->NA, NZ = S->NZ, M = S->M, T = S->T;
  uint8_t *dm[] = { D->Ar_j, D->Ar_rtn, D->Ar_rmv, D->Zr_j, D->Ac_j, D->Ac_rtn, D->Ac_rmv, D->Tc_j };
  const uint8_t *sm[] = { S->Ar_j, S->Ar_rtn, S->Ar_rmv, S->Zr_j, S->Ac_j, S->Ac_rtn, S->Ac_rmv, S->Tc_j };
  size_t mn[] = { NA, NA, NA, NZ, M, M, M, T };
  for (int i = 0; i < 8; i++) memcpy(dm[i], sm[i], lko_bsize((int)mn[i]));
  double *dd[] = { D->lf_An_ajdk, D->lf_Zn_ajdk, D->lf_AtTn, D->lf_ZtSn, D->lf_AnZt, D->lf_AnAt, D->lf_AtTAnZtSZn,
                   D->lf_AtTAnAtTAn, D->lf_AnZtSZnAt, D->lf_AnAtTAnAt, D->At_An_a1d1, D->An_a2d2_ZtSn, D->An_a2d2_AtTn,
                   D->a1d2_ZtSn, D->a3d2_ZtSn, D->a1d2_AtTn, D->a3d2_AtTn, D->et_Tn, D->et_Sn, D->et_An, D->T_AnAt_AnAt,
                   D->QR_ZSZ, D->QR_TAT, D->QC_ZSZ, D->QC_TAT, D->QR_ZSZ_nrm, D->QR_TAT_nrm, D->QC_ZSZ_nrm, D->QC_TAT_nrm };
  const double *sd[] = { S->lf_An_ajdk, S->lf_Zn_ajdk, S->lf_AtTn, S->lf_ZtSn, S->lf_AnZt, S->lf_AnAt, S->lf_AtTAnZtSZn,
                         S->lf_AtTAnAtTAn, S->lf_AnZtSZnAt, S->lf_AnAtTAnAt, S->At_An_a1d1, S->An_a2d2_ZtSn, S->An_a2d2_AtTn,
                         S->a1d2_ZtSn, S->a3d2_ZtSn, S->a1d2_AtTn, S->a3d2_AtTn, S->et_Tn, S->et_Sn, S->et_An, S->T_AnAt_AnAt,
                         S->QR_ZSZ, S->QR_TAT, S->QC_ZSZ, S->QC_TAT, S->QR_ZSZ_nrm, S->QR_TAT_nrm, S->QC_ZSZ_nrm, S->QC_TAT_nrm };
  size_t dn[] = { NA * LKO_AJDK, NZ * LKO_AJDK, M * T, M * T, NA * NZ, NA * NA, M * T, M * T, NA * T, NA * T, M,
                  NA * T, NA * T, T, T, T, T, T, T, M, T * NA, NA * T, NA * T, M * T, M * T, NA * T, NA * T, M * T, M * T };
  for (size_t i = 0; i < sizeof dd / sizeof dd[0]; i++) memcpy(dd[i], sd[i], sizeof(double) * dn[i]);
  D->et_An_a1d1 = S->et_An_a1d1;
}

/* Runs the scorebox of an .in text.  out gets row_num x col_num entries at
   [nr + nc*row_num]; returns 0, or -1 with a message in err. */
static int sb_run(const char *in_text, lko_sbout *B, char *err, int nerr, int *dims, int recompute)
{
  lko_result R0; memset(&R0, 0, sizeof R0);
  lko_state *S = (lko_state *)xcalloc(1, sizeof(lko_state)), *U = (lko_state *)xcalloc(1, sizeof(lko_state));
  int rc = -1;
  int *rord = NULL, *cord = NULL, nro = 0, nco = 0;
  if (cfg_parse(in_text, &S->c) || cfg_parse(in_text, &U->c)) { snprintf(err, nerr, "bad .in"); goto done; }
  {
    lko_cfg *c = &S->c;
    const int row_num = c->sb_row_num, col_num = c->sb_col_num;
    if (dims) { dims[0] = c->sb_row_max; dims[1] = row_num; dims[2] = c->sb_col_max; dims[3] = col_num; }
    if (row_num < 1 || col_num < 1) { snprintf(err, nerr, "scorebox grid needs row_num, col_num >= 1"); goto done; }
    if (!B) { rc = 0; goto done; }
    if (st_load(S, &R0) || st_load(U, &R0)) { snprintf(err, nerr, "%s", R0.err); goto done; }
    FILE *fp = fopen(c->sb_file, "r");
    if (!fp) { snprintf(err, nerr, "cannot open the scorebox out_xdrop file"); goto done; }
    int nl = 0; for (int ch; (ch = fgetc(fp)) != EOF;) nl += ch == '\n';
    fseek(fp, 0, SEEK_SET);
    rord = (int *)xcalloc((size_t)nl + 1, sizeof(int)); cord = (int *)xcalloc((size_t)nl + 1, sizeof(int));
    for (int l = 0; l < nl; l++) {
      int a = -1, b = -1;
      if (fscanf(fp, "%d %d", &a, &b) != 2) break;
      if (a >= 0) rord[nro++] = a;
      if (b >= 0) cord[nco++] = b;
    }
    fclose(fp);
    int *rkeep = (int *)xcalloc(row_num, sizeof(int)), *rdrop = (int *)xcalloc(row_num, sizeof(int));
    int *ckeep = (int *)xcalloc(col_num, sizeof(int)), *cdrop = (int *)xcalloc(col_num, sizeof(int));
    lko_scorebox_xdrop_array(popmask(S->Ac_j, S->M), c->sb_col_max, col_num, ckeep, cdrop);
    lko_scorebox_xdrop_array(popmask(S->Ar_j, S->NA), c->sb_row_max, row_num, rkeep, rdrop);
    const int rows_outer = !(c->sb_row_max > c->sb_col_max);   /* lakcluster_scorebox_cr, else _rc */
    const int n_out = rows_outer ? row_num : col_num, n_in = rows_outer ? col_num : row_num;
    int bad = 0;
    /* dexcluster_scorebox_rc / _cr (dexcluster_scorebox.c:67-165): D_ori along the outer loop,
       D_sub along the inner one; D_sub is loaded once and never copied from D_ori, so it
       carries its own drops from one inner loop to the next (and never the outer loop's);
       the first point's drops (excerpt_1) do not take effect */
    const int dcc = !strcmp(c->test_type, "dexcluster_scorebox");
    for (int o = 0; o < n_out && !bad && dcc; o++) {
      const int r_o = o == 0 ? rdrop[0] : (rows_outer ? rdrop[o] : 0), c_o = o == 0 ? cdrop[0] : (rows_outer ? 0 : cdrop[o]);
      if (sb_drop(S, rord, nro, cord, nco, r_o, c_o)) { bad = 1; break; }
      sb_dcc(S, o > 0);
      sb_record_x(S, B, rows_outer ? o : o * row_num, 1);
      for (int i = 1; i < n_in; i++) {
        const int r_i = rows_outer ? 0 : rdrop[i], c_i = rows_outer ? cdrop[i] : 0;
        if (sb_drop(U, rord, nro, cord, nco, r_i, c_i)) { bad = 1; break; }
        sb_dcc(U, 1);
        sb_record_x(U, B, rows_outer ? o + i * row_num : i + o * row_num, 1);
      }
    }
    for (int o = 0; o < n_out && !bad && !dcc; o++) {
      const int r_o = rows_outer ? rdrop[o] : 0, c_o = rows_outer ? 0 : cdrop[o];
      if (o == 0) {
        if (sb_drop(S, rord, nro, cord, nco, rdrop[0], cdrop[0])) { bad = 1; break; }
        sb_full(S);
      } else {
        if (sb_drop(S, rord, nro, cord, nco, r_o, c_o)) { bad = 1; break; }
        if (recompute) sb_full(S); else sb_update(S);
      }
      sb_record(S, B, rows_outer ? o : o * row_num);
      sb_copy(U, S);
      for (int i = 1; i < n_in; i++) {
        const int r_i = rows_outer ? 0 : rdrop[i], c_i = rows_outer ? cdrop[i] : 0;
        if (sb_drop(U, rord, nro, cord, nco, r_i, c_i)) { bad = 1; break; }
        if (recompute) sb_full(U); else sb_update(U);
        sb_record(U, B, rows_outer ? o + i * row_num : i + o * row_num);
      }
    }
    free(rkeep); free(rdrop); free(ckeep); free(cdrop);
    if (bad) { snprintf(err, nerr, "the out_xdrop order does not cover the active rows and columns"); goto done; }
    rc = 0;
  }
done:
  free(rord); free(cord);
  st_free(S); st_free(U); free(S); free(U);
  return rc;
}

int lko_scorebox(const char *in_text, int recompute, int *dims, int *rpop, int *cpop, int *irem, double *qr,
                 double *qc, char *err, int nerr)
{
  if (!rpop || !cpop || !irem || !qr || !qc) return sb_run(in_text, NULL, err, nerr, dims, recompute);
  lko_sbout B = { rpop, cpop, irem, qr, qc };
  return sb_run(in_text, &B, err, nerr, dims, recompute);
}

/* ------------------------------------------------------------------ */
/* stepwise access to the state, for stage-level comparisons in tests  */
/* ------------------------------------------------------------------ */
typedef struct { lko_state S; lko_result R; int iter; } lko_stepper;

void *lko_state_new(const char *in_text)
{
  lko_stepper *P = (lko_stepper *)xcalloc(1, sizeof(lko_stepper));
  if (cfg_parse(in_text, &P->S.c) || st_load(&P->S, &P->R)) { st_free(&P->S); free(P); return NULL; }
  if (!strcmp(P->S.c.test_type, "dexcluster_driver")) { st_ajdk(&P->S); st_ZtSn(&P->S); st_dcc(&P->S); return P; }
  st_ajdk(&P->S); st_ZtSn(&P->S); st_YnWt(&P->S); st_At_T(&P->S); st_QR(&P->S);
  st_single(&P->S); st_double(&P->S); st_flatten(&P->S);
  return P;
}

int lko_state_iterate(void *h)
{
  lko_stepper *P = (lko_stepper *)h; lko_state *S = &P->S;
  if (!(popmask(S->Ar_j, S->NA) > 0 || popmask(S->Ac_j, S->M) > 0)) return 1;
  int Irem = 0, rdrop = 0, cdrop = 0;
  const int dcc = !strcmp(S->c.test_type, "dexcluster_driver");
  st_nrm_ifT_cmb_x(S, dcc); st_xij(S, &Irem); st_mxA(S, P->iter, Irem, &rdrop, &cdrop);
  if (dcc) {
    memcpy(S->Ar_j, S->Ar_rtn, lko_bsize(S->NA)); memcpy(S->Ac_j, S->Ac_rtn, lko_bsize(S->M));
    st_ajdk(S); st_ZtSn(S); st_dcc(S);
  } else { st_lrup(S); st_rescore_tail(S); }
  P->iter++;
  return (popmask(S->Ar_j, S->NA) > 0 || popmask(S->Ac_j, S->M) > 0) ? 0 : 1;
}

/* copies the named array (oracle layout) into out; returns its length, or -1 */
long lko_state_get(void *h, const char *name, double *out, long n)
{
  lko_state *S = &((lko_stepper *)h)->S;
  size_t NA = S->NA, NZ = S->NZ, M = S->M, T = S->T;
  struct { const char *nm; const double *p; size_t len; } tab[] = {
    { "lf_An_ajdk", S->lf_An_ajdk, NA * LKO_AJDK }, { "lf_Zn_ajdk", S->lf_Zn_ajdk, NZ * LKO_AJDK },
    { "lf_AtTn", S->lf_AtTn, M * T }, { "lf_ZtSn", S->lf_ZtSn, M * T },
    { "lf_AnZt", S->lf_AnZt, NA * NZ }, { "lf_AnAt", S->lf_AnAt, NA * NA },
    { "lf_AtTAnZtSZn", S->lf_AtTAnZtSZn, M * T }, { "lf_AtTAnAtTAn", S->lf_AtTAnAtTAn, M * T },
    { "lf_AnZtSZnAt", S->lf_AnZtSZnAt, NA * T }, { "lf_AnAtTAnAt", S->lf_AnAtTAnAt, NA * T },
    { "At_An_a1d1", S->At_An_a1d1, M }, { "An_a2d2_ZtSn", S->An_a2d2_ZtSn, NA * T }, { "An_a2d2_AtTn", S->An_a2d2_AtTn, NA * T },
    { "a1d2_AtTn", S->a1d2_AtTn, T }, { "a3d2_AtTn", S->a3d2_AtTn, T }, { "a1d2_ZtSn", S->a1d2_ZtSn, T }, { "a3d2_ZtSn", S->a3d2_ZtSn, T },
    { "et_Tn", S->et_Tn, T }, { "et_Sn", S->et_Sn, T }, { "et_An", S->et_An, M }, { "et_An_a1d1", &S->et_An_a1d1, 1 },
    { "T_AnAt_AnAt", S->T_AnAt_AnAt, T * NA },
    { "QR_ZSZ", S->QR_ZSZ, NA * T }, { "QR_TAT", S->QR_TAT, NA * T }, { "QC_ZSZ", S->QC_ZSZ, M * T }, { "QC_TAT", S->QC_TAT, M * T },
    { "ajdk", S->ajdk, (size_t)S->pcols * LKO_AJDK }, { "A_p", S->A_p, (size_t)S->pcols },
  };
  struct { const char *nm; const uint8_t *m; int n; } mt[] = {
    { "mask_Ar_j", S->Ar_j, S->NA }, { "mask_Zr_j", S->Zr_j, S->NZ }, { "mask_Ac_j", S->Ac_j, S->M },
    { "mask_Ar_b", S->Ar_b, S->NA }, { "mask_Zr_b", S->Zr_b, S->NZ }, { "mask_Ac_b", S->Ac_b, S->M },
  };
  /* the bit matrices as 0/1 values, row-major (after any scramble) */
  if (!strcmp(name, "An_bits") || !strcmp(name, "At_bits") || !strcmp(name, "Zn_bits") || !strcmp(name, "Tt_bits") ||
      !strcmp(name, "St_bits")) {
    const bm_t *B = name[0] == 'Z' ? &S->Zn : name[0] == 'T' ? &S->Tt : name[0] == 'S' ? &S->St
                  : (name[1] == 'n' ? &S->An : &S->At);
    const long len = (long)B->nr * B->nc;
    if (out) for (long e = 0; e < MINI(len, n); e++) out[e] = bget(bm_row(B, (int)(e / B->nc)), e % B->nc);
    return len;
  }
  for (size_t i = 0; i < sizeof mt / sizeof mt[0]; i++) {
    if (strcmp(mt[i].nm, name)) continue;
    if (out) for (long j = 0; j < MINI((long)mt[i].n, n); j++) out[j] = bget(mt[i].m, j);
    return mt[i].n;
  }
  for (size_t i = 0; i < sizeof tab / sizeof tab[0]; i++) {
    if (strcmp(tab[i].nm, name)) continue;
    if (out) memcpy(out, tab[i].p, sizeof(double) * (size_t)MINI((long)tab[i].len, n));
    return (long)tab[i].len;
  }
  return -1;
}

void lko_state_free(void *h)
{
  lko_stepper *P = (lko_stepper *)h;
  if (!P) return;
  free(P->S.trace); free(P->S.xdrop_a); free(P->S.xdrop_b);
  st_free(&P->S); free(P);
}

/* get_An_ZtSn_ww with its xcalc (An_ZtSn_ww.c:44-70, xcalc.c:41-246) on caller buffers:
   L[i*ns + s] for n-row nlist[i]; out[alist[a]*ns + s] */
void lko_an_ztsn_ww_L(const uint8_t *A, size_t rbA, int nr_total, const int *alist, int na, const uint8_t *amask,
                      const double *L, const int *nlist, int nn, int n_total, int ns, int B, long long D_MLT,
                      double *out)
{
  lko_planes P; memset(&P, 0, sizeof P);
  lko_xcalc(L, nlist, nn, n_total, ns, B, D_MLT, &P, NULL);
  bm_t m; m.b = (uint8_t *)A; m.nr = nr_total; m.nc = n_total; m.rb = rbA;
  int *sl = (int *)xcalloc(ns + 1, sizeof(int));
  for (int s = 0; s < ns; s++) sl[s] = s;
  an_ztsn_ww(&m, alist, na, amask, &P, sl, ns, out, ns);
  free(sl); lko_planes_free(&P);
}
