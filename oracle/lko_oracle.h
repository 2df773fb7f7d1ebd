/*
 * lko_oracle.h -- CPU oracle for the lakcluster `bcc` elimination path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; nothing in the product links it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.
 *
 * It restates, in plain C, the arithmetic of the reference's
 * lakcluster_driver() path (adirangan/lakcluster_c, dir_c/lakcluster_driver.c:14-251)
 * for one bin, no Y/W genetic controls and the "YnWt" QR/QC strategies.
 * Every floating-point expression keeps the reference's evaluation order;
 * every function cites the reference file:line it restates.
 */
#ifndef LKO_ORACLE_H
#define LKO_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define LKO_POP 1920          /* bits per weighting block (lakcluster_header.h:268) */
#define LKO_POPB 240          /* bytes per weighting block */
#define LKO_AJDK 15           /* number of alpha^j D^k tables (lakcluster_header.h:77) */

/* bytes of one packed bit-row of n entries, padded to 16 then to 1920 bits
   (bsize, lakcluster_header.h:262) */
size_t lko_bsize(int n);
/* number of 1920-bit blocks (psize, lakcluster_header.h:263) */
int lko_psize(int n);

/* ---- scalar helpers exported for tests ---- */
/* get_xdrop (xdrop.c:5-64); gamma/gamma_type passed explicitly */
void lko_get_xdrop(double lrij, double lcij, double gamma, int gamma_type, int *rdrop, int *cdrop);
/* calc_A_ajdk (An_ajdk_v.c:5-30): out[nc + jk*pcols] */
void lko_calc_ajdk(const double *p, int pcols, double *out);
/* dQuickSort_xij (Quicksort.c:83-141) on d with up to six companion int arrays (may be NULL) */
unsigned int lko_quicksort_xij(double *d, int *i0, int *i1, int *i2, int *i3, int *i4, int *i5, int l, int r);

/* ---- kernels (compact outputs; lists are ascending original indices) ---- */
/* get_An_ajdk_v (An_ajdk_v.c:272-363): out[r*15+jk] for rows in rows[] */
void lko_an_ajdk_v(const uint8_t *X, size_t rb, const int *rows, int nrows,
                   const uint8_t *colmask, int ncols, const double *ajdk, double *out);
/* get_AnZt_vv__run (AnZt_vv.c:5-183), TYPE_pm.  If ajdk==NULL the plain +-1 product
   is formed (and xa,xz are ignored).  out[i*nz + j]. */
void lko_anzt_vv(const uint8_t *A, size_t rbA, const int *arows, int na,
                 const uint8_t *Z, size_t rbZ, const int *zrows, int nz,
                 const uint8_t *colmask, int ncols, const double *ajdk,
                 const double *xa, const double *xz, double *out);
/* get_xcalc (xcalc.c:41-246): quantises L[i*nw + j] (i over n-rows list, j over w list)
   into B bit planes.  Returns min/mlt; cz[j*nn + i] holds the integers (for tests). */
typedef struct {
  int B; int n_total; size_t rb;   /* planes are bit rows over the n dimension */
  int nw;                           /* number of active w */
  uint8_t *row0;                    /* bsize(n_total) bytes: the active n mask */
  uint8_t *planes;                  /* nw*B rows of rb bytes; plane (j,nb) holds bit nb */
  double min_d, max_d, mlt_d;
} lko_planes;
void lko_xcalc(const double *L, const int *nrows_list, int nn, int n_total, int nw,
               int B, long long D_MLT, lko_planes *out, unsigned long long *cz);
void lko_planes_free(lko_planes *p);

/* whole-run driver (lakcluster_driver.c:14-251); returns an opaque result handle */
void *lko_run_text(const char *in_text, const char *out_dir, int max_iter, int write_files);
/* the same elimination with the low-rank update replaced by recomputing every
   loop-subscore from scratch after each drop (a check of the update) */
void *lko_run_text_recompute(const char *in_text, int max_iter);
int lko_status(void *h, char *err, int n);
int lko_n_trace(void *h);
int lko_n_xdrop(void *h);
int lko_n_iter(void *h);
long long lko_ops(void *h);
void lko_times(void *h, double *t3);
void lko_get_trace(void *h, double *out);
void lko_get_drops(void *h, int *a, int *b);
void lko_free(void *h);

/* get_An_ZtSn_ww with its xcalc on caller buffers: L[i*ns + s] for n-row nlist[i];
   rows of A have rbA = lko_bsize(n_total) bytes; out[alist[a]*ns + s] */
void lko_an_ztsn_ww_L(const uint8_t *A, size_t rbA, int nr_total, const int *alist, int na, const uint8_t *amask,
                      const double *L, const int *nlist, int nn, int n_total, int ns, int B, long long D_MLT,
                      double *out);

/* stepwise access for stage-level comparisons: load + initial pass, one loop
   pass (returns 1 when no row or column remains), named arrays in the layouts
   of the restatement ([row*T + t], [a*NZ + z], [r*15 + jk]) */
void *lko_state_new(const char *in_text);
int lko_state_iterate(void *S);
long lko_state_get(void *S, const char *name, double *out, long n);
void lko_state_free(void *S);

/* lakcluster_scorebox (lakcluster_scorebox.c:5-271): grid entries at [nr + nc*row_num];
   dims = [row_max, row_num, col_max, col_num] (with NULL outputs only dims are read);
   recompute=1 recomputes every loop-subscore at each grid point instead of the
   low-rank update.  xval/xdrop: xcc_scorebox_xdrop_array (bcc_scorebox.c:282-308). */
int lko_scorebox(const char *in_text, int recompute, int *dims, int *rpop, int *cpop, int *irem, double *qr,
                 double *qc, char *err, int nerr);
void lko_scorebox_xdrop_array(int nxols, int x_max, int x_num, int *xval, int *xdrop);

/* dcc_sumscores_test (dcc_sumscores.c:700-747) on dcc_init's synthetic problem
   (dcc.c:1217-1275; glibc rand() from its initial state): several bins, Y/W genetic
   controls, GLOBAL_Ireq = 0.  Writes out_trace.txt / out_xdrop_a.txt / out_xdrop_b.txt
   into out_dir when it is not NULL; trace holds 6 doubles per line, the xdrop
   arrays 2 ints per line (lko_dcc_bins.c). */
typedef struct {
  int nbins, A_n_rows, A_n_cols, Z_n_rows, Y_n_cols, T_n_cols, B_MLT, niter, sparse, gamma_type;
  double mrand, gamma, kappa2;
} lko_dcc_params;
/* dexcluster_driver on the .in's per-bin files (several bins, Y/W controls): runs to the
   end and writes out_trace.txt / out_xdrop_a.txt / out_xdrop_b.txt into out_dir */
int lko_dcc_driver_bins(const char *in_text, const char *out_dir, int *n_trace, int *n_xdrop);
int lko_dcc_sumscores_test(const lko_dcc_params *p, const char *out_dir, double *trace, int *n_trace,
                           int *xdrop_a, int *xdrop_b, int *n_xdrop);

#ifdef __cplusplus
}
#endif
#endif
