/*
 * lkc.h -- C ABI of the MI355X (gfx950) implementation of lakcluster's `bcc`
 * loop-counting elimination (adirangan/lakcluster_c, dir_c/lakcluster_driver.c).
 *
 * The library owns all device memory.  Every entry point returns 0 on success
 * and a negative status on failure (the message is available from
 * lkc_last_error / lkc_bcc_error); nothing in the library calls exit().
 * One handle is one elimination; a handle may be used from one host thread at
 * a time, different handles are independent.
 *
 * Scope of the path (as in the reference driver with these settings): one bin
 * (GLOBAL_NBINS= 1), no Y/W genetic controls (GLOBAL_Y_n_cols= 0), QR/QC
 * strategies "YnWt", at most 8 covariate columns; bcc_scramble of the loaded A
 * (GLOBAL_scramble_num) and the lakcluster_scorebox replay are supported.  The
 * reference's dcc self-test (lkc_dcc_sumscores_test) runs several bins with Y/W
 * genetic controls on its synthetic problem.
 */
#ifndef LKC_H
#define LKC_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define LKC_OK 0
#define LKC_ERR_CONFIG -1     /* unknown key or unsupported setting in the .in text */
#define LKC_ERR_IO -2         /* cannot read or write a file */
#define LKC_ERR_SHAPE -3      /* input files disagree in shape */
#define LKC_ERR_HIP -4        /* HIP runtime error (no device, out of memory, launch failure) */
#define LKC_ERR_STATE -5      /* call out of order or bad argument */

typedef struct lkc_bcc lkc_bcc;

/* Last error message of the calling thread (load failures, kernel-level calls). */
const char *lkc_last_error(void);
/* Error message recorded in a handle. */
const char *lkc_bcc_error(const lkc_bcc *D);

/* ---- whole run ------------------------------------------------------------ */
/* lakcluster_driver (lakcluster_driver.c:14-251) driven by the key/value text of
   an .in file (updateglobals, updateglobals.c:5-199).  Writes out_trace.txt,
   out_xdrop_a.txt, out_xdrop_b.txt into out_dir (NULL: GLOBAL_DIR_NAME).  The bcc
   elimination whatever GLOBAL_TEST_TYPE says, as the reference function. */
int lkc_lakcluster_driver(const char *in_text, const char *out_dir);

/* dexcluster_driver (dexcluster_driver.c:5-57, GLOBAL_TEST_TYPE= dexcluster_driver at
   lakcluster_ver18.c:345): the dcc half-loop elimination, recomputing its subscores
   (dcc_An_ajdk, dcc_lf_ZtSn, dcc_lf_D_AtTn_ZtSn_vv, dcc_lf_TAnZtS_ww,
   dcc_wrap_dcc_halfloop) after each drop; the same three output files.  With
   GLOBAL_NBINS > 1 or Y/W files it runs lkc_dcc_driver_bins. */
int lkc_dexcluster_driver(const char *in_text, const char *out_dir);

/* dexcluster_driver with several bins and Y/W genetic controls: dcc_load (dcc.c:1157-1215)
   reads bin n's files from item n of the comma-separated GLOBAL_A_n_name_, A_t_name_,
   A_n_rind_, Z_n_name_, Z_t_name_, Z_n_rind_, Y_n_name_, Y_t_name_, W_n_name_, W_t_name_,
   T_n_name_, T_t_name_, S_n_name_, S_t_name_ lists (updateglobals.c:76-117) with the
   shared GLOBAL_A_n_cind / Y_n_cind / T_n_cind masks; A_p and Y_p come from the data
   (dcc_load_A_p, dcc.c:875-991).  The subscores of every bin pair are recomputed after
   each drop (dexcluster_driver.c:22-53); the same three output files, row indices over
   the bins in bin order; GLOBAL_A_p_name / GLOBAL_Y_p_name, when given, replace the data's
   frequencies.  Every bin names the same kinds of file.  Any number of bins
   from 1 to 64 (1 bin without Y gives what lkc_dexcluster_driver gives). */
int lkc_dcc_driver_bins(const char *in_text, const char *out_dir);

/* dcc_sumscores_test (dcc_sumscores.c:700-747; GLOBAL_TEST_TYPE= dcc_sumscores at
   lakcluster_ver18.c:343): the synthetic problem of dcc_init (dcc.c:1217-1275) drawn from
   the GLOBAL_TEST_* keys (GLOBAL_TEST_TYP2= error) with glibc rand()'s sequence from its
   initial state, GLOBAL_NBINS bins, Y/W genetic controls of GLOBAL_TEST_Y_n_cols columns;
   GLOBAL_TEST_niter passes of the dcc half-loop elimination (to the end when <= 0); writes
   out_trace.txt, out_xdrop_a.txt, out_xdrop_b.txt into out_dir (NULL: GLOBAL_DIR_NAME, else
   GLOBAL_DIR_BASE/dir_<GLOBAL_out_name>). */
int lkc_dcc_sumscores_test(const char *in_text, const char *out_dir);

/* bcc_sumscores_test (bcc_sumscores.c:781-880; GLOBAL_TEST_TYPE= bcc_sumscores at
   lakcluster_ver18.c:335): the synthetic problem of bcc_init (bcc.c:1542-1600; bcc_ajdk_init,
   bcc_single_init_M_An, bcc_init_A_p) drawn from the GLOBAL_TEST_* keys with glibc rand()
   from its initial state, then lakcluster_driver's stages for GLOBAL_TEST_niter passes (all
   of them when 0); the three output files.  One bin, GLOBAL_TEST_Y_n_cols= 0,
   GLOBAL_TEST_TYP2= error, the YnWt strategies; every covariate column must be drawn active.
   lkc_bcc_load on such an .in gives the same problem as a handle.  The reference's _uu
   error printouts are not produced. */
int lkc_bcc_sumscores_test(const char *in_text, const char *out_dir);

/* lakcluster_scorebox (lakcluster_scorebox.c:265-271, dispatched on GLOBAL_TEST_TYPE=
   lakcluster_scorebox at lakcluster_ver18.c:338): replays the elimination order of
   GLOBAL_scorebox_out_xdrop over the grid of GLOBAL_scorebox_row_num x _col_num
   (rows kept, columns kept) points and writes the S_handle_dmp files
   (S_init.c:79-89) scorebox_r<row_max>n<row_num>_c<col_max>n<col_num>_{A_rpop_j_total,
   A_cpop_j,Irem,QR_avg,QC_avg}.mda into out_dir (NULL: GLOBAL_DIR_NAME). */
int lkc_lakcluster_scorebox(const char *in_text, const char *out_dir);

/* dexcluster_scorebox (dexcluster_scorebox.c:221-227, dispatched on GLOBAL_TEST_TYPE=
   dexcluster_scorebox at lakcluster_ver18.c:346): the same grid replay on the dcc
   half-loop elimination (dexcluster_scorebox_rc / _cr, :67-165), the subscores recomputed
   from the masks at every point, and the same five .mda files.  As in the reference, the
   inner loop's problem (D_sub) is loaded once and never copied from the outer one's: it
   keeps its own drops across inner loops and never takes the outer loop's, and the first
   point's drops do not take effect.  lkc_lakcluster_scorebox on a dexcluster_scorebox .in
   runs the bcc replay, as the reference function would. */
int lkc_dexcluster_scorebox(const char *in_text, const char *out_dir);
/* xcc_scorebox_xdrop_array (bcc_scorebox.c:282-308): xval[x_num] index values from
   x_max down to 0, xdrop[x_num] the number of indices removed at each step */
int lkc_scorebox_xdrop_array(int nxols, int x_max, int x_num, int *xval, int *xdrop);

/* ---- elimination handle: the bcc stage API -------------------------------- */
/* bcc_load + bcc_init (bcc.c:1482-1540, 1369-1438): reads the .b16 inputs and
   masks, computes A_p and the A_ajdk tables, uploads everything to device 0. */
int lkc_bcc_load(const char *in_text, lkc_bcc **out);
/* the initial loop-subscores (lakcluster_driver.c:73-106): bcc_An_ajdk,
   bcc_lf_ZtSn, bcc_lf_YnWt, bcc_M_YnWt + bcc_lf_At_T_YnWt_S_Zn,
   bcc_lf_AnZt_S_WnYt, bcc_singlestudy_ww, bcc_doublestudy_ww, bcc_flattenloop */
int lkc_bcc_init_lf(lkc_bcc *D);
/* one pass of the elimination loop (lakcluster_driver.c:107-246): bcc_sumscores_
   nrm/ifT/cmb/xij/mxA, bcc_lrup_QR_YnWt_stage_0..3, bcc_lrup_QC_YnWt_stage_a0..c,
   bcc_lrup_mxdup, then the re-scoring stages.  *done is set to 1 once no row and
   no column remains active. */
int lkc_bcc_iterate(lkc_bcc *D, int *done);
/* init (if not yet done) then iterate until done or max_iter passes (max_iter<0: no limit) */
int lkc_bcc_run(lkc_bcc *D, int max_iter);
/* bcc_sumscores_dmp output files (bcc_sumscores.c:127-157 format of out_trace /
   out_xdrop_a / out_xdrop_b) */
int lkc_bcc_write_outputs(const lkc_bcc *D, const char *out_dir);
/* The scorebox on a freshly loaded handle (lakcluster_scorebox_rc / _cr,
   lakcluster_scorebox.c:96-263): S_init_bcc, then per grid point bcc_scorebox_srt +
   bcc_scorebox_mxA, the loop-subscores (recomputed at the first point, by the low-rank
   update after it; the inner loop from a copy of the outer state, bcc_copy), then
   bcc_sumscores_nrm/ifT/srt/cmb + bcc_scorebox_svalue + lakcluster_scorebox_excerpt_0. */
int lkc_bcc_scorebox(lkc_bcc *D);
/* its results: dims4 = {row_max, row_num, col_max, col_num}; each non-NULL array gets the
   row_num*col_num grid entries at [nr + nc*row_num] (the S_handle arrays) */
int lkc_bcc_get_scorebox(const lkc_bcc *D, int *dims4, int *rpop, int *cpop, int *irem, double *qr_avg,
                         double *qc_avg);
/* S_handle_dmp (S_init.c:79-89) into out_dir (NULL: GLOBAL_DIR_NAME) */
int lkc_bcc_write_scorebox(const lkc_bcc *D, const char *out_dir);
/* 1 when the handle's GLOBAL_TEST_TYPE is lakcluster_scorebox (lakcluster_ver18.c:337-338) */
int lkc_bcc_is_scorebox(const lkc_bcc *D);
/* 1 when the handle runs the dcc elimination (GLOBAL_TEST_TYPE= dexcluster_driver): init,
   iterate and run then follow dexcluster_driver (dexcluster_driver.c:22-53) */
int lkc_bcc_is_dcc(const lkc_bcc *D);

/* The individual stages, in the order lakcluster_driver calls them
   (lakcluster_driver.c:73-246; declared at lakcluster_header.h:413-502).  They act on
   the device-resident state of D.  The initial pass, in this order: */
int lkc_bcc_An_ajdk(lkc_bcc *D);                 /* bcc_lf.c:7-51 */
int lkc_bcc_lf_ZtSn(lkc_bcc *D);                 /* bcc_lf.c:53-101 */
int lkc_bcc_lf_YnWt(lkc_bcc *D);                 /* bcc_lf.c:223-302 */
int lkc_bcc_lf_At_T_YnWt_S_Zn(lkc_bcc *D);       /* bcc_lf.c:318-356 + 636-709 (bcc_M_YnWt inside) */
int lkc_bcc_lf_AnZt_S_WnYt(lkc_bcc *D);          /* bcc_lf.c:360-409 */
int lkc_bcc_singlestudy_ww(lkc_bcc *D);          /* bcc_flattenloop.c:753-831 */
int lkc_bcc_doublestudy_ww(lkc_bcc *D);          /* bcc_flattenloop.c:977-1029 */
int lkc_bcc_flattenloop(lkc_bcc *D);             /* bcc_flattenloop.c:1431-1889; ends the initial pass */

/* One pass of the loop, while lkc_bcc_active(D) (the driver's
   `while (D->A_rpop_j_total>0 || D->A_cpop_j>0)`), in this order; each returns
   LKC_ERR_STATE before the initial pass.  With one bin and GLOBAL_Ireq <= 0: */
int lkc_bcc_active(const lkc_bcc *D);            /* 1 while a row or a column is active */
int lkc_bcc_sumscores_nrm(lkc_bcc *D);           /* bcc_sumscores.c:550-779 */
int lkc_bcc_sumscores_ifT(lkc_bcc *D);           /* bcc_sumscores.c:483-548 */
int lkc_bcc_sumscores_srt(lkc_bcc *D);           /* bcc_sumscores.c:358-481: nothing to do with Ireq <= 0 */
int lkc_bcc_sumscores_cmb(lkc_bcc *D);           /* bcc_sumscores.c:324-356 */
int lkc_bcc_sumscores_xij(lkc_bcc *D);           /* bcc_sumscores.c:225-322 with dQuickSort_xij */
int lkc_bcc_sumscores_mxA(lkc_bcc *D);           /* bcc_sumscores.c:159-224 with _mxB, _mxC (37-125), bcc_lrup_mxset;
                                                    LKC_ERR_STATE unless _xij ran in this pass */
int lkc_bcc_sumscores_dmp(const lkc_bcc *D, const char *out_dir);   /* bcc_sumscores.c:127-157 */
int lkc_bcc_lrup_QR_YnWt_stage_0(lkc_bcc *D);    /* bcc_lrup.c:659-701 */
int lkc_bcc_lrup_QR_YnWt_stage_1(lkc_bcc *D);    /* bcc_lrup.c:702-771, with stage 3's subtraction */
int lkc_bcc_lrup_QR_YnWt_stage_3(lkc_bcc *D);    /* bcc_lrup.c:863-895: done inside stage 1 */
int lkc_bcc_lrup_QC_ZtSWn_stage_0(lkc_bcc *D);   /* bcc_lrup.c:896-938: no-op for QC_strategy YnWt */
int lkc_bcc_lrup_QC_ZtSWn_stage_1(lkc_bcc *D);   /* bcc_lrup.c:939-978: no-op for QC_strategy YnWt */
int lkc_bcc_lrup_QC_ZtSWn_stage_2(lkc_bcc *D);   /* bcc_lrup.c:979-1058: no-op for QC_strategy YnWt */
int lkc_bcc_lrup_QC_YnWt_stage_a0(lkc_bcc *D);   /* bcc_lrup.c:5-66 */
int lkc_bcc_lrup_QC_YnWt_stage_a1(lkc_bcc *D);   /* bcc_lrup.c:67-196 */
int lkc_bcc_lrup_QC_YnWt_stage_a2(lkc_bcc *D);   /* bcc_lrup.c:197-237 (the Grams quantised) */
int lkc_bcc_lrup_QC_YnWt_stage_a3(lkc_bcc *D);   /* bcc_lrup.c:238-304 */
int lkc_bcc_lrup_QC_YnWt_stage_b0(lkc_bcc *D);   /* bcc_lrup.c:305-362 */
int lkc_bcc_lrup_QC_YnWt_stage_b1(lkc_bcc *D);   /* bcc_lrup.c:363-408; on the matrix cores it also
                                                    forms stage b2's sums (the triples are not kept) */
int lkc_bcc_lrup_QC_YnWt_stage_b2(lkc_bcc *D);   /* bcc_lrup.c:409-482 */
int lkc_bcc_lrup_mxdup(lkc_bcc *D);              /* bcc_lrup.c:1128-1139 */
int lkc_bcc_M_mxset(lkc_bcc *D);                 /* bcc.c:1414-1438: the counts follow the masks
                                                    that lkc_bcc_lrup_mxdup uploads; no-op */
int lkc_bcc_M_ZtSWn(lkc_bcc *D);                 /* bcc_lf.c:176-222: no-op for QC_strategy YnWt */
int lkc_bcc_M_YnWt(lkc_bcc *D);                  /* bcc_lf.c:318-359: each Gram is quantised where it
                                                    is contracted (stage a2, lf_AnZt_S_WnYt); no-op */
int lkc_bcc_lf_AtTYn____WtSZn(lkc_bcc *D);       /* bcc_lf.c:566-635: no-op for QC_strategy YnWt */
int lkc_bcc_lrup_QC_YnWt_stage_c(lkc_bcc *D);    /* bcc_lrup.c:592-655 */
/* then lkc_bcc_lf_AnZt_S_WnYt, lkc_bcc_An_ajdk, lkc_bcc_lf_ZtSn, lkc_bcc_singlestudy_ww,
   lkc_bcc_doublestudy_ww, lkc_bcc_flattenloop (lakcluster_driver.c:225-245). */

/* Grouped: lkc_bcc_sumscores = nrm, ifT, srt, cmb, xij, mxA and bcc_lrup_mxdup;
   lkc_bcc_lrup = the QR and QC stages above through stage c, then bcc_lrup_mxdup if
   still pending (the update reads the _rtn / _rmv sets, not the active ones). */
int lkc_bcc_sumscores(lkc_bcc *D);
int lkc_bcc_lrup(lkc_bcc *D);

/* dQuickSort_xij (Quicksort.c:124-141) with one companion index array: sorts d[l..r] in
   place by the reference's first-element-pivot partition (idx permuted alongside), stops
   recursing at depth GLOBAL_recursion_limit = 32768 as the reference does, and returns the
   deepest level reached.  The order the elimination's score sort must reproduce; host only. */
unsigned int lkc_dquicksort_xij(double *d, int *idx, int l, int r);

/* get_xdrop (xdrop.c:5-64): the rows and columns one pass drops from lrij rows and
   lcij columns, with gamma and gamma_type passed in (GLOBAL_gamma, GLOBAL_gamma_type:
   0 log, 1 sqrt, 2 lin, 3 rows); either output may be NULL. */
int lkc_get_xdrop(double lrij, double lcij, double gamma, int gamma_type, int *rdrop, int *cdrop);
/* get_xdrop_length (xdrop.c:66-83; declared get_xdrop_total at lakcluster_header.h:748):
   the number of passes; negative on a non-finite count */
int lkc_get_xdrop_length(double lrij, double lcij, double gamma, int gamma_type);
/* get_xdrop_array (xdrop.c:85-111) into caller arrays of cap entries (each may be NULL):
   the drops and the counts kept after each pass; *length = the number of passes.  With
   every array NULL only *length is set; LKC_ERR_STATE when cap < *length. */
int lkc_get_xdrop_array(double lrij, double lcij, double gamma, int gamma_type, int cap, int *length,
                        int *rdrop, int *cdrop, int *rkeep, int *ckeep);

/* ---- results ------------------------------------------------------------- */
int lkc_bcc_n_iter(const lkc_bcc *D);
int lkc_bcc_n_trace(const lkc_bcc *D);
/* 6 doubles per line: iteration, rows, cols, mean QR, mean QC, Irem */
int lkc_bcc_get_trace(const lkc_bcc *D, double *out);
int lkc_bcc_n_xdrop(const lkc_bcc *D);
/* two ints per line each: a = (row, col) original indices, b = (row, col) loaded indices */
int lkc_bcc_get_xdrop(const lkc_bcc *D, int *a, int *b);
/* the reference's bit-operation accounting, summed over all stage calls */
long long lkc_bcc_ops(const lkc_bcc *D);
/* Score sorts (of a column or a row list) that ran the reference's quicksort on the host
   (dQuickSort_xij, Quicksort.c:83-141) instead of taking the device radix sort's order:
   ties or NaN in the dropped prefix, or ties anywhere in a list longer than the
   quicksort's recursion limit (GLOBAL_recursion_limit = 32768, lakcluster_ver18.c:68).
   Counted since load or the last reset. */
long lkc_bcc_sort_fallbacks(const lkc_bcc *D);
/* seconds: [load, init, loop] */
int lkc_bcc_times(const lkc_bcc *D, double *t3);
/* problem size: [rows of A, rows of Z, columns, covariate columns] */
int lkc_bcc_shape(const lkc_bcc *D, int *out4);
void lkc_bcc_free(lkc_bcc *D);
/* Back to the state right after lkc_bcc_load (all rows and columns active, no
   trace), keeping the device-resident inputs; the next run starts with init. */
int lkc_bcc_reset(lkc_bcc *D);
/* HIP-event timing of kernel groups on the handle's stream (off by default).
   lkc_bcc_kernel_stats reports, for a group name ("ww": the At_T_YnWt_S_Zn_ww
   contraction), the launches, the summed device milliseconds and the summed
   algorithmic work (int8 multiply-adds x2 of the exact integer contraction)
   since the last reset. */
int lkc_bcc_set_profiling(lkc_bcc *D, int on);
/* Implementation of the contractions (the centred AnZt_vv Grams, At_T_YnWt_S_Zn_ww
   and the AtTYn_vv terms of the update): 0 = int8 matrix cores (default), 1 = VALU
   popcount over packed bits (and the B_MLT bit planes).  Both are bit-exact. */
int lkc_bcc_set_kernels(lkc_bcc *D, int which);
int lkc_bcc_kernel_stats(const lkc_bcc *D, const char *name, long long *launches, double *ms, double *work);
/* Launch-shape autotuner (no counterpart in the reference, whose SSE loops have one form;
   for callers that run many problems of varying shape, such as the batch and serving
   front-ends).  The ww contraction (fused, or split over w steps and n parts) and the
   full-layout Grams (one block per tile, or split over the 1920-blocks) have forms with
   bit-identical results; with tuning on, the first launch of each shape class (sizes to
   half an octave) times every form with HIP events on the handle's stream and keeps the
   fastest; off, a cost model picks.  mode 0 off, 1 measure new classes (the default, or
   LKC_TUNE=0..3 in the environment), 2 use stored classes only, 3 measure and check that
   every form wrote the same bits as the first (lkc_tune_verify: launches compared and
   8-byte words that differed, 0 when the forms agree).  The store is shared by
   the process; lkc_tune_save / lkc_tune_load write and read it as text (return the number
   of classes).  lkc_tune_get: mode, stored classes, classes measured, and the summed
   device milliseconds the tuned forms saved over the cost model's at measurement. */
int lkc_tune_set(int mode);
int lkc_tune_get(int *mode, int *classes, long long *measured, double *saved_ms);
int lkc_tune_verify(long long *compared, long long *mismatched);
int lkc_tune_save(const char *path);
int lkc_tune_load(const char *path);
int lkc_tune_clear(void);
/* State inspection (no counterpart in the reference, which keeps these arrays in
   struct bcc_ajdk/bcc_single/bcc_double): copies the named score array, e.g.
   "lf_AnZt", "QC_TAT", "et_An", in row-major (row, column) order with the
   covariate last, into out (at most n values); returns its length or -1. */
long lkc_bcc_debug_get(lkc_bcc *D, const char *name, double *out, long n);

/* ---- single kernels on host buffers (copy in, run on device 0, copy out) --- */
/* Bit rows: row r of X starts at X + r*rb; bit i of a row is bit (7 - i%8) of byte i/8; 1 <=> +1.
   rb must be at least bsize(ncols) = ceil(round16(ncols)/1920)*240 bytes. */
size_t lkc_bsize(int ncols);
/* get_An_ajdk_v (An_ajdk_v.c:272-363): out[i*15 + jk] for rows[i] */
int lkc_An_ajdk_v(const uint8_t *X, size_t rb, int nrows_total, const int *rows, int nrows,
                  const uint8_t *colmask, int ncols, const double *ajdk, double *out);
/* get_AnZt_vv (AnZt_vv.c:5-223): out[i*nz + j] for (arows[i], zrows[j]); ajdk==NULL gives the
   plain +-1 product, else the centred D-weighted form with xa[i], xz[j] (the rows' lf_An_ajdk a1d1 values) */
int lkc_AnZt_vv(const uint8_t *A, size_t rbA, int na_total, const int *arows, int na,
                const uint8_t *Z, size_t rbZ, int nz_total, const int *zrows, int nz,
                const uint8_t *colmask, int ncols, const double *ajdk,
                const double *xa, const double *xz, double *out);
/* get_xcalc (xcalc.c:41-246): L[i*nw + j] for n-row nrows_list[i] (in [0, n_total)) and w j.
   Returns the quantised integers in the reference's n_total-long layout, cz[j*n_total +
   nrows_list[i]] (0 at positions no list entry names: the planes' clear bits), and
   [min, mlt]. */
int lkc_xcalc(const double *L, const int *nrows_list, int nn, int n_total, int nw, int B,
              long long D_MLT, unsigned long long *cz, double *min_mlt);
/* get_An_ZtSn_ww (An_ZtSn_ww.c:5-155) with its xcalc: L[i*ns + s] holds the (n-row
   nlist[i], s) values quantised to B-bit planes over n_total bits; A rows (rbA bytes)
   are contracted with the planes under amask; out[a*ns + s] for a in alist, 0 elsewhere. */
int lkc_An_ZtSn_ww(const uint8_t *A, size_t rbA, int nr_total, const int *alist, int na, const uint8_t *amask,
                   const double *L, const int *nlist, int nn, int n_total, int ns, int B, long long D_MLT, double *out);
/* No counterpart in the reference: lkc_An_ZtSn_ww for callers that cannot allocate device
   memory (code under HIP graph capture, serving loops that keep their buffers, harnesses that
   forbid hipMalloc).  Every pointer is device memory: A holds nr_total rows of
   lkc_bsize(n_total) bytes, amask lkc_bsize(n_total) bytes, alist/nlist na/nn row indices
   (which must lie in [0, nr_total) / [0, n_total): they are not read on the host), L the
   nn x ns values, out nr_total*ns doubles, out[a*ns + s] as lkc_An_ZtSn_ww (0 for rows not in
   alist).  The intermediates live in a fixed 32 MiB device array per GPU, held until the
   process ends; a call needing more (lkc's estimate: ns*n_total*4 + ns*B*bsize(n_total) +
   ns*nr_total*8 bytes plus about 10 KB) fails with LKC_ERR_STATE.  The work is enqueued on
   `stream` (a hipStream_t; NULL: the null stream) and the call returns without waiting.
   Calls that overlap on one GPU share the array and race; calls on one stream are ordered. */
int lkc_An_ZtSn_ww_noalloc(const uint8_t *A, int nr_total, const int *alist, int na, const uint8_t *amask,
                           const double *L, const int *nlist, int nn, int n_total, int ns, int B, long long D_MLT,
                           double *out, void *stream);
/* wrap_transpose (binary_read.c:174-269), on the device: Y (ncols rows of rbY >=
   bsize(nrows) bytes) receives the transpose of X (nrows rows of ncols bits), padding
   bits zero.  lkc_bcc_load does the same when a .in names only one of *_n / *_t. */
int lkc_transpose_bits(const uint8_t *X, size_t rbX, int nrows, int ncols, uint8_t *Y, size_t rbY);

/* ---- many eliminations on one GPU: permutation-null batches and serving ------
   (no counterpart in the reference, which runs one problem per process; the nulls of
   its MATLAB drivers, e.g. dir_m/lakcluster_uADZSZDA_shuffle_1.m:1-2, re-run the binary
   once per shuffle.)  A pool keeps the patients' bits on the device; a job takes rows of
   it as A (cases) and Z (controls), with an optional column mask, and runs the same
   elimination as lkc_bcc_run on a .in naming those rows and columns as masks. */
typedef struct lkc_pool lkc_pool;
/* rows: nrows patients of ncols bits (row stride rb >= bsize(ncols)); cov: ncov covariate
   columns as rows of nrows bits (stride rbc >= bsize(nrows)) */
int lkc_pool_create(const uint8_t *rows, size_t rb, int nrows, int ncols,
                    const uint8_t *cov, size_t rbc, int ncov, lkc_pool **out);
/* A synthetic pool drawn on the device: entry (r, j) of nrows x ncols bits is 1 with
   probability p_j = pmin + (pmax - pmin) U(seed, ~0, j), drawn as U(seed, r, j) < p_j, where U is
   a counter-based uniform (the splitmix64 finaliser of seed, row and column; every entry can be
   recomputed on its own); then block k (plant_m[k] rows listed in plant_rows, plant_mc[k]
   columns in plant_cols, both consecutive over k) is overwritten with plant_bits (row-major
   0/1 bytes).  cov as in lkc_pool_create. */
int lkc_pool_create_random(int nrows, int ncols, double pmin, double pmax, unsigned long long seed, int nplant,
                           const int *plant_m, const int *plant_mc, const int *plant_rows, const int *plant_cols,
                           const uint8_t *plant_bits, const uint8_t *cov, size_t rbc, int ncov, lkc_pool **out);
/* rows r0 .. r0+n-1 of the pool, lkc_bsize(ncols) bytes each, into out */
int lkc_pool_read(const lkc_pool *P, int r0, int n, uint8_t *out);
void lkc_pool_free(lkc_pool *P);
typedef struct lkc_job_params {
  double gamma; int gamma_type; int B_MLT; double kappa_squared; int sparse;
} lkc_job_params;
/* one job as a handle (then lkc_bcc_run / lkc_bcc_get_trace as usual) */
int lkc_bcc_load_job(const lkc_pool *P, const int *rowsA, int nA, const int *rowsZ, int nZ,
                     const uint8_t *colmask, const lkc_job_params *prm, lkc_bcc **out);
typedef struct lkc_job {
  const int *rowsA; int nA; const int *rowsZ; int nZ; const uint8_t *colmask;
  lkc_job_params params; double arrival_s;   /* seconds after the call starts */
} lkc_job;
typedef struct lkc_job_result {
  int status; int n_iter; long long ops; double arrival_s, start_s, end_s;
  long n_trace; double *trace;               /* 6 doubles per line, as lkc_bcc_get_trace */
  long n_xdrop; int *xdrop_a; int *xdrop_b;  /* 2 ints per line, as lkc_bcc_get_xdrop */
} lkc_job_result;
/* Runs the jobs, at most max_concurrent at a time, each admitted in order once it has
   arrived and its device-memory estimate fits hbm_budget_bytes (<= 0: 80% of the free
   device memory plus what the library's stream-ordered memory pool holds idle).  Results are allocated by the library; free them with
   lkc_batch_results_free.  Returns LKC_OK if the batch ran (per-job status inside). */
int lkc_batch_run(const lkc_pool *P, const lkc_job *jobs, int njobs, int max_concurrent,
                  double hbm_budget_bytes, lkc_job_result *results);
void lkc_batch_results_free(lkc_job_result *results, int njobs);
/* the device-memory estimate the admission uses for one job */
double lkc_job_bytes(int nA, int nZ, int ncols, int B_MLT);

#ifdef __cplusplus
}
#endif
#endif
