/* stage_loop.c -- a C caller of include/lkc.h that runs lakcluster_driver's loop stage by
   stage (lakcluster_driver.c:73-246) on one handle and the whole run (lkc_bcc_run) on a
   second handle of the same .in file, and requires the two traces and drop lists to be
   identical byte for byte.  It also checks the stage API's order rules: a loop stage
   before the initial pass and bcc_sumscores_mxA without bcc_sumscores_xij are refused.

   usage: stage_loop file.in [out_dir]
   With out_dir, bcc_sumscores_dmp writes the three output files there after every pass,
   as the reference driver does.  Exit status 0 when everything matches. */
#include "../../include/lkc.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CALL(x) do { int rc_ = (x); if (rc_) { \
  fprintf(stderr, "stage_loop: %s failed (%d): %s\n", #x, rc_, lkc_bcc_error(S)); return 1; } } while (0)

static char* read_file(const char* path)
{
  FILE* fp = fopen(path, "rb");
  if (!fp) return NULL;
  size_t cap = 1 << 16, n = 0, k;
  char* buf = malloc(cap);
  while (buf && (k = fread(buf + n, 1, cap - n - 1, fp)) > 0) {
    n += k;
    if (n + 1 >= cap) { cap *= 2; buf = realloc(buf, cap); }
  }
  fclose(fp);
  if (buf) buf[n] = 0;
  return buf;
}

static int initial_pass(lkc_bcc* S)
{
  CALL(lkc_bcc_An_ajdk(S));
  CALL(lkc_bcc_lf_ZtSn(S));
  CALL(lkc_bcc_lf_YnWt(S));
  CALL(lkc_bcc_lf_At_T_YnWt_S_Zn(S));
  CALL(lkc_bcc_lf_AnZt_S_WnYt(S));
  CALL(lkc_bcc_singlestudy_ww(S));
  CALL(lkc_bcc_doublestudy_ww(S));
  CALL(lkc_bcc_flattenloop(S));
  return 0;
}

static int one_pass(lkc_bcc* S, const char* out_dir)
{
  CALL(lkc_bcc_sumscores_nrm(S));
  CALL(lkc_bcc_sumscores_ifT(S));
  CALL(lkc_bcc_sumscores_srt(S));
  CALL(lkc_bcc_sumscores_cmb(S));
  CALL(lkc_bcc_sumscores_xij(S));
  CALL(lkc_bcc_sumscores_mxA(S));
  if (out_dir) CALL(lkc_bcc_sumscores_dmp(S, out_dir));
  CALL(lkc_bcc_lrup_QR_YnWt_stage_0(S));
  CALL(lkc_bcc_lrup_QR_YnWt_stage_1(S));
  CALL(lkc_bcc_lrup_QR_YnWt_stage_3(S));
  CALL(lkc_bcc_lrup_QC_ZtSWn_stage_0(S));
  CALL(lkc_bcc_lrup_QC_ZtSWn_stage_1(S));
  CALL(lkc_bcc_lrup_QC_ZtSWn_stage_2(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_a0(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_a1(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_a2(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_a3(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_b0(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_b1(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_b2(S));
  CALL(lkc_bcc_lrup_mxdup(S));
  CALL(lkc_bcc_M_mxset(S));
  CALL(lkc_bcc_M_ZtSWn(S));
  CALL(lkc_bcc_M_YnWt(S));
  CALL(lkc_bcc_lf_AtTYn____WtSZn(S));
  CALL(lkc_bcc_lrup_QC_YnWt_stage_c(S));
  CALL(lkc_bcc_lf_AnZt_S_WnYt(S));
  CALL(lkc_bcc_An_ajdk(S));
  CALL(lkc_bcc_lf_ZtSn(S));
  CALL(lkc_bcc_singlestudy_ww(S));
  CALL(lkc_bcc_doublestudy_ww(S));
  CALL(lkc_bcc_flattenloop(S));
  return 0;
}

/* the trace and drop lists of a handle, in malloc'd arrays */
static int results(const lkc_bcc* D, int* nt, double** tr, int* nx, int** xa, int** xb)
{
  *nt = lkc_bcc_n_trace(D); *nx = lkc_bcc_n_xdrop(D);
  if (*nt < 0 || *nx < 0) return 1;
  *tr = malloc(sizeof(double) * 6 * (*nt + 1));
  *xa = malloc(sizeof(int) * 2 * (*nx + 1)); *xb = malloc(sizeof(int) * 2 * (*nx + 1));
  if (lkc_bcc_get_trace(D, *tr) || lkc_bcc_get_xdrop(D, *xa, *xb)) return 1;
  return 0;
}

int main(int argc, char** argv)
{
  if (argc < 2) { fprintf(stderr, "usage: stage_loop file.in [out_dir]\n"); return 2; }
  char* text = read_file(argv[1]);
  if (!text) { fprintf(stderr, "stage_loop: cannot read %s\n", argv[1]); return 2; }
  const char* out_dir = argc > 2 ? argv[2] : NULL;

  lkc_bcc *R = NULL, *S = NULL;
  if (lkc_bcc_load(text, &R) || lkc_bcc_load(text, &S)) {
    fprintf(stderr, "stage_loop: load: %s\n", lkc_last_error());
    return 1;
  }
  /* order rules: nothing of the loop before the initial pass */
  if (lkc_bcc_sumscores_nrm(S) != LKC_ERR_STATE || lkc_bcc_lrup_QC_YnWt_stage_a0(S) != LKC_ERR_STATE) {
    fprintf(stderr, "stage_loop: a loop stage ran before the initial pass\n");
    return 1;
  }
  if (initial_pass(S)) return 1;
  /* _mxA needs the sort of its own pass */
  if (lkc_bcc_sumscores_mxA(S) != LKC_ERR_STATE) {
    fprintf(stderr, "stage_loop: bcc_sumscores_mxA ran without bcc_sumscores_xij\n");
    return 1;
  }
  int passes = 0;
  while (lkc_bcc_active(S)) {
    if (one_pass(S, out_dir)) return 1;
    passes++;
  }
  if (lkc_bcc_run(R, -1)) { fprintf(stderr, "stage_loop: run: %s\n", lkc_bcc_error(R)); return 1; }

  int nt1, nx1, nt2, nx2, *xa1, *xb1, *xa2, *xb2; double *tr1, *tr2;
  if (results(S, &nt1, &tr1, &nx1, &xa1, &xb1) || results(R, &nt2, &tr2, &nx2, &xa2, &xb2)) {
    fprintf(stderr, "stage_loop: cannot read the results\n");
    return 1;
  }
  int ok = nt1 == nt2 && nx1 == nx2 && passes == nt1 && lkc_bcc_n_iter(S) == lkc_bcc_n_iter(R) &&
           memcmp(tr1, tr2, sizeof(double) * 6 * nt1) == 0 && memcmp(xa1, xa2, sizeof(int) * 2 * nx1) == 0 &&
           memcmp(xb1, xb2, sizeof(int) * 2 * nx1) == 0;
  printf("stage_loop: %d passes stage by stage, %d in the whole run, %d drop lines: %s\n", passes, nt2, nx2,
         ok ? "identical" : "DIFFERENT");
  free(tr1); free(tr2); free(xa1); free(xa2); free(xb1); free(xb2);
  lkc_bcc_free(R); lkc_bcc_free(S); free(text);
  return ok ? 0 : 1;
}
