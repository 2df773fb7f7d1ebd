/*
 * lko_dcc_bins.c -- CPU oracle for the dcc half-loop elimination with several bins
 * and Y/W genetic controls, on the reference's own synthetic problem and on the
 * .in's per-bin files.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Restates, in plain C:
 *   - dcc_init (dcc.c:1217-1275): dcc_ajdk_init (dcc.c:117-174), dcc_single_init_M_An
 *     (dcc.c:543-619) with wrap_M_setup_test_excerpt_0/1 (An_ajdk_v.c:33-98),
 *     dcc_init_A_p (dcc.c:993-1011), dcc_init_QX (dcc.c:1047-1090); every draw is
 *     glibc rand() from its initial state, in the reference's order;
 *   - the half-loop subscores of every bin pair: dcc_An_ajdk, dcc_lf_ZtSn
 *     (dcc_lf.c:7-101), dcc_lf_D_AtTn_ZtSn_vv (dcc_lf.c:102-163; D_AtTn_ZtSn.c:5-310),
 *     dcc_lf_TAnZtS_ww (TAnZtS.c:5-540), dcc_wrap_dcc_halfloop (dcc_lf.c:379-530);
 *   - dcc_sumscores_nrm / ifT / srt / cmb / xij / mxA / mxB / mxC / dmp
 *     (dcc_sumscores.c:5-699), dcc_lrup_mxdup (dcc_lf.c:531-542), dcc_M_mxset
 *     (dcc.c:1092-1113);
 *   - the loop of dcc_sumscores_test (dcc_sumscores.c:700-747);
 *   - dcc_load (dcc.c:1157-1215): the per-bin file lists of updateglobals
 *     (updateglobals.c:76-117), dcc_ajdk_load's column masks (dcc.c:54-115),
 *     dcc_single_load_M_An (dcc.c:251-419), dcc_load_A_p (dcc.c:875-991), and the
 *     dexcluster_driver loop to the end (dexcluster_driver.c:5-57).
 * Bit matrices are kept whole ("a" indices); loaded-row and loaded-column masks
 * (*_b) and active masks (*_j, a subset) select what the reference's compacted
 * storage would hold.  Double expressions keep the reference's order (build with
 * -ffp-contract=off).
 */
#include "lko_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

/* the reference's macros (lakcluster_header.h:257-261): arguments are evaluated
   again in the chosen branch, so a draw inside them can be taken more than once */
#define maximum(A,B) ((A) > (B) ? (A) : (B))
#define minimum(A,B) ((A) < (B) ? (A) : (B))
#define rand01 ((double)rand()/(double)RAND_MAX)

#define NQ 4   /* the four half-loop terms, in the order TYnWtS, TAnZtS, TYnYtT, TAnAtT */
enum { Q_YW = 0, Q_AZ = 1, Q_YY = 2, Q_AA = 3 };

static void *zalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s ? s : 1); if (!p) { fprintf(stderr, "lko_dcc_bins: out of memory\n"); abort(); } return p; }
static inline int bg(const uint8_t *b, long i) { return (b[i >> 3] >> (7 - (i & 7))) & 1; }
static inline void bs(uint8_t *b, long i) { b[i >> 3] |= (uint8_t)(1u << (7 - (i & 7))); }
static inline int bpm(const uint8_t *b, long i) { return bg(b, i) ? 1 : -1; }   /* bget____ */
static long pop_m(const uint8_t *x, const uint8_t *m, size_t nb) { long s = 0; for (size_t i = 0; i < nb; i++) s += __builtin_popcount(x[i] & m[i]); return s; }
static long pop_xm(const uint8_t *x, const uint8_t *z, const uint8_t *m, size_t nb) { long s = 0; for (size_t i = 0; i < nb; i++) s += __builtin_popcount((x[i] ^ z[i]) & m[i]); return s; }
static int popn(const uint8_t *m, int n) { int s = 0; for (int i = 0; i < n; i++) s += bg(m, i); return s; }
static uint8_t *mask_new(int n) { return (uint8_t *)zalloc(lko_bsize(n) + 8, 1); }

/* sum_b w[b] * |m in block b|, blocks in order (the a0d1 / a2d1 sums, TAnZtS.c:152-157) */
static double mask_lf(const uint8_t *m, int pcols, const double *w)
{
  double s = 0;
  for (int b = 0; b < pcols; b++) s += (double)w[b] * (double)pop_m(m + (size_t)b * LKO_POPB, m + (size_t)b * LKO_POPB, LKO_POPB);
  return s;
}

typedef struct { uint8_t *b; int nr, nc; size_t rb; } dm;     /* nr packed rows of nc bits */
static dm dm_new(int nr, int nc) { dm m; m.nr = nr; m.nc = nc; m.rb = lko_bsize(nc); m.b = (uint8_t *)zalloc((size_t)(nr > 0 ? nr : 1) * m.rb + 8, 1); return m; }
static inline const uint8_t *row(const dm *m, int r) { return m->b + (size_t)r * m->rb; }

typedef struct { int *ix; int n; } il;
static il il_mask(const uint8_t *m, int n) { il L; L.ix = (int *)zalloc(n + 1, sizeof(int)); L.n = 0; for (int i = 0; i < n; i++) if (bg(m, i)) L.ix[L.n++] = i; return L; }

typedef struct {
  int nA, nZ;
  uint8_t *Ar_b, *Ar_j, *Ar_rtn, *Ar_rmv, *Zr_b, *Zr_j;
  dm An, At, Zn, Zt, Yn, Yt, Wn, Wt, Tn, Tt, Sn, St;
  double *lfA, *lfY, *lfZ, *lfW;                 /* lf_An_ajdk, lf_Yn_ajdk, lf_Zn_ajdk, lf_Wn_ajdk: [r*15 + jk] */
  double *AtTn, *YtTn, *ZtSn, *WtSn;             /* plain +-1 products [k*T + t] */
  double A_a0d1, A_a2d1, Y_a0d1, Y_a2d1;
  double *a1_AtTn, *a1_ZtSn, *a1_YtTn, *a1_WtSn, *et_Tn, *et_Sn;   /* [T] */
  lko_planes P_WtSn, P_ZtSn, P_YtTn, P_AtTn;     /* the quantised a0d1-weighted products */
  double *QRn[NQ];                               /* E->QR_*_nrm: [nb2*nA*T + c*nA + r] */
  int *g_a, *g_b;                                /* QR_index_global_mr_a / _b */
} dbin;

typedef struct {
  double *ww[NQ];   /* lf_T*_ww of stage 2: [r*T + t] over the rows of nb1 */
  double *qr[NQ];   /* F->QR_* after the half-loop: [r*T + tb], tb the loaded-covariate index */
  double *qc[NQ];   /* F->QC_*: [k*T + tb] over Y columns (TYnWtS, TYnYtT) or A columns */
} dpair;

typedef struct {
  lko_dcc_params p;
  int nbins, M, NY, T, B; long long D_MLT;
  uint8_t *Ac_b, *Ac_j, *Ac_rtn, *Ac_rmv, *Yc_b, *Yc_j, *Tc_b, *Tc_j;
  int Apc, Ypc; double *A_p, *A_ajdk, *Y_p, *Y_ajdk;
  int *tb;                                        /* loaded index of every covariate column */
  dbin *E; dpair *F;
  double *QCn[NQ];                                /* D->QC_*_nrm: [nbx*X*T + c*X + k], X = NY or M */
  int A_nrows_total, A_rpop_b_total, A_rpop_j_total, Z_rpop_b_total, Z_rpop_j_total, A_cpop_j, Y_cpop_j, T_cpop_j, Irem;
  double *QC_sv, *QR_sv; int *QC_a, *QC_b, *QC_j, *QR_a, *QR_b, *QR_j, *QR_nb, *QR_ga, *QR_gb;
  int n_trace; double *trace; int n_xdrop; int *xdrop_a, *xdrop_b;
} dcc_t;

static int qcols(const dcc_t *D, int q) { return (q == Q_YW || q == Q_YY) ? D->NY : D->M; }
/* the `bother` flag of a bin pair's term (dcc_lf.c:118-154, TAnZtS.c:409-416, dcc_sumscores.c:503-559) */
static int on(const dcc_t *D, int q, int nb1, int nb2)
{
  if (D->E[nb1].nA <= 0) return 0;
  switch (q) { case Q_YW: return D->NY > 0 && D->E[nb2].nZ > 0; case Q_AZ: return D->M > 0 && D->E[nb2].nZ > 0;
               case Q_YY: return D->NY > 0 && D->E[nb2].nA > 0; default: return D->M > 0 && D->E[nb2].nA > 0; }
}

/* wrap_M_setup_test_excerpt_0, "error" (An_ajdk_v.c:33-52): column-major draws, bit = draw > 1/2 */
static void gen_bits(int nrows, int ncols, dm *Xn, dm *Xt)
{
  *Xn = dm_new(nrows, ncols); *Xt = dm_new(ncols, nrows);
  double mrnd = maximum(0, minimum(1, 0.5));
  for (int nc = 0; nc < ncols; nc++) for (int nr = 0; nr < nrows; nr++) {
    double dtmp = rand01;
    if (dtmp > mrnd) { bs(Xn->b + (size_t)nr * Xn->rb, nc); bs(Xt->b + (size_t)nc * Xt->rb, nr); }
  }
}
/* a row/column mask pair as dcc_ajdk_init / dcc_single_init_M_An draw it */
static void gen_mask(int n, double mrnd, uint8_t *mb, uint8_t *mj)
{
  for (int nx = 0; nx < n; nx++) { double dtmp = rand01; if (dtmp > mrnd * mrnd) bs(mb, nx); if (dtmp > mrnd) bs(mj, nx); }
}

static void dcc_alloc(dcc_t *D);

/* dcc_init (dcc.c:1217-1275) */
static void dcc_init(dcc_t *D, const lko_dcc_params *p)
{
  memset(D, 0, sizeof *D); D->p = *p;
  D->nbins = p->nbins; D->M = p->A_n_cols; D->NY = p->Y_n_cols; D->T = p->T_n_cols;
  D->B = p->B_MLT; D->D_MLT = 1LL << p->B_MLT;
  srand(1);    /* glibc: the state rand() starts from when nothing seeds it */
  double mrnd = p->mrand;
  D->Ac_b = mask_new(D->M); D->Ac_j = mask_new(D->M); D->Ac_rtn = mask_new(D->M); D->Ac_rmv = mask_new(D->M);
  D->Yc_b = mask_new(D->NY); D->Yc_j = mask_new(D->NY); D->Tc_b = mask_new(D->T); D->Tc_j = mask_new(D->T);
  gen_mask(D->M, mrnd, D->Ac_b, D->Ac_j);
  gen_mask(D->NY, mrnd, D->Yc_b, D->Yc_j);
  gen_mask(D->T, mrnd, D->Tc_b, D->Tc_j);
  bs(D->Tc_b, 0); bs(D->Tc_j, 0); bs(D->Tc_b, D->T - 1); bs(D->Tc_j, D->T - 1);
  D->tb = (int *)zalloc(D->T + 1, sizeof(int));
  for (int t = 0, b = 0; t < D->T; t++) { D->tb[t] = b; if (bg(D->Tc_b, t)) b++; }
  D->E = (dbin *)zalloc(D->nbins, sizeof(dbin));
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    E->nA = maximum(0, p->A_n_rows + (nb % 3) - 1);
    E->Ar_b = mask_new(E->nA); E->Ar_j = mask_new(E->nA); E->Ar_rtn = mask_new(E->nA); E->Ar_rmv = mask_new(E->nA);
    gen_mask(E->nA, mrnd, E->Ar_b, E->Ar_j);
    gen_bits(E->nA, D->M, &E->An, &E->At);
    E->nZ = maximum(0, p->Z_n_rows + (nb % 5) - 2);
    E->Zr_b = mask_new(E->nZ); E->Zr_j = mask_new(E->nZ);
    gen_mask(E->nZ, mrnd, E->Zr_b, E->Zr_j);
    gen_bits(E->nZ, D->M, &E->Zn, &E->Zt);
    gen_bits(E->nA, D->NY, &E->Yn, &E->Yt);
    gen_bits(E->nZ, D->NY, &E->Wn, &E->Wt);
    /* covariate column 0 is the intercept: set in every row (dcc.c:601-602, 608-609) */
    gen_bits(E->nA, D->T, &E->Tn, &E->Tt);
    for (int r = 0; r < E->nA; r++) { bs(E->Tn.b + (size_t)r * E->Tn.rb, 0); bs(E->Tt.b, r); }
    gen_bits(E->nZ, D->T, &E->Sn, &E->St);
    for (int r = 0; r < E->nZ; r++) { bs(E->Sn.b + (size_t)r * E->Sn.rb, 0); bs(E->St.b, r); }
  }
  /* dcc_init_A_p: per 1920-column block p = maximum(0.1, minimum(0.9, rand01)), the
     macros drawing again in their chosen branches */
  D->Apc = lko_psize(D->M); D->A_p = (double *)zalloc(D->Apc, sizeof(double));
  for (int nc = 0; nc < D->Apc; nc++) D->A_p[nc] = maximum(0.1, minimum(0.9, rand01));
  if (p->sparse == 0) for (int nc = 0; nc < D->Apc; nc++) D->A_p[nc] = 0.5;
  D->A_ajdk = (double *)zalloc((size_t)LKO_AJDK * D->Apc, sizeof(double)); lko_calc_ajdk(D->A_p, D->Apc, D->A_ajdk);
  D->Ypc = lko_psize(D->NY); D->Y_p = (double *)zalloc(D->Ypc, sizeof(double));
  for (int nc = 0; nc < D->Ypc; nc++) D->Y_p[nc] = maximum(0.1, minimum(0.9, rand01));
  if (p->sparse == 0) for (int nc = 0; nc < D->Ypc; nc++) D->Y_p[nc] = 0.5;
  D->Y_ajdk = (double *)zalloc((size_t)LKO_AJDK * D->Ypc, sizeof(double)); lko_calc_ajdk(D->Y_p, D->Ypc, D->Y_ajdk);
  dcc_alloc(D);
}

/* dcc_X_nrows_total (dcc.c:1115-1131) and dcc_init_QX (dcc.c:1047-1090): global row
   indices over the bins, then every score and work array */
static void dcc_alloc(dcc_t *D)
{
  int mx_a = 0, mx_b = 0;
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    E->g_a = (int *)zalloc(E->nA + 1, sizeof(int)); E->g_b = (int *)zalloc(E->nA + 1, sizeof(int));
    for (int r = 0; r < E->nA; r++) { E->g_a[r] = mx_a++; if (bg(E->Ar_b, r)) E->g_b[r] = mx_b++; }
    D->A_nrows_total += E->nA; D->A_rpop_b_total += popn(E->Ar_b, E->nA); D->Z_rpop_b_total += popn(E->Zr_b, E->nZ);
    for (int q = 0; q < NQ; q++) E->QRn[q] = (double *)zalloc((size_t)E->nA * D->T * D->nbins, sizeof(double));
  }
  for (int q = 0; q < NQ; q++) D->QCn[q] = (double *)zalloc((size_t)qcols(D, q) * D->T * D->nbins * D->nbins, sizeof(double));
  D->F = (dpair *)zalloc((size_t)D->nbins * D->nbins, sizeof(dpair));
  for (int nb1 = 0; nb1 < D->nbins; nb1++) for (int nb2 = 0; nb2 < D->nbins; nb2++) {
    dpair *F = &D->F[nb1 + nb2 * D->nbins];
    for (int q = 0; q < NQ; q++) {
      F->ww[q] = (double *)zalloc((size_t)D->E[nb1].nA * D->T, sizeof(double));
      F->qr[q] = (double *)zalloc((size_t)D->E[nb1].nA * D->T, sizeof(double));
      F->qc[q] = (double *)zalloc((size_t)qcols(D, q) * D->T, sizeof(double));
    }
  }
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    E->lfA = (double *)zalloc((size_t)E->nA * LKO_AJDK, sizeof(double)); E->lfY = (double *)zalloc((size_t)E->nA * LKO_AJDK, sizeof(double));
    E->lfZ = (double *)zalloc((size_t)E->nZ * LKO_AJDK, sizeof(double)); E->lfW = (double *)zalloc((size_t)E->nZ * LKO_AJDK, sizeof(double));
    E->AtTn = (double *)zalloc((size_t)D->M * D->T, sizeof(double)); E->ZtSn = (double *)zalloc((size_t)D->M * D->T, sizeof(double));
    E->YtTn = (double *)zalloc((size_t)D->NY * D->T, sizeof(double)); E->WtSn = (double *)zalloc((size_t)D->NY * D->T, sizeof(double));
    E->a1_AtTn = (double *)zalloc(D->T, sizeof(double)); E->a1_ZtSn = (double *)zalloc(D->T, sizeof(double));
    E->a1_YtTn = (double *)zalloc(D->T, sizeof(double)); E->a1_WtSn = (double *)zalloc(D->T, sizeof(double));
    E->et_Tn = (double *)zalloc(D->T, sizeof(double)); E->et_Sn = (double *)zalloc(D->T, sizeof(double));
  }
  int ntot = D->M + D->A_nrows_total;
  D->trace = (double *)zalloc((size_t)ntot * 6, sizeof(double));
  D->xdrop_a = (int *)zalloc((size_t)ntot * 2, sizeof(int)); D->xdrop_b = (int *)zalloc((size_t)ntot * 2, sizeof(int));
  D->QC_sv = (double *)zalloc(D->M + 1, sizeof(double)); D->QR_sv = (double *)zalloc(D->A_nrows_total + 1, sizeof(double));
  D->QC_a = (int *)zalloc(D->M + 1, sizeof(int)); D->QC_b = (int *)zalloc(D->M + 1, sizeof(int)); D->QC_j = (int *)zalloc(D->M + 1, sizeof(int));
  int *q7[7]; for (int i = 0; i < 7; i++) q7[i] = (int *)zalloc(D->A_nrows_total + 1, sizeof(int));
  D->QR_a = q7[0]; D->QR_b = q7[1]; D->QR_j = q7[2]; D->QR_nb = q7[3]; D->QR_ga = q7[4]; D->QR_gb = q7[5]; free(q7[6]);
}

/* ------------------------------------------------------------------ */
/* dcc_load (dcc.c:1157-1215): the problem from the .in's per-bin files  */
/* ------------------------------------------------------------------ */
#define DL_MAXBINS 64
typedef struct { char *v[DL_MAXBINS]; } dlist;           /* per-bin paths, NULL: none */
typedef struct {
  int nbins, B, gamma_type, sparse, Ireq, A_ncols, Y_ncols, T_ncols, scramble_num;
  double gamma, kappa2;
  char xpre[4096];
  dlist A_n, A_t, A_rind, Z_n, Z_t, Z_rind, Y_n, Y_t, W_n, W_t, T_n, T_t, S_n, S_t;
  char *A_cind, *Y_cind, *T_cind, *A_p, *Y_p;
} dcfg;

static void dl_trim(char *s)
{
  size_t n = strlen(s);
  while (n && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\n' || s[n - 1] == '\r')) s[--n] = 0;
  size_t i = 0; while (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r') i++;
  if (i) memmove(s, s + i, strlen(s + i) + 1);
}
/* a name as updateglobals keeps it: GLOBAL_DIR_XPRE/name when the prefix was set before */
static char *dl_path(const dcfg *c, const char *v)
{
  if (!v[0]) return NULL;
  size_t n = strlen(c->xpre) + strlen(v) + 2;
  char *p = (char *)zalloc(n, 1);
  if (c->xpre[0]) snprintf(p, n, "%s/%s", c->xpre, v); else snprintf(p, n, "%s", v);
  return p;
}
static void dl_list(const dcfg *c, char *val, dlist *L)
{
  int nb = 0; char *q = val;
  while (nb < DL_MAXBINS) {
    char *cm = strchr(q, ',');
    if (cm) *cm = 0;
    char item[4096]; snprintf(item, sizeof item, "%s", q); dl_trim(item);
    L->v[nb++] = dl_path(c, item);
    if (!cm) break;
    q = cm + 1;
  }
}
/* the keys dexcluster_driver reads (updateglobals.c:5-186); others are accepted and ignored */
static int dcfg_parse(const char *text, dcfg *c)
{
  memset(c, 0, sizeof *c);
  c->B = 32; c->gamma = 0.125; c->sparse = 1;
  const char *p = text;
  while (*p) {
    const char *eq = strchr(p, '='); if (!eq) break;
    char name[256]; size_t nl = (size_t)(eq - p); if (nl > 255) nl = 255;
    memcpy(name, p, nl); name[nl] = 0; dl_trim(name);
    const char *v0 = eq + 1; const char *semi = strchr(v0, ';'); if (!semi) semi = v0 + strlen(v0);
    size_t vl = (size_t)(semi - v0);
    char *val = (char *)zalloc(vl + 1, 1); memcpy(val, v0, vl); dl_trim(val);
    p = (*semi) ? semi + 1 : semi;
    if (!strcmp(name, "END")) { free(val); break; }
    if (!strcmp(name, "GLOBAL_NBINS")) c->nbins = atoi(val);
    else if (!strcmp(name, "GLOBAL_B_MLT")) c->B = atoi(val);
    else if (!strcmp(name, "GLOBAL_gamma")) c->gamma = atof(val);
    else if (!strcmp(name, "GLOBAL_gamma_type")) c->gamma_type = atoi(val);
    else if (!strcmp(name, "GLOBAL_kappa_squared")) c->kappa2 = atof(val);
    else if (!strcmp(name, "GLOBAL_TEST_sparse")) c->sparse = atoi(val);
    else if (!strcmp(name, "GLOBAL_Ireq")) c->Ireq = atoi(val);
    else if (!strcmp(name, "GLOBAL_scramble_num")) c->scramble_num = atoi(val);
    else if (!strcmp(name, "GLOBAL_A_n_cols")) c->A_ncols = atoi(val);
    else if (!strcmp(name, "GLOBAL_Y_n_cols")) c->Y_ncols = atoi(val);
    else if (!strcmp(name, "GLOBAL_T_n_cols")) c->T_ncols = atoi(val);
    else if (!strcmp(name, "GLOBAL_DIR_XPRE")) snprintf(c->xpre, sizeof c->xpre, "%s", val);
    else if (!strcmp(name, "GLOBAL_A_n_name_")) dl_list(c, val, &c->A_n);
    else if (!strcmp(name, "GLOBAL_A_t_name_")) dl_list(c, val, &c->A_t);
    else if (!strcmp(name, "GLOBAL_A_n_rind_")) dl_list(c, val, &c->A_rind);
    else if (!strcmp(name, "GLOBAL_Z_n_name_")) dl_list(c, val, &c->Z_n);
    else if (!strcmp(name, "GLOBAL_Z_t_name_")) dl_list(c, val, &c->Z_t);
    else if (!strcmp(name, "GLOBAL_Z_n_rind_")) dl_list(c, val, &c->Z_rind);
    else if (!strcmp(name, "GLOBAL_Y_n_name_")) dl_list(c, val, &c->Y_n);
    else if (!strcmp(name, "GLOBAL_Y_t_name_")) dl_list(c, val, &c->Y_t);
    else if (!strcmp(name, "GLOBAL_W_n_name_")) dl_list(c, val, &c->W_n);
    else if (!strcmp(name, "GLOBAL_W_t_name_")) dl_list(c, val, &c->W_t);
    else if (!strcmp(name, "GLOBAL_T_n_name_")) dl_list(c, val, &c->T_n);
    else if (!strcmp(name, "GLOBAL_T_t_name_")) dl_list(c, val, &c->T_t);
    else if (!strcmp(name, "GLOBAL_S_n_name_")) dl_list(c, val, &c->S_n);
    else if (!strcmp(name, "GLOBAL_S_t_name_")) dl_list(c, val, &c->S_t);
    else if (!strcmp(name, "GLOBAL_A_n_cind")) c->A_cind = dl_path(c, val);
    else if (!strcmp(name, "GLOBAL_Y_n_cind")) c->Y_cind = dl_path(c, val);
    else if (!strcmp(name, "GLOBAL_T_n_cind")) c->T_cind = dl_path(c, val);
    else if (!strcmp(name, "GLOBAL_A_p_name")) c->A_p = dl_path(c, val);
    else if (!strcmp(name, "GLOBAL_Y_p_name")) c->Y_p = dl_path(c, val);
    else if (strncmp(name, "GLOBAL_", 7)) { fprintf(stderr, "lko_dcc_bins: unknown key %s\n", name); free(val); return -1; }
    free(val);
  }
  return 0;
}
static void dcfg_free(dcfg *c)
{
  dlist *L[] = { &c->A_n, &c->A_t, &c->A_rind, &c->Z_n, &c->Z_t, &c->Z_rind, &c->Y_n, &c->Y_t, &c->W_n, &c->W_t,
                 &c->T_n, &c->T_t, &c->S_n, &c->S_t };
  for (size_t i = 0; i < sizeof L / sizeof L[0]; i++) for (int nb = 0; nb < DL_MAXBINS; nb++) free(L[i]->v[nb]);
  free(c->A_cind); free(c->Y_cind); free(c->T_cind); free(c->A_p); free(c->Y_p);
}

/* .b16 (binary_read.c:5-55): a file holding an (fnr x fnc) matrix column by column gives
   fnc rows of fnr bits */
static int dl_read(const char *path, dm *out)
{
  FILE *fp = fopen(path, "rb");
  if (!fp) { fprintf(stderr, "lko_dcc_bins: cannot open %s\n", path); return -1; }
  int hdr[3];
  if (fread(hdr, sizeof(int), 3, fp) != 3) { fclose(fp); return -1; }
  int bitj = hdr[0] > 0 ? hdr[0] : 16, fnr = hdr[1], fnc = hdr[2];
  size_t wp = (size_t)(fnr + (bitj - (fnr % bitj)) % bitj) / 8;
  *out = dm_new(fnc, fnr);
  for (int r = 0; r < fnc; r++)
    if (wp && fread(out->b + (size_t)r * out->rb, 1, wp, fp) != wp) { fclose(fp); return -1; }
  fclose(fp);
  return 0;
}
/* an (n x 1) mask file, or all n set when none is named */
static int dl_mask(const char *path, int n, uint8_t *m)
{
  if (!path) { for (int i = 0; i < n; i++) bs(m, i); return 0; }
  dm t; if (dl_read(path, &t)) return -1;
  if (t.nr != 1 || t.nc != n) { fprintf(stderr, "lko_dcc_bins: mask %s has %d entries, expected %d\n", path, t.nc, n); free(t.b); return -1; }
  for (int i = 0; i < n; i++) if (bg(t.b, i)) bs(m, i);
  free(t.b);
  return 0;
}
/* X and its transpose from a pair of files, checked against the expected shape */
static int dl_pair(const char *pt, const char *pn, int nrows, int ncols, dm *Xn, dm *Xt)
{
  if (!pt || !pn) return -1;
  if (dl_read(pt, Xn) || dl_read(pn, Xt)) return -1;
  if (Xn->nr != nrows || Xn->nc != ncols || Xt->nr != ncols || Xt->nc != nrows) {
    fprintf(stderr, "lko_dcc_bins: %s / %s are %dx%d / %dx%d, expected %dx%d\n", pt, pn, Xn->nr, Xn->nc, Xt->nr, Xt->nc, nrows, ncols);
    return -1;
  }
  return 0;
}
/* the intercept covariate when no T/S file is named (dcc.c:374-384): one column, set in every row */
static void dl_intercept(int nrows, dm *Xn, dm *Xt)
{
  *Xn = dm_new(nrows, 1); *Xt = dm_new(1, nrows);
  for (int r = 0; r < nrows; r++) { bs(Xn->b + (size_t)r * Xn->rb, 0); bs(Xt->b, r); }
}
/* sizes from a .b16 header: the file's rows (fnr) and columns (fnc) */
static int dl_size(const char *path, int *fnr, int *fnc)
{
  FILE *fp = fopen(path, "rb"); if (!fp) return -1;
  int hdr[3]; int ok = fread(hdr, sizeof(int), 3, fp) == 3; fclose(fp);
  if (!ok) return -1;
  *fnr = hdr[1]; *fnc = hdr[2];
  return 0;
}

/* dcc_ajdk_load, dcc_single_load_M_An per bin, dcc_load_A_p (dcc.c:54-115, 251-419, 875-991) */
static int dcc_load(dcc_t *D, const dcfg *c)
{
  memset(D, 0, sizeof *D);
  D->p.nbins = c->nbins; D->p.B_MLT = c->B; D->p.gamma = c->gamma; D->p.gamma_type = c->gamma_type;
  D->p.kappa2 = c->kappa2; D->p.sparse = c->sparse; D->p.niter = 0;
  D->nbins = c->nbins; D->B = c->B; D->D_MLT = 1LL << c->B;
  int fnr = 0, fnc = 0;
  /* columns from bin 0's files: A (SNPs), Y (controls' SNPs), T (covariates) */
  if (!c->A_t.v[0] || dl_size(c->A_t.v[0], &fnr, &fnc)) return -1;
  D->M = fnr;
  D->NY = 0; if (c->Y_t.v[0]) { if (dl_size(c->Y_t.v[0], &fnr, &fnc)) return -1; D->NY = fnr; }
  D->T = 1; if (c->T_t.v[0]) { if (dl_size(c->T_t.v[0], &fnr, &fnc)) return -1; D->T = fnr; }
  /* every bin names the same kinds of file: Y (W) and T (S) in all bins or in none */
  for (int nb = 0; nb < D->nbins; nb++)
    if (!c->A_t.v[nb] || !c->A_n.v[nb] || (!c->Y_t.v[nb]) != (!c->Y_t.v[0]) || (!c->T_t.v[nb]) != (!c->T_t.v[0]) ||
        (!c->Z_t.v[nb]) != (!c->Z_n.v[nb])) {
      fprintf(stderr, "lko_dcc_bins: bin %d names a different set of files\n", nb); return -1;
    }
  if ((c->A_ncols && c->A_ncols != D->M) || (c->Y_ncols && c->Y_ncols != D->NY)) {
    fprintf(stderr, "lko_dcc_bins: column counts of the .in and the files differ\n"); return -1;
  }
  if (D->T > 8) return -1;
  D->Ac_b = mask_new(D->M); D->Ac_j = mask_new(D->M); D->Ac_rtn = mask_new(D->M); D->Ac_rmv = mask_new(D->M);
  D->Yc_b = mask_new(D->NY); D->Yc_j = mask_new(D->NY); D->Tc_b = mask_new(D->T); D->Tc_j = mask_new(D->T);
  if (dl_mask(c->A_cind, D->M, D->Ac_b) || dl_mask(c->Y_cind, D->NY, D->Yc_b)) return -1;
  /* covariates: the T_n_cind mask, else the first GLOBAL_T_n_cols of them (all when 0), as
     the one-bin loader takes them */
  if (c->T_cind) { if (dl_mask(c->T_cind, D->T, D->Tc_b)) return -1; }
  else { const int nt = c->T_ncols > 0 ? minimum(c->T_ncols, D->T) : D->T; for (int t = 0; t < nt; t++) bs(D->Tc_b, t); }
  memcpy(D->Ac_j, D->Ac_b, lko_bsize(D->M)); memcpy(D->Yc_j, D->Yc_b, lko_bsize(D->NY)); memcpy(D->Tc_j, D->Tc_b, lko_bsize(D->T));
  D->tb = (int *)zalloc(D->T + 1, sizeof(int));
  for (int t = 0, b = 0; t < D->T; t++) { D->tb[t] = b; if (bg(D->Tc_b, t)) b++; }
  D->E = (dbin *)zalloc(D->nbins, sizeof(dbin));
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    if (!c->A_t.v[nb] || dl_size(c->A_t.v[nb], &fnr, &fnc)) return -1;
    E->nA = fnc;
    if (dl_pair(c->A_t.v[nb], c->A_n.v[nb], E->nA, D->M, &E->An, &E->At)) return -1;
    E->nZ = 0;
    if (c->Z_t.v[nb]) { if (dl_size(c->Z_t.v[nb], &fnr, &fnc)) return -1; E->nZ = fnc; }
    if (E->nZ > 0) { if (dl_pair(c->Z_t.v[nb], c->Z_n.v[nb], E->nZ, D->M, &E->Zn, &E->Zt)) return -1; }
    else { E->Zn = dm_new(0, D->M); E->Zt = dm_new(D->M, 0); }
    if (D->NY > 0 && c->Y_t.v[nb]) { if (dl_pair(c->Y_t.v[nb], c->Y_n.v[nb], E->nA, D->NY, &E->Yn, &E->Yt)) return -1; }
    else { E->Yn = dm_new(E->nA, D->NY); E->Yt = dm_new(D->NY, E->nA); }
    if (D->NY > 0 && E->nZ > 0) { if (dl_pair(c->W_t.v[nb], c->W_n.v[nb], E->nZ, D->NY, &E->Wn, &E->Wt)) return -1; }
    else { E->Wn = dm_new(E->nZ, D->NY); E->Wt = dm_new(D->NY, E->nZ); }
    if (c->T_t.v[nb]) { if (dl_pair(c->T_t.v[nb], c->T_n.v[nb], E->nA, D->T, &E->Tn, &E->Tt)) return -1; }
    else if (D->T == 1) dl_intercept(E->nA, &E->Tn, &E->Tt);
    else return -1;
    if (E->nZ > 0 && c->S_t.v[nb]) { if (dl_pair(c->S_t.v[nb], c->S_n.v[nb], E->nZ, D->T, &E->Sn, &E->St)) return -1; }
    else if (D->T == 1 || E->nZ == 0) dl_intercept(E->nZ, &E->Sn, &E->St);
    else return -1;
    E->Ar_b = mask_new(E->nA); E->Ar_j = mask_new(E->nA); E->Ar_rtn = mask_new(E->nA); E->Ar_rmv = mask_new(E->nA);
    E->Zr_b = mask_new(E->nZ); E->Zr_j = mask_new(E->nZ);
    if (dl_mask(c->A_rind.v[nb], E->nA, E->Ar_b) || dl_mask(c->Z_rind.v[nb], E->nZ, E->Zr_b)) return -1;
    memcpy(E->Ar_j, E->Ar_b, lko_bsize(E->nA)); memcpy(E->Zr_j, E->Zr_b, lko_bsize(E->nZ));
  }
  /* dcc_load_A_p: per column the loaded rows' counts of A and Z (Y and W) over the bins,
     then per 1920-column block the frequency over the block's loaded columns */
  int arb = 0, zrb = 0;
  for (int nb = 0; nb < D->nbins; nb++) { arb += popn(D->E[nb].Ar_b, D->E[nb].nA); zrb += popn(D->E[nb].Zr_b, D->E[nb].nZ); }
  double *AZ = (double *)zalloc(D->M + 1, sizeof(double)), *YW = (double *)zalloc(D->NY + 1, sizeof(double));
  for (int nb = 0; nb < D->nbins; nb++) {
    const dbin *E = &D->E[nb];
    for (int k = 0; k < D->M; k++) {
      const double a = bg(D->Ac_b, k) ? (double)pop_m(row(&E->At, k), E->Ar_b, lko_bsize(E->nA)) : 0.0;
      const double z = (bg(D->Ac_b, k) && E->nZ > 0) ? (double)pop_m(row(&E->Zt, k), E->Zr_b, lko_bsize(E->nZ)) : 0.0;
      AZ[k] += a; AZ[k] += z;
    }
    for (int k = 0; k < D->NY; k++) {
      const double y = bg(D->Yc_b, k) ? (double)pop_m(row(&E->Yt, k), E->Ar_b, lko_bsize(E->nA)) : 0.0;
      const double w = (bg(D->Yc_b, k) && E->nZ > 0) ? (double)pop_m(row(&E->Wt, k), E->Zr_b, lko_bsize(E->nZ)) : 0.0;
      YW[k] += y; YW[k] += w;
    }
  }
  D->Apc = lko_psize(D->M); D->A_p = (double *)zalloc(D->Apc, sizeof(double));
  D->Ypc = lko_psize(D->NY); D->Y_p = (double *)zalloc(D->Ypc + 1, sizeof(double));
  const double *rs[2] = { AZ, YW }; const uint8_t *cm[2] = { D->Ac_b, D->Yc_b }; double *pp[2] = { D->A_p, D->Y_p };
  const int nx[2] = { D->M, D->NY }, npc[2] = { D->Apc, D->Ypc };
  for (int v = 0; v < 2; v++) for (int b = 0; b < npc[v]; b++) {
    pp[v][b] = 0.5;
    if (!c->sparse) continue;
    const int s0 = minimum(nx[v] - 1, maximum(0, b * LKO_POPB * 8)), s1 = minimum(nx[v] - 1, maximum(0, (b + 1) * LKO_POPB * 8 - 1));
    double tsum = 0, tnum = 0;
    for (int k = s0; k <= s1; k++) if (bg(cm[v], k)) { tsum += rs[v][k]; tnum += 1; }
    if (tnum == 0) pp[v][b] = 0.5;
    else { double q = (tsum) / (tnum * (arb + zrb)); pp[v][b] = q < 0.01 ? 0.01 : (q > 0.99 ? 0.99 : q); }
  }
  free(AZ); free(YW);
  /* GLOBAL_A_p_name / GLOBAL_Y_p_name (mda_read_d2_r8): the frequencies from a file instead */
  const char *pf[2] = { c->A_p, c->Y_p };
  for (int v = 0; v < 2; v++) {
    if (!pf[v]) continue;
    FILE *fp = fopen(pf[v], "rb"); int hdr[3];
    const int ok = fp && fread(hdr, sizeof(int), 3, fp) == 3 && hdr[1] == npc[v] &&
                   fread(pp[v], sizeof(double), (size_t)npc[v], fp) == (size_t)npc[v];
    if (fp) fclose(fp);
    if (!ok) { fprintf(stderr, "lko_dcc_bins: cannot read %s\n", pf[v]); return -1; }
  }
  D->A_ajdk = (double *)zalloc((size_t)LKO_AJDK * D->Apc, sizeof(double)); lko_calc_ajdk(D->A_p, D->Apc, D->A_ajdk);
  D->Y_ajdk = (double *)zalloc((size_t)LKO_AJDK * (D->Ypc + 1), sizeof(double)); lko_calc_ajdk(D->Y_p, D->Ypc, D->Y_ajdk);
  dcc_alloc(D);
  return 0;
}

/* dcc_M_mxset (dcc.c:1092-1113): active counts from the masks */
static void dcc_M_mxset(dcc_t *D)
{
  D->A_cpop_j = popn(D->Ac_j, D->M); D->Y_cpop_j = popn(D->Yc_j, D->NY); D->T_cpop_j = popn(D->Tc_j, D->T);
  D->A_rpop_j_total = 0; D->Z_rpop_j_total = 0;
  for (int nb = 0; nb < D->nbins; nb++) { D->A_rpop_j_total += popn(D->E[nb].Ar_j, D->E[nb].nA); D->Z_rpop_j_total += popn(D->E[nb].Zr_j, D->E[nb].nZ); }
}

/* get_An_ajdk_v (An_ajdk_v.c:272-363) over the active rows; zero elsewhere */
static void ajdk_rows(const dm *Xn, const uint8_t *rmask, const uint8_t *cmask, const double *ajdk, double *out)
{
  il r = il_mask(rmask, Xn->nr);
  memset(out, 0, sizeof(double) * (size_t)Xn->nr * LKO_AJDK);
  double *tmp = (double *)zalloc((size_t)r.n * LKO_AJDK + 1, sizeof(double));
  lko_an_ajdk_v(Xn->b, Xn->rb, r.ix, r.n, cmask, Xn->nc, ajdk, tmp);
  for (int i = 0; i < r.n; i++) memcpy(out + (size_t)r.ix[i] * LKO_AJDK, tmp + (size_t)i * LKO_AJDK, sizeof(double) * LKO_AJDK);
  free(tmp); free(r.ix);
}
/* dcc_An_ajdk (dcc_lf.c:7-51) */
static void st_An_ajdk(dcc_t *D)
{
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    ajdk_rows(&E->An, E->Ar_j, D->Ac_j, D->A_ajdk, E->lfA);
    ajdk_rows(&E->Yn, E->Ar_j, D->Yc_j, D->Y_ajdk, E->lfY);
    ajdk_rows(&E->Zn, E->Zr_j, D->Ac_j, D->A_ajdk, E->lfZ);
    ajdk_rows(&E->Wn, E->Zr_j, D->Yc_j, D->Y_ajdk, E->lfW);
  }
}

/* plain +-1 product AnZt_vv(Xt, Ct) (AnZt_vv.c:62-95 without A_ajdk) over the patients
   of rmask: out[k*T + t] for active columns k (kmask) and covariates t */
static void plain_to(const dcc_t *D, const dm *Xt, const uint8_t *kmask, const dm *Ct, const uint8_t *rmask, double *out)
{
  il kl = il_mask(kmask, Xt->nr), tl = il_mask(D->Tc_j, D->T);
  memset(out, 0, sizeof(double) * (size_t)Xt->nr * D->T);
  double *tmp = (double *)zalloc((size_t)kl.n * tl.n + 1, sizeof(double));
  lko_anzt_vv(Xt->b, Xt->rb, kl.ix, kl.n, Ct->b, Ct->rb, tl.ix, tl.n, rmask, Xt->nc, NULL, NULL, NULL, tmp);
  for (int i = 0; i < kl.n; i++) for (int j = 0; j < tl.n; j++) out[(size_t)kl.ix[i] * D->T + tl.ix[j]] = tmp[(size_t)i * tl.n + j];
  free(tmp); free(kl.ix); free(tl.ix);
}
/* dcc_lf_ZtSn (dcc_lf.c:53-100) */
static void st_ZtSn(dcc_t *D)
{
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    plain_to(D, &E->At, D->Ac_j, &E->Tt, E->Ar_j, E->AtTn);
    plain_to(D, &E->Yt, D->Yc_j, &E->Tt, E->Ar_j, E->YtTn);
    plain_to(D, &E->Zt, D->Ac_j, &E->St, E->Zr_j, E->ZtSn);
    plain_to(D, &E->Wt, D->Yc_j, &E->St, E->Zr_j, E->WtSn);
  }
}

/* -cpop_j + 2 popcount of covariate row t over the active patients (dtmp_etTn / lf_et_Tn) */
static double et_of(const dm *Ct, int t, const uint8_t *rmask, int nrows)
{
  return (double)(-(long long)popn(rmask, nrows) + 2 * pop_m(row(Ct, t), rmask, Ct->rb));
}

/* get_D_AtTn_ZtSn_vv / get_D_AtTn_AtTn_vv (D_AtTn_ZtSn.c:5-310) for one bin pair:
   out[k*T + tb[t]]; same!=0 adds the repeated-index correction of the _AtTn_AtTn form */
static void d_term(const dcc_t *D, const dm *Xt1, const uint8_t *r1, int n1, const double *L1,
                   const dm *Xt2, const dm *Ct2, const uint8_t *r2, int n2, const double *L2,
                   const dm *Ct1, const uint8_t *kmask, int ncols, const double *ajdk, int pc, int same, double *out)
{
  const double *aA = ajdk + 1 * (size_t)pc, *DA = ajdk + 5 * (size_t)pc;
  memset(out, 0, sizeof(double) * (size_t)ncols * D->T);
  il kl = il_mask(kmask, ncols), tl = il_mask(D->Tc_j, D->T);
  int cpop1 = popn(r1, n1);
  for (int t = 0; t < tl.n; t++) {
    int tt = tl.ix[t];
    double etS = et_of(Ct2, tt, r2, n2), etT = et_of(Ct1, tt, r1, n1);
    for (int k = 0; k < kl.n; k++) {
      int kk = kl.ix[k]; size_t o = (size_t)kk * D->T + tt;
      double a = aA[kk / LKO_POP], Dk = DA[kk / LKO_POP];
      double v = (L1[o] - etT * a) * Dk * (L2[o] - a * etS);
      if (same) {
        double etA = (double)(-(long long)cpop1 + 2 * pop_m(row(Xt1, kk), r1, Xt1->rb));
        v -= Dk * (cpop1 - 2 * etA * a + cpop1 * a * a);
      }
      out[(size_t)kk * D->T + D->tb[tt]] = v;
    }
  }
  (void)Xt2;
  free(kl.ix); free(tl.ix);
}

/* dcc_lf_D_AtTn_ZtSn_vv (dcc_lf.c:102-163) and the QC half of the half-loop */
static void st_D(dcc_t *D)
{
  for (int nb1 = 0; nb1 < D->nbins; nb1++) for (int nb2 = 0; nb2 < D->nbins; nb2++) {
    dbin *E1 = &D->E[nb1], *E2 = &D->E[nb2]; dpair *F = &D->F[nb1 + nb2 * D->nbins];
    for (int q = 0; q < NQ; q++) memset(F->qc[q], 0, sizeof(double) * (size_t)qcols(D, q) * D->T);
    if (on(D, Q_YW, nb1, nb2)) d_term(D, &E1->Yt, E1->Ar_j, E1->nA, E1->YtTn, &E2->Wt, &E2->St, E2->Zr_j, E2->nZ, E2->WtSn, &E1->Tt, D->Yc_j, D->NY, D->Y_ajdk, D->Ypc, 0, F->qc[Q_YW]);
    if (on(D, Q_YY, nb1, nb2)) d_term(D, &E1->Yt, E1->Ar_j, E1->nA, E1->YtTn, &E2->Yt, &E2->Tt, E2->Ar_j, E2->nA, E2->YtTn, &E1->Tt, D->Yc_j, D->NY, D->Y_ajdk, D->Ypc, nb1 == nb2, F->qc[Q_YY]);
    if (on(D, Q_AZ, nb1, nb2)) d_term(D, &E1->At, E1->Ar_j, E1->nA, E1->AtTn, &E2->Zt, &E2->St, E2->Zr_j, E2->nZ, E2->ZtSn, &E1->Tt, D->Ac_j, D->M, D->A_ajdk, D->Apc, 0, F->qc[Q_AZ]);
    if (on(D, Q_AA, nb1, nb2)) d_term(D, &E1->At, E1->Ar_j, E1->nA, E1->AtTn, &E2->At, &E2->Tt, E2->Ar_j, E2->nA, E2->AtTn, &E1->Tt, D->Ac_j, D->M, D->A_ajdk, D->Apc, nb1 == nb2, F->qc[Q_AA]);
  }
}

/* sum over the active patients of +-1 covariate bit times lf_*_ajdk[., AJDK_1_1] (TAnZtS.c:159-219) */
static void a1d1(const dcc_t *D, const dm *Ct, const uint8_t *rmask, int nrows, const double *lf, double *out)
{
  memset(out, 0, sizeof(double) * D->T);
  il tl = il_mask(D->Tc_j, D->T), rl = il_mask(rmask, nrows);
  for (int t = 0; t < tl.n; t++) {
    const uint8_t *Cr = row(Ct, tl.ix[t]);
    for (int i = 0; i < rl.n; i++) out[tl.ix[t]] += bpm(Cr, rl.ix[i]) * lf[(size_t)rl.ix[i] * LKO_AJDK + 6];
  }
  free(tl.ix); free(rl.ix);
}
/* An_a0d1_ZtSn_excerpt (An_ZtSn_ww.c:243-265) then xcalc (TAnZtS.c:266-288) */
static void quantise(const dcc_t *D, const double *lf, const uint8_t *kmask, int ncols, const double *ajdk, int pc, lko_planes *P)
{
  il kl = il_mask(kmask, ncols), tl = il_mask(D->Tc_j, D->T);
  double *L = (double *)zalloc((size_t)kl.n * tl.n + 1, sizeof(double));
  for (int k = 0; k < kl.n; k++) for (int t = 0; t < tl.n; t++)
    L[(size_t)k * tl.n + t] = ajdk[kl.ix[k] / LKO_POP + 5 * (size_t)pc] * lf[(size_t)kl.ix[k] * D->T + tl.ix[t]];
  lko_planes_free(P); memset(P, 0, sizeof *P);
  lko_xcalc(L, kl.ix, kl.n, ncols, tl.n, D->B, D->D_MLT, P, NULL);
  free(L); free(kl.ix); free(tl.ix);
}
/* get_An_ZtSn_ww (An_ZtSn_ww.c:44-70) of the rows of Xn against quantised planes: out[r*T + t] */
static void ztsn_ww(const dcc_t *D, const dm *Xn, const uint8_t *rmask, const uint8_t *cmask, const lko_planes *P, double *out)
{
  il rl = il_mask(rmask, Xn->nr), tl = il_mask(D->Tc_j, D->T);
  memset(out, 0, sizeof(double) * (size_t)Xn->nr * D->T);
  for (int i = 0; i < rl.n; i++) {
    const uint8_t *Ar = row(Xn, rl.ix[i]);
    long p0 = 0;
    for (size_t q = 0; q < P->rb; q++) p0 += __builtin_popcount(Ar[q] & P->row0[q] & cmask[q]) - __builtin_popcount((uint8_t)~Ar[q] & P->row0[q] & cmask[q]);
    double base = (double)p0 * P->min_d * P->mlt_d;
    for (int s = 0; s < tl.n; s++) {
      double tmp = 0; long long n2 = 1;
      for (int nb = 0; nb < P->B; nb++) {
        const uint8_t *pl = P->planes + ((size_t)s * P->B + nb) * P->rb;
        long l = 0;
        for (size_t q = 0; q < P->rb; q++) l += __builtin_popcount(Ar[q] & pl[q] & cmask[q]) - __builtin_popcount((uint8_t)~Ar[q] & pl[q] & cmask[q]);
        tmp += n2 * (double)l; n2 *= 2;
      }
      out[(size_t)rl.ix[i] * D->T + tl.ix[s]] = (double)(base + tmp) / P->mlt_d;
    }
  }
  free(rl.ix); free(tl.ix);
}

/* dcc_lf_TAnZtS_ww stages 0-2 (TAnZtS.c:5-540) and the QR half of the half-loop */
static void st_TAnZtS(dcc_t *D)
{
  const int T = D->T;
  il tl = il_mask(D->Tc_j, T);
  for (int nb = 0; nb < D->nbins; nb++) {      /* stage 0 */
    dbin *E = &D->E[nb];
    E->A_a0d1 = mask_lf(D->Ac_j, D->Apc, D->A_ajdk + 5 * (size_t)D->Apc);
    E->A_a2d1 = mask_lf(D->Ac_j, D->Apc, D->A_ajdk + 7 * (size_t)D->Apc);
    E->Y_a0d1 = mask_lf(D->Yc_j, D->Ypc, D->Y_ajdk + 5 * (size_t)D->Ypc);
    E->Y_a2d1 = mask_lf(D->Yc_j, D->Ypc, D->Y_ajdk + 7 * (size_t)D->Ypc);
    a1d1(D, &E->Tt, E->Ar_j, E->nA, E->lfA, E->a1_AtTn);
    a1d1(D, &E->St, E->Zr_j, E->nZ, E->lfZ, E->a1_ZtSn);
    a1d1(D, &E->Tt, E->Ar_j, E->nA, E->lfY, E->a1_YtTn);
    a1d1(D, &E->St, E->Zr_j, E->nZ, E->lfW, E->a1_WtSn);
    memset(E->et_Tn, 0, sizeof(double) * T); memset(E->et_Sn, 0, sizeof(double) * T);
    for (int t = 0; t < tl.n; t++) { E->et_Tn[tl.ix[t]] = et_of(&E->Tt, tl.ix[t], E->Ar_j, E->nA); E->et_Sn[tl.ix[t]] = et_of(&E->St, tl.ix[t], E->Zr_j, E->nZ); }
  }
  for (int nb = 0; nb < D->nbins; nb++) {      /* stage 1: quantise per bin */
    dbin *E = &D->E[nb];
    quantise(D, E->WtSn, D->Yc_j, D->NY, D->Y_ajdk, D->Ypc, &E->P_WtSn);
    quantise(D, E->ZtSn, D->Ac_j, D->M, D->A_ajdk, D->Apc, &E->P_ZtSn);
    quantise(D, E->YtTn, D->Yc_j, D->NY, D->Y_ajdk, D->Ypc, &E->P_YtTn);
    quantise(D, E->AtTn, D->Ac_j, D->M, D->A_ajdk, D->Apc, &E->P_AtTn);
  }
  for (int nb1 = 0; nb1 < D->nbins; nb1++) for (int nb2 = 0; nb2 < D->nbins; nb2++) {
    dbin *E1 = &D->E[nb1], *E2 = &D->E[nb2]; dpair *F = &D->F[nb1 + nb2 * D->nbins];
    int nA = E1->nA;
    double *x[NQ];
    for (int q = 0; q < NQ; q++) x[q] = (double *)zalloc((size_t)nA * T + 1, sizeof(double));
    ztsn_ww(D, &E1->Yn, E1->Ar_j, D->Yc_j, &E2->P_WtSn, x[Q_YW]);
    ztsn_ww(D, &E1->An, E1->Ar_j, D->Ac_j, &E2->P_ZtSn, x[Q_AZ]);
    ztsn_ww(D, &E1->Yn, E1->Ar_j, D->Yc_j, &E2->P_YtTn, x[Q_YY]);
    ztsn_ww(D, &E1->An, E1->Ar_j, D->Ac_j, &E2->P_AtTn, x[Q_AA]);
    /* stage 2 (get_dcc_TAnZtS_ww_stage_2) */
    for (int q = 0; q < NQ; q++) memset(F->ww[q], 0, sizeof(double) * (size_t)nA * T);
    il rl = il_mask(E1->Ar_j, nA);
    for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t];
      for (int i = 0; i < rl.n; i++) {
        int r = rl.ix[i]; size_t o = (size_t)r * T + tt; double x6 = E1->lfA[(size_t)r * LKO_AJDK + 6];
        double l = x[Q_AZ][o] - E2->a1_ZtSn[tt] - x6 * E2->et_Sn[tt] + E1->A_a2d1 * E2->et_Sn[tt];
        F->ww[Q_AZ][o] += +l;
      }
      for (int i = 0; i < rl.n; i++) {
        int r = rl.ix[i]; size_t o = (size_t)r * T + tt; double x6 = E1->lfA[(size_t)r * LKO_AJDK + 6];
        double l = x[Q_AA][o] - E2->a1_AtTn[tt] - x6 * E2->et_Tn[tt] + E1->A_a2d1 * E2->et_Tn[tt];
        F->ww[Q_AA][o] += +l;
      }
    }
    if (nb1 == nb2) for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t]; const uint8_t *Tr = row(&E1->Tt, tt);
      for (int i = 0; i < rl.n; i++) {
        int r = rl.ix[i]; size_t o = (size_t)r * T + tt;
        double l = bpm(Tr, r) * (E1->A_a0d1 - 2 * E1->lfA[(size_t)r * LKO_AJDK + 6] + E1->A_a2d1);
        F->ww[Q_AA][o] += -l;
      }
    }
    for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t];
      for (int i = 0; i < rl.n; i++) {
        int r = rl.ix[i]; size_t o = (size_t)r * T + tt; double y6 = E1->lfY[(size_t)r * LKO_AJDK + 6];
        double l = x[Q_YW][o] - E2->a1_WtSn[tt] - y6 * E2->et_Sn[tt] + E1->Y_a2d1 * E2->et_Sn[tt];
        F->ww[Q_YW][o] += +l;
      }
      for (int i = 0; i < rl.n; i++) {
        int r = rl.ix[i]; size_t o = (size_t)r * T + tt; double y6 = E1->lfY[(size_t)r * LKO_AJDK + 6];
        double l = x[Q_YY][o] - E2->a1_YtTn[tt] - y6 * E2->et_Tn[tt] + E1->Y_a2d1 * E2->et_Tn[tt];
        F->ww[Q_YY][o] += +l;
      }
    }
    if (nb1 == nb2) for (int t = 0; t < tl.n; t++) {
      int tt = tl.ix[t]; const uint8_t *Tr = row(&E1->Tt, tt);
      for (int i = 0; i < rl.n; i++) {
        int r = rl.ix[i]; size_t o = (size_t)r * T + tt;
        double l = bpm(Tr, r) * (E1->Y_a0d1 - 2 * E1->lfY[(size_t)r * LKO_AJDK + 6] + E1->Y_a2d1);
        F->ww[Q_YY][o] += -l;
      }
    }
    /* get_dcc_halfloop: the covariate index becomes the loaded one (terms without their
       rows or columns stay zero) */
    for (int q = 0; q < NQ; q++) {
      memset(F->qr[q], 0, sizeof(double) * (size_t)nA * T);
      if (!on(D, q, nb1, nb2)) continue;
      for (int t = 0; t < tl.n; t++) for (int i = 0; i < rl.n; i++)
        F->qr[q][(size_t)rl.ix[i] * T + D->tb[tl.ix[t]]] = F->ww[q][(size_t)rl.ix[i] * T + tl.ix[t]];
    }
    free(rl.ix);
    for (int q = 0; q < NQ; q++) free(x[q]);
  }
  free(tl.ix);
}

/* dcc_sumscores_nrm with GLOBAL_Ireq <= 0 (dcc_sumscores.c:569-699) */
static void st_nrm(dcc_t *D)
{
  const int T = D->T, nbins = D->nbins;
  long long At = D->A_rpop_j_total, Zt = D->Z_rpop_j_total;
  for (int nb1 = 0; nb1 < nbins; nb1++) for (int nb2 = 0; nb2 < nbins; nb2++) {
    int nbx = nb1 + nb2 * nbins; dbin *E = &D->E[nb1]; dpair *F = &D->F[nbx];
    long long lldr[NQ] = { Zt * D->Y_cpop_j, Zt * D->A_cpop_j, (At - 1) * D->Y_cpop_j, (At - 1) * D->A_cpop_j };
    long long lldc[NQ] = { At * Zt, At * Zt, At * (At - 1), At * (At - 1) };
    for (int q = 0; q < NQ; q++) {   /* QR rows, in the order TYnWtS, TYnYtT, TAnZtS, TAnAtT */
      int qq = q == 1 ? Q_YY : q == 2 ? Q_AZ : q;
      if (!on(D, qq, nb1, nb2)) continue;
      double *dst = E->QRn[qq] + (size_t)nb2 * E->nA * T;
      memset(dst, 0, sizeof(double) * (size_t)E->nA * T);
      double d = (double)lldr[qq]; if (fabs(d) < 1e-12) d = 1;
      for (int c = 0; c < T; c++) if (bg(D->Tc_j, c)) for (int r = 0; r < E->nA; r++) if (bg(E->Ar_j, r))
        dst[(size_t)c * E->nA + r] += F->qr[qq][(size_t)r * T + c] / d;
    }
    for (int q = 0; q < NQ; q++) {
      int qq = q == 1 ? Q_YY : q == 2 ? Q_AZ : q;
      if (!on(D, qq, nb1, nb2)) continue;
      int X = qcols(D, qq); const uint8_t *km = X == D->NY && (qq == Q_YW || qq == Q_YY) ? D->Yc_j : D->Ac_j;
      double *dst = D->QCn[qq] + (size_t)nbx * X * T;
      memset(dst, 0, sizeof(double) * (size_t)X * T);
      double d = (double)lldc[qq]; if (fabs(d) < 1e-12) d = 1;
      for (int c = 0; c < T; c++) if (bg(D->Tc_j, c)) for (int k = 0; k < X; k++) if (bg(km, k))
        dst[(size_t)c * X + k] += F->qc[qq][(size_t)k * T + c] / d;
    }
  }
  /* dra_sumx (ra_stats.c:106-109): the bins' terms added into bin 0, in bin order */
  for (int c = 0; c < T; c++) {
    if (!bg(D->Tc_j, c)) continue;
    for (int nb1 = 0; nb1 < nbins; nb1++) {
      dbin *E = &D->E[nb1]; size_t stride = (size_t)E->nA * T;
      const int order[NQ] = { Q_YW, Q_YY, Q_AZ, Q_AA };
      const int use[NQ] = { E->nZ > 0 && D->NY > 0, E->nZ > 0 && D->M > 0, D->NY > 0, D->M > 0 };   /* by term */
      if (E->nA <= 0) continue;
      for (int q = 0; q < NQ; q++) if (use[order[q]]) for (int r = 0; r < E->nA; r++) if (bg(E->Ar_b, r) && bg(E->Ar_j, r)) {
        double *x = E->QRn[order[q]] + (size_t)c * E->nA + r;
        for (int k = 1; k < nbins; k++) x[0] += x[k * stride];
      }
    }
    const int order[NQ] = { Q_YW, Q_YY, Q_AZ, Q_AA };
    const int ar = D->A_rpop_b_total > 0, zr = D->Z_rpop_b_total > 0;
    const int usec[NQ] = { ar && D->NY > 0 && zr, ar && D->M > 0 && zr, ar && D->NY > 0 && zr, ar && D->M > 0 };   /* by term */
    for (int q = 0; q < NQ; q++) {
      int qq = order[q], X = qcols(D, qq);
      if (!usec[qq]) continue;
      const uint8_t *mb = (qq == Q_YW || qq == Q_YY) ? D->Yc_b : D->Ac_b, *mj = (qq == Q_YW || qq == Q_YY) ? D->Yc_j : D->Ac_j;
      size_t stride = (size_t)X * T;
      for (int k = 0; k < X; k++) if (bg(mb, k) && bg(mj, k)) {
        double *x = D->QCn[qq] + (size_t)c * X + k;
        for (int j = 1; j < nbins * nbins; j++) x[0] += x[j * stride];
      }
    }
  }
}

/* dra_mds_nrm_s___m_m / dra_mds_pow_s___m_m (ra_stats.c:77-104) */
static void mds_nrm(double *x, int nrows, int ncols, double den, const uint8_t *rm, const uint8_t *cm)
{
  if (!(den > 0)) return;
  for (int r = 0; r < nrows; r++) if (bg(rm, r)) for (int c = 1; c < ncols; c++) if (bg(cm, c)) x[r] -= x[r + (size_t)c * nrows] / den;
}
static void mds_pow(double *x, int nrows, int ncols, double den, const uint8_t *rm, const uint8_t *cm)
{
  if (!(den > 0)) return;
  for (int r = 0; r < nrows; r++) {
    if (!bg(rm, r) || !(x[r] >= 0)) continue;
    x[r] *= x[r];
    for (int c = 1; c < ncols; c++) if (bg(cm, c)) x[r] -= x[r + (size_t)c * nrows] * x[r + (size_t)c * nrows] / den;
    x[r] = x[r] >= 0 ? sqrt(x[r]) : -sqrt(-x[r]);
  }
}
/* dcc_sumscores_ifT (dcc_sumscores.c:418-476): with GLOBAL_Ireq <= 0 only nb2 = 0 */
static void st_ifT(dcc_t *D)
{
  static const double scale[6] = { 1.00, 0.34, 0.19, 0.13, 0.10, 0.09 };   /* GLOBAL_kappa_squared_loop_scale_factor_ */
  int n_mds = maximum(0, D->T_cpop_j - 1); if (n_mds > 6) n_mds = 6;
  if (n_mds <= 0) return;
  double f = scale[n_mds - 1];
  f = D->p.kappa2;   /* the stray ';' after `if (GLOBAL_kappa_squared>0)` makes this unconditional */
  double den = n_mds * f;
  for (int nb1 = 0; nb1 < D->nbins; nb1++) {
    int nb2 = 0, nbx = nb1 + nb2 * D->nbins; dbin *E = &D->E[nb1];
    const int order[NQ] = { Q_YW, Q_YY, Q_AZ, Q_AA };
    const int use[NQ] = { E->nZ > 0 && D->NY > 0, E->nZ > 0 && D->M > 0, D->NY > 0, D->M > 0 };
    const int arb = D->A_rpop_b_total > 0, zrb = D->Z_rpop_b_total > 0;
    const int usec[NQ] = { arb && D->NY > 0 && zrb, arb && D->M > 0 && zrb, arb && D->NY > 0, arb && D->M > 0 };
    if (E->nA > 0) for (int q = 0; q < NQ; q++) if (use[order[q]]) mds_pow(E->QRn[order[q]] + (size_t)nb2 * E->nA * D->T, E->nA, D->T, den, E->Ar_j, D->Tc_j);
    for (int q = 0; q < NQ; q++) {
      int qq = order[q], X = qcols(D, qq);
      if (!usec[qq]) continue;
      mds_nrm(D->QCn[qq] + (size_t)nbx * X * D->T, X, D->T, den, (qq == Q_YW || qq == Q_YY) ? D->Yc_j : D->Ac_j, D->Tc_j);
    }
  }
}
/* dcc_sumscores_cmb (dcc_sumscores.c:243-278): first covariate column only */
static void st_cmb(dcc_t *D)
{
  if (D->M > 0 && D->A_rpop_b_total > 0 && D->Z_rpop_b_total > 0) for (int nbx = 0; nbx < D->nbins * D->nbins; nbx++) {
    double *a = D->QCn[Q_AA] + (size_t)nbx * D->M * D->T, *z = D->QCn[Q_AZ] + (size_t)nbx * D->M * D->T;
    for (int k = 0; k < D->M; k++) if (bg(D->Ac_j, k)) a[k] += -1.0 * z[k];
  }
  for (int nb1 = 0; nb1 < D->nbins; nb1++) {
    dbin *E = &D->E[nb1]; size_t blk = (size_t)E->nA * D->T;
    if (E->nA <= 0 || D->M <= 0) continue;
    if (E->nZ > 0 && D->NY > 0) for (int nb2 = 0; nb2 < D->nbins; nb2++) for (int r = 0; r < E->nA; r++) if (bg(E->Ar_j, r)) E->QRn[Q_AA][nb2 * blk + r] += +1.0 * E->QRn[Q_YW][nb2 * blk + r];
    if (E->nZ > 0) for (int nb2 = 0; nb2 < D->nbins; nb2++) for (int r = 0; r < E->nA; r++) if (bg(E->Ar_j, r)) E->QRn[Q_AA][nb2 * blk + r] += -1.0 * E->QRn[Q_AZ][nb2 * blk + r];
    if (D->NY > 0) for (int nb2 = 0; nb2 < D->nbins; nb2++) for (int r = 0; r < E->nA; r++) if (bg(E->Ar_j, r)) E->QRn[Q_AA][nb2 * blk + r] += -1.0 * E->QRn[Q_YY][nb2 * blk + r];
  }
}
/* dcc_sumscores_xij (dcc_sumscores.c:158-241) with Ireq = 0 ("empty low": Irow = Icol = 0) */
static void st_xij(dcc_t *D)
{
  D->Irem = 0; for (int nb = 0; nb < D->nbins; nb++) D->Irem += popn(D->E[nb].Ar_j, D->E[nb].nA) > 0 ? 1 : 0;
  if (D->M > 0 && D->A_rpop_b_total) {
    int na_b = 0, na_j = 0;
    for (int k = 0; k < D->M; k++) {
      if (!bg(D->Ac_b, k)) continue;
      if (bg(D->Ac_j, k)) { D->QC_sv[na_j] = D->QCn[Q_AA][k]; D->QC_a[na_j] = k; D->QC_b[na_j] = na_b; D->QC_j[na_j] = na_j; na_j++; }
      na_b++;
    }
    lko_quicksort_xij(D->QC_sv, D->QC_a, D->QC_b, D->QC_j, NULL, NULL, NULL, 0, D->A_cpop_j - 1);
  }
  if (D->A_rpop_b_total && D->M > 0) {
    D->A_nrows_total = 0; D->A_rpop_b_total = 0; D->A_rpop_j_total = 0;
    for (int nb = 0; nb < D->nbins; nb++) {
      dbin *E = &D->E[nb]; int ma_b = 0, ma_j = 0;
      for (int r = 0; r < E->nA; r++) {
        if (bg(E->Ar_b, r)) {
          if (bg(E->Ar_j, r)) {
            int i = D->A_rpop_j_total;
            D->QR_sv[i] = E->QRn[Q_AA][r]; D->QR_a[i] = r; D->QR_b[i] = ma_b; D->QR_j[i] = ma_j; D->QR_nb[i] = nb;
            D->QR_ga[i] = E->g_a[r]; D->QR_gb[i] = E->g_b[r];
            ma_j++; D->A_rpop_j_total++;
          }
          ma_b++; D->A_rpop_b_total++;
        }
        D->A_nrows_total++;
      }
    }
    lko_quicksort_xij(D->QR_sv, D->QR_a, D->QR_b, D->QR_j, D->QR_nb, D->QR_ga, D->QR_gb, 0, D->A_rpop_j_total - 1);
  }
}
/* dcc_sumscores_mxA / mxB / mxC (dcc_sumscores.c:5-156) */
static void st_mxA(dcc_t *D)
{
  int rdrop = 0, cdrop = 0; double lrij = D->A_rpop_j_total, lcij = D->A_cpop_j;
  lko_get_xdrop(D->A_rpop_j_total, D->A_cpop_j, D->p.gamma, D->p.gamma_type, &rdrop, &cdrop);
  size_t mb = lko_bsize(D->M);
  memcpy(D->Ac_rtn, D->Ac_j, mb); memset(D->Ac_rmv, 0, mb);
  for (int nc = 0; nc < (int)minimum(cdrop, lcij); nc++) { int k = D->QC_a[nc]; bs(D->Ac_rmv, k); D->Ac_rtn[k >> 3] &= (uint8_t)~(1u << (7 - (k & 7))); }
  for (int nb = 0; nb < D->nbins; nb++) { dbin *E = &D->E[nb]; memcpy(E->Ar_rtn, E->Ar_j, lko_bsize(E->nA)); memset(E->Ar_rmv, 0, lko_bsize(E->nA)); }
  for (int nr = 0; nr < (int)minimum(rdrop, lrij); nr++) {
    dbin *E = &D->E[D->QR_nb[nr]]; int r = D->QR_a[nr];
    bs(E->Ar_rmv, r); E->Ar_rtn[r >> 3] &= (uint8_t)~(1u << (7 - (r & 7)));
  }
  D->A_cpop_j = popn(D->Ac_j, D->M);
  /* mxC: the trace line, means over the sorted scores */
  double *tr = D->trace + (size_t)D->n_trace * 6, m = 0;
  tr[0] = D->n_trace; tr[1] = D->A_rpop_j_total; tr[2] = D->A_cpop_j;
  m = 0; for (int i = 0; i < D->A_rpop_j_total; i++) m += D->QR_sv[i]; m /= (double)D->A_rpop_j_total; if (!isfinite(m)) m = 1.0; tr[3] = m;
  m = 0; for (int i = 0; i < D->A_cpop_j; i++) m += D->QC_sv[i]; m /= (double)D->A_cpop_j; if (!isfinite(m)) m = 1.0; tr[4] = m;
  tr[5] = D->Irem; D->n_trace++;
  for (int nr = 0; nr < (int)minimum(rdrop, lrij); nr++) {
    D->xdrop_a[2 * D->n_xdrop] = D->QR_ga[nr]; D->xdrop_a[2 * D->n_xdrop + 1] = -1;
    D->xdrop_b[2 * D->n_xdrop] = D->QR_gb[nr]; D->xdrop_b[2 * D->n_xdrop + 1] = -1; D->n_xdrop++;
  }
  for (int nc = 0; nc < (int)minimum(cdrop, lcij); nc++) {
    D->xdrop_a[2 * D->n_xdrop] = -1; D->xdrop_a[2 * D->n_xdrop + 1] = D->QC_a[nc];
    D->xdrop_b[2 * D->n_xdrop] = -1; D->xdrop_b[2 * D->n_xdrop + 1] = D->QC_b[nc]; D->n_xdrop++;
  }
}
/* dcc_lrup_mxdup (dcc_lf.c:531-542) */
static void st_mxdup(dcc_t *D)
{
  memcpy(D->Ac_j, D->Ac_rtn, lko_bsize(D->M));
  for (int nb = 0; nb < D->nbins; nb++) memcpy(D->E[nb].Ar_j, D->E[nb].Ar_rtn, lko_bsize(D->E[nb].nA));
}

static void subscores(dcc_t *D) { st_An_ajdk(D); st_ZtSn(D); st_D(D); st_TAnZtS(D); }

/* dcc_sumscores_dmp (dcc_sumscores.c:90-118) */
static int write_outputs(const dcc_t *D, const char *dir)
{
  char p[4096]; FILE *f;
  snprintf(p, sizeof p, "%s/out_trace.txt", dir); if (!(f = fopen(p, "w"))) return -1;
  for (int i = 0; i < D->n_trace; i++) { const double *t = D->trace + (size_t)i * 6; fprintf(f, "%d %d %d %0.16f %0.16f %d\n", (int)t[0], (int)t[1], (int)t[2], t[3], t[4], (int)t[5]); }
  fclose(f);
  FILE *fa, *fb;
  snprintf(p, sizeof p, "%s/out_xdrop_a.txt", dir); if (!(fa = fopen(p, "w"))) return -1;
  snprintf(p, sizeof p, "%s/out_xdrop_b.txt", dir); if (!(fb = fopen(p, "w"))) { fclose(fa); return -1; }
  for (int i = 0; i < D->n_xdrop; i++) { fprintf(fa, "%d %d\n", D->xdrop_a[2 * i], D->xdrop_a[2 * i + 1]); fprintf(fb, "%d %d\n", D->xdrop_b[2 * i], D->xdrop_b[2 * i + 1]); }
  fclose(fa); fclose(fb);
  return 0;
}

static void dcc_free(dcc_t *D)
{
  for (int nb = 0; nb < D->nbins; nb++) {
    dbin *E = &D->E[nb];
    void *ps[] = { E->Ar_b, E->Ar_j, E->Ar_rtn, E->Ar_rmv, E->Zr_b, E->Zr_j, E->An.b, E->At.b, E->Zn.b, E->Zt.b, E->Yn.b, E->Yt.b, E->Wn.b, E->Wt.b,
                   E->Tn.b, E->Tt.b, E->Sn.b, E->St.b, E->lfA, E->lfY, E->lfZ, E->lfW, E->AtTn, E->YtTn, E->ZtSn, E->WtSn,
                   E->a1_AtTn, E->a1_ZtSn, E->a1_YtTn, E->a1_WtSn, E->et_Tn, E->et_Sn, E->g_a, E->g_b };
    for (size_t i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
    for (int q = 0; q < NQ; q++) free(E->QRn[q]);
    lko_planes_free(&E->P_WtSn); lko_planes_free(&E->P_ZtSn); lko_planes_free(&E->P_YtTn); lko_planes_free(&E->P_AtTn);
  }
  for (int x = 0; x < D->nbins * D->nbins; x++) for (int q = 0; q < NQ; q++) { free(D->F[x].ww[q]); free(D->F[x].qr[q]); free(D->F[x].qc[q]); }
  for (int q = 0; q < NQ; q++) free(D->QCn[q]);
  void *ps[] = { D->E, D->F, D->Ac_b, D->Ac_j, D->Ac_rtn, D->Ac_rmv, D->Yc_b, D->Yc_j, D->Tc_b, D->Tc_j, D->A_p, D->A_ajdk, D->Y_p, D->Y_ajdk, D->tb,
                 D->QC_sv, D->QR_sv, D->QC_a, D->QC_b, D->QC_j, D->QR_a, D->QR_b, D->QR_j, D->QR_nb, D->QR_ga, D->QR_gb, D->trace, D->xdrop_a, D->xdrop_b };
  for (size_t i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
}

int lko_dcc_sumscores_test(const lko_dcc_params *p, const char *out_dir, double *trace, int *n_trace, int *xdrop_a, int *xdrop_b, int *n_xdrop)
{
  if (!p || p->nbins < 1 || p->T_n_cols < 1 || p->A_n_cols < 1 || p->B_MLT < 1 || p->B_MLT > 32) return -1;
  dcc_t D;
  dcc_init(&D, p);
  dcc_M_mxset(&D);
  subscores(&D);
  int nl = 0;
  while ((p->niter <= 0 || nl < p->niter) && (D.A_rpop_j_total > 0 || D.A_cpop_j > 0)) {
    st_nrm(&D); st_ifT(&D); /* dcc_sumscores_srt: nothing with GLOBAL_Ireq <= 0 */ st_cmb(&D);
    st_xij(&D); st_mxA(&D);
    if (out_dir && write_outputs(&D, out_dir) != 0) { dcc_free(&D); return -2; }
    st_mxdup(&D); dcc_M_mxset(&D);
    subscores(&D);
    nl++;
  }
  if (n_trace) *n_trace = D.n_trace;
  if (n_xdrop) *n_xdrop = D.n_xdrop;
  if (trace) memcpy(trace, D.trace, sizeof(double) * 6 * D.n_trace);
  if (xdrop_a) memcpy(xdrop_a, D.xdrop_a, sizeof(int) * 2 * D.n_xdrop);
  if (xdrop_b) memcpy(xdrop_b, D.xdrop_b, sizeof(int) * 2 * D.n_xdrop);
  dcc_free(&D);
  return 0;
}

/* dexcluster_driver (dexcluster_driver.c:5-57) on the .in's files, with several bins and
   Y/W genetic controls: the loop to the end, the three output files in out_dir */
int lko_dcc_driver_bins(const char *in_text, const char *out_dir, int *n_trace, int *n_xdrop)
{
  if (!in_text || !out_dir) return -1;
  dcfg c;
  if (dcfg_parse(in_text, &c)) { dcfg_free(&c); return -1; }
  if (c.nbins < 1 || c.nbins > DL_MAXBINS || (c.Ireq > 0 && c.nbins > 1) || c.scramble_num > 0 || c.B < 1 || c.B > 32) {
    dcfg_free(&c); return -1;
  }
  dcc_t D;
  const int rc = dcc_load(&D, &c);
  dcfg_free(&c);
  if (rc) return -3;
  dcc_M_mxset(&D);
  subscores(&D);
  while (D.A_rpop_j_total > 0 || D.A_cpop_j > 0) {
    st_nrm(&D); st_ifT(&D); /* dcc_sumscores_srt: nothing with GLOBAL_Ireq <= 0 */ st_cmb(&D);
    st_xij(&D); st_mxA(&D);
    st_mxdup(&D); dcc_M_mxset(&D);
    subscores(&D);
  }
  const int wr = write_outputs(&D, out_dir);
  if (n_trace) *n_trace = D.n_trace;
  if (n_xdrop) *n_xdrop = D.n_xdrop;
  dcc_free(&D);
  return wr ? -2 : 0;
}
