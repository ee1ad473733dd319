/*
 * oracle.c -- CPU restatement of the dmrg_C superblock eigensolve path.
 *
 * TEST INFRASTRUCTURE ONLY (checker).  The product never loads this library; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math so that every product and sum is
 * evaluated in source order.  The reference itself is built with -Ofast
 * (lib/makefile:13); its Val bits are compared within a few ulp, Row/Col exactly.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_ZERO 1.0e-15 /* dmrg/DMRG_MAKE_ELEM_LL_LLLRRRRL.c:18, GET_HAM_LLLRRRRL.c:26 */

/* ------------------------------------------------------------ start vector */
static uint64_t splitmix64(uint64_t z) {
   z += 0x9E3779B97F4A7C15ULL;
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
   return z ^ (z >> 31);
}

/* Integers in [-5e7, 5e7), the distribution of rand()%100000000 - 50000000
 * (sml/LANCZOS_SLOW_SYM.c:82), drawn from a counter hash instead of rand(). */
void or_start_vector(int64_t n, uint64_t seed, double *out) {
   for (int64_t i = 0; i < n; i++) {
      uint64_t h = splitmix64(seed * 0xD1B54A32D192ED03ULL + (uint64_t)i);
      out[i] = (double)(int64_t)(h % 100000000ULL) - 50000000.0;
   }
}

/* -------------------------------------------------------------------- SpMV */
void or_spmv_sym(int32_t n, const int64_t *row, const int32_t *col, const double *val,
                 const double *x, double *y) {
   /* sml/MATRIX_VECTOR_PRODUCT_SYM.c:26-49 with p_threads = 1: one Temp array. */
   for (int32_t i = 0; i < n; i++) y[i] = 0.0;
   for (int32_t i = 0; i < n; i++) {
      int64_t last = row[i + 1] - 1;
      double t1 = val[last] * x[i];
      double t2 = x[i];
      for (int64_t j = row[i]; j < last; j++) {
         t1 = t1 + val[j] * x[col[j]];
         y[col[j]] = y[col[j]] + val[j] * t2;
      }
      y[i] = y[i] + t1;
   }
}

void or_spmv_full(int32_t n, const int64_t *row, const int32_t *col, const double *val,
                  const double *x, double *y) {
   /* sml/MATRIX_VECTOR_PRODUCT.c:13-19 */
   for (int32_t i = 0; i < n; i++) {
      double t = 0.0;
      for (int64_t j = row[i]; j < row[i + 1]; j++) t = t + val[j] * x[col[j]];
      y[i] = t;
   }
}

int64_t or_expand_sym(int32_t n, const int64_t *row, const int32_t *col, const double *val,
                      int64_t *frow, int32_t *fcol, double *fval) {
   int64_t *cnt = calloc((size_t)n + 1, sizeof(int64_t));
   for (int32_t i = 0; i < n; i++) {
      for (int64_t j = row[i]; j < row[i + 1]; j++) {
         cnt[i + 1]++;
         if (col[j] != i) cnt[col[j] + 1]++;
      }
   }
   for (int32_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
   int64_t nnz = cnt[n];
   if (frow) memcpy(frow, cnt, sizeof(int64_t) * ((size_t)n + 1));
   if (fcol && fval) {
      /* entries of row r: first the transposed ones (columns < r, ascending since
         source rows are visited in order), then row r's own lower-triangle entries */
      int64_t *pos = malloc(sizeof(int64_t) * (size_t)n);
      for (int32_t i = 0; i < n; i++) pos[i] = cnt[i];
      for (int32_t i = 0; i < n; i++) {
         for (int64_t j = row[i]; j < row[i + 1]; j++) {
            int32_t c = col[j];
            if (c == i) continue;
            /* element (i,c) with c<i goes to row c at column i */
            fcol[pos[c]] = i;
            fval[pos[c]] = val[j];
            pos[c]++;
         }
      }
      /* now row i's own lower entries: they have columns <= i, must precede the
         transposed ones (columns > i). Rebuild each row in order. */
      for (int32_t i = 0; i < n; i++) {
         int64_t own = row[i + 1] - row[i];
         int64_t s = cnt[i], trans = pos[i] - cnt[i];
         /* shift transposed block right by `own` */
         memmove(fcol + s + own, fcol + s, sizeof(int32_t) * (size_t)trans);
         memmove(fval + s + own, fval + s, sizeof(double) * (size_t)trans);
         memcpy(fcol + s, col + row[i], sizeof(int32_t) * (size_t)own);
         memcpy(fval + s, val + row[i], sizeof(double) * (size_t)own);
      }
      free(pos);
   }
   free(cnt);
   return nnz;
}

/* ---------------------------------------------------------------- assembly */
enum { K_LL = 0, K_LR, K_RR, K_RL, K_LLLR, K_LLRL, K_LLRR, K_LRRL, K_LRRR, K_RRRL, K_ZERO };
enum { S_LL = 0, S_LR = 1, S_RR = 2, S_RL = 3 };

/* slot each operator of a term acts on (include/dmrg.h:330-340 argument order) */
static const int op1_slot[11] = {S_LL, S_LR, S_RR, S_RL, S_LL, S_LL, S_LL, S_LR, S_LR, S_RR, -1};
static const int op2_slot[11] = {-1, -1, -1, -1, S_LR, S_RL, S_RR, S_RL, S_RR, S_RL, -1};

typedef struct {
   int32_t *acc_inv;   /* DMRG_A_BASIS.Inv_LLLRRRRL: row-local dedup map, -1 = absent */
   int32_t *e_col;     /* superblock column of each accumulated element */
   double *e_val;
   int32_t cap, num;
} or_acc;

static void acc_put(or_acc *a, int32_t inv, double v) {
   /* dmrg/DMRG_MAKE_ELEM_LL_LLLRRRRL.c:30-43: first insertion sets, later ones add */
   int32_t chk = a->acc_inv[inv];
   if (chk == -1) {
      if (a->num >= a->cap) {
         a->cap *= 2;
         a->e_col = realloc(a->e_col, sizeof(int32_t) * (size_t)a->cap);
         a->e_val = realloc(a->e_val, sizeof(double) * (size_t)a->cap);
      }
      a->e_col[a->num] = inv;
      a->e_val[a->num] = v;
      a->acc_inv[inv] = a->num;
      a->num++;
   } else {
      a->e_val[chk] = a->e_val[chk] + v;
   }
}

static int term_sign(const or_term *t, const int32_t *const *eles, const int q[4]) {
   /* sign rules of DMRG_MAKE_ELEM_{LLLR,LLRL,LLRR,LRRL,LRRR,RRRL}_LLLRRRRL.c:28-52 */
   if (t->sign_mode == 0) return 1;
   int p = 0;
   switch (t->kind) {
   case K_LLLR: p = eles[t->ele0][q[S_LL]]; break;
   case K_LLRL: p = eles[t->ele0][q[S_LL]] + eles[t->ele1][q[S_RR]] + eles[t->ele2][q[S_LR]]; break;
   case K_LLRR: p = eles[t->ele0][q[S_LL]] + eles[t->ele1][q[S_LR]]; break;
   case K_LRRL: p = eles[t->ele0][q[S_RR]] + eles[t->ele1][q[S_LR]]; break;
   case K_LRRR: p = eles[t->ele0][q[S_LR]]; break;
   case K_RRRL: p = eles[t->ele0][q[S_RR]]; break;
   default: return 1;
   }
   int even = (p % 2 == 0);
   if (t->sign_mode == 1) return even ? -1 : 1;
   return even ? 1 : -1;
}

static inline int64_t sb_index(const int q[4], int32_t dim_RR, int32_t d) {
   /* inv_sup of DMRG_MAKE_ELEM_LL_LLLRRRRL.c:26 */
   return (int64_t)q[S_LL] * d * dim_RR * d + (int64_t)q[S_LR] * dim_RR * d + (int64_t)q[S_RR] * d + q[S_RL];
}

/* MAKE_ELEMENT_HAM_LLLRRRRL for one row: run the term list in call order */
static void make_row(int32_t basis, int32_t dim_RR, int32_t d,
                     const int16_t *LL, const int16_t *LR, const int16_t *RR, const int16_t *RL,
                     const int32_t *inv_whole, const or_crs *ops, const int32_t *const *eles,
                     int32_t n_terms, const or_term *terms, or_acc *a) {
   int q[4] = {LL[basis], LR[basis], RR[basis], RL[basis]};
   a->num = 0;
   for (int32_t t = 0; t < n_terms; t++) {
      const or_term *tm = &terms[t];
      if (tm->kind == K_ZERO) {
         /* dmrg/DMRG_MAKE_ELEM_ZERO_LLLRRRRL.c:15-26 */
         int32_t inv = inv_whole[sb_index(q, dim_RR, d)];
         if (inv >= 0 && a->acc_inv[inv] == -1) acc_put(a, inv, 0.0);
         continue;
      }
      if (fabs(tm->coeff) < OR_ZERO) continue; /* DMRG_MAKE_ELEM_*.c:20-22 */
      const or_crs *m1 = &ops[tm->op1];
      int s1 = op1_slot[tm->kind];
      if (op2_slot[tm->kind] < 0) {
         /* DMRG_MAKE_ELEM_{LL,LR,RR,RL}_LLLRRRRL.c:24-44 */
         for (int64_t i = m1->row[q[s1]]; i < m1->row[q[s1] + 1]; i++) {
            int qq[4] = {q[0], q[1], q[2], q[3]};
            qq[s1] = m1->col[i];
            int32_t inv = inv_whole[sb_index(qq, dim_RR, d)];
            double val = tm->coeff * m1->val[i];
            if (fabs(val) > OR_ZERO && inv >= 0) acc_put(a, inv, val);
         }
      } else {
         /* DMRG_MAKE_ELEM_LLLR_LLLRRRRL.c:54-79 and siblings: op1 row outer, op2 row inner */
         const or_crs *m2 = &ops[tm->op2];
         int s2 = op2_slot[tm->kind];
         int sign = term_sign(tm, eles, q);
         for (int64_t i = m1->row[q[s1]]; i < m1->row[q[s1] + 1]; i++) {
            int c1 = m1->col[i];
            double v1 = m1->val[i];
            for (int64_t j = m2->row[q[s2]]; j < m2->row[q[s2] + 1]; j++) {
               int qq[4] = {q[0], q[1], q[2], q[3]};
               qq[s1] = c1;
               qq[s2] = m2->col[j];
               int32_t inv = inv_whole[sb_index(qq, dim_RR, d)];
               double val = (double)sign * tm->coeff * v1 * m2->val[j]; /* :63, left to right */
               if (fabs(val) > OR_ZERO && inv >= 0) acc_put(a, inv, val);
            }
         }
      }
   }
}

static void sort_cols(int32_t *c, double *v, int64_t len) {
   /* ascending insertion sort; columns are unique within a row
      (SORT_COLUMN_CRS1.c:5-17 gives the same order) */
   for (int64_t i = 1; i < len; i++) {
      int32_t kc = c[i];
      double kv = v[i];
      int64_t j = i - 1;
      while (j >= 0 && c[j] > kc) {
         c[j + 1] = c[j];
         v[j + 1] = v[j];
         j--;
      }
      c[j + 1] = kc;
      v[j + 1] = kv;
   }
}

int64_t or_assemble(int32_t n, int32_t dim_LL, int32_t dim_RR, int32_t d,
                    const int16_t *LL, const int16_t *LR, const int16_t *RR, const int16_t *RL,
                    const int32_t *inv_whole, int32_t n_ops, const or_crs *ops,
                    int32_t n_eles, const int32_t *const *eles, int32_t n_terms, const or_term *terms,
                    int64_t *row_out, int32_t *col_out, double *val_out) {
   (void)dim_LL; (void)n_ops; (void)n_eles;
   or_acc a;
   a.acc_inv = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
   for (int32_t i = 0; i < n; i++) a.acc_inv[i] = -1;
   a.cap = 256;
   a.e_col = malloc(sizeof(int32_t) * (size_t)a.cap);
   a.e_val = malloc(sizeof(double) * (size_t)a.cap);
   const int fill = (col_out != NULL);
   int64_t total = 0;
   if (!fill) row_out[0] = 0;
   for (int32_t basis = 0; basis < n; basis++) {
      make_row(basis, dim_RR, d, LL, LR, RR, RL, inv_whole, ops, eles, n_terms, terms, &a);
      /* keep rule of GET_HAM_LLLRRRRL.c:49 / :87 */
      int64_t k = fill ? row_out[basis] : 0;
      int64_t cnt = 0;
      for (int32_t e = 0; e < a.num; e++) {
         int32_t inv = a.e_col[e];
         double val = a.e_val[e];
         if ((0 <= inv && inv <= basis && fabs(val) > OR_ZERO) || inv == basis) {
            if (fill) {
               col_out[k + cnt] = inv;
               val_out[k + cnt] = val;
            }
            cnt++;
         }
         a.acc_inv[inv] = -1; /* reset loop of MAKE_ELEMENT_HAM_LLLRRRRL.c:82-90 */
      }
      if (fill) sort_cols(col_out + k, val_out + k, cnt);
      else row_out[basis + 1] = row_out[basis] + cnt;
      total += cnt;
   }
   free(a.acc_inv);
   free(a.e_col);
   free(a.e_val);
   return total;
}

/* ------------------------------------------------------ tridiagonal solver */
/* number of eigenvalues of T strictly below x (Sturm sequence of LDL^T pivots) */
static int sturm_count(int32_t k, const double *a, const double *b, double x) {
   int cnt = 0;
   double q = a[0] - x;
   if (q < 0) cnt++;
   for (int32_t i = 1; i < k; i++) {
      if (q == 0.0) q = 1e-300;
      q = (a[i] - x) - b[i - 1] * b[i - 1] / q;
      if (q < 0) cnt++;
   }
   return cnt;
}

int or_tridiag_lowest(int32_t k, const double *a, const double *b, double *eval, double *evec) {
   if (k <= 0) return -1;
   if (k == 1) {
      *eval = a[0];
      if (evec) evec[0] = 1.0;
      return 0;
   }
   /* Gershgorin interval */
   double lo = a[0], hi = a[0];
   for (int32_t i = 0; i < k; i++) {
      double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < k - 1 ? fabs(b[i]) : 0.0);
      if (a[i] - r < lo) lo = a[i] - r;
      if (a[i] + r > hi) hi = a[i] + r;
   }
   double span = hi - lo;
   lo -= 1e-12 * span + 1e-300;
   hi += 1e-12 * span + 1e-300;
   for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      if (sturm_count(k, a, b, mid) >= 1) hi = mid; else lo = mid;
   }
   double lam = 0.5 * (lo + hi);
   *eval = lam;
   if (!evec) return 0;
   /* inverse iteration: (T - lam I) x = y by Gaussian elimination with partial pivoting */
   double *dl = malloc(sizeof(double) * (size_t)k), *dd = malloc(sizeof(double) * (size_t)k);
   double *du = malloc(sizeof(double) * (size_t)k), *du2 = malloc(sizeof(double) * (size_t)k);
   int *piv = malloc(sizeof(int) * (size_t)k);
   double tiny = 1e-15 * (fabs(span) + 1e-300);
   for (int32_t i = 0; i < k; i++) {
      dd[i] = a[i] - lam;
      dl[i] = (i < k - 1) ? b[i] : 0.0;
      du[i] = (i < k - 1) ? b[i] : 0.0;
      du2[i] = 0.0;
   }
   for (int32_t i = 0; i < k - 1; i++) {
      if (fabs(dd[i]) >= fabs(dl[i])) {
         piv[i] = 0;
         if (dd[i] == 0.0) dd[i] = tiny;
         double f = dl[i] / dd[i];
         dl[i] = f;
         dd[i + 1] -= f * du[i];
         du2[i] = 0.0;
      } else {
         piv[i] = 1;
         double f = dd[i] / dl[i];
         dd[i] = dl[i];
         dl[i] = f;
         double t = du[i];
         du[i] = dd[i + 1];
         dd[i + 1] = t - f * dd[i + 1];
         if (i < k - 2) {
            du2[i] = du[i + 1];
            du[i + 1] = -f * du[i + 1];
         }
      }
   }
   if (dd[k - 1] == 0.0) dd[k - 1] = tiny;
   for (int32_t i = 0; i < k; i++) evec[i] = 1.0;
   for (int rep = 0; rep < 3; rep++) {
      /* forward: apply L^-1 with the recorded row swaps */
      for (int32_t i = 0; i < k - 1; i++) {
         if (piv[i] == 0) {
            evec[i + 1] -= dl[i] * evec[i];
         } else {
            double t = evec[i];
            evec[i] = evec[i + 1];
            evec[i + 1] = t - dl[i] * evec[i + 1];
         }
      }
      /* back substitution with U (diag dd, super du, super-super du2) */
      evec[k - 1] /= dd[k - 1];
      if (k > 1) evec[k - 2] = (evec[k - 2] - du[k - 2] * evec[k - 1]) / dd[k - 2];
      for (int32_t i = k - 3; i >= 0; i--)
         evec[i] = (evec[i] - du[i] * evec[i + 1] - du2[i] * evec[i + 2]) / dd[i];
      double nrm = 0.0;
      for (int32_t i = 0; i < k; i++) nrm += evec[i] * evec[i];
      nrm = 1.0 / sqrt(nrm);
      for (int32_t i = 0; i < k; i++) evec[i] *= nrm;
   }
   free(dl); free(dd); free(du); free(du2); free(piv);
   return 0;
}

/* ----------------------------------------------------------- BLAS-1 helpers */
static double dot(int64_t n, const double *x, const double *y) {
   double s = 0.0;
   for (int64_t i = 0; i < n; i++) s = s + x[i] * y[i];
   return s;
}
static void normalize(int64_t n, double *v) { /* sml/NORMALIZE.c:6-23 */
   double s = 1.0 / sqrt(dot(n, v, v));
   for (int64_t i = 0; i < n; i++) v[i] = s * v[i];
}

void or_crs_matvec(void *ctx, const double *x, double *y) {
   const or_crs_op *op = (const or_crs_op *)ctx;
   if (op->sym) or_spmv_sym(op->n, op->row, op->col, op->val, x, y);
   else or_spmv_full(op->n, op->row, op->col, op->val, x, y);
}

/* ----------------------------------------------------------------- Lanczos */
int or_lanczos_slow(int64_t n, or_matvec_fn mv, void *ctx, const double *start,
                    double acc, int min_step, int max_step, int type,
                    double *eig_val, double *eig_vec) {
   if (max_step <= min_step) return -2; /* sml/LANCZOS_SLOW_SYM.c:26-30 */
   double *v0 = malloc(sizeof(double) * (size_t)n), *v1 = malloc(sizeof(double) * (size_t)n);
   double *v2 = malloc(sizeof(double) * (size_t)n);
   double *alpha = calloc((size_t)max_step + 1, sizeof(double));
   double *beta = calloc((size_t)max_step + 1, sizeof(double));
   double *s = calloc((size_t)max_step + 1, sizeof(double));
   double E = 1e15, lam = 0.0;
   int conv = 0, k;

   memcpy(v0, start, sizeof(double) * (size_t)n);
   normalize(n, v0);                                   /* :98  */
   mv(ctx, v0, v1);                                    /* :99  */
   alpha[0] = dot(n, v0, v1);                          /* :100 */
   for (int64_t i = 0; i < n; i++) v1[i] = v1[i] - alpha[0] * v0[i];
   for (k = 1; k < max_step; k++) {                    /* :112 */
      memcpy(v2, v1, sizeof(double) * (size_t)n);
      beta[k] = sqrt(dot(n, v2, v2));                  /* :116 */
      normalize(n, v2);
      mv(ctx, v2, v1);                                 /* :120 */
      alpha[k] = dot(n, v1, v2);                       /* :122 */
      if (k >= min_step) {                             /* :125-181 */
         or_tridiag_lowest(k + 1, alpha, beta + 1, &lam, s);
         if (type == 1) conv = fabs(s[k]) < acc;
         else conv = fabs(E - lam) < acc;
         if (conv) break;
         E = lam;
      }
      for (int64_t i = 0; i < n; i++) v1[i] = v1[i] - alpha[k] * v2[i] - beta[k] * v0[i];
      double *t = v0; v0 = v2; v2 = t;                 /* COPY_DOUBLE1(V2,V0) :188 */
   }
   int ret = -1;
   if (conv) {
      /* replay from the same start (:193-261) */
      *eig_val = lam;
      memcpy(v0, start, sizeof(double) * (size_t)n);
      normalize(n, v0);
      for (int64_t i = 0; i < n; i++) eig_vec[i] = s[0] * v0[i];
      mv(ctx, v0, v1);
      for (int64_t i = 0; i < n; i++) v1[i] = v1[i] - alpha[0] * v0[i];
      for (int it = 1; it <= k; it++) {
         memcpy(v2, v1, sizeof(double) * (size_t)n);
         normalize(n, v2);
         for (int64_t i = 0; i < n; i++) eig_vec[i] = eig_vec[i] + s[it] * v2[i];
         if (it == k) break;
         mv(ctx, v2, v1);
         for (int64_t i = 0; i < n; i++) v1[i] = v1[i] - alpha[it] * v2[i] - beta[it] * v0[i];
         double *t = v0; v0 = v2; v2 = t;
      }
      normalize(n, eig_vec);
      ret = k;
   }
   free(v0); free(v1); free(v2); free(alpha); free(beta); free(s);
   return ret;
}

/* -------------------------------------------------- inverse iteration + CG */
typedef struct { or_matvec_fn mv; void *ctx; double shift; int64_t n; } shifted_op;
static void shifted_mv(void *c, const double *x, double *y) {
   shifted_op *s = (shifted_op *)c;
   s->mv(s->ctx, x, y);
   for (int64_t i = 0; i < s->n; i++) y[i] = y[i] + x[i] * s->shift;
}

/* sml/CONJUGATE_GRADIENT_SYM.c:8-123: solves A x = b, x (in/out) normalized first */
static int cg_solve(int64_t n, or_matvec_fn mv, void *ctx, const double *b, double *x, double acc, int max_step) {
   double *r = malloc(sizeof(double) * (size_t)n), *p = malloc(sizeof(double) * (size_t)n);
   double *y = malloc(sizeof(double) * (size_t)n);
   normalize(n, x);                                                /* :202 */
   mv(ctx, x, r);                                                  /* :204 */
   for (int64_t i = 0; i < n; i++) { r[i] = b[i] - r[i]; p[i] = r[i]; }
   int iter;
   for (iter = 1; iter < max_step; iter++) {                       /* :212 */
      mv(ctx, p, y);
      double rr = dot(n, r, r);
      double alpha = rr / dot(n, p, y);
      for (int64_t i = 0; i < n; i++) x[i] = x[i] + alpha * p[i];
      for (int64_t i = 0; i < n; i++) r[i] = r[i] - alpha * y[i];
      double t = dot(n, r, r);
      if (t < acc) break;                                          /* :236 */
      double beta = t / rr;
      for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
   }
   free(r); free(p); free(y);
   return iter;
}

/* sml/MINIMUM_RESIDUAL.c:7-150: MINRES from x as given; returns the iteration at which
 * |eta| < acc (x then holds the previous iterate, as the reference leaves it), or -1. */
int or_minimum_residual(int64_t n, or_matvec_fn mv, void *ctx, const double *b, double *x, double acc,
                        int max_step) {
   size_t sz = sizeof(double) * (size_t)n;
   double *v0 = calloc((size_t)n, sizeof(double)), *v1 = malloc(sz), *v2 = malloc(sz);
   double *w0 = calloc((size_t)n, sizeof(double)), *w1 = calloc((size_t)n, sizeof(double));
   double *w2 = malloc(sz), *x1 = malloc(sz), *t = malloc(sz);
   mv(ctx, x, v1);
   for (int64_t i = 0; i < n; i++) v1[i] = b[i] - v1[i];            /* :31-38 */
   double gamma0 = sqrt(dot(n, v1, v1)), eta = gamma0;              /* :39-41 */
   double s0 = 0, s1 = 0, c0 = 1, c1 = 1;
   int ret = -1;
   for (int iter = 0; iter < max_step; iter++) {
      double inv = 1.0 / gamma0;
      for (int64_t i = 0; i < n; i++) v1[i] = v1[i] * inv;         /* :48-51 */
      mv(ctx, v1, t);
      double delta = dot(n, t, v1);
      for (int64_t i = 0; i < n; i++) v2[i] = t[i] - delta * v1[i] - gamma0 * v0[i];
      double gamma1 = sqrt(dot(n, v2, v2));
      double a0 = c1 * delta - c0 * s1 * gamma0;                      /* :66-73 */
      double a1 = sqrt(a0 * a0 + gamma1 * gamma1);
      double a2 = s1 * delta + c0 * c1 * gamma0;
      double a3 = s0 * gamma0;
      double c2 = a0 / a1, s2 = gamma1 / a1;
      double ti = 1.0 / a1;
      for (int64_t i = 0; i < n; i++) {
         w2[i] = ti * (v1[i] - a3 * w0[i] - a2 * w1[i]);
         x1[i] = x[i] + c2 * eta * w2[i];
      }
      eta = -s2 * eta;
      if (fabs(eta) < acc) { ret = iter; break; }                     /* :78-99 */
      memcpy(v0, v1, sz); memcpy(v1, v2, sz); memcpy(w0, w1, sz); memcpy(w1, w2, sz); memcpy(x, x1, sz);
      s0 = s1; s1 = s2; c0 = c1; c1 = c2; gamma0 = gamma1;
   }
   free(v0); free(v1); free(v2); free(w0); free(w1); free(w2); free(x1); free(t);
   return ret;
}

double or_inverse_iteration(int64_t n, or_matvec_fn mv, void *ctx, double eig_val, double *vec,
                            double ii_acc, double ii_diag_add, int ii_max_step,
                            double cg_acc, int cg_max_step, int *cg_steps, int *ii_steps) {
   shifted_op sh = {mv, ctx, ii_diag_add - eig_val, n};   /* INVERSE_ITERATION_SYM.c:24 */
   double *un = malloc(sizeof(double) * (size_t)n), *tmp = malloc(sizeof(double) * (size_t)n);
   double err = 1000.0;
   int total = 0, iter;
   memcpy(un, vec, sizeof(double) * (size_t)n);            /* :26 */
   for (iter = 0; iter <= ii_max_step; iter++) {
      /* CONVERGE_CHECK_SYM(M', v, ii_diag_add) = sum |ii_diag_add*v - M'v| (:30) */
      shifted_mv(&sh, vec, tmp);
      err = 0.0;
      for (int64_t i = 0; i < n; i++) err = err + fabs(ii_diag_add * vec[i] - tmp[i]);
      if (err < ii_acc) break;
      total += cg_solve(n, shifted_mv, &sh, vec, un, cg_acc, cg_max_step); /* x0 = un (:57) */
      normalize(n, un);
      memcpy(vec, un, sizeof(double) * (size_t)n);
   }
   if (cg_steps) *cg_steps = total;
   if (ii_steps) *ii_steps = iter;
   free(un); free(tmp);
   return err;
}

int or_diagonalize_superblock_sym(int32_t n, const int64_t *row, const int32_t *col, const double *val,
                                  const double *start, double diag_acc, int min_step, int max_step,
                                  int lan_type, double ii_acc, double ii_diag_add, int ii_max_step,
                                  double cg_acc, int cg_max_step,
                                  double *gs_val, double *gs_error, double *gs_vec,
                                  int *lan_steps, int *cg_steps) {
   or_crs_op op = {n, row, col, val, 1};
   double e = 0.0;
   int k = or_lanczos_slow(n, or_crs_matvec, &op, start, diag_acc, min_step, max_step, lan_type, &e, gs_vec);
   if (lan_steps) *lan_steps = k;
   if (k < 0) return -1;
   *gs_error = or_inverse_iteration(n, or_crs_matvec, &op, e, gs_vec, ii_acc, ii_diag_add, ii_max_step,
                                    cg_acc, cg_max_step, cg_steps, NULL);
   *gs_val = e;   /* the Lanczos eigenvalue (DMRG_DIAGONALIZE_SUPERBLOCK_SYM.c:56) */
   return 0;
}

/* --------------------------------------------------------- matrix-free OBC */
void or_obc_matvec(void *ctx, const double *V, double *Out) {
   const or_obc_op *o = (const or_obc_op *)ctx;
   int32_t d = o->d, dim_RR = o->dim_RR;
   for (int32_t b = 0; b < o->n; b++) {
      int LL = o->LL[b], LR = o->LR[b], RR = o->RR[b], RL = o->RL[b];
      int LLLR = o->inv_LLLR[LL * d + LR];
      int LRRL = o->inv_LRRL[LR * d + RL];
      int RRRL = o->inv_RRRL[RR * d + RL];
      double val = 0.0;
      for (int64_t it = o->M_LLLR.row[LLLR]; it < o->M_LLLR.row[LLLR + 1]; it++) {  /* :29-38 */
         int c = o->M_LLLR.col[it];
         int64_t s = (int64_t)o->LL_LLLR[c] * d * dim_RR * d + (int64_t)o->LR_LLLR[c] * dim_RR * d + (int64_t)RR * d + RL;
         int32_t inv = o->inv_whole[s];
         if (inv >= 0) val = val + V[inv] * o->M_LLLR.val[it];
      }
      for (int64_t it = o->M_LRRL.row[LRRL]; it < o->M_LRRL.row[LRRL + 1]; it++) {  /* :41-50 */
         int c = o->M_LRRL.col[it];
         int64_t s = (int64_t)LL * d * dim_RR * d + (int64_t)o->LR_LRRL[c] * dim_RR * d + (int64_t)RR * d + o->RL_LRRL[c];
         int32_t inv = o->inv_whole[s];
         if (inv >= 0) val = val + V[inv] * o->M_LRRL.val[it];
      }
      int sign = (o->Ele_RR[RR] % 2 == 1) ? -1 : 1;                               /* :52-57 */
      if (o->M_LRRL_Sign.row) {
         for (int64_t it = o->M_LRRL_Sign.row[LRRL]; it < o->M_LRRL_Sign.row[LRRL + 1]; it++) {
            int c = o->M_LRRL_Sign.col[it];
            int64_t s = (int64_t)LL * d * dim_RR * d + (int64_t)o->LR_LRRL[c] * dim_RR * d + (int64_t)RR * d + o->RL_LRRL[c];
            int32_t inv = o->inv_whole[s];
            if (inv >= 0) val = val + V[inv] * o->M_LRRL_Sign.val[it] * sign;
         }
      }
      for (int64_t it = o->M_RRRL.row[RRRL]; it < o->M_RRRL.row[RRRL + 1]; it++) {  /* :72-81 */
         int c = o->M_RRRL.col[it];
         int64_t s = (int64_t)LL * d * dim_RR * d + (int64_t)LR * dim_RR * d + (int64_t)o->RR_RRRL[c] * d + o->RL_RRRL[c];
         int32_t inv = o->inv_whole[s];
         if (inv >= 0) val = val + V[inv] * o->M_RRRL.val[it];
      }
      Out[b] = val;
   }
}

/* ------------------------------------------------------ expectation values */
void or_v_m_q0(int slot, int32_t n, int32_t dim_RR, int32_t d, const int16_t *LL, const int16_t *LR,
               const int16_t *RR, const int16_t *RL, const int32_t *inv_whole, const int64_t *row,
               const int32_t *col, const double *val, const double *x, double *y) {
   for (int32_t b = 0; b < n; b++) {
      int q[4] = {LL[b], LR[b], RR[b], RL[b]};
      const int own = q[slot];
      double v = 0.0;
      for (int64_t j = row[own]; j < row[own + 1]; j++) {
         q[slot] = col[j];
         int32_t inv = inv_whole[sb_index(q, dim_RR, d)];
         if (inv >= 0) v = v + x[inv] * val[j];   /* DMRG_V_M_LL_Q0.c:24-26 */
      }
      y[b] = v;
   }
}

void or_v_m_sector(int slot, int32_t n_out, int32_t dim_RR, int32_t d, const int16_t *LL, const int16_t *LR,
                   const int16_t *RR, const int16_t *RL, const int32_t *inv_whole, const int64_t *row,
                   const int32_t *col, const double *val, const int32_t *ele_LL, const int32_t *ele_LR,
                   const int32_t *ele_RR, int32_t sign_flag, const double *x, double *y) {
   for (int32_t b = 0; b < n_out; b++) {
      int q[4] = {LL[b], LR[b], RR[b], RL[b]};
      int sign = 1;
      if (sign_flag && slot > 0) {   /* DMRG_V_M_RL_Q2.c:30-39: parity of the earlier slots */
         int e = ele_LL[q[0]];
         if (slot >= 2) e += ele_LR[q[1]];
         if (slot >= 3) e += ele_RR[q[2]];
         sign = (e % 2 == 0) ? 1 : -1;
      }
      const int own = q[slot];
      double v = 0.0;
      for (int64_t j = row[own]; j < row[own + 1]; j++) {
         q[slot] = col[j];
         int32_t inv = inv_whole[sb_index(q, dim_RR, d)];
         if (inv >= 0) v = v + sign * x[inv] * val[j];   /* DMRG_V_M_RL_Q2.c:44-49 */
      }
      y[b] = v;
   }
}

int or_density_matrix(int32_t n, int32_t dim_RR, int32_t d, int32_t dim_LLLR, const int16_t *LL_LLLR,
                      const int16_t *LR_LLLR, const int32_t *Sum_LLLR, const int16_t *RR, const int16_t *RL,
                      const int32_t *inv_whole, const double *vec, int32_t s0, int32_t db, double *out) {
   int64_t *base = calloc((size_t)dim_LLLR + 1, sizeof(int64_t));   /* Base_LLLR, :20-22 */
   for (int32_t s = 0; s < dim_LLLR; s++) base[s + 1] = base[s] + Sum_LLLR[s];
   int bad = base[dim_LLLR] != n || s0 < 0 || s0 + db > dim_LLLR;
   for (int64_t i = 0; i < (int64_t)db * db; i++) out[i] = 0.0;
   for (int32_t r = 0; r < db && !bad; r++) {
      const int32_t sr = s0 + r;
      for (int32_t c = r; c < db && !bad; c++) {
         const int32_t sc = s0 + c;
         if (Sum_LLLR[sr] != Sum_LLLR[sc]) { bad = 1; break; }
         double m = 0.0;
         for (int32_t k = 0; k < Sum_LLLR[sr]; k++) {   /* :52-83 */
            const int64_t pr = base[sr] + k, pc = base[sc] + k;
            if (RR[pr] != RR[pc] || RL[pr] != RL[pc]) { bad = 1; break; }
            int qr[4] = {LL_LLLR[sr], LR_LLLR[sr], RR[pr], RL[pr]};
            int qc[4] = {LL_LLLR[sc], LR_LLLR[sc], RR[pc], RL[pc]};
            const int32_t ir = inv_whole[sb_index(qr, dim_RR, d)], ic = inv_whole[sb_index(qc, dim_RR, d)];
            if (ir < 0 || ic < 0) { bad = 1; break; }
            m = m + vec[ir] * vec[ic];
         }
         out[(int64_t)r * db + c] = m;
      }
   }
   for (int32_t r = 0; r < db; r++)   /* :97-101 */
      for (int32_t c = r; c < db; c++) out[(int64_t)c * db + r] = out[(int64_t)r * db + c];
   free(base);
   return bad ? -1 : 0;
}

static void swap_i32(int32_t *a, int32_t *b) { int32_t t = *a; *a = *b; *b = t; }

int32_t or_system_select(int32_t nb, const int32_t *dim_block, const int32_t *qb1, const int32_t *qb2,
                         const double *w, int32_t max_dim, double *val_dist, double *sums, int32_t *kept_block,
                         int32_t *kept_basis) {
   int32_t dim = 0;
   for (int32_t b = 0; b < nb; b++) dim += dim_block[b];
   int32_t *blk = malloc(sizeof(int32_t) * dim), *bas = malloc(sizeof(int32_t) * dim);
   int32_t *q1 = malloc(sizeof(int32_t) * dim), *q2 = malloc(sizeof(int32_t) * dim);
   int64_t *first = malloc(sizeof(int64_t) * (nb + 1));
   first[0] = 0;
   for (int32_t b = 0; b < nb; b++) first[b + 1] = first[b] + dim_block[b];
   int32_t c = 0;
   for (int32_t b = 0; b < nb; b++)   /* :38-46 */
      for (int32_t i = 0; i < dim_block[b]; i++, c++) { blk[c] = b; bas[c] = i; q1[c] = qb1[b]; q2[c] = qb2[b]; }
#define VAL(k) (w[first[blk[k]] + bas[k]])
#define SWAP4(k) do { swap_i32(&blk[k], &blk[k + 1]); swap_i32(&bas[k], &bas[k + 1]); \
                      swap_i32(&q1[k], &q1[k + 1]); swap_i32(&q2[k], &q2[k + 1]); } while (0)
   for (int32_t i = 0; i < dim - 1; i++) {   /* :55-72 */
      int moved = 0;
      for (int32_t j = 0; j < dim - 1 - i; j++)
         if (VAL(j) < VAL(j + 1)) { SWAP4(j); moved = 1; }
      if (!moved) break;
   }
   double sum = 0;
   for (int32_t i = 0; i < dim; i++) { val_dist[i] = VAL(i); sum = sum + VAL(i); }   /* :77-81 */
   int32_t renorm = 0;
   const double zero = pow(10, -15);
   if (dim <= max_dim) {   /* :85-110 */
      for (int32_t i = dim; i >= 1; i--)
         if (VAL(i - 1) > zero) { renorm = i; break; }
   } else {
      for (int32_t i = max_dim; i >= 1; i--)
         if (fabs(VAL(i - 1) - VAL(i)) > zero && VAL(i - 1) > zero) { renorm = i; break; }
   }
   double kept = 0;
   for (int32_t i = 0; i < renorm; i++) kept = kept + VAL(i);   /* :116-120 */
   sums[0] = sum;
   sums[1] = sum - kept;
   for (int32_t i = 0; i < renorm - 1; i++) {   /* :124-139, by Q_Number2 */
      int moved = 0;
      for (int32_t j = 0; j < renorm - 1 - i; j++)
         if (q2[j] > q2[j + 1]) { SWAP4(j); moved = 1; }
      if (!moved) break;
   }
   for (int32_t i = 0; i < renorm - 1; i++) {   /* :141-158, by Q_Number1 */
      int moved = 0;
      for (int32_t j = 0; j < renorm - 1 - i; j++)
         if (q1[j] > q1[j + 1]) { SWAP4(j); moved = 1; }
      if (!moved) break;
   }
#undef VAL
#undef SWAP4
   for (int32_t i = 0; i < renorm; i++) { kept_block[i] = blk[i]; kept_basis[i] = bas[i]; }
   free(blk); free(bas); free(q1); free(q2); free(first);
   return renorm > 0 ? renorm : -1;
}

int64_t or_crs_crs_ccs_product(int32_t n1, const int64_t *r1, const int32_t *c1, const double *v1, int32_t n2,
                               const int64_t *r2, const int32_t *c2, const double *v2, int32_t n3, int32_t k,
                               const int64_t *cptr3, const int32_t *ridx3, const double *v3, int64_t *out_row,
                               int32_t *out_col, double *out_val) {
   const double zero = pow(10, -15);
   /* pass 1 (:33-63): W = M2 M3 column by column, kept entries in CCS */
   int64_t cap = (int64_t)n2 * k + 1, nw = 0;
   double *wv = malloc(sizeof(double) * cap), *dense = calloc((size_t)(n3 > n2 ? n3 : n2) + 1, sizeof(double));
   int32_t *wr = malloc(sizeof(int32_t) * cap);
   int64_t *wc = malloc(sizeof(int64_t) * ((size_t)k + 1));
   wc[0] = 0;
   for (int32_t j = 0; j < k; j++) {
      for (int64_t e = cptr3[j]; e < cptr3[j + 1]; e++) dense[ridx3[e]] = v3[e];
      for (int32_t i = 0; i < n2; i++) {
         double val = 0;
         for (int64_t e = r2[i]; e < r2[i + 1]; e++) val = val + dense[c2[e]] * v2[e];
         if (fabs(val) > zero) { wv[nw] = val; wr[nw] = i; nw++; }
      }
      for (int64_t e = cptr3[j]; e < cptr3[j + 1]; e++) dense[ridx3[e]] = 0.0;
      wc[j + 1] = nw;
   }
   /* pass 2 (:67-121): out = M1 W row by row */
   for (int32_t i = 0; i <= n2; i++) dense[i] = 0;
   int64_t nnz = 0;
   out_row[0] = 0;
   for (int32_t i = 0; i < n1; i++) {
      for (int64_t e = r1[i]; e < r1[i + 1]; e++) dense[c1[e]] = v1[e];
      for (int32_t j = 0; j < k; j++) {
         double val = 0;
         for (int64_t e = wc[j]; e < wc[j + 1]; e++) val = val + dense[wr[e]] * wv[e];
         if (fabs(val) > zero) {
            if (out_col) { out_col[nnz] = j; out_val[nnz] = val; }
            nnz++;
         }
      }
      for (int64_t e = r1[i]; e < r1[i + 1]; e++) dense[c1[e]] = 0.0;
      out_row[i + 1] = nnz;
   }
   free(wv); free(dense); free(wr); free(wc);
   return nnz;
}

/* ------------------------------------------------------ exact diagonalization */
/* sml/BINARY_SEARCH_LINT1.c: index of target in the ascending array, -1 when absent */
static int64_t bsearch_i64(const int64_t *a, int64_t lo, int64_t hi, int64_t target) {
   hi -= 1;
   while (lo <= hi) {
      int64_t mid = lo + (hi - lo) / 2;
      if (a[mid] == target) return mid;
      if (a[mid] < target) lo = mid + 1;
      else hi = mid - 1;
   }
   return -1;
}

static int ed_local(int64_t code, int32_t site, int32_t d) { /* exact/EXACT_FIND_SITE_STATE.c:3-15 */
   for (int32_t i = 0; i < site; i++) code /= d;
   return (int)(code % d);
}

static int64_t ed_pow(int32_t d, int32_t site) {
   int64_t p = 1;
   for (int32_t i = 0; i < site; i++) p *= d;
   return p;
}

typedef struct { int64_t code; double val; int64_t order; } ed_contrib;

static int ed_cmp(const void *a, const void *b) {
   const ed_contrib *x = a, *y = b;
   if (x->code != y->code) return x->code < y->code ? -1 : 1;
   return x->order < y->order ? -1 : (x->order > y->order);
}

/* one row: contributions in call order (EXACT_MAKE_ELEM_ON.c:18-29, EXACT_MAKE_ELEM_INTER.c:17-34),
 * grouped by target code, summed first-to-last.  Returns the number of distinct targets;
 * codes/vals hold them ascending. */
static int64_t ed_row(int64_t code, int32_t L, int32_t d, const or_crs *ops, const int32_t *n_ele,
                      int32_t n_terms, const or_ed_term *terms, ed_contrib **buf, int64_t *cap,
                      int64_t *codes, double *vals) {
   int64_t m = 0;
   (void)L;
   for (int32_t t = 0; t < n_terms; t++) {
      const or_ed_term *tm = &terms[t];
      const or_crs *a = &ops[tm->op1];
      const int l1 = ed_local(code, tm->site1, d);
      const int64_t p1 = ed_pow(d, tm->site1);
      if (tm->kind == 0) {
         for (int64_t i = a->row[l1]; i < a->row[l1 + 1]; i++) {
            if (m == *cap) { *cap *= 2; *buf = realloc(*buf, sizeof(ed_contrib) * (size_t)*cap); }
            (*buf)[m].code = code + (a->col[i] - l1) * p1;
            (*buf)[m].val = tm->coeff * a->val[i];
            (*buf)[m].order = m;
            m++;
         }
      } else {
         const or_crs *b = &ops[tm->op2];
         const int l2 = ed_local(code, tm->site2, d);
         const int64_t p2 = ed_pow(d, tm->site2);
         int sign = tm->sign;
         if (tm->par_lo < tm->par_hi) {
            int ne = 0;
            for (int32_t s = tm->par_lo; s < tm->par_hi; s++) ne += n_ele[ed_local(code, s, d)];
            if (ne % 2 == 0) sign = -sign;
         }
         for (int64_t i = a->row[l1]; i < a->row[l1 + 1]; i++) {
            for (int64_t j = b->row[l2]; j < b->row[l2 + 1]; j++) {
               if (m == *cap) { *cap *= 2; *buf = realloc(*buf, sizeof(ed_contrib) * (size_t)*cap); }
               (*buf)[m].code = code + (a->col[i] - l1) * p1 + (b->col[j] - l2) * p2;
               (*buf)[m].val = sign * tm->coeff * a->val[i] * b->val[j];
               (*buf)[m].order = m;
               m++;
            }
         }
      }
   }
   qsort(*buf, (size_t)m, sizeof(ed_contrib), ed_cmp);
   int64_t u = 0;
   for (int64_t i = 0; i < m; i++) {
      if (u > 0 && codes[u - 1] == (*buf)[i].code) {
         vals[u - 1] = vals[u - 1] + (*buf)[i].val;
      } else {
         codes[u] = (*buf)[i].code;
         vals[u] = (*buf)[i].val;
         u++;
      }
   }
   return u;
}

int64_t or_ed_assemble(int32_t L, int32_t d, int32_t n, const int64_t *basis, int32_t n_ops,
                       const or_crs *ops, const int32_t *n_ele, int32_t n_terms, const or_ed_term *terms,
                       int64_t *row_out, int32_t *col_out, double *val_out) {
   (void)n_ops;
   int64_t cap = 4096;
   ed_contrib *buf = malloc(sizeof(ed_contrib) * (size_t)cap);
   int64_t ucap = 4096;
   int64_t *codes = malloc(sizeof(int64_t) * (size_t)ucap);
   double *vals = malloc(sizeof(double) * (size_t)ucap);
   int64_t nnz = 0;
   if (!col_out) row_out[0] = 0;
   for (int32_t r = 0; r < n; r++) {
      /* the number of targets never exceeds the number of contributions */
      int64_t need = 0;
      for (int32_t t = 0; t < n_terms; t++) {
         const or_crs *a = &ops[terms[t].op1];
         int64_t k = a->row[ed_local(basis[r], terms[t].site1, d) + 1] - a->row[ed_local(basis[r], terms[t].site1, d)];
         if (terms[t].kind == 1) {
            const or_crs *b = &ops[terms[t].op2];
            const int l2 = ed_local(basis[r], terms[t].site2, d);
            k *= b->row[l2 + 1] - b->row[l2];
         }
         need += k;
      }
      if (need > ucap) {
         ucap = need;
         codes = realloc(codes, sizeof(int64_t) * (size_t)ucap);
         vals = realloc(vals, sizeof(double) * (size_t)ucap);
      }
      int64_t u = ed_row(basis[r], L, d, ops, n_ele, n_terms, terms, &buf, &cap, codes, vals);
      int64_t p = col_out ? row_out[r] : 0;
      for (int64_t i = 0; i < u; i++) {
         int64_t inv = bsearch_i64(basis, 0, n, codes[i]);
         if ((fabs(vals[i]) > OR_ZERO && inv >= 0) || inv == r) { /* GET_HAM.c:35,63 */
            if (col_out) { col_out[p] = (int32_t)inv; val_out[p] = vals[i]; }
            p++;
         }
      }
      if (!col_out) row_out[r + 1] = row_out[r] + p;
      nnz += col_out ? p - row_out[r] : p;
   }
   free(buf);
   free(codes);
   free(vals);
   return nnz;
}

void or_ed_v_m(int32_t d, int32_t site, const or_crs *op, int32_t n_out, const int64_t *basis_out,
               int32_t n_in, const int64_t *basis_in, const double *x, int32_t sign_flag,
               const int32_t *n_ele, double *y) {
   const int64_t p = ed_pow(d, site);
   for (int32_t i = 0; i < n_out; i++) {
      const int64_t code = basis_out[i];
      const int a = ed_local(code, site, d);
      int sign = 1;
      if (sign_flag) {
         int ne = 0;
         for (int32_t s = 0; s < site; s++) ne += n_ele[ed_local(code, s, d)];
         sign = (ne % 2 == 0) ? 1 : -1;   /* EXACT_V_M_Q2.c:35-42 */
      }
      double v = 0.0;
      for (int64_t j = op->row[a]; j < op->row[a + 1]; j++) {
         int64_t inv = bsearch_i64(basis_in, 0, n_in, code - (a - op->col[j]) * p);
         if (inv >= 0) v = v + x[inv] * op->val[j];
      }
      y[i] = v * sign;
   }
}

/* ------------------------------------------------------------ superblock basis */
/* main/dmrg/hubbard_vf/src/allocations/GET_BASIS_SUPERBLOCK.c:63-97: enumerate the product
 * space in (LL, LR, RR, RL) order, keep the target sector, then sort with the key order of
 * dmrg/DMRG_QUICK_SORT_BASIS_Q2.c:29-46 (Q1 for S^z only).  The keys are unique per state,
 * so any correct sort gives the reference's order. */
typedef struct {
   int32_t ele, sz;
   int16_t LL, LR, RR, RL;
} or_bstate;

static int or_bstate_cmp(const void *pa, const void *pb) {
   const or_bstate *a = (const or_bstate *)pa, *b = (const or_bstate *)pb;
   if (a->ele != b->ele) return a->ele < b->ele ? -1 : 1;
   if (a->sz != b->sz) return a->sz < b->sz ? -1 : 1;
   if (a->LL != b->LL) return a->LL < b->LL ? -1 : 1;
   if (a->LR != b->LR) return a->LR < b->LR ? -1 : 1;
   if (a->RR != b->RR) return a->RR < b->RR ? -1 : 1;
   if (a->RL != b->RL) return a->RL < b->RL ? -1 : 1;
   return 0;
}

int64_t or_superblock_basis(int32_t n_qn, int32_t dim_LL, const int32_t *sz_LL, const int32_t *ele_LL,
                            int32_t dim_RR, const int32_t *sz_RR, const int32_t *ele_RR, int32_t d,
                            const int32_t *sz_site, const int32_t *ele_site, int32_t target_sz,
                            int32_t target_ele, int16_t *LL, int16_t *LR, int16_t *RR, int16_t *RL) {
   int64_t n = 0;
   for (int32_t ll = 0; ll < dim_LL; ll++)
      for (int32_t lr = 0; lr < d; lr++)
         for (int32_t rr = 0; rr < dim_RR; rr++)
            for (int32_t rl = 0; rl < d; rl++) {
               const int32_t sz = sz_LL[ll] + sz_site[lr] + sz_RR[rr] + sz_site[rl];
               const int32_t ele = n_qn == 2 ? ele_LL[ll] + ele_site[lr] + ele_RR[rr] + ele_site[rl] : 0;
               if (sz == target_sz && (n_qn != 2 || ele == target_ele)) n++;
            }
   if (!LL || n == 0) return n;
   or_bstate *st = (or_bstate *)malloc(sizeof(or_bstate) * (size_t)n);
   int64_t k = 0;
   for (int32_t ll = 0; ll < dim_LL; ll++)
      for (int32_t lr = 0; lr < d; lr++)
         for (int32_t rr = 0; rr < dim_RR; rr++)
            for (int32_t rl = 0; rl < d; rl++) {
               const int32_t sz = sz_LL[ll] + sz_site[lr] + sz_RR[rr] + sz_site[rl];
               const int32_t ele = n_qn == 2 ? ele_LL[ll] + ele_site[lr] + ele_RR[rr] + ele_site[rl] : 0;
               if (sz != target_sz || (n_qn == 2 && ele != target_ele)) continue;
               or_bstate *b = &st[k++];
               b->ele = n_qn == 2 ? ele_LL[ll] + ele_site[lr] : 0;   /* Tot_Ele_LLLR */
               b->sz = sz_LL[ll] + sz_site[lr];                       /* Tot_Sz_LLLR  */
               b->LL = (int16_t)ll;
               b->LR = (int16_t)lr;
               b->RR = (int16_t)rr;
               b->RL = (int16_t)rl;
            }
   qsort(st, (size_t)n, sizeof(or_bstate), or_bstate_cmp);
   for (int64_t i = 0; i < n; i++) {
      LL[i] = st[i].LL;
      LR[i] = st[i].LR;
      RR[i] = st[i].RR;
      RL[i] = st[i].RL;
   }
   free(st);
   return n;
}
