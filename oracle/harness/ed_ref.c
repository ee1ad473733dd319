/*
 * ed_ref.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Exact-diagonalization golden data from the compiled reference.  Linked against
 * oracle/_ref/libdmrgref.a (the reference's exact/, onsite_op/ and sml/ compiled in
 * place), it builds one ED Hamiltonian with the reference's element makers
 * (exact/EXACT_MAKE_ELEM_ON.c, exact/EXACT_MAKE_ELEM_INTER.c), solves it with
 * EXACT_DIAGONALIZE_HAMILTONIAN and evaluates observables with the reference's
 * EXACT_EXPECTATION_* and EXACT_V_M_Q1/Q2/Q3, then dumps inputs and outputs.
 *
 * The reference ships one ED program (main/exact_diag/klm_vf, which does not compile);
 * its GET_HAM (src/GET_HAM.c:11-92) and MAKE_ELEMENT_HAM (src/MAKE_ELEMENT_HAM.c:11-76)
 * give the pattern this harness follows for the XXZ and Hubbard chains:
 *   onsite Hamiltonian on every site (SSD-weighted), nearest-neighbour couplings with
 *   the fermion sign taken from the row state's electron parity on the left site,
 *   the periodic wrap (parity of sites 0..L-2), then a zero-valued diagonal at site 0.
 * Coupling order per bond follows the DMRG programs' MAKE_ELEMENT_HAM_LLLRRRRL:
 *   XXZ      (main/dmrg/xxz_vf/.../MAKE_ELEMENT_HAM_LLLRRRRL.c):  S+S-, S-S+, SzSz
 *   Hubbard  (main/dmrg/hubbard_vf/.../MAKE_ELEMENT_HAM_LLLRRRRL.c): c+up cup, cup c+up,
 *            c+dn cdn, cdn c+dn, nup nup, ndn ndn
 *
 * Arguments key=value: model=xxz|hubbard L= spin= J_xy= J_z= h_z= D_z= t= U= V= mu=
 *   N= sz= (2*Sz) BC=OBC|PBC|SSD out=<file> method=Lanczos|Lanczos_Slow threads=
 *
 * Output (little endian): "DMRGED01", then
 *   i32 model(0 xxz, 1 hubbard), L, d, n, n_ops
 *   n_ops x CRS {i32 row_dim, col_dim; i64 nnz; i64 Row[row_dim+1]; i32 Col[nnz]; f64 Val[nnz]}
 *     op order  xxz:     Ham_On, Sp, Sm, Sz, Zero
 *               hubbard: Ham_On, CUp_D, CUp, CDown_D, CDown, NC_Up, NC_Down, Zero, NC, SzC
 *   i32 n_ele[d], i32 sz2[d]
 *   i64 basis[n]
 *   CRS Ham (full, as GET_HAM builds it)
 *   f64 E, f64 error, f64 vec[n]
 *   i32 n_obs; n_obs x {i32 kind, op1, op2, site(-1), origin, len; f64 out[len]}
 *       kind 0: EXACT_EXPECTATION_ONSITE(op1) over all sites
 *       kind 1: EXACT_EXPECTATION_INTERSITE_Q0(op1, op2, start=origin, end=L)
 *   i32 n_vm; n_vm x {i32 q, op, site, sign_flag, n_out; i64 basis_out[n_out]; f64 out[n_out]}
 *       EXACT_V_M_Q<q> of op at site from the ground-state sector (in) to basis_out
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "SML.h"
#include "exact.h"
#include "model.h"
#include "onsite.h"

/* reproducible srand(time(NULL)) inside sml/LANCZOS.c */
time_t __wrap_time(time_t *t) {
   if (t) *t = 1234567;
   return 1234567;
}

typedef struct {
   int model, L, spin, N, sz, threads;
   double J_xy, J_z, h_z, D_z, t, U, V, mu;
   char BC[16], method[32], out[512];
} Cfg;

static void parse(Cfg *c, int argc, char **argv) {
   memset(c, 0, sizeof(*c));
   c->L = 8; c->spin = 1; c->threads = 8; c->J_xy = 1.0; c->J_z = 1.0;
   c->t = -1.0; c->U = 4.0; c->N = -1;
   strcpy(c->BC, "OBC"); strcpy(c->method, "Lanczos"); strcpy(c->out, "ed.bin");
   for (int i = 1; i < argc; i++) {
      char *eq = strchr(argv[i], '=');
      if (!eq) { fprintf(stderr, "bad arg %s\n", argv[i]); exit(2); }
      *eq = 0;
      const char *k = argv[i], *v = eq + 1;
      if (!strcmp(k, "model")) c->model = !strcmp(v, "hubbard") ? 1 : 0;
      else if (!strcmp(k, "L")) c->L = atoi(v);
      else if (!strcmp(k, "spin")) c->spin = atoi(v);
      else if (!strcmp(k, "N")) c->N = atoi(v);
      else if (!strcmp(k, "sz")) c->sz = atoi(v);
      else if (!strcmp(k, "threads")) c->threads = atoi(v);
      else if (!strcmp(k, "J_xy")) c->J_xy = atof(v);
      else if (!strcmp(k, "J_z")) c->J_z = atof(v);
      else if (!strcmp(k, "h_z")) c->h_z = atof(v);
      else if (!strcmp(k, "D_z")) c->D_z = atof(v);
      else if (!strcmp(k, "t")) c->t = atof(v);
      else if (!strcmp(k, "U")) c->U = atof(v);
      else if (!strcmp(k, "V")) c->V = atof(v);
      else if (!strcmp(k, "mu")) c->mu = atof(v);
      else if (!strcmp(k, "BC")) snprintf(c->BC, sizeof c->BC, "%s", v);
      else if (!strcmp(k, "method")) snprintf(c->method, sizeof c->method, "%s", v);
      else if (!strcmp(k, "out")) snprintf(c->out, sizeof c->out, "%s", v);
      else { fprintf(stderr, "unknown key %s\n", k); exit(2); }
   }
   if (c->N < 0) c->N = c->L;
}

/* ------------------------------------------------------------------ output */
static FILE *F;
static void w32(int32_t x) { fwrite(&x, 4, 1, F); }
static void w64(int64_t x) { fwrite(&x, 8, 1, F); }
static void wf(double x) { fwrite(&x, 8, 1, F); }
static void wcrs(const CRS1 *M) {
   w32(M->row_dim); w32(M->col_dim);
   int64_t nnz = M->Row[M->row_dim];
   w64(nnz);
   fwrite(M->Row, 8, (size_t)M->row_dim + 1, F);
   for (int64_t i = 0; i < nnz; i++) w32(M->Col[i]);
   fwrite(M->Val, 8, (size_t)nnz, F);
}

/* ------------------------------------------------------------------ ops */
static CRS1 *zero_diag(int d) {
   CRS1 *M = GET_CRS1(d, d * d);
   for (int i = 0; i < d; i++) { M->Col[i] = i; M->Val[i] = 0.0; M->Row[i + 1] = i + 1; }
   M->row_dim = M->col_dim = d;
   return M;
}

enum { X_HAM, X_SP, X_SM, X_SZ, X_ZERO, X_N };
enum { H_HAM, H_CUPD, H_CUP, H_CDND, H_CDN, H_NUP, H_NDN, H_ZERO, H_NC, H_SZC, H_N };

static int n_ops;
static CRS1 *ops[16];
static int d, L;
static int n_ele[16], sz2[16];

static void build_ops(const Cfg *c) {
   if (c->model == 0) {
      d = c->spin + 1;
      /* onsite: D_z SzSz + h_z Sz (main/dmrg/xxz_vf/src/main/MAKE_ONSITE_HAM.c) */
      CRS1 *SzSz = GET_CRS1(d, d * d), *Sz = GET_CRS1(d, d * d);
      ONSITE_SZSZ_SZBASIS_HB(c->spin, SzSz, 1.0);
      ONSITE_SZ_SZBASIS_HB(c->spin, Sz, 1.0);
      for (long i = 0; i < SzSz->Row[d]; i++) SzSz->Val[i] = SzSz->Val[i] * c->D_z;
      for (long i = 0; i < Sz->Row[d]; i++) Sz->Val[i] = Sz->Val[i] * c->h_z;
      ops[X_HAM] = GET_CRS1(d, d * d);
      MATRIX_SUM_CRS1(SzSz, Sz, ops[X_HAM]);
      SORT_COLUMN_CRS1(ops[X_HAM], 1);
      ops[X_SP] = GET_CRS1(d, d * d); ONSITE_SP_SZBASIS_HB(c->spin, ops[X_SP], 1.0);
      ops[X_SM] = GET_CRS1(d, d * d); ONSITE_SM_SZBASIS_HB(c->spin, ops[X_SM], 1.0);
      ops[X_SZ] = GET_CRS1(d, d * d); ONSITE_SZ_SZBASIS_HB(c->spin, ops[X_SZ], 1.0);
      ops[X_ZERO] = zero_diag(d);
      n_ops = X_N;
      for (int a = 0; a < d; a++) { n_ele[a] = 0; sz2[a] = c->spin - 2 * a; }
   } else {
      d = 4;
      MODEL_1DHUBBARD_VF M;
      memset(&M, 0, sizeof M);
      M.dim_onsite = 4; M.h_z = c->h_z; M.U = c->U; M.mu = c->mu; M.t1 = c->t; M.V = c->V;
      for (int i = 0; i < H_N; i++) ops[i] = GET_CRS1(d, d * d);
      ONSITE_MAKE_HAM_SZBASIS_HUBBARD(&M, ops[H_HAM]);
      ONSITE_CUP_DAGGER_SZBASIS_HUBBARD(ops[H_CUPD], 1.0);
      ONSITE_CUP_SZBASIS_HUBBARD(ops[H_CUP], 1.0);
      ONSITE_CDOWN_DAGGER_SZBASIS_HUBBARD(ops[H_CDND], 1.0);
      ONSITE_CDOWN_SZBASIS_HUBBARD(ops[H_CDN], 1.0);
      ONSITE_NC_UP_SZBASIS_HUBBARD(ops[H_NUP], 1.0);
      ONSITE_NC_DOWN_SZBASIS_HUBBARD(ops[H_NDN], 1.0);
      FREE_CRS1(ops[H_ZERO]);
      ops[H_ZERO] = zero_diag(d);
      ONSITE_NC_SZBASIS_HUBBARD(ops[H_NC], 1.0);
      ONSITE_SZC_SZBASIS_HUBBARD(ops[H_SZC], 1.0);
      n_ops = H_N;
      for (int a = 0; a < d; a++) {
         n_ele[a] = ONSITE_FIND_SITE_ELE_SZBASIS_HUBBARD(a);
         sz2[a] = ONSITE_FIND_SITE_SZ_SZBASIS_HUBBARD(a);
      }
   }
}

/* ------------------------------------------------------------------ basis */
static long *sector_basis(int target_sz2, int target_ele, int use_ele, int *dim_out) {
   long tot = 1;
   for (int s = 0; s < L; s++) tot *= d;
   long cap = 1024, n = 0;
   long *B = malloc(sizeof(long) * cap);
   for (long code = 0; code < tot; code++) {
      int s2 = 0, e = 0;
      for (int s = 0; s < L; s++) {
         int a = EXACT_FIND_SITE_STATE(code, s, d);
         s2 += sz2[a];
         e += n_ele[a];
      }
      if (s2 == target_sz2 && (!use_ele || e == target_ele)) {
         if (n == cap) { cap *= 2; B = realloc(B, sizeof(long) * cap); }
         B[n++] = code;
      }
   }
   *dim_out = (int)n;
   return B;
}

/* ------------------------------------------------------------------ Hamiltonian */
static const Cfg *C;

static void make_row(long basis, EXACT_A_BASIS *A) {
   long elem_num = 0;
   for (int site = 0; site < L; site++)
      EXACT_MAKE_ELEM_ON(basis, site, d, ops[0], &elem_num, EXACT_SSD_COEFF(site, L, (char *)C->BC, "Onsite"), A);
   for (int site = 0; site < L - 1; site++) {
      double s = EXACT_SSD_COEFF(site, L, (char *)C->BC, "Intersite");
      if (C->model == 0) {
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[X_SP], ops[X_SM], &elem_num, 0.5 * C->J_xy * s, 1, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[X_SM], ops[X_SP], &elem_num, 0.5 * C->J_xy * s, 1, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[X_SZ], ops[X_SZ], &elem_num, C->J_z * s, 1, A);
      } else {
         int sign = (ONSITE_FIND_SITE_ELE_SZBASIS_HUBBARD(EXACT_FIND_SITE_STATE(basis, site, d)) % 2 == 1) ? 1 : -1;
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[H_CUPD], ops[H_CUP], &elem_num, C->t * s, sign, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[H_CUP], ops[H_CUPD], &elem_num, C->t * s, -1 * sign, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[H_CDND], ops[H_CDN], &elem_num, C->t * s, sign, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[H_CDN], ops[H_CDND], &elem_num, C->t * s, -1 * sign, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[H_NUP], ops[H_NUP], &elem_num, C->V * s, 1, A);
         EXACT_MAKE_ELEM_INTER(basis, site, site + 1, d, ops[H_NDN], ops[H_NDN], &elem_num, C->V * s, 1, A);
      }
   }
   if (strcmp(C->BC, "PBC") == 0 && L > 2) {
      if (C->model == 0) {
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[X_SP], ops[X_SM], &elem_num, 0.5 * C->J_xy, 1, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[X_SM], ops[X_SP], &elem_num, 0.5 * C->J_xy, 1, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[X_SZ], ops[X_SZ], &elem_num, C->J_z, 1, A);
      } else {
         int ne = 0;
         for (int site = 0; site < L - 1; site++)
            ne += ONSITE_FIND_SITE_ELE_SZBASIS_HUBBARD(EXACT_FIND_SITE_STATE(basis, site, d));
         int sign = (ne % 2 == 1) ? 1 : -1;
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[H_CUPD], ops[H_CUP], &elem_num, C->t, sign, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[H_CUP], ops[H_CUPD], &elem_num, C->t, -1 * sign, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[H_CDND], ops[H_CDN], &elem_num, C->t, sign, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[H_CDN], ops[H_CDND], &elem_num, C->t, -1 * sign, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[H_NUP], ops[H_NUP], &elem_num, C->V, 1, A);
         EXACT_MAKE_ELEM_INTER(basis, 0, L - 1, d, ops[H_NDN], ops[H_NDN], &elem_num, C->V, 1, A);
      }
   }
   EXACT_MAKE_ELEM_ON(basis, 0, d, C->model == 0 ? ops[X_ZERO] : ops[H_ZERO], &elem_num, 0.0, A);
   for (long i = 0; i < elem_num; i++) A->Check[i] = -1;
   A->elem_num = elem_num;
}

/* two passes with the keep rule (|val| > 1e-15 && inv >= 0) || inv == row, as
   main/exact_diag/klm_vf/src/GET_HAM.c:25-75; one thread keeps the order simple */
static CRS1 *get_ham(const long *B, int n) {
   EXACT_A_BASIS **A = EXACT_GET_A_BASIS(1, 10000);
   double zero = pow(10, -15);
   long *cnt = calloc((size_t)n + 1, sizeof(long));
   for (int r = 0; r < n; r++) {
      make_row(B[r], A[0]);
      for (long i = 0; i < A[0]->elem_num; i++) {
         long inv = BINARY_SEARCH_LINT1((long *)B, 0, n, A[0]->Basis[i]);
         double val = A[0]->Val[i];
         if ((fabs(val) > zero && inv >= 0) || inv == r) cnt[r + 1]++;
      }
   }
   for (int r = 0; r < n; r++) cnt[r + 1] += cnt[r];
   CRS1 *H = GET_CRS1(n, cnt[n]);
   for (int r = 0; r < n; r++) {
      make_row(B[r], A[0]);
      long p = cnt[r];
      for (long i = 0; i < A[0]->elem_num; i++) {
         long inv = BINARY_SEARCH_LINT1((long *)B, 0, n, A[0]->Basis[i]);
         double val = A[0]->Val[i];
         if ((fabs(val) > zero && inv >= 0) || inv == r) { H->Val[p] = val; H->Col[p] = (int)inv; p++; }
      }
      H->Row[r + 1] = p;
   }
   H->row_dim = H->col_dim = n;
   SORT_COLUMN_CRS1(H, 1);
   EXACT_FREE_A_BASIS(A, 1);
   free(cnt);
   return H;
}

int main(int argc, char **argv) {
   Cfg cfg;
   parse(&cfg, argc, argv);
   C = &cfg;
   L = cfg.L;
   build_ops(&cfg);
   int n;
   long *B = sector_basis(cfg.sz, cfg.N, cfg.model == 1, &n);
   CRS1 *H = get_ham(B, n);
   /* INVERSE_ITERATION adds and removes its shift on the diagonal in place
      (sml/INVERSE_ITERATION.c via DIAG_ADD_CRS1), so keep the assembled values */
   double *val0 = malloc(sizeof(double) * (size_t)H->Row[n]);
   memcpy(val0, H->Val, sizeof(double) * (size_t)H->Row[n]);

   EXACT_PARAMETER P;
   memset(&P, 0, sizeof P);
   P.diag_num = 1;
   P.est_max_row_elem_num = 10000;
   snprintf(P.Diag_Method, sizeof P.Diag_Method, "%s", cfg.method);
   P.diag_acc = pow(10, -14);  /* main/exact_diag/klm_vf/src/main.c:33-46 */
   P.diag_min_step = 1;
   P.diag_max_step = 800;
   P.cg_acc = pow(10, -7);
   P.cg_max_step = 1000;
   P.inv_iter_acc = pow(10, -9);
   P.inv_iter_diag_add = pow(10, -11);
   P.inv_iter_max_step = 3;
   EXACT_TIME T;
   memset(&T, 0, sizeof T);
   EXACT_HAM_INFO I;
   memset(&I, 0, sizeof I);
   I.Ham = H;
   I.Vector = GET_ARRAY_DOUBLE2(1, n);
   I.Value = GET_ARRAY_DOUBLE1(1);
   I.Error = GET_ARRAY_DOUBLE1(1);
   EXACT_DIAGONALIZE_HAMILTONIAN(&I, &P, &T, cfg.threads);
   double *vec = I.Vector[0];

   F = fopen(cfg.out, "wb");
   if (!F) { perror(cfg.out); return 1; }
   fwrite("DMRGED01", 1, 8, F);
   w32(cfg.model); w32(L); w32(d); w32(n); w32(n_ops);
   for (int i = 0; i < n_ops; i++) wcrs(ops[i]);
   for (int a = 0; a < d; a++) w32(n_ele[a]);
   for (int a = 0; a < d; a++) w32(sz2[a]);
   for (int i = 0; i < n; i++) w64(B[i]);
   double *val1 = H->Val;
   H->Val = val0;
   wcrs(H);
   H->Val = val1;
   wf(I.Value[0]); wf(I.Error[0]);
   fwrite(vec, 8, (size_t)n, F);

   /* observables: EXACT_EXPECTATION_ONSITE / _INTERSITE_Q0 */
   EXACT_BASIS_INFO BI;
   memset(&BI, 0, sizeof BI);
   BI.dim = n;
   BI.Basis = B;
   double *t1 = GET_ARRAY_DOUBLE1(n), *t2 = GET_ARRAY_DOUBLE1(n), *out = GET_ARRAY_DOUBLE1(L);
   /* onsite: {op}, intersite: {op1, op2, origin} */
   const int onsite_ops[2][2] = {{X_SZ, X_HAM}, {H_NC, H_SZC}};
   const int inter_ops[2][2][3] = {{{X_SZ, X_SZ, 0}, {X_SZ, X_SZ, L / 4}},
                                   {{H_NC, H_NC, 0}, {H_SZC, H_SZC, L / 4}}};
   w32(4);
   for (int k = 0; k < 2; k++) {
      int op = onsite_ops[cfg.model][k];
      EXACT_EXPECTATION_ONSITE(ops[op], out, vec, t1, d, L, cfg.threads, &BI);
      w32(0); w32(op); w32(-1); w32(-1); w32(0); w32(L);
      fwrite(out, 8, (size_t)L, F);
   }
   for (int k = 0; k < 2; k++) {
      int op1 = inter_ops[cfg.model][k][0], op2 = inter_ops[cfg.model][k][1], o = inter_ops[cfg.model][k][2];
      EXACT_EXPECTATION_INTERSITE_Q0(ops[op1], ops[op2], o, L, out, vec, t1, t2, d, cfg.threads, &BI);
      w32(1); w32(op1); w32(op2); w32(-1); w32(o); w32(L - o);
      fwrite(out, 8, (size_t)(L - o), F);
   }

   /* EXACT_V_M_Q1 / Q2 / Q3 from the ground-state sector to a neighbouring sector */
   if (cfg.model == 0) {
      /* S+ raises 2Sz by 2: sectors indexed by (2Sz + 2L*spin)/2 */
      int nout;
      long *Bo = sector_basis(cfg.sz + 2, 0, 0, &nout);
      int q_in = 0, q_out = 1;
      EXACT_WHOLE_BASIS_Q1 W;
      W.Dim = malloc(sizeof(int) * 2);
      W.Basis = malloc(sizeof(long *) * 2);
      W.Dim[q_in] = n; W.Basis[q_in] = B;
      W.Dim[q_out] = nout; W.Basis[q_out] = Bo;
      double *o1 = GET_ARRAY_DOUBLE1(nout);
      int sites[2] = {0, L / 2};
      w32(2);
      for (int k = 0; k < 2; k++) {
         EXACT_V_M_Q1(ops[X_SP], q_out, vec, q_in, o1, d, sites[k], cfg.threads, &W);
         w32(1); w32(X_SP); w32(sites[k]); w32(0); w32(nout);
         for (int i = 0; i < nout; i++) w64(Bo[i]);
         fwrite(o1, 8, (size_t)nout, F);
      }
   } else {
      /* c+up: N -> N+1, 2Sz -> 2Sz+1, with the Jordan-Wigner sign of sites < op_site */
      int nout;
      long *Bo = sector_basis(cfg.sz + 1, cfg.N + 1, 1, &nout);
      EXACT_WHOLE_BASIS_Q2 W2;
      W2.Dim = malloc(sizeof(int *) * 2);
      W2.Basis = malloc(sizeof(long **) * 2);
      for (int a = 0; a < 2; a++) {
         W2.Dim[a] = malloc(sizeof(int) * 2);
         W2.Basis[a] = malloc(sizeof(long *) * 2);
      }
      W2.Dim[0][0] = n; W2.Basis[0][0] = B;
      W2.Dim[1][1] = nout; W2.Basis[1][1] = Bo;
      EXACT_WHOLE_BASIS_Q3 W3;
      W3.Dim = malloc(sizeof(int **) * 2);
      W3.Basis = malloc(sizeof(long ***) * 2);
      for (int a = 0; a < 2; a++) {
         W3.Dim[a] = malloc(sizeof(int *) * 2);
         W3.Basis[a] = malloc(sizeof(long **) * 2);
         for (int b = 0; b < 2; b++) {
            W3.Dim[a][b] = malloc(sizeof(int) * 1);
            W3.Basis[a][b] = malloc(sizeof(long *) * 1);
         }
      }
      W3.Dim[0][0][0] = n; W3.Basis[0][0][0] = B;
      W3.Dim[1][1][0] = nout; W3.Basis[1][1][0] = Bo;
      double *o2 = GET_ARRAY_DOUBLE1(nout);
      int sites[2] = {0, L - 2};
      /* one thread: EXACT_V_M_Q2.c:22,25 / Q3.c:22,25 leave the loop counter `site` out of the
         OpenMP private list, so with several threads the parity loops race on it */
      w32(5);
      for (int k = 0; k < 2; k++) {
         EXACT_V_M_Q2(ops[H_CUPD], 1, 1, vec, 0, 0, o2, "Yes", n_ele, d, sites[k], 1, &W2);
         w32(2); w32(H_CUPD); w32(sites[k]); w32(1); w32(nout);
         for (int i = 0; i < nout; i++) w64(Bo[i]);
         fwrite(o2, 8, (size_t)nout, F);
      }
      EXACT_V_M_Q2(ops[H_CUPD], 1, 1, vec, 0, 0, o2, "No", n_ele, d, L / 2, 1, &W2);
      w32(2); w32(H_CUPD); w32(L / 2); w32(0); w32(nout);
      for (int i = 0; i < nout; i++) w64(Bo[i]);
      fwrite(o2, 8, (size_t)nout, F);
      for (int k = 0; k < 2; k++) {
         EXACT_V_M_Q3(ops[H_CUPD], 1, 1, 0, vec, 0, 0, 0, o2, k == 0 ? "Yes" : "No", n_ele, d, L / 2 + k,
                      1, &W3);
         w32(3); w32(H_CUPD); w32(L / 2 + k); w32(k == 0); w32(nout);
         for (int i = 0; i < nout; i++) w64(Bo[i]);
         fwrite(o2, 8, (size_t)nout, F);
      }
   }
   fclose(F);
   printf("n=%d nnz=%ld E=%.17g err=%.3e\n", n, H->Row[n], I.Value[0], I.Error[0]);
   return 0;
}
