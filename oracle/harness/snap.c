/*
 * Snapshot harness for the compiled reference (TEST INFRASTRUCTURE ONLY).
 *
 * This file is linked with the reference's own DMRG program sources
 * (main/dmrg/<model>_vf/src/** minus main.c) and with oracle/_ref/libdmrgref.so,
 * which is built from the reference's sml/, dmrg/, onsite_op/ and exact/ sources.
 * It replaces the program's hard-coded main() (e.g. main/dmrg/xxz_vf/src/main/main.c:11-61)
 * with one that takes the parameters from argv, and uses the linker's --wrap to
 * observe the superblock step without touching reference sources:
 *
 *   __wrap_time                          fixed value, so srand(time(NULL)) in
 *                                        sml/LANCZOS_SLOW_SYM.c:64 is reproducible
 *   __wrap_DMRG_MAKE_ELEM_*_LLLRRRRL     record the model's ordered term list
 *                                        (main/dmrg/<model>/src/main/MAKE_ELEMENT_HAM_LLLRRRRL.c)
 *   __wrap_GET_HAM_LLLRRRRL              dump basis + block operators + term list,
 *                                        and the assembled lower-triangle CRS
 *   __wrap_DMRG_DIAGONALIZE_SUPERBLOCK_SYM  log E, gs_error per step; dump GS vector
 *   __wrap_DIAGONALIZE_SUPERBLOCK (XXZ)     the same log line on the matrix-free path
 *
 * Build: oracle/Makefile (outputs into oracle/_ref/ only).  Compile with one of
 * -DSNAP_XXZ, -DSNAP_HUBBARD, -DSNAP_KLM and -I main/dmrg/<model>/include.
 *
 * Output file formats are documented in oracle/snapshot_format.md.
 */
#include "Header.h"
#include <stdint.h>

#if defined(SNAP_XXZ)
typedef MODEL_1DXXZ_VF MODEL_T;
#elif defined(SNAP_HUBBARD)
typedef MODEL_1DHUBBARD_VF MODEL_T;
#elif defined(SNAP_KLM)
typedef MODEL_1DKLM_VF MODEL_T;
#else
#error "define SNAP_XXZ, SNAP_HUBBARD or SNAP_KLM"
#endif

/* ------------------------------------------------------------------ options */
static const char *g_outdir = ".";
static int g_dump_every = 0;        /* dump a snapshot at every step whose index is listed */
static int g_dump_steps[64];
static int g_n_dump_steps = 0;
static int g_dump_ham = 1;          /* include the assembled CRS in the snapshot          */
static int g_dump_vec = 1;          /* write the GS vector                                */
static int g_dump_qn = 0;           /* write the basis quantum numbers (qn_<step>.bin)    */
static int g_dump_snap = 1;         /* write the superblock snapshot (snap_<step>.bin)    */
static int g_stop_after = -1;       /* exit(0) after this step has been dumped            */
static long g_time_value = 12345;   /* value returned by time()                           */
static FILE *g_energy_log = NULL;

time_t __wrap_time(time_t *t) {
   if (t) *t = (time_t)g_time_value;
   return (time_t)g_time_value;
}

static int step_selected(int step) {
   if (g_dump_every) return 1;
   for (int i = 0; i < g_n_dump_steps; i++) if (g_dump_steps[i] == step) return 1;
   return 0;
}

/* ------------------------------------------------------------ term recorder */
enum { K_LL = 0, K_LR, K_RR, K_RL, K_LLLR, K_LLRL, K_LLRR, K_LRRL, K_LRRR, K_RRRL, K_ZERO };
enum { ROLE_LL = 0, ROLE_RR = 1, ROLE_ON = 2 };

typedef struct {
   int kind;
   CRS1 *op1, *op2;
   int *ele[3];
   int ele_role[3];
   int sign_mode;   /* 0 none, 1 "A_B" order (even parity -> -1), 2 "B_A" order (even -> +1) */
   double coeff;
} REC_TERM;

static int g_recording = 0;
static REC_TERM g_terms[256];
static int g_n_terms = 0;

static int sign_mode_of(const char *flag, const char *type, const char *first, const char *second) {
   if (flag == NULL || strcmp(flag, "Yes") != 0) return 0;
   if (type && strcmp(type, first) == 0) return 1;
   if (type && strcmp(type, second) == 0) return 2;
   fprintf(stderr, "snap: unknown sign type %s\n", type ? type : "(null)");
   exit(2);
}

static void rec(int kind, CRS1 *a, CRS1 *b, double c, int mode,
                int *e0, int r0, int *e1, int r1, int *e2, int r2) {
   if (g_n_terms >= 256) { fprintf(stderr, "snap: too many terms\n"); exit(2); }
   REC_TERM *t = &g_terms[g_n_terms++];
   t->kind = kind; t->op1 = a; t->op2 = b; t->coeff = c; t->sign_mode = mode;
   t->ele[0] = e0; t->ele_role[0] = r0;
   t->ele[1] = e1; t->ele_role[1] = r1;
   t->ele[2] = e2; t->ele_role[2] = r2;
}

#define ONE_OP_WRAP(NAME, KIND)                                                          \
   void __real_##NAME(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, double, int *); \
   void __wrap_##NAME(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M,          \
                      double c, int *en) {                                               \
      if (g_recording) { rec(KIND, M, NULL, c, 0, NULL, 0, NULL, 0, NULL, 0); return; }  \
      __real_##NAME(B, A, Inv, M, c, en);                                                \
   }
ONE_OP_WRAP(DMRG_MAKE_ELEM_LL_LLLRRRRL, K_LL)
ONE_OP_WRAP(DMRG_MAKE_ELEM_LR_LLLRRRRL, K_LR)
ONE_OP_WRAP(DMRG_MAKE_ELEM_RR_LLLRRRRL, K_RR)
ONE_OP_WRAP(DMRG_MAKE_ELEM_RL_LLLRRRRL, K_RL)

void __real_DMRG_MAKE_ELEM_LLLR_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, CRS1 *, double, int *, int *, char[], char[]);
void __wrap_DMRG_MAKE_ELEM_LLLR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LL, CRS1 *M_LR, double c, int *LL_Ele, int *en, char Type[], char Sign_Flag[]) {
   if (g_recording) {
      rec(K_LLLR, M_LL, M_LR, c, sign_mode_of(Sign_Flag, Type, "LL_LR", "LR_LL"), LL_Ele, ROLE_LL, NULL, 0, NULL, 0);
      return;
   }
   __real_DMRG_MAKE_ELEM_LLLR_LLLRRRRL(B, A, Inv, M_LL, M_LR, c, LL_Ele, en, Type, Sign_Flag);
}
void __real_DMRG_MAKE_ELEM_LLRL_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, CRS1 *, double, int *, int *, int *, int *, char[], char[]);
void __wrap_DMRG_MAKE_ELEM_LLRL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LL, CRS1 *M_RL, double c, int *Ele_LL, int *Ele_RR, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
   if (g_recording) {
      rec(K_LLRL, M_LL, M_RL, c, sign_mode_of(Sign_Flag, Type, "LL_RL", "RL_LL"), Ele_LL, ROLE_LL, Ele_RR, ROLE_RR, Ele_LR, ROLE_ON);
      return;
   }
   __real_DMRG_MAKE_ELEM_LLRL_LLLRRRRL(B, A, Inv, M_LL, M_RL, c, Ele_LL, Ele_RR, Ele_LR, en, Type, Sign_Flag);
}
void __real_DMRG_MAKE_ELEM_LLRR_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, CRS1 *, double, int *, int *, int *, char[], char[]);
void __wrap_DMRG_MAKE_ELEM_LLRR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LL, CRS1 *M_RR, double c, int *Ele_LL, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
   if (g_recording) {
      rec(K_LLRR, M_LL, M_RR, c, sign_mode_of(Sign_Flag, Type, "LL_RR", "RR_LL"), Ele_LL, ROLE_LL, Ele_LR, ROLE_ON, NULL, 0);
      return;
   }
   __real_DMRG_MAKE_ELEM_LLRR_LLLRRRRL(B, A, Inv, M_LL, M_RR, c, Ele_LL, Ele_LR, en, Type, Sign_Flag);
}
void __real_DMRG_MAKE_ELEM_LRRL_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, CRS1 *, double, int *, int *, int *, char[], char[]);
void __wrap_DMRG_MAKE_ELEM_LRRL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LR, CRS1 *M_RL, double c, int *Ele_RR, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
   if (g_recording) {
      rec(K_LRRL, M_LR, M_RL, c, sign_mode_of(Sign_Flag, Type, "LR_RL", "RL_LR"), Ele_RR, ROLE_RR, Ele_LR, ROLE_ON, NULL, 0);
      return;
   }
   __real_DMRG_MAKE_ELEM_LRRL_LLLRRRRL(B, A, Inv, M_LR, M_RL, c, Ele_RR, Ele_LR, en, Type, Sign_Flag);
}
void __real_DMRG_MAKE_ELEM_LRRR_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, CRS1 *, double, int *, int *, char[], char[]);
void __wrap_DMRG_MAKE_ELEM_LRRR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LR, CRS1 *M_RR, double c, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
   if (g_recording) {
      rec(K_LRRR, M_LR, M_RR, c, sign_mode_of(Sign_Flag, Type, "LR_RR", "RR_LR"), Ele_LR, ROLE_ON, NULL, 0, NULL, 0);
      return;
   }
   __real_DMRG_MAKE_ELEM_LRRR_LLLRRRRL(B, A, Inv, M_LR, M_RR, c, Ele_LR, en, Type, Sign_Flag);
}
void __real_DMRG_MAKE_ELEM_RRRL_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, CRS1 *, CRS1 *, double, int *, int *, char[], char[]);
void __wrap_DMRG_MAKE_ELEM_RRRL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_RR, CRS1 *M_RL, double c, int *Ele_RR, int *en, char Type[], char Sign_Flag[]) {
   if (g_recording) {
      rec(K_RRRL, M_RR, M_RL, c, sign_mode_of(Sign_Flag, Type, "RR_RL", "RL_RR"), Ele_RR, ROLE_RR, NULL, 0, NULL, 0);
      return;
   }
   __real_DMRG_MAKE_ELEM_RRRL_LLLRRRRL(B, A, Inv, M_RR, M_RL, c, Ele_RR, en, Type, Sign_Flag);
}
void __real_DMRG_MAKE_ELEM_ZERO_LLLRRRRL(DMRG_BASIS_ONSITE *, DMRG_A_BASIS *, int *, int *);
void __wrap_DMRG_MAKE_ELEM_ZERO_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, int *en) {
   if (g_recording) { rec(K_ZERO, NULL, NULL, 0.0, 0, NULL, 0, NULL, 0, NULL, 0); return; }
   __real_DMRG_MAKE_ELEM_ZERO_LLLRRRRL(B, A, Inv, en);
}

/* ------------------------------------------------------------------ writers */
static void w_i32(FILE *f, int32_t v) { fwrite(&v, 4, 1, f); }
static void w_i64(FILE *f, int64_t v) { fwrite(&v, 8, 1, f); }
static void w_f64(FILE *f, double v) { fwrite(&v, 8, 1, f); }

static void w_crs(FILE *f, const CRS1 *M) {
   int64_t nnz = M->Row[M->row_dim];
   w_i32(f, M->row_dim);
   w_i32(f, M->col_dim);
   w_i64(f, nnz);
   fwrite(M->Row, 8, (size_t)M->row_dim + 1, f);
   fwrite(M->Col, 4, (size_t)nnz, f);
   fwrite(M->Val, 8, (size_t)nnz, f);
}

static int find_ptr(void **list, int n, void *p) {
   for (int i = 0; i < n; i++) if (list[i] == p) return i;
   return -1;
}

/* --------------------------------------------- superblock basis inputs (QN) */
/* Quantum numbers GET_BASIS_SUPERBLOCK (main/dmrg/<model>_vf/src/allocations/
 * GET_BASIS_SUPERBLOCK.c) builds the basis from, and the basis it built:
 *   "DMRGQN01", n_qn (1: S^z only, XXZ; 2: S^z and electrons), dim_LL, dim_RR, d,
 *   target_sz, target_ele, Sz_LL[dim_LL], Ele_LL[dim_LL], Sz_RR[dim_RR], Ele_RR[dim_RR],
 *   Sz_site[d], Ele_site[d] (int32; Ele_* zero when n_qn == 1), n, LL[n], LR[n], RR[n],
 *   RL[n] (int16). */
static void dump_qn(int step, DMRG_BASIS *B, BLOCK *System, BLOCK *Enviro, MODEL_T *Model, DMRG_STATUS *St) {
   char path[4096];
   snprintf(path, sizeof(path), "%s/qn_%04d.bin", g_outdir, step);
   FILE *f = fopen(path, "wb");
   if (!f) { perror(path); exit(2); }
   const int LLs = St->LL_site, RRs = St->RR_site, d = B->dim_onsite;
#if defined(SNAP_XXZ)
   const int n_qn = 1;
#else
   const int n_qn = 2;
#endif
   fwrite("DMRGQN01", 1, 8, f);
   w_i32(f, n_qn);
   w_i32(f, B->dim_LL);
   w_i32(f, B->dim_RR);
   w_i32(f, d);
   w_i32(f, B->tot_sz_LLLRRRRL);
   w_i32(f, n_qn == 2 ? B->tot_ele_LLLRRRRL : 0);
   for (int i = 0; i < B->dim_LL; i++) w_i32(f, System->Tot_Sz[LLs][i]);
   for (int i = 0; i < B->dim_LL; i++) {
#if defined(SNAP_XXZ)
      w_i32(f, 0);
#else
      w_i32(f, System->Tot_Ele[LLs][i]);
#endif
   }
   for (int i = 0; i < B->dim_RR; i++) w_i32(f, Enviro->Tot_Sz[RRs][i]);
   for (int i = 0; i < B->dim_RR; i++) {
#if defined(SNAP_XXZ)
      w_i32(f, 0);
#else
      w_i32(f, Enviro->Tot_Ele[RRs][i]);
#endif
   }
   for (int s = 0; s < d; s++) {
#if defined(SNAP_XXZ)
      w_i32(f, FIND_SITE_SZ(s, Model->spin));
#elif defined(SNAP_HUBBARD)
      w_i32(f, ONSITE_FIND_SITE_SZ_SZBASIS_HUBBARD(s));
#else
      w_i32(f, ONSITE_FIND_SITE_SZ_SZBASIS_KLM(s, Model->spin_loc));
#endif
   }
   for (int s = 0; s < d; s++) {
#if defined(SNAP_XXZ)
      w_i32(f, 0);
#elif defined(SNAP_HUBBARD)
      w_i32(f, ONSITE_FIND_SITE_ELE_SZBASIS_HUBBARD(s));
#else
      w_i32(f, ONSITE_FIND_SITE_ELE_SZBASIS_KLM(s, Model->spin_loc));
#endif
   }
   const int n = B->dim_LLLRRRRL;
   w_i32(f, n);
   fwrite(B->LL_LLLRRRRL, 2, (size_t)n, f);
   fwrite(B->LR_LLLRRRRL, 2, (size_t)n, f);
   fwrite(B->RR_LLLRRRRL, 2, (size_t)n, f);
   fwrite(B->RL_LLLRRRRL, 2, (size_t)n, f);
   fclose(f);
}

/* ------------------------------------------------------------ GET_HAM hook */
CRS1 *__real_GET_HAM_LLLRRRRL(DMRG_BASIS *, BLOCK *, BLOCK *, MODEL_T *, DMRG_STATUS *, DMRG_TIME *);
CRS1 *__wrap_GET_HAM_LLLRRRRL(DMRG_BASIS *Dmrg_Basis, BLOCK *System, BLOCK *Enviro, MODEL_T *Model, DMRG_STATUS *Dmrg_Status, DMRG_TIME *Dmrg_Time) {
   int step = Dmrg_Status->tot_iter_now;
   FILE *f = NULL;
   if (step_selected(step) && g_dump_qn) dump_qn(step, Dmrg_Basis, System, Enviro, Model, Dmrg_Status);
   if (step_selected(step) && g_dump_snap) {
      char path[4096];
      snprintf(path, sizeof(path), "%s/snap_%04d.bin", g_outdir, step);
      f = fopen(path, "wb");
      if (!f) { perror(path); exit(2); }

      /* Record the ordered term list by running the model's row generator once. */
      int p_save = Model->p_threads;
      Model->p_threads = 1;
      HAM_BOX **Box = GET_HAM_BOX(System, Enviro, Model, Dmrg_Status->LL_site, Dmrg_Status->RR_site);
      DMRG_A_BASIS dummyA;
      memset(&dummyA, 0, sizeof(dummyA));
      DMRG_BASIS_ONSITE dummyB;
      memset(&dummyB, 0, sizeof(dummyB));
      dummyB.dim_RR = Dmrg_Basis->dim_RR;
      dummyB.dim_onsite = Dmrg_Basis->dim_onsite;
      g_n_terms = 0;
      g_recording = 1;
      MAKE_ELEMENT_HAM_LLLRRRRL(&dummyB, &dummyA, Dmrg_Basis->Inv_LLLRRRRL, Box[0]);
      g_recording = 0;

      int n = Dmrg_Basis->dim_LLLRRRRL;
      fwrite("DMRGSNP1", 1, 8, f);
      w_i32(f, n);
      w_i32(f, Dmrg_Basis->dim_LL);
      w_i32(f, Dmrg_Basis->dim_RR);
      w_i32(f, Dmrg_Basis->dim_onsite);
      w_i32(f, Dmrg_Status->LL_site);
      w_i32(f, Dmrg_Status->RR_site);
      w_i32(f, step);
      fwrite(Dmrg_Basis->LL_LLLRRRRL, 2, (size_t)n, f);
      fwrite(Dmrg_Basis->LR_LLLRRRRL, 2, (size_t)n, f);
      fwrite(Dmrg_Basis->RR_LLLRRRRL, 2, (size_t)n, f);
      fwrite(Dmrg_Basis->RL_LLLRRRRL, 2, (size_t)n, f);

      /* distinct operators and parity arrays, in first-use order */
      void *ops[512]; int n_ops = 0;
      void *eles[64]; int ele_role[64]; int n_eles = 0;
      int t_op[256][2], t_ele[256][3];
      for (int t = 0; t < g_n_terms; t++) {
         CRS1 *o[2] = {g_terms[t].op1, g_terms[t].op2};
         for (int k = 0; k < 2; k++) {
            if (!o[k]) { t_op[t][k] = -1; continue; }
            int id = find_ptr(ops, n_ops, o[k]);
            if (id < 0) { id = n_ops; ops[n_ops++] = o[k]; }
            t_op[t][k] = id;
         }
         for (int k = 0; k < 3; k++) {
            if (!g_terms[t].ele[k]) { t_ele[t][k] = -1; continue; }
            int id = find_ptr(eles, n_eles, g_terms[t].ele[k]);
            if (id < 0) { id = n_eles; eles[n_eles] = g_terms[t].ele[k]; ele_role[n_eles] = g_terms[t].ele_role[k]; n_eles++; }
            t_ele[t][k] = id;
         }
      }
      w_i32(f, n_ops);
      for (int i = 0; i < n_ops; i++) w_crs(f, (CRS1 *)ops[i]);
      w_i32(f, n_eles);
      for (int i = 0; i < n_eles; i++) {
         int len = ele_role[i] == ROLE_LL ? Dmrg_Basis->dim_LL : ele_role[i] == ROLE_RR ? Dmrg_Basis->dim_RR : Dmrg_Basis->dim_onsite;
         w_i32(f, len);
         fwrite(eles[i], 4, (size_t)len, f);
      }
      w_i32(f, g_n_terms);
      for (int t = 0; t < g_n_terms; t++) {
         w_i32(f, g_terms[t].kind);
         w_i32(f, t_op[t][0]);
         w_i32(f, t_op[t][1]);
         w_i32(f, t_ele[t][0]);
         w_i32(f, t_ele[t][1]);
         w_i32(f, t_ele[t][2]);
         w_i32(f, g_terms[t].sign_mode);
         w_f64(f, g_terms[t].coeff);
      }
      FREE_HAM_BOX(Box, Model);
      Model->p_threads = p_save;
   }

   CRS1 *Ham = __real_GET_HAM_LLLRRRRL(Dmrg_Basis, System, Enviro, Model, Dmrg_Status, Dmrg_Time);

   if (f) {
      w_i32(f, g_dump_ham);
      if (g_dump_ham) w_crs(f, Ham);
#if defined(SNAP_XXZ)
      /* matrix-free inputs of main/dmrg/xxz_vf/src/allocations/DIAGONALIZE_SUPERBLOCK.c:15-17 */
      w_i32(f, 1);
      CRS1 *H_LLLR = GET_HAM_LLLR(System, Enviro, Dmrg_Basis, Dmrg_Status, Model);
      CRS1 *H_LRRL = GET_HAM_LRRL(System, Enviro, Dmrg_Basis, Dmrg_Status, Model);
      CRS1 *H_RRRL = GET_HAM_RRRL(System, Enviro, Dmrg_Basis, Dmrg_Status, Model);
      w_i32(f, Dmrg_Basis->dim_LLLR);
      fwrite(Dmrg_Basis->LL_LLLR, 2, (size_t)Dmrg_Basis->dim_LLLR, f);
      fwrite(Dmrg_Basis->LR_LLLR, 2, (size_t)Dmrg_Basis->dim_LLLR, f);
      w_i32(f, Dmrg_Basis->dim_LRRL);
      fwrite(Dmrg_Basis->LR_LRRL, 2, (size_t)Dmrg_Basis->dim_LRRL, f);
      fwrite(Dmrg_Basis->RL_LRRL, 2, (size_t)Dmrg_Basis->dim_LRRL, f);
      w_i32(f, Dmrg_Basis->dim_RRRL);
      fwrite(Dmrg_Basis->RR_RRRL, 2, (size_t)Dmrg_Basis->dim_RRRL, f);
      fwrite(Dmrg_Basis->RL_RRRL, 2, (size_t)Dmrg_Basis->dim_RRRL, f);
      w_crs(f, H_LLLR);
      w_crs(f, H_LRRL);
      w_crs(f, H_RRRL);
      FREE_CRS1(H_LLLR);
      FREE_CRS1(H_LRRL);
      FREE_CRS1(H_RRRL);
#else
      w_i32(f, 0);
#endif
      fclose(f);
   }
   return Ham;
}

/* --------------------------------------------------------- eigensolve hook */
void __real_DMRG_DIAGONALIZE_SUPERBLOCK_SYM(CRS1 *, DMRG_TIME *, DMRG_PARAMETER *, DMRG_STATUS *, int);
void __wrap_DMRG_DIAGONALIZE_SUPERBLOCK_SYM(CRS1 *Ham, DMRG_TIME *Time, DMRG_PARAMETER *Param, DMRG_STATUS *St, int p_threads) {
   double t0 = omp_get_wtime();
   __real_DMRG_DIAGONALIZE_SUPERBLOCK_SYM(Ham, Time, Param, St, p_threads);
   double t1 = omp_get_wtime();
   int step = St->tot_iter_now;
   int n = Ham->row_dim;
   if (g_energy_log) {
      fprintf(g_energy_log, "%d %d %d %d %d %.17g %.6e %.6f\n", step, St->sweep_now, St->LL_site, St->RR_site, n, St->gs_val, St->gs_error, t1 - t0);
      fflush(g_energy_log);
   }
   if (step_selected(step) && g_dump_vec) {
      char path[4096];
      snprintf(path, sizeof(path), "%s/gs_%04d.bin", g_outdir, step);
      FILE *f = fopen(path, "wb");
      if (!f) { perror(path); exit(2); }
      fwrite("DMRGGSV1", 1, 8, f);
      w_i32(f, n);
      w_f64(f, St->gs_val);
      w_f64(f, St->gs_error);
      fwrite(St->GS_Vec, 8, (size_t)n, f);
      fclose(f);
   }
   if (g_stop_after >= 0 && step >= g_stop_after) {
      if (g_energy_log) fclose(g_energy_log);
      exit(0);
   }
}

#if defined(SNAP_XXZ)
/* ------------------------------------------- matrix-free eigensolve hook (XXZ) */
/* main/dmrg/xxz_vf/src/allocations/DIAGONALIZE_SUPERBLOCK.c: taken instead of
 * DMRG_DIAGONALIZE_SUPERBLOCK_SYM when the superblock dimension reaches limit=
 * (GET_GROUND_STATE.c:15-21); logs the same energy line. */
void __real_DIAGONALIZE_SUPERBLOCK(DMRG_BASIS *, BLOCK *, BLOCK *, MODEL_T *, DMRG_PARAMETER *, DMRG_STATUS *, DMRG_TIME *);
void __wrap_DIAGONALIZE_SUPERBLOCK(DMRG_BASIS *B, BLOCK *System, BLOCK *Enviro, MODEL_T *Model, DMRG_PARAMETER *P,
                                   DMRG_STATUS *St, DMRG_TIME *Time) {
   double t0 = omp_get_wtime();
   __real_DIAGONALIZE_SUPERBLOCK(B, System, Enviro, Model, P, St, Time);
   double t1 = omp_get_wtime();
   if (g_energy_log) {
      fprintf(g_energy_log, "%d %d %d %d %d %.17g %.6e %.6f\n", St->tot_iter_now, St->sweep_now, St->LL_site, St->RR_site,
              B->dim_LLLRRRRL, St->gs_val, St->gs_error, t1 - t0);
      fflush(g_energy_log);
   }
}
#endif

/* ----------------------------------------------------- expectation values */
/* dmrg/DMRG_EXPECTATION_ONSITE.c and DMRG_EXPECTATION_INTERSITE_Q0.c: dump inputs and
 * the reference's outputs of every call (file expect_<step>_<k>.bin). */
static int g_expect_calls = 0;

static void dump_expect(int kind, CRS1 **M_CF, CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, int origin,
                        double *Vec, DMRG_BASIS *B, DMRG_STATUS *St) {
   char path[4096];
   snprintf(path, sizeof(path), "%s/expect_%04d_%02d.bin", g_outdir, St->tot_iter_now, g_expect_calls++);
   FILE *f = fopen(path, "wb");
   if (!f) { perror(path); exit(2); }
   int n = St->dim_LLLRRRRL, LL_site = St->LL_site, RR_site = St->RR_site;
   fwrite("DMRGEXP1", 1, 8, f);
   w_i32(f, kind);
   w_i32(f, n);
   w_i32(f, B->dim_LL);
   w_i32(f, B->dim_RR);
   w_i32(f, B->dim_onsite);
   w_i32(f, LL_site);
   w_i32(f, RR_site);
   w_i32(f, origin);
   fwrite(B->LL_LLLRRRRL, 2, (size_t)n, f);
   fwrite(B->LR_LLLRRRRL, 2, (size_t)n, f);
   fwrite(B->RR_LLLRRRRL, 2, (size_t)n, f);
   fwrite(B->RL_LLLRRRRL, 2, (size_t)n, f);
   w_i32(f, LL_site + 1);
   for (int s = 0; s <= LL_site; s++) w_crs(f, M_LL[s]);
   w_crs(f, M_On);
   w_i32(f, RR_site + 1);
   for (int s = 0; s <= RR_site; s++) w_crs(f, M_RR[s]);
   int ncf = (kind == 1 && origin <= LL_site) ? LL_site - origin + 1 : 0;
   w_i32(f, ncf);
   for (int r = 0; r < ncf; r++) w_crs(f, M_CF[r]);
   fwrite(Vec, 8, (size_t)n, f);
   int nout = (kind == 0) ? LL_site + RR_site + 4 : LL_site + RR_site + 4 - origin;
   w_i32(f, nout);
   fwrite(Out, 8, (size_t)nout, f);
   fclose(f);
}

void __real_DMRG_EXPECTATION_ONSITE(CRS1 **, CRS1 *, CRS1 **, double *, double *, double *, int, DMRG_BASIS *, DMRG_STATUS *);
void __wrap_DMRG_EXPECTATION_ONSITE(CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, double *Vec, double *Temp_V,
                                    int p_threads, DMRG_BASIS *B, DMRG_STATUS *St) {
   __real_DMRG_EXPECTATION_ONSITE(M_LL, M_On, M_RR, Out, Vec, Temp_V, p_threads, B, St);
   if (step_selected(St->tot_iter_now)) dump_expect(0, NULL, M_LL, M_On, M_RR, Out, 0, Vec, B, St);
}

void __real_DMRG_EXPECTATION_INTERSITE_Q0(CRS1 **, CRS1 **, CRS1 *, CRS1 **, double *, int, double *, double *, double *, int, DMRG_BASIS *, DMRG_STATUS *);
void __wrap_DMRG_EXPECTATION_INTERSITE_Q0(CRS1 **M_CF, CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, int origin,
                                          double *Vec, double *T1, double *T2, int p_threads, DMRG_BASIS *B,
                                          DMRG_STATUS *St) {
   __real_DMRG_EXPECTATION_INTERSITE_Q0(M_CF, M_LL, M_On, M_RR, Out, origin, Vec, T1, T2, p_threads, B, St);
   if (step_selected(St->tot_iter_now) && origin <= St->LL_site + 1)
      dump_expect(1, M_CF, M_LL, M_On, M_RR, Out, origin, Vec, B, St);
}

/* ---------------------------------------------------------------------- main */
static double argd(int argc, char **argv, const char *key, double def) {
   size_t k = strlen(key);
   for (int i = 1; i < argc; i++) if (strncmp(argv[i], key, k) == 0 && argv[i][k] == '=') return atof(argv[i] + k + 1);
   return def;
}
static const char *args(int argc, char **argv, const char *key, const char *def) {
   size_t k = strlen(key);
   for (int i = 1; i < argc; i++) if (strncmp(argv[i], key, k) == 0 && argv[i][k] == '=') return argv[i] + k + 1;
   return def;
}

int main(int argc, char **argv) {
   DMRG_PARAMETER *P = calloc(1, sizeof(*P));
   MODEL_T *M = calloc(1, sizeof(*M));

   g_outdir = args(argc, argv, "out", ".");
   g_time_value = (long)argd(argc, argv, "seed", 12345);
   g_dump_ham = (int)argd(argc, argv, "dump_ham", 1);
   g_dump_vec = (int)argd(argc, argv, "dump_vec", 1);
   g_dump_qn = (int)argd(argc, argv, "dump_qn", 0);
   g_dump_snap = (int)argd(argc, argv, "dump_snap", 1);
   g_stop_after = (int)argd(argc, argv, "stop_after", -1);
   const char *steps = args(argc, argv, "dump", "");
   if (strcmp(steps, "all") == 0) g_dump_every = 1;
   else {
      char buf[1024];
      snprintf(buf, sizeof(buf), "%s", steps);
      for (char *tok = strtok(buf, ","); tok && g_n_dump_steps < 64; tok = strtok(NULL, ","))
         g_dump_steps[g_n_dump_steps++] = atoi(tok);
   }
   char logpath[4096];
   snprintf(logpath, sizeof(logpath), "%s/energies.txt", g_outdir);
   g_energy_log = fopen(logpath, "w");

   strcpy(P->Initial_Guess, "No");
   strcpy(P->Enviro_Copy, "Yes");
   strcpy(P->Lan_Con, args(argc, argv, "lan_con", "Normal"));
   strcpy(P->II_Type, "CG");
   P->sweep = (int)argd(argc, argv, "sweep", 3);
   /* superblocks at or above this dimension take the matrix-free path (OBC/SSD),
    * main/dmrg/xxz_vf/src/allocations/GET_GROUND_STATE.c:15-21 */
   P->dim_LLLRRRRL_limit = (int)argd(argc, argv, "limit", (double)INT_MAX);
   P->max_dim_system = (int)argd(argc, argv, "m", 64);
   P->sp_LL = argd(argc, argv, "sp_LL", 0.5);
   P->param_iter = 1;
   P->param_iter_now = 0;
   P->diag_acc = argd(argc, argv, "diag_acc", 1e-14);
   P->diag_min_step = 1;
   P->diag_max_step = (int)argd(argc, argv, "diag_max_step", 800);
   P->inv_iter_acc = argd(argc, argv, "ii_acc", 1e-9);
   P->inv_iter_diag_add = argd(argc, argv, "ii_diag_add", 1e-11);
   P->inv_iter_max_step = (int)argd(argc, argv, "ii_max_step", 3);

   strcpy(M->BC, args(argc, argv, "bc", "OBC"));
   M->p_threads = (int)argd(argc, argv, "threads", 8);
   M->tot_site = (int)argd(argc, argv, "L", 16);
   M->tot_sz = (int)argd(argc, argv, "tot_sz", 0);
   M->cf_origin = 0;
#if defined(SNAP_XXZ)
   strcpy(P->Diag_Method, args(argc, argv, "method", "Lanczos_Slow"));
   P->cg_acc = argd(argc, argv, "cg_acc", 1e-6);
   P->cg_max_step = (int)argd(argc, argv, "cg_max_step", 1000);
   M->spin = (int)argd(argc, argv, "spin", 1);
   M->J_xy = argd(argc, argv, "J_xy", 1.0);
   M->J_z = argd(argc, argv, "J_z", 1.0);
   M->D_z = argd(argc, argv, "D_z", 0.0);
   M->h_z = argd(argc, argv, "h_z", 0.0);
#elif defined(SNAP_HUBBARD)
   strcpy(P->Diag_Method, args(argc, argv, "method", "Lanczos"));
   P->cg_acc = argd(argc, argv, "cg_acc", 1e-7);
   P->cg_max_step = (int)argd(argc, argv, "cg_max_step", 2000);
   M->tot_ele = (int)argd(argc, argv, "N", M->tot_site);
   M->t1 = argd(argc, argv, "t1", -1.0);
   M->t2 = argd(argc, argv, "t2", 0.0);
   M->U = argd(argc, argv, "U", 4.0);
   M->V = argd(argc, argv, "V", 0.0);
   M->h_z = argd(argc, argv, "h_z", 0.0);
   M->mu = argd(argc, argv, "mu", 0.0);
#elif defined(SNAP_KLM)
   strcpy(P->Diag_Method, args(argc, argv, "method", "Lanczos"));
   P->cg_acc = argd(argc, argv, "cg_acc", 1e-7);
   P->cg_max_step = (int)argd(argc, argv, "cg_max_step", 2000);
   M->tot_ele = (int)argd(argc, argv, "N", M->tot_site);
   M->spin_loc = (int)argd(argc, argv, "spin_loc", 1);
   M->t = argd(argc, argv, "t", -1.0);
   M->J = argd(argc, argv, "J", 2.0);
   M->D_z = argd(argc, argv, "D_z", 0.0);
   M->I_xy = argd(argc, argv, "I_xy", 0.0);
   M->I_z = argd(argc, argv, "I_z", 0.0);
   M->h_z = argd(argc, argv, "h_z", 0.0);
   M->mu = argd(argc, argv, "mu", 0.0);
#endif
   DMRG(M, P);
   if (g_energy_log) fclose(g_energy_log);
   free(P);
   free(M);
   return 0;
}
