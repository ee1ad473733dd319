/*
 * dmrg_c_amd.h -- C ABI of libdmrg_c_amd.so, the MI355X (gfx950) superblock eigensolve.
 *
 * Part 1 declares the reference's own entry points for the hot path, with the
 * reference's signatures and struct layouts, so a dmrg_C program links against this
 * library instead of libsml.a / libdmrg.a / libexact.a for these symbols and runs the
 * eigensolve on the GPU.  Each declaration cites the reference prototype it replaces.
 *
 * Part 2 declares new entry points (prefix dmrgx_) that return status codes instead of
 * calling exit(), take device-resident data, and expose the forms the reference lacks
 * (device assembly from a recorded term list, precision / batching options).
 *
 * Struct layouts are field-for-field those of include/SML.h, include/dmrg.h and
 * include/exact.h so the same objects can be passed.  If the reference headers were
 * included first, their definitions are used.
 */
#ifndef DMRG_C_AMD_H
#define DMRG_C_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== types */
#ifndef SML_h
/* include/SML.h:17-28 */
typedef struct {
   long max_val;
   int max_row;
   int row_dim;
   int col_dim;
   int *Col;
   long *Row;
   double *Val;
} CRS1;

/* include/SML.h:56-65 */
typedef struct {
   CRS1 *M;
   double acc;
   double *Vec;
   double *Out_Vec;
   int max_step;
   int p_threads;
} BOX_CG;

/* include/SML.h:67-86 */
typedef struct {
   CRS1 *M;
   double ii_acc;
   double ii_diag_add;
   int ii_max_step;
   double cg_acc;
   int cg_max_step;
   double error;
   double *Eig_Vec;
   double *Eig_Val;
   int p_threads;
   char Type[100];
} BOX_II;

/* include/SML.h:122-135 */
typedef struct {
   CRS1 *M;
   double *Eig_Vec;
   double *eig_val;
   double acc;
   int min_step;
   int max_step;
   int max_block_step;
   int p_threads;
   char Guess[100];
   char Type[100];
} BOX_LAN;

/* include/SML.h:43-54 */
typedef struct {
   long max_val;
   int max_col;
   int row_dim;
   int col_dim;
   long *Col;
   int *Row;
   double *Val;
} CCS1;
#endif /* SML_h */

#ifndef dmrg_h
/* include/dmrg.h:6-38 */
typedef struct {
   char Initial_Guess[1000];
   char Enviro_Copy[1000];
   char II_Type[1000];
   char Lan_Con[1000];
   int max_dim_system;
   int sweep;
   char Diag_Method[1000];
   double diag_acc;
   int diag_min_step;
   int diag_max_step;
   double cg_acc;
   int cg_max_step;
   double inv_iter_acc;
   double inv_iter_diag_add;
   int inv_iter_max_step;
   double sp_LL;
   int param_iter;
   int param_iter_now;
   int dim_LLLRRRRL_limit;
} DMRG_PARAMETER;

/* include/dmrg.h:40-51 */
typedef struct {
   double diag;
   double inv_iter;
   double make_ham;
   double make_basis;
   double make_dens_mat;
   double trans_main;
   double trans_exp;
   double total;
} DMRG_TIME;

/* include/dmrg.h:54-99 */
typedef struct {
   short *LL_LLLRRRRL;
   short *LR_LLLRRRRL;
   short *RL_LLLRRRRL;
   short *RR_LLLRRRRL;
   int *Inv_LLLRRRRL;
   int dim_LLLRRRRL;
   int dim_LL;
   int dim_RR;
   int dim_onsite;
   int tot_sz_LLLRRRRL;
   int tot_ele_LLLRRRRL;
   int tot_ele_1_LLLRRRRL;
   int tot_ele_2_LLLRRRRL;
   int tot_parity_LLLRRRRL;
   short *LL_LLLR;
   short *LR_LLLR;
   int **Inv_LLLR;
   int dim_LLLR;
   short *LR_LRRL;
   short *RL_LRRL;
   int **Inv_LRRL;
   int dim_LRRL;
   short *RR_RRRL;
   short *RL_RRRL;
   int **Inv_RRRL;
   int dim_RRRL;
   int *Sum_LLLR;
   int *Tot_Sz_LLLR;
   int *Tot_Ele_LLLR;
   int *Tot_Ele_1_LLLR;
   int *Tot_Ele_2_LLLR;
   int *Tot_Parity_LLLR;
} DMRG_BASIS;

/* include/dmrg.h:128-138 */
typedef struct {
   short *LL_LLLRRRRL;
   short *LR_LLLRRRRL;
   short *RL_LLLRRRRL;
   short *RR_LLLRRRRL;
   int *Inv_LLLRRRRL;
   double *Val_LLLRRRRL;
   int elem_num;
} DMRG_A_BASIS;

/* include/dmrg.h:198-208 */
typedef struct {
   int LL;
   int LR;
   int RR;
   int RL;
   int row;
   int dim_RR;
   int dim_onsite;
} DMRG_BASIS_ONSITE;

/* include/dmrg.h:236-267 */
typedef struct {
   char BC[1000];
   char Enviro_Copy[100];
   int LL_site;
   int RR_site;
   int max_dim_system;
   int dim_LLLRRRRL;
   int dim_LL;
   int dim_onsite;
   int dim_RR;
   int spin;
   int tot_sz;
   int tot_ele;
   int tot_ele_1;
   int tot_ele_2;
   int tot_parity;
   double *GS_Vec;
   double gs_val;
   double gs_error;
   double percent_LL;
   int sweep_now;
   int param_iter_now;
   int tot_iter_now;
   int sweep;
   int param_iter;
   int tot_iter;
} DMRG_STATUS;

/* include/dmrg.h:269-285 */
typedef struct {
   double *Eig_Vec;
   double *eig_val;
   double acc;
   int min_step;
   int max_step;
   int p_threads;
   char Guess[100];
   char Type[100];
   CRS1 *M_LLLR;
   CRS1 *M_LRRL;
   CRS1 *M_LRRL_Sign;
   CRS1 *M_RRRL;
   int *Ele_RR;
   DMRG_BASIS *Dmrg_Basis;
} DMRG_BOX_LAN;

/* include/dmrg.h:287-302 */
typedef struct {
   double acc;
   double diag_val;
   double *Vec;
   double *Out_Vec;
   int max_step;
   int p_threads;
   CRS1 *M_LLLR;
   CRS1 *M_LRRL;
   CRS1 *M_LRRL_Sign;
   CRS1 *M_RRRL;
   int *Ele_RR;
   DMRG_BASIS *Dmrg_Basis;
} DMRG_BOX_CG;

/* include/dmrg.h:304-326 */
typedef struct {
   double ii_acc;
   double ii_diag_add;
   int ii_max_step;
   double cg_acc;
   int cg_max_step;
   double *Eig_Vec;
   double *eig_val;
   double error;
   int p_threads;
   CRS1 *M_LLLR;
   CRS1 *M_LRRL;
   CRS1 *M_LRRL_Sign;
   CRS1 *M_RRRL;
   int *Ele_RR;
   DMRG_BASIS *Dmrg_Basis;
} DMRG_BOX_II;

/* include/dmrg.h:220-234 */
typedef struct {
   CCS1 *Trans_Matrix;
   CRS1 *Trans_Matrix_Dagger;
   double tr_error;
   double *Val_DM_Dist;
   double sum_val_dm;
   int dim_LLLR;
   int dim_renorm;
   int *Q_Number1;
   int *Q_Number2;
   int *Q_Number3;
} DMRG_SYSTEM_INFO;
#endif /* dmrg_h */

#ifndef exact_h
/* include/exact.h:14-37 */
typedef struct {
   int diag_num;
   int est_max_row_elem_num;
   char Diag_Method[1000];
   double diag_acc;
   int diag_min_step;
   int diag_max_step;
   double cg_acc;
   int cg_max_step;
   double inv_iter_acc;
   double inv_iter_diag_add;
   int inv_iter_max_step;
   int param_iter;
   int param_iter_now;
} EXACT_PARAMETER;

/* include/exact.h:39-51 */
typedef struct {
   double diag;
   double inv_iter;
   double make_ham;
   double find_dim;
   double make_basis;
   double exp_values;
   double cf;
   double sc;
   double total;
} EXACT_TIME;

/* include/exact.h:54-64 */
typedef struct {
   int dim;
   int diag_num;
   double **Vector;
   double *Value;
   double *Error;
   double mem_ham;
   CRS1 *Ham;
} EXACT_HAM_INFO;

/* include/exact.h:66-74 */
typedef struct {
   int dim;
   long *Basis;
   int tot_sz;
   int tot_ele;
   int tot_parity;
} EXACT_BASIS_INFO;

/* include/exact.h:76-87 */
typedef struct {
   long *Basis;
   long elem_num;
   long *Check;
   double *Val;
   int max_row;
   int dim_onsite;
   long dim_whole;
} EXACT_A_BASIS;

/* include/exact.h:89-111 */
typedef struct {
   int *Dim;
   long **Basis;
   int max_dim;
} EXACT_WHOLE_BASIS_Q1;
typedef struct {
   int **Dim;
   long ***Basis;
   int max_dim;
} EXACT_WHOLE_BASIS_Q2;
typedef struct {
   int ***Dim;
   long ****Basis;
   int max_dim;
} EXACT_WHOLE_BASIS_Q3;
#endif /* exact_h */

/* ========================================== Part 1: reference entry points */
/* Semantics follow the reference: void return, exit(1) with the reference's message
 * on the conditions it treats as fatal, warnings + ./SML_out/error.txt for the soft
 * ones, step counts appended to ./SML_out/<SOLVER>_Step.txt (set DMRGX_SML_OUT=0 to
 * skip the files).  Work runs on the current HIP device; p_threads is accepted and
 * ignored.  Guess == "No" starts from a counter-hash random vector seeded by
 * dmrgx_set_seed() (default 1) instead of srand(time(NULL)). */

/* include/SML.h:291 -- y = M x, M lower-triangle CRS with the diagonal stored last */
void MATRIX_VECTOR_PRODUCT_SYM(CRS1 *M, double *V1, double *Out, double **Temp, int p_threads);
/* include/SML.h:290 -- y = M x, full CRS */
void MATRIX_VECTOR_PRODUCT(CRS1 *M, double *V1, double *Out, int p_threads);
/* include/SML.h:270,272 */
void LANCZOS_SLOW_SYM(BOX_LAN *Box);
void LANCZOS_SYM(BOX_LAN *Box);
/* include/SML.h:269,271 (full CRS) */
void LANCZOS_SLOW(BOX_LAN *Box);
void LANCZOS(BOX_LAN *Box);
/* include/SML.h:266-267 */
void INVERSE_ITERATION(BOX_II *Box_II);
void INVERSE_ITERATION_SYM(BOX_II *Box_II);
/* include/SML.h:172-173 */
void CONJUGATE_GRADIENT(BOX_CG *Box);
void CONJUGATE_GRADIENT_SYM(BOX_CG *Box);
/* include/SML.h:293-294 -- MINRES on (M) x = Vec from x = Out_Vec;
 * like the reference, the iterate that first meets |eta| < acc is not copied to Out_Vec */
void MINIMUM_RESIDUAL(BOX_CG *Box);
void MINIMUM_RESIDUAL_SYM(BOX_CG *Box);
/* include/SML.h:174-175 */
double CONVERGE_CHECK(CRS1 *M, double *Eig_Vec, double eig_val, int p_threads);
double CONVERGE_CHECK_SYM(CRS1 *M, double *Eig_Vec, double eig_val, int p_threads);

/* include/dmrg.h:406 and :353 */
void DMRG_DIAGONALIZE_SUPERBLOCK_SYM(CRS1 *Ham, DMRG_TIME *Time, DMRG_PARAMETER *Param, DMRG_STATUS *Dmrg_Status, int p_threads);
void DMRG_DIAGONALIZE_SUPERBLOCK(CRS1 *Ham, DMRG_TIME *Time, DMRG_PARAMETER *Param, DMRG_STATUS *Dmrg_Status, int p_threads);

/* include/dmrg.h:401-405 -- matrix-free OBC superblock operator and its solvers */
void DMRG_MATRIX_VECTOR_PRODUCT_OBC(CRS1 *M_LLLR, CRS1 *M_LRRL, CRS1 *M_LRRL_Sign, CRS1 *M_RRRL, double *V, double *Out_V, int *Ele_RR, int p_threads, DMRG_BASIS *Dmrg_Basis);
void DMRG_LANCZOS_SLOW_OBC(DMRG_BOX_LAN *Box);
void DMRG_CONJUGATE_GRADIENT_OBC(DMRG_BOX_CG *Box);
void DMRG_INVERSE_ITERATION_OBC(DMRG_BOX_II *Box_II);
double DMRG_CONVERGE_CHECK_OBC(double eig_val, CRS1 *M_LLLR, CRS1 *M_LRRL, CRS1 *M_LRRL_Sign, CRS1 *M_RRRL, double *Eig_V, double *Temp_V, int *Ele_RR, int p_threads, DMRG_BASIS *Dmrg_Basis);

/* include/dmrg.h:367-370 -- one operator on one slot of the superblock state, y = O x */
void DMRG_V_M_LL_Q0(CRS1 *M_LL, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *Dmrg_Basis);
void DMRG_V_M_LR_Q0(CRS1 *M_LR, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *Dmrg_Basis);
void DMRG_V_M_RR_Q0(CRS1 *M_RR, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *Dmrg_Basis);
void DMRG_V_M_RL_Q0(CRS1 *M_RL, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *Dmrg_Basis);
/* include/dmrg.h:363-364 -- on-site values and correlation functions on the ground state */
void DMRG_EXPECTATION_ONSITE(CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, double *Vec, double *Temp_V,
                             int p_threads, DMRG_BASIS *Dmrg_Basis, DMRG_STATUS *Dmrg_Status);
void DMRG_EXPECTATION_INTERSITE_Q0(CRS1 **M_CF, CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, int origin,
                                   double *Vec, double *Temp_V1, double *Temp_V2, int p_threads,
                                   DMRG_BASIS *Dmrg_Basis, DMRG_STATUS *Dmrg_Status);

/* include/dmrg.h:330-340 -- superblock element makers.  Outside a recording they
 * behave as the reference (accumulate one row into A_Basis on the host).  Between
 * dmrgx_record_begin() and dmrgx_record_end() they only append to the term list. */
void DMRG_MAKE_ELEM_LL_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_LR_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_RR_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_RL_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_LLLR_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M_LL, CRS1 *M_LR, double coeef, int *LL_Ele, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_LLRL_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M_LL, CRS1 *M_RL, double coeef, int *Ele_LL, int *Ele_RR, int *Ele_LR, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_LLRR_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M_LL, CRS1 *M_RR, double coeef, int *Ele_LL, int *Ele_LR, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_LRRL_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M_LR, CRS1 *M_RL, double coeef, int *Ele_RR, int *Ele_LR, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_LRRR_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M_LR, CRS1 *M_RR, double coeef, int *Ele_LR, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_RRRL_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, CRS1 *M_RR, CRS1 *M_RL, double coeef, int *Ele_RR, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_ZERO_LLLRRRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int *Inv_LLLRRRRL, int *elem_num);

/* include/dmrg.h:342-350 (dmrg/DMRG_MAKE_ELEM_LL_LLLR.c ... DMRG_MAKE_ELEM_RRRL_RRRL.c) --
 * element makers of the sub-Hamiltonians the matrix-free path multiplies with
 * (GET_HAM_LLLR / _LRRL / _RRRL of the XXZ program): one row of the LLLR, LRRL or RRRL
 * sub-basis accumulated into A_Basis on the host, the reference's rule bit for bit. */
void DMRG_MAKE_ELEM_LL_LLLR(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_LLLR, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_LR_LLLR(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_LLLR, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_LLLR_LLLR(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_LLLR, CRS1 *M_LL, CRS1 *M_LR, double coeef, int *LL_Ele, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_LRRL_LRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_LRRL, CRS1 *M_LR, CRS1 *M_RL, double coeef, int *Ele_LR, int *elem_num, char Type[], char Sign_Flag[]);
void DMRG_MAKE_ELEM_RL_RRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_RRRL, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_RR_RRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_RRRL, CRS1 *M, double coeef, int *elem_num);
void DMRG_MAKE_ELEM_RRRL_RRRL(DMRG_BASIS_ONSITE *Basis, DMRG_A_BASIS *A_Basis, int **Inv_RRRL, CRS1 *M_RR, CRS1 *M_RL, double coeef, int *Ele_RR, int *elem_num, char Type[], char Sign_Flag[]);

/* ---- exact diagonalization (include/exact.h:113-125) */
/* include/exact.h:118 -- Lanczos ("Lanczos" | "Lanczos_Slow"), inverse iteration with CG and
 * the residual check on the full CRS Ham_Info->Ham, on the device.  Fills Vector[0],
 * Value[0] (the Lanczos eigenvalue), Error[0], dim, mem_ham and Time->diag / inv_iter. */
void EXACT_DIAGONALIZE_HAMILTONIAN(EXACT_HAM_INFO *Ham_Info, EXACT_PARAMETER *Param, EXACT_TIME *Time, int p_threads);
/* include/exact.h:119,122-124 -- one on-site operator at one site on an ED vector (device);
 * Q1..Q3 map between the quantum-number sectors of W_Basis, Q2/Q3 apply the fermion sign
 * (-1)^(electrons on sites < op_site) when Sign_Flag is "Yes". */
void EXACT_V_M_Q0(CRS1 *M_On, double *Vec, double *Out_Vec, int dim_onsite, int site, int p_threads, EXACT_BASIS_INFO *Basis_Info);
void EXACT_V_M_Q1(CRS1 *M_On, int qn_out, double *Vec, int qn_in, double *Out_Vec, int dim_onsite, int site, int p_threads, EXACT_WHOLE_BASIS_Q1 *W_Basis);
void EXACT_V_M_Q2(CRS1 *M_On, int qn1_out, int qn2_out, double *Vec, int qn1_in, int qn2_in, double *Out_Vec, char Sign_Flag[], int *N_Ele, int dim_onsite, int op_site, int p_threads, EXACT_WHOLE_BASIS_Q2 *W_Basis);
void EXACT_V_M_Q3(CRS1 *M_On, int qn1_out, int qn2_out, int qn3_out, double *Vec, int qn1_in, int qn2_in, int qn3_in, double *Out_Vec, char Sign_Flag[], int *N_Ele, int dim_onsite, int op_site, int p_threads, EXACT_WHOLE_BASIS_Q3 *W_Basis);
/* include/exact.h:120-121 */
void EXACT_EXPECTATION_ONSITE(CRS1 *M_On, double *Out, double *Vec, double *T_Vec, int dim_onsite, int tot_site, int p_threads, EXACT_BASIS_INFO *Basis_Info);
void EXACT_EXPECTATION_INTERSITE_Q0(CRS1 *M_O, CRS1 *M_R, int start, int end, double *Out, double *Vec, double *T_Vec1, double *T_Vec2, int dim_onsite, int p_threads, EXACT_BASIS_INFO *Basis_Info);
/* include/exact.h:113-117,125 -- host helpers of the programs' GET_HAM (row accumulation
 * into A_Basis, kept sorted by target code), with the reference's results */
EXACT_A_BASIS **EXACT_GET_A_BASIS(int p_threads, int max_row);
void EXACT_FREE_A_BASIS(EXACT_A_BASIS **A_Basis, int p_threads);
int EXACT_FIND_SITE_STATE(long basis, int site, int dim_onsite);
void EXACT_MAKE_ELEM_ON(long basis, int site, int dim_onsite, CRS1 *M_On, long *elem_num, double coeef, EXACT_A_BASIS *A_Basis);
void EXACT_MAKE_ELEM_INTER(long basis, int site1, int site2, int dim_onsite, CRS1 *M1, CRS1 *M2, long *elem_num, double coeef, int sign, EXACT_A_BASIS *A_Basis);
double EXACT_SSD_COEFF(int site, int tot_site, char BC[], char Inter_Name[]);

/* ================================================= Part 2: dmrgx_ entry points */
/* Status codes */
#define DMRGX_OK              0
#define DMRGX_ERR_ARG        -1   /* invalid argument (sizes, strings, NULL)         */
#define DMRGX_ERR_HIP        -2   /* HIP runtime error (message in dmrgx_last_error) */
#define DMRGX_ERR_NOCONV     -3   /* Lanczos did not converge within max_step        */
#define DMRGX_ERR_NODIAG     -4   /* a row of the matrix has no diagonal element     */
#define DMRGX_ERR_NOMEM      -5   /* device allocation failed                        */
#define DMRGX_ERR_NOGPU      -6   /* no usable HIP device                            */

const char *dmrgx_last_error(void);
int dmrgx_init(int device);
void dmrgx_set_seed(uint64_t seed);
/* Same splitmix64 counter hash as the Guess == "No" start vector: values are integers
 * in [-5e7, 5e7), the range of rand()%100000000 - 50000000 (sml/LANCZOS_SLOW_SYM.c:82). */
void dmrgx_start_vector(int64_t n, uint64_t seed, double *out);
/* Host routine of the Lanczos convergence checks, in place of dstev_ on the k x k
 * tridiagonal (sml/LANCZOS_SLOW_SYM.c:137, dmrg/DMRG_LANCZOS_SLOW_OBC.c:110): lam_out[j]
 * = lowest eigenvalue of the leading (j+1) x (j+1) block of T(a[0..k-1], b[0..k-2]) for
 * j = 0..k-1, computed step by step as the solvers do (each step bounded by the last);
 * vec_out (may be NULL): normalised eigenvector of the full T for lam_out[k-1].  Host
 * only: needs no GPU. */
int dmrgx_tridiag_lowest(int32_t k, const double *a, const double *b, double *lam_out, double *vec_out);

/* ---- device matrix (full CSR, columns ascending, resident in HBM) */
typedef struct dmrgx_matrix dmrgx_matrix;
/* Upload a host CRS1.  sym != 0: M is the lower triangle incl. diagonal (the layout
 * GET_HAM_LLLRRRRL produces) and is expanded to the full matrix on the device. */
int dmrgx_matrix_from_crs1(const CRS1 *M, int sym, dmrgx_matrix **out);
int dmrgx_matrix_from_arrays(int32_t n, const int64_t *row, const int32_t *col, const double *val,
                             int sym, dmrgx_matrix **out);
void dmrgx_matrix_free(dmrgx_matrix *A);
int64_t dmrgx_matrix_nnz(const dmrgx_matrix *A);
int32_t dmrgx_matrix_dim(const dmrgx_matrix *A);
/* Copy the full CSR back to the host (row: n+1, col/val: nnz). */
int dmrgx_matrix_download(const dmrgx_matrix *A, int64_t *row, int32_t *col, double *val);
/* y = A x with device pointers (x, y of length n), on the library stream. */
int dmrgx_spmv_device(const dmrgx_matrix *A, const double *x_dev, double *y_dev);
/* y = A x with host pointers. */
int dmrgx_spmv_host(const dmrgx_matrix *A, const double *x, double *y);
/* SpMV schedule of this matrix: lanes (threads) per row, a power of two 1..64; the cap on
 * workgroups (0 = one workgroup per row group); and the load variant 0..39.  Bits 3-5 are
 * the storage format the products read: 0 CSR, 1 value dictionary -- 8-byte entries
 * (column, index into a table of the matrix's distinct values), FP64 products only, when
 * the matrix has few distinct values -- 2 delta columns -- 2-byte differences of
 * consecutive columns in a row plus the row's first column, when every difference fits --
 * 3 sliced -- slices of 64/lanes rows read one slot column per lane, entry k*lanes + l of
 * each row in slot k of lane l, padded to the slice's longest row; lanes 1, 2 or 4 only --
 * 4 sliced with 16-bit column differences and a 16-bit index per slot into the value
 * dictionary's table (4 bytes per entry), when the matrix has at most 65536 distinct
 * values (else it reads format 3 with 16-bit columns, which gives the same bits); lanes 1,
 * 2 or 4 only.  A format the matrix cannot have reads CSR; formats are built on the
 * device at their first use.  Bit 0: for CSR two entries per lane per load (a 16-byte value pair and an
 * 8-byte column pair); for the sliced form 16-bit column differences per slot column
 * (0xFFFF marking a column kept whole in an exception list).  Bit 1: eight entries in
 * flight per lane instead of four (CSR, dictionary, delta); non-temporal matrix loads
 * (sliced; bit 0 is ignored by format 4).  Bit 2: each workgroup takes a contiguous range
 * of row groups and each XCD a contiguous share of the rows, instead of round-robin groups.
 * Bit 1, formats 1-2 against 0, format 4 against 3 and the sliced form's bit 0 never
 * change results; lanes, the sliced forms against the others, and CSR's bit 0 change
 * the summation order within a row, and cap and bit 2 the order of the dot products fused
 * into the solvers' products.  Defaults: lanes from the mean row length, cap 4096,
 * variant 2; dmrgx_matrix_autotune times every candidate on the device (product + fused
 * dot), then times the eight fastest as the solver runs them -- a Lanczos step and a CG
 * iteration with their fused vector kernels -- and keeps the fastest; *best_us_out is that
 * step-pair time.  Single and batched solves use the matrix's schedule. */
int dmrgx_matrix_set_schedule(dmrgx_matrix *A, int32_t lanes, int32_t grid_cap, int32_t load_variant);
int dmrgx_matrix_get_schedule(const dmrgx_matrix *A, int32_t *lanes, int32_t *grid_cap, int32_t *load_variant);
int dmrgx_matrix_autotune(dmrgx_matrix *A, int32_t *lanes_out, int32_t *grid_out, int32_t *load_out,
                          double *best_us_out);
/* Mean microseconds of one product + fused dot under the given schedule (reps launches
 * after one warm-up), leaving A's schedule unchanged. */
int dmrgx_matrix_time_schedule(dmrgx_matrix *A, int32_t lanes, int32_t grid_cap, int32_t load_variant,
                               int32_t reps, double *us_out);
/* Distinct values in A's value dictionary (built now if it has not been tried); 0 if A
 * keeps the plain form. */
int dmrgx_matrix_dict_values(dmrgx_matrix *A, int64_t *n_values);
/* Columns of A's sliced form with lanes (1, 2 or 4) lanes per row that the 16-bit column
 * form keeps whole (built now if it has not been); for the products' byte count. */
int dmrgx_matrix_sell16_exceptions(dmrgx_matrix *A, int32_t lanes, int64_t *exceptions);

/* ---- ground-state solve (Lanczos -> inverse iteration with CG), all on the device */
typedef struct {
   double diag_acc;         /* Lanczos convergence threshold                     */
   int diag_min_step;
   int diag_max_step;
   int lan_type;            /* 0 "Normal" |E_k - E_k-1| < acc, 1 "Tight" |s_k| < acc,
                                2 "Residual" (no reference counterpart): the Ritz pair's
                                residual bound beta_{k+1}|s_k| < ii_acc/(2 sqrt(n)), so that
                                the inverse iteration's residual test usually passes at once */
   int guess;               /* 0: counter-hash start vector (seed), 1: vec_inout    */
   uint64_t seed;
   double ii_acc;           /* inverse iteration: stop when sum|Hv - Ev| < ii_acc */
   double ii_diag_add;
   int ii_max_step;
   double cg_acc;           /* CG: stop when ||r||^2 < cg_acc                      */
   int cg_max_step;
   int skip_ii;             /* 1: Lanczos only                                    */
   int precision;           /* 0: FP64; 1: FP32 CRS values + FP32 SpMV in Lanczos,
                                FP64 recurrence scalars and FP64 II/CG refinement;
                                2 (single solves): FP32 Lanczos until its Ritz residual
                                bound falls to mixed_switch*max(1,|E|), then FP64
                                Lanczos from that Ritz vector until the bound is below
                                ii_acc/(2 sqrt(n)), then the FP64 residual test (and
                                FP64 II/CG if it fails)                              */
   double mixed_switch;     /* precision 2: hand-over residual (<= 0: 1e-3)          */
} dmrgx_solve_params;

typedef struct {
   double eig_val;
   double gs_error;         /* sum |Hv - Ev| after inverse iteration               */
   int lanczos_steps;
   int ii_steps;
   int cg_steps;
   int spmv_calls;
   double seconds;
   int lanczos_fp32_steps;  /* of lanczos_steps, those with FP32 values (precision 1, 2) */
} dmrgx_solve_stats;

void dmrgx_default_params(dmrgx_solve_params *p);
/* vec_inout: host array of length n; start vector when guess == 1, ground state out. */
int dmrgx_ground_state(const dmrgx_matrix *A, const dmrgx_solve_params *p, double *vec_inout,
                       dmrgx_solve_stats *stats);


/* ---- superblock assembly on the device from a recorded term list
 * Recording: between dmrgx_record_begin() and dmrgx_record_end() the DMRG_MAKE_ELEM_*
 * symbols of this library append (kind, operators, coefficient, sign rule, parity arrays)
 * to a term list instead of computing; run the model's MAKE_ELEMENT_HAM_LLLRRRRL once
 * (any row) to capture it.  The recorded pointers must stay valid until assembly. */
typedef struct dmrgx_terms dmrgx_terms;
typedef struct dmrgx_superblock dmrgx_superblock;
int dmrgx_record_begin(void);
dmrgx_terms *dmrgx_record_end(void);
void dmrgx_terms_free(dmrgx_terms *t);
int dmrgx_terms_count(const dmrgx_terms *t);

/* Term of the array form.  kind: 0 LL, 1 LR, 2 RR, 3 RL, 4 LLLR, 5 LLRL, 6 LLRR, 7 LRRL,
 * 8 LRRR, 9 RRRL, 10 ZERO (the element makers of include/dmrg.h:330-340, in that order).
 * op1/op2 index ops[]; ele0..2 index eles[] in the argument order of the element maker
 * (e.g. LRRL: Ele_RR, Ele_LR); sign_mode 0 = Sign_Flag "No", 1 = first Type ("LL_LR",
 * "LL_RL", "LL_RR", "LR_RL", "LR_RR", "RR_RL"), 2 = the reversed Type. */
typedef struct {
   int32_t kind, op1, op2, ele0, ele1, ele2, sign_mode, reserved;
   double coeff;
} dmrgx_term;

/* Upload a superblock (basis from GET_BASIS_SUPERBLOCK, operators, term list). */
int dmrgx_superblock_create(const DMRG_BASIS *B, const dmrgx_terms *T, dmrgx_superblock **out);
int dmrgx_superblock_create_arrays(int32_t n, int32_t dim_LL, int32_t dim_RR, int32_t dim_onsite,
                                   const int16_t *LL, const int16_t *LR, const int16_t *RR, const int16_t *RL,
                                   int32_t n_ops, const CRS1 *ops, int32_t n_eles, const int32_t *const *eles,
                                   const int32_t *ele_len, int32_t n_terms, const dmrgx_term *terms,
                                   dmrgx_superblock **out);
/* Superblock basis from the blocks' quantum numbers, as GET_BASIS_SUPERBLOCK of each DMRG
 * program builds it (main/dmrg/hubbard_vf/src/allocations/GET_BASIS_SUPERBLOCK.c:63-97,
 * xxz_vf/...:36-64, klm_vf/...:66-97, ordered as dmrg/DMRG_QUICK_SORT_BASIS_Q1/Q2.c): the
 * states (LL, LR, RR, RL) whose S^z (n_qn == 1) or S^z and electron numbers (n_qn == 2)
 * add up to the target, ordered by electrons of LL+LR, S^z of LL+LR, LL, LR, RR, RL.
 * sz_LL / ele_LL: Tot_Sz / Tot_Ele of the LL block's states (System->Tot_Sz[LL_site]),
 * likewise RR; sz_site / ele_site: of the on-site states (FIND_SITE_SZ,
 * ONSITE_FIND_SITE_SZ/ELE_SZBASIS_*); ele_* are ignored when n_qn == 1.  The target is the
 * program's (GET_BASIS_SUPERBLOCK.c:24-42 adjusts it per step; pass the adjusted value). */
typedef struct {
   int32_t n_qn, dim_LL, dim_RR, dim_onsite, target_sz, target_ele;
   const int32_t *sz_LL, *ele_LL, *sz_RR, *ele_RR, *sz_site, *ele_site;
} dmrgx_basis_qn;
/* *n_out = the sector dimension; with LL..RL given (n_out entries each) also the basis,
 * built on the device.  No counterpart of the whole-space enumeration and quicksort: the
 * order is produced directly, sector by sector. */
int dmrgx_superblock_basis(const dmrgx_basis_qn *q, int64_t *n_out, int16_t *LL, int16_t *LR, int16_t *RR,
                           int16_t *RL);
/* dmrgx_superblock_create_arrays with the basis built on the device from q. */
int dmrgx_superblock_create_qn(const dmrgx_basis_qn *q, int32_t n_ops, const CRS1 *ops, int32_t n_eles,
                               const int32_t *const *eles, const int32_t *ele_len, int32_t n_terms,
                               const dmrgx_term *terms, dmrgx_superblock **out);
int dmrgx_superblock_n(const dmrgx_superblock *S, int32_t *n);
/* Assemble on the device: the lower-triangle CRS of GET_HAM_LLLRRRRL (kept in the handle,
 * nnz in *nnz_lower) and, if full_out != NULL, its full symmetric expansion ready to solve. */
int dmrgx_superblock_assemble(dmrgx_superblock *S, int64_t *nnz_lower, dmrgx_matrix **full_out);
int dmrgx_superblock_download_lower(const dmrgx_superblock *S, int64_t *row, int32_t *col, double *val);
void dmrgx_superblock_free(dmrgx_superblock *S);

/* ---- Kronecker form of a superblock Hamiltonian: the operator of the CRS that
 * dmrgx_superblock_assemble builds (GET_HAM_LLLRRRRL, main/dmrg/<model>_vf/src/allocations/
 * GET_HAM_LLLRRRRL.c:11-114), applied without assembling it.  No reference symbol has this
 * form; it applies the superblock operator as the reference's matrix-free product does
 * (dmrg/DMRG_MATRIX_VECTOR_PRODUCT_OBC.c:18-83: sub-Hamiltonians of the left and right
 * halves plus the terms that couple them), built from the recorded term list instead of
 * the program's GET_HAM_LLLR/RRRL.  Per sector of the superblock basis (one LLLR quantum
 * number: a dense block of LLLR by RRRL states) the LL/LR/LLLR terms form a dense matrix
 * acting on the block's rows and the RR/RL/RRRL terms one acting on its columns; the
 * coupling terms are applied row by row through the superblock inverse.  Results agree with
 * the CRS product to rounding (the sums are in an order of their own).  The superblock S
 * must outlive the handle.  DMRGX_ERR_ARG if the basis is not a product of LLLR and RRRL
 * states in every sector (a basis from GET_BASIS_SUPERBLOCK always is). */
typedef struct dmrgx_kron dmrgx_kron;
int dmrgx_superblock_kron(const dmrgx_superblock *S, dmrgx_kron **out);
/* dimension, sectors, coupling terms, floating-point operations of the dense products per
 * product (any pointer may be null) */
int dmrgx_kron_info(const dmrgx_kron *K, int32_t *n, int32_t *sectors, int32_t *coupling_terms, double *flops);
/* kernel of the dense sector products: 0 FP64 vector units on 64 x 64 tiles; FP64 matrix
 * cores on 1: 64 x 64, 2: 128 x 128, 3: 64 x 128, 4: 128 x 64 tiles, 5: 64 x 64 tiles with
 * k-slabs of 32, 6: as 1 with at most 128 registers per lane, 7: as 2 with at most 256
 * (default 1); plus 8 for the same kernels with consecutive tiles (one sector's) dealt to
 * the workgroups of one XCD.  The results agree to rounding between the variants. */
int dmrgx_kron_set_kernel(dmrgx_kron *K, int32_t variant);
/* times every variant (5 products each after one warm-up) and keeps the fastest; the default
 * (variant 1) stays unless another is more than 2 % faster */
int dmrgx_kron_autotune(dmrgx_kron *K, int32_t *variant_out, double *us_out);
/* y = H x (host arrays of length n) */
int dmrgx_kron_spmv(const dmrgx_kron *K, const double *x, double *y);
/* dmrgx_ground_state on the Kronecker form (precision 0 only) */
int dmrgx_kron_ground_state(const dmrgx_kron *K, const dmrgx_solve_params *p, double *vec_inout,
                            dmrgx_solve_stats *stats);
void dmrgx_kron_free(dmrgx_kron *K);
/* Drop-in body of GET_HAM_LLLRRRRL (main/dmrg/<model>_vf/src/allocations/GET_HAM_LLLRRRRL.c:11):
 * assembles on the device, returns a malloc'ed CRS1 (free with FREE_CRS1) holding the
 * reference-layout lower triangle, and keeps the expanded device copy for the next
 * DMRG_DIAGONALIZE_SUPERBLOCK_SYM call on that CRS1 (no re-upload).  exit(1) on error. */
CRS1 *dmrgx_get_ham_lllrrrrl(const DMRG_BASIS *B, const dmrgx_terms *T);

/* ---- single-slot products between sectors: DMRG_V_M_{LL,LR,RR,RL}_Q1..Q4
 * (include/dmrg.h:369-382, dmrg/DMRG_V_M_*_Q1..Q4.c).  The reference passes a
 * DMRG_WHOLE_BASIS_Q1..Q4 (include/dmrg.h:140-200): per sector the states' LL/LR/RR/RL
 * arrays and one whole-space inverse table Inv_LLLRRRRL (dim_LL*d*dim_RR*d entries, each
 * -1 or the state's index within its own sector).  Neither the table's length nor the
 * input vector's is in those structs, so here the table goes to the device once, with its
 * block sizes, the sectors are registered one by one (the caller keeps the map from its
 * quantum-number indices qn1_out.. to the ids), and each product names its output sector
 * by id and gives the input vector's length.  Results are bit-identical to the reference's
 * loop.  An operator that reaches a state whose Inv lies at or past dim_in is an error
 * (the reference would read past V). */
typedef struct dmrgx_whole_basis dmrgx_whole_basis;
int dmrgx_whole_basis_create(int32_t dim_LL, int32_t dim_RR, int32_t dim_onsite, const int32_t *inv,
                             dmrgx_whole_basis **out);
/* One sector's states (LL_LLLRRRRL[qn..] etc., dim entries each); its id in *id (0, 1, ...). */
int dmrgx_whole_basis_add_sector(dmrgx_whole_basis *W, int32_t dim, const int16_t *LL, const int16_t *LR,
                                 const int16_t *RR, const int16_t *RL, int32_t *id);
/* Out_V (the output sector's dimension) = M acting on slot (0 LL, 1 LR, 2 RR, 3 RL) of V
 * (dim_in entries).  sign_flag != 0 is Sign_Flag "Yes": (-1)^(Ele_LL[LL]) for LR,
 * (-1)^(Ele_LL[LL] + Ele_LR[LR]) for RR, (-1)^(Ele_LL[LL] + Ele_LR[LR] + Ele_RR[RR]) for
 * RL (ele_LL: dim_LL entries, ele_LR: d, ele_RR: dim_RR; the ones a slot does not use may
 * be NULL); the LL product has no sign, as DMRG_V_M_LL_Q* has no Sign_Flag. */
int dmrgx_v_m_sector(dmrgx_whole_basis *W, int32_t slot, const CRS1 *M, int32_t sector_out, int32_t dim_in,
                     const double *V, const int32_t *ele_LL, const int32_t *ele_LR, const int32_t *ele_RR,
                     int32_t sign_flag, double *Out_V);
/* EXPECTATION_INTERSITE_Q2 of the XXZ program (main/dmrg/xxz_vf/src/expectations/
 * EXPECTATION_INTERSITE_Q2.c:11-67, the S^x-type correlation across the S^z-2 and S^z+2
 * sectors) on the device, its intermediate vectors kept there: Out[r] for
 * r = 0 .. RR_site + LL_site + 3 - origin, the sectors named by their ids (the input vector
 * has the input sector's dimension).  M_CF[0 .. LL_site - origin], M_LL[origin] (unless
 * origin == LL_site + 1), M_RR[0 .. RR_site].  Nothing is written when origin > LL_site + 1,
 * as in the reference.  Inner products are the library's deterministic reductions, so
 * entries agree with the reference to rounding. */
int dmrgx_expectation_intersite_sector(dmrgx_whole_basis *W, int32_t sector_in, int32_t sector_out1,
                                       int32_t sector_out2, const CRS1 *const *M_CF, const CRS1 *const *M_LL,
                                       const CRS1 *M_On, const CRS1 *const *M_RR, int32_t origin, int32_t LL_site,
                                       int32_t RR_site, const double *Vec, double *Out);
void dmrgx_whole_basis_free(dmrgx_whole_basis *W);

/* ---- reduced density matrices of the left half (truncation, first step)
 * dmrg/DMRG_GET_DENSITY_MATRIX.c:5-107, called per block by DMRG_GET_SYSTEM_INFO_Q1..Q3
 * (dmrg/DMRG_GET_SYSTEM_INFO_Q2.c:22-23).  Block b is the run of Dim_Block[b] LLLR states
 * (LL_LLLR/LR_LLLR, Sum_LLLR of the basis) after those of the blocks before it; its matrix
 * is rho[r][c] = sum_k Vec[state (r, k)] Vec[state (c, k)] over the partners k the block's
 * states share.  Entries are bit-identical to the reference's; the reference's checks
 * (equal Sum_LLLR and partners within a block, no missing state) are errors here too. */
/* All blocks at once: *n_out = sum of Dim_Block[b]^2; out (if not NULL) receives the
 * blocks' row-major matrices one after another. */
int dmrgx_density_matrices(const DMRG_BASIS *B, int32_t n_blocks, const int32_t *Dim_Block, const double *Vec,
                           int64_t *n_out, double *out);
/* include/dmrg.h:383 -- one block, rows malloc'ed as GET_ARRAY_DOUBLE2 does (free with
 * FREE_ARRAY_DOUBLE2); exit(1) on error. */
double **DMRG_GET_DENSITY_MATRIX(int block, int *Dim_Block, DMRG_BASIS *Dmrg_Basis, double *Vec, int p_threads);

/* ---- truncation: density matrices, their block eigendecompositions and the kept states
 * include/dmrg.h:390-391 (dmrg/DMRG_GET_SYSTEM_INFO_Q1.c, DMRG_GET_SYSTEM_INFO_Q2.c:7-236):
 * blocks of equal quantum numbers (DMRG_GET_DENSITY_MATRIX_INFO_Q1/Q2 and their order
 * checks), each block's density matrix and its eigendecomposition on the device in place
 * of LAPACK_DSYEV (blocks up to 192 states: the library's own solver, one workgroup per
 * block -- Householder tridiagonalisation in LDS, every eigenvalue by Sturm bisection, and
 * inverse iteration for the kept states' eigenvectors only; larger blocks: rocSOLVER
 * dsyevd), then the reference's selection: all
 * eigenvalues sorted descending (stable), the cut at max_dim moved down past values that
 * differ by 1e-15 or less, the truncation error, the kept states re-sorted by Q_Number2 and
 * then Q_Number1 (stable), and the transformation matrix (CCS1, columns = kept
 * eigenvectors over the LLLR states) with its transpose (CRS1).  Allocations are those of
 * GET_CCS1/GET_CRS1/GET_ARRAY_*: DMRG_FREE_SYSTEM_INFO frees the result.  The eigenvalues
 * (Val_DM_Dist), dim_renorm, the truncation error and the quantum numbers match the
 * reference to LAPACK's accuracy; each kept eigenvector may differ from LAPACK's by its
 * sign, and within a degenerate set by a rotation of the set.  Time in Dmrg_Time->make_dens_mat. */
DMRG_SYSTEM_INFO *DMRG_GET_SYSTEM_INFO_Q1(int *Q_Number, DMRG_BASIS *Dmrg_Basis, double *Vec, int max_dim,
                                          int p_threads, DMRG_TIME *Dmrg_Time);
DMRG_SYSTEM_INFO *DMRG_GET_SYSTEM_INFO_Q2(int *Q_Number1, int *Q_Number2, DMRG_BASIS *Dmrg_Basis, double *Vec,
                                          int max_dim, int p_threads, DMRG_TIME *Dmrg_Time);
/* The same with a status code: n_qn 1 (Q_Number2 ignored) or 2; *out = NULL on error. */
int dmrgx_get_system_info(int32_t n_qn, const int32_t *Q_Number1, const int32_t *Q_Number2, const DMRG_BASIS *B,
                          const double *Vec, int32_t max_dim, DMRG_SYSTEM_INFO **out);
/* No counterpart in the reference: which eigensolver the truncation calls above use for
 * blocks up to 192 states -- 0 (default) the library's own, 1 rocSOLVER dsyevd for every
 * block.  Process-wide; DMRGX_ERR_ARG for any other mode.  The own solver hands a block to
 * rocSOLVER by itself when one of its kept eigenvectors comes out linearly dependent on the
 * others. */
int dmrgx_set_block_eigensolver(int mode);
/* Frees what the three calls above return, as DMRG_FREE_SYSTEM_INFO (include/dmrg.h:357,
 * dmrg/DMRG_FREE_SYSTEM_INFO.c) does; the reference's own function frees it as well. */
void dmrgx_free_system_info(DMRG_SYSTEM_INFO *Info);
/* include/SML.h:181 (sml/CRS_CRS_CCS_PRODUCT.c:6-124) -- Out = M1 (M2 M3), M3 in CCS, entries
 * of magnitude <= 1e-15 dropped after each of the two passes, Out's columns ascending; the
 * renormalisation T^dagger O T of DMRG_TRANS_MAT_*.  Bit-identical to the reference.  Work and
 * Work_Vec are not written (the intermediate product stays on the device); exit(1) with the
 * reference's conditions: dimensions that do not chain, Work->max_col <= M3's columns,
 * Out->max_row <= M1's rows, and the kept entries of M2 M3 (or of Out) before the last
 * position it evaluates reaching Work->max_val (or Out->max_val) -- so an exactly full
 * Work or Out whose last position is dropped exits, as in the reference. */
void CRS_CRS_CCS_PRODUCT(CRS1 *M1, CRS1 *M2, CCS1 *M3, CRS1 *Out, CCS1 *Work, double *Work_Vec);

/* ---- exact diagonalization on the device
 * A problem is a sorted basis of base-d codes (site s is digit s, site 0 least
 * significant: exact/EXACT_FIND_SITE_STATE.c), on-site operators (d x d CRS) and a term
 * list the device assembler turns into the full CRS of GET_HAM
 * (main/exact_diag/klm_vf/src/GET_HAM.c:11-92), bit-identical to the reference's element
 * makers driven by the same terms.
 *
 * Term: kind 0 = coeff * op1 on site1 (EXACT_MAKE_ELEM_ON), kind 1 = s * coeff *
 * op1(site1) op2(site2) (EXACT_MAKE_ELEM_INTER) with integer sign s = sign, negated when
 * par_lo < par_hi and the row state holds an even number of electrons (n_ele[] of the
 * local states) on sites [par_lo, par_hi) -- the fermion sign rule of
 * main/exact_diag/klm_vf/src/MAKE_ELEMENT_HAM.c:31-37,44-64. */
typedef struct {
   int32_t kind, site1, site2, op1, op2, par_lo, par_hi, sign;
   double coeff;
} dmrgx_ed_term;

typedef struct dmrgx_ed dmrgx_ed;
/* basis: n sorted codes (host); ops: n_ops CRS1 of dimension d; n_ele: d entries (may be
 * NULL when no term has a parity range). */
int dmrgx_ed_create(int32_t L, int32_t d, int64_t n, const int64_t *basis, int32_t n_ops, const CRS1 *ops,
                    const int32_t *n_ele, int32_t n_terms, const dmrgx_ed_term *terms, dmrgx_ed **out);
/* Same with the basis enumerated on the device: all codes of d^L whose per-site sums of
 * qn1[] and (if qn2 != NULL) qn2[] equal target1 / target2, ascending. */
int dmrgx_ed_create_sector(int32_t L, int32_t d, const int32_t *qn1, int32_t target1, const int32_t *qn2,
                           int32_t target2, int32_t n_ops, const CRS1 *ops, const int32_t *n_ele,
                           int32_t n_terms, const dmrgx_ed_term *terms, dmrgx_ed **out);
int64_t dmrgx_ed_dim(const dmrgx_ed *E);
int dmrgx_ed_download_basis(const dmrgx_ed *E, int64_t *basis);
/* Full CRS of the Hamiltonian on the device (columns ascending), ready for
 * dmrgx_ground_state. */
int dmrgx_ed_assemble(const dmrgx_ed *E, dmrgx_matrix **out);
/* y = O x on device vectors: O = ops[op] of E_in at site, from E_in's basis to E_out's
 * (the same handle for a sector-preserving operator); sign_flag: fermion sign of sites
 * < site (EXACT_V_M_Q2.c:31-44).  E_out may carry no terms (only its basis is used). */
int dmrgx_ed_apply_onsite(const dmrgx_ed *E_out, const dmrgx_ed *E_in, int32_t op, int32_t site, int32_t sign_flag,
                          const double *x_dev, double *y_dev);
void dmrgx_ed_free(dmrgx_ed *E);

/* Models with no ED driver in the reference, written on the klm_vf pattern with each
 * DMRG program's coupling order:
 *   model 0: spin-S XXZ  sum J_z SzSz + J_xy/2 (S+S- + S-S+) + sum D_z Sz^2 + h_z Sz;
 *            ops 0 Ham_On, 1 S+, 2 S-, 3 Sz, 4 zero diagonal; sector 2*Sz_tot = sz
 *   model 1: Hubbard     sum t (c+_{i s} c_{i+1 s} + h.c.) + V (nup nup + ndn ndn)
 *            + sum U n_up n_dn + h_z Sz + mu n;  ops 0 Ham_On, 1 c+up, 2 cup, 3 c+dn, 4 cdn,
 *            5 nup, 6 ndn, 7 zero diagonal, 8 n, 9 Sz;  sector (N, 2*Sz_tot = sz)
 * bc: 0 OBC, 1 PBC, 2 SSD (EXACT_SSD_COEFF weights).  spin = 2S (model 0). */
typedef struct {
   int32_t model, L, spin, N, sz, bc;
   double J_xy, J_z, h_z, D_z, t, U, V, mu;
} dmrgx_ed_model;
int dmrgx_ed_model_build(const dmrgx_ed_model *m, dmrgx_ed **out);
/* Host only (no device work): what dmrgx_ed_model_build uploads.  Sizes come back in
 * d, n_ops, n_terms (and op_nnz[n_ops]); any array argument may be NULL.  Operators are
 * packed: op_row[k*(d+1) ..], op_col/op_val at the running sum of op_nnz.  targets:
 * {target1, target2, qn2 used (0|1)}. */
int dmrgx_ed_model_describe(const dmrgx_ed_model *m, int32_t *d, int32_t *n_ops, int64_t *op_nnz, int64_t *op_row,
                            int32_t *op_col, double *op_val, int32_t *n_terms, dmrgx_ed_term *terms, int32_t *n_ele,
                            int32_t *qn1, int32_t *qn2, int32_t *targets);

/* ---- batched solves and continuous batching
 * An engine runs many independent ground-state solves together: every Lanczos step (and
 * every CG iteration of the refinement) of all active requests is one launch per kernel,
 * rows of all requests side by side; each request keeps the blocks, lanes and summation
 * order of dmrgx_ground_state, so its result is bit-identical to a solve on its own.
 * Requests are admitted at chunk boundaries while the HBM budget allows. */
typedef struct dmrgx_engine dmrgx_engine;
typedef struct {
   int64_t hbm_budget_bytes;   /* device memory the active requests may hold            */
   int32_t max_active;         /* requests advancing together                           */
   int32_t chunk_ticks;        /* steps between host checks                             */
   int32_t basis_steps_estimate; /* Lanczos vectors assumed per request for admission   */
   int32_t use_graphs;         /* 1: a chunk whose launch sequence repeats the previous
                                  one's is captured once as a hipGraph and replayed        */
} dmrgx_engine_opts;
typedef struct {
   int64_t request_id;
   uint64_t tag;
   int32_t status;             /* DMRGX_OK or DMRGX_ERR_NOCONV                            */
   int32_t n;
   double eig_val, gs_error;
   int32_t lanczos_steps, ii_steps, cg_steps, spmv_calls;
   double t_submit, t_start, t_done;   /* steady-clock seconds                            */
} dmrgx_engine_result;
void dmrgx_engine_default_opts(dmrgx_engine_opts *o);
int dmrgx_engine_create(const dmrgx_engine_opts *o, dmrgx_engine **out);
/* Queue a solve of A (a dmrgx_matrix; take_ownership != 0 frees it with the request).
 * p->guess must be 0 (the start vector comes from p->seed). */
int dmrgx_engine_submit(dmrgx_engine *E, dmrgx_matrix *A, int take_ownership, const dmrgx_solve_params *p,
                        uint64_t tag, int64_t *request_id);
/* Admit and advance up to max_chunks chunks of steps. */
int dmrgx_engine_run(dmrgx_engine *E, int32_t max_chunks);
int32_t dmrgx_engine_pending(const dmrgx_engine *E);   /* queued + active */
int32_t dmrgx_engine_done(const dmrgx_engine *E);      /* finished, not yet polled */
/* Tag and dimension of the result dmrgx_engine_poll would return next: 1, or 0 if none. */
int dmrgx_engine_peek(const dmrgx_engine *E, uint64_t *tag, int32_t *n);
/* Oldest finished request: returns 1 and fills *res (and vec_out[0..n) if given), 0 if none. */
int dmrgx_engine_poll(dmrgx_engine *E, dmrgx_engine_result *res, double *vec_out);
void dmrgx_engine_free(dmrgx_engine *E);
/* Host-side seconds spent by the engine so far: [0] building and uploading the chunks'
 * launch descriptors, [1] launching, [2] waiting for each chunk, [3] Lanczos convergence
 * checks (with the Ritz vectors and first residual checks they start), [4] CG ends and
 * inverse-iteration rounds, [5] admitting requests; *chunks = chunks run. */
int dmrgx_engine_host_times(const dmrgx_engine *E, double *seconds6, int64_t *chunks);
/* Chunks launched eagerly, graphs captured, graph replays. */
int dmrgx_engine_graph_stats(const dmrgx_engine *E, int64_t *eager, int64_t *captures, int64_t *replays);
/* count solves through one engine; vec_out[i] (host, n_i doubles) and stats[i] may be NULL. */
int dmrgx_ground_state_batch(int32_t count, dmrgx_matrix *const *A, const dmrgx_solve_params *p,
                             double *const *vec_out, dmrgx_solve_stats *stats);

/* ---- measurement
 * Between begin and end, every CSR SpMV launch of the library that always runs in full is
 * bracketed by HIP events on the library's stream (the CG products are not: a launch past
 * convergence returns at once on the device); end returns the summed kernel time, the
 * launch count and the algorithmic bytes (nnz*(sizeof(val)+4) + (n+1)*8 + 2*n*8 per
 * launch). */
int dmrgx_profile_begin(void);
/* As dmrgx_profile_begin, bracketing only the first of every `every` SpMV launches: each
 * event pair costs the stream a few microseconds between kernels, so sampling keeps the
 * timed solves close to their unprofiled speed.  end then reports the sampled launches. */
int dmrgx_profile_begin_sampled(int32_t every);
int dmrgx_profile_end(double *spmv_ms, int64_t *spmv_launches, double *spmv_alg_bytes);
/* Plain streaming copy of count doubles (device pointers): traffic calibration for
 * rocprofv3 FETCH_SIZE / WRITE_SIZE. */
int dmrgx_copy_device(const double *src, double *dst, int64_t count);
int dmrgx_device_alloc(int64_t bytes, void **out);
int dmrgx_device_free(void *p);
/* Synchronous copies between caller host memory and device memory (bytes). */
int dmrgx_copy_to_device(void *dst_dev, const void *src_host, int64_t bytes);
int dmrgx_copy_to_host(void *dst_host, const void *src_dev, int64_t bytes);

/* ---- allocation-free entry point
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops and harnesses that forbid hipMalloc).  The
 * computation is CONVERGE_CHECK's (sml/CONVERGE_CHECK.c:7-23), *out = sum_i |eig_val v_i -
 * (H v)_i|, over a full CRS (int64 row pointers, int32 columns, FP64 values; n rows, nnz
 * entries) and a vector that are already in device memory; *out is host memory.  It calls
 * no hipMalloc/hipFree and no allocating library routine: the block partials live in a
 * fixed file-scope device array of 4128 doubles (33 KB) per GPU, held until the process
 * ends.  Same kernels and schedule as CONVERGE_CHECK on the same arrays, so the same bits.
 * Calls that overlap on one GPU share that array and race; the call returns after the
 * result has reached *out.  DMRGX_ERR_ARG for negative sizes or NULL arrays. */
int dmrgx_converge_check_noalloc(int32_t n, int64_t nnz, const int64_t *row_dev, const int32_t *col_dev,
                                 const double *val_dev, const double *vec_dev, double eig_val, double *out);

#ifdef __cplusplus
}
#endif
#endif /* DMRG_C_AMD_H */
