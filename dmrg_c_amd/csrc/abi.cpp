// extern "C" entry points of libdmrg_c_amd.so.
//
// Part 1: the reference's own symbols (include/SML.h, include/dmrg.h) with the reference's
// argument meaning and error behaviour: fatal conditions print the reference's message and
// exit(1); soft ones warn, append to ./SML_out/error.txt and continue; step counts are
// appended to ./SML_out/<SOLVER>_Step.txt (DMRGX_SML_OUT=0 turns the files off).
// Part 2: dmrgx_* with status codes.
#include <sys/stat.h>

#include <cstdarg>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "entry.h"

using namespace dmrgx;

namespace dmrgx {
const std::string &last_error();
void launch_start_vector(int64_t n, uint64_t seed, double *out, hipStream_t s);

std::recursive_mutex &api_mutex() {
  static std::recursive_mutex mu;
  return mu;
}

static bool sml_out_enabled() {
  const char *e = getenv("DMRGX_SML_OUT");
  return !(e && strcmp(e, "0") == 0);
}

void sml_append(const char *file, const char *fmt, ...) {
  if (!sml_out_enabled()) return;
  mkdir("SML_out", 0777);
  char path[256];
  snprintf(path, sizeof(path), "./SML_out/%s", file);
  FILE *f = fopen(path, "a+");
  if (!f) {
    printf("Can't open file.\n");
    exit(1);
  }
  va_list ap;
  va_start(ap, fmt);
  vfprintf(f, fmt, ap);
  va_end(ap);
  fclose(f);
}

void fatal(const char *fn, const char *fmt, ...) {
  printf("Error in %s\n", fn);
  va_list ap;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
  fflush(stdout);
  exit(1);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void check_square(const char *fn, const CRS1 *M) {
  if (M->row_dim != M->col_dim)
    fatal(fn, "The input matrix is not a square matirx(row=%d,col=%d)\n", M->row_dim, M->col_dim);
  if (M->row_dim <= 0 || M->col_dim <= 0)
    fatal(fn, "row_dim(%d) or col_dim(%d) is illegal value\n", M->row_dim, M->col_dim);
}

}  // namespace dmrgx

namespace {

// Lanczos entry shared by the four LANCZOS* symbols.
void lanczos_entry(const char *fn, const char *step_file, BOX_LAN *Box, bool sym, bool slow) {
  check_square(fn, Box->M);
  if (Box->max_step <= Box->min_step)
    fatal(fn, "Box->max_step=%d,Box->min_step=%d\n", Box->max_step, Box->min_step);
  const bool guess = strcmp(Box->Guess, "Yes") == 0;
  if (!guess && strcmp(Box->Guess, "No") != 0) fatal(fn, "Box->Guess=%s is error\n", Box->Guess);
  int type;
  if (strcmp(Box->Type, "Normal") == 0) type = 0;
  else if (strcmp(Box->Type, "Tight") == 0) type = 1;
  else fatal(fn, "Box->Type=%s is error\n", Box->Type);
  const int n = Box->M->row_dim;
  run_or_die(fn, [&] {
    auto A = upload_csr(n, Box->M->Row, Box->M->Col, Box->M->Val, sym);
    DevVec v(n);
    if (guess) v.up(Box->Eig_Vec);
    else launch_start_vector(n, ctx().seed, v.d.p, ctx().stream);
    LanczosOpts o;
    o.acc = Box->acc;
    o.min_step = Box->min_step;
    o.max_step = Box->max_step;
    o.max_block_step = slow ? 0 : Box->max_block_step;
    o.type = type;
    SolveOut out;
    if (n == 1) {
      Box->eig_val[0] = Box->M->Val[0];
      Box->Eig_Vec[0] = 1.0;
      sml_append(step_file, "%d\n", 0);
      return;
    }
    lanczos(op_of(*A), o, v.d.p, out);
    if (!out.lan_converged) {
      sml_append("error.txt", "Error in %s\nDoes not converge (max step is %d)\n", fn, Box->max_step);
      fatal(fn, "Does not converge (max step is %d)\n", Box->max_step);
    }
    Box->eig_val[0] = out.eig;
    v.down(Box->Eig_Vec);
    sml_append(step_file, "%d\n", out.lan_steps);
  });
}

void ii_entry(const char *fn, BOX_II *B, bool sym) {
  if (B->M->row_dim != B->M->col_dim)
    fatal(fn, "The input matrix is not a square matirx(row=%d,col=%d)\n", B->M->row_dim, B->M->col_dim);
  const bool mr = strcmp(B->Type, "MR") == 0;
  if (!mr && strcmp(B->Type, "CG") != 0) fatal(fn, "Invalid Box_II->Type=%s\n", B->Type);
  const int n = B->M->row_dim;
  run_or_die(fn, [&] {
    auto A = upload_csr(n, B->M->Row, B->M->Col, B->M->Val, sym);
    DevVec v(n);
    v.up(B->Eig_Vec);
    SolveOut out;
    inverse_iteration(op_of(*A), B->Eig_Val[0], v.d.p, B->ii_acc, B->ii_diag_add, B->ii_max_step, B->cg_acc,
                      B->cg_max_step, out, mr);
    v.down(B->Eig_Vec);
    B->error = out.err;
    if (out.ii_converged) {
      sml_append("Inv_iter_step.txt", "%d\n", out.ii_steps);
    } else {
      printf("Error in %s\n", fn);
      printf("Does not converge(%e > %e) (max step is %d)\n", out.err, B->ii_acc, B->ii_max_step);
      printf("continue...\n");
      sml_append("error.txt", "Error in %s\nDoes not converge(%e > %e) (max step is %d)\ncontinue...\n", fn,
                 out.err, B->ii_acc, B->ii_max_step);
    }
  });
}

// CONJUGATE_GRADIENT(_SYM) and MINIMUM_RESIDUAL(_SYM) on a BOX_CG
void cg_entry(const char *fn, BOX_CG *B, bool sym, bool mr = false) {
  check_square(fn, B->M);
  const int n = B->M->row_dim;
  run_or_die(fn, [&] {
    auto A = upload_csr(n, B->M->Row, B->M->Col, B->M->Val, sym);
    DevVec b(n), x(n);
    b.up(B->Vec);
    x.up(B->Out_Vec);
    SolveOut out;
    bool conv = false;
    int it = mr ? mr_solve(op_of(*A), 0.0, b.d.p, x.d.p, B->acc, B->max_step, &conv, out)
                : cg_solve(op_of(*A), 0.0, b.d.p, x.d.p, B->acc, B->max_step, &conv, out);
    x.down(B->Out_Vec);
    const char *sf = mr ? "MINIMUM_RESIDUAL_Step.txt"
                        : sym ? "CONJUGATE_GRADIENT_SYM_Step.txt" : "CONJUGATE_GRADIENT_Step.txt";
    if (!conv) {
      printf("Error in %s\nDoes not converge (max step is %d)\nContinue...\n", fn, B->max_step);
      sml_append("error.txt", "Error in %s\nDoes not converge (max step is %d)\nContinue...\n", fn, B->max_step);
    }
    sml_append(sf, "%d\n", it);
  });
}

double check_entry(const char *fn, CRS1 *M, double *v, double eig, bool sym) {
  double r = 0.0;
  run_or_die(fn, [&] {
    auto A = upload_csr(M->row_dim, M->Row, M->Col, M->Val, sym);
    DevVec d(M->row_dim);
    d.up(v);
    r = converge_check(op_of(*A), d.d.p, eig);
  });
  return r;
}

void diagonalize_entry(const char *fn, CRS1 *Ham, DMRG_TIME *Time, DMRG_PARAMETER *Param, DMRG_STATUS *St,
                       bool sym) {
  const bool lan = strcmp(Param->Diag_Method, "Lanczos") == 0;
  const bool slow = strcmp(Param->Diag_Method, "Lanczos_Slow") == 0;
  if (!lan && !slow) fatal("DMRG_DIAGONALIZE_HAMILTONIAN", "%s", "");
  int type;
  if (strcmp(Param->Lan_Con, "Normal") == 0) type = 0;
  else if (strcmp(Param->Lan_Con, "Tight") == 0) type = 1;
  else fatal(slow ? "LANCZOS_SLOW_SYM" : "LANCZOS_SYM", "Box->Type=%s is error\n", Param->Lan_Con);
  const bool mr = strcmp(Param->II_Type, "MR") == 0;
  if (!mr && strcmp(Param->II_Type, "CG") != 0)
    fatal("INVERSE_ITERATION_SYM", "Invalid Box_II->Type=%s\n", Param->II_Type);
  check_square(slow ? "LANCZOS_SLOW_SYM" : "LANCZOS_SYM", Ham);
  if (Param->diag_max_step <= Param->diag_min_step)
    fatal(slow ? "LANCZOS_SLOW_SYM" : "LANCZOS_SYM", "Box->max_step=%d,Box->min_step=%d\n",
          Param->diag_max_step, Param->diag_min_step);
  const int n = Ham->row_dim;
  run_or_die(fn, [&] {
    double t0 = now_s();
    std::unique_ptr<dmrgx_matrix> A(sym ? take_pending(Ham) : nullptr);
    if (!A) A = upload_csr(n, Ham->Row, Ham->Col, Ham->Val, sym);
    DevVec v(n);
    SolveOut out;
    double eig;
    if (n == 1) {
      eig = Ham->Val[0];
      double one = 1.0;
      v.up(&one);
    } else {
      launch_start_vector(n, ctx().seed, v.d.p, ctx().stream);  // Guess is forced to "No" (:24)
      LanczosOpts o;
      o.acc = Param->diag_acc;
      o.min_step = Param->diag_min_step;
      o.max_step = Param->diag_max_step;
      o.max_block_step = slow ? 0 : Param->diag_max_step;
      o.type = type;
      lanczos(op_of(*A), o, v.d.p, out);
      if (!out.lan_converged) {
        const char *ln = slow ? "LANCZOS_SLOW_SYM" : "LANCZOS_SYM";
        sml_append("error.txt", "Error in %s\nDoes not converge (max step is %d)\n", ln, Param->diag_max_step);
        fatal(ln, "Does not converge (max step is %d)\n", Param->diag_max_step);
      }
      eig = out.eig;
      sml_append(slow ? "LANCZOS_SLOW_SYM_Step.txt" : "LANCZOS_SYM_Step.txt", "%d\n", out.lan_steps);
    }
    double t1 = now_s();
    inverse_iteration(op_of(*A), eig, v.d.p, Param->inv_iter_acc, Param->inv_iter_diag_add, Param->inv_iter_max_step,
                      Param->cg_acc, Param->cg_max_step, out, mr);
    v.down(St->GS_Vec);
    double t2 = now_s();
    if (out.ii_converged) {
      sml_append("Inv_iter_step.txt", "%d\n", out.ii_steps);
    } else {
      printf("Error in INVERSE_ITERATION_SYM\nDoes not converge(%e > %e) (max step is %d)\ncontinue...\n", out.err,
             Param->inv_iter_acc, Param->inv_iter_max_step);
      sml_append("error.txt", "Error in INVERSE_ITERATION_SYM\nDoes not converge(%e > %e) (max step is %d)\ncontinue...\n",
                 out.err, Param->inv_iter_acc, Param->inv_iter_max_step);
    }
    St->gs_val = eig;       // the Lanczos eigenvalue (DMRG_DIAGONALIZE_SUPERBLOCK_SYM.c:56)
    St->gs_error = out.err;
    if (Time) {
      Time->diag = t1 - t0;
      Time->inv_iter = t2 - t1;
    }
  });
}

}  // namespace

namespace dmrgx {
std::unique_ptr<dmrgx_matrix> upload_csr(int32_t n, const int64_t *row, const int32_t *col, const double *val,
                                         bool sym) {
  if (n <= 0) throw Failure(DMRGX_ERR_ARG, "matrix dimension must be positive");
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int64_t nnz = row[n] - row[0];
  if (row[0] != 0 || nnz < 0) throw Failure(DMRGX_ERR_ARG, "CRS Row[0] must be 0 and Row non-decreasing");
  // host checks before anything reaches the device: the kernels index x and the output
  // rows with these arrays
  for (int32_t r = 0; r < n; r++)
    if (row[r + 1] < row[r]) throw Failure(DMRGX_ERR_ARG, "CRS Row must be non-decreasing");
  if (nnz > 0 && (!col || !val)) throw Failure(DMRGX_ERR_ARG, "CRS Col/Val missing");
  for (int64_t j = 0; j < nnz; j++)
    if (col[j] < 0 || col[j] >= n) throw Failure(DMRGX_ERR_ARG, "CRS column index out of range");
  auto A = std::make_unique<dmrgx_matrix>();
  if (sym) {
    DevBuf<int64_t> tr((size_t)n + 1);
    DevBuf<int32_t> tc(nnz);
    DevBuf<double> tv(nnz);
    ctx().copy_in(tr.p, row, sizeof(int64_t) * ((size_t)n + 1));
    ctx().copy_in(tc.p, col, sizeof(int32_t) * nnz);
    ctx().copy_in(tv.p, val, sizeof(double) * nnz);
    launch_expand_lower(n, nnz, tr.p, tc.p, tv.p, *A, st);
  } else {
    A->n = n;
    A->nnz = nnz;
    A->row.resize((size_t)n + 1);
    A->col.resize(nnz);
    A->val.resize(nnz);
    ctx().copy_in(A->row.p, row, sizeof(int64_t) * ((size_t)n + 1));
    ctx().copy_in(A->col.p, col, sizeof(int32_t) * nnz);
    ctx().copy_in(A->val.p, val, sizeof(double) * nnz);
    HIPCHK(hipStreamSynchronize(st));
  }
  default_schedule(*A);
  return A;
}
}  // namespace dmrgx

// ================================================================ Part 1
extern "C" {

void MATRIX_VECTOR_PRODUCT_SYM(CRS1 *M, double *V1, double *Out, double **Temp, int p_threads) {
  (void)Temp;
  (void)p_threads;
  run_or_die("MATRIX_VECTOR_PRODUCT_SYM", [&] {
    auto A = upload_csr(M->row_dim, M->Row, M->Col, M->Val, true);
    DevVec x(M->row_dim), y(M->row_dim);
    x.up(V1);
    launch_spmv(op_of(*A), x.d.p, y.d.p, ctx().stream);
    y.down(Out);
  });
}

void MATRIX_VECTOR_PRODUCT(CRS1 *M, double *V1, double *Out, int p_threads) {
  (void)p_threads;
  run_or_die("MATRIX_VECTOR_PRODUCT", [&] {
    auto A = upload_csr(M->row_dim, M->Row, M->Col, M->Val, false);
    DevVec x(M->row_dim), y(M->row_dim);
    x.up(V1);
    launch_spmv(op_of(*A), x.d.p, y.d.p, ctx().stream);
    y.down(Out);
  });
}

void LANCZOS_SLOW_SYM(BOX_LAN *Box) { lanczos_entry("LANCZOS_SLOW_SYM", "LANCZOS_SLOW_SYM_Step.txt", Box, true, true); }
void LANCZOS_SYM(BOX_LAN *Box) { lanczos_entry("LANCZOS_SYM", "LANCZOS_SYM_Step.txt", Box, true, false); }
void LANCZOS_SLOW(BOX_LAN *Box) { lanczos_entry("LANCZOS_SLOW", "LANCZOS_SLOW_Step.txt", Box, false, true); }
void LANCZOS(BOX_LAN *Box) { lanczos_entry("LANCZOS", "LANCZOS_Step.txt", Box, false, false); }

void INVERSE_ITERATION_SYM(BOX_II *B) { ii_entry("INVERSE_ITERATION_SYM", B, true); }
void INVERSE_ITERATION(BOX_II *B) { ii_entry("INVERSE_ITERATION", B, false); }
void CONJUGATE_GRADIENT_SYM(BOX_CG *B) { cg_entry("CONJUGATE_GRADIENT_SYM", B, true); }
void CONJUGATE_GRADIENT(BOX_CG *B) { cg_entry("CONJUGATE_GRADIENT", B, false); }
void MINIMUM_RESIDUAL_SYM(BOX_CG *B) { cg_entry("MINIMUM_RESIDUAL_SYM", B, true, true); }
void MINIMUM_RESIDUAL(BOX_CG *B) { cg_entry("MINIMUM_RESIDUAL", B, false, true); }

double CONVERGE_CHECK_SYM(CRS1 *M, double *Eig_Vec, double eig_val, int p_threads) {
  (void)p_threads;
  return check_entry("CONVERGE_CHECK_SYM", M, Eig_Vec, eig_val, true);
}
double CONVERGE_CHECK(CRS1 *M, double *Eig_Vec, double eig_val, int p_threads) {
  (void)p_threads;
  return check_entry("CONVERGE_CHECK", M, Eig_Vec, eig_val, false);
}

void DMRG_DIAGONALIZE_SUPERBLOCK_SYM(CRS1 *Ham, DMRG_TIME *Time, DMRG_PARAMETER *Param, DMRG_STATUS *St,
                                     int p_threads) {
  (void)p_threads;
  diagonalize_entry("DMRG_DIAGONALIZE_SUPERBLOCK_SYM", Ham, Time, Param, St, true);
}
void DMRG_DIAGONALIZE_SUPERBLOCK(CRS1 *Ham, DMRG_TIME *Time, DMRG_PARAMETER *Param, DMRG_STATUS *St,
                                 int p_threads) {
  (void)p_threads;
  diagonalize_entry("DMRG_DIAGONALIZE_SUPERBLOCK", Ham, Time, Param, St, false);
}

// ================================================================ Part 2
const char *dmrgx_last_error(void) { return last_error().c_str(); }

int dmrgx_init(int device) {
  return run_status([&] {
    if (device >= 0) HIPCHK(hipSetDevice(device));
    ensure_init();
  });
}

void dmrgx_set_seed(uint64_t seed) { ctx().seed = seed; }

static uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void dmrgx_start_vector(int64_t n, uint64_t seed, double *out) {
  for (int64_t i = 0; i < n; i++) {
    const uint64_t h = splitmix64(seed * 0xD1B54A32D192ED03ULL + (uint64_t)i);
    out[i] = (double)(int64_t)(h % 100000000ULL) - 50000000.0;
  }
}

int dmrgx_tridiag_lowest(int32_t k, const double *a, const double *b, double *lam_out, double *vec_out) {
  if (k < 1 || !a || (k > 1 && !b) || !lam_out) return DMRGX_ERR_ARG;
  RitzTrack track;
  for (int32_t j = 0; j < k; j++) lam_out[j] = track.next(j + 1, a, b);
  if (vec_out) tridiag_vector(k, a, b, lam_out[k - 1], vec_out);
  return DMRGX_OK;
}

int dmrgx_matrix_from_crs1(const CRS1 *M, int sym, dmrgx_matrix **out) {
  if (!M || !out) return DMRGX_ERR_ARG;
  return dmrgx_matrix_from_arrays(M->row_dim, M->Row, M->Col, M->Val, sym, out);
}

int dmrgx_matrix_from_arrays(int32_t n, const int64_t *row, const int32_t *col, const double *val, int sym,
                             dmrgx_matrix **out) {
  if (!row || !out || n <= 0) return DMRGX_ERR_ARG;
  return run_status([&] { *out = upload_csr(n, row, col, val, sym != 0).release(); });
}

void dmrgx_matrix_free(dmrgx_matrix *A) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  delete A;
}
int64_t dmrgx_matrix_nnz(const dmrgx_matrix *A) { return A ? A->nnz : -1; }
int32_t dmrgx_matrix_dim(const dmrgx_matrix *A) { return A ? A->n : -1; }

int dmrgx_matrix_download(const dmrgx_matrix *A, int64_t *row, int32_t *col, double *val) {
  if (!A) return DMRGX_ERR_ARG;
  return run_status([&] {
    hipStream_t st = ctx().stream;
    ctx().copy_out(row, A->row.p, sizeof(int64_t) * ((size_t)A->n + 1));
    ctx().copy_out(col, A->col.p, sizeof(int32_t) * A->nnz);
    ctx().copy_out(val, A->val.p, sizeof(double) * A->nnz);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int dmrgx_spmv_device(const dmrgx_matrix *A, const double *x_dev, double *y_dev) {
  if (!A) return DMRGX_ERR_ARG;
  return run_status([&] { launch_spmv(op_of(*A), x_dev, y_dev, ctx().stream); });
}

int dmrgx_matrix_set_schedule(dmrgx_matrix *A, int32_t lanes, int32_t grid_cap, int32_t load_variant) {
  if (!A || lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) || grid_cap < 0 || !valid_load(load_variant))
    return DMRGX_ERR_ARG;
  if (sliced_format(load_format(load_variant)) && lanes > kSellMaxLanes) return DMRGX_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  A->lanes = lanes;
  A->grid = grid_cap;
  A->load = load_variant;
  return DMRGX_OK;
}

int dmrgx_matrix_get_schedule(const dmrgx_matrix *A, int32_t *lanes, int32_t *grid_cap, int32_t *load_variant) {
  if (!A) return DMRGX_ERR_ARG;
  if (lanes) *lanes = A->lanes;
  if (grid_cap) *grid_cap = A->grid;
  if (load_variant) *load_variant = A->load;
  return DMRGX_OK;
}

// SpMV schedule autotuner: times y = A x with a fused dot product (HIP events on the library
// stream) for each power-of-two lanes-per-row count around the mean row length, each
// workgroup cap and each load variant (in the formats A can have),
// re-times the six fastest with more launches, and keeps the fastest.
namespace {
// Mean time of y = A x with a fused dot product under schedule (g, cap, v), as the solvers
// run their products (product, block partials, finish kernel), over reps launches after
// one warm-up launch; A's own schedule is restored.
static double time_schedule(dmrgx_matrix *A, int g, int cap, int v, int reps, const double *x, double *y, double *dot,
                     Reduce &R, hipEvent_t e0, hipEvent_t e1) {
  hipStream_t st = ctx().stream;
  const int g0 = A->lanes, c0 = A->grid, v0 = A->load;
  A->lanes = g;
  A->grid = cap;
  A->load = v;
  launch_spmv_dot(op_of(*A), x, y, dot, R, st);
  HIPCHK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++) launch_spmv_dot(op_of(*A), x, y, dot, R, st);
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  A->lanes = g0;
  A->grid = c0;
  A->load = v0;
  return ms * 1000.0 / reps;
}

struct TuneEnv {
  int32_t n;
  DevBuf<double> x, y, dot;
  Reduce R;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit TuneEnv(int32_t n_) : n(n_), x(n_), y(n_), dot(1) {
    launch_start_vector(n, 7, x.p, ctx().stream);
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
  }
  ~TuneEnv() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  double time(dmrgx_matrix *A, int g, int cap, int v, int reps) {
    return time_schedule(A, g, cap, v, reps, x.p, y.p, dot.p, R, e0, e1);
  }
  // Microseconds per (Lanczos step + CG iteration) run with the solver's fused kernels:
  // the product with its epilogue and the vector kernels that sum its block partials,
  // whose cost grows with the product's block count.  The numbers computed do not matter
  // (the CG runs on the unshifted operator with a zero tolerance, so it never stops).
  double time_steps(dmrgx_matrix *A, int g, int cap, int v, int reps) {
    hipStream_t st = ctx().stream;
    if (w.count < (size_t)n) {
      for (DevBuf<double> *b : {&w, &va, &vb, &ly, &p, &q, &xw, &r}) b->resize(n);
      status.resize(2);
    }
    const size_t ns = (size_t)reps + 2;
    if (sc.count < 4 * ns + 1) sc.resize(4 * ns + 1);
    double *alpha = sc.p, *beta = sc.p + ns, *rr = sc.p + 2 * ns, *pap = sc.p + 3 * ns, *nrm2 = sc.p + 4 * ns;
    const int g0 = A->lanes, c0 = A->grid, v0 = A->load;
    A->lanes = g;
    A->grid = cap;
    A->load = v;
    const Operator op = op_of(*A);
    launch_copy(n, x.p, w.p, st);
    launch_copy(n, x.p, p.p, st);
    launch_copy(n, x.p, r.p, st);
    launch_norm2(n, w.p, R, nrm2, st);
    launch_norm2(n, r.p, R, rr, st);
    HIPCHK(hipMemsetAsync(status.p, 0, 2 * sizeof(int), st));
    HIPCHK(hipEventRecord(e0, st));
    int nbl = 0;
    for (int k = 1; k <= reps; k++) {
      double *vk = (k & 1) ? va.p : vb.p, *vkm1 = (k & 1) ? vb.p : va.p;
      const int na = launch_lanczos_a_part(op, false, k, w.p, nrm2, beta, k == 1 ? nullptr : LB.partial.p,
                                           k == 1 ? 0 : nbl, vk, ly.p, alpha, LA, st);
      nbl = launch_lanczos_b_part(n, k, ly.p, vk, vkm1, alpha, beta, LA.partial.p, na, w.p, LB, st);
      const int nc = launch_cg_ap_part(op, 0.0, p.p, q.p, status.p, CA, st);
      const int nu = launch_cg_update_part(n, k, p.p, q.p, xw.p, r.p, pap, rr, status.p, CA.partial.p, nc, CB, st);
      launch_cg_dir_part(n, k, 0.0, r.p, rr, p.p, status.p, CB.partial.p, nu, st);
    }
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    A->lanes = g0;
    A->grid = c0;
    A->load = v0;
    return ms * 1000.0 / reps;
  }

 private:
  DevBuf<double> w, va, vb, ly, p, q, xw, r, sc;
  DevBuf<int> status;
  Reduce LA, LB, CA, CB;
};
}  // namespace

int dmrgx_matrix_time_schedule(dmrgx_matrix *A, int32_t lanes, int32_t grid_cap, int32_t load_variant,
                               int32_t reps, double *us_out) {
  if (!A || !us_out || lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) || grid_cap < 0 ||
      !valid_load(load_variant) || reps < 1)
    return DMRGX_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  return run_status([&] {
    TuneEnv env(A->n);
    *us_out = env.time(A, lanes, grid_cap, load_variant, reps);
  });
}

int dmrgx_matrix_dict_values(dmrgx_matrix *A, int64_t *n_values) {
  if (!A || !n_values) return DMRGX_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  return run_status([&] {
    ensure_dict(*A);
    *n_values = std::max<int64_t>(A->n_values, 0);
  });
}

int dmrgx_matrix_sell16_exceptions(dmrgx_matrix *A, int32_t lanes, int64_t *exceptions) {
  if (!A || !exceptions || (lanes != 1 && lanes != 2 && lanes != 4)) return DMRGX_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  return run_status([&] {
    ensure_sell(*A, lanes);
    *exceptions = ensure_sell16(*A, lanes);
  });
}

int dmrgx_matrix_autotune(dmrgx_matrix *A, int32_t *lanes_out, int32_t *grid_out, int32_t *load_out,
                          double *best_us_out) {
  if (!A) return DMRGX_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  return run_status([&] {
    TuneEnv env(A->n);
    const int base = pick_lanes(A->n, A->nnz);
    const int caps[] = {1024, 2048, 4096, 8192, 16384, 0};
    const int g_lo = std::max(1, base / 16), g_hi = std::min(64, base * 2);
    struct Cand {
      int g, cap, v;
      double us;
    };
    std::vector<Cand> cands;
    const bool has[3] = {true, ensure_dict(*A), ensure_delta(*A)};
    const bool has_idx = has[1] && A->n_values <= kSellIdxMaxValues;
    for (int v = 0; v < kLoadVariants; v++) {
      if (sliced_format(load_format(v))) {
        // sliced form: 1, 2 or 4 lanes per row, 32- or 16-bit columns, plain or
        // non-temporal matrix loads, plain or XCD-contiguous row ranges; with value
        // indices (bit 0 ignored) where the matrix has a small value table
        if (load_format(v) == kFmtSellIdx && ((v & kLoadPair) || !has_idx)) continue;
        for (int g = 1; g <= kSellMaxLanes; g *= 2)
          for (int cap : caps) {
            cands.push_back({g, cap, v, env.time(A, g, cap, v, 3)});
            // 16-bit columns: also with the matrix loads one pass group ahead
            if (load_format(v) == kFmtSell && (v & kLoadPair))
              cands.push_back({g, cap, v | kLoadPipe, env.time(A, g, cap, v | kLoadPipe, 3)});
          }
        continue;
      }
      if (has[load_format(v)])
        for (int g = g_lo; g <= g_hi; g *= 2)
          for (int cap : caps) cands.push_back({g, cap, v, env.time(A, g, cap, v, 3)});
    }
    // the leaders again, timed as the solver runs them: a Lanczos step and a CG iteration
    // with their fused vector kernels, several times over
    std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.us < b.us; });
    cands.resize(std::min<size_t>(cands.size(), 8));
    for (Cand &c : cands) c.us = env.time_steps(A, c.g, c.cap, c.v, 12);
    const Cand best = *std::min_element(cands.begin(), cands.end(),
                                        [](const Cand &a, const Cand &b) { return a.us < b.us; });
    A->lanes = best.g;
    A->grid = best.cap;
    A->load = best.v;
    // the sliced forms built for the candidates, except the one the schedule reads
    drop_sell_except(*A, sliced_format(load_format(best.v)) ? best.g : 0);
    if (lanes_out) *lanes_out = best.g;
    if (grid_out) *grid_out = best.cap;
    if (load_out) *load_out = best.v;
    if (best_us_out) *best_us_out = best.us;
  });
}

int dmrgx_spmv_host(const dmrgx_matrix *A, const double *x, double *y) {
  if (!A) return DMRGX_ERR_ARG;
  return run_status([&] {
    DevVec dx(A->n), dy(A->n);
    dx.up(x);
    launch_spmv(op_of(*A), dx.d.p, dy.d.p, ctx().stream);
    dy.down(y);
  });
}

void dmrgx_default_params(dmrgx_solve_params *p) {
  memset(p, 0, sizeof(*p));
  p->diag_acc = 1e-14;     // main/dmrg/hubbard_vf/src/main/main.c:40-42
  p->diag_min_step = 1;
  p->diag_max_step = 800;
  p->lan_type = 0;
  p->guess = 0;
  p->seed = 1;
  p->ii_acc = 1e-9;        // :47-49
  p->ii_diag_add = 1e-11;
  p->ii_max_step = 3;
  p->cg_acc = 1e-7;        // :44-45
  p->cg_max_step = 2000;
}

}  // extern "C"

namespace {
// The ground-state solve on an operator: an assembled matrix A (op = op_of(*A)) or the
// Kronecker form (A null: FP64 only).
int ground_state_op(const Operator &op, const dmrgx_matrix *A, const dmrgx_solve_params *p, double *vec_inout,
                    dmrgx_solve_stats *stats) {
  if (!p || !vec_inout) return DMRGX_ERR_ARG;
  if (p->diag_max_step <= p->diag_min_step) return DMRGX_ERR_ARG;
  return run_status([&] {
    const double t0 = now_s();
    hipStream_t st = ctx().stream;
    DevVec v(op.n);
    if (p->guess) v.up(vec_inout);
    else launch_start_vector(op.n, p->seed, v.d.p, st);
    SolveOut out;
    LanczosOpts o;
    o.acc = p->diag_acc;
    o.min_step = p->diag_min_step;
    o.max_step = p->diag_max_step;
    o.type = p->lan_type;
    if (p->lan_type < 0 || p->lan_type > 2) throw Failure(DMRGX_ERR_ARG, "lan_type must be 0 (Normal), 1 (Tight) or 2 (Residual)");
    if (p->lan_type == 2) {
      // "Residual": stop on the Ritz pair's residual bound, at which the L1 residual test
      // (at most sqrt(n) times the 2-norm) passes with room to spare
      o.type = 0;
      o.res_tol = 0.5 * p->ii_acc / std::sqrt((double)std::max<int32_t>(op.n, 1));
      o.res_abs = true;
    }
    const bool mixed = p->precision == 1;
    const bool two_phase = p->precision == 2;
    if (p->precision != 0 && !mixed && !two_phase)
      throw Failure(DMRGX_ERR_ARG,
                    "precision must be 0 (FP64), 1 (FP32 + FP64 refinement) or 2 (FP32 then FP64 Lanczos)");
    if ((mixed || two_phase) && !A) throw Failure(DMRGX_ERR_ARG, "the Kronecker form solves in FP64 only (precision 0)");
    if (mixed || two_phase) {
      ensure_fp32(*A);
      o.fp32 = true;
    }
    if (two_phase && p->lan_type == 2) throw Failure(DMRGX_ERR_ARG, "precision 2 has its own stopping rule (lan_type 0)");
    if (two_phase) o.res_tol = p->mixed_switch > 0.0 ? p->mixed_switch : 1e-3;
    if (op.n == 1) {
      double h;
      if (A) {
        HIPCHK(hipStreamSynchronize(st));   // the values may still be written on the library stream
        HIPCHK(hipMemcpy(&h, A->val.p, sizeof(double), hipMemcpyDeviceToHost));
      } else {   // the 1 x 1 operator applied to 1
        const double one = 1.0;
        v.up(&one);
        op.kron->product(v.d.p, st);
        HIPCHK(hipStreamSynchronize(st));   // the library's streams do not block on the null stream
        double hl, hc;
        HIPCHK(hipMemcpy(&h, op.kron->z.p, sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&hl, op.kron->zl.p, sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&hc, op.kron->zc.p, sizeof(double), hipMemcpyDeviceToHost));
        h = (h + hl) + hc;
      }
      out.eig = h;
      const double one = 1.0;
      v.up(&one);
      out.lan_converged = true;
    } else {
      lanczos(op, o, v.d.p, out);
    }
    if (!out.lan_converged) throw Failure(DMRGX_ERR_NOCONV, "Lanczos did not converge within diag_max_step");
    int fp32_steps = (mixed || two_phase) ? out.lan_steps : 0;
    if (two_phase && op.n > 1) {
      // FP64 Lanczos from the FP32 phase's Ritz vector, to a residual bound at which the
      // L1 residual test (at most sqrt(n) times the 2-norm) passes with room to spare
      LanczosOpts o2 = o;
      o2.fp32 = false;
      o2.res_tol = 0.5 * p->ii_acc / std::sqrt((double)op.n) / std::max(1.0, std::fabs(out.eig));
      SolveOut out2;
      lanczos(op, o2, v.d.p, out2);
      if (!out2.lan_converged) throw Failure(DMRGX_ERR_NOCONV, "Lanczos did not converge within diag_max_step");
      out.eig = out2.eig;
      out.lan_steps += out2.lan_steps;
      out.spmv += out2.spmv;
    }
    // mixed precision: the FP32 Lanczos value is off by the FP32 rounding of the matrix; the
    // inverse iteration's shift and its residual test take the FP64 Rayleigh quotient of the
    // Ritz vector instead (exact to the square of the vector's error)
    if (mixed && op.n > 1 && !p->skip_ii) {
      out.eig = rayleigh_quotient(op, v.d.p);
      out.spmv++;
    }
    if (!p->skip_ii)
      inverse_iteration(op, out.eig, v.d.p, p->ii_acc, p->ii_diag_add, p->ii_max_step, p->cg_acc, p->cg_max_step,
                        out, false, mixed || two_phase);
    // mixed precision: the FP32 Lanczos value is only a shift for the FP64 refinement;
    // the energy reported is the FP64 Rayleigh quotient of the refined vector (in the
    // two-phase form only if inverse iteration changed the FP64 Lanczos vector)
    if ((mixed || (two_phase && out.cg_steps > 0)) && op.n > 1) {
      out.eig = rayleigh_quotient(op, v.d.p);
      out.spmv++;
    }
    v.down(vec_inout);
    if (stats) {
      stats->eig_val = out.eig;
      stats->gs_error = out.err;
      stats->lanczos_steps = out.lan_steps;
      stats->ii_steps = out.ii_steps;
      stats->cg_steps = out.cg_steps;
      stats->spmv_calls = out.spmv;
      stats->seconds = now_s() - t0;
      stats->lanczos_fp32_steps = fp32_steps;
    }
  });
}

}  // namespace

extern "C" {

int dmrgx_ground_state(const dmrgx_matrix *A, const dmrgx_solve_params *p, double *vec_inout,
                       dmrgx_solve_stats *stats) {
  if (!A) return DMRGX_ERR_ARG;
  return ground_state_op(op_of(*A), A, p, vec_inout, stats);
}

int dmrgx_kron_ground_state(const dmrgx_kron *K, const dmrgx_solve_params *p, double *vec_inout,
                            dmrgx_solve_stats *stats) {
  if (!K) return DMRGX_ERR_ARG;
  return ground_state_op(op_of(kron_dev(K)), nullptr, p, vec_inout, stats);
}

int dmrgx_profile_begin_sampled(int32_t every) {
  if (every < 1) return DMRGX_ERR_ARG;
  return run_status([&] {
    Context &C = ctx();
    HIPCHK(hipStreamSynchronize(C.stream));
    C.prof = true;
    C.prof_used = 0;
    C.prof_bytes = 0.0;
    C.prof_every = every;
    C.prof_seen = 0;
  });
}

int dmrgx_profile_begin(void) { return dmrgx_profile_begin_sampled(1); }

int dmrgx_profile_end(double *spmv_ms, int64_t *spmv_launches, double *spmv_alg_bytes) {
  return run_status([&] {
    Context &C = ctx();
    HIPCHK(hipStreamSynchronize(C.stream));
    double tot = 0.0;
    for (size_t i = 0; i < C.prof_used; i++) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, C.prof_ev[i].first, C.prof_ev[i].second));
      tot += ms;
    }
    if (spmv_ms) *spmv_ms = tot;
    if (spmv_launches) *spmv_launches = (int64_t)C.prof_used;
    if (spmv_alg_bytes) *spmv_alg_bytes = C.prof_bytes;
    C.prof = false;
    C.prof_used = 0;
    C.prof_bytes = 0.0;
  });
}

int dmrgx_copy_device(const double *src, double *dst, int64_t count) {
  return run_status([&] {
    launch_copy(count, src, dst, ctx().stream);
    HIPCHK(hipStreamSynchronize(ctx().stream));
  });
}

int dmrgx_device_alloc(int64_t bytes, void **out) {
  return run_status([&] { HIPCHK(hipMalloc(out, (size_t)bytes)); });
}

int dmrgx_device_free(void *p) {
  return run_status([&] { HIPCHK(hipFree(p)); });
}

int dmrgx_copy_to_device(void *dst_dev, const void *src_host, int64_t bytes) {
  if ((!dst_dev || !src_host) && bytes > 0) return DMRGX_ERR_ARG;
  return run_status([&] {
    if (bytes > 0) HIPCHK(hipMemcpy(dst_dev, src_host, (size_t)bytes, hipMemcpyHostToDevice));
  });
}

int dmrgx_copy_to_host(void *dst_host, const void *src_dev, int64_t bytes) {
  if ((!dst_host || !src_dev) && bytes > 0) return DMRGX_ERR_ARG;
  return run_status([&] {
    if (bytes > 0) HIPCHK(hipMemcpy(dst_host, src_dev, (size_t)bytes, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
