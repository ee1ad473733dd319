// Operator renormalisation: CRS_CRS_CCS_PRODUCT (sml/CRS_CRS_CCS_PRODUCT.c:6-124), the
// product T^dagger O T that DMRG_TRANS_MAT_ONE/TWO/C_C apply to every block operator
// (dmrg/DMRG_TRANS_MAT_ONE.c:6-54 through DMRG_EXTEND_AND_REDUCE).
//
// The reference forms Out = M1 (M2 M3) in two passes, each entry a sum in a fixed order
// with entries of magnitude <= 1e-15 dropped:
//   pass 1  W(i, j)  = sum over the entries k of row i of M2, in stored order,
//                      of M3(col_k, j) * M2_k
//   pass 2  Out(i, j) = sum over the kept rows r of column j of W, ascending,
//                      of M1(i, r) * W(r, j)
// Here M3 is expanded to a dense device matrix, pass 1 is a CRS x dense product (one thread
// per entry of W, the row's entries in stored order, dropped entries set to 0), and pass 2
// is a CRS x dense product over the rows of M1 with their entries in ascending column
// order, skipping the zeros of W: the same nonzero terms in the same order as the
// reference's (the terms the reference adds for M1(i, r) = 0 are signed zeros, which do not
// change a sum).  The file is built without FMA contraction, so Out is the reference's bit
// for bit.  The dense result is compacted into the caller's CRS1 on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "entry.h"

namespace dmrgx {
namespace {

constexpr double kDropBelow = 1e-15;

// dense[r][j] = M3(r, j) from CCS (col_ptr over columns, row indices per entry)
__global__ __launch_bounds__(256) void k_ccs_to_dense(int32_t cols, int32_t rows, const int64_t *__restrict__ cptr,
                                                      const int32_t *__restrict__ ridx,
                                                      const double *__restrict__ val, double *__restrict__ dense) {
  const int32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= cols) return;
  for (int64_t e = cptr[j]; e < cptr[j + 1]; e++) dense[(int64_t)ridx[e] * cols + j] = val[e];
}

// out[i][j] = sum over the entries e of row i of A (stored order) of B[col_e][j] * A_e,
// dropped to 0 when |sum| <= 1e-15; with skip_zero, terms whose B value is 0 are left out.
// With kept set, the number of kept entries other than the last one, (rows-1, cols-1), is
// added to *kept (the count the reference's capacity check sees, CRS_CRS_CCS_PRODUCT.c:56).
template <bool kSkipZero>
__global__ __launch_bounds__(256) void k_crs_dense(int32_t rows, int32_t cols, const int64_t *__restrict__ rptr,
                                                   const int32_t *__restrict__ cidx, const double *__restrict__ aval,
                                                   const double *__restrict__ B, double *__restrict__ out,
                                                   unsigned long long *kept) {
  const int32_t j = blockIdx.x * 256 + threadIdx.x;
  unsigned int mine = 0;
  if (j < cols) {
    for (int32_t i = blockIdx.y; i < rows; i += gridDim.y) {
      double s = 0.0;
      for (int64_t e = rptr[i]; e < rptr[i + 1]; e++) {
        const double b = B[(int64_t)cidx[e] * cols + j];
        if (kSkipZero) {
          if (b != 0.0) s = s + b * aval[e];
        } else {
          s = s + b * aval[e];
        }
      }
      const bool keep = std::fabs(s) > kDropBelow;
      out[(int64_t)i * cols + j] = keep ? s : 0.0;
      if (keep && !(i == rows - 1 && j == cols - 1)) mine++;
    }
  }
  if (!kept) return;   // uniform over the launch
  __shared__ unsigned int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(kept, (unsigned long long)cnt);
}

void check_crs(const CRS1 *M, const char *what) {
  if (!M || !M->Row || M->row_dim < 0 || M->col_dim < 0) throw Failure(DMRGX_ERR_ARG, what);
  const int64_t nnz = M->Row[M->row_dim];
  if (nnz > 0 && (!M->Col || !M->Val)) throw Failure(DMRGX_ERR_ARG, what);
  for (int32_t r = 0; r < M->row_dim; r++)
    if (M->Row[r] < 0 || M->Row[r] > M->Row[r + 1]) throw Failure(DMRGX_ERR_ARG, what);
  for (int64_t e = 0; e < nnz; e++)
    if (M->Col[e] < 0 || M->Col[e] >= M->col_dim) throw Failure(DMRGX_ERR_ARG, what);
}

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

extern "C" {

void CRS_CRS_CCS_PRODUCT(CRS1 *M1, CRS1 *M2, CCS1 *M3, CRS1 *Out, CCS1 *Work, double *Work_Vec) {
  (void)Work_Vec;
  run_or_die("CRS_CRS_CCS_PRODUCT", [&] {
    if (!M1 || !M2 || !M3 || !Out || !Work) throw Failure(DMRGX_ERR_ARG, "NULL argument");
    // the reference's checks (sml/CRS_CRS_CCS_PRODUCT.c:8-28)
    if (M1->col_dim != M2->row_dim || M2->col_dim != M3->row_dim)
      throw Failure(DMRGX_ERR_ARG, "matrix dimensions do not chain");
    if (Work->max_col <= M3->col_dim) throw Failure(DMRGX_ERR_ARG, "Work has too few columns");
    if (Out->max_row <= M1->row_dim) throw Failure(DMRGX_ERR_ARG, "Out has too few rows");
    check_crs(M1, "M1 is not a valid CRS1");
    check_crs(M2, "M2 is not a valid CRS1");
    const int32_t n1 = M1->row_dim, n2 = M2->row_dim, n3 = M3->row_dim, k = M3->col_dim;
    if (!M3->Col || k < 0 || n3 < 0) throw Failure(DMRGX_ERR_ARG, "M3 is not a valid CCS1");
    const int64_t nnz3 = M3->Col[k];
    for (int32_t c = 0; c < k; c++)
      if (M3->Col[c] < 0 || M3->Col[c] > M3->Col[c + 1]) throw Failure(DMRGX_ERR_ARG, "M3 is not a valid CCS1");
    for (int64_t e = 0; e < nnz3; e++)
      if (M3->Row[e] < 0 || M3->Row[e] >= n3) throw Failure(DMRGX_ERR_ARG, "M3 is not a valid CCS1");
    Context &C = ctx();
    hipStream_t st = C.stream;
    // M1's rows in ascending column order (the order pass 2 adds them in); a column stored
    // twice in a row keeps its last value, as the reference's dense scatter of the row does
    const int64_t nnz2 = M2->Row[n2];
    std::vector<int64_t> hr1(n1 + 1, 0), perm;
    std::vector<int32_t> c1;
    std::vector<double> v1;
    for (int32_t r = 0; r < n1; r++) {
      const int64_t b = M1->Row[r], e = M1->Row[r + 1];
      perm.resize(e - b);
      std::iota(perm.begin(), perm.end(), b);
      std::stable_sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) { return M1->Col[x] < M1->Col[y]; });
      for (int64_t t = 0; t < e - b; t++) {
        if (t + 1 < e - b && M1->Col[perm[t + 1]] == M1->Col[perm[t]]) continue;
        c1.push_back(M1->Col[perm[t]]);
        v1.push_back(M1->Val[perm[t]]);
      }
      hr1[r + 1] = (int64_t)c1.size();
    }
    const int64_t nnz1 = (int64_t)c1.size();
    DevBuf<int64_t> r1(n1 + 1), r2(n2 + 1), p3(k + 1);
    DevBuf<int32_t> dc1(nnz1 > 0 ? nnz1 : 1), dc2(nnz2 > 0 ? nnz2 : 1), dr3(nnz3 > 0 ? nnz3 : 1);
    DevBuf<double> dv1(nnz1 > 0 ? nnz1 : 1), dv2(nnz2 > 0 ? nnz2 : 1), dv3(nnz3 > 0 ? nnz3 : 1);
    const int64_t nd3 = (int64_t)n3 * k, nw = (int64_t)n2 * k, no = (int64_t)n1 * k;
    DevBuf<double> d3(nd3 > 0 ? nd3 : 1), w(nw > 0 ? nw : 1), o(no > 0 ? no : 1);
    C.copy_in(r1.p, hr1.data(), sizeof(int64_t) * (n1 + 1));
    C.copy_in(r2.p, M2->Row, sizeof(int64_t) * (n2 + 1));
    C.copy_in(p3.p, M3->Col, sizeof(int64_t) * (k + 1));
    if (nnz1 > 0) {
      C.copy_in(dc1.p, c1.data(), sizeof(int32_t) * nnz1);
      C.copy_in(dv1.p, v1.data(), sizeof(double) * nnz1);
    }
    if (nnz2 > 0) {
      C.copy_in(dc2.p, M2->Col, sizeof(int32_t) * nnz2);
      C.copy_in(dv2.p, M2->Val, sizeof(double) * nnz2);
    }
    if (nnz3 > 0) {
      C.copy_in(dr3.p, M3->Row, sizeof(int32_t) * nnz3);
      C.copy_in(dv3.p, M3->Val, sizeof(double) * nnz3);
    }
    std::vector<double> ho(no > 0 ? no : 1);
    unsigned long long w_kept = 0;
    if (k > 0 && n1 > 0) {
      DevBuf<unsigned long long> dkept(1);
      HIPCHK(hipMemsetAsync(dkept.p, 0, sizeof(unsigned long long), st));
      if (nd3 > 0) HIPCHK(hipMemsetAsync(d3.p, 0, sizeof(double) * nd3, st));
      const unsigned gx = (unsigned)((k + 255) / 256);
      k_ccs_to_dense<<<gx, 256, 0, st>>>(k, n3, p3.p, dr3.p, dv3.p, d3.p);
      // grid y at most 65535; the kernels loop over the rows beyond it
      if (n2 > 0)
        k_crs_dense<false><<<dim3(gx, (unsigned)std::min(n2, 65535)), 256, 0, st>>>(n2, k, r2.p, dc2.p, dv2.p, d3.p,
                                                                                   w.p, dkept.p);
      k_crs_dense<true><<<dim3(gx, (unsigned)std::min(n1, 65535)), 256, 0, st>>>(n1, k, r1.p, dc1.p, dv1.p, w.p, o.p,
                                                                                nullptr);
      HIPCHK(hipGetLastError());
      C.copy_out(ho.data(), o.p, sizeof(double) * no);
      C.copy_out(&w_kept, dkept.p, sizeof(unsigned long long));
    } else {
      std::fill(ho.begin(), ho.end(), 0.0);
    }
    // The reference checks the kept count against the capacity before every entry it
    // evaluates (CRS_CRS_CCS_PRODUCT.c:56 for Work, :101 for Out), so it exits when the
    // entries kept before the last evaluated position reach max_val, even if nothing more
    // is stored after them.
    if (n2 > 0 && k > 0 && (long long)w_kept >= (long long)Work->max_val)
      throw Failure(DMRGX_ERR_ARG, "Work has too few elements");
    if (n1 > 0 && k > 0) {
      long long out_kept = 0;
      for (int64_t t = 0; t + 1 < no; t++) out_kept += ho[t] != 0.0;
      if (out_kept >= (long long)Out->max_val) throw Failure(DMRGX_ERR_ARG, "Out has too few elements");
    }
    // compaction into Out, columns ascending (sml/CRS_CRS_CCS_PRODUCT.c:98-121)
    int64_t e = 0;
    Out->Row[0] = 0;
    for (int32_t i = 0; i < n1; i++) {
      for (int32_t j = 0; j < k; j++) {
        const double v = ho[(int64_t)i * k + j];
        if (v != 0.0) {
          if (e >= Out->max_val) throw Failure(DMRGX_ERR_ARG, "Out has too few elements");
          Out->Val[e] = v;
          Out->Col[e] = j;
          e++;
        }
      }
      Out->Row[i + 1] = e;
    }
    Out->row_dim = n1;
    Out->col_dim = k;
  });
}

}  // extern "C"
