// Single-slot products from the state's sector into another sector:
// DMRG_V_M_{LL,LR,RR,RL}_Q1..Q4 (include/dmrg.h:369-382; dmrg/DMRG_V_M_LL_Q1.c:8-42,
// DMRG_V_M_LR_Q2.c:8-55, DMRG_V_M_RR_Q2.c:8-55, DMRG_V_M_RL_Q2.c:8-55 and the Q1/Q3/Q4
// siblings, which differ only in how the output sector is indexed).
//
//   Out[b] = sign_b * sum_{j in row q_slot(b) of M} V[Inv(q(b) with q_slot = col_j)] * M_j
//
// over the output sector's states b = (LL, LR, RR, RL), entries whose Inv is -1 skipped;
// sign_b = (-1)^(electrons of the slots before the acted one, in the order LL, LR, RR, RL)
// when the sign flag is set.  The reference's sector arrays and the whole-space inverse
// table carry neither the table's length nor the input vector's, so the table and the
// sectors are handed over once, with their sizes, to a dmrgx_whole_basis, and each product
// names its output sector by the id it got.  One thread per output state adds its row's
// entries in the reference's order, so the result is bit-identical to the reference loop.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "entry.h"
#include "reduce.h"

namespace dmrgx {
namespace {

struct SectorDev {
  int32_t n;
  DevBuf<int16_t> LL, LR, RR, RL;
};

__global__ __launch_bounds__(256) void k_v_m_sector(int32_t n_out, int32_t slot, int32_t dim_RR, int32_t d,
                                                    const int16_t *__restrict__ LL, const int16_t *__restrict__ LR,
                                                    const int16_t *__restrict__ RR, const int16_t *__restrict__ RL,
                                                    const int32_t *__restrict__ inv, const int64_t *__restrict__ row,
                                                    const int32_t *__restrict__ col, const double *__restrict__ val,
                                                    const int32_t *__restrict__ ele_LL,
                                                    const int32_t *__restrict__ ele_LR,
                                                    const int32_t *__restrict__ ele_RR, int32_t n_in,
                                                    const double *__restrict__ x, double *__restrict__ y,
                                                    int32_t *__restrict__ bad) {
  const int32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n_out) return;
  int q[4] = {LL[b], LR[b], RR[b], RL[b]};
  double sgn = 1.0;
  if (ele_LL) {
    int32_t e = 0;
    if (slot >= 1) e += ele_LL[q[0]];
    if (slot >= 2) e += ele_LR[q[1]];
    if (slot >= 3) e += ele_RR[q[2]];
    if (e % 2 != 0) sgn = -1.0;
  }
  const int own = q[slot];
  // the flat index with the acted slot's term left out: the gathered state of column c is
  // base + c * stride
  const int64_t stride = slot == 0 ? (int64_t)d * dim_RR * d : slot == 1 ? (int64_t)dim_RR * d : slot == 2 ? d : 1;
  q[slot] = 0;
  const int64_t base = (((int64_t)q[0] * d + q[1]) * dim_RR + q[2]) * d + q[3];
  double s = 0.0;
  bool out_of_range = false;
  int64_t j = row[own];
  const int64_t end = row[own + 1];
  // four entries' loads in flight at a time; the additions stay in entry order
  for (; j + 4 <= end; j += 4) {
    int32_t iv[4];
    double xv[4], vv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) iv[u] = inv[base + (int64_t)col[j + u] * stride];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      xv[u] = (iv[u] >= 0 && iv[u] < n_in) ? x[iv[u]] : 0.0;
      vv[u] = val[j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (iv[u] >= n_in) out_of_range = true;
      else if (iv[u] >= 0) s += sgn * xv[u] * vv[u];
    }
  }
  for (; j < end; j++) {
    const int32_t iv = inv[base + (int64_t)col[j] * stride];
    if (iv >= n_in) {
      out_of_range = true;
    } else if (iv >= 0) {
      s += sgn * x[iv] * val[j];
    }
  }
  y[b] = s;
  if (out_of_range) *bad = 1;
}

constexpr int kDotBlocks = 512;

// This block's share of sum_i a_i b_i (grid-stride, then the deterministic block sum).
__global__ __launch_bounds__(kBlock) void k_dot_partial(int64_t n, const double *__restrict__ a,
                                                        const double *__restrict__ b, double *__restrict__ partial) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += a[i] * b[i];
  block_partial(s, partial);
}

// *out = (accumulate ? *out : 0) + the partials' sum in block order.
__global__ __launch_bounds__(kBlock) void k_dot_finish(const double *__restrict__ partial, int nblk,
                                                       double *__restrict__ out, int accumulate) {
  const double t = finish_sum(partial, nblk);
  if (threadIdx.x == 0) *out = accumulate ? *out + t : t;
}

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

struct dmrgx_whole_basis {
  int32_t dim_LL, dim_RR, d;
  int64_t dim_whole;
  DevBuf<int32_t> inv;
  std::vector<std::unique_ptr<SectorDev>> sectors;
};

namespace {

// Device copy of a slot operator (validated like dmrgx_v_m_sector's).
struct DevSlotOp {
  DevBuf<int64_t> row;
  DevBuf<int32_t> col;
  DevBuf<double> val;
};

void upload_slot_op(const dmrgx_whole_basis &W, int32_t slot, const CRS1 *M, DevSlotOp &O) {
  const int32_t slot_dim = slot == 0 ? W.dim_LL : slot == 2 ? W.dim_RR : W.d;
  if (!M || !M->Row || M->row_dim < slot_dim) throw Failure(DMRGX_ERR_ARG, "operator has fewer rows than its block");
  const int64_t nnz = M->Row[slot_dim];
  for (int32_t r = 0; r < slot_dim; r++)
    if (M->Row[r] < 0 || M->Row[r] > M->Row[r + 1]) throw Failure(DMRGX_ERR_ARG, "operator row pointers not ascending");
  for (int64_t j = 0; j < nnz; j++)
    if (M->Col[j] < 0 || M->Col[j] >= slot_dim) throw Failure(DMRGX_ERR_ARG, "operator column outside its block");
  Context &C = ctx();
  O.row.resize(slot_dim + 1);
  O.col.resize(nnz > 0 ? nnz : 1);
  O.val.resize(nnz > 0 ? nnz : 1);
  C.copy_in(O.row.p, M->Row, sizeof(int64_t) * (slot_dim + 1));
  if (nnz > 0) {
    C.copy_in(O.col.p, M->Col, sizeof(int32_t) * nnz);
    C.copy_in(O.val.p, M->Val, sizeof(double) * nnz);
  }
}

// y (the output sector's dimension) = O on slot of x (dim_in entries), no sign.
void apply_slot(const dmrgx_whole_basis &W, int32_t slot, const CRS1 *M, int32_t sector_out, const double *x,
                int32_t dim_in, double *y, int32_t *bad) {
  DevSlotOp O;
  upload_slot_op(W, slot, M, O);
  const SectorDev &S = *W.sectors[sector_out];
  if (S.n == 0) return;
  k_v_m_sector<<<(unsigned)((S.n + 255) / 256), 256, 0, ctx().stream>>>(
      S.n, slot, W.dim_RR, W.d, S.LL.p, S.LR.p, S.RR.p, S.RL.p, W.inv.p, O.row.p, O.col.p, O.val.p, nullptr, nullptr,
      nullptr, dim_in, x, y, bad);
  HIPCHK(hipGetLastError());
  // O's buffers go back to the allocation cache here; whatever reuses them is queued on the
  // same stream, behind this kernel
}

void dot_into(int64_t n, const double *a, const double *b, double *partial, double *out, bool accumulate) {
  hipStream_t st = ctx().stream;
  k_dot_partial<<<kDotBlocks, kBlock, 0, st>>>(n, a, b, partial);
  k_dot_finish<<<1, kBlock, 0, st>>>(partial, kDotBlocks, out, accumulate ? 1 : 0);
  HIPCHK(hipGetLastError());
}

}  // namespace

extern "C" {

int dmrgx_whole_basis_create(int32_t dim_LL, int32_t dim_RR, int32_t dim_onsite, const int32_t *inv,
                             dmrgx_whole_basis **out) {
  if (!out || !inv) return DMRGX_ERR_ARG;
  *out = nullptr;
  return run_status([&] {
    if (dim_LL <= 0 || dim_RR <= 0 || dim_onsite <= 0 || dim_LL > 32767 || dim_RR > 32767 || dim_onsite > 32767)
      throw Failure(DMRGX_ERR_ARG, "block dimensions must lie in 1..32767");
    auto W = std::make_unique<dmrgx_whole_basis>();
    W->dim_LL = dim_LL;
    W->dim_RR = dim_RR;
    W->d = dim_onsite;
    W->dim_whole = (int64_t)dim_LL * dim_onsite * dim_RR * dim_onsite;
    for (int64_t i = 0; i < W->dim_whole; i++)
      if (inv[i] < -1) throw Failure(DMRGX_ERR_ARG, "inverse table entry below -1");
    W->inv.resize(W->dim_whole);
    ctx().copy_in(W->inv.p, inv, sizeof(int32_t) * W->dim_whole);
    *out = W.release();
  });
}

int dmrgx_whole_basis_add_sector(dmrgx_whole_basis *W, int32_t dim, const int16_t *LL, const int16_t *LR,
                                 const int16_t *RR, const int16_t *RL, int32_t *id) {
  if (!W || !id || dim < 0 || (dim > 0 && !(LL && LR && RR && RL))) return DMRGX_ERR_ARG;
  return run_status([&] {
    const int32_t dims[4] = {W->dim_LL, W->d, W->dim_RR, W->d};
    const int16_t *Q[4] = {LL, LR, RR, RL};
    for (int k = 0; k < 4; k++)
      for (int32_t i = 0; i < dim; i++)
        if (Q[k][i] < 0 || Q[k][i] >= dims[k]) throw Failure(DMRGX_ERR_ARG, "sector state outside its block dimension");
    auto S = std::make_unique<SectorDev>();
    S->n = dim;
    DevBuf<int16_t> *bufs[4] = {&S->LL, &S->LR, &S->RR, &S->RL};
    for (int k = 0; k < 4; k++) {
      bufs[k]->resize(dim > 0 ? dim : 1);
      if (dim > 0) ctx().copy_in(bufs[k]->p, Q[k], sizeof(int16_t) * dim);
    }
    *id = (int32_t)W->sectors.size();
    W->sectors.push_back(std::move(S));
  });
}

int dmrgx_v_m_sector(dmrgx_whole_basis *W, int32_t slot, const CRS1 *M, int32_t sector_out, int32_t dim_in,
                     const double *V, const int32_t *ele_LL, const int32_t *ele_LR, const int32_t *ele_RR,
                     int32_t sign_flag, double *Out_V) {
  if (!W || !M || slot < 0 || slot > 3 || dim_in < 0 || (dim_in > 0 && !V)) return DMRGX_ERR_ARG;
  return run_status([&] {
    if (sector_out < 0 || sector_out >= (int32_t)W->sectors.size())
      throw Failure(DMRGX_ERR_ARG, "no such output sector");
    const SectorDev &S = *W->sectors[sector_out];
    if (S.n > 0 && !Out_V) throw Failure(DMRGX_ERR_ARG, "output vector missing");
    const int32_t slot_dim = slot == 0 ? W->dim_LL : slot == 2 ? W->dim_RR : W->d;
    if (!M->Row || M->row_dim < slot_dim) throw Failure(DMRGX_ERR_ARG, "operator has fewer rows than its block");
    const int64_t nnz = M->Row[slot_dim];
    if (nnz < 0) throw Failure(DMRGX_ERR_ARG, "operator row pointers negative");
    for (int32_t r = 0; r < slot_dim; r++)
      if (M->Row[r] < 0 || M->Row[r] > M->Row[r + 1]) throw Failure(DMRGX_ERR_ARG, "operator row pointers not ascending");
    for (int64_t j = 0; j < nnz; j++)
      if (M->Col[j] < 0 || M->Col[j] >= slot_dim) throw Failure(DMRGX_ERR_ARG, "operator column outside its block");
    const bool sign = sign_flag != 0 && slot >= 1;
    if (sign && (!ele_LL || (slot >= 2 && !ele_LR) || (slot >= 3 && !ele_RR)))
      throw Failure(DMRGX_ERR_ARG, "electron numbers missing for the sign");
    if (S.n == 0) return;
    Context &C = ctx();
    hipStream_t st = C.stream;
    DevBuf<int64_t> row(slot_dim + 1);
    DevBuf<int32_t> col(nnz > 0 ? nnz : 1);
    DevBuf<double> val(nnz > 0 ? nnz : 1), x(dim_in > 0 ? dim_in : 1), y(S.n);
    C.copy_in(row.p, M->Row, sizeof(int64_t) * (slot_dim + 1));
    if (nnz > 0) {
      C.copy_in(col.p, M->Col, sizeof(int32_t) * nnz);
      C.copy_in(val.p, M->Val, sizeof(double) * nnz);
    }
    if (dim_in > 0) C.copy_in(x.p, V, sizeof(double) * dim_in);
    DevBuf<int32_t> eL, eLR, eR, bad(1);
    if (sign) {
      eL.resize(W->dim_LL);
      C.copy_in(eL.p, ele_LL, sizeof(int32_t) * W->dim_LL);
      if (slot >= 2) {
        eLR.resize(W->d);
        C.copy_in(eLR.p, ele_LR, sizeof(int32_t) * W->d);
      }
      if (slot >= 3) {
        eR.resize(W->dim_RR);
        C.copy_in(eR.p, ele_RR, sizeof(int32_t) * W->dim_RR);
      }
    }
    HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
    k_v_m_sector<<<(unsigned)((S.n + 255) / 256), 256, 0, st>>>(
        S.n, slot, W->dim_RR, W->d, S.LL.p, S.LR.p, S.RR.p, S.RL.p, W->inv.p, row.p, col.p, val.p,
        sign ? eL.p : nullptr, eLR.p, eR.p, dim_in, x.p, y.p, bad.p);
    HIPCHK(hipGetLastError());
    int32_t hbad = 0;
    C.copy_out(&hbad, bad.p, sizeof(int32_t));
    if (hbad) throw Failure(DMRGX_ERR_ARG, "the operator reaches a state whose index lies outside the input vector");
    C.copy_out(Out_V, y.p, sizeof(double) * S.n);
  });
}

int dmrgx_expectation_intersite_sector(dmrgx_whole_basis *W, int32_t sector_in, int32_t sector_out1,
                                       int32_t sector_out2, const CRS1 *const *M_CF, const CRS1 *const *M_LL,
                                       const CRS1 *M_On, const CRS1 *const *M_RR, int32_t origin, int32_t LL_site,
                                       int32_t RR_site, const double *Vec, double *Out) {
  if (!W || !M_CF || !M_On || !M_RR || !Vec || !Out || origin < 0 || LL_site < 0 || RR_site < 0)
    return DMRGX_ERR_ARG;
  return run_status([&] {
    const int32_t ns = (int32_t)W->sectors.size();
    if (sector_in < 0 || sector_in >= ns || sector_out1 < 0 || sector_out1 >= ns || sector_out2 < 0 ||
        sector_out2 >= ns)
      throw Failure(DMRGX_ERR_ARG, "no such sector");
    if (origin > LL_site + 1) return;   // EXPECTATION_INTERSITE_Q2.c:27-29
    if (origin <= LL_site && (!M_LL || !M_LL[origin])) throw Failure(DMRGX_ERR_ARG, "M_LL[origin] missing");
    const int32_t n_in = W->sectors[sector_in]->n, n1 = W->sectors[sector_out1]->n, n2 = W->sectors[sector_out2]->n;
    const int32_t n_out = RR_site + LL_site + 4 - origin;
    Context &C = ctx();
    hipStream_t st = C.stream;
    DevBuf<double> x(n_in > 0 ? n_in : 1), t1a(n1 > 0 ? n1 : 1), t1b(n2 > 0 ? n2 : 1), t2a(n1 > 0 ? n1 : 1),
        t2b(n2 > 0 ? n2 : 1), t2in(n_in > 0 ? n_in : 1), res(n_out), partial(kDotBlocks);
    DevBuf<int32_t> bad(1);
    HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
    if (n_in > 0) C.copy_in(x.p, Vec, sizeof(double) * n_in);
    // :31-37: the origin's operator into both neighbouring sectors
    if (origin == LL_site + 1) {
      apply_slot(*W, 1, M_On, sector_out1, x.p, n_in, t1a.p, bad.p);
      apply_slot(*W, 1, M_On, sector_out2, x.p, n_in, t1b.p, bad.p);
    } else {
      apply_slot(*W, 0, M_LL[origin], sector_out1, x.p, n_in, t1a.p, bad.p);
      apply_slot(*W, 0, M_LL[origin], sector_out2, x.p, n_in, t1b.p, bad.p);
    }
    // :39-43: sites origin..LL_site, the correlation operator in the input sector
    for (int32_t site = origin; site <= LL_site; site++) {
      const int32_t r = site - origin;
      if (!M_CF[r]) throw Failure(DMRGX_ERR_ARG, "M_CF[r] missing");
      apply_slot(*W, 0, M_CF[r], sector_in, x.p, n_in, t2in.p, bad.p);
      dot_into(n_in, x.p, t2in.p, partial.p, res.p + r, false);
    }
    // :45-54: LR and RL sites, both neighbouring sectors
    const int32_t slots[2] = {1, 3};
    for (int q = 0; q < 2; q++) {
      const int32_t r = LL_site + 1 + q - origin;
      apply_slot(*W, slots[q], M_On, sector_out1, x.p, n_in, t2a.p, bad.p);
      apply_slot(*W, slots[q], M_On, sector_out2, x.p, n_in, t2b.p, bad.p);
      dot_into(n1, t1a.p, t2a.p, partial.p, res.p + r, false);
      dot_into(n2, t1b.p, t2b.p, partial.p, res.p + r, true);
    }
    // :56-62: RR sites from RR_site down to 0
    for (int32_t site = RR_site; site >= 0; site--) {
      const int32_t r = RR_site + LL_site + 3 - site - origin;
      if (!M_RR[site]) throw Failure(DMRGX_ERR_ARG, "M_RR[site] missing");
      apply_slot(*W, 2, M_RR[site], sector_out1, x.p, n_in, t2a.p, bad.p);
      apply_slot(*W, 2, M_RR[site], sector_out2, x.p, n_in, t2b.p, bad.p);
      dot_into(n1, t1a.p, t2a.p, partial.p, res.p + r, false);
      dot_into(n2, t1b.p, t2b.p, partial.p, res.p + r, true);
    }
    int32_t hbad = 0;
    C.copy_out(&hbad, bad.p, sizeof(int32_t));
    if (hbad) throw Failure(DMRGX_ERR_ARG, "an operator reaches a state whose index lies outside the input vector");
    C.copy_out(Out, res.p, sizeof(double) * n_out);
  });
}

void dmrgx_whole_basis_free(dmrgx_whole_basis *W) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  delete W;
}

}  // extern "C"
