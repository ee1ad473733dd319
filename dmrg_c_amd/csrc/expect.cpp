// Expectation values on the superblock ground state: DMRG_V_M_{LL,LR,RR,RL}_Q0,
// DMRG_EXPECTATION_ONSITE and DMRG_EXPECTATION_INTERSITE_Q0 (include/dmrg.h:367-372,
// dmrg/DMRG_V_M_LL_Q0.c, dmrg/DMRG_EXPECTATION_ONSITE.c:3-28,
// dmrg/DMRG_EXPECTATION_INTERSITE_Q0.c:3-43).
//
// Each call uploads the basis (the dense superblock inverse is rebuilt on the device from
// it) and the vector once, then runs one single-slot product per site with the inner
// product fused into its epilogue; the per-site results stay on the device until the
// end.  Products and inner products use the library's deterministic reductions.
#include <cstring>

#include "entry.h"

namespace dmrgx {
void fill_inverse(int32_t n, int32_t dim_RR, int32_t d, const int16_t *LL, const int16_t *LR, const int16_t *RR,
                  const int16_t *RL, int32_t *inv, int64_t dim_whole, hipStream_t s);
}

using namespace dmrgx;

namespace {

struct DevBasis {
  int32_t n, dim_LL, dim_RR, d;
  DevBuf<int16_t> LL, LR, RR, RL;
  DevBuf<int32_t> inv;
};

std::unique_ptr<DevBasis> upload_basis(const DMRG_BASIS *B, int32_t n) {
  auto D = std::make_unique<DevBasis>();
  D->n = n;
  D->dim_LL = B->dim_LL;
  D->dim_RR = B->dim_RR;
  D->d = B->dim_onsite;
  if (n <= 0 || n > B->dim_LLLRRRRL || D->dim_LL <= 0 || D->dim_RR <= 0 || D->d <= 0)
    throw Failure(DMRGX_ERR_ARG, "bad superblock dimensions");
  const int dims[4] = {D->dim_LL, D->d, D->dim_RR, D->d};
  const short *Q[4] = {B->LL_LLLRRRRL, B->LR_LLLRRRRL, B->RR_LLLRRRRL, B->RL_LLLRRRRL};
  for (int k = 0; k < 4; k++)
    for (int32_t i = 0; i < n; i++)
      if (Q[k][i] < 0 || Q[k][i] >= dims[k]) throw Failure(DMRGX_ERR_ARG, "basis index outside its block dimension");
  hipStream_t st = ctx().stream;
  DevBuf<int16_t> *bufs[4] = {&D->LL, &D->LR, &D->RR, &D->RL};
  for (int k = 0; k < 4; k++) {
    bufs[k]->resize(n);
    ctx().copy_in(bufs[k]->p, Q[k], sizeof(int16_t) * n);
  }
  const int64_t dim_whole = (int64_t)D->dim_LL * D->d * D->dim_RR * D->d;
  D->inv.resize(dim_whole);
  fill_inverse(n, D->dim_RR, D->d, D->LL.p, D->LR.p, D->RR.p, D->RL.p, D->inv.p, dim_whole, st);
  return D;
}

struct DevOp {
  DevBuf<int64_t> row;
  DevBuf<int32_t> col;
  DevBuf<double> val;
};

// operator acting on a slot of dimension dim: rows >= dim, columns < dim
void upload_op(const CRS1 *M, int dim, DevOp &O) {
  if (!M || M->row_dim < dim) throw Failure(DMRGX_ERR_ARG, "operator has fewer rows than its block");
  const int64_t nnz = M->Row[dim];
  for (int64_t j = 0; j < nnz; j++)
    if (M->Col[j] < 0 || M->Col[j] >= dim) throw Failure(DMRGX_ERR_ARG, "operator column outside its block");
  O.row.resize((size_t)dim + 1);
  O.col.resize(nnz > 0 ? nnz : 1);
  O.val.resize(nnz > 0 ? nnz : 1);
  ctx().copy_in(O.row.p, M->Row, sizeof(int64_t) * ((size_t)dim + 1));
  if (nnz) {
    ctx().copy_in(O.col.p, M->Col, sizeof(int32_t) * nnz);
    ctx().copy_in(O.val.p, M->Val, sizeof(double) * nnz);
  }
}

SlotRows rows_of(const DevBasis &D, int slot, const DevOp &O) {
  return SlotRows{slot, D.dim_RR, D.d, D.LL.p, D.LR.p, D.RR.p, D.RL.p, D.inv.p, O.row.p, O.col.p, O.val.p};
}

int slot_dim(const DevBasis &D, int slot) { return slot == 0 ? D.dim_LL : slot == 2 ? D.dim_RR : D.d; }

void v_m(const char *fn, int slot, CRS1 *M, double *V, double *Out, int dim, DMRG_BASIS *B) {
  run_or_die(fn, [&] {
    auto D = upload_basis(B, dim);
    DevOp O;
    upload_op(M, slot_dim(*D, slot), O);
    hipStream_t st = ctx().stream;
    DevBuf<double> x(dim), y(dim);
    ctx().copy_in(x.p, V, sizeof(double) * dim);
    Reduce R;
    launch_slot_apply(dim, rows_of(*D, slot, O), x.p, y.p, nullptr, nullptr, R, st);
    ctx().copy_out(Out, y.p, sizeof(double) * dim);
    HIPCHK(hipStreamSynchronize(st));
  });
}

}  // namespace

extern "C" {

void DMRG_V_M_LL_Q0(CRS1 *M_LL, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *B) {
  (void)p_threads;
  v_m("DMRG_V_M_LL_Q0", 0, M_LL, V, Out_V, dim, B);
}
void DMRG_V_M_LR_Q0(CRS1 *M_LR, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *B) {
  (void)p_threads;
  v_m("DMRG_V_M_LR_Q0", 1, M_LR, V, Out_V, dim, B);
}
void DMRG_V_M_RR_Q0(CRS1 *M_RR, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *B) {
  (void)p_threads;
  v_m("DMRG_V_M_RR_Q0", 2, M_RR, V, Out_V, dim, B);
}
void DMRG_V_M_RL_Q0(CRS1 *M_RL, double *V, double *Out_V, int dim, int p_threads, DMRG_BASIS *B) {
  (void)p_threads;
  v_m("DMRG_V_M_RL_Q0", 3, M_RL, V, Out_V, dim, B);
}

void DMRG_EXPECTATION_ONSITE(CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, double *Vec, double *Temp_V,
                             int p_threads, DMRG_BASIS *B, DMRG_STATUS *St) {
  (void)p_threads;
  static const char *fn = "DMRG_EXPECTATION_ONSITE";
  run_or_die(fn, [&] {
    const int LL_site = St->LL_site, RR_site = St->RR_site, n = St->dim_LLLRRRRL;
    const int nout = LL_site + RR_site + 4;
    auto D = upload_basis(B, n);
    hipStream_t st = ctx().stream;
    DevBuf<double> v(n), t(n), res(nout);
    ctx().copy_in(v.p, Vec, sizeof(double) * n);
    Reduce R;
    DevOp O;
    // LL block sites (:11-14), LR (:17-18), RL (:21-22), RR block sites (:24-27)
    for (int site = 0; site <= LL_site; site++) {
      upload_op(M_LL[site], D->dim_LL, O);
      launch_slot_apply(n, rows_of(*D, 0, O), v.p, t.p, v.p, res.p + site, R, st);
    }
    upload_op(M_On, D->d, O);
    launch_slot_apply(n, rows_of(*D, 1, O), v.p, t.p, v.p, res.p + LL_site + 1, R, st);
    launch_slot_apply(n, rows_of(*D, 3, O), v.p, t.p, v.p, res.p + LL_site + 2, R, st);
    for (int site = RR_site; site >= 0; site--) {
      upload_op(M_RR[site], D->dim_RR, O);
      launch_slot_apply(n, rows_of(*D, 2, O), v.p, t.p, v.p, res.p + RR_site + LL_site + 3 - site, R, st);
    }
    ctx().copy_out(Out, res.p, sizeof(double) * nout);
    if (Temp_V) ctx().copy_out(Temp_V, t.p, sizeof(double) * n);
    HIPCHK(hipStreamSynchronize(st));
  });
}

void DMRG_EXPECTATION_INTERSITE_Q0(CRS1 **M_CF, CRS1 **M_LL, CRS1 *M_On, CRS1 **M_RR, double *Out, int origin,
                                   double *Vec, double *Temp_V1, double *Temp_V2, int p_threads, DMRG_BASIS *B,
                                   DMRG_STATUS *St) {
  (void)p_threads;
  static const char *fn = "DMRG_EXPECTATION_INTERSITE_Q0";
  const int LL_site = St->LL_site, RR_site = St->RR_site, n = St->dim_LLLRRRRL;
  if (origin > LL_site + 1) return;  // :10-12
  run_or_die(fn, [&] {
    const int nout = LL_site + RR_site + 4 - origin;
    auto D = upload_basis(B, n);
    hipStream_t st = ctx().stream;
    DevBuf<double> v(n), t1(n), t2(n), res(nout > 0 ? nout : 1);
    ctx().copy_in(v.p, Vec, sizeof(double) * n);
    Reduce R;
    DevOp O;
    // T1 = O_origin v (:14-19)
    if (origin == LL_site + 1) {
      upload_op(M_On, D->d, O);
      launch_slot_apply(n, rows_of(*D, 1, O), v.p, t1.p, nullptr, nullptr, R, st);
    } else {
      upload_op(M_LL[origin], D->dim_LL, O);
      launch_slot_apply(n, rows_of(*D, 0, O), v.p, t1.p, nullptr, nullptr, R, st);
    }
    // O*O inside the LL block (:22-26): <v | CF_r v>
    for (int site = origin; site <= LL_site; site++) {
      const int r = site - origin;
      upload_op(M_CF[r], D->dim_LL, O);
      launch_slot_apply(n, rows_of(*D, 0, O), v.p, t2.p, v.p, res.p + r, R, st);
    }
    // O*LR (:29-30), O*RL (:33-34), O*RR (:37-41): <T1 | O_site v>
    upload_op(M_On, D->d, O);
    launch_slot_apply(n, rows_of(*D, 1, O), v.p, t2.p, t1.p, res.p + LL_site + 1 - origin, R, st);
    launch_slot_apply(n, rows_of(*D, 3, O), v.p, t2.p, t1.p, res.p + LL_site + 2 - origin, R, st);
    for (int site = RR_site; site >= 0; site--) {
      const int r = RR_site + LL_site + 3 - site - origin;
      upload_op(M_RR[site], D->dim_RR, O);
      launch_slot_apply(n, rows_of(*D, 2, O), v.p, t2.p, t1.p, res.p + r, R, st);
    }
    ctx().copy_out(Out, res.p, sizeof(double) * nout);
    if (Temp_V1) ctx().copy_out(Temp_V1, t1.p, sizeof(double) * n);
    if (Temp_V2) ctx().copy_out(Temp_V2, t2.p, sizeof(double) * n);
    HIPCHK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
