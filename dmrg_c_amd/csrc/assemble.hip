// Superblock Hamiltonian assembly from a recorded term list.
//
// The model code (MAKE_ELEMENT_HAM_LLLRRRRL of each program) calls the element makers
// DMRG_MAKE_ELEM_*_LLLRRRRL once per row with the same ordered list of terms; only the
// row quantum numbers change.  Between dmrgx_record_begin() and dmrgx_record_end() this
// library's element makers record that list (kind, operator pointers, coefficient, sign
// rule, parity arrays) instead of computing; outside a recording they compute one row on
// the host exactly as dmrg/DMRG_MAKE_ELEM_*_LLLRRRRL.c do.
//
// The device assembler then builds, for every superblock row at once, the same
// lower-triangle CRS as GET_HAM_LLLRRRRL (main/dmrg/<model>_vf/src/allocations/
// GET_HAM_LLLRRRRL.c:11-114):
//   * one wavefront per row enumerates the row's contributions term by term in call order
//     (op1 row outer, op2 row inner, DMRG_MAKE_ELEM_LLLR_LLLRRRRL.c:54-79), drops those the
//     reference drops (inv < 0, |val| <= 1e-15), and those above the diagonal (they never
//     reach the lower triangle), and appends the rest to an LDS list in that order;
//   * the list is sorted by (column, position); each column's values are then summed
//     left to right in list order -- the reference's "first insertion sets, later ones add"
//     (DMRG_MAKE_ELEM_LL_LLLRRRRL.c:31-42) gives the same rounding;
//   * keep rule (0 <= inv <= row && |val| > 1e-15) || inv == row (GET_HAM_LLLRRRRL.c:49),
//     diagonal forced by a ZERO term (DMRG_MAKE_ELEM_ZERO_LLLRRRRL.c:15-26), columns
//     ascending (SORT_COLUMN_CRS1.c);
//   * two passes (count, then fill after an exclusive scan), like the reference.
// Products are evaluated left to right as sign*coeef*v1*v2 with contraction off, so Row,
// Col and Val come out bit-identical to the reference's.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "entry.h"

#pragma clang fp contract(off)

#include "superblock.h"

namespace dmrgx {
namespace {

__device__ __forceinline__ unsigned lanes_below(uint64_t mask) {
  return __popcll(mask & ((1ull << (threadIdx.x & 63)) - 1ull));
}

// One wavefront (one 64-thread workgroup) per superblock row.  CAP: LDS list capacity.
// FILL = false: counts[r] = kept entries; rows whose list exceeds CAP are appended to
// overflow (count pass) and redone by the next larger variant; FILL = true writes the rows
// whose lists fit and skips the others.  A count pass given scratch (stride CAP + 1 per
// row) also writes each fitting row's kept entries there and sets fitted[r], so that the
// fill pass for those rows is a copy (k_assemble_compact) instead of a second enumeration.
template <int CAP, bool FILL>
__global__ __launch_bounds__(64) void k_assemble(SbView sb, const int32_t *row_list, int32_t n_list,
                                                 int64_t *counts, const int64_t *row_ptr, int32_t *col_out,
                                                 double *val_out, int32_t *overflow, int32_t *n_overflow,
                                                 int32_t *scol = nullptr, double *sval = nullptr,
                                                 uint8_t *fitted = nullptr, unsigned *tcnt = nullptr,
                                                 int *bad = nullptr) {
  __shared__ uint64_t key[CAP];
  __shared__ double val[CAP];
  const int lane = threadIdx.x;
  const int32_t item = blockIdx.x;
  if (item >= n_list) return;
  const int32_t r = row_list ? row_list[item] : item;
  const int q[4] = {sb.LL[r], sb.LR[r], sb.RR[r], sb.RL[r]};
  int m = 0;  // list length (wave-uniform)
  for (int t = 0; t < sb.nt; t++) {
    const DTerm tm = sb.terms[t];
    if (tm.kind == K_ZERO) continue;  // handled after the sums (see file comment)
    const DOp a = sb.ops[tm.op1];
    const int64_t b1 = a.row[q[tm.s1]];
    const int64_t l1 = a.row[q[tm.s1] + 1] - b1;
    int64_t b2 = 0, l2 = 1;
    DOp b = a;
    if (tm.s2 >= 0) {
      b = sb.ops[tm.op2];
      b2 = b.row[q[tm.s2]];
      l2 = b.row[q[tm.s2] + 1] - b2;
    }
    const int64_t total = l1 * l2;
    if (total == 0) continue;
    const double sc = (double)term_sign(tm, sb.eles, q) * tm.coeff;
    for (int64_t base = 0; base < total; base += 64) {
      const int64_t k = base + lane;
      bool keep = false;
      int32_t col = 0;
      double v = 0.0;
      if (k < total) {
        int qq[4] = {q[0], q[1], q[2], q[3]};
        const int64_t i = (tm.s2 >= 0) ? k / l2 : k;
        qq[tm.s1] = a.col[b1 + i];
        double prod = sc * a.val[b1 + i];
        if (tm.s2 >= 0) {
          const int64_t j = k - i * l2;
          qq[tm.s2] = b.col[b2 + j];
          prod = prod * b.val[b2 + j];
        }
        const int32_t inv = sb.inv[sb_index(qq[0], qq[1], qq[2], qq[3], sb.dim_RR, sb.d)];
        col = inv;
        v = prod;
        keep = inv >= 0 && inv <= r && fabs(prod) > kZero;
      }
      const uint64_t mask = __ballot(keep);
      if (keep) {
        const int pos = m + (int)lanes_below(mask);
        if (pos < CAP) {
          key[pos] = ((uint64_t)(uint32_t)col << 32) | (uint32_t)pos;
          val[pos] = v;
        }
      }
      m += __popcll(mask);
    }
  }
  if (m > CAP) {
    // count pass: queue the row for the larger variant, or count it as an error there
    if (!FILL && lane == 0) {
      if (overflow) overflow[atomicAdd(n_overflow, 1)] = r;
      else atomicAdd(n_overflow, 1);
      if (fitted) fitted[r] = 0;
    }
    return;
  }
  const int64_t sbase = (int64_t)r * (CAP + 1);   // this row's scratch slot
  // bitonic sort of key[0..P) with P = next power of two >= m (padding sorts last)
  int P = 1;
  while (P < m) P <<= 1;
  for (int i = m + lane; i < P; i += 64) key[i] = ~0ull;
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int jx = i ^ stride;
        if (jx > i) {
          const uint64_t x = key[i], y = key[jx];
          const bool up = (i & size) == 0;
          if ((x > y) == up) {
            key[i] = y;
            key[jx] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  // segment heads: sum each column's values in list order; keep rule of GET_HAM_LLLRRRRL.c:49
  int out = 0;
  bool diag_seen = false;
  for (int base = 0; base < m; base += 64) {
    const int i = base + lane;
    bool keep = false, is_diag = false;
    int32_t c = 0;
    double s = 0.0;
    if (i < m) {
      c = (int32_t)(key[i] >> 32);
      const bool head = (i == 0) || ((int32_t)(key[i - 1] >> 32) != c);
      if (head) {
        s = val[(uint32_t)key[i]];
        for (int j = i + 1; j < m && (int32_t)(key[j] >> 32) == c; j++) s = s + val[(uint32_t)key[j]];
        is_diag = (c == r);
        keep = is_diag || fabs(s) > kZero;
      }
    }
    const uint64_t km = __ballot(keep);
    diag_seen = diag_seen || (__ballot(is_diag) != 0);
    if (FILL && keep) {
      const int64_t dst = row_ptr[r] + out + (int64_t)lanes_below(km);
      col_out[dst] = c;
      val_out[dst] = s;
    }
    if (!FILL && scol && keep) {
      const int64_t dst = sbase + out + (int64_t)lanes_below(km);
      scol[dst] = c;
      sval[dst] = s;
    }
    // the count pass also counts each column's strictly-lower entries for the expansion
    if (!FILL && tcnt && keep && c < r) atomicAdd(&tcnt[c], 1u);
    out += __popcll(km);
  }
  if (!diag_seen && sb.has_zero) {
    if (FILL && lane == 0) {
      col_out[row_ptr[r] + out] = r;
      val_out[row_ptr[r] + out] = 0.0;
    }
    if (!FILL && scol && lane == 0) {
      scol[sbase + out] = r;
      sval[sbase + out] = 0.0;
    }
    out++;
  }
  if (!FILL && lane == 0) {
    counts[r] = out;
    if (fitted) fitted[r] = 1;
    if (bad && !diag_seen && !sb.has_zero) atomicOr(bad, 1);
  }
}

// Fill pass of the rows whose kept entries a count pass left in scratch: one wavefront
// per row copies them to their place in the CRS.
__global__ __launch_bounds__(64) void k_assemble_compact(int32_t n, const uint8_t *__restrict__ fitted,
                                                         const int64_t *__restrict__ row_ptr, int64_t stride,
                                                         const int32_t *__restrict__ scol,
                                                         const double *__restrict__ sval,
                                                         int32_t *__restrict__ col_out,
                                                         double *__restrict__ val_out) {
  const int32_t r = blockIdx.x;
  if (r >= n || !fitted[r]) return;
  const int64_t b = row_ptr[r], len = row_ptr[r + 1] - b;
  const int64_t src = (int64_t)r * stride;
  for (int64_t i = threadIdx.x; i < len; i += 64) {
    col_out[b + i] = scol[src + i];
    val_out[b + i] = sval[src + i];
  }
}

__global__ __launch_bounds__(256) void k_fill_int(int64_t n, int32_t v, int32_t *out) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) out[i] = v;
}

__global__ __launch_bounds__(256) void k_inv_scatter(SbView sb, int32_t *inv) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= sb.n) return;
  inv[sb_index(sb.LL[r], sb.LR[r], sb.RR[r], sb.RL[r], sb.dim_RR, sb.d)] = r;
}

// List capacities of the three k_assemble variants: most rows fit the small one, whose 4 KiB
// of LDS lets a CU hold as many rows as it has wave slots; longer rows are redone by the
// larger variants.
constexpr int kCapSmall = 256;   // 4 KiB of LDS per row
constexpr int kCapMid = 1024;    // 16 KiB
constexpr int kCapLarge = 4096;  // 64 KiB
// Device scratch the count pass may use to keep its rows' entries (n * (kCapSmall + 1) * 12
// bytes: 0.45 GB at n = 144542, 6.1 GB at n = 1993680; held only during the assembly).
constexpr int64_t kScratchBytes = (int64_t)16 << 30;

// ------------------------------------------------------------------ host
int op_role_dim(int slot, const dmrgx_superblock &S) {
  return slot == S_LL ? S.dim_LL : slot == S_RR ? S.dim_RR : S.d;
}

}  // namespace

void launch_exclusive_scan(int32_t n, const int64_t *len, int64_t *out, hipStream_t s);

struct ArrOp {
  int32_t row_dim, col_dim;
  const int64_t *row;
  const int32_t *col;
  const double *val;
};

int64_t build_basis_device(const dmrgx_basis_qn &q, DevBuf<int16_t> &LL, DevBuf<int16_t> &LR, DevBuf<int16_t> &RR,
                           DevBuf<int16_t> &RL);

// The basis comes from the caller's arrays (LL..RL, host) or, with qn set, is built on the
// device from the blocks' quantum numbers (n, dim_LL, dim_RR, d are then taken from qn).
static dmrgx_superblock *create_superblock(int32_t n, int32_t dim_LL, int32_t dim_RR, int32_t d,
                                           const int16_t *LL, const int16_t *LR, const int16_t *RR,
                                           const int16_t *RL, const std::vector<ArrOp> &ops,
                                           const std::vector<std::pair<const int32_t *, int32_t>> &eles,
                                           const std::vector<DTerm> &terms_in, const dmrgx_basis_qn *qn = nullptr) {
  auto S = std::make_unique<dmrgx_superblock>();
  if (qn) {
    dim_LL = qn->dim_LL;
    dim_RR = qn->dim_RR;
    d = qn->dim_onsite;
    n = (int32_t)build_basis_device(*qn, S->LL, S->LR, S->RR, S->RL);
  }
  if (n <= 0 || dim_LL <= 0 || dim_RR <= 0 || d <= 0) throw Failure(DMRGX_ERR_ARG, "superblock dimensions must be positive");
  if ((int)terms_in.size() > kMaxTerms) throw Failure(DMRGX_ERR_ARG, "too many terms");
  if (!qn) {
    const int dims[4] = {dim_LL, d, dim_RR, d};
    const int16_t *Q[4] = {LL, LR, RR, RL};
    for (int k = 0; k < 4; k++)
      for (int32_t i = 0; i < n; i++)
        if (Q[k][i] < 0 || Q[k][i] >= dims[k]) throw Failure(DMRGX_ERR_ARG, "basis index outside its block dimension");
  }
  S->n = n;
  S->dim_LL = dim_LL;
  S->dim_RR = dim_RR;
  S->d = d;
  std::vector<DTerm> terms;
  // validate operators against the slots they act on: rows indexed by the slot's state,
  // columns become the slot's new state (so must stay inside the block dimension)
  std::vector<int> op_need_rows(ops.size(), 0), op_col_limit(ops.size(), 1 << 30);
  for (const DTerm &t0 : terms_in) {
    DTerm t = t0;
    if (t.kind < 0 || t.kind >= K_COUNT) throw Failure(DMRGX_ERR_ARG, "bad term kind");
    if (t.kind == K_ZERO) {
      S->has_zero = 1;
      continue;
    }
    if (std::fabs(t.coeff) < kZero) continue;  // the element makers return at once (:20-22)
    t.s1 = kSlot1[t.kind];
    t.s2 = kSlot2[t.kind];
    if (t.op1 < 0 || t.op1 >= (int)ops.size() || (t.s2 >= 0 && (t.op2 < 0 || t.op2 >= (int)ops.size())))
      throw Failure(DMRGX_ERR_ARG, "term refers to a missing operator");
    const int o[2] = {t.op1, t.s2 >= 0 ? t.op2 : -1};
    const int sl[2] = {t.s1, t.s2};
    for (int k = 0; k < 2; k++) {
      if (o[k] < 0) continue;
      op_need_rows[o[k]] = std::max(op_need_rows[o[k]], op_role_dim(sl[k], *S));
      op_col_limit[o[k]] = std::min(op_col_limit[o[k]], op_role_dim(sl[k], *S));
    }
    if (t.sign_mode != 0) {
      const int roles[K_COUNT][3] = {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1},
                                     {S_LL, -1, -1}, {S_LL, S_RR, S_LR}, {S_LL, S_LR, -1}, {S_RR, S_LR, -1},
                                     {S_LR, -1, -1}, {S_RR, -1, -1}, {-1, -1, -1}};
      const int e[3] = {t.e0, t.e1, t.e2};
      for (int k = 0; k < 3; k++) {
        if (roles[t.kind][k] < 0) continue;
        if (e[k] < 0 || e[k] >= (int)eles.size()) throw Failure(DMRGX_ERR_ARG, "term with a sign rule lacks its parity array");
        if (eles[e[k]].second < op_role_dim(roles[t.kind][k], *S))
          throw Failure(DMRGX_ERR_ARG, "parity array shorter than the block it indexes");
      }
    }
    terms.push_back(t);
  }
  for (size_t i = 0; i < ops.size(); i++) {
    if (op_need_rows[i] == 0) continue;
    const ArrOp &a = ops[i];
    if (a.row_dim < op_need_rows[i]) throw Failure(DMRGX_ERR_ARG, "operator has fewer rows than its block");
    const int64_t nnz = a.row[a.row_dim] - a.row[0];
    for (int64_t j = 0; j < nnz; j++)
      if (a.col[j] < 0 || a.col[j] >= op_col_limit[i]) throw Failure(DMRGX_ERR_ARG, "operator column outside its block");
  }
  S->nt = (int)terms.size();
  S->hterms = terms;
  S->hele.resize(eles.size());
  for (size_t i = 0; i < eles.size(); i++) S->hele[i].assign(eles[i].first, eles[i].first + eles[i].second);
  hipStream_t st = ctx().stream;
  auto up16 = [&](DevBuf<int16_t> &b, const int16_t *h) {
    b.resize(n);
    ctx().copy_in(b.p, h, sizeof(int16_t) * n);
  };
  if (!qn) {
    up16(S->LL, LL);
    up16(S->LR, LR);
    up16(S->RR, RR);
    up16(S->RL, RL);
  }
  std::vector<DOp> hops(std::max<size_t>(ops.size(), 1));
  S->op_row.resize(ops.size());
  S->op_col.resize(ops.size());
  S->op_val.resize(ops.size());
  for (size_t i = 0; i < ops.size(); i++) {
    const ArrOp &a = ops[i];
    const int64_t nnz = a.row[a.row_dim];
    S->op_row[i].resize((size_t)a.row_dim + 1);
    S->op_col[i].resize(std::max<int64_t>(nnz, 1));
    S->op_val[i].resize(std::max<int64_t>(nnz, 1));
    ctx().copy_in(S->op_row[i].p, a.row, sizeof(int64_t) * ((size_t)a.row_dim + 1));
    if (nnz) {
      ctx().copy_in(S->op_col[i].p, a.col, sizeof(int32_t) * nnz);
      ctx().copy_in(S->op_val[i].p, a.val, sizeof(double) * nnz);
    }
    hops[i] = DOp{S->op_row[i].p, S->op_col[i].p, S->op_val[i].p};
  }
  std::vector<const int32_t *> heles(std::max<size_t>(eles.size(), 1), nullptr);
  S->ele.resize(eles.size());
  for (size_t i = 0; i < eles.size(); i++) {
    S->ele[i].resize(std::max(eles[i].second, 1));
    ctx().copy_in(S->ele[i].p, eles[i].first, sizeof(int32_t) * eles[i].second);
    heles[i] = S->ele[i].p;
  }
  S->ops.resize(hops.size());
  ctx().copy_in(S->ops.p, hops.data(), sizeof(DOp) * hops.size());
  S->terms.resize(std::max<size_t>(terms.size(), 1));
  if (!terms.empty())
    ctx().copy_in(S->terms.p, terms.data(), sizeof(DTerm) * terms.size());
  S->ele_ptrs.resize(heles.size());
  ctx().copy_in(S->ele_ptrs.p, heles.data(), sizeof(const int32_t *) * heles.size());
  // dense superblock inverse, -1 where absent (GET_BASIS_SUPERBLOCK.c Inv fill)
  const int64_t dim_whole = (int64_t)dim_LL * d * dim_RR * d;
  if (dim_whole > (int64_t)1 << 34) throw Failure(DMRGX_ERR_ARG, "superblock product space too large for the dense inverse");
  S->inv.resize(dim_whole);
  k_fill_int<<<1024, 256, 0, st>>>(dim_whole, -1, S->inv.p);
  k_inv_scatter<<<(n + 255) / 256, 256, 0, st>>>(S->view(), S->inv.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));  // host arrays above go out of scope
  return S.release();
}

void fill_inverse(int32_t n, int32_t dim_RR, int32_t d, const int16_t *LL, const int16_t *LR, const int16_t *RR,
                  const int16_t *RL, int32_t *inv, int64_t dim_whole, hipStream_t s) {
  SbView v{};
  v.n = n;
  v.dim_RR = dim_RR;
  v.d = d;
  v.LL = LL;
  v.LR = LR;
  v.RR = RR;
  v.RL = RL;
  k_fill_int<<<1024, 256, 0, s>>>(dim_whole, -1, inv);
  k_inv_scatter<<<(n + 255) / 256, 256, 0, s>>>(v, inv);
  HIPCHK(hipGetLastError());
}

// Assemble the lower triangle into S->lrow/lcol/lval (device).  Returns nnz.
static int64_t assemble_lower(dmrgx_superblock &S) {
  hipStream_t st = ctx().stream;
  const int32_t n = S.n;
  const SbView v = S.view();
  DevBuf<int64_t> counts(n);
  DevBuf<int32_t> over1(n), over2(n), n_over(3);
  HIPCHK(hipMemsetAsync(n_over.p, 0, 3 * sizeof(int32_t), st));
  // The small-list count pass keeps its rows' entries in scratch (one slot of
  // kCapSmall + 1 entries per row) when that fits kScratchBytes; the fill pass of those rows
  // is then a copy.
  const int64_t stride = kCapSmall + 1;
  const bool keep = (int64_t)n * stride * (int64_t)(sizeof(int32_t) + sizeof(double)) <= kScratchBytes;
  DevBuf<int32_t> scol(keep ? (size_t)n * stride : 0);
  DevBuf<double> sval(keep ? (size_t)n * stride : 0);
  DevBuf<uint8_t> fitted(keep ? (size_t)n : 0);
  S.tcnt.resize(n);
  S.bad.resize(1);
  HIPCHK(hipMemsetAsync(S.tcnt.p, 0, sizeof(unsigned) * n, st));
  HIPCHK(hipMemsetAsync(S.bad.p, 0, sizeof(int), st));
  // count pass: small lists for every row, then the rows that overflowed with larger lists
  k_assemble<kCapSmall, false><<<n, 64, 0, st>>>(v, nullptr, n, counts.p, nullptr, nullptr, nullptr, over1.p,
                                                 n_over.p, scol.p, sval.p, fitted.p, S.tcnt.p, S.bad.p);
  HIPCHK(hipGetLastError());
  // row pointers right away: they hold unless some row overflowed (then they are redone),
  // so the common case reads the overflow count and nnz with one round trip
  S.lrow.resize((size_t)n + 1);
  launch_exclusive_scan(n, counts.p, S.lrow.p, st);
  int64_t *hb = reinterpret_cast<int64_t *>(ctx().host_stage(4));
  int *h = reinterpret_cast<int *>(hb + 1);
  HIPCHK(hipMemcpyAsync(hb, S.lrow.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h, n_over.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int n_mid = h[0];
  int64_t nnz = hb[0];
  int n_large = 0;
  if (n_mid > 0) {
    k_assemble<kCapMid, false><<<n_mid, 64, 0, st>>>(v, over1.p, n_mid, counts.p, nullptr, nullptr, nullptr,
                                                     over2.p, n_over.p + 1, nullptr, nullptr, nullptr, S.tcnt.p,
                                                     S.bad.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h + 1, n_over.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    n_large = h[1];
  }
  if (n_large > 0) {
    k_assemble<kCapLarge, false><<<n_large, 64, 0, st>>>(v, over2.p, n_large, counts.p, nullptr, nullptr,
                                                         nullptr, nullptr, n_over.p + 2, nullptr, nullptr, nullptr,
                                                         S.tcnt.p, S.bad.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h + 2, n_over.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h[2] > 0)
      throw Failure(DMRGX_ERR_ARG, "a superblock row has more than 4096 lower-triangle contributions");
  }
  if (n_mid > 0) {
    launch_exclusive_scan(n, counts.p, S.lrow.p, st);
    HIPCHK(hipMemcpyAsync(hb, S.lrow.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    nnz = hb[0];
  }
  S.lcol.resize(std::max<int64_t>(nnz, 1));
  S.lval.resize(std::max<int64_t>(nnz, 1));
  // fill pass: each variant writes the rows whose lists fit it and skips the others
  if (keep)
    k_assemble_compact<<<n, 64, 0, st>>>(n, fitted.p, S.lrow.p, stride, scol.p, sval.p, S.lcol.p, S.lval.p);
  else
    k_assemble<kCapSmall, true><<<n, 64, 0, st>>>(v, nullptr, n, nullptr, S.lrow.p, S.lcol.p, S.lval.p, nullptr,
                                                  nullptr);
  if (n_mid > 0)
    k_assemble<kCapMid, true><<<n_mid, 64, 0, st>>>(v, over1.p, n_mid, nullptr, S.lrow.p, S.lcol.p, S.lval.p,
                                                    nullptr, nullptr);
  if (n_large > 0)
    k_assemble<kCapLarge, true><<<n_large, 64, 0, st>>>(v, over2.p, n_large, nullptr, S.lrow.p, S.lcol.p,
                                                        S.lval.p, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  // no wait: whatever reads the triangle next is queued behind the fill on this stream
  // (and the scratch released here is only reused in stream order)
  S.lnnz = nnz;
  return nnz;
}

static dmrgx_matrix *expand_last(dmrgx_superblock &S) {
  auto A = std::make_unique<dmrgx_matrix>();
  launch_expand_lower(S.n, S.lnnz, S.lrow.p, S.lcol.p, S.lval.p, *A, ctx().stream, S.tcnt.p, S.bad.p);
  default_schedule(*A);
  return A.release();
}

// ------------------------------------------------------------- recorder state
namespace {
thread_local dmrgx_terms *g_rec = nullptr;
std::mutex g_cache_mu;
// one-shot hand-over of the device matrix assembled for a returned CRS1
struct Pending {
  const CRS1 *crs = nullptr;
  const long *row = nullptr;
  const int *col = nullptr;
  const double *val = nullptr;
  dmrgx_matrix *A = nullptr;
} g_pending;
}  // namespace

dmrgx_matrix *take_pending(const CRS1 *M) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  if (g_pending.A && g_pending.crs == M && g_pending.row == M->Row && g_pending.col == M->Col &&
      g_pending.val == M->Val && g_pending.A->n == M->row_dim) {
    dmrgx_matrix *A = g_pending.A;
    g_pending = Pending{};
    return A;
  }
  return nullptr;
}

static void set_pending(const CRS1 *M, dmrgx_matrix *A) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  if (g_pending.A) dmrgx_matrix_free(g_pending.A);
  g_pending = Pending{M, M->Row, M->Col, M->Val, A};
}

}  // namespace dmrgx

using namespace dmrgx;

// ------------------------------------------------------ host element makers
namespace {

[[noreturn]] void bad_type(const char *fn) {
  printf("Error in %s\n", fn);
  exit(1);
}

// sign of a two-operator term (DMRG_MAKE_ELEM_*_LLLRRRRL.c:28-52)
int sign_of(const char *fn, const char *flag, const char *type, const char *first, const char *second, int parity) {
  if (!flag || strcmp(flag, "Yes") != 0) return 1;
  const bool even = parity % 2 == 0;
  if (type && strcmp(type, first) == 0) return even ? -1 : 1;
  if (type && strcmp(type, second) == 0) return even ? 1 : -1;
  bad_type(fn);
}

int mode_of(const char *fn, const char *flag, const char *type, const char *first, const char *second) {
  if (!flag || strcmp(flag, "Yes") != 0) return 0;
  if (type && strcmp(type, first) == 0) return 1;
  if (type && strcmp(type, second) == 0) return 2;
  bad_type(fn);
}

void record(int kind, const CRS1 *a, const CRS1 *b, double c, int mode, const int *e0, int r0, const int *e1,
            int r1, const int *e2, int r2) {
  dmrgx_terms::T t;
  t.kind = kind;
  t.op1 = a;
  t.op2 = b;
  t.coeff = c;
  t.sign_mode = mode;
  t.ele[0] = e0;
  t.ele[1] = e1;
  t.ele[2] = e2;
  t.role[0] = r0;
  t.role[1] = r1;
  t.role[2] = r2;
  g_rec->t.push_back(t);
}

// one row of the reference element makers, accumulating into A_Basis
void host_elem(int kind, DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, const CRS1 *M1, const CRS1 *M2,
               double coeff, int sign, int *elem_num) {
  if (std::fabs(coeff) < kZero) return;
  int q[4] = {B->LL, B->LR, B->RR, B->RL};
  const int d = B->dim_onsite, dr = B->dim_RR;
  const int s1 = kSlot1[kind], s2 = kSlot2[kind];
  int ne = *elem_num;
  auto put = [&](const int qq[4], double v) {
    const long sup = (long)qq[0] * d * dr * d + (long)qq[1] * dr * d + (long)qq[2] * d + qq[3];
    const int inv = Inv[sup];
    if (!(std::fabs(v) > kZero && inv >= 0)) return;
    const int chk = A->Inv_LLLRRRRL[inv];
    if (chk == -1) {
      A->LL_LLLRRRRL[ne] = (short)qq[0];
      A->LR_LLLRRRRL[ne] = (short)qq[1];
      A->RR_LLLRRRRL[ne] = (short)qq[2];
      A->RL_LLLRRRRL[ne] = (short)qq[3];
      A->Val_LLLRRRRL[ne] = v;
      A->Inv_LLLRRRRL[inv] = ne;
      ne++;
    } else {
      A->Val_LLLRRRRL[chk] = A->Val_LLLRRRRL[chk] + v;
    }
  };
  for (long i = M1->Row[q[s1]]; i < M1->Row[q[s1] + 1]; i++) {
    int qq[4] = {q[0], q[1], q[2], q[3]};
    qq[s1] = M1->Col[i];
    if (s2 < 0) {
      put(qq, coeff * M1->Val[i]);
    } else {
      const double v1 = M1->Val[i];
      for (long j = M2->Row[q[s2]]; j < M2->Row[q[s2] + 1]; j++) {
        qq[s2] = M2->Col[j];
        put(qq, (double)sign * coeff * v1 * M2->Val[j]);
      }
    }
  }
  *elem_num = ne;
}

// Element makers of the matrix-free path's sub-Hamiltonians (dmrg/DMRG_MAKE_ELEM_LL_LLLR.c,
// _LR_LLLR, _LLLR_LLLR, _LRRL_LRRL, _RL_RRRL, _RR_RRRL, _RRRL_RRRL): the two-slot sub-basis
// (sa, sb) of LLLR (LL, LR), LRRL (LR, RL) or RRRL (RR, RL), its 2-D inverse Inv[q_sa][q_sb],
// one operator acting on slot a1 or two acting on a1 and a2; the same first-insertion /
// accumulation rule and 1e-15 drops as the superblock makers, products left to right.
// A state absent from the sub-basis (inverse -1) is skipped before its accumulator entry is
// read (the reference reads A_Basis->Inv_LLLRRRRL[-1] there and then discards it).
void host_sub_elem(int sa, int sb, int a1, int a2, DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv,
                   const CRS1 *M1, const CRS1 *M2, double coeff, int sign, int *elem_num) {
  if (std::fabs(coeff) < kZero) return;
  int q[4] = {B->LL, B->LR, B->RR, B->RL};
  short *field[4] = {A->LL_LLLRRRRL, A->LR_LLLRRRRL, A->RR_LLLRRRRL, A->RL_LLLRRRRL};
  int ne = *elem_num;
  auto put = [&](const int qq[4], double v) {
    const int inv = Inv[qq[sa]][qq[sb]];
    if (!(std::fabs(v) > kZero && inv >= 0)) return;
    const int chk = A->Inv_LLLRRRRL[inv];
    if (chk == -1) {
      field[sa][ne] = (short)qq[sa];
      field[sb][ne] = (short)qq[sb];
      A->Val_LLLRRRRL[ne] = v;
      A->Inv_LLLRRRRL[inv] = ne;
      ne++;
    } else {
      A->Val_LLLRRRRL[chk] = A->Val_LLLRRRRL[chk] + v;
    }
  };
  for (long i = M1->Row[q[a1]]; i < M1->Row[q[a1] + 1]; i++) {
    int qq[4] = {q[0], q[1], q[2], q[3]};
    qq[a1] = M1->Col[i];
    if (a2 < 0) {
      put(qq, coeff * M1->Val[i]);
    } else {
      const double v1 = M1->Val[i];
      for (long j = M2->Row[q[a2]]; j < M2->Row[q[a2] + 1]; j++) {
        qq[a2] = M2->Col[j];
        put(qq, (double)sign * coeff * v1 * M2->Val[j]);
      }
    }
  }
  *elem_num = ne;
}

}  // namespace

extern "C" {

// ---- sub-Hamiltonians of the matrix-free path (include/dmrg.h:342-350)
void DMRG_MAKE_ELEM_LL_LLLR(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_LLLR, CRS1 *M, double c, int *en) {
  host_sub_elem(0, 1, 0, -1, B, A, Inv_LLLR, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_LR_LLLR(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_LLLR, CRS1 *M, double c, int *en) {
  host_sub_elem(0, 1, 1, -1, B, A, Inv_LLLR, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_LLLR_LLLR(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_LLLR, CRS1 *M_LL, CRS1 *M_LR, double c,
                              int *LL_Ele, int *en, char Type[], char Sign_Flag[]) {
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int sg = sign_of("DMRG_MAKE_ELEM_LLLR_LLLR", Sign_Flag, Type, "LL_LR", "LR_LL", yes ? LL_Ele[B->LL] : 0);
  host_sub_elem(0, 1, 0, 1, B, A, Inv_LLLR, M_LL, M_LR, c, sg, en);
}
void DMRG_MAKE_ELEM_LRRL_LRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_LRRL, CRS1 *M_LR, CRS1 *M_RL, double c,
                              int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int sg = sign_of("DMRG_MAKE_ELEM_LRRL_LRRL", Sign_Flag, Type, "LR_RL", "RL_LR", yes ? Ele_LR[B->LR] : 0);
  host_sub_elem(1, 3, 1, 3, B, A, Inv_LRRL, M_LR, M_RL, c, sg, en);
}
void DMRG_MAKE_ELEM_RL_RRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_RRRL, CRS1 *M, double c, int *en) {
  host_sub_elem(2, 3, 3, -1, B, A, Inv_RRRL, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_RR_RRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_RRRL, CRS1 *M, double c, int *en) {
  host_sub_elem(2, 3, 2, -1, B, A, Inv_RRRL, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_RRRL_RRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int **Inv_RRRL, CRS1 *M_RR, CRS1 *M_RL, double c,
                              int *Ele_RR, int *en, char Type[], char Sign_Flag[]) {
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int sg = sign_of("DMRG_MAKE_ELEM_RRRL_RRRL", Sign_Flag, Type, "RR_RL", "RL_RR", yes ? Ele_RR[B->RR] : 0);
  host_sub_elem(2, 3, 2, 3, B, A, Inv_RRRL, M_RR, M_RL, c, sg, en);
}

void DMRG_MAKE_ELEM_LL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M, double c, int *en) {
  if (g_rec) return record(K_LL, M, nullptr, c, 0, nullptr, 0, nullptr, 0, nullptr, 0);
  host_elem(K_LL, B, A, Inv, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_LR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M, double c, int *en) {
  if (g_rec) return record(K_LR, M, nullptr, c, 0, nullptr, 0, nullptr, 0, nullptr, 0);
  host_elem(K_LR, B, A, Inv, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_RR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M, double c, int *en) {
  if (g_rec) return record(K_RR, M, nullptr, c, 0, nullptr, 0, nullptr, 0, nullptr, 0);
  host_elem(K_RR, B, A, Inv, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_RL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M, double c, int *en) {
  if (g_rec) return record(K_RL, M, nullptr, c, 0, nullptr, 0, nullptr, 0, nullptr, 0);
  host_elem(K_RL, B, A, Inv, M, nullptr, c, 1, en);
}
void DMRG_MAKE_ELEM_LLLR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LL, CRS1 *M_LR, double c,
                                  int *LL_Ele, int *en, char Type[], char Sign_Flag[]) {
  static const char *fn = "DMRG_MAKE_ELEM_LLLR_LLLRRRRL";
  if (g_rec) return record(K_LLLR, M_LL, M_LR, c, mode_of(fn, Sign_Flag, Type, "LL_LR", "LR_LL"), LL_Ele, ROLE_LL, nullptr, 0, nullptr, 0);
  if (std::fabs(c) < kZero) return;
  const int sg = sign_of(fn, Sign_Flag, Type, "LL_LR", "LR_LL", LL_Ele ? LL_Ele[B->LL] : 0);
  host_elem(K_LLLR, B, A, Inv, M_LL, M_LR, c, sg, en);
}
void DMRG_MAKE_ELEM_LLRL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LL, CRS1 *M_RL, double c,
                                  int *Ele_LL, int *Ele_RR, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
  static const char *fn = "DMRG_MAKE_ELEM_LLRL_LLLRRRRL";
  if (g_rec) return record(K_LLRL, M_LL, M_RL, c, mode_of(fn, Sign_Flag, Type, "LL_RL", "RL_LL"), Ele_LL, ROLE_LL, Ele_RR, ROLE_RR, Ele_LR, ROLE_ON);
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int p = yes ? Ele_LL[B->LL] + Ele_RR[B->RR] + Ele_LR[B->LR] : 0;
  host_elem(K_LLRL, B, A, Inv, M_LL, M_RL, c, sign_of(fn, Sign_Flag, Type, "LL_RL", "RL_LL", p), en);
}
void DMRG_MAKE_ELEM_LLRR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LL, CRS1 *M_RR, double c,
                                  int *Ele_LL, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
  static const char *fn = "DMRG_MAKE_ELEM_LLRR_LLLRRRRL";
  if (g_rec) return record(K_LLRR, M_LL, M_RR, c, mode_of(fn, Sign_Flag, Type, "LL_RR", "RR_LL"), Ele_LL, ROLE_LL, Ele_LR, ROLE_ON, nullptr, 0);
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int p = yes ? Ele_LL[B->LL] + Ele_LR[B->LR] : 0;
  host_elem(K_LLRR, B, A, Inv, M_LL, M_RR, c, sign_of(fn, Sign_Flag, Type, "LL_RR", "RR_LL", p), en);
}
void DMRG_MAKE_ELEM_LRRL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LR, CRS1 *M_RL, double c,
                                  int *Ele_RR, int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
  static const char *fn = "DMRG_MAKE_ELEM_LRRL_LLLRRRRL";
  if (g_rec) return record(K_LRRL, M_LR, M_RL, c, mode_of(fn, Sign_Flag, Type, "LR_RL", "RL_LR"), Ele_RR, ROLE_RR, Ele_LR, ROLE_ON, nullptr, 0);
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int p = yes ? Ele_RR[B->RR] + Ele_LR[B->LR] : 0;
  host_elem(K_LRRL, B, A, Inv, M_LR, M_RL, c, sign_of(fn, Sign_Flag, Type, "LR_RL", "RL_LR", p), en);
}
void DMRG_MAKE_ELEM_LRRR_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_LR, CRS1 *M_RR, double c,
                                  int *Ele_LR, int *en, char Type[], char Sign_Flag[]) {
  static const char *fn = "DMRG_MAKE_ELEM_LRRR_LLLRRRRL";
  if (g_rec) return record(K_LRRR, M_LR, M_RR, c, mode_of(fn, Sign_Flag, Type, "LR_RR", "RR_LR"), Ele_LR, ROLE_ON, nullptr, 0, nullptr, 0);
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int p = yes ? Ele_LR[B->LR] : 0;
  host_elem(K_LRRR, B, A, Inv, M_LR, M_RR, c, sign_of(fn, Sign_Flag, Type, "LR_RR", "RR_LR", p), en);
}
void DMRG_MAKE_ELEM_RRRL_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, CRS1 *M_RR, CRS1 *M_RL, double c,
                                  int *Ele_RR, int *en, char Type[], char Sign_Flag[]) {
  static const char *fn = "DMRG_MAKE_ELEM_RRRL_LLLRRRRL";
  if (g_rec) return record(K_RRRL, M_RR, M_RL, c, mode_of(fn, Sign_Flag, Type, "RR_RL", "RL_RR"), Ele_RR, ROLE_RR, nullptr, 0, nullptr, 0);
  if (std::fabs(c) < kZero) return;
  const bool yes = Sign_Flag && strcmp(Sign_Flag, "Yes") == 0;
  const int p = yes ? Ele_RR[B->RR] : 0;
  host_elem(K_RRRL, B, A, Inv, M_RR, M_RL, c, sign_of(fn, Sign_Flag, Type, "RR_RL", "RL_RR", p), en);
}
void DMRG_MAKE_ELEM_ZERO_LLLRRRRL(DMRG_BASIS_ONSITE *B, DMRG_A_BASIS *A, int *Inv, int *en) {
  if (g_rec) return record(K_ZERO, nullptr, nullptr, 0.0, 0, nullptr, 0, nullptr, 0, nullptr, 0);
  const int d = B->dim_onsite, dr = B->dim_RR;
  const long sup = (long)B->LL * d * dr * d + (long)B->LR * dr * d + (long)B->RR * d + B->RL;
  const int inv = Inv[sup];
  if (inv >= 0 && A->Inv_LLLRRRRL[inv] == -1) {
    const int ne = *en;
    A->LL_LLLRRRRL[ne] = (short)B->LL;
    A->LR_LLLRRRRL[ne] = (short)B->LR;
    A->RR_LLLRRRRL[ne] = (short)B->RR;
    A->RL_LLLRRRRL[ne] = (short)B->RL;
    A->Val_LLLRRRRL[ne] = 0.0;
    A->Inv_LLLRRRRL[inv] = ne;
    *en = ne + 1;
  }
}

// ------------------------------------------------------------- dmrgx_ API
int dmrgx_record_begin(void) {
  if (g_rec) return DMRGX_ERR_ARG;
  g_rec = new dmrgx_terms();
  return DMRGX_OK;
}

dmrgx_terms *dmrgx_record_end(void) {
  dmrgx_terms *t = g_rec;
  g_rec = nullptr;
  return t;
}

void dmrgx_terms_free(dmrgx_terms *t) { delete t; }
int dmrgx_terms_count(const dmrgx_terms *t) { return t ? (int)t->t.size() : -1; }

int dmrgx_superblock_create_arrays(int32_t n, int32_t dim_LL, int32_t dim_RR, int32_t d, const int16_t *LL,
                                   const int16_t *LR, const int16_t *RR, const int16_t *RL, int32_t n_ops,
                                   const CRS1 *ops, int32_t n_eles, const int32_t *const *eles,
                                   const int32_t *ele_len, int32_t n_terms, const dmrgx_term *terms,
                                   dmrgx_superblock **out) {
  if (!out || !LL || !LR || !RR || !RL || n_ops < 0 || n_eles < 0 || n_terms < 0) return DMRGX_ERR_ARG;
  return run_status([&] {
    std::vector<ArrOp> o(n_ops);
    for (int i = 0; i < n_ops; i++)
      o[i] = ArrOp{ops[i].row_dim, ops[i].col_dim, reinterpret_cast<const int64_t *>(ops[i].Row), ops[i].Col, ops[i].Val};
    std::vector<std::pair<const int32_t *, int32_t>> e(n_eles);
    for (int i = 0; i < n_eles; i++) e[i] = {eles[i], ele_len[i]};
    std::vector<DTerm> t(n_terms);
    for (int i = 0; i < n_terms; i++)
      t[i] = DTerm{terms[i].kind, terms[i].op1, terms[i].op2, -1, -1, terms[i].sign_mode, terms[i].ele0, terms[i].ele1,
                   terms[i].ele2, 0, terms[i].coeff};
    *out = create_superblock(n, dim_LL, dim_RR, d, LL, LR, RR, RL, o, e, t);
  });
}

int dmrgx_superblock_create_qn(const dmrgx_basis_qn *q, int32_t n_ops, const CRS1 *ops, int32_t n_eles,
                               const int32_t *const *eles, const int32_t *ele_len, int32_t n_terms,
                               const dmrgx_term *terms, dmrgx_superblock **out) {
  if (!q || !out || n_ops < 0 || n_eles < 0 || n_terms < 0) return DMRGX_ERR_ARG;
  return run_status([&] {
    std::vector<ArrOp> o(n_ops);
    for (int i = 0; i < n_ops; i++)
      o[i] = ArrOp{ops[i].row_dim, ops[i].col_dim, reinterpret_cast<const int64_t *>(ops[i].Row), ops[i].Col, ops[i].Val};
    std::vector<std::pair<const int32_t *, int32_t>> e(n_eles);
    for (int i = 0; i < n_eles; i++) e[i] = {eles[i], ele_len[i]};
    std::vector<DTerm> t(n_terms);
    for (int i = 0; i < n_terms; i++)
      t[i] = DTerm{terms[i].kind, terms[i].op1, terms[i].op2, -1, -1, terms[i].sign_mode, terms[i].ele0, terms[i].ele1,
                   terms[i].ele2, 0, terms[i].coeff};
    *out = create_superblock(0, 0, 0, 0, nullptr, nullptr, nullptr, nullptr, o, e, t, q);
  });
}

int dmrgx_superblock_n(const dmrgx_superblock *S, int32_t *n) {
  if (!S || !n) return DMRGX_ERR_ARG;
  *n = S->n;
  return DMRGX_OK;
}

int dmrgx_superblock_create(const DMRG_BASIS *B, const dmrgx_terms *T, dmrgx_superblock **out) {
  if (!B || !T || !out) return DMRGX_ERR_ARG;
  {
    std::vector<const CRS1 *> opp;
    std::vector<std::pair<const int *, int>> elp;
    std::vector<dmrgx_term> terms;
    auto op_id = [&](const CRS1 *m) {
      if (!m) return -1;
      for (size_t i = 0; i < opp.size(); i++)
        if (opp[i] == m) return (int)i;
      opp.push_back(m);
      return (int)opp.size() - 1;
    };
    auto ele_id = [&](const int *e, int role) {
      if (!e) return -1;
      const int len = role == ROLE_LL ? B->dim_LL : role == ROLE_RR ? B->dim_RR : B->dim_onsite;
      for (size_t i = 0; i < elp.size(); i++)
        if (elp[i].first == e) {
          elp[i].second = std::max(elp[i].second, len);
          return (int)i;
        }
      elp.push_back({e, len});
      return (int)elp.size() - 1;
    };
    for (const auto &t : T->t) {
      dmrgx_term x{};
      x.kind = t.kind;
      x.op1 = op_id(t.op1);
      x.op2 = op_id(t.op2);
      x.sign_mode = t.sign_mode;
      x.ele0 = t.sign_mode ? ele_id(t.ele[0], t.role[0]) : -1;
      x.ele1 = t.sign_mode ? ele_id(t.ele[1], t.role[1]) : -1;
      x.ele2 = t.sign_mode ? ele_id(t.ele[2], t.role[2]) : -1;
      x.coeff = t.coeff;
      terms.push_back(x);
    }
    std::vector<CRS1> ops(opp.size());
    for (size_t i = 0; i < opp.size(); i++) ops[i] = *opp[i];
    std::vector<const int32_t *> ep(elp.size());
    std::vector<int32_t> el(elp.size());
    for (size_t i = 0; i < elp.size(); i++) {
      ep[i] = elp[i].first;
      el[i] = elp[i].second;
    }
    return dmrgx_superblock_create_arrays(B->dim_LLLRRRRL, B->dim_LL, B->dim_RR, B->dim_onsite, B->LL_LLLRRRRL,
                                          B->LR_LLLRRRRL, B->RR_LLLRRRRL, B->RL_LLLRRRRL, (int)ops.size(),
                                          ops.data(), (int)ep.size(), ep.data(), el.data(), (int)terms.size(),
                                          terms.data(), out);
  }
}

int dmrgx_superblock_assemble(dmrgx_superblock *S, int64_t *nnz_lower, dmrgx_matrix **full_out) {
  if (!S) return DMRGX_ERR_ARG;
  return run_status([&] {
    const int64_t nnz = assemble_lower(*S);
    if (nnz_lower) *nnz_lower = nnz;
    if (full_out) *full_out = expand_last(*S);
  });
}

int dmrgx_superblock_download_lower(const dmrgx_superblock *S, int64_t *row, int32_t *col, double *val) {
  if (!S || S->lnnz < 0) return DMRGX_ERR_ARG;
  return run_status([&] {
    hipStream_t st = ctx().stream;
    ctx().copy_out(row, S->lrow.p, sizeof(int64_t) * ((size_t)S->n + 1));
    if (S->lnnz) {
      ctx().copy_out(col, S->lcol.p, sizeof(int32_t) * S->lnnz);
      ctx().copy_out(val, S->lval.p, sizeof(double) * S->lnnz);
    }
    HIPCHK(hipStreamSynchronize(st));
  });
}

void dmrgx_superblock_free(dmrgx_superblock *S) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  delete S;
}

CRS1 *dmrgx_get_ham_lllrrrrl(const DMRG_BASIS *B, const dmrgx_terms *T) {
  dmrgx_superblock *S = nullptr;
  int rc = dmrgx_superblock_create(B, T, &S);
  if (rc != DMRGX_OK) {
    printf("Error in GET_HAM_LLLRRRRL\n%s\n", dmrgx_last_error());
    exit(1);
  }
  int64_t nnz = 0;
  dmrgx_matrix *A = nullptr;
  rc = dmrgx_superblock_assemble(S, &nnz, &A);
  if (rc != DMRGX_OK) {
    printf("Error in GET_HAM_LLLRRRRL\n%s\n", dmrgx_last_error());
    exit(1);
  }
  // same allocation shape as GET_CRS1(dim, tot_elem_num + 1) (GET_HAM_LLLRRRRL.c:62)
  CRS1 *H = (CRS1 *)malloc(sizeof(CRS1));
  H->max_val = nnz + 1;
  H->max_row = S->n + 1;
  H->row_dim = S->n;
  H->col_dim = S->n;
  H->Val = (double *)calloc((size_t)nnz + 1, sizeof(double));
  H->Col = (int *)calloc((size_t)nnz + 1, sizeof(int));
  H->Row = (long *)calloc((size_t)S->n + 1, sizeof(long));
  if (!H->Val || !H->Col || !H->Row) {
    printf("Error in GET_HAM_LLLRRRRL\nhost allocation failed\n");
    exit(1);
  }
  rc = dmrgx_superblock_download_lower(S, reinterpret_cast<int64_t *>(H->Row), H->Col, H->Val);
  dmrgx_superblock_free(S);
  if (rc != DMRGX_OK) {
    printf("Error in GET_HAM_LLLRRRRL\n%s\n", dmrgx_last_error());
    exit(1);
  }
  set_pending(H, A);
  return H;
}

}  // extern "C"
