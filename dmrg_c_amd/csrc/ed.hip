// Exact diagonalization on the device: sector basis enumeration and Hamiltonian assembly.
//
// The reference's ED program builds its Hamiltonian row by row on the host
// (main/exact_diag/klm_vf/src/GET_HAM.c:11-92): for every basis code the model's
// MAKE_ELEMENT_HAM calls EXACT_MAKE_ELEM_ON / EXACT_MAKE_ELEM_INTER, which accumulate the
// row's contributions into an array kept sorted by target code (a quicksort after every
// insert, EXACT_MAKE_ELEM_ON.c:22); the targets are then found in the sorted basis by
// binary search and kept when (|val| > 1e-15 && inv >= 0) || inv == row.
//
// Here one wavefront assembles one row from a term list:
//   * lane s extracts the digit of site s once; the electron counts of the digits are
//     prefix-summed across the wave so any parity range is two shuffles;
//   * contributions are appended to LDS in the reference's call order (terms in order;
//     op1 row outer, op2 row inner, EXACT_MAKE_ELEM_INTER.c:17-18), values
//     coeff * v (on-site) or (sign * coeff) * v1 * v2 (intersite), evaluated left to right
//     with contraction off;
//   * the list is sorted by (target code, position) and each target's values are summed
//     left to right -- the reference's "first insertion sets, later ones add" (:23-27);
//   * each distinct target is looked up in the basis, the keep rule applied, and the row
//     written in ascending column order (the basis is ascending, so code order is column
//     order and SORT_COLUMN_CRS1 has nothing to do).
// Row, Col and Val come out bit-identical to the reference's element makers.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ed.h"

#pragma clang fp contract(off)

namespace dmrgx {
namespace {

constexpr double kZero = 1.0e-15;       // main/exact_diag/klm_vf/src/GET_HAM.c:20
constexpr int kMaxL = 63;               // sites held one per lane
constexpr int kPosBits = 12;            // key = code << 12 | list position
constexpr int64_t kMaxCode = (int64_t)1 << 51;
constexpr int kCapSmall = 256;          // 4 KiB of LDS per row
constexpr int kCapLarge = 4096;         // 64 KiB of LDS per row
constexpr int kTileThreads = 256, kTileCodes = kTileThreads * 16;

using DOp = dmrgx_ed::DOp;

struct EdView {
  int32_t L, d, nt;
  int64_t n;
  const int64_t *basis, *pw;
  const dmrgx_ed_term *terms;
  const DOp *ops;
  const int32_t *n_ele;
};

__device__ __forceinline__ unsigned lanes_below(uint64_t mask) {
  return __popcll(mask & ((1ull << (threadIdx.x & 63)) - 1ull));
}

template <int CAP, bool FILL>
__global__ __launch_bounds__(64) void k_ed_assemble(EdView v, const int32_t *row_list, int32_t n_list,
                                                    int64_t *counts, const int64_t *row_ptr, int32_t *col_out,
                                                    double *val_out, int32_t *overflow, int32_t *n_overflow) {
  __shared__ uint64_t key[CAP];
  __shared__ double val[CAP];
  const int lane = threadIdx.x;
  const int32_t item = blockIdx.x;
  if (item >= n_list) return;
  const int32_t r = row_list ? row_list[item] : item;
  const int64_t code = v.basis[r];
  // digit of site `lane` (exact/EXACT_FIND_SITE_STATE.c) and inclusive electron prefix
  int dig = 0, pre = 0;
  if (lane < v.L) {
    dig = (int)((code / v.pw[lane]) % v.d);
    pre = v.n_ele ? v.n_ele[dig] : 0;
  }
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(pre, off);
    if (lane >= off) pre += t;
  }
  int m = 0;
  for (int t = 0; t < v.nt; t++) {
    const dmrgx_ed_term tm = v.terms[t];
    const DOp a = v.ops[tm.op1];
    const int l1 = __shfl(dig, tm.site1);
    const int64_t b1 = a.row[l1], n1 = a.row[l1 + 1] - b1;
    const int64_t p1 = v.pw[tm.site1];
    DOp b = a;
    int l2 = 0;
    int64_t b2 = 0, n2 = 1, p2 = 0;
    double sc = tm.coeff;
    if (tm.kind == 1) {
      b = v.ops[tm.op2];
      l2 = __shfl(dig, tm.site2);
      b2 = b.row[l2];
      n2 = b.row[l2 + 1] - b2;
      p2 = v.pw[tm.site2];
      int s = tm.sign;
      if (tm.par_lo < tm.par_hi) {
        const int hi = __shfl(pre, tm.par_hi - 1);
        const int lo = tm.par_lo > 0 ? __shfl(pre, tm.par_lo - 1) : 0;
        if (((hi - lo) & 1) == 0) s = -s;  // even parity: MAKE_ELEMENT_HAM.c:31-37
      }
      sc = (double)s * tm.coeff;
    }
    const int64_t total = n1 * n2;
    for (int64_t base = 0; base < total; base += 64) {
      const int64_t k = base + lane;
      const bool have = k < total;
      int64_t tgt = 0;
      double prod = 0.0;
      if (have) {
        const int64_t i = k / n2;
        tgt = code + (int64_t)(a.col[b1 + i] - l1) * p1;
        prod = sc * a.val[b1 + i];
        if (tm.kind == 1) {
          const int64_t j = k - i * n2;
          tgt += (int64_t)(b.col[b2 + j] - l2) * p2;
          prod = prod * b.val[b2 + j];
        }
      }
      const uint64_t mask = __ballot(have);
      if (have) {
        const int pos = m + (int)lanes_below(mask);
        if (pos < CAP) {
          key[pos] = ((uint64_t)tgt << kPosBits) | (uint64_t)pos;
          val[pos] = prod;
        }
      }
      m += __popcll(mask);
    }
  }
  if (m > CAP) {
    if (!FILL && lane == 0) {
      if (overflow) overflow[atomicAdd(n_overflow, 1)] = r;
      else atomicAdd(n_overflow, 1);
    }
    return;
  }
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
  int out = 0;
  for (int base = 0; base < m; base += 64) {
    const int i = base + lane;
    bool keep = false;
    int64_t inv = -1;
    double s = 0.0;
    if (i < m) {
      const uint64_t c = key[i] >> kPosBits;
      const bool head = (i == 0) || ((key[i - 1] >> kPosBits) != c);
      if (head) {
        s = val[key[i] & ((1u << kPosBits) - 1)];
        for (int j = i + 1; j < m && (key[j] >> kPosBits) == c; j++) s = s + val[key[j] & ((1u << kPosBits) - 1)];
        inv = find_code(v.basis, v.n, (int64_t)c);
        keep = (fabs(s) > kZero && inv >= 0) || inv == r;  // GET_HAM.c:35,63
      }
    }
    const uint64_t km = __ballot(keep);
    if (FILL && keep) {
      const int64_t dst = row_ptr[r] + out + (int64_t)lanes_below(km);
      col_out[dst] = (int32_t)inv;
      val_out[dst] = s;
    }
    out += __popcll(km);
  }
  if (!FILL && lane == 0) counts[r] = out;
}

// ---------------------------------------------------------------- sector basis
struct QnTab {
  int32_t q1[16], q2[16];
  int32_t t1, t2, use2, d, L;
  int64_t total;
};

__device__ __forceinline__ bool in_sector(int64_t code, const QnTab &q) {
  int s1 = 0, s2 = 0;
  for (int k = 0; k < q.L; k++) {
    const int a = (int)(code % q.d);
    code /= q.d;
    s1 += q.q1[a];
    s2 += q.q2[a];
  }
  return s1 == q.t1 && (!q.use2 || s2 == q.t2);
}

// codes [tile*4096, tile*4096 + 4096): 16 consecutive codes per thread
__global__ __launch_bounds__(kTileThreads) void k_sector_count(QnTab q, int64_t *tile_cnt) {
  __shared__ int part[kTileThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kTileCodes + threadIdx.x * 16;
  int c = 0;
  for (int k = 0; k < 16; k++) c += (base + k < q.total) && in_sector(base + k, q);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kTileThreads / 64; w++) s += part[w];
    tile_cnt[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kTileThreads) void k_sector_fill(QnTab q, const int64_t *tile_off, int64_t *out) {
  __shared__ int wsum[kTileThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kTileCodes + threadIdx.x * 16;
  uint32_t bits = 0;
  for (int k = 0; k < 16; k++)
    if ((base + k < q.total) && in_sector(base + k, q)) bits |= 1u << k;
  const int c = __popc(bits);
  // exclusive scan of c over the block, in thread order
  int inc = c;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(inc, off);
    if (lane >= off) inc += t;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < w; k++) before += wsum[k];
  int64_t dst = tile_off[blockIdx.x] + before + inc - c;
  for (int k = 0; k < 16; k++)
    if (bits & (1u << k)) out[dst++] = base + k;
}

// ---------------------------------------------------------------- host
EdView view_of(const dmrgx_ed &E) {
  return EdView{E.L, E.d, (int32_t)E.hterms.size(), E.n, E.basis.p, E.pw.p, E.terms.p, E.ops.p,
                E.hn_ele.empty() ? nullptr : E.n_ele.p};
}

int64_t ipow_checked(int32_t d, int32_t L) {
  int64_t p = 1;
  for (int32_t s = 0; s < L; s++) {
    if (p > kMaxCode / d) throw Failure(DMRGX_ERR_ARG, "d^L too large for ED codes (limit 2^51)");
    p *= d;
  }
  return p;
}

// operators, terms and parity table; the basis is filled by the caller
std::unique_ptr<dmrgx_ed> ed_setup(int32_t L, int32_t d, int32_t n_ops, const CRS1 *ops, const int32_t *n_ele,
                                   int32_t n_terms, const dmrgx_ed_term *terms) {
  if (L <= 0 || L > kMaxL) throw Failure(DMRGX_ERR_ARG, "L must be in 1..63");
  if (d < 2 || d > 16) throw Failure(DMRGX_ERR_ARG, "dim_onsite must be in 2..16");
  if (n_ops < 0 || n_terms < 0 || (n_ops && !ops) || (n_terms && !terms))
    throw Failure(DMRGX_ERR_ARG, "bad operator or term arrays");
  ipow_checked(d, L);
  auto E = std::make_unique<dmrgx_ed>();
  E->L = L;
  E->d = d;
  for (int32_t i = 0; i < n_ops; i++) {
    const CRS1 &o = ops[i];
    if (o.row_dim != d || !o.Row || o.Row[0] != 0) throw Failure(DMRGX_ERR_ARG, "on-site operator must have d rows");
    const int64_t nnz = o.Row[d];
    if (nnz < 0 || (nnz && (!o.Col || !o.Val))) throw Failure(DMRGX_ERR_ARG, "bad on-site operator arrays");
    for (int32_t r = 0; r < d; r++)
      if (o.Row[r + 1] < o.Row[r]) throw Failure(DMRGX_ERR_ARG, "on-site operator Row not monotone");
    for (int64_t j = 0; j < nnz; j++)
      if (o.Col[j] < 0 || o.Col[j] >= d) throw Failure(DMRGX_ERR_ARG, "on-site operator column outside 0..d-1");
    E->hrow.emplace_back(o.Row, o.Row + d + 1);
    E->hcol.emplace_back(o.Col, o.Col + nnz);
    E->hval.emplace_back(o.Val, o.Val + nnz);
  }
  bool need_parity = false;
  for (int32_t t = 0; t < n_terms; t++) {
    const dmrgx_ed_term &x = terms[t];
    if (x.kind != 0 && x.kind != 1) throw Failure(DMRGX_ERR_ARG, "term kind must be 0 (on-site) or 1 (intersite)");
    if (x.site1 < 0 || x.site1 >= L || x.op1 < 0 || x.op1 >= n_ops) throw Failure(DMRGX_ERR_ARG, "term site/op out of range");
    if (x.kind == 1) {
      if (x.site2 < 0 || x.site2 >= L || x.op2 < 0 || x.op2 >= n_ops)
        throw Failure(DMRGX_ERR_ARG, "term site/op out of range");
      if (x.site2 == x.site1) throw Failure(DMRGX_ERR_ARG, "intersite term on one site");
      if (x.par_lo < 0 || x.par_hi > L || x.par_lo > x.par_hi) throw Failure(DMRGX_ERR_ARG, "bad parity range");
      need_parity |= x.par_lo < x.par_hi;
    }
    E->hterms.push_back(x);
  }
  if (need_parity && !n_ele) throw Failure(DMRGX_ERR_ARG, "parity range given without n_ele");
  if (n_ele) E->hn_ele.assign(n_ele, n_ele + d);
  hipStream_t st = ctx().stream;
  std::vector<int64_t> &pw = E->hpw;
  pw.assign(L + 1, 1);
  for (int s = 1; s <= L; s++) pw[s] = pw[s - 1] * d;
  E->pw.resize(L + 1);
  HIPCHK(hipMemcpyAsync(E->pw.p, pw.data(), sizeof(int64_t) * (L + 1), hipMemcpyHostToDevice, st));
  std::vector<DOp> hops(std::max(n_ops, 1));
  E->op_row.resize(n_ops);
  E->op_col.resize(n_ops);
  E->op_val.resize(n_ops);
  for (int32_t i = 0; i < n_ops; i++) {
    const size_t nnz = E->hcol[i].size();
    E->op_row[i].resize(d + 1);
    E->op_col[i].resize(std::max<size_t>(nnz, 1));
    E->op_val[i].resize(std::max<size_t>(nnz, 1));
    HIPCHK(hipMemcpyAsync(E->op_row[i].p, E->hrow[i].data(), sizeof(int64_t) * (d + 1), hipMemcpyHostToDevice, st));
    if (nnz) {
      HIPCHK(hipMemcpyAsync(E->op_col[i].p, E->hcol[i].data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(E->op_val[i].p, E->hval[i].data(), sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    }
    hops[i] = DOp{E->op_row[i].p, E->op_col[i].p, E->op_val[i].p};
  }
  E->ops.resize(hops.size());
  E->hops.assign(reinterpret_cast<const char *>(hops.data()), reinterpret_cast<const char *>(hops.data() + hops.size()));
  HIPCHK(hipMemcpyAsync(E->ops.p, E->hops.data(), sizeof(DOp) * hops.size(), hipMemcpyHostToDevice, st));
  E->terms.resize(std::max<size_t>(E->hterms.size(), 1));
  if (!E->hterms.empty())
    HIPCHK(hipMemcpyAsync(E->terms.p, E->hterms.data(), sizeof(dmrgx_ed_term) * E->hterms.size(),
                          hipMemcpyHostToDevice, st));
  E->n_ele.resize(d);
  if (n_ele) HIPCHK(hipMemcpyAsync(E->n_ele.p, E->hn_ele.data(), sizeof(int32_t) * d, hipMemcpyHostToDevice, st));
  // no wait: the host arrays copied from live in E until it is freed, and every user of the
  // device copies is queued behind the copies on this stream
  return E;
}

void ed_set_basis(dmrgx_ed &E, int64_t n, const int64_t *basis) {
  if (n <= 0 || n > INT32_MAX) throw Failure(DMRGX_ERR_ARG, "ED dimension must be in 1..2^31-1");
  const int64_t top = ipow_checked(E.d, E.L);
  for (int64_t i = 0; i < n; i++) {
    if (basis[i] < 0 || basis[i] >= top) throw Failure(DMRGX_ERR_ARG, "basis code outside 0..d^L-1");
    if (i && basis[i] <= basis[i - 1]) throw Failure(DMRGX_ERR_ARG, "basis codes must be strictly ascending");
  }
  E.n = n;
  E.basis.resize(n);
  ctx().copy_in(E.basis.p, basis, sizeof(int64_t) * n);
  HIPCHK(hipStreamSynchronize(ctx().stream));
}

void ed_enumerate(dmrgx_ed &E, const int32_t *qn1, int32_t t1, const int32_t *qn2, int32_t t2) {
  QnTab q{};
  for (int a = 0; a < E.d; a++) {
    q.q1[a] = qn1[a];
    q.q2[a] = qn2 ? qn2[a] : 0;
  }
  q.t1 = t1;
  q.t2 = t2;
  q.use2 = qn2 != nullptr;
  q.d = E.d;
  q.L = E.L;
  q.total = ipow_checked(E.d, E.L);
  if (q.total > ((int64_t)1 << 36)) throw Failure(DMRGX_ERR_ARG, "d^L above 2^36: enumerate the basis on the host");
  const int64_t tiles = (q.total + kTileCodes - 1) / kTileCodes;
  hipStream_t st = ctx().stream;
  DevBuf<int64_t> cnt(tiles), off(tiles + 1);
  k_sector_count<<<(unsigned)tiles, kTileThreads, 0, st>>>(q, cnt.p);
  HIPCHK(hipGetLastError());
  launch_exclusive_scan((int32_t)tiles, cnt.p, off.p, st);
  int64_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, off.p + tiles, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (n <= 0) throw Failure(DMRGX_ERR_ARG, "the requested sector is empty");
  if (n > INT32_MAX) throw Failure(DMRGX_ERR_ARG, "sector dimension above 2^31-1");
  E.n = n;
  E.basis.resize(n);
  k_sector_fill<<<(unsigned)tiles, kTileThreads, 0, st>>>(q, off.p, E.basis.p);
  HIPCHK(hipGetLastError());
  // no wait: the basis's users are queued behind the fill on this stream
}

std::unique_ptr<dmrgx_matrix> ed_assemble(const dmrgx_ed &E) {
  if (E.n <= 0) throw Failure(DMRGX_ERR_ARG, "ED problem has no basis");
  hipStream_t st = ctx().stream;
  const int32_t n = (int32_t)E.n;
  const EdView v = view_of(E);
  DevBuf<int64_t> counts(n);
  DevBuf<int32_t> over(n), n_over(2);
  HIPCHK(hipMemsetAsync(n_over.p, 0, 2 * sizeof(int32_t), st));
  k_ed_assemble<kCapSmall, false><<<n, 64, 0, st>>>(v, nullptr, n, counts.p, nullptr, nullptr, nullptr, over.p,
                                                     n_over.p);
  HIPCHK(hipGetLastError());
  // row pointers right away: they hold unless a row overflowed (then they are redone), so
  // the common case reads the overflow count and nnz in one round trip
  auto A = std::make_unique<dmrgx_matrix>();
  A->n = n;
  A->row.resize((size_t)n + 1);
  launch_exclusive_scan(n, counts.p, A->row.p, st);
  int64_t *hb = reinterpret_cast<int64_t *>(ctx().host_stage(4));
  int *h = reinterpret_cast<int *>(hb + 1);
  HIPCHK(hipMemcpyAsync(hb, A->row.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h, n_over.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  A->nnz = hb[0];
  const int n_large = h[0];
  if (n_large > 0) {
    k_ed_assemble<kCapLarge, false><<<n_large, 64, 0, st>>>(v, over.p, n_large, counts.p, nullptr, nullptr, nullptr,
                                                            nullptr, n_over.p + 1);
    HIPCHK(hipGetLastError());
    launch_exclusive_scan(n, counts.p, A->row.p, st);
    HIPCHK(hipMemcpyAsync(h + 1, n_over.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hb, A->row.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h[1] > 0) throw Failure(DMRGX_ERR_ARG, "an ED row has more than 4096 contributions");
    A->nnz = hb[0];
  }
  A->col.resize(std::max<int64_t>(A->nnz, 1));
  A->val.resize(std::max<int64_t>(A->nnz, 1));
  k_ed_assemble<kCapSmall, true><<<n, 64, 0, st>>>(v, nullptr, n, nullptr, A->row.p, A->col.p, A->val.p, nullptr,
                                                    nullptr);
  if (n_large > 0)
    k_ed_assemble<kCapLarge, true><<<n_large, 64, 0, st>>>(v, over.p, n_large, nullptr, A->row.p, A->col.p,
                                                           A->val.p, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  // no wait: the matrix's users are queued behind the fill on this stream (and the scratch
  // released here is only reused in stream order)
  default_schedule(*A);
  return A;
}

}  // namespace

// shared with ed_abi.cpp
std::unique_ptr<dmrgx_ed> ed_create(int32_t L, int32_t d, int32_t n_ops, const CRS1 *ops, const int32_t *n_ele,
                                    int32_t n_terms, const dmrgx_ed_term *terms) {
  return ed_setup(L, d, n_ops, ops, n_ele, n_terms, terms);
}
void ed_basis(dmrgx_ed &E, int64_t n, const int64_t *basis) { ed_set_basis(E, n, basis); }
void ed_sector(dmrgx_ed &E, const int32_t *qn1, int32_t t1, const int32_t *qn2, int32_t t2) {
  ed_enumerate(E, qn1, t1, qn2, t2);
}
std::unique_ptr<dmrgx_matrix> ed_matrix(const dmrgx_ed &E) { return ed_assemble(E); }

EdRows ed_rows(const dmrgx_ed &out, const dmrgx_ed &in, int32_t op, int32_t site, int32_t sign_flag) {
  if (op < 0 || op >= (int32_t)in.hrow.size()) throw Failure(DMRGX_ERR_ARG, "operator index out of range");
  if (site < 0 || site >= in.L || out.L != in.L || out.d != in.d) throw Failure(DMRGX_ERR_ARG, "site or lattice mismatch");
  if (sign_flag && in.hn_ele.empty()) throw Failure(DMRGX_ERR_ARG, "fermion sign requested without n_ele");
  int64_t pw = 1;
  for (int s = 0; s < site; s++) pw *= in.d;
  return EdRows{in.d, site, sign_flag ? 1 : 0, pw, in.n, out.basis.p, in.basis.p, in.op_row[op].p,
                in.op_col[op].p, in.op_val[op].p, in.n_ele.p};
}

}  // namespace dmrgx

using namespace dmrgx;

extern "C" {

int dmrgx_ed_create(int32_t L, int32_t d, int64_t n, const int64_t *basis, int32_t n_ops, const CRS1 *ops,
                    const int32_t *n_ele, int32_t n_terms, const dmrgx_ed_term *terms, dmrgx_ed **out) {
  if (!out || !basis) return DMRGX_ERR_ARG;
  return run_status([&] {
    auto E = ed_create(L, d, n_ops, ops, n_ele, n_terms, terms);
    ed_basis(*E, n, basis);
    *out = E.release();
  });
}

int dmrgx_ed_create_sector(int32_t L, int32_t d, const int32_t *qn1, int32_t target1, const int32_t *qn2,
                           int32_t target2, int32_t n_ops, const CRS1 *ops, const int32_t *n_ele, int32_t n_terms,
                           const dmrgx_ed_term *terms, dmrgx_ed **out) {
  if (!out || !qn1) return DMRGX_ERR_ARG;
  return run_status([&] {
    auto E = ed_create(L, d, n_ops, ops, n_ele, n_terms, terms);
    ed_sector(*E, qn1, target1, qn2, target2);
    *out = E.release();
  });
}

int64_t dmrgx_ed_dim(const dmrgx_ed *E) { return E ? E->n : -1; }

int dmrgx_ed_download_basis(const dmrgx_ed *E, int64_t *basis) {
  if (!E || !basis) return DMRGX_ERR_ARG;
  return run_status([&] {
    ctx().copy_out(basis, E->basis.p, sizeof(int64_t) * E->n);
    HIPCHK(hipStreamSynchronize(ctx().stream));
  });
}

int dmrgx_ed_assemble(const dmrgx_ed *E, dmrgx_matrix **out) {
  if (!E || !out) return DMRGX_ERR_ARG;
  return run_status([&] { *out = ed_matrix(*E).release(); });
}

int dmrgx_ed_apply_onsite(const dmrgx_ed *E_out, const dmrgx_ed *E_in, int32_t op, int32_t site, int32_t sign_flag,
                          const double *x_dev, double *y_dev) {
  if (!E_out || !E_in || !x_dev || !y_dev) return DMRGX_ERR_ARG;
  return run_status([&] {
    Reduce R;
    launch_ed_apply((int32_t)E_out->n, ed_rows(*E_out, *E_in, op, site, sign_flag), x_dev, y_dev, nullptr, nullptr, R,
                    ctx().stream);
    HIPCHK(hipStreamSynchronize(ctx().stream));
  });
}

void dmrgx_ed_free(dmrgx_ed *E) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  delete E;
}

}  // extern "C"
