// Device kernels of the superblock eigensolve for gfx950 (wave64, 256-thread workgroups).
//
// * CSR SpMV with G lanes per row (G in {4..64}, picked from the mean row length) and a
//   fused per-row epilogue (Lanczos normalise + alpha dot, CG A*p + shift + p.Ap, the
//   L1 residual of the inverse-iteration convergence check).
// * Streaming vector kernels for the Lanczos / CG updates with the next reduction fused.
// * Every reduction is deterministic: each block stores one partial with a plain store and
//   a one-block finish kernel sums the partials in index order.  (A last-block ticket with
//   agent-scope acquire/release atomics would make every block write back its L2 for the
//   other XCDs to see; with ~36k blocks per SpMV that costs more than the product.)
//   Results do not depend on scheduling, so a solve gives the same bits alone or next to
//   other work.
// * Lower-triangle -> full CSR expansion (counting, scan, scatter, per-row sort).

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "internal.h"
#include "reduce.h"

namespace dmrgx {
namespace {

// One block: sum of the nblk partials in index order, handed to f.finish().
template <class Fin>
__global__ __launch_bounds__(kBlock) void k_finish(Fin f, const double *partial, int nblk) {
  if (f.skip()) return;
  const double t = finish_sum(partial, nblk);
  if (threadIdx.x == 0) f.finish(t);
}

// ------------------------------------------------------------------ SpMV + epilogue
// Row sources: each returns this lane's share of sum_j H_rj x_j for row r.
template <class VT, int V = 0>
struct CsrRows {
  const int64_t *__restrict__ row;
  const int32_t *__restrict__ col;
  const VT *__restrict__ val;
  int64_t nnz;
  static constexpr bool kTimed = true;
  // algorithmic bytes of one y = H x: values + columns + row pointers + x once + y once
  double alg_bytes(int32_t n) const {
    return (double)nnz * (sizeof(VT) + 4) + (double)(n + 1) * 8 + 2.0 * n * 8;
  }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    return csr_lane_sum_v<G, V>(row, col, val, r, lane, x);
  }
};

template <int V>
struct DictRows {
  const int64_t *__restrict__ row;
  const uint64_t *__restrict__ ent;
  const double *__restrict__ table;
  int64_t nnz, n_values;
  static constexpr bool kTimed = true;
  // algorithmic bytes of one y = H x: packed entries + table once + row pointers + x + y
  double alg_bytes(int32_t n) const {
    return (double)nnz * 8 + (double)n_values * 8 + (double)(n + 1) * 8 + 2.0 * n * 8;
  }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    return dict_lane_sum_v<G, V>(row, ent, table, r, lane, x);
  }
};

template <class VT, int V>
struct DeltaRows {
  const int64_t *__restrict__ row;
  const int32_t *__restrict__ first;
  const uint16_t *__restrict__ dcol;
  const VT *__restrict__ val;
  int64_t nnz;
  static constexpr bool kTimed = true;
  // algorithmic bytes of one y = H x: values + 2-byte differences + row pointers + first
  // columns + x + y
  double alg_bytes(int32_t n) const {
    return (double)nnz * (sizeof(VT) + 2) + (double)(n + 1) * 8 + (double)n * 4 + 2.0 * n * 8;
  }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    return delta_lane_sum_v<G, V>(row, first, dcol, val, r, lane, x);
  }
};

// Sliced form with P lanes per row (sell_lane_sum); only launched with G == P.
template <class VT, int P, bool NT = false>
struct SellRows {
  const int64_t *__restrict__ sptr;
  const int32_t *__restrict__ slen;
  const int32_t *__restrict__ scol;
  const VT *__restrict__ sval;
  int64_t nnz;
  int32_t nslices;
  static constexpr bool kTimed = true;
  // algorithmic bytes of one y = H x: the entries' values and columns (padding not
  // counted) + row lengths + slice offsets + x once + y once
  double alg_bytes(int32_t n) const {
    return (double)nnz * (sizeof(VT) + 4) + (double)n * 4 + (double)(nslices + 1) * 8 + 2.0 * n * 8;
  }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    static_assert(G == P, "sliced rows are read with P lanes per row");
    return sell_lane_sum<P, VT, NT>(sptr, slen, scol, sval, r, lane, x);
  }
};

// The sliced form with 16-bit columns (sell16_lane_sum, or sell16_lane_sum_pipe with PIPE);
// only launched with G == P.
template <class VT, int P, bool NT = false, bool PIPE = false>
struct Sell16Rows {
  const int64_t *__restrict__ sptr;
  const int32_t *__restrict__ slen;
  const uint16_t *__restrict__ sd16;
  const int32_t *__restrict__ sfirst;
  const int64_t *__restrict__ septr;
  const int32_t *__restrict__ sexc;
  const VT *__restrict__ sval;
  int64_t nnz, nexc;
  int32_t nslices;
  static constexpr bool kTimed = true;
  // algorithmic bytes of one y = H x: values and 2-byte column differences of the entries
  // (padding not counted) + row lengths + slice offsets + per-sequence first column and
  // exception offset + the exceptions + x once + y once
  double alg_bytes(int32_t n) const {
    return (double)nnz * (sizeof(VT) + 2) + (double)n * 4 + (double)(nslices + 1) * 8 +
           (double)nslices * 64 * 12 + (double)nexc * 4 + 2.0 * n * 8;
  }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    static_assert(G == P, "sliced rows are read with P lanes per row");
    if constexpr (PIPE)
      return sell16_lane_sum_pipe<P, VT, NT>(sptr, slen, sd16, sfirst, septr, sexc, sval, r, lane, x);
    else
      return sell16_lane_sum<P, VT, NT>(sptr, slen, sd16, sfirst, septr, sexc, sval, r, lane, x);
  }
};

// The sliced form with 16-bit columns and 16-bit value indices (format 4, sellx_lane_sum);
// only launched with G == P.
template <class VT, int P, bool NT = false>
struct SellxRows {
  const int64_t *__restrict__ sptr;
  const int32_t *__restrict__ slen;
  const uint16_t *__restrict__ sd16;
  const int32_t *__restrict__ sfirst;
  const int64_t *__restrict__ septr;
  const int32_t *__restrict__ sexc;
  const uint16_t *__restrict__ sidx;
  const VT *__restrict__ table;
  int64_t nnz, nexc, nvalues;
  int32_t nslices;
  static constexpr bool kTimed = true;
  // algorithmic bytes of one y = H x: 2-byte value index and 2-byte column difference per
  // entry (padding not counted) + row lengths + slice offsets + per-sequence first column
  // and exception offset + the exceptions + the value table once + x once + y once
  double alg_bytes(int32_t n) const {
    return (double)nnz * 4 + (double)n * 4 + (double)(nslices + 1) * 8 + (double)nslices * 64 * 12 +
           (double)nexc * 4 + (double)nvalues * sizeof(VT) + 2.0 * n * 8;
  }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    static_assert(G == P, "sliced rows are read with P lanes per row");
    return sellx_lane_sum<P, VT, NT>(sptr, slen, sd16, sfirst, septr, sexc, sidx, table, r, lane, x);
  }
};

// Row groups of kBlock/G rows are dealt to the blocks (the grid may be capped, so a block
// takes several groups) round-robin, or with xcd set in contiguous ranges over XCD-mapped
// logical blocks; each thread's epilogue contributions are summed in the order of its
// groups, then per block, and the block's partial goes to its logical slot.
// Epilogues with kLateBegin load their scalar's inputs in prefetch() and finish it in
// begin() only after the block's first row group has been gathered.
template <class E, class = void>
struct MaySkip : std::false_type {};
template <class E>
struct MaySkip<E, std::void_t<decltype(E::kMaySkip)>> : std::bool_constant<E::kMaySkip> {};

template <class E, class = void>
struct LateBegin : std::false_type {};
template <class E>
struct LateBegin<E, std::void_t<decltype(E::kLateBegin)>> : std::bool_constant<E::kLateBegin> {};

template <int G, class Rows, class Epi>
__global__ __launch_bounds__(kBlock) void k_apply(int32_t n, Rows rows, const double *__restrict__ x, Epi epi,
                                                  double *partial, int xcd) {
  if (epi.skip()) return;
  constexpr int RPB = kBlock / G;
  const int g = threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int32_t lb = xcd ? xcd_logical(blockIdx.x, gridDim.x) : (int32_t)blockIdx.x;
  const GroupRange gr = group_range(lb, gridDim.x, (n + RPB - 1) / RPB, xcd != 0);
  double c = 0.0;
  int32_t gi = gr.start;
  if constexpr (LateBegin<Epi>::value) {
    typename Epi::Pre pf;
    epi.prefetch(pf);
    const bool have = gi < gr.end;   // the same for the whole block
    const int32_t r = gi * RPB + g;
    double s = 0.0;
    if (have && r < n) s = rows.template lane_sum<G>(r, lane, x);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    epi.begin(pf);
    if (have) {
      if (lane == 0 && r < n) c += epi.row(r, s);
      gi += gr.step;
    }
  } else {
    epi.begin();
  }
  for (; gi < gr.end; gi += gr.step) {
    const int32_t r = gi * RPB + g;
    double s = 0.0;
    if (r < n) s = rows.template lane_sum<G>(r, lane, x);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (lane == 0 && r < n) c += epi.row(r, s);
  }
  if constexpr (Epi::kReduce) block_partial_to(c, partial + lb);
}

struct EpiPlain {
  double *y;
  static constexpr bool kReduce = false;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ double row(int32_t r, double s) { y[r] = s; return 0.0; }
  __device__ void finish(double) {}
};

// K_a of a Lanczos step: v_k = w / ||w||, y = H v_k, alpha_k = <y, v_k>
// (sml/LANCZOS_SLOW_SYM.c:114-122 with NORMALIZE folded into the product).
struct EpiLanczosA {
  const double *w;
  double *nrm2;
  double *vk, *y, *alpha;
  int k;
  double shift;   // operator H + shift I (the minimum-residual solve of inverse iteration)
  double scale;
  // With pre set, |w|^2 is the sum of the previous k_lanczos_b's pre_n block partials, taken
  // by every wavefront exactly as the finish kernel would (block 0 also stores *nrm2 and
  // beta[k]); without, it is read from *nrm2.
  const double *pre = nullptr;
  int pre_n = 0;
  double *beta = nullptr;
  // the partials loaded ahead of the gathers (those past four per lane are read after them);
  // held by the kernel as a local, so that they stay in registers
  using Pre = WavePartialSum<1>;
  static constexpr bool kReduce = true;
  static constexpr bool kLateBegin = true;
  __device__ bool skip() const { return false; }
  __device__ void prefetch(Pre &ps) {
    if (pre) ps.load(pre, pre_n);
    else scale = *nrm2;
  }
  __device__ void begin(const Pre &ps) {
    if (!pre) {
      scale = 1.0 / sqrt(scale);
      return;
    }
    const double t = ps.sum(pre, pre_n);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *nrm2 = t;
      beta[k] = sqrt(t);
    }
    scale = 1.0 / sqrt(t);
  }
  __device__ double row(int32_t r, double s) {
    const double xs = scale * w[r];
    vk[r] = xs;
    const double yv = scale * s + shift * xs;
    y[r] = yv;
    return yv * xs;
  }
  __device__ void finish(double t) { alpha[k] = t; }
};

// Minimum-residual start (sml/MINIMUM_RESIDUAL.c:31-38): v = b - (H + shift) x, |v|^2
struct EpiMrInit {
  const double *b, *x;
  double *v, *nrm2;
  double shift;
  static constexpr bool kReduce = true;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ double row(int32_t r, double s) {
    const double rv = b[r] - (s + shift * x[r]);
    v[r] = rv;
    return rv * rv;
  }
  __device__ void finish(double t) { *nrm2 = t; }
};

// CG start (sml/CONJUGATE_GRADIENT_SYM.c:202-210): x = x0/||x0||, r = b - (H + shift) x, p = r
struct EpiCgInit {
  const double *nrm2, *x_in, *b;
  double *x_out, *r_out, *p_out, *rr0;
  double shift, scale;
  static constexpr bool kReduce = true;
  __device__ bool skip() const { return false; }
  __device__ void begin() { scale = 1.0 / sqrt(*nrm2); }
  __device__ double row(int32_t r, double s) {
    const double xs = scale * x_in[r];
    x_out[r] = xs;
    const double rv = b[r] - (scale * s + shift * xs);
    r_out[r] = rv;
    p_out[r] = rv;
    return rv * rv;
  }
  __device__ void finish(double t) { *rr0 = t; }
};

// CG product (CONJUGATE_GRADIENT_SYM.c:214-218, shift as DMRG_CONJUGATE_GRADIENT_OBC.c:40)
struct EpiCgAp {
  const double *p;
  double *y, *pap;
  const int *status;
  double shift;
  static constexpr bool kReduce = true;
  static constexpr bool kMaySkip = true;   // launches past convergence return at once
  __device__ bool skip() const { return status[0] != 0; }
  __device__ void begin() {}
  __device__ double row(int32_t r, double s) {
    const double yv = s + shift * p[r];
    y[r] = yv;
    return p[r] * yv;
  }
  __device__ void finish(double t) { *pap = t; }
};

// CONVERGE_CHECK_SYM (sml/CONVERGE_CHECK_SYM.c:24-33): sum_i |E v_i - (H v)_i|
struct EpiResidualL1 {
  const double *v;
  double *out;
  double eig;
  static constexpr bool kReduce = true;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ double row(int32_t r, double s) { return fabs(eig * v[r] - s); }
  __device__ void finish(double t) { *out = t; }
};

// y[r] = (O x)_r (if y) and sum_r u[r] (O x)_r (if u)
struct EpiStoreDot {
  double *y;
  const double *u;
  double *out;
  static constexpr bool kReduce = true;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ double row(int32_t r, double s) {
    if (y) y[r] = s;
    return u ? u[r] * s : 0.0;
  }
  __device__ void finish(double t) {
    if (out) *out = t;
  }
};

// Launches the product with its epilogue; with fin the reduction's finish kernel follows,
// without it the caller's next kernel sums the partials.  Returns the number of blocks
// (= partials).
// The launch with G lanes per row and the block partials at part (spmv_blocks(n, G,
// grid_cap) doubles when the epilogue reduces); timed = false keeps it out of the live SpMV
// timing.
template <int G, class Rows, class Epi>
int apply_rows_g(int32_t n, int grid_cap, const Rows &rows, const double *x, const Epi &epi, double *part,
                 hipStream_t s, int xcd, bool fin, bool timed) {
  const int blocks = spmv_blocks(n, G, grid_cap);
  if (blocks == 0) return 0;
  Context &C = ctx();
  std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
  // live timing brackets only products that always run in full (a skipped launch would
  // count its bytes over a few microseconds)
  if (timed && Rows::kTimed && !MaySkip<Epi>::value && C.prof_take()) {
    ev = C.prof_slot();
    HIPCHK(hipEventRecord(ev->first, s));
    C.prof_bytes += rows.alg_bytes(n);
  }
  k_apply<G, Rows, Epi><<<blocks, kBlock, 0, s>>>(n, rows, x, epi, part, xcd);
  HIPCHK(hipGetLastError());
  if (ev) HIPCHK(hipEventRecord(ev->second, s));
  if (Epi::kReduce && fin) {
    k_finish<Epi><<<1, kBlock, 0, s>>>(epi, part, blocks);
    HIPCHK(hipGetLastError());
  }
  return blocks;
}

template <class Rows, class Epi>
int apply_rows_into(int32_t n, int lanes, int grid_cap, const Rows &rows, const double *x, const Epi &epi,
                    double *part, hipStream_t s, int xcd, bool fin, bool timed) {
  switch (lanes) {
#define L_CASE(G) \
  case G: return apply_rows_g<G>(n, grid_cap, rows, x, epi, part, s, xcd, fin, timed);
    L_CASE(1) L_CASE(2) L_CASE(4) L_CASE(8) L_CASE(16) L_CASE(32) L_CASE(64)
#undef L_CASE
    default: throw Failure(DMRGX_ERR_ARG, "bad lanes per row");
  }
}

template <class Rows, class Epi>
int apply_rows(int32_t n, int lanes, int grid_cap, const Rows &rows, const double *x, const Epi &epi, Reduce *R,
               hipStream_t s, int xcd = 0, bool fin = true) {
  const int blocks = spmv_blocks(n, lanes, grid_cap);
  if (blocks == 0) return 0;
  double *part = nullptr;
  if (Epi::kReduce) {
    R->ensure(blocks);
    part = R->partial.p;
  }
  return apply_rows_into(n, lanes, grid_cap, rows, x, epi, part, s, xcd, fin, true);
}

template <class Epi>
int apply_dict(const Operator &op, const double *x, const Epi &epi, Reduce *R, hipStream_t s, bool fin) {
  const dmrgx_matrix &A = *op.A;
  const int xcd = (op.load & kLoadXcd) ? 1 : 0;
  switch (op.load & (kLoadXcd - 1)) {
#define V_CASE(V)                                                                                           \
  case V:                                                                                                   \
    return apply_rows(op.n, op.lanes, op.grid, DictRows<V>{A.row.p, A.ent.p, A.table.p, A.nnz, A.n_values}, x, \
                      epi, R, s, xcd, fin);
    V_CASE(0) V_CASE(1) V_CASE(2) V_CASE(3)
#undef V_CASE
    default: throw Failure(DMRGX_ERR_ARG, "bad SpMV load variant");
  }
}

template <class VT, class Epi>
int apply_delta(const Operator &op, const VT *val, const double *x, const Epi &epi, Reduce *R, hipStream_t s,
                bool fin) {
  const dmrgx_matrix &A = *op.A;
  const int xcd = (op.load & kLoadXcd) ? 1 : 0;
  switch (op.load & (kLoadXcd - 1)) {
#define V_CASE(V)                                                                                        \
  case V:                                                                                                \
    return apply_rows(op.n, op.lanes, op.grid, DeltaRows<VT, V>{A.row.p, A.first.p, A.dcol.p, val, A.nnz}, x,  \
                      epi, R, s, xcd, fin);
    V_CASE(0) V_CASE(1) V_CASE(2) V_CASE(3)
#undef V_CASE
    default: throw Failure(DMRGX_ERR_ARG, "bad SpMV load variant");
  }
}

template <class VT, class Epi>
int apply_csr(const Operator &op, const VT *val, const double *x, const Epi &epi, Reduce *R, hipStream_t s,
              bool fin) {
  const dmrgx_matrix &A = *op.A;
  const int xcd = (op.load & kLoadXcd) ? 1 : 0;
  switch (op.load & (kLoadXcd - 1)) {
#define V_CASE(V)                                                                                              \
  case V:                                                                                                      \
    return apply_rows(op.n, op.lanes, op.grid, CsrRows<VT, V>{A.row.p, A.col.p, val, A.nnz}, x, epi, R, s, xcd, \
                      fin);
    V_CASE(0) V_CASE(1) V_CASE(2) V_CASE(3)
#undef V_CASE
    default: throw Failure(DMRGX_ERR_ARG, "bad SpMV load variant");
  }
}

template <class VT, int P, class Epi>
int apply_sell_p(const Operator &op, const VT *sval, const double *x, const Epi &epi, Reduce *R, hipStream_t s,
                 bool fin) {
  const dmrgx_matrix &A = *op.A;
  constexpr int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  const int32_t nsl = (op.n + 64 / P - 1) / (64 / P);
  const int blocks = spmv_blocks(op.n, P, op.grid);
  double *part = nullptr;
  if (Epi::kReduce && blocks > 0) {
    R->ensure(blocks);
    part = R->partial.p;
  }
  const int xcd = (op.load & kLoadXcd) ? 1 : 0;
  if ((op.load & kLoadPair) && A.sell_nexc[lp] >= 0) {   // 16-bit columns (built by apply_op)
    if (op.load & kLoadPipe) {
      if (op.load & kLoadDeep) {
        const Sell16Rows<VT, P, true, true> rows{A.sell_ptr[lp].p, A.sell_len.p,      A.sell_d16[lp].p,
                                                 A.sell_first[lp].p, A.sell_eptr[lp].p, A.sell_exc[lp].p,
                                                 sval,              A.nnz,             A.sell_nexc[lp],
                                                 nsl};
        return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
      }
      const Sell16Rows<VT, P, false, true> rows{A.sell_ptr[lp].p, A.sell_len.p,      A.sell_d16[lp].p,
                                                A.sell_first[lp].p, A.sell_eptr[lp].p, A.sell_exc[lp].p,
                                                sval,              A.nnz,             A.sell_nexc[lp],
                                                nsl};
      return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
    }
    if (op.load & kLoadDeep) {
      const Sell16Rows<VT, P, true> rows{A.sell_ptr[lp].p, A.sell_len.p,   A.sell_d16[lp].p, A.sell_first[lp].p,
                                         A.sell_eptr[lp].p, A.sell_exc[lp].p, sval, A.nnz, A.sell_nexc[lp], nsl};
      return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
    }
    const Sell16Rows<VT, P> rows{A.sell_ptr[lp].p, A.sell_len.p,   A.sell_d16[lp].p, A.sell_first[lp].p,
                                 A.sell_eptr[lp].p, A.sell_exc[lp].p, sval, A.nnz, A.sell_nexc[lp], nsl};
    return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
  }
  if (op.load & kLoadDeep) {
    const SellRows<VT, P, true> rows{A.sell_ptr[lp].p, A.sell_len.p, A.sell_col[lp].p, sval, A.nnz, nsl};
    return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
  }
  const SellRows<VT, P> rows{A.sell_ptr[lp].p, A.sell_len.p, A.sell_col[lp].p, sval, A.nnz, nsl};
  return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
}

template <class VT, int P, class Epi>
int apply_sellx_p(const Operator &op, const VT *table, const double *x, const Epi &epi, Reduce *R, hipStream_t s,
                  bool fin) {
  const dmrgx_matrix &A = *op.A;
  constexpr int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  const int32_t nsl = (op.n + 64 / P - 1) / (64 / P);
  const int blocks = spmv_blocks(op.n, P, op.grid);
  double *part = nullptr;
  if (Epi::kReduce && blocks > 0) {
    R->ensure(blocks);
    part = R->partial.p;
  }
  const int xcd = (op.load & kLoadXcd) ? 1 : 0;
  if (op.load & kLoadDeep) {
    const SellxRows<VT, P, true> rows{A.sell_ptr[lp].p,  A.sell_len.p,     A.sell_d16[lp].p,
                                      A.sell_first[lp].p, A.sell_eptr[lp].p, A.sell_exc[lp].p,
                                      A.sell_idx16[lp].p, table,            A.nnz,
                                      A.sell_nexc[lp],    A.n_values,       nsl};
    return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
  }
  const SellxRows<VT, P> rows{A.sell_ptr[lp].p,  A.sell_len.p,     A.sell_d16[lp].p, A.sell_first[lp].p,
                              A.sell_eptr[lp].p, A.sell_exc[lp].p, A.sell_idx16[lp].p, table,
                              A.nnz,             A.sell_nexc[lp],  A.n_values,       nsl};
  return apply_rows_g<P>(op.n, op.grid, rows, x, epi, part, s, xcd, fin, true);
}

template <class Epi>
int apply_sellx(const Operator &op, bool fp32, const double *x, const Epi &epi, Reduce *R, hipStream_t s, bool fin) {
  const dmrgx_matrix &A = *op.A;
  if (fp32) {
    switch (op.lanes) {
      case 1: return apply_sellx_p<float, 1>(op, A.table32.p, x, epi, R, s, fin);
      case 2: return apply_sellx_p<float, 2>(op, A.table32.p, x, epi, R, s, fin);
      case 4: return apply_sellx_p<float, 4>(op, A.table32.p, x, epi, R, s, fin);
    }
  } else {
    switch (op.lanes) {
      case 1: return apply_sellx_p<double, 1>(op, A.table.p, x, epi, R, s, fin);
      case 2: return apply_sellx_p<double, 2>(op, A.table.p, x, epi, R, s, fin);
      case 4: return apply_sellx_p<double, 4>(op, A.table.p, x, epi, R, s, fin);
    }
  }
  throw Failure(DMRGX_ERR_ARG, "sliced form needs 1, 2 or 4 lanes per row");
}

template <class Epi>
int apply_sell(const Operator &op, bool fp32, const double *x, const Epi &epi, Reduce *R, hipStream_t s, bool fin) {
  const dmrgx_matrix &A = *op.A;
  const int lp = op.lanes == 1 ? 0 : (op.lanes == 2 ? 1 : 2);
  if (fp32) {
    const float *v = A.sell_val32[lp].p;
    switch (op.lanes) {
      case 1: return apply_sell_p<float, 1>(op, v, x, epi, R, s, fin);
      case 2: return apply_sell_p<float, 2>(op, v, x, epi, R, s, fin);
      case 4: return apply_sell_p<float, 4>(op, v, x, epi, R, s, fin);
    }
  } else {
    const double *v = A.sell_val[lp].p;
    switch (op.lanes) {
      case 1: return apply_sell_p<double, 1>(op, v, x, epi, R, s, fin);
      case 2: return apply_sell_p<double, 2>(op, v, x, epi, R, s, fin);
      case 4: return apply_sell_p<double, 4>(op, v, x, epi, R, s, fin);
    }
  }
  throw Failure(DMRGX_ERR_ARG, "sliced form needs 1, 2 or 4 lanes per row");
}

// Rows of a product computed beforehand (the Kronecker form writes H x into z): one lane per
// row returns z[r], so that the solvers' fused epilogues run on it unchanged.
struct PreRows {
  const double *z, *zl, *zc;   // the product is (z[r] + zl[r]) + zc[r]
  static constexpr bool kTimed = false;
  double alg_bytes(int32_t) const { return 0.0; }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__) const {
    return lane == 0 ? (z[r] + zl[r]) + zc[r] : 0.0;
  }
};

template <class Epi>
int apply_op(const Operator &op, bool fp32, const double *x, const Epi &epi, Reduce *R, hipStream_t s,
             bool fin = true) {
  if (op.obc) return apply_rows(op.n, op.lanes, op.grid, op.obc->rows(), x, epi, R, s, 0, fin);
  if (op.kron) {
    if (fp32) throw Failure(DMRGX_ERR_ARG, "the Kronecker form has no FP32 values");
    op.kron->product(x, s);
    return apply_rows(op.n, 1, op.grid, PreRows{op.kron->z.p, op.kron->zl.p, op.kron->zc.p}, x, epi, R, s, 0, fin);
  }
  const int fmt = load_format(op.load);
  if (fmt == kFmtSellIdx && ensure_sell(*op.A, op.lanes, false)) {
    ensure_sell16(*op.A, op.lanes);
    if (ensure_sellx(*op.A, op.lanes, fp32)) return apply_sellx(op, fp32, x, epi, R, s, fin);
    // no value table of at most 65536 values: format 3 with 16-bit columns (the same bits)
    Operator o3 = op;
    o3.load = (op.load & ~kLoadFmtMask) | kLoadSell | kLoadPair;
    ensure_sell(*op.A, op.lanes, fp32);
    return apply_sell(o3, fp32, x, epi, R, s, fin);
  }
  if (fmt == kFmtSell && ensure_sell(*op.A, op.lanes, fp32)) {
    if (op.load & kLoadPair) ensure_sell16(*op.A, op.lanes);
    return apply_sell(op, fp32, x, epi, R, s, fin);
  }
  const bool delta = fmt == kFmtDelta && ensure_delta(*op.A);
  if (fp32) {
    if (!op.A->val32.p) throw Failure(DMRGX_ERR_ARG, "FP32 values not prepared");
    if (delta) return apply_delta<float>(op, op.A->val32.p, x, epi, R, s, fin);
    if (fmt == kFmtCsr) return apply_csr<float>(op, op.A->val32.p, x, epi, R, s, fin);
    // FP32 values of a dictionary-format matrix are read as CSR, without paired-entry loads
    Operator csr = op;
    csr.load &= ~kLoadPair;
    return apply_csr<float>(csr, op.A->val32.p, x, epi, R, s, fin);
  }
  if (delta) return apply_delta<double>(op, op.A->val.p, x, epi, R, s, fin);
  if (fmt == kFmtDict && ensure_dict(*op.A)) return apply_dict(op, x, epi, R, s, fin);
  return apply_csr<double>(op, op.A->val.p, x, epi, R, s, fin);
}

// -------------------------------------------------------------- vector kernels
__global__ __launch_bounds__(kBlock) void k_norm2(int64_t n, const double *__restrict__ x, double *partial) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += x[i] * x[i];
  block_partial(s, partial);
}

struct FinStore {
  double *out;
  __device__ bool skip() const { return false; }
  __device__ void finish(double t) const { *out = t; }
};

// K_b of a Lanczos step (sml/LANCZOS_SLOW_SYM.c:183-186 and :116 of the next step):
// w = y - alpha_k v_k - beta_k v_{k-1}; nrm2 = <w,w>; beta_{k+1} = sqrt(nrm2)
// With pre set, alpha_k is the sum of the product's pre_n block partials, taken in every
// block exactly as its finish kernel would (block 0 stores alpha[k]).
__global__ __launch_bounds__(kBlock) void k_lanczos_b(int64_t n, int k, const double *__restrict__ y,
                                                      const double *__restrict__ vk,
                                                      const double *__restrict__ vkm1,
                                                      double *alpha, const double *beta,
                                                      double *__restrict__ w, double *partial,
                                                      const double *pre, int pre_n) {
  // four elements per pass with their loads issued together; sums in index order.  The
  // first pass's operands are loaded before alpha is known, so that their latency overlaps
  // the partial sum.
  const int64_t S = (int64_t)gridDim.x * kBlock;
  const int64_t first = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  double fy[4], fk[4], fm[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int64_t i = first + u * S;
    const bool in = i < n;
    fy[u] = in ? y[i] : 0.0;
    fk[u] = in ? vk[i] : 0.0;
    fm[u] = (in && k > 0) ? vkm1[i] : 0.0;
  }
  double a;
  if (pre) {
    __shared__ double bc;
    const double t = finish_sum(pre, pre_n);
    if (threadIdx.x == 0) {
      bc = t;
      if (blockIdx.x == 0) alpha[k] = t;
    }
    __syncthreads();
    a = bc;
  } else {
    a = alpha[k];
  }
  const double b = (k > 0) ? beta[k] : 0.0;
  double s = 0.0;
  for (int64_t i0 = first; i0 < n; i0 += 4 * S) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * S;
      if (i < n) {
        const bool pf = i0 == first;
        v[u] = (pf ? fy[u] : y[i]) - a * (pf ? fk[u] : vk[i]);
        if (k > 0) v[u] -= b * (pf ? fm[u] : vkm1[i]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * S;
      if (i < n) {
        w[i] = v[u];
        s += v[u] * v[u];
      }
    }
  }
  block_partial(s, partial);
}

struct FinLanczosB {
  double *nrm2, *beta;
  int k;
  __device__ bool skip() const { return false; }
  __device__ void finish(double t) const {
    *nrm2 = t;
    beta[k + 1] = sqrt(t);
  }
};

// Ritz vector sum_j c_j v_j over the stored Lanczos basis (sml/LANCZOS_SYM.c:195-202)
__global__ __launch_bounds__(kBlock) void k_ritz(int64_t n, int k, const double *const *vecs,
                                                 const double *coef, double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    double t = 0.0;
    for (int j = 0; j <= k; j++) t += coef[j] * vecs[j][i];
    out[i] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_scale_rnorm(int64_t n, const double *nrm2,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ y) {
  const double sc = 1.0 / sqrt(*nrm2);
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = sc * x[i];
}

__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const double *__restrict__ x,
                                                 double *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = x[i];
}

// CG update (CONJUGATE_GRADIENT_SYM.c:216-236): alpha = rr/pAp, x += alpha p, r -= alpha y,
// rr_new = <r,r>; the last block records convergence in status[0..1].
// With pre set, pAp is the sum of the product's pre_n block partials, taken in every block
// exactly as its finish kernel would (block 0 stores pap[it]).
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t n, int it, double acc,
                                                      const double *__restrict__ p,
                                                      const double *__restrict__ y,
                                                      double *__restrict__ x, double *__restrict__ r,
                                                      double *pap, const double *rr, const int *status,
                                                      double *partial, const double *pre, int pre_n,
                                                      CgEarly early) {
  if (status[0] != 0) return;
  // the first pass's operands are loaded before pAp is known (see k_lanczos_b)
  const int64_t S = (int64_t)gridDim.x * kBlock;
  const int64_t first = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  double fx[4], fp[4], fr[4], fy[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int64_t i = first + u * S;
    const bool in = i < n;
    fx[u] = in ? x[i] : 0.0;
    fp[u] = in ? p[i] : 0.0;
    fr[u] = in ? r[i] : 0.0;
    fy[u] = in ? y[i] : 0.0;
  }
  double pAp;
  if (pre) {
    __shared__ double bc;
    const double t = finish_sum(pre, pre_n);
    if (threadIdx.x == 0) {
      bc = t;
      if (blockIdx.x == 0) pap[it] = t;
    }
    __syncthreads();
    pAp = bc;
  } else {
    pAp = pap[it];
  }
  const double alpha = rr[it - 1] / pAp;
  double s = 0.0, e1 = 0.0, e2 = 0.0;
  for (int64_t i0 = first; i0 < n; i0 += 4 * S) {
    double rv[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * S;
      if (i < n) {
        const bool pf = i0 == first;
        xv[u] = (pf ? fx[u] : x[i]) + alpha * (pf ? fp[u] : p[i]);
        x[i] = xv[u];
        rv[u] = (pf ? fr[u] : r[i]) - alpha * (pf ? fy[u] : y[i]);
        r[i] = rv[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + u * S < n) s += rv[u] * rv[u];
    if (early.b) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t i = i0 + u * S;
        if (i < n) {
          e1 += fabs(early.b[i] - rv[u] - early.diag_add * xv[u]);
          e2 += xv[u] * xv[u];
        }
      }
    }
  }
  block_partial(s, partial);
  if (early.b) {   // uniform per launch; block_sum's LDS is reused, hence the barriers
    __syncthreads();
    block_partial(e1, early.partial);
    __syncthreads();
    block_partial_to(e2, early.partial + gridDim.x + blockIdx.x);
  }
}

struct FinCgUpdate {
  double *rr;
  int *status;
  int it;
  double acc;
  __device__ bool skip() const { return status[0] != 0; }
  __device__ void finish(double t) const {
    rr[it] = t;
    status[1] = it;
    if (t < acc) status[0] = 1;
  }
};

// CG direction (CONJUGATE_GRADIENT_SYM.c:257-262): p = r + (rr_it / rr_{it-1}) p.  With pre
// set, rr_it is the sum of k_cg_update's pre_n block partials, taken in every block exactly
// as its finish kernel would; block 0 stores rr[it] and the status (FinCgUpdate's work), and
// a converged iteration leaves p alone.
__global__ __launch_bounds__(kBlock) void k_cg_dir(int64_t n, int it, const double *__restrict__ r,
                                                   double *rr, double *__restrict__ p, int *status,
                                                   const double *pre, int pre_n, double acc, CgEarly early) {
  if (pre) {
    // block 0 of this very kernel may set status[0]: read it once per block, so that all
    // threads of a block take the same branch (the reduction below has barriers)
    __shared__ int done0;
    __shared__ double bc;
    if (threadIdx.x == 0) done0 = status[0];
    __syncthreads();
    if (done0 != 0) return;
    // this thread's first element, loaded before the partial sum
    const int64_t i1 = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    const double r1 = i1 < n ? r[i1] : 0.0, p1 = i1 < n ? p[i1] : 0.0;
    const double t = finish_sum(pre, pre_n);
    // the iterate test (CgEarly), block 0 only (a uniform branch: every thread of the block
    // takes part in the two sums)
    bool early_done = false;
    if (early.b && blockIdx.x == 0) {
      __syncthreads();   // block_sum's LDS is reused
      const double e1 = finish_sum(early.partial, pre_n);
      __syncthreads();
      const double e2 = finish_sum(early.partial + pre_n, pre_n);
      early_done = e2 > 0.0 && e1 < early.acc * sqrt(e2);   // valid in thread 0
    }
    if (threadIdx.x == 0) {
      bc = t;
      if (blockIdx.x == 0) {
        rr[it] = t;
        status[1] = it;
        if (t < acc || early_done) status[0] = 1;
      }
    }
    __syncthreads();
    if (bc < acc) return;
    const double beta = bc / rr[it - 1];
    if (i1 < n) p[i1] = r1 + beta * p1;
    for (int64_t i = i1 + (int64_t)gridDim.x * kBlock; i < n; i += (int64_t)gridDim.x * kBlock)
      p[i] = r[i] + beta * p[i];
    return;
  }
  if (status[0] != 0) return;
  const double beta = rr[it] / rr[it - 1];
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = r[i] + beta * p[i];
}

// ------------------------------------------------- lower triangle -> full CSR
// bad[0] != 0 if a row is empty, its last entry is not the diagonal, or a column is out
// of [0, row).  tcnt[c] counts the strictly-lower entries in column c.
// One wavefront per row (coalesced column reads); four rows per workgroup.
__global__ __launch_bounds__(kBlock) void k_tri_count(int32_t n, const int64_t *__restrict__ trow,
                                                      const int32_t *__restrict__ tcol,
                                                      int64_t *__restrict__ own,
                                                      unsigned *__restrict__ tcnt, int *bad) {
  const int32_t r = blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= n) return;
  const int64_t b = trow[r], e = trow[r + 1];
  if (lane == 0) own[r] = e - b;
  if (e <= b || tcol[e - 1] != r) {
    if (lane == 0) atomicOr(bad, 1);
    return;
  }
  for (int64_t j = b + lane; j < e - 1; j += 64) {
    const int32_t c = tcol[j];
    if (c < 0 || c >= r) {
      atomicOr(bad, 2);
      continue;
    }
    atomicAdd(&tcnt[c], 1u);
  }
}

__global__ __launch_bounds__(kBlock) void k_row_len(int32_t n, const int64_t *trow,
                                                    const unsigned *tcnt, int64_t *len) {
  const int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r < n) len[r] = (trow[r + 1] - trow[r]) + tcnt[r];
}

// exclusive scan of len[0..n) into out[0..n], 1024 elements per block
constexpr int kScanPer = 4;
constexpr int kScanChunk = kBlock * kScanPer;

__device__ int64_t block_excl_scan(int64_t v, int64_t *total) {
  __shared__ int64_t lds[kBlock];
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int64_t t = (threadIdx.x >= (unsigned)off) ? lds[threadIdx.x - off] : 0;
    __syncthreads();
    lds[threadIdx.x] += t;
    __syncthreads();
  }
  const int64_t incl = lds[threadIdx.x];
  *total = lds[kBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_blocks(int32_t n, const int64_t *len, int64_t *bsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int64_t s = 0;
  for (int q = 0; q < kScanPer; q++)
    if (base + q < n) s += len[base + q];
  int64_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(int32_t nb, int64_t *bsum) {
  // single workgroup: in-place exclusive scan of nb block sums
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int32_t base = 0; base < nb; base += kBlock) {
    const int32_t i = base + threadIdx.x;
    const int64_t v = (i < nb) ? bsum[i] : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan(v, &tot);
    if (i < nb) bsum[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(int32_t n, const int64_t *len, const int64_t *bsum,
                                                       int64_t *out) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int64_t v[kScanPer];
  int64_t s = 0;
  for (int q = 0; q < kScanPer; q++) {
    v[q] = (base + q < n) ? len[base + q] : 0;
    s += v[q];
  }
  int64_t tot;
  int64_t ex = block_excl_scan(s, &tot) + bsum[blockIdx.x];
  for (int q = 0; q < kScanPer; q++) {
    if (base + q < n) out[base + q] = ex;
    ex += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = ex;
}

// A transposed entry on its way to the tail of its new row: column and value in one
// 16-byte record, so that the scattered append touches one cache line, not two.
struct alignas(16) TailRec {
  int32_t col;
  int32_t pad;
  double val;
};

// Start of row c's tail in the record array: the strictly-lower entries of the columns
// before c (frow counts own + tail entries per row, trow only the own ones).
__device__ __forceinline__ int64_t tail_base(const int64_t *frow, const int64_t *trow, int32_t c) {
  return frow[c] - (trow[c] - trow[0]);
}

// One wavefront per row: copy the row's own (lower) entries to the full CSR, and append
// each of them, transposed, to the tail records of row c (position by an atomic counter
// per row; k_sort_tail puts the tails in column order).
__global__ __launch_bounds__(64) void k_tri_fill(int32_t n, const int64_t *__restrict__ trow,
                                                 const int32_t *__restrict__ tcol,
                                                 const double *__restrict__ tval,
                                                 const int64_t *__restrict__ frow,
                                                 unsigned *__restrict__ tfill,
                                                 int32_t *__restrict__ fcol,
                                                 double *__restrict__ fval,
                                                 TailRec *__restrict__ tail) {
  const int32_t r = blockIdx.x;
  if (r >= n) return;
  const int64_t b = trow[r], e = trow[r + 1];
  const int64_t dst = frow[r];
  for (int64_t j = b + threadIdx.x; j < e; j += 64) {
    fcol[dst + (j - b)] = tcol[j];
    fval[dst + (j - b)] = tval[j];
  }
  for (int64_t j = b + threadIdx.x; j < e - 1; j += 64) {
    const int32_t c = tcol[j];
    const int64_t pos = tail_base(frow, trow, c) + atomicAdd(&tfill[c], 1u);
    TailRec t;
    t.col = r;
    t.pad = 0;
    t.val = tval[j];
    tail[pos] = t;
  }
}

// One wavefront per row: sort the row's tail records by column (the columns are
// distinct) and write them after the row's own entries.  Tails up to kTailCap entries are
// bitonic-sorted in LDS; longer ones by one lane.
constexpr int kTailCap = 512;
__global__ __launch_bounds__(64) void k_sort_tail(int32_t n, const int64_t *__restrict__ trow,
                                                  const int64_t *__restrict__ frow,
                                                  const TailRec *__restrict__ tail,
                                                  int32_t *__restrict__ fcol,
                                                  double *__restrict__ fval) {
  __shared__ int32_t kc[kTailCap];
  __shared__ double kv[kTailCap];
  const int32_t r = blockIdx.x;
  if (r >= n) return;
  const int lane = threadIdx.x;
  const int64_t own = trow[r + 1] - trow[r];
  const int64_t b = frow[r] + own, e = frow[r + 1];
  const int m = (int)(e - b);
  const TailRec *src = tail + tail_base(frow, trow, r);
  if (m <= 0) return;
  if (m > kTailCap) {
    if (lane != 0) return;
    for (int64_t i = 0; i < m; i++) {
      const int32_t c0 = src[i].col;
      const double v0 = src[i].val;
      int64_t j = b + i - 1;
      while (j >= b && fcol[j] > c0) {
        fcol[j + 1] = fcol[j];
        fval[j + 1] = fval[j];
        j--;
      }
      fcol[j + 1] = c0;
      fval[j + 1] = v0;
    }
    return;
  }
  int P = 1;
  while (P < m) P <<= 1;
  for (int i = lane; i < P; i += 64) {
    if (i < m) {
      const TailRec t = src[i];
      kc[i] = t.col;
      kv[i] = t.val;
    } else {
      kc[i] = INT32_MAX;
      kv[i] = 0.0;
    }
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int jx = i ^ stride;
        if (jx > i) {
          const bool up = (i & size) == 0;
          const int32_t x = kc[i], y = kc[jx];
          if ((x > y) == up) {
            kc[i] = y;
            kc[jx] = x;
            const double t = kv[i];
            kv[i] = kv[jx];
            kv[jx] = t;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = lane; i < m; i += 64) {
    fcol[b + i] = kc[i];
    fval[b + i] = kv[i];
  }
}

__global__ __launch_bounds__(kBlock) void k_to_fp32(int64_t nnz, const double *__restrict__ v,
                                                    float *__restrict__ f) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * kBlock)
    f[i] = (float)v[i];
}

// Start vector: splitmix64 of a counter, integers in [-5e7, 5e7) as the
// rand()%100000000 - 50000000 of sml/LANCZOS_SLOW_SYM.c:82 (seeded, reproducible).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_start_vector(int64_t n, uint64_t seed, double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint64_t h = splitmix64(seed * 0xD1B54A32D192ED03ULL + (uint64_t)i);
    out[i] = (double)(int64_t)(h % 100000000ULL) - 50000000.0;
  }
}

inline int vec_grid(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (int)(b < kVecGrid ? (b > 0 ? b : 1) : kVecGrid);
}

}  // namespace

// ------------------------------------------------------------------ launchers
void Reduce::ensure(size_t blocks) {
  if (partial.count < blocks) partial.resize(blocks);
}

int pick_lanes(int32_t n, int64_t nnz) {
  const double avg = n > 0 ? (double)nnz / n : 0.0;
  int g = 4;
  while (g < 64 && g * 3 < avg) g <<= 1;
  return g;
}

void default_schedule(dmrgx_matrix &A) {
  A.lanes = pick_lanes(A.n, A.nnz);
  A.grid = kGridCap;
  // eight loads in flight per lane; the value dictionary only where the tuner finds it
  // faster (its table lookups cost more than the 4 bytes per entry they save on MI355X for
  // the superblocks measured)
  A.load = kLoadDeep;
}

int spmv_blocks(int32_t n, int lanes, int grid_cap) {
  const int rpb = kBlock / lanes;
  const int64_t groups = ((int64_t)n + rpb - 1) / rpb;
  return (int)(grid_cap > 0 && groups > grid_cap ? grid_cap : groups);
}

void launch_spmv(const Operator &op, const double *x, double *y, hipStream_t s) {
  apply_op(op, false, x, EpiPlain{y}, nullptr, s);
}

void launch_spmv_dot(const Operator &op, const double *x, double *y, double *dot, Reduce &R, hipStream_t s) {
  apply_op(op, false, x, EpiStoreDot{y, x, dot}, &R, s);
}

void launch_slot_apply(int32_t n, const SlotRows &rows, const double *x, double *y, const double *dot_with,
                       double *dot_out, Reduce &R, hipStream_t s) {
  apply_rows(n, 8, kGridCap, rows, x, EpiStoreDot{y, dot_with, dot_out}, &R, s);
}

void launch_ed_apply(int32_t n, const EdRows &rows, const double *x, double *y, const double *dot_with,
                     double *dot_out, Reduce &R, hipStream_t s) {
  // an on-site row has at most d entries: one lane per row
  apply_rows(n, 1, kGridCap, rows, x, EpiStoreDot{y, dot_with, dot_out}, &R, s);
}

void launch_norm2(int64_t n, const double *x, Reduce &R, double *out, hipStream_t s) {
  const int g = vec_grid(n);
  R.ensure(g);
  k_norm2<<<g, kBlock, 0, s>>>(n, x, R.partial.p);
  k_finish<FinStore><<<1, kBlock, 0, s>>>(FinStore{out}, R.partial.p, g);
  HIPCHK(hipGetLastError());
}

void launch_lanczos_a(const Operator &op, bool fp32, int k, const double *w, const double *nrm2, double shift,
                      double *vk, double *y, double *alpha, Reduce &R, hipStream_t s) {
  apply_op(op, fp32, w, EpiLanczosA{w, const_cast<double *>(nrm2), vk, y, alpha, k, shift, 0.0}, &R, s);
}

void launch_lanczos_b(int64_t n, int k, const double *y, const double *vk, const double *vkm1,
                      const double *alpha, double *beta, double *w, double *nrm2, Reduce &R,
                      hipStream_t s) {
  const int g = reduce_grid(n);
  R.ensure(g);
  k_lanczos_b<<<g, kBlock, 0, s>>>(n, k, y, vk, vkm1 ? vkm1 : vk, const_cast<double *>(alpha), beta, w,
                                   R.partial.p, nullptr, 0);
  k_finish<FinLanczosB><<<1, kBlock, 0, s>>>(FinLanczosB{nrm2, beta, k}, R.partial.p, g);
  HIPCHK(hipGetLastError());
}

int launch_lanczos_a_part(const Operator &op, bool fp32, int k, const double *w, double *nrm2, double *beta,
                          const double *pre, int pre_n, double *vk, double *y, double *alpha, Reduce &R,
                          hipStream_t s) {
  EpiLanczosA e{w, nrm2, vk, y, alpha, k, 0.0, 0.0};
  e.pre = pre;
  e.pre_n = pre_n;
  e.beta = beta;
  return apply_op(op, fp32, w, e, &R, s, false);
}

int launch_lanczos_b_part(int64_t n, int k, const double *y, const double *vk, const double *vkm1, double *alpha,
                          const double *beta, const double *pre, int pre_n, double *w, Reduce &R, hipStream_t s) {
  const int g = reduce_grid(n);
  R.ensure(g);
  k_lanczos_b<<<g, kBlock, 0, s>>>(n, k, y, vk, vkm1 ? vkm1 : vk, alpha, beta, w, R.partial.p, pre, pre_n);
  HIPCHK(hipGetLastError());
  return g;
}

void launch_lanczos_b_finish(int k, double *beta, double *nrm2, const Reduce &R, int nblk, hipStream_t s) {
  k_finish<FinLanczosB><<<1, kBlock, 0, s>>>(FinLanczosB{nrm2, beta, k}, R.partial.p, nblk);
  HIPCHK(hipGetLastError());
}

void launch_ritz(int64_t n, int k, const double *const *vecs_dev, const double *coef_dev, double *out,
                 hipStream_t s) {
  k_ritz<<<vec_grid(n), kBlock, 0, s>>>(n, k, vecs_dev, coef_dev, out);
  HIPCHK(hipGetLastError());
}

void launch_scale_by_rnorm(int64_t n, const double *nrm2, const double *x, double *y, hipStream_t s) {
  k_scale_rnorm<<<vec_grid(n), kBlock, 0, s>>>(n, nrm2, x, y);
  HIPCHK(hipGetLastError());
}

void launch_copy(int64_t n, const double *x, double *y, hipStream_t s) {
  k_copy<<<vec_grid(n), kBlock, 0, s>>>(n, x, y);
  HIPCHK(hipGetLastError());
}

void launch_rayleigh(const Operator &op, const double *v, double *out, Reduce &R, hipStream_t s) {
  apply_op(op, false, v, EpiStoreDot{nullptr, v, out}, &R, s);
}

void launch_residual_l1(const Operator &op, const double *v, double eig, double *out, Reduce &R,
                        hipStream_t s) {
  apply_op(op, false, v, EpiResidualL1{v, out, eig}, &R, s);
}

int residual_l1_default_blocks(int32_t n, int64_t nnz) { return spmv_blocks(n, pick_lanes(n, nnz), kGridCap); }

void launch_residual_l1_into(int32_t n, int64_t nnz, const int64_t *row, const int32_t *col, const double *val,
                             const double *v, double eig, double *out, double *partial, hipStream_t s) {
  // default_schedule's choices for a CSR matrix of these dimensions (lanes from the mean row
  // length, kGridCap workgroups, eight loads in flight, CSR), so that the result equals
  // launch_residual_l1 on a dmrgx_matrix of the same arrays bit for bit
  apply_rows_into(n, pick_lanes(n, nnz), kGridCap, CsrRows<double, kLoadDeep>{row, col, val, nnz}, v,
                  EpiResidualL1{v, out, eig}, partial, s, 0, true, false);
}

void launch_cg_init(const Operator &op, double shift, const double *nrm2, const double *x_in,
                    double *x_out, const double *b, double *r, double *p, double *rr0, Reduce &R,
                    hipStream_t s) {
  apply_op(op, false, x_in, EpiCgInit{nrm2, x_in, b, x_out, r, p, rr0, shift, 0.0}, &R, s);
}

void launch_cg_ap(const Operator &op, double shift, const double *p, double *y, double *pap,
                  const int *done, Reduce &R, hipStream_t s) {
  apply_op(op, false, p, EpiCgAp{p, y, pap, done, shift}, &R, s);
}

void launch_cg_update(int64_t n, int it, double acc, const double *p, const double *y, double *x,
                      double *r, const double *pap, double *rr, int *status, Reduce &R,
                      hipStream_t s) {
  const int g = reduce_grid(n);
  R.ensure(g);
  k_cg_update<<<g, kBlock, 0, s>>>(n, it, acc, p, y, x, r, const_cast<double *>(pap), rr, status, R.partial.p,
                                   nullptr, 0, CgEarly{});
  k_finish<FinCgUpdate><<<1, kBlock, 0, s>>>(FinCgUpdate{rr, status, it, acc}, R.partial.p, g);
  HIPCHK(hipGetLastError());
}

void launch_cg_dir(int64_t n, int it, const double *r, const double *rr, double *p, const int *status,
                   hipStream_t s) {
  k_cg_dir<<<vec_grid(n), kBlock, 0, s>>>(n, it, r, const_cast<double *>(rr), p, const_cast<int *>(status),
                                          nullptr, 0, 0.0, CgEarly{});
  HIPCHK(hipGetLastError());
}

int launch_cg_ap_part(const Operator &op, double shift, const double *p, double *y, const int *done, Reduce &R,
                      hipStream_t s) {
  return apply_op(op, false, p, EpiCgAp{p, y, nullptr, done, shift}, &R, s, false);
}

int launch_cg_update_part(int64_t n, int it, const double *p, const double *y, double *x, double *r, double *pap,
                          const double *rr, const int *status, const double *pre, int pre_n, Reduce &R,
                          hipStream_t s, const CgEarly &early) {
  const int g = reduce_grid(n);
  R.ensure(g);
  k_cg_update<<<g, kBlock, 0, s>>>(n, it, 0.0, p, y, x, r, pap, rr, status, R.partial.p, pre, pre_n, early);
  HIPCHK(hipGetLastError());
  return g;
}

void launch_cg_dir_part(int64_t n, int it, double acc, const double *r, double *rr, double *p, int *status,
                        const double *pre, int pre_n, hipStream_t s, const CgEarly &early) {
  k_cg_dir<<<vec_grid(n), kBlock, 0, s>>>(n, it, r, rr, p, status, pre, pre_n, acc, early);
  HIPCHK(hipGetLastError());
}

// ------------------------------------------------------------- minimum residual
// State of sml/MINIMUM_RESIDUAL.c:40-45 kept on the device: c0 c1 s0 s1 eta; status[0] = done,
// status[1] = iteration that converged.  Per iteration k, coef[5k..5k+4] =
// {1/alpha1, alpha3, alpha2, c2*eta, commit} for the vector update.
__global__ void k_mr_init(const double *nrm2, double *beta, double *mr, int *status) {
  const double g = sqrt(*nrm2);   // gamma0 = eta = |b - A x0| (:39-41)
  beta[0] = g;
  mr[0] = 1.0;
  mr[1] = 1.0;
  mr[2] = 0.0;
  mr[3] = 0.0;
  mr[4] = g;
  status[0] = 0;
  status[1] = -1;
}

// Givens step of :63-74 and the convergence test of :78
__global__ void k_mr_scalars(int k, double acc, const double *alpha, const double *beta, double *mr, int *status,
                             double *coef) {
  double *cf = coef + 5 * (size_t)k;
  if (status[0]) {
    cf[4] = -1.0;  // iteration not run
    return;
  }
  const double c0 = mr[0], c1 = mr[1], s0 = mr[2], s1 = mr[3], eta = mr[4];
  const double delta = alpha[k], gamma0 = beta[k], gamma1 = beta[k + 1];
  const double a0 = c1 * delta - c0 * s1 * gamma0;
  const double a1 = sqrt(a0 * a0 + gamma1 * gamma1);
  const double a2 = s1 * delta + c0 * c1 * gamma0;
  const double a3 = s0 * gamma0;
  const double c2 = a0 / a1, s2 = gamma1 / a1;
  cf[0] = 1.0 / a1;
  cf[1] = a3;
  cf[2] = a2;
  cf[3] = c2 * eta;
  const double eta2 = -s2 * eta;
  const bool done = fabs(eta2) < acc;
  // on convergence the reference returns before copying X1 into Out_Vec (:79-99)
  cf[4] = done ? 0.0 : 1.0;
  if (done) {
    status[0] = 1;
    status[1] = k;
  }
  mr[0] = c1;
  mr[1] = c2;
  mr[2] = s1;
  mr[3] = s2;
  mr[4] = eta2;
}

// W2 = (V1 - alpha3 W0 - alpha2 W1) / alpha1;  x += c2 eta W2 unless converged (:75-77, :100-106)
__global__ __launch_bounds__(kBlock) void k_mr_update(int64_t n, int k, const double *coef,
                                                      const double *__restrict__ v1, const double *__restrict__ w0,
                                                      const double *__restrict__ w1, double *__restrict__ w2,
                                                      double *__restrict__ x) {
  const double *cf = coef + 5 * (size_t)k;
  const double run = cf[4];
  if (run < 0.0) return;
  const double t = cf[0], a3 = cf[1], a2 = cf[2], ce = cf[3];
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double wv = t * (v1[i] - a3 * w0[i] - a2 * w1[i]);
    w2[i] = wv;
    if (run > 0.0) x[i] = x[i] + ce * wv;
  }
}

void launch_mr_init(const Operator &op, double shift, const double *b, const double *x, double *v, double *nrm2,
                    double *beta, double *mr, int *status, Reduce &R, hipStream_t s) {
  apply_op(op, false, x, EpiMrInit{b, x, v, nrm2, shift}, &R, s);
  k_mr_init<<<1, 1, 0, s>>>(nrm2, beta, mr, status);
  HIPCHK(hipGetLastError());
}

void launch_mr_step(int64_t n, int k, double acc, const double *alpha, const double *beta, double *mr, int *status,
                    double *coef, const double *v1, const double *w0, const double *w1, double *w2, double *x,
                    hipStream_t s) {
  k_mr_scalars<<<1, 1, 0, s>>>(k, acc, alpha, beta, mr, status, coef);
  k_mr_update<<<vec_grid(n), kBlock, 0, s>>>(n, k, coef, v1, w0, w1, w2, x);
  HIPCHK(hipGetLastError());
}

void launch_exclusive_scan(int32_t n, const int64_t *len, int64_t *out, hipStream_t s) {
  const int sb = (n + kScanChunk - 1) / kScanChunk;
  if (sb == 0) {
    HIPCHK(hipMemsetAsync(out, 0, sizeof(int64_t), s));
    return;
  }
  DevBuf<int64_t> bsum(sb);
  k_scan_blocks<<<sb, kBlock, 0, s>>>(n, len, bsum.p);
  k_scan_top<<<1, kBlock, 0, s>>>(sb, bsum.p);
  k_scan_apply<<<sb, kBlock, 0, s>>>(n, len, bsum.p, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
}

void launch_start_vector(int64_t n, uint64_t seed, double *out, hipStream_t s) {
  k_start_vector<<<vec_grid(n), kBlock, 0, s>>>(n, seed, out);
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------- value dictionary
// Open-addressing hash set of the values' bit patterns (linear probing, at most kProbe
// probes; a full neighbourhood sets *overflow and the build is retried with a larger
// table).  Inserts are vector atomicCAS on distinct slots; a slot that already holds the
// key is found with a plain load.
constexpr uint64_t kEmptyKey = ~0ull;   // a NaN pattern: NaN values make the build fail
constexpr int kProbe = 64;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_fill_u64(int64_t n, uint64_t v, uint64_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) out[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_dict_insert(int64_t nnz, const double *__restrict__ val,
                                                        unsigned long long *keys, uint64_t mask, int *overflow) {
  for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * kBlock) {
    const uint64_t key = (uint64_t)__double_as_longlong(val[j]);
    if (key == kEmptyKey) {
      atomicOr(overflow, 2);   // a value with the empty pattern: no dictionary
      continue;
    }
    uint64_t h = mix64(key) & mask;
    bool placed = false;
    for (int p = 0; p < kProbe && !placed; p++) {
      unsigned long long cur = __atomic_load_n(keys + h, __ATOMIC_RELAXED);
      if (cur == kEmptyKey) cur = atomicCAS(keys + h, (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (cur == kEmptyKey || cur == key) placed = true;
      else h = (h + 1) & mask;
    }
    if (!placed) atomicOr(overflow, 1);   // neighbourhood full: retry with a larger table
  }
}

__global__ __launch_bounds__(kBlock) void k_dict_occupied(int64_t cap, const unsigned long long *keys, int64_t *occ) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < cap; i += (int64_t)gridDim.x * kBlock)
    occ[i] = keys[i] != kEmptyKey ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_dict_table(int64_t cap, const unsigned long long *keys,
                                                       const int64_t *pos, double *table) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < cap; i += (int64_t)gridDim.x * kBlock)
    if (keys[i] != kEmptyKey) table[pos[i]] = __longlong_as_double((long long)keys[i]);
}

__global__ __launch_bounds__(kBlock) void k_dict_pack(int64_t nnz, const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const unsigned long long *__restrict__ keys, uint64_t mask,
                                                      const int64_t *__restrict__ pos, uint64_t *__restrict__ ent,
                                                      int *overflow) {
  for (int64_t j = blockIdx.x * (int64_t)kBlock + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * kBlock) {
    const uint64_t key = (uint64_t)__double_as_longlong(val[j]);
    uint64_t h = mix64(key) & mask;
    int p = 0;
    while (p < kProbe && keys[h] != key) {
      h = (h + 1) & mask;
      p++;
    }
    if (p == kProbe) {
      *overflow = 1;
      ent[j] = 0;
      continue;
    }
    ent[j] = (uint64_t)(uint32_t)col[j] | ((uint64_t)pos[h] << 32);
  }
}

// Returns the number of distinct values, or 0 if a table of up to max_cap slots could not
// hold them at half load (the matrix then keeps the plain form).
int64_t build_value_dict(dmrgx_matrix &A, hipStream_t s) {
  const int64_t nnz = A.nnz;
  if (nnz == 0) return 0;
  constexpr int64_t kMaxCap = int64_t(1) << 24;  // 128 MiB of keys at most
  int64_t cap = int64_t(1) << 16;
  while (cap < kMaxCap && cap < 2 * std::min<int64_t>(nnz, int64_t(1) << 20)) cap <<= 1;
  DevBuf<int> flag(1);
  for (;; cap <<= 1) {
    if (cap > kMaxCap) return 0;
    DevBuf<unsigned long long> keys(cap);
    k_fill_u64<<<vec_grid(cap), kBlock, 0, s>>>(cap, kEmptyKey, reinterpret_cast<uint64_t *>(keys.p));
    HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    k_dict_insert<<<vec_grid(nnz), kBlock, 0, s>>>(nnz, A.val.p, keys.p, (uint64_t)(cap - 1), flag.p);
    HIPCHK(hipGetLastError());
    int ovf = 0;
    HIPCHK(hipMemcpyAsync(&ovf, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ovf & 2) return 0;
    if (ovf) continue;
    DevBuf<int64_t> occ(cap), pos(cap + 1);
    k_dict_occupied<<<vec_grid(cap), kBlock, 0, s>>>(cap, keys.p, occ.p);
    HIPCHK(hipGetLastError());
    launch_exclusive_scan((int32_t)cap, occ.p, pos.p, s);
    int64_t nv = 0;
    HIPCHK(hipMemcpyAsync(&nv, pos.p + cap, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (2 * nv > cap) continue;            // keep probe runs short: at most half full
    if (nv > nnz / 4) return 0;            // too many distinct values to be worth a table
    A.table.resize(nv);
    A.ent.resize(nnz);
    k_dict_table<<<vec_grid(cap), kBlock, 0, s>>>(cap, keys.p, pos.p, A.table.p);
    k_dict_pack<<<vec_grid(nnz), kBlock, 0, s>>>(nnz, A.col.p, A.val.p, keys.p, (uint64_t)(cap - 1), pos.p,
                                                 A.ent.p, flag.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&ovf, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ovf) throw Failure(DMRGX_ERR_HIP, "value dictionary lookup failed after a complete build");
    return nv;
  }
}

// ---------------------------------------------------------------- delta form
// One wavefront per row: first column, then the differences of consecutive columns.
__global__ __launch_bounds__(kBlock) void k_delta_build(int32_t n, const int64_t *__restrict__ row,
                                                        const int32_t *__restrict__ col, int32_t *__restrict__ first,
                                                        uint16_t *__restrict__ dcol, int *too_wide) {
  const int32_t r = blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= n) return;
  const int64_t b = row[r], e = row[r + 1];
  if (lane == 0) first[r] = b < e ? col[b] : 0;
  for (int64_t j = b + lane; j < e; j += 64) {
    const int32_t d = j == b ? 0 : col[j] - col[j - 1];
    if (d < 0 || d > 65535) *too_wide = 1;
    dcol[j] = (uint16_t)(d < 0 || d > 65535 ? 0 : d);
  }
}

bool build_delta(dmrgx_matrix &A, hipStream_t s) {
  if (A.n == 0) return false;
  A.first.resize(A.n);
  A.dcol.resize(std::max<int64_t>(A.nnz, 1));
  DevBuf<int> flag(1);
  HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
  const int rows_per_block = kBlock / 64;
  k_delta_build<<<(A.n + rows_per_block - 1) / rows_per_block, kBlock, 0, s>>>(A.n, A.row.p, A.col.p, A.first.p,
                                                                               A.dcol.p, flag.p);
  HIPCHK(hipGetLastError());
  int wide = 0;
  HIPCHK(hipMemcpyAsync(&wide, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return wide == 0;
}

// ---------------------------------------------------------------- sliced form
// Row lengths, and per slice of C = 64/P rows the slots it takes: 64 per pass, passes =
// ceil(longest row / P).
template <int P>
__global__ __launch_bounds__(kBlock) void k_sell_width(int32_t n, int32_t nsl, const int64_t *__restrict__ row,
                                                       int32_t *__restrict__ slen, int64_t *__restrict__ slots) {
  constexpr int C = 64 / P;
  const int32_t sl = blockIdx.x * kBlock + threadIdx.x;
  if (sl >= nsl) return;
  int64_t mx = 0;
  for (int i = 0; i < C; i++) {
    const int32_t r = sl * C + i;
    if (r >= n) break;
    const int64_t len = row[r + 1] - row[r];
    slen[r] = (int32_t)len;
    mx = len > mx ? len : mx;
  }
  slots[sl] = (mx + P - 1) / P * 64;
}

// One wavefront (one block) per slice: its rows' entries are copied through LDS in tiles of
// kFillTile slot columns, so that the reads take each row's run of entries in one piece and
// the writes fill whole 64-slot columns.  Slot (k, i P + l) of a slice holds entry k P + l of
// its row i; a slot past the row's end holds the row's last column (0 for an empty row or
// a row past n) and the value 0.  scol may be null (values only).
constexpr int kFillTile = 16;
template <int P, class VT>
__global__ __launch_bounds__(64) void k_sell_fill(int32_t n, int32_t nsl, const int64_t *__restrict__ row,
                                                  const int32_t *__restrict__ col, const VT *__restrict__ val,
                                                  const int64_t *__restrict__ sptr, int32_t *__restrict__ scol,
                                                  VT *__restrict__ sval) {
  constexpr int C = 64 / P, E = kFillTile * P;   // rows per slice, a row's entries per tile
  __shared__ VT tv[kFillTile][64];
  __shared__ int32_t tc[kFillTile][64];
  const int lane = threadIdx.x;
  const int32_t sl = blockIdx.x;
  const int64_t b0 = sptr[sl];
  const int32_t w = (int32_t)((sptr[sl + 1] - b0) >> 6);
  // lane i < C: row i's first entry, length and last column
  int64_t rb = 0;
  int32_t rl = 0, rlast = 0;
  const int32_t r = sl * C + lane;
  if (lane < C && r < n) {
    rb = row[r];
    rl = (int32_t)(row[r + 1] - rb);
    rlast = rl > 0 ? col[rb + rl - 1] : 0;
  }
  for (int32_t k0 = 0; k0 < w; k0 += kFillTile) {
#pragma unroll 4
    for (int u = 0; u < kFillTile; u++) {   // C rows x E entries = 64 x kFillTile elements
      const int e = u * 64 + lane;
      const int i = e / E, jj = e % E;
      const int64_t ib = __shfl(rb, i, 64);
      const int32_t il = __shfl(rl, i, 64), ilast = __shfl(rlast, i, 64);
      const int32_t j = k0 * P + jj;
      VT v = VT(0);
      int32_t c = ilast;
      if (j < il) {
        v = val[ib + j];
        c = col[ib + j];
      }
      tv[jj / P][i * P + jj % P] = v;
      tc[jj / P][i * P + jj % P] = c;
    }
    __syncthreads();
    const int kn = min(kFillTile, w - k0);
    for (int k = 0; k < kn; k++) {
      const int64_t at = b0 + (int64_t)(k0 + k) * 64 + lane;
      sval[at] = tv[k][lane];
      if (scol) scol[at] = tc[k][lane];
    }
    __syncthreads();
  }
}

// 16-bit columns of the sliced form: one thread per sequence (slot column q = slice * 64 + j
// of a slice), walking its slots in order.  Count pass: the sequence's first column and how
// many of its differences do not fit in 0..65534; fill pass (eptr set): the differences and
// the exceptions.
__global__ __launch_bounds__(kBlock) void k_sell16_walk(int32_t nsl, const int64_t *__restrict__ sptr,
                                                        const int32_t *__restrict__ scol, int32_t *__restrict__ first,
                                                        int64_t *__restrict__ count, const int64_t *__restrict__ eptr,
                                                        uint16_t *__restrict__ d16, int32_t *__restrict__ exc) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= (int64_t)nsl * 64) return;
  const int64_t sl = q / 64, j = q % 64;
  const int64_t b0 = sptr[sl];
  const int64_t w = (sptr[sl + 1] - b0) / 64;
  int32_t prev = w > 0 ? scol[b0 + j] : 0;
  int64_t e = eptr ? eptr[q] : 0, cnt = 0;
  if (!eptr) first[q] = prev;
  else if (w > 0) d16[b0 + j] = 0;
  for (int64_t k = 1; k < w; k++) {
    const int64_t at = b0 + k * 64 + j;
    const int32_t c = scol[at];
    const int64_t d = (int64_t)c - prev;
    const bool wide = d < 0 || d > 65534;
    if (!eptr) {
      cnt += wide ? 1 : 0;
    } else if (wide) {
      d16[at] = 0xFFFF;
      exc[e++] = c;
    } else {
      d16[at] = (uint16_t)d;
    }
    prev = c;
  }
  if (!eptr) count[q] = cnt;
}

int64_t build_sell16(dmrgx_matrix &A, int P, hipStream_t s) {
  const int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  const int C = 64 / P;
  const int32_t nsl = (A.n + C - 1) / C;
  const int64_t nseq = (int64_t)nsl * 64;
  if (!A.sell_ptr[lp].p) throw Failure(DMRGX_ERR_ARG, "16-bit sliced columns need the sliced form first");
  if (nseq > INT32_MAX) throw Failure(DMRGX_ERR_ARG, "too many slices for the 16-bit sliced columns");
  A.sell_first[lp].resize((size_t)std::max<int64_t>(nseq, 1));
  A.sell_eptr[lp].resize((size_t)nseq + 1);
  A.sell_d16[lp].resize(A.sell_col[lp].count);
  DevBuf<int64_t> count((size_t)std::max<int64_t>(nseq, 1));
  const unsigned g = (unsigned)((nseq + kBlock - 1) / kBlock);
  if (nseq > 0) {
    k_sell16_walk<<<g, kBlock, 0, s>>>(nsl, A.sell_ptr[lp].p, A.sell_col[lp].p, A.sell_first[lp].p, count.p,
                                       nullptr, nullptr, nullptr);
    HIPCHK(hipGetLastError());
  }
  launch_exclusive_scan((int32_t)nseq, count.p, A.sell_eptr[lp].p, s);
  int64_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, A.sell_eptr[lp].p + nseq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // one entry past the last exception: the lane sums load an exception one ahead
  A.sell_exc[lp].resize((size_t)total + 1);
  HIPCHK(hipMemsetAsync(A.sell_exc[lp].p + total, 0, sizeof(int32_t), s));
  if (nseq > 0) {
    k_sell16_walk<<<g, kBlock, 0, s>>>(nsl, A.sell_ptr[lp].p, A.sell_col[lp].p, nullptr, nullptr,
                                       A.sell_eptr[lp].p, A.sell_d16[lp].p, A.sell_exc[lp].p);
    HIPCHK(hipGetLastError());
  }
  A.sell_nexc[lp] = total;
  return total;
}

// Format 4's value indices: one thread per row position of every slice (as k_sell_fill),
// the index half of each entry's dictionary word at its slot, 0 at padding slots.
template <int P>
__global__ __launch_bounds__(kBlock) void k_sellx_fill(int32_t n, int32_t nsl, const int64_t *__restrict__ row,
                                                       const uint64_t *__restrict__ ent,
                                                       const int64_t *__restrict__ sptr, uint16_t *__restrict__ sidx) {
  constexpr int C = 64 / P;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)nsl * C) return;
  const int32_t sl = (int32_t)(t / C), i = (int32_t)(t % C), r = (int32_t)t;
  const int64_t b0 = sptr[sl], slots = (sptr[sl + 1] - b0) / 64 * P;
  const int64_t b = r < n ? row[r] : 0, len = r < n ? row[r + 1] - b : 0;
  for (int64_t j = 0; j < slots; j++) {
    const int64_t at = b0 + (j / P) * 64 + (int64_t)i * P + j % P;
    sidx[at] = j < len ? (uint16_t)(ent[b + j] >> 32) : (uint16_t)0;
  }
}

__global__ __launch_bounds__(kBlock) void k_to_fp32_table(int64_t n, const double *__restrict__ a,
                                                          float *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    b[i] = (float)a[i];
}

void build_sellx(dmrgx_matrix &A, int P, bool fp32, hipStream_t s) {
  const int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  const int C = 64 / P;
  const int32_t nsl = (A.n + C - 1) / C;
  if (!A.sell_ptr[lp].p || A.n_values <= 0 || A.n_values > kSellIdxMaxValues || !A.ent.p)
    throw Failure(DMRGX_ERR_ARG, "sliced value indices need the sliced form and a table of at most 65536 values");
  if (!A.sell_idx16[lp].p) {
    A.sell_idx16[lp].resize(A.sell_col[lp].count);
    const unsigned gf = (unsigned)(((int64_t)nsl * C + kBlock - 1) / kBlock);
    if (nsl > 0) {
      switch (P) {
        case 1: k_sellx_fill<1><<<gf, kBlock, 0, s>>>(A.n, nsl, A.row.p, A.ent.p, A.sell_ptr[lp].p, A.sell_idx16[lp].p); break;
        case 2: k_sellx_fill<2><<<gf, kBlock, 0, s>>>(A.n, nsl, A.row.p, A.ent.p, A.sell_ptr[lp].p, A.sell_idx16[lp].p); break;
        default: k_sellx_fill<4><<<gf, kBlock, 0, s>>>(A.n, nsl, A.row.p, A.ent.p, A.sell_ptr[lp].p, A.sell_idx16[lp].p); break;
      }
      HIPCHK(hipGetLastError());
    }
  }
  if (fp32 && !A.table32.p) {
    A.table32.resize((size_t)A.n_values);
    k_to_fp32_table<<<vec_grid(A.n_values), kBlock, 0, s>>>(A.n_values, A.table.p, A.table32.p);
    HIPCHK(hipGetLastError());
  }
}

void build_sell(dmrgx_matrix &A, int P, bool fp32, hipStream_t s) {
  const int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  const int C = 64 / P;
  const int32_t nsl = (A.n + C - 1) / C;
  if (!A.sell_ptr[lp].p) {
    if (!A.sell_len.p) A.sell_len.resize(std::max(A.n, 1));
    DevBuf<int64_t> slots(std::max(nsl, 1));
    A.sell_ptr[lp].resize((size_t)nsl + 1);
    const unsigned gw = (unsigned)((nsl + kBlock - 1) / kBlock);
    if (nsl > 0) {
      switch (P) {
        case 1: k_sell_width<1><<<gw, kBlock, 0, s>>>(A.n, nsl, A.row.p, A.sell_len.p, slots.p); break;
        case 2: k_sell_width<2><<<gw, kBlock, 0, s>>>(A.n, nsl, A.row.p, A.sell_len.p, slots.p); break;
        default: k_sell_width<4><<<gw, kBlock, 0, s>>>(A.n, nsl, A.row.p, A.sell_len.p, slots.p); break;
      }
      HIPCHK(hipGetLastError());
    }
    launch_exclusive_scan(nsl, slots.p, A.sell_ptr[lp].p, s);
    int64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, A.sell_ptr[lp].p + nsl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    A.sell_col[lp].resize((size_t)std::max<int64_t>(total, 1));
    A.sell_val[lp].resize((size_t)std::max<int64_t>(total, 1));
    if (nsl > 0) {
      switch (P) {
        case 1: k_sell_fill<1, double><<<nsl, 64, 0, s>>>(A.n, nsl, A.row.p, A.col.p, A.val.p, A.sell_ptr[lp].p,
                                                         A.sell_col[lp].p, A.sell_val[lp].p); break;
        case 2: k_sell_fill<2, double><<<nsl, 64, 0, s>>>(A.n, nsl, A.row.p, A.col.p, A.val.p, A.sell_ptr[lp].p,
                                                         A.sell_col[lp].p, A.sell_val[lp].p); break;
        default: k_sell_fill<4, double><<<nsl, 64, 0, s>>>(A.n, nsl, A.row.p, A.col.p, A.val.p, A.sell_ptr[lp].p,
                                                          A.sell_col[lp].p, A.sell_val[lp].p); break;
      }
      HIPCHK(hipGetLastError());
    }
  }
  if (fp32 && !A.sell_val32[lp].p) {
    A.sell_val32[lp].resize(A.sell_val[lp].count);
    if (nsl > 0) {
      // the columns are in place already: only the FP32 values are written
      switch (P) {
        case 1: k_sell_fill<1, float><<<nsl, 64, 0, s>>>(A.n, nsl, A.row.p, A.col.p, A.val32.p, A.sell_ptr[lp].p,
                                                        nullptr, A.sell_val32[lp].p); break;
        case 2: k_sell_fill<2, float><<<nsl, 64, 0, s>>>(A.n, nsl, A.row.p, A.col.p, A.val32.p, A.sell_ptr[lp].p,
                                                        nullptr, A.sell_val32[lp].p); break;
        default: k_sell_fill<4, float><<<nsl, 64, 0, s>>>(A.n, nsl, A.row.p, A.col.p, A.val32.p, A.sell_ptr[lp].p,
                                                         nullptr, A.sell_val32[lp].p); break;
      }
      HIPCHK(hipGetLastError());
    }
  }
  HIPCHK(hipStreamSynchronize(s));
}

void launch_to_fp32(const dmrgx_matrix &A, hipStream_t s) {
  k_to_fp32<<<vec_grid(A.nnz), kBlock, 0, s>>>(A.nnz, A.val.p, A.val32.p);
  HIPCHK(hipGetLastError());
}

void launch_expand_lower(int32_t n, int64_t nnz_tri, const int64_t *trow, const int32_t *tcol,
                         const double *tval, dmrgx_matrix &A, hipStream_t s, const unsigned *tcnt_given,
                         const int *bad_given) {
  const int nb = (n + kBlock - 1) / kBlock;
  DevBuf<int64_t> own, len(n);
  DevBuf<unsigned> tcnt, tfill(n);
  DevBuf<int> bad;
  HIPCHK(hipMemsetAsync(tfill.p, 0, sizeof(unsigned) * n, s));
  const unsigned *cnt = tcnt_given;
  const int *badp = bad_given;
  if (!cnt) {
    own.resize(n);
    tcnt.resize(n);
    bad.resize(1);
    HIPCHK(hipMemsetAsync(tcnt.p, 0, sizeof(unsigned) * n, s));
    HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    k_tri_count<<<(n + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, s>>>(n, trow, tcol, own.p, tcnt.p, bad.p);
    HIPCHK(hipGetLastError());
    cnt = tcnt.p;
    badp = bad.p;
  }
  // row lengths and pointers before the structure check is read: one round trip for both
  k_row_len<<<nb, kBlock, 0, s>>>(n, trow, cnt, len.p);
  HIPCHK(hipGetLastError());
  A.row.resize((size_t)n + 1);
  const int sb = (n + kScanChunk - 1) / kScanChunk;
  DevBuf<int64_t> bsum(sb > 0 ? sb : 1);
  k_scan_blocks<<<sb, kBlock, 0, s>>>(n, len.p, bsum.p);
  k_scan_top<<<1, kBlock, 0, s>>>(sb, bsum.p);
  k_scan_apply<<<sb, kBlock, 0, s>>>(n, len.p, bsum.p, A.row.p);
  HIPCHK(hipGetLastError());
  int64_t *h = reinterpret_cast<int64_t *>(ctx().host_stage(2));
  HIPCHK(hipMemcpyAsync(h, A.row.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(h + 1, badp, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int hbad = *reinterpret_cast<const int *>(h + 1);
  if (hbad & 1) throw Failure(DMRGX_ERR_NODIAG, "a row has no diagonal element as its last entry");
  if (hbad & 2) throw Failure(DMRGX_ERR_ARG, "lower-triangle CRS has a column outside [0,row)");
  const int64_t nnz = h[0];
  A.n = n;
  A.nnz = nnz;
  A.col.resize(nnz);
  A.val.resize(nnz);
  // the transposed (strictly lower) entries: all but the triangle's own (trow[0] == 0)
  DevBuf<TailRec> tail(std::max<int64_t>(nnz - nnz_tri, 1));
  k_tri_fill<<<n, 64, 0, s>>>(n, trow, tcol, tval, A.row.p, tfill.p, A.col.p, A.val.p, tail.p);
  k_sort_tail<<<n, 64, 0, s>>>(n, trow, A.row.p, tail.p, A.col.p, A.val.p);
  HIPCHK(hipGetLastError());
  // no wait: the matrix's users are queued behind the fill on this stream (and the scratch
  // released here is only reused in stream order)
}

}  // namespace dmrgx
