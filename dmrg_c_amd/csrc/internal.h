// Internal declarations of libdmrg_c_amd.so (not installed, not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <unordered_map>
#include <string>
#include <vector>

#include "dmrg_c_amd.h"

namespace dmrgx {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kVecGrid = 1024;       // fixed grid of the streaming vector kernels
// Blocks of the vector kernels whose block partials the next kernel sums (Lanczos update,
// CG update): about four elements per thread, at most kVecGrid blocks.  Single and batched
// solves use the same count, so their partials (and scalars) are the same.
inline int reduce_grid(int64_t n) {
  const int64_t b = (n + 4 * kBlock - 1) / (4 * kBlock);
  return (int)(b < 1 ? 1 : (b > kVecGrid ? kVecGrid : b));
}
constexpr int kGridCap = 4096;       // default cap on SpMV workgroups (row groups round-robin)

// Error carried from the HIP layer up to the C entry points.
struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);
[[noreturn]] void hip_fail(hipError_t e, const char *what, const char *file, int line);

#define HIPCHK(expr)                                                   \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) ::dmrgx::hip_fail(e_, #expr, __FILE__, __LINE__); \
  } while (0)

// One process-wide context per device: stream, pinned staging, allocation cache.
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  uint64_t seed = 1;
  double *pinned = nullptr;        // host staging for scalars read back per chunk
  size_t pinned_cap = 0;
  std::vector<std::pair<size_t, void *>> free_list;  // cached device blocks
  size_t cached_bytes = 0;
  std::unordered_map<void *, size_t> live;            // block sizes of handed-out pointers

  void *alloc(size_t bytes);
  void release(void *p, size_t bytes);
  double *host_stage(size_t count);
  // Copies between device memory and caller (pageable) host memory through two pinned
  // chunk buffers: the DMA of one chunk overlaps the host copy of the other.  A plain
  // hipMemcpy from pageable memory runs at a few GB/s.  Both return after the data has
  // landed (copy_in: on the device, in stream order).
  void copy_out(void *dst, const void *dev_src, size_t bytes);
  void copy_in(void *dev_dst, const void *src, size_t bytes);
  // several host arrays to their device destinations through the same pinned chunks, as one
  // pipelined transfer (one wait at the end instead of one per array)
  struct CopySeg {
    void *dev_dst;
    const void *src;
    size_t bytes;
  };
  void copy_in_many(const std::vector<CopySeg> &segs);
  char *xfer_pin[2] = {nullptr, nullptr};
  hipEvent_t xfer_ev[2] = {nullptr, nullptr};
  // two events for the solvers' pipelined chunk read-backs
  hipEvent_t *step_events();
  hipEvent_t step_ev[2] = {nullptr, nullptr};
  // live kernel timing of the SpMV launches (dmrgx_profile_begin/end)
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
  size_t prof_used = 0;
  double prof_bytes = 0.0;
  int prof_every = 1;        // bracket one launch in every prof_every
  int64_t prof_seen = 0;     // launches seen since begin
  std::pair<hipEvent_t, hipEvent_t> *prof_slot();
  bool prof_take() { return prof && (prof_seen++ % prof_every) == 0; }
  void sizes_note(void *p, size_t bytes) { live[p] = bytes; }
  size_t sizes_take(void *p) {
    auto it = live.find(p);
    size_t b = it == live.end() ? 0 : it->second;
    if (it != live.end()) live.erase(it);
    return b;
  }
};

Context &ctx();          // lazily initialised on the current device
void ensure_init();

// RAII device buffer from the context cache.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t count = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), count(o.count) { o.p = nullptr; o.count = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { reset(); p = o.p; count = o.count; o.p = nullptr; o.count = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void resize(size_t n) {
    reset();
    if (n) { p = static_cast<T *>(ctx().alloc(n * sizeof(T))); count = n; }
  }
  void reset() {
    if (p) ctx().release(p, count * sizeof(T));
    p = nullptr;
    count = 0;
  }
};

}  // namespace dmrgx

// Full CSR resident on the device.  Columns ascending in every row.
struct dmrgx_matrix {
  int32_t n = 0;
  int64_t nnz = 0;
  dmrgx::DevBuf<int64_t> row;
  dmrgx::DevBuf<int32_t> col;
  dmrgx::DevBuf<double> val;
  dmrgx::DevBuf<float> val32;   // FP32 copy of val for the mixed-precision Lanczos
  // Value dictionary (built on first use): ent[j] = col[j] | idx << 32 with
  // table[idx] == val[j] bit for bit, so a product reads 8 bytes per entry instead of 12.
  dmrgx::DevBuf<uint64_t> ent;
  dmrgx::DevBuf<double> table;
  int64_t n_values = -1;        // distinct values; -1 not built yet, 0 not worth building
  // Delta form (built on first use): first[r] = first column of row r, dcol[j] = col[j] -
  // col[j-1] within a row (0 at a row start); only when every difference fits 16 bits.
  dmrgx::DevBuf<int32_t> first;
  dmrgx::DevBuf<uint16_t> dcol;
  int delta = -1;               // -1 not built yet, 0 not available, 1 built
  // Sliced form per lanes-per-row count P = 1, 2, 4 (index log2 P), built on first use:
  // slice offsets, padded columns and values (FP64, and FP32 once the mixed form asks);
  // sell_len[r] = entries of row r.
  dmrgx::DevBuf<int64_t> sell_ptr[3];
  dmrgx::DevBuf<int32_t> sell_col[3];
  dmrgx::DevBuf<double> sell_val[3];
  dmrgx::DevBuf<float> sell_val32[3];
  dmrgx::DevBuf<int32_t> sell_len;
  // 16-bit columns of the sliced form (per P, built on first use).  A sequence is one of
  // the 64 slot columns of a slice (the slots one lane of one row reads, in order).
  // sell_d16 holds at every slot the column's difference from the sequence's previous slot,
  // or 0xFFFF where that does not fit in 0..65534, in which case the column itself is the
  // sequence's next entry of sell_exc (from sell_eptr[sequence]); sell_first is each
  // sequence's first column.  sell_nexc: exceptions, -1 while not built.
  dmrgx::DevBuf<uint16_t> sell_d16[3];
  dmrgx::DevBuf<int32_t> sell_first[3];
  dmrgx::DevBuf<int64_t> sell_eptr[3];
  dmrgx::DevBuf<int32_t> sell_exc[3];
  int64_t sell_nexc[3] = {-1, -1, -1};
  // Value indices of the sliced form (format 4, per P): sell_idx16 = the slot's index into
  // the value dictionary's table (0 at padding slots); table32 = the table in FP32 for the
  // mixed form.  sell_idx_ok: -1 not tried, 0 not available (no table of at most 65536
  // values), 1 built.
  dmrgx::DevBuf<uint16_t> sell_idx16[3];
  dmrgx::DevBuf<float> table32;
  int sell_idx_ok[3] = {-1, -1, -1};
  int lanes = 8;                // lanes per row of the SpMV kernel
  int grid = dmrgx::kGridCap;   // cap on SpMV workgroups (0: one per row group)
  int load = 0;                 // load variant (kLoad* bits)
};

namespace dmrgx {

// Matrix-free superblock operator of dmrg/DMRG_MATRIX_VECTOR_PRODUCT_OBC.c:18-83: the row's
// contributions from M_LLLR, M_LRRL, M_LRRL_Sign and M_RRRL are one flattened list the
// lanes stride over.
struct ObcRows {
  int32_t dim_RR, d;
  const int16_t *LL, *LR, *RR, *RL;
  const int32_t *inv;
  const int32_t *inv_LLLR, *inv_LRRL, *inv_RRRL;
  const int16_t *LL_LLLR, *LR_LLLR, *LR_LRRL, *RL_LRRL, *RR_RRRL, *RL_RRRL;
  const int64_t *r1, *r2, *r3, *r4;
  const int32_t *c1, *c2, *c3, *c4;
  const double *v1, *v2, *v3, *v4;
  const int32_t *ele_RR;
  static constexpr bool kTimed = false;
  double alg_bytes(int32_t) const { return 0.0; }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    const int ll = LL[r], lr = LR[r], rr = RR[r], rl = RL[r];
    const int a = inv_LLLR[ll * d + lr], b = inv_LRRL[lr * d + rl], c = inv_RRRL[rr * d + rl];
    const int64_t s1 = r1[a], n1 = r1[a + 1] - s1;
    const int64_t s2 = r2[b], n2 = r2[b + 1] - s2;
    const int64_t s3 = r3[b], n3 = r3[b + 1] - s3;
    const int64_t s4 = r4[c], n4 = r4[c + 1] - s4;
    const double sign = (ele_RR[rr] % 2 == 1) ? -1.0 : 1.0;
    const int64_t base = (int64_t)ll * d * dim_RR * d + (int64_t)lr * dim_RR * d;
    double s = 0.0;
    for (int64_t t = lane; t < n1 + n2 + n3 + n4; t += G) {
      int64_t idx;
      double v;
      if (t < n1) {
        const int32_t cc = c1[s1 + t];
        idx = ((int64_t)LL_LLLR[cc] * d + LR_LLLR[cc]) * dim_RR * d + (int64_t)rr * d + rl;
        v = v1[s1 + t];
      } else if (t < n1 + n2) {
        const int64_t j = s2 + (t - n1);
        const int32_t cc = c2[j];
        idx = (int64_t)ll * d * dim_RR * d + (int64_t)LR_LRRL[cc] * dim_RR * d + (int64_t)rr * d + RL_LRRL[cc];
        v = v2[j];
      } else if (t < n1 + n2 + n3) {
        const int64_t j = s3 + (t - n1 - n2);
        const int32_t cc = c3[j];
        idx = (int64_t)ll * d * dim_RR * d + (int64_t)LR_LRRL[cc] * dim_RR * d + (int64_t)rr * d + RL_LRRL[cc];
        v = v3[j] * sign;
      } else {
        const int64_t j = s4 + (t - n1 - n2 - n3);
        const int32_t cc = c4[j];
        idx = base + (int64_t)RR_RRRL[cc] * d + RL_RRRL[cc];
        v = v4[j];
      }
      const int32_t iv = inv[idx];
      if (iv >= 0) s += x[iv] * v;
    }
    return s;
  }
};

// One operator acting on one slot (LL, LR, RR or RL) of the superblock state,
// dmrg/DMRG_V_M_LL_Q0.c:14-29 and its three siblings.
struct SlotRows {
  int32_t slot, dim_RR, d;
  const int16_t *LL, *LR, *RR, *RL;
  const int32_t *inv;
  const int64_t *row;
  const int32_t *col;
  const double *val;
  static constexpr bool kTimed = false;
  double alg_bytes(int32_t) const { return 0.0; }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    int q[4] = {LL[r], LR[r], RR[r], RL[r]};
    const int own = q[slot];
    double s = 0.0;
    for (int64_t j = row[own] + lane; j < row[own + 1]; j += G) {
      q[slot] = col[j];
      const int32_t iv = inv[(((int64_t)q[0] * d + q[1]) * dim_RR + q[2]) * d + q[3]];
      if (iv >= 0) s += x[iv] * val[j];
    }
    return s;
  }
};

// Load of the streamed matrix arrays: NT = true marks them non-temporal (evict first), so
// that x has the L2 to itself; which policy moves fewer bytes depends on the matrix, so the
// schedule autotuner picks it.
template <bool NT, class T>
__device__ __forceinline__ T stream_load(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// This lane's share of sum_j H_rj x_j over CSR row r: entries b+lane, b+lane+G, ... summed in
// that order.  Each pass issues the lane's next U (value, column) loads and their U x
// gathers together; positions past the row end load the row's last entry again and are
// masked out of the sum, so there is no branch between the loads.  The order of the
// additions is that of the plain loop whatever U and the load policy are, so the schedule
// never changes the result.
template <int G, int U, bool NT, class VT>
__device__ __forceinline__ double csr_lane_sum(const int64_t *__restrict__ row, const int32_t *__restrict__ col,
                                               const VT *__restrict__ val, int32_t r, int lane,
                                               const double *__restrict__ x) {
  double s = 0.0;
  const int64_t e = row[r + 1];
  const int64_t last = e - 1;
  for (int64_t j = row[r] + lane; j < e; j += U * G) {
    VT v[U];
    int32_t c[U];
    bool p[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      p[u] = j + u * G < e;
      const int64_t ju = p[u] ? j + u * G : last;
      v[u] = stream_load<NT>(val + ju);
      c[u] = stream_load<NT>(col + ju);
    }
    double xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < U; u++)
      if (p[u]) s += (double)v[u] * xv[u];
  }
  return s;
}

// Variants of the CSR product.  Bit 0 (CSR only): two entries per lane per load -- a
// 16-byte value pair and an 8-byte column pair -- which changes the summation order within
// a row; bit 1: eight entries in flight per lane instead of four (never changes a result).
// Bit 2: row groups in contiguous ranges per block, blocks mapped so that each XCD takes
// a contiguous share of the rows (with a capped grid this changes the order of the fused
// dot products).
constexpr int kLoadPair = 1;   // CSR: two entries per lane per load (changes the order)
constexpr int kLoadDeep = 2;
constexpr int kLoadXcd = 4;
// Bit 6, sliced form with 16-bit columns only: matrix loads one group of eight passes
// ahead of the x gathers (sell16_lane_sum_pipe); never changes a result.
constexpr int kLoadPipe = 64;
// Bits 3-5: storage format of the product.  CSR: 4-byte columns; dictionary (FP64 products
// only): 8-byte (column, value index) entries over a table of the distinct values; delta:
// 2-byte column differences within each row plus each row's first column.  Every format
// gives the CSR products bit for bit (same values, same order).
// Format 3, sliced: see sell_lane_sum; its lanes per row (1, 2 or 4) set the slice height,
// and its sums are in an order of their own.  For it bit 0 selects the 16-bit column form
// (sell16_lane_sum) and bit 1 non-temporal matrix loads; neither changes a result.
// Format 4, sliced with value indices: the sliced form's 16-bit columns and a 16-bit index
// per slot into the value dictionary's table (sellx_lane_sum; 4 bytes per entry), for a
// matrix with at most 65536 distinct values (ED Hamiltonians have a few dozen); bit 0 is
// ignored, bit 1 selects non-temporal loads.  Same values and order as format 3, so the
// same bits; a matrix without such a table reads format 3 with 16-bit columns.
constexpr int kFmtShift = 3;
constexpr int kFmtCsr = 0, kFmtDict = 1, kFmtDelta = 2, kFmtSell = 3, kFmtSellIdx = 4;
constexpr int kLoadDict = kFmtDict << kFmtShift;
constexpr int kLoadDelta = kFmtDelta << kFmtShift;
constexpr int kLoadSell = kFmtSell << kFmtShift;
constexpr int kLoadSellIdx = kFmtSellIdx << kFmtShift;
constexpr int kLoadFmtMask = 7 << kFmtShift;
constexpr int kLoadVariants = 5 << kFmtShift;   // 0 .. 39
constexpr int kSellMaxLanes = 4;
constexpr int64_t kSellIdxMaxValues = 65536;
__host__ __device__ inline int load_format(int load) { return (load & kLoadFmtMask) >> kFmtShift; }
// a load variant the products accept: bits 0-5 below kLoadVariants, bit 6 with format 3
__host__ __device__ inline bool valid_load(int load) {
  const int base = load & ~kLoadPipe;
  return load >= 0 && base < kLoadVariants && (!(load & kLoadPipe) || load_format(base) == kFmtSell);
}
__host__ __device__ inline bool sliced_format(int fmt) { return fmt == kFmtSell || fmt == kFmtSellIdx; }

// Row groups [start, end) with stride step that logical block lb of nblk takes out of
// ngroups: round-robin (lb, lb + nblk, ...) or one contiguous range.
struct GroupRange {
  int32_t start, end, step;
};
__device__ __forceinline__ GroupRange group_range(int32_t lb, int32_t nblk, int32_t ngroups, bool contiguous) {
  if (!contiguous) return GroupRange{lb, ngroups, nblk};
  const int32_t per = ngroups / nblk, rem = ngroups % nblk;
  const int32_t s = lb * per + min(lb, rem);
  return GroupRange{s, s + per + (lb < rem ? 1 : 0), 1};
}

// Logical id of block i of nblk when the hardware deals blocks round-robin over the 8 XCDs
// (block i runs on XCD i % 8): XCD j gets the contiguous ids [j q + min(j, r), ...), q =
// nblk / 8, r = nblk % 8, so with contiguous group ranges each XCD's L2 serves one band of
// rows and the x columns near it.
constexpr int kXcds = 8;
__device__ __forceinline__ int32_t xcd_logical(int32_t i, int32_t nblk) {
  const int32_t q = nblk / kXcds, r = nblk % kXcds, j = i % kXcds, t = i / kXcds;
  return j * q + min(j, r) + t;
}

// The same sum over the dictionary form: one 8-byte load per entry (column | value index
// << 32), the value from the table (a small array that stays in cache), additions in the
// same order, so the result equals csr_lane_sum's bit for bit.
template <int G, int U, bool NT>
__device__ __forceinline__ double dict_lane_sum(const int64_t *__restrict__ row, const uint64_t *__restrict__ ent,
                                                const double *__restrict__ table, int32_t r, int lane,
                                                const double *__restrict__ x) {
  double s = 0.0;
  const int64_t e = row[r + 1];
  const int64_t last = e - 1;
  for (int64_t j = row[r] + lane; j < e; j += U * G) {
    uint64_t q[U];
    bool p[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      p[u] = j + u * G < e;
      q[u] = stream_load<NT>(ent + (p[u] ? j + u * G : last));
    }
    double v[U], xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      v[u] = table[q[u] >> 32];
      xv[u] = x[(uint32_t)q[u]];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (p[u]) s += v[u] * xv[u];
  }
  return s;
}

// The same sum over the delta form: per pass each lane loads its U 2-byte column differences
// and values, the G lanes of the row turn the differences into columns with a prefix sum
// (carried from pass to pass from the row's first column), then gather x.  All G lanes of
// a row run the same passes (the shuffles need them); the additions are in the order of
// csr_lane_sum, so the result equals it bit for bit.
template <int G, int U, bool NT, class VT>
__device__ __forceinline__ double delta_lane_sum(const int64_t *__restrict__ row, const int32_t *__restrict__ first,
                                                 const uint16_t *__restrict__ dcol, const VT *__restrict__ val,
                                                 int32_t r, int lane, const double *__restrict__ x) {
  double s = 0.0;
  const int64_t b = row[r], e = row[r + 1];
  if (b >= e) return s;
  const int64_t last = e - 1;
  int32_t carry = first[r];
  for (int64_t j0 = b; j0 < e; j0 += U * G) {
    int32_t c[U];
    VT v[U];
    bool p[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t j = j0 + u * G + lane;
      p[u] = j < e;
      const int64_t jj = p[u] ? j : last;
      c[u] = p[u] ? (int32_t)stream_load<NT>(dcol + jj) : 0;
      v[u] = stream_load<NT>(val + jj);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
      for (int off = 1; off < G; off <<= 1) {
        const int32_t t = __shfl_up(c[u], off, G);
        if (lane >= off) c[u] += t;
      }
      c[u] += carry;
      carry = __shfl(c[u], G - 1, G);
    }
    double xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < U; u++)
      if (p[u]) s += (double)v[u] * xv[u];
  }
  return s;
}

template <int G, int V, class VT>
__device__ __forceinline__ double delta_lane_sum_v(const int64_t *__restrict__ row, const int32_t *__restrict__ first,
                                                   const uint16_t *__restrict__ dcol, const VT *__restrict__ val,
                                                   int32_t r, int lane, const double *__restrict__ x) {
  return delta_lane_sum<G, (V & kLoadDeep) ? 8 : 4, false>(row, first, dcol, val, r, lane, x);  // V < 4
}

template <int G, int V>
__device__ __forceinline__ double dict_lane_sum_v(const int64_t *__restrict__ row, const uint64_t *__restrict__ ent,
                                                  const double *__restrict__ table, int32_t r, int lane,
                                                  const double *__restrict__ x) {
  return dict_lane_sum<G, (V & kLoadDeep) ? 8 : 4, false>(row, ent, table, r, lane, x);  // V < 4
}

// CSR sum with two entries per lane per load: pair q = entries (2q, 2q+1), loaded as one
// 16-byte value pair and one 8-byte column pair (half the memory instructions of
// csr_lane_sum for the matrix arrays).  Lane l takes the row's full pairs q_lo + l,
// q_lo + l + G, ...; lane 0 adds the unpaired first entry (odd row start) before its pairs
// and the unpaired last entry (odd row end) after them.  A fixed order, different from
// csr_lane_sum's.
template <class VT> struct PairOf;
template <> struct PairOf<double> { using type = double2; };
template <> struct PairOf<float> { using type = float2; };

template <int G, int U, class VT>
__device__ __forceinline__ double csr_pair_lane_sum(const int64_t *__restrict__ row, const int32_t *__restrict__ col,
                                                    const VT *__restrict__ val, int32_t r, int lane,
                                                    const double *__restrict__ x) {
  using V2 = typename PairOf<VT>::type;
  double s = 0.0;
  const int64_t b = row[r], e = row[r + 1];
  if (b >= e) return s;
  const int64_t q_lo = (b + 1) >> 1, q_hi = e >> 1;   // full pairs [q_lo, q_hi)
  if ((b & 1) && lane == 0) s += (double)val[b] * x[col[b]];
  const V2 *__restrict__ v2 = reinterpret_cast<const V2 *>(val);
  const int2 *__restrict__ c2 = reinterpret_cast<const int2 *>(col);
  for (int64_t q0 = q_lo + lane; q0 < q_hi; q0 += U * G) {
    V2 v[U];
    int2 c[U];
    bool p[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      p[u] = q0 + u * G < q_hi;
      const int64_t q = p[u] ? q0 + u * G : q_hi - 1;
      v[u] = v2[q];
      c[u] = c2[q];
    }
    double x0[U], x1[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      x0[u] = x[c[u].x];
      x1[u] = x[c[u].y];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (p[u]) {
        s += (double)v[u].x * x0[u];
        s += (double)v[u].y * x1[u];
      }
  }
  if ((e & 1) && e - 1 >= 2 * q_lo && lane == 0) s += (double)val[e - 1] * x[col[e - 1]];
  return s;
}

template <int G, int V, class VT>
__device__ __forceinline__ double csr_lane_sum_v(const int64_t *__restrict__ row, const int32_t *__restrict__ col,
                                                 const VT *__restrict__ val, int32_t r, int lane,
                                                 const double *__restrict__ x) {
  if constexpr ((V & kLoadPair) != 0)
    return csr_pair_lane_sum<G, (V & kLoadDeep) ? 4 : 2, VT>(row, col, val, r, lane, x);  // V < 4
  else
    return csr_lane_sum<G, (V & kLoadDeep) ? 8 : 4, false>(row, col, val, r, lane, x);
}

// Sliced form (format 3): slices of C = 64/P consecutive rows, one wavefront per slice with P
// lanes per row.  Pass k of slice s holds entry k P + l of every row of the slice (lane l of
// the row) at sptr[s] + 64 k + (r mod C) P + l; a row shorter than the slice's widest is
// padded with its last column and the value 0, masked out of the sum by the row length.
// A wavefront's matrix loads are then contiguous 512-byte (values) and 256-byte (columns)
// segments, and one x gather instruction reads the columns of C neighbouring rows at the
// same position in their rows, which lie close together in superblock and ED matrices (the
// same terms acting on neighbouring basis states), instead of G consecutive columns of
// 64/G rows as the CSR forms do.  Lane l sums its entries in eight accumulators (pass k into
// accumulator k mod 8, eight passes' loads in flight), combined as
// ((a0 + a4) + (a2 + a6)) + ((a1 + a5) + (a3 + a7)); the P lanes of a row are then combined by
// the kernel's shuffles.  A fixed order, different from the CSR forms'.
// With NT the matrix loads are non-temporal (streamed past the caches), which leaves L2 and
// the Infinity Cache to x when the matrix is far larger than they are.
constexpr int kSellAcc = 8;
template <int P, class VT, bool NT = false>
__device__ __forceinline__ double sell_lane_sum(const int64_t *__restrict__ sptr, const int32_t *__restrict__ slen,
                                                const int32_t *__restrict__ scol, const VT *__restrict__ sval,
                                                int32_t r, int lane, const double *__restrict__ x) {
  constexpr int C = 64 / P;
  const int32_t s = r / C;
  const int64_t b0 = sptr[s];
  const int64_t base = b0 + (int64_t)(r % C) * P + lane;
  const int32_t w = (int32_t)((sptr[s + 1] - b0) >> 6);
  const int32_t len = slen[r];
  double acc[kSellAcc];
#pragma unroll
  for (int t = 0; t < kSellAcc; t++) acc[t] = 0.0;
  for (int32_t k = 0; k < w; k += kSellAcc) {
    VT v[kSellAcc];
    int32_t c[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int64_t at = base + (int64_t)min(k + t, w - 1) * 64;
      v[t] = stream_load<NT>(sval + at);
      c[t] = stream_load<NT>(scol + at);
    }
    double xv[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) xv[t] = x[c[t]];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++)
      if ((k + t) * P + lane < len) acc[t] += (double)v[t] * xv[t];
  }
  return ((acc[0] + acc[4]) + (acc[2] + acc[6])) + ((acc[1] + acc[5]) + (acc[3] + acc[7]));
}

// sell_lane_sum over the 16-bit column form: the same slots, values, order and masks, the
// columns rebuilt from the sequence's first column and the differences (an 0xFFFF slot
// takes the sequence's next exception, which is loaded one exception ahead so that its
// latency hides behind the slots in between).  10 bytes per FP64 entry instead of 12, and
// the same sums bit for bit.
template <int P, class VT, bool NT = false>
__device__ __forceinline__ double sell16_lane_sum(const int64_t *__restrict__ sptr, const int32_t *__restrict__ slen,
                                                  const uint16_t *__restrict__ sd16,
                                                  const int32_t *__restrict__ sfirst,
                                                  const int64_t *__restrict__ septr,
                                                  const int32_t *__restrict__ sexc, const VT *__restrict__ sval,
                                                  int32_t r, int lane, const double *__restrict__ x) {
  constexpr int C = 64 / P;
  const int32_t s = r / C;
  const int j = (r % C) * P + lane;
  const int64_t b0 = sptr[s];
  const int64_t base = b0 + j;
  const int32_t w = (int32_t)((sptr[s + 1] - b0) >> 6);
  const int32_t len = slen[r];
  const int64_t q = (int64_t)s * 64 + j;
  int32_t c = sfirst[q];
  int64_t e = septr[q];
  int32_t en = sexc[e];   // sell_exc holds one entry past the last exception
  double acc[kSellAcc];
#pragma unroll
  for (int t = 0; t < kSellAcc; t++) acc[t] = 0.0;
  for (int32_t k = 0; k < w; k += kSellAcc) {
    VT v[kSellAcc];
    uint16_t d[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int64_t at = base + (int64_t)min(k + t, w - 1) * 64;
      v[t] = stream_load<NT>(sval + at);
      d[t] = stream_load<NT>(sd16 + at);
    }
    int32_t cc[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int32_t kk = k + t;
      if (kk > 0 && kk < w) {
        if (d[t] == 0xFFFF) {
          c = en;
          en = sexc[++e];
        } else {
          c += d[t];
        }
      }
      cc[t] = c;
    }
    double xv[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) xv[t] = x[cc[t]];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++)
      if ((k + t) * P + lane < len) acc[t] += (double)v[t] * xv[t];
  }
  return ((acc[0] + acc[4]) + (acc[2] + acc[6])) + ((acc[1] + acc[5]) + (acc[3] + acc[7]));
}

// sell16_lane_sum with the matrix loads one pass group ahead: the values and column
// differences of slots k+8 .. k+15 are requested before the x gathers of slots k .. k+7,
// so that a wavefront's next matrix loads are in flight while it waits for its gathers.
// Same slots, values, columns, order and masks: the same sums bit for bit.  The group
// ahead is loaded only while there is one (w is uniform over a wavefront, which holds one
// slice), so every load stays inside the slice.
template <int P, class VT, bool NT = false>
__device__ __forceinline__ double sell16_lane_sum_pipe(const int64_t *__restrict__ sptr,
                                                       const int32_t *__restrict__ slen,
                                                       const uint16_t *__restrict__ sd16,
                                                       const int32_t *__restrict__ sfirst,
                                                       const int64_t *__restrict__ septr,
                                                       const int32_t *__restrict__ sexc,
                                                       const VT *__restrict__ sval, int32_t r, int lane,
                                                       const double *__restrict__ x) {
  constexpr int C = 64 / P;
  const int32_t s = r / C;
  const int j = (r % C) * P + lane;
  const int64_t b0 = sptr[s];
  const int64_t base = b0 + j;
  const int32_t w = (int32_t)((sptr[s + 1] - b0) >> 6);
  const int32_t len = slen[r];
  const int64_t q = (int64_t)s * 64 + j;
  int32_t c = sfirst[q];
  int64_t e = septr[q];
  int32_t en = sexc[e];
  double acc[kSellAcc];
#pragma unroll
  for (int t = 0; t < kSellAcc; t++) acc[t] = 0.0;
  VT v[kSellAcc];
  uint16_t d[kSellAcc];
  if (w > 0) {   // an empty slice has no slots to load
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int64_t at = base + (int64_t)min(t, w - 1) * 64;
      v[t] = stream_load<NT>(sval + at);
      d[t] = stream_load<NT>(sd16 + at);
    }
  }
  for (int32_t k = 0; k < w; k += kSellAcc) {
    int32_t cc[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int32_t kk = k + t;
      if (kk > 0 && kk < w) {
        if (d[t] == 0xFFFF) {
          c = en;
          en = sexc[++e];
        } else {
          c += d[t];
        }
      }
      cc[t] = c;
    }
    VT vn[kSellAcc];
    uint16_t dn[kSellAcc];
    const bool ahead = k + kSellAcc < w;
    if (ahead) {
#pragma unroll
      for (int t = 0; t < kSellAcc; t++) {
        const int64_t at = base + (int64_t)min(k + kSellAcc + t, w - 1) * 64;
        vn[t] = stream_load<NT>(sval + at);
        dn[t] = stream_load<NT>(sd16 + at);
      }
    }
    double xv[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) xv[t] = x[cc[t]];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++)
      if ((k + t) * P + lane < len) acc[t] += (double)v[t] * xv[t];
    if (ahead) {
#pragma unroll
      for (int t = 0; t < kSellAcc; t++) {
        v[t] = vn[t];
        d[t] = dn[t];
      }
    }
  }
  return ((acc[0] + acc[4]) + (acc[2] + acc[6])) + ((acc[1] + acc[5]) + (acc[3] + acc[7]));
}

// sell16_lane_sum with each slot's value read through its 16-bit index into the value
// table (format 4): 2 + 2 bytes per slot; table[idx] equals the slot's value bit for bit,
// so the sums are sell16_lane_sum's.  The table (a few dozen to 65536 values) stays in the
// caches.
template <int P, class VT, bool NT = false>
__device__ __forceinline__ double sellx_lane_sum(const int64_t *__restrict__ sptr, const int32_t *__restrict__ slen,
                                                 const uint16_t *__restrict__ sd16,
                                                 const int32_t *__restrict__ sfirst,
                                                 const int64_t *__restrict__ septr,
                                                 const int32_t *__restrict__ sexc,
                                                 const uint16_t *__restrict__ sidx, const VT *__restrict__ table,
                                                 int32_t r, int lane, const double *__restrict__ x) {
  constexpr int C = 64 / P;
  const int32_t s = r / C;
  const int j = (r % C) * P + lane;
  const int64_t b0 = sptr[s];
  const int64_t base = b0 + j;
  const int32_t w = (int32_t)((sptr[s + 1] - b0) >> 6);
  const int32_t len = slen[r];
  const int64_t q = (int64_t)s * 64 + j;
  int32_t c = sfirst[q];
  int64_t e = septr[q];
  int32_t en = sexc[e];
  double acc[kSellAcc];
#pragma unroll
  for (int t = 0; t < kSellAcc; t++) acc[t] = 0.0;
  for (int32_t k = 0; k < w; k += kSellAcc) {
    uint16_t ix[kSellAcc], d[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int64_t at = base + (int64_t)min(k + t, w - 1) * 64;
      ix[t] = stream_load<NT>(sidx + at);
      d[t] = stream_load<NT>(sd16 + at);
    }
    int32_t cc[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      const int32_t kk = k + t;
      if (kk > 0 && kk < w) {
        if (d[t] == 0xFFFF) {
          c = en;
          en = sexc[++e];
        } else {
          c += d[t];
        }
      }
      cc[t] = c;
    }
    VT v[kSellAcc];
    double xv[kSellAcc];
#pragma unroll
    for (int t = 0; t < kSellAcc; t++) {
      v[t] = table[ix[t]];
      xv[t] = x[cc[t]];
    }
#pragma unroll
    for (int t = 0; t < kSellAcc; t++)
      if ((k + t) * P + lane < len) acc[t] += (double)v[t] * xv[t];
  }
  return ((acc[0] + acc[4]) + (acc[2] + acc[6])) + ((acc[1] + acc[5]) + (acc[3] + acc[7]));
}

// Binary search of an ascending int64 array: index of target or -1
// (sml/BINARY_SEARCH_LINT1.c, iterative).
__device__ __forceinline__ int64_t find_code(const int64_t *__restrict__ a, int64_t n, int64_t target) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    const int64_t v = a[mid];
    if (v == target) return mid;
    if (v < target) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

// One on-site operator at one site of an ED state (exact/EXACT_V_M_Q0.c:20-35, Q1-Q3):
// row r of the output basis is code c; the operator row of c's local state a at `site`
// sends c to c - (a - col) d^site, looked up in the input basis; with sign_flag the
// result is negated when sites < site hold an odd number of electrons (EXACT_V_M_Q2.c:31-44).
struct EdRows {
  int32_t d, site, sign_flag;
  int64_t pw, n_in;
  const int64_t *basis_out, *basis_in;
  const int64_t *row;
  const int32_t *col;
  const double *val;
  const int32_t *n_ele;
  static constexpr bool kTimed = false;
  double alg_bytes(int32_t) const { return 0.0; }
  template <int G>
  __device__ __forceinline__ double lane_sum(int32_t r, int lane, const double *__restrict__ x) const {
    const int64_t code = basis_out[r];
    const int a = (int)((code / pw) % d);
    double s = 0.0;
    for (int64_t j = row[a] + lane; j < row[a + 1]; j += G) {
      const int64_t inv = find_code(basis_in, n_in, code - (int64_t)(a - col[j]) * pw);
      if (inv >= 0) s += x[inv] * val[j];
    }
    if (sign_flag) {
      int ne = 0;
      int64_t c = code;
      for (int k = 0; k < site; k++) {
        ne += n_ele[c % d];
        c /= d;
      }
      if (ne & 1) s = -s;
    }
    return s;
  }
};

// Device arrays of the matrix-free operator (uploaded from DMRG_BASIS + the three
// sub-Hamiltonians, obc.cpp).
struct ObcDev {
  int32_t n = 0, dim_LL = 0, dim_RR = 0, d = 0, lanes = 8;
  DevBuf<int16_t> LL, LR, RR, RL, LL_LLLR, LR_LLLR, LR_LRRL, RL_LRRL, RR_RRRL, RL_RRRL;
  DevBuf<int32_t> inv, inv_LLLR, inv_LRRL, inv_RRRL, ele_RR;
  DevBuf<int64_t> r1, r2, r3, r4;
  DevBuf<int32_t> c1, c2, c3, c4;
  DevBuf<double> v1, v2, v3, v4;
  ObcRows rows() const {
    return ObcRows{dim_RR, d, LL.p, LR.p, RR.p, RL.p, inv.p, inv_LLLR.p, inv_LRRL.p, inv_RRRL.p,
                   LL_LLLR.p, LR_LLLR.p, LR_LRRL.p, RL_LRRL.p, RR_RRRL.p, RL_RRRL.p,
                   r1.p, r2.p, r3.p, r4.p, c1.p, c2.p, c3.p, c4.p, v1.p, v2.p, v3.p, v4.p, ele_RR.p};
  }
};

// Eigenpairs of a truncation step's block density matrices up to kBlockEigMax (blockeig.hip),
// one workgroup per block, in two stages around the selection of the kept states.
// values(): block j (dimension dim[j]) row-major at a_off[j] in A (read only); launches the
// tridiagonal reduction and the eigenvalues and waits for them: block j's eigenvalues
// ascending at Wg.p + r_off[j].  vectors(): the eigenvectors of eigenvalues kept[j] (indices
// into block j's ascending list, ascending), as rows: X[x_off[j] + t * dim[j] ...] for the
// t-th of kept[j].  False when a vector came out dependent on the others (the caller solves
// another way).
constexpr int32_t kBlockEigMax = 192;   // packed lower triangle, n (n + 1) / 2 doubles, in LDS
struct BlockEig {
  int32_t nblk = 0;
  std::vector<int32_t> n;
  std::vector<int64_t> r_off, p_off;
  DevBuf<double> H, Dg, Eg, HVg, Wg;   // reflectors (packed), T's diagonal and subdiagonal, h, eigenvalues
  void values(int32_t count, const int32_t *dim, const int64_t *a_off, const double *A, hipStream_t s);
  bool vectors(const std::vector<std::vector<int32_t>> &kept, std::vector<double> &X, std::vector<int64_t> &x_off,
               hipStream_t s);
};

// Kronecker form of a superblock Hamiltonian (kron.hip): per sector of the superblock basis
// (rows with one LLLR quantum number: a dense da x db block of the vector, LLLR states by
// RRRL states) the LLLR-only terms as a dense da x da matrix M_L and the RRRL-only terms as
// a dense db x db matrix M_R, so that their part of H x is M_L X + X M_R^T per sector; the
// terms that couple the two halves are applied row by row from the term list.
struct KronSector {
  int64_t off;        // first superblock row
  int32_t da, db;     // LLLR states, RRRL states
  int64_t ml, mr;     // offsets of the sector's M_L and M_R in KronDev::ML / MR
  int32_t a0, b0;     // first of its LLLR / RRRL states in the flat state lists
};
// Tile shapes of the dense sector products and the product kernels over them: variant 0
// the FP64 vector form on 64 x 64 tiles, 1-4 the FP64 matrix cores on 64 x 64, 128 x 128,
// 64 x 128 and 128 x 64 tiles, 5 the matrix cores on 64 x 64 tiles with k-slabs of 32, 6 as 1
// held to four wavefronts per SIMD, 7 as 2 held to two.
constexpr int kKronShapes = 4;
inline constexpr int kKronTM[kKronShapes] = {64, 128, 64, 128};
inline constexpr int kKronTN[kKronShapes] = {64, 128, 128, 64};
constexpr int kKronVariants = 8;
constexpr int kKronDefault = 1;   // 64 x 64 tiles, FP64 matrix cores, default occupancy, no XCD mapping
inline constexpr int kKronShape[kKronVariants] = {0, 0, 1, 2, 3, 0, 0, 1};
struct KronDev {
  int32_t n = 0, nsec = 0, ncross = 0, na = 0, nb = 0;
  int32_t ntiles[kKronShapes] = {0, 0, 0, 0};
  int variant = kKronDefault;             // product kernel (kKronShape of it picks the tiles)
  int xcd = 0;                            // 1: consecutive tiles to the blocks of one XCD
  const dmrgx_superblock *sb = nullptr;   // borrowed: basis, inverse, operators, term list
  std::vector<KronSector> hsec;
  DevBuf<KronSector> sec;
  DevBuf<double> ML, MR;                  // dense sector blocks, row-major
  DevBuf<int32_t> tiles[kKronShapes];     // per shape: (sector, tile row, tile column, left/right)
  DevBuf<int32_t> cross;                  // indices of the coupling terms in the term list
  // per coupling term: its factor, and the LLLR / RRRL states each state reaches with the
  // term's operator on that half, with coefficients (k_kron_cross)
  DevBuf<double> ct, lval, rval;
  DevBuf<int32_t> lptr, lidx, rptr, ridx;
  // flat LLLR / RRRL state of (LL d + LR) / (RR d + RL) (-1: none) and each flat state's sector
  DevBuf<int32_t> aidx, bidx, asec, bsec;
  mutable DevBuf<double> z, zl, zc;       // H x of the last product is (z + zl) + zc: right, left, coupling
  mutable hipStream_t side = nullptr;     // the coupling terms' stream (default priority)
  mutable hipEvent_t fork = nullptr, join = nullptr;
  KronDev() = default;
  KronDev(const KronDev &) = delete;
  KronDev &operator=(const KronDev &) = delete;
  ~KronDev() {
    if (side) {
      (void)hipStreamSynchronize(side);
      (void)hipStreamDestroy(side);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
  }
  double flops = 0.0;                     // of the two dense products per H x
  // z = H x on stream s
  void product(const double *x, hipStream_t s) const;
};

// The operator a solver applies: an assembled full CSR, the matrix-free OBC form, or the
// Kronecker form.
struct Operator {
  int32_t n = 0;
  int lanes = 8;
  int grid = kGridCap;
  int load = 0;
  const dmrgx_matrix *A = nullptr;
  const ObcDev *obc = nullptr;
  const KronDev *kron = nullptr;
};
inline Operator op_of(const dmrgx_matrix &A) { return Operator{A.n, A.lanes, A.grid, A.load, &A, nullptr, nullptr}; }
inline Operator op_of(const ObcDev &O) { return Operator{O.n, O.lanes, kGridCap, 0, nullptr, &O, nullptr}; }
inline Operator op_of(const KronDev &K) { return Operator{K.n, 1, kGridCap, 0, nullptr, nullptr, &K}; }
const KronDev &kron_dev(const dmrgx_kron *K);

// ---- kernel launchers (kernels.hip)
struct Reduce {              // deterministic two-kernel reduction workspace
  DevBuf<double> partial;    // one slot per block, summed in order by a finish kernel
  void ensure(size_t blocks);
};

int pick_lanes(int32_t n, int64_t nnz);
// CONVERGE_CHECK's product over caller-held device CSR arrays (default schedule of a matrix
// of these dimensions), block partials at partial (residual_l1_default_blocks(n, nnz)
// doubles), sum_i |eig v_i - (H v)_i| at *out; allocates nothing.
int residual_l1_default_blocks(int32_t n, int64_t nnz);
void launch_residual_l1_into(int32_t n, int64_t nnz, const int64_t *row, const int32_t *col, const double *val,
                             const double *v, double eig, double *out, double *partial, hipStream_t s);
// lanes from the mean row length, the default workgroup cap, the dictionary for large A
void default_schedule(dmrgx_matrix &A);
int spmv_blocks(int32_t n, int lanes, int grid_cap);

void launch_spmv(const Operator &op, const double *x, double *y, hipStream_t s);
// y = A x and *dot = <x, y>: the shape of the solver's products (reduction + finish kernel)
void launch_spmv_dot(const Operator &op, const double *x, double *y, double *dot, Reduce &R, hipStream_t s);
// Full CSR from a lower triangle (diagonal last in every row).  A producer that already
// counted the strictly-lower entries per column (tcnt) and checked the rows' diagonals
// (bad, nonzero if a row lacks it) passes them; otherwise they are computed here.
void launch_expand_lower(int32_t n, int64_t nnz_tri, const int64_t *trow, const int32_t *tcol,
                         const double *tval, dmrgx_matrix &A, hipStream_t s,
                         const unsigned *tcnt_given = nullptr, const int *bad_given = nullptr);
void launch_to_fp32(const dmrgx_matrix &A, hipStream_t s);
// value dictionary of A (ent, table); returns the number of distinct values, 0 if none built
int64_t build_value_dict(dmrgx_matrix &A, hipStream_t s);
// delta form of A (first, dcol); returns false if a column difference needs more than 16 bits
bool build_delta(dmrgx_matrix &A, hipStream_t s);
void launch_start_vector(int64_t n, uint64_t seed, double *out, hipStream_t s);

// Lanczos step kernels.  Scalars live in device arrays indexed by step.
void launch_norm2(int64_t n, const double *x, Reduce &R, double *out, hipStream_t s);
void launch_lanczos_a(const Operator &op, bool fp32, int k, const double *w, const double *nrm2, double shift,
                      double *vk, double *y, double *alpha, Reduce &R, hipStream_t s);
void launch_lanczos_b(int64_t n, int k, const double *y, const double *vk, const double *vkm1,
                      const double *alpha, double *beta, double *w, double *nrm2, Reduce &R,
                      hipStream_t s);
void launch_ritz(int64_t n, int k, const double *const *vecs_dev, const double *coef_dev, double *out,
                 hipStream_t s);
void launch_scale_by_rnorm(int64_t n, const double *nrm2, const double *x, double *y, hipStream_t s);
void launch_copy(int64_t n, const double *x, double *y, hipStream_t s);

// minimum-residual solve of (H + shift) x = b: start, and one Givens update per Lanczos step
void launch_mr_init(const Operator &op, double shift, const double *b, const double *x, double *v, double *nrm2,
                    double *beta, double *mr, int *status, Reduce &R, hipStream_t s);
void launch_mr_step(int64_t n, int k, double acc, const double *alpha, const double *beta, double *mr, int *status,
                    double *coef, const double *v1, const double *w0, const double *w1, double *w2, double *x,
                    hipStream_t s);

// *out = <v, A v> in FP64 (the Rayleigh quotient of a normalised v)
void launch_rayleigh(const Operator &op, const double *v, double *out, Reduce &R, hipStream_t s);

// inverse iteration / CG kernels
void launch_residual_l1(const Operator &op, const double *v, double eig, double *out, Reduce &R,
                        hipStream_t s);
void launch_cg_init(const Operator &op, double shift, const double *nrm2, const double *x_in,
                    double *x_out, const double *b, double *r, double *p, double *rr0, Reduce &R,
                    hipStream_t s);
void launch_cg_ap(const Operator &op, double shift, const double *p, double *y, double *pap,
                  const int *done, Reduce &R, hipStream_t s);
void launch_cg_update(int64_t n, int it, double acc, const double *p, const double *y, double *x,
                      double *r, const double *pap, double *rr, int *status, Reduce &R,
                      hipStream_t s);
void launch_cg_dir(int64_t n, int it, const double *r, const double *rr, double *p, const int *status,
                   hipStream_t s);
// The same steps without their finish kernels: each consumer sums the producer's block
// partials itself, in the finish kernel's order (bit-identical scalars); block 0 stores the
// scalar the finish kernel would have stored.  Return the number of partials written.
int launch_lanczos_a_part(const Operator &op, bool fp32, int k, const double *w, double *nrm2, double *beta,
                          const double *pre, int pre_n, double *vk, double *y, double *alpha, Reduce &R,
                          hipStream_t s);
int launch_lanczos_b_part(int64_t n, int k, const double *y, const double *vk, const double *vkm1, double *alpha,
                          const double *beta, const double *pre, int pre_n, double *w, Reduce &R, hipStream_t s);
void launch_lanczos_b_finish(int k, double *beta, double *nrm2, const Reduce &R, int nblk, hipStream_t s);
int launch_cg_ap_part(const Operator &op, double shift, const double *p, double *y, const int *done, Reduce &R,
                      hipStream_t s);
// The mixed form's test of the CG iterate (inverse iteration with an inexact inner solve):
// with r = b - (H + d - eig) x the recursively updated residual, H u - eig u =
// (b - r - d x) / |x| for u = x / |x|, so sum_i |b_i - r_i - d x_i| / |x| is the L1 residual
// that the inverse iteration's convergence test takes of the normalised iterate.  When b is
// set, the CG update kernel also writes the block partials of sum |b - r - d x| and of |x|^2
// (partial[0..nblk) and partial[nblk..2 nblk)), and the direction kernel's block 0 ends the
// solve (status[0] = 1) once their quotient is below acc.
struct CgEarly {
  const double *b = nullptr;
  double diag_add = 0.0, acc = 0.0;
  double *partial = nullptr;
};
int launch_cg_update_part(int64_t n, int it, const double *p, const double *y, double *x, double *r, double *pap,
                          const double *rr, const int *status, const double *pre, int pre_n, Reduce &R,
                          hipStream_t s, const CgEarly &early = CgEarly{});
void launch_cg_dir_part(int64_t n, int it, double acc, const double *r, double *rr, double *p, int *status,
                        const double *pre, int pre_n, hipStream_t s, const CgEarly &early = CgEarly{});

// single-slot operator products: y = O x (y may be null) and, if dot_with != null,
// *dot_out = <dot_with, O x> (deterministic reduction)
void launch_slot_apply(int32_t n, const SlotRows &rows, const double *x, double *y, const double *dot_with,
                       double *dot_out, Reduce &R, hipStream_t s);
// the same for an on-site operator on an ED state (n = rows of the output basis)
void launch_ed_apply(int32_t n, const EdRows &rows, const double *x, double *y, const double *dot_with,
                     double *dot_out, Reduce &R, hipStream_t s);

// ---- host tridiagonal solver (tridiag.cpp)
// lowest eigenvalue of T(a[0..k-1], b[0..k-2]); upper_hint: an upper bound or +inf;
// lower_guess: a value probably below it, or NaN
double tridiag_lowest_value(int k, const double *a, const double *b, double upper_hint,
                            double lower_guess = std::numeric_limits<double>::quiet_NaN());
// Lowest eigenvalues of the successive Lanczos tridiagonals T_1, T_2, ... (T_k is the
// leading block of T_{k+1}, so by interlacing each value is at most the previous one: that
// bounds the search from above, and twice the last drop gives a guess below it).
struct RitzTrack {
  double last = std::numeric_limits<double>::infinity();
  double drop = std::numeric_limits<double>::quiet_NaN();
  double next(int k, const double *a, const double *b);
};
// eigenvector of T for eigenvalue lam, normalised
void tridiag_vector(int k, const double *a, const double *b, double lam, double *v);

// ---- solvers (solver.hip)
struct SolveOut {
  double eig = 0, err = 0;
  int lan_steps = 0, ii_steps = 0, cg_steps = 0, spmv = 0;
  bool lan_converged = false, cg_all_converged = true, ii_converged = false;
};
struct LanczosOpts {
  double acc = 1e-14;
  int min_step = 1, max_step = 800, max_block_step = 0;  // 0: no restart
  int type = 0;        // 0 Normal, 1 Tight
  bool fp32 = false;
  bool dense_small = true;  // n < 1000: residual-bound stop standing in for the dense dsyev_ branch
  // > 0: stop instead when the Ritz pair's residual bound beta_{k+1}|s_k| falls to
  // res_tol * max(1, |E_k|), or to res_tol with res_abs (type and acc are then not used)
  double res_tol = 0.0;
  bool res_abs = false;
};
// vec: device in/out (start on input, Ritz vector on output)
void lanczos(const Operator &A, const LanczosOpts &o, double *vec, SolveOut &out);
// Inverse iteration (shift folded into the operator), CG inner solver; vec device in/out.
// mr: minimum residual ("MR") instead of conjugate gradient ("CG") for the inner solves;
// rq_shift (the mixed form): after each inner solve the residual test and the next shift
// take the refined vector's FP64 Rayleigh quotient instead of the starting eigenvalue, and
// an inner CG solve also ends as soon as its normalised iterate passes half the residual
// test (CgEarly); the test after the solve decides as before
void inverse_iteration(const Operator &A, double eig, double *vec, double ii_acc, double ii_diag_add,
                       int ii_max_step, double cg_acc, int cg_max_step, SolveOut &out, bool mr = false,
                       bool rq_shift = false);
// the fraction of ii_acc at which the mixed form's inner solve stops early
constexpr double kCgEarlyFraction = 0.5;
// CG solve of (A + shift I) x = b with x normalised first; returns iterations, sets converged.
// early (optional): diag_add and acc of the iterate test above; b and partial are set here.
int cg_solve(const Operator &A, double shift, const double *b, double *x, double acc, int max_step,
             bool *converged, SolveOut &out, const CgEarly *early = nullptr);
// Minimum-residual solve of (A + shift I) x = b from x as given (not normalised); x in/out.
int mr_solve(const Operator &A, double shift, const double *b, double *x, double acc, int max_step,
             bool *converged, SolveOut &out);
double converge_check(const Operator &A, const double *v, double eig);
// <v, A v> / <v, v> in FP64
double rayleigh_quotient(const Operator &A, const double *v);
// FP32 copy of the values for the mixed-precision Lanczos (made once per matrix)
void ensure_fp32(const dmrgx_matrix &A);
// Builds A's value dictionary if it has not been tried; true if A has one (a matrix with
// more distinct values than a table worth caching keeps the plain form).
bool ensure_dict(const dmrgx_matrix &A);
// Builds A's delta form if it has not been tried; true if A has one.
bool ensure_delta(const dmrgx_matrix &A);
// Builds A's sliced form for P lanes per row (1, 2, 4) if not built (and its FP32 values
// with fp32); false for any other P.
bool ensure_sell(const dmrgx_matrix &A, int P, bool fp32 = false);
void build_sell(dmrgx_matrix &A, int P, bool fp32, hipStream_t s);
// the sliced form's 16-bit columns for P lanes per row (after ensure_sell); exceptions
int64_t ensure_sell16(const dmrgx_matrix &A, int P);
int64_t build_sell16(dmrgx_matrix &A, int P, hipStream_t s);
// format 4's value indices for P lanes per row (after ensure_sell16); false if A has no
// value table of at most kSellIdxMaxValues values.  fp32: also the FP32 table.
bool ensure_sellx(const dmrgx_matrix &A, int P, bool fp32);
void build_sellx(dmrgx_matrix &A, int P, bool fp32, hipStream_t s);
// frees the sliced forms other than P's (all with P = 0)
void drop_sell_except(const dmrgx_matrix &A, int P);
// A's load variant restricted to the formats it has (a format it cannot have reads CSR).
int effective_load(const dmrgx_matrix &A);

// device matrix handed over by dmrgx_get_ham_lllrrrrl for the CRS1 it returned (assemble.hip)
dmrgx_matrix *take_pending(const CRS1 *M);
void launch_exclusive_scan(int32_t n, const int64_t *len, int64_t *out, hipStream_t s);

// upload helpers (abi.cpp)
std::unique_ptr<dmrgx_matrix> upload_csr(int32_t n, const int64_t *row, const int32_t *col,
                                         const double *val, bool sym);

}  // namespace dmrgx
