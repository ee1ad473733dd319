// Batched ground-state engine: many independent solves advance together, one launch per
// kernel per step for all of them.
//
// Each request (a device matrix + solve parameters) occupies a slot and goes through the
// phases of dmrgx_ground_state -- Lanczos (sml/LANCZOS_SLOW_SYM.c:57-274 with the basis
// kept in HBM), then inverse iteration with CG (sml/INVERSE_ITERATION_SYM.c:9-116,
// sml/CONJUGATE_GRADIENT_SYM.c:8-123) -- but the per-step kernels are segmented: a launch
// covers the rows of every slot in that phase, a block belongs to exactly one slot, and
// each slot's block partials are summed in block order by a finish kernel (one block per
// slot).  A slot gets
// the same blocks, the same lanes per row and the same summation order as the single-solve
// kernels, so its result is bit-identical to dmrgx_ground_state on the same matrix, alone
// or next to any other requests.
//
// Per tick (one Lanczos step of every Lanczos slot, one CG iteration of every CG slot):
//   k_seg_spmv<G, LANCZOS_A>  v_k = w/|w|, y = H v_k, alpha_k       (one launch per lanes bin)
//   k_seg_lanczos_b           w = y - alpha v_k - beta v_{k-1}, |w|^2, beta_{k+1}
//   k_seg_spmv<G, CG_AP>      y = (H + shift) p, p.y
//   k_seg_cg_update           x += a p, r -= a y, |r|^2, done flag
//   k_seg_cg_dir              p = r + b p
// The per-tick descriptors of all slots are built on the host and copied once per chunk of
// ticks; the host reads the recurrence scalars once per chunk (one gather kernel, one copy),
// runs the tridiagonal checks, and moves slots between phases.  Chunk c + 1 is launched
// before chunk c's read-back is processed, so that host work overlaps the device: a slot
// that converged in chunk c has run unused steps in chunk c + 1 (CG kernels skip them on
// the device; Lanczos slots close to convergence get only the steps their convergence rate
// predicts), which never changes a result.  New requests are admitted at chunk boundaries
// while the HBM budget allows (continuous batching).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <unordered_map>

#include "entry.h"
#include "reduce.h"

namespace dmrgx {
namespace {

enum Mode { LANCZOS_A = 0, CG_AP = 1 };

// One slot's view for one launch of a segmented kernel.
struct SegDesc {
  const int64_t *row;
  const int32_t *col;
  const double *val;
  const float *val32;
  const uint64_t *ent;    // value dictionary form (FP64 products)
  const double *table;
  const int32_t *first;   // delta form
  const uint16_t *dcol;
  const int64_t *sptr;    // sliced form (lanes per row 1, 2 or 4)
  const int32_t *slen, *scol;
  const double *sval;
  const float *sval32;
  const uint16_t *sd16;   // sliced form's 16-bit columns (load bit 0, format 4)
  const int32_t *sfirst, *sexc;
  const int64_t *septr;
  const uint16_t *sidx;   // format 4: value indices into stab / stab32
  const double *stab;
  const float *stab32;
  int32_t n, fp32;
  int32_t blk_lo, nblk;   // this slot's blocks in the launch
  int32_t k, load;        // step (Lanczos) or iteration (CG); SpMV load variant
  double shift, acc;
  const double *x;        // LANCZOS_A: w      CG_AP: p
  double *o1, *o2;        // LANCZOS_A: v_k, y CG_AP: y
  const double *o3;       // lanczos_b: v_{k-1} (null at k = 0); cg_update/dir: p
  double *s1, *s2, *s3;   // LANCZOS_A: nrm2, alpha       lanczos_b: alpha, beta, nrm2
                          // CG_AP: pap                   cg_update: pap, rr      cg_dir: rr
  double *v1, *v2;        // cg_update: x, r
  int *status;            // CG: {done, last iteration}
  const double *eb;       // cg_update, mixed form: b of the iterate test (CgEarly), else null
  double eadd, eacc;      //   its diag_add and threshold
};

// Segment (slot) of block b of a launch: map[b / kMapStride] is the segment of block
// kMapStride * (b / kMapStride); the later ones start at desc[s].blk_lo, scanned forward.
// The map costs one entry per kMapStride blocks instead of one per block.
constexpr int kMapStride = 64;
__device__ __forceinline__ int block_segment(const SegDesc *desc, const int32_t *map, int nseg, int b) {
  int s = map[b / kMapStride];
  while (s + 1 < nseg && desc[s + 1].blk_lo <= b) s++;
  return s;
}

enum FinKind { FIN_LANCZOS_A = 0, FIN_CG_AP = 1, FIN_LANCZOS_B = 2, FIN_CG_UPDATE = 3 };

// One block per segment: its partials summed in block order (the single-solve k_finish
// over the segment's blocks), then the kernel's scalar update.
// For FIN_CG_UPDATE with the iterate test, partial2 holds the update kernel's two extra
// partial arrays (tot = that launch's block count apart), summed as k_cg_dir's block 0 does.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_seg_finish(const SegDesc *desc, const double *partial,
                                                       const double *partial2, int tot) {
  const SegDesc d = desc[blockIdx.x];
  if ((KIND == FIN_CG_AP || KIND == FIN_CG_UPDATE) && d.status[0] != 0) return;
  const double t = finish_sum(partial + d.blk_lo, d.nblk);
  bool early_done = false;
  if (KIND == FIN_CG_UPDATE && d.eb) {
    __syncthreads();   // block_sum's LDS is reused
    const double e1 = finish_sum(partial2 + d.blk_lo, d.nblk);
    __syncthreads();
    const double e2 = finish_sum(partial2 + tot + d.blk_lo, d.nblk);
    early_done = e2 > 0.0 && e1 < d.eacc * sqrt(e2);   // valid in thread 0
  }
  if (threadIdx.x != 0) return;
  if (KIND == FIN_LANCZOS_A) {
    *d.s2 = t;                 // alpha_k
  } else if (KIND == FIN_CG_AP) {
    *d.s1 = t;                 // pAp
  } else if (KIND == FIN_LANCZOS_B) {
    *d.s3 = t;                 // |w|^2
    d.s2[1] = sqrt(t);         // beta_{k+1}
  } else {
    d.s2[d.k] = t;             // rr_it
    d.status[1] = d.k;
    if (t < d.acc || early_done) d.status[0] = 1;
  }
}

template <int G, int MODE>
__global__ __launch_bounds__(kBlock) void k_seg_spmv(const SegDesc *desc, const int32_t *map, int nseg,
                                                    double *partial) {
  const int seg = block_segment(desc, map, nseg, blockIdx.x);
  const SegDesc d = desc[seg];
  if (MODE == CG_AP && d.status[0] != 0) return;
  constexpr int RPB = kBlock / G;
  const int g = threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const double scale = (MODE == LANCZOS_A) ? 1.0 / sqrt(*d.s1) : 0.0;
  double c = 0.0;
  // row groups over the slot's logical blocks, as k_apply deals them (the partial of
  // logical block lb is at blk_lo + lb, where the finish kernel sums it in lb order)
  const int32_t lb = blockIdx.x - d.blk_lo;
  const GroupRange gr = group_range(lb, d.nblk, (d.n + RPB - 1) / RPB, (d.load & kLoadXcd) != 0);
  const int fmt = load_format(d.load);
  for (int32_t gi = gr.start; gi < gr.end; gi += gr.step) {
    const int32_t r = gi * RPB + g;
    double s = 0.0;
    if (r < d.n && sliced_format(fmt)) {
      if constexpr (G <= kSellMaxLanes) {
        const bool nt = (d.load & kLoadDeep) != 0;
        if (d.sidx) {
          if (d.fp32)
            s = nt ? sellx_lane_sum<G, float, true>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sidx,
                                                    d.stab32, r, lane, d.x)
                   : sellx_lane_sum<G, float>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sidx, d.stab32, r,
                                              lane, d.x);
          else
            s = nt ? sellx_lane_sum<G, double, true>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sidx,
                                                     d.stab, r, lane, d.x)
                   : sellx_lane_sum<G, double>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sidx, d.stab, r,
                                               lane, d.x);
        } else if (d.sd16) {
          if (d.fp32)
            s = nt ? sell16_lane_sum<G, float, true>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sval32, r,
                                                     lane, d.x)
                   : sell16_lane_sum<G, float>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sval32, r,
                                               lane, d.x);
          else
            s = nt ? sell16_lane_sum<G, double, true>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sval, r,
                                                      lane, d.x)
                   : sell16_lane_sum<G, double>(d.sptr, d.slen, d.sd16, d.sfirst, d.septr, d.sexc, d.sval, r, lane,
                                                d.x);
        } else if (d.fp32) {
          s = nt ? sell_lane_sum<G, float, true>(d.sptr, d.slen, d.scol, d.sval32, r, lane, d.x)
                 : sell_lane_sum<G, float>(d.sptr, d.slen, d.scol, d.sval32, r, lane, d.x);
        } else {
          s = nt ? sell_lane_sum<G, double, true>(d.sptr, d.slen, d.scol, d.sval, r, lane, d.x)
                 : sell_lane_sum<G, double>(d.sptr, d.slen, d.scol, d.sval, r, lane, d.x);
        }
      }
    } else if (r < d.n) {
      switch (d.load & (kLoadXcd - 1)) {
#define V_CASE(V)                                                                              \
  case V:                                                                                      \
    if (fmt == kFmtDelta) {                                                                    \
      if (d.fp32) s = delta_lane_sum_v<G, V>(d.row, d.first, d.dcol, d.val32, r, lane, d.x);   \
      else s = delta_lane_sum_v<G, V>(d.row, d.first, d.dcol, d.val, r, lane, d.x);          \
    } else if (d.fp32 && fmt == kFmtCsr) {                                                     \
      s = csr_lane_sum_v<G, V>(d.row, d.col, d.val32, r, lane, d.x);                           \
    } else if (d.fp32) { /* FP32 values of a dictionary matrix: CSR without paired loads */    \
      s = csr_lane_sum_v<G, (V & ~kLoadPair)>(d.row, d.col, d.val32, r, lane, d.x);            \
    } else if (fmt == kFmtDict) {                                                              \
      s = dict_lane_sum_v<G, V>(d.row, d.ent, d.table, r, lane, d.x);                          \
    } else {                                                                                   \
      s = csr_lane_sum_v<G, V>(d.row, d.col, d.val, r, lane, d.x);                             \
    }                                                                                          \
    break;
        V_CASE(0) V_CASE(1) V_CASE(2) V_CASE(3)
#undef V_CASE
      }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (lane == 0 && r < d.n) {
      if (MODE == LANCZOS_A) {
        const double xs = scale * d.x[r];
        d.o1[r] = xs;
        const double yv = scale * s + d.shift * xs;
        d.o2[r] = yv;
        c += yv * xs;
      } else {
        const double yv = s + d.shift * d.x[r];
        d.o1[r] = yv;
        c += d.x[r] * yv;
      }
    }
  }
  block_partial(c, partial);
}

// Vector kernels: a slot gets vec_grid(n) blocks and the single-solve grid-stride pattern.
__global__ __launch_bounds__(kBlock) void k_seg_lanczos_b(const SegDesc *desc, const int32_t *map, int nseg,
                                                          double *partial) {
  const int seg = block_segment(desc, map, nseg, blockIdx.x);
  const SegDesc d = desc[seg];
  const int lb = blockIdx.x - d.blk_lo;
  const double a = *d.s1;
  const double bb = (d.k > 0) ? *d.s2 : 0.0;
  const double *vkm1 = d.o3 ? d.o3 : d.o1;
  double s = 0.0;
  const int64_t S = (int64_t)d.nblk * kBlock;
  for (int64_t i0 = lb * (int64_t)kBlock + threadIdx.x; i0 < d.n; i0 += 4 * S) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * S;
      if (i < d.n) {
        v[u] = d.o2[i] - a * d.o1[i];
        if (d.k > 0) v[u] -= bb * vkm1[i];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * S;
      if (i < d.n) {
        const_cast<double *>(d.x)[i] = v[u];
        s += v[u] * v[u];
      }
    }
  }
  block_partial(s, partial);
}

// With d.eb set (the mixed form) also the iterate test's partials (CgEarly, as k_cg_update):
// sum |b - r - diag_add x| at partial2[b], |x|^2 at partial2[gridDim.x + b].
__global__ __launch_bounds__(kBlock) void k_seg_cg_update(const SegDesc *desc, const int32_t *map, int nseg,
                                                          double *partial, double *partial2) {
  const int seg = block_segment(desc, map, nseg, blockIdx.x);
  const SegDesc d = desc[seg];
  if (d.status[0] != 0) return;
  const int lb = blockIdx.x - d.blk_lo;
  const int it = d.k;
  const double alpha = d.s2[it - 1] / d.s1[it];
  double s = 0.0, e1 = 0.0, e2 = 0.0;
  const int64_t S = (int64_t)d.nblk * kBlock;
  for (int64_t i0 = lb * (int64_t)kBlock + threadIdx.x; i0 < d.n; i0 += 4 * S) {
    double rv[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * S;
      if (i < d.n) {
        xv[u] = d.v1[i] + alpha * d.o3[i];
        d.v1[i] = xv[u];
        rv[u] = d.v2[i] - alpha * d.o1[i];
        d.v2[i] = rv[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + u * S < d.n) s += rv[u] * rv[u];
    if (d.eb) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t i = i0 + u * S;
        if (i < d.n) {
          e1 += fabs(d.eb[i] - rv[u] - d.eadd * xv[u]);
          e2 += xv[u] * xv[u];
        }
      }
    }
  }
  block_partial(s, partial);
  if (d.eb) {   // uniform per block; block_sum's LDS is reused, hence the barriers
    __syncthreads();
    block_partial(e1, partial2);
    __syncthreads();
    block_partial_to(e2, partial2 + gridDim.x + blockIdx.x);
  }
}

// One block per piece: copies count 8-byte words from src to dst + off, so that the host
// reads every slot's new scalars with one device-to-host copy.
struct GatherDesc {
  const double *src;
  int64_t off;
  int32_t count, pad;
};
__global__ __launch_bounds__(kBlock) void k_gather_words(const GatherDesc *desc, double *dst) {
  const GatherDesc d = desc[blockIdx.x];
  for (int i = threadIdx.x; i < d.count; i += kBlock) dst[d.off + i] = d.src[i];
}

__global__ __launch_bounds__(kBlock) void k_seg_cg_dir(const SegDesc *desc, const int32_t *map, int nseg) {
  const int seg = block_segment(desc, map, nseg, blockIdx.x);
  const SegDesc d = desc[seg];
  if (d.status[0] != 0) return;
  const int lb = blockIdx.x - d.blk_lo;
  const int it = d.k;
  const double beta = d.s2[it] / d.s2[it - 1];
  for (int64_t i = lb * (int64_t)kBlock + threadIdx.x; i < d.n; i += (int64_t)d.nblk * kBlock)
    d.v1[i] = d.v2[i] + beta * d.v1[i];
}

// the vector kernels' blocks per slot: the single-solve count (same partials, same scalars)
int vec_blocks(int64_t n) { return reduce_grid(n); }

double now_sec() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Lanczos basis of one slot, chunks of 32 vectors (as solver.hip).
struct SlotBasis {
  int64_t n = 0;
  std::vector<DevBuf<double>> chunks;
  double *vec(int k) {
    const size_t c = (size_t)k / 32;
    while (chunks.size() <= c) chunks.emplace_back((size_t)n * 32);
    return chunks[c].p + (size_t)(k % 32) * n;
  }
};

}  // namespace

// P_WAIT: the slot's next step depends on a scalar its last launches compute (the L1
// residual of an inverse-iteration round, or the mixed form's Rayleigh quotient); the
// engine reads those scalars for all waiting slots together.
enum Phase { P_LANCZOS, P_CG, P_WAIT, P_DONE };
// residual; final Rayleigh quotient; mixed-form shift before the first inner solve and
// after each later one
enum Wait { W_RESID = 1, W_RQ = 2, W_SHIFT = 3, W_REFRESH = 4 };

struct Slot {
  int64_t id = 0;
  uint64_t tag = 0;
  const dmrgx_matrix *A = nullptr;
  dmrgx_matrix *owned = nullptr;
  dmrgx_solve_params p{};
  bool mixed = false;
  Phase phase = P_LANCZOS;
  int epoch = 0;          // bumped at every phase change (stale read-backs are ignored)
  Wait wait = W_RESID;
  bool wait_gathered = false;   // the awaited scalars ride on a launched chunk's read-back
  double dE_prev = INFINITY, dE_last = INFINITY;   // last two |E_k - E_{k-1}| of the checks
  int status = DMRGX_OK;
  int32_t n = 0;
  int lanes = 8;
  size_t bytes = 0;
  // device vectors
  DevBuf<double> w, y, vec, un, r, pv, cy, xw;
  SlotBasis V;
  DevBuf<double> scal;   // alpha[KS] beta[KS] nrm2 | rr[CS] pap[CS] | residual
  DevBuf<int> cgst;      // CG {done, last iteration}
  int KS = 0, CS = 0;
  // host Lanczos state (as solver.hip:lanczos)
  int k_next = 0;        // next Lanczos step to compute
  int k_read = 0;        // steps whose scalars the host has checked
  double E = 1e15, lam = 0.0;
  RitzTrack track;       // lowest Ritz values of the growing tridiagonal
  std::vector<double> ha, hb, s;
  // II / CG state
  double eig = 0.0, err = 1000.0;
  int ii_iter = 0, cg_it = 0;
  SolveOut out;
  double t_submit = 0, t_start = 0, t_done = 0;
  double *alpha() { return scal.p; }
  double *beta() { return scal.p + KS; }
  double *nrm2() { return scal.p + 2 * KS; }
  double *rr() { return scal.p + 2 * KS + 1; }
  double *pap() { return scal.p + 2 * KS + 1 + CS; }
  double *resid() { return scal.p + 2 * KS + 1 + 2 * CS; }
  double *rq() { return scal.p + 2 * KS + 2 + 2 * CS; }   // <v, H v>, <v, v>
  std::vector<const double *> ritz_ptrs;                   // kept until the slot ends
};

}  // namespace dmrgx

namespace dmrgx {
// A launched chunk's read-back: the slots whose scalars it carries (with their phase epoch
// at launch), the Lanczos steps and CG iterations it ran, and where each slot's words are
// in the pinned read-back area.
struct Inflight {
  std::vector<Slot *> L, Cg, W;   // valid only while launching (slots may retire later)
  std::vector<int64_t> idL, idC, idW;
  std::vector<int> epL, epC, epW;
  std::vector<int> k_start, k_end, cg_end;
  std::vector<size_t> off;   // L: 2 (k_end - k_start) words, then Cg: 1 word, then W: 1 or 2
  int area = 0;
};
}  // namespace dmrgx

struct dmrgx_engine {
  dmrgx_engine_opts opts{};
  int64_t next_id = 1;
  size_t bytes_in_use = 0;
  std::deque<std::unique_ptr<dmrgx::Slot>> pending;
  std::vector<std::unique_ptr<dmrgx::Slot>> active;
  std::deque<std::unique_ptr<dmrgx::Slot>> done;
  // launch scratch
  dmrgx::DevBuf<dmrgx::SegDesc> ddesc;
  dmrgx::DevBuf<int32_t> dblk;
  dmrgx::DevBuf<double> partial;
  dmrgx::DevBuf<double> partial2;   // the mixed form's CG iterate-test partials
  dmrgx::DevBuf<dmrgx::GatherDesc> dgather;  // per-chunk scalar read-back
  dmrgx::DevBuf<double> dstage;
  // host seconds: building + uploading descriptors, launching, waiting for the chunk,
  // Lanczos checks (with the Ritz/inverse-iteration starts they trigger), CG ends and
  // inverse-iteration rounds, admission; then the number of chunks
  double host_t[6] = {0, 0, 0, 0, 0, 0};
  int64_t chunks = 0;
  // Two pinned areas, used by consecutive chunks in turn (a chunk is launched before the
  // previous one's read-back is processed): uploads, read-backs, and the read-back events.
  void *up_pin[2] = {nullptr, nullptr};
  size_t up_bytes[2] = {0, 0};
  void *rb_pin[2] = {nullptr, nullptr};
  size_t rb_bytes[2] = {0, 0};
  hipEvent_t rb_ev[2] = {nullptr, nullptr};
  int area = 0;
  std::unique_ptr<dmrgx::Inflight> fl;   // launched, read-back not processed yet (slots by id)
  // hipGraph of the last chunk launch sequence seen twice in a row (kernels only: the
  // descriptor uploads and the read-back stay outside), replayed while it repeats
  std::vector<int64_t> prev_sig, graph_sig;
  hipGraphExec_t graph = nullptr;
  int64_t n_eager = 0, n_captures = 0, n_replays = 0;
  ~dmrgx_engine() {
    (void)hipStreamSynchronize(dmrgx::ctx().stream);
    if (graph) (void)hipGraphExecDestroy(graph);
    for (int i = 0; i < 2; i++) {
      if (up_pin[i]) (void)hipHostFree(up_pin[i]);
      if (rb_pin[i]) (void)hipHostFree(rb_pin[i]);
      if (rb_ev[i]) (void)hipEventDestroy(rb_ev[i]);
    }
  }
};

namespace dmrgx {
namespace {

size_t slot_bytes(const dmrgx_engine &E, const Slot &s) {
  // eight work vectors, the scalars and an estimate of the Lanczos basis
  const size_t n = (size_t)s.n;
  const size_t steps = (size_t)std::min(s.p.diag_max_step + 1, std::max(E.opts.basis_steps_estimate, 1));
  size_t b = n * 8 * (8 + steps) + (size_t)(2 * s.KS + 2 * s.CS + 2) * 8;
  if (s.owned) b += (size_t)s.owned->nnz * 12 + (size_t)(s.n + 1) * 8;
  return b;
}

// The slot matrix's load variant, restricted to the formats the matrix has.
int slot_load(const Slot &s) { return effective_load(*s.A); }

// the sliced-form arrays of A for its lanes per row (null when A has none)
void set_sell(SegDesc &a, const dmrgx_matrix &A) {
  const int lp = A.lanes == 1 ? 0 : (A.lanes == 2 ? 1 : (A.lanes == 4 ? 2 : -1));
  if (lp < 0) return;
  a.sptr = A.sell_ptr[lp].p;
  a.slen = A.sell_len.p;
  a.scol = A.sell_col[lp].p;
  a.sval = A.sell_val[lp].p;
  a.sval32 = A.sell_val32[lp].p;
  const int fmt = load_format(A.load);
  if ((fmt == kFmtSellIdx || (fmt == kFmtSell && (A.load & kLoadPair))) && A.sell_nexc[lp] >= 0) {   // 16-bit columns
    a.sd16 = A.sell_d16[lp].p;
    a.sfirst = A.sell_first[lp].p;
    a.septr = A.sell_eptr[lp].p;
    a.sexc = A.sell_exc[lp].p;
  }
  if (fmt == kFmtSellIdx && A.sell_idx_ok[lp] == 1) {   // value indices (made at slot start)
    a.sidx = A.sell_idx16[lp].p;
    a.stab = A.table.p;
    a.stab32 = A.table32.p;
  }
}

// Start a slot: start vector and |w|^2 (solver.hip:lanczos before its first step).
void start_slot(Slot &s) {
  hipStream_t st = ctx().stream;
  const int64_t n = s.n;
  s.w.resize(n);
  s.y.resize(n);
  s.vec.resize(n);
  s.V.n = n;
  s.KS = s.p.diag_max_step + 2;
  s.CS = std::max(s.p.cg_max_step, 1) + 1;
  s.scal.resize(2 * (size_t)s.KS + 2 * (size_t)s.CS + 4);
  s.cgst.resize(2);
  s.ha.assign(s.KS, 0.0);
  s.hb.assign(s.KS, 0.0);
  s.s.assign(s.KS, 0.0);
  if (s.mixed) ensure_fp32(*s.A);
  if (load_format(s.A->load) == kFmtDict) ensure_dict(*s.A);
  if (load_format(s.A->load) == kFmtDelta) ensure_delta(*s.A);
  const int fmt = load_format(s.A->load);
  if (fmt == kFmtSell) {
    ensure_sell(*s.A, s.A->lanes, s.mixed);
    if (s.A->load & kLoadPair) ensure_sell16(*s.A, s.A->lanes);
  } else if (fmt == kFmtSellIdx) {
    // format 4, or format 3 with 16-bit columns where the matrix has no small value table
    ensure_sell(*s.A, s.A->lanes, false);
    ensure_sell16(*s.A, s.A->lanes);
    if (!ensure_sellx(*s.A, s.A->lanes, s.mixed)) ensure_sell(*s.A, s.A->lanes, s.mixed);
  }
  launch_start_vector(n, s.p.seed, s.vec.p, st);
  launch_copy(n, s.vec.p, s.w.p, st);
  Reduce R;
  launch_norm2(n, s.w.p, R, s.nrm2(), st);
  HIPCHK(hipGetLastError());
  s.t_start = now_sec();
}

void finish_slot(Slot &s, int status) {
  s.phase = P_DONE;
  s.epoch++;
  s.status = status;
  s.t_done = now_sec();
  // release the work vectors; vec (the result) stays until polled
  s.w.reset();
  s.y.reset();
  s.un.reset();
  s.r.reset();
  s.pv.reset();
  s.cy.reset();
  s.xw.reset();
  s.V.chunks.clear();
  s.scal.reset();
  s.cgst.reset();
}

// CG start for the current II iteration (solver.hip:cg_solve before its loop)
void start_cg(Slot &s) {
  hipStream_t st = ctx().stream;
  const int64_t n = s.n;
  Reduce R;
  const double shift = s.p.ii_diag_add - s.eig;
  HIPCHK(hipMemsetAsync(s.cgst.p, 0, 2 * sizeof(int), st));
  launch_norm2(n, s.un.p, R, s.nrm2(), st);
  launch_cg_init(op_of(*s.A), shift, s.nrm2(), s.un.p, s.xw.p, s.vec.p, s.r.p, s.pv.p, s.rr(), R, st);
  s.out.spmv++;
  s.cg_it = 1;
  s.phase = P_CG;
  s.epoch++;
}

// End of the inverse iteration (solver.hip:inverse_iteration after its loop): counters, and
// for the mixed form the FP64 Rayleigh quotient, read back with the other waiting slots.
void ii_finalize(Slot &s) {
  if (!s.p.skip_ii) {
    s.out.ii_steps = s.ii_iter;
    s.out.err = s.err;
  }
  if (s.mixed && s.n > 1) {
    Reduce R;
    launch_rayleigh(op_of(*s.A), s.vec.p, s.rq(), R, ctx().stream);
    launch_norm2(s.n, s.vec.p, R, s.rq() + 1, ctx().stream);
    s.phase = P_WAIT;
    s.epoch++;
    s.wait = W_RQ;
    return;
  }
  s.out.eig = s.eig;
  finish_slot(s, DMRGX_OK);
}

// Inverse-iteration step after Lanczos or after a CG solve (solver.hip:inverse_iteration):
// launches the residual check and waits for it, or ends the iteration.
void ii_advance(Slot &s) {
  if (!s.p.skip_ii && s.ii_iter <= s.p.ii_max_step) {
    Reduce R;
    launch_residual_l1(op_of(*s.A), s.vec.p, s.eig, s.resid(), R, ctx().stream);
    s.phase = P_WAIT;
    s.epoch++;
    s.wait = W_RESID;
    return;
  }
  ii_finalize(s);
}

void begin_ii(Slot &s);

// A waiting slot's scalars arrived: residual -> done or the next CG solve; Rayleigh
// quotient -> the mixed form's shift (inverse iteration starts) or done.
void wait_resolved(Slot &s, const double *h) {
  if (s.wait == W_RESID) {
    s.err = h[0];
    s.out.spmv++;
    if (s.err < s.p.ii_acc) {
      s.out.ii_converged = true;
      ii_finalize(s);
    } else {
      start_cg(s);
    }
  } else if (s.wait == W_SHIFT) {
    s.eig = h[0] / h[1];
    s.out.spmv++;
    begin_ii(s);
  } else if (s.wait == W_REFRESH) {
    s.eig = h[0] / h[1];
    s.out.spmv++;
    ii_advance(s);
  } else {
    s.eig = h[0] / h[1];
    s.out.spmv++;
    s.out.eig = s.eig;
    finish_slot(s, DMRGX_OK);
  }
}

// A CG solve ended (converged or out of steps): x -> un, normalise, vec = un, next II round.
void cg_finished(Slot &s, int done_it) {
  hipStream_t st = ctx().stream;
  const int64_t n = s.n;
  Reduce R;
  const bool conv = done_it >= 0;
  if (!conv) s.out.cg_all_converged = false;
  s.out.cg_steps += conv ? done_it : s.p.cg_max_step;
  s.out.spmv += conv ? done_it : std::max(s.p.cg_max_step - 1, 0);
  launch_copy(n, s.xw.p, s.un.p, st);
  launch_norm2(n, s.un.p, R, s.nrm2(), st);
  launch_scale_by_rnorm(n, s.nrm2(), s.un.p, s.un.p, st);
  launch_copy(n, s.un.p, s.vec.p, st);
  s.ii_iter++;
  if (s.mixed && s.ii_iter <= s.p.ii_max_step) {
    // mixed form: the next residual test and shift at the refined vector's FP64 Rayleigh
    // quotient (solver.hip:inverse_iteration with rq_shift)
    launch_rayleigh(op_of(*s.A), s.vec.p, s.rq(), R, st);
    launch_norm2(n, s.vec.p, R, s.rq() + 1, st);
    s.phase = P_WAIT;
    s.epoch++;
    s.wait = W_REFRESH;
    return;
  }
  ii_advance(s);
}

// First inverse-iteration round from the Ritz vector in vec.
void begin_ii(Slot &s) {
  launch_copy(s.n, s.vec.p, s.un.p, ctx().stream);
  s.ii_iter = 0;
  ii_advance(s);
}

// The Lanczos run of this slot converged at step kconv: Ritz vector, then inverse iteration.
void lanczos_finished(Slot &s, int kconv) {
  hipStream_t st = ctx().stream;
  const int64_t n = s.n;
  tridiag_vector(kconv + 1, s.ha.data(), s.hb.data() + 1, s.lam, s.s.data());
  std::vector<const double *> &ptrs = s.ritz_ptrs;   // host source of an asynchronous copy
  ptrs.resize(kconv + 1);
  for (int j = 0; j <= kconv; j++) ptrs[j] = s.V.vec(j);
  DevBuf<const double *> dptr(kconv + 1);
  DevBuf<double> dcoef(kconv + 1);
  HIPCHK(hipMemcpyAsync(dptr.p, ptrs.data(), sizeof(double *) * (kconv + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dcoef.p, s.s.data(), sizeof(double) * (kconv + 1), hipMemcpyHostToDevice, st));
  launch_ritz(n, kconv, dptr.p, dcoef.p, s.vec.p, st);
  Reduce R;
  launch_norm2(n, s.vec.p, R, s.nrm2(), st);
  launch_scale_by_rnorm(n, s.nrm2(), s.vec.p, s.vec.p, st);
  // no synchronisation: the basis memory goes back to the allocator and is reused only by
  // work queued after the Ritz kernel on the same stream
  s.V.chunks.clear();
  s.eig = s.lam;
  s.out.lan_steps = kconv;   // steps after the first, as solver.hip:lanczos counts them
  s.out.lan_converged = true;
  if (s.p.skip_ii) {
    ii_advance(s);
    return;
  }
  s.un.resize(n);
  s.r.resize(n);
  s.pv.resize(n);
  s.cy.resize(n);
  s.xw.resize(n);
  if (s.mixed && n > 1) {
    // the shift of the mixed form's inverse iteration: the FP64 Rayleigh quotient of the
    // Ritz vector (dmrgx_ground_state does the same), read back with the waiting slots
    Reduce R;
    launch_rayleigh(op_of(*s.A), s.vec.p, s.rq(), R, st);
    launch_norm2(n, s.vec.p, R, s.rq() + 1, st);
    s.phase = P_WAIT;
    s.epoch++;
    s.wait = W_SHIFT;
    return;
  }
  begin_ii(s);
}

// Host checks of the steps computed since the last chunk (solver.hip:lanczos, no restarts).
// alpha_new[j] = alpha_{k_read + j}, beta_new[j] = beta_{k_read + 1 + j} for the steps
// k_read .. k_to - 1, read back since the last check.
void lanczos_check(Slot &s, int k_to, const double *alpha_new, const double *beta_new) {
  const bool small = s.n < 1000;
  for (int k = s.k_read; k < k_to; k++) {
    s.ha[k] = alpha_new[k - s.k_read];
    s.hb[k + 1] = beta_new[k - s.k_read];
  }
  for (int k = s.k_read; k < k_to; k++) {
    s.k_read = k + 1;
    s.out.spmv++;
    if (k == 0) continue;  // step 0 is the start of the recurrence (solver.hip: block step 0)
    const bool check = small || k >= s.p.diag_min_step;
    if (!check) continue;
    s.lam = s.track.next(k + 1, s.ha.data(), s.hb.data() + 1);
    bool conv;
    const double bnext = s.hb[k + 1];
    const double tscale = std::max(1.0, std::fabs(s.lam));
    if (small) {
      tridiag_vector(k + 1, s.ha.data(), s.hb.data() + 1, s.lam, s.s.data());
      conv = !(bnext > 1e-12 * tscale) || bnext * std::fabs(s.s[k]) <= 1e-12 * tscale;
    } else if (!(bnext > 1e-14 * tscale)) {
      conv = true;
    } else if (s.p.lan_type == 1) {
      tridiag_vector(k + 1, s.ha.data(), s.hb.data() + 1, s.lam, s.s.data());
      conv = std::fabs(s.s[k]) < s.p.diag_acc;
    } else {
      conv = std::fabs(s.E - s.lam) < s.p.diag_acc;
    }
    if (conv) {
      lanczos_finished(s, k);
      return;
    }
    s.dE_prev = s.dE_last;
    s.dE_last = std::fabs(s.E - s.lam);
    s.E = s.lam;
  }
  if (k_to >= s.p.diag_max_step) {
    s.out.lan_steps = k_to - 1;
    s.out.eig = s.lam;
    finish_slot(s, DMRGX_ERR_NOCONV);
  }
}

// Lay out one segmented launch: blocks of each slot consecutively, desc index = slot order.
struct LaunchPlan {
  std::vector<SegDesc> desc;
  std::vector<int32_t> blk;
};

// n_map: block maps of all launches; max_blocks: the largest launch (one partial per block)
void ensure_scratch(dmrgx_engine &E, size_t n_desc, size_t n_map, size_t max_blocks) {
  if (E.ddesc.count < n_desc) E.ddesc.resize(std::max<size_t>(n_desc, 2 * E.ddesc.count));
  if (E.dblk.count < n_map) E.dblk.resize(std::max<size_t>(n_map, 2 * E.dblk.count));
  if (E.partial.count < max_blocks) E.partial.resize(std::max<size_t>(max_blocks, 2 * E.partial.count));
  if (E.partial2.count < 2 * max_blocks) E.partial2.resize(std::max<size_t>(2 * max_blocks, 2 * E.partial2.count));
}

template <int MODE>
void launch_seg_spmv(int G, int nblk, const SegDesc *desc, const int32_t *map, int nseg, double *partial,
                     hipStream_t st) {
  switch (G) {
#define G_CASE(g) \
  case g: k_seg_spmv<g, MODE><<<nblk, kBlock, 0, st>>>(desc, map, nseg, partial); break;
    G_CASE(1) G_CASE(2) G_CASE(4) G_CASE(8) G_CASE(16) G_CASE(32) G_CASE(64)
#undef G_CASE
    default: throw Failure(DMRGX_ERR_ARG, "bad lanes per row");
  }
}

// One chunk of ticks over all active slots.
void *grow_pinned(void *&p, size_t &have, size_t want) {
  if (have < want) {
    const size_t nb = std::max(want, 2 * have);
    if (p) HIPCHK(hipHostFree(p));
    p = nullptr;
    have = 0;
    HIPCHK(hipHostMalloc(&p, nb, hipHostMallocDefault));
    have = nb;
  }
  return p;
}

// Lanczos steps to launch for a slot in the next chunk.  The chunk is launched before the
// previous read-back is checked, so a slot about to converge would run steps that are
// never used; when the last two energy changes predict convergence (geometric decrease)
// the slot gets only the predicted steps plus two.  The prediction never changes a
// result, only how many unused steps run.
int lan_budget_of(const Slot &s, int ticks) {
  if (s.n < 1000 || s.p.lan_type == 1) return ticks;
  if (!(std::isfinite(s.dE_last) && std::isfinite(s.dE_prev)) || !(s.dE_last > 0.0) || !(s.dE_last < s.dE_prev))
    return ticks;
  const double need = std::ceil(std::log(s.p.diag_acc / s.dE_last) / std::log(s.dE_last / s.dE_prev));
  const double unchecked = (double)(s.k_next - s.k_read);
  return (int)std::min<double>(ticks, std::max(0.0, need + 2.0 - unchecked));
}

// Launches one chunk of ticks for every Lanczos and CG slot and the read-back of their new
// scalars (and of every waiting slot's) into a pinned area, without waiting for it.
std::unique_ptr<Inflight> launch_chunk(dmrgx_engine &E, int ticks) {
  hipStream_t st = ctx().stream;
  double tm = now_sec();
  auto lap = [&](int i) {
    const double t = now_sec();
    E.host_t[i] += t - tm;
    tm = t;
  };
  auto F = std::make_unique<Inflight>();
  std::vector<Slot *> &L = F->L, &Cg = F->Cg, &W = F->W;
  for (auto &sp : E.active) {
    if (sp->phase == P_LANCZOS) L.push_back(sp.get());
    else if (sp->phase == P_CG) Cg.push_back(sp.get());
    else if (sp->phase == P_WAIT && !sp->wait_gathered) W.push_back(sp.get());
  }
  if (L.empty() && Cg.empty() && W.empty()) return nullptr;
  E.chunks++;
  std::vector<int> lan_budget(L.size());
  for (size_t i = 0; i < L.size(); i++) lan_budget[i] = lan_budget_of(*L[i], ticks);
  // lanes bins (each bin one SpMV launch per tick)
  std::map<int, std::vector<int>> binL, binC;
  for (size_t i = 0; i < L.size(); i++) binL[L[i]->lanes].push_back((int)i);
  for (size_t i = 0; i < Cg.size(); i++) binC[Cg[i]->lanes].push_back((int)i);
  // Build every launch of the chunk: descriptors + block tables, one upload.
  struct Launch {
    int kind;  // 0 spmv-A, 1 lanczos_b, 2 spmv-cg, 3 cg_update, 4 cg_dir
    int G, nblk, nseg;
    size_t desc_off, blk_off;
  };
  std::vector<Launch> launches;
  std::vector<SegDesc> hdesc;
  std::vector<int32_t> hblk;
  auto add_launch = [&](int kind, int G, const std::vector<SegDesc> &segs, const std::vector<int> &nb) {
    Launch l{kind, G, 0, (int)segs.size(), hdesc.size(), hblk.size()};
    int lo = 0;
    for (size_t i = 0; i < segs.size(); i++) {
      SegDesc d = segs[i];
      d.blk_lo = lo;
      d.nblk = nb[i];
      hdesc.push_back(d);
      lo += nb[i];
    }
    l.nblk = lo;
    // coarse block -> segment map, one entry per kMapStride blocks
    size_t seg = 0;
    int seg_end = segs.empty() ? 0 : nb[0];
    for (int b = 0; b < lo; b += kMapStride) {
      while (b >= seg_end) seg_end += nb[++seg];
      hblk.push_back((int32_t)seg);
    }
    if (lo > 0) launches.push_back(l);
  };
  std::vector<int> lan_k(L.size()), cg_it(Cg.size());
  for (size_t i = 0; i < L.size(); i++) lan_k[i] = L[i]->k_next;
  for (size_t i = 0; i < Cg.size(); i++) cg_it[i] = Cg[i]->cg_it;
  for (int t = 0; t < ticks; t++) {
    bool any = false;
    // Lanczos: A then B per bin
    for (auto &kv : binL) {
      std::vector<SegDesc> segA, segB;
      std::vector<int> nbA, nbB;
      for (int i : kv.second) {
        Slot &s = *L[i];
        const int k = lan_k[i];
        if (k >= s.p.diag_max_step || k - s.k_next >= lan_budget[i]) continue;
        SegDesc a{};
        a.row = s.A->row.p;
        a.col = s.A->col.p;
        a.val = s.A->val.p;
        a.val32 = s.A->val32.p;
        a.n = s.n;
        a.fp32 = s.mixed ? 1 : 0;
        a.load = slot_load(s);
        a.ent = s.A->ent.p;
        a.table = s.A->table.p;
        a.first = s.A->first.p;
        a.dcol = s.A->dcol.p;
        set_sell(a, *s.A);
        a.k = k;
        a.shift = 0.0;
        a.x = s.w.p;
        a.o1 = s.V.vec(k);
        a.o2 = s.y.p;
        a.s1 = s.nrm2();
        a.s2 = s.alpha() + k;
        segA.push_back(a);
        nbA.push_back(spmv_blocks(s.n, kv.first, s.A->grid));
        SegDesc b{};
        b.n = s.n;
        b.k = k;
        b.x = s.w.p;
        b.o1 = s.V.vec(k);
        b.o2 = s.y.p;
        b.o3 = k > 0 ? s.V.vec(k - 1) : nullptr;
        b.s1 = s.alpha() + k;
        b.s2 = s.beta() + k;
        b.s3 = s.nrm2();
        segB.push_back(b);
        nbB.push_back(vec_blocks(s.n));
        lan_k[i] = k + 1;
        any = true;
      }
      add_launch(0, kv.first, segA, nbA);
      add_launch(1, 0, segB, nbB);
    }
    for (auto &kv : binC) {
      std::vector<SegDesc> sp, su, sd;
      std::vector<int> nbp, nbv;
      for (int i : kv.second) {
        Slot &s = *Cg[i];
        const int it = cg_it[i];
        if (it >= s.p.cg_max_step) continue;
        SegDesc a{};
        a.row = s.A->row.p;
        a.col = s.A->col.p;
        a.val = s.A->val.p;
        a.n = s.n;
        a.k = it;
        a.load = slot_load(s);
        a.ent = s.A->ent.p;
        a.table = s.A->table.p;
        a.first = s.A->first.p;
        a.dcol = s.A->dcol.p;
        set_sell(a, *s.A);
        a.shift = s.p.ii_diag_add - s.eig;
        a.x = s.pv.p;
        a.o1 = s.cy.p;
        a.s1 = s.pap() + it;
        a.status = s.cgst.p;
        sp.push_back(a);
        nbp.push_back(spmv_blocks(s.n, kv.first, s.A->grid));
        SegDesc u{};
        u.n = s.n;
        u.k = it;
        u.acc = s.p.cg_acc;
        u.o1 = s.cy.p;
        u.o3 = s.pv.p;
        u.s1 = s.pap();
        u.s2 = s.rr();
        u.v1 = s.xw.p;
        u.v2 = s.r.p;
        u.status = s.cgst.p;
        if (s.mixed) {   // the iterate test of the mixed form's inner solves (CgEarly)
          u.eb = s.vec.p;
          u.eadd = s.p.ii_diag_add;
          u.eacc = kCgEarlyFraction * s.p.ii_acc;
        }
        su.push_back(u);
        SegDesc dd{};
        dd.n = s.n;
        dd.k = it;
        dd.s2 = s.rr();
        dd.v1 = s.pv.p;
        dd.v2 = s.r.p;
        dd.status = s.cgst.p;
        sd.push_back(dd);
        nbv.push_back(vec_blocks(s.n));
        cg_it[i] = it + 1;
        any = true;
      }
      add_launch(2, kv.first, sp, nbp);
      add_launch(3, 0, su, nbv);
      add_launch(4, 0, sd, nbv);
    }
    if (!any) break;
  }
  // upload: descriptors, block maps and the read-back gather list through this chunk's area
  for (size_t i = 0; i < L.size(); i++) {
    F->idL.push_back(L[i]->id);
    F->epL.push_back(L[i]->epoch);
    F->k_start.push_back(L[i]->k_next);
    F->k_end.push_back(lan_k[i]);
    L[i]->k_next = lan_k[i];
  }
  for (size_t i = 0; i < Cg.size(); i++) {
    F->idC.push_back(Cg[i]->id);
    F->epC.push_back(Cg[i]->epoch);
    F->cg_end.push_back(cg_it[i]);
    Cg[i]->cg_it = cg_it[i];
  }
  for (Slot *w : W) {
    F->idW.push_back(w->id);
    F->epW.push_back(w->epoch);
    w->wait_gathered = true;
  }
  std::vector<size_t> &off = F->off;
  off.assign(L.size() + Cg.size() + W.size() + 1, 0);
  std::vector<GatherDesc> gd;
  for (size_t i = 0; i < L.size(); i++) {
    const Slot &sl = *L[i];
    const int32_t c = F->k_end[i] - F->k_start[i];
    off[i + 1] = off[i] + 2 * (size_t)c;
    if (c == 0) continue;
    gd.push_back({sl.scal.p + F->k_start[i], (int64_t)off[i], c, 0});
    gd.push_back({sl.scal.p + sl.KS + F->k_start[i] + 1, (int64_t)(off[i] + c), c, 0});
  }
  for (size_t i = 0; i < Cg.size(); i++) {
    const size_t o = off[L.size() + i];
    off[L.size() + i + 1] = o + 1;   // {done, last iteration} in 8 bytes
    gd.push_back({reinterpret_cast<const double *>(Cg[i]->cgst.p), (int64_t)o, 1, 0});
  }
  for (size_t i = 0; i < W.size(); i++) {
    const size_t o = off[L.size() + Cg.size() + i];
    const bool rq = W[i]->wait == W_RQ || W[i]->wait == W_SHIFT || W[i]->wait == W_REFRESH;
    off[L.size() + Cg.size() + i + 1] = o + (rq ? 2 : 1);
    gd.push_back({rq ? W[i]->rq() : W[i]->resid(), (int64_t)o, rq ? 2 : 1, 0});
  }
  const int area = E.area;
  E.area ^= 1;
  F->area = area;
  size_t max_blocks = 0;
  for (const Launch &l : launches) max_blocks = std::max(max_blocks, (size_t)l.nblk);
  ensure_scratch(E, hdesc.size(), hblk.size(), max_blocks);
  if (E.dgather.count < gd.size()) E.dgather.resize(std::max(gd.size(), 2 * E.dgather.count));
  const size_t words = std::max<size_t>(off.back(), 1);
  if (E.dstage.count < words) E.dstage.resize(std::max(words, 2 * E.dstage.count));
  const size_t bd = sizeof(SegDesc) * hdesc.size(), bm = sizeof(int32_t) * hblk.size();
  const size_t bg = sizeof(GatherDesc) * gd.size();
  char *pin = static_cast<char *>(grow_pinned(E.up_pin[area], E.up_bytes[area], bd + bm + bg + 64));
  const size_t og = (bd + bm + 15) / 16 * 16;   // GatherDesc holds 8-byte fields
  if (og + bg > E.up_bytes[area]) pin = static_cast<char *>(grow_pinned(E.up_pin[area], E.up_bytes[area], og + bg));
  std::memcpy(pin, hdesc.data(), bd);
  std::memcpy(pin + bd, hblk.data(), bm);
  std::memcpy(pin + og, gd.data(), bg);
  if (bd) HIPCHK(hipMemcpyAsync(E.ddesc.p, pin, bd, hipMemcpyHostToDevice, st));
  if (bm) HIPCHK(hipMemcpyAsync(E.dblk.p, pin + bd, bm, hipMemcpyHostToDevice, st));
  if (bg) HIPCHK(hipMemcpyAsync(E.dgather.p, pin + og, bg, hipMemcpyHostToDevice, st));
  lap(0);
  auto run_launches = [&]() {
    for (const Launch &l : launches) {
      const SegDesc *dd = E.ddesc.p + l.desc_off;
      const int32_t *bb = E.dblk.p + l.blk_off;
      switch (l.kind) {
        case 0:
          launch_seg_spmv<LANCZOS_A>(l.G, l.nblk, dd, bb, l.nseg, E.partial.p, st);
          k_seg_finish<FIN_LANCZOS_A><<<l.nseg, kBlock, 0, st>>>(dd, E.partial.p, nullptr, 0);
          break;
        case 1:
          k_seg_lanczos_b<<<l.nblk, kBlock, 0, st>>>(dd, bb, l.nseg, E.partial.p);
          k_seg_finish<FIN_LANCZOS_B><<<l.nseg, kBlock, 0, st>>>(dd, E.partial.p, nullptr, 0);
          break;
        case 2:
          launch_seg_spmv<CG_AP>(l.G, l.nblk, dd, bb, l.nseg, E.partial.p, st);
          k_seg_finish<FIN_CG_AP><<<l.nseg, kBlock, 0, st>>>(dd, E.partial.p, nullptr, 0);
          break;
        case 3:
          k_seg_cg_update<<<l.nblk, kBlock, 0, st>>>(dd, bb, l.nseg, E.partial.p, E.partial2.p);
          k_seg_finish<FIN_CG_UPDATE><<<l.nseg, kBlock, 0, st>>>(dd, E.partial.p, E.partial2.p, l.nblk);
          break;
        case 4: k_seg_cg_dir<<<l.nblk, kBlock, 0, st>>>(dd, bb, l.nseg); break;
      }
    }
  };
  // the launch sequence: kernels, grids, descriptor offsets and the scratch buffers they use
  std::vector<int64_t> sig;
  sig.reserve(6 * launches.size() + 3);
  for (const Launch &l : launches) {
    sig.push_back(l.kind);
    sig.push_back(l.G);
    sig.push_back(l.nblk);
    sig.push_back(l.nseg);
    sig.push_back((int64_t)l.desc_off);
    sig.push_back((int64_t)l.blk_off);
  }
  sig.push_back((int64_t)(intptr_t)E.ddesc.p);
  sig.push_back((int64_t)(intptr_t)E.dblk.p);
  sig.push_back((int64_t)(intptr_t)E.partial.p);
  sig.push_back((int64_t)(intptr_t)E.partial2.p);
  if (E.opts.use_graphs && !launches.empty() && E.graph && sig == E.graph_sig) {
    HIPCHK(hipGraphLaunch(E.graph, st));
    E.n_replays++;
  } else if (E.opts.use_graphs && !launches.empty() && sig == E.prev_sig) {
    // the same sequence twice in a row: capture it (kernel launches only) and replay it
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      run_launches();
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ie);
    if (E.graph) HIPCHK(hipGraphExecDestroy(E.graph));
    E.graph = ex;
    E.graph_sig = sig;
    HIPCHK(hipGraphLaunch(E.graph, st));
    E.n_captures++;
  } else {
    run_launches();
    E.n_eager++;
  }
  E.prev_sig.swap(sig);
  HIPCHK(hipGetLastError());
  double *h = static_cast<double *>(grow_pinned(E.rb_pin[area], E.rb_bytes[area], sizeof(double) * words));
  if (!gd.empty()) {
    k_gather_words<<<(int)gd.size(), kBlock, 0, st>>>(E.dgather.p, E.dstage.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, E.dstage.p, sizeof(double) * off.back(), hipMemcpyDeviceToHost, st));
  }
  if (!E.rb_ev[area]) HIPCHK(hipEventCreateWithFlags(&E.rb_ev[area], hipEventDisableTiming));
  HIPCHK(hipEventRecord(E.rb_ev[area], st));
  lap(1);
  return F;
}

// Waits for a launched chunk's read-back and advances its slots: Lanczos checks, CG ends,
// waiting slots' residuals and Rayleigh quotients.  Slots whose phase changed since the
// launch (a later chunk already saw them converge) skip their now unused scalars.
void process_chunk(dmrgx_engine &E, const Inflight &F) {
  double tm = now_sec();
  auto lap = [&](int i) {
    const double t = now_sec();
    E.host_t[i] += t - tm;
    tm = t;
  };
  HIPCHK(hipEventSynchronize(E.rb_ev[F.area]));
  lap(2);
  const double *h = static_cast<const double *>(E.rb_pin[F.area]);
  const std::vector<double> stage(h, h + F.off.back());   // the phase changes below launch work
  // slots are looked up by id among the active ones: a slot that finished after the launch
  // may have been retired and polled since
  std::unordered_map<int64_t, Slot *> live;
  for (auto &sp : E.active) live[sp->id] = sp.get();
  auto find = [&](int64_t id) -> Slot * {
    auto it = live.find(id);
    return it == live.end() ? nullptr : it->second;
  };
  const size_t nl = F.idL.size(), nc = F.idC.size();
  for (size_t i = 0; i < nl; i++) {
    Slot *sp = find(F.idL[i]);
    if (!sp) continue;
    Slot &sl = *sp;
    if (sl.phase != P_LANCZOS || sl.epoch != F.epL[i] || sl.k_read != F.k_start[i]) continue;
    const size_t c = (size_t)(F.k_end[i] - F.k_start[i]);
    lanczos_check(sl, F.k_end[i], stage.data() + F.off[i], stage.data() + F.off[i] + c);
  }
  lap(3);
  for (size_t i = 0; i < nc; i++) {
    Slot *sp = find(F.idC[i]);
    if (!sp) continue;
    Slot &sl = *sp;
    if (sl.phase != P_CG || sl.epoch != F.epC[i]) continue;
    int hs[2];
    std::memcpy(hs, stage.data() + F.off[nl + i], sizeof(hs));
    if (hs[0]) cg_finished(sl, hs[1]);
    else if (F.cg_end[i] >= sl.p.cg_max_step) cg_finished(sl, -1);
  }
  for (size_t i = 0; i < F.idW.size(); i++) {
    Slot *sp = find(F.idW[i]);
    if (!sp) continue;
    Slot &sl = *sp;
    sl.wait_gathered = false;
    if (sl.phase != P_WAIT || sl.epoch != F.epW[i]) continue;
    wait_resolved(sl, stage.data() + F.off[nl + nc + i]);
  }
  lap(4);
}

void admit(dmrgx_engine &E) {
  while (!E.pending.empty()) {
    Slot &s = *E.pending.front();
    if ((int)E.active.size() >= E.opts.max_active) break;
    if (!E.active.empty() && E.bytes_in_use + s.bytes > (size_t)E.opts.hbm_budget_bytes) break;
    E.bytes_in_use += s.bytes;
    auto sp = std::move(E.pending.front());
    E.pending.pop_front();
    if (sp->n == 1) {
      // 1x1: the eigenpair is the element itself (as dmrgx_ground_state)
      sp->vec.resize(1);
      double h;
      HIPCHK(hipStreamSynchronize(ctx().stream));   // the library's streams do not block on the null stream
      HIPCHK(hipMemcpy(&h, sp->A->val.p, sizeof(double), hipMemcpyDeviceToHost));
      const double one = 1.0;
      HIPCHK(hipMemcpy(sp->vec.p, &one, sizeof(double), hipMemcpyHostToDevice));
      sp->eig = sp->out.eig = h;
      sp->out.lan_converged = true;
      sp->t_start = now_sec();
      finish_slot(*sp, DMRGX_OK);
    } else {
      start_slot(*sp);
    }
    E.active.push_back(std::move(sp));
  }
}

void retire(dmrgx_engine &E) {
  for (size_t i = 0; i < E.active.size();) {
    if (E.active[i]->phase == P_DONE) {
      E.bytes_in_use -= std::min(E.bytes_in_use, E.active[i]->bytes);
      E.done.push_back(std::move(E.active[i]));
      E.active.erase(E.active.begin() + i);
    } else {
      i++;
    }
  }
}

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

extern "C" {

void dmrgx_engine_default_opts(dmrgx_engine_opts *o) {
  memset(o, 0, sizeof(*o));
  o->hbm_budget_bytes = (int64_t)200 << 30;   // of the 288 GB of one MI355X
  o->max_active = 1024;
  o->chunk_ticks = 16;
  o->basis_steps_estimate = 200;
  o->use_graphs = 1;
}

int dmrgx_engine_create(const dmrgx_engine_opts *o, dmrgx_engine **out) {
  if (!out) return DMRGX_ERR_ARG;
  return run_status([&] {
    ensure_init();
    auto E = std::make_unique<dmrgx_engine>();
    if (o) E->opts = *o;
    else dmrgx_engine_default_opts(&E->opts);
    if (E->opts.max_active <= 0 || E->opts.chunk_ticks <= 0 || E->opts.hbm_budget_bytes <= 0)
      throw Failure(DMRGX_ERR_ARG, "engine options must be positive");
    *out = E.release();
  });
}

int dmrgx_engine_submit(dmrgx_engine *E, dmrgx_matrix *A, int take_ownership, const dmrgx_solve_params *p,
                        uint64_t tag, int64_t *request_id) {
  if (!E || !A || !p) return DMRGX_ERR_ARG;
  if (p->diag_max_step <= p->diag_min_step || (p->precision != 0 && p->precision != 1) || p->guess ||
      p->lan_type < 0 || p->lan_type > 1)
    return DMRGX_ERR_ARG;
  return run_status([&] {
    auto s = std::make_unique<Slot>();
    s->id = E->next_id++;
    s->tag = tag;
    s->A = A;
    s->owned = take_ownership ? A : nullptr;
    s->p = *p;
    s->mixed = p->precision == 1;
    s->n = A->n;
    s->lanes = A->lanes;
    s->KS = p->diag_max_step + 2;
    s->CS = std::max(p->cg_max_step, 1) + 1;
    s->t_submit = now_sec();
    s->bytes = slot_bytes(*E, *s);
    if (request_id) *request_id = s->id;
    E->pending.push_back(std::move(s));
  });
}

int dmrgx_engine_run(dmrgx_engine *E, int32_t max_chunks) {
  if (!E) return DMRGX_ERR_ARG;
  int rc = run_status([&] {
    // chunk c + 1 is launched before chunk c's read-back is processed, so the host's checks
    // and phase changes overlap the device's next chunk
    for (int c = 0; c < max_chunks; c++) {
      const double t0 = now_sec();
      admit(*E);
      E->host_t[5] += now_sec() - t0;
      retire(*E);   // requests finished at admission (1 x 1) leave the active set here
      if (E->active.empty() && !E->fl) break;
      std::unique_ptr<Inflight> next = launch_chunk(*E, E->opts.chunk_ticks);
      if (!next && !E->fl) break;
      if (E->fl) process_chunk(*E, *E->fl);
      E->fl = std::move(next);
      retire(*E);
    }
    // the last chunk stays in flight; the next call (or the engine's end) processes it
    admit(*E);
    retire(*E);
  });
  return rc;
}

int dmrgx_engine_host_times(const dmrgx_engine *E, double *seconds6, int64_t *chunks) {
  if (!E) return DMRGX_ERR_ARG;
  if (seconds6)
    for (int i = 0; i < 6; i++) seconds6[i] = E->host_t[i];
  if (chunks) *chunks = E->chunks;
  return DMRGX_OK;
}

int dmrgx_engine_graph_stats(const dmrgx_engine *E, int64_t *eager, int64_t *captures, int64_t *replays) {
  if (!E) return DMRGX_ERR_ARG;
  if (eager) *eager = E->n_eager;
  if (captures) *captures = E->n_captures;
  if (replays) *replays = E->n_replays;
  return DMRGX_OK;
}

int32_t dmrgx_engine_pending(const dmrgx_engine *E) {
  return E ? (int32_t)(E->pending.size() + E->active.size()) : -1;
}

int32_t dmrgx_engine_done(const dmrgx_engine *E) { return E ? (int32_t)E->done.size() : -1; }

int dmrgx_engine_peek(const dmrgx_engine *E, uint64_t *tag, int32_t *n) {
  if (!E) return DMRGX_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  if (E->done.empty()) return 0;
  if (tag) *tag = E->done.front()->tag;
  if (n) *n = E->done.front()->n;
  return 1;
}

int dmrgx_engine_poll(dmrgx_engine *E, dmrgx_engine_result *res, double *vec_out) {
  if (!E || !res) return DMRGX_ERR_ARG;
  int got = 0;
  int rc = run_status([&] {
    if (E->done.empty()) return;
    auto s = std::move(E->done.front());
    E->done.pop_front();
    memset(res, 0, sizeof(*res));
    res->request_id = s->id;
    res->tag = s->tag;
    res->status = s->status;
    res->n = s->n;
    res->eig_val = s->out.eig;
    res->gs_error = s->out.err;
    res->lanczos_steps = s->out.lan_steps;
    res->ii_steps = s->out.ii_steps;
    res->cg_steps = s->out.cg_steps;
    res->spmv_calls = s->out.spmv;
    res->t_submit = s->t_submit;
    res->t_start = s->t_start;
    res->t_done = s->t_done;
    if (vec_out && s->vec.p) {
      ctx().copy_out(vec_out, s->vec.p, sizeof(double) * s->n);
      HIPCHK(hipStreamSynchronize(ctx().stream));
    }
    if (s->owned) dmrgx_matrix_free(s->owned);
    got = 1;
  });
  return rc != DMRGX_OK ? rc : got;
}

void dmrgx_engine_free(dmrgx_engine *E) {
  if (!E) return;
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  for (auto *q : {&E->pending, &E->done})
    for (auto &s : *q)
      if (s->owned) dmrgx_matrix_free(s->owned);
  for (auto &s : E->active)
    if (s->owned) dmrgx_matrix_free(s->owned);
  delete E;
}

int dmrgx_ground_state_batch(int32_t count, dmrgx_matrix *const *A, const dmrgx_solve_params *p,
                             double *const *vec_out, dmrgx_solve_stats *stats) {
  if (count < 0 || (count && (!A || !p))) return DMRGX_ERR_ARG;
  dmrgx_engine *E = nullptr;
  dmrgx_engine_opts o;
  dmrgx_engine_default_opts(&o);
  int rc = dmrgx_engine_create(&o, &E);
  if (rc != DMRGX_OK) return rc;
  for (int32_t i = 0; i < count && rc == DMRGX_OK; i++) rc = dmrgx_engine_submit(E, A[i], 0, p, (uint64_t)i, nullptr);
  const double t0 = now_sec();
  while (rc == DMRGX_OK && dmrgx_engine_pending(E) > 0) rc = dmrgx_engine_run(E, 64);
  int worst = rc;
  dmrgx_engine_result r;
  uint64_t next_tag = 0;
  while (rc == DMRGX_OK && dmrgx_engine_peek(E, &next_tag, nullptr) == 1) {
    // the vector goes straight to the caller's buffer for this tag
    const int i = (int)next_tag;
    int got = dmrgx_engine_poll(E, &r, vec_out ? vec_out[i] : nullptr);
    if (got < 0) {
      rc = got;
      break;
    }
    if (r.status != DMRGX_OK && worst == DMRGX_OK) worst = r.status;
    if (stats) {
      stats[i].eig_val = r.eig_val;
      stats[i].gs_error = r.gs_error;
      stats[i].lanczos_steps = r.lanczos_steps;
      stats[i].ii_steps = r.ii_steps;
      stats[i].cg_steps = r.cg_steps;
      stats[i].spmv_calls = r.spmv_calls;
      stats[i].seconds = r.t_done - t0;
      stats[i].lanczos_fp32_steps = p && p->precision == 1 ? r.lanczos_steps : 0;
    }
  }
  dmrgx_engine_free(E);
  return rc != DMRGX_OK ? rc : worst;
}

}  // extern "C"
