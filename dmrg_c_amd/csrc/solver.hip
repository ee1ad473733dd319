// Device-resident ground-state drivers: Lanczos (stored basis, host tridiagonal checks
// every chunk of steps), inverse iteration with CG, convergence check.
//
// Algorithms follow sml/LANCZOS_SLOW_SYM.c:57-274, sml/LANCZOS_SYM.c:58-258,
// sml/INVERSE_ITERATION_SYM.c:9-166, sml/CONJUGATE_GRADIENT_SYM.c:8-123 and
// sml/CONVERGE_CHECK_SYM.c:7-44.  Differences, all without effect on the converged result:
//  * the Lanczos vectors are kept in HBM, so the eigenvector is one pass over them
//    instead of the reference's replay of the whole recurrence (LANCZOS_SLOW_SYM.c:193-261,
//    which recomputes the same vectors);
//  * alpha/beta are read back once per chunk of steps; steps past the converged one are
//    computed and discarded;
//  * the lowest Ritz value comes from Sturm bisection (tridiag.cpp) instead of dstev_;
//  * the inverse-iteration shift is applied inside the operator (as
//    dmrg/DMRG_CONJUGATE_GRADIENT_OBC.c:29,40 does) instead of DIAG_ADD_CRS1 on the matrix;
//  * below n = 1000, where the reference calls dense dsyev_ (LANCZOS_SLOW_SYM.c:38), the
//    Lanczos run stops on the residual bound beta_{k+1}|s_k| <= 1e-12 max(1,|E|) (or on an
//    invariant subspace), which gives the extremal eigenvalue to working precision.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>

#include "internal.h"

namespace dmrgx {

namespace {

// Lanczos basis stored in chunks of 32 vectors, allocated as the recurrence grows.
struct Basis {
  int64_t n;
  static constexpr int kPer = 32;
  std::vector<DevBuf<double>> chunks;
  explicit Basis(int64_t n_) : n(n_) {}
  double *vec(int k) {
    const size_t c = (size_t)k / kPer;
    while (chunks.size() <= c) chunks.emplace_back((size_t)n * kPer);
    return chunks[c].p + (size_t)(k % kPer) * n;
  }
};

void ritz(Basis &V, int k, const std::vector<double> &s, double *out, hipStream_t st) {
  std::vector<const double *> ptrs(k + 1);
  for (int j = 0; j <= k; j++) ptrs[j] = V.vec(j);
  DevBuf<const double *> dptr(k + 1);
  DevBuf<double> dcoef(k + 1);
  HIPCHK(hipMemcpyAsync(dptr.p, ptrs.data(), sizeof(double *) * (k + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dcoef.p, s.data(), sizeof(double) * (k + 1), hipMemcpyHostToDevice, st));
  launch_ritz(V.n, k, dptr.p, dcoef.p, out, st);
  HIPCHK(hipStreamSynchronize(st));  // host vectors above go out of scope
}

double read_scalar(const double *d, hipStream_t st) {
  double *h = ctx().host_stage(1);
  HIPCHK(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return h[0];
}

}  // namespace

// Steps (iterations) queued behind the chunk whose read-back the host examines: enough
// device work to cover the host's check and next launch, and all that runs past
// convergence.  A Lanczos step costs a product; a CG iteration past convergence is skipped
// on the device and costs only its launches.
constexpr int kLanczosAhead = 2;
constexpr int kCgAhead = 4;
// Longest CG chunk: |r|^2 does not fall geometrically, so long predictions overshoot, and
// every iteration launched past convergence still costs the host its three launches.
constexpr int kCgChunkMax = 16;

void lanczos(const Operator &A, const LanczosOpts &o, double *vec, SolveOut &out) {
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int64_t n = A.n;
  const bool small = o.dense_small && n < 1000;
  const int max_step = o.max_step;
  const int block = (o.max_block_step > 0 && o.max_block_step < max_step) ? o.max_block_step : max_step;
  const bool restarts = (block < max_step);

  DevBuf<double> w(n), y(n);
  DevBuf<double> sc(2 * ((size_t)block + 2) + 1);
  double *alpha = sc.p, *beta = sc.p + block + 2, *nrm2 = sc.p + 2 * (block + 2);
  Reduce R;
  Basis V(n);
  Reduce RA, RB;   // product partials, vector-update partials (alternating consumers)
  std::vector<double> ha(block + 2), hb(block + 2), s(block + 2);

  double E = 1e15;  // sml/LANCZOS_SYM.c:77
  int total_step = 0;
  launch_copy(n, vec, w.p, st);
  launch_norm2(n, w.p, R, nrm2, st);

  while (true) {
    // step 0 of this block: sml/LANCZOS_SLOW_SYM.c:98-106
    launch_lanczos_a(A, o.fp32, 0, w.p, nrm2, 0.0, V.vec(0), y.p, alpha, R, st);
    launch_lanczos_b(n, 0, y.p, V.vec(0), nullptr, alpha, beta, w.p, nrm2, R, st);
    out.spmv++;
    int kconv = -1, chunk = small ? 16 : 8;
    // first and last |E_k - E_{k-1}| of the chunk (with res_tol: the residual bounds) and
    // the value the prediction aims at
    double dE_first = 0.0, dE_last = 0.0, res_target = o.acc;
    int n_dE = 0;
    double lam = E;
    RitzTrack track;  // T of this block
    bool block_done = false;
    // Chunks of steps are pipelined: the next chunk is launched before the host examines
    // the previous one's alpha/beta, so the device does not wait for the host.  Steps
    // launched past convergence compute for nothing; the chunk lengths keep them few.
    int enext = 1, etotal = total_step;   // next step to launch, steps launched
    struct Pending {
      int k0, count, slot;
    };
    Pending pend[2];
    int npend = 0, nslot = 0;
    const size_t span = 2 * ((size_t)block + 2);   // alpha | beta per read-back slot
    double *hbase = C.host_stage(2 * span);
    hipEvent_t *ev = C.step_events();
    auto room = [&] { return std::min(block - enext, max_step - etotal); };
    auto enqueue = [&](int c) {
      // two kernels per step: each sums the other's block partials itself (the scalar the
      // finish kernel would store, same order); the chunk's last step gets the finish
      // kernel so that the host reads beta_{k+1}
      int nb = 0;
      for (int k = enext; k < enext + c; k++) {
        const bool first = k == enext;
        const int na = launch_lanczos_a_part(A, o.fp32, k, w.p, nrm2, beta, first ? nullptr : RB.partial.p,
                                             first ? 0 : nb, V.vec(k), y.p, alpha, RA, st);
        nb = launch_lanczos_b_part(n, k, y.p, V.vec(k), V.vec(k - 1), alpha, beta, RA.partial.p, na, w.p, RB, st);
        if (k == enext + c - 1) launch_lanczos_b_finish(k, beta, nrm2, RB, nb, st);
      }
      double *h = hbase + nslot * span;
      HIPCHK(hipMemcpyAsync(h, alpha, sizeof(double) * (enext + c), hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(h + block + 2, beta, sizeof(double) * (enext + c + 1), hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(ev[nslot], st));
      pend[npend++] = Pending{enext, c, nslot};
      nslot ^= 1;
      enext += c;
      etotal += c;
    };
    if (room() > 0) enqueue(std::min(chunk, room()));
    while (npend > 0) {
      // a short chunk queued behind the one the host waits for keeps the device busy
      // while the host examines it; it is all that runs past convergence
      if (npend == 1 && room() > 0) enqueue(std::min(kLanczosAhead, room()));
      const Pending q = pend[0];
      pend[0] = pend[1];
      npend--;
      HIPCHK(hipEventSynchronize(ev[q.slot]));
      const double *h = hbase + q.slot * span;
      const int next = q.k0, c = q.count;
      std::copy(h, h + next + c, ha.begin());
      std::copy(h + block + 2, h + block + 2 + next + c + 1, hb.begin());
      for (int k = next; k < next + c; k++) {
        total_step++;
        out.spmv++;
        const bool check = small || k >= o.min_step || (restarts && k == block - 1);
        if (!check) continue;
        // T_{k+1}: diag ha[0..k], off hb[1..k]
        lam = track.next(k + 1, ha.data(), hb.data() + 1);
        bool conv;
        const double bnext = hb[k + 1];
        const double tscale = std::max(1.0, std::fabs(lam));
        if (small) {
          // dense-branch equivalent: residual bound of the Ritz pair, or invariant subspace
          tridiag_vector(k + 1, ha.data(), hb.data() + 1, lam, s.data());
          conv = !(bnext > 1e-12 * tscale) || bnext * std::fabs(s[k]) <= 1e-12 * tscale;
        } else if (!(bnext > 1e-14 * tscale)) {
          conv = true;  // invariant subspace: the Ritz values are exact
        } else if (o.res_tol > 0.0) {
          tridiag_vector(k + 1, ha.data(), hb.data() + 1, lam, s.data());
          const double res = bnext * std::fabs(s[k]);
          const double target = o.res_abs ? o.res_tol : o.res_tol * tscale;
          conv = res <= target;
          if (!conv) {   // progress measure of the chunk prediction below
            if (n_dE++ == 0) dE_first = res;
            dE_last = res;
            res_target = target;
          }
        } else if (o.type == 1) {
          tridiag_vector(k + 1, ha.data(), hb.data() + 1, lam, s.data());
          conv = std::fabs(s[k]) < o.acc;     // "Tight": sml/LANCZOS_SLOW_SYM.c:140-146
        } else {
          conv = std::fabs(E - lam) < o.acc;  // "Normal": sml/LANCZOS_SLOW_SYM.c:147-153
        }
        if (conv) {
          kconv = k;
          break;
        }
        if (restarts && k == block - 1) {
          block_done = true;
          break;
        }
        if (o.res_tol <= 0.0 && E < 1e14) {   // a real previous value (E starts at 1e15)
          if (n_dE++ == 0) dE_first = std::fabs(E - lam);
          dE_last = std::fabs(E - lam);
        }
        E = lam;
      }
      if (kconv >= 0 || block_done) break;
      // Next chunk: the steps the chunk's energy changes predict ("Normal" convergence is
      // geometric) plus two, less the steps already queued behind this chunk -- or
      // doubling while no prediction holds.
      const int queued = (npend > 0) ? pend[0].count : 0;
      int nc;
      if (!small && (o.res_tol > 0.0 || o.type != 1) && n_dE >= 3 && dE_last > 0.0 && dE_last < dE_first) {
        const double ratio = std::pow(dE_last / dE_first, 1.0 / (n_dE - 1));   // mean ratio per step
        const double need = std::ceil(std::log(res_target / dE_last) / std::log(ratio)) + 2.0;
        nc = (int)std::min(32.0, need - queued);
      } else {
        chunk = std::min(chunk * 2, 32);
        nc = chunk;
      }
      n_dE = 0;
      if (nc > 0 && room() > 0) enqueue(std::min(nc, room()));
    }
    if (kconv >= 0) {
      tridiag_vector(kconv + 1, ha.data(), hb.data() + 1, lam, s.data());
      ritz(V, kconv, s, vec, st);
      launch_norm2(n, vec, R, nrm2, st);
      launch_scale_by_rnorm(n, nrm2, vec, vec, st);
      HIPCHK(hipStreamSynchronize(st));
      out.eig = lam;
      out.lan_steps = total_step;
      out.lan_converged = true;
      return;
    }
    if (!restarts || total_step >= max_step) break;
    // restart from the Ritz vector of this block (sml/LANCZOS_SYM.c:249-256)
    const int kb = block - 1;
    tridiag_vector(kb + 1, ha.data(), hb.data() + 1, lam, s.data());
    ritz(V, kb, s, w.p, st);
    launch_norm2(n, w.p, R, nrm2, st);
  }
  out.lan_steps = total_step;
  out.lan_converged = false;
}

double converge_check(const Operator &A, const double *v, double eig) {
  Context &C = ctx();
  Reduce R;
  DevBuf<double> res(1);
  launch_residual_l1(A, v, eig, res.p, R, C.stream);
  return read_scalar(res.p, C.stream);
}

double rayleigh_quotient(const Operator &A, const double *v) {
  Context &C = ctx();
  Reduce R;
  DevBuf<double> res(2);
  launch_rayleigh(A, v, res.p, R, C.stream);
  launch_norm2(A.n, v, R, res.p + 1, C.stream);
  double *h = C.host_stage(2);
  HIPCHK(hipMemcpyAsync(h, res.p, 2 * sizeof(double), hipMemcpyDeviceToHost, C.stream));
  HIPCHK(hipStreamSynchronize(C.stream));
  return h[0] / h[1];
}

void ensure_fp32(const dmrgx_matrix &A) {
  if (A.val32.p) return;
  // the FP32 values are a cache of the FP64 ones; the matrix itself is not changed
  dmrgx_matrix &M = const_cast<dmrgx_matrix &>(A);
  M.val32.resize(std::max<int64_t>(A.nnz, 1));
  launch_to_fp32(M, ctx().stream);
  HIPCHK(hipStreamSynchronize(ctx().stream));
}

bool ensure_dict(const dmrgx_matrix &A) {
  if (A.n_values < 0) {
    // the dictionary is a cache of the FP64 values; the matrix itself is not changed
    dmrgx_matrix &M = const_cast<dmrgx_matrix &>(A);
    M.n_values = build_value_dict(M, ctx().stream);
    if (M.n_values == 0) {
      M.ent.reset();
      M.table.reset();
    }
  }
  return A.n_values > 0;
}

bool ensure_delta(const dmrgx_matrix &A) {
  if (A.delta < 0) {
    // like the dictionary, a cache derived from the columns; the matrix is not changed
    dmrgx_matrix &M = const_cast<dmrgx_matrix &>(A);
    M.delta = build_delta(M, ctx().stream) ? 1 : 0;
    if (!M.delta) {
      M.first.reset();
      M.dcol.reset();
    }
  }
  return A.delta > 0;
}

bool ensure_sell(const dmrgx_matrix &A, int P, bool fp32) {
  if (P != 1 && P != 2 && P != 4) return false;
  const int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  if (fp32) ensure_fp32(A);
  if (!A.sell_ptr[lp].p || (fp32 && !A.sell_val32[lp].p)) {
    // a cache derived from the CSR arrays; the matrix itself is not changed
    build_sell(const_cast<dmrgx_matrix &>(A), P, fp32, ctx().stream);
  }
  return true;
}

int64_t ensure_sell16(const dmrgx_matrix &A, int P) {
  const int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  if (A.sell_nexc[lp] < 0) build_sell16(const_cast<dmrgx_matrix &>(A), P, ctx().stream);
  return A.sell_nexc[lp];
}

bool ensure_sellx(const dmrgx_matrix &A, int P, bool fp32) {
  const int lp = P == 1 ? 0 : (P == 2 ? 1 : 2);
  if (A.sell_idx_ok[lp] == 0) return false;
  dmrgx_matrix &M = const_cast<dmrgx_matrix &>(A);
  if (A.sell_idx_ok[lp] < 0) {
    if (!ensure_dict(A) || A.n_values > kSellIdxMaxValues) {
      M.sell_idx_ok[lp] = 0;
      return false;
    }
    build_sellx(M, P, false, ctx().stream);
    M.sell_idx_ok[lp] = 1;
  }
  if (fp32 && !A.table32.p) build_sellx(M, P, true, ctx().stream);
  return true;
}

void drop_sell_except(const dmrgx_matrix &A, int P) {
  dmrgx_matrix &M = const_cast<dmrgx_matrix &>(A);
  for (int lp = 0; lp < 3; lp++)
    if ((1 << lp) != P) {
      M.sell_ptr[lp].reset();
      M.sell_col[lp].reset();
      M.sell_val[lp].reset();
      M.sell_val32[lp].reset();
      M.sell_d16[lp].reset();
      M.sell_first[lp].reset();
      M.sell_eptr[lp].reset();
      M.sell_exc[lp].reset();
      M.sell_nexc[lp] = -1;
      M.sell_idx16[lp].reset();
      M.sell_idx_ok[lp] = -1;
    }
  if (P != 1 && P != 2 && P != 4) M.sell_len.reset();
}

int effective_load(const dmrgx_matrix &A) {
  const int fmt = load_format(A.load);
  const bool ok = fmt == kFmtCsr || (fmt == kFmtDict && A.n_values > 0) || (fmt == kFmtDelta && A.delta > 0) ||
                  (sliced_format(fmt) && A.lanes <= kSellMaxLanes);
  // paired-entry loads are a choice of the CSR format: its fallback for another format drops them
  return ok ? A.load : (A.load & ~(kLoadFmtMask | kLoadPair));
}

int cg_solve(const Operator &A, double shift, const double *b, double *x, double acc, int max_step,
             bool *converged, SolveOut &out, const CgEarly *early) {
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int64_t n = A.n;
  DevBuf<double> r(n), p(n), y(n), xw(n);
  Reduce RE;   // the iterate test's two block-partial arrays
  CgEarly ce;
  if (early) {
    RE.ensure(2 * (size_t)reduce_grid(n));
    ce = CgEarly{b, early->diag_add, early->acc, RE.partial.p};
  }
  const int slots = std::max(max_step, 1) + 1;
  DevBuf<double> sc(2 * (size_t)slots + 1);
  double *rr = sc.p, *pap = sc.p + slots, *nrm2 = sc.p + 2 * slots;
  DevBuf<int> status(2);
  HIPCHK(hipMemsetAsync(status.p, 0, 2 * sizeof(int), st));
  Reduce R, RA, RB;
  launch_norm2(n, x, R, nrm2, st);
  launch_cg_init(A, shift, nrm2, x, xw.p, b, r.p, p.p, rr, R, st);
  out.spmv++;
  // Chunks of iterations are pipelined: the next chunk is launched before the previous
  // one's read-back ({done, last iteration} and the last |r|^2) is examined, so the device
  // does not wait for the host.  Iterations launched past convergence are skipped on the
  // device (every kernel checks the flag).
  int it = 1, done_it = -1;
  double *hst = C.host_stage(8);
  hipEvent_t *ev = C.step_events();
  struct Pending {
    int first, count;
    double *h;   // {done, last iteration} | |r|^2 at the chunk's end | |r|^2 before it
    int slot;
  };
  Pending pend[2];
  int npend = 0, nslot = 0;
  auto enqueue = [&](int c) {
    Pending q{it, c, hst + 4 * nslot, nslot};
    nslot ^= 1;
    // three kernels per iteration, each summing the previous one's block partials itself
    // (pAp in the update, |r|^2 and the convergence flag in the direction kernel)
    for (int i = it; i < it + c; i++) {
      const int na = launch_cg_ap_part(A, shift, p.p, y.p, status.p, RA, st);
      const int nb =
          launch_cg_update_part(n, i, p.p, y.p, xw.p, r.p, pap, rr, status.p, RA.partial.p, na, RB, st, ce);
      launch_cg_dir_part(n, i, acc, r.p, rr, p.p, status.p, RB.partial.p, nb, st, ce);
    }
    HIPCHK(hipMemcpyAsync(q.h, status.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(q.h + 1, rr + it + c - 1, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(q.h + 2, rr + it - 1, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ev[q.slot], st));
    pend[npend++] = q;
    it += c;
  };
  // iterations the |r|^2 decrease over a finished chunk predicts are still needed, plus two
  // (launches past convergence cost little, a chunk that stops short costs a round trip)
  auto predict = [&](const Pending &q) {
    const double rr_last = q.h[1], rr_prev = q.h[2];
    const double ratio = (rr_prev > 0.0 && rr_last > 0.0) ? std::pow(rr_last / rr_prev, 1.0 / q.count) : 1.0;
    if (ratio > 0.0 && ratio < 0.999 && rr_last > acc)
      return (int)std::ceil(std::log(acc / rr_last) / std::log(ratio)) + 2;
    return -1;   // no prediction: the chunk length doubles
  };
  int chunk = 4;
  if (it < max_step) enqueue(std::min(chunk, max_step - it));
  while (npend > 0) {
    // a short chunk queued behind the one the host waits for keeps the device busy while
    // the host examines it (at most two are queued)
    if (npend == 1 && it < max_step) enqueue(std::min(kCgAhead, max_step - it));
    const Pending q = pend[0];
    pend[0] = pend[1];
    npend--;
    HIPCHK(hipEventSynchronize(ev[q.slot]));
    const int *hs = reinterpret_cast<const int *>(q.h);
    if (hs[0]) {
      done_it = hs[1];
      break;
    }
    const int need = predict(q);
    const int queued = (npend > 0) ? pend[0].count : 0;
    int nc;
    if (need > 0) {
      nc = std::min(kCgChunkMax, need - queued);
    } else {
      chunk = std::min(chunk * 2, kCgChunkMax);
      nc = chunk;
    }
    if (nc > 0 && it < max_step) enqueue(std::min(nc, max_step - it));
  }
  launch_copy(n, xw.p, x, st);
  HIPCHK(hipStreamSynchronize(st));
  *converged = done_it >= 0;
  const int iters = (done_it >= 0) ? done_it : max_step;
  out.spmv += (done_it >= 0) ? done_it : std::max(max_step - 1, 0);
  return iters;
}

// sml/MINIMUM_RESIDUAL.c:7-150: MINRES through the Lanczos recurrence of H + shift with
// Givens rotations; the Lanczos step is the library's fused pair of kernels, the rotation
// scalars stay on the device and the host reads the convergence flag once per chunk.
int mr_solve(const Operator &A, double shift, const double *b, double *x, double acc, int max_step,
             bool *converged, SolveOut &out) {
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int64_t n = A.n;
  const int slots = std::max(max_step, 1) + 1;
  DevBuf<double> vraw(n), t(n), vn0(n), vn1(n), wbuf(3 * (size_t)n);
  DevBuf<double> sc(2 * (size_t)slots + 1 + 5), coef(5 * (size_t)slots);
  double *alpha = sc.p, *beta = sc.p + slots, *nrm2 = sc.p + 2 * slots, *mr = nrm2 + 1;
  DevBuf<int> status(2);
  HIPCHK(hipMemsetAsync(wbuf.p, 0, sizeof(double) * 3 * (size_t)n, st));  // W0 = W1 = 0 (:20-21)
  Reduce R;
  launch_mr_init(A, shift, b, x, vraw.p, nrm2, beta, mr, status.p, R, st);
  out.spmv++;
  double *vprev = vn0.p, *vcur = vn1.p;
  double *W[3] = {wbuf.p, wbuf.p + n, wbuf.p + 2 * n};
  int it = 0, chunk = 4, done_it = -1;
  int *hs = reinterpret_cast<int *>(C.host_stage(1));
  while (it < max_step) {
    const int c = std::min(chunk, max_step - it);
    for (int k = it; k < it + c; k++) {
      // V1 = V1 / gamma0, T = (H + shift) V1, delta = <T, V1>, V2 = T - delta V1 - gamma0 V0 (:46-62)
      launch_lanczos_a(A, false, k, vraw.p, nrm2, shift, vcur, t.p, alpha, R, st);
      launch_lanczos_b(n, k, t.p, vcur, k > 0 ? vprev : nullptr, alpha, beta, vraw.p, nrm2, R, st);
      // w_j lives in W[j % 3]: W0 = w_{k-2}, W1 = w_{k-1}, W2 = w_k (zero before they exist)
      launch_mr_step(n, k, acc, alpha, beta, mr, status.p, coef.p, vcur, W[(k + 1) % 3], W[(k + 2) % 3],
                     W[k % 3], x, st);
      std::swap(vprev, vcur);
    }
    HIPCHK(hipMemcpyAsync(hs, status.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (hs[0]) {
      done_it = hs[1];
      break;
    }
    it += c;
    chunk = std::min(chunk * 2, 64);
  }
  HIPCHK(hipStreamSynchronize(st));
  *converged = done_it >= 0;
  const int iters = (done_it >= 0) ? done_it : max_step;
  out.spmv += (done_it >= 0) ? done_it + 1 : max_step;
  return iters;
}

void inverse_iteration(const Operator &A, double eig, double *vec, double ii_acc, double ii_diag_add,
                       int ii_max_step, double cg_acc, int cg_max_step, SolveOut &out, bool mr, bool rq_shift) {
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int64_t n = A.n;
  double shift = ii_diag_add - eig;  // sml/INVERSE_ITERATION_SYM.c:24
  DevBuf<double> un(n), nrm2(1);
  Reduce R;
  launch_copy(n, vec, un.p, st);
  double err = 1000.0;
  int iter;
  out.ii_converged = false;
  for (iter = 0; iter <= ii_max_step; iter++) {
    err = converge_check(A, vec, eig);
    out.spmv++;
    if (err < ii_acc) {
      out.ii_converged = true;
      break;
    }
    bool conv = false;
    CgEarly early;
    early.diag_add = ii_diag_add;
    early.acc = kCgEarlyFraction * ii_acc;
    out.cg_steps += mr ? mr_solve(A, shift, vec, un.p, cg_acc, cg_max_step, &conv, out)
                       : cg_solve(A, shift, vec, un.p, cg_acc, cg_max_step, &conv, out, rq_shift ? &early : nullptr);
    if (!conv) out.cg_all_converged = false;
    launch_norm2(n, un.p, R, nrm2.p, st);
    launch_scale_by_rnorm(n, nrm2.p, un.p, un.p, st);
    launch_copy(n, un.p, vec, st);
    if (rq_shift && iter < ii_max_step) {
      // the next residual test and shift at the refined vector's FP64 Rayleigh quotient
      eig = rayleigh_quotient(A, vec);
      out.spmv++;
      shift = ii_diag_add - eig;
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  out.err = err;
  out.ii_steps = iter;
}

}  // namespace dmrgx
