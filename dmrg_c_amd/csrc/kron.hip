// Kronecker form of the superblock Hamiltonian: the operator of the assembled CRS
// (GET_HAM_LLLRRRRL from the recorded term list) applied without assembling it.
//
// The superblock basis (GET_BASIS_SUPERBLOCK: states ordered by the LLLR quantum numbers,
// then LL, LR, RR, RL) falls into sectors of one LLLR quantum number, and a sector holds
// every (LL, LR) state of that number combined with every (RR, RL) state of the
// complementary one, LLLR-major.  A sector's part of a vector is therefore a dense row-major
// da x db matrix X (da LLLR states, db RRRL states).  A term of the list that acts only on
// LL and LR (kinds LL, LR, LLLR) keeps the sector and acts on the rows of X; one that acts
// only on RR and RL (RR, RL, RRRL) acts on its columns.  Summed per sector they are dense
// matrices M_L (da x da) and M_R (db x db), built once from the term list with the element
// makers' rules (operator rows of the row state, sign rules of the row's parities,
// contributions with |value| <= 1e-15 or outside the superblock basis dropped), and
//     (H x)_sector = M_L X + X M_R^T + (coupling terms)
// where the coupling terms (LLRL, LLRR, LRRL, LRRR: one operator on each half) are applied
// row by row from the term list through the superblock inverse, as the matrix-free product of
// dmrg/DMRG_MATRIX_VECTOR_PRODUCT_OBC.c:18-83 applies its M_LRRL.
//
// A product is two kernels: the coupling terms write z row by row, then one LDS-tiled FP64
// kernel whose workgroups each take a 64 x 64 tile of M_L X (written to zl) or of X M_R^T
// (added to z); the solvers' epilogues read z + zl.  The sums are in an order of
// their own (the CRS sums every term's contributions to an entry first), so results agree
// with the CRS product to rounding.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "superblock.h"

#pragma clang fp contract(off)

namespace dmrgx {
namespace {


struct KronBuild {
  int32_t d, nstate_a, nstate_b;
  const int16_t *aLL, *aLR, *bRR, *bRL;   // flat state lists (sector order)
  const int32_t *asec, *bsec;             // sector of each flat state
  const int32_t *aidx, *bidx;             // (LL*d+LR) -> flat a-state, (RR*d+RL) -> flat b-state, -1 if none
  const KronSector *sec;
};

// One thread per LLLR state: its row of M_L, from the terms acting on LL and LR only, in
// term order (op1 row outer, op2 row inner, as DMRG_MAKE_ELEM_LLLR_LLLRRRRL.c:54-79).
__global__ __launch_bounds__(256) void k_kron_left(KronBuild kb, const DTerm *terms, const int32_t *ids, int nt,
                                                   const DOp *ops, const int32_t *const *eles, double *ML) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kb.nstate_a) return;
  const int s = kb.asec[i];
  const KronSector S = kb.sec[s];
  const int32_t ap = i - S.a0;
  double *row = ML + S.ml + (int64_t)ap * S.da;
  const int q[4] = {kb.aLL[i], kb.aLR[i], 0, 0};
  for (int u = 0; u < nt; u++) {
    const DTerm t = terms[ids[u]];
    const DOp a = ops[t.op1];
    const double sc = (double)term_sign(t, eles, q) * t.coeff;
    const int64_t b1 = a.row[q[t.s1]], e1 = a.row[q[t.s1] + 1];
    for (int64_t i1 = b1; i1 < e1; i1++) {
      int qq[2] = {q[0], q[1]};
      qq[t.s1] = a.col[i1];
      const double p1 = sc * a.val[i1];
      if (t.s2 < 0) {
        const int32_t c = kb.aidx[qq[0] * kb.d + qq[1]];
        if (c >= 0 && kb.asec[c] == s && fabs(p1) > 1.0e-15) row[c - S.a0] += p1;
        continue;
      }
      const DOp b = ops[t.op2];
      const int64_t b2 = b.row[q[t.s2]], e2 = b.row[q[t.s2] + 1];
      for (int64_t i2 = b2; i2 < e2; i2++) {
        qq[t.s2] = b.col[i2];
        const double p = p1 * b.val[i2];
        const int32_t c = kb.aidx[qq[0] * kb.d + qq[1]];
        if (c >= 0 && kb.asec[c] == s && fabs(p) > 1.0e-15) row[c - S.a0] += p;
      }
    }
  }
}

// One thread per RRRL state: its row of M_R, from the terms acting on RR and RL only.
__global__ __launch_bounds__(256) void k_kron_right(KronBuild kb, const DTerm *terms, const int32_t *ids, int nt,
                                                    const DOp *ops, const int32_t *const *eles, double *MR) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kb.nstate_b) return;
  const int s = kb.bsec[i];
  const KronSector S = kb.sec[s];
  const int32_t bp = i - S.b0;
  double *row = MR + S.mr + (int64_t)bp * S.db;
  const int q[4] = {0, 0, kb.bRR[i], kb.bRL[i]};
  for (int u = 0; u < nt; u++) {
    const DTerm t = terms[ids[u]];
    const DOp a = ops[t.op1];
    const double sc = (double)term_sign(t, eles, q) * t.coeff;
    const int64_t b1 = a.row[q[t.s1]], e1 = a.row[q[t.s1] + 1];
    for (int64_t i1 = b1; i1 < e1; i1++) {
      int qq[4] = {q[0], q[1], q[2], q[3]};
      qq[t.s1] = a.col[i1];
      const double p1 = sc * a.val[i1];
      if (t.s2 < 0) {
        const int32_t c = kb.bidx[qq[2] * kb.d + qq[3]];
        if (c >= 0 && kb.bsec[c] == s && fabs(p1) > 1.0e-15) row[c - S.b0] += p1;
        continue;
      }
      const DOp b = ops[t.op2];
      const int64_t b2 = b.row[q[t.s2]], e2 = b.row[q[t.s2] + 1];
      for (int64_t i2 = b2; i2 < e2; i2++) {
        qq[t.s2] = b.col[i2];
        const double p = p1 * b.val[i2];
        const int32_t c = kb.bidx[qq[2] * kb.d + qq[3]];
        if (c >= 0 && kb.bsec[c] == s && fabs(p) > 1.0e-15) row[c - S.b0] += p;
      }
    }
  }
}

// z = coupling terms applied to x.  A coupling term has one operator on each half (op1 on
// LL or LR, op2 on RR or RL) and a sign that is a product of a factor of the row's LLLR state
// and one of its RRRL state, so per term the builder lists, for every LLLR state a', the
// states a it reaches with op1 and their coefficients (left parity sign x op1 value), and
// likewise for every RRRL state b' (right parity sign x op2 value); the term's own factor is
// its coefficient times the sign rule's constant.  One workgroup per LLLR state a' of the
// output, its threads over the sector's RRRL states b': for each term (in term order), op1
// entry and op2 entry (in row order, as the element makers enumerate them) the contribution
// c_t c_l c_r x[a, b] is added when (a, b) is a superblock state (a and b in one sector) and
// |c_t c_l c_r| > 1e-15, as the element makers drop it.
struct KronCross {
  int nt, na, nb;
  const double *ct;                         // per term
  const int32_t *lptr, *lidx, *rptr, *ridx;  // term t: lptr[t (na + 1) + a'] ... into lidx / lval
  const double *lval, *rval;
  const int32_t *asec, *bsec;
  const KronSector *sec;
};
__device__ __forceinline__ void kron_cross_row(const KronCross &kc, int32_t i, const double *__restrict__ x,
                                               double *__restrict__ z) {
  const KronSector S = kc.sec[kc.asec[i]];   // i: flat LLLR state of the output rows
  double *zr = z + S.off + (int64_t)(i - S.a0) * S.db;
  for (int32_t bl = threadIdx.x; bl < S.db; bl += 256) {
    const int32_t bf = S.b0 + bl;
    double acc = 0.0;
    for (int t = 0; t < kc.nt; t++) {
      const double c = kc.ct[t];
      const int32_t *lp = kc.lptr + (int64_t)t * (kc.na + 1), *rp = kc.rptr + (int64_t)t * (kc.nb + 1);
      for (int32_t e = lp[i]; e < lp[i + 1]; e++) {
        const int32_t a = kc.lidx[e];
        const int sa = kc.asec[a];
        const double pl = c * kc.lval[e];
        for (int32_t f = rp[bf]; f < rp[bf + 1]; f++) {
          const int32_t b = kc.ridx[f];
          if (kc.bsec[b] != sa) continue;
          const double p = pl * kc.rval[f];
          if (!(fabs(p) > 1.0e-15)) continue;
          const KronSector T = kc.sec[sa];
          acc += p * x[T.off + (int64_t)(a - T.a0) * T.db + (b - T.b0)];
        }
      }
    }
    zr[bl] = acc;
  }
}

__global__ __launch_bounds__(256) void k_kron_cross(KronCross kc, const double *__restrict__ x, double *__restrict__ z) {
  kron_cross_row(kc, blockIdx.x, x, z);
}

// One TM x TN tile of one of a sector's two dense products: the left one (M_L X, k over
// the sector's LLLR states, written to zl) or the right one (X M_R^T, k over its RRRL
// states, written to z).  k-slabs of kKt staged in LDS, double
// buffered: the next slab's global loads are issued before the current slab's products
// (kKt 16, or 32 for a longer product per slab behind the loads).
// Outside the sector the operands are zero.  MFMA: the four wavefronts split the tile 2 x 2,
// each computing its (TM/2) x (TN/2) part as blocks of v_mfma_f64_16x16x4 (A fragment: lane
// l holds A[l % 16][l / 16]; B fragment: B[l / 16][l % 16]; result register v of lane l:
// D[4 v + l / 16][l % 16], as a one-hot scan of the instruction on gfx950 shows).  Without
// (64 x 64 only): each thread a 4 x 4 block of FP64 fused multiply-adds (rows ty + 16 i,
// columns tx + 16 j).
typedef double kron_v4d __attribute__((ext_vector_type(4)));

template <int TM, int TN, bool MFMA, int kKt>
__device__ __forceinline__ void kron_tile_body(const KronSector *sec, const int32_t *tiles, int32_t nt,
                                               const double *__restrict__ ML, const double *__restrict__ MR,
                                               const double *__restrict__ x, double *__restrict__ z,
                                               double *__restrict__ zl, int xcd) {
  static_assert(MFMA || (TM == 64 && TN == 64), "the vector form computes 64 x 64 tiles");
  constexpr int PA = kKt * TM / 256, PB = kKt * TN / 256;   // staged elements per thread
  constexpr int BM = TM / 32, BN = TN / 32;                 // 16 x 16 blocks per wavefront
  __shared__ double As[2][kKt][TM + 1];   // As[buffer][k][row]
  __shared__ double Bs[2][kKt][TN + 1];   // Bs[buffer][k][col]
  // xcd: consecutive tiles (one sector's, in list order) to the blocks of one XCD, so that
  // a sector's operands are fetched into one XCD's L2
  const int32_t lt = xcd ? xcd_logical(blockIdx.x, nt) : (int32_t)blockIdx.x;
  const int32_t *tl = tiles + 4 * (int64_t)lt;
  const KronSector S = sec[tl[0]];
  const int32_t r0 = tl[1] * TM, c0 = tl[2] * TN;
  const bool right = tl[3] != 0;
  const int32_t K = right ? S.db : S.da;
  const double *X = x + S.off;
  const double *Aop = right ? X : ML + S.ml;           // A[i][k] = Aop[(r0 + i) lda + k]
  const int64_t lda = right ? S.db : S.da;
  const double *Bop = right ? MR + S.mr : X;           // left: B[k][j] = X[k db + c0 + j]
  double ra[PA], rb[PB];                               // right: B[k][j] = M_R[(c0 + j) db + k]
  // staging geometry, fixed over the k loop: A element u of this thread is row
  // tid / kKt + (256 / kKt) u of the tile at column tid % kKt of the slab (right B the same
  // with rows of M_R); left B element u is k-row tid / TN + (256 / TN) u at column tid % TN
  const int kA = threadIdx.x % kKt, iA = threadIdx.x / kKt;
  const double *pA = Aop + (int64_t)(r0 + iA) * lda + kA;
  const int64_t sA = (int64_t)(256 / kKt) * lda;
  const int kB = right ? kA : threadIdx.x / TN;
  const int jB = right ? iA : threadIdx.x % TN;
  const double *pB = right ? Bop + (int64_t)(c0 + jB) * S.db + kA : Bop + (int64_t)kB * S.db + c0 + jB;
  const int64_t sB = right ? (int64_t)(256 / kKt) * S.db : (int64_t)(256 / TN) * S.db;
  auto fetch = [&](int32_t k0) {
#pragma unroll
    for (int u = 0; u < PA; u++) {
      const bool ok = r0 + iA + (256 / kKt) * u < S.da && k0 + kA < K;
      ra[u] = ok ? pA[u * sA + k0] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PB; u++) {
      if (right) {
        const bool ok = c0 + jB + (256 / kKt) * u < S.db && k0 + kA < K;
        rb[u] = ok ? pB[u * sB + k0] : 0.0;
      } else {
        const bool ok = c0 + jB < S.db && k0 + kB + (256 / TN) * u < K;
        rb[u] = ok ? pB[u * sB + (int64_t)k0 * S.db] : 0.0;
      }
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int u = 0; u < PA; u++) {
      const int e = threadIdx.x + 256 * u;
      As[buf][e % kKt][e / kKt] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < PB; u++) {
      const int e = threadIdx.x + 256 * u;
      if (right) Bs[buf][e % kKt][e / kKt] = rb[u];
      else Bs[buf][e / TN][e % TN] = rb[u];
    }
  };
  const int wave = threadIdx.x / 64, l = threadIdx.x % 64;
  const int wr = (wave / 2) * (TM / 2), wc = (wave % 2) * (TN / 2);   // MFMA: this wavefront's part
  const int li = l % 16, lk = l / 16;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  kron_v4d macc[BM][BN];
  double vacc[4][4];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) macc[i][j] = kron_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) vacc[i][j] = 0.0;
  fetch(0);
  stage(0);
  __syncthreads();
  int buf = 0;
  for (int32_t k0 = 0; k0 < K; k0 += kKt) {
    const bool more = k0 + kKt < K;
    if (more) fetch(k0 + kKt);
    if constexpr (MFMA) {
#pragma unroll
      for (int kk = 0; kk < kKt; kk += 4) {
        double a[BM], b[BN];
#pragma unroll
        for (int i = 0; i < BM; i++) a[i] = As[buf][kk + lk][wr + 16 * i + li];
#pragma unroll
        for (int j = 0; j < BN; j++) b[j] = Bs[buf][kk + lk][wc + 16 * j + li];
#pragma unroll
        for (int i = 0; i < BM; i++)
#pragma unroll
          for (int j = 0; j < BN; j++) macc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], macc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll 2
      for (int k = 0; k < kKt; k++) {
        double av[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) av[i] = As[buf][k][ty + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; j++) bv[j] = Bs[buf][k][tx + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) vacc[i][j] = fma(av[i], bv[j], vacc[i][j]);
      }
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  double *Z = right ? z + S.off : zl + S.off;
  auto put = [&](int32_t rr, int32_t cc, double v) {
    if (rr < S.da && cc < S.db) {
      const int64_t at = (int64_t)rr * S.db + cc;
      Z[at] = v;
    }
  };
  if constexpr (MFMA) {
#pragma unroll
    for (int bi = 0; bi < BM; bi++)
#pragma unroll
      for (int bj = 0; bj < BN; bj++)
#pragma unroll
        for (int v = 0; v < 4; v++) put(r0 + wr + 16 * bi + 4 * v + lk, c0 + wc + 16 * bj + li, macc[bi][bj][v]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) put(r0 + ty + 16 * i, c0 + tx + 16 * j, vacc[i][j]);
  }
}

#define KRON_LAUNCH_ARGS                                                                     \
  const KronSector *sec, const int32_t *tiles, int32_t nt, const double *ML, const double *MR, \
      const double *x, double *z, double *zl, int xcd

template <int TM, int TN, bool MFMA, int kKt = 16>
__global__ __launch_bounds__(256) void k_kron_tile(KRON_LAUNCH_ARGS) {
  kron_tile_body<TM, TN, MFMA, kKt>(sec, tiles, nt, ML, MR, x, z, zl, xcd);
}
// the 64 x 64 matrix-core tile held to at most 128 registers per lane, four wavefronts per
// SIMD
__global__ __launch_bounds__(256, 4) void k_kron_tile_o4(KRON_LAUNCH_ARGS) {
  kron_tile_body<64, 64, true, 16>(sec, tiles, nt, ML, MR, x, z, zl, xcd);
}

// the 128 x 128 matrix-core tile held to two wavefronts per SIMD (at most 256 registers per
// lane): half the operand traffic per output of the 64 x 64 tile
__global__ __launch_bounds__(256, 2) void k_kron_tile_o2(KRON_LAUNCH_ARGS) {
  kron_tile_body<128, 128, true, 16>(sec, tiles, nt, ML, MR, x, z, zl, xcd);
}

__global__ __launch_bounds__(256) void k_zero(int64_t n, double *p) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0.0;
}

bool left_kind(int k) { return k == K_LL || k == K_LR || k == K_LLLR; }
bool right_kind(int k) { return k == K_RR || k == K_RL || k == K_RRRL; }

}  // namespace

// The coupling terms (to zc) on a side stream, beside the dense tiles (to z and zl) on the
// caller's stream: the tiles are queued first and hold four workgroups per CU, the coupling
// terms' workgroups, which need no LDS, fill the room beside them.
void KronDev::product(const double *x, hipStream_t s) const {
  const KronCross kc{ncross, na, nb, ct.p, lptr.p, lidx.p, rptr.p, ridx.p, lval.p, rval.p, asec.p, bsec.p, sec.p};
  const int shape = kKronShape[variant];
  const int nt = ntiles[shape];
  if (!side) {
    HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  }
  if (na > 0) HIPCHK(hipEventRecord(fork, s));   // x is ready, and the last product's zc has been read
  if (nt > 0) {
    const int32_t *tl = tiles[shape].p;
    // live timing (dmrgx_profile_begin_sampled): HIP events around the dense products, their
    // floating-point operations counted where the CRS products count bytes
    Context &C = ctx();
    std::pair<hipEvent_t, hipEvent_t> *ev = nullptr;
    if (C.prof_take()) {
      ev = C.prof_slot();
      HIPCHK(hipEventRecord(ev->first, s));
      C.prof_bytes += flops;
    }
#define KRON_ARGS sec.p, tl, nt, ML.p, MR.p, x, z.p, zl.p, xcd
    switch (variant) {
      case 0: k_kron_tile<64, 64, false><<<nt, 256, 0, s>>>(KRON_ARGS); break;
      case 1: k_kron_tile<64, 64, true><<<nt, 256, 0, s>>>(KRON_ARGS); break;
      case 2: k_kron_tile<128, 128, true><<<nt, 256, 0, s>>>(KRON_ARGS); break;
      case 3: k_kron_tile<64, 128, true><<<nt, 256, 0, s>>>(KRON_ARGS); break;
      case 4: k_kron_tile<128, 64, true><<<nt, 256, 0, s>>>(KRON_ARGS); break;
      case 5: k_kron_tile<64, 64, true, 32><<<nt, 256, 0, s>>>(KRON_ARGS); break;
      case 6: k_kron_tile_o4<<<nt, 256, 0, s>>>(KRON_ARGS); break;
      default: k_kron_tile_o2<<<nt, 256, 0, s>>>(KRON_ARGS); break;
    }
#undef KRON_ARGS
    if (ev) HIPCHK(hipEventRecord(ev->second, s));
    HIPCHK(hipGetLastError());
  }
  if (na > 0) {
    HIPCHK(hipStreamWaitEvent(side, fork, 0));
    k_kron_cross<<<na, 256, 0, side>>>(kc, x, zc.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(join, side));
    HIPCHK(hipStreamWaitEvent(s, join, 0));   // zc complete before the epilogue
  }
}

}  // namespace dmrgx

using namespace dmrgx;

struct dmrgx_kron {
  KronDev K;
};

namespace dmrgx {
const KronDev &kron_dev(const dmrgx_kron *K) { return K->K; }
}  // namespace dmrgx

extern "C" {

int dmrgx_superblock_kron(const dmrgx_superblock *S, dmrgx_kron **out) {
  if (!S || !out) return DMRGX_ERR_ARG;
  *out = nullptr;
  return run_status([&] {
    auto H = std::make_unique<dmrgx_kron>();
    KronDev &K = H->K;
    const int32_t n = S->n, d = S->d;
    K.n = n;
    K.sb = S;
    hipStream_t st = ctx().stream;
    // the basis (host copy) and its sectors
    std::vector<int16_t> LL(n), LR(n), RR(n), RL(n);
    ctx().copy_out(LL.data(), S->LL.p, sizeof(int16_t) * n);
    ctx().copy_out(LR.data(), S->LR.p, sizeof(int16_t) * n);
    ctx().copy_out(RR.data(), S->RR.p, sizeof(int16_t) * n);
    ctx().copy_out(RL.data(), S->RL.p, sizeof(int16_t) * n);
    std::vector<int16_t> aLL, aLR, bRR, bRL;
    std::vector<int32_t> asec, bsec;
    int64_t r = 0;
    while (r < n) {
      KronSector sc{};
      sc.off = r;
      // db: rows until (LL, LR) changes; da: runs of db rows with the same RRRL sequence
      int64_t e = r;
      while (e < n && LL[e] == LL[r] && LR[e] == LR[r]) e++;
      sc.db = (int32_t)(e - r);
      int64_t a = r;
      while (a < n && a + sc.db <= n && RR[a] == RR[r] && RL[a] == RL[r]) a += sc.db;
      sc.da = (int32_t)((a - r) / sc.db);
      // every row of the sector must be (LLLR state i, RRRL state j) at off + i db + j
      // the product check: every row run of the sector has one (LL, LR) and the first run's
      // RRRL sequence
      for (int32_t i = 0; i < sc.da; i++) {
        const int64_t q0 = r + (int64_t)i * sc.db;
        int same = std::memcmp(&RR[q0], &RR[r], sizeof(int16_t) * sc.db) == 0 &&
                   std::memcmp(&RL[q0], &RL[r], sizeof(int16_t) * sc.db) == 0;
        const int16_t ll = LL[q0], lr = LR[q0];
        for (int32_t j = 1; j < sc.db; j++) same &= (int)(LL[q0 + j] == ll) & (int)(LR[q0 + j] == lr);
        if (!same) throw Failure(DMRGX_ERR_ARG, "superblock basis is not a product of LLLR and RRRL states per sector");
      }
      sc.a0 = (int32_t)aLL.size();
      sc.b0 = (int32_t)bRR.size();
      for (int32_t i = 0; i < sc.da; i++) {
        aLL.push_back(LL[r + (int64_t)i * sc.db]);
        aLR.push_back(LR[r + (int64_t)i * sc.db]);
        asec.push_back(K.nsec);
      }
      for (int32_t j = 0; j < sc.db; j++) {
        bRR.push_back(RR[r + j]);
        bRL.push_back(RL[r + j]);
        bsec.push_back(K.nsec);
      }
      K.hsec.push_back(sc);
      K.nsec++;
      r += (int64_t)sc.da * sc.db;
    }
    // state -> flat index maps (each LLLR / RRRL state belongs to one sector)
    std::vector<int32_t> aidx((size_t)S->dim_LL * d, -1), bidx((size_t)S->dim_RR * d, -1);
    for (size_t i = 0; i < aLL.size(); i++) {
      int32_t &m = aidx[(size_t)aLL[i] * d + aLR[i]];
      if (m >= 0) throw Failure(DMRGX_ERR_ARG, "an LLLR state appears in two sectors");
      m = (int32_t)i;
    }
    for (size_t i = 0; i < bRR.size(); i++) {
      int32_t &m = bidx[(size_t)bRR[i] * d + bRL[i]];
      if (m >= 0) throw Failure(DMRGX_ERR_ARG, "an RRRL state appears in two sectors");
      m = (int32_t)i;
    }
    K.na = (int32_t)aLL.size();
    K.nb = (int32_t)bRR.size();
    int64_t ml = 0, mr = 0;
    std::vector<int32_t> tiles[kKronShapes];
    for (int s = 0; s < K.nsec; s++) {
      KronSector &sc = K.hsec[s];
      sc.ml = ml;
      sc.mr = mr;
      ml += (int64_t)sc.da * sc.da;
      mr += (int64_t)sc.db * sc.db;
      K.flops += 2.0 * sc.da * sc.db * (double)(sc.da + sc.db);
      for (int sh = 0; sh < kKronShapes; sh++)
        for (int ph = 0; ph < 2; ph++)   // the left and the right product of each tile
          for (int32_t ti = 0; ti * kKronTM[sh] < sc.da; ti++)
            for (int32_t tj = 0; tj * kKronTN[sh] < sc.db; tj++) {
              tiles[sh].push_back(s);
              tiles[sh].push_back(ti);
              tiles[sh].push_back(tj);
              tiles[sh].push_back(ph);
            }
    }
    // longest first (k extent of the tile's product: da for the left, db for the right one),
    // so that the short tiles fill in behind the long ones instead of the long ones ending last
    for (int sh = 0; sh < kKronShapes; sh++) {
      std::vector<std::array<int32_t, 4>> t(tiles[sh].size() / 4);
      std::memcpy(t.data(), tiles[sh].data(), sizeof(int32_t) * tiles[sh].size());
      auto klen = [&](const std::array<int32_t, 4> &a) { return a[3] ? K.hsec[a[0]].db : K.hsec[a[0]].da; };
      std::stable_sort(t.begin(), t.end(), [&](const auto &a, const auto &b) { return klen(a) > klen(b); });
      std::memcpy(tiles[sh].data(), t.data(), sizeof(int32_t) * tiles[sh].size());
      K.ntiles[sh] = (int32_t)t.size();
    }
    // the term list split by the halves it acts on
    std::vector<int32_t> left, right, cross;
    for (size_t i = 0; i < S->hterms.size(); i++) {
      const int k = S->hterms[i].kind;
      if (left_kind(k)) left.push_back((int32_t)i);
      else if (right_kind(k)) right.push_back((int32_t)i);
      else cross.push_back((int32_t)i);
    }
    K.ncross = (int32_t)cross.size();
    // the coupling terms' per-state lists (see k_kron_cross), from host copies of the operators
    {
      auto op_host = [&](int id, std::vector<int64_t> &r, std::vector<int32_t> &c, std::vector<double> &v) {
        r.resize(S->op_row[id].count);
        ctx().copy_out(r.data(), S->op_row[id].p, sizeof(int64_t) * r.size());
        const int64_t nnz = r.back();
        c.resize(std::max<int64_t>(nnz, 1));
        v.resize(std::max<int64_t>(nnz, 1));
        if (nnz > 0) {
          ctx().copy_out(c.data(), S->op_col[id].p, sizeof(int32_t) * nnz);
          ctx().copy_out(v.data(), S->op_val[id].p, sizeof(double) * nnz);
        }
      };
      std::vector<double> ct;
      std::vector<int32_t> lptr, lidx, rptr, ridx;
      std::vector<double> lval, rval;
      for (int32_t ci : cross) {
        const DTerm &t = S->hterms[ci];
        std::vector<int64_t> r1, r2;
        std::vector<int32_t> c1, c2;
        std::vector<double> v1, v2;
        op_host(t.op1, r1, c1, v1);
        op_host(t.op2, r2, c2, v2);
        auto par = [&](int e, int state) { return S->hele[e][state]; };
        // parity of the row's LLLR part and of its RRRL part in the term's sign rule
        auto left_parity = [&](int ll, int lr) {
          if (t.sign_mode == 0) return 0;
          switch (t.kind) {
            case K_LLRL: return par(t.e0, ll) + par(t.e2, lr);
            case K_LLRR: return par(t.e0, ll) + par(t.e1, lr);
            case K_LRRL: return par(t.e1, lr);
            case K_LRRR: return par(t.e0, lr);
            default: return 0;
          }
        };
        auto right_parity = [&](int rr) {
          if (t.sign_mode == 0) return 0;
          switch (t.kind) {
            case K_LLRL: return par(t.e1, rr);
            case K_LRRL: return par(t.e0, rr);
            default: return 0;
          }
        };
        ct.push_back((t.sign_mode == 1 ? -1.0 : 1.0) * t.coeff);
        const int32_t base_l = (int32_t)lidx.size(), base_r = (int32_t)ridx.size();
        (void)base_l;
        (void)base_r;
        for (int32_t i = 0; i < K.na; i++) {
          lptr.push_back((int32_t)lidx.size());
          const int q[2] = {aLL[i], aLR[i]};
          const double sg = (left_parity(q[0], q[1]) % 2 == 0) ? 1.0 : -1.0;
          for (int64_t e = r1[q[t.s1]]; e < r1[q[t.s1] + 1]; e++) {
            int qq[2] = {q[0], q[1]};
            qq[t.s1] = c1[e];
            const int32_t a = aidx[(size_t)qq[0] * d + qq[1]];
            if (a < 0) continue;
            lidx.push_back(a);
            lval.push_back(sg * v1[e]);
          }
        }
        lptr.push_back((int32_t)lidx.size());
        for (int32_t i = 0; i < K.nb; i++) {
          rptr.push_back((int32_t)ridx.size());
          const int q[4] = {0, 0, bRR[i], bRL[i]};
          const double sg = (right_parity(q[2]) % 2 == 0) ? 1.0 : -1.0;
          for (int64_t e = r2[q[t.s2]]; e < r2[q[t.s2] + 1]; e++) {
            int qq[4] = {q[0], q[1], q[2], q[3]};
            qq[t.s2] = c2[e];
            const int32_t b = bidx[(size_t)qq[2] * d + qq[3]];
            if (b < 0) continue;
            ridx.push_back(b);
            rval.push_back(sg * v2[e]);
          }
        }
        rptr.push_back((int32_t)ridx.size());
      }
      if ((int64_t)lidx.size() > INT32_MAX / 2 || (int64_t)ridx.size() > INT32_MAX / 2)
        throw Failure(DMRGX_ERR_ARG, "coupling-term lists too long");
      auto upl = [](auto &buf, const auto &h) {
        using T = typename std::decay_t<decltype(h)>::value_type;
        buf.resize(std::max<size_t>(h.size(), 1));
        if (!h.empty()) ctx().copy_in(buf.p, h.data(), sizeof(T) * h.size());
      };
      upl(K.ct, ct);
      upl(K.lptr, lptr);
      upl(K.lidx, lidx);
      upl(K.lval, lval);
      upl(K.rptr, rptr);
      upl(K.ridx, ridx);
      upl(K.rval, rval);
    }
    auto up = [](auto &buf, const auto &h) {
      using T = typename std::decay_t<decltype(h)>::value_type;
      buf.resize(std::max<size_t>(h.size(), 1));
      if (!h.empty()) ctx().copy_in(buf.p, h.data(), sizeof(T) * h.size());
    };
    up(K.sec, K.hsec);
    for (int sh = 0; sh < kKronShapes; sh++) up(K.tiles[sh], tiles[sh]);
    up(K.cross, cross);
    DevBuf<int16_t> daLL, daLR, dbRR, dbRL;
    DevBuf<int32_t> dleft, dright;
    up(daLL, aLL);
    up(daLR, aLR);
    up(dbRR, bRR);
    up(dbRL, bRL);
    up(K.asec, asec);
    up(K.bsec, bsec);
    up(K.aidx, aidx);
    up(K.bidx, bidx);
    up(dleft, left);
    up(dright, right);
    K.ML.resize(std::max<int64_t>(ml, 1));
    K.MR.resize(std::max<int64_t>(mr, 1));
    k_zero<<<1024, 256, 0, st>>>(ml, K.ML.p);
    k_zero<<<1024, 256, 0, st>>>(mr, K.MR.p);
    const KronBuild kb{d, K.na, K.nb, daLL.p, daLR.p, dbRR.p, dbRL.p, K.asec.p, K.bsec.p, K.aidx.p, K.bidx.p, K.sec.p};
    if (K.na > 0)
      k_kron_left<<<(K.na + 255) / 256, 256, 0, st>>>(kb, S->terms.p, dleft.p, (int)left.size(), S->ops.p,
                                                      S->ele_ptrs.p, K.ML.p);
    if (K.nb > 0)
      k_kron_right<<<(K.nb + 255) / 256, 256, 0, st>>>(kb, S->terms.p, dright.p, (int)right.size(), S->ops.p,
                                                       S->ele_ptrs.p, K.MR.p);
    HIPCHK(hipGetLastError());
    K.z.resize(std::max<int32_t>(n, 1));
    K.zl.resize(std::max<int32_t>(n, 1));
    K.zc.resize(std::max<int32_t>(n, 1));
    HIPCHK(hipStreamSynchronize(st));   // the build's own buffers go out of scope
    *out = H.release();
  });
}

int dmrgx_kron_info(const dmrgx_kron *K, int32_t *n, int32_t *sectors, int32_t *coupling_terms, double *flops) {
  if (!K) return DMRGX_ERR_ARG;
  if (n) *n = K->K.n;
  if (sectors) *sectors = K->K.nsec;
  if (coupling_terms) *coupling_terms = K->K.ncross;
  if (flops) *flops = K->K.flops;
  return DMRGX_OK;
}

int dmrgx_kron_set_kernel(dmrgx_kron *K, int32_t variant) {
  if (!K || variant < 0 || variant >= 2 * kKronVariants) return DMRGX_ERR_ARG;
  K->K.variant = variant % kKronVariants;
  K->K.xcd = variant / kKronVariants;
  return DMRGX_OK;
}

int dmrgx_kron_autotune(dmrgx_kron *K, int32_t *variant_out, double *us_out) {
  if (!K) return DMRGX_ERR_ARG;
  return run_status([&] {
    KronDev &D = K->K;
    hipStream_t st = ctx().stream;
    DevBuf<double> x(std::max<int32_t>(D.n, 1));
    launch_start_vector(D.n, 7, x.p, st);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int best = D.variant, best_xcd = D.xcd;
    float best_ms = 1e30f, dflt_ms = 1e30f;
    for (int c = 0; c < 2 * kKronVariants; c++) {
      const int v = c % kKronVariants, xm = c / kKronVariants;
      D.variant = v;
      D.xcd = xm;
      D.product(x.p, st);
      HIPCHK(hipEventRecord(e0, st));
      for (int r = 0; r < 5; r++) D.product(x.p, st);
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (c == kKronDefault) dflt_ms = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = v;
        best_xcd = xm;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the default kernel stays unless another is more than 2 % faster: the five products'
    // timing varies by about 1 %, and between the two 64 x 64 kernels, which time within
    // that, whole solves favour the default-occupancy one
    if (dflt_ms <= 1.02f * best_ms) {
      best_ms = dflt_ms;
      best = kKronDefault % kKronVariants;
      best_xcd = kKronDefault / kKronVariants;
    }
    D.variant = best;
    D.xcd = best_xcd;
    if (variant_out) *variant_out = best + kKronVariants * best_xcd;
    if (us_out) *us_out = best_ms * 1000.0 / 5;
  });
}

int dmrgx_kron_spmv(const dmrgx_kron *K, const double *x, double *y) {
  if (!K || !x || !y) return DMRGX_ERR_ARG;
  return run_status([&] {
    const int32_t n = K->K.n;
    DevBuf<double> dx(n);
    ctx().copy_in(dx.p, x, sizeof(double) * n);
    K->K.product(dx.p, ctx().stream);
    std::vector<double> zl(n), zc(n);
    ctx().copy_out(y, K->K.z.p, sizeof(double) * n);
    ctx().copy_out(zl.data(), K->K.zl.p, sizeof(double) * n);
    ctx().copy_out(zc.data(), K->K.zc.p, sizeof(double) * n);
    for (int32_t i = 0; i < n; i++) y[i] = (y[i] + zl[i]) + zc[i];   // as the solvers' epilogue sums them
  });
}

void dmrgx_kron_free(dmrgx_kron *K) {
  if (!K) return;
  (void)hipStreamSynchronize(ctx().stream);
  delete K;
}

}  // extern "C"
