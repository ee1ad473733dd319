// Superblock handle internals shared by the CRS assembler (assemble.hip) and the
// Kronecker form of the same operator (kron.hip).
#pragma once

#include <vector>

#include "entry.h"

namespace dmrgx {

enum { K_LL = 0, K_LR, K_RR, K_RL, K_LLLR, K_LLRL, K_LLRR, K_LRRL, K_LRRR, K_RRRL, K_ZERO, K_COUNT };
enum { S_LL = 0, S_LR = 1, S_RR = 2, S_RL = 3 };
enum { ROLE_LL = 0, ROLE_RR = 1, ROLE_ON = 2 };
inline constexpr int kSlot1[K_COUNT] = {S_LL, S_LR, S_RR, S_RL, S_LL, S_LL, S_LL, S_LR, S_LR, S_RR, -1};
inline constexpr int kSlot2[K_COUNT] = {-1, -1, -1, -1, S_LR, S_RL, S_RR, S_RL, S_RR, S_RL, -1};

constexpr double kZero = 1.0e-15;  // dmrg/DMRG_MAKE_ELEM_LL_LLLRRRRL.c:18
constexpr int kMaxTerms = 256;

struct DTerm {
  int32_t kind, op1, op2, s1, s2, sign_mode, e0, e1, e2, pad;
  double coeff;
};
struct DOp {
  const int64_t *row;
  const int32_t *col;
  const double *val;
};
struct SbView {
  int32_t n, dim_RR, d, nt, has_zero;
  const int16_t *LL, *LR, *RR, *RL;
  const int32_t *inv;
  const DTerm *terms;
  const DOp *ops;
  const int32_t *const *eles;
};

}  // namespace dmrgx

// --------------------------------------------------------------------- handles
struct dmrgx_terms {
  struct T {
    int kind;
    const CRS1 *op1, *op2;
    double coeff;
    int sign_mode;
    const int *ele[3];
    int role[3];
  };
  std::vector<T> t;
};

struct dmrgx_superblock {
  int32_t n = 0, dim_LL = 0, dim_RR = 0, d = 0, nt = 0, has_zero = 0;
  dmrgx::DevBuf<int16_t> LL, LR, RR, RL;
  dmrgx::DevBuf<int32_t> inv;
  std::vector<dmrgx::DevBuf<int64_t>> op_row;
  std::vector<dmrgx::DevBuf<int32_t>> op_col;
  std::vector<dmrgx::DevBuf<double>> op_val;
  std::vector<dmrgx::DevBuf<int32_t>> ele;
  dmrgx::DevBuf<dmrgx::DOp> ops;
  dmrgx::DevBuf<dmrgx::DTerm> terms;
  dmrgx::DevBuf<const int32_t *> ele_ptrs;
  std::vector<dmrgx::DTerm> hterms;   // host copy of terms (without ZERO and zero-coefficient terms)
  std::vector<std::vector<int32_t>> hele;   // host copies of the parity arrays
  // last assembled lower triangle
  dmrgx::DevBuf<int64_t> lrow;
  dmrgx::DevBuf<int32_t> lcol;
  dmrgx::DevBuf<double> lval;
  int64_t lnnz = -1;
  // counted by the assembly for the expansion: strictly-lower entries per column, and a
  // flag (bit 0) for rows without a diagonal entry
  dmrgx::DevBuf<unsigned> tcnt;
  dmrgx::DevBuf<int> bad;
  dmrgx::SbView view() const {
    return dmrgx::SbView{n, dim_RR, d, nt, has_zero, LL.p, LR.p, RR.p, RL.p, inv.p, terms.p, ops.p, ele_ptrs.p};
  }
};

namespace dmrgx {

__device__ __forceinline__ int64_t sb_index(int q0, int q1, int q2, int q3, int dim_RR, int d) {
  return (int64_t)q0 * d * dim_RR * d + (int64_t)q1 * dim_RR * d + (int64_t)q2 * d + q3;
}

__device__ __forceinline__ int term_sign(const DTerm &t, const int32_t *const *eles, const int q[4]) {
  // sign rules of DMRG_MAKE_ELEM_{LLLR,LLRL,LLRR,LRRL,LRRR,RRRL}_LLLRRRRL.c:28-52
  if (t.sign_mode == 0) return 1;
  int p = 0;
  switch (t.kind) {
    case K_LLLR: p = eles[t.e0][q[S_LL]]; break;
    case K_LLRL: p = eles[t.e0][q[S_LL]] + eles[t.e1][q[S_RR]] + eles[t.e2][q[S_LR]]; break;
    case K_LLRR: p = eles[t.e0][q[S_LL]] + eles[t.e1][q[S_LR]]; break;
    case K_LRRL: p = eles[t.e0][q[S_RR]] + eles[t.e1][q[S_LR]]; break;
    case K_LRRR: p = eles[t.e0][q[S_LR]]; break;
    case K_RRRL: p = eles[t.e0][q[S_RR]]; break;
    default: return 1;
  }
  const bool even = (p % 2 == 0);
  return (t.sign_mode == 1) ? (even ? -1 : 1) : (even ? 1 : -1);
}

}  // namespace dmrgx
