// Superblock basis from the blocks' quantum numbers: the part of GET_BASIS_SUPERBLOCK
// (main/dmrg/hubbard_vf/src/allocations/GET_BASIS_SUPERBLOCK.c:63-97, xxz_vf ...:36-64,
// klm_vf ...:66-97) that lists the states (LL, LR, RR, RL) of the target sector in the
// order DMRG_QUICK_SORT_BASIS_Q1/Q2 give them (dmrg/DMRG_QUICK_SORT_BASIS_Q2.c:29-46):
// electrons of LL+LR, then S^z of LL+LR, then LL, LR, RR, RL, ascending.
//
// The reference enumerates the whole product space (dim_LL * d * dim_RR * d states) and
// sorts the kept ones.  That order has a product structure that makes both unnecessary:
// the left half (LL, LR) of a state fixes its left quantum numbers, the target fixes the
// right ones, and every left state of a class pairs with every right state of the
// complementary class.  So the basis is a sequence of sectors, one per left class in
// ascending (electrons, S^z) order, each the left states of the class in (LL, LR) order
// times the right states of the complementary class in (RR, RL) order, right index
// fastest.  The host groups the dim_LL * d left and dim_RR * d right states into classes
// (small); one device thread per superblock row then finds its sector and writes its four
// indices.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <utility>
#include <vector>

#include "entry.h"

namespace dmrgx {
namespace {

struct BasisSector {
  int64_t off;       // first superblock row of the sector
  int32_t dB;        // right states per left state
  int32_t a0, b0;    // first left / right state of the sector in the packed lists
};

// One thread per superblock row: binary search of its sector, then the left state of the
// row (row / dB within the sector) and the right state (row % dB).
__global__ __launch_bounds__(256) void k_basis_fill(int64_t n, int32_t n_sec, const BasisSector *__restrict__ sec,
                                                    const int16_t *__restrict__ a_half,
                                                    const int16_t *__restrict__ b_half, int16_t *__restrict__ LL,
                                                    int16_t *__restrict__ LR, int16_t *__restrict__ RR,
                                                    int16_t *__restrict__ RL) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  int32_t lo = 0, hi = n_sec - 1;   // last sector with off <= i
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (sec[mid].off <= i) lo = mid;
    else hi = mid - 1;
  }
  const BasisSector s = sec[lo];
  const int64_t k = i - s.off;
  const int32_t ia = s.a0 + (int32_t)(k / s.dB), ib = s.b0 + (int32_t)(k % s.dB);
  LL[i] = a_half[2 * ia];
  LR[i] = a_half[2 * ia + 1];
  RR[i] = b_half[2 * ib];
  RL[i] = b_half[2 * ib + 1];
}

}  // namespace

// The sector list of the basis; returns n.  a_half / b_half: (LL, LR) / (RR, RL) pairs of
// the sectors' left / right states, packed per class.
int64_t basis_sectors(const dmrgx_basis_qn &q, std::vector<BasisSector> &sec, std::vector<int16_t> &a_half,
                      std::vector<int16_t> &b_half) {
  if (q.n_qn != 1 && q.n_qn != 2) throw Failure(DMRGX_ERR_ARG, "n_qn must be 1 (S^z) or 2 (S^z and electrons)");
  if (q.dim_LL <= 0 || q.dim_RR <= 0 || q.dim_onsite <= 0)
    throw Failure(DMRGX_ERR_ARG, "basis dimensions must be positive");
  if (q.dim_LL > 32767 || q.dim_RR > 32767 || q.dim_onsite > 32767)
    throw Failure(DMRGX_ERR_ARG, "basis dimensions above 32767 do not fit the 16-bit basis arrays");
  if (!q.sz_LL || !q.sz_RR || !q.sz_site || (q.n_qn == 2 && (!q.ele_LL || !q.ele_RR || !q.ele_site)))
    throw Failure(DMRGX_ERR_ARG, "quantum-number array missing");
  const bool two = q.n_qn == 2;
  // classes keyed (electrons, S^z), in ascending order; members in enumeration order
  using Key = std::pair<int64_t, int64_t>;
  std::map<Key, std::vector<int16_t>> left, right;
  for (int32_t ll = 0; ll < q.dim_LL; ll++)
    for (int32_t lr = 0; lr < q.dim_onsite; lr++) {
      const Key k{two ? (int64_t)q.ele_LL[ll] + q.ele_site[lr] : 0, (int64_t)q.sz_LL[ll] + q.sz_site[lr]};
      auto &v = left[k];
      v.push_back((int16_t)ll);
      v.push_back((int16_t)lr);
    }
  for (int32_t rr = 0; rr < q.dim_RR; rr++)
    for (int32_t rl = 0; rl < q.dim_onsite; rl++) {
      const Key k{two ? (int64_t)q.ele_RR[rr] + q.ele_site[rl] : 0, (int64_t)q.sz_RR[rr] + q.sz_site[rl]};
      auto &v = right[k];
      v.push_back((int16_t)rr);
      v.push_back((int16_t)rl);
    }
  sec.clear();
  a_half.clear();
  b_half.clear();
  int64_t n = 0;
  for (const auto &kv : left) {
    const Key want{two ? (int64_t)q.target_ele - kv.first.first : 0, (int64_t)q.target_sz - kv.first.second};
    const auto it = right.find(want);
    if (it == right.end()) continue;
    const int32_t dA = (int32_t)(kv.second.size() / 2), dB = (int32_t)(it->second.size() / 2);
    sec.push_back(BasisSector{n, dB, (int32_t)(a_half.size() / 2), (int32_t)(b_half.size() / 2)});
    a_half.insert(a_half.end(), kv.second.begin(), kv.second.end());
    b_half.insert(b_half.end(), it->second.begin(), it->second.end());
    n += (int64_t)dA * dB;
  }
  if (n > INT32_MAX) throw Failure(DMRGX_ERR_ARG, "superblock sector dimension above 2^31 - 1");
  return n;
}

// Device basis arrays of the sector (LL, LR, RR, RL resized to n); returns n.
int64_t build_basis_device(const dmrgx_basis_qn &q, DevBuf<int16_t> &LL, DevBuf<int16_t> &LR, DevBuf<int16_t> &RR,
                           DevBuf<int16_t> &RL) {
  std::vector<BasisSector> sec;
  std::vector<int16_t> a_half, b_half;
  const int64_t n = basis_sectors(q, sec, a_half, b_half);
  if (n == 0) throw Failure(DMRGX_ERR_ARG, "no state of the blocks reaches the target quantum numbers");
  Context &C = ctx();
  DevBuf<BasisSector> dsec(sec.size());
  DevBuf<int16_t> da(a_half.size()), db(b_half.size());
  C.copy_in(dsec.p, sec.data(), sizeof(BasisSector) * sec.size());
  C.copy_in(da.p, a_half.data(), sizeof(int16_t) * a_half.size());
  C.copy_in(db.p, b_half.data(), sizeof(int16_t) * b_half.size());
  LL.resize(n);
  LR.resize(n);
  RR.resize(n);
  RL.resize(n);
  k_basis_fill<<<(unsigned)((n + 255) / 256), 256, 0, C.stream>>>(n, (int32_t)sec.size(), dsec.p, da.p, db.p, LL.p,
                                                                   LR.p, RR.p, RL.p);
  HIPCHK(hipGetLastError());
  return n;
}

}  // namespace dmrgx

using namespace dmrgx;

int dmrgx_superblock_basis(const dmrgx_basis_qn *q, int64_t *n_out, int16_t *LL, int16_t *LR, int16_t *RR,
                           int16_t *RL) {
  if (!q || !n_out) return DMRGX_ERR_ARG;
  if ((LL || LR || RR || RL) && !(LL && LR && RR && RL)) return DMRGX_ERR_ARG;
  return run_status([&] {
    if (!LL) {
      std::vector<BasisSector> sec;
      std::vector<int16_t> a_half, b_half;
      *n_out = basis_sectors(*q, sec, a_half, b_half);
      return;
    }
    DevBuf<int16_t> dLL, dLR, dRR, dRL;
    const int64_t n = build_basis_device(*q, dLL, dLR, dRR, dRL);
    Context &C = ctx();
    C.copy_out(LL, dLL.p, sizeof(int16_t) * n);
    C.copy_out(LR, dLR.p, sizeof(int16_t) * n);
    C.copy_out(RR, dRR.p, sizeof(int16_t) * n);
    C.copy_out(RL, dRL.p, sizeof(int16_t) * n);
    *n_out = n;
  });
}
