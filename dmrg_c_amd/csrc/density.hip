// Reduced density matrices of the left half (LL, LR) of the superblock ground state:
// DMRG_GET_DENSITY_MATRIX (dmrg/DMRG_GET_DENSITY_MATRIX.c:5-107), called once per block by
// DMRG_GET_SYSTEM_INFO_Q1..Q3 (dmrg/DMRG_GET_SYSTEM_INFO_Q2.c:22-23).
//
// A block is a run of Dim_Block[b] consecutive LLLR states (LL_LLLR / LR_LLLR) with the same
// quantum numbers.  Every state of a block has the same Sum_LLLR partners (RR, RL), listed
// at RR/RL_LLLRRRRL[Base_LLLR[s] + k] with Base_LLLR the prefix sums of Sum_LLLR, so with
//   Psi_b[r][k] = Vec[Inv(LL_LLLR[s_r], LR_LLLR[s_r], RR_k, RL_k)]
// the block's matrix is rho_b[r][c] = sum_k Psi_b[r][k] Psi_b[c][k], added in k order.
// The host checks the reference's conditions (same Sum and partners within a block, no
// missing state) and builds the O(n) index array of Psi; the device gathers Psi into a
// dense row-major copy and forms every block's matrix with an LDS-tiled product whose
// entries each add their k terms in order with separate multiplies and adds (this file
// is built without FMA contraction), so the result is the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "entry.h"

namespace dmrgx {
namespace {

constexpr int kDmTile = 32;   // output tile edge
constexpr int kDmK = 32;      // k terms staged per pass

struct DmTile {
  int64_t psi_off;   // first element of the block's Psi (row-major, db x sum)
  int64_t out_off;   // first element of the block's matrix (row-major, db x db)
  int32_t db, sum, r0, c0;
};

__global__ __launch_bounds__(256) void k_dm_gather(int64_t n, const int32_t *__restrict__ idx,
                                                   const double *__restrict__ vec, double *__restrict__ psi) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) psi[i] = vec[idx[i]];
}

// One workgroup per tile (r0, c0) with c0 >= r0 of one block.  Thread (ty, tx) owns the
// outputs (r0 + ty + 8 i, c0 + tx), i = 0..3; it writes them and, off the diagonal tile,
// their mirror images.
__global__ __launch_bounds__(256) void k_dm_syrk(const DmTile *__restrict__ tiles, const double *__restrict__ psi,
                                                 double *__restrict__ out) {
  __shared__ double As[kDmTile][kDmK + 1];
  __shared__ double Bs[kDmTile][kDmK + 1];
  const DmTile T = tiles[blockIdx.x];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const double *P = psi + T.psi_off;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < T.sum; k0 += kDmK) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int rr = ty + 8 * i, k = k0 + tx;
      const int ra = T.r0 + rr, rb = T.c0 + rr;
      As[rr][tx] = (ra < T.db && k < T.sum) ? P[(int64_t)ra * T.sum + k] : 0.0;
      Bs[rr][tx] = (rb < T.db && k < T.sum) ? P[(int64_t)rb * T.sum + k] : 0.0;
    }
    __syncthreads();
    const int kn = min(kDmK, T.sum - k0);
    for (int kk = 0; kk < kn; kk++) {
      const double b = Bs[tx][kk];
#pragma unroll
      for (int i = 0; i < 4; i++) acc[i] = acc[i] + As[ty + 8 * i][kk] * b;
    }
    __syncthreads();
  }
  const int c = T.c0 + tx;
  if (c >= T.db) return;
  double *O = out + T.out_off;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = T.r0 + ty + 8 * i;
    if (r >= T.db) continue;
    O[(int64_t)r * T.db + c] = acc[i];
    if (T.c0 != T.r0) O[(int64_t)c * T.db + r] = acc[i];
  }
}

// Checks and the Psi index array for blocks [b0, b1).  Returns the tiles; idx gets the Vec
// index of every Psi element, packed block after block.
std::vector<DmTile> dm_plan(const DMRG_BASIS &B, int32_t b0, int32_t b1, const int32_t *Dim_Block,
                            std::vector<int32_t> &idx, std::vector<int64_t> &out_off) {
  const int32_t dim_LLLR = B.dim_LLLR, n = B.dim_LLLRRRRL;
  if (dim_LLLR <= 0 || n <= 0 || !B.LL_LLLR || !B.LR_LLLR || !B.Sum_LLLR || !B.RR_LLLRRRRL || !B.RL_LLLRRRRL ||
      !B.Inv_LLLRRRRL)
    throw Failure(DMRGX_ERR_ARG, "basis without its LLLR sub-basis, partners or inverse table");
  const int32_t d = B.dim_onsite, dim_RR = B.dim_RR, dim_LL = B.dim_LL;
  std::vector<int64_t> base(dim_LLLR + 1, 0);
  for (int32_t s = 0; s < dim_LLLR; s++) {
    if (B.Sum_LLLR[s] < 0) throw Failure(DMRGX_ERR_ARG, "negative Sum_LLLR");
    base[s + 1] = base[s] + B.Sum_LLLR[s];
  }
  if (base[dim_LLLR] != n) throw Failure(DMRGX_ERR_ARG, "Sum_LLLR does not add up to dim_LLLRRRRL");
  int64_t first = 0;
  for (int32_t b = 0; b < b0; b++) first += Dim_Block[b];
  idx.clear();
  out_off.assign(1, 0);
  std::vector<DmTile> tiles;
  for (int32_t b = b0; b < b1; b++) {
    const int32_t db = Dim_Block[b];
    if (db <= 0 || first + db > dim_LLLR) throw Failure(DMRGX_ERR_ARG, "block outside the LLLR states");
    const int32_t s0 = (int32_t)first, sum = B.Sum_LLLR[s0];
    const int64_t psi_off = (int64_t)idx.size();
    for (int32_t r = 0; r < db; r++) {
      const int32_t s = s0 + r;
      if (B.Sum_LLLR[s] != sum) throw Failure(DMRGX_ERR_ARG, "Sum_LLLR differs within a block");
      const int LL = B.LL_LLLR[s], LR = B.LR_LLLR[s];
      if (LL < 0 || LL >= dim_LL || LR < 0 || LR >= d) throw Failure(DMRGX_ERR_ARG, "LLLR state outside its blocks");
      for (int32_t k = 0; k < sum; k++) {
        const int RR = B.RR_LLLRRRRL[base[s] + k], RL = B.RL_LLLRRRRL[base[s] + k];
        if (RR != B.RR_LLLRRRRL[base[s0] + k] || RL != B.RL_LLLRRRRL[base[s0] + k])
          throw Failure(DMRGX_ERR_ARG, "partner states differ within a block");
        if (RR < 0 || RR >= dim_RR || RL < 0 || RL >= d) throw Failure(DMRGX_ERR_ARG, "partner state outside its blocks");
        const int32_t inv = B.Inv_LLLRRRRL[(((int64_t)LL * d + LR) * dim_RR + RR) * d + RL];
        if (inv < 0 || inv >= n) throw Failure(DMRGX_ERR_ARG, "state missing from the inverse table");
        idx.push_back(inv);
      }
    }
    for (int32_t r0 = 0; r0 < db; r0 += kDmTile)
      for (int32_t c0 = r0; c0 < db; c0 += kDmTile) tiles.push_back(DmTile{psi_off, out_off.back(), db, sum, r0, c0});
    out_off.push_back(out_off.back() + (int64_t)db * db);
    first += db;
  }
  return tiles;
}

// Device matrices of blocks [b0, b1) into host memory at out (packed row-major blocks).
void density_blocks(const DMRG_BASIS &B, int32_t b0, int32_t b1, const int32_t *Dim_Block, const double *Vec,
                    double *out) {
  std::vector<int32_t> idx;
  std::vector<int64_t> out_off;
  const std::vector<DmTile> tiles = dm_plan(B, b0, b1, Dim_Block, idx, out_off);
  const int64_t n_out = out_off.back(), n_psi = (int64_t)idx.size();
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int32_t n = B.dim_LLLRRRRL;
  DevBuf<double> vec(n), psi(n_psi > 0 ? n_psi : 1), dout(n_out > 0 ? n_out : 1);
  DevBuf<int32_t> didx(n_psi > 0 ? n_psi : 1);
  DevBuf<DmTile> dt(tiles.size());
  C.copy_in(vec.p, Vec, sizeof(double) * n);
  if (n_psi > 0) C.copy_in(didx.p, idx.data(), sizeof(int32_t) * n_psi);
  C.copy_in(dt.p, tiles.data(), sizeof(DmTile) * tiles.size());
  if (n_psi > 0) k_dm_gather<<<(unsigned)((n_psi + 255) / 256), 256, 0, st>>>(n_psi, didx.p, vec.p, psi.p);
  // an empty partner list leaves its block's matrix zero, as the reference's does
  k_dm_syrk<<<(unsigned)tiles.size(), 256, 0, st>>>(dt.p, psi.p, dout.p);
  HIPCHK(hipGetLastError());
  C.copy_out(out, dout.p, sizeof(double) * n_out);
}

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

extern "C" {

int dmrgx_density_matrices(const DMRG_BASIS *B, int32_t n_blocks, const int32_t *Dim_Block, const double *Vec,
                           int64_t *n_out, double *out) {
  if (!B || n_blocks <= 0 || !Dim_Block || !n_out) return DMRGX_ERR_ARG;
  return run_status([&] {
    int64_t total = 0, rows = 0;
    for (int32_t b = 0; b < n_blocks; b++) {
      if (Dim_Block[b] <= 0) throw Failure(DMRGX_ERR_ARG, "block dimension must be positive");
      total += (int64_t)Dim_Block[b] * Dim_Block[b];
      rows += Dim_Block[b];
    }
    if (rows > B->dim_LLLR) throw Failure(DMRGX_ERR_ARG, "blocks hold more states than dim_LLLR");
    *n_out = total;
    if (!out) return;
    if (!Vec) throw Failure(DMRGX_ERR_ARG, "state vector missing");
    density_blocks(*B, 0, n_blocks, Dim_Block, Vec, out);
  });
}

double **DMRG_GET_DENSITY_MATRIX(int block, int *Dim_Block, DMRG_BASIS *Dmrg_Basis, double *Vec, int p_threads) {
  (void)p_threads;
  double **M = nullptr;
  run_or_die("DMRG_GET_DENSITY_MATRIX", [&] {
    if (!Dim_Block || !Dmrg_Basis || !Vec || block < 0) throw Failure(DMRGX_ERR_ARG, "bad arguments");
    const int32_t db = Dim_Block[block];
    if (db <= 0) throw Failure(DMRGX_ERR_ARG, "block dimension must be positive");
    std::vector<double> dense((size_t)db * db);
    density_blocks(*Dmrg_Basis, block, block + 1, Dim_Block, Vec, dense.data());
    // rows allocated one by one, as GET_ARRAY_DOUBLE2 does, so FREE_ARRAY_DOUBLE2 frees them
    M = static_cast<double **>(std::malloc(sizeof(double *) * db));
    if (!M) throw Failure(DMRGX_ERR_NOMEM, "host allocation failed");
    for (int32_t r = 0; r < db; r++) {
      M[r] = static_cast<double *>(std::malloc(sizeof(double) * db));
      if (!M[r]) throw Failure(DMRGX_ERR_NOMEM, "host allocation failed");
      std::memcpy(M[r], dense.data() + (size_t)r * db, sizeof(double) * db);
    }
  });
  return M;
}

}  // extern "C"
