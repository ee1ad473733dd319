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

#include <atomic>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <utility>
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

// Copies block matrices into a batch of p x p matrices: slot j holds the block whose matrix
// starts at src_off[j] (dimension dim[j]) in its leading corner, -1 on the rest of the
// diagonal and 0 elsewhere.  The padded matrix is block-diagonal, so its eigenpairs are the
// block's plus p - dim[j] pairs of eigenvalue -1 (below every eigenvalue of a density
// matrix), which the eigensolver lists first.
__global__ __launch_bounds__(256) void k_dm_pad(int32_t p, const int64_t *__restrict__ src_off,
                                                const int32_t *__restrict__ dim, const double *__restrict__ src,
                                                double *__restrict__ dst) {
  const int32_t j = blockIdx.y;
  const int32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= p * p) return;
  const int32_t r = e / p, c = e % p, db = dim[j];
  double v = 0.0;
  if (r < db && c < db) v = src[src_off[j] + (int64_t)r * db + c];
  else if (r == c) v = -1.0;
  dst[(int64_t)j * p * p + e] = v;
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

// Device matrices of blocks [b0, b1) (packed row-major blocks, block b at out_off[b - b0]).
void density_device(const DMRG_BASIS &B, int32_t b0, int32_t b1, const int32_t *Dim_Block, const double *Vec,
                    DevBuf<double> &dout, std::vector<int64_t> &out_off) {
  std::vector<int32_t> idx;
  const std::vector<DmTile> tiles = dm_plan(B, b0, b1, Dim_Block, idx, out_off);
  const int64_t n_out = out_off.back(), n_psi = (int64_t)idx.size();
  Context &C = ctx();
  hipStream_t st = C.stream;
  const int32_t n = B.dim_LLLRRRRL;
  DevBuf<double> vec(n), psi(n_psi > 0 ? n_psi : 1);
  DevBuf<int32_t> didx(n_psi > 0 ? n_psi : 1);
  DevBuf<DmTile> dt(tiles.size());
  dout.resize(n_out > 0 ? n_out : 1);
  C.copy_in_many({{vec.p, Vec, sizeof(double) * n},
                  {didx.p, idx.data(), sizeof(int32_t) * (size_t)std::max<int64_t>(n_psi, 0)},
                  {dt.p, tiles.data(), sizeof(DmTile) * tiles.size()}});
  if (n_psi > 0) k_dm_gather<<<(unsigned)((n_psi + 255) / 256), 256, 0, st>>>(n_psi, didx.p, vec.p, psi.p);
  // an empty partner list leaves its block's matrix zero, as the reference's does
  k_dm_syrk<<<(unsigned)tiles.size(), 256, 0, st>>>(dt.p, psi.p, dout.p);
  HIPCHK(hipGetLastError());
  // the temporaries go back to the allocation cache only after the kernels that read them
  HIPCHK(hipStreamSynchronize(st));
}

// Host copy of the matrices of blocks [b0, b1) at out.
void density_blocks(const DMRG_BASIS &B, int32_t b0, int32_t b1, const int32_t *Dim_Block, const double *Vec,
                    double *out) {
  DevBuf<double> dout;
  std::vector<int64_t> out_off;
  density_device(B, b0, b1, Dim_Block, Vec, dout, out_off);
  ctx().copy_out(out, dout.p, sizeof(double) * out_off.back());
}

// Side streams for the batched eigensolves (default priority), each with its own rocBLAS
// handle and workspace, created once per process; the batches of one call run concurrently
// and the library stream waits for all of them through events.
constexpr int kSolverStreams = 8;
struct SolverLane {
  hipStream_t stream = nullptr;
  rocblas_handle handle = nullptr;
  hipEvent_t done = nullptr;
};

// A lane is published (stored in the table) only once its stream, event and handle all
// exist and the handle is bound to the stream; a failure part way leaves the slot empty
// and releases what was made.
SolverLane &solver_lane(int i) {
  static SolverLane lanes[kSolverStreams];
  SolverLane &L = lanes[i];
  if (L.handle) return L;
  SolverLane T;
  auto release = [&T]() {
    if (T.handle) (void)rocblas_destroy_handle(T.handle);
    if (T.done) (void)hipEventDestroy(T.done);
    if (T.stream) (void)hipStreamDestroy(T.stream);
  };
  if (hipStreamCreateWithFlags(&T.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&T.done, hipEventDisableTiming) != hipSuccess) {
    release();
    throw Failure(DMRGX_ERR_HIP, "solver stream creation failed");
  }
  if (rocblas_create_handle(&T.handle) != rocblas_status_success) {
    T.handle = nullptr;
    release();
    throw Failure(DMRGX_ERR_HIP, "rocblas_create_handle failed");
  }
  if (rocblas_set_stream(T.handle, T.stream) != rocblas_status_success) {
    release();
    throw Failure(DMRGX_ERR_HIP, "rocblas_set_stream failed");
  }
  L = T;
  return L;
}

// Waits for every side stream used so far when it goes out of scope, whatever the exit
// path, so that no batched solve still runs on buffers the caller's scope releases.
struct LaneGuard {
  size_t used = 0;
  ~LaneGuard() {
    for (size_t k = 0; k < std::min(used, (size_t)kSolverStreams); k++)
      (void)hipStreamSynchronize(solver_lane((int)k).stream);
  }
};

// 0: blocks up to kBlockEigMax through the library's own eigensolver; 1: every block
// through rocSOLVER (dmrgx_set_block_eigensolver)
std::atomic<int> g_block_eig_mode{0};
int block_eig_mode() { return g_block_eig_mode.load(std::memory_order_relaxed); }

int32_t padded_block_size(int32_t d) {
  if (d <= 16) return 16;
  if (d <= 512) return (d + 31) / 32 * 32;
  return (d + 63) / 64 * 64;
}

// One eigenpair of a block, as DMRG_GET_SYSTEM_INFO_Q2.c:33-46 lists them.
struct Pair {
  double val;
  int32_t block, basis, q1, q2;
};

template <class T>
T *host_array(int64_t n) {   // GET_ARRAY_*: zeroed, freed with free()
  T *p = static_cast<T *>(std::calloc((size_t)(n > 0 ? n : 1), sizeof(T)));
  if (!p) throw Failure(DMRGX_ERR_NOMEM, "host allocation failed");
  return p;
}

DMRG_SYSTEM_INFO *system_info(int32_t n_qn, const int32_t *Q1, const int32_t *Q2, const DMRG_BASIS &B,
                              const double *Vec, int32_t max_dim) {
  const int32_t dim_LLLR = B.dim_LLLR;
  if (n_qn != 1 && n_qn != 2) throw Failure(DMRGX_ERR_ARG, "n_qn must be 1 or 2");
  if (dim_LLLR <= 0) throw Failure(DMRGX_ERR_ARG, "dim_LLLR must be positive");
  if (!Q1 || (n_qn == 2 && !Q2) || !Vec) throw Failure(DMRGX_ERR_ARG, "quantum numbers or vector missing");
  auto q2 = [&](int32_t s) { return n_qn == 2 ? Q2[s] : 0; };
  // DMRG_GET_DENSITY_MATRIX_INFO_Q1/Q2: order checks, then the blocks in ascending
  // (Q_Number1, Q_Number2) order
  // (DMRG_GET_DENSITY_MATRIX_INFO_Q2.c:18-35): pairwise order, and the number of runs of
  // equal (Q1, Q2) must equal the number of distinct pairs (:69-74), so that each block's
  // states are contiguous
  int32_t runs = 1;
  for (int32_t s = 0; s + 1 < dim_LLLR; s++) {
    const bool bad = n_qn == 1 ? Q1[s] > Q1[s + 1]
                               : ((Q1[s] > Q1[s + 1] && q2(s) == q2(s + 1)) || (Q1[s] == Q1[s + 1] && q2(s) > q2(s + 1)));
    if (bad) throw Failure(DMRGX_ERR_ARG, "LLLR quantum numbers out of order");
    if (Q1[s] != Q1[s + 1] || q2(s) != q2(s + 1)) runs++;
  }
  std::map<std::pair<int32_t, int32_t>, int32_t> count;
  for (int32_t s = 0; s < dim_LLLR; s++) count[{Q1[s], q2(s)}]++;
  if ((size_t)runs != count.size())
    throw Failure(DMRGX_ERR_ARG, "LLLR quantum-number blocks not contiguous (runs != distinct pairs)");
  std::vector<int32_t> dim_block, qb1, qb2;
  for (const auto &kv : count) {
    dim_block.push_back(kv.second);
    qb1.push_back(kv.first.first);
    qb2.push_back(kv.first.second);
  }
  const int32_t nb = (int32_t)dim_block.size();
  DevBuf<double> dm;
  std::vector<int64_t> off;
  density_device(B, 0, nb, dim_block.data(), Vec, dm, off);
  // eigenpairs of every block in place (LAPACK_DSYEV: ascending eigenvalues; rows of the
  // row-major result are the eigenvectors, the matrix being symmetric)
  Context &C = ctx();
  std::vector<int64_t> row_base(nb + 1, 0);
  for (int32_t b = 0; b < nb; b++) row_base[b + 1] = row_base[b] + dim_block[b];
  // per block: eigenvalues ascending (hw, at row_base); the eigenvectors the transformation
  // matrix takes: vec[row_base[b] + i] points at eigenvector i of block b (rows of hv for
  // the rocSOLVER blocks, of X for the own solver's kept states)
  std::vector<double> hw(dim_LLLR), hv;
  std::vector<const double *> vec(dim_LLLR, nullptr);
  // rocSOLVER for the listed blocks, by batches of equal padded size: one batched solve per
  // size instead of one solve per block; sizes padded to 16, then to multiples of 32 up to
  // 512 and of 64 above, so that a block just past a size step costs at most ~25 % more
  // work, not the 8x of a power-of-two step
  auto library_solve = [&](const std::vector<int32_t> &list) {
    std::map<int32_t, std::vector<int32_t>> buckets;
    for (int32_t b : list) buckets[padded_block_size(dim_block[b])].push_back(b);
    struct Batch {
      int32_t p;
      std::vector<int32_t> blocks;
      DevBuf<double> a, w, e;
      DevBuf<rocblas_int> info;
    };
    std::vector<std::unique_ptr<Batch>> batches;
    // every batch padded on the library stream first
    {
      std::vector<std::unique_ptr<DevBuf<int64_t>>> keep_so;
      std::vector<std::unique_ptr<DevBuf<int32_t>>> keep_sd;
      for (const auto &kv : buckets) {
        auto Bt = std::make_unique<Batch>();
        Bt->p = kv.first;
        Bt->blocks = kv.second;
        const int32_t p = Bt->p, cnt = (int32_t)Bt->blocks.size();
        std::vector<int64_t> so(cnt);
        std::vector<int32_t> sd(cnt);
        for (int32_t j = 0; j < cnt; j++) {
          so[j] = off[Bt->blocks[j]];
          sd[j] = dim_block[Bt->blocks[j]];
        }
        keep_so.push_back(std::make_unique<DevBuf<int64_t>>(cnt));
        keep_sd.push_back(std::make_unique<DevBuf<int32_t>>(cnt));
        C.copy_in(keep_so.back()->p, so.data(), sizeof(int64_t) * cnt);
        C.copy_in(keep_sd.back()->p, sd.data(), sizeof(int32_t) * cnt);
        Bt->a.resize((int64_t)cnt * p * p);
        Bt->w.resize((int64_t)cnt * p);
        Bt->e.resize((int64_t)cnt * p);
        Bt->info.resize(cnt);
        k_dm_pad<<<dim3((unsigned)((p * p + 255) / 256), (unsigned)cnt), 256, 0, C.stream>>>(
            p, keep_so.back()->p, keep_sd.back()->p, dm.p, Bt->a.p);
        HIPCHK(hipGetLastError());
        batches.push_back(std::move(Bt));
      }
      // the offset and size arrays go back to the allocation cache after the padding kernels
      HIPCHK(hipStreamSynchronize(C.stream));
    }
    // then the batched eigensolves side by side, batch i on side stream i mod kSolverStreams
    LaneGuard lanes_used;   // declared after batches: destroyed (streams drained) before them
    for (size_t i = 0; i < batches.size(); i++) {
      Batch &Bt = *batches[i];
      SolverLane &L = solver_lane((int)(i % kSolverStreams));
      lanes_used.used = i + 1;
      const int32_t p = Bt.p, cnt = (int32_t)Bt.blocks.size();
      if (rocsolver_dsyevd_strided_batched(L.handle, rocblas_evect_original, rocblas_fill_upper, p, Bt.a.p, p,
                                           (rocblas_stride)p * p, Bt.w.p, p, Bt.e.p, p, Bt.info.p,
                                           cnt) != rocblas_status_success)
        throw Failure(DMRGX_ERR_HIP, "rocsolver_dsyevd_strided_batched failed");
      HIPCHK(hipEventRecord(L.done, L.stream));
      HIPCHK(hipStreamWaitEvent(C.stream, L.done, 0));
    }
    hv.resize(off.back());
    for (const auto &Bt : batches) {
      const int32_t p = Bt->p, cnt = (int32_t)Bt->blocks.size();
      std::vector<double> ha((int64_t)cnt * p * p), hwp((int64_t)cnt * p);
      std::vector<rocblas_int> hinfo(cnt);
      C.copy_out(hinfo.data(), Bt->info.p, sizeof(rocblas_int) * cnt);
      C.copy_out(hwp.data(), Bt->w.p, sizeof(double) * cnt * p);
      C.copy_out(ha.data(), Bt->a.p, sizeof(double) * cnt * p * p);
      for (int32_t j = 0; j < cnt; j++) {
        if (hinfo[j] != 0) throw Failure(DMRGX_ERR_HIP, "block eigendecomposition did not converge");
        const int32_t b = Bt->blocks[j], db = dim_block[b], pad = p - db;
        const double *wj = hwp.data() + (int64_t)j * p, *aj = ha.data() + (int64_t)j * p * p;
        // the pad eigenvalues (-1) come first; eigenvector i of the result is row i of the
        // row-major view (column i of the column-major output)
        for (int32_t i = 0; i < pad; i++)
          if (wj[i] > -0.5) throw Failure(DMRGX_ERR_HIP, "padded eigenvalues out of place");
        for (int32_t i = 0; i < db; i++) {
          hw[row_base[b] + i] = wj[pad + i];
          std::memcpy(hv.data() + off[b] + (int64_t)i * db, aj + (int64_t)(pad + i) * p, sizeof(double) * db);
          vec[row_base[b] + i] = hv.data() + off[b] + (int64_t)i * db;
        }
      }
    }
  };
  // DMRG_GET_SYSTEM_INFO_Q2.c:33-112: every eigenpair in stable descending order, then the
  // cut, moved down past (nearly) equal values
  std::vector<Pair> p;
  int32_t dim_renorm = 0;
  const double zero = 1e-15;
  auto select = [&]() {
    p.clear();
    p.reserve(dim_LLLR);
    for (int32_t b = 0; b < nb; b++)
      for (int32_t i = 0; i < dim_block[b]; i++) p.push_back(Pair{hw[row_base[b] + i], b, i, qb1[b], qb2[b]});
    std::stable_sort(p.begin(), p.end(), [](const Pair &a, const Pair &c) { return a.val > c.val; });
    dim_renorm = 0;
    if (dim_LLLR <= max_dim) {
      for (int32_t i = dim_LLLR; i >= 1; i--)
        if (p[i - 1].val > zero) {
          dim_renorm = i;
          break;
        }
    } else {
      for (int32_t i = max_dim; i >= 1; i--)
        if (std::fabs(p[i - 1].val - p[i].val) > zero && p[i - 1].val > zero) {
          dim_renorm = i;
          break;
        }
    }
  };
  // blocks up to kBlockEigMax: the library's own solver (blockeig.hip), eigenvalues of every
  // block first, eigenvectors of the kept states only; larger blocks (and every block when
  // the own solver is switched off): rocSOLVER
  std::vector<int32_t> small, large;
  for (int32_t b = 0; b < nb; b++)
    (dim_block[b] <= kBlockEigMax && block_eig_mode() == 0 ? small : large).push_back(b);
  BlockEig own;
  if (!small.empty()) {
    std::vector<int32_t> sd(small.size());
    std::vector<int64_t> so(small.size());
    for (size_t j = 0; j < small.size(); j++) {
      sd[j] = dim_block[small[j]];
      so[j] = off[small[j]];
    }
    own.values((int32_t)small.size(), sd.data(), so.data(), dm.p, C.stream);
    std::vector<double> ws(own.r_off.back());
    C.copy_out(ws.data(), own.Wg.p, sizeof(double) * ws.size());
    for (size_t j = 0; j < small.size(); j++)
      std::memcpy(hw.data() + row_base[small[j]], ws.data() + own.r_off[j], sizeof(double) * sd[j]);
  }
  if (!large.empty()) library_solve(large);
  select();
  std::vector<double> X;   // the own solver's kept eigenvectors (rows), pointed at by vec
  std::vector<int64_t> x_off;
  if (!small.empty()) {
    std::vector<int32_t> slot(nb, -1);
    for (size_t j = 0; j < small.size(); j++) slot[small[j]] = (int32_t)j;
    std::vector<std::vector<int32_t>> kept(small.size());
    for (int32_t i = 0; i < dim_renorm; i++)
      if (slot[p[i].block] >= 0) kept[slot[p[i].block]].push_back(p[i].basis);
    for (auto &k : kept) std::sort(k.begin(), k.end());
    if (own.vectors(kept, X, x_off, C.stream)) {
      for (size_t j = 0; j < small.size(); j++) {
        const int32_t b = small[j], db = dim_block[b];
        for (size_t t = 0; t < kept[j].size(); t++)
          vec[row_base[b] + kept[j][t]] = X.data() + x_off[j] + (int64_t)t * db;
      }
    } else {   // a kept vector came out dependent on the others: rocSOLVER for these blocks
      library_solve(small);
      select();
    }
  }
  DMRG_SYSTEM_INFO *S = static_cast<DMRG_SYSTEM_INFO *>(std::calloc(1, sizeof(DMRG_SYSTEM_INFO)));
  if (!S) throw Failure(DMRGX_ERR_NOMEM, "host allocation failed");
  std::unique_ptr<DMRG_SYSTEM_INFO, void (*)(DMRG_SYSTEM_INFO *)> guard(S, dmrgx_free_system_info);
  S->Val_DM_Dist = host_array<double>(dim_LLLR);
  S->sum_val_dm = 0;
  for (int32_t i = 0; i < dim_LLLR; i++) {
    S->Val_DM_Dist[i] = p[i].val;
    S->sum_val_dm = S->sum_val_dm + p[i].val;
  }
  if (dim_renorm <= 0) throw Failure(DMRGX_ERR_ARG, "no state kept (dim_renorm <= 0)");
  double kept = 0;
  for (int32_t i = 0; i < dim_renorm; i++) kept = kept + p[i].val;
  S->tr_error = S->sum_val_dm - kept;
  // :124-158: kept states by Q_Number2, then by Q_Number1, both stable
  std::stable_sort(p.begin(), p.begin() + dim_renorm, [](const Pair &a, const Pair &c) { return a.q2 < c.q2; });
  std::stable_sort(p.begin(), p.begin() + dim_renorm, [](const Pair &a, const Pair &c) { return a.q1 < c.q1; });
  // :159-205: transformation matrix (CCS1) and its transpose (CRS1) from the eigenvectors
  int64_t nnz = 0;
  for (int32_t i = 0; i < dim_renorm; i++) nnz += dim_block[p[i].block];
  CCS1 *T = static_cast<CCS1 *>(std::calloc(1, sizeof(CCS1)));
  if (!T) throw Failure(DMRGX_ERR_NOMEM, "host allocation failed");
  S->Trans_Matrix = T;
  T->max_val = nnz;
  T->max_col = dim_renorm + 1;
  T->Val = host_array<double>(nnz);
  T->Col = host_array<long>(dim_renorm + 1);
  T->Row = host_array<int>(nnz);
  CRS1 *Td = static_cast<CRS1 *>(std::calloc(1, sizeof(CRS1)));
  if (!Td) throw Failure(DMRGX_ERR_NOMEM, "host allocation failed");
  S->Trans_Matrix_Dagger = Td;
  Td->max_val = nnz;
  Td->max_row = dim_renorm + 1;
  Td->Val = host_array<double>(nnz);
  Td->Col = host_array<int>(nnz);
  Td->Row = host_array<long>(dim_renorm + 1);
  S->Q_Number1 = host_array<int>(dim_renorm);
  S->Q_Number2 = host_array<int>(n_qn == 2 ? dim_renorm : 1);
  S->Q_Number3 = host_array<int>(1);
  S->dim_renorm = dim_renorm;
  S->dim_LLLR = dim_LLLR;
  for (int32_t i = 0; i < dim_renorm; i++) {
    S->Q_Number1[i] = p[i].q1;
    if (n_qn == 2) S->Q_Number2[i] = p[i].q2;
  }
  if (n_qn == 1) S->Q_Number2[0] = INT_MIN;
  S->Q_Number3[0] = INT_MIN;
  int64_t k = 0;
  for (int32_t i = 0; i < dim_renorm; i++) {
    const int32_t b = p[i].block, db = dim_block[b];
    const double *v = vec[row_base[b] + p[i].basis];
    for (int32_t j = 0; j < db; j++, k++) {
      T->Val[k] = v[j];
      T->Row[k] = (int)(row_base[b] + j);
      Td->Val[k] = v[j];
      Td->Col[k] = (int)(row_base[b] + j);
    }
    T->Col[i + 1] = k;
    Td->Row[i + 1] = k;
  }
  T->row_dim = dim_LLLR;
  T->col_dim = dim_renorm;
  Td->row_dim = dim_renorm;
  Td->col_dim = dim_LLLR;
  guard.release();
  return S;
}

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

extern "C" {

int dmrgx_set_block_eigensolver(int mode) {
  return run_status([&] {
    if (mode != 0 && mode != 1) throw Failure(DMRGX_ERR_ARG, "block eigensolver mode must be 0 or 1");
    g_block_eig_mode.store(mode, std::memory_order_relaxed);
  });
}

void dmrgx_free_system_info(DMRG_SYSTEM_INFO *Info) {
  if (!Info) return;
  if (Info->Trans_Matrix) {
    std::free(Info->Trans_Matrix->Val);
    std::free(Info->Trans_Matrix->Col);
    std::free(Info->Trans_Matrix->Row);
    std::free(Info->Trans_Matrix);
  }
  if (Info->Trans_Matrix_Dagger) {
    std::free(Info->Trans_Matrix_Dagger->Val);
    std::free(Info->Trans_Matrix_Dagger->Col);
    std::free(Info->Trans_Matrix_Dagger->Row);
    std::free(Info->Trans_Matrix_Dagger);
  }
  std::free(Info->Val_DM_Dist);
  std::free(Info->Q_Number1);
  std::free(Info->Q_Number2);
  std::free(Info->Q_Number3);
  std::free(Info);
}

int dmrgx_get_system_info(int32_t n_qn, const int32_t *Q_Number1, const int32_t *Q_Number2, const DMRG_BASIS *B,
                          const double *Vec, int32_t max_dim, DMRG_SYSTEM_INFO **out) {
  if (!B || !out) return DMRGX_ERR_ARG;
  *out = nullptr;
  return run_status([&] { *out = system_info(n_qn, Q_Number1, Q_Number2, *B, Vec, max_dim); });
}

DMRG_SYSTEM_INFO *DMRG_GET_SYSTEM_INFO_Q1(int *Q_Number, DMRG_BASIS *Dmrg_Basis, double *Vec, int max_dim,
                                          int p_threads, DMRG_TIME *Dmrg_Time) {
  (void)p_threads;
  DMRG_SYSTEM_INFO *S = nullptr;
  const double t0 = now_s();
  run_or_die("DMRG_GET_SYSTEM_INFO_Q1", [&] {
    if (!Dmrg_Basis) throw Failure(DMRGX_ERR_ARG, "basis missing");
    S = system_info(1, Q_Number, nullptr, *Dmrg_Basis, Vec, max_dim);
  });
  if (Dmrg_Time) Dmrg_Time->make_dens_mat = now_s() - t0;
  return S;
}

DMRG_SYSTEM_INFO *DMRG_GET_SYSTEM_INFO_Q2(int *Q_Number1, int *Q_Number2, DMRG_BASIS *Dmrg_Basis, double *Vec,
                                          int max_dim, int p_threads, DMRG_TIME *Dmrg_Time) {
  (void)p_threads;
  DMRG_SYSTEM_INFO *S = nullptr;
  const double t0 = now_s();
  run_or_die("DMRG_GET_SYSTEM_INFO_Q2", [&] {
    if (!Dmrg_Basis) throw Failure(DMRGX_ERR_ARG, "basis missing");
    S = system_info(2, Q_Number1, Q_Number2, *Dmrg_Basis, Vec, max_dim);
  });
  if (Dmrg_Time) Dmrg_Time->make_dens_mat = now_s() - t0;
  return S;
}

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
