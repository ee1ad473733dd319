// Eigenpairs of the block density matrices of a truncation step, for blocks up to
// kBlockEigMax (dmrg/DMRG_GET_SYSTEM_INFO_Q2.c:22-46 calls LAPACK's dsyev per block through
// sml/LAPACK_DSYEV.c and keeps the eigenvectors of the largest eigenvalues only).  The route
// is LAPACK's dsyevx one (dsytrd, dstebz, dstein, dormtr), with eigenvectors formed for the
// kept states only, in two launches around the host's selection:
//  1. k_eig_values, one workgroup per block.  The block's lower triangle, packed, in LDS;
//     Householder reduction to tridiagonal form, last row first: for i = n-1 .. 2 the
//     reflector H_i = I - v v^T / h, v from row i's first i entries, zeroes them below the
//     subdiagonal, and the leading i x i part B becomes B - v w^T - w v^T with p = B v / h,
//     w = p - (v.p / 2h) v; v stays in row i.  Out: the reflectors, T's diagonal and
//     subdiagonal, its Gershgorin interval.  Then k_eig_bisect, one wavefront per
//     eigenvalue over all blocks: bisection on the Sturm count (pivots of the LDL^T
//     factorisation of T - x I, floored at LAPACK's pivmin as in dstebz), 64 counts per
//     step, to an absolute width of 2 eps ||T||: the eigenvalues ascending.
//  2. k_eig_vectors, one workgroup per block with kept states.  Per kept eigenvalue (one
//     thread each): three steps of inverse iteration on T - lambda I (Gaussian elimination
//     with partial pivoting, pivots below eps ||T|| raised to it, as LAPACK's dlagtf/dlagts;
//     the pivots kept as reciprocals),
//     from a fixed pseudo-random start; modified Gram-Schmidt over the block's kept vectors
//     in ascending eigenvalue order (a second pass when a vector lost more than half its
//     norm to it), one wavefront per vector; then k_eig_back: x = H_{n-1} ... H_2 y, one
//     wavefront per eight vectors, the reflectors in LDS.
// Eigenvalues agree with dsyev's to ~eps ||block||; eigenvectors up to sign, and up to a
// rotation within a set of (nearly) equal eigenvalues.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace dmrgx {
namespace {

constexpr int kEigThreads = 256;
constexpr int kLdsDoubles = 160 * 1024 / 8;             // LDS of one workgroup, in doubles
constexpr int kVecFixed = 4 * kBlockEigMax + 8;         // stage 2's LDS besides the vectors (d, e, nrm, onrm, red)
constexpr int64_t kVecLdsMax = kLdsDoubles - kVecFixed;  // vectors (n x k) held in LDS up to this
static_assert(kBlockEigMax * (kBlockEigMax + 1) / 2 + kBlockEigMax <= kLdsDoubles, "reflectors must fit in LDS");

__device__ __forceinline__ int tri(int r) { return r * (r + 1) / 2; }

// max of v over the workgroup (every thread gets it); red: kEigThreads / 64 doubles
template <int kThreads = kEigThreads>
__device__ double block_max_all(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __syncthreads();
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  double t = red[0];
  for (int k = 1; k < kThreads / 64; k++) t = fmax(t, red[k]);
  return t;
}

template <int kCtrl>
__device__ __forceinline__ double dpp(double x) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)v, kCtrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), kCtrl, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// sum over the 16 lanes of a row of the wavefront (every lane of the row gets it): lane
// pairs l^1, l^2, then the mirrors within 8 and within 16 lanes
__device__ __forceinline__ double sum16(double x) {
  x += dpp<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += dpp<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += dpp<0x141>(x);   // row_half_mirror
  x += dpp<0x140>(x);   // row_mirror
  return x;
}

// the value of x in lane l ^ 32 (l ^ 16): the gfx950 lane swaps, which exchange the upper
// half of their first operand with the lower half of their second (rows 1 and 3 with rows 0
// and 2 of 16 lanes); both operands x, the partner's value is the second result below the
// swapped boundary and the first above it
__device__ __forceinline__ double swap32(double x) {
  const long long v = __double_as_longlong(x);
  const auto lo = __builtin_amdgcn_permlane32_swap((int)v, (int)v, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((int)(v >> 32), (int)(v >> 32), false, false);
  const bool up = threadIdx.x & 32;
  const int l = up ? lo[0] : lo[1], h = up ? hi[0] : hi[1];
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)h << 32) | (unsigned)l));
}

__device__ __forceinline__ double swap16(double x) {
  const long long v = __double_as_longlong(x);
  const auto lo = __builtin_amdgcn_permlane16_swap((int)v, (int)v, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((int)(v >> 32), (int)(v >> 32), false, false);
  const bool up = threadIdx.x & 16;
  const int l = up ? lo[0] : lo[1], h = up ? hi[0] : hi[1];
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)h << 32) | (unsigned)l));
}

// sum over the wavefront, every lane gets it (all 64 lanes active)
__device__ __forceinline__ double wave_sum(double v) {
  v = sum16(v);
  v += swap16(v);
  v += swap32(v);
  return v;
}

// Eight sums over the wavefront at once: in, lane l's partial sums of eight quantities; out,
// the full sum of quantity ((l >> 5) & 1) * 4 + ((l >> 4) & 1) * 2 + ((l >> 3) & 1), held by
// lane l (lanes 32 a + 16 b + 8 c + 0..7 hold quantity 4 a + 2 b + c).  The set is halved at
// the first three exchanges (lanes l ^ 32, l ^ 16, and l with 15 - l in a row of 16), the
// rest summed within 8 lanes (all 64 lanes active).
__device__ __forceinline__ double wave_sum8(const double *x, int lane) {
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
  double e4[4], e2[2];
#pragma unroll
  for (int q = 0; q < 4; q++) e4[q] = (b5 ? x[q + 4] : x[q]) + swap32(b5 ? x[q] : x[q + 4]);
#pragma unroll
  for (int q = 0; q < 2; q++) e2[q] = (b4 ? e4[q + 2] : e4[q]) + swap16(b4 ? e4[q] : e4[q + 2]);
  double e1 = (b3 ? e2[1] : e2[0]) + dpp<0x140>(b3 ? e2[0] : e2[1]);   // row_mirror: l and 15 - l
  e1 += dpp<0xB1>(e1);
  e1 += dpp<0x4E>(e1);
  e1 += dpp<0x141>(e1);   // row_half_mirror: l and 7 - l
  return e1;
}

// the value of x in lane l (uniform l)
__device__ __forceinline__ double read_lane(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// 1 / q to ~1 ulp: the hardware estimate and one Newton step (|q| >= the pivot floor, so
// no overflow)
__device__ __forceinline__ double recip(double q) {
  double r = __builtin_amdgcn_rcp(q);
  const double e = fma(-q, r, 1.0);
  return fma(r, e, r);
}

struct ValJob {
  int32_t n;
  int64_t a_off;   // the block's row-major matrix in A
  int64_t p_off;   // its packed lower triangle (reflectors) in H: tri(n) entries
  int64_t r_off;   // its first row in D, E, HV, W
};

// Stage 1's threads as a 32 x 16 grid, (tr, tc) = (tid / 16, tid % 16): thread (tr, tc)
// holds the block's entries (tr + 32 a, tc + 16 b), b <= 2 a + 1, in registers, slot
// a (a + 1) + b: below the diagonal for b < 2 a, and for b = 2 a (b = 2 a + 1) when tr >= tc
// (tr >= tc + 16); the 16 threads of one tr are one row of 16 lanes of a wavefront.
constexpr int kValThreads = 512;
constexpr int kGridR = 32, kGridC = 16;
constexpr int kSlabs = kBlockEigMax / kGridR;          // row slabs (column slabs: 2 kSlabs)
constexpr int kSlots = kSlabs * (kSlabs + 1);
static_assert(kBlockEigMax % kGridR == 0 && kGridR * kGridC == kValThreads, "stage 1 thread grid");
__host__ __device__ constexpr int slot(int a, int b) { return a * (a + 1) + b; }

__global__ __launch_bounds__(kValThreads) void k_eig_values(const ValJob *jobs, const double *__restrict__ A,
                                                            double *__restrict__ H, double *__restrict__ D,
                                                            double *__restrict__ E, double *__restrict__ HV,
                                                            double *__restrict__ BND) {
  extern __shared__ double sm[];
  const ValJob J = jobs[blockIdx.x];
  const int n = J.n, tid = threadIdx.x, lane = tid & 63;
  const int tr = tid / kGridC, tc = tid % kGridC, na = (n + kGridR - 1) / kGridR;
  // vb: row i of the current step (two buffers, alternating); pr, pc: partial products
  double *d = sm, *e = d + n, *hv = e + n, *vb = hv + n, *pw = vb + 2 * n, *pr = pw + n, *pc = pr + n,
         *red = pc + kGridR * n;
  double Lr[kSlots];
#pragma unroll
  for (int a = 0; a < kSlabs; a++)
#pragma unroll
    for (int b = 0; b <= 2 * a + 1; b++) {
      const int r = tr + kGridR * a, c = tc + kGridC * b;
      Lr[slot(a, b)] = (a < na && r < n && c <= r) ? A[J.a_off + (int64_t)r * n + c] : 0.0;
    }
  // weights of the slots b = 2 a and b = 2 a + 1: the diagonal and below in the row
  // products, strictly below in the column products (else above the diagonal, not an entry)
  const double fd0 = tr >= tc ? 1.0 : 0.0, fs0 = tr > tc ? 1.0 : 0.0;
  const double fd1 = tr >= tc + kGridC ? 1.0 : 0.0, fs1 = tr > tc + kGridC ? 1.0 : 0.0;
  for (int i = n - 1; i >= 2; i--) {
    double *vv = vb + (i & 1) * n;
    const int ai = i / kGridR, ti = i % kGridR;
    if (tr == ti) {   // row i's entries c < i, from the threads holding them
#pragma unroll
      for (int a = 0; a < kSlabs; a++)
        if (a == ai)
#pragma unroll
          for (int b = 0; b <= 2 * a + 1; b++) {
            const int c = tc + kGridC * b;
            if (c < i) vv[c] = Lr[slot(a, b)];
          }
    }
    __syncthreads();
    double sg = 0.0;
    for (int k = lane; k < i - 1; k += 64) sg += vv[k] * vv[k];
    const double sig = wave_sum(sg);   // every wavefront, from LDS
    const double xl = vv[i - 1];
    if (sig == 0.0) {   // row i already tridiagonal
      if (tid == 0) {
        e[i] = xl;
        hv[i] = 0.0;
      }
      continue;
    }
    const double alpha = xl >= 0.0 ? -sqrt(sig + xl * xl) : sqrt(sig + xl * xl);
    const double h = alpha * alpha - xl * alpha, vl = xl - alpha;
    if (tid == 0) {
      e[i] = alpha;
      hv[i] = h;
    }
    // v at this thread's rows and columns: row i's entries, v_{i-1} = xl - alpha, zero from i on
    double vr[kSlabs], vc[2 * kSlabs];
#pragma unroll
    for (int a = 0; a < kSlabs; a++) {
      const int r = tr + kGridR * a;
      vr[a] = (a < na && r < i) ? (r == i - 1 ? vl : vv[r]) : 0.0;
    }
#pragma unroll
    for (int b = 0; b < 2 * kSlabs; b++) {
      const int c = tc + kGridC * b;
      vc[b] = c < i ? (c == i - 1 ? vl : vv[c]) : 0.0;
    }
    if (tr == ti && tc == (i - 1) % kGridC) {   // v kept in row i
#pragma unroll
      for (int a = 0; a < kSlabs; a++)
        if (a == ai)
#pragma unroll
          for (int b = 0; b <= 2 * a + 1; b++)
            if (b == (i - 1) / kGridC) Lr[slot(a, b)] = vl;
    }
    // p = B v / h, B = S + D + S^T: row sums over (D + S) and column sums over S, per thread,
    // then the rows' over the 16 threads of a row group and the columns' over the 32 groups
    double pcol[2 * kSlabs];
#pragma unroll
    for (int b = 0; b < 2 * kSlabs; b++) pcol[b] = 0.0;
#pragma unroll
    for (int a = 0; a < kSlabs; a++) {
      if (a >= na) continue;
      double prow = 0.0;
#pragma unroll
      for (int b = 0; b <= 2 * a + 1; b++) {
        const double x = Lr[slot(a, b)];
        const double wd = b < 2 * a ? 1.0 : (b == 2 * a ? fd0 : fd1);
        const double ws = b < 2 * a ? 1.0 : (b == 2 * a ? fs0 : fs1);
        prow += wd * x * vc[b];
        pcol[b] += ws * x * vr[a];
      }
      const double s = sum16(prow);
      const int r = tr + kGridR * a;
      if (tc == 0 && r < i) pr[r] = s;
    }
#pragma unroll
    for (int b = 0; b < 2 * kSlabs; b++) {
      const int c = tc + kGridC * b;
      if (c < i) pc[tr * n + c] = pcol[b];
    }
    __syncthreads();
    for (int j = tid; j < i; j += kValThreads) {
      double s = pr[j];
      for (int t = 0; t < kGridR; t++) s += pc[t * n + j];
      pw[j] = s / h;
    }
    __syncthreads();
    // K = v.p / 2h (every wavefront), w = p - K v at this thread's rows and columns
    double vp = 0.0;
    for (int k = lane; k < i; k += 64) vp += (k == i - 1 ? vl : vv[k]) * pw[k];
    const double K = wave_sum(vp) / (2.0 * h);
    double wc[2 * kSlabs];
#pragma unroll
    for (int b = 0; b < 2 * kSlabs; b++) {
      const int c = tc + kGridC * b;
      wc[b] = c < i ? pw[c] - K * vc[b] : 0.0;
    }
    // B -= v w^T + w v^T (rows and columns from i on: v = w = 0, unchanged)
#pragma unroll
    for (int a = 0; a < kSlabs; a++) {
      if (a >= na) continue;
      const int r = tr + kGridR * a;
      const double wr = r < i ? pw[r] - K * vr[a] : 0.0;
#pragma unroll
      for (int b = 0; b <= 2 * a + 1; b++) Lr[slot(a, b)] -= vr[a] * wc[b] + wr * vc[b];
    }
  }
  // T: diagonal d, e[k] = T[k][k-1] (k >= 1); the reflectors (rows) to H
#pragma unroll
  for (int a = 0; a < kSlabs; a++)
#pragma unroll
    for (int b = 0; b <= 2 * a + 1; b++) {
      const int r = tr + kGridR * a, c = tc + kGridC * b;
      if (a >= na || r >= n || c > r) continue;
      H[J.p_off + tri(r) + c] = Lr[slot(a, b)];
      if (r == c) d[r] = Lr[slot(a, b)];
      if (r == 1 && c == 0) e[1] = Lr[slot(a, b)];
    }
  if (tid == 0) {
    e[0] = 0.0;
    hv[0] = 0.0;
    if (n > 1) hv[1] = 0.0;
  }
  __syncthreads();
  for (int k = tid; k < n; k += kValThreads) {
    D[J.r_off + k] = d[k];
    E[J.r_off + k] = e[k];
    HV[J.r_off + k] = hv[k];
  }
  // Gershgorin interval, pivot floor and bisection width for k_eig_bisect
  double lo_k = DBL_MAX, hi_k = -DBL_MAX, e2max = 0.0;
  for (int k = tid; k < n; k += kValThreads) {
    const double r = fabs(e[k]) + (k + 1 < n ? fabs(e[k + 1]) : 0.0);
    lo_k = fmin(lo_k, d[k] - r);
    hi_k = fmax(hi_k, d[k] + r);
    e2max = fmax(e2max, e[k] * e[k]);
  }
  const double gl0 = -block_max_all<kValThreads>(-lo_k, red), gu0 = block_max_all<kValThreads>(hi_k, red);
  e2max = block_max_all<kValThreads>(e2max, red);
  if (tid == 0) {
    const double tnorm = fmax(fabs(gl0), fabs(gu0));
    const double pivmin = DBL_MIN * fmax(1.0, e2max);
    const double widen = 2.0 * DBL_EPSILON * tnorm * n + 2.0 * pivmin;
    double *bnd = BND + 4 * (int64_t)blockIdx.x;
    bnd[0] = gl0 - widen;
    bnd[1] = gu0 + widen;
    bnd[2] = pivmin;
    bnd[3] = fmax(2.0 * DBL_EPSILON * tnorm, DBL_MIN);
  }
}

// Every eigenvalue of the tridiagonals, one wavefront per eigenvalue (kEigThreads / 64 per
// workgroup): eigenvalue t of block j is the smallest x with more than t eigenvalues at or
// below it.  Each step counts at 64 equally spaced interior points of [lo, hi] (lane l at
// lo + (l + 1) (hi - lo) / 65, each lane its own pivot recurrence, pivots floored at the
// block's pivmin), and the first point whose count exceeds t and the one before it bracket
// the eigenvalue: ~9 steps to a width of 2 eps ||T||.
constexpr int kBisWaves = kEigThreads / 64;
__global__ __launch_bounds__(kEigThreads) void k_eig_bisect(const int32_t *__restrict__ EJ, int64_t total,
                                                            const ValJob *__restrict__ jobs,
                                                            const double *__restrict__ D, const double *__restrict__ E,
                                                            const double *__restrict__ BND, double *__restrict__ W) {
  __shared__ double sd[kBisWaves][kBlockEigMax], se[kBisWaves][kBlockEigMax];
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t g = (int64_t)blockIdx.x * kBisWaves + wv;   // this wavefront's eigenvalue
  const bool live = g < total;
  int32_t j = 0, n = 0, t = 0;
  if (live) {
    j = EJ[g];
    n = jobs[j].n;
    t = (int32_t)(g - jobs[j].r_off);
    for (int k = lane; k < n; k += 64) {
      sd[wv][k] = D[jobs[j].r_off + k];
      const double ek = E[jobs[j].r_off + k];
      se[wv][k] = ek * ek;
    }
  }
  __syncthreads();
  if (!live) return;
  const double *d = sd[wv], *e2 = se[wv];
  double lo = BND[4 * (int64_t)j], hi = BND[4 * (int64_t)j + 1];
  const double pivmin = BND[4 * (int64_t)j + 2], atol = BND[4 * (int64_t)j + 3];
  for (int it = 0; it < 64; it++) {
    const double w = hi - lo;
    if (w <= atol || !(lo + w * (1.0 / 65.0) > lo)) break;
    const double x = lo + w * ((double)(lane + 1) * (1.0 / 65.0));
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = q < 0.0;
    for (int k = 1; k < n; k++) {
      q = (d[k] - x) - e2[k] * recip(q);
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
    const unsigned long long above = __ballot(cnt > t);
    const int f = above ? __ffsll((long long)above) - 1 : 64;   // first point past eigenvalue t
    const double nlo = f > 0 ? __shfl(x, f - 1) : lo;
    const double nhi = f < 64 ? __shfl(x, f) : hi;
    lo = fmin(nlo, nhi);
    hi = fmax(nhi, lo);
  }
  if (lane == 0) W[jobs[j].r_off + t] = 0.5 * (lo + hi);
}

struct VecJob {
  int32_t n, k;
  int64_t p_off, r_off;   // as ValJob
  int64_t i_off;          // the block's k kept eigenvalue indices (ascending) in IDX
  int64_t x_off;          // its k eigenvectors, row-major k x n, in X
  int64_t w_off;          // workspace: 3 n k doubles in U, and n k in Y when not in LDS
};

// deterministic start vector entry in [-1, 1)
__device__ __forceinline__ double start_entry(uint32_t c, uint32_t idx) {
  uint32_t h = c * 0x9E3779B1u ^ (idx + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)(h >> 8) * (1.0 / 8388608.0) - 1.0;
}

template <bool kInLds>
__global__ __launch_bounds__(kEigThreads) void k_eig_vectors(const VecJob *jobs, const double *__restrict__ H,
                                                             const double *__restrict__ D, const double *__restrict__ E,
                                                             const double *__restrict__ HV, const double *__restrict__ W,
                                                             const int32_t *__restrict__ IDX, double *__restrict__ U,
                                                             double *__restrict__ Yg, double *__restrict__ X,
                                                             int *fail) {
  extern __shared__ double sm[];
  const VecJob J = jobs[blockIdx.x];
  const int n = J.n, k = J.k, tid = threadIdx.x;
  double *d = sm, *e = d + n, *nrm = e + n, *red = nrm + 2 * kBlockEigMax;   // nrm, onrm: kBlockEigMax each
  double *Y = kInLds ? red + 8 : Yg + J.w_off;   // Y[c * k + t]: component c of vector t
  double *Ut = U + 3 * J.w_off;                  // U[(3c + j) k + t]: row c of vector t's factor
  __shared__ int bad, again;
  for (int c = tid; c < n; c += kEigThreads) {
    d[c] = D[J.r_off + c];
    e[c] = E[J.r_off + c];
  }
  if (tid == 0) {
    bad = 0;
    again = 0;
  }
  double tn = 0.0;
  for (int c = tid; c < n; c += kEigThreads) tn = fmax(tn, fabs(d[c]) + fabs(e[c]) + (c + 1 < n ? fabs(e[c + 1]) : 0.0));
  const double tiny = fmax(DBL_EPSILON * block_max_all(tn, red), DBL_MIN);   // barrier: d, e complete
  // inverse iteration, one vector per thread
  for (int t = tid; t < k; t += kEigThreads) {
    const int idx = IDX[J.i_off + t];
    const double lam = W[J.r_off + idx];
    for (int c = 0; c < n; c++) Y[(int64_t)c * k + t] = start_entry((uint32_t)c, (uint32_t)idx);
    for (int it = 0; it < 3; it++) {
      double p0 = d[0] - lam, p1 = n > 1 ? e[1] : 0.0, r = Y[t];
      for (int c = 0; c + 1 < n; c++) {
        const double l = e[c + 1], a = d[c + 1] - lam, s = c + 2 < n ? e[c + 2] : 0.0;
        const double rn = Y[(int64_t)(c + 1) * k + t];
        double *u = Ut + (int64_t)3 * c * k + t;
        if (fabs(p0) >= fabs(l)) {
          const double rp = recip(fabs(p0) < tiny ? copysign(tiny, p0) : p0);
          const double m = l * rp;
          u[0] = rp;   // the pivot's reciprocal
          u[k] = p1;
          u[2 * k] = 0.0;
          Y[(int64_t)c * k + t] = r;
          p0 = a - m * p1;
          p1 = s;
          r = rn - m * r;
        } else {
          const double rl = recip(fabs(l) < tiny ? copysign(tiny, l) : l);
          const double m = p0 * rl;
          u[0] = rl;
          u[k] = a;
          u[2 * k] = s;
          Y[(int64_t)c * k + t] = rn;
          p0 = p1 - m * a;
          p1 = -m * s;
          r = r - m * rn;
        }
      }
      // back substitution, then scale to max |y| = 1
      double x1 = r * recip(fabs(p0) < tiny ? copysign(tiny, p0) : p0), x2 = 0.0;
      Y[(int64_t)(n - 1) * k + t] = x1;
      double big = fabs(x1);
      for (int c = n - 2; c >= 0; c--) {
        const double *u = Ut + (int64_t)3 * c * k + t;
        const double x = (Y[(int64_t)c * k + t] - u[k] * x1 - u[2 * k] * x2) * u[0];
        Y[(int64_t)c * k + t] = x;
        big = fmax(big, fabs(x));
        x2 = x1;
        x1 = x;
      }
      if (!(big > 0.0) || !(big <= DBL_MAX)) {
        atomicAdd(&bad, 1);
        break;
      }
      const double sc = 1.0 / big;
      for (int c = 0; c < n; c++) Y[(int64_t)c * k + t] *= sc;
    }
  }
  __syncthreads();
  if (bad) {
    if (tid == 0) atomicAdd(fail, 1);
    return;
  }
  // modified Gram-Schmidt in ascending eigenvalue order, one wavefront per vector (lane l
  // holds components l, l + 64, l + 128); nrm[j]: |y_j|^2 once y_j is final, onrm[t]: |y_t|^2
  // before the pass
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);   // wavefront index, uniform
  constexpr int kWaves = kEigThreads / 64;
  double *onrm = nrm + kBlockEigMax;
  auto load3 = [&](int t, double *y) {
    for (int r = 0; r < 3; r++) {
      const int c = lane + 64 * r;
      y[r] = c < n ? Y[(int64_t)c * k + t] : 0.0;
    }
  };
  auto store3 = [&](int t, const double *y) {
    for (int r = 0; r < 3; r++) {
      const int c = lane + 64 * r;
      if (c < n) Y[(int64_t)c * k + t] = y[r];
    }
  };
  for (int pass = 0; pass < 2; pass++) {
    for (int t = wv; t < k; t += kWaves) {
      double y[3];
      load3(t, y);
      const double s = wave_sum(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
      if (lane == 0) {
        onrm[t] = s;
        if (t == 0) nrm[0] = s;
      }
    }
    __syncthreads();
    for (int j = 0; j + 1 < k; j++) {
      const double nj = nrm[j];
      double yj[3];
      load3(j, yj);
      // this wavefront's vectors j + 1 + wv + 4 g, eight at a time
      for (int g0 = 0; j + 1 + wv + kWaves * g0 < k; g0 += 8) {
        double a[8][3], dt[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int t = j + 1 + wv + kWaves * (g0 + u);
          if (t < k) load3(t, a[u]);
          else a[u][0] = a[u][1] = a[u][2] = 0.0;
          dt[u] = yj[0] * a[u][0] + yj[1] * a[u][1] + yj[2] * a[u][2];
        }
        const double fl = wave_sum8(dt, lane) / nj;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int t = j + 1 + wv + kWaves * (g0 + u);
          const double f = read_lane(fl, (u >> 2) * 32 + ((u >> 1) & 1) * 16 + (u & 1) * 8);
          for (int r = 0; r < 3; r++) a[u][r] = a[u][r] - f * yj[r];
          if (t < k) store3(t, a[u]);
        }
        if (wv == 0 && g0 == 0) {   // vector j + 1 is final: its norm, and what the pass took
          const int t = j + 1;
          const double sa = wave_sum(a[0][0] * a[0][0] + a[0][1] * a[0][1] + a[0][2] * a[0][2]);
          if (lane == 0) {
            nrm[t] = sa;
            if (!(sa > 1e-20 * onrm[t])) atomicAdd(&bad, 1);
            else if (sa < 0.5 * onrm[t]) again = 1;
          }
        }
      }
      __syncthreads();
    }
    if (bad || !again) break;
    __syncthreads();
    if (tid == 0) again = 0;
  }
  if (bad) {
    if (tid == 0) atomicAdd(fail, 1);
    return;
  }
  // unit length, out as row t of the block's k x n result in X (k_eig_back then applies the
  // reflectors there)
  for (int t = wv; t < k; t += kWaves) {
    double y[3];
    load3(t, y);
    const double sc = 1.0 / sqrt(wave_sum(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]));
    for (int r = 0; r < 3; r++) {
      const int c = lane + 64 * r;
      if (c < n) X[J.x_off + (int64_t)t * n + c] = y[r] * sc;
    }
  }
}

// x = H_{n-1} ... H_2 y in place in X for kChunk of a block's kept vectors: one wavefront per
// workgroup, the block's reflectors and h in LDS, the vectors in registers over the
// reflectors (lane l: components l, l + 64, l + 128)
constexpr int kChunk = 8;   // wave_sum8
struct BackJob {
  int32_t job, t0;   // the block (index into the VecJob list) and its first vector
};

__global__ __launch_bounds__(64) void k_eig_back(const BackJob *__restrict__ bj, const VecJob *__restrict__ jobs,
                                                 const double *__restrict__ H, const double *__restrict__ HV,
                                                 double *__restrict__ X) {
  extern __shared__ double sm[];
  const BackJob B = bj[blockIdx.x];
  const VecJob J = jobs[B.job];
  const int n = J.n, k = J.k, lane = threadIdx.x, t0 = B.t0;
  double *Hs = sm, *hs = sm + tri(n);
  for (int q = lane; q < tri(n); q += 64) Hs[q] = H[J.p_off + q];
  for (int c = lane; c < n; c += 64) hs[c] = HV[J.r_off + c];
  __syncthreads();
  double y[kChunk][3];
#pragma unroll
  for (int u = 0; u < kChunk; u++)
    for (int r = 0; r < 3; r++) {
      const int c = lane + 64 * r;
      y[u][r] = (t0 + u < k && c < n) ? X[J.x_off + (int64_t)(t0 + u) * n + c] : 0.0;
    }
  for (int i = 2; i < n; i++) {
    const double h = hs[i];
    if (h == 0.0) continue;
    const double *v = Hs + tri(i);
    double vr[3];
    for (int r = 0; r < 3; r++) {
      const int c = lane + 64 * r;
      vr[r] = c < i ? v[c] : 0.0;
    }
    double dot[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; u++) dot[u] = vr[0] * y[u][0] + vr[1] * y[u][1] + vr[2] * y[u][2];
    const double fl = wave_sum8(dot, lane) / h;
#pragma unroll
    for (int u = 0; u < kChunk; u++) {
      const double f = read_lane(fl, (u >> 2) * 32 + ((u >> 1) & 1) * 16 + (u & 1) * 8);
      for (int r = 0; r < 3; r++) y[u][r] -= f * vr[r];
    }
  }
#pragma unroll
  for (int u = 0; u < kChunk; u++) {
    if (t0 + u >= k) continue;
    for (int r = 0; r < 3; r++) {
      const int c = lane + 64 * r;
      if (c < n) X[J.x_off + (int64_t)(t0 + u) * n + c] = y[u][r];
    }
  }
}

}  // namespace

void BlockEig::values(int32_t count, const int32_t *dim, const int64_t *a_off, const double *A, hipStream_t s) {
  nblk = count;
  n.assign(dim, dim + count);
  r_off.assign(count + 1, 0);
  p_off.assign(count + 1, 0);
  if (count <= 0) return;
  std::vector<ValJob> jobs(count);
  int32_t nmax = 1;
  for (int32_t j = 0; j < count; j++) {
    if (dim[j] < 1 || dim[j] > kBlockEigMax) throw Failure(DMRGX_ERR_ARG, "block eigensolver: dimension out of range");
    r_off[j + 1] = r_off[j] + dim[j];
    p_off[j + 1] = p_off[j] + (int64_t)dim[j] * (dim[j] + 1) / 2;
    jobs[j] = ValJob{dim[j], a_off[j], p_off[j], r_off[j]};
    nmax = std::max(nmax, dim[j]);
  }
  H.resize(p_off[count]);
  Dg.resize(r_off[count]);
  Eg.resize(r_off[count]);
  HVg.resize(r_off[count]);
  Wg.resize(r_off[count]);
  // one upload: the job list, then the map from eigenvalue to block
  const size_t oe = (sizeof(ValJob) * count + 255) / 256 * 256;
  std::vector<char> up(oe + sizeof(int32_t) * r_off[count]);
  std::memcpy(up.data(), jobs.data(), sizeof(ValJob) * count);
  int32_t *ej = reinterpret_cast<int32_t *>(up.data() + oe);
  for (int32_t j = 0; j < count; j++) std::fill(ej + r_off[j], ej + r_off[j + 1], j);
  DevBuf<char> dup(up.size());
  ctx().copy_in(dup.p, up.data(), up.size());
  const ValJob *dj = reinterpret_cast<const ValJob *>(dup.p);
  const int32_t *dej = reinterpret_cast<const int32_t *>(dup.p + oe);
  const size_t lds = sizeof(double) * ((size_t)(7 + kGridR) * nmax + kValThreads / 64);   // d e hv vb(2) pw pr pc(32) red
  HIPCHK(hipFuncSetAttribute((const void *)k_eig_values, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  DevBuf<double> bnd(4 * (size_t)count);
  k_eig_values<<<count, kValThreads, lds, s>>>(dj, A, H.p, Dg.p, Eg.p, HVg.p, bnd.p);
  HIPCHK(hipGetLastError());
  const int64_t total = r_off[count];
  k_eig_bisect<<<(unsigned)((total + kBisWaves - 1) / kBisWaves), kEigThreads, 0, s>>>(dej, total, dj, Dg.p, Eg.p,
                                                                                      bnd.p, Wg.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));   // the job list goes back to the allocation cache
}

bool BlockEig::vectors(const std::vector<std::vector<int32_t>> &kept, std::vector<double> &X,
                       std::vector<int64_t> &x_off, hipStream_t s) {
  x_off.assign(nblk + 1, 0);
  std::vector<VecJob> lds_jobs, glob_jobs;
  std::vector<int32_t> idx;
  int64_t w = 0, wg = 0;
  for (int32_t j = 0; j < nblk; j++) {
    const int32_t nj = n[j], kj = (int32_t)kept[j].size();
    x_off[j + 1] = x_off[j] + (int64_t)kj * nj;
    if (kj == 0) continue;
    VecJob V{nj, kj, p_off[j], r_off[j], (int64_t)idx.size(), x_off[j], 0};
    idx.insert(idx.end(), kept[j].begin(), kept[j].end());
    if ((int64_t)nj * kj <= kVecLdsMax) {
      V.w_off = w;
      w += (int64_t)nj * kj;
      lds_jobs.push_back(V);
    } else {
      V.w_off = wg;
      wg += (int64_t)nj * kj;
      glob_jobs.push_back(V);
    }
  }
  X.resize(x_off[nblk]);
  if (idx.empty()) return true;
  for (auto &V : glob_jobs) V.w_off += w;   // after the LDS jobs' factor workspace
  // the reflectors' jobs, kChunk vectors per workgroup
  std::vector<VecJob> all(lds_jobs);
  all.insert(all.end(), glob_jobs.begin(), glob_jobs.end());
  std::vector<BackJob> back;
  int32_t nmax = 1;
  for (size_t j = 0; j < all.size(); j++) {
    nmax = std::max(nmax, all[j].n);
    for (int32_t t0 = 0; t0 < all[j].k; t0 += kChunk) back.push_back(BackJob{(int32_t)j, t0});
  }
  // one upload for every job list and the kept indices
  std::vector<char> up;
  auto put = [&](const void *src, size_t bytes) {
    const size_t o = (up.size() + 255) / 256 * 256;
    up.resize(o + std::max<size_t>(bytes, 1));
    if (bytes) std::memcpy(up.data() + o, src, bytes);
    return o;
  };
  const size_t o_idx = put(idx.data(), sizeof(int32_t) * idx.size());
  const size_t o_l = put(lds_jobs.data(), sizeof(VecJob) * lds_jobs.size());
  const size_t o_g = put(glob_jobs.data(), sizeof(VecJob) * glob_jobs.size());
  const size_t o_a = put(all.data(), sizeof(VecJob) * all.size());
  const size_t o_b = put(back.data(), sizeof(BackJob) * back.size());
  DevBuf<char> dup(up.size());
  Context &C = ctx();
  C.copy_in(dup.p, up.data(), up.size());
  const int32_t *di = reinterpret_cast<const int32_t *>(dup.p + o_idx);
  const VecJob *dl = reinterpret_cast<const VecJob *>(dup.p + o_l), *dg = reinterpret_cast<const VecJob *>(dup.p + o_g);
  const VecJob *da = reinterpret_cast<const VecJob *>(dup.p + o_a);
  const BackJob *db = reinterpret_cast<const BackJob *>(dup.p + o_b);
  // X, then one double whose first word counts the failed blocks: one read-back for both
  DevBuf<double> U(3 * (w + wg)), Yg(std::max<int64_t>(w + wg, 1)), Xd(x_off[nblk] + 1);
  int *fail = reinterpret_cast<int *>(Xd.p + x_off[nblk]);
  HIPCHK(hipMemsetAsync(fail, 0, sizeof(double), s));
  // LDS: the fixed part, then the vectors (LDS form)
  auto area = [](const std::vector<VecJob> &js, bool vecs) {
    int64_t m = 0;
    for (const auto &V : js) m = std::max(m, vecs ? (int64_t)V.n * V.k : 0);
    return sizeof(double) * (size_t)(kVecFixed + m);
  };
  if (!lds_jobs.empty()) {
    const size_t lds = area(lds_jobs, true);
    HIPCHK(hipFuncSetAttribute((const void *)k_eig_vectors<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    k_eig_vectors<true><<<(unsigned)lds_jobs.size(), kEigThreads, lds, s>>>(dl, H.p, Dg.p, Eg.p, HVg.p, Wg.p, di,
                                                                            U.p, Yg.p, Xd.p, fail);
    HIPCHK(hipGetLastError());
  }
  if (!glob_jobs.empty()) {
    const size_t lds = area(glob_jobs, false);
    HIPCHK(hipFuncSetAttribute((const void *)k_eig_vectors<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    k_eig_vectors<false><<<(unsigned)glob_jobs.size(), kEigThreads, lds, s>>>(dg, H.p, Dg.p, Eg.p, HVg.p, Wg.p, di,
                                                                             U.p, Yg.p, Xd.p, fail);
    HIPCHK(hipGetLastError());
  }
  const size_t lds_b = sizeof(double) * ((size_t)nmax * (nmax + 1) / 2 + nmax);
  HIPCHK(hipFuncSetAttribute((const void *)k_eig_back, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
  k_eig_back<<<(unsigned)back.size(), 64, lds_b, s>>>(db, da, H.p, HVg.p, Xd.p);
  HIPCHK(hipGetLastError());
  X.resize(x_off[nblk] + 1);
  C.copy_out(X.data(), Xd.p, sizeof(double) * X.size());   // through pinned memory, after the kernels
  int h = 0;
  std::memcpy(&h, X.data() + x_off[nblk], sizeof(int));
  X.resize(x_off[nblk]);
  return h == 0;
}

}  // namespace dmrgx
