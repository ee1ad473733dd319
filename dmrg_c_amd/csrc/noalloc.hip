// The allocation-free entry point: CONVERGE_CHECK over CSR arrays and a vector the caller
// already holds in device memory, for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops that forbid hipMalloc).
//
// Its only device memory beyond the caller's inputs is `scratch`, a file-scope array of
// kScratchDoubles doubles that every GPU the library runs on holds until the process ends:
// the block partials of the residual product (at most kGridCap, one per workgroup, each
// fully written by its workgroup before the finish kernel reads it, so nothing has to be
// cleared between calls) and the finished sum.  All calls on one GPU share that GPU's copy:
// calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "entry.h"

namespace dmrgx {
namespace {

constexpr int kScratchPartials = kGridCap;              // workgroups of the default schedule
constexpr int kScratchDoubles = kScratchPartials + 32;  // + the result, padded to 256 bytes

static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` per device, found once per device with that device current.
// The host-side name of the array is a stand-in, not a device address.
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

double *scratch_on_current_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Failure(DMRGX_ERR_ARG, "device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Failure(DMRGX_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

extern "C" {

int dmrgx_converge_check_noalloc(int32_t n, int64_t nnz, const int64_t *row_dev, const int32_t *col_dev,
                                 const double *val_dev, const double *vec_dev, double eig_val, double *out) {
  return run_status([&] {
    if (n < 0 || nnz < 0 || !out) throw Failure(DMRGX_ERR_ARG, "bad size or NULL output");
    if (n > 0 && (!row_dev || !vec_dev || (nnz > 0 && (!col_dev || !val_dev))))
      throw Failure(DMRGX_ERR_ARG, "NULL device array");
    Context &C = ctx();   // the library's stream (created at initialisation; no device memory)
    double *sp = scratch_on_current_device();
    if (n == 0) {
      *out = 0.0;
      return;
    }
    const int blocks = residual_l1_default_blocks(n, nnz);
    if (blocks > kScratchPartials)
      throw Failure(DMRGX_ERR_ARG, "dmrgx_converge_check_noalloc: more workgroups than the scratch holds");
    launch_residual_l1_into(n, nnz, row_dev, col_dev, val_dev, vec_dev, eig_val, sp + kScratchPartials, sp,
                            C.stream);
    HIPCHK(hipMemcpyAsync(out, sp + kScratchPartials, sizeof(double), hipMemcpyDeviceToHost, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
  });
}

}  // extern "C"
