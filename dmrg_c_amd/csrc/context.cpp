// Process-wide device context: one non-blocking stream at default priority, pinned
// staging for scalars, and a cache of device blocks so repeated solves do not pay
// hipMalloc/hipFree on the critical path.
#include <cstdarg>
#include <cstdlib>
#include <mutex>

#include "internal.h"

namespace dmrgx {

namespace {
Context g_ctx;
std::mutex g_init_mu;
thread_local std::string g_last_error;
constexpr size_t kCacheLimit = size_t(16) << 30;  // bytes kept for reuse
}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }
const std::string &last_error() { return g_last_error; }

void hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  throw Failure(e == hipErrorOutOfMemory ? DMRGX_ERR_NOMEM : DMRGX_ERR_HIP, buf);
}

void ensure_init() {
  if (g_ctx.stream) return;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ctx.stream) return;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw Failure(DMRGX_ERR_NOGPU, "no HIP device visible: libdmrg_c_amd runs on the GPU only");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipSetDevice(dev));
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  g_ctx.device = dev;
}

Context &ctx() {
  ensure_init();
  return g_ctx;
}

void *Context::alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  size_t best = (size_t)-1, best_i = 0;
  for (size_t i = 0; i < free_list.size(); i++) {
    const size_t sz = free_list[i].first;
    if (sz >= bytes && sz <= bytes + bytes / 4 + 4096 && sz < best) {
      best = sz;
      best_i = i;
    }
  }
  if (best != (size_t)-1) {
    void *p = free_list[best_i].second;
    cached_bytes -= best;
    free_list.erase(free_list.begin() + best_i);
    // remember the block's true size through the size map
    sizes_note(p, best);
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    HIPCHK(hipStreamSynchronize(stream));
    for (auto &f : free_list) (void)hipFree(f.second);
    free_list.clear();
    cached_bytes = 0;
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      char buf[128];
      snprintf(buf, sizeof(buf), "device allocation of %zu bytes failed", bytes);
      throw Failure(DMRGX_ERR_NOMEM, buf);
    }
  }
  sizes_note(p, bytes);
  return p;
}

void Context::release(void *p, size_t) {
  if (!p) return;
  const size_t bytes = sizes_take(p);
  if (cached_bytes + bytes <= kCacheLimit) {
    free_list.emplace_back(bytes, p);
    cached_bytes += bytes;
  } else {
    (void)hipStreamSynchronize(stream);
    (void)hipFree(p);
  }
}

std::pair<hipEvent_t, hipEvent_t> *Context::prof_slot() {
  if (prof_used == prof_ev.size()) {
    std::pair<hipEvent_t, hipEvent_t> p;
    HIPCHK(hipEventCreate(&p.first));
    HIPCHK(hipEventCreate(&p.second));
    prof_ev.push_back(p);
  }
  return &prof_ev[prof_used++];
}

double *Context::host_stage(size_t count) {
  if (count > pinned_cap) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    size_t cap = 4096;
    while (cap < count) cap *= 2;
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&pinned), cap * sizeof(double), hipHostMallocDefault));
    pinned_cap = cap;
  }
  return pinned;
}

}  // namespace dmrgx
