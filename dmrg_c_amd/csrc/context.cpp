// Process-wide device context: one non-blocking stream at default priority, pinned
// staging for scalars, and a cache of device blocks so repeated solves do not pay
// hipMalloc/hipFree on the critical path.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "internal.h"

namespace dmrgx {

namespace {
Context g_ctx;
std::mutex g_init_mu;
thread_local std::string g_last_error;
// Bytes of freed blocks kept for reuse.  A block returned beyond this is freed with
// hipFree, which waits for the device to go idle; a batch of solves that all finish at
// different times (the 256-problem ED batch releases ~60 GB of Lanczos bases) would then
// stall the host at every release and lose the overlap of host checks with the device.
// An allocation that fails frees the whole cache and retries, so a large cache never
// keeps a caller from memory it needs.
constexpr size_t kCacheLimit = size_t(96) << 30;
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

namespace {
constexpr size_t kXferChunk = (size_t)1 << 20;   // bytes per pinned chunk buffer
}

static void xfer_init(Context &C) {
  if (C.xfer_pin[0]) return;
  for (int i = 0; i < 2; i++) {
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&C.xfer_pin[i]), kXferChunk, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&C.xfer_ev[i], hipEventDisableTiming));
  }
}

void Context::copy_out(void *dst, const void *dev_src, size_t bytes) {
  if (bytes == 0) return;
  xfer_init(*this);
  char *out = static_cast<char *>(dst);
  const char *src = static_cast<const char *>(dev_src);
  const size_t chunks = (bytes + kXferChunk - 1) / kXferChunk;
  auto len_of = [&](size_t k) { return std::min(kXferChunk, bytes - k * kXferChunk); };
  for (size_t k = 0; k <= chunks; k++) {
    if (k < chunks) {
      // chunk k's buffer was last used by chunk k - 2, whose host copy is done
      HIPCHK(hipMemcpyAsync(xfer_pin[k & 1], src + k * kXferChunk, len_of(k), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipEventRecord(xfer_ev[k & 1], stream));
    }
    if (k > 0) {
      const size_t j = k - 1;
      HIPCHK(hipEventSynchronize(xfer_ev[j & 1]));
      std::memcpy(out + j * kXferChunk, xfer_pin[j & 1], len_of(j));
    }
  }
}

void Context::copy_in(void *dev_dst, const void *src, size_t bytes) {
  if (bytes == 0) return;
  xfer_init(*this);
  char *dst = static_cast<char *>(dev_dst);
  const char *in = static_cast<const char *>(src);
  const size_t chunks = (bytes + kXferChunk - 1) / kXferChunk;
  for (size_t k = 0; k < chunks; k++) {
    const size_t len = std::min(kXferChunk, bytes - k * kXferChunk);
    if (k >= 2) HIPCHK(hipEventSynchronize(xfer_ev[k & 1]));   // chunk k - 2's DMA has read the buffer
    std::memcpy(xfer_pin[k & 1], in + k * kXferChunk, len);
    HIPCHK(hipMemcpyAsync(dst + k * kXferChunk, xfer_pin[k & 1], len, hipMemcpyHostToDevice, stream));
    HIPCHK(hipEventRecord(xfer_ev[k & 1], stream));
  }
  // the buffers are free again (and the data on the device) once the last DMAs are done
  HIPCHK(hipEventSynchronize(xfer_ev[(chunks - 1) & 1]));
  if (chunks > 1) HIPCHK(hipEventSynchronize(xfer_ev[chunks & 1]));
}

void Context::copy_in_many(const std::vector<CopySeg> &segs) {
  size_t total = 0;
  for (const CopySeg &g : segs) total += g.bytes;
  if (total == 0) return;
  xfer_init(*this);
  // the segments' bytes in order, kXferChunk at a time; a chunk's pieces go to their own
  // destinations
  size_t seg = 0, in_seg = 0;   // next byte to stage: segment seg, offset in_seg
  const size_t chunks = (total + kXferChunk - 1) / kXferChunk;
  for (size_t k = 0; k < chunks; k++) {
    if (k >= 2) HIPCHK(hipEventSynchronize(xfer_ev[k & 1]));   // chunk k - 2's DMAs have read the buffer
    char *pin = xfer_pin[k & 1];
    size_t fill = 0;
    while (fill < kXferChunk && seg < segs.size()) {
      const CopySeg &g = segs[seg];
      const size_t len = std::min(kXferChunk - fill, g.bytes - in_seg);
      if (len > 0) {
        std::memcpy(pin + fill, static_cast<const char *>(g.src) + in_seg, len);
        HIPCHK(hipMemcpyAsync(static_cast<char *>(g.dev_dst) + in_seg, pin + fill, len, hipMemcpyHostToDevice, stream));
        fill += len;
        in_seg += len;
      }
      if (in_seg == g.bytes) {
        seg++;
        in_seg = 0;
      }
    }
    HIPCHK(hipEventRecord(xfer_ev[k & 1], stream));
  }
  HIPCHK(hipEventSynchronize(xfer_ev[(chunks - 1) & 1]));
  if (chunks > 1) HIPCHK(hipEventSynchronize(xfer_ev[chunks & 1]));
}

hipEvent_t *Context::step_events() {
  if (!step_ev[0])
    for (int i = 0; i < 2; i++) HIPCHK(hipEventCreateWithFlags(&step_ev[i], hipEventDisableTiming));
  return step_ev;
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
