// Helpers shared by the extern "C" entry points (abi.cpp, obc.cpp, expect.cpp, ed_abi.cpp).
#pragma once

#include <mutex>

#include "internal.h"

namespace dmrgx {

// One API call at a time: the library keeps one stream and one allocation cache per process.
std::recursive_mutex &api_mutex();

// Append to ./SML_out/<file> as the reference's solvers do (DMRGX_SML_OUT=0 disables).
void sml_append(const char *file, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// The reference's fatal path: "Error in <fn>", the message, exit(1).
[[noreturn]] void fatal(const char *fn, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
double now_s();
void check_square(const char *fn, const CRS1 *M);

// Run `body`; on a library failure print it and exit(1) (the drop-in has no status channel).
template <class F>
void run_or_die(const char *fn, F &&body) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  try {
    body();
  } catch (const Failure &f) {
    fatal(fn, "%s\n", f.what());
  }
}

// Run `body` for a dmrgx_* entry point; failures become status codes.
template <class F>
int run_status(F &&body) {
  std::lock_guard<std::recursive_mutex> lk(api_mutex());
  try {
    body();
    return DMRGX_OK;
  } catch (const Failure &f) {
    set_last_error(f.what());
    return f.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return DMRGX_ERR_HIP;
  }
}

// Device copy of a host vector, synchronised on download.
struct DevVec {
  DevBuf<double> d;
  int64_t n;
  explicit DevVec(int64_t n_) : d(n_), n(n_) {}
  void up(const double *h) { ctx().copy_in(d.p, h, sizeof(double) * n); }
  void down(double *h) { ctx().copy_out(h, d.p, sizeof(double) * n); }
};

}  // namespace dmrgx
