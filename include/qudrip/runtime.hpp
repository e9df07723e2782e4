// qudrip/runtime.hpp -- process-wide GPU context of the drop-in API.
//
// The reference has no device; the drop-in headers keep its call surface and
// route the heavy work through the C ABI (include/qudrip_hip.h).  One
// context (device QUDRIP_DEVICE, default 0, one HIP stream) serves every
// object of the process, like the reference's single-threaded, non-reentrant
// use of its shared index cursors.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>

#include "../qudrip_hip.h"

namespace qudrip {
namespace detail {

struct Runtime {
  qd_ctx* ctx = nullptr;
  Runtime() {
    const char* dev = std::getenv("QUDRIP_DEVICE");
    int rc = qd_ctx_create(dev ? std::atoi(dev) : 0, &ctx);
    if (rc != QD_OK) {
      std::fprintf(stderr, "qudrip: no usable GPU (qd_ctx_create status %d)\n", rc);
      std::abort();
    }
  }
  ~Runtime() { qd_ctx_destroy(ctx); }
};

inline qd_ctx* ctx() {
  static Runtime rt;
  return rt.ctx;
}

inline void check(int rc, const char* what) {
  if (rc != QD_OK) throw std::runtime_error(std::string(what) + ": " + qd_last_error(ctx()));
}

// owning device buffer of complex128 entries
struct DeviceBuffer {
  double* p = nullptr;
  std::size_t count = 0;  // complex entries
  DeviceBuffer() = default;
  explicit DeviceBuffer(std::size_t n) : count(n) {
    void* q = nullptr;
    check(qd_malloc(ctx(), sizeof(double) * 2 * (n ? n : 1), &q), "qd_malloc");
    p = static_cast<double*>(q);
  }
  ~DeviceBuffer() {
    if (p) qd_free(ctx(), p);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

}  // namespace detail

// Seed of the seeded shuffle_state (the reference seeds from random_device,
// utils.hpp:131-132, so its runs are not reproducible; these are).
inline unsigned long long& shuffle_seed() {
  static unsigned long long s = 1;
  return s;
}

}  // namespace qudrip
