// Runtime of the library: device context, stream, caching device allocator, pinned staging,
// error reporting.  One process drives one GPU (the current HIP device).
#pragma once
#include <stdio.h>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <string>

namespace zk {

// Status codes returned by the zkg_* extension API (0 = success).
enum Status : int {
  ZK_OK = 0,
  ZK_EINVAL = 1,     // bad argument (size, curve id, window, ...)
  ZK_EHIP = 2,       // HIP runtime error
  ZK_ENOMEM = 3,     // device allocation failed
  ZK_ENOTROOT = 4,   // NTT generator does not have order 2^m
};

void set_error(const std::string &msg);
const char *last_error();

// Abort with a message, mirroring the reference's assert() on failure
// (lib/cbits/curves/g1/proj/bn128_G1_proj.c:517, curves/poly/mont/bn128_poly_mont.c:458).
[[noreturn]] void fatal(const char *where, const char *msg);

#define ZK_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::zk::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return ::zk::ZK_EHIP;                                                       \
    }                                                                             \
  } while (0)

#define ZK_TRY(expr)                  \
  do {                                \
    int s_ = (expr);                  \
    if (s_ != ::zk::ZK_OK) return s_; \
  } while (0)

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  std::recursive_mutex lock;  // serialises users of the shared stream and pools
};

// Initialises (once) and returns the context for the current device.
int context(Context **out);

// Caching device allocator: blocks are kept per exact size and reused.
int dev_alloc(void **p, size_t bytes);
void dev_free(void *p);          // returns the block to the cache
void dev_trim();                 // releases cached blocks back to HIP

// RAII device buffer
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) dev_free(p);
  }
  int alloc(size_t n) {
    if (p && bytes >= n) return ZK_OK;
    if (p) dev_free(p);
    p = nullptr;
    bytes = 0;
    int s = dev_alloc(&p, n ? n : 256);
    if (s == ZK_OK) bytes = n;
    return s;
  }
  template <class T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

// Synchronises a stream when it leaves scope unless released.  dev_free returns a block to the
// cache at once, with no record of the stream still using it, so no DevBuf may be destroyed
// while work queued on a stream still reads or writes it: a function that queues work and can
// return early declares this guard after its DevBufs (destructors run in reverse order, so the
// stream is drained before any of them is freed) and releases it after its own final sync.
struct StreamSyncGuard {
  hipStream_t s;
  explicit StreamSyncGuard(hipStream_t st) : s(st) {}
  StreamSyncGuard(const StreamSyncGuard &) = delete;
  StreamSyncGuard &operator=(const StreamSyncGuard &) = delete;
  ~StreamSyncGuard() {
    if (s) (void)hipStreamSynchronize(s);
  }
  void release() { s = nullptr; }
};

// Host -> device / device -> host copies of caller (pageable) memory through a pinned
// staging ring, on the library stream.  Synchronous with respect to the host on return.
int copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
int copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);
// the same staging between an open file and device memory (raw bytes, sequential)
int file_to_device(FILE *f, void *dst, size_t bytes, hipStream_t s);
int device_to_file(FILE *f, const void *src, size_t bytes, hipStream_t s);

// Event-timed region helper for kernels launched on the library stream.
struct KernelTimer {
  hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace zk
