#include "runtime.hpp"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <vector>

namespace zk {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

void fatal(const char *where, const char *msg) {
  fprintf(stderr, "zkgpu: %s: %s\n", where, msg && *msg ? msg : "unknown error");
  fflush(stderr);
  abort();
}

static Context g_ctx;
static std::once_flag g_once;
static int g_init_status = ZK_OK;

int context(Context **out) {
  std::call_once(g_once, [] {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
      set_error("hipGetDevice failed: no usable GPU");
      g_init_status = ZK_EHIP;
      return;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
      set_error("hipGetDeviceProperties failed");
      g_init_status = ZK_EHIP;
      return;
    }
    g_ctx.device = dev;
    g_ctx.num_cus = prop.multiProcessorCount;
    // default priority on purpose: mixed stream priorities are avoided
    if (hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking) != hipSuccess) {
      set_error("hipStreamCreate failed");
      g_init_status = ZK_EHIP;
      return;
    }
  });
  if (g_init_status != ZK_OK) return g_init_status;
  *out = &g_ctx;
  return ZK_OK;
}

// ---------------------------------------------------------------------------------------
// caching allocator

static std::mutex g_alloc_lock;
static std::multimap<size_t, void *> g_free_blocks;  // size -> block
static std::map<void *, size_t> g_live;               // block -> size

static size_t round_size(size_t n) {
  // 2 MiB granularity above 2 MiB, powers of two below: keeps the cache hit rate high
  if (n <= (2u << 20)) {
    size_t r = 256;
    while (r < n) r <<= 1;
    return r;
  }
  return (n + (2u << 20) - 1) & ~((size_t)(2u << 20) - 1);
}

int dev_alloc(void **p, size_t bytes) {
  size_t sz = round_size(bytes);
  {
    std::lock_guard<std::mutex> g(g_alloc_lock);
    auto it = g_free_blocks.find(sz);
    if (it != g_free_blocks.end()) {
      *p = it->second;
      g_free_blocks.erase(it);
      g_live[*p] = sz;
      return ZK_OK;
    }
  }
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, sz);
  if (e != hipSuccess) {
    // release the cache and retry once
    dev_trim();
    (void)hipGetLastError();
    e = hipMalloc(&q, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      char buf[128];
      snprintf(buf, sizeof buf, "hipMalloc(%zu) failed", sz);
      set_error(buf);
      return ZK_ENOMEM;
    }
  }
  std::lock_guard<std::mutex> g(g_alloc_lock);
  g_live[q] = sz;
  *p = q;
  return ZK_OK;
}

void dev_free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(g_alloc_lock);
  auto it = g_live.find(p);
  if (it == g_live.end()) return;
  g_free_blocks.emplace(it->second, p);
  g_live.erase(it);
}

void dev_trim() {
  std::vector<void *> blocks;
  {
    std::lock_guard<std::mutex> g(g_alloc_lock);
    for (auto &kv : g_free_blocks) blocks.push_back(kv.second);
    g_free_blocks.clear();
  }
  for (void *b : blocks) (void)hipFree(b);
}

// ---------------------------------------------------------------------------------------
// pinned staging for pageable caller buffers

static const size_t kStageBytes = 32u << 20;
static void *g_stage[2] = {nullptr, nullptr};
static hipEvent_t g_stage_ev[2] = {nullptr, nullptr};
static std::mutex g_stage_lock;

static int ensure_stage() {
  if (g_stage[0]) return ZK_OK;
  for (int i = 0; i < 2; i++) {
    ZK_HIP(hipHostMalloc(&g_stage[i], kStageBytes, hipHostMallocDefault));
    ZK_HIP(hipEventCreateWithFlags(&g_stage_ev[i], hipEventDisableTiming));
  }
  return ZK_OK;
}

// fill(stage, off, n) writes bytes [off, off + n) of the source into the pinned stage; the two
// stages alternate so that filling one overlaps the copy out of the other
template <class Fill>
static int stage_h2d(void *dst, size_t bytes, hipStream_t s, Fill fill) {
  if (!bytes) return ZK_OK;
  std::lock_guard<std::mutex> g(g_stage_lock);
  ZK_TRY(ensure_stage());
  size_t off = 0;
  int k = 0;
  bool used[2] = {false, false};
  while (off < bytes) {
    size_t n = bytes - off < kStageBytes ? bytes - off : kStageBytes;
    if (used[k]) ZK_HIP(hipEventSynchronize(g_stage_ev[k]));
    ZK_TRY(fill(g_stage[k], off, n));
    ZK_HIP(hipMemcpyAsync((char *)dst + off, g_stage[k], n, hipMemcpyHostToDevice, s));
    ZK_HIP(hipEventRecord(g_stage_ev[k], s));
    used[k] = true;
    off += n;
    k ^= 1;
  }
  ZK_HIP(hipStreamSynchronize(s));
  return ZK_OK;
}

// drain(stage, off, n) consumes bytes [off, off + n) of the device source; the copy of the next
// chunk into the other stage runs while the current one is drained
template <class Drain>
static int stage_d2h(const void *src, size_t bytes, hipStream_t s, Drain drain) {
  if (!bytes) return ZK_OK;
  std::lock_guard<std::mutex> g(g_stage_lock);
  ZK_TRY(ensure_stage());
  auto issue = [&](size_t off, int k) -> int {
    const size_t n = bytes - off < kStageBytes ? bytes - off : kStageBytes;
    ZK_HIP(hipMemcpyAsync(g_stage[k], (const char *)src + off, n, hipMemcpyDeviceToHost, s));
    ZK_HIP(hipEventRecord(g_stage_ev[k], s));
    return ZK_OK;
  };
  ZK_TRY(issue(0, 0));
  int k = 0;
  for (size_t off = 0; off < bytes; off += kStageBytes, k ^= 1) {
    const size_t n = bytes - off < kStageBytes ? bytes - off : kStageBytes;
    if (off + n < bytes) ZK_TRY(issue(off + n, k ^ 1));
    ZK_HIP(hipEventSynchronize(g_stage_ev[k]));
    ZK_TRY(drain(g_stage[k], off, n));
  }
  ZK_HIP(hipStreamSynchronize(s));
  return ZK_OK;
}

int copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
  return stage_h2d(dst, bytes, s, [&](void *stage, size_t off, size_t n) {
    memcpy(stage, (const char *)src + off, n);
    return ZK_OK;
  });
}

int copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
  return stage_d2h(src, bytes, s, [&](const void *stage, size_t off, size_t n) {
    memcpy((char *)dst + off, stage, n);
    return ZK_OK;
  });
}

int file_to_device(FILE *f, void *dst, size_t bytes, hipStream_t s) {
  return stage_h2d(dst, bytes, s, [&](void *stage, size_t, size_t n) {
    if (fread(stage, 1, n, f) != n) {
      set_error("flat array read: short read");
      return ZK_EINVAL;
    }
    return ZK_OK;
  });
}

int device_to_file(FILE *f, const void *src, size_t bytes, hipStream_t s) {
  return stage_d2h(src, bytes, s, [&](const void *stage, size_t, size_t n) {
    if (fwrite(stage, 1, n, f) != n) {
      set_error("flat array write: short write");
      return ZK_EINVAL;
    }
    return ZK_OK;
  });
}

}  // namespace zk
