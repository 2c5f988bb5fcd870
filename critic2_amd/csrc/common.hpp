// Shared host-side plumbing of libc2gpu: context, error reporting and a
// grow-only device buffer cache.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/c2gpu.h"

struct c2gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  // named grow-only device allocations reused across calls
  struct Buf {
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Buf> bufs;
  // per-kernel device timing (c2gpu_ctx_timing): events recorded around
  // each launch while enabled, resolved at the end of every API call
  bool timing = false;
  struct Pending {
    const char *name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> evpool;
  struct KTime {
    double ms = 0;
    int64_t launches = 0;
    double bytes = 0;  // algorithmic bytes reported by the launching code
  };
  std::map<std::string, KTime> ktime;
  // batches: worker contexts (one stream each) and the tuned workgroups per
  // CU of the work-queue NCI kernel (0 = not tuned yet)
  std::vector<c2gpu_ctx *> subs;
  int queue_wg_per_cu = 0;
  // pinned host staging for the library's own device -> host read-backs
  // (copy_to_host): the copy engine writes it, so the CPU never reads a
  // device address itself
  void *pin = nullptr;
  size_t pin_bytes = 0;
  // the surface weights of the last c2gpu_yt call (context buffers
  // yt_ias, yt_woff, yt_wc, yt_wb, yt_ww), for c2gpu_yt_multipoles, as
  // critic2 keeps them in its YT checkpoint for later yt_weights calls
  struct YtState {
    bool valid = false;
    int n[3] = {0, 0, 0};
    int nattr = 0;
    uint32_t nias = 0;
    uint64_t nent = 0;  // weight entries over all surface points
    const uint32_t *ias = nullptr;
    uint32_t *wb = nullptr;  // basin ids, remapped by c2gpu_yt_remap
    const uint64_t *woff = nullptr;
    const uint16_t *wc = nullptr;
    const double *ww = nullptr;
  } yt;
};

namespace c2 {

struct Error {
  int code;
  std::string msg;
};

inline void fail(int code, const std::string &msg) { throw Error{code, msg}; }

#define C2_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      ::c2::fail(C2GPU_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Returns a device buffer of at least `bytes` bytes registered under `name`
// in the context (grow-only; contents are not preserved on growth).
inline void *buffer(c2gpu_ctx *ctx, const std::string &name, size_t bytes) {
  auto &b = ctx->bufs[name];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    C2_HIP(hipStreamSynchronize(ctx->stream));
    C2_HIP(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    fail(C2GPU_ERR_ALLOC, "hipMalloc(" + std::to_string(want) + " B) for " + name + ": " +
                              hipGetErrorString(e));
  }
  b.bytes = want;
  return b.ptr;
}

template <class T>
inline T *buf(c2gpu_ctx *ctx, const std::string &name, size_t count) {
  return static_cast<T *>(buffer(ctx, name, count * sizeof(T)));
}

// Frees a named buffer (after the stream's work), so the next call of
// buffer() for it allocates afresh.
inline void release(c2gpu_ctx *ctx, const std::string &name) {
  auto it = ctx->bufs.find(name);
  if (it == ctx->bufs.end()) return;
  if (it->second.ptr) {
    C2_HIP(hipStreamSynchronize(ctx->stream));
    C2_HIP(hipFree(it->second.ptr));
  }
  ctx->bufs.erase(it);
}

// A buffer of at least min_bytes and at most want_bytes: as much of
// want_bytes as the device's memory allows (the buffer's own bytes count as
// available), keeping `spare` bytes free for later allocations.  A held
// buffer at least that large is kept as it is; otherwise the old contents
// are not kept.
inline void *buffer_within(c2gpu_ctx *ctx, const std::string &name, size_t min_bytes, size_t want_bytes,
                           size_t spare) {
  auto it = ctx->bufs.find(name);
  const size_t own = (it != ctx->bufs.end() && it->second.ptr) ? it->second.bytes : 0;
  if (own >= want_bytes) return it->second.ptr;
  size_t fr = 0, tot = 0;
  C2_HIP(hipMemGetInfo(&fr, &tot));
  size_t bytes = fr + own > spare ? fr + own - spare : 0;
  if (bytes > want_bytes) bytes = want_bytes;
  bytes &= ~(size_t)255;
  if (bytes < min_bytes) bytes = min_bytes;
  if (own >= bytes) return it->second.ptr;  // already as large as it can get
  release(ctx, name);
  return buffer(ctx, name, bytes);
}

inline hipEvent_t event_get(c2gpu_ctx *ctx) {
  if (!ctx->evpool.empty()) {
    hipEvent_t e = ctx->evpool.back();
    ctx->evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  C2_HIP(hipEventCreate(&e));
  return e;
}

// Accumulate the recorded launch times (waits for the recorded events).
inline void flush_timing(c2gpu_ctx *ctx) {
  for (auto &p : ctx->pending) {
    C2_HIP(hipEventSynchronize(p.b));
    float ms = 0;
    C2_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto &k = ctx->ktime[p.name];
    k.ms += ms;
    k.launches += 1;
    ctx->evpool.push_back(p.a);
    ctx->evpool.push_back(p.b);
  }
  ctx->pending.clear();
}

// Credit algorithmic bytes to a kernel's entry (only while timing is on).
inline void add_kernel_bytes(c2gpu_ctx *ctx, const char *name, double bytes) {
  if (ctx->timing) ctx->ktime[name].bytes += bytes;
}

// Kernel launch on the context's stream, timed when timing is enabled.
#define C2_LAUNCH(ctx, kern, grid, block, lds, ...)                            \
  do {                                                                         \
    hipEvent_t a_ = nullptr, b_ = nullptr;                                     \
    if ((ctx)->timing) {                                                       \
      a_ = ::c2::event_get(ctx);                                               \
      b_ = ::c2::event_get(ctx);                                               \
      C2_HIP(hipEventRecord(a_, (ctx)->stream));                               \
    }                                                                          \
    hipLaunchKernelGGL(kern, grid, block, lds, (ctx)->stream, __VA_ARGS__);    \
    C2_HIP(hipGetLastError());                                                 \
    if ((ctx)->timing) {                                                       \
      C2_HIP(hipEventRecord(b_, (ctx)->stream));                               \
      (ctx)->pending.push_back({#kern, a_, b_});                               \
    }                                                                          \
  } while (0)

inline void release_all(c2gpu_ctx *ctx) {
  for (auto &kv : ctx->bufs)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  ctx->bufs.clear();
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  ctx->pin = nullptr;
  ctx->pin_bytes = 0;
}

// The context's pinned staging buffer, at least `bytes` long (grow-only).
inline void *pinned(c2gpu_ctx *ctx, size_t bytes) {
  if (ctx->pin && ctx->pin_bytes >= bytes) return ctx->pin;
  if (ctx->pin) {
    C2_HIP(hipStreamSynchronize(ctx->stream));
    C2_HIP(hipHostFree(ctx->pin));
    ctx->pin = nullptr;
    ctx->pin_bytes = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  hipError_t e = hipHostMalloc(&ctx->pin, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    ctx->pin = nullptr;
    fail(C2GPU_ERR_ALLOC, "hipHostMalloc(" + std::to_string(want) + " B): " + hipGetErrorString(e));
  }
  ctx->pin_bytes = want;
  return ctx->pin;
}

// Device -> host copy of the library's own results (counts, offsets, small
// tables) into any host memory, through the pinned staging buffer in chunks
// of at most 16 MiB; returns after the data has arrived.
inline void copy_to_host(c2gpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
  const size_t chunk = (size_t)16 << 20;
  char *d = static_cast<char *>(dst);
  const char *s = static_cast<const char *>(src);
  void *pin = pinned(ctx, bytes < chunk ? bytes : chunk);
  for (size_t done = 0; done < bytes;) {
    const size_t n = bytes - done < chunk ? bytes - done : chunk;
    C2_HIP(hipMemcpyAsync(pin, s + done, n, hipMemcpyDeviceToHost, ctx->stream));
    C2_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(d + done, pin, n);
    done += n;
  }
  if (bytes == 0) C2_HIP(hipStreamSynchronize(ctx->stream));
}

// Runs f(), converting exceptions into a status code and ctx->err.
template <class F>
inline int guarded(c2gpu_ctx *ctx, F &&f) {
  if (!ctx) return C2GPU_ERR_ARG;
  try {
    ctx->err.clear();
    C2_HIP(hipSetDevice(ctx->device));
    f();
    flush_timing(ctx);
    return C2GPU_OK;
  } catch (const Error &e) {
    ctx->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return C2GPU_ERR_ARG;
  }
}

// With C2GPU_PTR_DEVICE the caller promises device memory: verify it, so a
// host pointer is rejected with C2GPU_ERR_ARG instead of faulting the GPU.
inline void require_device_ptr(const void *p, const char *what) {
  if (!p) fail(C2GPU_ERR_ARG, std::string(what) + " is NULL");
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(C2GPU_ERR_ARG, std::string(what) + " is not a device pointer (C2GPU_PTR_DEVICE set)");
  }
  if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)
    fail(C2GPU_ERR_ARG, std::string(what) + " is not device memory (C2GPU_PTR_DEVICE set)");
}

// Workgroups of `block` threads covering n work-items.  One dispatch holds
// fewer than 2^32 work-items (the grid size is a 32-bit field of the
// dispatch packet); larger launches must loop over a bounded grid.
inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g * (int64_t)block >= (int64_t(1) << 32))
    fail(C2GPU_ERR_ARG, "launch of " + std::to_string(n) + " work-items exceeds one dispatch (2^32)");
  return (unsigned)g;
}

// Number of resident threads used by persistent (grid-stride) kernels.
inline int persistent_threads(int device, int block) {
  int ncu = 256;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  return ncu * 8 * block;
}

}  // namespace c2
