// Host helpers shared by the HIP translation units: timers, scalar
// read-back and 64-bit exclusive scans.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>

#include "common.hpp"

namespace c2 {

struct Timer {
  hipStream_t s;
  hipEvent_t a, b;
  explicit Timer(hipStream_t st) : s(st) {
    C2_HIP(hipEventCreate(&a));
    C2_HIP(hipEventCreate(&b));
    C2_HIP(hipEventRecord(a, s));
  }
  double stop() {
    C2_HIP(hipEventRecord(b, s));
    C2_HIP(hipEventSynchronize(b));
    float ms = 0;
    C2_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

// One value computed on the device, read back through the pinned staging
// buffer (waits for the stream).
template <class T>
inline T read_scalar(c2gpu_ctx *ctx, const T *d) {
  T h;
  copy_to_host(ctx, &h, d, sizeof(T));
  return h;
}

struct ToU64 {
  __host__ __device__ uint64_t operator()(uint32_t v) const { return (uint64_t)v; }
};

// Exclusive prefix sum of 32-bit counts into 64-bit offsets (the scan
// accumulates in 64 bits: totals exceed 2^32 on large grids).
inline void exclusive_scan(c2gpu_ctx *ctx, const char *tag, const uint32_t *in, uint64_t *out,
                           int64_t n) {
  hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> it(in, ToU64{});
  size_t tmp = 0;
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, (int)n, ctx->stream));
  void *t = buffer(ctx, std::string("scan_tmp_") + tag, tmp);
  C2_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, it, out, (int)n, ctx->stream));
}

// Sum of 32-bit counts in 64 bits.
inline uint64_t sum_u32(c2gpu_ctx *ctx, const char *tag, const uint32_t *in, int64_t n) {
  hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> it(in, ToU64{});
  uint64_t *d_out = static_cast<uint64_t *>(buffer(ctx, std::string("sum_out_") + tag, 8));
  size_t tmp = 0;
  C2_HIP(hipcub::DeviceReduce::Sum(nullptr, tmp, it, d_out, (int)n, ctx->stream));
  void *t = buffer(ctx, std::string("sum_tmp_") + tag, tmp);
  C2_HIP(hipcub::DeviceReduce::Sum(t, tmp, it, d_out, (int)n, ctx->stream));
  return read_scalar(ctx, d_out);
}

// Indices i with flag[i] != 0, in increasing order; returns the count.
inline uint32_t compact_flags(c2gpu_ctx *ctx, const char *tag, const uint8_t *flag, uint32_t n,
                              uint32_t *out) {
  uint32_t *d_n = static_cast<uint32_t *>(buffer(ctx, std::string("compact_n_") + tag, 4));
  hipcub::CountingInputIterator<uint32_t> it(0);
  size_t tmp = 0;
  C2_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp, it, flag, out, d_n, (int)n, ctx->stream));
  void *t = buffer(ctx, std::string("compact_tmp_") + tag, tmp);
  C2_HIP(hipcub::DeviceSelect::Flagged(t, tmp, it, flag, out, d_n, (int)n, ctx->stream));
  return read_scalar(ctx, d_n);
}

}  // namespace c2
