// The 5^q table of the decimal conversions (number.hpp, format.hpp): built
// once per process on the host, uploaded once per context.
#pragma once
#include <cstdint>
#include <vector>

struct c2gpu_ctx;

namespace c2 {

struct Pow5Host {
  std::vector<uint64_t> t;  // (hi, lo) of T(q) per q from POW5_QMIN
  std::vector<int16_t> s;   // s(q)
};

const Pow5Host &pow5_host();
void pow5_device(c2gpu_ctx *ctx, const uint64_t *&t, const int16_t *&s);

}  // namespace c2
