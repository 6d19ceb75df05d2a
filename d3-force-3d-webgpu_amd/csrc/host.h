// host.h -- host-side helpers shared by the single-simulation and batch engines.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/d3fgpu.h"
#include "common.h"
#include "forces.h"

namespace fg {

constexpr float BIG = 3e38f;  // stands for infinity on the device (passes/util.js:5-9)

// A JS number as the device's f32: NaN -> 0, +-inf -> +-3e38 (util.js:24-28).
inline float finite_f32(double v) {
  if (!(v == v)) return 0.0f;
  if (v > BIG) return BIG;
  if (v < -BIG) return -BIG;
  return (float)v;
}

// One device allocation, freed with its owner.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void alloc(size_t b) {
    free();
    if (b == 0) b = 16;
    FG_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
  ~DevBuf() { free(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// The incident links of every node (CSR) from a link descriptor, with d3's defaults for
// missing arrays (passes/link.js:65-101).  Throws ArgError on a node index out of range.
void link_csr(uint32_t n, const fg_force_desc& d, std::vector<uint32_t>& row, std::vector<LinkEntry>& ent);

}  // namespace fg
