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

// A link descriptor's arrays copied to the device for link_csr_device (freed with it).
struct LinkUpload {
  DevBuf l, st, di, bi;
  LinkUpload(hipStream_t s, const fg_force_desc& d) {
    const size_t m = d.count;
    if (m) {
      l.alloc(sizeof(uint32_t) * 2 * m);
      FG_HIP(hipMemcpyAsync(l.p, d.links, sizeof(uint32_t) * 2 * m, hipMemcpyHostToDevice, s));
    }
    copy(s, st, d.strengths, m);
    copy(s, di, d.distances, m);
    copy(s, bi, d.bias, m);
  }
  static void copy(hipStream_t s, DevBuf& b, const double* src, size_t m) {
    if (!src || !m) return;
    b.alloc(sizeof(double) * m);
    FG_HIP(hipMemcpyAsync(b.p, src, sizeof(double) * m, hipMemcpyHostToDevice, s));
  }
  const uint32_t* links() const { return l.as<uint32_t>(); }
  const double* strengths() const { return st.as<double>(); }
  const double* distances() const { return di.as<double>(); }
  const double* bias() const { return bi.as<double>(); }
};

}  // namespace fg
