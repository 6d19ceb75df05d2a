// api.hip -- the C ABI (include/d3fgpu.h): one simulation context on one GPU.
//
// Replaces the reference engine (src/gpu/engine.js:73-428).  The state lives in HBM as
// structure-of-arrays f32 (x, y, z, vx, vy, vz; two velocity buffers so that passes that
// read other nodes' velocities -- link, small-graph collide -- write the other one), pins as
// f32 targets plus a byte mask.  Each force descriptor gets a pass (passes/index.js:8-14)
// that keeps its device arrays until its (id, version) changes (engine.js:314-330, 357).
// A tick runs the passes in order, then the integrator (engine.js:376-383).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/d3fgpu.h"
#include "common.h"
#include "forces.h"
#include "host.h"
#include "sort.h"
#include "tree.h"

using namespace fg;

namespace {

constexpr uint32_t GRID_MIN = 8192;  // collide through the grid from this many nodes (collide.js:21-22)
constexpr double EXACT_THETA2 = 1e-6;  // all-pairs many-body below this theta^2 (manyBody.js:20-21)
constexpr int MAX_PENDING = 3;          // busy() threshold (engine.js:11-13)

void upload_f32(hipStream_t s, DevBuf& b, const double* src, size_t count) {
  std::vector<float> tmp(count > 0 ? count : 1);
  for (size_t i = 0; i < count; ++i) tmp[i] = src ? finite_f32(src[i]) : 0.0f;
  b.alloc(sizeof(float) * tmp.size());
  FG_HIP(hipMemcpyAsync(b.p, tmp.data(), sizeof(float) * count, hipMemcpyHostToDevice, s));
  FG_HIP(hipStreamSynchronize(s));
}

struct Pass {
  uint64_t id = 0, version = ~0ull;
  int type = 0;
  fg_force_desc d{};
  DevBuf a, b;          // per-node arrays (strengths / radii / targets)
  DevBuf row, ent;      // link CSR
  DevBuf snap;          // link: the snapshot pos + vel (float4 per node)
  uint32_t m = 0;
  float rmax = 0;
  GridClasses grid_classes{0.0f, 0.0f, 0.0f, 0, 0.0f};
  std::unique_ptr<TreeBuffers> tree;
  std::unique_ptr<GridWork> grid;
  DevBuf center;
  DevBuf pairs;         // exact many-body: the per-part sums when the sources are split
  PairsPlan pairs_plan;
  uint32_t pairs_n = 0;  // node count pairs_plan was made for
};

// HIP events around the phases of each tick, summed on request.
struct EventTimer : PhaseTimer {
  hipStream_t stream = nullptr;
  struct Rec {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  hipEvent_t open[FG_PHASE_COUNT] = {};
  std::vector<hipEvent_t> pool;
  double ms[FG_PHASE_COUNT] = {};
  uint64_t launches[FG_PHASE_COUNT] = {};

  hipEvent_t get() {
    hipEvent_t e;
    if (!pool.empty()) {
      e = pool.back();
      pool.pop_back();
    } else {
      FG_HIP(hipEventCreate(&e));
    }
    return e;
  }
  void begin(int phase) override {
    hipEvent_t e = get();
    FG_HIP(hipEventRecord(e, stream));
    open[phase] = e;
  }
  void end(int phase) override {
    hipEvent_t e = get();
    FG_HIP(hipEventRecord(e, stream));
    recs.push_back(Rec{phase, open[phase], e});
    open[phase] = nullptr;
  }
  void collect() {
    for (auto& r : recs) {
      float t = 0;
      FG_HIP(hipEventSynchronize(r.b));
      FG_HIP(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.phase] += t;
      launches[r.phase] += 1;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    recs.clear();
  }
  void reset() {
    collect();
    for (int i = 0; i < FG_PHASE_COUNT; ++i) ms[i] = 0, launches[i] = 0;
  }
  ~EventTimer() {
    for (auto& r : recs) (void)hipEventDestroy(r.a), (void)hipEventDestroy(r.b);
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

struct fg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n = 0;
  int ndim = 2;
  DevBuf pos[3], vel[2][3], fix[3], mask;
  int cur = 0;  // which velocity buffer is current
  bool pinned_any = false;
  std::vector<std::unique_ptr<Pass>> passes;
  uint32_t seed = 0;
  uint64_t steps = 0;
  uint64_t epoch = 1;
  int radix_bits = 8;
  uint32_t walk_packed_max = WALK_PACKED_BY_MODE;
  uint32_t walk_tile = 0;
  uint32_t walk_solo_max = WALK_SOLO_BY_DIMS;
  uint32_t walk_wide_max = WALK_WIDE_BY_DIMS;
  uint32_t walk_group = 0;
  uint32_t walk_pool_k = WALK_POOL_BY_MODE;
  uint32_t walk_exact_max = WALK_EXACT_MAX;
  uint32_t walk_pool_cap = 0xffffffffu;
  bool fuse_link = true;  // a link pass in a one-workgroup tree build's launch (D3FGPU_FUSE_LINK=0: not)
  bool reduce_aligned = true;
  uint32_t reduce_block = 0;
  const Pass* center_fused = nullptr;  // this tick's last pass, a center left to the integrator's launch
  // test hooks, read at creation, so that a host's fallbacks can be driven without a real
  // failure: D3FGPU_CAPACITY=N makes fg_fits refuse more than N nodes; D3FGPU_FAIL=enqueue:K
  // or wait:K makes the K-th fg_enqueue, or the K-th fg_sync / fg_readback_wait, report a
  // HIP failure (the device is not touched)
  uint64_t capacity = 0;
  int fail_kind = 0;  // 1: enqueue, 2: wait
  uint64_t fail_at = 0, fail_calls = 0;
  std::string err;
  // readback
  float* staging = nullptr;
  size_t staging_n = 0;
  hipEvent_t rb_event = nullptr;
  uint64_t rb_ticket = 0, rb_epoch = 0;
  bool rb_inflight = false;
  // AoS mirror for fg_buffers
  DevBuf aos_pos, aos_vel;
  bool mirror = false;
  std::vector<hipEvent_t> pending;
  fg_stats stats{};
  std::unique_ptr<EventTimer> timer;
  PhaseTimer* T() { return timer.get(); }

  float* P(int a) { return pos[a].as<float>(); }
  float* V(int a) { return vel[cur][a].as<float>(); }
  float* W(int a) { return vel[cur ^ 1][a].as<float>(); }
};

namespace {

__global__ void pack_aos(uint32_t n, const float* x, const float* y, const float* z, float4* out) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = make_float4(x[i], y[i], z[i], 0.0f);
}

void set_device(fg_ctx* c) { FG_HIP(hipSetDevice(c->device)); }

int fail(fg_ctx* c, int status, const std::string& msg) {
  if (c) c->err = msg;
  return status;
}

template <typename F>
int guarded(fg_ctx* c, F&& f) {
  if (!c) return FG_ERR_INVALID;
  try {
    set_device(c);
    return f();
  } catch (const ArgError& e) {
    return fail(c, e.status, e.what());
  } catch (const HipError& e) {
    return fail(c, e.code == hipErrorOutOfMemory ? FG_ERR_NOMEM : FG_ERR_HIP, e.what());
  } catch (const std::bad_alloc&) {
    return fail(c, FG_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return fail(c, FG_ERR_HIP, e.what());
  }
}

void zero3(hipStream_t s, DevBuf* b, uint32_t n) {
  for (int a = 0; a < 3; ++a) FG_HIP(hipMemsetAsync(b[a].p, 0, sizeof(float) * (n ? n : 1), s));
}

void build_links(fg_ctx* c, Pass& p, const fg_force_desc& d) {
  const uint32_t n = c->n, m = d.count;
  if (m && !d.links) throw ArgError(FG_ERR_INVALID, "link: links array is null");
  p.row.alloc(sizeof(uint32_t) * ((size_t)n + 1));
  p.ent.alloc(sizeof(LinkEntry) * (m ? 2 * (size_t)m : 1));
  LinkUpload up(c->stream, d);
  link_csr_device(c->stream, n, m, up.links(), up.strengths(), up.distances(), up.bias(), p.row.as<uint32_t>(),
                  p.ent.as<LinkEntry>());
  p.m = m;
}

void update_pass(fg_ctx* c, Pass& p, const fg_force_desc& d) {
  bool fresh = p.version != d.version || p.id != d.id;
  p.d = d;
  p.d.links = nullptr;
  p.d.strengths = p.d.radii = p.d.distances = p.d.bias = p.d.targets = nullptr;
  uint32_t n = c->n;
  switch (d.type) {
    case FG_FORCE_MANYBODY:
      if (fresh) upload_f32(c->stream, p.a, d.strengths, d.strengths ? n : 0);
      if (!d.strengths && fresh) upload_f32(c->stream, p.a, nullptr, n);
      break;
    case FG_FORCE_LINK:
      if (fresh) build_links(c, p, d);
      break;
    case FG_FORCE_COLLIDE:
      if (fresh) {
        upload_f32(c->stream, p.a, d.radii, n);
        p.rmax = 0;
        std::vector<float> r(n > 0 ? n : 1, 0.0f);
        for (uint32_t i = 0; d.radii && i < n; ++i) r[i] = finite_f32(d.radii[i]), p.rmax = std::max(p.rmax, r[i]);
        p.grid_classes = grid_classes(r.data(), n);
      }
      break;
    case FG_FORCE_RADIAL:
      if (fresh) {
        upload_f32(c->stream, p.a, d.strengths, n);
        upload_f32(c->stream, p.b, d.radii, n);
      }
      break;
    case FG_FORCE_POSITION:
      if (d.axis < 0 || d.axis > 2) throw ArgError(FG_ERR_INVALID, "position: axis must be 0, 1 or 2");
      if (fresh) {
        upload_f32(c->stream, p.a, d.strengths, n);
        upload_f32(c->stream, p.b, d.targets, n);
      }
      break;
    case FG_FORCE_CENTER:
      if (!p.center.p) {
        p.center.alloc(sizeof(CenterWork));
        clear_device(p.center.p, sizeof(CenterWork));
      }
      break;
    default:
      throw ArgError(FG_ERR_UNSUPPORTED, "unknown force type " + std::to_string(d.type));
  }
  p.version = d.version;
  p.id = d.id;
}

// Enqueues one force pass.  defer_center: a center pass computes its shift only and
// returns its work area, for the integrator to apply the shift (it must be the tick's last
// pass); otherwise null.
// fuse: for a many-body pass, the center pass whose mean its tree's box kernel takes; for
// that center pass itself, its own address (its shift is then ready).
// link: for a many-body pass whose tree build can take it (tree_link_pass), the link pass
// right before it, which the build's first launch runs (its buffer swap included).
const CenterWork* run_pass(fg_ctx* c, Pass& p, StepParams sp, bool defer_center = false, const Pass* fuse = nullptr,
                           const Pass* link = nullptr) {
  const uint32_t n = c->n;
  const int nd = c->ndim;
  float* pos[3] = {c->P(0), c->P(1), c->P(2)};
  switch (p.type) {
    case FG_FORCE_MANYBODY: {
      float* vel[3] = {c->V(0), c->V(1), c->V(2)};
      if (!(p.d.theta2 > EXACT_THETA2)) {
        PhaseScope ps(c->T(), FG_PHASE_PAIRS);
        if (p.pairs_n != n) {
          p.pairs_plan = pairs_plan(n, device_cus());
          p.pairs_n = n;
          const size_t need = pairs_scratch_bytes(p.pairs_plan, n);
          if (need > p.pairs.bytes) p.pairs.alloc(need);
        }
        launch_manybody_pairs(c->stream, n, nd, pos, p.a.as<float>(), vel, finite_f32(p.d.distance_min2),
                              finite_f32(p.d.distance_max2), sp, p.pairs_plan, p.pairs.as<float>());
      } else {
        if (!p.tree) p.tree.reset(new TreeBuffers());
        TreeParams tp{p.d.theta2, p.d.distance_min2, p.d.distance_max2, c->walk_packed_max, c->walk_tile,
                      c->walk_solo_max, c->walk_wide_max, c->walk_group, c->reduce_block, c->reduce_aligned,
                      c->walk_pool_k, c->walk_pool_cap, c->walk_exact_max};
        if (link) {  // the link pass, from the current velocities into the other set, in the build's launch
          const LinkFuse lf{c->V(0), c->V(1), c->V(2), c->W(0), c->W(1), c->W(2), link->row.as<uint32_t>(),
                            link->ent.as<LinkEntry>(), sp.alpha, sp.step, sp.seed};
          CenterFuse cf;
          if (fuse) cf = CenterFuse{fuse->center.as<CenterWork>(), fuse->d.x, fuse->d.y, fuse->d.z, fuse->d.strength};
          tree_build(c->stream, *p.tree, n, nd, pos[0], pos[1], pos[2], p.a.as<float>(), tp, c->radix_bits, c->T(),
                     fuse ? &cf : nullptr, &lf);
          c->cur ^= 1;
          tree_forces(c->stream, *p.tree, n, nd, c->V(0), c->V(1), c->V(2), tp, sp, c->T());
          break;
        }
        CenterFuse cf;
        if (fuse) cf = CenterFuse{fuse->center.as<CenterWork>(), fuse->d.x, fuse->d.y, fuse->d.z, fuse->d.strength};
        tree_manybody(c->stream, *p.tree, n, nd, pos[0], pos[1], pos[2], p.a.as<float>(), vel[0],
                      vel[1], vel[2], tp, sp, c->radix_bits, c->T(), fuse ? &cf : nullptr);
      }
      break;
    }
    case FG_FORCE_LINK: {
      PhaseScope ps(c->T(), FG_PHASE_LINK);
      if (p.d.iterations > 0 && p.m && p.snap.bytes < sizeof(float4) * (size_t)n) p.snap.alloc(sizeof(float4) * (size_t)n);
      for (int k = 0; k < p.d.iterations && p.m; ++k) {
        const float* vin[3] = {c->V(0), c->V(1), c->V(2)};
        float* vout[3] = {c->W(0), c->W(1), c->W(2)};
        launch_link(c->stream, n, nd, pos, vin, vout, p.row.as<uint32_t>(), p.ent.as<LinkEntry>(),
                    p.snap.as<float4>(), sp);
        c->cur ^= 1;
      }
      break;
    }
    case FG_FORCE_COLLIDE: {
      if (!(p.rmax > 0)) break;
      PhaseScope ps(c->T(), FG_PHASE_COLLIDE);
      for (int k = 0; k < p.d.iterations; ++k) {
        if (n >= GRID_MIN) {
          if (!p.grid) p.grid.reset(new GridWork());
          float* vel[3] = {c->V(0), c->V(1), c->V(2)};
          launch_collide_grid(c->stream, *p.grid, n, nd, pos, vel, p.a.as<float>(), p.grid_classes,
                              (float)p.d.strength, c->radix_bits, sp, k, p.d.iterations);
        } else {
          const float* vin[3] = {c->V(0), c->V(1), c->V(2)};
          float* vout[3] = {c->W(0), c->W(1), c->W(2)};
          launch_collide_pairs(c->stream, n, nd, pos, vin, vout, p.a.as<float>(), (float)p.d.strength, sp);
          c->cur ^= 1;
        }
      }
      break;
    }
    case FG_FORCE_CENTER: {
      PhaseScope ps(c->T(), FG_PHASE_CENTER);
      const bool ready = fuse == &p;  // the many-body box kernel took the mean
      if (defer_center && n <= CENTER_FUSE_MAX && !ready) {  // the integrator's launch does it all
        c->center_fused = &p;
        return nullptr;
      }
      launch_center(c->stream, n, nd, pos, p.center.as<CenterWork>(), p.d.x, p.d.y, p.d.z, p.d.strength,
                    !defer_center, !ready);
      if (defer_center) return p.center.as<CenterWork>();
      break;
    }
    case FG_FORCE_RADIAL: {
      PhaseScope ps(c->T(), FG_PHASE_RADIAL);
      float* vel[3] = {c->V(0), c->V(1), c->V(2)};
      launch_radial(c->stream, n, nd, pos, vel, p.b.as<float>(), p.a.as<float>(), finite_f32(p.d.x),
                    finite_f32(p.d.y), finite_f32(p.d.z), sp);
      break;
    }
    case FG_FORCE_POSITION:
      if (p.d.axis < nd) {
        PhaseScope ps(c->T(), FG_PHASE_POSITION);
        launch_position(c->stream, n, c->P(p.d.axis), c->V(p.d.axis), p.b.as<float>(), p.a.as<float>(), sp);
      }
      break;
    default: break;
  }
  return nullptr;
}

void refresh_mirror(fg_ctx* c) {
  if (!c->mirror || !c->n) return;
  dim3 g(ceil_div(c->n, 256)), b(256);
  hipLaunchKernelGGL(pack_aos, g, b, 0, c->stream, c->n, c->P(0), c->P(1), c->P(2), c->aos_pos.as<float4>());
  hipLaunchKernelGGL(pack_aos, g, b, 0, c->stream, c->n, c->V(0), c->V(1), c->V(2), c->aos_vel.as<float4>());
  FG_LAUNCH_CHECK();
}

void retire_pending(fg_ctx* c) {
  while (!c->pending.empty() && hipEventQuery(c->pending.front()) == hipSuccess) {
    (void)hipEventDestroy(c->pending.front());
    c->pending.erase(c->pending.begin());
  }
}

void ensure_staging(fg_ctx* c) {
  size_t need = 6 * (size_t)(c->n ? c->n : 1);
  if (c->staging && c->staging_n >= need) return;
  if (c->staging) (void)hipHostFree(c->staging);
  c->staging = nullptr;
  FG_HIP(hipHostMalloc((void**)&c->staging, sizeof(float) * need, hipHostMallocDefault));
  c->staging_n = need;
}

void copy_state_async(fg_ctx* c) {
  ensure_staging(c);
  size_t n = c->n;
  for (int a = 0; a < 3; ++a) {
    FG_HIP(hipMemcpyAsync(c->staging + a * n, c->P(a), sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipMemcpyAsync(c->staging + (3 + a) * n, c->V(a), sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  }
}

// The staged readback (float, one array per axis) into the caller's (x, y, z) doubles per
// node.  The copy is bound by the host's memory bandwidth, of which one thread reaches a
// small part: from LAND_SPLIT nodes up it is split over up to LAND_THREADS threads (1M
// nodes: 20 ms on one).
constexpr size_t LAND_SPLIT = 65536;
constexpr unsigned LAND_THREADS = 16;

void land(fg_ctx* c, double* pos, double* vel) {
  const size_t n = c->n;
  const int nd = c->ndim;
  const float* st = c->staging;
  auto part = [=](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i)
      for (int a = 0; a < 3; ++a) {
        const bool on = a < nd;
        if (pos) pos[3 * i + a] = on ? (double)st[a * n + i] : 0.0;
        if (vel) vel[3 * i + a] = on ? (double)st[(3 + a) * n + i] : 0.0;
      }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t parts = n < LAND_SPLIT ? 1 : std::min<size_t>(hw ? hw : 1, LAND_THREADS);
  const size_t chunk = (n + parts - 1) / parts;
  std::vector<std::thread> threads;
  size_t inline_from = parts;  // chunks from here on run on this thread (no thread to spare)
  for (size_t k = 1; k < parts; ++k) {
    const size_t lo = k * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    try {
      threads.emplace_back(part, lo, hi);
    } catch (const std::system_error&) {
      inline_from = k;
      break;
    }
  }
  part(0, std::min(n, chunk));
  for (size_t k = inline_from; k < parts; ++k) part(k * chunk, std::min(n, k * chunk + chunk));
  for (auto& t : threads) t.join();
}

// ---- the sort that allocates nothing (fg_radix_sort_pairs_noalloc) -------------------
//
// Its work -- the ping-pong keys and values, digit histograms, look-back words, tile tickets,
// epoch and error counter of a u32-key sort of at most NOALLOC_MAX_PAIRS pairs -- lives in
// one file-scope device array, which each GPU holds from when the library's code is loaded
// on it until the process ends.  Host code never uses the array's name as an address: the
// device's own address of it is looked up once per device (at fg_create) and kept.
constexpr uint32_t NOALLOC_MAX_PAIRS = 1u << 20;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_KEYS = 0;
constexpr size_t NA_VALS = NA_KEYS + align256(sizeof(uint32_t) * NOALLOC_MAX_PAIRS);
constexpr size_t NA_HIST = NA_VALS + align256(sizeof(uint32_t) * NOALLOC_MAX_PAIRS);
constexpr size_t NA_LOOK = NA_HIST + align256(sizeof(uint32_t) * 8 * 2048);
constexpr uint64_t NA_LOOK_WORDS = sort_u32_lookback_words(NOALLOC_MAX_PAIRS);
constexpr size_t NA_TICK = NA_LOOK + align256(sizeof(uint64_t) * NA_LOOK_WORDS);
constexpr size_t NA_EPOCH = NA_TICK + align256(sizeof(uint32_t) * 16);
constexpr size_t NA_ERR = NA_EPOCH + align256(sizeof(uint32_t) * 4);
constexpr size_t NA_BYTES = NA_ERR + align256(sizeof(uint32_t) * 4);

}  // namespace

static __device__ __align__(256) uint64_t scratch[NA_BYTES / sizeof(uint64_t)];

namespace {

__global__ void clear_words(uint32_t* __restrict__ p, uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

constexpr int NA_DEVICES = 64;
std::mutex na_mu;
uint8_t* na_base[NA_DEVICES] = {};  // each device's address of `scratch`, once set up

// Once per device, with that device current: look up the array's device address and zero
// the look-back words and the epoch (later sorts tell stale words apart by the epoch, which
// every sort advances).  Synchronous, so that a context on another stream of the same device
// finds the array ready.
hipError_t noalloc_setup(int device, hipStream_t stream) {
  if (device < 0 || device >= NA_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(na_mu);
  if (na_base[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  uint8_t* base = (uint8_t*)p;
  const uint64_t words = (NA_ERR - NA_LOOK) / sizeof(uint32_t);
  hipLaunchKernelGGL(clear_words, dim3(1024), dim3(256), 0, stream, (uint32_t*)(base + NA_LOOK), words);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  na_base[device] = base;
  return hipSuccess;
}

uint8_t* noalloc_base(int device) {
  std::lock_guard<std::mutex> lock(na_mu);
  return device >= 0 && device < NA_DEVICES ? na_base[device] : nullptr;
}

}  // namespace

extern "C" {

int fg_abi_version(void) { return FG_ABI_VERSION; }

int fg_device_count(int* count) {
  if (!count) return FG_ERR_INVALID;
  int k = 0;
  if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
  *count = k;
  return FG_OK;
}

int fg_create(fg_ctx** out, int device) {
  if (!out) return FG_ERR_INVALID;
  *out = nullptr;
  int k = 0;
  if (hipGetDeviceCount(&k) != hipSuccess || k <= 0) return FG_ERR_NODEVICE;
  if (device < 0 || device >= k) return FG_ERR_INVALID;
  fg_ctx* c = new (std::nothrow) fg_ctx();
  if (!c) return FG_ERR_NOMEM;
  c->device = device;
  // experiment overrides (the tuned defaults apply otherwise)
  if (const char* v = getenv("D3FGPU_WALK_PACKED_MAX")) c->walk_packed_max = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_TILE")) c->walk_tile = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_SOLO")) c->walk_solo_max = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_WIDE")) c->walk_wide_max = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_GROUP")) c->walk_group = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_POOL_K")) c->walk_pool_k = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_EXACT")) c->walk_exact_max = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_WALK_POOL_CAP")) c->walk_pool_cap = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_FUSE_LINK")) c->fuse_link = atoi(v) != 0;
  if (const char* v = getenv("D3FGPU_REDUCE_ALIGNED")) c->reduce_aligned = atoi(v) != 0;
  if (const char* v = getenv("D3FGPU_REDUCE_BLOCK")) c->reduce_block = (uint32_t)atoi(v);
  if (const char* v = getenv("D3FGPU_RADIX_BITS")) c->radix_bits = atoi(v) == 11 ? 11 : 8;
  if (const char* v = getenv("D3FGPU_CAPACITY")) c->capacity = strtoull(v, nullptr, 10);
  if (const char* v = getenv("D3FGPU_FAIL")) {
    const char* colon = strchr(v, ':');
    if (colon) {
      c->fail_kind = strncmp(v, "enqueue", (size_t)(colon - v)) == 0 ? 1 : strncmp(v, "wait", (size_t)(colon - v)) == 0 ? 2 : 0;
      c->fail_at = strtoull(colon + 1, nullptr, 10);
    }
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FG_ERR_HIP;
  }
  if (hipEventCreateWithFlags(&c->rb_event, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FG_ERR_HIP;
  }
  if (noalloc_setup(device, c->stream) != hipSuccess) {
    (void)hipEventDestroy(c->rb_event);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return FG_ERR_HIP;
  }
  *out = c;
  return FG_OK;
}

int fg_destroy(fg_ctx* c) {
  if (!c) return FG_ERR_INVALID;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->passes.clear();
  for (hipEvent_t e : c->pending) (void)hipEventDestroy(e);
  c->pending.clear();
  if (c->staging) (void)hipHostFree(c->staging);
  if (c->rb_event) (void)hipEventDestroy(c->rb_event);
  for (int a = 0; a < 3; ++a) {
    c->pos[a].free();
    c->vel[0][a].free();
    c->vel[1][a].free();
    c->fix[a].free();
  }
  c->mask.free();
  c->aos_pos.free();
  c->aos_vel.free();
  (void)hipStreamDestroy(c->stream);
  delete c;
  return FG_OK;
}

const char* fg_last_error(const fg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int fg_fits(fg_ctx* c, uint32_t n, int32_t ndim, int32_t* fits) {
  return guarded(c, [&] {
    if (!fits || ndim < 1 || ndim > 3) return FG_ERR_INVALID;
    size_t free_b = 0, total_b = 0;
    FG_HIP(hipMemGetInfo(&free_b, &total_b));
    // state, pins, tree (keys, sort scratch, leaves, links, records, sums), grid, link CSR
    double per_node = 4.0 * 9 + 1 + 8 * 2 + 4 * 2 + 16 + 4 * 3 + 16 + 64 + 40 + 16 + 8 + 64;
    *fits = (n < 0x7fffffffu && per_node * n < 0.8 * (double)free_b) ? 1 : 0;
    if (c->capacity && n > c->capacity) *fits = 0;
    return FG_OK;
  });
}

int fg_set_nodes(fg_ctx* c, uint32_t n, int32_t ndim, const double* pos, const double* vel,
                 const double* fix, const uint8_t* fixmask) {
  return guarded(c, [&] {
    if (ndim < 1 || ndim > 3) return fail(c, FG_ERR_INVALID, "ndim must be 1, 2 or 3");
    if (n && !pos) return fail(c, FG_ERR_INVALID, "positions are required");
    FG_HIP(hipStreamSynchronize(c->stream));
    c->passes.clear();
    c->n = n;
    c->ndim = ndim;
    c->cur = 0;
    c->steps = 0;
    c->epoch++;
    c->rb_inflight = false;
    c->stats = fg_stats{};
    size_t b = sizeof(float) * (n ? n : 1);
    for (int a = 0; a < 3; ++a) {
      c->pos[a].alloc(b);
      c->vel[0][a].alloc(b);
      c->vel[1][a].alloc(b);
      c->fix[a].alloc(b);
    }
    c->mask.alloc(n ? n : 1);
    if (c->mirror) {
      c->aos_pos.alloc(16 * (size_t)(n ? n : 1));
      c->aos_vel.alloc(16 * (size_t)(n ? n : 1));
    }
    std::vector<float> h(6 * (size_t)(n ? n : 1), 0.0f);
    std::vector<uint8_t> m(n ? n : 1, 0);
    c->pinned_any = false;
    for (uint32_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        bool on = a < ndim;
        h[a * (size_t)n + i] = on ? finite_f32(pos[3 * (size_t)i + a]) : 0.0f;
        h[(3 + a) * (size_t)n + i] = (on && vel) ? finite_f32(vel[3 * (size_t)i + a]) : 0.0f;
      }
    for (int a = 0; a < 3; ++a) {
      FG_HIP(hipMemcpyAsync(c->pos[a].p, h.data() + a * (size_t)n, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
      FG_HIP(hipMemcpyAsync(c->vel[0][a].p, h.data() + (3 + a) * (size_t)n, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
      FG_HIP(hipStreamSynchronize(c->stream));
    }
    zero3(c->stream, c->vel[1], n);
    std::fill(h.begin(), h.end(), 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
      uint8_t mm = fixmask ? (uint8_t)(fixmask[i] & ((1u << ndim) - 1u)) : 0;
      m[i] = mm;
      if (mm) c->pinned_any = true;
      for (int a = 0; a < 3; ++a)
        if ((mm >> a) & 1) h[a * (size_t)n + i] = fix ? finite_f32(fix[3 * (size_t)i + a]) : 0.0f;
    }
    for (int a = 0; a < 3; ++a)
      FG_HIP(hipMemcpyAsync(c->fix[a].p, h.data() + a * (size_t)n, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipMemcpyAsync(c->mask.p, m.data(), n, hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    refresh_mirror(c);
    return FG_OK;
  });
}

static int update_vec(fg_ctx* c, DevBuf* dst, uint32_t first, uint32_t count, const double* src) {
  return guarded(c, [&] {
    if (!src || (uint64_t)first + count > c->n) return fail(c, FG_ERR_INVALID, "range out of bounds");
    std::vector<float> h(3 * (size_t)(count ? count : 1));
    for (uint32_t i = 0; i < count; ++i)
      for (int a = 0; a < 3; ++a) h[a * (size_t)count + i] = a < c->ndim ? finite_f32(src[3 * (size_t)i + a]) : 0.0f;
    for (int a = 0; a < 3; ++a)
      FG_HIP(hipMemcpyAsync(dst[a].as<float>() + first, h.data() + a * (size_t)count, sizeof(float) * count,
                            hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    c->epoch++;
    return FG_OK;
  });
}

int fg_update_positions(fg_ctx* c, uint32_t first, uint32_t count, const double* pos) {
  if (!c) return FG_ERR_INVALID;
  return update_vec(c, c->pos, first, count, pos);
}

int fg_update_velocities(fg_ctx* c, uint32_t first, uint32_t count, const double* vel) {
  if (!c) return FG_ERR_INVALID;
  return update_vec(c, c->vel[c->cur], first, count, vel);
}

int fg_update_pins(fg_ctx* c, uint32_t first, uint32_t count, const double* fix, const uint8_t* fixmask) {
  return guarded(c, [&] {
    if (!fixmask || (uint64_t)first + count > c->n) return fail(c, FG_ERR_INVALID, "range out of bounds");
    std::vector<float> h(3 * (size_t)(count ? count : 1), 0.0f);
    std::vector<uint8_t> m(count ? count : 1);
    for (uint32_t i = 0; i < count; ++i) {
      uint8_t mm = (uint8_t)(fixmask[i] & ((1u << c->ndim) - 1u));
      m[i] = mm;
      if (mm) c->pinned_any = true;
      for (int a = 0; a < 3; ++a)
        if ((mm >> a) & 1) h[a * (size_t)count + i] = fix ? finite_f32(fix[3 * (size_t)i + a]) : 0.0f;
    }
    for (int a = 0; a < 3; ++a)
      FG_HIP(hipMemcpyAsync(c->fix[a].as<float>() + first, h.data() + a * (size_t)count, sizeof(float) * count,
                            hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipMemcpyAsync(c->mask.as<uint8_t>() + first, m.data(), count, hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    return FG_OK;
  });
}

int fg_set_forces(fg_ctx* c, uint32_t k, const fg_force_desc* descs) {
  return guarded(c, [&] {
    if (k && !descs) return fail(c, FG_ERR_INVALID, "descriptors are null");
    for (uint32_t i = 0; i < k; ++i)
      if (descs[i].type < FG_FORCE_CENTER || descs[i].type > FG_FORCE_RADIAL)
        return fail(c, FG_ERR_UNSUPPORTED, "unsupported force type " + std::to_string(descs[i].type));
    FG_HIP(hipStreamSynchronize(c->stream));
    std::vector<std::unique_ptr<Pass>> next;
    for (uint32_t i = 0; i < k; ++i) {
      std::unique_ptr<Pass> p;
      for (auto& q : c->passes)
        if (q && q->id == descs[i].id && q->type == descs[i].type) {
          p = std::move(q);
          break;
        }
      if (!p) {
        p.reset(new Pass());
        p->type = descs[i].type;
      }
      update_pass(c, *p, descs[i]);
      next.push_back(std::move(p));
    }
    c->passes = std::move(next);
    return FG_OK;
  });
}

int fg_set_seed(fg_ctx* c, uint32_t seed) {
  if (!c) return FG_ERR_INVALID;
  c->seed = seed;
  return FG_OK;
}

namespace {
// The test hook's failure for call kind `kind`, when this is the call it names.
bool injected(fg_ctx* c, int kind) {
  return c->fail_kind == kind && ++c->fail_calls == c->fail_at;
}
}  // namespace

int fg_enqueue(fg_ctx* c, uint32_t steps, const double* alphas, double decay) {
  return guarded(c, [&] {
    if (injected(c, 1)) return fail(c, FG_ERR_HIP, "injected failure (D3FGPU_FAIL)");
    if (steps && !alphas) return fail(c, FG_ERR_INVALID, "alphas are null");
    if (!c->n) return FG_OK;
    const uint8_t* mask = c->pinned_any ? c->mask.as<uint8_t>() : nullptr;
    for (uint32_t s = 0; s < steps; ++s) {
      StepParams sp{(float)alphas[s], (float)decay, (uint32_t)c->steps, c->seed};
      PhaseScope tick(c->T(), FG_PHASE_TICK);
      // a center pass last in the list leaves its shift to the integrator (one launch fewer)
      const CenterWork* shift = nullptr;
      c->center_fused = nullptr;
      // the first center pass, when a many-body pass before it builds a tree with the box
      // kernel: that kernel reads the same positions (only center moves nodes during the
      // forces) and takes the center's mean on its way, one launch fewer
      int fuse_mb = -1, fuse_ct = -1;
      for (size_t i = 0; i < c->passes.size(); ++i) {
        const Pass& q = *c->passes[i];
        if (q.type == FG_FORCE_CENTER) {
          if (fuse_mb >= 0) fuse_ct = (int)i;
          break;
        }
        if (q.type == FG_FORCE_MANYBODY && fuse_mb < 0 && q.d.theta2 > EXACT_THETA2 && tree_box_pass(c->n))
          fuse_mb = (int)i;
      }
      const Pass* fused = fuse_ct >= 0 ? c->passes[fuse_ct].get() : nullptr;
      // a link pass of one iteration right before a many-body pass, on a graph small enough
      // for the one-launch link pass: the tree build's first launch (the one-workgroup build,
      // or the box kernel) runs it in workgroups of its own on CUs it leaves idle (link only
      // changes velocities, the build reads positions), one launch fewer
      int link_mb = -1;
      for (size_t i = 0; i + 1 < c->passes.size() && c->fuse_link; ++i) {
        const Pass& a = *c->passes[i];
        const Pass& b = *c->passes[i + 1];
        if (a.type == FG_FORCE_LINK && a.d.iterations == 1 && a.m && b.type == FG_FORCE_MANYBODY &&
            b.d.theta2 > EXACT_THETA2 && tree_link_pass(c->n)) {
          link_mb = (int)i + 1;
          break;
        }
      }
      for (size_t i = 0; i < c->passes.size(); ++i) {
        if ((int)i + 1 == link_mb) continue;  // run by the next pass's build
        shift = run_pass(c, *c->passes[i], sp, i + 1 == c->passes.size(),
                         (int)i == fuse_mb || (int)i == fuse_ct ? fused : nullptr,
                         (int)i == link_mb ? c->passes[i - 1].get() : nullptr);
      }
      float* pos[3] = {c->P(0), c->P(1), c->P(2)};
      float* vel[3] = {c->V(0), c->V(1), c->V(2)};
      float* fix[3] = {c->fix[0].as<float>(), c->fix[1].as<float>(), c->fix[2].as<float>()};
      {
        PhaseScope ps(c->T(), FG_PHASE_INTEGRATE);
        if (const Pass* cp = c->center_fused)
          launch_center_integrate(c->stream, c->n, c->ndim, pos, vel, mask, fix, sp, cp->d.x, cp->d.y, cp->d.z,
                                  cp->d.strength);
        else
          launch_integrate(c->stream, c->n, c->ndim, pos, vel, mask, fix, sp, shift);
      }
      c->steps++;
    }
    refresh_mirror(c);
    hipEvent_t e;
    FG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    FG_HIP(hipEventRecord(e, c->stream));
    c->pending.push_back(e);
    retire_pending(c);
    return FG_OK;
  });
}

int fg_apply_forces(fg_ctx* c, double alpha) {
  return guarded(c, [&] {
    if (!c->n) return FG_OK;
    StepParams sp{(float)alpha, 1.0f, (uint32_t)c->steps, c->seed};
    for (auto& p : c->passes) run_pass(c, *p, sp);
    return FG_OK;
  });
}

int fg_busy(fg_ctx* c, int32_t* busy) {
  return guarded(c, [&] {
    if (!busy) return FG_ERR_INVALID;
    retire_pending(c);
    *busy = (int32_t)c->pending.size() >= MAX_PENDING ? 1 : 0;
    return FG_OK;
  });
}

int fg_sync(fg_ctx* c) {
  return guarded(c, [&] {
    if (injected(c, 2)) return fail(c, FG_ERR_HIP, "injected failure (D3FGPU_FAIL)");
    FG_HIP(hipStreamSynchronize(c->stream));
    retire_pending(c);
    return FG_OK;
  });
}

int fg_readback(fg_ctx* c, double* pos, double* vel) {
  return guarded(c, [&] {
    if (!c->n) return FG_OK;
    copy_state_async(c);
    FG_HIP(hipStreamSynchronize(c->stream));
    retire_pending(c);
    land(c, pos, vel);
    return FG_OK;
  });
}

int fg_readback_start(fg_ctx* c, uint64_t* ticket) {
  return guarded(c, [&] {
    if (!ticket) return FG_ERR_INVALID;
    if (c->rb_inflight) return fail(c, FG_ERR_STATE, "a readback is already in flight");
    if (c->n) {
      copy_state_async(c);
      FG_HIP(hipEventRecord(c->rb_event, c->stream));
    }
    c->rb_inflight = true;
    c->rb_epoch = c->epoch;
    *ticket = ++c->rb_ticket;
    return FG_OK;
  });
}

int fg_readback_wait(fg_ctx* c, uint64_t ticket) {
  return guarded(c, [&] {
    if (!c->rb_inflight || ticket != c->rb_ticket) return fail(c, FG_ERR_STATE, "no such readback in flight");
    if (c->n) FG_HIP(hipEventSynchronize(c->rb_event));
    if (injected(c, 2)) return fail(c, FG_ERR_HIP, "injected failure (D3FGPU_FAIL)");
    return FG_OK;
  });
}

int fg_readback_fetch(fg_ctx* c, uint64_t ticket, double* pos, double* vel) {
  return guarded(c, [&] {
    if (!c->rb_inflight || ticket != c->rb_ticket) return fail(c, FG_ERR_STATE, "no such readback in flight");
    if (c->n) FG_HIP(hipEventSynchronize(c->rb_event));
    c->rb_inflight = false;
    if (c->rb_epoch != c->epoch) return fail(c, FG_ERR_STATE, "the state was replaced after the readback started");
    land(c, pos, vel);
    return FG_OK;
  });
}

int fg_buffers(fg_ctx* c, fg_buffer_info* out) {
  return guarded(c, [&] {
    if (!out) return FG_ERR_INVALID;
    if (!c->mirror) {
      c->mirror = true;
      c->aos_pos.alloc(16 * (size_t)(c->n ? c->n : 1));
      c->aos_vel.alloc(16 * (size_t)(c->n ? c->n : 1));
      refresh_mirror(c);
      FG_HIP(hipStreamSynchronize(c->stream));
    }
    out->positions = c->aos_pos.p;
    out->velocities = c->aos_vel.p;
    out->count = c->n;
    out->stride = 16;
    out->epoch = c->epoch;
    return FG_OK;
  });
}

int fg_debug_copy_from_device(fg_ctx* c, const void* device_src, uint64_t bytes, void* host_dst) {
  return guarded(c, [&] {
    if (!device_src || !host_dst) return FG_ERR_INVALID;
    FG_HIP(hipStreamSynchronize(c->stream));
    FG_HIP(hipMemcpy(host_dst, device_src, (size_t)bytes, hipMemcpyDeviceToHost));
    return FG_OK;
  });
}

int fg_get_stats(fg_ctx* c, fg_stats* out) {
  return guarded(c, [&] {
    if (!out) return FG_ERR_INVALID;
    FG_HIP(hipStreamSynchronize(c->stream));
    fg_stats s{};
    s.steps = c->steps;
    s.sort_bits = (uint32_t)c->radix_bits;
    for (auto& p : c->passes) {
      if (p->tree && p->tree->counters) {
        uint32_t h[16];
        FG_HIP(hipMemcpy(h, p->tree->counters, sizeof(h), hipMemcpyDeviceToHost));
        s.walk_max_stack = h[8];
        s.interactions = (uint64_t)h[0] | ((uint64_t)h[1] << 32);
        s.opened = (uint64_t)h[2] | ((uint64_t)h[3] << 32);
        s.lane_visits = (uint64_t)h[6] | ((uint64_t)h[7] << 32);
        s.walk_errors = h[4];
        if (const uint32_t nb = p->tree->slot_blocks) {  // the wide walk's counts, per workgroup
          std::vector<uint32_t> sl((size_t)8 * nb);
          FG_HIP(hipMemcpy(sl.data(), p->tree->slots, sizeof(uint32_t) * sl.size(), hipMemcpyDeviceToHost));
          for (uint32_t b = 0; b < nb; ++b) {
            const uint32_t* q = &sl[(size_t)8 * b];
            const uint64_t op = (uint64_t)q[2] | ((uint64_t)q[3] << 32);
            s.interactions += (uint64_t)q[0] | ((uint64_t)q[1] << 32);
            s.opened += op;
            s.lane_visits += op;  // one point per opening
            s.walk_errors += q[4];
            s.walk_max_stack = std::max(s.walk_max_stack, q[5]);
          }
        }
        uint32_t e[4] = {0, 0, 0, 0};
        if (p->tree->sort.errors) FG_HIP(hipMemcpy(e, p->tree->sort.errors, sizeof(e), hipMemcpyDeviceToHost));
        s.walk_errors += e[0];
        break;
      }
    }
    *out = s;
    return FG_OK;
  });
}

int fg_timing_enable(fg_ctx* c, int32_t on) {
  return guarded(c, [&] {
    if (on) {
      if (!c->timer) {
        c->timer.reset(new EventTimer());
        c->timer->stream = c->stream;
      }
      c->timer->reset();
    } else if (c->timer) {
      FG_HIP(hipStreamSynchronize(c->stream));
      c->timer.reset();
    }
    return FG_OK;
  });
}

int fg_get_timing(fg_ctx* c, fg_timing* out) {
  return guarded(c, [&] {
    if (!out) return FG_ERR_INVALID;
    *out = fg_timing{};
    if (!c->timer) return FG_OK;
    c->timer->collect();
    for (int i = 0; i < FG_PHASE_COUNT; ++i) out->ms[i] = c->timer->ms[i], out->launches[i] = c->timer->launches[i];
    return FG_OK;
  });
}

int fg_radix_sort_pairs(fg_ctx* c, uint32_t n, const uint32_t* keys, const uint32_t* values,
                        uint32_t first_bit, uint32_t bits, uint32_t* keys_out, uint32_t* values_out) {
  return guarded(c, [&] {
    if (n && (!keys || !values || !keys_out || !values_out)) return FG_ERR_INVALID;
    if (first_bit >= 32 || bits == 0 || first_bit + bits > 32) return fail(c, FG_ERR_INVALID, "bit range outside the key");
    if (!n) return FG_OK;
    DevBuf k, v;
    k.alloc(sizeof(uint32_t) * n);
    v.alloc(sizeof(uint32_t) * n);
    FG_HIP(hipMemcpyAsync(k.p, keys, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    FG_HIP(hipMemcpyAsync(v.p, values, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    SortWork w;
    uint32_t *ok, *ov;
    radix_sort_u32(c->stream, w, c->radix_bits, k.as<uint32_t>(), v.as<uint32_t>(), n, first_bit, bits, &ok, &ov);
    FG_HIP(hipMemcpyAsync(keys_out, ok, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipMemcpyAsync(values_out, ov, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    uint32_t e = 0;
    FG_HIP(hipMemcpy(&e, w.errors, sizeof(e), hipMemcpyDeviceToHost));
    if (e) return fail(c, FG_ERR_HIP, "sort: look-back wait bound reached");
    return FG_OK;
  });
}

int fg_radix_sort_pairs_noalloc(fg_ctx* c, uint32_t n, const uint32_t* keys, const uint32_t* values,
                                uint32_t first_bit, uint32_t bits, uint32_t* keys_out, uint32_t* values_out) {
  return guarded(c, [&] {
    if (n && (!keys || !values || !keys_out || !values_out)) return FG_ERR_INVALID;
    if (first_bit >= 32 || bits == 0 || first_bit + bits > 32) return fail(c, FG_ERR_INVALID, "bit range outside the key");
    if (n > NOALLOC_MAX_PAIRS)
      return fail(c, FG_ERR_INVALID, "fg_radix_sort_pairs_noalloc: " + std::to_string(n) + " pairs exceed the " +
                                         std::to_string(NOALLOC_MAX_PAIRS) + " its scratch holds");
    if (!n) return FG_OK;
    uint8_t* base = noalloc_base(c->device);
    if (!base) return fail(c, FG_ERR_STATE, "fg_radix_sort_pairs_noalloc: scratch not set up on this device");
    const size_t bytes = sizeof(uint32_t) * n;
    if (keys_out != keys) FG_HIP(hipMemcpyAsync(keys_out, keys, bytes, hipMemcpyDeviceToDevice, c->stream));
    if (values_out != values) FG_HIP(hipMemcpyAsync(values_out, values, bytes, hipMemcpyDeviceToDevice, c->stream));
    uint32_t* errors = (uint32_t*)(base + NA_ERR);
    hipLaunchKernelGGL(clear_words, dim3(1), dim3(64), 0, c->stream, errors, (uint64_t)4);
    FG_LAUNCH_CHECK();
    SortWork w;
    w.owned = false;
    w.keys_alt = base + NA_KEYS;
    w.vals_alt = (uint32_t*)(base + NA_VALS);
    w.ghist = (uint32_t*)(base + NA_HIST);
    w.lookback = (uint64_t*)(base + NA_LOOK);
    w.tickets = (uint32_t*)(base + NA_TICK);
    w.epoch = (uint32_t*)(base + NA_EPOCH);
    w.errors = errors;
    w.cap_n = NOALLOC_MAX_PAIRS;
    w.cap_key_bytes = sizeof(uint32_t);
    w.lookback_words = NA_LOOK_WORDS;
    uint32_t *ok, *ov;
    radix_sort_u32(c->stream, w, c->radix_bits, keys_out, values_out, n, first_bit, bits, &ok, &ov);
    if (ok != keys_out) FG_HIP(hipMemcpyAsync(keys_out, ok, bytes, hipMemcpyDeviceToDevice, c->stream));
    if (ov != values_out) FG_HIP(hipMemcpyAsync(values_out, ov, bytes, hipMemcpyDeviceToDevice, c->stream));
    uint32_t e = 0;
    FG_HIP(hipMemcpyAsync(&e, errors, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    FG_HIP(hipStreamSynchronize(c->stream));
    if (e) return fail(c, FG_ERR_HIP, "sort: look-back wait bound reached");
    return FG_OK;
  });
}

int fg_debug_tree(fg_ctx* c, float* cube, int32_t* shift, uint32_t* codes, uint32_t* order,
                  uint32_t* leaf_depth, uint32_t* left, uint32_t* right, uint32_t* depth,
                  uint32_t* kids, double* com, double* value) {
  return guarded(c, [&] {
    TreeBuffers* t = nullptr;
    for (auto& p : c->passes)
      if (p->tree && p->tree->skeys) {
        t = p->tree.get();
        break;
      }
    if (!t) return fail(c, FG_ERR_STATE, "no many-body tree has been built");
    FG_HIP(hipStreamSynchronize(c->stream));
    uint32_t n = c->n, inner = n > 1 ? n - 1 : 0;
    const int L = c->ndim == 3 ? 20 : 24, half = L / 2 * c->ndim;
    if (cube || shift) {
      CubeInfo ci;
      FG_HIP(hipMemcpy(&ci, t->cube, sizeof(ci), hipMemcpyDeviceToHost));
      if (cube) cube[0] = ci.origin[0], cube[1] = ci.origin[1], cube[2] = ci.origin[2], cube[3] = ci.size;
      if (shift) *shift = ci.shift;
    }
    std::vector<uint32_t> ord(n ? n : 1);
    FG_HIP(hipMemcpy(ord.data(), t->order, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (order) memcpy(order, ord.data(), sizeof(uint32_t) * n);
    if (codes) {
      std::vector<uint64_t> sk(n ? n : 1);
      FG_HIP(hipMemcpy(sk.data(), t->skeys, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
      for (uint32_t s = 0; s < n; ++s) {
        uint64_t key = sk[s];
        uint32_t hi = (uint32_t)(key >> half), lo = (uint32_t)(key & ((1ull << half) - 1));
        codes[2 * ord[s]] = hi << (32 - half);
        codes[2 * ord[s] + 1] = lo << (32 - half);
      }
    }
    if (leaf_depth) FG_HIP(hipMemcpy(leaf_depth, t->leaf_depth, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (inner && (left || right || depth)) {
      std::vector<uint4> ln(inner);
      FG_HIP(hipMemcpy(ln.data(), t->links, sizeof(uint4) * inner, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < inner; ++i) {
        if (left) left[i] = ln[i].x;
        if (right) right[i] = ln[i].y;
        if (depth) depth[i] = ln[i].z;
      }
    }
    if (inner && (kids || com || value)) {
      std::vector<NodeRec> r(inner);
      FG_HIP(hipMemcpy(r.data(), t->rec, sizeof(NodeRec) * inner, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < inner; ++i) {
        if (kids) memcpy(kids + 8 * (size_t)i, r[i].kids, sizeof(r[i].kids));
        if (com) com[3 * i] = r[i].cx, com[3 * i + 1] = r[i].cy, com[3 * i + 2] = r[i].cz;
        if (value) value[i] = r[i].value;
      }
    }
    return FG_OK;
  });
}

}  // extern "C"
