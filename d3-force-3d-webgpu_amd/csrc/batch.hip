// batch.hip -- the batch engine behind fg_batch_* (include/d3fgpu.h).
//
// Jobs are resident: each owns one device allocation holding its positions, two velocity
// sets, pins and force arrays, and a record in the device job table (BJob).  The jobs of
// one dimensionality form a group with three index spaces (nodes; jobs with a many-body
// force; jobs with a collide force).  Adding or removing a job rebuilds the spaces' segment
// tables and tiles, never moves a job's state.  A tick is
//
//   advance (every job's alpha, step and status, on the device)
//   per group: link iterations, the segmented tree + persistent walk, collide iterations
//              on the grid, center, radial, x/y/z, integrate
//
// captured once per layout into a hipGraph and replayed.  Jobs finish on the device after
// the tick whose alpha falls below alphaMin; the host sees it when it reads the table.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/d3fgpu.h"
#include "batch.h"
#include "common.h"
#include "forces.h"
#include "host.h"
#include "sort.h"

using namespace fg;

namespace {

constexpr uint32_t MAX_SLOTS = 4096;

// Bytes of scratch per point of each space (what admission charges a job for its share).
constexpr uint64_t NODE_SCRATCH = 24;   // segment of the node, tiles, link snapshot
constexpr uint64_t TREE_SCRATCH = 240;  // codes, sort buffers, leaves, links, records, sums
constexpr uint64_t GRID_SCRATCH = 64;   // slots, sort buffers, snapshot, table

// One resident job on the host side.
struct JobHost {
  uint64_t id = 0;
  uint32_t slot = 0;
  uint32_t n = 0;
  int ndim = 0;
  void* mem = nullptr;  // the job's one device allocation
  uint64_t charge = 0;  // bytes charged against the budget
  BJob rec{};           // the record as submitted (the device copy evolves)
  bool tree = false, grid = false;
};

struct Group {
  int ndim = 0;
  BSpace node, tree, grid;
  DevBuf node_seg_of, node_seg_job, node_tile;
  DevBuf tree_seg_of, tree_seg_job, tree_tile, tree_walk;
  DevBuf grid_seg_of, grid_seg_job;
  DevBuf snap;  // link snapshot, one float4 per node-space index
  BatchTree btree;
  BatchGrid bgrid;
  BatchWork work;
  uint32_t max_link = 0, max_coll = 0;
  bool center = false, radial = false, pos[3] = {false, false, false};
};

// The layout fields of a job record, patched on the device when the spaces change.
struct LayoutPatch {
  uint32_t slot;
  uint32_t f[8];  // n, node_off, tree_off, grid_off, node_tiles_first, node_tiles, tree_tiles_first, tree_tiles
};

__global__ void patch_layout(BJob* jobs, const LayoutPatch* p, uint32_t count) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= count) return;
  BJob& B = jobs[p[k].slot];
  B.n = p[k].f[0];
  B.node_off = p[k].f[1];
  B.tree_off = p[k].f[2];
  B.grid_off = p[k].f[3];
  B.node_tiles_first = p[k].f[4];
  B.node_tiles = p[k].f[5];
  B.tree_tiles_first = p[k].f[6];
  B.tree_tiles = p[k].f[7];
}

// seg_of[i] for a space whose segments start at starts[0..segs) (ascending).
__global__ void fill_seg_of(uint32_t total, const uint32_t* __restrict__ starts, uint32_t segs,
                            uint32_t* __restrict__ seg_of) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  uint32_t lo = 0, hi = segs;  // the last start <= i
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (starts[mid] <= i) lo = mid;
    else hi = mid;
  }
  seg_of[i] = lo;
}

__global__ void copy_f32(uint32_t n, const float* __restrict__ a, float* __restrict__ b) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

struct fg_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  uint64_t budget = 0, charged = 0;
  BJob* table = nullptr;  // device job table [MAX_SLOTS]
  std::vector<JobHost*> jobs;  // submission order
  std::vector<JobHost*> slot_job;
  uint32_t slots_used = 0;     // highest slot in use + 1
  uint64_t next_id = 1;
  Group groups[3];
  bool dirty = true;
  bool use_graphs = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t graph_nodes = 0;
  BatchTune tune;
  uint64_t ticks = 0, graph_builds = 0;
  std::string err;
  std::vector<BJob> host_table;

  ~fg_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (JobHost* j : jobs) {
      if (j->mem) (void)hipFree(j->mem);
      delete j;
    }
    if (table) (void)hipFree(table);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int bfail(fg_batch* b, int status, const std::string& msg) {
  if (b) b->err = msg;
  return status;
}

template <typename F>
int bguard(fg_batch* b, F&& f) {
  if (!b) return FG_ERR_INVALID;
  try {
    FG_HIP(hipSetDevice(b->device));
    return f();
  } catch (const ArgError& e) {
    return bfail(b, e.status, e.what());
  } catch (const HipError& e) {
    return bfail(b, e.code == hipErrorOutOfMemory ? FG_ERR_NOMEM : FG_ERR_HIP, e.what());
  } catch (const std::bad_alloc&) {
    return bfail(b, FG_ERR_NOMEM, "host memory exhausted");
  } catch (const std::exception& e) {
    return bfail(b, FG_ERR_HIP, e.what());
  }
}

void drop_graph(fg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_nodes = 0;
}

// ---- job description -> device allocation + record -----------------------------------

int force_rank(const fg_force_desc& d) {
  switch (d.type) {
    case FG_FORCE_LINK: return 0;
    case FG_FORCE_MANYBODY: return 1;
    case FG_FORCE_COLLIDE: return 2;
    case FG_FORCE_CENTER: return 3;
    case FG_FORCE_RADIAL: return 4;
    case FG_FORCE_POSITION: return 5 + (d.axis >= 0 && d.axis <= 2 ? d.axis : 0);
    default: return -1;
  }
}

// Host image of one job's allocation: the arrays and their offsets.
struct JobImage {
  std::vector<uint8_t> bytes;
  size_t off_p[3], off_v[2][3], off_fix[3], off_mask = 0, off_mb = 0, off_row = 0, off_ent = 0,
      off_radius = 0, off_rr = 0, off_rw = 0, off_goal[3] = {0, 0, 0}, off_weight[3] = {0, 0, 0};
  bool mask = false, mb = false, link = false, collide = false, radial = false, pos[3] = {false, false, false};
  float rmax = 0;
  GridClasses gc{0.0f, 0.0f, 0.0f, 0};
  uint32_t link_iters = 0, coll_iters = 0;
  const fg_force_desc* fd[8] = {};
  uint64_t charge = 0;
  bool tree = false, grid = false;
};

void check_job(const fg_job_desc* j) {
  if (!j) throw ArgError(FG_ERR_INVALID, "job is null");
  if (j->n < 1) throw ArgError(FG_ERR_INVALID, "a job needs at least one node");
  if (j->ndim < 1 || j->ndim > 3) throw ArgError(FG_ERR_INVALID, "ndim must be 1, 2 or 3");
  if (!j->pos) throw ArgError(FG_ERR_INVALID, "positions are required");
  if (j->nforces && !j->forces) throw ArgError(FG_ERR_INVALID, "forces are null");
  int last = -1;
  for (uint32_t k = 0; k < j->nforces; ++k) {
    const int r = force_rank(j->forces[k]);
    if (r < 0) throw ArgError(FG_ERR_UNSUPPORTED, "unsupported force type " + std::to_string(j->forces[k].type));
    if (j->forces[k].type == FG_FORCE_POSITION && (j->forces[k].axis < 0 || j->forces[k].axis > 2))
      throw ArgError(FG_ERR_INVALID, "position: axis must be 0, 1 or 2");
    if (r <= last)
      throw ArgError(FG_ERR_UNSUPPORTED,
                     "batched jobs apply forces in the order link, manyBody, collide, center, radial, x, y, z, "
                     "each at most once");
    last = r;
  }
}

// Lays out (and, with fill, converts) everything a job owns.
void job_image(const fg_job_desc* j, JobImage& im, bool fill) {
  const uint32_t n = j->n;
  const size_t fb = align256(sizeof(float) * n);
  size_t at = 0;
  for (int a = 0; a < 3; ++a) im.off_p[a] = at, at += fb;
  for (int c = 0; c < 2; ++c)
    for (int a = 0; a < 3; ++a) im.off_v[c][a] = at, at += fb;
  bool pinned = false;
  if (j->fixmask)
    for (uint32_t i = 0; i < n && !pinned; ++i) pinned = (j->fixmask[i] & ((1u << j->ndim) - 1u)) != 0;
  im.mask = pinned;
  if (pinned) {
    for (int a = 0; a < 3; ++a) im.off_fix[a] = at, at += fb;
    im.off_mask = at, at += align256(n);
  }
  for (uint32_t k = 0; k < j->nforces; ++k) {
    const fg_force_desc& d = j->forces[k];
    im.fd[force_rank(d)] = &d;
    switch (d.type) {
      case FG_FORCE_MANYBODY:
        im.mb = true;
        im.off_mb = at, at += fb;
        break;
      case FG_FORCE_LINK:
        im.link = true;
        im.link_iters = d.iterations > 0 ? (uint32_t)d.iterations : 0u;
        if (d.count && !d.links) throw ArgError(FG_ERR_INVALID, "link: links array is null");
        im.off_row = at, at += align256(sizeof(uint32_t) * ((size_t)n + 1));
        im.off_ent = at, at += align256(sizeof(LinkEntry) * (d.count ? 2 * (size_t)d.count : 1));
        break;
      case FG_FORCE_COLLIDE:
        im.collide = true;
        im.coll_iters = d.iterations > 0 ? (uint32_t)d.iterations : 0u;
        im.rmax = 0;
        for (uint32_t i = 0; d.radii && i < n; ++i) im.rmax = std::max(im.rmax, finite_f32(d.radii[i]));
        im.off_radius = at, at += fb;
        break;
      case FG_FORCE_RADIAL:
        im.radial = true;
        im.off_rr = at, at += fb;
        im.off_rw = at, at += fb;
        break;
      case FG_FORCE_POSITION:
        im.pos[d.axis] = true;
        im.off_goal[d.axis] = at, at += fb;
        im.off_weight[d.axis] = at, at += fb;
        break;
      default: break;
    }
  }
  im.tree = im.mb && n >= 2;
  im.grid = im.collide && im.rmax > 0 && im.coll_iters > 0;
  im.charge = at + NODE_SCRATCH * n + (im.tree ? TREE_SCRATCH * n : 0) + (im.grid ? GRID_SCRATCH * n : 0);
  if (!fill) return;
  im.bytes.assign(at > 0 ? at : 256, 0);
  auto F = [&](size_t off) { return (float*)(im.bytes.data() + off); };
  for (uint32_t i = 0; i < n; ++i)
    for (int a = 0; a < j->ndim; ++a) {
      F(im.off_p[a])[i] = finite_f32(j->pos[3 * (size_t)i + a]);
      F(im.off_v[0][a])[i] = j->vel ? finite_f32(j->vel[3 * (size_t)i + a]) : 0.0f;
    }
  if (pinned) {
    for (uint32_t i = 0; i < n; ++i) {
      const uint8_t m = (uint8_t)(j->fixmask[i] & ((1u << j->ndim) - 1u));
      im.bytes[im.off_mask + i] = m;
      for (int a = 0; a < 3; ++a)
        if ((m >> a) & 1u) F(im.off_fix[a])[i] = j->fix ? finite_f32(j->fix[3 * (size_t)i + a]) : 0.0f;
    }
  }
  for (uint32_t k = 0; k < j->nforces; ++k) {
    const fg_force_desc& d = j->forces[k];
    switch (d.type) {
      case FG_FORCE_MANYBODY:
        for (uint32_t i = 0; i < n; ++i) F(im.off_mb)[i] = d.strengths ? finite_f32(d.strengths[i]) : 0.0f;
        break;
      case FG_FORCE_LINK:  // built on the device once the image is resident (link_csr_device)
        break;
      case FG_FORCE_COLLIDE:
        for (uint32_t i = 0; i < n; ++i) F(im.off_radius)[i] = d.radii ? finite_f32(d.radii[i]) : 0.0f;
        im.gc = grid_classes(F(im.off_radius), n);
        break;
      case FG_FORCE_RADIAL:
        for (uint32_t i = 0; i < n; ++i) {
          F(im.off_rr)[i] = d.radii ? finite_f32(d.radii[i]) : 0.0f;
          F(im.off_rw)[i] = d.strengths ? finite_f32(d.strengths[i]) : 0.0f;
        }
        break;
      case FG_FORCE_POSITION:
        for (uint32_t i = 0; i < n; ++i) {
          F(im.off_goal[d.axis])[i] = d.targets ? finite_f32(d.targets[i]) : 0.0f;
          F(im.off_weight[d.axis])[i] = d.strengths ? finite_f32(d.strengths[i]) : 0.0f;
        }
        break;
      default: break;
    }
  }
}

BJob job_record(const fg_job_desc* j, const JobImage& im, uint8_t* base) {
  BJob r{};
  auto P = [&](size_t off) { return (float*)(base + off); };
  for (int a = 0; a < 3; ++a) {
    r.p[a] = P(im.off_p[a]);
    r.v[0][a] = P(im.off_v[0][a]);
    r.v[1][a] = P(im.off_v[1][a]);
    r.fix[a] = im.mask ? P(im.off_fix[a]) : nullptr;
  }
  r.mask = im.mask ? base + im.off_mask : nullptr;
  r.n = j->n;
  r.ndim = (uint32_t)j->ndim;
  r.alpha = j->alpha;
  r.alpha_min = j->alpha_min;
  r.alpha_decay = j->alpha_decay;
  r.alpha_target = j->alpha_target;
  r.decay = (float)(1.0 - j->velocity_decay);
  r.status = BJOB_RUNNING;
  r.max_ticks = j->max_ticks;
  r.seed = j->seed;
  if (im.mb) {
    const fg_force_desc& d = *im.fd[1];
    r.forces |= BF_MANYBODY;
    r.mb_strength = P(im.off_mb);
    r.theta2 = d.theta2;
    r.dmin2 = d.distance_min2;
    r.dmax2 = d.distance_max2;
  }
  if (im.link) {
    r.forces |= BF_LINK;
    r.link_row = (const uint32_t*)(base + im.off_row);
    r.link_ent = (const LinkEntry*)(base + im.off_ent);
    r.link_iters = im.link_iters;
  }
  r.vcur = r.link_iters & 1u;
  if (im.grid) {
    const fg_force_desc& d = *im.fd[2];
    r.forces |= BF_COLLIDE;
    r.radius = P(im.off_radius);
    r.coll_iters = im.coll_iters;
    r.coll = im.gc;
    r.coll_strength = (float)d.strength;
  }
  if (im.fd[3]) {
    const fg_force_desc& d = *im.fd[3];
    r.forces |= BF_CENTER;
    r.center[0] = d.x, r.center[1] = d.y, r.center[2] = d.z;
    r.center_strength = d.strength;
  }
  if (im.radial) {
    const fg_force_desc& d = *im.fd[4];
    r.forces |= BF_RADIAL;
    r.radial_r = P(im.off_rr);
    r.radial_w = P(im.off_rw);
    r.radial_c[0] = finite_f32(d.x), r.radial_c[1] = finite_f32(d.y), r.radial_c[2] = finite_f32(d.z);
  }
  for (int a = 0; a < 3; ++a)
    if (im.pos[a]) {
      r.forces |= BF_POS_X << a;
      r.pos_goal[a] = P(im.off_goal[a]);
      r.pos_weight[a] = P(im.off_weight[a]);
    }
  for (int a = 0; a < 3; ++a) r.cube_box[a] = 0xffffffffu, r.cube_box[3 + a] = 0u;
  return r;
}

// ---- layout ------------------------------------------------------------------------

void upload(hipStream_t s, DevBuf& b, const void* src, size_t bytes) {
  if (b.bytes < bytes || !b.p) b.alloc(bytes);
  if (bytes) FG_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
}

void build_space(hipStream_t s, BSpace& sp, DevBuf& seg_of, DevBuf& seg_job, const std::vector<uint32_t>& starts,
                 const std::vector<uint32_t>& jobs, uint32_t total) {
  sp.total = total;
  sp.segs = (uint32_t)jobs.size();
  upload(s, seg_job, jobs.data(), sizeof(uint32_t) * jobs.size());
  if (seg_of.bytes < sizeof(uint32_t) * total || !seg_of.p) seg_of.alloc(sizeof(uint32_t) * (total ? total : 1));
  sp.seg_job = seg_job.as<uint32_t>();
  sp.seg_of = seg_of.as<uint32_t>();
  if (!total) return;
  DevBuf st;
  st.alloc(sizeof(uint32_t) * starts.size());
  FG_HIP(hipMemcpyAsync(st.p, starts.data(), sizeof(uint32_t) * starts.size(), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(fill_seg_of, dim3(ceil_div(total, 256)), dim3(256), 0, s, total, st.as<uint32_t>(),
                     (uint32_t)starts.size(), sp.seg_of);
  FG_LAUNCH_CHECK();
  FG_HIP(hipStreamSynchronize(s));
}

void rebuild_layout(fg_batch* b) {
  FG_HIP(hipStreamSynchronize(b->stream));
  drop_graph(b);
  std::vector<LayoutPatch> patches;
  for (int gi = 0; gi < 3; ++gi) {
    Group& g = b->groups[gi];
    g.ndim = gi + 1;
    std::vector<uint32_t> ns, nj, ts, tj, gs, gj;
    std::vector<BTile> ntile, ttile;
    std::vector<uint2> walk;
    uint32_t nt = 0, tt = 0, gt = 0, tree_max = 0;
    g.max_link = g.max_coll = 0;
    g.center = g.radial = false;
    g.pos[0] = g.pos[1] = g.pos[2] = false;
    for (JobHost* j : b->jobs) {
      if (j->ndim != g.ndim) continue;
      LayoutPatch lp{};
      lp.slot = j->slot;
      lp.f[0] = j->n;
      // node space
      lp.f[1] = nt;
      lp.f[4] = (uint32_t)ntile.size();
      for (uint32_t f = 0; f < j->n; f += BTILE)
        ntile.push_back(BTile{(uint32_t)nj.size(), nt + f, std::min(BTILE, j->n - f), 0});
      lp.f[5] = (uint32_t)ntile.size() - lp.f[4];
      ns.push_back(nt);
      nj.push_back(j->slot);
      nt += j->n;
      // tree space
      lp.f[2] = tt;
      lp.f[6] = (uint32_t)ttile.size();
      if (j->tree) {
        for (uint32_t f = 0; f < j->n; f += BTILE)
          ttile.push_back(BTile{(uint32_t)tj.size(), tt + f, std::min(BTILE, j->n - f), 0});
        for (uint32_t f = 0; f < j->n; f += WAVE) walk.push_back(make_uint2((uint32_t)tj.size(), tt + f));
        tree_max = std::max(tree_max, j->n);
        ts.push_back(tt);
        tj.push_back(j->slot);
        tt += j->n;
      }
      lp.f[7] = (uint32_t)ttile.size() - lp.f[6];
      // grid space
      lp.f[3] = gt;
      if (j->grid) {
        gs.push_back(gt);
        gj.push_back(j->slot);
        gt += j->n;
        g.max_coll = std::max(g.max_coll, j->rec.coll_iters);
      }
      g.max_link = std::max(g.max_link, j->rec.link_iters);
      g.center |= (j->rec.forces & BF_CENTER) != 0;
      g.radial |= (j->rec.forces & BF_RADIAL) != 0;
      for (int a = 0; a < 3; ++a) g.pos[a] |= (j->rec.forces & (BF_POS_X << a)) != 0;
      patches.push_back(lp);
    }
    build_space(b->stream, g.node, g.node_seg_of, g.node_seg_job, ns, nj, nt);
    build_space(b->stream, g.tree, g.tree_seg_of, g.tree_seg_job, ts, tj, tt);
    build_space(b->stream, g.grid, g.grid_seg_of, g.grid_seg_job, gs, gj, gt);
    upload(b->stream, g.node_tile, ntile.data(), sizeof(BTile) * ntile.size());
    upload(b->stream, g.tree_tile, ttile.data(), sizeof(BTile) * ttile.size());
    upload(b->stream, g.tree_walk, walk.data(), sizeof(uint2) * walk.size());
    g.node.tiles = (uint32_t)ntile.size();
    g.node.tile = g.node_tile.as<BTile>();
    g.tree.tiles = (uint32_t)ttile.size();
    g.tree.tile = g.tree_tile.as<BTile>();
    g.tree.walk_tiles = (uint32_t)walk.size();
    g.tree.max_n = tree_max;
    g.tree.walk = g.tree_walk.as<uint2>();
    // scratch sized before any capture
    if (g.max_link && g.snap.bytes < sizeof(float4) * (size_t)nt) g.snap.alloc(sizeof(float4) * (size_t)nt);
    if (g.tree.total) g.btree.reserve(g.tree.total);
    if (g.grid.total) g.bgrid.reserve(g.grid.total);
    if (g.node.tiles) g.work.reserve(g.node.tiles);
    FG_HIP(hipStreamSynchronize(b->stream));
  }
  if (!patches.empty()) {
    DevBuf pb;
    pb.alloc(sizeof(LayoutPatch) * patches.size());
    FG_HIP(hipMemcpyAsync(pb.p, patches.data(), sizeof(LayoutPatch) * patches.size(), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(patch_layout, dim3(ceil_div(patches.size(), 256)), dim3(256), 0, b->stream, b->table,
                       pb.as<LayoutPatch>(), (uint32_t)patches.size());
    FG_LAUNCH_CHECK();
    FG_HIP(hipStreamSynchronize(b->stream));
  }
  b->dirty = false;
}

// ---- the tick ----------------------------------------------------------------------

void encode_tick(fg_batch* b) {
  batch_advance(b->stream, b->table, b->slots_used, nullptr);
  for (int gi = 2; gi >= 0; --gi) {
    Group& g = b->groups[gi];
    if (!g.node.total) continue;
    for (uint32_t it = 0; it < g.max_link; ++it)
      batch_link(b->stream, g.ndim, g.node, b->table, it, g.snap.as<float4>());
    if (g.tree.total) batch_manybody(b->stream, g.ndim, g.btree, g.tree, b->table, b->tune, b->cus, nullptr);
    for (uint32_t it = 0; it < g.max_coll; ++it)
      batch_collide(b->stream, g.ndim, g.bgrid, g.grid, b->table, it, b->tune.radix_bits);
    if (g.center) batch_center(b->stream, g.ndim, g.work, g.node, b->table);
    if (g.radial) batch_radial(b->stream, g.ndim, g.node, b->table);
    for (int a = 0; a < 3; ++a)
      if (g.pos[a]) batch_position(b->stream, g.node, b->table, a);
    batch_integrate(b->stream, g.ndim, g.node, b->table);
  }
}

void ensure_graph(fg_batch* b) {
  if (b->exec) return;
  FG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  try {
    encode_tick(b);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  FG_HIP(hipStreamEndCapture(b->stream, &b->graph));
  FG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  FG_HIP(hipGraphGetNodes(b->graph, nullptr, &b->graph_nodes));
  b->graph_builds++;
}

void run_ticks(fg_batch* b, uint32_t ticks) {
  if (b->dirty) rebuild_layout(b);
  if (b->jobs.empty() || !ticks) return;
  if (b->use_graphs) {
    ensure_graph(b);
    for (uint32_t t = 0; t < ticks; ++t) FG_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    for (uint32_t t = 0; t < ticks; ++t) encode_tick(b);
  }
  b->ticks += ticks;
}

void read_table(fg_batch* b) {
  FG_HIP(hipStreamSynchronize(b->stream));
  b->host_table.resize(b->slots_used);
  if (b->slots_used)
    FG_HIP(hipMemcpy(b->host_table.data(), b->table, sizeof(BJob) * b->slots_used, hipMemcpyDeviceToHost));
}

JobHost* find_job(fg_batch* b, uint64_t id) {
  for (JobHost* j : b->jobs)
    if (j->id == id) return j;
  return nullptr;
}

uint32_t walk_errors(fg_batch* b, uint64_t* interactions, uint32_t* max_stack = nullptr) {
  uint32_t e = 0, deepest = 0;
  uint64_t inter = 0;
  for (Group& g : b->groups) {
    if (!g.tree.total || !g.btree.counters) continue;
    uint32_t h[16];
    FG_HIP(hipMemcpy(h, g.btree.counters, sizeof(h), hipMemcpyDeviceToHost));
    inter += (uint64_t)h[0] | ((uint64_t)h[1] << 32);
    e += h[4];
    deepest = std::max(deepest, h[8]);
    uint32_t se[4] = {0, 0, 0, 0};
    if (g.btree.sort.errors) {
      FG_HIP(hipMemcpy(se, g.btree.sort.errors, sizeof(se), hipMemcpyDeviceToHost));
      e += se[0];
    }
  }
  if (interactions) *interactions = inter;
  if (max_stack) *max_stack = deepest;
  return e;
}

// ---- autotuner ---------------------------------------------------------------------

std::string shape_key(fg_batch* b) {
  std::ostringstream k;
  for (Group& g : b->groups) {
    if (!g.node.total) continue;
    auto lg = [](uint64_t v) { int r = 0; while ((1ull << (r + 1)) <= v) ++r; return r; };
    k << "d" << g.ndim << ":n" << lg(g.node.total) << ":j" << lg(g.node.segs) << ":t" << lg(g.tree.total + 1)
      << ":g" << lg(g.grid.total + 1) << ";";
  }
  return k.str();
}

bool cache_lookup(const char* path, const std::string& key, BatchTune& t) {
  if (!path) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string k;
    BatchTune c;
    if (in >> k >> c.radix_bits >> c.walk_waves >> c.walk_stack >> c.walk_blocks_per_cu && k == key) {
      t = c;
      return true;
    }
  }
  return false;
}

void cache_store(const char* path, const std::string& key, const BatchTune& t) {
  if (!path) return;
  std::ofstream f(path, std::ios::app);
  f << key << " " << t.radix_bits << " " << t.walk_waves << " " << t.walk_stack << " " << t.walk_blocks_per_cu << "\n";
}

// Every resident job's positions and current velocities plus the job table.
struct Snapshot {
  DevBuf data, table;
  std::vector<std::pair<const float*, float*>> pairs;
  std::vector<uint32_t> sizes;
};

void snapshot_take(fg_batch* b, Snapshot& s) {
  uint64_t floats = 0;
  for (JobHost* j : b->jobs) floats += 6ull * j->n;
  s.data.alloc(sizeof(float) * (floats ? floats : 1));
  s.table.alloc(sizeof(BJob) * (b->slots_used ? b->slots_used : 1));
  float* at = s.data.as<float>();
  s.pairs.clear();
  s.sizes.clear();
  for (JobHost* j : b->jobs) {
    for (int a = 0; a < 3; ++a) {
      s.pairs.push_back({j->rec.p[a], at});
      s.sizes.push_back(j->n);
      at += j->n;
      s.pairs.push_back({j->rec.v[0][a], at});
      s.sizes.push_back(j->n);
      at += j->n;
    }
  }
  for (size_t k = 0; k < s.pairs.size(); ++k)
    hipLaunchKernelGGL(copy_f32, dim3(ceil_div(s.sizes[k], 256)), dim3(256), 0, b->stream, s.sizes[k],
                       s.pairs[k].first, s.pairs[k].second);
  FG_LAUNCH_CHECK();
  if (b->slots_used)
    FG_HIP(hipMemcpyAsync(s.table.p, b->table, sizeof(BJob) * b->slots_used, hipMemcpyDeviceToDevice, b->stream));
  FG_HIP(hipStreamSynchronize(b->stream));
}

void snapshot_restore(fg_batch* b, Snapshot& s) {
  FG_HIP(hipStreamSynchronize(b->stream));
  for (size_t k = 0; k < s.pairs.size(); ++k)
    hipLaunchKernelGGL(copy_f32, dim3(ceil_div(s.sizes[k], 256)), dim3(256), 0, b->stream, s.sizes[k],
                       (const float*)s.pairs[k].second, (float*)s.pairs[k].first);
  FG_LAUNCH_CHECK();
  if (b->slots_used)
    FG_HIP(hipMemcpyAsync(b->table, s.table.p, sizeof(BJob) * b->slots_used, hipMemcpyDeviceToDevice, b->stream));
  FG_HIP(hipStreamSynchronize(b->stream));
}

// Mean ms per tick of `tune` on the current jobs (from the snapshot); < 0 when the
// configuration cannot run or its walks hit their bounds.
double time_tune(fg_batch* b, Snapshot& s, const BatchTune& t) {
  b->tune = t;
  drop_graph(b);
  double ms = -1.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  try {
    snapshot_restore(b, s);
    run_ticks(b, 1);  // warm-up and graph capture
    FG_HIP(hipEventCreate(&e0));
    FG_HIP(hipEventCreate(&e1));
    const uint32_t K = 3;
    FG_HIP(hipEventRecord(e0, b->stream));
    run_ticks(b, K);
    FG_HIP(hipEventRecord(e1, b->stream));
    FG_HIP(hipEventSynchronize(e1));
    float t_ms = 0;
    FG_HIP(hipEventElapsedTime(&t_ms, e0, e1));
    ms = walk_errors(b, nullptr) ? -1.0 : t_ms / K;
  } catch (const ArgError&) {
    ms = -1.0;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return ms;
}

}  // namespace

extern "C" {

int fg_batch_create(fg_batch** out, int device, uint64_t hbm_budget) {
  if (!out) return FG_ERR_INVALID;
  *out = nullptr;
  int k = 0;
  if (hipGetDeviceCount(&k) != hipSuccess || k <= 0) return FG_ERR_NODEVICE;
  if (device < 0 || device >= k) return FG_ERR_INVALID;
  fg_batch* b = new (std::nothrow) fg_batch();
  if (!b) return FG_ERR_NOMEM;
  b->device = device;
  int rc = bguard(b, [&] {
    FG_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    FG_HIP(hipGetDeviceProperties(&prop, device));
    b->cus = prop.multiProcessorCount;
    size_t free_b = 0, total_b = 0;
    FG_HIP(hipMemGetInfo(&free_b, &total_b));
    b->budget = hbm_budget ? hbm_budget : (uint64_t)(0.8 * (double)free_b);
    FG_HIP(hipMalloc(&b->table, sizeof(BJob) * MAX_SLOTS));
    clear_device(b->table, sizeof(BJob) * MAX_SLOTS);
    b->slot_job.assign(MAX_SLOTS, nullptr);
    if (const char* v = getenv("D3FGPU_RADIX_BITS")) b->tune.radix_bits = atoi(v) == 11 ? 11 : 8;
    return FG_OK;
  });
  if (rc != FG_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return FG_OK;
}

int fg_batch_destroy(fg_batch* b) {
  if (!b) return FG_ERR_INVALID;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
  return FG_OK;
}

const char* fg_batch_last_error(const fg_batch* b) { return b ? b->err.c_str() : "null batch"; }

int fg_batch_job_bytes(const fg_job_desc* job, uint64_t* bytes) {
  if (!bytes) return FG_ERR_INVALID;
  try {
    check_job(job);
    JobImage im;
    job_image(job, im, false);
    *bytes = im.charge;
    return FG_OK;
  } catch (const ArgError& e) {
    return e.status;
  } catch (...) {
    return FG_ERR_INVALID;
  }
}

int fg_batch_submit(fg_batch* b, const fg_job_desc* job, uint64_t* id) {
  return bguard(b, [&] {
    if (!id) return bfail(b, FG_ERR_INVALID, "id is null");
    check_job(job);
    JobImage im;
    job_image(job, im, false);
    if (b->charged + im.charge > b->budget)
      return bfail(b, FG_ERR_NOMEM, "the job does not fit the batch's HBM budget now");
    uint32_t slot = MAX_SLOTS;
    for (uint32_t s = 0; s < MAX_SLOTS; ++s)
      if (!b->slot_job[s]) {
        slot = s;
        break;
      }
    if (slot == MAX_SLOTS) return bfail(b, FG_ERR_NOMEM, "the batch holds its maximum number of jobs");
    job_image(job, im, true);
    std::unique_ptr<JobHost> j(new JobHost());
    FG_HIP(hipMalloc(&j->mem, im.bytes.size()));
    try {
      FG_HIP(hipMemcpyAsync(j->mem, im.bytes.data(), im.bytes.size(), hipMemcpyHostToDevice, b->stream));
      if (im.link) {
        const fg_force_desc& d = *im.fd[0];
        LinkUpload up(b->stream, d);
        link_csr_device(b->stream, job->n, d.count, up.links(), up.strengths(), up.distances(), up.bias(),
                        (uint32_t*)((uint8_t*)j->mem + im.off_row), (LinkEntry*)((uint8_t*)j->mem + im.off_ent));
      }
      j->rec = job_record(job, im, (uint8_t*)j->mem);
      FG_HIP(hipMemcpyAsync(b->table + slot, &j->rec, sizeof(BJob), hipMemcpyHostToDevice, b->stream));
      FG_HIP(hipStreamSynchronize(b->stream));
    } catch (...) {
      (void)hipFree(j->mem);
      throw;
    }
    j->id = b->next_id++;
    j->slot = slot;
    j->n = job->n;
    j->ndim = job->ndim;
    j->charge = im.charge;
    j->tree = im.tree;
    j->grid = im.grid;
    b->charged += im.charge;
    b->slot_job[slot] = j.get();
    b->slots_used = std::max(b->slots_used, slot + 1);
    *id = j->id;
    b->jobs.push_back(j.release());
    b->dirty = true;
    return FG_OK;
  });
}

int fg_batch_step(fg_batch* b, uint32_t ticks) {
  return bguard(b, [&] {
    run_ticks(b, ticks);
    return FG_OK;
  });
}

int fg_batch_sync(fg_batch* b) {
  return bguard(b, [&] {
    FG_HIP(hipStreamSynchronize(b->stream));
    return FG_OK;
  });
}

int fg_batch_jobs(fg_batch* b, fg_job_info* out, uint32_t cap, uint32_t* count) {
  return bguard(b, [&] {
    if (!count || (cap && !out)) return bfail(b, FG_ERR_INVALID, "null output");
    read_table(b);
    uint32_t k = 0;
    for (JobHost* j : b->jobs) {
      if (k < cap) {
        const BJob& r = b->host_table[j->slot];
        fg_job_info& o = out[k];
        o.id = j->id;
        o.status = (r.status == BJOB_FINISHING || r.status == BJOB_DONE) ? FG_JOB_DONE : FG_JOB_RUNNING;
        o.ticks = r.ticks;
        o.alpha = r.alpha;
        o.n = j->n;
        o.ndim = j->ndim;
      }
      ++k;
    }
    *count = k;
    return FG_OK;
  });
}

int fg_batch_read(fg_batch* b, uint64_t id, double* pos, double* vel) {
  return bguard(b, [&] {
    JobHost* j = find_job(b, id);
    if (!j) return bfail(b, FG_ERR_INVALID, "no such job");
    FG_HIP(hipStreamSynchronize(b->stream));
    const size_t n = j->n;
    std::vector<float> h(6 * n);
    for (int a = 0; a < 3; ++a) {
      FG_HIP(hipMemcpyAsync(h.data() + a * n, j->rec.p[a], sizeof(float) * n, hipMemcpyDeviceToHost, b->stream));
      FG_HIP(hipMemcpyAsync(h.data() + (3 + a) * n, j->rec.v[0][a], sizeof(float) * n, hipMemcpyDeviceToHost, b->stream));
    }
    FG_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        const bool on = a < j->ndim;
        if (pos) pos[3 * i + a] = on ? (double)h[a * n + i] : 0.0;
        if (vel) vel[3 * i + a] = on ? (double)h[(3 + a) * n + i] : 0.0;
      }
    return FG_OK;
  });
}

int fg_batch_release(fg_batch* b, uint64_t id) {
  return bguard(b, [&] {
    JobHost* j = find_job(b, id);
    if (!j) return bfail(b, FG_ERR_INVALID, "no such job");
    FG_HIP(hipStreamSynchronize(b->stream));
    drop_graph(b);
    BJob zero{};
    FG_HIP(hipMemcpyAsync(b->table + j->slot, &zero, sizeof(BJob), hipMemcpyHostToDevice, b->stream));
    FG_HIP(hipStreamSynchronize(b->stream));
    FG_HIP(hipFree(j->mem));
    b->charged -= j->charge;
    b->slot_job[j->slot] = nullptr;
    while (b->slots_used > 0 && !b->slot_job[b->slots_used - 1]) --b->slots_used;
    b->jobs.erase(std::find(b->jobs.begin(), b->jobs.end(), j));
    delete j;
    b->dirty = true;
    return FG_OK;
  });
}

int fg_batch_set_tune(fg_batch* b, const fg_batch_tune* t) {
  return bguard(b, [&] {
    if (!t) return bfail(b, FG_ERR_INVALID, "tune is null");
    if (t->radix_bits != 8 && t->radix_bits != 11) return bfail(b, FG_ERR_INVALID, "radix_bits must be 8 or 11");
    if (t->walk_waves != 2 && t->walk_waves != 4 && t->walk_waves != 8)
      return bfail(b, FG_ERR_INVALID, "walk_waves must be 2, 4 or 8");
    if (t->walk_stack != 16 && t->walk_stack != 32 && t->walk_stack != 48 && t->walk_stack != 64 &&
        t->walk_stack != 96)
      return bfail(b, FG_ERR_INVALID, "walk_stack must be 16, 32, 48, 64 or 96");
    FG_HIP(hipStreamSynchronize(b->stream));
    b->tune.radix_bits = t->radix_bits;
    b->tune.walk_waves = t->walk_waves;
    b->tune.walk_stack = t->walk_stack;
    b->tune.walk_blocks_per_cu = t->walk_blocks_per_cu > 0 ? t->walk_blocks_per_cu : 0;
    drop_graph(b);
    return FG_OK;
  });
}

int fg_batch_get_tune(fg_batch* b, fg_batch_tune* t) {
  if (!b || !t) return FG_ERR_INVALID;
  t->radix_bits = b->tune.radix_bits;
  t->walk_waves = b->tune.walk_waves;
  t->walk_stack = b->tune.walk_stack;
  t->walk_blocks_per_cu = b->tune.walk_blocks_per_cu;
  return FG_OK;
}

int fg_batch_autotune(fg_batch* b, const char* cache_path, fg_batch_tune* chosen) {
  return bguard(b, [&] {
    if (b->dirty) rebuild_layout(b);
    BatchTune best = b->tune;
    const std::string key = shape_key(b);
    if (!b->jobs.empty() && !cache_lookup(cache_path, key, best)) {
      Snapshot snap;
      snapshot_take(b, snap);
      const uint64_t ticks0 = b->ticks;
      double best_ms = time_tune(b, snap, best);
      // coordinate search: digit width, then workgroup size, then stack depth
      auto consider = [&](BatchTune t) {
        double ms = time_tune(b, snap, t);
        if (ms > 0 && (best_ms < 0 || ms < best_ms)) best_ms = ms, best = t;
      };
      for (int rb : {8, 11})
        if (rb != best.radix_bits) {
          BatchTune t = best;
          t.radix_bits = rb;
          consider(t);
        }
      for (int w : {2, 4, 8})
        if (w != best.walk_waves) {
          BatchTune t = best;
          t.walk_waves = w;
          consider(t);
        }
      for (int k : {32, 48, 64})
        if (k != best.walk_stack) {
          BatchTune t = best;
          t.walk_stack = k;
          consider(t);
        }
      snapshot_restore(b, snap);
      b->ticks = ticks0;
      if (best_ms < 0) best = BatchTune();
      cache_store(cache_path, key, best);
    }
    b->tune = best;
    drop_graph(b);
    if (chosen) {
      chosen->radix_bits = best.radix_bits;
      chosen->walk_waves = best.walk_waves;
      chosen->walk_stack = best.walk_stack;
      chosen->walk_blocks_per_cu = best.walk_blocks_per_cu;
    }
    return FG_OK;
  });
}

int fg_batch_use_graphs(fg_batch* b, int32_t on) {
  return bguard(b, [&] {
    FG_HIP(hipStreamSynchronize(b->stream));
    b->use_graphs = on != 0;
    drop_graph(b);
    return FG_OK;
  });
}

int fg_batch_get_stats(fg_batch* b, fg_batch_stats* out) {
  return bguard(b, [&] {
    if (!out) return bfail(b, FG_ERR_INVALID, "null output");
    read_table(b);
    fg_batch_stats s{};
    for (JobHost* j : b->jobs) {
      const uint32_t st = b->host_table[j->slot].status;
      if (st == BJOB_RUNNING) s.jobs_running++;
      else s.jobs_done++;
      s.nodes += j->n;
    }
    s.bytes = b->charged;
    s.budget = b->budget;
    s.ticks = b->ticks;
    s.graph_builds = b->graph_builds;
    s.walk_errors = walk_errors(b, &s.interactions, &s.walk_max_stack);
    s.launches_per_tick = (uint32_t)b->graph_nodes;
    *out = s;
    return FG_OK;
  });
}

}  // extern "C"
