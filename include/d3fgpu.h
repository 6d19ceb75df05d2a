/*
 * d3fgpu.h -- C ABI of the MI355X physics backend for forceSimulationGPU.
 *
 * One context = one simulation resident on one GPU (positions, velocities, pins, the
 * force passes and their device buffers).  The JavaScript host (an N-API addon) and the
 * Python bindings call only these functions; no C++ or torch type crosses this line.
 *
 * The entry points replace the engine object of the reference, src/gpu/engine.js
 * (createEngine, engine.js:73-428), and the plugin contract behind it: a force's
 * gpu() descriptor {type, version, ...} (docs/differences.md:71-78) becomes an
 * fg_force_desc, and passes[type].create/update/encode (passes/index.js:8-14) happen
 * inside fg_set_forces and fg_enqueue.
 *
 * Conventions
 *   - Every function returns FG_OK (0) or a negative FG_ERR_* status; the message of the
 *     last failure on a context is fg_last_error(ctx).
 *   - Host arrays are owned by the caller and are copied before the call returns
 *     (as queue.writeBuffer copies, passes/util.js:13-20).
 *   - Node vectors on the host are interleaved doubles, 3 per node (x, y, z), whatever
 *     the number of dimensions; unused axes are ignored on input and 0 on output.
 *   - Calls on one context come from one thread.  Contexts are independent.
 */
#ifndef D3FGPU_H
#define D3FGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_ABI_VERSION 1

/* status codes */
#define FG_OK 0
#define FG_ERR_INVALID (-1)     /* bad argument (null pointer, out of range, wrong order) */
#define FG_ERR_HIP (-2)         /* a HIP runtime call failed (device lost, launch failure) */
#define FG_ERR_NOMEM (-3)       /* device or pinned host memory exhausted */
#define FG_ERR_UNSUPPORTED (-4) /* force type or size this backend does not run */
#define FG_ERR_STATE (-5)       /* call not valid in the context's current state */
#define FG_ERR_NODEVICE (-6)    /* no usable GPU */

/* Force types: the `type` strings of the reference descriptors (passes/index.js:8-14). */
enum fg_force_type {
  FG_FORCE_CENTER = 1,   /* src/center.js:49-51   -> passes/center.js   */
  FG_FORCE_COLLIDE = 2,  /* src/collide.js:134-136 -> passes/collide.js, collideGrid.js */
  FG_FORCE_LINK = 3,     /* src/link.js:127-129   -> passes/link.js     */
  FG_FORCE_MANYBODY = 4, /* src/manyBody.js:142-144 -> passes/manyBody.js, barnesHut.js */
  FG_FORCE_POSITION = 5, /* src/x.js:43-45 (y.js, z.js) -> passes/position.js */
  FG_FORCE_RADIAL = 6    /* src/radial.js:69-71   -> passes/radial.js   */
};

/*
 * One force, as its gpu() descriptor describes it.  Array members are read only when
 * (id, version) differs from what the context last saw for that id; scalars are
 * refreshed on every fg_set_forces call (engine.js:357 refreshes them every step).
 */
typedef struct fg_force_desc {
  int32_t type;          /* enum fg_force_type */
  int32_t iterations;    /* link, collide: passes per tick (>= 0) */
  uint64_t id;           /* identity of the force object: a retained id keeps its pass */
  uint64_t version;      /* bumped by the force whenever its per-node/per-link arrays change */
  double x, y, z;        /* center, radial: the centre */
  double strength;       /* center, collide: the scalar strength */
  double theta2;         /* manyBody: theta^2 (<= 1e-6 selects the exact all-pairs sum) */
  double distance_min2;  /* manyBody */
  double distance_max2;  /* manyBody (may be +inf) */
  int32_t axis;          /* position: 0 = forceX, 1 = forceY, 2 = forceZ */
  uint32_t count;        /* link: number of links m */
  const uint32_t* links; /* link: 2*m node indices, (source, target) per link */
  const double* strengths; /* manyBody[n], link[m], radial[n], position[n] */
  const double* radii;     /* collide[n]; radial: radiuses[n] */
  const double* distances; /* link[m] */
  const double* bias;      /* link[m] */
  const double* targets;   /* position[n] */
} fg_force_desc;

typedef struct fg_ctx fg_ctx;

/* Device-resident state for rendering without readback (engine.js:418-420). */
typedef struct fg_buffer_info {
  void* positions;   /* device pointer: float4 per node, xyz used (AoS mirror) */
  void* velocities;  /* device pointer: float4 per node, xyz used */
  uint32_t count;
  uint32_t stride;   /* 16 */
  uint64_t epoch;    /* changes when the pointers are replaced (set_nodes, destroy) */
} fg_buffer_info;

/* Counters of the last step that ran a many-body tree (debug and roofline). */
typedef struct fg_stats {
  uint64_t steps;            /* steps run since set_nodes */
  uint64_t interactions;     /* body-source evaluations in the last walk (0 if not counted) */
  uint64_t opened;           /* wave-level node openings in the last walk */
  uint64_t lane_visits;      /* sum over openings of the lanes (points) that opened the node */
  uint32_t sort_bits;        /* bits per radix digit used by the tree sort */
  uint32_t walk_errors;      /* walks that hit their stack or iteration bound (must be 0) */
  uint32_t walk_max_stack;   /* deepest stack any wave of the last walk used */
} fg_stats;

/* ---- library ------------------------------------------------------------------- */

/* FG_ABI_VERSION of the loaded library. */
int fg_abi_version(void);

/* Number of usable GPUs (0 when none). Probe for isWebGPUAvailable (device.js:6-18). */
int fg_device_count(int* count);

/* ---- context lifetime (createEngine / destroy, engine.js:73, 422-427) ------------ */

int fg_create(fg_ctx** out, int device);
int fg_destroy(fg_ctx* ctx);
const char* fg_last_error(const fg_ctx* ctx);

/* Whether `n` nodes in `ndim` dimensions fit this device (engine.fits, engine.js:334-338).
 * Test hooks, read by fg_create from the environment: D3FGPU_CAPACITY=N makes fg_fits refuse
 * more than N nodes (the reference fakes device limits for its oversize test,
 * test/gpu-simulation-test.js:384-406); D3FGPU_FAIL=enqueue:K or wait:K makes the K-th
 * fg_enqueue, or the K-th fg_sync / fg_readback_wait, return FG_ERR_HIP without touching
 * the device, so that a host's failure handling can be exercised. */
int fg_fits(fg_ctx* ctx, uint32_t n, int32_t ndim, int32_t* fits);

/* ---- nodes (setNodes / uploadAll / uploadPositions / uploadPinned, engine.js:115-169, 267-310) */

/*
 * Replace the node set.  pos: 3n (required), vel: 3n (NULL = zero), fix: 3n pinned
 * coordinates (NULL = none), fixmask: n bytes, bit a = axis a pinned (NULL = none).
 * Forces must be set again afterwards.  n = 0 frees the buffers.
 */
int fg_set_nodes(fg_ctx* ctx, uint32_t n, int32_t ndim, const double* pos, const double* vel,
                 const double* fix, const uint8_t* fixmask);
/* Overwrite positions [first, first + count) (user edits between steps). */
int fg_update_positions(fg_ctx* ctx, uint32_t first, uint32_t count, const double* pos);
/* Overwrite velocities [first, first + count). */
int fg_update_velocities(fg_ctx* ctx, uint32_t first, uint32_t count, const double* vel);
/* Overwrite pins [first, first + count): fix 3*count, fixmask count. */
int fg_update_pins(fg_ctx* ctx, uint32_t first, uint32_t count, const double* fix,
                   const uint8_t* fixmask);

/* ---- forces (setForces + pass.update, engine.js:314-330, 357) -------------------- */

/* The forces, in application order.  Passes of ids no longer listed are released. */
int fg_set_forces(fg_ctx* ctx, uint32_t k, const fg_force_desc* descs);

/* Seed of the coincident-node jiggle (engine.setSeed, engine.js:341-343). */
int fg_set_seed(fg_ctx* ctx, uint32_t seed);

/* ---- stepping (enqueue, engine.js:352-391) -------------------------------------- */

/*
 * Queue `steps` ticks: tick s applies every force with alpha = alphas[s], then the
 * integrator with velocity multiplier `decay` (1 - velocityDecay).  Returns at once.
 */
int fg_enqueue(fg_ctx* ctx, uint32_t steps, const double* alphas, double decay);

/* Apply every force once with `alpha` and no integrator (forces-only probe: the
 * velocity increment of one tick is read back with fg_readback). */
int fg_apply_forces(fg_ctx* ctx, double alpha);

/* True while 3 or more queued submissions are unfinished (busy, engine.js:394-396). */
int fg_busy(fg_ctx* ctx, int32_t* busy);

/* Block until every queued step has finished (idle, engine.js:402-404). */
int fg_sync(fg_ctx* ctx);

/* ---- readback (readback / land, engine.js:198-239) ------------------------------ */

/* Wait for queued work and copy the state: pos, vel 3n doubles each (either may be NULL). */
int fg_readback(fg_ctx* ctx, double* pos, double* vel);

/*
 * Asynchronous readback: start copies the state into pinned staging and returns a
 * ticket; wait blocks (callable from a worker thread) until those copies are done;
 * fetch converts them into the caller's arrays.  One readback is in flight at a time.
 * fetch fails with FG_ERR_STATE if the nodes were replaced since start (epoch).
 */
int fg_readback_start(fg_ctx* ctx, uint64_t* ticket);
int fg_readback_wait(fg_ctx* ctx, uint64_t ticket);
int fg_readback_fetch(fg_ctx* ctx, uint64_t ticket, double* pos, double* vel);

/* Device pointers of the state (gpuBuffers, engine.js:418-420).  Keeps an AoS float4
 * mirror up to date after every queued step from the first call on. */
int fg_buffers(fg_ctx* ctx, fg_buffer_info* out);

/* Test and debugging aid, with no counterpart in the reference: copies `bytes` from device
 * memory of this context's GPU (for instance what fg_buffers points to) into host memory,
 * after the queued work.  A renderer would read those buffers with its own HIP calls. */
int fg_debug_copy_from_device(fg_ctx* ctx, const void* device_src, uint64_t bytes, void* host_dst);

int fg_get_stats(fg_ctx* ctx, fg_stats* out);

/* ---- timing (HIP events on the context's stream) -------------------------------- */

/* Phases the timer reports, each the sum of its launches' event-to-event times. */
enum fg_phase {
  FG_PHASE_TICK = 0,       /* a whole tick: every pass and the integrator */
  FG_PHASE_TREE_BUILD = 1, /* bbox, keys, leaves, Karras build, reduction (sort excluded) */
  FG_PHASE_TREE_SORT = 2,  /* the radix sort of the Morton codes */
  FG_PHASE_WALK = 3,       /* the tree walk kernel */
  FG_PHASE_PAIRS = 4,      /* the exact all-pairs many-body kernel */
  FG_PHASE_LINK = 5,
  FG_PHASE_COLLIDE = 6,
  FG_PHASE_CENTER = 7,     /* empty up to 1,024 nodes when center is the last force: see INTEGRATE */
  FG_PHASE_RADIAL = 8,
  FG_PHASE_POSITION = 9,
  FG_PHASE_INTEGRATE = 10, /* the integrator; up to 1,024 nodes with center the last force, the
                              same launch finds and applies the center shift */
  FG_PHASE_COUNT = 16
};

typedef struct fg_timing {
  double ms[FG_PHASE_COUNT];        /* summed milliseconds since timing was enabled */
  uint64_t launches[FG_PHASE_COUNT];
} fg_timing;

/* Start (on != 0, clearing the sums) or stop recording events around every phase. */
int fg_timing_enable(fg_ctx* ctx, int32_t on);
/* Wait for queued work and return the sums. */
int fg_get_timing(fg_ctx* ctx, fg_timing* out);

/* ---- kernels on their own (the reference tests drive sort.js the same way,
 *      test/gpu-sort-test.js:9-38) ----------------------------------------------- */

/*
 * Stable LSD radix sort of (key, value) u32 pairs on bits [first_bit, first_bit + bits)
 * of the key, ascending, ties in input order (createSort, sort.js:223-273).
 * Host arrays of n entries; runs on ctx's device.
 */
int fg_radix_sort_pairs(fg_ctx* ctx, uint32_t n, const uint32_t* keys, const uint32_t* values,
                        uint32_t first_bit, uint32_t bits, uint32_t* keys_out,
                        uint32_t* values_out);

/*
 * fg_radix_sort_pairs for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same sort and result, but keys, values, keys_out and values_out are DEVICE
 * pointers of n u32 each (keys_out may equal keys and values_out values: in place), and the
 * call allocates nothing: its work lives in a fixed array of about 14 MB that the library
 * holds on each GPU it runs on until the process ends.  n is at most 1048576 (FG_ERR_INVALID
 * beyond).  Queued on ctx's stream; returns once the sort is done.  Every call on one GPU,
 * from any context, shares that array: calls that overlap on one GPU (other threads, other
 * contexts' streams) race and must be serialised by the caller.
 */
int fg_radix_sort_pairs_noalloc(fg_ctx* ctx, uint32_t n, const uint32_t* keys,
                                const uint32_t* values, uint32_t first_bit, uint32_t bits,
                                uint32_t* keys_out, uint32_t* values_out);

/*
 * The many-body tree of the last step that built one, for parity checks against the
 * reference algorithm (barnesHut.js:21-47).  Any pointer may be NULL.
 *   cube: origin x, y, z and size (4 floats); shift: LEVELS - log2(size)
 *   codes: 2n u32, (high, low) Morton halves per node (barnesHut.js:352-366)
 *   order: n u32, node index at each sorted position
 *   leaf_depth: n u32, d3 depth of each sorted point's leaf cell
 *   left, right: n-1 u32 children of each binary node; bit 31 set = sorted point
 *   depth: n-1 u32, d3 depth of the cells each binary node stands for
 *   kids: 8(n-1) u32, the d3 children of each binary node (0xffffffff = none)
 *   com: 3(n-1) doubles, value: n-1 doubles (centre of mass, strength at `depth`)
 */
int fg_debug_tree(fg_ctx* ctx, float* cube, int32_t* shift, uint32_t* codes, uint32_t* order,
                  uint32_t* leaf_depth, uint32_t* left, uint32_t* right, uint32_t* depth,
                  uint32_t* kids, double* com, double* value);


/* ---- batches: many independent layouts advanced by one launch sequence ------------
 *
 * No counterpart in the reference (it runs one simulation per engine).  A batch holds
 * jobs of any size and dimensionality resident on one GPU; each tick of the batch ticks
 * every running job once, with one sequence of launches per dimensionality (segmented
 * sort and tree, one persistent walk over a queue of 64-point tiles of all jobs), the
 * tick captured in a hipGraph.  Each job keeps d3's own schedule (alpha, alphaMin,
 * alphaDecay, alphaTarget, velocityDecay; simulation.js:54-73), advanced on the device,
 * and finishes after the tick whose alpha falls below alphaMin.  A job's result does not
 * depend on the other jobs: alone or in any mix it advances bit for bit the same.
 *
 * A job's forces are applied in the order link, manyBody, collide, center, radial, x, y,
 * z (each at most once); descriptors in another order are refused.  Collide always runs
 * on the grid, manyBody always on the tree.
 */
typedef struct fg_batch fg_batch;

typedef struct fg_job_desc {
  uint32_t n;                  /* nodes (>= 1) */
  int32_t ndim;                /* 1, 2 or 3 */
  const double* pos;           /* 3n, required */
  const double* vel;           /* 3n or NULL (zero) */
  const double* fix;           /* 3n or NULL */
  const uint8_t* fixmask;      /* n or NULL: bit a = axis a pinned */
  uint32_t nforces;
  const fg_force_desc* forces;
  double alpha;                /* starting alpha (1) */
  double alpha_min;            /* 0.001 */
  double alpha_decay;          /* 1 - 0.001^(1/300) */
  double alpha_target;         /* 0 */
  double velocity_decay;       /* d3's velocityDecay (0.4): velocities are multiplied by 1 - it */
  uint32_t seed;               /* jiggle seed */
  uint32_t max_ticks;          /* stop after this many ticks as well (0: no limit) */
} fg_job_desc;

/* Job states */
#define FG_JOB_RUNNING 1
#define FG_JOB_DONE 3

typedef struct fg_job_info {
  uint64_t id;
  uint32_t status;   /* FG_JOB_RUNNING or FG_JOB_DONE */
  uint32_t ticks;    /* ticks run */
  double alpha;      /* alpha of the last tick run */
  uint32_t n;
  int32_t ndim;
} fg_job_info;

/* Tunables of the batched tick (fg_batch_autotune chooses them). */
typedef struct fg_batch_tune {
  int32_t radix_bits;          /* 8 or 11 bits per radix digit */
  int32_t walk_waves;          /* waves per walk workgroup: 2, 4 or 8 */
  int32_t walk_stack;          /* walk stack entries per wave kept in LDS: 16, 32, 48, 64 or 96
                                  (deeper entries spill to global memory, up to 192) */
  int32_t walk_blocks_per_cu;  /* persistent walk workgroups per CU, 0 = as many as fit */
} fg_batch_tune;

typedef struct fg_batch_stats {
  uint32_t jobs_running, jobs_done;
  uint64_t nodes;              /* resident nodes */
  uint64_t bytes;              /* device bytes held for resident jobs and the tick's scratch */
  uint64_t budget;             /* the HBM budget */
  uint64_t ticks;              /* batched ticks run */
  uint64_t graph_builds;       /* tick graphs captured (one per layout or tune change) */
  uint64_t interactions;       /* body-source evaluations of the last tick's walks */
  uint32_t walk_errors;        /* walks that hit their stack or iteration bound (must be 0) */
  uint32_t launches_per_tick;  /* kernel launches in one tick */
  uint32_t walk_max_stack;     /* deepest stack any wave of the last tick's walks used */
} fg_batch_stats;

/* hbm_budget: device bytes the batch may hold (0 = 80% of the free memory). */
int fg_batch_create(fg_batch** out, int device, uint64_t hbm_budget);
int fg_batch_destroy(fg_batch* b);
const char* fg_batch_last_error(const fg_batch* b);
/* Device bytes a job costs while resident (its state and arrays plus its share of the
 * tick's scratch); what admission charges against the budget. */
int fg_batch_job_bytes(const fg_job_desc* job, uint64_t* bytes);
/* Make a job resident (copies every array).  FG_ERR_NOMEM: the budget has no room now. */
int fg_batch_submit(fg_batch* b, const fg_job_desc* job, uint64_t* id);
/* Queue `ticks` ticks of every running job; returns at once. */
int fg_batch_step(fg_batch* b, uint32_t ticks);
int fg_batch_sync(fg_batch* b);
/* Waits for queued ticks; the resident jobs in submission order. */
int fg_batch_jobs(fg_batch* b, fg_job_info* out, uint32_t cap, uint32_t* count);
/* Waits for queued ticks; a job's state, 3n doubles each (either may be NULL). */
int fg_batch_read(fg_batch* b, uint64_t id, double* pos, double* vel);
/* Removes a job (running or done) and frees its memory. */
int fg_batch_release(fg_batch* b, uint64_t id);
int fg_batch_set_tune(fg_batch* b, const fg_batch_tune* t);
int fg_batch_get_tune(fg_batch* b, fg_batch_tune* t);
/* Times candidate tunables on the resident jobs (their state is saved and restored, so
 * results are unchanged), keeps the fastest that runs without walk errors and records it
 * in the text file cache_path under the batch's shape (NULL: no file).  A shape found in
 * the file is taken from it without timing. */
int fg_batch_autotune(fg_batch* b, const char* cache_path, fg_batch_tune* chosen);
/* Capture ticks in a hipGraph (on != 0, the default) or launch them one by one. */
int fg_batch_use_graphs(fg_batch* b, int32_t on);
int fg_batch_get_stats(fg_batch* b, fg_batch_stats* out);

/* ---- serving: a request loop over a batch (continuous batching) -----------------
 *
 * No counterpart in the reference.  A server owns a batch and a worker thread.  Submitted
 * jobs are copied and queued; between rounds of ticks the worker admits queued jobs in
 * arrival order while they fit the HBM budget (a job that cannot fit even an empty batch
 * fails), so jobs arriving while others run join the running batch as memory frees up.
 * After every round it reads back and releases the jobs that finished (alpha <
 * alphaMin), which wakes their waiters.  Latency is finish time minus submit time.
 */
typedef struct fg_server fg_server;

typedef struct fg_serve_config {
  uint64_t hbm_budget;       /* device bytes for resident jobs and scratch (0: 80% of free) */
  uint32_t ticks_per_round;  /* ticks between admission and retirement passes (0: 10) */
  int32_t autotune;          /* nonzero: tune each new batch shape (fg_batch_autotune) */
  const char* tune_cache;    /* the autotuner's cache file (NULL: none) */
} fg_serve_config;

#define FG_SERVE_QUEUED 0
#define FG_SERVE_RUNNING 1
#define FG_SERVE_FINISHED 2
#define FG_SERVE_FAILED 3

typedef struct fg_serve_result {
  uint64_t id;
  uint32_t status;           /* FG_SERVE_* */
  uint32_t n;
  int32_t ndim;
  uint32_t ticks;            /* ticks the job ran */
  double submitted_s, admitted_s, finished_s;  /* on the server's clock (seconds) */
} fg_serve_result;

typedef struct fg_serve_stats {
  uint64_t submitted, admitted, finished, failed;
  uint32_t queued, running, peak_running;
  double p50_latency_s, p99_latency_s, mean_latency_s, max_latency_s;
  double layouts_per_s;      /* finished jobs / (last finish - first submission) */
  uint64_t rounds, ticks;
  uint64_t peak_bytes, budget;
} fg_serve_stats;

int fg_serve_create(fg_server** out, int device, const fg_serve_config* cfg);
/* Stops the loop; jobs not yet finished are dropped. */
int fg_serve_destroy(fg_server* s);
const char* fg_serve_last_error(const fg_server* s);
/* Copies the job and queues it; returns at once with its id. */
int fg_serve_submit(fg_server* s, const fg_job_desc* job, uint64_t* id);
/* Blocks until the job finished (or timeout_s passed: FG_ERR_STATE), then copies its
 * final state (3n doubles each, either may be NULL) and forgets it. */
int fg_serve_wait(fg_server* s, uint64_t id, double timeout_s, double* pos, double* vel,
                  fg_serve_result* result);
/* Blocks until no job is queued or running (or timeout_s passed: FG_ERR_STATE). */
int fg_serve_drain(fg_server* s, double timeout_s);
int fg_serve_get_stats(fg_server* s, fg_serve_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* D3FGPU_H */
