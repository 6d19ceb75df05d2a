// addon.cc -- N-API binding of the C ABI (include/d3fgpu.h) for the Node.js host.
//
// Most JS calls map to one fg_* call.  The exception is the node mirror: the JS engine
// (js/engine.js) packs the simulation's node objects into one Float64Array of STATE doubles
// per node and hands it over whole; this file keeps, per engine, the shadow of what the
// device was last told (positions and pins), uploads only the runs of nodes that differ
// from it (load / push), and merges a readback into the packed state without overwriting
// nodes the caller moved since the last exchange (land).  Waiting happens off the JS
// thread: readbackAsync and idleAsync run fg_readback_wait / fg_sync in napi_async_work
// and settle a Promise.
#include <node_api.h>
#include <stdint.h>
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/d3fgpu.h"

namespace {

#define NAPI_CALL(env, call)                                                   \
  do {                                                                         \
    napi_status s_ = (call);                                                   \
    if (s_ != napi_ok) {                                                       \
      napi_throw_error((env), nullptr, "d3fgpu addon: N-API call failed: " #call); \
      return nullptr;                                                          \
    }                                                                          \
  } while (0)

napi_value throw_status(napi_env env, fg_ctx* c, int rc, const char* what) {
  std::string msg = std::string("d3fgpu ") + what + " failed (" + std::to_string(rc) + ")";
  if (c) msg += ": " + std::string(fg_last_error(c));
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
  return true;
}

// The packed node state: x, y, z, vx, vy, vz, fx, fy, fz per node (fx..fz NaN: not pinned).
constexpr size_t STATE = 9;
// The shadow keeps x, y, z, fx, fy, fz per node: what the device holds (or will once the
// uploads land), compared bit for bit.
constexpr size_t SHADOW = 6;

// The JS handle owns a box, so that destroy() can free the context before the handle is
// collected and the finalizer then finds nothing to free.
struct Box {
  fg_ctx* c;
  std::vector<double> shadow;
  uint32_t n = 0;
  int nd = 2;
};

Box* box_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((Box*)p)->c) {
    napi_throw_error(env, nullptr, "d3fgpu: not a live engine handle");
    return nullptr;
  }
  return (Box*)p;
}

fg_ctx* ctx_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((Box*)p)->c) {
    napi_throw_error(env, nullptr, "d3fgpu: not a live engine handle");
    return nullptr;
  }
  return ((Box*)p)->c;
}

double num(napi_env env, napi_value v, double dflt = 0.0) {
  double d = dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_number) napi_get_value_double(env, v, &d);
  return d;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_undefined || t == napi_null;
}

// Data pointer and element count of a typed array of the given type (nullptr if absent).
template <typename T>
T* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* count) {
  *count = 0;
  if (is_nullish(env, v)) return nullptr;
  bool ok = false;
  napi_is_typedarray(env, v, &ok);
  if (!ok) return nullptr;
  napi_typedarray_type type;
  size_t length;
  void* data;
  napi_value buf;
  size_t off;
  napi_get_typedarray_info(env, v, &type, &length, &data, &buf, &off);
  if (type != want) return nullptr;
  *count = length;
  return (T*)data;
}

napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok) napi_get_undefined(env, &v);
  return v;
}

void finalize_ctx(napi_env, void* data, void*) {
  Box* b = (Box*)data;
  if (b && b->c) fg_destroy(b->c);
  delete b;
}

// ---- calls --------------------------------------------------------------------------

napi_value DeviceCount(napi_env env, napi_callback_info) {
  int k = 0;
  fg_device_count(&k);
  napi_value r;
  NAPI_CALL(env, napi_create_int32(env, k, &r));
  return r;
}

napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_ctx* c = nullptr;
  int rc = fg_create(&c, (int)num(env, argv[0], 0));
  if (rc != FG_OK) return throw_status(env, nullptr, rc, "create");
  napi_value r;
  Box* b = new Box{c};
  NAPI_CALL(env, napi_create_external(env, b, finalize_ctx, nullptr, &r));
  return r;
}

// destroy: frees the context now; the handle's finalizer then has nothing to do.
napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p) return nullptr;
  Box* b = (Box*)p;
  if (b->c) fg_destroy(b->c);
  b->c = nullptr;
  return nullptr;
}

napi_value Fits(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int32_t f = 0;
  int rc = fg_fits(c, (uint32_t)num(env, argv[1]), (int32_t)num(env, argv[2], 2), &f);
  if (rc != FG_OK) return throw_status(env, c, rc, "fits");
  napi_value r;
  NAPI_CALL(env, napi_get_boolean(env, f != 0, &r));
  return r;
}

napi_value SetNodes(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  get_args(env, info, 7, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  uint32_t n = (uint32_t)num(env, argv[1]);
  size_t k;
  double* pos = typed<double>(env, argv[3], napi_float64_array, &k);
  double* vel = typed<double>(env, argv[4], napi_float64_array, &k);
  double* fix = typed<double>(env, argv[5], napi_float64_array, &k);
  uint8_t* mask = typed<uint8_t>(env, argv[6], napi_uint8_array, &k);
  int rc = fg_set_nodes(c, n, (int32_t)num(env, argv[2], 2), pos, vel, fix, mask);
  if (rc != FG_OK) return throw_status(env, c, rc, "setNodes");
  return nullptr;
}

napi_value UpdatePositions(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  size_t k;
  double* pos = typed<double>(env, argv[2], napi_float64_array, &k);
  int rc = fg_update_positions(c, (uint32_t)num(env, argv[1]), (uint32_t)(k / 3), pos);
  if (rc != FG_OK) return throw_status(env, c, rc, "updatePositions");
  return nullptr;
}

napi_value UpdatePins(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  size_t k, km;
  double* fix = typed<double>(env, argv[2], napi_float64_array, &k);
  uint8_t* mask = typed<uint8_t>(env, argv[3], napi_uint8_array, &km);
  int rc = fg_update_pins(c, (uint32_t)num(env, argv[1]), (uint32_t)km, fix, mask);
  if (rc != FG_OK) return throw_status(env, c, rc, "updatePins");
  return nullptr;
}

// ---- the node mirror -------------------------------------------------------------------

bool same_bits(const double* a, const double* b, int k) { return memcmp(a, b, sizeof(double) * (size_t)k) == 0; }

// load(handle, state, nDim): every node goes up (a new node set, or after the host moved the
// nodes itself); the shadow becomes the state.
napi_value Load(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  Box* b = box_of(env, argv[0]);
  if (!b) return nullptr;
  size_t k;
  const double* st = typed<double>(env, argv[1], napi_float64_array, &k);
  const int nd = (int)num(env, argv[2], 2);
  const uint32_t n = (uint32_t)(k / STATE);
  std::vector<double> pos(3 * (size_t)n, 0.0), vel(3 * (size_t)n, 0.0), fix(3 * (size_t)n, 0.0);
  std::vector<uint8_t> mask(n, 0);
  b->shadow.assign(SHADOW * (size_t)n, 0.0);
  for (uint32_t i = 0; i < n; ++i) {
    const double* s = st + STATE * i;
    double* sh = b->shadow.data() + SHADOW * i;
    for (int a = 0; a < 3; ++a) {
      const bool used = a < nd;
      pos[3 * i + a] = used ? s[a] : 0.0;
      vel[3 * i + a] = used ? s[3 + a] : 0.0;
      const bool pinned = used && s[6 + a] == s[6 + a];
      fix[3 * i + a] = pinned ? s[6 + a] : 0.0;
      if (pinned) mask[i] |= (uint8_t)(1u << a);
      sh[a] = pos[3 * i + a];
      sh[3 + a] = used ? s[6 + a] : NAN;
    }
  }
  b->n = n;
  b->nd = nd;
  int rc = fg_set_nodes(b->c, n, nd, pos.data(), vel.data(), fix.data(), mask.data());
  if (rc != FG_OK) return throw_status(env, b->c, rc, "load");
  return nullptr;
}

// push(handle, state) -> nodes uploaded: positions and pins that differ from the shadow go
// up, one fg_update_positions / fg_update_pins call per run of consecutive changed nodes.
napi_value Push(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  Box* b = box_of(env, argv[0]);
  if (!b) return nullptr;
  size_t k;
  const double* st = typed<double>(env, argv[1], napi_float64_array, &k);
  if (k / STATE != b->n) {
    napi_throw_error(env, nullptr, "d3fgpu push: the state has another node count than the last load");
    return nullptr;
  }
  const int nd = b->nd;
  uint32_t moved = 0;
  std::vector<double> run_pos, run_fix;
  std::vector<uint8_t> run_mask;
  uint32_t pos_first = 0, pin_first = 0;
  bool in_pos = false, in_pin = false;
  auto flush_pos = [&](uint32_t end) {
    if (!in_pos) return 0;
    in_pos = false;
    moved += end - pos_first;
    return fg_update_positions(b->c, pos_first, end - pos_first, run_pos.data());
  };
  auto flush_pin = [&](uint32_t end) {
    if (!in_pin) return 0;
    in_pin = false;
    return fg_update_pins(b->c, pin_first, end - pin_first, run_fix.data(), run_mask.data());
  };
  int rc = FG_OK;
  for (uint32_t i = 0; i <= b->n && rc == FG_OK; ++i) {
    bool dpos = false, dpin = false;
    double p[3] = {0, 0, 0}, f[3] = {NAN, NAN, NAN};
    if (i < b->n) {
      const double* s = st + STATE * i;
      for (int a = 0; a < nd; ++a) p[a] = s[a], f[a] = s[6 + a];
      double* sh = b->shadow.data() + SHADOW * i;
      dpos = !same_bits(p, sh, 3);
      dpin = !same_bits(f, sh + 3, 3);
      if (dpos) memcpy(sh, p, sizeof(p));
      if (dpin) memcpy(sh + 3, f, sizeof(f));
    }
    if (dpos) {
      if (!in_pos) in_pos = true, pos_first = i, run_pos.clear();
      run_pos.insert(run_pos.end(), p, p + 3);
    } else {
      rc = flush_pos(i);
    }
    if (rc != FG_OK) break;
    if (dpin) {
      if (!in_pin) in_pin = true, pin_first = i, run_fix.clear(), run_mask.clear();
      uint8_t m = 0;
      for (int a = 0; a < 3; ++a) {
        const bool pinned = f[a] == f[a];
        run_fix.push_back(pinned ? f[a] : 0.0);
        if (pinned) m |= (uint8_t)(1u << a);
      }
      run_mask.push_back(m);
    } else {
      rc = flush_pin(i);
    }
  }
  if (rc != FG_OK) return throw_status(env, b->c, rc, "push");
  napi_value r;
  NAPI_CALL(env, napi_create_uint32(env, moved, &r));
  return r;
}

// land(handle, state, pos, vel): a readback's positions and velocities into the packed
// state.  A node whose position in the state still equals the shadow has not been moved by
// the caller since the last exchange: it takes the device's position (and so does the
// shadow); a node the caller moved keeps its position, which the next push sends up.
// Velocities always come from the device.
napi_value Land(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  Box* b = box_of(env, argv[0]);
  if (!b) return nullptr;
  size_t k, kp, kv;
  double* st = typed<double>(env, argv[1], napi_float64_array, &k);
  const double* pos = typed<double>(env, argv[2], napi_float64_array, &kp);
  const double* vel = typed<double>(env, argv[3], napi_float64_array, &kv);
  if (k / STATE != b->n || kp < 3 * (size_t)b->n || kv < 3 * (size_t)b->n) {
    napi_throw_error(env, nullptr, "d3fgpu land: arrays do not match the node count");
    return nullptr;
  }
  const int nd = b->nd;
  for (uint32_t i = 0; i < b->n; ++i) {
    double* s = st + STATE * i;
    double* sh = b->shadow.data() + SHADOW * i;
    double p[3] = {0, 0, 0};
    for (int a = 0; a < nd; ++a) p[a] = s[a];
    if (same_bits(p, sh, 3))
      for (int a = 0; a < nd; ++a) s[a] = sh[a] = pos[3 * i + a];
    for (int a = 0; a < nd; ++a) s[3 + a] = vel[3 * i + a];
  }
  return nullptr;
}

// One force descriptor from its JS object (arrays are borrowed for the duration of a call).
void parse_desc(napi_env env, napi_value o, fg_force_desc& d) {
  memset(&d, 0, sizeof(d));
  d.type = (int32_t)num(env, prop(env, o, "type"));
  d.id = (uint64_t)num(env, prop(env, o, "id"));
  d.version = (uint64_t)num(env, prop(env, o, "version"));
  d.iterations = (int32_t)num(env, prop(env, o, "iterations"), 1);
  d.x = num(env, prop(env, o, "x"));
  d.y = num(env, prop(env, o, "y"));
  d.z = num(env, prop(env, o, "z"));
  d.strength = num(env, prop(env, o, "strength"), 1.0);
  d.theta2 = num(env, prop(env, o, "theta2"), 0.81);
  d.distance_min2 = num(env, prop(env, o, "distanceMin2"), 1.0);
  d.distance_max2 = num(env, prop(env, o, "distanceMax2"), 1e308 * 10);
  d.axis = (int32_t)num(env, prop(env, o, "axis"));
  size_t k;
  d.links = typed<uint32_t>(env, prop(env, o, "links"), napi_uint32_array, &k);
  d.count = (uint32_t)(k / 2);
  d.strengths = typed<double>(env, prop(env, o, "strengths"), napi_float64_array, &k);
  d.radii = typed<double>(env, prop(env, o, "radii"), napi_float64_array, &k);
  d.distances = typed<double>(env, prop(env, o, "distances"), napi_float64_array, &k);
  d.bias = typed<double>(env, prop(env, o, "bias"), napi_float64_array, &k);
  d.targets = typed<double>(env, prop(env, o, "targets"), napi_float64_array, &k);
}

// setForces(handle, [desc...]): desc = {type, id, version, iterations, x, y, z, strength,
// theta2, distanceMin2, distanceMax2, axis, links (Uint32Array), strengths, radii,
// distances, bias, targets (Float64Array)}; arrays may be null when the version is unchanged.
napi_value SetForces(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  uint32_t len = 0;
  napi_get_array_length(env, argv[1], &len);
  std::vector<fg_force_desc> descs(len);
  for (uint32_t i = 0; i < len; ++i) {
    napi_value o;
    napi_get_element(env, argv[1], i, &o);
    parse_desc(env, o, descs[i]);
  }
  int rc = fg_set_forces(c, len, descs.data());
  if (rc != FG_OK) return throw_status(env, c, rc, "setForces");
  return nullptr;
}

napi_value SetSeed(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  fg_set_seed(c, (uint32_t)num(env, argv[1]));
  return nullptr;
}

napi_value Enqueue(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  size_t k;
  double* a = typed<double>(env, argv[1], napi_float64_array, &k);
  int rc = fg_enqueue(c, (uint32_t)k, a, num(env, argv[2], 0.6));
  if (rc != FG_OK) return throw_status(env, c, rc, "enqueue");
  return nullptr;
}

napi_value Busy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int32_t b = 0;
  int rc = fg_busy(c, &b);
  if (rc != FG_OK) return throw_status(env, c, rc, "busy");
  napi_value r;
  NAPI_CALL(env, napi_get_boolean(env, b != 0, &r));
  return r;
}

napi_value ReadbackSync(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  size_t k;
  int rc = fg_readback(c, typed<double>(env, argv[1], napi_float64_array, &k),
                       typed<double>(env, argv[2], napi_float64_array, &k));
  if (rc != FG_OK) return throw_status(env, c, rc, "readback");
  return nullptr;
}

// ---- asynchronous waits ---------------------------------------------------------------

struct AsyncJob {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref pos_ref = nullptr, vel_ref = nullptr;
  fg_ctx* c = nullptr;
  uint64_t ticket = 0;
  bool readback = false;
  int rc = FG_OK;
};

void job_execute(napi_env, void* data) {
  AsyncJob* j = (AsyncJob*)data;
  j->rc = j->readback ? fg_readback_wait(j->c, j->ticket) : fg_sync(j->c);
}

void job_complete(napi_env env, napi_status status, void* data) {
  AsyncJob* j = (AsyncJob*)data;
  napi_value result;
  if (status == napi_ok && j->rc == FG_OK && j->readback) {
    napi_value pv, vv;
    napi_get_reference_value(env, j->pos_ref, &pv);
    napi_get_reference_value(env, j->vel_ref, &vv);
    size_t k;
    int rc = fg_readback_fetch(j->c, j->ticket, typed<double>(env, pv, napi_float64_array, &k),
                               typed<double>(env, vv, napi_float64_array, &k));
    // a readback that predates new uploads (epoch) resolves false: nothing landed
    napi_get_boolean(env, rc == FG_OK, &result);
    napi_resolve_deferred(env, j->deferred, result);
  } else if (status == napi_ok && j->rc == FG_OK) {
    napi_get_boolean(env, true, &result);
    napi_resolve_deferred(env, j->deferred, result);
  } else {
    std::string msg = std::string("d3fgpu wait failed: ") + fg_last_error(j->c);
    napi_value m, err;
    napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, nullptr, m, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  if (j->pos_ref) napi_delete_reference(env, j->pos_ref);
  if (j->vel_ref) napi_delete_reference(env, j->vel_ref);
  napi_delete_async_work(env, j->work);
  delete j;
}

napi_value queue_job(napi_env env, AsyncJob* j, const char* name) {
  napi_value promise, rname;
  NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
  NAPI_CALL(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &rname));
  NAPI_CALL(env, napi_create_async_work(env, nullptr, rname, job_execute, job_complete, j, &j->work));
  NAPI_CALL(env, napi_queue_async_work(env, j->work));
  return promise;
}

// readbackAsync(handle, Float64Array pos, Float64Array vel) -> Promise<boolean>
napi_value ReadbackAsync(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  AsyncJob* j = new AsyncJob();
  j->c = c;
  j->readback = true;
  int rc = fg_readback_start(c, &j->ticket);
  if (rc != FG_OK) {
    delete j;
    return throw_status(env, c, rc, "readbackAsync");
  }
  napi_create_reference(env, argv[1], 1, &j->pos_ref);
  napi_create_reference(env, argv[2], 1, &j->vel_ref);
  return queue_job(env, j, "d3fgpu.readback");
}

napi_value IdleAsync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  AsyncJob* j = new AsyncJob();
  j->c = c;
  return queue_job(env, j, "d3fgpu.idle");
}

napi_value Buffers(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  fg_buffer_info b;
  int rc = fg_buffers(c, &b);
  if (rc != FG_OK) return throw_status(env, c, rc, "buffers");
  napi_value o, v;
  NAPI_CALL(env, napi_create_object(env, &o));
  NAPI_CALL(env, napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)b.positions, &v));
  napi_set_named_property(env, o, "positions", v);
  NAPI_CALL(env, napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)b.velocities, &v));
  napi_set_named_property(env, o, "velocities", v);
  napi_create_uint32(env, b.count, &v);
  napi_set_named_property(env, o, "count", v);
  napi_create_uint32(env, b.stride, &v);
  napi_set_named_property(env, o, "stride", v);
  napi_create_double(env, (double)b.epoch, &v);
  napi_set_named_property(env, o, "epoch", v);
  return o;
}

// readDevice(handle, pointer (BigInt), floats) -> Float32Array: a copy of device memory,
// for tests of what gpuBuffers() points to (fg_debug_copy_from_device).
napi_value ReadDevice(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  get_args(env, info, 3, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  uint64_t addr = 0;
  bool lossless = false;
  NAPI_CALL(env, napi_get_value_bigint_uint64(env, argv[1], &addr, &lossless));
  const size_t count = (size_t)num(env, argv[2]);
  napi_value buf, arr;
  void* data = nullptr;
  NAPI_CALL(env, napi_create_arraybuffer(env, sizeof(float) * count, &data, &buf));
  int rc = fg_debug_copy_from_device(c, (const void*)(uintptr_t)addr, sizeof(float) * count, data);
  if (rc != FG_OK) return throw_status(env, c, rc, "readDevice");
  NAPI_CALL(env, napi_create_typedarray(env, napi_float32_array, count, buf, 0, &arr));
  return arr;
}

napi_value Stats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  fg_stats s;
  int rc = fg_get_stats(c, &s);
  if (rc != FG_OK) return throw_status(env, c, rc, "stats");
  napi_value o, v;
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_create_double(env, (double)s.steps, &v);
  napi_set_named_property(env, o, "steps", v);
  napi_create_double(env, (double)s.interactions, &v);
  napi_set_named_property(env, o, "interactions", v);
  napi_create_double(env, (double)s.walk_errors, &v);
  napi_set_named_property(env, o, "walkErrors", v);
  return o;
}

// ---- batches and serving -------------------------------------------------------------

// A job from {n, nDim, pos, vel?, fix?, fixMask?, forces: [desc...], alpha, alphaMin,
// alphaDecay, alphaTarget, velocityDecay, seed, maxTicks}; `descs` holds the forces.
void parse_job(napi_env env, napi_value o, fg_job_desc& j, std::vector<fg_force_desc>& descs) {
  memset(&j, 0, sizeof(j));
  j.n = (uint32_t)num(env, prop(env, o, "n"));
  j.ndim = (int32_t)num(env, prop(env, o, "nDim"), 2);
  size_t k;
  j.pos = typed<double>(env, prop(env, o, "pos"), napi_float64_array, &k);
  j.vel = typed<double>(env, prop(env, o, "vel"), napi_float64_array, &k);
  j.fix = typed<double>(env, prop(env, o, "fix"), napi_float64_array, &k);
  j.fixmask = typed<uint8_t>(env, prop(env, o, "fixMask"), napi_uint8_array, &k);
  napi_value fs = prop(env, o, "forces");
  uint32_t len = 0;
  napi_get_array_length(env, fs, &len);
  descs.resize(len);
  for (uint32_t i = 0; i < len; ++i) {
    napi_value f;
    napi_get_element(env, fs, i, &f);
    parse_desc(env, f, descs[i]);
  }
  j.nforces = len;
  j.forces = descs.data();
  j.alpha = num(env, prop(env, o, "alpha"), 1.0);
  j.alpha_min = num(env, prop(env, o, "alphaMin"), 0.001);
  j.alpha_decay = num(env, prop(env, o, "alphaDecay"), 1.0 - 0.9772372209558107);
  j.alpha_target = num(env, prop(env, o, "alphaTarget"), 0.0);
  j.velocity_decay = num(env, prop(env, o, "velocityDecay"), 0.4);
  j.seed = (uint32_t)num(env, prop(env, o, "seed"));
  j.max_ticks = (uint32_t)num(env, prop(env, o, "maxTicks"));
}

struct BatchBox {
  fg_batch* b;
};
struct ServerBox {
  fg_server* s;
};

void finalize_batch(napi_env, void* data, void*) {
  BatchBox* b = (BatchBox*)data;
  if (b && b->b) fg_batch_destroy(b->b);
  delete b;
}
void finalize_server(napi_env, void* data, void*) {
  ServerBox* b = (ServerBox*)data;
  if (b && b->s) fg_serve_destroy(b->s);
  delete b;
}

fg_batch* batch_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((BatchBox*)p)->b) {
    napi_throw_error(env, nullptr, "d3fgpu: not a live batch handle");
    return nullptr;
  }
  return ((BatchBox*)p)->b;
}
fg_server* server_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((ServerBox*)p)->s) {
    napi_throw_error(env, nullptr, "d3fgpu: not a live server handle");
    return nullptr;
  }
  return ((ServerBox*)p)->s;
}

napi_value batch_fail(napi_env env, fg_batch* b, int rc, const char* what) {
  std::string msg = std::string("d3fgpu ") + what + " failed (" + std::to_string(rc) + ")";
  if (b) msg += ": " + std::string(fg_batch_last_error(b));
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

// batchCreate(device, budgetBytes) -> handle
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_batch* b = nullptr;
  int rc = fg_batch_create(&b, (int)num(env, argv[0]), (uint64_t)num(env, argv[1]));
  if (rc != FG_OK) return batch_fail(env, nullptr, rc, "batchCreate");
  napi_value r;
  NAPI_CALL(env, napi_create_external(env, new BatchBox{b}, finalize_batch, nullptr, &r));
  return r;
}

napi_value BatchDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p) return nullptr;
  BatchBox* b = (BatchBox*)p;
  if (b->b) fg_batch_destroy(b->b);
  b->b = nullptr;
  return nullptr;
}

// batchSubmit(handle, job) -> id (a number); null when the budget has no room now
napi_value BatchSubmit(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_batch* b = batch_of(env, argv[0]);
  if (!b) return nullptr;
  fg_job_desc j;
  std::vector<fg_force_desc> descs;
  parse_job(env, argv[1], j, descs);
  uint64_t id = 0;
  int rc = fg_batch_submit(b, &j, &id);
  napi_value r;
  if (rc == FG_ERR_NOMEM) {
    napi_get_null(env, &r);
    return r;
  }
  if (rc != FG_OK) return batch_fail(env, b, rc, "batchSubmit");
  NAPI_CALL(env, napi_create_double(env, (double)id, &r));
  return r;
}

napi_value BatchStep(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_batch* b = batch_of(env, argv[0]);
  if (!b) return nullptr;
  int rc = fg_batch_step(b, (uint32_t)num(env, argv[1], 1));
  if (rc != FG_OK) return batch_fail(env, b, rc, "batchStep");
  return nullptr;
}

// batchJobs(handle) -> [{id, done, ticks, alpha}]
napi_value BatchJobs(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_batch* b = batch_of(env, argv[0]);
  if (!b) return nullptr;
  std::vector<fg_job_info> jobs(4096);
  uint32_t count = 0;
  int rc = fg_batch_jobs(b, jobs.data(), (uint32_t)jobs.size(), &count);
  if (rc != FG_OK) return batch_fail(env, b, rc, "batchJobs");
  napi_value arr, o, v;
  NAPI_CALL(env, napi_create_array_with_length(env, count, &arr));
  for (uint32_t i = 0; i < count && i < jobs.size(); ++i) {
    napi_create_object(env, &o);
    napi_create_double(env, (double)jobs[i].id, &v);
    napi_set_named_property(env, o, "id", v);
    napi_get_boolean(env, jobs[i].status == FG_JOB_DONE, &v);
    napi_set_named_property(env, o, "done", v);
    napi_create_uint32(env, jobs[i].ticks, &v);
    napi_set_named_property(env, o, "ticks", v);
    napi_create_double(env, jobs[i].alpha, &v);
    napi_set_named_property(env, o, "alpha", v);
    napi_set_element(env, arr, i, o);
  }
  return arr;
}

// batchRead(handle, id, Float64Array pos, Float64Array vel)
napi_value BatchRead(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  get_args(env, info, 4, argv);
  fg_batch* b = batch_of(env, argv[0]);
  if (!b) return nullptr;
  size_t k;
  int rc = fg_batch_read(b, (uint64_t)num(env, argv[1]), typed<double>(env, argv[2], napi_float64_array, &k),
                         typed<double>(env, argv[3], napi_float64_array, &k));
  if (rc != FG_OK) return batch_fail(env, b, rc, "batchRead");
  return nullptr;
}

napi_value BatchRelease(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_batch* b = batch_of(env, argv[0]);
  if (!b) return nullptr;
  int rc = fg_batch_release(b, (uint64_t)num(env, argv[1]));
  if (rc != FG_OK) return batch_fail(env, b, rc, "batchRelease");
  return nullptr;
}

// batchAutotune(handle, cachePath?) -> {radixBits, walkWaves, walkStack}
napi_value BatchAutotune(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_batch* b = batch_of(env, argv[0]);
  if (!b) return nullptr;
  std::string path;
  if (!is_nullish(env, argv[1])) {
    size_t len = 0;
    napi_get_value_string_utf8(env, argv[1], nullptr, 0, &len);
    path.resize(len + 1);
    napi_get_value_string_utf8(env, argv[1], &path[0], len + 1, &len);
    path.resize(len);
  }
  fg_batch_tune t;
  int rc = fg_batch_autotune(b, path.empty() ? nullptr : path.c_str(), &t);
  if (rc != FG_OK) return batch_fail(env, b, rc, "batchAutotune");
  napi_value o, v;
  NAPI_CALL(env, napi_create_object(env, &o));
  napi_create_int32(env, t.radix_bits, &v);
  napi_set_named_property(env, o, "radixBits", v);
  napi_create_int32(env, t.walk_waves, &v);
  napi_set_named_property(env, o, "walkWaves", v);
  napi_create_int32(env, t.walk_stack, &v);
  napi_set_named_property(env, o, "walkStack", v);
  return o;
}

// serveCreate(device, {budget, ticksPerRound}) -> handle
napi_value ServeCreate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_serve_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  if (!is_nullish(env, argv[1])) {
    cfg.hbm_budget = (uint64_t)num(env, prop(env, argv[1], "budget"));
    cfg.ticks_per_round = (uint32_t)num(env, prop(env, argv[1], "ticksPerRound"));
  }
  fg_server* s = nullptr;
  int rc = fg_serve_create(&s, (int)num(env, argv[0]), &cfg);
  if (rc != FG_OK) return throw_status(env, nullptr, rc, "serveCreate");
  napi_value r;
  NAPI_CALL(env, napi_create_external(env, new ServerBox{s}, finalize_server, nullptr, &r));
  return r;
}

napi_value ServeDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p) return nullptr;
  ServerBox* b = (ServerBox*)p;
  if (b->s) fg_serve_destroy(b->s);
  b->s = nullptr;
  return nullptr;
}

// serveSubmit(handle, job) -> id
napi_value ServeSubmit(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  get_args(env, info, 2, argv);
  fg_server* s = server_of(env, argv[0]);
  if (!s) return nullptr;
  fg_job_desc j;
  std::vector<fg_force_desc> descs;
  parse_job(env, argv[1], j, descs);
  uint64_t id = 0;
  int rc = fg_serve_submit(s, &j, &id);
  if (rc != FG_OK) {
    std::string msg = std::string("d3fgpu serveSubmit failed (") + std::to_string(rc) + "): " + fg_serve_last_error(s);
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
  }
  napi_value r;
  NAPI_CALL(env, napi_create_double(env, (double)id, &r));
  return r;
}

struct ServeWait {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref pos_ref = nullptr, vel_ref = nullptr;
  fg_server* s = nullptr;
  uint64_t id = 0;
  double timeout = 0;
  double *pos = nullptr, *vel = nullptr;
  fg_serve_result r{};
  int rc = FG_OK;
  std::string err;
};

void serve_execute(napi_env, void* data) {
  ServeWait* w = (ServeWait*)data;
  w->rc = fg_serve_wait(w->s, w->id, w->timeout, w->pos, w->vel, &w->r);
  if (w->rc != FG_OK) w->err = fg_serve_last_error(w->s);
}

void serve_complete(napi_env env, napi_status status, void* data) {
  ServeWait* w = (ServeWait*)data;
  if (status == napi_ok && w->rc == FG_OK) {
    napi_value o, v;
    napi_create_object(env, &o);
    napi_create_uint32(env, w->r.ticks, &v);
    napi_set_named_property(env, o, "ticks", v);
    napi_create_double(env, w->r.submitted_s, &v);
    napi_set_named_property(env, o, "submitted", v);
    napi_create_double(env, w->r.admitted_s, &v);
    napi_set_named_property(env, o, "admitted", v);
    napi_create_double(env, w->r.finished_s, &v);
    napi_set_named_property(env, o, "finished", v);
    napi_resolve_deferred(env, w->deferred, o);
  } else {
    std::string msg = std::string("d3fgpu layout failed: ") + w->err;
    napi_value m, e;
    napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, nullptr, m, &e);
    napi_reject_deferred(env, w->deferred, e);
  }
  if (w->pos_ref) napi_delete_reference(env, w->pos_ref);
  if (w->vel_ref) napi_delete_reference(env, w->vel_ref);
  napi_delete_async_work(env, w->work);
  delete w;
}

// serveWaitAsync(handle, id, timeoutSeconds, Float64Array pos, Float64Array vel) ->
// Promise<{ticks, submitted, admitted, finished}>; the arrays receive the final state
napi_value ServeWaitAsync(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  get_args(env, info, 5, argv);
  fg_server* s = server_of(env, argv[0]);
  if (!s) return nullptr;
  ServeWait* w = new ServeWait();
  w->s = s;
  w->id = (uint64_t)num(env, argv[1]);
  w->timeout = num(env, argv[2], 600.0);
  size_t k;
  w->pos = typed<double>(env, argv[3], napi_float64_array, &k);
  w->vel = typed<double>(env, argv[4], napi_float64_array, &k);
  if (w->pos) napi_create_reference(env, argv[3], 1, &w->pos_ref);
  if (w->vel) napi_create_reference(env, argv[4], 1, &w->vel_ref);
  napi_value promise, rname;
  NAPI_CALL(env, napi_create_promise(env, &w->deferred, &promise));
  NAPI_CALL(env, napi_create_string_utf8(env, "d3fgpu.serveWait", NAPI_AUTO_LENGTH, &rname));
  NAPI_CALL(env, napi_create_async_work(env, nullptr, rname, serve_execute, serve_complete, w, &w->work));
  NAPI_CALL(env, napi_queue_async_work(env, w->work));
  return promise;
}

// serveStats(handle) -> {submitted, finished, failed, queued, running, p50, p99, layoutsPerSecond}
napi_value ServeStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  get_args(env, info, 1, argv);
  fg_server* s = server_of(env, argv[0]);
  if (!s) return nullptr;
  fg_serve_stats st;
  int rc = fg_serve_get_stats(s, &st);
  if (rc != FG_OK) return throw_status(env, nullptr, rc, "serveStats");
  napi_value o, v;
  NAPI_CALL(env, napi_create_object(env, &o));
  const struct {
    const char* name;
    double value;
  } fields[] = {{"submitted", (double)st.submitted}, {"finished", (double)st.finished},
                {"failed", (double)st.failed},       {"queued", (double)st.queued},
                {"running", (double)st.running},     {"peakRunning", (double)st.peak_running},
                {"p50", st.p50_latency_s},          {"p99", st.p99_latency_s},
                {"layoutsPerSecond", st.layouts_per_s}};
  for (const auto& f : fields) {
    napi_create_double(env, f.value, &v);
    napi_set_named_property(env, o, f.name, v);
  }
  return o;
}

napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"deviceCount", nullptr, DeviceCount, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fits", nullptr, Fits, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setNodes", nullptr, SetNodes, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"load", nullptr, Load, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"push", nullptr, Push, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"land", nullptr, Land, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"updatePositions", nullptr, UpdatePositions, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"updatePins", nullptr, UpdatePins, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setForces", nullptr, SetForces, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setSeed", nullptr, SetSeed, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"enqueue", nullptr, Enqueue, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"busy", nullptr, Busy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"readbackSync", nullptr, ReadbackSync, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"readbackAsync", nullptr, ReadbackAsync, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"idleAsync", nullptr, IdleAsync, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"buffers", nullptr, Buffers, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"stats", nullptr, Stats, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"readDevice", nullptr, ReadDevice, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchCreate", nullptr, BatchCreate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchDestroy", nullptr, BatchDestroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchSubmit", nullptr, BatchSubmit, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchStep", nullptr, BatchStep, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchJobs", nullptr, BatchJobs, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchRead", nullptr, BatchRead, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchRelease", nullptr, BatchRelease, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"batchAutotune", nullptr, BatchAutotune, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"serveCreate", nullptr, ServeCreate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"serveDestroy", nullptr, ServeDestroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"serveSubmit", nullptr, ServeSubmit, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"serveWaitAsync", nullptr, ServeWaitAsync, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"serveStats", nullptr, ServeStats, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
