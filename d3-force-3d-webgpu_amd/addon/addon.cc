// addon.cc -- N-API binding of the C ABI (include/d3fgpu.h) for the Node.js host.
//
// A thin layer: every JS call maps to one fg_* call.  The JS engine (js/engine.js) keeps the
// reference engine's logic (edit diffing, readback landing, busy/idle, src/gpu/engine.js);
// this file only converts arguments.  Waiting happens off the JS thread: readbackAsync and
// idleAsync run fg_readback_wait / fg_sync in napi_async_work and settle a Promise.
#include <node_api.h>
#include <stdint.h>
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

// The JS handle owns a box, so that destroy() can free the context before the handle is
// collected and the finalizer then finds nothing to free.
struct Box {
  fg_ctx* c;
};

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
    fg_force_desc& d = descs[i];
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

napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"deviceCount", nullptr, DeviceCount, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fits", nullptr, Fits, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setNodes", nullptr, SetNodes, nullptr, nullptr, nullptr, napi_default, nullptr},
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
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
