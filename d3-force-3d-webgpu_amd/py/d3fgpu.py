"""Python bindings of the MI355X physics backend's C ABI (include/d3fgpu.h), over ctypes.

The same calls a ctypes-based host would make: one `Engine` wraps one fg_ctx (one
simulation on one GPU) and mirrors the reference engine's methods (src/gpu/engine.js:
setNodes, setForces, enqueue, readback, sync, buffers, fits, setSeed).  Loading fails
loudly when lib/libd3fgpu.so has not been built: there is no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libd3fgpu.so")

FG_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of memory", -4: "unsupported",
          -5: "bad state", -6: "no device"}

CENTER, COLLIDE, LINK, MANYBODY, POSITION, RADIAL = 1, 2, 3, 4, 5, 6

_d = C.POINTER(C.c_double)
_u = C.POINTER(C.c_uint32)
_f = C.POINTER(C.c_float)
_u8 = C.POINTER(C.c_uint8)
_i = C.POINTER(C.c_int32)


class ForceDesc(C.Structure):
    """fg_force_desc."""
    _fields_ = [
        ("type", C.c_int32), ("iterations", C.c_int32), ("id", C.c_uint64), ("version", C.c_uint64),
        ("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("strength", C.c_double),
        ("theta2", C.c_double), ("distance_min2", C.c_double), ("distance_max2", C.c_double),
        ("axis", C.c_int32), ("count", C.c_uint32), ("links", _u), ("strengths", _d), ("radii", _d),
        ("distances", _d), ("bias", _d), ("targets", _d),
    ]


class BufferInfo(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("velocities", C.c_void_p), ("count", C.c_uint32),
                ("stride", C.c_uint32), ("epoch", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("interactions", C.c_uint64), ("opened", C.c_uint64),
                ("lane_visits", C.c_uint64), ("sort_bits", C.c_uint32), ("walk_errors", C.c_uint32),
                ("walk_max_stack", C.c_uint32)]


class Timing(C.Structure):
    _fields_ = [("ms", C.c_double * 16), ("launches", C.c_uint64 * 16)]


class JobDesc(C.Structure):
    """fg_job_desc."""
    _fields_ = [("n", C.c_uint32), ("ndim", C.c_int32), ("pos", _d), ("vel", _d), ("fix", _d), ("fixmask", _u8),
                ("nforces", C.c_uint32), ("forces", C.POINTER(ForceDesc)), ("alpha", C.c_double),
                ("alpha_min", C.c_double), ("alpha_decay", C.c_double), ("alpha_target", C.c_double),
                ("velocity_decay", C.c_double), ("seed", C.c_uint32), ("max_ticks", C.c_uint32)]


class JobInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_uint32), ("ticks", C.c_uint32), ("alpha", C.c_double),
                ("n", C.c_uint32), ("ndim", C.c_int32)]


class BatchTune(C.Structure):
    _fields_ = [("radix_bits", C.c_int32), ("walk_waves", C.c_int32), ("walk_stack", C.c_int32),
                ("walk_blocks_per_cu", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("jobs_running", C.c_uint32), ("jobs_done", C.c_uint32), ("nodes", C.c_uint64),
                ("bytes", C.c_uint64), ("budget", C.c_uint64), ("ticks", C.c_uint64),
                ("graph_builds", C.c_uint64), ("interactions", C.c_uint64), ("walk_errors", C.c_uint32),
                ("launches_per_tick", C.c_uint32), ("walk_max_stack", C.c_uint32)]


JOB_RUNNING, JOB_DONE = 1, 3


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("ticks_per_round", C.c_uint32), ("autotune", C.c_int32),
                ("tune_cache", C.c_char_p)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_uint32), ("n", C.c_uint32), ("ndim", C.c_int32),
                ("ticks", C.c_uint32), ("submitted_s", C.c_double), ("admitted_s", C.c_double),
                ("finished_s", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("finished", C.c_uint64),
                ("failed", C.c_uint64), ("queued", C.c_uint32), ("running", C.c_uint32),
                ("peak_running", C.c_uint32), ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double),
                ("mean_latency_s", C.c_double), ("max_latency_s", C.c_double), ("layouts_per_s", C.c_double),
                ("rounds", C.c_uint64), ("ticks", C.c_uint64), ("peak_bytes", C.c_uint64), ("budget", C.c_uint64)]


SERVE_QUEUED, SERVE_RUNNING, SERVE_FINISHED, SERVE_FAILED = 0, 1, 2, 3

PHASES = ["tick", "tree_build", "tree_sort", "walk", "pairs", "link", "collide", "center", "radial",
          "position", "integrate"]

# every entry point of include/d3fgpu.h: name -> (restype, argtypes)
SIGNATURES = {
    "fg_abi_version": (C.c_int, []),
    "fg_device_count": (C.c_int, [_i]),
    "fg_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "fg_destroy": (C.c_int, [C.c_void_p]),
    "fg_last_error": (C.c_char_p, [C.c_void_p]),
    "fg_fits": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, _i]),
    "fg_set_nodes": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32, _d, _d, _d, _u8]),
    "fg_update_positions": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _d]),
    "fg_update_velocities": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _d]),
    "fg_update_pins": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _d, _u8]),
    "fg_set_forces": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(ForceDesc)]),
    "fg_set_seed": (C.c_int, [C.c_void_p, C.c_uint32]),
    "fg_enqueue": (C.c_int, [C.c_void_p, C.c_uint32, _d, C.c_double]),
    "fg_apply_forces": (C.c_int, [C.c_void_p, C.c_double]),
    "fg_busy": (C.c_int, [C.c_void_p, _i]),
    "fg_sync": (C.c_int, [C.c_void_p]),
    "fg_readback": (C.c_int, [C.c_void_p, _d, _d]),
    "fg_readback_start": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "fg_readback_wait": (C.c_int, [C.c_void_p, C.c_uint64]),
    "fg_readback_fetch": (C.c_int, [C.c_void_p, C.c_uint64, _d, _d]),
    "fg_buffers": (C.c_int, [C.c_void_p, C.POINTER(BufferInfo)]),
    "fg_debug_copy_from_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "fg_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "fg_timing_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "fg_get_timing": (C.c_int, [C.c_void_p, C.POINTER(Timing)]),
    "fg_radix_sort_pairs": (C.c_int, [C.c_void_p, C.c_uint32, _u, _u, C.c_uint32, C.c_uint32, _u, _u]),
    "fg_radix_sort_pairs_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                              C.c_uint32, C.c_void_p, C.c_void_p]),
    "fg_debug_tree": (C.c_int, [C.c_void_p, _f, _i, _u, _u, _u, _u, _u, _u, _u, _d, _d]),
    "fg_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_uint64]),
    "fg_batch_destroy": (C.c_int, [C.c_void_p]),
    "fg_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "fg_batch_job_bytes": (C.c_int, [C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "fg_batch_submit": (C.c_int, [C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "fg_batch_step": (C.c_int, [C.c_void_p, C.c_uint32]),
    "fg_batch_sync": (C.c_int, [C.c_void_p]),
    "fg_batch_jobs": (C.c_int, [C.c_void_p, C.POINTER(JobInfo), C.c_uint32, C.POINTER(C.c_uint32)]),
    "fg_batch_read": (C.c_int, [C.c_void_p, C.c_uint64, _d, _d]),
    "fg_batch_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "fg_batch_set_tune": (C.c_int, [C.c_void_p, C.POINTER(BatchTune)]),
    "fg_batch_get_tune": (C.c_int, [C.c_void_p, C.POINTER(BatchTune)]),
    "fg_batch_autotune": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(BatchTune)]),
    "fg_batch_use_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
    "fg_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "fg_serve_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(ServeConfig)]),
    "fg_serve_destroy": (C.c_int, [C.c_void_p]),
    "fg_serve_last_error": (C.c_char_p, [C.c_void_p]),
    "fg_serve_submit": (C.c_int, [C.c_void_p, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
    "fg_serve_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, _d, _d, C.POINTER(ServeResult)]),
    "fg_serve_drain": (C.c_int, [C.c_void_p, C.c_double]),
    "fg_serve_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServeStats)]),
}

_lib = None


def lib():
    """The loaded backend library (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"backend library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class FgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{ERRORS.get(status, status)}: {msg}")
        self.status = status


def ptr(a, t=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def device_count():
    k = C.c_int32(0)
    lib().fg_device_count(C.byref(k))
    return k.value


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        assert a.shape[0] == n
    return a


class Force:
    """A force as its gpu() descriptor describes it; keeps its arrays alive."""
    _next_id = 1

    def __init__(self, type, **kw):
        self.type = type
        self.id = Force._next_id
        Force._next_id += 1
        self.version = 1
        self.kw = kw

    def bump(self, **kw):
        self.kw.update(kw)
        self.version += 1

    def desc(self):
        k = self.kw
        d = ForceDesc()
        d.type, d.id, d.version = self.type, self.id, self.version
        d.iterations = int(k.get("iterations", 1))
        d.x, d.y, d.z = (float(k.get(a, 0.0)) for a in "xyz")
        d.strength = float(k.get("strength", 1.0))
        d.theta2 = float(k.get("theta2", 0.81))
        d.distance_min2 = float(k.get("distance_min2", 1.0))
        d.distance_max2 = float(k.get("distance_max2", np.inf))
        d.axis = int(k.get("axis", 0))
        keep = []
        for name in ("strengths", "radii", "distances", "bias", "targets"):
            if k.get(name) is not None:
                arr = _f64(k[name])
                keep.append(arr)
                setattr(d, name, ptr(arr))
        if k.get("links") is not None:
            links = np.ascontiguousarray(k["links"], dtype=np.uint32).reshape(-1)
            keep.append(links)
            d.links = ptr(links, C.c_uint32)
            d.count = len(links) // 2
        return d, keep


def many_body(strengths, theta=0.9, distance_min=1.0, distance_max=np.inf):
    return Force(MANYBODY, strengths=strengths, theta2=theta * theta, distance_min2=distance_min ** 2,
                 distance_max2=distance_max ** 2)


def link(src, dst, strengths, distances, bias, iterations=1):
    links = np.stack([np.asarray(src, np.uint32), np.asarray(dst, np.uint32)], 1)
    return Force(LINK, links=links, strengths=strengths, distances=distances, bias=bias, iterations=iterations)


def collide(radii, strength=1.0, iterations=1):
    return Force(COLLIDE, radii=radii, strength=strength, iterations=iterations)


def center(x=0.0, y=0.0, z=0.0, strength=1.0):
    return Force(CENTER, x=x, y=y, z=z, strength=strength)


def radial(radii, strengths, x=0.0, y=0.0, z=0.0):
    return Force(RADIAL, radii=radii, strengths=strengths, x=x, y=y, z=z)


def position(axis, targets, strengths):
    return Force(POSITION, axis=axis, targets=targets, strengths=strengths)


class Engine:
    """One simulation on one GPU (the reference's engine object)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().fg_create(C.byref(h), device)
        if rc != FG_OK:
            raise FgError(rc, "fg_create failed")
        self.h = h
        self.n = 0
        self.ndim = 2

    def close(self):
        if self.h:
            lib().fg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != FG_OK:
            raise FgError(rc, lib().fg_last_error(self.h).decode())

    def fits(self, n, ndim):
        out = C.c_int32(0)
        self._check(lib().fg_fits(self.h, n, ndim, C.byref(out)))
        return bool(out.value)

    def set_nodes(self, pos, vel=None, fix=None, fixmask=None, ndim=3):
        pos = _f64(pos)
        n = len(pos)
        pos3 = np.zeros((n, 3))
        pos3[:, :pos.shape[1]] = pos
        vel3 = None
        if vel is not None:
            vel = _f64(vel)
            vel3 = np.zeros((n, 3))
            vel3[:, :vel.shape[1]] = vel
        fix3 = None if fix is None else np.ascontiguousarray(_f64(fix).reshape(n, 3))
        mask = None if fixmask is None else np.ascontiguousarray(fixmask, dtype=np.uint8)
        self._check(lib().fg_set_nodes(self.h, n, ndim, ptr(pos3), ptr(vel3), ptr(fix3), ptr(mask, C.c_uint8)))
        self.n, self.ndim = n, ndim

    def update_positions(self, first, pos):
        pos = np.ascontiguousarray(_f64(pos).reshape(-1, 3))
        self._check(lib().fg_update_positions(self.h, first, len(pos), ptr(pos)))

    def update_velocities(self, first, vel):
        vel = np.ascontiguousarray(_f64(vel).reshape(-1, 3))
        self._check(lib().fg_update_velocities(self.h, first, len(vel), ptr(vel)))

    def update_pins(self, first, fix, fixmask):
        fix = np.ascontiguousarray(_f64(fix).reshape(-1, 3))
        mask = np.ascontiguousarray(fixmask, dtype=np.uint8)
        self._check(lib().fg_update_pins(self.h, first, len(mask), ptr(fix), ptr(mask, C.c_uint8)))

    def set_forces(self, forces):
        descs, keep = [], []
        for f in forces:
            d, k = f.desc()
            descs.append(d)
            keep.extend(k)
        arr = (ForceDesc * max(1, len(descs)))(*descs)
        self._check(lib().fg_set_forces(self.h, len(descs), arr))

    def set_seed(self, seed):
        self._check(lib().fg_set_seed(self.h, seed & 0xffffffff))

    def enqueue(self, alphas, decay=0.6):
        a = _f64(np.atleast_1d(alphas))
        self._check(lib().fg_enqueue(self.h, len(a), ptr(a), decay))

    def apply_forces(self, alpha):
        self._check(lib().fg_apply_forces(self.h, alpha))

    def busy(self):
        b = C.c_int32(0)
        self._check(lib().fg_busy(self.h, C.byref(b)))
        return bool(b.value)

    def sync(self):
        self._check(lib().fg_sync(self.h))

    def readback(self, out=None):
        """Positions and velocities as (n, 3) float64 arrays; out: a (pos, vel) pair of such
        arrays (C-contiguous) to fill instead of new ones."""
        if out is None:
            pos = np.zeros((self.n, 3))
            vel = np.zeros((self.n, 3))
        else:
            pos, vel = out
            for a in (pos, vel):
                if a.shape != (self.n, 3) or a.dtype != np.float64 or not a.flags.c_contiguous:
                    raise ValueError("readback: out arrays must be C-contiguous float64 of shape (n, 3)")
        self._check(lib().fg_readback(self.h, ptr(pos), ptr(vel)))
        return pos, vel

    def readback_async(self):
        t = C.c_uint64(0)
        self._check(lib().fg_readback_start(self.h, C.byref(t)))
        return t.value

    def readback_wait(self, ticket):
        self._check(lib().fg_readback_wait(self.h, ticket))

    def readback_fetch(self, ticket):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        self._check(lib().fg_readback_fetch(self.h, ticket, ptr(pos), ptr(vel)))
        return pos, vel

    def buffers(self):
        b = BufferInfo()
        self._check(lib().fg_buffers(self.h, C.byref(b)))
        return b

    def copy_from_device(self, pointer, count, dtype=np.float32):
        """`count` elements of device memory at `pointer` (e.g. a buffers() pointer)."""
        out = np.zeros(count, dtype)
        self._check(lib().fg_debug_copy_from_device(self.h, C.c_void_p(pointer), out.nbytes,
                                                    out.ctypes.data_as(C.c_void_p)))
        return out

    def stats(self):
        s = Stats()
        self._check(lib().fg_get_stats(self.h, C.byref(s)))
        return s

    def timing_enable(self, on=True):
        self._check(lib().fg_timing_enable(self.h, 1 if on else 0))

    def timing(self):
        t = Timing()
        self._check(lib().fg_get_timing(self.h, C.byref(t)))
        return {name: (t.ms[i], int(t.launches[i])) for i, name in enumerate(PHASES)}

    def radix_sort_pairs(self, keys, values, first_bit=0, bits=32):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        values = np.ascontiguousarray(values, dtype=np.uint32)
        n = len(keys)
        ko, vo = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        self._check(lib().fg_radix_sort_pairs(self.h, n, ptr(keys, C.c_uint32), ptr(values, C.c_uint32),
                                              first_bit, bits, ptr(ko, C.c_uint32), ptr(vo, C.c_uint32)))
        return ko, vo

    def radix_sort_pairs_noalloc(self, n, keys, values, keys_out, values_out, first_bit=0, bits=32):
        """The allocation-free sort on device buffers: keys, values, keys_out and values_out are
        device addresses (ints) of n u32 each."""
        self._check(lib().fg_radix_sort_pairs_noalloc(self.h, n, C.c_void_p(keys), C.c_void_p(values), first_bit,
                                                      bits, C.c_void_p(keys_out), C.c_void_p(values_out)))

    def debug_tree(self):
        n = self.n
        inner = max(1, n - 1)
        cube = np.zeros(4, np.float32)
        shift = C.c_int32(0)
        codes = np.zeros(2 * n, np.uint32)
        order = np.zeros(n, np.uint32)
        leaf_depth = np.zeros(n, np.uint32)
        left, right, depth = (np.zeros(inner, np.uint32) for _ in range(3))
        kids = np.zeros(8 * inner, np.uint32)
        com = np.zeros(3 * inner)
        value = np.zeros(inner)
        u = C.c_uint32
        self._check(lib().fg_debug_tree(self.h, ptr(cube, C.c_float), C.byref(shift), ptr(codes, u), ptr(order, u),
                                        ptr(leaf_depth, u), ptr(left, u), ptr(right, u), ptr(depth, u),
                                        ptr(kids, u), ptr(com), ptr(value)))
        return dict(cube=cube, shift=shift.value, hi=codes[0::2].copy(), lo=codes[1::2].copy(), order=order,
                    leaf_depth=leaf_depth, left=left[:n - 1], right=right[:n - 1], depth=depth[:n - 1],
                    kids=kids.reshape(inner, 8)[:n - 1], com=com.reshape(inner, 3)[:n - 1], value=value[:n - 1])


ALPHA_DECAY = 1 - 0.001 ** (1 / 300)


class Job:
    """One layout for a batch: nodes, forces (in the order link, manyBody, collide, center,
    radial, x, y, z) and d3's schedule."""

    def __init__(self, pos, forces, ndim=None, vel=None, fix=None, fixmask=None, alpha=1.0, alpha_min=0.001,
                 alpha_decay=ALPHA_DECAY, alpha_target=0.0, velocity_decay=0.4, seed=0, max_ticks=0):
        pos = _f64(pos)
        self.n = len(pos)
        self.ndim = ndim or pos.shape[1]
        self.pos3 = np.zeros((self.n, 3))
        self.pos3[:, :pos.shape[1]] = pos
        self.vel3 = None
        if vel is not None:
            vel = _f64(vel)
            self.vel3 = np.zeros((self.n, 3))
            self.vel3[:, :vel.shape[1]] = vel
        self.fix3 = None if fix is None else np.ascontiguousarray(_f64(fix).reshape(self.n, 3))
        self.mask = None if fixmask is None else np.ascontiguousarray(fixmask, dtype=np.uint8)
        self.forces = list(forces)
        self.alpha, self.alpha_min, self.alpha_decay = alpha, alpha_min, alpha_decay
        self.alpha_target, self.velocity_decay = alpha_target, velocity_decay
        self.seed, self.max_ticks = seed, max_ticks

    def desc(self):
        """(fg_job_desc, objects that must stay alive while it is used)."""
        descs, keep = [], [self]
        for f in self.forces:
            d, k = f.desc()
            descs.append(d)
            keep.extend(k)
        arr = (ForceDesc * max(1, len(descs)))(*descs)
        keep.append(arr)
        j = JobDesc(n=self.n, ndim=self.ndim, pos=ptr(self.pos3), vel=ptr(self.vel3), fix=ptr(self.fix3),
                    fixmask=ptr(self.mask, C.c_uint8), nforces=len(descs), forces=arr, alpha=self.alpha,
                    alpha_min=self.alpha_min, alpha_decay=self.alpha_decay, alpha_target=self.alpha_target,
                    velocity_decay=self.velocity_decay, seed=self.seed & 0xffffffff, max_ticks=self.max_ticks)
        return j, keep

    def bytes(self):
        j, keep = self.desc()
        out = C.c_uint64(0)
        rc = lib().fg_batch_job_bytes(C.byref(j), C.byref(out))
        if rc != FG_OK:
            raise FgError(rc, "fg_batch_job_bytes failed")
        return out.value


class Batch:
    """Many layouts resident on one GPU, ticked together (fg_batch_*)."""

    def __init__(self, device=0, budget=0):
        h = C.c_void_p()
        rc = lib().fg_batch_create(C.byref(h), device, budget)
        if rc != FG_OK:
            raise FgError(rc, "fg_batch_create failed")
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            lib().fg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != FG_OK:
            raise FgError(rc, lib().fg_batch_last_error(self.h).decode())

    def submit(self, job):
        """The job's id; raises FgError with status -3 when the budget has no room."""
        j, keep = job.desc()
        out = C.c_uint64(0)
        self._check(lib().fg_batch_submit(self.h, C.byref(j), C.byref(out)))
        self.sizes[out.value] = job.n
        return out.value

    def try_submit(self, job):
        """The job's id, or None when the budget has no room now."""
        try:
            return self.submit(job)
        except FgError as e:
            if e.status == -3:
                return None
            raise

    def step(self, ticks=1):
        self._check(lib().fg_batch_step(self.h, ticks))

    def sync(self):
        self._check(lib().fg_batch_sync(self.h))

    def jobs(self):
        cap = 4096
        arr = (JobInfo * cap)()
        k = C.c_uint32(0)
        self._check(lib().fg_batch_jobs(self.h, arr, cap, C.byref(k)))
        return [dict(id=arr[i].id, status=arr[i].status, ticks=arr[i].ticks, alpha=arr[i].alpha, n=arr[i].n,
                     ndim=arr[i].ndim) for i in range(min(k.value, cap))]

    def read(self, job_id):
        n = self.sizes[job_id]
        pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
        self._check(lib().fg_batch_read(self.h, job_id, ptr(pos), ptr(vel)))
        return pos, vel

    def release(self, job_id):
        self._check(lib().fg_batch_release(self.h, job_id))
        self.sizes.pop(job_id, None)

    def set_tune(self, radix_bits=8, walk_waves=4, walk_stack=48, walk_blocks_per_cu=0):
        t = BatchTune(radix_bits, walk_waves, walk_stack, walk_blocks_per_cu)
        self._check(lib().fg_batch_set_tune(self.h, C.byref(t)))

    def tune(self):
        t = BatchTune()
        self._check(lib().fg_batch_get_tune(self.h, C.byref(t)))
        return dict(radix_bits=t.radix_bits, walk_waves=t.walk_waves, walk_stack=t.walk_stack,
                    walk_blocks_per_cu=t.walk_blocks_per_cu)

    def autotune(self, cache_path=None):
        t = BatchTune()
        path = cache_path.encode() if cache_path else None
        self._check(lib().fg_batch_autotune(self.h, path, C.byref(t)))
        return dict(radix_bits=t.radix_bits, walk_waves=t.walk_waves, walk_stack=t.walk_stack,
                    walk_blocks_per_cu=t.walk_blocks_per_cu)

    def use_graphs(self, on=True):
        self._check(lib().fg_batch_use_graphs(self.h, 1 if on else 0))

    def stats(self):
        s = BatchStats()
        self._check(lib().fg_batch_get_stats(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in BatchStats._fields_}


class Server:
    """A request loop over a batch (fg_serve_*): submit layouts at any time, wait for them."""

    def __init__(self, device=0, budget=0, ticks_per_round=10, autotune=False, tune_cache=None):
        self._cache = tune_cache.encode() if tune_cache else None
        cfg = ServeConfig(budget, ticks_per_round, 1 if autotune else 0, self._cache)
        h = C.c_void_p()
        rc = lib().fg_serve_create(C.byref(h), device, C.byref(cfg))
        if rc != FG_OK:
            raise FgError(rc, "fg_serve_create failed")
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            lib().fg_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != FG_OK:
            raise FgError(rc, lib().fg_serve_last_error(self.h).decode())

    def submit(self, job):
        j, keep = job.desc()
        out = C.c_uint64(0)
        self._check(lib().fg_serve_submit(self.h, C.byref(j), C.byref(out)))
        self.sizes[out.value] = job.n
        return out.value

    def wait(self, job_id, timeout=600.0):
        """(positions, velocities, result) of a finished job."""
        n = self.sizes[job_id]
        pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
        r = ServeResult()
        self._check(lib().fg_serve_wait(self.h, job_id, timeout, ptr(pos), ptr(vel), C.byref(r)))
        self.sizes.pop(job_id, None)
        return pos, vel, {name: getattr(r, name) for name, _ in ServeResult._fields_}

    def drain(self, timeout=600.0):
        self._check(lib().fg_serve_drain(self.h, timeout))

    def stats(self):
        st = ServeStats()
        self._check(lib().fg_serve_get_stats(self.h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in ServeStats._fields_}
