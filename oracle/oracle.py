"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_u = C.POINTER(C.c_uint32)
_i64 = C.POINTER(C.c_int64)
_u64 = C.POINTER(C.c_uint64)
I = C.c_int32


class ForceDesc(C.Structure):
    """Mirror of fg_force_desc (include/d3fgpu.h)."""
    _fields_ = [
        ("type", C.c_int32), ("iterations", C.c_int32), ("id", C.c_uint64), ("version", C.c_uint64),
        ("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("strength", C.c_double),
        ("theta2", C.c_double), ("distance_min2", C.c_double), ("distance_max2", C.c_double),
        ("axis", C.c_int32), ("count", C.c_uint32), ("links", _u), ("strengths", _d), ("radii", _d),
        ("distances", _d), ("bias", _d), ("targets", _d),
    ]


def _declare(L):
    L.or_lcg.restype = C.c_double
    L.or_lcg.argtypes = [_u64]
    L.or_gpu_jiggle.restype = C.c_float
    L.or_gpu_jiggle.argtypes = [C.c_uint32] * 5
    L.or_phyllotaxis.argtypes = [I, I, _d, _d, _d]
    L.or_make_graph.restype = I
    L.or_make_graph.argtypes = [I, C.c_uint32, _u, _u]
    L.or_manybody.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, C.c_double, C.c_double, C.c_double,
                              C.c_double, _u64, _i64]
    L.or_manybody_exact.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, C.c_double, C.c_double, C.c_double]
    L.or_link_seq.argtypes = [I, I, _d, _d, _d, _d, _d, _d, I, _u, _u, _d, _d, _d, I, C.c_double, _u64]
    L.or_link_jacobi.argtypes = [I, I, _d, _d, _d, _d, _d, _d, I, _u, _u, _d, _d, _d, I, C.c_double]
    L.or_link_defaults.argtypes = [I, I, _u, _u, _d, _d, _d]
    L.or_collide_seq.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, C.c_double, I, _u64]
    L.or_collide_jacobi.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, C.c_double, I]
    L.or_center.argtypes = [I, I, _d, _d, _d, C.c_double, C.c_double, C.c_double, C.c_double]
    L.or_radial.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, _d, C.c_double, C.c_double, C.c_double,
                            C.c_double]
    L.or_position.argtypes = [I, _d, _d, _d, _d, C.c_double]
    L.or_integrate.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, _d, _d, C.c_double]
    L.or_sim_create.restype = C.c_void_p
    L.or_sim_create.argtypes = [I, I, _d, _d, _d, _d, _d, _d, _d, _d, _d]
    L.or_sim_destroy.argtypes = [C.c_void_p]
    L.or_sim_set_forces.restype = I
    L.or_sim_set_forces.argtypes = [C.c_void_p, I, C.POINTER(ForceDesc), I]
    L.or_sim_set_params.argtypes = [C.c_void_p] + [C.c_double] * 5
    L.or_sim_tick.argtypes = [C.c_void_p, I]
    L.or_sim_apply_forces.argtypes = [C.c_void_p, C.c_double]
    L.or_sim_alpha.restype = C.c_double
    L.or_sim_alpha.argtypes = [C.c_void_p]
    L.or_sim_get.argtypes = [C.c_void_p, _d, _d, _d, _d, _d, _d]
    L.or_sim_set_state.argtypes = [C.c_void_p, _d, _d, _d, _d, _d, _d]
    L.or_sim_stat.restype = C.c_int64
    L.or_sim_stat.argtypes = [C.c_void_p, I]
    L.or_gt_cube.restype = I
    L.or_gt_cube.argtypes = [I, I, _f, _f, _f, _f]
    L.or_gt_keys.argtypes = [I, I, _f, _f, _f, _f, I, _u, _u]
    L.or_gt_order.argtypes = [I, _u, _u, _u]
    L.or_gt_build.argtypes = [I, I, _u, _u, _u, _u, _u, _u]
    L.or_gt_collapse.argtypes = [I, I, _u, _u, _u, _u]
    L.or_gt_cells.argtypes = [I, I, _u, _d, _d, _d, _d, _u, _u, _u, _u, _d, _d]


def P(a, t=C.c_double):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def dv(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- inputs --------------------------------------------------------------------------

def lcg_stream(count, seed=1):
    s = C.c_uint64(seed)
    return np.array([lib().or_lcg(C.byref(s)) for _ in range(count)])


def phyllotaxis(n, ndim):
    x, y, z = (np.zeros(n) for _ in range(3))
    lib().or_phyllotaxis(n, ndim, P(x), P(y), P(z))
    return x, y, z


def make_graph(n, seed=42):
    mmax = n - 1 + (n + 1) // 2 + 1
    src = np.zeros(max(1, mmax), np.uint32)
    dst = np.zeros(max(1, mmax), np.uint32)
    m = lib().or_make_graph(n, seed, P(src, C.c_uint32), P(dst, C.c_uint32))
    return src[:m].copy(), dst[:m].copy()


def link_defaults(n, src, dst):
    m = len(src)
    bias, strengths, distances = np.zeros(m), np.zeros(m), np.zeros(m)
    lib().or_link_defaults(n, m, P(src, C.c_uint32), P(dst, C.c_uint32), P(bias), P(strengths), P(distances))
    return bias, strengths, distances


# ---- single forces (arrays updated in place) -------------------------------------------

def manybody(pos, vel, strengths, alpha, theta2=0.81, dmin2=1.0, dmax2=np.inf, ndim=3, rng=1):
    x, y, z = (dv(pos[:, a]) for a in range(3))
    vx, vy, vz = (dv(vel[:, a]) for a in range(3))
    st = np.zeros(2, np.int64)
    s = C.c_uint64(rng)
    lib().or_manybody(len(x), ndim, P(x), P(y), P(z), P(vx), P(vy), P(vz), P(dv(strengths)), alpha,
                      theta2, dmin2, dmax2, C.byref(s), P(st, C.c_int64))
    return np.stack([vx, vy, vz], 1), st


def manybody_exact(pos, vel, strengths, alpha, dmin2=1.0, dmax2=np.inf, ndim=3):
    x, y, z = (dv(pos[:, a]) for a in range(3))
    vx, vy, vz = (dv(vel[:, a]) for a in range(3))
    lib().or_manybody_exact(len(x), ndim, P(x), P(y), P(z), P(vx), P(vy), P(vz), P(dv(strengths)), alpha,
                            dmin2, dmax2)
    return np.stack([vx, vy, vz], 1)


# ---- the reference GPU tree -----------------------------------------------------------

def gt_tree(pos32, ndim, strengths=None):
    """Cube, codes, order, topology and kids of the reference GPU tree over f32 positions."""
    n = len(pos32)
    f = [np.ascontiguousarray(pos32[:, a], dtype=np.float32) for a in range(3)]
    cube = np.zeros(4, np.float32)
    L = lib()
    shift = L.or_gt_cube(n, ndim, P(f[0], C.c_float), P(f[1], C.c_float), P(f[2], C.c_float), P(cube, C.c_float))
    hi, lo = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    L.or_gt_keys(n, ndim, P(f[0], C.c_float), P(f[1], C.c_float), P(f[2], C.c_float), P(cube, C.c_float),
                 shift, P(hi, C.c_uint32), P(lo, C.c_uint32))
    order = np.zeros(n, np.uint32)
    L.or_gt_order(n, P(hi, C.c_uint32), P(lo, C.c_uint32), P(order, C.c_uint32))
    shi, slo = hi[order].copy(), lo[order].copy()
    inner = max(1, n - 1)
    left, right, depth = (np.zeros(inner, np.uint32) for _ in range(3))
    leaf_depth = np.zeros(n, np.uint32)
    kids = np.zeros(8 * inner, np.uint32)
    L.or_gt_build(n, ndim, P(shi, C.c_uint32), P(slo, C.c_uint32), P(left, C.c_uint32), P(right, C.c_uint32),
                  P(depth, C.c_uint32), P(leaf_depth, C.c_uint32))
    L.or_gt_collapse(n, ndim, P(left, C.c_uint32), P(right, C.c_uint32), P(depth, C.c_uint32), P(kids, C.c_uint32))
    out = dict(cube=cube, shift=shift, hi=hi, lo=lo, order=order, left=left, right=right, depth=depth,
               leaf_depth=leaf_depth, kids=kids.reshape(inner, 8))
    if strengths is not None and n > 1:
        com = np.zeros(3 * inner)
        value = np.zeros(inner)
        p64 = [dv(pos32[:, a].astype(np.float64)) for a in range(3)]
        L.or_gt_cells(n, ndim, P(order, C.c_uint32), P(p64[0]), P(p64[1]), P(p64[2]), P(dv(strengths)),
                      P(left, C.c_uint32), P(right, C.c_uint32), P(depth, C.c_uint32),
                      P(leaf_depth, C.c_uint32), P(com), P(value))
        out["com"] = com.reshape(inner, 3)
        out["value"] = value
    return out


# ---- whole simulation -----------------------------------------------------------------

class Sim:
    """FP64 d3-force-3d simulation (or, with jacobi=True, the GPU's Jacobi link/collide)."""

    def __init__(self, pos, vel=None, fix=None, ndim=3):
        pos = np.asarray(pos, np.float64)
        n = len(pos)
        self.n, self.ndim = n, ndim
        vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, np.float64)
        cols = [dv(pos[:, a]) for a in range(3)] + [dv(vel[:, a]) for a in range(3)]
        fixc = [None] * 3 if fix is None else [dv(fix[:, a]) for a in range(3)]
        self._h = C.c_void_p(lib().or_sim_create(n, ndim, *[P(c) for c in cols], *[P(c) for c in fixc]))
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_sim_destroy(self._h)
            self._h = None

    def set_forces(self, descs, jacobi=False):
        arr = (ForceDesc * len(descs))(*descs)
        rc = lib().or_sim_set_forces(self._h, len(descs), arr, 1 if jacobi else 0)
        if rc != 0:
            raise ValueError("bad force list")

    def set_params(self, alpha=1.0, alpha_min=0.001, alpha_decay=None, alpha_target=0.0, decay=0.6):
        if alpha_decay is None:
            alpha_decay = 1 - 0.001 ** (1 / 300)
        lib().or_sim_set_params(self._h, alpha, alpha_min, alpha_decay, alpha_target, decay)

    def tick(self, k=1):
        lib().or_sim_tick(self._h, k)

    def apply_forces(self, alpha):
        lib().or_sim_apply_forces(self._h, alpha)

    @property
    def alpha(self):
        return lib().or_sim_alpha(self._h)

    def state(self):
        cols = [np.zeros(self.n) for _ in range(6)]
        lib().or_sim_get(self._h, *[P(c) for c in cols])
        return np.stack(cols[:3], 1), np.stack(cols[3:], 1)

    def set_state(self, pos, vel):
        cols = [dv(pos[:, a]) for a in range(3)] + [dv(vel[:, a]) for a in range(3)]
        lib().or_sim_set_state(self._h, *[P(c) for c in cols])

    def stat(self, which):
        return lib().or_sim_stat(self._h, which)
