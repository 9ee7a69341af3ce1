"""ctypes binding of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built from ``smoke_oracle.cpp``).  Only the test
suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
use it, and only as the checker.  Arrays are numpy arrays in the reference's
layouts: density/mask shaped (nz, ny, nx), velocity components shaped
(nz+1, ny+1, nx+1), x fastest (project/smokeSimulation.cu:146-147, :267).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

PROJ_REFERENCE, PROJ_RBSOR, PROJ_JACOBI = 0, 1, 2
OBST_LAST_WINS, OBST_UNION = 0, 1
STAGE_SCENE, STAGE_FORCE, STAGE_PROJECT, STAGE_ADVECT_VEL, STAGE_ADVECT_DENS = range(5)

F_SELF, F_XM, F_XP, F_YM, F_YP, F_ZM, F_ZP, F_ACTIVE = 1, 2, 4, 8, 16, 32, 64, 128


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("gravity", ctypes.c_float),
        ("buoyancy", ctypes.c_float),
        ("density_gate", ctypes.c_float),
        ("max_velocity", ctypes.c_float),
        ("sor_omega", ctypes.c_double),
        ("jacobi_omega", ctypes.c_double),
        ("projection", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("obstacle_mode", ctypes.c_int),
        ("warm_start", ctypes.c_int),
        ("vorticity", ctypes.c_float),
        ("obstacle_velocity", ctypes.c_int),
    ]


DEFAULT_PARAMS = dict(
    gravity=-9.82, buoyancy=2.0, density_gate=0.0, max_velocity=3.0,
    sor_omega=1.9, jacobi_omega=1.0, projection=PROJ_REFERENCE, iterations=30,
    obstacle_mode=OBST_LAST_WINS, warm_start=0, vorticity=0.0, obstacle_velocity=0,
)

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run make -C oracle)")
    lib = ctypes.CDLL(_LIB_PATH)
    vp, i, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    sig = {
        "orc_create": (vp, [i, i, i, i]),
        "orc_destroy": (None, [vp]),
        "orc_set_threads": (None, [i]),
        "orc_max_threads": (i, []),
        "orc_set_params": (None, [vp, ctypes.POINTER(OrcParams)]),
        "orc_add_object": (i, [vp, i, f, f, f, f, f, f, f]),
        "orc_move_object": (i, [vp, i, f, f, f]),
        "orc_set_object_velocity": (i, [vp, i, f, f, f]),
        "orc_set_solids": (None, [vp, u8p]),
        "orc_rebuild_mask": (None, [vp]),
        "orc_set_density": (None, [vp, i, dp]),
        "orc_get_density": (None, [vp, i, dp]),
        "orc_set_velocity": (None, [vp, i, i, dp]),
        "orc_get_velocity": (None, [vp, i, i, dp]),
        "orc_get_pressure": (None, [vp, dp]),
        "orc_get_rhs": (None, [vp, dp]),
        "orc_get_mask": (None, [vp, u8p]),
        "orc_get_flags": (None, [vp, u8p]),
        "orc_index_now": (i, [vp]),
        "orc_stage": (None, [vp, i, f]),
        "orc_step": (None, [vp, f]),
        "orc_steps": (None, [vp, f, i]),
        "orc_plane_box": (i, [fp, fp, fp, fp, i, fp]),
        "orc_proxy_geometry": (i, [fp, fp, fp, fp, i, fp, ctypes.POINTER(ctypes.c_int16),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u8ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def set_threads(n):
    load().orc_set_threads(int(n))


def max_threads():
    return load().orc_max_threads()


class Oracle:
    """One simulation of the oracle, FP64 (real_bits=64) or FP32 (32)."""

    def __init__(self, nx, ny, nz, real_bits=64, params=None):
        self.lib = load()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.real_bits = real_bits
        self.h = self.lib.orc_create(real_bits, self.nx, self.ny, self.nz)
        if not self.h:
            raise ValueError("oracle: invalid dims or precision")
        self.params = dict(DEFAULT_PARAMS)
        if params:
            self.set_params(**params)
        else:
            self._push_params()

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.orc_destroy(h)
            self.h = None

    @property
    def cell_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def face_shape(self):
        return (self.nz + 1, self.ny + 1, self.nx + 1)

    def _push_params(self):
        p = OrcParams(**{k: self.params[k] for k, _ in OrcParams._fields_})
        self.lib.orc_set_params(self.h, ctypes.byref(p))

    def set_params(self, **kw):
        for k in kw:
            if k not in self.params:
                raise KeyError(k)
        self.params.update(kw)
        self._push_params()

    def add_obstacle(self, x, y, z, r, vel=(0.0, 0.0, 0.0)):
        return self.lib.orc_add_object(self.h, 0, x, y, z, vel[0], vel[1], vel[2], r)

    def add_source(self, x, y, z, r):
        return self.lib.orc_add_object(self.h, 1, x, y, z, 0.0, 0.0, 0.0, r)

    def move_object(self, oid, x, y, z):
        if self.lib.orc_move_object(self.h, oid, x, y, z) != 0:
            raise IndexError(oid)

    def set_object_velocity(self, oid, vx, vy, vz):
        if self.lib.orc_set_object_velocity(self.h, oid, vx, vy, vz) != 0:
            raise IndexError(oid)

    def set_solids(self, solid):
        a = np.ascontiguousarray(solid, dtype=np.uint8).reshape(self.cell_shape)
        self.lib.orc_set_solids(self.h, _u8ptr(a))

    def set_density(self, half, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.cell_shape)
        self.lib.orc_set_density(self.h, half, _dptr(a))

    def get_density(self, half):
        a = np.empty(self.cell_shape, dtype=np.float64)
        self.lib.orc_get_density(self.h, half, _dptr(a))
        return a

    def set_velocity(self, comp, half, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.face_shape)
        self.lib.orc_set_velocity(self.h, comp, half, _dptr(a))

    def get_velocity(self, comp, half):
        a = np.empty(self.face_shape, dtype=np.float64)
        self.lib.orc_get_velocity(self.h, comp, half, _dptr(a))
        return a

    def get_pressure(self):
        a = np.empty(self.cell_shape, dtype=np.float64)
        self.lib.orc_get_pressure(self.h, _dptr(a))
        return a

    def get_rhs(self):
        a = np.empty(self.cell_shape, dtype=np.float64)
        self.lib.orc_get_rhs(self.h, _dptr(a))
        return a

    def get_mask(self):
        a = np.empty(self.cell_shape, dtype=np.uint8)
        self.lib.orc_get_mask(self.h, _u8ptr(a))
        return a

    def get_flags(self):
        a = np.empty(self.cell_shape, dtype=np.uint8)
        self.lib.orc_get_flags(self.h, _u8ptr(a))
        return a

    def rebuild_mask(self):
        self.lib.orc_rebuild_mask(self.h)

    @property
    def index_now(self):
        return self.lib.orc_index_now(self.h)

    def stage(self, stage, dt):
        self.lib.orc_stage(self.h, int(stage), float(dt))

    def step(self, dt, n=1):
        self.lib.orc_steps(self.h, float(dt), int(n))

    def output_density(self):
        """Density returned by the last step: half [1 - index_now]."""
        return self.get_density(1 - self.index_now)

    def state(self):
        """Every field of both halves, plus mask and flags."""
        out = {"mask": self.get_mask(), "flags": self.get_flags(), "now": self.index_now}
        for h in (0, 1):
            out[f"rho{h}"] = self.get_density(h)
            for k, name in enumerate("uvw"):
                out[f"{name}{h}"] = self.get_velocity(k, h)
        return out


# ------------------------------------------------------------------ step norms

def divergence_f32(u, v, w):
    """Per-cell divergence of face fields (nz+1, ny+1, nx+1) in FP32 with the
    right-hand side's operand order, -u0 + u1 + -v0 + v1 + -w0 + w1, left to
    right (project/smokeSimulation.cu:379-381); shape (nz, ny, nx)."""
    f = np.float32
    u = np.asarray(u, dtype=f)
    v = np.asarray(v, dtype=f)
    w = np.asarray(w, dtype=f)
    u0, u1 = u[:-1, :-1, :-1], u[:-1, :-1, 1:]
    v0, v1 = v[:-1, :-1, :-1], v[:-1, 1:, :-1]
    w0, w1 = w[:-1, :-1, :-1], w[1:, :-1, :-1]
    return ((((-u0 + u1) + -v0) + v1) + -w0) + w1


def step_norms(sim, dt, rhs=None):
    """Norms of a state (the reduction smoke_get_step_norms performs), from the
    fields of `sim` (the oracle or a GPU simulation: get_density,
    get_velocity, get_flags, index_now): divergence of half [now] and `rhs`
    (the projection's right-hand side, or None) over relaxed cells, summed in
    FP64; max |u|, |v|, |w| over the faces of the output half [1 - now] that
    stages write (x >= 1 for u, y >= 1 for v, z >= 1 for w, within the cell
    range); density total and max; non-finite count."""
    now = sim.index_now
    out_h = 1 - now
    flags = sim.get_flags()
    act = (flags & F_ACTIVE) != 0
    vel_now = [sim.get_velocity(k, now) for k in range(3)]
    div = divergence_f32(*vel_now).astype(np.float64)
    rho = np.asarray(sim.get_density(out_h), dtype=np.float32)
    uo, vo, wo = (np.asarray(sim.get_velocity(k, out_h), dtype=np.float32) for k in range(3))
    faces = [uo[:-1, :-1, 1:-1], vo[:-1, 1:-1, :-1], wo[1:-1, :-1, :-1]]
    res = {"div_l2": float(np.sqrt(np.sum(div[act] ** 2))),
           "rhs_l2": -1.0 if rhs is None else float(np.sqrt(np.sum(
               np.asarray(rhs, dtype=np.float64)[act] ** 2))),
           "density_sum": float(np.sum(rho.astype(np.float64)))}
    m = np.abs(rho[~np.isnan(rho)])
    res["max_density"] = float(m.max()) if m.size else 0.0
    bad = int(np.count_nonzero(~np.isfinite(rho)))
    for name, a in zip(("max_u", "max_v", "max_w"), faces):
        m = np.abs(a[~np.isnan(a)])
        res[name] = float(m.max()) if m.size else 0.0
        bad += int(np.count_nonzero(~np.isfinite(a)))
    res["nonfinite"] = bad
    res["cfl"] = float(np.float64(np.float32(dt)) * max(res["max_u"], res["max_v"], res["max_w"]))
    return res


# ------------------------------------------------ proxy-plane slicing geometry

PROXY_MAX_VERTS_PER_PLANE = 13
PROXY_MAX_INDICES_PER_PLANE = 36


def _f3(v):
    a = np.ascontiguousarray(v, dtype=np.float32).reshape(3)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def plane_box(box_pos, box_size, plane_pos, plane_normal, sort=False):
    """planeBoxIntersections (and optionally sortPoints): float32 (n, 3), the
    mean last."""
    lib = load()
    keep = [_f3(v) for v in (box_pos, box_size, plane_pos, plane_normal)]
    out = np.empty((13, 3), np.float32)
    n = lib.orc_plane_box(*(k[1] for k in keep), int(bool(sort)),
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out[:n].copy()


def proxy_geometry(box_pos, box_dims, view_origin, view_dir, steps):
    """generateProxyGeometry: (vertices float32 (n, 3), indices int16 (m,),
    planes int32 (steps, 2) of (num_triangles, triangle_offset))."""
    lib = load()
    keep = [_f3(v) for v in (box_pos, box_dims, view_origin, view_dir)]
    steps = int(steps)
    verts = np.empty((max(steps, 1) * PROXY_MAX_VERTS_PER_PLANE, 3), np.float32)
    idx = np.empty(max(steps, 1) * PROXY_MAX_INDICES_PER_PLANE, np.int16)
    planes = np.empty((max(steps, 1), 2), np.int32)
    nv, ni = ctypes.c_int(0), ctypes.c_int(0)
    rc = lib.orc_proxy_geometry(*(k[1] for k in keep), steps,
                                verts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                                planes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                ctypes.byref(nv), ctypes.byref(ni))
    if rc != 0:
        raise ValueError("proxy_geometry: steps must be >= 1")
    return verts[:nv.value].copy(), idx[:ni.value].copy(), planes[:steps].copy()
