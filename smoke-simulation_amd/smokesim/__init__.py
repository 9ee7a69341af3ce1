"""Python binding of libsmokesim.so, the MI355X smoke time step.

A thin ctypes layer over the C ABI declared in ``include/smoke.h``; every
method maps to one entry point.  Arrays are numpy arrays in the reference's
layouts (project/smokeSimulation.cu:146-147, :267): density, masks and
pressure shaped (nz, ny, nx); velocity components shaped (nz+1, ny+1, nx+1);
x fastest.

The library is required: importing this module on a machine where
``libsmokesim.so`` is missing raises, and nothing here falls back to CPU code.
"""
import ctypes
import os

import numpy as np

from . import scenes  # noqa: F401  (re-exported helper module)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmokesim.so")

SMOKE_OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_NONFINITE = 1, 2, 3, 4, 5
PROJ_REFERENCE, PROJ_RBSOR, PROJ_JACOBI, PROJ_REFINED = 0, 1, 2, 3
STORAGE_FP32, STORAGE_BF16, STORAGE_BF16_VELOCITY, STORAGE_FP16, STORAGE_FP16_VELOCITY = 0, 1, 2, 3, 4
OBSTACLES_LAST_WINS, OBSTACLES_UNION = 0, 1
STAGE_SCENE, STAGE_FORCE, STAGE_PROJECT, STAGE_ADVECT_VEL, STAGE_ADVECT_DENS = range(5)
STAGE_COUNT = 5
STAGE_NAMES = ("scene", "force", "project", "advect_velocity", "advect_density")
MAX_OBJECTS = 64
READBACK_F32, READBACK_F16, READBACK_BF16 = 0, 1, 2

F_SELF, F_XM, F_XP, F_YM, F_YP, F_ZM, F_ZP, F_ACTIVE = 1, 2, 4, 8, 16, 32, 64, 128


class SmokeError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"smoke status {status}: {msg}")
        self.status = status


class Params(ctypes.Structure):
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
        ("inner_sweeps", ctypes.c_int),
        ("vorticity", ctypes.c_float),
        ("tolerance", ctypes.c_double),
        ("obstacle_velocity", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JacobiPlan(ctypes.Structure):
    _fields_ = [("sweeps_per_launch", ctypes.c_int), ("tile", ctypes.c_int), ("zc", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TILE_32x32, TILE_64x16, TILE_64x32, TILE_32x64 = 0, 1, 2, 3


class StepNorms(ctypes.Structure):
    _fields_ = [
        ("div_l2", ctypes.c_double), ("rhs_l2", ctypes.c_double),
        ("density_sum", ctypes.c_double), ("cfl", ctypes.c_double),
        ("max_u", ctypes.c_float), ("max_v", ctypes.c_float), ("max_w", ctypes.c_float),
        ("max_density", ctypes.c_float),
        ("nonfinite", ctypes.c_longlong), ("step", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ProxyPlane(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_int), ("triangle_offset", ctypes.c_int)]


PROXY_MAX_STEPS = 4096
PROXY_VERTICES_PER_PLANE = 13
PROXY_INDICES_PER_PLANE = 36

_lib = None


def _declare(lib):
    vp = ctypes.c_void_p
    st = ctypes.c_int
    i, u, f = ctypes.c_int, ctypes.c_uint, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    dp = ctypes.POINTER(ctypes.c_double)
    llp = ctypes.POINTER(ctypes.c_longlong)
    ip = ctypes.POINTER(ctypes.c_int)
    up = ctypes.POINTER(ctypes.c_uint)
    pp = ctypes.POINTER(Params)
    sig = {
        "smoke_last_error": (ctypes.c_char_p, []),
        "smoke_version": (ctypes.c_char_p, []),
        "smoke_device_info": (st, [ctypes.c_char_p, ctypes.c_size_t]),
        "smoke_params_default": (st, [pp]),
        "smoke_create": (st, [u, u, u, fp, ctypes.POINTER(vp)]),
        "smoke_create_ex": (st, [u, u, u, fp, i, ctypes.POINTER(vp)]),
        "smoke_get_storage": (i, [vp]),
        "smoke_destroy": (None, [vp]),
        "smoke_get_dims": (st, [vp, up, up, up]),
        "smoke_footprint_bytes": (ctypes.c_size_t, [u, u, u]),
        "smoke_footprint_bytes_ex": (ctypes.c_size_t, [u, u, u, i, i]),
        "smoke_get_solve_info": (st, [vp, dp, ip]),
        "smoke_snapshot_bytes": (ctypes.c_size_t, [vp]),
        "smoke_snapshot": (st, [vp, vp, ctypes.c_size_t]),
        "smoke_restore": (st, [vp, vp, ctypes.c_size_t]),
        "smoke_set_params": (st, [vp, pp]),
        "smoke_get_params": (st, [vp, pp]),
        "smoke_set_jacobi_plan": (st, [vp, ctypes.POINTER(JacobiPlan)]),
        "smoke_get_jacobi_plan": (st, [vp, ctypes.POINTER(JacobiPlan)]),
        "smoke_autotune": (st, [vp, ctypes.POINTER(JacobiPlan), dp]),
        "smoke_add_obstacle": (st, [vp, f, f, f, f, f, f, f, ip]),
        "smoke_add_source": (st, [vp, f, f, f, f, ip]),
        "smoke_move_object": (st, [vp, i, f, f, f]),
        "smoke_set_object_velocity": (st, [vp, i, f, f, f]),
        "smoke_set_solids": (st, [vp, u8p]),
        "smoke_set_density": (st, [vp, i, fp]),
        "smoke_get_density": (st, [vp, i, fp]),
        "smoke_set_velocity": (st, [vp, i, i, fp]),
        "smoke_get_velocity": (st, [vp, i, i, fp]),
        "smoke_get_mask": (st, [vp, u8p]),
        "smoke_get_flags": (st, [vp, u8p]),
        "smoke_get_pressure": (st, [vp, fp]),
        "smoke_index_now": (i, [vp]),
        "smoke_step": (st, [vp, f, fp]),
        "smoke_step_async": (st, [vp, f]),
        "smoke_set_step_graph": (st, [vp, ctypes.c_int]),
        "smoke_run_stage": (st, [vp, i, f]),
        "smoke_read_density": (st, [vp, fp]),
        "smoke_synchronize": (st, [vp]),
        "smoke_readback_begin": (st, [vp, i]),
        "smoke_readback_wait": (st, [vp, ctypes.POINTER(vp), llp]),
        "smoke_timing_enable": (st, [vp, i]),
        "smoke_timing_read": (st, [vp, dp, llp, llp]),
        "smoke_timing_reset": (st, [vp]),
        "smoke_timing_step_ms": (st, [vp, dp, i, ip]),
        "smoke_timing_sweeps": (st, [vp, dp, llp]),
        "smoke_norms_enable": (st, [vp, i]),
        "smoke_get_step_norms": (st, [vp, ctypes.POINTER(StepNorms)]),
        "smoke_get_step_norms_noalloc": (st, [vp, vp]),
        "smoke_batch_create": (st, [ctypes.POINTER(vp), i, fp, ctypes.POINTER(vp)]),
        "smoke_batch_destroy": (None, [vp]),
        "smoke_batch_step": (st, [vp, i]),
        "smoke_batch_synchronize": (st, [vp]),
        "smoke_batch_set_jacobi_plan": (st, [vp, ctypes.POINTER(JacobiPlan)]),
        "smoke_batch_get_jacobi_plan": (st, [vp, ctypes.POINTER(JacobiPlan)]),
        "smoke_batch_autotune": (st, [vp, ctypes.POINTER(JacobiPlan), dp]),
        "smoke_batch_step_ms": (st, [vp, dp, i, ip]),
        "smoke_batch_cells": (ctypes.c_longlong, [vp]),
        "smoke_batch_interior_cells": (ctypes.c_longlong, [vp]),
        "smoke_batch_time_sweeps": (st, [vp, i, dp, ip]),
        "smoke_server_create": (st, [ctypes.c_size_t, i, ctypes.POINTER(vp)]),
        "smoke_server_destroy": (None, [vp]),
        "smoke_server_submit": (st, [vp, ctypes.POINTER(Request), llp]),
        "smoke_server_tick": (st, [vp, ip, ip]),
        "smoke_server_result": (st, [vp, ctypes.c_longlong, ip, fp, dp, dp]),
        "smoke_server_stats_get": (st, [vp, ctypes.POINTER(ServerStats)]),
        "smoke_server_tick_ms": (st, [vp, dp, ip, i, ip]),
        "smoke_proxy_geometry": (st, [fp, fp, fp, fp, i, fp, ctypes.POINTER(ctypes.c_int16),
                                      ctypes.POINTER(ProxyPlane), ip, ip]),
        "smoke_proxy_geometry_device": (st, [fp, fp, fp, fp, i, vp, vp, vp, vp, vp]),
        "smoke_legacy_device_info": (st, []),
        "smoke_legacy_initialize": (st, [fp, u, u, u]),
        "smoke_legacy_delete": (st, []),
        "smoke_legacy_simulate": (st, [fp, f]),
        "smoke_legacy_add_obstacle": (i, [f, f, f, f, f, f, f]),
        "smoke_legacy_add_source": (i, [f, f, f, f]),
        "smoke_legacy_update_object_pos": (st, [i, f, f, f]),
        "smoke_legacy_buoyancy_ptr": (fp, []),
        "smoke_legacy_gravity_ptr": (fp, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load():
    """Load libsmokesim.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libsmokesim.so not found at {LIB_PATH}; build it with "
                "`make -C smoke-simulation_amd` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        _declare(lib)
        _lib = lib
    return _lib


def _check(status):
    if status != SMOKE_OK:
        raise SmokeError(status, load().smoke_last_error().decode())


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u8ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def device_info():
    buf = ctypes.create_string_buffer(512)
    _check(load().smoke_device_info(buf, len(buf)))
    return buf.value.decode()


def default_params():
    p = Params()
    _check(load().smoke_params_default(ctypes.byref(p)))
    return p.as_dict()


def footprint_bytes(nx, ny, nz, storage=STORAGE_FP32, refined=False):
    return int(load().smoke_footprint_bytes_ex(nx, ny, nz, int(storage), 1 if refined else 0))


class Simulation:
    """One simulation on the current device (smoke_create_ex ... smoke_destroy).
    storage: STORAGE_FP32 (the drop-in) or a 16-bit form (STORAGE_FP16,
    STORAGE_BF16; *_VELOCITY keeps the density in FP32)."""

    def __init__(self, nx, ny, nz, density=None, params=None, storage=STORAGE_FP32):
        self.lib = load()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        h = ctypes.c_void_p()
        dptr = None
        if density is not None:
            d = np.ascontiguousarray(density, dtype=np.float32).reshape(self.cell_shape)
            dptr = _fptr(d)
        _check(self.lib.smoke_create_ex(self.nx, self.ny, self.nz, dptr, int(storage),
                                        ctypes.byref(h)))
        self.h = h
        self.storage = int(storage)
        if params:
            self.set_params(**params)

    def solve_info(self):
        """(relative residual, refinements) of the last REFINED solve."""
        rel, n = ctypes.c_double(), ctypes.c_int()
        _check(self.lib.smoke_get_solve_info(self.h, ctypes.byref(rel), ctypes.byref(n)))
        return rel.value, n.value

    def snapshot(self):
        """The whole simulation state as bytes (smoke_snapshot)."""
        n = int(self.lib.smoke_snapshot_bytes(self.h))
        buf = (ctypes.c_char * n)()
        _check(self.lib.smoke_snapshot(self.h, buf, n))
        return bytes(buf)

    def restore(self, blob):
        """Load a snapshot taken from a simulation of the same dims and storage."""
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        _check(self.lib.smoke_restore(self.h, buf, len(blob)))

    def close(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.smoke_destroy(h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def cell_shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def face_shape(self):
        return (self.nz + 1, self.ny + 1, self.nx + 1)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    # params
    def get_params(self):
        p = Params()
        _check(self.lib.smoke_get_params(self.h, ctypes.byref(p)))
        return p.as_dict()

    def set_params(self, **kw):
        cur = self.get_params()
        for k in kw:
            if k not in cur:
                raise KeyError(k)
        cur.update(kw)
        p = Params(**cur)
        _check(self.lib.smoke_set_params(self.h, ctypes.byref(p)))

    def get_jacobi_plan(self):
        p = JacobiPlan()
        _check(self.lib.smoke_get_jacobi_plan(self.h, ctypes.byref(p)))
        return p.as_dict()

    def set_jacobi_plan(self, **kw):
        cur = self.get_jacobi_plan()
        for k in kw:
            if k not in cur:
                raise KeyError(k)
        cur.update(kw)
        p = JacobiPlan(**cur)
        _check(self.lib.smoke_set_jacobi_plan(self.h, ctypes.byref(p)))

    def autotune(self):
        """Pick the fastest fused-Jacobi plan for this grid and iteration
        count; returns (plan dict, ms per sweep)."""
        p = JacobiPlan()
        ms = ctypes.c_double()
        _check(self.lib.smoke_autotune(self.h, ctypes.byref(p), ctypes.byref(ms)))
        return p.as_dict(), ms.value

    # scene
    def add_obstacle(self, x, y, z, r, vel=(0.0, 0.0, 0.0)):
        oid = ctypes.c_int()
        _check(self.lib.smoke_add_obstacle(self.h, x, y, z, vel[0], vel[1], vel[2], r,
                                           ctypes.byref(oid)))
        return oid.value

    def add_source(self, x, y, z, r):
        oid = ctypes.c_int()
        _check(self.lib.smoke_add_source(self.h, x, y, z, r, ctypes.byref(oid)))
        return oid.value

    def move_object(self, oid, x, y, z):
        _check(self.lib.smoke_move_object(self.h, int(oid), x, y, z))

    def set_object_velocity(self, oid, vx, vy, vz):
        _check(self.lib.smoke_set_object_velocity(self.h, int(oid), vx, vy, vz))

    def set_solids(self, solid):
        a = np.ascontiguousarray(solid, dtype=np.uint8).reshape(self.cell_shape)
        _check(self.lib.smoke_set_solids(self.h, _u8ptr(a)))

    # fields
    def set_density(self, half, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.cell_shape)
        _check(self.lib.smoke_set_density(self.h, int(half), _fptr(a)))

    def get_density(self, half):
        a = np.empty(self.cell_shape, dtype=np.float32)
        _check(self.lib.smoke_get_density(self.h, int(half), _fptr(a)))
        return a

    def set_velocity(self, comp, half, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self.face_shape)
        _check(self.lib.smoke_set_velocity(self.h, int(comp), int(half), _fptr(a)))

    def get_velocity(self, comp, half):
        a = np.empty(self.face_shape, dtype=np.float32)
        _check(self.lib.smoke_get_velocity(self.h, int(comp), int(half), _fptr(a)))
        return a

    def get_mask(self):
        a = np.empty(self.cell_shape, dtype=np.uint8)
        _check(self.lib.smoke_get_mask(self.h, _u8ptr(a)))
        return a

    def get_flags(self):
        a = np.empty(self.cell_shape, dtype=np.uint8)
        _check(self.lib.smoke_get_flags(self.h, _u8ptr(a)))
        return a

    def get_pressure(self):
        a = np.empty(self.cell_shape, dtype=np.float32)
        _check(self.lib.smoke_get_pressure(self.h, _fptr(a)))
        return a

    @property
    def index_now(self):
        return self.lib.smoke_index_now(self.h)

    # stepping
    def step(self, dt, readback=True):
        if readback:
            out = np.empty(self.cell_shape, dtype=np.float32)
            _check(self.lib.smoke_step(self.h, float(dt), _fptr(out)))
            return out
        _check(self.lib.smoke_step(self.h, float(dt), None))
        return None

    def step_async(self, dt):
        _check(self.lib.smoke_step_async(self.h, float(dt)))

    def set_step_graph(self, on=True):
        """Capture steps into a HIP graph per parity and replay them."""
        _check(self.lib.smoke_set_step_graph(self.h, 1 if on else 0))

    def run_stage(self, stage, dt):
        _check(self.lib.smoke_run_stage(self.h, int(stage), float(dt)))

    def read_density(self):
        out = np.empty(self.cell_shape, dtype=np.float32)
        _check(self.lib.smoke_read_density(self.h, _fptr(out)))
        return out

    def synchronize(self):
        _check(self.lib.smoke_synchronize(self.h))

    def readback_begin(self, fmt=READBACK_F32):
        _check(self.lib.smoke_readback_begin(self.h, int(fmt)))
        self._rb_formats = getattr(self, "_rb_formats", []) + [int(fmt)]

    def readback_wait(self):
        """(density array copied out of the pinned buffer, step number)."""
        ptr = ctypes.c_void_p()
        step = ctypes.c_longlong()
        _check(self.lib.smoke_readback_wait(self.h, ctypes.byref(ptr), ctypes.byref(step)))
        fmt = self._rb_formats.pop(0)
        n = self.cells
        if fmt == READBACK_F32:
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)), (n,))
        else:
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint16)), (n,))
            if fmt == READBACK_F16:
                a = a.view(np.float16)
        return a.copy().reshape(self.cell_shape), step.value

    # timing
    def timing_enable(self, on=True):
        _check(self.lib.smoke_timing_enable(self.h, 1 if on else 0))

    def timing_reset(self):
        _check(self.lib.smoke_timing_reset(self.h))

    def timing_read(self):
        ms = (ctypes.c_double * STAGE_COUNT)()
        launches = (ctypes.c_longlong * STAGE_COUNT)()
        steps = ctypes.c_longlong()
        _check(self.lib.smoke_timing_read(self.h, ms, launches, ctypes.byref(steps)))
        return {
            "steps": steps.value,
            "stage_ms": {STAGE_NAMES[k]: ms[k] for k in range(STAGE_COUNT)},
            "stage_launches": {STAGE_NAMES[k]: launches[k] for k in range(STAGE_COUNT)},
        }

    def timing_sweeps(self):
        """(total ms, launches) of the projection sweep loops of timed steps."""
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(self.lib.smoke_timing_sweeps(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def timing_step_ms(self, cap=1 << 16):
        buf = (ctypes.c_double * cap)()
        n = ctypes.c_int()
        _check(self.lib.smoke_timing_step_ms(self.h, buf, cap, ctypes.byref(n)))
        return np.array(buf[: n.value], dtype=np.float64)

    def norms_enable(self, on=True):
        """Every following step ends with the norms reduction (in-stream)."""
        _check(self.lib.smoke_norms_enable(self.h, 1 if on else 0))

    def step_norms(self):
        """Norms of the current state (smoke_get_step_norms) as a dict; a
        non-finite field is reported in 'nonfinite' and 'status'
        (ERR_NONFINITE), not raised."""
        n = StepNorms()
        st = self.lib.smoke_get_step_norms(self.h, ctypes.byref(n))
        if st not in (SMOKE_OK, ERR_NONFINITE):
            _check(st)
        out = n.as_dict()
        out["status"] = st
        return out

    def state(self):
        """Every field of both halves, plus mask and flags (blocking)."""
        out = {"mask": self.get_mask(), "flags": self.get_flags(), "now": self.index_now}
        for h in (0, 1):
            out[f"rho{h}"] = self.get_density(h)
            for k, name in enumerate("uvw"):
                out[f"{name}{h}"] = self.get_velocity(k, h)
        return out


class Batch:
    """Simulations advanced together, one launch per stage (smoke_batch_*)."""

    def __init__(self, sims, dts):
        self.lib = load()
        self.sims = list(sims)  # keeps the members alive
        n = len(self.sims)
        handles = (ctypes.c_void_p * n)(*[s.h.value for s in self.sims])
        d = np.ascontiguousarray(dts, dtype=np.float32)
        if d.shape != (n,):
            raise ValueError("one dt per simulation")
        h = ctypes.c_void_p()
        _check(self.lib.smoke_batch_create(handles, n, _fptr(d), ctypes.byref(h)))
        self.h = h

    def close(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.smoke_batch_destroy(h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def cells(self):
        return int(self.lib.smoke_batch_cells(self.h))

    def step(self, use_graph=False):
        _check(self.lib.smoke_batch_step(self.h, 1 if use_graph else 0))

    def synchronize(self):
        _check(self.lib.smoke_batch_synchronize(self.h))

    def get_jacobi_plan(self):
        p = JacobiPlan()
        _check(self.lib.smoke_batch_get_jacobi_plan(self.h, ctypes.byref(p)))
        return p.as_dict()

    def set_jacobi_plan(self, **kw):
        cur = self.get_jacobi_plan()
        cur.update(kw)
        p = JacobiPlan(**cur)
        _check(self.lib.smoke_batch_set_jacobi_plan(self.h, ctypes.byref(p)))

    def autotune(self):
        p = JacobiPlan()
        ms = ctypes.c_double()
        _check(self.lib.smoke_batch_autotune(self.h, ctypes.byref(p), ctypes.byref(ms)))
        return p.as_dict(), ms.value

    @property
    def interior_cells(self):
        return int(self.lib.smoke_batch_interior_cells(self.h))

    def time_sweeps(self, reps=3):
        """(ms per sweep launch, launches per sequence) of the batch's sweeps."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(self.lib.smoke_batch_time_sweeps(self.h, int(reps), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def step_ms(self, cap=1 << 16):
        buf = (ctypes.c_double * cap)()
        n = ctypes.c_int()
        _check(self.lib.smoke_batch_step_ms(self.h, buf, cap, ctypes.byref(n)))
        return np.array(buf[: n.value], dtype=np.float64)


class Request(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint), ("ny", ctypes.c_uint), ("nz", ctypes.c_uint),
        ("steps", ctypes.c_int), ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float), ("buoyancy", ctypes.c_float),
        ("density_gate", ctypes.c_float), ("max_velocity", ctypes.c_float),
        ("solid", ctypes.POINTER(ctypes.c_uint8)),
        ("n_sources", ctypes.c_int), ("sources", ctypes.POINTER(ctypes.c_float)),
        ("n_obstacles", ctypes.c_int), ("obstacles", ctypes.POINTER(ctypes.c_float)),
        ("velocity", ctypes.POINTER(ctypes.c_float) * 3),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong), ("admitted", ctypes.c_longlong),
        ("completed", ctypes.c_longlong), ("ticks", ctypes.c_longlong),
        ("cell_updates", ctypes.c_longlong), ("gpu_ms", ctypes.c_double),
        ("hbm_used", ctypes.c_size_t), ("hbm_budget", ctypes.c_size_t),
        ("hbm_peak", ctypes.c_size_t), ("active", ctypes.c_int), ("queued", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


REQUEST_QUEUED, REQUEST_RUNNING, REQUEST_DONE = 0, 1, 2


class Server:
    """Request loop with continuous batching (smoke_server_*)."""

    def __init__(self, hbm_budget_bytes, iterations):
        self.lib = load()
        h = ctypes.c_void_p()
        _check(self.lib.smoke_server_create(int(hbm_budget_bytes), int(iterations), ctypes.byref(h)))
        self.h = h
        self.dims = {}

    def close(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.smoke_server_destroy(h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, scene, steps):
        """Queue a scene (smokesim.scenes format: dims, solid, sources,
        obstacles, velocity, params, dt) for `steps` steps; returns its id."""
        nx, ny, nz = scene["dims"]
        p = scene["params"]
        keep = []
        q = Request()
        q.nx, q.ny, q.nz, q.steps, q.dt = nx, ny, nz, int(steps), float(scene["dt"])
        q.gravity, q.buoyancy = p["gravity"], p["buoyancy"]
        q.density_gate, q.max_velocity = p["density_gate"], p["max_velocity"]
        if scene.get("solid") is not None:
            a = np.ascontiguousarray(scene["solid"], dtype=np.uint8)
            keep.append(a)
            q.solid = _u8ptr(a)
        for field, cnt in (("sources", "n_sources"), ("obstacles", "n_obstacles")):
            items = scene.get(field) or []
            if items:
                a = np.ascontiguousarray(np.array(items, dtype=np.float32).reshape(-1, 4))
                keep.append(a)
                setattr(q, field, _fptr(a))
            setattr(q, cnt, len(items))
        for k, v in enumerate(scene.get("velocity") or []):
            a = np.ascontiguousarray(v, dtype=np.float32)
            keep.append(a)
            q.velocity[k] = _fptr(a)
        rid = ctypes.c_longlong()
        _check(self.lib.smoke_server_submit(self.h, ctypes.byref(q), ctypes.byref(rid)))
        self.dims[rid.value] = (nz, ny, nx)
        return rid.value

    def tick(self):
        a, q = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.smoke_server_tick(self.h, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def result(self, rid):
        """(state, density or None, queue_ms, run_ms)."""
        st = ctypes.c_int()
        qms, rms = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.smoke_server_result(self.h, rid, ctypes.byref(st), None,
                                            ctypes.byref(qms), ctypes.byref(rms)))
        if st.value != REQUEST_DONE:
            return st.value, None, None, None
        out = np.empty(self.dims[rid], dtype=np.float32)
        _check(self.lib.smoke_server_result(self.h, rid, ctypes.byref(st), _fptr(out), None, None))
        return st.value, out, qms.value, rms.value

    def stats(self):
        s = ServerStats()
        _check(self.lib.smoke_server_stats_get(self.h, ctypes.byref(s)))
        return s.as_dict()

    def tick_ms(self, cap=1 << 20):
        ms = (ctypes.c_double * cap)()
        req = (ctypes.c_int * cap)()
        n = ctypes.c_int()
        _check(self.lib.smoke_server_tick_ms(self.h, ms, req, cap, ctypes.byref(n)))
        return (np.array(ms[: n.value], dtype=np.float64),
                np.array(req[: n.value], dtype=np.int64))


def _vec3(v):
    a = np.ascontiguousarray(v, dtype=np.float32).reshape(3)
    return a, _fptr(a)


def proxy_geometry(box_position, box_dimensions, view_origin, view_direction, steps):
    """Proxy-plane slicing geometry of one view (smoke_proxy_geometry):
    (vertices float32 (n, 3), indices int16 (m,), planes int32 (steps, 2) of
    (num_triangles, triangle_offset))."""
    lib = load()
    keep = [_vec3(v) for v in (box_position, box_dimensions, view_origin, view_direction)]
    steps = int(steps)
    cap = max(steps, 1)
    verts = np.empty((cap * PROXY_VERTICES_PER_PLANE, 3), np.float32)
    idx = np.empty(cap * PROXY_INDICES_PER_PLANE, np.int16)
    planes = np.empty((cap, 2), np.int32)
    nv, ni = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib.smoke_proxy_geometry(*(k[1] for k in keep), steps, _fptr(verts),
                                    idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                                    planes.ctypes.data_as(ctypes.POINTER(ProxyPlane)),
                                    ctypes.byref(nv), ctypes.byref(ni)))
    return verts[:nv.value].copy(), idx[:ni.value].copy(), planes[:steps].copy()


def proxy_geometry_device(box_position, box_dimensions, view_origin, view_direction, steps,
                          vertices, indices, planes, counts, stream=None):
    """smoke_proxy_geometry_device on device pointers (ints), enqueued on
    `stream` (a hipStream_t as int, None = the default stream)."""
    lib = load()
    keep = [_vec3(v) for v in (box_position, box_dimensions, view_origin, view_direction)]
    _check(lib.smoke_proxy_geometry_device(*(k[1] for k in keep), int(steps), vertices, indices,
                                           planes, counts, stream))
