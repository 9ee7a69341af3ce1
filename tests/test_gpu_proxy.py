"""Proxy-plane slicing geometry on the GPU against the oracle.

The kernel (proxy_kernels.hip) restates BoundingBox::generateProxyGeometry
(project/boundingBox.cpp:133-183) in FP32 with glm's operation order, so the
vertices are compared bit for bit with the oracle's (NaN vertices, which the
reference produces for planes that touch an edge lying in them or miss the
box, compared as NaN), and the 16-bit indices and the per-plane
(num_triangles, triangle_offset) table exactly."""
import ctypes

import numpy as np
import pytest

import oracle as orc
import smokesim

pytestmark = pytest.mark.gpu


def same(gpu, ref):
    gv, gi, gp = gpu
    ov, oi, op = ref
    assert gv.shape == ov.shape and gi.shape == oi.shape and gp.shape == op.shape
    assert np.array_equal(gi, oi) and np.array_equal(gp, op)
    nan = np.isnan(ov)
    assert np.array_equal(np.isnan(gv), nan)
    assert np.array_equal(gv[~nan].view(np.uint32), ov[~nan].view(np.uint32))


def random_view(seed):
    rng = np.random.default_rng(seed)
    dims = (rng.integers(16, 160, 3) * rng.choice([0.5, 1.0, 2.0])).astype(np.float32)
    box = np.zeros(3, np.float32) if seed % 3 else rng.uniform(-20, 20, 3).astype(np.float32)
    origin = rng.uniform(-300, 300, 3).astype(np.float32)
    d = rng.normal(size=3).astype(np.float32)
    if seed % 7 == 0:
        d[rng.integers(0, 3)] = 0.0  # view parallel to a box face
    return box, dims, origin, d


@pytest.mark.parametrize("steps", [1, 2, 20, 63, 64, 65, 100, 200, 255, 256, 257, 1000, 1023, 1024, 1025,
                                   2500, 4096])
def test_random_views_bitwise(steps):
    for seed in range(12):
        view = random_view(1000 * steps + seed)
        same(smokesim.proxy_geometry(*view, steps), orc.proxy_geometry(*view, steps))


def test_axis_aligned_and_degenerate_views():
    dims = (80.0, 80.0, 80.0)
    cases = [
        ((0, 0, 0), dims, (0, 0, 0), (0, 0, -1)),       # rectangles
        ((0, 0, 0), dims, (0, 0, 0), (1, 0, 0)),
        ((0, 0, 0), dims, (5, -3, 2), (1, 1, 1)),       # hexagons
        ((30, 0, 0), dims, (0, 0, 0), (-1, 0, 0)),      # planes that miss a shifted box
        ((0, 0, 0), (2, 2, 2), (0, 0, 0), (1, 1, 0)),    # edges lying in planes (NaN points)
        ((0, 0, 0), dims, (0, 0, 0), (0, 0, 0)),        # zero direction: NaN throughout
    ]
    for steps in (1, 3, 64, 100):
        for c in cases:
            same(smokesim.proxy_geometry(*c, steps), orc.proxy_geometry(*c, steps))


def test_device_entry_point_matches_host_entry_point():
    # Device buffers straight from the HIP runtime (the entry point takes raw
    # device pointers and a stream; NULL = the default stream).
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    d2h = 2  # hipMemcpyDeviceToHost
    view = random_view(77)
    steps = 300
    sizes = [steps * smokesim.PROXY_VERTICES_PER_PLANE * 12, steps * smokesim.PROXY_INDICES_PER_PLANE * 2,
             steps * 8, 8]
    ptrs = []
    try:
        for n in sizes:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), n) == 0
            ptrs.append(p.value)
        smokesim.proxy_geometry_device(*view, steps, *ptrs, None)
        assert hip.hipDeviceSynchronize() == 0
        counts = np.empty(2, np.int32)
        assert hip.hipMemcpy(counts.ctypes.data, ptrs[3], 8, d2h) == 0
        nv, ni = int(counts[0]), int(counts[1])
        gv = np.empty((nv, 3), np.float32)
        gi = np.empty(ni, np.int16)
        gp = np.empty((steps, 2), np.int32)
        for arr, ptr in ((gv, ptrs[0]), (gi, ptrs[1]), (gp, ptrs[2])):
            assert hip.hipMemcpy(arr.ctypes.data, ptr, arr.nbytes, d2h) == 0
    finally:
        for p in ptrs:
            hip.hipFree(p)
    same((gv, gi, gp), smokesim.proxy_geometry(*view, steps))


def test_argument_checks():
    v = ((0, 0, 0), (1, 1, 1), (0, 0, 0), (0, 0, 1))
    for steps in (0, -1, smokesim.PROXY_MAX_STEPS + 1):
        with pytest.raises(smokesim.SmokeError):
            smokesim.proxy_geometry(*v, steps)
    lib = smokesim.load()
    f3 = (ctypes.c_float * 3)(0, 0, 1)
    st = lib.smoke_proxy_geometry(None, f3, f3, f3, 4, None, None, None, None, None)
    assert st != smokesim.SMOKE_OK
