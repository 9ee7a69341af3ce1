"""The reference's nine-function interface (project/smokeSimulation.cuh:4-17)
through the library's legacy entry points, driven the way the reference app
drives it (project/main.cpp:253-293, :609-618, :832-833, :889-896), against
the FP32 oracle with the reference's default parameters.  Bit-for-bit."""
import ctypes

import numpy as np
import pytest

import oracle as orc
import smokesim

pytestmark = pytest.mark.gpu


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_legacy_interface_runs_the_reference_scene():
    lib = smokesim.load()
    n = 80
    grid = np.zeros((n, n, n), dtype=np.float32)   # BoundingBox::m_grid, zero-filled
    assert lib.smoke_legacy_initialize(fptr(grid), n, n, n) == 0
    assert lib.smoke_legacy_device_info() == 0
    src = lib.smoke_legacy_add_source(40.0, 40.0, 40.0, 5.0)       # main.cpp:288-289
    obs = lib.smoke_legacy_add_obstacle(60.0, 10.0, 60.0, 0.0, 0.0, 0.0, 13.0)  # main.cpp:291
    assert (src, obs) == (0, 1)

    o = orc.Oracle(n, n, n, 32)     # reference defaults: g -9.82, alpha 2, M 3, 30 iterations
    assert o.add_source(40.0, 40.0, 40.0, 5.0) == 0
    assert o.add_obstacle(60.0, 10.0, 60.0, 13.0) == 1

    # first tick at dt = 0.01 (main.cpp:293), then 20 Hz ticks with the
    # source circling and the sliders moved (main.cpp:609-618, :832-833).
    dts = [0.01] + [0.05] * 7
    for k, dt in enumerate(dts):
        if k == 3:
            lib.smoke_legacy_gravity_ptr()[0] = 7.5
            lib.smoke_legacy_buoyancy_ptr()[0] = 4.0
            o.set_params(gravity=7.5, buoyancy=4.0)
        if k >= 2:
            ang = 0.3 * k
            x, z = 40.0 + 20.0 * np.cos(ang), 40.0 + 20.0 * np.sin(ang)
            assert lib.smoke_legacy_update_object_pos(src, x, 40.0, z) == 0
            o.move_object(src, x, 40.0, z)
        assert lib.smoke_legacy_simulate(fptr(grid), dt) == 0
        o.step(dt)
        expect = o.output_density().astype(np.float32)
        assert np.array_equal(grid.view(np.uint32), expect.view(np.uint32)), f"tick {k}"
    assert grid.max() > 0.0
    assert lib.smoke_legacy_update_object_pos(99, 0.0, 0.0, 0.0) != 0
    assert lib.smoke_legacy_delete() == 0
    assert lib.smoke_legacy_simulate(fptr(grid), 0.05) != 0   # no volume
