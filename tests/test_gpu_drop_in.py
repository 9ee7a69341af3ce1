"""The drop-in claim, tested with a compiled C++ caller: tests/drop_in/
app_driver.cpp includes include/smokeSimulation.cuh, is built with the host
compiler alone (no HIP), links libsmokesim.so and drives the nine reference
functions the way the reference app does -- initializeVolume from a global's
constructor before main (project/main.cpp:87-88, boundingBox.cpp:30),
getGPUProperties, addSmokeSource, addObstacle, a first update at dt 0.01
(main.cpp:257-293), ticks with updateObjectPos and slider writes through
getGravity/getBuoyancy (:609-618, :832-833, :889-896), deleteVolume at the
end (boundingBox.cpp:397-402).  The density each simulate() returned equals
the FP32 oracle driven with the same inputs, bit for bit, on every tick."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "drop_in", "app_driver")
N = 80
TICKS = 12


def test_compiled_caller_matches_the_oracle(tmp_path):
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", os.path.dirname(DRIVER)], check=True, timeout=300)
    out = tmp_path / "density.bin"
    r = subprocess.run([DRIVER, str(TICKS), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "device 0" in r.stdout          # getGPUProperties printed the device
    rec = np.fromfile(out, dtype=np.float32).reshape(TICKS, 6 + N * N * N)
    o = orc.Oracle(N, N, N, 32)            # reference defaults: g -9.82, alpha 2, M 3, RB-SOR x30
    assert o.add_source(40.0, 40.0, 40.0, 5.0) == 0
    assert o.add_obstacle(60.0, 10.0, 60.0, 13.0) == 1
    for k in range(TICKS):
        x, y, z, dt, grav, buoy = (float(v) for v in rec[k, :6])
        if k > 0:
            o.move_object(0, x, y, z)
        o.set_params(gravity=grav, buoyancy=buoy)
        o.step(dt)
        got = rec[k, 6:].reshape(N, N, N)
        want = o.output_density().astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"tick {k}"
    assert rec[-1, 6:].max() > 0.0
    assert rec[4, 4] == np.float32(7.5) and rec[4, 5] == np.float32(4.0)
