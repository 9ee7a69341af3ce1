"""Checkpoint / resume (smoke_snapshot, smoke_restore): a simulation restored
from a snapshot continues bit for bit as the original, for every projection
and both field storages; mismatched or corrupt snapshots are refused."""
import numpy as np
import pytest

import smokesim
from helpers import assert_bitwise_equal, gpu_for
from smokesim import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("projection,storage", [
    (smokesim.PROJ_REFERENCE, smokesim.STORAGE_FP32),
    (smokesim.PROJ_RBSOR, smokesim.STORAGE_FP32),
    (smokesim.PROJ_JACOBI, smokesim.STORAGE_FP32),
    (smokesim.PROJ_REFINED, smokesim.STORAGE_BF16),
    (smokesim.PROJ_REFINED, smokesim.STORAGE_FP16),
    (smokesim.PROJ_JACOBI, smokesim.STORAGE_FP16_VELOCITY),
])
def test_restore_continues_bitwise(projection, storage):
    proj0 = smokesim.PROJ_JACOBI if projection == smokesim.PROJ_REFINED else projection
    sc = scenes.test_scene((29, 23, 19), seed=90, projection=proj0, iterations=12, n_sources=2,
                           n_obstacles=2)
    a = gpu_for(sc, storage=storage)
    if projection == smokesim.PROJ_REFINED:
        a.set_params(projection=smokesim.PROJ_REFINED, inner_sweeps=6, iterations=12)
    a.set_jacobi_plan(sweeps_per_launch=2, zc=16)
    for _ in range(3):
        a.step(sc["dt"], readback=False)
    blob = a.snapshot()
    b = smokesim.Simulation(29, 23, 19, storage=storage)
    b.restore(blob)
    assert b.get_params() == a.get_params() and b.get_jacobi_plan() == a.get_jacobi_plan()
    assert_bitwise_equal(b.state(), a.state(), "restored state")
    a.move_object(0, 10.0, 12.0, 9.0)  # the object list travels with the snapshot
    b.move_object(0, 10.0, 12.0, 9.0)
    for _ in range(3):
        a.step(sc["dt"], readback=False)
        b.step(sc["dt"], readback=False)
    assert_bitwise_equal(b.state(), a.state(), "after 3 more steps")
    a.close()
    b.close()


def test_restore_refuses_mismatch_and_corruption():
    a = smokesim.Simulation(16, 12, 10)
    blob = a.snapshot()
    with pytest.raises(smokesim.SmokeError):
        smokesim.Simulation(16, 12, 11).restore(blob)
    with pytest.raises(smokesim.SmokeError):
        smokesim.Simulation(16, 12, 10, storage=smokesim.STORAGE_BF16).restore(blob)
    with pytest.raises(smokesim.SmokeError):
        a.restore(blob[: len(blob) // 2])
    bad = bytearray(blob)
    bad[0:8] = b"XXXXXXXX"
    with pytest.raises(smokesim.SmokeError):
        a.restore(bytes(bad))
    a.restore(blob)  # the intact blob still loads
    a.close()


def test_restore_refuses_corrupt_object_lists():
    # A blob whose object list breaks what add_object enforces (more than
    # SMOKE_MAX_OBJECTS of a type, or an unknown type) is refused before any
    # state changes; the intact blob still loads.
    dims = (16, 12, 10)
    a = smokesim.Simulation(*dims)
    for k in range(smokesim.MAX_OBJECTS):
        a.add_obstacle(1.0 + k % 8, 2.0 + k // 8, 3.0, 1.0)
    a.add_source(8.0, 3.0, 5.0, 2.0)
    blob = a.snapshot()
    n = smokesim.MAX_OBJECTS + 1
    head = len(blob) - smokesim.footprint_bytes(*dims) - 32 * n   # objects are 32 bytes
    src_type = head + 32 * (n - 1) + 28
    assert blob[src_type:src_type + 4] == (1).to_bytes(4, "little")
    bad = bytearray(blob)
    bad[src_type:src_type + 4] = (0).to_bytes(4, "little")       # 65 obstacles
    with pytest.raises(smokesim.SmokeError):
        a.restore(bytes(bad))
    bad[src_type:src_type + 4] = (7).to_bytes(4, "little")       # unknown type
    with pytest.raises(smokesim.SmokeError):
        a.restore(bytes(bad))
    a.restore(blob)
    again = a.snapshot()
    assert again[head:head + 32 * n] == blob[head:head + 32 * n]   # the object list as saved
    assert a.get_mask().sum() < 16 * 11 * 10                       # the obstacles are in the mask
    a.close()
