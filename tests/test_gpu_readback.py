"""Asynchronous, double-buffered density read-back (smoke_readback_begin /
smoke_readback_wait): every step's density arrives intact even though the
next steps are already enqueued (their source fill writes both density halves
and must wait for the pack), in FP32 bit-for-bit and in FP16/BF16 equal to a
round-to-nearest-even conversion of the oracle's FP32 density."""
import numpy as np
import pytest

import smokesim
from helpers import gpu_for, oracle_for
from smokesim import scenes

pytestmark = pytest.mark.gpu


def to_bf16_bits(a):
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
    nan = (b & 0x7FFFFFFF) > 0x7F800000
    r[nan] = ((b[nan] >> 16) | 0x40).astype(np.uint16)
    return r


@pytest.mark.parametrize("fmt", [smokesim.READBACK_F32, smokesim.READBACK_F16,
                                 smokesim.READBACK_BF16])
def test_async_readback_delivers_every_step(fmt):
    dims = (40, 36, 32)
    sc = scenes.test_scene(dims, seed=31, projection=smokesim.PROJ_JACOBI, iterations=20,
                           n_sources=3, n_obstacles=1)
    g = gpu_for(sc)
    o = oracle_for(sc, 32)
    expected, got = {}, {}
    for k in range(1, 9):
        # move a source every step so a late pack would see the next fill
        for sim in (g, o):
            sim.move_object(1, 8.0 + k, 5.0, 12.0)
        g.step_async(sc["dt"])
        g.readback_begin(fmt)
        o.step(sc["dt"])
        expected[k] = o.output_density().astype(np.float32)
        if k >= 2:
            arr, step = g.readback_wait()
            got[step] = arr
    arr, step = g.readback_wait()
    got[step] = arr
    assert sorted(got) == list(range(1, 9))
    for k in range(1, 9):
        if fmt == smokesim.READBACK_F32:
            assert np.array_equal(got[k].view(np.uint32), expected[k].view(np.uint32)), k
        elif fmt == smokesim.READBACK_F16:
            assert np.array_equal(got[k].view(np.uint16),
                                  expected[k].astype(np.float16).view(np.uint16)), k
        else:
            assert np.array_equal(got[k], to_bf16_bits(expected[k]).reshape(got[k].shape)), k


def test_readback_state_errors():
    g = smokesim.Simulation(12, 10, 8)
    with pytest.raises(smokesim.SmokeError):
        g.readback_wait()
    with pytest.raises(smokesim.SmokeError):
        g.readback_begin(7)
    g.step_async(0.05)
    g.readback_begin()
    g.readback_begin()
    with pytest.raises(smokesim.SmokeError):
        g.readback_begin()
    g.readback_wait()
    g.readback_wait()
    g.close()
