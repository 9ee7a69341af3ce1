"""Fused Jacobi launches on wide planes.

The fused sweeps address a block's span of planes (its z-chunk plus slack)
with 32-bit buffer offsets below 2^31 bytes, so the z-chunk of a plan is
capped for wide planes, and a plane too wide for any fused block runs one
sweep per launch.  The z-chunk and the sweeps per launch only change the
speed: both cases give the FP32 oracle's bits."""
import numpy as np
import pytest

import smokesim
from helpers import assert_bitwise_equal, gpu_for, oracle_for
from smokesim import scenes

pytestmark = pytest.mark.gpu


def wide_scene(nx, ny, nz, seed):
    sc = scenes.test_scene((nx, ny, nz), seed=seed, projection=smokesim.PROJ_JACOBI, iterations=8,
                           box_fraction=0.0, n_sources=2, n_obstacles=1, noise=0.05)
    return sc


@pytest.mark.parametrize("dims,zc", [((1790, 1790, 7), 1000), ((3400, 3400, 4), 64)],
                         ids=["capped-zchunk", "single-sweep-launches"])
def test_wide_planes_bitwise(dims, zc):
    sc = wide_scene(*dims, seed=71)
    g = gpu_for(sc)
    g.set_jacobi_plan(sweeps_per_launch=4, zc=zc)
    o = oracle_for(sc, 32)
    g.step(sc["dt"], readback=False)
    o.step(sc["dt"])
    assert_bitwise_equal(g.state(), o.state(), f"{dims} zc={zc}")
    g.close()
