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


@pytest.mark.parametrize("dims,zc,tile", [((1790, 1790, 7), 1000, smokesim.TILE_32x32),
                                          ((1790, 1790, 7), 1000, smokesim.TILE_64x32),
                                          ((3400, 3400, 4), 64, smokesim.TILE_32x32)],
                         ids=["capped-zchunk", "capped-zchunk-two-rows", "single-sweep-launches"])
def test_wide_planes_bitwise(dims, zc, tile):
    sc = wide_scene(*dims, seed=71)
    g = gpu_for(sc)
    g.set_jacobi_plan(sweeps_per_launch=4, zc=zc, tile=tile)
    o = oracle_for(sc, 32)
    g.step(sc["dt"], readback=False)
    o.step(sc["dt"])
    assert_bitwise_equal(g.state(), o.state(), f"{dims} zc={zc}")
    g.close()


def test_wide_planes_refined_solve():
    # The refined solve's FP16 sweeps use the same 32-bit buffer spans: a
    # capped z-chunk gives the bits of a short one; a plane too wide for any
    # chunk is refused when the refined solve is chosen.
    sc = wide_scene(1790, 1790, 7, seed=72)
    states = []
    for zc, tile in ((1000, smokesim.TILE_32x32), (2, smokesim.TILE_64x32)):
        g = gpu_for(sc, storage=smokesim.STORAGE_FP16)
        g.set_params(projection=smokesim.PROJ_REFINED, inner_sweeps=4, iterations=8)
        g.set_jacobi_plan(sweeps_per_launch=4, zc=zc, tile=tile)
        g.step(sc["dt"], readback=False)
        st = g.state()
        st["pressure"] = g.get_pressure()
        states.append(st)
        g.close()
    for k, v in states[0].items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v.view(np.uint8), states[1][k].view(np.uint8)), k
        else:
            assert v == states[1][k], k
    sc = wide_scene(3400, 3400, 4, seed=73)
    g = gpu_for(sc, storage=smokesim.STORAGE_FP16)
    with pytest.raises(smokesim.SmokeError, match="too wide"):
        g.set_params(projection=smokesim.PROJ_REFINED, inner_sweeps=4, iterations=8)
    g.close()
