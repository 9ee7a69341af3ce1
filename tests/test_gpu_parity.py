"""Parity of the HIP step (libsmokesim.so) with the oracle.

The FP32 GPU kernels evaluate every value with the oracle's FP32 operations in
the same order (no contraction, IEEE division), so the bar is bit-for-bit
equality of every field of both ping-pong halves, the mask and the flags,
after every stage and every step.  Against the FP64 oracle the FP32 step is
held to a stated tolerance: relative L2 <= 1e-4 for density and <= 1e-3 for
each velocity component after the steps run here."""
import numpy as np
import pytest

import oracle as orc
import smokesim
from helpers import assert_bitwise_equal, gpu_for, oracle_for, random_density, rel_l2
from smokesim import scenes

pytestmark = pytest.mark.gpu

REF, RB, JAC = smokesim.PROJ_REFERENCE, smokesim.PROJ_RBSOR, smokesim.PROJ_JACOBI

CASES = [
    # dims, seed, projection, iterations, steps, extra params
    ((16, 16, 16), 0, REF, 20, 5, {}),
    ((24, 20, 28), 1, RB, 15, 5, {}),
    ((33, 17, 21), 2, JAC, 30, 5, {}),
    ((3, 3, 3), 3, REF, 5, 3, {}),
    ((4, 5, 3), 4, JAC, 4, 3, {}),
    ((65, 9, 12), 5, JAC, 25, 4, {"jacobi_omega": 0.8}),
    ((20, 31, 18), 6, RB, 10, 4, {"warm_start": 1}),
    ((18, 22, 20), 7, JAC, 12, 4, {"warm_start": 1}),
    ((25, 25, 25), 8, REF, 30, 4, {"obstacle_mode": smokesim.OBSTACLES_UNION}),
    ((21, 19, 17), 9, JAC, 0, 3, {}),
    ((30, 26, 22), 10, REF, 8, 4, {"density_gate": 0.0, "max_velocity": 3.0,
                                   "gravity": -9.82, "buoyancy": 2.0}),
]


def case_id(c):
    dims, seed, proj, iters, steps, extra = c
    return f"{dims[0]}x{dims[1]}x{dims[2]}-p{proj}-it{iters}-s{seed}" + ("-" + "-".join(extra) if extra else "")


def build(case, n_obstacles=2):
    dims, seed, proj, iters, steps, extra = case
    sc = scenes.test_scene(dims, seed=seed, projection=proj, iterations=iters,
                           n_obstacles=n_obstacles, n_sources=3)
    sc["params"].update(extra)
    dens = random_density(dims, seed) * np.float32(0.5)
    return sc, dens


@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_steps_bitwise_equal_fp32_oracle(case):
    sc, dens = build(case)
    steps = case[4]
    g = gpu_for(sc, dens)
    o = oracle_for(sc, 32, dens)
    assert_bitwise_equal(g.state(), o.state(), "initial")
    for k in range(steps):
        out = g.step(sc["dt"])
        o.step(sc["dt"])
        assert np.array_equal(out.view(np.uint32),
                              o.output_density().astype(np.float32).view(np.uint32)), f"step {k}"
        assert_bitwise_equal(g.state(), o.state(), f"after step {k}")
    if sc["params"]["projection"] != REF:
        p = g.get_pressure()
        po = o.get_pressure().astype(np.float32)
        assert np.array_equal(p.view(np.uint32), po.view(np.uint32))


@pytest.mark.parametrize("case", CASES[:3], ids=case_id)
def test_stages_bitwise_equal_fp32_oracle(case):
    sc, dens = build(case)
    g = gpu_for(sc, dens)
    o = oracle_for(sc, 32, dens)
    for step in range(2):
        for stage in range(smokesim.STAGE_COUNT):
            g.run_stage(stage, sc["dt"])
            o.stage(stage, sc["dt"])
            assert_bitwise_equal(g.state(), o.state(),
                                 f"step {step} stage {smokesim.STAGE_NAMES[stage]}")


@pytest.mark.parametrize("case", CASES[:3], ids=case_id)
def test_fp32_gpu_within_tolerance_of_fp64_oracle(case):
    sc, dens = build(case)
    steps = 10
    g = gpu_for(sc, dens)
    o = oracle_for(sc, 64, dens)
    for _ in range(steps):
        g.step(sc["dt"], readback=False)
    o.step(sc["dt"], steps)
    past = 1 - g.index_now
    assert rel_l2(g.read_density(), o.output_density()) <= 1e-4
    assert np.abs(g.read_density() - o.output_density()).max() <= 1e-4
    for k in range(3):
        assert rel_l2(g.get_velocity(k, past), o.get_velocity(k, past)) <= 1e-3


def test_moving_objects_and_parameter_changes():
    dims = (28, 24, 26)
    sc = scenes.test_scene(dims, seed=11, projection=REF, iterations=10, n_obstacles=2,
                           n_sources=2)
    g = gpu_for(sc)
    o = oracle_for(sc, 32)
    for k in range(6):
        # move one source and one obstacle every step, and flip the projection
        for sim in (g, o):
            sim.move_object(0, 5.0 + k, 4.0, 10.0 + 0.5 * k)
            sim.move_object(2, 14.0 - k, 12.0, 13.0)
            sim.set_params(projection=[REF, RB, JAC][k % 3], iterations=6 + k)
        g.step(0.04)
        o.step(0.04)
        assert_bitwise_equal(g.state(), o.state(), f"step {k}")


@pytest.mark.parametrize("dims,seed,noise,dt", [((70, 11, 10), 40, 0.5, 0.05),
                                                 ((67, 18, 9), 41, 40.0, 0.1),
                                                 ((130, 9, 6), 42, 8.0, 0.3),
                                                 ((129, 66, 35), 43, 6.0, 0.25)])
def test_velocity_advection_tiles_bitwise(dims, seed, noise, dt):
    # Velocity advection stages 64 x 4 x 2 cell tiles with a one-cell halo:
    # grids wider than one tile with partial tiles on every axis, and
    # back-traces longer than one cell (the last two cases: up to 2.7 cells
    # per step at their noise and dt) that read outside the tile.
    sc = scenes.test_scene(dims, seed=seed, projection=JAC, iterations=4, noise=noise)
    sc["dt"] = dt
    g, o = gpu_for(sc), oracle_for(sc, 32)
    for k in range(2):
        g.step(dt, readback=False)
        o.step(dt)
        assert_bitwise_equal(g.state(), o.state(), f"step {k}")


def test_max_objects_and_errors():
    g = smokesim.Simulation(8, 8, 8)
    for i in range(smokesim.MAX_OBJECTS):
        g.add_source(4.0, 4.0, 4.0, 1.0)
    with pytest.raises(smokesim.SmokeError):
        g.add_source(4.0, 4.0, 4.0, 1.0)
    with pytest.raises(smokesim.SmokeError):
        g.move_object(1000, 0, 0, 0)
    with pytest.raises(smokesim.SmokeError):
        g.set_params(projection=7)
    with pytest.raises(smokesim.SmokeError):
        g.step(float("nan"))
    with pytest.raises(smokesim.SmokeError):
        g.run_stage(9, 0.05)
    # the simulation is still usable after rejected calls
    g.step(0.05)


def test_footprint_matches_allocation_budget():
    n = smokesim.footprint_bytes(64, 48, 40)
    assert n > 0
    with smokesim.Simulation(64, 48, 40):
        pass


PLANS = [(t, tile, zc) for t in (1, 2, 3, 4, 5, 6)
         for tile in (smokesim.TILE_32x32, smokesim.TILE_64x16, smokesim.TILE_64x32, smokesim.TILE_32x64)
         for zc in (1, 7, 128)]


@pytest.mark.parametrize("plan", PLANS, ids=lambda p: f"T{p[0]}-tile{p[1]}-zc{p[2]}")
def test_every_jacobi_plan_is_bitwise_equal(plan):
    """Fused Jacobi sweeps (temporal blocking) give the single-sweep bits for
    every sweeps-per-launch, tile and z-chunk, plain and damped."""
    t, tile, zc = plan
    dims = (37, 29, 23)
    sc = scenes.test_scene(dims, seed=21, projection=JAC, iterations=13, n_obstacles=2,
                           n_sources=2)
    sc["params"]["jacobi_omega"] = 0.9 if (t + zc) % 2 else 1.0
    g = gpu_for(sc)
    g.set_jacobi_plan(sweeps_per_launch=t, tile=tile, zc=zc)
    o = oracle_for(sc, 32)
    for k in range(2):
        g.step(sc["dt"], readback=False)
        o.step(sc["dt"])
    assert_bitwise_equal(g.state(), o.state(), f"plan {plan}")
    p = g.get_pressure()
    assert np.array_equal(p.view(np.uint32), o.get_pressure().astype(np.float32).view(np.uint32))


def test_jacobi_plan_validation():
    g = smokesim.Simulation(8, 8, 8)
    for bad in (dict(sweeps_per_launch=0), dict(sweeps_per_launch=7), dict(tile=5), dict(zc=0)):
        with pytest.raises(smokesim.SmokeError):
            g.set_jacobi_plan(**bad)
    assert g.get_jacobi_plan() == {"sweeps_per_launch": 4, "tile": 0, "zc": 64}


def test_autotune_picks_a_plan_and_keeps_results_bitwise():
    dims = (48, 40, 36)
    sc = scenes.test_scene(dims, seed=22, projection=JAC, iterations=17, n_obstacles=1,
                           n_sources=2)
    g = gpu_for(sc)
    plan, ms = g.autotune()
    assert 1 <= plan["sweeps_per_launch"] <= 6 and ms > 0.0
    assert g.get_jacobi_plan() == plan
    o = oracle_for(sc, 32)
    for _ in range(2):
        g.step(sc["dt"], readback=False)
        o.step(sc["dt"])
    assert_bitwise_equal(g.state(), o.state(), f"autotuned plan {plan}")


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 5, 3), (65, 9, 12), (9, 70, 6), (130, 33, 5)],
                         ids=lambda d: f"{d[0]}x{d[1]}x{d[2]}")
def test_two_row_tiles_on_small_and_thin_grids(dims):
    """The two-row tiles (one block covers 64x32 or 32x64 cells) on grids
    smaller or thinner than a tile, every fused sweep count, z-chunks of 1
    plane and of the whole depth: the FP32 oracle's bits."""
    sc = scenes.test_scene(dims, seed=23, projection=JAC, iterations=11, n_obstacles=1,
                           n_sources=2)
    o = oracle_for(sc, 32)
    o.step(sc["dt"], 2)
    want = o.state()
    for tile in (smokesim.TILE_64x32, smokesim.TILE_32x64):
        for t in (2, 5, 6):
            for zc in (1, dims[2]):
                g = gpu_for(sc)
                g.set_jacobi_plan(sweeps_per_launch=t, tile=tile, zc=zc)
                for _ in range(2):
                    g.step(sc["dt"], readback=False)
                assert_bitwise_equal(g.state(), want, f"tile {tile} T={t} zc={zc}")
                g.close()
