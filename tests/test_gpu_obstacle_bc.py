"""Moving-obstacle velocity boundary condition on the GPU against the oracle.

The reference stores each obstacle's velocity (smokeSimulation.cu:45-52) and
never reads it (:299-301); the oracle's obstacle_velocity_bc defines the
condition (pinned in test_oracle_physics.py by a brute-force face-ownership
restatement and by the flow a sphere induces in still fluid).  The condition
only assigns stored velocities, so FP32 fields are compared bit for bit with
the FP32 oracle after every step while the obstacles move."""
import numpy as np
import pytest

import smokesim
from helpers import assert_bitwise_equal, gpu_for, oracle_for, random_density, rel_l2
from smokesim import scenes

pytestmark = pytest.mark.gpu

REF, RB, JAC = smokesim.PROJ_REFERENCE, smokesim.PROJ_RBSOR, smokesim.PROJ_JACOBI
LAST, UNION = smokesim.OBSTACLES_LAST_WINS, smokesim.OBSTACLES_UNION

CASES = [
    # dims, seed, projection, iterations, obstacle mode
    ((16, 16, 16), 50, REF, 12, LAST),
    ((24, 20, 18), 51, RB, 10, UNION),
    ((33, 17, 21), 52, JAC, 20, UNION),
    ((20, 22, 19), 53, JAC, 16, LAST),
]


def build(dims, seed, proj, iters, mode):
    sc = scenes.test_scene(dims, seed=seed, projection=proj, iterations=iters, n_sources=2,
                           n_obstacles=3)
    sc["params"].update(obstacle_velocity=1, obstacle_mode=mode)
    return sc, random_density(dims, seed) * np.float32(0.5)


def obstacle_ids(sc):
    # Sources and obstacles share ids in insertion order (obstacles first).
    return list(range(len(sc["obstacles"])))


def motion(step, k):
    rng = np.random.default_rng(1000 * step + k)
    return rng.uniform(-1.5, 1.5, 3).astype(np.float32), rng.uniform(-0.8, 0.8, 3).astype(np.float32)


@pytest.mark.parametrize("dims,seed,proj,iters,mode", CASES,
                         ids=[f"{c[0][0]}x{c[0][1]}x{c[0][2]}-p{c[2]}-m{c[4]}" for c in CASES])
def test_moving_obstacles_bitwise(dims, seed, proj, iters, mode):
    sc, dens = build(dims, seed, proj, iters, mode)
    g, o = gpu_for(sc, dens), oracle_for(sc, 32, dens)
    for step in range(4):
        for k, oid in enumerate(obstacle_ids(sc)):
            dpos, vel = motion(step, k)
            x, y, z, _ = sc["obstacles"][k]
            pos = [float(np.float32(p + d)) for p, d in zip((x, y, z), dpos)]
            for sim in (g, o):
                sim.move_object(oid, *pos)
                sim.set_object_velocity(oid, *(float(t) for t in vel))
        g.step(sc["dt"], readback=False)
        o.step(sc["dt"], 1)
        assert_bitwise_equal(g.state(), o.state(), f"step {step}")
    g.close()


def test_condition_is_applied_in_the_scene_stage():
    sc, dens = build((18, 15, 14), 54, JAC, 8, UNION)
    g, o = gpu_for(sc, dens), oracle_for(sc, 32, dens)
    for sim in (g, o):
        sim.set_object_velocity(0, 0.5, -0.25, 1.0)
    g.run_stage(smokesim.STAGE_SCENE, sc["dt"])
    o.stage(smokesim.STAGE_SCENE, sc["dt"])
    assert_bitwise_equal(g.state(), o.state(), "scene stage")
    g.close()


def test_bf16_fields_with_moving_obstacles_within_tolerance():
    sc, dens = build((20, 18, 16), 55, JAC, 20, UNION)
    g = gpu_for(sc, dens, storage=smokesim.STORAGE_BF16)
    o = oracle_for(sc, 64, dens)
    for sim in (g, o):
        sim.set_object_velocity(0, 1.0, 0.0, -0.5)
    for _ in range(3):
        g.step(sc["dt"], readback=False)
    o.step(sc["dt"], 3)
    gs, os_ = g.state(), o.state()
    n = gs["now"]
    assert rel_l2(gs[f"rho{n}"], os_[f"rho{n}"]) <= 5e-3
    for c in "uvw":
        assert rel_l2(gs[f"{c}{n}"], os_[f"{c}{n}"]) <= 2e-2, c
    g.close()


def test_restore_continues_bitwise_with_the_condition():
    sc, dens = build((17, 14, 13), 56, JAC, 10, LAST)
    a = gpu_for(sc, dens)
    a.set_object_velocity(0, 0.75, 0.5, -0.25)
    a.step(sc["dt"], readback=False)
    b = smokesim.Simulation(*sc["dims"])
    b.restore(a.snapshot())
    for _ in range(2):
        a.step(sc["dt"], readback=False)
        b.step(sc["dt"], readback=False)
    assert_bitwise_equal(b.state(), a.state(), "restored")
    a.close()
    b.close()


def test_argument_checks_and_batch_refusal():
    g = smokesim.Simulation(12, 10, 8)
    oid = g.add_obstacle(6.0, 5.0, 4.0, 2.0)
    with pytest.raises(smokesim.SmokeError):
        g.set_params(obstacle_velocity=2)
    with pytest.raises(smokesim.SmokeError):
        g.set_object_velocity(oid + 5, 1.0, 0.0, 0.0)
    with pytest.raises(smokesim.SmokeError):
        g.set_object_velocity(oid, float("nan"), 0.0, 0.0)
    g.set_params(obstacle_velocity=1, projection=JAC, iterations=4)
    with pytest.raises(smokesim.SmokeError):
        smokesim.Batch([g], [0.05])
    g.set_params(obstacle_velocity=0)
    smokesim.Batch([g], [0.05]).close()
    g.close()
