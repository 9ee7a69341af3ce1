"""Step graphs (smoke_set_step_graph): steps replayed from a HIP graph give
the bits of eager steps, through changes that force a new capture (a moved
source, moved obstacle, new parameters, a new plan, another dt) and with the
program's own eager work (density reads, state reads) between the replays;
for the reference projection, Jacobi with an odd launch count (the pressure
parity alternates with the field parity) and the refined solve."""
import numpy as np
import pytest

import smokesim
from helpers import assert_bitwise_equal, gpu_for
from smokesim import scenes

pytestmark = pytest.mark.gpu

CASES = [
    # projection, iterations, plan (T, tile, zc), extra params
    (smokesim.PROJ_REFERENCE, 6, None, {}),
    (smokesim.PROJ_JACOBI, 5, (2, smokesim.TILE_32x32, 7), {}),
    (smokesim.PROJ_JACOBI, 10, (5, smokesim.TILE_64x32, 64), {}),
    (smokesim.PROJ_RBSOR, 4, None, {"warm_start": 1}),
    (smokesim.PROJ_REFINED, 8, (4, smokesim.TILE_32x32, 16), {"inner_sweeps": 4}),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"p{c[0]}-it{c[1]}")
def test_graph_steps_equal_eager_steps(case):
    proj, its, plan, extra = case
    sc = scenes.test_scene((29, 21, 17), seed=33, projection=proj, iterations=its,
                           n_sources=2, n_obstacles=2)
    sc["params"].update(extra)
    sims = [gpu_for(sc), gpu_for(sc)]
    sims[1].set_step_graph(True)
    if plan:
        for s in sims:
            s.set_jacobi_plan(sweeps_per_launch=plan[0], tile=plan[1], zc=plan[2])
    dt = sc["dt"]
    for k in range(9):
        if k == 3:
            for s in sims:
                s.move_object(0, 10.0, 8.0, 9.0)   # the first obstacle
        if k == 4:
            for s in sims:
                s.move_object(2, 14.0, 4.0, 6.0)   # the first source
        if k == 5:
            for s in sims:
                s.set_params(iterations=its + 2)
        if k == 6:
            dt = sc["dt"] * 0.5
        outs = [s.step(dt) for s in sims]
        assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32)), f"step {k}"
        if k % 2:
            assert_bitwise_equal(sims[1].state(), sims[0].state(), f"after step {k}")
            assert np.array_equal(sims[0].get_pressure().view(np.uint32),
                                  sims[1].get_pressure().view(np.uint32)), f"pressure {k}"
    for _ in range(3):  # replays back to back, then one read
        for s in sims:
            s.step_async(dt)
    assert_bitwise_equal(sims[1].state(), sims[0].state(), "after back-to-back replays")
    sims[1].set_step_graph(False)
    outs = [s.step(dt) for s in sims]
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32)), "after disabling"
    for s in sims:
        s.close()
