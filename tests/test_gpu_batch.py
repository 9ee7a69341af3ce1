"""Batched stepping (smoke_batch_*): uneven simulations advanced with one
launch per stage give, bit for bit, what each gives alone (on the GPU and in
the FP32 oracle), eagerly and replayed from a captured HIP graph (several
replays with eager work and synchronisations in between), for every Jacobi
plan and the autotuned one."""
import numpy as np
import pytest

import smokesim
from helpers import assert_bitwise_equal, gpu_for, oracle_for
from smokesim import scenes

pytestmark = pytest.mark.gpu

DIMS = [(17, 23, 19), (40, 33, 21), (26, 48, 30), (9, 8, 12), (35, 18, 44), (64, 20, 24)]


def members(iterations=11, omega=1.0):
    out = []
    for k, dims in enumerate(DIMS):
        sc = scenes.test_scene(dims, seed=40 + k, projection=smokesim.PROJ_JACOBI,
                               iterations=iterations, n_sources=1 + k % 3, n_obstacles=k % 2)
        sc["params"]["jacobi_omega"] = omega
        sc["dt"] = 0.02 + 0.01 * k
        out.append(sc)
    return out


def run_alone(sc, steps):
    g = gpu_for(sc)
    for _ in range(steps):
        g.step(sc["dt"], readback=False)
    st = g.state()
    g.close()
    return st


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("omega", [1.0, 0.85])
def test_batch_equals_alone_and_oracle(use_graph, omega):
    scs = members(omega=omega)
    sims = [gpu_for(sc) for sc in scs]
    steps = 4
    with smokesim.Batch(sims, [sc["dt"] for sc in scs]) as b:
        for k in range(steps):
            b.step(use_graph=use_graph)
            if use_graph and k == 1:
                # eager work and synchronisation between graph replays
                b.synchronize()
                _ = sims[0].get_density(0)
        b.synchronize()
        states = [s.state() for s in sims]
    for sc, st in zip(scs, states):
        assert_bitwise_equal(st, run_alone(sc, steps), f"{sc['name']} batch vs alone")
        o = oracle_for(sc, 32)
        o.step(sc["dt"], steps)
        assert_bitwise_equal(st, o.state(), f"{sc['name']} batch vs oracle")


@pytest.mark.parametrize("plan", [(2, 0, 32), (3, 1, 7), (5, 0, 128), (1, 0, 64), (6, 1, 1),
                                  (5, 2, 64), (4, 3, 9), (6, 2, 256), (2, 2, 1)])
def test_batch_plans(plan):
    t, tile, zc = plan
    scs = members(iterations=13)
    sims = [gpu_for(sc) for sc in scs]
    with smokesim.Batch(sims, [sc["dt"] for sc in scs]) as b:
        b.set_jacobi_plan(sweeps_per_launch=t, tile=tile, zc=zc)
        for _ in range(2):
            b.step(use_graph=True)
        b.synchronize()
        states = [s.state() for s in sims]
    for sc, st in zip(scs, states):
        o = oracle_for(sc, 32)
        o.step(sc["dt"], 2)
        assert_bitwise_equal(st, o.state(), f"{sc['name']} plan {plan}")


def test_batch_autotune_and_readback():
    scs = members(iterations=9)
    sims = [gpu_for(sc) for sc in scs]
    with smokesim.Batch(sims, [sc["dt"] for sc in scs]) as b:
        plan, ms = b.autotune()
        assert ms > 0
        b.step()
        sims[2].readback_begin()
        b.step(use_graph=True)
        b.step(use_graph=True)
        arr, step = sims[2].readback_wait()
        assert step == 1
        b.synchronize()
        states = [s.state() for s in sims]
    o = oracle_for(scs[2], 32)
    o.step(scs[2]["dt"], 1)
    assert np.array_equal(arr.view(np.uint32), o.output_density().astype(np.float32).view(np.uint32))
    for sc, st in zip(scs, states):
        o = oracle_for(sc, 32)
        o.step(sc["dt"], 3)
        assert_bitwise_equal(st, o.state(), f"{sc['name']} autotuned {plan}")


def test_batch_rules():
    scs = members()
    sims = [gpu_for(sc) for sc in scs]
    dts = [sc["dt"] for sc in scs]
    sims[1].set_params(iterations=12)
    with pytest.raises(smokesim.SmokeError):
        smokesim.Batch(sims, dts)
    sims[1].set_params(iterations=11)
    with pytest.raises(smokesim.SmokeError):
        smokesim.Batch(sims + [sims[0]], dts + [0.05])
    b = smokesim.Batch(sims, dts)
    with pytest.raises(smokesim.SmokeError):
        sims[0].step(0.05)
    with pytest.raises(smokesim.SmokeError):
        sims[0].move_object(0, 1.0, 1.0, 1.0)
    with pytest.raises(smokesim.SmokeError):
        smokesim.Batch(sims[:2], dts[:2])        # already in a batch
    b.step()
    sims[3].close()                               # destroying a member breaks the batch
    with pytest.raises(smokesim.SmokeError):
        b.step()
    b.close()
    sims[0].step(0.05)                            # free again after the batch is gone


def test_batch_64_workload_members_match_oracle():
    """The benchmark's 64-simulation batch: every member equals its solo GPU
    run after two batched graph steps; four members also checked against the
    FP32 oracle."""
    scs = scenes.batch_64()
    sims = [gpu_for(sc) for sc in scs]
    with smokesim.Batch(sims, [sc["dt"] for sc in scs]) as b:
        b.autotune()
        b.step(use_graph=True)
        b.step(use_graph=True)
        b.synchronize()
        dens = [s.read_density() for s in sims]
    for s in sims:
        s.close()
    for k, sc in enumerate(scs):
        g = gpu_for(sc)
        g.step(sc["dt"], readback=False)
        alone = g.step(sc["dt"])
        g.close()
        assert np.array_equal(dens[k].view(np.uint32), alone.view(np.uint32)), sc["name"]
    for k in (0, 17, 42, 63):
        o = oracle_for(scs[k], 32)
        o.step(scs[k]["dt"], 2)
        assert np.array_equal(dens[k].view(np.uint32),
                              o.output_density().astype(np.float32).view(np.uint32)), scs[k]["name"]


def test_batch_long_backtraces():
    # The batched advection stages (velocity and density columns marched
    # through an LDS ring) with back-traces longer than one cell, which read
    # outside the ring: large velocity noise and dt, uneven sizes with partial
    # columns on every axis.
    scs = []
    for k, dims in enumerate([(70, 13, 37), (33, 66, 20), (9, 8, 12)]):
        sc = scenes.test_scene(dims, seed=60 + k, projection=smokesim.PROJ_JACOBI, iterations=5,
                               noise=8.0)
        sc["dt"] = 0.25 + 0.05 * k
        scs.append(sc)
    sims = [gpu_for(sc) for sc in scs]
    with smokesim.Batch(sims, [sc["dt"] for sc in scs]) as b:
        for _ in range(2):
            b.step(use_graph=False)
        b.synchronize()
        states = [s.state() for s in sims]
    for sc, st in zip(scs, states):
        o = oracle_for(sc, 32)
        o.step(sc["dt"], 2)
        assert_bitwise_equal(st, o.state(), f"{sc['name']} long back-traces")

