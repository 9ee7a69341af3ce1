"""Vorticity confinement on the GPU against the oracle.

The reference lists vorticity confinement as future work (README.md:33) and
has no code for it; the oracle's vorticity_confinement defines it (pinned by
the closed-form cases in test_oracle_physics.py).  FP32 fields: bit-for-bit
equality with the FP32 oracle after every step, under every projection.
BF16 fields: relative L2 within the BF16 bars of the mixed-precision tests.
Against the FP64 oracle the FP32 step is held to relative L2 <= 1e-4 for
density and <= 1e-3 for each velocity component."""
import numpy as np
import pytest

import smokesim
from helpers import assert_bitwise_equal, gpu_for, oracle_for, random_density, rel_l2
from smokesim import scenes

pytestmark = pytest.mark.gpu

REF, RB, JAC = smokesim.PROJ_REFERENCE, smokesim.PROJ_RBSOR, smokesim.PROJ_JACOBI

CASES = [
    # dims, seed, projection, iterations, eps
    ((16, 16, 16), 40, REF, 12, 1.5),
    ((24, 20, 28), 41, RB, 10, 0.5),
    ((33, 17, 21), 42, JAC, 20, 3.0),
    ((3, 3, 3), 43, REF, 4, 2.0),
    ((65, 9, 12), 44, JAC, 16, 1.0),
]


def build(dims, seed, proj, iters, eps):
    sc = scenes.test_scene(dims, seed=seed, projection=proj, iterations=iters, noise=0.5,
                           n_sources=2, n_obstacles=2)
    sc["params"]["vorticity"] = eps
    return sc, random_density(dims, seed) * np.float32(0.5)


@pytest.mark.parametrize("dims,seed,proj,iters,eps", CASES,
                         ids=[f"{c[0][0]}x{c[0][1]}x{c[0][2]}-p{c[2]}-eps{c[4]}" for c in CASES])
def test_vorticity_step_bitwise(dims, seed, proj, iters, eps):
    sc, dens = build(dims, seed, proj, iters, eps)
    g, o = gpu_for(sc, dens), oracle_for(sc, 32, dens)
    for k in range(3):
        g.step(sc["dt"], readback=False)
        o.step(sc["dt"], 1)
        assert_bitwise_equal(g.state(), o.state(), f"step {k}")
    g.close()


def test_vorticity_force_stage_bitwise():
    sc, dens = build((19, 14, 13), 45, JAC, 8, 2.0)
    g, o = gpu_for(sc, dens), oracle_for(sc, 32, dens)
    for stage in (smokesim.STAGE_SCENE, smokesim.STAGE_FORCE):
        g.run_stage(stage, sc["dt"])
        o.stage(stage, sc["dt"])
        assert_bitwise_equal(g.state(), o.state(), f"stage {stage}")
    g.close()


def test_vorticity_changes_the_flow_and_tracks_fp64():
    sc, dens = build((24, 22, 20), 46, JAC, 30, 2.0)
    g = gpu_for(sc, dens)
    o64 = oracle_for(sc, 64, dens)
    off = oracle_for(dict(sc, params=dict(sc["params"], vorticity=0.0)), 64, dens)
    for _ in range(3):
        g.step(sc["dt"], readback=False)
    o64.step(sc["dt"], 3)
    off.step(sc["dt"], 3)
    gs, os_, fs = g.state(), o64.state(), off.state()
    n = gs["now"]
    assert rel_l2(gs[f"rho{n}"], os_[f"rho{n}"]) <= 1e-4
    for c in "uvw":
        assert rel_l2(gs[f"{c}{n}"], os_[f"{c}{n}"]) <= 1e-3, c
    # The confinement term is not a no-op on this flow.
    assert rel_l2(os_[f"u{n}"], fs[f"u{n}"]) > 1e-3
    g.close()


@pytest.mark.parametrize("projection", [JAC, smokesim.PROJ_REFINED])
def test_vorticity_bf16_fields_within_tolerance(projection):
    sc, dens = build((20, 18, 16), 47, JAC, 20, 1.5)
    g = gpu_for(sc, dens, storage=smokesim.STORAGE_BF16)
    g.set_params(projection=projection)
    o = oracle_for(sc, 64, dens)
    for _ in range(3):
        g.step(sc["dt"], readback=False)
    o.step(sc["dt"], 3)
    gs, os_ = g.state(), o.state()
    n = gs["now"]
    assert rel_l2(gs[f"rho{n}"], os_[f"rho{n}"]) <= 5e-3
    for c in "uvw":
        assert rel_l2(gs[f"{c}{n}"], os_[f"{c}{n}"]) <= 2e-2, c
    g.close()


def test_vorticity_argument_checks_and_batch_refusal():
    g = smokesim.Simulation(12, 10, 8)
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(smokesim.SmokeError):
            g.set_params(vorticity=bad)
    g.set_params(vorticity=1.0, projection=JAC, iterations=4)
    with pytest.raises(smokesim.SmokeError):
        smokesim.Batch([g], [0.05])
    g.set_params(vorticity=0.0)
    smokesim.Batch([g], [0.05]).close()  # the same simulation batches without it
    g.close()
