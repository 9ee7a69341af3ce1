"""Shared test helpers: build the same scene in the oracle and on the GPU and
compare every field of both ping-pong halves."""
import numpy as np

import oracle as orc
from smokesim import scenes

FIELDS = ["rho0", "rho1", "u0", "u1", "v0", "v1", "w0", "w1"]


def oracle_for(scene, real_bits, density=None):
    nx, ny, nz = scene["dims"]
    o = orc.Oracle(nx, ny, nz, real_bits)
    scenes.apply(o, scene, density)
    return o


def gpu_for(scene, density=None, storage=0):
    import smokesim
    nx, ny, nz = scene["dims"]
    s = smokesim.Simulation(nx, ny, nz, storage=storage)
    scenes.apply(s, scene, density)
    return s


def bit_mismatch(a, b):
    """Number of entries whose float32 bit patterns differ."""
    a32 = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    b32 = np.ascontiguousarray(b, dtype=np.float32).view(np.uint32)
    return int(np.count_nonzero(a32 != b32))


def assert_bitwise_equal(gpu_state, orc_state, context=""):
    assert gpu_state["now"] == orc_state["now"], context
    assert np.array_equal(gpu_state["mask"], orc_state["mask"]), f"mask differs {context}"
    assert np.array_equal(gpu_state["flags"], orc_state["flags"]), f"flags differ {context}"
    for k in FIELDS:
        n = bit_mismatch(gpu_state[k], orc_state[k])
        if n:
            diff = np.abs(gpu_state[k].astype(np.float64) - orc_state[k])
            idx = np.unravel_index(np.argmax(diff), diff.shape)
            raise AssertionError(
                f"{k}: {n} entries differ bitwise {context}; max |diff| {diff.max():.3e} at "
                f"(z,y,x)={idx}: gpu {gpu_state[k][idx]!r} oracle {orc_state[k][idx]!r}")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def random_density(dims, seed):
    nx, ny, nz = dims
    r = scenes.uniform_array(scenes.splitmix64(0xD5 ^ seed), nx * ny * nz)
    return r.reshape(nz, ny, nx).astype(np.float32)
