"""The allocation-free entry point (smoke_get_step_norms_noalloc): its record,
written to caller device memory through the library's static scratch, equals
smoke_get_step_norms's bit for bit, on two calls in a row, the second with a
smaller grid (the scratch keeps the first call's partials, which must not
leak into the second), and on a state with non-finite values."""
import ctypes

import numpy as np
import pytest

import smokesim
from helpers import gpu_for
from smokesim import scenes

pytestmark = pytest.mark.gpu

FIELDS = [k for k, _ in smokesim.StepNorms._fields_]


class Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        vp = ctypes.c_void_p
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [vp]
        self.lib.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]

    def malloc(self, n):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), n) == 0
        return p

    def read_norms(self, dev):
        out = smokesim.StepNorms()
        assert self.lib.hipMemcpy(ctypes.byref(out), dev, ctypes.sizeof(out), 2) == 0  # D2H
        return out


def noalloc(sim, hip, dev):
    smokesim._check(sim.lib.smoke_get_step_norms_noalloc(sim.h, dev))
    sim.synchronize()
    return hip.read_norms(dev)


def same(rec, ref):
    """Every field bit for bit (NaN sums included)."""
    for k in FIELDS:
        a, b = getattr(rec, k), ref[k]
        if isinstance(a, float):
            assert np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64), (k, a, b)
        else:
            assert a == b, (k, a, b)


def test_noalloc_equals_the_allocating_entry_point():
    hip = Hip()
    dev = hip.malloc(ctypes.sizeof(smokesim.StepNorms))
    try:
        for dims, seed, steps in [((70, 41, 33), 31, 3), ((19, 14, 11), 32, 2)]:
            sc = scenes.test_scene(dims, seed=seed, projection=2, iterations=10, noise=0.3)
            g = gpu_for(sc)
            for _ in range(steps):
                g.step_async(sc["dt"])
            assert hip.lib.hipMemset(dev, 0xFF, ctypes.sizeof(smokesim.StepNorms)) == 0
            rec = noalloc(g, hip, dev)
            ref = g.step_norms()
            assert ref["status"] == smokesim.SMOKE_OK and rec.step == steps
            same(rec, ref)
            g.close()
        # non-finite values show in the record's count
        sc = scenes.test_scene((23, 17, 12), seed=33, projection=2, iterations=6)
        g = gpu_for(sc)
        g.step(sc["dt"], readback=False)
        h = 1 - g.index_now
        rho = g.get_density(h)
        rho[4, 4, 4] = np.nan
        g.set_density(h, rho)
        rec = noalloc(g, hip, dev)
        ref = g.step_norms()
        assert ref["status"] == smokesim.ERR_NONFINITE and rec.nonfinite == 1
        same(rec, ref)
        g.close()
    finally:
        hip.lib.hipFree(dev)
