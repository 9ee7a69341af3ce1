"""Deterministic synthetic scenes for the smoke step.

Every random quantity comes from SplitMix64 streams keyed by
(workload id, simulation id, purpose), so the GPU runs, the CPU oracle and the
tests see identical inputs on any machine.  Units are cells and seconds.

Workloads (sizes and physics of the benchmark configurations):
  * ``reference_80``  -- 80^3, solid floor, obstacle sphere r=10 at the centre,
    source sphere r=6 at (40,8,40), velocity noise +-0.01, net buoyancy
    +9.8/density (gravity 7.8 + buoyancy 2.0, both reachable with the
    reference's own sliders, project/main.cpp:832-833), density gate 0.001,
    clamp M=4, dt=0.05, 40 red-black SOR iterations, 100 steps.
  * ``grid_256``  -- 256^3, solid floor plus seeded axis-aligned boxes (edges
    4..32) until >= 5 % of cells are solid, 4 seeded sources (r in [4,10],
    centres in the lower third), same physics, 100 Jacobi sweeps, dt=0.05.
  * ``grid_512``  -- as grid_256 at 512^3 with 8 sources.
"""
import numpy as np

GOLDEN = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1

WORKLOAD_IDS = {"reference_80": 0, "grid_256": 1, "grid_512": 2, "batch_64": 3, "serve": 4,
                "test": 99}
PURPOSE_VELOCITY = 1
PURPOSE_BOXES = 2
PURPOSE_SOURCES = 3
PURPOSE_DIMS = 4
PURPOSE_DT = 5


def splitmix64(x):
    """One SplitMix64 output for state x (Python int)."""
    z = (x + GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def stream_seed(workload, sim_id, purpose):
    wid = WORKLOAD_IDS.get(workload, workload) if isinstance(workload, str) else int(workload)
    return splitmix64((wid << 40) ^ (int(sim_id) << 20) ^ int(purpose))


class Stream:
    """Scalar SplitMix64 stream."""

    def __init__(self, seed):
        self.state = int(seed) & M64

    def next_u64(self):
        self.state = (self.state + GOLDEN) & M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 40) / float(1 << 24))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)


def uniform_array(seed, n):
    """n uniforms in [0,1) (24-bit), the i-th output of the SplitMix64
    sequence started at `seed`, computed vectorised."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def velocity_noise(dims, amplitude, seed):
    """u, v, w of shape (nz+1, ny+1, nx+1), uniform in [-a, a), float32."""
    nx, ny, nz = dims
    shape = (nz + 1, ny + 1, nx + 1)
    n = shape[0] * shape[1] * shape[2]
    out = []
    for k in range(3):
        r = uniform_array(splitmix64(seed ^ (k + 1)), n)
        out.append(((2.0 * r - 1.0) * amplitude).astype(np.float32).reshape(shape))
    return out


def floor_solid(dims):
    nx, ny, nz = dims
    s = np.zeros((nz, ny, nx), dtype=np.uint8)
    s[:, 0, :] = 1
    return s


def seeded_boxes(dims, fraction, seed, edge=(4, 32)):
    """Solid mask (nonzero = solid): the y = 0 floor plus axis-aligned boxes
    added until at least `fraction` of all cells are solid."""
    nx, ny, nz = dims
    s = floor_solid(dims)
    rng = Stream(seed)
    target = fraction * s.size
    boxes = []
    count = int(s.sum())
    guard = 0
    while count < target and guard < 100000:
        guard += 1
        ex = min(rng.integer(*edge), nx)
        ey = min(rng.integer(*edge), ny)
        ez = min(rng.integer(*edge), nz)
        x0 = rng.integer(0, nx - ex)
        y0 = rng.integer(0, ny - ey)
        z0 = rng.integer(0, nz - ez)
        sl = s[z0:z0 + ez, y0:y0 + ey, x0:x0 + ex]
        count += int(sl.size - sl.sum())
        sl[...] = 1
        boxes.append((x0, y0, z0, ex, ey, ez))
    return s, boxes


def seeded_sources(dims, count, seed, radius=(4.0, 10.0)):
    """Source spheres (x, y, z, r): radius uniform in `radius`, centres
    uniform in the lower third of the grid."""
    nx, ny, nz = dims
    rng = Stream(seed)
    out = []
    for _ in range(count):
        r = rng.uniform(*radius)
        x = rng.uniform(min(r, nx / 2), max(nx - r, nx / 2))
        y = rng.uniform(1.0, max(ny / 3.0, 1.0))
        z = rng.uniform(min(r, nz / 2), max(nz - r, nz / 2))
        out.append((float(np.float32(x)), float(np.float32(y)), float(np.float32(z)),
                    float(np.float32(r))))
    return out


PHYSICS = dict(gravity=7.8, buoyancy=2.0, density_gate=0.001, max_velocity=4.0)


def reference_80():
    """80^3 reference-size workload (see module docstring)."""
    dims = (80, 80, 80)
    seed = stream_seed("reference_80", 0, PURPOSE_VELOCITY)
    return dict(
        name="reference_80", dims=dims, solid=None,
        obstacles=[(40.0, 40.0, 40.0, 10.0)], sources=[(40.0, 8.0, 40.0, 6.0)],
        velocity=velocity_noise(dims, 0.01, seed),
        params=dict(PHYSICS, projection=0, iterations=40), dt=0.05, steps=100)


def grid(n, sources, workload, projection=2, iterations=100, dims=None):
    """Cubic benchmark grid (grid_256 / grid_512)."""
    dims = dims or (n, n, n)
    solid, boxes = seeded_boxes(dims, 0.05, stream_seed(workload, 0, PURPOSE_BOXES))
    return dict(
        name=workload, dims=dims, solid=solid, boxes=boxes, obstacles=[],
        sources=seeded_sources(dims, sources, stream_seed(workload, 0, PURPOSE_SOURCES)),
        velocity=velocity_noise(dims, 0.01, stream_seed(workload, 0, PURPOSE_VELOCITY)),
        params=dict(PHYSICS, projection=projection, iterations=iterations), dt=0.05, steps=50)


def grid_256(projection=2, iterations=100):
    return grid(256, 4, "grid_256", projection, iterations)


def grid_512(projection=2, iterations=100):
    return grid(512, 8, "grid_512", projection, iterations)


def batch_member(sim_id, workload="batch_64", lo=32, hi=160, iterations=60):
    """One simulation of the batched workload: each axis uniform in [lo, hi],
    seeded boxes covering a uniform 0-10 % of cells, 1-3 sources (r in
    [3, 8]), dt uniform in [0.02, 0.08], Jacobi x iterations."""
    rng = Stream(stream_seed(workload, sim_id, PURPOSE_DIMS))
    dims = (rng.integer(lo, hi), rng.integer(lo, hi), rng.integer(lo, hi))
    frac = rng.uniform(0.0, 0.10)
    n_src = rng.integer(1, 3)
    dt = float(np.float32(rng.uniform(0.02, 0.08)))
    solid, boxes = seeded_boxes(dims, frac, stream_seed(workload, sim_id, PURPOSE_BOXES))
    return dict(
        name=f"{workload}[{sim_id}]", dims=dims, solid=solid, boxes=boxes, obstacles=[],
        sources=seeded_sources(dims, n_src, stream_seed(workload, sim_id, PURPOSE_SOURCES),
                               radius=(3.0, 8.0)),
        velocity=velocity_noise(dims, 0.01, stream_seed(workload, sim_id, PURPOSE_VELOCITY)),
        params=dict(PHYSICS, projection=2, iterations=iterations), dt=dt, steps=50)


def serve_request(req_id, workload="serve", lo=48, hi=256):
    """One request of the serving workload: each axis log-uniform in
    [lo, hi], 20-200 steps, 1-3 sources in the lower third, solid floor,
    starting at rest (no noise, so requests can be generated as they arrive),
    dt = 0.05.  Returns (scene, steps)."""
    rng = Stream(stream_seed(workload, req_id, PURPOSE_DIMS))
    dims = tuple(int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
    steps = rng.integer(20, 200)
    n_src = rng.integer(1, 3)
    sc = dict(
        name=f"{workload}[{req_id}]", dims=dims, solid=None, obstacles=[],
        sources=seeded_sources(dims, n_src, stream_seed(workload, req_id, PURPOSE_SOURCES),
                               radius=(3.0, 8.0)),
        velocity=[], params=dict(PHYSICS, projection=2, iterations=60), dt=0.05)
    return sc, steps


def batch_64(count=64, iterations=60):
    """The batched workload: `count` independent simulations."""
    return [batch_member(i, iterations=iterations) for i in range(count)]


def test_scene(dims, seed=0, box_fraction=0.05, n_sources=2, n_obstacles=1, projection=0,
               iterations=20, noise=0.01):
    """Small randomised scene for parity tests."""
    nx, ny, nz = dims
    solid = None
    if box_fraction > 0:
        edge = (1, max(2, min(dims) // 3))
        solid, _ = seeded_boxes(dims, box_fraction, stream_seed("test", seed, PURPOSE_BOXES), edge)
    rng = Stream(stream_seed("test", seed, PURPOSE_SOURCES))
    sources, obstacles = [], []
    rmax = max(1.5, min(dims) / 4.0)
    for _ in range(n_sources):
        sources.append((rng.uniform(0, nx), rng.uniform(0, ny / 2), rng.uniform(0, nz),
                        rng.uniform(1.0, rmax)))
    for _ in range(n_obstacles):
        obstacles.append((rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz),
                          rng.uniform(1.0, rmax)))
    sources = [tuple(float(np.float32(v)) for v in s) for s in sources]
    obstacles = [tuple(float(np.float32(v)) for v in s) for s in obstacles]
    return dict(
        name=f"test_{nx}x{ny}x{nz}_s{seed}", dims=dims, solid=solid, obstacles=obstacles,
        sources=sources,
        velocity=velocity_noise(dims, noise, stream_seed("test", seed, PURPOSE_VELOCITY)),
        params=dict(PHYSICS, projection=projection, iterations=iterations), dt=0.05, steps=5)


def apply(sim, scene, density=None):
    """Load a scene into a Simulation (GPU) or an Oracle (CPU): both expose the
    same set_solids / add_obstacle / add_source / set_velocity / set_params."""
    if scene.get("solid") is not None:
        sim.set_solids(scene["solid"])
    for (x, y, z, r) in scene["obstacles"]:
        sim.add_obstacle(x, y, z, r)
    for (x, y, z, r) in scene["sources"]:
        sim.add_source(x, y, z, r)
    for k, a in enumerate(scene["velocity"]):
        sim.set_velocity(k, 0, a)
    if density is not None:
        sim.set_density(0, density)
    sim.set_params(**scene["params"])
    return sim
