/*
 * smoke.h -- C ABI of the MI355X smoke time step (libsmokesim.so).
 *
 * Plain C: opaque handle, POD structs, pointers and sizes, status codes.  No
 * HIP or C++ types cross this boundary.  Each entry point names the reference
 * interface it replaces (project/smokeSimulation.cuh / .cu of
 * RasmusAlmryd/smoke-simulation).  The reference's own nine C++ functions are
 * kept, unchanged in name and meaning, by the header-only compat layer
 * smokeSimulation.cuh in this directory, which calls the smoke_legacy_*
 * entry points below.
 *
 * Layouts at the boundary are the reference's:
 *   density, solid mask : nx*ny*nz, c = x + y*nx + z*nx*ny      (:267)
 *   velocity component  : (nx+1)*(ny+1)*(nz+1),
 *                         f = x + y*(nx+1) + z*(nx+1)*(ny+1)     (:146-147)
 * Units are cells and seconds (the reference has no dx).
 *
 * Threading: calls on one handle are serialised by the library; different
 * handles are independent.  Every call that returns smoke_status leaves a
 * message for smoke_last_error() (per thread) when it fails.
 */
#ifndef SMOKE_H
#define SMOKE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smoke_sim smoke_sim;

typedef enum smoke_status {
    SMOKE_OK = 0,
    SMOKE_ERR_INVALID = 1, /* bad argument (dims, id, size, NULL) */
    SMOKE_ERR_HIP = 2,     /* HIP runtime error */
    SMOKE_ERR_NOMEM = 3,   /* device allocation failed */
    SMOKE_ERR_STATE = 4,   /* call not valid in the current state */
    SMOKE_ERR_NONFINITE = 5 /* a field holds NaN or Inf (smoke_get_step_norms) */
} smoke_status;

/* Pressure projection.  REFERENCE is the reference's in-place red-black SOR on
 * face velocities (smokeSimulation.cu:356-394, :795-801).  RBSOR and JACOBI
 * solve the same linear system in pressure form (P, right-hand side div u*,
 * then u -= grad P on fluid-fluid faces).  REFINED is the mixed-precision
 * solve of that system: an FP32 residual-correction loop whose corrections
 * take `inner_sweeps` FP16 Jacobi sweeps each, FP64 norms on the device; it
 * runs ceil(iterations / inner_sweeps) refinements, fewer once the relative
 * residual reaches `tolerance` (> 0).  Cold start only; planes of at most
 * 10956549 cells (nx+1 padded to 32, times ny+1; smoke_set_params refuses
 * REFINED on wider ones). */
typedef enum smoke_projection {
    SMOKE_PROJ_REFERENCE = 0,
    SMOKE_PROJ_RBSOR = 1,
    SMOKE_PROJ_JACOBI = 2,
    SMOKE_PROJ_REFINED = 3
} smoke_projection;

/* Storage of density and u, v, w (chosen at creation).  The 16-bit forms
 * keep FP32 arithmetic (widen on load, round to nearest even on store);
 * pressure and divergence stay FP32.  BF16 / FP16: density and velocity in
 * BF16 (8 significant bits, FP32's range) or IEEE binary16 (11 significant
 * bits, magnitudes up to 65504: the clamp bounds the speeds far below that;
 * a value beyond it becomes Inf, which smoke_get_step_norms reports).
 * *_VELOCITY: u, v, w in 16 bits and the density in FP32.  Simulations with
 * 16-bit velocity use the pressure forms (RBSOR, JACOBI, REFINED), not
 * REFERENCE, and do not join batches. */
typedef enum smoke_storage {
    SMOKE_STORAGE_FP32 = 0,
    SMOKE_STORAGE_BF16 = 1,
    SMOKE_STORAGE_BF16_VELOCITY = 2,
    SMOKE_STORAGE_FP16 = 3,
    SMOKE_STORAGE_FP16_VELOCITY = 4
} smoke_storage;

/* How several obstacles combine into the interior mask.  LAST_WINS is the
 * reference's fillObstacle (smokeSimulation.cu:295-311); UNION marks a cell
 * solid if any obstacle covers it. */
typedef enum smoke_obstacle_mode {
    SMOKE_OBSTACLES_LAST_WINS = 0,
    SMOKE_OBSTACLES_UNION = 1
} smoke_obstacle_mode;

/* Step parameters.  Defaults (smoke_params_default) are the reference's. */
typedef struct smoke_params {
    float gravity;      /* smokeSimulation.cu:28   -9.82; getGravity()        */
    float buoyancy;     /* smokeSimulation.cu:29    2.0;  getBuoyancy()       */
    float density_gate; /* buoyancy only where density > gate; <= 0: no gate  */
    float max_velocity; /* smokeSimulation.cu:39    3.0 (speed clamp M)       */
    double sor_omega;   /* smokeSimulation.cu:38    1.9                       */
    double jacobi_omega;/* Jacobi damping, 1.0 = plain Jacobi                 */
    int projection;     /* smoke_projection, default REFERENCE                */
    int iterations;     /* smokeSimulation.cu:797  30: red-black iterations
                           (two half-sweeps each) or Jacobi sweeps            */
    int obstacle_mode;  /* smoke_obstacle_mode, default LAST_WINS             */
    int warm_start;     /* pressure forms: start from the last step's P       */
    int inner_sweeps;   /* REFINED: FP16 sweeps per refinement (2..1000, even),
                           default 20                                         */
    float vorticity;    /* vorticity-confinement strength eps (>= 0); 0 = off,
                           the reference (README.md:33 lists it as future
                           work)                                              */
    double tolerance;   /* REFINED: stop at this relative residual; 0: run all */
    int obstacle_velocity; /* 1: moving-obstacle velocity boundary condition:
                           each step, after the mask is built, every face of
                           [now] next to a cell an obstacle made solid takes
                           that obstacle's velocity (the high-side cell's
                           obstacle when both are).  0 = off, the reference
                           (it stores obstacle velocities and never reads
                           them, smokeSimulation.cu:45-52, :299-301)         */
} smoke_params;

/* Step stages, in execution order (for stage-level testing and timing). */
typedef enum smoke_stage {
    SMOKE_STAGE_SCENE = 0,       /* swap halves, sources, solid mask      */
    SMOKE_STAGE_FORCE = 1,       /* buoyancy + speed clamp                */
    SMOKE_STAGE_PROJECT = 2,     /* pressure projection                   */
    SMOKE_STAGE_ADVECT_VEL = 3,  /* velocity advection [now] -> [past]    */
    SMOKE_STAGE_ADVECT_DENS = 4, /* density advection  [now] -> [past]    */
    SMOKE_STAGE_COUNT = 5
} smoke_stage;

/* ----------------------------------------------------------------- basics */

/* Message of the last failed call on this thread ("" if none). */
const char* smoke_last_error(void);

/* Library version string. */
const char* smoke_version(void);

/* Replaces getGPUProperties (smokeSimulation.cu:62-85): writes a one-line
 * description of device 0 into buf.  SMOKE_ERR_HIP if no GPU is visible. */
smoke_status smoke_device_info(char* buf, size_t len);

/* Reference defaults: gravity -9.82, buoyancy 2.0, no gate, M = 3, omega 1.9,
 * REFERENCE projection, 30 iterations, LAST_WINS, cold start. */
smoke_status smoke_params_default(smoke_params* out);

/* ------------------------------------------------------------- lifecycle */

/* Replaces initializeVolume (smokeSimulation.cu:124-238).  Allocates both
 * ping-pong halves of density and u, v, w on the current device, copies
 * `density` (nx*ny*nz floats, may be NULL for zeros) into half 0, zeroes the
 * rest, and sets the solid mask to the y = 0 floor.  Requires n >= 3 per
 * axis. */
smoke_status smoke_create(unsigned nx, unsigned ny, unsigned nz, const float* density,
                          smoke_sim** out);

/* smoke_create with a field storage (smoke_storage).  No reference
 * counterpart: the mixed-precision forms' 16-bit fields. */
smoke_status smoke_create_ex(unsigned nx, unsigned ny, unsigned nz, const float* density,
                             int storage, smoke_sim** out);
int smoke_get_storage(const smoke_sim* sim);

/* Replaces deleteVolume (smokeSimulation.cu:240-248).  NULL is a no-op. */
void smoke_destroy(smoke_sim* sim);

smoke_status smoke_get_dims(const smoke_sim* sim, unsigned* nx, unsigned* ny, unsigned* nz);

/* Device bytes held by one simulation of the given size (budgeting). */
size_t smoke_footprint_bytes(unsigned nx, unsigned ny, unsigned nz);
/* ... with a storage and, if `refined`, the REFINED solver's buffers. */
size_t smoke_footprint_bytes_ex(unsigned nx, unsigned ny, unsigned nz, int storage, int refined);

/* REFINED: relative residual ||r|| / ||div|| (FP64) after the last step's
 * solve and the refinements it ran.  Waits for the simulation's work. */
smoke_status smoke_get_solve_info(smoke_sim* sim, double* rel_residual, int* refinements);

/* Checkpoint / resume (no reference counterpart: the reference's state lives
 * in module globals).  A snapshot is a host blob holding every device field
 * of both ping-pong halves, the masks, the ping-pong parities, the parameters,
 * the Jacobi plan and the object list; restoring it into a simulation of the
 * same dims and storage continues bit for bit as the original would.  Both
 * wait for the simulation's work. */
size_t smoke_snapshot_bytes(const smoke_sim* sim);
smoke_status smoke_snapshot(smoke_sim* sim, void* host, size_t bytes);
smoke_status smoke_restore(smoke_sim* sim, const void* host, size_t bytes);

/* Replaces the globals behind getGravity/getBuoyancy (smokeSimulation.cu:28-36)
 * and the compile-time constants (:38-39, :797). */
smoke_status smoke_set_params(smoke_sim* sim, const smoke_params* p);
smoke_status smoke_get_params(const smoke_sim* sim, smoke_params* out);

/* Launch plan of the fused Jacobi sweeps (SMOKE_PROJ_JACOBI): how many sweeps
 * one launch fuses (1..6), the thread-block tile and the z-chunk length.
 * Every plan gives bit-identical results; only the speed differs.  No
 * reference counterpart (the reference launches one kernel per half-sweep,
 * smokeSimulation.cu:797-801). */
/* Tiles 64x32 and 32x64 give each thread two rows of its column (fewer halo
 * cells per written cell, one 1024-thread block per CU instead of two). */
typedef enum smoke_tile {
    SMOKE_TILE_32x32 = 0,
    SMOKE_TILE_64x16 = 1,
    SMOKE_TILE_64x32 = 2,
    SMOKE_TILE_32x64 = 3
} smoke_tile;
typedef struct smoke_jacobi_plan {
    int sweeps_per_launch; /* 1..6 (default 4) */
    int tile;              /* smoke_tile (default 32x32) */
    int zc;                /* z planes per block, >= 1 (default 64) */
} smoke_jacobi_plan;
smoke_status smoke_set_jacobi_plan(smoke_sim* sim, const smoke_jacobi_plan* plan);
smoke_status smoke_get_jacobi_plan(const smoke_sim* sim, smoke_jacobi_plan* out);

/* Autotuner: times the projection's sweep sequence (the simulation's current
 * iteration count, or 100 if that is 0) on this simulation's buffers for
 * every candidate plan -- 2..6 sweeps per launch x the four tiles x z-chunks
 * of 32, 64 and 128 planes and of the interior's nz-2 planes split into 1..4
 * equal chunks -- and keeps the fastest.  With the REFINED projection it
 * times the refined solve instead (FP16 sweep launches of 2, 4 or 6 sweeps,
 * one-row 32x32 or two-row 64x32 tiles, the same z-chunks).  Results are identical for
 * every plan, so tuning never changes answers.  Resets the pressure kept for
 * warm starts to zero.  `chosen` (may be NULL) receives the plan and
 * `ms_per_sweep` (may be NULL) its measured time per sweep. */
smoke_status smoke_autotune(smoke_sim* sim, smoke_jacobi_plan* chosen, double* ms_per_sweep);

/* ---------------------------------------------------------------- scene */

/* Replace addObstacle / addSmokeSource (smokeSimulation.cu:88-104).  Ids are
 * shared by both kinds and count up from 0.  At most SMOKE_MAX_OBJECTS of each
 * kind (the reference silently overflows its device buffers beyond 3). */
#define SMOKE_MAX_OBJECTS 64
smoke_status smoke_add_obstacle(smoke_sim* sim, float x, float y, float z, float vx, float vy,
                                float vz, float r, int* id);
smoke_status smoke_add_source(smoke_sim* sim, float x, float y, float z, float r, int* id);

/* Replaces updateObjectPos (smokeSimulation.cu:106-109), with a bounds
 * check. */
smoke_status smoke_move_object(smoke_sim* sim, int id, float x, float y, float z);

/* Sets the velocity an object was added with (addObstacle's vx, vy, vz,
 * smokeSimulation.cu:88-97); read by the moving-obstacle boundary condition.
 * No reference counterpart: the reference cannot change it after adding. */
smoke_status smoke_set_object_velocity(smoke_sim* sim, int id, float vx, float vy, float vz);

/* Base solid mask, nx*ny*nz bytes, nonzero = solid.  Obstacles are combined
 * with it each step (reference: the mask is the y = 0 floor only,
 * smokeSimulation.cu:200-207, which is also the initial value here). */
smoke_status smoke_set_solids(smoke_sim* sim, const uint8_t* solid);

/* ---------------------------------------------------------------- fields */

/* Density of half h (0 or 1): nx*ny*nz floats. */
smoke_status smoke_set_density(smoke_sim* sim, int half, const float* src);
smoke_status smoke_get_density(smoke_sim* sim, int half, float* dst);

/* Velocity component comp (0 = u, 1 = v, 2 = w) of half h:
 * (nx+1)*(ny+1)*(nz+1) floats. */
smoke_status smoke_set_velocity(smoke_sim* sim, int comp, int half, const float* src);
smoke_status smoke_get_velocity(smoke_sim* sim, int comp, int half, float* dst);

/* Solid mask (1 = fluid, as the reference's dev_s) and cell flags
 * (bit 0 self fluid, bits 1-6 fluid neighbour x-,x+,y-,y+,z-,z+, bit 7 cell
 * relaxed by the projection), nx*ny*nz bytes each, as of the last scene
 * stage. */
smoke_status smoke_get_mask(smoke_sim* sim, uint8_t* dst);
smoke_status smoke_get_flags(smoke_sim* sim, uint8_t* dst);

/* Pressure of the last pressure-form projection, nx*ny*nz floats. */
smoke_status smoke_get_pressure(smoke_sim* sim, float* dst);

/* Ping-pong parity (indexNow, smokeSimulation.cu:707): after a step the
 * output density is half 1 - index_now. */
int smoke_index_now(const smoke_sim* sim);

/* ------------------------------------------------------------------ step */

/* Replaces simulate (smokeSimulation.cu:774-819): one full step with time
 * step dt, then, if density_out is not NULL, a blocking copy of the new
 * density (nx*ny*nz floats) into it. */
smoke_status smoke_step(smoke_sim* sim, float dt, float* density_out);

/* Enqueues one full step on the simulation's stream and returns. */
smoke_status smoke_step_async(smoke_sim* sim, float dt);

/* Step graphs (off by default): with enable != 0, smoke_step / _async capture
 * a step into a HIP graph, one per ping-pong parity, and replay it -- one
 * launch per step instead of one per kernel, for grids small enough that
 * launches bound the step.  Same bits as eager steps.  A graph is captured
 * again after any change of parameters, plan or objects, or of dt; the mask
 * refresh after an obstacle or solid change runs before the replay.  Steps
 * with timing or step norms enabled run eagerly.  No reference counterpart
 * (the reference launches every kernel, smokeSimulation.cu:774-819). */
smoke_status smoke_set_step_graph(smoke_sim* sim, int enable);

/* Enqueues one stage (smoke_stage) of a step; five stages in order equal one
 * step. */
smoke_status smoke_run_stage(smoke_sim* sim, int stage, float dt);

/* Blocking copy of the density produced by the last step (half
 * 1 - index_now). */
smoke_status smoke_read_density(smoke_sim* sim, float* dst);

/* Waits for all work enqueued on the simulation's stream. */
smoke_status smoke_synchronize(smoke_sim* sim);

/* ----------------------------------------------- asynchronous read-back */
/* The reference copies the density to the host synchronously after every
 * step (smokeSimulation.cu:814) and names that copy a bottleneck.  These two
 * calls overlap it with the following steps: begin() packs the density of the
 * last enqueued step on the device (the next step waits only for this pack)
 * and starts its copy into one of two pinned host buffers (used in turn);
 * wait() blocks until the oldest started copy has landed and returns its
 * buffer, which stays valid until the second begin() after the one that
 * produced it.  At most two read-backs may be pending: a third begin()
 * returns SMOKE_ERR_STATE until wait() is called.  Formats: FP32, or FP16 /
 * BF16 (half the bytes; round-to-nearest-even conversion of the FP32
 * density). */
typedef enum smoke_readback_format {
    SMOKE_READBACK_F32 = 0,
    SMOKE_READBACK_F16 = 1,
    SMOKE_READBACK_BF16 = 2
} smoke_readback_format;
smoke_status smoke_readback_begin(smoke_sim* sim, int format);
/* *host: nx*ny*nz values of the chosen format, x fastest; *step: 1-based
 * number of the step the density belongs to. */
smoke_status smoke_readback_wait(smoke_sim* sim, const void** host, long long* step);

/* ---------------------------------------------------------------- timing */

/* Per-stage device timing with HIP events recorded on the simulation's
 * stream.  While enabled, every step accumulates, per stage, the elapsed time
 * and the number of kernel launches of that stage. */
smoke_status smoke_timing_enable(smoke_sim* sim, int enable);
/* ms[SMOKE_STAGE_COUNT], launches[SMOKE_STAGE_COUNT] totals since the last
 * reset, and the number of timed steps. Synchronises the stream. */
smoke_status smoke_timing_read(smoke_sim* sim, double* ms, long long* launches, long long* steps);
smoke_status smoke_timing_reset(smoke_sim* sim);
/* Time between the first and the last pressure-sweep launch of every timed
 * step (the projection's sweep loop, the step's dominant kernel) and the
 * number of sweep launches in it, totals since the last reset. */
smoke_status smoke_timing_sweeps(smoke_sim* sim, double* ms, long long* launches);
/* Step latencies (ms) of the last `cap` timed steps, oldest first; *count
 * receives how many were written. */
smoke_status smoke_timing_step_ms(smoke_sim* sim, double* ms, int cap, int* count);

/* ------------------------------------------------------------ step norms */
/* Norms of the state a step leaves, reduced on the device in FP64 (sums) and
 * FP32 (maxima) in a fixed order, so they depend only on the fields and the
 * grid dims.  The reference computes none of them: it clamps face speeds
 * without reporting them (smokeSimulation.cu:339-349) and checks only the
 * final copy for errors (:815-818).  div_l2 is the velocity divergence over
 * the cells the projection relaxes (flags bit 7) after the projection (half
 * index_now, which advection reads and leaves alone), in the right-hand
 * side's operand order (-u0 + u1 + -v0 + v1 + -w0 + w1, :379-381); rhs_l2 the
 * same before it (the projection's right-hand side; REFERENCE computes it only
 * while norms are enabled, else -1).  Speeds and the density are those of the
 * step's output (half 1 - index_now); a component's faces on the grid walls
 * of its own axis (u at x = 0 and x = nx, ...) are never written and are left
 * out.  cfl = last dt x the largest speed: how many cells a face value's
 * back-trace can span next step. */
typedef struct smoke_step_norms {
    double div_l2;         /* sqrt(sum div^2), after the projection        */
    double rhs_l2;         /* sqrt(sum div^2), before it; -1 if unknown     */
    double density_sum;    /* sum of the output density                     */
    double cfl;            /* last dt * max(max_u, max_v, max_w)            */
    float max_u, max_v, max_w; /* max |u|, |v|, |w| over the output faces   */
    float max_density;     /* max |density|                                 */
    long long nonfinite;   /* NaN or Inf values in output density, u, v, w  */
    long long step;        /* steps enqueued when the norms were taken      */
} smoke_step_norms;
/* enable != 0: every following step ends with the reduction (two kernels on
 * the simulation's stream, no synchronisation; counted in the density
 * advection stage's timing), so norms cost no extra wait per step. */
smoke_status smoke_norms_enable(smoke_sim* sim, int enable);
/* Norms of the current state (the last step's, taken in-step when enabled,
 * otherwise reduced now).  Waits for the simulation's work.  Returns
 * SMOKE_ERR_NONFINITE, with *out filled, when nonfinite > 0. */
smoke_status smoke_get_step_norms(smoke_sim* sim, smoke_step_norms* out);
/* smoke_get_step_norms for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * reference counterpart.  Enqueues the reduction of the current state on the
 * simulation's stream and writes the record to `out_device` (caller-owned
 * device memory), without waiting and without allocating: its scratch (the
 * partial records) is a static device array of the library, one 192 KiB copy
 * per GPU, large enough for every grid (the reduction's first pass never
 * exceeds 4096 workgroups).  The record equals smoke_get_step_norms's bit for
 * bit; it never reuses an in-step result.  Read it after smoke_synchronize
 * (a non-finite field shows as nonfinite > 0; no status).  All calls on one
 * GPU share the scratch: calls that overlap on one GPU (simulations on
 * different streams) race, so serialise them per GPU.  SMOKE_ERR_STATE for a
 * simulation in a batch. */
smoke_status smoke_get_step_norms_noalloc(smoke_sim* sim, smoke_step_norms* out_device);

/* --------------------------------------------------------------- batches */
/* Many independent simulations of different sizes advanced together: one
 * kernel launch per stage for all of them (persistent kernels over a
 * (simulation, tile) work list), optionally replayed from a captured HIP
 * graph.  No reference counterpart (the reference runs one global
 * simulation).  Every simulation gives the same bits in a batch as alone.
 *
 * Requirements: at most SMOKE_MAX_BATCH simulations on the current device,
 * all with the Jacobi projection, cold start, and equal iteration count and
 * jacobi_omega; at most 8 sources each.  While a simulation belongs to a
 * batch its scene is frozen and it is stepped only through the batch (the
 * per-simulation step, scene and parameter calls return SMOKE_ERR_STATE);
 * field reads wait for the batch.  dts: one time step per simulation. */
#define SMOKE_MAX_BATCH 256
typedef struct smoke_batch smoke_batch;
smoke_status smoke_batch_create(smoke_sim* const* sims, int count, const float* dts,
                                smoke_batch** out);
void smoke_batch_destroy(smoke_batch* batch);
/* Enqueues one step of every simulation; use_graph != 0 captures the step
 * into a HIP graph on first use and replays it afterwards. */
smoke_status smoke_batch_step(smoke_batch* batch, int use_graph);
smoke_status smoke_batch_synchronize(smoke_batch* batch);
smoke_status smoke_batch_set_jacobi_plan(smoke_batch* batch, const smoke_jacobi_plan* plan);
smoke_status smoke_batch_get_jacobi_plan(const smoke_batch* batch, smoke_jacobi_plan* out);
/* Autotuner of the batched sweep sequence (2..6 sweeps per launch x the four
 * tiles x z-chunks of 32, 64, 128 and 256 planes); invalidates a captured
 * graph. */
smoke_status smoke_batch_autotune(smoke_batch* batch, smoke_jacobi_plan* chosen, double* ms_per_sweep);
/* Step latencies (ms, HIP events around each step on the batch's stream) of
 * the last `cap` steps, oldest first. */
smoke_status smoke_batch_step_ms(smoke_batch* batch, double* ms, int cap, int* count);
/* Cells of all simulations of the batch (for cell-updates/s). */
long long smoke_batch_cells(const smoke_batch* batch);
/* Interior cells (n-2 per axis) of all simulations: the sweeps' domain. */
long long smoke_batch_interior_cells(const smoke_batch* batch);
/* Times the batch's pressure-sweep sequence with its current plan (HIP
 * events on the batch's stream, one untimed sequence first, then `reps`):
 * average time per sweep launch and launches per sequence.  The sweeps run on
 * the last step's right-hand side; each step starts its solve from zero, so
 * this changes no later result. */
smoke_status smoke_batch_time_sweeps(smoke_batch* batch, int reps, double* ms_per_launch,
                                     int* launches);

/* --------------------------------------------------------------- serving */
/* A request loop over a stream of independent simulation requests
 * (continuous batching): each tick admits queued requests in arrival order
 * while their device footprint fits the HBM budget, steps every resident
 * request once in one batch (one launch per stage; replayed from a HIP graph
 * while the membership is unchanged), and retires requests that have run all
 * their steps (their final density is copied out and their memory freed).
 * A request's result is the same bits alone and in any mix.  No reference
 * counterpart.  All requests use the Jacobi projection with the server's
 * iteration count. */
typedef struct smoke_server smoke_server;
typedef struct smoke_request {
    unsigned nx, ny, nz;
    int steps;              /* >= 1 */
    float dt;
    float gravity, buoyancy, density_gate, max_velocity;
    const uint8_t* solid;   /* nx*ny*nz bytes, nonzero = solid; NULL: floor only */
    int n_sources;          /* 0..8 */
    const float* sources;   /* x, y, z, r per source */
    int n_obstacles;        /* 0..SMOKE_MAX_OBJECTS */
    const float* obstacles; /* x, y, z, r per obstacle */
    const float* velocity[3]; /* initial u, v, w ((nx+1)(ny+1)(nz+1)); NULL: zero */
} smoke_request;
typedef enum smoke_request_state {
    SMOKE_REQUEST_QUEUED = 0,
    SMOKE_REQUEST_RUNNING = 1,
    SMOKE_REQUEST_DONE = 2
} smoke_request_state;
typedef struct smoke_server_stats {
    long long submitted, admitted, completed, ticks;
    long long cell_updates;  /* cells x steps of all completed and running work */
    double gpu_ms;           /* GPU time of all batched steps */
    size_t hbm_used, hbm_budget, hbm_peak;
    int active, queued;
} smoke_server_stats;
smoke_status smoke_server_create(size_t hbm_budget_bytes, int iterations, smoke_server** out);
void smoke_server_destroy(smoke_server* server);
/* Copies the request's inputs and queues it; *id numbers requests from 0. */
smoke_status smoke_server_submit(smoke_server* server, const smoke_request* req, long long* id);
/* One scheduler iteration (admit, step, retire).  Returns how many requests
 * are running and queued after it. */
smoke_status smoke_server_tick(smoke_server* server, int* active, int* queued);
/* State of a request; when DONE and density is not NULL, copies its final
 * density (nx*ny*nz floats).  queue_ms / run_ms (may be NULL): host time from
 * submission to admission and from admission to completion. */
smoke_status smoke_server_result(smoke_server* server, long long id, int* state, float* density,
                                 double* queue_ms, double* run_ms);
smoke_status smoke_server_stats_get(smoke_server* server, smoke_server_stats* out);
/* GPU time (ms) of each of the last `cap` ticks' batched step and the number
 * of requests it advanced, oldest first. */
smoke_status smoke_server_tick_ms(smoke_server* server, double* ms, int* requests, int cap,
                                  int* count);

/* ------------------------------------------ proxy-plane slicing geometry */
/* The renderer's view-aligned slices of the volume's bounding box,
 * BoundingBox::generateProxyGeometry (project/boundingBox.cpp:133-183), on
 * the GPU: `steps` planes perpendicular to the view direction, from the
 * nearest to the farthest corner of a box of `box_dimensions` centred at the
 * origin (getMinMaxPoints, :231-241), each cut with the box at
 * `box_position` (planeBoxIntersections, :285-354), its corners ordered by
 * pseudoangle about their mean (sortPoints, :259-272) and fanned into
 * triangles (teselate, :243-255).  Output as the reference's three arrays:
 * vertices (3 floats each; per plane its corners, then their mean), 16-bit
 * triangle indices into the whole vertex array, and one smoke_proxy_plane
 * per plane (m_planeIndexing).  Bit-identical to the sequential CPU
 * computation in FP32.  A plane that misses the box yields its (NaN) mean
 * and no triangles (the reference's teselate is undefined there). */
typedef struct smoke_proxy_plane {
    int num_triangles;    /* corners of the cut (PlaneIndexValue, boundingBox.h:22-25) */
    int triangle_offset;  /* first triangle of the plane in the index array   */
} smoke_proxy_plane;

#define SMOKE_PROXY_MAX_STEPS 4096
#define SMOKE_PROXY_VERTICES_PER_PLANE 13  /* capacity per plane, worst case */
#define SMOKE_PROXY_INDICES_PER_PLANE 36

/* Host arrays: vertices >= 3*13*steps floats, indices >= 36*steps,
 * planes >= steps; *n_vertices and *n_indices receive the counts used.
 * Replaces the body of generateProxyGeometry; synchronous. */
smoke_status smoke_proxy_geometry(const float box_position[3], const float box_dimensions[3],
                                  const float view_origin[3], const float view_direction[3],
                                  int steps, float* vertices, int16_t* indices,
                                  smoke_proxy_plane* planes, int* n_vertices, int* n_indices);
/* Device arrays of the same capacities (e.g. buffers mapped from the
 * renderer), and `counts` = 2 device ints (vertices, indices).  Enqueued on
 * `stream` (a hipStream_t; NULL = the default stream); returns without
 * waiting.  Allocates nothing. */
smoke_status smoke_proxy_geometry_device(const float box_position[3], const float box_dimensions[3],
                                         const float view_origin[3], const float view_direction[3],
                                         int steps, float* vertices, int16_t* indices,
                                         smoke_proxy_plane* planes, int* counts, void* stream);

/* ------------------------------------------------------ legacy interface */
/* Process-global single simulation behind the reference's nine functions
 * (smokeSimulation.cuh:4-17).  Used by smokeSimulation.cuh; errors are
 * returned here and turned into the reference's print-and-exit there. */
smoke_status smoke_legacy_device_info(void);                                   /* getGPUProperties */
smoke_status smoke_legacy_initialize(float* grid, unsigned w, unsigned h, unsigned d); /* initializeVolume */
smoke_status smoke_legacy_delete(void);                                         /* deleteVolume */
smoke_status smoke_legacy_simulate(float* grid, float dt);                      /* simulate */
int smoke_legacy_add_obstacle(float x, float y, float z, float vx, float vy, float vz, float r); /* addObstacle */
int smoke_legacy_add_source(float x, float y, float z, float r);                /* addSmokeSource */
smoke_status smoke_legacy_update_object_pos(int id, float x, float y, float z); /* updateObjectPos */
float* smoke_legacy_buoyancy_ptr(void);                                          /* getBuoyancy */
float* smoke_legacy_gravity_ptr(void);                                           /* getGravity */

#ifdef __cplusplus
}
#endif

#endif /* SMOKE_H */
