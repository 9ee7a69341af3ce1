// Internal declarations shared by the host runtime and the HIP kernels of
// libsmokesim.so.  Not part of the public ABI (include/smoke.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "smoke.h"

namespace smk {

// One padded layout for every field.  Cell (x,y,z) and the faces on its low
// x/y/z sides share the index x + px*y + plane*z, so a kernel addresses
// density, mask, pressure and u/v/w with one index.  Rows are padded to a
// multiple of 32 elements (128 B of FP32) so every row starts on a cache-line
// boundary; planes hold ny+1 rows and there are nz+1 planes, the extents of
// the reference's staggered arrays (smokeSimulation.cu:146-147).
struct Grid {
    int nx, ny, nz;
    int px;     // row pitch in elements, >= nx+1, multiple of 32
    int py;     // rows per plane, ny+1
    int plane;  // px*py
    int elems;  // plane*(nz+1)
    __host__ __device__ __forceinline__ int idx(int x, int y, int z) const
    {
        return x + px * y + plane * z;
    }
};

// Thread block of the per-cell kernels (64 x 4: one wave64 per row segment).
constexpr int kBlockX = 64;
constexpr int kBlockY = 4;

// 16-bit storage elements of the mixed-precision forms (arithmetic stays
// FP32): BF16 (the upper half of an FP32 word) and IEEE binary16.
struct bf16s {
    uint16_t bits;
};
struct f16s {
    uint16_t bits;
};

// Device-side state of the mixed-precision pressure solve of one simulation
// (refine_kernels.hip): FP64 norms, the power-of-two scale of the current
// correction and the convergence decision.
struct RefineCtl {
    double bnorm2;     // ||b||^2 of this step (b = divergence on active cells)
    double rnorm2;     // ||r||^2 at the last residual pass
    double rel;        // sqrt(rnorm2 / bnorm2)
    float s, sinv;     // correction scale 2^k and its inverse
    unsigned qmax;     // max |r| as FP32 bits (order-preserving for r >= 0)
    int done;          // relative residual reached: later passes of the step idle
    int refinements;   // refinements run this step
    int pad;
};

// Cell flag bits (one byte per cell).
enum : uint8_t {
    F_SELF = 1,
    F_XM = 2,
    F_XP = 4,
    F_YM = 8,
    F_YP = 16,
    F_ZM = 32,
    F_ZP = 64,
    F_ACTIVE = 128,
};
constexpr uint8_t F_NBRS = F_XM | F_XP | F_YM | F_YP | F_ZM | F_ZP;

// Spheres passed by value as a kernel argument: x, y, z, radius.
struct SphereList {
    int count;
    float4 s[SMOKE_MAX_OBJECTS];
};

struct ForceArgs {
    float gravity, buoyancy, dt, max_v2, gate;
    int gate_on;
};

// Launch plan of the fused Jacobi kernel: sweeps fused per launch (2..6),
// thread-block tile and z-chunk length.  Every plan gives bit-identical
// results; the choice only changes speed.
// The fused sweeps address one block's span of planes with 32-bit buffer
// offsets below 2^31 bytes: a z-chunk of zc planes plus at most
// kFusedSlackPlanes more.  Largest usable zc for a plane of `plane` elements
// (< 1: no fused launch fits; the sweeps then run one per launch).
constexpr int kFusedSlackPlanes = 48;
__host__ __device__ inline int fused_max_zc(long long plane)
{
    const long long planes = 2147483647LL / (plane * 4LL);
    return (int)(planes > 1000000 ? 1000000 : planes) - kFusedSlackPlanes;
}

struct JacobiPlan {
    // Tiles 64x32 and 32x64 give each thread two rows of its column.
    enum Tile { TILE_32x32 = 0, TILE_64x16 = 1, TILE_64x32 = 2, TILE_32x64 = 3 };
    int sweeps_per_launch = 4;
    int tile = TILE_32x32;
    int zc = 64;
};
inline int tile_width(int tile) { return (tile == JacobiPlan::TILE_32x32 || tile == JacobiPlan::TILE_32x64) ? 32 : 64; }
inline int tile_height(int tile)
{
    return tile == JacobiPlan::TILE_64x16 ? 16 : (tile == JacobiPlan::TILE_32x64 ? 64 : 32);
}

// ------------------------------------------------------------- batching

// One simulation as the batched kernels see it (device memory).  `now` and
// `pcur` are flipped on the device at the end of every batched step, so a
// captured step replays correctly.
struct SimDesc {
    Grid g;
    float* rho[2];
    float* u[2];
    float* v[2];
    float* w[2];
    float* P[2];
    float* d;
    const uint8_t* flags;
    ForceArgs force;  // includes this simulation's dt
    int now;          // ping-pong parity (indexNow)
    int pcur;         // P buffer holding the latest pressure
};

// Consecutive z planes per thread in the per-cell streaming stages (force,
// divergence, gradient, density advection), single and batched.
constexpr int kColZ = 4;

// Planes per column of the velocity advection (a block of kBlockX x kBlockY
// cells marches up z through kAdvChunkZ planes).
constexpr int kAdvChunkZ = 16;

// Work list of one batched stage: simulation s owns virtual blocks
// [start[s], start[s+1]) arranged as a gx[s] x gy[s] x (rest) block grid.
constexpr int kMaxBatch = 256;
struct StageTable {
    int start[kMaxBatch + 1];
    int gx[kMaxBatch];
    int gy[kMaxBatch];
};

// Source fills of a batch: item k fills one source sphere of simulation
// sim[k] over the box lo[k] + [0, ext[k]) (clipped to the interior).
constexpr int kMaxBatchSources = kMaxBatch * 8;
struct SourceTable {
    int count;
    int start[kMaxBatchSources + 1];  // virtual blocks per item (prefix)
    int sim[kMaxBatchSources];
    float4 sph[kMaxBatchSources];
    int4 lo[kMaxBatchSources];   // x, y, z of the box corner
    int4 ext[kMaxBatchSources];  // ex, ey, ez, blocks in x
};

enum BatchStage { BS_FORCE = 0, BS_DIV = 1, BS_GRAD = 2, BS_ADV_V = 3, BS_ADV_D = 4 };

// Batched launchers (batch_kernels.hip).  `grid` is the persistent grid size
// (a multiple of 8); `lp` is the parity of this launch within the step's
// sweep sequence (input P = P[pcur ^ lp]).
hipError_t batch_sources(const SimDesc* sims, const SourceTable* tab, int total_blocks, int grid,
                         hipStream_t st);
hipError_t batch_cells(int stage, const SimDesc* sims, const StageTable* tab, int nsims,
                       int total_blocks, int grid, int lp, hipStream_t st);
hipError_t batch_jacobi(const SimDesc* sims, const StageTable* tab, int nsims, int total_blocks,
                        int grid, const JacobiPlan& plan, int sweeps, int lp, int zero_in,
                        float omega, int plain, hipStream_t st);
hipError_t batch_flip(SimDesc* sims, int nsims, int pflip, hipStream_t st);
// Resident blocks per CU of the batched kernels (for sizing persistent grids).
int batch_cells_blocks_per_cu();
// Velocity advection of every simulation in columns (kBlockX x kBlockY x
// kAdvChunkZ cells over [1, n-1]^3, table built with those columns).
hipError_t batch_advect_velocity(const SimDesc* sims, const StageTable* tab, int nsims,
                                 int total_blocks, int grid, hipStream_t st);
int batch_advect_velocity_blocks_per_cu();
// Density advection of every simulation in columns (kBlockX x kBlockY x
// kAdvChunkZ cells over [1, n-2]^3, x from 0; table built with those columns).
hipError_t batch_advect_density(const SimDesc* sims, const StageTable* tab, int nsims,
                                int total_blocks, int grid, hipStream_t st);
int batch_advect_density_blocks_per_cu();
int batch_jacobi_blocks_per_cu(const JacobiPlan& plan, int sweeps);

// ------------------------------------------------------------ launchers
// Every launcher enqueues on `st` and returns hipGetLastError().

hipError_t fill_f32(float* p, long long n, float value, hipStream_t st);
hipError_t fill_u8(uint8_t* p, long long n, uint8_t value, hipStream_t st);

// Contiguous (reference layout) <-> padded layout.
hipError_t unpack_cells_f32(const Grid& g, const float* src, float* dst, hipStream_t st);
hipError_t pack_cells_f32(const Grid& g, const float* src, float* dst, hipStream_t st);
hipError_t unpack_faces_f32(const Grid& g, const float* src, float* dst, hipStream_t st);
hipError_t pack_faces_f32(const Grid& g, const float* src, float* dst, hipStream_t st);
hipError_t unpack_cells_solid(const Grid& g, const uint8_t* solid, uint8_t* base, hipStream_t st);
hipError_t pack_cells_u8(const Grid& g, const uint8_t* src, uint8_t* dst, hipStream_t st);
// Density to contiguous FP16 (format 1) or BF16 (format 2), round to nearest even.
hipError_t pack_cells_half(const Grid& g, const float* src, uint16_t* dst, int format,
                           hipStream_t st);
hipError_t floor_mask(const Grid& g, uint8_t* base, hipStream_t st);

// Scene: solid mask from base + obstacles, then flags; sources.
hipError_t build_mask(const Grid& g, const uint8_t* base, uint8_t* s, const SphereList& obst,
                      int mode, hipStream_t st);
hipError_t build_flags(const Grid& g, const uint8_t* s, uint8_t* flags, hipStream_t st);
hipError_t fill_source(const Grid& g, float* rho0, float* rho1, float4 sph, hipStream_t st,
                       int* launched);
// Conservative bounding box of a source sphere clipped to the interior:
// false if it holds no cell.
bool source_box(const Grid& g, float4 sph, int lo[3], int ext[3]);

// Force: buoyancy + clamp on [now].
hipError_t force(const Grid& g, float* u, float* v, float* w, const float* rho,
                 const uint8_t* flags, const ForceArgs& a, hipStream_t st);

// Moving-obstacle velocity boundary condition (an extension; the reference
// stores obstacle velocities and never reads them).  build_owner records per
// cell the index in the obstacle list of the obstacle that made it solid
// (kNoOwner: none), under the same rule as build_mask; obstacle_velocity_bc
// then sets every face of [now] next to an owned cell to the owner's velocity
// (the high-side cell's owner first).  `vel` holds (vx, vy, vz, 0) per
// obstacle, in obstacle-list order.
constexpr uint8_t kNoOwner = 255;
hipError_t build_owner(const Grid& g, uint8_t* owner, const SphereList& obst, int mode,
                       hipStream_t st);
hipError_t obstacle_velocity_bc(const Grid& g, float* u, float* v, float* w, const uint8_t* owner,
                                const SphereList& vel, hipStream_t st);

// Proxy-plane slicing geometry (proxy_kernels.hip): one workgroup builds
// every plane of one view.  With `done` non-NULL the kernel stores `seq`
// there (system-scope release) once every output is visible to the host.
struct ProxyView {
    float box_pos[3], box_dims[3], origin[3], dir[3];
};
hipError_t proxy_geometry(const ProxyView& view, int steps, float* verts, int16_t* indices,
                          smoke_proxy_plane* planes, int* counts, hipStream_t st,
                          unsigned* done = nullptr, unsigned seq = 0);

// Vorticity confinement on [now], after the force stage (an extension; three
// passes over four FP32 padded scratch fields).
struct VortBufs {
    float* mag;   // |omega| per cell
    float* f[3];  // omega, then the confinement force, per cell
};
hipError_t vorticity(const Grid& g, float* u, float* v, float* w, const uint8_t* flags,
                     const VortBufs& b, float eps, float dt, hipStream_t st);

// Projection.
hipError_t project_reference_half(const Grid& g, float* u, float* v, float* w,
                                  const uint8_t* flags, int colour, double omega, hipStream_t st);
hipError_t divergence(const Grid& g, const float* u, const float* v, const float* w,
                      const uint8_t* flags, float* d, hipStream_t st);
hipError_t jacobi_sweep(const Grid& g, const float* pin, float* pout, const float* d,
                        const uint8_t* flags, float omega, int plain, int zero_in, hipStream_t st);
// `sweeps` (1..6) Jacobi sweeps in one launch (1: the single-sweep kernel).
hipError_t jacobi_fused(const Grid& g, const float* pin, float* pout, const float* d,
                        const uint8_t* flags, const JacobiPlan& plan, int sweeps, float omega,
                        int plain, int zero_in, hipStream_t st);
hipError_t rbsor_half(const Grid& g, float* p, const float* d, const uint8_t* flags, int colour,
                      float omega, hipStream_t st);
hipError_t mask_pressure(const Grid& g, float* p, const uint8_t* flags, hipStream_t st);
hipError_t gradient(const Grid& g, float* u, float* v, float* w, const float* p,
                    const uint8_t* flags, hipStream_t st);

// Advection.
hipError_t advect_velocity(const Grid& g, const float* u0, const float* v0, const float* w0,
                           float* u1, float* v1, float* w1, const uint8_t* flags, float dt,
                           hipStream_t st);
hipError_t advect_density(const Grid& g, const float* r0, float* r1, const float* u,
                          const float* v, const float* w, const uint8_t* flags, float dt,
                          hipStream_t st);

// 16-bit stored fields H (bf16s or f16s; FP32 arithmetic): the same stages
// and conversions, with density stored as H or as FP32 (the split forms).
// The step-norms reduction is declared with the FP32 one below.
#define SMK_HALF_STORAGE_DECLS(H)                                                                  \
    hipError_t unpack_cells_f32(const Grid& g, const float* src, H* dst, hipStream_t st);          \
    hipError_t pack_cells_f32(const Grid& g, const H* src, float* dst, hipStream_t st);            \
    hipError_t unpack_faces_f32(const Grid& g, const float* src, H* dst, hipStream_t st);          \
    hipError_t pack_faces_f32(const Grid& g, const H* src, float* dst, hipStream_t st);            \
    hipError_t pack_cells_half(const Grid& g, const H* src, uint16_t* dst, int format,             \
                               hipStream_t st);                                                    \
    hipError_t fill_source(const Grid& g, H* rho0, H* rho1, float4 sph, hipStream_t st,            \
                           int* launched);                                                         \
    hipError_t force(const Grid& g, H* u, H* v, H* w, const H* rho, const uint8_t* flags,          \
                     const ForceArgs& a, hipStream_t st);                                          \
    hipError_t force(const Grid& g, H* u, H* v, H* w, const float* rho, const uint8_t* flags,      \
                     const ForceArgs& a, hipStream_t st);                                          \
    hipError_t obstacle_velocity_bc(const Grid& g, H* u, H* v, H* w, const uint8_t* owner,         \
                                    const SphereList& vel, hipStream_t st);                        \
    hipError_t vorticity(const Grid& g, H* u, H* v, H* w, const uint8_t* flags,                    \
                         const VortBufs& b, float eps, float dt, hipStream_t st);                  \
    hipError_t divergence(const Grid& g, const H* u, const H* v, const H* w,                       \
                          const uint8_t* flags, float* d, hipStream_t st);                         \
    hipError_t gradient(const Grid& g, H* u, H* v, H* w, const float* p, const uint8_t* flags,     \
                        hipStream_t st);                                                           \
    hipError_t advect_velocity(const Grid& g, const H* u0, const H* v0, const H* w0, H* u1,        \
                               H* v1, H* w1, const uint8_t* flags, float dt, hipStream_t st);      \
    hipError_t advect_density(const Grid& g, const H* r0, H* r1, const H* u, const H* v,           \
                              const H* w, const uint8_t* flags, float dt, hipStream_t st);         \
    hipError_t advect_density(const Grid& g, const float* r0, float* r1, const H* u, const H* v,   \
                              const H* w, const uint8_t* flags, float dt, hipStream_t st);
SMK_HALF_STORAGE_DECLS(bf16s)
SMK_HALF_STORAGE_DECLS(f16s)
#undef SMK_HALF_STORAGE_DECLS

// ------------------------------------------------------------ step norms
// Result record of the norms reduction (norms_kernels.hip), read back by the
// host: FP64 sums of div(u[now])^2 and d^2 over relaxed cells and of the
// output density; FP32 maxima of |u|, |v|, |w| over live faces of the output
// half and of |density|; the count of non-finite output values.
struct StepNorms {
    double div2, rhs2, rho_sum;
    float umax, vmax, wmax, rho_max;
    unsigned long long nonfinite;
};
// Workgroups of the first reduction pass (at most): the scratch holds that
// many partial records plus the reduced one (norms_scratch_bytes()).
constexpr int kNormsBlocks = 4096;
size_t norms_scratch_bytes();
// Enqueues the reduction (partials, fixed-order total, then the public
// record written to `out`, device memory).  rhs = 0: d is not read (no
// right-hand side of this step is known; rhs_l2 = -1).  dt and step fill the
// record's cfl and step.
hipError_t step_norms(const Grid& g, const float* rho, const float* u, const float* v,
                      const float* w, const float* un, const float* vn, const float* wn,
                      const float* d, int rhs, const uint8_t* flags, void* scratch,
                      smoke_step_norms* out, float dt, long long step, hipStream_t st);
#define SMK_HALF_NORMS_DECLS(H)                                                                    \
    hipError_t step_norms(const Grid& g, const H* rho, const H* u, const H* v, const H* w,         \
                          const H* un, const H* vn, const H* wn, const float* d, int rhs,          \
                          const uint8_t* flags, void* scratch, smoke_step_norms* out, float dt,    \
                          long long step, hipStream_t st);                                         \
    hipError_t step_norms(const Grid& g, const float* rho, const H* u, const H* v, const H* w,     \
                          const H* un, const H* vn, const H* wn, const float* d, int rhs,          \
                          const uint8_t* flags, void* scratch, smoke_step_norms* out, float dt,    \
                          long long step, hipStream_t st);
SMK_HALF_NORMS_DECLS(bf16s)
SMK_HALF_NORMS_DECLS(f16s)
#undef SMK_HALF_NORMS_DECLS
// Device address (on `device`) of the library's static norms scratch, for
// the allocation-free entry point; resolved once per device.
hipError_t noalloc_norms_scratch(int device, void** out);
size_t noalloc_norms_scratch_bytes();

// ------------------------------------------- mixed-precision pressure solve
// Device buffers of the refinement loop (refine_kernels.hip), padded layout.
// e and qi hold FP16 values, two cells per 32-bit word.  The sweeps take the
// diagonal inverse from the flags byte (a table in LDS).
struct RefineBufs {
    float* q = nullptr;            // negated FP32 residual of the current refinement
    uint32_t* e[2] = {nullptr, nullptr};  // FP16 correction, ping-pong
    uint32_t* qi = nullptr;        // FP16 inv * r * s
    RefineCtl* ctl = nullptr;
    double* part = nullptr;        // per-block FP64 partial sums of r^2
    unsigned* pmax = nullptr;      // per-block max |r| (FP32 bits)
};
// First-stage partials of the two-stage norm reduction (at most).
constexpr int kRefineStage2 = 1024;
// Blocks (partials) of a residual pass over g; the partial buffers hold
// refine_parts(g) + kRefineStage2 entries.
int refine_parts(const Grid& g);
// Norms of d (the residual of x = 0) and the step's control reset.
hipError_t refine_norm0(const Grid& g, const float* d, const uint8_t* flags, const RefineBufs& b,
                        double tol, hipStream_t st);
// `sweeps` (2, 4 or 6) FP16 sweeps e[ein] -> e[eout]; `first`: e starts at 0
// and the scaled right-hand side is formed from the FP32 q.  two_rows: tiles
// of 64 rows with two rows per thread (else 32 rows, one per thread); every
// choice gives the same bits.
hipError_t refine_sweeps(const Grid& g, const RefineBufs& b, const float* q, const uint8_t* flags,
                         int first, int ein, int eout, int sweeps, int zc, int two_rows,
                         hipStream_t st);
// x_new = x_old + e[ein]/s and the next residual (norms, control update for
// refinement `stage`).
hipError_t refine_resid(const Grid& g, const float* xo, float* xn, const float* d,
                        const uint8_t* flags, const RefineBufs& b, int ein, int xzero, int last,
                        int stage, double tol, hipStream_t st);

}  // namespace smk
