// Device-side arithmetic of one smoke step, shared by the single-simulation
// kernels (step_kernels.hip) and the batched persistent kernels
// (batch_kernels.hip), so a simulation gives the same bits alone and in a
// batch.
//
// Every function evaluates each output with the same operations, in the same
// order, as the CPU oracle (oracle/smoke_oracle.cpp), which restates
// project/smokeSimulation.cu.  Files including this header are compiled with
// -ffp-contract=off (no multiply-add is fused behind the source's back) and
// HIP's correctly rounded FP32 division, so the results are bit-identical to
// the oracle's FP32 mode.
//
// The *_at functions take the cell/face coordinates a thread owns and check
// their own range; the caller maps threads to coordinates.
#pragma once

#include "smoke_internal.h"

namespace smk {
namespace dev {

// Thread block of the per-cell kernels: one wave64 covers 64 consecutive x of
// one row (one contiguous 256-byte run of a 128-byte-aligned FP32 row).
constexpr int BX = kBlockX;
constexpr int BY = kBlockY;

// Sphere membership as fillSmoke/fillObstacle test it
// (smokeSimulation.cu:265, :304): integer cell index, float centre, squares
// and sums in float, strict comparison with r*r.
__device__ __forceinline__ bool inside(float4 s, int x, int y, int z)
{
    float dx = (float)x - s.x;
    float dy = (float)y - s.y;
    float dz = (float)z - s.z;
    float dist = dx * dx + dy * dy + dz * dz;
    return dist < s.w * s.w;
}

__device__ __forceinline__ bool is_interior(const Grid& g, int x, int y, int z)
{
    return x >= 1 && x <= g.nx - 2 && y >= 1 && y <= g.ny - 2 && z >= 1 && z <= g.nz - 2;
}

// Field storage: FP32, or BF16 / FP16 (mixed-precision forms) widened to FP32
// on load (exactly) and rounded to nearest even on store.  The step bodies
// below are templates over the element type; with float they are plain loads
// and stores.
__device__ __forceinline__ float ldf(const float* __restrict__ p, int i) { return p[i]; }
__device__ __forceinline__ float ldf(const bf16s* __restrict__ p, int i)
{
    return __uint_as_float((uint32_t)p[i].bits << 16);
}
__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    const uint32_t b = __float_as_uint(f);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x0040u);  // quiet NaN
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ void stf(float* __restrict__ p, int i, float v) { p[i] = v; }
__device__ __forceinline__ void stf(bf16s* __restrict__ p, int i, float v)
{
    p[i].bits = bf16_rne(v);
}
// An FP32 value the compiler must treat as opaque: the FP16 conversion that
// follows stays a conversion of this FP32 value.  Without it the compiler may
// fuse the multiply or add that produced the value into the conversion
// (v_fma_mix*_f16), which rounds the exact result once to FP16 instead of
// the FP32 result, so the same source would give different bits in
// different kernels.
#ifndef SMK_OPAQUE_F32
#define SMK_OPAQUE_F32(x) asm volatile("" : "+v"(x))
#endif
__device__ __forceinline__ _Float16 to_f16(float v)
{
    SMK_OPAQUE_F32(v);
    return (_Float16)v;
}
// binary16: the hardware conversions (FP32 -> FP16 rounds to nearest even,
// keeps subnormals, overflows to Inf and keeps NaN).
__device__ __forceinline__ float ldf(const f16s* __restrict__ p, int i)
{
    return (float)__builtin_bit_cast(_Float16, p[i].bits);
}
__device__ __forceinline__ void stf(f16s* __restrict__ p, int i, float v)
{
    p[i].bits = __builtin_bit_cast(uint16_t, to_f16(v));
}

// fillSmoke (smokeSimulation.cu:251-273) for one cell the caller knows to be
// interior: density 1.0 in both halves if the cell is inside the source.
template <typename T>
__device__ __forceinline__ void source_at(const Grid& g, T* __restrict__ r0, T* __restrict__ r1,
                                          float4 sph, int x, int y, int z)
{
    if (inside(sph, x, y, z)) {
        int i = g.idx(x, y, z);
        stf(r0, i, 1.0f);
        stf(r1, i, 1.0f);
    }
}

// integrate (smokeSimulation.cu:315-329) and velocityConfinement (:331-352)
// fused per staggered index: the buoyancy range is x in [0,nx-1],
// y in [1,ny-1], z in [0,nz-1]; the clamp range x,y,z in [1,n-1] lies inside
// it and sees the post-buoyancy v of the same index.  All loads of the
// per-cell bodies are issued unconditionally (every address is inside the
// padded field for the cells the range checks admit), so they do not wait
// for the flags; only the stores are predicated.
//
// Each body is split into a load phase and an apply phase; the *_col
// versions run the load phase of NZ cells of one (x, y) column first, so a
// thread keeps all their loads in flight, then apply them in z order (the
// cells of a stage are independent).  *_at is the one-cell case.
struct ForceIn {
    float v, r, u, w;
    uint8_t fl;
    bool ok;
};
template <typename T, typename TR>
__device__ __forceinline__ ForceIn force_load(const Grid& g, const T* __restrict__ u,
                                              const T* __restrict__ v, const T* __restrict__ w,
                                              const TR* __restrict__ rho,
                                              const uint8_t* __restrict__ flags, int x, int y, int z)
{
    ForceIn in{0.0f, 0.0f, 0.0f, 0.0f, 0, false};
    if (x >= g.nx || y < 1 || y >= g.ny || z >= g.nz) return in;
    const int i = g.idx(x, y, z);
    in.ok = true;
    in.fl = flags[i];
    in.v = ldf(v, i);
    in.r = ldf(rho, i);
    in.u = ldf(u, i);
    in.w = ldf(w, i);
    return in;
}
template <typename T>
__device__ __forceinline__ void force_apply(const Grid& g, T* __restrict__ u, T* __restrict__ v,
                                            T* __restrict__ w, const ForceArgs& a,
                                            const ForceIn& in, int x, int y, int z)
{
    if (!in.ok) return;
    const int i = g.idx(x, y, z);
    const float r = in.r, uu = in.u, ww = in.w;
    float vv = in.v;
    bool write_v = false;
    if ((in.fl & F_SELF) && (in.fl & F_YM)) {
        if (!a.gate_on || r > a.gate) {
            float b = a.buoyancy * (r - 0.0f);
            vv = vv + (r * a.gravity * a.dt + b * a.dt);
            write_v = true;
        }
    }
    if (x >= 1 && z >= 1) {
        float L = uu * uu + vv * vv + ww * ww;
        if (L * a.dt > a.max_v2) {
            stf(u, i, uu * (a.max_v2 / (L * a.dt)));
            vv = vv * (a.max_v2 / (L * a.dt));
            stf(w, i, ww * (a.max_v2 / (L * a.dt)));
            write_v = true;
        }
    }
    if (write_v) stf(v, i, vv);
}
template <int NZ, typename T, typename TR>
__device__ __forceinline__ void force_col(const Grid& g, T* __restrict__ u, T* __restrict__ v,
                                          T* __restrict__ w, const TR* __restrict__ rho,
                                          const uint8_t* __restrict__ flags, const ForceArgs& a,
                                          int x, int y, int z0)
{
    ForceIn in[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) in[j] = force_load(g, u, v, w, rho, flags, x, y, z0 + j);
#pragma unroll
    for (int j = 0; j < NZ; ++j) force_apply(g, u, v, w, a, in[j], x, y, z0 + j);
}
template <typename T, typename TR>
__device__ __forceinline__ void force_at(const Grid& g, T* __restrict__ u, T* __restrict__ v,
                                         T* __restrict__ w, const TR* __restrict__ rho,
                                         const uint8_t* __restrict__ flags, const ForceArgs& a,
                                         int x, int y, int z)
{
    force_col<1>(g, u, v, w, rho, flags, a, x, y, z);
}

// Colour-o cells of row (y,z): x = 2t + 1 + ((y+z+o+1) & 1) enumerates the
// interior x with (x+y+z+o) even, as smokeSimulation.cu:362 does.
__device__ __forceinline__ int colour_x(int t, int y, int z, int o)
{
    return 2 * t + 1 + ((y + z + o + 1) & 1);
}

// Right-hand side of the pressure form: d = div(u*) on active cells, in the
// reference's operand order (smokeSimulation.cu:379-381); 0 elsewhere.
template <int NZ, typename T>
__device__ __forceinline__ void divergence_col(const Grid& g, const T* __restrict__ u,
                                               const T* __restrict__ v, const T* __restrict__ w,
                                               const uint8_t* __restrict__ flags,
                                               float* __restrict__ d, int x, int y, int z0)
{
    if (x < 1 || x > g.nx - 2 || y < 1 || y > g.ny - 2) return;
    float dv[NZ];
    bool act[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        const int z = z0 + j;
        dv[j] = 0.0f;
        act[j] = false;
        if (z < 1 || z > g.nz - 2) continue;
        const int i = g.idx(x, y, z);
        const float u0 = ldf(u, i), u1 = ldf(u, i + 1), v0 = ldf(v, i), v1 = ldf(v, i + g.px),
                    w0 = ldf(w, i), w1 = ldf(w, i + g.plane);
        act[j] = (flags[i] & F_ACTIVE) != 0;
        dv[j] = -u0 + u1 + -v0 + v1 + -w0 + w1;
    }
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        const int z = z0 + j;
        if (z < 1 || z > g.nz - 2) continue;
        d[g.idx(x, y, z)] = act[j] ? dv[j] : 0.0f;
    }
}
template <typename T>
__device__ __forceinline__ void divergence_at(const Grid& g, const T* __restrict__ u,
                                              const T* __restrict__ v, const T* __restrict__ w,
                                              const uint8_t* __restrict__ flags,
                                              float* __restrict__ d, int x, int y, int z)
{
    divergence_col<1>(g, u, v, w, flags, d, x, y, z);
}

// RN(1/acc) for acc = 1..6: compile-time constants (correctly rounded by the
// compiler) picked with selects, so no memory is touched.
__device__ __forceinline__ float inv_acc(int acc)
{
    float r = 1.0f;
    r = (acc == 2) ? 0.5f : r;
    r = (acc == 3) ? (1.0f / 3.0f) : r;
    r = (acc == 4) ? 0.25f : r;
    r = (acc == 5) ? 0.2f : r;
    r = (acc == 6) ? (1.0f / 6.0f) : r;
    return r;
}

// Diagonal inverse of a cell, 0 for cells the projection does not relax.
__device__ __forceinline__ float cell_inv(uint8_t fl)
{
    return (fl & F_ACTIVE) ? inv_acc(__popc(fl & F_NBRS)) : 0.0f;
}

// Neighbour sum of the pressure forms at an active cell: all six neighbours
// left to right (x-, x+, y-, y+, z-, z+), as the oracle defines it.  The
// pressure is exactly +0 on every cell that is not active (ring, solid and
// isolated cells are written 0 and never relaxed), so this is the sum over the
// fluid neighbours.
__device__ __forceinline__ float sum6(float a0, float a1, float a2, float a3, float a4, float a5)
{
    float sum = a0 + a1;
    sum = sum + a2;
    sum = sum + a3;
    sum = sum + a4;
    sum = sum + a5;
    return sum;
}

// The pressure forms' update of an active cell, G = (sum - d) * inv, as the
// oracle defines it: one fused multiply-add fma(sum, inv, 0 - d*inv).  The
// negated scaled right-hand side 0 - d*inv is a per-cell constant of a solve.
__device__ __forceinline__ float neg_scaled_rhs(float d, float inv) { return 0.0f - d * inv; }
__device__ __forceinline__ float pressure_update(float sum, float inv, float d)
{
    return __builtin_fmaf(sum, inv, neg_scaled_rhs(d, inv));
}

// One (damped) Jacobi sweep of acc_i P_i - sum_n s_n P_n = -d_i at an interior
// cell; inactive cells are written 0.  zero_in: the input is all zeros (cold
// start), so it is not read.
__device__ __forceinline__ void jacobi_at(const Grid& g, const float* __restrict__ pin,
                                          float* __restrict__ pout, const float* __restrict__ d,
                                          const uint8_t* __restrict__ flags, float omega,
                                          int plain, int zero_in, int x, int y, int z)
{
    if (x < 1 || x > g.nx - 2 || y < 1 || y > g.ny - 2 || z < 1 || z > g.nz - 2) return;
    int i = g.idx(x, y, z);
    float inv = cell_inv(flags[i]);
    float sum = 0.0f, pi = 0.0f;
    if (!zero_in) {
        sum = sum6(pin[i - 1], pin[i + 1], pin[i - g.px], pin[i + g.px], pin[i - g.plane],
                   pin[i + g.plane]);
        pi = pin[i];
    }
    const float G = pressure_update(sum, inv, d[i]);
    float out = plain ? G : pi + omega * (G - pi);
    pout[i] = (inv > 0.0f) ? out : 0.0f;
}

// Planes in flight per thread of the z-streaming loops of the fused sweeps:
// one loop iteration (one plane, one barrier) is much shorter than the HBM
// latency under load, so loads are issued kPrefetch iterations ahead.
constexpr int kPrefetch = 3;
// Unroll factor of those loops: the register windows rotate with period 3 (z
// neighbours), T+1 (per-plane diagonal and right-hand side) and kPrefetch
// (the prefetch ring), so after the lcm of the three every register is back
// in its place and the unrolled body needs no copies.
__host__ __device__ constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
__host__ __device__ constexpr int lcm_c(int a, int b) { return a / gcd_c(a, b) * b; }
__host__ __device__ constexpr int z_unroll(int T) { return lcm_c(lcm_c(3, T + 1), kPrefetch); }

// The value held by the previous / next lane of the wave64 (DPP wave_shr:1 /
// wave_shl:1 row moves); lane 0 / lane 63 receive 0.  The fused sweeps take
// their x neighbours this way instead of through LDS: a wave covers whole
// tile rows, and the lanes at a row's ends are halo columns, whose values
// never reach the written region.
__device__ __forceinline__ uint32_t lane_prev(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ uint32_t lane_next(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_prev(float v)
{
    return __uint_as_float(lane_prev(__float_as_uint(v)));
}
__device__ __forceinline__ float lane_next(float v)
{
    return __uint_as_float(lane_next(__float_as_uint(v)));
}

// LDS of one fused-Jacobi block: per level, two planes of the grown tile,
// padded by one on each side (edge threads read the never-written pad, which
// only reaches halo columns that are not written out; a level's plane is
// contiguous, so the lanes of a row read consecutive banks); and the diagonal
// inverse of every flags byte.
template <int TX, int TY, int T>
struct FusedLds {
    float v[T][2][TY + 2][TX + 2];
    float inv[256];
};

// Raw buffer access (gfx9 buffer descriptor: a uniform base and a size in
// bytes; a load at or past the size returns 0).  Offsets are 32-bit and
// unsigned, so the address needs no per-lane 64-bit arithmetic, and an offset
// pushed out of range turns a load into a 0 with no per-lane test.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0));
}
__device__ __forceinline__ uint8_t buf_u8(__amdgpu_buffer_rsrc_t r, uint32_t byte_off)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, (int)byte_off, 0, 0);
}
__device__ __forceinline__ void buf_st_f32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)byte_off, 0, 0);
}

// Planes of slack a fused-sweep block's buffer offsets need beyond its z-chunk
// of zc planes: T below and T above the chunk, the loop's padding to a whole
// unroll (< z_unroll(T)) and the prefetch, plus T more so that the offsets of
// lanes outside the grid (which start at 2^31) stay past the span.  The host
// bounds zc with kFusedSlackPlanes (fused_max_zc, smoke_internal.h).
__host__ __device__ constexpr int fused_slack_planes(int T) { return 3 * T + z_unroll(T) + kPrefetch + 1; }
static_assert(fused_slack_planes(2) <= kFusedSlackPlanes && fused_slack_planes(3) <= kFusedSlackPlanes &&
                  fused_slack_planes(4) <= kFusedSlackPlanes && fused_slack_planes(5) <= kFusedSlackPlanes &&
                  fused_slack_planes(6) <= kFusedSlackPlanes,
              "z-chunk slack of the fused sweeps");

// T fused Jacobi sweeps (temporal blocking) for one tile (tile_x, tile_y) and
// z-chunk tile_z of a TX x TY tile, run by TX x (TY / R) threads: each thread
// holds R adjacent tile rows of one column.  The block computes the tile
// grown by T cells on each side in x and y and writes only the inner
// (TX-2T) x (TY-2T) columns, whose values after T sweeps depend only on data
// inside the grown tile.  Each thread streams its cells in z: level t (sweep
// t) at plane k-t is computed from level t-1 at planes k-t-1, k-t, k-t+1 kept
// in registers and from the four x/y neighbours of level t-1 at plane k-t (x
// from the adjacent lanes; y from the thread's own other rows, or from LDS
// for the rows beyond its first and last, double buffered, one barrier per
// plane).  Every value is computed with the single-sweep arithmetic, so
// results are bit-identical to T single sweeps.  Planes are prefetched into
// registers kPrefetch iterations ahead.  The block reads planes [plo, phi) of
// P, d and flags through buffer descriptors over exactly that span: planes
// outside it and lanes outside the grid (offsets from 2^31) read 0, so the
// loads need no validity selects.  Requires zc <= fused_max_zc(plane, T).
// All threads of the block must call this with the same tile (the barriers
// are block-wide).
//
// Level t is valid only on tile rows [t, TY - t) (a row needs its two
// neighbour rows of level t-1), and valid rows read only valid rows.  A wave
// holds R * 64 / TX whole rows, so a wave none of whose rows is valid at
// level t skips that level and every level above: it runs
// jacobi_fused_levels with NL < T computed levels (a wave-uniform choice).
// Every variant runs the same plane loop with the same barriers, stores its
// levels 0..NL into LDS for the rows above and below, and writes no output
// (written rows lie in waves with NL = T).
template <int TX, int TY, int T, bool DAMPED, int R, int NL>
__device__ __forceinline__ void jacobi_fused_levels(const Grid& g, const float* __restrict__ pin,
                                                    float* __restrict__ pout,
                                                    const float* __restrict__ d,
                                                    const uint8_t* __restrict__ flags, int zc,
                                                    int zero_in, float omega, int tile_x,
                                                    int tile_y, int tile_z,
                                                    FusedLds<TX, TY, T>& sh)
{
    static_assert(NL >= 0 && NL <= T, "levels of a wave");
    static_assert(R >= 1 && TY % R == 0, "whole tile rows per thread");
    const int tx = threadIdx.x;
    const int r0 = (int)threadIdx.y * R;  // first tile row of this thread
    constexpr int OX = TX - 2 * T, OY = TY - 2 * T;
    const int x = tile_x * OX + tx - T + 1;
    const int y0 = tile_y * OY + r0 - T + 1;
    const int z0 = tile_z * zc + 1;
    const int z1 = min(z0 + zc, g.nz - 1);
    const bool x_in = x >= 0 && x < g.nx;
    const bool x_writes = tx >= T && tx < TX - T && x <= g.nx - 2;
    const int plo = max(z0 - T, 0), phi = min(z1 + T, g.nz);
    const uint32_t plane = (uint32_t)g.plane;
    const uint32_t span = (uint32_t)(phi - plo) * plane;  // elements
    const size_t first = (size_t)plo * (size_t)g.plane;
    // zero_in: P reads as 0 everywhere (an empty descriptor), with no branch.
    const __amdgpu_buffer_rsrc_t rp = buf_rsrc(pin + first, zero_in ? 0u : span * 4u);
    const __amdgpu_buffer_rsrc_t rd = buf_rsrc(d + first, span * 4u);
    const __amdgpu_buffer_rsrc_t rf = buf_rsrc(flags + first, span);
    const __amdgpu_buffer_rsrc_t ro = buf_rsrc(pout + first, span * 4u);
    uint32_t lane_b[R], lane_e[R];  // byte / element offsets in plane plo
    bool writes[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int y = y0 + j, r = r0 + j;
        const bool in_grid = x_in && y >= 0 && y < g.ny;
        writes[j] = x_writes && r >= T && r < TY - T && y <= g.ny - 2;
        const uint32_t cell = in_grid ? (uint32_t)g.idx(x, y, 0) : 0u;
        lane_b[j] = in_grid ? cell * 4u : 0x80000000u;
        lane_e[j] = in_grid ? cell : 0x80000000u;
    }
    for (int f = tx + TX * (int)threadIdx.y; f < 256; f += TX * (TY / R)) sh.inv[f] = cell_inv((uint8_t)f);
    __syncthreads();

    float w[R][T][3];  // level t at planes (newest-2, newest-1, newest)
    // Inverse and negated scaled right-hand side 0 - d*inv of planes k, k-1,
    // ..., k-T.  On cells that are not active inv = 0 and d = 0, so nq = +0
    // and the plain update fma(sum, 0, +0) is +0 with no select.
    float iq[R][T + 1], dq[R][T + 1];
#pragma unroll
    for (int j = 0; j < R; ++j) {
#pragma unroll
        for (int t = 0; t < T; ++t) w[j][t][0] = w[j][t][1] = w[j][t][2] = 0.0f;
#pragma unroll
        for (int t = 0; t <= T; ++t) iq[j][t] = dq[j][t] = 0.0f;
    }

    // Prefetch ring: plane k + s in slot s.  Planes below plo wrap to offsets
    // near 2^32 and planes at or past phi lie past the span: both read 0.
    float np[R][kPrefetch], nd[R][kPrefetch];
    uint8_t nf[R][kPrefetch];
    auto load = [&](int kn, int s) {
        const uint32_t dk = (uint32_t)(kn - plo) * plane;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            np[j][s] = buf_f32(rp, lane_b[j] + dk * 4u);
            nd[j][s] = buf_f32(rd, lane_b[j] + dk * 4u);
            nf[j][s] = buf_u8(rf, lane_e[j] + dk);
        }
    };
#pragma unroll
    for (int s = 0; s < kPrefetch; ++s) load(z0 - T + s, s);
    // The trip count is rounded up to a multiple of the unroll factor; the
    // extra iterations load nothing in range and write nothing (zo >= z1).
    constexpr int U = z_unroll(T);
    const int kend = z0 - T + ((z1 - z0 + 2 * T + U - 1) / U) * U;
    for (int kb = z0 - T; kb < kend; kb += U)
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = kb + u;
        float p0[R], dd[R];
        uint8_t fl[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            p0[j] = np[j][0];
            dd[j] = nd[j][0];
            fl[j] = nf[j][0];
#pragma unroll
            for (int s = 0; s + 1 < kPrefetch; ++s) {
                np[j][s] = np[j][s + 1];
                nd[j][s] = nd[j][s + 1];
                nf[j][s] = nf[j][s + 1];
            }
        }
        load(k + kPrefetch, kPrefetch - 1);
#pragma unroll
        for (int j = 0; j < R; ++j) {
#pragma unroll
            for (int t = T; t > 0; --t) {
                iq[j][t] = iq[j][t - 1];
                dq[j][t] = dq[j][t - 1];
            }
            iq[j][0] = sh.inv[fl[j]];
            dq[j][0] = neg_scaled_rhs(dd[j], iq[j][0]);
            w[j][0][0] = w[j][0][1];
            w[j][0][1] = w[j][0][2];
            w[j][0][2] = p0[j];
        }
        const int cur = k & 1, prev = cur ^ 1;
        float newv[R][T + 1];
#pragma unroll
        for (int j = 0; j < R; ++j) newv[j][0] = p0[j];
#pragma unroll
        for (int t = 1; t <= NL; ++t) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                // x neighbours of level t-1 at plane k-t from the adjacent
                // lanes (w[j][t-1][1] holds that plane for every lane); y
                // neighbours from this thread's rows j-1 / j+1, or from LDS
                // (tile row r lives in LDS row r + 1).
                const float ym = (j > 0) ? w[j > 0 ? j - 1 : 0][t - 1][1]
                                         : sh.v[t - 1][prev][r0][tx + 1];
                const float yp = (j + 1 < R) ? w[j + 1 < R ? j + 1 : 0][t - 1][1]
                                             : sh.v[t - 1][prev][r0 + R + 1][tx + 1];
                const float sum = sum6(lane_prev(w[j][t - 1][1]), lane_next(w[j][t - 1][1]), ym,
                                       yp, w[j][t - 1][0], w[j][t - 1][2]);
                const float G = __builtin_fmaf(sum, iq[j][t], dq[j][t]);
                float v;
                if (DAMPED) {
                    const float pc = w[j][t - 1][1];
                    v = pc + omega * (G - pc);
                    v = (iq[j][t] > 0.0f) ? v : 0.0f;
                } else {
                    v = G;
                }
                newv[j][t] = v;
            }
            if (t < NL) {
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    w[j][t][0] = w[j][t][1];
                    w[j][t][1] = w[j][t][2];
                    w[j][t][2] = newv[j][t];
                }
            }
        }
        const int zo = k - T;
        if (NL == T && zo >= z0 && zo < z1) {
#pragma unroll
            for (int j = 0; j < R; ++j)
                if (writes[j]) buf_st_f32(ro, lane_b[j] + (uint32_t)(zo - plo) * plane * 4u, newv[j][NL]);
        }
#pragma unroll
        for (int t = 0; t < T && t <= NL; ++t)
#pragma unroll
            for (int j = 0; j < R; ++j) sh.v[t][cur][r0 + j + 1][tx + 1] = newv[j][t];
        __syncthreads();
    }
}

// The wave's level count NL (see jacobi_fused_levels) as a template argument.
template <int TX, int TY, int T, bool DAMPED, int R, int K>
__device__ __forceinline__ void jacobi_fused_dispatch(int nl, const Grid& g,
                                                      const float* __restrict__ pin,
                                                      float* __restrict__ pout,
                                                      const float* __restrict__ d,
                                                      const uint8_t* __restrict__ flags, int zc,
                                                      int zero_in, float omega, int tile_x,
                                                      int tile_y, int tile_z,
                                                      FusedLds<TX, TY, T>& sh)
{
    if constexpr (K == 0) {
        jacobi_fused_levels<TX, TY, T, DAMPED, R, 0>(g, pin, pout, d, flags, zc, zero_in, omega,
                                                     tile_x, tile_y, tile_z, sh);
    } else {
        if (nl == K)
            jacobi_fused_levels<TX, TY, T, DAMPED, R, K>(g, pin, pout, d, flags, zc, zero_in,
                                                         omega, tile_x, tile_y, tile_z, sh);
        else
            jacobi_fused_dispatch<TX, TY, T, DAMPED, R, K - 1>(nl, g, pin, pout, d, flags, zc,
                                                               zero_in, omega, tile_x, tile_y,
                                                               tile_z, sh);
    }
}

// Block of TX x (TY / R) threads; R tile rows per thread.
template <int TX, int TY, int T, bool DAMPED, int R = 1>
__device__ __forceinline__ void jacobi_fused_tile(const Grid& g, const float* __restrict__ pin,
                                                  float* __restrict__ pout,
                                                  const float* __restrict__ d,
                                                  const uint8_t* __restrict__ flags, int zc,
                                                  int zero_in, float omega, int tile_x, int tile_y,
                                                  int tile_z, FusedLds<TX, TY, T>& sh)
{
    static_assert(64 % TX == 0, "a wave holds whole tile rows");
    constexpr int TPW = 64 / TX;  // thread rows per wave
    constexpr int RPW = TPW * R;  // tile rows per wave
    // First tile row of this thread's wave (uniform across the wave) and the
    // highest level with a valid row in the wave: t <= last row, t <= TY-1-first row.
    const int r0 = __builtin_amdgcn_readfirstlane((int)threadIdx.y / TPW * TPW) * R;
    const int nl = min(T, min(r0 + RPW - 1, TY - 1 - r0));
    jacobi_fused_dispatch<TX, TY, T, DAMPED, R, T>(nl, g, pin, pout, d, flags, zc, zero_in, omega,
                                                   tile_x, tile_y, tile_z, sh);
}

// u -= grad P on every face whose two cells are fluid.  Faces that touch no
// active cell see P_b - P_a = 0 and are skipped (x,y,z >= 1 suffices).
template <int NZ, typename T>
__device__ __forceinline__ void gradient_col(const Grid& g, T* __restrict__ u, T* __restrict__ v,
                                             T* __restrict__ w, const float* __restrict__ p,
                                             const uint8_t* __restrict__ flags, int x, int y, int z0)
{
    if (x < 1 || x > g.nx - 1 || y < 1 || y > g.ny - 1) return;
    float pc[NZ], pxm[NZ], pym[NZ], pzm[NZ], uu[NZ], vv[NZ], ww[NZ];
    uint8_t fl[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        const int z = z0 + j;
        fl[j] = 0;
        pc[j] = pxm[j] = pym[j] = pzm[j] = uu[j] = vv[j] = ww[j] = 0.0f;
        if (z < 1 || z > g.nz - 1) continue;
        const int i = g.idx(x, y, z);
        fl[j] = flags[i];
        pc[j] = p[i];
        pxm[j] = p[i - 1];
        pym[j] = p[i - g.px];
        pzm[j] = p[i - g.plane];
        uu[j] = ldf(u, i);
        vv[j] = ldf(v, i);
        ww[j] = ldf(w, i);
    }
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        if (!(fl[j] & F_SELF)) continue;  // also every cell outside the z range
        const int i = g.idx(x, y, z0 + j);
        if (fl[j] & F_XM) stf(u, i, uu[j] - (pc[j] - pxm[j]));
        if (fl[j] & F_YM) stf(v, i, vv[j] - (pc[j] - pym[j]));
        if (fl[j] & F_ZM) stf(w, i, ww[j] - (pc[j] - pzm[j]));
    }
}
template <typename T>
__device__ __forceinline__ void gradient_at(const Grid& g, T* __restrict__ u, T* __restrict__ v,
                                            T* __restrict__ w, const float* __restrict__ p,
                                            const uint8_t* __restrict__ flags, int x, int y, int z)
{
    gradient_col<1>(g, u, v, w, p, flags, x, y, z);
}

// avgU/avgV/avgW (smokeSimulation.cu:409-447): 8-point means in the
// reference's summation order.  The *_p forms take the index of the cell in
// plane z (ic) and of the same cell in plane z-1 (im) separately, so that the
// two planes can come from different slots of an LDS ring.
template <typename T>
__device__ __forceinline__ float avgU_p(const T* __restrict__ a, int im, int ic, int px)
{
    float t = ldf(a, im) + ldf(a, im + 1) + ldf(a, im - px) + ldf(a, im + 1 - px) + ldf(a, ic) +
              ldf(a, ic + 1) + ldf(a, ic - px) + ldf(a, ic + 1 - px);
    return t / 8.0f;
}
template <typename T>
__device__ __forceinline__ float avgV_p(const T* __restrict__ a, int im, int ic, int px)
{
    float t = ldf(a, im) + ldf(a, im - 1) + ldf(a, im + px) + ldf(a, im - 1 + px) + ldf(a, ic) +
              ldf(a, ic - 1) + ldf(a, ic + px) + ldf(a, ic - 1 + px);
    return t / 8.0f;
}
template <typename T>
__device__ __forceinline__ float avgW_p(const T* __restrict__ a, int im, int ic, int px)
{
    float t = ldf(a, ic) + ldf(a, ic - 1) + ldf(a, ic - px) + ldf(a, ic - 1 - px) + ldf(a, im) +
              ldf(a, im - 1) + ldf(a, im - px) + ldf(a, im - 1 - px);
    return t / 8.0f;
}
template <typename T>
__device__ __forceinline__ float avgU(const T* __restrict__ a, int i, int px, int pl)
{
    return avgU_p(a, i - pl, i, px);
}
template <typename T>
__device__ __forceinline__ float avgV(const T* __restrict__ a, int i, int px, int pl)
{
    return avgV_p(a, i - pl, i, px);
}
template <typename T>
__device__ __forceinline__ float avgW(const T* __restrict__ a, int i, int px, int pl)
{
    return avgW_p(a, i - pl, i, px);
}

// sampleSmoke (smokeSimulation.cu:451-484), split in two so that a kernel can
// issue the loads of several samples together: Tri holds the clamped
// trilinear weights and the four row offsets of the corner pairs (x0, x0+1).
struct Tri {
    float xw0, xw1, yw0, yw1, zw0, zw1;
    int r00, r10, r01, r11;  // idx(x0, y0|y1, z0|z1): first index y, second z
    int x0, y0, y1, z0, z1;  // the corner coordinates (for windowed reads)
    bool xsame;              // x1 == x0 (position clamped at the last cell)
};

// Clamp to [1, n-1] with the cell dims and form weights and indices.
__device__ __forceinline__ Tri tri_setup(const Grid& g, float px_, float py_, float pz_, float dx,
                                         float dy, float dz)
{
    Tri t;
    float x = fmaxf(fminf(px_, (float)(g.nx - 1)), 1.0f);
    float y = fmaxf(fminf(py_, (float)(g.ny - 1)), 1.0f);
    float z = fmaxf(fminf(pz_, (float)(g.nz - 1)), 1.0f);
    int x0 = (int)fminf(floorf(x - dx), (float)(g.nx - 1));
    t.xw1 = (x - dx) - (float)x0;
    t.xw0 = 1.0f - t.xw1;
    int x1 = (int)fminf((float)(x0 + 1), (float)(g.nx - 1));
    int y0 = (int)fminf(floorf(y - dy), (float)(g.ny - 1));
    t.yw1 = (y - dy) - (float)y0;
    t.yw0 = 1.0f - t.yw1;
    int y1 = (int)fminf((float)(y0 + 1), (float)(g.ny - 1));
    int z0 = (int)fminf(floorf(z - dz), (float)(g.nz - 1));
    t.zw1 = (z - dz) - (float)z0;
    t.zw0 = 1.0f - t.zw1;
    int z1 = (int)fminf((float)(z0 + 1), (float)(g.nz - 1));
    t.r00 = g.idx(x0, y0, z0);
    t.r10 = g.idx(x0, y1, z0);
    t.r01 = g.idx(x0, y0, z1);
    t.r11 = g.idx(x0, y1, z1);
    t.x0 = x0;
    t.y0 = y0;
    t.y1 = y1;
    t.z0 = z0;
    t.z1 = z1;
    t.xsame = x1 == x0;
    return t;
}

// Two consecutive floats from a 4-byte-aligned address (one dwordx2 load).
// a[r+1] exists for every row offset r of a cell x0 <= nx-1: rows are
// px >= nx+1 long.
__device__ __forceinline__ float2 ld2(const float* __restrict__ a, int r)
{
    float2 v;
    __builtin_memcpy(&v, a + r, sizeof v);
    return v;
}
__device__ __forceinline__ float2 ld2(const bf16s* __restrict__ a, int r)
{
    uint32_t w;  // both values with one 4-byte load (2-byte aligned)
    __builtin_memcpy(&w, a + r, sizeof w);
    return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}
__device__ __forceinline__ float2 ld2(const f16s* __restrict__ a, int r)
{
    uint32_t w;
    __builtin_memcpy(&w, a + r, sizeof w);
    return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)),
                       (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)));
}

// The eight corners in the reference's order (:473-481), from the four
// x-pairs at rows (y0|y1, z0|z1).
__device__ __forceinline__ float tri_mix(const Tri& t, float2 p00, float2 p10, float2 p01,
                                         float2 p11)
{
    const float f000 = p00.x, f100 = t.xsame ? p00.x : p00.y;
    const float f010 = p10.x, f110 = t.xsame ? p10.x : p10.y;
    const float f001 = p01.x, f101 = t.xsame ? p01.x : p01.y;
    const float f011 = p11.x, f111 = t.xsame ? p11.x : p11.y;
    return t.xw0 * t.yw0 * t.zw0 * f000 + t.xw1 * t.yw0 * t.zw0 * f100 +
           t.xw0 * t.yw1 * t.zw0 * f010 + t.xw1 * t.yw1 * t.zw0 * f110 +
           t.xw0 * t.yw0 * t.zw1 * f001 + t.xw1 * t.yw0 * t.zw1 * f101 +
           t.xw0 * t.yw1 * t.zw1 * f011 + t.xw1 * t.yw1 * t.zw1 * f111;
}
template <typename T>
__device__ __forceinline__ float tri_eval(const T* __restrict__ a, const Tri& t)
{
    return tri_mix(t, ld2(a, t.r00), ld2(a, t.r10), ld2(a, t.r01), ld2(a, t.r11));
}

// velocityAdvectionU/V/W (smokeSimulation.cu:527-615) fused per staggered
// index over x,y,z in [1,n-1]: each component keeps its own range and
// fluid-fluid condition; read [now], write [past].  The three averages are
// shared between the components that use them (avgW feeds both U and V with
// the same arguments, and so on), and all loads of a phase are issued
// together: first the own components and the averages, then the corners of
// all three samples.  Every address read lies inside the allocation (faces up
// to x = nx, y = ny; clamped sample corners), so the loads are unconditional
// and only the stores are predicated.
template <typename T>
__device__ __forceinline__ void advect_velocity_at(const Grid& g, const T* __restrict__ u0,
                                                   const T* __restrict__ v0,
                                                   const T* __restrict__ w0, T* __restrict__ u1,
                                                   T* __restrict__ v1, T* __restrict__ w1,
                                                   const uint8_t* __restrict__ flags, float dt,
                                                   int x, int y, int z)
{
    if (x < 1 || x > g.nx - 1 || y < 1 || y > g.ny - 1 || z < 1 || z > g.nz - 1) return;
    int i = g.idx(x, y, z);
    uint8_t fl = flags[i];
    const bool self = (fl & F_SELF) != 0;
    const bool doU = self && y < g.ny - 1 && z < g.nz - 1 && (fl & F_XM);
    const bool doV = self && x < g.nx - 1 && z < g.nz - 1 && (fl & F_YM);
    const bool doW = self && x < g.nx - 1 && y < g.ny - 1 && (fl & F_ZM);
    if (!(doU || doV || doW)) return;
    const int px = g.px, pl = g.plane;
    const float uu = ldf(u0, i), vv = ldf(v0, i), ww = ldf(w0, i);
    const float aU = avgU(u0, i, px, pl);
    const float aV = avgV(v0, i, px, pl);
    const float aW = avgW(w0, i, px, pl);
    const float xf = (float)x, yf = (float)y, zf = (float)z;
    const Tri tu = tri_setup(g, xf - dt * uu, (yf + 0.5f) - dt * aV, (zf + 0.5f) - dt * aW, 0.0f,
                             0.5f, 0.5f);
    const Tri tv = tri_setup(g, (xf + 0.5f) - dt * aU, yf - dt * vv, (zf + 0.5f) - dt * aW, 0.5f,
                             0.0f, 0.5f);
    const Tri tw = tri_setup(g, (xf + 0.5f) - dt * aU, (yf + 0.5f) - dt * aV, zf - dt * ww, 0.5f,
                             0.5f, 0.0f);
    const float nu = tri_eval(u0, tu);
    const float nv = tri_eval(v0, tv);
    const float nw = tri_eval(w0, tw);
    if (doU) stf(u1, i, nu);
    if (doV) stf(v1, i, nv);
    if (doW) stf(w1, i, nw);
}

// LDS ring of the z-marching advection kernels: NR planes of the window
// [x0-XL, x0+AX+XL-1] x [y0-1, y0+AY] of NF fields (as float; BF16 converts
// exactly), window plane zw >= 0 in slot zw % NR (NR a power of two, taken
// as zw & (NR - 1)).  x0 - XL is even (see
// RingPlane).
template <int AX, int AY, int NR, int NF, int XL = 1>
struct AdvRing {
    static_assert(NR > 0 && (NR & (NR - 1)) == 0, "ring slots: a power of two");
    static constexpr int WX = AX + 2 * XL, WY = AY + 2, PN = WX * WY;
    static constexpr int kPairs = PN / 2;  // WX is even: a pair never straddles a row
    static constexpr int kPerThread = (kPairs + BX * BY - 1) / (BX * BY);
    float f[NF][NR * PN];
};

// One window plane in registers, on its way to its ring slot: as x-pairs
// (one 8-byte FP32 or 4-byte BF16 load each).  The window starts at an even
// x (x0 - XL = BX*m or BX*m - 2) and WX is even, so a pair starts at an even
// x and never straddles a row; rows hold px >= nx + 2 elements when nx is
// even (px is a multiple of 32), so the pair of a window cell x <= nx lies
// inside its row.  Its second element may be past nx, and pairs left of
// x = 0 are not loaded: such cells are never read.
template <int AX, int AY, int NR, int NF, int XL>
struct RingPlane {
    float2 v[NF][AdvRing<AX, AY, NR, NF, XL>::kPerThread];
};

template <int AX, int AY, int NR, int NF, int XL, typename T>
__device__ __forceinline__ RingPlane<AX, AY, NR, NF, XL> ring_load(const Grid& g,
                                                                   const T* const (&src)[NF],
                                                                   int wx0, int wy0, int zw)
{
    using R = AdvRing<AX, AY, NR, NF, XL>;
    const int tid = (int)(threadIdx.x + BX * threadIdx.y);
    RingPlane<AX, AY, NR, NF, XL> r;
#pragma unroll
    for (int j = 0; j < R::kPerThread; ++j) {
        const int k = 2 * (tid + j * (BX * BY));
        const int lx = k % R::WX, ly = k / R::WX;
        const int xw = wx0 + lx, yw = wy0 + ly;
#pragma unroll
        for (int f = 0; f < NF; ++f) r.v[f][j] = make_float2(0.0f, 0.0f);
        if (k < R::PN && xw >= 0 && xw <= g.nx && yw <= g.ny) {
            const int i = g.idx(xw, yw, zw);
#pragma unroll
            for (int f = 0; f < NF; ++f) r.v[f][j] = ld2(src[f], i);
        }
    }
    return r;
}

template <int AX, int AY, int NR, int NF, int XL>
__device__ __forceinline__ void ring_store(AdvRing<AX, AY, NR, NF, XL>& ring,
                                           const RingPlane<AX, AY, NR, NF, XL>& r, int zw)
{
    using R = AdvRing<AX, AY, NR, NF, XL>;
    const int tid = (int)(threadIdx.x + BX * threadIdx.y);
    const int base = (zw & (NR - 1)) * R::PN;
#pragma unroll
    for (int j = 0; j < R::kPerThread; ++j) {
        const int k = 2 * (tid + j * (BX * BY));
        if (k < R::PN) {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                ring.f[f][base + k] = r.v[f][j].x;
                ring.f[f][base + k + 1] = r.v[f][j].y;
            }
        }
    }
}

// One sample from the ring when its corners lie in the window in x and y and
// in the ring's planes [zlo, zlo + 2], else from global memory: the same
// eight values either way.
template <int WX, int WY, int PN, int NR, typename T>
__device__ __forceinline__ float tri_ring(const T* __restrict__ a, const float* s, const Tri& t,
                                          int wx0, int wy0, int zlo)
{
    const int lx = t.x0 - wx0, ly0 = t.y0 - wy0, ly1 = t.y1 - wy0;
    if (lx >= 0 && lx <= WX - 2 && ly0 >= 0 && ly1 <= WY - 1 && t.z0 >= zlo && t.z1 <= zlo + 2) {
        const int p0 = (t.z0 & (NR - 1)) * PN, p1 = (t.z1 & (NR - 1)) * PN;
        const int l00 = p0 + lx + WX * ly0, l10 = p0 + lx + WX * ly1, l01 = p1 + lx + WX * ly0,
                  l11 = p1 + lx + WX * ly1;
        return tri_mix(t, make_float2(s[l00], s[l00 + 1]), make_float2(s[l10], s[l10 + 1]),
                       make_float2(s[l01], s[l01 + 1]), make_float2(s[l11], s[l11 + 1]));
    }
    return tri_eval(a, t);
}

// advect_velocity_at over a column of BX x AY cells at (x0, y0), z in
// [z0, min(z0 + ZC, nz) - 1], with a block of BX x AY threads (AY == BY):
// the block marches up z keeping window planes z-1, z, z+1 of u, v, w in a
// 4-slot LDS ring; while plane z is computed, window plane z+2 is loaded into
// registers and then stored into the slot of plane z-2 (last read in the
// previous step, before the step's barrier).  Own components, averages and
// in-window samples read LDS; a back-trace longer than one cell reads global
// memory.  Each window plane is read from memory once per column, and a
// thread has one cell per plane.  Window rows/columns past (ny, nx) are left
// at zero and never read; window planes stay at or below nz.  Same values,
// same arithmetic as advect_velocity_at: the results are identical.
template <int AY, int ZC, typename T>
__device__ __forceinline__ void advect_velocity_march(const Grid& g, const T* __restrict__ u0,
                                                      const T* __restrict__ v0,
                                                      const T* __restrict__ w0,
                                                      T* __restrict__ u1, T* __restrict__ v1,
                                                      T* __restrict__ w1,
                                                      const uint8_t* __restrict__ flags, float dt,
                                                      int x0, int y0, int z0,
                                                      AdvRing<BX, AY, 4, 3>& ring)
{
    static_assert(AY == BY, "one cell per thread and plane");
    constexpr int NR = 4;
    using R = AdvRing<BX, AY, NR, 3>;
    using P = RingPlane<BX, AY, NR, 3, 1>;
    const T* const src[3] = {u0, v0, w0};
    const int wx0 = x0 - 1, wy0 = y0 - 1;
    const int x = x0 + (int)threadIdx.x, y = y0 + (int)threadIdx.y;
    const int zlast = min(z0 + ZC, g.nz) - 1;  // the grid keeps z0 <= nz - 1
    const bool col = x <= g.nx - 1 && y <= g.ny - 1;
    {
        const P a = ring_load<BX, AY, NR, 3, 1>(g, src, wx0, wy0, z0 - 1);
        const P b = ring_load<BX, AY, NR, 3, 1>(g, src, wx0, wy0, z0);
        const P c = ring_load<BX, AY, NR, 3, 1>(g, src, wx0, wy0, z0 + 1);
        ring_store(ring, a, z0 - 1);
        ring_store(ring, b, z0);
        ring_store(ring, c, z0 + 1);
    }
    uint8_t fl_next = col ? flags[g.idx(x, y, z0)] : (uint8_t)0;
    __syncthreads();
    const float xf = (float)x, yf = (float)y;
    const int l = (x - wx0) + R::WX * (y - wy0);
    for (int z = z0; z <= zlast; ++z) {
        const bool more = z < zlast;  // uniform
        P nxt;
        if (more) nxt = ring_load<BX, AY, NR, 3, 1>(g, src, wx0, wy0, z + 2);
        const uint8_t fl = fl_next;
        if (more && col) fl_next = flags[g.idx(x, y, z + 1)];
        const bool self = (fl & F_SELF) != 0;
        const bool doU = self && y < g.ny - 1 && z < g.nz - 1 && (fl & F_XM);
        const bool doV = self && x < g.nx - 1 && z < g.nz - 1 && (fl & F_YM);
        const bool doW = self && x < g.nx - 1 && y < g.ny - 1 && (fl & F_ZM);
        if (doU || doV || doW) {
            const int pc = (z & (NR - 1)) * R::PN + l, pm = ((z - 1) & (NR - 1)) * R::PN + l;
            const float uu = ring.f[0][pc], vv = ring.f[1][pc], ww = ring.f[2][pc];
            const float aU = avgU_p(ring.f[0], pm, pc, R::WX);
            const float aV = avgV_p(ring.f[1], pm, pc, R::WX);
            const float aW = avgW_p(ring.f[2], pm, pc, R::WX);
            const float zf = (float)z;
            const Tri tu = tri_setup(g, xf - dt * uu, (yf + 0.5f) - dt * aV,
                                     (zf + 0.5f) - dt * aW, 0.0f, 0.5f, 0.5f);
            const Tri tv = tri_setup(g, (xf + 0.5f) - dt * aU, yf - dt * vv,
                                     (zf + 0.5f) - dt * aW, 0.5f, 0.0f, 0.5f);
            const Tri tw = tri_setup(g, (xf + 0.5f) - dt * aU, (yf + 0.5f) - dt * aV,
                                     zf - dt * ww, 0.5f, 0.5f, 0.0f);
            const float nu = tri_ring<R::WX, R::WY, R::PN, NR>(u0, ring.f[0], tu, wx0, wy0, z - 1);
            const float nv = tri_ring<R::WX, R::WY, R::PN, NR>(v0, ring.f[1], tv, wx0, wy0, z - 1);
            const float nw = tri_ring<R::WX, R::WY, R::PN, NR>(w0, ring.f[2], tw, wx0, wy0, z - 1);
            const int i = g.idx(x, y, z);
            if (doU) stf(u1, i, nu);
            if (doV) stf(v1, i, nv);
            if (doW) stf(w1, i, nw);
        }
        if (more) ring_store(ring, nxt, z + 2);
        __syncthreads();
    }
}

// advectSmoke (smokeSimulation.cu:617-638): cell velocity from the
// post-advection faces, back-trace in double rounded to float.
// The sample's corners are clamped into the field, so it is read
// unconditionally too (a NaN position clamps to a valid cell as well); the
// column version issues the face loads of all NZ cells, then the corner loads
// of all NZ samples, then stores.
template <int NZ, typename T, typename TV>
__device__ __forceinline__ void advect_density_col(const Grid& g, const T* __restrict__ r0,
                                                   T* __restrict__ r1, const TV* __restrict__ u,
                                                   const TV* __restrict__ v,
                                                   const TV* __restrict__ w,
                                                   const uint8_t* __restrict__ flags, float dt,
                                                   int x, int y, int z0)
{
    if (x < 1 || x > g.nx - 2 || y < 1 || y > g.ny - 2) return;
    float fu0[NZ], fu1[NZ], fv0[NZ], fv1[NZ], fw0[NZ], fw1[NZ];
    bool self[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        const int z = z0 + j;
        self[j] = false;
        fu0[j] = fu1[j] = fv0[j] = fv1[j] = fw0[j] = fw1[j] = 0.0f;
        if (z < 1 || z > g.nz - 2) continue;
        const int i = g.idx(x, y, z);
        self[j] = (flags[i] & F_SELF) != 0;
        fu0[j] = ldf(u, i);
        fu1[j] = ldf(u, i + 1);
        fv0[j] = ldf(v, i);
        fv1[j] = ldf(v, i + g.px);
        fw0[j] = ldf(w, i);
        fw1[j] = ldf(w, i + g.plane);
    }
    float s[NZ];
#pragma unroll
    for (int j = 0; j < NZ; ++j) {
        const int z = z0 + j;
        float ut = (fu0[j] + fu1[j]) / 2.0f;
        float vt = (fv0[j] + fv1[j]) / 2.0f;
        float wt = (fw0[j] + fw1[j]) / 2.0f;
        float xt = (float)((double)x + 0.5 - (double)(ut * dt));
        float yt = (float)((double)y + 0.5 - (double)(vt * dt));
        float zt = (float)((double)z + 0.5 - (double)(wt * dt));
        s[j] = tri_eval(r0, tri_setup(g, xt, yt, zt, 0.5f, 0.5f, 0.5f));
    }
#pragma unroll
    for (int j = 0; j < NZ; ++j)
        if (self[j]) stf(r1, g.idx(x, y, z0 + j), s[j]);
}
// advect_density_col over a column of BX x AY cells at (x0, y0), z in
// [z0, min(z0 + ZC, nz - 1) - 1], with a block of BX x AY threads
// (AY == BY; x0 a multiple of BX, lane x = cell x, lane 0 of the first
// column idle), marching up z like advect_velocity_march: window planes z-1,
// z, z+1 of the density (x in [x0-2, x0+BX+1], y in [y0-1, y0+AY]: the column
// plus a one-cell halo, widened to an even start) in a 4-slot LDS ring, plane z+2 and the faces of the next cell loaded while
// plane z is computed.  Samples whose corners lie in the ring read LDS, a
// back-trace longer than one cell reads global memory.  Same values, same
// arithmetic as advect_density_col: the results are identical.
template <int AY, int ZC, typename T, typename TV>
__device__ __forceinline__ void advect_density_march(const Grid& g, const T* __restrict__ r0,
                                                     T* __restrict__ r1, const TV* __restrict__ u,
                                                     const TV* __restrict__ v,
                                                     const TV* __restrict__ w,
                                                     const uint8_t* __restrict__ flags, float dt,
                                                     int x0, int y0, int z0,
                                                     AdvRing<BX, AY, 4, 1, 2>& ring)
{
    static_assert(AY == BY, "one cell per thread and plane");
    constexpr int NR = 4;
    using R = AdvRing<BX, AY, NR, 1, 2>;
    using P = RingPlane<BX, AY, NR, 1, 2>;
    const T* const src[1] = {r0};
    const int wx0 = x0 - 2, wy0 = y0 - 1;
    const int x = x0 + (int)threadIdx.x, y = y0 + (int)threadIdx.y;
    const int zlast = min(z0 + ZC, g.nz - 1) - 1;  // the grid keeps z0 <= nz - 2
    const bool col = x >= 1 && x <= g.nx - 2 && y <= g.ny - 2;
    // Faces and flags of cell z (loaded one plane ahead).
    struct Faces {
        float u0, u1, v0, v1, w0, w1;
        bool self;
    };
    auto faces = [&](int z) {
        Faces f = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, false};
        if (col) {
            const int i = g.idx(x, y, z);
            f.self = (flags[i] & F_SELF) != 0;
            f.u0 = ldf(u, i);
            f.u1 = ldf(u, i + 1);
            f.v0 = ldf(v, i);
            f.v1 = ldf(v, i + g.px);
            f.w0 = ldf(w, i);
            f.w1 = ldf(w, i + g.plane);
        }
        return f;
    };
    {
        const P a = ring_load<BX, AY, NR, 1, 2>(g, src, wx0, wy0, z0 - 1);
        const P b = ring_load<BX, AY, NR, 1, 2>(g, src, wx0, wy0, z0);
        const P c = ring_load<BX, AY, NR, 1, 2>(g, src, wx0, wy0, z0 + 1);
        ring_store(ring, a, z0 - 1);
        ring_store(ring, b, z0);
        ring_store(ring, c, z0 + 1);
    }
    Faces fn = faces(z0);
    __syncthreads();
    for (int z = z0; z <= zlast; ++z) {
        const bool more = z < zlast;  // uniform
        P nxt;
        if (more) nxt = ring_load<BX, AY, NR, 1, 2>(g, src, wx0, wy0, z + 2);
        const Faces fc = fn;
        if (more) fn = faces(z + 1);
        if (fc.self) {
            float ut = (fc.u0 + fc.u1) / 2.0f;
            float vt = (fc.v0 + fc.v1) / 2.0f;
            float wt = (fc.w0 + fc.w1) / 2.0f;
            float xt = (float)((double)x + 0.5 - (double)(ut * dt));
            float yt = (float)((double)y + 0.5 - (double)(vt * dt));
            float zt = (float)((double)z + 0.5 - (double)(wt * dt));
            const Tri t = tri_setup(g, xt, yt, zt, 0.5f, 0.5f, 0.5f);
            stf(r1, g.idx(x, y, z),
                tri_ring<R::WX, R::WY, R::PN, NR>(r0, ring.f[0], t, wx0, wy0, z - 1));
        }
        if (more) ring_store(ring, nxt, z + 2);
        __syncthreads();
    }
}

template <typename T, typename TV>
__device__ __forceinline__ void advect_density_at(const Grid& g, const T* __restrict__ r0,
                                                  T* __restrict__ r1, const TV* __restrict__ u,
                                                  const TV* __restrict__ v, const TV* __restrict__ w,
                                                  const uint8_t* __restrict__ flags, float dt,
                                                  int x, int y, int z)
{
    advect_density_col<1>(g, r0, r1, u, v, w, flags, dt, x, y, z);
}

// Vorticity confinement, an extension: the reference lists it as future work
// (README.md:33) and has no code for it, so the oracle's
// vorticity_confinement (oracle/smoke_oracle.cpp) defines it and these
// bodies follow its operations in its order.  With cell-centred velocities
// c = mean of a cell's two faces, omega = curl c by central differences on
// interior fluid cells (0 elsewhere), eta = grad |omega|, N = eta/|eta| where
// |eta| > kVortMinGrad, F = eps (N x omega); every face between two fluid
// cells gains dt * (mean of its two cells' F).  Three passes: curl (omega
// kept in the F buffers, |omega| in mag), force (F over omega in place: a
// cell reads only its own omega) and face update.
constexpr float kVortMinGrad = 1e-6f;

template <typename T>
__device__ __forceinline__ float vort_cen(const T* __restrict__ a, int i, int step)
{
    return (ldf(a, i) + ldf(a, i + step)) * 0.5f;
}
// omega at interior cell i (every face read lies inside the padded field).
template <typename T>
__device__ __forceinline__ void vort_curl(const Grid& g, const T* __restrict__ u,
                                          const T* __restrict__ v, const T* __restrict__ w, int i,
                                          float& ox, float& oy, float& oz)
{
    const int X = 1, Y = g.px, Z = g.plane;
    ox = (vort_cen(w, i + Y, Z) - vort_cen(w, i - Y, Z)) * 0.5f -
         (vort_cen(v, i + Z, Y) - vort_cen(v, i - Z, Y)) * 0.5f;
    oy = (vort_cen(u, i + Z, X) - vort_cen(u, i - Z, X)) * 0.5f -
         (vort_cen(w, i + X, Z) - vort_cen(w, i - X, Z)) * 0.5f;
    oz = (vort_cen(v, i + X, Y) - vort_cen(v, i - X, Y)) * 0.5f -
         (vort_cen(u, i + Y, X) - vort_cen(u, i - Y, X)) * 0.5f;
}
// Pass 1 at cell (x,y,z) of [0,nx)x[0,ny)x[0,nz): writes mag and omega (zeros
// off the interior fluid cells).
template <typename T>
__device__ __forceinline__ void vort_curl_at(const Grid& g, const T* __restrict__ u,
                                             const T* __restrict__ v, const T* __restrict__ w,
                                             const uint8_t* __restrict__ flags,
                                             float* __restrict__ mag, float* __restrict__ f0,
                                             float* __restrict__ f1, float* __restrict__ f2, int x,
                                             int y, int z)
{
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int i = g.idx(x, y, z);
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, m = 0.0f;
    if (is_interior(g, x, y, z) && (flags[i] & F_SELF)) {
        vort_curl(g, u, v, w, i, ox, oy, oz);
        m = sqrtf(ox * ox + oy * oy + oz * oz);
    }
    mag[i] = m;
    f0[i] = ox;
    f1[i] = oy;
    f2[i] = oz;
}
// Pass 2 at interior fluid cells: F = eps (N x omega), or 0 where |eta| is at
// rounding-noise level.
__device__ __forceinline__ void vort_force_at(const Grid& g, const uint8_t* __restrict__ flags,
                                              const float* __restrict__ mag,
                                              float* __restrict__ f0, float* __restrict__ f1,
                                              float* __restrict__ f2, float eps, int x, int y,
                                              int z)
{
    if (!is_interior(g, x, y, z)) return;
    const int i = g.idx(x, y, z);
    if (!(flags[i] & F_SELF)) return;
    const float ex = (mag[i + 1] - mag[i - 1]) * 0.5f;
    const float ey = (mag[i + g.px] - mag[i - g.px]) * 0.5f;
    const float ez = (mag[i + g.plane] - mag[i - g.plane]) * 0.5f;
    const float len = sqrtf(ex * ex + ey * ey + ez * ez);
    if (!(len > kVortMinGrad)) {
        f0[i] = 0.0f;
        f1[i] = 0.0f;
        f2[i] = 0.0f;
        return;
    }
    const float nx = ex / len, ny = ey / len, nz = ez / len;
    const float ox = f0[i], oy = f1[i], oz = f2[i];
    f0[i] = eps * (ny * oz - nz * oy);
    f1[i] = eps * (nz * ox - nx * oz);
    f2[i] = eps * (nx * oy - ny * ox);
}
// Pass 3 at the low faces of cell (x,y,z): u/v/w += dt * mean F of the two
// fluid cells the face separates.
template <typename T>
__device__ __forceinline__ void vort_apply_at(const Grid& g, T* __restrict__ u, T* __restrict__ v,
                                              T* __restrict__ w, const uint8_t* __restrict__ flags,
                                              const float* __restrict__ f0,
                                              const float* __restrict__ f1,
                                              const float* __restrict__ f2, float dt, int x, int y,
                                              int z)
{
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int i = g.idx(x, y, z);
    const uint8_t fl = flags[i];
    if (!(fl & F_SELF)) return;
    if (fl & F_XM) stf(u, i, ldf(u, i) + dt * ((f0[i - 1] + f0[i]) * 0.5f));
    if (fl & F_YM) stf(v, i, ldf(v, i) + dt * ((f1[i - g.px] + f1[i]) * 0.5f));
    if (fl & F_ZM) stf(w, i, ldf(w, i) + dt * ((f2[i - g.plane] + f2[i]) * 0.5f));
}

}  // namespace dev
}  // namespace smk
