// HIP kernels of one smoke time step for one simulation, for gfx950: FP32
// fields, or BF16 fields with FP32 arithmetic (the mixed-precision form; the
// per-field kernels are templates over the storage element).
//
// The arithmetic lives in step_device.h (shared with the batched kernels);
// the kernels here map threads to cells.  Thread blocks are 64 x 4 x 1: one
// wave64 covers 64 consecutive x of one row starting at a multiple of 64 (a
// stage that starts at x = 1 idles lane 0 of a row's first wave), so a
// wave's own-cell accesses are one contiguous 256-byte run (FP32) on two
// 128-byte lines of a 128-byte-aligned row, shared with no other wave;
// blockIdx.z walks z planes (kZ planes per thread in the streaming stages).
// Compiled with -ffp-contract=off: results are bit-identical to the oracle's
// FP32 mode.
#include <algorithm>

#include "step_device.h"

namespace smk {

namespace {

using dev::BX;
using dev::BY;

inline dim3 grid_for(int wx, int wy, int wz)
{
    return dim3((unsigned)((wx + BX - 1) / BX), (unsigned)((wy + BY - 1) / BY), (unsigned)wz);
}

// Cells per thread of the streaming stages (force, divergence, gradient,
// density advection): kZ consecutive planes of one column, so each thread
// has kZ cells' loads in flight (the *_col bodies of step_device.h).
constexpr int kZ = kColZ;
inline dim3 grid_col(int wx, int wy, int wz)
{
    return dim3((unsigned)((wx + BX - 1) / BX), (unsigned)((wy + BY - 1) / BY),
                (unsigned)((wz + kZ - 1) / kZ));
}

// ------------------------------------------------------------ utilities

__global__ void k_fill_f32(float* __restrict__ p, long long n, float value)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, long long n, uint8_t value)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = value;
}

// Contiguous index of the reference layout with row length rx and rows ry.
__device__ __forceinline__ long long cidx(int x, int y, int z, int rx, int ry)
{
    return x + (long long)rx * (y + (long long)ry * z);
}

// FP32 values of the reference layout <-> a padded field stored as T.
template <bool TO_PADDED, typename T>
__global__ void k_convert(Grid g, const void* __restrict__ src_, void* __restrict__ dst_, int wx,
                          int wy, int wz)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= wx || y >= wy || z >= wz) return;
    int i = g.idx(x, y, z);
    long long c = cidx(x, y, z, wx, wy);
    if (TO_PADDED)
        dev::stf(static_cast<T*>(dst_), i, static_cast<const float*>(src_)[c]);
    else
        static_cast<float*>(dst_)[c] = dev::ldf(static_cast<const T*>(src_), i);
}

template <bool TO_PADDED, typename T>
__global__ void k_repack(Grid g, const T* __restrict__ src, T* __restrict__ dst, int wx, int wy,
                         int wz)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= wx || y >= wy || z >= wz) return;
    int i = g.idx(x, y, z);
    long long c = cidx(x, y, z, wx, wy);
    if (TO_PADDED) dst[i] = src[c];
    else dst[c] = src[i];
}

// Padded FP32 cells -> contiguous 16-bit values: format 1 FP16 (hardware
// conversion, round to nearest even), format 2 BF16 (round to nearest even on
// the FP32 bits; NaN stays a quiet NaN).
template <typename T>
__global__ void k_pack_half(Grid g, const T* __restrict__ src, uint16_t* __restrict__ dst, int format)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const float f = dev::ldf(src, g.idx(x, y, z));
    uint16_t h;
    if (format == 1) {
        h = __builtin_bit_cast(uint16_t, dev::to_f16(f));
    } else {
        h = dev::bf16_rne(f);
    }
    dst[cidx(x, y, z, g.nx, g.ny)] = h;
}

__global__ void k_unpack_solid(Grid g, const uint8_t* __restrict__ solid, uint8_t* __restrict__ base)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    base[g.idx(x, y, z)] = solid[cidx(x, y, z, g.nx, g.ny)] ? 0 : 1;
}

// Initial mask: fluid everywhere, solid floor y = 0 (smokeSimulation.cu:200-207).
__global__ void k_floor_mask(Grid g, uint8_t* __restrict__ base)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    base[g.idx(x, y, z)] = (y == 0) ? 0 : 1;
}

// --------------------------------------------------------------- scene

// fillObstacle (smokeSimulation.cu:289-313) over a base mask: interior cells
// only; LAST_WINS keeps the reference's overwrite-per-obstacle result.
__global__ void k_mask(Grid g, const uint8_t* __restrict__ base, uint8_t* __restrict__ s,
                       SphereList obst, int mode)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    int i = g.idx(x, y, z);
    uint8_t v = base[i];
    if (obst.count > 0 && dev::is_interior(g, x, y, z)) {
        bool solid;
        if (mode == SMOKE_OBSTACLES_LAST_WINS) {
            solid = dev::inside(obst.s[obst.count - 1], x, y, z);
        } else {
            solid = false;
            for (int k = 0; k < obst.count; ++k) solid = solid || dev::inside(obst.s[k], x, y, z);
        }
        if (solid) v = 0;
    }
    s[i] = v;
}

__global__ void k_flags(Grid g, const uint8_t* __restrict__ s, uint8_t* __restrict__ flags)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    int i = g.idx(x, y, z);
    uint8_t fl = s[i] ? F_SELF : 0;
    if (x > 0 && s[i - 1]) fl |= F_XM;
    if (x < g.nx - 1 && s[i + 1]) fl |= F_XP;
    if (y > 0 && s[i - g.px]) fl |= F_YM;
    if (y < g.ny - 1 && s[i + g.px]) fl |= F_YP;
    if (z > 0 && s[i - g.plane]) fl |= F_ZM;
    if (z < g.nz - 1 && s[i + g.plane]) fl |= F_ZP;
    if (dev::is_interior(g, x, y, z) && (fl & F_SELF) && (fl & F_NBRS)) fl |= F_ACTIVE;
    flags[i] = fl;
}

// Owner of each cell for the moving-obstacle boundary condition: the index
// of the obstacle whose sphere makes it solid under k_mask's rule (LAST_WINS:
// the last obstacle; UNION: the last one containing the cell), or kNoOwner.
__global__ void k_owner(Grid g, uint8_t* __restrict__ owner, SphereList obst, int mode)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    uint8_t who = kNoOwner;
    if (obst.count > 0 && dev::is_interior(g, x, y, z)) {
        if (mode == SMOKE_OBSTACLES_LAST_WINS) {
            if (dev::inside(obst.s[obst.count - 1], x, y, z)) who = (uint8_t)(obst.count - 1);
        } else {
            for (int k = 0; k < obst.count; ++k)
                if (dev::inside(obst.s[k], x, y, z)) who = (uint8_t)k;
        }
    }
    owner[g.idx(x, y, z)] = who;
}

// The low faces of cell (x,y,z) next to an owned cell take the owner's
// velocity component (this cell's owner first, else the low neighbour's).
template <typename T>
__global__ void k_obstacle_bc(Grid g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                              const uint8_t* __restrict__ owner, SphereList vel)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int i = g.idx(x, y, z);
    const uint8_t self = owner[i];
    if (x >= 1) {
        const uint8_t o = self != kNoOwner ? self : owner[i - 1];
        if (o != kNoOwner) dev::stf(u, i, vel.s[o].x);
    }
    if (y >= 1) {
        const uint8_t o = self != kNoOwner ? self : owner[i - g.px];
        if (o != kNoOwner) dev::stf(v, i, vel.s[o].y);
    }
    if (z >= 1) {
        const uint8_t o = self != kNoOwner ? self : owner[i - g.plane];
        if (o != kNoOwner) dev::stf(w, i, vel.s[o].z);
    }
}

// fillSmoke (smokeSimulation.cu:251-273) restricted to the bounding box of
// one source (clipped to the interior by the launcher).
template <typename T>
__global__ void k_source(Grid g, T* __restrict__ r0, T* __restrict__ r1, float4 sph, int x0, int y0,
                         int z0, int ex, int ey)
{
    int x = x0 + blockIdx.x * BX + threadIdx.x;
    int y = y0 + blockIdx.y * BY + threadIdx.y;
    int z = z0 + blockIdx.z;
    if (x >= x0 + ex || y >= y0 + ey) return;
    dev::source_at(g, r0, r1, sph, x, y, z);
}

// ---------------------------------------------------------- step stages

template <typename T, typename TR>
__global__ void k_force(Grid g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                        const TR* __restrict__ rho, const uint8_t* __restrict__ flags, ForceArgs a)
{
    dev::force_col<kZ>(g, u, v, w, rho, flags, a, blockIdx.x * BX + threadIdx.x,
                       blockIdx.y * BY + threadIdx.y + 1, blockIdx.z * kZ);
}

// Vorticity confinement (an extension with no reference code; the bodies in
// step_device.h define it): curl, force and face passes over one simulation.
template <typename T>
__global__ void k_vort_curl(Grid g, const T* __restrict__ u, const T* __restrict__ v,
                            const T* __restrict__ w, const uint8_t* __restrict__ flags,
                            VortBufs b)
{
    dev::vort_curl_at(g, u, v, w, flags, b.mag, b.f[0], b.f[1], b.f[2],
                      blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y, blockIdx.z);
}

__global__ void k_vort_force(Grid g, const uint8_t* __restrict__ flags, VortBufs b, float eps)
{
    dev::vort_force_at(g, flags, b.mag, b.f[0], b.f[1], b.f[2], eps,
                       blockIdx.x * BX + threadIdx.x + 1, blockIdx.y * BY + threadIdx.y + 1,
                       blockIdx.z + 1);
}

template <typename T>
__global__ void k_vort_apply(Grid g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                             const uint8_t* __restrict__ flags, VortBufs b, float dt)
{
    dev::vort_apply_at(g, u, v, w, flags, b.f[0], b.f[1], b.f[2], dt,
                       blockIdx.x * BX + threadIdx.x, blockIdx.y * BY + threadIdx.y, blockIdx.z);
}

// divergence (smokeSimulation.cu:356-394): one red-black SOR half-sweep in
// place on the face velocities.  Same-colour cells share no face, so the
// in-place update is race-free, and any assignment of the colour's cells to
// threads gives the same bits.  The (column pair, row) index is flat over the
// block's lanes (no idle lanes when a row holds fewer pairs than a wave), and
// each thread takes kRefZ consecutive planes, loading every cell's six faces
// and flags before it updates any (one wait for memory; every interior cell's
// faces lie inside the arrays).
constexpr int kRefZ = 2;

__global__ __launch_bounds__(BX * BY) void k_project_ref_half(
    Grid g, float* __restrict__ u, float* __restrict__ v, float* __restrict__ w,
    const uint8_t* __restrict__ flags, int colour, double omega, int pairs)
{
    const int f = (int)blockIdx.x * (BX * BY) + (int)threadIdx.y * BX + (int)threadIdx.x;
    if (f >= pairs * (g.ny - 2)) return;
    const int t = f % pairs, y = f / pairs + 1;
    const int z0 = (int)blockIdx.y * kRefZ + 1;
    float u0[kRefZ], u1[kRefZ], v0[kRefZ], v1[kRefZ], w0[kRefZ], w1[kRefZ];
    uint8_t fl[kRefZ];
    int ic[kRefZ];
#pragma unroll
    for (int j = 0; j < kRefZ; ++j) {
        const int z = z0 + j, x = dev::colour_x(t, y, z, colour);
        fl[j] = 0;
        ic[j] = 0;
        u0[j] = u1[j] = v0[j] = v1[j] = w0[j] = w1[j] = 0.0f;
        if (z > g.nz - 2 || x > g.nx - 2) continue;
        const int i = g.idx(x, y, z);
        ic[j] = i;
        u0[j] = u[i];
        u1[j] = u[i + 1];
        v0[j] = v[i];
        v1[j] = v[i + g.px];
        w0[j] = w[i];
        w1[j] = w[i + g.plane];
        fl[j] = flags[i];
    }
#pragma unroll
    for (int j = 0; j < kRefZ; ++j) {
        const uint8_t c = fl[j];
        if (!(c & F_ACTIVE)) continue;  // solid (:365), no fluid neighbour (:377), or no cell
        const int i = ic[j];
        int sx0 = (c >> 1) & 1, sx1 = (c >> 2) & 1;
        int sy0 = (c >> 3) & 1, sy1 = (c >> 4) & 1;
        int sz0 = (c >> 5) & 1, sz1 = (c >> 6) & 1;
        int acc = sx0 + sx1 + sy0 + sy1 + sz0 + sz1;
        float dv = -u0[j] + u1[j] + -v0[j] + v1[j] + -w0[j] + w1[j];
        float p = (float)((double)(-dv / (float)acc) * omega);
        u[i] = u0[j] - p * (float)sx0;
        u[i + 1] = u1[j] + p * (float)sx1;
        v[i] = v0[j] - p * (float)sy0;
        v[i + g.px] = v1[j] + p * (float)sy1;
        w[i] = w0[j] - p * (float)sz0;
        w[i + g.plane] = w1[j] + p * (float)sz1;
    }
}

template <typename T>
__global__ void k_divergence(Grid g, const T* __restrict__ u, const T* __restrict__ v,
                             const T* __restrict__ w, const uint8_t* __restrict__ flags,
                             float* __restrict__ d)
{
    dev::divergence_col<kZ>(g, u, v, w, flags, d, blockIdx.x * BX + threadIdx.x,
                            blockIdx.y * BY + threadIdx.y + 1, blockIdx.z * kZ + 1);
}

__global__ void k_jacobi(Grid g, const float* __restrict__ pin, float* __restrict__ pout,
                         const float* __restrict__ d, const uint8_t* __restrict__ flags,
                         float omega, int plain, int zero_in)
{
    dev::jacobi_at(g, pin, pout, d, flags, omega, plain, zero_in, blockIdx.x * BX + threadIdx.x,
                   blockIdx.y * BY + threadIdx.y + 1, blockIdx.z + 1);
}

// Tile order of the tiled kernels.  Workgroup b runs on XCD b % 8 (8 XCDs,
// each with its own L2): the tiles are renumbered so that each XCD gets one
// contiguous run of tiles (x fastest), and tiles that share halo cells run on
// the same L2 at the same time.  The renumbering is a permutation (the tail
// beyond a multiple of 8 keeps its index).
__device__ __forceinline__ int xcd_tile_order()
{
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int nb = nbx * nby * (int)gridDim.z;
    const int b = (int)blockIdx.x + nbx * ((int)blockIdx.y + nby * (int)blockIdx.z);
    const int full = nb & ~7;
    return (b < full) ? (b & 7) * (full >> 3) + (b >> 3) : b;
}

// T fused Jacobi sweeps, one TX x TY tile per block of TX x (TY / R) threads
// (dev::jacobi_fused_tile).  One row per thread: at most 64 VGPRs, two
// 1024-thread blocks per CU; two rows per thread: at most 128 VGPRs.
template <int TX, int TY, int T, bool DAMPED, int R>
__global__ __launch_bounds__(TX * TY / R, 8 / R) void k_jacobi_fused(
    Grid g, const float* __restrict__ pin, float* __restrict__ pout, const float* __restrict__ d,
    const uint8_t* __restrict__ flags, int zc, int zero_in, float omega)
{
    __shared__ dev::FusedLds<TX, TY, T> sh;
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int L = xcd_tile_order();
    dev::jacobi_fused_tile<TX, TY, T, DAMPED, R>(g, pin, pout, d, flags, zc, zero_in, omega,
                                                 L % nbx, (L / nbx) % nby, L / (nbx * nby), sh);
}

// One red-black SOR half-sweep of the pressure form, in place.
__global__ __launch_bounds__(BX * BY) void k_rbsor_half(Grid g, float* __restrict__ p,
                                                    const float* __restrict__ d,
                                                    const uint8_t* __restrict__ flags, int colour,
                                                    float omega, int pairs)
{
    // The same thread layout as k_project_ref_half: a flat (column pair, row)
    // index and kRefZ consecutive planes per thread, every load issued before
    // any update.  The neighbours are of the other colour, which this launch
    // leaves alone, and every interior cell's neighbours lie inside the array.
    const int f = (int)blockIdx.x * (BX * BY) + (int)threadIdx.y * BX + (int)threadIdx.x;
    if (f >= pairs * (g.ny - 2)) return;
    const int t = f % pairs, y = f / pairs + 1;
    const int z0 = (int)blockIdx.y * kRefZ + 1;
    float nb[kRefZ][6], pc[kRefZ], di[kRefZ];
    uint8_t fl[kRefZ];
    int ic[kRefZ];
#pragma unroll
    for (int j = 0; j < kRefZ; ++j) {
        const int z = z0 + j, x = dev::colour_x(t, y, z, colour);
        fl[j] = 0;
        ic[j] = 0;
        pc[j] = di[j] = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) nb[j][k] = 0.0f;
        if (z > g.nz - 2 || x > g.nx - 2) continue;
        const int i = g.idx(x, y, z);
        ic[j] = i;
        nb[j][0] = p[i - 1];
        nb[j][1] = p[i + 1];
        nb[j][2] = p[i - g.px];
        nb[j][3] = p[i + g.px];
        nb[j][4] = p[i - g.plane];
        nb[j][5] = p[i + g.plane];
        pc[j] = p[i];
        di[j] = d[i];
        fl[j] = flags[i];
    }
#pragma unroll
    for (int j = 0; j < kRefZ; ++j) {
        if (!(fl[j] & F_ACTIVE)) continue;
        const float G = dev::pressure_update(
            dev::sum6(nb[j][0], nb[j][1], nb[j][2], nb[j][3], nb[j][4], nb[j][5]),
            dev::inv_acc(__popc(fl[j] & F_NBRS)), di[j]);
        p[ic[j]] = pc[j] + omega * (G - pc[j]);
    }
}

// Warm start: zero the pressure of cells that are not active.
__global__ void k_mask_pressure(Grid g, float* __restrict__ p, const uint8_t* __restrict__ flags)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    int i = g.idx(x, y, z);
    if (!(flags[i] & F_ACTIVE)) p[i] = 0.0f;
}

template <typename T>
__global__ void k_gradient(Grid g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                           const float* __restrict__ p, const uint8_t* __restrict__ flags)
{
    dev::gradient_col<kZ>(g, u, v, w, p, flags, blockIdx.x * BX + threadIdx.x,
                          blockIdx.y * BY + threadIdx.y + 1, blockIdx.z * kZ + 1);
}

// Velocity advection: a block of BX x BY threads marches a column of BX x BY
// cells through kAdvZC planes (dev::advect_velocity_march), in XCD tile order.
constexpr int kAdvZC = kAdvChunkZ;

template <typename T>
__global__ __launch_bounds__(BX * BY) void k_advect_velocity(
    Grid g, const T* __restrict__ u0, const T* __restrict__ v0, const T* __restrict__ w0,
    T* __restrict__ u1, T* __restrict__ v1, T* __restrict__ w1, const uint8_t* __restrict__ flags,
    float dt)
{
    __shared__ dev::AdvRing<BX, BY, 4, 3> ring;
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int L = xcd_tile_order();
    dev::advect_velocity_march<BY, kAdvZC>(g, u0, v0, w0, u1, v1, w1, flags, dt,
                                           1 + BX * (L % nbx), 1 + BY * ((L / nbx) % nby),
                                           1 + kAdvZC * (L / (nbx * nby)), ring);
}

// Density advection: a block of BX x BY threads marches a column of BX x BY
// cells through kAdvZC planes (dev::advect_density_march), in XCD tile order.
template <typename T, typename TV>
__global__ __launch_bounds__(BX * BY) void k_advect_density(
    Grid g, const T* __restrict__ r0, T* __restrict__ r1, const TV* __restrict__ u,
    const TV* __restrict__ v, const TV* __restrict__ w, const uint8_t* __restrict__ flags, float dt)
{
    __shared__ dev::AdvRing<BX, BY, 4, 1, 2> ring;
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int L = xcd_tile_order();
    dev::advect_density_march<BY, kAdvZC>(g, r0, r1, u, v, w, flags, dt, BX * (L % nbx),
                                          1 + BY * ((L / nbx) % nby),
                                          1 + kAdvZC * (L / (nbx * nby)), ring);
}

}  // namespace

// ------------------------------------------------------------- launchers

hipError_t fill_f32(float* p, long long n, float value, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, value);
    return hipGetLastError();
}

hipError_t fill_u8(uint8_t* p, long long n, uint8_t value, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, st, p, n, value);
    return hipGetLastError();
}

hipError_t unpack_cells_f32(const Grid& g, const float* src, float* dst, hipStream_t st)
{
    hipLaunchKernelGGL((k_repack<true, float>), grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st,
                       g, src, dst, g.nx, g.ny, g.nz);
    return hipGetLastError();
}

hipError_t pack_cells_f32(const Grid& g, const float* src, float* dst, hipStream_t st)
{
    hipLaunchKernelGGL((k_repack<false, float>), grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st,
                       g, src, dst, g.nx, g.ny, g.nz);
    return hipGetLastError();
}

hipError_t unpack_faces_f32(const Grid& g, const float* src, float* dst, hipStream_t st)
{
    hipLaunchKernelGGL((k_repack<true, float>), grid_for(g.nx + 1, g.ny + 1, g.nz + 1),
                       dim3(BX, BY), 0, st, g, src, dst, g.nx + 1, g.ny + 1, g.nz + 1);
    return hipGetLastError();
}

hipError_t pack_faces_f32(const Grid& g, const float* src, float* dst, hipStream_t st)
{
    hipLaunchKernelGGL((k_repack<false, float>), grid_for(g.nx + 1, g.ny + 1, g.nz + 1),
                       dim3(BX, BY), 0, st, g, src, dst, g.nx + 1, g.ny + 1, g.nz + 1);
    return hipGetLastError();
}

hipError_t unpack_cells_solid(const Grid& g, const uint8_t* solid, uint8_t* base, hipStream_t st)
{
    hipLaunchKernelGGL(k_unpack_solid, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, solid,
                       base);
    return hipGetLastError();
}

hipError_t pack_cells_u8(const Grid& g, const uint8_t* src, uint8_t* dst, hipStream_t st)
{
    hipLaunchKernelGGL((k_repack<false, uint8_t>), grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st,
                       g, src, dst, g.nx, g.ny, g.nz);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_pack_half(const Grid& g, const T* src, uint16_t* dst, int format,
                                   hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_half<T>, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, src, dst,
                       format);
    return hipGetLastError();
}
hipError_t pack_cells_half(const Grid& g, const float* src, uint16_t* dst, int format,
                           hipStream_t st)
{
    return launch_pack_half(g, src, dst, format, st);
}


hipError_t floor_mask(const Grid& g, uint8_t* base, hipStream_t st)
{
    hipLaunchKernelGGL(k_floor_mask, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, base);
    return hipGetLastError();
}

hipError_t build_mask(const Grid& g, const uint8_t* base, uint8_t* s, const SphereList& obst,
                      int mode, hipStream_t st)
{
    hipLaunchKernelGGL(k_mask, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, base, s, obst,
                       mode);
    return hipGetLastError();
}

hipError_t build_owner(const Grid& g, uint8_t* owner, const SphereList& obst, int mode,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_owner, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, owner, obst,
                       mode);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_obstacle_bc(const Grid& g, T* u, T* v, T* w, const uint8_t* owner,
                                     const SphereList& vel, hipStream_t st)
{
    hipLaunchKernelGGL(k_obstacle_bc<T>, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, u, v,
                       w, owner, vel);
    return hipGetLastError();
}
hipError_t obstacle_velocity_bc(const Grid& g, float* u, float* v, float* w, const uint8_t* owner,
                                const SphereList& vel, hipStream_t st)
{
    return launch_obstacle_bc(g, u, v, w, owner, vel, st);
}

hipError_t build_flags(const Grid& g, const uint8_t* s, uint8_t* flags, hipStream_t st)
{
    hipLaunchKernelGGL(k_flags, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, s, flags);
    return hipGetLastError();
}

bool source_box(const Grid& g, float4 sph, int lo[3], int ext[3])
{
    // Any cell with d^2 < r^2 lies within ceil(|r|) + 1 of the centre on each
    // axis; clipped to the interior [1, n-2].
    double r = fabs((double)sph.w);
    double c[3] = {(double)sph.x, (double)sph.y, (double)sph.z};
    int n[3] = {g.nx, g.ny, g.nz};
    for (int k = 0; k < 3; ++k) {
        if (!(c[k] == c[k]) || !(r == r)) return false;  // NaN: nothing inside
        double l = std::floor(c[k] - r) - 1.0, h = std::ceil(c[k] + r) + 1.0;
        int lk = l < 1.0 ? 1 : (l > n[k] - 2 ? n[k] - 1 : (int)l);
        int hk = h > n[k] - 2 ? n[k] - 2 : (h < 1.0 ? 0 : (int)h);
        if (lk > hk) return false;
        lo[k] = lk;
        ext[k] = hk - lk + 1;
    }
    return true;
}

template <typename T>
static hipError_t launch_source(const Grid& g, T* rho0, T* rho1, float4 sph, hipStream_t st,
                               int* launched)
{
    *launched = 0;
    int lo[3], ext[3];
    if (!source_box(g, sph, lo, ext)) return hipSuccess;
    hipLaunchKernelGGL(k_source<T>, grid_for(ext[0], ext[1], ext[2]), dim3(BX, BY), 0, st, g, rho0,
                       rho1, sph, lo[0], lo[1], lo[2], ext[0], ext[1]);
    *launched = 1;
    return hipGetLastError();
}

hipError_t fill_source(const Grid& g, float* rho0, float* rho1, float4 sph, hipStream_t st,
                       int* launched)
{
    return launch_source(g, rho0, rho1, sph, st, launched);
}

template <typename T, typename TR>
static hipError_t launch_force(const Grid& g, T* u, T* v, T* w, const TR* rho, const uint8_t* flags,
                               const ForceArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL((k_force<T, TR>), grid_col(g.nx, g.ny - 1, g.nz), dim3(BX, BY), 0, st, g, u,
                       v, w, rho, flags, a);
    return hipGetLastError();
}
hipError_t force(const Grid& g, float* u, float* v, float* w, const float* rho,
                 const uint8_t* flags, const ForceArgs& a, hipStream_t st)
{
    return launch_force(g, u, v, w, rho, flags, a, st);
}

template <typename T>
static hipError_t launch_vorticity(const Grid& g, T* u, T* v, T* w, const uint8_t* flags,
                                   const VortBufs& b, float eps, float dt, hipStream_t st)
{
    hipLaunchKernelGGL(k_vort_curl<T>, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, u, v, w,
                       flags, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vort_force, grid_for(g.nx - 2, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st, g,
                       flags, b, eps);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vort_apply<T>, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, u, v, w,
                       flags, b, dt);
    return hipGetLastError();
}
hipError_t vorticity(const Grid& g, float* u, float* v, float* w, const uint8_t* flags,
                     const VortBufs& b, float eps, float dt, hipStream_t st)
{
    return launch_vorticity(g, u, v, w, flags, b, eps, dt, st);
}

hipError_t project_reference_half(const Grid& g, float* u, float* v, float* w,
                                  const uint8_t* flags, int colour, double omega, hipStream_t st)
{
    const int pairs = (g.nx - 2 + 1) / 2;  // cells of one colour per row, at most
    const long long items = (long long)pairs * (g.ny - 2);
    const dim3 grid((unsigned)((items + BX * BY - 1) / (BX * BY)),
                    (unsigned)((g.nz - 2 + kRefZ - 1) / kRefZ));
    hipLaunchKernelGGL(k_project_ref_half, grid, dim3(BX, BY), 0, st, g, u, v, w, flags, colour,
                       omega, pairs);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_divergence(const Grid& g, const T* u, const T* v, const T* w,
                                    const uint8_t* flags, float* d, hipStream_t st)
{
    hipLaunchKernelGGL(k_divergence<T>, grid_col(g.nx - 1, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st,
                       g, u, v, w, flags, d);
    return hipGetLastError();
}
hipError_t divergence(const Grid& g, const float* u, const float* v, const float* w,
                      const uint8_t* flags, float* d, hipStream_t st)
{
    return launch_divergence(g, u, v, w, flags, d, st);
}

hipError_t jacobi_sweep(const Grid& g, const float* pin, float* pout, const float* d,
                        const uint8_t* flags, float omega, int plain, int zero_in, hipStream_t st)
{
    hipLaunchKernelGGL(k_jacobi, grid_for(g.nx - 1, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st, g,
                       pin, pout, d, flags, omega, plain, zero_in);
    return hipGetLastError();
}

template <int TX, int TY, int R, int T>
static hipError_t launch_fused(const Grid& g, const float* pin, float* pout, const float* d,
                               const uint8_t* flags, int zc, int zero_in, float omega, bool damped,
                               hipStream_t st)
{
    constexpr int OX = TX - 2 * T, OY = TY - 2 * T;
    const int m[3] = {g.nx - 2, g.ny - 2, g.nz - 2};
    dim3 grid((unsigned)((m[0] + OX - 1) / OX), (unsigned)((m[1] + OY - 1) / OY),
              (unsigned)((m[2] + zc - 1) / zc));
    const dim3 block(TX, TY / R);
    if (damped)
        hipLaunchKernelGGL((k_jacobi_fused<TX, TY, T, true, R>), grid, block, 0, st, g, pin, pout,
                           d, flags, zc, zero_in, omega);
    else
        hipLaunchKernelGGL((k_jacobi_fused<TX, TY, T, false, R>), grid, block, 0, st, g, pin,
                           pout, d, flags, zc, zero_in, omega);
    return hipGetLastError();
}

template <int TX, int TY, int R>
static hipError_t launch_fused_t(const Grid& g, const float* pin, float* pout, const float* d,
                                 const uint8_t* flags, int t, int zc, int zero_in, float omega,
                                 bool damped, hipStream_t st)
{
    switch (t) {
    case 2: return launch_fused<TX, TY, R, 2>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 3: return launch_fused<TX, TY, R, 3>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 4: return launch_fused<TX, TY, R, 4>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 5: return launch_fused<TX, TY, R, 5>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 6: return launch_fused<TX, TY, R, 6>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t jacobi_fused(const Grid& g, const float* pin, float* pout, const float* d,
                        const uint8_t* flags, const JacobiPlan& plan, int sweeps, float omega,
                        int plain, int zero_in, hipStream_t st)
{
    if (sweeps == 1)
        return jacobi_sweep(g, pin, pout, d, flags, omega, plain, zero_in, st);
    if (plan.zc < 1) return hipErrorInvalidValue;
    // The z-chunk is a speed parameter (every chunk gives the same bits): cap
    // it so a block's span of planes fits the 32-bit buffer offsets.
    const int zc = std::min(plan.zc, fused_max_zc(g.plane));
    if (zc < 1) return hipErrorInvalidValue;  // enqueue_jacobi runs single sweeps then
    const bool damped = !plain;
    switch (plan.tile) {
    case JacobiPlan::TILE_32x32:
        return launch_fused_t<32, 32, 1>(g, pin, pout, d, flags, sweeps, zc, zero_in, omega, damped, st);
    case JacobiPlan::TILE_64x16:
        return launch_fused_t<64, 16, 1>(g, pin, pout, d, flags, sweeps, zc, zero_in, omega, damped, st);
    case JacobiPlan::TILE_64x32:
        return launch_fused_t<64, 32, 2>(g, pin, pout, d, flags, sweeps, zc, zero_in, omega, damped, st);
    case JacobiPlan::TILE_32x64:
        return launch_fused_t<32, 64, 2>(g, pin, pout, d, flags, sweeps, zc, zero_in, omega, damped, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t rbsor_half(const Grid& g, float* p, const float* d, const uint8_t* flags, int colour,
                      float omega, hipStream_t st)
{
    const int pairs = (g.nx - 2 + 1) / 2;
    const long long items = (long long)pairs * (g.ny - 2);
    const dim3 grid((unsigned)((items + BX * BY - 1) / (BX * BY)),
                    (unsigned)((g.nz - 2 + kRefZ - 1) / kRefZ));
    hipLaunchKernelGGL(k_rbsor_half, grid, dim3(BX, BY), 0, st, g, p, d, flags, colour, omega,
                       pairs);
    return hipGetLastError();
}

hipError_t mask_pressure(const Grid& g, float* p, const uint8_t* flags, hipStream_t st)
{
    hipLaunchKernelGGL(k_mask_pressure, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, p,
                       flags);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_gradient(const Grid& g, T* u, T* v, T* w, const float* p,
                                  const uint8_t* flags, hipStream_t st)
{
    hipLaunchKernelGGL(k_gradient<T>, grid_col(g.nx, g.ny - 1, g.nz - 1), dim3(BX, BY), 0, st,
                       g, u, v, w, p, flags);
    return hipGetLastError();
}
hipError_t gradient(const Grid& g, float* u, float* v, float* w, const float* p,
                    const uint8_t* flags, hipStream_t st)
{
    return launch_gradient(g, u, v, w, p, flags, st);
}

template <typename T>
static hipError_t launch_advect_velocity(const Grid& g, const T* u0, const T* v0, const T* w0,
                                         T* u1, T* v1, T* w1, const uint8_t* flags, float dt,
                                         hipStream_t st)
{
    const dim3 grid((unsigned)((g.nx - 1 + BX - 1) / BX), (unsigned)((g.ny - 1 + BY - 1) / BY),
                    (unsigned)((g.nz - 1 + kAdvZC - 1) / kAdvZC));
    hipLaunchKernelGGL(k_advect_velocity<T>, grid, dim3(BX, BY), 0, st, g, u0, v0, w0, u1, v1, w1,
                       flags, dt);
    return hipGetLastError();
}
hipError_t advect_velocity(const Grid& g, const float* u0, const float* v0, const float* w0,
                           float* u1, float* v1, float* w1, const uint8_t* flags, float dt,
                           hipStream_t st)
{
    return launch_advect_velocity(g, u0, v0, w0, u1, v1, w1, flags, dt, st);
}

template <typename T, typename TV>
static hipError_t launch_advect_density(const Grid& g, const T* r0, T* r1, const TV* u, const TV* v,
                                        const TV* w, const uint8_t* flags, float dt, hipStream_t st)
{
    const dim3 grid((unsigned)((g.nx - 1 + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY),
                    (unsigned)((g.nz - 2 + kAdvZC - 1) / kAdvZC));
    hipLaunchKernelGGL((k_advect_density<T, TV>), grid, dim3(BX, BY), 0, st, g, r0, r1, u, v, w,
                       flags, dt);
    return hipGetLastError();
}
hipError_t advect_density(const Grid& g, const float* r0, float* r1, const float* u,
                          const float* v, const float* w, const uint8_t* flags, float dt,
                          hipStream_t st)
{
    return launch_advect_density(g, r0, r1, u, v, w, flags, dt, st);
}

// The stages and conversions for 16-bit stored fields H (declared in
// smoke_internal.h), density stored as H or as FP32.
#define SMK_HALF_STORAGE_DEFS(H)                                                                   \
    hipError_t pack_cells_half(const Grid& g, const H* src, uint16_t* dst, int format,             \
                               hipStream_t st)                                                     \
    {                                                                                              \
        return launch_pack_half(g, src, dst, format, st);                                          \
    }                                                                                              \
    hipError_t unpack_cells_f32(const Grid& g, const float* src, H* dst, hipStream_t st)           \
    {                                                                                              \
        hipLaunchKernelGGL((k_convert<true, H>), grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st,  \
                           g, (const void*)src, (void*)dst, g.nx, g.ny, g.nz);                     \
        return hipGetLastError();                                                                  \
    }                                                                                              \
    hipError_t pack_cells_f32(const Grid& g, const H* src, float* dst, hipStream_t st)             \
    {                                                                                              \
        hipLaunchKernelGGL((k_convert<false, H>), grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, \
                           g, (const void*)src, (void*)dst, g.nx, g.ny, g.nz);                     \
        return hipGetLastError();                                                                  \
    }                                                                                              \
    hipError_t unpack_faces_f32(const Grid& g, const float* src, H* dst, hipStream_t st)           \
    {                                                                                              \
        hipLaunchKernelGGL((k_convert<true, H>), grid_for(g.nx + 1, g.ny + 1, g.nz + 1),           \
                           dim3(BX, BY), 0, st, g, (const void*)src, (void*)dst, g.nx + 1,         \
                           g.ny + 1, g.nz + 1);                                                    \
        return hipGetLastError();                                                                  \
    }                                                                                              \
    hipError_t pack_faces_f32(const Grid& g, const H* src, float* dst, hipStream_t st)             \
    {                                                                                              \
        hipLaunchKernelGGL((k_convert<false, H>), grid_for(g.nx + 1, g.ny + 1, g.nz + 1),          \
                           dim3(BX, BY), 0, st, g, (const void*)src, (void*)dst, g.nx + 1,         \
                           g.ny + 1, g.nz + 1);                                                    \
        return hipGetLastError();                                                                  \
    }                                                                                              \
    hipError_t fill_source(const Grid& g, H* rho0, H* rho1, float4 sph, hipStream_t st,            \
                           int* launched)                                                          \
    {                                                                                              \
        return launch_source(g, rho0, rho1, sph, st, launched);                                    \
    }                                                                                              \
    hipError_t force(const Grid& g, H* u, H* v, H* w, const H* rho, const uint8_t* flags,          \
                     const ForceArgs& a, hipStream_t st)                                           \
    {                                                                                              \
        return launch_force(g, u, v, w, rho, flags, a, st);                                        \
    }                                                                                              \
    hipError_t force(const Grid& g, H* u, H* v, H* w, const float* rho, const uint8_t* flags,      \
                     const ForceArgs& a, hipStream_t st)                                           \
    {                                                                                              \
        return launch_force(g, u, v, w, rho, flags, a, st);                                        \
    }                                                                                              \
    hipError_t obstacle_velocity_bc(const Grid& g, H* u, H* v, H* w, const uint8_t* owner,         \
                                    const SphereList& vel, hipStream_t st)                         \
    {                                                                                              \
        return launch_obstacle_bc(g, u, v, w, owner, vel, st);                                     \
    }                                                                                              \
    hipError_t vorticity(const Grid& g, H* u, H* v, H* w, const uint8_t* flags,                    \
                         const VortBufs& b, float eps, float dt, hipStream_t st)                   \
    {                                                                                              \
        return launch_vorticity(g, u, v, w, flags, b, eps, dt, st);                                \
    }                                                                                              \
    hipError_t divergence(const Grid& g, const H* u, const H* v, const H* w,                       \
                          const uint8_t* flags, float* d, hipStream_t st)                          \
    {                                                                                              \
        return launch_divergence(g, u, v, w, flags, d, st);                                        \
    }                                                                                              \
    hipError_t gradient(const Grid& g, H* u, H* v, H* w, const float* p, const uint8_t* flags,     \
                        hipStream_t st)                                                            \
    {                                                                                              \
        return launch_gradient(g, u, v, w, p, flags, st);                                          \
    }                                                                                              \
    hipError_t advect_velocity(const Grid& g, const H* u0, const H* v0, const H* w0, H* u1,        \
                               H* v1, H* w1, const uint8_t* flags, float dt, hipStream_t st)       \
    {                                                                                              \
        return launch_advect_velocity(g, u0, v0, w0, u1, v1, w1, flags, dt, st);                   \
    }                                                                                              \
    hipError_t advect_density(const Grid& g, const H* r0, H* r1, const H* u, const H* v,           \
                              const H* w, const uint8_t* flags, float dt, hipStream_t st)          \
    {                                                                                              \
        return launch_advect_density(g, r0, r1, u, v, w, flags, dt, st);                           \
    }                                                                                              \
    hipError_t advect_density(const Grid& g, const float* r0, float* r1, const H* u, const H* v,   \
                              const H* w, const uint8_t* flags, float dt, hipStream_t st)          \
    {                                                                                              \
        return launch_advect_density(g, r0, r1, u, v, w, flags, dt, st);                           \
    }
SMK_HALF_STORAGE_DEFS(bf16s)
SMK_HALF_STORAGE_DEFS(f16s)
#undef SMK_HALF_STORAGE_DEFS

}  // namespace smk
