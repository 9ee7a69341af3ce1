// HIP kernels of one smoke time step, FP32, for gfx950.
//
// Each kernel evaluates every output value with the same operations, in the
// same order, as the CPU oracle (oracle/smoke_oracle.cpp), which restates
// project/smokeSimulation.cu.  This file is compiled with -ffp-contract=off so
// no multiply-add is fused behind the source's back; FP32 division is the
// correctly rounded one (HIP's default).  Results are therefore bit-identical
// to the oracle's FP32 mode.
//
// Thread blocks are 64 x 4 x 1: one wave64 covers 64 consecutive x of one row,
// so every field access of a wave is one contiguous 256-byte run (FP32) of a
// 128-byte-aligned row.  blockIdx.z walks z planes.
#include "smoke_internal.h"

namespace smk {

namespace {

constexpr int BX = 64;
constexpr int BY = 4;

inline dim3 grid_for(int wx, int wy, int wz)
{
    return dim3((unsigned)((wx + BX - 1) / BX), (unsigned)((wy + BY - 1) / BY), (unsigned)wz);
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
__global__ void k_pack_half(Grid g, const float* __restrict__ src, uint16_t* __restrict__ dst,
                            int format)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const float f = src[g.idx(x, y, z)];
    uint16_t h;
    if (format == 1) {
        const _Float16 v = (_Float16)f;
        h = __builtin_bit_cast(uint16_t, v);
    } else {
        const uint32_t b = __float_as_uint(f);
        if ((b & 0x7fffffffu) > 0x7f800000u) h = (uint16_t)((b >> 16) | 0x0040u);
        else h = (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
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
    if (obst.count > 0 && is_interior(g, x, y, z)) {
        bool solid;
        if (mode == SMOKE_OBSTACLES_LAST_WINS) {
            solid = inside(obst.s[obst.count - 1], x, y, z);
        } else {
            solid = false;
            for (int k = 0; k < obst.count; ++k) solid = solid || inside(obst.s[k], x, y, z);
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
    if (is_interior(g, x, y, z) && (fl & F_SELF) && (fl & F_NBRS)) fl |= F_ACTIVE;
    flags[i] = fl;
}

// fillSmoke (smokeSimulation.cu:251-273) restricted to the bounding box of
// one source: every interior cell inside it gets density 1.0 in both halves.
__global__ void k_source(Grid g, float* __restrict__ r0, float* __restrict__ r1, float4 sph,
                         int x0, int y0, int z0, int ex, int ey)
{
    int x = x0 + blockIdx.x * BX + threadIdx.x;
    int y = y0 + blockIdx.y * BY + threadIdx.y;
    int z = z0 + blockIdx.z;
    if (x >= x0 + ex || y >= y0 + ey) return;
    if (inside(sph, x, y, z)) {
        int i = g.idx(x, y, z);
        r0[i] = 1.0f;
        r1[i] = 1.0f;
    }
}

// --------------------------------------------------------------- force

// integrate (smokeSimulation.cu:315-329) and velocityConfinement (:331-352)
// fused per staggered index: the buoyancy range is x in [0,nx-1],
// y in [1,ny-1], z in [0,nz-1]; the clamp range x,y,z in [1,n-1] lies inside
// it and sees the post-buoyancy v of the same index.
__global__ void k_force(Grid g, float* __restrict__ u, float* __restrict__ v, float* __restrict__ w,
                        const float* __restrict__ rho, const uint8_t* __restrict__ flags, ForceArgs a)
{
    int x = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z;
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    int i = g.idx(x, y, z);
    uint8_t fl = flags[i];
    float vv = v[i];
    bool write_v = false;
    if ((fl & F_SELF) && (fl & F_YM)) {
        float r = rho[i];
        if (!a.gate_on || r > a.gate) {
            float b = a.buoyancy * (r - 0.0f);
            vv = vv + (r * a.gravity * a.dt + b * a.dt);
            write_v = true;
        }
    }
    if (x >= 1 && z >= 1) {
        float uu = u[i];
        float ww = w[i];
        float L = uu * uu + vv * vv + ww * ww;
        if (L * a.dt > a.max_v2) {
            u[i] = uu * (a.max_v2 / (L * a.dt));
            vv = vv * (a.max_v2 / (L * a.dt));
            w[i] = ww * (a.max_v2 / (L * a.dt));
            write_v = true;
        }
    }
    if (write_v) v[i] = vv;
}

// ----------------------------------------------------------- projection

// Colour-o cells of row (y,z): x = 2t + 1 + ((y+z+o+1) & 1) enumerates the
// interior x with (x+y+z+o) even, as smokeSimulation.cu:362 does.
__device__ __forceinline__ int colour_x(int t, int y, int z, int o)
{
    return 2 * t + 1 + ((y + z + o + 1) & 1);
}

// divergence (smokeSimulation.cu:356-394): one red-black SOR half-sweep in
// place on the face velocities.  Same-colour cells share no face, so the
// in-place update is race-free.
__global__ void k_project_ref_half(Grid g, float* __restrict__ u, float* __restrict__ v,
                                   float* __restrict__ w, const uint8_t* __restrict__ flags,
                                   int colour, double omega)
{
    int t = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (y > g.ny - 2 || z > g.nz - 2) return;
    int x = colour_x(t, y, z, colour);
    if (x > g.nx - 2) return;
    int i = g.idx(x, y, z);
    uint8_t fl = flags[i];
    if (!(fl & F_ACTIVE)) return;  // solid (:365) or no fluid neighbour (:377)
    int sx0 = (fl >> 1) & 1, sx1 = (fl >> 2) & 1;
    int sy0 = (fl >> 3) & 1, sy1 = (fl >> 4) & 1;
    int sz0 = (fl >> 5) & 1, sz1 = (fl >> 6) & 1;
    int acc = sx0 + sx1 + sy0 + sy1 + sz0 + sz1;
    float dv = -u[i] + u[i + 1] + -v[i] + v[i + g.px] + -w[i] + w[i + g.plane];
    float p = (float)((double)(-dv / (float)acc) * omega);
    u[i] -= p * (float)sx0;
    u[i + 1] += p * (float)sx1;
    v[i] -= p * (float)sy0;
    v[i + g.px] += p * (float)sy1;
    w[i] -= p * (float)sz0;
    w[i + g.plane] += p * (float)sz1;
}

// Right-hand side of the pressure form: d = div(u*) on active cells, in the
// reference's operand order; 0 elsewhere.  Interior cells only.
__global__ void k_divergence(Grid g, const float* __restrict__ u, const float* __restrict__ v,
                             const float* __restrict__ w, const uint8_t* __restrict__ flags,
                             float* __restrict__ d)
{
    int x = blockIdx.x * BX + threadIdx.x + 1;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (x > g.nx - 2 || y > g.ny - 2 || z > g.nz - 2) return;
    int i = g.idx(x, y, z);
    float dv = 0.0f;
    if (flags[i] & F_ACTIVE)
        dv = -u[i] + u[i + 1] + -v[i] + v[i + g.px] + -w[i] + w[i + g.plane];
    d[i] = dv;
}

__device__ __forceinline__ float nbr_sum(const float* __restrict__ p, int i, uint8_t fl, int px,
                                         int plane)
{
    float sum = 0.0f;
    if (fl & F_XM) sum += p[i - 1];
    if (fl & F_XP) sum += p[i + 1];
    if (fl & F_YM) sum += p[i - px];
    if (fl & F_YP) sum += p[i + px];
    if (fl & F_ZM) sum += p[i - plane];
    if (fl & F_ZP) sum += p[i + plane];
    return sum;
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

// The pressure is +0 on every cell that is not active (ring, solid and
// isolated cells are written 0 and never relaxed), and a running sum that
// starts at +0 is never -0, so adding all six neighbours unmasked gives the
// same bits as adding only the fluid ones.
__device__ __forceinline__ float sum6(float a0, float a1, float a2, float a3, float a4, float a5)
{
    float sum = 0.0f;
    sum = sum + a0;
    sum = sum + a1;
    sum = sum + a2;
    sum = sum + a3;
    sum = sum + a4;
    sum = sum + a5;
    return sum;
}

// One (damped) Jacobi sweep of acc_i P_i - sum_n s_n P_n = -d_i over the
// interior; inactive cells are written 0.  zero_in: the input is all zeros
// (cold start), so it is not read.
__global__ void k_jacobi(Grid g, const float* __restrict__ pin, float* __restrict__ pout,
                         const float* __restrict__ d, const uint8_t* __restrict__ flags,
                         float omega, int plain, int zero_in)
{
    int x = blockIdx.x * BX + threadIdx.x + 1;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (x > g.nx - 2 || y > g.ny - 2 || z > g.nz - 2) return;
    int i = g.idx(x, y, z);
    float inv = cell_inv(flags[i]);
    float sum = 0.0f, pi = 0.0f;
    if (!zero_in) {
        sum = sum6(pin[i - 1], pin[i + 1], pin[i - g.px], pin[i + g.px], pin[i - g.plane],
                   pin[i + g.plane]);
        pi = pin[i];
    }
    float G = (sum - d[i]) * inv;
    float out = plain ? G : pi + omega * (G - pi);
    pout[i] = (inv > 0.0f) ? out : 0.0f;
}

// T fused Jacobi sweeps (temporal blocking).  A TX x TY block owns one z-chunk
// of a tile of columns; it computes the tile grown by T cells on each side in
// x and y and writes only the inner (TX-2T) x (TY-2T) columns, whose values
// after T sweeps depend only on data inside the grown tile.  Each thread holds
// one column and streams it in z: level t (sweep t) at plane k-t is computed
// from level t-1 at planes k-t-1, k-t, k-t+1 kept in registers and from the
// four x/y neighbours of level t-1 at plane k-t held in LDS (double buffered,
// one barrier per plane).  Every value is computed with the single-sweep
// arithmetic, so results are bit-identical to T launches of k_jacobi.
// The next plane is prefetched into registers one iteration ahead.
template <int TX, int TY, int T, bool DAMPED>
__global__ __launch_bounds__(TX * TY) void k_jacobi_fused(
    Grid g, const float* __restrict__ pin, float* __restrict__ pout, const float* __restrict__ d,
    const uint8_t* __restrict__ flags, int zc, int zero_in, float omega)
{
    // Padded by one on each side: edge threads read the (never written) pad,
    // which only reaches halo columns that are not written out.
    __shared__ float lds[T][2][TY + 2][TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    constexpr int OX = TX - 2 * T, OY = TY - 2 * T;
    // Workgroup b runs on XCD b % 8 (8 XCDs, each with its own L2).  Renumber
    // so that each XCD gets one contiguous run of tiles (x fastest): tiles
    // that share halo columns then run on the same L2 at the same time.  The
    // renumbering is a permutation of the tiles (the tail beyond a multiple
    // of 8 keeps its index).
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int nb = nbx * nby * (int)gridDim.z;
    const int b = (int)blockIdx.x + nbx * ((int)blockIdx.y + nby * (int)blockIdx.z);
    const int full = nb & ~7;
    const int L = (b < full) ? (b & 7) * (full >> 3) + (b >> 3) : b;
    const int x = (L % nbx) * OX + tx - T + 1;
    const int y = ((L / nbx) % nby) * OY + ty - T + 1;
    const int z0 = (L / (nbx * nby)) * zc + 1;
    const int z1 = min(z0 + zc, g.nz - 1);
    const bool in_grid = x >= 0 && x < g.nx && y >= 0 && y < g.ny;
    const bool writes = tx >= T && tx < TX - T && ty >= T && ty < TY - T && x <= g.nx - 2 &&
                        y <= g.ny - 2;
    const int base = in_grid ? g.idx(x, y, 0) : 0;

    float w[T][3];  // level t at planes (newest-2, newest-1, newest)
    float iq[T + 1], dq[T + 1];  // inverse and rhs of planes k, k-1, ..., k-T
#pragma unroll
    for (int t = 0; t < T; ++t) w[t][0] = w[t][1] = w[t][2] = 0.0f;
#pragma unroll
    for (int t = 0; t <= T; ++t) iq[t] = dq[t] = 0.0f;

    float np = 0.0f, nd = 0.0f;
    uint8_t nf = 0;
    {
        const int k = z0 - T;
        if (in_grid && k >= 0 && k < g.nz) {
            const int i = base + k * g.plane;
            np = zero_in ? 0.0f : pin[i];
            nf = flags[i];
            nd = d[i];
        }
    }
    for (int k = z0 - T; k < z1 + T; ++k) {
        const float p0 = np, dd = nd;
        const uint8_t fl = nf;
        {
            const int kn = k + 1;
            np = 0.0f;
            nd = 0.0f;
            nf = 0;
            if (in_grid && kn >= 0 && kn < g.nz && kn < z1 + T) {
                const int i = base + kn * g.plane;
                np = zero_in ? 0.0f : pin[i];
                nf = flags[i];
                nd = d[i];
            }
        }
#pragma unroll
        for (int t = T; t > 0; --t) {
            iq[t] = iq[t - 1];
            dq[t] = dq[t - 1];
        }
        iq[0] = cell_inv(fl);
        dq[0] = dd;
        w[0][0] = w[0][1];
        w[0][1] = w[0][2];
        w[0][2] = p0;
        const int cur = k & 1, prev = cur ^ 1;
        float newv[T + 1];
        newv[0] = p0;
#pragma unroll
        for (int t = 1; t <= T; ++t) {
            const float sum = sum6(lds[t - 1][prev][ty + 1][tx], lds[t - 1][prev][ty + 1][tx + 2],
                                   lds[t - 1][prev][ty][tx + 1], lds[t - 1][prev][ty + 2][tx + 1],
                                   w[t - 1][0], w[t - 1][2]);
            const float G = (sum - dq[t]) * iq[t];
            float v;
            if (DAMPED) {
                const float pc = w[t - 1][1];
                v = pc + omega * (G - pc);
            } else {
                v = G;
            }
            v = (iq[t] > 0.0f) ? v : 0.0f;
            newv[t] = v;
            if (t < T) {
                w[t][0] = w[t][1];
                w[t][1] = w[t][2];
                w[t][2] = v;
            }
        }
        const int zo = k - T;
        if (writes && zo >= z0 && zo < z1) pout[base + zo * g.plane] = newv[T];
#pragma unroll
        for (int t = 0; t < T; ++t) lds[t][cur][ty + 1][tx + 1] = newv[t];
        __syncthreads();
    }
}

// One red-black SOR half-sweep of the pressure form, in place.
__global__ void k_rbsor_half(Grid g, float* __restrict__ p, const float* __restrict__ d,
                             const uint8_t* __restrict__ flags, int colour, float omega)
{
    int t = blockIdx.x * BX + threadIdx.x;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (y > g.ny - 2 || z > g.nz - 2) return;
    int x = colour_x(t, y, z, colour);
    if (x > g.nx - 2) return;
    int i = g.idx(x, y, z);
    uint8_t fl = flags[i];
    if (!(fl & F_ACTIVE)) return;
    float G = (nbr_sum(p, i, fl, g.px, g.plane) - d[i]) * inv_acc(__popc(fl & F_NBRS));
    p[i] = p[i] + omega * (G - p[i]);
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

// u -= grad P on every face whose two cells are fluid.  Faces that touch no
// active cell see P_b - P_a = 0 and are skipped (x,y,z >= 1 suffices).
__global__ void k_gradient(Grid g, float* __restrict__ u, float* __restrict__ v,
                           float* __restrict__ w, const float* __restrict__ p,
                           const uint8_t* __restrict__ flags)
{
    int x = blockIdx.x * BX + threadIdx.x + 1;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (x > g.nx - 1 || y > g.ny - 1 || z > g.nz - 1) return;
    int i = g.idx(x, y, z);
    uint8_t fl = flags[i];
    if (!(fl & F_SELF)) return;
    float pc = p[i];
    if (fl & F_XM) u[i] = u[i] - (pc - p[i - 1]);
    if (fl & F_YM) v[i] = v[i] - (pc - p[i - g.px]);
    if (fl & F_ZM) w[i] = w[i] - (pc - p[i - g.plane]);
}

// ------------------------------------------------------------ advection

// avgU/avgV/avgW (smokeSimulation.cu:409-447): 8-point means in the
// reference's summation order.
__device__ __forceinline__ float avgU(const float* __restrict__ a, int i, int px, int pl)
{
    float t = a[i - pl] + a[i + 1 - pl] + a[i - px - pl] + a[i + 1 - px - pl] + a[i] + a[i + 1] +
              a[i - px] + a[i + 1 - px];
    return t / 8.0f;
}
__device__ __forceinline__ float avgV(const float* __restrict__ a, int i, int px, int pl)
{
    float t = a[i - pl] + a[i - 1 - pl] + a[i + px - pl] + a[i - 1 + px - pl] + a[i] + a[i - 1] +
              a[i + px] + a[i - 1 + px];
    return t / 8.0f;
}
__device__ __forceinline__ float avgW(const float* __restrict__ a, int i, int px, int pl)
{
    float t = a[i] + a[i - 1] + a[i - px] + a[i - 1 - px] + a[i - pl] + a[i - 1 - pl] +
              a[i - px - pl] + a[i - 1 - px - pl];
    return t / 8.0f;
}

// sampleSmoke (smokeSimulation.cu:451-484), split in two so that a kernel can
// issue the loads of several samples together: Tri holds the clamped
// trilinear weights and the four row offsets of the corner pairs (x0, x0+1).
struct Tri {
    float xw0, xw1, yw0, yw1, zw0, zw1;
    int r00, r10, r01, r11;  // idx(x0, y0|y1, z0|z1): first index y, second z
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

// The eight corners in the reference's order (:473-481).
__device__ __forceinline__ float tri_eval(const float* __restrict__ a, const Tri& t)
{
    const float2 p00 = ld2(a, t.r00), p10 = ld2(a, t.r10), p01 = ld2(a, t.r01),
                 p11 = ld2(a, t.r11);
    const float f000 = p00.x, f100 = t.xsame ? p00.x : p00.y;
    const float f010 = p10.x, f110 = t.xsame ? p10.x : p10.y;
    const float f001 = p01.x, f101 = t.xsame ? p01.x : p01.y;
    const float f011 = p11.x, f111 = t.xsame ? p11.x : p11.y;
    return t.xw0 * t.yw0 * t.zw0 * f000 + t.xw1 * t.yw0 * t.zw0 * f100 +
           t.xw0 * t.yw1 * t.zw0 * f010 + t.xw1 * t.yw1 * t.zw0 * f110 +
           t.xw0 * t.yw0 * t.zw1 * f001 + t.xw1 * t.yw0 * t.zw1 * f101 +
           t.xw0 * t.yw1 * t.zw1 * f011 + t.xw1 * t.yw1 * t.zw1 * f111;
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
__global__ void k_advect_velocity(Grid g, const float* __restrict__ u0, const float* __restrict__ v0,
                                  const float* __restrict__ w0, float* __restrict__ u1,
                                  float* __restrict__ v1, float* __restrict__ w1,
                                  const uint8_t* __restrict__ flags, float dt)
{
    int x = blockIdx.x * BX + threadIdx.x + 1;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (x > g.nx - 1 || y > g.ny - 1 || z > g.nz - 1) return;
    int i = g.idx(x, y, z);
    uint8_t fl = flags[i];
    const bool self = (fl & F_SELF) != 0;
    const bool doU = self && y < g.ny - 1 && z < g.nz - 1 && (fl & F_XM);
    const bool doV = self && x < g.nx - 1 && z < g.nz - 1 && (fl & F_YM);
    const bool doW = self && x < g.nx - 1 && y < g.ny - 1 && (fl & F_ZM);
    if (!(doU || doV || doW)) return;
    const int px = g.px, pl = g.plane;
    const float uu = u0[i], vv = v0[i], ww = w0[i];
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
    if (doU) u1[i] = nu;
    if (doV) v1[i] = nv;
    if (doW) w1[i] = nw;
}

// advectSmoke (smokeSimulation.cu:617-638): cell velocity from the
// post-advection faces, back-trace in double rounded to float.
__global__ void k_advect_density(Grid g, const float* __restrict__ r0, float* __restrict__ r1,
                                 const float* __restrict__ u, const float* __restrict__ v,
                                 const float* __restrict__ w, const uint8_t* __restrict__ flags,
                                 float dt)
{
    int x = blockIdx.x * BX + threadIdx.x + 1;
    int y = blockIdx.y * BY + threadIdx.y + 1;
    int z = blockIdx.z + 1;
    if (x > g.nx - 2 || y > g.ny - 2 || z > g.nz - 2) return;
    int i = g.idx(x, y, z);
    if (!(flags[i] & F_SELF)) return;
    float ut = (u[i] + u[i + 1]) / 2.0f;
    float vt = (v[i] + v[i + g.px]) / 2.0f;
    float wt = (w[i] + w[i + g.plane]) / 2.0f;
    float xt = (float)((double)x + 0.5 - (double)(ut * dt));
    float yt = (float)((double)y + 0.5 - (double)(vt * dt));
    float zt = (float)((double)z + 0.5 - (double)(wt * dt));
    r1[i] = tri_eval(r0, tri_setup(g, xt, yt, zt, 0.5f, 0.5f, 0.5f));
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

hipError_t pack_cells_half(const Grid& g, const float* src, uint16_t* dst, int format,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_half, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, src, dst,
                       format);
    return hipGetLastError();
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

hipError_t build_flags(const Grid& g, const uint8_t* s, uint8_t* flags, hipStream_t st)
{
    hipLaunchKernelGGL(k_flags, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, s, flags);
    return hipGetLastError();
}

hipError_t fill_source(const Grid& g, float* rho0, float* rho1, float4 sph, hipStream_t st,
                       int* launched)
{
    *launched = 0;
    // Conservative bounding box: any cell with d^2 < r^2 lies within
    // ceil(|r|) + 1 of the centre on each axis; clipped to the interior.
    double r = fabs((double)sph.w);
    double c[3] = {(double)sph.x, (double)sph.y, (double)sph.z};
    int n[3] = {g.nx, g.ny, g.nz};
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        if (!(c[k] == c[k]) || !(r == r)) return hipSuccess;  // NaN: nothing inside
        double l = std::floor(c[k] - r) - 1.0, h = std::ceil(c[k] + r) + 1.0;
        lo[k] = l < 1.0 ? 1 : (l > n[k] - 2 ? n[k] - 1 : (int)l);
        hi[k] = h > n[k] - 2 ? n[k] - 2 : (h < 1.0 ? 0 : (int)h);
        if (lo[k] > hi[k]) return hipSuccess;
    }
    int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    hipLaunchKernelGGL(k_source, grid_for(ex, ey, ez), dim3(BX, BY), 0, st, g, rho0, rho1, sph,
                       lo[0], lo[1], lo[2], ex, ey);
    *launched = 1;
    return hipGetLastError();
}

hipError_t force(const Grid& g, float* u, float* v, float* w, const float* rho,
                 const uint8_t* flags, const ForceArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(k_force, grid_for(g.nx, g.ny - 1, g.nz), dim3(BX, BY), 0, st, g, u, v, w,
                       rho, flags, a);
    return hipGetLastError();
}

hipError_t project_reference_half(const Grid& g, float* u, float* v, float* w,
                                  const uint8_t* flags, int colour, double omega, hipStream_t st)
{
    int tx = (g.nx - 2 + 1) / 2;
    hipLaunchKernelGGL(k_project_ref_half, grid_for(tx, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st, g,
                       u, v, w, flags, colour, omega);
    return hipGetLastError();
}

hipError_t divergence(const Grid& g, const float* u, const float* v, const float* w,
                      const uint8_t* flags, float* d, hipStream_t st)
{
    hipLaunchKernelGGL(k_divergence, grid_for(g.nx - 2, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st, g,
                       u, v, w, flags, d);
    return hipGetLastError();
}

hipError_t jacobi_sweep(const Grid& g, const float* pin, float* pout, const float* d,
                        const uint8_t* flags, float omega, int plain, int zero_in, hipStream_t st)
{
    hipLaunchKernelGGL(k_jacobi, grid_for(g.nx - 2, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st, g,
                       pin, pout, d, flags, omega, plain, zero_in);
    return hipGetLastError();
}

template <int TX, int TY, int T>
static hipError_t launch_fused(const Grid& g, const float* pin, float* pout, const float* d,
                               const uint8_t* flags, int zc, int zero_in, float omega, bool damped,
                               hipStream_t st)
{
    constexpr int OX = TX - 2 * T, OY = TY - 2 * T;
    const int m[3] = {g.nx - 2, g.ny - 2, g.nz - 2};
    dim3 grid((unsigned)((m[0] + OX - 1) / OX), (unsigned)((m[1] + OY - 1) / OY),
              (unsigned)((m[2] + zc - 1) / zc));
    if (damped)
        hipLaunchKernelGGL((k_jacobi_fused<TX, TY, T, true>), grid, dim3(TX, TY), 0, st, g, pin, pout,
                           d, flags, zc, zero_in, omega);
    else
        hipLaunchKernelGGL((k_jacobi_fused<TX, TY, T, false>), grid, dim3(TX, TY), 0, st, g, pin,
                           pout, d, flags, zc, zero_in, omega);
    return hipGetLastError();
}

template <int TX, int TY>
static hipError_t launch_fused_t(const Grid& g, const float* pin, float* pout, const float* d,
                                 const uint8_t* flags, int t, int zc, int zero_in, float omega,
                                 bool damped, hipStream_t st)
{
    switch (t) {
    case 2: return launch_fused<TX, TY, 2>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 3: return launch_fused<TX, TY, 3>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 4: return launch_fused<TX, TY, 4>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 5: return launch_fused<TX, TY, 5>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
    case 6: return launch_fused<TX, TY, 6>(g, pin, pout, d, flags, zc, zero_in, omega, damped, st);
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
    const bool damped = !plain;
    switch (plan.tile) {
    case JacobiPlan::TILE_32x32:
        return launch_fused_t<32, 32>(g, pin, pout, d, flags, sweeps, plan.zc, zero_in, omega, damped, st);
    case JacobiPlan::TILE_64x16:
        return launch_fused_t<64, 16>(g, pin, pout, d, flags, sweeps, plan.zc, zero_in, omega, damped, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t rbsor_half(const Grid& g, float* p, const float* d, const uint8_t* flags, int colour,
                      float omega, hipStream_t st)
{
    int tx = (g.nx - 2 + 1) / 2;
    hipLaunchKernelGGL(k_rbsor_half, grid_for(tx, g.ny - 2, g.nz - 2), dim3(BX, BY), 0, st, g, p, d,
                       flags, colour, omega);
    return hipGetLastError();
}

hipError_t mask_pressure(const Grid& g, float* p, const uint8_t* flags, hipStream_t st)
{
    hipLaunchKernelGGL(k_mask_pressure, grid_for(g.nx, g.ny, g.nz), dim3(BX, BY), 0, st, g, p,
                       flags);
    return hipGetLastError();
}

hipError_t gradient(const Grid& g, float* u, float* v, float* w, const float* p,
                    const uint8_t* flags, hipStream_t st)
{
    hipLaunchKernelGGL(k_gradient, grid_for(g.nx - 1, g.ny - 1, g.nz - 1), dim3(BX, BY), 0, st, g,
                       u, v, w, p, flags);
    return hipGetLastError();
}

hipError_t advect_velocity(const Grid& g, const float* u0, const float* v0, const float* w0,
                           float* u1, float* v1, float* w1, const uint8_t* flags, float dt,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_advect_velocity, grid_for(g.nx - 1, g.ny - 1, g.nz - 1), dim3(BX, BY), 0,
                       st, g, u0, v0, w0, u1, v1, w1, flags, dt);
    return hipGetLastError();
}

hipError_t advect_density(const Grid& g, const float* r0, float* r1, const float* u,
                          const float* v, const float* w, const uint8_t* flags, float dt,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_advect_density, grid_for(g.nx - 2, g.ny - 2, g.nz - 2), dim3(BX, BY), 0,
                       st, g, r0, r1, u, v, w, flags, dt);
    return hipGetLastError();
}

}  // namespace smk
