// Batched smoke step: many simulations of different sizes advance together
// with one launch per stage.  Each launch is a persistent grid that walks a
// flattened work list of (simulation, block) "virtual blocks"; the arithmetic
// is the single-simulation arithmetic of step_device.h, so every simulation
// gives the same bits in a batch as alone.
//
// Virtual blocks are handed to XCDs in contiguous runs: workgroup b runs on
// XCD b % 8, so the grid's workgroups of XCD j take the j-th eighth of the
// work list, where neighbouring blocks of one simulation (sharing rows and
// halo columns) meet in the same L2.
#include "step_device.h"

namespace smk {

namespace {

using dev::BX;
using dev::BY;

// Virtual blocks of this workgroup: the XCD's contiguous share of [0, total)
// walked with a stride of the XCD's workgroup count.  gridDim.x is a multiple
// of 8 (launchers guarantee it).
struct VbRange {
    int begin, end, stride;
};
__device__ __forceinline__ VbRange vb_range(int total)
{
    const int xcd = (int)blockIdx.x & 7;
    const int per_xcd = (int)gridDim.x >> 3;
    const int chunk = (total + 7) >> 3;
    VbRange r;
    r.begin = xcd * chunk + ((int)blockIdx.x >> 3);
    r.end = min((xcd + 1) * chunk, total);
    r.stride = per_xcd;
    return r;
}

// Simulation owning virtual block vb: last s with start[s] <= vb (binary
// search over the prefix table; uniform across the workgroup).
__device__ __forceinline__ int find_owner(const int* __restrict__ start, int n, int vb)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= vb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(BX * BY) void kb_sources(const SimDesc* __restrict__ sims,
                                                      const SourceTable* __restrict__ tab)
{
    const int total = tab->start[tab->count];
    const VbRange r = vb_range(total);
    for (int vb = r.begin; vb < r.end; vb += r.stride) {
        const int k = find_owner(tab->start, tab->count, vb);
        const int lb = vb - tab->start[k];
        const int4 lo = tab->lo[k], ext = tab->ext[k];
        const int gy = (ext.y + BY - 1) / BY;
        const int bx = lb % ext.w, by = (lb / ext.w) % gy, bz = lb / (ext.w * gy);
        const int x = lo.x + bx * BX + (int)threadIdx.x;
        const int y = lo.y + by * BY + (int)threadIdx.y;
        const int z = lo.z + bz;
        if (x >= lo.x + ext.x || y >= lo.y + ext.y) continue;
        const SimDesc& s = sims[tab->sim[k]];
        dev::source_at(s.g, s.rho[0], s.rho[1], tab->sph[k], x, y, z);
    }
}

// Per-cell stages (force, divergence, gradient), kColZ planes per thread.  The
// coordinate offsets match the single-simulation launches: x from 0 (lane x
// = cell x), y from 1, z from 0 for the force and from 1 for the others.
template <int STAGE>
__global__ __launch_bounds__(BX * BY) void kb_cells(const SimDesc* __restrict__ sims,
                                                    const StageTable* __restrict__ tab, int nsims,
                                                    int lp)
{
    const int total = tab->start[nsims];
    const VbRange r = vb_range(total);
    for (int vb = r.begin; vb < r.end; vb += r.stride) {
        const int si = find_owner(tab->start, nsims, vb);
        const int lb = vb - tab->start[si];
        const int gx = tab->gx[si], gy = tab->gy[si];
        const int bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
        const SimDesc& s = sims[si];
        const int n = s.now;
        const int x = bx * BX + (int)threadIdx.x;
        const int y = by * BY + (int)threadIdx.y + 1;
        const int z0 = bz * kColZ + (STAGE == BS_FORCE ? 0 : 1);
        if (STAGE == BS_FORCE)
            dev::force_col<kColZ>(s.g, s.u[n], s.v[n], s.w[n], s.rho[n], s.flags, s.force, x, y,
                                  z0);
        else if (STAGE == BS_DIV)
            dev::divergence_col<kColZ>(s.g, s.u[n], s.v[n], s.w[n], s.flags, s.d, x, y, z0);
        else
            dev::gradient_col<kColZ>(s.g, s.u[n], s.v[n], s.w[n], s.P[s.pcur ^ lp], s.flags, x, y,
                                     z0);
    }
}

// Velocity advection over every column of every simulation
// (dev::advect_velocity_march: BX x BY cells marched through kAdvChunkZ
// planes); the barrier after a column keeps the next column's ring stores
// from overwriting planes still being read.  Every thread of a workgroup
// walks the same columns.
__global__ __launch_bounds__(BX * BY) void kb_adv_v(const SimDesc* __restrict__ sims,
                                                    const StageTable* __restrict__ tab, int nsims)
{
    __shared__ dev::AdvRing<BX, BY, 4, 3> ring;
    const int total = tab->start[nsims];
    const VbRange r = vb_range(total);
    for (int vb = r.begin; vb < r.end; vb += r.stride) {
        const int si = find_owner(tab->start, nsims, vb);
        const int lb = vb - tab->start[si];
        const int gx = tab->gx[si], gy = tab->gy[si];
        const SimDesc& s = sims[si];
        const int n = s.now, p = 1 - s.now;
        dev::advect_velocity_march<BY, kAdvChunkZ>(
            s.g, s.u[n], s.v[n], s.w[n], s.u[p], s.v[p], s.w[p], s.flags, s.force.dt,
            1 + BX * (lb % gx), 1 + BY * ((lb / gx) % gy), 1 + kAdvChunkZ * (lb / (gx * gy)), ring);
        __syncthreads();
    }
}

// Density advection over every column of every simulation
// (dev::advect_density_march), as kb_adv_v.
__global__ __launch_bounds__(BX * BY) void kb_adv_d(const SimDesc* __restrict__ sims,
                                                    const StageTable* __restrict__ tab, int nsims)
{
    __shared__ dev::AdvRing<BX, BY, 4, 1, 2> ring;
    const int total = tab->start[nsims];
    const VbRange r = vb_range(total);
    for (int vb = r.begin; vb < r.end; vb += r.stride) {
        const int si = find_owner(tab->start, nsims, vb);
        const int lb = vb - tab->start[si];
        const int gx = tab->gx[si], gy = tab->gy[si];
        const SimDesc& s = sims[si];
        const int n = s.now, p = 1 - s.now;
        dev::advect_density_march<BY, kAdvChunkZ>(
            s.g, s.rho[n], s.rho[p], s.u[p], s.v[p], s.w[p], s.flags, s.force.dt, BX * (lb % gx),
            1 + BY * ((lb / gx) % gy), 1 + kAdvChunkZ * (lb / (gx * gy)), ring);
        __syncthreads();
    }
}

// Fused Jacobi sweeps over every tile of every simulation (TX x TY tiles,
// R tile rows per thread).
template <int TX, int TY, int T, bool DAMPED, int R>
__global__ __launch_bounds__(TX * TY / R, 8 / R) void kb_jacobi(const SimDesc* __restrict__ sims,
                                                     const StageTable* __restrict__ tab, int nsims,
                                                     int lp, int zc, int zero_in, float omega)
{
    __shared__ dev::FusedLds<TX, TY, T> sh;
    const int total = tab->start[nsims];
    const VbRange r = vb_range(total);
    for (int vb = r.begin; vb < r.end; vb += r.stride) {
        const int si = find_owner(tab->start, nsims, vb);
        const int lb = vb - tab->start[si];
        const int gx = tab->gx[si], gy = tab->gy[si];
        const SimDesc& s = sims[si];
        const int in = s.pcur ^ lp;
        dev::jacobi_fused_tile<TX, TY, T, DAMPED, R>(s.g, s.P[in], s.P[in ^ 1], s.d, s.flags, zc,
                                                     zero_in, omega, lb % gx, (lb / gx) % gy,
                                                     lb / (gx * gy), sh);
    }
}

// Single sweeps (the remainder of K when it is 1), one cell per thread.
__global__ __launch_bounds__(BX * BY) void kb_jacobi1(const SimDesc* __restrict__ sims,
                                                      const StageTable* __restrict__ tab,
                                                      int nsims, int lp, int zero_in, float omega,
                                                      int plain)
{
    const int total = tab->start[nsims];
    const VbRange r = vb_range(total);
    for (int vb = r.begin; vb < r.end; vb += r.stride) {
        const int si = find_owner(tab->start, nsims, vb);
        const int lb = vb - tab->start[si];
        const int gx = tab->gx[si], gy = tab->gy[si];
        const int bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
        const SimDesc& s = sims[si];
        const int in = s.pcur ^ lp;
        // The interior table's blocks hold kColZ planes each.
#pragma unroll
        for (int j = 0; j < kColZ; ++j)
            dev::jacobi_at(s.g, s.P[in], s.P[in ^ 1], s.d, s.flags, omega, plain, zero_in,
                           bx * BX + (int)threadIdx.x, by * BY + (int)threadIdx.y + 1,
                           bz * kColZ + j + 1);
    }
}

__global__ void kb_flip(SimDesc* __restrict__ sims, int nsims, int pflip)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nsims) return;
    sims[i].now ^= 1;
    sims[i].pcur ^= pflip;
}

template <int TX, int TY, int R, int T>
hipError_t launch_bj(const SimDesc* sims, const StageTable* tab, int nsims, int grid, int lp,
                     int zc, int zero_in, float omega, int plain, hipStream_t st)
{
    const dim3 block(TX, TY / R);
    if (plain)
        hipLaunchKernelGGL((kb_jacobi<TX, TY, T, false, R>), dim3((unsigned)grid), block, 0, st,
                           sims, tab, nsims, lp, zc, zero_in, omega);
    else
        hipLaunchKernelGGL((kb_jacobi<TX, TY, T, true, R>), dim3((unsigned)grid), block, 0, st,
                           sims, tab, nsims, lp, zc, zero_in, omega);
    return hipGetLastError();
}

template <int TX, int TY, int R>
hipError_t launch_bj_t(int t, const SimDesc* sims, const StageTable* tab, int nsims, int grid,
                       int lp, int zc, int zero_in, float omega, int plain, hipStream_t st)
{
    switch (t) {
    case 2: return launch_bj<TX, TY, R, 2>(sims, tab, nsims, grid, lp, zc, zero_in, omega, plain, st);
    case 3: return launch_bj<TX, TY, R, 3>(sims, tab, nsims, grid, lp, zc, zero_in, omega, plain, st);
    case 4: return launch_bj<TX, TY, R, 4>(sims, tab, nsims, grid, lp, zc, zero_in, omega, plain, st);
    case 5: return launch_bj<TX, TY, R, 5>(sims, tab, nsims, grid, lp, zc, zero_in, omega, plain, st);
    case 6: return launch_bj<TX, TY, R, 6>(sims, tab, nsims, grid, lp, zc, zero_in, omega, plain, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename K>
int occupancy(K kernel, int threads)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1)
        n = 1;
    return n;
}

// Resident blocks per CU of the plain T-sweep kernel of one tile.
template <int TX, int TY, int R>
int bj_occupancy(int t)
{
    constexpr int threads = TX * TY / R;
    switch (t) {
    case 2: return occupancy(kb_jacobi<TX, TY, 2, false, R>, threads);
    case 3: return occupancy(kb_jacobi<TX, TY, 3, false, R>, threads);
    case 4: return occupancy(kb_jacobi<TX, TY, 4, false, R>, threads);
    case 5: return occupancy(kb_jacobi<TX, TY, 5, false, R>, threads);
    default: return occupancy(kb_jacobi<TX, TY, 6, false, R>, threads);
    }
}

}  // namespace

hipError_t batch_sources(const SimDesc* sims, const SourceTable* tab, int total_blocks, int grid,
                         hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(kb_sources, dim3((unsigned)grid), dim3(BX, BY), 0, st, sims, tab);
    return hipGetLastError();
}

hipError_t batch_cells(int stage, const SimDesc* sims, const StageTable* tab, int nsims,
                       int total_blocks, int grid, int lp, hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    const dim3 gd((unsigned)grid), bd(BX, BY);
    switch (stage) {
    case BS_FORCE: hipLaunchKernelGGL(kb_cells<BS_FORCE>, gd, bd, 0, st, sims, tab, nsims, lp); break;
    case BS_DIV: hipLaunchKernelGGL(kb_cells<BS_DIV>, gd, bd, 0, st, sims, tab, nsims, lp); break;
    case BS_GRAD: hipLaunchKernelGGL(kb_cells<BS_GRAD>, gd, bd, 0, st, sims, tab, nsims, lp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t batch_jacobi(const SimDesc* sims, const StageTable* tab, int nsims, int total_blocks,
                        int grid, const JacobiPlan& plan, int sweeps, int lp, int zero_in,
                        float omega, int plain, hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    if (sweeps == 1) {
        hipLaunchKernelGGL(kb_jacobi1, dim3((unsigned)grid), dim3(BX, BY), 0, st, sims, tab, nsims,
                           lp, zero_in, omega, plain);
        return hipGetLastError();
    }
    if (plan.zc < 1) return hipErrorInvalidValue;
    switch (plan.tile) {
    case JacobiPlan::TILE_32x32:
        return launch_bj_t<32, 32, 1>(sweeps, sims, tab, nsims, grid, lp, plan.zc, zero_in, omega,
                                      plain, st);
    case JacobiPlan::TILE_64x16:
        return launch_bj_t<64, 16, 1>(sweeps, sims, tab, nsims, grid, lp, plan.zc, zero_in, omega,
                                      plain, st);
    case JacobiPlan::TILE_64x32:
        return launch_bj_t<64, 32, 2>(sweeps, sims, tab, nsims, grid, lp, plan.zc, zero_in, omega,
                                      plain, st);
    case JacobiPlan::TILE_32x64:
        return launch_bj_t<32, 64, 2>(sweeps, sims, tab, nsims, grid, lp, plan.zc, zero_in, omega,
                                      plain, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t batch_flip(SimDesc* sims, int nsims, int pflip, hipStream_t st)
{
    hipLaunchKernelGGL(kb_flip, dim3((unsigned)((nsims + 63) / 64)), dim3(64), 0, st, sims, nsims,
                       pflip);
    return hipGetLastError();
}

int batch_cells_blocks_per_cu() { return occupancy(kb_cells<BS_GRAD>, BX * BY); }

hipError_t batch_advect_velocity(const SimDesc* sims, const StageTable* tab, int nsims,
                                 int total_blocks, int grid, hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(kb_adv_v, dim3((unsigned)grid), dim3(BX, BY), 0, st, sims, tab, nsims);
    return hipGetLastError();
}

int batch_advect_velocity_blocks_per_cu() { return occupancy(kb_adv_v, BX * BY); }

hipError_t batch_advect_density(const SimDesc* sims, const StageTable* tab, int nsims,
                                int total_blocks, int grid, hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(kb_adv_d, dim3((unsigned)grid), dim3(BX, BY), 0, st, sims, tab, nsims);
    return hipGetLastError();
}

int batch_advect_density_blocks_per_cu() { return occupancy(kb_adv_d, BX * BY); }

int batch_jacobi_blocks_per_cu(const JacobiPlan& plan, int sweeps)
{
    if (sweeps == 1) return occupancy(kb_jacobi1, BX * BY);
    switch (plan.tile) {
    case JacobiPlan::TILE_64x16: return bj_occupancy<64, 16, 1>(sweeps);
    case JacobiPlan::TILE_64x32: return bj_occupancy<64, 32, 2>(sweeps);
    case JacobiPlan::TILE_32x64: return bj_occupancy<32, 64, 2>(sweeps);
    default: return bj_occupancy<32, 32, 1>(sweeps);
    }
}

}  // namespace smk
