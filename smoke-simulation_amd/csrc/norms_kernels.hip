// Step norms of one simulation (smoke_get_step_norms): FP64 sums of squares
// of the velocity divergence before and after the projection, the density
// total, FP32 maxima of |u|, |v|, |w| (the CFL number's speed) and of the
// density, and a count of non-finite values, reduced on the device in a fixed
// order, so the result depends only on the fields and the grid dims.
//
// Pass 1 (k_norms_part): a 1-D grid of at most kNormsBlocks (4096) workgroups of
// 64 x 4 threads walks the rows (y, z) of the cell range [0,nx) x [0,ny) x
// [0,nz); lane x takes cells x, x + 64, ... of its row.  Each thread sums in
// its own fixed order, each wave reduces its 64 lanes with a butterfly of
// lane exchanges (every lane ends with the same, fixed-order total), and the
// four waves combine in LDS in wave order.  Pass 2 (k_norms_final), one
// workgroup, sums the per-workgroup partials in a fixed order the same way.
//
// What is read, per cell (x, y, z):
//   * the output half [past]: density (every cell); u at x >= 1, v at y >= 1,
//     w at z >= 1 (the faces every stage can write; faces on the grid's low
//     and high walls of their own axis are never touched, reference
//     smokeSimulation.cu:336, :364, :532, :562, :592);
//   * on cells the projection relaxes (flags bit 7): the divergence of the
//     projected velocity [now], in the right-hand side's operand order
//     (-u0 + u1 + -v0 + v1 + -w0 + w1, smokeSimulation.cu:379-381), and the
//     right-hand side d of the projection (the divergence before it).
#include <algorithm>
#include <mutex>

#include "step_device.h"

namespace smk {

namespace {

using dev::BX;
using dev::BY;

using NormAcc = StepNorms;

__device__ __forceinline__ void acc_zero(NormAcc& a)
{
    a.div2 = a.rhs2 = a.rho_sum = 0.0;
    a.umax = a.vmax = a.wmax = a.rho_max = 0.0f;
    a.nonfinite = 0;
}

// a += b in a fixed order (sums) / order-free (max of magnitudes; NaN is
// ignored by fmaxf and counted in nonfinite instead).
__device__ __forceinline__ void acc_add(NormAcc& a, const NormAcc& b)
{
    a.div2 = a.div2 + b.div2;
    a.rhs2 = a.rhs2 + b.rhs2;
    a.rho_sum = a.rho_sum + b.rho_sum;
    a.umax = fmaxf(a.umax, b.umax);
    a.vmax = fmaxf(a.vmax, b.vmax);
    a.wmax = fmaxf(a.wmax, b.wmax);
    a.rho_max = fmaxf(a.rho_max, b.rho_max);
    a.nonfinite = a.nonfinite + b.nonfinite;
}

// Butterfly over the 64 lanes of a wave: after step m every lane holds the
// sum of its 2m-lane group, combined as (lower half) + (upper half), so all
// lanes end with the same value, formed in the same order.
__device__ __forceinline__ NormAcc wave_reduce(NormAcc a)
{
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        NormAcc o;
        o.div2 = __shfl_xor(a.div2, m);
        o.rhs2 = __shfl_xor(a.rhs2, m);
        o.rho_sum = __shfl_xor(a.rho_sum, m);
        o.umax = __shfl_xor(a.umax, m);
        o.vmax = __shfl_xor(a.vmax, m);
        o.wmax = __shfl_xor(a.wmax, m);
        o.rho_max = __shfl_xor(a.rho_max, m);
        o.nonfinite = __shfl_xor(a.nonfinite, m);
        if (lane & m) {
            NormAcc lo = o;
            acc_add(lo, a);
            a = lo;
        } else {
            acc_add(a, o);
        }
    }
    return a;
}

// Block of BX x BY threads: wave results combined in wave order by thread 0.
__device__ __forceinline__ NormAcc block_reduce(NormAcc a, NormAcc (&sh)[BY])
{
    a = wave_reduce(a);
    if (threadIdx.x == 0) sh[threadIdx.y] = a;
    __syncthreads();
    NormAcc r = sh[0];
#pragma unroll
    for (int k = 1; k < BY; ++k) acc_add(r, sh[k]);
    return r;
}

__device__ __forceinline__ void count_bad(NormAcc& a, float v)
{
    a.nonfinite += __builtin_isfinite(v) ? 0ull : 1ull;
}

template <typename T, typename TR>
__global__ __launch_bounds__(BX * BY) void k_norms_part(
    Grid g, const TR* __restrict__ rho, const T* __restrict__ u, const T* __restrict__ v,
    const T* __restrict__ w, const T* __restrict__ un, const T* __restrict__ vn,
    const T* __restrict__ wn, const float* __restrict__ d, int rhs, const uint8_t* __restrict__ flags,
    NormAcc* __restrict__ part)
{
    __shared__ NormAcc sh[BY];
    NormAcc a;
    acc_zero(a);
    const int rows = g.ny * g.nz;
    for (int r = (int)(blockIdx.x * BY + threadIdx.y); r < rows; r += (int)(gridDim.x * BY)) {
        const int y = r % g.ny, z = r / g.ny;
#pragma unroll 2
        for (int x = (int)threadIdx.x; x < g.nx; x += BX) {
            const int i = g.idx(x, y, z);
            const float q = dev::ldf(rho, i);
            a.rho_sum = a.rho_sum + (double)q;
            a.rho_max = fmaxf(a.rho_max, fabsf(q));
            count_bad(a, q);
            if (x >= 1) {
                const float f = dev::ldf(u, i);
                a.umax = fmaxf(a.umax, fabsf(f));
                count_bad(a, f);
            }
            if (y >= 1) {
                const float f = dev::ldf(v, i);
                a.vmax = fmaxf(a.vmax, fabsf(f));
                count_bad(a, f);
            }
            if (z >= 1) {
                const float f = dev::ldf(w, i);
                a.wmax = fmaxf(a.wmax, fabsf(f));
                count_bad(a, f);
            }
            if (flags[i] & F_ACTIVE) {
                const float dv = -dev::ldf(un, i) + dev::ldf(un, i + 1) + -dev::ldf(vn, i) +
                                 dev::ldf(vn, i + g.px) + -dev::ldf(wn, i) +
                                 dev::ldf(wn, i + g.plane);
                a.div2 = a.div2 + (double)dv * (double)dv;
                if (rhs) {
                    const double b = (double)d[i];
                    a.rhs2 = a.rhs2 + b * b;
                }
            }
        }
    }
    const NormAcc t = block_reduce(a, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(BX * BY) void k_norms_final(const NormAcc* __restrict__ part, int n,
                                                        NormAcc* __restrict__ out)
{
    __shared__ NormAcc sh[BY];
    NormAcc a;
    acc_zero(a);
    const int t = (int)(threadIdx.x + BX * threadIdx.y);
    for (int k = t; k < n; k += BX * BY) acc_add(a, part[k]);
    const NormAcc r = block_reduce(a, sh);
    if (t == 0) *out = r;
}

// The public record: square roots of the sums (FP64), the CFL number from the
// step's dt, and the counts.  Both entry points produce it with this kernel.
__global__ void k_norms_publish(const NormAcc* __restrict__ r, smoke_step_norms* __restrict__ out,
                                float dt, int rhs, long long step)
{
    if (threadIdx.x != 0) return;
    const NormAcc a = *r;
    smoke_step_norms o;
    o.div_l2 = sqrt(a.div2);
    o.rhs_l2 = rhs ? sqrt(a.rhs2) : -1.0;
    o.density_sum = a.rho_sum;
    o.cfl = (double)dt * (double)fmaxf(a.umax, fmaxf(a.vmax, a.wmax));
    o.max_u = a.umax;
    o.max_v = a.vmax;
    o.max_w = a.wmax;
    o.max_density = a.rho_max;
    o.nonfinite = (long long)a.nonfinite;
    o.step = step;
    *out = o;
}

template <typename T, typename TR>
hipError_t launch_norms(const Grid& g, const TR* rho, const T* u, const T* v, const T* w,
                        const T* un, const T* vn, const T* wn, const float* d, int rhs,
                        const uint8_t* flags, void* scratch, smoke_step_norms* out, float dt,
                        long long step, hipStream_t st)
{
    const int rows = g.ny * g.nz;
    const int blocks = std::min(kNormsBlocks, (rows + BY - 1) / BY);
    NormAcc* part = static_cast<NormAcc*>(scratch);
    NormAcc* res = part + kNormsBlocks;
    hipLaunchKernelGGL((k_norms_part<T, TR>), dim3((unsigned)blocks), dim3(BX, BY), 0, st, g, rho, u, v,
                       w, un, vn, wn, d, rhs, flags, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(BX, BY), 0, st, part, blocks, res);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_norms_publish, dim3(1), dim3(64), 0, st, res, out, dt, rhs, step);
    return hipGetLastError();
}

}  // namespace

size_t norms_scratch_bytes() { return (size_t)(kNormsBlocks + 1) * sizeof(NormAcc); }

// Scratch of the allocation-free entry point (smoke_get_step_norms_noalloc):
// the partial records and the reduced record, one copy per GPU the library
// runs on, held until the process ends (192 KiB).  Every grid fits, because
// the first pass never launches more than kNormsBlocks workgroups.  All calls
// on one GPU share it: calls that overlap on one GPU race.
static __device__ __align__(256) NormAcc scratch[kNormsBlocks + 1];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

hipError_t noalloc_norms_scratch(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // hipGetSymbolAddress resolves the copy of the current device.
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

size_t noalloc_norms_scratch_bytes() { return sizeof(scratch); }

hipError_t step_norms(const Grid& g, const float* rho, const float* u, const float* v,
                      const float* w, const float* un, const float* vn, const float* wn,
                      const float* d, int rhs, const uint8_t* flags, void* scratch,
                      smoke_step_norms* out, float dt, long long step, hipStream_t st)
{
    return launch_norms(g, rho, u, v, w, un, vn, wn, d, rhs, flags, scratch, out, dt, step, st);
}

#define SMK_HALF_NORMS_DEFS(H)                                                                     \
    hipError_t step_norms(const Grid& g, const H* rho, const H* u, const H* v, const H* w,         \
                          const H* un, const H* vn, const H* wn, const float* d, int rhs,          \
                          const uint8_t* flags, void* scratch, smoke_step_norms* out, float dt,    \
                          long long step, hipStream_t st)                                          \
    {                                                                                              \
        return launch_norms(g, rho, u, v, w, un, vn, wn, d, rhs, flags, scratch, out, dt, step,    \
                            st);                                                                   \
    }                                                                                              \
    hipError_t step_norms(const Grid& g, const float* rho, const H* u, const H* v, const H* w,     \
                          const H* un, const H* vn, const H* wn, const float* d, int rhs,          \
                          const uint8_t* flags, void* scratch, smoke_step_norms* out, float dt,    \
                          long long step, hipStream_t st)                                          \
    {                                                                                              \
        return launch_norms(g, rho, u, v, w, un, vn, wn, d, rhs, flags, scratch, out, dt, step,    \
                            st);                                                                   \
    }
SMK_HALF_NORMS_DEFS(bf16s)
SMK_HALF_NORMS_DEFS(f16s)
#undef SMK_HALF_NORMS_DEFS

}  // namespace smk
