// Test harness: force-included (-include) when the library's HIP kernel
// sources are compiled as plain host C++ for the CPU emulator.  Kernels then
// run on host threads (emu_runtime.cpp) and "device" memory is malloc'ed, so
// AddressSanitizer checks every kernel access against its allocation and the
// results can be compared with the oracle on small grids.  Never part of the
// shipped library: smokesim only loads the gfx950 build.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>

#undef __shared__
#define __shared__ static  // one block runs at a time, so a static is per-block LDS
#undef __launch_bounds__
#define __launch_bounds__(...)

struct EmuIdx {
    unsigned x, y, z;
};
extern thread_local EmuIdx threadIdx, blockIdx, blockDim, gridDim;

void emu_syncthreads();
#define __syncthreads() emu_syncthreads()

inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline unsigned __float_as_uint(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float __uint_as_float(unsigned u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Lane exchange of a block run on host threads: every thread of the block
// publishes its value and reads its partner's (both sides of a barrier).
unsigned long long emu_exchange(unsigned long long bits, int partner);

// DPP row moves used by the kernels: wave_shr:1 (0x138) and wave_shl:1 (0x130)
// within each wave64; a lane without a source gets 0 (bound_ctrl) or `old`.
int emu_dpp(int old, int src, int ctrl, bool bound_ctrl);
#define __builtin_amdgcn_update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl) \
    emu_dpp((old), (src), (ctrl), (bound_ctrl))
template <typename T>
T __shfl_xor(T v, int mask, int width = 64)
{
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    const int lane = t % width;
    const int partner = (lane ^ mask) < width ? t - lane + (lane ^ mask) : t;
    unsigned long long b = 0;
    std::memcpy(&b, &v, sizeof(T));
    b = emu_exchange(b, partner);
    T r;
    std::memcpy(&r, &b, sizeof(T));
    return r;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

hipError_t emu_last_error();
#define hipGetLastError() emu_last_error()

// Runs (or, on a capturing stream, records) one launch.
void emu_launch(const char* name, std::function<void()> fn, dim3 grid, dim3 block, hipStream_t st);
#undef hipLaunchKernelGGL
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) \
    emu_launch(#K, [=]() { K(__VA_ARGS__); }, dim3(G), dim3(B), ST)

template <class F>
hipError_t emu_occupancy(int* n, F, int threads, size_t)
{
    *n = threads >= 1024 ? 2 : 2048 / threads;
    return hipSuccess;
}
#define hipOccupancyMaxActiveBlocksPerMultiprocessor emu_occupancy
