// Mixed-precision pressure solve for gfx950: an FP32 residual-correction loop
// around FP16 Jacobi sweeps, with the global norms reduced in FP64.
//
// The pressure equation is the Jacobi form's acc*P - sum_n P_n = -d on the
// active cells (d = divergence).  With r = d + acc*x - sum_n x_n (the negated
// residual of the current FP32 iterate x), one refinement
//   1. scales r by s = 2^k so that max|r*s| lies in [1/2, 1),
//   2. runs n Jacobi sweeps on acc*e - sum_n e_n = -r*s from e = 0 in packed
//      FP16 (two cells per 32-bit word, v_pk_* arithmetic):
//        e <- inv * sum_n e_n - inv*r*s,   inv = 1/acc (0 off the active set),
//   3. adds x += e / s in FP32 and forms the next r (one fused pass).
// In exact arithmetic refinement with Jacobi inner sweeps from e = 0 is the
// plain Jacobi iteration (x + e_k obeys the same recurrence as x_k), so R
// refinements of n sweeps equal R*n Jacobi sweeps; only the rounding differs.
// |e| <= n * max|inv*r*s| <= n, far inside FP16's range.
//
// Norms: every residual pass writes one FP64 partial sum of r^2 and one
// max|r| per block; a one-block kernel reduces them in a fixed order (so the
// decision is deterministic), sets the scale of the next correction and, with
// a tolerance, the `done` flag once ||r|| / ||d|| <= tol.  Later kernels of the
// step read `done` and idle, so the loop needs no host round trip and can be
// captured.
#include "step_device.h"

namespace smk {

namespace {

using dev::BX;
using dev::BY;

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float h2f(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

// Block reduction of an FP64 sum and an FP32 max (blocks of BX*BY threads), in
// a fixed order; thread 0 stores the block's pair.
__device__ __forceinline__ void block_partials(double s2, float m, double* __restrict__ part,
                                               unsigned* __restrict__ pmax)
{
    __shared__ double ss[BX * BY / 64];
    __shared__ float sm[BX * BY / 64];
    for (int off = 32; off > 0; off >>= 1) {
        s2 += __shfl_xor(s2, off, 64);
        m = fmaxf(m, __shfl_xor(m, off, 64));
    }
    const int t = (int)(threadIdx.x + blockDim.x * threadIdx.y);
    if ((t & 63) == 0) {
        ss[t >> 6] = s2;
        sm[t >> 6] = m;
    }
    __syncthreads();
    if (t == 0) {
        double a = 0.0;
        float b = 0.0f;
        for (int k = 0; k < BX * BY / 64; ++k) {
            a += ss[k];
            b = fmaxf(b, sm[k]);
        }
        const int blk = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
        part[blk] = a;
        pmax[blk] = __float_as_uint(b);
    }
}

// Planes per thread of the norm and residual passes (one block: 64 x 4 cells
// x kZB planes, one partial): enough independent loads per thread to cover
// the memory latency, and enough blocks to fill the GPU.
constexpr int kZB = 4;

// Norms of the right-hand side (the first residual, x = 0): r = d on active
// cells.
__global__ __launch_bounds__(BX * BY) void k_ref_norm0(Grid g, const float* __restrict__ d,
                                                       const uint8_t* __restrict__ flags,
                                                       double* __restrict__ part,
                                                       unsigned* __restrict__ pmax)
{
    const int x = blockIdx.x * BX + threadIdx.x;  // lane x = cell x: lane 0 idles
    const int y = blockIdx.y * BY + threadIdx.y + 1;
    const int z0 = blockIdx.z * kZB + 1;
    double s2 = 0.0;
    float m = 0.0f;
    if (x >= 1 && x <= g.nx - 2 && y <= g.ny - 2) {
        float r[kZB];
#pragma unroll
        for (int k = 0; k < kZB; ++k) {
            const int z = z0 + k;
            r[k] = 0.0f;
            if (z <= g.nz - 2) {
                const int i = g.idx(x, y, z);
                r[k] = (flags[i] & F_ACTIVE) ? d[i] : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < kZB; ++k) {
            s2 += (double)r[k] * (double)r[k];
            m = fmaxf(m, fabsf(r[k]));
        }
    }
    block_partials(s2, m, part, pmax);
}

// x_new = x_old + e/s on every interior cell (x_old = 0 when XZERO) and the
// next negated residual r = (d + acc*x_new) - sum_n x_new_n on active cells,
// with its partial norms; LAST skips the r store.  Neighbour values are
// recomputed with the same two operations, so they equal the neighbours' own
// x_new.  A thread walks kZB planes of one column: every load of the column
// and of its x/y neighbours is issued first (plane indices clamped into the
// padded field, so none is predicated), then the planes are computed and the
// stores issued last, so the loads stay in flight together.  Once `done` is
// set the pass only copies x_old.
template <bool XZERO, bool LAST>
__global__ __launch_bounds__(BX * BY, 8) void k_ref_resid(Grid g, const float* __restrict__ xo,
                                                       float* __restrict__ xn,
                                                       const float* __restrict__ d,
                                                       const uint8_t* __restrict__ flags,
                                                       const uint16_t* __restrict__ e,
                                                       float* __restrict__ q,
                                                       const RefineCtl* __restrict__ ctl,
                                                       double* __restrict__ part,
                                                       unsigned* __restrict__ pmax)
{
    const int done = ctl->done;
    const float sinv = ctl->sinv;
    const int x = blockIdx.x * BX + threadIdx.x;  // lane x = cell x: lane 0 idles
    const int y = blockIdx.y * BY + threadIdx.y + 1;
    const int z0 = blockIdx.z * kZB + 1;
    const int nk = min(kZB, g.nz - 1 - z0);  // planes of this block (>= 1)
    double s2 = 0.0;
    float m = 0.0f;
    if (x >= 1 && x <= g.nx - 2 && y <= g.ny - 2) {
        const int i0 = g.idx(x, y, z0);
        if (done) {
#pragma unroll
            for (int k = 0; k < kZB; ++k)
                if (k < nk) xn[i0 + k * g.plane] = XZERO ? 0.0f : xo[i0 + k * g.plane];
        } else {
            const int row = g.idx(x, y, 0);
            auto plane_of = [&](int z) { return row + min(z, g.nz) * g.plane; };
            // Column x_old and e at planes z0-1 .. z0+kZB, neighbours at z0 .. z0+kZB-1.
            float cxo[kZB + 2], ce[kZB + 2];
#pragma unroll
            for (int k = 0; k < kZB + 2; ++k) {
                const int j = plane_of(z0 - 1 + k);
                cxo[k] = XZERO ? 0.0f : xo[j];
                ce[k] = h2f(e[j]);
            }
            float nxo[kZB][4], ne[kZB][4], dd[kZB];
            uint8_t fl[kZB];
#pragma unroll
            for (int k = 0; k < kZB; ++k) {
                const int i = plane_of(z0 + k);
                const int nb[4] = {i - 1, i + 1, i - g.px, i + g.px};
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    nxo[k][n] = XZERO ? 0.0f : xo[nb[n]];
                    ne[k][n] = h2f(e[nb[n]]);
                }
                dd[k] = d[i];
                fl[k] = flags[i];
            }
            float col[kZB + 2];
#pragma unroll
            for (int k = 0; k < kZB + 2; ++k) col[k] = cxo[k] + ce[k] * sinv;
            float rr[kZB];
#pragma unroll
            for (int k = 0; k < kZB; ++k) {
                rr[k] = 0.0f;
                if (k < nk && (fl[k] & F_ACTIVE)) {
                    float X[4];
#pragma unroll
                    for (int n = 0; n < 4; ++n) X[n] = nxo[k][n] + ne[k][n] * sinv;
                    const float sum = dev::sum6(X[0], X[1], X[2], X[3], col[k], col[k + 2]);
                    const float r = (dd[k] + (float)__popc(fl[k] & F_NBRS) * col[k + 1]) - sum;
                    rr[k] = r;
                    s2 += (double)r * (double)r;
                    m = fmaxf(m, fabsf(r));
                }
            }
#pragma unroll
            for (int k = 0; k < kZB; ++k) {
                if (k >= nk) break;
                const int i = i0 + k * g.plane;
                xn[i] = col[k + 1];
                if (!LAST) q[i] = rr[k];
            }
        }
    }
    block_partials(s2, m, part, pmax);
}

// First stage of the partials reduction: block b reduces the fixed chunk
// [b*chunk, (b+1)*chunk) in a fixed order.
__global__ __launch_bounds__(256) void k_ref_reduce(const double* __restrict__ part,
                                                    const unsigned* __restrict__ pmax, int n,
                                                    int chunk, double* __restrict__ part2,
                                                    unsigned* __restrict__ pmax2,
                                                    const RefineCtl* __restrict__ ctl, int stage)
{
    __shared__ double ss[256];
    __shared__ unsigned sm[256];
    if (stage > 0 && ctl->done) return;  // uniform
    const int t = (int)threadIdx.x;
    const int b0 = (int)blockIdx.x * chunk, b1 = min(b0 + chunk, n);
    double a = 0.0;
    unsigned m = 0;
    for (int k = b0 + t; k < b1; k += 256) {
        a += part[k];
        m = max(m, pmax[k]);
    }
    ss[t] = a;
    sm[t] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] += ss[t + w];
            sm[t] = max(sm[t], sm[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part2[blockIdx.x] = ss[0];
        pmax2[blockIdx.x] = sm[0];
    }
}

// Fixed-order reduction of the first-stage partials and the control update
// (one block of 1024 threads).  stage 0: norms of d, a new step (resets
// done/refinements); stage > 0: the residual after a refinement.
__global__ __launch_bounds__(1024) void k_ref_finish(RefineCtl* __restrict__ ctl,
                                                     const double* __restrict__ part,
                                                     const unsigned* __restrict__ pmax, int n,
                                                     int stage, double tol)
{
    __shared__ double ss[1024];
    __shared__ unsigned sm[1024];
    const int t = (int)threadIdx.x;
    if (stage > 0 && ctl->done) return;  // uniform: keep the converged state
    double a = 0.0;
    unsigned b = 0;
    for (int k = t; k < n; k += 1024) {
        a += part[k];
        b = max(b, pmax[k]);
    }
    ss[t] = a;
    sm[t] = b;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] += ss[t + w];
            sm[t] = max(sm[t], sm[t + w]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    const double r2 = ss[0];
    const unsigned qmax = sm[0];
    if (stage == 0) {
        ctl->bnorm2 = r2;
        ctl->refinements = 0;
        ctl->done = 0;
    } else {
        ctl->refinements = stage;
    }
    ctl->rnorm2 = r2;
    const double bn = ctl->bnorm2;
    ctl->rel = bn > 0.0 ? sqrt(r2 / bn) : 0.0;
    ctl->qmax = qmax;
    // s = 2^k with max|r| * s in [1/2, 1): k = 126 - biased exponent.
    int k = 0;
    if (qmax != 0) {
        k = 126 - (int)(qmax >> 23);
        k = k < -126 ? -126 : (k > 126 ? 126 : k);
    }
    ctl->s = __uint_as_float((unsigned)(k + 127) << 23);
    ctl->sinv = __uint_as_float((unsigned)(127 - k) << 23);
    if (qmax == 0 || (tol > 0.0 && ctl->rel <= tol)) ctl->done = 1;
}

// LDS of one FP16 sweep block: per level, two planes of the grown tile of
// TXP x TY cell pairs, padded by one pair on each side.
template <int TXP, int TY, int T>
struct HalfLds {
    uint32_t v[T][2][TY + 2][TXP + 2];
    uint16_t itab[256];  // FP16 diagonal inverse of each flags byte
};

// T fused FP16 Jacobi sweeps on the correction for one tile, temporal blocking
// as in the FP32 kernel (step_device.h: jacobi_fused_levels), with each thread
// owning a pair of adjacent columns (x0, x0+1), x0 even, in R adjacent tile
// rows.  The grown tile is 2*TXP cells wide and TY rows high; the inner
// 2*TXP - 2T columns and TY - 2T rows are written (T even, so the inner
// region is whole pairs).  Pairs cover x in [0, px): the cells beyond the
// interior have inv = 0 and r = 0, so they stay e = 0.  The diagonal inverse
// of a cell is the FP16 rounding of cell_inv(flags byte), from a 256-entry
// LDS table: the lane loads its pair's two flags bytes (one byte per cell
// instead of a two-byte FP16 stream).  Flags beyond the grid (padding
// columns, row and plane) are 0 from creation on (only k_flags writes flags,
// over the grid), so they give inv = 0 there.  FIRST: e starts at 0 and the
// scaled right-hand side inv*r*s is formed from the FP32 r and written for
// the later launches of the refinement.
//
// Loads and stores go through buffer descriptors over the block's planes
// [plo, phi): a plane outside the span or a lane outside the grid (offsets
// from 2^31) reads 0, as the cells there must.  A wave none of whose rows is
// valid at level t skips that level and the ones above (NL < T), as in the
// FP32 kernel.
template <int TXP, int TY, int T, bool FIRST, int R, int NL>
__device__ __forceinline__ void ref_sweep_levels(const Grid& g, const uint32_t* __restrict__ ein,
                                                 uint32_t* __restrict__ eout,
                                                 const float* __restrict__ q,
                                                 uint32_t* __restrict__ qi,
                                                 const uint16_t* __restrict__ fl2, float s, int zc,
                                                 int tile_x, int tile_y, int tile_z,
                                                 HalfLds<TXP, TY, T>& sh)
{
    static_assert(NL >= 0 && NL <= T, "levels of a wave");
    const int tx = threadIdx.x;
    const int r0 = (int)threadIdx.y * R;  // first tile row of this thread
    constexpr int OXC = 2 * TXP - 2 * T, OY = TY - 2 * T;
    const int x0 = tile_x * OXC + 2 * tx - T;
    const int z0 = tile_z * zc + 1;
    const int z1 = min(z0 + zc, g.nz - 1);
    const int plo = max(z0 - T, 0), phi = min(z1 + T, g.nz + 1);
    const uint32_t wplane = (uint32_t)g.plane >> 1;       // pairs per plane
    const uint32_t span = (uint32_t)(phi - plo) * wplane;  // pairs
    const size_t firstw = (size_t)plo * wplane;
    const __amdgpu_buffer_rsrc_t rf = dev::buf_rsrc(fl2 + firstw, span * 2u);
    const __amdgpu_buffer_rsrc_t re = dev::buf_rsrc(ein + firstw, FIRST ? 0u : span * 4u);
    const __amdgpu_buffer_rsrc_t rq = dev::buf_rsrc(qi + firstw, span * 4u);
    const __amdgpu_buffer_rsrc_t rr = dev::buf_rsrc(q + 2 * firstw, FIRST ? span * 8u : 0u);
    const __amdgpu_buffer_rsrc_t ro = dev::buf_rsrc(eout + firstw, span * 4u);
    uint32_t o4[R];  // byte offset of the lane's pair word in plane plo (0x80000000: outside)
    uint32_t o8[R];  // the same for the FP32 pair of the FIRST launch's residual
    bool writes[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int r = r0 + j;
        const int y = tile_y * OY + r - T + 1;
        const bool in_grid = x0 >= 0 && x0 < g.px && y >= 0 && y <= g.ny;
        writes[j] = tx >= T / 2 && tx < TXP - T / 2 && r >= T && r < TY - T && x0 >= 0 &&
                    x0 < g.px && y >= 1 && y <= g.ny - 2;
        o4[j] = in_grid ? (uint32_t)(g.idx(x0, y, 0) >> 1) * 4u : 0x80000000u;
        o8[j] = in_grid ? (uint32_t)(g.idx(x0, y, 0) >> 1) * 8u : 0x80000000u;
    }

    const h2 zero2 = {(_Float16)0.0f, (_Float16)0.0f};
    h2 w[R][T][3];
    h2 iq[R][T + 1], qq[R][T + 1];
#pragma unroll
    for (int j = 0; j < R; ++j) {
#pragma unroll
        for (int t = 0; t < T; ++t) w[j][t][0] = w[j][t][1] = w[j][t][2] = zero2;
#pragma unroll
        for (int t = 0; t <= T; ++t) iq[j][t] = qq[j][t] = zero2;
    }

    // Prefetch ring (plane k + i in slot i), dev::kPrefetch planes deep.  The
    // ring holds loaded values only (FIRST: the raw FP32 r); anything computed
    // from them is computed when the plane is consumed, so no wait is forced
    // at the load.  Offsets: 4 bytes per pair word, 2 per flags pair, 8 per
    // FP32 pair (a lane outside the grid stays past every span).
    constexpr int D = dev::kPrefetch;
    uint32_t ne[R][D], nq[R][D];
    uint16_t nfl[R][D];  // the pair's two flags bytes
    float nr0[R][D], nr1[R][D];
    auto load = [&](int kn, int i) {
        const uint32_t dk = (uint32_t)(kn - plo) * wplane;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint32_t o = o4[j] + dk * 4u;
            nfl[j][i] = __builtin_amdgcn_raw_buffer_load_b16(rf, (int)(o >> 1), 0, 0);
            if (FIRST) {
                const auto v = __builtin_amdgcn_raw_buffer_load_b64(rr, (int)(o8[j] + dk * 8u), 0, 0);
                // Elements through scalars: a bit cast of v[1] itself reads
                // element 0 with this compiler.
                const uint32_t v0 = v[0], v1 = v[1];
                nr0[j][i] = __builtin_bit_cast(float, v0);
                nr1[j][i] = __builtin_bit_cast(float, v1);
            } else {
                ne[j][i] = __builtin_amdgcn_raw_buffer_load_b32(re, (int)o, 0, 0);
                nq[j][i] = __builtin_amdgcn_raw_buffer_load_b32(rq, (int)o, 0, 0);
            }
        }
    };
#pragma unroll
    for (int i = 0; i < D; ++i) load(z0 - T + i, i);
    // Unrolled by lcm(3, T+1) so the register windows rotate without copies;
    // the padded iterations load and write nothing.
    constexpr int U = dev::z_unroll(T);
    const int kend = z0 - T + ((z1 - z0 + 2 * T + U - 1) / U) * U;
    for (int kb = z0 - T; kb < kend; kb += U)
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int k = kb + u;
        h2 qv[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint16_t fp = nfl[j][0];  // 0 outside: inv = 0
            const h2 ii = {__builtin_bit_cast(_Float16, sh.itab[fp & 0xff]),
                           __builtin_bit_cast(_Float16, sh.itab[fp >> 8])};
            const h2 p0 = FIRST ? zero2 : __builtin_bit_cast(h2, ne[j][0]);
            if (FIRST) {
                qv[j].x = dev::to_f16(nr0[j][0] * s * (float)ii.x);
                qv[j].y = dev::to_f16(nr1[j][0] * s * (float)ii.y);
            } else {
                qv[j] = __builtin_bit_cast(h2, nq[j][0]);
            }
#pragma unroll
            for (int i = 0; i + 1 < D; ++i) {
                ne[j][i] = ne[j][i + 1];
                nfl[j][i] = nfl[j][i + 1];
                nq[j][i] = nq[j][i + 1];
                nr0[j][i] = nr0[j][i + 1];
                nr1[j][i] = nr1[j][i + 1];
            }
#pragma unroll
            for (int t = T; t > 0; --t) {
                iq[j][t] = iq[j][t - 1];
                qq[j][t] = qq[j][t - 1];
            }
            iq[j][0] = ii;
            qq[j][0] = qv[j];
            w[j][0][0] = w[j][0][1];
            w[j][0][1] = w[j][0][2];
            w[j][0][2] = p0;
        }
        load(k + D, D - 1);
        const int cur = k & 1, prev = cur ^ 1;
        h2 newv[R][T + 1];
#pragma unroll
        for (int j = 0; j < R; ++j) newv[j][0] = w[j][0][2];
#pragma unroll
        for (int t = 1; t <= NL; ++t) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                // x neighbours from the adjacent lanes (DPP); y neighbours from
                // this thread's rows j-1 / j+1, or from LDS (tile row r lives
                // in LDS row r + 1).
                const h2 C = w[j][t - 1][1];
                const uint32_t cw = __builtin_bit_cast(uint32_t, C);
                const h2 Lp = __builtin_bit_cast(h2, dev::lane_prev(cw));
                const h2 Rp = __builtin_bit_cast(h2, dev::lane_next(cw));
                const h2 Up = (j > 0) ? w[j > 0 ? j - 1 : 0][t - 1][1]
                                      : __builtin_bit_cast(h2, sh.v[t - 1][prev][r0][tx + 1]);
                const h2 Dp = (j + 1 < R) ? w[j + 1 < R ? j + 1 : 0][t - 1][1]
                                          : __builtin_bit_cast(h2, sh.v[t - 1][prev][r0 + R + 1][tx + 1]);
                const h2 xm = {Lp.y, C.x};
                const h2 xp = {C.y, Rp.x};
                // Tree order, the z+ value (level t-1 of this iteration) last:
                // the chain through the levels is one add and one fma per level.
                const h2 part = ((xm + xp) + (Up + Dp)) + w[j][t - 1][0];
                const h2 sum = part + w[j][t - 1][2];
                newv[j][t] = __builtin_elementwise_fma(iq[j][t], sum, -qq[j][t]);
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
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (NL == T && writes[j] && zo >= z0 && zo < z1)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, newv[j][NL]), ro,
                                                      (int)(o4[j] + (uint32_t)(zo - plo) * wplane * 4u), 0, 0);
            // The scaled right-hand side for the later launches of the
            // refinement, stored as its plane is consumed: a store issued
            // among the prefetch loads would make every later wait drain all
            // outstanding loads.
            if (FIRST && writes[j] && k >= z0 && k < z1)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, qv[j]), rq,
                                                      (int)(o4[j] + (uint32_t)(k - plo) * wplane * 4u), 0, 0);
        }
#pragma unroll
        for (int t = 0; t < T && t <= NL; ++t)
#pragma unroll
            for (int j = 0; j < R; ++j)
                sh.v[t][cur][r0 + j + 1][tx + 1] = __builtin_bit_cast(uint32_t, newv[j][t]);
        __syncthreads();
    }
}

template <int TXP, int TY, int T, bool FIRST, int R, int K>
__device__ __forceinline__ void ref_sweep_dispatch(int nl, const Grid& g, const uint32_t* __restrict__ ein,
                                                   uint32_t* __restrict__ eout,
                                                   const float* __restrict__ q,
                                                   uint32_t* __restrict__ qi,
                                                   const uint16_t* __restrict__ fl2, float s, int zc,
                                                   int tile_x, int tile_y, int tile_z,
                                                   HalfLds<TXP, TY, T>& sh)
{
    if constexpr (K == 0) {
        ref_sweep_levels<TXP, TY, T, FIRST, R, 0>(g, ein, eout, q, qi, fl2, s, zc, tile_x, tile_y,
                                                  tile_z, sh);
    } else {
        if (nl == K)
            ref_sweep_levels<TXP, TY, T, FIRST, R, K>(g, ein, eout, q, qi, fl2, s, zc, tile_x,
                                                      tile_y, tile_z, sh);
        else
            ref_sweep_dispatch<TXP, TY, T, FIRST, R, K - 1>(nl, g, ein, eout, q, qi, fl2, s, zc,
                                                            tile_x, tile_y, tile_z, sh);
    }
}

// One tile per block of TXP x (TY / R) threads.  Requires zc <=
// fused_max_zc(plane) (32-bit buffer offsets).
template <int TXP, int TY, int T, bool FIRST, int R>
__global__ __launch_bounds__(TXP * TY / R, 8 / R) void k_ref_sweeps(Grid g, const uint32_t* __restrict__ ein,
                                                                 uint32_t* __restrict__ eout,
                                                                 const float* __restrict__ q,
                                                                 uint32_t* __restrict__ qi,
                                                                 const uint16_t* __restrict__ fl2,
                                                                 const RefineCtl* __restrict__ ctl, int zc)
{
    static_assert(T % 2 == 0, "pairs: the halo must be whole pairs");
    static_assert(R >= 1 && TY % R == 0, "whole tile rows per thread");
    static_assert(64 % TXP == 0, "a wave holds whole thread rows");
    if (ctl->done) return;  // uniform across the grid
    __shared__ HalfLds<TXP, TY, T> sh;
    constexpr int NT = TXP * TY / R;  // threads
    {
        uint32_t* flat = &sh.v[0][0][0][0];
        const int n = T * 2 * (TY + 2) * (TXP + 2);
        for (int k = (int)(threadIdx.x + TXP * threadIdx.y); k < n; k += NT) flat[k] = 0u;
        for (int f = (int)(threadIdx.x + TXP * threadIdx.y); f < 256; f += NT)
            sh.itab[f] = __builtin_bit_cast(uint16_t, dev::to_f16(dev::cell_inv((uint8_t)f)));
    }
    __syncthreads();
    // XCD-contiguous tile order, as in the FP32 fused kernel.
    const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
    const int nb = nbx * nby * (int)gridDim.z;
    const int b = (int)blockIdx.x + nbx * ((int)blockIdx.y + nby * (int)blockIdx.z);
    const int full = nb & ~7;
    const int L = (b < full) ? (b & 7) * (full >> 3) + (b >> 3) : b;
    const int tile_x = L % nbx, tile_y = (L / nbx) % nby, tile_z = L / (nbx * nby);
    const float s = FIRST ? ctl->s : 0.0f;
    // Levels the wave needs: its tile rows r0w .. r0w + RPW - 1.
    constexpr int TPW = 64 / TXP, RPW = TPW * R;
    const int r0w = __builtin_amdgcn_readfirstlane((int)threadIdx.y / TPW * TPW) * R;
    const int nl = min(T, min(r0w + RPW - 1, TY - 1 - r0w));
    ref_sweep_dispatch<TXP, TY, T, FIRST, R, T>(nl, g, ein, eout, q, qi, fl2, s, zc, tile_x, tile_y,
                                                tile_z, sh);
}

inline dim3 resid_grid(const Grid& g)
{
    return dim3((unsigned)((g.nx - 1 + BX - 1) / BX), (unsigned)((g.ny - 2 + BY - 1) / BY),
                (unsigned)((g.nz - 2 + kZB - 1) / kZB));
}

// Two-stage fixed-order reduction of n partials (second stage in the
// finish kernel), then the control update.
hipError_t reduce_and_finish(const RefineBufs& b, int n, int stage, double tol, hipStream_t st)
{
    const int chunk = (n + kRefineStage2 - 1) / kRefineStage2;
    const int na = (n + chunk - 1) / chunk;
    double* part2 = b.part + n;
    unsigned* pmax2 = b.pmax + n;
    hipLaunchKernelGGL(k_ref_reduce, dim3((unsigned)na), dim3(256), 0, st, (const double*)b.part,
                       (const unsigned*)b.pmax, n, chunk, part2, pmax2, (const RefineCtl*)b.ctl,
                       stage);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ref_finish, dim3(1), dim3(1024), 0, st, b.ctl, (const double*)part2,
                       (const unsigned*)pmax2, na, stage, tol);
    return hipGetLastError();
}

// Tiles of 32 column pairs x 32 rows, one row per thread, or x 64 rows, two
// rows per thread (1024 threads either way).
template <int T, bool FIRST, int R>
hipError_t launch_sweeps(const Grid& g, const RefineBufs& b, const float* q, const uint8_t* flags,
                         int ein, int eout, int zc, hipStream_t st)
{
    // Two flags bytes per 16-bit word: the pair (x0, x0+1) of word x0/2.
    const uint16_t* flags2 = reinterpret_cast<const uint16_t*>(flags);
    constexpr int TXP = 32, TY = 32 * R;
    constexpr int OXC = 2 * TXP - 2 * T, OY = TY - 2 * T;
    dim3 grid((unsigned)((g.px + OXC - 1) / OXC), (unsigned)((g.ny - 2 + OY - 1) / OY),
              (unsigned)((g.nz - 2 + zc - 1) / zc));
    hipLaunchKernelGGL((k_ref_sweeps<TXP, TY, T, FIRST, R>), grid, dim3(TXP, TY / R), 0, st, g,
                       (const uint32_t*)b.e[ein], b.e[eout], q, b.qi, flags2, b.ctl,
                       zc);
    return hipGetLastError();
}

template <int R>
hipError_t launch_sweeps_t(const Grid& g, const RefineBufs& b, const float* q, const uint8_t* flags,
                           int first, int ein, int eout, int sweeps, int zc, hipStream_t st)
{
    switch (sweeps * 2 + (first ? 1 : 0)) {
    case 4: return launch_sweeps<2, false, R>(g, b, q, flags, ein, eout, zc, st);
    case 5: return launch_sweeps<2, true, R>(g, b, q, flags, ein, eout, zc, st);
    case 8: return launch_sweeps<4, false, R>(g, b, q, flags, ein, eout, zc, st);
    case 9: return launch_sweeps<4, true, R>(g, b, q, flags, ein, eout, zc, st);
    case 12: return launch_sweeps<6, false, R>(g, b, q, flags, ein, eout, zc, st);
    case 13: return launch_sweeps<6, true, R>(g, b, q, flags, ein, eout, zc, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int refine_parts(const Grid& g)
{
    const dim3 gr = resid_grid(g);
    return (int)(gr.x * gr.y * gr.z);
}

hipError_t refine_norm0(const Grid& g, const float* d, const uint8_t* flags, const RefineBufs& b,
                        double tol, hipStream_t st)
{
    const dim3 grid = resid_grid(g);
    hipLaunchKernelGGL(k_ref_norm0, grid, dim3(BX, BY), 0, st, g, d, flags, b.part, b.pmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return reduce_and_finish(b, (int)(grid.x * grid.y * grid.z), 0, tol, st);
}

hipError_t refine_sweeps(const Grid& g, const RefineBufs& b, const float* q, const uint8_t* flags,
                         int first, int ein, int eout, int sweeps, int zc, int two_rows,
                         hipStream_t st)
{
    // The z-chunk is a speed parameter: cap it so a block's span of planes
    // fits the 32-bit buffer offsets (planes too wide for any chunk: error).
    zc = std::min(zc, fused_max_zc(g.plane));
    if (zc < 1) return hipErrorInvalidValue;
    return two_rows ? launch_sweeps_t<2>(g, b, q, flags, first, ein, eout, sweeps, zc, st)
                    : launch_sweeps_t<1>(g, b, q, flags, first, ein, eout, sweeps, zc, st);
}

hipError_t refine_resid(const Grid& g, const float* xo, float* xn, const float* d,
                        const uint8_t* flags, const RefineBufs& b, int ein, int xzero, int last,
                        int stage, double tol, hipStream_t st)
{
    const dim3 grid = resid_grid(g);
    const uint16_t* e = reinterpret_cast<const uint16_t*>(b.e[ein]);
    if (xzero && last)
        hipLaunchKernelGGL((k_ref_resid<true, true>), grid, dim3(BX, BY), 0, st, g, xo, xn, d, flags,
                           e, b.q, (const RefineCtl*)b.ctl, b.part, b.pmax);
    else if (xzero)
        hipLaunchKernelGGL((k_ref_resid<true, false>), grid, dim3(BX, BY), 0, st, g, xo, xn, d,
                           flags, e, b.q, (const RefineCtl*)b.ctl, b.part, b.pmax);
    else if (last)
        hipLaunchKernelGGL((k_ref_resid<false, true>), grid, dim3(BX, BY), 0, st, g, xo, xn, d,
                           flags, e, b.q, (const RefineCtl*)b.ctl, b.part, b.pmax);
    else
        hipLaunchKernelGGL((k_ref_resid<false, false>), grid, dim3(BX, BY), 0, st, g, xo, xn, d,
                           flags, e, b.q, (const RefineCtl*)b.ctl, b.part, b.pmax);
    hipError_t e2 = hipGetLastError();
    if (e2 != hipSuccess) return e2;
    return reduce_and_finish(b, (int)(grid.x * grid.y * grid.z), stage, tol, st);
}

}  // namespace smk
