// k_sort.hip -- exclusive scans and a stable LSD radix sort of (key, id) records on gfx950.
//
// Radix sort: 8-bit digits over the significant bits of the keys only. Each pass is
//   histogram (per 2048-key tile, LDS atomics)  ->  exclusive scan in digit-major order
//   ->  scatter, where each tile ranks its keys stably with 64-lane ballots (8 ballots give the
//   lanes holding the same digit), stages them digit-grouped in LDS, and writes contiguous runs.
// Stability matches the stable argsort of gpu/mitm_gpu2.py:101-104.
#include "tree_dev.h"

namespace {

constexpr int SB = 256;         // scan block
constexpr int SI = 8;           // items per thread
constexpr int ST = SB * SI;     // 2048 per block
static_assert(SI == 8, "the vector path moves 8 items");

constexpr int RB = 256;         // radix block
constexpr int RI = 8;           // items per thread
constexpr int RT = RB * RI;     // 2048-key tiles
constexpr int RADIX = 256;

__device__ __forceinline__ u64 wave_incl_scan(u64 v) { return wave_incl_sum64(v); }

// Block-wide exclusive scan of one value per thread (SB threads = 4 waves). Returns the
// exclusive prefix; *total gets the block sum.
__device__ __forceinline__ u64 block_excl_scan(u64 v, u64 *total) {
    __shared__ u64 s_w[SB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 inc = wave_incl_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    u64 pre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < SB / 64; ++q) {
        if (q < w) pre += s_w[q];
        tot += s_w[q];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(SB) k_scan_reduce(const T *__restrict__ in, size_t n, u64 *__restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * ST;
    u64 s = 0;
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        const size_t e = base + (size_t)k * SB + threadIdx.x;
        if (e < n) s += (u64)in[e];
    }
    u64 tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// out[e] = exclusive prefix of in[] (+ base offset of the block); the last block writes out[n].
template <typename T>
__global__ void __launch_bounds__(SB) k_scan_final(const T *__restrict__ in, size_t n, const u64 *__restrict__ boff,
                                                   u64 *__restrict__ out) {
    const size_t base = (size_t)blockIdx.x * ST + (size_t)threadIdx.x * SI;  // thread-major items
    // a thread's SI items are one run: 16-byte loads and stores when the run is whole and aligned
    const bool vec = base + SI <= n && ((((uintptr_t)in | (uintptr_t)out) & 15) == 0);
    u64 v[SI];
    u64 s = 0;
    if (vec) {
        if constexpr (sizeof(T) == 4) {
            const uint4 a = ((const uint4 *)(in + base))[0], b = ((const uint4 *)(in + base))[1];
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        } else {
#pragma unroll
            for (int k = 0; k < SI / 2; ++k) {
                const ulonglong2 a = ((const ulonglong2 *)(in + base))[k];
                v[2 * k] = a.x;
                v[2 * k + 1] = a.y;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < SI; ++k) {
            const size_t e = base + k;
            v[k] = e < n ? (u64)in[e] : 0;
        }
    }
#pragma unroll
    for (int k = 0; k < SI; ++k) s += v[k];
    u64 tot;
    u64 pre = block_excl_scan(s, &tot) + (boff ? boff[blockIdx.x] : 0);
    if (vec) {
#pragma unroll
        for (int k = 0; k < SI / 2; ++k) {
            const u64 a = pre;
            pre += v[2 * k];
            ((ulonglong2 *)(out + base))[k] = make_ulonglong2(a, pre);
            pre += v[2 * k + 1];
        }
    } else {
#pragma unroll
        for (int k = 0; k < SI; ++k) {
            const size_t e = base + k;
            if (e < n) out[e] = pre;
            pre += v[k];
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SB - 1) out[n] = pre;
}

template <typename T>
void scan_impl(cf_ctx *ctx, const T *in, u64 *out, size_t n) {
    const size_t nb = (n + ST - 1) / ST;
    if (nb <= 1) {
        CF_LAUNCH(ctx, "k_scan_final", (k_scan_final<T>), dim3(1), dim3(SB), 0, in, n, (const u64 *)nullptr, out);
        return;
    }
    DBuf<u64> part(ctx, nb), part_off(ctx, nb + 1);
    CF_LAUNCH(ctx, "k_scan_reduce", (k_scan_reduce<T>), dim3((unsigned)nb), dim3(SB), 0, in, n, part.p);
    scan_impl<u64>(ctx, part.p, part_off.p, nb);
    CF_LAUNCH(ctx, "k_scan_final", (k_scan_final<T>), dim3((unsigned)nb), dim3(SB), 0, in, n, (const u64 *)part_off.p, out);
}

// ---- radix sort -----------------------------------------------------------------------------
__global__ void __launch_bounds__(RB) k_radix_hist(const u64 *__restrict__ keys, size_t n, int shift,
                                                   u32 *__restrict__ hist, size_t ntiles) {
    __shared__ u32 h[RADIX];
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        h[threadIdx.x] = 0;
        __syncthreads();
        const size_t base = tile * RT;
#pragma unroll
        for (int k = 0; k < RI; ++k) {
            const size_t e = base + (size_t)k * RB + threadIdx.x;
            if (e < n) atomicAdd(&h[(keys[e] >> shift) & (RADIX - 1)], 1u);
        }
        __syncthreads();
        hist[(size_t)threadIdx.x * ntiles + tile] = h[threadIdx.x];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RB) k_radix_scatter(const u64 *__restrict__ kin, const u64 *__restrict__ vin, size_t n,
                                                      int shift, const u64 *__restrict__ off, size_t ntiles,
                                                      u64 *__restrict__ kout, u64 *__restrict__ vout) {
    __shared__ u32 s_cnt[RADIX];
    __shared__ u32 s_wc[RB / 64][RADIX];
    __shared__ u32 s_loc[RADIX];
    __shared__ u64 s_goff[RADIX];
    __shared__ u64 s_k[RT];
    __shared__ u64 s_v[RT];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const size_t base = tile * RT;
        const int tcount = (int)((n - base) < (size_t)RT ? (n - base) : (size_t)RT);
        s_goff[threadIdx.x] = off[(size_t)threadIdx.x * ntiles + tile];
        s_cnt[threadIdx.x] = 0;
        u64 k[RI], v[RI];
        u32 rk[RI];
#pragma unroll
        for (int it = 0; it < RI; ++it) {
            const int le = it * RB + threadIdx.x;
            const bool valid = le < tcount;
            k[it] = valid ? kin[base + le] : 0;
            v[it] = valid ? vin[base + le] : 0;
            const u32 d = (u32)((k[it] >> shift) & (RADIX - 1));
            u64 peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const u64 bb = __ballot(bit);
                peers &= bit ? bb : ~bb;
            }
            const u32 rw = (u32)__popcll(peers & lt);
            const u32 cw = (u32)__popcll(peers);
#pragma unroll
            for (int q = 0; q < RB / 64; ++q) s_wc[q][threadIdx.x] = 0;
            __syncthreads();
            if (valid && rw == cw - 1) s_wc[w][d] = cw;
            __syncthreads();
            if (valid) {
                u32 pre = s_cnt[d];
                for (int q = 0; q < w; ++q) pre += s_wc[q][d];
                rk[it] = pre + rw;
            }
            __syncthreads();
            u32 add = 0;
#pragma unroll
            for (int q = 0; q < RB / 64; ++q) add += s_wc[q][threadIdx.x];
            s_cnt[threadIdx.x] += add;
            __syncthreads();
        }
        u64 tot;
        const u64 loc = block_excl_scan((u64)s_cnt[threadIdx.x], &tot);
        s_loc[threadIdx.x] = (u32)loc;
        __syncthreads();
#pragma unroll
        for (int it = 0; it < RI; ++it) {
            const int le = it * RB + threadIdx.x;
            if (le < tcount) {
                const u32 d = (u32)((k[it] >> shift) & (RADIX - 1));
                const u32 at = s_loc[d] + rk[it];
                s_k[at] = k[it];
                s_v[at] = v[it];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < RI; ++it) {
            const int le = it * RB + threadIdx.x;
            if (le < tcount) {
                const u64 key = s_k[le];
                const u32 d = (u32)((key >> shift) & (RADIX - 1));
                const u64 pos = s_goff[d] + (u64)(le - (int)s_loc[d]);
                kout[pos] = key;
                vout[pos] = s_v[le];
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(SB) k_max(const u64 *__restrict__ a, size_t n, unsigned long long *__restrict__ out) {
    u64 m = 0;
    for (size_t e = (size_t)blockIdx.x * SB + threadIdx.x; e < n; e += (size_t)gridDim.x * SB) m = a[e] > m ? a[e] : m;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const u64 o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, (unsigned long long)m);
}

}  // namespace

u64 max_u64(cf_ctx *ctx, const u64 *a, size_t n) {
    DBuf<u64> m(ctx, 1);
    CF_HIP(hipMemsetAsync(m.p, 0, sizeof(u64), ctx->stream));
    if (n) {
        size_t g = (n + SB * 8 - 1) / (SB * 8);
        g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
        CF_LAUNCH(ctx, "k_max", k_max, dim3((unsigned)g), dim3(SB), 0, a, n, (unsigned long long *)m.p);
    }
    return d2h_u64(ctx, m.p);
}

// The same scan with its block partials in a caller's workspace (scan_ws_words(n) words) instead
// of pool allocations.
size_t scan_ws_words(size_t n) {
    const size_t nb = (n + ST - 1) / ST;
    return nb <= 1 ? 0 : 2 * nb + 1 + scan_ws_words(nb);
}

void scan_u64_ws(cf_ctx *ctx, const u64 *in, u64 *out, size_t n, u64 *ws) {
    const size_t nb = (n + ST - 1) / ST;
    if (nb <= 1) {
        CF_LAUNCH(ctx, "k_scan_final", (k_scan_final<u64>), dim3(1), dim3(SB), 0, in, n, (const u64 *)nullptr, out);
        return;
    }
    u64 *part = ws, *part_off = ws + nb;
    CF_LAUNCH(ctx, "k_scan_reduce", (k_scan_reduce<u64>), dim3((unsigned)nb), dim3(SB), 0, in, n, part);
    scan_u64_ws(ctx, part, part_off, nb, ws + 2 * nb + 1);
    CF_LAUNCH(ctx, "k_scan_final", (k_scan_final<u64>), dim3((unsigned)nb), dim3(SB), 0, in, n, (const u64 *)part_off, out);
}

u64 scan_u32_to_u64(cf_ctx *ctx, const u32 *in, u64 *out, size_t n) {
    scan_impl<u32>(ctx, in, out, n);
    return d2h_u64(ctx, out + n);
}

u64 scan_u64(cf_ctx *ctx, const u64 *in, u64 *out, size_t n) {
    scan_impl<u64>(ctx, in, out, n);
    return d2h_u64(ctx, out + n);
}

void scan_u32_to_u64_async(cf_ctx *ctx, const u32 *in, u64 *out, size_t n) { scan_impl<u32>(ctx, in, out, n); }

void radix_sort_pairs(cf_ctx *ctx, u64 *keys, u64 *vals, size_t n, int bits) {
    if (n <= 1 || bits <= 0) return;
    const int passes = (bits + 7) / 8;
    const size_t ntiles = (n + RT - 1) / RT;
    DBuf<u64> k2(ctx, n), v2(ctx, n);
    DBuf<u32> hist(ctx, ntiles * RADIX);
    DBuf<u64> off(ctx, ntiles * RADIX + 1);
    u64 *ka = keys, *va = vals, *kb = k2.p, *vb = v2.p;
    const unsigned grid = (unsigned)(ntiles < 65536 ? ntiles : 65536);
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = ps * 8;
        CF_LAUNCH(ctx, "k_radix_hist", k_radix_hist, dim3(grid), dim3(RB), 0, ka, n, shift, hist.p, ntiles);
        scan_impl<u32>(ctx, hist.p, off.p, ntiles * RADIX);
        CF_LAUNCH(ctx, "k_radix_scatter", k_radix_scatter, dim3(grid), dim3(RB), 0, ka, va, n, shift, off.p, ntiles, kb, vb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        CF_HIP(hipMemcpyAsync(keys, ka, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(vals, va, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
    }
}
