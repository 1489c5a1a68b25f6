// k_join.hip -- equal-key join of two key-sorted record sides on gfx950.
//
// Semantics of gpu/mitm_gpu2.py:107-119 (join_sorted): for every I row, in I order, the run
// [lo, hi) of D rows with an equal key, emitted in D order. count -> scan -> write, so the
// output order is exactly the reference's (I-major, then D position).
#include "cf_internal.h"

namespace {
constexpr int BLK = 256;

__device__ __forceinline__ u64 lower_bound_u64(const u64 *__restrict__ a, u64 n, u64 v) {
    u64 lo = 0, hi = n;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ u64 upper_bound_from(const u64 *__restrict__ a, u64 lo, u64 n, u64 v) {
    u64 hi = n;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(BLK) k_join_count(const u64 *__restrict__ Dk, u64 nD, const u64 *__restrict__ Ik, u64 nI,
                                                    u64 *__restrict__ lo, u64 *__restrict__ cnt) {
    for (u64 i = (u64)blockIdx.x * BLK + threadIdx.x; i < nI; i += (u64)gridDim.x * BLK) {
        const u64 k = Ik[i];
        const u64 l = lower_bound_u64(Dk, nD, k);
        u64 c = 0;
        if (l < nD && Dk[l] == k) c = upper_bound_from(Dk, l, nD, k) - l;
        lo[i] = l;
        cnt[i] = c;
    }
}

__global__ void __launch_bounds__(BLK) k_join_write(const u64 *__restrict__ Did, const u64 *__restrict__ Iid, u64 nI,
                                                    const u64 *__restrict__ lo, const u64 *__restrict__ cnt,
                                                    const u64 *__restrict__ off, u64 *__restrict__ od,
                                                    u64 *__restrict__ oi, u64 *__restrict__ okey,
                                                    const u64 *__restrict__ Ik) {
    for (u64 i = (u64)blockIdx.x * BLK + threadIdx.x; i < nI; i += (u64)gridDim.x * BLK) {
        const u64 c = cnt[i];
        if (!c) continue;
        const u64 o = off[i], l = lo[i], iid = Iid[i];
        for (u64 q = 0; q < c; ++q) {
            od[o + q] = Did[l + q];
            oi[o + q] = iid;
            if (okey) okey[o + q] = Ik[i];
        }
    }
}

inline unsigned grid_for(u64 n) {
    u64 g = (n + BLK - 1) / BLK;
    if (g < 1) g = 1;
    return (unsigned)(g > (1u << 20) ? (1u << 20) : g);
}
}  // namespace

u64 join_sorted_count(cf_ctx *ctx, const u64 *Dk, size_t nD, const u64 *Ik, size_t nI, u64 *lo, u64 *cnt, u64 *off) {
    if (!nI) return 0;
    CF_LAUNCH(ctx, "k_join_count", k_join_count, dim3(grid_for(nI)), dim3(BLK), 0, Dk, (u64)nD, Ik, (u64)nI, lo, cnt);
    return scan_u64(ctx, cnt, off, nI);
}

void join_sorted_write(cf_ctx *ctx, const u64 *Did, const u64 *Iid, size_t nI, const u64 *lo, const u64 *cnt,
                       const u64 *off, u64 *od, u64 *oi, u64 *okey, const u64 *Ik) {
    if (!nI) return;
    CF_LAUNCH(ctx, "k_join_write", k_join_write, dim3(grid_for(nI)), dim3(BLK), 0, Did, Iid, (u64)nI, lo, cnt, off, od,
                       oi, okey, Ik);
}
