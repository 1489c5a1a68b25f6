// k_join.hip -- equal-key join of two key-sorted record sides on gfx950.
//
// Semantics of gpu/mitm_gpu2.py:107-119 (join_sorted): for every I row, in I order, the run
// [lo, hi) of D rows with an equal key, emitted in D order. count -> scan -> write, so the
// output order is exactly the reference's (I-major, then D position).
#include <map>
#include <mutex>

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

// Writes nothing when the join's total (off[nI]) exceeds cap.
__global__ void __launch_bounds__(BLK) k_join_write(const u64 *__restrict__ Did, const u64 *__restrict__ Iid, u64 nI,
                                                    const u64 *__restrict__ lo, const u64 *__restrict__ cnt,
                                                    const u64 *__restrict__ off, u64 *__restrict__ od,
                                                    u64 *__restrict__ oi, u64 *__restrict__ okey,
                                                    const u64 *__restrict__ Ik, u64 cap) {
    if (off[nI] > cap) return;
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
                       oi, okey, Ik, ~0ULL);
}

// ---- the join without allocation ---------------------------------------------------------------
// cf_join_sorted_noalloc keeps its per-row run starts, counts and offsets, and the scan's block
// partials, in this file-scope array: one copy per device, resident while the library is loaded.
// In host code `scratch` names a host-side stand-in, not device memory: its device address comes
// from hipGetSymbolAddress, once per device (join_scratch_bind, called when a context is created),
// and only that address is passed to kernels.
namespace {
constexpr u64 NOALLOC_MAX_I = (u64)CF_JOIN_NOALLOC_MAX_I;
constexpr size_t scan_words_bound(size_t n) {   // scan_ws_words, evaluated at compile time
    return (n + 2047) / 2048 <= 1 ? 0 : 2 * ((n + 2047) / 2048) + 1 + scan_words_bound((n + 2047) / 2048);
}
constexpr size_t SCRATCH_WORDS = 3 * NOALLOC_MAX_I + 1 + scan_words_bound(NOALLOC_MAX_I);
static __device__ __align__(256) u64 scratch[SCRATCH_WORDS];

std::mutex g_scratch_mu;
std::map<int, u64 *> g_scratch_dev;   // device -> address of its scratch
}  // namespace

u64 *join_scratch_bind(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_dev.find(device);
    if (it != g_scratch_dev.end()) return it->second;
    int cur = -1;
    CF_HIP(hipGetDevice(&cur));
    CF_REQUIRE(cur == device, "join scratch bound with another device current");
    void *p = nullptr;
    CF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CF_REQUIRE(p != nullptr, "no device address for the join scratch");
    g_scratch_dev[device] = (u64 *)p;
    return (u64 *)p;
}

extern "C" int cf_join_sorted_noalloc(cf_ctx *ctx, const uint64_t *Dk, const uint64_t *Did, uint64_t nD,
                                      const uint64_t *Ik, const uint64_t *Iid, uint64_t nI, uint64_t *out_d,
                                      uint64_t *out_i, uint64_t cap, uint64_t *npairs) {
    API_BEGIN
    CF_REQUIRE(ctx && npairs, "NULL argument");
    CF_REQUIRE(ctx->join_scratch, "the context has no join scratch");
    CF_REQUIRE(nI <= NOALLOC_MAX_I, "nI above CF_JOIN_NOALLOC_MAX_I: join the I side in slices");
    CF_REQUIRE(cap == 0 || (out_d && out_i), "NULL output with a non-zero capacity");
    static_assert(SCRATCH_WORDS * sizeof(u64) <= (64u << 20), "the join scratch stays modest");
    *npairs = 0;
    if (nD == 0 || nI == 0) return CF_OK;
    u64 *lo = ctx->join_scratch, *cnt = lo + NOALLOC_MAX_I, *off = cnt + NOALLOC_MAX_I;
    u64 *ws = off + NOALLOC_MAX_I + 1;
    CF_LAUNCH(ctx, "k_join_count", k_join_count, dim3(grid_for(nI)), dim3(BLK), 0, Dk, (u64)nD, Ik, (u64)nI, lo, cnt);
    scan_u64_ws(ctx, cnt, off, nI, ws);
    CF_LAUNCH(ctx, "k_join_write", k_join_write, dim3(grid_for(nI)), dim3(BLK), 0, Did, Iid, (u64)nI, lo, cnt, off,
              out_d, out_i, (u64 *)nullptr, Ik, (u64)cap);
    const u64 tot = d2h_u64(ctx, off + nI);
    *npairs = tot;
    CF_REQUIRE(tot <= cap, "the output capacity is below the pair count (*npairs); nothing was written");
    API_END
}
