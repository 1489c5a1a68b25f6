// k_join.hip -- equal-key join of two key-sorted record sides on gfx950.
//
// Semantics of gpu/mitm_gpu2.py:107-119 (join_sorted): for every I row, in I order, the run
// [lo, hi) of D rows with an equal key, emitted in D order. Pairs are counted per span of 512 I rows,
// the span counts scanned, and each span with pairs recomputes its rows' runs and writes them from
// its offset, so the output order is exactly the reference's (I-major, then D position) and no
// per-row array is stored.
#include <map>
#include <mutex>

#include "cf_internal.h"
#include "tree_dev.h"

namespace {
constexpr int BLK = 256;
constexpr int JW = 8;            // 64-key chunks per span: a span is 512 consecutive I rows, one wave's
constexpr u64 SPAN = 64 * JW;

__device__ __forceinline__ u64 lower_bound_in(const u64 *__restrict__ a, u64 lo, u64 hi, u64 v) {
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ u64 upper_bound_in(const u64 *__restrict__ a, u64 lo, u64 hi, u64 v) {
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// chunk_bounds: the D run of each key of one 64-row chunk (lane = row): lb = first D row with key >= k,
// cnt = rows with key == k. Called by whole waves. Ascending chunks (the sorted I side) place their
// keys in a window of 64 D keys held one per lane; a lane whose key lies beyond the window searches
// the rest of D itself. A chunk out of order falls back to a binary search per lane, so any I order
// gets exact bounds.
// The D window of a wave: the 64 D rows from L0 (fewer at the end of D) in lane order; ok once it
// is loaded. Wave-uniform except W.
struct Win {
    u64 L0, W, navail;
    bool ok;
};

__device__ __forceinline__ u64 rl64(u64 v, int l) {
    return ((u64)(u32)__builtin_amdgcn_readlane((int)(v >> 32), l) << 32) |
           (u64)(u32)__builtin_amdgcn_readlane((int)(u32)v, l);
}

// `lv`: the chunk's last row inside the I side (lanes past it repeat its key and count nothing).
// `from` (wave-uniform): a D row known to be at or below the first key's bound (the previous chunk's
// last bound when its last key is at most this chunk's first, else 0); `keep`: the window is known to
// start at or below that bound as well. The window is kept while it reaches the chunk's first key
// (a span of sorted I rows usually falls within one window), else reloaded from `from`, and when all
// of those rows lie below the chunk too, a 64-ary search over the rest finds the bound. A lane's
// bounds are then the window rows below (and at or below) its key: the rows below the chunk's first
// key count for every lane, and the few rows between its first and last key are read out of the
// window one by one (no lane shuffles).
__device__ __forceinline__ void chunk_bounds(const u64 *__restrict__ Dk, u64 nD, u64 k, int lv, u64 from, bool keep,
                                             Win &w, u64 &lb, u64 &cnt) {
    const int lane = threadIdx.x & 63;
    const bool valid = lane <= lv;
    const u64 k0 = rl64(k, 0), klast = rl64(k, lv);
    if (!valid) k = klast;
    const u64 prev = ((u64)wave_shift_up1((u32)(k >> 32)) << 32) | wave_shift_up1((u32)k);
    if (!__all(lane == 0 || prev <= k)) {
        lb = lower_bound_in(Dk, 0, nD, k);
        cnt = valid ? upper_bound_in(Dk, lb, nD, k) - lb : 0;
        w.ok = false;
        return;
    }
    if (!(keep && w.ok && (w.navail < 64 || __popcll(__ballot(w.W < k0)) < 64))) {
        // invariant of the search: D rows below lo hold keys < k0, rows at or above hi keys >= k0
        w.L0 = from;
        w.navail = nD - w.L0 < 64 ? nD - w.L0 : 64;
        w.W = (u64)lane < w.navail ? Dk[w.L0 + lane] : 0;
        if (w.navail == 64 && __popcll(__ballot(w.W < k0)) == 64) {
            u64 lo = w.L0 + 64, hi = nD;
            while (hi - lo > 64) {
                const u64 step = (hi - lo + 63) >> 6;
                const u64 p = lo + (u64)lane * step;
                const int c = __popcll(__ballot(p < hi && Dk[p] < k0));
                if (c == 0) {
                    hi = lo;
                    break;
                }
                const u64 last = lo + (u64)(c - 1) * step;   // the last probe below k0
                lo = last + 1;
                if (last + step < hi) hi = last + step;      // the first probe at or above k0
            }
            if (hi > lo) lo += __popcll(__ballot(lo + lane < hi && Dk[lo + lane] < k0));
            w.L0 = lo;
            w.navail = nD - w.L0 < 64 ? nD - w.L0 : 64;
            w.W = (u64)lane < w.navail ? Dk[w.L0 + lane] : 0;
        }
        w.ok = true;
    }
    const bool in = (u64)lane < w.navail;
    const u32 c0 = (u32)__popcll(__ballot(in && w.W < k0));
    const u32 e = (u32)__popcll(__ballot(in && w.W <= klast));
    u32 a = c0, b = c0;   // window rows below k, and at or below k
    for (u32 j = c0; j < e; ++j) {
        const u64 wj = rl64(w.W, (int)j);
        a += wj < k;
        b += wj <= k;
    }
    lb = a < 64 ? w.L0 + a : lower_bound_in(Dk, w.L0 + 64, nD, k);
    const u64 ub = b < 64 ? w.L0 + b : upper_bound_in(Dk, lb > w.L0 + 64 ? lb : w.L0 + 64, nD, k);
    cnt = valid ? ub - lb : 0;
}

// Pairs per span: span s covers I rows [s * SPAN, (s + 1) * SPAN); one wave per span.
__global__ void __launch_bounds__(BLK) k_join_count(const u64 *__restrict__ Dk, u64 nD, const u64 *__restrict__ Ik, u64 nI,
                                                    u64 nspans, u64 *__restrict__ scnt) {
    const int lane = threadIdx.x & 63;
    const u64 w0 = ((u64)blockIdx.x * BLK + threadIdx.x) >> 6, nw = ((u64)gridDim.x * BLK) >> 6;
    for (u64 sp = w0; sp < nspans; sp += nw) {
        u64 tot = 0, from = 0, last = 0;
        Win win{0, 0, 0, false};
        u64 kr[JW];   // the span's keys, loaded ahead of the searches
#pragma unroll
        for (int j = 0; j < JW; ++j) {
            const u64 i = sp * SPAN + (u64)j * 64 + lane;
            kr[j] = i < nI ? Ik[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < JW; ++j) {
            const u64 i0 = sp * SPAN + (u64)j * 64;
            if (i0 >= nI) break;
            const u64 i = i0 + lane;
            const u64 k = kr[j];
            u64 lb, c;
            const int lv = nI - i0 < 64 ? (int)(nI - i0) - 1 : 63;   // the chunk's last row
            const bool after = rl64(k, 0) >= last;
            chunk_bounds(Dk, nD, k, lv, after ? from : 0, after, win, lb, c);
            tot += c;
            from = rl64(lb, lv);
            last = rl64(k, lv);
        }
        tot = wave_incl_sum64(tot);
        if (lane == 63) scnt[sp] = tot;
    }
}

// Writes span s's pairs from soff[s] on, in I order and D order within a row; spans without pairs
// read nothing but their offsets. Writes nothing when the join's total (soff[nspans]) exceeds cap.
__global__ void __launch_bounds__(BLK) k_join_write(const u64 *__restrict__ Dk, u64 nD, const u64 *__restrict__ Ik,
                                                    const u64 *__restrict__ Did, const u64 *__restrict__ Iid, u64 nI,
                                                    u64 nspans, const u64 *__restrict__ soff, u64 *__restrict__ od,
                                                    u64 *__restrict__ oi, u64 *__restrict__ okey, u64 cap) {
    if (soff[nspans] > cap) return;
    const int lane = threadIdx.x & 63;
    const u64 w0 = ((u64)blockIdx.x * BLK + threadIdx.x) >> 6, nw = ((u64)gridDim.x * BLK) >> 6;
    for (u64 sp = w0; sp < nspans; sp += nw) {
        u64 base = soff[sp];
        if (soff[sp + 1] == base) continue;
        u64 from = 0, last = 0;
        Win win{0, 0, 0, false};
        for (int j = 0; j < JW; ++j) {
            const u64 i0 = sp * SPAN + (u64)j * 64;
            if (i0 >= nI) break;
            const u64 i = i0 + lane;
            const bool valid = i < nI;
            const u64 k = valid ? Ik[i] : 0;
            u64 lb, c;
            const int lv = nI - i0 < 64 ? (int)(nI - i0) - 1 : 63;
            const bool after = rl64(k, 0) >= last;
            chunk_bounds(Dk, nD, k, lv, after ? from : 0, after, win, lb, c);
            from = rl64(lb, lv);
            last = rl64(k, lv);
            const u64 incl = wave_incl_sum64(c);
            const u64 o = base + incl - c;
            if (c) {
                const u64 iid = Iid[i];
                for (u64 q = 0; q < c; ++q) {
                    od[o + q] = Did[lb + q];
                    oi[o + q] = iid;
                    if (okey) okey[o + q] = k;
                }
            }
            base += rl64(incl, 63);
        }
    }
}

inline unsigned grid_for_spans(u64 nspans) {
    u64 g = (nspans + BLK / 64 - 1) / (BLK / 64);
    if (g < 1) g = 1;
    return (unsigned)(g > (1u << 18) ? (1u << 18) : g);
}
}  // namespace

size_t join_spans(size_t nI) { return (nI + SPAN - 1) / SPAN; }

u64 join_sorted_count(cf_ctx *ctx, const u64 *Dk, size_t nD, const u64 *Ik, size_t nI, u64 *scnt, u64 *soff) {
    if (!nI || !nD) return 0;
    const u64 ns = join_spans(nI);
    CF_LAUNCH(ctx, "k_join_count", k_join_count, dim3(grid_for_spans(ns)), dim3(BLK), 0, Dk, (u64)nD, Ik, (u64)nI,
              ns, scnt);
    return scan_u64(ctx, scnt, soff, ns);
}

void join_sorted_write(cf_ctx *ctx, const u64 *Dk, size_t nD, const u64 *Ik, const u64 *Did, const u64 *Iid, size_t nI,
                       const u64 *soff, u64 *od, u64 *oi, u64 *okey) {
    if (!nI || !nD) return;
    const u64 ns = join_spans(nI);
    CF_LAUNCH(ctx, "k_join_write", k_join_write, dim3(grid_for_spans(ns)), dim3(BLK), 0, Dk, (u64)nD, Ik, Did, Iid,
              (u64)nI, ns, soff, od, oi, okey, ~0ULL);
}

// ---- the join without allocation ---------------------------------------------------------------
// cf_join_sorted_noalloc keeps its per-span pair counts and offsets, and the scan's block partials,
// in this file-scope array: one copy per device, resident while the library is loaded.
// In host code `scratch` names a host-side stand-in, not device memory: its device address comes
// from hipGetSymbolAddress, once per device (join_scratch_bind, called when a context is created),
// and only that address is passed to kernels.
namespace {
constexpr u64 NOALLOC_MAX_I = (u64)CF_JOIN_NOALLOC_MAX_I;
constexpr size_t scan_words_bound(size_t n) {   // scan_ws_words, evaluated at compile time
    return (n + 2047) / 2048 <= 1 ? 0 : 2 * ((n + 2047) / 2048) + 1 + scan_words_bound((n + 2047) / 2048);
}
constexpr size_t NOALLOC_SPANS = (NOALLOC_MAX_I + SPAN - 1) / SPAN;
constexpr size_t SCRATCH_WORDS = 2 * NOALLOC_SPANS + 1 + scan_words_bound(NOALLOC_SPANS);
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
    const u64 ns = join_spans(nI);
    u64 *scnt = ctx->join_scratch, *soff = scnt + NOALLOC_SPANS, *ws = soff + NOALLOC_SPANS + 1;
    CF_LAUNCH(ctx, "k_join_count", k_join_count, dim3(grid_for_spans(ns)), dim3(BLK), 0, Dk, (u64)nD, Ik, (u64)nI, ns,
              scnt);
    scan_u64_ws(ctx, scnt, soff, ns, ws);
    CF_LAUNCH(ctx, "k_join_write", k_join_write, dim3(grid_for_spans(ns)), dim3(BLK), 0, Dk, (u64)nD, Ik, Did, Iid,
              (u64)nI, ns, (const u64 *)soff, out_d, out_i, (u64 *)nullptr, (u64)cap);
    const u64 tot = d2h_u64(ctx, soff + ns);
    *npairs = tot;
    CF_REQUIRE(tot <= cap, "the output capacity is below the pair count (*npairs); nothing was written");
    API_END
}
