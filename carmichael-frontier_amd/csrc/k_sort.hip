// k_sort.hip -- exclusive scans and a stable LSD radix sort of (key, id) records on gfx950.
//
// Radix passes: 8-bit digits over the significant bits of the keys only. Each pass is
//   histogram (per 2048-key tile, LDS atomics)  ->  exclusive scan in digit-major order
//   ->  scatter, where each tile ranks its keys stably with 64-lane ballots (8 ballots give the
//   lanes holding the same digit), stages them digit-grouped in LDS, and writes contiguous runs.
// The sort takes the top bits off in one or two such passes (most significant first, the second
// within each of the first's buckets) and finishes every bucket in LDS; a plain LSD sort covers
// what that cannot take.
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

// ---- radix passes ---------------------------------------------------------------------------
// A pass splits its keys into tiles of at most RT keys; tile t's count of digit d sits at
// c0 + d * cs in the digit-major count array, so one exclusive scan of that array gives every
// (tile, digit) its output offset. Flat: tiles of the whole array. Segmented: tiles of each segment
// (a run of keys equal in their higher digits, as a previous pass left them), laid out segment by
// segment, so the scan keeps each segment inside its own range.
struct FlatGeo {
    size_t n, ntiles;
    __device__ bool tile(u64 t, u64 &base, int &count, u64 &c0, u64 &cs) const {
        base = t * RT;
        count = (int)((n - base) < (size_t)RT ? (n - base) : (size_t)RT);
        c0 = t;
        cs = ntiles;
        return true;
    }
};
struct SegGeo {
    const u64 *sstart, *send;
    const u32 *sfirst, *sn;   // first tile and tile count of each segment; sfirst[S] = all tiles
    int S;
    __device__ bool tile(u64 t, u64 &base, int &count, u64 &c0, u64 &cs) const {
        if (t >= sfirst[S]) {   // beyond the last tile: zero counts at c0 + d * cs
            c0 = t * RADIX;
            cs = 1;
            return false;
        }
        int lo = 0, hi = S - 1;  // the last segment whose first tile is at or below t (it has tiles)
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sfirst[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const u64 F = sfirst[lo], local = t - F;
        base = sstart[lo] + local * RT;
        const u64 left = send[lo] - base;
        count = (int)(left < (u64)RT ? left : (u64)RT);
        c0 = (u64)RADIX * F + local;
        cs = sn[lo];
        return true;
    }
};

template <class G>
__global__ void __launch_bounds__(RB) k_radix_hist(const u64 *__restrict__ keys, int shift, u32 *__restrict__ hist,
                                                   size_t ntiles, G g) {
    __shared__ u32 h[RADIX];
    // the scatter's XCD-aware order: neighbouring tiles' counts (adjacent words) meet in one L2
    const size_t per = (ntiles + 7) / 8;
    for (size_t b = blockIdx.x; b < 8 * per; b += gridDim.x) {
        const size_t tile = (b & 7) * per + (b >> 3);
        if (tile >= ntiles) continue;
        u64 base, c0, cs;
        int tcount = 0;
        const bool live = g.tile(tile, base, tcount, c0, cs);
        h[threadIdx.x] = 0;
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < RI; ++k) {
                const int le = k * RB + threadIdx.x;
                if (le < tcount) atomicAdd(&h[(keys[base + le] >> shift) & (RADIX - 1)], 1u);
            }
        }
        __syncthreads();
        hist[c0 + (u64)threadIdx.x * cs] = h[threadIdx.x];
        __syncthreads();
    }
}

template <class G>
__global__ void __launch_bounds__(RB) k_radix_scatter(const u64 *__restrict__ kin, const u64 *__restrict__ vin,
                                                      int shift, const u64 *__restrict__ off, size_t ntiles, G g,
                                                      u64 *__restrict__ kout, u64 *__restrict__ vout) {
    __shared__ u32 s_cnt[RADIX];
    __shared__ u32 s_wc[RB / 64][RADIX];
    __shared__ u32 s_loc[RADIX];
    __shared__ u64 s_goff[RADIX];
    __shared__ u64 s_k[RT];
    __shared__ u64 s_v[RT];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const u64 lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
    // XCD-aware order: blocks are dealt to the 8 dies round-robin, so die x takes the tiles
    // [x * per, (x + 1) * per) in order, and a digit's runs from neighbouring tiles (adjacent in the
    // output) meet in that die's L2 instead of reaching memory as partial lines from two dies
    const size_t per = (ntiles + 7) / 8;
    for (size_t b = blockIdx.x; b < 8 * per; b += gridDim.x) {
        const size_t tile = (b & 7) * per + (b >> 3);
        u64 base, c0, cs;
        int tcount = 0;
        if (tile >= ntiles || !g.tile(tile, base, tcount, c0, cs)) continue;   // uniform over the block
        s_goff[threadIdx.x] = off[c0 + (u64)threadIdx.x * cs];
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

// The segments of the next pass: segment b = the keys whose previous-pass digit is b (offA in the
// flat pass's layout, ntA tiles), or the whole array when offA is null (S = 1).
__global__ void __launch_bounds__(SB) k_seg_table(const u64 *__restrict__ offA, size_t ntA, int S, size_t n,
                                                  u64 *__restrict__ sstart, u64 *__restrict__ send,
                                                  u32 *__restrict__ sfirst, u32 *__restrict__ sn) {
    const int b = threadIdx.x;
    u64 st = 0, en = 0;
    if (b < S) {
        st = offA ? offA[(size_t)b * ntA] : 0;
        en = offA ? offA[(size_t)(b + 1) * ntA] : n;
    }
    const u64 nt = (en - st + RT - 1) / RT;
    u64 tot;
    const u64 f = block_excl_scan(nt, &tot);
    if (b < S) {
        sstart[b] = st;
        send[b] = en;
        sfirst[b] = (u32)f;
        sn[b] = (u32)nt;
    }
    if (b == 0) sfirst[S] = (u32)tot;
}

// ---- per-bucket sort in LDS ------------------------------------------------------------------
// After the MSD passes, bucket q holds the keys whose top m bits equal q, in input order. One
// block sorts a bucket of at most BCAP keys by its remaining `lowbits` bits. Each key is packed
// with its position in the bucket as (low bits << BIB) | position: the packed words are distinct
// and their order is the stable order, so any exact sort of them is the stable sort; the position
// carries each id to its key's place.
//   k_lds_sort: the top GB bits of the low bits pick a group; LDS atomics give each word a
//   slot, a scan of the group counts places the groups, and each word's rank inside its group is
//   the number of smaller words there (groups hold a few words when the keys spread). A bucket with
//   a group above GMAX words (keys bunched on a few values) is marked and left to
//   k_lds_sort_lsd: 8-bit LSD passes in LDS, where wave w owns positions [w * R * 64, (w + 1) * R * 64),
//   its lanes rank equal digits with ballots and keep a running count per digit across their
//   rounds, and one block scan per pass orders the waves.
// A bucket above BCAP sets bit 0 of *flag and is copied through unsorted; a marked bucket sets bit 1.
constexpr int BT = 512;
constexpr int BW = BT / 64;
constexpr int BI = 16;
constexpr int BCAP = BT * BI;   // 8192
constexpr int BIB = 13;         // 2^BIB = BCAP
constexpr int GB = 11;          // group bits of the counting pass
constexpr int NG = 1 << GB;
constexpr u32 GMAX = 64;
static_assert((1 << BIB) == BCAP, "positions fill BIB bits");
static_assert(NG % BT == 0, "each thread scans NG / BT group counts");

// Bucket q's range [start, start + len) in the MSD output (the whole array when off is null).
__device__ __forceinline__ void bucket_range(u32 q, const u64 *__restrict__ off, const u32 *__restrict__ sfirst,
                                             const u32 *__restrict__ sn, u64 n, int m2, u64 &start, u64 &len) {
    start = 0;
    len = n;
    if (!off) return;
    const u32 b1 = q >> m2, d2 = q & ((1u << m2) - 1);
    const u64 F = sfirst[b1], N = sn[b1];
    if (N == 0) {
        len = 0;
        return;
    }
    start = off[(u64)RADIX * F + (u64)d2 * N];
    len = off[(u64)RADIX * F + (u64)(d2 + 1) * N] - start;
}

__global__ void __launch_bounds__(SB) k_msd_bucket_max(const u64 *__restrict__ off, const u32 *__restrict__ sfirst,
                                                   const u32 *__restrict__ sn, u32 nb, int m2,
                                                   unsigned long long *__restrict__ out) {
    const u32 q = blockIdx.x * SB + threadIdx.x;
    u64 start = 0, len = 0;
    if (q < nb) bucket_range(q, off, sfirst, sn, 0, m2, start, len);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const u64 o = __shfl_xor(len, d, 64);
        len = o > len ? o : len;
    }
    if ((threadIdx.x & 63) == 0 && len) atomicMax(out, (unsigned long long)len);
}

__global__ void __launch_bounds__(BT) k_lds_sort(const u64 *kin, const u64 *vin, u64 *kout, u64 *vout,
                                                     const u64 *__restrict__ off, const u32 *__restrict__ sfirst,
                                                     const u32 *__restrict__ sn, u64 n, int m2, int lowbits,
                                                     unsigned long long *__restrict__ flag, u32 *__restrict__ mark) {
    __shared__ u64 s_v[BCAP];
    __shared__ u32 s_c[NG + 1];   // group counts, then group starts
    __shared__ u32 s_ws[BW];
    __shared__ u32 s_big;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const u32 q = blockIdx.x;
    u64 start, len;
    bucket_range(q, off, sfirst, sn, n, m2, start, len);
    if (len == 0) return;
    if (len > (u64)BCAP) {
        if (tid == 0) atomicOr(flag, 1ULL);
        if (kin != kout)
            for (u64 e = tid; e < len; e += BT) {
                kout[start + e] = kin[start + e];
                vout[start + e] = vin[start + e];
            }
        return;
    }
    // 32-bit positions against bucket-based pointers
    const u32 ln = (u32)len;
    const u64 *kb = kin + start, *vb = vin + start;
    u64 *kob = kout + start, *vob = vout + start;
    const int R = (int)((ln + BT - 1) / BT);
    const u32 wbase = (u32)(w * R * 64);
    const u64 lowmask = lowbits ? (~0ULL >> (64 - lowbits)) : 0;
    const int gb = lowbits < GB ? lowbits : GB;
    const int gsh = BIB + lowbits - gb;
    const u32 gmask = (1u << gb) - 1;
    for (int x = tid; x <= NG; x += BT) s_c[x] = 0;
    if (tid == 0) s_big = 0;
    __syncthreads();
    u64 v[BI], idr[BI];
    u32 slot[BI];
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        const u32 e = wbase + (u32)(it * 64 + lane);
        if (it < R && e < ln) {
            v[it] = ((kb[e] & lowmask) << BIB) | e;
            idr[it] = vb[e];
            slot[it] = atomicAdd(&s_c[(u32)(v[it] >> gsh) & gmask], 1u);
        }
    }
    __syncthreads();
    {   // group starts: thread t scans groups [t * NG / BT, (t + 1) * NG / BT)
        constexpr int PER = NG / BT;
        u32 c[PER], sum = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            c[j] = s_c[tid * PER + j];
            sum += c[j];
        }
        const u32 inc = wave_incl_sum(sum);
        if (lane == 63) s_ws[w] = inc;
        __syncthreads();
        u32 pre = inc - sum;
        for (int x = 0; x < w; ++x) pre += s_ws[x];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            s_c[tid * PER + j] = pre;
            pre += c[j];
        }
        if (tid == BT - 1) s_c[NG] = pre;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BI; ++it)
        if (it < R && wbase + (u32)(it * 64 + lane) < ln) s_v[s_c[(u32)(v[it] >> gsh) & gmask] + slot[it]] = v[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        if (it < R && wbase + (u32)(it * 64 + lane) < ln) {
            const u32 g = (u32)(v[it] >> gsh) & gmask, g0 = s_c[g], g1 = s_c[g + 1];
            if (g1 - g0 > GMAX) {
                s_big = 1;
            } else {
                u32 rank = 0;
                for (u32 j = g0; j < g1; ++j) rank += s_v[j] < v[it];
                slot[it] = g0 + rank;
            }
        }
    }
    __syncthreads();
    if (s_big) {
        if (tid == 0) {
            mark[q] = 1;
            atomicOr(flag, 2ULL);
        }
        return;
    }
    // keys, then ids, each staged at its final place and written in order (reads of the bucket are
    // all done: kin and kout, vin and vout may be one array)
#pragma unroll
    for (int it = 0; it < BI; ++it)
        if (it < R && wbase + (u32)(it * 64 + lane) < ln) s_v[slot[it]] = v[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        const u32 e = wbase + (u32)(it * 64 + lane);
        if (it < R && e < ln) kob[e] = ((u64)q << lowbits) | (s_v[e] >> BIB);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BI; ++it)
        if (it < R && wbase + (u32)(it * 64 + lane) < ln) s_v[slot[it]] = idr[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        const u32 e = wbase + (u32)(it * 64 + lane);
        if (it < R && e < ln) vob[e] = s_v[e];
    }
}

__global__ void __launch_bounds__(BT) k_lds_sort_lsd(const u64 *kin, const u64 *vin, u64 *kout, u64 *vout,
                                                   const u64 *__restrict__ off, const u32 *__restrict__ sfirst,
                                                   const u32 *__restrict__ sn, u64 n, int m2, int lowbits,
                                                   const u32 *__restrict__ mark) {
    __shared__ u64 s_v[BCAP];
    __shared__ u32 s_wc[BW][RADIX];
    __shared__ u32 s_ws[BW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const u32 q = blockIdx.x;
    if (!mark[q]) return;
    u64 start, len;
    bucket_range(q, off, sfirst, sn, n, m2, start, len);
    const u32 ln = (u32)len;   // marked buckets hold at most BCAP keys
    const u64 *kb = kin + start, *vb = vin + start;
    u64 *kob = kout + start, *vob = vout + start;
    const int R = (int)((ln + BT - 1) / BT);
    const u32 wbase = (u32)(w * R * 64);
    const u64 lowmask = lowbits ? (~0ULL >> (64 - lowbits)) : 0;
    const u64 lt = lane ? (~0ULL >> (64 - lane)) : 0ULL;
    u64 v[BI];
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        const u32 e = wbase + (u32)(it * 64 + lane);
        v[it] = (it < R && e < ln) ? (((kb[e] & lowmask) << BIB) | e) : 0;
    }
    const int passes = (lowbits + 7) / 8;
    for (int ps = 0; ps < passes; ++ps) {
        const int sh = BIB + 8 * ps;
        for (int x = tid; x < BW * RADIX; x += BT) (&s_wc[0][0])[x] = 0;
        __syncthreads();
        u32 rk[BI];
#pragma unroll
        for (int it = 0; it < BI; ++it) {
            rk[it] = 0;
            if (it < R) {
                const bool valid = wbase + (u32)(it * 64 + lane) < ln;
                const u32 d = (u32)(v[it] >> sh) & (RADIX - 1);
                u64 peers = __ballot(valid);
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const bool bit = (d >> b) & 1u;
                    const u64 bb = __ballot(bit);
                    peers &= bit ? bb : ~bb;
                }
                const u32 rw = (u32)__popcll(peers & lt), cw = (u32)__popcll(peers);
                // every lane of the wave reads the digit's running count before its last holder adds
                // the round's run: LDS operations of one wave complete in issue order
                const u32 run = valid ? s_wc[w][d] : 0;
                rk[it] = run + rw;
                __builtin_amdgcn_wave_barrier();
                if (valid && rw == cw - 1) s_wc[w][d] = run + cw;
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        // digit d: the waves' counts become their offsets within the digit, then the digit's base
        u32 tot = 0;
        if (tid < RADIX) {
#pragma unroll
            for (int x = 0; x < BW; ++x) {
                const u32 c = s_wc[x][tid];
                s_wc[x][tid] = tot;
                tot += c;
            }
        }
        const u32 inc = wave_incl_sum(tot);
        if (lane == 63) s_ws[w] = inc;
        __syncthreads();
        if (tid < RADIX) {
            u32 base = inc - tot;
            for (int x = 0; x < w; ++x) base += s_ws[x];
#pragma unroll
            for (int x = 0; x < BW; ++x) s_wc[x][tid] += base;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < BI; ++it) {
            if (it < R && wbase + (u32)(it * 64 + lane) < ln) {
                const u32 d = (u32)(v[it] >> sh) & (RADIX - 1);
                s_v[s_wc[w][d] + rk[it]] = v[it];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < BI; ++it) {
            const u32 e = wbase + (u32)(it * 64 + lane);
            if (it < R && e < ln) v[it] = s_v[e];
        }
        __syncthreads();
    }
    // with no pass every key of the bucket is equal and the input order is the order
    u64 idv[BI];
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        const u32 e = wbase + (u32)(it * 64 + lane);
        idv[it] = 0;
        if (it < R && e < ln) {
            idv[it] = vb[(u32)v[it] & (BCAP - 1)];
            v[it] = ((u64)q << lowbits) | (v[it] >> BIB);
        }
    }
    __syncthreads();   // every id is read before any is written (vin and vout may be one array)
#pragma unroll
    for (int it = 0; it < BI; ++it) {
        const u32 e = wbase + (u32)(it * 64 + lane);
        if (it < R && e < ln) {
            kob[e] = v[it];
            vob[e] = idv[it];
        }
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

namespace {
// Blocks of a pass over ntiles tiles: a multiple of 8 (one tile per block up to 65536 blocks).
unsigned scatter_grid(size_t ntiles) {
    const size_t g = 8 * ((ntiles + 7) / 8);
    return (unsigned)(g < 65536 ? g : 65536);
}

// Stable LSD radix sort in place: one flat pass per 8 bits of `bits`.
void lsd_sort_pairs(cf_ctx *ctx, u64 *keys, u64 *vals, size_t n, int bits) {
    if (n <= 1 || bits <= 0) return;
    const int passes = (bits + 7) / 8;
    const size_t ntiles = (n + RT - 1) / RT;
    DBuf<u64> k2(ctx, n), v2(ctx, n);
    DBuf<u32> hist(ctx, ntiles * RADIX);
    DBuf<u64> off(ctx, ntiles * RADIX + 1);
    u64 *ka = keys, *va = vals, *kb = k2.p, *vb = v2.p;
    const unsigned grid = scatter_grid(ntiles);
    const FlatGeo g{n, ntiles};
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = ps * 8;
        CF_LAUNCH(ctx, "k_radix_hist", k_radix_hist<FlatGeo>, dim3(grid), dim3(RB), 0, ka, shift, hist.p, ntiles, g);
        scan_impl<u32>(ctx, hist.p, off.p, ntiles * RADIX);
        CF_LAUNCH(ctx, "k_radix_scatter", k_radix_scatter<FlatGeo>, dim3(grid), dim3(RB), 0, ka, va, shift, off.p,
                  ntiles, g, kb, vb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        CF_HIP(hipMemcpyAsync(keys, ka, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(vals, va, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
    }
}

constexpr size_t MSD_MEAN = 4096;   // mean bucket the MSD bits aim at (half of BCAP)

// The largest of the 2^m buckets a segmented pass's offsets describe.
u64 max_bucket(cf_ctx *ctx, const u64 *off, const u32 *sfirst, const u32 *sn, int m, int m2) {
    DBuf<u64> mx(ctx, 1);
    CF_HIP(hipMemsetAsync(mx.p, 0, sizeof(u64), ctx->stream));
    const u32 nb = 1u << m;
    CF_LAUNCH(ctx, "k_msd_bucket_max", k_msd_bucket_max, dim3((nb + SB - 1) / SB), dim3(SB), 0, off, sfirst, sn, nb, m2,
              (unsigned long long *)mx.p);
    return d2h_u64(ctx, mx.p);
}

// The 2^m buckets sorted from (kin, vin) into (kout, vout); *flag gets bit 0 for a bucket above
// BCAP (copied through). Buckets the counting pass marks are finished by the LSD kernel here.
// Returns the flag word (one read-back).
u64 bucket_sort(cf_ctx *ctx, const u64 *kin, const u64 *vin, u64 *kout, u64 *vout, const u64 *off, const u32 *sfirst,
                 const u32 *sn, size_t n, int m, int m2, int lowbits, u64 *flag) {
    const unsigned nb = 1u << m;
    DBuf<u32> mark(ctx, nb);
    CF_HIP(hipMemsetAsync(mark.p, 0, nb * sizeof(u32), ctx->stream));
    CF_LAUNCH(ctx, "k_lds_sort", k_lds_sort, dim3(nb), dim3(BT), 0, kin, vin, kout, vout, off, sfirst, sn, (u64)n,
              m2, lowbits, (unsigned long long *)flag, mark.p);
    const u64 f = d2h_u64(ctx, flag);
    if (f & 2)
        CF_LAUNCH(ctx, "k_lds_sort_lsd", k_lds_sort_lsd, dim3(nb), dim3(BT), 0, kin, vin, kout, vout, off, sfirst, sn,
                  (u64)n, m2, lowbits, (const u32 *)mark.p);
    return f;
}
constexpr int MSD_MAX_BITS = 16;    // two passes of at most 8 bits
}  // namespace

// Stable sort of (kin, vin) by key into (kout, vout) (which may be the inputs themselves), for keys
// below 2^bits. The top m bits (mean bucket at most MSD_MEAN) are split off by one or two stable
// MSD passes, the second segmented by the first's digit, and each bucket is then sorted in LDS by
// k_lds_sort: three passes over HBM. Stability makes the result the same as the LSD sort's, which
// takes over when the MSD bits would exceed two passes or a bucket overflows the LDS.
void sort_pairs_into(cf_ctx *ctx, const u64 *kin, const u64 *vin, u64 *kout, u64 *vout, size_t n, int bits) {
    const bool inplace = kin == kout;
    CF_REQUIRE(inplace == (vin == vout), "keys and ids are both sorted in place or both into new arrays");
    if (!inplace && n) {
        if (n == 1 || bits <= 0) {
            CF_HIP(hipMemcpyAsync(kout, kin, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
            CF_HIP(hipMemcpyAsync(vout, vin, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
            return;
        }
    }
    if (n <= 1 || bits <= 0) return;
    int m = 0;
    // below 2^24 keys a bucket costs little, so the first try aims lower and leaves room for keys
    // bunched up to 8x their mean (the D sides of the divisor families bunch about 4x)
    const size_t mean = n < ((size_t)1 << 24) ? MSD_MEAN / 4 : MSD_MEAN;
    while ((n >> m) > mean) ++m;
    if (m > MSD_MAX_BITS && (n >> MSD_MAX_BITS) <= MSD_MEAN + MSD_MEAN / 2) m = MSD_MAX_BITS;
    m = std::max(m, bits + BIB - 64);   // the low bits and a position fit one word
    m = std::min(m, bits);
    if (m > MSD_MAX_BITS) {
        if (!inplace) {
            CF_HIP(hipMemcpyAsync(kout, kin, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
            CF_HIP(hipMemcpyAsync(vout, vin, n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
        }
        lsd_sort_pairs(ctx, kout, vout, n, bits);
        return;
    }
    DBuf<u64> flag(ctx, 1);
    CF_HIP(hipMemsetAsync(flag.p, 0, sizeof(u64), ctx->stream));
    if (m == 0) {   // one bucket: no more keys than the target mean
        if (bucket_sort(ctx, kin, vin, kout, vout, nullptr, nullptr, nullptr, n, 0, 0, bits, flag.p) & 1)
            lsd_sort_pairs(ctx, kout, vout, n, bits);
        return;
    }
    const int mcap = std::min(MSD_MAX_BITS, bits);
    bool overflow = false;
    for (int attempt = 0;; ++attempt) {
        const int m2 = std::min(m, 8), m1 = m - m2, S = 1 << m1;
        const int lowbits = bits - m;
        DBuf<u64> sstart(ctx, S), send(ctx, S);
        DBuf<u32> sfirst(ctx, S + 1), sn(ctx, S);
        DBuf<u64> tk, tv;
        if (m1 > 0 || inplace) {
            tk = DBuf<u64>(ctx, n);
            tv = DBuf<u64>(ctx, n);
        }
        const u64 *ak = kin, *av = vin;   // pass B's input
        DBuf<u64> offA;
        size_t ntA = 0;
        if (m1 > 0) {   // pass A: flat, top m1 bits, into the scratch pair
            ntA = (n + RT - 1) / RT;
            DBuf<u32> histA(ctx, ntA * RADIX);
            offA = DBuf<u64>(ctx, ntA * RADIX + 1);
            const unsigned grid = scatter_grid(ntA);
            const FlatGeo g{n, ntA};
            CF_LAUNCH(ctx, "k_radix_hist", k_radix_hist<FlatGeo>, dim3(grid), dim3(RB), 0, kin, bits - m1, histA.p,
                      ntA, g);
            scan_impl<u32>(ctx, histA.p, offA.p, ntA * RADIX);
            CF_LAUNCH(ctx, "k_radix_scatter", k_radix_scatter<FlatGeo>, dim3(grid), dim3(RB), 0, kin, vin, bits - m1,
                      (const u64 *)offA.p, ntA, g, tk.p, tv.p);
            ak = tk.p;
            av = tv.p;
        }
        CF_LAUNCH(ctx, "k_seg_table", k_seg_table, dim3(1), dim3(SB), 0, (const u64 *)(m1 > 0 ? offA.p : nullptr), ntA,
                  S, n, sstart.p, send.p, sfirst.p, sn.p);
        // pass B: segmented, the next m2 bits; into the outputs, or the scratch pair when sorting in
        // place with one MSD pass (pass B cannot write over its own input)
        const size_t ntB = (n + RT - 1) / RT + (size_t)S;
        DBuf<u32> histB(ctx, ntB * RADIX);
        DBuf<u64> offB(ctx, ntB * RADIX + 1);
        const unsigned grid = scatter_grid(ntB);
        const SegGeo g{sstart.p, send.p, sfirst.p, sn.p, S};
        CF_LAUNCH(ctx, "k_radix_hist", k_radix_hist<SegGeo>, dim3(grid), dim3(RB), 0, ak, lowbits, histB.p, ntB, g);
        scan_impl<u32>(ctx, histB.p, offB.p, ntB * RADIX);
        if (m < mcap && attempt == 0) {
            // keys bunched at this depth (a bucket past the LDS): take more top bits and start again
            const u64 big = max_bucket(ctx, offB.p, sfirst.p, sn.p, m, m2);
            if (big > (u64)BCAP) {
                int grow = 1;
                while ((big >> grow) > MSD_MEAN) ++grow;
                m = std::min(mcap, m + grow);
                continue;
            }
        }
        u64 *ok = kout, *ov = vout;
        if (m1 == 0 && inplace) {
            ok = tk.p;
            ov = tv.p;
        }
        CF_LAUNCH(ctx, "k_radix_scatter", k_radix_scatter<SegGeo>, dim3(grid), dim3(RB), 0, ak, av, lowbits,
                  (const u64 *)offB.p, ntB, g, ok, ov);
        overflow = bucket_sort(ctx, ok, ov, kout, vout, offB.p, sfirst.p, sn.p, n, m, m2, lowbits, flag.p) & 1;
        break;
    }
    if (overflow) lsd_sort_pairs(ctx, kout, vout, n, bits);
}

void radix_sort_pairs(cf_ctx *ctx, u64 *keys, u64 *vals, size_t n, int bits) {
    sort_pairs_into(ctx, keys, vals, keys, vals, n, bits);
}
