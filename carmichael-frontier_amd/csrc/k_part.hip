// k_part.hip -- radix-partitioned residue join on gfx950.
//
// Keys of one solve live in one key space (instance j owns [koff_j, koff_j + M_j), so keys of
// different instances never collide). The space is cut into NB = 2^lgNB (about a third of the D
// records) fine buckets by multiplicative hashing (the top bits of key * 0x9E3779B97F4A7C15).
// Hashing, not key order, because keys are far from uniform: for small r the products prod(D)
// stay below M and are their own residues, all in a sliver of [0, M), and a batch mixes moduli
// from 1e11 to 1e17. The top lgP bits of the hash give P <= 256 coarse partitions.
//
// Probe table: per bucket 8 fingerprint slots of 16 bits (16 bytes) for its first 8 D records; a
// probe reads its bucket's 16 bytes and touches the D records only on a fingerprint hit (then the
// exact key is recomputed and checked) or when the bucket holds more than 8 records.
//
//  D side: records are counted per fine bucket and scattered into bucket order (a counting sort
//          on b: no full key sort is needed, a probe scans its bucket, 2-4 records on average).
//  I side: leaves are scattered by partition as 8-byte records (bucket offset, fingerprint, leaf
//          index): each block ranks its 2048 leaves by partition in LDS and writes one contiguous
//          run per partition into slots sized from the expected share.
//  Probe:  the partitioned I records are read once, in partition order, with an XCD-aware tile
//          order (each of the 8 XCDs walks its own contiguous slice of tiles) so the D buckets
//          of the partitions in flight stay in that XCD's L2. Every equal key emits a pair.
// HBM traffic per I record: 8 B written by the scatter + 8 B read by the probe (plus the 16-byte
// table bucket it looks up).
// Join semantics: every equal-key (D, I) pair (gpu/mitm_gpu2.py:107-119).
#include "tree_dev.h"

namespace {

constexpr int PB = 256;            // threads per block
constexpr int LPT = 8;             // consecutive leaves per thread in the leaf kernels
constexpr int LT = PB * LPT;       // 2048 leaves per tile
constexpr int PI = 16;             // records per thread in the probe
constexpr int PT = PB * PI;        // 4096 records per probe tile
constexpr int NXCD = 8;            // MI355X: 8 accelerator complex dies, blocks dealt round-robin

struct LeafSrc {
    const SideDesc *descs;
    const Node *nodes;
    size_t nn;
    const u64 *off;   // exclusive scan of leaves per node, nn + 1 entries
    u64 lo, n;        // leaves [lo, lo + n)
    int depth;        // r - 1
    const u64 *tp;    // per tile of LT leaves: owners of its first and last leaf (2 entries per tile)
};

// Hashing of a key: h = key * 0x9E3779B97F4A7C15 (odd multiplier, so a bijection on u64).
//   fine bucket    b = top lgNB bits of h          (lgNB <= 28: bits 36..63)
//   partition      p = top lgP bits of h           (the first lgP bits of b)
//   fingerprint    f = bits 20..31 of h            (independent of b)
// An I record is 8 bytes: low word (b mod 2^(lgNB - lgP)) << 12 | f, high word the leaf's index
// in its chunk; the exact key and id of a leaf are recomputed from the resident node tree only
// when its fingerprint hits the D table. Table slots hold 0x8000 | f (0 = empty).
struct KeyHash {
    int lgNB, lgP;
};
__device__ __forceinline__ u64 khash(u64 key) { return key * 0x9E3779B97F4A7C15ULL; }
__device__ __forceinline__ u64 hbucket(u64 h, int lgNB) { return h >> (64 - lgNB); }
__device__ __forceinline__ u32 hfp(u64 h) { return (u32)(h >> 20) & 0xFFFu; }
constexpr int TSLOTS = 8;

// Tile set-up: the nodes owning leaves [j0, j1) are p0 .. p0 + np - 1 (every last-level node owns at
// least one leaf, so np <= LT); their leaf offsets go to LDS: s_off[q] = off[p0 + q], q <= np.
__device__ __forceinline__ void tile_setup(const LeafSrc &L, size_t t, u64 *s_off, size_t &p0, u32 &np) {
    p0 = L.tp[2 * t];
    np = (u32)(L.tp[2 * t + 1] - p0 + 1);
    const u32 nc = np <= (u32)LT ? np : (u32)LT;
    for (u32 i = threadIdx.x; i <= nc; i += PB) s_off[i] = L.off[p0 + i];
    __syncthreads();
}

// The leaves [jt, jt + LPT) of one thread, walked in order: the owning node is found once (binary
// search of the tile's offsets in LDS), then advanced as the leaf index crosses node boundaries.
struct LeafWalk {
    const LeafSrc &L;
    const u64 *s_off;
    size_t p0;
    u32 np, q;
    u64 lo_off, hi_off;
    Node nd;
    const SideDesc *s;

    __device__ __forceinline__ LeafWalk(const LeafSrc &L_, const u64 *s_off_, size_t p0_, u32 np_, u64 jt)
        : L(L_), s_off(s_off_), p0(p0_), np(np_) {
        u32 a = 0, b = np - 1;
        while (a < b) {
            const u32 mid = (a + b + 1) >> 1;
            if (off_at(mid) <= jt) a = mid; else b = mid - 1;
        }
        q = a;
        load();
    }
    __device__ __forceinline__ u64 off_at(u32 i) const { return np <= (u32)LT ? s_off[i] : L.off[p0 + i]; }
    __device__ __forceinline__ void load() {
        lo_off = off_at(q);
        hi_off = off_at(q + 1);
        nd = L.nodes[p0 + q];
        s = &L.descs[nd.meta >> 8];
    }
    // key of leaf j (j >= every earlier j of this walk)
    __device__ __forceinline__ void eval_key(u64 j, u64 &key) {
        while (j >= hi_off) {
            ++q;
            load();
        }
        const int x = child_index(*s, nd, (int)(j - lo_off));
        const u64 xm = cf_mulmod_m(nd.xm, s->pR[x], s->mc);
        const u64 x2 = nd.x2 * s->p[x];
        key = cf_key(xm, x2, s->mc) + s->koff;
    }
    // key, id and instance of leaf j (j >= every earlier j of this walk)
    __device__ __forceinline__ void eval(u64 j, u64 &key, u64 &id, u32 &inst) {
        while (j >= hi_off) {
            ++q;
            load();
        }
        inst = nd.meta >> 8;
        const int x = child_index(*s, nd, (int)(j - lo_off));
        const u64 xm = cf_mulmod_m(nd.xm, s->pR[x], s->mc);
        const u64 x2 = nd.x2 * s->p[x];
        key = cf_key(xm, x2, s->mc) + s->koff;
        id = id_update(*s, nd.id, x, L.depth);
    }
};

// Owners of the first and the last leaf of every LT-leaf tile of [lo, lo + n): one binary search
// per lane, all tiles at once, so no leaf kernel starts a tile with a serial search.
__global__ void __launch_bounds__(PB) k_tile_parents(const u64 *__restrict__ off, size_t nn, u64 lo, u64 n,
                                                     u64 *__restrict__ tp) {
    const u64 ntiles = (n + LT - 1) / LT;
    for (u64 i = (u64)blockIdx.x * PB + threadIdx.x; i < 2 * ntiles; i += (u64)gridDim.x * PB) {
        const u64 t = i >> 1;
        const u64 j0 = lo + t * LT;
        const u64 j = (i & 1) ? (j0 + LT < lo + n ? j0 + LT : lo + n) - 1 : j0;
        tp[i] = find_parent(off, 0, nn - 1, j);
    }
}

// ---- D side: counting sort by fine bucket ---------------------------------------------------
__global__ void __launch_bounds__(PB) k_bucket_count(LeafSrc L, int lgNB, u32 *__restrict__ bcount) {
    __shared__ u64 s_off[LT + 1];
    const size_t ntiles = (L.n + LT - 1) / LT;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * LT, jend = L.lo + L.n;
        const u64 j1 = j0 + LT < jend ? j0 + LT : jend;
        size_t p0;
        u32 np;
        tile_setup(L, t, s_off, p0, np);
        const u64 jt = j0 + (u64)threadIdx.x * LPT;
        if (jt < j1) {
            LeafWalk w(L, s_off, p0, np, jt);
            for (int k = 0; k < LPT && jt + k < j1; ++k) {
                u64 key, id;
                u32 inst;
                w.eval(jt + k, key, id, inst);
                atomicAdd(&bcount[hbucket(khash(key), lgNB)], 1u);
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(PB) k_bucket_scatter(LeafSrc L, int lgNB, const u64 *__restrict__ bstart,
                                                       u32 *__restrict__ bcur, u64 *__restrict__ Dk,
                                                       u64 *__restrict__ Did) {
    __shared__ u64 s_off[LT + 1];
    const size_t ntiles = (L.n + LT - 1) / LT;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * LT, jend = L.lo + L.n;
        const u64 j1 = j0 + LT < jend ? j0 + LT : jend;
        size_t p0;
        u32 np;
        tile_setup(L, t, s_off, p0, np);
        const u64 jt = j0 + (u64)threadIdx.x * LPT;
        if (jt < j1) {
            LeafWalk w(L, s_off, p0, np, jt);
            for (int k = 0; k < LPT && jt + k < j1; ++k) {
                u64 key, id;
                u32 inst;
                w.eval(jt + k, key, id, inst);
                const u64 b = hbucket(khash(key), lgNB);
                const u64 at = bstart[b] + atomicAdd(&bcur[b], 1u);
                Dk[at] = key;
                Did[at] = id;
            }
        }
        __syncthreads();
    }
}

// ---- I side: partition by the top bits of the key hash --------------------------------------
__global__ void __launch_bounds__(PB) k_part_count(LeafSrc L, int lgP, u32 *__restrict__ pcount) {
    __shared__ u64 s_off[LT + 1];
    __shared__ u32 h[256];
    const int P = 1 << lgP;
    const size_t ntiles = (L.n + LT - 1) / LT;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * LT, jend = L.lo + L.n;
        const u64 j1 = j0 + LT < jend ? j0 + LT : jend;
        h[threadIdx.x] = 0;
        size_t p0;
        u32 np;
        tile_setup(L, t, s_off, p0, np);
        const u64 jt = j0 + (u64)threadIdx.x * LPT;
        if (jt < j1) {
            LeafWalk w(L, s_off, p0, np, jt);
            for (int k = 0; k < LPT && jt + k < j1; ++k) {
                u64 key, id;
                u32 inst;
                w.eval(jt + k, key, id, inst);
                atomicAdd(&h[(u32)(khash(key) >> (64 - lgP))], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < P && h[threadIdx.x]) atomicAdd(&pcount[threadIdx.x], h[threadIdx.x]);
        __syncthreads();
    }
}

// Partition p owns record slots [pbase[p], pbase[p + 1]); pcur[p] counts the slots taken. With a
// capacity layout a partition that would run past its end sets *ovf and writes nothing for that
// tile; the caller then redoes the chunk with exact counts as the layout.
__global__ void __launch_bounds__(PB) k_part_scatter(LeafSrc L, KeyHash kh, const u64 *__restrict__ pbase,
                                                     u32 *__restrict__ pcur, u32 *__restrict__ ovf,
                                                     u64 *__restrict__ rec) {
    __shared__ u64 s_off[LT + 1];
    __shared__ u32 h[256];
    __shared__ u32 loc[256];
    __shared__ u64 gbase[256];
    __shared__ u32 s_w[PB / 64];
    __shared__ u64 s_r[LT];
    __shared__ u8_t s_p[LT];
    constexpr u64 NONE = ~0ULL;
    const int P = 1 << kh.lgP;
    const u64 bomask = (1ULL << (kh.lgNB - kh.lgP)) - 1;
    const size_t ntiles = (L.n + LT - 1) / LT;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * LT, jend = L.lo + L.n;
        const u64 j1 = j0 + LT < jend ? j0 + LT : jend;
        const int cnt = (int)(j1 - j0);
        h[threadIdx.x] = 0;
        size_t p0;
        u32 np;
        tile_setup(L, t, s_off, p0, np);
        u64 r[LPT];
        u32 part[LPT], rank[LPT];
        const u64 jt = j0 + (u64)threadIdx.x * LPT;
        if (jt < j1) {
            LeafWalk w(L, s_off, p0, np, jt);
#pragma unroll
            for (int k = 0; k < LPT; ++k) {
                if (jt + k < j1) {
                    u64 key, id;
                    u32 inst;
                    w.eval_key(jt + k, key);
                    const u64 hh = khash(key);
                    part[k] = (u32)(hh >> (64 - kh.lgP));
                    const u32 lo = (u32)((hbucket(hh, kh.lgNB) & bomask) << 12) | hfp(hh);
                    r[k] = ((u64)(jt + k - L.lo) << 32) | lo;
                    rank[k] = atomicAdd(&h[part[k]], 1u);
                }
            }
        }
        __syncthreads();
        {   // local exclusive offsets of the partitions within the tile; one reservation per partition
            const u32 v = h[threadIdx.x];
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            u32 inc = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const u32 o = __shfl_up(inc, d, 64);
                if (lane >= d) inc += o;
            }
            if (lane == 63) s_w[wv] = inc;
            __syncthreads();
            u32 pre = 0;
            for (int q = 0; q < wv; ++q) pre += s_w[q];
            loc[threadIdx.x] = pre + inc - v;
            u64 gb = NONE;
            if (threadIdx.x < (unsigned)P && v) {
                const u64 r0 = atomicAdd(&pcur[threadIdx.x], v);
                const u64 b0 = pbase[threadIdx.x] + r0;
                if (b0 + v <= pbase[threadIdx.x + 1]) gb = b0;
                else atomicOr(ovf, 1u);
            }
            gbase[threadIdx.x] = gb;
        }
        __syncthreads();
        if (jt < j1) {
#pragma unroll
            for (int k = 0; k < LPT; ++k) {
                if (jt + k < j1) {
                    const u32 at = loc[part[k]] + rank[k];
                    s_r[at] = r[k];
                    s_p[at] = (u8_t)part[k];
                }
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += PB) {
            const u32 pi = s_p[q];
            if (gbase[pi] == NONE) continue;
            rec[gbase[pi] + (u64)(q - (int)loc[pi])] = s_r[q];
        }
        __syncthreads();
    }
}

// ---- probe table: 8 fingerprint slots of 16 bits per bucket ---------------------------------
__global__ void __launch_bounds__(PB) k_table_build(const u64 *__restrict__ bstart, const u64 *__restrict__ Dk, u64 NB,
                                                   unsigned short *__restrict__ T) {
    for (u64 b = (u64)blockIdx.x * PB + threadIdx.x; b < NB; b += (u64)gridDim.x * PB) {
        const u64 lo = bstart[b], cnt = bstart[b + 1] - lo;
        u32 f[TSLOTS];
#pragma unroll
        for (int k = 0; k < TSLOTS; ++k) f[k] = (u64)k < cnt ? (0x8000u | hfp(khash(Dk[lo + k]))) : 0u;
        *(uint4 *)(T + b * TSLOTS) = make_uint4(f[0] | (f[1] << 16), f[2] | (f[3] << 16), f[4] | (f[5] << 16),
                                                f[6] | (f[7] << 16));
    }
}

// Instance of a key: the last koff <= key (koffs ascending, koffs[0] = 0).
__device__ __forceinline__ u32 inst_of_key(const u64 *__restrict__ koffs, int ninst, u64 key) {
    int a = 0, b = ninst - 1;
    while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (koffs[mid] <= key) a = mid; else b = mid - 1;
    }
    return (u32)a;
}

// Exact key and id of leaf lo + leaf (rare: only on a fingerprint hit).
constexpr int PROBE_BATCH = 8;     // records per thread whose loads are issued before any compare

__device__ __forceinline__ void leaf_exact(const LeafSrc &L, u64 leaf, u64 &key, u64 &id) {
    const u64 j = L.lo + leaf;
    const size_t p = find_parent(L.off, 0, L.nn - 1, j);
    const Node nd = L.nodes[p];
    const SideDesc &s = L.descs[nd.meta >> 8];
    const int x = child_index(s, nd, (int)(j - L.off[p]));
    const u64 xm = cf_mulmod_m(nd.xm, s.pR[x], s.mc);
    const u64 x2 = nd.x2 * s.p[x];
    key = cf_key(xm, x2, s.mc) + s.koff;
    id = id_update(s, nd.id, x, L.depth);
}

// Probe, pass 1: each record reads its bucket's 16 bytes; a record whose fingerprint is among the
// bucket's slots, or whose bucket holds more than 8 D records, is a candidate. Candidates of a
// 4096-record tile are placed with one global reservation per tile (a block scan of per-thread
// hit counts), into cand (capacity = nslots, so the list cannot overflow).
// Does nothing if the scatter overflowed (*ovf).
__global__ void __launch_bounds__(PB) k_part_probe(KeyHash kh, const u64 *__restrict__ rec, const u64 *__restrict__ pbase,
                                                   const u32 *__restrict__ pfill, const u32 *__restrict__ ovf,
                                                   const unsigned short *__restrict__ T,
                                                   unsigned long long *__restrict__ ncand, u64 *__restrict__ cand) {
    __shared__ u64 s_base[257];
    __shared__ u64 s_end[256];
    __shared__ u32 s_w[PB / 64];
    __shared__ u64 s_gbase;
    if (*ovf) return;
    const int P = 1 << kh.lgP;
    const int boshift = kh.lgNB - kh.lgP;
    for (int i = threadIdx.x; i <= P; i += PB) s_base[i] = pbase[i];
    for (int i = threadIdx.x; i < P; i += PB) s_end[i] = pbase[i] + pfill[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u64 nslots = s_base[P];
    const u64 ntiles = (nslots + PT - 1) / PT;
    // XCD-aware order: block b runs on XCD b % 8; give each XCD a contiguous slice of tiles
    const u32 G = gridDim.x;
    const u32 xcd = blockIdx.x % NXCD;
    const u32 nb_x = G / NXCD + (xcd < G % NXCD ? 1 : 0);     // blocks on this XCD
    const u32 kb = blockIdx.x / NXCD;                         // rank of this block on its XCD
    const u64 t_lo = ntiles * xcd / NXCD, t_hi = ntiles * (xcd + 1) / NXCD;
    for (u64 t = t_lo + kb; t < t_hi; t += nb_x) {
        const u64 e0 = t * PT;
        int p = 0;
        {
            int a = 0, b = P - 1;
            while (a < b) {
                const int mid = (a + b + 1) >> 1;
                if (s_base[mid] <= e0) a = mid; else b = mid - 1;
            }
            p = a;
        }
        u32 hmask = 0;
        for (int it0 = 0; it0 < PI; it0 += PROBE_BATCH) {
            u32 lo[PROBE_BATCH];
            u64 b[PROBE_BATCH];
            bool ok[PROBE_BATCH];
            uint4 tv[PROBE_BATCH];
#pragma unroll
            for (int q = 0; q < PROBE_BATCH; ++q) {
                const u64 e = e0 + (u64)(it0 + q) * PB + threadIdx.x;
                int pp = p;
                while (pp + 1 < P && e >= s_base[pp + 1]) ++pp;
                ok[q] = e < nslots && e < s_end[pp];
                lo[q] = ok[q] ? (u32)rec[e] : 0u;
                b[q] = ((u64)pp << boshift) | (lo[q] >> 12);
            }
#pragma unroll
            for (int q = 0; q < PROBE_BATCH; ++q) tv[q] = *(const uint4 *)(T + b[q] * TSLOTS);
#pragma unroll
            for (int q = 0; q < PROBE_BATCH; ++q) {
                const u32 fs = 0x8000u | (lo[q] & 0xFFFu);
                const uint4 v = tv[q];
                const u32 w[4] = {v.x, v.y, v.z, v.w};
                bool hit = (v.w >> 16) != 0;     // more than 8 D records in the bucket
#pragma unroll
                for (int k = 0; k < 4; ++k) hit |= ((w[k] & 0xFFFFu) == fs) | ((w[k] >> 16) == fs);
                if (hit && ok[q]) hmask |= 1u << (it0 + q);
            }
        }
        // place the tile's candidates: block exclusive scan of the per-thread counts
        const u32 c = (u32)__popc(hmask);
        u32 inc = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const u32 o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) s_w[wv] = inc;
        __syncthreads();
        u32 pre = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < PB / 64; ++q) {
            if (q < wv) pre += s_w[q];
            tot += s_w[q];
        }
        if (threadIdx.x == 0) s_gbase = tot ? atomicAdd(ncand, (unsigned long long)tot) : 0;
        __syncthreads();
        u64 at = s_gbase + pre + inc - c;
        while (hmask) {
            const int it = __ffs(hmask) - 1;
            hmask &= hmask - 1;
            cand[at++] = e0 + (u64)it * PB + threadIdx.x;
        }
        __syncthreads();
    }
}

// Probe, pass 2: each candidate's exact key and id are recomputed from the node tree and compared
// with the D records of its bucket; every equal key emits a pair.
__global__ void __launch_bounds__(PB) k_probe_candidates(LeafSrc L, KeyHash kh, const u64 *__restrict__ rec,
                                                         const u64 *__restrict__ pbase,
                                                         const unsigned long long *__restrict__ ncand,
                                                         const u64 *__restrict__ cand, const u64 *__restrict__ koffs,
                                                         int ninst, const unsigned short *__restrict__ T,
                                                         const u64 *__restrict__ bstart, const u64 *__restrict__ Dk,
                                                         const u64 *__restrict__ Did,
                                                         unsigned long long *__restrict__ npairs, u64 cap,
                                                         u64 *__restrict__ out_d, u64 *__restrict__ out_i,
                                                         u64 *__restrict__ out_inst) {
    __shared__ u64 s_base[257];
    const int P = 1 << kh.lgP;
    const int boshift = kh.lgNB - kh.lgP;
    for (int i = threadIdx.x; i <= P; i += PB) s_base[i] = pbase[i];
    __syncthreads();
    const u64 nc = *ncand;
    for (u64 c = (u64)blockIdx.x * PB + threadIdx.x; c < nc; c += (u64)gridDim.x * PB) {
        const u64 e = cand[c];
        int a = 0, bb = P - 1;
        while (a < bb) {
            const int mid = (a + bb + 1) >> 1;
            if (s_base[mid] <= e) a = mid; else bb = mid - 1;
        }
        const u64 r = rec[e];
        const u64 b = ((u64)a << boshift) | ((u32)r >> 12);
        const u32 fs = 0x8000u | ((u32)r & 0xFFFu);
        const uint4 v = *(const uint4 *)(T + b * TSLOTS);
        const u32 w[4] = {v.x, v.y, v.z, v.w};
        u64 key, iid;
        leaf_exact(L, r >> 32, key, iid);
        const u64 lo = bstart[b], hi = bstart[b + 1];
        for (u64 q = lo; q < hi; ++q) {
            if (q < lo + TSLOTS && ((w[(q - lo) >> 1] >> (16 * ((q - lo) & 1))) & 0xFFFFu) != fs) continue;
            if (Dk[q] != key) continue;
            const unsigned long long at = atomicAdd(npairs, 1ULL);
            if (at < cap) {
                out_d[at] = Did[q];
                out_i[at] = iid;
                out_inst[at] = inst_of_key(koffs, ninst, key);
            }
        }
    }
}

// Per-instance record totals from node boundaries: nodes of one instance are contiguous, so
// instance j's leaves are off[last+1] - off[first].
__global__ void __launch_bounds__(PB) k_inst_bounds(const Node *__restrict__ nodes, const u64 *__restrict__ off, size_t nn,
                                                    u64 *__restrict__ out) {
    for (size_t j = (size_t)blockIdx.x * PB + threadIdx.x; j < nn; j += (size_t)gridDim.x * PB) {
        const u32 inst = nodes[j].meta >> 8;
        // the first and the last node of an instance may be different lanes: add atomically
        if (j == 0 || (nodes[j - 1].meta >> 8) != inst)
            atomicAdd((unsigned long long *)&out[inst], (unsigned long long)(0 - off[j]));
        if (j == nn - 1 || (nodes[j + 1].meta >> 8) != inst)
            atomicAdd((unsigned long long *)&out[inst], (unsigned long long)off[j + 1]);
    }
}

__global__ void __launch_bounds__(PB) k_fill_u64(u64 *__restrict__ a, size_t n, u64 v) {
    for (size_t j = (size_t)blockIdx.x * PB + threadIdx.x; j < n; j += (size_t)gridDim.x * PB) a[j] = v;
}
__global__ void __launch_bounds__(PB) k_fill_u32(u32 *__restrict__ a, size_t n, u32 v) {
    for (size_t j = (size_t)blockIdx.x * PB + threadIdx.x; j < n; j += (size_t)gridDim.x * PB) a[j] = v;
}

inline unsigned grid_for(u64 n, u64 per, unsigned cap = 1u << 20) {
    u64 g = (n + per - 1) / per;
    if (g < 1) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

}  // namespace

void fill_u64(cf_ctx *ctx, u64 *a, size_t n, u64 v) {
    if (!n) return;
    CF_LAUNCH(ctx, "k_fill_u64", k_fill_u64, dim3(grid_for(n, PB, 65536)), dim3(PB), 0, a, n, v);
}

void fill_u32(cf_ctx *ctx, u32 *a, size_t n, u32 v) {
    if (!n) return;
    CF_LAUNCH(ctx, "k_fill_u32", k_fill_u32, dim3(grid_for(n, PB, 65536)), dim3(PB), 0, a, n, v);
}

// Leaf source of [lo, lo + n) with its tile owners computed into tp (2 * ceil(n / LT) entries).
static LeafSrc leaf_src(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                        int depth, DBuf<u64> &tp) {
    const u64 ntiles = (n + LT - 1) / LT;
    tp = DBuf<u64>(ctx, 2 * ntiles);
    CF_LAUNCH(ctx, "k_tile_parents", k_tile_parents, dim3(grid_for(2 * ntiles, PB, 65536)), dim3(PB), 0, off, nn, lo, n,
              tp.p);
    return LeafSrc{descs, nodes, nn, off, lo, n, depth, tp.p};
}

void part_bucket_count(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                       int depth, int lgNB, u32 *bcount) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    CF_LAUNCH(ctx, "k_bucket_count", k_bucket_count, dim3(grid_for(n, LT)), dim3(PB), 0, L, lgNB, bcount);
}

void part_bucket_scatter(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo,
                         u64 n, int depth, int lgNB, const u64 *bstart, u32 *bcur, u64 *Dk, u64 *Did) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    CF_LAUNCH(ctx, "k_bucket_scatter", k_bucket_scatter, dim3(grid_for(n, LT)), dim3(PB), 0, L, lgNB, bstart, bcur, Dk,
              Did);
}

void part_table_build(cf_ctx *ctx, const u64 *bstart, const u64 *Dk, u64 NB, unsigned short *T) {
    CF_LAUNCH(ctx, "k_table_build", k_table_build, dim3(grid_for(NB, PB, 65536)), dim3(PB), 0, bstart, Dk, NB, T);
}

void part_count(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                int depth, int lgP, u32 *pcount) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    CF_LAUNCH(ctx, "k_part_count", k_part_count, dim3(grid_for(n, LT)), dim3(PB), 0, L, lgP, pcount);
}

void part_scatter(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                  int depth, int lgNB, int lgP, const u64 *pbase, u32 *pcur, u32 *ovf, u64 *rec) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    CF_LAUNCH(ctx, "k_part_scatter", k_part_scatter, dim3(grid_for(n, LT)), dim3(PB), 0, L, KeyHash{lgNB, lgP}, pbase,
              pcur, ovf, rec);
}

void part_probe(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, int depth,
                int lgNB, int lgP, const u64 *rec, const u64 *pbase, const u32 *pfill, u64 nslots, const u32 *ovf,
                const u64 *koffs, int ninst, const unsigned short *T, const u64 *bstart, const u64 *Dk, const u64 *Did,
                unsigned long long *npairs, u64 cap, u64 *od, u64 *oi, u64 *oinst) {
    if (!nslots) return;
    const KeyHash kh{lgNB, lgP};
    DBuf<u64> ncand(ctx, 1), cand(ctx, nslots);
    fill_u64(ctx, ncand.p, 1, 0);
    // the tile order gives each XCD a slice of tiles, so every XCD needs at least one block:
    // the grid is a non-zero multiple of the XCD count
    const u64 ntiles = (nslots + PT - 1) / PT;
    unsigned g = (unsigned)(ntiles < 8192 ? ntiles : 8192);
    g = g < NXCD ? NXCD : g - g % NXCD;
    CF_LAUNCH(ctx, "k_part_probe", k_part_probe, dim3(g), dim3(PB), 0, kh, rec, pbase, pfill, ovf, T,
              (unsigned long long *)ncand.p, cand.p);
    LeafSrc L{descs, nodes, nn, off, lo, 0, depth, nullptr};
    CF_LAUNCH(ctx, "k_probe_candidates", k_probe_candidates, dim3(1024), dim3(PB), 0, L, kh, rec, pbase,
              (const unsigned long long *)ncand.p, cand.p, koffs, ninst, T, bstart, Dk, Did, npairs, cap, od, oi, oinst);
}

void inst_bounds(cf_ctx *ctx, const Node *nodes, const u64 *off, size_t nn, u64 *out) {
    if (!nn) return;
    CF_LAUNCH(ctx, "k_inst_bounds", k_inst_bounds, dim3(grid_for(nn, PB, 65536)), dim3(PB), 0, nodes, off, nn, out);
}
