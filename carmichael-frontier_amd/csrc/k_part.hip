// k_part.hip -- radix-partitioned residue join on gfx950.
//
// Keys of one solve live in one key space (instance j owns [koff_j, koff_j + M_j), so keys of
// different instances never collide). Every leaf (a D or an I record) is one residue step below
// a node of its enumeration tree; the step is the 32-bit scaled one (cf_redc32_mul) when every
// prime of the instance is below 2^32, so a leaf's residue state is (xm * 2^-32 mod m, x2 mod 2^a)
// on both sides and equal residues still meet.
//
// Hash of a leaf: v = (xm + rot32(x2 mod 2^a) + koff) * 0x9E3779B97F4A7C15, then h = (v ^ v >> 32) *
// 0xD6E8FEB86659FD93 -- no CRT step is needed to place a record, only to compare one exactly. The
// second round breaks up the structure a single multiplicative hash keeps from consecutive leaves'
// keys (products sharing all but their last prime), which piled one block's records of a large D
// side into a few partitions. h gives the record's partition, hash range,
// bucket and fingerprint (keyhash.h); NB is chosen for about 2 D records per bucket, and a
// partition's buckets are contiguous, so its slice of the probe table is NB * 16 / P bytes (about
// 2 MB).
//
// Probe table T: 8 slots of 16 bits per bucket (16 bytes), slot k holding 0x8000 | f of the
// bucket's k-th D record (f cut to the fpb fingerprint bits an I record keeps, at most 15); slot 7
// holds T_OVERFLOW instead when the bucket holds more than 8 records (rare at 2 per bucket: such a
// bucket makes every probe of it a candidate). Bucket filter F: a few bits per bucket (below); the
// probe keeps its partition's slice of F in LDS and reads T only for records the filter passes.
//
//  D side: records counted per bucket and scattered into bucket order (DR: the D leaf index; a
//          record's exact key is recomputed from it when a candidate meets it); the scatter writes
//          the table slots as it goes; F is built from T.
//  I side: leaves scattered by partition as 8-byte records (low word: bucket offset << fpb | f,
//          high word: leaf index in the chunk): each block ranks a tile of 4096 leaves by partition
//          in LDS and writes one run per partition into slots sized from the expected share.
//  Probe:  pass 1 reads the records in partition order: each of the 8 XCDs takes tiles of its own
//          contiguous slice from a queue, so an XCD works on one partition at a time and that
//          partition's table slice stays in its L2. Pass 2 recomputes candidates exactly.
// HBM traffic per I record: 8 B written by the scatter + 8 B read by the probe.
// Join semantics: every equal-key (D, I) pair (gpu/mitm_gpu2.py:107-119).
#include <type_traits>

#include "tree_dev.h"

namespace {

constexpr int PB = 256;            // threads per block (other kernels)
constexpr int LPB = 512;           // threads per block of the leaf kernels
constexpr int LPT = 8;             // leaves per thread in the leaf kernels
constexpr int LT = LPB * LPT;      // 4096 leaves per tile
constexpr int NXCD = 8;            // MI355X: 8 accelerator complex dies, blocks dealt round-robin
constexpr int TSLOTS = CF_TABLE_SLOTS;   // table slots per bucket
// A DR entry: the D record's leaf index (relative to its table's first leaf), above it the record's
// fingerprint (the fpb bits its table slot holds, 15 bits of room) and 4 more bits of its hash (hash
// bits 12-15, used by neither the bucket nor the fingerprint), so that a candidate recomputes only
// the keys of records that agree with it on fpb + 4 hash bits.
constexpr int DR_FP_SHIFT = 40;
constexpr int DR_EX_SHIFT = 55;
constexpr u32 DR_FP_MASK = 0x7FFFu;
constexpr unsigned short T_OVERFLOW = 0x0001;   // slot 7 of a bucket of more than 8 records (no valid bit)
constexpr u64 DR_LEAF_MASK = ((u64)1 << DR_FP_SHIFT) - 1;
__device__ __forceinline__ u32 hash_ex(u64 h) { return (u32)(h >> 12) & 15u; }
// A window-built D table holds fewer than 2^28 records, so its records' leaf indices leave their top
// 4 bits to carry hash_ex from the scatter to the window sort.
constexpr int DREC_EX_SHIFT = 28;

// Block barrier for LDS traffic only: orders the block's LDS accesses without waiting for its
// outstanding global loads and stores (__syncthreads is also a workgroup memory fence, which on this
// target waits for every vector memory operation in flight, prefetches included). For kernels whose
// threads never read global data another thread of the block wrote.
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct LeafSrc {
    const SideDesc *descs;
    const LeafNode *nodes;
    size_t nn;
    const u64 *off;   // exclusive scan of leaves per node, nn + 1 entries
    u64 lo, n;        // leaves [lo, lo + n)
    int depth;        // r - 1
    const u64 *tp;    // per tile of LT leaves: owners of its first and last leaf (2 entries per tile)
};

// residue state of the leaf choosing index x below node nd
__device__ __forceinline__ void leaf_residue(const SideDesc &s, const LeafNode &nd, int x, u64 &xm, u64 &x2) {
    if (s.leaf32) {
        const u32 p = (u32)s.p[x];
        xm = cf_redc32_mul(nd.xm, p, s.mc.m, (u32)s.mc.mneg);
        x2 = nd.x2 * (u64)p;
    } else {
        xm = cf_mulmod_m(nd.xm, s.pR[x], s.mc);
        x2 = nd.x2 * s.p[x];
    }
}
__device__ __forceinline__ u64 leaf_hash(u64 xm, u64 x2, u64 mask2, u64 koff) {
    const u64 x2m = x2 & mask2;
    const u64 v = (xm + ((x2m << 32) | (x2m >> 32)) + koff) * 0x9E3779B97F4A7C15ULL;
    return (v ^ (v >> 32)) * 0xD6E8FEB86659FD93ULL;
}

// ---- leaf tiles ------------------------------------------------------------------------------
// A tile is LT consecutive leaves [j0, j0 + cnt). Lane i of the block takes leaves j0 + i + LPB * k,
// so the 64 lanes of a wave work on 64 neighbouring leaves. The tile's parent nodes are
// p0 .. p0 + np - 1 (every last-level node owns at least one leaf, so np <= LT). Set-up builds the
// owner map own[i] = parent of leaf j0 + i (each node's first leaf is marked with its number, a
// running maximum fills the rest) and, for a tile whose parents all belong to one instance with
// primes below 2^32 and number at most NCAP (nearly every tile), stages in LDS what a leaf needs:
// the parents' residue state, next index and first child number, the instance's primes and its
// modulus constants. A leaf is then a few LDS reads and one 32-bit Montgomery step; other tiles
// read nodes and descriptors from global memory.
constexpr int NCAP = 2048;         // parents staged per tile
constexpr int PCAP = 1024;         // primes staged (pools hold at most 1024)

struct TileLds {
    u64 xm[NCAP];                  // xm and x2 (32 KB) double as the scatter's staging area
    u64 x2[NCAP];
    int kb[NCAP];                  // child number of leaf j0 under parent q, i.e. j0 - off[p0 + q]
    int nx[NCAP];                  // parent's next index
    u32 prime[PCAP];
    unsigned short own[LT];
    u32 w[LPB / 64];
    u64 m, mask2, koff, minv;
    u32 mneg32;
    int sideI, inst;
    int narrow;                    // the staged instance's 2^a is at most 2^32 (x2 needs 32 bits)
    int fastok;                    // the staged instance takes the fast path (leaf32, n <= PCAP)
    u32 meta_first, meta_last;     // instance fields of the tile's first and last parent
};

struct Tile {
    u64 j0;
    u32 cnt;
    size_t p0;
    u32 np;
    bool fast;
};

// OWN_ZEROED: the caller keeps own[] all zero between tiles (it clears it once a tile's leaves are
// hashed), so a parent marks its first leaf in the same pass that stages it. The instance's primes
// and constants are staged only when a fast tile's instance differs from the last one staged
// (S.inst, -1 before the first tile: the caller sets it).
template <bool OWN_ZEROED = false>
__device__ __forceinline__ Tile tile_setup(const LeafSrc &L, u64 t, TileLds &S) {
    Tile T;
    T.j0 = L.lo + t * LT;
    const u64 jend = L.lo + L.n;
    T.cnt = (u32)((T.j0 + LT < jend ? T.j0 + LT : jend) - T.j0);
    T.p0 = L.tp[2 * t];
    T.np = (u32)(L.tp[2 * t + 1] - T.p0 + 1);
    const u32 i0 = L.nodes[T.p0].meta >> 8, i1 = L.nodes[T.p0 + T.np - 1].meta >> 8;
    const SideDesc &sd = L.descs[i0];
    T.fast = i0 == i1 && T.np <= (u32)NCAP && sd.leaf32 && sd.n <= PCAP;
    const bool restage = T.fast && S.inst != (int)i0;   // read before any thread rewrites S.inst
    if (!OWN_ZEROED)
        for (u32 i = threadIdx.x; i < (u32)LT; i += LPB) S.own[i] = 0;
    if (T.fast) {
        for (u32 q = threadIdx.x; q < T.np; q += LPB) {
            const LeafNode &nd = L.nodes[T.p0 + q];
            const u64 o = L.off[T.p0 + q];
            S.xm[q] = nd.xm;
            S.x2[q] = nd.x2;
            S.nx[q] = node_nxt(nd);
            S.kb[q] = (int)(long long)(T.j0 - o);
            if (OWN_ZEROED && q) S.own[o - T.j0] = (unsigned short)q;
        }
        if (restage) {
            for (int x = threadIdx.x; x < sd.n; x += LPB) S.prime[x] = (u32)sd.p[x];
            if (threadIdx.x == 0) {
                S.m = sd.mc.m;
                S.mask2 = sd.mc.mask2;
                S.koff = sd.koff;
                S.minv = sd.mc.minv;
                S.mneg32 = (u32)sd.mc.mneg;
                S.sideI = sd.side == SIDE_I;
                S.narrow = sd.mc.mask2 <= 0xFFFFFFFFull;
            }
        }
    } else if (OWN_ZEROED) {
        for (u32 q = 1 + threadIdx.x; q < T.np; q += LPB) S.own[L.off[T.p0 + q] - T.j0] = (unsigned short)q;
    }
    lds_sync();
    if (!OWN_ZEROED) {
        if (T.fast)
            for (u32 q = 1 + threadIdx.x; q < T.np; q += LPB) S.own[-S.kb[q]] = (unsigned short)q;
        else
            for (u32 q = 1 + threadIdx.x; q < T.np; q += LPB) S.own[L.off[T.p0 + q] - T.j0] = (unsigned short)q;
    }
    if (restage && threadIdx.x == 0) S.inst = (int)i0;
    if (!OWN_ZEROED) lds_sync();
    // running maximum over own: LPT entries per thread, then a block scan of the thread maxima
    u32 run = 0;
    unsigned short v[LPT];
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        v[k] = S.own[threadIdx.x * LPT + k];
        run = run > v[k] ? run : v[k];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u32 inc = wave_incl_max(run);
    if (lane == 63) S.w[wv] = inc;
    lds_sync();
    u32 pre = 0;
    for (int q = 0; q < wv; ++q) pre = pre > S.w[q] ? pre : S.w[q];
    const u32 ex = wave_shift_up1(inc);
    u32 cur = pre > ex ? pre : ex;
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        cur = cur > v[k] ? cur : v[k];
        S.own[threadIdx.x * LPT + k] = (unsigned short)cur;
    }
    lds_sync();
    return T;
}

// Parents of a tile fetched ahead into registers (the scatter's pipeline): the tile's first and
// last parent, and for q = threadIdx.x < np, parent p0 + q's residue state, next index, instance
// field and first leaf.
struct TilePF {
    u64 p0, p1;
    u64 xm, x2, off;
    u32 pos, meta;
};
__device__ __forceinline__ void tile_fetch(const LeafSrc &L, u64 p0, u64 p1, TilePF &f) {
    f.p0 = p0;
    f.p1 = p1;
    const u64 q = p0 + threadIdx.x;
    if (q <= p1) {
        const LeafNode *nd = L.nodes + q;
        f.xm = nd->xm;
        f.x2 = nd->x2;
        f.pos = nd->pos;
        f.meta = nd->meta;
        f.off = L.off[q];
    }
}
// A tile's owner pair as a vector load (lane k holds tp[2t + (k & 1)]), so that waiting for it does
// not wait for scalar loads; lanes 0 and 1 are read out when it is used.
__device__ __forceinline__ u64 tile_owner_load(const LeafSrc &L, u64 t) { return L.tp[2 * t + (threadIdx.x & 1)]; }
__device__ __forceinline__ u64 lane_u64(u64 v, int lane) {
    return ((u64)(u32)__builtin_amdgcn_readlane((int)(v >> 32), lane) << 32) |
           (u64)(u32)__builtin_amdgcn_readlane((int)(u32)v, lane);
}

// Tile set-up from fetched parents (owner map kept zero between tiles by the caller). Whether the
// tile takes the fast path is known from the staged instance unless the tile's instance differs;
// only then is the descriptor read (and the primes and constants restaged).
__device__ __forceinline__ Tile tile_setup_pf(const LeafSrc &L, u64 t, TileLds &S, const TilePF &f) {
    Tile T;
    T.j0 = L.lo + t * LT;
    const u64 jend = L.lo + L.n;
    T.cnt = (u32)((T.j0 + LT < jend ? T.j0 + LT : jend) - T.j0);
    T.p0 = f.p0;
    T.np = (u32)(f.p1 - f.p0 + 1);
    {
        const u32 q = threadIdx.x;
        if (q < T.np) {
            if (q < (u32)NCAP) {
                S.xm[q] = f.xm;
                S.x2[q] = f.x2;
                S.nx[q] = (int)(short)(f.pos & 0xFFFFu);
                S.kb[q] = (int)(long long)(T.j0 - f.off);
            }
            if (q) S.own[f.off - T.j0] = (unsigned short)q;
            if (q == 0) S.meta_first = f.meta;
            if (q == T.np - 1) S.meta_last = f.meta;
        }
        // parents beyond the block's threads (a tile of many short subtrees): read here
        for (u32 qq = LPB + threadIdx.x; qq < T.np; qq += LPB) {
            const LeafNode &nd = L.nodes[T.p0 + qq];
            const u64 o = L.off[T.p0 + qq];
            if (qq < (u32)NCAP) {
                S.xm[qq] = nd.xm;
                S.x2[qq] = nd.x2;
                S.nx[qq] = node_nxt(nd);
                S.kb[qq] = (int)(long long)(T.j0 - o);
            }
            S.own[o - T.j0] = (unsigned short)qq;
            if (qq == T.np - 1) S.meta_last = nd.meta;
        }
    }
    lds_sync();
    const u32 i0 = S.meta_first >> 8, i1 = S.meta_last >> 8;
    bool fastok;
    if (i0 == i1 && (int)i0 != S.inst) {
        // a new instance: its descriptor decides the path; primes and constants restaged
        const SideDesc &sd = L.descs[i0];
        fastok = sd.leaf32 && sd.n <= PCAP;
        if (fastok) {
            for (int x = threadIdx.x; x < sd.n; x += LPB) S.prime[x] = (u32)sd.p[x];
            if (threadIdx.x == 0) {
                S.m = sd.mc.m;
                S.mask2 = sd.mc.mask2;
                S.koff = sd.koff;
                S.minv = sd.mc.minv;
                S.mneg32 = (u32)sd.mc.mneg;
                S.sideI = sd.side == SIDE_I;
                S.narrow = sd.mc.mask2 <= 0xFFFFFFFFull;
            }
        }
        lds_sync();   // every thread has read S.inst and S.fastok above
        if (threadIdx.x == 0) {
            S.inst = (int)i0;
            S.fastok = fastok;
        }
    } else {
        fastok = S.fastok != 0;
    }
    T.fast = i0 == i1 && T.np <= (u32)NCAP && fastok;
    // running maximum over own: LPT entries per thread, then a block scan of the thread maxima
    u32 run = 0;
    unsigned short v[LPT];
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        v[k] = S.own[threadIdx.x * LPT + k];
        run = run > v[k] ? run : v[k];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u32 inc = wave_incl_max(run);
    if (lane == 63) S.w[wv] = inc;
    lds_sync();
    u32 pre = 0;
    for (int q = 0; q < wv; ++q) pre = pre > S.w[q] ? pre : S.w[q];
    const u32 ex = wave_shift_up1(inc);
    u32 cur = pre > ex ? pre : ex;
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        cur = cur > v[k] ? cur : v[k];
        S.own[threadIdx.x * LPT + k] = (unsigned short)cur;
    }
    lds_sync();
    return T;
}

// Before a kernel's first tile_setup: no instance staged yet, owner map clear.
__device__ __forceinline__ void tile_lds_init(TileLds &S) {
    if (threadIdx.x == 0) {
        S.inst = -1;
        S.fastok = 0;
    }
    for (u32 i = threadIdx.x; i < (u32)LT; i += LPB) S.own[i] = 0;
    lds_sync();
}

// Modulus constants of a fast tile, read once from LDS.
struct FastCtx {
    u64 m, mask2, koff;
    u32 mneg32;
    bool sideI, narrow;
    __device__ __forceinline__ explicit FastCtx(const TileLds &S)
        : m(S.m), mask2(S.mask2), koff(S.koff), mneg32(S.mneg32), sideI(S.sideI != 0), narrow(S.narrow != 0) {}
};

// Fast tile: residue state of leaf j0 + i (any i < LT whose owner entry is valid) and its hash.
// NARROW (2^a <= 2^32): only the low word of x2 is carried (x2 is exact mod 2^a, which is all that
// the hash and the key read).
template <bool NARROW = false>
__device__ __forceinline__ u64 leaf_fast(const TileLds &S, const FastCtx &c, u32 i, u64 &xm, u64 &x2, int &x,
                                         u32 &q) {
    q = S.own[i];
    const int k = (int)i + S.kb[q];
    x = c.sideI ? S.nx[q] + k : S.nx[q] - k;
    const u32 p = S.prime[x];
    xm = cf_redc32_mul(S.xm[q], p, c.m, c.mneg32);
    x2 = NARROW ? (u64)((u32)S.x2[q] * p) : S.x2[q] * (u64)p;
    return leaf_hash(xm, x2, c.mask2, c.koff);
}

// Generic tile: the same from global memory.
struct LeafRef {
    LeafNode nd;
    const SideDesc *s;
    int x;
};
__device__ __forceinline__ LeafRef leaf_ref(const LeafSrc &L, const Tile &T, const TileLds &S, u32 i) {
    LeafRef r;
    const u32 q = S.own[i];
    r.nd = L.nodes[T.p0 + q];
    r.s = &L.descs[r.nd.meta >> 8];
    r.x = child_index(*r.s, r.nd, (int)(T.j0 + i - L.off[T.p0 + q]));
    return r;
}
__device__ __forceinline__ u64 leaf_generic(const LeafSrc &L, const Tile &T, const TileLds &S, u32 i, u64 &xm,
                                            u64 &x2, LeafRef &r) {
    r = leaf_ref(L, T, S, i);
    leaf_residue(*r.s, r.nd, r.x, xm, x2);
    return leaf_hash(xm, x2, r.s->mc.mask2, r.s->koff);
}

// Hashes of the LPT leaves of this lane on a fast tile; FULL: the tile holds LT leaves, else lanes
// past its end repeat its last leaf.
template <bool NARROW, bool FULL>
__device__ __forceinline__ void fast_hashes(const Tile &T, const TileLds &S, const FastCtx &c, u64 (&h)[LPT]) {
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        const u32 i = threadIdx.x + LPB * k;
        u64 xm, x2;
        int x;
        u32 q;
        h[k] = leaf_fast<NARROW>(S, c, FULL || i < T.cnt ? i : T.cnt - 1, xm, x2, x, q);
    }
}

// Hashes of the LPT leaves of this lane (FULL as in fast_hashes).
template <bool FULL = false>
__device__ __forceinline__ void tile_hashes(const LeafSrc &L, const Tile &T, const TileLds &S, u64 (&h)[LPT]) {
    if (T.fast) {
        const FastCtx c(S);
        if (c.narrow) fast_hashes<true, FULL>(T, S, c, h);
        else fast_hashes<false, FULL>(T, S, c, h);
    } else {
        for (int k = 0; k < LPT; ++k) {
            const u32 i = threadIdx.x + LPB * k;
            u64 xm, x2;
            LeafRef r;
            h[k] = leaf_generic(L, T, S, i < T.cnt ? i : T.cnt - 1, xm, x2, r);
        }
    }
}

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

// ---- plan records ------------------------------------------------------------------------------
// Records (key, id) of leaves [lo, lo + n) of a plan's last level, record j at j - lo: the leaf tiles
// above, with the tile's parents' partial ids staged beside their residue state. On a fast tile a
// leaf is a few LDS reads, one 32-bit Montgomery step and the exact key's CRT step; each lane writes
// neighbouring leaves, so the stores are whole lines.
template <bool NARROW>
__device__ __forceinline__ void fast_records(const LeafSrc &L, const Tile &T, const TileLds &S, const u64 *s_id,
                                             const SideDesc &sd, u64 *__restrict__ keys, u64 *__restrict__ ids) {
    const FastCtx c(S);
    const u64 jb = T.j0 - L.lo;
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
        const u32 i = threadIdx.x + LPB * k;
        if (i < T.cnt) {
            u64 xm, x2;
            int x;
            u32 q;
            leaf_fast<NARROW>(S, c, i, xm, x2, x, q);
            keys[jb + i] = xm + c.m * (((x2 - xm) * S.minv) & c.mask2) + c.koff;
            ids[jb + i] = id_update(sd, s_id[q], x, L.depth);
        }
    }
}

// Block k takes tiles [k * tpb, (k + 1) * tpb) with the scatter's pipeline: the next tile's parents
// (and their ids) are fetched into registers while this tile's records are computed and written.
__global__ void __launch_bounds__(LPB) k_leaf_records(LeafSrc L, u64 tpb, u64 *__restrict__ keys,
                                                      u64 *__restrict__ ids) {
    __shared__ TileLds S;
    __shared__ u64 s_id[NCAP];
    tile_lds_init(S);
    const u64 ntiles = (L.n + LT - 1) / LT;
    const u64 t0 = (u64)blockIdx.x * tpb;
    const u64 t1 = t0 + tpb < ntiles ? t0 + tpb : ntiles;
    TilePF pf{};
    u64 pid = 0;   // partial id of parent pf.p0 + threadIdx.x
    u64 tpv = 0;
    auto fetch = [&](u64 v) {
        const u64 p0 = lane_u64(v, 0), p1 = lane_u64(v, 1);
        tile_fetch(L, p0, p1, pf);
        if (p0 + threadIdx.x <= p1) pid = L.nodes[p0 + threadIdx.x].id;
    };
    if (t0 < t1) fetch(tile_owner_load(L, t0));
    if (t0 + 1 < t1) tpv = tile_owner_load(L, t0 + 1);
    for (u64 t = t0; t < t1; ++t) {
        if (threadIdx.x <= pf.p1 - pf.p0 && threadIdx.x < (u32)NCAP) s_id[threadIdx.x] = pid;
        const Tile T = tile_setup_pf(L, t, S, pf);
        if (t + 1 < t1) {
            fetch(tpv);
            if (t + 2 < t1) tpv = tile_owner_load(L, t + 2);
        }
        if (T.fast) {
            // parents past the block's threads: ids from memory; every parent's residue times 2^32,
            // so that the leaf's 32-bit step gives the exact residue (the join compares residues all
            // scaled by 2^-32; records carry keys)
            const SideDesc &sd = L.descs[S.inst];
            for (u32 q = threadIdx.x; q < T.np; q += LPB) {
                if (q >= (u32)LPB) s_id[q] = L.nodes[T.p0 + q].id;
                S.xm[q] = cf_mulmod_m(S.xm[q], sd.s32, sd.mc);
            }
            lds_sync();
            if (S.narrow) fast_records<true>(L, T, S, s_id, sd, keys, ids);
            else fast_records<false>(L, T, S, s_id, sd, keys, ids);
        } else {
            const u64 jb = T.j0 - L.lo;
            for (int k = 0; k < LPT; ++k) {
                const u32 i = threadIdx.x + LPB * k;
                if (i < T.cnt) {
                    const LeafRef r = leaf_ref(L, T, S, i);
                    const u64 xm = cf_mulmod_m(r.nd.xm, r.s->pR[r.x], r.s->mc), x2 = r.nd.x2 * r.s->p[r.x];
                    keys[jb + i] = cf_key(xm, x2, r.s->mc) + r.s->koff;
                    ids[jb + i] = id_update(*r.s, r.nd.id, r.x, L.depth);
                }
            }
        }
        lds_sync();   // the tile's LDS is read; its owner map goes back to zero for the next set-up
#pragma unroll
        for (int k = 0; k < LPT; ++k) S.own[threadIdx.x * LPT + k] = 0;
        lds_sync();
    }
}

// ---- D side: counting sort by bucket --------------------------------------------------------
__global__ void __launch_bounds__(LPB) k_bucket_count(LeafSrc L, KeyHash kh, u32 *__restrict__ bcount) {
    __shared__ TileLds S;
    tile_lds_init(S);
    const u64 ntiles = (L.n + LT - 1) / LT;
    for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile T = tile_setup(L, t, S);
        u64 h[LPT];
        tile_hashes(L, T, S, h);
#pragma unroll
        for (int k = 0; k < LPT; ++k)
            if (threadIdx.x + LPB * k < T.cnt) atomicAdd(&bcount[hbucket(h[k], kh)], 1u);
        lds_sync();
    }
}

// Scatter into bucket order; the rank a record takes in its bucket is also its table slot.
__global__ void __launch_bounds__(LPB) k_bucket_scatter(LeafSrc L, KeyHash kh, const u64 *__restrict__ bstart,
                                                       u32 *__restrict__ bcur, u64 *__restrict__ DR,
                                                       unsigned short *__restrict__ T) {
    __shared__ TileLds S;
    const u32 fmask = (1u << kh.fpb) - 1u;
    tile_lds_init(S);
    const u64 ntiles = (L.n + LT - 1) / LT;
    for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tile_setup(L, t, S);
        for (int k = 0; k < LPT; ++k) {
            const u32 i = threadIdx.x + LPB * k;
            if (i >= tl.cnt) continue;
            u64 xm, x2, h;
            if (tl.fast) {
                const FastCtx c(S);
                int x;
                u32 q;
                h = leaf_fast(S, c, i, xm, x2, x, q);
            } else {
                LeafRef r;
                h = leaf_generic(L, tl, S, i, xm, x2, r);
            }
            const u64 b = hbucket(h, kh);
            const u32 slot = atomicAdd(&bcur[b], 1u);
            const u64 b0 = bstart[b];
            // the leaf relative to the source's first, and its fingerprint
            DR[b0 + slot] = (tl.j0 - L.lo + i) | (u64)(hfp(h) & fmask) << DR_FP_SHIFT | (u64)hash_ex(h) << DR_EX_SHIFT;
            if (slot < (u32)TSLOTS) {
                const bool over = slot == TSLOTS - 1 && bstart[b + 1] - b0 > (u64)TSLOTS;
                T[b * TSLOTS + slot] = over ? T_OVERFLOW : (unsigned short)(0x8000u | (hfp(h) & fmask));
            }
        }
        lds_sync();
    }
}

// ---- I side: partition by the top bits of the hash -------------------------------------------
// Two passes over the same leaves and no global atomics: block k owns tiles [k * tpb, (k+1) * tpb)
// and first counts its leaves per partition into H[p * nblk + k]; an exclusive scan of H in that
// order gives O[p * nblk + k], where block k's run of partition p starts, so the second pass writes
// every record to its final slot, one contiguous run per partition and block.
__global__ void __launch_bounds__(LPB) k_part_hist(LeafSrc L, int lgP, u64 tpb, u32 nblk, u32 *__restrict__ H) {
    __shared__ TileLds S;
    __shared__ u32 hist[512];
    const int P = 1 << lgP;
    if (threadIdx.x < 512) hist[threadIdx.x] = 0;
    tile_lds_init(S);
    const u64 ntiles = (L.n + LT - 1) / LT;
    const u64 t0 = (u64)blockIdx.x * tpb;
    const u64 t1 = t0 + tpb < ntiles ? t0 + tpb : ntiles;
    for (u64 t = t0; t < t1; ++t) {
        const Tile T = tile_setup(L, t, S);
        u64 h[LPT];
        tile_hashes(L, T, S, h);
#pragma unroll
        for (int k = 0; k < LPT; ++k)
            if (threadIdx.x + LPB * k < T.cnt) atomicAdd(&hist[hpart(h[k], lgP)], 1u);
        lds_sync();
    }
    lds_sync();
    if (threadIdx.x < (unsigned)P) H[(u64)threadIdx.x * nblk + blockIdx.x] = hist[threadIdx.x];
}

// Second pass. Exact layout (O non-null): block k's run of partition p starts at O[p * nblk + k].
// Capacity layout (O null): block k owns the segment of 2^capl slots at (p * nblk + k) << capl for
// each partition p and writes its fill into fill[p * nblk + k]; a segment that would overflow sets
// *ovf (its excess records are not written), and the caller redoes the chunk with the exact layout.
// The capacity layout needs no first pass: its segments are sized for the expected share
// (hashing spreads records evenly) with a margin of about 8 standard deviations.
// Slots are 32-bit: a chunk holds at most 2^32 records, and the caller takes the capacity layout
// only when it has fewer than 2^32 slots (~0 marks a record past its segment's end).
//
// Per partition the block keeps its next slot gb and the room left in its segment rm (no limit in
// the exact layout), in registers of the scan waves (wave w, lane l: partitions 256w + 4l .. +3; one
// wave for up to 256 partitions, two for 512); per tile, the scan waves scan the partition counts
// (wave 1 adds wave 0's total, which it sums itself) and write one 64-bit word
// lc[p] = gb << 32 | min(room, 0xFFFF) << 16 | the partition's offset in the tile. A ranked record is staged at its partition's offset plus its
// rank, with its partition and place in the tile in the staged high word; the store side reads the
// staged records in order (so the lanes of a wave read the words of only a few partitions) and
// takes each record's slot from its partition's word.
// DSIDE: the D side's records, whose leaf indices (below 2^28) carry hash_ex in their top 4 bits.
// LGMAXP: partitions at most 2^LGMAXP (8 or 9); the staged high word holds the partition from bit 16
// and, on the D side, hash_ex above it.
template <bool SPLIT = false, bool DSIDE = false, int LGMAXP = 8>
__device__ __forceinline__ void part_scatter_body(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk,
                                                  const u64 *__restrict__ O, int capl, u32 *__restrict__ fill,
                                                  u32 *__restrict__ ovf, uint2 *__restrict__ rec,
                                                  u32 *__restrict__ rhi) {
    constexpr u32 MAXP = 1u << LGMAXP, PM = MAXP - 1;
    constexpr u32 NSW = MAXP / 256;                // scan waves
    __shared__ TileLds S;
    __shared__ u32 hist[MAXP];
    __shared__ u64 lc[MAXP];
    __shared__ u32 s_ovf;
    // once a tile's leaves are hashed its parents' staged state is dead: the records are staged over
    // xm and x2 (32 KB)
    uint2 *s_r = (uint2 *)S.xm;
    static_assert(sizeof(uint2) * LT <= sizeof(u64) * 2 * NCAP, "staging fits over xm, x2");
    const int P = 1 << kh.lgP;
    const u32 fmask = (1u << kh.fpb) - 1u;
    const bool exact = O != nullptr;
    const u32 cap = 1u << capl;
    constexpr int PPL = 256 / 64;   // partitions per lane of a scan wave
    const bool scanw = threadIdx.x < 64 * NSW;
    const u32 pb = (threadIdx.x >> 6) * 256 + (threadIdx.x & 63) * PPL;   // this lane's first partition
    u32 gb[PPL], rm[PPL], hv[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
        const u32 p = pb + j;
        const u64 g = (u64)p * nblk + blockIdx.x;
        gb[j] = scanw && p < (u32)P ? (exact ? (u32)O[g] : (u32)(g << capl)) : 0u;
        rm[j] = exact ? 0xFFFFu : cap;
        hv[j] = 0;
    }
    if (threadIdx.x == 0) s_ovf = 0;
    tile_lds_init(S);
    const u64 ntiles = (L.n + LT - 1) / LT;
    const u64 t0 = (u64)blockIdx.x * tpb;
    const u64 t1 = t0 + tpb < ntiles ? t0 + tpb : ntiles;
    // two-deep pipeline: tile t + 1's parents are fetched while tile t is ranked and written, and
    // tile t + 2's owner pair one tile earlier still
    TilePF pf{};
    u64 tpv = 0;
    if (t0 < t1) {
        const u64 v0 = tile_owner_load(L, t0);
        tile_fetch(L, lane_u64(v0, 0), lane_u64(v0, 1), pf);
    }
    if (t0 + 1 < t1) tpv = tile_owner_load(L, t0 + 1);
    // the same wait for the first tile's parents (every entry to the loop finds them in registers)
    asm volatile("" ::"v"(pf.xm), "v"(pf.x2), "v"(pf.off), "v"(pf.pos), "v"(pf.meta), "v"(tpv));
    // one tile: its leaves hashed, ranked, staged and written; FULL: the tile holds LT leaves (every
    // tile but the last of the chunk), so no lane needs a bound check. Full tiles and the last one
    // run in loops of their own.
    auto tile = [&](u64 t, auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        if (threadIdx.x < MAXP) hist[threadIdx.x] = 0;
        const Tile T = tile_setup_pf(L, t, S, pf);
        {
            // each leaf's record word (bucket offset in the partition << fpb | fingerprint) and partition
            u32 lo[LPT], pt[LPT];
            {
                u64 h[LPT];
                tile_hashes<FULL>(L, T, S, h);
#pragma unroll
                for (int k = 0; k < LPT; ++k) {
                    const u32 hi = (u32)(h[k] >> 32);
                    pt[k] = hi >> (32 - kh.lgP);
                    lo[k] = (hoff32(hi, kh) << kh.fpb) | (hfp(h[k]) & fmask);
                    if constexpr (DSIDE) pt[k] |= hash_ex(h[k]) << LGMAXP;   // staged above the partition
                }
            }
            if (t + 1 < t1) {
                tile_fetch(L, lane_u64(tpv, 0), lane_u64(tpv, 1), pf);
                if (t + 2 < t1) tpv = tile_owner_load(L, t + 2);
            }
            u32 rank[LPT];
#pragma unroll
            for (int k = 0; k < LPT; ++k)
                if (FULL || threadIdx.x + LPB * k < T.cnt) rank[k] = atomicAdd(&hist[pt[k] & PM], 1u);
            lds_sync();   // also: every lane is done with the staged parents and the owner map
#pragma unroll
            for (int k = 0; k < LPT; ++k) S.own[threadIdx.x * LPT + k] = 0;   // clear for the next tile
            // tile-local exclusive offsets of the partitions: the scan waves scan the counts, PPL per lane
            if (scanw) {
                u32 sum = 0;
#pragma unroll
                for (int j = 0; j < PPL; ++j) {
                    hv[j] = hist[pb + j];
                    sum += hv[j];
                }
                u32 off = wave_incl_sum(sum) - sum;
                if (NSW > 1 && threadIdx.x >= 64) {
                    // partitions 0 .. 255 come first: wave 1 sums wave 0's counts itself
                    const u32 l = threadIdx.x & 63;
                    u32 s0 = 0;
#pragma unroll
                    for (int j = 0; j < PPL; ++j) s0 += hist[l * PPL + j];
                    off += (u32)__builtin_amdgcn_readlane((int)wave_incl_sum(s0), 63);
                }
#pragma unroll
                for (int j = 0; j < PPL; ++j) {
                    const u32 p = pb + j;
                    if (p < (u32)P)
                        lc[p] = ((u64)gb[j] << 32) | ((rm[j] < 0xFFFFu ? rm[j] : 0xFFFFu) << 16) | off;
                    off += hv[j];
                }
            }
            lds_sync();
            // staging: each record at its partition's run in the tile (offset: the low half of lc)
            const u32 *lc32 = (const u32 *)lc;
#pragma unroll
            for (int k = 0; k < LPT; ++k) {
                const u32 i = threadIdx.x + LPB * k;
                if (FULL || i < T.cnt) {
                    const u32 at = (lc32[2 * (pt[k] & PM)] & 0xFFFFu) + rank[k];
                    s_r[at] = make_uint2(lo[k], (pt[k] << 16) | i);
                }
            }
            lds_sync();
            // The next tile's prefetched parents and owner pair are waited for here, before this
            // tile's stores: loads and stores share one counter that this target treats as
            // completing out of order, so a wait for a load issued before pending stores waits for
            // the stores as well. Waited for after them (at the next tile's set-up), every tile would
            // wait for its own stores to reach memory; waited for here, the stores drain while the
            // next tile is hashed.
            asm volatile("" ::"v"(pf.xm), "v"(pf.x2), "v"(pf.off), "v"(pf.pos), "v"(pf.meta), "v"(tpv));
            // every staged record and its partition's word read first, then the stores (staged
            // places past the tile's end hold stale words: their partition is masked to stay in
            // lc, and they are not stored)
            const u32 jb = (u32)(T.j0 - L.lo);
            uint2 r[LPT];
            u64 w[LPT];
#pragma unroll
            for (int k = 0; k < LPT; ++k) r[k] = s_r[threadIdx.x + LPB * k];
#pragma unroll
            for (int k = 0; k < LPT; ++k) w[k] = lc[(r[k].y >> 16) & PM];
            u32 over = 0;
#pragma unroll
            for (int k = 0; k < LPT; ++k) {
                const u32 q = threadIdx.x + LPB * k;
                if (FULL || q < T.cnt) {
                    const u32 rk = q - ((u32)w[k] & 0xFFFFu);
                    const bool fits = rk < ((u32)w[k] >> 16);
                    over |= !fits;
                    if (exact || fits) {
                        const u32 d = (u32)(w[k] >> 32) + rk;
                        if (SPLIT) {
                            ((u32 *)rec)[d] = r[k].x;
                            rhi[d] = jb + (r[k].y & 0xFFFFu);
                        } else {
                            rec[d] = make_uint2(r[k].x, (jb + (r[k].y & 0xFFFFu)) |
                                                            (DSIDE ? (r[k].y >> (16 + LGMAXP)) << DREC_EX_SHIFT : 0u));
                        }
                    }
                }
            }
            if (over) s_ovf = 1;
        }
        if (scanw) {
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                gb[j] += hv[j];
                if (!exact) rm[j] = rm[j] > hv[j] ? rm[j] - hv[j] : 0u;
            }
        }
        lds_sync();
    };
    const u64 tfull = t1 == ntiles && L.n % LT ? t1 - 1 : t1;   // tiles [t0, tfull) are full
    u64 t = t0;
    for (; t < tfull; ++t) tile(t, std::true_type{});
    for (; t < t1; ++t) tile(t, std::false_type{});
    lds_sync();
    if (!exact) {
        if (scanw) {
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
                const u32 p = pb + j;
                if (p < (u32)P) fill[(u64)p * nblk + blockIdx.x] = cap - rm[j];
            }
        }
        if (threadIdx.x == 0 && s_ovf) atomicOr(ovf, 1u);
    }
}

// I side and D side: the same records from separate kernels, so that profiles tell them apart. Two
// blocks per CU (four waves per SIMD: at most 128 VGPRs).
__global__ void __launch_bounds__(LPB) __attribute__((amdgpu_waves_per_eu(4))) k_part_scatter(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk, const u64 *__restrict__ O,
                                                     int capl, u32 *__restrict__ fill, u32 *__restrict__ ovf,
                                                     uint2 *__restrict__ rec) {
    part_scatter_body<false>(L, kh, tpb, nblk, O, capl, fill, ovf, rec, nullptr);
}
__global__ void __launch_bounds__(LPB) __attribute__((amdgpu_waves_per_eu(4))) k_part_scatter_d(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk, const u64 *__restrict__ O,
                                                     int capl, u32 *__restrict__ fill, u32 *__restrict__ ovf,
                                                     uint2 *__restrict__ rec) {
    part_scatter_body<false, true>(L, kh, tpb, nblk, O, capl, fill, ovf, rec, nullptr);
}
// The I side with its records split into two arrays of words (bucket words, then leaf indices), for
// chunks probed against several tables: the probe then reads 4 bytes per record instead of 8.
__global__ void __launch_bounds__(LPB) __attribute__((amdgpu_waves_per_eu(4))) k_part_scatter_split(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk, const u64 *__restrict__ O,
                                                     int capl, u32 *__restrict__ fill, u32 *__restrict__ ovf,
                                                     uint2 *__restrict__ rec, u32 *__restrict__ rhi) {
    part_scatter_body<true>(L, kh, tpb, nblk, O, capl, fill, ovf, rec, rhi);
}
// The same three for geometries of 512 partitions (large D sides).
__global__ void __launch_bounds__(LPB) __attribute__((amdgpu_waves_per_eu(4))) k_part_scatter_p512(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk, const u64 *__restrict__ O,
                                                     int capl, u32 *__restrict__ fill, u32 *__restrict__ ovf,
                                                     uint2 *__restrict__ rec) {
    part_scatter_body<false, false, 9>(L, kh, tpb, nblk, O, capl, fill, ovf, rec, nullptr);
}
__global__ void __launch_bounds__(LPB) __attribute__((amdgpu_waves_per_eu(4))) k_part_scatter_d_p512(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk, const u64 *__restrict__ O,
                                                     int capl, u32 *__restrict__ fill, u32 *__restrict__ ovf,
                                                     uint2 *__restrict__ rec) {
    part_scatter_body<false, true, 9>(L, kh, tpb, nblk, O, capl, fill, ovf, rec, nullptr);
}
__global__ void __launch_bounds__(LPB) __attribute__((amdgpu_waves_per_eu(4))) k_part_scatter_split_p512(LeafSrc L, KeyHash kh, u64 tpb, u32 nblk, const u64 *__restrict__ O,
                                                     int capl, u32 *__restrict__ fill, u32 *__restrict__ ovf,
                                                     uint2 *__restrict__ rec, u32 *__restrict__ rhi) {
    part_scatter_body<true, false, 9>(L, kh, tpb, nblk, O, capl, fill, ovf, rec, rhi);
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

// Exact key and id of leaf lo + leaf (rare: only on a fingerprint hit), and its hash.
__device__ __forceinline__ void leaf_exact(const LeafSrc &L, u64 leaf, u64 &key, u64 &id, u64 *hash = nullptr) {
    const u64 j = L.lo + leaf;
    // the tile owner map (when given) brackets the parent to the tile's few nodes
    const size_t p = L.tp ? find_parent(L.off, L.tp[2 * (leaf / LT)], L.tp[2 * (leaf / LT) + 1], j)
                          : find_parent(L.off, 0, L.nn - 1, j);
    const LeafNode nd = L.nodes[p];
    const SideDesc &s = L.descs[nd.meta >> 8];
    const int x = child_index(s, nd, (int)(j - L.off[p]));
    u64 xm, x2;
    leaf_residue(s, nd, x, xm, x2);
    key = cf_key(xm, x2, s.mc) + s.koff;
    id = id_update(s, nd.id, x, L.depth);
    if (hash) *hash = leaf_hash(xm, x2, s.mc.mask2, s.koff);
}

// Is the slot pattern fs2 (0x8000 | f in both halves of a word) among the 8 slots of a bucket? Slots
// hold 0x8000 | f with f cut to the record's fingerprint bits (slot 7 of an overflowing bucket holds
// T_OVERFLOW, without the valid bit, so it matches nothing), and a slot matches when it XOR the
// pattern is zero, so the 8 slots reduce to one packed 16-bit minimum.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bool bucket_hit(const uint4 &a, u32 fs2) {
    const u16x2 x0 = __builtin_bit_cast(u16x2, a.x ^ fs2);
    const u16x2 x1 = __builtin_bit_cast(u16x2, a.y ^ fs2);
    const u16x2 x2 = __builtin_bit_cast(u16x2, a.z ^ fs2);
    const u16x2 x3 = __builtin_bit_cast(u16x2, a.w ^ fs2);
    const u16x2 m = __builtin_elementwise_min(__builtin_elementwise_min(x0, x1), __builtin_elementwise_min(x2, x3));
    return (m.x == 0) | (m.y == 0);
}

constexpr int PROBE_BATCH = 8;     // records per thread whose loads are issued before any compare

// Bucket filter: a blocked Bloom filter of FB bits per bucket (4 <= FB <= 16) in 32-bit words. Each
// partition's slice starts on a word: partition p owns words [p * PW, (p + 1) * PW), PW =
// ceil(K * FB / 32), and its bucket k belongs to word p * PW + k * FB / 32 (rounded down), so a word
// serves about 32 / FB consecutive buckets of one partition, which share its 32 bits: each of their
// D records sets two, at its fingerprint bits 0-4 and 5-9. An I record whose two bits are not both set matches no D record of its bucket, so the
// probe rejects it without reading the table. At 2 D records per bucket about 16% of the records
// that miss pass at FB = 8 (22% with one bit each in 8-bit per-bucket fields), 14% at FB = 9 and
// 31% at FB = 5. A bucket with more records than slots has every record's bits set, and the
// records it passes meet the overflow flag of its slot 7.
__device__ __forceinline__ u32 filter_bit_mask(u32 fp) { return (1u << (fp & 31u)) | (1u << ((fp >> 5) & 31u)); }
CF_HD u64 filter_slice_words(u64 K, int FB) { return (K * (u64)FB + 31) >> 5; }
// the first bucket (within its partition) of word wl of a partition's slice
__device__ __forceinline__ u64 filter_bucket_lo(u64 wl, int FB) { return (wl * 32 + (u64)FB - 1) / (u64)FB; }

// The filter of a table bucketed by global atomics: one thread per word, over the records of its
// buckets (their fingerprints from DR).
__global__ void __launch_bounds__(PB) k_bucket_filter(const u64 *__restrict__ DR, const u64 *__restrict__ bstart,
                                                      u64 K, int FB, u32 *__restrict__ F, u64 nwords) {
    const u64 PW = filter_slice_words(K, FB);
    for (u64 w = (u64)blockIdx.x * PB + threadIdx.x; w < nwords; w += (u64)gridDim.x * PB) {
        const u64 part = w / PW, wl = w - part * PW;
        const u64 k1 = filter_bucket_lo(wl + 1, FB) < K ? filter_bucket_lo(wl + 1, FB) : K;
        u32 word = 0;
        for (u64 k = filter_bucket_lo(wl, FB); k < k1; ++k) {
            const u64 b = part * K + k;
            for (u64 q = bstart[b]; q < bstart[b + 1]; ++q) word |= filter_bit_mask((u32)(DR[q] >> DR_FP_SHIFT) & DR_FP_MASK);
        }
        F[w] = word;
    }
}

// Probe, pass 1: each record reads its bucket's 8 slots (16 bytes); a record whose fingerprint is
// among them, or whose bucket overflows (slot 7 flagged: more D records than slots), is a
// candidate. With a bucket filter (FB > 0) the block keeps its partition's slice of the filter in
// LDS and a record reads the table only if its filter bit is set.
// A tile is up to SPT consecutive segments of one partition (a partition's nblk segments are
// consecutive in the layout); each wave takes whole segments, so a segment's partition and start
// are uniform across the wave. Candidates are placed by the wave that finds them into slots it
// reserves CAND_CHUNK at a time from *ncand (slots it leaves unused hold CAND_NONE), within
// cand[0, ccap); *ncand counts every reserved slot, so the caller sees an overflow and repeats with
// room. Tiles come from a queue per XCD (xq[8], zero at launch)
// over that XCD's contiguous slice of tiles, so an XCD works through its partitions one after
// another and the table slice of the partition in hand stays in its L2; every block leaves once its
// XCD's queue is past the slice.
constexpr int PBP = 1024;                       // threads per probe block (16 waves)
constexpr u64 CAND_NONE = ~0ull;                // an unused candidate slot
constexpr u32 CAND_CHUNK = 256;                 // candidate slots a wave reserves at a time
constexpr int SPW = 2;                          // segments a wave probes together (one round)
constexpr int GPW = 2;                          // such groups per wave per tile
constexpr int SPT = GPW * SPW * (PBP / 64);     // segments per tile
constexpr int FWMAX = 38912;                    // filter words a block holds (152 KB of LDS)

__device__ __forceinline__ u64 seg_start(const SegLayout &S, u64 g) { return S.O ? S.O[g] : (g << S.capl); }

// Raw buffer descriptor over [base, base + bytes): loads at byte offsets past the end return zero.
// The third dword is the gfx9 descriptor's format word (32-bit data format).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, u64 bytes) {
    const int n = bytes > 0x7FFFFFFFull ? 0x7FFFFFFF : (int)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, n, 0x00020000);
}
__device__ __forceinline__ u64 uniform_u64(u64 v) {
    return ((u64)(u32)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
           (u64)(u32)__builtin_amdgcn_readfirstlane((int)(u32)v);
}

__global__ void __launch_bounds__(PBP) k_part_probe(KeyHash kh, SegLayout S, const u32 *__restrict__ recs,
                                                    const unsigned short *__restrict__ T,
                                                    const u32 *__restrict__ F, int FB, u32 *__restrict__ xq,
                                                    unsigned long long *__restrict__ ncand, u64 ccap,
                                                    u64 *__restrict__ cand, u32 tag) {
    __shared__ u32 sF[FWMAX];
    __shared__ u64 s_nx[2];
    const int lane = threadIdx.x & 63;
    const u32 wv = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int P = 1 << kh.lgP;
    const u64 tpp = (S.nblk + SPT - 1) / SPT;   // tiles per partition
    const u64 ntiles = tpp * (u64)P;
    const u32 xcd = blockIdx.x % NXCD;
    const u64 t_lo = ntiles * xcd / NXCD, t_hi = ntiles * (xcd + 1) / NXCD;
    const u32 fmask = (1u << kh.fpb) - 1u;
    const u64 K = kh_k(kh);
    const u32 rsz = S.hi ? 4u : 8u;   // bytes per record in the bucket-word stream
    u64 cur_part = ~0ull;
    // the wave's reserved candidate slots [cbase, cbase + cleft): one counter update per CAND_CHUNK
    // candidates instead of one per round that has any
    u64 cbase = 0;
    u32 cleft = 0;
    if (threadIdx.x == 0) s_nx[0] = t_lo + atomicAdd(&xq[xcd], 1u);
    lds_sync();
    u64 t = uniform_u64(s_nx[0]);
    for (u32 it = 1;; ++it) {
        if (t >= t_hi) break;
        // the next tile's index is taken while this one is probed and read after the barrier that
        // ends it; of two slots, the one written here was last read before the previous barrier
        if (threadIdx.x == 0) s_nx[it & 1] = t_lo + atomicAdd(&xq[xcd], 1u);
        const u64 part = t / tpp;
        if (FB > 0 && part != cur_part) {
            // this block's first tile of a partition: its filter slice into LDS (the barrier above
            // ordered every read of the previous slice before this)
            // (every load of the slice is issued before the first LDS store waits on one)
            const u32 nw = (u32)filter_slice_words(K, FB);
            const u64 w0 = part * nw;
            constexpr int FPT = 16;   // words in flight per thread (a batch of FPT * PBP words)
            for (u32 i0 = 0; i0 < nw; i0 += FPT * PBP) {
                u32 fw[FPT];
#pragma unroll
                for (int j = 0; j < FPT; ++j) {
                    const u32 i = i0 + threadIdx.x + (u32)j * PBP;
                    fw[j] = i < nw ? F[w0 + i] : 0u;
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < FPT; ++j) {
                    const u32 i = i0 + threadIdx.x + (u32)j * PBP;
                    if (i < nw) sF[i] = fw[j];
                }
            }
            cur_part = part;
            lds_sync();
        }
        const u64 gp = part * S.nblk;            // the partition's first segment
        const u64 gend = gp + S.nblk;
        // the partition's table slice: byte offset hoff * 16 of bucket hoff
        const __amdgpu_buffer_rsrc_t tr = buf_rsrc(T + part * K * TSLOTS, K * TSLOTS * 2);
        // fills and starts of the wave's GPW * SPW segments of this tile, one load each: lane l holds
        // those of segment l mod (GPW * SPW), read out as wave-uniform values (a descriptor held in
        // vector registers would put every record load in a loop over its distinct values)
        static_assert(GPW * SPW <= 64, "one lane per segment of the wave");
        u32 fv;
        u64 sv;
        {
            const int l = lane % (GPW * SPW);
            const u64 g = gp + (t - part * tpp) * SPT + ((u64)(l / SPW) * (PBP / 64) + wv) * SPW + (u64)(l % SPW);
            fv = g < gend ? S.fill[g] : 0u;
            sv = g < gend ? seg_start(S, g) : 0;
        }
        for (int grp = 0; grp < GPW; ++grp) {
            // a segment's records through a descriptor that ends at its fill
            u32 fills[SPW];
            u64 starts[SPW];
            __amdgpu_buffer_rsrc_t rr[SPW];
            u32 rounds_end = 0;
#pragma unroll
            for (int sj = 0; sj < SPW; ++sj) {
                fills[sj] = (u32)__builtin_amdgcn_readlane((int)fv, grp * SPW + sj);
                starts[sj] = lane_u64(sv, grp * SPW + sj);
                rr[sj] = buf_rsrc((const u8_t *)recs + starts[sj] * rsz, (u64)fills[sj] * rsz);
                rounds_end = fills[sj] > rounds_end ? fills[sj] : rounds_end;
            }
            // the wave's SPW segments are probed together, 64 * PROBE_BATCH slots of each per round
            // (one round for a capacity segment), so one round trip serves all of them
            for (u32 v0 = 0; v0 < rounds_end; v0 += 64 * PROBE_BATCH) {
                constexpr int NB2 = SPW * PROBE_BATCH;
                static_assert(NB2 <= 32, "one mask bit per record of a round");
                u32 lo[NB2];
                uint4 ta[NB2];
                u32 valid = 0;
#pragma unroll
                for (int e = 0; e < NB2; ++e) {
                    const int sj = e / PROBE_BATCH, q = e % PROBE_BATCH;
                    const u32 v = v0 + (u32)q * 64 + lane;
                    // past the fill the descriptor returns 0: bucket 0 of the partition, masked below;
                    // non-temporal (cache policy 2): the records stream through L2 once and should not
                    // evict the partition's table slice
                    lo[e] = (u32)__builtin_amdgcn_raw_buffer_load_b32(rr[sj], (int)(v * rsz), 0, 2);
                    valid |= (u32)(v < fills[sj]) << e;
                }
                u32 look = valid;
                if (FB > 0) {
                    // the bucket's word of the slice, two bits by the fingerprint (filter_bit_mask)
                    u32 pass = 0;
#pragma unroll
                    for (int e = 0; e < NB2; ++e) {
                        const u32 w = sF[((lo[e] >> kh.fpb) * (u32)FB) >> 5];
                        const u32 m = filter_bit_mask(lo[e] & fmask);
                        pass |= (u32)((w & m) == m) << e;
                    }
                    look &= pass;
                }
                // buckets; a record not looked up takes an offset past the slice (the load returns
                // zeros without touching memory and cannot match). Every load of the round is issued
                // before the first compare waits on one.
#pragma unroll
                for (int e = 0; e < NB2; ++e) {
                    const u32 off = (look >> e) & 1u ? (lo[e] >> kh.fpb) * (2u * TSLOTS) : 0x7FFFFFF0u;
                    const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(tr, (int)off, 0, 0);
                    ta[e] = make_uint4(x[0], x[1], x[2], x[3]);
                }
                __builtin_amdgcn_sched_barrier(0);
                // a fingerprint among the bucket's slots, or the overflow mark in slot 7, is a candidate
                u32 hmask = 0;
#pragma unroll
                for (int e = 0; e < NB2; ++e) {
                    const u32 fs2 = (0x8000u | (lo[e] & fmask)) * 0x10001u;
                    hmask |= (u32)(bucket_hit(ta[e], fs2) | ((ta[e].w >> 16) == T_OVERFLOW)) << e;
                }
                hmask &= look;
                if (__any(hmask != 0)) {
                    const u32 c = (u32)__popc(hmask);
                    const u32 inc = wave_incl_sum(c);
                    const u32 total = (u32)__builtin_amdgcn_readlane((int)inc, 63);
                    if (total > cleft) {
                        // the rest of the current reservation is marked unused; a new one is taken
                        for (u32 i = lane; i < cleft; i += 64)
                            if (cbase + i < ccap) cand[cbase + i] = CAND_NONE;
                        const u32 want = total > CAND_CHUNK ? total : CAND_CHUNK;
                        unsigned long long nb = 0;
                        if (lane == 0) nb = atomicAdd(ncand, (unsigned long long)want);
                        cbase = uniform_u64(__shfl(nb, 0, 64));
                        cleft = want;
                    }
                    u64 at = cbase + inc - c;
                    cbase += total;
                    cleft -= total;
                    while (hmask) {
                        const int e = __ffs(hmask) - 1;
                        hmask &= hmask - 1;
                        if (at < ccap)
                            cand[at] = (starts[e / PROBE_BATCH] + v0 + (u64)(e % PROBE_BATCH) * 64 + lane) | (u64)tag << 32;
                        ++at;
                    }
                }
            }
        }
        lds_sync();
        t = uniform_u64(s_nx[it & 1]);
    }
    for (u32 i = lane; i < cleft; i += 64)
        if (cbase + i < ccap) cand[cbase + i] = CAND_NONE;
}

// ---- D side, partitioned: the D records were scattered by partition like the I side ({bucket word,
// leaf}, 8 bytes). A partition's K buckets are cut into windows of WIN = 2^wb buckets
// (about 2048 D records each); two more passes bucket them with no per-record global atomic:
//   window scatter: each block ranks a tile of its partition's records by window in LDS and
//          reserves room per window with one atomic per (tile, window) in a fixed-capacity region
//          of CAPW records per window, then writes one run per window; a window whose records do
//          not fit sets *ovf (heavily repeated keys: the caller then buckets this group with
//          per-record global atomics instead)
//   window sort: one block per window loads its records into LDS, counting-sorts them by bucket,
//          and writes the window's bucket starts, its records in bucket order ({key, leaf}), its
//          table slots and its filter words, all contiguous
// The window starts in DR are the exclusive scan of the window counts (windows are contiguous
// bucket ranges in partition order, so that order is bucket order).
constexpr int CAPW = 3072;             // records a window holds
constexpr int WPT = 4;                 // D records of a window-scatter tile: 4 segments, one per wave

struct DWin {
    int wb;        // log2 buckets per window
    u32 R;         // windows per partition
};

__global__ void __launch_bounds__(PB) k_dwin_scatter(KeyHash kh, DWin dw, SegLayout S, const uint2 *__restrict__ recs,
                                                     u32 *__restrict__ wcount, uint2 *__restrict__ wreg,
                                                     u32 *__restrict__ ovf) {
    __shared__ u32 hist[256];
    __shared__ u32 loc[256];
    __shared__ u32 woff[256];
    __shared__ u32 s_w[PB / 64];
    __shared__ uint2 s_r[WPT * 512];
    __shared__ u8_t s_win[WPT * 512];
    __shared__ u32 s_ovf;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u64 tpp = (S.nblk + WPT - 1) / WPT;
    const u64 part = blockIdx.x / tpp;
    const u64 g = part * S.nblk + (blockIdx.x - part * tpp) * WPT + wv;
    const bool has = g < (part + 1) * S.nblk;
    const u32 fill = has ? S.fill[g] : 0u;
    const u64 st = has ? seg_start(S, g) : 0;
    if (threadIdx.x < 256) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_ovf = 0;
    lds_sync();
    // a segment holds at most 512 records (capacity layout) -- the exact layout's larger segments
    // are taken 512 at a time by the loop below
    for (u32 c0 = 0;; c0 += 512) {
        const u32 cn = c0 < fill ? (fill - c0 < 512u ? fill - c0 : 512u) : 0u;
        uint2 r[8];
        u32 rank[8], win[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u32 v = (u32)k * 64 + lane;
            if (v < cn) {
                r[k] = recs[st + c0 + v];
                win[k] = (r[k].x >> kh.fpb) >> dw.wb;
                rank[k] = atomicAdd(&hist[win[k]], 1u);
            }
        }
        lds_sync();
        // tile offsets of the windows (block scan of hist), and one reservation per window present
        const u32 v = threadIdx.x < 256 ? hist[threadIdx.x] : 0u;
        {
            const u32 inc = wave_incl_sum(v);
            if (lane == 63) s_w[wv] = inc;
            lds_sync();
            u32 pre = 0;
            for (int q = 0; q < wv; ++q) pre += s_w[q];
            if (threadIdx.x < 256) {
                loc[threadIdx.x] = pre + inc - v;
                woff[threadIdx.x] = v ? atomicAdd(&wcount[part * dw.R + threadIdx.x], v) : 0u;
            }
        }
        lds_sync();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u32 vv = (u32)k * 64 + lane;
            if (vv < cn) {
                const u32 at = loc[win[k]] + rank[k] + 0u;
                s_r[at] = r[k];
                s_win[at] = (u8_t)win[k];
            }
        }
        lds_sync();
        u32 tot = 0;
        for (int q = 0; q < PB / 64; ++q) tot += s_w[q];
        for (u32 q = threadIdx.x; q < tot; q += PB) {
            const u32 w = s_win[q];
            const u32 o = woff[w] + (q - loc[w]);
            if (o >= (u32)CAPW) {
                s_ovf = 1;
                continue;
            }
            wreg[((u64)part * dw.R + w) * CAPW + o] = s_r[q];
        }
        lds_sync();
        if (threadIdx.x < 256) hist[threadIdx.x] = 0;
        // every wave of the block runs the same number of rounds: the largest of its segments
        const u32 more = __syncthreads_or(c0 + 512 < fill);
        if (!more) break;
    }
    if (threadIdx.x == 0 && s_ovf) atomicOr(ovf, 1u);
}

constexpr int DSB = 512;               // threads of a window-sort block
__global__ void __launch_bounds__(DSB) k_dwin_sort(KeyHash kh, DWin dw, const u32 *__restrict__ wcount,
                                                   const u64 *__restrict__ wbase, const uint2 *__restrict__ wreg,
                                                   u64 *__restrict__ bstart, u64 *__restrict__ DR,
                                                   unsigned short *__restrict__ T, u32 *__restrict__ F, int FB,
                                                   u32 *__restrict__ kerr) {
    constexpr int WMAXB = 4096;        // buckets a window may have
    __shared__ uint2 s_r[CAPW];
    __shared__ unsigned short s_idx[CAPW];
    __shared__ u32 pos[WMAXB + 1];     // per-bucket counts, then (scanned in place) bucket starts
    __shared__ u32 s_w[DSB / 64];
    // the window's filter words, built from the records as they are counted (up to FWW words: a
    // window of 1024 buckets at up to 15 bits each; a larger one takes the per-word loop below)
    constexpr u32 FWW = 512;
    __shared__ u32 s_f[FWW];
    u32 *cnt = pos;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const u64 pw = blockIdx.x;
    const u64 part = pw / dw.R, w = pw - part * dw.R;
    const u64 K = kh_k(kh);
    const u32 WIN = 1u << dw.wb;
    const u64 b_lo = w * WIN;                                  // first bucket, within the partition
    const u32 nb = (u32)((b_lo + WIN <= K ? b_lo + WIN : K) - b_lo);
    const u32 n = wcount[pw] < (u32)CAPW ? wcount[pw] : (u32)CAPW;
    const u64 base = wbase[pw];
    const u64 gb = part * K + b_lo;                            // first bucket, global
    // the window's filter words: from fw0 (its first bit fb0 bits into it), nfw of them
    const u64 fw0 = FB > 0 ? part * filter_slice_words(K, FB) + ((b_lo * (u64)FB) >> 5) : 0;
    const u32 fb0 = (u32)((b_lo * (u64)FB) & 31);
    const u32 nfw = FB > 0 ? (fb0 + nb * (u32)FB + 31) >> 5 : 0;
    const bool fl = FB > 0 && nfw <= FWW;
    const u32 fmask = (1u << kh.fpb) - 1u;
    for (u32 i = threadIdx.x; i <= nb; i += DSB) cnt[i] = 0;
    if (fl)
        for (u32 i = threadIdx.x; i < nfw; i += DSB) s_f[i] = 0;
    for (u32 i = threadIdx.x; i < n; i += DSB) s_r[i] = wreg[pw * CAPW + i];
    lds_sync();
    constexpr int RPT = (CAPW + DSB - 1) / DSB;
    u32 rk[RPT];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const u32 i = threadIdx.x + (u32)k * DSB;
        if (i < n) {
            // a record's window is its bucket offset >> wb, so its bucket is in the window; a record
            // that is not (inconsistent input) is reported through kerr and the window left unsorted
            const u32 x = s_r[i].x;
            const u32 bi = (x >> kh.fpb) - (u32)b_lo;
            if (bi < nb) {
                rk[k] = atomicAdd(&cnt[bi], 1u);
                if (fl) atomicOr(&s_f[(fb0 + bi * (u32)FB) >> 5], filter_bit_mask(x & fmask));
            } else {
                bad = true;
            }
        }
    }
    if (__syncthreads_or(bad)) {
        // the window's buckets are written empty (no table slot, no filter bit, every bucket
        // starting at the window's base), so no later read follows an unwritten start or record
        for (u32 i = threadIdx.x; i < nb; i += DSB) {
            bstart[gb + i] = base;
            *(uint4 *)(T + (gb + i) * TSLOTS) = make_uint4(0u, 0u, 0u, 0u);
        }
        if (pw + 1 == gridDim.x && threadIdx.x == 0) bstart[gb + nb] = base + n;
        if (threadIdx.x == 0) atomicOr(kerr, KERR_DWINDOW);
        return;
    }
    lds_sync();
    // exclusive scan of cnt[0, nb) into pos (DSB threads, consecutive runs of PER entries)
    constexpr int PER = WMAXB / DSB;
    u32 sum = 0, cv[PER];
    const u32 b0 = threadIdx.x * PER;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        cv[k] = b0 + k < nb ? cnt[b0 + k] : 0u;
        sum += cv[k];
    }
    const u32 inc = wave_incl_sum(sum);
    if (lane == 63) s_w[wv] = inc;
    lds_sync();
    u32 run = inc - sum;
    for (int q = 0; q < wv; ++q) run += s_w[q];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const u32 i = b0 + k;
        if (i < nb) {
            pos[i] = run;
            bstart[gb + i] = base + run;
            run += cv[k];
        }
    }
    if (threadIdx.x == DSB - 1) pos[nb] = run;
    // the last window of the last partition closes bstart
    if (pw + 1 == gridDim.x && threadIdx.x == DSB - 1) bstart[gb + nb] = base + run;
    lds_sync();
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const u32 i = threadIdx.x + (u32)k * DSB;
        if (i < n) s_idx[pos[(s_r[i].x >> kh.fpb) - (u32)b_lo] + rk[k]] = (unsigned short)i;
    }
    lds_sync();
    // records (D leaves, fingerprints and hash_ex) in bucket order
    const u32 fpm = (1u << kh.fpb) - 1u;
    for (u32 k = threadIdx.x; k < n; k += DSB) {
        const uint2 r = s_r[s_idx[k]];
        DR[base + k] = (u64)(r.y & ((1u << DREC_EX_SHIFT) - 1u)) | (u64)(r.x & fpm) << DR_FP_SHIFT |
                       (u64)(r.y >> DREC_EX_SHIFT) << DR_EX_SHIFT;
    }
    // table slots: one bucket per thread, 16 bytes each
    for (u32 i = threadIdx.x; i < nb; i += DSB) {
        const u32 c = pos[i + 1] - pos[i];
        u32 sl[TSLOTS];
#pragma unroll
        for (int q = 0; q < TSLOTS; ++q)
            sl[q] = (u32)q < c ? (0x8000u | (s_r[s_idx[pos[i] + q]].x & fmask)) : 0u;
        if (c > (u32)TSLOTS) sl[TSLOTS - 1] = T_OVERFLOW;
        static_assert(TSLOTS == 8, "a bucket is one 16-byte store");
        *(uint4 *)(T + (gb + i) * TSLOTS) =
            make_uint4(sl[0] | (sl[1] << 16), sl[2] | (sl[3] << 16), sl[4] | (sl[5] << 16), sl[6] | (sl[7] << 16));
    }
    // filter words of the window's buckets; a word shared with a neighbouring window is merged
    // atomically
    if (fl) {
        const bool open_end = ((fb0 + nb * (u32)FB) & 31u) != 0;   // the last word may hold the next window's bits
        for (u32 j = threadIdx.x; j < nfw; j += DSB) {
            const bool edge = (j == 0 && fb0 != 0) || (j + 1 == nfw && open_end);
            if (edge) atomicOr(&F[fw0 + j], s_f[j]);
            else F[fw0 + j] = s_f[j];
        }
    } else if (FB > 0) {
        for (u32 j = threadIdx.x; j < nfw; j += DSB) {
            const u64 wd = fw0 + j;
            // the word's buckets as offsets from the window's first, in 32-bit arithmetic: its first
            // bit is r0 bits past the window's first (-32 < r0 < nb * FB), so its buckets are
            // [ceil(r0 / FB), ceil((r0 + 32) / FB)), cut to the window
            const int r0 = (int)(j * 32) - (int)fb0;
            const u32 a = r0 <= 0 ? 0u : ((u32)r0 + (u32)FB - 1u) / (u32)FB;
            const u32 z0 = ((u32)(r0 + 32) + (u32)FB - 1u) / (u32)FB;
            const u32 z = z0 < nb ? z0 : nb;
            u32 word = 0;
            for (u32 i = a; i < z; ++i)
                for (u32 e = pos[i]; e < pos[i + 1]; ++e) word |= filter_bit_mask(s_r[s_idx[e]].x & fmask);
            // a word that may hold another window's buckets is merged atomically
            const bool edge = r0 < 0 || z0 > nb;
            if (edge) atomicOr(&F[wd], word);
            else F[wd] = word;
        }
    }
}

// Probe, pass 2, over the candidates of every table a chunk was probed against (a candidate's tag
// names its table): the D records of each candidate's bucket that agree with the candidate on the
// fingerprint and hash_ex have their exact keys recomputed from their leaves; the candidate's own
// key is recomputed once, at the first record with its fingerprint. Every equal key emits a pair.
__global__ void __launch_bounds__(PB) k_probe_candidates(LeafSrc L, LeafSrc LD0, KeyHash kh,
                                                         const u32 *__restrict__ recs,
                                                         SegLayout S,
                                                         const unsigned long long *__restrict__ ncand, u64 ccap,
                                                         const u64 *__restrict__ cand, const u64 *__restrict__ koffs,
                                                         int ninst, const DTabDev *__restrict__ tabs,
                                                         unsigned long long *__restrict__ npairs, u64 cap,
                                                         u64 *__restrict__ out_d, u64 *__restrict__ out_i,
                                                         u64 *__restrict__ out_inst, u32 *__restrict__ kerr) {
    const u64 nc = *ncand < ccap ? *ncand : ccap;
    // candidates past the buffer were never stored: report it rather than lose their pairs
    if (blockIdx.x == 0 && threadIdx.x == 0 && *ncand > ccap) atomicOr(kerr, KERR_CAND);
    for (u64 c = (u64)blockIdx.x * PB + threadIdx.x; c < nc; c += (u64)gridDim.x * PB) {
        const u64 raw = cand[c];
        if (raw == CAND_NONE) continue;
        const u64 e = raw & 0xFFFFFFFFull;
        const DTabDev tb = tabs[raw >> 32];
        const u64 *__restrict__ bstart = tb.bstart;
        const u64 *__restrict__ DR = tb.DR;
        LeafSrc LD = LD0;
        LD.lo = tb.lo;
        LD.n = tb.n;
        LD.tp = tb.tp;
        u64 g;
        if (S.O) {   // the last segment starting at or before e (it is not empty)
            u64 a = 0, bb = S.nseg - 1;
            while (a < bb) {
                const u64 mid = (a + bb + 1) >> 1;
                if (S.O[mid] <= e) a = mid; else bb = mid - 1;
            }
            g = a;
        } else {
            g = e >> S.capl;
        }
        const uint2 r = S.hi ? make_uint2(recs[e], S.hi[e]) : ((const uint2 *)recs)[e];
        const u64 b = (g / S.nblk) * kh_k(kh) + (r.x >> kh.fpb);
        const u64 fp = r.x & ((1u << kh.fpb) - 1u);
        u64 key = 0, iid = 0;
        u32 ex = 0;
        bool have = false;
        const u64 lo = bstart[b], hi = bstart[b + 1];
        for (u64 q = lo; q < hi; ++q) {
            const u64 dr = DR[q];
            if (((dr >> DR_FP_SHIFT) & DR_FP_MASK) != fp) continue;
            if (!have) {
                u64 h;
                leaf_exact(L, r.y, key, iid, &h);
                ex = hash_ex(h);
                have = true;
            }
            if ((u32)(dr >> DR_EX_SHIFT) != ex) continue;
            u64 dkey, did;
            leaf_exact(LD, dr & DR_LEAF_MASK, dkey, did);
            if (dkey != key) continue;
            const unsigned long long at = atomicAdd(npairs, 1ULL);
            if (at < cap) {
                out_d[at] = did;
                out_i[at] = iid;
                out_inst[at] = inst_of_key(koffs, ninst, key);
            }
        }
    }
}

// Per-instance record totals from node boundaries: nodes of one instance are contiguous and in
// instance order, so instance i's nodes are [lo, hi) with lo and hi found by binary search on the
// instance field, and its leaves are off[hi] - off[lo]. One thread per instance.
__global__ void __launch_bounds__(PB) k_inst_bounds(const LeafNode *__restrict__ nodes, const u64 *__restrict__ off,
                                                    size_t nn, int ninst, u64 *__restrict__ out) {
    const int i = blockIdx.x * PB + threadIdx.x;
    if (i >= ninst) return;
    size_t bounds[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {   // first node whose instance is >= i + e
        size_t a = 0, b = nn;
        while (a < b) {
            const size_t mid = (a + b) >> 1;
            if ((int)(nodes[mid].meta >> 8) < i + e) a = mid + 1; else b = mid;
        }
        bounds[e] = a;
    }
    if (bounds[1] > bounds[0]) out[i] += off[bounds[1]] - off[bounds[0]];
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
static LeafSrc leaf_src(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                        int depth, DBuf<u64> &tp) {
    const u64 ntiles = (n + LT - 1) / LT;
    tp = DBuf<u64>(ctx, 2 * ntiles);
    CF_LAUNCH(ctx, "k_tile_parents", k_tile_parents, dim3(grid_for(2 * ntiles, PB, 65536)), dim3(PB), 0, off, nn, lo, n,
              tp.p);
    return LeafSrc{descs, nodes, nn, off, lo, n, depth, tp.p};
}

void leaf_records(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                  int depth, u64 *keys, u64 *ids) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    const u64 ntiles = (n + LT - 1) / LT;
    const u64 nblk = ntiles < 1024 ? ntiles : 1024, tpb = (ntiles + nblk - 1) / nblk;
    CF_LAUNCH(ctx, "k_leaf_records", k_leaf_records, dim3((unsigned)((ntiles + tpb - 1) / tpb)), dim3(LPB), 0, L, tpb,
              keys, ids);
}

void part_bucket_count(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                       int depth, const KeyHash &kh, u32 *bcount) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    CF_LAUNCH(ctx, "k_bucket_count", k_bucket_count, dim3(grid_for(n, LT)), dim3(LPB), 0, L, kh, bcount);
}

void part_bucket_scatter(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo,
                         u64 n, int depth, const KeyHash &kh, const u64 *bstart, u32 *bcur, void *DR, unsigned short *T) {
    if (!n) return;
    DBuf<u64> tp;
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    CF_LAUNCH(ctx, "k_bucket_scatter", k_bucket_scatter, dim3(grid_for(n, LT)), dim3(LPB), 0, L, kh, bstart, bcur,
              (u64 *)DR, T);
}

// D side, partitioned (after part_scatter_cap / part_scatter_exact of the D side): window geometry,
// window scatter and window sort.
bool dwin_geometry(const KeyHash &kh, u64 nD, int &wb, u32 &R) {
    // about 2048 D records per window at the group's records per bucket
    const u64 NB = kh_nb(kh);
    const double dpb = NB ? (double)nD / (double)NB : 1.0;
    wb = 12;
    while (wb > 4 && (double)((u64)1 << wb) * dpb > 2048.0) --wb;
    const u64 K = kh_k(kh);
    const u64 r = (K + ((u64)1 << wb) - 1) >> wb;
    R = (u32)r;
    return r <= 256;
}

void part_d_windows(cf_ctx *ctx, const KeyHash &kh, int wb, u32 R, const SegLayout &S, const u64 *rec,
                    u32 *wcount, void *wreg, u32 *ovf) {
    const int P = 1 << kh.lgP;
    fill_u32(ctx, wcount, (size_t)P * R, 0);
    fill_u32(ctx, ovf, 1, 0);
    if (!S.nseg) return;
    const u64 tpp = (S.nblk + WPT - 1) / WPT;
    CF_LAUNCH(ctx, "k_dwin_scatter", k_dwin_scatter, dim3((unsigned)(tpp * P)), dim3(PB), 0, kh, DWin{wb, R}, S,
              (const uint2 *)rec, wcount, (uint2 *)wreg, ovf);
}

void part_d_sort(cf_ctx *ctx, const KeyHash &kh, int wb, u32 R, const u32 *wcount, const u64 *wbase,
                 const void *wreg, u64 *bstart, void *DR, unsigned short *T, u32 *F, int FB, u32 *kerr) {
    const int P = 1 << kh.lgP;
    CF_LAUNCH(ctx, "k_dwin_sort", k_dwin_sort, dim3((unsigned)((u64)P * R)), dim3(DSB), 0, kh, DWin{wb, R}, wcount,
              wbase, (const uint2 *)wreg, bstart, (u64 *)DR, T, F, FB, kerr);
}

u64 dwin_capacity() { return CAPW; }

// The partition scatter for the layout's side and record form, and for the geometry's partition count
// (the 512-partition kernels only where the geometry has more than 256).
static void launch_part_scatter(cf_ctx *ctx, const LeafSrc &L, const KeyHash &kh, u64 tpb, u32 nblk, const u64 *O,
                                int capl, u32 *fill, u32 *ovf, u64 *rec, u32 *hi, bool dside) {
    const bool wide = kh.lgP > 8;
    if (dside && wide)
        CF_LAUNCH(ctx, "k_part_scatter_d", k_part_scatter_d_p512, dim3(nblk), dim3(LPB), 0, L, kh, tpb, nblk, O, capl,
                  fill, ovf, (uint2 *)rec);
    else if (dside)
        CF_LAUNCH(ctx, "k_part_scatter_d", k_part_scatter_d, dim3(nblk), dim3(LPB), 0, L, kh, tpb, nblk, O, capl, fill,
                  ovf, (uint2 *)rec);
    else if (hi && wide)
        CF_LAUNCH(ctx, "k_part_scatter_split", k_part_scatter_split_p512, dim3(nblk), dim3(LPB), 0, L, kh, tpb, nblk, O,
                  capl, fill, ovf, (uint2 *)rec, hi);
    else if (hi)
        CF_LAUNCH(ctx, "k_part_scatter_split", k_part_scatter_split, dim3(nblk), dim3(LPB), 0, L, kh, tpb, nblk, O, capl,
                  fill, ovf, (uint2 *)rec, hi);
    else if (wide)
        CF_LAUNCH(ctx, "k_part_scatter", k_part_scatter_p512, dim3(nblk), dim3(LPB), 0, L, kh, tpb, nblk, O, capl, fill,
                  ovf, (uint2 *)rec);
    else
        CF_LAUNCH(ctx, "k_part_scatter", k_part_scatter, dim3(nblk), dim3(LPB), 0, L, kh, tpb, nblk, O, capl, fill, ovf,
                  (uint2 *)rec);
}

// Capacity layout of a chunk of n I leaves: blocks of tpb tiles whose expected share per partition
// is a set fraction of a 2^9-slot segment (default 70%: mean ~350, segment ~ mean + 8 sd).
constexpr int CAPL = 9;
static void cap_geometry(u64 n, const KeyHash &kh, const Tuning &tn, u64 &tpb, u32 &nblk) {
    const u64 ntiles = (n + LT - 1) / LT;
    const u64 pct = tn.seg_fill_pct > 0 && tn.seg_fill_pct <= 100 ? (u64)tn.seg_fill_pct : 70;
    tpb = (pct << CAPL << kh.lgP) / (100 * (u64)LT);
    if (tpb < 1) tpb = 1;
    nblk = (u32)((ntiles + tpb - 1) / tpb);
}

static u64 cap_slots(u64 n, const KeyHash &kh, const Tuning &tn) {
    u64 tpb;
    u32 nblk;
    cap_geometry(n, kh, tn, tpb, nblk);
    return ((u64)nblk << kh.lgP) << CAPL;
}
// slots are 32-bit in the scatter: a layout of 2^32 slots or more is not taken
static bool cap_fits(u64 n, const KeyHash &kh, const Tuning &tn) { return cap_slots(n, kh, tn) < ((u64)1 << 32); }

u64 part_cap_slots(u64 n, const KeyHash &kh, const Tuning &tn) { return cap_fits(n, kh, tn) ? cap_slots(n, kh, tn) : n; }

void part_scatter_cap(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                      int depth, const KeyHash &kh, const Tuning &tn, IChunkLayout &lay, u64 *rec, u32 *ovf,
                      DBuf<u64> &tp, bool dside, u64 hi_at) {
    CF_REQUIRE(!(dside && hi_at), "D records are {word, leaf} pairs");
    lay.hi = hi_at ? (u32 *)rec + hi_at : nullptr;
    if (!cap_fits(n, kh, tn)) {
        // too many segments for 32-bit slots (a chunk near 2^32 leaves at a low fill): the exact
        // layout directly, never overflowing
        fill_u32(ctx, ovf, 1, 0);
        (void)leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
        part_scatter_exact(ctx, descs, nodes, nn, off, lo, n, depth, kh, lay, rec, tp.p, dside, hi_at);
        return;
    }
    u64 tpb;
    u32 nblk;
    cap_geometry(n, kh, tn, tpb, nblk);
    lay.O.reset();
    lay.nblk = nblk;
    lay.nseg = (u64)nblk << kh.lgP;
    lay.capl = CAPL;
    lay.fill = DBuf<u32>(ctx, lay.nseg);
    fill_u32(ctx, ovf, 1, 0);
    LeafSrc L = leaf_src(ctx, descs, nodes, nn, off, lo, n, depth, tp);
    launch_part_scatter(ctx, L, kh, tpb, nblk, nullptr, CAPL, lay.fill.p, ovf, rec, lay.hi, dside);
}

void part_scatter_exact(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo,
                        u64 n, int depth, const KeyHash &kh, IChunkLayout &lay, u64 *rec, const u64 *tp,
                        bool dside, u64 hi_at) {
    CF_REQUIRE(!(dside && hi_at), "D records are {word, leaf} pairs");
    CF_REQUIRE(!hi_at || hi_at >= n, "the split records need room for every record");
    lay.hi = hi_at ? (u32 *)rec + hi_at : nullptr;
    const int P = 1 << kh.lgP;
    LeafSrc L{descs, nodes, nn, off, lo, n, depth, tp};
    const u64 ntiles = (n + LT - 1) / LT;
    // about 8 resident blocks per CU over 256 CUs, each with a long contiguous leaf range
    const u64 tpb = (ntiles + 16383) / 16384;
    const u32 nblk = (u32)((ntiles + tpb - 1) / tpb);
    lay.nblk = nblk;
    lay.nseg = (u64)P * nblk;
    lay.capl = 0;
    lay.fill = DBuf<u32>(ctx, lay.nseg);
    lay.O = DBuf<u64>(ctx, lay.nseg + 1);
    CF_LAUNCH(ctx, "k_part_hist", k_part_hist, dim3(nblk), dim3(LPB), 0, L, kh.lgP, tpb, nblk, lay.fill.p);
    scan_u32_to_u64_async(ctx, lay.fill.p, lay.O.p, lay.nseg);
    launch_part_scatter(ctx, L, kh, tpb, nblk, lay.O.p, 0, nullptr, nullptr, rec, lay.hi, dside);
}

u64 cand_margin(const Tuning &tn) {
    // unused slots the probe may leave: a partial reservation per wave of the grid
    const u64 g = tn.probe_blocks > 0 ? (u64)tn.probe_blocks : 256;
    return g * (PBP / 64) * CAND_CHUNK;
}

int filter_bits(const KeyHash &kh) {
    // the widest filter whose partition slice fits in LDS
    const u64 K = kh_k(kh);
    for (int fb = 16; fb >= 4; --fb)
        if (filter_slice_words(K, fb) <= (u64)FWMAX) return fb;
    return 0;
}

u64 filter_words(const KeyHash &kh, int FB) { return FB ? ((u64)1 << kh.lgP) * filter_slice_words(kh_k(kh), FB) : 0; }

void part_filter(cf_ctx *ctx, const u64 *DR, const u64 *bstart, const KeyHash &kh, int FB, u32 *F) {
    const u64 nw = filter_words(kh, FB);
    if (!nw) return;
    CF_LAUNCH(ctx, "k_bucket_filter", k_bucket_filter, dim3(grid_for(nw, PB, 65536)), dim3(PB), 0, DR, bstart,
              kh_k(kh), FB, F, nw);
}

void part_probe_pass1(cf_ctx *ctx, const KeyHash &kh, const SegLayout &S, const u64 *rec, const unsigned short *T,
                      const u32 *F, int FB, u32 *xq, unsigned long long *ncand, u64 ccap, u64 *cand,
                      const Tuning &tn, u32 tag, bool reset) {
    fill_u32(ctx, xq, NXCD, 0);
    if (reset) fill_u64(ctx, (u64 *)ncand, 1, 0);
    if (!S.nseg) return;
    // tiles are dealt per XCD, so every XCD needs at least one block: the grid is a non-zero
    // multiple of the XCD count; by default one 1024-thread block per CU (the filter slice takes
    // most of a CU's LDS)
    const u64 ntiles = (u64)((S.nblk + SPT - 1) / SPT) << kh.lgP;
    const u64 gmax = tn.probe_blocks > 0 ? (u64)tn.probe_blocks : 256;
    unsigned g = (unsigned)(ntiles < gmax ? ntiles : gmax);
    g = g < NXCD ? NXCD : g - g % NXCD;
    CF_LAUNCH(ctx, "k_part_probe", k_part_probe, dim3(g), dim3(PBP), 0, kh, S, (const u32 *)rec, T, F, FB, xq,
              ncand, ccap, cand, tag);
}

void part_probe_pass2(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo,
                      int depth, const KeyHash &kh, const SegLayout &S, const u64 *rec, const u64 *tp,
                      const DLeaves &dl, const DTabDev *tabs, const unsigned long long *ncand, u64 ccap,
                      const u64 *cand, const u64 *koffs, int ninst, unsigned long long *npairs, u64 cap, u64 *od,
                      u64 *oi, u64 *oinst, u32 *kerr) {
    LeafSrc L{descs, nodes, nn, off, lo, 0, depth, tp};
    LeafSrc LD{dl.descs, dl.nodes, dl.nn, dl.off, 0, 0, depth, nullptr};
    CF_LAUNCH(ctx, "k_probe_candidates", k_probe_candidates, dim3(1024), dim3(PB), 0, L, LD, kh, (const u32 *)rec, S,
              ncand, ccap, cand, koffs, ninst, tabs, npairs, cap, od, oi, oinst, kerr);
}

// Table descriptors into device memory, up to 32 per launch (as kernel arguments).
struct DTabBatch {
    DTabDev t[32];
    u32 n;
};
__global__ void __launch_bounds__(64) k_set_dtabs(DTabBatch b, DTabDev *__restrict__ out) {
    if (threadIdx.x < b.n) out[threadIdx.x] = b.t[threadIdx.x];
}
void set_dtabs(cf_ctx *ctx, const DTabDev *h, size_t n, DTabDev *d) {
    for (size_t i = 0; i < n; i += 32) {
        DTabBatch b{};
        b.n = (u32)(n - i < 32 ? n - i : 32);
        for (u32 k = 0; k < b.n; ++k) b.t[k] = h[i + k];
        CF_LAUNCH(ctx, "k_set_dtabs", k_set_dtabs, dim3(1), dim3(64), 0, b, d + i);
    }
}

void inst_bounds(cf_ctx *ctx, const LeafNode *nodes, const u64 *off, size_t nn, int ninst, u64 *out) {
    if (!nn || ninst <= 0) return;
    CF_LAUNCH(ctx, "k_inst_bounds", k_inst_bounds, dim3((unsigned)((ninst + PB - 1) / PB)), dim3(PB), 0, nodes, off, nn,
              ninst, out);
}
