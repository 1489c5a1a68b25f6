// k_part.hip -- radix-partitioned residue join on gfx950.
//
// Keys of one solve live in one key space (instance j owns [koff_j, koff_j + M_j)). Instance j
// owns the fine buckets [bbase_j, bbase_j + nb_j), nb_j near half its D count, reached by
// multiplicative hashing  b = bbase_j + mulhi(key * 0x9E3779B97F4A7C15, nb_j). Hashing, not key
// order, because the D keys are far from uniform: for small r the products prod(D) stay below M
// and are their own residues, all in a sliver of [0, M). The top bits of b give P <= 128 coarse
// partitions. Sizing the buckets per instance keeps them at a few records however uneven the
// moduli and sides of a batch are.
//
//  D side: records are counted per fine bucket and scattered into bucket order (a counting sort
//          on b: no full key sort is needed, a probe scans its bucket, 2-4 records on average).
//  I side: leaves are counted per partition (LDS histogram), then scattered by partition: each
//          block ranks its 2048 leaves by partition in LDS and writes one contiguous run per
//          partition (16 leaves on average with 128 partitions: a 128-byte run per array).
//  Probe:  the partitioned I records are read once, in partition order, with an XCD-aware tile
//          order (each of the 8 XCDs walks its own contiguous slice of tiles) so the D buckets
//          of the partitions in flight stay in that XCD's L2. Every equal key emits a pair.
// Algorithmic HBM traffic per I record: 16 B written by the scatter + 16 B read by the probe.
// Join semantics: every equal-key (D, I) pair (gpu/mitm_gpu2.py:107-119).
#include "tree_dev.h"

namespace {

constexpr int PB = 256;            // threads per block
constexpr int PI = 16;             // leaves per thread (count and probe passes)
constexpr int PT = PB * PI;        // 4096 leaves per tile
constexpr int SI = 8;              // leaves per thread in the scatter (bounds its registers and LDS)
constexpr int ST = PB * SI;        // 2048 leaves per scatter tile
constexpr int NXCD = 8;            // MI355X: 8 accelerator complex dies, blocks dealt round-robin

struct LeafSrc {
    const SideDesc *descs;
    const Node *nodes;
    size_t nn;
    const u64 *off;   // exclusive scan of leaves per node, nn + 1 entries
    u64 lo, n;        // leaves [lo, lo + n)
    int depth;        // r - 1
};

__device__ __forceinline__ u64 bucket_of(const SideDesc &s, u64 key) {
    return s.bbase + cf_mulhi(key * 0x9E3779B97F4A7C15ULL, s.nb);
}

__device__ __forceinline__ void leaf_eval(const LeafSrc &L, size_t p, u64 j, u64 &key, u64 &id, u32 &inst) {
    const Node nd = L.nodes[p];
    inst = nd.meta >> 8;
    const SideDesc &s = L.descs[inst];
    const int x = child_index(s, nd, (int)(j - L.off[p]));
    const u64 xm = cf_mulmod_m(nd.xm, s.pR[x], s.mc);
    const u64 x2 = nd.x2 * s.p[x];
    key = cf_key(xm, x2, s.mc) + s.koff;
    id = id_update(s, nd.id, x, L.depth);
}

// Parents of the first and last leaf of a tile, shared through LDS.
__device__ __forceinline__ void tile_parents(const LeafSrc &L, u64 j0, u64 j1, size_t *p0, size_t *p1) {
    __shared__ size_t s_p[2];
    if (threadIdx.x == 0) s_p[0] = find_parent(L.off, 0, L.nn - 1, j0);
    if (threadIdx.x == 1) s_p[1] = find_parent(L.off, 0, L.nn - 1, j1 - 1);
    __syncthreads();
    *p0 = s_p[0];
    *p1 = s_p[1];
    __syncthreads();
}

// ---- D side: counting sort by fine bucket ---------------------------------------------------
__global__ void __launch_bounds__(PB) k_bucket_count(LeafSrc L, u32 *__restrict__ bcount) {
    const size_t ntiles = (L.n + PB - 1) / PB;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * PB, jend = L.lo + L.n;
        const u64 j1 = j0 + PB < jend ? j0 + PB : jend;
        size_t p0, p1;
        tile_parents(L, j0, j1, &p0, &p1);
        const u64 j = j0 + threadIdx.x;
        if (j < j1) {
            u64 key, id;
            u32 inst;
            leaf_eval(L, find_parent(L.off, p0, p1, j), j, key, id, inst);
            atomicAdd(&bcount[bucket_of(L.descs[inst], key)], 1u);
        }
    }
}

__global__ void __launch_bounds__(PB) k_bucket_scatter(LeafSrc L, const u64 *__restrict__ bstart,
                                                       u32 *__restrict__ bcur, u64 *__restrict__ Dk,
                                                       u64 *__restrict__ Did) {
    const size_t ntiles = (L.n + PB - 1) / PB;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * PB, jend = L.lo + L.n;
        const u64 j1 = j0 + PB < jend ? j0 + PB : jend;
        size_t p0, p1;
        tile_parents(L, j0, j1, &p0, &p1);
        const u64 j = j0 + threadIdx.x;
        if (j < j1) {
            u64 key, id;
            u32 inst;
            leaf_eval(L, find_parent(L.off, p0, p1, j), j, key, id, inst);
            const u64 b = bucket_of(L.descs[inst], key);
            const u64 at = bstart[b] + atomicAdd(&bcur[b], 1u);
            Dk[at] = key;
            Did[at] = id;
        }
    }
}

// ---- I side: partition by the top bits of the fine bucket -----------------------------------
__global__ void __launch_bounds__(PB) k_part_count(LeafSrc L, int pshift, int P, u32 *__restrict__ pcount) {
    __shared__ u32 h[256];
    const size_t ntiles = (L.n + PT - 1) / PT;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * PT, jend = L.lo + L.n;
        const u64 j1 = j0 + PT < jend ? j0 + PT : jend;
        h[threadIdx.x] = 0;
        size_t p0, p1;
        tile_parents(L, j0, j1, &p0, &p1);
        size_t pp = p0;
#pragma unroll 4
        for (int it = 0; it < PI; ++it) {
            const u64 j = j0 + (u64)it * PB + threadIdx.x;
            if (j < j1) {
                pp = find_parent(L.off, pp, p1, j);
                u64 key, id;
                u32 inst;
                leaf_eval(L, pp, j, key, id, inst);
                atomicAdd(&h[(u32)(bucket_of(L.descs[inst], key) >> pshift)], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < P && h[threadIdx.x]) atomicAdd(&pcount[threadIdx.x], h[threadIdx.x]);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(PB) k_part_scatter(LeafSrc L, int pshift, int P, const u64 *__restrict__ pstart,
                                                     u32 *__restrict__ pcur, u64 *__restrict__ Ik, u64 *__restrict__ Iid) {
    __shared__ u32 h[256];
    __shared__ u32 loc[256];
    __shared__ u64 gbase[256];
    __shared__ u32 s_w[PB / 64];
    __shared__ u64 s_k[ST];
    __shared__ u64 s_i[ST];
    __shared__ u8_t s_p[ST];
    const size_t ntiles = (L.n + ST - 1) / ST;
    for (size_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const u64 j0 = L.lo + t * ST, jend = L.lo + L.n;
        const u64 j1 = j0 + ST < jend ? j0 + ST : jend;
        const int cnt = (int)(j1 - j0);
        h[threadIdx.x] = 0;
        size_t p0, p1;
        tile_parents(L, j0, j1, &p0, &p1);
        u64 key[SI], id[SI];
        u32 part[SI], rank[SI];
        size_t pp = p0;
#pragma unroll
        for (int it = 0; it < SI; ++it) {
            const u64 j = j0 + (u64)it * PB + threadIdx.x;
            if (j < j1) {
                u32 inst;
                pp = find_parent(L.off, pp, p1, j);
                leaf_eval(L, pp, j, key[it], id[it], inst);
                part[it] = (u32)(bucket_of(L.descs[inst], key[it]) >> pshift);
                rank[it] = atomicAdd(&h[part[it]], 1u);
            }
        }
        __syncthreads();
        // local exclusive offsets of the partitions within the tile (block scan of 256 counts),
        // and one global reservation per non-empty partition
        {
            const u32 v = h[threadIdx.x];
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            u32 inc = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const u32 o = __shfl_up(inc, d, 64);
                if (lane >= d) inc += o;
            }
            if (lane == 63) s_w[w] = inc;
            __syncthreads();
            u32 pre = 0;
            for (int q = 0; q < w; ++q) pre += s_w[q];
            loc[threadIdx.x] = pre + inc - v;
            gbase[threadIdx.x] = (threadIdx.x < (unsigned)P && v) ? pstart[threadIdx.x] + atomicAdd(&pcur[threadIdx.x], v) : 0;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < SI; ++it) {
            const int le = it * PB + threadIdx.x;
            if (le < cnt) {
                const u32 at = loc[part[it]] + rank[it];
                s_k[at] = key[it];
                s_i[at] = id[it];
                s_p[at] = (u8_t)part[it];
            }
        }
        __syncthreads();
        // staged slot q belongs to partition s_p[q]; its runs are contiguous in the output
        for (int q = threadIdx.x; q < cnt; q += PB) {
            const u32 pi = s_p[q];
            const u64 at = gbase[pi] + (u64)(q - (int)loc[pi]);
            Ik[at] = s_k[q];
            Iid[at] = s_i[q];
        }
        __syncthreads();
    }
}

// ---- probe: partitioned I records against the bucketed D records ----------------------------
// Instance of a key: the last koff <= key (koffs ascending, koffs[0] = 0).
__device__ __forceinline__ u32 inst_of_key(const u64 *__restrict__ koffs, int ninst, u64 key) {
    int a = 0, b = ninst - 1;
    while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (koffs[mid] <= key) a = mid; else b = mid - 1;
    }
    return (u32)a;
}

__global__ void __launch_bounds__(PB) k_part_probe(const u64 *__restrict__ Ik, const u64 *__restrict__ Iid,
                                                   const SideDesc *__restrict__ descs, const u64 *__restrict__ koffs,
                                                   int ninst, u64 nI,
                                                   const u64 *__restrict__ bstart, const u64 *__restrict__ Dk,
                                                   const u64 *__restrict__ Did, unsigned long long *__restrict__ npairs,
                                                   u64 cap, u64 *__restrict__ out_d, u64 *__restrict__ out_i,
                                                   u64 *__restrict__ out_inst) {
    const u64 ntiles = (nI + PT - 1) / PT;
    // XCD-aware order: block b runs on XCD b % 8; give each XCD a contiguous slice of tiles
    const u32 G = gridDim.x;
    const u32 xcd = blockIdx.x % NXCD;
    const u32 nb_x = G / NXCD + (xcd < G % NXCD ? 1 : 0);     // blocks on this XCD
    const u32 k = blockIdx.x / NXCD;                          // rank of this block on its XCD
    const u64 t_lo = ntiles * xcd / NXCD, t_hi = ntiles * (xcd + 1) / NXCD;
    for (u64 t = t_lo + k; t < t_hi; t += nb_x) {
        const u64 e0 = t * PT;
#pragma unroll 4
        for (int it = 0; it < PI; ++it) {
            const u64 e = e0 + (u64)it * PB + threadIdx.x;
            if (e >= nI) break;
            const u64 key = Ik[e];
            const u32 inst = ninst == 1 ? 0u : inst_of_key(koffs, ninst, key);
            const u64 b = bucket_of(descs[inst], key);
            const u64 lo = bstart[b], hi = bstart[b + 1];
            for (u64 q = lo; q < hi; ++q) {
                if (Dk[q] == key) {
                    const unsigned long long at = atomicAdd(npairs, 1ULL);
                    if (at < cap) {
                        out_d[at] = Did[q];
                        out_i[at] = Iid[e];
                        out_inst[at] = inst;
                    }
                }
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

void part_bucket_count(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                       int depth, u32 *bcount) {
    if (!n) return;
    LeafSrc L{descs, nodes, nn, off, lo, n, depth};
    CF_LAUNCH(ctx, "k_bucket_count", k_bucket_count, dim3(grid_for(n, PB)), dim3(PB), 0, L, bcount);
}

void part_bucket_scatter(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo,
                         u64 n, int depth, const u64 *bstart, u32 *bcur, u64 *Dk, u64 *Did) {
    if (!n) return;
    LeafSrc L{descs, nodes, nn, off, lo, n, depth};
    CF_LAUNCH(ctx, "k_bucket_scatter", k_bucket_scatter, dim3(grid_for(n, PB)), dim3(PB), 0, L, bstart, bcur, Dk, Did);
}

void part_count(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                int depth, int pshift, int P, u32 *pcount) {
    if (!n) return;
    LeafSrc L{descs, nodes, nn, off, lo, n, depth};
    CF_LAUNCH(ctx, "k_part_count", k_part_count, dim3(grid_for(n, PT)), dim3(PB), 0, L, pshift, P, pcount);
}

void part_scatter(cf_ctx *ctx, const SideDesc *descs, const Node *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                  int depth, int pshift, int P, const u64 *pstart, u32 *pcur, u64 *Ik, u64 *Iid) {
    if (!n) return;
    LeafSrc L{descs, nodes, nn, off, lo, n, depth};
    CF_LAUNCH(ctx, "k_part_scatter", k_part_scatter, dim3(grid_for(n, ST)), dim3(PB), 0, L, pshift, P, pstart, pcur,
              Ik, Iid);
}

void part_probe(cf_ctx *ctx, const u64 *Ik, const u64 *Iid, const SideDesc *descs, const u64 *koffs, int ninst, u64 nI,
                const u64 *bstart,
                const u64 *Dk, const u64 *Did, unsigned long long *npairs, u64 cap, u64 *od, u64 *oi, u64 *oinst) {
    if (!nI) return;
    // enough blocks to fill the chip several times over, a multiple of the XCD count
    const u64 ntiles = (nI + PT - 1) / PT;
    // the tile order gives each XCD a slice of tiles, so every XCD needs at least one block:
    // the grid is a non-zero multiple of the XCD count
    unsigned g = (unsigned)(ntiles < 8192 ? ntiles : 8192);
    g = g < NXCD ? NXCD : g - g % NXCD;
    CF_LAUNCH(ctx, "k_part_probe", k_part_probe, dim3(g), dim3(PB), 0, Ik, Iid, descs, koffs, ninst, nI, bstart, Dk, Did, npairs,
              cap, od, oi, oinst);
}

void inst_bounds(cf_ctx *ctx, const Node *nodes, const u64 *off, size_t nn, u64 *out) {
    if (!nn) return;
    CF_LAUNCH(ctx, "k_inst_bounds", k_inst_bounds, dim3(grid_for(nn, PB, 65536)), dim3(PB), 0, nodes, off, nn, out);
}
