// k_gen.hip -- r-subset residue generation on gfx950.
//
// Two generators:
//  * tree expansion (filtered sides): the r-subsets admitted by the size bounds are the leaves
//    of a prefix tree. Because every bound is monotone in the next index (products grow with the
//    index, the stratum count never falls), the admissible children of a node form one
//    contiguous index range starting at the node's next candidate. A level is expanded by
//      count (binary search of the range end)  ->  exclusive scan  ->  write (one thread per child)
//    so work is balanced across lanes however lopsided the subtrees are, and outputs come out
//    in the depth-first (lexicographic) order of the reference's recursive enumerators:
//      I side (ascending, prod < Icap, cheapest completion, stratum): src/mitm64_idump_stream2.cpp:47-57
//      D side (descending, prod >= Dmin reach prune):                 src/mitm64_ddump.cpp:29-44
//  * rank ranges (unfiltered, compatibility generate()): one lane per combinatorial-number-system
//    rank, lexicographic unranking as gpu/mitm_gpu2.py:40-52.
#include "tree_dev.h"

namespace {

constexpr int BLK = 256;


// Number of primes with index in [x+1, x+t] at or above split_idx: the fewest "large" primes any
// completion after choosing x can contain (indices only grow).
__device__ __forceinline__ int min_large_after(int x, int t, int split_idx) {
    int lo = x + 1 > split_idx ? x + 1 : split_idx;
    int c = x + t - lo + 1;
    return c > 0 ? c : 0;
}

// Tables a level's kernels read: the bound table, the primes and their Montgomery forms. GTab reads
// them from the instance's descriptor in global memory; LTab from copies a block staged in LDS (one
// row of the bound table, the one every child of the launch reads).
struct GTab {
    const SideDesc &s;
    __device__ __forceinline__ U128 bound(int t, int x) const { return s.bound[(size_t)t * s.n + x]; }
    __device__ __forceinline__ u64 p(int x) const { return s.p[x]; }
    __device__ __forceinline__ u64 pR(int x) const { return s.pR[x]; }
};
struct LTab {
    const U128 *brow;
    const u64 *p_, *pR_;
    __device__ __forceinline__ U128 bound(int, int x) const { return brow[x]; }
    __device__ __forceinline__ u64 p(int x) const { return p_[x]; }
    __device__ __forceinline__ u64 pR(int x) const { return pR_[x]; }
};

// Is index x an admissible choice at a node with product A, `nl` large primes so far, t more
// choices to follow?  I side: A * a_x * ... * a_{x+t} <= Icap - 1 and the stratum can still hold.
template <class Tab>
__device__ __forceinline__ bool valid_i(const SideDesc &s, const Tab &tb, U128 A, int nl, int x, int t) {
    const U128 h = tb.bound(t, x);
    if (u128_gt(A, h)) return false;
    const int nlx = nl + (x >= s.split_idx);
    return nlx + min_large_after(x, t, s.split_idx) <= s.hmax;
}
// D side: A * b_i * b_{i-1} * ... * b_{i-t} >= Dmin, i.e. A >= G[t][i].
template <class Tab>
__device__ __forceinline__ bool valid_d(const Tab &tb, U128 A, int i, int t) {
    return !u128_lt(A, tb.bound(t, i));
}

// Children of a depth-`depth` node: returns the count; the children are nxt, nxt+1, ... (I) or
// nxt, nxt-1, ... (D).
template <class Tab>
__device__ int child_count(const SideDesc &s, const Tab &tb, const Node &nd, int depth) {
    const int t = s.r - depth - 1;
    const U128 A = {nd.A_lo, nd.A_hi};
    const int nxt = node_nxt(nd);
    if (s.side == SIDE_I) {
        const int nl = (int)(nd.meta & 0xFFu);
        int lo = nxt, hi = s.n - t;   // candidates [lo, hi)
        if (lo >= hi || !valid_i(s, tb, A, nl, lo, t)) return 0;
        // largest valid x: valid on [lo, v], invalid on (v, hi)
        int a = lo, b = hi - 1;
        while (a < b) {
            int mid = (a + b + 1) >> 1;
            if (valid_i(s, tb, A, nl, mid, t)) a = mid; else b = mid - 1;
        }
        return a - lo + 1;
    } else {
        int lo = t, hi = nxt;          // candidates hi, hi-1, ..., lo
        if (hi < lo || !valid_d(tb, A, hi, t)) return 0;
        int a = lo, b = hi;            // smallest valid i
        while (a < b) {
            int mid = (a + b) >> 1;
            if (valid_d(tb, A, mid, t)) b = mid; else a = mid + 1;
        }
        return hi - a + 1;
    }
}

// Child number k of a depth-`depth` node.
template <class Tab>
__device__ __forceinline__ Node make_child(const SideDesc &s, const Tab &tb, const Node &nd, int depth, int k) {
    const int x = (s.side == SIDE_I) ? node_nxt(nd) + k : node_nxt(nd) - k;
    const u64 p = tb.p(x);
    Node c;
    const U128 A = u128_mul64(U128{nd.A_lo, nd.A_hi}, p);
    c.A_lo = A.lo;
    c.A_hi = A.hi;
    c.xm = cf_mulmod_m(nd.xm, tb.pR(x), s.mc);
    c.x2 = nd.x2 * p;
    c.id = id_update(s, nd.id, x, depth);
    const int nn = (s.side == SIDE_I) ? x + 1 : x - 1;
    const u32 first = depth == 0 ? (u32)x : (nd.pos >> 16);
    c.pos = ((u32)(unsigned short)(short)nn) | (first << 16);
    c.meta = (nd.meta & ~0xFFu) | ((nd.meta & 0xFFu) + (s.side == SIDE_I && x >= s.split_idx));
    return c;
}

__global__ void __launch_bounds__(BLK) k_expand_count(const SideDesc *__restrict__ descs, const Node *__restrict__ nodes,
                                                      size_t nn, int depth, u32 *__restrict__ cnt) {
    for (size_t j = (size_t)blockIdx.x * BLK + threadIdx.x; j < nn; j += (size_t)gridDim.x * BLK) {
        const Node nd = nodes[j];
        const SideDesc &s = descs[nd.meta >> 8];
        cnt[j] = (u32)child_count(s, GTab{s}, nd, depth);
    }
}


// A level is written in tiles of EC = BLK * EPT outputs, one block of BLK threads each (EPT outputs
// per thread): EPT = 4 on large levels, so a block's staged tables serve 1024 outputs; EPT = 1 on
// small ones (the top of the trees, where tiles span many instances and take the global path), so
// no thread runs several searches of global tables one after another.
constexpr int NSTAGE = 512;      // instances with at most this many primes have their tables staged
constexpr u64 EPT4_MIN = 1u << 20;   // outputs of a level written with EPT = 4

// Owner of the first output of every ec-output tile: parent p's outputs are [off[p], off[p+1]), so
// it owns the tiles that start there (a parent has at most a pool's worth of children, so a few
// tiles at most). One thread per parent, one coalesced pass over off.
__global__ void __launch_bounds__(BLK) k_mark_owners(const u64 *__restrict__ off, size_t nn, u64 ec,
                                                     u64 *__restrict__ tp) {
    for (size_t p = (size_t)blockIdx.x * BLK + threadIdx.x; p < nn; p += (size_t)gridDim.x * BLK) {
        const u64 a = off[p], b = off[p + 1];
        if (b <= a) continue;
        for (u64 t = (a + ec - 1) / ec; t <= (b - 1) / ec; ++t) tp[t] = p;
    }
}

// One level down with the next level's counts: each child is written with the number of its own
// children, so the next level needs no separate count pass. A tile's parents lie between the owner
// of its first output and the owner of the next tile's first; each parent with children marks its
// first output's slot in LDS, and a running maximum over the slots gives every output its parent.
// When the tile's parents belong to one instance of at most NSTAGE primes (nearly every tile), its
// primes, their Montgomery forms and the one bound-table row the children's counts read are staged
// in LDS, so the counts' binary searches read LDS rather than global memory.
// OutT: Node, or LeafNode for the last interior level (the children's products are not kept).
__device__ __forceinline__ void put_node(Node *out, const Node &c) { *out = c; }
__device__ __forceinline__ void put_node(LeafNode *out, const Node &c) {
    LeafNode l;
    l.xm = c.xm;
    l.x2 = c.x2;
    l.id = c.id;
    l.pos = c.pos;
    l.meta = c.meta;
    *out = l;
}
template <int EPT, class OutT>
__global__ void __launch_bounds__(BLK) k_expand_write_count(const SideDesc *__restrict__ descs,
                                                           const Node *__restrict__ nodes, size_t nn,
                                                           const u64 *__restrict__ off,
                                                           const u64 *__restrict__ tp, size_t nout, int depth,
                                                           OutT *__restrict__ out, u32 *__restrict__ cnt) {
    __shared__ U128 s_b[NSTAGE];
    __shared__ u64 s_p[NSTAGE], s_pr[NSTAGE];
    constexpr int EC = BLK * EPT;
    __shared__ u32 s_own[EC];
    __shared__ u32 s_w[BLK / 64];
    constexpr u64 MARKCAP = 8 * EC;   // parents per tile resolved through LDS marks (more: binary search)
    const u64 j0 = (u64)blockIdx.x * EC;
    const u32 cnt_t = (u32)(nout - j0 < (u64)EC ? nout - j0 : (u64)EC);
    const size_t p0 = tp[blockIdx.x];
    const size_t p1 = blockIdx.x + 1 < gridDim.x ? tp[blockIdx.x + 1] : nn - 1;
    const u64 np = p1 - p0 + 1;
    const u32 i0 = nodes[p0].meta >> 8, i1 = nodes[p1].meta >> 8;
    const SideDesc &s0 = descs[i0];
    const bool staged = i0 == i1 && s0.n <= NSTAGE;
    const bool marked = np <= MARKCAP;
    if (staged) {
        const int t = s0.r - depth - 2;   // the children (depth + 1) count with row r - (depth + 1) - 1
        for (int x = threadIdx.x; x < s0.n; x += BLK) {
            s_b[x] = s0.bound[(size_t)t * s0.n + x];
            s_p[x] = s0.p[x];
            s_pr[x] = s0.pR[x];
        }
    }
    if (marked) {
#pragma unroll
        for (int k = 0; k < EPT; ++k) s_own[threadIdx.x * EPT + k] = 0;
        __syncthreads();
        for (u64 q = 1 + threadIdx.x; q < np; q += BLK) {
            const u64 a = off[p0 + q];
            if (a < j0 + cnt_t && off[p0 + q + 1] > a) s_own[a - j0] = (u32)q;
        }
        __syncthreads();
        // running maximum: EPT slots per thread, then a block scan of the thread maxima
        u32 v[EPT], run = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            v[k] = s_own[threadIdx.x * EPT + k];
            run = run > v[k] ? run : v[k];
        }
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const u32 inc = wave_incl_max(run);
        if (lane == 63) s_w[wv] = inc;
        __syncthreads();
        u32 cur = 0;
        for (int q = 0; q < wv; ++q) cur = cur > s_w[q] ? cur : s_w[q];
        const u32 ex = wave_shift_up1(inc);
        cur = cur > ex ? cur : ex;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            cur = cur > v[k] ? cur : v[k];
            s_own[threadIdx.x * EPT + k] = cur;
        }
    }
    __syncthreads();
    // every parent is loaded before any child is built, so the EPT loads are in flight together
    // (slots past the level's end repeat its last output and are not stored)
    Node nd[EPT];
    int kc[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const u32 i = threadIdx.x + BLK * k < cnt_t ? threadIdx.x + BLK * k : cnt_t - 1;
        const u64 j = j0 + i;
        const size_t p = marked ? p0 + s_own[i] : find_parent(off, p0, p1, j);
        nd[k] = nodes[p];
        kc[k] = (int)(j - off[p]);
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const u32 i = threadIdx.x + BLK * k;
        Node c;
        u32 cc;
        if (staged) {
            const LTab tb{s_b, s_p, s_pr};
            c = make_child(s0, tb, nd[k], depth, kc[k]);
            cc = (u32)child_count(s0, tb, c, depth + 1);
        } else {
            const SideDesc &s = descs[nd[k].meta >> 8];
            c = make_child(s, GTab{s}, nd[k], depth, kc[k]);
            cc = (u32)child_count(s, GTab{s}, c, depth + 1);
        }
        if (i < cnt_t) {
            put_node(out + j0 + i, c);
            cnt[j0 + i] = cc;
        }
    }
}

// Nodes as LeafNodes (the roots of a radius-1 side are its last interior level).
__global__ void __launch_bounds__(BLK) k_leaf_nodes(const Node *__restrict__ nodes, size_t nn,
                                                    LeafNode *__restrict__ out) {
    for (size_t j = (size_t)blockIdx.x * BLK + threadIdx.x; j < nn; j += (size_t)gridDim.x * BLK) put_node(out + j, nodes[j]);
}

// Per first index: the leaf offsets where its nodes start and end. The nodes of one first index
// are one contiguous run (depth-first order), so the node opening a run writes lo[first] = off[j]
// and the node closing it writes hi[first] = off[j + 1]; its leaves are hi - lo (0 - 0 for a first
// index with no node). One streaming pass, no atomics (a per-node atomic on the run's counter
// serializes every node of a run).
__global__ void k_first_bounds(const LeafNode *__restrict__ nodes, const u64 *__restrict__ off, size_t nn,
                               u64 *__restrict__ lo, u64 *__restrict__ hi) {
    for (size_t j = (size_t)blockIdx.x * BLK + threadIdx.x; j < nn; j += (size_t)gridDim.x * BLK) {
        const u32 f = nodes[j].pos >> 16;
        if (j == 0 || (nodes[j - 1].pos >> 16) != f) lo[f] = off[j];
        if (j + 1 == nn || (nodes[j + 1].pos >> 16) != f) hi[f] = off[j + 1];
    }
}

// Compatibility generator: lane per CNS rank in [lo, lo + count), lexicographic unranking,
// key = mult * prod mod M, id = indices packed id_bits wide (gpu/mitm_gpu2.py:36-64).
__global__ void __launch_bounds__(BLK) k_gen_rank(const SideDesc *__restrict__ descs, u64 lo, size_t count,
                                                  u64 *__restrict__ keys, u64 *__restrict__ ids) {
    const SideDesc &s = descs[0];
    const int n = s.n, R = s.r, stride = R + 1;
    for (size_t t = (size_t)blockIdx.x * BLK + threadIdx.x; t < count; t += (size_t)gridDim.x * BLK) {
        u64 rank = lo + t;
        u64 xm = s.xm0, x2 = s.x20, id = 0;
        int start = 0;
        for (int pos = 0; pos < R; ++pos) {
            const int remaining = R - pos - 1;
            int c = start;
            for (; c < n - 1; ++c) {
                const u64 cntc = s.binom[(size_t)(n - 1 - c) * stride + remaining];
                if (rank < cntc) break;
                rank -= cntc;
            }
            start = c + 1;
            xm = cf_mulmod_m(xm, s.pR[c], s.mc);
            x2 *= s.p[c];
            id |= (u64)c << (s.id_bits * pos);
        }
        keys[t] = cf_key(xm, x2, s.mc) + s.koff;
        ids[t] = id;
    }
}

// Compatibility generator, prefix-shared. In lexicographic order the records of consecutive ranks
// come in runs that share their first r - 1 indices while the last index counts up, so a wave takes
// 64 consecutive prefixes (ranks pr0 + q among the (r - 1)-subsets of {0 .. n-2}): each lane
// unranks its prefix and multiplies out its residue state and id once; the rank of its run's first
// record is sum_pos C(n - start_pos, r - pos) - C(n - c_pos, r - pos) (hockey stick), and its run
// is n - 1 - c_{r-2} records long. The wave then writes the records of [lo, lo + count) its runs
// cover 64 at a time, contiguously: a lane finds the run of its record by a binary search over the
// lanes' run starts and takes that lane's state with shuffles, then one residue step for the last
// index. Every lane of the wave runs every step (no shuffle reads an inactive lane); only the stores
// are masked.
__global__ void __launch_bounds__(BLK) k_gen_prefix(const SideDesc *__restrict__ descs, u64 lo, u64 count, u64 pr0,
                                                    u64 npre, u64 *__restrict__ keys, u64 *__restrict__ ids) {
    const SideDesc &s = descs[0];
    const int n = s.n, R = s.r, stride = R + 1, K = R - 1;
    const int lane = threadIdx.x & 63;
    const u64 hi = lo + count;
    const u64 waves = (u64)gridDim.x * (BLK / 64);
    for (u64 wv = (u64)blockIdx.x * (BLK / 64) + (threadIdx.x >> 6); wv * 64 < npre; wv += waves) {
        const u64 q = wv * 64 + (u64)lane;
        const bool has = q < npre;
        u64 rank = pr0 + (has ? q : 0);
        u64 xm = s.xm0, x2 = s.x20, id = 0, r0 = 0;
        int start = 0;
        for (int pos = 0; pos < K; ++pos) {
            const int rem = K - 1 - pos;
            int c = start;
            for (; c < n - 2; ++c) {
                const u64 cnt = s.binom[(size_t)(n - 2 - c) * stride + rem];
                if (rank < cnt) break;
                rank -= cnt;
            }
            r0 += s.binom[(size_t)(n - start) * stride + (R - pos)] - s.binom[(size_t)(n - c) * stride + (R - pos)];
            start = c + 1;
            xm = cf_mulmod_m(xm, s.pR[c], s.mc);
            x2 *= s.p[c];
            id |= (u64)c << (s.id_bits * pos);
        }
        // this lane's run clipped to [lo, hi), as output positions [sb, se); lanes past the last prefix
        // start after every output
        const u64 a = r0 > lo ? r0 : lo;
        const u64 e0 = r0 + (u64)(n - start);
        const u64 b = e0 < hi ? e0 : hi;
        const u64 sb = has ? a - lo : ~0ull;
        const u64 se = has ? b - lo : 0;
        const u64 w_lo = (u64)__shfl((long long)sb, 0, 64);
        u64 w_hi = se;
        for (int d = 32; d >= 1; d >>= 1) {
            const u64 o = (u64)__shfl_xor((long long)w_hi, d, 64);
            w_hi = o > w_hi ? o : w_hi;
        }
        for (u64 o0 = w_lo; o0 < w_hi; o0 += 64) {
            const u64 o = o0 + (u64)lane < w_hi ? o0 + (u64)lane : w_hi - 1;
            int j = 0;
            for (int d = 32; d >= 1; d >>= 1) {
                const u64 sj = (u64)__shfl((long long)sb, j + d, 64);
                if (sj <= o) j += d;
            }
            const u64 jxm = (u64)__shfl((long long)xm, j, 64);
            const u64 jx2 = (u64)__shfl((long long)x2, j, 64);
            const u64 jid = (u64)__shfl((long long)id, j, 64);
            const u64 jr0 = (u64)__shfl((long long)r0, j, 64);
            const int jst = __shfl(start, j, 64);
            const int last = jst + (int)(o + lo - jr0);
            const u64 kxm = cf_mulmod_m(jxm, s.pR[last], s.mc);
            const u64 kx2 = jx2 * s.p[last];
            if (o0 + (u64)lane < w_hi) {
                keys[o] = cf_key(kxm, kx2, s.mc) + s.koff;
                ids[o] = jid | ((u64)last << (s.id_bits * K));
            }
        }
    }
}

inline unsigned grid_for(size_t n, size_t per_block, unsigned cap = 1u << 20) {
    size_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

}  // namespace

void launch_expand_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, int depth, u32 *cnt) {
    if (!nn) return;
    CF_LAUNCH(ctx, "k_expand_count", k_expand_count, dim3(grid_for(nn, BLK)), dim3(BLK), 0, d_desc, nodes, nn, depth, cnt);
}


template <class OutT>
static void expand_write_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, const u64 *off,
                               size_t nout, int depth, OutT *out, u32 *cnt) {
    if (!nout) return;
    const bool wide = nout >= EPT4_MIN;
    const u64 ec = wide ? 4 * BLK : BLK;
    const u64 ntiles = (nout + ec - 1) / ec;
    DBuf<u64> tp(ctx, ntiles);
    CF_LAUNCH(ctx, "k_mark_owners", k_mark_owners, dim3(grid_for(nn, BLK, 65536)), dim3(BLK), 0, off, nn, ec, tp.p);
    if (wide)
        CF_LAUNCH(ctx, "k_expand_write_count", (k_expand_write_count<4, OutT>), dim3((unsigned)ntiles), dim3(BLK), 0,
                  d_desc, nodes, nn, off, tp.p, nout, depth, out, cnt);
    else
        CF_LAUNCH(ctx, "k_expand_write_count", (k_expand_write_count<1, OutT>), dim3((unsigned)ntiles), dim3(BLK), 0,
                  d_desc, nodes, nn, off, tp.p, nout, depth, out, cnt);
}
void launch_expand_write_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, const u64 *off,
                               size_t nout, int depth, Node *out, u32 *cnt) {
    expand_write_count(ctx, d_desc, nodes, nn, off, nout, depth, out, cnt);
}
void launch_expand_write_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, const u64 *off,
                               size_t nout, int depth, LeafNode *out, u32 *cnt) {
    expand_write_count(ctx, d_desc, nodes, nn, off, nout, depth, out, cnt);
}
void launch_leaf_nodes(cf_ctx *ctx, const Node *nodes, size_t nn, LeafNode *out) {
    if (!nn) return;
    CF_LAUNCH(ctx, "k_leaf_nodes", k_leaf_nodes, dim3(grid_for(nn, BLK, 65536)), dim3(BLK), 0, nodes, nn, out);
}

void launch_first_bounds(cf_ctx *ctx, const LeafNode *nodes, const u64 *off, size_t nn, u64 *lo, u64 *hi) {
    if (!nn) return;
    CF_LAUNCH(ctx, "k_first_bounds", k_first_bounds, dim3(grid_for(nn, BLK, 65536)), dim3(BLK), 0, nodes, off, nn, lo,
              hi);
}

void launch_gen_rank(cf_ctx *ctx, const SideDesc *d_desc, u64 lo, size_t count, u64 *keys, u64 *ids) {
    if (!count) return;
    CF_LAUNCH(ctx, "k_gen_rank", k_gen_rank, dim3(grid_for(count, BLK)), dim3(BLK), 0, d_desc, lo, count, keys, ids);
}

namespace {
typedef unsigned __int128 u128;
u128 binom_exact(int a, int b) {
    if (b < 0 || b > a) return 0;
    u128 c = 1;
    for (int j = 0; j < b; ++j) c = c * (u128)(a - j) / (u128)(j + 1);
    return c;
}
// the k-subset of {0 .. m-1} of lexicographic rank `rank`
void lex_unrank(u128 rank, int m, int k, int *out) {
    int start = 0;
    for (int pos = 0; pos < k; ++pos) {
        int c = start;
        for (;; ++c) {
            const u128 cnt = binom_exact(m - 1 - c, k - 1 - pos);
            if (rank < cnt) break;
            rank -= cnt;
        }
        out[pos] = c;
        start = c + 1;
    }
}
// the lexicographic rank of the ascending k-subset c of {0 .. m-1}
u128 lex_rank(const int *c, int m, int k) {
    u128 r = 0;
    int start = 0;
    for (int pos = 0; pos < k; ++pos) {
        r += binom_exact(m - start, k - pos) - binom_exact(m - c[pos], k - pos);
        start = c[pos] + 1;
    }
    return r;
}
}  // namespace

void launch_gen_prefix(cf_ctx *ctx, const SideDesc *d_desc, int n, int r, u64 lo, size_t count, u64 *keys, u64 *ids) {
    if (!count) return;
    CF_REQUIRE(r >= 1 && n >= r && n <= 1024 && r <= 16, "rank-range generation needs 1 <= r <= n <= 1024");
    CF_REQUIRE((u128)lo + count <= binom_exact(n, r), "rank range beyond C(n, r)");
    // the prefixes (first r - 1 indices, a subset of {0 .. n-2}) of the range's first and last records
    int c0[16], c1[16];
    lex_unrank(lo, n, r, c0);
    lex_unrank((u128)lo + count - 1, n, r, c1);
    const u64 pr0 = (u64)lex_rank(c0, n - 1, r - 1);
    const u64 npre = (u64)(lex_rank(c1, n - 1, r - 1) - pr0 + 1);
    const u64 waves = (npre + 63) / 64;
    CF_LAUNCH(ctx, "k_gen_prefix", k_gen_prefix, dim3(grid_for(waves, BLK / 64, 1u << 16)), dim3(BLK), 0, d_desc, lo,
              (u64)count, pr0, npre, keys, ids);
}

// ---- host: per-side tables ---------------------------------------------------------------
typedef unsigned __int128 u128;

static inline u128 div128(u128 a, u128 b) {
    // 64-bit division when both fit (the common case), else the 128-bit routine
    if ((a >> 64) == 0 && (b >> 64) == 0) return (u128)((u64)a / (u64)b);
    return a / b;
}

// Host tables of one side, appended to buf (u64 words, U128 entries 16-byte aligned):
//   p[n] primes, pR[n] = p * 2^64 mod m, bound[r*n], binom[(n+1)*(r+1)];
// off[0..3] receive the word offsets of the four arrays. Fills every field of d except the
// four device pointers.
void side_tables_host(const SideSpec &s, SideDesc &d, std::vector<u64> &buf, size_t off[4]) {
    const int n = (int)s.primes.size();
    const int r = s.r;
    CF_REQUIRE(n >= 0 && n <= 1024, "pool size must be <= 1024");
    CF_REQUIRE(r >= 1 && r <= 16, "radius must be in [1, 16]");
    CF_REQUIRE(s.M >= 1, "modulus must be >= 1");
    d = SideDesc{};
    d.mc = cf_modctx(s.M);
    d.n = n;
    d.r = r;
    d.side = s.side;
    d.id_mode = s.id_mode;
    d.id_bits = s.id_bits;
    d.split_idx = s.split_idx;
    d.hmax = s.hmax;
    d.koff = s.koff;
    const u64 mult = s.mult % s.M;
    d.xm0 = d.mc.m == 1 ? 0 : mult % d.mc.m;
    d.x20 = mult;
    d.s32 = d.mc.m == 1 ? 0 : (u64)(((unsigned __int128)1 << 96) % d.mc.m);
    // one side's records take the 32-bit leaf step when its primes are below 2^32 (a join sets this
    // again from both sides: the scaled residues it compares must share one step)
    d.leaf32 = 1;
    for (u64 q : s.primes) d.leaf32 &= q < (1ULL << 32);
    const int nn = n ? n : 1;
    auto take = [&](size_t words) {
        size_t at = (buf.size() + 1) & ~(size_t)1;   // 16-byte alignment for every array
        buf.resize(at + words, 0);
        return at;
    };
    off[0] = take(nn);
    off[1] = take(nn);
    off[2] = take((size_t)r * nn * 2);
    off[3] = take((size_t)(n + 1) * (r + 1));
    for (int i = 0; i < n; ++i) {
        buf[off[0] + i] = s.primes[i];
        buf[off[1] + i] = d.mc.m == 1 ? 0 : cf_montmul_factor(s.primes[i], d.mc);
    }
    // bound table, exact: I: floor((Icap - 1) / W), W = a_x .. a_{x+t}; D: ceil(Dmin / W),
    // W = b_i .. b_{i-t}; the window grows one prime per t. Entries whose window leaves the pool
    // are never consulted (the candidate ranges exclude them) and are set to reject.
    const u128 ALL = ~(u128)0;
    const u128 top = s.cap ? s.cap - 1 : 0;
    for (int x = 0; x < n; ++x) {
        u128 w = 1;
        bool over = false;
        for (int tt = 0; tt < r; ++tt) {
            const int y = s.side == SIDE_I ? x + tt : x - tt;
            u128 v;
            if (y < 0 || y >= n) {
                v = s.side == SIDE_I ? 0 : ALL;
            } else {
                const u64 q = s.primes[y];
                if (!over && w > ALL / q) over = true;
                if (!over) w *= q;
                if (s.side == SIDE_I) {
                    v = over ? 0 : div128(top, w);
                } else {
                    if (over) {
                        v = s.cap ? 1 : 0;
                    } else {
                        const u128 qq = div128(s.cap, w);
                        v = qq + (qq * w != s.cap);
                    }
                }
            }
            const size_t e = off[2] + 2 * ((size_t)tt * nn + x);
            buf[e] = (u64)v;
            buf[e + 1] = (u64)(v >> 64);
        }
    }
    for (int a = 0; a <= n; ++a) {
        buf[off[3] + (size_t)a * (r + 1)] = 1;
        for (int b = 1; b <= r && a > 0; ++b)
            buf[off[3] + (size_t)a * (r + 1) + b] =
                buf[off[3] + (size_t)(a - 1) * (r + 1) + b - 1] + buf[off[3] + (size_t)(a - 1) * (r + 1) + b];
    }
}

void side_tables_bind(SideDesc &d, const u64 *dev, const size_t off[4]) {
    d.p = dev + off[0];
    d.pR = dev + off[1];
    d.bound = (const U128 *)(dev + off[2]);
    d.binom = dev + off[3];
}

void build_side_tables(cf_ctx *ctx, const SideSpec &s, SideTables &t) {
    std::vector<u64> buf;
    size_t off[4];
    side_tables_host(s, t.desc, buf, off);
    t.p = DBuf<u64>(ctx, buf.size());
    h2d(ctx, t.p.p, buf.data(), buf.size());
    side_tables_bind(t.desc, t.p.p, off);
}
