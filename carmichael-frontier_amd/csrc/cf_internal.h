// cf_internal.h -- shared declarations of libcfjoin (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "keyhash.h"
#include "modarith.h"

// ---- errors -----------------------------------------------------------------------------------
void cf_set_error(const std::string &msg);
struct CfError {
    int code;
    std::string msg;
};
#define CF_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw CfError{-3, std::string(#call) + ": " + hipGetErrorString(e_)};           \
    } while (0)
#define CF_REQUIRE(cond, msg)                        \
    do {                                             \
        if (!(cond)) throw CfError{-1, (msg)};       \
    } while (0)

extern thread_local std::string g_err;
#include "../../include/cfjoin.h"
#define API_BEGIN try {
#define API_END                                            \
    }                                                      \
    catch (const CfError &e) {                             \
        g_err = e.msg;                                     \
        return e.code;                                     \
    }                                                      \
    catch (const std::bad_alloc &) {                       \
        g_err = "host out of memory";                      \
        return CF_ENOMEM;                                  \
    }                                                      \
    catch (const std::exception &e) {                      \
        g_err = e.what();                                  \
        return CF_EINTERNAL;                               \
    }                                                      \
    return CF_OK;

// ---- context: one device, one stream, a caching allocator ---------------------------------
struct cf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::multimap<size_t, void *> free_blocks;    // cached blocks by size
    std::map<void *, size_t> live;                // live allocations and their block size
    size_t cached_bytes = 0;
    std::mutex mu;
    void *pinned = nullptr;                       // small pinned staging buffer for scalars
    size_t pinned_bytes = 0;
    // kernel timing with HIP events on this context's stream (enabled by cf_prof_enable)
    bool prof_on = false;
    std::string prof_filter;                      // empty: every kernel
    struct ProfRec { const char *name; hipEvent_t a, b; };
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_free;
    std::map<std::string, std::pair<double, uint64_t>> prof_totals;   // name -> (ms, launches)
    // stream capture into a HIP graph: while set, blocks freed go to the arena (reused only by
    // later allocations of the same capture) and every block handed out is recorded in it, so
    // nothing a graph node touches returns to the shared pool while the graph lives
    struct CaptureArena *arena = nullptr;
    // branch of the capture being recorded (0: the capturing stream; k > 0: side stream k, forked
    // from it and joined back): a block freed on one branch is reused only on that branch, whose
    // stream order keeps the reuse safe while branches run side by side
    int arena_branch = 0;
    size_t live_bytes = 0;                        // bytes handed out and not yet freed
    size_t peak_live = 0;                         // largest live footprint since the last reset
    u64 *join_scratch = nullptr;                  // this device's scratch of cf_join_sorted_noalloc
};

struct CaptureArena {
    std::map<int, std::multimap<size_t, void *>> free;   // freed during the capture, per branch
    std::map<void *, size_t> owned;                      // every block the captured work uses
};
// return an arena's blocks to its context's pool (the graph that used them is gone)
void arena_release(cf_ctx *ctx, CaptureArena *a);

void cf_prof_begin(cf_ctx *ctx, const char *name, hipEvent_t *ev);
void cf_prof_end(cf_ctx *ctx, const char *name, hipEvent_t ev);

// Launch on the context's stream; bracketed by HIP events when profiling selects the kernel.
#define CF_LAUNCH(ctx, name, kern, grid, block, shmem, ...)                           \
    do {                                                                             \
        hipEvent_t ev_a_ = nullptr;                                                  \
        cf_prof_begin((ctx), (name), &ev_a_);                                        \
        hipLaunchKernelGGL(kern, grid, block, shmem, (ctx)->stream, __VA_ARGS__);    \
        CF_HIP(hipGetLastError());                                                   \
        cf_prof_end((ctx), (name), ev_a_);                                           \
    } while (0)

void *ctx_alloc(cf_ctx *ctx, size_t bytes);
void ctx_free(cf_ctx *ctx, void *p);
void ctx_release(cf_ctx *ctx);

// RAII device buffer from the context pool.
template <typename T>
struct DBuf {
    cf_ctx *ctx = nullptr;
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(cf_ctx *c, size_t count) : ctx(c), n(count) { p = (T *)ctx_alloc(c, count * sizeof(T)); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        reset();
        ctx = o.ctx; p = o.p; n = o.n; o.p = nullptr; o.n = 0;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (p) ctx_free(ctx, p);
        p = nullptr;
        n = 0;
    }
    T *release() { T *q = p; p = nullptr; n = 0; return q; }
};

template <typename T>
void h2d(cf_ctx *ctx, T *dst, const T *src, size_t n) {
    if (n) CF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}
template <typename T>
void d2h(cf_ctx *ctx, T *dst, const T *src, size_t n) {
    if (n) CF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
}
u64 d2h_u64(cf_ctx *ctx, const u64 *src);     // one scalar, synchronous

// ---- enumeration tree ---------------------------------------------------------------------
// The r-subsets of one side are the leaves of a tree whose depth-d nodes are the d-prefixes
// that still admit a valid completion. The bounds make the admissible children of every node
// a contiguous index range that starts at the node's `nxt`, so nodes are expanded level by
// level with a count -> scan -> write pass and no per-thread DFS.
enum : int { SIDE_I = 0, SIDE_D = 1 };
enum : int {
    ID_COLEX = 0,      // sum_j C(i_j, j + 1) over ascending positions
    ID_PACK_ASC = 1,   // index of position d at bits [bits*d, bits*d + bits)
    ID_PACK_DESC = 2,  // descending choice order; position d lands at field r-1-d (6-bit D ids)
    ID_MASK = 3,       // 64-bit deletion mask
};

struct __align__(16) Node {
    u64 A_lo, A_hi;  // exact product of the chosen primes
    u64 xm, x2;      // residue state (odd part, 2-adic part), includes the side multiplier
    u64 id;          // partial record id
    u32 pos;         // bits 0..15: first candidate index of the next choice as int16 (I: last+1,
                     // D: last-1); bits 16..31: the index chosen at depth 0
    u32 meta;        // bits 0..7: primes chosen above the stratum split; bits 8..31: instance
};
static_assert(sizeof(Node) == 48, "node layout");

// A node of a tree's last interior level, as the leaf kernels read it: the node without its exact
// product, which decides only the node's own child count (taken when the level is written).
struct __align__(16) LeafNode {
    u64 xm, x2;      // as Node
    u64 id;
    u32 pos, meta;
};
static_assert(sizeof(LeafNode) == 32, "leaf-node layout");

// Device-side description of one instance side (I or D). Arrays live in device memory.
struct SideDesc {
    ModCtx mc;
    const u64 *p;       // primes (I ascending; D ascending base)
    const u64 *pR;      // p * 2^64 mod m
    const U128 *bound;  // I: H[t*n + x] = floor((Icap-1) / (a_x..a_{x+t})); D: G[t*n + i] = ceil(Dmin / (b_i..b_{i-t}))
    const u64 *binom;   // colex table C(x, j) at [x*(r+1) + j], x <= n, j <= r (ID_COLEX only)
    u64 koff;           // offset added to every key (disjoint key ranges across a batch)
    int n, r, side, id_mode, id_bits, split_idx, hmax;
    int leaf32;         // 1: every prime of the instance (both sides) is below 2^32, so join leaves take the
                        // 32-bit scaled step (cf_redc32_mul); 0: the plain 64-bit step
    u64 xm0, x20;       // initial residue state (the side multiplier)
    u64 s32;            // 2^96 mod m: cf_mulmod_m(x, s32) = x * 2^32 mod m, which the 32-bit step
                        // (x * p * 2^-32) turns into the exact x * p
};

// Host-owned copy of one side's device tables (one allocation holding all four arrays).
struct SideTables {
    SideDesc desc{};
    DBuf<u64> p;
};

struct SideSpec {  // host description of one side
    int side = SIDE_I;
    std::vector<u64> primes;
    int r = 0;
    u64 M = 1;
    u64 mult = 1;                    // key = mult * prod mod M
    unsigned __int128 cap = 0;       // I: Icap (exclusive); D: Dmin (inclusive)
    int split_idx = 1 << 30, hmax = 1 << 30;
    int id_mode = ID_COLEX, id_bits = 8;
    u64 koff = 0;
};

void build_side_tables(cf_ctx *ctx, const SideSpec &s, SideTables &t);
void side_tables_host(const SideSpec &s, SideDesc &d, std::vector<u64> &buf, size_t off[4]);
void side_tables_bind(SideDesc &d, const u64 *dev, const size_t off[4]);

// A filtered enumeration plan: the tree of one side expanded to its last level.
struct cf_plan {
    cf_ctx *ctx = nullptr;
    SideSpec spec;
    SideTables tab;
    DBuf<SideDesc> d_desc;
    DBuf<LeafNode> leafp;  // depth r-1 nodes
    size_t nleafp = 0;
    DBuf<u32> cnt;         // leaves per depth r-1 node
    DBuf<u64> off;         // exclusive scan of cnt, nleafp + 1 entries
    u64 total = 0;
    bool empty = false;
    std::vector<u64> first_counts;
    bool have_first = false;
};
void plan_expand(cf_plan *pl);
void plan_write_range(cf_plan *pl, u64 L0, u64 L1, u64 *keys, u64 *ids);
int bitlen(u64 v);

// ---- kernels (host launchers) -------------------------------------------------------------
// exclusive scan of u32 or u64 counts into u64 offsets; returns the total (synchronous)
u64 scan_u32_to_u64(cf_ctx *ctx, const u32 *in, u64 *out, size_t n);
u64 scan_u64(cf_ctx *ctx, const u64 *in, u64 *out, size_t n);
// device address of the no-allocation join's scratch on `device` (which must be current); looked up
// once per device
u64 *join_scratch_bind(int device);
// exclusive scan of u64 with the block partials in ws (scan_ws_words(n) words; no allocation)
size_t scan_ws_words(size_t n);
void scan_u64_ws(cf_ctx *ctx, const u64 *in, u64 *out, size_t n, u64 *ws);
void scan_u32_to_u64_async(cf_ctx *ctx, const u32 *in, u64 *out, size_t n);   // no host sync

// stable LSD radix sort of (key, val) by the low `bits` bits of key; in-place via scratch
void radix_sort_pairs(cf_ctx *ctx, u64 *keys, u64 *vals, size_t n, int bits);
// the same stable sort from (kin, vin) into (kout, vout); both pairs may be one
void sort_pairs_into(cf_ctx *ctx, const u64 *kin, const u64 *vin, u64 *kout, u64 *vout, size_t n, int bits);

// tree expansion
void launch_expand_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, int depth, u32 *cnt);
// one level down (children of nodes[0, nn) with exclusive child scan off) together with each child's
// own child count (cnt, nout entries)
void launch_expand_write_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, const u64 *off,
                               size_t nout, int depth, Node *out, u32 *cnt);
// the same for the last interior level, written as LeafNodes
void launch_expand_write_count(cf_ctx *ctx, const SideDesc *d_desc, const Node *nodes, size_t nn, const u64 *off,
                               size_t nout, int depth, LeafNode *out, u32 *cnt);
// nodes (roots of radius 1) as LeafNodes
void launch_leaf_nodes(cf_ctx *ctx, const Node *nodes, size_t nn, LeafNode *out);
// records (key, id) of leaves [lo, lo + n) of a last level, record j written at j - lo (leaf tiles
// with their parents staged in LDS)
void leaf_records(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                  int depth, u64 *keys, u64 *ids);
void launch_gen_rank(cf_ctx *ctx, const SideDesc *d_desc, u64 lo, size_t count, u64 *keys, u64 *ids);
// the same records with runs sharing their first r - 1 indices computed once (one wave per 64
// consecutive prefixes); rank range checked against C(n, r)
void launch_gen_prefix(cf_ctx *ctx, const SideDesc *d_desc, int n, int r, u64 lo, size_t count, u64 *keys, u64 *ids);
// per first index x: lo[x], hi[x] = leaf offsets where its nodes start and end (zeroed by the caller)
void launch_first_bounds(cf_ctx *ctx, const LeafNode *nodes, const u64 *off, size_t nn, u64 *lo, u64 *hi);
u64 max_u64(cf_ctx *ctx, const u64 *a, size_t n);   // synchronous

// radix-partitioned join (k_part.hip)
void fill_u64(cf_ctx *ctx, u64 *a, size_t n, u64 v);
void fill_u32(cf_ctx *ctx, u32 *a, size_t n, u32 v);
void part_bucket_count(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                       int depth, const KeyHash &kh, u32 *bcount);
// DR: the n D records in bucket order (leaf index relative to lo | fingerprint << 40); T:
// CF_TABLE_SLOTS u16 slots per bucket, zeroed
void part_bucket_scatter(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                         int depth, const KeyHash &kh, const u64 *bstart, u32 *bcur, void *DR, unsigned short *T);
// validated side specs of an instance (the preconditions of cf_plan_create_d / _i); throw on a violation
// the modulus an instance's keys live in: key_M = M for M < 2^64, else a divisor of M below 2^64
// (wide: the join's pairs are then checked modulo M); mult = P0 mod key_M
struct KeyMod {
    unsigned __int128 M = 1, P0 = 0;
    u64 key_M = 1, mult = 0;
    bool wide = false;
};
KeyMod key_mod(const cf_instance &in);
SideSpec spec_of_d(const cf_instance &in, u64 koff, const KeyMod *km = nullptr);
SideSpec spec_of_i(const cf_instance &in, u64 koff, const KeyMod *km = nullptr);

// Where one chunk's I records live: nseg segments, partition-major (nblk per partition); segment g
// holds fill[g] records from slot O[g] (exact layout) or g << capl (capacity layout, O null).
struct SegLayout {
    const u64 *O;
    const u32 *fill;
    u64 nseg;
    u32 nblk;
    int capl;
    const u32 *hi;   // I side: the records' leaf indices, apart from their words (null: {word, leaf} pairs)
};
struct IChunkLayout {   // the buffers behind a SegLayout
    DBuf<u64> O;
    DBuf<u32> fill;
    u64 nseg = 0;
    u32 nblk = 0;
    int capl = 0;
    u32 *hi = nullptr;
    SegLayout view() const { return SegLayout{O.p, fill.p, nseg, nblk, capl, hi}; }
};
// slots a chunk of n leaves needs in the capacity layout
// launch shapes and geometry targets that change speed, never results (cf_tuning, 0 = default)
struct Tuning {
    int slice_kib = 0;       // probe-table slice per partition (default 2048)
    int d_per_bucket = 0;    // D records per bucket (default 2)
    int seg_fill_pct = 0;    // expected fill of a capacity segment, percent (default 70)
    int probe_blocks = 0;    // probe grid, 1024-thread blocks (default 256: one per CU)
};
u64 part_cap_slots(u64 n, const KeyHash &kh, const Tuning &tn);
// one pass into fixed-size segments; *ovf set if a segment overflowed (then use the exact layout);
// a chunk whose segments would number 2^32 slots or more is laid out exactly at once (and
// part_cap_slots returns n for it)
// tp receives the chunk's tile owner map (2 entries per 2048 leaves), which the exact layout and
// pass 2 take as given
// Records are {bucket word, leaf} pairs, or, with hi_at (I side only), two arrays of 32-bit words:
// bucket words from rec on, leaf indices from word hi_at on (rec holding at least 2 * hi_at words);
// the probe reads only the first. dside: the D side's records (the same records; a kernel of its own,
// so that profiles tell the sides apart).
void part_scatter_cap(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo, u64 n,
                      int depth, const KeyHash &kh, const Tuning &tn, IChunkLayout &lay, u64 *rec, u32 *ovf,
                      DBuf<u64> &tp, bool dside = false, u64 hi_at = 0);
// count, scan and scatter: n records exactly, in partition order
void part_scatter_exact(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo,
                        u64 n, int depth, const KeyHash &kh, IChunkLayout &lay, u64 *rec, const u64 *tp,
                        bool dside = false, u64 hi_at = 0);
// D side bucketed from its partitioned records: windows of 2^wb buckets within each partition
// (dwin_geometry: false if a partition would need more than 256); part_d_windows moves the records
// into fixed-capacity window regions (wreg: P * R * dwin_capacity() records of 8 bytes; *ovf set if a
// window overflowed); after the caller's exclusive scan of wcount into wbase (P * R + 1 entries),
// part_d_sort writes bstart (NB + 1), DR in bucket order (D leaf index | fingerprint << 40: a
// record's exact key is recomputed from its leaf when a candidate with its fingerprint meets it),
// the table slots T (every bucket) and the filter F (zeroed, FB bits per bucket);
// kerr gets KERR_DWINDOW if a window holds a record outside its buckets (inconsistent input; the
// window's buckets are then written empty)
constexpr u32 KERR_DWINDOW = 1;   // a D record outside its window's buckets
constexpr u32 KERR_CAND = 2;      // more candidates reserved than the candidate buffer holds
bool dwin_geometry(const KeyHash &kh, u64 nD, int &wb, u32 &R);
u64 dwin_capacity();
void part_d_windows(cf_ctx *ctx, const KeyHash &kh, int wb, u32 R, const SegLayout &S, const u64 *rec,
                    u32 *wcount, void *wreg, u32 *ovf);
void part_d_sort(cf_ctx *ctx, const KeyHash &kh, int wb, u32 R, const u32 *wcount, const u64 *wbase,
                 const void *wreg, u64 *bstart, void *DR, unsigned short *T, u32 *F, int FB, u32 *kerr);
// the D leaves of a radius group, for recomputing a pair's D id in pass 2
struct DLeaves {
    const SideDesc *descs;
    const LeafNode *nodes;
    size_t nn;
    const u64 *off;
};
// One D table as pass 2 reads it: bucket starts, records, and the D leaves [lo, lo + n) its records
// refer to (relative to lo), with their tile owner map (or null).
struct DTabDev {
    const u64 *bstart;
    const u64 *DR;
    u64 lo, n;
    const u64 *tp;
};
// n table descriptors into device memory (by kernels whose arguments carry them, so a captured solve
// replays them)
void set_dtabs(cf_ctx *ctx, const DTabDev *h, size_t n, DTabDev *d);
// xq: 8 u32 queue counters; ncand: 1 counter; cand: ccap slots
// rec: 8-byte I records (low word bucket offset << fpb | fingerprint, high word leaf index)
// bucket filter of a D table: bits per bucket for the partition geometry (0: none fits in LDS), its
// size in 32-bit words, and its construction from the table slots and bucket starts
int filter_bits(const KeyHash &kh);
// candidate slots the probe may leave unused (added to the candidate buffer)
u64 cand_margin(const Tuning &tn);
u64 filter_words(const KeyHash &kh, int FB);
void part_filter(cf_ctx *ctx, const u64 *DR, const u64 *bstart, const KeyHash &kh, int FB, u32 *F);
// candidates are the records' slots tagged with the table (slot | tag << 32); reset: *ncand cleared
// first (the first table of a chunk), else the candidates follow those already counted
void part_probe_pass1(cf_ctx *ctx, const KeyHash &kh, const SegLayout &S, const u64 *rec, const unsigned short *T,
                      const u32 *F, int FB, u32 *xq, unsigned long long *ncand, u64 ccap, u64 *cand,
                      const Tuning &tn, u32 tag, bool reset);
void part_probe_pass2(cf_ctx *ctx, const SideDesc *descs, const LeafNode *nodes, size_t nn, const u64 *off, u64 lo,
                      int depth, const KeyHash &kh, const SegLayout &S, const u64 *rec, const u64 *tp,
                      const DLeaves &dl, const DTabDev *tabs, const unsigned long long *ncand, u64 ccap,
                      const u64 *cand, const u64 *koffs, int ninst, unsigned long long *npairs, u64 cap, u64 *od,
                      u64 *oi, u64 *oinst, u32 *kerr);
// moduli of 2^64 and above: the instance is joined modulo wide_divisor(M) (the largest divisor of M
// in [2^40, 2^62]) and its pairs are checked exactly modulo M afterwards (k_wide.hip)
struct WideMod {
    u64 M_lo, M_hi;    // M
    u64 P_lo, P_hi;    // P0 mod M
    u64 on;            // 1: the instance's M is 2^64 or more; 0: its pairs are kept as they are
};
u64 wide_divisor(unsigned __int128 M);
// keep[e] = pair e is exact modulo its instance's M; the kept pairs are written in order to od, oi,
// oj and their count to *count (pos: n + 1 words of scan)
void wide_check_pairs(cf_ctx *ctx, const u64 *pd, const u64 *pi, const u64 *pj, u64 n, const SideDesc *dd,
                      const SideDesc *id, const WideMod *wm, u32 *keep, u64 *pos, u64 *od, u64 *oi, u64 *oj,
                      unsigned long long *count);
// per-instance leaf totals of a frontier added into out[0, ninst) (nodes of an instance are contiguous)
void inst_bounds(cf_ctx *ctx, const LeafNode *nodes, const u64 *off, size_t nn, int ninst, u64 *out);

// joins
// equal-key join of key-sorted sides: pairs per span of I rows (scnt[join_spans(nI)]), their scan
// (soff, one more word), then the pairs; returns the pair total
size_t join_spans(size_t nI);
u64 join_sorted_count(cf_ctx *ctx, const u64 *Dk, size_t nD, const u64 *Ik, size_t nI, u64 *scnt, u64 *soff);
void join_sorted_write(cf_ctx *ctx, const u64 *Dk, size_t nD, const u64 *Ik, const u64 *Did, const u64 *Iid, size_t nI,
                       const u64 *soff, u64 *od, u64 *oi, u64 *okey);
