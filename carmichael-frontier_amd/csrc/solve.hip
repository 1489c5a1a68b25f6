// solve.hip -- whole-instance and batched solves (host orchestration of the gfx950 kernels).
//
// Partitioned solve (default) of one or many instances (the device half of
// gpu/mitm_gpu_mstar_driver.py:184-243 for every instance of a batch):
//   1. D side of every instance: enumeration tree expanded level by level for all instances of
//      one radius at once; instances live in disjoint key ranges [koff, koff + M) of one key
//      space, so cross-instance keys never collide;
//   2. D records counting-sorted into fine key buckets (k_part.hip);
//   3. I side expanded likewise; its leaves, in chunks, are counted per coarse partition,
//      scattered by partition and probed against the D buckets of their partition.
// Sort-merge solve (CF_SOLVE_SORTMERGE): the reference's shape -- I chunks materialized, radix
// sorted and joined against the sorted D side with the searchsorted semantics of
// gpu/mitm_gpu2.py:107-119.
#include <algorithm>
#include <cmath>
#include <array>
#include <map>

#include "cf_internal.h"

typedef unsigned __int128 u128;

namespace {

struct Frontier {
    DBuf<LeafNode> nodes;   // the last interior level
    size_t nn = 0;
    DBuf<u32> cnt;
    DBuf<u64> off;
    u64 total = 0;
};

// Values the host reads back from the device during a solve (level sizes, overflow flags,
// candidate and pair counts) decide allocation sizes and control flow. A tape records them in an
// eager solve and plays them back while the same solve is captured into a HIP graph, where nothing
// may be read back; the solve is deterministic, so the captured work is the recorded work.
struct Tape {
    enum Mode { OFF, RECORD, REPLAY } mode = OFF;
    std::vector<u64> v;
    size_t pos = 0;
    bool replay() const { return mode == REPLAY; }
};

u64 host_value(cf_ctx *ctx, Tape *t, const u64 *dev) {
    if (t && t->replay()) {
        CF_REQUIRE(t->pos < t->v.size(), "captured solve asked for more values than were recorded");
        return t->v[t->pos++];
    }
    const u64 x = d2h_u64(ctx, dev);
    if (t && t->mode == Tape::RECORD) t->v.push_back(x);
    return x;
}
u64 host_value_u32(cf_ctx *ctx, Tape *t, const u32 *dev) {
    if (t && t->replay()) return host_value(ctx, t, nullptr);
    u32 x = 0;
    d2h(ctx, &x, dev, 1);
    if (t && t->mode == Tape::RECORD) t->v.push_back(x);
    return x;
}

// Expand the given depth-0 roots (one radius r, on the device) to the last level.
Frontier expand_to_last(cf_ctx *ctx, const SideDesc *d_desc, const Node *roots, size_t nroots, int r, Tape *tape) {
    Frontier f;
    if (!nroots) return f;
    DBuf<Node> cur;
    const Node *curp = roots;
    size_t ncur = nroots;
    // the roots' child counts; every level below gets its counts from the pass that writes it; the
    // last interior level is written as LeafNodes
    DBuf<u32> c(ctx, ncur);
    launch_expand_count(ctx, d_desc, curp, ncur, 0, c.p);
    DBuf<LeafNode> last;
    for (int depth = 0; depth + 1 < r; ++depth) {
        DBuf<u64> o(ctx, ncur + 1);
        scan_u32_to_u64_async(ctx, c.p, o.p, ncur);
        const u64 tot = host_value(ctx, tape, o.p + ncur);
        if (tot == 0) return f;
        DBuf<u32> cn(ctx, tot);
        if (depth + 2 < r) {
            DBuf<Node> nx(ctx, tot);
            launch_expand_write_count(ctx, d_desc, curp, ncur, o.p, tot, depth, nx.p, cn.p);
            cur = std::move(nx);
            curp = cur.p;
        } else {
            last = DBuf<LeafNode>(ctx, tot);
            launch_expand_write_count(ctx, d_desc, curp, ncur, o.p, tot, depth, last.p, cn.p);
        }
        c = std::move(cn);
        ncur = tot;
    }
    if (!last.p) {   // r == 1: the roots are the last level
        last = DBuf<LeafNode>(ctx, ncur);
        launch_leaf_nodes(ctx, roots, ncur, last.p);
    }
    f.cnt = std::move(c);
    f.off = DBuf<u64>(ctx, ncur + 1);
    scan_u32_to_u64_async(ctx, f.cnt.p, f.off.p, ncur);
    f.total = host_value(ctx, tape, f.off.p + ncur);
    f.nodes = std::move(last);
    f.nn = ncur;
    return f;
}

Node root_node(const SideDesc &sd, int side, int inst) {
    Node nd{};
    nd.A_lo = 1;
    nd.A_hi = 0;
    nd.xm = sd.xm0;
    nd.x2 = sd.x20;
    nd.id = 0;
    const int nxt = side == SIDE_I ? 0 : sd.n - 1;
    nd.pos = (u32)(unsigned short)(short)nxt;
    nd.meta = (u32)inst << 8;
    return nd;
}

}  // namespace

// The modulus the keys of an instance live in: M itself below 2^64, else its largest divisor in
// [2^40, 2^62] (the pairs are then checked exactly modulo M after the join).
KeyMod key_mod(const cf_instance &in) {
    KeyMod k;
    k.M = ((u128)in.M_hi << 64) | in.M;
    CF_REQUIRE(k.M >= 1 && k.M < ((u128)1 << 126), "need 1 <= M < 2^126");
    k.P0 = (((u128)in.P0modM_hi << 64) | in.P0modM) % k.M;
    k.wide = in.M_hi != 0;
    k.key_M = k.wide ? wide_divisor(k.M) : in.M;
    k.mult = k.wide ? (u64)(k.P0 % k.key_M) : in.P0modM;
    return k;
}

// Validated side specs of one instance (the preconditions of cf_plan_create_d / _i).
SideSpec spec_of_d(const cf_instance &in, u64 koff, const KeyMod *km) {
    CF_REQUIRE(in.nB >= 1 && in.nB <= 64 && in.r >= 1 && in.r <= 16, "need nB <= 64 base primes and 1 <= r <= 16");
    const KeyMod k = km ? *km : key_mod(in);
    for (int i = 1; i < in.nB; ++i) CF_REQUIRE(in.B[i - 1] < in.B[i], "base must be strictly ascending");
    CF_REQUIRE(in.d_id_mode == ID_PACK_DESC || in.d_id_mode == ID_MASK, "D ids are 6-bit packed or deletion masks");
    if (in.d_id_mode == ID_PACK_DESC) CF_REQUIRE(in.r <= 10, "6-bit packed ids need r <= 10");
    u128 w = 1;
    for (int j = 0; j < in.r && j < in.nB; ++j) {
        const u64 q = in.B[in.nB - 1 - j];
        CF_REQUIRE(w <= (~(u128)0) / q, "products of r base primes exceed 128 bits");
        w *= q;
    }
    SideSpec s;
    s.side = SIDE_D;
    s.primes.assign(in.B, in.B + in.nB);
    s.r = in.r;
    s.M = k.key_M;
    s.mult = 1;
    s.cap = ((u128)in.dmin_hi << 64) | in.dmin_lo;
    s.id_mode = in.d_id_mode;
    s.id_bits = 6;
    s.koff = koff;
    return s;
}

SideSpec spec_of_i(const cf_instance &in, u64 koff, const KeyMod *km) {
    CF_REQUIRE(in.r >= 1 && in.r <= 16 && in.nI >= 0 && in.nI <= 1024, "need 1 <= r <= 16 and n <= 1024");
    const KeyMod k = km ? *km : key_mod(in);
    for (int i = 1; i < in.nI; ++i) CF_REQUIRE(in.INS[i - 1] < in.INS[i], "INS must be strictly ascending");
    CF_REQUIRE(in.i_id_mode == ID_COLEX || in.i_id_mode == ID_PACK_ASC, "I ids are colex ranks or 8-bit packed");
    if (in.i_id_mode == ID_PACK_ASC) CF_REQUIRE(in.nI <= 256 && in.r <= 8, "8-bit packed ids need n <= 256 and r <= 8");
    if (in.i_id_mode == ID_COLEX) {
        u128 c = 1;
        bool over = false;
        for (int j = 0; j < in.r && j < in.nI; ++j) {
            c = c * (u128)(in.nI - j) / (u128)(j + 1);
            if (c > (u128)~0ULL) over = true;
        }
        CF_REQUIRE(!over, "rank ids would overflow u64 (C(n, r) >= 2^64)");
    }
    SideSpec s;
    s.side = SIDE_I;
    s.primes.assign(in.INS, in.INS + in.nI);
    s.r = in.r;
    s.M = k.key_M;
    s.mult = k.mult;
    s.cap = ((u128)in.icap_hi << 64) | in.icap_lo;
    s.split_idx = in.split_idx;
    s.hmax = in.hmax;
    s.id_mode = in.i_id_mode;
    s.id_bits = 8;
    s.koff = koff;
    return s;
}

namespace {

struct BatchOut {
    DBuf<u64> pd, pi, pinst;          // pairs: D id, I id, instance
    DBuf<u64> dcount, icount;         // records per instance
    DBuf<u64> cnt_pairs;              // pairs found (may exceed the capacity of pd)
    DBuf<u32> kerr;
    u64 npairs = 0;
    std::vector<u64> nD, nI;
};

struct PhaseClock {
    cf_ctx *ctx;
    std::vector<hipEvent_t> ev;
    explicit PhaseClock(cf_ctx *c) : ctx(c) {}
    void mark() {
        hipEvent_t e;
        CF_HIP(hipEventCreate(&e));
        CF_HIP(hipEventRecord(e, ctx->stream));
        ev.push_back(e);
    }
    double ms(size_t a, size_t b) {
        float t = 0;
        CF_HIP(hipEventSynchronize(ev[b]));
        CF_HIP(hipEventElapsedTime(&t, ev[a], ev[b]));
        return t;
    }
    ~PhaseClock() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// Bucket geometry of one radius group (keyhash.h: kh_choose).
struct Partitioning {
    KeyHash kh{};
    u64 NB = 0;
    int P() const { return 1 << kh.lgP; }
};

Partitioning assign_buckets(u64 nD, const Tuning &tn) {
    Partitioning pt;
    const u64 slice = tn.slice_kib > 0 ? (u64)tn.slice_kib << 10 : (u64)2 << 20;
    CF_REQUIRE(kh_choose(nD, pt.kh, tn.d_per_bucket, slice),
               "more than 2^24 buckets per partition for one radius group: split the batch");
    pt.NB = kh_nb(pt.kh);
    return pt;
}

}  // namespace

// A prepared batch: every instance validated, its exact bound tables computed on the host and
// uploaded in one allocation; solving it is device work only.
struct BatchGraph;
struct cf_batch {
    cf_ctx *ctx = nullptr;
    int n = 0;
    std::vector<u64> koff;
    std::vector<SideDesc> dd, id;
    std::vector<bool> i_live, d_live;
    std::map<int, std::vector<int>> groups;
    DBuf<u64> store;
    DBuf<SideDesc> d_desc, i_desc;
    DBuf<u64> koffs;
    std::vector<WideMod> wide;                   // per instance: M and P0 mod M when M >= 2^64
    int nwide = 0;
    DBuf<WideMod> wide_dev;
    std::map<int, DBuf<Node>> droots, iroots;    // per radius: the enumeration roots of live sides
    std::map<int, size_t> nd_roots, ni_roots;
    u64 pair_cap = 1 << 16;                      // pair capacity of the next solve (grows to fit)
    Tuning tune;                                 // cf_batch_set_tuning
    BatchGraph *graph = nullptr;                 // the captured solve, once cf_batch_capture ran
    // side streams (default priority) and fork/join events of the captured solve: each radius
    // group is captured on its own branch, so the replayed groups run side by side
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> cap_events;
    ~cf_batch();
};

namespace {

void batch_prepare(cf_batch *b, const cf_instance *insts, int n) {
    cf_ctx *ctx = b->ctx;
    b->n = n;
    b->koff.resize(n);
    // key offsets: instance j owns keys [koff_j, koff_j + M_j) (M_j the modulus of its keys); the
    // union must stay below 2^64 - 1
    std::vector<KeyMod> km(n);
    b->wide.assign(n, WideMod{});
    b->nwide = 0;
    u128 K = 0;
    for (int j = 0; j < n; ++j) {
        km[j] = key_mod(insts[j]);
        b->koff[j] = (u64)K;
        K += km[j].key_M;
        if (km[j].wide) {
            b->wide[j] = WideMod{(u64)km[j].M, (u64)(km[j].M >> 64), (u64)km[j].P0, (u64)(km[j].P0 >> 64), 1};
            b->nwide += 1;
        }
    }
    CF_REQUIRE(K < (u128)~0ULL, "the batch's moduli sum to 2^64 or more: split the batch");
    b->dd.resize(n);
    b->id.resize(n);
    b->i_live.assign(n, true);
    b->d_live.assign(n, true);
    std::vector<u64> buf;
    std::vector<std::array<size_t, 4>> offd(n), offi(n);
    for (int j = 0; j < n; ++j) {
        SideSpec sd = spec_of_d(insts[j], b->koff[j], &km[j]);
        SideSpec si = spec_of_i(insts[j], b->koff[j], &km[j]);
        side_tables_host(sd, b->dd[j], buf, offd[j].data());
        side_tables_host(si, b->id[j], buf, offi[j].data());
        // the 32-bit scaled leaf step needs every prime of the instance (both sides) below 2^32
        bool small = true;
        for (u64 q : sd.primes) small &= q < (1ULL << 32);
        for (u64 q : si.primes) small &= q < (1ULL << 32);
        b->dd[j].leaf32 = b->id[j].leaf32 = small ? 1 : 0;
        if (si.cap <= 1 || (int)si.primes.size() < si.r) b->i_live[j] = false;
        if ((int)sd.primes.size() < sd.r) b->d_live[j] = false;
        b->groups[insts[j].r].push_back(j);
    }
    b->store = DBuf<u64>(ctx, buf.size() ? buf.size() : 1);
    h2d(ctx, b->store.p, buf.data(), buf.size());
    for (int j = 0; j < n; ++j) {
        side_tables_bind(b->dd[j], b->store.p, offd[j].data());
        side_tables_bind(b->id[j], b->store.p, offi[j].data());
    }
    b->d_desc = DBuf<SideDesc>(ctx, n);
    b->i_desc = DBuf<SideDesc>(ctx, n);
    b->koffs = DBuf<u64>(ctx, n);
    h2d(ctx, b->d_desc.p, b->dd.data(), n);
    h2d(ctx, b->i_desc.p, b->id.data(), n);
    h2d(ctx, b->koffs.p, b->koff.data(), n);
    if (b->nwide) {
        b->wide_dev = DBuf<WideMod>(ctx, n);
        h2d(ctx, b->wide_dev.p, b->wide.data(), n);
    }
    // enumeration roots, resident for every solve of the batch
    for (auto &g : b->groups) {
        std::vector<Node> rd, ri;
        for (int j : g.second) {
            if (b->d_live[j]) rd.push_back(root_node(b->dd[j], SIDE_D, j));
            if (b->i_live[j]) ri.push_back(root_node(b->id[j], SIDE_I, j));
        }
        b->nd_roots[g.first] = rd.size();
        b->ni_roots[g.first] = ri.size();
        b->droots[g.first] = DBuf<Node>(ctx, rd.size() ? rd.size() : 1);
        b->iroots[g.first] = DBuf<Node>(ctx, ri.size() ? ri.size() : 1);
        h2d(ctx, b->droots[g.first].p, rd.data(), rd.size());
        h2d(ctx, b->iroots[g.first].p, ri.data(), ri.size());
    }
    CF_HIP(hipStreamSynchronize(ctx->stream));
}

// I leaves per chunk: as many as fit comfortably (the record index in a chunk is 32 bits); one
// chunk lets each XCD's probe stay inside one partition for a long stretch
u64 default_chunk() {
    size_t free_b = 0, total_b = 0;
    CF_HIP(hipMemGetInfo(&free_b, &total_b));
    return std::min<u64>((u64)1 << 31, std::max<u64>((u64)free_b / 48, (u64)1 << 24));
}

// One table of D records: the D leaves [lo, lo + n) of a radius group bucketed by hash (bucket
// starts, records in bucket order, slots, filter) for the probe.
struct DTable {
    u64 lo = 0, n = 0;
    DBuf<u64> bstart, DR;
    DBuf<unsigned short> T;
    DBuf<u32> F;
    DBuf<u64> tp;     // tile owner map of the range's leaves (window build), or empty
};

// The D leaves of a group in nr ranges of at most m records: one range while the window build takes
// the whole side (fewer than 2^28 records, whose leaf indices leave 4 bits of a 32-bit word to the
// hash bits the records carry, and at most 256 windows per partition), else the fewest
// equal ranges it takes one by one. The global-atomic build (d_part false) takes one range of any size.
void d_ranges(u64 nD, const Tuning &tn, bool d_part, u64 &nr, u64 &m) {
    nr = 1;
    m = nD;
    if (!d_part || nD == 0) return;
    for (;;) {
        m = (nD + nr - 1) / nr;
        const Partitioning q = assign_buckets(m, tn);
        int wb = 0;
        u32 R = 0;
        if (m < ((u64)1 << 28) && dwin_geometry(q.kh, m, wb, R)) return;
        if (nr >= 65536) break;
        nr = nr < 8 ? nr + 1 : nr + nr / 8;
    }
    nr = 1;   // no range size fits: the global-atomic build over the whole side
    m = nD;
}

// Builds the table of D leaves [lo, lo + m): partitioned like the I side (keys alongside), moved into
// bucket windows and counting-sorted one window per block in LDS; by per-record global atomics when
// the window build does not fit (d_part false, or a window overflowing on heavily repeated keys).
// DR's leaf references are relative to lo.
void build_dtable(cf_ctx *ctx, const SideDesc *d_desc, const Frontier &fd, int depth, const Partitioning &pt, int FB,
                  bool d_part, u64 lo, u64 m, const Tuning &tn, DTable &t, u32 *ovf, u32 *kerr, Tape *tape) {
    const KeyHash &kh = pt.kh;
    CF_REQUIRE(m < ((u64)1 << 40), "a D table holds fewer than 2^40 records (DR keeps 40-bit leaf indices)");
    t.lo = lo;
    t.n = m;
    t.bstart = DBuf<u64>(ctx, pt.NB + 1);
    t.DR = DBuf<u64>(ctx, m ? m : 1);
    t.T = DBuf<unsigned short>(ctx, pt.NB * CF_TABLE_SLOTS);
    t.F = DBuf<u32>(ctx, std::max<u64>(filter_words(kh, FB), 1));
    if (!m) return;
    fill_u32(ctx, t.F.p, filter_words(kh, FB), 0);
    int wb = 0;
    u32 R = 0;
    if (d_part && m < ((u64)1 << 28) && dwin_geometry(kh, m, wb, R)) {
        IChunkLayout dl;
        const u64 slots = std::max<u64>(part_cap_slots(m, kh, tn), m);
        DBuf<u64> drec(ctx, slots);
        part_scatter_cap(ctx, d_desc, fd.nodes.p, fd.nn, fd.off.p, lo, m, depth, kh, tn, dl, drec.p, ovf, t.tp, true);
        if (host_value_u32(ctx, tape, ovf))
            part_scatter_exact(ctx, d_desc, fd.nodes.p, fd.nn, fd.off.p, lo, m, depth, kh, dl, drec.p, t.tp.p, true);
        const u64 nw = (u64)pt.P() * R;
        DBuf<u32> wcount(ctx, nw);
        DBuf<u64> wreg(ctx, nw * dwin_capacity()), wbase(ctx, nw + 1);
        part_d_windows(ctx, kh, wb, R, dl.view(), drec.p, wcount.p, wreg.p, ovf);
        if (!host_value_u32(ctx, tape, ovf)) {
            scan_u32_to_u64_async(ctx, wcount.p, wbase.p, nw);
            part_d_sort(ctx, kh, wb, R, wcount.p, wbase.p, wreg.p, t.bstart.p, t.DR.p, t.T.p, t.F.p, FB, kerr);
            return;
        }
        t.tp.reset();
    }
    fill_u32(ctx, (u32 *)t.T.p, pt.NB * CF_TABLE_SLOTS / 2, 0);
    DBuf<u32> cnt(ctx, pt.NB);
    fill_u32(ctx, cnt.p, pt.NB, 0);
    part_bucket_count(ctx, d_desc, fd.nodes.p, fd.nn, fd.off.p, lo, m, depth, kh, cnt.p);
    scan_u32_to_u64_async(ctx, cnt.p, t.bstart.p, pt.NB);
    fill_u32(ctx, cnt.p, pt.NB, 0);   // reused as the scatter cursors
    part_bucket_scatter(ctx, d_desc, fd.nodes.p, fd.nn, fd.off.p, lo, m, depth, kh, t.bstart.p, cnt.p, t.DR.p, t.T.p);
    part_filter(ctx, t.DR.p, t.bstart.p, kh, FB, t.F.p);
}

// One radius group's branch of a captured solve: while it lives, launches go to the group's side
// stream (forked from the capturing stream) and blocks the group frees are reused only on that
// branch; on leaving it records the group's completion event, which the capturing stream waits on.
// Declared first in the group's scope, so the group's buffers are freed while the branch is set.
// Inactive (no fork) for an eager solve.
struct GroupBranch {
    cf_ctx *ctx;
    hipStream_t main = nullptr;
    hipEvent_t done = nullptr;
    GroupBranch(cf_ctx *c, hipStream_t side, hipEvent_t fork, hipEvent_t done_ev, int branch) : ctx(c) {
        if (!side) return;
        main = ctx->stream;
        done = done_ev;
        CF_HIP(hipStreamWaitEvent(side, fork, 0));
        ctx->stream = side;
        ctx->arena_branch = branch;
    }
    ~GroupBranch() {
        if (!main) return;
        (void)hipEventRecord(done, ctx->stream);
        ctx->stream = main;
        ctx->arena_branch = 0;
    }
};

// tape: null for a plain eager solve; RECORD to note every value read back; REPLAY while the
// stream is being captured (no value is read back and no timing is taken)
void solve_part(cf_batch *b, BatchOut &out, cf_stats *st, u64 max_chunk, bool d_part = true, Tape *tape = nullptr) {
    cf_ctx *ctx = b->ctx;
    const int n = b->n;
    const auto &groups = b->groups;
    const DBuf<SideDesc> &d_desc = b->d_desc;
    const DBuf<SideDesc> &i_desc = b->i_desc;
    const DBuf<u64> &koffs = b->koffs;
    const bool capturing = tape && tape->replay();

    PhaseClock clk(ctx);
    if (!capturing) clk.mark();  // 0
    out.dcount = DBuf<u64>(ctx, n);
    out.icount = DBuf<u64>(ctx, n);
    DBuf<u64> &dcount = out.dcount, &icount = out.icount;
    const u64 chunk = max_chunk ? max_chunk : default_chunk();
    CF_REQUIRE(chunk <= ((u64)1 << 32), "max_chunk above 2^32 leaves");
    u64 cap = b->pair_cap;
    out.cnt_pairs = DBuf<u64>(ctx, 1);
    out.kerr = DBuf<u32>(ctx, 1);           // set by a kernel that met an inconsistent record
    DBuf<u64> &cnt_pairs = out.cnt_pairs;
    DBuf<u32> &kerr = out.kerr;
    fill_u32(ctx, kerr.p, 1, 0);
    double ms_d = 0, ms_i = 0;
    u64 parents_d = 0, parents_i = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        out.pd = DBuf<u64>(ctx, cap);
        out.pi = DBuf<u64>(ctx, cap);
        out.pinst = DBuf<u64>(ctx, cap);
        fill_u64(ctx, cnt_pairs.p, 1, 0);
        fill_u64(ctx, dcount.p, n, 0);
        fill_u64(ctx, icount.p, n, 0);
        ms_d = ms_i = 0;
        parents_d = parents_i = 0;
        // each radius group: its D side is bucketed into its own probe table, then its I side is
        // generated, partitioned and probed against it. Eager: one group after another (largest
        // radius first). Captured:
        // each group on its own branch (side stream), all forked from and joined back to the
        // capturing stream, so a replay runs the groups side by side.
        const bool fork = capturing && groups.size() > 1;
        hipEvent_t ev_fork = nullptr;
        std::vector<hipEvent_t> ev_done;
        if (fork) {
            const size_t ns = std::min<size_t>(groups.size(), 4);
            while (b->side.size() < ns) {
                hipStream_t st = nullptr;
                CF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
                b->side.push_back(st);
            }
            auto mk = [&]() {
                hipEvent_t e = nullptr;
                CF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                b->cap_events.push_back(e);
                return e;
            };
            ev_fork = mk();
            for (size_t k = 0; k < groups.size(); ++k) ev_done.push_back(mk());
            CF_HIP(hipEventRecord(ev_fork, ctx->stream));
        }
        size_t gi = 0;
        // largest radius first: a replay's kernels are submitted branch after branch, so the group
        // with the most work (the largest radius in the family batches) starts at once instead of
        // after the submission of every other group's launches
        for (auto git = groups.rbegin(); git != groups.rend(); ++git) {
            auto &g = *git;
            const size_t si = gi % std::max<size_t>(b->side.size(), 1);
            GroupBranch branch(ctx, fork ? b->side[si] : nullptr, ev_fork, fork ? ev_done[gi] : nullptr, (int)si + 1);
            ++gi;
            DBuf<u32> xq(ctx, 8), ovf(ctx, 1);
            DBuf<u64> flags(ctx, 1);                // candidate count of the last probe
            PhaseClock gc(ctx);
            if (!capturing) gc.mark();
            Frontier fd = expand_to_last(ctx, d_desc.p, b->droots.at(g.first).p, b->nd_roots.at(g.first), g.first, tape);
            const u64 nD = fd.total;
            if (nD) inst_bounds(ctx, fd.nodes.p, fd.off.p, fd.nn, n, dcount.p);
            // D side: ranges of D leaves, each bucketed into a table of its own; every range has the
            // same bucket geometry, so an I record's partition, bucket and fingerprint hold for all
            u64 nr = 1, m_rng = nD;
            d_ranges(nD, b->tune, d_part, nr, m_rng);
            const Partitioning pt = assign_buckets(m_rng, b->tune);
            const KeyHash &kh = pt.kh;
            // bucket filter: the probe tests it in LDS before reading the table
            const int FB = nD ? filter_bits(kh) : 0;
            std::vector<DTable> tabs(nD ? nr : 0);
            for (u64 k = 0; k < tabs.size(); ++k) {
                const u64 lo = k * m_rng;
                build_dtable(ctx, d_desc.p, fd, g.first - 1, pt, FB, d_part, lo, std::min<u64>(m_rng, nD - lo), b->tune,
                             tabs[k], ovf.p, kerr.p, tape);
            }
            // the tables as pass 2 reads them
            DBuf<DTabDev> dtabs(ctx, std::max<size_t>(tabs.size(), 1));
            {
                std::vector<DTabDev> h(tabs.size());
                for (size_t k = 0; k < tabs.size(); ++k)
                    h[k] = DTabDev{tabs[k].bstart.p, tabs[k].DR.p, tabs[k].lo, tabs[k].n, tabs[k].tp.p};
                set_dtabs(ctx, h.data(), h.size(), dtabs.p);
            }
            if (!capturing) gc.mark();
            // ---- I side: expand, then per chunk of leaves: partition (one scatter for every table),
            // probe each table, check the candidates exactly
            Frontier f = expand_to_last(ctx, i_desc.p, b->iroots.at(g.first).p, b->ni_roots.at(g.first), g.first, tape);
            if (f.total) inst_bounds(ctx, f.nodes.p, f.off.p, f.nn, n, icount.p);
            parents_d += fd.nn;
            parents_i += f.nn;
            if (f.total && nD) {
                DBuf<u64> rec;
                DBuf<u64> cand;
                DBuf<u64> tp;   // the chunk's tile owner map
                IChunkLayout lay;
                u64 cap_m = 0, ccap = 0;
                for (u64 L = 0; L < f.total; L += chunk) {
                    const u64 m = std::min<u64>(chunk, f.total - L);
                    const u64 need = std::max<u64>(part_cap_slots(m, kh, b->tune), m);
                    if (cap_m < need) {
                        cap_m = need;
                        rec = DBuf<u64>(ctx, cap_m);
                    }
                    // candidates: fingerprint hits and records of overflowing buckets (~1e-4 of the records)
                    const u64 cwant = std::max<u64>(m >> 5, 1 << 16) + tabs.size() * cand_margin(b->tune);
                    if (ccap < cwant) {
                        ccap = cwant;
                        cand = DBuf<u64>(ctx, ccap);
                    }
                    // one pass into segments sized for the expected share; the probe that follows
                    // is wasted only if a segment overflowed (heavily repeated keys), and then the
                    // chunk is laid out exactly and probed again
                    // records split into bucket words and leaf indices when the chunk is probed against
                    // several tables (every probe then reads 4 bytes a record; the scatter writes its
                    // two streams less efficiently, which one probe does not repay)
                    const u64 split_at = tabs.size() >= 2 ? cap_m : 0;
                    part_scatter_cap(ctx, i_desc.p, f.nodes.p, f.nn, f.off.p, L, m, g.first - 1, kh, b->tune, lay, rec.p,
                                     ovf.p, tp, false, split_at);
                    // every table's candidates into one list (tagged with the table), then one pass 2
                    auto probe = [&](size_t k) {
                        part_probe_pass1(ctx, kh, lay.view(), rec.p, tabs[k].T.p, tabs[k].F.p, FB, xq.p,
                                         (unsigned long long *)flags.p, ccap, cand.p, b->tune, (u32)k, k == 0);
                    };
                    probe(0);
                    if (host_value_u32(ctx, tape, ovf.p)) {
                        // a segment overflowed (heavily repeated keys): the chunk laid out exactly
                        part_scatter_exact(ctx, i_desc.p, f.nodes.p, f.nn, f.off.p, L, m, g.first - 1, kh, lay, rec.p,
                                           tp.p, false, split_at);
                        probe(0);
                    }
                    for (size_t k = 1; k < tabs.size(); ++k) probe(k);
                    u64 nc = host_value(ctx, tape, flags.p);
                    while (nc > ccap) {
                        // more candidates than room (heavily repeated keys): probe again with room for
                        // all. The count includes the unused tails of partial reservations, and which
                        // wave takes which tile varies from run to run. A wave leaves unused fewer slots
                        // than the candidates of the round that makes it reserve again, plus at most one
                        // partial reservation at the end of each probe, so a run over the tables needs at
                        // most 2 * (candidates) + one cand_margin per table, and nc >= candidates: one
                        // rerun suffices; the loop guards it
                        ccap = 2 * nc + tabs.size() * cand_margin(b->tune);
                        cand = DBuf<u64>(ctx, ccap);
                        for (size_t k = 0; k < tabs.size(); ++k) probe(k);
                        nc = host_value(ctx, tape, flags.p);
                    }
                    const DLeaves dleaves{d_desc.p, fd.nodes.p, fd.nn, fd.off.p};
                    part_probe_pass2(ctx, i_desc.p, f.nodes.p, f.nn, f.off.p, L, g.first - 1, kh, lay.view(), rec.p,
                                     tp.p, dleaves, dtabs.p, (const unsigned long long *)flags.p, ccap, cand.p, koffs.p,
                                     n, (unsigned long long *)cnt_pairs.p, cap, out.pd.p, out.pi.p, out.pinst.p, kerr.p);
                }
            }
            if (!capturing) {
                gc.mark();
                ms_d += gc.ms(0, 1);
                ms_i += gc.ms(1, 2);
            }
        }
        for (hipEvent_t e : ev_done) CF_HIP(hipStreamWaitEvent(ctx->stream, e, 0));
        const u64 np = host_value(ctx, tape, cnt_pairs.p);
        const u64 ke = host_value_u32(ctx, tape, kerr.p);
        if (ke & KERR_DWINDOW) throw CfError{CF_EINTERNAL, "a D record fell outside its hash range's buckets"};
        if (ke & KERR_CAND) throw CfError{CF_EINTERNAL, "the probe's candidate list was truncated"};
        out.npairs = np;
        if (np <= cap) break;
        cap = np;
    }
    b->pair_cap = std::max<u64>(b->pair_cap, out.npairs);
    if (b->nwide && out.npairs) {
        // instances of M >= 2^64 were joined modulo a divisor of M: keep the pairs exact modulo M
        const u64 np = out.npairs;
        DBuf<u32> keep(ctx, np);
        DBuf<u64> pos(ctx, np + 1), od(ctx, np), oi(ctx, np), oj(ctx, np);
        wide_check_pairs(ctx, out.pd.p, out.pi.p, out.pinst.p, np, d_desc.p, i_desc.p, b->wide_dev.p, keep.p, pos.p,
                         od.p, oi.p, oj.p, (unsigned long long *)cnt_pairs.p);
        out.pd = std::move(od);
        out.pi = std::move(oi);
        out.pinst = std::move(oj);
        out.npairs = host_value(ctx, tape, cnt_pairs.p);
    }
    if (capturing) return;
    clk.mark();  // 1
    out.nD.assign(n, 0);
    out.nI.assign(n, 0);
    d2h(ctx, out.nD.data(), dcount.p, n);
    d2h(ctx, out.nI.data(), icount.p, n);
    if (st) {
        st->nD = 0;
        st->nI = 0;
        for (int j = 0; j < n; ++j) {
            st->nD += out.nD[j];
            st->nI += out.nI[j];
        }
        st->npairs = out.npairs;
        st->peak_bytes = ctx->peak_live;
        st->ms_total = clk.ms(0, 1);
        st->ms_gen_d = ms_d;            // D expansion, bucketing and probe table
        st->ms_sort_d = 0;
        st->ms_gen_i = ms_i;            // I expansion, partition scatter and probe
        st->ms_sort_i = 0;
        st->ms_join = 0;
        st->nD_parents = parents_d;
        st->nI_parents = parents_i;
    }
}

}  // namespace

// A batch's solve captured into a HIP graph: every buffer its nodes touch lives in the arena
// until the batch is destroyed; out holds the pair and count buffers each replay refills.
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CaptureArena arena;
    BatchOut out;
    u64 npairs = 0;
    u64 chunk = 0;
    // pinned host block a replay's counts are read back into with one synchronization:
    // [pair count, error flags, nD of every instance, nI of every instance]
    u64 *hdr = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;   // around each replay's graph launch
};

static void graph_destroy(cf_ctx *ctx, BatchGraph *G) {
    (void)hipStreamSynchronize(ctx->stream);
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    // the output buffers are arena blocks: hand them back with the rest of the arena
    G->out.pd.release();
    G->out.pi.release();
    G->out.pinst.release();
    G->out.dcount.release();
    G->out.icount.release();
    G->out.cnt_pairs.release();
    G->out.kerr.release();
    arena_release(ctx, &G->arena);
    if (G->hdr) (void)hipHostFree(G->hdr);
    if (G->t0) (void)hipEventDestroy(G->t0);
    if (G->t1) (void)hipEventDestroy(G->t1);
    delete G;
}

cf_batch::~cf_batch() {
    if (graph) graph_destroy(ctx, graph);
    for (hipEvent_t e : cap_events) (void)hipEventDestroy(e);
    for (hipStream_t st : side) (void)hipStreamDestroy(st);
}

extern "C" {

int cf_batch_footprint(cf_batch *b, uint64_t *bytes, uint64_t *nD_total, uint64_t *nI_total) {
    API_BEGIN
    CF_REQUIRE(b && bytes, "NULL argument");
    // Expand both trees of every radius group (counts only) and price the group's two phases: the D
    // build (tables, its temporary records and the D parents) and the I phase (tables, the I records
    // in their capacity layout and the I parents); groups run one after another.
    cf_ctx *ctx = b->ctx;
    u64 peak = 0, sd = 0, si = 0;
    for (auto &g : b->groups) {
        u64 nD = 0, nnD = 0, nI = 0, nnI = 0;
        {
            Frontier fd = expand_to_last(ctx, b->d_desc.p, b->droots.at(g.first).p, b->nd_roots.at(g.first), g.first,
                                         nullptr);
            nD = fd.total;
            nnD = fd.nn;
        }
        {
            Frontier fi = expand_to_last(ctx, b->i_desc.p, b->iroots.at(g.first).p, b->ni_roots.at(g.first), g.first,
                                         nullptr);
            nI = fi.total;
            nnI = fi.nn;
        }
        sd += nD;
        si += nI;
        u64 nr = 1, m = nD;
        d_ranges(nD, b->tune, true, nr, m);
        const Partitioning pt = assign_buckets(m, b->tune);
        const u64 tables = 8 * nD + 30 * pt.NB * nr;        // DR, T, bstart, counters, filter of every range
        // + one range's partitioned D records and window regions + D parents (kept through the I phase
        // for pass 2)
        const u64 d_build = tables + 8 * std::max<u64>(part_cap_slots(m, pt.kh, b->tune), m) + 16 * m + 60 * nnD;
        const u64 i_phase = tables + 8 * std::max<u64>(part_cap_slots(nI, pt.kh, b->tune), nI) +
                            60 * (nnI + nnD) + nI / 4 + nD / 128;
        peak = std::max(peak, std::max(d_build, i_phase));
    }
    *bytes = peak + (64u << 20);
    if (nD_total) *nD_total = sd;
    if (nI_total) *nI_total = si;
    API_END
}

int cf_batch_set_tuning(cf_batch *b, const cf_tuning *t) {
    API_BEGIN
    CF_REQUIRE(b && t, "NULL argument");
    CF_REQUIRE(!b->graph, "the batch is already captured: tune before cf_batch_capture");
    CF_REQUIRE(t->slice_kib >= 0 && t->slice_kib <= (1 << 20) && t->d_per_bucket >= 0 && t->d_per_bucket <= 12 &&
                   t->seg_fill_pct >= 0 && t->seg_fill_pct <= 90 && t->probe_blocks >= 0 && t->probe_blocks <= 65536,
               "tuning values out of range");
    b->tune.slice_kib = t->slice_kib;
    b->tune.d_per_bucket = t->d_per_bucket;
    b->tune.seg_fill_pct = t->seg_fill_pct;
    b->tune.probe_blocks = t->probe_blocks;
    API_END
}

int cf_batch_capture(cf_batch *b, uint64_t max_chunk) {
    API_BEGIN
    CF_REQUIRE(b, "NULL argument");
    CF_REQUIRE(!b->graph, "the batch is already captured");
    cf_ctx *ctx = b->ctx;
    const u64 chunk = max_chunk ? max_chunk : default_chunk();
    {   // settles the pair capacity, so the recorded solve makes a single attempt
        BatchOut warm;
        solve_part(b, warm, nullptr, chunk);
    }
    Tape tape;
    tape.mode = Tape::RECORD;
    {
        BatchOut rec;
        solve_part(b, rec, nullptr, chunk, true, &tape);
    }
    CF_HIP(hipStreamSynchronize(ctx->stream));
    std::unique_ptr<BatchGraph> G(new BatchGraph());
    G->chunk = chunk;
    tape.mode = Tape::REPLAY;
    tape.pos = 0;
    ctx->arena = &G->arena;
    CF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    try {
        solve_part(b, G->out, nullptr, chunk, true, &tape);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        ctx->arena = nullptr;
        graph_destroy(ctx, G.release());
        throw;
    }
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &G->graph);
    ctx->arena = nullptr;
    if (ec != hipSuccess || tape.pos != tape.v.size()) {
        graph_destroy(ctx, G.release());
        CF_HIP(ec);
        throw CfError{CF_EINTERNAL, "the captured solve read back fewer values than the recorded one"};
    }
    G->npairs = G->out.npairs;
    const hipError_t ei = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        graph_destroy(ctx, G.release());
        CF_HIP(ei);
    }
    hipError_t eh = hipHostMalloc((void **)&G->hdr, (2 + 2 * (size_t)b->n) * sizeof(u64), hipHostMallocDefault);
    if (eh != hipSuccess) G->hdr = nullptr;
    if (eh == hipSuccess) eh = hipEventCreate(&G->t0);
    if (eh == hipSuccess) eh = hipEventCreate(&G->t1);
    if (eh != hipSuccess) {
        graph_destroy(ctx, G.release());
        CF_HIP(eh);
    }
    b->graph = G.release();
    API_END
}

int cf_batch_create(cf_ctx *ctx, const cf_instance *insts, int n, cf_batch **out) {
    API_BEGIN
    CF_REQUIRE(ctx && insts && n >= 1 && out, "NULL argument");
    std::unique_ptr<cf_batch> b(new cf_batch());
    b->ctx = ctx;
    batch_prepare(b.get(), insts, n);
    *out = b.release();
    API_END
}

int cf_batch_solve(cf_batch *b, uint64_t max_chunk, uint64_t **out_d, uint64_t **out_i, uint64_t **out_inst,
                   uint64_t *npairs, uint64_t *inst_counts, cf_stats *stats) {
    API_BEGIN
    CF_REQUIRE(b && out_d && out_i && out_inst && npairs, "NULL argument");
    *out_d = *out_i = *out_inst = nullptr;
    *npairs = 0;
    b->ctx->peak_live = b->ctx->live_bytes;
    BatchGraph *G = b->graph;
    if (G && (!max_chunk || max_chunk == G->chunk)) {
        // replay the captured solve: one graph launch, then the few results are read back. The pair
        // count of every replay equals the captured one (checked below), so the pair copies are
        // queued with the counts and one synchronization serves all of them.
        cf_ctx *ctx = b->ctx;
        const u64 npc = G->npairs;
        DBuf<u64> pd, pi, pj;
        if (npc) {
            pd = DBuf<u64>(ctx, npc);
            pi = DBuf<u64>(ctx, npc);
            pj = DBuf<u64>(ctx, npc);
        }
        CF_HIP(hipEventRecord(G->t0, ctx->stream));
        CF_HIP(hipGraphLaunch(G->exec, ctx->stream));
        CF_HIP(hipEventRecord(G->t1, ctx->stream));
        u64 *h = G->hdr;
        h[1] = 0;
        CF_HIP(hipMemcpyAsync(h, G->out.cnt_pairs.p, sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipMemcpyAsync(h + 1, G->out.kerr.p, sizeof(u32), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipMemcpyAsync(h + 2, G->out.dcount.p, b->n * sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipMemcpyAsync(h + 2 + b->n, G->out.icount.p, b->n * sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
        if (npc) {
            CF_HIP(hipMemcpyAsync(pd.p, G->out.pd.p, npc * 8, hipMemcpyDeviceToDevice, ctx->stream));
            CF_HIP(hipMemcpyAsync(pi.p, G->out.pi.p, npc * 8, hipMemcpyDeviceToDevice, ctx->stream));
            CF_HIP(hipMemcpyAsync(pj.p, G->out.pinst.p, npc * 8, hipMemcpyDeviceToDevice, ctx->stream));
        }
        CF_HIP(hipStreamSynchronize(ctx->stream));
        const u64 np = h[0];
        const u32 ke = (u32)h[1];
        float ms = 0;
        CF_HIP(hipEventElapsedTime(&ms, G->t0, G->t1));
        if (ke & KERR_DWINDOW) throw CfError{CF_EINTERNAL, "a D record fell outside its hash range's buckets"};
        if (ke & KERR_CAND) throw CfError{CF_EINTERNAL, "the probe's candidate list was truncated"};
        if (np != npc) throw CfError{CF_EINTERNAL, "a replay of the captured solve found a different pair count"};
        const u64 *nD = h + 2, *nI = h + 2 + b->n;
        if (np) {
            *out_d = pd.release();
            *out_i = pi.release();
            *out_inst = pj.release();
        }
        *npairs = np;
        if (inst_counts)
            for (int j = 0; j < b->n; ++j) {
                inst_counts[2 * j] = nD[j];
                inst_counts[2 * j + 1] = nI[j];
            }
        if (stats) {
            *stats = cf_stats{};
            for (int j = 0; j < b->n; ++j) {
                stats->nD += nD[j];
                stats->nI += nI[j];
            }
            stats->npairs = np;
            stats->ms_total = ms;
        }
        return CF_OK;
    }
    BatchOut bo;
    solve_part(b, bo, stats, max_chunk);
    if (bo.npairs) {
        *out_d = bo.pd.release();
        *out_i = bo.pi.release();
        *out_inst = bo.pinst.release();
    }
    *npairs = bo.npairs;
    if (inst_counts)
        for (int j = 0; j < b->n; ++j) {
            inst_counts[2 * j] = bo.nD[j];
            inst_counts[2 * j + 1] = bo.nI[j];
        }
    API_END
}

int cf_batch_destroy(cf_batch *b) {
    API_BEGIN
    delete b;
    API_END
}

int cf_solve_batch(cf_ctx *ctx, const cf_instance *insts, int n, uint64_t **out_d, uint64_t **out_i,
                   uint64_t **out_inst, uint64_t *npairs, uint64_t *inst_counts, cf_stats *stats) {
    cf_batch *b = nullptr;
    int rc = cf_batch_create(ctx, insts, n, &b);
    if (rc) return rc;
    rc = cf_batch_solve(b, 0, out_d, out_i, out_inst, npairs, inst_counts, stats);
    cf_batch_destroy(b);
    return rc;
}

int cf_solve(cf_ctx *ctx, const cf_instance *inst, int mode, uint64_t max_chunk, uint64_t **out_d, uint64_t **out_i,
             uint64_t *npairs, cf_stats *stats) {
    API_BEGIN
    CF_REQUIRE(ctx && inst && out_d && out_i && npairs, "NULL argument");
    CF_REQUIRE(mode == CF_SOLVE_PARTITION || mode == CF_SOLVE_SORTMERGE || mode == CF_SOLVE_PARTITION_DGLOBAL,
               "unknown solve mode");
    *out_d = nullptr;
    *out_i = nullptr;
    *npairs = 0;
    ctx->peak_live = ctx->live_bytes;
    CF_REQUIRE(mode != CF_SOLVE_SORTMERGE || inst->M_hi == 0, "the sort-merge mode needs M < 2^64");
    if (mode != CF_SOLVE_SORTMERGE) {
        cf_batch b;
        b.ctx = ctx;
        batch_prepare(&b, inst, 1);
        BatchOut bo;
        solve_part(&b, bo, stats, max_chunk, mode == CF_SOLVE_PARTITION);
        if (bo.npairs) {
            *out_d = bo.pd.release();
            *out_i = bo.pi.release();
        }
        *npairs = bo.npairs;
        return CF_OK;
    }
    // ---- sort-merge: D generated and sorted; I streamed in chunks through sort + sorted join
    hipEvent_t ev[6];
    for (auto &e : ev) CF_HIP(hipEventCreate(&e));
    CF_HIP(hipEventRecord(ev[0], ctx->stream));
    cf_plan *pd = nullptr;
    int rc = cf_plan_create_d(ctx, inst->B, inst->nB, inst->r, inst->M, inst->dmin_lo, inst->dmin_hi, inst->d_id_mode, &pd);
    if (rc) throw CfError{rc, g_err};
    std::unique_ptr<cf_plan> hold_d(pd);
    const u64 nD = pd->empty ? 0 : pd->total;
    DBuf<u64> Dk(ctx, nD ? nD : 1), Did(ctx, nD ? nD : 1);
    plan_write_range(pd, 0, nD, Dk.p, Did.p);
    hold_d.reset();
    CF_HIP(hipEventRecord(ev[1], ctx->stream));
    const int kbits = bitlen(inst->M - 1 ? inst->M - 1 : 1);
    radix_sort_pairs(ctx, Dk.p, Did.p, nD, kbits);
    CF_HIP(hipEventRecord(ev[2], ctx->stream));
    cf_plan *pi = nullptr;
    rc = cf_plan_create_i(ctx, inst->INS, inst->nI, inst->r, inst->M, inst->P0modM, inst->icap_lo, inst->icap_hi,
                          inst->split_idx, inst->hmax, inst->i_id_mode, &pi);
    if (rc) throw CfError{rc, g_err};
    std::unique_ptr<cf_plan> hold_i(pi);
    const u64 nI = pi->empty ? 0 : pi->total;
    float ms_gen_i = 0, ms_sort_i = 0, ms_join = 0;
    CF_HIP(hipEventRecord(ev[3], ctx->stream));
    std::vector<std::pair<DBuf<u64>, DBuf<u64>>> parts;
    u64 total_pairs = 0;
    if (nI && nD) {
        const u64 chunk = max_chunk ? max_chunk : (u64)1 << 28;
        const u64 cs = std::min<u64>(chunk, nI);
        DBuf<u64> Ik(ctx, cs), Iid(ctx, cs), scnt(ctx, join_spans(cs)), soff(ctx, join_spans(cs) + 1);
        hipEvent_t a, b, c, d;
        CF_HIP(hipEventCreate(&a));
        CF_HIP(hipEventCreate(&b));
        CF_HIP(hipEventCreate(&c));
        CF_HIP(hipEventCreate(&d));
        for (u64 L = 0; L < nI; L += chunk) {
            const u64 m = std::min<u64>(chunk, nI - L);
            CF_HIP(hipEventRecord(a, ctx->stream));
            plan_write_range(pi, L, L + m, Ik.p, Iid.p);
            CF_HIP(hipEventRecord(b, ctx->stream));
            radix_sort_pairs(ctx, Ik.p, Iid.p, m, kbits);
            CF_HIP(hipEventRecord(c, ctx->stream));
            const u64 tot = join_sorted_count(ctx, Dk.p, nD, Ik.p, m, scnt.p, soff.p);
            if (tot) {
                DBuf<u64> od(ctx, tot), oi(ctx, tot);
                join_sorted_write(ctx, Dk.p, nD, Ik.p, Did.p, Iid.p, m, soff.p, od.p, oi.p, nullptr);
                parts.emplace_back(std::move(od), std::move(oi));
                total_pairs += tot;
            }
            CF_HIP(hipEventRecord(d, ctx->stream));
            CF_HIP(hipEventSynchronize(d));
            float t1, t2, t3;
            CF_HIP(hipEventElapsedTime(&t1, a, b));
            CF_HIP(hipEventElapsedTime(&t2, b, c));
            CF_HIP(hipEventElapsedTime(&t3, c, d));
            ms_gen_i += t1;
            ms_sort_i += t2;
            ms_join += t3;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        (void)hipEventDestroy(c);
        (void)hipEventDestroy(d);
    }
    if (total_pairs) {
        DBuf<u64> od(ctx, total_pairs), oi(ctx, total_pairs);
        u64 at = 0;
        for (auto &pr : parts) {
            CF_HIP(hipMemcpyAsync(od.p + at, pr.first.p, pr.first.n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
            CF_HIP(hipMemcpyAsync(oi.p + at, pr.second.p, pr.second.n * sizeof(u64), hipMemcpyDeviceToDevice, ctx->stream));
            at += pr.first.n;
        }
        parts.clear();
        *out_d = od.release();
        *out_i = oi.release();
    }
    CF_HIP(hipEventRecord(ev[4], ctx->stream));
    CF_HIP(hipEventSynchronize(ev[4]));
    *npairs = total_pairs;
    if (stats) {
        float a, b, c, t;
        CF_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        CF_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        CF_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
        CF_HIP(hipEventElapsedTime(&t, ev[0], ev[4]));
        stats->nD = nD;
        stats->nI = nI;
        stats->npairs = total_pairs;
        stats->peak_bytes = ctx->peak_live;
        stats->ms_total = t;
        stats->ms_gen_d = a;
        stats->ms_sort_d = b;
        stats->ms_gen_i = c + ms_gen_i;
        stats->ms_sort_i = ms_sort_i;
        stats->ms_join = ms_join;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    API_END
}

}  // extern "C"
