// capi.hip -- the C ABI of libcfjoin (declared in include/cfjoin.h): context, device pool,
// enumeration plans and the one-call instance solve.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>

#include "../../include/cfjoin.h"
#include "cf_internal.h"

thread_local std::string g_err;
void cf_set_error(const std::string &msg) { g_err = msg; }



typedef unsigned __int128 u128;

// ---- device pool ------------------------------------------------------------------------------
static size_t pool_round(size_t bytes) {
    if (bytes < 512) return 512;
    if (bytes <= (1u << 20)) {
        size_t s = 512;
        while (s < bytes) s <<= 1;
        return s;
    }
    const size_t g = 2u << 20;
    return (bytes + g - 1) / g * g;
}


// the live footprint (bytes handed out and not yet freed, cached blocks included when reused) and its peak
static void note_live(cf_ctx *ctx, long long delta) {
    ctx->live_bytes = (size_t)((long long)ctx->live_bytes + delta);
    ctx->peak_live = std::max(ctx->peak_live, ctx->live_bytes);
}

void *ctx_alloc(cf_ctx *ctx, size_t bytes) {
    const size_t sz = pool_round(bytes);
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->arena) {
        auto &fl = ctx->arena->free[ctx->arena_branch];
        auto a = fl.lower_bound(sz);
        if (a != fl.end() && a->first <= sz + sz / 2) {
            void *p = a->second;
            ctx->live[p] = a->first;
            note_live(ctx, (long long)a->first);
            fl.erase(a);
            return p;
        }
    }
    auto it = ctx->free_blocks.lower_bound(sz);
    if (it != ctx->free_blocks.end() && it->first <= sz + sz / 2) {
        void *p = it->second;
        const size_t bs = it->first;
        ctx->free_blocks.erase(it);
        ctx->cached_bytes -= bs;
        ctx->live[p] = bs;
        note_live(ctx, (long long)bs);
        if (ctx->arena) ctx->arena->owned[p] = bs;
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess && ctx->arena) {
        (void)hipGetLastError();
        throw CfError{CF_ENOMEM, "device out of memory allocating " + std::to_string(sz) + " bytes during capture"};
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // give the cache back to the device and retry once
        for (auto &kv : ctx->free_blocks) (void)hipFree(kv.second);
        ctx->free_blocks.clear();
        ctx->cached_bytes = 0;
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw CfError{CF_ENOMEM, "device out of memory allocating " + std::to_string(sz) + " bytes"};
        }
    }
    ctx->live[p] = sz;
    note_live(ctx, (long long)sz);
    if (ctx->arena) ctx->arena->owned[p] = sz;
    return p;
}

void ctx_free(cf_ctx *ctx, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->live.find(p);
    if (it == ctx->live.end()) return;
    note_live(ctx, -(long long)it->second);
    if (ctx->arena) {
        ctx->arena->free[ctx->arena_branch].emplace(it->second, p);
        ctx->arena->owned[p] = it->second;
        ctx->live.erase(it);
        return;
    }
    ctx->free_blocks.emplace(it->second, p);
    ctx->cached_bytes += it->second;
    ctx->live.erase(it);
}

void arena_release(cf_ctx *ctx, CaptureArena *a) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto &kv : a->owned) {
        auto it = ctx->live.find(kv.first);
        if (it != ctx->live.end()) {
            note_live(ctx, -(long long)it->second);
            ctx->live.erase(it);
        }
        ctx->free_blocks.emplace(kv.second, kv.first);
        ctx->cached_bytes += kv.second;
    }
    a->owned.clear();
    a->free.clear();
}

void ctx_release(cf_ctx *ctx) {
    CF_HIP(hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto &kv : ctx->free_blocks) CF_HIP(hipFree(kv.second));
    ctx->free_blocks.clear();
    ctx->cached_bytes = 0;
}

u64 d2h_u64(cf_ctx *ctx, const u64 *src) {
    CF_HIP(hipMemcpyAsync(ctx->pinned, src, sizeof(u64), hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    u64 v;
    std::memcpy(&v, ctx->pinned, sizeof(u64));
    return v;
}

int bitlen(u64 v) { return v ? 64 - __builtin_clzll(v) : 0; }

// ---- kernel timing ---------------------------------------------------------------------------
static bool prof_selected(cf_ctx *ctx, const char *name) {
    // no events inside a captured graph: replays do not time them reliably
    return ctx->prof_on && !ctx->arena && (ctx->prof_filter.empty() || ctx->prof_filter == name);
}
static hipEvent_t prof_event(cf_ctx *ctx) {
    if (!ctx->prof_free.empty()) {
        hipEvent_t e = ctx->prof_free.back();
        ctx->prof_free.pop_back();
        return e;
    }
    hipEvent_t e;
    CF_HIP(hipEventCreate(&e));
    return e;
}
void cf_prof_begin(cf_ctx *ctx, const char *name, hipEvent_t *ev) {
    if (!prof_selected(ctx, name)) return;
    *ev = prof_event(ctx);
    CF_HIP(hipEventRecord(*ev, ctx->stream));
}
void cf_prof_end(cf_ctx *ctx, const char *name, hipEvent_t ev) {
    if (!prof_selected(ctx, name) || !ev) return;
    hipEvent_t b = prof_event(ctx);
    CF_HIP(hipEventRecord(b, ctx->stream));
    ctx->prof_pending.push_back({name, ev, b});
}
static void prof_drain(cf_ctx *ctx) {
    if (ctx->prof_pending.empty()) return;
    CF_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &r : ctx->prof_pending) {
        float ms = 0;
        CF_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto &t = ctx->prof_totals[r.name];
        t.first += ms;
        t.second += 1;
        ctx->prof_free.push_back(r.a);
        ctx->prof_free.push_back(r.b);
    }
    ctx->prof_pending.clear();
}

// ---- plans ----------------------------------------------------------------------------------


void plan_expand(cf_plan *pl) {
    cf_ctx *ctx = pl->ctx;
    build_side_tables(ctx, pl->spec, pl->tab);
    pl->d_desc = DBuf<SideDesc>(ctx, 1);
    h2d(ctx, pl->d_desc.p, &pl->tab.desc, 1);
    const SideDesc &sd = pl->tab.desc;
    Node root{};
    root.A_lo = 1;
    root.A_hi = 0;
    root.xm = sd.xm0;
    root.x2 = sd.x20;
    root.id = 0;
    const int nxt = pl->spec.side == SIDE_I ? 0 : sd.n - 1;
    root.pos = (u32)(unsigned short)(short)nxt;
    root.meta = 0;
    DBuf<Node> rootd(ctx, 1);
    h2d(ctx, rootd.p, &root, 1);
    if (sd.n == 0 || sd.n < sd.r) { pl->empty = true; return; }
    // level by level, each pass writing a level and counting its nodes' children; the last interior
    // level as LeafNodes
    DBuf<Node> cur;
    const Node *curp = rootd.p;
    size_t ncur = 1;
    DBuf<u32> c(ctx, 1);
    launch_expand_count(ctx, pl->d_desc.p, curp, 1, 0, c.p);
    DBuf<LeafNode> last;
    for (int depth = 0; depth + 1 < sd.r; ++depth) {
        DBuf<u64> o(ctx, ncur + 1);
        const u64 tot = scan_u32_to_u64(ctx, c.p, o.p, ncur);
        if (tot == 0) { pl->empty = true; return; }
        DBuf<u32> cn(ctx, tot);
        if (depth + 2 < sd.r) {
            DBuf<Node> nx(ctx, tot);
            launch_expand_write_count(ctx, pl->d_desc.p, curp, ncur, o.p, tot, depth, nx.p, cn.p);
            cur = std::move(nx);
            curp = cur.p;
        } else {
            last = DBuf<LeafNode>(ctx, tot);
            launch_expand_write_count(ctx, pl->d_desc.p, curp, ncur, o.p, tot, depth, last.p, cn.p);
        }
        c = std::move(cn);
        ncur = tot;
    }
    if (!last.p) {   // r == 1: the root is the last level
        last = DBuf<LeafNode>(ctx, 1);
        launch_leaf_nodes(ctx, rootd.p, 1, last.p);
    }
    pl->off = DBuf<u64>(ctx, ncur + 1);
    pl->total = scan_u32_to_u64(ctx, c.p, pl->off.p, ncur);
    pl->cnt = std::move(c);
    pl->leafp = std::move(last);
    pl->nleafp = ncur;
    if (pl->total == 0) pl->empty = true;
}

static void plan_first_counts(cf_plan *pl) {
    if (pl->have_first) return;
    const int n = pl->tab.desc.n;
    pl->first_counts.assign(n, 0);
    if (!pl->empty) {
        if (pl->tab.desc.r == 1) {
            // the root is the only last-level node; its leaves are the first indices themselves
            u32 c0 = 0;
            d2h(pl->ctx, &c0, pl->cnt.p, 1);
            for (u32 k = 0; k < c0; ++k) {
                const int x = pl->spec.side == SIDE_I ? (int)k : n - 1 - (int)k;
                pl->first_counts[x] = 1;
            }
        } else {
            DBuf<u64> fb(pl->ctx, 2 * (size_t)n);
            CF_HIP(hipMemsetAsync(fb.p, 0, 2 * (size_t)n * sizeof(u64), pl->ctx->stream));
            launch_first_bounds(pl->ctx, pl->leafp.p, pl->off.p, pl->nleafp, fb.p, fb.p + n);
            std::vector<u64> h(2 * (size_t)n);
            d2h(pl->ctx, h.data(), fb.p, 2 * (size_t)n);
            for (int x = 0; x < n; ++x) pl->first_counts[x] = h[n + x] - h[x];
        }
    }
    pl->have_first = true;
}

void plan_write_range(cf_plan *pl, u64 L0, u64 L1, u64 *keys, u64 *ids) {
    if (L1 <= L0) return;
    leaf_records(pl->ctx, pl->d_desc.p, pl->leafp.p, pl->nleafp, pl->off.p, L0, L1 - L0, pl->tab.desc.r - 1, keys, ids);
}

extern "C" {

int cf_version(void) { return 1; }
const char *cf_last_error(void) { return g_err.c_str(); }

int cf_ctx_create(int device, cf_ctx **out) {
    API_BEGIN
    CF_REQUIRE(out != nullptr, "out is NULL");
    int ndev = 0;
    CF_HIP(hipGetDeviceCount(&ndev));
    CF_REQUIRE(device >= 0 && device < ndev, "no such device");
    CF_HIP(hipSetDevice(device));
    u64 *scratch = join_scratch_bind(device);
    cf_ctx *c = new cf_ctx();
    c->device = device;
    CF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CF_HIP(hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault));
    c->pinned_bytes = 4096;
    c->join_scratch = scratch;
    *out = c;
    API_END
}

int cf_ctx_destroy(cf_ctx *ctx) {
    API_BEGIN
    if (!ctx) return CF_OK;
    CF_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &kv : ctx->free_blocks) (void)hipFree(kv.second);
    for (auto &kv : ctx->live) (void)hipFree(kv.first);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto &r : ctx->prof_pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : ctx->prof_free) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    API_END
}

int cf_alloc(cf_ctx *ctx, uint64_t bytes, void **out_dev) {
    API_BEGIN
    CF_REQUIRE(ctx && out_dev, "NULL argument");
    *out_dev = ctx_alloc(ctx, bytes);
    API_END
}

int cf_free(cf_ctx *ctx, void *p_dev) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    ctx_free(ctx, p_dev);
    API_END
}

int cf_pool_release(cf_ctx *ctx) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    ctx_release(ctx);
    API_END
}

int cf_pool_stats(cf_ctx *ctx, uint64_t *live_bytes, uint64_t *cached_bytes) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    size_t live = 0;
    for (auto &kv : ctx->live) live += kv.second;
    if (live_bytes) *live_bytes = live;
    if (cached_bytes) *cached_bytes = ctx->cached_bytes;
    API_END
}

int cf_mem_info(cf_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    size_t f = 0, t = 0;
    CF_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    API_END
}

int cf_memcpy_h2d(cf_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    if (bytes) CF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_memcpy_d2h(cf_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    if (bytes) CF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_memcpy_d2d(cf_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    if (bytes) CF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    API_END
}

int cf_synchronize(cf_ctx *ctx) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_prof_enable(cf_ctx *ctx, const char *kernel) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    ctx->prof_on = true;
    ctx->prof_filter = kernel ? kernel : "";
    API_END
}

int cf_prof_disable(cf_ctx *ctx) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    prof_drain(ctx);
    ctx->prof_on = false;
    API_END
}

int cf_prof_reset(cf_ctx *ctx) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    prof_drain(ctx);
    ctx->prof_totals.clear();
    API_END
}

int cf_prof_read(cf_ctx *ctx, const char *kernel, double *ms, uint64_t *launches) {
    API_BEGIN
    CF_REQUIRE(ctx && kernel, "NULL argument");
    prof_drain(ctx);
    auto it = ctx->prof_totals.find(kernel);
    if (ms) *ms = it == ctx->prof_totals.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == ctx->prof_totals.end() ? 0 : it->second.second;
    API_END
}

int cf_prof_dump(cf_ctx *ctx, char *buf, uint64_t size) {
    API_BEGIN
    CF_REQUIRE(ctx && buf && size, "NULL argument");
    prof_drain(ctx);
    std::string text;
    for (auto &kv : ctx->prof_totals)
        text += kv.first + " " + std::to_string(kv.second.first) + " " + std::to_string(kv.second.second) + "\n";
    const size_t m = std::min<size_t>(text.size(), size - 1);
    std::memcpy(buf, text.data(), m);
    buf[m] = 0;
    API_END
}

int cf_gen_subsets(cf_ctx *ctx, const uint64_t *primes, int n, int r, int bits, uint64_t M, uint64_t mult,
                   uint64_t lo, uint64_t count, uint64_t *keys_dev, uint64_t *ids_dev) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    CF_REQUIRE(r >= 1 && r <= 10 && r * bits <= 64, "need 1 <= r <= 10 and r*bits <= 64");
    CF_REQUIRE(bits >= 1 && n >= 0 && n <= (1 << std::min(bits, 10)), "pool does not fit the index width");
    CF_REQUIRE(M >= 1, "modulus must be >= 1");
    if (count == 0) return CF_OK;
    CF_REQUIRE(n >= r, "fewer primes than r");
    SideSpec s;
    s.side = SIDE_I;
    s.primes.assign(primes, primes + n);
    s.r = r;
    s.M = M;
    s.mult = mult;
    s.cap = 0;
    s.id_mode = ID_PACK_ASC;
    s.id_bits = bits;
    SideTables t;
    build_side_tables(ctx, s, t);
    DBuf<SideDesc> dd(ctx, 1);
    h2d(ctx, dd.p, &t.desc, 1);
    launch_gen_prefix(ctx, dd.p, n, r, lo, count, keys_dev, ids_dev);
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_sort_pairs(cf_ctx *ctx, uint64_t *keys_dev, uint64_t *ids_dev, uint64_t n) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    if (n <= 1) return CF_OK;
    // only the significant bits of the largest key need a pass
    radix_sort_pairs(ctx, keys_dev, ids_dev, n, bitlen(max_u64(ctx, keys_dev, n)));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_sort_pairs_into(cf_ctx *ctx, const uint64_t *keys_dev, const uint64_t *ids_dev, uint64_t *keys_out_dev,
                       uint64_t *ids_out_dev, uint64_t n, uint64_t key_bound) {
    API_BEGIN
    CF_REQUIRE(ctx, "NULL context");
    if (n == 0) return CF_OK;
    CF_REQUIRE(keys_dev && ids_dev && keys_out_dev && ids_out_dev, "NULL array");
    CF_REQUIRE((keys_dev == keys_out_dev) == (ids_dev == ids_out_dev), "sort both arrays in place or neither");
    const int bits = key_bound ? bitlen(key_bound - 1) : bitlen(max_u64(ctx, keys_dev, n));
    sort_pairs_into(ctx, keys_dev, ids_dev, keys_out_dev, ids_out_dev, n, bits);
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_join_sorted(cf_ctx *ctx, const uint64_t *Dk, const uint64_t *Did, uint64_t nD, const uint64_t *Ik,
                   const uint64_t *Iid, uint64_t nI, uint64_t **out_d, uint64_t **out_i, uint64_t *npairs) {
    API_BEGIN
    CF_REQUIRE(ctx && out_d && out_i && npairs, "NULL argument");
    *out_d = nullptr;
    *out_i = nullptr;
    *npairs = 0;
    if (nD == 0 || nI == 0) return CF_OK;
    const size_t ns = join_spans(nI);
    DBuf<u64> scnt(ctx, ns), soff(ctx, ns + 1);
    const u64 tot = join_sorted_count(ctx, Dk, nD, Ik, nI, scnt.p, soff.p);
    if (tot == 0) return CF_OK;
    DBuf<u64> od(ctx, tot), oi(ctx, tot);
    join_sorted_write(ctx, Dk, nD, Ik, Did, Iid, nI, soff.p, od.p, oi.p, nullptr);
    CF_HIP(hipStreamSynchronize(ctx->stream));
    *out_d = od.release();
    *out_i = oi.release();
    *npairs = tot;
    API_END
}

int cf_plan_create_i(cf_ctx *ctx, const uint64_t *INS, int n, int r, uint64_t M, uint64_t P0modM, uint64_t icap_lo,
                     uint64_t icap_hi, int split_idx, int hmax, int id_mode, cf_plan **out) {
    API_BEGIN
    CF_REQUIRE(ctx && out, "NULL argument");
    CF_REQUIRE(r >= 1 && r <= 16 && n >= 0 && n <= 1024, "need 1 <= r <= 16 and n <= 1024");
    CF_REQUIRE(M >= 1, "need 1 <= M < 2^64");
    for (int i = 1; i < n; ++i) CF_REQUIRE(INS[i - 1] < INS[i], "INS must be strictly ascending");
    CF_REQUIRE(id_mode == ID_COLEX || id_mode == ID_PACK_ASC, "I ids are colex ranks or 8-bit packed");
    if (id_mode == ID_PACK_ASC) CF_REQUIRE(n <= 256 && r <= 8, "8-bit packed ids need n <= 256 and r <= 8");
    if (id_mode == ID_COLEX) {
        u128 c = 1;
        bool over = false;
        for (int j = 0; j < r; ++j) {
            // C(n, r) = prod (n - j) / (j + 1), exact at every step
            c = c * (u128)(n - j) / (u128)(j + 1);
            if (c > (u128)~0ULL) over = true;
        }
        CF_REQUIRE(!over, "rank ids would overflow u64 (C(n, r) >= 2^64)");
    }
    std::unique_ptr<cf_plan> pl(new cf_plan());
    pl->ctx = ctx;
    pl->spec.side = SIDE_I;
    pl->spec.primes.assign(INS, INS + n);
    pl->spec.r = r;
    pl->spec.M = M;
    pl->spec.mult = P0modM;
    pl->spec.cap = ((u128)icap_hi << 64) | icap_lo;
    pl->spec.split_idx = split_idx;
    pl->spec.hmax = hmax;
    pl->spec.id_mode = id_mode;
    pl->spec.id_bits = 8;
    if (pl->spec.cap <= 1) {
        build_side_tables(ctx, pl->spec, pl->tab);
        pl->empty = true;
    } else {
        plan_expand(pl.get());
    }
    *out = pl.release();
    API_END
}

int cf_plan_create_d(cf_ctx *ctx, const uint64_t *B, int nB, int r, uint64_t M, uint64_t dmin_lo, uint64_t dmin_hi,
                     int id_mode, cf_plan **out) {
    API_BEGIN
    CF_REQUIRE(ctx && out, "NULL argument");
    CF_REQUIRE(nB >= 1 && nB <= 64 && r >= 1 && r <= 16, "need nB <= 64 base primes and 1 <= r <= 16");
    CF_REQUIRE(M >= 1, "need 1 <= M < 2^64");
    for (int i = 1; i < nB; ++i) CF_REQUIRE(B[i - 1] < B[i], "base must be strictly ascending");
    CF_REQUIRE(id_mode == ID_PACK_DESC || id_mode == ID_MASK, "D ids are 6-bit packed or deletion masks");
    if (id_mode == ID_PACK_DESC) CF_REQUIRE(r <= 10, "6-bit packed ids need r <= 10");
    // exact products: the r largest base primes must multiply below 2^128
    {
        u128 w = 1;
        for (int j = 0; j < r && j < nB; ++j) {
            const u64 q = B[nB - 1 - j];
            CF_REQUIRE(w <= (~(u128)0) / q, "products of r base primes exceed 128 bits");
            w *= q;
        }
    }
    std::unique_ptr<cf_plan> pl(new cf_plan());
    pl->ctx = ctx;
    pl->spec.side = SIDE_D;
    pl->spec.primes.assign(B, B + nB);
    pl->spec.r = r;
    pl->spec.M = M;
    pl->spec.mult = 1;
    pl->spec.cap = ((u128)dmin_hi << 64) | dmin_lo;
    pl->spec.id_mode = id_mode;
    pl->spec.id_bits = 6;
    plan_expand(pl.get());
    *out = pl.release();
    API_END
}

int cf_plan_total(cf_plan *plan, uint64_t *total) {
    API_BEGIN
    CF_REQUIRE(plan && total, "NULL argument");
    *total = plan->empty ? 0 : plan->total;
    API_END
}

int cf_plan_first_counts(cf_plan *plan, uint64_t *counts) {
    API_BEGIN
    CF_REQUIRE(plan && counts, "NULL argument");
    plan_first_counts(plan);
    std::copy(plan->first_counts.begin(), plan->first_counts.end(), counts);
    API_END
}

int cf_plan_write(cf_plan *plan, int first_lo, int first_hi, uint64_t *keys, uint64_t *ids, uint64_t *written) {
    API_BEGIN
    CF_REQUIRE(plan, "NULL plan");
    const int n = plan->tab.desc.n;
    first_lo = std::max(first_lo, 0);
    first_hi = std::min(first_hi, n);
    u64 L0 = 0, L1 = 0;
    if (!plan->empty && first_lo < first_hi) {
        if (first_lo == 0 && first_hi == n) {
            L0 = 0;
            L1 = plan->total;
        } else {
            plan_first_counts(plan);
            const auto &fc = plan->first_counts;
            if (plan->spec.side == SIDE_I) {
                for (int x = 0; x < first_lo; ++x) L0 += fc[x];
                L1 = L0;
                for (int x = first_lo; x < first_hi; ++x) L1 += fc[x];
            } else {  // descending first indices
                for (int x = n - 1; x >= first_hi; --x) L0 += fc[x];
                L1 = L0;
                for (int x = first_hi - 1; x >= first_lo; --x) L1 += fc[x];
            }
        }
        plan_write_range(plan, L0, L1, keys, ids);
        CF_HIP(hipStreamSynchronize(plan->ctx->stream));
    }
    if (written) *written = L1 - L0;
    API_END
}

int cf_plan_destroy(cf_plan *plan) {
    API_BEGIN
    delete plan;
    API_END
}

}  // extern "C"
