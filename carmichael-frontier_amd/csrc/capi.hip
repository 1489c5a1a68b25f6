// capi.hip -- the C ABI of libcfjoin (declared in include/cfjoin.h): context, device pool,
// enumeration plans and the one-call instance solve.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>

#include "../../include/cfjoin.h"
#include "cf_internal.h"

static thread_local std::string g_err;
void cf_set_error(const std::string &msg) { g_err = msg; }

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

static size_t g_peak_live = 0;

void *ctx_alloc(cf_ctx *ctx, size_t bytes) {
    const size_t sz = pool_round(bytes);
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->free_blocks.lower_bound(sz);
    if (it != ctx->free_blocks.end() && it->first <= sz + sz / 2) {
        void *p = it->second;
        const size_t bs = it->first;
        ctx->free_blocks.erase(it);
        ctx->cached_bytes -= bs;
        ctx->live[p] = bs;
        return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
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
    size_t live = 0;
    for (auto &kv : ctx->live) live += kv.second;
    g_peak_live = std::max(g_peak_live, live);
    return p;
}

void ctx_free(cf_ctx *ctx, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->live.find(p);
    if (it == ctx->live.end()) return;
    ctx->free_blocks.emplace(it->second, p);
    ctx->cached_bytes += it->second;
    ctx->live.erase(it);
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

static int bitlen(u64 v) { return v ? 64 - __builtin_clzll(v) : 0; }

// ---- plans ----------------------------------------------------------------------------------
struct cf_plan {
    cf_ctx *ctx = nullptr;
    SideSpec spec;
    SideTables tab;
    DBuf<SideDesc> d_desc;
    DBuf<Node> leafp;      // depth r-1 nodes
    size_t nleafp = 0;
    DBuf<u32> cnt;         // leaves per depth r-1 node
    DBuf<u64> off;         // exclusive scan of cnt, nleafp + 1 entries
    u64 total = 0;
    bool empty = false;
    std::vector<u64> first_counts;
    bool have_first = false;
};

static void plan_expand(cf_plan *pl) {
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
    DBuf<Node> cur(ctx, 1);
    h2d(ctx, cur.p, &root, 1);
    size_t ncur = 1;
    if (sd.n == 0 || sd.n < sd.r) { pl->empty = true; return; }
    for (int depth = 0; depth + 1 < sd.r; ++depth) {
        DBuf<u32> c(ctx, ncur);
        launch_expand_count(ctx, pl->d_desc.p, cur.p, ncur, depth, c.p);
        DBuf<u64> o(ctx, ncur + 1);
        const u64 tot = scan_u32_to_u64(ctx, c.p, o.p, ncur);
        if (tot == 0) { pl->empty = true; return; }
        DBuf<Node> nx(ctx, tot);
        launch_expand_write(ctx, pl->d_desc.p, cur.p, ncur, o.p, tot, depth, nx.p);
        cur = std::move(nx);
        ncur = tot;
    }
    pl->cnt = DBuf<u32>(ctx, ncur);
    launch_expand_count(ctx, pl->d_desc.p, cur.p, ncur, sd.r - 1, pl->cnt.p);
    pl->off = DBuf<u64>(ctx, ncur + 1);
    pl->total = scan_u32_to_u64(ctx, pl->cnt.p, pl->off.p, ncur);
    pl->leafp = std::move(cur);
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
            DBuf<u64> fc(pl->ctx, n);
            CF_HIP(hipMemsetAsync(fc.p, 0, n * sizeof(u64), pl->ctx->stream));
            launch_first_counts(pl->ctx, pl->leafp.p, pl->cnt.p, pl->nleafp, fc.p);
            d2h(pl->ctx, pl->first_counts.data(), fc.p, n);
        }
    }
    pl->have_first = true;
}

static void plan_write_range(cf_plan *pl, u64 L0, u64 L1, u64 *keys, u64 *ids) {
    if (L1 <= L0) return;
    launch_leaf_write(pl->ctx, pl->d_desc.p, pl->leafp.p, pl->nleafp, pl->off.p, L0, L1 - L0, pl->tab.desc.r - 1, keys,
                      ids);
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
    cf_ctx *c = new cf_ctx();
    c->device = device;
    CF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CF_HIP(hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault));
    c->pinned_bytes = 4096;
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
    launch_gen_rank(ctx, dd.p, lo, count, keys_dev, ids_dev);
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

int cf_join_sorted(cf_ctx *ctx, const uint64_t *Dk, const uint64_t *Did, uint64_t nD, const uint64_t *Ik,
                   const uint64_t *Iid, uint64_t nI, uint64_t **out_d, uint64_t **out_i, uint64_t *npairs) {
    API_BEGIN
    CF_REQUIRE(ctx && out_d && out_i && npairs, "NULL argument");
    *out_d = nullptr;
    *out_i = nullptr;
    *npairs = 0;
    if (nD == 0 || nI == 0) return CF_OK;
    DBuf<u64> lo(ctx, nI), cnt(ctx, nI), off(ctx, nI + 1);
    const u64 tot = join_sorted_count(ctx, Dk, nD, Ik, nI, lo.p, cnt.p, off.p);
    if (tot == 0) return CF_OK;
    DBuf<u64> od(ctx, tot), oi(ctx, tot);
    join_sorted_write(ctx, Did, Iid, nI, lo.p, cnt.p, off.p, od.p, oi.p, nullptr, Ik);
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

int cf_solve(cf_ctx *ctx, const cf_instance *inst, uint64_t max_chunk, uint64_t **out_d, uint64_t **out_i,
             uint64_t *npairs, cf_stats *stats) {
    API_BEGIN
    CF_REQUIRE(ctx && inst && out_d && out_i && npairs, "NULL argument");
    *out_d = nullptr;
    *out_i = nullptr;
    *npairs = 0;
    g_peak_live = 0;
    hipEvent_t ev[6];
    for (auto &e : ev) CF_HIP(hipEventCreate(&e));
    CF_HIP(hipEventRecord(ev[0], ctx->stream));
    // D side: generate, sort
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
    // I side: plan, then stream leaf ranges of at most max_chunk records through sort + join
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
        DBuf<u64> Ik(ctx, cs), Iid(ctx, cs), lo(ctx, cs), cnt(ctx, cs), off(ctx, cs + 1);
        for (u64 L = 0; L < nI; L += chunk) {
            const u64 m = std::min<u64>(chunk, nI - L);
            hipEvent_t a, b, c, d;
            CF_HIP(hipEventCreate(&a)); CF_HIP(hipEventCreate(&b)); CF_HIP(hipEventCreate(&c)); CF_HIP(hipEventCreate(&d));
            CF_HIP(hipEventRecord(a, ctx->stream));
            plan_write_range(pi, L, L + m, Ik.p, Iid.p);
            CF_HIP(hipEventRecord(b, ctx->stream));
            radix_sort_pairs(ctx, Ik.p, Iid.p, m, kbits);
            CF_HIP(hipEventRecord(c, ctx->stream));
            const u64 tot = join_sorted_count(ctx, Dk.p, nD, Ik.p, m, lo.p, cnt.p, off.p);
            if (tot) {
                DBuf<u64> od(ctx, tot), oi(ctx, tot);
                join_sorted_write(ctx, Did.p, Iid.p, m, lo.p, cnt.p, off.p, od.p, oi.p, nullptr, Ik.p);
                parts.emplace_back(std::move(od), std::move(oi));
                total_pairs += tot;
            }
            CF_HIP(hipEventRecord(d, ctx->stream));
            CF_HIP(hipEventSynchronize(d));
            float t1, t2, t3;
            CF_HIP(hipEventElapsedTime(&t1, a, b)); CF_HIP(hipEventElapsedTime(&t2, b, c)); CF_HIP(hipEventElapsedTime(&t3, c, d));
            ms_gen_i += t1; ms_sort_i += t2; ms_join += t3;
            (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); (void)hipEventDestroy(d);
        }
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
        stats->peak_bytes = g_peak_live;
        stats->ms_total = t;
        stats->ms_gen_d = a;
        stats->ms_sort_d = b;
        stats->ms_gen_i = c + ms_gen_i;   // I plan expansion + leaf writes
        stats->ms_sort_i = ms_sort_i;
        stats->ms_join = ms_join;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    API_END
}

}  // extern "C"
