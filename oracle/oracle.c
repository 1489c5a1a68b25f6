/* oracle.c -- CPU restatement of the exchange-MITM residue join.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it (through ctypes, as liboracle.so). The product library
 * (libcfjoin.so) never links, calls or falls back to it.
 *
 * Each function restates one reference routine on the host with exact u128 arithmetic:
 *   or_gen_rank     gpu/mitm_gpu2.py:25-65 (gen_records) + :79-98 (generate)
 *   or_ddump        src/mitm64_ddump.cpp:29-44 (rec), :46-55 (main) -- product-filtered D side
 *   or_idump        src/mitm64_idump_stream2.cpp:31-57 (rec), :75-84 (workers) -- filtered I side
 *   or_join_sorted  gpu/mitm_gpu2.py:107-119 (join_sorted) on key-sorted sides
 *   or_cpu_join     src/mitm64.cpp:51-121 (hash the D residues, T threads striped over the first
 *                   insertion index run the I DFS and probe) over the same record sets as the
 *                   two dumpers, emitting every equal-residue pair
 * How each is pinned (golden streams of the compiled reference tools, logged counts, brute force)
 * is stated in tests/test_oracle.py.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned long long u64;
typedef unsigned __int128 u128;

static inline u64 mulmod(u64 a, u64 b, u64 M) { return (u64)(((u128)a * b) % M); }
static inline u128 mk128(u64 lo, u64 hi) { return ((u128)hi << 64) | lo; }

/* ------------------------------------------------------------------------------------------ */
/* gen_records: lexicographic combinatorial-number-system unranking, key = mult * prod mod M,  */
/* id = indices packed `bits` wide, lowest position first (gpu/mitm_gpu2.py:36-64).           */
long long or_gen_rank(const u64 *primes, int n, int r, int bits, u64 M, u64 mult, u64 lo,
                      u64 count, u64 *keys, u64 *ids) {
    if (r < 1 || r > 10 || r * bits > 64 || M == 0) return -1;
    const int stride = r + 1;
    u64 *binom = (u64 *)calloc((size_t)(n + 1) * stride, sizeof(u64));
    if (!binom) return -2;
    for (int a = 0; a <= n; ++a) binom[(size_t)a * stride] = 1;            /* mitm_gpu2.py:70-76 */
    for (int a = 1; a <= n; ++a)
        for (int b = 1; b <= (a < r ? a : r); ++b)
            binom[(size_t)a * stride + b] = binom[(size_t)(a - 1) * stride + b - 1] + binom[(size_t)(a - 1) * stride + b];
    for (u64 t = 0; t < count; ++t) {
        u64 rank = lo + t;
        int idx[10], start = 0;
        for (int pos = 0; pos < r; ++pos) {
            int remaining = r - pos - 1;
            for (int c = start; c <= n - 1; ++c) {
                u64 cnt = binom[(size_t)(n - 1 - c) * stride + remaining];
                if (rank < cnt) { idx[pos] = c; start = c + 1; break; }
                rank -= cnt;
            }
        }
        u64 acc = mult % M, id = 0;
        for (int pos = 0; pos < r; ++pos) {
            acc = mulmod(acc, primes[idx[pos]] % M, M);
            id |= ((u64)idx[pos]) << (pos * bits);
        }
        keys[t] = acc;
        ids[t] = id;
    }
    free(binom);
    return (long long)count;
}

/* ------------------------------------------------------------------------------------------ */
/* D side: subsets of the ascending base B with prod >= Dmin, chosen in DESCENDING index order */
/* with the reach prune "partial * (t largest still available) < Dmin -> break"; key =        */
/* prod mod M, id = 6-bit packed ascending indices or a 64-bit deletion mask                  */
/* (src/mitm64_ddump.cpp:29-44). Products wrap in u128 exactly as the reference's do.         */
typedef struct {
    u64 M; int R; u128 Dmin; const u64 *B; int bitmask;
    u64 *keys, *ids; u64 cap, total;
} dctx_t;

static void drec(dctx_t *c, int hi, int depth, u128 prod, u64 res, int *idx) {
    if (depth == c->R) {
        u64 id = 0;
        if (c->bitmask) { for (int j = 0; j < c->R; ++j) id |= 1ULL << idx[j]; }
        else { for (int j = c->R - 1, pos = 0; j >= 0; --j, ++pos) id |= ((u64)idx[j]) << (6 * pos); }
        if (c->total < c->cap) { c->keys[c->total] = res % c->M; c->ids[c->total] = id; }
        c->total++;
        return;
    }
    const int t = c->R - depth - 1;
    for (int i = hi; i >= t; --i) {
        u128 np = prod * c->B[i];
        u128 best = np;
        for (int j = 0; j < t; ++j) best *= c->B[i - 1 - j];
        if (best < c->Dmin) break;
        idx[depth] = i;
        drec(c, i - 1, depth + 1, np, mulmod(res, c->B[i] % c->M, c->M), idx);
    }
}

/* Returns the record count; writes the first min(count, cap) records. */
long long or_ddump(u64 M, int r, u64 dmin_lo, u64 dmin_hi, const u64 *B, int nB, int bitmask,
                   u64 *keys, u64 *ids, u64 cap) {
    if (nB < 1 || nB > 64 || r < 1 || r > 16 || M == 0) return -1;
    for (int i = 1; i < nB; ++i) if (B[i] <= B[i - 1]) return -4;
    dctx_t c = {M, r, mk128(dmin_lo, dmin_hi), B, bitmask || r > 10, keys, ids, cap, 0};
    int idx[16];
    drec(&c, nB - 1, 0, (u128)1, 1, idx);
    return (long long)c.total;
}

/* ------------------------------------------------------------------------------------------ */
/* I side: ascending subsets of INS with prod < Icap (exclusive), the cheapest-completion     */
/* break, loop bound i <= n-1-t and at most hmax indices >= split_idx; key = P0modM * prod    */
/* mod M; id = colex rank sum C(i_j, j+1) or 8-bit packed (src/mitm64_idump_stream2.cpp).     */
typedef struct {
    u64 M, P0modM; int R, n, split_idx, hmax, rank; u128 Icap; const u64 *INS;
    u128 (*CB)[17];
} ictx_t;

static inline int mul_ge(u128 a, u128 b, u128 cap) {
    u128 r;
    if (__builtin_mul_overflow(a, b, &r)) return 1;
    return r >= cap;
}
static inline int np_window_ge(const ictx_t *c, u128 np, int a, int t) {
    u128 v = np;
    for (int j = 0; j < t; ++j) if (__builtin_mul_overflow(v, (u128)c->INS[a + j], &v)) return 1;
    return v >= c->Icap;
}

typedef void (*leaf_fn)(void *user, u64 key, u64 id);

static void irec(const ictx_t *c, int start, int depth, u128 prod, u64 resI, u64 id, int nlarge,
                 leaf_fn emit, void *user) {
    if (depth == c->R) { emit(user, mulmod(c->P0modM, resI, c->M), id); return; }
    const int t = c->R - depth - 1;
    for (int i = start; i <= c->n - 1 - t; ++i) {
        if (i >= c->split_idx && nlarge + 1 > c->hmax) break;
        if (mul_ge(prod, c->INS[i], c->Icap)) break;
        u128 np = prod * c->INS[i];
        if (t > 0 && np_window_ge(c, np, i + 1, t)) break;
        u64 nid = c->rank ? id + (u64)c->CB[i][depth + 1] : (id | ((u64)i << (8 * depth)));
        irec(c, i + 1, depth + 1, np, mulmod(resI, c->INS[i] % c->M, c->M), nid,
             nlarge + (i >= c->split_idx), emit, user);
    }
}

/* one first index i0, as the stream2 workers do (src/mitm64_idump_stream2.cpp:76-82) */
static void itop(const ictx_t *c, int i0, leaf_fn emit, void *user) {
    const int tt = c->R - 1;
    if (i0 > c->n - 1 - tt) return;
    if (i0 >= c->split_idx && 1 > c->hmax) return;
    u128 np = (u128)c->INS[i0];
    if (np >= c->Icap) return;
    if (tt > 0 && np_window_ge(c, np, i0 + 1, tt)) return;
    irec(c, i0 + 1, 1, np, c->INS[i0] % c->M, (u64)i0, (i0 >= c->split_idx), emit, user);
}

static u128 (*make_cb(void))[17] {
    u128 (*CB)[17] = (u128 (*)[17])calloc(1025, sizeof(*CB));
    if (!CB) return NULL;
    for (int n = 0; n <= 1024; ++n) {
        CB[n][0] = 1;
        for (int k = 1; k < 17; ++k) CB[n][k] = n ? CB[n - 1][k - 1] + CB[n - 1][k] : 0;
    }
    return CB;
}

typedef struct { u64 *keys, *ids; u64 cap, total; } sink_t;
static void sink_put(void *user, u64 key, u64 id) {
    sink_t *s = (sink_t *)user;
    if (s->total < s->cap) { s->keys[s->total] = key; s->ids[s->total] = id; }
    s->total++;
}

static int ictx_init(ictx_t *c, u64 M, int r, u64 P0modM, u64 icap_lo, u64 icap_hi, const u64 *INS,
                     int n, int split_idx, int hmax, int rank) {
    if (M == 0 || r < 1 || r > 16 || n > 1024 || n < 0) return -1;
    if (!rank && (n > 256 || r > 8)) return -3;
    for (int i = 1; i < n; ++i) if (INS[i] <= INS[i - 1]) return -4;
    c->M = M; c->P0modM = P0modM; c->R = r; c->n = n; c->split_idx = split_idx; c->hmax = hmax;
    c->rank = rank; c->Icap = mk128(icap_lo, icap_hi); c->INS = INS;
    c->CB = make_cb();
    if (!c->CB) return -2;
    if (rank && c->CB[n][r] > (u128)~0ULL) { free(c->CB); return -5; }
    return 0;
}

/* Single-threaded, first indices in ascending order: the record order of the reference dumper
 * run with T = 1. Records for first indices in [i_lo, i_hi). Returns the count. */
long long or_idump(u64 M, int r, u64 P0modM, u64 icap_lo, u64 icap_hi, const u64 *INS, int n,
                   int split_idx, int hmax, int rank, int i_lo, int i_hi, u64 *keys, u64 *ids, u64 cap) {
    ictx_t c;
    int rc = ictx_init(&c, M, r, P0modM, icap_lo, icap_hi, INS, n, split_idx, hmax, rank);
    if (rc) return rc;
    sink_t s = {keys, ids, cap, 0};
    for (int i0 = i_lo; i0 < i_hi && i0 < n; ++i0) itop(&c, i0, sink_put, &s);
    free(c.CB);
    return (long long)s.total;
}

/* ------------------------------------------------------------------------------------------ */
/* join_sorted: for every I row (in order) the run of equal D keys, emitted D-position order   */
/* (gpu/mitm_gpu2.py:107-119). Both sides key-sorted. Returns the pair count.                 */
static u64 lower_bound(const u64 *a, u64 n, u64 v) {
    u64 lo = 0, hi = n;
    while (lo < hi) { u64 mid = lo + (hi - lo) / 2; if (a[mid] < v) lo = mid + 1; else hi = mid; }
    return lo;
}
long long or_join_sorted(const u64 *Dk, const u64 *Did, u64 nD, const u64 *Ik, const u64 *Iid, u64 nI,
                         u64 *od, u64 *oi, u64 cap) {
    u64 tot = 0;
    for (u64 i = 0; i < nI; ++i) {
        u64 p = lower_bound(Dk, nD, Ik[i]);
        for (; p < nD && Dk[p] == Ik[i]; ++p) {
            if (tot < cap) { od[tot] = Did[p]; oi[tot] = Iid[i]; }
            ++tot;
        }
    }
    return (long long)tot;
}

/* ------------------------------------------------------------------------------------------ */
/* CPU join in the reference's shape (src/mitm64.cpp:51-121): build a hash of the D residues   */
/* on one thread, then T threads take first insertion indices i0 = tid (mod T), run the I DFS */
/* and probe every leaf. Unlike mitm64 (which keeps only the max-prod D per residue,           */
/* src/mitm64.cpp:56), every D record is kept so the output is the full equal-residue pair set */
/* that the GPU join emits. Record sets are those of the two dumpers (Icap exclusive).         */
typedef struct { u64 key, id; int64_t next; } dent_t;
typedef struct {
    dent_t *ent; u64 nent, capent;
    int64_t *head; u64 *slotkey; u64 mask;
} dhash_t;

static inline u64 hslot(u64 k, u64 mask) { return (k * 0x9E3779B97F4A7C15ULL >> 17) & mask; }

static void dh_insert(dhash_t *h, u64 key, u64 id) {
    if (h->nent == h->capent) {
        h->capent = h->capent ? 2 * h->capent : 1024;
        h->ent = (dent_t *)realloc(h->ent, h->capent * sizeof(dent_t));
    }
    u64 s = hslot(key, h->mask);
    while (h->head[s] >= 0 && h->slotkey[s] != key) s = (s + 1) & h->mask;
    if (h->head[s] < 0) h->slotkey[s] = key;
    h->ent[h->nent].key = key;
    h->ent[h->nent].id = id;
    h->ent[h->nent].next = h->head[s];
    h->head[s] = (int64_t)h->nent;
    h->nent++;
}

typedef struct {
    const ictx_t *c; const dhash_t *h; int tid, T;
    u64 *pd, *pi; u64 npairs, cap; u64 leaves;
} worker_t;

static void probe_put(void *user, u64 key, u64 id) {
    worker_t *w = (worker_t *)user;
    w->leaves++;
    const dhash_t *h = w->h;
    u64 s = hslot(key, h->mask);
    while (h->head[s] >= 0) {
        if (h->slotkey[s] == key) {
            for (int64_t e = h->head[s]; e >= 0; e = h->ent[e].next) {
                if (w->npairs == w->cap) {
                    w->cap = w->cap ? 2 * w->cap : 64;
                    w->pd = (u64 *)realloc(w->pd, w->cap * sizeof(u64));
                    w->pi = (u64 *)realloc(w->pi, w->cap * sizeof(u64));
                }
                w->pd[w->npairs] = h->ent[e].id;
                w->pi[w->npairs] = id;
                w->npairs++;
            }
            return;
        }
        s = (s + 1) & h->mask;
    }
}

static void *worker_main(void *arg) {
    worker_t *w = (worker_t *)arg;
    for (int i0 = w->tid; i0 < w->c->n; i0 += w->T) itop(w->c, i0, probe_put, w);
    return NULL;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* stats[0]=D leaves, stats[1]=I leaves, stats[2]=pairs; times[0]=D build s, times[1]=I scan s.
 * Pairs (D id, I id) are written up to cap, sorted by neither side (thread-merge order). */
long long or_cpu_join(u64 M, int r, u64 P0modM, u64 icap_lo, u64 icap_hi, u64 dmin_lo, u64 dmin_hi,
                      const u64 *B, int nB, const u64 *INS, int n, int split_idx, int hmax, int rank,
                      int T, u64 *out_d, u64 *out_i, u64 cap, u64 *stats, double *times) {
    if (T < 1) T = 1;
    if (nB < 1 || nB > 64) return -1;
    double t0 = now_s();
    /* D side: count first to size the table, then insert (one thread, as recD) */
    long long nD = or_ddump(M, r, dmin_lo, dmin_hi, B, nB, 0, NULL, NULL, 0);
    if (nD < 0) return nD;
    u64 *dk = (u64 *)malloc((nD ? nD : 1) * sizeof(u64)), *di = (u64 *)malloc((nD ? nD : 1) * sizeof(u64));
    or_ddump(M, r, dmin_lo, dmin_hi, B, nB, 0, dk, di, (u64)nD);
    dhash_t h = {0};
    u64 slots = 1024;
    while (slots < 2 * (u64)nD) slots <<= 1;
    h.mask = slots - 1;
    h.head = (int64_t *)malloc(slots * sizeof(int64_t));
    h.slotkey = (u64 *)malloc(slots * sizeof(u64));
    for (u64 s = 0; s < slots; ++s) h.head[s] = -1;
    for (long long j = 0; j < nD; ++j) dh_insert(&h, dk[j], di[j]);
    free(dk); free(di);
    double t1 = now_s();

    ictx_t c;
    int rc = ictx_init(&c, M, r, P0modM, icap_lo, icap_hi, INS, n, split_idx, hmax, rank);
    if (rc) { free(h.head); free(h.slotkey); free(h.ent); return rc; }
    worker_t *ws = (worker_t *)calloc(T, sizeof(worker_t));
    pthread_t *th = (pthread_t *)calloc(T, sizeof(pthread_t));
    for (int t = 0; t < T; ++t) {
        ws[t].c = &c; ws[t].h = &h; ws[t].tid = t; ws[t].T = T;
        pthread_create(&th[t], NULL, worker_main, &ws[t]);
    }
    u64 leaves = 0, np = 0;
    for (int t = 0; t < T; ++t) {
        pthread_join(th[t], NULL);
        leaves += ws[t].leaves;
        for (u64 j = 0; j < ws[t].npairs; ++j, ++np)
            if (np < cap) { out_d[np] = ws[t].pd[j]; out_i[np] = ws[t].pi[j]; }
        free(ws[t].pd); free(ws[t].pi);
    }
    double t2 = now_s();
    free(ws); free(th); free(c.CB);
    free(h.head); free(h.slotkey); free(h.ent);
    if (stats) { stats[0] = (u64)nD; stats[1] = leaves; stats[2] = np; }
    if (times) { times[0] = t1 - t0; times[1] = t2 - t1; }
    return (long long)np;
}

/* ------------------------------------------------------------------------------------------ */
/* The CPU join with the roles swapped, for a D side far larger than the I side: the I records   */
/* go into the hash (key -> every id with that key), then T threads sweep the D subsets, thread  */
/* t taking the largest index i = nB-1-t, nB-1-t-T, ... and running the D enumeration below it   */
/* (same order, prune, key and ids as or_ddump), probing each D key. Every (D id, I id) pair of  */
/* equal keys is emitted. Same pair set as or_cpu_join (src/mitm64.cpp:51-121 hashes D instead). */
typedef struct { dhash_t *h; } isink_t;
static void isink_put(void *user, u64 key, u64 id) { dh_insert(((isink_t *)user)->h, key, id); }

typedef struct {
    u64 M; int R; u128 Dmin; const u64 *B; int nB; int bitmask;
    const dhash_t *h; int tid, T;
    u64 *pd, *pi; u64 npairs, cap; u64 leaves;
} dsweep_t;

static void dsweep_emit(dsweep_t *w, u64 key, u64 id) {
    const dhash_t *h = w->h;
    u64 s = hslot(key, h->mask);
    while (h->head[s] >= 0) {
        if (h->slotkey[s] == key) {
            for (int64_t e = h->head[s]; e >= 0; e = h->ent[e].next) {
                if (w->npairs == w->cap) {
                    w->cap = w->cap ? 2 * w->cap : 64;
                    w->pd = (u64 *)realloc(w->pd, w->cap * sizeof(u64));
                    w->pi = (u64 *)realloc(w->pi, w->cap * sizeof(u64));
                }
                w->pd[w->npairs] = id;
                w->pi[w->npairs] = h->ent[e].id;
                w->npairs++;
            }
            return;
        }
        s = (s + 1) & h->mask;
    }
}

static void dsweep_rec(dsweep_t *w, int hi, int depth, u128 prod, u64 res, int *idx) {
    if (depth == w->R) {
        u64 id = 0;
        if (w->bitmask) { for (int j = 0; j < w->R; ++j) id |= 1ULL << idx[j]; }
        else { for (int j = w->R - 1, pos = 0; j >= 0; --j, ++pos) id |= ((u64)idx[j]) << (6 * pos); }
        w->leaves++;
        dsweep_emit(w, res % w->M, id);
        return;
    }
    const int t = w->R - depth - 1;
    for (int i = hi; i >= t; --i) {
        u128 np = prod * w->B[i];
        u128 best = np;
        for (int j = 0; j < t; ++j) best *= w->B[i - 1 - j];
        if (best < w->Dmin) break;
        idx[depth] = i;
        dsweep_rec(w, i - 1, depth + 1, np, mulmod(res, w->B[i] % w->M, w->M), idx);
        if (depth == 0) break;   /* depth 0: this thread's one largest index */
    }
}

static void *dsweep_main(void *arg) {
    dsweep_t *w = (dsweep_t *)arg;
    int idx[16];
    for (int top = w->nB - 1 - w->tid; top >= w->R - 1; top -= w->T) dsweep_rec(w, top, 0, (u128)1, 1 % w->M, idx);
    return NULL;
}

/* stats[0]=D leaves, stats[1]=I leaves, stats[2]=pairs; times[0]=I hash s, times[1]=D sweep s. */
long long or_join_dsweep(u64 M, int r, u64 P0modM, u64 icap_lo, u64 icap_hi, u64 dmin_lo, u64 dmin_hi,
                         const u64 *B, int nB, const u64 *INS, int n, int split_idx, int hmax, int rank,
                         int T, u64 *out_d, u64 *out_i, u64 cap, u64 *stats, double *times) {
    if (T < 1) T = 1;
    if (nB < 1 || nB > 64 || r < 1 || r > 16 || M == 0) return -1;
    for (int i = 1; i < nB; ++i) if (B[i] <= B[i - 1]) return -4;
    double t0 = now_s();
    long long nI = or_idump(M, r, P0modM, icap_lo, icap_hi, INS, n, split_idx, hmax, rank, 0, n, NULL, NULL, 0);
    if (nI < 0) return nI;
    ictx_t c;
    int rc = ictx_init(&c, M, r, P0modM, icap_lo, icap_hi, INS, n, split_idx, hmax, rank);
    if (rc) return rc;
    dhash_t h = {0};
    u64 slots = 1024;
    while (slots < 2 * (u64)nI) slots <<= 1;
    h.mask = slots - 1;
    h.head = (int64_t *)malloc(slots * sizeof(int64_t));
    h.slotkey = (u64 *)malloc(slots * sizeof(u64));
    h.capent = (u64)(nI ? nI : 1);
    h.ent = (dent_t *)malloc(h.capent * sizeof(dent_t));
    if (!h.head || !h.slotkey || !h.ent) { free(h.head); free(h.slotkey); free(h.ent); free(c.CB); return -2; }
    for (u64 s = 0; s < slots; ++s) h.head[s] = -1;
    isink_t is = {&h};
    for (int i0 = 0; i0 < n; ++i0) itop(&c, i0, isink_put, &is);
    free(c.CB);
    double t1 = now_s();
    dsweep_t *ws = (dsweep_t *)calloc(T, sizeof(dsweep_t));
    pthread_t *th = (pthread_t *)calloc(T, sizeof(pthread_t));
    for (int t = 0; t < T; ++t) {
        ws[t] = (dsweep_t){M, r, mk128(dmin_lo, dmin_hi), B, nB, r > 10, &h, t, T, NULL, NULL, 0, 0, 0};
        pthread_create(&th[t], NULL, dsweep_main, &ws[t]);
    }
    u64 leaves = 0, np = 0;
    for (int t = 0; t < T; ++t) {
        pthread_join(th[t], NULL);
        leaves += ws[t].leaves;
        for (u64 j = 0; j < ws[t].npairs; ++j, ++np)
            if (np < cap) { out_d[np] = ws[t].pd[j]; out_i[np] = ws[t].pi[j]; }
        free(ws[t].pd); free(ws[t].pi);
    }
    double t2 = now_s();
    free(ws); free(th);
    free(h.head); free(h.slotkey); free(h.ent);
    if (stats) { stats[0] = leaves; stats[1] = h.nent; stats[2] = np; }
    if (times) { times[0] = t1 - t0; times[1] = t2 - t1; }
    return (long long)np;
}
