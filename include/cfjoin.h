/* cfjoin.h -- C ABI of libcfjoin.so, the MI355X (gfx950) exchange-MITM residue join.
 *
 * The reference reaches its GPU path through two Python modules (gpu/mitm_gpu2.py,
 * gpu/igen_gpu.py) that JIT CUDA kernels through CuPy, plus two C++ record dumpers reached through
 * pipes (src/mitm64_ddump.cpp, src/mitm64_idump_stream2.cpp). This library replaces the device
 * side of all four; each entry point below names the reference routine it stands in for.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative status: -1 a violated precondition (the
 *    reference raises AssertionError for the same conditions), -2 out of device memory, -3 a HIP
 *    runtime error, -4 an internal consistency check failed. cf_last_error() returns the message of
 *    the last failure on the calling thread.
 *  - Pointers named *_dev are device addresses (from cf_alloc or any HIP allocation on the
 *    context's device); all others are host addresses. Sizes are element counts unless named *bytes.
 *  - One context = one device + one HIP stream. Calls on one context are ordered; a context must
 *    not be used from two threads at once.
 *  - Keys are canonical residues in [0, M), M < 2^64. Ids follow the reference encodings
 *    (packed indices, 64-bit deletion masks or colex ranks).
 */
#ifndef CFJOIN_H
#define CFJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cf_ctx cf_ctx;
typedef struct cf_plan cf_plan;
typedef struct cf_batch cf_batch;

#define CF_OK 0
#define CF_EINVAL -1
#define CF_ENOMEM -2
#define CF_EHIP -3
#define CF_EINTERNAL -4

/* id encodings (cf_plan_create_*) */
#define CF_ID_COLEX 0     /* sum_j C(i_j, j+1): src/mitm64_idump_stream2.cpp:54 (IDUMP_RANK=1) */
#define CF_ID_PACK_ASC 1  /* 8-bit packed ascending: src/mitm64_idump_stream2.cpp:55 */
#define CF_ID_PACK_DESC 2 /* 6-bit packed ascending indices of a descending D choice: src/mitm64_ddump.cpp:33 */
#define CF_ID_MASK 3      /* 64-bit deletion mask: src/mitm64_ddump.cpp:32 */

/* ---- version, errors ------------------------------------------------------------------- */
int cf_version(void);
const char *cf_last_error(void);

/* ---- context and device memory ----------------------------------------------------------
 * Replace the CuPy plumbing the drivers use: cp.asarray / cp.asnumpy / .get()
 * (gpu/mitm_gpu_mstar_driver.py:112,224,254), the memory pool and free_all_blocks()
 * (gpu/mitm_gpu_mstar_driver.py:114,238), cp.cuda.Device().mem_info (:75). */
int cf_ctx_create(int device, cf_ctx **out);
int cf_ctx_destroy(cf_ctx *ctx);
int cf_alloc(cf_ctx *ctx, uint64_t bytes, void **out_dev);
int cf_free(cf_ctx *ctx, void *p_dev);
int cf_pool_release(cf_ctx *ctx);                                   /* free_all_blocks() */
int cf_pool_stats(cf_ctx *ctx, uint64_t *live_bytes, uint64_t *cached_bytes);
int cf_mem_info(cf_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);
int cf_memcpy_h2d(cf_ctx *ctx, void *dst_dev, const void *src, uint64_t bytes);
int cf_memcpy_d2h(cf_ctx *ctx, void *dst, const void *src_dev, uint64_t bytes);  /* synchronous */
int cf_memcpy_d2d(cf_ctx *ctx, void *dst_dev, const void *src_dev, uint64_t bytes);
int cf_synchronize(cf_ctx *ctx);

/* ---- kernel timing (no counterpart in the reference; its drivers time phases on the host,
 * gpu/mitm_gpu_mstar_driver.py:55,241) -------------------------------------------------------
 * While enabled, launches of the named kernel (all kernels if kernel == NULL) are bracketed by
 * HIP events on the context's stream; cf_prof_read returns the summed device time and the
 * launch count since the last reset. */
int cf_prof_enable(cf_ctx *ctx, const char *kernel);
int cf_prof_disable(cf_ctx *ctx);
int cf_prof_reset(cf_ctx *ctx);
int cf_prof_read(cf_ctx *ctx, const char *kernel, double *ms, uint64_t *launches);
/* every timed kernel as text lines "name total_ms launches" (truncated to size - 1 bytes) */
int cf_prof_dump(cf_ctx *ctx, char *buf, uint64_t size);

/* ---- r-subsets by rank range ------------------------------------------------------------
 * Replaces gen_records + generate() (gpu/mitm_gpu2.py:25-98): records of the r-combinations of
 * primes[0..n) with combinatorial-number-system rank in [lo, lo+count), in rank order;
 * key = mult * prod mod M, id = the r indices packed `bits` wide, position 0 lowest.
 * Preconditions as gpu/mitm_gpu2.py:83-88 except that any M < 2^64 is accepted. */
int cf_gen_subsets(cf_ctx *ctx, const uint64_t *primes, int n, int r, int bits, uint64_t M, uint64_t mult,
                   uint64_t lo, uint64_t count, uint64_t *keys_dev, uint64_t *ids_dev);

/* ---- sort one side ----------------------------------------------------------------------
 * Replaces sort_side() (gpu/mitm_gpu2.py:101-104): stable sort of (keys, ids) by key, in place. */
int cf_sort_pairs(cf_ctx *ctx, uint64_t *keys_dev, uint64_t *ids_dev, uint64_t n);
/* The same sort into new arrays (sort_side returns new arrays, gpu/mitm_gpu2.py:101-104): the
 * inputs are left unchanged; the output pair may also be the input pair. key_bound: 0, or a bound
 * every key is below (the modulus the keys were reduced by), which spares the pass that finds the
 * largest key; a key at or above a non-zero bound leaves the order undefined. */
int cf_sort_pairs_into(cf_ctx *ctx, const uint64_t *keys_dev, const uint64_t *ids_dev, uint64_t *keys_out_dev,
                       uint64_t *ids_out_dev, uint64_t n, uint64_t key_bound);

/* ---- join two sorted sides --------------------------------------------------------------
 * Replaces join_sorted() (gpu/mitm_gpu2.py:107-119): every (D, I) with equal key, I-major then
 * D order. Output arrays are allocated from the context pool (release with cf_free); when the
 * join is empty both are NULL. */
int cf_join_sorted(cf_ctx *ctx, const uint64_t *Dk_dev, const uint64_t *Did_dev, uint64_t nD,
                   const uint64_t *Ik_dev, const uint64_t *Iid_dev, uint64_t nI,
                   uint64_t **out_d_dev, uint64_t **out_i_dev, uint64_t *npairs);

/* ---- join two sorted sides without allocating ---------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): cf_join_sorted with the
 * outputs supplied by the caller and every intermediate in a static per-device scratch of the
 * library (32 KiB, resident while the library is loaded); calls no allocator.
 * Requires nI <= CF_JOIN_NOALLOC_MAX_I (join a longer sorted I side in slices: the pairs of a slice
 * are that slice's part of the I-major output). Writes the pairs to out_d_dev / out_i_dev in
 * cf_join_sorted's order and sets *npairs; if *npairs exceeds cap, nothing is written and
 * CF_EINVAL is returned with *npairs holding the count needed. Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (two streams, two contexts or two threads) race. */
#define CF_JOIN_NOALLOC_MAX_I (1u << 20)
int cf_join_sorted_noalloc(cf_ctx *ctx, const uint64_t *Dk_dev, const uint64_t *Did_dev, uint64_t nD,
                           const uint64_t *Ik_dev, const uint64_t *Iid_dev, uint64_t nI,
                           uint64_t *out_d_dev, uint64_t *out_i_dev, uint64_t cap, uint64_t *npairs);

/* ---- filtered enumeration plans ---------------------------------------------------------
 * cf_plan_create_i replaces the igen kernel and its host set-up (gpu/igen_gpu.py:30-111) and
 * the record set of src/mitm64_idump_stream2.cpp:47-57: ascending r-subsets I of INS[0..n)
 * with prod(I) < Icap (Icap = icap_hi:icap_lo), at most hmax indices >= split_idx;
 * key = P0modM * prod(I) mod M. Preconditions of gpu/igen_gpu.py:103-105 (1 <= r <= 16,
 * n <= 1024, INS strictly ascending); Icap <= 1 gives an empty plan.
 * cf_plan_create_d replaces src/mitm64_ddump.cpp:29-55: r-subsets D of the ascending base
 * B[0..nB), nB <= 64, with prod(D) >= Dmin, key = prod(D) mod M, enumerated in the dumper's
 * descending order (id_mode CF_ID_PACK_DESC or CF_ID_MASK).
 * Creating a plan expands the enumeration tree down to the last level; records are written
 * in the reference's depth-first order. */
int cf_plan_create_i(cf_ctx *ctx, const uint64_t *INS, int n, int r, uint64_t M, uint64_t P0modM,
                     uint64_t icap_lo, uint64_t icap_hi, int split_idx, int hmax, int id_mode, cf_plan **out);
int cf_plan_create_d(cf_ctx *ctx, const uint64_t *B, int nB, int r, uint64_t M, uint64_t dmin_lo,
                     uint64_t dmin_hi, int id_mode, cf_plan **out);
int cf_plan_total(cf_plan *plan, uint64_t *total);
/* Records below each first (depth-0) index, counts[0..n): the per-first-index counts of the
 * count pass of gpu/igen_gpu.py:125-127. */
int cf_plan_first_counts(cf_plan *plan, uint64_t *counts);
/* Writes the records whose first index lies in [first_lo, first_hi), in order: one write pass
 * of gpu/igen_gpu.py:135-139 (there the count/write equality check, here *written). */
int cf_plan_write(cf_plan *plan, int first_lo, int first_hi, uint64_t *keys_dev, uint64_t *ids_dev,
                  uint64_t *written);
int cf_plan_destroy(cf_plan *plan);

/* ---- whole instances ---------------------------------------------------------------------
 * The device half of gpu/mitm_gpu_mstar_driver.py:184-243 (build_D with DFILTER, I generation,
 * sort, join against every D chunk) for one instance per call (cf_solve) or for many instances
 * in one set of launches (cf_solve_batch): every equal-residue pair (D id, I id). Output arrays
 * come from the context pool (release with cf_free); all NULL when there is no pair. */
typedef struct cf_instance {
    uint64_t M;
    int32_t r;
    int32_t nB;
    const uint64_t *B;        /* base, ascending, nB <= 64 */
    int32_t nI;
    int32_t hmax;
    const uint64_t *INS;      /* insertion pool, ascending, nI <= 1024 */
    uint64_t P0modM;
    uint64_t icap_lo, icap_hi;  /* exclusive bound on prod(I) */
    uint64_t dmin_lo, dmin_hi;  /* inclusive bound on prod(D) */
    int32_t split_idx;
    int32_t i_id_mode;        /* CF_ID_COLEX or CF_ID_PACK_ASC (8 bits) */
    int32_t d_id_mode;        /* CF_ID_PACK_DESC (6 bits) or CF_ID_MASK */
    int32_t pad;
    uint64_t M_hi;            /* M = M_hi * 2^64 + M: moduli up to 2^126 (see below) */
    uint64_t P0modM_hi;       /* P0 mod M = P0modM_hi * 2^64 + P0modM */
} cf_instance;
/* Moduli of 2^64 and above (M_hi != 0), which the reference's runner leaves unsearched
 * (gpu/run_mstar_family.py:70-73): keys are residues modulo M1, the largest divisor of M in
 * [2^40, 2^62] (M factored by trial division below 2^20, the cofactor kept whole), and every pair of
 * that join is checked exactly modulo M on the device before it is returned, so the pair set is the
 * join's modulo M. Such an instance takes the partitioned solve modes only; a batch still needs
 * the sum of its instances' key moduli (M1 for these) below 2^64. */

typedef struct cf_stats {
    uint64_t nD, nI, npairs;
    uint64_t peak_bytes;      /* largest pool footprint during the call */
    double ms_total;          /* device time, first kernel to pairs ready */
    double ms_gen_d, ms_sort_d, ms_gen_i, ms_sort_i, ms_join;
    uint64_t nD_parents, nI_parents;   /* last-level tree nodes (the leaves' parents) of each side */
} cf_stats;

/* mode: CF_SOLVE_PARTITION -- D records counting-sorted into hash buckets behind a fingerprint
 *       table; I leaves, in chunks of at most max_chunk (0: as many as fit in device memory, at
 *       most 2^31), partitioned by hash range and probed against the table slice of their
 *       partition; CF_SOLVE_PARTITION_DGLOBAL -- the same with the D side bucketed by per-record
 *       global atomics (the path very large D sides take); CF_SOLVE_SORTMERGE -- I chunks radix
 *       sorted and joined as join_sorted(). Same pair set. */
#define CF_SOLVE_PARTITION 0
#define CF_SOLVE_SORTMERGE 1
#define CF_SOLVE_PARTITION_DGLOBAL 2
int cf_solve(cf_ctx *ctx, const cf_instance *inst, int mode, uint64_t max_chunk, uint64_t **out_d_dev,
             uint64_t **out_i_dev, uint64_t *npairs, cf_stats *stats);

/* n instances of any radii in one solve: each instance's keys occupy their own range of one key
 * space, so one D sort, one probe table and one set of launches per radius serve the batch.
 * out_inst_dev[k] is the index (into insts) of pair k; inst_counts (host, 2n entries) receives
 * nD, nI of every instance. Requires the sum of the moduli to stay below 2^64 - 1. */
int cf_solve_batch(cf_ctx *ctx, const cf_instance *insts, int n, uint64_t **out_d_dev, uint64_t **out_i_dev,
                   uint64_t **out_inst_dev, uint64_t *npairs, uint64_t *inst_counts, cf_stats *stats);

/* The same in two steps: cf_batch_create validates the instances and puts their descriptors and
 * exact bound tables on the device (the host-side instance construction the reference's drivers
 * do before any GPU work, gpu/mitm_gpu_mstar_driver.py:184-208, gpu/igen_gpu.py:91-111);
 * cf_batch_solve is then device work only and may be called any number of times. Instances may be
 * freed after cf_batch_create. max_chunk bounds the I records partitioned at once (0: as many as
 * fit, at most 2^31). */
int cf_batch_create(cf_ctx *ctx, const cf_instance *insts, int n, cf_batch **out);
int cf_batch_solve(cf_batch *batch, uint64_t max_chunk, uint64_t **out_d_dev, uint64_t **out_i_dev,
                   uint64_t **out_inst_dev, uint64_t *npairs, uint64_t *inst_counts, cf_stats *stats);
int cf_batch_destroy(cf_batch *batch);

/* Speed knobs of a batch's solves; results never depend on them (0 keeps the default). Chosen by
 * the autotuner (cfjoin/autotune.py), which times candidates on the batch itself. No counterpart
 * in the reference, whose kernels run with fixed launch shapes (gpu/mitm_gpu2.py:93-97). */
typedef struct cf_tuning {
    int32_t slice_kib;        /* probe-table slice per partition, KiB (default 2048) */
    int32_t d_per_bucket;     /* D records per table bucket (default 2, at most 12) */
    int32_t seg_fill_pct;     /* expected fill of an I capacity segment, percent (default 70, at most 90) */
    int32_t probe_blocks;     /* probe grid, 1024-thread blocks (default 256: one per CU) */
} cf_tuning;
int cf_batch_set_tuning(cf_batch *batch, const cf_tuning *tuning);

/* Device memory a solve of the batch needs at its peak (bytes, an estimate with headroom) and its
 * exact record totals, from the enumeration trees' level counts (no leaf is generated). Used to
 * admit work by memory footprint; no counterpart in the reference. */
int cf_batch_footprint(cf_batch *batch, uint64_t *bytes, uint64_t *nD_total, uint64_t *nI_total);

/* The batch's persistent launch: one eager solve notes every value the host reads back during a
 * solve (level sizes, candidate and pair counts), then the whole solve is captured into a HIP
 * graph with those values, and every later cf_batch_solve (with max_chunk 0 or the captured one)
 * is a single graph launch followed by the read-back of the pairs -- no host round trip inside
 * the solve. The graph keeps its device buffers until cf_batch_destroy. The reference launches
 * every kernel of every instance from Python (gpu/mitm_gpu_mstar_driver.py:209-243); this has
 * no counterpart there. */
int cf_batch_capture(cf_batch *batch, uint64_t max_chunk);

/* ---- integer issue rate ---------------------------------------------------------------------
 * No counterpart in the reference. Measures this GPU's vector-ALU issue rate for one integer
 * instruction (the denominator of the integer-ALU roofline the benchmark reports): every SIMD of
 * every CU runs independent chains of the instruction for `iters` loop trips; returns the wave-level
 * instructions executed and the elapsed device time. */
#define CF_PEAK_ADD_U32 0
#define CF_PEAK_MUL_LO_U32 1
#define CF_PEAK_MUL_HI_U32 2
#define CF_PEAK_MAD_U64_U32 3
int cf_peak_valu(cf_ctx *ctx, int op, uint64_t iters, uint64_t *wave_instructions, double *ms);

/* ---- serving loop (continuous batching) -------------------------------------------------------
 * Requests (instances) are submitted as they arrive and solved by `workers` threads, each with its
 * own stream on the device: a free worker takes everything that has arrived (up to max_batch) and
 * solves it as one batch; a batch starts only while the priced footprints (cf_batch_footprint) of
 * the batches on the device stay within mem_budget bytes (0: half the free device memory), or
 * alone if it is larger. Each result equals the request's solve on its own. The reference runs
 * instances strictly one after another (gpu/run_mstar_family.py); this has no counterpart there. */
typedef struct cf_server cf_server;
typedef struct cf_served {
    uint64_t tag;
    int32_t status;           /* 0, or the failed call's status (error holds the message) */
    int32_t batch_size;       /* requests solved in the same batch */
    uint64_t nD, nI, npairs;
    const uint64_t *pairs_d;  /* host arrays, valid until cf_server_release(tag) */
    const uint64_t *pairs_i;
    double t_submit_us, t_start_us, t_done_us;   /* since cf_server_create */
    char error[160];
} cf_served;
int cf_server_create(int device, int workers, int max_batch, uint64_t mem_budget, cf_server **out);
/* the instance (and its prime lists) is copied; a tag stays in use from its submission until
 * cf_server_release, and submitting a tag in use fails with CF_EINVAL */
int cf_server_submit(cf_server *server, const cf_instance *inst, uint64_t tag);
/* 1: *out holds the next finished request; 0: none finished within timeout_ms */
int cf_server_next(cf_server *server, int timeout_ms, cf_served *out);
/* frees the result's arrays and frees its tag; fails with CF_EINVAL for a tag whose result was
 * not handed out yet (an unknown tag is a no-op) */
int cf_server_release(cf_server *server, uint64_t tag);
int cf_server_pending(cf_server *server, uint64_t *queued, uint64_t *outstanding);
/* finishes every submitted request, then stops the workers */
int cf_server_destroy(cf_server *server);

#ifdef __cplusplus
}
#endif
#endif
