// keyhash.h -- bucket geometry of the partitioned join, shared by host and device code.
//
// A leaf's 64-bit hash h decides where its record goes:
//   partition p = top lgP bits of h        (P = 2^lgP probe partitions)
//   range     d = top lgR bits of h        (R = 2^lgR hash ranges, lgR >= lgP; each partition is
//                                           2^(lgR - lgP) consecutive ranges)
//   bucket    b = d * KR + ((next 32 - lgR bits of h) * KR >> (32 - lgR))
// so range d owns exactly the buckets [d * KR, (d + 1) * KR) and partition p the buckets
// [p * K, (p + 1) * K), K = KR << (lgR - lgP): a bucket never straddles a range or a partition,
// and the D side can be counting-sorted one range at a time. NB = KR << lgR buckets in all.
//   fingerprint f = bits 17..31 of h (15 bits, below every bit that places the record)
#pragma once
#include "modarith.h"

struct KeyHash {
    int lgP;    // 4 <= lgP <= CF_MAX_LGP
    int lgR;    // lgP <= lgR <= 24
    int fpb;    // fingerprint bits kept in an I record (<= 15)
    int pad;
    u64 KR;     // buckets per range
};

constexpr u64 CF_RANGE_BUCKETS_LDS = 8192;   // most buckets per hash range
constexpr int CF_RANGE_BITS_LDS = 12;         // most hash-range bits (2^12 ranges)
constexpr int CF_TABLE_SLOTS = 8;             // 16-bit fingerprint slots per bucket (16 bytes)
constexpr int CF_MAX_LGP = 9;                 // most partition bits (512 partitions; the scatter's
                                              // 512-partition kernels take geometries above 256)

CF_HD u64 kh_nb(const KeyHash &kh) { return kh.KR << kh.lgR; }
CF_HD u64 kh_k(const KeyHash &kh) { return kh.KR << (kh.lgR - kh.lgP); }
CF_HD u32 hpart(u64 h, int lgP) { return (u32)(h >> (64 - lgP)); }
CF_HD u32 hrange(u64 h, const KeyHash &kh) { return (u32)(h >> (64 - kh.lgR)); }
CF_HD u64 hbucket(u64 h, const KeyHash &kh) {
    const u32 r = (u32)(h >> 32) & ((1u << (32 - kh.lgR)) - 1u);
    return (u64)hrange(h, kh) * kh.KR + (((u64)r * kh.KR) >> (32 - kh.lgR));
}
// bucket offset within the record's partition, < K
CF_HD u32 hoff(u64 h, const KeyHash &kh) { return (u32)(hbucket(h, kh) - (u64)hpart(h, kh.lgP) * kh_k(kh)); }
CF_HD u32 hfp(u64 h) { return (u32)(h >> 17) & 0x7FFFu; }
// The same offset from the high word of h in 32-bit arithmetic: the range within the partition
// times KR, plus floor(r * KR / 2^(32 - lgR)) taken as the high word of (r << lgR) * KR (r has
// 32 - lgR bits, so r << lgR fits 32 bits; KR < 2^32 and the offset < K <= 2^24).
CF_HD u32 umulhi32(u32 a, u32 b) { return (u32)(((u64)a * b) >> 32); }
CF_HD u32 hoff32(u32 hi, const KeyHash &kh) {
    const u32 rl = (hi >> (32 - kh.lgR)) & ((1u << (kh.lgR - kh.lgP)) - 1u);
    return rl * (u32)kh.KR + umulhi32(hi << kh.lgR, (u32)kh.KR);
}

// Geometry for a radius group with nD D records: about 2 D records per bucket by default, so an
// 8-slot bucket rarely overflows (Poisson tail P(X > 8) ~ 2.4e-4); P = 2^lgP probe partitions, as few
// as keep one partition's slice of the 16-byte-per-bucket table within about 2 MB by default (an
// XCD's L2 holds it while the partition is probed), 16 <= P <= 256; as few hash ranges as keep a
// range within 8192 buckets (at most 2^12 ranges). An I record keeps the bucket offset within its
// partition and 32 - that many bits of fingerprint, at most 15. Returns false if a partition would
// need more than 2^24 buckets.
inline int kh_bitlen(u64 v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}
// d_per_bucket and slice_bytes are the tunable targets (defaults 2 and 2 MiB). A D side too large
// for 2^24 buckets per partition at that many records per bucket takes the fewest more per bucket
// that fit (such buckets overflow more often, which costs candidates, not results).
inline bool kh_choose_at(u64 nD, KeyHash &kh, u64 dpb, u64 slice_bytes) {
    const u64 want = nD / dpb > 16 ? nD / dpb : 16;
    int lgP = 4;
    while (lgP < 8 && want * (2 * CF_TABLE_SLOTS) / ((u64)1 << lgP) > slice_bytes) ++lgP;
    // 512 partitions shorten the scatters' runs (8 records per partition and tile): taken only when
    // 256 would leave slices above 1.5 times the target (large D sides; a batch's radius group just
    // above the target keeps 256)
    if (lgP == 8 && CF_MAX_LGP > 8 && 2 * (want * (2 * CF_TABLE_SLOTS) / 256) > 3 * slice_bytes) lgP = 9;
    int lgR = lgP;
    while (lgR < CF_RANGE_BITS_LDS && ((want + ((u64)1 << lgR) - 1) >> lgR) > CF_RANGE_BUCKETS_LDS) ++lgR;
    const u64 KR = (want + ((u64)1 << lgR) - 1) >> lgR;
    const int obits = kh_bitlen((KR << (lgR - lgP)) - 1);
    kh.lgP = lgP;
    kh.lgR = lgR;
    kh.KR = KR;
    kh.fpb = 32 - obits < 15 ? 32 - obits : 15;
    kh.pad = 0;
    return obits <= 24;
}
inline bool kh_choose(u64 nD, KeyHash &kh, int d_per_bucket = 2, u64 slice_bytes = 2u << 20) {
    for (u64 dpb = d_per_bucket > 0 ? (u64)d_per_bucket : 2; dpb <= 64; ++dpb)
        if (kh_choose_at(nD, kh, dpb, slice_bytes)) return true;
    return false;
}
