// modarith.h -- exact residue arithmetic mod an even 64-bit modulus, host and device.
//
// Every q in the pool Q(M) is an odd prime with (q - 1) | M, so M is even and plain
// Montgomery reduction does not apply. Write M = 2^a * m with m odd. A running product is kept
// as the pair
//     xm = (value mod m)        in [0, m), plain (not Montgomery) form
//     x2 = (value mod 2^64)     whose low a bits are value mod 2^a
// Multiplying by a prime p uses the table entry pR = (p mod m) * 2^64 mod m, so one Montgomery
// product REDC(xm * pR) = xm * p mod m keeps xm in plain form (Montgomery 1985, "Modular
// multiplication without trial division"); x2 takes a wrapping 64-bit product. The canonical key
// in [0, M) is recovered by Garner's CRT step:
//     key = xm + m * (((x2 - xm) * m^-1) mod 2^a).
// Only 64-bit multiply-low and multiply-high are used on the device: no 128/64 division.
//
// The join's leaves take one cheaper step, cf_redc32_mul: for a prime p < 2^32 a 32-bit Montgomery
// step gives x * p * 2^-32 mod m from five 32x32 multiplies. Every leaf of an instance, D side and
// I side alike, is exactly one such step below a plain-form node, so all of them carry the same
// unit factor 2^-32 and equal residues stay equal: the join compares these scaled residues.
#pragma once
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

typedef uint64_t u64;
typedef unsigned int u32;
typedef unsigned char u8_t;

#ifdef __HIPCC__
#define CF_HD __host__ __device__ __forceinline__
#else
#define CF_HD inline
#endif

CF_HD u64 cf_mulhi(u64 a, u64 b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (u64)(((unsigned __int128)a * b) >> 64);
#endif
}

// 128-bit unsigned value as two 64-bit halves; products are exact while the caller's bound
// guarantees they stay below 2^128.
struct U128 {
    u64 lo, hi;
};

CF_HD U128 u128_mul64(U128 a, u64 b) {
    U128 r;
    r.lo = a.lo * b;
    r.hi = a.hi * b + cf_mulhi(a.lo, b);
    return r;
}
CF_HD bool u128_gt(U128 a, U128 b) { return a.hi > b.hi || (a.hi == b.hi && a.lo > b.lo); }
CF_HD bool u128_lt(U128 a, U128 b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// Per-modulus constants.
struct ModCtx {
    u64 M;      // modulus, 1 <= M < 2^64
    u64 m;      // odd part of M
    u64 mneg;   // -m^-1 mod 2^64
    u64 minv;   // m^-1 mod 2^64 (only its low a bits are used)
    u64 mask2;  // 2^a - 1
};

// Montgomery reduction: (hi:lo) < m * 2^64  ->  (hi:lo) * 2^-64 mod m, in [0, m).
CF_HD u64 cf_redc(u64 lo, u64 hi, u64 m, u64 mneg) {
    const u64 u = lo * mneg;
    const u64 uh = cf_mulhi(u, m);
    // lo + lo(u*m) == 0 mod 2^64, so the low half carries exactly when lo != 0
    const u64 s = hi + uh;
    const u64 c1 = s < hi;
    const u64 s2 = s + (lo != 0);
    const u64 c2 = s2 < s;
    return (c1 | c2 | (s2 >= m)) ? s2 - m : s2;
}

// xm * p mod m for xm < m, given pR = p * 2^64 mod m.
CF_HD u64 cf_mulmod_m(u64 xm, u64 pR, const ModCtx &c) {
    return cf_redc(xm * pR, cf_mulhi(xm, pR), c.m, c.mneg);
}

// x * p * 2^-32 mod m for x < m and p < 2^32, with mneg32 = -m^-1 mod 2^32 (m odd).
// T = x * p < m * 2^32 is held as hi = T >> 32 and the low word of lo; q makes T + q * m divisible
// by 2^32, and (T + q * m) / 2^32 < 2m, which can pass 2^64 when m > 2^63 (caught by s < hi).
CF_HD u64 cf_redc32_mul(u64 x, u32 p, u64 m, u32 mneg32) {
    const u64 lo = (u64)(u32)x * p;
    const u64 hi = (u64)(u32)(x >> 32) * p + (lo >> 32);
    const u32 q = (u32)lo * mneg32;
    const u64 c = ((u64)q * (u32)m + (u32)lo) >> 32;
    const u64 t = (u64)q * (u32)(m >> 32) + c;
    const u64 s = hi + t;
    return (s < hi || s >= m) ? s - m : s;
}

// Canonical residue in [0, M) from (value mod m, value mod 2^64).
CF_HD u64 cf_key(u64 xm, u64 x2, const ModCtx &c) {
    const u64 t = ((x2 - xm) * c.minv) & c.mask2;
    return xm + c.m * t;
}

// ---- host-side set-up (exact, may use 128-bit division: host only) ---------------------------
inline u64 cf_inv64(u64 m) {  // m odd: m^-1 mod 2^64 by Newton's iteration
    u64 x = m;                  // correct to 3 bits
    for (int i = 0; i < 5; ++i) x *= 2 - m * x;
    return x;
}

inline ModCtx cf_modctx(u64 M) {
    ModCtx c;
    c.M = M;
    int a = 0;
    u64 m = M;
    while (m && !(m & 1)) { m >>= 1; ++a; }
    c.m = m;
    c.minv = cf_inv64(m);
    c.mneg = (u64)0 - c.minv;
    c.mask2 = (a >= 64) ? ~0ULL : ((1ULL << a) - 1);
    return c;
}

// p * 2^64 mod m
inline u64 cf_montmul_factor(u64 p, const ModCtx &c) {
    unsigned __int128 R = ((unsigned __int128)1 << 64) % c.m;
    return (u64)(((unsigned __int128)(p % c.m) * (u64)R) % c.m);
}
