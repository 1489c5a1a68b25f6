// Host check of the device residue arithmetic (carmichael-frontier_amd/csrc/modarith.h) against
// exact 128-bit integer arithmetic: running products by the plain Montgomery step and Garner's
// key, the 32-bit scaled leaf step, and the 128-bit helpers. Edge moduli include 1, powers of two,
// 2^63 + 1 and odd parts just below 2^64. Exit status 0 = all agree.
#include "modarith.h"

#include <cstdio>
#include <random>

typedef unsigned __int128 u128;

int main() {
    std::mt19937_64 g(1);
    const u64 Ms[] = {1768248177696000ULL, 155016423578016000ULL, 1, 2, 3, 1ULL << 63, (1ULL << 63) + 1, ~0ULL,
                      ~0ULL - 1, 12, 97, 1ULL << 40, 0xFFFFFFFFFFFFFFC5ULL, 1008, 0xFFFFFFFFULL * 2};
    long bad = 0, tot = 0;
    for (u64 M : Ms)
        for (int rep = 0; rep < 2000; ++rep) {
            const ModCtx c = cf_modctx(M);
            const u64 mult = g() % M;
            u64 xm = mult % c.m, x2 = mult, ref = mult % M;
            for (int k = 0; k < 10; ++k) {
                const u64 p = (k & 1) ? (g() % 200000) | 1 : g();
                xm = cf_mulmod_m(xm, cf_montmul_factor(p, c), c);
                x2 *= p;
                ref = (u64)(((u128)ref * (p % M)) % M);
                ++tot;
                if (cf_key(xm, x2, c) != ref && bad++ < 5) printf("plain step M=%llu k=%d\n", (unsigned long long)M, k);
            }
            // 32-bit scaled step: x * p * 2^-32 mod m, for x < m and p < 2^32
            for (int k = 0; k < 8; ++k) {
                const u64 x = g() % c.m;
                const u32 p = (k == 0) ? 0xFFFFFFFFu : (k == 1 ? 1u : (u32)g());
                const u64 got = cf_redc32_mul(x, p, c.m, (u32)c.mneg);
                // check got * 2^32 == x * p (mod m) and got < m
                const u64 lhs = (u64)(((u128)got << 32) % c.m);
                const u64 rhs = (u64)((u128)x * p % c.m);
                ++tot;
                if ((got >= c.m && c.m > 1) || (c.m == 1 && got != 0) || lhs != rhs) {
                    if (bad++ < 5) printf("redc32 M=%llu x=%llu p=%u got=%llu\n", (unsigned long long)M,
                                          (unsigned long long)x, p, (unsigned long long)got);
                }
            }
        }
    // largest odd parts: x = m - 1 and p = 2^32 - 1 push (T + q m) / 2^32 past 2^64
    for (u64 m : {~0ULL, ~0ULL - 2, (1ULL << 63) + 1, 0xFFFFFFFFFFFFFFC5ULL}) {
        const ModCtx c = cf_modctx(m);
        for (u64 x : {m - 1, m - 2, (u64)1, (u64)0})
            for (u32 p : {0xFFFFFFFFu, 0xFFFFFFFBu, 3u}) {
                const u64 got = cf_redc32_mul(x, p, c.m, (u32)c.mneg);
                ++tot;
                if (got >= m || (u64)(((u128)got << 32) % m) != (u64)((u128)x * p % m)) bad++;
            }
    }
    for (int i = 0; i < 100000; ++i) {
        const u128 a = ((u128)(g() >> 40) << 64) | g();
        const u64 b = g() >> 20;
        const U128 R = u128_mul64(U128{(u64)a, (u64)(a >> 64)}, b);
        const u128 r = a * b;
        ++tot;
        if (R.lo != (u64)r || R.hi != (u64)(r >> 64)) bad++;
    }
    printf("tot=%ld bad=%ld\n", tot, bad);
    return bad != 0;
}
