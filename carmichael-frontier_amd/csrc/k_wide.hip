// k_wide.hip -- moduli of 2^64 and above.
//
// Keys are u64 throughout the join. An instance whose modulus M is 2^64 or more is joined modulo
// M1, a divisor of M below 2^64: P0 * prod(I) = prod(D) (mod M) implies the same congruence modulo
// every divisor of M, so the pairs modulo M1 contain every pair modulo M. Each pair the join emits
// is then checked exactly modulo M (128-bit residues of prod(D) and P0 * prod(I), rebuilt from the
// pair's ids and the instance's prime tables) and only the exact ones are kept. With M1 >= 2^40 the
// extra pairs are few (about nD * nI / M1), so the check costs nothing next to the join.
#include "cf_internal.h"

typedef unsigned __int128 u128;

// ---- host: the divisor ------------------------------------------------------------------------
namespace {
void divisor_search(const std::vector<std::pair<u128, int>> &f, size_t i, u128 d, u128 cap, u128 &best) {
    if (i == f.size()) {
        if (d > best) best = d;
        return;
    }
    u128 x = d;
    for (int e = 0; e <= f[i].second; ++e) {
        divisor_search(f, i + 1, x, cap, best);
        if (e == f[i].second || x > cap / f[i].first) break;
        x *= f[i].first;
    }
}
}  // namespace

u64 wide_divisor(u128 M) {
    // M = (prime powers found by trial division below 2^20) * cofactor, the cofactor taken as one
    // more factor of exponent 1 (prime or not, it divides M); then the largest divisor <= 2^62
    std::vector<std::pair<u128, int>> f;
    u128 c = M;
    for (u64 q = 2; q < ((u64)1 << 20) && (u128)q * q <= c; q += (q == 2 ? 1 : 2)) {
        if (c % q) continue;
        int e = 0;
        while (c % q == 0) {
            c /= q;
            ++e;
        }
        f.push_back({(u128)q, e});
    }
    if (c > 1) f.push_back({c, 1});
    u128 best = 0;
    divisor_search(f, 0, 1, (u128)1 << 62, best);
    CF_REQUIRE(best >= ((u128)1 << 40), "a modulus of 2^64 or more needs a divisor between 2^40 and 2^62");
    return (u64)best;
}

// ---- device: the exact check ------------------------------------------------------------------
namespace {
constexpr int WB = 256;

// x * p mod M for x < M < 2^126 and p >= 1: double and add over the bits of p (pairs are few)
__device__ u128 mulmod_wide(u128 x, u64 p, u128 M) {
    u128 acc = 0;
    for (int b = 63 - __builtin_clzll(p); b >= 0; --b) {
        acc <<= 1;
        if (acc >= M) acc -= M;
        if ((p >> b) & 1) {
            acc += x;
            if (acc >= M) acc -= M;
        }
    }
    return acc;
}

__global__ void __launch_bounds__(WB) k_wide_check(const u64 *__restrict__ pd, const u64 *__restrict__ pi,
                                                  const u64 *__restrict__ pj, u64 n, const SideDesc *__restrict__ dd,
                                                  const SideDesc *__restrict__ id, const WideMod *__restrict__ wm,
                                                  u32 *__restrict__ keep) {
    for (u64 e = (u64)blockIdx.x * WB + threadIdx.x; e < n; e += (u64)gridDim.x * WB) {
        const u64 j = pj[e];
        const WideMod w = wm[j];
        if (!w.on) {
            keep[e] = 1;
            continue;
        }
        const u128 M = ((u128)w.M_hi << 64) | w.M_lo;
        const SideDesc &D = dd[j];
        const SideDesc &I = id[j];
        // prod(D) mod M: the D id holds the chosen base indices (6-bit fields or a bit mask)
        u128 xd = 1;
        const u64 dv = pd[e];
        if (D.id_mode == ID_MASK) {
            for (u64 m = dv; m; m &= m - 1) xd = mulmod_wide(xd, D.p[__builtin_ctzll(m)], M);
        } else {
            for (int q = 0; q < D.r; ++q) xd = mulmod_wide(xd, D.p[(dv >> (D.id_bits * q)) & 63u], M);
        }
        // P0 * prod(I) mod M: 8-bit fields, or a colex rank decoded greedily from the top position
        u128 xi = ((u128)w.P_hi << 64) | w.P_lo;
        const u64 iv = pi[e];
        if (I.id_mode == ID_PACK_ASC) {
            for (int q = 0; q < I.r; ++q) xi = mulmod_wide(xi, I.p[(iv >> (I.id_bits * q)) & 255u], M);
        } else {
            u64 v = iv;
            int x = I.n - 1;
            for (int t = I.r - 1; t >= 0; --t) {
                while (x > t && I.binom[(size_t)x * (I.r + 1) + t + 1] > v) --x;
                v -= I.binom[(size_t)x * (I.r + 1) + t + 1];
                xi = mulmod_wide(xi, I.p[x], M);
                --x;
            }
        }
        keep[e] = xd == xi ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(WB) k_wide_compact(const u64 *__restrict__ pd, const u64 *__restrict__ pi,
                                                    const u64 *__restrict__ pj, u64 n, const u32 *__restrict__ keep,
                                                    const u64 *__restrict__ pos, u64 *__restrict__ od,
                                                    u64 *__restrict__ oi, u64 *__restrict__ oj,
                                                    unsigned long long *__restrict__ count) {
    for (u64 e = (u64)blockIdx.x * WB + threadIdx.x; e < n; e += (u64)gridDim.x * WB) {
        if (keep[e]) {
            const u64 at = pos[e];
            od[at] = pd[e];
            oi[at] = pi[e];
            oj[at] = pj[e];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = pos[n];
}

unsigned wgrid(u64 n) {
    const u64 g = (n + WB - 1) / WB;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}
}  // namespace

void wide_check_pairs(cf_ctx *ctx, const u64 *pd, const u64 *pi, const u64 *pj, u64 n, const SideDesc *dd,
                      const SideDesc *id, const WideMod *wm, u32 *keep, u64 *pos, u64 *od, u64 *oi, u64 *oj,
                      unsigned long long *count) {
    if (!n) return;
    CF_LAUNCH(ctx, "k_wide_check", k_wide_check, dim3(wgrid(n)), dim3(WB), 0, pd, pi, pj, n, dd, id, wm, keep);
    scan_u32_to_u64_async(ctx, keep, pos, n);
    CF_LAUNCH(ctx, "k_wide_compact", k_wide_compact, dim3(wgrid(n)), dim3(WB), 0, pd, pi, pj, n, keep, pos, od, oi, oj,
              count);
}
