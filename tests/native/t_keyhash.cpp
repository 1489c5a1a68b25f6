// Host check of the join's bucket geometry (carmichael-frontier_amd/csrc/keyhash.h): for the
// geometries chosen for D sides of 1 to 10^10 records, every hash lands in a bucket of its own hash
// range and of its own partition, its offset in the partition fits the I record beside the
// fingerprint and equals the 32-bit form the leaf kernels compute (hoff32), and the last and first
// hashes of every range stay on their side of the boundary.
// A bucket map scaled per partition with K = ceil(NB / P) buckets each, K not a multiple of the
// ranges per partition, lets a range's last hashes reach the next range's first bucket: that map
// must fail the same boundary check. Exit status 0 = all hold.
#include "keyhash.h"

#include <cstdio>
#include <random>

static long bad = 0, tot = 0;

static void check(const KeyHash &kh, u64 h) {
    const u64 b = hbucket(h, kh);
    const u64 d = hrange(h, kh), p = hpart(h, kh.lgP), K = kh_k(kh);
    const u32 o = hoff(h, kh);
    ++tot;
    const bool ok = hoff32((u32)(h >> 32), kh) == o && b < kh_nb(kh) && b >= d * kh.KR && b < (d + 1) * kh.KR && b >= p * K && b < (p + 1) * K &&
                    o < K && ((u64)o << kh.fpb) < (1ULL << 32) && (d >> (kh.lgR - kh.lgP)) == p;
    if (!ok && bad++ < 5)
        printf("lgP=%d lgR=%d KR=%llu h=%016llx b=%llu\n", kh.lgP, kh.lgR, (unsigned long long)kh.KR,
               (unsigned long long)h, (unsigned long long)b);
}

// the partition-scaled map: bucket = p * K + (next 32 - lgP bits) * K >> (32 - lgP)
static u64 bucket_per_partition(u64 h, int lgP, u64 K) {
    const u32 r = (u32)(h >> 32) & ((1u << (32 - lgP)) - 1u);
    return (u64)hpart(h, lgP) * K + (((u64)r * K) >> (32 - lgP));
}

int main() {
    std::mt19937_64 g(7);
    const u64 sizes[] = {1, 10, 80, 1000, 8198, 122155, 333461, 1624827, 2082166, 5907791, 8340300, 18426621,
                         92521742, 200000000, 1000000000, 4426165368ULL, 10000000000ULL};
    int planted_caught = 0, planted_tried = 0;
    for (u64 nD : sizes) {
        KeyHash kh{};
        if (!kh_choose(nD, kh)) {
            printf("no geometry for nD=%llu\n", (unsigned long long)nD);
            return 1;
        }
        for (int i = 0; i < 200000; ++i) check(kh, g());
        const u64 R = 1ULL << kh.lgR;
        const u64 Kp = (kh_nb(kh) + (1ULL << kh.lgP) - 1) / (1ULL << kh.lgP) + 1;   // not a multiple
        bool planted_bad = false;
        for (u64 d = 0; d < R; ++d) {
            const u64 first = d << (64 - kh.lgR);
            const u64 last = first + ((1ULL << (64 - kh.lgR)) - 1);
            check(kh, first);
            check(kh, last);
            // planted map: the first bucket of range d + 1 must lie above every bucket of range d
            if (d + 1 < R && bucket_per_partition(last, kh.lgP, Kp) >=
                                 bucket_per_partition(last + 1, kh.lgP, Kp))
                planted_bad = true;
        }
        if (kh.lgR > kh.lgP) {
            ++planted_tried;
            planted_caught += planted_bad;
        }
    }
    printf("tot=%ld bad=%ld planted %d/%d\n", tot, bad, planted_caught, planted_tried);
    return (bad != 0 || planted_tried == 0 || planted_caught == 0) ? 1 : 0;
}
