"""Moduli of 2^64 and above on the GPU, against the exact Python join of the oracle.

The instances are the neighbourhood-census rows of 65 to 74 bits (cfjoin/data/workloads.json,
family nbhd_wide; the reference's runner leaves them unsearched, gpu/run_mstar_family.py:70-73), each
with P0 chosen so that a known pair (D0, I0) must meet modulo M ("planted"), and again with P0 moved
by M1, the divisor of M the keys are joined modulo, so that the same pair meets modulo M1 but not
modulo M ("decoy"): the device must return the first and drop the second. The library finds M1 itself
(the largest divisor of M in [2^40, 2^62]); the test recomputes it the same way to build the decoy.
"""
from math import prod

import numpy as np
import pytest

import oracle as O
from cfjoin import instance as I
from cfjoin import pipeline

pytestmark = pytest.mark.gpu


def say(*a):
    print("CASE", *a, flush=True)


def key_divisor(M):
    """The largest divisor of M in [2^40, 2^62] (M factored by trial division below 2^20)."""
    f, c, q = [], M, 2
    while q < (1 << 20) and q * q <= c:
        if c % q == 0:
            e = 0
            while c % q == 0:
                c //= q
                e += 1
            f.append((q, e))
        q += 1 if q == 2 else 2
    if c > 1:
        f.append((c, 1))
    best = [0]

    def go(i, d):
        if i == len(f):
            best[0] = max(best[0], d)
            return
        for e in range(f[i][1] + 1):
            if d * f[i][0] ** e > 1 << 62:
                break
            go(i + 1, d * f[i][0] ** e)
    go(0, 1)
    assert best[0] >= 1 << 40
    return best[0]


def wide_instance(row, U, r, P0, n_ins=None):
    D = int(row["D"])
    base = I.family_instance(D, int(row["cap"]), U, r)
    INS = base.INS if n_ins is None else base.INS[:n_ins]
    return I.Instance(M=D, r=r, B0=base.B0, INS=INS, P0=P0, Icap=base.Icap, Dmin=base.Dmin, split=base.split,
                      hmax=base.hmax, tag=f"wide{D.bit_length()}_r{r}")


def planted(row, U, r, n_ins=None):
    """P0 making (the r largest base primes, the r smallest pool primes) meet modulo M."""
    base = wide_instance(row, U, r, 1, n_ins)
    D0, I0 = base.B0[-r:], base.INS[:r]
    return prod(D0) * pow(prod(I0), -1, base.M) % base.M


def device_pairs(inst, res):
    d_dec, i_dec = pipeline.decoders(inst)
    return {(tuple(sorted(d_dec(int(a)))), tuple(sorted(i_dec(int(b))))) for a, b in zip(res.pairs_d, res.pairs_i)}


def check(inst, res):
    nD, nI, pairs = O.exact_join(inst.M, inst.r, inst.P0, inst.Icap, inst.Dmin, inst.B0, inst.INS, inst.split_idx,
                                 inst.hmax)
    assert (res.nD, res.nI) == (nD, nI)
    got = device_pairs(inst, res)
    assert len(got) == len(res.pairs_d) and got == pairs
    return pairs


@pytest.mark.parametrize("k", [0, 2, 4, 6, 8])
def test_wide_modulus_planted_and_decoy(workloads, k):
    fam = workloads["families"]["nbhd_wide"]
    row, U = fam["rows"][k], int(fam["U"])
    M = int(row["D"])
    P0 = planted(row, U, 3)
    exact = wide_instance(row, U, 3, P0)
    decoy = wide_instance(row, U, 3, (P0 + key_divisor(M)) % M)
    say(f"wide M={M} ({M.bit_length()} bits) r=3 |INS|={len(exact.INS)} M1={key_divisor(M)}")
    need = (tuple(range(len(exact.B0) - 3, len(exact.B0))), (0, 1, 2))
    res = pipeline.solve(exact)
    assert need in check(exact, res)
    res = pipeline.solve(decoy)
    assert need not in check(decoy, res)
    # both in one batch with an ordinary instance: each result as alone
    both, _ = pipeline.solve_batch([exact, decoy, I.c1_instance(0, 100)])
    assert need in device_pairs(exact, both[0]) and need not in device_pairs(decoy, both[1])
    assert device_pairs(exact, both[0]) == device_pairs(exact, pipeline.solve(exact))


def test_wide_modulus_id_encodings_and_graph(workloads):
    """8-bit packed I ids (pool cut to 200 primes) and deletion-mask D ids (r = 11 on a 16-prime base
    and a 20-prime pool), and the exact check inside a captured graph, replayed twice."""
    fam = workloads["families"]["nbhd_wide"]
    row, U = fam["rows"][6], int(fam["U"])
    M = int(row["D"])
    packed = wide_instance(row, U, 3, planted(row, U, 3, n_ins=200), n_ins=200)
    assert not packed.rank_ids
    say(f"wide M={M} r=3 packed I ids")
    check(packed, pipeline.solve(packed))
    base = wide_instance(row, U, 3, 1)
    B, INS = base.B0[:16], base.INS[:20]
    P0 = prod(B[-11:]) * pow(prod(INS[:11]), -1, M) % M
    big = I.Instance(M=M, r=11, B0=B, INS=INS, P0=P0, Icap=prod(INS[-11:]) + 1, Dmin=1, split=None, hmax=11,
                     tag="wide_r11")
    say(f"wide M={M} r=11 mask D ids, colex I ids")
    pairs = check(big, pipeline.solve(big))
    assert (tuple(range(5, 16)), tuple(range(11))) in pairs
    b = pipeline.Batch([packed, big])
    try:
        eager, _ = b.solve()
        b.capture()
        for _ in range(2):
            rep, _ = b.solve()
            for x, y in zip(eager, rep):
                assert np.array_equal(x.pair_array(), y.pair_array()) and (x.nD, x.nI) == (y.nD, y.nI)
            pipeline.solve(I.c1_instance(1, 100))    # eager work between the replays
    finally:
        b.close()


def test_insertion_bound_beyond_128_bits_host_exact():
    """The logged M0 instance at r = 14 (insertion bound of 134 bits; logged D 13 | I 23, no pair)
    through the host-exact join, then with P0 planting a pair, against the exact Python join."""
    from cfjoin import hostexact
    inst = I.anchor_instance("M0", 14)
    assert inst.Icap.bit_length() == 134
    say("host-exact M0 r=14")
    res = hostexact.solve(inst)
    assert (res.nD, res.nI, len(res.pairs_d)) == (13, 23, 0)
    P0 = prod(inst.B0[-14:]) * pow(prod(inst.INS[:14]), -1, inst.M) % inst.M
    plant = I.Instance(M=inst.M, r=14, B0=inst.B0, INS=inst.INS, P0=P0, Icap=inst.Icap, Dmin=inst.Dmin,
                       split=inst.split, hmax=inst.hmax, tag="M0_r14_planted")
    pairs = check(plant, hostexact.solve(plant))
    assert (tuple(range(50, 64)), tuple(range(14))) in pairs
