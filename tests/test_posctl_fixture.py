"""CPU checks of the positive-control fixture (tests/golden/posctl_pairs.json) and its recipe.

The fixture's pair sets come from the CPU oracle. Here they are pinned by what the mathematics
fixes, independently of the oracle's join: every pair has exactly equal residues (checked in
Python integers), every pair respects the instance's product bounds, the pair whose completion is
the published incumbent is present, and that incumbent passes the certificate check. The
composition fixture (full D side, unpruned I side, from the oracle's roles-swapped join) is also
checked against the oracle's ordinary join on the bounded instance: the bounded pairs must be
exactly the composition's pairs that meet the bounds.
"""
import json
import os
from math import prod

import pytest

import oracle as O
from cfjoin import instance as I
from cfjoin import pipeline
from cfjoin import verify as V

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "posctl_pairs.json")


@pytest.fixture(scope="module")
def posctl():
    with open(GOLDEN) as f:
        return json.load(f)


def test_control_recipe_reaches_the_incumbents():
    radii = {}
    for name in I.CONTROLS:
        c = I.control_instance(name)
        ok, n, _ = V.verify_certificate(c.factors)
        assert ok and n == c.N and len(c.factors) == 64
        # the completion of (D_need, I_need) is the incumbent, and it lies within the bounds
        assert c.inst.P0 * prod(c.I_need) == c.N * prod(c.D_need)
        assert prod(c.I_need) < c.inst.Icap and prod(c.D_need) >= c.inst.Dmin
        assert set(c.I_need) <= set(c.inst.INS) and set(c.D_need) <= set(c.inst.B0)
        radii[name] = c.inst.r
    assert radii == {"N": 6, "N_prev": 7, "N0": 8, "N149": 9}


@pytest.mark.parametrize("name", list(I.CONTROLS))
def test_solve_fixture_pairs_are_exact(posctl, name):
    c = I.control_instance(name)
    inst = c.inst
    fx = posctl["solve"][name]
    assert (fx["M"], fx["r"], fx["nINS"]) == (inst.M, inst.r, len(inst.INS))
    d_dec, i_dec = pipeline.decoders(inst)
    best, need_seen = None, False
    for dv, iv in fx["pairs"]:
        d, i = d_dec(dv), i_dec(iv)
        D = [inst.B0[x] for x in d]
        Iq = [inst.INS[x] for x in i]
        assert len(set(d)) == len(set(i)) == inst.r
        assert (inst.P0 * prod(Iq) - prod(D)) % inst.M == 0
        assert prod(Iq) < inst.Icap and prod(D) >= inst.Dmin
        cmp = V.completion(inst.B0, inst.INS, d, i, inst.P0, inst.M)
        if cmp is not None:
            best = cmp[1] if best is None else min(best, cmp[1])
        need_seen |= sorted(D) == c.D_need and sorted(Iq) == c.I_need
    assert need_seen and best == int(fx["min_completion"]) <= c.N
    assert best == c.N or name == "N149"       # the 149-digit incumbent has a smaller completion


@pytest.mark.parametrize("name", ["N", "N_prev"])
def test_composition_fixture_against_bounded_join(posctl, name):
    c = I.control_instance(name, nins=30)
    inst = c.inst
    fx = posctl["compose"][name]
    assert fx["INS"] == inst.INS
    comp = {(tuple(D), tuple(Iq)) for D, Iq in fx["pairs"]}
    assert len(comp) == len(fx["pairs"])
    for D, Iq in comp:
        assert (inst.P0 * prod(Iq) - prod(D)) % inst.M == 0
    assert (tuple(c.D_need), tuple(c.I_need)) in comp
    res = O.cpu_join(inst.M, inst.r, inst.P0 % inst.M, inst.Icap, inst.Dmin, inst.B0, inst.INS, None, None, False,
                     threads=4)
    d_dec, i_dec = pipeline.decoders(inst)
    bounded = {(tuple(sorted(inst.B0[x] for x in d_dec(int(a)))), tuple(inst.INS[x] for x in i_dec(int(b))))
               for a, b in zip(res["pairs_d"], res["pairs_i"])}
    assert len(bounded) == len(res["pairs_d"])
    assert bounded == {(D, Iq) for D, Iq in comp if prod(Iq) < inst.Icap and prod(D) >= inst.Dmin}
