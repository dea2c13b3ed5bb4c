"""Pins of the interaction lists (AFMM2DTree.cpp:24-967).

topology_L4.json holds the lists the reference's own assignment statements
produce (tests/golden/make_topology_fixture.py replays them); the oracle must
reproduce them exactly.  Closed-form counts and symmetry cover deeper trees.
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "topology_L4.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_oracle_topology_matches_reference_tables(oracle_mod, gold, level):
    nbr, inner, outer = oracle_mod.topology(level)
    np.testing.assert_array_equal(nbr, np.array(gold["neighbor"][level]))
    np.testing.assert_array_equal(inner, np.array(gold["inner"][level]))
    np.testing.assert_array_equal(outer, np.array(gold["outer"][level]))


def test_pair_counts_and_symmetry(oracle_mod):
    expect = {2: 156, 3: 1116, 4: 5628, 5: 25020, 6: 105276, 7: 431676}
    for level, cnt in expect.items():
        nbr, inner, outer = oracle_mod.topology(level)
        il = np.concatenate([inner, outer], axis=1)
        assert int((il >= 0).sum()) == cnt
        assert int((il >= 0).sum(axis=1).max()) == (27 if level >= 3 else 12)
        pairs = {(k, b) for k in range(il.shape[0]) for b in il[k] if b >= 0}
        assert all((b, k) in pairs for k, b in pairs)
        npairs = {(k, b) for k in range(nbr.shape[0]) for b in nbr[k] if b >= 0}
        assert all((b, k) in npairs for k, b in npairs)
        if level == 7:
            assert len(npairs) + nbr.shape[0] == 145924
    nbr3, _, _ = oracle_mod.topology(3)
    assert len({(k, b) for k in range(64) for b in nbr3[k] if b >= 0}) + 64 == 484
