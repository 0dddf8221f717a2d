"""Host side of the product's Barnes-Hut path (no GPU): the flattened octree
built behind cp_bh_build has the oracle's structure (node count, depth, root
box) on the reference's known-answer grid and on random clouds."""
import numpy as np
import pytest

from cpanel_amd.case import Context
from oracle import vp_oracle as V


@pytest.fixture(scope="module")
def hctx():
    return Context(-1)


def grid27():
    return np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], float)


@pytest.mark.parametrize("case", ["grid", "r700", "r5000", "r10", "flat"])
def test_tree_structure_matches_oracle(hctx, case):
    rng = np.random.default_rng(len(case))
    if case == "grid":
        x = grid27()
    elif case == "flat":  # a sheet (zero extent in z), like a fresh wake
        x = rng.normal(size=(3000, 3))
        x[:, 2] = 0.0
    else:
        x = rng.normal(size=(int(case[1:]), 3))
    a = rng.normal(size=x.shape)
    r = np.full(len(x), 0.05)
    for mm in (10, 4):
        t = hctx.bh_tree(x, a, r, max_members=mm)
        o = V.BHTree(x, a, r, max_members=mm)
        ti, oi = t.info(), o.info()
        assert ti["n_nodes"] == oi["n_nodes"] and ti["levels"] == oi["levels"]
        assert np.array_equal(ti["origin"], oi["origin"]) and np.array_equal(ti["half"], oi["half"])


def test_host_only_context_refuses_eval(hctx):
    x = grid27()
    t = hctx.bh_tree(x, x, np.full(27, 0.1))
    with pytest.raises(Exception):
        t.eval(1)
