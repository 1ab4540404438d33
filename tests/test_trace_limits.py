"""The traversal stack's bound and the renderer's depth check (vh_trace.h stackNeed / kMaxTreeDepth,
enforced by vh_renderer_create and vh_renderer_update), on the CPU.

The traversal (traceStep / waveStep) pushes the inner children a node's boxes hit (at most 4) and
pops the next one; a pop may skip entries, never add any. So a tree of depth D (levels of 4-wide
nodes, the root on level 1) needs at most 3 (D - 1) + 4 entries, and the renderer refuses trees
deeper than the depth whose bound fills the stack. Checked here: the library's limits agree with
that bound, and a simulation of the push/pop policy over random trees (every box hit: culling only
shrinks the stack) never holds more than the bound at any node.
"""
import ctypes as C
import os
import random

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "verion_amd", "libverion_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libverion_hip.so not built")


def limits():
    from verion_amd.backend import library

    out = np.zeros(4, np.uint32)
    assert library().vh_trace_limits(out.ctypes.data_as(C.POINTER(C.c_uint32))) == 0
    return {"stack": int(out[0]), "short": int(out[1]), "max_depth": int(out[2]), "leaf_max": int(out[3])}


def need(depth):
    return 3 * (depth + 1) + 4  # the renderer's bound (a margin of 6 over 3 (depth - 1) + 4)


def test_the_accepted_depth_is_the_deepest_whose_bound_fits_the_stack():
    L = limits()
    assert L["stack"] == 96 and L["short"] == 12 and L["leaf_max"] == 4
    assert 0 < L["short"] < L["stack"]
    assert need(L["max_depth"]) <= L["stack"] < need(L["max_depth"] + 1)
    # the procedural workload's LBVH (depth 16) and the many-lights one (15) are well inside
    assert L["max_depth"] >= 16


def simulate(depth, rng, budget=200_000, full=False):
    """DFS over a random 4-wide tree of `depth` levels built on the fly: each inner node has 1-4
    children, inner ones below the last level; every box is hit; the pushes of a node come in a
    random order (the sort by distance). Returns the largest stack seen at each level."""
    worst = {}
    stack = [1]  # levels of pending nodes; the root is level 1
    visits = 0
    while stack and visits < budget:
        level = stack.pop()
        visits += 1
        here = len(stack)  # entries pending when this node is visited
        kids = 4 if full else rng.randint(1, 4)
        inner = 0
        if level < depth:
            inner = kids if full else rng.randint(0, kids)
        stack.extend([level + 1] * inner)
        worst[level] = max(worst.get(level, 0), here + inner)
    return worst


@pytest.mark.parametrize("depth", [1, 2, 3, 5, 8])
def test_simulated_traversal_stays_within_the_bound_on_whole_trees(depth):
    rng = random.Random(depth)
    for trial in range(20):
        worst = simulate(depth, rng, full=trial == 0)
        for level, sp in worst.items():
            assert sp <= 3 * (level - 1) + 4, (level, sp)
        assert max(worst.values()) <= 3 * (depth - 1) + 4 <= need(depth)
    # a full tree (every node 4 inner children) reaches the bound of its deepest pushing level
    full = simulate(depth, rng, full=True)
    if depth > 1:
        assert max(full.values()) == 3 * (depth - 2) + 4


def test_simulated_traversal_of_the_deepest_accepted_tree_fits_the_stack():
    L = limits()
    rng = random.Random(7)
    for trial in range(4):
        worst = simulate(L["max_depth"], rng, budget=50_000, full=trial == 0)
        assert max(worst.values()) <= L["stack"]
        for level, sp in worst.items():
            assert sp <= 3 * (level - 1) + 4
