"""Greedy candidate selection for the learned planner (CPU)."""
import numpy as np

from gputimeseriescompressionoptimizer_amd.select_candidates import greedy


def test_greedy_picks_best_single_tree_first_then_complements():
    sizes = np.array([[10, 5, 7], [10, 9, 3], [10, 9, 8.0]])
    # tree 2 has the smallest geometric total alone; tree 1 then helps block 0
    assert greedy(sizes, 2) == [2, 1]


def test_greedy_matches_brute_force_on_small_pools():
    r = np.random.default_rng(0)
    for _ in range(20):
        sizes = np.exp2(r.uniform(4, 12, (30, 6)))
        pick = greedy(sizes, 1)
        brute = int(np.argmin(np.log2(sizes).sum(0)))
        assert pick == [brute]
        assert len(set(greedy(sizes, 6))) == 6
