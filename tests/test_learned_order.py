"""Launch orders of the cost model (matrixgraph_amd.learned.order_by_cost), on the CPU."""
import numpy as np

from matrixgraph_amd.learned import order_by_cost


def test_ungrouped_order_is_longest_first_and_stable_within_buckets():
    cost = np.array([1.0, 100.0, 3.0, 100.0, 1.0, 1000.0])
    o = order_by_cost(cost)
    assert list(o) == [5, 1, 3, 2, 0, 4]
    assert sorted(o) == list(range(len(cost)))


def test_grouped_order_keeps_groups_adjacent_and_ranks_them_by_summed_cost():
    # Three groups (the same tile range of one graph in several panels), interleaved.
    groups = np.array([7, 9, 7, 4, 9, 7, 4])
    cost = np.array([10.0, 1.0, 10.0, 300.0, 1.0, 10.0, 1.0])  # sums: 7 -> 30, 9 -> 2, 4 -> 301
    o = order_by_cost(cost, groups)
    assert sorted(o) == list(range(len(cost)))
    seq = groups[o]
    # Each group is one contiguous run, heaviest group first, members in their current order.
    assert list(seq) == [4, 4, 7, 7, 7, 9, 9]
    assert list(o) == [3, 6, 0, 2, 5, 1, 4]


def test_grouped_order_ties_keep_first_appearance():
    groups = np.array([2, 1, 2, 1])
    cost = np.array([5.0, 5.0, 5.0, 5.0])
    assert list(order_by_cost(cost, groups)) == [0, 2, 1, 3]
