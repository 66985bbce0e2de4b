"""The leaf-size tuner's walk over its ladder (pibem_amd.leaf_walk), on
cost functions with a known minimum: it finds an interior minimum and
brackets it, stops at the ladder's end for monotone costs (not bracketed),
searches both sides of the start, and measures every leaf size at most
once."""
from collections import Counter

from pibem_amd import LEAF_LADDER, leaf_walk


def _walk(cost, start=20):
    seen = Counter()

    def measure(leaf):
        seen[leaf] += 1
        return cost(leaf)

    best, bracketed = leaf_walk(measure, start)
    return best, bracketed, seen


def test_interior_minimum_bracketed():
    for target in (12, 20, 48, 64, 128):
        best, bracketed, seen = _walk(lambda l: abs(l - target))
        assert best == target and bracketed
        assert max(seen.values()) == 1


def test_monotone_costs_end_of_ladder():
    best, bracketed, _ = _walk(lambda l: l)
    assert best == LEAF_LADDER[0] and not bracketed
    best, bracketed, _ = _walk(lambda l: -l)
    assert best == LEAF_LADDER[-1] and not bracketed


def test_walk_stops_after_three_misses():
    # two valleys: the walk from 20 takes the nearer one and stops at the
    # ladder's end (three misses would be needed to stop earlier)
    cost = {8: 5, 12: 4, 16: 3, 20: 4, 24: 5, 32: 6, 40: 7, 48: 1, 64: 2}
    best, bracketed, seen = _walk(lambda l: cost.get(l, 9))
    assert best == 16 and bracketed and 48 not in seen and 8 in seen


def test_walk_crosses_a_flat_start():
    # the neighbours of 20 are within the 1% gain, the real minimum is at 48
    cost = {16: 14.98, 20: 14.04, 24: 13.99, 32: 14.25, 40: 13.06, 48: 12.91, 64: 13.06, 80: 13.84, 96: 14.5}
    best, bracketed, seen = _walk(lambda l: cost.get(l, 15.0))
    assert best == 48 and bracketed


def test_walk_leaves_a_start_cheaper_than_both_neighbours():
    # 20 is a shade cheaper than 24 and 16 (noise), the minimum is at 96: the
    # walk still goes towards the cheaper neighbour and finds it
    cost = {16: 14.38, 20: 13.35, 24: 13.38, 32: 12.93, 40: 11.58, 48: 11.01, 64: 10.29, 80: 10.85, 96: 9.56,
            128: 10.55, 160: 13.99}
    best, bracketed, seen = _walk(lambda l: cost.get(l, 15.0))
    assert best == 96 and bracketed


def test_start_at_the_minimum_stays():
    # a true minimum at the start: three misses on each side end the walk
    cost = {8: 7, 12: 6, 16: 5, 20: 4, 24: 5.5, 32: 6.5}
    best, bracketed, seen = _walk(lambda l: cost.get(l, 9))
    assert best == 20 and bracketed and 48 not in seen


def test_noisy_neighbour_does_not_hide_its_side():
    # 24 timed high once (16.48 ms, about 13.6 otherwise), so the walk goes
    # towards 16 first; the walk from 20 upwards still finds 96
    cost = {8: 28.21, 12: 17.36, 16: 14.62, 20: 13.55, 24: 16.48, 32: 12.58, 40: 11.27, 48: 10.94, 64: 10.19,
            80: 10.46, 96: 9.61, 128: 10.51, 160: 13.90}
    best, bracketed, seen = _walk(lambda l: cost[l])
    assert best == 96 and bracketed and max(seen.values()) == 1


def test_walk_crosses_a_plateau():
    # equal costs at 32 and 40 do not end the walk: 64 is found
    cost = {16: 9, 20: 8, 24: 7, 32: 6, 40: 6, 48: 5, 64: 4, 80: 4.5, 96: 5}
    best, bracketed, seen = _walk(lambda l: cost.get(l, 10))
    assert best == 64 and bracketed and 160 not in seen


def test_noise_does_not_move_the_choice():
    # 80 is 0.5% cheaper than 64: within the 1% gain the walk asks, 64 stays
    cost = {16: 96.6, 20: 86.1, 24: 81.6, 32: 76.9, 40: 67.0, 48: 62.3, 64: 55.12, 80: 54.83, 96: 56.3, 128: 64.6}
    best, bracketed, seen = _walk(lambda l: cost.get(l, 99.0))
    assert best == 64 and bracketed and 96 in seen
