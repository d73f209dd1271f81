"""The schedule of the GPU's incidence-list colouring rounds (k_jpi_*) equals the Jones-Plassmann
contend / commit rounds of src/solver_velocity.hpp:305-355 (oracle/pdref.c jpColor): same colours,
same round count, same uncoloured total entering each round.

Both schedules are restated here in plain Python from their definitions and run on random edge
sets (parallel edges, edges of one body, an edge whose two ends are the same body, bodies with
more than 62 edges, and round caps that bind), and the incidence schedule is run on the oracle's
own contacts and manifolds, against the oracle's colours and round statistics.
"""
import random

import numpy as np
import pytest

from oracle_binding import oracle_sim


def priority(idx):
    z = (idx + 0x9E3779B9) & 0xFFFFFFFF
    z = ((z ^ (z >> 16)) * 0x21F0AAAD) & 0xFFFFFFFF
    z = ((z ^ (z >> 15)) * 0x735A2D97) & 0xFFFFFFFF
    z ^= z >> 15
    return ((z & 0x7FFFFFFF) << 32) | idx


def first_fit(forbidden):
    c = 0
    while c < 62 and forbidden & (1 << c):
        c += 1
    return c


def contend_commit(ends, num_bodies, colour):
    """Rounds as the reference runs them: each uncoloured edge holding the highest priority at all
    its bodies takes the first colour free at them."""
    colour = list(colour)
    mask = [0] * num_bodies
    remaining, prev, rounds, uncol = 1, -1, 0, 0
    while rounds < num_bodies + 2 and remaining > 0:
        winner = [-1] * num_bodies
        live = [e for e in range(len(ends)) if colour[e] == -1]
        uncol += len(live)
        for e in live:
            for x in ends[e]:
                if x >= 0:
                    winner[x] = max(winner[x], priority(e))
        rem = 0
        for e in live:
            a, b = ends[e]
            if winner[a] != priority(e) or (b >= 0 and winner[b] != priority(e)):
                rem += 1
                continue
            c = first_fit(mask[a] | (mask[b] if b >= 0 else 0))
            colour[e] = c
            mask[a] |= 1 << c
            if b >= 0:
                mask[b] |= 1 << c
        rounds += 1
        if rem == prev:
            break
        prev, remaining = rem, rem
    return colour, rounds, uncol


def incidence_rounds(ends, num_bodies, colour):
    """The GPU's schedule: per-body lists sorted by descending priority, a cursor per body, an edge
    coloured in the round after it became the cursor edge of the last of its distinct bodies."""
    colour = list(colour)
    entering = [e for e in range(len(ends)) if colour[e] == -1]

    def bodies(e):
        a, b = ends[e]
        return [a] if b < 0 or b == a else [a, b]

    lists = [[] for _ in range(num_bodies)]
    for e in entering:
        for x in bodies(e):
            lists[x].append(e)
    for lst in lists:
        lst.sort(key=priority, reverse=True)
    cursor, mask, arrived = [0] * num_bodies, [0] * num_bodies, {}

    def arrive(e, out):
        arrived[e] = arrived.get(e, 0) + 1
        if arrived[e] == len(bodies(e)):
            out.append(e)

    frontier = []
    for lst in lists:
        if lst:
            arrive(lst[0], frontier)
    rounds, uncol, remaining, size, prev = 0, 0, 1, len(entering), -1
    while rounds < num_bodies + 2 and remaining > 0:
        uncol += size
        rounds += 1
        nxt = []
        for e in frontier:
            bs = bodies(e)
            forbidden = 0
            for x in bs:
                forbidden |= mask[x]
            c = first_fit(forbidden)
            colour[e] = c
            for x in bs:
                mask[x] |= 1 << c
                cursor[x] += 1
                if cursor[x] < len(lists[x]):
                    arrive(lists[x][cursor[x]], nxt)
        rem = size - len(frontier)
        assert not (rem == prev and rem > 0), "a round coloured no edge"
        prev, remaining, size, frontier = rem, rem, rem, nxt
    return colour, rounds, uncol


@pytest.mark.parametrize("seed", range(6))
def test_schedules_agree_on_random_edge_sets(seed):
    rng = random.Random(seed)
    for _ in range(60):
        nb = rng.randint(1, 24)
        ends = []
        for _ in range(rng.randint(0, 160)):
            a = rng.randrange(nb)
            u = rng.random()
            ends.append((a, -1 if u < 0.15 else (a if u < 0.2 else rng.randrange(nb))))
        start = [-1 if rng.random() < 0.85 else -2 for _ in ends]
        assert incidence_rounds(ends, nb, start) == contend_commit(ends, nb, start)


def test_schedules_agree_past_62_colours_and_at_the_round_cap():
    # one body with 100 edges to 100 others: colours 0..61 then 62 repeated, 100 rounds
    star = [(0, k) for k in range(1, 101)]
    a = incidence_rounds(star, 101, [-1] * 100)
    assert a == contend_commit(star, 101, [-1] * 100)
    assert a[1] == 100 and sorted(a[0])[-1] == 62 and a[0].count(62) == 38
    # two bodies joined by 30 parallel edges: the cap (4 rounds) leaves 26 uncoloured
    pair = [(0, 1)] * 30
    b = incidence_rounds(pair, 2, [-1] * 30)
    assert b == contend_commit(pair, 2, [-1] * 30)
    assert b[1] == 4 and b[0].count(-1) == 26


@pytest.fixture(scope="module")
def dense_bed():
    from scenes import apply_scene, make_scene

    sc = make_scene("bed_periodic", 300, 9)
    s = apply_scene(oracle_sim(300), sc)
    for _ in range(3):
        s.step(sc["dt"])
    return s


def test_incidence_schedule_on_the_oracle_contacts(dense_bed):
    ids, _ = dense_bed.debug_contacts()
    ends = [tuple(x) for x in ids.tolist()]
    col, rounds, uncol = incidence_rounds(ends, dense_bed.get_stats()["num_particles"],
                                          [-1] * len(ends))
    st = dense_bed.get_stats()
    assert np.array_equal(np.array(col), dense_bed.debug_contact_colors())
    assert (rounds, uncol) == (st["rounds_pos"], st["uncolored_pos"])
    assert rounds > 3


def test_incidence_schedule_on_the_oracle_manifolds(dense_bed):
    ids, _ = dense_bed.debug_manifolds()
    ocol = dense_bed.debug_manifold_colors()
    ends = [(ra, rb) for _, _, _, ra, rb in ids.tolist()]
    start = [-1 if c != -2 else -2 for c in ocol.tolist()]
    col, rounds, uncol = incidence_rounds(ends, dense_bed.num_particles(), start)
    st = dense_bed.get_stats()
    assert np.array_equal(np.array(col), ocol)
    assert (rounds, uncol) == (st["rounds_vel"], st["uncolored_vel"])
