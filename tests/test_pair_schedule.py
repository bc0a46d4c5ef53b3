"""The pair schedule of the pair-once kernel's diagonal workgroups
(lennard-jones-cuda_amd/csrc/lj_force_n3.hip: pair_tile, tri_pair, tri_tiles),
enumerated on the CPU: every unordered pair of a group of 256 is met exactly
once, no particle meets itself, and the rotating j sums end on the lane that
owns their slot.  The same steps, lane by lane, as the kernel's loops."""
import itertools

SUB = 64  # particles per sub-tile; sub-tile k of a group = particles 64 k + [0, 64)


def dual(meet, p1, p2, i1, i2, steps, half_last):
    """Two streams: at step t lane l's particle of sub-tile i1 meets slot
    (l + t) mod 64 of sub-tile p1, that of i2 slot (l + t) mod 64 of p2; at the
    last step of a triangle (t = 32) only lanes 0-31 count."""
    for t in steps:
        for lane in range(SUB):
            if half_last and t == 32 and lane >= 32:
                continue
            for isub, jsub in ((i1, p1), (i2, p2)):
                meet(SUB * isub + lane, SUB * jsub + (lane + t) % SUB)


def group_schedule():
    met = []
    meet = lambda i, j: met.append((i, j))
    # pair-tile wave: i = sub-tiles 0 and 1 (two per lane), j = sub-tiles 2, 3
    for st in (2, 3):
        dual(meet, st, st, 0, 1, range(64), False)
    # triangle wave: 1. triangles of 1 and 3, 2. 0 x 1 and 2 x 3, 3. triangles of 0 and 2
    dual(meet, 1, 3, 1, 3, range(1, 33), True)
    dual(meet, 1, 3, 0, 2, range(64), False)
    dual(meet, 0, 2, 0, 2, range(1, 33), True)
    return met


def test_every_pair_of_a_group_once():
    met = group_schedule()
    assert all(i != j for i, j in met)
    seen = {}
    for i, j in met:
        key = (min(i, j), max(i, j))
        seen[key] = seen.get(key, 0) + 1
    every = set(itertools.combinations(range(4 * SUB), 2))
    assert set(seen) == every and all(c == 1 for c in seen.values())
    assert len(met) == 256 * 255 // 2


def rotate(slots, k):
    """Lane l receives the register of lane (l + k) mod 64 (ds_bpermute)."""
    return [slots[(lane + k) % SUB] for lane in range(SUB)]


def test_rotating_sums_end_on_their_lane():
    # a 64-step pass (pair tile, cross phase): lane l starts with slot l
    slots = list(range(SUB))
    for t in range(64):
        assert slots == [(lane + t) % SUB for lane in range(SUB)]
        slots = rotate(slots, 1)
    assert slots == list(range(SUB))
    # a triangle (t = 1..32, the sums start at zero on slot l + 1), then the
    # alignment by 31 lanes
    slots = [(lane + 1) % SUB for lane in range(SUB)]
    for t in range(1, 33):
        assert slots == [(lane + t) % SUB for lane in range(SUB)]
        slots = rotate(slots, 1)
    assert rotate(slots, 31) == list(range(SUB))


def step_items(ng):
    """The time step's workgroups (n3_step_items): diagonal pairs (g, g + half)
    first, a last odd group alone, then the off-diagonal group pairs."""
    half = (ng + 1) // 2
    items = [("diag", g, g + half if g + half < ng else None) for g in range(half)]
    items += [("pair", g, h) for g in range(ng) for h in range(g + 1, ng)]
    return items


def test_step_items_cover_every_group_pair_once():
    for ng in (1, 2, 3, 40, 51, 63, 64, 80):
        items = step_items(ng)
        assert len(items) == ng * (ng - 1) // 2 + (ng + 1) // 2
        diag = [g for kind, a, b in items if kind == "diag" for g in (a, b) if g is not None]
        assert sorted(diag) == list(range(ng))
        pairs = [(a, b) for kind, a, b in items if kind == "pair"]
        assert sorted(pairs) == list(itertools.combinations(range(ng), 2))
    assert len(step_items(64)) == 2048  # N = 16384: exactly the resident workgroups
