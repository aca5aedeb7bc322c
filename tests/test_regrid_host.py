"""Host grid generation of the regrid (adg_host_make_grids, the
AmrCore::MakeNewGrids steps the reference's Mesh uses) on the CPU: closed-
form layouts for a single tagged cell and a tagged band, and on seeded
random tags in 2D and 3D with three levels the invariants a layout must
hold -- every tagged cell with its n_error_buf buffer refined, boxes
disjoint, aligned to the blocking factor, inside the domain, no longer than
max_grid_size, each new level nested in the one below by n_proper cells, the
same result on every call.  (AMReX's own box layouts are not reproduced
cell for cell: AMReX is not available to compare with.)"""
import itertools

import numpy as np
import pytest

import pyadg


def _cells(box, D):
    lo, hi = box
    return itertools.product(*[range(lo[d], hi[d] + 1) for d in range(D)])


def _covered(levels_l, D):
    s = set()
    for b in levels_l:
        s.update(_cells(b, D))
    return s


def test_single_tag_closed_form():
    # tag (5,7) on 16^2, buffer 1 -> 3x3 cells [4..6]x[6..8] -> coarsened by
    # 4/2 = 2 -> [2..3]x[3..4] -> refined x2 then x2: [8..15]x[12..19]
    out = pyadg.make_grids(2, [16, 16], [[((0, 0), (15, 15))]], [[(5, 7)]], 1)
    assert out == [[((0, 0), (15, 15))], [((8, 12), (15, 19))]]


def test_band_closed_form():
    tags = [(x, y) for x in range(16) for y in range(4, 8)]
    out = pyadg.make_grids(2, [16, 16], [[((0, 0), (15, 15))]], [tags], 1)
    assert out == [[((0, 0), (15, 15))], [((0, 4), (31, 19))]]


def test_no_tags_no_fine_level():
    out = pyadg.make_grids(2, [16, 16], [[((0, 0), (15, 15))]], [[]], 2)
    assert out == [[((0, 0), (15, 15))]]


def _check_layout(D, dom0, out, tags, per, bf=4, mgs=None, nbuf=1, nprop=1):
    mgs = mgs or (128 if D == 2 else 32)
    for l in range(1, len(out)):
        dom = [n << l for n in dom0]
        boxes = out[l]
        cells = [set(_cells(b, D)) for b in boxes]
        # disjoint
        assert sum(len(c) for c in cells) == len(set().union(*cells)), f"level {l} overlaps"
        for lo, hi in boxes:
            for d in range(D):
                assert lo[d] % bf == 0 and (hi[d] + 1) % bf == 0, "blocking factor"
                assert 0 <= lo[d] <= hi[d] < dom[d], "inside the domain"
                assert hi[d] - lo[d] + 1 <= mgs, "max_grid_size"
    # buffered tags of level l are refined at level l+1
    for l, tl in enumerate(tags):
        if l + 1 >= len(out):
            assert len(tl) == 0
            continue
        fine = _covered(out[l + 1], D)
        dom = [n << l for n in dom0]
        for t in tl:
            for o in itertools.product(range(-nbuf, nbuf + 1), repeat=D):
                c = [t[d] + o[d] for d in range(D)]
                if any((c[d] < 0 or c[d] >= dom[d]) and not per[d] for d in range(D)):
                    continue
                c = [c[d] % dom[d] for d in range(D)]
                for ch in itertools.product(range(2), repeat=D):
                    assert tuple(2 * c[d] + ch[d] for d in range(D)) in fine, (l, t)
    # nesting: level l+1 coarsened and grown by n_proper lies in level l
    for l in range(1, len(out) - 1):
        crse = _covered(out[l], D)
        dom = [n << l for n in dom0]
        for b in out[l + 1]:
            for c in _cells(b, D):
                cc = [c[d] // 2 for d in range(D)]
                for o in itertools.product(range(-nprop, nprop + 1), repeat=D):
                    g = [cc[d] + o[d] for d in range(D)]
                    if any((g[d] < 0 or g[d] >= dom[d]) and not per[d] for d in range(D)):
                        continue
                    g = tuple(g[d] % dom[d] for d in range(D))
                    assert g in crse, f"level {l + 1} not nested"


@pytest.mark.parametrize("D,seed,periodic", [(2, 1, True), (2, 2, False), (3, 3, True), (3, 4, False)])
def test_random_tags_three_levels(D, seed, periodic):
    rng = np.random.default_rng(seed)
    n0 = 16
    dom0 = [n0] * D
    per = [1 if periodic else 0] * D
    L0 = [(tuple([0] * D), tuple([n0 - 1] * D))]
    # an old level 1 in the middle, tags in two blobs on level 0 and inside level 1
    L1 = [(tuple([8] * D), tuple([23] * D))]
    t0 = set()
    for c in (np.array([3] * D), np.array([11] * D)):
        for _ in range(12):
            t0.add(tuple(int(v) for v in np.clip(c + rng.integers(-2, 3, D), 0, n0 - 1)))
    t1 = set()
    for _ in range(20):
        t1.add(tuple(int(v) for v in rng.integers(10, 22, D)))
    t0, t1 = sorted(t0), sorted(t1)
    out = pyadg.make_grids(D, dom0, [L0, L1], [t0, t1], 2, periodic=per)
    assert len(out) == 3
    _check_layout(D, dom0, out, [t0, t1], per)
    again = pyadg.make_grids(D, dom0, [L0, L1], [t0, t1], 2, periodic=per)
    assert again == out


def test_max_grid_size_split():
    """A fully tagged 64^2 level 0 gives a 128^2 level 1 split into boxes of
    at most 32 (max_grid_size 32)."""
    tags = [(x, y) for x in range(64) for y in range(64)]
    out = pyadg.make_grids(2, [64, 64], [[((0, 0), (63, 63))]], [tags], 1, max_grid_size=32)
    assert len(out[1]) == 16
    for lo, hi in out[1]:
        assert hi[0] - lo[0] + 1 == 32 and hi[1] - lo[1] + 1 == 32
