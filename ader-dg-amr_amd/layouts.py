"""Patch layouts (level index space, refinement ratio 2) for the host AMR
substrate: uniform tilings and the Kelvin-Helmholtz band hierarchy."""


def tiles(n, mg, D=2, lo0=None, hi0=None):
    """Boxes of edge mg covering [lo0, hi0] (default the whole n^D domain), x fastest."""
    import itertools
    lo0 = lo0 or [0] * D
    hi0 = hi0 or [n - 1] * D
    axes = [list(range(lo0[d], hi0[d] + 1, mg)) for d in range(D)]
    out = []
    for rev in itertools.product(*reversed(axes)):
        lo = tuple(reversed(rev))
        hi = tuple(min(lo[d] + mg - 1, hi0[d]) for d in range(D))
        out.append((lo, hi))
    return out


def single_level(n, mg=None, D=2):
    return [tiles(n, mg or n, D)]


def two_level_centre(n, mg=None, D=2):
    """Level 0 n^D; level 1 covers the central half (ratio-aligned)."""
    l0 = tiles(n, mg or n, D)
    lo = [n // 2] * D
    hi = [n + n // 2 - 1] * D
    l1 = tiles(2 * n, mg or n, D, lo, hi)
    return [l0, l1]


def kh_bands(n0, nlev=3, mg=32, D=2):
    """Kelvin-Helmholtz layout: level l refines full-width bands around the
    two shear layers y=0.25, 0.75, narrowing with level."""
    levels = [tiles(n0, mg, D)]
    for l in range(1, nlev):
        n = n0 << l
        w = max(4, n // (4 << l))  # half width in level cells, multiple of 2
        w += w % 2
        boxes = []
        for yc in (n // 4, 3 * n // 4):
            lo = [0, yc - w]
            hi = [n - 1, yc + w - 1]
            boxes += tiles(n, mg, D, lo, hi)
        levels.append(boxes)
    return levels


def batch_patches(npatch, nmin, nmax, seed, row_width=4096):
    """Square patches of edge n ~ U{nmin..nmax} (seeded) shelf-packed into
    rows of a bounding index space, for a batch of independent periodic
    problems.  Returns (levels, dom, edges)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    edges = rng.integers(nmin, nmax + 1, size=npatch)
    boxes, x, y, row_h = [], 0, 0, 0
    for n in edges:
        n = int(n)
        if x + n > row_width:
            x, y, row_h = 0, y + row_h, 0
        boxes.append(((x, y), (x + n - 1, y + n - 1)))
        x += n
        row_h = max(row_h, n)
    dom = [row_width, y + row_h]
    return [boxes], dom, [int(n) for n in edges]
