"""Parity bars shared by the GPU tests (imported by them, not collected).

mode_err is the per-(variable, mode) relative error of a modal array
[Q][modes][cells]: for each variable q and mode n, max over cells of
|got - ref| divided by max over cells of |ref[q, n]|.  A mode that is far
below its variable's size carries only the round-off of the larger terms
it is summed from, so its denominator is floored at 1 % of the scale of
the variable's group: density, momentum (every component together, so a
component that is zero by symmetry is measured on the flow's momentum) or
energy; the group scales are themselves floored at 1e-6 of the largest
(a state at rest has no momentum scale of its own).  Below that floor the
bar is therefore 1e-14 of the group scale."""
import numpy as np

FLOOR = 1e-2


def group_scales(ref):
    Q = ref.shape[0]
    D = Q - 2
    s0 = np.abs(ref[:, 0]).reshape(Q, -1).max(axis=1)
    top = float(np.abs(ref).max())
    rho, mom, en = s0[0], s0[1:1 + D].max(), s0[D + 1]
    return [max(v, 1e-6 * top, 1e-300) for v in [rho] + [mom] * D + [en]]


def mode_err(got, ref):
    """max over (q, n) of the per-mode relative error, and the worst (q, n)."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    sc = group_scales(ref)
    worst, at = 0.0, None
    for q in range(ref.shape[0]):
        for n in range(ref.shape[1]):
            den = max(np.abs(ref[q, n]).max(), FLOOR * sc[q])
            e = np.abs(got[q, n] - ref[q, n]).max() / den
            if e > worst:
                worst, at = e, (q, n)
    return worst, at


def uncovered(levels):
    """Per level: slot indices (x fastest inside each grown box, one ghost
    layer) of the valid cells no finer box covers, and the cell volume
    factor 2^-D l relative to level 0."""
    D = len(levels[0][0][0])
    out = []
    for l, lev in enumerate(levels):
        fine = levels[l + 1] if l + 1 < len(levels) else []
        slots, off = [], 0
        for lo, hi in lev:
            g = [hi[d] - lo[d] + 3 for d in range(D)]
            n = [hi[d] - lo[d] + 1 for d in range(D)]
            for idx in np.ndindex(*reversed(n)):
                iv = [lo[d] + idx[D - 1 - d] for d in range(D)]
                if any(all(flo[d] <= 2 * iv[d] <= fhi[d] for d in range(D)) for flo, fhi in fine):
                    continue
                s, st = off, 1
                for d in range(D):
                    s += (iv[d] - lo[d] + 1) * st
                    st *= g[d]
                slots.append(s)
            off += int(np.prod(g))
        out.append((np.array(slots, dtype=np.int64), 2.0 ** (-D * l)))
    return out


def density_l1_diff(a_levels, b_levels, levels):
    """Composite L1 norm (cell averages, uncovered cells) of the density
    difference a - b relative to that of b."""
    num = den = 0.0
    for l, (sl, w) in enumerate(uncovered(levels)):
        num += w * np.abs(a_levels[l][0, 0, sl] - b_levels[l][0, 0, sl]).sum()
        den += w * np.abs(b_levels[l][0, 0, sl]).sum()
    return num / den
