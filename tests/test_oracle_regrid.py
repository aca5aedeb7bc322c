"""Pins of the oracle's regrid pieces on the CPU.
- remake (RemakeLevel / MakeNewLevelFromCoarse through FillPatchTwoLevels
  with L2ProjInterp, AmrDG_Level.tpp:52-224): cells the old fine level held
  are copied bit for bit; a new fine cell is the L2 projection of its
  parent's polynomial onto the child, checked against numpy quadrature of
  the parent polynomial (child index, P_cf, the mass matrix); the children's
  mass equals the parent's.
- tags (AMR_tag_cell_refinement, AmrDG_TagCellRef.tpp:4-52): the vortex
  geometry counted by numpy, the density-gradient criterion on a step, the
  TVB indicator on a ramp (the threshold tvb_c * amr_c * h^2 on either side)."""
import itertools

import numpy as np
import pytest
from numpy.polynomial import legendre as npl

from layouts import single_level, tiles

G = 1.4


def _mi(T, D):
    return T.get("mi_s").reshape(-1, D).astype(int)


def _leg(k, x):
    e = np.zeros(k + 1)
    e[-1] = 1.0
    return npl.legval(x, e)


@pytest.mark.parametrize("D,P", [(2, 2), (2, 3), (3, 2)])
def test_remake_new_cells_are_projections(orc, D, P):
    n = 8
    T = orc.Tables(P, D)
    Lold = single_level(n, 4, D)
    Ho = orc.Hier(D, Lold, [n] * D, [0] * D, [1] * D)
    rng = np.random.default_rng(10 + 3 * D + P)
    Uo = orc.alloc_levels(T, Ho)
    Uo[0][:] = rng.standard_normal(Uo[0].shape)
    Uo[0][0, 0] += 3.0
    new = Lold + [tiles(2 * n, 4, D, [4] * D, [11] * D)]
    Hn = orc.Hier(D, new, [n] * D, [0] * D, [1] * D)
    Un = orc.remake(T, Ho, Uo, Hn)
    assert np.array_equal(Un[0], Uo[0])
    mi = _mi(T, D)
    xg, wg = npl.leggauss(P + 2)
    dxf = 1.0 / (2 * n)
    vm = Hn.valid_mask(1)
    mass_f = np.zeros(T.Q)
    for s in np.nonzero(vm)[0]:
        c = Hn.slot_centre(1, int(s))
        iv = np.floor(c / dxf + 1e-9).astype(int)
        civ = iv // 2
        # the parent's slot on level 0 (one patch per 4^D tile, grown by 1)
        ps = None
        for t in np.nonzero(Ho.valid_mask(0))[0]:
            if np.all(np.floor(Ho.slot_centre(0, int(t)) * n + 1e-9).astype(int) == civ):
                ps = int(t)
                break
        child = [iv[d] - 2 * civ[d] for d in range(D)]
        expect = np.zeros((T.Q, T.Np_s))
        for pt in itertools.product(range(P + 2), repeat=D):
            xf = np.array([xg[k] for k in pt])
            w = np.prod([wg[k] for k in pt])
            xc = (xf + (2 * np.array(child) - 1)) / 2.0   # the child's point in the parent's coordinates
            phic = np.array([np.prod([_leg(m[d], xc[d]) for d in range(D)]) for m in mi])
            phif = np.array([np.prod([_leg(m[d], xf[d]) for d in range(D)]) for m in mi])
            uc = Uo[0][:, :, ps] @ phic
            expect += np.outer(uc, phif) * w
        norm = np.array([np.prod([2.0 / (2 * m[d] + 1) for d in range(D)]) for m in mi])
        expect /= norm
        assert np.abs(Un[1][:, :, s] - expect).max() < 1e-13 * np.abs(expect).max()
        mass_f += Un[1][:, 0, s]
    # mass: the children's mode 0 average to the parents'
    nfine = int(vm.sum())
    assert nfine == 8 ** D
    mass_c = np.zeros(T.Q)
    for t in np.nonzero(Ho.valid_mask(0))[0]:
        c = np.floor(Ho.slot_centre(0, int(t)) * n + 1e-9).astype(int)
        if np.all((c >= 2) & (c <= 5)):
            mass_c += Uo[0][:, 0, t]
    assert np.abs(mass_f / 2 ** D - mass_c).max() < 1e-12 * np.abs(mass_c).max()


def test_remake_copies_old_fine_cells(orc):
    D, P, n = 2, 2, 8
    T = orc.Tables(P, D)
    Lold = single_level(n, 4) + [tiles(16, 4, 2, [4, 4], [7, 7])]
    Ho = orc.Hier(D, Lold, [n, n], [0, 0], [1, 1])
    rng = np.random.default_rng(3)
    Uo = orc.alloc_levels(T, Ho)
    for u in Uo:
        u[:] = rng.standard_normal(u.shape)
    new = single_level(n, 4) + [tiles(16, 4, 2, [4, 4], [11, 7])]   # overlaps the old box
    Hn = orc.Hier(D, new, [n, n], [0, 0], [1, 1])
    Un = orc.remake(T, Ho, Uo, Hn)
    old_cells = {}
    for s in np.nonzero(Ho.valid_mask(1))[0]:
        old_cells[tuple(np.floor(Ho.slot_centre(1, int(s)) * 16 + 1e-9).astype(int))] = int(s)
    ncopy = 0
    for s in range(Hn.nslot(1)):
        iv = tuple(np.floor(Hn.slot_centre(1, s) * 16 + 1e-9).astype(int) % 16)
        if iv in old_cells:
            assert np.array_equal(Un[1][:, :, s], Uo[1][:, :, old_cells[iv]])
            ncopy += 1
    assert ncopy >= 16   # the old box's cells, plus ghosts that fall on them


def test_vortex_tags_counted(orc):
    n, D, P = 32, 2, 1
    T = orc.Tables(P, D)
    H = orc.Hier(D, single_level(n, 16), [n, n], [0, 0], [10, 10])
    Uw = orc.alloc_levels(T, H)
    c = 1.4
    tags = orc.tag_cells(T, H, Uw, 0, 0, c, params=(5.0, 5.0))
    xs = (np.arange(n) + 0.5) * 10.0 / n
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    expect = int((np.sqrt((X - 5) ** 2 + (Y - 5) ** 2) <= c ** 2).sum())
    assert int(tags.sum()) == expect > 0


def test_density_gradient_tags_on_a_step(orc):
    n, D, P = 16, 2, 1
    T = orc.Tables(P, D)
    H = orc.Hier(D, single_level(n), [n, n], [0, 0], [1, 1])
    Uw = orc.alloc_levels(T, H)
    for s in np.nonzero(H.valid_mask(0))[0]:
        x = H.slot_centre(0, int(s))[0]
        Uw[0][0, 0, s] = 1.0 if x < 0.5 else 0.125
        Uw[0][3, 0, s] = 2.5
    tags = orc.tag_cells(T, H, Uw, 0, 2, 0.5)
    # columns 7 and 8 beside the step x = 1/2 and columns 15 and 0 beside
    # the periodic wrap: jumps of 0.875 relative to 1 (> 0.5) and to 0.125
    cols = sorted({int(np.floor(H.slot_centre(0, int(s))[0] * n)) for s, t in
                   zip(np.nonzero(H.valid_mask(0))[0], tags) if t})
    assert cols == [0, 7, 8, 15]
    assert int(tags.sum()) == 4 * n
    assert int(orc.tag_cells(T, H, Uw, 0, 2, 8.0).sum()) == 0


def test_tvb_tags_threshold(orc):
    """The ramp of the limiter pins: |Du| = 0.3 in column 2 tagged when
    tvb_c * amr_c * h^2 < 0.3 (h = 1/8), not when above."""
    n, D, P = 8, 2, 1
    rho = [1.0, 1.0, 0.8, 0.5, 0.125, 0.125, 0.125, 0.125]
    T = orc.Tables(P, D)
    H = orc.Hier(D, single_level(n), [n, n], [0, 0], [1, 1])
    for amr_c, expect in ((0.3 * 64 * 1.01, 0), (0.3 * 64 * 0.99, n)):
        Uw = orc.alloc_levels(T, H)
        kx = [k for k, m in enumerate(_mi(T, D)) if m[0] == 1 and m.sum() == 1][0]
        for s in np.nonzero(H.valid_mask(0))[0]:
            i = int(np.floor(H.slot_centre(0, int(s))[0] * n))
            Uw[0][0, 0, s] = rho[i]
            Uw[0][3, 0, s] = 1.0 / (G - 1.0)
            if i == 2:
                Uw[0][0, kx, s] = -0.3
        assert int(orc.tag_cells(T, H, Uw, 0, 1, amr_c, tvb_c=1.0).sum()) == expect
