"""Reflux rule of the oracle.  The reference (L2ProjInterp::reflux,
AmrDG_Interpolater.tpp:39-69) corrects a coarse cell only where one of its
faces is flagged in its own box's cf_face_b_coarse, and that mask is
ownership-filtered (AmrDG_Level.tpp:330-343): a coarse-fine face that is
the low face of the uncovered cell's box is owned by the neighbouring box,
so that cell keeps its uncorrected coarse flux and the composite totals
change.  The conservative rule corrects every cell the registers reach.
Pinned here by what the mathematics fixes: with the conservative rule the
composite totals of a periodic problem stay constant to rounding; where no
coarse-fine face lies on a coarse box's low edge the two rules agree."""
import numpy as np

from layouts import tiles, two_level_centre

VORTEX = [5.0, 5.0, 5.0, 1.0, 0.0]
# level 0: 16^2 in 8^2 boxes; level 1: two 8^2 fine boxes side by side whose
# upper coarse-fine face (coarse row 8) is the low face of the coarse boxes
# above it
STEP_ON_BOX_EDGE = [tiles(16, 8, 2), [((8, 8), (15, 15)), ((16, 8), (23, 15))]]


def _run(orc, levels, conservative, nsteps=2, P=2):
    T = orc.Tables(P, 2)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [10.0, 10.0], conservative_reflux=conservative)
    U = orc.project_ic(T, H, 0, VORTEX)
    orc.average_down(T, H, U)
    orc.fill_patch(T, H, U)
    m0 = orc.totals(T, H, U)
    Dt = orc.set_dt(T, H, U, 0.5)
    orc.evolve(T, H, U, 0.0, Dt, 10.0, 0.5, nsteps_max=nsteps)
    m = orc.totals(T, H, U)
    drift = max(abs(m[q] - m0[q]) / max(abs(m0[q]), 1.0) for q in range(4))
    return U, H, drift


def test_reference_rule_loses_mass_on_a_box_edge(orc):
    _, _, drift = _run(orc, STEP_ON_BOX_EDGE, False)
    assert drift > 1e-7


def test_conservative_rule_conserves(orc):
    _, _, drift = _run(orc, STEP_ON_BOX_EDGE, True)
    assert drift < 1e-13


def test_rules_agree_off_box_edges(orc):
    levels = two_level_centre(16, 16)
    Ua, H, da = _run(orc, levels, False)
    Ub, _, db = _run(orc, levels, True)
    assert da < 1e-13 and db < 1e-13
    for l in range(2):
        v = H.valid_mask(l)
        assert np.array_equal(Ua[l][:, :, v], Ub[l][:, :, v]), f"level {l}"


def _planted_state(orc, cfl_uncovered):
    """Vortex on two levels with one covered coarse cell given a negative
    density at spatial node 0 (mode 1 of density made large), as the L2
    projection of a discontinuity can leave it."""
    levels = two_level_centre(16, 16)
    T = orc.Tables(2, 2)
    H = orc.Hier(2, levels, [16, 16], [0, 0], [10.0, 10.0], cfl_uncovered=cfl_uncovered)
    U = orc.project_ic(T, H, 0, VORTEX)
    orc.average_down(T, H, U)
    orc.fill_patch(T, H, U)
    s = int(np.flatnonzero(H.valid_mask(0))[0]) + 18 * 9 + 9   # coarse cell (9, 9): under the fine level
    assert H.valid_mask(0)[s]
    return T, H, U, s


def test_cfl_over_uncovered_cells(orc):
    """The reference's set_Dt reads covered coarse cells: a negative node
    density there collapses Dt.  Over the uncovered cells only, Dt is the
    one of the state without the planted cell."""
    T, H, U, s = _planted_state(orc, False)
    dt_clean = orc.set_dt(T, H, U, 0.5)
    U[0][0, 1, s] = -5.0 * U[0][0, 0, s]       # node 0 at x1 = +0.77: rho0 - 3.9 rho0 < 0
    dt_ref = orc.set_dt(T, H, U, 0.5)
    assert dt_ref < 0.1 * dt_clean
    T2, H2, U2, s2 = _planted_state(orc, True)
    assert s2 == s
    dt_unc_clean = orc.set_dt(T2, H2, U2, 0.5)
    U2[0][0, 1, s] = -5.0 * U2[0][0, 0, s]
    assert orc.set_dt(T2, H2, U2, 0.5) == dt_unc_clean
    # on a healthy state the finest level sets Dt under either rule
    assert dt_unc_clean == dt_clean
