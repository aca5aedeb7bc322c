"""Initial states of the oracle that have no reference counterpart, pinned
by what their definitions fix."""
import numpy as np

from layouts import single_level

KH = [0.01, 1e-3, 7]


def _proj(orc, n, case, prm):
    T = orc.Tables(3, 2)
    H = orc.Hier(2, single_level(n, n), [n, n], [0, 0], [1, 1])
    U = orc.project_ic(T, H, case, prm)[0]
    return U[:, :, H.valid_mask(0)].reshape(U.shape[0], U.shape[1], n, n)   # [q][mode][y][x]


def test_fixed_grid_noise_equals_cell_noise_on_its_grid(orc):
    """Case 7 with the noise grid equal to level 0 is case 1 on level 0."""
    assert np.array_equal(_proj(orc, 16, 7, KH + [16]), _proj(orc, 16, 1, KH))


def test_fixed_grid_noise_is_resolution_independent(orc):
    """The cell means of case 7 on 32^2 average, 2x2, to its means on 16^2
    (the noise is constant on the 16^2 cells); with the per-cell noise of
    case 1 they do not."""
    def coarse_means(U):
        m = U[:, 0]
        return 0.25 * (m[:, 0::2, 0::2] + m[:, 1::2, 0::2] + m[:, 0::2, 1::2] + m[:, 1::2, 1::2])
    fine7, crse7 = _proj(orc, 32, 7, KH + [16]), _proj(orc, 16, 7, KH + [16])
    assert np.abs(coarse_means(fine7) - crse7[:, 0]).max() < 1e-10
    fine1, crse1 = _proj(orc, 32, 1, KH), _proj(orc, 16, 1, KH)
    assert np.abs(coarse_means(fine1) - crse1[:, 0]).max() > 1e-5
