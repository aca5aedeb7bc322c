"""Pins of the oracle's TVB limiter (Limiter_w / Limiter_linear_tvb,
AmrDG.h:1657-1855; minmodB / minmod, AmrDG_Limiter.tpp:41-84; pde_EV_L/R,
Compressible_Euler.cpp:138-277) on the CPU:
- the characteristic matrices: L R = I and A R = R diag(u-c, u, u+c, u..)
  with A the flux Jacobian of the Euler flux written from the physics
  (complex-step derivative), in 2D and 3D and every direction;
- a density ramp and step at rest, whose characteristic decomposition is the
  density alone, limited to the closed-form minmod slopes (the 1/2 of the
  mean differences, the common-sign rule, the zero at an extremum, the
  zeroing of the higher modes, mode 0 untouched);
- the TVB threshold C M h^2 on either side of the slope it lets through;
- a linear state unchanged away from the periodic wrap."""
import numpy as np
import pytest

from layouts import single_level

G = 1.4


def flux_c(D, d, u):
    """Euler flux in complex arithmetic (for the complex-step Jacobian)."""
    rho, m, E = u[0], u[1:D + 1], u[D + 1]
    p = (G - 1.0) * (E - 0.5 * (m * m).sum() / rho)
    f = np.zeros(D + 2, dtype=complex)
    f[0] = m[d]
    f[1:D + 1] = m * m[d] / rho
    f[1 + d] += p
    f[D + 1] = (E + p) * m[d] / rho
    return f


@pytest.mark.parametrize("D", [2, 3])
def test_characteristic_matrices(orc, D):
    rng = np.random.default_rng(3 + D)
    for _ in range(4):
        rho, p = rng.uniform(0.2, 2.0), rng.uniform(0.1, 3.0)
        vel = rng.uniform(-1.5, 1.5, D)
        u = np.array([rho, *(rho * vel), p / (G - 1) + 0.5 * rho * vel @ vel])
        c = np.sqrt(G * p / rho)
        for d in range(D):
            Lm, Rm = orc.ev_matrices(D, d, u)
            A = np.zeros((D + 2, D + 2))
            for j in range(D + 2):
                uc = u.astype(complex)
                uc[j] += 1e-30j
                A[:, j] = flux_c(D, d, uc).imag / 1e-30
            lam = np.array([vel[d] - c, vel[d], vel[d] + c] + [vel[d]] * (D - 1))
            assert np.abs(Lm @ Rm - np.eye(D + 2)).max() < 1e-14
            assert np.abs(A @ Rm - Rm * lam).max() < 1e-13 * max(1.0, np.abs(A).max())


def _state_at_rest(orc, D, P, n, rho_means):
    """One periodic level n^D on [0,1]^D, zero velocity, pressure 1, cell
    means of density rho_means[i] along x (uniform in the other directions)."""
    T = orc.Tables(P, D)
    levels = single_level(n, None, D)
    H = orc.Hier(D, levels, [n] * D, [0] * D, [1] * D)
    Uw = orc.alloc_levels(T, H)
    valid = H.valid_mask(0)
    ix = {}
    for s in np.nonzero(valid)[0]:
        c = H.slot_centre(0, int(s))
        i = int(np.floor(c[0] * n))
        ix[int(s)] = i
        Uw[0][0, 0, s] = rho_means[i]
        Uw[0][D + 1, 0, s] = 1.0 / (G - 1.0)
    return T, H, Uw, ix


def _mode(T, D, axis):
    """Index of the mode linear in `axis` (degree 0 in the others)."""
    mi = T.get("mi_s").reshape(-1, D).astype(int)
    for k, m in enumerate(mi):
        if m[axis] == 1 and m.sum() == 1:
            return k
    raise AssertionError


@pytest.mark.parametrize("D,P", [(2, 1), (2, 2), (3, 1)])
def test_ramp_limited_to_minmod_slopes(orc, D, P):
    n = 8
    rho = [1.0, 1.0, 0.8, 0.5, 0.125, 0.125, 0.125, 0.125]
    T, H, Uw, ix = _state_at_rest(orc, D, P, n, rho)
    kx, ky = _mode(T, D, 0), _mode(T, D, 1)
    hi = T.Np_s - 1 if P >= 2 else None   # a mode of degree 2 (P >= 2)
    for s, i in ix.items():
        if i == 2:   # steeper than both neighbours' differences: limited to -0.1
            Uw[0][0, kx, s] = -0.3
            Uw[0][0, ky, s] = 0.05          # no variation in y: an extremum, -> 0
            if hi is not None:
                Uw[0][0, hi, s] = 0.01      # -> 0
        if i == 3:   # gentler than both: kept, with its higher modes
            Uw[0][0, kx, s] = -0.12
            if hi is not None:
                Uw[0][0, hi, s] = 0.02
        if i == 4:   # at the foot of the step: minmod(-0.2, -0.1875, 0) = 0
            Uw[0][0, kx, s] = -0.2
    before = [u.copy() for u in Uw]
    nlim = orc.limiter_tvb(T, H, Uw, 0, 0.0, [1.0])
    U = Uw[0]
    cells = {i: [s for s, j in ix.items() if j == i] for i in range(n)}
    for s in cells[2]:
        assert U[0, kx, s] == pytest.approx(-0.1, abs=1e-15)
        assert np.all(U[:, 1:kx, s] == 0) and np.all(U[:, kx + 1:, s] == 0)
        assert np.all(U[1:, kx, s] == 0)
        assert np.array_equal(U[:, 0, s], before[0][:, 0, s])
    for s in cells[3]:
        assert np.array_equal(U[:, :, s], before[0][:, :, s])
    for s in cells[4]:
        assert np.all(U[:, 1:, s] == 0)
    assert nlim == 2 * n ** (D - 1)   # the cells of columns 2 and 4
    valid = H.valid_mask(0)
    assert np.array_equal(U[:, 0, valid], before[0][:, 0, valid])   # no cell mean moves


def test_tvb_threshold(orc):
    """|Du| = 0.3 on a level with h = 1/8: kept when C M h^2 >= 0.3,
    limited below (C the level's coefficient)."""
    n, D, P = 8, 2, 1
    rho = [1.0, 1.0, 0.8, 0.5, 0.125, 0.125, 0.125, 0.125]
    for C, M, limited in ((1.0, 0.3 * 64 * 1.01, False), (1.0, 0.3 * 64 * 0.99, True),
                          (0.5, 0.3 * 64 * 2.02, False), (0.5, 0.3 * 64 * 1.98, True)):
        T, H, Uw, ix = _state_at_rest(orc, D, P, n, rho)
        kx = _mode(T, D, 0)
        for s, i in ix.items():
            if i == 2:
                Uw[0][0, kx, s] = -0.3
        nlim = orc.limiter_tvb(T, H, Uw, 0, M, [C])
        got = [Uw[0][0, kx, s] for s, i in ix.items() if i == 2]
        assert nlim == (n if limited else 0)
        assert np.allclose(got, -0.1 if limited else -0.3, rtol=0, atol=1e-15)


def test_linear_state_unchanged(orc):
    """rho = 1 + 0.4 x with its exact linear modes, at rest: every cell away
    from the periodic wrap keeps its modes (to round-off)."""
    n, D, P = 16, 2, 2
    T = orc.Tables(P, D)
    H = orc.Hier(D, single_level(n), [n, n], [0, 0], [1, 1])
    Uw = orc.alloc_levels(T, H)
    kx = _mode(T, D, 0)
    dx = 1.0 / n
    interior = []
    for s in np.nonzero(H.valid_mask(0))[0]:
        c = H.slot_centre(0, int(s))
        Uw[0][0, 0, s] = 1.0 + 0.4 * c[0]
        Uw[0][0, kx, s] = 0.4 * dx / 2.0
        Uw[0][3, 0, s] = 1.0 / (G - 1.0)
        if 0 < int(np.floor(c[0] * n)) < n - 1:
            interior.append(int(s))
    before = Uw[0].copy()
    orc.limiter_tvb(T, H, Uw, 0, 0.0, [1.0])
    assert np.abs(Uw[0][:, :, interior] - before[:, :, interior]).max() < 1e-15
