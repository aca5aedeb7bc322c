"""The oracle's composite Lp norms (LpNorm_DG_AMR, AmrDG.h:1399-1655) pinned
by brute force: numpy's Gauss-Legendre rules, numpy's Legendre series for
the modal basis, the isentropic-vortex density written from its formula,
summed cell by cell over the uncovered cells of a two-level hierarchy.
The L1 norm uses the (P+1)-point rule per dimension, the L2 norm the
2(P+1)-point rule (L2Norm_DG_AMR, AmrDG.h:1587-1655), so a rule of the wrong
size, a dropped weight or volume factor, a covered cell counted or a
missing root fails the comparison."""
import ctypes as C
import itertools

import numpy as np
import pytest
from numpy.polynomial import legendre as npl

from layouts import two_level_centre

G = 1.4
VORTEX = [5.0, 5.0, 5.0, 1.0, 0.0]   # beta, centre x, centre y, u_inf, v_inf
L = 10.0


def multi_indices(P, D):
    """Total-degree multi-indices, first index slowest."""
    return [mi for mi in itertools.product(range(P + 1), repeat=D) if sum(mi) <= P]


def vortex_rho(x, y, t):
    beta, xc, yc, u, v = VORTEX
    xs = np.mod(x - u * t, L)
    ys = np.mod(y - v * t, L)
    r2 = (xs - xc) ** 2 + (ys - yc) ** 2
    return (1.0 - (G - 1.0) * beta ** 2 / (8.0 * G * np.pi ** 2) * np.exp(1.0 - r2)) ** (1.0 / (G - 1.0))


def brute_force(T, H, Uw, levels, P, p, t, nq=None):
    D = 2
    nq = nq or (P + 1) * (2 if p == 2 else 1)
    xg, wg = npl.leggauss(nq)
    mis = multi_indices(P, D)
    tot, vol_amr = 0.0, 0.0
    for l in range(len(levels)):
        dx = L / (H.dom_n[0] * 2 ** l)
        fine = levels[l + 1] if l + 1 < len(levels) else []
        valid = H.valid_mask(l)
        for s in np.nonzero(valid)[0]:
            c = H.slot_centre(l, int(s))
            iv = np.floor(c / dx).astype(int)
            if any(all(lo[d] <= 2 * iv[d] <= hi[d] for d in range(D)) for lo, hi in fine):
                continue
            vol_amr += dx * dx
            cell = 0.0
            for a in range(nq):
                for b in range(nq):
                    xi = (xg[a], xg[b])
                    uh = 0.0
                    for n, mi in enumerate(mis):
                        ph = 1.0
                        for d in range(D):
                            e = np.zeros(mi[d] + 1)
                            e[-1] = 1.0
                            ph *= npl.legval(xi[d], e)
                        uh += Uw[l][0, n, s] * ph
                    ue = vortex_rho(c[0] + 0.5 * dx * xi[0], c[1] + 0.5 * dx * xi[1], t)
                    cell += abs(ue - uh) ** p * wg[a] * wg[b]
            tot += cell * dx * dx / 4.0
    return (tot / vol_amr) ** (1.0 / p)


def test_l2_nodes_are_gauss_legendre(orc):
    for P in (1, 2, 3):
        x = orc.l2_nodes(P)
        N = 2 * (P + 1)
        assert x.size == N
        ref = np.sort(npl.leggauss(N)[0])
        assert np.abs(np.sort(x) - ref).max() < 1e-15
        # the reference's order: [x1, -x1, x2, -x2, ...], x1 > x2 > ... > 0
        pos = x[0::2]
        assert np.all(x[1::2] == -pos) and np.all(np.diff(pos) < 0) and pos[-1] > 0


@pytest.mark.parametrize("P", [1, 2, 3])
def test_product_l2_rule_matches_oracle(adg, orc, P):
    """The library's L2 node table (host side of adg_create) is the oracle's rule."""
    fn = adg.lib().adg_host_table
    fn.argtypes = [C.c_int, C.c_int, C.c_char_p, C.POINTER(C.c_double), C.c_int]
    fn.restype = C.c_int
    n = fn(P, 2, b"nodes2", None, 0)
    got = np.zeros(n)
    fn(P, 2, b"nodes2", got.ctypes.data_as(C.POINTER(C.c_double)), n)
    ref = orc.l2_nodes(P)
    assert np.abs(got - ref).max() <= 2e-16
    w = np.zeros(n)
    fn(P, 2, b"weights2", w.ctypes.data_as(C.POINTER(C.c_double)), n)
    xr, wr = npl.leggauss(n)
    order = np.argsort(got)
    assert np.abs(w[order] - wr).max() < 1e-14


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("P", [1, 2])
def test_lp_norm_brute_force(orc, P, p):
    levels = two_level_centre(4, 4)
    T = orc.Tables(P, 2)
    H = orc.Hier(2, levels, [4, 4], [0, 0], [L, L])
    Uw = orc.project_ic(T, H, 0, VORTEX)
    rng = np.random.default_rng(7 + P)
    for u in Uw:   # modal perturbation so that every mode enters the error
        u += 1e-3 * rng.standard_normal(u.shape)
    t = 0.3
    got = orc.lp_error(T, H, Uw, 0, p, 0, VORTEX, t)
    ref = brute_force(T, H, Uw, levels, P, p, t)
    assert abs(got - ref) <= 1e-12 * ref, (got, ref)


def test_l2_rule_size_is_visible(orc):
    """On the same data the (P+1)-point rule gives a different L2 value, so
    the brute-force comparison above does pin the size of the rule."""
    P = 1
    levels = two_level_centre(4, 4)
    T = orc.Tables(P, 2)
    H = orc.Hier(2, levels, [4, 4], [0, 0], [L, L])
    Uw = orc.project_ic(T, H, 0, VORTEX)
    l2 = orc.lp_error(T, H, Uw, 0, 2, 0, VORTEX, 0.0)
    small = brute_force(T, H, Uw, levels, P, 2, 0.0, nq=P + 1)
    assert abs(l2 - small) > 1e-3 * l2
