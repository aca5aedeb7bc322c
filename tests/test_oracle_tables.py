"""Pins of the oracle's basis, quadrature and element matrices (CPU).

Each check is fixed by something other than the oracle's own formula: a
library routine (numpy Legendre), a closed form, an algebraic identity or
brute-force quadrature, chosen so that a dropped term, a wrong sign, index
or transposed operand fails.
"""
import json
import os
from math import comb

import numpy as np
import pytest
from numpy.polynomial import legendre as npl

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [(1, 2), (2, 2), (3, 2), (1, 3), (2, 3)]


def leg(k, x):
    return npl.legval(x, [0] * k + [1])


def dleg(k, x):
    return npl.legval(x, npl.legder([0] * k + [1])) if k > 0 else np.zeros_like(np.asarray(x, float))


def mats(T, name, *shape):
    return T.get(name).reshape(shape)


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"P{c[0]}D{c[1]}")
def tab(request, orc):
    P, D = request.param
    return orc.Tables(P, D)


def test_sizes(tab):
    P, D = tab.P, tab.D
    assert tab.Np_s == comb(P + D, D)
    assert tab.Np_st == comb(P + D + 1, D + 1)
    assert tab.qMp_st == (P + 1) ** (D + 1)
    assert tab.qMp_st_bd == (P + 1) ** D


def test_nodes_weights_vs_numpy(tab):
    # AmrDG.h:225-249: paired order [x1,-x1,x2,-x2,...,0]
    with open(os.path.join(HERE, "golden", "leggauss.json")) as f:
        ref = json.load(f)[str(tab.N)]
    x, w = tab.get("nodes"), tab.get("weights")
    assert np.allclose(np.sort(x), ref["nodes"], rtol=0, atol=2e-16)
    assert np.allclose(w[np.argsort(x)], ref["weights"], rtol=0, atol=1e-15)
    assert x[0] > 0 and x[1] == -x[0]
    if tab.N % 2:
        assert x[-1] == 0.0
    assert abs(w.sum() - 2.0) < 1e-15
    for k in range(2 * tab.N):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs((w * x ** k).sum() - exact) < 1e-14


def test_multi_index_order(tab):
    # MultiIndex<P,D>::generate, AmrDG.h:76-121: first index slowest, total degree <= P
    mi_s = tab.get("mi_s").reshape(tab.Np_s, tab.D).astype(int)
    mi_st = tab.get("mi_st").reshape(tab.Np_st, tab.D + 1).astype(int)
    import itertools

    def enum(dims):
        return [c for c in itertools.product(range(tab.P + 1), repeat=dims) if sum(c) <= tab.P]

    assert [tuple(r) for r in mi_s] == enum(tab.D)
    assert [tuple(r) for r in mi_st] == enum(tab.D + 1)
    if tab.P == 2 and tab.D == 2:
        assert [tuple(r) for r in mi_s] == [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 0)]


def test_vandermonde_and_pinv(tab):
    x = tab.get("nodes")
    mi_st = tab.get("mi_st").reshape(tab.Np_st, tab.D + 1).astype(int)
    V = mats(tab, "V", tab.qMp_st, tab.Np_st)
    N, Dt = tab.N, tab.D + 1
    for i in range(tab.qMp_st):
        idx = [(i // N ** (Dt - 1 - d)) % N for d in range(Dt)]  # dim 0 slowest, time fastest
        for j in range(tab.Np_st):
            ref = np.prod([leg(mi_st[j][d], x[idx[d]]) for d in range(Dt)])
            assert abs(V[i, j] - ref) < 1e-14
    Vinv = mats(tab, "Vinv", tab.Np_st, tab.qMp_st)
    assert np.abs(Vinv @ V - np.eye(tab.Np_st)).max() < 1e-13
    ls = np.linalg.solve(V.T @ V, V.T)
    assert np.abs(Vinv - ls).max() < 1e-13


def test_predictor_matrices_closed_form(tab):
    P, D, Ns, Nst = tab.P, tab.D, tab.Np_s, tab.Np_st
    mi_s = tab.get("mi_s").reshape(Ns, D).astype(int)
    mi_st = tab.get("mi_st").reshape(Nst, D + 1).astype(int)
    xg, wg = npl.leggauss(P + 2)

    def ip(a, b):  # int P_a P'_b
        return float((wg * leg(a, xg) * dleg(b, xg)).sum())

    def mass(k):
        return 2.0 / (2 * k + 1)

    Mhw = mats(tab, "Mk_h_w", Nst, Nst)
    Mpred = mats(tab, "Mk_pred", Nst, Ns)
    Sk = mats(tab, "Sk_pred", D, Nst, Nst)
    for j in range(Nst):
        for i in range(Nst):
            ms = np.prod([mass(mi_st[j][d]) if mi_st[j][d] == mi_st[i][d] else 0.0 for d in range(D)])
            # AmrDG_ElementMatrix.tpp:35-36: M^s_ji (P_tj(1) P_ti(1) - int P_ti P'_tj)
            assert abs(Mhw[j, i] - ms * (1.0 - ip(mi_st[i][D], mi_st[j][D]))) < 1e-14
            tm = mass(mi_st[j][D]) if mi_st[j][D] == mi_st[i][D] else 0.0
            for d in range(D):
                sd = np.prod([ip(mi_st[j][a], mi_st[i][a]) if a == d else
                              (mass(mi_st[j][a]) if mi_st[j][a] == mi_st[i][a] else 0.0) for a in range(D)])
                assert abs(Sk[d, j, i] - tm * sd) < 1e-13
        for i in range(Ns):
            ms = np.prod([mass(mi_st[j][d]) if mi_st[j][d] == mi_s[i][d] else 0.0 for d in range(D)])
            assert abs(Mpred[j, i] - (-1) ** mi_st[j][D] * ms) < 1e-15
    Minv = mats(tab, "Mk_h_w_inv", Nst, Nst)
    assert np.abs(Minv @ Mhw - np.eye(Nst)).max() < 1e-13
    SV = mats(tab, "Sk_predVinv", D, Nst, tab.qMp_st)
    Vinv = mats(tab, "Vinv", Nst, tab.qMp_st)
    for d in range(D):
        assert np.abs(SV[d] - Sk[d] @ Vinv).max() < 1e-13


def test_corrector_matrices(tab):
    D, Ns = tab.D, tab.Np_s
    mi_s = tab.get("mi_s").reshape(Ns, D).astype(int)
    Mk = mats(tab, "Mk_corr", Ns, Ns)
    assert np.allclose(Mk, np.diag([np.prod([2.0 / (2 * k + 1) for k in r]) for r in mi_s]), rtol=0, atol=1e-16)
    Sc = mats(tab, "Sk_corr", D, Ns, tab.qMp_st)
    assert np.abs(Sc[:, 0, :]).max() == 0.0          # dval[0] = 0: mass row is exactly zero
    Mm = mats(tab, "Mkbdm", D, Ns, tab.qMp_st_bd)
    Mp = mats(tab, "Mkbdp", D, Ns, tab.qMp_st_bd)
    w = tab.get("weights")
    N = tab.N
    wbd = np.array([np.prod([w[(i // N ** (D - 1 - p)) % N] for p in range(D)]) for i in range(tab.qMp_st_bd)])
    for d in range(D):
        assert np.array_equal(Mm[d, 0], Mp[d, 0])
        assert np.allclose(Mp[d, 0], wbd, rtol=0, atol=1e-16)
        for k in range(Ns):  # P_k(-1) = (-1)^k
            assert np.array_equal(Mm[d, k], (-1) ** mi_s[k][d] * Mp[d, k])
    # integration of a space-time polynomial: sum_i Sk_corr(j,i) f(x_i) = int dphi_j f
    x = tab.get("nodes")
    for d in range(D):
        for j in range(Ns):
            f = []
            Dt = D + 1
            for i in range(tab.qMp_st):
                idx = [(i // N ** (Dt - 1 - a)) % N for a in range(Dt)]
                f.append(x[idx[0]] ** 2 * (1 + x[idx[-1]]))
            got = (Sc[d, j] * np.array(f)).sum()
            xg, wg = npl.leggauss(6)
            ref = 2.0  # time integral of (1 + t)
            for a in range(D):
                k = mi_s[j][a]
                g = dleg(k, xg) if a == d else leg(k, xg)
                h = xg ** 2 if a == 0 else np.ones_like(xg)
                ref *= (wg * g * h).sum()
            assert abs(got - ref) < 1e-13


def test_projection_exact_and_inverse(tab):
    """Scatter reproduces the coarse polynomial on every child; gather after
    scatter is the identity (AmrDG_Interpolater.tpp:157-231,344-398)."""
    D, Ns = tab.D, tab.Np_s
    mi_s = tab.get("mi_s").reshape(Ns, D).astype(int)
    Minv = mats(tab, "Minv", Ns, Ns)
    Pcf = mats(tab, "P_cf", 2 ** D, Ns, Ns)
    Pfc = mats(tab, "P_fc", 2 ** D, Ns, Ns)
    rng = np.random.default_rng(3)
    uc = rng.standard_normal(Ns)
    back = np.zeros(Ns)
    pts = rng.uniform(-1, 1, size=(7, D))
    for c in range(2 ** D):
        assert np.abs(Pfc[c] - Pcf[c].T).max() < 1e-15
        uf = Minv @ Pcf[c] @ uc
        sh = np.array([0.5 if (c >> (D - 1 - d)) & 1 else -0.5 for d in range(D)])
        for p in pts:
            xc = 0.5 * p + sh
            vc = sum(uc[n] * np.prod([leg(mi_s[n][d], xc[d]) for d in range(D)]) for n in range(Ns))
            vf = sum(uf[n] * np.prod([leg(mi_s[n][d], p[d]) for d in range(D)]) for n in range(Ns))
            assert abs(vc - vf) < 1e-12
        back += Pfc[c] @ uf
    assert np.abs(Minv @ back / 2 ** D - uc).max() < 1e-13


def test_flux_projection_brute_force(tab):
    """P_flux_fc: a fine sub-face integral of (fine node values) projected on
    coarse modes equals direct quadrature of phi_coarse * f on the sub-face."""
    P, D, Ns, N = tab.P, tab.D, tab.Np_s, tab.N
    mi_s = tab.get("mi_s").reshape(Ns, D).astype(int)
    Pl = mats(tab, "P_flux_low", D, 2 ** (D - 1), Ns, tab.qMp_st_bd)
    Ph = mats(tab, "P_flux_high", D, 2 ** (D - 1), Ns, tab.qMp_st_bd)
    x = tab.get("nodes")
    for d in range(D):
        free = [a for a in range(D) if a != d]
        for k in range(2 ** (D - 1)):
            up = {a: (k >> b) & 1 for b, a in enumerate(free)}
            f = []
            for m in range(tab.qMp_st_bd):
                idx = [(m // N ** (D - 1 - p)) % N for p in range(D)]
                # f = (1 + xi_first_free) (1 + tau^2 / 2), fine-face local coordinates
                f.append((1.0 + x[idx[0]]) * (1.0 + 0.5 * x[idx[-1]] ** 2))
            f = np.array(f)
            for r in range(Ns):
                if any(mi_s[r][a] > 0 for a in free[1:]):
                    continue
                xg, wg = npl.leggauss(8)
                a0 = free[0]
                shift = 0.5 if up[a0] else -0.5
                s_int = (wg * leg(mi_s[r][a0], 0.5 * xg + shift) * (1.0 + xg)).sum()
                t_int = 2.0 + 0.5 * (2.0 / 3.0)  # int (1 + t^2/2) dt
                extra = 2.0 ** (len(free) - 1)
                for sgn, M in ((-1, Pl), (1, Ph)):
                    ref = (sgn ** mi_s[r][d]) * s_int * t_int * extra
                    assert abs((M[d, k, r] * f).sum() - ref) < 1e-12


def test_golden_regression(orc):
    """The oracle keeps reproducing its committed fixtures."""
    for P, D in ((1, 2), (2, 2), (3, 2), (2, 3)):
        g = np.load(os.path.join(HERE, "golden", f"oracle_p{P}_d{D}.npz"))
        T = orc.Tables(P, D)
        for name in g.files:
            if name.startswith("step_"):
                continue
            assert np.array_equal(T.get(name), g[name]), name
