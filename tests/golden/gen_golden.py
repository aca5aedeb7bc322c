"""Writes the committed fixtures under tests/golden/.

- leggauss.json: Gauss-Legendre nodes/weights from numpy (an independent
  library routine) for N = 2..5, in numpy's ascending order.
- thesis_l1.json: the density L1 errors the reference publishes for the
  isentropic vortex (Doc/CSE_MSc_Thesis_Fabiano_Sasselli.pdf, Table 5.2,
  p. 40), as data.
- oracle_p{P}_d{D}.npz: element matrices and one vortex ADER step of the
  oracle (regression fixture: the oracle must keep reproducing them).

Run: python tests/golden/gen_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ader-dg-amr_amd")]


def main():
    lg = {}
    for n in range(2, 6):
        x, w = np.polynomial.legendre.leggauss(n)
        lg[str(n)] = {"nodes": x.tolist(), "weights": w.tolist()}
    with open(os.path.join(HERE, "leggauss.json"), "w") as f:
        json.dump(lg, f, indent=1)

    thesis = {
        "source": "Doc/CSE_MSc_Thesis_Fabiano_Sasselli.pdf Table 5.2 p.40; isentropic vortex [0,10]^2, "
                  "beta=5, u_inf=(1,0), T=10, single level",
        "N": [8, 16, 32, 64, 128],
        "dg_p2_L1_rho": [4.4952e-03, 7.5640e-04, 6.3484e-05, 7.7768e-06, 1.1154e-06],
        "dg_p3_L1_rho": [1.4067e-03, 9.1312e-05, 4.0289e-06, 1.8033e-07, 9.2438e-09],
    }
    with open(os.path.join(HERE, "thesis_l1.json"), "w") as f:
        json.dump(thesis, f, indent=1)

    import pyoracle as O
    from layouts import single_level
    names = ["nodes", "weights", "V", "Vinv", "Mk_h_w", "Mk_h_w_inv", "Mk_pred", "Sk_pred", "Sk_predVinv",
             "Mk_corr", "Sk_corr", "Mkbdm", "Mkbdp", "quadmat", "Minv", "P_cf", "P_fc", "P_flux_low",
             "P_flux_high", "mi_s", "mi_st"]
    for P, D in ((1, 2), (2, 2), (3, 2), (2, 3)):
        T = O.Tables(P, D)
        out = {n: T.get(n) for n in names}
        if D == 2:
            H = O.Hier(2, single_level(8), [8, 8], [0, 0], [10, 10])
            prm = [5.0, 5.0, 5.0, 1.0, 0.0]
            Uw = O.project_ic(T, H, 0, prm)
            O.fill_patch(T, H, Uw)
            Dt = O.set_dt(T, H, Uw, 0.5)
            out["step_U0"] = Uw[0].copy()
            out["step_Dt"] = np.array([Dt])
            O.ader(T, H, Uw, Dt)
            out["step_U1"] = Uw[0].copy()
        np.savez_compressed(os.path.join(HERE, f"oracle_p{P}_d{D}.npz"), **out)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
