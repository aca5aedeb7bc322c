// adg_tables.hpp -- basis, Gauss-Legendre quadrature and element matrices of
// the modal ADER-DG scheme, computed once on the host in FP64 and staged to
// the device.  Restates the reference's QuadratureGaussLegendre<P> and
// BasisLegendre<P> (NumericalMethod/AmrDG/AmrDG.h:53-353) and
// set_vandermat / set_ref_element_matrix (AmrDG_ElementMatrix.tpp:7-198),
// L2ProjInterp::interp_proj_mat / flux_proj_mat (AmrDG_Interpolater.tpp:85-231).
#pragma once

#include <array>
#include <string>
#include <vector>

namespace adg {

struct ElementTables {
    int P = 0, D = 0, Q = 0, N = 0;
    int Np_s = 0, Np_st = 0, qMp_s = 0, qMp_st = 0, qMp_st_bd = 0;
    std::vector<std::array<int, 4>> mi_s, mi_st;   // total-degree multi-indices
    std::vector<double> nodes, weights;            // paired order [x1,-x1,...,0]
    // the 2N-point Gauss-Legendre rule of the L2 norm (AmrDG.h:1587-1655),
    // paired order [x1,-x1,x2,-x2,...]
    std::vector<double> nodes2, weights2;

    // Reference matrices, row-major with the reference's (row, col) meaning.
    std::vector<double> V;            // [qMp_st][Np_st]
    std::vector<double> Vinv;         // [Np_st][qMp_st]
    std::vector<double> Mk_h_w, Mk_h_w_inv;  // [Np_st][Np_st]
    std::vector<double> Mk_pred;      // [Np_st][Np_s]
    std::vector<double> Sk_predVinv;  // [D][Np_st][qMp_st]
    std::vector<double> Sk_corr;      // [D][Np_s][qMp_st]
    std::vector<double> Mkbdm, Mkbdp; // [D][Np_s][qMp_st_bd]
    std::vector<double> Mk_corr;      // [Np_s][Np_s] (diagonal)
    std::vector<double> quadmat;      // [Np_s][qMp_s]
    std::vector<double> Minv;         // [Np_s][Np_s] (diagonal)
    std::vector<double> P_cf, P_fc;   // [2^D][Np_s][Np_s]
    std::vector<double> P_flux_low, P_flux_high;  // [D][2^(D-1)][Np_s][qMp_st_bd]

    // Device forms.
    std::vector<double> A0;   // Mk_h_w_inv * Mk_pred            [Np_st][Np_s]
    std::vector<double> B;    // -Mk_h_w_inv * Sk_predVinv[d]    [D][Np_st][qMp_st]
    std::vector<double> Vbd;  // trace evaluation                [D][2][qMp_st_bd][Np_st]
    std::vector<double> phi0; // spatial basis at spatial node 0 [Np_s]

    int node_idx(int m, int d, int Dtot) const;
};

ElementTables build_element_tables(int P, int D);

// A operands of the 16x16x4 MFMA contractions of the predictor, in fragment
// order: fragment f holds 64 values, lane l = row (l%16) + 16 * kidx (l/16).
// Rows/columns beyond the matrix are zero.  Offsets count fragments.
struct Fragments {
    int A0 = 0, V = 0, B = 0, S = 0, T = 0, BT = 0, total = 0;
    bool tail = false;   // B rows past the last full 16-row block as 4x4x1 16-block fragments
    int MB2 = 0, KSU = 0, NB = 0, KSH = 0, MBS = 0, MBD = 0;
    std::vector<double> data;  // [total][64]
};
// fp64_rows selects the accumulator row order of the FP64 MFMA (row = 4r+g)
// instead of the FP32 one (row = 4g+r).  node_lane selects the fragment set
// of the node-lane predictor (permuted output rows, K step (nb, r) over nodes
// 16nb + 4k + r) instead of the quad-gather one (K step over the nodes the
// accumulator of a nodal product puts in register r).
Fragments build_fragments(const ElementTables& t, bool fp64_rows, bool node_lane);

// Named access for tests (returns empty vector for unknown names).
const std::vector<double>* table_by_name(const ElementTables& t, const std::string& name);

}  // namespace adg
