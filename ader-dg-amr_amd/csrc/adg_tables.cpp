// adg_tables.cpp -- host construction of the element tables (see header).
//
// The quadrature nodes follow the reference exactly (Newton on P_N from a
// Taylor-series cosine start, AmrDG.h:198-249) because the node order and
// last bits feed every matrix.  The matrices are then built from the
// definitions: the Vandermonde pseudo-inverse by normal equations
// (V^T V)^-1 V^T (equal to the reference's SVD pinv for full column rank),
// Mk_h_w^-1 by Gauss-Jordan elimination, the spatial stiffness by Gauss
// quadrature with Legendre derivatives from the three-term recurrence.
#include "adg_tables.hpp"

#include <cmath>
#include <stdexcept>

namespace adg {
namespace {

// P_n(x) and P'_n(x) by the Bonnet recurrence.
void legendre_pd(int n, double x, double* p, double* dp) {
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (n == 0) { *p = 1.0; *dp = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        double d2 = k * p1 + x * d1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    (void)d0;
    *p = p1;
    *dp = d1;
}
double leg(int n, double x) { double p, d; legendre_pd(n, x, &p, &d); return p; }
double dleg(int n, double x) { double p, d; legendre_pd(n, x, &p, &d); return d; }

double taylor_cos(double x) {
    double r = 1.0, term = 1.0;
    for (int n = 1; n <= 30; ++n) {
        term *= -x * x / ((2 * n - 1) * (2 * n));
        r += term;
    }
    return r;
}

void multi_indices(int P, int dims, std::vector<std::array<int, 4>>& out) {
    // first index slowest, each index bounded by the remaining degree budget
    std::array<int, 4> idx{0, 0, 0, 0};
    struct Rec {
        static void go(int P, int dims, int d, int left, std::array<int, 4>& idx,
                       std::vector<std::array<int, 4>>& out) {
            if (d == dims) { out.push_back(idx); return; }
            for (int v = 0; v <= left; ++v) { idx[d] = v; go(P, dims, d + 1, left - v, idx, out); }
            idx[d] = 0;
        }
    };
    Rec::go(P, dims, 0, P, idx, out);
}

// Solve A X = B (A n x n, B n x k) by Gauss-Jordan with partial pivoting.
std::vector<double> solve(std::vector<double> A, std::vector<double> B, int n, int k) {
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
        if (A[piv * n + c] == 0.0) throw std::runtime_error("singular matrix");
        if (piv != c) {
            for (int j = 0; j < n; ++j) std::swap(A[c * n + j], A[piv * n + j]);
            for (int j = 0; j < k; ++j) std::swap(B[c * k + j], B[piv * k + j]);
        }
        double inv = 1.0 / A[c * n + c];
        for (int j = 0; j < n; ++j) A[c * n + j] *= inv;
        for (int j = 0; j < k; ++j) B[c * k + j] *= inv;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = A[r * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) A[r * n + j] -= f * A[c * n + j];
            for (int j = 0; j < k; ++j) B[r * k + j] -= f * B[c * k + j];
        }
    }
    return B;
}

std::vector<double> matmul(const std::vector<double>& A, const std::vector<double>& B, int m, int k, int n) {
    std::vector<double> C((size_t)m * n, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += A[(size_t)i * k + l] * B[(size_t)l * n + j];
            C[(size_t)i * n + j] = s;
        }
    return C;
}

}  // namespace

int ElementTables::node_idx(int m, int d, int Dtot) const {
    int div = 1;
    for (int i = 0; i < Dtot - 1 - d; ++i) div *= N;
    return (m / div) % N;
}

ElementTables build_element_tables(int P, int D) {
    ElementTables t;
    t.P = P; t.D = D; t.Q = D + 2; t.N = P + 1;
    const int N = t.N;
    // Gauss-Legendre nodes in the reference's paired order (AmrDG.h:198-237).
    t.nodes.assign(N, 0.0);
    t.weights.assign(N, 0.0);
    {
        const double pi = 3.14159265358979323846;
        int at = 0;
        for (int i = 1; i <= N / 2; ++i) {
            double x = taylor_cos(pi * (i - 0.25) / (N + 0.5));
            for (int it = 0; it < 30; ++it) {
                double dx = leg(N, x) / dleg(N, x);
                x -= dx;
                if (std::fabs(dx) < 1e-16) break;
            }
            t.nodes[at++] = x;
            t.nodes[at++] = -x;
        }
        if (N % 2) t.nodes[at] = 0.0;
        for (int q = 0; q < N; ++q) {
            double x = t.nodes[q], d = dleg(N, x);
            t.weights[q] = 2.0 / ((1.0 - x * x) * d * d);
        }
    }
    // The L2 norm's rule: 2N nodes from the asymptotic start
    // cos(theta) (1 - 1/(8 N2^2) + 1/(8 N2^3) - ...) of AmrDG.h:1595-1609,
    // Newton on P_N2 to a step of 1e-15 (AmrDG_Quadrature.tpp:249-268),
    // weights 2 / ((1 - x^2) P'_N2(x)^2).
    {
        const double pi = 3.14159265358979323846;
        const int N2 = 2 * N;
        const double n2 = (double)N2;
        t.nodes2.assign(N2, 0.0);
        t.weights2.assign(N2, 0.0);
        int at = 0;
        for (int i = 1; i <= N2 / 2; ++i) {
            const double th = pi * (i - 0.25) / (n2 + 0.5);
            double x;
            if (i <= (int)((1.0 / 3.0) * n2)) {
                const double s = std::sin(th);
                x = (1.0 - 0.125 / (n2 * n2) + 0.125 / (n2 * n2 * n2) -
                     (1.0 / 384.0) / (n2 * n2 * n2 * n2) * (39.0 - 28.0 / (s * s))) * std::cos(th);
            } else {
                x = (1.0 - 1.0 / (8.0 * n2 * n2) + 1.0 / (8.0 * n2 * n2 * n2)) * std::cos(th);
            }
            for (int it = 0; it < 1000; ++it) {
                const double xn = x - leg(N2, x) / dleg(N2, x);
                const double e = std::fabs(xn - x);
                x = xn;
                if (e <= 1e-15) break;
            }
            t.nodes2[at++] = x;
            t.nodes2[at++] = -x;
        }
        for (int q = 0; q < N2; ++q) {
            const double x = t.nodes2[q], d = dleg(N2, x);
            t.weights2[q] = 2.0 / ((1.0 - x * x) * d * d);
        }
    }
    multi_indices(P, D, t.mi_s);
    multi_indices(P, D + 1, t.mi_st);
    t.Np_s = (int)t.mi_s.size();
    t.Np_st = (int)t.mi_st.size();
    t.qMp_s = 1; for (int d = 0; d < D; ++d) t.qMp_s *= N;
    t.qMp_st = t.qMp_s * N;
    t.qMp_st_bd = t.qMp_s;
    const int Ns = t.Np_s, Nst = t.Np_st, Ms = t.qMp_s, Mst = t.qMp_st, Mbd = t.qMp_st_bd;

    auto nval = [&](int k, int q) { return leg(k, t.nodes[q]); };
    auto ndval = [&](int k, int q) { return dleg(k, t.nodes[q]); };

    // Vandermonde (AmrDG_ElementMatrix.tpp:7-19) and its least-squares inverse.
    t.V.assign((size_t)Mst * Nst, 0.0);
    for (int i = 0; i < Mst; ++i)
        for (int j = 0; j < Nst; ++j) {
            double phi = 1.0;
            for (int d = 0; d <= D; ++d) phi *= nval(t.mi_st[j][d], t.node_idx(i, d, D + 1));
            t.V[(size_t)i * Nst + j] = phi;
        }
    {
        std::vector<double> VtV((size_t)Nst * Nst, 0.0), Vt((size_t)Nst * Mst);
        for (int i = 0; i < Mst; ++i)
            for (int j = 0; j < Nst; ++j) Vt[(size_t)j * Mst + i] = t.V[(size_t)i * Nst + j];
        VtV = matmul(Vt, t.V, Nst, Mst, Nst);
        t.Vinv = solve(VtV, Vt, Nst, Mst);
    }

    // 1D mass of Legendre modes and the spatial / temporal pieces.
    auto mass1 = [](int k) { return 2.0 / (2.0 * k + 1.0); };
    auto Ms_st = [&](int j, int i) {  // spatial mass between two space-time modes
        double m = 1.0;
        for (int d = 0; d < D; ++d) m *= (t.mi_st[j][d] == t.mi_st[i][d]) ? mass1(t.mi_st[j][d]) : 0.0;
        return m;
    };
    // int_{-1}^{1} P_a(x) P'_b(x) dx by N-point Gauss (exact, degree <= 2P-1)
    auto int_p_dp = [&](int a, int b) {
        double s = 0.0;
        for (int q = 0; q < N; ++q) s += t.weights[q] * nval(a, q) * ndval(b, q);
        return s;
    };

    // Predictor matrices (AmrDG_ElementMatrix.tpp:30-58).
    t.Mk_h_w.assign((size_t)Nst * Nst, 0.0);
    t.Mk_pred.assign((size_t)Nst * Ns, 0.0);
    std::vector<double> Sk_pred((size_t)D * Nst * Nst, 0.0);
    for (int j = 0; j < Nst; ++j) {
        int tj = t.mi_st[j][D];
        for (int i = 0; i < Nst; ++i) {
            int ti = t.mi_st[i][D];
            // phi_t(1) = 1; time stiffness int P_ti P'_tj
            t.Mk_h_w[(size_t)j * Nst + i] = Ms_st(j, i) * (1.0 - int_p_dp(ti, tj));
            double tmass = (ti == tj) ? mass1(tj) : 0.0;
            for (int d = 0; d < D; ++d) {
                // int_{[-1,1]^D} phi_j d/dx_d phi_i
                double s = 1.0;
                for (int a = 0; a < D; ++a) {
                    int ja = t.mi_st[j][a], ia = t.mi_st[i][a];
                    if (a == d) s *= int_p_dp(ja, ia);
                    else s *= (ja == ia) ? mass1(ja) : 0.0;
                }
                Sk_pred[((size_t)d * Nst + j) * Nst + i] = tmass * s;
            }
        }
        double pt_m1 = (tj % 2 == 0) ? 1.0 : -1.0;  // P_tj(-1)
        for (int i = 0; i < Ns; ++i) {
            double m = 1.0;
            for (int d = 0; d < D; ++d) m *= (t.mi_st[j][d] == t.mi_s[i][d]) ? mass1(t.mi_st[j][d]) : 0.0;
            t.Mk_pred[(size_t)j * Ns + i] = pt_m1 * m;
        }
    }
    t.Sk_predVinv.assign((size_t)D * Nst * Mst, 0.0);
    for (int d = 0; d < D; ++d) {
        std::vector<double> Sd(Sk_pred.begin() + (size_t)d * Nst * Nst, Sk_pred.begin() + (size_t)(d + 1) * Nst * Nst);
        std::vector<double> SV = matmul(Sd, t.Vinv, Nst, Nst, Mst);
        std::copy(SV.begin(), SV.end(), t.Sk_predVinv.begin() + (size_t)d * Nst * Mst);
    }
    {
        std::vector<double> I((size_t)Nst * Nst, 0.0);
        for (int i = 0; i < Nst; ++i) I[(size_t)i * Nst + i] = 1.0;
        t.Mk_h_w_inv = solve(t.Mk_h_w, I, Nst, Nst);
    }
    t.A0 = matmul(t.Mk_h_w_inv, t.Mk_pred, Nst, Nst, Ns);
    t.B.assign((size_t)D * Nst * Mst, 0.0);
    for (int d = 0; d < D; ++d) {
        std::vector<double> Sd(t.Sk_predVinv.begin() + (size_t)d * Nst * Mst,
                               t.Sk_predVinv.begin() + (size_t)(d + 1) * Nst * Mst);
        std::vector<double> Bd = matmul(t.Mk_h_w_inv, Sd, Nst, Nst, Mst);
        for (size_t i = 0; i < Bd.size(); ++i) t.B[(size_t)d * Nst * Mst + i] = -Bd[i];
    }

    // Corrector matrices (AmrDG_ElementMatrix.tpp:69-131).
    t.Mk_corr.assign((size_t)Ns * Ns, 0.0);
    for (int j = 0; j < Ns; ++j) {
        double m = 1.0;
        for (int d = 0; d < D; ++d) m *= mass1(t.mi_s[j][d]);
        t.Mk_corr[(size_t)j * Ns + j] = m;
    }
    t.Sk_corr.assign((size_t)D * Ns * Mst, 0.0);
    t.Mkbdm.assign((size_t)D * Ns * Mbd, 0.0);
    t.Mkbdp.assign((size_t)D * Ns * Mbd, 0.0);
    std::vector<double> wbd(Mbd, 1.0);
    for (int i = 0; i < Mbd; ++i)
        for (int pos = 0; pos < D; ++pos) wbd[i] *= t.weights[t.node_idx(i, pos, D)];
    for (int d = 0; d < D; ++d) {
        for (int i = 0; i < Mst; ++i) {
            double w = 1.0;
            for (int a = 0; a <= D; ++a) w *= t.weights[t.node_idx(i, a, D + 1)];
            for (int j = 0; j < Ns; ++j) {
                double g = 1.0;
                for (int a = 0; a < D; ++a) {
                    int q = t.node_idx(i, a, D + 1);
                    g *= (a == d) ? ndval(t.mi_s[j][a], q) : nval(t.mi_s[j][a], q);
                }
                t.Sk_corr[((size_t)d * Ns + j) * Mst + i] = g * w;
            }
        }
        for (int i = 0; i < Mbd; ++i)
            for (int j = 0; j < Ns; ++j) {
                double pm = 1.0, pp = 1.0;
                for (int a = 0; a < D; ++a) {
                    int k = t.mi_s[j][a];
                    if (a == d) { pm *= (k % 2 == 0) ? 1.0 : -1.0; pp *= 1.0; }
                    else {
                        int pos = a < d ? a : a - 1;
                        double v = nval(k, t.node_idx(i, pos, D));
                        pm *= v; pp *= v;
                    }
                }
                t.Mkbdm[((size_t)d * Ns + j) * Mbd + i] = pm * wbd[i];
                t.Mkbdp[((size_t)d * Ns + j) * Mbd + i] = pp * wbd[i];
            }
    }
    t.quadmat.assign((size_t)Ns * Ms, 0.0);
    for (int i = 0; i < Ms; ++i) {
        double w = 1.0;
        for (int d = 0; d < D; ++d) w *= t.weights[t.node_idx(i, d, D)];
        for (int j = 0; j < Ns; ++j) {
            double phi = 1.0;
            for (int d = 0; d < D; ++d) phi *= nval(t.mi_s[j][d], t.node_idx(i, d, D));
            t.quadmat[(size_t)j * Ms + i] = phi * w;
        }
    }
    // Face traces of the space-time basis (get_H_from_H_w_bd, AmrDG.tpp:437-493):
    // Vbd[d][side][m][n] = P_{n_d}(+-1) * prod_{free a} P_{n_a}(node(m, pos(a))).
    t.Vbd.assign((size_t)D * 2 * Mbd * Nst, 0.0);
    for (int d = 0; d < D; ++d)
        for (int side = 0; side < 2; ++side)
            for (int m = 0; m < Mbd; ++m)
                for (int n = 0; n < Nst; ++n) {
                    double phi = 1.0;
                    for (int a = 0; a <= D; ++a) {
                        int k = t.mi_st[n][a];
                        if (a == d) phi *= (side == 1 || k % 2 == 0) ? 1.0 : -1.0;
                        else {
                            int pos = a < d ? a : a - 1;
                            phi *= nval(k, t.node_idx(m, pos, D));
                        }
                    }
                    t.Vbd[(((size_t)d * 2 + side) * Mbd + m) * Nst + n] = phi;
                }
    t.phi0.assign(Ns, 1.0);
    for (int n = 0; n < Ns; ++n)
        for (int d = 0; d < D; ++d) t.phi0[n] *= nval(t.mi_s[n][d], 0);

    // L2 projection between levels (AmrDG_Interpolater.tpp:157-231).
    t.Minv.assign((size_t)Ns * Ns, 0.0);
    for (int j = 0; j < Ns; ++j) t.Minv[(size_t)j * Ns + j] = 1.0 / t.Mk_corr[(size_t)j * Ns + j];
    const int nch = 1 << D;
    t.P_cf.assign((size_t)nch * Ns * Ns, 0.0);
    t.P_fc.assign((size_t)nch * Ns * Ns, 0.0);
    auto sub = [&](int k, int q, bool upper) { return leg(k, 0.5 * t.nodes[q] + (upper ? 0.5 : -0.5)); };
    for (int c = 0; c < nch; ++c)
        for (int j = 0; j < Ns; ++j)
            for (int m = 0; m < Ns; ++m) {
                double scf = 0.0, sfc = 0.0;
                for (int q = 0; q < Ms; ++q) {
                    double pm = 1.0, pj = 1.0;
                    for (int d = 0; d < D; ++d) {
                        bool up = ((c >> (D - 1 - d)) & 1) != 0;
                        int qd = t.node_idx(q, d, D);
                        pm *= sub(t.mi_s[m][d], qd, up);
                        pj *= sub(t.mi_s[j][d], qd, up);
                    }
                    scf += t.quadmat[(size_t)j * Ms + q] * pm;
                    sfc += t.quadmat[(size_t)m * Ms + q] * pj;
                }
                t.P_cf[((size_t)c * Ns + j) * Ns + m] = scf;
                t.P_fc[((size_t)c * Ns + j) * Ns + m] = sfc;
            }
    // Fine-face -> coarse-face moment projections (AmrDG_Interpolater.tpp:85-155).
    const int nfc = 1 << (D - 1);
    t.P_flux_low.assign((size_t)D * nfc * Ns * Mbd, 0.0);
    t.P_flux_high.assign((size_t)D * nfc * Ns * Mbd, 0.0);
    for (int d = 0; d < D; ++d)
        for (int k = 0; k < nfc; ++k)
            for (int r = 0; r < Ns; ++r)
                for (int m = 0; m < Mbd; ++m) {
                    double pl = 1.0, ph = 1.0;
                    int bit = 0;
                    for (int a = 0; a < D; ++a) {
                        int kk = t.mi_s[r][a];
                        if (a == d) { pl *= (kk % 2 == 0) ? 1.0 : -1.0; continue; }
                        bool up = ((k >> bit) & 1) != 0;
                        bit++;
                        int pos = a < d ? a : a - 1;
                        double v = sub(kk, t.node_idx(m, pos, D), up);
                        pl *= v; ph *= v;
                    }
                    size_t at = (((size_t)d * nfc + k) * Ns + r) * Mbd + m;
                    t.P_flux_low[at] = pl * wbd[m];
                    t.P_flux_high[at] = ph * wbd[m];
                }
    return t;
}

Fragments build_fragments(const ElementTables& t, bool fp64_rows, bool node_lane) {
    Fragments f;
    const int D = t.D, Ns = t.Np_s, Nst = t.Np_st, Mst = t.qMp_st, Mbd = t.qMp_st_bd;
    auto cdiv = [](int a, int b) { return (a + b - 1) / b; };
    f.MB2 = cdiv(Nst, 16); f.KSU = cdiv(Ns, 4); f.NB = cdiv(Mst, 16);
    f.KSH = cdiv(Nst, 4); f.MBS = cdiv(Ns, 16); f.MBD = cdiv(Mbd, 16);
    std::vector<double>& v = f.data;
    // accumulator row held by register r of lane group k
    auto acc_row = [&](int k, int r) { return fp64_rows ? 4 * r + k : 4 * k + r; };
    // Output-row permutation.  Quad layout: identity.  Node-lane layout: the
    // physical row i of a fragment carries logical row 4r + g where (g, r) is
    // the lane group and register the accumulator puts row i in, so register
    // r of lane group g holds logical row 4r + g in both precisions.
    auto lrow = [&](int i) {
        if (!node_lane || fp64_rows) return i;
        return 4 * (i & 3) + (i >> 2);
    };
    // node of K step (nb, r), lane group k, in the products with nodal data
    auto knode = [&](int nb, int r, int k) { return 16 * nb + (node_lane ? 4 * k + r : acc_row(k, r)); };
    auto frag = [&](auto&& value) {  // value(row_in_block, kidx) -> double
        for (int l = 0; l < 64; ++l) v.push_back(value(l % 16, l / 16));
    };
    // C0 = A0 * U_w: rows Np_st (blocks mb), K = U_w modes 4ks+k
    f.A0 = (int)(v.size() / 64);
    for (int mb = 0; mb < f.MB2; ++mb)
        for (int ks = 0; ks < f.KSU; ++ks)
            frag([&](int i, int k) {
                int row = 16 * mb + lrow(i), col = 4 * ks + k;
                return (row < Nst && col < Ns) ? t.A0[(size_t)row * Ns + col] : 0.0;
            });
    // H = V * H_w: rows nodes (blocks nb), K = H_w modes 4ks+k.  The node-lane
    // kernel uses it as the B operand of the swapped product (same registers).
    f.V = (int)(v.size() / 64);
    for (int nb = 0; nb < f.NB; ++nb)
        for (int ks = 0; ks < f.KSH; ++ks)
            frag([&](int i, int k) {
                int node = 16 * nb + i, mode = 4 * ks + k;
                return (node < Mst && mode < Nst) ? t.V[(size_t)node * Nst + mode] : 0.0;
            });
    // FP32 node-lane form with 4 modes past the last full 16-row block (p = 3
    // in 2D: 20 = 16 + 4): those rows go through the 4x4x1 16-block MFMA,
    // whose block b = lane/4 contracts one node, instead of a 16-row block
    // of which 12 rows would be padding
    f.tail = node_lane && !fp64_rows && f.MB2 >= 2 && Nst - 16 * (f.MB2 - 1) == 4;
    const int MBfull = f.tail ? f.MB2 - 1 : f.MB2;
    // H_w update: rows modes (blocks mb), K step (nb, r) takes knode(nb, r, k)
    f.B = (int)(v.size() / 64);
    for (int mb = 0; mb < MBfull; ++mb)
        for (int d = 0; d < D; ++d)
            for (int nb = 0; nb < f.NB; ++nb)
                for (int r = 0; r < 4; ++r)
                    frag([&](int i, int k) {
                        int mode = 16 * mb + lrow(i), node = knode(nb, r, k);
                        return (mode < Nst && node < Mst) ? t.B[((size_t)d * Nst + mode) * Mst + node] : 0.0;
                    });
    // corrector volume term: rows spatial modes, same K order as B
    f.S = (int)(v.size() / 64);
    for (int mb = 0; mb < f.MBS; ++mb)
        for (int d = 0; d < D; ++d)
            for (int nb = 0; nb < f.NB; ++nb)
                for (int r = 0; r < 4; ++r)
                    frag([&](int i, int k) {
                        int row = 16 * mb + lrow(i), node = knode(nb, r, k);
                        return (row < Ns && node < Mst) ? t.Sk_corr[((size_t)d * Ns + row) * Mst + node] : 0.0;
                    });
    // face traces: rows face nodes (blocks mbd) per (d, side), K = H_w modes
    f.T = (int)(v.size() / 64);
    for (int ds = 0; ds < 2 * D; ++ds)
        for (int mbd = 0; mbd < f.MBD; ++mbd)
            for (int ks = 0; ks < f.KSH; ++ks)
                frag([&](int i, int k) {
                    int row = 16 * mbd + lrow(i), mode = 4 * ks + k;
                    return (row < Mbd && mode < Nst) ? t.Vbd[((size_t)ds * Mbd + row) * Nst + mode] : 0.0;
                });
    // tail rows: lane (k, i) of K step (d, nb, r) holds row 16 (MB2-1) + (i & 3)
    // of B_d at node 16 nb + 4 k + r, the node its 4x4 block contracts
    f.BT = (int)(v.size() / 64);
    if (f.tail)
        for (int d = 0; d < D; ++d)
            for (int nb = 0; nb < f.NB; ++nb)
                for (int r = 0; r < 4; ++r)
                    frag([&](int i, int k) {
                        int mode = 16 * (f.MB2 - 1) + (i & 3), node = 16 * nb + 4 * k + r;
                        return (mode < Nst && node < Mst) ? t.B[((size_t)d * Nst + mode) * Mst + node] : 0.0;
                    });
    f.total = (int)(v.size() / 64);
    return f;
}

const std::vector<double>* table_by_name(const ElementTables& t, const std::string& n) {
    if (n == "nodes") return &t.nodes;
    if (n == "weights") return &t.weights;
    if (n == "nodes2") return &t.nodes2;
    if (n == "weights2") return &t.weights2;
    if (n == "V") return &t.V;
    if (n == "Vinv") return &t.Vinv;
    if (n == "Mk_h_w") return &t.Mk_h_w;
    if (n == "Mk_h_w_inv") return &t.Mk_h_w_inv;
    if (n == "Mk_pred") return &t.Mk_pred;
    if (n == "Sk_predVinv") return &t.Sk_predVinv;
    if (n == "Sk_corr") return &t.Sk_corr;
    if (n == "Mkbdm") return &t.Mkbdm;
    if (n == "Mkbdp") return &t.Mkbdp;
    if (n == "Mk_corr") return &t.Mk_corr;
    if (n == "quadmat") return &t.quadmat;
    if (n == "Minv") return &t.Minv;
    if (n == "P_cf") return &t.P_cf;
    if (n == "P_fc") return &t.P_fc;
    if (n == "P_flux_low") return &t.P_flux_low;
    if (n == "P_flux_high") return &t.P_flux_high;
    if (n == "A0") return &t.A0;
    if (n == "B") return &t.B;
    if (n == "Vbd") return &t.Vbd;
    return nullptr;
}

}  // namespace adg
