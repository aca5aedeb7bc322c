// adg_oracle.cpp -- CPU restatement of the reference AmrDG ADER-DG hot path.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py; never by
// the product.  Built with -ffp-contract=off so that every expression is
// rounded as the reference's non-FMA x86-64 build rounds it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root, prefix Source/).  Where the reference relies
// on a third-party routine (Eigen JacobiSVD / inverse, AMReX FluxRegister and
// FillPatch) the published algorithm is restated and the difference is noted.

#include "adg_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Small dense matrix (row-major, like Eigen's (r,c) indexing used in the
// reference).
struct Mat {
    int r = 0, c = 0;
    std::vector<double> a;
    Mat() = default;
    Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
    double& operator()(int i, int j) { return a[(size_t)i * c + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

Mat matmul(const Mat& A, const Mat& B) {
    Mat C(A.r, B.c);
    for (int i = 0; i < A.r; ++i)
        for (int j = 0; j < B.c; ++j) {
            double s = 0.0;
            for (int k = 0; k < A.c; ++k) s += A(i, k) * B(k, j);
            C(i, j) = s;
        }
    return C;
}

// Inverse by LU with partial pivoting (what Eigen's MatrixXd::inverse() uses
// for dynamic sizes: PartialPivLU).  AmrDG_ElementMatrix.tpp:58.
Mat lu_inverse(const Mat& A) {
    const int n = A.r;
    Mat LU = A;
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(LU(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(LU(i, k)) > best) { best = std::fabs(LU(i, k)); p = i; }
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(LU(k, j), LU(p, j));
            std::swap(perm[k], perm[p]);
        }
        for (int i = k + 1; i < n; ++i) {
            LU(i, k) /= LU(k, k);
            for (int j = k + 1; j < n; ++j) LU(i, j) -= LU(i, k) * LU(k, j);
        }
    }
    Mat inv(n, n);
    for (int col = 0; col < n; ++col) {
        std::vector<double> x(n);
        for (int i = 0; i < n; ++i) x[i] = (perm[i] == col) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) x[i] -= LU(i, j) * x[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) x[i] -= LU(i, j) * x[j];
            x[i] /= LU(i, i);
        }
        for (int i = 0; i < n; ++i) inv(i, col) = x[i];
    }
    return inv;
}

// Pseudo-inverse V * S^-1 * U^T from a thin SVD, no truncation
// (AmrDG_ElementMatrix.tpp:22-25, AmrDG_Interpolater.tpp:187-190).  Eigen's
// two-sided JacobiSVD is replaced by the one-sided (Hestenes) Jacobi SVD; both
// converge to the same factorisation up to rounding, so the result agrees
// with the reference's Vinv to rounding only (last-bit parity unpinned).
Mat svd_pinv(const Mat& A) {
    const int m = A.r, n = A.c;  // m >= n
    Mat W = A;                   // columns rotated into U*S
    Mat V(n, n);
    for (int i = 0; i < n; ++i) V(i, i) = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < m; ++i) {
                    alpha += W(i, p) * W(i, p);
                    beta += W(i, q) * W(i, q);
                    gamma += W(i, p) * W(i, q);
                }
                if (gamma == 0.0) continue;
                double conv = std::fabs(gamma) / std::sqrt(alpha * beta);
                off = std::max(off, conv);
                if (conv < 1e-300) continue;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) /
                           (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
                for (int i = 0; i < m; ++i) {
                    double wp = W(i, p), wq = W(i, q);
                    W(i, p) = cs * wp - sn * wq;
                    W(i, q) = sn * wp + cs * wq;
                }
                for (int i = 0; i < n; ++i) {
                    double vp = V(i, p), vq = V(i, q);
                    V(i, p) = cs * vp - sn * vq;
                    V(i, q) = sn * vp + cs * vq;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<double> sig(n);
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < m; ++i) s += W(i, j) * W(i, j);
        sig[j] = std::sqrt(s);
    }
    Mat Pinv(n, m);  // V * S^-1 * U^T, U = W / sig
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < m; ++k) {
            double s = 0;
            for (int j = 0; j < n; ++j) s += V(i, j) * (1.0 / sig[j]) * (W(k, j) / sig[j]);
            Pinv(i, k) = s;
        }
    return Pinv;
}

// ---------------------------------------------------------------------------
// Compile-time quadrature helpers of QuadratureGaussLegendre<P>,
// AmrDG.h:153-222, evaluated at run time with the same expression order.
double cx_cos(double x) {  // AmrDG.h:160-168 (30-term Taylor series)
    double result = 1.0, term = 1.0;
    for (int n = 1; n <= 30; ++n) {
        term *= -x * x / ((2 * n - 1) * (2 * n));
        result += term;
    }
    return result;
}
double cx_legendre(int n, double x) {  // AmrDG.h:171-180 (Bonnet)
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}
double cx_dlegendre(int n, double x) {  // AmrDG.h:183-195
    if (n == 0) return 0.0;
    double p0 = 1.0, p1 = x, dp1 = 1.0;
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        double dp2 = k * p1 + x * dp1;
        p0 = p1;
        p1 = p2;
        dp1 = dp2;
    }
    return dp1;
}

constexpr double kPi = 3.14159265358979323846;

int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// Total-degree multi-index enumeration, first index slowest
// (MultiIndex<P,D>::generate, AmrDG.h:76-121).
void gen_multi(int P, int D, int depth, int budget, std::array<int, 4>& cur,
               std::vector<std::array<int, 4>>& out) {
    if (depth == D) { out.push_back(cur); return; }
    for (int v = 0; v <= budget; ++v) {
        cur[depth] = v;
        gen_multi(P, D, depth + 1, budget - v, cur, out);
    }
    cur[depth] = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
struct orc_tables {
    int P, D, Q, N, Np_s, Np_st, qMp_s, qMp_st, qMp_st_bd, qMp_s_bd;
    std::vector<std::array<int, 4>> mi_s, mi_st;
    double nodes[8], weights[8];
    double val[8][8], dval[8][8], bd_val[8][2], slo[8][8], shi[8][8];
    std::vector<std::vector<double>> xs;  // xi_ref_quad_s
    Mat V, Vinv, Mk_h_w, Mk_h_w_inv, Mk_pred, Mk_corr, quadmat, M, Minv;
    std::vector<Mat> Sk_pred, Sk_predVinv, Sk_corr, Mkbdm, Mkbdp, P_cf, P_fc;
    std::vector<std::vector<Mat>> Pfl, Pfh;
    std::vector<double> wbd;  // quad_weights_st_bd (same for m and p)

    int node_idx(int m, int d, int Dtot) const {  // AmrDG.h:337-341
        int div = 1;
        for (int i = 0; i < Dtot - 1 - d; ++i) div *= N;
        return (m / div) % N;
    }
    static int bd_free_pos(int a, int dfix) {  // AmrDG.h:345-348
        if (a == dfix) return -1;
        return (a < dfix) ? a : a - 1;
    }
    // phi_t(j, tau) via Legendre<P> (AmrDG_Basis.tpp:4-15, 66-71)
    double phi_t(int j, double tau) const {
        double v[8];
        v[0] = 1.0;
        if (P >= 1) v[1] = tau;
        for (int n = 2; n <= P; ++n) v[n] = ((2 * n - 1) * tau * v[n - 1] - (n - 1) * v[n - 2]) / n;
        return v[mi_st[j][D]];
    }
    // refMat_phiphi over the D spatial indices (AmrDG_ElementMatrix.tpp:200-213)
    double phiphi(const std::array<int, 4>& mj, const std::array<int, 4>& mi) const {
        double m = 1.0;
        for (int d = 0; d < D; ++d)
            m *= (double)(mi[d] == mj[d] ? 1 : 0) * (2.0 / (2.0 * (double)mj[d] + 1.0));
        return m;
    }
    // refMat_phiDphi (AmrDG_ElementMatrix.tpp:215-254), libstdc++ special math
    double phiDphi(const std::array<int, 4>& mj, const std::array<int, 4>& mi, int dim) const {
        double sum = 0.0;
        for (int q = 0; q < qMp_s; ++q) {
            double w = 1.0, phi = 1.0, dphi = 1.0;
            for (int d = 0; d < D; ++d) phi *= std::legendre(mj[d], xs[q][d]);
            for (int a = 0; a < D; ++a) {
                if (a != dim) dphi *= std::legendre(mi[a], xs[q][a]);
                else dphi *= (std::assoc_legendre(mi[dim], 1, xs[q][dim])) /
                             (std::sqrt(1.0 - std::pow(xs[q][dim], 2.0)));
            }
            for (int d = 0; d < D; ++d) w *= 2.0 / std::pow(std::assoc_legendre(N, 1, xs[q][d]), 2.0);
            sum += (phi * dphi * w);
        }
        return sum;
    }
    double tphitphi(int j, int i) const {  // AmrDG_ElementMatrix.tpp:256-271
        int ti = mi_st[i][D], tj = mi_st[j][D];
        return (double)(ti == tj ? 1 : 0) * (2.0 / (2.0 * (double)tj + 1.0));
    }
    double tphiDtphi(int j, int i) const {  // AmrDG_ElementMatrix.tpp:273-305
        double s = 0.0;
        for (int q = 0; q < N; ++q) {
            double x = nodes[q];
            double w = 1.0;
            w *= 2.0 / (double)std::pow((double)std::assoc_legendre(N, 1, x), 2.0);
            double phi_val = std::legendre(mi_st[i][D], x);
            int n_t = mi_st[j][D];
            double dphi_val = 0.0;
            if (n_t >= 1) {
                double pp = 1.0, pc = x;
                dphi_val = 1.0;
                for (int nn = 2; nn <= n_t; ++nn) {
                    double pn = ((2 * nn - 1) * x * pc - (nn - 1) * pp) / nn;
                    dphi_val = nn * pc + x * dphi_val;
                    pp = pc;
                    pc = pn;
                }
            }
            s += (phi_val * dphi_val * w);
        }
        return s;
    }
};

static void build_tables(orc_tables* t) {
    const int P = t->P, D = t->D, N = t->N;
    // nodes, weights, value tables: AmrDG.h:198-318
    {
        int idx = 0;
        for (int i = 1; i <= N / 2; ++i) {
            double theta = kPi * (i - 0.25) / (N + 0.5);
            double x = cx_cos(theta);
            for (int it = 0; it < 30; ++it) {
                double pv = cx_legendre(N, x), dpv = cx_dlegendre(N, x);
                double dx = pv / dpv;
                x -= dx;
                if (std::fabs(dx) < 1e-16) break;
            }
            t->nodes[idx++] = x;
            t->nodes[idx++] = -x;
        }
        if (N % 2 != 0) t->nodes[idx] = 0.0;
        for (int q = 0; q < N; ++q) {
            double x = t->nodes[q], dp = cx_dlegendre(N, x);
            t->weights[q] = 2.0 / ((1.0 - x * x) * dp * dp);
        }
        for (int k = 0; k <= P; ++k) {
            for (int q = 0; q < N; ++q) {
                t->val[k][q] = cx_legendre(k, t->nodes[q]);
                t->dval[k][q] = cx_dlegendre(k, t->nodes[q]);
                t->slo[k][q] = cx_legendre(k, 0.5 * t->nodes[q] - 0.5);
                t->shi[k][q] = cx_legendre(k, 0.5 * t->nodes[q] + 0.5);
            }
            t->bd_val[k][0] = cx_legendre(k, -1.0);
            t->bd_val[k][1] = cx_legendre(k, 1.0);
        }
    }
    // basis multi-indices
    {
        std::array<int, 4> cur{0, 0, 0, 0};
        gen_multi(P, D, 0, P, cur, t->mi_s);
        gen_multi(P, D + 1, 0, P, cur, t->mi_st);
    }
    t->Np_s = (int)t->mi_s.size();
    t->Np_st = (int)t->mi_st.size();
    t->qMp_s = ipow(N, D);
    t->qMp_st = ipow(N, D + 1);
    t->qMp_st_bd = ipow(N, D);
    t->qMp_s_bd = ipow(N, D - 1);
    t->xs.assign(t->qMp_s, std::vector<double>(D));
    for (int m = 0; m < t->qMp_s; ++m)
        for (int d = 0; d < D; ++d) t->xs[m][d] = t->nodes[t->node_idx(m, d, D)];

    const int Ns = t->Np_s, Nst = t->Np_st, Mst = t->qMp_st, Mbd = t->qMp_st_bd, Ms = t->qMp_s;

    // Vandermonde + pinv: AmrDG_ElementMatrix.tpp:7-26
    t->V = Mat(Mst, Nst);
    for (int i = 0; i < Mst; ++i)
        for (int j = 0; j < Nst; ++j) {
            double phi = 1.0;
            for (int d = 0; d < D + 1; ++d) phi *= t->val[t->mi_st[j][d]][t->node_idx(i, d, D + 1)];
            t->V(i, j) = phi;
        }
    t->Vinv = svd_pinv(t->V);

    // predictor matrices: AmrDG_ElementMatrix.tpp:30-58
    t->Mk_h_w = Mat(Nst, Nst);
    t->Mk_pred = Mat(Nst, Ns);
    t->Sk_pred.assign(D, Mat(Nst, Nst));
    for (int j = 0; j < Nst; ++j) {
        for (int i = 0; i < Nst; ++i) {
            t->Mk_h_w(j, i) = t->phiphi(t->mi_st[j], t->mi_st[i]) *
                              ((t->phi_t(j, 1.0) * t->phi_t(i, 1.0)) - t->tphiDtphi(j, i));
            for (int d = 0; d < D; ++d)
                t->Sk_pred[d](j, i) = t->tphitphi(j, i) * t->phiDphi(t->mi_st[j], t->mi_st[i], d);
        }
        for (int i = 0; i < Ns; ++i)
            t->Mk_pred(j, i) = t->phi_t(j, -1.0) * t->phiphi(t->mi_st[j], t->mi_s[i]);
    }
    t->Sk_predVinv.clear();
    for (int d = 0; d < D; ++d) t->Sk_predVinv.push_back(matmul(t->Sk_pred[d], t->Vinv));
    t->Mk_h_w_inv = lu_inverse(t->Mk_h_w);

    // corrector matrices: AmrDG_ElementMatrix.tpp:69-131
    t->Mk_corr = Mat(Ns, Ns);
    for (int j = 0; j < Ns; ++j)
        for (int i = 0; i < Ns; ++i) t->Mk_corr(j, i) = t->phiphi(t->mi_s[j], t->mi_s[i]);
    t->Sk_corr.assign(D, Mat(Ns, Mst));
    t->Mkbdm.assign(D, Mat(Ns, Mbd));
    t->Mkbdp.assign(D, Mat(Ns, Mbd));
    t->wbd.assign(Mbd, 0.0);
    for (int d = 0; d < D; ++d) {
        for (int i = 0; i < Mst; ++i) {
            double w = 1.0;
            for (int dd = 0; dd < D + 1; ++dd) w *= t->weights[t->node_idx(i, dd, D + 1)];
            for (int j = 0; j < Ns; ++j) {
                double dphi = 1.0;
                for (int a = 0; a < D; ++a) {
                    int qa = t->node_idx(i, a, D + 1);
                    dphi *= (a != d) ? t->val[t->mi_s[j][a]][qa] : t->dval[t->mi_s[j][a]][qa];
                }
                t->Sk_corr[d](j, i) = dphi * w;
            }
        }
        for (int i = 0; i < Mbd; ++i) {
            double wm = 1.0, wp = 1.0;
            for (int pos = 0; pos < D; ++pos) {
                int qi = t->node_idx(i, pos, D);
                wm *= t->weights[qi];
                wp *= t->weights[qi];
            }
            t->wbd[i] = wm;
            for (int j = 0; j < Ns; ++j) {
                double pm = 1.0, pp = 1.0;
                for (int a = 0; a < D; ++a) {
                    if (a == d) {
                        pm *= t->bd_val[t->mi_s[j][a]][0];
                        pp *= t->bd_val[t->mi_s[j][a]][1];
                    } else {
                        int qa = t->node_idx(i, orc_tables::bd_free_pos(a, d), D);
                        pm *= t->val[t->mi_s[j][a]][qa];
                        pp *= t->val[t->mi_s[j][a]][qa];
                    }
                }
                t->Mkbdm[d](j, i) = pm * wm;
                t->Mkbdp[d](j, i) = pp * wp;
            }
        }
    }
    // spatial projection matrix: AmrDG_ElementMatrix.tpp:154-168
    t->quadmat = Mat(Ns, Ms);
    for (int i = 0; i < Ms; ++i) {
        double w = 1.0;
        for (int d = 0; d < D; ++d) w *= t->weights[t->node_idx(i, d, D)];
        for (int j = 0; j < Ns; ++j) {
            double phi = 1.0;
            for (int d = 0; d < D; ++d) phi *= t->val[t->mi_s[j][d]][t->node_idx(i, d, D)];
            t->quadmat(j, i) = phi * w;
        }
    }
    // L2 projection matrices: AmrDG_Interpolater.tpp:157-231
    t->M = Mat(Ns, Ns);
    for (int j = 0; j < Ns; ++j)
        for (int n = 0; n < Ns; ++n) t->M(j, n) = t->phiphi(t->mi_s[j], t->mi_s[n]);
    t->Minv = svd_pinv(t->M);
    const int nch = 1 << D;
    t->P_cf.assign(nch, Mat(Ns, Ns));
    t->P_fc.assign(nch, Mat(Ns, Ns));
    for (int l = 0; l < nch; ++l) {
        bool hi[3];
        for (int d = 0; d < D; ++d) hi[d] = ((l >> (D - 1 - d)) & 1) != 0;
        for (int j = 0; j < Ns; ++j)
            for (int m = 0; m < Ns; ++m) {
                double scf = 0.0, sfc = 0.0;
                for (int q = 0; q < Ms; ++q) {
                    double pm = 1.0, pj = 1.0;
                    for (int d = 0; d < D; ++d) {
                        int qd = t->node_idx(q, d, D);
                        const double (*tbl)[8] = hi[d] ? t->shi : t->slo;
                        pm *= tbl[t->mi_s[m][d]][qd];
                        pj *= tbl[t->mi_s[j][d]][qd];
                    }
                    scf += t->quadmat(j, q) * pm;
                    sfc += t->quadmat(m, q) * pj;
                }
                t->P_cf[l](j, m) = scf;
                t->P_fc[l](j, m) = sfc;
            }
    }
    // face projection matrices: AmrDG_Interpolater.tpp:85-155
    const int nfc = 1 << (D - 1);
    t->Pfl.assign(D, std::vector<Mat>(nfc, Mat(Ns, Mbd)));
    t->Pfh.assign(D, std::vector<Mat>(nfc, Mat(Ns, Mbd)));
    for (int d = 0; d < D; ++d)
        for (int k = 0; k < nfc; ++k) {
            bool up[3] = {false, false, false};
            int bc = 0;
            for (int dir = 0; dir < D; ++dir) {
                if (dir == d) continue;
                up[dir] = ((k >> bc) & 1) != 0;
                bc++;
            }
            for (int r = 0; r < Ns; ++r)
                for (int m = 0; m < Mbd; ++m) {
                    double pl = 1.0, ph = 1.0;
                    for (int dim = 0; dim < D; ++dim) {
                        if (dim == d) {
                            pl *= t->bd_val[t->mi_s[r][dim]][0];
                            ph *= t->bd_val[t->mi_s[r][dim]][1];
                        } else {
                            int qa = t->node_idx(m, orc_tables::bd_free_pos(dim, d), D);
                            const double (*tbl)[8] = up[dim] ? t->shi : t->slo;
                            pl *= tbl[t->mi_s[r][dim]][qa];
                            ph *= tbl[t->mi_s[r][dim]][qa];
                        }
                    }
                    t->Pfl[d][k](r, m) = pl * t->wbd[m];
                    t->Pfh[d][k](r, m) = ph * t->wbd[m];
                }
        }
}

// ---------------------------------------------------------------------------
// Pointwise model: Compressible_Euler.h:370-604, Compressible_Euler.cpp:50-70
static const double kLimit = 10 * std::numeric_limits<double>::epsilon();

static double pressure_(int D, const double* u, double gamma) {  // .h:562-589
    double p = 0.0;
    for (int d = 0; d < D; ++d) p += (std::pow(u[d + 1], 2.0) / u[0]);
    p *= (-0.5);
    p += u[D + 1];
    p *= (gamma - 1.0);
    return std::max(p, kLimit);
}
static double soundspeed_(int D, const double* u, double gamma) {  // .h:591-604
    double prs = pressure_(D, u, gamma);
    double rho = std::max(u[0], kLimit);
    return std::sqrt(gamma * (prs / rho));
}
static double flux_(int D, int d, int q, const double* u, double prs) {  // .h:370-487
    if (q == 0) return u[d + 1];
    if (q == D + 1) return (u[D + 1] + prs) * (u[d + 1] / u[0]);
    if (q == d + 1) return (std::pow(u[d + 1], 2.0) / u[0]) + prs;
    return u[d + 1] * u[q] / u[0];
}
static double dflux_(int D, int d, const double* u, double gamma) {  // .h:489-518
    double c = soundspeed_(D, u, gamma);
    double un = 0.0;
    for (int dd = 0; dd < D; ++dd) un += ((u[dd + 1] / u[0]) * (dd == d ? 1.0 : 0.0));
    return std::max({std::fabs(un - c), std::fabs(un), std::fabs(un + c)});
}
static double llf_(double uL, double uR, double fL, double fR, double dfL, double dfR) {
    // AmrDG_NumericalFlux.tpp:34-40
    double C = std::max(std::fabs(dfL), std::fabs(dfR));
    return 0.5 * ((fL + fR) - C * (uR - uL));
}

// ---------------------------------------------------------------------------
// Hierarchy: boxes per level, slot layout, ghost sources, C-F masks.
struct OPatch {
    int lo[3], hi[3], n[3], g[3];
    long off;
    long foff[3];
};
struct OLevel {
    std::vector<OPatch> p;
    long nslot = 0, nfslot[3] = {0, 0, 0};
    int dom[3] = {1, 1, 1};
    double dx[3] = {1, 1, 1};
    std::vector<long> owner;        // domain cell -> valid slot (or -1)
    std::vector<int> valid;         // slot -> 1/0
    // ghost sources: kind 0 copy(src slot) ; kind 1 scatter(coarse slot, child)
    std::vector<long> gslot, gsrc;
    std::vector<int> gkind, gchild, gmask;   // gmask: mirrored directions of a transmissive boundary ghost
    std::vector<int> bcoarse[3], bfine[3], cidx[3];
};

struct orc_hier {
    int D, nlev;
    double plo[3], phi[3];
    int per[3];
    int conservative_reflux = 0;   // 0: the reference's reflux rule
    int cfl_uncovered = 0;         // 0: the reference's set_Dt over every valid cell
    std::vector<OLevel> L;

    long cell_lin(int l, const int* iv) const {
        const OLevel& lv = L[l];
        long s = 0;
        for (int d = D - 1; d >= 0; --d) s = s * lv.dom[d] + iv[d];
        return s;
    }
    // wrap periodic; return false if outside a non-periodic domain
    bool wrap(int l, int* iv) const {
        for (int d = 0; d < D; ++d) {
            int n = L[l].dom[d];
            if (iv[d] < 0 || iv[d] >= n) {
                if (!per[d]) return false;
                iv[d] = ((iv[d] % n) + n) % n;
            }
        }
        return true;
    }
    long find_valid(int l, const int* iv_in) const {
        int iv[3] = {iv_in[0], iv_in[1], iv_in[2]};
        if (!wrap(l, iv)) return -1;
        return L[l].owner[cell_lin(l, iv)];
    }
    long slot_of(int l, int b, const int* iv) const {  // iv within grown box
        const OPatch& p = L[l].p[b];
        long s = 0;
        for (int d = D - 1; d >= 0; --d) s = s * p.g[d] + (iv[d] - p.lo[d] + 1);
        return p.off + s;
    }
    void slot_decode(int l, long slot, int* b_out, int* iv) const {
        const OLevel& lv = L[l];
        int b = 0;
        while (b + 1 < (int)lv.p.size() && lv.p[b + 1].off <= slot) ++b;
        const OPatch& p = lv.p[b];
        long loc = slot - p.off;
        for (int d = 0; d < 3; ++d) iv[d] = 0;
        for (int d = 0; d < D; ++d) {
            iv[d] = (int)(loc % p.g[d]) + p.lo[d] - 1;
            loc /= p.g[d];
        }
        *b_out = b;
    }
    // face slot of face iv (face index along d) in patch b
    long fslot_of(int l, int b, int d, const int* iv) const {
        const OPatch& p = L[l].p[b];
        long s = 0;
        for (int dd = D - 1; dd >= 0; --dd) {
            int fn = p.n[dd] + (dd == d ? 1 : 0);
            s = s * fn + (iv[dd] - p.lo[dd]);
        }
        return p.foff[d] + s;
    }
    bool covered_by_finer(int l, const int* iv) const {  // coarse cell under level l+1
        if (l + 1 >= nlev) return false;
        for (const OPatch& f : L[l + 1].p) {
            bool in = true;
            for (int d = 0; d < D; ++d) {
                int clo = f.lo[d] >> 1, chi = f.hi[d] >> 1;
                if (iv[d] < clo || iv[d] > chi) in = false;
            }
            if (in) return true;
        }
        return false;
    }
};

static void build_hier(orc_hier* h) {
    const int D = h->D;
    for (int l = 0; l < h->nlev; ++l) {
        OLevel& lv = h->L[l];
        long off = 0, foff[3] = {0, 0, 0};
        for (OPatch& p : lv.p) {
            for (int d = 0; d < 3; ++d) { p.n[d] = 1; p.g[d] = 1; }
            for (int d = 0; d < D; ++d) { p.n[d] = p.hi[d] - p.lo[d] + 1; p.g[d] = p.n[d] + 2; }
            p.off = off;
            off += (long)p.g[0] * p.g[1] * p.g[2];
            for (int d = 0; d < D; ++d) {
                long nf = 1;
                for (int dd = 0; dd < D; ++dd) nf *= p.n[dd] + (dd == d ? 1 : 0);
                p.foff[d] = foff[d];
                foff[d] += nf;
            }
        }
        lv.nslot = off;
        for (int d = 0; d < 3; ++d) lv.nfslot[d] = foff[d];
        long ndom = 1;
        for (int d = 0; d < D; ++d) ndom *= lv.dom[d];
        lv.owner.assign(ndom, -1);
        lv.valid.assign(lv.nslot, 0);
        for (int b = 0; b < (int)lv.p.size(); ++b) {
            const OPatch& p = lv.p[b];
            int iv[3] = {0, 0, 0};
            for (iv[2] = (D > 2 ? p.lo[2] : 0); iv[2] <= (D > 2 ? p.hi[2] : 0); ++iv[2])
                for (iv[1] = p.lo[1]; iv[1] <= p.hi[1]; ++iv[1])
                    for (iv[0] = p.lo[0]; iv[0] <= p.hi[0]; ++iv[0]) {
                        long s = h->slot_of(l, b, iv);
                        lv.owner[h->cell_lin(l, iv)] = s;
                        lv.valid[s] = 1;
                    }
        }
    }
    // ghost sources: AMReX FillPatchSingleLevel / FillPatchTwoLevels semantics
    // with the L2ProjInterp mapper (AmrDG_Level.tpp:172-224,
    // AmrDG_Interpolater.tpp:251-281,344-364,400-423).
    for (int l = 0; l < h->nlev; ++l) {
        OLevel& lv = h->L[l];
        for (long s = 0; s < lv.nslot; ++s) {
            if (lv.valid[s]) continue;
            int b, iv[3];
            h->slot_decode(l, s, &b, iv);
            // outside a non-periodic boundary: transmissive (outflow) ghost as
            // the mirror image across the boundary face (the reference's own
            // non-periodic modal BC is a stub, Compressible_Euler.h:58-72)
            int mv[3] = {iv[0], iv[1], iv[2]}, mask = 0;
            for (int d = 0; d < D; ++d) {
                const int n = lv.dom[d];
                if (h->per[d]) continue;
                if (iv[d] < 0) { mv[d] = -1 - iv[d]; mask |= 1 << d; }
                if (iv[d] >= n) { mv[d] = 2 * n - 1 - iv[d]; mask |= 1 << d; }
            }
            long src = h->find_valid(l, mv);
            if (src >= 0) {
                lv.gslot.push_back(s); lv.gsrc.push_back(src);
                lv.gkind.push_back(0); lv.gchild.push_back(0); lv.gmask.push_back(mask);
                continue;
            }
            if (l == 0) {  // no source (cannot happen on a level-0 cover)
                lv.gslot.push_back(s); lv.gsrc.push_back(-1);
                lv.gkind.push_back(2); lv.gchild.push_back(0); lv.gmask.push_back(0);
                continue;
            }
            int civ[3] = {0, 0, 0};
            int child = 0;
            for (int d = 0; d < D; ++d) {
                int c = (mv[d] >= 0) ? mv[d] / 2 : -((-mv[d] + 1) / 2);  // coarsen (floor)
                civ[d] = c;
                int bit = mv[d] - c * 2;
                child += bit * (1 << (D - 1 - d));
            }
            long csrc = h->find_valid(l - 1, civ);
            lv.gslot.push_back(s); lv.gsrc.push_back(csrc);
            lv.gkind.push_back(csrc >= 0 ? 1 : 2); lv.gchild.push_back(child); lv.gmask.push_back(mask);
        }
    }
    // C-F face masks: AmrDG_Level.tpp:240-414
    for (int l = 0; l < h->nlev; ++l) {
        OLevel& lv = h->L[l];
        // cf_mask (coverage by level l+1) on valid + FillBoundary ghosts
        std::vector<int> cfm(lv.nslot, 0), vm(lv.nslot, 0);
        for (long s = 0; s < lv.nslot; ++s) {
            int b, iv[3];
            h->slot_decode(l, s, &b, iv);
            if (lv.valid[s]) {
                cfm[s] = h->covered_by_finer(l, iv) ? 1 : 0;
                vm[s] = 1;
            } else {
                int w[3] = {iv[0], iv[1], iv[2]};
                bool indom = h->wrap(l, w);
                long src = indom ? lv.owner[h->cell_lin(l, w)] : -1;
                if (src >= 0) {
                    cfm[s] = h->covered_by_finer(l, w) ? 1 : 0;
                    vm[s] = 1;
                } else if (!indom) {
                    // fill_boundary_ghosts: cf_mask copies adjacent valid, valid_mask = 1
                    int a[3] = {iv[0], iv[1], iv[2]};
                    for (int d = 0; d < D; ++d) a[d] = std::min(std::max(a[d], 0), lv.dom[d] - 1);
                    cfm[s] = h->covered_by_finer(l, a) ? 1 : 0;
                    vm[s] = 1;
                }
            }
        }
        for (int d = 0; d < D; ++d) {
            lv.bcoarse[d].assign(lv.nfslot[d], 0);
            lv.bfine[d].assign(lv.nfslot[d], 0);
            lv.cidx[d].assign(lv.nfslot[d], 0);
            for (int b = 0; b < (int)lv.p.size(); ++b) {
                const OPatch& p = lv.p[b];
                int hi[3] = {p.hi[0], p.hi[1], D > 2 ? p.hi[2] : 0};
                int lo[3] = {p.lo[0], p.lo[1], D > 2 ? p.lo[2] : 0};
                hi[d] += 1;
                int iv[3];
                for (iv[2] = lo[2]; iv[2] <= hi[2]; ++iv[2])
                    for (iv[1] = lo[1]; iv[1] <= hi[1]; ++iv[1])
                        for (iv[0] = lo[0]; iv[0] <= hi[0]; ++iv[0]) {
                            int L_[3] = {iv[0], iv[1], iv[2]};
                            L_[d] -= 1;
                            long fs = h->fslot_of(l, b, d, iv);
                            long sl = h->slot_of(l, b, L_), sr = h->slot_of(l, b, iv);
                            if (l + 1 < h->nlev && L_[d] >= p.lo[d]) {  // ownership rule
                                int ml = cfm[sl], mr = cfm[sr];
                                if (ml != mr) lv.bcoarse[d][fs] = (ml == 0) ? 1 : -1;
                            }
                            if (l > 0) {
                                int vl = vm[sl], vr = vm[sr];
                                if (vl != vr) {
                                    lv.bfine[d][fs] = (vl == 0) ? 1 : -1;
                                    int ci = 0, bp = 0;
                                    for (int dir = 0; dir < D; ++dir) {
                                        if (dir == d) continue;
                                        if (iv[dir] % 2 != 0) ci |= (1 << bp);
                                        bp++;
                                    }
                                    lv.cidx[d][fs] = ci;
                                }
                            }
                        }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Per-level scratch mirroring the reference MultiFabs.
namespace {
struct Scratch {
    std::vector<double> Hw, H, F, Hm, Hp, Fm, Fp, DFm, DFp, rhs;
    std::vector<double> Fnum[3], Fic[3], Fif[3];
};

inline double& A4(std::vector<double>& v, long ncomp_stride_slots, int q, int ncomp, int c, long s) {
    return v[((long)q * ncomp + c) * ncomp_stride_slots + s];
}

// set_predictor: AmrDG.tpp:495-537 (grown box, flat index copy)
void set_predictor(const orc_tables* t, long ns, const double* Uw, double* Hw) {
    const int Q = t->Q, Ns = t->Np_s, Nst = t->Np_st;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s)
        for (int n = 0; n < Nst; ++n)
            for (int q = 0; q < Q; ++q)
                Hw[((long)q * Nst + n) * ns + s] = (n < Ns) ? Uw[((long)q * Ns + n) * ns + s] : 0.0;
}

// get_H_from_H_w: AmrDG.tpp:387-435 (n outer, += per m)
void get_H(const orc_tables* t, long ns, const double* Hw, double* H) {
    const int Q = t->Q, Nst = t->Np_st, M = t->qMp_st, D = t->D;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s) {
        for (int q = 0; q < Q; ++q)
            for (int m = 0; m < M; ++m) H[((long)q * M + m) * ns + s] = 0.0;
        for (int n = 0; n < Nst; ++n)
            for (int q = 0; q < Q; ++q)
                for (int m = 0; m < M; ++m) {
                    double phi = 1.0;
                    for (int d = 0; d < D + 1; ++d) phi *= t->val[t->mi_st[n][d]][t->node_idx(m, d, D + 1)];
                    H[((long)q * M + m) * ns + s] += Hw[((long)q * Nst + n) * ns + s] * phi;
                }
    }
}

// get_H_from_H_w_bd: AmrDG.tpp:437-493
void get_H_bd(const orc_tables* t, long ns, const double* Hw, double* Hb, int dfix, int side) {
    const int Q = t->Q, Nst = t->Np_st, M = t->qMp_st_bd, D = t->D;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s) {
        for (int q = 0; q < Q; ++q)
            for (int m = 0; m < M; ++m) Hb[((long)q * M + m) * ns + s] = 0.0;
        for (int n = 0; n < Nst; ++n)
            for (int q = 0; q < Q; ++q)
                for (int m = 0; m < M; ++m) {
                    double phi = 1.0;
                    for (int a = 0; a <= D; ++a) {
                        if (a == dfix) phi *= t->bd_val[t->mi_st[n][a]][side];
                        else phi *= t->val[t->mi_st[n][a]][t->node_idx(m, orc_tables::bd_free_pos(a, dfix), D)];
                    }
                    Hb[((long)q * M + m) * ns + s] += Hw[((long)q * Nst + n) * ns + s] * phi;
                }
    }
}

// AmrDG::flux -> pde_flux: AmrDG.h:1167-1205
void flux_vol(const orc_tables* t, long ns, const double* H, double* F, int d, double gamma) {
    const int Q = t->Q, M = t->qMp_st, D = t->D;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s)
        for (int m = 0; m < M; ++m) {
            double u[5];
            for (int q = 0; q < Q; ++q) u[q] = H[((long)q * M + m) * ns + s];
            double prs = pressure_(D, u, gamma);
            for (int q = 0; q < Q; ++q) F[((long)q * M + m) * ns + s] = flux_(D, d, q, u, prs);
        }
}

// AmrDG::flux_bd -> pde_flux + pde_dflux: AmrDG.h:1207-1251
void flux_bd(const orc_tables* t, long ns, const double* Hb, double* Fb, double* DFb, int d, double gamma) {
    const int Q = t->Q, M = t->qMp_st_bd, D = t->D;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s)
        for (int m = 0; m < M; ++m) {
            double u[5];
            for (int q = 0; q < Q; ++q) u[q] = Hb[((long)q * M + m) * ns + s];
            double prs = pressure_(D, u, gamma);
            double df = dflux_(D, d, u, gamma);
            for (int q = 0; q < Q; ++q) {
                Fb[((long)q * M + m) * ns + s] = flux_(D, d, q, u, prs);
                DFb[((long)q * M + m) * ns + s] = df;
            }
        }
}

// update_H_w: AmrDG.tpp:744-827 (per q; rhs reset; grown box)
void update_H_w(const orc_tables* t, long ns, const double* Uw, const double* F, double* Hw,
                const double* dx, double Dt) {
    const int Q = t->Q, Ns = t->Np_s, Nst = t->Np_st, M = t->qMp_st, D = t->D;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s) {
        double rhs[128];
        for (int q = 0; q < Q; ++q) {
            for (int n = 0; n < Nst; ++n) {
                rhs[n] = 0.0;
                Hw[((long)q * Nst + n) * ns + s] = 0.0;
                for (int m = 0; m < Ns; ++m) rhs[n] += t->Mk_pred(n, m) * Uw[((long)q * Ns + m) * ns + s];
            }
            for (int d = 0; d < D; ++d)
                for (int n = 0; n < Nst; ++n)
                    for (int m = 0; m < M; ++m)
                        rhs[n] -= ((Dt / dx[d]) * t->Sk_predVinv[d](n, m) *
                                   F[(((long)d * Q + q) * M + m) * ns + s]);
            for (int n = 0; n < Nst; ++n)
                for (int m = 0; m < Nst; ++m) Hw[((long)q * Nst + n) * ns + s] += t->Mk_h_w_inv(n, m) * rhs[m];
        }
    }
}
}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

orc_tables* orc_tables_create(int P, int D) {
    if (P < 1 || P > 3 || D < 2 || D > 3) return nullptr;
    orc_tables* t = new orc_tables();
    t->P = P; t->D = D; t->Q = D + 2; t->N = P + 1;
    build_tables(t);
    return t;
}
void orc_tables_destroy(orc_tables* t) { delete t; }
void orc_tables_sizes(const orc_tables* t, int* o) {
    o[0] = t->Np_s; o[1] = t->Np_st; o[2] = t->qMp_s; o[3] = t->qMp_st;
    o[4] = t->qMp_st_bd; o[5] = t->N; o[6] = t->Q;
}

static int put(const Mat& m, double* out, int cap, int at) {
    for (size_t i = 0; i < m.a.size(); ++i)
        if (at + (int)i < cap) out[at + i] = m.a[i];
    return at + (int)m.a.size();
}

int orc_tables_get(const orc_tables* t, const char* nm, double* out, int cap) {
    std::string n(nm);
    int at = 0;
    auto put1 = [&](double v) { if (at < cap) out[at] = v; ++at; };
    if (n == "nodes") { for (int i = 0; i < t->N; ++i) put1(t->nodes[i]); return at; }
    if (n == "weights") { for (int i = 0; i < t->N; ++i) put1(t->weights[i]); return at; }
    if (n == "val" || n == "dval" || n == "shifted_lo" || n == "shifted_hi") {
        for (int k = 0; k <= t->P; ++k)
            for (int q = 0; q < t->N; ++q)
                put1(n == "val" ? t->val[k][q] : n == "dval" ? t->dval[k][q]
                     : n == "shifted_lo" ? t->slo[k][q] : t->shi[k][q]);
        return at;
    }
    if (n == "bd_val") { for (int k = 0; k <= t->P; ++k) { put1(t->bd_val[k][0]); put1(t->bd_val[k][1]); } return at; }
    if (n == "mi_s") { for (auto& e : t->mi_s) for (int d = 0; d < t->D; ++d) put1(e[d]); return at; }
    if (n == "mi_st") { for (auto& e : t->mi_st) for (int d = 0; d <= t->D; ++d) put1(e[d]); return at; }
    if (n == "V") return put(t->V, out, cap, 0);
    if (n == "Vinv") return put(t->Vinv, out, cap, 0);
    if (n == "Mk_h_w") return put(t->Mk_h_w, out, cap, 0);
    if (n == "Mk_h_w_inv") return put(t->Mk_h_w_inv, out, cap, 0);
    if (n == "Mk_pred") return put(t->Mk_pred, out, cap, 0);
    if (n == "Mk_corr") return put(t->Mk_corr, out, cap, 0);
    if (n == "quadmat") return put(t->quadmat, out, cap, 0);
    if (n == "Minv") return put(t->Minv, out, cap, 0);
    const std::vector<Mat>* v = nullptr;
    if (n == "Sk_pred") v = &t->Sk_pred;
    if (n == "Sk_predVinv") v = &t->Sk_predVinv;
    if (n == "Sk_corr") v = &t->Sk_corr;
    if (n == "Mkbdm") v = &t->Mkbdm;
    if (n == "Mkbdp") v = &t->Mkbdp;
    if (n == "P_cf") v = &t->P_cf;
    if (n == "P_fc") v = &t->P_fc;
    if (v) { for (const Mat& m : *v) at = put(m, out, cap, at); return at; }
    if (n == "P_flux_low" || n == "P_flux_high") {
        const auto& vv = (n == "P_flux_low") ? t->Pfl : t->Pfh;
        for (const auto& row : vv) for (const Mat& m : row) at = put(m, out, cap, at);
        return at;
    }
    return -1;
}

orc_hier* orc_hier_create(int D, int nlev, const int* npatch, const int* boxes, const int* dom_n,
                          const double* prob_lo, const double* prob_hi, const int* periodic) {
    orc_hier* h = new orc_hier();
    h->D = D; h->nlev = nlev;
    for (int d = 0; d < 3; ++d) {
        h->plo[d] = d < D ? prob_lo[d] : 0.0;
        h->phi[d] = d < D ? prob_hi[d] : 1.0;
        h->per[d] = d < D ? periodic[d] : 1;
    }
    h->L.resize(nlev);
    int at = 0;
    for (int l = 0; l < nlev; ++l) {
        OLevel& lv = h->L[l];
        for (int d = 0; d < D; ++d) {
            lv.dom[d] = dom_n[d] << l;
            lv.dx[d] = (h->phi[d] - h->plo[d]) / lv.dom[d];
        }
        for (int b = 0; b < npatch[l]; ++b) {
            OPatch p{};
            for (int d = 0; d < 3; ++d) { p.lo[d] = 0; p.hi[d] = 0; }
            for (int d = 0; d < D; ++d) { p.lo[d] = boxes[at + d]; p.hi[d] = boxes[at + D + d]; }
            at += 2 * D;
            lv.p.push_back(p);
        }
    }
    build_hier(h);
    return h;
}
void orc_hier_destroy(orc_hier* h) { delete h; }
void orc_hier_set_conservative_reflux(orc_hier* h, int on) { h->conservative_reflux = on != 0; }
void orc_hier_set_cfl_uncovered(orc_hier* h, int on) { h->cfl_uncovered = on != 0; }
int orc_hier_nslot(const orc_hier* h, int l) { return (int)h->L[l].nslot; }
int orc_hier_nfslot(const orc_hier* h, int l, int d) { return (int)h->L[l].nfslot[d]; }
void orc_hier_valid_mask(const orc_hier* h, int l, int* out) {
    for (long s = 0; s < h->L[l].nslot; ++s) out[s] = h->L[l].valid[s];
}
void orc_hier_face_mask(const orc_hier* h, int l, int d, int kind, int* out) {
    const std::vector<int>& v = kind == 0 ? h->L[l].bcoarse[d] : kind == 1 ? h->L[l].bfine[d] : h->L[l].cidx[d];
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
}
void orc_hier_slot_centre(const orc_hier* h, int l, int slot, double* x) {
    int b, iv[3];
    h->slot_decode(l, slot, &b, iv);
    for (int d = 0; d < h->D; ++d) x[d] = h->plo[d] + (iv[d] + 0.5) * h->L[l].dx[d];
}

// ---- pointwise model wrappers
double orc_pressure(int D, const double* u, double g) { return pressure_(D, u, g); }
double orc_pde_flux(int D, int d, int q, const double* u, double g) { return flux_(D, d, q, u, pressure_(D, u, g)); }
double orc_pde_dflux(int D, int d, const double* u, double g) { return dflux_(D, d, u, g); }
double orc_cfl_lambda(int D, int d, const double* u, double g) {  // .cpp:57-70
    double c = soundspeed_(D, u, g);
    return std::fabs(u[d + 1] / u[0]) + c;
}
double orc_llf(double uL, double uR, double fL, double fR, double dfL, double dfR) {
    return llf_(uL, uR, fL, fR, dfL, dfR);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Initial conditions (harness).  The vortex restates Compressible_Euler.h:
// 148-187 with (u_inf, v_inf, beta, centre) as parameters; the KH state with
// seeded cell noise and the uniform state have no reference implementation.
static uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// uniform [0,1) per (seed, level, cell) -- same generator as the product host
static double cell_noise(uint64_t seed, int lev, const int* iv) {
    uint64_t k = splitmix(seed ^ 0x5eedull);
    k = splitmix(k ^ (uint64_t)(uint32_t)lev);
    k = splitmix(k ^ (uint64_t)(uint32_t)iv[0]);
    k = splitmix(k ^ (uint64_t)(uint32_t)iv[1]);
    k = splitmix(k ^ (uint64_t)(uint32_t)iv[2]);
    return (double)(k >> 11) * (1.0 / 9007199254740992.0);
}
// uniform [0,1) per (seed, perturbation mode k)
static double mode_noise(uint64_t seed, int k) {
    uint64_t h = splitmix(seed ^ 0x6b68ull);
    h = splitmix(h ^ (uint64_t)(uint32_t)k);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static void ic_state(int D, int ic_case, const double* prm, double gamma, const double* x,
                     int lev, const int* iv, double* u) {
    for (int q = 0; q < 5; ++q) u[q] = 0.0;
    if (ic_case == 0) {
        double beta = prm[0], xc = prm[1], yc = prm[2], uinf = prm[3], vinf = prm[4];
        double r = std::sqrt(((x[0] - xc) * (x[0] - xc) + (x[1] - yc) * (x[1] - yc)));
        double rho = std::pow(1.0 - (((gamma - 1.0) * std::pow(beta, 2.0)) /
                                     (8.0 * gamma * std::pow(M_PI, 2))) * std::exp(1.0 - std::pow(r, 2.0)),
                              (1.0 / (gamma - 1.0)));
        double prs = std::pow(rho, gamma);
        double du1 = -(x[1] - yc) * (beta / (M_PI * 2.0)) * std::exp((1.0 - std::pow(r, 2.0)) / 2.0);
        double du2 = (x[0] - xc) * (beta / (M_PI * 2.0)) * std::exp((1.0 - std::pow(r, 2.0)) / 2.0);
        double u1 = uinf + du1, u2 = vinf + du2;
        double e = (prs / (rho * (gamma - 1.0))) + 0.5 * (std::pow(u1, 2.0) + std::pow(u2, 2.0));
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2; u[3] = rho * e;
    } else if (ic_case == 1) {
        double amp = prm[0], namp = prm[1];
        uint64_t seed = (uint64_t)prm[2];
        double y = x[1];
        bool in = std::fabs(y - 0.5) < 0.25;
        double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        double u2 = amp * std::sin(4.0 * M_PI * x[0]) + namp * (2.0 * cell_noise(seed, lev, iv) - 1.0);
        double E = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2; u[D + 1] = E;
    } else if (ic_case == 7) {
        // KH as case 1, noise constant on the cells of an N x N grid of the
        // unit square
        double amp = prm[0], namp = prm[1];
        uint64_t seed = (uint64_t)prm[2];
        int N = (int)prm[3];
        int nv[3] = {0, 0, 0};
        for (int d = 0; d < 2; ++d) nv[d] = std::min(N - 1, std::max(0, (int)std::floor(x[d] * N)));
        bool in = std::fabs(x[1] - 0.5) < 0.25;
        double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        double u2 = amp * std::sin(4.0 * M_PI * x[0]) + namp * (2.0 * cell_noise(seed, 0, nv) - 1.0);
        double E = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2; u[D + 1] = E;
    } else if (ic_case == 4) {
        // KH with v = sum_{k=1..8} a_k sin(2 pi k x), a_k = amax * r_k
        uint64_t seed = (uint64_t)prm[0];
        double amax = prm[1];
        bool in = std::fabs(x[1] - 0.5) < 0.25;
        double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        double u2 = 0.0;
        for (int k = 1; k <= 8; ++k) u2 += amax * mode_noise(seed, k) * std::sin(2.0 * M_PI * k * x[0]);
        double E = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2; u[D + 1] = E;
    } else if (ic_case == 6) {
        // Compressible_Euler.h:237-256: shear layer with a Gaussian-modulated
        // v perturbation of amplitude w0 and width sigma about both layers
        double w0 = prm[0], sig = prm[1], y = x[1];
        bool in = (y > 0.25) && (y < 0.75);
        double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        double g = std::exp(-(y - 0.25) * (y - 0.25) / (2.0 * sig * sig)) +
                   std::exp(-(y - 0.75) * (y - 0.75) / (2.0 * sig * sig));
        double u2 = w0 * std::sin(4.0 * M_PI * x[0]) * g;
        double E = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2; u[D + 1] = E;
    } else if (ic_case == 5) {
        // smooth periodic density wave (unit-cube domain) advected by (1, 0.5, 0.25)
        double prod = 1.0;
        for (int d = 0; d < D; ++d) prod *= std::sin(2.0 * M_PI * x[d]);
        double rho = 1.0 + prm[0] * prod, prs = prm[1];
        const double vel[3] = {1.0, 0.5, 0.25};
        double k = 0.0;
        u[0] = rho;
        for (int d = 0; d < D; ++d) { u[d + 1] = rho * vel[d]; k += vel[d] * vel[d]; }
        u[D + 1] = prs / (gamma - 1.0) + 0.5 * rho * k;
    } else if (ic_case == 3) {
        // radial / spherical shock tube centred in the unit-cell domain
        // (2D form of Compressible_Euler.h:336-357 with parameters)
        double r2 = 0.0;
        for (int d = 0; d < D; ++d) r2 += (x[d] - prm[5 + d]) * (x[d] - prm[5 + d]);
        bool in = std::sqrt(r2) <= prm[0];
        u[0] = in ? prm[1] : prm[3];
        u[D + 1] = (in ? prm[2] : prm[4]) / (gamma - 1.0);
    } else {
        double rho = prm[0];
        double vel[3] = {prm[1], prm[2], D > 2 ? prm[3] : 0.0};
        double prs = prm[D == 3 ? 4 : 3];
        double k = 0;
        u[0] = rho;
        for (int d = 0; d < D; ++d) { u[d + 1] = rho * vel[d]; k += vel[d] * vel[d]; }
        u[D + 1] = prs / (gamma - 1.0) + 0.5 * rho * k;
    }
}

extern "C" void orc_project_ic(const orc_tables* t, const orc_hier* h, int lev, int ic_case,
                               const double* prm, double gamma, double* Uw) {
    // AmrDG.h:782-836: sum_m IC(x(xi_m)) * quadmat(n,m) / refMat_phiphi(n,n)
    const OLevel& lv = h->L[lev];
    const int D = t->D, Q = t->Q, Ns = t->Np_s, Ms = t->qMp_s;
    const long ns = lv.nslot;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s) {
        int b, iv[3];
        h->slot_decode(lev, s, &b, iv);
        std::vector<double> uq((size_t)Ms * 5);
        for (int m = 0; m < Ms; ++m) {
            double x[3];
            for (int d = 0; d < D; ++d)
                x[d] = (lv.dx[d] / 2.0) * t->xs[m][d] + (h->plo[d] + (iv[d] + 0.5) * lv.dx[d]);
            ic_state(D, ic_case, prm, gamma, x, lev, iv, &uq[(size_t)m * 5]);
        }
        for (int q = 0; q < Q; ++q)
            for (int n = 0; n < Ns; ++n) {
                double sum = 0.0;
                for (int m = 0; m < Ms; ++m) sum += uq[(size_t)m * 5 + q] * t->quadmat(n, m);
                Uw[((long)q * Ns + n) * ns + s] = sum / t->phiphi(t->mi_s[n], t->mi_s[n]);
            }
    }
}

// ---------------------------------------------------------------------------
// Stages.
extern "C" void orc_predictor(const orc_tables* t, const orc_hier* h, int lev, const double* Uw,
                              double Dt, double gamma, double* Hw_out) {
    // AmrDG.h:1100-1119: set_predictor, then exactly P Picard iterations
    const long ns = h->L[lev].nslot;
    const int Q = t->Q, M = t->qMp_st, D = t->D;
    std::vector<double> H((size_t)Q * M * ns), F((size_t)D * Q * M * ns);
    set_predictor(t, ns, Uw, Hw_out);
    for (int it = 0; it < t->P; ++it) {
        get_H(t, ns, Hw_out, H.data());
        for (int d = 0; d < D; ++d) flux_vol(t, ns, H.data(), F.data() + (size_t)d * Q * M * ns, d, gamma);
        update_H_w(t, ns, Uw, F.data(), Hw_out, h->L[lev].dx, Dt);
    }
}

namespace {
// numflux: AmrDG.tpp:539-650 for one level and direction
void numflux(const orc_tables* t, const orc_hier* h, int lev, int d, const double* Hm, const double* Hp,
             const double* Fm, const double* Fp, const double* DFm, const double* DFp, double* Fnum,
             double* Fic, double* Fif, double Dt) {
    const OLevel& lv = h->L[lev];
    const int D = t->D, Q = t->Q, M = t->qMp_st_bd, Ns = t->Np_s;
    const long ns = lv.nslot, nf = lv.nfslot[d];
    double dvol = 1.0;  // Mesh::get_dvol(lev,d): face area
    for (int dd = 0; dd < D; ++dd)
        if (dd != d) dvol *= lv.dx[dd];
    double jac = (Dt / 2.0) * (dvol / std::pow(2.0, D - 1));
    std::fill(Fic, Fic + (size_t)Q * Ns * nf, 0.0);
    std::fill(Fif, Fif + (size_t)Q * Ns * nf, 0.0);
    const bool multi = h->nlev > 1;
    for (int b = 0; b < (int)lv.p.size(); ++b) {
        const OPatch& p = lv.p[b];
        int lo[3] = {p.lo[0], p.lo[1], D > 2 ? p.lo[2] : 0};
        int hi[3] = {p.hi[0], p.hi[1], D > 2 ? p.hi[2] : 0};
        hi[d] += 1;
        int nface[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
        long total = (long)nface[0] * nface[1] * nface[2];
#pragma omp parallel for schedule(static)
        for (long f = 0; f < total; ++f) {
            int iv[3] = {lo[0] + (int)(f % nface[0]), lo[1] + (int)((f / nface[0]) % nface[1]),
                         lo[2] + (int)(f / ((long)nface[0] * nface[1]))};
            int Lv[3] = {iv[0], iv[1], iv[2]};
            Lv[d] -= 1;
            long fs = h->fslot_of(lev, b, d, iv);
            long sl = h->slot_of(lev, b, Lv), sr = h->slot_of(lev, b, iv);
            for (int m = 0; m < M; ++m)
                for (int q = 0; q < Q; ++q) {
                    long il = ((long)q * M + m) * ns + sl, ir = ((long)q * M + m) * ns + sr;
                    Fnum[((long)q * M + m) * nf + fs] = llf_(Hp[il], Hm[ir], Fp[il], Fm[ir], DFp[il], DFm[ir]);
                }
            if (!multi) continue;
            for (int n = 0; n < Ns; ++n) {
                if (lev < h->nlev - 1) {
                    int bc = lv.bcoarse[d][fs];
                    if (bc != 0) {
                        const Mat& mat = (bc == 1) ? t->Mkbdp[d] : t->Mkbdm[d];
                        for (int q = 0; q < Q; ++q) {
                            double sum = 0.0;
                            for (int m = 0; m < M; ++m) sum += Fnum[((long)q * M + m) * nf + fs] * mat(n, m);
                            Fic[((long)q * Ns + n) * nf + fs] = sum * jac;
                        }
                    }
                }
                if (lev > 0) {
                    int bf = lv.bfine[d][fs];
                    if (bf != 0) {
                        const Mat& Pm = (bf == 1) ? t->Pfh[d][lv.cidx[d][fs]] : t->Pfl[d][lv.cidx[d][fs]];
                        for (int q = 0; q < Q; ++q) {
                            double sum = 0.0;
                            for (int m = 0; m < M; ++m) sum += Pm(n, m) * Fnum[((long)q * M + m) * nf + fs];
                            Fif[((long)q * Ns + n) * nf + fs] = sum * jac;
                        }
                    }
                }
            }
        }
    }
}

// update_U_w: AmrDG.tpp:652-742 (valid cells)
void update_U_w(const orc_tables* t, const orc_hier* h, int lev, double* Uw, const double* F,
                double* const* Fnum, double Dt) {
    const OLevel& lv = h->L[lev];
    const int D = t->D, Q = t->Q, Ns = t->Np_s, M = t->qMp_st, Mb = t->qMp_st_bd;
    const long ns = lv.nslot;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < ns; ++s) {
        if (!lv.valid[s]) continue;
        int b, iv[3];
        h->slot_decode(lev, s, &b, iv);
        for (int q = 0; q < Q; ++q) {
            double rhs[64];
            for (int n = 0; n < Ns; ++n) rhs[n] = 0.0 + (t->Mk_corr(n, n) * Uw[((long)q * Ns + n) * ns + s]);
            for (int d = 0; d < D; ++d) {
                double Sn = (Dt / lv.dx[d]);
                for (int n = 0; n < Ns; ++n)
                    for (int m = 0; m < M; ++m)
                        rhs[n] += Sn * (t->Sk_corr[d](n, m) * F[(((long)d * Q + q) * M + m) * ns + s]);
                double Mb_n = (Dt / lv.dx[d]);
                int hiv[3] = {iv[0], iv[1], iv[2]};
                hiv[d] += 1;
                long nf = lv.nfslot[d];
                long flo = h->fslot_of(lev, b, d, iv), fhi = h->fslot_of(lev, b, d, hiv);
                for (int n = 0; n < Ns; ++n)
                    for (int m = 0; m < Mb; ++m)
                        rhs[n] -= (Mb_n * (t->Mkbdp[d](n, m) * Fnum[d][((long)q * Mb + m) * nf + fhi] -
                                           t->Mkbdm[d](n, m) * Fnum[d][((long)q * Mb + m) * nf + flo]));
            }
            for (int n = 0; n < Ns; ++n) {
                rhs[n] /= t->Mk_corr(n, n);
                Uw[((long)q * Ns + n) * ns + s] = rhs[n];
            }
        }
    }
}

// Flux registers + reflux.  AMReX FluxRegister semantics as used by the
// reference (AmrDG.h:1138-1153 CrseAdd(-1)/FineAdd(+1); AmrDG_Level.tpp:
// 416-446 Reflux with unit volume) followed by L2ProjInterp::reflux
// (AmrDG_Interpolater.tpp:7-72).  Register of fine level lf for direction d
// on each fine box side: reg = sum(fine child moments) - coarse moment.
void flux_correction(const orc_tables* t, const orc_hier* h, double** Uw,
                     std::vector<Scratch>& sc) {
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    for (int lf = h->nlev - 1; lf > 0; --lf) {
        const int lc = lf - 1;
        const OLevel& fl = h->L[lf];
        const OLevel& cl = h->L[lc];
        // registers: key (d, face coarse index vector) -> value.  Stored per
        // fine box / side as in AMReX, then scattered onto coarse face arrays.
        // corrF[d]: per coarse face slot of every coarse patch face array.
        std::vector<std::vector<double>> corrF(D);
        for (int d = 0; d < D; ++d) corrF[d].assign((size_t)Q * Ns * cl.nfslot[d], 0.0);
        for (int B = 0; B < (int)fl.p.size(); ++B) {
            const OPatch& fp = fl.p[B];
            int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) { clo[d] = fp.lo[d] >> 1; chi[d] = fp.hi[d] >> 1; }
            for (int d = 0; d < D; ++d)
                for (int side = 0; side < 2; ++side) {
                    int fidx = side == 0 ? clo[d] : chi[d] + 1;
                    int tlo[3] = {clo[0], clo[1], clo[2]}, thi[3] = {chi[0], chi[1], chi[2]};
                    tlo[d] = thi[d] = fidx;
                    int iv[3];
                    for (iv[2] = tlo[2]; iv[2] <= thi[2]; ++iv[2])
                        for (iv[1] = tlo[1]; iv[1] <= thi[1]; ++iv[1])
                            for (iv[0] = tlo[0]; iv[0] <= thi[0]; ++iv[0]) {
                                double reg[5 * 64];
                                for (int i = 0; i < Q * Ns; ++i) reg[i] = 0.0;
                                // images of this coarse face along periodic d
                                int imgs[2] = {iv[d], -99999};
                                int nimg = 1;
                                if (h->per[d]) {
                                    if (iv[d] == 0) imgs[nimg++] = cl.dom[d];
                                    else if (iv[d] == cl.dom[d]) imgs[nimg++] = 0;
                                }
                                // CrseAdd(-1): every coarse patch face array holding an image
                                for (int ii = 0; ii < nimg; ++ii) {
                                    int fv[3] = {iv[0], iv[1], iv[2]};
                                    fv[d] = imgs[ii];
                                    for (int cb = 0; cb < (int)cl.p.size(); ++cb) {
                                        const OPatch& cp = cl.p[cb];
                                        bool in = true;
                                        for (int dd = 0; dd < D; ++dd) {
                                            int hi_ = cp.hi[dd] + (dd == d ? 1 : 0);
                                            if (fv[dd] < cp.lo[dd] || fv[dd] > hi_) in = false;
                                        }
                                        if (!in) continue;
                                        long fs = h->fslot_of(lc, cb, d, fv);
                                        for (int q = 0; q < Q; ++q)
                                            for (int n = 0; n < Ns; ++n)
                                                reg[q * Ns + n] += -1.0 * sc[lc].Fic[d][((long)q * Ns + n) * cl.nfslot[d] + fs];
                                    }
                                }
                                // FineAdd(+1): children faces in this fine box
                                int nfc = 1 << (D - 1);
                                for (int c = 0; c < nfc; ++c) {
                                    int fv[3] = {2 * iv[0], 2 * iv[1], 2 * iv[2]};
                                    int bp = 0;
                                    for (int dd = 0; dd < D; ++dd) {
                                        if (dd == d) continue;
                                        fv[dd] += (c >> bp) & 1;
                                        bp++;
                                    }
                                    long fs = h->fslot_of(lf, B, d, fv);
                                    for (int q = 0; q < Q; ++q)
                                        for (int n = 0; n < Ns; ++n)
                                            reg[q * Ns + n] += 1.0 * sc[lf].Fif[d][((long)q * Ns + n) * fl.nfslot[d] + fs];
                                }
                                // Reflux: copyTo/plusTo onto coarse face arrays (with images)
                                for (int ii = 0; ii < nimg; ++ii) {
                                    int fv[3] = {iv[0], iv[1], iv[2]};
                                    fv[d] = imgs[ii];
                                    for (int cb = 0; cb < (int)cl.p.size(); ++cb) {
                                        const OPatch& cp = cl.p[cb];
                                        bool in = true;
                                        for (int dd = 0; dd < D; ++dd) {
                                            int hi_ = cp.hi[dd] + (dd == d ? 1 : 0);
                                            if (fv[dd] < cp.lo[dd] || fv[dd] > hi_) in = false;
                                        }
                                        if (!in) continue;
                                        long fs = h->fslot_of(lc, cb, d, fv);
                                        for (int i = 0; i < Q * Ns; ++i)
                                            corrF[d][(size_t)i * cl.nfslot[d] + fs] += reg[i];
                                    }
                                }
                            }
                }
        }
        // correction = -(F(hi) - F(lo)) per direction; apply where the coarse
        // cell touches a b_coarse face of its own patch (ownership quirk kept).
        double vol = 1.0;
        for (int d = 0; d < D; ++d) vol *= cl.dx[d];
        double inv_jac = std::pow(2.0, D) / vol;
        const long ns = cl.nslot;
        double* U = Uw[lc];
        for (int cb = 0; cb < (int)cl.p.size(); ++cb) {
            const OPatch& cp = cl.p[cb];
            int iv[3];
            for (iv[2] = (D > 2 ? cp.lo[2] : 0); iv[2] <= (D > 2 ? cp.hi[2] : 0); ++iv[2])
                for (iv[1] = cp.lo[1]; iv[1] <= cp.hi[1]; ++iv[1])
                    for (iv[0] = cp.lo[0]; iv[0] <= cp.hi[0]; ++iv[0]) {
                        bool at_if = false;
                        double corr[5 * 64];
                        for (int i = 0; i < Q * Ns; ++i) corr[i] = 0.0;
                        for (int d = 0; d < D; ++d) {
                            int hv[3] = {iv[0], iv[1], iv[2]};
                            hv[d] += 1;
                            long flo = h->fslot_of(lc, cb, d, iv), fhi = h->fslot_of(lc, cb, d, hv);
                            if (cl.bcoarse[d][flo] != 0 || cl.bcoarse[d][fhi] != 0) at_if = true;
                            for (int i = 0; i < Q * Ns; ++i)
                                corr[i] += -1.0 * (corrF[d][(size_t)i * cl.nfslot[d] + fhi] -
                                                   corrF[d][(size_t)i * cl.nfslot[d] + flo]) / 1.0;
                        }
                        if (!at_if && !h->conservative_reflux) continue;
                        long s = h->slot_of(lc, cb, iv);
                        for (int q = 0; q < Q; ++q) {
                            double mv[64];
                            for (int n = 0; n < Ns; ++n) {
                                double a = 0.0;
                                for (int k = 0; k < Ns; ++k) a += t->Minv(n, k) * corr[q * Ns + k];
                                mv[n] = a;
                            }
                            for (int n = 0; n < Ns; ++n) U[((long)q * Ns + n) * ns + s] += inv_jac * mv[n];
                        }
                    }
        }
    }
}
}  // namespace

extern "C" void orc_ader(const orc_tables* t, const orc_hier* h, double** Uw, double Dt, double gamma,
                         double* Fnum_out) {
    // AmrDG::ADER, AmrDG.h:1071-1165
    const int D = t->D, Q = t->Q, Nst = t->Np_st, M = t->qMp_st, Mb = t->qMp_st_bd, Ns = t->Np_s;
    std::vector<Scratch> sc(h->nlev);
    for (int l = h->nlev - 1; l >= 0; --l) {
        const OLevel& lv = h->L[l];
        const long ns = lv.nslot;
        Scratch& S = sc[l];
        S.Hw.assign((size_t)Q * Nst * ns, 0.0);
        S.H.assign((size_t)Q * M * ns, 0.0);
        S.F.assign((size_t)D * Q * M * ns, 0.0);
        for (auto* v : {&S.Hm, &S.Hp, &S.Fm, &S.Fp, &S.DFm, &S.DFp}) v->assign((size_t)Q * Mb * ns, 0.0);
        set_predictor(t, ns, Uw[l], S.Hw.data());
        for (int it = 0; it < t->P; ++it) {
            get_H(t, ns, S.Hw.data(), S.H.data());
            for (int d = 0; d < D; ++d) flux_vol(t, ns, S.H.data(), S.F.data() + (size_t)d * Q * M * ns, d, gamma);
            update_H_w(t, ns, Uw[l], S.F.data(), S.Hw.data(), lv.dx, Dt);
        }
        get_H(t, ns, S.Hw.data(), S.H.data());
        for (int d = 0; d < D; ++d) {
            flux_vol(t, ns, S.H.data(), S.F.data() + (size_t)d * Q * M * ns, d, gamma);
            get_H_bd(t, ns, S.Hw.data(), S.Hm.data(), d, 0);
            flux_bd(t, ns, S.Hm.data(), S.Fm.data(), S.DFm.data(), d, gamma);
            get_H_bd(t, ns, S.Hw.data(), S.Hp.data(), d, 1);
            flux_bd(t, ns, S.Hp.data(), S.Fp.data(), S.DFp.data(), d, gamma);
            S.Fnum[d].assign((size_t)Q * Mb * lv.nfslot[d], 0.0);
            S.Fic[d].assign((size_t)Q * Ns * lv.nfslot[d], 0.0);
            S.Fif[d].assign((size_t)Q * Ns * lv.nfslot[d], 0.0);
            numflux(t, h, l, d, S.Hm.data(), S.Hp.data(), S.Fm.data(), S.Fp.data(), S.DFm.data(),
                    S.DFp.data(), S.Fnum[d].data(), S.Fic[d].data(), S.Fif[d].data(), Dt);
        }
        double* fn[3] = {S.Fnum[0].data(), S.Fnum[1].data(), D > 2 ? S.Fnum[2].data() : nullptr};
        update_U_w(t, h, l, Uw[l], S.F.data(), fn, Dt);
        // free the big volume buffers early
        std::vector<double>().swap(S.H);
        std::vector<double>().swap(S.F);
        std::vector<double>().swap(S.Hw);
    }
    if (h->nlev > 1) flux_correction(t, h, Uw, sc);
    if (Fnum_out) {
        size_t at = 0;
        for (int l = 0; l < h->nlev; ++l)
            for (int d = 0; d < D; ++d) {
                std::copy(sc[l].Fnum[d].begin(), sc[l].Fnum[d].end(), Fnum_out + at);
                at += sc[l].Fnum[d].size();
            }
    }
}

extern "C" void orc_average_down(const orc_tables* t, const orc_hier* h, double** Uw) {
    // AmrDG_Level.tpp:227-238 -> L2ProjInterp::average_down / amr_gather
    // (AmrDG_Interpolater.tpp:283-342,366-398,425-448)
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    const int nch = 1 << D;
    for (int lf = h->nlev - 1; lf > 0; --lf) {
        const int lc = lf - 1;
        const OLevel& fl = h->L[lf];
        const OLevel& cl = h->L[lc];
        for (int B = 0; B < (int)fl.p.size(); ++B) {
            const OPatch& fp = fl.p[B];
            int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) { clo[d] = fp.lo[d] >> 1; chi[d] = fp.hi[d] >> 1; }
            int iv[3];
            for (iv[2] = clo[2]; iv[2] <= chi[2]; ++iv[2])
                for (iv[1] = clo[1]; iv[1] <= chi[1]; ++iv[1])
                    for (iv[0] = clo[0]; iv[0] <= chi[0]; ++iv[0]) {
                        long cs = cl.owner[h->cell_lin(lc, iv)];
                        for (int q = 0; q < Q; ++q) {
                            double sum[64];
                            for (int n = 0; n < Ns; ++n) sum[n] = 0.0;
                            for (int l = 0; l < nch; ++l) {
                                int fv[3] = {2 * iv[0], 2 * iv[1], 2 * iv[2]};
                                for (int d = 0; d < D; ++d) fv[d] += (l >> (D - 1 - d)) & 1;
                                long fs = h->slot_of(lf, B, fv);
                                for (int n = 0; n < Ns; ++n) {
                                    double a = 0.0;
                                    for (int k = 0; k < Ns; ++k)
                                        a += t->P_fc[l](n, k) * Uw[lf][((long)q * Ns + k) * fl.nslot + fs];
                                    sum[n] += a;
                                }
                            }
                            for (int n = 0; n < Ns; ++n) {
                                double a = 0.0;
                                for (int k = 0; k < Ns; ++k) a += t->Minv(n, k) * sum[k];
                                Uw[lc][((long)q * Ns + n) * cl.nslot + cs] = (1.0 / nch) * a;
                            }
                        }
                    }
        }
    }
}

extern "C" void orc_fill_patch(const orc_tables* t, const orc_hier* h, double** Uw) {
    // AmrDG.h:1005-1015 -> AMR_FillPatch per level (ghost cells only; sources
    // are valid cells so the level order does not matter)
    const int Q = t->Q, Ns = t->Np_s;
    for (int l = 0; l < h->nlev; ++l) {
        const OLevel& lv = h->L[l];
        const long ns = lv.nslot;
        const long ng = (long)lv.gslot.size();
#pragma omp parallel for schedule(static)
        for (long g = 0; g < ng; ++g) {
            long s = lv.gslot[g], src = lv.gsrc[g];
            if (lv.gkind[g] == 0) {
                for (int q = 0; q < Q; ++q)
                    for (int n = 0; n < Ns; ++n)
                        Uw[l][((long)q * Ns + n) * ns + s] = Uw[l][((long)q * Ns + n) * ns + src];
            } else if (lv.gkind[g] == 1) {
                const long cns = h->L[l - 1].nslot;
                const Mat& Pc = t->P_cf[lv.gchild[g]];
                for (int q = 0; q < Q; ++q) {
                    double tmp[64];
                    for (int j = 0; j < Ns; ++j) {
                        double a = 0.0;
                        for (int k = 0; k < Ns; ++k) a += Pc(j, k) * Uw[l - 1][((long)q * Ns + k) * cns + src];
                        tmp[j] = a;
                    }
                    for (int n = 0; n < Ns; ++n) {
                        double a = 0.0;
                        for (int k = 0; k < Ns; ++k) a += t->Minv(n, k) * tmp[k];
                        Uw[l][((long)q * Ns + n) * ns + s] = a;
                    }
                }
            }
            // transmissive boundary ghost: mirror image, P_k(-xi) = (-1)^k P_k(xi)
            const int mask = lv.gmask[g];
            if (mask && lv.gkind[g] != 2)
                for (int q = 0; q < Q; ++q)
                    for (int n = 0; n < Ns; ++n) {
                        int par = 0;
                        for (int d = 0; d < t->D; ++d)
                            if (mask & (1 << d)) par += t->mi_s[n][d];
                        if (par & 1) Uw[l][((long)q * Ns + n) * ns + s] = -Uw[l][((long)q * Ns + n) * ns + s];
                    }
        }
    }
}

extern "C" double orc_set_dt(const orc_tables* t, const orc_hier* h, double** Uw, double c_dt,
                             double gamma) {
    // AmrDG.h:1253-1357: lambda at spatial node m=0, CFL = 1/((2P+1)D)
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    double dt_min = 1.0;
    std::vector<double> dtl;
    for (int l = 0; l < h->nlev; ++l) {
        const OLevel& lv = h->L[l];
        double dx_avg = 0.0;
        for (int d = 0; d < D; ++d) dx_avg += lv.dx[d];
        dx_avg /= (double)D;
        double phi0[64];
        for (int n = 0; n < Ns; ++n) {
            double phi = 1.0;
            for (int d = 0; d < D; ++d) phi *= t->val[t->mi_s[n][d]][t->node_idx(0, d, D)];
            phi0[n] = phi;
        }
        const long ns = lv.nslot;
        double lvmin = std::numeric_limits<double>::infinity();
        bool any = false;
#pragma omp parallel for reduction(min : lvmin) schedule(static)
        for (long s = 0; s < ns; ++s) {
            if (!lv.valid[s]) continue;
            if (h->cfl_uncovered && l + 1 < h->nlev) {
                int b, iv[3];
                h->slot_decode(l, s, &b, iv);
                if (h->covered_by_finer(l, iv)) continue;
            }
            double u[5];
            for (int q = 0; q < Q; ++q) {
                double a = 0.0;
                for (int n = 0; n < Ns; ++n) a += Uw[l][((long)q * Ns + n) * ns + s] * phi0[n];
                u[q] = a;
            }
            double lmax = orc_cfl_lambda(D, 0, u, gamma);
            for (int d = 1; d < D; ++d) lmax = std::max(lmax, orc_cfl_lambda(D, d, u, gamma));
            double CFL = (1.0 / (2.0 * (double)t->P + 1.0)) * (1.0 / (double)D);
            double dt = CFL * (dx_avg / lmax);
            lvmin = std::min(lvmin, dt);
        }
        for (long s = 0; s < ns && !any; ++s) any = lv.valid[s] != 0;
        dtl.push_back(any ? lvmin : 1.0);
    }
    if (!dtl.empty()) dt_min = *std::min_element(dtl.begin(), dtl.end());
    return c_dt * dt_min;
}

extern "C" long orc_limiter_tvb(const orc_tables* t, const orc_hier* h, double** Uw, int l, double M,
                                const double* tvb_c, double gamma);

extern "C" int orc_evolve(const orc_tables* t, const orc_hier* h, double** Uw, double* t_io,
                          double* Dt_io, double T, double c_dt, double gamma, int nsteps_max) {
    // AmrDG::evolve (AmrDG.h:853-1055) without regrid, limiter and plots
    int n = 0;
    double tt = *t_io, Dt = *Dt_io;
    while (tt < T && (nsteps_max <= 0 || n < nsteps_max)) {
        orc_ader(t, h, Uw, Dt, gamma, nullptr);
        orc_average_down(t, h, Uw);
        orc_fill_patch(t, h, Uw);
        n += 1;
        tt += Dt;
        Dt = orc_set_dt(t, h, Uw, c_dt, gamma);
        if (T - tt < Dt) Dt = T - tt;
    }
    *t_io = tt;
    *Dt_io = Dt;
    return n;
}

// NewtonRhapson (AmrDG_Quadrature.tpp:249-268): Newton on P_n with the
// derivative from assoc_legendre(n,1,x)/sqrt(1-x^2), step tolerance 1e-15.
static void orc_newton_rhapson(double& x, int n) {
    for (int it = 0; it < 1000; ++it) {
        double df = (std::assoc_legendre(n, 1, x)) / (std::sqrt(1.0 - std::pow(x, 2.0)));
        double f = std::legendre(n, x);
        double x_new = x - (f / df);
        double err = std::abs(x_new - x);
        x = x_new;
        if (err <= 1e-15) break;
    }
}

// The 2(P+1)-point rule of L2Norm_DG_AMR (AmrDG.h:1587-1655): asymptotic
// starting values, Newton, order [x1,-x1,x2,-x2,...]; points of the D-fold
// product with dimension 0 slowest.
static std::vector<std::vector<double>> orc_l2_points(int N, int D) {
    std::vector<double> gl;
    for (int i = 1; i <= (int)(N / 2); ++i) {
        double theta = kPi * (i - 0.25) / ((double)N + 0.5), xiq;
        if ((1 <= i) && (i <= (int)((1.0 / 3.0) * (double)N))) {
            xiq = (1 - 0.125 * (1.0 / std::pow(N, 2)) + 0.125 * (1.0 / std::pow(N, 3)) -
                   (1.0 / 384.0) * (1.0 / std::pow(N, 4)) * (39.0 - 28.0 * (1.0 / std::pow(std::sin(theta), 2)))) *
                  std::cos(theta);
        } else {
            xiq = (1.0 - (1.0 / (8.0 * std::pow((double)N, 2))) + (1.0 / (8.0 * std::pow((double)N, 3)))) *
                  std::cos(theta);
        }
        orc_newton_rhapson(xiq, N);
        gl.push_back(xiq);
        gl.push_back(-xiq);
    }
    const int M = ipow(N, D);
    std::vector<std::vector<double>> pts(M, std::vector<double>(D));
    for (int m = 0; m < M; ++m) {
        int div = M;
        for (int d = 0; d < D; ++d) {
            div /= N;
            pts[m][d] = gl[(m / div) % N];
        }
    }
    return pts;
}

extern "C" int orc_l2_nodes(int P, double* out) {
    // the 1D node set of the L2 rule (2(P+1) values), for tests
    const int N = 2 * (P + 1);
    std::vector<std::vector<double>> pts = orc_l2_points(N, 1);
    for (int i = 0; i < N; ++i) out[i] = pts[i][0];
    return N;
}

extern "C" int orc_evolve_lim(const orc_tables* t, const orc_hier* h, double** Uw, double* t_io,
                              double* Dt_io, double T, double c_dt, double gamma, int nsteps_max, double M,
                              const double* tvb_c, int t_limit, long* n_limited) {
    // AmrDG::evolve (AmrDG.h:930-1030) without regrid and plots, with the
    // TVB limiter on the finest level after the ADER step every t_limit
    // steps (AmrDG.h:990-997), before the average down
    int n = 0;
    long lim = 0;
    double tt = *t_io, Dt = *Dt_io;
    while (tt < T && (nsteps_max <= 0 || n < nsteps_max)) {
        orc_ader(t, h, Uw, Dt, gamma, nullptr);
        if (t_limit > 0 && (n % t_limit == 0)) lim += orc_limiter_tvb(t, h, Uw, h->nlev - 1, M, tvb_c, gamma);
        orc_average_down(t, h, Uw);
        orc_fill_patch(t, h, Uw);
        n += 1;
        tt += Dt;
        Dt = orc_set_dt(t, h, Uw, c_dt, gamma);
        if (T - tt < Dt) Dt = T - tt;
    }
    *t_io = tt;
    *Dt_io = Dt;
    if (n_limited) *n_limited = lim;
    return n;
}

extern "C" double orc_lp_error(const orc_tables* t, const orc_hier* h, double** Uw, int qq, int p,
                               int ic_case, const double* prm, double gamma, double time) {
    // LpNorm_DG_AMR (AmrDG.h:1399-1578): L1Norm_DG_AMR on the (P+1)^D
    // solution points, L2Norm_DG_AMR on the (2(P+1))^D points of its own rule
    // (AmrDG.h:1580-1655); weights 2/assoc_legendre(N,1,x)^2 per dimension.
    // The exact solution is the IC advected by the background velocity
    // (periodic shift), which reduces to the reference's IC comparison at time 0.
    const int D = t->D, Ns = t->Np_s;
    const int N = p == 2 ? 2 * t->N : t->N;
    const std::vector<std::vector<double>> pts = p == 2 ? orc_l2_points(N, D) : t->xs;
    const int Ms = (int)pts.size();
    double total = 0.0, Vamr = 0.0;
    for (int l = 0; l < h->nlev; ++l) {
        const OLevel& lv = h->L[l];
        double vol = 1.0;
        for (int d = 0; d < D; ++d) vol *= lv.dx[d];
        const long ns = lv.nslot;
        double acc = 0.0;
        long ncell = 0;
#pragma omp parallel for reduction(+ : acc, ncell) schedule(static)
        for (long s = 0; s < ns; ++s) {
            if (!lv.valid[s]) continue;
            int b, iv[3];
            h->slot_decode(l, s, &b, iv);
            if (h->covered_by_finer(l, iv)) continue;
            ncell++;
            double cell = 0.0;
            for (int m = 0; m < Ms; ++m) {
                double w = 1.0;
                for (int d = 0; d < D; ++d) w *= 2.0 / std::pow(std::assoc_legendre(N, 1, pts[m][d]), 2);
                double uh = 0.0;
                for (int n = 0; n < Ns; ++n) {
                    double phi = 1.0;
                    for (int d = 0; d < D; ++d) phi *= cx_legendre(t->mi_s[n][d], pts[m][d]);
                    uh += Uw[l][((long)qq * Ns + n) * ns + s] * phi;
                }
                double x[3], u[5];
                for (int d = 0; d < D; ++d) {
                    x[d] = (lv.dx[d] / 2.0) * pts[m][d] + (h->plo[d] + (iv[d] + 0.5) * lv.dx[d]);
                    if ((ic_case == 0 && d < 2) || ic_case == 5) {
                        // translate by the background velocity on the periodic domain
                        const double vel5[3] = {1.0, 0.5, 0.25};
                        double L = h->phi[d] - h->plo[d];
                        double v = ic_case == 0 ? prm[3 + d] : vel5[d];
                        double sh = x[d] - v * time - h->plo[d];
                        sh = sh - L * std::floor(sh / L);
                        x[d] = h->plo[d] + sh;
                    }
                }
                ic_state(D, ic_case, prm, gamma, x, l, iv, u);
                cell += std::pow(std::fabs(u[qq] - uh), (double)p) * w;
            }
            acc += cell * (vol / std::pow(2.0, D));
        }
        total += acc;
        Vamr += vol * (double)ncell;
    }
    return std::pow(total / Vamr, 1.0 / (double)p);
}

// ---------------------------------------------------------------------------
// TVB limiter: Limiter_w (AmrDG.h:1657-1770), Limiter_linear_tvb
// (AmrDG.h:1772-1855), minmodB / minmod (AmrDG_Limiter.tpp:41-84) and the
// characteristic matrices pde_EV_Lmatrix / pde_EV_Rmatrix
// (Compressible_Euler.cpp:138-277), evaluated at U_w mode 0 (the arrays the
// limiter passes are the modal ones, component m = 0).

// pde_EV_Rmatrix (Compressible_Euler.cpp:138-205), R[row][col]
static void ev_right(int D, int d, const double* u, double gamma, double R[5][5]) {
    const double c = soundspeed_(D, u, gamma);
    const double u1 = u[1] / u[0], u2 = u[2] / u[0], u3 = D == 3 ? u[3] / u[0] : 0.0;
    const double ek = D == 2 ? 0.5 * (std::pow(u1, 2) + std::pow(u2, 2))
                             : 0.5 * (std::pow(u1, 2) + std::pow(u2, 2) + std::pow(u3, 2));
    const double hh = (std::pow(c, 2) / (gamma - 1.0)) + ek;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) R[i][j] = 0.0;
    if (D == 2) {
        if (d == 0) {
            const double r[4][4] = {{1.0, 1.0, 1.0, 0.0}, {u1 - c, u1, u1 + c, 0.0}, {u2, u2, u2, -1.0},
                                    {hh - c * u1, ek, hh + c * u1, -u2}};
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) R[i][j] = r[i][j];
        } else {
            const double r[4][4] = {{1.0, 1.0, 1.0, 0.0}, {u1, u1, u1, 1.0}, {u2 - c, u2, u2 + c, 0.0},
                                    {hh - c * u2, ek, hh + c * u2, u1}};
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) R[i][j] = r[i][j];
        }
    } else {
        if (d == 0) {
            const double r[5][5] = {{1.0, 1.0, 1.0, 0.0, 0.0}, {u1 - c, u1, u1 + c, 0.0, 0.0},
                                    {u2, u2, u2, -1.0, 0.0}, {u3, u3, u3, 0.0, 1.0},
                                    {hh - c * u1, ek, hh + c * u1, -u2, u3}};
            for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) R[i][j] = r[i][j];
        } else if (d == 1) {
            const double r[5][5] = {{1.0, 1.0, 1.0, 0.0, 0.0}, {u1, u1, u1, 1.0, 0.0},
                                    {u2 - c, u2, u2 + c, 0.0, 0.0}, {u3, u3, u3, 0.0, -1.0},
                                    {hh - c * u2, ek, hh + c * u2, u1, -u3}};
            for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) R[i][j] = r[i][j];
        } else {
            const double r[5][5] = {{1.0, 1.0, 1.0, 0.0, 0.0}, {u1, u1, u1, -1.0, 0.0},
                                    {u2, u2, u2, 0.0, 1.0}, {u3 - c, u3, u3 + c, 0.0, 0.0},
                                    {hh - c * u3, ek, hh + c * u3, -u1, u2}};
            for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) R[i][j] = r[i][j];
        }
    }
}

// pde_EV_Lmatrix (Compressible_Euler.cpp:207-277), L[row][col]
static void ev_left(int D, int d, const double* u, double gamma, double L[5][5]) {
    const double c = soundspeed_(D, u, gamma);
    const double u1 = u[1] / u[0], u2 = u[2] / u[0], u3 = D == 3 ? u[3] / u[0] : 0.0;
    const double ek = D == 2 ? 0.5 * (std::pow(u1, 2) + std::pow(u2, 2))
                             : 0.5 * (std::pow(u1, 2) + std::pow(u2, 2) + std::pow(u3, 2));
    const double b = 1.0 / (2.0 * std::pow(c, 2));
    const double g = gamma - 1.0;
    const double ph = g * ek;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) L[i][j] = 0.0;
    if (D == 2) {
        if (d == 0) {
            const double l[4][4] = {{b * (ph + c * u1), -b * (g * u1 + c), -b * g * u2, b * g},
                                    {1.0 - 2.0 * b * ph, 2.0 * b * g * u1, 2.0 * b * g * u2, -2.0 * b * g},
                                    {b * (ph - c * u1), -b * (g * u1 - c), -b * (g * u2), b * g},
                                    {u2, 0.0, -1.0, 0.0}};
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) L[i][j] = l[i][j];
        } else {
            const double l[4][4] = {{b * (ph + c * u2), -b * g * u1, -b * (g * u2 + c), b * g},
                                    {1.0 - 2.0 * b * ph, 2.0 * b * g * u1, 2.0 * b * g * u2, -2.0 * b * g},
                                    {b * (ph - c * u2), -b * (g * u1), -b * (g * u2 - c), b * g},
                                    {-u1, 1.0, 0.0, 0.0}};
            for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) L[i][j] = l[i][j];
        }
    } else {
        if (d == 0) {
            const double l[5][5] = {
                {b * (ph + c * u1), -b * (g * u1 + c), -b * g * u2, -b * g * u3, b * g},
                {1.0 - 2.0 * b * ph, 2.0 * b * g * u1, 2.0 * b * g * u2, 2.0 * b * g * u3, -2.0 * b * g},
                {b * (ph - c * u1), -b * (g * u1 - c), -b * (g * u2), -b * g * u3, b * g},
                {u2, 0.0, -1.0, 0.0, 0.0},
                {-u3, 0.0, 0.0, 1.0, 0.0}};
            for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) L[i][j] = l[i][j];
        } else if (d == 1) {
            const double l[5][5] = {
                {b * (ph + c * u2), -b * g * u1, -b * (g * u2 + c), -b * g * u3, b * g},
                {1.0 - 2.0 * b * ph, 2.0 * b * g * u1, 2.0 * b * g * u2, 2.0 * b * g * u3, -2.0 * b * g},
                {b * (ph - c * u2), -b * (g * u1), -b * (g * u2 - c), -b * g * u3, b * g},
                {-u1, 1.0, 0.0, 0.0, 0.0},
                {u3, 0.0, 0.0, -1.0, 0.0}};
            for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) L[i][j] = l[i][j];
        } else {
            const double l[5][5] = {
                {b * (ph + c * u3), -b * g * u1, -b * g * u2, -b * (g * u3 + c), b * g},
                {1.0 - 2.0 * b * ph, 2.0 * b * g * u1, 2.0 * b * g * u2, 2.0 * b * g * u3, -2.0 * b * g},
                {b * (ph - c * u3), -b * (g * u1), -b * g * u2, -b * (g * u3 - c), b * g},
                {u1, -1.0, 0.0, 0.0, 0.0},
                {-u2, 0.0, 1.0, 0.0, 0.0}};
            for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j) L[i][j] = l[i][j];
        }
    }
}

// minmod (AmrDG_Limiter.tpp:62-84): the smallest magnitude with the common
// sign (signbit), or 0 when the three disagree
static double minmod_(double a1, double a2, double a3) {
    const bool same = (std::signbit(a1) == std::signbit(a2)) && (std::signbit(a2) == std::signbit(a3));
    const int sign = std::signbit(a1) == std::signbit(-1) ? -1 : +1;
    if (same) return sign * std::min({std::abs(a1), std::abs(a2), std::abs(a3)});
    return 0;
}
// minmodB (AmrDG_Limiter.tpp:41-60): a1 kept when |a1| <= C M h^2
static double minmodB_(double a1, double a2, double a3, double C, double M, double h, bool* troubled) {
    if (std::abs(a1) <= C * M * std::pow(h, 2.0)) {
        *troubled = false;
        return a1;
    }
    *troubled = true;
    return minmod_(a1, a2, a3);
}

extern "C" void orc_fill_boundary(const orc_tables* t, const orc_hier* h, double** Uw, int l) {
    // MultiFab::FillBoundary(periodicity) of one level: the ghost cells a
    // valid cell of the same level (directly or through a periodic image)
    // covers; coarse-fine and physical-boundary ghosts keep their values
    const int Q = t->Q, Ns = t->Np_s;
    const OLevel& lv = h->L[l];
    const long ns = lv.nslot;
    for (size_t g = 0; g < lv.gslot.size(); ++g) {
        if (lv.gkind[g] != 0 || lv.gmask[g] != 0) continue;
        const long s = lv.gslot[g], src = lv.gsrc[g];
        for (int q = 0; q < Q; ++q)
            for (int n = 0; n < Ns; ++n) Uw[l][((long)q * Ns + n) * ns + s] = Uw[l][((long)q * Ns + n) * ns + src];
    }
}

extern "C" long orc_limiter_tvb(const orc_tables* t, const orc_hier* h, double** Uw, int l, double M,
                                const double* tvb_c, double gamma) {
    // Limiter_w on level l: FillBoundary, Limiter_linear_tvb on every valid
    // cell in place (it writes modes n >= 1 only and reads the neighbours'
    // mode 0, so the cell order does not matter), FillBoundary; returns the
    // number of cells limited.
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    const OLevel& lv = h->L[l];
    const long ns = lv.nslot;
    orc_fill_boundary(t, h, Uw, l);
    double hmin = lv.dx[0];
    for (int d = 1; d < D; ++d) hmin = std::min(hmin, lv.dx[d]);
    int lin_mode_idx[3] = {0, 0, 0};   // AmrDG.tpp:174-184
    for (int d = 0; d < D; ++d)
        for (int n = 0; n < Ns; ++n) {
            bool lin = t->mi_s[n][d] == 1, others = true;
            for (int dd = 0; dd < D; ++dd)
                if (dd != d && t->mi_s[n][dd] != 0) others = false;
            if (lin && others) { lin_mode_idx[d] = n; break; }
        }
    auto U = [&](int q, int n, long s) -> double& { return Uw[l][((long)q * Ns + n) * ns + s]; };
    long count = 0;
    for (long s = 0; s < ns; ++s) {
        if (!lv.valid[s]) continue;
        int b, iv[3];
        h->slot_decode(l, s, &b, iv);
        double res_limit[3][5] = {};
        double vw[5][3];
        bool troubled = false;
        for (int li = 0; li < D; ++li) {
            const int sm = lin_mode_idx[li];
            int sh[3] = {0, 0, 0}, dd = 0;
            for (int d = 0; d < D; ++d)
                if (t->mi_s[sm][d] == 1) { sh[d] = 1; dd = d; break; }
            int ivm[3] = {iv[0] - sh[0], iv[1] - sh[1], iv[2] - sh[2]};
            int ivp[3] = {iv[0] + sh[0], iv[1] + sh[1], iv[2] + sh[2]};
            const long sm1 = h->slot_of(l, b, ivm), sp1 = h->slot_of(l, b, ivp);
            double u0[5];
            for (int q = 0; q < Q; ++q) u0[q] = U(q, 0, s);
            double Lm[5][5], Rm[5][5];
            ev_left(D, dd, u0, gamma, Lm);
            for (int q = 0; q < Q; ++q) {
                double Dm = 0.0, Dp = 0.0, Du = 0.0;
                for (int r = 0; r < Q; ++r) {
                    Dm += 0.5 * Lm[q][r] * (U(r, 0, s) - U(r, 0, sm1));
                    Dp += 0.5 * Lm[q][r] * (U(r, 0, sp1) - U(r, 0, s));
                    Du += Lm[q][r] * U(r, sm, s);
                }
                bool tf = false;
                vw[q][li] = minmodB_(Du, Dm, Dp, tvb_c[l], M, hmin, &tf);
                troubled = troubled || (vw[q][li] != Du);
            }
            ev_right(D, dd, u0, gamma, Rm);
            for (int q = 0; q < Q; ++q) {
                double sum = 0.0;
                for (int r = 0; r < Q; ++r) sum += Rm[q][r] * vw[r][li];
                res_limit[li][q] = sum;
            }
        }
        if (troubled) {
            for (int q = 0; q < Q; ++q) {
                for (int n = 1; n < Ns; ++n) U(q, n, s) = 0.0;
                for (int li = 0; li < D; ++li) U(q, lin_mode_idx[li], s) = res_limit[li][q];
            }
            ++count;
        }
    }
    orc_fill_boundary(t, h, Uw, l);
    return count;
}

// AMR_tag_cell_refinement (AmrDG_TagCellRef.tpp:4-52): FillBoundary, then
// pde_tag_cell_refinement per valid cell (Compressible_Euler.h:605-682);
// one byte per valid cell of level l, patch by patch, x fastest.
//   crit 0 isentropic_vortex: r <= amr_c^2 about (prm[0], prm[1]), periodic distance
//   crit 1 kelvin_helmolz_instability: minmodB's troubled test on every
//          characteristic slope, threshold tvb_c * amr_c * h^2
//   crit 2 density gradient (the build's own criterion, no reference
//          counterpart): a face neighbour's cell-mean density differs by more
//          than amr_c times the cell's own
extern "C" long orc_tag_cells(const orc_tables* t, const orc_hier* h, double** Uw, int l, int crit, double amr_c,
                              double tvb_c, const double* prm, double gamma, unsigned char* out) {
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    const OLevel& lv = h->L[l];
    const long ns = lv.nslot;
    orc_fill_boundary(t, h, Uw, l);
    double hmin = lv.dx[0];
    for (int d = 1; d < D; ++d) hmin = std::min(hmin, lv.dx[d]);
    int lmi[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d)
        for (int n = 0; n < Ns; ++n) {
            bool lin = t->mi_s[n][d] == 1, others = true;
            for (int dd = 0; dd < D; ++dd)
                if (dd != d && t->mi_s[n][dd] != 0) others = false;
            if (lin && others) { lmi[d] = n; break; }
        }
    auto U = [&](int q, int n, long s) { return Uw[l][((long)q * Ns + n) * ns + s]; };
    long k = 0, ntag = 0;
    for (int b = 0; b < (int)lv.p.size(); ++b) {
        const OPatch& p = lv.p[b];
        int iv[3] = {0, 0, 0};
        for (iv[2] = (D > 2 ? p.lo[2] : 0); iv[2] <= (D > 2 ? p.hi[2] : 0); ++iv[2])
            for (iv[1] = p.lo[1]; iv[1] <= p.hi[1]; ++iv[1])
                for (iv[0] = p.lo[0]; iv[0] <= p.hi[0]; ++iv[0]) {
                    const long s = h->slot_of(l, b, iv);
                    bool tag = false;
                    if (crit == 0) {
                        const double Lx = h->phi[0] - h->plo[0], Ly = h->phi[1] - h->plo[1];
                        const double x = h->plo[0] + (iv[0] + 0.5) * lv.dx[0];
                        const double y = h->plo[1] + (iv[1] + 0.5) * lv.dx[1];
                        double dxv = x - prm[0], dyv = y - prm[1];
                        if (dxv > Lx / 2.0) dxv -= Lx;
                        if (dxv < -Lx / 2.0) dxv += Lx;
                        if (dyv > Ly / 2.0) dyv -= Ly;
                        if (dyv < -Ly / 2.0) dyv += Ly;
                        const double r = std::sqrt(dxv * dxv + dyv * dyv);
                        tag = r <= std::pow(amr_c, 2);
                    } else if (crit == 1) {
                        double u0[5];
                        for (int q = 0; q < Q; ++q) u0[q] = U(q, 0, s);
                        for (int d = 0; d < D && !tag; ++d) {
                            int ivm[3] = {iv[0], iv[1], iv[2]}, ivp[3] = {iv[0], iv[1], iv[2]};
                            ivm[d] -= 1;
                            ivp[d] += 1;
                            const long sm = h->slot_of(l, b, ivm), sp = h->slot_of(l, b, ivp);
                            double Lm[5][5], Rm[5][5];
                            ev_left(D, d, u0, gamma, Lm);
                            (void)Rm;
                            for (int q = 0; q < Q && !tag; ++q) {
                                double Du = 0.0, Dm = 0.0, Dp = 0.0;
                                for (int r = 0; r < Q; ++r) {
                                    Du += Lm[q][r] * U(r, lmi[d], s);
                                    Dm += 0.5 * Lm[q][r] * (U(r, 0, s) - U(r, 0, sm));
                                    Dp += 0.5 * Lm[q][r] * (U(r, 0, sp) - U(r, 0, s));
                                }
                                bool troubled = false;
                                minmodB_(Du, Dm, Dp, tvb_c, amr_c, hmin, &troubled);
                                if (troubled) tag = true;
                            }
                        }
                    } else {
                        const double r0 = U(0, 0, s);
                        for (int d = 0; d < D; ++d) {
                            int ivm[3] = {iv[0], iv[1], iv[2]}, ivp[3] = {iv[0], iv[1], iv[2]};
                            ivm[d] -= 1;
                            ivp[d] += 1;
                            const double j = std::max(std::fabs(U(0, 0, h->slot_of(l, b, ivp)) - r0),
                                                      std::fabs(r0 - U(0, 0, h->slot_of(l, b, ivm))));
                            if (j > amr_c * r0) tag = true;
                        }
                    }
                    out[k++] = tag ? 1 : 0;
                    ntag += tag;
                }
    }
    return ntag;
}

// RemakeLevel / MakeNewLevelFromCoarse (AmrDG_Level.tpp:52-80,120-170) for a
// new layout: level 0 keeps its boxes and is copied; every slot (valid and
// ghost) of a new level l >= 1 takes, through FillPatchTwoLevels /
// InterpFromCoarseLevel with L2ProjInterp (AmrDG_Level.tpp:126-224,
// AmrDG_Interpolater.tpp:251-281,344-364), the old level-l valid cell at its
// (periodic) place, else the projection of its parent in the new level l-1;
// outside a non-periodic boundary the mirror image of the cell inside.
// Levels are remade coarse to fine.  Returns 0, or -1 when a new cell has
// no parent (a layout that does not nest).
extern "C" int orc_remake(const orc_tables* t, const orc_hier* ho, double** Uo, const orc_hier* hn, double** Un) {
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    if (ho->L[0].nslot != hn->L[0].nslot) return -1;
    {
        const long ns = hn->L[0].nslot;
        for (long i = 0; i < (long)Q * Ns * ns; ++i) Un[0][i] = Uo[0][i];
    }
    for (int l = 1; l < hn->nlev; ++l) {
        const OLevel& lv = hn->L[l];
        const long ns = lv.nslot, cns = hn->L[l - 1].nslot;
        const bool had = l < ho->nlev;
        const long ons = had ? ho->L[l].nslot : 0;
        for (long s = 0; s < ns; ++s) {
            int b, iv[3];
            hn->slot_decode(l, s, &b, iv);
            int mv[3] = {iv[0], iv[1], iv[2]}, mask = 0;
            for (int d = 0; d < D; ++d) {
                const int n = lv.dom[d];
                if (hn->per[d]) continue;
                if (iv[d] < 0) { mv[d] = -1 - iv[d]; mask |= 1 << d; }
                if (iv[d] >= n) { mv[d] = 2 * n - 1 - iv[d]; mask |= 1 << d; }
            }
            const long os = had ? ho->find_valid(l, mv) : -1;
            double v[5][64];
            if (os >= 0) {
                for (int q = 0; q < Q; ++q)
                    for (int n = 0; n < Ns; ++n) v[q][n] = Uo[l][((long)q * Ns + n) * ons + os];
            } else {
                int civ[3] = {0, 0, 0}, child = 0;
                for (int d = 0; d < D; ++d) {
                    const int c = (mv[d] >= 0) ? mv[d] / 2 : -((-mv[d] + 1) / 2);
                    civ[d] = c;
                    child += (mv[d] - c * 2) * (1 << (D - 1 - d));
                }
                const long cs = hn->find_valid(l - 1, civ);
                if (cs < 0) return -1;
                const Mat& Pc = t->P_cf[child];
                for (int q = 0; q < Q; ++q) {
                    double tmp[64];
                    for (int j = 0; j < Ns; ++j) {
                        double a = 0.0;
                        for (int k = 0; k < Ns; ++k) a += Pc(j, k) * Un[l - 1][((long)q * Ns + k) * cns + cs];
                        tmp[j] = a;
                    }
                    for (int n = 0; n < Ns; ++n) {
                        double a = 0.0;
                        for (int k = 0; k < Ns; ++k) a += t->Minv(n, k) * tmp[k];
                        v[q][n] = a;
                    }
                }
            }
            for (int q = 0; q < Q; ++q)
                for (int n = 0; n < Ns; ++n) {
                    int par = 0;
                    for (int d = 0; d < D; ++d)
                        if (mask & (1 << d)) par += t->mi_s[n][d];
                    Un[l][((long)q * Ns + n) * ns + s] = (par & 1) ? -v[q][n] : v[q][n];
                }
        }
    }
    return 0;
}

extern "C" void orc_ev_matrices(int D, int d, const double* u, double gamma, double* Lout, double* Rout) {
    // the characteristic matrices at state u, row-major Q x Q (for tests)
    double Lm[5][5], Rm[5][5];
    ev_left(D, d, u, gamma, Lm);
    ev_right(D, d, u, gamma, Rm);
    const int Q = D + 2;
    for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j) {
            Lout[i * Q + j] = Lm[i][j];
            Rout[i * Q + j] = Rm[i][j];
        }
}

extern "C" void orc_totals(const orc_tables* t, const orc_hier* h, double** Uw, double* out) {
    const int D = t->D, Q = t->Q, Ns = t->Np_s;
    for (int q = 0; q < Q; ++q) out[q] = 0.0;
    for (int l = 0; l < h->nlev; ++l) {
        const OLevel& lv = h->L[l];
        double vol = 1.0;
        for (int d = 0; d < D; ++d) vol *= lv.dx[d];
        for (long s = 0; s < lv.nslot; ++s) {
            if (!lv.valid[s]) continue;
            int b, iv[3];
            h->slot_decode(l, s, &b, iv);
            if (h->covered_by_finer(l, iv)) continue;
            for (int q = 0; q < Q; ++q) out[q] += vol * Uw[l][((long)q * Ns + 0) * lv.nslot + s];
        }
    }
}

// ---------------------------------------------------------------------------
// Learned warm start: predictor with a chosen iteration count, and the
// feature / target definition of the surrogate (no counterpart in the
// reference; the product's collection and inference kernels compute the same
// quantities).
extern "C" void orc_predictor_iters(const orc_tables* t, const orc_hier* h, int lev, const double* Uw,
                                    double Dt, double gamma, int iters, double* Hw_out) {
    const long ns = h->L[lev].nslot;
    const int Q = t->Q, M = t->qMp_st, D = t->D;
    std::vector<double> H((size_t)Q * M * ns), F((size_t)D * Q * M * ns);
    set_predictor(t, ns, Uw, Hw_out);
    for (int it = 0; it < iters; ++it) {
        get_H(t, ns, Hw_out, H.data());
        for (int d = 0; d < D; ++d) flux_vol(t, ns, H.data(), F.data() + (size_t)d * Q * M * ns, d, gamma);
        update_H_w(t, ns, Uw, F.data(), Hw_out, h->L[lev].dx, Dt);
    }
}

extern "C" void orc_mlp_dims(const orc_tables* t, int* out) {
    const int nin = t->Q * t->Np_s + 1 + t->D, nout = t->Q * t->Np_st;
    out[0] = nin;
    out[1] = 16 * ((nin + 15) / 16);
    out[2] = 128;
    out[3] = nout;
    out[4] = 16 * ((nout + 15) / 16);
}

// x = [U[q][k] / rho0 (q-major), nu, tau_0..tau_{D-1}], nu the local CFL
// number sum_d tau_d (|u_d| + c) of the cell-average state.
extern "C" void orc_mlp_features(const orc_tables* t, const double* U, const double* tau, double gamma,
                                 double* x) {
    const int Q = t->Q, Ns = t->Np_s, D = t->D;
    const double rho0 = U[0];
    for (int q = 0; q < Q; ++q)
        for (int k = 0; k < Ns; ++k) x[q * Ns + k] = U[q * Ns + k] / rho0;
    double ke = 0.0;
    for (int d = 0; d < D; ++d) ke += U[(d + 1) * Ns] * U[(d + 1) * Ns];
    const double prs = (gamma - 1.0) * (U[(D + 1) * Ns] - 0.5 * ke / rho0);
    const double c = std::sqrt(gamma * prs / rho0);
    double nu = 0.0;
    for (int d = 0; d < D; ++d) nu += tau[d] * (std::fabs(U[(d + 1) * Ns] / rho0) + c);
    x[Q * Ns] = nu;
    for (int d = 0; d < D; ++d) x[Q * Ns + 1 + d] = tau[d];
}

// y = (H - A0 U) / (rho0 max_d tau_d), A0 = Mk_h_w^-1 Mk_pred, q-major.
extern "C" void orc_mlp_target(const orc_tables* t, const double* U, const double* Hst, const double* tau,
                               double* y) {
    const int Q = t->Q, Ns = t->Np_s, Nst = t->Np_st, D = t->D;
    double tmax = 0.0;
    for (int d = 0; d < D; ++d) tmax = std::max(tmax, tau[d]);
    const double sc = 1.0 / (U[0] * tmax);
    for (int q = 0; q < Q; ++q)
        for (int n = 0; n < Nst; ++n) {
            double c0 = 0.0;
            for (int j = 0; j < Nst; ++j) {
                double r = 0.0;
                for (int m = 0; m < Ns; ++m) r += t->Mk_pred(j, m) * U[q * Ns + m];
                c0 += t->Mk_h_w_inv(n, j) * r;
            }
            y[q * Nst + n] = (Hst[q * Nst + n] - c0) * sc;
        }
}
