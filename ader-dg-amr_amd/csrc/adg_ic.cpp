// adg_ic.cpp -- initial conditions: L2 projection of a pointwise state onto
// the modal basis (AmrDG::set_initial_condition / set_initial_condition_U_w,
// NumericalMethod/AmrDG/AmrDG.h:782-836), evaluated on the host.
//
// Pointwise states:
//  0  isentropic vortex of Compressible_Euler::pde_IC
//     (Model/CompressibleEuler/Compressible_Euler.h:148-187) with the centre,
//     strength and background velocity as parameters;
//  1  Kelvin-Helmholtz shear layer rho=2/u=+0.5 for |y-0.5|<0.25 else
//     rho=1/u=-0.5, p=2.5, v = amp sin(4 pi x) + noise_amp*(2r-1) with r a
//     seeded per-cell uniform number (no counterpart in the reference);
//  2  uniform state;
//  3  radial / spherical shock tube centred in the domain (2D: the
//     reference's radial_shock_tube, Compressible_Euler.h:336-357, with the
//     radius and the two states as parameters; 3D has no reference IC);
//  5  smooth periodic density wave rho = 1 + amp prod_d sin(2 pi x_d / L_d),
//     velocity (1, 0.5, 0.25), pressure p0 (params: amp, p0; no counterpart
//     in the reference);
//  4  Kelvin-Helmholtz shear layer as in 1 with the seeded multi-mode
//     perturbation v = sum_{k=1..8} a_k sin(2 pi k x), a_k = amax * r_k,
//     r_k a seeded uniform number (no counterpart in the reference);
//  7  Kelvin-Helmholtz as in 1 with the noise constant on the cells of a
//     fixed N x N grid of the unit square (params: amp, noise_amp, seed, N):
//     the same initial state at every resolution, so that runs at different
//     resolutions can be compared (no counterpart in the reference);
//  6  the reference's kelvin_helmolz_instability state
//     (Compressible_Euler.h:237-256): v = w0 sin(4 pi x) times Gaussians of
//     width sigma about y = 0.25 and 0.75 (params: w0, sigma; the reference
//     uses w0 = 0.2, sigma = 0.05/sqrt 2).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "adg_hier.hpp"
#include "adg_tables.hpp"

namespace adg {

namespace {
uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
}  // namespace

// uniform [0,1) for (seed, level, cell index)
double cell_uniform(uint64_t seed, int lev, const int* iv) {
    uint64_t k = mix64(seed ^ 0x5eedull);
    k = mix64(k ^ (uint64_t)(uint32_t)lev);
    k = mix64(k ^ (uint64_t)(uint32_t)iv[0]);
    k = mix64(k ^ (uint64_t)(uint32_t)iv[1]);
    k = mix64(k ^ (uint64_t)(uint32_t)iv[2]);
    return (double)(k >> 11) * (1.0 / 9007199254740992.0);
}

// uniform [0,1) for (seed, mode k) of the multi-mode KH perturbation
double mode_uniform(uint64_t seed, int k) {
    uint64_t h = mix64(seed ^ 0x6b68ull);
    h = mix64(h ^ (uint64_t)(uint32_t)k);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void pointwise_state(int D, int ic_case, const double* prm, double gamma, const double* x, const double* plo,
                     const double* phi, int lev, const int* iv, double* u) {
    for (int q = 0; q < 5; ++q) u[q] = 0.0;
    if (ic_case == 0) {
        const double beta = prm[0], xc = prm[1], yc = prm[2], uinf = prm[3], vinf = prm[4];
        const double r = std::sqrt((x[0] - xc) * (x[0] - xc) + (x[1] - yc) * (x[1] - yc));
        const double rho = std::pow(1.0 - (((gamma - 1.0) * std::pow(beta, 2.0)) /
                                           (8.0 * gamma * std::pow(M_PI, 2))) * std::exp(1.0 - std::pow(r, 2.0)),
                                    (1.0 / (gamma - 1.0)));
        const double prs = std::pow(rho, gamma);
        const double du1 = -(x[1] - yc) * (beta / (M_PI * 2.0)) * std::exp((1.0 - std::pow(r, 2.0)) / 2.0);
        const double du2 = (x[0] - xc) * (beta / (M_PI * 2.0)) * std::exp((1.0 - std::pow(r, 2.0)) / 2.0);
        const double u1 = uinf + du1, u2 = vinf + du2;
        const double e = (prs / (rho * (gamma - 1.0))) + 0.5 * (std::pow(u1, 2.0) + std::pow(u2, 2.0));
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2; u[3] = rho * e;
    } else if (ic_case == 1) {
        const double amp = prm[0], namp = prm[1];
        const uint64_t seed = (uint64_t)prm[2];
        const bool in = std::fabs(x[1] - 0.5) < 0.25;
        const double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        const double u2 = amp * std::sin(4.0 * M_PI * x[0]) + namp * (2.0 * cell_uniform(seed, lev, iv) - 1.0);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2;
        u[D + 1] = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
    } else if (ic_case == 7) {
        const double amp = prm[0], namp = prm[1];
        const uint64_t seed = (uint64_t)prm[2];
        const int N = (int)prm[3];
        int nv[3] = {0, 0, 0};
        for (int d = 0; d < 2; ++d) nv[d] = std::min(N - 1, std::max(0, (int)std::floor(x[d] * N)));
        const bool in = std::fabs(x[1] - 0.5) < 0.25;
        const double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        const double u2 = amp * std::sin(4.0 * M_PI * x[0]) + namp * (2.0 * cell_uniform(seed, 0, nv) - 1.0);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2;
        u[D + 1] = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
    } else if (ic_case == 4) {
        const uint64_t seed = (uint64_t)prm[0];
        const double amax = prm[1];
        const bool in = std::fabs(x[1] - 0.5) < 0.25;
        const double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        double u2 = 0.0;
        for (int k = 1; k <= 8; ++k) u2 += amax * mode_uniform(seed, k) * std::sin(2.0 * M_PI * k * x[0]);
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2;
        u[D + 1] = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
    } else if (ic_case == 6) {
        const double w0 = prm[0], sigma = prm[1];
        const double y = x[1];
        const bool in = (y > 0.25) && (y < 0.75);
        const double rho = in ? 2.0 : 1.0, u1 = in ? 0.5 : -0.5, prs = 2.5;
        const double two_s2 = 2.0 * sigma * sigma;
        const double env = std::exp(-(y - 0.25) * (y - 0.25) / two_s2) + std::exp(-(y - 0.75) * (y - 0.75) / two_s2);
        const double u2 = w0 * std::sin(4.0 * M_PI * x[0]) * env;
        u[0] = rho; u[1] = rho * u1; u[2] = rho * u2;
        u[D + 1] = (prs / (gamma - 1.0)) + 0.5 * rho * (u1 * u1 + u2 * u2);
    } else if (ic_case == 5) {
        // smooth periodic density wave advected by a uniform flow
        double prod = 1.0;
        for (int d = 0; d < D; ++d) prod *= std::sin(2.0 * M_PI * (x[d] - plo[d]) / (phi[d] - plo[d]));
        const double rho = 1.0 + prm[0] * prod, prs = prm[1];
        const double vel[3] = {1.0, 0.5, 0.25};
        double k = 0.0;
        u[0] = rho;
        for (int d = 0; d < D; ++d) { u[d + 1] = rho * vel[d]; k += vel[d] * vel[d]; }
        u[D + 1] = prs / (gamma - 1.0) + 0.5 * rho * k;
    } else if (ic_case == 2) {
        const double rho = prm[0];
        const double vel[3] = {prm[1], prm[2], D > 2 ? prm[3] : 0.0};
        const double prs = prm[D == 3 ? 4 : 3];
        double k = 0.0;
        u[0] = rho;
        for (int d = 0; d < D; ++d) { u[d + 1] = rho * vel[d]; k += vel[d] * vel[d]; }
        u[D + 1] = prs / (gamma - 1.0) + 0.5 * rho * k;
    } else {
        double r2 = 0.0;
        for (int d = 0; d < D; ++d) {
            const double c = 0.5 * (plo[d] + phi[d]);
            r2 += (x[d] - c) * (x[d] - c);
        }
        const bool in = std::sqrt(r2) <= prm[0];
        const double rho = in ? prm[1] : prm[3], prs = in ? prm[2] : prm[4];
        u[0] = rho;
        u[D + 1] = prs / (gamma - 1.0);
    }
}

// Projects onto the modes of the slots of patch p (ghosts included) of its
// level, into out ([q][n][slot of the level]).  For independent patches the
// coordinates are local to the patch, whose domain is [plo, plo + n dx).
void project_ic_patch(const ElementTables& t, const Hierarchy& H, int p, int ic_case, const double* prm, double gamma,
                      std::vector<double>& out) {
    const int D = t.D, Q = t.Q, Ns = t.Np_s, Ms = t.qMp_s;
    const PatchLayout& P = H.patch[p];
    const LevelLayout& L = H.lev[P.lev];
    const long ns = L.nslot;
    std::vector<double> xi((size_t)Ms * D);
    for (int m = 0; m < Ms; ++m)
        for (int d = 0; d < D; ++d) xi[(size_t)m * D + d] = t.nodes[t.node_idx(m, d, D)];
    double plo[3] = {H.plo[0], H.plo[1], H.plo[2]}, phi[3] = {H.phi[0], H.phi[1], H.phi[2]}, shift[3] = {0, 0, 0};
    if (H.independent)
        for (int d = 0; d < D; ++d) {
            shift[d] = P.lo[d] * L.dx[d];
            phi[d] = plo[d] + P.n[d] * L.dx[d];
        }
    const int glo[3] = {P.lo[0] - 1, P.lo[1] - 1, D > 2 ? P.lo[2] - 1 : 0};
    const int ghi[3] = {P.hi[0] + 1, P.hi[1] + 1, D > 2 ? P.hi[2] + 1 : 0};
    int iv[3];
    std::vector<double> st((size_t)Ms * 5);
    for (iv[2] = glo[2]; iv[2] <= ghi[2]; ++iv[2])
        for (iv[1] = glo[1]; iv[1] <= ghi[1]; ++iv[1])
            for (iv[0] = glo[0]; iv[0] <= ghi[0]; ++iv[0]) {
                const long s = H.slot_of(p, iv) - L.slot_off;
                for (int m = 0; m < Ms; ++m) {
                    double x[3] = {0, 0, 0};
                    for (int d = 0; d < D; ++d)
                        x[d] = (L.dx[d] / 2.0) * xi[(size_t)m * D + d] + (H.plo[d] + (iv[d] + 0.5) * L.dx[d]) - shift[d];
                    pointwise_state(D, ic_case, prm, gamma, x, plo, phi, P.lev, iv, &st[(size_t)m * 5]);
                }
                for (int q = 0; q < Q; ++q)
                    for (int n = 0; n < Ns; ++n) {
                        double sum = 0.0;
                        for (int m = 0; m < Ms; ++m) sum += st[(size_t)m * 5 + q] * t.quadmat[(size_t)n * Ms + m];
                        out[((size_t)q * Ns + n) * ns + s] = sum / t.Mk_corr[(size_t)n * Ns + n];
                    }
            }
}

// Projects onto the modes of every slot of level lev: [q][n][slot].
void project_ic(const ElementTables& t, const Hierarchy& H, int lev, int ic_case, const double* prm, double gamma,
                std::vector<double>& out) {
    const LevelLayout& L = H.lev[lev];
    out.assign((size_t)t.Q * t.Np_s * L.nslot, 0.0);
    for (int p = L.p_begin; p < L.p_end; ++p) project_ic_patch(t, H, p, ic_case, prm, gamma, out);
}

}  // namespace adg
