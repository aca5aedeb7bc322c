// adg_kernels.hpp -- gfx950 kernels of the ADER-DG hot path (FP64 form).
//
// Data layout in HBM (structure of arrays, cells contiguous):
//   Uw   [Q][Np_s][NS]            modal solution, every slot of every level
//   Vol  [Q][Np_s][NS]            corrector volume term sum_d (Dt/dx_d) Sk_corr[d] F_d
//   Tr   [D][2][Q][qMp_bd][NS]    predictor traces on the faces (side 0: xi=-1, 1: +1)
//   Fnum [d]: [Q][qMp_bd][NF_d]   LLF flux at the face quadrature nodes
//   Mom  [Q][Np_s][NM]            flux-register face moments
// Element matrices live in one read-only buffer; every index into it is
// wave-uniform, so the loads are scalar (SGPR operands of v_fma_f64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace adg {

constexpr int kMaxLev = 8;

__host__ __device__ constexpr int binom(int n, int k) {
    int r = 1;
    for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
    return r;
}
__host__ __device__ constexpr int ipow_c(int b, int e) { return e == 0 ? 1 : b * ipow_c(b, e - 1); }

template <int P, int D>
struct Dims {
    static constexpr int Q = D + 2;
    static constexpr int N = P + 1;
    static constexpr int Ns = binom(P + D, D);
    static constexpr int Nst = binom(P + D + 1, D + 1);
    static constexpr int Ms = ipow_c(N, D);
    static constexpr int Mst = ipow_c(N, D + 1);
    static constexpr int Mbd = ipow_c(N, D);
    static constexpr int NCH = 1 << D;
    static constexpr int NFC = 1 << (D - 1);
    static constexpr int EPW = 64 / Q;  // elements per wavefront (lane = element * Q + q)
};

// Offsets (in doubles) of every table inside the device table buffer.
struct TabOff {
    int A0, B, V, Skc, Vbd, Mkbdm, Mkbdp, Mkc, Minv, Pcf, Pfc, Pfl, Pfh, phi0, total;
};

struct LevelGeo {
    double dx[kMaxLev][3];
    double area[kMaxLev][3];   // face area of direction d (product of the other dx)
    double vol[kMaxLev];
    double dx_avg[kMaxLev];
};

struct TimeState {
    double t;
    double Dt;
    unsigned long long dtmin_bits;
    double pad;
};

static constexpr double kNumLimit = 10.0 * 2.220446049250313e-16;  // 10 * DBL_EPSILON

// ---------------------------------------------------------------------------
// Euler pointwise model (Compressible_Euler.h:370-604, .cpp:50-70).
template <int D>
__device__ __forceinline__ double pressure(const double* s, double rinv, double gm1) {
    double k = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) k += s[a + 1] * s[a + 1] * rinv;
    double p = (-0.5 * k + s[D + 1]) * gm1;
    return fmax(p, kNumLimit);
}

// Flux component q of direction d, branch-free in q (q varies across lanes).
template <int D>
__device__ __forceinline__ double flux_q(int d, int q, const double* s, double rinv, double prs) {
    const double md = s[d + 1];
    const double ud = md * rinv;
    double sq = s[0];
#pragma unroll
    for (int a = 1; a < D + 2; ++a) sq = (q == a) ? s[a] : sq;
    double mom = ud * sq + ((q == d + 1) ? prs : 0.0);
    double en = (s[D + 1] + prs) * ud;
    return (q == 0) ? md : ((q == D + 1) ? en : mom);
}

// ---------------------------------------------------------------------------
// Broadcast of the value held by lane (base + src) of each element group.
template <int Q>
__device__ __forceinline__ double group_bcast(double v, int src, int lane_base) {
    if constexpr (Q == 4) {
        // DPP quad_perm broadcast inside each quad of lanes: no LDS traffic.
        int lo = __double2loint(v), hi = __double2hiint(v);
        int rlo, rhi;
        switch (src) {
            case 0:
                rlo = __builtin_amdgcn_mov_dpp(lo, 0x00, 0xf, 0xf, false);
                rhi = __builtin_amdgcn_mov_dpp(hi, 0x00, 0xf, 0xf, false);
                break;
            case 1:
                rlo = __builtin_amdgcn_mov_dpp(lo, 0x55, 0xf, 0xf, false);
                rhi = __builtin_amdgcn_mov_dpp(hi, 0x55, 0xf, 0xf, false);
                break;
            case 2:
                rlo = __builtin_amdgcn_mov_dpp(lo, 0xaa, 0xf, 0xf, false);
                rhi = __builtin_amdgcn_mov_dpp(hi, 0xaa, 0xf, 0xf, false);
                break;
            default:
                rlo = __builtin_amdgcn_mov_dpp(lo, 0xff, 0xf, 0xf, false);
                rhi = __builtin_amdgcn_mov_dpp(hi, 0xff, 0xf, 0xf, false);
                break;
        }
        return __hiloint2double(rhi, rlo);
    } else {
        return __shfl(v, lane_base + src, 64);
    }
}

// ---------------------------------------------------------------------------
// Predictor: set_predictor guess, P Picard iterations
//   H_w <- A0 U_w + sum_d (Dt/dx_d) B_d F_d(V H_w),  B_d = -Mk_h_w^-1 Sk_predVinv[d]
// then the final nodal evaluation: corrector volume term and face traces.
struct PredArgs {
    const double* __restrict__ tab;
    TabOff to;
    const double* __restrict__ Uw;
    double* __restrict__ Vol;
    double* __restrict__ Tr;
    double* __restrict__ Hw_dbg;   // nullable
    const uint8_t* __restrict__ slot_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NS;
    long slot_begin;
    long nelem;
    double gamma;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_predictor(PredArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Nst = Dm::Nst, Mst = Dm::Mst, Mbd = Dm::Mbd, EPW = Dm::EPW;
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int el = lane / Q, q = lane - (lane / Q) * Q;
    const bool lane_ok = lane < EPW * Q;
    if (!lane_ok) { el = EPW - 1; q = Q - 1; }
    long e = wave * EPW + el;
    const bool active = lane_ok && e < a.nelem;
    if (e >= a.nelem) e = a.nelem - 1;
    const long slot = a.slot_begin + e;
    const int lane_base = lane - q;
    const int lev = a.slot_lev[slot];
    const double Dt = a.ts->Dt;
    double cd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) cd[d] = Dt / a.geo.dx[lev][d];
    const double gm1 = a.gamma - 1.0;
    const long NS = a.NS;

    const double* __restrict__ A0 = a.tab + a.to.A0;
    const double* __restrict__ B = a.tab + a.to.B;
    const double* __restrict__ V = a.tab + a.to.V;
    const double* __restrict__ Skc = a.tab + a.to.Skc;
    const double* __restrict__ Vbd = a.tab + a.to.Vbd;

    double u[Ns];
#pragma unroll
    for (int n = 0; n < Ns; ++n) u[n] = a.Uw[((long)q * Ns + n) * NS + slot];
    double c[Nst];
#pragma unroll
    for (int n = 0; n < Nst; ++n) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < Ns; ++k) s = fma(A0[n * Ns + k], u[k], s);
        c[n] = s;
    }
    double hw[Nst];
#pragma unroll
    for (int n = 0; n < Nst; ++n) hw[n] = (n < Ns) ? u[n] : 0.0;  // AmrDG.tpp:522-533 flat copy

    for (int it = 0; it < P; ++it) {
        double acc[Nst];
#pragma unroll
        for (int n = 0; n < Nst; ++n) acc[n] = c[n];
        for (int m = 0; m < Mst; ++m) {
            double h = 0.0;
#pragma unroll
            for (int n = 0; n < Nst; ++n) h = fma(V[m * Nst + n], hw[n], h);
            double s[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k) s[k] = group_bcast<Q>(h, k, lane_base);
            const double rinv = 1.0 / s[0];
            const double prs = pressure<D>(s, rinv, gm1);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double g = cd[d] * flux_q<D>(d, q, s, rinv, prs);
                const double* __restrict__ Bd = B + (long)d * Nst * Mst + m;
#pragma unroll
                for (int n = 0; n < Nst; ++n) acc[n] = fma(Bd[n * Mst], g, acc[n]);
            }
        }
#pragma unroll
        for (int n = 0; n < Nst; ++n) hw[n] = acc[n];
    }

    // final evaluation: corrector volume term sum_d (Dt/dx_d) Sk_corr[d] F_d
    double vol[Ns];
#pragma unroll
    for (int k = 0; k < Ns; ++k) vol[k] = 0.0;
    for (int m = 0; m < Mst; ++m) {
        double h = 0.0;
#pragma unroll
        for (int n = 0; n < Nst; ++n) h = fma(V[m * Nst + n], hw[n], h);
        double s[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) s[k] = group_bcast<Q>(h, k, lane_base);
        const double rinv = 1.0 / s[0];
        const double prs = pressure<D>(s, rinv, gm1);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double g = cd[d] * flux_q<D>(d, q, s, rinv, prs);
            const double* __restrict__ Sd = Skc + (long)d * Ns * Mst + m;
#pragma unroll
            for (int k = 0; k < Ns; ++k) vol[k] = fma(Sd[k * Mst], g, vol[k]);
        }
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < Ns; ++k) a.Vol[((long)q * Ns + k) * NS + slot] = vol[k];
        if (a.Hw_dbg) {
#pragma unroll
            for (int n = 0; n < Nst; ++n) a.Hw_dbg[((long)q * Nst + n) * NS + slot] = hw[n];
        }
    }
    // face traces (get_H_from_H_w_bd) for every direction and side
    for (int ds = 0; ds < 2 * D; ++ds) {
        const double* __restrict__ Vb = Vbd + (long)ds * Mbd * Nst;
        for (int mb = 0; mb < Mbd; ++mb) {
            double tr = 0.0;
#pragma unroll
            for (int n = 0; n < Nst; ++n) tr = fma(Vb[mb * Nst + n], hw[n], tr);
            if (active) a.Tr[(((long)ds * Q + q) * Mbd + mb) * NS + slot] = tr;
        }
    }
}

// ---------------------------------------------------------------------------
// Face flux: LLF at every face quadrature node (AmrDG_NumericalFlux.tpp:4-41)
// from the traces of the left (side +) and right (side -) cells.
struct FaceArgs {
    const double* __restrict__ Tr;
    double* __restrict__ Fnum[3];
    const int* __restrict__ face_lr[3];
    long NF[3];
    long NS;
    double gamma;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_faceflux(FaceArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Mbd = Dm::Mbd;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0;
    long tot = 0;
#pragma unroll
    for (int dd = 0; dd < D; ++dd) {
        long nd = a.NF[dd] * Mbd;
        if (t >= tot && t < tot + nd) d = dd;
        tot += nd;
    }
    if (t >= tot) return;
    long base = 0;
    for (int dd = 0; dd < d; ++dd) base += a.NF[dd] * Mbd;
    long r = t - base;
    const long NFd = a.NF[d];
    const int mb = (int)(r / NFd);
    const long f = r - (long)mb * NFd;
    const long sl = a.face_lr[d][2 * f], sr = a.face_lr[d][2 * f + 1];
    const long NS = a.NS;
    double uL[Q], uR[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uL[q] = a.Tr[(((long)(2 * d + 1) * Q + q) * Mbd + mb) * NS + sl];
        uR[q] = a.Tr[(((long)(2 * d + 0) * Q + q) * Mbd + mb) * NS + sr];
    }
    const double gm1 = a.gamma - 1.0;
    const double riL = 1.0 / uL[0], riR = 1.0 / uR[0];
    const double pL = pressure<D>(uL, riL, gm1), pR = pressure<D>(uR, riR, gm1);
    const double cL = sqrt(a.gamma * (pL / fmax(uL[0], kNumLimit)));
    const double cR = sqrt(a.gamma * (pR / fmax(uR[0], kNumLimit)));
    const double unL = uL[d + 1] * riL, unR = uR[d + 1] * riR;
    const double dfL = fmax(fmax(fabs(unL - cL), fabs(unL)), fabs(unL + cL));
    const double dfR = fmax(fmax(fabs(unR - cR), fabs(unR)), fabs(unR + cR));
    const double C = fmax(fabs(dfL), fabs(dfR));
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double fL = flux_q<D>(d, q, uL, riL, pL);
        const double fR = flux_q<D>(d, q, uR, riR, pR);
        a.Fnum[d][((long)q * Mbd + mb) * NFd + f] = 0.5 * ((fL + fR) - C * (uR[q] - uL[q]));
    }
}

// ---------------------------------------------------------------------------
// Corrector on valid cells (AmrDG.tpp:652-742).
struct CorrArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;
    const double* __restrict__ Vol;
    const double* __restrict__ Fnum[3];
    const int* __restrict__ cell_face[3][2];
    const int* __restrict__ valid_list;
    const uint8_t* __restrict__ slot_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NF[3];
    long NS;
    long nvalid;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_corrector(CorrArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Mbd = Dm::Mbd;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nvalid * Q) return;
    const int q = (int)(t / a.nvalid);
    const long i = t - (long)q * a.nvalid;
    const long slot = a.valid_list[i];
    const int lev = a.slot_lev[slot];
    const double Dt = a.ts->Dt;
    const long NS = a.NS;
    const double* __restrict__ Mkc = a.tab + a.to.Mkc;
    const double* __restrict__ Mbm = a.tab + a.to.Mkbdm;
    const double* __restrict__ Mbp = a.tab + a.to.Mkbdp;
    double r[Ns];
#pragma unroll
    for (int k = 0; k < Ns; ++k)
        r[k] = Mkc[k] * a.Uw[((long)q * Ns + k) * NS + slot] + a.Vol[((long)q * Ns + k) * NS + slot];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double cdd = Dt / a.geo.dx[lev][d];
        const long flo = a.cell_face[d][0][i], fhi = a.cell_face[d][1][i];
        const double* __restrict__ F = a.Fnum[d] + (long)q * Mbd * a.NF[d];
        for (int mb = 0; mb < Mbd; ++mb) {
            const double Fh = F[(long)mb * a.NF[d] + fhi];
            const double Fl = F[(long)mb * a.NF[d] + flo];
#pragma unroll
            for (int k = 0; k < Ns; ++k)
                r[k] -= cdd * (Mbp[((long)d * Ns + k) * Mbd + mb] * Fh - Mbm[((long)d * Ns + k) * Mbd + mb] * Fl);
        }
    }
#pragma unroll
    for (int k = 0; k < Ns; ++k) a.Uw[((long)q * Ns + k) * NS + slot] = r[k] / Mkc[k];
}

// ---------------------------------------------------------------------------
// Flux-register face moments: coarse faces J * Mkbd{p|m}[d] Fnum, fine faces
// J * P_flux_{high|low}[d][child] Fnum (AmrDG.tpp:604-647).
struct MomArgs {
    const double* __restrict__ tab;
    TabOff to;
    const double* __restrict__ Fnum[3];
    double* __restrict__ Mom;
    const int* __restrict__ mom_face;
    const uint8_t* __restrict__ mom_d;
    const uint8_t* __restrict__ mom_kind;
    const uint8_t* __restrict__ mom_sel;
    const uint8_t* __restrict__ mom_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NF[3];
    long NM;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_moments(MomArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Mbd = Dm::Mbd, NFC = Dm::NFC;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.NM * Q) return;
    const int q = (int)(t / a.NM);
    const long j = t - (long)q * a.NM;
    const int d = a.mom_d[j], lev = a.mom_lev[j], sel = a.mom_sel[j];
    const long f = a.mom_face[j];
    const double jac = (a.ts->Dt / 2.0) * (a.geo.area[lev][d] / (double)(1 << (D - 1)));
    const double* __restrict__ M;
    if (a.mom_kind[j] == 0) {
        M = a.tab + ((sel & 1) ? a.to.Mkbdp : a.to.Mkbdm) + (long)d * Ns * Mbd;
    } else {
        const int child = sel >> 1;
        M = a.tab + ((sel & 1) ? a.to.Pfh : a.to.Pfl) + ((long)d * NFC + child) * Ns * Mbd;
    }
    const double* __restrict__ F = a.Fnum[d] + (long)q * Mbd * a.NF[d] + f;
    double s[Ns];
#pragma unroll
    for (int n = 0; n < Ns; ++n) s[n] = 0.0;
    for (int mb = 0; mb < Mbd; ++mb) {
        const double fv = F[(long)mb * a.NF[d]];
#pragma unroll
        for (int n = 0; n < Ns; ++n) s[n] = fma(M[n * Mbd + mb], fv, s[n]);
    }
#pragma unroll
    for (int n = 0; n < Ns; ++n) a.Mom[((long)q * Ns + n) * a.NM + j] = s[n] * jac;
}

// Reflux: u_c += (2^D / vol_c) * Minv * sum(coef * moment) on corrected cells
// (AmrDG_Interpolater.tpp:7-72).
struct RefluxArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;
    const double* __restrict__ Mom;
    const int* __restrict__ rc_slot;
    const int* __restrict__ rc_start;
    const int* __restrict__ rc_count;
    const uint8_t* __restrict__ rc_lev;
    const int* __restrict__ re_mom;
    const double* __restrict__ re_coef;
    LevelGeo geo;
    long NS, NM, nrc;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_reflux(RefluxArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nrc * Q) return;
    const int q = (int)(t / a.nrc);
    const long r = t - (long)q * a.nrc;
    const long slot = a.rc_slot[r];
    const int lev = a.rc_lev[r];
    const double inv_jac = (double)(1 << D) / a.geo.vol[lev];
    const double* __restrict__ Minv = a.tab + a.to.Minv;
    double corr[Ns];
#pragma unroll
    for (int n = 0; n < Ns; ++n) corr[n] = 0.0;
    const int st = a.rc_start[r], cnt = a.rc_count[r];
    for (int e = 0; e < cnt; ++e) {
        const long j = a.re_mom[st + e];
        const double cf = a.re_coef[st + e];
#pragma unroll
        for (int n = 0; n < Ns; ++n) corr[n] += cf * a.Mom[((long)q * Ns + n) * a.NM + j];
    }
#pragma unroll
    for (int n = 0; n < Ns; ++n) a.Uw[((long)q * Ns + n) * a.NS + slot] += inv_jac * (Minv[n] * corr[n]);
}

// ---------------------------------------------------------------------------
// Average down: u_c = 2^-D Minv sum_c P_fc[c] u_f[c] (AmrDG_Interpolater.tpp:366-398).
struct AvgArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;
    const int* __restrict__ coarse;
    const int* __restrict__ fine;
    long NS, ngroup;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_avgdown(AvgArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, NCH = Dm::NCH;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ngroup * Q) return;
    const int q = (int)(t / a.ngroup);
    const long g = t - (long)q * a.ngroup;
    const double* __restrict__ Pfc = a.tab + a.to.Pfc;
    const double* __restrict__ Minv = a.tab + a.to.Minv;
    double s[Ns];
#pragma unroll
    for (int n = 0; n < Ns; ++n) s[n] = 0.0;
    for (int c = 0; c < NCH; ++c) {
        const long fs = a.fine[g * NCH + c];
        double uf[Ns];
#pragma unroll
        for (int k = 0; k < Ns; ++k) uf[k] = a.Uw[((long)q * Ns + k) * a.NS + fs];
#pragma unroll
        for (int n = 0; n < Ns; ++n) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < Ns; ++k) v = fma(Pfc[(c * Ns + n) * Ns + k], uf[k], v);
            s[n] += v;
        }
    }
    const long cs = a.coarse[g];
#pragma unroll
    for (int n = 0; n < Ns; ++n) a.Uw[((long)q * Ns + n) * a.NS + cs] = (1.0 / NCH) * (Minv[n] * s[n]);
}

// Ghost fill: same-level / periodic copy, or u_f = Minv P_cf[child] u_c
// (AmrDG_Interpolater.tpp:344-364).
struct GhostArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;
    const int* __restrict__ g_slot;
    const int* __restrict__ g_src;
    const uint8_t* __restrict__ g_kind;
    const uint8_t* __restrict__ g_child;
    long NS, nghost;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_ghosts(GhostArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nghost * Q) return;
    const int q = (int)(t / a.nghost);
    const long g = t - (long)q * a.nghost;
    const long s = a.g_slot[g], src = a.g_src[g];
    double uc[Ns];
#pragma unroll
    for (int k = 0; k < Ns; ++k) uc[k] = a.Uw[((long)q * Ns + k) * a.NS + src];
    if (a.g_kind[g] == 0) {
#pragma unroll
        for (int k = 0; k < Ns; ++k) a.Uw[((long)q * Ns + k) * a.NS + s] = uc[k];
        return;
    }
    const double* __restrict__ Pcf = a.tab + a.to.Pcf + (long)a.g_child[g] * Ns * Ns;
    const double* __restrict__ Minv = a.tab + a.to.Minv;
#pragma unroll
    for (int n = 0; n < Ns; ++n) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < Ns; ++k) v = fma(Pcf[n * Ns + k], uc[k], v);
        a.Uw[((long)q * Ns + n) * a.NS + s] = Minv[n] * v;
    }
}

// ---------------------------------------------------------------------------
// CFL: lambda at spatial node 0 (AmrDG.h:1300-1314), min over valid cells by
// wavefront shuffle + one atomicMin per wavefront on the IEEE bits (positive
// doubles order like unsigned integers).
struct CflArgs {
    const double* __restrict__ tab;
    TabOff to;
    const double* __restrict__ Uw;
    const int* __restrict__ valid_list;
    const uint8_t* __restrict__ slot_lev;
    TimeState* ts;
    LevelGeo geo;
    long NS, nvalid;
    double gamma, cfl;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_cfl(CflArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dt = 1.0e300;
    if (i < a.nvalid) {
        const long slot = a.valid_list[i];
        const int lev = a.slot_lev[slot];
        const double* __restrict__ phi0 = a.tab + a.to.phi0;
        double s[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double v = 0.0;
#pragma unroll
            for (int n = 0; n < Ns; ++n) v += a.Uw[((long)q * Ns + n) * a.NS + slot] * phi0[n];
            s[q] = v;
        }
        const double rinv = 1.0 / s[0];
        const double prs = pressure<D>(s, rinv, a.gamma - 1.0);
        const double c = sqrt(a.gamma * (prs / fmax(s[0], kNumLimit)));
        double lmax = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) lmax = fmax(lmax, fabs(s[d + 1] / s[0]) + c);
        dt = a.cfl * (a.geo.dx_avg[lev] / lmax);
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
    if ((threadIdx.x & 63) == 0 && dt < 1.0e300)
        atomicMin(&a.ts->dtmin_bits, (unsigned long long)__double_as_longlong(dt));
}

// Finalize: optional t += Dt, then Dt = c_dt * min, clamped to T - t.
__global__ void k_dt_finalize(TimeState* ts, double c_dt, double T, int advance) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (advance) ts->t = ts->t + ts->Dt;
    double m = __longlong_as_double((long long)ts->dtmin_bits);
    double Dt = c_dt * m;
    if (T - ts->t < Dt) Dt = T - ts->t;
    ts->Dt = Dt;
    ts->dtmin_bits = 0x7ff0000000000000ull;  // +inf for the next reduction
}

// Conserved totals: per-block partial sums of vol * Uw[q][0] (deterministic
// host-side final sum).
struct TotArgs {
    const double* __restrict__ Uw;
    const int* __restrict__ list;
    const uint8_t* __restrict__ slot_lev;
    double* __restrict__ partial;  // [Q][nblocks]
    LevelGeo geo;
    long NS, n;
    int Ns, Q;
};

__global__ void __launch_bounds__(256) k_totals(TotArgs a) {
    __shared__ double red[256];
    for (int q = 0; q < a.Q; ++q) {
        long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
        double v = 0.0;
        if (i < a.n) {
            long s = a.list[i];
            v = a.geo.vol[a.slot_lev[s]] * a.Uw[((long)q * a.Ns) * a.NS + s];
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.partial[(long)q * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

}  // namespace adg
