// adg_kernels.hpp -- gfx950 kernels of the ADER-DG hot path.
//
// Data layout in HBM:
//   Uw   [Q][Np_s][NS]  FP64        modal solution, every slot of every level
//   Vol  per slot, [MBS][4][4][Q]   corrector volume term sum_d (Dt/dx_d) Sk_corr[d] F_d
//   Tr   per slot, [2D][MBD][4][4][Q] predictor traces on the faces
//                                    (side 0: xi_d=-1, side 1: xi_d=+1)
//   Fnum [d]: [NF_d][Q][qMp_st_bd]   LLF flux at the face quadrature nodes (face-major)
//   Mom  [Q][Np_s][NM]  FP64         flux-register face moments
// Vol and Tr are stored in the order the 16x16 MFMA accumulators hold them
// (row = 16*block + 4*lane_group + register), so a wavefront writes each
// element's block as contiguous 16*Q values.  Vol/Tr/Fnum use the form's
// element type T (double, or float in the FP32-MFMA form); U_w, reflux and
// all reductions stay FP64.
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
__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

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
    static constexpr int EPW = 64 / Q;   // VALU predictor: elements per wavefront
    static constexpr int EPM = 16 / Q;   // MFMA predictor: elements per 16 columns
    // MFMA blocking
    static constexpr int MB2 = cdiv(Nst, 16);  // row blocks of the H_w update
    static constexpr int KSU = cdiv(Ns, 4);    // K steps over U_w modes
    static constexpr int NB = cdiv(Mst, 16);   // node blocks
    static constexpr int KSH = cdiv(Nst, 4);   // K steps over H_w modes
    static constexpr int MBS = cdiv(Ns, 16);   // row blocks of the corrector volume term
    static constexpr int MBD = cdiv(Mbd, 16);  // row blocks of a face trace
};

// Row of a 16x16x4 MFMA accumulator held by register r of lane group g
// (g = lane/16), measured on gfx950: FP32 row = 4g + r, FP64 row = 4r + g.
template <typename T>
__host__ __device__ constexpr int acc_row(int g, int r) {
    return sizeof(T) == 8 ? 4 * r + g : 4 * g + r;
}

// Layouts, slot (or face) index fastest except for the traces:
//   Tr  [ds][mb][slot][q]   face traces, ds = 2*d + side (0 = lo, 1 = hi);
//                           the Q variables of a node innermost: a predictor
//                           tile's store covers its elements' variables as
//                           one contiguous run, and the face flux reads a
//                           node's state as one run
//   Vol [q][k][slot]        corrector volume term, like U_w
//   Fnum[d][q][mb][face]    numerical flux at the face quadrature nodes
template <int P, int D>
__host__ __device__ __forceinline__ long tr_idx(int ds, int mb, int q, long slot, long NS) {
    using Dm = Dims<P, D>;
    return ((long)(ds * Dm::Mbd + mb) * NS + slot) * Dm::Q + q;
}
template <int P, int D>
__host__ __device__ __forceinline__ long vol_idx(int k, int q, long slot, long NS) {
    return ((long)q * Dims<P, D>::Ns + k) * NS + slot;
}

// Offsets (in doubles) of every table inside the device table buffer.
struct TabOff {
    int A0, B, V, Skc, Vbd, Mkbdm, Mkbdp, Mkc, Minv, Pcf, Pfc, Pfl, Pfh, phi0, sgn;
    int Mkbdpf;       // Mkbdp rounded to FP32 (two per double entry), the face moments of the FP32 form
    // Lp-norm rules: [0] the N^D solution points (L1), [1] the (2N)^D points
    // of the L2 norm; basis values [M][Np_s], weights [M], coordinates [M][D]
    struct LpRule { int Vs, ws, xs, M; } lp[2];
    int total;
};
// Offsets (in 64-entry fragments) of the MFMA A-operand fragment sets.
struct FragOff {
    int A0, V, B, S, T, BT, total;
    int tail;   // the B rows past the last full 16-row block use the BT fragments
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
    unsigned int done;   // CFL blocks finished in the current reduction
    unsigned int nstep;  // steps advanced since the counter was last reset (the limiter's n)
};

static constexpr double kNumLimit = 10.0 * 2.220446049250313e-16;  // 10 * DBL_EPSILON

// ---------------------------------------------------------------------------
// Euler pointwise model (Compressible_Euler.h:370-604, .cpp:50-70).
template <int D, typename T>
__device__ __forceinline__ T pressure(const T* s, T rinv, T gm1) {
    T k = T(0);
#pragma unroll
    for (int a = 0; a < D; ++a) k += s[a + 1] * s[a + 1] * rinv;
    T p = (T(-0.5) * k + s[D + 1]) * gm1;
    return fmax(p, T(kNumLimit));
}

// 1/x: IEEE division in FP64; the hardware reciprocal (1 ulp) in the FP32
// form, whose arithmetic is FP32 round-off throughout.
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// Flux component q of direction d, branch-free in q (q varies across lanes).
// sq is the conserved variable q itself (the lane's own value), which keeps
// the compiler from turning a select chain into a private-memory lookup.
template <int D, typename T>
__device__ __forceinline__ T flux_q(int d, int q, const T* s, T sq, T rinv, T prs) {
    const T md = s[d + 1];
    const T ud = md * rinv;
    T mom = ud * sq + ((q == d + 1) ? prs : T(0));
    T en = (s[D + 1] + prs) * ud;
    return (q == 0) ? md : ((q == D + 1) ? en : mom);
}

// ---------------------------------------------------------------------------
// DPP quad_perm broadcast of lane `src` of every quad (no LDS traffic).
template <int SRC>
__device__ __forceinline__ int quad_bcast_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, SRC * 0x55, 0xf, 0xf, false);
}
template <int SRC>
__device__ __forceinline__ double quad_bcast(double v) {
    return __hiloint2double(quad_bcast_i<SRC>(__double2hiint(v)), quad_bcast_i<SRC>(__double2loint(v)));
}
template <int SRC>
__device__ __forceinline__ float quad_bcast(float v) {
    return __int_as_float(quad_bcast_i<SRC>(__float_as_int(v)));
}

// Gather the Q conserved variables of this lane's element: lanes src_base+k.
template <int Q, typename T>
__device__ __forceinline__ void gather_state(T v, int src_base, T* s) {
    if constexpr (Q == 4) {
        s[0] = quad_bcast<0>(v);
        s[1] = quad_bcast<1>(v);
        s[2] = quad_bcast<2>(v);
        s[3] = quad_bcast<3>(v);
    } else {
#pragma unroll
        for (int k = 0; k < Q; ++k) s[k] = __shfl(v, src_base + k, 64);
    }
}

// ---------------------------------------------------------------------------
// VALU predictor (FP64): lane = (element, q); set_predictor guess, P Picard
// iterations  H_w <- A0 U_w + sum_d (Dt/dx_d) B_d F_d(V H_w),
// B_d = -Mk_h_w^-1 Sk_predVinv[d]; then the corrector volume term and traces.
// Element matrices are read through wave-uniform (scalar) loads.
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
    long slot_begin;   // with slots: the first list index, else the first slot
    long nelem;
    double gamma;
    int iters;   // Picard iterations (P)
    const int* __restrict__ slots;   // nullable: the slots to predict (a list), else a slot range
};

// Slot of the e-th element of a predictor launch: from the slot list when
// there is one, else from the contiguous range.
__device__ __forceinline__ long pred_slot(const int* __restrict__ slots, long begin, long e) {
    return slots ? (long)slots[begin + e] : begin + e;
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_predictor(PredArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Nst = Dm::Nst, Mst = Dm::Mst, Mbd = Dm::Mbd, EPW = Dm::EPW;
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int el = lane / Q, q = lane - (lane / Q) * Q;
    const bool lane_ok = lane < EPW * Q;
    if (!lane_ok) { el = EPW - 1; q = 0; }
    long e = wave * EPW + el;
    const bool active = lane_ok && e < a.nelem;
    if (e >= a.nelem) e = a.nelem - 1;
    const long slot = pred_slot(a.slots, a.slot_begin, e);
    const int lane_base = lane_ok ? lane - q : lane;
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

    for (int it = 0; it < a.iters; ++it) {
        double acc[Nst];
#pragma unroll
        for (int n = 0; n < Nst; ++n) acc[n] = c[n];
        for (int m = 0; m < Mst; ++m) {
            double h = 0.0;
#pragma unroll
            for (int n = 0; n < Nst; ++n) h = fma(V[m * Nst + n], hw[n], h);
            double s[Q];
            if constexpr (Q == 4) gather_state<Q>(h, lane_base, s);
            else for (int k = 0; k < Q; ++k) s[k] = __shfl(h, lane_base + k, 64);
            const double rinv = 1.0 / s[0];
            const double prs = pressure<D>(s, rinv, gm1);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double g = cd[d] * flux_q<D>(d, q, s, h, rinv, prs);
                const double* __restrict__ Bd = B + (long)d * Nst * Mst + m;
#pragma unroll
                for (int n = 0; n < Nst; ++n) acc[n] = fma(Bd[n * Mst], g, acc[n]);
            }
        }
#pragma unroll
        for (int n = 0; n < Nst; ++n) hw[n] = acc[n];
    }

    double vol[Ns];
#pragma unroll
    for (int k = 0; k < Ns; ++k) vol[k] = 0.0;
    for (int m = 0; m < Mst; ++m) {
        double h = 0.0;
#pragma unroll
        for (int n = 0; n < Nst; ++n) h = fma(V[m * Nst + n], hw[n], h);
        double s[Q];
        if constexpr (Q == 4) gather_state<Q>(h, lane_base, s);
        else for (int k = 0; k < Q; ++k) s[k] = __shfl(h, lane_base + k, 64);
        const double rinv = 1.0 / s[0];
        const double prs = pressure<D>(s, rinv, gm1);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double g = cd[d] * flux_q<D>(d, q, s, h, rinv, prs);
            const double* __restrict__ Sd = Skc + (long)d * Ns * Mst + m;
#pragma unroll
            for (int k = 0; k < Ns; ++k) vol[k] = fma(Sd[k * Mst], g, vol[k]);
        }
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < Ns; ++k) a.Vol[vol_idx<P, D>(k, q, slot, NS)] = vol[k];
        if (a.Hw_dbg) {
#pragma unroll
            for (int n = 0; n < Nst; ++n) a.Hw_dbg[((long)q * Nst + n) * NS + slot] = hw[n];
        }
    }
    for (int ds = 0; ds < 2 * D; ++ds) {
        const double* __restrict__ Vb = Vbd + (long)ds * Mbd * Nst;
        for (int mb = 0; mb < Mbd; ++mb) {
            double tr = 0.0;
#pragma unroll
            for (int n = 0; n < Nst; ++n) tr = fma(Vb[mb * Nst + n], hw[n], tr);
            if (active) a.Tr[tr_idx<P, D>(ds, mb, q, slot, NS)] = tr;
        }
    }
}

// ---------------------------------------------------------------------------
// MFMA predictor.  One wavefront owns 16 matrix columns = EPM elements x Q
// variables (column c = Q*e + q).  Every contraction is a 16x16x4 MFMA with
// the element matrix as the A operand (pre-swizzled fragments, staged once
// per persistent workgroup in LDS) and the per-element data as B:
//   C0 = A0 * U_w                       (Nst x Ns)
//   H  = V * H_w                        nodal state, 16 nodes per block
//   g  = (Dt/dx_d) F_d(H)               pointwise, in the accumulator layout
//   H_w' = C0 + sum_d B_d * g_d          K ordered so that the accumulator
//                                        registers of H feed B directly
// The H_w' accumulators are re-laid as B operands through a per-wave LDS
// buffer.  Accumulator rows follow acc_row<T>(lane group, register).  Final stage: Vol = sum_d Sk_corr[d] g_d and the face traces.
template <typename T>
struct MF;
template <>
struct MF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct MF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // 16 blocks of 4x4x1: block b = lane/4 takes A row lane%4 and B column
    // lane%4 of its lanes; lane 4b + j, register i holds D_b(i, j)
    static __device__ __forceinline__ v4 mma4(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
    }
};

// Work dealing of the persistent element kernels (predictors, warm start):
// whole rounds give each workgroup nwave * unit consecutive elements (one
// unit per wavefront, full cache lines); the remainder, less than a round, is
// cut into units dealt kq per workgroup, so that the last partial round runs
// on every workgroup (a few wavefronts each, on distinct SIMDs) instead of on
// a few whole workgroups while the rest of the GPU idles.
struct Deal {
    long nfull;   // whole rounds
    long rest0;   // first element of the remainder
    long nrwb;    // units in the remainder
    long kq;      // remainder units per workgroup
};
__device__ __forceinline__ Deal make_deal(long nelem, long unit, int nwave) {
    Deal d;
    const long round = (long)gridDim.x * nwave * unit;
    d.nfull = nelem / round;
    d.rest0 = d.nfull * round;
    d.nrwb = (nelem - d.rest0 + unit - 1) / unit;
    d.kq = (d.nrwb + gridDim.x - 1) / gridDim.x;
    return d;
}
// First element of iteration it of this wavefront; false when it has none.
__device__ __forceinline__ bool deal_next(const Deal& d, long it, int wave, long unit, int nwave, long* e0) {
    if (it < d.nfull) {
        *e0 = ((it * gridDim.x + blockIdx.x) * nwave + wave) * unit;
        return true;
    }
    if (it == d.nfull && wave < d.kq) {
        const long rb = (long)blockIdx.x * d.kq + wave;
        if (rb < d.nrwb) {
            *e0 = d.rest0 + rb * unit;
            return true;
        }
    }
    return false;
}

// Row stride (elements) of the node-lane predictor's per-wave LDS tile:
// 16 nodes plus padding that keeps 16-byte alignment of each row.
template <typename T>
__host__ __device__ constexpr int nl_tile_stride() {
    return sizeof(T) == 8 ? 18 : 20;
}
// Directions the tile holds at once: all D, except FP64 in 3D, where one
// direction at a time keeps the p = 2 fragments and eight tiles within the
// LDS of a CU.
template <typename T>
__host__ __device__ constexpr int nl_tile_dirs(int D) {
    return (sizeof(T) == 8 && D == 3) ? 1 : D;
}

// Workgroup copy of nbytes (a multiple of 16) from global memory to LDS with
// eight 16-byte loads in flight per thread before their stores, so that the
// staging of element matrices costs one memory latency, not one per pass.
__device__ __forceinline__ void stage_to_lds(void* dst, const void* src, int nbytes) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    const uint4* s = reinterpret_cast<const uint4*>(src);
    const int n = nbytes >> 4, step = (int)blockDim.x;
    int i = threadIdx.x;
    for (; i + 7 * step < n; i += 8 * step) {
        uint4 r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] = s[i + k * step];
#pragma unroll
        for (int k = 0; k < 8; ++k) d[i + k * step] = r[k];
    }
    for (; i < n; i += step) d[i] = s[i];
}

// max that propagates NaN from either operand (fmax drops a NaN operand,
// which would let a residual of NaN pass an acceptance test)
template <typename T>
__device__ __forceinline__ T nanmax(T a, T b) {
    return (a > b || a != a) ? a : b;
}

// v + v(lane ^ 16) and v + v(lane ^ 32) with the row swaps of gfx950
// (v_permlane16_swap / v_permlane32_swap: VALU, no LDS traffic).  The swap
// of v with itself gives every lane the value of the even and of the odd
// row of its pair (the lower and the upper half); their sum is the sum with
// the partner lane (addition commutes, so the bits are those of
// v + __shfl_xor(v, 16 or 32)).
__device__ __forceinline__ float sum_xor16(float v) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float sum_xor32(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename T>
struct PredMArgs {
    const T* __restrict__ frags;   // [nfrag][64]
    FragOff fo;
    const double* __restrict__ Uw;
    T* __restrict__ Vol;
    T* __restrict__ Tr;
    double* __restrict__ Hw_dbg;   // nullable
    const uint8_t* __restrict__ slot_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NS;
    long slot_begin;
    long nelem;
    double gamma;
    int iters;                               // Picard iterations of the plain path (P)
    // learned warm start (node-lane kernel only; warm == nullptr: plain)
    const float* __restrict__ warm;          // H_w start increment over A0 U_w, [q*Np_st+n][slot]
    float tol;                               // accept when max|H_k - H_{k-1}| <= tol rho0;
                                             // < 0: exactly learned_iters iterations, no check
    int learned_iters;                       // Picard iterations tried from the learned start
    unsigned long long* __restrict__ stats;  // [elements, elements that fell back]
    const int* __restrict__ slots;           // nullable: the slots to predict, slot_begin its first index
};

// Quad-gather K step r of node block nb contracts nodes 16 nb + acc_row(k, r)
// (lane groups k); it holds a node when the smallest of them is below Mst.
// Steps without one multiply zero fragments by zero fluxes; they are skipped
// where the last block holds at most 4 nodes (3D p = 2: the 81st node is
// alone in the sixth block), elsewhere the branches cost more registers
// than the few steps they save.
template <typename T, int Mst>
__device__ __forceinline__ bool quad_kstep_live(int nb, int r) {
    if constexpr (Mst % 16 == 0 || Mst % 16 > 4) return true;
    else return 16 * nb + acc_row<T>(0, r) < Mst;
}

// LEARN = false compiles the learned start out (a.warm is ignored): the plain
// form then holds fewer registers (no learned-start arrays live across the
// Picard loop).  The plain arithmetic is the same in both instantiations.
template <typename T, int P, int D, bool LEARN>
__device__ __forceinline__ void predictor_mfma_body(const PredMArgs<T>& a) {
    using Dm = Dims<P, D>;
    using M = MF<T>;
    using v4 = typename M::v4;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Nst = Dm::Nst, Mst = Dm::Mst, Mbd = Dm::Mbd, EPM = Dm::EPM;
    constexpr int MB2 = Dm::MB2, KSU = Dm::KSU, NB = Dm::NB, KSH = Dm::KSH, MBS = Dm::MBS, MBD = Dm::MBD;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* fr = reinterpret_cast<T*>(smem_raw);
    const int nfr = a.fo.total * 64;
    stage_to_lds(fr, a.frags, nfr * (int)sizeof(T));
    __shared__ unsigned long long wg_stats[2];
    if (threadIdx.x == 0) { wg_stats[0] = 0ull; wg_stats[1] = 0ull; }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    T* xch = fr + nfr + wave * (MB2 * 16 * 16);
    const int g = lane >> 4, c = lane & 15;
    const bool col_ok = c < EPM * Q;
    const int e_loc = col_ok ? c / Q : 0;
    const int q = col_ok ? c - e_loc * Q : 0;
    const int src_base = col_ok ? (lane - q) : lane;
    const T gm1 = T(a.gamma - 1.0);
    const Deal deal = make_deal(a.nelem, EPM, nwave);
    long eb;
    for (long it = 0; deal_next(deal, it, wave, EPM, nwave, &eb); ++it) {
        const long e = eb + e_loc;
        const bool active = col_ok && e < a.nelem;
        const long slot = pred_slot(a.slots, a.slot_begin, e < a.nelem ? e : a.nelem - 1);
        const int lev = a.slot_lev[slot];
        const double Dt = a.ts->Dt;
        T cd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) cd[d] = T(Dt / a.geo.dx[lev][d]);

        // U_w as B operand (modes 4ks+g), then C0 = A0 U_w and the initial
        // guess H_w[mode] = U_w[mode] for mode < Np_s (AmrDG.tpp:522-533).
        T hb[KSH];
#pragma unroll
        for (int ks = 0; ks < KSH; ++ks) {
            const int k = 4 * ks + g;
            hb[ks] = (k < Ns) ? T(a.Uw[((long)q * Ns + k) * a.NS + slot]) : T(0);
        }
        // C0 is recomputed from U_w every iteration: KSU B registers instead
        // of MB2 accumulators stay live across the iterations.
        T hu[KSU];
#pragma unroll
        for (int ks = 0; ks < KSU; ++ks) hu[ks] = hb[ks];

        // C0 = A0 U_w into acc
        auto c0_into = [&](v4 (&acc)[MB2]) {
#pragma unroll
            for (int mb = 0; mb < MB2; ++mb) acc[mb] = v4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KSU; ++ks)
#pragma unroll
                for (int mb = 0; mb < MB2; ++mb)
                    acc[mb] = M::mma(fr[(a.fo.A0 + mb * KSU + ks) * 64 + lane], hu[ks], acc[mb]);
        };
        // accumulator layout -> B operand layout (mode 4ks+g).  FP64: register
        // r of block mb holds mode 16mb+4r+g = 4(4mb+r)+g already; FP32
        // (mode 16mb+4g+r) goes through the per-wave LDS buffer.
        auto relay = [&](const v4 (&acc)[MB2]) {
            if constexpr (sizeof(T) == 8) {
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) hb[ks] = acc[ks >> 2][ks & 3];
            } else {
#pragma unroll
                for (int mb = 0; mb < MB2; ++mb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) xch[(16 * mb + acc_row<T>(g, r)) * 16 + c] = acc[mb][r];
                wave_lds_sync();
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) hb[ks] = xch[(4 * ks + g) * 16 + c];
                wave_lds_sync();
            }
        };
        // one Picard iteration H_w <- C0 + sum_d B_d g_d(H_w)
        auto picard = [&]() {
            v4 acc[MB2];
            c0_into(acc);
#pragma unroll 1
            for (int nb = 0; nb < NB; ++nb) {
                v4 h = v4{0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) h = M::mma(fr[(a.fo.V + nb * KSH + ks) * 64 + lane], hb[ks], h);
                T gd[D][4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (!quad_kstep_live<T, Mst>(nb, r)) {
#pragma unroll
                        for (int d = 0; d < D; ++d) gd[d][r] = T(0);
                        continue;
                    }
                    T s[Q];
                    gather_state<Q>(h[r], src_base, s);
                    const bool ok = col_ok && (16 * nb + acc_row<T>(g, r)) < Mst;
                    if (!ok) s[0] = T(1);
                    const T rinv = recip(s[0]);
                    const T prs = pressure<D>(s, rinv, gm1);
#pragma unroll
                    for (int d = 0; d < D; ++d) gd[d][r] = ok ? cd[d] * flux_q<D>(d, q, s, h[r], rinv, prs) : T(0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (!quad_kstep_live<T, Mst>(nb, r)) continue;
#pragma unroll
                    for (int d = 0; d < D; ++d)
#pragma unroll
                        for (int mb = 0; mb < MB2; ++mb)
                            acc[mb] = M::mma(fr[(a.fo.B + ((mb * D + d) * NB + nb) * 4 + r) * 64 + lane], gd[d][r],
                                             acc[mb]);
                }
            }
            relay(acc);
        };

        int nplain = a.iters;
        bool keep_learned = false;   // this column's element passed the check
        T hl[KSH];
        if (LEARN && a.warm) {
            // learned start H_0 = A0 U_w + Delta and Picard iterations until
            // the update of an element passes the tolerance; the decision is
            // per element (its data only), elements that do not pass are
            // recomputed by the plain predictor from its own start
            const T inv_sc = T(1.0 / fabs(a.Uw[slot]));   // 1 / rho0 of the column's element
            {
                v4 acc[MB2];
                c0_into(acc);
                relay(acc);
            }
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) {
                const int mode = 4 * ks + g;
                if (col_ok && mode < Nst) hb[ks] += T(a.warm[((long)q * Nst + mode) * a.NS + slot]);
            }
            bool pass_e = false;
            if (a.tol < 0.f) {
                for (int k = 0; k < a.learned_iters; ++k) picard();   // diagnostic: no check
                pass_e = true;
            } else {
                for (int k = 0; k < a.learned_iters; ++k) {
                    T prev[KSH];
#pragma unroll
                    for (int ks = 0; ks < KSH; ++ks) prev[ks] = hb[ks];
                    picard();
                    T r = T(0);
#pragma unroll
                    for (int ks = 0; ks < KSH; ++ks) r = nanmax(r, T(fabs(hb[ks] - prev[ks])));
                    r = col_ok ? r * inv_sc : T(0);
                    // max over the element's lanes: 4 lane groups x its Q columns
                    r = nanmax(r, __shfl_xor(r, 16, 64));
                    r = nanmax(r, __shfl_xor(r, 32, 64));
                    T re = r;
#pragma unroll
                    for (int j = 0; j < Q; ++j) re = nanmax(re, __shfl(r, src_base + j, 64));
                    if (pass_e) {
#pragma unroll
                        for (int ks = 0; ks < KSH; ++ks) hb[ks] = prev[ks];
                    } else {
                        pass_e = re <= T(a.tol);   // NaN fails
                    }
                    if (__all(pass_e)) break;
                }
            }
            // one representative lane per element (lane group 0, variable 0)
            const bool rep = g == 0 && q == 0 && active;
            const unsigned long long n_el = __popcll(__ballot(rep));
            const unsigned long long n_fb = __popcll(__ballot(rep && !pass_e));
            if (lane == 0 && n_el) {   // workgroup totals in LDS, one global update per workgroup
                atomicAdd(&wg_stats[0], n_el);
                if (n_fb) atomicAdd(&wg_stats[1], n_fb);
            }
            if (__all(pass_e)) {
                nplain = 0;
            } else {
                keep_learned = pass_e;
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) {
                    hl[ks] = hb[ks];
                    hb[ks] = ks < KSU ? hu[ks] : T(0);
                }
            }
        }
        for (int it = 0; it < nplain; ++it) picard();
        if (keep_learned) {
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) hb[ks] = hl[ks];
        }

        // final evaluation: corrector volume term
        v4 vol[MBS];
#pragma unroll
        for (int mbs = 0; mbs < MBS; ++mbs) vol[mbs] = v4{0, 0, 0, 0};
#pragma unroll 1
        for (int nb = 0; nb < NB; ++nb) {
            v4 h = v4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) h = M::mma(fr[(a.fo.V + nb * KSH + ks) * 64 + lane], hb[ks], h);
            T gd[D][4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (!quad_kstep_live<T, Mst>(nb, r)) {
#pragma unroll
                    for (int d = 0; d < D; ++d) gd[d][r] = T(0);
                    continue;
                }
                T s[Q];
                gather_state<Q>(h[r], src_base, s);
                const bool ok = col_ok && (16 * nb + acc_row<T>(g, r)) < Mst;
                if (!ok) s[0] = T(1);
                const T rinv = recip(s[0]);
                const T prs = pressure<D>(s, rinv, gm1);
#pragma unroll
                for (int d = 0; d < D; ++d) gd[d][r] = ok ? cd[d] * flux_q<D>(d, q, s, h[r], rinv, prs) : T(0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (!quad_kstep_live<T, Mst>(nb, r)) continue;
#pragma unroll
                for (int d = 0; d < D; ++d)
#pragma unroll
                    for (int mbs = 0; mbs < MBS; ++mbs)
                        vol[mbs] = M::mma(fr[(a.fo.S + ((mbs * D + d) * NB + nb) * 4 + r) * 64 + lane], gd[d][r],
                                          vol[mbs]);
            }
        }
        if (active) {
#pragma unroll
            for (int mbs = 0; mbs < MBS; ++mbs)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int mode = 16 * mbs + acc_row<T>(g, r);
                    if (mode < Ns) a.Vol[vol_idx<P, D>(mode, q, slot, a.NS)] = vol[mbs][r];
                }
            if (a.Hw_dbg) {
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) {
                    const int mode = 4 * ks + g;
                    if (mode < Nst) a.Hw_dbg[((long)q * Nst + mode) * a.NS + slot] = double(hb[ks]);
                }
            }
        }
        // face traces: Tr[ds][mb] = Vbd[ds] * H_w
#pragma unroll 1
        for (int ds = 0; ds < 2 * D; ++ds)
#pragma unroll
            for (int mbd = 0; mbd < MBD; ++mbd) {
                v4 t = v4{0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks)
                    t = M::mma(fr[(a.fo.T + (ds * MBD + mbd) * KSH + ks) * 64 + lane], hb[ks], t);
                if (active) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int fn = 16 * mbd + acc_row<T>(g, r);
                        if (fn < Mbd) a.Tr[tr_idx<P, D>(ds, fn, q, slot, a.NS)] = t[r];
                    }
                }
            }
    }
    if (LEARN && a.warm) {
        __syncthreads();
        if (threadIdx.x == 0 && wg_stats[0]) {
            atomicAdd(a.stats, wg_stats[0]);
            if (wg_stats[1]) atomicAdd(a.stats + 1, wg_stats[1]);
        }
    }
}

// 2D: capped at 128 VGPRs so that two 512-thread workgroups (4 waves per
// SIMD) share a CU; the 3D tiles need more registers than that cap allows
// without spilling and run uncapped.
template <typename T, int P, int D, bool LEARN>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_predictor_mfma(PredMArgs<T> a) {
    predictor_mfma_body<T, P, D, LEARN>(a);
}
template <typename T, int P, int D, bool LEARN>
__global__ void __launch_bounds__(512) k_predictor_mfma_wide(PredMArgs<T> a) {
    predictor_mfma_body<T, P, D, LEARN>(a);
}
template <typename T, int P, int D, bool LEARN>
inline const void* predictor_mfma_kernel() {
    if constexpr (D == 2) return (const void*)k_predictor_mfma<T, P, D, LEARN>;
    else return (const void*)k_predictor_mfma_wide<T, P, D, LEARN>;
}

// ---------------------------------------------------------------------------
// Node-lane MFMA predictor.  Q = 4 (2D): a wavefront owns 4 elements; its
// 16 matrix columns are the (element, variable) pairs, column acc_row<T>(e, q),
// so that the nodal product with swapped operands leaves the 4 variables of
// one (element, node) in the 4 accumulator registers of one lane.  Q = 5
// (3D): a wavefront owns 3 elements; lane group e < 3 holds variables 0-3 of
// element e (column acc_row<T>(e, q)), lane group 3 holds the energies of
// the 3 elements (column acc_row<T>(3, e)) and one padding column; a lane
// of group e fetches its energy from lane group 3 (one shuffle per node and
// element) and evaluates the flux of (element e, node) once:
//   C0    = A0 U_w                         rows modes, columns (e, q)
//   H^T   = H_w^T V^T                      lane (e = g, node = c), register q
//   g_d   = (Dt/dx_d) F_d(H)               once per (element, node)
//   H_w'  = C0 + sum_d B_d g_d             g_d re-laid through a per-wave LDS
//                                          tile as the B operand
// Register r of lane group g of every modal accumulator holds logical row
// 16*blk + 4r + g (the fragment rows of A0, B, S and T are permuted for that),
// which is the operand layout of K step 4*blk + r of the next nodal product.
// Final pass: Vol = sum_d S_d g_d and the face traces Tr = Vbd H_w.
template <typename T, int P, int D, bool LEARN>
__device__ __forceinline__ void predictor_nl_body(const PredMArgs<T>& a) {
    using Dm = Dims<P, D>;
    using M = MF<T>;
    using v4 = typename M::v4;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Nst = Dm::Nst, Mst = Dm::Mst, Mbd = Dm::Mbd;
    constexpr int MB2 = Dm::MB2, KSU = Dm::KSU, NB = Dm::NB, KSH = Dm::KSH, MBS = Dm::MBS, MBD = Dm::MBD;
    constexpr int TST = nl_tile_stride<T>();
    constexpr bool FULL = Mst % 16 == 0;
    // FP32 with 4 modes past the last full 16-row block (p = 3: 20 = 16 + 4):
    // those rows use the 4x4x1 16-block MFMA (a quarter of the cycles of a
    // 16x16x4 step whose block would be 12/16 padding); block b = 4g + c/4 of
    // K step (d, nb, s) contracts node 16 nb + 4 g + s, so its B operand is the
    // same gv[d][s] register and the 4 lane groups hold partial sums over
    // disjoint node sets, added at the end of the iteration
    constexpr bool TAIL = sizeof(T) == 4 && MB2 >= 2 && Nst - 16 * (MB2 - 1) == 4;
    constexpr int MBF = TAIL ? MB2 - 1 : MB2;
    constexpr int EW = Q == 4 ? 4 : 3;   // elements per wavefront
    constexpr int TD = nl_tile_dirs<T>(D);
    static_assert(Q == 4 || Q == 5, "node-lane layout for 2D and 3D Euler");
    static_assert(D % TD == 0, "tile rounds cover the directions");
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T* fr = reinterpret_cast<T*>(smem_raw);
    const int nfr = a.fo.total * 64;
    stage_to_lds(fr, a.frags, nfr * (int)sizeof(T));
    __syncthreads();

    __shared__ unsigned long long wg_stats[2];
    if (threadIdx.x == 0) { wg_stats[0] = 0ull; wg_stats[1] = 0ull; }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    T* tile = fr + nfr + wave * (TD * 16 * TST);
    const int g = lane >> 4, c = lane & 15;
    // column c of the modal data is row acc_row<T>(cg, cr) of the nodal
    // accumulator and holds variable cq of element ce (ce = 3: padding)
    const int cg = sizeof(T) == 8 ? (c & 3) : (c >> 2);
    const int cr = sizeof(T) == 8 ? (c >> 2) : (c & 3);
    const int ce = (Q == 4 || cg < 3) ? cg : cr;
    const int cq = (Q == 4 || cg < 3) ? cr : 4;
    // column of variable q of element e
    auto col_of = [](int e, int q) { return (Q == 4 || q < 4) ? acc_row<T>(e, q) : acc_row<T>(3, e); };
    const T gm1 = T(a.gamma - 1.0);
    const Deal deal = make_deal(a.nelem, EW, nwave);
    long e0;
    for (long it = 0; deal_next(deal, it, wave, EW, nwave, &e0); ++it) {
        const bool active = ce < EW && e0 + ce < a.nelem;
        const long slot = pred_slot(a.slots, a.slot_begin, active ? e0 + ce : a.nelem - 1);
        const long slot_f = pred_slot(a.slots, a.slot_begin, e0 + g < a.nelem ? e0 + g : a.nelem - 1);
        const int lev = a.slot_lev[slot_f];
        const double Dt = a.ts->Dt;
        T cd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) cd[d] = T(Dt / a.geo.dx[lev][d]);

        T hu[KSU];
#pragma unroll
        for (int ks = 0; ks < KSU; ++ks) {
            const int k = 4 * ks + g;
            hu[ks] = (k < Ns) ? T(a.Uw[((long)cq * Ns + k) * a.NS + slot]) : T(0);
        }
        v4 c0[MB2];
#pragma unroll
        for (int mb = 0; mb < MB2; ++mb) {
            c0[mb] = v4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KSU; ++ks)
                c0[mb] = M::mma(fr[(a.fo.A0 + mb * KSU + ks) * 64 + lane], hu[ks], c0[mb]);
        }
        // initial guess H_w[mode] = U_w[mode] for mode < Np_s (AmrDG.tpp:522-533)
        T hb[KSH];
#pragma unroll
        for (int ks = 0; ks < KSH; ++ks) hb[ks] = ks < KSU ? hu[ks] : T(0);

        // nodal state of node block nb -> (Dt/dx_d) F_d in the LDS tile
        // [d][column][node] (TD directions per round); returns the B
        // operands of the 4 K steps
        auto nodal_flux = [&](int nb, T (&gv)[D][4]) {
            v4 h = v4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) h = M::mma(hb[ks], fr[(a.fo.V + nb * KSH + ks) * 64 + lane], h);
            T s[Q];
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] = h[q];
            bool ok = FULL || (16 * nb + c < Mst);
            if constexpr (Q == 5) {
                // the energy of element g is register g of lane group 3
                const T en0 = __shfl(h[0], 48 + c, 64), en1 = __shfl(h[1], 48 + c, 64);
                const T en2 = __shfl(h[2], 48 + c, 64);
                s[4] = g == 0 ? en0 : (g == 1 ? en1 : en2);
                ok = ok && g < 3;
            }
            if (!ok) s[0] = T(1);
            const T rinv = recip(s[0]);
            const T prs = pressure<D>(s, rinv, gm1);
            T f[D][Q];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const T ud = s[d + 1] * rinv;
                f[d][0] = s[d + 1];
#pragma unroll
                for (int b = 0; b < D; ++b) f[d][b + 1] = ud * s[b + 1] + ((b == d) ? prs : T(0));
                f[d][D + 1] = (s[D + 1] + prs) * ud;
            }
#pragma unroll
            for (int d0 = 0; d0 < D; d0 += TD) {
#pragma unroll
                for (int dd = 0; dd < TD; ++dd) {
                    const int d = d0 + dd;
                    if (Q == 4 || g < 3) {
#pragma unroll
                        for (int q = 0; q < Q; ++q) tile[(dd * 16 + col_of(g, q)) * TST + c] = ok ? cd[d] * f[d][q] : T(0);
                    } else {
                        tile[(dd * 16 + acc_row<T>(3, 3)) * TST + c] = T(0);   // padding column
                    }
                }
                wave_lds_sync();
#pragma unroll
                for (int dd = 0; dd < TD; ++dd)
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) gv[d0 + dd][s4] = tile[(dd * 16 + c) * TST + 4 * g + s4];
                wave_lds_sync();
            }
        };

        // one Picard iteration H_w <- C0 + sum_d B_d g_d(H_w)
        auto picard = [&]() {
            v4 acc[MB2];
#pragma unroll
            for (int mb = 0; mb < MB2; ++mb) acc[mb] = c0[mb];
            v4 tacc = v4{0, 0, 0, 0};
#pragma unroll 1
            for (int nb = 0; nb < NB; ++nb) {
                T gv[D][4];
                nodal_flux(nb, gv);
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    if (!FULL && 16 * nb + s4 >= Mst) continue;   // K step without a node (3D p = 2: 81 = 5 x 16 + 1)
#pragma unroll
                    for (int d = 0; d < D; ++d) {
#pragma unroll
                        for (int mb = 0; mb < MBF; ++mb)
                            acc[mb] = M::mma(fr[(a.fo.B + ((mb * D + d) * NB + nb) * 4 + s4) * 64 + lane], gv[d][s4],
                                             acc[mb]);
                        if constexpr (TAIL)
                            tacc = M::mma4(fr[(a.fo.BT + (d * NB + nb) * 4 + s4) * 64 + lane], gv[d][s4], tacc);
                    }
                }
            }
            if constexpr (TAIL) {
                // rows 16(MB2-1) + i summed over the 4 lane groups; register 0
                // of lane group g of the last block holds row 16(MB2-1) + g
#pragma unroll
                for (int i = 0; i < 4; ++i) tacc[i] = sum_xor32(sum_xor16(tacc[i]));
                const T tv = g == 0 ? tacc[0] : g == 1 ? tacc[1] : g == 2 ? tacc[2] : tacc[3];
                acc[MB2 - 1][0] = c0[MB2 - 1][0] + tv;
            }
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) hb[ks] = acc[ks >> 2][ks & 3];
        };

        int nplain = a.iters;
        bool keep_learned = false;   // this column's element passed the check
        T hl[KSH];
        if (LEARN && a.warm) {
            // learned start H_0 = A0 U_w + Delta, then Picard iterations until
            // the update of an element passes the tolerance.  The decision is
            // per element (it depends on the element's data only, so a ghost
            // copy and its source cell always agree and face fluxes stay
            // conservative); elements that do not pass are recomputed by the
            // plain predictor from its own start (the fallback).
            const T inv_sc = T(1.0 / fabs(a.Uw[slot]));   // 1 / rho0 of the column's element
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) {
                const int mode = 4 * ks + g;
                hb[ks] = c0[ks >> 2][ks & 3] + (mode < Nst ? T(a.warm[((long)cq * Nst + mode) * a.NS + slot]) : T(0));
            }
            bool pass_e = false;
            if (a.tol < 0.f) {
                // diagnostic: exactly learned_iters iterations, no check
                for (int k = 0; k < a.learned_iters; ++k) picard();
                pass_e = true;
            } else {
                // an element that passes keeps the iterate it passed with
                // while the rest of its wavefront iterates on, so its result
                // does not depend on the other elements of the wavefront
                for (int k = 0; k < a.learned_iters; ++k) {
                    T prev[KSH];
#pragma unroll
                    for (int ks = 0; ks < KSH; ++ks) prev[ks] = hb[ks];
                    picard();
                    T r = T(0);
#pragma unroll
                    for (int ks = 0; ks < KSH; ++ks) r = nanmax(r, T(fabs(hb[ks] - prev[ks])));
                    r *= inv_sc;
                    if constexpr (Q == 4) {
                        // max over the 16 lanes of the element: its 4 columns x 4 lane groups
                        constexpr int QX0 = sizeof(T) == 8 ? 4 : 1, QX1 = sizeof(T) == 8 ? 8 : 2;
                        r = nanmax(r, __shfl_xor(r, QX0, 64));
                        r = nanmax(r, __shfl_xor(r, QX1, 64));
                        r = nanmax(r, __shfl_xor(r, 16, 64));
                        r = nanmax(r, __shfl_xor(r, 32, 64));
                    } else {
                        // max over the 4 lane groups, then over the element's 5 columns
                        r = nanmax(r, __shfl_xor(r, 16, 64));
                        r = nanmax(r, __shfl_xor(r, 32, 64));
                        const int eq = ce < EW ? ce : 0;
                        T re = r;
#pragma unroll
                        for (int q = 0; q < Q; ++q) re = nanmax(re, __shfl(r, (lane & 48) | col_of(eq, q), 64));
                        r = ce < EW ? re : T(0);   // the padding column always passes
                    }
                    if (pass_e) {
#pragma unroll
                        for (int ks = 0; ks < KSH; ++ks) hb[ks] = prev[ks];
                    } else {
                        pass_e = r <= T(a.tol);   // NaN fails
                    }
                    if (__all(pass_e)) break;
                }
            }
            // one representative lane per element (lane group 0, variable 0)
            const bool rep = g == 0 && cq == 0 && active;
            const unsigned long long n_el = __popcll(__ballot(rep));
            const unsigned long long n_fb = __popcll(__ballot(rep && !pass_e));
            if (lane == 0 && n_el) {   // workgroup totals in LDS, one global update per workgroup
                atomicAdd(&wg_stats[0], n_el);
                if (n_fb) atomicAdd(&wg_stats[1], n_fb);
            }
            if (__all(pass_e)) {
                nplain = 0;
            } else {
                keep_learned = pass_e;
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) {
                    hl[ks] = hb[ks];
                    hb[ks] = ks < KSU ? hu[ks] : T(0);
                }
            }
        }
        for (int it = 0; it < nplain; ++it) picard();
        if (keep_learned) {
#pragma unroll
            for (int ks = 0; ks < KSH; ++ks) hb[ks] = hl[ks];
        }

        // final evaluation: corrector volume term
        v4 vol[MBS];
#pragma unroll
        for (int mbs = 0; mbs < MBS; ++mbs) vol[mbs] = v4{0, 0, 0, 0};
#pragma unroll 1
        for (int nb = 0; nb < NB; ++nb) {
            T gv[D][4];
            nodal_flux(nb, gv);
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                if (!FULL && 16 * nb + s4 >= Mst) continue;
#pragma unroll
                for (int d = 0; d < D; ++d)
#pragma unroll
                    for (int mbs = 0; mbs < MBS; ++mbs)
                        vol[mbs] = M::mma(fr[(a.fo.S + ((mbs * D + d) * NB + nb) * 4 + s4) * 64 + lane], gv[d][s4],
                                          vol[mbs]);
            }
        }
        if (active) {
#pragma unroll
            for (int mbs = 0; mbs < MBS; ++mbs)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int mode = 16 * mbs + 4 * r + g;
                    if (mode < Ns) a.Vol[vol_idx<P, D>(mode, cq, slot, a.NS)] = vol[mbs][r];
                }
            if (a.Hw_dbg) {
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks) {
                    const int mode = 4 * ks + g;
                    if (mode < Nst) a.Hw_dbg[((long)cq * Nst + mode) * a.NS + slot] = double(hb[ks]);
                }
            }
        }
        // face traces: Tr[ds][node] = Vbd[ds] * H_w
#pragma unroll 1
        for (int ds = 0; ds < 2 * D; ++ds)
#pragma unroll
            for (int mbd = 0; mbd < MBD; ++mbd) {
                v4 t = v4{0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KSH; ++ks)
                    t = M::mma(fr[(a.fo.T + (ds * MBD + mbd) * KSH + ks) * 64 + lane], hb[ks], t);
                if (active) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int fn = 16 * mbd + 4 * r + g;
                        if (fn < Mbd) a.Tr[tr_idx<P, D>(ds, fn, cq, slot, a.NS)] = t[r];
                    }
                }
            }
    }
    if (LEARN && a.warm) {
        __syncthreads();
        if (threadIdx.x == 0 && wg_stats[0]) {
            atomicAdd(a.stats, wg_stats[0]);
            if (wg_stats[1]) atomicAdd(a.stats + 1, wg_stats[1]);
        }
    }
}

template <typename T, int P, int D, int WPB, bool LEARN>
__global__ void __launch_bounds__(64 * WPB) k_predictor_nl(PredMArgs<T> a) {
    // 3D p = 3 fragments exceed the LDS: no body
    if constexpr (D == 2 || P <= 2) predictor_nl_body<T, P, D, LEARN>(a);
}
template <typename T, int P, int D, int WPB>
inline const void* predictor_nl_kernel(bool learn) {
    return learn ? (const void*)k_predictor_nl<T, P, D, WPB, true> : (const void*)k_predictor_nl<T, P, D, WPB, false>;
}

// ---------------------------------------------------------------------------
// Face moment row k of direction d: S_k = sum_m Mkbdp[d](k,m) Fnum[m], one
// FMA chain over the face nodes in a fixed order.  The corrector of both
// neighbours of a face and the flux registers all call this, so they see
// bitwise equal moments: the left cell (its + face) subtracts S, the right
// cell (its - face) adds (-1)^{k_d} S since Mkbdm[d](k,m) = (-1)^{k_d}
// Mkbdp[d](k,m) exactly, and the mode-0 rows telescope.
template <typename T, int Mbd>
__device__ __forceinline__ T face_moment(const T* __restrict__ Mrow, const T* F) {
    T s = T(0);
#pragma unroll
    for (int mb = 0; mb < Mbd; ++mb) s = fma(Mrow[mb], F[mb], s);
    return s;
}
// Mkbdp in the arithmetic type of the face moments (the FP32 copy is the
// FP64 table rounded to nearest once, on the host)
template <typename T>
__device__ __forceinline__ const T* moment_rows(const double* __restrict__ tab, const TabOff& to) {
    if constexpr (sizeof(T) == 8) return tab + to.Mkbdp;
    else return reinterpret_cast<const float*>(tab + to.Mkbdpf);
}

// ---------------------------------------------------------------------------
// Face flux: LLF at every face quadrature node (AmrDG_NumericalFlux.tpp:4-41)
// from the traces of the left (side +) and right (side -) cells, computed
// once per face.  One thread per (direction, face node, face), face index
// fastest: the trace reads and the flux writes of a wavefront are
// consecutive slots / faces of one SoA row.
template <typename T>
struct FaceArgs {
    const T* __restrict__ Tr;
    T* __restrict__ Fnum[3];
    const int* __restrict__ face_lr[3];
    long NF[3];
    long NS;
    double gamma;
};

template <typename T, int P, int D>
__global__ void __launch_bounds__(256) k_faceflux(FaceArgs<T> a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Mbd = Dm::Mbd;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0;
#pragma unroll
    for (int dd = 0; dd < D - 1; ++dd)
        if (d == dd && t >= a.NF[dd] * Mbd) { t -= a.NF[dd] * Mbd; d = dd + 1; }
    const long nf = a.NF[d];
    if (t >= nf * Mbd) return;
    const int mb = (int)(t / nf);
    const long f = t - (long)mb * nf;
    const long sl = a.face_lr[d][2 * f], sr = a.face_lr[d][2 * f + 1];
    const long NS = a.NS;
    T uL[Q], uR[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uL[q] = a.Tr[tr_idx<P, D>(2 * d + 1, mb, q, sl, NS)];
        uR[q] = a.Tr[tr_idx<P, D>(2 * d + 0, mb, q, sr, NS)];
    }
    const T gm = T(a.gamma), gm1 = T(a.gamma - 1.0);
    const T riL = recip(uL[0]), riR = recip(uR[0]);
    const T pL = pressure<D>(uL, riL, gm1), pR = pressure<D>(uR, riR, gm1);
    const T cL = sqrt(gm * (pL / fmax(uL[0], T(kNumLimit))));
    const T cR = sqrt(gm * (pR / fmax(uR[0], T(kNumLimit))));
    const T unL = uL[d + 1] * riL, unR = uR[d + 1] * riR;
    const T dfL = fmax(fmax(fabs(unL - cL), fabs(unL)), fabs(unL + cL));
    const T dfR = fmax(fmax(fabs(unR - cR), fabs(unR)), fabs(unR + cR));
    const T C = fmax(fabs(dfL), fabs(dfR));
    T* __restrict__ out = a.Fnum[d] + (long)mb * nf + f;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const T fL = flux_q<D>(d, q, uL, uL[q], riL, pL);
        const T fR = flux_q<D>(d, q, uR, uR[q], riR, pR);
        out[(long)q * Mbd * nf] = T(0.5) * ((fL + fR) - C * (uR[q] - uL[q]));
    }
}

// ---------------------------------------------------------------------------
// Corrector on valid cells (AmrDG.tpp:652-742):
//   U_w <- [M U_w + Vol - sum_d (Dt/dx_d) (S_hi - (-1)^{k_d} S_lo)] / M
// with each face product subtracted / added on its own, in FP64.
template <typename T>
struct CorrArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;
    const T* __restrict__ Vol;
    const T* __restrict__ Fnum[3];
    const int* __restrict__ cell_face[3][2];
    const int* __restrict__ valid_list;
    const uint8_t* __restrict__ slot_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NF[3];
    long NS;
    long nvalid;
};

// Element of a row whose base is wavefront-uniform, at a 32-bit per-lane
// byte offset: the global access takes the scalar-base + vector-offset form,
// with no 64-bit address arithmetic per access.
template <typename T>
__device__ __forceinline__ T& row_at(T* base, unsigned byte_off) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ const T& row_at(const T* base, unsigned byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
__host__ __device__ __forceinline__ long pad_wave(long n) { return (n + 63) & ~63L; }

// One thread per (q, valid cell), the cells padded to whole wavefronts so
// that q is wavefront-uniform: every array offset is a uniform row plus the
// 32-bit byte offset of the slot or face (slots and faces below 2^29).
template <typename T, int P, int D>
__global__ void __launch_bounds__(256) k_corrector(CorrArgs<T> a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Mbd = Dm::Mbd;
    // directions whose face values are held in registers at once (all, but
    // one at a time for the 3D FP64 form: 2 x 27 doubles per direction)
    constexpr int DPRE = (D == 3 && sizeof(T) == 8) ? 1 : D;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nvp = pad_wave(a.nvalid);
    if (t >= nvp * Q) return;
    const int q = __builtin_amdgcn_readfirstlane((int)(t / nvp));
    const long i = t - (long)q * nvp;
    if (i >= a.nvalid) return;
    const long NS = a.NS;
    const double* __restrict__ Mkc = a.tab + a.to.Mkc;
    const double* __restrict__ sgn = a.tab + a.to.sgn;
    // every load of the thread is issued before its arithmetic (two memory
    // round trips: the lists, then the modes and face values)
    unsigned flo[D], fhi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        flo[d] = (unsigned)a.cell_face[d][0][i] * (unsigned)sizeof(T);
        fhi[d] = (unsigned)a.cell_face[d][1][i] * (unsigned)sizeof(T);
    }
    const long slot = a.valid_list[i];
    const int lev = a.slot_lev[slot];
    const unsigned sb = (unsigned)slot * 8u, sbt = (unsigned)slot * (unsigned)sizeof(T);
    double* __restrict__ Uq = a.Uw + (long)q * Ns * NS;
    const T* __restrict__ Vq = a.Vol + (long)q * Ns * NS;
    double u[Ns], mk[Ns];
    T vol[Ns];
#pragma unroll
    for (int k = 0; k < Ns; ++k) {
        u[k] = row_at(Uq + k * NS, sb);
        vol[k] = row_at(Vq + k * NS, sbt);
        mk[k] = Mkc[k];   // held: the stores to Uw below may alias the table for the compiler
    }
    T Fh[DPRE][Mbd], Fl[DPRE][Mbd];
    auto load_dir = [&](int d, T* fh, T* fl) {
        const long nf = a.NF[d];
        const T* __restrict__ Fq = a.Fnum[d] + (long)q * Mbd * nf;
#pragma unroll
        for (int mb = 0; mb < Mbd; ++mb) { fh[mb] = row_at(Fq + mb * nf, fhi[d]); fl[mb] = row_at(Fq + mb * nf, flo[d]); }
    };
#pragma unroll
    for (int d = 0; d < DPRE; ++d) load_dir(d, Fh[d], Fl[d]);
    __builtin_amdgcn_sched_barrier(0);
    const double Dt = a.ts->Dt;
    double r[Ns];
#pragma unroll
    for (int k = 0; k < Ns; ++k) r[k] = mk[k] * u[k] + double(vol[k]);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int s = DPRE == D ? d : 0;
        if (DPRE < D && d > 0) load_dir(d, Fh[0], Fl[0]);
        const double cdd = Dt / a.geo.dx[lev][d];
        const T* __restrict__ Mbp = moment_rows<T>(a.tab, a.to) + (long)d * Ns * Mbd;
#pragma unroll
        for (int k = 0; k < Ns; ++k) {
            const T sh = face_moment<T, Mbd>(Mbp + k * Mbd, Fh[s]);
            const T sl = face_moment<T, Mbd>(Mbp + k * Mbd, Fl[s]);
            r[k] -= cdd * double(sh);
            r[k] += cdd * (sgn[d * Ns + k] * double(sl));
        }
    }
#pragma unroll
    for (int k = 0; k < Ns; ++k) row_at(Uq + k * NS, sb) = r[k] / mk[k];
}

// H_w of the copy ghosts from their sources (the kept H_w only: the step
// never reads a copy ghost's predictor result).  One thread per (ghost, row).
__global__ void __launch_bounds__(256) k_copy_slots(double* __restrict__ buf, int nrow, long NS,
                                                   const int* __restrict__ dst, const int* __restrict__ src, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nrow) return;
    const int r = (int)(t / n);
    const long g = t - (long)r * n;
    buf[(long)r * NS + dst[g]] = buf[(long)r * NS + src[g]];
}

// ---------------------------------------------------------------------------
// Flux-register face moments (FP64): coarse faces J * Mkbd{p|m}[d] Fnum =
// J (+-1)^{k_d} S, fine faces J * P_flux_{high|low}[d][child] Fnum whose mode
// 0 row equals the quadrature weights, so it is taken from S as well
// (AmrDG.tpp:604-647).
// Mode n of a register moment from the face's node fluxes F (already loaded).
template <typename T, int P, int D>
__device__ __forceinline__ double register_moment_of(const double* __restrict__ tab, const TabOff& to, const T* F,
                                                     int d, int kind, int sel, int n, double jac) {
    using Dm = Dims<P, D>;
    constexpr int Ns = Dm::Ns, Mbd = Dm::Mbd, NFC = Dm::NFC;
    const T* __restrict__ Mbp = moment_rows<T>(tab, to) + (long)d * Ns * Mbd;
    if (kind == 0 || n == 0) {
        const double sv = double(face_moment<T, Mbd>(Mbp + n * Mbd, F));
        return jac * ((kind == 1 || (sel & 1)) ? sv : tab[to.sgn + d * Ns + n] * sv);
    }
    const double* __restrict__ M = tab + ((sel & 1) ? to.Pfh : to.Pfl) + ((long)d * NFC + (sel >> 1)) * Ns * Mbd;
    double s = 0.0;
#pragma unroll
    for (int mb = 0; mb < Mbd; ++mb) s = fma(M[n * Mbd + mb], double(F[mb]), s);
    return jac * s;
}

template <typename T, int P, int D>
__device__ __forceinline__ double register_moment(const double* __restrict__ tab, const TabOff& to,
                                                  const T* __restrict__ Fnum_d, long nf, int d, int kind, int sel,
                                                  long f, int q, int n, double jac) {
    constexpr int Mbd = Dims<P, D>::Mbd;
    T F[Mbd];
    const T* __restrict__ pf = Fnum_d + (long)q * Mbd * nf + f;
#pragma unroll
    for (int mb = 0; mb < Mbd; ++mb) F[mb] = pf[mb * nf];
    return register_moment_of<T, P, D>(tab, to, F, d, kind, sel, n, jac);
}

template <typename T>
struct MomArgs {
    const double* __restrict__ tab;
    TabOff to;
    const T* __restrict__ Fnum[3];
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

template <typename T, int P, int D>
__global__ void __launch_bounds__(256) k_moments(MomArgs<T> a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.NM * Q * Ns) return;
    const int qn = (int)(t / a.NM);
    const long j = t - (long)qn * a.NM;
    const int q = qn / Ns, n = qn - q * Ns;
    const int d = a.mom_d[j], lev = a.mom_lev[j];
    const double jac = (a.ts->Dt / 2.0) * (a.geo.area[lev][d] / (double)(1 << (D - 1)));
    a.Mom[((long)q * Ns + n) * a.NM + j] =
        register_moment<T, P, D>(a.tab, a.to, a.Fnum[d], a.NF[d], d, a.mom_kind[j], a.mom_sel[j], a.mom_face[j], q, n, jac);
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
    if (t >= a.nrc * Q * Ns) return;
    const int qn = (int)(t / a.nrc);
    const long r = t - (long)qn * a.nrc;
    const int q = qn / Ns, n = qn - q * Ns;
    const long slot = a.rc_slot[r];
    const double inv_jac = (double)(1 << D) / a.geo.vol[a.rc_lev[r]];
    double corr = 0.0;
    const int st = a.rc_start[r], cnt = a.rc_count[r];
    for (int e = 0; e < cnt; ++e) corr += a.re_coef[st + e] * a.Mom[((long)q * Ns + n) * a.NM + a.re_mom[st + e]];
    a.Uw[((long)q * Ns + n) * a.NS + slot] += inv_jac * (a.tab[a.to.Minv + n] * corr);
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
__device__ __forceinline__ void avgdown_item(const double* __restrict__ tab, const TabOff& to, double* __restrict__ Uw,
                                             const int* __restrict__ coarse, const int* __restrict__ fine,
                                             long NS, long ngroup, long t);

template <int P, int D>
__global__ void __launch_bounds__(256) k_avgdown(AvgArgs a) {
    avgdown_item<P, D>(a.tab, a.to, a.Uw, a.coarse, a.fine, a.NS, a.ngroup,
                       (long)blockIdx.x * blockDim.x + threadIdx.x);
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
__device__ __forceinline__ void ghost_item(const double* __restrict__ tab, const TabOff& to, double* __restrict__ Uw,
                                           const int* __restrict__ g_slot, const int* __restrict__ g_src,
                                           const uint8_t* __restrict__ g_kind, const uint8_t* __restrict__ g_child,
                                           long NS, long nghost, long t);

template <int P, int D>
__global__ void __launch_bounds__(256) k_ghosts(GhostArgs a) {
    ghost_item<P, D>(a.tab, a.to, a.Uw, a.g_slot, a.g_src, a.g_kind, a.g_child, a.NS, a.nghost,
                     (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// CFL: lambda at spatial node 0 (AmrDG.h:1300-1314), min over valid cells.
// Each thread takes cfl_k cells (at most kCflBlocks blocks), the block
// minimum goes through a wavefront shuffle and LDS, and one atomicMin per
// block on the IEEE bits (positive doubles order like unsigned integers)
// publishes it: same-address device atomics serialise (one per block of 256
// cells had cost ~90 us per step on 524 288 cells, 2 048 blocks).
constexpr long kCflBlocks = 1024;
__host__ __device__ __forceinline__ int cfl_per_thread(long nvalid) {
    const long k = (nvalid + kCflBlocks * 256 - 1) / (kCflBlocks * 256);
    return (int)(k < 1 ? 1 : k);
}

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
    int cfl_k;   // cells per thread
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_cfl(CflArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    double dt = 1.0e300;
    for (int j = 0; j < a.cfl_k; ++j) {
        const long i = ((long)blockIdx.x * a.cfl_k + j) * blockDim.x + threadIdx.x;
        if (i >= a.nvalid) break;
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
        dt = fmin(dt, a.cfl * (a.geo.dx_avg[lev] / lmax));
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
    __shared__ double wmin[4];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmin(m, wmin[w]);
        if (m < 1.0e300) atomicMin(&a.ts->dtmin_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// Finalize: optional t += Dt, then Dt = c_dt * min, clamped to T - t.
__global__ void k_dt_finalize(TimeState* ts, double c_dt, double T, int advance) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (advance) {
        ts->t = ts->t + ts->Dt;
        ts->nstep = ts->nstep + 1u;
    }
    double m = __longlong_as_double((long long)ts->dtmin_bits);
    double Dt = c_dt * m;
    if (T - ts->t < Dt) Dt = T - ts->t;
    ts->Dt = Dt;
    ts->dtmin_bits = 0x7ff0000000000000ull;  // +inf for the next reduction
}

// ---------------------------------------------------------------------------
// Fused step epilogue.  Device functions of the reflux, average-down, ghost
// and CFL work items, and two kernels that run disjoint work items side by
// side so that one step needs few launches:
//   k_reflux_avg : reflux rows (moments computed inline) || average down of
//                  the finest level (coarse cells written are disjoint: the
//                  reflux touches uncovered cells only)
//   k_ghost_cfl  : ghost fill || CFL reduction; the last CFL block to finish
//                  advances t and sets the next Dt (no extra launch).
template <typename T>
struct EpiArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;
    const T* __restrict__ Fnum[3];
    long NF[3];
    // reflux rows
    const int* __restrict__ mom_face;
    const uint8_t* __restrict__ mom_d;
    const uint8_t* __restrict__ mom_kind;
    const uint8_t* __restrict__ mom_sel;
    const uint8_t* __restrict__ mom_lev;
    const int* __restrict__ rc_slot;
    const int* __restrict__ rc_start;
    const int* __restrict__ rc_count;
    const uint8_t* __restrict__ rc_lev;
    const int* __restrict__ re_mom;
    const double* __restrict__ re_coef;
    long nrc;
    // average down groups
    const int* __restrict__ ad_coarse;
    const int* __restrict__ ad_fine;
    long ngroup;
    // ghosts
    const int* __restrict__ g_slot;
    const int* __restrict__ g_src;
    const uint8_t* __restrict__ g_kind;
    const uint8_t* __restrict__ g_child;
    long nghost;
    // CFL
    const int* __restrict__ valid_list;
    const uint8_t* __restrict__ slot_lev;
    long nvalid;
    int cfl_k;   // cells per CFL thread
    unsigned ncfl_blocks;
    double cfl, c_dt, Tend;
    TimeState* ts;
    LevelGeo geo;
    long NS;
    double gamma;
    unsigned nblk_a;   // blocks of the first work type
};

// Threads of the reflux rows: 2D one per (row, q, n), a short latency-bound
// item; 3D one per (row, q) for every mode, so that the node fluxes of each
// register entry are gathered once, not once per mode (3D rows are many and
// their entries' fluxes 27 scattered values each).
__host__ __device__ __forceinline__ long reflux_threads(long nrc, int Q, int Ns, int D) {
    return nrc * Q * (D == 3 ? 1 : Ns);
}

// Every mode sums its entries in row order with the same arithmetic in both
// mappings.
template <typename T, int P, int D>
__device__ __forceinline__ void reflux_item(const EpiArgs<T>& a, long t) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Mbd = Dm::Mbd;
    constexpr int NPT = D == 3 ? Ns : 1;   // modes per thread
    if (t >= a.nrc * Q * (Ns / NPT)) return;
    const int qn = (int)(t / a.nrc);
    const long r = t - (long)qn * a.nrc;
    const int q = NPT == Ns ? qn : qn / Ns, n0 = NPT == Ns ? 0 : qn - q * Ns;
    const long slot = a.rc_slot[r];
    const int lev = a.rc_lev[r];
    const double inv_jac = (double)(1 << D) / a.geo.vol[lev];
    const double Dt = a.ts->Dt;
    double corr[NPT];
#pragma unroll
    for (int j = 0; j < NPT; ++j) corr[j] = 0.0;
    const int st = a.rc_start[r], cnt = a.rc_count[r];
    for (int e = 0; e < cnt; ++e) {
        const long j = a.re_mom[st + e];
        const int d = a.mom_d[j], ml = a.mom_lev[j], kind = a.mom_kind[j], sel = a.mom_sel[j];
        const double jac = (Dt / 2.0) * (a.geo.area[ml][d] / (double)(1 << (D - 1)));
        const double coef = a.re_coef[st + e];
        T F[Mbd];
        const T* __restrict__ pf = a.Fnum[d] + (long)q * Mbd * a.NF[d] + a.mom_face[j];
#pragma unroll
        for (int mb = 0; mb < Mbd; ++mb) F[mb] = pf[mb * a.NF[d]];
#pragma unroll
        for (int k = 0; k < NPT; ++k)
            corr[k] += coef * register_moment_of<T, P, D>(a.tab, a.to, F, d, kind, sel, n0 + k, jac);
    }
#pragma unroll
    for (int k = 0; k < NPT; ++k)
        a.Uw[((long)q * Ns + n0 + k) * a.NS + slot] += inv_jac * (a.tab[a.to.Minv + n0 + k] * corr[k]);
}

// Threads of the average down: groups padded to whole wavefronts, so that the
// variable of a wavefront is uniform.
__host__ __device__ __forceinline__ long avg_groups_padded(long ngroup) { return (ngroup + 63) & ~63L; }

// Children of an average-down group held in registers at once: the largest
// divisor of 2^D whose modes fit 48 doubles.
template <int NCH, int Ns>
__host__ __device__ constexpr int avg_child_block() {
    return NCH * Ns <= 48 ? NCH : (NCH / 2) * Ns <= 48 ? NCH / 2 : (NCH / 4) * Ns <= 48 ? NCH / 4 : 1;
}

// Threads of an average down over ngroup groups.  2D: one per (group, q, n),
// a short latency-bound item (the launch is small); 3D: one per (group, q)
// computing every mode, so that each child mode is read once per thread
// (8 children x Np_s modes would otherwise be re-read by Np_s threads).
__host__ __device__ __forceinline__ long avg_threads(long ngroup, int Q, int Ns, int D) {
    return avg_groups_padded(ngroup) * Q * (D == 3 ? 1 : Ns);
}

// u_c[n] = 2^-D Minv[n] sum_c sum_k P_fc[c](n,k) u_f[c][k], group fastest and
// padded to whole wavefronts: (q, n) or q is wavefront-uniform and P_fc is
// read with scalar loads; every mode n sums its children in order, one FMA
// chain per child, in both mappings.
template <int P, int D>
__device__ __forceinline__ void avgdown_item(const double* __restrict__ tab, const TabOff& to, double* __restrict__ Uw,
                                             const int* __restrict__ coarse, const int* __restrict__ fine,
                                             long NS, long ngroup, long t) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, NCH = Dm::NCH;
    constexpr int NPT = D == 3 ? Ns : 1;   // modes per thread
    const long gp = avg_groups_padded(ngroup);
    if (t >= gp * Q * (Ns / NPT)) return;
    const int qn = __builtin_amdgcn_readfirstlane((int)(t / gp));
    const long g = t - (long)qn * gp;
    if (g >= ngroup) return;
    const int q = NPT == Ns ? qn : qn / Ns, n0 = NPT == Ns ? 0 : qn - q * Ns;
    const double* __restrict__ Pfc = tab + to.Pfc;
    long fs[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) fs[c] = fine[g * NCH + c];
    const long cs = coarse[g];
    // children held at once: one memory round trip per block that fits 48 doubles
    constexpr int CB = NPT == 1 ? NCH : avg_child_block<NCH, Ns>();
    double s[NPT];
#pragma unroll
    for (int j = 0; j < NPT; ++j) s[j] = 0.0;
#pragma unroll 1
    for (int c0 = 0; c0 < NCH; c0 += CB) {
        double u[CB][Ns];
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
            for (int k = 0; k < Ns; ++k) u[c][k] = Uw[((long)q * Ns + k) * NS + fs[c0 + c]];
#pragma unroll
        for (int j = 0; j < NPT; ++j)
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < Ns; ++k) v = fma(Pfc[((c0 + c) * Ns + n0 + j) * Ns + k], u[c][k], v);
                s[j] += v;
            }
    }
#pragma unroll
    for (int j = 0; j < NPT; ++j)
        Uw[((long)q * Ns + n0 + j) * NS + cs] = (1.0 / NCH) * (tab[to.Minv + n0 + j] * s[j]);
}

// Threads of a ghost fill over nghost ghosts: 2D one per (ghost, q, n), a
// short latency-bound item; 3D one per (ghost, q) for every mode, so that an
// interpolated ghost reads its parent's modes once, not once per mode.
__host__ __device__ __forceinline__ long ghost_threads(long nghost, int Q, int Ns, int D) {
    return nghost * Q * (D == 3 ? 1 : Ns);
}

// ghost slot mode n: copy, or Minv[n] sum_k P_cf[child](n,k) u_c[k], the same
// arithmetic per mode in both mappings
template <int P, int D>
__device__ __forceinline__ void ghost_item(const double* __restrict__ tab, const TabOff& to, double* __restrict__ Uw,
                                           const int* __restrict__ g_slot, const int* __restrict__ g_src,
                                           const uint8_t* __restrict__ g_kind, const uint8_t* __restrict__ g_child,
                                           long NS, long nghost, long t) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    constexpr int NPT = D == 3 ? Ns : 1;   // modes per thread
    if (t >= nghost * Q * (Ns / NPT)) return;
    const int qn = (int)(t / nghost);
    const long g = t - (long)qn * nghost;
    const int q = NPT == Ns ? qn : qn / Ns, n0 = NPT == Ns ? 0 : qn - q * Ns;
    const long s = g_slot[g], src = g_src[g];
    const int kind = g_kind[g], mask = kind >> 2;
    double v[NPT];
    if ((kind & 3) == 0) {
#pragma unroll
        for (int j = 0; j < NPT; ++j) v[j] = Uw[((long)q * Ns + n0 + j) * NS + src];
    } else {
        const double* __restrict__ Pcf = tab + to.Pcf + (long)g_child[g] * Ns * Ns;
        double u[Ns];
#pragma unroll
        for (int k = 0; k < Ns; ++k) u[k] = Uw[((long)q * Ns + k) * NS + src];
#pragma unroll
        for (int j = 0; j < NPT; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < Ns; ++k) acc = fma(Pcf[(n0 + j) * Ns + k], u[k], acc);
            v[j] = tab[to.Minv + n0 + j] * acc;
        }
    }
    // transmissive boundary: mirror image, mode n times (-1)^{k_d} per mirrored d
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (mask & (1 << d)) {
#pragma unroll
            for (int j = 0; j < NPT; ++j) v[j] *= tab[to.sgn + d * Ns + n0 + j];
        }
#pragma unroll
    for (int j = 0; j < NPT; ++j) Uw[((long)q * Ns + n0 + j) * NS + s] = v[j];
}

template <typename T, int P, int D>
__global__ void __launch_bounds__(256) k_reflux_avg(EpiArgs<T> a) {
    const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x < a.nblk_a) {
        reflux_item<T, P, D>(a, t0);
    } else {
        const long t = t0 - (long)a.nblk_a * blockDim.x;
        avgdown_item<P, D>(a.tab, a.to, a.Uw, a.ad_coarse, a.ad_fine, a.NS, a.ngroup, t);
    }
}

template <typename T, int P, int D>
__global__ void __launch_bounds__(256) k_ghost_cfl(EpiArgs<T> a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    if (blockIdx.x < a.nblk_a) {
        const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
        ghost_item<P, D>(a.tab, a.to, a.Uw, a.g_slot, a.g_src, a.g_kind, a.g_child, a.NS, a.nghost, t);
        return;
    }
    // CFL (AmrDG.h:1300-1314), cfl_k cells per thread, + last-block finalize
    // (AmrDG.h:1018-1029)
    double dt = 1.0e300;
    for (int j = 0; j < a.cfl_k; ++j) {
        const long i = ((long)(blockIdx.x - a.nblk_a) * a.cfl_k + j) * blockDim.x + threadIdx.x;
        if (i >= a.nvalid) break;
        const long slot = a.valid_list[i];
        const int lev = a.slot_lev[slot];
        const double* __restrict__ phi0 = a.tab + a.to.phi0;
        double u[Q][Ns];   // all modes in one memory round trip
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int n = 0; n < Ns; ++n) u[q][n] = a.Uw[((long)q * Ns + n) * a.NS + slot];
        __builtin_amdgcn_sched_barrier(0);
        double s[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double v = 0.0;
#pragma unroll
            for (int n = 0; n < Ns; ++n) v += u[q][n] * phi0[n];
            s[q] = v;
        }
        const double rinv = 1.0 / s[0];
        const double prs = pressure<D>(s, rinv, a.gamma - 1.0);
        const double c = sqrt(a.gamma * (prs / fmax(s[0], kNumLimit)));
        double lmax = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) lmax = fmax(lmax, fabs(s[d + 1] / s[0]) + c);
        dt = fmin(dt, a.cfl * (a.geo.dx_avg[lev] / lmax));
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
    __shared__ double wmin[4];
    __shared__ unsigned last;
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        // one thread per block publishes the block minimum, then counts the
        // block in: the device-scope fence (an L2 write-back on a multi-XCD
        // part) is paid once per block, not once per wavefront
        double m = wmin[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmin(m, wmin[w]);
        if (m < 1.0e300) atomicMin(&a.ts->dtmin_bits, (unsigned long long)__double_as_longlong(m));
        __threadfence();
        last = (atomicAdd(&a.ts->done, 1u) == a.ncfl_blocks - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        const unsigned long long bits = atomicAdd(&a.ts->dtmin_bits, 0ull);
        const double tn = a.ts->t + a.ts->Dt;
        double Dt = a.c_dt * __longlong_as_double((long long)bits);
        if (a.Tend - tn < Dt) Dt = a.Tend - tn;
        a.ts->t = tn;
        a.ts->Dt = Dt;
        a.ts->nstep = a.ts->nstep + 1u;
        a.ts->dtmin_bits = 0x7ff0000000000000ull;
        a.ts->done = 0u;
        __threadfence();
    }
}

// ---------------------------------------------------------------------------
// TVB limiter of the finest level (Limiter_w + Limiter_linear_tvb,
// AmrDG.h:1657-1855; minmodB / minmod, AmrDG_Limiter.tpp:41-84; the
// characteristic matrices pde_EV_Lmatrix / pde_EV_Rmatrix,
// Compressible_Euler.cpp:138-277, evaluated at U_w mode 0), in place, one
// thread per valid cell.  A cell reads its mode 0 and linear modes and its
// neighbours' mode 0: a neighbour that is a same-level ghost is read at its
// source cell (what FillBoundary copies there), a coarse-fine or boundary
// ghost keeps the value of the last ghost fill, as in the reference.  Mode 0
// is never written, so the cells are independent.  The arithmetic is not
// contracted into FMAs and follows the reference's expression order, so the
// threshold, sign and minimum decisions see the same rounded values.
struct LimArgs {
    double* __restrict__ Uw;
    const int* __restrict__ cells;   // finest-level valid slots
    const int* __restrict__ nbr;     // [d][side][cell] neighbour slots (same-level ghosts resolved)
    unsigned long long* __restrict__ count;   // cells limited (accumulated)
    const TimeState* __restrict__ ts;
    long NS, n;
    double thr;      // AMR_TVB_C[lev] * M * h^2, h = min dx of the level
    double gamma;
    int t_limit;     // limit when nstep % t_limit == 0
    int lin[3];      // lin_mode_idx (AmrDG.tpp:174-184)
};

template <int D>
__device__ __forceinline__ void tvb_ev(int d, const double* u, double g_, double L[D + 2][D + 2],
                                       double R[D + 2][D + 2]) {
#pragma clang fp contract(off)
    constexpr int Q = D + 2;
    // Soundspeed / Pressure with the clamps (Compressible_Euler.h:562-604)
    double prs = 0.0;
    for (int e = 0; e < D; ++e) prs += (u[e + 1] * u[e + 1]) / u[0];
    prs *= (-0.5);
    prs += u[D + 1];
    prs *= (g_ - 1.0);
    prs = fmax(prs, kNumLimit);
    const double c = sqrt(g_ * (prs / fmax(u[0], kNumLimit)));
    const double u1 = u[1] / u[0], u2 = u[2] / u[0], u3 = D == 3 ? u[3] / u[0] : 0.0;
    const double ek = D == 2 ? 0.5 * ((u1 * u1) + (u2 * u2)) : 0.5 * ((u1 * u1) + (u2 * u2) + (u3 * u3));
    const double hh = ((c * c) / (g_ - 1.0)) + ek;
    const double b = 1.0 / (2.0 * (c * c));
    const double gm = g_ - 1.0;
    const double ph = gm * ek;
    for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j) { L[i][j] = 0.0; R[i][j] = 0.0; }
    const double ud = d == 0 ? u1 : (d == 1 ? u2 : u3);
    // rows 0..2 of L and columns 0..2 of R: acoustic and entropy fields
    L[0][0] = b * (ph + c * ud);
    L[1][0] = 1.0 - 2.0 * b * ph;
    L[2][0] = b * (ph - c * ud);
    const double uv[3] = {u1, u2, u3};
    for (int e = 0; e < D; ++e) {
        L[0][1 + e] = e == d ? -b * (gm * uv[e] + c) : -b * gm * uv[e];
        L[1][1 + e] = 2.0 * b * gm * uv[e];
        // the reference writes -b*(g*u) for the first component past the
        // direction (u2 for x, u1 for y and z) and -b*g*u for the other
        L[2][1 + e] = e == d ? -b * (gm * uv[e] - c)
                             : (e == (d == 0 ? 1 : 0) ? -b * (gm * uv[e]) : -b * gm * uv[e]);
    }
    L[0][D + 1] = b * gm;
    L[1][D + 1] = -2.0 * b * gm;
    L[2][D + 1] = b * gm;
    R[0][0] = 1.0; R[0][1] = 1.0; R[0][2] = 1.0;
    for (int e = 0; e < D; ++e) {
        R[1 + e][0] = e == d ? uv[e] - c : uv[e];
        R[1 + e][1] = uv[e];
        R[1 + e][2] = e == d ? uv[e] + c : uv[e];
    }
    R[D + 1][0] = hh - c * ud;
    R[D + 1][1] = ek;
    R[D + 1][2] = hh + c * ud;
    // shear fields (rows / columns 3 and 4), as the reference lists them
    if (D == 2) {
        if (d == 0) {
            L[3][0] = u2; L[3][2] = -1.0;
            R[2][3] = -1.0; R[3][3] = -u2;
        } else {
            L[3][0] = -u1; L[3][1] = 1.0;
            R[1][3] = 1.0; R[3][3] = u1;
        }
    } else {
        if (d == 0) {
            L[3][0] = u2; L[3][2] = -1.0;
            L[4][0] = -u3; L[4][3] = 1.0;
            R[2][3] = -1.0; R[Q - 1][3] = -u2;
            R[3][4] = 1.0; R[Q - 1][4] = u3;
        } else if (d == 1) {
            L[3][0] = -u1; L[3][1] = 1.0;
            L[4][0] = u3; L[4][3] = -1.0;
            R[1][3] = 1.0; R[Q - 1][3] = u1;
            R[3][4] = -1.0; R[Q - 1][4] = -u3;
        } else {
            L[3][0] = u1; L[3][1] = -1.0;
            L[4][0] = -u2; L[4][2] = 1.0;
            R[1][3] = -1.0; R[Q - 1][3] = -u1;
            R[2][4] = 1.0; R[Q - 1][4] = u2;
        }
    }
}

__device__ __forceinline__ double tvb_minmod(double a1, double a2, double a3) {
    const bool same = (signbit(a1) == signbit(a2)) && (signbit(a2) == signbit(a3));
    const double sign = signbit(a1) ? -1.0 : 1.0;
    if (same) return sign * fmin(fmin(fabs(a1), fabs(a2)), fabs(a3));
    return 0.0;
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_limiter_tvb(LimArgs a) {
#pragma clang fp contract(off)
    using Dm = Dims<P, D>;
    constexpr int Q = D + 2, Ns = Dm::Ns;
    __shared__ unsigned int nlim;
    if (threadIdx.x == 0) nlim = 0u;
    __syncthreads();
    if (a.t_limit > 1 && (a.ts->nstep % (unsigned)a.t_limit) != 0u) return;   // uniform over the grid
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) {
        const long s = a.cells[i];
        double u0[Q];
        for (int q = 0; q < Q; ++q) u0[q] = a.Uw[((long)q * Ns) * a.NS + s];
        double res[D][Q];
        bool troubled = false;
        for (int d = 0; d < D; ++d) {
            const long sm = a.nbr[(long)(2 * d) * a.n + i], sp = a.nbr[(long)(2 * d + 1) * a.n + i];
            double um[Q], up[Q], ul[Q];
            for (int q = 0; q < Q; ++q) {
                um[q] = a.Uw[((long)q * Ns) * a.NS + sm];
                up[q] = a.Uw[((long)q * Ns) * a.NS + sp];
                ul[q] = a.Uw[((long)q * Ns + a.lin[d]) * a.NS + s];
            }
            double L[Q][Q], R[Q][Q];
            tvb_ev<D>(d, u0, a.gamma, L, R);
            double vw[Q];
            for (int q = 0; q < Q; ++q) {
                double Dmv = 0.0, Dpv = 0.0, Du = 0.0;
                for (int r = 0; r < Q; ++r) {
                    Dmv += 0.5 * L[q][r] * (u0[r] - um[r]);
                    Dpv += 0.5 * L[q][r] * (up[r] - u0[r]);
                    Du += L[q][r] * ul[r];
                }
                vw[q] = fabs(Du) <= a.thr ? Du : tvb_minmod(Du, Dmv, Dpv);
                troubled = troubled || (vw[q] != Du);
            }
            for (int q = 0; q < Q; ++q) {
                double sum = 0.0;
                for (int r = 0; r < Q; ++r) sum += R[q][r] * vw[r];
                res[d][q] = sum;
            }
        }
        if (troubled) {
            for (int q = 0; q < Q; ++q)
                for (int n = 1; n < Ns; ++n) {
                    double v = 0.0;
                    for (int d = 0; d < D; ++d)
                        if (n == a.lin[d]) v = res[d][q];
                    a.Uw[((long)q * Ns + n) * a.NS + s] = v;
                }
            atomicAdd(&nlim, 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && nlim) atomicAdd(a.count, (unsigned long long)nlim);
}

// ---------------------------------------------------------------------------
// Refinement tags of one level (AmrDG::AMR_tag_cell_refinement,
// AmrDG_TagCellRef.tpp:4-52), one thread per valid cell, one byte out:
//   crit 1: the reference's Kelvin-Helmholtz indicator
//           (Compressible_Euler::pde_tag_cell_refinement,
//           Compressible_Euler.h:645-679): some characteristic slope
//           |L_d U_lin,d| exceeds AMR_TVB_C[lev] * amr_c[lev] * h^2 (the
//           troubled branch of minmodB);
//   crit 2: density gradient (no counterpart in the reference): the cell
//           mean density differs from a face neighbour's by more than
//           amr_c[lev] times its own.
// Neighbours are read at the slots beside the cell (ghost slots hold the
// last ghost fill, which FillBoundary would reproduce; a same-level copy is
// read at its source cell, which holds the same values).
struct TagArgs {
    const double* __restrict__ Uw;
    const int* __restrict__ valid_list;
    const int* __restrict__ cell_face[3][2];
    const int* __restrict__ face_lr[3];
    uint8_t* __restrict__ out;
    long NS, vbegin, n;
    int crit;
    double thr;      // crit 1: AMR_TVB_C * amr_c * h^2; crit 2: amr_c
    double gamma;
    int lin[3];
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_tag(TagArgs a) {
#pragma clang fp contract(off)
    using Dm = Dims<P, D>;
    constexpr int Q = D + 2, Ns = Dm::Ns;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const long v = a.vbegin + i;
    const long s = a.valid_list[v];
    bool tag = false;
    if (a.crit == 1) {
        double u0[Q];
        for (int q = 0; q < Q; ++q) u0[q] = a.Uw[((long)q * Ns) * a.NS + s];
        for (int d = 0; d < D && !tag; ++d) {
            double ul[Q];
            for (int q = 0; q < Q; ++q) ul[q] = a.Uw[((long)q * Ns + a.lin[d]) * a.NS + s];
            double L[Q][Q], R[Q][Q];
            tvb_ev<D>(d, u0, a.gamma, L, R);
            for (int q = 0; q < Q; ++q) {
                double Du = 0.0;
                for (int r = 0; r < Q; ++r) Du += L[q][r] * ul[r];
                if (!(fabs(Du) <= a.thr)) tag = true;
            }
        }
    } else {
        const double r0 = a.Uw[s];
        for (int d = 0; d < D; ++d) {
            const long sl = a.face_lr[d][2 * (long)a.cell_face[d][0][v]];
            const long sr = a.face_lr[d][2 * (long)a.cell_face[d][1][v] + 1];
            const double j = fmax(fabs(a.Uw[sr] - r0), fabs(r0 - a.Uw[sl]));
            if (j > a.thr * r0) tag = true;
        }
    }
    a.out[i] = tag ? 1 : 0;
}

// RemakeLevel / MakeNewLevelFromCoarse data of one new level
// (AmrDG_Level.tpp:52-80,120-170 through FillPatchTwoLevels /
// InterpFromCoarseLevel with the L2ProjInterp mapper): every slot (valid and
// ghost) of the new level takes the old level's valid cell at the same place
// (kind 0, read from the old field), else the L2 projection of the new
// coarse level's parent cell (kind 1, AmrDG_Interpolater.tpp:344-364);
// kind >> 2 marks the mirrored directions of a transmissive boundary ghost.
// One thread per (entry, variable).
struct RemakeArgs {
    const double* __restrict__ tab;
    TabOff to;
    double* __restrict__ Uw;            // new field
    const double* __restrict__ Uw_old;  // old field
    const int* __restrict__ slot;
    const int* __restrict__ src;
    const uint8_t* __restrict__ kind;
    const uint8_t* __restrict__ child;
    long NS, NS_old, n;
};

template <int P, int D>
__global__ void __launch_bounds__(256) k_remake(RemakeArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = D + 2, Ns = Dm::Ns;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n * Q) return;
    const int q = (int)(t / a.n);
    const long e = t - (long)q * a.n;
    const long s = a.slot[e], src = a.src[e];
    const int kind = a.kind[e], mask = kind >> 2;
    double v[Ns];
    if ((kind & 3) == 0) {
        for (int n = 0; n < Ns; ++n) v[n] = a.Uw_old[((long)q * Ns + n) * a.NS_old + src];
    } else {
        const double* __restrict__ Pcf = a.tab + a.to.Pcf + (long)a.child[e] * Ns * Ns;
        double u[Ns];
        for (int k = 0; k < Ns; ++k) u[k] = a.Uw[((long)q * Ns + k) * a.NS + src];
        for (int n = 0; n < Ns; ++n) {
            double acc = 0.0;
            for (int k = 0; k < Ns; ++k) acc = fma(Pcf[n * Ns + k], u[k], acc);
            v[n] = a.tab[a.to.Minv + n] * acc;
        }
    }
    for (int d = 0; d < D; ++d)
        if (mask & (1 << d))
            for (int n = 0; n < Ns; ++n) v[n] *= a.tab[a.to.sgn + d * Ns + n];
    for (int n = 0; n < Ns; ++n) a.Uw[((long)q * Ns + n) * a.NS + s] = v[n];
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

// Second stage of the totals on the device: out[q] = sum over b of
// partial[q][b], added in the order b = 0, 1, ... (the order of the host
// loop that reads the same partials back), so both give the same bits.
// One workgroup of 256 threads: the partials pass through LDS in chunks of
// 256 per variable, and thread q adds its variable's chunk.
__global__ void __launch_bounds__(256) k_totals_finish(const double* __restrict__ partial, int nb, int Q,
                                                       double* __restrict__ out) {
    __shared__ double chunk[5][256];
    double acc = 0.0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int b = b0 + (int)threadIdx.x;
        for (int q = 0; q < Q; ++q) chunk[q][threadIdx.x] = b < nb ? partial[(long)q * nb + b] : 0.0;
        __syncthreads();
        if ((int)threadIdx.x < Q) {
            const int m = min(256, nb - b0);
            for (int i = 0; i < m; ++i) acc += chunk[threadIdx.x][i];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < Q) out[threadIdx.x] = acc;
}

}  // namespace adg

namespace adg {
// ---------------------------------------------------------------------------
// Composite Lp error of one conserved variable over the uncovered cells
// (LpNorm_DG_AMR, AmrDG.h:1399-1578): (P+1)^D Gauss-Legendre points per cell,
// cell weight vol/2^D, against the exact solution at time t.  Exact
// solutions: the isentropic vortex (initial condition 0) and the density
// wave (5) translated by their uniform velocity on the periodic domain, and
// the uniform state (2).
struct LpArgs {
    const double* __restrict__ tab;
    TabOff to;
    const double* __restrict__ Uw;
    const int* __restrict__ list;        // uncovered cells
    const double* __restrict__ centre;   // their cell centres [n][3]
    const uint8_t* __restrict__ slot_lev;
    double* __restrict__ partial;
    LevelGeo geo;
    long NS, n;
    int q, p, ic;
    double prm[8];
    double gamma, t;
    double plo[3], phi[3];
};

template <int D>
__device__ void exact_state(const LpArgs& a, const double* xin, double* u) {
    double x[3] = {xin[0], xin[1], D > 2 ? xin[2] : 0.0};
    const double vel5[3] = {1.0, 0.5, 0.25};
    for (int d = 0; d < D; ++d) {
        const double v = a.ic == 0 ? (d == 0 ? a.prm[3] : (d == 1 ? a.prm[4] : 0.0)) : (a.ic == 5 ? vel5[d] : 0.0);
        const double L = a.phi[d] - a.plo[d];
        double sh = x[d] - v * a.t - a.plo[d];
        sh -= L * floor(sh / L);
        x[d] = a.plo[d] + sh;
    }
    for (int q = 0; q < 5; ++q) u[q] = 0.0;
    const double g = a.gamma;
    if (a.ic == 0) {
        const double beta = a.prm[0], xc = a.prm[1], yc = a.prm[2];
        const double r2 = (x[0] - xc) * (x[0] - xc) + (x[1] - yc) * (x[1] - yc);
        const double rho = pow(1.0 - ((g - 1.0) * beta * beta / (8.0 * g * M_PI * M_PI)) * exp(1.0 - r2), 1.0 / (g - 1.0));
        const double prs = pow(rho, g);
        const double f = (beta / (2.0 * M_PI)) * exp((1.0 - r2) / 2.0);
        const double u1 = a.prm[3] - (x[1] - yc) * f, u2 = a.prm[4] + (x[0] - xc) * f;
        u[0] = rho;
        u[1] = rho * u1;
        u[2] = rho * u2;
        u[D + 1] = prs / (g - 1.0) + 0.5 * rho * (u1 * u1 + u2 * u2);
    } else if (a.ic == 5) {
        double prod = 1.0;
        for (int d = 0; d < D; ++d) prod *= sin(2.0 * M_PI * (x[d] - a.plo[d]) / (a.phi[d] - a.plo[d]));
        const double rho = 1.0 + a.prm[0] * prod;
        double k = 0.0;
        u[0] = rho;
        for (int d = 0; d < D; ++d) { u[d + 1] = rho * vel5[d]; k += vel5[d] * vel5[d]; }
        u[D + 1] = a.prm[1] / (g - 1.0) + 0.5 * rho * k;
    } else {
        const double rho = a.prm[0];
        const double vel[3] = {a.prm[1], a.prm[2], D > 2 ? a.prm[3] : 0.0};
        double k = 0.0;
        u[0] = rho;
        for (int d = 0; d < D; ++d) { u[d + 1] = rho * vel[d]; k += vel[d] * vel[d]; }
        u[D + 1] = a.prm[D == 3 ? 4 : 3] / (g - 1.0) + 0.5 * rho * k;
    }
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_lp_error(LpArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Ns = Dm::Ns;
    __shared__ double red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double cell = 0.0;
    if (i < a.n) {
        const long s = a.list[i];
        const int lev = a.slot_lev[s];
        const TabOff::LpRule& r = a.to.lp[a.p == 2 ? 1 : 0];
        const int Ms = r.M;
        const double* Vs = a.tab + r.Vs;
        const double* ws = a.tab + r.ws;
        const double* xs = a.tab + r.xs;
        double U[Ns];
        for (int n = 0; n < Ns; ++n) U[n] = a.Uw[((long)a.q * Ns + n) * a.NS + s];
        for (int m = 0; m < Ms; ++m) {
            double uh = 0.0;
            for (int n = 0; n < Ns; ++n) uh += U[n] * Vs[m * Ns + n];
            double x[3] = {0.0, 0.0, 0.0}, ue[5];
            for (int d = 0; d < D; ++d) x[d] = 0.5 * a.geo.dx[lev][d] * xs[m * D + d] + a.centre[i * 3 + d];
            exact_state<D>(a, x, ue);
            const double e = fabs(ue[a.q] - uh);
            cell += (a.p == 1 ? e : pow(e, (double)a.p)) * ws[m];
        }
        cell *= a.geo.vol[lev] / (double)(1 << D);
    }
    red[threadIdx.x] = cell;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[blockIdx.x] = red[0];
}

// Layout probe of the 16x16x4 MFMA builtins (one wavefront, in-bounds):
// mode 0: A[lane] = lane+1, B = 1 on lanes 0..15 (kidx 0) else 0
// mode 1: A = 1 on lanes 0..15 else 0, B[lane] = lane+1
// out[lane*4 + r] = accumulator register r of lane.
template <typename T>
__global__ void k_probe_mfma(T* out, int mode) {
    using M = MF<T>;
    const int l = threadIdx.x & 63;
    T a, b;
    if (mode == 0) { a = T(l + 1); b = (l < 16) ? T(1) : T(0); }
    else { a = (l < 16) ? T(1) : T(0); b = T(l + 1); }
    typename M::v4 c = {0, 0, 0, 0};
    c = M::mma(a, b, c);
    for (int r = 0; r < 4; ++r) out[l * 4 + r] = c[r];
}
}  // namespace adg
