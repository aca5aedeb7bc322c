// adg_mlp.hpp -- learned warm start of the space-time predictor (no
// counterpart in the reference): a 3-layer MLP with GELU maps each element's
// U_w modes, local CFL number and Dt/dx to a starting guess for H_w.
//
// Surrogate definition (per element, tau_d = Dt/dx_d, rho0 = U_w[rho][0]):
//   x = [U_w/rho0 (q-major, Q*Np_s values), nu, tau_0 .. tau_{D-1}],
//       nu = sum_d tau_d (|u_d| + c) at the cell-average state;
//   y = (H_w* - A0 U_w) / (rho0 max_d tau_d), q-major (Q*Np_st values),
//       H_w* the converged predictor, A0 = Mk_h_w^-1 Mk_pred;
//   x and y are standardised per column (dataset mean / deviation).
// Network: h1 = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2), y = W3 h2 + b3,
// widths nin_p -> 128 -> 128 -> nout_p (nin_p, nout_p = widths padded to 16),
// gelu(a) = a/2 (1 + erf(a/sqrt 2)) (erf to 1.5e-7).  Every matrix product runs on the
// 16x16x16 BF16 MFMA with FP32 accumulation; operands are rounded to BF16
// (round to nearest even) as they are loaded, master weights stay FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "adg_kernels.hpp"

namespace adg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpHidden = 128;

// BF16 bits of an FP32 value, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short bf16_bits(float v) {
    unsigned u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ bf16x4 pack_bf16(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)bf16_bits(a);
    r[1] = (short)bf16_bits(b);
    r[2] = (short)bf16_bits(c);
    r[3] = (short)bf16_bits(d);
    return r;
}
// The same rounding (to nearest, ties to even) by the gfx950 conversion
// instruction v_cvt_pk_bf16_f32, two values per instruction.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x4 pack_bf16_hw(float a, float b, float c, float d) {
    const bf16x2_t lo = __builtin_convertvector(f32x2_t{a, b}, bf16x2_t);
    const bf16x2_t hi = __builtin_convertvector(f32x2_t{c, d}, bf16x2_t);
    typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(bf16x4, u32x2_t{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)});
}
// C += A B, 16x16x16: A[i][4k'+t] on lane i+16k' (t = 0..3 packed),
// B[4k'+t][j] on lane j+16k', C row 4g+r of lane group g in register r.
__device__ __forceinline__ f32x4 mma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, formula
// 7.1.26: erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2),
// t = 1 / (1 + p x) for x >= 0, odd extension; |error| <= 1.5e-7, far below
// the BF16 resolution of the activations it feeds.  About a third of the
// instructions of the library erf, which dominates the inference kernel.
__device__ __forceinline__ float erf_as(float x) {
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
    float y = fmaf(t, 1.061405429f, -1.453152027f);
    y = fmaf(t, y, 1.421413741f);
    y = fmaf(t, y, -0.284496736f);
    y = fmaf(t, y, 0.254829592f);
    y = 1.0f - t * y * __expf(-ax * ax);
    return copysignf(y, x);
}
__device__ __forceinline__ float gelu_f(float a) { return 0.5f * a * (1.0f + erf_as(a * 0.70710678118654752f)); }
// Twice the same GELU, a + |a| erf(|a|/sqrt 2), for the inference kernel: no
// sign transfer, one FMA for the final combination; the factor 1/2 is folded
// into the packed weights of the layer that consumes it (an exact power-of-two
// scaling of BF16 values).
__device__ __forceinline__ float gelu_w(float a) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f * 0.70710678118654752f, fabsf(a), 1.0f));
    float y = fmaf(t, 1.061405429f, -1.453152027f);
    y = fmaf(t, y, 1.421413741f);
    y = fmaf(t, y, -0.284496736f);
    y = fmaf(t, y, 0.254829592f);
    const float r = a * 0.84932180028801904f;   // exp(-a^2/2) = 2^(-r^2), r = a sqrt(log2(e)/2)
    const float e = fmaf(-t * y, __builtin_amdgcn_exp2f(-r * r), 1.0f);   // erf(|a|/sqrt 2)
    return fmaf(fabsf(a), e, a);
}
__device__ __forceinline__ float dgelu_f(float a) {
    return 0.5f * (1.0f + erf_as(a * 0.70710678118654752f)) + a * 0.39894228040143268f * __expf(-0.5f * a * a);
}

// ---------------------------------------------------------------------------
// Tiled GEMM C = op(A) op(B), M, N, K multiples of 16; one wavefront per
// 16x16 output tile and K chunk.  op(A) = A (M x K, A[i*lda+k]) or A^T
// (A[k*lda+i]); op(B) = B (K x N, B[k*ldb+j]) or B^T (B[j*ldb+k]).
// Epilogues: 0 store; 1 C = acc + bias[j], C2 = gelu(C) (C2 nullable); 2 atomic add
// (split K); 3 C = acc * gelu'(aux[i][j]).
struct GemmArgs {
    const float* A;
    long lda;
    int transA;
    const float* B;
    long ldb;
    int transB;
    float* C;
    long ldc;
    int M, N, K, ksplit, epi;
    const float* bias;
    float* C2;
    const float* aux;
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs a) {
    const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int tm_n = a.M >> 4, tn_n = a.N >> 4;
    const long ntile = (long)tm_n * tn_n;
    if (w >= ntile * a.ksplit) return;   // whole wavefronts exit together
    const int kc = (int)(w / ntile);
    const long tile = w - (long)kc * ntile;
    const int tm = (int)(tile / tn_n), tn = (int)(tile - (long)tm * tn_n);
    const int lane = threadIdx.x & 63, i = lane & 15, kg = lane >> 4;
    const int ks_n = a.K >> 4;
    const int ks0 = (int)((long)ks_n * kc / a.ksplit), ks1 = (int)((long)ks_n * (kc + 1) / a.ksplit);
    const long row = 16 * tm + i, col = 16 * tn + i;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s = ks0; s < ks1; ++s) {
        const long k0 = 16 * s + 4 * kg;
        float av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            av[t] = a.transA ? a.A[(k0 + t) * a.lda + row] : a.A[row * a.lda + k0 + t];
            bv[t] = a.transB ? a.B[col * a.ldb + k0 + t] : a.B[(k0 + t) * a.ldb + col];
        }
        acc = mma_bf16(pack_bf16(av[0], av[1], av[2], av[3]), pack_bf16(bv[0], bv[1], bv[2], bv[3]), acc);
    }
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long ii = 16 * tm + 4 * g + r, jj = 16 * tn + c;
        float v = acc[r];
        float* dst = a.C + ii * a.ldc + jj;
        if (a.epi == 0) {
            *dst = v;
        } else if (a.epi == 1) {
            v += a.bias[jj];
            *dst = v;
            if (a.C2) a.C2[ii * a.ldc + jj] = gelu_f(v);
        } else if (a.epi == 2) {
            atomicAdd(dst, v);
        } else {
            *dst = v * dgelu_f(a.aux[ii * a.ldc + jj]);
        }
    }
}

// g = 2 (yp - y) / (n nout) on the first nout columns (0 on padding); the
// summed squared error is accumulated into *loss (FP64) per block.
__global__ void __launch_bounds__(256) k_mlp_lossgrad(const float* __restrict__ yp, const float* __restrict__ y,
                                                      float* __restrict__ g, long n, int nout_p, int nout,
                                                      double* __restrict__ loss) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (t < n * nout_p) {
        const int col = (int)(t % nout_p);
        const float e = col < nout ? yp[t] - y[t] : 0.f;
        g[t] = 2.f * e / (float)((double)n * nout);
        e2 = (double)e * e;
    }
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (ws[0] + ws[1] + ws[2] + ws[3]) / ((double)n * nout));
}

// out[j] += sum over rows [r0, r0 + 256) of G[r][j]  (bias gradients)
__global__ void __launch_bounds__(256) k_colsum(const float* __restrict__ G, long n, int ncol, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long r0 = (long)blockIdx.y * 256, r1 = r0 + 256 < n ? r0 + 256 : n;
    float s = 0.f;
    for (long r = r0; r < r1; ++r) s += G[r * ncol + j];
    atomicAdd(out + j, s);
}

// per-column sum and sum of squares (FP64) of an n x ncol row-major array
__global__ void __launch_bounds__(256) k_colstats(const float* __restrict__ X, long n, int ncol,
                                                  double* __restrict__ sum, double* __restrict__ sq) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncol) return;
    const long r0 = (long)blockIdx.y * 1024, r1 = r0 + 1024 < n ? r0 + 1024 : n;
    double s = 0.0, s2 = 0.0;
    for (long r = r0; r < r1; ++r) {
        const double v = X[r * ncol + j];
        s += v;
        s2 += v * v;
    }
    atomicAdd(sum + j, s);
    atomicAdd(sq + j, s2);
}

__global__ void __launch_bounds__(256) k_fill(float* __restrict__ p, long n, float v) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    w[t] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// Random minibatch: row b takes dataset row hash(seed, step, b) mod ndata,
// standardised with the column statistics.
__device__ __forceinline__ uint64_t mix64_d(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__global__ void __launch_bounds__(256) k_gather_batch(const float* __restrict__ X, const float* __restrict__ Y, long ndata,
                                                      int nin_p, int nout_p, const float* __restrict__ norm,
                                                      uint64_t seed, long step, long batch, float* __restrict__ Xb,
                                                      float* __restrict__ Yb) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int w = nin_p + nout_p;
    if (t >= batch * w) return;
    const long b = t / w;
    const int col = (int)(t - b * w);
    const long r = (long)(mix64_d(seed ^ mix64_d((uint64_t)step * 0x100000001ull + (uint64_t)b)) % (uint64_t)ndata);
    const float* mux = norm;
    const float* sdx = norm + nin_p;
    const float* muy = norm + 2 * nin_p;
    const float* sdy = muy + nout_p;
    if (col < nin_p) Xb[b * nin_p + col] = (X[r * nin_p + col] - mux[col]) / sdx[col];
    else {
        const int o = col - nin_p;
        Yb[b * nout_p + o] = (Y[r * nout_p + o] - muy[o]) / sdy[o];
    }
}

// ---------------------------------------------------------------------------
// Surrogate features and targets of every valid element, from U_w and the
// converged predictor H_w (FP64 [q][Np_st][slot]); one thread per element.
struct CollectArgs {
    const double* __restrict__ tab;
    TabOff to;
    const double* __restrict__ Uw;
    const double* __restrict__ Hw;
    const int* __restrict__ valid_list;
    const uint8_t* __restrict__ slot_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NS, nvalid;
    double gamma;
    float* __restrict__ X;   // rows [nvalid][nin_p]
    float* __restrict__ Y;   // rows [nvalid][nout_p]
    int nin_p, nout_p;
};

template <int P, int D>
__device__ __forceinline__ void mlp_features(const double* U, const double* tau, double gamma, double* x) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    const double rho0 = U[0];
    for (int q = 0; q < Q; ++q)
        for (int k = 0; k < Ns; ++k) x[q * Ns + k] = U[q * Ns + k] / rho0;
    double ke = 0.0;
    for (int d = 0; d < D; ++d) ke += U[(d + 1) * Ns] * U[(d + 1) * Ns];
    const double prs = (gamma - 1.0) * (U[(D + 1) * Ns] - 0.5 * ke / rho0);
    const double c = sqrt(gamma * prs / rho0);
    double nu = 0.0;
    for (int d = 0; d < D; ++d) nu += tau[d] * (fabs(U[(d + 1) * Ns] / rho0) + c);
    x[Q * Ns] = nu;
    for (int d = 0; d < D; ++d) x[Q * Ns + 1 + d] = tau[d];
}

template <int P, int D>
__global__ void __launch_bounds__(64) k_mlp_collect(CollectArgs a) {
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns, Nst = Dm::Nst;
    constexpr int NIN = Q * Ns + 1 + D;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nvalid) return;
    const long slot = a.valid_list[i];
    const int lev = a.slot_lev[slot];
    const double Dt = a.ts->Dt;
    double tau[D], tmax = 0.0;
    for (int d = 0; d < D; ++d) { tau[d] = Dt / a.geo.dx[lev][d]; tmax = fmax(tmax, tau[d]); }
    double U[Q * Ns];
    for (int f = 0; f < Q * Ns; ++f) U[f] = a.Uw[(long)f * a.NS + slot];
    double x[NIN];
    mlp_features<P, D>(U, tau, a.gamma, x);
    const double* A0 = a.tab + a.to.A0;
    const double sc = 1.0 / (U[0] * tmax);
    double y[Q * Nst];
    for (int q = 0; q < Q; ++q)
        for (int n = 0; n < Nst; ++n) {
            double c0 = 0.0;
            for (int m = 0; m < Ns; ++m) c0 += A0[n * Ns + m] * U[q * Ns + m];
            y[q * Nst + n] = (a.Hw[((long)q * Nst + n) * a.NS + slot] - c0) * sc;
        }
    // a row with a value that is not a finite float (a predictor that did
    // not converge on a non-physical state) is stored as zeros
    bool ok = true;
    for (int f = 0; f < NIN; ++f) ok = ok && fabs(x[f]) < 1e30;
    for (int o = 0; o < Q * Nst; ++o) ok = ok && fabs(y[o]) < 1e30;
    float* xr = a.X + i * a.nin_p;
    for (int f = 0; f < a.nin_p; ++f) xr[f] = (ok && f < NIN) ? (float)x[f] : 0.f;
    float* yr = a.Y + i * a.nout_p;
    for (int o = 0; o < a.nout_p; ++o) yr[o] = (ok && o < Q * Nst) ? (float)y[o] : 0.f;
}

// ---------------------------------------------------------------------------
// Inference: one wavefront per 16 elements (matrix columns), weights as the A
// operand (pre-packed BF16 fragments staged in LDS), activations as the B
// operand; the accumulator of output block u (lane group g, register r =
// unit 16u + 4g + r, column = element) is the B operand of K step u of the
// next layer, so activations never leave registers.  Writes the warm-start
// increment Delta = rho0 tau_max (mu_y + sd_y y) per output, [o][slot].
struct WarmArgs {
    const uint4* __restrict__ stage;   // packed BF16 fragments [nfrag][64] (8 B each), then the constants
    int f1, f2, f3, nfrag;             // first fragment of each layer
    int stage16;                       // 16-byte units of the staged block
    const double* __restrict__ Uw;
    const uint8_t* __restrict__ slot_lev;
    const TimeState* __restrict__ ts;
    LevelGeo geo;
    long NS, slot_begin, nelem;
    double gamma;
    float* __restrict__ warm;          // [nout][NS]
    const int* __restrict__ slots;     // nullable: the predictor's slot list, slot_begin its first index
};

// Warm-start constants, staged after the fragments: mu_x, 1/sd_x [nin_p],
// mu_y, sd_y [nout_p], b1, b2 [hidden], b3 [nout_p] (zero past nout), padded
// to 16 bytes; in LDS they are followed by the cell sizes of every level
// (kMaxLev x 3 doubles).
template <int NINP, int NOUTP>
__host__ __device__ constexpr int warm_const_floats() {
    return 2 * NINP + 3 * NOUTP + 2 * kMlpHidden;
}
__host__ __device__ constexpr int warm_const_pad(int nfloats) { return (nfloats * 4 + 15) / 16 * 16; }
template <int NINP, int NOUTP>
__host__ __device__ constexpr int warm_const_bytes() {
    return warm_const_pad(warm_const_floats<NINP, NOUTP>()) + (kMaxLev * 3 * 8 + 15) / 16 * 16;
}

// Builds the constant block (layout above) from the normalisation vector
// [mu_x, sd_x, mu_y, sd_y] and the biases; one thread per float.
__global__ void __launch_bounds__(256) k_mlp_warm_consts(const float* __restrict__ norm, const float* __restrict__ b1,
                                                         const float* __restrict__ b2, const float* __restrict__ b3,
                                                         int nin_p, int nout_p, int nout, int nh, int npad,
                                                         float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nnorm = 2 * nin_p + 2 * nout_p, nconst = nnorm + 2 * nh + nout_p;
    if (i >= npad) return;
    float v = 0.f;
    if (i < nnorm) v = (i >= nin_p && i < 2 * nin_p) ? 1.f / norm[i] : norm[i];
    else if (i < nnorm + nh) v = b1[i - nnorm];
    else if (i < nnorm + 2 * nh) v = b2[i - nnorm - nh];
    else if (i < nconst) v = (i - nnorm - 2 * nh < nout) ? b3[i - nnorm - 2 * nh] : 0.f;
    out[i] = v;
}

// Copies n16 16-byte units from global memory to LDS with the global-to-LDS
// loads of gfx950 (global_load_lds_dwordx4): lane l of a wavefront writes
// bytes [16 l, 16 l + 16) past the wavefront's (uniform) LDS address, so no
// registers hold the data in flight.  Visible after s_waitcnt vmcnt(0) and a
// workgroup barrier.
__device__ __forceinline__ void stage_to_lds_dma(unsigned char* dst, const uint4* src, int n16) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (int b = wave * 64; b < n16; b += nwave * 64)
        if (b + lane < n16)
            __builtin_amdgcn_global_load_lds(src + b + lane, (__attribute__((address_space(3))) void*)(dst + (long)b * 16),
                                             16, 0, 0);
}

// NT column tiles of 16 columns per wavefront share every fragment read.  A
// tile holds ELT = 16 elements, or ELT = 8 with columns 8-15 repeating 0-7
// (twice the wavefronts, each with half the activation work: the kernel is
// latency-bound at the sizes of a step, not MFMA-bound).  Layers 1 and 2 keep
// twice the GELU (their consumers' weights are packed at half scale), and
// every accumulator starts at its bias.  The hidden layers have no scheduling
// barrier between their 16-row blocks, so one block's MFMAs overlap the
// previous block's GELU (bitwise the same output, 2-6 % shorter launches than
// with a barrier per block); the output layer keeps one per block.
template <int P, int D, int NINP, int NOUTP, int NT, int ELT>
__device__ __forceinline__ void mlp_warm_body(const WarmArgs& a) {
    static_assert(ELT == 16 || ELT == 8, "8 or 16 elements per tile");
    using Dm = Dims<P, D>;
    constexpr int Q = Dm::Q, Ns = Dm::Ns;
    constexpr int NIN = Q * Ns + 1 + D, NOUT = Q * Dm::Nst, NH = kMlpHidden;
    constexpr int KS1 = NINP / 16, KS2 = NH / 16, UT3 = NOUTP / 16;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const bf16x4* fr = reinterpret_cast<const bf16x4*>(smem_raw);
    const float* cst = reinterpret_cast<const float*>(smem_raw + (long)a.nfrag * 64 * 8);
    double* ldx = reinterpret_cast<double*>(smem_raw + (long)a.nfrag * 64 * 8 +
                                            warm_const_pad(warm_const_floats<NINP, NOUTP>()));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    constexpr int NNORM = 2 * NINP + 2 * NOUTP;
    const float* mux = cst;
    const float* rsdx = cst + NINP;
    const float* muy = cst + 2 * NINP;
    const float* sdy = muy + NOUTP;
    const float* b1 = cst + NNORM;
    const float* b2 = b1 + NH;
    const float* b3 = b2 + NH;
    const long per_wave = (long)ELT * NT;
    const int ce = c & (ELT - 1);   // element of column c
    const Deal deal = make_deal(a.nelem, per_wave, nwave);

    long slot[NT];
    bool active[NT];
    int levs[NT];
    double avgs[NT][Q], uf[NT][KS1][4];
    auto load = [&](long eb) {   // every global load of a tile's features, one memory round trip
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) {
            const long e = eb + (long)ELT * tt + ce;
            active[tt] = e < a.nelem && c < ELT;   // columns past ELT repeat, not stored
            slot[tt] = pred_slot(a.slots, a.slot_begin, e < a.nelem ? e : a.nelem - 1);
            levs[tt] = a.slot_lev[slot[tt]];
            const double* us = a.Uw + slot[tt];
#pragma unroll
            for (int q = 0; q < Q; ++q) avgs[tt][q] = us[(long)q * Ns * a.NS];
            // row 16 s + 4 g + t: the lane part 4 g apart, the rest uniform
            const double* ug = us + (long)(4 * g) * a.NS;
#pragma unroll
            for (int s = 0; s < KS1; ++s)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    uf[tt][s][t] = 16 * s + 4 * g + t < Q * Ns ? ug[(long)(16 * s + t) * a.NS] : 0.0;
        }
    };
    // fragments and constants by DMA, then the first tile's features: one
    // round trip covers both
    stage_to_lds_dma(smem_raw, a.stage, a.stage16);
    long eb;
    bool have = deal_next(deal, 0, wave, per_wave, nwave, &eb);
    if (have) load(eb);
    for (int i = threadIdx.x; i < kMaxLev * 3; i += blockDim.x) ldx[i] = a.geo.dx[i / 3][i % 3];
    // the global-to-LDS loads count in vmcnt: all of them have landed
    // (0x0f70 = vmcnt(0) expcnt(7) lgkmcnt(15) in the gfx9 encoding)
    __builtin_amdgcn_s_waitcnt(0x0f70);
    __syncthreads();
    const double Dt = a.ts->Dt;
    for (long it = 0; have; ++it) {
        float sc[NT];
        bf16x4 xb[NT][KS1];
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) {
            const int lev = levs[tt];
            double tau[D], tmax = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) { tau[d] = Dt / ldx[lev * 3 + d]; tmax = fmax(tmax, tau[d]); }
            const double* avg = avgs[tt];
            const double rho0 = avg[0];
            const double rrho = 1.0 / rho0;
            double ke = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) ke += avg[d + 1] * avg[d + 1];
            const double prs = (a.gamma - 1.0) * (avg[D + 1] - 0.5 * ke * rrho);
            const double cs = sqrt(a.gamma * prs * rrho);
            double nu = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) nu += tau[d] * (fabs(avg[d + 1] * rrho) + cs);
            sc[tt] = (float)(rho0 * tmax);
            // layer 1 input: features 16 s + 4 g + t of element c
#pragma unroll
            for (int s = 0; s < KS1; ++s) {
                float v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int f = 16 * s + 4 * g + t;
                    double xv = 0.0;
                    if (f < Q * Ns) xv = uf[tt][s][t] * rrho;
                    else if (f == Q * Ns) xv = nu;
                    else if (f < NIN) xv = tau[f - Q * Ns - 1];
                    v[t] = f < NIN ? ((float)xv - mux[f]) * rsdx[f] : 0.f;
                }
                xb[tt][s] = pack_bf16_hw(v[0], v[1], v[2], v[3]);
            }
        }
        bf16x4 h1[NT][KS2];
#pragma unroll
        for (int u = 0; u < KS2; ++u) {
            f32x4 acc[NT];
            const f32x4 bias = *reinterpret_cast<const f32x4*>(b1 + 16 * u + 4 * g);
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) acc[tt] = bias;
#pragma unroll
            for (int s = 0; s < KS1; ++s) {
                const bf16x4 w = fr[(a.f1 + u * KS1 + s) * 64 + lane];
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) acc[tt] = mma_bf16(w, xb[tt][s], acc[tt]);
            }
#pragma unroll
            for (int tt = 0; tt < NT; ++tt)
                h1[tt][u] = pack_bf16_hw(gelu_w(acc[tt][0]), gelu_w(acc[tt][1]), gelu_w(acc[tt][2]), gelu_w(acc[tt][3]));
        }
        bf16x4 h2[NT][KS2];
#pragma unroll
        for (int u = 0; u < KS2; ++u) {
            f32x4 acc[NT];
            const f32x4 bias = *reinterpret_cast<const f32x4*>(b2 + 16 * u + 4 * g);
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) acc[tt] = bias;
#pragma unroll
            for (int s = 0; s < KS2; ++s) {
                const bf16x4 w = fr[(a.f2 + u * KS2 + s) * 64 + lane];
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) acc[tt] = mma_bf16(w, h1[tt][s], acc[tt]);
            }
#pragma unroll
            for (int tt = 0; tt < NT; ++tt)
                h2[tt][u] = pack_bf16_hw(gelu_w(acc[tt][0]), gelu_w(acc[tt][1]), gelu_w(acc[tt][2]), gelu_w(acc[tt][3]));
        }
#pragma unroll
        for (int u = 0; u < UT3; ++u) {
            f32x4 acc[NT];
            const f32x4 bias = *reinterpret_cast<const f32x4*>(b3 + 16 * u + 4 * g);
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) acc[tt] = bias;
#pragma unroll
            for (int s = 0; s < KS2; ++s) {
                const bf16x4 w = fr[(a.f3 + u * KS2 + s) * 64 + lane];
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) acc[tt] = mma_bf16(w, h2[tt][s], acc[tt]);
            }
            __builtin_amdgcn_sched_barrier(0);
            const f32x4 my = *reinterpret_cast<const f32x4*>(muy + 16 * u + 4 * g);
            const f32x4 sy = *reinterpret_cast<const f32x4*>(sdy + 16 * u + 4 * g);
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) {
                if (!active[tt]) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = 16 * u + 4 * g + r;
                    if (o < NOUT) a.warm[(long)o * a.NS + slot[tt]] = sc[tt] * fmaf(sy[r], acc[tt][r], my[r]);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // the next tile's loads stay behind this tile's stores
        have = deal_next(deal, it + 1, wave, per_wave, nwave, &eb);
        if (have) load(eb);
    }
}

// One tile per wavefront capped at 128 VGPRs (two 512-thread workgroups
// per CU), or two tiles per wavefront sharing each fragment read.
template <int P, int D, int NINP, int NOUTP>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_mlp_warm1(WarmArgs a) {
    mlp_warm_body<P, D, NINP, NOUTP, 1, 16>(a);
}
template <int P, int D, int NINP, int NOUTP>
__global__ void __launch_bounds__(512) k_mlp_warm2(WarmArgs a) {
    mlp_warm_body<P, D, NINP, NOUTP, 2, 16>(a);
}
template <int P, int D, int NINP, int NOUTP>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_mlp_warm_half(WarmArgs a) {
    mlp_warm_body<P, D, NINP, NOUTP, 1, 8>(a);
}

// Packs layer weights W [rows][cols] (FP32, row-major), times scale (a power
// of two, so exact), into A-operand fragments: fragment (u, s), lane i + 16k
// holds W[16u + i][16s + 4k + t].
__global__ void __launch_bounds__(64) k_mlp_pack(const float* __restrict__ W, int rows, int cols, float scale,
                                                  uint2* __restrict__ out) {
    const int f = blockIdx.x, lane = threadIdx.x;
    const int ks = cols / 16;
    const int u = f / ks, s = f - u * ks;
    const int i = lane & 15, k = lane >> 4;
    const float* w = W + (long)(16 * u + i) * cols + 16 * s + 4 * k;
    uint2 v;
    v.x = (unsigned)bf16_bits(scale * w[0]) | ((unsigned)bf16_bits(scale * w[1]) << 16);
    v.y = (unsigned)bf16_bits(scale * w[2]) | ((unsigned)bf16_bits(scale * w[3]) << 16);
    out[(long)f * 64 + lane] = v;
    (void)rows;
}

}  // namespace adg
