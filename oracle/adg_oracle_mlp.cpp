// adg_oracle_mlp.cpp -- CPU reference of the warm-start MLP (TEST
// INFRASTRUCTURE ONLY: loaded by tests/, never by the product).
//
//   a1 = W1 x + b1, h1 = gelu(a1); a2 = W2 h1 + b2, h2 = gelu(a2);
//   yp = W3 h2 + b3;  loss = mean over rows and the first nout columns of
//   (yp - y)^2;  gelu(a) = a/2 (1 + erf(a / sqrt 2)).
// With bf16 = 1 every matrix-product operand (weights, activations and
// back-propagated gradients) is rounded to bfloat16 (round to nearest even)
// as the MFMA kernels round them; sums are accumulated in double.  With
// bf16 = 0 the arithmetic is plain double (used for finite differences).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
inline double bf(double v, int on) {
    if (!on) return v;
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}
inline double gelu(double a) { return 0.5 * a * (1.0 + std::erf(a * M_SQRT1_2)); }
inline double dgelu(double a) { return 0.5 * (1.0 + std::erf(a * M_SQRT1_2)) + a * std::exp(-0.5 * a * a) * 0.5 * M_2_SQRTPI * M_SQRT1_2; }
}  // namespace

extern "C" double orc_mlp_loss_grad(int nin_p, int nh, int nout_p, int nout, int n, const float* par,
                                    const float* x, const float* y, int bf16, double* grad, double* ypred) {
    const float* W1 = par;
    const float* b1 = W1 + (size_t)nh * nin_p;
    const float* W2 = b1 + nh;
    const float* b2 = W2 + (size_t)nh * nh;
    const float* W3 = b2 + nh;
    const float* b3 = W3 + (size_t)nout_p * nh;
    const size_t npar = (size_t)nh * nin_p + nh + (size_t)nh * nh + nh + (size_t)nout_p * nh + nout_p;
    if (grad) std::memset(grad, 0, npar * sizeof(double));
    double* gW1 = grad;
    double* gb1 = grad ? gW1 + (size_t)nh * nin_p : nullptr;
    double* gW2 = grad ? gb1 + nh : nullptr;
    double* gb2 = grad ? gW2 + (size_t)nh * nh : nullptr;
    double* gW3 = grad ? gb2 + nh : nullptr;
    double* gb3 = grad ? gW3 + (size_t)nout_p * nh : nullptr;
    std::vector<double> a1(nh), h1(nh), a2(nh), h2(nh), yp(nout_p), g3(nout_p), d2(nh), d1(nh);
    double loss = 0.0;
    const double inv = 1.0 / ((double)n * nout);
    for (int r = 0; r < n; ++r) {
        const float* xr = x + (size_t)r * nin_p;
        for (int j = 0; j < nh; ++j) {
            double s = 0.0;
            for (int k = 0; k < nin_p; ++k) s += bf(W1[(size_t)j * nin_p + k], bf16) * bf(xr[k], bf16);
            a1[j] = s + b1[j];
            h1[j] = gelu(a1[j]);
        }
        for (int j = 0; j < nh; ++j) {
            double s = 0.0;
            for (int k = 0; k < nh; ++k) s += bf(W2[(size_t)j * nh + k], bf16) * bf(h1[k], bf16);
            a2[j] = s + b2[j];
            h2[j] = gelu(a2[j]);
        }
        for (int j = 0; j < nout_p; ++j) {
            double s = 0.0;
            for (int k = 0; k < nh; ++k) s += bf(W3[(size_t)j * nh + k], bf16) * bf(h2[k], bf16);
            yp[j] = s + b3[j];
            if (ypred) ypred[(size_t)r * nout_p + j] = yp[j];
        }
        for (int j = 0; j < nout_p; ++j) {
            const double e = j < nout ? yp[j] - y[(size_t)r * nout_p + j] : 0.0;
            loss += e * e * inv;
            g3[j] = 2.0 * e * inv;
        }
        if (!grad) continue;
        for (int j = 0; j < nout_p; ++j) {
            gb3[j] += g3[j];
            for (int k = 0; k < nh; ++k) gW3[(size_t)j * nh + k] += bf(g3[j], bf16) * bf(h2[k], bf16);
        }
        for (int k = 0; k < nh; ++k) {
            double s = 0.0;
            for (int j = 0; j < nout_p; ++j) s += bf(g3[j], bf16) * bf(W3[(size_t)j * nh + k], bf16);
            d2[k] = s * dgelu(a2[k]);
        }
        for (int j = 0; j < nh; ++j) {
            gb2[j] += d2[j];
            for (int k = 0; k < nh; ++k) gW2[(size_t)j * nh + k] += bf(d2[j], bf16) * bf(h1[k], bf16);
        }
        for (int k = 0; k < nh; ++k) {
            double s = 0.0;
            for (int j = 0; j < nh; ++j) s += bf(d2[j], bf16) * bf(W2[(size_t)j * nh + k], bf16);
            d1[k] = s * dgelu(a1[k]);
        }
        for (int j = 0; j < nh; ++j) {
            gb1[j] += d1[j];
            for (int k = 0; k < nin_p; ++k) gW1[(size_t)j * nin_p + k] += bf(d1[j], bf16) * bf(xr[k], bf16);
        }
    }
    return loss;
}
