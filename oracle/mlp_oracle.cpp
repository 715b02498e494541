// CPU oracle of the learned-proposal MLP: TEST INFRASTRUCTURE ONLY (the
// checker for the GPU forward/backward kernels; nothing in the product
// links it).  The MLP has no counterpart in the reference; its definition is
// the one documented in oracle.h.  FP64 throughout; with bf16 != 0 each
// matrix product's operands are first rounded to bfloat16 (round to nearest
// even of the FP32 value), the rounding the GPU kernels apply when they
// stage operands for the matrix cores.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

const int kIn = 8, kH = 128, kOut = 2;
const int kDims[6] = {kIn, kH, kH, kH, kH, kOut};

// offsets of W_l (row-major [in][out]) and b_l in the packed parameters
void layout(int l, size_t &w, size_t &b) {
    size_t o = 0;
    for (int k = 0; k < l; ++k) o += size_t(kDims[k]) * kDims[k + 1] + kDims[k + 1];
    w = o;
    b = o + size_t(kDims[l]) * kDims[l + 1];
}

double bf16(double v) {
    float f = float(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) != 0x7F800000u) u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return double(f);
}

double gelu(double z) { return 0.5 * z * (1.0 + std::erf(z / std::sqrt(2.0))); }
double gelu_d(double z) {
    const double cdf = 0.5 * (1.0 + std::erf(z / std::sqrt(2.0)));
    const double pdf = std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
    return cdf + z * pdf;
}

struct Net {
    // per layer l = 0..4: input activations a[l] (n x dims[l]) and
    // pre-activations z[l] (n x dims[l+1])
    std::vector<double> a[5], z[5];
};

void forward(const float *params, const float *x, int32_t n, int q, Net &N) {
    N.a[0].assign(size_t(n) * kIn, 0.0);
    for (size_t i = 0; i < N.a[0].size(); ++i) N.a[0][i] = x[i];
    for (int l = 0; l < 5; ++l) {
        const int din = kDims[l], dout = kDims[l + 1];
        size_t wo, bo;
        layout(l, wo, bo);
        std::vector<double> W(size_t(din) * dout), A(N.a[l]);
        for (size_t i = 0; i < W.size(); ++i) W[i] = q ? bf16(params[wo + i]) : double(params[wo + i]);
        if (q)
            for (double &v : A) v = bf16(v);
        N.z[l].assign(size_t(n) * dout, 0.0);
        for (int32_t s = 0; s < n; ++s)
            for (int o = 0; o < dout; ++o) {
                double acc = 0.0;
                for (int k = 0; k < din; ++k) acc += A[size_t(s) * din + k] * W[size_t(k) * dout + o];
                N.z[l][size_t(s) * dout + o] = acc + double(params[bo + o]);
            }
        if (l < 4) {
            N.a[l + 1].resize(N.z[l].size());
            for (size_t i = 0; i < N.z[l].size(); ++i) N.a[l + 1][i] = gelu(N.z[l][i]);
        }
    }
}

}  // namespace

extern "C" {

void or_mlp_forward(const float *params, const float *x, int32_t n, int bf16_operands, double *out) {
    Net N;
    forward(params, x, n, bf16_operands, N);
    std::memcpy(out, N.z[4].data(), sizeof(double) * size_t(n) * kOut);
}

double or_mlp_grad(const float *params, const float *x, const float *t, int32_t n, int bf16_operands,
                   double *grad) {
    const int q = bf16_operands;
    Net N;
    forward(params, x, n, q, N);
    std::memset(grad, 0, sizeof(double) * OR_MLP_PARAMS);
    // loss = mean_s sum_o (out - t)^2
    double loss = 0.0;
    std::vector<double> dz(size_t(n) * kOut);
    for (size_t i = 0; i < dz.size(); ++i) {
        const double d = N.z[4][i] - double(t[i]);
        loss += d * d;
        dz[i] = 2.0 * d / double(n);
    }
    loss /= double(n);
    for (int l = 4; l >= 0; --l) {
        const int din = kDims[l], dout = kDims[l + 1];
        size_t wo, bo;
        layout(l, wo, bo);
        std::vector<double> A(N.a[l]), D(dz), W(size_t(din) * dout);
        for (size_t i = 0; i < W.size(); ++i) W[i] = q ? bf16(params[wo + i]) : double(params[wo + i]);
        if (q) {
            for (double &v : A) v = bf16(v);
            for (double &v : D) v = bf16(v);
        }
        for (int k = 0; k < din; ++k)  // dW = A^T dz
            for (int o = 0; o < dout; ++o) {
                double acc = 0.0;
                for (int32_t s = 0; s < n; ++s) acc += A[size_t(s) * din + k] * D[size_t(s) * dout + o];
                grad[wo + size_t(k) * dout + o] = acc;
            }
        for (int o = 0; o < dout; ++o) {  // db = column sums of dz (unrounded)
            double acc = 0.0;
            for (int32_t s = 0; s < n; ++s) acc += dz[size_t(s) * dout + o];
            grad[bo + o] = acc;
        }
        if (l == 0) break;
        std::vector<double> dzp(size_t(n) * din);  // dz of the layer below: (dz W^T) * gelu'(z)
        for (int32_t s = 0; s < n; ++s)
            for (int k = 0; k < din; ++k) {
                double acc = 0.0;
                for (int o = 0; o < dout; ++o) acc += D[size_t(s) * dout + o] * W[size_t(k) * dout + o];
                dzp[size_t(s) * din + k] = acc * gelu_d(N.z[l - 1][size_t(s) * din + k]);
            }
        dz.swap(dzp);
    }
    return loss;
}

}  // extern "C"
